/* o_gcmc.c -- oracle: serial grand-canonical Monte Carlo (test infrastructure only). */
#include "o_internal.h"
