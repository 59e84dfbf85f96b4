/* o_internal.h -- oracle internals (test infrastructure only). */
#ifndef MAPP_ORACLE_INTERNAL_H
#define MAPP_ORACLE_INTERNAL_H
#include "oracle.h"

struct o_eam {
    int ne, nr, nrho, nf, dmd;
    double dr, drho;
    double* r_phi; /* [ne][ne][nr][4], symmetric copies */
    double* rho;   /* [ne][ne][nr][4]; rho[i][j]: element i's block, position j */
    double* F;     /* [ne][nrho][nf] */
    double* rc;    /* [ne][ne] */
};

#define OPHI(e, i, j) ((double(*)[4])((e)->r_phi + ((size_t)(i) * (e)->ne + (j)) * (size_t)(e)->nr * 4))
#define ORHO(e, i, j) ((double(*)[4])((e)->rho + ((size_t)(i) * (e)->ne + (j)) * (size_t)(e)->nr * 4))
#define OF(e, i) ((e)->F + (size_t)(i) * (e)->nrho * (e)->nf)

/* Periodic ghost ("phantom") bookkeeping shared by the MD and DMD systems:
 * restates Update::list (src/exchange.h:833-866, 1021-1032) for one rank. */
typedef struct {
    int nlcl, nall, cap;
    double* s;       /* [nall][3] fractional coordinates (ghosts shifted) */
    double* x;       /* [nall][3] Cartesian */
    int* src;        /* [nall] atom the ghost was copied from (-1 for locals) */
    int* owner;      /* [nall] owning local atom */
    int8_t* img;     /* [nall][3] accumulated image shift */
    int8_t* sdim;    /* [nall] dimension of the last shift */
    int8_t* sdir;    /* [nall] +1 or -1 */
} o_ghosts;

typedef struct {
    double H[3][3], B[3][3];
    double h[6], b[6];     /* packed lower triangles as in Atoms::update_H */
    double depth_inv[3];
    double vol;
} o_box;

void o_box_set(o_box* bx, const double* H);
void o_x2s(const o_box* bx, const double* x, double* s);     /* with wrap */
void o_s2x(const o_box* bx, const double* s, double* x);
void o_ghosts_build(o_ghosts* g, const o_box* bx, int nlcl, const double* s_lcl, double max_cut);
void o_ghosts_refresh(o_ghosts* g, const o_box* bx, const double* x_lcl);
void o_ghosts_free(o_ghosts* g);
double o_rsq(const double* a, const double* b);
/* Cell<1>::DO (src/cell.h:125-233): calls visit(ctx, i, j) for every local i
 * and every j != i in the 27 stencil cells, in the reference's order. */
void o_cell_do(const o_ghosts* g, const o_box* bx, double max_cut,
               void (*pre)(void*, int), void (*visit)(void*, int, int), void (*post)(void*, int), void* ctx);

/* read-only view of an MD system's atoms (locals then ghosts) and half list */
void o_md_view(const o_md* md, const double** x, const double** s, const uint8_t** elem,
               const int64_t** row, const int32_t** col);

/* EAM spline evaluation at r (p clamped as src/ff_eam.cpp:117-128) */
static inline void o_spline_pm(double v, double inv, int n, size_t* m, double* p)
{
    double pp = v * inv;
    size_t mm = (size_t)pp;
    if (mm > (size_t)(n - 2)) mm = (size_t)(n - 2);
    pp -= (double)mm;
    if (pp > 1.0) pp = 1.0;
    *m = mm; *p = pp;
}
#endif
