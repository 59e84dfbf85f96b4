/* o_gcmc.c -- oracle restatement of single-cell grand-canonical Monte Carlo
 * with the EAM exchange-energy hooks. TEST INFRASTRUCTURE ONLY: the product
 * path never calls this file.
 *
 * Follows, for one process (s_lo = 0, s_hi = 1, every trial point is local):
 *   GCMC::init / box_setup                    src/gcmc.cpp:86-150
 *   SGCMC ctor (27-cell stencil, m = 1)       src/sgcmc.cpp:19-78
 *   SGCMC::box_setup / xchng binning          src/sgcmc.cpp:114-206
 *   find_cell_no / find_cell_coord            src/sgcmc.cpp:232-251
 *   SGCMC::attmpt                             src/sgcmc.cpp:255-327
 *   ins_succ / del_succ (cell-list relink)    src/sgcmc.cpp:331-444
 *   prep_s_x_buff                             src/sgcmc.cpp:448-497
 *   next_iatm / next_jatm_reg / next_jatm_self src/sgcmc.cpp:509-640
 *   reset_tag                                 src/sgcmc.cpp:677-682
 *   ForceFieldEAM::init_xchng                 src/ff_eam.cpp:509-578
 *   pre_xchng_energy / xchng_energy / post    src/ff_eam.cpp:323-483
 *   Random                                    src/random.cpp:7-68
 * Sums run in the reference's order (cells of the stencil in its order, atoms
 * of a cell in ascending index), so dU is the reference's to the last bit up
 * to libm's exp/sqrt.
 *
 * Pinned by tests/test_oracle_gcmc.py: dU of an accepted insertion equals the
 * total-energy difference of two independent oracle energy calls (brute force
 * over periodic images through o_md), deletion dU likewise, the acceptance
 * rule against its closed form, and the RNG draw order against the stream. */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "o_internal.h"

enum { MODE_INS = 0, MODE_DEL = 1 };

struct o_gcmc {
    const o_eam* ff;
    o_box bx;
    o_rng rng;
    int n, cap, gas, ngas;
    double cut, kbT, beta, lambda, sigma, z_fac, vol;
    double cut_s[3], lo_ph[3], hi_ph[3], cell_size[3];
    int ncell[3], bcell[3], ncells, max_trials;
    int rel[27][3];
    double* cut_sq; /* [ne*ne] r_c^2 */
    double *s, *x, *v, *rho, *F, *rho_x, *F_x;
    uint8_t* elem;
    int *tag, *cell, *next, *head;
    double* s_tr[3];
    double* tx;  /* [max_trials][3] Cartesian trial images */
    int* tcell;  /* [max_trials][3] cell coordinates */
    int ntrial;
};

static inline double hval(const double* c, double p) { return ((c[3] * p + c[2]) * p + c[1]) * p + c[0]; }

/* F(rho) with the linear extrapolation beyond rho_max (src/ff_eam.cpp:395-403) */
static double f_embed(const o_eam* e, int el, double rho, int extrapolate)
{
    size_t m;
    double p;
    o_spline_pm(rho, 1.0 / e->drho, e->nrho, &m, &p);
    const double* c = OF(e, el) + m * e->nf;
    double out = hval(c, p);
    double rho_max = (double)e->nrho * e->drho;
    if (extrapolate && rho > rho_max)
        out += ((3.0 * c[3] * p + 2.0 * c[2]) * p + c[1]) * (1.0 / e->drho) * (rho - rho_max);
    return out;
}

static int cell_no(const o_gcmc* g, const double* s)
{
    int c = 0;
    for (int i = 0; i < 3; i++) {
        int k = (int)((s[i] - 0.0) / g->cell_size[i]);
        if (k > g->ncell[i] - 1) k = g->ncell[i] - 1;
        c += g->bcell[i] * k;
    }
    return c;
}

static void cell_coord(const o_gcmc* g, const double* s, int* cc)
{
    for (int i = 0; i < 3; i++) {
        if (s[i] < 0.0)
            cc[i] = -(int)((0.0 - s[i]) / g->cell_size[i]) - 1;
        else if (1.0 <= s[i]) {
            int k = (int)((s[i] - 0.0) / g->cell_size[i]);
            cc[i] = k > g->ncell[i] - 1 ? k : g->ncell[i] - 1;
        } else {
            int k = (int)((s[i] - 0.0) / g->cell_size[i]);
            cc[i] = k < g->ncell[i] - 1 ? k : g->ncell[i] - 1;
        }
    }
}

static void gcmc_load(o_gcmc* g, int natms, const double* x, const uint8_t* elem, const double* v);

o_gcmc* o_gcmc_create(const o_eam* ff, int natms, const double* H, const double* x, const uint8_t* elem,
                      int gas_elem, double mu, double T, double kB, double hP, const double* masses, int seed,
                      int cap)
{
    if (cap < natms) cap = natms;
    o_gcmc* g = (o_gcmc*)calloc(1, sizeof(o_gcmc));
    g->ff = ff;
    g->n = natms;
    g->cap = cap;
    g->gas = gas_elem;
    o_box_set(&g->bx, H);
    o_rng_init(&g->rng, seed);
    int ne = ff->ne;
    g->cut_sq = (double*)malloc(sizeof(double) * ne * ne);
    for (int i = 0; i < ne * ne; i++) g->cut_sq[i] = ff->rc[i] * ff->rc[i];
    /* GCMC::init */
    g->cut = ff->rc[gas_elem * ne + 0];
    for (int i = 1; i < ne; i++)
        if (ff->rc[gas_elem * ne + i] > g->cut) g->cut = ff->rc[gas_elem * ne + i];
    double m_gas = masses[gas_elem];
    g->kbT = kB * T;
    g->beta = 1.0 / g->kbT;
    g->lambda = hP / sqrt(2.0 * M_PI * g->kbT * m_gas);
    g->sigma = sqrt(g->kbT / m_gas);
    g->z_fac = 1.0;
    for (int i = 0; i < 3; i++) g->z_fac /= g->lambda;
    g->z_fac *= exp(g->beta * mu);
    g->vol = 1.0;
    for (int i = 0; i < 3; i++) g->vol *= g->bx.H[i][i];
    /* GCMC::box_setup + SGCMC::box_setup */
    g->max_trials = 1;
    int ntot = 0;
    g->ncells = 1;
    for (int i = 0; i < 3; i++) {
        double t = 0.0;
        for (int j = i; j < 3; j++) t += g->bx.B[j][i] * g->bx.B[j][i];
        g->cut_s[i] = sqrt(t) * g->cut;
        g->lo_ph[i] = 0.0 - g->cut_s[i];
        g->hi_ph[i] = 1.0 + g->cut_s[i];
        int per = 1 + (int)floor(g->hi_ph[i]) + -(int)floor(g->lo_ph[i]);
        g->max_trials *= per;
        ntot += per;
        g->cell_size[i] = g->cut_s[i] / 1.0;
        g->ncell[i] = (int)((1.0 - 0.0) / g->cell_size[i]) + 1;
        g->bcell[i] = g->ncells;
        g->ncells *= g->ncell[i];
    }
    g->s_tr[0] = (double*)malloc(sizeof(double) * ntot);
    g->s_tr[1] = g->s_tr[0] + 1 + (int)floor(g->hi_ph[0]) - (int)floor(g->lo_ph[0]);
    g->s_tr[2] = g->s_tr[1] + 1 + (int)floor(g->hi_ph[1]) - (int)floor(g->lo_ph[1]);
    g->tx = (double*)malloc(sizeof(double) * 3 * g->max_trials);
    g->tcell = (int*)malloc(sizeof(int) * 3 * g->max_trials);
    /* stencil offsets, first coordinate fastest; for m = 1 every offset of
     * the 3x3x3 block passes the sum test of the constructor */
    int k = 0;
    for (int c2 = -1; c2 <= 1; c2++)
        for (int c1 = -1; c1 <= 1; c1++)
            for (int c0 = -1; c0 <= 1; c0++) {
                g->rel[k][0] = c0; g->rel[k][1] = c1; g->rel[k][2] = c2;
                k++;
            }
    size_t cd = (size_t)cap > 0 ? (size_t)cap : 1;
    g->s = (double*)malloc(sizeof(double) * 3 * cd);
    g->x = (double*)malloc(sizeof(double) * 3 * cd);
    g->v = (double*)calloc(3 * cd, sizeof(double));
    g->rho = (double*)calloc(cd, sizeof(double));
    g->F = (double*)calloc(cd, sizeof(double));
    g->rho_x = (double*)calloc(cd, sizeof(double));
    g->F_x = (double*)calloc(cd, sizeof(double));
    g->elem = (uint8_t*)malloc(cd);
    g->tag = (int*)malloc(sizeof(int) * cd);
    g->cell = (int*)malloc(sizeof(int) * cd);
    g->next = (int*)malloc(sizeof(int) * cd);
    g->head = (int*)malloc(sizeof(int) * g->ncells);
    gcmc_load(g, natms, x, elem, NULL);
    return g;
}

/* GCMC::xchng's set-up over the current atoms (src/sgcmc.cpp xchng,
 * src/pgcmc.cpp:204-282): x2s of the locals with the wrap, the densities and
 * F of init_xchng, the cell lists and the gas count; v travels with the atoms
 * (NULL: zero). The random stream is not touched. */
static void gcmc_load(o_gcmc* g, int natms, const double* x, const uint8_t* elem, const double* v)
{
    const o_eam* ff = g->ff;
    int ne = ff->ne;
    int gas_elem = g->gas;
    double H[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) H[3 * i + j] = g->bx.H[i][j];
    g->n = natms;
    memcpy(g->elem, elem, natms);
    if (v) memcpy(g->v, v, sizeof(double) * 3 * natms);
    else memset(g->v, 0, sizeof(double) * 3 * natms);

    /* init_xchng runs over the half list of a fresh rebuild: reuse the MD
     * oracle for the ghosts and the list, then accumulate rho as written
     * there (left-associated rsq, r_c^2 test, F without extrapolation). */
    o_md* md = o_md_create(ff, natms, H, x, elem, 0.5);
    const double *xa, *sa;
    const uint8_t* ea;
    const int64_t* row;
    const int32_t* col;
    o_md_view(md, &xa, &sa, &ea, &row, &col);
    memcpy(g->x, xa, sizeof(double) * 3 * natms);
    memcpy(g->s, sa, sizeof(double) * 3 * natms);
    double dr_inv = 1.0 / ff->dr;
    for (int i = 0; i < natms; i++) g->rho[i] = 0.0;
    for (int i = 0; i < natms; i++) {
        int ei = ea[i];
        const double* xi = xa + 3 * i;
        for (int64_t q = row[i]; q < row[i + 1]; q++) {
            int j = col[q];
            int ej = ea[j];
            const double* xj = xa + 3 * j;
            double d0 = xi[0] - xj[0], d1 = xi[1] - xj[1], d2 = xi[2] - xj[2];
            double rsq = d0 * d0 + d1 * d1 + d2 * d2;
            if (rsq < g->cut_sq[ei * ne + ej]) {
                double r = sqrt(rsq), p;
                size_t m;
                o_spline_pm(r, dr_inv, ff->nr, &m, &p);
                g->rho[i] += hval(ORHO(ff, ej, ei)[m], p);
                if (j < natms) g->rho[j] += hval(ORHO(ff, ei, ej)[m], p);
            }
        }
        g->F[i] = f_embed(ff, ei, g->rho[i], 0);
    }
    o_md_free(md);

    /* xchng: bin the locals, descending, so every cell list ascends */
    for (int c = 0; c < g->ncells; c++) g->head[c] = -1;
    for (int i = natms - 1; i > -1; i--) {
        g->cell[i] = cell_no(g, g->s + 3 * i);
        g->next[i] = g->head[g->cell[i]];
        g->head[g->cell[i]] = i;
    }
    g->ngas = 0;
    for (int i = 0; i < natms; i++)
        if (g->elem[i] == gas_elem) g->ngas++;
}

int o_gcmc_reload(o_gcmc* g, int natms, const double* x, const uint8_t* elem, const double* v)
{
    if (natms < 0 || natms > g->cap) return -1;
    gcmc_load(g, natms, x, elem, v);
    return 0;
}

void o_gcmc_free(o_gcmc* g)
{
    if (!g) return;
    free(g->cut_sq); free(g->s_tr[0]); free(g->tx); free(g->tcell);
    free(g->s); free(g->x); free(g->v); free(g->rho); free(g->F); free(g->rho_x); free(g->F_x);
    free(g->elem); free(g->tag); free(g->cell); free(g->next); free(g->head);
    free(g);
}

/* prep_s_x_buff: periodic images of the trial point inside the padded box,
 * per dimension s, s+1, ... then s-1, s-2, ...; first dimension fastest. */
static void prep_trials(o_gcmc* g, const double* sb)
{
    int per[3];
    g->ntrial = 1;
    for (int i = 0; i < 3; i++) {
        int no = 0;
        double s = sb[i];
        for (double t = s; t < g->hi_ph[i]; t++)
            if (g->lo_ph[i] <= t && t < g->hi_ph[i]) g->s_tr[i][no++] = t;
        for (double t = s - 1.0; g->lo_ph[i] <= t; t--)
            if (g->lo_ph[i] <= t && t < g->hi_ph[i]) g->s_tr[i][no++] = t;
        per[i] = no;
        g->ntrial *= no;
    }
    int cnt[3] = {0, 0, 0};
    for (int k = 0; k < g->ntrial; k++) {
        double b[3];
        for (int j = 0; j < 3; j++) b[j] = g->s_tr[j][cnt[j]];
        cell_coord(g, b, g->tcell + 3 * k);
        o_s2x(&g->bx, b, g->tx + 3 * k);
        cnt[0]++;
        for (int j = 0; j < 2; j++)
            if (cnt[j] == per[j]) { cnt[j] = 0; cnt[j + 1]++; }
    }
}

/* pre_xchng_energy + xchng_energy: returns delta_u, leaves rho_x/F_x/tag
 * set for post_xchng_energy and the trial atom's density in *rho_iatm. */
static double exchange_energy(o_gcmc* g, int mode, int del_idx, double* rho_iatm_out)
{
    const o_eam* e = g->ff;
    int ne = e->ne, gas = g->gas, n = g->n;
    double dr_inv = 1.0 / e->dr;
    double c0 = mode == MODE_DEL ? -1.0 : 1.0;
    for (int i = 0; i < n; i++) g->tag[i] = -1;  /* reset_tag */
    for (int i = 0; i < n; i++) g->rho_x[i] = g->rho[i];
    double en = 0.0, rho_iatm = 0.0;
    for (int it = 0; it < g->ntrial; it++) {
        int iatm = (it == 0 && mode == MODE_DEL) ? del_idx : n + it;
        const double* ix = g->tx + 3 * it;
        const int* ic = g->tcell + 3 * it;
        for (int q = 0; q < 27; q++) {
            int jc[3], lcl = 1, jcell = 0;
            for (int d = 0; d < 3 && lcl; d++) {
                jc[d] = ic[d] + g->rel[q][d];
                jcell += g->bcell[d] * jc[d];
                if (jc[d] < 0 || jc[d] > g->ncell[d] - 1) lcl = 0;
            }
            if (!lcl) continue;
            for (int j = g->head[jcell]; j != -1; j = g->next[j]) {
                if (j == iatm) continue;
                int ej = g->elem[j];
                double rsq = o_rsq(ix, g->x + 3 * j);
                if (rsq >= g->cut_sq[gas * ne + ej]) continue;
                g->tag[j] = 0;
                double r = sqrt(rsq), p;
                size_t m;
                o_spline_pm(r, dr_inv, e->nr, &m, &p);
                en += hval(OPHI(e, gas, ej)[m], p) / r;  /* j is always local here */
                g->rho_x[j] += c0 * hval(ORHO(e, gas, ej)[m], p);
                rho_iatm += hval(ORHO(e, ej, gas)[m], p);
            }
        }
        if (it == 0) {  /* next_jatm_self: the trial point's own images */
            for (int k = 1; k < g->ntrial; k++) {
                double rsq = o_rsq(ix, g->tx + 3 * k);
                if (rsq >= g->cut_sq[gas * ne + gas]) continue;
                double r = sqrt(rsq), p;
                size_t m;
                o_spline_pm(r, dr_inv, e->nr, &m, &p);
                en += 0.5 * hval(OPHI(e, gas, gas)[m], p) / r;
                rho_iatm += hval(ORHO(e, gas, gas)[m], p);
            }
        }
    }
    double en0 = 0.0;
    for (int i = 0; i < n; i++)
        if (g->tag[i] == 0) {
            g->F_x[i] = f_embed(e, g->elem[i], g->rho_x[i], 1);
            en0 += g->F_x[i] - g->F[i];
        }
    en += en0 * c0;
    en += f_embed(e, gas, rho_iatm, 1);
    *rho_iatm_out = rho_iatm;
    return en;
}

static void post_energy(o_gcmc* g, int inserted, double rho_iatm)
{
    for (int i = 0; i < g->n; i++)
        if (g->tag[i] == 0) { g->rho[i] = g->rho_x[i]; g->F[i] = g->F_x[i]; }
    if (inserted) {
        int k = g->n - 1;
        g->rho[k] = rho_iatm;
        g->F[k] = f_embed(g->ff, g->gas, rho_iatm, 1);
    }
}

static void move_atom(o_gcmc* g, int to, int from)
{
    memcpy(g->s + 3 * to, g->s + 3 * from, sizeof(double) * 3);
    memcpy(g->x + 3 * to, g->x + 3 * from, sizeof(double) * 3);
    memcpy(g->v + 3 * to, g->v + 3 * from, sizeof(double) * 3);
    g->elem[to] = g->elem[from];
    g->rho[to] = g->rho[from]; g->F[to] = g->F[from];
    g->rho_x[to] = g->rho_x[from]; g->F_x[to] = g->F_x[from];
    g->tag[to] = g->tag[from];
    g->cell[to] = g->cell[from];
    g->next[to] = g->next[from];
}

/* ins_succ: append at the end of the arrays and at the tail of its cell */
static void insert_atom(o_gcmc* g, const double* sb, const double* vel)
{
    int k = g->n;
    memcpy(g->x + 3 * k, g->tx, sizeof(double) * 3);
    memcpy(g->v + 3 * k, vel, sizeof(double) * 3);
    memcpy(g->s + 3 * k, sb, sizeof(double) * 3);
    g->elem[k] = (uint8_t)g->gas;
    g->tag[k] = -1;
    int c = 0;
    for (int i = 0; i < 3; i++) c += g->bcell[i] * g->tcell[i];
    g->cell[k] = c;
    int* link = g->head + c;
    while (*link != -1) link = g->next + *link;
    *link = k;
    g->next[k] = -1;
    g->n++;
    g->ngas++;
}

/* del_succ: unlink, move the last atom into the hole and relink it so every
 * cell list stays ascending */
static void delete_atom(o_gcmc* g, int del)
{
    int last = g->n - 1;
    int* link = g->head + g->cell[del];
    while (*link != del) link = g->next + *link;
    *link = g->next[del];
    if (del != last) {
        int* first_after = g->head + g->cell[last];
        while (*first_after < del) first_after = g->next + *first_after;
        int* to_last = first_after;
        while (*to_last != last) to_last = g->next + *to_last;
        *to_last = -1;
        g->next[last] = *first_after;
        *first_after = del;
        move_atom(g, del, last);
    }
    g->n--;
    g->ngas--;
}

int o_gcmc_attempts(o_gcmc* g, int nattempts, int* decisions, double* dU)
{
    for (int a = 0; a < nattempts; a++) {
        double sb[3];
        int mode, del_idx = -1;
        decisions[a] = 0;
        dU[a] = 0.0;
        if (o_rng_uniform(&g->rng) < 0.5) {
            mode = MODE_INS;
            for (int i = 0; i < 3; i++) sb[i] = o_rng_uniform(&g->rng);
        } else {
            if (g->ngas == 0) continue;
            mode = MODE_DEL;
            int igas = (int)(g->ngas * o_rng_uniform(&g->rng));
            int cnt = -1;
            for (del_idx = 0; cnt != igas; del_idx++)
                if (g->elem[del_idx] == g->gas) cnt++;
            del_idx--;
            memcpy(sb, g->s + 3 * del_idx, sizeof(sb));
        }
        prep_trials(g, sb);
        double rho_iatm;
        double du = exchange_energy(g, mode, del_idx, &rho_iatm);
        dU[a] = du;
        if (mode == MODE_INS) {
            double fac = g->z_fac * g->vol / (((double)g->ngas + 1.0) * exp(g->beta * du));
            if (o_rng_uniform(&g->rng) < fac) {
                if (g->n == g->cap) return -1;
                double vel[3];
                for (int i = 0; i < 3; i++) vel[i] = o_rng_gaussian(&g->rng) * g->sigma;
                insert_atom(g, sb, vel);
                post_energy(g, 1, rho_iatm);
                decisions[a] = 1;
            } else
                decisions[a] = -1;
        } else {
            double fac = (double)g->ngas * exp(g->beta * du) / (g->z_fac * g->vol);
            if (o_rng_uniform(&g->rng) < fac) {
                delete_atom(g, del_idx);
                post_energy(g, 0, rho_iatm);
                decisions[a] = 2;
            } else
                decisions[a] = -2;
        }
    }
    return 0;
}

int o_gcmc_natms(const o_gcmc* g) { return g->n; }
int o_gcmc_ngas(const o_gcmc* g) { return g->ngas; }
int o_gcmc_rng_seed(const o_gcmc* g) { return g->rng.seed; }

void o_gcmc_info(const o_gcmc* g, double* out8, int* iout4)
{
    out8[0] = g->z_fac; out8[1] = g->vol; out8[2] = g->beta; out8[3] = g->lambda;
    out8[4] = g->sigma; out8[5] = g->cut; out8[6] = g->cut_s[0]; out8[7] = g->cell_size[0];
    iout4[0] = g->max_trials; iout4[1] = g->ncell[0]; iout4[2] = g->ncell[1]; iout4[3] = g->ncell[2];
}

void o_gcmc_state(const o_gcmc* g, double* x, double* s, uint8_t* elem, double* rho, double* F, double* v)
{
    int n = g->n;
    if (x) memcpy(x, g->x, sizeof(double) * 3 * n);
    if (s) memcpy(s, g->s, sizeof(double) * 3 * n);
    if (elem) memcpy(elem, g->elem, n);
    if (rho) memcpy(rho, g->rho, sizeof(double) * n);
    if (F) memcpy(F, g->F, sizeof(double) * n);
    if (v) memcpy(v, g->v, sizeof(double) * 3 * n);
}
