/*
 * o_dmd.c -- oracle: DMD (Gaussian-smeared EAM) neighbor list, free energy,
 * forces on x and alpha, and the per-component chemical-potential sums.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle).
 */
#include "o_internal.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define O_PI_IN_SQ 0.564189583547756286948079451561 /* 1/sqrt(pi) */

struct o_dmd {
    const o_eam* ff;
    o_box bx;
    o_ghosts g;
    int n, c_dim, N;
    double skin, max_cut, scl, kBT;
    double xi[32], wi[32];
    double* cut;     /* [ne*ne] r_c */
    double* cut_sk;  /* [ne*ne] r_c + skin */
    double* mu0;     /* [ne] */
    uint8_t* elem_lcl;
    double *alpha_lcl, *c_lcl;   /* [n*c_dim] */
    uint8_t* elem;   /* [nall*c_dim] */
    double *alpha, *c;           /* [nall*c_dim] */
    int64_t* row;    /* [n*c_dim + 1] */
    int32_t* col;    /* component indices j*c_dim + jc */
    int64_t npairs, cap;
    int32_t** tl;    /* per-component scratch lists while visiting site i */
    int* tl_n;
    int tl_cap;
};

static inline double hval(const double* c, double p) { return ((c[3] * p + c[2]) * p + c[1]) * p + c[0]; }

/* ghost copies of the per-component arrays follow the owner */
static void dmd_sync_ghost_state(o_dmd* d)
{
    int cd = d->c_dim;
    d->elem = (uint8_t*)realloc(d->elem, (size_t)d->g.nall * cd + 1);
    d->alpha = (double*)realloc(d->alpha, sizeof(double) * ((size_t)d->g.nall * cd + 1));
    d->c = (double*)realloc(d->c, sizeof(double) * ((size_t)d->g.nall * cd + 1));
    for (int k = 0; k < d->g.nall; k++) {
        int o = d->g.owner[k];
        for (int ic = 0; ic < cd; ic++) {
            d->elem[k * cd + ic] = d->elem_lcl[o * cd + ic];
            d->alpha[k * cd + ic] = d->alpha_lcl[o * cd + ic];
            d->c[k * cd + ic] = d->c_lcl[o * cd + ic];
        }
    }
}

typedef struct { o_dmd* d; } dctx;

/* NeighborDMD::create_list mid() (src/neighbor_dmd.cpp:150-167) */
static void dmd_visit(void* v, int i, int j)
{
    o_dmd* d = ((dctx*)v)->d;
    if (i >= j) return;
    int cd = d->c_dim, ne = d->ff->ne;
    double r = sqrt(o_rsq(d->g.x + 3 * i, d->g.x + 3 * j));
    const double* ci = d->c + (size_t)i * cd;
    const double* cj = d->c + (size_t)j * cd;
    const double* ai = d->alpha + (size_t)i * cd;
    const double* aj = d->alpha + (size_t)j * cd;
    const uint8_t* ei = d->elem + (size_t)i * cd;
    const uint8_t* ej = d->elem + (size_t)j * cd;
    for (int ic = 0; ic < cd && ci[ic] >= 0.0; ic++)
        for (int jc = 0; jc < cd && cj[jc] >= 0.0; jc++)
            if (r - d->scl * sqrt(ai[ic] * ai[ic] + aj[jc] * aj[jc]) < d->cut_sk[ei[ic] * ne + ej[jc]]) {
                if (d->tl_n[ic] == d->tl_cap) {
                    d->tl_cap = d->tl_cap * 2 + 256;
                    for (int q = 0; q < cd; q++) d->tl[q] = (int32_t*)realloc(d->tl[q], sizeof(int32_t) * d->tl_cap);
                }
                d->tl[ic][d->tl_n[ic]++] = j * cd + jc;
            }
}

static void dmd_post(void* v, int i)
{
    o_dmd* d = ((dctx*)v)->d;
    int cd = d->c_dim;
    for (int ic = 0; ic < cd; ic++) {
        int64_t need = d->npairs + d->tl_n[ic];
        if (need > d->cap) {
            d->cap = need * 2 + 1024;
            d->col = (int32_t*)realloc(d->col, sizeof(int32_t) * d->cap);
        }
        memcpy(d->col + d->npairs, d->tl[ic], sizeof(int32_t) * d->tl_n[ic]);
        d->npairs += d->tl_n[ic];
        d->row[(size_t)i * cd + ic + 1] = d->npairs;
        d->tl_n[ic] = 0;
    }
}

static void dmd_rebuild(o_dmd* d, const double* x_lcl)
{
    double* s = (double*)malloc(sizeof(double) * 3 * (d->n > 0 ? d->n : 1));
    for (int i = 0; i < d->n; i++) o_x2s(&d->bx, x_lcl + 3 * i, s + 3 * i);
    o_ghosts_build(&d->g, &d->bx, d->n, s, d->max_cut);
    free(s);
    dmd_sync_ghost_state(d);
    d->row[0] = 0;
    d->npairs = 0;
    dctx c = {d};
    o_cell_do(&d->g, &d->bx, d->max_cut, NULL, dmd_visit, dmd_post, &c);
}

o_dmd* o_dmd_create(const o_eam* ff, int natms, int c_dim, int nquad,
                    const double* H, const double* x, const uint8_t* elem,
                    const double* alpha, const double* c, double skin,
                    double kB, double T, double hP, const double* masses)
{
    o_dmd* d = (o_dmd*)calloc(1, sizeof(o_dmd));
    int ne = ff->ne;
    d->ff = ff;
    d->n = natms;
    d->c_dim = c_dim;
    d->N = nquad;
    d->skin = skin;
    o_quadrature_hg(nquad, d->xi, d->wi);
    d->scl = d->xi[nquad - 1];
    o_box_set(&d->bx, H);
    d->kBT = kB * T;
    d->cut = (double*)malloc(sizeof(double) * ne * ne);
    d->cut_sk = (double*)malloc(sizeof(double) * ne * ne);
    d->mu0 = (double*)malloc(sizeof(double) * ne);
    double ff_max_cut = 0.0;
    /* ForceFieldDMD::pre_init (src/ff_dmd.cpp:157-204) */
    for (int i = 0; i < ne; i++) {
        double lambda = hP / sqrt(2.0 * M_PI * d->kBT * masses[i]);
        d->mu0[i] = d->kBT * 1.5 * (2.0 * log(lambda) - 1.0 - log(M_PI));
        for (int j = 0; j <= i; j++) {
            double t = ff->rc[i * ne + j];
            if (t > ff_max_cut) ff_max_cut = t;
            d->cut[i * ne + j] = d->cut[j * ne + i] = t;
            d->cut_sk[i * ne + j] = d->cut_sk[j * ne + i] = t + skin;
        }
    }
    double max_alpha = 0.0;
    for (int k = 0; k < natms * c_dim; k++)
        if (c[k] != -1.0 && alpha[k] > max_alpha) max_alpha = alpha[k];
    /* atoms->max_cut (src/dynamic_dmd.cpp:46) */
    d->max_cut = ff_max_cut + skin + d->scl * sqrt(2.0) * max_alpha;
    d->elem_lcl = (uint8_t*)malloc((size_t)natms * c_dim + 1);
    memcpy(d->elem_lcl, elem, (size_t)natms * c_dim);
    d->alpha_lcl = (double*)malloc(sizeof(double) * ((size_t)natms * c_dim + 1));
    d->c_lcl = (double*)malloc(sizeof(double) * ((size_t)natms * c_dim + 1));
    memcpy(d->alpha_lcl, alpha, sizeof(double) * natms * c_dim);
    memcpy(d->c_lcl, c, sizeof(double) * natms * c_dim);
    d->row = (int64_t*)malloc(sizeof(int64_t) * ((size_t)natms * c_dim + 1));
    d->tl = (int32_t**)calloc(c_dim, sizeof(int32_t*));
    d->tl_n = (int*)calloc(c_dim, sizeof(int));
    dmd_rebuild(d, x);
    return d;
}

void o_dmd_free(o_dmd* d)
{
    if (!d) return;
    o_ghosts_free(&d->g);
    for (int q = 0; q < d->c_dim; q++) free(d->tl[q]);
    free(d->tl); free(d->tl_n);
    free(d->cut); free(d->cut_sk); free(d->mu0); free(d->elem_lcl); free(d->alpha_lcl); free(d->c_lcl);
    free(d->elem); free(d->alpha); free(d->c); free(d->row); free(d->col);
    free(d);
}

int64_t o_dmd_npairs(const o_dmd* d) { return d->npairs; }

void o_dmd_pairs(const o_dmd* d, int32_t* pi, int32_t* pj, int8_t* img)
{
    int cd = d->c_dim;
    for (int i = 0; i < d->n * cd; i++)
        for (int64_t k = d->row[i]; k < d->row[i + 1]; k++) {
            int j = d->col[k];
            int js = j / cd;
            pi[k] = i;
            pj[k] = d->g.owner[js] * cd + j % cd;
            memcpy(img + 3 * k, d->g.img + 3 * js, 3);
        }
}

/* new state between list builds: ghosts follow their owners */
/* new positions, alpha and c; the ghosts and the list are rebuilt, so any
 * displacement (also across the box) is allowed */
void o_dmd_set_state(o_dmd* d, const double* x, const double* alpha, const double* c)
{
    memcpy(d->alpha_lcl, alpha, sizeof(double) * d->n * d->c_dim);
    memcpy(d->c_lcl, c, sizeof(double) * d->n * d->c_dim);
    dmd_rebuild(d, x);
}

/* Smeared phi, rho_ij, rho_ji and their r/alpha derivatives by N-point
 * Gauss-Hermite over the odd extension of r*f(r) (src/ff_eam_dmd.cpp:195-253) */
static void smear(const o_dmd* d, int ei, int ej, double r, double a, double* rp, double* dr_, double* da_)
{
    const o_eam* e = d->ff;
    double dr_inv = 1.0 / e->dr;
    double H[3][3] = {{0}};
    for (int l = 0; l < d->N; l++) {
        double xl = d->xi[l], xl2 = xl * xl;
        double u = r - xl * a;
        size_t m;
        double p;
        o_spline_pm(fabs(u), dr_inv, e->nr, &m, &p);
        const double* tabs[3] = {OPHI(e, ei, ej)[m], ORHO(e, ei, ej)[m], ORHO(e, ej, ei)[m]};
        for (int k = 0; k < 3; k++) {
            double t = hval(tabs[k], p);
            if (u < 0.0) t *= -1.0;
            t *= d->wi[l];
            H[k][0] += t;
            H[k][1] += t * xl;
            H[k][2] += t * xl2;
        }
    }
    double r_inv = 1.0 / r, r2_inv = r_inv * r_inv;
    double a_inv = 1.0 / a, a2_inv = a_inv * a_inv;
    for (int k = 0; k < 3; k++) {
        rp[k] = O_PI_IN_SQ * r_inv * H[k][0];
        dr_[k] = -r2_inv * (rp[k] + 2.0 * O_PI_IN_SQ * a_inv * H[k][1]);
        da_[k] = -a2_inv * (rp[k] - 2.0 * O_PI_IN_SQ * r_inv * H[k][2]);
    }
}

void o_dmd_pair_eval(const o_dmd* d, int ei, int ej, double r, double alpha_ij, double* out9)
{
    smear(d, ei, ej, r, alpha_ij, out9, out9 + 3, out9 + 6);
}

static double ent(double x) { return x == 0.0 ? 0.0 : x * log(x); }

static void femb(const o_dmd* d, int el, double rho, double* F, double* dF)
{
    const o_eam* e = d->ff;
    double drho_inv = 1.0 / e->drho, rho_max = (double)e->nrho * e->drho;
    size_t m;
    double p;
    o_spline_pm(rho, drho_inv, e->nrho, &m, &p);
    const double* c = OF(e, el) + m * 5;
    double t0 = (((c[4] * p + c[3]) * p + c[2]) * p + c[1]) * p + c[0];
    double t1 = (((4.0 * c[4] * p + 3.0 * c[3]) * p + 2.0 * c[2]) * p + c[1]) * drho_inv;
    if (rho > rho_max) t0 += t1 * (rho - rho_max);
    *F = t0;
    *dF = t1;
}

/* ForceFieldEAMDMD::__force_calc (src/ff_eam_dmd.cpp:126-360) followed by
 * the reductions of ForceFieldDMD::derivative (src/ff_dmd.cpp:220-237) */
void o_dmd_force(o_dmd* d, double* fx, double* falpha, double* mu, double* out8)
{
    int cd = d->c_dim, ne = d->ff->ne;
    int n = d->n * cd;
    int nall = d->g.nall * cd;
    const double* x = d->g.x;
    const double* c = d->c;
    const double* al = d->alpha;
    const uint8_t* el = d->elem;
    double* rp = (double*)calloc((size_t)d->npairs * 3 + 1, sizeof(double));
    double* drp = (double*)calloc((size_t)d->npairs * 3 + 1, sizeof(double));
    double* dap = (double*)calloc((size_t)d->npairs * 3 + 1, sizeof(double));
    double* rho = (double*)calloc((size_t)n + 1, sizeof(double));
    double* dE = (double*)calloc((size_t)nall + 1, sizeof(double));
    double v[9] = {0};
    for (int i = 0; i < 3 * d->n; i++) fx[i] = 0.0;
    for (int i = 0; i < n; i++) { falpha[i] = 0.0; mu[i] = 0.0; }
    for (int i = 0; i < n; i++) {
        double ci = c[i];
        if (ci < 0.0) continue;
        int ei = el[i];
        for (int64_t k = d->row[i]; k < d->row[i + 1]; k++) {
            int j = d->col[k];
            int ej = el[j];
            double r = sqrt(o_rsq(x + 3 * (i / cd), x + 3 * (j / cd)));
            double a = sqrt(al[i] * al[i] + al[j] * al[j]);
            if (r - a * d->xi[d->N - 1] >= d->cut[ei * ne + ej]) continue;
            smear(d, ei, ej, r, a, rp + 3 * k, drp + 3 * k, dap + 3 * k);
            rho[i] += c[j] * rp[3 * k + 2];
            if (j < n) {
                rho[j] += ci * rp[3 * k + 1];
                v[0] += ci * c[j] * rp[3 * k];
            } else
                v[0] += 0.5 * ci * c[j] * rp[3 * k];
        }
        double F, dF;
        femb(d, ei, rho[i], &F, &dF);
        rho[i] = F;
        dE[i] = dF;
        mu[i] = F;
        v[0] += ci * (F + d->mu0[ei] - 3.0 * d->kBT * log(al[i]));
        v[7] += ent(ci);
    }
    for (int i = 0; i < d->n; i++) {
        double cv = 1.0;
        for (int ic = 0; ic < cd; ic++)
            if (c[i * cd + ic] > 0.0) cv -= c[i * cd + ic];
        v[7] += ent(cv);
    }
    v[0] += d->kBT * v[7];
    /* ghost update of dE (src/ff_eam_dmd.cpp:302) */
    for (int k = d->n; k < d->g.nall; k++)
        for (int ic = 0; ic < cd; ic++) dE[k * cd + ic] = dE[d->g.src[k] * cd + ic];
    double f_i[3] = {0, 0, 0};
    for (int i = 0; i < n; i++) {
        if (i % cd == 0) f_i[0] = f_i[1] = f_i[2] = 0.0;
        const double* xi = x + 3 * (i / cd);
        double ci = c[i], ai = al[i], dEi = dE[i];
        double fa = 0.0, mui = 0.0;
        for (int64_t k = d->row[i]; k < d->row[i + 1]; k++) {
            int j = d->col[k];
            const double* xj = x + 3 * (j / cd);
            double dx[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]};
            double fp = -(drp[3 * k + 2] * dEi + drp[3 * k + 1] * dE[j] + drp[3 * k]) * ci * c[j];
            double ap = -(dap[3 * k + 2] * dEi + dap[3 * k + 1] * dE[j] + dap[3 * k]) * ci * c[j];
            mui += c[j] * (rp[3 * k] + rp[3 * k + 1] * dE[j]);
            if (j < n) mu[j] += ci * (rp[3 * k] + rp[3 * k + 2] * dEi);
            for (int q = 0; q < 3; q++) f_i[q] += fp * dx[q];
            fa += ai * ap;
            if (j < n) {
                for (int q = 0; q < 3; q++) fx[3 * (j / cd) + q] += -fp * dx[q];
                falpha[j] += al[j] * ap;
            } else
                fp *= 0.5;
            /* DyadicV(-fp, dx, W): W += dx_a dx_b (-fp) */
            v[1] += dx[0] * dx[0] * -fp; v[2] += dx[0] * dx[1] * -fp; v[3] += dx[0] * dx[2] * -fp;
            v[4] += dx[1] * dx[1] * -fp; v[5] += dx[1] * dx[2] * -fp; v[6] += dx[2] * dx[2] * -fp;
        }
        fa += 3.0 * d->kBT * ci / ai;
        falpha[i] += fa;
        mu[i] += mui;
        if ((i + 1) % cd == 0)
            for (int q = 0; q < 3; q++) fx[3 * (i / cd) + q] += f_i[q];
    }
    out8[0] = v[0];
    for (int q = 1; q < 7; q++) out8[q] = -v[q];
    out8[7] = v[7];
    free(rp); free(drp); free(dap); free(rho); free(dE);
}

/* ForceFieldEAMDMD::__energy_calc (src/ff_eam_dmd.cpp:835-953) */
double o_dmd_energy(o_dmd* d)
{
    int cd = d->c_dim, ne = d->ff->ne;
    int n = d->n * cd;
    const o_eam* e = d->ff;
    const double* x = d->g.x;
    const double* c = d->c;
    const double* al = d->alpha;
    const uint8_t* el = d->elem;
    double dr_inv = 1.0 / e->dr;
    double* rho = (double*)calloc((size_t)n + 1, sizeof(double));
    double en = 0.0, s = 0.0;
    for (int i = 0; i < n; i++) {
        int ei = el[i];
        for (int64_t k = d->row[i]; k < d->row[i + 1]; k++) {
            int j = d->col[k];
            int ej = el[j];
            double r = sqrt(o_rsq(x + 3 * (i / cd), x + 3 * (j / cd)));
            double a = sqrt(al[i] * al[i] + al[j] * al[j]);
            if (r - a * d->xi[d->N - 1] >= d->cut[ei * ne + ej]) continue;
            double arr[3] = {0, 0, 0};
            for (int l = 0; l < d->N; l++) {
                double u = r - d->xi[l] * a;
                size_t m;
                double p;
                o_spline_pm(fabs(u), dr_inv, e->nr, &m, &p);
                const double* tabs[3] = {OPHI(e, ei, ej)[m], ORHO(e, ei, ej)[m], ORHO(e, ej, ei)[m]};
                for (int q = 0; q < 3; q++) {
                    double t = hval(tabs[q], p);
                    if (u < 0.0) t *= -1.0;
                    arr[q] += d->wi[l] * t;
                }
            }
            double t0 = O_PI_IN_SQ * (1.0 / r);
            for (int q = 0; q < 3; q++) arr[q] *= t0;
            rho[i] += c[j] * arr[2];
            if (j < n) {
                rho[j] += c[i] * arr[1];
                en += c[i] * c[j] * arr[0];
            } else
                en += 0.5 * c[i] * c[j] * arr[0];
        }
        if (c[i] < 0.0) continue;
        double F, dF;
        femb(d, ei, rho[i], &F, &dF);
        if (c[i] != 0.0) en += c[i] * (F + d->mu0[ei] - 3.0 * d->kBT * log(al[i]));
        s += ent(c[i]);
    }
    for (int i = 0; i < d->n; i++) {
        double cv = 1.0;
        for (int ic = 0; ic < cd; ic++)
            if (c[i * cd + ic] > 0.0) cv -= c[i * cd + ic];
        s += ent(cv);
    }
    en += d->kBT * s;
    free(rho);
    return en;
}

/* ForceFieldEAMDMD::calc_Q with theta (src/ff_eam_dmd.cpp:1611-1633): the
 * barrier of the exchange path between two sites, z0 its position, alpha_Q_sq
 * the inverse width, theta the derivative term the Jacobian uses. */
void o_dmd_calc_q(double gamma_i, double gamma_j, double mu_ji, double* Q, double* alpha_Q_sq, double* theta)
{
    double a = 5.0 * (gamma_i - gamma_j - 6.0 * mu_ji);
    double xi = -2.0 * (gamma_i + gamma_j);
    double delta_sqrt = sqrt((xi - a) * (xi - a) - 8.0 * a * gamma_i);
    double z0 = -4.0 * gamma_i / ((xi - a) - delta_sqrt);
    double dz0 = 30.0 * z0 * (1.0 - z0) / delta_sqrt;
    double dQ = z0 * z0 * z0 * (6.0 * z0 * z0 - 15.0 * z0 + 10.0);
    double q = z0 * z0 * (1.0 - z0) * (1.0 - z0) * ((1.0 - z0) * gamma_i + z0 * gamma_j);
    q += dQ * mu_ji;
    *Q = q;
    *alpha_Q_sq = 1.0 / ((1.0 - z0) * gamma_i + z0 * gamma_j);
    *theta = *alpha_Q_sq * (gamma_i - gamma_j) * dz0 - dQ;
}

/* The state ForceFieldEAMDMD::calc_mu leaves (src/ff_eam_dmd.cpp:1082-1282)
 * for c_d_calc and __J: rho_phi of every half-list pair (the smeared r*phi,
 * rho at j, rho at i; zero for pairs beyond the cut), dE and ddE = F''(rho),
 * the raw chemical potentials of local and ghost components and the vacancy
 * concentration cv of every site. */
typedef struct {
    double* rp;   /* [npairs*3] */
    double* dE;   /* [nall*c_dim] */
    double* ddE;  /* [nall*c_dim] */
    double* mu;   /* [nall*c_dim] */
    double* cv;   /* [nall sites] */
    double* rho;  /* [n*c_dim] local densities */
} dmd_static;

static void dmd_static_free(dmd_static* s)
{
    free(s->rp); free(s->dE); free(s->ddE); free(s->mu); free(s->cv); free(s->rho);
}

static void dmd_static_build(o_dmd* d, dmd_static* s)
{
    int cd = d->c_dim, ne = d->ff->ne;
    int n = d->n * cd, nall = d->g.nall * cd;
    const o_eam* e = d->ff;
    const double* x = d->g.x;
    const double* c = d->c;
    const double* al = d->alpha;
    const uint8_t* el = d->elem;
    s->rp = (double*)calloc((size_t)d->npairs * 3 + 1, sizeof(double));
    s->dE = (double*)calloc((size_t)nall + 1, sizeof(double));
    s->ddE = (double*)calloc((size_t)nall + 1, sizeof(double));
    s->mu = (double*)calloc((size_t)nall + 1, sizeof(double));
    s->cv = (double*)calloc((size_t)d->g.nall + 1, sizeof(double));
    double* rho = (double*)calloc((size_t)n + 1, sizeof(double));
    double dr9[9];
    for (int i = 0; i < n; i++) {
        if (c[i] < 0.0) continue;
        int ei = el[i];
        for (int64_t k = d->row[i]; k < d->row[i + 1]; k++) {
            int j = d->col[k];
            int ej = el[j];
            double r = sqrt(o_rsq(x + 3 * (i / cd), x + 3 * (j / cd)));
            double a = sqrt(al[i] * al[i] + al[j] * al[j]);
            if (r - a * d->xi[d->N - 1] >= d->cut[ei * ne + ej]) continue;
            smear(d, ei, ej, r, a, s->rp + 3 * k, dr9, dr9 + 3);
        }
    }
    /* rho, E, dE, ddE (src/ff_eam_dmd.cpp:1145-1180) */
    double drho_inv = 1.0 / e->drho, rho_max = (double)e->nrho * e->drho;
    for (int i = 0; i < n; i++) {
        if (c[i] < 0.0) continue;
        for (int64_t k = d->row[i]; k < d->row[i + 1]; k++) {
            int j = d->col[k];
            rho[i] += c[j] * s->rp[3 * k + 2];
            if (j < n) rho[j] += c[i] * s->rp[3 * k + 1];
        }
        size_t m;
        double p;
        o_spline_pm(rho[i], drho_inv, e->nrho, &m, &p);
        const double* cf = OF(e, el[i]) + m * 5;
        double E = (((cf[4] * p + cf[3]) * p + cf[2]) * p + cf[1]) * p + cf[0];
        double dE = (((4.0 * cf[4] * p + 3.0 * cf[3]) * p + 2.0 * cf[2]) * p + cf[1]) * drho_inv;
        s->ddE[i] = ((12.0 * cf[4] * p + 6.0 * cf[3]) * p + 2.0 * cf[2]) * drho_inv * drho_inv;
        if (rho[i] > rho_max) E += dE * (rho[i] - rho_max);
        s->mu[i] = E;
        s->dE[i] = dE;
    }
    /* ghost copies (update(dE_ptr), src/ff_eam_dmd.cpp:1182) */
    for (int k = d->n; k < d->g.nall; k++)
        for (int ic = 0; ic < cd; ic++) {
            s->dE[k * cd + ic] = s->dE[d->g.src[k] * cd + ic];
            s->ddE[k * cd + ic] = s->ddE[d->g.src[k] * cd + ic];
        }
    /* mu_i += c_j (rho_phi_0 + rho_phi_1 dE_j), mu_j += c_i (rho_phi_0 + rho_phi_2 dE_i) */
    for (int i = 0; i < n; i++)
        for (int64_t k = d->row[i]; k < d->row[i + 1]; k++) {
            int j = d->col[k];
            s->mu[i] += c[j] * (s->rp[3 * k] + s->rp[3 * k + 1] * s->dE[j]);
            if (j < n) s->mu[j] += c[i] * (s->rp[3 * k] + s->rp[3 * k + 2] * s->dE[i]);
        }
    /* update(f_c) */
    for (int k = d->n; k < d->g.nall; k++)
        for (int ic = 0; ic < cd; ic++) s->mu[k * cd + ic] = s->mu[d->g.src[k] * cd + ic];
    for (int k = 0; k < d->g.nall; k++) {
        double v = 1.0;
        for (int ic = 0; ic < cd; ic++)
            if (c[k * cd + ic] > 0.0) v -= c[k * cd + ic];
        s->cv[k] = v;
    }
    s->rho = rho;
}

/* The exchange terms of one second-list pair (src/ff_eam_dmd.cpp:1397-1462):
 * dc_ij and M_IJ[0..3]; returns 0 when (i, j) is not in the second list. */
static int dmd_exchange(const o_dmd* d, const dmd_static* s, int i, int j, const double* rsq_crd,
                        const double* zeta, const double* c1, double beta, double* dc, double* M)
{
    int cd = d->c_dim;
    const double* x = d->g.x;
    const double* c = d->c;
    const double* al = d->alpha;
    int ei = d->elem[i];
    if (d->elem[j] != ei) return 0;
    double rsq = o_rsq(x + 3 * (i / cd), x + 3 * (j / cd));
    if (rsq >= rsq_crd[ei]) return 0;
    double alpha_i = al[i], alpha_sq_i = alpha_i * alpha_i;
    double alpha_j = al[j], alpha_sq_j = alpha_j * alpha_j;
    double gamma_i = rsq / (zeta[ei] * alpha_sq_i);
    double gamma_j = rsq / (zeta[ei] * alpha_sq_j);
    double mu_ji = beta * (s->mu[j] - s->mu[i]);
    double Q, gamma_ij_inv, theta_i;
    o_dmd_calc_q(gamma_i, gamma_j, mu_ji, &Q, &gamma_ij_inv, &theta_i);
    double coef0 = -c1[ei] * rsq * gamma_ij_inv * exp(-Q + 0.5 * mu_ji);
    double theta_j = theta_i + 1.0;
    double cvi = s->cv[i / cd], cvj = s->cv[j / cd];
    *dc = coef0 * (c[i] * cvj * exp(-0.5 * mu_ji) / (alpha_sq_i * alpha_i) -
                   c[j] * cvi * exp(0.5 * mu_ji) / (alpha_sq_j * alpha_j));
    double d_i = exp(-0.5 * mu_ji) / (alpha_sq_i * alpha_i);
    double d_j = exp(0.5 * mu_ji) / (alpha_sq_j * alpha_j);
    M[0] = beta * (theta_i * c[i] * cvj * d_i - theta_j * c[j] * cvi * d_j);
    M[1] = d_i;
    M[2] = d_j;
    M[3] = coef0;
    return 1;
}

static void dmd_exchange_params(const o_dmd* d, const double* r_crd, const double* masses, double* rsq_crd,
                                double* c1, double* beta)
{
    for (int e = 0; e < d->ff->ne && e < 8; e++) {
        rsq_crd[e] = r_crd[e] * r_crd[e];
        c1[e] = sqrt(0.5 * d->kBT / masses[e]) / M_PI;  /* set_temp (src/ff_eam_dmd.cpp:1062-1066) */
    }
    *beta = 1.0 / d->kBT;  /* src/ff_dmd.cpp:166 */
}

/* ForceFieldDMD::c_d_calc (src/ff_dmd.cpp:449-456) with
 * ForceFieldEAMDMD::__c_d_calc (src/ff_eam_dmd.cpp:1359-1473): calc_mu's
 * chemical potentials, the second list of NeighborDMD::create_2nd_list
 * (src/neighbor_dmd.cpp:224-273: same element, RSQ < r_crd^2, from the half
 * list) and the exchange rates. c_d [n*c_dim] of the local components;
 * returns max |c_d|. */
double o_dmd_c_d(o_dmd* d, const double* r_crd, const double* zeta, const double* masses, double* c_d)
{
    int cd = d->c_dim;
    int n = d->n * cd;
    dmd_static s;
    dmd_static_build(d, &s);
    double rsq_crd[8], c1[8], beta, M[4];
    dmd_exchange_params(d, r_crd, masses, rsq_crd, c1, &beta);
    for (int i = 0; i < n; i++) c_d[i] = 0.0;
    double norm = 0.0;
    for (int i = 0; i < n; i++) {
        if (d->c[i] < 0.0) continue;
        for (int64_t k = d->row[i]; k < d->row[i + 1]; k++) {
            int j = d->col[k];
            double dc;
            if (!dmd_exchange(d, &s, i, j, rsq_crd, zeta, c1, beta, &dc, M)) continue;
            c_d[i] += dc;
            if (j < n) c_d[j] -= dc;
        }
        if (fabs(c_d[i]) > norm) norm = fabs(c_d[i]);
    }
    dmd_static_free(&s);
    return norm;
}

/* ForceFieldEAMDMD::__J(x, Ax) (src/ff_eam_dmd.cpp:1475-1575) at the current
 * state: the derivative of the rates c_d with respect to the concentrations,
 * applied to xc [n*c_dim] (ghost values follow their owners: update(x_ptr)). */
void o_dmd_J(o_dmd* d, const double* r_crd, const double* zeta, const double* masses, const double* xc, double* Ax)
{
    int cd = d->c_dim;
    int n = d->n * cd, nall = d->g.nall * cd;
    const double* c = d->c;
    dmd_static s;
    dmd_static_build(d, &s);
    double* xa = (double*)calloc((size_t)nall + 1, sizeof(double));
    double* A1 = (double*)calloc((size_t)nall + 1, sizeof(double));
    double* b0 = (double*)calloc((size_t)nall + 1, sizeof(double));
    double* xv = (double*)calloc((size_t)d->g.nall + 1, sizeof(double));
    memcpy(xa, xc, sizeof(double) * n);
    for (int k = d->n; k < d->g.nall; k++)
        for (int ic = 0; ic < cd; ic++) xa[k * cd + ic] = xa[d->g.src[k] * cd + ic];
    /* A1_i = c_i ddE_i sum_j rho_{j->i} x_j: the change of F'(rho_i), times c_i */
    for (int i = 0; i < n; i++) {
        double ci_ddEi = c[i] * s.ddE[i];
        for (int64_t k = d->row[i]; k < d->row[i + 1]; k++) {
            int j = d->col[k];
            A1[i] += ci_ddEi * s.rp[3 * k + 2] * xa[j];
            if (j < n) A1[j] += c[j] * s.ddE[j] * s.rp[3 * k + 1] * xa[i];
        }
    }
    for (int k = d->n; k < d->g.nall; k++)
        for (int ic = 0; ic < cd; ic++) A1[k * cd + ic] = A1[d->g.src[k] * cd + ic];
    /* b0_i: the change of the chemical potential mu_i */
    for (int i = 0; i < n; i++)
        for (int64_t k = d->row[i]; k < d->row[i + 1]; k++) {
            int j = d->col[k];
            double t0 = s.rp[3 * k] + s.dE[i] * s.rp[3 * k + 2] + s.dE[j] * s.rp[3 * k + 1];
            b0[i] += s.rp[3 * k + 1] * A1[j] + t0 * xa[j];
            if (j < n) b0[j] += s.rp[3 * k + 2] * A1[i] + t0 * xa[i];
        }
    for (int k = d->n; k < d->g.nall; k++)
        for (int ic = 0; ic < cd; ic++) b0[k * cd + ic] = b0[d->g.src[k] * cd + ic];
    for (int k = 0; k < d->g.nall; k++)
        for (int ic = 0; ic < cd; ic++)
            if (c[k * cd + ic] >= 0.0) xv[k] += xa[k * cd + ic];
    double rsq_crd[8], c1[8], beta, M[4];
    dmd_exchange_params(d, r_crd, masses, rsq_crd, c1, &beta);
    for (int i = 0; i < n; i++) Ax[i] = 0.0;
    for (int i = 0; i < n; i++) {
        if (c[i] < 0.0) continue;
        for (int64_t k = d->row[i]; k < d->row[i + 1]; k++) {
            int j = d->col[k];
            double dc;
            if (!dmd_exchange(d, &s, i, j, rsq_crd, zeta, c1, beta, &dc, M)) continue;
            double t = M[0] * (b0[j] - b0[i]);
            t += M[1] * (s.cv[j / cd] * xa[i] - c[i] * xv[j / cd]);
            t -= M[2] * (s.cv[i / cd] * xa[j] - c[j] * xv[i / cd]);
            t *= M[3];
            Ax[i] += t;
            if (j < n) Ax[j] -= t;
        }
    }
    free(xa); free(A1); free(b0); free(xv);
    dmd_static_free(&s);
}

/* __GMRES (src/gmres.h:40-60, src/gmres.cpp:32-211): restarted-free GMRES with
 * one classical Gram-Schmidt pass and Givens rotations, zero initial guess.
 * The matrix enters as a callback y = A(x). */
typedef void (*o_matvec)(void* ctx, const double* x, double* y);

typedef struct {
    int m, n;
    double* Q;       /* [m+1][n] Krylov basis */
    double* H;       /* packed columns: column i has i+1 entries at H + i(i+1)/2 */
    double* b_hat;   /* [m+1] rotated right-hand side */
    double* cs;      /* [m+1][2] rotations */
    double* y;       /* [m+1] */
} o_gmres;

static void gmres_init(o_gmres* g, int m, int n)
{
    g->m = m;
    g->n = n;
    g->Q = (double*)calloc((size_t)(m + 1) * n + 1, sizeof(double));
    g->H = (double*)calloc((size_t)m * (m + 1) / 2 + 1, sizeof(double));
    g->b_hat = (double*)calloc(m + 2, sizeof(double));
    g->cs = (double*)calloc(2 * (m + 2), sizeof(double));
    g->y = (double*)calloc(m + 2, sizeof(double));
}

static void gmres_free(o_gmres* g)
{
    free(g->Q); free(g->H); free(g->b_hat); free(g->cs); free(g->y);
}

/* back substitution and x = sum y_k q_k (src/gmres.cpp:113-211); returns |y| */
static double gmres_solve_y(o_gmres* g, int nvecs, double* x)
{
    for (int i = nvecs - 1; i > -1; i--) {
        g->y[i] = g->b_hat[i];
        for (int j = i + 1; j < nvecs; j++) g->y[i] -= g->H[(size_t)j * (j + 1) / 2 + i] * g->y[j];
        g->y[i] /= g->H[(size_t)i * (i + 1) / 2 + i];
    }
    for (int i = 0; i < g->n; i++) {
        x[i] = 0.0;
        for (int k = 0; k < nvecs; k++) x[i] += g->y[k] * g->Q[(size_t)k * g->n + i];
    }
    double s = 0.0;
    for (int k = 0; k < nvecs; k++) s += g->y[k] * g->y[k];
    return sqrt(s);
}

/* GMRES::solve: x ~ A^-1 b within |residual| < tol in at most m products;
 * b is overwritten (it holds the products). Returns 1 when the tolerance was
 * met; *norm = |x|. */
static int gmres_solve(o_gmres* g, o_matvec A, void* ctx, double* b, double tol, double* norm, double* x)
{
    int n = g->n, m = g->m;
    /* calc(Ax, x): first basis vector (src/gmres.cpp:32-46) */
    double s = 0.0;
    for (int i = 0; i < n; i++) s += b[i] * b[i];
    double bn = sqrt(s), inv = 1.0 / bn;
    if (!(bn > 0.0)) { /* b = 0: x = 0 (the reference divides by zero here) */
        for (int i = 0; i < n; i++) x[i] = 0.0;
        *norm = 0.0;
        return 1;
    }
    for (int i = 0; i < n; i++) x[i] = g->Q[i] = inv * b[i];
    g->b_hat[0] = bn;
    for (int it = 0; it < m; it++) {
        A(ctx, x, b);
        /* calc(iter, Ax, x) (src/gmres.cpp:50-112) */
        double* h = g->H + (size_t)it * (it + 1) / 2;
        int iv = it + 1;
        for (int j = 0; j < iv; j++) h[j] = 0.0;
        for (int i = 0; i < n; i++)
            for (int j = 0; j < iv; j++) h[j] += b[i] * g->Q[(size_t)j * n + i];
        double* qn = g->Q + (size_t)iv * n;
        double ns = 0.0;
        for (int i = 0; i < n; i++) {
            qn[i] = b[i];
            for (int j = 0; j < iv; j++) qn[i] -= h[j] * g->Q[(size_t)j * n + i];
            ns += qn[i] * qn[i];
        }
        double nrm = sqrt(ns), ninv = 1.0 / nrm;
        for (int i = 0; i < n; i++) x[i] = qn[i] *= ninv;
        for (int i = 0; i < it; i++) {
            double t = g->cs[2 * i] * h[i] - g->cs[2 * i + 1] * h[i + 1];
            h[i + 1] = g->cs[2 * i + 1] * h[i] + g->cs[2 * i] * h[i + 1];
            h[i] = t;
        }
        double t = sqrt(h[it] * h[it] + nrm * nrm);
        g->cs[2 * it] = h[it] / t;
        g->cs[2 * it + 1] = -nrm / t;
        h[it] = g->cs[2 * it] * h[it] - g->cs[2 * it + 1] * nrm;
        g->b_hat[it + 1] = g->cs[2 * it + 1] * g->b_hat[it];
        g->b_hat[it] *= g->cs[2 * it];
        if (fabs(g->b_hat[it + 1]) < tol) {
            *norm = gmres_solve_y(g, it + 1, x);
            return 1;
        }
    }
    *norm = gmres_solve_y(g, m, x);
    return 0;
}

typedef struct { const double* A; int n; } dense_ctx;

static void dense_mv(void* v, const double* x, double* y)
{
    const dense_ctx* c = (const dense_ctx*)v;
    for (int i = 0; i < c->n; i++) {
        double s = 0.0;
        for (int j = 0; j < c->n; j++) s += c->A[(size_t)i * c->n + j] * x[j];
        y[i] = s;
    }
}

int o_gmres_dense(int n, int m, const double* A, const double* b, double tol, double* x, double* norm)
{
    o_gmres g;
    gmres_init(&g, m, n);
    double* bb = (double*)malloc(sizeof(double) * n + 1);
    memcpy(bb, b, sizeof(double) * n);
    dense_ctx c = {A, n};
    int ok = gmres_solve(&g, dense_mv, &c, bb, tol, norm, x);
    free(bb);
    gmres_free(&g);
    return ok;
}

typedef struct {
    o_dmd* d;
    const double *r_crd, *zeta, *masses;
    double beta;
} newton_ctx;

/* the Jacobian of F = c / beta + a - c_d (src/dae_imp.cpp:163-172) */
static void newton_mv(void* v, const double* x, double* y)
{
    newton_ctx* t = (newton_ctx*)v;
    int n = t->d->n * t->d->c_dim;
    o_dmd_J(t->d, t->r_crd, t->zeta, t->masses, x, y);
    double bi = 1.0 / t->beta;
    for (int i = 0; i < n; i++) y[i] = -y[i] + bi * x[i];
}

static void newton_set_c(o_dmd* d, const double* c)
{
    memcpy(d->c_lcl, c, sizeof(double) * d->n * d->c_dim);
    dmd_sync_ghost_state(d);
}

/* DAEImplicit::newton (src/dae_imp.cpp:146-206) with update_c
 * (src/dae_imp.cpp:89-138) for F(c) = c / beta + a - c_d(c) = 0 from the
 * current c, GMRES of at most max_gmres products per step. Tolerances from
 * a_tol sqrt(nc_dof) (src/dae.cpp:128; nc_dof = natms c_dim) and err_fac.
 * On success c holds the solution; otherwise c is restored. Returns 1 on
 * convergence; *iters = Newton steps counted as the reference counts them. */
int o_dmd_newton(o_dmd* d, const double* r_crd, const double* zeta, const double* masses, double beta,
                 const double* a, double a_tol, double err_fac, int max_newton, int max_gmres, int* iters)
{
    int n = d->n * d->c_dim;
    double a_tol_sqrt = a_tol * sqrt((double)n);
    double* c0 = (double*)malloc(sizeof(double) * n + 1);
    double* c = (double*)malloc(sizeof(double) * n + 1);
    double* F = (double*)malloc(sizeof(double) * n + 1);
    double* cd = (double*)malloc(sizeof(double) * n + 1);
    double* del = (double*)malloc(sizeof(double) * n + 1);
    memcpy(c0, d->c_lcl, sizeof(double) * n);
    memcpy(c, d->c_lcl, sizeof(double) * n);
    o_gmres g;
    gmres_init(&g, max_gmres, n);
    newton_ctx ctx = {d, r_crd, zeta, masses, beta};
    double beta_inv = 1.0 / beta;
    /* F_norm */
#define F_NORM()                                                   \
    do {                                                           \
        o_dmd_c_d(d, r_crd, zeta, masses, cd);                     \
        for (int i_ = 0; i_ < n; i_++) F[i_] = beta_inv * c[i_] + a[i_] - cd[i_]; \
    } while (0)
    double res_tol = 0.005 * a_tol_sqrt / err_fac;
    double denom = 0.1 * a_tol_sqrt / err_fac;
    double norm = 1.0, delta = 0.0, delta_prev = 0.0, ratio = 1.0, R = 1.0;
    int iter = 0, converge = 0, diverge = 0;
    F_NORM();
    while (iter < max_newton && !converge && !diverge) {
        gmres_solve(&g, newton_mv, &ctx, F, res_tol, &norm, del);
        /* update_c: largest r <= 1 keeping 0 <= c - r del <= 1 */
        double r = 1.0;
        for (int i = 0; i < n; i++) {
            if (c[i] >= 0.0) {
                double tmp = c[i] - r * del[i];
                if (tmp > 1.0) {
                    r = (c[i] - 1.0) / del[i];
                    while (c[i] - r * del[i] > 1.0) r = nextafter(r, 0.0);
                }
                if (tmp < 0.0) {
                    r = c[i] / del[i];
                    while (c[i] - r * del[i] < 0.0) r = nextafter(r, 0.0);
                }
            } else
                del[i] = 0.0;
        }
        for (int i = 0; i < n; i++) {
            volatile double v = c[i] - r * del[i];
            v += 1.0;
            v -= 1.0;
            c[i] = v;
        }
        ratio = r;
        newton_set_c(d, c);
        delta = fabs(ratio * norm / denom);
        if (iter) R = fmax(0.3 * R, delta / delta_prev);
        if (iter > 1 && R * delta < 1.0) {
            F_NORM();
            converge = 1;
            continue;
        }
        if (iter > 1 && delta / delta_prev > 2.0) {
            diverge = 1;
            continue;
        }
        delta_prev = delta;
        F_NORM();
        iter++;
    }
#undef F_NORM
    if (!converge) newton_set_c(d, c0);
    *iters = iter;
    gmres_free(&g);
    free(c0); free(c); free(F); free(cd); free(del);
    return converge;
}

void o_dmd_get_c(const o_dmd* d, double* c)
{
    memcpy(c, d->c_lcl, sizeof(double) * d->n * d->c_dim);
}

/* The gradient per unit c that ForceFieldEAMDMD::calc_mu leaves in f and
 * f_alpha (src/ff_eam_dmd.cpp:1189-1278; fcoef of src/ff_eam_dmd.cpp:995-1046):
 * over the half list, pair terms (1/r dphi/dr + dE_i (1/r) drho_ji/dr + dE_j
 * (1/r) drho_ij/dr) weighted by c_j fcoef_i on i and by -c_i fcoef_j on a local
 * j; alpha_i (d/dalpha terms) c_j on both, then -3 kT / alpha_i per present
 * component. g [n][3] per site, ga [n*c_dim]. */
void o_dmd_unit_grad(o_dmd* d, double* g, double* ga)
{
    int cd = d->c_dim, ne = d->ff->ne;
    int n = d->n * cd;
    const double* x = d->g.x;
    const double* c = d->c;
    const double* al = d->alpha;
    dmd_static s;
    dmd_static_build(d, &s);
    double* fcoef = (double*)calloc((size_t)n + 1, sizeof(double));
    for (int site = 0; site < d->n; site++) {
        double ct = 0.0;
        for (int q = 0; q < cd; q++)
            if (c[site * cd + q] > 0.0) ct += c[site * cd + q];
        for (int q = 0; q < cd; q++) {
            int k = site * cd + q;
            if (cd == 1) fcoef[k] = c[k] >= 0.0 ? 1.0 : 0.0;
            else if (ct == 0.0) fcoef[k] = c[k] == 0.0 ? 1.0 : 0.0;
            else fcoef[k] = c[k] >= 0.0 ? c[k] / ct : 0.0;
        }
    }
    for (int k = 0; k < d->n * 3; k++) g[k] = 0.0;
    for (int k = 0; k < n; k++) ga[k] = 0.0;
    double rp[3], dr[3], da[3];
    for (int i = 0; i < n; i++) {
        if (c[i] < 0.0) continue;
        int ei = d->elem[i];
        const double* xi = x + 3 * (i / cd);
        for (int64_t k = d->row[i]; k < d->row[i + 1]; k++) {
            int j = d->col[k];
            int ej = d->elem[j];
            const double* xj = x + 3 * (j / cd);
            double r = sqrt(o_rsq(xi, xj));
            double a = sqrt(al[i] * al[i] + al[j] * al[j]);
            if (r - a * d->xi[d->N - 1] >= d->cut[ei * ne + ej]) continue;
            smear(d, ei, ej, r, a, rp, dr, da);
            double gp = dr[2] * s.dE[i] + dr[1] * s.dE[j] + dr[0];
            double ap = da[2] * s.dE[i] + da[1] * s.dE[j] + da[0];
            double dx[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]};
            for (int q = 0; q < 3; q++) g[3 * (i / cd) + q] += gp * c[j] * fcoef[i] * dx[q];
            ga[i] += al[i] * ap * c[j];
            if (j < n) {
                for (int q = 0; q < 3; q++) g[3 * (j / cd) + q] -= gp * c[i] * fcoef[j] * dx[q];
                ga[j] += al[j] * ap * c[i];
            }
        }
    }
    for (int i = 0; i < n; i++)
        if (c[i] >= 0.0) ga[i] -= 3.0 * d->kBT / al[i];
    free(fcoef);
    dmd_static_free(&s);
}

/* ---------------- second derivatives: the Newton stage over x and alpha ---------------- */

/* 1 drops the second pass of __J (the d dE terms); 2 flips the sign of the
 * (1/a d/da)^2 term; 0 is the restatement. Mutants exist to show that the
 * finite-difference pin of o_dmd_Jxa fails on them. */
int o_jxa_mutation = 0;

/* smear with second derivatives (src/ff_eam_dmd.cpp:420-495): moments of
 * xi^0..xi^4 over the N Gauss-Hermite points, then out[18] = rho_phi[3],
 * (1/r) d/dr [3], (1/a) d/da [3], ((1/r) d/dr)^2 [3], (1/a) d/da (1/r) d/dr
 * [3], ((1/a) d/da)^2 [3], each in the order phi(ei,ej), rho(ei,ej),
 * rho(ej,ei). The derivatives are those of the continuous Gaussian
 * convolution written through the moments, as the reference writes them. */
static void smear2(const o_dmd* d, int ei, int ej, double r, double a, double* o)
{
    const o_eam* e = d->ff;
    double dr_inv = 1.0 / e->dr;
    double H[3][5] = {{0}};
    for (int l = 0; l < d->N; l++) {
        double xl = d->xi[l];
        double u = r - xl * a;
        size_t m;
        double p;
        o_spline_pm(fabs(u), dr_inv, e->nr, &m, &p);
        const double* tabs[3] = {OPHI(e, ei, ej)[m], ORHO(e, ei, ej)[m], ORHO(e, ej, ei)[m]};
        for (int k = 0; k < 3; k++) {
            double t = hval(tabs[k], p);
            if (u < 0.0) t *= -1.0;
            t *= d->wi[l];
            double w = 1.0;
            for (int q = 0; q < 5; q++, w *= xl) H[k][q] += t * w;
        }
    }
    double r_inv = 1.0 / r, r2_inv = r_inv * r_inv, rsq = r * r;
    double a_inv = 1.0 / a, a2_inv = a_inv * a_inv, a4_inv = a2_inv * a2_inv, asq = a * a;
    for (int k = 0; k < 3; k++) {
        double rp = O_PI_IN_SQ * r_inv * H[k][0];
        double dr = -r2_inv * (rp + 2.0 * O_PI_IN_SQ * a_inv * H[k][1]);
        double da = -a2_inv * (rp - 2.0 * O_PI_IN_SQ * r_inv * H[k][2]);
        o[k] = rp;
        o[3 + k] = dr;
        o[6 + k] = da;
        o[9 + k] = -r2_inv * (3.0 * dr - 2.0 * da);
        o[12 + k] = -r2_inv * a2_inv * (3.0 * rp + 3.0 * rsq * dr + asq * da + 4.0 * O_PI_IN_SQ * a_inv * H[k][3]);
        o[15 + k] = -a4_inv * (3.0 * rp + 6.0 * asq * da - 4.0 * O_PI_IN_SQ * r_inv * H[k][4]);
        if (o_jxa_mutation == 2) o[15 + k] = -o[15 + k];
    }
}

void o_dmd_pair_eval2(const o_dmd* d, int ei, int ej, double r, double alpha_ij, double* out18)
{
    smear2(d, ei, ej, r, alpha_ij, out18);
}

/* fcoef of src/ff_eam_dmd.cpp:560-602: 1 for one component; else c_i over the
 * sum of positive c of the site, or 1 for the zero components of an empty site */
static void dmd_fcoef(const o_dmd* d, double* fcoef)
{
    int cd = d->c_dim;
    const double* c = d->c;
    for (int site = 0; site < d->n; site++) {
        double ct = 0.0;
        for (int q = 0; q < cd; q++)
            if (c[site * cd + q] > 0.0) ct += c[site * cd + q];
        for (int q = 0; q < cd; q++) {
            int k = site * cd + q;
            if (cd == 1) fcoef[k] = c[k] >= 0.0 ? 1.0 : 0.0;
            else if (ct == 0.0) fcoef[k] = c[k] == 0.0 ? 1.0 : 0.0;
            else fcoef[k] = c[k] >= 0.0 ? c[k] / ct : 0.0;
        }
    }
}

/* ForceFieldEAMDMD::__prepJ_n_res + __J (src/ff_eam_dmd.cpp:364-831) at the
 * current state: (Jdx, Jda) = J (dx, da) with J = -d f / d(x, alpha) at fixed
 * c, f the gradient per unit c of o_dmd_unit_grad. Over the half list: the
 * pair terms with dE fixed and the change of dE (deltadE = F''(rho) drho)
 * collected in the first pass, its pair terms in the second; the ghost
 * copies of dx, da and deltadE follow their owners. dx [natms][3], da and
 * Jda [natms*c_dim]. */
void o_dmd_Jxa(o_dmd* d, const double* dx, const double* da, double* Jdx, double* Jda)
{
    int cd = d->c_dim, ne = d->ff->ne;
    int n = d->n * cd, nall = d->g.nall * cd;
    const double* x = d->g.x;
    const double* c = d->c;
    const double* al = d->alpha;
    dmd_static s;
    dmd_static_build(d, &s);
    double* fcoef = (double*)calloc((size_t)n + 1, sizeof(double));
    dmd_fcoef(d, fcoef);
    double* dxa = (double*)calloc((size_t)d->g.nall * 3 + 1, sizeof(double));
    double* daa = (double*)calloc((size_t)nall + 1, sizeof(double));
    double* ddE = (double*)calloc((size_t)nall + 1, sizeof(double));
    for (int k = 0; k < d->g.nall; k++) {
        int o = d->g.owner[k];
        for (int q = 0; q < 3; q++) dxa[3 * k + q] = dx[3 * o + q];
        for (int ic = 0; ic < cd; ic++) daa[k * cd + ic] = da[o * cd + ic];
    }
    for (int k = 0; k < 3 * d->n; k++) Jdx[k] = 0.0;
    for (int k = 0; k < n; k++) Jda[k] = 0.0;
    double o[18];
    /* prepJ_n_res's F'' is zero beyond rho_max (src/ff_eam_dmd.cpp:545-553) */
    double rho_max = (double)d->ff->nrho * d->ff->drho;
    for (int i = 0; i < n; i++)
        if (c[i] >= 0.0 && s.rho[i] > rho_max) s.ddE[i] = 0.0;
    for (int k = d->n; k < d->g.nall; k++)
        for (int ic = 0; ic < cd; ic++) s.ddE[k * cd + ic] = s.ddE[d->g.src[k] * cd + ic];
    for (int pass = 0; pass < 2; pass++) {
        if (pass == 1) {
            if (o_jxa_mutation == 1) break;
            for (int k = d->n; k < d->g.nall; k++)
                for (int ic = 0; ic < cd; ic++) ddE[k * cd + ic] = ddE[d->g.src[k] * cd + ic];
        }
        for (int i = 0; i < n; i++) {
            if (c[i] < 0.0) continue;
            int ei = d->elem[i];
            const double* xi = x + 3 * (i / cd);
            const double* dxi = dxa + 3 * (i / cd);
            double ai = al[i], dai = daa[i];
            for (int64_t k = d->row[i]; k < d->row[i + 1]; k++) {
                int j = d->col[k];
                int ej = d->elem[j];
                const double* xj = x + 3 * (j / cd);
                const double* dxj = dxa + 3 * (j / cd);
                double r = sqrt(o_rsq(xi, xj));
                double a = sqrt(ai * ai + al[j] * al[j]);
                if (r - a * d->xi[d->N - 1] >= d->cut[ei * ne + ej]) continue;
                smear2(d, ei, ej, r, a, o);
                double v[3] = {xi[0] - xj[0], xi[1] - xj[1], xi[2] - xj[2]};
                double dv[3] = {dxi[0] - dxj[0], dxi[1] - dxj[1], dxi[2] - dxj[2]};
                double aj = al[j], daj = daa[j];
                if (pass == 0) {
                    double A = v[0] * dv[0] + v[1] * dv[1] + v[2] * dv[2];
                    double B = ai * dai + aj * daj;
                    double w[6];
                    for (int q = 1; q < 6; q++) w[q] = o[3 * q] + o[3 * q + 2] * s.dE[i] + o[3 * q + 1] * s.dE[j];
                    /* w[1] (1/r)d/dr, w[2] (1/a)d/da, w[3] rr, w[4] ra, w[5] aa */
                    double pa0 = -(w[4] * A + w[5] * B), pa1 = -w[2];
                    double df = -(w[4] * B + w[3] * A), fp = -w[1];
                    ddE[i] += (A * o[3 + 2] + B * o[6 + 2]) * s.ddE[i] * c[j];
                    for (int q = 0; q < 3; q++) Jdx[3 * (i / cd) + q] += (df * v[q] + fp * dv[q]) * c[j] * fcoef[i];
                    Jda[i] += c[j] * (pa0 * ai + pa1 * dai);
                    if (j < n) {
                        ddE[j] += (A * o[3 + 1] + B * o[6 + 1]) * s.ddE[j] * c[i];
                        for (int q = 0; q < 3; q++)
                            Jdx[3 * (j / cd) + q] -= (df * v[q] + fp * dv[q]) * c[i] * fcoef[j];
                        Jda[j] += c[i] * (pa0 * aj + pa1 * daj);
                    }
                } else {
                    double df = -(o[3 + 2] * ddE[i] + o[3 + 1] * ddE[j]);
                    double pa = -(o[6 + 2] * ddE[i] + o[6 + 1] * ddE[j]);
                    for (int q = 0; q < 3; q++) Jdx[3 * (i / cd) + q] += df * c[j] * fcoef[i] * v[q];
                    Jda[i] += ai * pa * c[j];
                    if (j < n) {
                        for (int q = 0; q < 3; q++) Jdx[3 * (j / cd) + q] -= df * c[i] * fcoef[j] * v[q];
                        Jda[j] += aj * pa * c[i];
                    }
                }
            }
        }
    }
    for (int i = 0; i < n; i++)
        if (c[i] >= 0.0) Jda[i] -= 3.0 * d->kBT * da[i] / (al[i] * al[i]);
    free(fcoef); free(dxa); free(daa); free(ddE);
    dmd_static_free(&s);
}

typedef struct { o_dmd* d; double *gx, *ga; } jxa_ctx;

static void jxa_mv(void* v, const double* in, double* out)
{
    jxa_ctx* q = (jxa_ctx*)v;
    int n3 = 3 * q->d->n;
    o_dmd_Jxa(q->d, in, in + n3, out, out + n3);
}

/* the residual [3 natms + natms c_dim] and its norm */
static double jxa_residual(o_dmd* d, double* f)
{
    int n3 = 3 * d->n, n = d->n * d->c_dim;
    o_dmd_unit_grad(d, f, f + n3);
    double s = 0.0;
    for (int k = 0; k < n3 + n; k++) s += f[k] * f[k];
    return sqrt(s);
}

/* The Newton stage of DAE::min_error_false (src/dae.cpp:294-325): while
 * |f| > a_tol sqrt(3 natms + natms c_dim) and fewer than max_newton steps,
 * GMRES (max_gmres products, tolerance 0.005 a_tol sqrt(...)) solves J h = f
 * from zero; r is the largest fraction <= 1 keeping every alpha positive
 * (MinDMDHandler::max_alpha_lcl_alpha, src/min_cg_dmd.h:347-360); x and
 * alpha take h whole when r = 1, else r/2 h; then the list is rebuilt and f
 * recomputed. Returns the steps taken; *res the final |f|. */
int o_dmd_newton_xa(o_dmd* d, double a_tol, int max_newton, int max_gmres, double* res_out)
{
    int n3 = 3 * d->n, n = d->n * d->c_dim, nv = n3 + n;
    double* f = (double*)calloc((size_t)nv + 1, sizeof(double));
    double* b = (double*)calloc((size_t)nv + 1, sizeof(double));
    double* h = (double*)calloc((size_t)nv + 1, sizeof(double));
    double* xl = (double*)malloc(sizeof(double) * (n3 + 1));
    o_gmres g;
    gmres_init(&g, max_gmres, nv);
    jxa_ctx ctx = {d, NULL, NULL};
    double tol_s = a_tol * sqrt((double)nv);
    double res = jxa_residual(d, f), norm;
    int istep = 0;
    for (; istep < max_newton && res / tol_s > 1.0; istep++) {
        memcpy(b, f, sizeof(double) * nv);
        gmres_solve(&g, jxa_mv, &ctx, b, 0.005 * tol_s, &norm, h);
        double r = 1.0, hmax = 0.0;
        for (int k = 0; k < n; k++) {
            double ha = h[n3 + k];
            if (ha < 0.0) r = fmin(-d->alpha_lcl[k] / ha, r);
            hmax = fmax(hmax, fabs(ha));
        }
        r = fmin(r, fabs(INFINITY / hmax));
        double st = r == 1.0 ? 1.0 : 0.5 * r;
        memcpy(xl, d->g.x, sizeof(double) * n3);
        for (int k = 0; k < n3; k++) xl[k] += st * h[k];
        for (int k = 0; k < n; k++) d->alpha_lcl[k] += st * h[n3 + k];
        dmd_rebuild(d, xl);
        res = jxa_residual(d, f);
    }
    *res_out = res;
    gmres_free(&g);
    free(f); free(b); free(h); free(xl);
    return istep;
}

/* the local positions [natms][3] (as the list holds them) and alpha [natms*c_dim] */
void o_dmd_get_xa(const o_dmd* d, double* x, double* alpha)
{
    memcpy(x, d->g.x, sizeof(double) * 3 * d->n);
    memcpy(alpha, d->alpha_lcl, sizeof(double) * d->n * d->c_dim);
}
