/*
 * o_md.c -- oracle: periodic box, ghost images, Cell<1> binning, the MD
 * half neighbor list, two-pass EAM force/energy/virial, the skin check and
 * the NVT (velocity Verlet + Nose-Hoover chain) driver.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle). Built with -ffp-contract=off so
 * that every product is rounded before it is added, as in the reference.
 */
#include "o_internal.h"
#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <stdlib.h>
#include <string.h>

/* ---------------- box (src/atoms.cpp:187-194, src/xmath.h:1403-1485) ---------------- */
void o_box_set(o_box* bx, const double* H)
{
    memset(bx, 0, sizeof(*bx));
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) bx->H[i][j] = j <= i ? H[i * 3 + j] : 0.0;
    double(*h)[3] = bx->H;
    double(*b)[3] = bx->B;
    /* lower-triangular inverse, row by row */
    b[0][0] = 1.0 / h[0][0];
    b[1][1] = 1.0 / h[1][1];
    b[1][0] = -b[1][1] * (b[0][0] * h[1][0]);
    b[2][2] = 1.0 / h[2][2];
    b[2][1] = -b[2][2] * (b[1][1] * h[2][1]);
    b[2][0] = -b[2][2] * (b[0][0] * h[2][0] + b[1][0] * h[2][1]);
    bx->vol = h[0][0] * (h[1][1] * h[2][2]);
    bx->depth_inv[0] = sqrt(b[0][0] * b[0][0] + (b[1][0] * b[1][0] + b[2][0] * b[2][0]));
    bx->depth_inv[1] = sqrt(b[1][1] * b[1][1] + b[2][1] * b[2][1]);
    bx->depth_inv[2] = fabs(b[2][2]);
    /* packed by columns of the lower triangle */
    bx->h[0] = h[0][0]; bx->h[1] = h[1][0]; bx->h[2] = h[2][0];
    bx->h[3] = h[1][1]; bx->h[4] = h[2][1]; bx->h[5] = h[2][2];
    bx->b[0] = b[0][0]; bx->b[1] = b[1][0]; bx->b[2] = b[2][0];
    bx->b[3] = b[1][1]; bx->b[4] = b[2][1]; bx->b[5] = b[2][2];
}

static double wrap01(double s)
{
    while (s < 0.0) ++s;
    while (s >= 1.0) --s;
    return s;
}

/* x -> s = x B with wrap into [0,1) per component (src/xmath.h:2660-2700) */
void o_x2s(const o_box* bx, const double* x, double* s)
{
    const double* b = bx->b;
    double s0 = b[0] * x[0] + (b[1] * x[1] + b[2] * x[2]);
    s0 = wrap01(s0);
    double s1 = b[3] * x[1] + b[4] * x[2];
    s1 = wrap01(s1);
    double s2 = b[5] * x[2];
    s2 = wrap01(s2);
    s[0] = s0; s[1] = s1; s[2] = s2;
}

/* s -> x = s H (src/xmath.h:2636-2657) */
void o_s2x(const o_box* bx, const double* s, double* x)
{
    const double* h = bx->h;
    double x0 = h[0] * s[0] + (h[1] * s[1] + h[2] * s[2]);
    double x1 = h[3] * s[1] + h[4] * s[2];
    double x2 = h[5] * s[2];
    x[0] = x0; x[1] = x1; x[2] = x2;
}

/* dx0^2 + (dx1^2 + dx2^2), the association of src/xmath.h:2464-2477 */
double o_rsq(const double* a, const double* b)
{
    double d0 = a[0] - b[0], d1 = a[1] - b[1], d2 = a[2] - b[2];
    return d0 * d0 + (d1 * d1 + d2 * d2);
}

/* ---------------- ghosts (src/exchange.h:501-1096, one rank) ---------------- */
static void ghosts_grow(o_ghosts* g, int need)
{
    if (need <= g->cap) return;
    int cap = need + need / 2 + 64;
    g->s = (double*)realloc(g->s, sizeof(double) * 3 * cap);
    g->x = (double*)realloc(g->x, sizeof(double) * 3 * cap);
    g->src = (int*)realloc(g->src, sizeof(int) * cap);
    g->owner = (int*)realloc(g->owner, sizeof(int) * cap);
    g->img = (int8_t*)realloc(g->img, 3 * cap);
    g->sdim = (int8_t*)realloc(g->sdim, cap);
    g->sdir = (int8_t*)realloc(g->sdir, cap);
    g->cap = cap;
}

static void ghost_copy(o_ghosts* g, int from, int d, int dir)
{
    int k = g->nall++;
    memcpy(g->s + 3 * k, g->s + 3 * from, sizeof(double) * 3);
    if (dir > 0) ++g->s[3 * k + d]; else --g->s[3 * k + d];
    g->src[k] = from;
    g->owner[k] = g->owner[from];
    memcpy(g->img + 3 * k, g->img + 3 * from, 3);
    g->img[3 * k + d] += dir;
    g->sdim[k] = (int8_t)d;
    g->sdir[k] = (int8_t)dir;
}

/* Per dimension: atoms within max_cut*depth_inv of the low face are copied
 * with s+1, then atoms within it of the high face with s-1; thin boxes repeat
 * the copy on the freshly made ghosts (ncomms layers). */
void o_ghosts_build(o_ghosts* g, const o_box* bx, int nlcl, const double* s_lcl, double max_cut)
{
    g->nlcl = nlcl;
    g->nall = nlcl;
    ghosts_grow(g, nlcl);
    memcpy(g->s, s_lcl, sizeof(double) * 3 * nlcl);
    for (int i = 0; i < nlcl; i++) {
        g->src[i] = -1; g->owner[i] = i;
        g->img[3 * i] = g->img[3 * i + 1] = g->img[3 * i + 2] = 0;
        g->sdim[i] = -1; g->sdir[i] = 0;
    }
    int* snd = NULL;
    int snd_cap = 0;
    for (int d = 0; d < 3; d++) {
        double cut_s = max_cut * bx->depth_inv[d];
        int ncomms = (int)(cut_s / 1.0) + 1;
        int last = g->nall;
        for (int dir = 0; dir < 2; dir++) {
            double bnd = dir == 0 ? 0.0 + cut_s : 1.0 - cut_s;
            int lo = 0, hi = last;
            for (int ic = 0; ic < ncomms; ic++) {
                int ns = 0;
                for (int i = lo; i < hi; i++) {
                    double v = g->s[3 * i + d];
                    int send = dir == 0 ? (v < bnd) : (v >= bnd);
                    if (send) {
                        if (ns == snd_cap) { snd_cap = snd_cap * 2 + 1024; snd = (int*)realloc(snd, sizeof(int) * snd_cap); }
                        snd[ns++] = i;
                    }
                }
                ghosts_grow(g, g->nall + ns);
                for (int k = 0; k < ns; k++) ghost_copy(g, snd[k], d, dir == 0 ? 1 : -1);
                lo = g->nall - ns;
                hi = g->nall;
            }
        }
    }
    free(snd);
    for (int i = 0; i < g->nall; i++) o_s2x(bx, g->s + 3 * i, g->x + 3 * i);
}

/* Update::update_w_x: ghost x = source x +/- row d of H (src/exchange.h:797-827, 959-967) */
void o_ghosts_refresh(o_ghosts* g, const o_box* bx, const double* x_lcl)
{
    memcpy(g->x, x_lcl, sizeof(double) * 3 * g->nlcl);
    for (int k = g->nlcl; k < g->nall; k++) {
        const double* xs = g->x + 3 * g->src[k];
        double* xk = g->x + 3 * k;
        int d = g->sdim[k];
        xk[0] = xs[0]; xk[1] = xs[1]; xk[2] = xs[2];
        for (int c = 0; c <= d; c++) {
            if (g->sdir[k] > 0) xk[c] += bx->H[d][c]; else xk[c] -= bx->H[d][c];
        }
    }
}

void o_ghosts_free(o_ghosts* g)
{
    free(g->s); free(g->x); free(g->src); free(g->owner); free(g->img); free(g->sdim); free(g->sdir);
    memset(g, 0, sizeof(*g));
}

/* ---------------- Cell<1>::DO (src/cell.h:38-233) ---------------- */
void o_cell_do(const o_ghosts* g, const o_box* bx, double max_cut,
               void (*pre)(void*, int), void (*visit)(void*, int, int), void (*post)(void*, int), void* ctx)
{
    const int M = 1;
    double cut_s[3], csz[3];
    int nc[3], den[3], ncells = 1;
    for (int d = 0; d < 3; d++) {
        cut_s[d] = max_cut * bx->depth_inv[d];
        csz[d] = cut_s[d] / (double)M;
        nc[d] = (int)((1.0 - 0.0) / csz[d]) + 1 + 2 * M;
        den[d] = ncells;
        ncells *= nc[d];
    }
    int rel[27];
    {
        int cnt[3] = {-1, -1, -1};
        for (int k = 0; k < 27; k++) {
            rel[k] = cnt[0] * den[0] + cnt[1] * den[1] + cnt[2] * den[2];
            cnt[0]++;
            for (int d = 0; d < 2; d++)
                if (cnt[d] == M + 1) { cnt[d] = -M; cnt[d + 1]++; }
        }
    }
    int* head = (int*)malloc(sizeof(int) * ncells);
    for (int c = 0; c < ncells; c++) head[c] = -1;
    int* next = (int*)malloc(sizeof(int) * (g->nall > 0 ? g->nall : 1));
    int* cell = (int*)malloc(sizeof(int) * (g->nall > 0 ? g->nall : 1));
    for (int i = g->nall - 1; i > g->nlcl - 1; i--) {
        const double* s = g->s + 3 * i;
        int c = 0;
        for (int d = 0; d < 3; d++) {
            int k;
            if (s[d] < 0.0) {
                k = (int)((s[d] - (0.0 - cut_s[d])) / csz[d]);
                if (k > M - 1) k = M - 1;
            } else if (1.0 <= s[d]) {
                k = (int)((s[d] - 0.0) / csz[d]);
                if (k < nc[d] - 2 * M - 1) k = nc[d] - 2 * M - 1;
                if (k > nc[d] - 1 - M) k = nc[d] - 1 - M;
                k += M;
            } else {
                k = (int)((s[d] - 0.0) / csz[d]);
                if (k > nc[d] - 2 * M - 1) k = nc[d] - 2 * M - 1;
                k += M;
            }
            c += den[d] * k;
        }
        cell[i] = c;
        next[i] = head[c];
        head[c] = i;
    }
    for (int i = g->nlcl - 1; i > -1; i--) {
        const double* s = g->s + 3 * i;
        int c = 0;
        for (int d = 0; d < 3; d++) {
            int k = (int)((s[d] - 0.0) / csz[d]);
            if (k > nc[d] - 2 * M - 1) k = nc[d] - 2 * M - 1;
            c += den[d] * (k + M);
        }
        cell[i] = c;
        next[i] = head[c];
        head[c] = i;
    }
    for (int i = 0; i < g->nlcl; i++) {
        if (pre) pre(ctx, i);
        for (int k = 0; k < 27; k++) {
            for (int j = head[cell[i] + rel[k]]; j != -1; j = next[j])
                if (i != j) visit(ctx, i, j);
        }
        if (post) post(ctx, i);
    }
    free(head); free(next); free(cell);
}

/* ---------------- MD system ---------------- */
struct o_md {
    const o_eam* ff;
    o_box bx;
    o_ghosts g;
    int n;
    double skin, max_cut;
    double* cut_sq;    /* [ne*ne] */
    double* cut_sk_sq; /* [ne*ne] */
    uint8_t* elem;     /* [nall] after ghosts */
    uint8_t* elem_lcl;
    double* x0;
    int64_t* row;      /* [n+1] CSR of the half list */
    int32_t* col;
    int64_t npairs, cap;
    double *rho, *dF, *drhoi, *drhoj;
    int64_t nscratch;
    double* thr_rows; /* per-thread rho and f rows of the threaded force */
    size_t thr_cap;
};

typedef struct { o_md* md; const double* xi; int ei; } list_ctx;

static void list_visit(void* v, int i, int j)
{
    list_ctx* c = (list_ctx*)v;
    o_md* md = c->md;
    if (i < j) {
        int ne = md->ff->ne;
        if (o_rsq(c->xi, md->g.x + 3 * j) < md->cut_sk_sq[c->ei * ne + md->elem[j]]) {
            if (md->npairs == md->cap) {
                md->cap = md->cap * 2 + 1024;
                md->col = (int32_t*)realloc(md->col, sizeof(int32_t) * md->cap);
            }
            md->col[md->npairs++] = j;
        }
    }
}

static void list_pre(void* v, int i)
{
    list_ctx* c = (list_ctx*)v;
    c->xi = c->md->g.x + 3 * i;
    c->ei = c->md->elem[i];
}

static void list_post(void* v, int i)
{
    list_ctx* c = (list_ctx*)v;
    c->md->row[i + 1] = c->md->npairs;
}

/* ---------------- threads (CPU baseline only) ----------------
 * With o_threads > 1 (o_set_threads) the list build and the EAM force run on
 * that many OpenMP threads: the same rows in the same order, the same
 * arithmetic per pair, sums of per-thread rows in a different order. The
 * checker itself always runs with one thread. */
static int o_threads = 1;

void o_set_threads(int n)
{
#ifdef _OPENMP
    o_threads = n > 1 ? (n < 16 ? n : 16) : 1;
#else
    (void)n;
    o_threads = 1;
#endif
}

int o_get_threads(void) { return o_threads; }

/* Cell<1>::DO's traversal for one local atom i (the sequential loop of
 * o_cell_do): calls emit(ctx, j) for every j of the half list in list order */
typedef struct {
    int* head; int* next; int* cell; int rel[27];
} o_cells;

static void cells_build(o_cells* cs, const o_ghosts* g, const o_box* bx, double max_cut)
{
    const int M = 1;
    double cut_s[3], csz[3];
    int nc[3], den[3], ncells = 1;
    for (int d = 0; d < 3; d++) {
        cut_s[d] = max_cut * bx->depth_inv[d];
        csz[d] = cut_s[d] / (double)M;
        nc[d] = (int)((1.0 - 0.0) / csz[d]) + 1 + 2 * M;
        den[d] = ncells;
        ncells *= nc[d];
    }
    int cnt[3] = {-1, -1, -1};
    for (int k = 0; k < 27; k++) {
        cs->rel[k] = cnt[0] * den[0] + cnt[1] * den[1] + cnt[2] * den[2];
        cnt[0]++;
        for (int d = 0; d < 2; d++)
            if (cnt[d] == M + 1) { cnt[d] = -M; cnt[d + 1]++; }
    }
    cs->head = (int*)malloc(sizeof(int) * ncells);
    for (int c = 0; c < ncells; c++) cs->head[c] = -1;
    cs->next = (int*)malloc(sizeof(int) * (g->nall > 0 ? g->nall : 1));
    cs->cell = (int*)malloc(sizeof(int) * (g->nall > 0 ? g->nall : 1));
    for (int i = g->nall - 1; i > g->nlcl - 1; i--) {
        const double* sv = g->s + 3 * i;
        int c = 0;
        for (int d = 0; d < 3; d++) {
            int k;
            if (sv[d] < 0.0) {
                k = (int)((sv[d] - (0.0 - cut_s[d])) / csz[d]);
                if (k > M - 1) k = M - 1;
            } else if (1.0 <= sv[d]) {
                k = (int)((sv[d] - 0.0) / csz[d]);
                if (k < nc[d] - 2 * M - 1) k = nc[d] - 2 * M - 1;
                if (k > nc[d] - 1 - M) k = nc[d] - 1 - M;
                k += M;
            } else {
                k = (int)((sv[d] - 0.0) / csz[d]);
                if (k > nc[d] - 2 * M - 1) k = nc[d] - 2 * M - 1;
                k += M;
            }
            c += den[d] * k;
        }
        cs->cell[i] = c;
        cs->next[i] = cs->head[c];
        cs->head[c] = i;
    }
    for (int i = g->nlcl - 1; i > -1; i--) {
        const double* sv = g->s + 3 * i;
        int c = 0;
        for (int d = 0; d < 3; d++) {
            int k = (int)((sv[d] - 0.0) / csz[d]);
            if (k > nc[d] - 2 * M - 1) k = nc[d] - 2 * M - 1;
            c += den[d] * (k + M);
        }
        cs->cell[i] = c;
        cs->next[i] = cs->head[c];
        cs->head[c] = i;
    }
}

static void cells_free(o_cells* cs) { free(cs->head); free(cs->next); free(cs->cell); }

/* the half list by threads: rows counted, offsets summed, rows filled, each
 * row in the order of the sequential traversal */
static void list_build_threaded(o_md* md)
{
    o_cells cs;
    cells_build(&cs, &md->g, &md->bx, md->max_cut);
    const int n = md->n, ne = md->ff->ne;
    const double* x = md->g.x;
    int64_t* cnt = (int64_t*)calloc((size_t)n + 1, sizeof(int64_t));
#pragma omp parallel for schedule(dynamic, 64) num_threads(o_threads)
    for (int i = 0; i < n; i++) {
        int64_t c = 0;
        const int ei = md->elem[i];
        for (int k = 0; k < 27; k++)
            for (int j = cs.head[cs.cell[i] + cs.rel[k]]; j != -1; j = cs.next[j])
                if (i < j && o_rsq(x + 3 * i, x + 3 * j) < md->cut_sk_sq[ei * ne + md->elem[j]]) c++;
        cnt[i + 1] = c;
    }
    md->row[0] = 0;
    for (int i = 0; i < n; i++) md->row[i + 1] = md->row[i] + cnt[i + 1];
    free(cnt);
    md->npairs = md->row[n];
    if (md->cap < md->npairs) {
        md->cap = md->npairs + 1024;
        md->col = (int32_t*)realloc(md->col, sizeof(int32_t) * md->cap);
    }
#pragma omp parallel for schedule(dynamic, 64) num_threads(o_threads)
    for (int i = 0; i < n; i++) {
        int64_t c = md->row[i];
        const int ei = md->elem[i];
        for (int k = 0; k < 27; k++)
            for (int j = cs.head[cs.cell[i] + cs.rel[k]]; j != -1; j = cs.next[j])
                if (i < j && o_rsq(x + 3 * i, x + 3 * j) < md->cut_sk_sq[ei * ne + md->elem[j]]) md->col[c++] = j;
    }
    cells_free(&cs);
}

/* rebuild: wrap locals (x2s_lcl), ghosts (Update::list), Cell<1>::DO with
 * s2x_all, NeighborMD::create_list (src/neighbor_md.cpp:82-152), store x0 */
static void md_rebuild(o_md* md, const double* x_lcl)
{
    double* s = (double*)malloc(sizeof(double) * 3 * (md->n > 0 ? md->n : 1));
    for (int i = 0; i < md->n; i++) o_x2s(&md->bx, x_lcl + 3 * i, s + 3 * i);
    o_ghosts_build(&md->g, &md->bx, md->n, s, md->max_cut);
    free(s);
    md->elem = (uint8_t*)realloc(md->elem, md->g.nall > 0 ? md->g.nall : 1);
    for (int k = 0; k < md->g.nall; k++) md->elem[k] = md->elem_lcl[md->g.owner[k]];
    if (o_threads > 1) {
        list_build_threaded(md);
    } else {
        md->row[0] = 0;
        md->npairs = 0;
        list_ctx c = {md, NULL, 0};
        o_cell_do(&md->g, &md->bx, md->max_cut, list_pre, list_visit, list_post, &c);
    }
    memcpy(md->x0, md->g.x, sizeof(double) * 3 * md->n);
}

o_md* o_md_create(const o_eam* ff, int natms, const double* H, const double* x,
                  const uint8_t* elem, double skin)
{
    o_md* md = (o_md*)calloc(1, sizeof(o_md));
    md->ff = ff;
    md->n = natms;
    md->skin = skin;
    o_box_set(&md->bx, H);
    int ne = ff->ne;
    md->cut_sq = (double*)malloc(sizeof(double) * ne * ne);
    md->cut_sk_sq = (double*)malloc(sizeof(double) * ne * ne);
    md->max_cut = 0.0;
    /* ForceFieldMD::pre_init (src/ff_md.cpp:80-94) */
    for (int i = 0; i < ne; i++)
        for (int j = 0; j <= i; j++) {
            double t = ff->rc[i * ne + j];
            if (t > md->max_cut) md->max_cut = t;
            md->cut_sq[i * ne + j] = md->cut_sq[j * ne + i] = t * t;
            t += skin;
            md->cut_sk_sq[i * ne + j] = md->cut_sk_sq[j * ne + i] = t * t;
        }
    md->max_cut += skin; /* atoms->max_cut = ff->max_cut + skin (src/dynamic_md.cpp:41) */
    md->elem_lcl = (uint8_t*)malloc(natms > 0 ? natms : 1);
    memcpy(md->elem_lcl, elem, natms);
    md->x0 = (double*)malloc(sizeof(double) * 3 * (natms > 0 ? natms : 1));
    md->row = (int64_t*)malloc(sizeof(int64_t) * (natms + 1));
    md_rebuild(md, x);
    return md;
}

void o_md_free(o_md* md)
{
    if (!md) return;
    o_ghosts_free(&md->g);
    free(md->cut_sq); free(md->cut_sk_sq); free(md->elem); free(md->elem_lcl); free(md->x0);
    free(md->row); free(md->col); free(md->rho); free(md->dF); free(md->drhoi); free(md->drhoj);
    free(md->thr_rows);
    free(md);
}

int o_md_natms(const o_md* md) { return md->n; }
int o_md_nghosts(const o_md* md) { return md->g.nall - md->g.nlcl; }
int64_t o_md_npairs(const o_md* md) { return md->npairs; }

void o_md_pairs(const o_md* md, int32_t* pi, int32_t* pj, int8_t* img)
{
    for (int i = 0; i < md->n; i++)
        for (int64_t k = md->row[i]; k < md->row[i + 1]; k++) {
            int j = md->col[k];
            pi[k] = i;
            pj[k] = md->g.owner[j];
            img[3 * k] = md->g.img[3 * j];
            img[3 * k + 1] = md->g.img[3 * j + 1];
            img[3 * k + 2] = md->g.img[3 * j + 2];
        }
}

void o_md_get_x(const o_md* md, double* x) { memcpy(x, md->g.x, sizeof(double) * 3 * md->n); }

void o_md_set_x(o_md* md, const double* x) { memcpy(md->g.x, x, sizeof(double) * 3 * md->n); }

/* DynamicMD::decide (src/dynamic_md.cpp:90-105) */
static int md_decide(const o_md* md)
{
    double skin_sq = 0.25 * md->skin * md->skin;
    for (int i = 0; i < md->n; i++)
        if (o_rsq(md->x0 + 3 * i, md->g.x + 3 * i) > skin_sq) return 0;
    return 1;
}

int o_md_update(o_md* md)
{
    if (md_decide(md)) {
        o_ghosts_refresh(&md->g, &md->bx, md->g.x);
        return 0;
    }
    double* x = (double*)malloc(sizeof(double) * 3 * (md->n > 0 ? md->n : 1));
    memcpy(x, md->g.x, sizeof(double) * 3 * md->n);
    md_rebuild(md, x);
    free(x);
    return 1;
}

static inline double hval(const double* c, double p) { return ((c[3] * p + c[2]) * p + c[1]) * p + c[0]; }
static inline double hder(const double* c, double p, double inv) { return ((3.0 * c[3] * p + 2.0 * c[2]) * p + c[1]) * inv; }

/* ForceFieldEAM::__force_calc (src/ff_eam.cpp:65-240): pass 1 gives rho,
 * the pair part and F(rho_i) in half-list order; the F' of every atom is then
 * known and pass 2 adds the embedding forces. */
static void eam_force(o_md* md, double* f, double* v7, double* rho_out)
{
    const o_eam* e = md->ff;
    int ne = e->ne;
    int n = md->n;
    const double* x = md->g.x;
    const uint8_t* el = md->elem;
    double dr_inv = 1.0 / e->dr, drho_inv = 1.0 / e->drho;
    double rho_max = (double)e->nrho * e->drho;
    if (md->nscratch < md->npairs || !md->rho) {
        free(md->drhoi); free(md->drhoj);
        md->nscratch = md->npairs > 0 ? md->npairs : 1;
        md->drhoi = (double*)malloc(sizeof(double) * md->nscratch);
        md->drhoj = (double*)malloc(sizeof(double) * md->nscratch);
    }
    md->rho = (double*)realloc(md->rho, sizeof(double) * (n > 0 ? n : 1));
    md->dF = (double*)realloc(md->dF, sizeof(double) * (md->g.nall > 0 ? md->g.nall : 1));
    double* rho = md->rho;
    double* dF = md->dF;
    for (int i = 0; i < 3 * n; i++) f[i] = 0.0;
    for (int k = 0; k < 7; k++) v7[k] = 0.0;
    for (int i = 0; i < n; i++) rho[i] = 0.0;
    for (int i = 0; i < n; i++) {
        int ei = el[i];
        for (int64_t k = md->row[i]; k < md->row[i + 1]; k++) {
            int j = md->col[k];
            int ej = el[j];
            double dx0 = x[3 * i] - x[3 * j];
            double dx1 = x[3 * i + 1] - x[3 * j + 1];
            double dx2 = x[3 * i + 2] - x[3 * j + 2];
            double rsq = dx0 * dx0 + dx1 * dx1 + dx2 * dx2;
            md->drhoi[k] = md->drhoj[k] = 0.0;
            if (rsq >= md->cut_sq[ei * ne + ej]) continue;
            double r = sqrt(rsq);
            double r_inv = 1.0 / r;
            size_t m;
            double p;
            o_spline_pm(r, dr_inv, e->nr, &m, &p);
            const double* c = ORHO(e, ej, ei)[m];
            double rho_i = hval(c, p), drho_i = hder(c, p, dr_inv);
            c = ORHO(e, ei, ej)[m];
            double rho_j = hval(c, p), drho_j = hder(c, p, dr_inv);
            c = OPHI(e, ei, ej)[m];
            double z2p = hder(c, p, dr_inv), z2 = hval(c, p);
            double phi = z2 * r_inv;
            double dphi = z2p * r_inv - phi * r_inv;
            rho[i] += rho_i;
            if (j < n) rho[j] += rho_j;
            double fp = -dphi * r_inv;
            f[3 * i] += fp * dx0; f[3 * i + 1] += fp * dx1; f[3 * i + 2] += fp * dx2;
            if (j < n) { f[3 * j] -= fp * dx0; f[3 * j + 1] -= fp * dx1; f[3 * j + 2] -= fp * dx2; }
            else { fp *= 0.5; phi *= 0.5; }
            v7[0] += phi;
            v7[1] -= fp * dx0 * dx0; v7[2] -= fp * dx0 * dx1; v7[3] -= fp * dx0 * dx2;
            v7[4] -= fp * dx1 * dx1; v7[5] -= fp * dx1 * dx2; v7[6] -= fp * dx2 * dx2;
            md->drhoi[k] = -drho_i * r_inv;
            md->drhoj[k] = -drho_j * r_inv;
        }
        size_t m;
        double p;
        o_spline_pm(rho[i], drho_inv, e->nrho, &m, &p);
        const double* c = OF(e, ei) + m * e->nf;
        double t1 = hder(c, p, drho_inv);
        double t0 = hval(c, p);
        if (rho[i] > rho_max) t0 += t1 * (rho[i] - rho_max);
        v7[0] += t0;
        dF[i] = t1;
    }
    /* Update::update_wo_x(dF): ghosts take their owner's F' (src/exchange.h:1074-1096) */
    for (int k = n; k < md->g.nall; k++) dF[k] = dF[md->g.src[k]];
    for (int i = 0; i < n; i++) {
        for (int64_t k = md->row[i]; k < md->row[i + 1]; k++) {
            if (md->drhoi[k] == 0.0 && md->drhoj[k] == 0.0) continue;
            int j = md->col[k];
            double fp = dF[i] * md->drhoi[k] + dF[j] * md->drhoj[k];
            double dx0 = x[3 * i] - x[3 * j];
            double dx1 = x[3 * i + 1] - x[3 * j + 1];
            double dx2 = x[3 * i + 2] - x[3 * j + 2];
            f[3 * i] += dx0 * fp; f[3 * i + 1] += dx1 * fp; f[3 * i + 2] += dx2 * fp;
            if (j < n) { f[3 * j] -= dx0 * fp; f[3 * j + 1] -= dx1 * fp; f[3 * j + 2] -= dx2 * fp; }
            else fp *= 0.5;
            v7[1] -= fp * dx0 * dx0; v7[2] -= fp * dx0 * dx1; v7[3] -= fp * dx0 * dx2;
            v7[4] -= fp * dx1 * dx1; v7[5] -= fp * dx1 * dx2; v7[6] -= fp * dx2 * dx2;
        }
    }
    if (rho_out) memcpy(rho_out, rho, sizeof(double) * n);
}

/* eam_force on o_threads threads: the same per-pair arithmetic over the same
 * half list; rho, f and the sums go to per-thread rows that are added at the
 * end (so F(rho_i) follows the complete density of every atom) */
static void eam_force_threaded(o_md* md, double* f, double* v7, double* rho_out)
{
    const o_eam* e = md->ff;
    const int ne = e->ne, n = md->n, T = o_threads;
    const double* x = md->g.x;
    const uint8_t* el = md->elem;
    const double dr_inv = 1.0 / e->dr, drho_inv = 1.0 / e->drho;
    const double rho_max = (double)e->nrho * e->drho;
    if (md->nscratch < md->npairs || !md->rho) {
        free(md->drhoi); free(md->drhoj);
        md->nscratch = md->npairs > 0 ? md->npairs : 1;
        md->drhoi = (double*)malloc(sizeof(double) * md->nscratch);
        md->drhoj = (double*)malloc(sizeof(double) * md->nscratch);
    }
    md->rho = (double*)realloc(md->rho, sizeof(double) * (n > 0 ? n : 1));
    md->dF = (double*)realloc(md->dF, sizeof(double) * (md->g.nall > 0 ? md->g.nall : 1));
    double* rho = md->rho;
    double* dF = md->dF;
    const size_t w = (size_t)(n > 0 ? n : 1);
    if (md->thr_cap < (size_t)T * 4 * w) {
        free(md->thr_rows);
        md->thr_cap = (size_t)T * 4 * w;
        md->thr_rows = (double*)malloc(sizeof(double) * md->thr_cap);
    }
    double* rho_t = md->thr_rows;
    double* f_t = md->thr_rows + (size_t)T * w;
    double v_t[16 * 8];
    memset(v_t, 0, sizeof v_t);
#pragma omp parallel num_threads(T)
    {
#ifdef _OPENMP
        const int t = omp_get_thread_num();
#else
        const int t = 0;
#endif
        double* rt = rho_t + (size_t)t * w;
        double* ft = f_t + (size_t)t * 3 * w;
        double* vt = v_t + (size_t)t * 8;
        memset(rt, 0, sizeof(double) * w);
        memset(ft, 0, sizeof(double) * 3 * w);
#pragma omp barrier
#pragma omp for schedule(dynamic, 64)
        for (int i = 0; i < n; i++) {
            const int ei = el[i];
            for (int64_t k = md->row[i]; k < md->row[i + 1]; k++) {
                const int j = md->col[k];
                const int ej = el[j];
                const double dx0 = x[3 * i] - x[3 * j];
                const double dx1 = x[3 * i + 1] - x[3 * j + 1];
                const double dx2 = x[3 * i + 2] - x[3 * j + 2];
                const double rsq = dx0 * dx0 + dx1 * dx1 + dx2 * dx2;
                md->drhoi[k] = md->drhoj[k] = 0.0;
                if (rsq >= md->cut_sq[ei * ne + ej]) continue;
                const double r = sqrt(rsq);
                const double r_inv = 1.0 / r;
                size_t m;
                double p;
                o_spline_pm(r, dr_inv, e->nr, &m, &p);
                const double* c = ORHO(e, ej, ei)[m];
                const double rho_i = hval(c, p), drho_i = hder(c, p, dr_inv);
                c = ORHO(e, ei, ej)[m];
                const double rho_j = hval(c, p), drho_j = hder(c, p, dr_inv);
                c = OPHI(e, ei, ej)[m];
                const double z2p = hder(c, p, dr_inv), z2 = hval(c, p);
                double phi = z2 * r_inv;
                const double dphi = z2p * r_inv - phi * r_inv;
                rt[i] += rho_i;
                if (j < n) rt[j] += rho_j;
                double fp = -dphi * r_inv;
                ft[3 * i] += fp * dx0; ft[3 * i + 1] += fp * dx1; ft[3 * i + 2] += fp * dx2;
                if (j < n) { ft[3 * j] -= fp * dx0; ft[3 * j + 1] -= fp * dx1; ft[3 * j + 2] -= fp * dx2; }
                else { fp *= 0.5; phi *= 0.5; }
                vt[0] += phi;
                vt[1] -= fp * dx0 * dx0; vt[2] -= fp * dx0 * dx1; vt[3] -= fp * dx0 * dx2;
                vt[4] -= fp * dx1 * dx1; vt[5] -= fp * dx1 * dx2; vt[6] -= fp * dx2 * dx2;
                md->drhoi[k] = -drho_i * r_inv;
                md->drhoj[k] = -drho_j * r_inv;
            }
        }
        /* complete densities, then F(rho_i) and F' */
#pragma omp for schedule(static)
        for (int i = 0; i < n; i++) {
            double sum = 0.0;
            for (int q = 0; q < T; q++) sum += rho_t[(size_t)q * w + i];
            rho[i] = sum;
            size_t m;
            double p;
            o_spline_pm(sum, drho_inv, e->nrho, &m, &p);
            const double* c = OF(e, el[i]) + m * e->nf;
            const double t1 = hder(c, p, drho_inv);
            double t0 = hval(c, p);
            if (sum > rho_max) t0 += t1 * (sum - rho_max);
            vt[7] += t0;
            dF[i] = t1;
        }
        /* a ghost's source may be another ghost: take the local owner's F' */
#pragma omp for schedule(static)
        for (int k = n; k < md->g.nall; k++) dF[k] = dF[md->g.owner[k]];
#pragma omp for schedule(dynamic, 64)
        for (int i = 0; i < n; i++) {
            for (int64_t k = md->row[i]; k < md->row[i + 1]; k++) {
                if (md->drhoi[k] == 0.0 && md->drhoj[k] == 0.0) continue;
                const int j = md->col[k];
                double fp = dF[i] * md->drhoi[k] + dF[j] * md->drhoj[k];
                const double dx0 = x[3 * i] - x[3 * j];
                const double dx1 = x[3 * i + 1] - x[3 * j + 1];
                const double dx2 = x[3 * i + 2] - x[3 * j + 2];
                ft[3 * i] += dx0 * fp; ft[3 * i + 1] += dx1 * fp; ft[3 * i + 2] += dx2 * fp;
                if (j < n) { ft[3 * j] -= dx0 * fp; ft[3 * j + 1] -= dx1 * fp; ft[3 * j + 2] -= dx2 * fp; }
                else fp *= 0.5;
                vt[1] -= fp * dx0 * dx0; vt[2] -= fp * dx0 * dx1; vt[3] -= fp * dx0 * dx2;
                vt[4] -= fp * dx1 * dx1; vt[5] -= fp * dx1 * dx2; vt[6] -= fp * dx2 * dx2;
            }
        }
#pragma omp for schedule(static)
        for (int i = 0; i < 3 * n; i++) {
            double sum = 0.0;
            for (int q = 0; q < T; q++) sum += f_t[(size_t)q * 3 * w + i];
            f[i] = sum;
        }
    }
    for (int k = 0; k < 7; k++) v7[k] = 0.0;
    for (int q = 0; q < T; q++) {
        for (int k = 0; k < 7; k++) v7[k] += v_t[8 * q + k];
        v7[0] += v_t[8 * q + 7];
    }
    if (rho_out) memcpy(rho_out, rho, sizeof(double) * n);
}

/* ForceFieldMD::force_calc (src/ff_md.cpp:125-144): out7[1..6] = -W (F_H) */
void o_md_force(o_md* md, double* f, double* out7, double* rho_out)
{
    double v7[7];
    if (o_threads > 1)
        eam_force_threaded(md, f, v7, rho_out);
    else
        eam_force(md, f, v7, rho_out);
    out7[0] = v7[0];
    for (int k = 1; k < 7; k++) out7[k] = -v7[k];
}

/* ForceFieldEAM::__energy_calc (src/ff_eam.cpp:244-318) */
double o_md_energy(o_md* md)
{
    const o_eam* e = md->ff;
    int ne = e->ne, n = md->n;
    const double* x = md->g.x;
    const uint8_t* el = md->elem;
    double dr_inv = 1.0 / e->dr, drho_inv = 1.0 / e->drho;
    double rho_max = (double)e->nrho * e->drho;
    double* rho = (double*)calloc(n > 0 ? n : 1, sizeof(double));
    double en = 0.0;
    for (int i = 0; i < n; i++) {
        int ei = el[i];
        for (int64_t k = md->row[i]; k < md->row[i + 1]; k++) {
            int j = md->col[k];
            int ej = el[j];
            double dx0 = x[3 * i] - x[3 * j];
            double dx1 = x[3 * i + 1] - x[3 * j + 1];
            double dx2 = x[3 * i + 2] - x[3 * j + 2];
            double rsq = dx0 * dx0 + dx1 * dx1 + dx2 * dx2;
            if (rsq < md->cut_sq[ei * ne + ej]) {
                double r = sqrt(rsq);
                size_t m;
                double p;
                o_spline_pm(r, dr_inv, e->nr, &m, &p);
                double phi = hval(OPHI(e, ei, ej)[m], p) / r;
                rho[i] += hval(ORHO(e, ej, ei)[m], p);
                if (j < n) {
                    rho[j] += hval(ORHO(e, ei, ej)[m], p);
                    en += phi;
                } else
                    en += 0.5 * phi;
            }
        }
        size_t m;
        double p;
        o_spline_pm(rho[i], drho_inv, e->nrho, &m, &p);
        const double* c = OF(e, ei) + m * e->nf;
        double t0 = hval(c, p);
        if (rho[i] > rho_max) t0 += hder(c, p, drho_inv) * (rho[i] - rho_max);
        en += t0;
    }
    free(rho);
    return en;
}

/* ---------------- NVT (src/md_nvt.cpp, src/thermostat.cpp:47-143) ---------------- */
/* Test hook: a nonzero value plants a known mistake in the thermostat so that
 * the tests can show that their checks catch it. 1 swaps ddt and ddt2; 2 flips
 * the sign of the coupling -xi_{j+1} xi_j in the forward sweep; 3 drives xi_2
 * by N_f xi_1^2 in the reverse sweep as well (the textbook chain, where the
 * reference's reverse sweep has xi_1^2). 0 is the reference's thermostat. */
int o_nhc_mutation = 0;

typedef struct { double ddt, ddt2, freq_sq; double eta[3]; int L, niters; } o_nhc;

/* ThermostatNHC(dt, t_relax, L = 3, niters) (src/thermostat.cpp:19-36) */
static void nhc_init(o_nhc* t, double dt, double t_relax, int niters)
{
    t->L = 3;
    t->niters = niters;
    t->ddt = dt / (double)niters;
    t->ddt2 = 0.5 * dt / (double)niters;
    if (o_nhc_mutation == 1) {
        double tmp = t->ddt;
        t->ddt = t->ddt2;
        t->ddt2 = tmp;
    }
    t->freq_sq = 1.0 / (t_relax * t_relax);
    t->eta[0] = t->eta[1] = t->eta[2] = 0.0;
}

static double nhc_f(double x)
{
    if (fabs(x) < 0.015625) return 1.0 + x / 2.0 + x * x / 6.0 + x * x * x / 24.0;
    return (exp(x) - 1.0) / x;
}

/* ThermostatNHC::operator() (src/thermostat.cpp:47-104), L > 1 branch.
 * mid (may be NULL) receives the chain after the reverse sweep of the last
 * iteration, the state at which the velocities are scaled. */
static double nhc_step(o_nhc* t, double T_r, int ndof, double* mid)
{
    double rescale = 1.0, prev, tmp;
    int L = t->L;
    double* e = t->eta;
    double sgn = o_nhc_mutation == 2 ? -1.0 : 1.0;
    double rev0 = o_nhc_mutation == 3 ? (double)ndof : 1.0;
    for (int it = 0; it < t->niters; it++) {
        e[L - 1] += t->ddt2 * (e[L - 2] * e[L - 2] - t->freq_sq);
        prev = -e[L - 1];
        for (int ich = L - 2; ich > 0; ich--) {
            double drive = (ich == 1 ? rev0 : 1.0) * e[ich - 1] * e[ich - 1];
            e[ich] += t->ddt2 * nhc_f(prev * t->ddt2) * (prev * e[ich] + drive - t->freq_sq);
            prev = -e[ich];
        }
        e[0] += t->ddt2 * nhc_f(prev * t->ddt2) * (prev * e[0] + t->freq_sq * (T_r - 1.0));
        prev = -e[0];
        if (mid) memcpy(mid, e, sizeof(double) * 3);
        tmp = exp(t->ddt * prev);
        rescale *= tmp;
        T_r *= tmp * tmp;
        prev = (T_r - 1.0) * t->freq_sq;
        e[0] += t->ddt2 * nhc_f(-e[1] * t->ddt2) * (-sgn * e[1] * e[0] + prev);
        prev = ndof * e[0] * e[0] - t->freq_sq;
        for (int ich = 1; ich < L - 1; ich++) {
            e[ich] += t->ddt2 * nhc_f(-e[ich + 1] * t->ddt2) * (-sgn * e[ich + 1] * e[ich] + prev);
            prev = e[ich] * e[ich] - t->freq_sq;
        }
        e[L - 1] += t->ddt2 * prev;
    }
    return rescale;
}

double o_nhc_run(double T_r, int ndof, double dt, double t_relax, int niters, int ncalls, double* hist)
{
    o_nhc th;
    nhc_init(&th, dt, t_relax, niters);
    double scale = 1.0;
    for (int k = 0; k < ncalls; k++) {
        double r = nhc_step(&th, T_r, ndof, NULL);
        scale *= r;
        T_r *= r * r;
        if (hist) {
            hist[4 * k] = T_r;
            memcpy(hist + 4 * k + 1, th.eta, sizeof(double) * 3);
        }
    }
    return scale;
}

static void kin_tensor(const o_md* md, const double* v, const double* masses, double* mvv)
{
    for (int k = 0; k < 6; k++) mvv[k] = 0.0;
    for (int i = 0; i < md->n; i++) {
        double m = masses[md->elem_lcl[i]];
        const double* u = v + 3 * i;
        mvv[0] += u[0] * u[0] * m; mvv[1] += u[1] * u[0] * m; mvv[2] += u[2] * u[0] * m;
        mvv[3] += u[1] * u[1] * m; mvv[4] += u[2] * u[1] * m; mvv[5] += u[2] * u[2] * m;
    }
}

/* the thermostat's integral of xi_j over one call spanning tcall (one
 * iteration): the call scales the velocities by exp(-tcall xi_1) at its mid
 * state, so it adds tcall * xi(mid) */
static void eta_accumulate(double* eta_int, double tcall, const double* mid)
{
    for (int k = 0; k < 3; k++) eta_int[k] += tcall * mid[k];
}

int o_md_nvt_run_hist(o_md* md, double* v, const double* masses, double kB, double T,
                      double dt, int nsteps, double* out7, double* T_part_out, double* hist)
{
    int n = md->n;
    double dt2 = dt / 2.0;
    int ndof = n * 3;
    o_nhc th;
    /* ThermostatNHC(dt/2, 100 dt, 3, 1) (src/md_nvt.cpp:19) */
    nhc_init(&th, dt / 2.0, 100.0 * dt, 1);
    double mid[3], eta_int[3] = {0.0, 0.0, 0.0};
    double* f = (double*)malloc(sizeof(double) * 3 * (n > 0 ? n : 1));
    double mvv[6];
    kin_tensor(md, v, masses, mvv);
    double T_part = (mvv[0] + (mvv[3] + mvv[5])) / (ndof * kB);
    o_md_force(md, f, out7, NULL);
    int nrebuild = 0;
    double fac, fac_x_d = 1.0;
    for (int step = 0; step < nsteps; step++) {
        fac_x_d *= fac = nhc_step(&th, T_part / T, ndof, mid);
        eta_accumulate(eta_int, dt2, mid);
        fac *= fac;
        for (int k = 0; k < 6; k++) mvv[k] *= fac;
        T_part *= fac;
        /* update_x_d__x (src/md_nvt.cpp:70-102) */
        double dx_l[3] = {0.0, 0.0, 0.0};
        double* x = md->g.x;
        for (int i = 0; i < n; i++) {
            double m = masses[md->elem_lcl[i]];
            for (int d = 0; d < 3; d++) {
                v[3 * i + d] = v[3 * i + d] * fac_x_d + f[3 * i + d] * dt2 / m;
                dx_l[d] += v[3 * i + d] * dt;
                x[3 * i + d] += v[3 * i + d] * dt;
            }
        }
        for (int d = 0; d < 3; d++) dx_l[d] /= (double)n;
        for (int i = 0; i < n; i++)
            for (int d = 0; d < 3; d++) x[3 * i + d] -= dx_l[d];
        nrebuild += o_md_update(md);
        o_md_force(md, f, out7, NULL);
        /* update_x_d (src/md_nvt.cpp:106-128) */
        for (int i = 0; i < n; i++) {
            double m = masses[md->elem_lcl[i]];
            for (int d = 0; d < 3; d++) v[3 * i + d] += f[3 * i + d] * dt2 / m;
        }
        kin_tensor(md, v, masses, mvv);
        T_part = (mvv[0] + (mvv[3] + mvv[5])) / (ndof * kB);
        fac_x_d = fac = nhc_step(&th, T_part / T, ndof, mid);
        eta_accumulate(eta_int, dt2, mid);
        fac *= fac;
        for (int k = 0; k < 6; k++) mvv[k] *= fac;
        T_part *= fac;
        if (hist) {
            double* h = hist + 8 * step;
            h[0] = T_part;
            h[1] = out7[0];
            memcpy(h + 2, th.eta, sizeof(double) * 3);
            memcpy(h + 5, eta_int, sizeof(double) * 3);
        }
    }
    for (int i = 0; i < 3 * n; i++) v[i] *= fac_x_d;
    if (T_part_out) *T_part_out = T_part;
    free(f);
    return nrebuild;
}

int o_md_nvt_run(o_md* md, double* v, const double* masses, double kB, double T,
                 double dt, int nsteps, double* out7, double* T_part_out)
{
    return o_md_nvt_run_hist(md, v, masses, kB, T, dt, nsteps, out7, T_part_out, NULL);
}

/* MDMuVT::run (src/md_muvt.cpp:122-280), the branch without fixed degrees of
 * freedom: MDNVT's step (src/md_nvt.cpp:70-128) where, after every nevery-th
 * first half step, gcmc.xchng(false, nattempts) takes the place of the skin
 * check (src/md_muvt.cpp:172-195). At one process PGCMC draws, decides and
 * orders the atoms exactly as SGCMC does (one concurrent attempt per round,
 * the same Random(seed) stream, s drawn in [0, 1), the same acceptance
 * expressions and add/del order: src/pgcmc.cpp:562-735,1163-1266 against
 * src/sgcmc.cpp:251-444), so the exchange runs the oracle's SGCMC chain,
 * created once per run (gcmc.init) and reloaded with the current atoms at
 * each exchange (x2s with the wrap, init_xchng). fin_xchng rebuilds the list
 * from the wrapped positions; ndof_part = 3 natms follows the count. The
 * exchange's update of mvv and T_part (src/md_muvt.cpp:192-194) is
 * overwritten by update_x_d before anything reads it, so it is not kept.
 * *pmd is replaced by the system after each exchange; v holds cap atoms.
 * counts: accepted insertions, accepted deletions, insertion trials, deletion
 * trials. hist [nsteps][3]: T_part, potential energy, gas atoms. Returns the
 * list builds (rebuilds and exchanges), or -1 when an insertion would exceed
 * cap. */
int o_md_muvt_run(o_md** pmd, double* v, int cap, const double* masses, double kB, double hP, double T, double dt,
                  double mu, int gas, int seed, int nevery, int nattempts, int nsteps, double* out7,
                  double* T_part_out, int* natms_out, int* counts, double* hist)
{
    o_md* md = *pmd;
    const o_eam* ff = md->ff;
    int n = md->n;
    double dt2 = dt / 2.0;
    int ndof = n * 3;
    double skin = md->skin;
    double H9[9];
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) H9[3 * i + j] = md->bx.H[i][j];
    o_nhc th;
    nhc_init(&th, dt / 2.0, 100.0 * dt, 1);
    size_t cd = (size_t)(cap > n ? cap : n);
    double* f = (double*)malloc(sizeof(double) * 3 * cd);
    double* xs = (double*)malloc(sizeof(double) * 3 * cd);
    double* ss = (double*)malloc(sizeof(double) * 3 * cd);
    double* rs = (double*)malloc(sizeof(double) * cd);
    double* Fs = (double*)malloc(sizeof(double) * cd);
    uint8_t* es = (uint8_t*)malloc(cd);
    int* dec = (int*)malloc(sizeof(int) * (nattempts > 0 ? nattempts : 1));
    double* du = (double*)malloc(sizeof(double) * (nattempts > 0 ? nattempts : 1));
    for (int k = 0; k < 4; k++) counts[k] = 0;
    /* PGCMC gcmc(...); gcmc.init() at the start of run */
    o_gcmc* g = o_gcmc_create(ff, n, H9, md->g.x, md->elem_lcl, gas, mu, T, kB, hP, masses, seed, (int)cd);
    double mvv[6];
    kin_tensor(md, v, masses, mvv);
    double T_part = (mvv[0] + (mvv[3] + mvv[5])) / (ndof * kB);
    o_md_force(md, f, out7, NULL);
    int nbuild = 0, status = 0;
    double fac, fac_x_d = 1.0;
    for (int step = 0; step < nsteps; step++) {
        fac_x_d *= fac = nhc_step(&th, T_part / T, ndof, NULL);
        fac *= fac;
        for (int k = 0; k < 6; k++) mvv[k] *= fac;
        T_part *= fac;
        double dx_l[3] = {0.0, 0.0, 0.0};
        double* x = md->g.x;
        for (int i = 0; i < n; i++) {
            double m = masses[md->elem_lcl[i]];
            for (int d = 0; d < 3; d++) {
                v[3 * i + d] = v[3 * i + d] * fac_x_d + f[3 * i + d] * dt2 / m;
                dx_l[d] += v[3 * i + d] * dt;
                x[3 * i + d] += v[3 * i + d] * dt;
            }
        }
        for (int d = 0; d < 3; d++) dx_l[d] /= (double)n;
        for (int i = 0; i < n; i++)
            for (int d = 0; d < 3; d++) x[3 * i + d] -= dx_l[d];
        if (nevery <= 0 || (step + 1) % nevery) {
            nbuild += o_md_update(md);
        } else {
            o_gcmc_reload(g, n, md->g.x, md->elem_lcl, v);
            if (o_gcmc_attempts(g, nattempts, dec, du) != 0) {
                status = -1;
                break;
            }
            for (int k = 0; k < nattempts; k++) {
                if (dec[k] == 1) counts[0]++;
                if (dec[k] == 2) counts[1]++;
                if (dec[k] == 1 || dec[k] == -1) counts[2]++; else counts[3]++;
            }
            n = o_gcmc_natms(g);
            o_gcmc_state(g, xs, ss, es, rs, Fs, v);
            ndof = 3 * n;
            o_md_free(md);
            md = o_md_create(ff, n, H9, xs, es, skin);
            *pmd = md;
            nbuild++;
        }
        o_md_force(md, f, out7, NULL);
        for (int i = 0; i < n; i++) {
            double m = masses[md->elem_lcl[i]];
            for (int d = 0; d < 3; d++) v[3 * i + d] += f[3 * i + d] * dt2 / m;
        }
        kin_tensor(md, v, masses, mvv);
        T_part = (mvv[0] + (mvv[3] + mvv[5])) / (ndof * kB);
        fac_x_d = fac = nhc_step(&th, T_part / T, ndof, NULL);
        fac *= fac;
        for (int k = 0; k < 6; k++) mvv[k] *= fac;
        T_part *= fac;
        if (hist) {
            int ng = 0;
            for (int i = 0; i < n; i++) ng += md->elem_lcl[i] == gas;
            hist[3 * step] = T_part;
            hist[3 * step + 1] = out7[0];
            hist[3 * step + 2] = (double)ng;
        }
    }
    if (status == 0)
        for (int i = 0; i < 3 * n; i++) v[i] *= fac_x_d;
    if (T_part_out) *T_part_out = T_part;
    if (natms_out) *natms_out = n;
    o_gcmc_free(g);
    free(f); free(xs); free(ss); free(rs); free(Fs); free(es); free(dec); free(du);
    return status == 0 ? nbuild : -1;
}

/* AtomsMD::create_T (src/atoms_md.cpp:130-160), one rank (seed + 0) */
void o_create_T(int natms, const uint8_t* elem, const double* masses, double kB,
                double T, int seed, double* v)
{
    o_rng r;
    o_rng_init(&r, seed);
    for (int i = 0; i < natms; i++) {
        double fac = sqrt(kB * T / masses[elem[i]]);
        for (int d = 0; d < 3; d++) v[3 * i + d] = o_rng_gaussian(&r) * fac;
    }
}

void o_md_view(const o_md* md, const double** x, const double** s, const uint8_t** elem,
               const int64_t** row, const int32_t** col)
{
    *x = md->g.x;
    *s = md->g.s;
    *elem = md->elem;
    *row = md->row;
    *col = md->col;
}

/* ---------------- NST (src/md_nst.cpp) ---------------- */

/* MDMath::f, df, ddf (src/md_nst.cpp:843-871): phi(x) = (e^x - 1) / x and its
 * first two derivatives, by their Taylor series below |x| = 1/64 */
static double mdm_f(double x)
{
    if (fabs(x) < 0.015625) return 1.0 + x / 2.0 + x * x / 6.0 + x * x * x / 24.0;
    return (exp(x) - 1.0) / x;
}
static double mdm_df(double x)
{
    if (fabs(x) < 0.015625) return 0.5 + x / 3.0 + x * x / 8.0 + x * x * x / 30.0;
    double e = exp(x);
    return (e - (e - 1.0) / x) / x;
}
static double mdm_ddf(double x)
{
    if (fabs(x) < 0.015625) return 1.0 / 3.0 + x / 4.0 + x * x / 10.0 + x * x * x / 36.0;
    double e = exp(x);
    return (e - 2.0 * ((e - (e - 1.0) / x) / x)) / x;
}

static const double MDM_SQRT_EPS = 1.4901161193847656e-08; /* sqrt(DBL_EPSILON) */

/* first and second divided differences of exp and of phi, switching to the
 * derivative at the midpoint when the nodes are closer than the rounding
 * allows (src/md_nst.cpp:890-901, 960-986) */
static double dd_exp(double x, double y, double fx, double fy)
{
    if (fabs(x - y) < MDM_SQRT_EPS) return exp(0.5 * (x + y));
    return (fx - fy) / (x - y);
}
static double dd_phi(double x, double y, double fx, double fy)
{
    double m = 0.5 * (x + y);
    if (fabs(x - y) < sqrt(2.220446049250313e-16 * mdm_f(m) / mdm_ddf(m))) return mdm_df(m);
    return (fx - fy) / (x - y);
}
static double dd2_phi(double x, double y, double z, double f_yz, double f_zx, double f_xy)
{
    double m = (x + y + z) / 3.0, d2 = mdm_ddf(m);
    double lim = sqrt(2.220446049250313e-16 * mdm_f(m) / d2);
    if (fabs(x - z) > lim) return (f_xy - f_yz) / (x - z);
    if (fabs(y - x) > lim) return (f_yz - f_zx) / (y - x);
    if (fabs(z - y) > lim) return (f_zx - f_xy) / (z - y);
    return d2;
}
static double dd2_exp(double x, double y, double z, double g_yz, double g_zx, double g_xy)
{
    if (fabs(x - z) > MDM_SQRT_EPS) return (g_xy - g_yz) / (x - z);
    if (fabs(y - x) > MDM_SQRT_EPS) return (g_yz - g_zx) / (y - x);
    if (fabs(z - y) > MDM_SQRT_EPS) return (g_zx - g_xy) / (z - y);
    return exp((x + y + z) / 3.0);
}

/* MDMath::calc (src/md_nst.cpp:887-1002) for a lower-triangular A (row-major
 * [9]): E = exp(Z0), F0 = dt2 phi(Z0), F1 = dt phi(Z1) with Z0 = -dt2 (A + a I)
 * and Z1 = dt A, each from the diagonal values and the divided differences of
 * exp and phi (the functions of a triangular matrix, Opitz) */
void o_nst_calc(double a, double dt, double dt2, const double* A, double* E, double* F0, double* F1)
{
    double z0[3], z1[3], f0[3], f1[3], e0[3];
    for (int i = 0; i < 3; i++) {
        z0[i] = -dt2 * (A[4 * i] + a);
        z1[i] = dt * A[4 * i];
        f0[i] = mdm_f(z0[i]);
        f1[i] = mdm_f(z1[i]);
        e0[i] = exp(z0[i]);
    }
    for (int k = 0; k < 9; k++) E[k] = F0[k] = F1[k] = 0.0;
    /* pairs (1,0), (2,0), (2,1) */
    const int pi[3] = {1, 2, 2}, pj[3] = {0, 0, 1};
    double d0[3] = {0, 0, 0}, d1[3] = {0, 0, 0}, de[3] = {0, 0, 0};
    for (int p = 0; p < 3; p++) {
        int i = pi[p], j = pj[p];
        if (A[3 * i + j] == 0.0) continue;
        d0[p] = dd_phi(z0[i], z0[j], f0[i], f0[j]);
        d1[p] = dd_phi(z1[i], z1[j], f1[i], f1[j]);
        de[p] = dd_exp(z0[i], z0[j], e0[i], e0[j]);
    }
    for (int i = 0; i < 3; i++) {
        F0[4 * i] = f0[i] * dt2;
        F1[4 * i] = f1[i] * dt;
        E[4 * i] = e0[i];
    }
    for (int p = 0; p < 3; p++) {
        int i = pi[p], j = pj[p];
        double aij = A[3 * i + j];
        F0[3 * i + j] = -aij * dt2 * dt2 * d0[p];
        F1[3 * i + j] = aij * dt * dt * d1[p];
        E[3 * i + j] = -aij * dt2 * de[p];
    }
    double a10 = A[3], a21 = A[7];
    if (a10 == 0.0 || a21 == 0.0) return;
    /* second order term of entry (2,0): z21 z10 [x, y, z] */
    F0[6] += dt2 * dt2 * dt2 * a21 * a10 * dd2_phi(z0[0], z0[1], z0[2], d0[2], d0[1], d0[0]);
    F1[6] += dt * dt * dt * a21 * a10 * dd2_phi(z1[0], z1[1], z1[2], d1[2], d1[1], d1[0]);
    E[6] += dt2 * dt2 * a21 * a10 * dd2_exp(z0[0], z0[1], z0[2], de[2], de[1], de[0]);
}

/* row vector times lower-triangular matrix: out_j = sum_{i >= j} v_i M_ij */
static void v_mlt(const double* v, const double* M, double* out)
{
    double o0 = v[0] * M[0] + v[1] * M[3] + v[2] * M[6];
    double o1 = v[1] * M[4] + v[2] * M[7];
    double o2 = v[2] * M[8];
    out[0] = o0; out[1] = o1; out[2] = o2;
}

static int lt_voigt(int i, int j) { return i + j * 3 - j * (j + 1) / 2; }

/* 1 flips the sign of the stress term of the box kick; 0 is the restatement.
 * The mutant exists to show that the physical checks of the NST fail on it. */
int o_nst_mutation = 0;

typedef struct {
    double T, kB, dt, dt2, ndof;
    int sdof[9];
    double S_dev[9], prefac[9], tau[9], VH[9], B_ref[9];
    double s_hyd, vol_ref, T_baro;
    int ndof_baro;
} nst_state;

/* MDNST::update_V_H (src/md_nst.cpp:320-351): the lower triangle of
 * L^T S_dev L with L = B_ref H and S_dev as pre_init leaves it (its lower
 * triangle; the upper stays zero), then the box-velocity kicks */
static void nst_update_VH(nst_state* s, const o_box* bx, const double* S_pe, const double* mvv, double T_part)
{
    double L[9] = {0}, SL[9] = {0}, M2[9] = {0};
    for (int i = 0; i < 3; i++)
        for (int j = 0; j <= i; j++) {
            double v = 0.0;
            for (int k = j; k <= i; k++) v += s->B_ref[3 * i + k] * bx->H[k][j];
            L[3 * i + j] = v;
        }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j < 3; j++) {
            double v = 0.0;
            for (int k = 0; k < 3; k++) v += s->S_dev[3 * i + k] * L[3 * k + j];
            SL[3 * i + j] = v;
        }
    for (int i = 0; i < 3; i++)
        for (int j = 0; j <= i; j++) {
            double v = 0.0;
            for (int k = 0; k < 3; k++) v += L[3 * k + i] * SL[3 * k + j];
            M2[3 * i + j] = v;
        }
    for (int i = 0; i < 3; i++)
        if (s->sdof[4 * i]) s->VH[4 * i] += (s->kB * T_part + bx->vol * s->s_hyd) * s->prefac[4 * i] * s->dt2;
    s->T_baro = 0.0;
    for (int j = 0; j < 3; j++)
        for (int i = j; i < 3; i++) {
            int k = 3 * i + j;
            if (!s->sdof[k]) continue;
            const double sg = o_nst_mutation == 1 ? -1.0 : 1.0;
            s->VH[k] += (s->vol_ref * M2[k] - sg * bx->vol * S_pe[k] + mvv[lt_voigt(i, j)]) * s->prefac[k] * s->dt2;
            s->T_baro += s->VH[k] * s->VH[k] * s->tau[k] * s->tau[k];
        }
    s->T_baro *= s->ndof * s->T / s->ndof_baro;
}

/* S_pe (lower triangle, row-major [9]) from ForceFieldMD::force_calc's F_H:
 * S_pe = -F_H / vol (src/ff_md.cpp:134-136) */
static void nst_spe(const double* out7, double vol, double* S_pe)
{
    for (int k = 0; k < 9; k++) S_pe[k] = 0.0;
    for (int j = 0; j < 3; j++)
        for (int i = j; i < 3; i++) S_pe[3 * i + j] = -out7[1 + lt_voigt(i, j)] / vol;
}

/* MDNST::run (src/md_nst.cpp:376-460) with update_x_d__x__x_d
 * (src/md_nst.cpp:80-153) and pre_init (src/md_nst.cpp:291-319) on one rank:
 * S9 the target stress (row-major, lower triangle read, NaN = free
 * component), tau9 the relaxation times (NULL: 1000 dt each), thermostats
 * NHC(dt/2, 100 dt) on the particles and NHC(dt/2, 1000 dt) on the box, both
 * L = 3 and one iteration; every nreset steps (0: never) the reference box
 * B_ref and volume follow the current ones. H9 gets the final box; hist, when
 * given, [nsteps][14]: T_part, pe, H00, H10, H11, H20, H21, H22 and S_pe
 * (xx, yx, zx, yy, zy, zz) after each step. Returns the number of list
 * rebuilds, or -1 when the box left [vol0 / 4, 4 vol0] and the run stopped. */
int o_md_nst_run(o_md* md, double* v, const double* masses, double kB, double T, double dt, const double* S9,
                 const double* tau9, int nreset, int nsteps, double* out7, double* T_part_out, double* H9,
                 double* hist)
{
    int n = md->n;
    nst_state s;
    memset(&s, 0, sizeof s);
    s.T = T; s.kB = kB; s.dt = dt; s.dt2 = dt / 2.0; s.ndof = 3.0 * n;
    o_nhc tp, tb;
    nhc_init(&tp, dt / 2.0, 100.0 * dt, 1);
    nhc_init(&tb, dt / 2.0, 1000.0 * dt, 1);
    double mid[3];
    double mvv[6];
    kin_tensor(md, v, masses, mvv);
    double T_part = (mvv[0] + (mvv[3] + mvv[5])) / (s.ndof * kB);
    /* pre_init */
    double S[9] = {0};
    for (int j = 0; j < 3; j++)
        for (int i = j; i < 3; i++) {
            int k = 3 * i + j;
            s.tau[k] = tau9 ? tau9[k] : 1000.0 * dt;
            if (isnan(S9[k])) {
                s.sdof[k] = 0;
                S[k] = 0.0;
            } else {
                s.sdof[k] = 1;
                S[k] = S9[k];
                s.prefac[k] = 1.0 / (s.tau[k] * s.tau[k] * s.ndof * kB * T);
                s.ndof_baro++;
            }
            s.S_dev[k] = S[k];
        }
    s.T_baro = 0.0; /* V_H starts at zero */
    s.s_hyd = (S[0] + S[4] + S[8]) / 3.0;
    for (int i = 0; i < 3; i++) s.S_dev[4 * i] = S[4 * i] - s.s_hyd;
    /* run */
    double* f = (double*)malloc(sizeof(double) * 3 * (n > 0 ? n : 1));
    double S_pe[9];
    o_md_force(md, f, out7, NULL);
    nst_spe(out7, md->bx.vol, S_pe);
    for (int k = 0; k < 9; k++) s.B_ref[k] = md->bx.B[k / 3][k % 3];
    s.vol_ref = md->bx.vol;
    int nrebuild = 0;
    double fac, fac_x_d = 1.0;
    const double vol0 = md->bx.vol;
    for (int step = 0; step < nsteps; step++) {
        /* a box that has left [vol0 / 4, 4 vol0] (or is not finite) ends the
         * run: -1 is returned */
        if (!(md->bx.vol > 0.25 * vol0 && md->bx.vol < 4.0 * vol0)) {
            nrebuild = -1;
            break;
        }
        fac = nhc_step(&tb, s.T_baro / T, s.ndof_baro, mid);
        for (int k = 0; k < 9; k++) s.VH[k] *= fac;
        s.T_baro *= fac * fac;
        fac_x_d *= fac = nhc_step(&tp, T_part / T, (int)s.ndof, mid);
        fac *= fac;
        for (int k = 0; k < 6; k++) mvv[k] *= fac;
        T_part *= fac;
        nst_update_VH(&s, &md->bx, S_pe, mvv, T_part);
        /* update_x_d__x__x_d */
        {
            double E[9], F0[9], F1[9];
            o_nst_calc((s.VH[0] + s.VH[4] + s.VH[8]) / s.ndof, dt, s.dt2, s.VH, E, F0, F1);
            double* x = md->g.x;
            double dx[3] = {0.0, 0.0, 0.0};
            for (int i = 0; i < n; i++) {
                double m_inv = 1.0 / masses[md->elem_lcl[i]];
                double a[3], b[3], v0[3];
                v_mlt(v + 3 * i, E, a);
                v_mlt(f + 3 * i, F0, b);
                for (int d = 0; d < 3; d++) v[3 * i + d] = fac_x_d * a[d] + m_inv * b[d];
                v_mlt(x + 3 * i, s.VH, v0);
                for (int d = 0; d < 3; d++) v0[d] += v[3 * i + d];
                v_mlt(v0, F1, v0);
                for (int d = 0; d < 3; d++) {
                    x[3 * i + d] += v0[d];
                    dx[d] += v0[d];
                }
            }
            for (int d = 0; d < 3; d++) dx[d] /= (double)n;
            for (int i = 0; i < n; i++)
                for (int d = 0; d < 3; d++) x[3 * i + d] -= dx[d];
            /* H <- H (I + F1 V_H) */
            double M[9] = {0}, Hn[9] = {0};
            for (int i = 0; i < 3; i++)
                for (int j = 0; j <= i; j++) {
                    double t = 0.0;
                    for (int k = j; k <= i; k++) t += F1[3 * i + k] * s.VH[3 * k + j];
                    M[3 * i + j] = t + (i == j ? 1.0 : 0.0);
                }
            for (int i = 0; i < 3; i++)
                for (int j = 0; j <= i; j++) {
                    double t = 0.0;
                    for (int k = j; k <= i; k++) t += md->bx.H[i][k] * M[3 * k + j];
                    Hn[3 * i + j] = t;
                }
            o_box_set(&md->bx, Hn);
            nrebuild += o_md_update(md);
            o_md_force(md, f, out7, NULL);
            nst_spe(out7, md->bx.vol, S_pe);
            for (int i = 0; i < n; i++) {
                double m_inv = 1.0 / masses[md->elem_lcl[i]];
                double a[3], b[3];
                v_mlt(v + 3 * i, E, a);
                v_mlt(f + 3 * i, F0, b);
                for (int d = 0; d < 3; d++) v[3 * i + d] = a[d] + m_inv * b[d];
            }
            kin_tensor(md, v, masses, mvv);
            T_part = (mvv[0] + (mvv[3] + mvv[5])) / (s.ndof * kB);
        }
        nst_update_VH(&s, &md->bx, S_pe, mvv, T_part);
        fac_x_d = fac = nhc_step(&tp, T_part / T, (int)s.ndof, mid);
        fac *= fac;
        for (int k = 0; k < 6; k++) mvv[k] *= fac;
        T_part *= fac;
        fac = nhc_step(&tb, s.T_baro / T, s.ndof_baro, mid);
        for (int k = 0; k < 9; k++) s.VH[k] *= fac;
        s.T_baro *= fac * fac;
        if (nreset && (step + 1) % nreset == 0) {
            for (int k = 0; k < 9; k++) s.B_ref[k] = md->bx.B[k / 3][k % 3];
            s.vol_ref = md->bx.vol;
        }
        if (hist) {
            double* h = hist + 14 * step;
            h[0] = T_part;
            h[1] = out7[0];
            h[2] = md->bx.H[0][0]; h[3] = md->bx.H[1][0]; h[4] = md->bx.H[1][1];
            h[5] = md->bx.H[2][0]; h[6] = md->bx.H[2][1]; h[7] = md->bx.H[2][2];
            for (int j = 0; j < 3; j++)
                for (int i = j; i < 3; i++) h[8 + lt_voigt(i, j)] = S_pe[3 * i + j];
        }
    }
    for (int i = 0; i < 3 * n; i++) v[i] *= fac_x_d;
    if (T_part_out) *T_part_out = T_part;
    if (H9)
        for (int k = 0; k < 9; k++) H9[k] = md->bx.H[k / 3][k % 3];
    free(f);
    return nrebuild;
}
