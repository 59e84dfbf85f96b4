/*
 * o_tables.c -- oracle: spline coefficients, DYNAMO setfl/fs readers,
 * Gauss-Hermite nodes, Park-Miller generator.
 *
 * TEST INFRASTRUCTURE ONLY (the oracle): loaded by tests/, smoke() and the
 * cpu_baseline leg of bench.py; never by the product library.
 */
#include "o_internal.h"
#include <ctype.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* Cubic Hermite coefficients on a uniform grid; slopes from the 5-point
 * stencil, one-sided at the ends. Restates src/import_eam.cpp:62-83. */
void o_interpolate4(size_t n, double (*s)[4])
{
    s[0][1] = s[1][0] - s[0][0];
    s[1][1] = 0.5 * (s[2][0] - s[0][0]);
    s[n - 2][1] = 0.5 * (s[n - 1][0] - s[n - 3][0]);
    s[n - 1][1] = s[n - 1][0] - s[n - 2][0];
    for (size_t k = 2; k < n - 2; k++)
        s[k][1] = ((s[k - 2][0] - s[k + 2][0]) + 8.0 * (s[k + 1][0] - s[k - 1][0])) / 12.0;
    for (size_t k = 0; k + 1 < n; k++) {
        double d = s[k + 1][0] - s[k][0];
        s[k][2] = 3.0 * d - 2.0 * s[k][1] - s[k + 1][1];
        s[k][3] = s[k][1] + s[k + 1][1] - 2.0 * d;
    }
    s[n - 1][2] = 0.0;
    s[n - 1][3] = 0.0;
}

/* Quartic variant used for the DMD embedding function: same cubic part
 * plus c4 = (c2[k+1]-c2[k])/6 - c3[k]/2. Restates src/import_eam.cpp:87-115. */
void o_interpolate5(size_t n, double (*s)[5])
{
    s[0][1] = s[1][0] - s[0][0];
    s[1][1] = 0.5 * (s[2][0] - s[0][0]);
    s[n - 2][1] = 0.5 * (s[n - 1][0] - s[n - 3][0]);
    s[n - 1][1] = s[n - 1][0] - s[n - 2][0];
    for (size_t k = 2; k < n - 2; k++)
        s[k][1] = ((s[k - 2][0] - s[k + 2][0]) + 8.0 * (s[k + 1][0] - s[k - 1][0])) / 12.0;
    for (size_t k = 0; k + 1 < n; k++) {
        double d = s[k + 1][0] - s[k][0];
        s[k][2] = 3.0 * d - 2.0 * s[k][1] - s[k + 1][1];
        s[k][3] = s[k][1] + s[k + 1][1] - 2.0 * d;
    }
    s[n - 1][2] = 0.0;
    s[n - 1][3] = 0.0;
    for (size_t k = 0; k + 1 < n; k++)
        s[k][4] = (s[k + 1][2] - s[k][2]) / 6.0 - 0.5 * s[k][3];
    s[n - 1][4] = 0.0;
}

/* r_c[i][j] = (last knot index with a non-zero r*phi_ij, rho_ij or rho_ji) * dr.
 * Restates src/import_eam.h:52-64 (the header cutoff is ignored). */
static void set_cutoff(o_eam* e)
{
    int ne = e->ne;
    for (int i = 0; i < ne; i++)
        for (int j = 0; j <= i; j++) {
            size_t k = (size_t)e->nr - 1;
            while (OPHI(e, i, j)[k][0] == 0.0 && ORHO(e, i, j)[k][0] == 0.0 &&
                   ORHO(e, j, i)[k][0] == 0.0 && k >= 1)
                k--;
            e->rc[i * ne + j] = e->rc[j * ne + i] = (double)k * e->dr;
        }
}

static o_eam* alloc_eam(int ne, int nr, int nrho, double dr, double drho, int dmd)
{
    o_eam* e = (o_eam*)calloc(1, sizeof(o_eam));
    e->ne = ne; e->nr = nr; e->nrho = nrho; e->dr = dr; e->drho = drho; e->dmd = dmd;
    e->nf = dmd ? 5 : 4;
    e->r_phi = (double*)calloc((size_t)ne * ne * nr * 4, sizeof(double));
    e->rho = (double*)calloc((size_t)ne * ne * nr * 4, sizeof(double));
    e->F = (double*)calloc((size_t)ne * nrho * e->nf, sizeof(double));
    e->rc = (double*)calloc((size_t)ne * ne, sizeof(double));
    return e;
}

/* Splines every table, derives r_c, and for DMD multiplies the density
 * knots by r_k = k*dr and re-splines (src/ff_eam_dmd.cpp:80-94). */
static void finish_eam(o_eam* e)
{
    int ne = e->ne;
    for (int i = 0; i < ne; i++) {
        if (e->nf == 4) o_interpolate4((size_t)e->nrho, (double(*)[4])OF(e, i));
        else o_interpolate5((size_t)e->nrho, (double(*)[5])OF(e, i));
        for (int j = 0; j < ne; j++) o_interpolate4((size_t)e->nr, ORHO(e, i, j));
    }
    for (int i = 0; i < ne; i++)
        for (int j = 0; j < ne; j++) o_interpolate4((size_t)e->nr, OPHI(e, i, j));
    set_cutoff(e);
    if (e->dmd) {
        for (int i = 0; i < ne; i++)
            for (int j = 0; j < ne; j++) {
                double(*s)[4] = ORHO(e, i, j);
                for (int k = 0; k < e->nr; k++) s[k][0] *= (double)k * e->dr;
                o_interpolate4((size_t)e->nr, s);
            }
    }
}

o_eam* o_eam_from_knots(int ne, int nr, int nrho, double dr, double drho,
                        const double* F, const double* rho, const double* r_phi,
                        int setfl_alias, int dmd)
{
    o_eam* e = alloc_eam(ne, nr, nrho, dr, drho, dmd);
    for (int i = 0; i < ne; i++) {
        for (int k = 0; k < nrho; k++) OF(e, i)[k * e->nf] = F[(size_t)i * nrho + k];
        for (int j = 0; j < ne; j++) {
            int jj = setfl_alias ? 0 : j;
            for (int k = 0; k < nr; k++) {
                ORHO(e, i, j)[k][0] = rho[((size_t)i * ne + jj) * nr + k];
                OPHI(e, i, j)[k][0] = r_phi[((size_t)i * ne + j) * nr + k];
            }
        }
    }
    finish_eam(e);
    return e;
}

/* ---- DYNAMO reader: line/token handling of src/import.cpp:92-240 ---- */
typedef struct { char* buf; char* cur; int line_no; } reader;

static char* next_line(reader* r)
{
    if (!r->cur || !*r->cur) return NULL;
    char* s = r->cur;
    char* nl = strchr(s, '\n');
    if (nl) { *nl = '\0'; r->cur = nl + 1; } else r->cur = s + strlen(s);
    r->line_no++;
    return s;
}

static int tokens(char* line, char** args, int max)
{
    char* h = strchr(line, '#');
    if (h) *h = '\0';
    int n = 0;
    char* p = line;
    while (*p) {
        while (*p && isspace((unsigned char)*p)) p++;
        if (!*p) break;
        if (n < max) args[n] = p;
        n++;
        while (*p && !isspace((unsigned char)*p)) p++;
        if (*p) *p++ = '\0';
    }
    return n;
}

/* src/import_eam.cpp:150-166: numbers may span lines; a line may not carry
 * more numbers than are still expected. */
static int read_doubles(reader* r, double* out, size_t n, char* err, int errlen)
{
    size_t i = 0;
    char* args[4096];
    while (i < n) {
        char* l = next_line(r);
        if (!l) { snprintf(err, errlen, "file ended immaturely"); return -1; }
        int na = tokens(l, args, 4096);
        if (na > 4096 || (size_t)na + i > n) { snprintf(err, errlen, "file has extra arguments"); return -1; }
        for (int k = 0; k < na; k++) out[i++] = atof(args[k]);
    }
    return 0;
}

o_eam* o_eam_read(const char* path, int fmt, int dmd, char* err, int errlen)
{
    FILE* fp = fopen(path, "rb");
    if (!fp) { snprintf(err, errlen, "cannot open %s", path); return NULL; }
    fseek(fp, 0, SEEK_END);
    long sz = ftell(fp);
    fseek(fp, 0, SEEK_SET);
    reader r;
    r.buf = (char*)malloc((size_t)sz + 1);
    if (fread(r.buf, 1, (size_t)sz, fp) != (size_t)sz) { fclose(fp); free(r.buf); snprintf(err, errlen, "read error"); return NULL; }
    r.buf[sz] = '\0';
    fclose(fp);
    r.cur = r.buf; r.line_no = 0;
    o_eam* e = NULL;
    double* buf = NULL;
    char* args[64];
    for (int k = 0; k < 3; k++)
        if (!next_line(&r)) { snprintf(err, errlen, "file ended immaturely"); goto fail; }
    char* l = next_line(&r);
    if (!l) { snprintf(err, errlen, "file ended immaturely"); goto fail; }
    int na = tokens(l, args, 64);
    if (na < 2) { snprintf(err, errlen, "invalid element line"); goto fail; }
    int ne = na - 1;
    l = next_line(&r);
    if (!l) { snprintf(err, errlen, "file ended immaturely"); goto fail; }
    size_t nrho, nr;
    double drho, dr, rc_hdr;
    if (sscanf(l, " %zu %lf %zu %lf %lf", &nrho, &drho, &nr, &dr, &rc_hdr) != 5) {
        snprintf(err, errlen, "invalid grid line"); goto fail;
    }
    if (nrho < 5 || nr < 5 || drho <= 0.0 || dr <= 0.0) { snprintf(err, errlen, "invalid grid"); goto fail; }
    e = alloc_eam(ne, (int)nr, (int)nrho, dr, drho, dmd);
    size_t nblock = fmt == 1 ? nrho + nr * (size_t)ne : nrho + nr;
    size_t npair = (size_t)ne * (ne + 1) / 2 * nr;
    buf = (double*)malloc(sizeof(double) * (nblock > npair ? nblock : npair));
    for (int i = 0; i < ne; i++) {
        if (!next_line(&r)) { snprintf(err, errlen, "file ended immaturely"); goto fail; }
        if (read_doubles(&r, buf, nblock, err, errlen)) goto fail;
        for (size_t k = 0; k < nrho; k++) OF(e, i)[k * e->nf] = buf[k];
        if (fmt == 1) {
            for (int j = 0; j < ne; j++)
                for (size_t k = 0; k < nr; k++) ORHO(e, i, j)[k][0] = buf[nrho + j * nr + k];
        } else {
            for (int j = 0; j < ne; j++)
                for (size_t k = 0; k < nr; k++) ORHO(e, i, j)[k][0] = buf[nrho + k];
        }
    }
    if (read_doubles(&r, buf, npair, err, errlen)) goto fail;
    size_t ipos = 0;
    for (int i = 0; i < ne; i++)
        for (int j = 0; j <= i; j++) {
            for (size_t k = 0; k < nr; k++) {
                OPHI(e, i, j)[k][0] = buf[ipos];
                OPHI(e, j, i)[k][0] = buf[ipos];
                ipos++;
            }
        }
    free(buf);
    free(r.buf);
    finish_eam(e);
    return e;
fail:
    free(buf);
    free(r.buf);
    if (e) o_eam_free(e);
    return NULL;
}

void o_eam_free(o_eam* e)
{
    if (!e) return;
    free(e->r_phi); free(e->rho); free(e->F); free(e->rc); free(e);
}

void o_eam_info(const o_eam* e, int* ne, int* nr, int* nrho, double* dr, double* drho)
{
    *ne = e->ne; *nr = e->nr; *nrho = e->nrho; *dr = e->dr; *drho = e->drho;
}

void o_eam_cut(const o_eam* e, double* rc)
{
    memcpy(rc, e->rc, sizeof(double) * e->ne * e->ne);
}

void o_eam_coefs(const o_eam* e, int which, int i, int j, double* out)
{
    if (which == 0) memcpy(out, OPHI(e, i, j), sizeof(double) * 4 * e->nr);
    else if (which == 1) memcpy(out, ORHO(e, i, j), sizeof(double) * 4 * e->nr);
    else memcpy(out, OF(e, i), sizeof(double) * e->nf * e->nrho);
}

/* Physicists' Gauss-Hermite rule by Newton iteration on the Hermite
 * polynomial written in u = x^2. Restates src/xmath.cpp:440-589. */
void o_quadrature_hg(int n, double* x, double* w)
{
    int m = n / 2 + 1;
    double* p1 = (double*)calloc((size_t)m, sizeof(double));
    double* p2 = (double*)calloc((size_t)m, sizeof(double));
    double* pc = (double*)calloc((size_t)m, sizeof(double));
    double* dpc = (double*)calloc((size_t)m, sizeof(double));
    p1[0] = 1.0;
    pc[0] = 1.0;
    /* H_i = 2x H_{i-1} - 2(i-1) H_{i-2}, stored by powers of u = x^2
     * (odd polynomials carry an implicit factor x). */
    for (int i = 1; i < n + 1; i++) {
        double ii = (double)i;
        int mi = i / 2 + 1;
        if (i % 2 == 0) {
            for (int j = 0; j < mi; j++) pc[j] = -2.0 * (ii - 1.0) * p2[j];
            for (int j = 1; j < mi; j++) pc[j] += 2.0 * p1[j - 1];
        } else {
            for (int j = 0; j < mi; j++) pc[j] = 2.0 * p1[j] - 2.0 * (ii - 1.0) * p2[j];
        }
        for (int j = 0; j < mi; j++) { p2[j] = p1[j]; p1[j] = pc[j]; }
    }
    m = n / 2 + 1;
    for (int i = 1; i < m; i++) dpc[i - 1] = pc[i] * (double)i;
    double tol = 2.220446049250313e-16;
    int ord = m;
    int icurs = n / 2;
    if (n % 2 == 1) { x[icurs] = 0.0; icurs++; }
    double a = pc[m - 1];
    for (int i = 0; i < m; i++) pc[i] /= a;
    for (int i = 1; i < m; i++) dpc[i - 1] /= a;
    while (ord > 1) {
        double u0 = 0.0, f = 1.0, df, du = 0.0, t0, t1;
        int iter = 50;
        while (fabs(f) > tol && iter) {
            u0 += du;
            df = f = 0.0;
            t0 = 1.0;
            for (int j = 0; j < ord - 1; j++) { f += pc[j] * t0; df += dpc[j] * t0; t0 *= u0; }
            f += pc[ord - 1] * t0;
            du = -f / df;
            iter--;
        }
        x[icurs] = sqrt(u0);
        t0 = 1.0; t1 = 0.0;
        for (int i = 0; i < m; i++) { t1 += p2[i] * t0; t0 *= u0; }
        if (n % 2 == 0) t1 *= sqrt(u0);
        w[icurs] = 1.0 / (t1 * t1);
        /* deflate the root u0 out of the monic polynomial */
        t0 = pc[ord - 2];
        pc[ord - 2] = 1.0;
        dpc[ord - 2] = (double)(ord - 1);
        for (int i = ord - 3; i > -1; i--) {
            t1 = pc[i];
            pc[i] = t0 + u0 * pc[i + 1];
            dpc[i] = pc[i + 1] * (double)(i + 1);
            t0 = t1;
        }
        ord--;
        icurs++;
    }
    double t0 = sqrt(M_PI) / (double)n;
    for (int i = 0; i < n - 1; i++) t0 *= (double)(2 * (i + 1));
    if (n % 2 == 1) w[n / 2] = t0 / (p2[0] * p2[0]);
    for (int i = 0; i < n / 2; i++) {
        icurs--;
        w[icurs] *= t0;
        w[i] = w[icurs];
        x[i] = -x[icurs];
    }
    free(p1); free(p2); free(pc); free(dpc);
}

/* Park-Miller minimal standard generator with Schrage's factorisation and
 * a Marsaglia polar Gaussian that caches its second value.
 * Restates src/random.cpp:7-68 (the divide by 2^31-1 is a true division). */
void o_rng_init(o_rng* r, int seed) { r->seed = seed; r->gauss_chk = 0; r->reserved = 0.0; }

double o_rng_uniform(o_rng* r)
{
    int hi = r->seed / 127773;
    int lo = r->seed % 127773;
    int t = 16807 * lo - 2836 * hi;
    r->seed = t > 0 ? t : t + 2147483647;
    return r->seed * 1.0 / 2147483647;
}

double o_rng_gaussian(o_rng* r)
{
    if (r->gauss_chk) { r->gauss_chk = 0; return r->reserved; }
    r->gauss_chk = 1;
    double u = 0.0, v = 0.0, s = 1.0;
    while (s >= 1.0) {
        u = 2.0 * o_rng_uniform(r) - 1.0;
        v = 2.0 * o_rng_uniform(r) - 1.0;
        s = u * u + v * v;
    }
    double f = sqrt(-2.0 * log(s) / s);
    r->reserved = v * f;
    return u * f;
}

/* ImportEAM::interpolate(arr, n, p, k) (src/import_eam.cpp:9-52): the cubic of
 * interval k through the knots with 5-point slopes (one-sided at the ends,
 * linear past the last knot), at offset p */
static double funcfl_interp(const double* a, size_t n, double p, size_t k)
{
    double c0 = a[k], c1, c2, c3;
    if (k == 0) {
        c1 = a[1] - a[0];
        c2 = (-a[0] + 2.0 * a[1] - a[2]) / 2.0;
        c3 = (a[0] - 2.0 * a[1] + a[2]) / 2.0;
    } else if (k == 1) {
        c1 = 0.5 * (a[2] - a[0]);
        c2 = (11.0 * a[0] - 28.0 * a[1] + 24.0 * a[2] - 8.0 * a[3] + a[4]) / 12.0;
        c3 = (-5.0 * a[0] + 16.0 * a[1] - 18.0 * a[2] + 8.0 * a[3] - a[4]) / 12.0;
    } else if (k == n - 2) {
        c1 = 0.5 * (a[n - 1] - a[n - 3]);
        c2 = a[n - 3] - 3.0 * a[n - 2] + 2.0 * a[n - 1];
        c3 = 0.5 * (-a[n - 3] + 3.0 * a[n - 2] - 2.0 * a[n - 1]);
    } else if (k == n - 1) {
        return c0 + (a[n - 1] - a[n - 2]) * p;
    } else {
        c1 = ((a[k - 2] - a[k + 2]) + 8.0 * (a[k + 1] - a[k - 1])) / 12.0;
        c2 = (-2.0 * a[k - 2] + 15.0 * a[k - 1] - 28.0 * a[k] + 21.0 * a[k + 1] - 6.0 * a[k + 2]) / 12.0;
        c3 = (a[k - 2] - 7.0 * a[k - 1] + 16.0 * a[k] - 17.0 * a[k + 1] + 7.0 * a[k + 2]) / 12.0;
    }
    return ((c3 * p + c2) * p + c1) * p + c0;
}

/* position of r on a grid of n knots spaced d: interval l, offset p
 * (src/import_eam.h:289-296) */
static size_t funcfl_pos(double r, double d, size_t n, double* p)
{
    double q = r / d;
    size_t l = (size_t)q;
    if (l > n - 2) l = n - 2;
    q -= (double)l;
    *p = q < 1.0 ? q : 1.0;
    return l;
}

/* ImportEAM::funcfl (src/import_eam.h:68-344): one DYNAMO funcfl file per
 * element. Each file: a comment line, the Z / mass / lattice line, then
 * "nrho drho nr dr rc" and F [nrho], Z [nr], rho [nr] in free format. The
 * files are resampled onto the coarsest spacing over the longest range
 * (nr = (n_r - 1) dr / dr_max rounded, likewise nrho), r phi_ij =
 * 27.2 * 0.529 Z_i Z_j, rho_ij = rho_i. As the reference does, Z is read at
 * interval k of the output grid while p is the offset within the file's own
 * interval; on one grid the two coincide. */
o_eam* o_eam_read_funcfl(const char* const* paths, int nfiles, int dmd, char* err, int errlen)
{
    if (nfiles < 1 || nfiles > 8) { snprintf(err, errlen, "1 to 8 funcfl files"); return NULL; }
    size_t nrs[8], nrhos[8];
    double drs[8], drhos[8];
    double* Fv[8] = {0};
    double* Zv[8] = {0};
    double* Rv[8] = {0};
    o_eam* e = NULL;
    double *F = NULL, *rho = NULL, *rphi = NULL;
    int ok = 0;
    for (int i = 0; i < nfiles; i++) {
        FILE* fp = fopen(paths[i], "rb");
        if (!fp) { snprintf(err, errlen, "cannot open %s", paths[i]); goto done; }
        fseek(fp, 0, SEEK_END);
        long sz = ftell(fp);
        fseek(fp, 0, SEEK_SET);
        reader r;
        r.buf = (char*)malloc((size_t)sz + 1);
        size_t got = fread(r.buf, 1, (size_t)sz, fp);
        fclose(fp);
        r.buf[got] = '\0';
        r.cur = r.buf; r.line_no = 0;
        char* l = NULL;
        if (!next_line(&r) || !next_line(&r) || !(l = next_line(&r))) {
            snprintf(err, errlen, "%s file ended immaturely", paths[i]); free(r.buf); goto done;
        }
        double rc_hdr;
        if (sscanf(l, " %zu %lf %zu %lf %lf", &nrhos[i], &drhos[i], &nrs[i], &drs[i], &rc_hdr) != 5) {
            snprintf(err, errlen, "invalid line in %s", paths[i]); free(r.buf); goto done;
        }
        if (nrhos[i] < 5 || nrs[i] < 5 || drhos[i] <= 0.0 || drs[i] <= 0.0) {
            snprintf(err, errlen, "invalid grid in %s", paths[i]); free(r.buf); goto done;
        }
        size_t nv = nrhos[i] + 2 * nrs[i];
        double* buf = (double*)malloc(sizeof(double) * nv);
        if (read_doubles(&r, buf, nv, err, errlen)) { free(buf); free(r.buf); goto done; }
        free(r.buf);
        Fv[i] = buf;
        Zv[i] = buf + nrhos[i];
        Rv[i] = buf + nrhos[i] + nrs[i];
    }
    {
        double maxr = 0.0, maxrho = 0.0, maxdr = 0.0, maxdrho = 0.0;
        for (int i = 0; i < nfiles; i++) {
            double a = (double)(nrs[i] - 1) * drs[i], b = (double)(nrhos[i] - 1) * drhos[i];
            if (a > maxr) maxr = a;
            if (b > maxrho) maxrho = b;
            if (drs[i] > maxdr) maxdr = drs[i];
            if (drhos[i] > maxdrho) maxdrho = drhos[i];
        }
        size_t nr = (size_t)(maxr / maxdr + 0.5), nrho = (size_t)(maxrho / maxdrho + 0.5);
        double dr = maxdr, drho = maxdrho;
        int ne = nfiles;
        for (int i = 0; i < ne; i++)
            if (nr > nrs[i]) {  /* the reference would read past this file's Z table */
                snprintf(err, errlen, "funcfl: %s has fewer r knots than the common grid", paths[i]);
                goto done;
            }
        F = (double*)calloc((size_t)ne * nrho, sizeof(double));
        rho = (double*)calloc((size_t)ne * ne * nr, sizeof(double));
        rphi = (double*)calloc((size_t)ne * ne * nr, sizeof(double));
        double p;
        for (int i = 0; i < ne; i++)
            for (int j = 0; j <= i; j++)
                for (size_t k = 0; k < nr; k++) {
                    double r = (double)k * dr;
                    funcfl_pos(r, drs[i], nrs[i], &p);
                    double t0 = funcfl_interp(Zv[i], nrs[i], p, k);
                    funcfl_pos(r, drs[j], nrs[j], &p);
                    double t1 = funcfl_interp(Zv[j], nrs[j], p, k);
                    double v = 27.2 * 0.529 * t0 * t1;
                    rphi[((size_t)i * ne + j) * nr + k] = v;
                    rphi[((size_t)j * ne + i) * nr + k] = v;
                }
        for (int i = 0; i < ne; i++)
            for (size_t k = 0; k < nr; k++) {
                size_t l = funcfl_pos((double)k * dr, drs[i], nrs[i], &p);
                rho[((size_t)i * ne) * nr + k] = funcfl_interp(Rv[i], nrs[i], p, l);
            }
        for (int i = 0; i < ne; i++)
            for (size_t k = 0; k < nrho; k++) {
                size_t l = funcfl_pos((double)k * drho, drhos[i], nrhos[i], &p);
                F[(size_t)i * nrho + k] = funcfl_interp(Fv[i], nrhos[i], p, l);
            }
        e = o_eam_from_knots(ne, (int)nr, (int)nrho, dr, drho, F, rho, rphi, 1, dmd);
        ok = 1;
    }
done:
    for (int i = 0; i < nfiles; i++) free(Fv[i]);
    free(F); free(rho); free(rphi);
    return ok ? e : NULL;
}
