/*
 * cntc_oracle.c -- CPU restatement (FP64) of CONTACT's inner contact solve.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP engine.  Loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by the product.
 *
 * Follows the reference with its MKL-FFT configuration (WITH_MKLFFT defined), which is the
 * configuration that produced the reference's golden outputs:
 *   - every VecAijPj product is an FFT convolution (src/m_aijpj.f90:344-347);
 *   - NormCG and TangCG use the FFT approximate-inverse preconditioner
 *     (src/m_snorm.f90:70-98, src/m_solvpt.f90:1877-1938).
 * The FFT here is a plain mixed-radix DFT written for this file (any length); the padded
 * sizes are the reference's 2*opt_fft_size(n) for products and 2*cf_mx x 2*cf_my for the
 * preconditioner, so only rounding separates it from the MKL transform.
 *
 * Not restated: ConvexGS (G=2, and the reference's switch to it when contact reaches the
 * trailing-edge column in steady rolling); such cases return CNTC_err_other.
 * Fatal paths (abort_run) become error codes.
 */
#include "cntc_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define CNTC_ERR_NORM (-27)
#define CNTC_ERR_TANG (-28)
#define CNTC_ERR_OTHER (-99)

static const double PI = 3.14159265358979323846264338327950288;
static const double TINY = 1e-20;   /* m_globals.f90:37 */

/* ------------------------------------------------------------------------------------------ */
/* small helpers                                                                                */
/* ------------------------------------------------------------------------------------------ */

static double *dalloc(size_t n) { return (double *)calloc(n ? n : 1, sizeof(double)); }
static int *ialloc(size_t n) { return (int *)calloc(n ? n : 1, sizeof(int)); }
static double sqr(double x) { return x * x; }
static double dmax(double a, double b) { return a > b ? a : b; }
static double dmin(double a, double b) { return a < b ? a : b; }
static int imax(int a, int b) { return a > b ? a : b; }
static int imin(int a, int b) { return a < b ? a : b; }
/* Fortran single-precision literal promoted to double, e.g. 0.60 -> (double)0.6f */
static double sp(float x) { return (double)x; }
/* Fortran nint */
static int nint(double x) { return (int)(x >= 0 ? floor(x + 0.5) : -floor(-x + 0.5)); }
/* Fortran sign(a,b) */
static double fsign(double a, double b) { return b >= 0 ? fabs(a) : -fabs(a); }

/* ------------------------------------------------------------------------------------------ */
/* FFT of any length (recursive mixed radix, smallest prime factor first)                       */
/* ------------------------------------------------------------------------------------------ */

typedef struct { double re, im; } cpx;

static cpx cmul(cpx a, cpx b) { cpx c = {a.re * b.re - a.im * b.im, a.re * b.im + a.im * b.re}; return c; }

typedef struct { int n; cpx *w; } twtab;    /* w[k] = exp(-2 pi i k / n) */

static _Thread_local twtab tw_cache[16];   /* per thread: the solver may run on several threads */
static _Thread_local int tw_ncache = 0;

static const cpx *twiddles(int n)
{
    for (int i = 0; i < tw_ncache; i++)
        if (tw_cache[i].n == n) return tw_cache[i].w;
    int slot = tw_ncache < 16 ? tw_ncache++ : (int)(rand() % 16);
    if (tw_cache[slot].w) free(tw_cache[slot].w);
    tw_cache[slot].n = n;
    tw_cache[slot].w = (cpx *)malloc(sizeof(cpx) * (size_t)n);
    for (int k = 0; k < n; k++) {
        double a = -2.0 * PI * (double)k / (double)n;
        tw_cache[slot].w[k].re = cos(a);
        tw_cache[slot].w[k].im = sin(a);
    }
    return tw_cache[slot].w;
}

static int smallest_factor(int n)
{
    if (n % 4 == 0) return 4;
    if (n % 2 == 0) return 2;
    for (int p = 3; p * p <= n; p += 2)
        if (n % p == 0) return p;
    return n;
}

/* out[k] = sum_j in[j*is] * exp(sign 2 pi i jk/n); wN = twiddles of the top-level length N,
   with n | N; scratch has room for n values */
static void fft_rec(int n, int sign, const cpx *in, int is, cpx *out, const cpx *wN, int N,
                    cpx *scratch)
{
    if (n == 1) { out[0] = in[0]; return; }
    int p = smallest_factor(n);
    int m = n / p;
    for (int r = 0; r < p; r++)
        fft_rec(m, sign, in + (size_t)r * is, is * p, scratch + (size_t)r * m, wN, N,
                out + (size_t)r * m);
    /* combine: X[k + m q] = sum_r W_n^{r(k + m q)} F_r[k] */
    int step = N / n;
    for (int k = 0; k < m; k++) {
        for (int q = 0; q < p; q++) {
            int kk = k + m * q;
            cpx acc = {0.0, 0.0};
            for (int r = 0; r < p; r++) {
                long e = ((long)r * kk) % n;
                cpx w = wN[(size_t)e * step];
                if (sign > 0) w.im = -w.im;
                cpx t = cmul(w, scratch[(size_t)r * m + k]);
                acc.re += t.re;
                acc.im += t.im;
            }
            out[kk] = acc;
        }
    }
}

static void fft1(int n, int sign, const cpx *in, int is, cpx *out, cpx *work)
{
    const cpx *w = twiddles(n);
    fft_rec(n, sign, in, is, out, w, n, work);
}

/* in-place 2-D complex FFT of a[ly][lx] (x fastest) */
static void fft2(int lx, int ly, int sign, cpx *a)
{
    int nmax = imax(lx, ly);
    cpx *tmp = (cpx *)malloc(sizeof(cpx) * (size_t)nmax);
    cpx *out = (cpx *)malloc(sizeof(cpx) * (size_t)nmax);
    cpx *work = (cpx *)malloc(sizeof(cpx) * (size_t)nmax);
    for (int iy = 0; iy < ly; iy++) {
        fft1(lx, sign, a + (size_t)iy * lx, 1, out, work);
        memcpy(a + (size_t)iy * lx, out, sizeof(cpx) * (size_t)lx);
    }
    for (int ix = 0; ix < lx; ix++) {
        for (int iy = 0; iy < ly; iy++) tmp[iy] = a[(size_t)iy * lx + ix];
        fft1(ly, sign, tmp, 1, out, work);
        for (int iy = 0; iy < ly; iy++) a[(size_t)iy * lx + ix] = out[iy];
    }
    free(tmp);
    free(out);
    free(work);
}

void orc_fft(int n, const double *re_in, const double *im_in, double *re_out, double *im_out,
             int sign)
{
    cpx *in = (cpx *)malloc(sizeof(cpx) * (size_t)n);
    cpx *out = (cpx *)malloc(sizeof(cpx) * (size_t)n);
    cpx *work = (cpx *)malloc(sizeof(cpx) * (size_t)n);
    for (int i = 0; i < n; i++) { in[i].re = re_in[i]; in[i].im = im_in ? im_in[i] : 0.0; }
    fft1(n, sign < 0 ? -1 : 1, in, 1, out, work);
    for (int i = 0; i < n; i++) { re_out[i] = out[i].re; im_out[i] = out[i].im; }
    free(in);
    free(out);
    free(work);
}

/* ------------------------------------------------------------------------------------------ */
/* opt_fft_size, src/m_aijpj.f90:1022-1119                                                      */
/* ------------------------------------------------------------------------------------------ */

int orc_opt_fft_size(int fft_size)
{
    /* trading patterns (d2,d3,d5,d7,d11, num, den, maxfac) of the reference table; the size
       starts at the next power of two and factors are traded while the product stays >= n */
    static const int kpat[9][8] = {
        {-2, 1, 0, 0, 0, 4, 3, 3},   {-5, 3, 0, 0, 0, 32, 27, 3}, {-1, -1, 1, 0, 0, 6, 5, 5},
        {-4, 1, 1, 0, 0, 16, 15, 5}, {0, -3, 2, 0, 0, 27, 25, 5}, {-3, 0, 0, 1, 0, 8, 7, 7},
        {1, -1, -1, 1, 0, 15, 14, 7}, {-1, 2, -1, 0, 0, 10, 9, 5}, {-2, -1, 0, 0, 1, 12, 11, 11}};
    const int use_fac = 7;
    int k2 = (int)ceil(log(1.0 * fft_size) / log(2.0));
    long prod = 1L << k2;
    int kf[5] = {k2, 0, 0, 0, 0};
    int changed = 1;
    while (changed) {
        changed = 0;
        for (int ip = 0; ip < 9; ip++) {
            if (kpat[ip][5] <= 0 || kpat[ip][7] > use_fac) continue;
            for (;;) {
                int ok = 1;
                for (int f = 0; f < 5; f++)
                    if (kf[f] + kpat[ip][f] < 0) ok = 0;
                if (!ok || (long)kpat[ip][6] * prod < (long)kpat[ip][5] * fft_size) break;
                for (int f = 0; f < 5; f++) kf[f] += kpat[ip][f];
                prod = (long)kpat[ip][6] * prod / kpat[ip][5];
                changed = 1;
            }
        }
    }
    return (int)prod;
}

/* ------------------------------------------------------------------------------------------ */
/* influence coefficients, src/m_visc.f90:431-604                                               */
/* ------------------------------------------------------------------------------------------ */

/* index of cf(ix,iy,ik,jk) in a block array of size (2mx)(2my) per (ik,jk) pair */
#define CFIDX(ix, iy, mx, my) ((size_t)((iy) + (my)) * (size_t)(2 * (mx)) + (size_t)((ix) + (mx)))

void orc_elascf_pcwcns(double akv, double nuv, int mx, int my, double dx, double dy,
                       double xshft, double yshft, double *cf)
{
    size_t nb = (size_t)(2 * mx) * (size_t)(2 * my);
    int nx1 = 2 * mx + 1, ny1 = 2 * my + 1;
    /* work arrays indexed (ix+mx, iy+my) for ix in [-mx,mx], iy in [-my,my] */
#define W(ix, iy) ((size_t)((iy) + my) * (size_t)nx1 + (size_t)((ix) + mx))
    double *r = dalloc((size_t)nx1 * ny1), *x = dalloc((size_t)nx1 * ny1),
           *y = dalloc((size_t)nx1 * ny1), *xly = dalloc((size_t)nx1 * ny1),
           *ylx = dalloc((size_t)nx1 * ny1);
    double e1 = (1.0 - nuv) / PI, e2 = 1.0 / PI, e3 = nuv / PI;
    double tolx = dx * 1e-10, tolx2 = dx * 1e-20;
    for (size_t i = 0; i < 9 * nb; i++) cf[i] = 0.0;
    int iy0 = -my;
    if (my == 1) iy0 = -my + 1;
    for (int iy = iy0; iy <= my; iy++)
        for (int ix = -mx; ix <= mx; ix++) {
            x[W(ix, iy)] = (double)(float)ix * dx + xshft - 0.5 * dx;
            y[W(ix, iy)] = (double)(float)iy * dy + yshft - 0.5 * dy;
            r[W(ix, iy)] = sqrt(sqr(x[W(ix, iy)]) + sqr(y[W(ix, iy)]));
        }
    double *c11 = cf + 0 * nb, *c12 = cf + 1 * nb, *c13 = cf + 2 * nb;
    double *c21 = cf + 3 * nb, *c22 = cf + 4 * nb, *c23 = cf + 5 * nb;
    double *c31 = cf + 6 * nb, *c32 = cf + 7 * nb, *c33 = cf + 8 * nb;
    for (int iy = iy0; iy <= my - 1; iy++)
        for (int ix = -mx; ix <= mx - 1; ix++) {
            size_t k = CFIDX(ix, iy, mx, my);
            c21[k] = c21[k] - e3 * r[W(ix, iy)];
            c21[k] = c21[k] + e3 * r[W(ix, iy + 1)];
            c21[k] = c21[k] + e3 * r[W(ix + 1, iy)];
            c21[k] = c21[k] - e3 * r[W(ix + 1, iy + 1)];
        }
    if (fabs(akv) >= 1e-6) {
        double *alr = dalloc((size_t)nx1 * ny1), *atxy = dalloc((size_t)nx1 * ny1),
               *atyx = dalloc((size_t)nx1 * ny1);
        for (int iy = iy0; iy <= my; iy++)
            for (int ix = -mx; ix <= mx; ix++) {
                size_t w = W(ix, iy);
                alr[w] = log(r[w]);
                atyx[w] = atan(y[w] / (x[w] + tolx2));
                atxy[w] = atan(x[w] / (y[w] + tolx2));
            }
        for (int iy = iy0; iy <= my - 1; iy++)
            for (int ix = -mx; ix <= mx - 1; ix++) {
                size_t k = CFIDX(ix, iy, mx, my);
                size_t w00 = W(ix, iy), w01 = W(ix, iy + 1), w10 = W(ix + 1, iy),
                       w11 = W(ix + 1, iy + 1);
                c13[k] = c13[k] - e2 * akv * (y[w00] * alr[w00] + x[w00] * atyx[w00]);
                c13[k] = c13[k] + e2 * akv * (y[w01] * alr[w01] + x[w01] * atyx[w01]);
                c13[k] = c13[k] + e2 * akv * (y[w10] * alr[w10] + x[w10] * atyx[w10]);
                c13[k] = c13[k] - e2 * akv * (y[w11] * alr[w11] + x[w11] * atyx[w11]);
                c23[k] = c23[k] - e2 * akv * (x[w00] * alr[w00] + y[w00] * atxy[w00]);
                c23[k] = c23[k] + e2 * akv * (x[w01] * alr[w01] + y[w01] * atxy[w01]);
                c23[k] = c23[k] + e2 * akv * (x[w10] * alr[w10] + y[w10] * atxy[w10]);
                c23[k] = c23[k] - e2 * akv * (x[w11] * alr[w11] + y[w11] * atxy[w11]);
            }
        free(alr);
        free(atxy);
        free(atyx);
    }
    for (int iy = iy0; iy <= my - 1; iy++)
        for (int ix = -mx; ix <= mx; ix++) {
            size_t w0 = W(ix, iy), w1 = W(ix, iy + 1);
            if (dmin(fabs(y[w0] + r[w0]), fabs(y[w1] + r[w1])) < tolx)
                xly[w0] = 0.0;
            else
                xly[w0] = x[w0] * log((y[w1] + r[w1]) / (y[w0] + r[w0]));
        }
    for (int iy = iy0; iy <= my; iy++)
        for (int ix = -mx; ix <= mx - 1; ix++) {
            size_t w0 = W(ix, iy), w1 = W(ix + 1, iy);
            if (dmin(fabs(x[w0] + r[w0]), fabs(x[w1] + r[w1])) < tolx)
                ylx[w0] = 0.0;
            else
                ylx[w0] = y[w0] * log((x[w1] + r[w1]) / (x[w0] + r[w0]));
        }
    for (int iy = iy0; iy <= my - 1; iy++)
        for (int ix = -mx; ix <= mx - 1; ix++) {
            size_t k = CFIDX(ix, iy, mx, my);
            c33[k] = c33[k] - e1 * (xly[W(ix, iy)] + ylx[W(ix, iy)]);
            c33[k] = c33[k] + e1 * (xly[W(ix + 1, iy)] + ylx[W(ix, iy + 1)]);
            c11[k] = c11[k] - (e1 * xly[W(ix, iy)] + e2 * ylx[W(ix, iy)]);
            c11[k] = c11[k] + (e1 * xly[W(ix + 1, iy)] + e2 * ylx[W(ix, iy + 1)]);
            c22[k] = c22[k] - (e1 * ylx[W(ix, iy)] + e2 * xly[W(ix, iy)]);
            c22[k] = c22[k] + (e1 * ylx[W(ix, iy + 1)] + e2 * xly[W(ix + 1, iy)]);
        }
    for (int iy = iy0; iy <= my - 1; iy++)
        for (int ix = -mx; ix <= mx - 1; ix++) {
            size_t k = CFIDX(ix, iy, mx, my);
            c31[k] = -c13[k];
            c32[k] = -c23[k];
            c12[k] = c21[k];
        }
#undef W
    free(r);
    free(x);
    free(y);
    free(xly);
    free(ylx);
}

/* ------------------------------------------------------------------------------------------ */
/* problem data                                                                                 */
/* ------------------------------------------------------------------------------------------ */

typedef struct {
    int defined;
    int cf_mx, cf_my;
    double dx, dy, dq;
    int nt_cpl;
    double ga, ga_inv;
    double *cf;                 /* 9 blocks of (2cf_mx)(2cf_my) */
    /* cached spectra of the blocks at the padded size of the current grid */
    int sp_mx, sp_my, sp_lx, sp_ly;
    cpx *spec[9];
} inflcf;

typedef struct {
    int *el;
    int *row1st, *rowlst;       /* 1-based ix per row */
    int ixmin, ixmax, iymin, iymax;
} eldiv;

struct orc_problem {
    /* grid */
    int mx, my, npot;
    double dx, dy, dxdy, xc1, yc1;
    double *x, *y;
    /* hertz */
    double hz_a1, hz_b1, hz_aa, hz_bb, hz_rho, hz_cp;
    /* material */
    double ga, nu, ak;
    /* influence coefficients */
    inflcf cs, cv, csv, ms;
    /* grid functions: [0]=x, [1]=y, [2]=n */
    double *hs[3];
    double *ps[3], *pv[3], *ss[2], *facdt;
    int *lix;          /* per element: 0-based ix of its leading-edge position when facdt < 0.9999, else -1 */
    eldiv igs, igv;
    int have_solution;
    /* kinematics and results */
    double pen, fntrue, fxrel, fyrel, fcntc[3], mztrue;
    double cksi, ceta, cphi, chi, dq, veloc, dt, muscal, fstat;
    int itnorm, ittang, itcg, itgs;
    int ierror;
    int gd_stagnated;
    int used_stdygs;
    int use_cnvxgs;     /* the tangential solver of this case is ConvexGS (set by stang) */
    /* sensitivities sens(iout, iin), column-major as in the reference (outpt%sens) */
    double sens[16];
    /* initial estimates of a sensitivity solve: the stopping tests measure the change from
       them (is_sens, src/m_solvpn.f90:94-97, src/m_solvpt.f90:1910-1914, 2895-2899) */
    const double *p0[3];
};

enum { IOUT_FN = 0, IOUT_FX = 1, IOUT_FY = 2, IOUT_MZ = 3 };
enum { IIN_PEN = 0, IIN_KSI = 1, IIN_ETA = 2, IIN_PHI = 3 };
#define SENS(P, io, ii) ((P)->sens[(io) + 4 * (ii)])

static void eldiv_alloc(eldiv *e, int npot, int my)
{
    free(e->el);
    free(e->row1st);
    free(e->rowlst);
    e->el = ialloc((size_t)npot);
    e->row1st = ialloc((size_t)my);
    e->rowlst = ialloc((size_t)my);
}

/* areas: row extents of the contact area, src/m_gridfunc.f90:582-637 */
static void areas(eldiv *e, int nx, int ny)
{
    e->ixmin = nx;
    e->ixmax = 1;
    e->iymin = ny;
    e->iymax = 1;
    for (int iy = 1; iy <= ny; iy++) {
        int i0 = (iy - 1) * nx;
        int first = 1;
        while (e->el[first + i0 - 1] <= ORC_EXTER && first != nx) first++;
        int last = 0;
        for (int ix = first; ix <= nx; ix++)
            if (e->el[ix + i0 - 1] >= ORC_ADHES) last = ix;
        e->row1st[iy - 1] = first;
        e->rowlst[iy - 1] = last;
        if (first <= last) {
            e->ixmin = imin(e->ixmin, first);
            e->ixmax = imax(e->ixmax, last);
            e->iymin = imin(e->iymin, iy);
            e->iymax = imax(e->iymax, iy);
        }
    }
}

static void count_el(const int *el, int n, int *nadh, int *nslip, int *nplast, int *nexter)
{
    *nadh = *nslip = *nplast = *nexter = 0;
    for (int i = 0; i < n; i++) {
        if (el[i] == ORC_ADHES) (*nadh)++;
        else if (el[i] == ORC_SLIP) (*nslip)++;
        else if (el[i] == ORC_PLAST) (*nplast)++;
        else (*nexter)++;
    }
}

/* ------------------------------------------------------------------------------------------ */
/* masked grid-function operations, src/m_gridfunc.f90:936-1650                                 */
/* ------------------------------------------------------------------------------------------ */

enum { ALLELM = -9, ALLINT = -8 };

static int in_mask(int iigs, int el) { return iigs == ALLELM || (iigs == ALLINT && el >= ORC_ADHES); }

static double gf_dot(int iigs, const int *el, int n, const double *a, const double *b)
{
    double s = 0.0;
    for (int i = 0; i < n; i++)
        if (in_mask(iigs, el[i])) s += a[i] * b[i];
    return s;
}
static double gf_sum(int iigs, const int *el, int n, const double *a)
{
    double s = 0.0;
    for (int i = 0; i < n; i++)
        if (in_mask(iigs, el[i])) s += a[i];
    return s;
}
/* rms over 1 or 2 components, src/m_gridfunc.f90:1467-1507 */
static double gf_rms(int iigs, const int *el, int n, const double *a, const double *b)
{
    double s = 0.0;
    long cnt = 0;
    for (int i = 0; i < n; i++)
        if (in_mask(iigs, el[i])) { s += a[i] * a[i]; cnt++; }
    if (b)
        for (int i = 0; i < n; i++)
            if (in_mask(iigs, el[i])) { s += b[i] * b[i]; cnt++; }
    return sqrt(s / (double)(cnt > 1 ? cnt : 1));
}
/* rms of (a - a0, b - b0): gf3_copy + gf3_axpy(-1, ps0) + gf3_rms of the is_sens tests */
static double gf_rms_diff(int iigs, const int *el, int n, const double *a, const double *a0,
                          const double *b, const double *b0)
{
    double s = 0.0;
    long cnt = 0;
    for (int i = 0; i < n; i++)
        if (in_mask(iigs, el[i])) { s += sqr(a[i] + -1.0 * a0[i]); cnt++; }
    if (b)
        for (int i = 0; i < n; i++)
            if (in_mask(iigs, el[i])) { s += sqr(b[i] + -1.0 * b0[i]); cnt++; }
    return sqrt(s / (double)(cnt > 1 ? cnt : 1));
}
static void gf_proj_avg(const int *el, int n, double *a)     /* AllInt, src/m_gridfunc.f90:1325-1370 */
{
    double avg = 0.0;
    int nval = 0;
    for (int i = 0; i < n; i++)
        if (el[i] >= ORC_ADHES) { avg += a[i]; nval++; }
    avg = avg / (double)imax(1, nval);
    for (int i = 0; i < n; i++)
        if (el[i] >= ORC_ADHES) a[i] -= avg;
}

/* ------------------------------------------------------------------------------------------ */
/* convolution, src/m_aijpj.f90:99-451, 712-1015                                                */
/* ------------------------------------------------------------------------------------------ */

static void inflcf_free(inflcf *c)
{
    free(c->cf);
    c->cf = NULL;
    for (int b = 0; b < 9; b++) { free(c->spec[b]); c->spec[b] = NULL; }
    c->defined = 0;
    c->sp_mx = c->sp_my = 0;
}

static void inflcf_new(inflcf *c, int mx, int my, double dx, double dy)
{
    inflcf_free(c);
    c->cf_mx = mx;
    c->cf_my = my;
    c->dx = dx;
    c->dy = dy;
    c->cf = dalloc((size_t)9 * (2 * mx) * (2 * my));
    c->defined = 1;
    c->nt_cpl = 1;
}

/* spectrum of block b at padded size lx x ly for products on an mx x my grid; the kernel is
   placed by wrap-around (offset d at index d mod L), which equals the reference's placement at
   fx+d (src/m_aijpj.f90:893-901) up to the half-period shift it applies to input and output */
static const cpx *block_spectrum(inflcf *c, int b, int mx, int my)
{
    int lx = 2 * orc_opt_fft_size(mx), ly = 2 * orc_opt_fft_size(my);
    if (c->sp_mx != mx || c->sp_my != my) {
        for (int k = 0; k < 9; k++) { free(c->spec[k]); c->spec[k] = NULL; }
        c->sp_mx = mx;
        c->sp_my = my;
        c->sp_lx = lx;
        c->sp_ly = ly;
    }
    if (!c->spec[b]) {
        cpx *a = (cpx *)calloc((size_t)lx * ly, sizeof(cpx));
        int fx = lx / 2, fy = ly / 2;
        for (int iy = -imin(fy, my); iy <= imin(fy, my) - 1; iy++)
            for (int ix = -imin(fx, mx); ix <= imin(fx, mx) - 1; ix++) {
                if (ix < -c->cf_mx || ix >= c->cf_mx || iy < -c->cf_my || iy >= c->cf_my) continue;
                int jx = (ix + lx) % lx, jy = (iy + ly) % ly;
                a[(size_t)jy * lx + jx].re =
                    c->cf[(size_t)b * (2 * c->cf_mx) * (2 * c->cf_my) + CFIDX(ix, iy, c->cf_mx, c->cf_my)];
            }
        fft2(lx, ly, -1, a);
        c->spec[b] = a;
    }
    return c->spec[b];
}

static double cf_at(const inflcf *c, int ik, int jk, int ix, int iy)
{
    if (ix < -c->cf_mx || ix >= c->cf_mx || iy < -c->cf_my || iy >= c->cf_my) return 0.0;
    return c->cf[(size_t)(ik * 3 + jk) * (2 * c->cf_mx) * (2 * c->cf_my) +
                 CFIDX(ix, iy, c->cf_mx, c->cf_my)];
}

/* u = sum_j cf(i - j) p(j) on the whole grid via FFT; result written (or added) where mask */
static void fft_block(orc_problem *P, inflcf *c, int ik, int jk, const double *p, double *u,
                      const int *el, int iigs, int ladd)
{
    int mx = P->mx, my = P->my;
    const cpx *cs = block_spectrum(c, ik * 3 + jk, mx, my);
    int lx = c->sp_lx, ly = c->sp_ly;
    cpx *a = (cpx *)calloc((size_t)lx * ly, sizeof(cpx));
    for (int iy = 0; iy < my; iy++)
        for (int ix = 0; ix < mx; ix++) a[(size_t)iy * lx + ix].re = p[iy * mx + ix];
    fft2(lx, ly, -1, a);
    for (size_t k = 0; k < (size_t)lx * ly; k++) a[k] = cmul(a[k], cs[k]);
    fft2(lx, ly, +1, a);
    double scale = 1.0 / ((double)lx * (double)ly);
    for (int iy = 0; iy < my; iy++)
        for (int ix = 0; ix < mx; ix++) {
            int ii = iy * mx + ix;
            if (!in_mask(iigs, el[ii])) continue;
            double v = a[(size_t)iy * lx + ix].re * scale;
            u[ii] = ladd ? u[ii] + v : v;
        }
    free(a);
}

/* gf3_VecAijPj with the FFT path, src/m_aijpj.f90:258-387; ikarg/jkarg in the reference's
   codes (1..3, -2 tangential, -3 all) */
static void vecaijpj(orc_problem *P, const int *el, int iigs, double **U, int ikarg, double **Pv,
                     int jkarg, inflcf *c)
{
    int ik0, ik1, jk0, jk1;
    if (ikarg == -3) { ik0 = 1; ik1 = 3; }
    else if (ikarg == -2) { ik0 = 1; ik1 = 2; }
    else { ik0 = ikarg; ik1 = ikarg; }
    if (jkarg == -3) { jk0 = 1; jk1 = 3; }
    else if (jkarg == -2) { jk0 = 1; jk1 = 2; }
    else { jk0 = jkarg; jk1 = jkarg; }
    int n = P->npot;
    for (int ik = ik0; ik <= ik1; ik++) {
        int ladd = 0;
        for (int jk = jk0; jk <= jk1; jk++) {
            if (!c->nt_cpl && (ik * jk == 3 || ik * jk == 6)) {
                if (!ladd)
                    for (int i = 0; i < n; i++)
                        if (in_mask(iigs, el[i])) U[ik - 1][i] = 0.0;
            } else {
                fft_block(P, c, ik - 1, jk - 1, Pv[jk - 1], U[ik - 1], el, iigs, ladd);
                ladd = 1;
            }
        }
    }
    for (int ik = ik0; ik <= ik1; ik++)
        for (int i = 0; i < n; i++)
            if (in_mask(iigs, el[i])) U[ik - 1][i] *= c->ga_inv;
}

/* gf3_AijPj: one element by direct summation over the row extents of the contact area,
   src/m_aijpj.f90:99-254 (ii, ik 1-based; pel = element division of the tractions) */
static double aijpj(const orc_problem *P, int ii, int ik, double **p, int jkarg, const inflcf *c,
                    const eldiv *pel)
{
    int jk0, jk1;
    if (jkarg == -3) {
        jk0 = 1; jk1 = 3;
        if (!c->nt_cpl && ik <= 2) jk1 = 2;
        if (!c->nt_cpl && ik == 3) jk0 = 3;
    } else if (jkarg == -2) {
        jk0 = 1; jk1 = 2;
        if (!c->nt_cpl && ik == 3) jk1 = 0;
    } else {
        jk0 = jkarg; jk1 = jkarg;
        if (!c->nt_cpl && ik <= 2) jk1 = imin(2, jk1);
        if (!c->nt_cpl && ik == 3) jk0 = 3;
    }
    int mx = P->mx, my = P->my;
    int ix = (ii - 1) % mx + 1, iy = (ii - 1) / mx + 1;
    double partsum = 0.0;
    for (int jy = 1; jy <= my; jy++)
        for (int jk = jk0; jk <= jk1; jk++) {
            int jh2 = (jy - 1) * mx;
            double rowsum = 0.0;
            for (int jx = imax(1, pel->row1st[jy - 1] - 1); jx <= imin(mx, pel->rowlst[jy - 1] + 1); jx++)
                rowsum += cf_at(c, ik - 1, jk - 1, ix - jx, iy - jy) * p[jk - 1][jx + jh2 - 1];
            partsum += rowsum;
        }
    return partsum * c->ga_inv;
}

void orc_conv_block(int mx, int my, const double *cfblock, const double *p, double *u, int method)
{
    if (method == 0) {
        for (int iy = 0; iy < my; iy++)
            for (int ix = 0; ix < mx; ix++) {
                double s = 0.0;
                for (int jy = 0; jy < my; jy++)
                    for (int jx = 0; jx < mx; jx++)
                        s += cfblock[CFIDX(ix - jx, iy - jy, mx, my)] * p[jy * mx + jx];
                u[iy * mx + ix] = s;
            }
        return;
    }
    orc_problem P;
    memset(&P, 0, sizeof P);
    P.mx = mx;
    P.my = my;
    P.npot = mx * my;
    inflcf c;
    memset(&c, 0, sizeof c);
    inflcf_new(&c, mx, my, 1.0, 1.0);
    memcpy(c.cf + 8 * (size_t)(2 * mx) * (2 * my), cfblock, sizeof(double) * (size_t)(2 * mx) * (2 * my));
    int *el = ialloc((size_t)mx * my);
    fft_block(&P, &c, 2, 2, p, u, el, ALLELM, 0);
    free(el);
    inflcf_free(&c);
}

/* fft_makePrec for diagonal block kk (0-based), src/m_aijpj.f90:457-708 */
static void make_prec(inflcf *c, int kk, inflcf *m)
{
    int fx = c->cf_mx, fy = c->cf_my, lx = 2 * fx, ly = 2 * fy;
    m->nt_cpl = c->nt_cpl;
    cpx *a = (cpx *)calloc((size_t)lx * ly, sizeof(cpx));
    for (int iy = -fy; iy <= fy - 1; iy++)
        for (int ix = -fx; ix <= fx - 1; ix++)
            a[(size_t)((iy + ly) % ly) * lx + (size_t)((ix + lx) % lx)].re = cf_at(c, kk, kk, ix, iy);
    fft2(lx, ly, -1, a);
    for (size_t k = 0; k < (size_t)lx * ly; k++) {
        double d = a[k].re * a[k].re + a[k].im * a[k].im;
        cpx inv = {a[k].re / d, -a[k].im / d};
        a[k] = inv;
    }
    fft2(lx, ly, +1, a);
    double scale = 1.0 / (4.0 * fx * fy);
    size_t nb = (size_t)(2 * m->cf_mx) * (2 * m->cf_my);
    for (int iy = -fy; iy <= fy - 1; iy++)
        for (int ix = -fx; ix <= fx - 1; ix++)
            m->cf[(size_t)(kk * 4) * nb + CFIDX(ix, iy, m->cf_mx, m->cf_my)] =
                c->ga * c->ga * a[(size_t)((iy + ly) % ly) * lx + (size_t)((ix + lx) % lx)].re * scale;
    free(a);
    /* the spectrum cache of m is stale now */
    free(m->spec[kk * 4]);
    m->spec[kk * 4] = NULL;
}

void orc_make_prec_block(int mx, int my, double ga, const double *cfblock, double *mblock)
{
    inflcf c, m;
    memset(&c, 0, sizeof c);
    memset(&m, 0, sizeof m);
    inflcf_new(&c, mx, my, 1.0, 1.0);
    inflcf_new(&m, mx, my, 1.0, 1.0);
    size_t nb = (size_t)(2 * mx) * (2 * my);
    memcpy(c.cf + 8 * nb, cfblock, sizeof(double) * nb);
    c.ga = ga;
    make_prec(&c, 2, &m);
    memcpy(mblock, m.cf + 8 * nb, sizeof(double) * nb);
    inflcf_free(&c);
    inflcf_free(&m);
}

/* ------------------------------------------------------------------------------------------ */
/* Hertz, src/m_hertz.f90:267-499 with complete elliptic integrals by Carlson's duplication     */
/* method (B.C. Carlson, Numer. Algorithms 10 (1995) 13-26) instead of ellip_bd's series        */
/* ------------------------------------------------------------------------------------------ */

static double carlson_rf(double x, double y, double z)
{
    double a0 = (x + y + z) / 3.0;
    double q = pow(3.0 * 1e-17, -1.0 / 6.0) * dmax(fabs(a0 - x), dmax(fabs(a0 - y), fabs(a0 - z)));
    double a = a0, f = 1.0;
    for (int it = 0; it < 60 && q * f >= fabs(a); it++) {
        double sx = sqrt(x), sy = sqrt(y), sz = sqrt(z);
        double lam = sx * sy + sx * sz + sy * sz;
        x = 0.25 * (x + lam);
        y = 0.25 * (y + lam);
        z = 0.25 * (z + lam);
        a = 0.25 * (a + lam);
        f *= 0.25;
    }
    double X = (a - x) / a, Y = (a - y) / a, Z = -(X + Y);
    double e2 = X * Y - Z * Z, e3 = X * Y * Z;
    return (1.0 - e2 / 10.0 + e3 / 14.0 + e2 * e2 / 24.0 - 3.0 * e2 * e3 / 44.0) / sqrt(a);
}

static double carlson_rd(double x, double y, double z)
{
    double a0 = (x + y + 3.0 * z) / 5.0;
    double q = pow(0.25 * 1e-17, -1.0 / 6.0) * dmax(fabs(a0 - x), dmax(fabs(a0 - y), fabs(a0 - z)));
    double a = a0, f = 1.0, sum = 0.0;
    for (int it = 0; it < 60 && q * f >= fabs(a); it++) {
        double sx = sqrt(x), sy = sqrt(y), sz = sqrt(z);
        double lam = sx * sy + sx * sz + sy * sz;
        sum += f / (sz * (z + lam));
        x = 0.25 * (x + lam);
        y = 0.25 * (y + lam);
        z = 0.25 * (z + lam);
        a = 0.25 * (a + lam);
        f *= 0.25;
    }
    double X = (a - x) / a, Y = (a - y) / a, Z = -(X + Y) / 3.0;
    double e2 = X * Y - 6.0 * Z * Z, e3 = (3.0 * X * Y - 8.0 * Z * Z) * Z;
    double e4 = 3.0 * (X * Y - Z * Z) * Z * Z, e5 = X * Y * Z * Z * Z;
    double s = 1.0 - 3.0 * e2 / 14.0 + e3 / 6.0 + 9.0 * e2 * e2 / 88.0 - 3.0 * e4 / 22.0 -
               9.0 * e2 * e3 / 52.0 + 3.0 * e5 / 26.0;
    return f * s / (a * sqrt(a)) + 3.0 * sum;
}

/* B(m) = int cos^2/Delta, D(m) = int sin^2/Delta with mc = 1-m; K = B + D, E = B + mc D */
void orc_ellip_bd(double mc, double *b, double *d)
{
    double m = 1.0 - mc;
    if (m < 1.11e-16) { *b = PI / 4.0; *d = PI / 4.0; return; }
    double mcc = mc < 1.11e-32 ? 1.11e-32 : mc;
    double kk = carlson_rf(0.0, mcc, 1.0);
    *d = carlson_rd(0.0, mcc, 1.0) / 3.0;
    *b = kk - *d;
}

/* eli, src/m_hertz.f90:446-499 */
static double eli(double elmodk, double *ell_k, double *ell_e)
{
    double mc = 1.0 - elmodk * elmodk, b, d;
    orc_ellip_bd(mc, &b, &d);
    *ell_k = b + d;
    *ell_e = b + mc * d;
    return (*ell_e - b) / b;
}

/* bisec, src/m_hertz.f90:386-442 */
static int bisec(double *ellipk, double *ellipe, double epshz, double *x, double xlft, double xrght,
                 double y)
{
    double xl = xlft, xr = xrght;
    double elr = eli(xr, ellipk, ellipe), ell = eli(xl, ellipk, ellipe);
    *x = 0.5 * (xl + xr);
    while (fabs(xr - xl) > epshz) {
        if ((y - elr) * (y - ell) > 0.0) return -1;
        *x = 0.5 * (xl + xr);
        double elx = eli(*x, ellipk, ellipe);
        if ((y - elr) * (y - elx) <= 0.0) { xl = *x; ell = elx; }
        else { xr = *x; elr = elx; }
    }
    eli(*x, ellipk, ellipe);
    return 0;
}

/* hzcalc3d, src/m_hertz.f90:267-384 */
static int hzcalc3d(double e_star, double epshz, int ipotcn, double *a1, double *b1, double *aa,
                    double *bb, int ic_norm, double *pen, double *fntrue, double *cp, double *rho)
{
    const double penmin = 1e-9;
    double ellipk, ellipe, g, sqrtg, x, y;
    if (ipotcn == -1) {
        if (dmin(*a1, *b1) < TINY) return -1;
        int zbla = (*b1 <= *a1);
        y = zbla ? *b1 / *a1 : *a1 / *b1;
        if (bisec(&ellipk, &ellipe, epshz, &x, 0.0, 1.0, y) != 0) return -1;
        double elmodk = x;
        g = sqrt(dmax(1e-40, 1.0 - elmodk * elmodk));
        sqrtg = sqrt(g);
        *rho = 2.0 / (*a1 + *b1);
        if (ic_norm == 1) {
            *cp = pow((3.0 * dmax(0.0, *fntrue) * *rho * ellipe) / (4.0 * PI * e_star * sqrtg), 1.0 / 3.0);
            *pen = 2.0 * sqr(*cp * sqrtg) * ellipk / (*rho * ellipe);
        } else {
            *cp = sqrt((dmax(0.0, *pen) * *rho * ellipe) / (2.0 * ellipk * sqrtg * sqrtg));
            *fntrue = 4.0 * PI * pow(*cp, 3) * e_star * sqrtg / (3.0 * *rho * ellipe);
        }
        if (zbla) { *aa = *cp * sqrtg; *bb = *cp / sqrtg; }
        else { *aa = *cp / sqrtg; *bb = *cp * sqrtg; }
    } else {
        int zbla = (*bb <= *aa);
        g = zbla ? *bb / *aa : *aa / *bb;
        sqrtg = sqrt(g);
        double elmodk = sqrt(1.0 - g * g);
        y = eli(elmodk, &ellipk, &ellipe);
        *cp = sqrt(*aa * *bb);
        if (ic_norm == 1) {
            *rho = 4.0 * PI * pow(*cp, 3) * e_star * sqrtg / (3.0 * *fntrue * ellipe);
            *pen = 2.0 * sqr(*cp * sqrtg) * ellipk / (*rho * ellipe);
        } else {
            *pen = dmax(*pen, penmin);
            *rho = 2.0 * sqr(*cp * sqrtg) * ellipk / (*pen * ellipe);
            *fntrue = 4.0 * PI * pow(*cp, 3) * e_star * sqrtg / (3.0 * *rho * ellipe);
        }
        double apb = 2.0 / *rho, ama = apb / (y + 1.0), ami = y * ama;
        if (zbla) { *a1 = ami; *b1 = ama; }
        else { *a1 = ama; *b1 = ami; }
    }
    return 0;
}

/* elliptic_to_curv, src/m_hertz.f90:95-213: the curvature b1 at which the Hertz ellipse of
   curvature a1 has the semi-axis ratio aob (halving/doubling from b1 = a1, then regula falsi) */
static int elliptic_to_curv(double e_star, double epshz, int ic_norm, double *a1, double aob,
                            double *b1, double *aa, double *bb, double *pen, double *fntrue,
                            double *cp, double *rho)
{
    int iter = 0;
    double epsaob = 10.0 * epshz, b1low, b1high, aoblow, aobhigh;
    *b1 = *a1;
    if (hzcalc3d(e_star, epshz, -1, a1, b1, aa, bb, ic_norm, pen, fntrue, cp, rho) != 0) return -1;
    if (*aa < TINY && *bb < TINY) return 0;
    if (*aa / *bb > aob) {
        b1high = *b1;
        aobhigh = *aa / *bb;
        while (iter <= 100 && *aa / *bb > aob) {
            iter++;
            b1high = *b1;
            aobhigh = *aa / *bb;
            *b1 = *b1 * 0.5;
            if (hzcalc3d(e_star, epshz, -1, a1, b1, aa, bb, ic_norm, pen, fntrue, cp, rho) != 0) return -1;
        }
        b1low = *b1;
        aoblow = *aa / *bb;
    } else {
        b1low = *b1;
        aoblow = *aa / *bb;
        while (iter <= 100 && *aa / *bb < aob) {
            iter++;
            b1low = *b1;
            aoblow = *aa / *bb;
            *b1 = *b1 * 2.0;
            if (hzcalc3d(e_star, epshz, -1, a1, b1, aa, bb, ic_norm, pen, fntrue, cp, rho) != 0) return -1;
        }
        b1high = *b1;
        aobhigh = *aa / *bb;
    }
    while (iter <= 100 && fabs(*aa / *bb - aob) > epsaob * aob) {
        iter++;
        double daob = aob - aoblow;
        double rico = (aobhigh - aoblow) / (b1high - b1low);
        *b1 = b1low + daob / rico;
        if (hzcalc3d(e_star, epshz, -1, a1, b1, aa, bb, ic_norm, pen, fntrue, cp, rho) != 0) return -1;
        if (*aa / *bb > aob) {
            b1high = *b1;
            aobhigh = *aa / *bb;
        } else {
            b1low = *b1;
            aoblow = *aa / *bb;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* defaults                                                                                     */
/* ------------------------------------------------------------------------------------------ */

void orc_default_input(orc_input *inp)
{
    memset(inp, 0, sizeof *inp);
    /* module-3 defaults: src/m_hierarch_data.f90:966-1014, src/m_global_data.f90:336-350 */
    inp->tang = 3;
    inp->norm = 0;
    inp->pvtime = 2;
    inp->iestim = 0;
    inp->gausei = 0;
    inp->omegah = (double)0.90f;    /* src/m_hierarch_data.f90:1736-1737 */
    inp->omegas = 1.0;
    inp->poiss1 = inp->poiss2 = 0.28;
    inp->gg1 = inp->gg2 = 1.0;
    inp->fstat = 0.30;
    inp->ipotcn = 3;
    inp->mx = inp->my = 1;
    inp->dx = inp->dy = 1.0;
    inp->hz_scale = 1.0;
    inp->ibase = 1;
    inp->veloc = 1e3;
    inp->dq = -1.0;     /* < 0: dq = dx */
    /* solver defaults: src/m_hierarch_data.f90:1728-1747 (single-precision literals) */
    inp->maxgs = 999;
    inp->maxin = 20;
    inp->maxnr = 25;
    inp->maxout = 1;
    inp->eps = 1e-5;
    inp->fdecay = sp(0.90f);
    inp->betath = sp(0.10f);
    inp->kdowfb = 1;
    inp->d_ifc = 1.0;
    inp->d_lin = 1.0;
    inp->d_cns = 1.0;
    inp->d_slp = 1.0;
    inp->pow_s = sp(0.90f);
    inp->sens = 1;          /* src/m_hierarch_data.f90:981 */
    inp->mxsens = 30;       /* src/m_hierarch_data.f90:1733-1735 */
    inp->epsens = 1e-4;
}

/* gd_meth/kdown as derived by cntc_setsolverflags for G=5 (src/contact_addon.f90:1240-1247);
   solv_init's gd_meth=1, kdown=3 otherwise */
static void gd_method(const orc_input *inp, int *gd_meth, int *kdown)
{
    if (inp->gausei != 5) { *gd_meth = 1; *kdown = 3; return; }
    if (inp->fdecay >= 0.999) { *gd_meth = 1; *kdown = 3; }
    else if (inp->fdecay <= 0.001) { *gd_meth = 2; *kdown = imax(1, nint(-inp->fdecay)); }
    else { *gd_meth = 3; *kdown = 3; }
}

/* ------------------------------------------------------------------------------------------ */
/* NormCG, src/m_solvpn.f90:24-461                                                              */
/* ------------------------------------------------------------------------------------------ */

static int normcg(orc_problem *P, int ic_norm, int maxcg, double eps, const double *hstot,
                  double *pen, double fntrue, int *itcg_out, double *err_out)
{
    int npot = P->npot;
    int *el = P->igs.el;
    double *ps = P->ps[2];
    double dxdy = P->dxdy;
    int numinn = npot <= 150 ? 3 : (npot <= 400 ? 2 : 1);
    double *rhs = dalloc(npot), *res = dalloc(npot), *r_prv = dalloc(npot), *dd = dalloc(npot),
           *z = dalloc(npot), *v = dalloc(npot), *q = dalloc(npot);
    double *Z3[3] = {NULL, NULL, z}, *R3[3] = {NULL, NULL, res}, *Q3[3] = {NULL, NULL, q},
           *V3[3] = {NULL, NULL, v}, *D3[3] = {NULL, NULL, dd}, *P3[3] = {NULL, NULL, ps};
    int itcg = 0, ret = 0;
    double err = 0.0, davg = 0.0;

    if (ic_norm == 1) *pen = 0.0;
    for (int i = 0; i < npot; i++) rhs[i] = *pen - hstot[i];
    int ncon = 0;
    for (int i = 0; i < npot; i++)
        if (el[i] >= ORC_ADHES) ncon++;

    if (ic_norm == 0) {
        double hsmin = hstot[0];
        for (int i = 1; i < npot; i++) hsmin = dmin(hsmin, hstot[i]);
        hsmin -= *pen;
        if (hsmin >= 0.0) {
            for (int i = 0; i < npot; i++) { ps[i] = 0.0; el[i] = ORC_EXTER; }
            *itcg_out = 0;
            *err_out = 0.0;
            goto done;
        }
    } else {
        if (ncon <= 0) {
            double hsmin = hstot[0], hsmax = hstot[0];
            for (int i = 1; i < npot; i++) { hsmin = dmin(hsmin, hstot[i]); hsmax = dmax(hsmax, hstot[i]); }
            double htrsh = hsmin + 0.1 * dmax(hsmax - hsmin, 1e-10);
            for (int i = 0; i < npot; i++)
                if (hstot[i] < htrsh) { el[i] = ORC_ADHES; ncon++; }
            areas(&P->igs, P->mx, P->my);
        }
        double fk = dxdy * gf_sum(ALLELM, el, npot, ps);
        if (fabs(fk) < sp(1e-3f) * fntrue) {
            double pn = fntrue / (dxdy * (double)ncon);
            for (int i = 0; i < npot; i++)
                if (el[i] >= ORC_ADHES) ps[i] = pn;
        } else {
            for (int i = 0; i < npot; i++)
                if (el[i] >= ORC_ADHES) ps[i] *= fntrue / fk;
        }
    }

    /* res = rhs - A ps on the interior */
    vecaijpj(P, el, ALLINT, R3, 3, P3, 3, &P->cs);
    for (int i = 0; i < npot; i++)
        if (el[i] >= ORC_ADHES) res[i] = -res[i] + rhs[i];
    if (ic_norm == 1) gf_proj_avg(el, npot, res);

    double rz1 = 0.0, rz2 = 0.0, rms_xk = 1.0, rms_upd = 2.0 * eps * 1.0, rms_upd1 = 0.0;
    double alpha = 0.0, beta;
    int itinn = 0, lchanged = 0;
    while ((lchanged || rms_upd > eps * rms_xk) && itcg < maxcg) {
        itcg++;
        itinn++;
        vecaijpj(P, el, ALLINT, Z3, 3, R3, 3, &P->ms);
        if (ic_norm == 1) gf_proj_avg(el, npot, z);
        rz1 = rz2;
        rz2 = gf_dot(ALLINT, el, npot, z, res);
        if (itcg <= 1 || rz1 < TINY) {
            for (int i = 0; i < npot; i++)
                if (el[i] >= ORC_ADHES) v[i] = z[i];
        } else {
            double rrprv = gf_dot(ALLINT, el, npot, z, r_prv);
            beta = dmax(0.0, (rz2 - rrprv) / dmax(TINY, rz1));
            for (int i = 0; i < npot; i++)
                if (el[i] >= ORC_ADHES) v[i] = beta * v[i] + z[i];
        }
        if (ic_norm == 1) gf_proj_avg(el, npot, v);
        vecaijpj(P, el, ALLINT, Q3, 3, V3, 3, &P->cs);
        if (ic_norm == 1) gf_proj_avg(el, npot, q);
        double rv = gf_dot(ALLINT, el, npot, res, v);
        double vav = gf_dot(ALLINT, el, npot, q, v);
        if (fabs(vav) > 1e-32 && ncon == 1) alpha = rv / vav;
        else alpha = rv / dmax(TINY, vav);
        for (int i = 0; i < npot; i++)
            if (el[i] >= ORC_ADHES) ps[i] += alpha * v[i];
        rms_upd = fabs(alpha) * gf_rms(ALLINT, el, npot, v, NULL);
        if (itcg == 1) rms_upd1 = rms_upd;
        if (itcg <= 3 || itcg % 10 == 0)
            rms_xk = P->p0[2] ? gf_rms_diff(ALLINT, el, npot, ps, P->p0[2], NULL, NULL)
                              : gf_rms(ALLINT, el, npot, ps, NULL);
        memcpy(r_prv, res, sizeof(double) * npot);

        if (itinn < numinn && rms_upd >= eps * rms_xk) {
            for (int i = 0; i < npot; i++)
                if (el[i] >= ORC_ADHES) res[i] -= alpha * q[i];
        } else {
            int lchg_negpn = 0;
            for (int i = 0; i < npot; i++)
                if (el[i] >= ORC_ADHES && ps[i] < 0.0) {
                    el[i] = ORC_EXTER;
                    ncon--;
                    ps[i] = 0.0;
                    lchg_negpn = 1;
                }
            if (ncon <= 0) {
                double hsmin = hstot[0];
                for (int i = 1; i < npot; i++) hsmin = dmin(hsmin, hstot[i]);
                for (int i = 0; i < npot; i++)
                    if (hstot[i] <= hsmin + 1e-5) {
                        el[i] = ORC_ADHES;
                        ncon++;
                        ps[i] = 0.0;
                        lchg_negpn = 1;
                    }
            }
            if (ic_norm == 1 && lchg_negpn) {
                double fk = dxdy * gf_sum(ALLELM, el, npot, ps);
                if (fabs(fk) < sp(1e-3f) * fntrue) {
                    for (int i = 0; i < npot; i++)
                        if (el[i] >= ORC_ADHES) ps[i] = 1.0;
                    fk = gf_sum(ALLELM, el, npot, ps);
                }
                for (int i = 0; i < npot; i++)
                    if (el[i] >= ORC_ADHES) ps[i] *= fntrue / fk;
            }
            if (lchg_negpn) areas(&P->igs, P->mx, P->my);
            vecaijpj(P, el, ALLELM, D3, 3, P3, 3, &P->cs);
            for (int i = 0; i < npot; i++) dd[i] -= rhs[i];
            if (ic_norm == 1) {
                davg = gf_sum(ALLINT, el, npot, dd) / (double)ncon;
                gf_proj_avg(el, npot, dd);
            }
            for (int i = 0; i < npot; i++) res[i] = (el[i] >= ORC_ADHES) ? -dd[i] : 0.0;
            int lchg_intpen = 0;
            for (int i = 0; i < npot; i++)
                if (el[i] == ORC_EXTER && dd[i] - davg < 0.0) {
                    el[i] = ORC_ADHES;
                    ncon++;
                    res[i] = -(dd[i] - davg);
                    lchg_intpen = 1;
                }
            for (int i = 0; i < npot; i++)
                if (el[i] == ORC_EXTER) v[i] = 0.0;
            if (lchg_intpen) areas(&P->igs, P->mx, P->my);
            itinn = 0;
            lchanged = lchg_intpen || lchg_negpn;
        }
    }
    if (ic_norm == 1) *pen = davg;
    {
        double conv = 1.0;
        if (rms_upd * rms_upd1 > 0.0 && itcg > 1) conv = exp(log(rms_upd / rms_upd1) / (itcg - 1));
        if (rms_upd > rms_xk && conv > 1.0 && itcg >= maxcg) ret = CNTC_ERR_NORM; /* abort_run */
    }
    err = rms_upd;
    *itcg_out = itcg;
    *err_out = err;
done:
    free(rhs); free(res); free(r_prv); free(dd); free(z); free(v); free(q);
    return ret;
}

/* sens_norm, src/m_solvpn.f90:531-650: dFn/dpen by one NormCG at a perturbed penetration
   (N=0, maxit mxsens, eps epsens), started from the converged pressures and element
   division, which are restored afterwards */
static int sens_norm(orc_problem *P, const orc_input *inp, const double *hstot, const double *unn)
{
    int npot = P->npot, *el = P->igs.el;
    double *ps = P->ps[2];
    double pentru = -1e20;                          /* gf3_max(AllInt, unn), :571 */
    for (int i = 0; i < npot; i++)
        if (el[i] >= ORC_ADHES && unn[i] > pentru) pentru = unn[i];
    double dpen = 0.0067 * pentru;
    double penloc = P->pen + dpen;
    int *el0 = ialloc(npot);
    double *ps0 = dalloc(npot);
    memcpy(el0, el, sizeof(int) * npot);
    memcpy(ps0, ps, sizeof(double) * npot);
    P->p0[2] = ps0;
    int it = 0;
    double err = 0.0;
    int rc = normcg(P, 0, inp->mxsens, inp->epsens, hstot, &penloc, P->fntrue, &it, &err);
    P->p0[2] = NULL;
    double fnloc = P->dxdy * gf_sum(ALLELM, el, npot, ps);
    SENS(P, IOUT_FN, IIN_PEN) = (fnloc - P->fntrue) / dpen;
    memcpy(el, el0, sizeof(int) * npot);
    memcpy(ps, ps0, sizeof(double) * npot);
    areas(&P->igs, P->mx, P->my);
    free(el0);
    free(ps0);
    return rc;
}

/* ------------------------------------------------------------------------------------------ */
/* snorm, src/m_snorm.f90:31-378                                                                */
/* ------------------------------------------------------------------------------------------ */

static int snorm(orc_problem *P, const orc_input *inp, int *itnorm_out)
{
    int npot = P->npot;
    int *el = P->igs.el;
    double *ps = P->ps[2];
    areas(&P->igs, P->mx, P->my);
    int use_vecfft = npot >= 700;
    make_prec(&P->cs, 2, &P->ms);

    double *hstot = dalloc(npot), *htang = dalloc(npot), *tmp = dalloc(npot), *unn = dalloc(npot);
    memcpy(hstot, P->hs[2], sizeof(double) * npot);
    if (P->cs.nt_cpl) {
        double *H3[3] = {NULL, NULL, htang};
        vecaijpj(P, el, ALLELM, H3, 3, P->ps, -2, &P->cs);
        for (int i = 0; i < npot; i++) hstot[i] += htang[i];
    }
    int itnorm = 0, itcg = 0, zready, ret = 0;
    int smllpn = 0, smlldd = 0;
    do {
        itnorm++;
        zready = 1;
        for (int i = 0; i < npot; i++)
            if (el[i] < ORC_ADHES) ps[i] = 0.0;
        areas(&P->igs, P->mx, P->my);
        int it;
        double errpn;
        int rc = normcg(P, inp->norm, inp->maxgs, inp->eps, hstot, &P->pen, P->fntrue, &it, &errpn);
        if (rc) { ret = rc; itcg += it; break; }
        itcg += it;
        int newext = 0;
        smllpn = 0;
        for (int i = 0; i < npot; i++)
            if (el[i] >= ORC_ADHES) {
                if (ps[i] < -errpn) { el[i] = ORC_EXTER; newext++; ps[i] = 0.0; }
                if (fabs(ps[i]) < errpn) smllpn++;
            }
        if (newext > 0) zready = 0;
        if (zready) {
            areas(&P->igs, P->mx, P->my);
            double *U3[3] = {NULL, NULL, unn};
            if (use_vecfft) vecaijpj(P, el, ALLELM, U3, 3, P->ps, 3, &P->cs);
            int ix = imax(1, P->mx / 2), iy = imax(1, P->my / 2);
            int ii = ix + (iy - 1) * P->mx;
            for (int i = 0; i < npot; i++) tmp[i] = 1.0;
            double *T3[3] = {NULL, NULL, tmp};
            double tol = fabs(errpn * aijpj(P, ii, 3, T3, 3, &P->cs, &P->igs));
            smlldd = 0;
            int newcon = 0;
            for (int i = 0; i < npot; i++)
                if (el[i] == ORC_EXTER && hstot[i] - P->pen < 0.0) {
                    double dd;
                    if (use_vecfft) dd = hstot[i] - P->pen + unn[i];
                    else dd = hstot[i] - P->pen + aijpj(P, i + 1, 3, P->ps, 3, &P->cs, &P->igs);
                    if (dd < -tol) { el[i] = ORC_ADHES; newcon++; }
                    if (fabs(dd) < tol) smlldd++;
                }
            if (newcon != 0) zready = 0;
        }
        if (it >= inp->maxgs) zready = 0;
    } while (!zready && itnorm < inp->maxin);
    if (!zready && !ret) itnorm = -1;
    for (int i = 0; i < npot; i++)
        if (ps[i] < 0.0 && el[i] >= ORC_ADHES) { ps[i] = 0.0; el[i] = ORC_EXTER; }
    P->itcg = itcg;
    areas(&P->igs, P->mx, P->my);
    if (inp->norm == 0) P->fntrue = P->dxdy * gf_sum(ALLELM, el, npot, ps);
    /* sensitivity dFn/dpen (:357-368); below 700 elements unn is formed here (:359-361) */
    if (!ret && inp->sens >= 2) {
        if (!use_vecfft) {
            double *U3[3] = {NULL, NULL, unn};
            vecaijpj(P, el, ALLELM, U3, 3, P->ps, 3, &P->cs);
        }
        ret = sens_norm(P, inp, hstot, unn);
    } else {
        SENS(P, IOUT_FN, IIN_PEN) = 0.0;
    }
    free(hstot); free(htang); free(tmp); free(unn);
    *itnorm_out = itnorm;
    return ret;
}

/* ------------------------------------------------------------------------------------------ */
/* elmtrc / plstrc without plasticity, src/m_solvpt.f90:3258-3807                               */
/* ------------------------------------------------------------------------------------------ */

/* With tau_c0 = taucv = 1e20 and k_tau = 0 the plastic branch always reports a violated
   constraint (taucs = 1e20 >= mu*pn) and only moves the state on to Adhes; its other results
   are overwritten by the next branch. */
static void elmtrc(int *el, const double coef[2][2], double eps, double omegah, double omegas,
                   double pr[3], double mus, double s[2])
{
    const int maxnr = 10;
    const double prc = 0.0001;
    const double taucs = 1e20;
    double pold1 = pr[0], pold2 = pr[1];
    double si0[2], si[2] = {0.0, 0.0}, dp[2], trc_bound = 0.0;
    si0[0] = s[0] - coef[0][0] * pr[0] - coef[0][1] * pr[1];
    si0[1] = s[1] - coef[1][0] * pr[0] - coef[1][1] * pr[1];
    int itry = 0, violat = 1;
    while (itry <= 3 && violat == 1) {
        itry++;
        if (*el == ORC_ADHES) {
            trc_bound = mus * pr[2];
            double detinv = 1.0 / (coef[0][0] * coef[1][1] - coef[0][1] * coef[1][0]);
            dp[0] = (-coef[1][1] * s[0] + coef[0][1] * s[1]) * detinv;
            dp[1] = (coef[1][0] * s[0] - coef[0][0] * s[1]) * detinv;
            pr[0] = pold1 + omegah * dp[0];
            pr[1] = pold2 + omegah * dp[1];
            double pabs2 = sqrt(pr[0] * pr[0] + pr[1] * pr[1]);
            if (pabs2 <= dmin(trc_bound, taucs)) {
                violat = 0;
            } else {
                double pabs1 = sqrt(sqr(pold1 + dp[0]) + sqr(pold2 + dp[1]));
                if (pabs1 <= dmin(trc_bound, taucs)) {
                    pr[0] = pr[0] * dmin(trc_bound, taucs) / pabs2;
                    pr[1] = pr[1] * dmin(trc_bound, taucs) / pabs2;
                    violat = 0;
                } else {
                    violat = 1;
                }
            }
            si[0] = si0[0] + coef[0][0] * pr[0] + coef[0][1] * pr[1];
            si[1] = si0[1] + coef[1][0] * pr[0] + coef[1][1] * pr[1];
        } else if (*el == ORC_SLIP) {
            violat = 0;
            trc_bound = mus * pr[2];
            double sabs = sqrt(si0[0] * si0[0] + si0[1] * si0[1]);
            pr[0] = -trc_bound * si0[0] / sabs;
            pr[1] = -trc_bound * si0[1] / sabs;
            si[0] = si0[0] + coef[0][0] * pr[0] + coef[0][1] * pr[1];
            si[1] = si0[1] + coef[1][0] * pr[0] + coef[1][1] * pr[1];
            double fk[2], gradf[2][2];
            fk[0] = pr[1] * si[0] - pr[0] * si[1];
            fk[1] = pr[0] * pr[0] + pr[1] * pr[1] - trc_bound * trc_bound;
            int itnr = 0;
            while (itnr == 0 || (itnr < maxnr && fabs(fk[0]) + fabs(fk[1]) >= prc * eps * trc_bound)) {
                itnr++;
                gradf[0][0] = -2.0 * coef[1][0] * pr[0] - si0[1] + (coef[0][0] - coef[1][1]) * pr[1];
                gradf[0][1] = 2.0 * coef[0][1] * pr[1] + si0[0] + (coef[0][0] - coef[1][1]) * pr[0];
                gradf[1][0] = 2.0 * pr[0];
                gradf[1][1] = 2.0 * pr[1];
                double det = gradf[0][0] * gradf[1][1] - gradf[0][1] * gradf[1][0];
                if (det == 0.0) {
                    itnr = maxnr;
                } else {
                    dp[0] = -(gradf[1][1] * fk[0] - gradf[0][1] * fk[1]) / det;
                    dp[1] = -(-gradf[1][0] * fk[0] + gradf[0][0] * fk[1]) / det;
                    pr[0] += dp[0];
                    pr[1] += dp[1];
                }
                si[0] = si0[0] + coef[0][0] * pr[0] + coef[0][1] * pr[1];
                si[1] = si0[1] + coef[1][0] * pr[0] + coef[1][1] * pr[1];
                fk[0] = pr[1] * si[0] - pr[0] * si[1];
                fk[1] = pr[0] * pr[0] + pr[1] * pr[1] - trc_bound * trc_bound;
            }
            double alph0 = atan2(pold2, pold1), alph1 = atan2(pr[1], pr[0]);
            double dalph = alph1 - alph0;
            if (dalph < -PI) dalph += 2.0 * PI;
            if (dalph > PI) dalph -= 2.0 * PI;
            alph1 = alph0 + omegas * dalph;
            pr[0] = trc_bound * cos(alph1);
            pr[1] = trc_bound * sin(alph1);
            if (fabs(pr[0]) > fabs(pr[1])) {
                if (pr[0] * si[0] > 0.0) violat = 1;
            } else {
                if (pr[1] * si[1] > 0.0) violat = 1;
            }
        } else {
            violat = 1;   /* plastic branch, see above */
        }
        if (itry <= 3 && violat == 1) {
            if (*el == ORC_ADHES) *el = ORC_SLIP;      /* k_tau=0, trc_bound <= taucs */
            else if (*el == ORC_SLIP) *el = ORC_PLAST;
            else *el = ORC_ADHES;
        }
    }
    s[0] = si[0];
    s[1] = si[1];
}

/* ------------------------------------------------------------------------------------------ */
/* GDsteady, src/gdsteady.f90                                                                   */
/* ------------------------------------------------------------------------------------------ */

typedef struct {
    int mx, my, npot;
    int gd_meth, kdown, kdowfb;
    double fdecay, betath, d_ifc, d_lin, d_cns, d_slp, pow_s;
} gdpar;

/* compute_dp, src/gdsteady.f90:610-635 */
static void compute_dp(int mx, int my, const double *px, const double *py, double *dpx, double *dpy)
{
    for (int iy = 0; iy < my; iy++) {
        int ii = mx - 1 + iy * mx;
        dpx[ii] = px[ii];
        dpy[ii] = py[ii];
        for (int ix = mx - 2; ix >= 0; ix--) {
            ii = ix + iy * mx;
            dpx[ii] = px[ii] - px[ii + 1];
            dpy[ii] = py[ii] - py[ii + 1];
        }
    }
}

/* apply_trcbnd, src/gdsteady.f90:807-847 */
static void apply_trcbnd(int mx, int my, const int *el, const double *g, double *dpx, double *dpy,
                         double *px, double *py)
{
    for (int iy = 0; iy < my; iy++) {
        int ii = mx - 1 + iy * mx;
        px[ii] = 0.0;
        py[ii] = 0.0;
        for (int ix = mx - 2; ix >= 0; ix--) {
            ii = ix + iy * mx;
            if (el[ii] <= ORC_EXTER) {
                px[ii] = 0.0;
                py[ii] = 0.0;
            } else {
                px[ii] = px[ii + 1] + dpx[ii];
                py[ii] = py[ii + 1] + dpy[ii];
                double ptabs = sqrt(px[ii] * px[ii] + py[ii] * py[ii]);
                if (el[ii] == ORC_SLIP || ptabs >= g[ii]) {
                    px[ii] = px[ii] * g[ii] / ptabs;
                    py[ii] = py[ii] * g[ii] / ptabs;
                }
            }
            dpx[ii] = px[ii] - px[ii + 1];
            dpy[ii] = py[ii] - py[ii + 1];
        }
    }
}

/* project_searchdir, src/gdsteady.f90:639-803 */
static void project_searchdir(const gdpar *G, const int *el, const double *g, int itgd, int *it_fb,
                              double betaj, const double *tx, const double *ty, double *dvx,
                              double *dvy, double *vx, double *vy, double fac_v)
{
    int mx = G->mx, my = G->my, imeth, kdown;
    if (betaj > G->betath || itgd - *it_fb <= 1) {
        imeth = G->gd_meth;
        kdown = G->kdown;
    } else {
        imeth = 2;
        kdown = G->kdowfb;
        *it_fb = itgd;
    }
    double *dvx_in = dalloc(mx), *dvy_in = dalloc(mx);
    for (int iy = 0; iy < my; iy++) {
        int ii0 = iy * mx;
        int ii = ii0 + mx - 1;
        vx[ii] = 0.0;
        vy[ii] = 0.0;
        if (imeth == 2) {
            memcpy(dvx_in, dvx + ii0, sizeof(double) * mx);
            memcpy(dvy_in, dvy + ii0, sizeof(double) * mx);
        }
        for (int ix = mx - 2; ix >= 0; ix--) {
            ii = ii0 + ix;
            if (imeth == 1 || imeth == 3) {
                if (el[ii] == ORC_ADHES) {
                    double f = imeth == 3 ? G->fdecay : 1.0;
                    if (imeth == 1) {
                        vx[ii] = vx[ii + 1] + dvx[ii];
                        vy[ii] = vy[ii + 1] + dvy[ii];
                    } else {
                        vx[ii] = f * vx[ii + 1] + dvx[ii];
                        vy[ii] = f * vy[ii + 1] + dvy[ii];
                        dvx[ii] = vx[ii] - vx[ii + 1];
                        dvy[ii] = vy[ii] - vy[ii + 1];
                    }
                } else if (el[ii] == ORC_SLIP) {
                    double vt = tx[ii] * dvx[ii] + ty[ii] * dvy[ii];
                    vt = fsign(1.0, vt) * dmin(fabs(vt), fac_v * g[ii]);
                    vx[ii] = tx[ii] * vt;
                    vy[ii] = ty[ii] * vt;
                    dvx[ii] = vx[ii] - vx[ii + 1];
                    dvy[ii] = vy[ii] - vy[ii + 1];
                } else {
                    vx[ii] = 0.0;
                    vy[ii] = 0.0;
                    dvx[ii] = -vx[ii + 1];
                    dvy[ii] = -vy[ii + 1];
                }
            } else {
                int ixf = ix + 1; /* 1-based */
                if (el[ii] == ORC_ADHES) {
                    if (ixf + kdown <= mx) {
                        vx[ii] = vx[ii + 1] + dvx_in[ix] - dvx_in[ix + kdown];
                        vy[ii] = vy[ii + 1] + dvy_in[ix] - dvy_in[ix + kdown];
                    } else {
                        vx[ii] = vx[ii + 1] + dvx_in[ix];
                        vy[ii] = vy[ii + 1] + dvy_in[ix];
                    }
                    dvx[ii] = vx[ii] - vx[ii + 1];
                    dvy[ii] = vy[ii] - vy[ii + 1];
                } else if (el[ii] == ORC_SLIP) {
                    double vt = tx[ii] * dvx_in[ix] + ty[ii] * dvy_in[ix];
                    vt = fsign(1.0, vt) * dmin(fabs(vt), fac_v * g[ii]);
                    vx[ii] = tx[ii] * vt;
                    vy[ii] = ty[ii] * vt;
                    dvx[ii] = vx[ii] - vx[ii + 1];
                    dvy[ii] = vy[ii] - vy[ii + 1];
                    for (int k = 1; k <= kdown; k++)
                        if (ixf + k <= mx) { dvx_in[ix + k] = 0.0; dvy_in[ix + k] = 0.0; }
                } else {
                    vx[ii] = 0.0;
                    vy[ii] = 0.0;
                    dvx[ii] = -vx[ii + 1];
                    dvy[ii] = -vy[ii + 1];
                    for (int k = 1; k <= kdown; k++)
                        if (ixf + k <= mx) { dvx_in[ix + k] = 0.0; dvy_in[ix + k] = 0.0; }
                }
            }
        }
    }
    free(dvx_in);
    free(dvy_in);
}

/* compute_diagscaling, src/gdsteady.f90:901-1042 (elnew/fac_s keep their values between
   elements, as the reference's locals do) */
static void compute_diagscaling(const gdpar *G, const int *el, const double *mus,
                                const double coef[2][2], double **ps, double **ss, double *dscl)
{
    const double tiny_err = 1e-6, epselm = 1e-6, omgscl = 1.0;
    int mx = G->mx;
    int elnew = ORC_EXTER;
    double fac_s = 0.0;
    for (int ii = 0; ii < G->npot; ii++) {
        if (el[ii] == ORC_SLIP) {
            double pr[3] = {ps[0][ii], ps[1][ii], ps[2][ii]};
            double si[2] = {ss[0][ii], ss[1][ii]};
            double th_p0 = atan2(pr[1], pr[0]);
            double th_s0 = atan2(-si[1], -si[0]);
            double err_0 = th_s0 - th_p0;
            if (fabs(err_0) >= PI) err_0 = err_0 - nint(err_0 / (2.0 * PI)) * 2.0 * PI;
            if (fabs(err_0) <= tiny_err) {
                pr[0] = mus[ii] * pr[2] * cos(th_p0 + 0.1);
                pr[1] = mus[ii] * pr[2] * sin(th_p0 + 0.1);
                si[0] = ss[0][ii] + coef[0][0] * (pr[0] - ps[0][ii]) + coef[0][1] * (pr[1] - ps[1][ii]);
                si[1] = ss[1][ii] + coef[1][0] * (pr[0] - ps[0][ii]) + coef[1][1] * (pr[1] - ps[1][ii]);
                th_p0 = atan2(pr[1], pr[0]);
                th_s0 = atan2(-si[1], -si[0]);
                err_0 = th_s0 - th_p0;
                if (fabs(err_0) >= PI) err_0 = err_0 - nint(err_0 / (2.0 * PI)) * 2.0 * PI;
            }
            if (fabs(err_0) > tiny_err) {
                elnew = el[ii];
                elmtrc(&elnew, coef, epselm, 1.0, 1.0, pr, mus[ii], si);
                if (elnew == ORC_SLIP) {
                    double th_p1 = atan2(pr[1], pr[0]);
                    double th_s1 = atan2(-si[1], -si[0]);
                    double upd_p = th_p1 - th_p0;
                    if (fabs(upd_p) >= PI) upd_p = upd_p - nint(upd_p / (2.0 * PI)) * 2.0 * PI;
                    double upd_s = th_s1 - th_s0;
                    if (fabs(upd_s) >= PI) upd_s = upd_s - nint(upd_s / (2.0 * PI)) * 2.0 * PI;
                    fac_s = fabs(upd_s) / fabs(err_0);
                    (void)upd_p;
                }
            }
        }
        if (el[ii] == ORC_SLIP && elnew == ORC_SLIP) {
            dscl[ii] = G->d_slp * pow(fac_s, G->pow_s);
        } else if (el[ii] == ORC_ADHES || elnew == ORC_ADHES) {
            int ix = ii % mx + 1;
            int jx = ix;
            while (jx > 1 && el[ii - ix + jx - 1] == ORC_ADHES) jx--;
            double dscl_new;
            if (G->d_cns < G->d_ifc)
                dscl_new = dmax(G->d_cns, G->d_ifc + (ix - jx - 1) * dmin(0.0, G->d_lin));
            else
                dscl_new = dmin(G->d_cns, G->d_ifc + (ix - jx - 1) * dmax(0.0, G->d_lin));
            dscl[ii] = dmin(50.0 * dscl[ii], omgscl * dscl_new + (1.0 - omgscl) * dscl[ii]);
        } else {
            dscl[ii] = 0.0;
        }
    }
}


/* perform_linesearch, src/gdsteady.f90:1046-1462.  Returns 0, or -1 when no bracket was found
   (the reference aborts the run there).  Tables are 1-based as in the reference. */
static int perform_linesearch(const gdpar *G, const int *el, const double *g, const double *dscl,
                              const double *dpx, const double *dpy, double **ss, const double *dvx,
                              const double *dvy, const double *vx, const double *vy,
                              const double *qx, const double *qy, const double *txold,
                              const double *tyold, double alpha0, double *alpha_out, int *j_out)
{
    enum { MAX_J = 99 };
    const double tiny = 1e-12, eps_j = 0.001;
    int mx = G->mx, my = G->my, npot = G->npot;
    double *dpnx = dalloc(npot), *dpny = dalloc(npot), *psnx = dalloc(npot), *psny = dalloc(npot);
    double tbl_alpha[MAX_J + 2], tbl_rho[MAX_J + 2], tbl_drhoda[MAX_J + 2];
    int j = 0, stop_j = 0, has_bracket = 0, prev_bracket;
    int itb_a = 0, itb_b = 0, itb_x = 0, itb_w = 0, itb_v = 0, itb_u;
    double alpha_j = 0.0, alpha_prv = -1.0, dst_cur = 0.0, dst_prv1 = 0.0, dst_prv2 = 0.0, ddrho;
    const double *sx = ss[0], *sy = ss[1];
    while (!stop_j) {
        double rhonew = 0.0, c0adh = 0.0, c1adh = 0.0, c0slp = 0.0, c1slp = 0.0, c2slp = 0.0;
        for (int i = 0; i < npot; i++) {
            dpnx[i] = dpx[i];
            dpny[i] = dpy[i];
            if (el[i] >= ORC_ADHES) {
                dpnx[i] += alpha_j * dvx[i];
                dpny[i] += alpha_j * dvy[i];
            }
        }
        apply_trcbnd(mx, my, el, g, dpnx, dpny, psnx, psny);
        for (int ii = 0; ii < npot; ii++) {
            double vt = vx[ii] * txold[ii] + vy[ii] * tyold[ii];
            double dth_da = (g[ii] * vt) / (dmax(tiny, g[ii] * g[ii]) + alpha_j * alpha_j * vt * vt);
            double theta = atan2(psny[ii], psnx[ii]);
            double nxn = cos(theta), nyn = sin(theta), txn = -nyn, tyn = nxn;
            double sn = sx[ii] * nxn + sy[ii] * nyn;
            int elnew = el[ii];
            if (elnew == ORC_SLIP && sn > 0.0) elnew = ORC_ADHES;
            double st = sx[ii] * txn + sy[ii] * tyn, rt = -st;
            double qn = qx[ii] * nxn + qy[ii] * nyn, qt = qx[ii] * txn + qy[ii] * tyn;
            double d2 = dscl[ii] * dscl[ii];
            if (elnew == ORC_ADHES) {
                c0adh += 2.0 * d2 * (sx[ii] * qx[ii] + sy[ii] * qy[ii]);
                c1adh += 2.0 * d2 * (qx[ii] * qx[ii] + qy[ii] * qy[ii]);
                rhonew += d2 * (sqr(sx[ii] + alpha_j * qx[ii]) + sqr(sy[ii] + alpha_j * qy[ii]));
            } else if (elnew == ORC_SLIP) {
                c0slp -= 2.0 * d2 * (rt * qt - rt * sn * dth_da);
                c1slp -= 2.0 * d2 * (-qt * qt + qt * sn * dth_da - rt * qn * dth_da);
                c2slp -= 2.0 * d2 * (qt * qn * dth_da);
                rhonew += d2 * sqr(st + alpha_j * qt);
            }
        }
        double drho_da = c0adh + c0slp + alpha_j * (c1adh + c1slp) + alpha_j * alpha_j * c2slp;

        /* insert into the table sorted on alpha */
        int itb = 1;
        while (itb <= j && alpha_j > tbl_alpha[itb]) itb++;
        itb_u = itb;
        for (int k = j; k >= itb; k--) {
            tbl_alpha[k + 1] = tbl_alpha[k];
            tbl_rho[k + 1] = tbl_rho[k];
            tbl_drhoda[k + 1] = tbl_drhoda[k];
        }
        if (has_bracket) {
            if (itb_a >= itb_u) itb_a++;
            if (itb_b >= itb_u) itb_b++;
            if (itb_x >= itb_u) itb_x++;
            if (itb_w >= itb_u) itb_w++;
            if (itb_v >= itb_u) itb_v++;
        }
        j++;
        tbl_alpha[itb] = alpha_j;
        tbl_rho[itb] = rhonew;
        tbl_drhoda[itb] = drho_da;

        prev_bracket = has_bracket;
        has_bracket = 0;
        int ibrack = 1;
        while (ibrack < j - 1 && !has_bracket) {
            has_bracket = tbl_rho[ibrack + 1] <= dmin(tbl_rho[ibrack], tbl_rho[ibrack + 2]);
            if (!has_bracket) ibrack++;
        }
        if (has_bracket && !prev_bracket) {
            itb_a = ibrack;
            itb_x = ibrack + 1;
            itb_b = ibrack + 2;
            if (tbl_rho[itb_a] < tbl_rho[itb_b]) { itb_w = itb_a; itb_v = itb_b; }
            else { itb_w = itb_b; itb_v = itb_a; }
            dst_prv1 = tbl_alpha[itb_x] - tbl_alpha[itb_w];
            dst_prv2 = tbl_alpha[itb_w] - tbl_alpha[itb_v];
        } else if (has_bracket) {
            if (tbl_rho[itb_u] <= tbl_rho[itb_x]) {
                if (tbl_alpha[itb_u] >= tbl_alpha[itb_x]) itb_a = itb_x;
                else itb_b = itb_x;
                itb_v = itb_w;
                itb_w = itb_x;
                itb_x = itb_u;
            } else {
                if (tbl_alpha[itb_u] < tbl_alpha[itb_x]) itb_a = itb_u;
                else itb_b = itb_u;
                if (tbl_rho[itb_u] <= tbl_rho[itb_w] || tbl_alpha[itb_w] == tbl_alpha[itb_x]) {
                    itb_v = itb_w;
                    itb_w = itb_u;
                } else if (tbl_rho[itb_u] <= tbl_rho[itb_v] || tbl_alpha[itb_v] == tbl_alpha[itb_x] ||
                           tbl_alpha[itb_v] == tbl_alpha[itb_w]) {
                    itb_v = itb_u;
                }
            }
            dst_prv2 = dst_prv1;
            dst_prv1 = dst_cur;
        }

        alpha_prv = alpha_j;
        if (has_bracket) {
            double xmid = 0.5 * (tbl_alpha[itb_a] + tbl_alpha[itb_b]);
            double tolx = eps_j * fabs(tbl_alpha[itb_x]) + tiny;
            int ldone = fabs(tbl_alpha[itb_x] - xmid) <= 2.0 * tolx - 0.5 * (tbl_alpha[itb_b] - tbl_alpha[itb_a]);
            if (!ldone) {
                int use_parab;
                if (fabs(dst_prv2) <= tolx) {
                    use_parab = 0;
                } else {
                    double br = (tbl_alpha[itb_x] - tbl_alpha[itb_w]) * (tbl_rho[itb_x] - tbl_rho[itb_v]);
                    double bq = (tbl_alpha[itb_x] - tbl_alpha[itb_v]) * (tbl_rho[itb_x] - tbl_rho[itb_w]);
                    double bp = (tbl_alpha[itb_x] - tbl_alpha[itb_v]) * bq - (tbl_alpha[itb_x] - tbl_alpha[itb_w]) * br;
                    bq = 2.0 * (bq - br);
                    if (bq > 0.0) bp = -bp;
                    bq = fabs(bq);
                    if (fabs(bp) >= fabs(0.5 * bq * dst_prv2) || bp <= bq * (tbl_alpha[itb_a] - tbl_alpha[itb_x]) ||
                        bp >= bq * (tbl_alpha[itb_b] - tbl_alpha[itb_x])) {
                        use_parab = 0;
                    } else {
                        use_parab = 1;
                        dst_cur = bp / bq;
                        alpha_j = tbl_alpha[itb_x] + dst_cur;
                        if (alpha_j - tbl_alpha[itb_a] < 2.0 * tolx || tbl_alpha[itb_b] - alpha_j < 2.0 * tolx)
                            dst_cur = tolx * fsign(1.0, xmid - tbl_alpha[itb_x]);
                    }
                }
                if (!use_parab) {
                    if (tbl_alpha[itb_x] >= xmid) dst_prv1 = tbl_alpha[itb_a] - tbl_alpha[itb_x];
                    else dst_prv1 = tbl_alpha[itb_b] - tbl_alpha[itb_x];
                    dst_cur = 0.381966 * dst_prv1;
                }
                alpha_j = tbl_alpha[itb_x] + dst_cur;
            }
        } else if (j == 1) {
            alpha_j = alpha_j + 0.6 * alpha0;
            ddrho = 10.0;
        } else if (tbl_rho[j] < tbl_rho[1]) {
            int it2 = j - 1;
            double dalph = tbl_alpha[it2 + 1] - tbl_alpha[it2];
            if (j < 3) {
                ddrho = (tbl_drhoda[it2 + 1] - tbl_drhoda[it2]) / dalph;
                alpha_j = tbl_alpha[it2 + 1] + dmin(3.0 * dalph, -tbl_drhoda[it2] / ddrho);
            } else {
                alpha_j = tbl_alpha[it2 + 1] + 3.0 * dalph;
            }
        } else {
            double dalph = tbl_alpha[2] - tbl_alpha[1];
            if (j < 3) {
                ddrho = (tbl_drhoda[2] - tbl_drhoda[1]) / dalph;
                alpha_j = tbl_alpha[1] + dmin(-3.0 * dalph, tbl_drhoda[1] / ddrho);
            } else {
                alpha_j = tbl_alpha[1] - 3.0 * dalph;
            }
        }
        stop_j = (j >= MAX_J || fabs(alpha_j - alpha_prv) < eps_j * dmax(fabs(alpha_j), fabs(alpha_prv)));
    }
    (void)ddrho;
    int ret = has_bracket ? 0 : -1;
    if (j >= MAX_J) {
        int best = 1;
        for (int k = 2; k <= MAX_J; k++)
            if (tbl_rho[k] < tbl_rho[best]) best = k;
        alpha_j = tbl_alpha[best];
    }
    free(dpnx); free(dpny); free(psnx); free(psny);
    *alpha_out = alpha_j;
    *j_out = j;
    return ret;
}

/* gdsteady, src/gdsteady.f90:9-606.  Returns 0 or CNTC_ERR_TANG (no bracket). */
static int gdsteady(orc_problem *P, const gdpar *G, int maxgd, double eps, double **ws,
                    const double *mus, int *itgd_out, double *err_out, int *lstagn)
{
    const double tiny = 1e-12, fac_v = 1000.0;
    int mx = P->mx, npot = P->npot, *el = P->igs.el;
    double *px = P->ps[0], *py = P->ps[1], *pn = P->ps[2];
    double *sx = P->ss[0], *sy = P->ss[1];
    double coefs[2][2];
    coefs[0][0] = P->cs.ga_inv * cf_at(&P->cs, 0, 0, 0, 0);
    coefs[0][1] = P->cs.ga_inv * cf_at(&P->cs, 0, 1, 0, 0);
    coefs[1][0] = P->cs.ga_inv * cf_at(&P->cs, 1, 0, 0, 0);
    coefs[1][1] = P->cs.ga_inv * cf_at(&P->cs, 1, 1, 0, 0);
    int *igsold = ialloc(npot), *igsopt = ialloc(npot);
    double *g = dalloc(npot), *dpx = dalloc(npot), *dpy = dalloc(npot), *dscl = dalloc(npot),
           *nxv = dalloc(npot), *nyv = dalloc(npot), *txv = dalloc(npot), *tyv = dalloc(npot),
           *rx = dalloc(npot), *ry = dalloc(npot), *dvx = dalloc(npot), *dvy = dalloc(npot),
           *vx = dalloc(npot), *vy = dalloc(npot), *qx = dalloc(npot), *qy = dalloc(npot),
           *dqx = dalloc(npot), *dqy = dalloc(npot), *poptx = dalloc(npot), *popty = dalloc(npot),
           *poldx = dalloc(npot), *poldy = dalloc(npot);
    double *DP[3] = {dpx, dpy, NULL}, *DV[3] = {dvx, dvy, NULL}, *Q[3] = {qx, qy, NULL};
    int nadh, nslip, nplast, nexter, ret = 0;
    count_el(el, npot, &nadh, &nslip, &nplast, &nexter);
    double facnel = (double)sqrtf((float)npot / (float)(nadh + nslip + nplast));
    for (int i = 0; i < npot; i++) g[i] = mus[i] * pn[i];
    for (int i = 0; i < npot; i++) dscl[i] = 1.0;

    int itgd = 0, it_fb = -99, lchanged = 0;
    double dif = 2.0, difid = 1.0, dif1 = 0.0, beta = 1.0, alpha = 0.0, alpha0 = 0.0, alpha1;
    *lstagn = 0;

    for (int i = 0; i < npot; i++) {
        double th = atan2(py[i], px[i]);
        nxv[i] = cos(th); nyv[i] = sin(th); txv[i] = -nyv[i]; tyv[i] = nxv[i];
    }
    compute_dp(mx, P->my, px, py, dpx, dpy);
    for (int i = 0; i < npot; i++) { sx[i] = 0.0; sy[i] = 0.0; }
    vecaijpj(P, el, ALLINT, P->ss, -2, DP, -2, &P->cs);
    for (int i = 0; i < npot; i++)
        if (el[i] >= ORC_ADHES) { sx[i] += ws[0][i]; sy[i] += ws[1][i]; }
    compute_diagscaling(G, el, mus, coefs, P->ps, P->ss, dscl);

#define RESIDUAL()                                                                                 \
    do {                                                                                           \
        for (int i = 0; i < npot; i++) {                                                           \
            rx[i] = 0.0; ry[i] = 0.0;                                                              \
            if (el[i] == ORC_ADHES) { rx[i] = -sx[i] * dscl[i]; ry[i] = -sy[i] * dscl[i]; }        \
            else if (el[i] == ORC_SLIP) {                                                          \
                double st = txv[i] * sx[i] + tyv[i] * sy[i];                                       \
                rx[i] = -st * txv[i] * dscl[i]; ry[i] = -st * tyv[i] * dscl[i];                    \
            }                                                                                      \
        }                                                                                          \
    } while (0)
    RESIDUAL();
    {
        double mxr = 0.0;
        for (int i = 0; i < npot; i++) mxr = dmax(mxr, dmax(fabs(rx[i]), fabs(ry[i])));
        if (mxr < tiny) dif = 0.0;
    }

    while ((lchanged || dif > difid) && itgd < maxgd) {
        itgd++;
        memcpy(dvx, rx, sizeof(double) * npot);
        memcpy(dvy, ry, sizeof(double) * npot);
        project_searchdir(G, el, g, itgd, &it_fb, beta, txv, tyv, dvx, dvy, vx, vy, fac_v);
        vecaijpj(P, el, ALLINT, Q, -2, DV, -2, &P->cs);
        if (itgd <= 1) {
            alpha0 = 0.0;
            alpha1 = 0.0;
            for (int i = 0; i < npot; i++) {
                alpha0 += rx[i] * rx[i] + ry[i] * ry[i];
                alpha1 += (rx[i] * qx[i] + ry[i] * qy[i]) * dscl[i];
            }
            alpha0 = alpha0 / alpha1;
        } else {
            alpha0 = 0.5 * (alpha0 + alpha);
        }
        int its_j;
        if (perform_linesearch(G, el, g, dscl, dpx, dpy, P->ss, dvx, dvy, vx, vy, qx, qy, txv, tyv,
                               alpha0, &alpha, &its_j) != 0) {
            ret = CNTC_ERR_TANG;
            break;
        }
        for (int i = 0; i < npot; i++) { dqx[i] = dscl[i] * qx[i]; dqy[i] = dscl[i] * qy[i]; }
        {
            double a = 0.0, b = 0.0;
            for (int i = 0; i < npot; i++)
                if (el[i] >= ORC_ADHES) { a += dqx[i] * dqx[i] + dqy[i] * dqy[i]; b += rx[i] * rx[i] + ry[i] * ry[i]; }
            beta = alpha * sqrt(a) / sqrt(b);
        }
        lchanged = 0;
        memcpy(igsold, el, sizeof(int) * npot);
        memcpy(poldx, px, sizeof(double) * npot);
        memcpy(poldy, py, sizeof(double) * npot);
        for (int i = 0; i < npot; i++)
            if (el[i] >= ORC_ADHES) { dpx[i] += alpha * dvx[i]; dpy[i] += alpha * dvy[i]; }
        apply_trcbnd(mx, P->my, el, g, dpx, dpy, px, py);
        for (int i = 0; i < npot; i++) {
            double th = atan2(py[i], px[i]);
            nxv[i] = cos(th); nyv[i] = sin(th); txv[i] = -nyv[i]; tyv[i] = nxv[i];
        }
        for (int i = 0; i < npot; i++) { sx[i] = 0.0; sy[i] = 0.0; }
        vecaijpj(P, el, ALLINT, P->ss, -2, DP, -2, &P->cs);
        for (int i = 0; i < npot; i++)
            if (el[i] >= ORC_ADHES) { sx[i] += ws[0][i]; sy[i] += ws[1][i]; }
        /* solve_elmtrc: only the elements that were in slip are used below */
        for (int i = 0; i < npot; i++) {
            if (igsold[i] != ORC_SLIP) continue;
            double pr[3] = {px[i], py[i], pn[i]}, si[2] = {sx[i], sy[i]};
            int elnew = el[i];
            elmtrc(&elnew, coefs, 1e-6, 1.0, 1.0, pr, mus[i], si);
            igsopt[i] = elnew;
            poptx[i] = pr[0];
            popty[i] = pr[1];
        }
        for (int i = 0; i < npot; i++)
            if (el[i] == ORC_ADHES) {
                double ptabs = sqrt(px[i] * px[i] + py[i] * py[i]);
                if (ptabs >= g[i] - tiny) { el[i] = ORC_SLIP; lchanged = 1; }
            }
        for (int i = 0; i < npot; i++)
            if (igsold[i] == ORC_SLIP) {
                double snrm = nxv[i] * sx[i] + nyv[i] * sy[i];
                double cosdth = (nxv[i] * poptx[i] + nyv[i] * popty[i]) / g[i];
                if ((snrm > 0.0 && (dscl[i] >= 0.001 || igsopt[i] == ORC_ADHES)) || cosdth <= -0.5) {
                    el[i] = ORC_ADHES;
                    lchanged = 1;
                }
            }
        compute_diagscaling(G, el, mus, coefs, P->ps, P->ss, dscl);
        RESIDUAL();
        for (int i = 0; i < npot; i++)
            if (el[i] >= ORC_ADHES) { poldx[i] -= px[i]; poldy[i] -= py[i]; }
        double facdif;
        if (beta >= 0.1) facdif = 1.0;
        else if (beta >= 0.001) facdif = 0.1 / beta;
        else facdif = 100.0;
        dif = facnel * gf_rms(ALLELM, el, npot, poldx, poldy) * facdif;
        difid = eps * dmax(1e-6, facnel * gf_rms(ALLELM, el, npot, px, py));
        if (itgd == 1) dif1 = dif;
    }
#undef RESIDUAL
    if (!ret) {
        double resrms = gf_rms(ALLINT, el, npot, rx, ry);
        double res_dp = resrms / dmax(coefs[0][0], coefs[1][1]);
        double conv = 1.0;
        if (dif * dif1 > 0.0 && itgd > 1) conv = exp(log(dif / dif1) / (itgd - 1));
        if (lchanged && itgd >= maxgd) *lstagn = 1;
        else if (dif > difid && conv > 1.0 && itgd >= maxgd) *lstagn = 1;
        else if (conv < 1.0 - tiny && res_dp > 5.0 * difid / (1.0 - conv)) { itgd = -itgd; *lstagn = 1; }
    }
    *itgd_out = itgd;
    *err_out = dif;
    free(igsold); free(igsopt); free(g); free(dpx); free(dpy); free(dscl); free(nxv); free(nyv);
    free(txv); free(tyv); free(rx); free(ry); free(dvx); free(dvy); free(vx); free(vy); free(qx);
    free(qy); free(dqx); free(dqy); free(poptx); free(popty); free(poldx); free(poldy);
    return ret;
}

/* ------------------------------------------------------------------------------------------ */
/* SteadyGS, src/m_solvpt.f90:2825-3254 (chi = 0, no plasticity, no third body)                 */
/* ------------------------------------------------------------------------------------------ */

/* Gauss-Seidel sweeps over the rows, each row from the trailing edge (ix = 1) to the leading
   edge (ix = mx).  Per interior element: the slip of the other elements by direct summation
   over the differences dp (gf3_AijPj), a local 2x2 problem (plstrc) with the coefficients of
   the element relative to the first upstream non-adhesion element, then the tractions of the
   upstream part of the row are rebuilt from dp with the traction bound applied. */
/* info: 0, or 1 when maxgs was reached, 2 when also the mean rate of convergence exceeds
   0.997, 3 when it exceeds 1 (:3205-3234) */
static int stdygs(orc_problem *P, double **ws, const double *mus, int k, double eps, int maxgs,
                  double omegah, double omegas, int *itgs_out, double *err_out, int *info)
{
    int mx = P->mx, my = P->my, npot = P->npot, *el = P->igs.el;
    double *px = P->ps[0], *py = P->ps[1], *pn = P->ps[2];
    double *sx = P->ss[0], *sy = P->ss[1];
    double *dpx = dalloc(npot), *dpy = dalloc(npot);
    double *DP[3] = {dpx, dpy, NULL};
    const inflcf *cs = &P->cs;
    /* dp = backward differences along x; the leading-edge column keeps p (:2893-2908) */
    for (int iy = 0; iy < my; iy++)
        for (int ix = 1; ix <= mx; ix++) {
            int ii = ix - 1 + iy * mx;
            if (ix != mx) { dpx[ii] = px[ii] - px[ii + 1]; dpy[ii] = py[ii] - py[ii + 1]; }
            else { dpx[ii] = px[ii]; dpy[ii] = py[ii]; }
        }
    int nadh, nslip, nplast, nexter;
    count_el(el, npot, &nadh, &nslip, &nplast, &nexter);
    double facnel = (double)sqrtf((float)npot / (float)(nadh + nslip + nplast));
    int itgs = 0;
    double dif = 2.0, difid = 1.0, dif1 = 0.0;
    while (dif >= difid && itgs < maxgs) {
        itgs++;
        dif = 0.0;
        for (int iy = 1; iy <= my; iy++) {
            for (int ix = 1; ix <= mx; ix++) {
                int ii = ix + (iy - 1) * mx;          /* 1-based */
                if (el[ii - 1] < ORC_ADHES) continue;
                /* first upstream element that is not in adhesion (:2972-2977) */
                int jx = ix - 1;
                while (jx >= 1 && el[jx + (iy - 1) * mx - 1] == ORC_ADHES && jx != 1) jx--;
                double coef[2][2];
                for (int a = 0; a < 2; a++)
                    for (int b = 0; b < 2; b++)
                        coef[a][b] = cs->ga_inv * (cf_at(cs, a, b, 0, 0) - cf_at(cs, a, b, jx - ix, 0));
                double s[2];
                s[0] = ws[0][ii - 1] + aijpj(P, ii, 1, DP, -2, cs, &P->igs);
                s[1] = ws[1][ii - 1] + aijpj(P, ii, 2, DP, -2, cs, &P->igs);
                double pr[3] = {px[ii - 1], py[ii - 1], pn[ii - 1]};
                elmtrc(&el[ii - 1], coef, eps, omegah, omegas, pr, mus[ii - 1], s);
                dif += sqr(pr[0] - px[ii - 1]) + sqr(pr[1] - py[ii - 1]);
                dpx[ii - 1] += pr[0] - px[ii - 1];
                dpy[ii - 1] += pr[1] - py[ii - 1];
                px[ii - 1] = pr[0];
                py[ii - 1] = pr[1];
                sx[ii - 1] = s[0];
                sy[ii - 1] = s[1];
                /* rebuild the upstream part of the row (:3086-3127) */
                for (int kx = ix - 1; kx >= 1; kx--) {
                    int jj = kx + (iy - 1) * mx - 1, jn = jj + 1;
                    if (el[jj] == ORC_ADHES) {
                        px[jj] = px[jn] + dpx[jj];
                        py[jj] = py[jn] + dpy[jj];
                        double ptabs = sqrt(px[jj] * px[jj] + py[jj] * py[jj]);
                        double ptbnd = mus[jj] * pn[jj];
                        if (ptabs > ptbnd) {
                            px[jj] = px[jj] * ptbnd / ptabs;
                            py[jj] = py[jj] * ptbnd / ptabs;
                            dpx[jj] = px[jj] - px[jn];
                            dpy[jj] = py[jj] - py[jn];
                        }
                    } else if (el[jj] == ORC_SLIP || el[jj] == ORC_PLAST) {
                        dpx[jj] = px[jj] - px[jn];
                        dpy[jj] = py[jj] - py[jn];
                    } else if (el[jn] >= ORC_ADHES) {
                        dpx[jj] = -px[jn];
                        dpy[jj] = -py[jn];
                    }
                }
            }
        }
        dif = sqrt(dif / (2.0 * k));
        double rms = P->p0[0] ? gf_rms_diff(ALLELM, el, npot, px, P->p0[0], py, P->p0[1])
                              : gf_rms(ALLELM, el, npot, px, py);
        difid = eps * dmax(1e-6, facnel * rms);
        if (itgs == 1) dif1 = dif;
    }
    double conv = 1.0;
    if (dif * dif1 != 0.0 && itgs > 1) conv = exp(log(dif / dif1) / (itgs - 1));
    *info = 0;
    if (itgs >= maxgs) *info = 1;
    if (itgs >= maxgs && conv > sp(0.997f)) *info = 2;
    if (itgs >= maxgs && conv > 1.0) *info = 3;
    *itgs_out = itgs;
    *err_out = dif;
    free(dpx);
    free(dpy);
    return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* TangCG, src/m_solvpt.f90:1841-2442                                                           */
/* ------------------------------------------------------------------------------------------ */

static int tangcg(orc_problem *P, int maxcg, double eps, double **ws, const double *mus,
                  int *itcg_out, double *err_out)
{
    const int num_inn = 4;
    const double small = 1e-6;
    int npot = P->npot, *el = P->igs.el, my = P->my;
    double *px = P->ps[0], *py = P->ps[1], *pn = P->ps[2], *sx = P->ss[0], *sy = P->ss[1];
    double c11 = cf_at(&P->cs, 0, 0, 0, 0), c22 = cf_at(&P->cs, 1, 1, 0, 0);
    double c12 = cf_at(&P->cs, 0, 1, 0, 0), c21 = cf_at(&P->cs, 1, 0, 0, 0);
    double ga = P->cs.ga;
    if (dmax(fabs(c12), fabs(c21)) > sp(1e-8f) * dmin(c11, c22)) return CNTC_ERR_OTHER;
    int use_fftprec = 1;
    make_prec(&P->cs, 0, &P->ms);
    make_prec(&P->cs, 1, &P->ms);
    double *g = dalloc(npot), *nn = dalloc(npot), *nx = dalloc(npot), *ny = dalloc(npot),
           *tx = dalloc(npot), *ty = dalloc(npot), *rx = dalloc(npot), *ry = dalloc(npot),
           *zx = dalloc(npot), *zy = dalloc(npot), *vx = dalloc(npot), *vy = dalloc(npot),
           *qx = dalloc(npot), *qy = dalloc(npot), *poldx = dalloc(npot), *poldy = dalloc(npot);
    double *R[3] = {rx, ry, NULL}, *Z[3] = {zx, zy, NULL}, *V[3] = {vx, vy, NULL}, *Q[3] = {qx, qy, NULL};
    int nadh, nslip, nplast, nexter;
    count_el(el, npot, &nadh, &nslip, &nplast, &nexter);
    double facnel = (double)sqrtf((float)npot / (float)(nadh + nslip + nplast));
    for (int i = 0; i < npot; i++) g[i] = mus[i] * pn[i];
    int itcg = 0, it_inn = 0, lchanged = 0;
    double dif = 2.0, difid = 1.0, difinn = 0.0, trsinn = 0.0, alpha = 0.0, beta;
#define NT()                                                                                       \
    for (int i = 0; i < npot; i++) {                                                               \
        nn[i] = atan2(py[i], px[i]); nx[i] = cos(nn[i]); ny[i] = sin(nn[i]);                       \
        tx[i] = -ny[i]; ty[i] = nx[i];                                                             \
    }
    NT();
    for (int i = 0; i < npot; i++) { sx[i] = 0.0; sy[i] = 0.0; }
    vecaijpj(P, el, ALLINT, P->ss, -2, P->ps, -2, &P->cs);
    for (int i = 0; i < npot; i++)
        if (el[i] >= ORC_ADHES) { sx[i] += ws[0][i]; sy[i] += ws[1][i]; }
#define RSET()                                                                                     \
    for (int i = 0; i < npot; i++) {                                                               \
        rx[i] = -sx[i]; ry[i] = -sy[i];                                                            \
        if (el[i] == ORC_SLIP) {                                                                   \
            double perp = tx[i] * rx[i] + ty[i] * ry[i];                                           \
            rx[i] = perp * tx[i]; ry[i] = perp * ty[i];                                            \
        }                                                                                          \
    }
    RSET();
    while ((lchanged || dif > difid) && itcg < maxcg) {
        itcg++;
        it_inn++;
        if (use_fftprec) {
            if (nslip + nplast > 3 * nadh && my > 1) use_fftprec = 0;
            if (itcg >= maxcg / 2) use_fftprec = 0;
            if (!use_fftprec) lchanged = 1;
        }
        if (use_fftprec) {
            vecaijpj(P, el, ALLINT, Z, 1, R, 1, &P->ms);
            vecaijpj(P, el, ALLINT, Z, 2, R, 2, &P->ms);
        } else {
            memcpy(zx, rx, sizeof(double) * npot);
            memcpy(zy, ry, sizeof(double) * npot);
        }
        for (int i = 0; i < npot; i++) {
            if (el[i] == ORC_SLIP) {
                double snrm = -sx[i] * nx[i] - sy[i] * ny[i];
                zx[i] = zx[i] / (c11 + ga * snrm / g[i]);
                zy[i] = zy[i] / (c22 + ga * snrm / g[i]);
            } else {
                zx[i] = zx[i] / c11;
                zy[i] = zy[i] / c22;
            }
        }
        for (int i = 0; i < npot; i++)
            if (el[i] == ORC_SLIP) {
                double perp = tx[i] * zx[i] + ty[i] * zy[i];
                zx[i] = perp * tx[i]; zy[i] = perp * ty[i];
            }
        if (itcg <= 1 || lchanged) {
            memcpy(vx, zx, sizeof(double) * npot);
            memcpy(vy, zy, sizeof(double) * npot);
        } else {
            double zq = gf_dot(ALLELM, el, npot, zx, qx) + gf_dot(ALLELM, el, npot, zy, qy);
            double vq = gf_dot(ALLELM, el, npot, vx, qx) + gf_dot(ALLELM, el, npot, vy, qy);
            if (fabs(zq) < 1e-60 || fabs(vq) < small * fabs(zq)) beta = 0.0;
            else beta = -zq / vq;
            for (int i = 0; i < npot; i++) { vx[i] = beta * vx[i] + zx[i]; vy[i] = beta * vy[i] + zy[i]; }
        }
        vecaijpj(P, el, ALLINT, Q, -2, V, -2, &P->cs);
        for (int i = 0; i < npot; i++)
            if (el[i] == ORC_SLIP) {
                double perp = tx[i] * qx[i] + ty[i] * qy[i];
                qx[i] = perp * tx[i]; qy[i] = perp * ty[i];
            }
        for (int i = 0; i < npot; i++)
            if (el[i] == ORC_SLIP) {
                double snrm = (nx[i] * sx[i] + ny[i] * sy[i]) / g[i];
                qx[i] -= snrm * vx[i];
                qy[i] -= snrm * vy[i];
            }
        double rv = gf_dot(ALLELM, el, npot, rx, vx) + gf_dot(ALLELM, el, npot, ry, vy);
        double vq = gf_dot(ALLELM, el, npot, vx, qx) + gf_dot(ALLELM, el, npot, vy, qy);
        if (fabs(rv) < 1e-60) alpha = 0.0;
        else if (fabs(vq) < small * fabs(rv)) alpha = 1.0;
        else alpha = rv / vq;
        memcpy(poldx, px, sizeof(double) * npot);
        memcpy(poldy, py, sizeof(double) * npot);
        for (int i = 0; i < npot; i++)
            if (el[i] >= ORC_ADHES) { px[i] += alpha * vx[i]; py[i] += alpha * vy[i]; }
        for (int i = 0; i < npot; i++)
            if (el[i] == ORC_SLIP) {
                double ptabs = sqrt(px[i] * px[i] + py[i] * py[i]);
                px[i] = px[i] * g[i] / ptabs;
                py[i] = py[i] * g[i] / ptabs;
            }
        dif = alpha * facnel * gf_rms(ALLELM, el, npot, vx, vy);
        difinn += dif;
        if (itcg <= 5 || itcg % 5 == 1) {
            double ptang = facnel * (P->p0[0] ? gf_rms_diff(ALLELM, el, npot, px, P->p0[0], py, P->p0[1])
                                              : gf_rms(ALLELM, el, npot, px, py));
            difid = eps * dmax(1e-6, ptang);
            trsinn = sp(0.01f) * ptang;
        }
        if (it_inn >= num_inn || dif <= difid || difinn > trsinn) {
            lchanged = 0;
            for (int i = 0; i < npot; i++)
                if (el[i] == ORC_ADHES) {
                    double ptabs = sqrt(px[i] * px[i] + py[i] * py[i]);
                    if (ptabs > g[i]) {
                        el[i] = ORC_SLIP;
                        px[i] = px[i] * g[i] / ptabs;
                        py[i] = py[i] * g[i] / ptabs;
                        lchanged = 1;
                    }
                }
            vecaijpj(P, el, ALLINT, P->ss, -2, P->ps, -2, &P->cs);
            for (int i = 0; i < npot; i++)
                if (el[i] >= ORC_ADHES) { sx[i] += ws[0][i]; sy[i] += ws[1][i]; }
            for (int i = 0; i < npot; i++)
                if (el[i] == ORC_SLIP) {
                    double snrm = sx[i] * px[i] + sy[i] * py[i];
                    if (snrm > 0.0) { el[i] = ORC_ADHES; lchanged = 1; }
                }
            for (int i = 0; i < npot; i++) { poldx[i] -= px[i]; poldy[i] -= py[i]; }
            dif = facnel * gf_rms(ALLELM, el, npot, poldx, poldy);
            it_inn = 0;
            difinn = 0.0;
            if (lchanged) count_el(el, npot, &nadh, &nslip, &nplast, &nexter);
        }
        if ((lchanged || dif > difid) && itcg < maxcg) {
            if (it_inn <= 0) {
                NT();
                RSET();
            } else {
                for (int i = 0; i < npot; i++)
                    if (el[i] >= ORC_ADHES) { rx[i] -= alpha * qx[i]; ry[i] -= alpha * qy[i]; }
            }
        }
    }
#undef NT
#undef RSET
    *itcg_out = itcg;
    *err_out = dif;
    free(g); free(nn); free(nx); free(ny); free(tx); free(ty); free(rx); free(ry); free(zx);
    free(zy); free(vx); free(vy); free(qx); free(qy); free(poldx); free(poldy);
    return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* stang (Coulomb, creepages prescribed), src/m_stang.f90:28-951, src/m_solvpt.f90:51-493       */
/* ------------------------------------------------------------------------------------------ */

/* sxbnd, src/m_leadedge.f90:92-332: per element the fraction facdt of the time step spent in
   the contact area and, for elements entering it (facdt < 0.9999), the leading-edge position
   (the interior element ix0 before the exterior) whose displacement ubnd replaces the previous
   time's; the edge lies fxdfac = 1 element beyond ix0 (2 with SteadyGS, :215-224) */
static int set_facdt(orc_problem *P, int tang, double fxdfac)
{
    int mx = P->mx, my = P->my, *el = P->igs.el;
    int is_roll = tang == 2 || tang == 3;
    int ixinc = 1;
    for (int iy = 0; iy < my; iy++) {
        /* leading-edge transition positions in this row */
        int nb = 0;
        int xb[4096];
        for (int ix = 1; ix <= mx - 1; ix++) {
            int ii = ix + iy * mx;
            if (el[ii - 1] >= ORC_ADHES && el[ii] <= ORC_EXTER && nb < 4096) xb[nb++] = ix;
        }
        if (el[mx - 1 + iy * mx] >= ORC_ADHES && nb < 4096) xb[nb++] = mx;
        int j = 0;
        for (int ix = 1; ix <= mx; ix++) {
            while (j < nb && ix > xb[j]) j++;
            int ii = ix - 1 + iy * mx;
            P->lix[ii] = -1;
            if (!is_roll) P->facdt[ii] = 1.0;
            else if (el[ii] <= ORC_EXTER) P->facdt[ii] = 0.0;
            else if (ix + 2 > mx) P->facdt[ii] = 1.0;
            else {
                double xbnd = P->x[xb[j] - 1 + iy * mx] + ixinc * fxdfac * P->dx;
                P->facdt[ii] = dmin(1.0, ixinc * (xbnd - P->x[ii]) / P->dq);
                if (P->facdt[ii] < 0.9999) P->lix[ii] = xb[j] - 1;
            }
        }
    }
    return 0;
}

/* subnd, src/m_leadedge.f90:336-394: the tangential displacement difference at the leading edge
   of element ii's row, interpolated between ix0 and ix0 + 1 from the previous tractions pv (all
   components, direct sums with the coefficients cs and pv's element division igv) */
static double subnd_of(orc_problem *P, int ii, int ik, double fxdfac, double **p, const eldiv *pel)
{
    int mx = P->mx, iy = ii / mx, ix0 = P->lix[ii] + 1;   /* 1-based */
    if (ix0 + 2 > mx) return 0.0;
    int i0 = ix0 + iy * mx;                                /* 1-based index of ix0 */
    double u0 = aijpj(P, i0, ik, p, -3, &P->cs, pel);
    double u1 = aijpj(P, i0 + 1, ik, p, -3, &P->cs, pel);
    return (1.0 - fxdfac) * u0 + fxdfac * u1;
}

static double subnd(orc_problem *P, int ii, int ik, double fxdfac)
{
    return subnd_of(P, ii, ik, fxdfac, P->pv, &P->igv);
}

/* cnvxgs, src/m_solvpt.f90:2446-2821 (chi = 0, no plasticity, no third body): point
   Gauss-Seidel over the interior elements in grid order, each element's shift by direct
   summation with the current tractions.  Transient (T=1, 2): s = ws + A(cs) p.  Steady rolling
   (T=3): s = ws + A(csv) p, and near the leading edge (lix >= 0) s = ws + A(cs) p - ubnd with
   ubnd interpolated from A(cs) p at the sweep's start (subnd, src/m_leadedge.f90:336-394).
   After 1000 sweeps the slip elements are skipped on odd sweeps (:2603). */
static int cnvxgs(orc_problem *P, double **ws, const double *mus, int k, double eps, int maxgs,
                  double omegah, double omegas, int is_ssrol, int *itgs_out, double *err_out,
                  int *info)
{
    int npot = P->npot, *el = P->igs.el;
    double *px = P->ps[0], *py = P->ps[1], *pn = P->ps[2];
    double *sx = P->ss[0], *sy = P->ss[1];
    double coefs[2][2], coefsv[2][2];
    for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++) {
            coefs[a][b] = P->cs.ga_inv * cf_at(&P->cs, a, b, 0, 0);
            coefsv[a][b] = P->csv.ga_inv * cf_at(&P->csv, a, b, 0, 0);
        }
    int *iel = ialloc(npot), n = 0;
    for (int i = 0; i < npot; i++)
        if (el[i] >= ORC_ADHES) iel[n++] = i;
    double *ubx = dalloc(npot), *uby = dalloc(npot);
    int nadh, nslip, nplast, nexter;
    count_el(el, npot, &nadh, &nslip, &nplast, &nexter);
    double facnel = (double)sqrtf((float)npot / (float)(nadh + nslip + nplast));
    int itgs = 0;
    double dif = 2.0, difid = 1.0, dif1 = 0.0;
    while (dif >= difid && itgs < maxgs) {
        itgs++;
        dif = 0.0;
        if (is_ssrol)
            for (int i = 0; i < n; i++) {
                int ii = iel[i];
                if (P->lix[ii] < 0) continue;
                ubx[ii] = subnd_of(P, ii, 1, 1.0, P->ps, &P->igs);
                uby[ii] = subnd_of(P, ii, 2, 1.0, P->ps, &P->igs);
            }
        for (int i = 0; i < n; i++) {
            int ii = iel[i];
            if (!(itgs <= 1000 || itgs % 2 == 0 || el[ii] == ORC_ADHES)) continue;
            int zledge = is_ssrol && P->lix[ii] >= 0;
            double s[2];
            if (!is_ssrol) {
                s[0] = ws[0][ii] + aijpj(P, ii + 1, 1, P->ps, -2, &P->cs, &P->igs);
                s[1] = ws[1][ii] + aijpj(P, ii + 1, 2, P->ps, -2, &P->cs, &P->igs);
            } else if (!zledge) {
                s[0] = ws[0][ii] + aijpj(P, ii + 1, 1, P->ps, -2, &P->csv, &P->igs);
                s[1] = ws[1][ii] + aijpj(P, ii + 1, 2, P->ps, -2, &P->csv, &P->igs);
            } else {
                s[0] = ws[0][ii] + aijpj(P, ii + 1, 1, P->ps, -2, &P->cs, &P->igs) - ubx[ii];
                s[1] = ws[1][ii] + aijpj(P, ii + 1, 2, P->ps, -2, &P->cs, &P->igs) - uby[ii];
            }
            double pr[3] = {px[ii], py[ii], pn[ii]};
            elmtrc(&el[ii], (!is_ssrol || zledge) ? coefs : coefsv, eps, omegah, omegas, pr, mus[ii], s);
            dif += sqr(px[ii] - pr[0]) + sqr(py[ii] - pr[1]);
            px[ii] = pr[0];
            py[ii] = pr[1];
            sx[ii] = s[0];
            sy[ii] = s[1];
        }
        dif = sqrt(dif / (2.0 * k));
        double rms = P->p0[0] ? gf_rms_diff(ALLELM, el, npot, px, P->p0[0], py, P->p0[1])
                              : gf_rms(ALLELM, el, npot, px, py);
        difid = eps * dmax(1e-6, facnel * rms);
        if (itgs == 1) dif1 = dif;
    }
    double conv = 1.0;
    if (dif * dif1 != 0.0 && itgs > 1) conv = exp(log(dif / dif1) / (itgs - 1));
    *info = 0;
    if (itgs >= maxgs) *info = 1;
    if (itgs >= maxgs && conv > sp(0.997f)) *info = 2;
    if (itgs >= maxgs && conv > 1.0) *info = 3;
    *itgs_out = itgs;
    *err_out = dif;
    free(iel); free(ubx); free(uby);
    return 0;
}

/* tang_solver, src/m_solvpt.f90:382-493: GDsteady for T=3 with G=5, restarted by SteadyGS when
   it stagnates; SteadyGS for other G in steady rolling; TangCG for T=1, 2.  *it counts the
   iterations of both solvers of a fall-back.  *info is set by SteadyGS only (TangCG and
   GDsteady leave it). */
static int tang_solver(orc_problem *P, const orc_input *inp, double **ws, const double *mus, int k,
                       double omegah, double omegas, const gdpar *G, int maxit, double eps, int *it,
                       double *err, int *info)
{
    int rc, it1 = 0;
    areas(&P->igs, P->mx, P->my);
    if (P->use_cnvxgs) {
        rc = cnvxgs(P, ws, mus, k, eps, maxit, omegah, omegas, inp->tang == 3, &it1, err, info);
        *it = abs(it1);
    } else if (inp->tang == 3 && inp->gausei == 5) {
        int lstagn = 0;
        rc = gdsteady(P, G, maxit, eps, ws, mus, &it1, err, &lstagn);
        P->gd_stagnated = lstagn;
        *it = abs(it1);
        if (!rc && lstagn) {
            rc = stdygs(P, ws, mus, k, eps, maxit, omegah, omegas, &it1, err, info);
            P->used_stdygs = 1;
            *it += abs(it1);
        }
    } else if (inp->tang == 3) {
        rc = stdygs(P, ws, mus, k, eps, maxit, omegah, omegas, &it1, err, info);
        P->used_stdygs = 1;
        *it = abs(it1);
    } else {
        rc = tangcg(P, maxit, eps, ws, mus, &it1, err);
        *it = abs(it1);
    }
    return rc;
}

/* Kalker's creep coefficients C11, C22, C23 (Kalker 1967 / 1990 table, as tabulated in
   src/m_hertz.f90:1158-1179) for g = 0.1, 0.2, ..., 1.0 and poisson ratios 0, 1/4, 1/2.  Case 0
   holds g = a/b <= 1, case 1 holds g = b/a < 1.  The reference stores them from single-precision
   literals, kept here as floats. */
static const float KALKER_C[2][10][3][3] = {   /* [case][g row][C11, C22, C23][nu] */
    {{{2.51f, 3.31f, 4.85f}, {2.51f, 2.52f, 2.53f}, {.334f, .473f, .731f}},
     {{2.59f, 3.37f, 4.81f}, {2.59f, 2.63f, 2.66f}, {.483f, .603f, .809f}},
     {{2.68f, 3.44f, 4.80f}, {2.68f, 2.75f, 2.81f}, {.607f, .715f, .889f}},
     {{2.78f, 3.53f, 4.82f}, {2.78f, 2.88f, 2.98f}, {.720f, .823f, .977f}},
     {{2.88f, 3.62f, 4.83f}, {2.88f, 3.01f, 3.14f}, {.827f, .929f, 1.07f}},
     {{2.98f, 3.72f, 4.91f}, {2.98f, 3.14f, 3.31f}, {.930f, 1.03f, 1.18f}},
     {{3.09f, 3.81f, 4.97f}, {3.09f, 3.28f, 3.48f}, {1.03f, 1.14f, 1.29f}},
     {{3.19f, 3.91f, 5.05f}, {3.19f, 3.41f, 3.65f}, {1.13f, 1.25f, 1.40f}},
     {{3.29f, 4.01f, 5.12f}, {3.29f, 3.54f, 3.82f}, {1.23f, 1.36f, 1.51f}},
     {{3.40f, 4.12f, 5.20f}, {3.40f, 3.67f, 3.98f}, {1.33f, 1.47f, 1.63f}}},
    {{{10.7f, 11.7f, 12.9f}, {10.7f, 12.8f, 16.0f}, {12.2f, 14.6f, 18.0f}},
     {{6.96f, 7.78f, 8.82f}, {6.96f, 8.14f, 9.79f}, {5.72f, 6.63f, 7.89f}},
     {{5.57f, 6.34f, 7.34f}, {5.57f, 6.40f, 7.51f}, {3.79f, 4.32f, 5.01f}},
     {{4.84f, 5.57f, 6.57f}, {4.84f, 5.48f, 6.31f}, {2.88f, 3.24f, 3.70f}},
     {{4.37f, 5.10f, 6.11f}, {4.37f, 4.90f, 5.56f}, {2.35f, 2.62f, 2.96f}},
     {{4.06f, 4.78f, 5.80f}, {4.06f, 4.50f, 5.04f}, {2.01f, 2.23f, 2.50f}},
     {{3.82f, 4.54f, 5.58f}, {3.82f, 4.21f, 4.67f}, {1.76f, 1.95f, 2.18f}},
     {{3.65f, 4.36f, 5.42f}, {3.65f, 3.99f, 4.39f}, {1.58f, 1.75f, 1.94f}},
     {{3.51f, 4.22f, 5.30f}, {3.51f, 3.81f, 4.16f}, {1.44f, 1.59f, 1.77f}},
     {{3.40f, 4.12f, 5.20f}, {3.40f, 3.67f, 3.98f}, {1.33f, 1.47f, 1.63f}}}};

/* linrol, src/m_hertz.f90:1147-1247: asymptotic expressions for slender ellipses (g < 0.101),
   otherwise linear interpolation in g between table rows and quadratic (Lagrange)
   interpolation of 1/C in nu */
void orc_linrol(double nu, double aob, double *cxx, double *cyy, double *cyz)
{
    int icase;
    double g;
    if (aob <= 1.0) { g = aob; icase = 0; }
    else { g = 1.0 / aob; icase = 1; }
    if (icase == 0 && g < 0.101) {
        *cxx = PI * PI / 4.0 / (1.0 - nu);
        *cyy = PI * PI / 4.0;
        *cyz = PI * sqrt(g) / 3.0 / (1.0 - nu) * (1.0 + nu * (log(16.0 / g) - 5.0));
    } else if (icase == 1 && g < 0.101) {
        double al = log(16.0 / (g * g));
        *cxx = 2.0 * PI / (al - 2.0 * nu) / g * (1.0 + (3.0 - log(4.0)) / (al - 2.0 * nu));
        *cyy = 2.0 * PI / g * (1.0 + (1.0 - nu) * (3.0 - log(4.0)) / ((1.0 - nu) * al + 2.0 * nu)) /
               ((1.0 - nu) * al + 2.0 * nu);
        *cyz = 2.0 * PI / 3.0 / pow(g, 1.5) / ((1.0 - nu) * al - 2.0 + 4.0 * nu);
    } else {
        int irow = (int)(10.0 * g - 0.005);        /* 1-based row below g */
        double fac = 10.0 * g - (double)irow;
        double c[3][3], cc[3];
        for (int inu = 0; inu < 3; inu++)
            for (int ic = 0; ic < 3; ic++)
                c[inu][ic] = fac * (double)KALKER_C[icase][irow][ic][inu] +
                             (1.0 - fac) * (double)KALKER_C[icase][irow - 1][ic][inu];
        for (int ic = 0; ic < 3; ic++)
            cc[ic] = (nu - 0.25) * (nu - 0.50) * 8.0 / c[0][ic] - (nu - 0.50) * nu * 16.0 / c[1][ic] +
                     nu * (nu - 0.25) * 8.0 / c[2][ic];
        *cxx = 1.0 / cc[0];
        *cyy = 1.0 / cc[1];
        *cyz = 1.0 / cc[2];
    }
}

/* char_creep, src/m_solvpt.f90:497-540: creepages that give a force of order mu*Fn in the
   linear theory, for the effective radius and half-axes of the current contact area */
static void char_creep(orc_problem *P, double *ch_ksi, double *ch_eta, double *ch_phi)
{
    int nadh, nslip, nplast, nexter;
    count_el(P->igs.el, P->npot, &nadh, &nslip, &nplast, &nexter);
    int ncon = nadh + nslip + nplast;
    double c = sqrt((double)(float)ncon * P->dx * P->dy / PI);
    areas(&P->igs, P->mx, P->my);
    double a = (double)imax(1, P->igs.ixmax - P->igs.ixmin + 1) * P->dx * 0.5;
    double b = (double)imax(1, P->igs.iymax - P->igs.iymin + 1) * P->dy * 0.5;
    double c11, c22, c23;
    orc_linrol(P->nu, a / b, &c11, &c22, &c23);
    *ch_ksi = P->fstat * P->fntrue / (c * c * P->ga * c11);
    *ch_eta = P->fstat * P->fntrue / (c * c * P->ga * c22);
    *ch_phi = P->fstat * P->fntrue / (c * c * c * P->ga * c23);
}

/* sens_tang, src/m_solvpt.f90:544-796: d(Fx, Fy)/(mu Fn) and dMz per creepage by one
   tang_solver call each (is_sens: maxit mxsens, eps epsens; src/m_solvpt.f90:382-493) on the
   fixed right-hand side plus the creepage change, started from the converged solution, which
   is restored after each */
static int sens_tang(orc_problem *P, const orc_input *inp, double **ws, const double *mus, int k,
                     double omegah, double omegas, const gdpar *G)
{
    int npot = P->npot, *el = P->igs.el;
    double dxdy = P->dxdy;
    double den = P->muscal * P->fntrue;
    double fxrel = dxdy * gf_sum(ALLELM, el, npot, P->ps[0]) / den;
    double fyrel = dxdy * gf_sum(ALLELM, el, npot, P->ps[1]) / den;
    double sy = 0.0, sx = 0.0;
    for (int i = 0; i < npot; i++) sy += P->ps[0][i] * P->y[i];
    for (int i = 0; i < npot; i++) sx += P->ps[1][i] * P->x[i];
    double mztrue = dxdy * (-sy + sx);
    double ch[3];
    char_creep(P, &ch[0], &ch[1], &ch[2]);
    int *el0 = ialloc(npot);
    double *save[4], *wsloc[2] = {dalloc(npot), dalloc(npot)};
    double *st[4] = {P->ps[0], P->ps[1], P->ss[0], P->ss[1]};
    for (int q = 0; q < 4; q++) { save[q] = dalloc(npot); memcpy(save[q], st[q], sizeof(double) * npot); }
    memcpy(el0, el, sizeof(int) * npot);
    int stagn0 = P->gd_stagnated, used0 = P->used_stdygs, ret = 0;
    for (int icrp = 0; icrp < 3; icrp++) {
        double dcrp = 0.0001 * ch[icrp];
        /* wstot = wsfix1 plus the creepages that solvpt adds when forces are prescribed
           (:600-614) */
        for (int i = 0; i < npot; i++) { wsloc[0][i] = ws[0][i]; wsloc[1][i] = ws[1][i]; }
        for (int i = 0; i < npot; i++) {
            if (el[i] < ORC_ADHES) continue;
            if (inp->force3 >= 1) wsloc[0][i] = wsloc[0][i] + P->facdt[i] * P->cksi * P->dq;
            if (inp->force3 >= 2) wsloc[1][i] = wsloc[1][i] + P->facdt[i] * P->ceta * P->dq;
        }
        for (int i = 0; i < npot; i++) {
            if (el[i] < ORC_ADHES) continue;
            double f = P->facdt[i];
            if (icrp == 0) wsloc[0][i] = wsloc[0][i] + f * dcrp * P->dq;
            else if (icrp == 1) wsloc[1][i] = wsloc[1][i] + f * dcrp * P->dq;
            else {
                wsloc[0][i] = wsloc[0][i] - f * dcrp * P->y[i] * P->dq;
                wsloc[1][i] = wsloc[1][i] + f * dcrp * P->x[i] * P->dq;
            }
        }
        P->p0[0] = save[0];
        P->p0[1] = save[1];
        int it = 0, info = 0, rc;
        double err = 0.0;
        rc = tang_solver(P, inp, wsloc, mus, k, omegah, omegas, G, inp->mxsens, inp->epsens, &it, &err, &info);
        P->p0[0] = P->p0[1] = NULL;
        if (rc) { ret = rc; }
        else {
            double fxloc = dxdy * gf_sum(ALLELM, el, npot, P->ps[0]) / den;
            double fyloc = dxdy * gf_sum(ALLELM, el, npot, P->ps[1]) / den;
            double a = 0.0, b = 0.0;
            for (int i = 0; i < npot; i++) a += P->ps[0][i] * P->y[i];
            for (int i = 0; i < npot; i++) b += P->ps[1][i] * P->x[i];
            double mzloc = dxdy * (-a + b);
            SENS(P, IOUT_FX, IIN_KSI + icrp) = (fxloc - fxrel) / dcrp;
            SENS(P, IOUT_FY, IIN_KSI + icrp) = (fyloc - fyrel) / dcrp;
            SENS(P, IOUT_MZ, IIN_KSI + icrp) = (mzloc - mztrue) / dcrp;
        }
        memcpy(el, el0, sizeof(int) * npot);
        for (int q = 0; q < 4; q++) memcpy(st[q], save[q], sizeof(double) * npot);
        areas(&P->igs, P->mx, P->my);
        if (ret) break;
    }
    P->gd_stagnated = stagn0;
    P->used_stdygs = used0;
    free(el0);
    for (int q = 0; q < 4; q++) free(save[q]);
    free(wsloc[0]);
    free(wsloc[1]);
    return ret;
}

/* solvpt, src/m_solvpt.f90:51-378.  F=0: one tang_solver call on wsfix1.  F=1 (Fx prescribed)
   or F=2 (Fx and Fy): wstot = wsfix1 + facdt (cksi, ceta) dq, then Newton-Raphson on cksi
   (and ceta) with the sensitivities d(Fx, Fy)/(mu Fn) per creepage, which are refreshed
   from each half step; *it sums the solver iterations. */
static int solvpt(orc_problem *P, const orc_input *inp, double **wsfix, const double *mus, int k,
                  double omegah, double omegas, const gdpar *G, int *itgs_out, double *err_out)
{
    int npot = P->npot, *el = P->igs.el, F = inp->force3;
    double dxdy = P->dxdy, den = P->muscal * P->fntrue;
    double *wstot[2] = {dalloc(npot), dalloc(npot)};
    memcpy(wstot[0], wsfix[0], sizeof(double) * npot);
    memcpy(wstot[1], wsfix[1], sizeof(double) * npot);
    for (int i = 0; i < npot; i++) {
        if (el[i] < ORC_ADHES) continue;
        if (F >= 1) wstot[0][i] = wstot[0][i] + P->facdt[i] * P->cksi * P->dq;
        if (F >= 2) wstot[1][i] = wstot[1][i] + P->facdt[i] * P->ceta * P->dq;
    }
    int itgs = 0, info1 = 0, it = 0, rc;
    double err = 0.0;
    rc = tang_solver(P, inp, wstot, mus, k, omegah, omegas, G, inp->maxgs, inp->eps, &it, &err, &info1);
    double fxkp1 = dxdy * gf_sum(ALLELM, el, npot, P->ps[0]) / den;
    double fykp1 = dxdy * gf_sum(ALLELM, el, npot, P->ps[1]) / den;
    itgs += it;
    int nadh, nslip, nplast, nexter;
    count_el(el, npot, &nadh, &nslip, &nplast, &nexter);
    if (!rc && F >= 1 && info1 <= 1) {
        int itnr = 0;
        double df = fabs(inp->fxrel - fxkp1);
        if (F >= 2) df = df + fabs(inp->fyrel - fykp1);
        while (!rc && df > inp->eps && itnr < inp->maxnr && info1 <= 1) {
            itnr++;
            double dfx = inp->fxrel - fxkp1, dfy = inp->fyrel - fykp1, dcksi = 0.0, dceta = 0.0;
            if (F == 1) {
                dceta = 0.0;
                if (fabs(SENS(P, IOUT_FX, IIN_KSI)) > sp(1e-6f)) dcksi = dfx / SENS(P, IOUT_FX, IIN_KSI);
                else dcksi = sp(0.00003f);
            } else {
                double det = SENS(P, IOUT_FX, IIN_KSI) * SENS(P, IOUT_FY, IIN_ETA) -
                             SENS(P, IOUT_FY, IIN_KSI) * SENS(P, IOUT_FX, IIN_ETA);
                if (det > inp->eps) {
                    dcksi = (SENS(P, IOUT_FY, IIN_ETA) * dfx - SENS(P, IOUT_FY, IIN_KSI) * dfy) / det;
                    dceta = (-SENS(P, IOUT_FX, IIN_ETA) * dfx + SENS(P, IOUT_FX, IIN_KSI) * dfy) / det;
                } else {
                    dcksi = 0.000003;
                    dceta = 0.000003;
                }
            }
            for (int ifxy = 1; ifxy <= F && !rc; ifxy++) {
                double fxk = fxkp1, fyk = fykp1;
                if (ifxy == 1) {
                    P->cksi = P->cksi + dcksi;
                    for (int i = 0; i < npot; i++)
                        if (el[i] >= ORC_ADHES) wstot[0][i] = wstot[0][i] + P->facdt[i] * dcksi * P->dq;
                } else {
                    P->ceta = P->ceta + dceta;
                    for (int i = 0; i < npot; i++)
                        if (el[i] >= ORC_ADHES) wstot[1][i] = wstot[1][i] + P->facdt[i] * dceta * P->dq;
                }
                rc = tang_solver(P, inp, wstot, mus, k, omegah, omegas, G, inp->maxgs, inp->eps, &it,
                                 &err, &info1);
                fxkp1 = dxdy * gf_sum(ALLELM, el, npot, P->ps[0]) / den;
                fykp1 = dxdy * gf_sum(ALLELM, el, npot, P->ps[1]) / den;
                itgs += it;
                count_el(el, npot, &nadh, &nslip, &nplast, &nexter);
                double dfxk = fxkp1 - fxk, dfyk = fykp1 - fyk;
                double area = (double)(nadh + nslip + nplast) * dxdy;
                if (ifxy == 1) {
                    if (nadh > 0) {
                        double dpxavg = dfxk * P->muscal * P->fntrue / area;
                        if (SENS(P, IOUT_FX, IIN_KSI) == 0.0 || fabs(dpxavg) > 10.0 * err)
                            SENS(P, IOUT_FX, IIN_KSI) = dfxk / dcksi;
                        double dpyavg = dfyk * P->muscal * P->fntrue / area;
                        if (fabs(dpyavg) > 10.0 * err) SENS(P, IOUT_FY, IIN_KSI) = dfyk / dcksi;
                    } else if (fabs(SENS(P, IOUT_FX, IIN_KSI)) < sp(1e-6f)) {
                        SENS(P, IOUT_FX, IIN_KSI) = fxkp1 / P->cksi;
                    }
                } else {
                    if (nadh > 0) {
                        double dpxavg = dfxk * P->muscal * P->fntrue / area;
                        if (fabs(dpxavg) > 10.0 * err) SENS(P, IOUT_FX, IIN_ETA) = dfxk / dceta;
                        double dpyavg = dfyk * P->muscal * P->fntrue / area;
                        if (SENS(P, IOUT_FY, IIN_ETA) == 0.0 || fabs(dpyavg) > 10.0 * err)
                            SENS(P, IOUT_FY, IIN_ETA) = dfyk / dceta;
                    } else if (fabs(SENS(P, IOUT_FY, IIN_ETA)) < sp(1e-6f)) {
                        SENS(P, IOUT_FY, IIN_ETA) = fykp1 / P->ceta;
                    }
                }
                df = fabs(inp->fxrel - fxkp1);
                if (F >= 2) df = df + fabs(inp->fyrel - fykp1);
            }
        }
        /* the reference's error estimate adds the slip elements as the state code Slip = 2
           (`slip` in src/m_solvpt.f90:366 names that constant) */
        err = err + 2.0 * df * P->muscal * P->fntrue / ((double)(nadh + ORC_SLIP + nplast) * dxdy);
    } else {
        SENS(P, IOUT_FX, IIN_KSI) = 0.0;
        SENS(P, IOUT_FY, IIN_KSI) = 0.0;
        SENS(P, IOUT_FX, IIN_ETA) = 0.0;
        SENS(P, IOUT_FY, IIN_ETA) = 0.0;
    }
    free(wstot[0]);
    free(wstot[1]);
    *itgs_out = itgs;
    *err_out = err;
    return rc;
}

static int stang(orc_problem *P, const orc_input *inp, int *ittang_out)
{
    int npot = P->npot, mx = P->mx, my = P->my, *el = P->igs.el;
    double *mus = dalloc(npot);
    for (int i = 0; i < npot; i++) mus[i] = P->fstat;
    int k = 0;
    for (int i = 0; i < npot; i++)
        if (el[i] >= ORC_ADHES) k++;
    if (k <= 0) { free(mus); return CNTC_ERR_OTHER; }
    /* the solver (src/m_stang.f90:145-190): ConvexGS for G=2, and for steady rolling with
       contact in the trailing-edge column (no exterior element to start the row's walk) */
    int icount = 0;
    for (int iy = 0; iy < my; iy++)
        if (el[iy * mx] >= ORC_ADHES) icount++;
    P->use_cnvxgs = inp->gausei == 2 || (inp->tang == 3 && icount > 0);
    /* sxbnd: the leading edge 1 dx beyond the last interior element, 2 dx for SteadyGS */
    int is_stdygs = inp->tang == 3 && !P->use_cnvxgs && inp->gausei != 5;
    int rc = set_facdt(P, inp->tang, is_stdygs ? 2.0 : 1.0);
    if (rc) { free(mus); return rc; }
    /* relaxation parameters, src/m_stang.f90:192-223 */
    double omegah = inp->omegah, omegas = inp->omegas;
    if (inp->gausei == 0 || inp->gausei == 4 || inp->gausei == 5) {
        if (k <= 25) { omegah = 1.0; omegas = 1.0; }
        else if (inp->tang == 3 && P->use_cnvxgs) { omegah = 0.5; omegas = 0.5; }
        else if (inp->tang == 3) {
            if (P->dx / P->dy <= 5.0) { omegah = 0.9; omegas = 1.0; }
            else if (P->dx / P->dy <= 15.0) { omegah = 0.8; omegas = 0.8; }
            else { omegah = 0.8; omegas = 0.6; }
        } else { omegah = 0.5; omegas = 0.5; }
    }

    /* stang_rhs, src/m_stang.f90:749-951 */
    double *ws[2] = {dalloc(npot), dalloc(npot)};
    double *usn[3] = {dalloc(npot), dalloc(npot), NULL}, *uvn[3] = {dalloc(npot), dalloc(npot), NULL},
           *uvt[3] = {dalloc(npot), dalloc(npot), NULL};
    areas(&P->igs, mx, my);
    vecaijpj(P, el, ALLINT, usn, -2, P->ps, 3, &P->cs);
    if (inp->tang != 3) vecaijpj(P, el, ALLINT, uvn, -2, P->pv, 3, &P->cv);
    else vecaijpj(P, el, ALLINT, uvn, -2, P->ps, 3, &P->cv);
    if (inp->tang != 3) vecaijpj(P, el, ALLINT, uvt, -2, P->pv, -2, &P->cv);
    areas(&P->igv, mx, my);
    for (int ik = 0; ik < 2; ik++)
        for (int i = 0; i < npot; i++)
            if (el[i] >= ORC_ADHES) {
                double wsrig = -P->facdt[i] * P->hs[ik][i];
                if (inp->tang == 3) ws[ik][i] = wsrig + usn[ik][i] - uvn[ik][i];
                else if (P->lix[i] < 0) ws[ik][i] = wsrig + usn[ik][i] - uvn[ik][i] - uvt[ik][i];
                else ws[ik][i] = wsrig + usn[ik][i] - subnd(P, i, ik + 1, 1.0);   /* (1b) */
            }

    gdpar G;
    G.mx = mx; G.my = my; G.npot = npot;
    gd_method(inp, &G.gd_meth, &G.kdown);
    G.kdowfb = inp->kdowfb; G.fdecay = inp->fdecay; G.betath = inp->betath;
    G.d_ifc = inp->d_ifc; G.d_lin = inp->d_lin; G.d_cns = inp->d_cns; G.d_slp = inp->d_slp;
    G.pow_s = inp->pow_s;

    int ittang = 0, itgs = 0, zready = 0, ret = 0;
    double *tmp[3] = {dalloc(npot), dalloc(npot), NULL};
    while (!zready && ittang < inp->maxin) {
        ittang++;
        zready = 1;
        int it = 0;
        double errpt = 0.0;
        rc = solvpt(P, inp, ws, mus, k, omegah, omegas, &G, &it, &errpt);
        itgs += it;
        if (rc) { ret = rc; break; }
        int newins = 0;
        double tol = sqrt(2.0) * errpt;
        for (int i = 0; i < npot; i++)
            if (el[i] == ORC_ADHES) {
                double pabs = sqrt(P->ps[0][i] * P->ps[0][i] + P->ps[1][i] * P->ps[1][i]);
                if (pabs >= mus[i] * P->ps[2][i] + tol) {
                    newins++;
                    el[i] = ORC_SLIP;
                    P->ps[0][i] = P->ps[0][i] * mus[i] * P->ps[2][i] / pabs;
                    P->ps[1][i] = P->ps[1][i] * mus[i] * P->ps[2][i] / pabs;
                }
            }
        if (newins != 0) zready = 0;
        if (zready) {
            areas(&P->igs, mx, my);
            int ii = imax(1, mx / 2) + (imax(1, my / 2) - 1) * mx;
            for (int i = 0; i < npot; i++) { tmp[0][i] = 1.0; tmp[1][i] = 1.0; }
            double tol1 = errpt * 2.0 * fabs(aijpj(P, ii, 1, tmp, 1, &P->cs, &P->igs));
            double tol2 = errpt * 2.0 * fabs(aijpj(P, ii, 2, tmp, 2, &P->cs, &P->igs));
            int newadh = 0;
            for (int i = 0; i < npot; i++)
                if (el[i] == ORC_SLIP) {
                    double ww = P->ss[0][i] * P->ps[0][i] + P->ss[1][i] * P->ps[1][i];
                    double tl = tol1 * fabs(P->ps[0][i]) + tol2 * fabs(P->ps[1][i]) +
                                errpt * (fabs(P->ss[0][i]) + fabs(P->ss[1][i]));
                    if (ww > tl) { el[i] = ORC_ADHES; newadh++; }
                }
            if (newadh != 0) zready = 0;
        }
    }
    if (!zready && !ret) ittang = -1;
    P->itgs = itgs;
    *ittang_out = ittang;
    /* sensitivities to the creepages, src/m_stang.f90:724-731 */
    if (!ret && inp->sens >= 3) ret = sens_tang(P, inp, ws, mus, k, omegah, omegas, &G);
    free(mus); free(ws[0]); free(ws[1]); free(usn[0]); free(usn[1]); free(uvn[0]); free(uvn[1]);
    free(uvt[0]); free(uvt[1]); free(tmp[0]); free(tmp[1]);
    return ret;
}

/* ------------------------------------------------------------------------------------------ */
/* case set-up and driver: contac/panprc, src/m_scontc.f90:37-553                               */
/* ------------------------------------------------------------------------------------------ */

orc_problem *orc_new(void) { return (orc_problem *)calloc(1, sizeof(orc_problem)); }

static void free_grid_arrays(orc_problem *p)
{
    free(p->x); free(p->y); p->x = p->y = NULL;
    for (int k = 0; k < 3; k++) {
        free(p->hs[k]); free(p->ps[k]); free(p->pv[k]);
        p->hs[k] = p->ps[k] = p->pv[k] = NULL;
    }
    for (int k = 0; k < 2; k++) { free(p->ss[k]); p->ss[k] = NULL; }
    free(p->facdt); p->facdt = NULL;
    free(p->lix); p->lix = NULL;
    free(p->igs.el); free(p->igs.row1st); free(p->igs.rowlst);
    free(p->igv.el); free(p->igv.row1st); free(p->igv.rowlst);
    memset(&p->igs, 0, sizeof p->igs);
    memset(&p->igv, 0, sizeof p->igv);
}

void orc_free(orc_problem *p)
{
    if (!p) return;
    free_grid_arrays(p);
    inflcf_free(&p->cs); inflcf_free(&p->cv); inflcf_free(&p->csv); inflcf_free(&p->ms);
    free(p);
}

/* sgencr, src/m_visc.f90:127-376 (regenerated each case: C3=2 is not a reuse setting) */
static void sgencr(orc_problem *P, const orc_input *inp)
{
    int mx = P->mx, my = P->my;
    int is_roll = inp->tang == 2 || inp->tang == 3;
    inflcf_new(&P->cs, mx, my, P->dx, P->dy);
    inflcf_new(&P->cv, mx, my, P->dx, P->dy);
    inflcf_new(&P->csv, mx, my, P->dx, P->dy);
    inflcf_new(&P->ms, mx, my, P->dx, P->dy);
    inflcf *all[4] = {&P->cs, &P->cv, &P->csv, &P->ms};
    for (int k = 0; k < 4; k++) { all[k]->ga = P->ga; all[k]->ga_inv = 1.0 / P->ga; }
    double cc = is_roll ? cos(P->chi) : 0.0, sc = is_roll ? sin(P->chi) : 0.0;
    if (fabs(P->ak) < 1e-6)
        for (int k = 0; k < 3; k++) all[k]->nt_cpl = 0;
    double facdq = 1.0, facdqi = 1.0 / facdq;
    P->cv.dq = P->dq;
    P->csv.dq = P->dq;
    orc_elascf_pcwcns(P->ak, P->nu, mx, my, P->dx, P->dy, 0.0, 0.0, P->cs.cf);
    size_t nb = (size_t)(2 * mx) * (2 * my);
    if (!is_roll) {
        memcpy(P->cv.cf, P->cs.cf, sizeof(double) * 9 * nb);
        memcpy(P->csv.cf, P->cs.cf, sizeof(double) * 9 * nb);
    } else {
        orc_elascf_pcwcns(P->ak, P->nu, mx, my, P->dx, P->dy, cc * facdq * P->dq, sc * facdq * P->dq, P->cv.cf);
        for (int jk = 0; jk < 3; jk++) memcpy(P->cv.cf + (size_t)(6 + jk) * nb, P->cs.cf + (size_t)(6 + jk) * nb, sizeof(double) * nb);
        for (int jk = 0; jk < 3; jk++)
            for (size_t k = 0; k < nb; k++) {
                for (int ik = 0; ik < 2; ik++) {
                    size_t b = (size_t)(ik * 3 + jk) * nb + k;
                    P->csv.cf[b] = facdqi * (P->cs.cf[b] - P->cv.cf[b]);
                    P->cv.cf[b] = P->cs.cf[b] - P->csv.cf[b];
                }
                P->csv.cf[(size_t)(6 + jk) * nb + k] = P->cs.cf[(size_t)(6 + jk) * nb + k];
            }
    }
}

/* eldiv0, src/m_sdis.f90:818-1012 */
static void eldiv0(orc_problem *P, const orc_input *inp)
{
    const int maxit = 50;
    const double facpen = sp(0.60f), reltol = sp(0.01f);
    int npot = P->npot;
    double *hs = P->hs[2];
    double fnscal = P->fntrue / P->ga;
    int iimin = 0;
    for (int i = 1; i < npot; i++)
        if (hs[i] < hs[iimin]) iimin = i;
    double hsmin = hs[iimin];
    double pentru = P->pen - hsmin;
    if (inp->norm == 1 && P->my == 1) {
        double rm;
        if (inp->ibase == 1) rm = 0.5 / dmax(1e-6, inp->prmudf[0]);
        else rm = 1.0;
        double cdy = 0.35 - 0.05 * (log(P->dy) - log(200.0));
        pentru = pow(fnscal / P->dy * (1.0 - P->nu) / 0.6 / cdy / pow(rm, 0.1), 0.926);
        P->pen = pentru + hsmin;
    } else if (inp->norm == 1) {
        double rmn = 1.0;
        double fac = rmn * 0.75 * sqrt(PI) * (1.0 - P->nu);
        double penmin = 0.0, fnmin = 0.0;
        double penmax = fac * fnscal / sqrt(P->dxdy);
        int ncnmax = 0;
        for (int i = 0; i < npot; i++)
            if (hs[i] - hsmin < facpen * penmax) ncnmax++;
        double fnmax = penmax * sqrt(P->dxdy * ncnmax) / fac;
        int iter = 0;
        while (iter < maxit && fabs(fnmax - fnmin) > reltol * fnscal) {
            iter++;
            double penmid = 0.5 * (penmax + penmin);
            int ncnmid = 0;
            for (int i = 0; i < npot; i++)
                if (hs[i] - hsmin < facpen * penmid) ncnmid++;
            double fnmid = penmid * sqrt(P->dxdy * ncnmid) / fac;
            if (fnmid < fnscal) { penmin = penmid; fnmin = fnmid; }
            else { penmax = penmid; fnmax = fnmid; }
        }
        pentru = penmax;
        P->pen = pentru + hsmin;
    }
    for (int i = 0; i < npot; i++) P->igs.el[i] = (hs[i] - hsmin < facpen * pentru) ? ORC_ADHES : ORC_EXTER;
}

int orc_calculate(orc_problem *P, const orc_input *inp)
{
    P->ierror = 0;
    P->itnorm = P->ittang = P->itcg = P->itgs = 0;
    P->gd_stagnated = 0;
    P->used_stdygs = 0;
    if (inp->tang < 0 || inp->tang > 3) return P->ierror = CNTC_ERR_OTHER;
    /* combin_mater, src/m_hierarch_data.f90:1543-1553 */
    P->ga = 2.0 / (1.0 / inp->gg1 + 1.0 / inp->gg2);
    P->nu = P->ga * (inp->poiss1 / inp->gg1 + inp->poiss2 / inp->gg2) / 2.0;
    P->ak = (P->ga / 4.0) * ((1.0 - 2.0 * inp->poiss1) / inp->gg1 - (1.0 - 2.0 * inp->poiss2) / inp->gg2);
    P->fstat = inp->fstat;
    P->muscal = inp->fstat;
    P->pen = inp->pen;
    P->fntrue = inp->fn;
    P->cksi = inp->cksi; P->ceta = inp->ceta; P->cphi = inp->cphi;
    P->chi = inp->chi; P->veloc = inp->veloc;

    /* grid: ipotcn=3 directly, Hertzian options via hzsol + potcon_hertz */
    int mx = inp->mx, my = inp->my;
    double xc1 = inp->xc1, yc1 = inp->yc1, dx = inp->dx, dy = inp->dy;
    double prmudf[6];
    memcpy(prmudf, inp->prmudf, sizeof prmudf);
    int ibase = inp->ibase;
    int hertz2d = inp->ipotcn == -4 || inp->ipotcn == -5;
    if (inp->bound != 0 && inp->bound != 2) return P->ierror = CNTC_ERR_OTHER;
    if (inp->bound == 2 && !(inp->ipotcn <= -1 && inp->ipotcn >= -5)) return P->ierror = CNTC_ERR_OTHER;
    if (hertz2d && inp->norm != 1) return P->ierror = CNTC_ERR_OTHER;
    if (inp->ipotcn <= -1 && inp->ipotcn >= -5) {
        /* hzsol, src/m_hertz.f90:29-90 */
        double e_star = P->ga / (1.0 - P->nu);
        P->hz_a1 = inp->hz_a1; P->hz_b1 = inp->hz_b1; P->hz_aa = inp->hz_aa; P->hz_bb = inp->hz_bb;
        int herr = 0;
        if (hertz2d) {
            /* hzcalc2d, src/m_hertz.f90:217-263 */
            double fn_2d = P->fntrue / (2.0 * P->hz_bb), rx;
            P->hz_b1 = 0.0;
            if (inp->ipotcn == -4) {
                rx = 1.0 / (2.0 * P->hz_a1);
                P->hz_aa = sqrt(4.0 * rx * fn_2d / (PI * e_star));
            } else {
                rx = PI * P->hz_aa * P->hz_aa * e_star / (4.0 * fn_2d);
                P->hz_a1 = 1.0 / (2.0 * rx);
            }
            P->hz_rho = 2.0 / (P->hz_a1 + P->hz_b1);
            P->hz_cp = sqrt(4.0 * P->hz_aa * P->hz_bb / PI);
        } else if (inp->ipotcn == -2) {
            herr = elliptic_to_curv(e_star, sp(1e-9f), inp->norm, &P->hz_a1, inp->hz_aob, &P->hz_b1,
                                    &P->hz_aa, &P->hz_bb, &P->pen, &P->fntrue, &P->hz_cp, &P->hz_rho);
        } else {
            herr = hzcalc3d(e_star, sp(1e-9f), inp->ipotcn, &P->hz_a1, &P->hz_b1, &P->hz_aa, &P->hz_bb,
                            inp->norm, &P->pen, &P->fntrue, &P->hz_cp, &P->hz_rho);
        }
        if (herr != 0) return P->ierror = CNTC_ERR_OTHER;
        ibase = 1;
        memset(prmudf, 0, sizeof prmudf);
        prmudf[0] = P->hz_a1;
        prmudf[2] = P->hz_b1;
        /* potcon_hertz, src/m_hierarch_data.f90:1889-1959 */
        double xl = -inp->hz_scale * dmax(1e-9, P->hz_aa), yl = -inp->hz_scale * dmax(1e-9, P->hz_bb);
        dx = (-xl - xl) / mx;
        dy = (-yl - yl) / my;
        if (hertz2d) {
            int n_int = my;
            if (inp->hz_scale > 1.0) {
                n_int = imax(1, (int)lround(my / inp->hz_scale));
                if ((my - n_int) % 2 == 1) n_int = n_int + 1;
            }
            dy = 2.0 * dmax(1e-9, P->hz_bb) / n_int;
            yl = -my * dy / 2.0;
        }
        xc1 = xl + 0.5 * dx;
        yc1 = yl + 0.5 * dy;
    } else if (inp->ipotcn != 3) {
        return P->ierror = CNTC_ERR_OTHER;
    }
    int newgrid = (mx != P->mx || my != P->my || !P->x);
    if (newgrid && (inp->pvtime != 2 || inp->iestim != 0) && P->x) return P->ierror = CNTC_ERR_OTHER;
    if (newgrid) {
        free_grid_arrays(P);
        P->mx = mx; P->my = my; P->npot = mx * my;
        int n = P->npot;
        P->x = dalloc(n); P->y = dalloc(n);
        for (int k = 0; k < 3; k++) { P->hs[k] = dalloc(n); P->ps[k] = dalloc(n); P->pv[k] = dalloc(n); }
        for (int k = 0; k < 2; k++) P->ss[k] = dalloc(n);
        P->facdt = dalloc(n);
        P->lix = (int *)calloc((size_t)n, sizeof(int));
        eldiv_alloc(&P->igs, n, my);
        eldiv_alloc(&P->igv, n, my);
    }
    P->dx = dx; P->dy = dy; P->dxdy = dx * dy; P->xc1 = xc1; P->yc1 = yc1;
    int npot = P->npot;
    for (int iy = 1; iy <= my; iy++)
        for (int ix = 1; ix <= mx; ix++) {
            int ii = ix - 1 + (iy - 1) * mx;
            P->x[ii] = xc1 + (ix - 1) * dx;
            P->y[ii] = yc1 + (iy - 1) * dy;
        }

    /* check_roll_stepsize, src/m_sdis.f90:125-204: steady rolling with SteadyGS or GDsteady
       uses dq = dx; ConvexGS (G=2) keeps dq */
    int is_roll = inp->tang == 2 || inp->tang == 3;
    P->dq = inp->dq > 0.0 ? inp->dq : dx;
    if (inp->tang == 3 && inp->gausei != 2) {
        if (fabs(P->chi) > 0.01 && fabs(P->chi - PI) > 0.01) P->chi = 0.0;
        P->dq = dx;
    }
    if (!is_roll) { P->chi = 0.0; P->dq = 1.0; P->veloc = 1.0; }
    if (fabs(P->chi) > 0.01) return P->ierror = CNTC_ERR_OTHER;   /* chi=0 only */
    P->dt = P->dq / dmax(1e-10, P->veloc);

    /* set_prev_data, src/m_sdis.f90:208-317 */
    if (inp->pvtime == 0) {
        memcpy(P->igv.el, P->igs.el, sizeof(int) * npot);
        for (int k = 0; k < 3; k++) memcpy(P->pv[k], P->ps[k], sizeof(double) * npot);
    } else {
        for (int i = 0; i < npot; i++) P->igv.el[i] = ORC_EXTER;
        for (int k = 0; k < 3; k++)
            for (int i = 0; i < npot; i++) P->pv[k][i] = 0.0;
    }

    sgencr(P, inp);

    /* set_norm_rhs, src/m_sdis.f90:321-494 */
    for (int i = 0; i < npot; i++) {
        if (ibase == 1)
            P->hs[2][i] = prmudf[0] * P->x[i] * P->x[i] + prmudf[1] * P->x[i] * P->y[i] +
                          prmudf[2] * P->y[i] * P->y[i] + prmudf[3] * P->x[i] + prmudf[4] * P->y[i] + prmudf[5];
        else if (ibase == 9)
            P->hs[2][i] = inp->h_elem[i];
        else
            return P->ierror = CNTC_ERR_OTHER;
    }
    /* set_tang_rhs, src/m_sdis.f90:498-583 (no spin offset; E digit 0-3) */
    {
        double xofs = is_roll ? cos(P->chi) * P->dq * (1.0 / 6.0) : 0.0;
        double yofs = is_roll ? sin(P->chi) * P->dq * (1.0 / 6.0) : 0.0;
        double facx = inp->rztang == 3 ? 0.0 : 1.0, facy = inp->rztang == 2 ? 0.0 : 1.0;
        int use_ex = inp->rztang == 1 && inp->exrhs;
        for (int i = 0; i < npot; i++) {
            double ex = use_ex ? inp->exrhs[i] : 0.0, ey = use_ex ? inp->exrhs[npot + i] : 0.0;
            P->hs[0][i] = facx * -(P->y[i] + yofs - 0.0) * P->cphi + ex;
            P->hs[1][i] = facy * (P->x[i] + xofs - 0.0) * P->cphi + ey;
            if (inp->force3 == 0) P->hs[0][i] += P->cksi;   /* F=0: cksi prescribed */
            if (inp->force3 <= 1) P->hs[1][i] += P->ceta;   /* F<=1: ceta prescribed */
            P->hs[0][i] *= -P->dq;
            P->hs[1][i] *= -P->dq;
        }
    }

    /* init_curr_data, src/m_sdis.f90:587-768 */
    int ihertz = inp->ipotcn < 0 ? -inp->ipotcn : 0;
    if (inp->iestim == 0) {
        for (int i = 0; i < npot; i++) P->ps[2][i] = 0.0;
    }
    if (inp->iestim == 0 || inp->tang == 0) {
        for (int i = 0; i < npot; i++) { P->ps[0][i] = 0.0; P->ps[1][i] = 0.0; }
    }
    int use_iestim_hertz = inp->iestim == 0 && inp->bound == 0 && ihertz >= 1;
    if (inp->bound == 2 || use_iestim_hertz) {
        if (ihertz <= 3) {
            double pnmax = 3.0 * P->fntrue / (2.0 * PI * P->hz_aa * P->hz_bb);
            for (int i = 0; i < npot; i++) {
                double pnfac = 1.0 - sqr(P->x[i] / P->hz_aa) - sqr(P->y[i] / P->hz_bb);
                P->ps[2][i] = pnmax * sqrt(dmax(0.0, pnfac));
            }
        } else {
            /* 2D elliptical traction bound, uniform in y */
            double pnmax = P->fntrue / (PI * P->hz_aa * P->hz_bb);
            for (int i = 0; i < npot; i++) {
                double pnfac = 1.0 - sqr(P->x[i] / P->hz_aa);
                if (fabs(P->y[i]) > P->hz_bb) pnfac = 0.0;
                P->ps[2][i] = pnmax * sqrt(dmax(0.0, pnfac));
            }
        }
        for (int i = 0; i < npot; i++) P->igs.el[i] = P->ps[2][i] > sp(1e-20f) ? ORC_ADHES : ORC_EXTER;
    } else if (inp->iestim == 0) {
        eldiv0(P, inp);
    }
    if (inp->iestim == 1) {
        /* regularise the tractions (:722-735): exterior 0, slip on the traction bound */
        for (int i = 0; i < npot; i++) {
            if (P->igs.el[i] <= ORC_EXTER) {
                P->ps[0][i] = P->ps[1][i] = P->ps[2][i] = 0.0;
            } else if (P->igs.el[i] == ORC_SLIP) {
                double pabs = dmax(1e-10, sqrt(sqr(P->ps[0][i]) + sqr(P->ps[1][i])));
                P->ps[0][i] = P->muscal * P->ps[2][i] * P->ps[0][i] / pabs;
                P->ps[1][i] = P->muscal * P->ps[2][i] * P->ps[1][i] / pabs;
            }
        }
    }
    for (int i = 0; i < npot; i++)
        if (P->hs[2][i] > sp(1e29f)) P->igs.el[i] = ORC_EXTER;
    areas(&P->igs, mx, my);
    /* initial sensitivities (:753-764); eldiv0's dFn/dpen estimate for N=1 is always replaced
       in snorm (src/m_snorm.f90:357-368) and is not formed */
    if (inp->iestim == 0)
        for (int io = IOUT_FX; io <= IOUT_MZ; io++) SENS(P, io, IIN_PEN) = 0.0;
    if (inp->iestim == 0 && (inp->norm != 1 || ihertz != 0)) SENS(P, IOUT_FN, IIN_PEN) = 0.0;
    if (inp->iestim == 0 || inp->iestim == 2) {
        if (inp->force3 >= 1) P->cksi = 1e-6;
        if (inp->force3 == 2) P->ceta = 0.0;
    }
    if (inp->iestim == 0 || inp->iestim == 2)
        for (int ii = IIN_KSI; ii <= IIN_PHI; ii++)
            for (int io = IOUT_FN; io <= IOUT_MZ; io++) SENS(P, io, ii) = 0.0;

    /* panprc, src/m_scontc.f90:356-553 */
    int maxout = fabs(P->ak) < 1e-4 ? 1 : 10;
    int itout = 0, itnorm_tot = 0, ittang_tot = 0;
    double difid = 1.0, dif = 200.0 * difid;
    double *po[3] = {dalloc(npot), dalloc(npot), dalloc(npot)};
    for (int k = 0; k < 3; k++) memcpy(po[k], P->ps[k], sizeof(double) * npot);
    int rc = 0;
    while (dif > difid && itout < maxout && itnorm_tot >= 0 && ittang_tot >= 0) {
        itout++;
        int it = 0;
        /* B=2: the Hertzian pressure is the normal solution (src/m_scontc.f90:409-411) */
        rc = inp->bound == 2 ? 0 : snorm(P, inp, &it);
        if (rc) break;
        itnorm_tot = it >= 0 ? itnorm_tot + it : -1;
        if (inp->tang == 3) memcpy(P->igv.el, P->igs.el, sizeof(int) * npot);
        int ncon = 0;
        for (int i = 0; i < npot; i++) {
            if (P->igs.el[i] <= ORC_EXTER) { P->ps[0][i] = P->ps[1][i] = P->ps[2][i] = 0.0; }
            else ncon++;
        }
        if (inp->tang == 0 || ncon <= 0) {
            for (int i = 0; i < npot; i++) { P->ss[0][i] = P->ss[1][i] = 0.0; }
            dif = 0.0;
        } else {
            rc = stang(P, inp, &it);
            if (rc) break;
            ittang_tot = it >= 0 ? ittang_tot + it : -1;
            for (int k = 0; k < 3; k++)
                for (int i = 0; i < npot; i++) po[k][i] -= P->ps[k][i];
            dif = sqrt(0.0 + sqr(gf_rms(ALLINT, P->igs.el, npot, po[0], po[1])) * 2.0 / 3.0 +
                       sqr(gf_rms(ALLINT, P->igs.el, npot, po[2], NULL)) / 3.0);
            difid = 5.0 * inp->eps *
                    sqrt(sqr(gf_rms(ALLINT, P->igs.el, npot, P->ps[0], P->ps[1])) * 2.0 / 3.0 +
                         sqr(gf_rms(ALLINT, P->igs.el, npot, P->ps[2], NULL)) / 3.0);
            for (int k = 0; k < 3; k++) memcpy(po[k], P->ps[k], sizeof(double) * npot);
        }
    }
    for (int k = 0; k < 3; k++) free(po[k]);
    P->itnorm = itnorm_tot;
    P->ittang = ittang_tot;

    /* soutpt forces, src/m_soutpt.f90:424-450 */
    if (inp->norm == 0) P->fntrue = P->dxdy * gf_sum(ALLELM, P->igs.el, npot, P->ps[2]);
    P->fxrel = P->dxdy * gf_sum(ALLELM, P->igs.el, npot, P->ps[0]) / (P->fntrue * P->muscal + TINY);
    P->fyrel = P->dxdy * gf_sum(ALLELM, P->igs.el, npot, P->ps[1]) / (P->fntrue * P->muscal + TINY);
    P->fcntc[0] = P->fxrel * (P->fntrue * P->muscal + TINY);
    P->fcntc[1] = P->fyrel * (P->fntrue * P->muscal + TINY);
    P->fcntc[2] = P->fntrue;
    {
        double a = 0.0, b = 0.0;
        for (int i = 0; i < npot; i++) { a += P->ps[0][i] * P->y[i]; b += P->ps[1][i] * P->x[i]; }
        P->mztrue = -P->dxdy * a + P->dxdy * b;
        if (fabs(P->mztrue) < 0.5 * inp->eps * (P->muscal * P->fntrue + TINY)) P->mztrue = 0.0;
    }
    P->have_solution = 1;
    if (rc) return P->ierror = rc;
    if (P->itnorm < 0) return P->ierror = CNTC_ERR_NORM;
    if (P->ittang < 0) return P->ierror = CNTC_ERR_TANG;
    return P->ierror = 0;
}

/* ------------------------------------------------------------------------------------------ */
/* getters                                                                                      */
/* ------------------------------------------------------------------------------------------ */

/* displacement differences of the last solution, soutpt (src/m_soutpt.f90:371-397): us = A ps
   on the interior (AllInt, matfil_surf < 2), 0 elsewhere (init_curr_data clears us,
   src/m_sdis.f90:624) */
void orc_get_displacements(orc_problem *P, double *un, double *ux, double *uy)
{
    int n = P->npot;
    double *U[3] = {ux, uy, un};
    for (int k = 0; k < 3; k++)
        for (int i = 0; i < n; i++) U[k][i] = 0.0;
    vecaijpj(P, P->igs.el, ALLINT, U, 3, P->ps, -3, &P->cs);
    vecaijpj(P, P->igs.el, ALLINT, U, -2, P->ps, -3, &P->cs);
}

int orc_get_mx(const orc_problem *p) { return p->mx; }
int orc_get_my(const orc_problem *p) { return p->my; }
void orc_get_grid(const orc_problem *p, double *xc1, double *yc1, double *dx, double *dy)
{
    *xc1 = p->xc1; *yc1 = p->yc1; *dx = p->dx; *dy = p->dy;
}
void orc_get_forces(const orc_problem *p, double *fn, double *fx, double *fy, double *mz)
{
    *fn = p->fcntc[2]; *fx = p->fcntc[0]; *fy = p->fcntc[1]; *mz = p->mztrue;
}
void orc_get_tractions(const orc_problem *p, double *pn, double *px, double *py)
{
    memcpy(pn, p->ps[2], sizeof(double) * p->npot);
    memcpy(px, p->ps[0], sizeof(double) * p->npot);
    memcpy(py, p->ps[1], sizeof(double) * p->npot);
}
void orc_get_slip(const orc_problem *p, double *sx, double *sy)
{
    memcpy(sx, p->ss[0], sizeof(double) * p->npot);
    memcpy(sy, p->ss[1], sizeof(double) * p->npot);
}
void orc_get_eldiv(const orc_problem *p, int *el) { memcpy(el, p->igs.el, sizeof(int) * p->npot); }
void orc_get_h(const orc_problem *p, double *h) { memcpy(h, p->hs[2], sizeof(double) * p->npot); }
double orc_get_pen(const orc_problem *p) { return p->pen; }
void orc_get_hertz(const orc_problem *p, double *a1, double *b1, double *aa, double *bb, double *rho,
                   double *cp)
{
    *a1 = p->hz_a1; *b1 = p->hz_b1; *aa = p->hz_aa; *bb = p->hz_bb; *rho = p->hz_rho; *cp = p->hz_cp;
}
void orc_get_iters(const orc_problem *p, int *itnorm, int *ittang, int *itcg, int *itgs)
{
    *itnorm = p->itnorm; *ittang = p->ittang; *itcg = p->itcg; *itgs = p->itgs;
}
int orc_get_stagnated(const orc_problem *p) { return p->gd_stagnated + 2 * p->used_stdygs; }
void orc_get_sens(const orc_problem *p, double *sens) { memcpy(sens, p->sens, sizeof p->sens); }
void orc_get_creepages(const orc_problem *p, double *cksi, double *ceta, double *cphi)
{
    *cksi = p->cksi;
    *ceta = p->ceta;
    *cphi = p->cphi;
}
