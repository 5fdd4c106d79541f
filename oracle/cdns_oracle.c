/* CPU oracle for the channel-dns per-step hot path -- TEST INFRASTRUCTURE.
 *
 * A plain-C restatement of the reference solver's algorithm (hesse-x/channel-dns,
 * serial single-rank layout).  Each function cites the reference file:line it
 * follows.  FFTW (the reference's transform library, absent here) is replaced by
 * an exact-definition mixed-radix DFT written below; the reference only relies on
 * FFTW's documented unnormalised r2c/c2r/c2c/REDFT00 definitions.
 *
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * this library.  It is a checker, never part of the product path.
 */
#include "cdns_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

static char g_err[256];
const char *orc_last_error(void) { return g_err; }
int orc_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

/* Complex helpers.  Division follows Smith's algorithm, the scheme libgcc uses for
 * std::complex<double> division; division by a real divides both parts. */
static inline cplx mk(double re, double im) { return CMPLX(re, im); }
static inline cplx cdivr(cplx z, double d) { return mk(creal(z) / d, cimag(z) / d); }
static inline cplx cmulr(cplx z, double d) { return mk(creal(z) * d, cimag(z) * d); }
static inline cplx cmul(cplx a, cplx b) {
    return mk(creal(a) * creal(b) - cimag(a) * cimag(b), creal(a) * cimag(b) + cimag(a) * creal(b));
}
static cplx cdiv(cplx a, cplx b) {
    double a_re = creal(a), a_im = cimag(a), c = creal(b), d = cimag(b);
    if (fabs(c) >= fabs(d)) {
        double r = d / c, den = c + d * r;
        return mk((a_re + a_im * r) / den, (a_im - a_re * r) / den);
    }
    double r = c / d, den = c * r + d;
    return mk((a_re * r + a_im) / den, (a_im * r - a_re) / den);
}
/* (i*k)*v written as in src/core/complex_ops.hpp:6-8. */
static inline cplx fd(double k, cplx v) { return mk(-k * cimag(v), k * creal(v)); }

/* ------------------------------------------------------------------ FFT ---- */
/* Mixed-radix decimation-in-time DFT with exact-angle twiddles (long double).
 * forward: X_k = sum_n x_n e^{-2 pi i k n / N};  backward: e^{+...}; unnormalised. */
typedef struct {
    int n, nf, f[40];
    cplx *w; /* w[k] = exp(-2 pi i k / n) */
} fftplan;

static void plan_init(fftplan *p, int n) {
    p->n = n;
    p->nf = 0;
    int m = n;
    static const int pref[] = {4, 2, 3, 5, 7};
    for (int i = 0; i < 5 && m > 1; ++i)
        while (m % pref[i] == 0) { p->f[p->nf++] = pref[i]; m /= pref[i]; }
    for (int q = 11; m > 1; q += 2)
        while (m % q == 0) { p->f[p->nf++] = q; m /= q; }
    p->w = (cplx *)malloc(sizeof(cplx) * (size_t)(n > 0 ? n : 1));
    const long double twopi = 6.283185307179586476925286766559005768L;
    for (int k = 0; k < n; ++k) {
        long double a = twopi * (long double)k / (long double)n;
        p->w[k] = mk((double)cosl(a), -(double)sinl(a));
    }
}
static void plan_free(fftplan *p) { free(p->w); p->w = NULL; }

static void fft_rec(const cplx *in, int is, cplx *out, int n, const int *f, const fftplan *pl,
                    int wstep, int sign, cplx *tmp) {
    if (n == 1) { out[0] = in[0]; return; }
    const int p = f[0], m = n / p, N = pl->n;
    for (int q = 0; q < p; ++q) fft_rec(in + (size_t)q * is, is * p, out + (size_t)q * m, m, f + 1, pl, wstep * p, sign, tmp);
    for (int k = 0; k < m; ++k) {
        for (int q = 0; q < p; ++q) {
            int idx = (int)(((long long)q * k * wstep) % N);
            cplx tw = pl->w[idx];
            if (sign > 0) tw = conj(tw);
            tmp[q] = cmul(out[(size_t)q * m + k], tw);
        }
        for (int r = 0; r < p; ++r) {
            cplx s = 0;
            for (int q = 0; q < p; ++q) {
                int idx = (int)(((long long)q * r * m * wstep) % N);
                cplx tw = pl->w[idx];
                if (sign > 0) tw = conj(tw);
                s += cmul(tmp[q], tw);
            }
            out[(size_t)k + (size_t)r * m] = s;
        }
    }
}
/* out may not alias in. tmp needs >= largest factor entries. */
static void fft(const fftplan *pl, const cplx *in, cplx *out, int sign, cplx *tmp) {
    fft_rec(in, 1, out, pl->n, pl->f, pl, 1, sign, tmp);
}

/* REDFT00 (FFTW's DCT-I): Y_k = x_0 + (-1)^k x_N + 2 sum_{j=1}^{N-1} x_j cos(pi j k / N),
 * for a complex line of N+1 points, through the even extension of length 2N. */
static void redft00(const fftplan *p2n, cplx *line, int N, cplx *a, cplx *b, cplx *tmp) {
    if (N == 0) return;
    for (int j = 0; j <= N; ++j) a[j] = line[j];
    for (int j = 1; j < N; ++j) a[2 * N - j] = line[j];
    fft(p2n, a, b, -1, tmp);
    for (int k = 0; k <= N; ++k) line[k] = b[k];
}

/* ------------------------------------------------------------- context ---- */
enum { F_VEL = 0, F_VEL_PREV, F_NL, F_NL_PREV, F_P, F_PBC, F_PBC_PREV, F_VBC, F_COUNT };

struct orc_ctx {
    orc_config cfg;
    int NXH, NZH, nzk, nkx, nj, zoff;
    size_t S, BND, MODE;
    double *WAVX, *WAVZ, *WL, *WD, *WR;
    fftplan px2, pz2, py2, px, pz, py; /* fine x/z, 2*NY2 DCT; coarse x/z, 2*NY */
    int maxf;
    /* solver tables (eqsolve_pv.hpp:58-75) */
    double g_v, pressure_a, pressure_bcoef;
    double *wn_p, *wn_v, *ad_p, *ad_v;
    cplx *PI[2], *VI[2], *ID[2], *PB, *VB, *SIG0;
    double *u_dpdx, *w_dpdz, qx_unit, qz_unit;
    /* simulation state */
    cplx *st[F_COUNT];
    cplx *curl_ws;
    cplx *body; /* constant body force added to the nonlinear term, or NULL */
    double dpdx, dpdz, time;
    int setup_done;
};

static size_t sidx(const orc_ctx *c, int izi, int iy, int kx) {
    return ((size_t)izi * c->nj + iy) * c->nkx + kx;
}
static size_t bidx(const orc_ctx *c, int izi, int wall, int kx) {
    return ((size_t)izi * 2 + wall) * c->nkx + kx;
}

static int max_factor(const fftplan *p) {
    int m = 1;
    for (int i = 0; i < p->nf; ++i) if (p->f[i] > m) m = p->f[i];
    return m;
}

orc_ctx *orc_create(const orc_config *cfg) {
    const orc_config *m = cfg;
    if (m->NXS < 0 || m->NYS < 3 || m->NZS < 0 || 2 * m->NXS > m->NX2 || 2 * m->NXS > m->NX ||
        2 * m->NZS > m->NZ2 || 2 * m->NZS > m->NZ || m->NY2 < m->NYS || m->NY < m->NYS) {
        snprintf(g_err, sizeof g_err, "invalid mesh sizes");
        return NULL;
    }
    orc_ctx *c = (orc_ctx *)calloc(1, sizeof(orc_ctx));
    c->cfg = *cfg;
    c->NXH = m->NX / 2;
    c->NZH = m->NZ / 2;
    c->nzk = 2 * m->NZS + 1;
    c->nkx = m->NXS + 1;
    c->nj = m->NYS + 1;
    c->zoff = m->NZS;
    c->S = (size_t)c->nzk * c->nj * c->nkx;
    c->BND = (size_t)c->nzk * 2 * c->nkx;
    c->MODE = (size_t)c->nzk * c->nkx;
    const double PI = acos(-1.0);
    /* transforms.cpp:71-108 */
    const double ALPHAX = 2.0 * PI / m->XL, ALPHAZ = 2.0 * PI / m->ZL;
    c->WAVX = (double *)calloc(c->nkx, sizeof(double));
    for (int k = 0; k <= m->NXS; ++k) c->WAVX[k] = (0 + k) * ALPHAX;
    c->WAVZ = (double *)calloc(c->nzk, sizeof(double));
    for (int k = -m->NZS; k <= m->NZS; ++k) c->WAVZ[k + c->zoff] = k * ALPHAZ;
    c->WL = (double *)calloc(c->nj, sizeof(double));
    c->WD = (double *)calloc(c->nj, sizeof(double));
    c->WR = (double *)calloc(c->nj, sizeof(double));
    for (int iy = m->NYS; iy >= 2; --iy) {
        c->WL[iy] = 1.0 / (4.0 * iy * (iy - 1));
        c->WD[iy] = -1.0 / (2.0 * (iy * iy - 1));
        c->WR[iy] = 1.0 / (4.0 * iy * (iy + 1));
    }
    c->WL[2] = 2.0 * c->WL[2];
    for (int iy = m->NYS - 1; iy <= m->NYS; ++iy) c->WD[iy] = 0.0;
    for (int iy = m->NYS - 3; iy <= m->NYS; ++iy) if (iy >= 0) c->WR[iy] = 0.0;
    plan_init(&c->px2, m->NX2);
    plan_init(&c->pz2, m->NZ2);
    plan_init(&c->py2, 2 * m->NY2);
    plan_init(&c->px, m->NX);
    plan_init(&c->pz, m->NZ);
    plan_init(&c->py, 2 * m->NY);
    c->maxf = 1;
    const fftplan *pls[6] = {&c->px2, &c->pz2, &c->py2, &c->px, &c->pz, &c->py};
    for (int i = 0; i < 6; ++i) if (max_factor(pls[i]) > c->maxf) c->maxf = max_factor(pls[i]);
    c->wn_p = (double *)calloc(c->MODE, sizeof(double));
    c->wn_v = (double *)calloc(c->MODE, sizeof(double));
    c->ad_p = (double *)calloc(c->S, sizeof(double));
    c->ad_v = (double *)calloc(c->S, sizeof(double));
    for (int i = 0; i < 2; ++i) {
        c->PI[i] = (cplx *)calloc(c->S, sizeof(cplx));
        c->VI[i] = (cplx *)calloc(3 * c->S, sizeof(cplx));
        c->ID[i] = (cplx *)calloc(c->BND, sizeof(cplx));
    }
    c->PB = (cplx *)calloc(c->S, sizeof(cplx));
    c->VB = (cplx *)calloc(3 * c->S, sizeof(cplx));
    c->SIG0 = (cplx *)calloc(c->BND, sizeof(cplx));
    c->u_dpdx = (double *)calloc(c->nj, sizeof(double));
    c->w_dpdz = (double *)calloc(c->nj, sizeof(double));
    c->pressure_a = m->a_pre0;
    c->pressure_bcoef = m->b_pre0;
    size_t sz[F_COUNT] = {3 * c->S, 3 * c->S, 3 * c->S, 3 * c->S, c->S, c->BND, c->BND, 3 * c->BND};
    for (int f = 0; f < F_COUNT; ++f) c->st[f] = (cplx *)calloc(sz[f], sizeof(cplx));
    c->curl_ws = (cplx *)calloc(3 * c->S, sizeof(cplx));
    /* initial_bc.cpp:95-101: wall-normal wall velocity on the global zero mode */
    c->st[F_VBC][bidx(c, c->zoff, 0, 0) * 3 + 1] = m->V_Microblow0;
    c->st[F_VBC][bidx(c, c->zoff, 1, 0) * 3 + 1] = m->V_Microblow0;
    return c;
}

void orc_destroy(orc_ctx *c) {
    if (!c) return;
    free(c->WAVX); free(c->WAVZ); free(c->WL); free(c->WD); free(c->WR);
    plan_free(&c->px2); plan_free(&c->pz2); plan_free(&c->py2);
    plan_free(&c->px); plan_free(&c->pz); plan_free(&c->py);
    free(c->wn_p); free(c->wn_v); free(c->ad_p); free(c->ad_v);
    for (int i = 0; i < 2; ++i) { free(c->PI[i]); free(c->VI[i]); free(c->ID[i]); }
    free(c->PB); free(c->VB); free(c->SIG0); free(c->u_dpdx); free(c->w_dpdz);
    for (int f = 0; f < F_COUNT; ++f) free(c->st[f]);
    free(c->curl_ws);
    free(c->body);
    free(c);
}

void orc_tables(const orc_ctx *c, double *WAVX, double *WAVZ, double *WL, double *WD, double *WR) {
    if (WAVX) memcpy(WAVX, c->WAVX, sizeof(double) * c->nkx);
    if (WAVZ) memcpy(WAVZ, c->WAVZ, sizeof(double) * c->nzk);
    if (WL) memcpy(WL, c->WL, sizeof(double) * c->nj);
    if (WD) memcpy(WD, c->WD, sizeof(double) * c->nj);
    if (WR) memcpy(WR, c->WR, sizeof(double) * c->nj);
}

/* -------------------------------------------------- fine 3-D transform ---- */
/* channel_xyzfft2 / xyzfft2_impl, serial layout (transforms.cpp:588-741). */
static void c2r_line(const fftplan *p, const cplx *X, int nin, double *out, cplx *a, cplx *b, cplx *tmp) {
    /* FFTW c2r: Hermitian extension of X[0..n/2]; the imaginary parts of the DC and
     * (even n) Nyquist inputs do not contribute. Only X[0..nin-1] may be nonzero. */
    const int n = p->n;
    for (int k = 0; k < n; ++k) a[k] = 0;
    a[0] = creal(X[0]);
    for (int k = 1; k < nin && 2 * k <= n; ++k) {
        if (2 * k == n) { a[k] = creal(X[k]); continue; }
        a[k] = X[k];
        a[n - k] = conj(X[k]);
    }
    fft(p, a, b, +1, tmp);
    for (int k = 0; k < n; ++k) out[k] = creal(b[k]);
}
static void r2c_line(const fftplan *p, const double *x, int nout, cplx *X, cplx *a, cplx *b, cplx *tmp) {
    const int n = p->n;
    for (int k = 0; k < n; ++k) a[k] = x[k];
    fft(p, a, b, -1, tmp);
    for (int k = 0; k < nout; ++k) X[k] = b[k];
    X[0] = creal(X[0]); /* r2c stores a real DC term */
    if (n % 2 == 0 && nout > n / 2) X[n / 2] = creal(X[n / 2]);
}

#define SCRATCH_DECL                                                                 \
    cplx *a_ = (cplx *)malloc(sizeof(cplx) * (size_t)(scratch_n + 8));               \
    cplx *b_ = (cplx *)malloc(sizeof(cplx) * (size_t)(scratch_n + 8));               \
    cplx *l_ = (cplx *)malloc(sizeof(cplx) * (size_t)(scratch_n + 8));               \
    cplx *t_ = (cplx *)malloc(sizeof(cplx) * (size_t)(c->maxf + 8));
#define SCRATCH_FREE free(a_); free(b_); free(l_); free(t_);

static void fine_backward(orc_ctx *c, double *UP2, const cplx *U, int stride) {
    const orc_config *m = &c->cfg;
    const int NY2 = m->NY2, NZ2 = m->NZ2, NX2 = m->NX2, NYS = m->NYS, NZS = m->NZS;
    const int nkx = c->nkx, nzk = c->nzk, NY21 = NY2 + 1;
    const int scratch_n = 2 * (NX2 > NZ2 ? NX2 : NZ2) + 4 * NY21;
    /* y stage: UY2[iz][iy][kx] (transforms.cpp:692-707) */
    cplx *Y = (cplx *)calloc((size_t)nzk * NY21 * nkx, sizeof(cplx));
#pragma omp parallel
    {
        SCRATCH_DECL
#pragma omp for schedule(static)
        for (int t = 0; t < nzk * nkx; ++t) {
            const int izi = t / nkx, kx = t % nkx;
            for (int iy = 0; iy < NY21; ++iy) l_[iy] = 0;
            for (int iy = 0; iy <= NYS && iy < NY21; ++iy) {
                cplx v = U[sidx(c, izi, iy, kx) * stride];
                if (iy == 0) v = v + v;
                l_[iy] = v;
            }
            redft00(&c->py2, l_, NY2, a_, b_, t_);
            for (int iy = 0; iy < NY21; ++iy) Y[((size_t)izi * NY21 + iy) * nkx + kx] = l_[iy];
        }
        SCRATCH_FREE
    }
    /* z stage: expand retained kz x0.5, backward c2c (transforms.cpp:715-732) */
    cplx *X = (cplx *)calloc((size_t)NZ2 * NY21 * nkx, sizeof(cplx));
#pragma omp parallel
    {
        SCRATCH_DECL
#pragma omp for schedule(static)
        for (int t = 0; t < NY21 * nkx; ++t) {
            const int iy = t / nkx, kx = t % nkx;
            for (int iz = 0; iz < NZ2; ++iz) {
                const int retained = iz <= NZS || iz >= NZ2 - NZS;
                cplx v = 0;
                if (retained) {
                    const int kz = iz <= NZS ? iz : iz - NZ2;
                    v = cmulr(Y[((size_t)(kz + NZS) * NY21 + iy) * nkx + kx], 0.5);
                    if (iz == NZ2 - NZS && iz <= NZS)
                        v = cmulr(Y[((size_t)(0) * NY21 + iy) * nkx + kx], 0.5);
                }
                l_[iz] = v;
            }
            fft(&c->pz2, l_, a_, +1, t_);
            for (int iz = 0; iz < NZ2; ++iz) X[((size_t)iz * NY21 + iy) * nkx + kx] = a_[iz];
        }
        SCRATCH_FREE
    }
    free(Y);
    /* x stage: c2r per (iz, iy) line (transforms.cpp:598-613, 734-740) */
#pragma omp parallel
    {
        SCRATCH_DECL
#pragma omp for schedule(static)
        for (int t = 0; t < NZ2 * NY21; ++t) {
            c2r_line(&c->px2, X + (size_t)t * nkx, nkx, UP2 + (size_t)t * NX2, a_, b_, t_);
        }
        SCRATCH_FREE
    }
    free(X);
}

static void fine_forward(orc_ctx *c, const double *UP2, cplx *U, int stride) {
    const orc_config *m = &c->cfg;
    const int NY2 = m->NY2, NZ2 = m->NZ2, NX2 = m->NX2, NYS = m->NYS, NZS = m->NZS;
    const int nkx = c->nkx, nzk = c->nzk, NY21 = NY2 + 1;
    const int scratch_n = 2 * (NX2 > NZ2 ? NX2 : NZ2) + 4 * NY21;
    const double invNX = 1.0 / NX2, invNZ = 1.0 / NZ2;
    cplx *X = (cplx *)calloc((size_t)NZ2 * NY21 * nkx, sizeof(cplx));
    /* x: r2c, x1/NX2 on kx<=NXS, no Nyquist halving (transforms.cpp:643, 652-657) */
#pragma omp parallel
    {
        SCRATCH_DECL
#pragma omp for schedule(static)
        for (int t = 0; t < NZ2 * NY21; ++t) {
            cplx *Xl = X + (size_t)t * nkx;
            r2c_line(&c->px2, UP2 + (size_t)t * NX2, nkx, Xl, a_, b_, t_);
            for (int kx = 0; kx < nkx; ++kx) Xl[kx] = cmulr(Xl[kx], invNX);
        }
        SCRATCH_FREE
    }
    /* z: forward c2c, x1/NZ2, retained kz scattered (transforms.cpp:658-664) */
    cplx *Y = (cplx *)calloc((size_t)nzk * NY21 * nkx, sizeof(cplx));
#pragma omp parallel
    {
        SCRATCH_DECL
#pragma omp for schedule(static)
        for (int t = 0; t < NY21 * nkx; ++t) {
            const int iy = t / nkx, kx = t % nkx;
            for (int iz = 0; iz < NZ2; ++iz) l_[iz] = X[((size_t)iz * NY21 + iy) * nkx + kx];
            fft(&c->pz2, l_, a_, -1, t_);
            for (int iz = 0; iz <= NZS; ++iz)
                Y[((size_t)(iz + NZS) * NY21 + iy) * nkx + kx] = cmulr(a_[iz], invNZ);
            for (int iz = -NZS; iz <= -1; ++iz)
                Y[((size_t)(iz + NZS) * NY21 + iy) * nkx + kx] = cmulr(a_[NZ2 + iz], invNZ);
        }
        SCRATCH_FREE
    }
    free(X);
    /* y: REDFT00, x1/(2 NY2) at iy=0, x1/NY2 for 1..NYS (transforms.cpp:673-681) */
#pragma omp parallel
    {
        SCRATCH_DECL
#pragma omp for schedule(static)
        for (int t = 0; t < nzk * nkx; ++t) {
            const int izi = t / nkx, kx = t % nkx;
            for (int iy = 0; iy < NY21; ++iy) l_[iy] = Y[((size_t)izi * NY21 + iy) * nkx + kx];
            redft00(&c->py2, l_, NY2, a_, b_, t_);
            for (int iy = 0; iy <= NYS; ++iy) {
                const double scale = iy == 0 ? 1.0 / (2.0 * NY2) : 1.0 / NY2;
                U[sidx(c, izi, iy, kx) * stride] = cmulr(l_[iy], scale);
            }
        }
        SCRATCH_FREE
    }
    free(Y);
}

int orc_xyzfft2(orc_ctx *c, char dir, double *phys, double *spec, int stride) {
    if (stride < 1) { snprintf(g_err, sizeof g_err, "spectral stride must be positive"); return -1; }
    if (dir == 'B') fine_backward(c, phys, (const cplx *)spec, stride);
    else if (dir == 'F') fine_forward(c, phys, (cplx *)spec, stride);
    else { snprintf(g_err, sizeof g_err, "direction must be F or B"); return -1; }
    return 0;
}

/* ------------------------------------------------ coarse 3-D transform ---- */
/* channel_xyzfft serial path (transforms.cpp:452-582). */
static void coarse_forward(orc_ctx *c, const double *UP, cplx *U, int stride) {
    const orc_config *m = &c->cfg;
    const int NX = m->NX, NY = m->NY, NZ = m->NZ, NXS = m->NXS, NYS = m->NYS, NZS = m->NZS;
    const int NXH = c->NXH, NZH = c->NZH, nkx = c->nkx, nzk = c->nzk, NY1 = NY + 1;
    const int fullx = NXS == NXH, fullz = NZS == NZH;
    const int scratch_n = 2 * (NX > NZ ? NX : NZ) + 4 * NY1;
    const double invNX = 1.0 / NX, invNZ = 1.0 / NZ;
    cplx *X = (cplx *)calloc((size_t)NZ * NY1 * nkx, sizeof(cplx));
#pragma omp parallel
    {
        SCRATCH_DECL
#pragma omp for schedule(static)
        for (int t = 0; t < NZ * NY1; ++t) {
            cplx *Xl = X + (size_t)t * nkx;
            r2c_line(&c->px, UP + (size_t)t * NX, nkx, Xl, a_, b_, t_);
            for (int kx = 0; kx <= NXS; ++kx) {
                Xl[kx] = cmulr(Xl[kx], invNX);
                if (fullx && kx == NXH) Xl[kx] = cmulr(Xl[kx], 0.5);
            }
        }
        SCRATCH_FREE
    }
    cplx *Y = (cplx *)calloc((size_t)nzk * NY1 * nkx, sizeof(cplx));
#pragma omp parallel
    {
        SCRATCH_DECL
#pragma omp for schedule(static)
        for (int t = 0; t < NY1 * nkx; ++t) {
            const int iy = t / nkx, kx = t % nkx;
            for (int iz = 0; iz < NZ; ++iz) l_[iz] = X[((size_t)iz * NY1 + iy) * nkx + kx];
            fft(&c->pz, l_, a_, -1, t_);
            for (int iz = 0; iz <= NZS; ++iz)
                Y[((size_t)(iz + NZS) * NY1 + iy) * nkx + kx] = cmulr(a_[iz], invNZ);
            for (int iz = -NZS; iz <= -1; ++iz)
                Y[((size_t)(iz + NZS) * NY1 + iy) * nkx + kx] = cmulr(a_[NZ + iz], invNZ);
            if (fullz) {
                cplx v = cmulr(cmulr(a_[NZH], invNZ), 0.5);
                Y[((size_t)(NZH + NZS) * NY1 + iy) * nkx + kx] = v;
                Y[((size_t)(-NZH + NZS) * NY1 + iy) * nkx + kx] = v;
            }
        }
        SCRATCH_FREE
    }
    free(X);
#pragma omp parallel
    {
        SCRATCH_DECL
#pragma omp for schedule(static)
        for (int t = 0; t < nzk * nkx; ++t) {
            const int izi = t / nkx, kx = t % nkx;
            for (int iy = 0; iy < NY1; ++iy) l_[iy] = Y[((size_t)izi * NY1 + iy) * nkx + kx];
            redft00(&c->py, l_, NY, a_, b_, t_);
            for (int iy = 0; iy <= NYS; ++iy) {
                const double scale = iy == 0 ? 1.0 / (2.0 * NY) : 1.0 / NY;
                U[sidx(c, izi, iy, kx) * stride] = cmulr(l_[iy], scale);
            }
        }
        SCRATCH_FREE
    }
    free(Y);
}

static void coarse_backward(orc_ctx *c, double *UP, const cplx *U, int stride) {
    const orc_config *m = &c->cfg;
    const int NX = m->NX, NY = m->NY, NZ = m->NZ, NXS = m->NXS, NYS = m->NYS, NZS = m->NZS;
    const int NXH = c->NXH, NZH = c->NZH, nkx = c->nkx, nzk = c->nzk, NY1 = NY + 1;
    const int fullx = NXS == NXH, fullz = NZS == NZH;
    const int scratch_n = 2 * (NX > NZ ? NX : NZ) + 4 * NY1;
    const int LDX = NXH + 1;
    cplx *Y = (cplx *)calloc((size_t)nzk * NY1 * nkx, sizeof(cplx));
#pragma omp parallel
    {
        SCRATCH_DECL
#pragma omp for schedule(static)
        for (int t = 0; t < nzk * nkx; ++t) {
            const int izi = t / nkx, kx = t % nkx;
            for (int iy = 0; iy < NY1; ++iy) l_[iy] = 0;
            for (int iy = 0; iy <= NYS && iy < NY1; ++iy) {
                cplx v = U[sidx(c, izi, iy, kx) * stride];
                if (iy == 0) v = v + v;
                l_[iy] = v;
            }
            redft00(&c->py, l_, NY, a_, b_, t_);
            for (int iy = 0; iy < NY1; ++iy) Y[((size_t)izi * NY1 + iy) * nkx + kx] = l_[iy];
        }
        SCRATCH_FREE
    }
    cplx *X = (cplx *)calloc((size_t)NZ * NY1 * LDX, sizeof(cplx));
#pragma omp parallel
    {
        SCRATCH_DECL
#pragma omp for schedule(static)
        for (int t = 0; t < NY1 * nkx; ++t) {
            const int iy = t / nkx, kx = t % nkx;
            for (int iz = 0; iz < NZ; ++iz) {
                const int retained = iz <= NZS || iz >= NZ - NZS;
                cplx v = 0;
                if (retained) {
                    const int kz = iz <= NZS ? iz : iz - NZ;
                    v = cmulr(Y[((size_t)(kz + NZS) * NY1 + iy) * nkx + kx], 0.5);
                    if (fullz && iz == NZH)
                        v = cmulr(Y[((size_t)(NZH + NZS) * NY1 + iy) * nkx + kx], 0.5) +
                            cmulr(Y[((size_t)(-NZH + NZS) * NY1 + iy) * nkx + kx], 0.5);
                }
                l_[iz] = v;
            }
            fft(&c->pz, l_, a_, +1, t_);
            for (int iz = 0; iz < NZ; ++iz) {
                cplx v = a_[iz];
                if (fullx && kx == NXH) v = cmulr(v, 2.0);
                X[((size_t)iz * NY1 + iy) * LDX + kx] = v;
            }
        }
        SCRATCH_FREE
    }
    free(Y);
#pragma omp parallel
    {
        SCRATCH_DECL
#pragma omp for schedule(static)
        for (int t = 0; t < NZ * NY1; ++t)
            c2r_line(&c->px, X + (size_t)t * LDX, LDX, UP + (size_t)t * NX, a_, b_, t_);
        SCRATCH_FREE
    }
    free(X);
}

int orc_xyzfft(orc_ctx *c, char dir, double *phys, double *spec, int stride) {
    if (stride < 1) { snprintf(g_err, sizeof g_err, "spectral stride must be positive"); return -1; }
    if (dir == 'B') coarse_backward(c, phys, (const cplx *)spec, stride);
    else if (dir == 'F') coarse_forward(c, phys, (cplx *)spec, stride);
    else { snprintf(g_err, sizeof g_err, "direction must be F or B"); return -1; }
    return 0;
}

/* ----------------------------------------------------------- vector ops ---- */
/* derivative_y_line (vector_ops.cpp:17-25), writing out[y*os], reading in[y*is]. */
static void dy_line(int ny, const cplx *in, size_t is, cplx *out, size_t os) {
    out[(size_t)ny * os] = 0;
    if (!ny) return;
    out[(size_t)(ny - 1) * os] = cmulr(in[(size_t)ny * is], (double)(2 * ny));
    for (int y = ny - 2; y >= 0; --y)
        out[(size_t)y * os] = out[(size_t)(y + 2) * os] + cmulr(in[(size_t)(y + 1) * is], (double)(2 * (y + 1)));
    out[0] = cmulr(out[0], 0.5);
}

void orc_derivative(orc_ctx *c, int axis, double *outd, const double *ind) {
    cplx *out = (cplx *)outd;
    const cplx *in = (const cplx *)ind;
    const int NYS = c->cfg.NYS;
    /* vector_ops.cpp:40-72 */
#pragma omp parallel for schedule(static)
    for (int izi = 0; izi < c->nzk; ++izi)
        for (int kx = 0; kx < c->nkx; ++kx) {
            if (axis == 0 || axis == 2) {
                const double k = axis == 0 ? c->WAVX[kx] : c->WAVZ[izi];
                for (int iy = 0; iy <= NYS; ++iy) out[sidx(c, izi, iy, kx)] = fd(k, in[sidx(c, izi, iy, kx)]);
            } else {
                dy_line(NYS, in + sidx(c, izi, 0, kx), c->nkx, out + sidx(c, izi, 0, kx), c->nkx);
            }
        }
}

void orc_gradient(orc_ctx *c, double *outd, const double *ind) {
    cplx *out = (cplx *)outd;
    const cplx *in = (const cplx *)ind;
    const int NYS = c->cfg.NYS;
    const size_t col = (size_t)c->nkx;
    /* vector_ops.cpp:73-88 */
#pragma omp parallel for schedule(static)
    for (int izi = 0; izi < c->nzk; ++izi)
        for (int kx = 0; kx < c->nkx; ++kx) {
            dy_line(NYS, in + sidx(c, izi, 0, kx), col, out + sidx(c, izi, 0, kx) * 3 + 1, 3 * col);
            for (int y = 0; y <= NYS; ++y) {
                out[sidx(c, izi, y, kx) * 3 + 0] = fd(c->WAVX[kx], in[sidx(c, izi, y, kx)]);
                out[sidx(c, izi, y, kx) * 3 + 2] = fd(c->WAVZ[izi], in[sidx(c, izi, y, kx)]);
            }
        }
}

void orc_divergence(orc_ctx *c, double *outd, const double *ind) {
    cplx *out = (cplx *)outd;
    const cplx *in = (const cplx *)ind;
    const int NYS = c->cfg.NYS;
    const size_t col = (size_t)c->nkx;
    /* vector_ops.cpp:90-105 */
#pragma omp parallel for schedule(static)
    for (int izi = 0; izi < c->nzk; ++izi)
        for (int kx = 0; kx < c->nkx; ++kx) {
            dy_line(NYS, in + sidx(c, izi, 0, kx) * 3 + 1, 3 * col, out + sidx(c, izi, 0, kx), col);
            for (int y = 0; y <= NYS; ++y) {
                const size_t s = sidx(c, izi, y, kx);
                out[s] = ((cplx)0 + fd(c->WAVX[kx], in[s * 3 + 0])) + out[s] + fd(c->WAVZ[izi], in[s * 3 + 2]);
            }
        }
}

void orc_curl(orc_ctx *c, double *outd, const double *ind) {
    cplx *out = (cplx *)outd;
    const cplx *in = (const cplx *)ind;
    const int NYS = c->cfg.NYS;
    /* enqueue_curl (vector_ops.cpp:112-151) */
#pragma omp parallel for schedule(static)
    for (int izi = 0; izi < c->nzk; ++izi) {
        const double kz = c->WAVZ[izi];
#define O(y, x, k) out[sidx(c, izi, y, x) * 3 + (k)]
#define IN_(y, x, k) in[sidx(c, izi, y, x) * 3 + (k)]
        for (int x = 0; x < c->nkx; ++x) { O(NYS, x, 0) = 0; O(NYS, x, 2) = 0; }
        if (NYS > 0) {
            const double top = (double)(2 * NYS);
            for (int x = 0; x < c->nkx; ++x) {
                O(NYS - 1, x, 0) = cmulr(IN_(NYS, x, 2), top);
                O(NYS - 1, x, 2) = cmulr(IN_(NYS, x, 0), top);
            }
            for (int y = NYS - 2; y >= 0; --y)
                for (int x = 0; x < c->nkx; ++x) {
                    const double scale = (double)(2 * (y + 1));
                    O(y, x, 0) = O(y + 2, x, 0) + cmulr(IN_(y + 1, x, 2), scale);
                    O(y, x, 2) = O(y + 2, x, 2) + cmulr(IN_(y + 1, x, 0), scale);
                }
            for (int x = 0; x < c->nkx; ++x) { O(0, x, 0) = cmulr(O(0, x, 0), 0.5); O(0, x, 2) = cmulr(O(0, x, 2), 0.5); }
        }
        for (int y = 0; y <= NYS; ++y)
            for (int x = 0; x < c->nkx; ++x) {
                const double kx = c->WAVX[x];
                O(y, x, 0) = O(y, x, 0) - fd(kz, IN_(y, x, 1));
                O(y, x, 1) = fd(kz, IN_(y, x, 0)) - fd(kx, IN_(y, x, 2));
                O(y, x, 2) = fd(kx, IN_(y, x, 1)) - O(y, x, 2);
            }
#undef O
#undef IN_
    }
}

void orc_divergence_boundary(orc_ctx *c, double *outd, const double *ind) {
    cplx *out = (cplx *)outd;
    const cplx *in = (const cplx *)ind;
    /* vector_ops.cpp:159-176 */
#pragma omp parallel for schedule(static)
    for (int izi = 0; izi < c->nzk; ++izi)
        for (int x = 0; x < c->nkx; ++x) {
            cplx lower = 0, upper = 0;
            for (int y = 0; y <= c->cfg.NYS; ++y) {
                const size_t s = sidx(c, izi, y, x);
                const cplx tangential = fd(c->WAVX[x], in[s * 3 + 0]) + fd(c->WAVZ[izi], in[s * 3 + 2]);
                const cplx normal = cmulr(in[s * 3 + 1], (double)(y * y));
                lower += tangential + normal;
                upper += (y % 2 == 0) ? tangential - normal : -tangential + normal;
            }
            out[bidx(c, izi, 0, x)] = lower;
            out[bidx(c, izi, 1, x)] = upper;
        }
}

double orc_flow_rate(const orc_ctx *c, const double *fieldd, int stride, int comp) {
    const cplx *f = (const cplx *)fieldd;
    double q = 0.0;
    /* vector_ops.cpp:186-191, simulation.cpp:104-109 */
    for (int iy = 0; iy <= c->cfg.NYS; iy += 2)
        q -= creal(f[sidx(c, c->zoff, iy, 0) * stride + comp]) / (double)(iy * iy - 1);
    return q;
}

static size_t fine_phys_size(const orc_ctx *c) {
    return (size_t)c->cfg.NZ2 * (c->cfg.NY2 + 1) * c->cfg.NX2;
}

void orc_cross_product(orc_ctx *c, double *outd, const double *ad, const double *bd) {
    /* enqueue_cross_product (vector_ops.cpp:250-272) */
    const size_t n = fine_phys_size(c);
    double *ap[3], *bp[3], *pr[3];
    for (int k = 0; k < 3; ++k) {
        ap[k] = (double *)malloc(n * sizeof(double));
        bp[k] = (double *)malloc(n * sizeof(double));
        pr[k] = (double *)malloc(n * sizeof(double));
    }
    for (int k = 0; k < 3; ++k) {
        fine_backward(c, ap[k], (const cplx *)ad + k, 3);
        fine_backward(c, bp[k], (const cplx *)bd + k, 3);
    }
#pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n; ++i) {
        const double a0 = ap[0][i], a1 = ap[1][i], a2 = ap[2][i];
        const double b0 = bp[0][i], b1 = bp[1][i], b2 = bp[2][i];
        pr[0][i] = a1 * b2 - a2 * b1;
        pr[1][i] = a2 * b0 - a0 * b2;
        pr[2][i] = a0 * b1 - a1 * b0;
    }
    for (int k = 0; k < 3; ++k) fine_forward(c, pr[k], (cplx *)outd + k, 3);
    for (int k = 0; k < 3; ++k) { free(ap[k]); free(bp[k]); free(pr[k]); }
}

void orc_nonlinear_term(orc_ctx *c, double *nld, const double *veld, double *pbcd) {
    /* TimeStepper::enqueue_nonlinear_term (timestepper.cpp:19-43) */
    cplx *nl = (cplx *)nld, *bc = (cplx *)pbcd;
    orc_curl(c, (double *)c->curl_ws, veld);
    orc_cross_product(c, nld, veld, (const double *)c->curl_ws);
    if (c->body)
        for (size_t s = 0; s < 3 * c->S; ++s) nl[s] += c->body[s];
    const double Re = c->cfg.Re;
#pragma omp parallel for schedule(static)
    for (int izi = 0; izi < c->nzk; ++izi) {
        for (int x = 0; x < c->nkx; ++x) { bc[bidx(c, izi, 0, x)] = 0; bc[bidx(c, izi, 1, x)] = 0; }
        const cplx czre = mk(0.0, c->WAVZ[izi] / Re);
        for (int y = 0; y <= c->cfg.NYS; ++y)
            for (int x = 0; x < c->nkx; ++x) {
                const cplx cxre = cdivr(mk(0.0, c->WAVX[x]), Re);
                const size_t s = sidx(c, izi, y, x) * 3;
                const cplx wall = cmul(cxre, c->curl_ws[s + 2]) - cmul(czre, c->curl_ws[s + 0]) + nl[s + 1];
                bc[bidx(c, izi, 0, x)] += wall;
                bc[bidx(c, izi, 1, x)] += (y % 2 == 0) ? -wall : wall;
            }
    }
}

/* --------------------------------------------------------------- solver ---- */
/* solve_tau_column (eqsolve_pv.cpp:16-60); one (kx,kz) column, rhs/out/ad stride 1. */
static void solve_tau_col(const orc_ctx *c, int ny, double g, double a, double b, int zero_mode,
                          cplx lower, cplx upper, double w, const cplx *rhs, cplx *out, const double *ad) {
    const double *WL = c->WL, *WD = c->WD, *WR = c->WR;
    const int normalize = g == 0.0 && a == 0.0 && zero_mode;
    cplx correction = 0;
    if (normalize) {
        for (int y = 2; y <= ny; y += 2) correction += cdivr(rhs[y], (double)(y * y - 1));
        if (b != 0.0) correction += cdivr(lower + upper, 2.0 * b);
    }
#define R(y) ((normalize && (y) == 0) ? correction : rhs[y])
    out[0] = cdivr(lower + upper, 2.0);
    out[1] = cdivr(lower - upper, 2.0);
    for (int iy = 2; iy <= ny - 2; ++iy)
        out[iy] = cmulr(R(iy - 2), WL[iy]) + cmulr(R(iy), WD[iy]) + cmulr(R(iy + 2), WR[iy]);
    out[ny - 1] = cmulr(R(ny - 3), WL[ny - 1]) + cmulr(R(ny - 1), WD[ny - 1]);
    out[ny] = cmulr(R(ny - 2), WL[ny]) + cmulr(R(ny), WD[ny]);
#undef R
    for (int iy = ny - 2; iy >= 2; --iy) out[iy] = out[iy] - cmulr(out[iy + 2], WR[iy] * w * ad[iy + 2]);
    double ad01[2] = {0, 0};
    for (int iy = ny; iy >= ny - 1; --iy) {
        const int q = iy % 2;
        ad01[q] = (a == 0.0 && b == 0.0) ? (double)((1 - q) * iy * iy + q) : a + b * iy * iy;
    }
    for (int iy = ny - 2; iy >= 0; --iy) {
        const int q = iy % 2;
        const double func = (a == 0.0 && b == 0.0) ? (double)((1 - q) * iy * iy + q) : a + b * iy * iy;
        out[q] = out[q] - cmulr(out[iy + 2], ad01[q] * ad[iy + 2]);
        ad01[q] = func - WL[iy + 2] * w * ad01[q] * ad[iy + 2];
    }
    if (normalize) { out[0] = 0; ad01[0] = 1.0; }
    out[0] = cdivr(out[0], ad01[0]);
    out[1] = cdivr(out[1], ad01[1]);
    for (int iy = 2; iy <= ny; ++iy) out[iy] = cmulr(out[iy] - cmulr(out[iy - 2], WL[iy] * w), ad[iy]);
}

/* build_wn_ad (eqsolve_pv.cpp:148-168) */
static void build_wn_ad(orc_ctx *c, double *wn, double *ad, double g) {
    const int NYS = c->cfg.NYS;
    for (int izi = 0; izi < c->nzk; ++izi)
        for (int ix = 0; ix < c->nkx; ++ix)
            wn[izi * c->nkx + ix] = -(c->WAVX[ix] * c->WAVX[ix] + c->WAVZ[izi] * c->WAVZ[izi] + g);
#pragma omp parallel for schedule(static)
    for (int izi = 0; izi < c->nzk; ++izi)
        for (int ix = 0; ix < c->nkx; ++ix) {
            const double w = wn[izi * c->nkx + ix];
            ad[sidx(c, izi, NYS, ix)] = 1.0 / (1.0 + c->WD[NYS] * w);
            ad[sidx(c, izi, NYS - 1, ix)] = 1.0 / (1.0 + c->WD[NYS - 1] * w);
            for (int iy = NYS - 2; iy >= 2; --iy) {
                const double fc = c->WR[iy] * c->WL[iy + 2];
                const double den = 1.0 + (c->WD[iy] - fc * w * ad[sidx(c, izi, iy + 2, ix)]) * w;
                ad[sidx(c, izi, iy, ix)] = 1.0 / den;
            }
        }
}

/* Extract column (izi,kx) of a strided field into col[0..NYS]. */
static void get_col(const orc_ctx *c, const cplx *f, int stride, int comp, int izi, int kx, cplx *col) {
    for (int y = 0; y < c->nj; ++y) col[y] = f[sidx(c, izi, y, kx) * stride + comp];
}
static void get_rcol(const orc_ctx *c, const double *f, int izi, int kx, double *col) {
    for (int y = 0; y < c->nj; ++y) col[y] = f[sidx(c, izi, y, kx)];
}

/* EqSolvePV::solve_tau (eqsolve_pv.cpp:170-188) */
static void solve_tau_field(orc_ctx *c, cplx *out, int ostride, int ocomp, const cplx *rhs, double g,
                            double a, double b, const cplx *bc, int bstride, int bcomp, int which) {
    const double *wn = which ? c->wn_v : c->wn_p;
    const double *ad = which ? c->ad_v : c->ad_p;
    const int ny = c->cfg.NYS, nj = c->nj;
#pragma omp parallel
    {
        cplx *r = (cplx *)malloc(sizeof(cplx) * nj), *o = (cplx *)malloc(sizeof(cplx) * nj);
        double *adc = (double *)malloc(sizeof(double) * nj);
#pragma omp for schedule(static)
        for (int t = 0; t < c->nzk * c->nkx; ++t) {
            const int izi = t / c->nkx, x = t % c->nkx;
            get_col(c, rhs, 1, 0, izi, x, r);
            get_rcol(c, ad, izi, x, adc);
            solve_tau_col(c, ny, g, a, b, izi == c->zoff && x == 0, bc[bidx(c, izi, 0, x) * bstride + bcomp],
                          bc[bidx(c, izi, 1, x) * bstride + bcomp], wn[izi * c->nkx + x], r, o, adc);
            for (int y = 0; y < nj; ++y) out[sidx(c, izi, y, x) * ostride + ocomp] = o[y];
        }
        free(r); free(o); free(adc);
    }
}

void orc_solve_tau(orc_ctx *c, double *out, const double *rhs, double g, double a, double b,
                   const double *bc, int which) {
    solve_tau_field(c, (cplx *)out, 1, 0, (const cplx *)rhs, g, a, b, (const cplx *)bc, 1, 0, which);
}

/* solve_coupled_base (eqsolve_pv.cpp:241-264) */
static void solve_coupled_base(orc_ctx *c, cplx *pressure, const cplx *rhs_p, const cplx *pbc,
                               cplx *vel, const cplx *rhs_v, const cplx *vbc) {
    const orc_config *m = &c->cfg;
    solve_tau_field(c, pressure, 1, 0, rhs_p, m->G_pre, c->pressure_a, c->pressure_bcoef, pbc, 1, 0, 0);
    cplx *grad = (cplx *)malloc(sizeof(cplx) * 3 * c->S);
    cplx *forcing = (cplx *)malloc(sizeof(cplx) * c->S);
    orc_gradient(c, (double *)grad, (const double *)pressure);
    for (int k = 0; k < 3; ++k) {
        for (size_t s = 0; s < c->S; ++s) forcing[s] = cmulr(grad[s * 3 + k] - rhs_v[s * 3 + k], m->Re);
        solve_tau_field(c, vel, 3, k, forcing, c->g_v, m->a_V[k], m->b_V[k], vbc, 3, k, 1);
    }
    free(grad);
    free(forcing);
}

/* solve_modes (eqsolve_pv.cpp:311-450), one (kx,kz) lane at a time. */
void orc_solve_modes(orc_ctx *c, double *pd, const double *rpd, const double *pbcd, double *vd,
                     const double *rvd, const double *vbcd, int correct_sigma) {
    const orc_config *m = &c->cfg;
    cplx *pressure = (cplx *)pd, *velocity = (cplx *)vd;
    const cplx *rhs_p = (const cplx *)rpd, *pbc = (const cplx *)pbcd, *rhs_v = (const cplx *)rvd,
               *vbc = (const cplx *)vbcd;
    const int fuse_rhs = rhs_p == NULL;
    const int ny = m->NYS, nj = c->nj;
    const double Re = m->Re;
#pragma omp parallel
    {
        cplx *buf = (cplx *)malloc(sizeof(cplx) * nj * 8);
        double *adp = (double *)malloc(sizeof(double) * nj), *adv = (double *)malloc(sizeof(double) * nj);
        cplx *pc = buf, *dy = buf + nj, *forcing = buf + 2 * nj;
        cplx *rv[3] = {buf + 3 * nj, buf + 4 * nj, buf + 5 * nj};
        cplx *vc[3] = {rv[0], rv[1], rv[2]};
        cplx *rp = dy, *prhs = buf + 6 * nj;
#pragma omp for schedule(static)
        for (int t = 0; t < c->nzk * c->nkx; ++t) {
            const int izi = t / c->nkx, x = t % c->nkx;
            const int zero = izi == c->zoff && x == 0;
            const double kx = c->WAVX[x], kz = c->WAVZ[izi];
            for (int k = 0; k < 3; ++k) get_col(c, rhs_v, 3, k, izi, x, rv[k]);
            if (fuse_rhs) {
                rp[ny] = 0;
                rp[ny - 1] = cmulr(rv[1][ny], (double)(2 * ny));
                for (int y = ny - 2; y >= 0; --y) rp[y] = rp[y + 2] + cmulr(rv[1][y + 1], (double)(2 * (y + 1)));
                rp[0] = cmulr(rp[0], 0.5);
                for (int y = 0; y <= ny; ++y) rp[y] = ((cplx)0 + fd(kx, rv[0][y])) + rp[y] + fd(kz, rv[2][y]);
            } else {
                get_col(c, rhs_p, 1, 0, izi, x, prhs);
            }
            get_rcol(c, c->ad_p, izi, x, adp);
            get_rcol(c, c->ad_v, izi, x, adv);
            solve_tau_col(c, ny, m->G_pre, c->pressure_a, c->pressure_bcoef, zero, pbc[bidx(c, izi, 0, x)],
                          pbc[bidx(c, izi, 1, x)], c->wn_p[izi * c->nkx + x], fuse_rhs ? rp : prhs, pc, adp);
            dy[ny] = 0;
            dy[ny - 1] = cmulr(pc[ny], (double)(2 * ny));
            for (int y = ny - 2; y >= 0; --y) dy[y] = dy[y + 2] + cmulr(pc[y + 1], (double)(2 * (y + 1)));
            dy[0] = cmulr(dy[0], 0.5);
            const cplx tail0 = rv[1][ny], tail1 = rv[1][ny - 1];
            for (int k = 0; k < 3; ++k) {
                for (int y = 0; y <= ny; ++y) {
                    const cplx grad = k == 0 ? fd(kx, pc[y]) : (k == 1 ? dy[y] : fd(kz, pc[y]));
                    forcing[y] = cmulr(grad - rv[k][y], Re);
                }
                solve_tau_col(c, ny, c->g_v, m->a_V[k], m->b_V[k], zero, vbc[bidx(c, izi, 0, x) * 3 + k],
                              vbc[bidx(c, izi, 1, x) * 3 + k], c->wn_v[izi * c->nkx + x], forcing, vc[k], adv);
            }
            cplx div0 = 0, div1 = 0, sigma[2] = {0, 0};
            for (int y = 0; y <= ny; ++y) {
                const cplx tang = fd(kx, vc[0][y]) + fd(kz, vc[2][y]);
                const cplx normal = cmulr(vc[1][y], (double)(y * y));
                div0 += tang + normal;
                div1 += y % 2 == 0 ? tang - normal : -tang + normal;
            }
            const cplx c0 = cmul(div1, c->ID[1][bidx(c, izi, 0, x)]) - cmul(div0, c->ID[1][bidx(c, izi, 1, x)]);
            const cplx c1 = cmul(div0, c->ID[0][bidx(c, izi, 1, x)]) - cmul(div1, c->ID[0][bidx(c, izi, 0, x)]);
            if (correct_sigma) {
                const cplx ptop = pc[ny] + (cmul(c0, c->PI[0][sidx(c, izi, ny, x)]) + cmul(c1, c->PI[1][sidx(c, izi, ny, x)]));
#define VT(y) (vc[1][y] + (cmul(c0, c->VI[0][sidx(c, izi, y, x) * 3 + 1]) + cmul(c1, c->VI[1][sidx(c, izi, y, x) * 3 + 1])))
                const double w = c->wn_v[izi * c->nkx + x];
                sigma[0] = cdiv(cmulr(VT(ny), -w) - cmulr(tail0, Re), c->SIG0[bidx(c, izi, 0, x)]);
                sigma[1] = cdiv(cmulr(VT(ny - 1), -w) + cmulr(cmulr(ptop, 2.0 * ny) - tail1, Re), c->SIG0[bidx(c, izi, 1, x)]);
#undef VT
            }
            for (int y = 0; y <= ny; ++y) {
                const int q = (y + 1) % 2;
                const size_t s = sidx(c, izi, y, x);
                cplx p = pc[y] + (cmul(c0, c->PI[0][s]) + cmul(c1, c->PI[1][s]));
                if (correct_sigma) p += cmul(sigma[q], c->PB[s]);
                pressure[s] = p;
                for (int k = 0; k < 3; ++k) {
                    cplx v = vc[k][y] + (cmul(c0, c->VI[0][s * 3 + k]) + cmul(c1, c->VI[1][s * 3 + k]));
                    if (correct_sigma) v += cmul(sigma[k == 1 ? 1 - q : q], c->VB[s * 3 + k]);
                    velocity[s * 3 + k] = v;
                }
            }
        }
        free(buf); free(adp); free(adv);
    }
}

/* The staged composition the reference's tests compare against solve_modes
 * (tests/test_eqsolve.cpp:17-47, with apply_sigma_correction eqsolve_pv.cpp:511-533). */
void orc_solve_staged(orc_ctx *c, double *pd, const double *rpd, const double *pbcd, double *vd,
                      const double *rvd, const double *vbcd, int correct) {
    cplx *pressure = (cplx *)pd, *velocity = (cplx *)vd;
    const cplx *rhs_v = (const cplx *)rvd;
    solve_coupled_base(c, pressure, (const cplx *)rpd, (const cplx *)pbcd, velocity, rhs_v, (const cplx *)vbcd);
    cplx *div = (cplx *)malloc(sizeof(cplx) * c->BND);
    orc_divergence_boundary(c, (double *)div, vd);
    const int ny = c->cfg.NYS;
    const double Re = c->cfg.Re;
    for (int izi = 0; izi < c->nzk; ++izi)
        for (int x = 0; x < c->nkx; ++x) {
            const cplx d0 = div[bidx(c, izi, 0, x)], d1 = div[bidx(c, izi, 1, x)];
            const cplx c0 = cmul(d1, c->ID[1][bidx(c, izi, 0, x)]) - cmul(d0, c->ID[1][bidx(c, izi, 1, x)]);
            const cplx c1 = cmul(d0, c->ID[0][bidx(c, izi, 1, x)]) - cmul(d1, c->ID[0][bidx(c, izi, 0, x)]);
            for (int y = 0; y <= ny; ++y) {
                const size_t s = sidx(c, izi, y, x);
                pressure[s] += cmul(c0, c->PI[0][s]) + cmul(c1, c->PI[1][s]);
                for (int k = 0; k < 3; ++k)
                    velocity[s * 3 + k] += cmul(c0, c->VI[0][s * 3 + k]) + cmul(c1, c->VI[1][s * 3 + k]);
            }
        }
    free(div);
    if (!correct) return;
    for (int izi = 0; izi < c->nzk; ++izi)
        for (int x = 0; x < c->nkx; ++x) {
            const double w = c->wn_v[izi * c->nkx + x];
            cplx sigma[2];
            sigma[0] = cdiv(cmulr(velocity[sidx(c, izi, ny, x) * 3 + 1], -w) - cmulr(rhs_v[sidx(c, izi, ny, x) * 3 + 1], Re),
                            c->SIG0[bidx(c, izi, 0, x)]);
            sigma[1] = cdiv(cmulr(velocity[sidx(c, izi, ny - 1, x) * 3 + 1], -w) +
                                cmulr(cmulr(pressure[sidx(c, izi, ny, x)], 2.0 * ny) - rhs_v[sidx(c, izi, ny - 1, x) * 3 + 1], Re),
                            c->SIG0[bidx(c, izi, 1, x)]);
            for (int y = 0; y <= ny; ++y) {
                const int q = (y + 1) % 2;
                const size_t s = sidx(c, izi, y, x);
                pressure[s] += cmul(sigma[q], c->PB[s]);
                velocity[s * 3 + 0] += cmul(sigma[q], c->VB[s * 3 + 0]);
                velocity[s * 3 + 1] += cmul(sigma[1 - q], c->VB[s * 3 + 1]);
                velocity[s * 3 + 2] += cmul(sigma[q], c->VB[s * 3 + 2]);
            }
        }
}

/* build_influence_coefficients (eqsolve_pv.cpp:266-303) */
static void build_influence(orc_ctx *c) {
    cplx *rhs_p = (cplx *)calloc(c->S, sizeof(cplx));
    cplx *rhs_v = (cplx *)calloc(3 * c->S, sizeof(cplx));
    cplx *bc_v = (cplx *)calloc(3 * c->BND, sizeof(cplx));
    cplx *bc_p = (cplx *)calloc(c->BND, sizeof(cplx));
    for (int i = 0; i < 2; ++i) {
        for (int izi = 0; izi < c->nzk; ++izi)
            for (int x = 0; x < c->nkx; ++x) {
                bc_p[bidx(c, izi, 0, x)] = (double)(1 - i);
                bc_p[bidx(c, izi, 1, x)] = (double)i;
            }
        solve_coupled_base(c, c->PI[i], rhs_p, bc_p, c->VI[i], rhs_v, bc_v);
        orc_divergence_boundary(c, (double *)c->ID[i], (const double *)c->VI[i]);
    }
    for (int izi = 0; izi < c->nzk; ++izi)
        for (int x = 0; x < c->nkx; ++x) {
            cplx den = cmul(c->ID[0][bidx(c, izi, 0, x)], c->ID[1][bidx(c, izi, 1, x)]) -
                       cmul(c->ID[0][bidx(c, izi, 1, x)], c->ID[1][bidx(c, izi, 0, x)]);
            const int root = x == 0 && izi == c->zoff;
            if (root) den = c->cfg.a_V[1] == 0.0 ? mk(1.0, 0.0) : cdiv(mk(0.25, 0.0), c->ID[0][bidx(c, c->zoff, 0, 0)]);
            for (int wall = 0; wall < 2; ++wall)
                for (int i = 0; i < 2; ++i)
                    c->ID[i][bidx(c, izi, wall, x)] =
                        root ? (wall == 0 ? den : 0) : cdiv(c->ID[i][bidx(c, izi, wall, x)], den);
        }
    free(rhs_p); free(rhs_v); free(bc_v); free(bc_p);
}

/* build_b_problem (eqsolve_pv.cpp:452-477) */
static void build_b_problem(orc_ctx *c) {
    const int NYS = c->cfg.NYS;
    cplx *rhs_p = (cplx *)calloc(c->S, sizeof(cplx));
    for (int izi = 0; izi < c->nzk; ++izi)
        for (int x = 0; x < c->nkx; ++x) {
            rhs_p[sidx(c, izi, 0, x)] = (double)(NYS - 1);
            for (int iy = 1; iy < NYS; ++iy) rhs_p[sidx(c, izi, iy, x)] = (double)(2 * (NYS - ((iy + 1) % 2)));
        }
    cplx *bc_p = (cplx *)calloc(c->BND, sizeof(cplx));
    cplx *rhs_v = (cplx *)calloc(3 * c->S, sizeof(cplx));
    cplx *bc_v = (cplx *)calloc(3 * c->BND, sizeof(cplx));
    orc_solve_modes(c, (double *)c->PB, (const double *)rhs_p, (const double *)bc_p, (double *)c->VB,
                    (const double *)rhs_v, (const double *)bc_v, 0);
    const double Re = c->cfg.Re;
    for (int izi = 0; izi < c->nzk; ++izi)
        for (int x = 0; x < c->nkx; ++x) {
            const double w = c->wn_v[izi * c->nkx + x];
            c->SIG0[bidx(c, izi, 0, x)] = Re + cmulr(c->VB[sidx(c, izi, NYS, x) * 3 + 1], w);
            c->SIG0[bidx(c, izi, 1, x)] = Re + cmulr(c->VB[sidx(c, izi, NYS - 1, x) * 3 + 1], w) -
                                          cmulr(c->PB[sidx(c, izi, NYS, x)], Re * (2.0 * NYS));
        }
    c->SIG0[bidx(c, c->zoff, 1, 0)] = 1.0;
    free(rhs_p); free(bc_p); free(rhs_v); free(bc_v);
}

/* build_pressure_gradient_responses (eqsolve_pv.cpp:479-493) */
static void build_pg_responses(orc_ctx *c) {
    const orc_config *m = &c->cfg;
    cplx *rhs = (cplx *)calloc(c->S, sizeof(cplx)), *solved = (cplx *)calloc(c->S, sizeof(cplx));
    cplx *bc = (cplx *)calloc(c->BND, sizeof(cplx));
    rhs[sidx(c, c->zoff, 0, 0)] = m->Re;
    solve_tau_field(c, solved, 1, 0, rhs, c->g_v, m->a_V[0], m->b_V[0], bc, 1, 0, 1);
    c->qx_unit = orc_flow_rate(c, (const double *)solved, 1, 0);
    for (int iy = 0; iy <= m->NYS; ++iy) c->u_dpdx[iy] = creal(solved[sidx(c, c->zoff, iy, 0)]);
    solve_tau_field(c, solved, 1, 0, rhs, c->g_v, m->a_V[2], m->b_V[2], bc, 1, 0, 1);
    c->qz_unit = orc_flow_rate(c, (const double *)solved, 1, 0);
    for (int iy = 0; iy <= m->NYS; ++iy) c->w_dpdz[iy] = creal(solved[sidx(c, c->zoff, iy, 0)]);
    free(rhs); free(solved); free(bc);
}

/* EqSolvePV::setup (eqsolve_pv.cpp:495-509) */
void orc_setup(orc_ctx *c, double g_v) {
    const orc_config *m = &c->cfg;
    c->g_v = g_v;
    build_wn_ad(c, c->wn_p, c->ad_p, m->G_pre);
    build_wn_ad(c, c->wn_v, c->ad_v, g_v);
    c->pressure_a = 1.0;
    c->pressure_bcoef = 0.0;
    build_influence(c);
    build_b_problem(c);
    c->pressure_a = m->a_pre0;
    c->pressure_bcoef = m->b_pre0;
    build_pg_responses(c);
    c->setup_done = 1;
}

/* g of the current velocity tables (NaN before any setup). */
double orc_setup_g(const orc_ctx *c) { return c->setup_done ? c->g_v : NAN; }

void orc_get_table(const orc_ctx *c, int which, double *out) {
    const size_t S = c->S, B = c->BND, M = c->MODE;
    switch (which) {
    case 0: memcpy(out, c->wn_p, M * sizeof(double)); break;
    case 1: memcpy(out, c->wn_v, M * sizeof(double)); break;
    case 2: memcpy(out, c->ad_p, S * sizeof(double)); break;
    case 3: memcpy(out, c->ad_v, S * sizeof(double)); break;
    case 4: memcpy(out, c->PI[0], S * sizeof(cplx)); break;
    case 5: memcpy(out, c->PI[1], S * sizeof(cplx)); break;
    case 6: memcpy(out, c->VI[0], 3 * S * sizeof(cplx)); break;
    case 7: memcpy(out, c->VI[1], 3 * S * sizeof(cplx)); break;
    case 8: memcpy(out, c->ID[0], B * sizeof(cplx)); break;
    case 9: memcpy(out, c->ID[1], B * sizeof(cplx)); break;
    case 10: memcpy(out, c->PB, S * sizeof(cplx)); break;
    case 11: memcpy(out, c->VB, 3 * S * sizeof(cplx)); break;
    case 12: memcpy(out, c->SIG0, B * sizeof(cplx)); break;
    case 13: memcpy(out, c->u_dpdx, c->nj * sizeof(double)); break;
    case 14: memcpy(out, c->w_dpdz, c->nj * sizeof(double)); break;
    case 15: out[0] = c->qx_unit; out[1] = c->qz_unit; break;
    default: break;
    }
}

/* ----------------------------------------------------------- simulation ---- */
/* step_coefficients (simulation.cpp:36-45) */
void orc_step_coefficients(int it, double dt, double Re, double out[5]) {
    const int eff = it < 3 ? it : 2;
    const double a0 = (double)eff / dt, a1 = (double)(1 - eff) / (2.0 * dt);
    out[0] = a0; out[1] = a1; out[2] = (double)eff; out[3] = (double)(1 - eff); out[4] = (a0 + a1) * Re;
}

static size_t field_size(const orc_ctx *c, int f) {
    switch (f) {
    case F_P: return c->S;
    case F_PBC: case F_PBC_PREV: return c->BND;
    case F_VBC: return 3 * c->BND;
    default: return 3 * c->S;
    }
}
void orc_set_state(orc_ctx *c, int f, const double *d) { memcpy(c->st[f], d, field_size(c, f) * sizeof(cplx)); }
void orc_get_state(const orc_ctx *c, int f, double *d) { memcpy(d, c->st[f], field_size(c, f) * sizeof(cplx)); }

/* correct_zero_mode (timestepper.cpp:60-68) */
static void correct_zero_mode(orc_ctx *c, cplx *vel, const cplx *vbc) {
    const cplx lower = vbc[bidx(c, c->zoff, 0, 0) * 3 + 1], upper = vbc[bidx(c, c->zoff, 1, 0) * 3 + 1];
    vel[sidx(c, c->zoff, 0, 0) * 3 + 1] = cmulr(lower + upper, 0.5);
    for (int iy = 1; iy <= c->cfg.NYS; ++iy) vel[sidx(c, c->zoff, iy, 0) * 3 + 1] = 0;
}

/* correct_flow (simulation.cpp:111-140) */
static void correct_flow(orc_ctx *c) {
    const orc_config *m = &c->cfg;
    cplx *vel = c->st[F_VEL];
    const double qx = orc_flow_rate(c, (const double *)vel, 3, 0);
    const double qz = orc_flow_rate(c, (const double *)vel, 3, 2);
    const double theta = m->THETA * acos(-1.0) / 180.0;
    if (m->INDEX_GRAD_PRE == 1) {
        c->dpdx = -cos(theta);
        c->dpdz = -sin(theta);
    } else {
        c->dpdx = (cos(theta) - qx) / c->qx_unit;
        c->dpdz = (sin(theta) - qz) / c->qz_unit;
    }
    for (int iy = 0; iy <= m->NYS; ++iy) {
        vel[sidx(c, c->zoff, iy, 0) * 3 + 0] += c->dpdx * c->u_dpdx[iy];
        vel[sidx(c, c->zoff, iy, 0) * 3 + 2] += c->dpdz * c->w_dpdz[iy];
    }
}

void orc_step(orc_ctx *c, int it) {
    const orc_config *m = &c->cfg;
    double co[5];
    orc_step_coefficients(it, m->DT, m->Re, co);
    if (it < 3) orc_setup(c, co[4]);
    c->time += m->DT;
    /* enqueue_serial_step (simulation.cpp:142-171) */
    orc_nonlinear_term(c, (double *)c->st[F_NL], (const double *)c->st[F_VEL], (double *)c->st[F_PBC]);
    for (size_t s = 0; s < c->BND; ++s)
        c->st[F_PBC_PREV][s] = cmulr(c->st[F_PBC][s], co[2]) + cmulr(c->st[F_PBC_PREV][s], co[3]);
    for (size_t s = 0; s < 3 * c->S; ++s)
        c->st[F_NL_PREV][s] = cmulr(c->st[F_VEL][s], co[0]) + cmulr(c->st[F_VEL_PREV][s], co[1]) +
                              cmulr(c->st[F_NL][s], co[2]) + cmulr(c->st[F_NL_PREV][s], co[3]);
    cplx *t = c->st[F_VEL_PREV]; c->st[F_VEL_PREV] = c->st[F_VEL]; c->st[F_VEL] = t;
    orc_solve_modes(c, (double *)c->st[F_P], NULL, (const double *)c->st[F_PBC_PREV], (double *)c->st[F_VEL],
                    (const double *)c->st[F_NL_PREV], (const double *)c->st[F_VBC], 1);
    correct_zero_mode(c, c->st[F_VEL], c->st[F_VBC]);
    t = c->st[F_NL_PREV]; c->st[F_NL_PREV] = c->st[F_NL]; c->st[F_NL] = t;
    t = c->st[F_PBC_PREV]; c->st[F_PBC_PREV] = c->st[F_PBC]; c->st[F_PBC] = t;
    correct_flow(c);
}

/* Constant body force f (spectral vector, reference layout) added to the nonlinear term of
 * every step, so that it enters the RHS with weight b0 + b1 = 1 and the pressure wall
 * condition through f_y.  Not in the reference (its only forcing is the mean gradient);
 * used by the discrete adjoint's gradient with respect to a forcing.  NULL removes it. */
void orc_set_body_force(orc_ctx *c, const double *f) {
    if (!f) { free(c->body); c->body = NULL; return; }
    if (!c->body) c->body = (cplx *)calloc(3 * c->S, sizeof(cplx));
    memcpy(c->body, f, 3 * c->S * sizeof(cplx));
}

void orc_run(orc_ctx *c, int nsteps) {
    for (int it = 1; it <= nsteps; ++it) orc_step(c, it);
}
void orc_flow_gradients(const orc_ctx *c, double out[2]) { out[0] = c->dpdx; out[1] = c->dpdz; }
double orc_time(const orc_ctx *c) { return c->time; }

/* ---------------------------------------------------------- mt19937_64 ---- */
/* The 64-bit Mersenne Twister of the C++ standard library (initial_bc.hpp:116). */
void orc_mt64_seed(orc_mt64 *g, uint64_t seed) {
    g->mt[0] = seed;
    for (int i = 1; i < 312; ++i)
        g->mt[i] = 6364136223846793005ULL * (g->mt[i - 1] ^ (g->mt[i - 1] >> 62)) + (uint64_t)i;
    g->idx = 312;
}
uint64_t orc_mt64_next(orc_mt64 *g) {
    if (g->idx >= 312) {
        for (int i = 0; i < 312; ++i) {
            uint64_t x = (g->mt[i] & 0xFFFFFFFF80000000ULL) | (g->mt[(i + 1) % 312] & 0x7FFFFFFFULL);
            uint64_t xa = x >> 1;
            if (x & 1ULL) xa ^= 0xB5026F5AA96619E9ULL;
            g->mt[i] = g->mt[(i + 156) % 312] ^ xa;
        }
        g->idx = 0;
    }
    uint64_t y = g->mt[g->idx++];
    y ^= (y >> 29) & 0x5555555555555555ULL;
    y ^= (y << 17) & 0x71D67FFFEDA60000ULL;
    y ^= (y << 37) & 0xFFF7EEE000000000ULL;
    y ^= (y >> 43);
    return y;
}

/* ---------------------------------------------------------- statistics / diagnostics ---- */

/* Statistics::reduce plane sums (statistics.cpp:54-90): four interleaved lanes per x row, lanes
 * combined as (l0 + l1) + (l2 + l3); p = raw pressure - |u|^2/2.  Physical arrays [iz][iy][ix]. */
void orc_plane_moments(const orc_ctx *c, const double *u, const double *v, const double *w, const double *pr,
                       int higher, double *sums, int *finite) {
    const int NX = c->cfg.NX, NZ = c->cfg.NZ, NL = c->cfg.NY + 1;
#pragma omp parallel for schedule(static)
    for (int iy = 0; iy < NL; ++iy) {
        double acc[20][4];
        memset(acc, 0, sizeof acc);
        int ok = 1;
        for (int iz = 0; iz < NZ; ++iz) {
            const size_t base = ((size_t)iz * NL + iy) * NX;
            for (int ix = 0; ix < NX; ix += 4)
                for (int lane = 0; lane < 4 && ix + lane < NX; ++lane) {
                    const size_t i = base + ix + lane;
                    const double U = u[i], V = v[i], W = w[i];
                    const double uu = U * U, vv = V * V, ww = W * W;
                    const double raw = pr[i];
                    const double P = raw - 0.5 * (uu + vv + ww);
                    ok &= isfinite(U) && isfinite(V) && isfinite(W) && isfinite(raw) && isfinite(P);
                    acc[0][lane] += U; acc[1][lane] += V; acc[2][lane] += W; acc[3][lane] += P;
                    acc[4][lane] += uu; acc[5][lane] += vv; acc[6][lane] += ww;
                    if (higher) {
                        acc[7][lane] += U * V; acc[8][lane] += V * W; acc[9][lane] += W * U;
                        acc[10][lane] += P * U; acc[11][lane] += P * V; acc[12][lane] += P * W;
                        acc[13][lane] += P * P;
                        acc[14][lane] += uu * U; acc[15][lane] += vv * V; acc[16][lane] += ww * W;
                        acc[17][lane] += uu * uu; acc[18][lane] += vv * vv; acc[19][lane] += ww * ww;
                    }
                }
        }
        for (int m = 0; m < 20; ++m) {
            sums[(size_t)iy * 20 + m] = (acc[m][0] + acc[m][1]) + (acc[m][2] + acc[m][3]);
            ok &= isfinite(sums[(size_t)iy * 20 + m]);
        }
        finite[iy] = ok;
    }
}

/* Diagnostics weights (diagnostics.cpp:17-33): Clenshaw-Curtis on [-1, 1], divided by 2. */
void orc_cc_weights(const orc_ctx *c, double *wt) {
    const int n = c->cfg.NY;
    const double pi = acos(-1.0);
    for (int j = 0; j <= n; ++j) {
        const double theta = pi * j / n;
        double weight;
        if (j == 0 || j == n) weight = 1.0 / (n * n - (n % 2 == 0 ? 1.0 : 0.0));
        else {
            double value = 1.0;
            for (int k = 1; k <= (n - 1) / 2; ++k) value -= 2.0 * cos(2 * k * theta) / (4.0 * k * k - 1.0);
            if (n % 2 == 0) value -= cos(n * theta) / (n * n - 1.0);
            weight = 2.0 * value / n;
        }
        wt[j] = weight / 2.0;
    }
}

/* Diagnostics::evaluate (diagnostics.cpp:39-75): out = {EKT, TKE, divergence L2} from the plane
 * sums of the physical fields and the spectral velocity (reference layout, interleaved). */
int orc_diagnostics(orc_ctx *c, const double *sums, const int *plane_finite, const double *velocity,
                    int spectral_finite, double out[3]) {
    const int NX = c->cfg.NX, NZ = c->cfg.NZ, NL = c->cfg.NY + 1;
    double *div = (double *)calloc(c->S * 2, sizeof(double));
    double *phys = (double *)malloc((size_t)NX * NZ * NL * sizeof(double));
    double *wt = (double *)malloc(NL * sizeof(double));
    orc_divergence(c, div, velocity);
    coarse_backward(c, phys, (const cplx *)div, 1);
    orc_cc_weights(c, wt);
    int finite = spectral_finite;
    const double inv_plane = 1.0 / ((double)NX * NZ);
    double ekt = 0.0, tke = 0.0, div_square = 0.0;
    for (int iy = 0; iy < NL; ++iy) {
        const double *p = sums + (size_t)iy * 20;
        finite &= plane_finite[iy];
        ekt += p[5];
        double variance = 0.0;
        for (int k = 0; k < 3; ++k) {
            const double mean = p[k] * inv_plane;
            const double raw = p[4 + k] * inv_plane - mean * mean;
            finite &= isfinite(raw);
            variance += raw > 0.0 ? raw : 0.0;
        }
        tke += 0.5 * wt[iy] * variance;
        for (int iz = 0; iz < NZ; ++iz)
            for (int ix = 0; ix < NX; ++ix) {
                const double d = phys[((size_t)iz * NL + iy) * NX + ix];
                finite &= isfinite(d);
                div_square += wt[iy] * inv_plane * d * d;
            }
    }
    out[0] = ekt;
    out[1] = tke;
    out[2] = sqrt(div_square);
    finite &= isfinite(out[0]) && isfinite(out[1]) && isfinite(out[2]);
    free(div);
    free(phys);
    free(wt);
    return finite;
}
