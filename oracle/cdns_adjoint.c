/* Tangent-linear and discrete adjoint of the channel-dns time step -- TEST INFRASTRUCTURE.
 *
 * The reference has no adjoint; this file differentiates the oracle's own step
 * (cdns_oracle.c, which restates src/sim/simulation.cpp:142-171) operation by operation and
 * transposes each linear piece exactly, so that <TL dx, l> = <dx, ADJ l> holds to round-off.
 * Inner products are real: complex arrays are vectors of (re, im) pairs, physical fields
 * plain real vectors.  Transposes of the pieces:
 *   fine transforms     backward^T = D_s * forward * D_p, forward^T = D_p' * backward * D_s'
 *                       (diagonal weights of the x, z and y steps; transforms.cpp:641-741)
 *   curl                (i k)^T = -i k, Chebyshev derivative^T (vector_ops.cpp:17-25, 112-151)
 *   cross product       l_a = b x l_p, l_b = l_p x a
 *   pressure wall cond. wall sums transposed (timestepper.cpp:26-43)
 *   tau solve           each sweep of solve_tau_column (eqsolve_pv.cpp:16-60) run backwards
 *   solve_modes         influence and sigma rank-2 updates transposed (eqsolve_pv.cpp:311-450)
 *   zero mode, flow     overwritten entries get zero adjoint; constant-flux dpdx transposed
 *                       (timestepper.cpp:60-68, simulation.cpp:111-140)
 * Only tests/ load this library; the product path never does.
 */
#include "cdns_oracle.h"

#include <complex.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

struct orc_adj {
    orc_ctx *c;
    orc_config m;
    int nzk, nkx, nj, zoff;
    size_t S, BND, MODE, PH;
    double *WAVX, *WAVZ, *WL, *WD, *WR;
    /* solver tables of the current setup */
    double g_tables;
    double *wn_p, *wn_v, *ad_p, *ad_v, *u_dpdx, *w_dpdz, qunit[2];
    cplx *PI[2], *VI[2], *ID[2], *PB, *VB, *SIG0;
    /* workspaces */
    cplx *w1, *w2, *w3, *w4;  /* vector S */
    double *ph[10];           /* fine physical scalars; ph[9] is a transform temp */
};
typedef struct orc_adj orc_adj;

static size_t sidx(const orc_adj *a, int izi, int iy, int kx) { return ((size_t)izi * a->nj + iy) * a->nkx + kx; }
static size_t bidx(const orc_adj *a, int izi, int wall, int kx) { return ((size_t)izi * 2 + wall) * a->nkx + kx; }
static inline cplx ik(double k, cplx v) { return CMPLX(-k * cimag(v), k * creal(v)); } /* (i k) v */
static inline cplx ikT(double k, cplx v) { return CMPLX(k * cimag(v), -k * creal(v)); } /* (-i k) v */

orc_adj *orc_adj_create(orc_ctx *c, const orc_config *m) {
    orc_adj *a = (orc_adj *)calloc(1, sizeof(orc_adj));
    a->c = c;
    a->m = *m;
    a->nzk = 2 * m->NZS + 1;
    a->nkx = m->NXS + 1;
    a->nj = m->NYS + 1;
    a->zoff = m->NZS;
    a->S = (size_t)a->nzk * a->nj * a->nkx;
    a->BND = (size_t)a->nzk * 2 * a->nkx;
    a->MODE = (size_t)a->nzk * a->nkx;
    a->PH = (size_t)m->NZ2 * (m->NY2 + 1) * m->NX2;
    a->WAVX = (double *)calloc(a->nkx, sizeof(double));
    a->WAVZ = (double *)calloc(a->nzk, sizeof(double));
    a->WL = (double *)calloc(a->nj, sizeof(double));
    a->WD = (double *)calloc(a->nj, sizeof(double));
    a->WR = (double *)calloc(a->nj, sizeof(double));
    orc_tables(c, a->WAVX, a->WAVZ, a->WL, a->WD, a->WR);
    a->wn_p = (double *)calloc(a->MODE, sizeof(double));
    a->wn_v = (double *)calloc(a->MODE, sizeof(double));
    a->ad_p = (double *)calloc(a->S, sizeof(double));
    a->ad_v = (double *)calloc(a->S, sizeof(double));
    a->u_dpdx = (double *)calloc(a->nj, sizeof(double));
    a->w_dpdz = (double *)calloc(a->nj, sizeof(double));
    for (int i = 0; i < 2; ++i) {
        a->PI[i] = (cplx *)calloc(a->S, sizeof(cplx));
        a->VI[i] = (cplx *)calloc(3 * a->S, sizeof(cplx));
        a->ID[i] = (cplx *)calloc(a->BND, sizeof(cplx));
    }
    a->PB = (cplx *)calloc(a->S, sizeof(cplx));
    a->VB = (cplx *)calloc(3 * a->S, sizeof(cplx));
    a->SIG0 = (cplx *)calloc(a->BND, sizeof(cplx));
    a->w1 = (cplx *)calloc(3 * a->S, sizeof(cplx));
    a->w2 = (cplx *)calloc(3 * a->S, sizeof(cplx));
    a->w3 = (cplx *)calloc(3 * a->S, sizeof(cplx));
    a->w4 = (cplx *)calloc(3 * a->S, sizeof(cplx));
    for (int i = 0; i < 10; ++i) a->ph[i] = (double *)calloc(a->PH, sizeof(double));
    a->g_tables = NAN;
    return a;
}

void orc_adj_destroy(orc_adj *a) {
    if (!a) return;
    free(a->WAVX); free(a->WAVZ); free(a->WL); free(a->WD); free(a->WR);
    free(a->wn_p); free(a->wn_v); free(a->ad_p); free(a->ad_v); free(a->u_dpdx); free(a->w_dpdz);
    for (int i = 0; i < 2; ++i) { free(a->PI[i]); free(a->VI[i]); free(a->ID[i]); }
    free(a->PB); free(a->VB); free(a->SIG0);
    free(a->w1); free(a->w2); free(a->w3); free(a->w4);
    for (int i = 0; i < 10; ++i) free(a->ph[i]);
    free(a);
}

/* Make the context's tables those of g (running its setup if they are not) and cache them. */
static void ensure_setup(orc_adj *a, double g) {
    if (orc_setup_g(a->c) != g) orc_setup(a->c, g);
    if (a->g_tables == g) return;
    orc_get_table(a->c, 0, a->wn_p);
    orc_get_table(a->c, 1, a->wn_v);
    orc_get_table(a->c, 2, a->ad_p);
    orc_get_table(a->c, 3, a->ad_v);
    orc_get_table(a->c, 4, (double *)a->PI[0]);
    orc_get_table(a->c, 5, (double *)a->PI[1]);
    orc_get_table(a->c, 6, (double *)a->VI[0]);
    orc_get_table(a->c, 7, (double *)a->VI[1]);
    orc_get_table(a->c, 8, (double *)a->ID[0]);
    orc_get_table(a->c, 9, (double *)a->ID[1]);
    orc_get_table(a->c, 10, (double *)a->PB);
    orc_get_table(a->c, 11, (double *)a->VB);
    orc_get_table(a->c, 12, (double *)a->SIG0);
    orc_get_table(a->c, 13, a->u_dpdx);
    orc_get_table(a->c, 14, a->w_dpdz);
    orc_get_table(a->c, 15, a->qunit);
    a->g_tables = g;
}

/* The step of iteration `it` uses the setup of g(it); setup runs on it = 1, 2 only
 * (simulation.cpp:176-177), so it >= 2 shares g(2). */
static double g_of(const orc_adj *a, int it) {
    double co[5];
    orc_step_coefficients(it < 2 ? it : 2, a->m.DT, a->m.Re, co);
    return co[4];
}

/* ------------------------------------------------------ transposed transforms -- */
static double cy(const orc_adj *a, int l) { return (l == 0 || l == a->m.NY2) ? 1.0 : 2.0; }
static double sy(const orc_adj *a, int j) { return j == 0 ? 1.0 / (2.0 * a->m.NY2) : 1.0 / a->m.NY2; }
static double bwy(const orc_adj *a, int j) { return j == a->m.NY2 ? 0.5 : 1.0; }
static double wx(const orc_adj *a, int k) { return (k == 0 || 2 * k == a->m.NX2) ? 1.0 : 2.0; }
static int collision(const orc_adj *a) { return a->m.NZ2 - a->m.NZS <= a->m.NZS; }

/* (fine backward)^T for one scalar field: phys (real) -> spec (reference layout, stride 1). */
void orc_adj_fine_backward_T(orc_adj *a, const double *phys, double *specd) {
    const orc_config *m = &a->m;
    double *t = a->ph[9];
    const size_t row = (size_t)m->NX2;
    for (int iz = 0; iz < m->NZ2; ++iz)
        for (int l = 0; l <= m->NY2; ++l) {
            const size_t o = ((size_t)iz * (m->NY2 + 1) + l) * row;
            const double s = 1.0 / cy(a, l);
            for (int ix = 0; ix < m->NX2; ++ix) t[o + ix] = phys[o + ix] * s;
        }
    orc_xyzfft2(a->c, 'F', t, specd, 1);
    cplx *spec = (cplx *)specd;
    for (int izi = 0; izi < a->nzk; ++izi)
        for (int j = 0; j < a->nj; ++j)
            for (int k = 0; k < a->nkx; ++k) {
                const double f = (double)m->NX2 * wx(a, k) * (double)m->NZ2 * bwy(a, j) / sy(a, j);
                spec[sidx(a, izi, j, k)] *= f;
                if (collision(a) && izi == 2 * m->NZS) spec[sidx(a, izi, j, k)] = 0;
            }
}

/* (fine forward)^T for one scalar field: spec -> phys. */
void orc_adj_fine_forward_T(orc_adj *a, const double *specd, double *phys) {
    const orc_config *m = &a->m;
    const cplx *spec = (const cplx *)specd;
    cplx *t = a->w4;
    for (int izi = 0; izi < a->nzk; ++izi)
        for (int j = 0; j < a->nj; ++j)
            for (int k = 0; k < a->nkx; ++k) {
                const double f = (1.0 / m->NX2) / wx(a, k) * (1.0 / m->NZ2) * sy(a, j) / bwy(a, j);
                t[sidx(a, izi, j, k)] = spec[sidx(a, izi, j, k)] * f;
            }
    if (collision(a))
        for (int j = 0; j < a->nj; ++j)
            for (int k = 0; k < a->nkx; ++k) {
                t[sidx(a, 0, j, k)] += t[sidx(a, 2 * m->NZS, j, k)];
                t[sidx(a, 2 * m->NZS, j, k)] = 0;
            }
    orc_xyzfft2(a->c, 'B', phys, (double *)t, 1);
    const size_t row = (size_t)m->NX2;
    for (int iz = 0; iz < m->NZ2; ++iz)
        for (int l = 0; l <= m->NY2; ++l) {
            const size_t o = ((size_t)iz * (m->NY2 + 1) + l) * row;
            const double s = cy(a, l);
            for (int ix = 0; ix < m->NX2; ++ix) phys[o + ix] *= s;
        }
}

/* ------------------------------------------------------- curl and its transpose -- */
/* Transpose of the Chebyshev derivative d = D f, d_j = c_j sum_{m > j, m - j odd} 2 m f_m,
 * c_0 = 1/2 (vector_ops.cpp:17-25):  (D^T g)_m = 2 m sum_{j < m, m - j odd} c_j g_j. */
static void dyT_col(int ny, const cplx *g, size_t st, cplx *o, size_t os) {
    cplx acc[2] = {0, 0};
    o[0] = 0;
    for (int m = 1; m <= ny; ++m) {
        const int j = m - 1;
        acc[j & 1] += (j == 0 ? 0.5 : 1.0) * g[(size_t)j * st];
        o[(size_t)m * os] = 2.0 * m * acc[j & 1];
    }
}

/* curl^T: lw (vector) -> lu (vector), both reference layout. */
void orc_adj_curl_T(orc_adj *a, const double *lwd, double *lud) {
    const cplx *lw = (const cplx *)lwd;
    cplx *lu = (cplx *)lud;
    const int ny = a->m.NYS;
    cplx *col = (cplx *)malloc(sizeof(cplx) * a->nj * 2);
    for (int izi = 0; izi < a->nzk; ++izi)
        for (int k = 0; k < a->nkx; ++k) {
            const double kx = a->WAVX[k], kz = a->WAVZ[izi];
            const cplx *base = lw + sidx(a, izi, 0, k) * 3;
            const size_t st = (size_t)a->nkx * 3;
            /* lu = -i kz ly - D^T lz ; lv = i kz lx - i kx lz ; lw = D^T lx + i kx ly */
            dyT_col(ny, base + 2, st, col, 1);
            dyT_col(ny, base + 0, st, col + a->nj, 1);
            for (int j = 0; j <= ny; ++j) {
                const size_t s = sidx(a, izi, j, k) * 3;
                const cplx lx = lw[s], ly = lw[s + 1], lz = lw[s + 2];
                lu[s + 0] = ikT(kz, ly) - col[j];
                lu[s + 1] = ik(kz, lx) - ik(kx, lz);
                lu[s + 2] = col[a->nj + j] + ik(kx, ly);
            }
        }
    free(col);
}

/* -------------------------------------------------- nonlinear term: TL and ADJ -- */
static void phys3(orc_adj *a, const cplx *v, double *o0, double *o1, double *o2) {
    /* orc_xyzfft2 with stride 3 reads component `offset` through the pointer offset */
    orc_xyzfft2(a->c, 'B', o0, (double *)(v + 0), 3);
    orc_xyzfft2(a->c, 'B', o1, (double *)(v + 1), 3);
    orc_xyzfft2(a->c, 'B', o2, (double *)(v + 2), 3);
}

/* Pressure wall condition B from curl w and nonlinear term N (timestepper.cpp:26-43):
 * B0 = sum_j [i kx w_z/Re - i kz w_x/Re + N_y]_j, B1 = sum_j (-1)^(j+1) [...]_j. */
static void wall_condition(orc_adj *a, const cplx *w, const cplx *n, cplx *B) {
    const double Re = a->m.Re;
    for (int izi = 0; izi < a->nzk; ++izi)
        for (int k = 0; k < a->nkx; ++k) {
            cplx b0 = 0, b1 = 0;
            for (int j = 0; j < a->nj; ++j) {
                const size_t s = sidx(a, izi, j, k) * 3;
                const cplx v = ik(a->WAVX[k] / Re, w[s + 2]) - ik(a->WAVZ[izi] / Re, w[s + 0]) + n[s + 1];
                b0 += v;
                b1 += (j % 2 == 0) ? -v : v;
            }
            B[bidx(a, izi, 0, k)] = b0;
            B[bidx(a, izi, 1, k)] = b1;
        }
}
static void wall_condition_T(orc_adj *a, const cplx *lB, cplx *lw, cplx *ln) {
    const double Re = a->m.Re;
    for (int izi = 0; izi < a->nzk; ++izi)
        for (int k = 0; k < a->nkx; ++k) {
            const cplx l0 = lB[bidx(a, izi, 0, k)], l1 = lB[bidx(a, izi, 1, k)];
            for (int j = 0; j < a->nj; ++j) {
                const size_t s = sidx(a, izi, j, k) * 3;
                const cplx lv = l0 + ((j % 2 == 0) ? -l1 : l1);
                ln[s + 1] += lv;
                lw[s + 2] += ikT(a->WAVX[k] / Re, lv);
                lw[s + 0] -= ikT(a->WAVZ[izi] / Re, lv);
            }
        }
}

/* dN = F( B(du) x W + U x B(curl du) ) for base velocity u. */
static void nl_tl(orc_adj *a, const cplx *u, const cplx *du, cplx *dN) {
    const size_t PH = a->PH;
    double **p = a->ph;
    cplx *w = a->w1, *dw = a->w2;
    orc_curl(a->c, (double *)w, (const double *)u);
    orc_curl(a->c, (double *)dw, (const double *)du);
    phys3(a, u, p[0], p[1], p[2]);   /* U */
    phys3(a, w, p[3], p[4], p[5]);   /* W */
    /* dU in p[6..7..8] one at a time is not enough: form the product component-wise */
    double *dU[3], *dW[3];
    for (int i = 0; i < 3; ++i) {
        dU[i] = (double *)malloc(PH * sizeof(double));
        dW[i] = (double *)malloc(PH * sizeof(double));
    }
    phys3(a, du, dU[0], dU[1], dU[2]);
    phys3(a, dw, dW[0], dW[1], dW[2]);
    for (int c = 0; c < 3; ++c) {
        const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
        double *o = p[6];
        for (size_t i = 0; i < PH; ++i)
            o[i] = (dU[c1][i] * p[3 + c2][i] - dU[c2][i] * p[3 + c1][i]) +
                   (p[c1][i] * dW[c2][i] - p[c2][i] * dW[c1][i]);
        orc_xyzfft2(a->c, 'F', o, (double *)(dN + c), 3);
    }
    for (int i = 0; i < 3; ++i) { free(dU[i]); free(dW[i]); }
}

/* (lu) += [dN/du]^T lN for base velocity u. */
static void nl_adj(orc_adj *a, const cplx *u, const cplx *lN, cplx *lu) {
    const size_t PH = a->PH;
    double **p = a->ph;
    cplx *w = a->w1, *spec = a->w2, *lw = a->w3;
    orc_curl(a->c, (double *)w, (const double *)u);
    phys3(a, u, p[0], p[1], p[2]); /* U */
    phys3(a, w, p[3], p[4], p[5]); /* W */
    /* L = F^T lN, per component */
    cplx *one = (cplx *)malloc(a->S * sizeof(cplx));
    for (int c = 0; c < 3; ++c) {
        for (size_t s = 0; s < a->S; ++s) one[s] = lN[s * 3 + c];
        orc_adj_fine_forward_T(a, (const double *)one, p[6 + c]);
    }
    /* lU = W x L, lW = L x U; each transformed back with B^T */
    double *t = (double *)malloc(PH * sizeof(double));
    memset(lw, 0, 3 * a->S * sizeof(cplx));
    for (int c = 0; c < 3; ++c) {
        const int c1 = (c + 1) % 3, c2 = (c + 2) % 3;
        for (size_t i = 0; i < PH; ++i) t[i] = p[3 + c1][i] * p[6 + c2][i] - p[3 + c2][i] * p[6 + c1][i];
        orc_adj_fine_backward_T(a, t, (double *)one);
        for (size_t s = 0; s < a->S; ++s) lu[s * 3 + c] += one[s];
        for (size_t i = 0; i < PH; ++i) t[i] = p[6 + c1][i] * p[c2][i] - p[6 + c2][i] * p[c1][i];
        orc_adj_fine_backward_T(a, t, (double *)one);
        for (size_t s = 0; s < a->S; ++s) lw[s * 3 + c] = one[s];
    }
    orc_adj_curl_T(a, (const double *)lw, (double *)spec);
    for (size_t s = 0; s < 3 * a->S; ++s) lu[s] += spec[s];
    free(one);
    free(t);
}

/* --------------------------------------------------------- transposed tau solve -- */
/* Transpose of solve_tau_col for one column: given l (adjoint of out[0..ny]) write
 * lr (adjoint of rhs[0..ny]) and the adjoints of the two wall values. */
static void tau_col_T(const orc_adj *a, int ny, double g, double ca, double cb, int zero_mode, double w,
                      const double *ad, cplx *l, cplx *lr, cplx *llower, cplx *lupper) {
    const double *WL = a->WL, *WD = a->WD, *WR = a->WR;
    const int normalize = g == 0.0 && ca == 0.0 && zero_mode;
    /* closure coefficients of the forward (data independent) */
    double kappa[4096];
    double ad01[2] = {0, 0}, ad01_final[2];
    for (int iy = ny; iy >= ny - 1; --iy) {
        const int q = iy % 2;
        ad01[q] = (ca == 0.0 && cb == 0.0) ? (double)((1 - q) * iy * iy + q) : ca + cb * iy * iy;
    }
    for (int iy = ny - 2; iy >= 0; --iy) {
        const int q = iy % 2;
        const double func = (ca == 0.0 && cb == 0.0) ? (double)((1 - q) * iy * iy + q) : ca + cb * iy * iy;
        kappa[iy] = ad01[q] * ad[iy + 2];
        ad01[q] = func - WL[iy + 2] * w * ad01[q] * ad[iy + 2];
    }
    ad01_final[0] = normalize ? 1.0 : ad01[0];
    ad01_final[1] = ad01[1];
    /* 6^T forward sweep out[iy] = (out[iy] - out[iy-2] WL w) ad[iy], iy = 2..ny */
    for (int iy = ny; iy >= 2; --iy) {
        const cplx t = l[iy] * ad[iy];
        l[iy - 2] -= WL[iy] * w * t;
        l[iy] = t;
    }
    /* 5^T */
    l[0] /= ad01_final[0];
    l[1] /= ad01_final[1];
    if (normalize) l[0] = 0;
    /* 4^T closure out[q] -= kappa[m-2] out[m] */
    for (int m = 2; m <= ny; ++m) l[m] -= kappa[m - 2] * l[m % 2];
    /* 3^T backward sweep out[iy] -= out[iy+2] WR[iy] w ad[iy+2], iy = ny-2..2 */
    for (int iy = 2; iy <= ny - 2; ++iy) l[iy + 2] -= WR[iy] * w * ad[iy + 2] * l[iy];
    /* 2^T rows iy = 2..ny: out[iy] = WL[iy] R(iy-2) + WD[iy] R(iy) + WR[iy] R(iy+2) (truncated) */
    cplx *lR = lr;
    for (int y = 0; y <= ny; ++y) lR[y] = 0;
    for (int iy = 2; iy <= ny; ++iy) {
        lR[iy - 2] += WL[iy] * l[iy];
        lR[iy] += WD[iy] * l[iy];
        if (iy <= ny - 2) lR[iy + 2] += WR[iy] * l[iy];
    }
    cplx lcorr = 0;
    if (normalize) {
        lcorr = lR[0];
        lR[0] = 0;
        for (int y = 2; y <= ny; y += 2) lR[y] += lcorr / (double)(y * y - 1);
    }
    /* 1^T out[0] = (lo + up)/2, out[1] = (lo - up)/2 */
    *llower = 0.5 * (l[0] + l[1]);
    *lupper = 0.5 * (l[0] - l[1]);
    if (normalize && cb != 0.0) {
        *llower += lcorr / (2.0 * cb);
        *lupper += lcorr / (2.0 * cb);
    }
}

/* Transpose of the fused solve_modes (correct_sigma = 1) restricted to the velocity output:
 * lV (vector) -> lr (vector RHS adjoint), lbc (pressure wall values adjoint). */
void orc_adj_solve_T(orc_adj *a, const double *lVd, double *lrd, double *lbcd) {
    const orc_config *m = &a->m;
    ensure_setup(a, orc_setup_g(a->c));
    const cplx *lV = (const cplx *)lVd;
    cplx *lr = (cplx *)lrd, *lbc = (cplx *)lbcd;
    const int ny = m->NYS, nj = a->nj;
    const double Re = m->Re;
    cplx *buf = (cplx *)calloc((size_t)nj * 12, sizeof(cplx));
    cplx *lvc[3] = {buf, buf + nj, buf + 2 * nj};
    cplx *lf = buf + 3 * nj, *lpc = buf + 4 * nj, *ldy = buf + 5 * nj, *lrp = buf + 6 * nj, *tmp = buf + 7 * nj;
    cplx *lrk[3] = {buf + 8 * nj, buf + 9 * nj, buf + 10 * nj};
    double *adp = (double *)malloc(sizeof(double) * nj), *adv = (double *)malloc(sizeof(double) * nj);
    for (int izi = 0; izi < a->nzk; ++izi)
        for (int x = 0; x < a->nkx; ++x) {
            const int zero = izi == a->zoff && x == 0;
            const double kx = a->WAVX[x], kz = a->WAVZ[izi];
            const double wv = a->wn_v[izi * a->nkx + x], wp = a->wn_p[izi * a->nkx + x];
            for (int y = 0; y < nj; ++y) {
                adp[y] = a->ad_p[sidx(a, izi, y, x)];
                adv[y] = a->ad_v[sidx(a, izi, y, x)];
            }
            /* (i) outputs V_k,y = vc + c0 VI0 + c1 VI1 + sigma[.] VB */
            cplx lc0 = 0, lc1 = 0, lsig[2] = {0, 0};
            for (int k = 0; k < 3; ++k)
                for (int y = 0; y <= ny; ++y) {
                    const size_t s = sidx(a, izi, y, x);
                    const cplx l = lV[s * 3 + k];
                    lvc[k][y] = l;
                    lc0 += conj(a->VI[0][s * 3 + k]) * l;
                    lc1 += conj(a->VI[1][s * 3 + k]) * l;
                    const int q = (y + 1) % 2;
                    lsig[k == 1 ? 1 - q : q] += conj(a->VB[s * 3 + k]) * l;
                }
            for (int y = 0; y < nj; ++y) lpc[y] = 0, lrk[0][y] = 0, lrk[1][y] = 0, lrk[2][y] = 0;
            /* (h) sigma */
            {
                const cplx inv0 = 1.0 / a->SIG0[bidx(a, izi, 0, x)], inv1 = 1.0 / a->SIG0[bidx(a, izi, 1, x)];
                const cplx ls0 = conj(inv0) * lsig[0], ls1 = conj(inv1) * lsig[1];
                const cplx lVTn = -wv * ls0, lVTn1 = -wv * ls1, lptop = 2.0 * ny * Re * ls1;
                lrk[1][ny] += -Re * ls0;
                lrk[1][ny - 1] += -Re * ls1;
                const size_t sn = sidx(a, izi, ny, x), sn1 = sidx(a, izi, ny - 1, x);
                lvc[1][ny] += lVTn;
                lc0 += conj(a->VI[0][sn * 3 + 1]) * lVTn;
                lc1 += conj(a->VI[1][sn * 3 + 1]) * lVTn;
                lvc[1][ny - 1] += lVTn1;
                lc0 += conj(a->VI[0][sn1 * 3 + 1]) * lVTn1;
                lc1 += conj(a->VI[1][sn1 * 3 + 1]) * lVTn1;
                lpc[ny] += lptop;
                lc0 += conj(a->PI[0][sn]) * lptop;
                lc1 += conj(a->PI[1][sn]) * lptop;
            }
            /* (g) c0 = div1 ID1[0] - div0 ID1[1]; c1 = div0 ID0[1] - div1 ID0[0] */
            const cplx A = a->ID[1][bidx(a, izi, 0, x)], Bc = a->ID[1][bidx(a, izi, 1, x)];
            const cplx Cc = a->ID[0][bidx(a, izi, 1, x)], Dc = a->ID[0][bidx(a, izi, 0, x)];
            const cplx ldiv1 = conj(A) * lc0 - conj(Dc) * lc1;
            const cplx ldiv0 = -conj(Bc) * lc0 + conj(Cc) * lc1;
            /* (f) wall divergence */
            for (int y = 0; y <= ny; ++y) {
                const double sgn = (y % 2 == 0) ? 1.0 : -1.0;
                const cplx lt = ldiv0 + sgn * ldiv1, ln = ldiv0 - sgn * ldiv1;
                lvc[0][y] += ikT(kx, lt);
                lvc[2][y] += ikT(kz, lt);
                lvc[1][y] += (double)(y * y) * ln;
            }
            /* (e)^T (d)^T velocity tau solves and forcing = Re (grad - r) */
            for (int y = 0; y < nj; ++y) ldy[y] = 0;
            for (int k = 0; k < 3; ++k) {
                cplx llo, lup;
                tau_col_T(a, ny, a->g_tables, m->a_V[k], m->b_V[k], zero, wv, adv, lvc[k], lf, &llo, &lup);
                for (int y = 0; y <= ny; ++y) {
                    const cplx lg = Re * lf[y];
                    lrk[k][y] -= Re * lf[y];
                    if (k == 0) lpc[y] += ikT(kx, lg);
                    else if (k == 1) ldy[y] += lg;
                    else lpc[y] += ikT(kz, lg);
                }
            }
            /* (c)^T dy = D pc */
            dyT_col(ny, ldy, 1, tmp, 1);
            for (int y = 0; y <= ny; ++y) lpc[y] += tmp[y];
            /* (b)^T pressure tau solve */
            cplx lplo, lpup;
            tau_col_T(a, ny, m->G_pre, m->a_pre0, m->b_pre0, zero, wp, adp, lpc, lrp, &lplo, &lpup);
            lbc[bidx(a, izi, 0, x)] = lplo;
            lbc[bidx(a, izi, 1, x)] = lpup;
            /* (a)^T rp = i kx r0 + D r1 + i kz r2 */
            dyT_col(ny, lrp, 1, tmp, 1);
            for (int y = 0; y <= ny; ++y) {
                lrk[0][y] += ikT(kx, lrp[y]);
                lrk[1][y] += tmp[y];
                lrk[2][y] += ikT(kz, lrp[y]);
            }
            for (int k = 0; k < 3; ++k)
                for (int y = 0; y <= ny; ++y) lr[sidx(a, izi, y, x) * 3 + k] = lrk[k][y];
        }
    free(buf);
    free(adp);
    free(adv);
}

/* ------------------------------------------------------------------ one step -- */
/* Step state X = (u, u_prev, N_prev, B_prev) as four arrays in reference layouts
 * (vector S, vector S, vector S, boundary BND). */

/* Tangent-linear step of iteration `it` at base state X (body force enters additively, so
 * df is its direction).  Writes dXo (may not alias dX). */
void orc_adj_tl_step(orc_adj *a, int it, const double *const Xd[4], const double *const dXd[4],
                     const double *dfd, double *const dXod[4]) {
    const orc_config *m = &a->m;
    const size_t S3 = 3 * a->S;
    const cplx *u = (const cplx *)Xd[0];
    const cplx *du = (const cplx *)dXd[0], *dup = (const cplx *)dXd[1], *dnp = (const cplx *)dXd[2],
               *dbp = (const cplx *)dXd[3];
    const cplx *df = (const cplx *)dfd;
    cplx *duo = (cplx *)dXod[0], *dupo = (cplx *)dXod[1], *dno = (cplx *)dXod[2], *dbo = (cplx *)dXod[3];
    double co[5];
    orc_step_coefficients(it, m->DT, m->Re, co);
    ensure_setup(a, g_of(a, it));
    cplx *dN = (cplx *)calloc(S3, sizeof(cplx)), *dw = (cplx *)calloc(S3, sizeof(cplx));
    cplx *dB = (cplx *)calloc(a->BND, sizeof(cplx)), *dbext = (cplx *)calloc(a->BND, sizeof(cplx));
    cplx *dr = (cplx *)calloc(S3, sizeof(cplx)), *dP = (cplx *)calloc(a->S, sizeof(cplx));
    cplx *zbc = (cplx *)calloc(3 * a->BND, sizeof(cplx));
    nl_tl(a, u, du, dN);
    if (df)
        for (size_t s = 0; s < S3; ++s) dN[s] += df[s];
    orc_curl(a->c, (double *)dw, (const double *)du);
    wall_condition(a, dw, dN, dB);
    for (size_t s = 0; s < a->BND; ++s) dbext[s] = co[2] * dB[s] + co[3] * dbp[s];
    for (size_t s = 0; s < S3; ++s) dr[s] = co[0] * du[s] + co[1] * dup[s] + co[2] * dN[s] + co[3] * dnp[s];
    /* the solve with the context's tables; zero velocity wall values (they are constants) */
    orc_solve_modes(a->c, (double *)dP, NULL, (const double *)dbext, (double *)duo, (const double *)dr,
                    (const double *)zbc, 1);
    for (int y = 0; y < a->nj; ++y) duo[sidx(a, a->zoff, y, 0) * 3 + 1] = 0;
    if (m->INDEX_GRAD_PRE != 1) {
        const double qx = orc_flow_rate(a->c, (const double *)duo, 3, 0);
        const double qz = orc_flow_rate(a->c, (const double *)duo, 3, 2);
        const double dpx = -qx / a->qunit[0], dpz = -qz / a->qunit[1];
        for (int y = 0; y < a->nj; ++y) {
            duo[sidx(a, a->zoff, y, 0) * 3 + 0] += dpx * a->u_dpdx[y];
            duo[sidx(a, a->zoff, y, 0) * 3 + 2] += dpz * a->w_dpdz[y];
        }
    }
    memcpy(dupo, du, S3 * sizeof(cplx));
    memcpy(dno, dN, S3 * sizeof(cplx));
    memcpy(dbo, dB, a->BND * sizeof(cplx));
    free(dN); free(dw); free(dB); free(dbext); free(dr); free(dP); free(zbc);
}

/* Adjoint step: L = (lu, lup, lnp, lbp) of the state after iteration `it` in, the adjoint of
 * the state before it out (in place); lf (optional) accumulates the body-force gradient. */
void orc_adj_step(orc_adj *a, int it, const double *const Xd[4], double *const Ld[4], double *lfd) {
    const orc_config *m = &a->m;
    const size_t S3 = 3 * a->S;
    const cplx *u = (const cplx *)Xd[0];
    cplx *lu = (cplx *)Ld[0], *lup = (cplx *)Ld[1], *lnp = (cplx *)Ld[2], *lbp = (cplx *)Ld[3];
    cplx *lf = (cplx *)lfd;
    double co[5];
    orc_step_coefficients(it, m->DT, m->Re, co);
    ensure_setup(a, g_of(a, it));
    cplx *ls = (cplx *)calloc(S3, sizeof(cplx)), *lr = (cplx *)calloc(S3, sizeof(cplx));
    cplx *lbext = (cplx *)calloc(a->BND, sizeof(cplx)), *lN = (cplx *)calloc(S3, sizeof(cplx));
    cplx *lB = (cplx *)calloc(a->BND, sizeof(cplx)), *lw = (cplx *)calloc(S3, sizeof(cplx));
    cplx *lnew = (cplx *)calloc(S3, sizeof(cplx));
    /* flow correction^T and zero mode^T on the adjoint of the new velocity */
    memcpy(ls, lu, S3 * sizeof(cplx));
    if (m->INDEX_GRAD_PRE != 1) {
        double gx = 0, gz = 0;
        for (int y = 0; y < a->nj; ++y) {
            gx += creal(lu[sidx(a, a->zoff, y, 0) * 3 + 0]) * a->u_dpdx[y];
            gz += creal(lu[sidx(a, a->zoff, y, 0) * 3 + 2]) * a->w_dpdz[y];
        }
        /* dpdx = -Q(u*_x)/qx_unit with Q(f) = -sum_{j even} Re f_j / (j^2 - 1) */
        for (int y = 0; y < a->nj; y += 2) {
            const double dq = -1.0 / (double)(y * y - 1);
            ls[sidx(a, a->zoff, y, 0) * 3 + 0] += -(dq / a->qunit[0]) * gx;
            ls[sidx(a, a->zoff, y, 0) * 3 + 2] += -(dq / a->qunit[1]) * gz;
        }
    }
    for (int y = 0; y < a->nj; ++y) ls[sidx(a, a->zoff, y, 0) * 3 + 1] = 0;
    orc_adj_solve_T(a, (const double *)ls, (double *)lr, (double *)lbext);
    /* r = a0 u + a1 up + b0 N + b1 np ; bext = b0 B + b1 bp */
    for (size_t s = 0; s < S3; ++s) {
        lnew[s] = co[0] * lr[s] + lup[s]; /* the new state's u_prev is u */
        lN[s] = co[2] * lr[s] + lnp[s];   /* the new state's N_prev is N */
    }
    for (size_t s = 0; s < a->BND; ++s) lB[s] = co[2] * lbext[s] + lbp[s];
    for (size_t s = 0; s < S3; ++s) {
        lup[s] = co[1] * lr[s];
        lnp[s] = co[3] * lr[s];
    }
    for (size_t s = 0; s < a->BND; ++s) lbp[s] = co[3] * lbext[s];
    /* B = wall(w, N) */
    wall_condition_T(a, lB, lw, lN);
    if (lf)
        for (size_t s = 0; s < S3; ++s) lf[s] += lN[s];
    /* N = NL(u) (+ f): lu += NL'^T lN ; w = curl u from the wall condition: lu += curl^T lw */
    nl_adj(a, u, lN, lnew);
    orc_adj_curl_T(a, (const double *)lw, (double *)ls);
    for (size_t s = 0; s < S3; ++s) lu[s] = lnew[s] + ls[s];
    free(ls); free(lr); free(lbext); free(lN); free(lB); free(lw); free(lnew);
}

/* ----------------------------------------------------------------- objectives -- */
/* 0: volume kinetic energy 1/2 int |u|^2 dV; 1: mean wall shear
 * 1/2 (du/dy(-1) - du/dy(+1)) of the plane mean = -sum_{j even} j^2 Re u_x(0,0,j). */
double orc_adj_objective(orc_adj *a, int which, const double *veld, double *gradd) {
    const orc_config *m = &a->m;
    const cplx *u = (const cplx *)veld;
    cplx *g = (cplx *)gradd;
    const int nj = a->nj;
    if (g) memset(g, 0, 3 * a->S * sizeof(cplx));
    if (which == 1) {
        double J = 0;
        for (int j = 0; j < nj; j += 2) {
            J -= (double)(j * j) * creal(u[sidx(a, a->zoff, j, 0) * 3 + 0]);
            if (g) g[sidx(a, a->zoff, j, 0) * 3 + 0] = -(double)(j * j);
        }
        return J;
    }
    /* I_mn = int_{-1}^{1} T_m T_n dy = 1/(1-(m+n)^2) + 1/(1-(m-n)^2), m+n even */
    double *Imn = (double *)malloc(sizeof(double) * nj * nj);
    for (int p = 0; p < nj; ++p)
        for (int q = 0; q < nj; ++q)
            Imn[p * nj + q] = ((p + q) % 2) ? 0.0
                                          : 1.0 / (1.0 - (double)(p + q) * (p + q)) + 1.0 / (1.0 - (double)(p - q) * (p - q));
    double J = 0;
    for (int izi = 0; izi < a->nzk; ++izi)
        for (int k = 0; k < a->nkx; ++k) {
            const double ck = (k == 0 ? 1.0 : 2.0) * m->XL * m->ZL;
            for (int c = 0; c < 3; ++c)
                for (int p = 0; p < nj; ++p) {
                    cplx acc = 0;
                    for (int q = p % 2; q < nj; q += 2) acc += Imn[p * nj + q] * u[sidx(a, izi, q, k) * 3 + c];
                    const cplx up = u[sidx(a, izi, p, k) * 3 + c];
                    J += 0.5 * ck * (creal(up) * creal(acc) + cimag(up) * cimag(acc));
                    if (g) g[sidx(a, izi, p, k) * 3 + c] = ck * acc;
                }
        }
    free(Imn);
    return J;
}

/* ------------------------------------------------------------------- horizon -- */
/* orc_step runs setup itself on it = 1, 2; later steps use whatever tables the context
 * holds, which the adjoint of step 1 may have switched to g(1). */
static void forward_step(orc_adj *a, int it) {
    if (it >= 3 && orc_setup_g(a->c) != g_of(a, it)) orc_setup(a->c, g_of(a, it));
    orc_step(a->c, it);
}

/* Gradient of J(u(T)) after nsteps steps (it = 1..nsteps) from u0 with body force f (NULL:
 * none), with checkpoints every kc steps and recomputation.  Writes grad_u0 and grad_f (if
 * non-NULL) and returns J.  The context's state is overwritten. */
double orc_adj_run(orc_adj *a, const double *u0, const double *f, int nsteps, int kc, int objective,
                   double *grad_u0, double *grad_f) {
    orc_ctx *c = a->c;
    const size_t S3 = 3 * a->S, BND = a->BND;
    const int nck = (nsteps + kc - 1) / kc;
    /* checkpoint = (u, u_prev, N_prev, B_prev) */
    double **ck = (double **)calloc((size_t)nck * 4, sizeof(double *));
    orc_set_body_force(c, f);
    double *zero = (double *)calloc(2 * S3, sizeof(double));
    orc_set_state(c, 0, u0);
    orc_set_state(c, 1, zero);
    orc_set_state(c, 3, zero);
    orc_set_state(c, 6, zero);
    const size_t sz[4] = {S3, S3, S3, BND};
    const int fid[4] = {0, 1, 3, 6};
    for (int it = 1; it <= nsteps; ++it) {
        if ((it - 1) % kc == 0) {
            const int k = (it - 1) / kc;
            for (int i = 0; i < 4; ++i) {
                ck[k * 4 + i] = (double *)malloc(sz[i] * 2 * sizeof(double));
                orc_get_state(c, fid[i], ck[k * 4 + i]);
            }
        }
        forward_step(a, it);
    }
    double *uT = (double *)malloc(S3 * 2 * sizeof(double));
    orc_get_state(c, 0, uT);
    double *L[4];
    for (int i = 0; i < 4; ++i) L[i] = (double *)calloc(sz[i] * 2, sizeof(double));
    const double J = orc_adj_objective(a, objective, uT, L[0]);
    if (grad_f) memset(grad_f, 0, S3 * 2 * sizeof(double));
    /* states of one segment */
    double **seg = (double **)calloc((size_t)kc * 4, sizeof(double *));
    for (int i = 0; i < kc * 4; ++i) seg[i] = (double *)malloc(sz[i % 4] * 2 * sizeof(double));
    for (int k = nck - 1; k >= 0; --k) {
        const int it0 = k * kc + 1, it1 = (k + 1) * kc < nsteps ? (k + 1) * kc : nsteps;
        for (int i = 0; i < 4; ++i) orc_set_state(c, fid[i], ck[k * 4 + i]);
        /* the forward steps of this segment need the tables of their own g */
        for (int it = it0; it <= it1; ++it) {
            for (int i = 0; i < 4; ++i) orc_get_state(c, fid[i], seg[(it - it0) * 4 + i]);
            forward_step(a, it);
        }
        for (int it = it1; it >= it0; --it) {
            const double *X[4] = {seg[(it - it0) * 4 + 0], seg[(it - it0) * 4 + 1], seg[(it - it0) * 4 + 2],
                                  seg[(it - it0) * 4 + 3]};
            orc_adj_step(a, it, X, L, grad_f);
        }
    }
    memcpy(grad_u0, L[0], S3 * 2 * sizeof(double));
    for (int i = 0; i < kc * 4; ++i) free(seg[i]);
    free(seg);
    for (int i = 0; i < nck * 4; ++i) free(ck[i]);
    free(ck);
    for (int i = 0; i < 4; ++i) free(L[i]);
    free(uT);
    free(zero);
    orc_set_body_force(c, NULL);
    return J;
}
