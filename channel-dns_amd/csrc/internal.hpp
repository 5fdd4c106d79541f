// Kernel argument blocks and launcher declarations shared by the backend's sources.
#pragma once

#include "common.hpp"

namespace cdns {

// y stages (dense Chebyshev apply).  coef: backward [j][NLH]; forward even/odd [NLH][nje|njo].
struct YArgs {
    const zd *spec[6]; // ybwd inputs: spectral component planes
    zd *spec_out[6];   // yfwd outputs
    const zd *add[6];  // yfwd: planes added to the outputs (a body force), or null
    void *y;           // Y fields, field f at y + f*YS (element type by precision)
    const double *coef, *coef2;
    int nj, nje, njo, JS, NQ, NLH, NY2, nwj, ng;
    size_t YS;
    int mfma;                   // use the matrix-core kernels when the tables fit
    int has_mid, rows_m, njp, KL; // matrix-core shape (set by their launchers)
    int parts, ncg;               // matrix-core grid: parts per column group, column groups
    // backward with the curl fused in (launch_ybwd_curl): spec[0..2] = u, v, w; outputs the six
    // fields u, v, w, wx, wy, wz and (if bcw) the curl part of the pressure wall condition
    const double *coefd;          // T'_j(y_l) [j][NLH]
    const double *wavx, *wavz;
    double Re;
    int nzk;
    zd *bcw;
    int out32; // forward: spec_out holds single-precision complex planes (the mixed step's N)
};

// y stages as a DCT-I through complex FFTs of length L = 2 NY2 (transforms_dct.hip).
struct DctArgs {
    const zd *spec[6]; // backward inputs (curl: u, v, w)
    zd *spec_out[6];   // forward outputs
    const zd *add[6];  // forward: planes added to the outputs, or null
    void *y;           // Y fields, field f at y + f*YS
    size_t YS;
    const void *tw;    // L roots exp(-2 pi i k / L), element type by precision
    const double *wavx, *wavz;
    double Re;
    zd *bcw;           // curl: the curl part of the pressure wall condition [q][2], or null
    int ny, JS, NQ, nzk, NY2, L, nf, curl, CH;
    int out32;         // forward: spec_out holds single-precision complex planes
};
bool dct_supported(int L);
bool launch_ybwd_dct(const DctArgs &a, bool fp32, hipStream_t st);
bool launch_yfwd_dct(const DctArgs &a, bool fp32, hipStream_t st);

// z stages (FFT over kz).
struct ZArgs {
    const void *in;
    void *out;
    size_t in_fs, out_fs; // field strides, complex elements
    const void *tw;
    FftPlan plan;
    int nkx, nzk, NZS, NQ, KT, NL;
    int rf; // use the register-resident FFT when available
    int nkxg; // register z forward: kx groups per (field, l)
    int pair; // FFT rows on lane pairs (even factors)
    int coarse; // coarse-grid transform: at NZS = N/2 the kz = +-NZS pair shares slot NZS
};

// Layout of the z-physical fields produced by the x forward stages and read by the z forward:
// line-major [l][kx][iz] (each z line contiguous), point p = l NZ2 + iz.  The z backward writes
// the row-major [l][iz][kx] that the x backward stages read.
__host__ __device__ inline size_t zline_index(int p, int kx, int nkx, int NZ2) {
    const int l = p / NZ2, iz = p - l * NZ2;
    return ((size_t)l * nkx + kx) * NZ2 + iz;
}

// x stages (FFT over kx) and the cross product.
struct XArgs {
    const void *in;
    void *out;
    size_t in_fs, out_fs;
    const void *tw;
    FftPlan plan;
    int nkx, NL, NZ2, P;
    int rf;
    int pair; // product: rows of the register FFT on lane pairs (even factors)
    int coarse; // coarse-grid transform: at NXS = N/2 the Nyquist coefficient counts twice
    double *phys; // reference-layout physical field for the single-field stages
    // product a x b: input fields a0 a1 a2 b0 b1 b2 and the three output fields
    const void *fin[6];
    void *fout[3];
};

void launch_ybwd(const YArgs &a, int nfields, bool fp32, hipStream_t st);
// matrix-core variants; return false when a table does not fit the LDS budget
bool launch_ybwd_mfma(YArgs a, int nf, bool fp32, hipStream_t st);
bool launch_yfwd_mfma(YArgs a, int nf, bool fp32, hipStream_t st);
bool launch_ybwd_curl(YArgs a, bool fp32, hipStream_t st);
void launch_yfwd(const YArgs &a, int nfields, bool fp32, hipStream_t st);
void launch_zbwd(const ZArgs &a, int nfields, bool fp32, hipStream_t st);
void launch_zfwd(const ZArgs &a, int nfields, bool fp32, hipStream_t st);
void launch_xprod(const XArgs &a, bool fp32, hipStream_t st);
void launch_xbwd_phys(const XArgs &a, bool fp32, hipStream_t st);
void launch_xfwd_phys(const XArgs &a, bool fp32, hipStream_t st);
// register-resident two-pass FFT variants; return false for lengths without an instantiation
bool rf_supported(int n);
bool launch_xprod_rf(const XArgs &a, bool fp32, hipStream_t st);
bool launch_xbwd_phys_rf(const XArgs &a, bool fp32, hipStream_t st);
bool launch_xfwd_phys_rf(const XArgs &a, bool fp32, hipStream_t st);
bool launch_zbwd_rf(const ZArgs &a, int nf, bool fp32, hipStream_t st);
bool launch_zfwd_rf(const ZArgs &a, int nf, bool fp32, hipStream_t st);

// ----------------------------------------------------------- spectral ----
struct LayoutArgs {
    // reference layout element (izi, j, kx, c) at ref[((izi*nrow + j)*nkx + kx)*stride + c]
    const zd *ref_in;
    zd *ref_out;
    int stride, ncomp, nrow; // nrow: NYS+1 for fields, 2 for boundaries
    const zd *plane_in[3];
    zd *plane_out[3];
    int nkx, nzk, JS; // internal column stride (JS for fields, 2 for boundaries)
};
void launch_ref_to_planes(const LayoutArgs &a, hipStream_t st);
void launch_planes_to_ref(const LayoutArgs &a, hipStream_t st);
void launch_real_planes_to_ref(const double *plane, double *ref, int nrow, int nkx, int nzk, int JS, hipStream_t st);
// n complex entries: single-precision plane -> FP64 plane, and back
void launch_widen(const void *src, zd *dst, size_t n, hipStream_t st);
void launch_narrow(const zd *src, void *dst, size_t n, hipStream_t st);

struct CurlArgs {
    const zd *u[3];
    zd *w[3];
    zd *bcw; // [q][2] curl part of the pressure wall condition, or null
    const double *WAVX, *WAVZ;
    double Re;
    int ny, JS, NQ, nzk;
};
void launch_curl(const CurlArgs &a, int E, hipStream_t st);
// pressure wall condition = curl part + sum_j (+-1)^j N_y,j (timestepper.cpp:26-43)
void launch_wall_sum(const zd *nly, const zd *bcw, zd *pbc, int ny, int JS, int NQ, int E, hipStream_t st);

// ------------------------------------------------------------- solver ----
enum SolveMode { MODE_STEP = 0, MODE_GIVEN = 1, MODE_INFLUENCE = 2, MODE_BPROBLEM = 3 };

struct SolveArgs {
    int mode;
    int ny, JS, NQ, nzk, nkx;
    const double *WAVX, *WAVZ, *WL, *WD, *WR;
    double Re, G_pre, g_v, pa, pb; // pressure BC a/b currently in force
    double aV[3], bV[3];
    // rhs sources
    const zd *u[3], *up[3], *nl[3], *nlp[3]; // step: a0 u + a1 up + b0 nl + b1 nlp
    int nl32;                                // step: nl, nlp hold single-precision complex planes
    double a0, a1, b0, b1;
    const zd *rv[3]; // given velocity rhs planes
    const zd *rp;    // given pressure rhs plane (null: fused divergence)
    const zd *pbc;   // given pressure BC [q][2]
    const zd *vbc;   // velocity BC [c][q][2] (3 planes of NQ*2)
    const zd *bcw;   // step: curl part of the pressure wall condition [q][2]
    zd *bc_prev;     // step: in B^{n-1}, out B^n [q][2]
    zd *bc_ext;      // step: out b0 B^n + b1 B^{n-1} [q][2]
    int influence_index;
    int correct, sigma;
    // tables
    const double *ad_p, *ad_v;
    const zd *PI[2], *VI[2][3], *ID[2], *PB, *VB[3], *SIG0;
    // compact form of the influence and B responses (used when u and w share their wall
    // conditions): P and v are real, u = i kx phi, w = i kz phi with phi real
    int compact;
    const double *PIr[2], *VIv[2], *VIf[2], *PBr, *VBv, *VBf;
    const double *u_unit, *w_unit;
    const double *q_unit;
    double grad_x, grad_z, qx_mean, qz_mean;
    int grad_pre;
    // outputs
    zd *P, *V[3];
    zd *div_out;  // influence: raw wall divergence [q][2]
    zd *sig_out;  // b problem: sigma0 [q][2]
    double *flow; // step: dpdx, dpdz
};
// Discrete adjoint of the step (no reference counterpart): the transposed solve with the RHS,
// history and wall-condition transposes (k_adj_solve), the output weights and curl^T of the
// transposed nonlinear term (k_adj_combine), and the terminal objective (k_objective).
struct AdjArgs {
    int ny, JS, NQ, nzk, nkx, NX2, NY2, collide, which;
    const double *WAVX, *WAVZ, *WL, *WD, *WR;
    double Re, G_pre, g_v, pa, pb, XL, ZL;
    double aV[3], bV[3];
    double a0, a1, b0, b1;
    int grad_pre;
    const double *ad_p, *ad_v;
    const zd *PI[2], *VI[2][3], *ID[2], *VB[3], *SIG0;
    const double *u_unit, *w_unit, *q_unit;
    // adjoint state (in/out): velocity, previous velocity, previous nonlinear term, previous
    // pressure wall condition [q][2]; lf accumulates the body-force gradient (or null)
    zd *lu[3], *lup[3], *lnp[3], *lbp, *lf[3];
    zd *lNs[3]; // out: weighted adjoint of the nonlinear term (input of the transposed synthesis)
    zd *lw[3];  // out: adjoint of the curl from the wall condition
    const zd *lUs[3], *lWs[3]; // transposed-product results (inputs of k_adj_combine)
    zd *grad[3];
    double *Jpart;
};
void launch_adj_solve(const AdjArgs &a, int E, hipStream_t st);
void launch_adj_combine(const AdjArgs &a, int E, hipStream_t st);
void launch_adj_fold(const AdjArgs &a, hipStream_t st);
void launch_objective(const AdjArgs &a, int E, hipStream_t st);

// diagnostics on the coarse physical grid
constexpr int kPlaneSums = 21; // the 20 sums of Statistics::reduce and the plane sum of div^2
struct DiagArgs {
    const zd *u, *v, *w; // spectral planes
    zd *out;
    const double *WAVX, *WAVZ;
    int NQ, nzk, JS, ny;
    const double *pu, *pv, *pw, *pp, *pdiv; // physical [iz][iy][ix]; pdiv may be null
    int NX, NZ, NL, higher;
    double *sums; // [NL][kPlaneSums]
    int *finite;  // [NL]
};
void launch_divergence(const DiagArgs &a, int E, hipStream_t st);
void launch_planes_finite(const zd *planes, size_t nplanes, int NQ, int JS, int ny, int *flags, int nblocks,
                          hipStream_t st);
void launch_plane_moments(const DiagArgs &a, hipStream_t st);

void launch_solve(const SolveArgs &a, int E, hipStream_t st);
// real tables of the compact form from the complex setup tables
void launch_compact_tables(const SolveArgs &a, double *PIr0, double *PIr1, double *VIv0, double *VIv1,
                           double *VIf0, double *VIf1, double *PBr, double *VBv, double *VBf, hipStream_t st);
void launch_build_ad(const double *WAVX, const double *WAVZ, const double *WL, const double *WD,
                     const double *WR, double g, int ny, int JS, int NQ, int nzk, double *ad,
                     double *wn, hipStream_t st);
void launch_normalize_influence(zd *ID0, zd *ID1, int NQ, int root_q, double aV1, hipStream_t st);
void launch_pg_response(const double *WAVX, const double *WAVZ, const double *WL, const double *WD,
                        const double *WR, const double *ad_v, double g_v, double Re, const double *aV,
                        const double *bV, int ny, int JS, int root_q, double *u_unit, double *w_unit,
                        double *q_unit, int E, hipStream_t st);

} // namespace cdns
