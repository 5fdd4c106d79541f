// y stages as dense Chebyshev matrix applies on the matrix cores (16x16x4 MFMA, FP64 or FP32).
//
// Backward (spectral -> Y):  Y[l][q] = sum_j B[j][l] S[q][j], split by parity of j:
//   E[l'] = sum_{j even} B[j][l'] S_j,  O[l'] = sum_{j odd} B[j][l'] S_j,
//   Y[l'] = E + O,  Y[NY2 - l'] = E - O  for l' < NY2/2; for even NY2 the middle row
//   l' = NY2/2 only has even terms and is summed on the vector ALUs.
//   MFMA: D[l'][col] += A[l'][j] * B[j][col], A = the Chebyshev table (LDS), B = the spectral
//   data (one real column per (q, re|im), read straight from HBM).
// Forward (Y -> spectral):  S[q][j] = sum_{l'} F_p[l'][m] (Y[l'] +- Y[NY2 - l'])[q], j = 2m + p.
//   MFMA: D[col][m] += A[col][l'] * B[l'][m], A = the folded data, B = the table (LDS); each
//   lane ends with the re/im pair of one (q, j) and writes it as one complex.
// A wave owns 8*CT complex columns; its work is split over row parts (backward) or j parts
// (forward) so that each part's table fits in 64 KB of LDS and the grid has enough waves.
//
// MFMA operand layout (16x16x4): A[i][k] at lane i + 16k, B[k][n] at lane n + 16k.  The
// result D[i][n] sits in lane n + 16*kk, register r with i = 4*kk + r for FP32 and
// i = 4*r + kk for FP64 (Mfma<R>::row).  The forward kernel orders its data columns so that
// each lane ends with (re, im) pairs of the same column in registers (0,1) and (2,3).
#include "internal.hpp"
#include "regfft.hpp"

#include <algorithm>

namespace cdns {

// row(kk, r): row index i of D held in register r of a lane with lane / 16 = kk.
// qcol/pcol: complex column (of 8) and part (0 re, 1 im) of data column c (forward kernel);
// qout(kk, h): the complex column whose (re, im) a lane holds in registers (2h, 2h + 1).
template <typename R> struct Mfma;
template <> struct Mfma<double> {
    typedef double v4 __attribute__((ext_vector_type(4)));
    __device__ static inline v4 mma(double a, double b, v4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    __device__ static inline int row(int kk, int r) { return 4 * r + kk; }
    __device__ static inline int qcol(int c) { return (c & 3) + 4 * (c >> 3); }
    __device__ static inline int pcol(int c) { return (c >> 2) & 1; }
    __device__ static inline int qout(int kk, int h) { return kk + 4 * h; }
};
template <> struct Mfma<float> {
    typedef float v4 __attribute__((ext_vector_type(4)));
    __device__ static inline v4 mma(float a, float b, v4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    __device__ static inline int row(int kk, int r) { return 4 * kk + r; }
    __device__ static inline int qcol(int c) { return c >> 1; }
    __device__ static inline int pcol(int c) { return c & 1; }
    __device__ static inline int qout(int kk, int h) { return 2 * kk + h; }
};

constexpr int kLdsBudget = 64 * 1024;
constexpr int kWaves = 8;

// LDS table fill: dst[e] = R(src[i]) where idx(e, i) holds, else 0, for e < n.  Eight loads per
// thread are issued before any is used (unconditional, at a clamped index), so a block's fill
// costs about one global-memory latency per eight rows instead of one per row.
template <typename R, typename Idx>
__device__ inline void fill_lds(R *dst, const double *src, int n, Idx idx) {
    constexpr int U = 8;
    for (int e0 = threadIdx.x; e0 < n; e0 += U * (int)blockDim.x) {
        double v[U];
        bool ok[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int e = e0 + u * (int)blockDim.x;
            int i = 0;
            ok[u] = e < n && idx(e, i);
            v[u] = src[ok[u] ? i : 0];
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int e = e0 + u * (int)blockDim.x;
            if (e < n) dst[e] = ok[u] ? R(v[u]) : R(0);
        }
    }
}
// fused backward: 4-wavefront blocks; one row tile (16 l') per wavefront (FP32 at Re_tau=180:
// 112 VGPRs, 4 waves per SIMD, 50.9 against 53.6 us with two tiles on 8-wavefront blocks), two
// for long FP32 columns (see launch_ybwd_curl); the loads are double-buffered; LDS budget of its
// two tables
constexpr int kYCurlWaves = 4;
constexpr int kYCurlLds = 64 * 1024;
// forward: waves per block (the parts' tables are shared by all of a block's column groups)
constexpr int kYFwdWaves = 8;

// ------------------------------------------------------------------ backward --
// grid: (column groups of 8 waves, row parts, fields); TRP row tiles (16 l' each) per part.
template <typename R, int TRP, int CT>
__global__ void __launch_bounds__(64 * kWaves) k_ybwd_mfma(YArgs a) {
    using M = Mfma<R>;
    using v4 = typename M::v4;
    constexpr int PITCH = TRP * 16;
    extern __shared__ unsigned char smem_raw[];
    R *Bt = reinterpret_cast<R *>(smem_raw); // [njp][PITCH]
    R *midc = Bt + (size_t)a.njp * PITCH;    // [njp], middle-row coefficients
    // logical block: the parts of one column group are neighbours on one XCD (shared columns)
    const int L = xcd_block((int)gridDim.x);
    const int part = L % a.parts, cgf = L / a.parts, cg = cgf % a.ncg, f = cgf / a.ncg;
    const int l0 = part * PITCH;
    const bool do_mid = a.has_mid && part == a.parts - 1;
    const int NLH = a.NLH, nj = a.nj, rows = a.rows_m;
    auto tidx = [&](int e, int &i) {
        const int j = e / PITCH, lp = l0 + (e - j * PITCH);
        i = j * NLH + lp;
        return j < nj && lp < rows;
    };
    fill_lds(Bt, a.coef, a.njp * PITCH, tidx);
    if (do_mid)
        fill_lds(midc, a.coef, a.njp, [&](int j, int &i) {
            i = j * NLH + NLH - 1;
            return j < nj;
        });
    __syncthreads();

    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int col = lane & 15, k = lane >> 4, hh = col & 1;
    const int qb = (cg * kWaves + wv) * 8 * CT;
    // FP32: lane (col, k) loads the whole complex S[q][8s + 2k + (col & 1)], q = qb + 8c + col/2,
    // so that each load instruction reads eight columns x 128 contiguous bytes; the MFMA operands
    // are one part of the even-j and one of the odd-j value, its own and its pair partner's (DPP
    // lane swap).  FP64: lane (col, k) loads part col & 1 of S[q][8s + 2k] and S[q][8s + 2k + 1]
    // (two 8-byte loads), which measures faster there than the 16-byte loads and lane swaps.
    constexpr bool kFused = sizeof(R) == 4;
    const zd *src[CT];
    const double *srcp[CT];
    bool okq[CT];
#pragma unroll
    for (int c = 0; c < CT; ++c) {
        const int q = qb + 8 * c + (col >> 1);
        okq[c] = q < a.NQ;
        src[c] = a.spec[f] + (size_t)(okq[c] ? q : 0) * a.JS;
        srcp[c] = reinterpret_cast<const double *>(a.spec[f]) + ((size_t)(okq[c] ? q : 0) * a.JS) * 2 + (col & 1);
    }
    v4 accE[CT][TRP], accO[CT][TRP];
    R mid[CT];
#pragma unroll
    for (int c = 0; c < CT; ++c) {
        mid[c] = R(0);
#pragma unroll
        for (int t = 0; t < TRP; ++t) {
            accE[c][t] = v4{R(0), R(0), R(0), R(0)};
            accO[c][t] = v4{R(0), R(0), R(0), R(0)};
        }
    }
    // k-step s covers j = 8s .. 8s+7: this lane's even j = 8s + 2k, odd j = 8s + 2k + 1.  Every
    // load is unconditional (column and j clamped into the plane) and j >= nj is masked by a
    // multiply: a select or a load under a branch makes the compiler wait for every load in
    // flight, the next step's prefetch included.  Columns past NQ compute garbage that is not
    // stored.
    const int nsteps = a.njp / 8;
    struct Buf { // raw loads of one step
        zd v[CT];
        double e[CT], o[CT];
    };
    auto load = [&](Buf &b, int s) {
        const int je = 8 * s + 2 * k;
#pragma unroll
        for (int c = 0; c < CT; ++c) {
            if constexpr (kFused) {
                const int j = je + hh;
                b.v[c] = src[c][j < nj ? j : 0];
            } else {
                const int jo = je + 1;
                b.e[c] = srcp[c][2 * (je < nj ? je : 0)];
                b.o[c] = srcp[c][2 * (jo < nj ? jo : 0)];
            }
        }
    };
    auto compute = [&](const Buf &b, int s) {
        const int je = 8 * s + 2 * k;
        const double mkE = je < nj ? 1.0 : 0.0, mkO = je + 1 < nj ? 1.0 : 0.0;
        R cE[CT], cO[CT];
#pragma unroll
        for (int c = 0; c < CT; ++c) {
            if constexpr (kFused) {
                const R mk = R(hh ? mkO : mkE);
                const R vx = R(b.v[c].x) * mk, vy = R(b.v[c].y) * mk;
                const R px = pair_swap(vx), py = pair_swap(vy);
                // even lanes: Re S[j even] (own), Re S[j odd] (partner); odd lanes: Im S[j even]
                // (partner), Im S[j odd] (own)
                cE[c] = hh ? py : vx;
                cO[c] = hh ? vy : px;
            } else {
                cE[c] = R(b.e[c] * mkE);
                cO[c] = R(b.o[c] * mkO);
            }
        }
        const R *rowE = Bt + (size_t)je * PITCH + col, *rowO = rowE + PITCH;
#pragma unroll
        for (int t = 0; t < TRP; ++t) {
            const R aE = rowE[16 * t], aO = rowO[16 * t];
#pragma unroll
            for (int c = 0; c < CT; ++c) {
                accE[c][t] = M::mma(aE, cE[c], accE[c][t]);
                accO[c][t] = M::mma(aO, cO[c], accO[c][t]);
            }
        }
        if (do_mid) {
            const R mc = midc[je];
#pragma unroll
            for (int c = 0; c < CT; ++c) mid[c] += mc * cE[c];
        }
    };
    // two buffers: the loads of a step land in registers the running step does not read; loads
    // past the last step read clamped addresses and are not used
    Buf b0, b1;
    load(b0, 0);
    for (int s = 0; s < nsteps; s += 2) {
        load(b1, s + 1);
        compute(b0, s);
        load(b0, s + 2);
        if (s + 1 < nsteps) compute(b1, s + 1);
    }
    R *Y = reinterpret_cast<R *>(a.y) + (size_t)f * a.YS * 2;
    const int NY2 = a.NY2;
#pragma unroll
    for (int c = 0; c < CT; ++c) {
        const int q = qb + 8 * c + (col >> 1);
        if (do_mid) {
            R v = mid[c];
            v += __shfl_xor(v, 16);
            v += __shfl_xor(v, 32);
            if (k == 0 && okq[c]) Y[((size_t)(NLH - 1) * a.NQ + q) * 2 + (col & 1)] = v;
        }
        if (!okq[c]) continue;
#pragma unroll
        for (int t = 0; t < TRP; ++t)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int lp = l0 + 16 * t + M::row(k, r);
                if (lp < rows) {
                    const R e = accE[c][t][r], o = accO[c][t][r];
                    Y[((size_t)lp * a.NQ + q) * 2 + (col & 1)] = e + o;
                    Y[((size_t)(NY2 - lp) * a.NQ + q) * 2 + (col & 1)] = e - o;
                }
            }
    }
}

// ---------------------------------------------------- backward with the curl --
// The velocity's y backward with vector_ops::enqueue_curl (src/solver/vector_ops.cpp:112-151)
// and the curl part of the pressure wall condition (src/solver/timestepper.cpp:26-43) fused in.
// The y derivative is applied as a second synthesis matrix, T'_j(y_l) (the exact derivative of
// the Chebyshev series at the Gauss-Lobatto points, what the recurrence of vector_ops.cpp:17-25
// followed by the synthesis gives), so the six physical fields
//   u, v, w,  wx = w' - i kz v,  wy = i kz u - i kx w,  wz = i kx v - u'
// come from the three velocity planes and the vorticity planes are never formed.  Per column
// group: five matrix applies (T on u, v, w; T' on u, w).  T'_j(-y) = (-1)^(j+1) T'_j(y): the
// derivative rows l and NY2 - l are O + E and O - E.
// Wall condition (FP64 in both precisions, from the loaded values): with v(+-1) = sum (+-1)^j v_j
// and u'(+-1) = sum (+-1)^(j+1) j^2 u_j,
//   bcw(wall 0) = (-k^2 v(1) - i kx u'(1) - i kz w'(1)) / Re,  bcw(wall 1) = -(same at y = -1),
// which is sum_j t_j and sum_j (-1)^(j+1) t_j of t_j = (i kx/Re) wz_j - (i kz/Re) wx_j.
// grid: (column groups of W waves x 8 complex columns, row parts); TRP row tiles per part.
template <typename R, int TRP, int W>
__global__ void __launch_bounds__(64 * W) k_ybwd_curl(YArgs a) {
    using M = Mfma<R>;
    using v4 = typename M::v4;
    constexpr int PITCH = TRP * 16;
    constexpr bool kFused = sizeof(R) == 4; // 16-byte loads + lane-pair swaps, as k_ybwd_mfma
    extern __shared__ unsigned char smem_raw[];
    R *Bt = reinterpret_cast<R *>(smem_raw); // [njp][PITCH] T_j(y_l)
    R *Dt = Bt + (size_t)a.njp * PITCH;      // [njp][PITCH] T'_j(y_l)
    R *midc = Dt + (size_t)a.njp * PITCH;    // [njp] T_j(0)
    R *midd = midc + a.njp;                  // [njp] T'_j(0)
    const int L = xcd_block((int)gridDim.x);
    const int part = L % a.parts, cg = L / a.parts;
    const int l0 = part * PITCH;
    const bool do_mid = a.has_mid && part == a.parts - 1;
    const bool do_wall = a.bcw != nullptr && part == 0;
    const int NLH = a.NLH, nj = a.nj, rows = a.rows_m;
    auto tidx = [&](int e, int &i) {
        const int j = e / PITCH, lp = l0 + (e - j * PITCH);
        i = j * NLH + lp;
        return j < nj && lp < rows;
    };
    auto midx = [&](int j, int &i) {
        i = j * NLH + NLH - 1;
        return j < nj;
    };
    fill_lds(Bt, a.coef, a.njp * PITCH, tidx);
    fill_lds(Dt, a.coefd, a.njp * PITCH, tidx);
    if (do_mid) {
        fill_lds(midc, a.coef, a.njp, midx);
        fill_lds(midd, a.coefd, a.njp, midx);
    }
    __syncthreads();

    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int col = lane & 15, k = lane >> 4, hh = col & 1;
    const int q = (cg * W + wv) * 8 + (col >> 1);
    const bool okq = q < a.NQ;
    // byte offset of this lane's first value in every velocity plane (the planes share a layout):
    // loads are a uniform plane base plus this 32-bit offset plus a constant
    const unsigned colb = (unsigned)((okq ? q : 0) * a.JS) * 16u;
    const char *pl[3];
#pragma unroll
    for (int f = 0; f < 3; ++f) pl[f] = reinterpret_cast<const char *>(a.spec[f]);
    const unsigned lane_off = kFused ? colb + (unsigned)(2 * k + hh) * 16u : colb + (unsigned)(2 * k) * 16u + (unsigned)hh * 8u;
    // accumulators: 0 U, 1 V, 2 W (T), 3 U', 4 W' (T')
    v4 accE[5][TRP], accO[5][TRP];
    R mid[5];
#pragma unroll
    for (int f = 0; f < 5; ++f) {
        mid[f] = R(0);
#pragma unroll
        for (int t = 0; t < TRP; ++t) {
            accE[f][t] = v4{R(0), R(0), R(0), R(0)};
            accO[f][t] = v4{R(0), R(0), R(0), R(0)};
        }
    }
    // wall sums v(1), v(-1), u'(1), u'(-1), w'(1), w'(-1): FP64 own part (FP64 kernel) or
    // complex (FP32 kernel, which holds whole complex values)
    double ws[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    zd wc[6] = {};
    const int nsteps = a.njp / 8;
    // two load buffers, so that the loads of step s + 1 land in registers that step s does not
    // read (no copies that would wait for them)
    struct Buf { // raw loads; j >= nj (inside the column's padding, j < njp <= JS) masked in the last step
        zd v[3];
        double e[3], o[3];
    };
    auto load = [&](Buf &b, int s) {
        // steps past the last (prefetch) reload the last step, inside the column
        const unsigned o = lane_off + (unsigned)(s < nsteps ? s : nsteps - 1) * 128u;
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            if constexpr (kFused) {
                b.v[f] = *reinterpret_cast<const zd *>(pl[f] + o);
            } else {
                b.e[f] = *reinterpret_cast<const double *>(pl[f] + o);
                b.o[f] = *reinterpret_cast<const double *>(pl[f] + o + 16u);
            }
        }
    };
    auto compute = [&](const Buf &b, int s) {
        const int je = 8 * s + 2 * k;
        zd v[3];
        double e[3], o[3];
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            v[f] = b.v[f];
            e[f] = b.e[f];
            o[f] = b.o[f];
        }
        if (s == nsteps - 1) { // the only step that can reach j >= nj
            const double mkE = je < nj ? 1.0 : 0.0, mkO = je + 1 < nj ? 1.0 : 0.0;
#pragma unroll
            for (int f = 0; f < 3; ++f) {
                if constexpr (kFused) {
                    const double mk = hh ? mkO : mkE;
                    v[f] = {v[f].x * mk, v[f].y * mk};
                } else {
                    e[f] *= mkE;
                    o[f] *= mkO;
                }
            }
        }
        if (do_wall) {
            if constexpr (kFused) {
                const int j = je + hh;
                const double j2 = double(j) * double(j), sj = hh ? -1.0 : 1.0; // (-1)^j
                wc[0] = wc[0] + v[1];
                wc[1] = wc[1] + sj * v[1];
                wc[2] = wc[2] + j2 * v[0];
                wc[3] = wc[3] + (-sj * j2) * v[0];
                wc[4] = wc[4] + j2 * v[2];
                wc[5] = wc[5] + (-sj * j2) * v[2];
            } else {
                const double e2 = double(je) * double(je), o2 = double(je + 1) * double(je + 1);
                ws[0] += e[1] + o[1];
                ws[1] += e[1] - o[1];
                ws[2] += e2 * e[0] + o2 * o[0];
                ws[3] += o2 * o[0] - e2 * e[0];
                ws[4] += e2 * e[2] + o2 * o[2];
                ws[5] += o2 * o[2] - e2 * e[2];
            }
        }
        R cE[3], cO[3];
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            if constexpr (kFused) {
                const R vx = R(v[f].x), vy = R(v[f].y);
                const R px = pair_swap(vx), py = pair_swap(vy);
                cE[f] = hh ? py : vx;
                cO[f] = hh ? vy : px;
            } else {
                cE[f] = R(e[f]);
                cO[f] = R(o[f]);
            }
        }
        const R *rowE = Bt + (size_t)je * PITCH + col, *rowO = rowE + PITCH;
        const R *drE = Dt + (size_t)je * PITCH + col, *drO = drE + PITCH;
#pragma unroll
        for (int t = 0; t < TRP; ++t) {
            const R aE = rowE[16 * t], aO = rowO[16 * t], dE = drE[16 * t], dO = drO[16 * t];
#pragma unroll
            for (int f = 0; f < 3; ++f) {
                accE[f][t] = M::mma(aE, cE[f], accE[f][t]);
                accO[f][t] = M::mma(aO, cO[f], accO[f][t]);
            }
            accE[3][t] = M::mma(dE, cE[0], accE[3][t]);
            accO[3][t] = M::mma(dO, cO[0], accO[3][t]);
            accE[4][t] = M::mma(dE, cE[2], accE[4][t]);
            accO[4][t] = M::mma(dO, cO[2], accO[4][t]);
        }
        if (do_mid) {
            // T_j(0) vanishes for odd j, T'_j(0) for even j
            const R mc = midc[je], md = midd[je + 1];
#pragma unroll
            for (int f = 0; f < 3; ++f) mid[f] += mc * cE[f];
            mid[3] += md * cO[0];
            mid[4] += md * cO[2];
        }
    };
    // loads past the last step reread the last step (inside the column) and are not used: no
    // branch around a load, so the compiler counts the loads in flight exactly
    Buf b0, b1;
    load(b0, 0);
    for (int s = 0; s < nsteps; s += 2) {
        load(b1, s + 1);
        compute(b0, s);
        load(b0, s + 2);
        if (s + 1 < nsteps) compute(b1, s + 1);
    }

    const int kxi = okq ? q / a.nzk : 0, kzi = okq ? q - kxi * a.nzk : 0;
    const double kxd = a.wavx[kxi], kzd = a.wavz[kzi];
    if (do_wall) {
        zd sum[6];
        if constexpr (kFused) {
#pragma unroll
            for (int i = 0; i < 6; ++i) {
                zd v = wc[i];
                v.x += __shfl_xor(v.x, 16);
                v.y += __shfl_xor(v.y, 16);
                v.x += __shfl_xor(v.x, 32);
                v.y += __shfl_xor(v.y, 32);
                v.x += __shfl_xor(v.x, 1);
                v.y += __shfl_xor(v.y, 1);
                sum[i] = v;
            }
        } else {
#pragma unroll
            for (int i = 0; i < 6; ++i) {
                double v = ws[i];
                v += __shfl_xor(v, 16);
                v += __shfl_xor(v, 32);
                const double o = __shfl_xor(v, 1);
                sum[i] = hh ? zd{o, v} : zd{v, o};
            }
        }
        if (okq && k == 0 && hh == 0) {
            const double k2 = kxd * kxd + kzd * kzd, inv = 1.0 / a.Re;
            // -k^2 v - i kx u' - i kz w'
            const zd wp = {-k2 * sum[0].x + kxd * sum[2].y + kzd * sum[4].y,
                           -k2 * sum[0].y - kxd * sum[2].x - kzd * sum[4].x};
            const zd wm = {-k2 * sum[1].x + kxd * sum[3].y + kzd * sum[5].y,
                           -k2 * sum[1].y - kxd * sum[3].x - kzd * sum[5].x};
            a.bcw[(size_t)q * 2 + 0] = {wp.x * inv, wp.y * inv};
            a.bcw[(size_t)q * 2 + 1] = {-wm.x * inv, -wm.y * inv};
        }
    }

    // i k X in this lane's part: sg k (X's other part), which the pair partner holds
    const R kx = R(kxd), kz = R(kzd), sg = hh ? R(1) : R(-1);
    R *Y = reinterpret_cast<R *>(a.y);
    const size_t FS = a.YS * 2;
    const int NY2 = a.NY2;
    if (do_mid) {
        R m[5];
#pragma unroll
        for (int f = 0; f < 5; ++f) {
            R v = mid[f];
            v += __shfl_xor(v, 16);
            v += __shfl_xor(v, 32);
            m[f] = v;
        }
        const R su = pair_swap(m[0]), sv = pair_swap(m[1]), sw = pair_swap(m[2]);
        const R out[6] = {m[0], m[1], m[2], m[4] - sg * kz * sv, sg * (kz * su - kx * sw), sg * kx * sv - m[3]};
        if (k == 0 && okq) {
            const size_t o = ((size_t)(NLH - 1) * a.NQ + q) * 2 + hh;
#pragma unroll
            for (int f = 0; f < 6; ++f) Y[f * FS + o] = out[f];
        }
    }
#pragma unroll
    for (int t = 0; t < TRP; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            const int lp = l0 + 16 * t + M::row(k, r);
            R up[3], um[3];
#pragma unroll
            for (int f = 0; f < 3; ++f) {
                up[f] = accE[f][t][r] + accO[f][t][r];
                um[f] = accE[f][t][r] - accO[f][t][r];
            }
            const R dup = accO[3][t][r] + accE[3][t][r], dum = accO[3][t][r] - accE[3][t][r];
            const R dwp = accO[4][t][r] + accE[4][t][r], dwm = accO[4][t][r] - accE[4][t][r];
            R sp[3], sm[3];
#pragma unroll
            for (int f = 0; f < 3; ++f) {
                sp[f] = pair_swap(up[f]);
                sm[f] = pair_swap(um[f]);
            }
            if (okq && lp < rows) {
                const size_t op = ((size_t)lp * a.NQ + q) * 2 + hh, om = ((size_t)(NY2 - lp) * a.NQ + q) * 2 + hh;
#pragma unroll
                for (int f = 0; f < 3; ++f) {
                    Y[f * FS + op] = up[f];
                    Y[f * FS + om] = um[f];
                }
                Y[3 * FS + op] = dwp - sg * kz * sp[1];
                Y[3 * FS + om] = dwm - sg * kz * sm[1];
                Y[4 * FS + op] = sg * (kz * sp[0] - kx * sp[2]);
                Y[4 * FS + om] = sg * (kz * sm[0] - kx * sm[2]);
                Y[5 * FS + op] = sg * kx * sp[1] - dup;
                Y[5 * FS + om] = sg * kx * sm[1] - dum;
            }
        }
}

// ------------------------------------------------------------------- forward --
// grid: (column groups of W waves, j parts, fields); JTP j tiles (16 m each) per parity per part.
// (W = 16 puts the two parts of the Re_tau=180 grid in one round of blocks instead of 1.02
// rounds of 8-wave blocks, but measures 3 us slower in FP64 and equal in FP32.)
template <typename R, int JTP, int CT, int W>
__global__ void __launch_bounds__(64 * W) k_yfwd_mfma(YArgs a) {
    using M = Mfma<R>;
    using v4 = typename M::v4;
    constexpr int PITCH = JTP * 16;
    extern __shared__ unsigned char smem_raw[];
    R *Fe = reinterpret_cast<R *>(smem_raw); // [KL][PITCH]
    R *Fo = Fe + (size_t)a.KL * PITCH;       // [KL][PITCH]
    R *Fm = Fo + (size_t)a.KL * PITCH;       // [PITCH], middle row (even j only)
    const int L = xcd_block((int)gridDim.x);
    const int part = L % a.parts, cgf = L / a.parts, cg = cgf % a.ncg, f = cgf / a.ncg;
    const int m0 = part * PITCH;
    const int rows = a.rows_m, NLH = a.NLH, nje = a.nje, njo = a.njo;
    fill_lds(Fe, a.coef, a.KL * PITCH, [&](int e, int &i) {
        const int lp = e / PITCH, m = m0 + (e - lp * PITCH);
        i = lp * nje + m;
        return lp < rows && m < nje;
    });
    fill_lds(Fo, a.coef2, a.KL * PITCH, [&](int e, int &i) {
        const int lp = e / PITCH, m = m0 + (e - lp * PITCH);
        i = lp * njo + m;
        return lp < rows && m < njo;
    });
    fill_lds(Fm, a.coef, PITCH, [&](int e, int &i) {
        const int m = m0 + e;
        i = (NLH - 1) * nje + m;
        return a.has_mid && m < nje;
    });
    __syncthreads();

    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int col = lane & 15, k = lane >> 4;
    const int qb = (cg * W + wv) * 8 * CT;
    const R *Y = reinterpret_cast<const R *>(a.y) + (size_t)f * a.YS * 2;
    const int NQ = a.NQ, NY2 = a.NY2;
    int qa[CT];
    bool oka[CT];
    const int pa = M::pcol(col);
#pragma unroll
    for (int c = 0; c < CT; ++c) {
        qa[c] = qb + 8 * c + M::qcol(col);
        oka[c] = qa[c] < NQ;
    }
    v4 accE[CT][JTP], accO[CT][JTP];
#pragma unroll
    for (int c = 0; c < CT; ++c)
#pragma unroll
        for (int t = 0; t < JTP; ++t) {
            accE[c][t] = v4{R(0), R(0), R(0), R(0)};
            accO[c][t] = v4{R(0), R(0), R(0), R(0)};
        }
    // k-step s: l' = 4s + k
    const int nsteps = a.KL / 4;
    // loads are unconditional, so that the compiler counts the loads in flight exactly and no
    // select waits for them: a row l' >= rows reads row rows - 1 (its table rows are zero) and a
    // column q >= NQ reads column 0 (its outputs are not written)
    // byte offsets within the field (a field is < 4 GB): loads are the uniform field base plus a
    // 32-bit row offset plus this lane's column offset
    const char *Yb = reinterpret_cast<const char *>(Y);
    const unsigned rowb = (unsigned)NQ * 2u * (unsigned)sizeof(R);
    unsigned ob[CT];
#pragma unroll
    for (int c = 0; c < CT; ++c) ob[c] = ((unsigned)(oka[c] ? qa[c] : 0) * 2u + (unsigned)pa) * (unsigned)sizeof(R);
    struct Buf {
        R y1[CT], y2[CT];
    };
    auto load = [&](Buf &b, int s) {
        int lp = 4 * s + k;
        lp = lp < rows ? lp : rows - 1;
        const unsigned r1 = (unsigned)lp * rowb, r2 = (unsigned)(NY2 - lp) * rowb;
#pragma unroll
        for (int c = 0; c < CT; ++c) {
            b.y1[c] = *reinterpret_cast<const R *>(Yb + (r1 + ob[c]));
            b.y2[c] = *reinterpret_cast<const R *>(Yb + (r2 + ob[c]));
        }
    };
    auto compute = [&](const Buf &b, int s) {
        R sv[CT], dv[CT];
#pragma unroll
        for (int c = 0; c < CT; ++c) {
            sv[c] = b.y1[c] + b.y2[c];
            dv[c] = b.y1[c] - b.y2[c];
        }
        const int lp = 4 * s + k;
        const R *rowE = Fe + (size_t)lp * PITCH + col, *rowO = Fo + (size_t)lp * PITCH + col;
#pragma unroll
        for (int t = 0; t < JTP; ++t) {
            const R be = rowE[16 * t], bo = rowO[16 * t];
#pragma unroll
            for (int c = 0; c < CT; ++c) {
                accE[c][t] = M::mma(sv[c], be, accE[c][t]);
                accO[c][t] = M::mma(dv[c], bo, accO[c][t]);
            }
        }
    };
    // four load buffers, the loads running two steps ahead into registers that the steps in
    // between do not read; the step count is a multiple of 4 (KL is a multiple of 16), so every
    // compute is unconditional and the compiler cannot sink a buffer's loads to its use
    Buf b0, b1, b2, b3;
    load(b0, 0);
    load(b1, 1);
    for (int s = 0; s < nsteps; s += 4) {
        load(b2, s + 2);
        load(b3, s + 3);
        __builtin_amdgcn_sched_barrier(0);
        compute(b0, s);
        compute(b1, s + 1);
        __builtin_amdgcn_sched_barrier(0);
        load(b0, s + 4);
        load(b1, s + 5);
        __builtin_amdgcn_sched_barrier(0);
        compute(b2, s + 2);
        compute(b3, s + 3);
        __builtin_amdgcn_sched_barrier(0);
    }
    // this lane holds (re, im) of complex column qout(k, h) in registers (2h, 2h + 1), for
    // m = m0 + 16t + col
#pragma unroll
    for (int c = 0; c < CT; ++c) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int q = qb + 8 * c + M::qout(k, h);
            if (q >= NQ) continue;
            cpx<R> ym = {R(0), R(0)};
            if (a.has_mid) {
                const R *pm = Y + ((size_t)(NLH - 1) * NQ + q) * 2;
                ym = {pm[0], pm[1]};
            }
            zd *out = a.spec_out[f] + (size_t)q * a.JS;
            cpx<float> *out32 = reinterpret_cast<cpx<float> *>(a.spec_out[f]) + (size_t)q * a.JS;
            const zd *add = a.add[f] ? a.add[f] + (size_t)q * a.JS : nullptr;
            auto put = [&](int j, zd v) {
                if (add) v = v + add[j];
                if (a.out32) out32[j] = {float(v.x), float(v.y)};
                else out[j] = v;
            };
#pragma unroll
            for (int t = 0; t < JTP; ++t) {
                const int m = m0 + 16 * t + col;
                const R fm = Fm[16 * t + col];
                if (m < nje) {
                    const R re = accE[c][t][2 * h] + fm * ym.x, im = accE[c][t][2 * h + 1] + fm * ym.y;
                    put(2 * m, zd{double(re), double(im)});
                }
                if (m < njo) put(2 * m + 1, zd{double(accO[c][t][2 * h]), double(accO[c][t][2 * h + 1])});
            }
        }
    }
}

// --------------------------------------------------------------------- launch --
// 1-D grids of (parts x column groups x fields) blocks, decoded by xcd_block.
template <typename R, int TRP, int CT> static void go_ybwd(YArgs a, int parts, int nf, hipStream_t st) {
    const size_t sm = ((size_t)a.njp * TRP * 16 + a.njp) * sizeof(R);
    const int cols = 8 * CT * kWaves;
    a.parts = parts;
    a.ncg = (a.NQ + cols - 1) / cols;
    hipLaunchKernelGGL((k_ybwd_mfma<R, TRP, CT>), dim3(a.parts * a.ncg * nf, 1), dim3(64 * kWaves), sm, st, a);
}
template <typename R, int JTP, int CT, int W = kYFwdWaves>
static void go_yfwd(YArgs a, int parts, int nf, hipStream_t st) {
    const size_t sm = ((size_t)2 * a.KL * JTP * 16 + JTP * 16) * sizeof(R);
    const int cols = 8 * CT * W;
    a.parts = parts;
    a.ncg = (a.NQ + cols - 1) / cols;
    hipLaunchKernelGGL((k_yfwd_mfma<R, JTP, CT, W>), dim3(a.parts * a.ncg * nf, 1), dim3(64 * W), sm, st, a);
}

static void y_shape(YArgs &a) {
    a.has_mid = (a.NY2 % 2 == 0) ? 1 : 0;
    a.rows_m = a.NLH - a.has_mid;
    a.njp = (a.nj + 7) / 8 * 8;
    // forward table rows: a multiple of 16, so that the forward kernel's k-steps (4 rows each) come
    // in groups of four
    a.KL = (a.rows_m + 15) / 16 * 16;
}

// largest tile count per part, at most `cap`, whose table fits the LDS budget
static int tiles_per_part(size_t bytes_per_tile, size_t fixed, int cap) {
    int t = cap;
    while (t > 1 && fixed + t * bytes_per_tile > (size_t)kLdsBudget) --t;
    return t;
}

bool launch_ybwd_mfma(YArgs a, int nf, bool fp32, hipStream_t st) {
    y_shape(a);
    const size_t esz = fp32 ? 4 : 8;
    const int TR = (a.rows_m + 15) / 16;
    if ((size_t)a.njp * 16 * esz + a.njp * esz > (size_t)kLdsBudget) return false;
    const int fit = std::min(tiles_per_part((size_t)a.njp * 16 * esz, a.njp * esz, 4), TR);
    const int parts = (TR + fit - 1) / fit;
    const int trp = (TR + parts - 1) / parts; // balanced parts; the kernel masks rows >= rows_m
    if (fp32) {
        switch (trp) {
        case 1: go_ybwd<float, 1, 1>(a, parts, nf, st); break;
        case 2: go_ybwd<float, 2, 1>(a, parts, nf, st); break;
        case 3: go_ybwd<float, 3, 1>(a, parts, nf, st); break;
        default: go_ybwd<float, 4, 1>(a, parts, nf, st); break;
        }
    } else {
        switch (trp) {
        case 1: go_ybwd<double, 1, 1>(a, parts, nf, st); break;
        case 2: go_ybwd<double, 2, 1>(a, parts, nf, st); break;
        case 3: go_ybwd<double, 3, 1>(a, parts, nf, st); break;
        default: go_ybwd<double, 4, 1>(a, parts, nf, st); break;
        }
    }
    return true;
}

template <typename R, int TRP, int W> static void go_ybwd_curl(YArgs a, int parts, hipStream_t st) {
    const size_t sm = ((size_t)2 * a.njp * TRP * 16 + 2 * a.njp) * sizeof(R);
    a.parts = parts;
    a.ncg = (a.NQ + 8 * W - 1) / (8 * W);
    if (sm > 64 * 1024) { // above the default dynamic LDS limit of a launch (the CU has 160 KB)
        static bool raised = false;
        if (!raised) {
            (void)hipFuncSetAttribute(reinterpret_cast<const void *>(&k_ybwd_curl<R, TRP, W>),
                                      hipFuncAttributeMaxDynamicSharedMemorySize, (int)sm);
            raised = true;
        }
    }
    hipLaunchKernelGGL((k_ybwd_curl<R, TRP, W>), dim3(a.parts * a.ncg), dim3(64 * W), sm, st, a);
}

bool launch_ybwd_curl(YArgs a, bool fp32, hipStream_t st) {
    y_shape(a);
    const size_t esz = fp32 ? 4 : 8;
    if (!a.coefd || a.nj > a.NY2 || (size_t)2 * a.njp * esz * 17 > (size_t)kYCurlLds) return false;
    const int TR = (a.rows_m + 15) / 16;
    // FP32 columns longer than 128 entries: two row tiles per wavefront (Re_tau=395, 200-entry
    // columns: 243 against 287 us with one); otherwise one (Re_tau=180: 49 against 53 us)
    if (fp32 && a.njp > 128 && (size_t)2 * a.njp * esz * 33 <= (size_t)kYCurlLds) {
        go_ybwd_curl<float, 2, kYCurlWaves>(a, (TR + 1) / 2, st);
        return true;
    }
    if (fp32) go_ybwd_curl<float, 1, kYCurlWaves>(a, TR, st);
    else go_ybwd_curl<double, 1, kYCurlWaves>(a, TR, st);
    return true;
}

// compute units of the device current at the first call (256 if the query fails); the static's
// initialisation runs once, also when several host threads launch
static int device_cus() {
    static const int cus = [] {
        int dev = 0, n = 0;
        if (hipGetDevice(&dev) == hipSuccess &&
            hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) == hipSuccess && n > 0)
            return n;
        return 256;
    }();
    return cus;
}

bool launch_yfwd_mfma(YArgs a, int nf, bool fp32, hipStream_t st) {
    y_shape(a);
    const size_t esz = fp32 ? 4 : 8;
    const int JT = (a.nje + 15) / 16;
    if ((size_t)2 * a.KL * 16 * esz + 16 * esz > (size_t)kLdsBudget) return false;
    const int fit = std::min(tiles_per_part((size_t)2 * a.KL * 16 * esz + 16 * esz, 0, 2), JT);
    const int parts = (JT + fit - 1) / fit;
    const int jtp = (JT + parts - 1) / parts;
    // a grid of fewer 8-wavefront blocks of two column sets than the device has CUs (small meshes:
    // 51 blocks at 64x65x64) runs one column set per wavefront in 2-wavefront blocks, eight times
    // the blocks; each column's matrix operations are the same either way
    if (parts * ((a.NQ + 8 * 2 * kYFwdWaves - 1) / (8 * 2 * kYFwdWaves)) * nf < device_cus()) {
        if (fp32) {
            if (jtp == 1) go_yfwd<float, 1, 1, 2>(a, parts, nf, st);
            else go_yfwd<float, 2, 1, 2>(a, parts, nf, st);
        } else {
            if (jtp == 1) go_yfwd<double, 1, 1, 2>(a, parts, nf, st);
            else go_yfwd<double, 2, 1, 2>(a, parts, nf, st);
        }
        return true;
    }
    if (fp32) {
        if (jtp == 1) go_yfwd<float, 1, 2>(a, parts, nf, st);
        else go_yfwd<float, 2, 2>(a, parts, nf, st);
    } else {
        // two column sets per wavefront: each table read feeds four MFMAs, and the Re_tau=180
        // grid runs in half a round of wavefronts instead of 1.02 rounds
        if (jtp == 1) go_yfwd<double, 1, 2>(a, parts, nf, st);
        else go_yfwd<double, 2, 2>(a, parts, nf, st);
    }
    return true;
}

} // namespace cdns
