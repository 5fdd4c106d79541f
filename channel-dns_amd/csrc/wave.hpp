// Wavefront-level column algebra: one (kx,kz) Chebyshev column per 64-lane wavefront.
//
// Lane i holds 2*CH consecutive coefficients j = 2*CH*i + 2*c + p (slot c, parity p), so each
// parity chain (j even / j odd, the two decoupled halves of every Chebyshev recurrence) has
// CH consecutive chain entries per lane.  Sequential recurrences of the reference become
// affine scans across the chain, carried out with DPP lane moves and v_readlane only.
#pragma once

#include "common.hpp"

namespace cdns {

__device__ inline int lane_id() { return threadIdx.x & 63; }

// Cross-lane moves without the LDS crossbar (GFX9 dpp_ctrl encodings, gfx950 permlane swaps):
// row shifts/rotations within 16-lane rows, whole-wave shifts by one lane, row broadcasts
// (lane 15 of a row into the next row, lane 31 into rows 2 and 3), and v_readlane for values at
// a wave-uniform lane.
namespace dppc {
constexpr int row_shl(int n) { return 0x100 + n; }
constexpr int row_shr(int n) { return 0x110 + n; }
constexpr int row_ror(int n) { return 0x120 + n; }
constexpr int wave_shl1 = 0x130; // lane i <- lane i + 1
constexpr int wave_shr1 = 0x138; // lane i <- lane i - 1
constexpr int row_bcast15 = 0x142;
constexpr int row_bcast31 = 0x143;
} // namespace dppc

// Cross-lane moves are convergent: the compiler keeps them where the source puts them.  A call
// written as an operand of ?: becomes a branch, though, so callers compute every cross-lane value
// into a variable first (lanes that skip a branch are disabled sources and read as zero).

// lanes without a source lane, and lanes of rows outside ROWS, receive `old`
template <int CTRL, int ROWS = 0xF> __device__ inline double dpp_mov(double old, double v) {
    const int lo = __builtin_amdgcn_update_dpp(__double2loint(old), __double2loint(v), CTRL, ROWS, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(__double2hiint(old), __double2hiint(v), CTRL, ROWS, 0xF, false);
    return __hiloint2double(hi, lo);
}
template <int CTRL, int ROWS = 0xF> __device__ inline zd dpp_mov(zd old, zd v) {
    return {dpp_mov<CTRL, ROWS>(old.x, v.x), dpp_mov<CTRL, ROWS>(old.y, v.y)};
}
// lanes without a source lane receive zero (bound_ctrl)
template <int CTRL> __device__ inline double dpp_zero(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, true);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, true);
    return __hiloint2double(hi, lo);
}
template <int CTRL> __device__ inline zd dpp_zero(zd v) { return {dpp_zero<CTRL>(v.x), dpp_zero<CTRL>(v.y)}; }

// value of v at lane l (l wave-uniform)
__device__ inline double lane_read(double v, int l) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}
__device__ inline zd lane_read(zd v, int l) { return {lane_read(v.x, l), lane_read(v.y, l)}; }

// value at lane - 1 (zero at lane 0) and at lane + 1 (zero at lane 63)
template <typename T> __device__ inline T lane_prev(T v, int) { return dpp_zero<dppc::wave_shr1>(v); }
template <typename T> __device__ inline T lane_next(T v, int) { return dpp_zero<dppc::wave_shl1>(v); }

// v(row 2m) + v(row 2m+1) and v(lanes 0-31) + v(lanes 32-63), the same sum in every lane
__device__ inline double pair_rows_sum(double v) {
    const auto lo = __builtin_amdgcn_permlane16_swap(__double2loint(v), __double2loint(v), false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap(__double2hiint(v), __double2hiint(v), false, false);
    return __hiloint2double(hi[0], lo[0]) + __hiloint2double(hi[1], lo[1]);
}
__device__ inline double halves_sum(double v) {
    const auto lo = __builtin_amdgcn_permlane32_swap(__double2loint(v), __double2loint(v), false, false);
    const auto hi = __builtin_amdgcn_permlane32_swap(__double2hiint(v), __double2hiint(v), false, false);
    return __hiloint2double(hi[0], lo[0]) + __hiloint2double(hi[1], lo[1]);
}

// Sum over the 64 lanes, returned in every lane (butterfly: row rotations, then the row-pair and
// half-wave swaps; every lane adds the same operands in the same order).
template <int NB> __device__ inline void wave_sum(double (&v)[NB]) {
#pragma unroll
    for (int n = 0; n < NB; ++n) v[n] += dpp_zero<dppc::row_ror(8)>(v[n]);
#pragma unroll
    for (int n = 0; n < NB; ++n) v[n] += dpp_zero<dppc::row_ror(4)>(v[n]);
#pragma unroll
    for (int n = 0; n < NB; ++n) v[n] += dpp_zero<dppc::row_ror(2)>(v[n]);
#pragma unroll
    for (int n = 0; n < NB; ++n) v[n] += dpp_zero<dppc::row_ror(1)>(v[n]);
#pragma unroll
    for (int n = 0; n < NB; ++n) v[n] = pair_rows_sum(v[n]);
#pragma unroll
    for (int n = 0; n < NB; ++n) v[n] = halves_sum(v[n]);
}
__device__ inline double wave_sum(double v) {
    double t[1] = {v};
    wave_sum<1>(t);
    return t[0];
}
__device__ inline zd wave_sum(zd v) {
    double t[2] = {v.x, v.y};
    wave_sum<2>(t);
    return {t[0], t[1]};
}

template <int CH> __device__ inline int jof(int lane, int c, int p) { return 2 * CH * lane + 2 * c + p; }

// Chain neighbours (same parity): value at chain index k+1 / k-1, zero beyond the ends.
template <int CH, typename T> __device__ inline T chain_next(const T (&v)[CH], int c, int lane) {
    if (c + 1 < CH) return v[c + 1];
    return lane_next(v[0], lane);
}
template <int CH, typename T> __device__ inline T chain_prev(const T (&v)[CH], int c, int lane) {
    if (c > 0) return v[c - 1];
    return lane_prev(v[CH - 1], lane);
}

__device__ inline zd axpy(double a, zd x, zd b) { return {b.x + a * x.x, b.y + a * x.y}; }
__device__ inline double axpy(double a, double x, double b) { return b + a * x; }

// Inclusive scans across the 64 lanes of affine maps y -> B[n] + A y, NB right-hand sides sharing
// the multipliers A.  Prefix: lane i ends with map_i o ... o map_0; suffix: map_i o ... o map_63.
// Hillis-Steele with DPP row shifts inside each 16-lane row (shifted-in B are zero, A one), then
// the row totals carried across rows: by row broadcasts for the prefix, by reading the first lane
// of each higher row for the suffix (DPP has no downward broadcast).  The final A is not kept.
template <int NB> __device__ inline void lane_scan_prefix(double &A, double (&B)[NB]) {
#define CDNS_SCAN_STEP(D)                                                                                        \
    {                                                                                                            \
        const double A2 = dpp_mov<dppc::row_shr(D)>(1.0, A);                                                     \
        _Pragma("unroll") for (int n = 0; n < NB; ++n) B[n] = axpy(A, dpp_zero<dppc::row_shr(D)>(B[n]), B[n]);   \
        A = A * A2;                                                                                              \
    }
    CDNS_SCAN_STEP(1)
    CDNS_SCAN_STEP(2)
    CDNS_SCAN_STEP(4)
    CDNS_SCAN_STEP(8)
#undef CDNS_SCAN_STEP
    {   // rows 1 and 3 take row 0 / row 2
        const double A2 = dpp_mov<dppc::row_bcast15, 0xA>(1.0, A);
#pragma unroll
        for (int n = 0; n < NB; ++n) B[n] = axpy(A, dpp_mov<dppc::row_bcast15, 0xA>(0.0, B[n]), B[n]);
        A = A * A2;
    }
    {   // rows 2 and 3 take rows 0-1
#pragma unroll
        for (int n = 0; n < NB; ++n) B[n] = axpy(A, dpp_mov<dppc::row_bcast31, 0xC>(0.0, B[n]), B[n]);
    }
}
template <int NB> __device__ inline void lane_scan_suffix(double &A, double (&B)[NB], int lane) {
#define CDNS_SCAN_STEP(D)                                                                                        \
    {                                                                                                            \
        const double A2 = dpp_mov<dppc::row_shl(D)>(1.0, A);                                                     \
        _Pragma("unroll") for (int n = 0; n < NB; ++n) B[n] = axpy(A, dpp_zero<dppc::row_shl(D)>(B[n]), B[n]);   \
        A = A * A2;                                                                                              \
    }
    CDNS_SCAN_STEP(1)
    CDNS_SCAN_STEP(2)
    CDNS_SCAN_STEP(4)
    CDNS_SCAN_STEP(8)
#undef CDNS_SCAN_STEP
    // carries into rows 2, 1, 0: C2 = T3, C1 = T2 o T3, C0 = T1 o C1 (T_r = row r total at lane 16 r)
    const double A1 = lane_read(A, 16), A2 = lane_read(A, 32), A3 = lane_read(A, 48);
    const double C1A = A2 * A3;
    const int row = lane >> 4;
    const double Ac = row == 0 ? A1 * C1A : (row == 1 ? C1A : (row == 2 ? A3 : 0.0));
#pragma unroll
    for (int n = 0; n < NB; ++n) {
        const double B1 = lane_read(B[n], 16), B2 = lane_read(B[n], 32), B3 = lane_read(B[n], 48);
        const double C1 = axpy(A2, B3, B2), C0 = axpy(A1, C1, B1);
        const double Bc = row == 0 ? C0 : (row == 1 ? C1 : (row == 2 ? B3 : 0.0));
        B[n] = axpy(A, Bc, B[n]);
    }
}

// Inclusive suffix sums across the 64 lanes (NB independent sums).
template <int NB> __device__ inline void lane_sum_suffix(double (&B)[NB], int lane) {
#pragma unroll
    for (int D = 1; D <= 8; D *= 2) {
#pragma unroll
        for (int n = 0; n < NB; ++n) {
            double t;
            if (D == 1) t = dpp_zero<dppc::row_shl(1)>(B[n]);
            else if (D == 2) t = dpp_zero<dppc::row_shl(2)>(B[n]);
            else if (D == 4) t = dpp_zero<dppc::row_shl(4)>(B[n]);
            else t = dpp_zero<dppc::row_shl(8)>(B[n]);
            B[n] += t;
        }
    }
    const int row = lane >> 4;
#pragma unroll
    for (int n = 0; n < NB; ++n) {
        const double B1 = lane_read(B[n], 16), B2 = lane_read(B[n], 32), B3 = lane_read(B[n], 48);
        const double C1 = B2 + B3, C0 = B1 + C1;
        B[n] += row == 0 ? C0 : (row == 1 ? C1 : (row == 2 ? B3 : 0.0));
    }
}

// x_k = b_k[n] + a_k x_{k+1}, x beyond the top = 0 (suffix affine scans along one chain, NB
// right-hand sides sharing a).
template <int CH, int NB>
__device__ inline void suffix_affine_n(const double (&a)[CH], const double (&b)[NB][CH], double (&x)[NB][CH], int lane) {
    double A = 1.0, B[NB];
#pragma unroll
    for (int n = 0; n < NB; ++n) B[n] = 0.0;
#pragma unroll
    for (int c = CH - 1; c >= 0; --c) {
#pragma unroll
        for (int n = 0; n < NB; ++n) B[n] = axpy(a[c], B[n], b[n][c]);
        A = a[c] * A;
    }
    lane_scan_suffix<NB>(A, B, lane);
    if constexpr (CH == 1) {
#pragma unroll
        for (int n = 0; n < NB; ++n) x[n][0] = B[n];
    } else {
#pragma unroll
        for (int n = 0; n < NB; ++n) {
            double carry = lane_next(B[n], lane);
#pragma unroll
            for (int c = CH - 1; c >= 0; --c) {
                carry = axpy(a[c], carry, b[n][c]);
                x[n][c] = carry;
            }
        }
    }
}

// y_k = b_k[n] + a_k y_{k-1}, y before the bottom = 0 (prefix affine scans along one chain).
template <int CH, int NB>
__device__ inline void prefix_affine_n(const double (&a)[CH], const double (&b)[NB][CH], double (&y)[NB][CH], int lane) {
    double A = 1.0, B[NB];
#pragma unroll
    for (int n = 0; n < NB; ++n) B[n] = 0.0;
#pragma unroll
    for (int c = 0; c < CH; ++c) {
#pragma unroll
        for (int n = 0; n < NB; ++n) B[n] = axpy(a[c], B[n], b[n][c]);
        A = a[c] * A;
    }
    lane_scan_prefix<NB>(A, B);
    if constexpr (CH == 1) {
#pragma unroll
        for (int n = 0; n < NB; ++n) y[n][0] = B[n];
    } else {
#pragma unroll
        for (int n = 0; n < NB; ++n) {
            double carry = lane_prev(B[n], lane);
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                carry = axpy(a[c], carry, b[n][c]);
                y[n][c] = carry;
            }
        }
    }
}

// complex right-hand sides as (re, im) pairs of the real scans
template <int CH, int NR>
__device__ inline void suffix_affine_c(const double (&a)[CH], const zd (&b)[NR][CH], zd (&x)[NR][CH], int lane) {
    double bb[2 * NR][CH], xx[2 * NR][CH];
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            bb[2 * r][c] = b[r][c].x;
            bb[2 * r + 1][c] = b[r][c].y;
        }
    suffix_affine_n<CH, 2 * NR>(a, bb, xx, lane);
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
        for (int c = 0; c < CH; ++c) x[r][c] = {xx[2 * r][c], xx[2 * r + 1][c]};
}
template <int CH, int NR>
__device__ inline void prefix_affine_c(const double (&a)[CH], const zd (&b)[NR][CH], zd (&y)[NR][CH], int lane) {
    double bb[2 * NR][CH], yy[2 * NR][CH];
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            bb[2 * r][c] = b[r][c].x;
            bb[2 * r + 1][c] = b[r][c].y;
        }
    prefix_affine_n<CH, 2 * NR>(a, bb, yy, lane);
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
        for (int c = 0; c < CH; ++c) y[r][c] = {yy[2 * r][c], yy[2 * r + 1][c]};
}

// single right-hand side forms
template <int CH, typename T>
__device__ inline void suffix_affine(const double (&a)[CH], const T (&b)[CH], T (&x)[CH], int lane) {
    if constexpr (sizeof(T) == sizeof(double)) {
        double bb[1][CH], xx[1][CH];
#pragma unroll
        for (int c = 0; c < CH; ++c) bb[0][c] = b[c];
        suffix_affine_n<CH, 1>(a, bb, xx, lane);
#pragma unroll
        for (int c = 0; c < CH; ++c) x[c] = xx[0][c];
    } else {
        const zd(&bz)[1][CH] = reinterpret_cast<const zd(&)[1][CH]>(b);
        zd(&xz)[1][CH] = reinterpret_cast<zd(&)[1][CH]>(x);
        suffix_affine_c<CH, 1>(a, bz, xz, lane);
    }
}
template <int CH, typename T>
__device__ inline void prefix_affine(const double (&a)[CH], const T (&b)[CH], T (&y)[CH], int lane) {
    if constexpr (sizeof(T) == sizeof(double)) {
        double bb[1][CH], yy[1][CH];
#pragma unroll
        for (int c = 0; c < CH; ++c) bb[0][c] = b[c];
        prefix_affine_n<CH, 1>(a, bb, yy, lane);
#pragma unroll
        for (int c = 0; c < CH; ++c) y[c] = yy[0][c];
    } else {
        const zd(&bz)[1][CH] = reinterpret_cast<const zd(&)[1][CH]>(b);
        zd(&yz)[1][CH] = reinterpret_cast<zd(&)[1][CH]>(y);
        prefix_affine_c<CH, 1>(a, bz, yz, lane);
    }
}

// Inclusive suffix sums along chains: NR complex chains at once.
template <int CH, int NR> __device__ inline void suffix_sum_c(const zd (&b)[NR][CH], zd (&x)[NR][CH], int lane) {
    double B[2 * NR];
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        zd s = {0.0, 0.0};
#pragma unroll
        for (int c = CH - 1; c >= 0; --c) s = s + b[r][c];
        B[2 * r] = s.x;
        B[2 * r + 1] = s.y;
    }
    lane_sum_suffix<2 * NR>(B, lane);
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        if constexpr (CH == 1) {
            x[r][0] = {B[2 * r], B[2 * r + 1]};
        } else {
            zd carry = {lane_next(B[2 * r], lane), lane_next(B[2 * r + 1], lane)};
#pragma unroll
            for (int c = CH - 1; c >= 0; --c) {
                carry = carry + b[r][c];
                x[r][c] = carry;
            }
        }
    }
}
template <int CH> __device__ inline void suffix_sum(const zd (&b)[CH], zd (&x)[CH], int lane) {
    const zd(&bz)[1][CH] = reinterpret_cast<const zd(&)[1][CH]>(b);
    zd(&xz)[1][CH] = reinterpret_cast<zd(&)[1][CH]>(x);
    suffix_sum_c<CH, 1>(bz, xz, lane);
}

// Chebyshev derivative coefficients d = D f (vector_ops.cpp:17-25):
// d_j = sum_{k>j, k-j odd} 2k f_k, then d_0 /= 2.  f zero beyond ny.
template <int CH> __device__ inline void cheb_dy(const zd (&f)[2][CH], zd (&d)[2][CH], int lane, int ny) {
    zd g[2][CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) {
        // g[0]: odd k (-> d at even j = 2i needs odd k >= 2i+1); g[1]: even k (-> d at odd j)
        const int j0 = jof<CH>(lane, c, 0);
        g[0][c] = (double)(2 * (j0 + 1)) * f[1][c];
        g[1][c] = (double)(2 * j0) * f[0][c];
    }
    zd s[2][CH];
    suffix_sum_c<CH, 2>(g, s, lane);
#pragma unroll
    for (int c = 0; c < CH; ++c) {
        const int j0 = jof<CH>(lane, c, 0), j1 = j0 + 1;
        d[0][c] = j0 <= ny ? s[0][c] : zd{0.0, 0.0};
        const zd nx = chain_next<CH>(s[1], c, lane); // even k >= 2i+2
        d[1][c] = j1 <= ny ? nx : zd{0.0, 0.0};
    }
    if (lane == 0) d[0][0] = 0.5 * d[0][0];
}

// Loads are unconditional (index clamped to ny, the value then zeroed beyond ny), so a wavefront
// issues all of them before its first wait.
template <int CH> __device__ inline void load_col(const zd *__restrict__ col, int lane, int ny, zd (&v)[2][CH]) {
#pragma unroll
    for (int c = 0; c < CH; ++c)
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            const int j = jof<CH>(lane, c, p);
            const zd t = col[j <= ny ? j : ny];
            v[p][c] = j <= ny ? t : zd{0.0, 0.0};
        }
}
template <int CH> __device__ inline void load_rcol(const double *__restrict__ col, int lane, int ny, double (&v)[2][CH]) {
#pragma unroll
    for (int c = 0; c < CH; ++c)
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            const int j = jof<CH>(lane, c, p);
            const double t = col[j <= ny ? j : ny];
            v[p][c] = j <= ny ? t : 0.0;
        }
}
template <int CH> __device__ inline void store_col(zd *col, int lane, int ny, const zd (&v)[2][CH]) {
#pragma unroll
    for (int c = 0; c < CH; ++c)
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            const int j = jof<CH>(lane, c, p);
            if (j <= ny) col[j] = v[p][c];
        }
}

// Value of element j of a lane-distributed column, broadcast to all lanes.
template <int CH> __device__ inline zd col_at(const zd (&v)[2][CH], int j) {
    const int lane = j / (2 * CH), r = j - lane * 2 * CH, c = r >> 1, p = r & 1;
    zd mine = {0.0, 0.0};
#pragma unroll
    for (int cc = 0; cc < CH; ++cc)
#pragma unroll
        for (int pp = 0; pp < 2; ++pp)
            if (cc == c && pp == p) mine = v[pp][cc];
    return lane_read(mine, lane);
}
template <int CH> __device__ inline double col_at(const double (&v)[2][CH], int j) {
    const int lane = j / (2 * CH), r = j - lane * 2 * CH, c = r >> 1, p = r & 1;
    double mine = 0.0;
#pragma unroll
    for (int cc = 0; cc < CH; ++cc)
#pragma unroll
        for (int pp = 0; pp < 2; ++pp)
            if (cc == c && pp == p) mine = v[pp][cc];
    return lane_read(mine, lane);
}

} // namespace cdns
