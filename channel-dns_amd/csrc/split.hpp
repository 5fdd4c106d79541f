// Split-parity column layout: one (kx,kz) Chebyshev column per 64-lane wavefront, the even
// chain (j = 0, 2, 4, ...) in lanes 0-31 and the odd chain (j = 1, 3, ...) in lanes 32-63.
// Lane l holds NE consecutive entries of its chain: chain index k = NE (l mod 32) + e,
// coefficient j = 2k + l/32.
//
// Every Chebyshev recurrence of the solver couples only entries of one parity, so one scan over
// 32-lane halves serves both chains at once: DPP shifts inside 16-lane rows, then one carry
// between the two rows of a half (a row broadcast upward for prefix scans; for suffix scans the
// first lane of the upper row, spread over its row and moved down by a row-pair swap).
#pragma once

#include "wave.hpp"

namespace cdns {
namespace sp {

__device__ inline int parity(int lane) { return lane >> 5; }
template <int NE> __device__ inline int jof(int lane, int e) { return 2 * (NE * (lane & 31) + e) + (lane >> 5); }

// row_newbcast:0 on a 64-bit value: every lane of a 16-lane row receives the row's first lane
__device__ inline double row_first(double v) {
    const long long b = __builtin_bit_cast(long long, v);
    const long long r = __builtin_amdgcn_update_dpp(b, b, 0x150, 0xF, 0xF, false);
    return __builtin_bit_cast(double, r);
}
// value from the odd row of the lane's row pair (rows 1 -> 0, 3 -> 2; rows 1 and 3 keep their own)
__device__ inline double from_odd_row(double v) {
    const auto lo = __builtin_amdgcn_permlane16_swap(__double2loint(v), __double2loint(v), false, false);
    const auto hi = __builtin_amdgcn_permlane16_swap(__double2hiint(v), __double2hiint(v), false, false);
    return __hiloint2double(hi[1], lo[1]);
}
// even half (lanes 0-31) receives lo's upper half, odd half (32-63) receives hi's lower half
__device__ inline double cross_halves(double lo, double hi, int lane) {
    const auto l = __builtin_amdgcn_permlane32_swap(__double2loint(lo), __double2loint(hi), false, false);
    const auto h = __builtin_amdgcn_permlane32_swap(__double2hiint(lo), __double2hiint(hi), false, false);
    // output 0: lo with its upper half replaced by hi's lower half; output 1: hi with its lower
    // half replaced by lo's upper half
    const double a = __hiloint2double(h[1], l[1]), b = __hiloint2double(h[0], l[0]);
    return lane < 32 ? a : b;
}

// Chain neighbours inside each half: entry k+1 / k-1, zero beyond the chain's ends.  Cross-lane
// values are always computed by all lanes before any lane-dependent choice (a conditional
// operand would become a branch that disabled lanes do not take).
template <int NE, typename T> __device__ inline T chain_next(const T (&v)[NE], int e, int lane) {
    if (e + 1 < NE) return v[e + 1];
    const T n = lane_next(v[0], lane);
    return (lane & 31) == 31 ? T{} : n;
}
template <int NE, typename T> __device__ inline T chain_prev(const T (&v)[NE], int e, int lane) {
    if (e > 0) return v[e - 1];
    const T n = lane_prev(v[NE - 1], lane);
    return (lane & 31) == 0 ? T{} : n;
}

// Sum over each 32-lane half (one sum per chain), returned in every lane of the half.
template <int NB> __device__ inline void half_sum(double (&v)[NB]) {
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
}

// Inclusive scans over each half of affine maps y -> B[n] + A y (NB right-hand sides sharing A).
// Prefix: lane i ends with map_i o ... o map_(first lane of its half); suffix: map_i o ... o
// map_(last lane of its half).  The final A is not kept.
template <int NB> __device__ inline void half_scan_prefix(double &A, double (&B)[NB]) {
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
    // rows 1 and 3 take the totals of rows 0 and 2
#pragma unroll
    for (int n = 0; n < NB; ++n) B[n] = axpy(A, dpp_mov<dppc::row_bcast15, 0xA>(0.0, B[n]), B[n]);
}
template <int NB> __device__ inline void half_scan_suffix(double &A, double (&B)[NB], int lane) {
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
    // rows 0 and 2 take the totals of rows 1 and 3 (held by their first lanes)
    const double Am = ((lane >> 4) & 1) ? 0.0 : A;
#pragma unroll
    for (int n = 0; n < NB; ++n) B[n] = axpy(Am, from_odd_row(row_first(B[n])), B[n]);
}

// x_k = b_k[n] + a_k x_{k+1} along each chain, x beyond the top = 0.
template <int NE, int NB>
__device__ inline void suffix_affine_n(const double (&a)[NE], const double (&b)[NB][NE], double (&x)[NB][NE], int lane) {
    double A = 1.0, B[NB];
#pragma unroll
    for (int n = 0; n < NB; ++n) B[n] = 0.0;
#pragma unroll
    for (int e = NE - 1; e >= 0; --e) {
#pragma unroll
        for (int n = 0; n < NB; ++n) B[n] = axpy(a[e], B[n], b[n][e]);
        A = a[e] * A;
    }
    sp::half_scan_suffix<NB>(A, B, lane);
    // x at entry 0 is the inclusive scan; higher entries need the carry from the next lane
#pragma unroll
    for (int n = 0; n < NB; ++n) {
        const double next = lane_next(B[n], lane); // evaluated by every lane, then masked
        double carry = (lane & 31) == 31 ? 0.0 : next;
#pragma unroll
        for (int e = NE - 1; e >= 1; --e) {
            carry = axpy(a[e], carry, b[n][e]);
            x[n][e] = carry;
        }
        x[n][0] = B[n];
    }
}

// y_k = b_k[n] + a_k y_{k-1} along each chain, y before the bottom = 0.
template <int NE, int NB>
__device__ inline void prefix_affine_n(const double (&a)[NE], const double (&b)[NB][NE], double (&y)[NB][NE], int lane) {
    double A = 1.0, B[NB];
#pragma unroll
    for (int n = 0; n < NB; ++n) B[n] = 0.0;
#pragma unroll
    for (int e = 0; e < NE; ++e) {
#pragma unroll
        for (int n = 0; n < NB; ++n) B[n] = axpy(a[e], B[n], b[n][e]);
        A = a[e] * A;
    }
    sp::half_scan_prefix<NB>(A, B);
#pragma unroll
    for (int n = 0; n < NB; ++n) {
        const double prev = lane_prev(B[n], lane);
        double carry = (lane & 31) == 0 ? 0.0 : prev;
#pragma unroll
        for (int e = 0; e < NE - 1; ++e) {
            carry = axpy(a[e], carry, b[n][e]);
            y[n][e] = carry;
        }
        y[n][NE - 1] = B[n];
    }
}

// complex right-hand sides as (re, im) pairs of the real scans
template <int NE, int NR>
__device__ inline void suffix_affine_c(const double (&a)[NE], const zd (&b)[NR][NE], zd (&x)[NR][NE], int lane) {
    double bb[2 * NR][NE], xx[2 * NR][NE];
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            bb[2 * r][e] = b[r][e].x;
            bb[2 * r + 1][e] = b[r][e].y;
        }
    sp::suffix_affine_n<NE, 2 * NR>(a, bb, xx, lane);
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
        for (int e = 0; e < NE; ++e) x[r][e] = {xx[2 * r][e], xx[2 * r + 1][e]};
}
template <int NE, int NR>
__device__ inline void prefix_affine_c(const double (&a)[NE], const zd (&b)[NR][NE], zd (&y)[NR][NE], int lane) {
    double bb[2 * NR][NE], yy[2 * NR][NE];
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            bb[2 * r][e] = b[r][e].x;
            bb[2 * r + 1][e] = b[r][e].y;
        }
    sp::prefix_affine_n<NE, 2 * NR>(a, bb, yy, lane);
#pragma unroll
    for (int r = 0; r < NR; ++r)
#pragma unroll
        for (int e = 0; e < NE; ++e) y[r][e] = {yy[2 * r][e], yy[2 * r + 1][e]};
}

// Inclusive suffix sums along each chain (one complex chain per half).
template <int NE> __device__ inline void suffix_sum(const zd (&b)[NE], zd (&x)[NE], int lane) {
    zd s = {0.0, 0.0};
#pragma unroll
    for (int e = NE - 1; e >= 0; --e) s = s + b[e];
    double B[2] = {s.x, s.y};
#pragma unroll
    for (int D = 1; D <= 8; D *= 2) {
#pragma unroll
        for (int n = 0; n < 2; ++n) {
            double t;
            if (D == 1) t = dpp_zero<dppc::row_shl(1)>(B[n]);
            else if (D == 2) t = dpp_zero<dppc::row_shl(2)>(B[n]);
            else if (D == 4) t = dpp_zero<dppc::row_shl(4)>(B[n]);
            else t = dpp_zero<dppc::row_shl(8)>(B[n]);
            B[n] += t;
        }
    }
    const bool lower_row = !((lane >> 4) & 1);
#pragma unroll
    for (int n = 0; n < 2; ++n) {
        const double t = from_odd_row(row_first(B[n]));
        B[n] += lower_row ? t : 0.0;
    }
    const zd next = {lane_next(B[0], lane), lane_next(B[1], lane)};
    zd carry = (lane & 31) == 31 ? zd{0.0, 0.0} : next;
#pragma unroll
    for (int e = NE - 1; e >= 1; --e) {
        carry = carry + b[e];
        x[e] = carry;
    }
    x[0] = {B[0], B[1]};
}

// Chebyshev derivative coefficients d = D f (vector_ops.cpp:17-25):
// d_j = sum_{k>j, k-j odd} 2k f_k, then d_0 /= 2.  f zero beyond ny.
// Even j take the odd chain's suffix sums at the same chain index, odd j = 2k+1 the even
// chain's at k+1: one suffix sum per half, then the halves exchange.
template <int NE> __device__ inline void cheb_dy(const zd (&f)[NE], zd (&d)[NE], int lane, int ny) {
    zd g[NE], s[NE];
#pragma unroll
    for (int e = 0; e < NE; ++e) g[e] = (double)(2 * jof<NE>(lane, e)) * f[e];
    sp::suffix_sum<NE>(g, s, lane);
#pragma unroll
    for (int e = 0; e < NE; ++e) {
        const zd sn = sp::chain_next<NE>(s, e, lane);
        const zd v = {cross_halves(s[e].x, sn.x, lane), cross_halves(s[e].y, sn.y, lane)};
        // lanes 0-31 now hold the odd chain's sums, lanes 32-63 the even chain's shifted by one
        d[e] = jof<NE>(lane, e) <= ny ? v : zd{0.0, 0.0};
    }
    if (lane == 0) d[0] = 0.5 * d[0];
}

// Loads are unconditional (index clamped to ny, the value then zeroed beyond ny).
template <int NE> __device__ inline void load_col(const zd *__restrict__ col, int lane, int ny, zd (&v)[NE]) {
#pragma unroll
    for (int e = 0; e < NE; ++e) {
        const int j = jof<NE>(lane, e);
        const zd t = col[j <= ny ? j : ny];
        v[e] = j <= ny ? t : zd{0.0, 0.0};
    }
}
// The same from a single-precision column, widened.
template <int NE> __device__ inline void load_col(const cpx<float> *__restrict__ col, int lane, int ny, zd (&v)[NE]) {
#pragma unroll
    for (int e = 0; e < NE; ++e) {
        const int j = jof<NE>(lane, e);
        const cpx<float> t = col[j <= ny ? j : ny];
        v[e] = j <= ny ? zd{double(t.x), double(t.y)} : zd{0.0, 0.0};
    }
}
template <int NE> __device__ inline void load_rcol(const double *__restrict__ col, int lane, int ny, double (&v)[NE]) {
#pragma unroll
    for (int e = 0; e < NE; ++e) {
        const int j = jof<NE>(lane, e);
        const double t = col[j <= ny ? j : ny];
        v[e] = j <= ny ? t : 0.0;
    }
}
template <int NE> __device__ inline void store_col(zd *col, int lane, int ny, const zd (&v)[NE]) {
#pragma unroll
    for (int e = 0; e < NE; ++e) {
        const int j = jof<NE>(lane, e);
        if (j <= ny) col[j] = v[e];
    }
}

// Value of coefficient j, broadcast to all lanes.
template <int NE, typename T> __device__ inline T col_at(const T (&v)[NE], int j) {
    const int k = j >> 1, lane = 32 * (j & 1) + k / NE, e = k - (k / NE) * NE;
    T mine = T{};
#pragma unroll
    for (int ee = 0; ee < NE; ++ee)
        if (ee == e) mine = v[ee];
    return lane_read(mine, lane);
}

} // namespace sp
} // namespace cdns
