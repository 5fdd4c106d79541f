// Wavefront-level column algebra: one (kx,kz) Chebyshev column per 64-lane wavefront.
//
// Lane i holds 2*CH consecutive coefficients j = 2*CH*i + 2*c + p (slot c, parity p), so each
// parity chain (j even / j odd, the two decoupled halves of every Chebyshev recurrence) has
// CH consecutive chain entries per lane.  Sequential recurrences of the reference become
// affine scans across the chain (DPP row shifts within 16-lane rows, row totals carried by
// v_readlane).
#pragma once

#include "common.hpp"

namespace cdns {

__device__ inline int lane_id() { return threadIdx.x & 63; }

// Cross-lane moves without the LDS crossbar: DPP row shifts (within 16-lane rows) and
// v_readlane for values at a wave-uniform lane.  GFX9 dpp_ctrl encodings.
namespace dppc {
constexpr int row_shl(int n) { return 0x100 + n; }
constexpr int row_shr(int n) { return 0x110 + n; }
} // namespace dppc

// lanes without a source lane (shifted past their row) receive `old`
template <int CTRL> __device__ inline double dpp_mov(double old, double v) {
    const int lo = __builtin_amdgcn_update_dpp(__double2loint(old), __double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(__double2hiint(old), __double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
template <int CTRL> __device__ inline zd dpp_mov(zd old, zd v) {
    return {dpp_mov<CTRL>(old.x, v.x), dpp_mov<CTRL>(old.y, v.y)};
}
// value of v at lane l (l wave-uniform)
__device__ inline double lane_read(double v, int l) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), l);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), l);
    return __hiloint2double(hi, lo);
}
__device__ inline zd lane_read(zd v, int l) { return {lane_read(v.x, l), lane_read(v.y, l)}; }

// value at lane - 1 (zero at lane 0) and at lane + 1 (zero at lane 63)
template <typename T> __device__ inline T lane_prev(T v, int lane) {
    T o = dpp_mov<dppc::row_shr(1)>(T{}, v);
    const T r15 = lane_read(v, 15), r31 = lane_read(v, 31), r47 = lane_read(v, 47);
    if (lane == 16) o = r15;
    if (lane == 32) o = r31;
    if (lane == 48) o = r47;
    return o;
}
template <typename T> __device__ inline T lane_next(T v, int lane) {
    T o = dpp_mov<dppc::row_shl(1)>(T{}, v);
    const T r16 = lane_read(v, 16), r32 = lane_read(v, 32), r48 = lane_read(v, 48);
    if (lane == 15) o = r16;
    if (lane == 31) o = r32;
    if (lane == 47) o = r48;
    return o;
}

__device__ inline double wave_sum(double v) {
    // prefix sums within rows: lane 15 of each row holds the row's sum
    v += dpp_mov<dppc::row_shr(1)>(0.0, v);
    v += dpp_mov<dppc::row_shr(2)>(0.0, v);
    v += dpp_mov<dppc::row_shr(4)>(0.0, v);
    v += dpp_mov<dppc::row_shr(8)>(0.0, v);
    return (lane_read(v, 15) + lane_read(v, 31)) + (lane_read(v, 47) + lane_read(v, 63));
}
__device__ inline zd wave_sum(zd v) { return {wave_sum(v.x), wave_sum(v.y)}; }

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

// Inclusive scans across the 64 lanes of affine maps y -> B + A y.  Prefix: lane i ends with
// map_i o ... o map_0; suffix: map_i o ... o map_63.  Hillis-Steele with DPP row shifts inside
// each 16-lane row, then the row totals (read at wave-uniform lanes) carried across rows.
template <typename T> __device__ inline void lane_scan_prefix(double &A, T &B, int lane) {
#define CDNS_SCAN_STEP(D)                                                                                        \
    {                                                                                                            \
        const double A2 = dpp_mov<dppc::row_shr(D)>(1.0, A);                                                     \
        const T B2 = dpp_mov<dppc::row_shr(D)>(T{}, B);                                                          \
        B = axpy(A, B2, B);                                                                                      \
        A = A * A2;                                                                                              \
    }
    CDNS_SCAN_STEP(1)
    CDNS_SCAN_STEP(2)
    CDNS_SCAN_STEP(4)
    CDNS_SCAN_STEP(8)
#undef CDNS_SCAN_STEP
    const double A0 = lane_read(A, 15), A1 = lane_read(A, 31), A2 = lane_read(A, 47);
    const T B0 = lane_read(B, 15), B1 = lane_read(B, 31), B2 = lane_read(B, 47);
    const int row = lane >> 4;
    double Ac = 1.0;
    T Bc = T{};
    if (row >= 1) {
        Ac = A0;
        Bc = B0;
    }
    if (row >= 2) {
        Bc = axpy(A1, Bc, B1);
        Ac = A1 * Ac;
    }
    if (row >= 3) {
        Bc = axpy(A2, Bc, B2);
        Ac = A2 * Ac;
    }
    B = axpy(A, Bc, B);
    A = A * Ac;
}
template <typename T> __device__ inline void lane_scan_suffix(double &A, T &B, int lane) {
#define CDNS_SCAN_STEP(D)                                                                                        \
    {                                                                                                            \
        const double A2 = dpp_mov<dppc::row_shl(D)>(1.0, A);                                                     \
        const T B2 = dpp_mov<dppc::row_shl(D)>(T{}, B);                                                          \
        B = axpy(A, B2, B);                                                                                      \
        A = A * A2;                                                                                              \
    }
    CDNS_SCAN_STEP(1)
    CDNS_SCAN_STEP(2)
    CDNS_SCAN_STEP(4)
    CDNS_SCAN_STEP(8)
#undef CDNS_SCAN_STEP
    const double A1 = lane_read(A, 16), A2 = lane_read(A, 32), A3 = lane_read(A, 48);
    const T B1 = lane_read(B, 16), B2 = lane_read(B, 32), B3 = lane_read(B, 48);
    const int row = lane >> 4;
    double Ac = 1.0;
    T Bc = T{};
    if (row <= 2) {
        Ac = A3;
        Bc = B3;
    }
    if (row <= 1) {
        Bc = axpy(A2, Bc, B2);
        Ac = A2 * Ac;
    }
    if (row == 0) {
        Bc = axpy(A1, Bc, B1);
        Ac = A1 * Ac;
    }
    B = axpy(A, Bc, B);
    A = A * Ac;
}

// x_k = b_k + a_k x_{k+1}, x beyond the top = 0 (suffix affine scan along one chain).
template <int CH, typename T>
__device__ inline void suffix_affine(const double (&a)[CH], const T (&b)[CH], T (&x)[CH], int lane) {
    double A = 1.0;
    T B = T{};
#pragma unroll
    for (int c = CH - 1; c >= 0; --c) {
        B = axpy(a[c], B, b[c]);
        A = a[c] * A;
    }
    lane_scan_suffix(A, B, lane);
    T carry = lane_next(B, lane);
#pragma unroll
    for (int c = CH - 1; c >= 0; --c) {
        carry = axpy(a[c], carry, b[c]);
        x[c] = carry;
    }
}

// y_k = b_k + a_k y_{k-1}, y before the bottom = 0 (prefix affine scan along one chain).
template <int CH, typename T>
__device__ inline void prefix_affine(const double (&a)[CH], const T (&b)[CH], T (&y)[CH], int lane) {
    double A = 1.0;
    T B = T{};
#pragma unroll
    for (int c = 0; c < CH; ++c) {
        B = axpy(a[c], B, b[c]);
        A = a[c] * A;
    }
    lane_scan_prefix(A, B, lane);
    T carry = lane_prev(B, lane);
#pragma unroll
    for (int c = 0; c < CH; ++c) {
        carry = axpy(a[c], carry, b[c]);
        y[c] = carry;
    }
}

// Inclusive suffix sum along one chain.
template <int CH> __device__ inline void suffix_sum(const zd (&b)[CH], zd (&x)[CH], int lane) {
    double a[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) a[c] = 1.0;
    suffix_affine<CH, zd>(a, b, x, lane);
}

// Chebyshev derivative coefficients d = D f (vector_ops.cpp:17-25):
// d_j = sum_{k>j, k-j odd} 2k f_k, then d_0 /= 2.  f zero beyond ny.
template <int CH> __device__ inline void cheb_dy(const zd (&f)[2][CH], zd (&d)[2][CH], int lane, int ny) {
    zd g[2][CH];
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            const int j = jof<CH>(lane, c, p);
            g[p][c] = (double)(2 * j) * f[p][c];
        }
    zd so[CH], se[CH];
    suffix_sum<CH>(g[1], so, lane); // odd k >= 2i+1  -> d at even j = 2i
    suffix_sum<CH>(g[0], se, lane); // even k >= 2i   -> d at odd j = 2i+1 needs k >= 2i+2
#pragma unroll
    for (int c = 0; c < CH; ++c) {
        const int j0 = jof<CH>(lane, c, 0), j1 = j0 + 1;
        d[0][c] = j0 <= ny ? so[c] : zd{0.0, 0.0};
        const zd nx = chain_next<CH>(se, c, lane);
        d[1][c] = j1 <= ny ? nx : zd{0.0, 0.0};
    }
    if (lane == 0) d[0][0] = 0.5 * d[0][0];
}

template <int CH> __device__ inline void load_col(const zd *__restrict__ col, int lane, int ny, zd (&v)[2][CH]) {
#pragma unroll
    for (int c = 0; c < CH; ++c)
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            const int j = jof<CH>(lane, c, p);
            v[p][c] = j <= ny ? col[j] : zd{0.0, 0.0};
        }
}
template <int CH> __device__ inline void load_rcol(const double *__restrict__ col, int lane, int ny, double (&v)[2][CH]) {
#pragma unroll
    for (int c = 0; c < CH; ++c)
#pragma unroll
        for (int p = 0; p < 2; ++p) {
            const int j = jof<CH>(lane, c, p);
            v[p][c] = j <= ny ? col[j] : 0.0;
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
