// Wavefront-level column algebra: one (kx,kz) Chebyshev column per 64-lane wavefront.
//
// Lane i holds 2*CH consecutive coefficients j = 2*CH*i + 2*c + p (slot c, parity p), so each
// parity chain (j even / j odd, the two decoupled halves of every Chebyshev recurrence) has
// CH consecutive chain entries per lane.  Sequential recurrences of the reference become
// affine scans across the chain (log2(64) = 6 shuffle steps).
#pragma once

#include "common.hpp"

namespace cdns {

__device__ inline int lane_id() { return threadIdx.x & 63; }

__device__ inline double shdown(double v, int d) { return __shfl_down(v, d, 64); }
__device__ inline double shup(double v, int d) { return __shfl_up(v, d, 64); }
__device__ inline double shidx(double v, int s) { return __shfl(v, s, 64); }
__device__ inline zd shdown(zd v, int d) { return {shdown(v.x, d), shdown(v.y, d)}; }
__device__ inline zd shup(zd v, int d) { return {shup(v.x, d), shup(v.y, d)}; }
__device__ inline zd shidx(zd v, int s) { return {shidx(v.x, s), shidx(v.y, s)}; }

__device__ inline double wave_sum(double v) {
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
    return v;
}
__device__ inline zd wave_sum(zd v) { return {wave_sum(v.x), wave_sum(v.y)}; }

template <int CH> __device__ inline int jof(int lane, int c, int p) { return 2 * CH * lane + 2 * c + p; }

// Chain neighbours (same parity): value at chain index k+1 / k-1, zero beyond the ends.
template <int CH, typename T> __device__ inline T chain_next(const T (&v)[CH], int c, int lane) {
    if (c + 1 < CH) return v[c + 1];
    T o = shdown(v[0], 1);
    if (lane == 63) o = T{};
    return o;
}
template <int CH, typename T> __device__ inline T chain_prev(const T (&v)[CH], int c, int lane) {
    if (c > 0) return v[c - 1];
    T o = shup(v[CH - 1], 1);
    if (lane == 0) o = T{};
    return o;
}

__device__ inline zd axpy(double a, zd x, zd b) { return {b.x + a * x.x, b.y + a * x.y}; }
__device__ inline double axpy(double a, double x, double b) { return b + a * x; }

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
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const double A2 = shdown(A, d);
        const T B2 = shdown(B, d);
        if (lane + d < 64) {
            B = axpy(A, B2, B);
            A = A * A2;
        }
    }
    T carry = shdown(B, 1);
    if (lane == 63) carry = T{};
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
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const double A2 = shup(A, d);
        const T B2 = shup(B, d);
        if (lane >= d) {
            B = axpy(A, B2, B);
            A = A * A2;
        }
    }
    T carry = shup(B, 1);
    if (lane == 0) carry = T{};
#pragma unroll
    for (int c = 0; c < CH; ++c) {
        carry = axpy(a[c], carry, b[c]);
        y[c] = carry;
    }
}

// Inclusive suffix sum along one chain.
template <int CH> __device__ inline void suffix_sum(const zd (&b)[CH], zd (&x)[CH], int lane) {
    zd B = {0.0, 0.0};
#pragma unroll
    for (int c = CH - 1; c >= 0; --c) B = B + b[c];
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const zd B2 = shdown(B, d);
        if (lane + d < 64) B = B + B2;
    }
    zd carry = shdown(B, 1);
    if (lane == 63) carry = zd{0.0, 0.0};
#pragma unroll
    for (int c = CH - 1; c >= 0; --c) {
        carry = carry + b[c];
        x[c] = carry;
    }
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
    return shidx(mine, lane);
}
template <int CH> __device__ inline double col_at(const double (&v)[2][CH], int j) {
    const int lane = j / (2 * CH), r = j - lane * 2 * CH, c = r >> 1, p = r & 1;
    double mine = 0.0;
#pragma unroll
    for (int cc = 0; cc < CH; ++cc)
#pragma unroll
        for (int pp = 0; pp < 2; ++pp)
            if (cc == c && pp == p) mine = v[pp][cc];
    return shidx(mine, lane);
}

} // namespace cdns
