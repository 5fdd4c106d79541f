// Internal types, layout description and device helpers of the gfx950 backend.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace cdns {

// ------------------------------------------------------------- complex ----
template <typename R> struct alignas(2 * sizeof(R)) cpx {
    R x, y;
};

template <typename R> __host__ __device__ inline cpx<R> mkc(R x, R y) { return cpx<R>{x, y}; }
template <typename R> __host__ __device__ inline cpx<R> operator+(cpx<R> a, cpx<R> b) { return {a.x + b.x, a.y + b.y}; }
template <typename R> __host__ __device__ inline cpx<R> operator-(cpx<R> a, cpx<R> b) { return {a.x - b.x, a.y - b.y}; }
template <typename R> __host__ __device__ inline cpx<R> operator-(cpx<R> a) { return {-a.x, -a.y}; }
template <typename R> __host__ __device__ inline cpx<R> operator*(R s, cpx<R> a) { return {s * a.x, s * a.y}; }
template <typename R> __host__ __device__ inline cpx<R> cmul(cpx<R> a, cpx<R> b) {
    return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename R> __host__ __device__ inline cpx<R> conjc(cpx<R> a) { return {a.x, -a.y}; }
// (i k) v, written as the reference writes it (src/core/complex_ops.hpp:6-8).
template <typename R> __host__ __device__ inline cpx<R> ikmul(R k, cpx<R> v) { return {-k * v.y, k * v.x}; }
// i v and -i v
template <typename R> __host__ __device__ inline cpx<R> mul_i(cpx<R> v) { return {-v.y, v.x}; }
// Smith's complex division.
__host__ __device__ inline cpx<double> cdiv(cpx<double> a, cpx<double> b) {
    if (fabs(b.x) >= fabs(b.y)) {
        double r = b.y / b.x, den = b.x + b.y * r;
        return {(a.x + a.y * r) / den, (a.y - a.x * r) / den};
    }
    double r = b.x / b.y, den = b.x * r + b.y;
    return {(a.x * r + a.y) / den, (a.y * r - a.x) / den};
}

using zd = cpx<double>;

// -------------------------------------------------------------- geometry --
// Internal device layout.
//   spectral column q = kx*nzk + (kz+NZS), column stride JS complex; component planes.
//   y-physical (Y):  [l][q]          l in [0, NL = NY2+1)
//   z-physical (Z):  [l][iz][kx]     iz in [0, NZ2)
struct Geometry {
    int NX, NY, NZ, NXS, NYS, NZS, NX2, NY2, NZ2;
    int nkx, nzk, nj, JS, NQ, NL, NLH, E;
    size_t S;  // complex entries of one spectral component plane = NQ*JS
    size_t YS; // complex entries of one Y field = NL*NQ
    size_t ZS; // complex entries of one Z field = NL*NZ2*nkx
};

constexpr int MAXST = 12; // FFT stages

struct FftPlan {
    int n, nst, rad[MAXST];
};

// --------------------------------------------------------- device FFT ----
// Small DFTs y_m = sum_q v_q exp(sign*2*pi*i*q*m/RAD), sign = -1 forward, +1 backward.
namespace dft_const {
constexpr double C3 = -0.5, S3 = 8.66025403784438596588e-01;
constexpr double C5_1 = 3.09016994374947451263e-01, S5_1 = 9.51056516295153531182e-01;
constexpr double C5_2 = -8.09016994374947451263e-01, S5_2 = 5.87785252292473137103e-01;
constexpr double C7_1 = 6.23489801858733483364e-01, S7_1 = 7.81831482468029803634e-01;
constexpr double C7_2 = -2.22520933956314392876e-01, S7_2 = 9.74927912181823619342e-01;
constexpr double C7_3 = -9.00968867902419145999e-01, S7_3 = 4.33883739117558120402e-01;
constexpr double R2 = 7.07106781186547572737e-01;
} // namespace dft_const

template <typename R, int RAD> struct Dft;

template <typename R> struct Dft<R, 2> {
    __device__ static inline void run(cpx<R> *v, int) {
        cpx<R> a = v[0], b = v[1];
        v[0] = a + b;
        v[1] = a - b;
    }
};
template <typename R> struct Dft<R, 4> {
    __device__ static inline void run(cpx<R> *v, int sign) {
        cpx<R> a = v[0] + v[2], b = v[0] - v[2], c = v[1] + v[3], d = v[1] - v[3];
        cpx<R> id = sign > 0 ? mul_i(d) : -mul_i(d);
        v[0] = a + c;
        v[2] = a - c;
        v[1] = b + id;
        v[3] = b - id;
    }
};
template <typename R> struct Dft<R, 8> {
    __device__ static inline void run(cpx<R> *v, int sign) {
        cpx<R> e[4] = {v[0], v[2], v[4], v[6]}, o[4] = {v[1], v[3], v[5], v[7]};
        Dft<R, 4>::run(e, sign);
        Dft<R, 4>::run(o, sign);
        const R r2 = (R)dft_const::R2;
        const R s = sign > 0 ? R(1) : R(-1);
        // w^m, w = exp(sign*2*pi*i/8)
        cpx<R> t1 = cmul(o[1], mkc<R>(r2, s * r2));
        cpx<R> t2 = sign > 0 ? mul_i(o[2]) : -mul_i(o[2]);
        cpx<R> t3 = cmul(o[3], mkc<R>(-r2, s * r2));
        v[0] = e[0] + o[0];
        v[4] = e[0] - o[0];
        v[1] = e[1] + t1;
        v[5] = e[1] - t1;
        v[2] = e[2] + t2;
        v[6] = e[2] - t2;
        v[3] = e[3] + t3;
        v[7] = e[3] - t3;
    }
};
template <typename R> struct Dft<R, 3> {
    __device__ static inline void run(cpx<R> *v, int sign) {
        const R c = (R)dft_const::C3, s = (R)(sign > 0 ? dft_const::S3 : -dft_const::S3);
        cpx<R> t = v[1] + v[2], d = v[1] - v[2];
        cpx<R> u = v[0] + c * t;
        cpx<R> w = s * mul_i(d);
        v[0] = v[0] + t;
        v[1] = u + w;
        v[2] = u - w;
    }
};
template <typename R> struct Dft<R, 5> {
    __device__ static inline void run(cpx<R> *v, int sign) {
        using namespace dft_const;
        const R sg = sign > 0 ? R(1) : R(-1);
        cpx<R> t1 = v[1] + v[4], t2 = v[2] + v[3], d1 = v[1] - v[4], d2 = v[2] - v[3];
        cpx<R> x0 = v[0];
        cpx<R> a1 = x0 + (R)C5_1 * t1 + (R)C5_2 * t2;
        cpx<R> a2 = x0 + (R)C5_2 * t1 + (R)C5_1 * t2;
        cpx<R> b1 = (sg * (R)S5_1) * mul_i(d1) + (sg * (R)S5_2) * mul_i(d2);
        cpx<R> b2 = (sg * (R)S5_2) * mul_i(d1) - (sg * (R)S5_1) * mul_i(d2);
        v[0] = x0 + t1 + t2;
        v[1] = a1 + b1;
        v[4] = a1 - b1;
        v[2] = a2 + b2;
        v[3] = a2 - b2;
    }
};
template <typename R> struct Dft<R, 7> {
    __device__ static inline void run(cpx<R> *v, int sign) {
        using namespace dft_const;
        const R sg = sign > 0 ? R(1) : R(-1);
        cpx<R> t1 = v[1] + v[6], t2 = v[2] + v[5], t3 = v[3] + v[4];
        cpx<R> d1 = mul_i(v[1] - v[6]), d2 = mul_i(v[2] - v[5]), d3 = mul_i(v[3] - v[4]);
        cpx<R> x0 = v[0];
        // cos(2 pi q m / 7) and sin for m = 1, 2, 3
        cpx<R> a1 = x0 + (R)C7_1 * t1 + (R)C7_2 * t2 + (R)C7_3 * t3;
        cpx<R> a2 = x0 + (R)C7_2 * t1 + (R)C7_3 * t2 + (R)C7_1 * t3;
        cpx<R> a3 = x0 + (R)C7_3 * t1 + (R)C7_1 * t2 + (R)C7_2 * t3;
        cpx<R> b1 = (sg * (R)S7_1) * d1 + (sg * (R)S7_2) * d2 + (sg * (R)S7_3) * d3;
        cpx<R> b2 = (sg * (R)S7_2) * d1 - (sg * (R)S7_3) * d2 - (sg * (R)S7_1) * d3;
        cpx<R> b3 = (sg * (R)S7_3) * d1 - (sg * (R)S7_1) * d2 + (sg * (R)S7_2) * d3;
        v[0] = x0 + t1 + t2 + t3;
        v[1] = a1 + b1;
        v[6] = a1 - b1;
        v[2] = a2 + b2;
        v[5] = a2 - b2;
        v[3] = a3 + b3;
        v[4] = a3 - b3;
    }
};

// One Stockham stage over `nlines` contiguous lines of length N held in LDS (in place:
// every thread loads its butterflies into registers before any thread stores).
// tw[k] = exp(-2 pi i k / N).
template <typename R, int RAD, int IT>
__device__ inline void fft_stage(cpx<R> *buf, int nlines, int N, int Ns, const cpx<R> *__restrict__ tw, int sign) {
    const int nb = N / RAD, total = nlines * nb;
    cpx<R> v[IT][RAD];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int item = threadIdx.x + it * blockDim.x;
        if (item < total) {
            const int ln = item / nb, t = item - ln * nb;
            const cpx<R> *s = buf + (size_t)ln * N;
#pragma unroll
            for (int r = 0; r < RAD; ++r) v[it][r] = s[t + r * nb];
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int item = threadIdx.x + it * blockDim.x;
        if (item < total) {
            const int ln = item / nb, t = item - ln * nb, k = t % Ns;
            const int step = (N / (Ns * RAD)) * k;
#pragma unroll
            for (int r = 1; r < RAD; ++r) {
                cpx<R> w = tw[r * step];
                if (sign > 0) w.y = -w.y;
                v[it][r] = cmul(v[it][r], w);
            }
            Dft<R, RAD>::run(v[it], sign);
            const int base = (t / Ns) * Ns * RAD + k;
            cpx<R> *d = buf + (size_t)ln * N;
#pragma unroll
            for (int r = 0; r < RAD; ++r) d[base + r * Ns] = v[it][r];
        }
    }
    __syncthreads();
}

// Unnormalised DFT of `nlines` contiguous LDS lines (natural order in and out).
template <typename R, int IT>
__device__ inline void fft_block(cpx<R> *buf, int nlines, const FftPlan &p, const cpx<R> *__restrict__ tw, int sign) {
    int Ns = 1;
    for (int s = 0; s < p.nst; ++s) {
        switch (p.rad[s]) {
        case 2: fft_stage<R, 2, IT>(buf, nlines, p.n, Ns, tw, sign); break;
        case 3: fft_stage<R, 3, IT>(buf, nlines, p.n, Ns, tw, sign); break;
        case 4: fft_stage<R, 4, IT>(buf, nlines, p.n, Ns, tw, sign); break;
        case 5: fft_stage<R, 5, IT>(buf, nlines, p.n, Ns, tw, sign); break;
        case 7: fft_stage<R, 7, IT>(buf, nlines, p.n, Ns, tw, sign); break;
        default: fft_stage<R, 8, IT>(buf, nlines, p.n, Ns, tw, sign); break;
        }
        Ns *= p.rad[s];
    }
}

// Smallest radix of a plan (determines butterflies per stage).
inline int plan_min_radix(const FftPlan &p) {
    int m = 64;
    for (int s = 0; s < p.nst; ++s) m = p.rad[s] < m ? p.rad[s] : m;
    return m;
}

// ------------------------------------------------------------ host util --
struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
#define CDNS_HIP(x)                                                                                  \
    do {                                                                                             \
        hipError_t e_ = (x);                                                                         \
        if (e_ != hipSuccess)                                                                        \
            throw ::cdns::HipError(std::string(#x) + ": " + hipGetErrorString(e_));                  \
    } while (0)

// Workgroups are dealt round-robin to the MI355X's 8 XCDs, each with its own L2.  Blocks that
// read the same data should run on one XCD at about the same time: for a 1-D grid of nblocks,
// this returns a logical block id such that each XCD receives one contiguous, in-order range.
__device__ inline int xcd_block(int nblocks) {
    constexpr int X = 8;
    const int b = blockIdx.x;
    const int per = (nblocks + X - 1) / X;
    const int tall = nblocks % X == 0 ? X : nblocks % X;
    const int xcd = b % X, local = b / X;
    return xcd < tall ? xcd * per + local : tall * per + (xcd - tall) * (per - 1) + local;
}

} // namespace cdns
