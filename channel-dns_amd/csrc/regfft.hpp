// Register-resident small DFTs with compile-time twiddles, and the two-pass (N = N1 x N2)
// line FFT built on them: each thread of a lane group does N2- and N1-point DFTs in
// registers, with one LDS transpose between the passes.
#pragma once

#include "common.hpp"

#include <type_traits>
#include <utility>

namespace cdns {

// ------------------------------------------------- compile-time trigonometry --
namespace ct {
constexpr double kPiOver2 = 1.57079632679489661923132169163975144;

constexpr double sin_taylor(double x) { // |x| <= pi/4
    double x2 = x * x, term = x, sum = x;
    for (int i = 1; i < 14; ++i) {
        term *= -x2 / ((2.0 * i) * (2.0 * i + 1.0));
        sum += term;
    }
    return sum;
}
constexpr double cos_taylor(double x) {
    double x2 = x * x, term = 1.0, sum = 1.0;
    for (int i = 1; i < 14; ++i) {
        term *= -x2 / ((2.0 * i - 1.0) * (2.0 * i));
        sum += term;
    }
    return sum;
}
// cos(2 pi k / N) (which = 0) or sin(2 pi k / N) (which = 1), with exact quadrant reduction.
constexpr double trig(long long k, long long N, int which) {
    long long m = (4 * k) % (4 * N);
    if (m < 0) m += 4 * N;
    const long long q = m / N, r = m - q * N; // angle = (q + r/N) pi/2
    double c, s;
    if (2 * r <= N) {
        const double x = kPiOver2 * (double)r / (double)N;
        c = cos_taylor(x);
        s = sin_taylor(x);
    } else {
        const double x = kPiOver2 * (double)(N - r) / (double)N;
        c = sin_taylor(x);
        s = cos_taylor(x);
    }
    if (r == 0) { c = 1.0; s = 0.0; }
    double cc = c, ss = s;
    if (q == 1) { cc = -s; ss = c; }
    else if (q == 2) { cc = -c; ss = -s; }
    else if (q == 3) { cc = s; ss = -c; }
    return which == 0 ? cc : ss;
}
} // namespace ct

// exp(S * 2 pi i K / N)
template <int K, int N, int S> struct Tw {
    static constexpr double c = ct::trig(K, N, 0);
    static constexpr double s = S * ct::trig(K, N, 1);
};

// Ordering of one wavefront's LDS accesses: a lane group's transpose buffer belongs to one
// wavefront, whose LDS instructions execute in program order, so the FFT passes need only a
// compiler-level fence, not a workgroup barrier that would make the other wavefronts wait.
__device__ inline void lsync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

template <int I> using ic = std::integral_constant<int, I>;
template <int B, int E, typename F> __device__ inline void sfor(F &&f) {
    if constexpr (B < E) {
        f(ic<B>{});
        sfor<B + 1, E>(f);
    }
}

template <typename R, int K, int N, int S> __device__ inline cpx<R> twmul(cpx<R> v) {
    if constexpr (K % N == 0) return v;
    else if constexpr ((4 * K) % N == 0) {
        // multiples of pi/2: exact rotations
        constexpr int q = (((4 * K) / N) % 4 + 4) % 4;
        constexpr int qs = S > 0 ? q : (4 - q) % 4;
        if constexpr (qs == 1) return {-v.y, v.x};
        else if constexpr (qs == 2) return {-v.x, -v.y};
        else return {v.y, -v.x};
    } else {
        constexpr R c = (R)Tw<K, N, S>::c, s = (R)Tw<K, N, S>::s;
        return {v.x * c - v.y * s, v.x * s + v.y * c};
    }
}

constexpr int small_factor(int n) {
    if (n % 4 == 0 && n != 8 && n > 4) return 4;
    for (int p : {2, 3, 5, 7})
        if (n % p == 0) return p;
    return n;
}

// y_m = sum_q v_q exp(S 2 pi i q m / N), in registers, fully unrolled.
template <typename R, int N, int S> struct RDft {
    __device__ static inline void run(cpx<R> (&v)[N]) {
        if constexpr (N == 1) {
            return;
        } else if constexpr (N == 2 || N == 3 || N == 4 || N == 5 || N == 7 || N == 8) {
            Dft<R, N>::run(v, S);
        } else {
            constexpr int p = small_factor(N), m = N / p;
            cpx<R> sub[p][m];
            sfor<0, p>([&](auto r) {
                sfor<0, m>([&](auto j) { sub[r][j] = v[r + p * j]; });
                RDft<R, m, S>::run(sub[r]);
            });
            sfor<0, m>([&](auto k) {
                cpx<R> t[p];
                sfor<0, p>([&](auto r) { t[r] = twmul<R, decltype(r)::value * decltype(k)::value, N, S>(sub[r][k]); });
                Dft<R, p>::run(t, S);
                sfor<0, p>([&](auto qq) { v[k + m * qq] = t[qq]; });
            });
        }
    }
};

// RDft for inputs that are zero from index NZ on (those entries are not read): each sub-DFT of
// the recursion is pruned to its nonzero inputs, and the base transforms with a zero tail become
// direct sums with compile-time twiddles.
template <typename R, int N, int S, int NZ> struct RDftP {
    __device__ static inline void run(cpx<R> (&v)[N]) {
        if constexpr (NZ >= N) {
            RDft<R, N, S>::run(v);
        } else if constexpr (NZ <= 0) {
            sfor<0, N>([&](auto k) { v[k] = cpx<R>{R(0), R(0)}; });
        } else if constexpr (NZ == 1) {
            sfor<1, N>([&](auto k) { v[k] = v[0]; });
        } else if constexpr (N == 2 || N == 3 || N == 4 || N == 5 || N == 7 || N == 8) {
            cpx<R> out[N];
            sfor<0, N>([&](auto k) {
                constexpr int K = decltype(k)::value;
                cpx<R> acc = v[0];
                sfor<1, NZ>([&](auto q) {
                    constexpr int Q = decltype(q)::value;
                    acc = acc + twmul<R, (Q * K) % N, N, S>(v[Q]);
                });
                out[K] = acc;
            });
            sfor<0, N>([&](auto k) { v[k] = out[k]; });
        } else {
            constexpr int p = small_factor(N), m = N / p;
            constexpr int NZP = NZ < p ? NZ : p; // sub-transforms r < NZP have a nonzero input
            cpx<R> sub[p][m];
            sfor<0, NZP>([&](auto r) {
                constexpr int RR = decltype(r)::value;
                constexpr int NZR = (NZ - RR + p - 1) / p;
                sfor<0, m>([&](auto j) {
                    constexpr int J = decltype(j)::value;
                    if constexpr (J < NZR) sub[RR][J] = v[RR + p * J];
                    else sub[RR][J] = cpx<R>{R(0), R(0)};
                });
                RDftP<R, m, S, NZR>::run(sub[RR]);
            });
            sfor<0, m>([&](auto k) {
                constexpr int K = decltype(k)::value;
                cpx<R> t[p];
                sfor<0, p>([&](auto r) {
                    constexpr int RR = decltype(r)::value;
                    if constexpr (RR < NZP) t[RR] = twmul<R, RR * K, N, S>(sub[RR][K]);
                    else t[RR] = cpx<R>{R(0), R(0)};
                });
                RDftP<R, p, S, NZP>::run(t);
                sfor<0, p>([&](auto qq) { v[K + m * qq] = t[qq]; });
            });
        }
    }
};

// Lane-group layout of the two-pass FFT of one line of N = N1*N2 points.
//   pass A: thread t < N1 holds x[t + N1*n2], n2 < N2; DFT over n2; times w^(t*k2); to LDS T[t][k2]
//   pass B: thread t < N2 holds T[n1][t], n1 < N1; DFT over n1 -> X[t + N2*k1], k1 < N1
// The output of pass B (index t + N2*k1) is the pass-A input layout of the transposed plan
// (N1' = N2, N2' = N1), so a forward FFT can follow an inverse one without data movement.
template <int N1, int N2> struct TwoPass {
    static constexpr int N = N1 * N2;
    static constexpr int PA = N2 + 1;  // LDS row pitch of pass-A output (conflict-free rows)
    static constexpr int LDS = N1 * PA; // complex entries per group buffer
};

// Pass A + transpose + pass B.  `x` holds pass-A input on threads t < N1 and receives pass-B
// output on threads t < N2 (arrays sized max(N1, N2)).  `tw` is the N-point root table
// (exp(-2 pi i k / N)) in LDS, `T` the group's transpose buffer.  Sync is the caller's barrier.
template <typename R, int N1, int N2, int S, int M, typename Sync>
__device__ __attribute__((always_inline)) inline void two_pass(cpx<R> (&x)[M], int t, const cpx<R> *tw, cpx<R> *T, Sync sync) {
    constexpr int N = N1 * N2, PA = N2 + 1;
    if (t < N1) {
        cpx<R> a[N2];
        sfor<0, N2>([&](auto j) { a[j] = x[j]; });
        RDft<R, N2, S>::run(a);
        sfor<0, N2>([&](auto k2) {
            constexpr int K2 = decltype(k2)::value;
            cpx<R> w = tw[(t * K2) % N];
            if (S > 0) w.y = -w.y;
            const cpx<R> v = a[K2];
            T[t * PA + K2] = {v.x * w.x - v.y * w.y, v.x * w.y + v.y * w.x};
        });
    }
    sync();
    if (t < N2) {
        cpx<R> b[N1];
        sfor<0, N1>([&](auto n1) { b[n1] = T[n1 * PA + t]; });
        RDft<R, N1, S>::run(b);
        sfor<0, N1>([&](auto k1) { x[k1] = b[k1]; });
    }
    sync();
}


// two_pass with the pass-A input x[n2] zero for n2 >= NZA (a zero-padded line: the pass-A DFTs are
// pruned to their nonzero inputs).
template <typename R, int N1, int N2, int S, int NZA, int M, typename Sync>
__device__ __attribute__((always_inline)) inline void two_pass_pruned(cpx<R> (&x)[M], int t, const cpx<R> *tw, cpx<R> *T,
                                                                     Sync sync) {
    constexpr int N = N1 * N2, PA = N2 + 1;
    if (t < N1) {
        cpx<R> a[N2];
        sfor<0, N2>([&](auto j) {
            constexpr int J = decltype(j)::value;
            if constexpr (J < NZA) a[J] = x[J];
            else a[J] = cpx<R>{R(0), R(0)};
        });
        RDftP<R, N2, S, NZA>::run(a);
        sfor<0, N2>([&](auto k2) {
            constexpr int K2 = decltype(k2)::value;
            cpx<R> w = tw[t * K2]; // t K2 <= (N1 - 1)(N2 - 1) < N
            if (S > 0) w.y = -w.y;
            const cpx<R> v = a[K2];
            T[t * PA + K2] = {v.x * w.x - v.y * w.y, v.x * w.y + v.y * w.x};
        });
    }
    sync();
    if (t < N2) {
        cpx<R> b[N1];
        sfor<0, N1>([&](auto n1) { b[n1] = T[n1 * PA + t]; });
        RDft<R, N1, S>::run(b);
        sfor<0, N1>([&](auto k1) { x[k1] = b[k1]; });
    }
    sync();
}

// ------------------------------------------------------- lane-pair DFTs (2H points) --
// The two lanes of a pair (lane ^ 1) hold H points each of one DFT of length 2H; a radix-2
// butterfly across the pair (DPP swap of adjacent lanes) joins two register DFTs of length H.
__device__ inline float pair_swap(float v) {
    return __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, false));
}
__device__ inline double pair_swap(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0xB1, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0xB1, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}
template <typename R> __device__ inline cpx<R> pair_swap(cpx<R> v) { return {pair_swap(v.x), pair_swap(v.y)}; }
// component-wise select (a select of whole structs can be lowered through scratch memory)
template <typename R> __device__ inline cpx<R> csel(bool c, cpx<R> a, cpx<R> b) {
    return {c ? a.x : b.x, c ? a.y : b.y};
}

// v (exp(S 2 pi i K / N) - 1), the twiddle step of the pair butterflies applied as a correction
template <typename R, int K, int N, int S> __device__ inline cpx<R> twmul_m1(cpx<R> v) {
    if constexpr ((4 * K) % N == 0) {
        const cpx<R> w = twmul<R, K, N, S>(v);
        return {w.x - v.x, w.y - v.y};
    } else {
        constexpr R c = (R)(Tw<K, N, S>::c - 1.0), s = (R)Tw<K, N, S>::s;
        return {v.x * c - v.y * s, v.x * s + v.y * c};
    }
}

// DIT: lane h holds x[2m + h], m < H; on return it holds X[H h + k], k < H.
// Radix-2 across the pair: lane 1 twiddles its odd-half DFT (v + h v (W^K - 1)), then lane 0
// forms E + W O = v + p and lane 1 E - W O = p - v, both as p + sg v with sg = 1 - 2h.
template <typename R, int H, int S> __device__ inline void pair_dit(cpx<R> (&v)[H], int h) {
    RDft<R, H, S>::run(v);
    const R hd = R(h), sg = R(1 - 2 * h);
    sfor<0, H>([&](auto k) {
        constexpr int K = decltype(k)::value;
        if constexpr (K > 0) {
            const cpx<R> d = twmul_m1<R, K, 2 * H, S>(v[K]);
            v[K] = {v[K].x + hd * d.x, v[K].y + hd * d.y};
        }
        const cpx<R> p = pair_swap(v[K]);
        v[K] = {p.x + sg * v[K].x, p.y + sg * v[K].y};
    });
}
// DIF: lane h holds x[H h + m], m < H; on return it holds X[2k + h], k < H.
// Lane 0 forms v + p, lane 1 (p - v) W^M: both p + sg v, then lane 1's twiddle as a correction.
template <typename R, int H, int S> __device__ inline void pair_dif(cpx<R> (&v)[H], int h) {
    const R hd = R(h), sg = R(1 - 2 * h);
    sfor<0, H>([&](auto m) {
        constexpr int M = decltype(m)::value;
        const cpx<R> p = pair_swap(v[M]);
        v[M] = {p.x + sg * v[M].x, p.y + sg * v[M].y};
        if constexpr (M > 0) {
            const cpx<R> d = twmul_m1<R, M, 2 * H, S>(v[M]);
            v[M] = {v[M].x + hd * d.x, v[M].y + hd * d.y};
        }
    });
    RDft<R, H, S>::run(v);
}

// Two-pass FFT of N = N1 N2 (N1, N2 even) on lane pairs: row thread t (pass A: t < N1, pass B:
// t < N2) is the pair (t, h); the pass-A rows sit in T at a pitch of PA entries.  Pass A input:
// DIT layout x[m] = line[t + N1 (2m + h)] or, with AIN_HALVES, line[t + N1 (H2 h + m)].  Pass B
// output: x[k] = X[t + N2 (H1 h + k)], which is the halves input layout of the transposed plan.
// x has max(N1, N2) / 2 entries.
template <typename R, int N1, int N2, int S, bool AIN_HALVES, int PA = N2 + 1, int M, typename Sync>
__device__ __attribute__((always_inline)) inline void two_pass_pair(cpx<R> (&x)[M], int t, int h, const cpx<R> *tw, cpx<R> *T, Sync sync) {
    constexpr int N = N1 * N2, H1 = N1 / 2, H2 = N2 / 2;
    static_assert(N1 % 2 == 0 && N2 % 2 == 0, "pair FFT needs even factors");
    if (t < N1) {
        cpx<R> a[H2];
        sfor<0, H2>([&](auto j) { a[j] = x[j]; });
        int k2base, k2step;
        if constexpr (AIN_HALVES) {
            pair_dif<R, H2, S>(a, h);
            k2base = h;
            k2step = 2;
        } else {
            pair_dit<R, H2, S>(a, h);
            k2base = H2 * h;
            k2step = 1;
        }
        sfor<0, H2>([&](auto k) {
            constexpr int K = decltype(k)::value;
            const int k2 = k2base + k2step * K;
            cpx<R> w = tw[(t * k2) % N];
            if (S > 0) w.y = -w.y;
            const cpx<R> v = a[K];
            T[t * PA + k2] = {v.x * w.x - v.y * w.y, v.x * w.y + v.y * w.x};
        });
    }
    sync();
    if (t < N2) {
        cpx<R> b[H1];
        sfor<0, H1>([&](auto m) { b[m] = T[(2 * decltype(m)::value + h) * PA + t]; });
        pair_dit<R, H1, S>(b, h);
        sfor<0, H1>([&](auto k) { x[k] = b[k]; });
    }
    sync();
}

// Two independent transforms of the same plan side by side (x0 through T0, x1 through T1): each
// pass runs both, so their arithmetic interleaves and the pair shares one barrier per pass and
// one root-table read per element.  Row pitch PA (entries) of the pass-A output.
template <typename R, int N1, int N2, int S, bool AIN_HALVES, int PA, int M, typename Sync>
__device__ __attribute__((always_inline)) inline void two_pass_pair2(cpx<R> (&x0)[M], cpx<R> (&x1)[M], int t, int h,
                                                                     const cpx<R> *tw, cpx<R> *T0, cpx<R> *T1,
                                                                     Sync sync) {
    constexpr int N = N1 * N2, H1 = N1 / 2, H2 = N2 / 2;
    static_assert(N1 % 2 == 0 && N2 % 2 == 0, "pair FFT needs even factors");
    if (t < N1) {
        cpx<R> a0[H2], a1[H2];
        sfor<0, H2>([&](auto j) {
            a0[j] = x0[j];
            a1[j] = x1[j];
        });
        int k2base, k2step;
        if constexpr (AIN_HALVES) {
            pair_dif<R, H2, S>(a0, h);
            pair_dif<R, H2, S>(a1, h);
            k2base = h;
            k2step = 2;
        } else {
            pair_dit<R, H2, S>(a0, h);
            pair_dit<R, H2, S>(a1, h);
            k2base = H2 * h;
            k2step = 1;
        }
        sfor<0, H2>([&](auto k) {
            constexpr int K = decltype(k)::value;
            const int k2 = k2base + k2step * K;
            cpx<R> w = tw[(t * k2) % N];
            if (S > 0) w.y = -w.y;
            const cpx<R> v0 = a0[K], v1 = a1[K];
            T0[t * PA + k2] = {v0.x * w.x - v0.y * w.y, v0.x * w.y + v0.y * w.x};
            T1[t * PA + k2] = {v1.x * w.x - v1.y * w.y, v1.x * w.y + v1.y * w.x};
        });
    }
    sync();
    if (t < N2) {
        cpx<R> b0[H1], b1[H1];
        sfor<0, H1>([&](auto m) {
            b0[m] = T0[(2 * decltype(m)::value + h) * PA + t];
            b1[m] = T1[(2 * decltype(m)::value + h) * PA + t];
        });
        pair_dit<R, H1, S>(b0, h);
        pair_dit<R, H1, S>(b1, h);
        sfor<0, H1>([&](auto k) {
            x0[k] = b0[k];
            x1[k] = b1[k];
        });
    }
    sync();
}

} // namespace cdns
