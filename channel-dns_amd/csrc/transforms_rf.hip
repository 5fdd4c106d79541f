// x and z stages with register-resident two-pass FFTs (N = N1 x N2, one LDS transpose per FFT).
//
// One wavefront per workgroup; the wavefront is split into lane groups of G >= max(N1, N2)
// lanes, one FFT line per group.  The x stage keeps the six synthesised physical lines of a
// point in registers, forms u x w there and transforms the products back, so nothing but
// the NXS+1 retained coefficients per line touches HBM.  Sizes without an instantiation use
// the LDS Stockham kernels of transforms.hip.
#include "internal.hpp"
#include "regfft.hpp"
#include <type_traits>

namespace cdns {

template <int N1, int N2> struct Geo {
    static constexpr int N = N1 * N2;
    static constexpr int MX = N1 > N2 ? N1 : N2;
    static constexpr int G = MX <= 8 ? 8 : (MX <= 16 ? 16 : 32);
    static constexpr int NG = 64 / G;
    static constexpr int L1 = TwoPass<N1, N2>::LDS, L2 = TwoPass<N2, N1>::LDS;
    static constexpr int LB = (L1 > L2 ? (L1 > N ? L1 : N) : (L2 > N ? L2 : N));
};

__device__ inline void wsync() { __syncthreads(); }

template <typename R>
__device__ inline cpx<R> herm_pack(const cpx<R> *A, const cpx<R> *B, int k, int N, int K) {
    // value at slot k of the full-length array whose inverse DFT is a + i b for the real
    // lines a, b with half spectra A, B (FFTW c2r: imaginary parts of DC/Nyquist ignored)
    int src = k;
    bool mirror = false;
    if (k > K) {
        if (N - k > K) return {R(0), R(0)};
        mirror = true;
        src = N - k;
    }
    cpx<R> av = A[src], bv = B ? B[src] : cpx<R>{R(0), R(0)};
    if (src == 0 || 2 * src == N) { av.y = R(0); bv.y = R(0); }
    if (mirror) { av.y = -av.y; bv.y = -bv.y; }
    return {av.x - bv.y, av.y + bv.x};
}

// herm_pack without branches: both loads are issued for every lane (index clamped to the
// retained range) and the value is selected afterwards, so a wavefront's loads of several lines
// can all be in flight at once.
// Element `row + i` of a field given by its kernel-argument base: a 32-bit byte offset from a
// wavefront-uniform base, so the load takes the scalar-base addressing form.
template <typename T> __device__ inline T ld_row(const T *base, unsigned row, int i) {
    const unsigned off = (row + (unsigned)i) * (unsigned)sizeof(T);
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + off);
}

template <typename R>
__device__ inline cpx<R> herm_pack_nb(const cpx<R> *A, const cpx<R> *B, unsigned row, int k, int N, int K, bool use) {
    const bool mirror = k > K, zero = !use || (mirror && N - k > K);
    int src = mirror ? N - k : k;
    src = src < 0 ? 0 : (src > K ? K : src);
    const cpx<R> av = ld_row(A, row, src), bv = ld_row(B, row, src);
    const bool real_end = src == 0 || 2 * src == N;
    const R ay = real_end ? R(0) : (mirror ? -av.y : av.y), by = real_end ? R(0) : (mirror ? -bv.y : bv.y);
    return zero ? cpx<R>{R(0), R(0)} : cpx<R>{av.x - by, ay + bv.x};
}

template <typename R, int N>
__device__ inline void load_roots(cpx<R> *tw, const void *src) {
    const cpx<R> *s = reinterpret_cast<const cpx<R> *>(src);
    for (int i = threadIdx.x; i < N; i += blockDim.x) tw[i] = s[i];
}

// ------------------------------------------------------- x stage + product --
template <typename R, int N1, int N2>
__global__ void __launch_bounds__(64) k_xprod_rf(XArgs a) {
    using Gm = Geo<N1, N2>;
    constexpr int N = Gm::N, MX = Gm::MX, G = Gm::G, NG = Gm::NG, LB = Gm::LB;
    __shared__ cpx<R> tw[N];
    __shared__ cpx<R> buf[NG][LB];
    const int lane = threadIdx.x, g = lane / G, t = lane - g * G;
    load_roots<R, N>(tw, a.tw);
    __syncthreads();
    const int np = a.NL * a.NZ2, K = a.nkx - 1;
    const int p = blockIdx.x * NG + g;
    const bool valid = p < np;
    cpx<R> *B = buf[g];
    cpx<R> ph[3][MX];
    sfor<0, 3>([&](auto line) {
        constexpr int LN = decltype(line)::value;
        cpx<R> x[MX];
        const cpx<R> *A = reinterpret_cast<const cpx<R> *>(a.fin[2 * LN]) + (size_t)p * a.nkx;
        const cpx<R> *Bv = reinterpret_cast<const cpx<R> *>(a.fin[2 * LN + 1]) + (size_t)p * a.nkx;
        sfor<0, MX>([&](auto j) { x[j] = cpx<R>{R(0), R(0)}; });
        if (valid && t < N1) sfor<0, N2>([&](auto n2) { x[n2] = herm_pack(A, Bv, t + N1 * decltype(n2)::value, N, K); });
        two_pass<R, N1, N2, +1>(x, t, tw, B, wsync);
        sfor<0, MX>([&](auto j) { ph[LN][j] = x[j]; });
    });
    // u x w at the N1 points n = t + N2*k1 held by this thread
    cpx<R> F[MX], Fz[MX];
    sfor<0, MX>([&](auto j) {
        F[j] = {R(0), R(0)};
        Fz[j] = {R(0), R(0)};
    });
    if (t < N2) {
        sfor<0, N1>([&](auto k1) {
            const R u = ph[0][k1].x, v = ph[0][k1].y, w = ph[1][k1].x, ox = ph[1][k1].y, oy = ph[2][k1].x,
                    oz = ph[2][k1].y;
            F[k1] = {v * oz - w * oy, w * ox - u * oz};
            Fz[k1] = {u * oy - v * ox, R(0)};
        });
    }
    const R half = R(0.5) / R(N);
    cpx<R> *Nx = reinterpret_cast<cpx<R> *>(a.fout[0]), *Ny = reinterpret_cast<cpx<R> *>(a.fout[1]),
           *Nz = reinterpret_cast<cpx<R> *>(a.fout[2]);
    // forward of (Nx + i Ny): transposed plan, input already in place
    two_pass<R, N2, N1, -1>(F, t, tw, B, wsync);
    if (t < N1) sfor<0, N2>([&](auto k1) { B[t + N1 * decltype(k1)::value] = F[k1]; });
    wsync();
    if (valid)
        for (int k = t; k <= K; k += G) {
            const cpx<R> f = B[k], gm = conjc(B[k == 0 ? 0 : N - k]);
            const cpx<R> s = f + gm, d = f - gm;
            Nx[zline_index(p, k, a.nkx, a.NZ2)] = {half * s.x, half * s.y};
            Ny[zline_index(p, k, a.nkx, a.NZ2)] = {half * d.y, -half * d.x};
        }
    wsync();
    two_pass<R, N2, N1, -1>(Fz, t, tw, B, wsync);
    if (t < N1) sfor<0, N2>([&](auto k1) { B[t + N1 * decltype(k1)::value] = Fz[k1]; });
    wsync();
    if (valid)
        for (int k = t; k <= K; k += G) {
            const cpx<R> f = B[k], gm = conjc(B[k == 0 ? 0 : N - k]);
            const cpx<R> s = f + gm;
            Nz[zline_index(p, k, a.nkx, a.NZ2)] = {half * s.x, half * s.y};
        }
}

// ---------------------------------------------- x stage + product, lane pairs --
// Same work as k_xprod_rf for even N1, N2, with each row of the two-pass FFT on a lane pair
// (radix-2 butterfly across the pair), so a lane holds half as many points: a group of 2 max(N1,N2)
// lanes per (l, iz) point.  This halves the registers of the three physical lines that the
// cross product needs at once (the FP64 kernel then runs several waves per SIMD).
template <int N1, int N2> struct PGeo {
    static constexpr int N = N1 * N2;
    static constexpr int MX = N1 > N2 ? N1 : N2, MH = MX / 2;
    static constexpr int G = 2 * MX <= 16 ? 16 : (2 * MX <= 32 ? 32 : 64);
    static constexpr int NG = 64 / G;
    static constexpr int LB = Geo<N1, N2>::LB;
};

// points per block: 8 in FP64, 16 in FP32, so that each output row segment [l][kx][iz..] is 128 B
template <typename R> constexpr int kXWaves = sizeof(R) == 8 ? 4 : 8;

// Row pitch (entries) of a lane-pair transpose with n entries per row: n itself when n = 2 or 6
// (mod 8), so that the pass-A writes of each 8-lane group hit distinct 16-byte bank groups
// (14: pitch 14), else n + 1 (24: 25; 16: 17).  Entries of one x-product transpose buffer.
constexpr int xpitch(int n) { return (n % 8 == 2 || n % 8 == 6) ? n : n + 1; }
template <int N1, int N2> constexpr int xbuf_entries() {
    constexpr int a = N1 * xpitch(N2), b = N2 * xpitch(N1), n = N1 * N2;
    return a > b ? (a > n ? a : n) : (b > n ? b : n);
}

template <typename R, int N1, int N2>
__global__ void __launch_bounds__(64 * kXWaves<R>) k_xprod_pair(XArgs a) {
    using Gm = PGeo<N1, N2>;
    constexpr int W = kXWaves<R>;
    constexpr int N = Gm::N, MH = Gm::MH, G = Gm::G, NG = Gm::NG, H1 = N1 / 2, H2 = N2 / 2;
    constexpr int PB = W * NG; // points per block
    constexpr int KP = (N / 2 + G) / G; // retained kx <= K <= N/2: entries k = r + G i, i < KP
    // two transpose buffers per point: the two forward transforms run side by side, each pass
    // once for both
    constexpr int PI = xpitch(N2), PF = xpitch(N1), XB = xbuf_entries<N1, N2>();
    __shared__ cpx<R> tw[N];
    __shared__ cpx<R> buf[PB][2 * XB];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane / G, r = lane - g * G, t = r >> 1, h = r & 1;
    const int gb = wv * NG + g;
    const int np = a.NL * a.NZ2, K = a.nkx - 1;
    const int p0 = blockIdx.x * PB, p = p0 + gb;
    const bool valid = p < np;
    cpx<R> *B0 = buf[gb], *B1 = buf[gb] + XB;
    // each line's loads are branch-free, so they are all in flight before the first use; the
    // first line's loads also overlap the root table's.  Issuing all three lines' loads before
    // the first transform costs some 50 more VGPRs in FP64
    cpx<R> ph[3][MH];
    const unsigned prow = (unsigned)(valid ? p : 0) * (unsigned)a.nkx;
    auto load_line = [&](auto line) {
        constexpr int LN = decltype(line)::value;
        const cpx<R> *A = reinterpret_cast<const cpx<R> *>(a.fin[2 * LN]);
        const cpx<R> *Bv = reinterpret_cast<const cpx<R> *>(a.fin[2 * LN + 1]);
        sfor<0, MH>([&](auto m) {
            constexpr int M = decltype(m)::value;
            if constexpr (M < H2)
                ph[LN][M] = herm_pack_nb(A, Bv, prow, t + N1 * (2 * M + h), N, K, valid && t < N1);
            else
                ph[LN][M] = cpx<R>{R(0), R(0)};
        });
    };
    if constexpr (sizeof(R) == 8 && MH <= 7) {
        // FP64 rows of up to 14 points: lines 0 and 1 side by side, then line 2 (156 VGPRs for
        // 196-point lines, still 3 waves per SIMD; the same in FP32 costs two waves per SIMD)
        load_line(std::integral_constant<int, 0>{});
        load_line(std::integral_constant<int, 1>{});
        load_roots<R, N>(tw, a.tw);
        __syncthreads();
        two_pass_pair2<R, N1, N2, +1, false, PI>(ph[0], ph[1], t, h, tw, B0, B1, lsync);
        load_line(std::integral_constant<int, 2>{});
        two_pass_pair<R, N1, N2, +1, false, PI>(ph[2], t, h, tw, B0, lsync);
    } else {
        sfor<0, 3>([&](auto line) {
            constexpr int LN = decltype(line)::value;
            load_line(line);
            if constexpr (LN == 0) {
                load_roots<R, N>(tw, a.tw);
                __syncthreads();
            }
            two_pass_pair<R, N1, N2, +1, false, PI>(ph[LN], t, h, tw, LN == 1 ? B1 : B0, lsync);
        });
    }
    // u x w at the points n = t + N2 (H1 h + k) held by this lane
    cpx<R> F[MH], Fz[MH];
    sfor<0, MH>([&](auto j) {
        F[j] = {R(0), R(0)};
        Fz[j] = {R(0), R(0)};
    });
    if (t < N2) {
        sfor<0, H1>([&](auto k) {
            const R u = ph[0][k].x, v = ph[0][k].y, w = ph[1][k].x, ox = ph[1][k].y, oy = ph[2][k].x, oz = ph[2][k].y;
            F[k] = {v * oz - w * oy, w * ox - u * oz};
            Fz[k] = {u * oy - v * ox, R(0)};
        });
    }
    const R half = R(0.5) / R(N);
    // forwards of (Nx + i Ny) and of Nz side by side: transposed plan, halves input layout
    // already in place
    two_pass_pair2<R, N2, N1, -1, true, PF>(F, Fz, t, h, tw, B0, B1, lsync);
    if (t < N1)
        sfor<0, H2>([&](auto k) {
            B0[t + N1 * (H2 * h + decltype(k)::value)] = F[k];
            B1[t + N1 * (H2 * h + decltype(k)::value)] = Fz[k];
        });
    lsync();
    cpx<R> sx[KP], sy[KP], sz[KP];
    sfor<0, KP>([&](auto i) {
        const int k = r + G * decltype(i)::value;
        if (k <= K) {
            const cpx<R> f = B0[k], gm = conjc(B0[k == 0 ? 0 : N - k]);
            const cpx<R> s = f + gm, d = f - gm;
            sx[i] = {half * s.x, half * s.y};
            sy[i] = {half * d.y, -half * d.x};
            const cpx<R> fz = B1[k], gz = conjc(B1[k == 0 ? 0 : N - k]);
            const cpx<R> s2 = fz + gz;
            sz[i] = {half * s2.x, half * s2.y};
        }
    });
    // all of the group's reads precede its writes (one wavefront, LDS in program order)
    lsync();
    sfor<0, KP>([&](auto i) {
        const int k = r + G * decltype(i)::value;
        if (k <= K) {
            B0[k] = sx[i];
            B0[a.nkx + k] = sy[i];
            B0[2 * a.nkx + k] = sz[i];
        }
    });
    // the block's retained coefficients go out as rows [l][kx][iz0 .. iz0 + PB) of the line-major
    // layout, gathered after a block barrier: thread i writes point j = i mod PB of rows i / PB,
    // i / PB + 64 W / PB, ...
    __syncthreads();
    const int wj = threadIdx.x % PB, wr = threadIdx.x / PB, pw = p0 + wj;
    if (pw < np) {
        const int wl = pw / a.NZ2;
        const size_t wbase = ((size_t)wl * a.nkx) * a.NZ2 + (pw - wl * a.NZ2);
        for (int fo = 0; fo < 3; ++fo) {
            cpx<R> *out = reinterpret_cast<cpx<R> *>(a.fout[fo]) + wbase;
            for (int k = wr; k < a.nkx; k += 64 * W / PB) out[(size_t)k * a.NZ2] = buf[wj][fo * a.nkx + k];
        }
    }
}

// ------------------------------------------- single-field x stages (physical) --
// Two points share one complex line: a + i b.
template <typename R, int N1, int N2>
__global__ void __launch_bounds__(64) k_xbwd_phys_rf(XArgs a) {
    using Gm = Geo<N1, N2>;
    constexpr int N = Gm::N, MX = Gm::MX, G = Gm::G, NG = Gm::NG, LB = Gm::LB;
    __shared__ cpx<R> tw[N];
    __shared__ cpx<R> buf[NG][LB];
    const int lane = threadIdx.x, g = lane / G, t = lane - g * G;
    load_roots<R, N>(tw, a.tw);
    __syncthreads();
    const int np = a.NL * a.NZ2, K = a.nkx - 1;
    const int pa = 2 * (blockIdx.x * NG + g), pb = pa + 1;
    const cpx<R> *Z = reinterpret_cast<const cpx<R> *>(a.in);
    cpx<R> x[MX];
    sfor<0, MX>([&](auto j) { x[j] = cpx<R>{R(0), R(0)}; });
    if (t < N1 && pa < np) {
        const cpx<R> *A = Z + (size_t)pa * a.nkx;
        const cpx<R> *Bv = pb < np ? Z + (size_t)pb * a.nkx : nullptr;
        sfor<0, N2>([&](auto n2) {
            const int k = t + N1 * decltype(n2)::value;
            x[n2] = herm_pack(A, Bv, k, N, K);
            // coarse grid, NXS = N/2: the Nyquist coefficient doubled (transforms.cpp:576-578)
            if (a.coarse && 2 * k == N) x[n2] = {R(2) * x[n2].x, R(2) * x[n2].y};
        });
    }
    two_pass<R, N1, N2, +1>(x, t, tw, buf[g], wsync);
    if (t < N2) {
        sfor<0, N1>([&](auto k1) {
            const int n = t + N2 * decltype(k1)::value;
            if (pa < np) {
                const int l = pa / a.NZ2, iz = pa - l * a.NZ2;
                a.phys[((size_t)iz * a.NL + l) * N + n] = double(x[k1].x);
            }
            if (pb < np) {
                const int l = pb / a.NZ2, iz = pb - l * a.NZ2;
                a.phys[((size_t)iz * a.NL + l) * N + n] = double(x[k1].y);
            }
        });
    }
}

template <typename R, int N1, int N2>
__global__ void __launch_bounds__(64) k_xfwd_phys_rf(XArgs a) {
    using Gm = Geo<N1, N2>;
    constexpr int N = Gm::N, MX = Gm::MX, G = Gm::G, NG = Gm::NG, LB = Gm::LB;
    __shared__ cpx<R> tw[N];
    __shared__ cpx<R> buf[NG][LB];
    const int lane = threadIdx.x, g = lane / G, t = lane - g * G;
    load_roots<R, N>(tw, a.tw);
    __syncthreads();
    const int np = a.NL * a.NZ2, K = a.nkx - 1;
    const int pa = 2 * (blockIdx.x * NG + g), pb = pa + 1;
    cpx<R> x[MX];
    sfor<0, MX>([&](auto j) { x[j] = cpx<R>{R(0), R(0)}; });
    if (t < N2) {
        sfor<0, N1>([&](auto j) {
            const int n = t + N2 * decltype(j)::value;
            R re = R(0), im = R(0);
            if (pa < np) {
                const int l = pa / a.NZ2, iz = pa - l * a.NZ2;
                re = R(a.phys[((size_t)iz * a.NL + l) * N + n]);
            }
            if (pb < np) {
                const int l = pb / a.NZ2, iz = pb - l * a.NZ2;
                im = R(a.phys[((size_t)iz * a.NL + l) * N + n]);
            }
            x[j] = {re, im};
        });
    }
    two_pass<R, N2, N1, -1>(x, t, tw, buf[g], wsync);
    cpx<R> *B = buf[g];
    if (t < N1) sfor<0, N2>([&](auto k1) { B[t + N1 * decltype(k1)::value] = x[k1]; });
    wsync();
    const R half = R(0.5) / R(N);
    cpx<R> *ZN = reinterpret_cast<cpx<R> *>(a.out);
    for (int k = t; k <= K; k += G) {
        const cpx<R> f = B[k], gm = conjc(B[k == 0 ? 0 : N - k]);
        const cpx<R> s = f + gm, d = f - gm;
        // coarse grid, NXS = N/2: the Nyquist coefficient halved (transforms.cpp:496-497)
        const R hk = (a.coarse && 2 * k == N) ? R(0.5) * half : half;
        if (pa < np) ZN[zline_index(pa, k, a.nkx, a.NZ2)] = {hk * s.x, hk * s.y};
        if (pb < np) ZN[zline_index(pb, k, a.nkx, a.NZ2)] = {hk * d.y, -hk * d.x};
    }
}

// ------------------------------------------------------------------ z stages --
// Backward: a block owns KB (16, or 8 for the largest lengths) consecutive kx lines of one
// (field, l), one lane group per line; its output (Z[l][iz][kx], kx contiguous) goes through an LDS
// tile [iz][kx] so that HBM sees whole KB-line rows (256 B in FP64).  Forward: the x stages write
// the line-major layout [l][kx][iz], so one wavefront reads its lines as contiguous runs.
template <typename R, int N1, int N2> struct ZTile {
    using Gm = Geo<N1, N2>;
    static constexpr int N = Gm::N;
    static constexpr int lds_of(int kb) {
        return (kb * Gm::LB > N * (kb + 1) ? kb * Gm::LB : N * (kb + 1)) + N; // tile + roots
    }
    // 16 lines per block when the tile fits 64 KB of LDS, else 8
    static constexpr int KB = lds_of(16) * (int)sizeof(cpx<R>) <= 65536 ? 16 : 8;
    static constexpr int WAVES = KB / Gm::NG, PITCH = KB + 1;
    static constexpr int LDS = lds_of(KB) - N; // complex entries of the tile
};
template <typename R, int N1, int N2> constexpr int kZThreads = 64 * ZTile<R, N1, N2>::WAVES;

// Tile store of a z backward block: lds[iz * PITCH + li] -> Z[iz * nkx + li], li < nl.  A full tile
// (nl = KB, all blocks but a plane's last) gives each thread one li and a fixed row stride, addressed
// as 32-bit offsets from the block's base: no division and no 64-bit address arithmetic per element.
template <typename R, int N, int KB, int PITCH, int NT>
__device__ inline void store_ztile(cpx<R> *Z, const cpx<R> *lds, int nl, int nkx) {
    static_assert(NT % KB == 0, "whole rows per block");
    if (nl == KB) {
        constexpr int RS = NT / KB;
        const int li = threadIdx.x % KB, r0 = threadIdx.x / KB;
        char *zb = reinterpret_cast<char *>(Z);
        const unsigned es = (unsigned)sizeof(cpx<R>), os = (unsigned)(RS * nkx) * es;
        const unsigned o = (unsigned)(r0 * nkx + li) * es;
        sfor<0, (N + RS - 1) / RS>([&](auto i) {
            constexpr int I = decltype(i)::value;
            const int iz = r0 + I * RS;
            if (I * RS + RS <= N || iz < N)
                *reinterpret_cast<cpx<R> *>(zb + o + (unsigned)I * os) = lds[iz * PITCH + li];
        });
        return;
    }
    for (int e = threadIdx.x; e < N * nl; e += NT) {
        const int iz = e / nl, li = e - iz * nl;
        Z[(size_t)iz * nkx + li] = lds[iz * PITCH + li];
    }
}

// backward: slot s of the kz axis <- retained kz, FFT over the slots, Z[l][iz][kx]
template <typename R, int N1, int N2>
__global__ void __launch_bounds__((kZThreads<R, N1, N2>)) k_zbwd_rf(ZArgs a) {
    using Gm = Geo<N1, N2>;
    using T = ZTile<R, N1, N2>;
    constexpr int kZLines = T::KB;
    constexpr int N = Gm::N, MX = Gm::MX, G = Gm::G, NG = Gm::NG, LB = Gm::LB, PITCH = T::PITCH;
    __shared__ cpx<R> tw[N];
    __shared__ cpx<R> lds[T::LDS];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane / G, t = lane - g * G;
    const int line = wv * NG + g;
    load_roots<R, N>(tw, a.tw);
    __syncthreads();
    const int kx0 = blockIdx.x * kZLines, kx = kx0 + line, l = blockIdx.y, f = blockIdx.z;
    const int nl = min(kZLines, a.nkx - kx0);
    const bool valid = line < nl;
    const int NZS = a.NZS;
    const cpx<R> *Y = reinterpret_cast<const cpx<R> *>(a.in) + (size_t)f * a.in_fs + (size_t)l * a.NQ +
                      (size_t)(valid ? kx : 0) * a.nzk;
    cpx<R> x[MX];
    sfor<0, MX>([&](auto j) { x[j] = cpx<R>{R(0), R(0)}; });
    if (valid && t < N1) {
        sfor<0, N2>([&](auto n2) {
            const int s = t + N1 * decltype(n2)::value;
            // slot s holds kz = s (s <= NZS) or s - N (s >= N - NZS); at an overlap the slot
            // N - NZS takes kz = -NZS (transforms.cpp:717-726)
            int kz = 1 << 30;
            if (s <= NZS) kz = s;
            else if (s >= N - NZS) kz = s - N;
            if (s == N - NZS && s <= NZS) kz = -NZS;
            if (kz != (1 << 30)) x[n2] = Y[kz + NZS];
            if (a.coarse && 2 * s == N && s == NZS) x[n2] = Y[2 * NZS] + Y[0];
        });
    }
    two_pass<R, N1, N2, +1>(x, t, tw, lds + line * LB, wsync);
    // two_pass ends with a barrier: the tile may now overwrite the transpose buffers
    if (valid && t < N2) sfor<0, N1>([&](auto k1) { lds[(t + N2 * decltype(k1)::value) * PITCH + line] = x[k1]; });
    __syncthreads();
    cpx<R> *Z = reinterpret_cast<cpx<R> *>(a.out) + (size_t)f * a.out_fs + (size_t)l * N * a.nkx + kx0;
    store_ztile<R, N, kZLines, PITCH, kZThreads<R, N1, N2>>(Z, lds, nl, a.nkx);
}

// forward: line (f, l, kx) over iz; keep the retained kz, x 1/NZ2.
template <typename R, int N1, int N2>
__global__ void __launch_bounds__(64) k_zfwd_rf(ZArgs a) {
    using Gm = Geo<N1, N2>;
    constexpr int N = Gm::N, MX = Gm::MX, G = Gm::G, NG = Gm::NG, LB = Gm::LB;
    // the root table is read from global memory (cache-resident), not staged in LDS: the transpose
    // buffers alone set the blocks per CU
    const cpx<R> *tw = reinterpret_cast<const cpx<R> *>(a.tw);
    __shared__ cpx<R> buf[NG][LB];
    const int lane = threadIdx.x, g = lane / G, t = lane - g * G;
    // neighbouring kx groups of one (field, l) read the same cache lines: keep them on one XCD
    const int L = xcd_block((int)gridDim.x);
    const int kxg = L % a.nkxg, lf = L / a.nkxg, l = lf % a.NL, f = lf / a.NL;
    const int kx = kxg * NG + g;
    const bool valid = kx < a.nkx;
    const int NZS = a.NZS;
    // line (f, l, kx) of the line-major layout: N contiguous values
    const cpx<R> *Z = reinterpret_cast<const cpx<R> *>(a.in) + (size_t)f * a.in_fs +
                      ((size_t)l * a.nkx + (valid ? kx : 0)) * N;
    cpx<R> x[MX];
    sfor<0, MX>([&](auto j) { x[j] = cpx<R>{R(0), R(0)}; });
    if (valid && t < N1) sfor<0, N2>([&](auto n2) { x[n2] = Z[t + N1 * decltype(n2)::value]; });
    two_pass<R, N1, N2, -1>(x, t, tw, buf[g], lsync);
    if (valid && t < N2) {
        const R inv = R(1.0 / N);
        cpx<R> *Y = reinterpret_cast<cpx<R> *>(a.out) + (size_t)f * a.out_fs + (size_t)l * a.NQ + (size_t)kx * a.nzk;
        sfor<0, N1>([&](auto k1) {
            const int s = t + N2 * decltype(k1)::value;
            // when N = 2 NZS the slot NZS serves both kz = NZS and kz = -NZS, halved on the
            // coarse grid (transforms.cpp:507-510)
            const R sc = (a.coarse && 2 * s == N) ? R(0.5) * inv : inv;
            if (s <= NZS) Y[s + NZS] = sc * x[k1];
            if (s >= N - NZS) Y[s - N + NZS] = sc * x[k1];
        });
    }
}

// -------------------------------------------------------- z backward, lane pairs --
// The z backward above with each FFT row on a lane pair (radix-2 across the pair): half the data
// registers per lane, two lines per wavefront, twice the waves per CU.  (For the forward stage the
// one-lane rows stay faster: with two lines per wavefront each row read is only 2 kx wide.)
template <typename R, int N1, int N2> struct ZTileP {
    using Gm = PGeo<N1, N2>;
    static constexpr int N = Gm::N;
    static constexpr int lds_of(int kb) {
        return (kb * Gm::LB > N * (kb + 1) ? kb * Gm::LB : N * (kb + 1)) + N; // tile + roots
    }
    static constexpr int KB = lds_of(16) * (int)sizeof(cpx<R>) <= 65536 ? 16 : 8;
    static constexpr int WAVES = KB / Gm::NG, PITCH = KB + 1;
    static constexpr int LDS = lds_of(KB) - N;
};
template <typename R, int N1, int N2> constexpr int kZPThreads = 64 * ZTileP<R, N1, N2>::WAVES;

template <typename R, int N1, int N2>
__global__ void __launch_bounds__((kZPThreads<R, N1, N2>)) k_zbwd_pair(ZArgs a) {
    using Gm = PGeo<N1, N2>;
    using T = ZTileP<R, N1, N2>;
    constexpr int kZLines = T::KB;
    constexpr int N = Gm::N, MH = Gm::MH, G = Gm::G, NG = Gm::NG, LB = Gm::LB, PITCH = T::PITCH;
    constexpr int H1 = N1 / 2, H2 = N2 / 2;
    __shared__ cpx<R> tw[N];
    __shared__ cpx<R> lds[T::LDS];
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane / G, r = lane - g * G, t = r >> 1, h = r & 1;
    const int line = wv * NG + g;
    const int kx0 = blockIdx.x * kZLines, kx = kx0 + line, l = blockIdx.y, f = blockIdx.z;
    const int nl = min(kZLines, a.nkx - kx0);
    const bool valid = line < nl;
    const int NZS = a.NZS;
    const cpx<R> *Y = reinterpret_cast<const cpx<R> *>(a.in) + (size_t)f * a.in_fs + (size_t)l * a.NQ +
                      (size_t)(valid ? kx : 0) * a.nzk;
    cpx<R> x[MH];
    sfor<0, MH>([&](auto j) { x[j] = cpx<R>{R(0), R(0)}; });
    if (valid && t < N1) {
        sfor<0, H2>([&](auto m) {
            const int s = t + N1 * (2 * decltype(m)::value + h);
            // slot s holds kz = s (s <= NZS) or s - N (s >= N - NZS); at an overlap the slot
            // N - NZS takes kz = -NZS (transforms.cpp:717-726)
            int kz = 1 << 30;
            if (s <= NZS) kz = s;
            else if (s >= N - NZS) kz = s - N;
            if (s == N - NZS && s <= NZS) kz = -NZS;
            if (kz != (1 << 30)) x[m] = Y[kz + NZS];
            // coarse grid, NZS = N/2: the slot holds X(NZS) + X(-NZS) (transforms.cpp:561-564)
            if (a.coarse && 2 * s == N && s == NZS) x[m] = Y[2 * NZS] + Y[0];
        });
    }
    load_roots<R, N>(tw, a.tw); // after the line loads: both latencies overlap
    __syncthreads();
    // a plane's last block can hold fewer lines than waves x lines per wave: a wavefront whose lines
    // all lie past the plane's end skips the transform (its passes synchronise only within the wave)
    if (wv * NG < nl) two_pass_pair<R, N1, N2, +1, false>(x, t, h, tw, lds + line * LB, lsync);
    __syncthreads(); // every line's transposes are done: the tile may overwrite their buffers
    if (valid && t < N2)
        sfor<0, H1>([&](auto k) { lds[(t + N2 * (H1 * h + decltype(k)::value)) * PITCH + line] = x[k]; });
    __syncthreads();
    cpx<R> *Z = reinterpret_cast<cpx<R> *>(a.out) + (size_t)f * a.out_fs + (size_t)l * N * a.nkx + kx0;
    store_ztile<R, N, kZLines, PITCH, kZPThreads<R, N1, N2>>(Z, lds, nl, a.nkx);
}

// ------------------------------------------------------------------ launch --
template <typename R, int N1, int N2> static void go_xprod(const XArgs &a, hipStream_t st) {
    if constexpr (N1 % 2 == 0 && N2 % 2 == 0) {
        // the three retained rows of a point are staged in its transpose buffer
        if (a.pair && 3 * a.nkx <= 2 * xbuf_entries<N1, N2>()) {
            constexpr int PB = kXWaves<R> * PGeo<N1, N2>::NG;
            hipLaunchKernelGGL((k_xprod_pair<R, N1, N2>), dim3((a.NL * a.NZ2 + PB - 1) / PB), dim3(64 * kXWaves<R>),
                               0, st, a);
            return;
        }
    }
    using Gm = Geo<N1, N2>;
    hipLaunchKernelGGL((k_xprod_rf<R, N1, N2>), dim3((a.NL * a.NZ2 + Gm::NG - 1) / Gm::NG), dim3(64), 0, st, a);
}
template <typename R, int N1, int N2> static void go_xbwd(const XArgs &a, hipStream_t st) {
    using Gm = Geo<N1, N2>;
    hipLaunchKernelGGL((k_xbwd_phys_rf<R, N1, N2>), dim3((a.NL * a.NZ2 + 2 * Gm::NG - 1) / (2 * Gm::NG)), dim3(64), 0,
                       st, a);
}
template <typename R, int N1, int N2> static void go_xfwd(const XArgs &a, hipStream_t st) {
    using Gm = Geo<N1, N2>;
    hipLaunchKernelGGL((k_xfwd_phys_rf<R, N1, N2>), dim3((a.NL * a.NZ2 + 2 * Gm::NG - 1) / (2 * Gm::NG)), dim3(64), 0,
                       st, a);
}
template <typename R, int N1, int N2> static void go_zbwd(const ZArgs &a, int nf, hipStream_t st) {
    if constexpr (N1 % 2 == 0 && N2 % 2 == 0) {
        if (a.pair) {
            using T = ZTileP<R, N1, N2>;
            hipLaunchKernelGGL((k_zbwd_pair<R, N1, N2>), dim3((a.nkx + T::KB - 1) / T::KB, a.NL, nf),
                               dim3(64 * T::WAVES), 0, st, a);
            return;
        }
    }
    using T = ZTile<R, N1, N2>;
    hipLaunchKernelGGL((k_zbwd_rf<R, N1, N2>), dim3((a.nkx + T::KB - 1) / T::KB, a.NL, nf), dim3(64 * T::WAVES), 0,
                       st, a);
}
template <typename R, int N1, int N2> static void go_zfwd(const ZArgs &a, int nf, hipStream_t st) {
    using Gm = Geo<N1, N2>;
    ZArgs b = a;
    b.nkxg = (a.nkx + Gm::NG - 1) / Gm::NG;
    hipLaunchKernelGGL((k_zfwd_rf<R, N1, N2>), dim3(b.nkxg * a.NL * nf), dim3(64), 0, st, b);
}

// Supported lengths: N = N1 * N2 with N1, N2 <= 16; the z stages also take the longer lines
// below (factors up to 24: one line per lane group of 32, or per wavefront on lane pairs), which
// the x product cannot hold (its three physical lines would need 288 VGPRs per lane).
#define CDNS_RF_ZONLY_SIZES(X) X(288, 16, 18) X(384, 16, 24)
#define CDNS_RF_SIZES(X) X(196, 14, 14) X(48, 6, 8) X(96, 8, 12) X(32, 4, 8) X(64, 8, 8) X(128, 8, 16) \
    X(144, 12, 12) X(192, 12, 16) X(256, 16, 16) X(36, 6, 6) X(72, 8, 9) X(120, 10, 12) X(160, 10, 16)

bool rf_supported(int n) {
#define X(NN, A, B) if (n == NN) return true;
    CDNS_RF_SIZES(X)
#undef X
    return false;
}

bool launch_xprod_rf(const XArgs &a, bool fp32, hipStream_t st) {
    const int n = a.plan.n;
#define X(NN, A, B) case NN: if (fp32) go_xprod<float, A, B>(a, st); else go_xprod<double, A, B>(a, st); return true;
    switch (n) { CDNS_RF_SIZES(X) X(384, 16, 24) default: return false; }
#undef X
}
bool launch_xbwd_phys_rf(const XArgs &a, bool fp32, hipStream_t st) {
    const int n = a.plan.n;
#define X(NN, A, B) case NN: if (fp32) go_xbwd<float, A, B>(a, st); else go_xbwd<double, A, B>(a, st); return true;
    switch (n) { CDNS_RF_SIZES(X) default: return false; }
#undef X
}
bool launch_xfwd_phys_rf(const XArgs &a, bool fp32, hipStream_t st) {
    const int n = a.plan.n;
#define X(NN, A, B) case NN: if (fp32) go_xfwd<float, A, B>(a, st); else go_xfwd<double, A, B>(a, st); return true;
    switch (n) { CDNS_RF_SIZES(X) default: return false; }
#undef X
}
bool launch_zbwd_rf(const ZArgs &a, int nf, bool fp32, hipStream_t st) {
    const int n = a.plan.n;
#define X(NN, A, B) case NN: if (fp32) go_zbwd<float, A, B>(a, nf, st); else go_zbwd<double, A, B>(a, nf, st); return true;
    switch (n) { CDNS_RF_SIZES(X) CDNS_RF_ZONLY_SIZES(X) default: return false; }
#undef X
}
bool launch_zfwd_rf(const ZArgs &a, int nf, bool fp32, hipStream_t st) {
    const int n = a.plan.n;
#define X(NN, A, B) case NN: if (fp32) go_zfwd<float, A, B>(a, nf, st); else go_zfwd<double, A, B>(a, nf, st); return true;
    switch (n) { CDNS_RF_SIZES(X) CDNS_RF_ZONLY_SIZES(X) default: return false; }
#undef X
}

} // namespace cdns
