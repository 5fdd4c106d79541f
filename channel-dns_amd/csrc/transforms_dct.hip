// y stages as a DCT-I computed through complex FFTs of length L = 2 NY2 (register-resident
// two-pass FFTs), the alternative to the matrix-core Chebyshev apply of
// transforms_y.hip.  With N = NY2:
//
// Backward (transforms.cpp:692-707): y_l = sum_j c_j cos(pi j l / N), l = 0..N, for the column
//   c_0..c_NYS (the j = N term at half weight, the REDFT00 end point once the reference's doubled
//   j = 0 and the halved result are folded in).  With X the inverse DFT of length L of the column
//   zero-padded beyond NYS, cos = (e^{i t} + e^{-i t}) / 2 gives y_l = (X_l + X_{L-l}) / 2.
// Forward (transforms.cpp:673-682): S'_j = Y_0 + (-1)^j Y_N + 2 sum_{l=1}^{N-1} Y_l cos(pi j l / N) =
//   X_j + X_{L-j}, X the DFT of (Y_0/2, Y_1, ..., Y_{N-1}, Y_N/2, 0, ...); S_0 = S'_0 / (2N),
//   S_j = S'_j / N for 1 <= j <= NYS.
//
// A workgroup owns 8 (FP64) or 16 (FP32) consecutive columns q of one field, so that each row of
// the Y field is written or read as a 128-byte segment; each FFT line sits on a lane group of the
// wavefront.  The Y side goes through an LDS tile [l][column] that reuses the FFT buffers; the
// spectral side is read and written column-contiguous.  The fields of one column group are
// neighbouring blocks on one XCD, so the columns they share stay in that XCD's L2.
//
// The backward with the curl (launch_ybwd_dct_curl) forms the spectral vorticity of each column
// in FP64 before its transform: wx = D w - i kz v, wy = i kz u - i kx w, wz = i kx v - D u
// (vector_ops.cpp:112-151, D the Chebyshev derivative recurrence of vector_ops.cpp:17-25), and the
// curl part of the pressure wall condition (timestepper.cpp:26-43) from closed-form wall values:
// at y = +1, (1/Re)(-(kx^2 + kz^2) v(1) - i kx u'(1) - i kz w'(1)) with v(1) = sum v_j and
// u'(1) = sum j^2 u_j; at wall 1 minus the same at y = -1 (v(-1) = sum (-1)^j v_j,
// u'(-1) = sum (-1)^(j+1) j^2 u_j).
#include "internal.hpp"
#include "regfft.hpp"
#include "wave.hpp"

namespace cdns {

// Column q's value of field f of the backward with the curl, FP64 in the (lane, parity) layout of
// wave.hpp, plus (f == 1) the curl part of the pressure wall condition.
template <int CH>
__device__ inline void curl_field(const DctArgs &a, int f, int q, bool valid, zd (&s)[2][CH], int lane) {
    const int ny = a.ny;
    const size_t off = (size_t)(valid ? q : 0) * a.JS;
    const int kxi = (valid ? q : 0) / a.nzk, kzi = (valid ? q : 0) - kxi * a.nzk;
    const double kx = a.wavx[kxi], kz = a.wavz[kzi];
    if (f < 3) {
        load_col<CH>(a.spec[f] + off, lane, ny, s);
    } else if (f == 3) { // wx = D w - i kz v
        zd v[2][CH], w[2][CH], dw[2][CH];
        load_col<CH>(a.spec[1] + off, lane, ny, v);
        load_col<CH>(a.spec[2] + off, lane, ny, w);
        cheb_dy<CH>(w, dw, lane, ny);
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int c = 0; c < CH; ++c) s[p][c] = dw[p][c] - ikmul(kz, v[p][c]);
    } else if (f == 4) { // wy = i kz u - i kx w
        zd u[2][CH], w[2][CH];
        load_col<CH>(a.spec[0] + off, lane, ny, u);
        load_col<CH>(a.spec[2] + off, lane, ny, w);
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int c = 0; c < CH; ++c) s[p][c] = ikmul(kz, u[p][c]) - ikmul(kx, w[p][c]);
    } else { // wz = i kx v - D u
        zd u[2][CH], v[2][CH], du[2][CH];
        load_col<CH>(a.spec[0] + off, lane, ny, u);
        load_col<CH>(a.spec[1] + off, lane, ny, v);
        cheb_dy<CH>(u, du, lane, ny);
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int c = 0; c < CH; ++c) s[p][c] = ikmul(kx, v[p][c]) - du[p][c];
    }
    if (f == 1 && a.bcw) {
        zd u[2][CH], w[2][CH];
        load_col<CH>(a.spec[0] + off, lane, ny, u);
        load_col<CH>(a.spec[2] + off, lane, ny, w);
        zd vp = {0.0, 0.0}, vm = {0.0, 0.0}, up = {0.0, 0.0}, um = {0.0, 0.0}, wp = {0.0, 0.0}, wm = {0.0, 0.0};
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                const double j = (double)jof<CH>(lane, c, p), j2 = j * j;
                const double sg = p == 0 ? 1.0 : -1.0; // (-1)^j
                vp = vp + s[p][c];
                vm = vm + sg * s[p][c];
                up = up + j2 * u[p][c];
                um = um - sg * j2 * u[p][c];
                wp = wp + j2 * w[p][c];
                wm = wm - sg * j2 * w[p][c];
            }
        vp = wave_sum(vp);
        vm = wave_sum(vm);
        up = wave_sum(up);
        um = wave_sum(um);
        wp = wave_sum(wp);
        wm = wave_sum(wm);
        if (lane == 0 && valid) {
            const double k2 = kx * kx + kz * kz, ire = 1.0 / a.Re;
            const zd top = ire * (-k2 * vp - ikmul(kx, up) - ikmul(kz, wp));
            const zd bot = ire * (-k2 * vm - ikmul(kx, um) - ikmul(kz, wm));
            a.bcw[(size_t)q * 2 + 0] = top;
            a.bcw[(size_t)q * 2 + 1] = -bot;
        }
    }
}

// Each FFT row sits on one lane, two or four lines per wavefront, the root table is read from
// global memory (cached), and pass A is pruned to the nonzero part of the line: the backward line
// holds the NYS + 1 coefficients of L = 2 NY2 slots, the forward line the NY2 + 1 samples.  NZA is
// the count of pass-A inputs per row that can be nonzero (n = t + N1 n2 with n2 < NZA), a
// compile-time bound on the line's nonzero prefix.
template <typename R, int N1, int N2> struct Dct2Geo {
    static constexpr int L = N1 * N2, NL = L / 2 + 1;
    static constexpr int MX = N1 > N2 ? N1 : N2;
    static constexpr int G = MX <= 16 ? 16 : 32; // lanes per line
    static constexpr int NG = 64 / G;            // lines per wavefront
    static constexpr int LB0 = TwoPass<N1, N2>::LDS;
    static constexpr int LB = LB0 > L ? LB0 : L;
    static constexpr int W = sizeof(R) == 8 ? 4 : 8; // wavefronts per block
    static constexpr int QB = W * NG;                // columns per block
    static constexpr int NLG = (NL + G - 1) / G;
    static constexpr int BUF = QB * LB > NL * QB ? QB * LB : NL * QB;
};

template <typename R, int N1, int N2, int CH, int NZA>
__global__ void __launch_bounds__((64 * Dct2Geo<R, N1, N2>::W)) k_ybwd_dct2(DctArgs a) {
    using D = Dct2Geo<R, N1, N2>;
    constexpr int L = D::L, NL = D::NL, G = D::G, NG = D::NG, LB = D::LB, QB = D::QB, MX = D::MX, NLG = D::NLG;
    __shared__ cpx<R> buf[D::BUF];
    const cpx<R> *tw = reinterpret_cast<const cpx<R> *>(a.tw);
    const int Lb = xcd_block((int)gridDim.x);
    const int f = Lb % a.nf, cg = Lb / a.nf;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane / G, t = lane - g * G;
    const int line = wv * NG + g;
    cpx<R> *B = buf + line * LB;
    const int nj = a.ny + 1;
    const R half_end = (a.ny == a.NY2) ? R(0.5) : R(1); // REDFT00 end point, present when NYS = NY2
    // the wavefront's NG spectral columns, FP64, one after the other with every lane taking part
    for (int gg = 0; gg < NG; ++gg) {
        const int qq = cg * QB + wv * NG + gg;
        const bool vq = qq < a.NQ;
        zd s[2][CH];
        if (a.curl) curl_field<CH>(a, f, qq, vq, s, lane);
        else load_col<CH>(a.spec[f] + (size_t)(vq ? qq : 0) * a.JS, lane, a.ny, s);
        cpx<R> *Bg = buf + (wv * NG + gg) * LB;
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                const int j = jof<CH>(lane, c, p);
                if (j < nj) {
                    const R sc = (j == a.NY2) ? half_end : R(1);
                    Bg[j] = vq ? cpx<R>{sc * R(s[p][c].x), sc * R(s[p][c].y)} : cpx<R>{R(0), R(0)};
                }
            }
    }
    lsync();
    cpx<R> x[MX];
    sfor<0, MX>([&](auto m) { x[m] = cpx<R>{R(0), R(0)}; });
    if (t < N1)
        sfor<0, NZA>([&](auto m) {
            const int n = t + N1 * decltype(m)::value;
            const cpx<R> v = B[n < nj ? n : 0];
            x[decltype(m)::value] = n < nj ? v : cpx<R>{R(0), R(0)};
        });
    lsync();
    two_pass_pruned<R, N1, N2, +1, NZA>(x, t, tw, B, lsync);
    if (t < N2) sfor<0, N1>([&](auto k) { B[t + N2 * decltype(k)::value] = x[k]; });
    lsync();
    cpx<R> yv[NLG];
#pragma unroll
    for (int i = 0; i < NLG; ++i) {
        const int l = t + G * i;
        const int lc = l <= L / 2 ? l : 0;
        const cpx<R> pv = B[lc], mv = B[lc == 0 ? 0 : L - lc];
        yv[i] = {R(0.5) * (pv.x + mv.x), R(0.5) * (pv.y + mv.y)};
    }
    __syncthreads(); // every line is done with its buffer: the tile may overwrite them
#pragma unroll
    for (int i = 0; i < NLG; ++i) {
        const int l = t + G * i;
        if (l < NL) buf[l * QB + line] = yv[i];
    }
    __syncthreads();
    const int q0 = cg * QB, nq = min(QB, a.NQ - q0);
    cpx<R> *Y = reinterpret_cast<cpx<R> *>(a.y) + (size_t)f * a.YS + q0;
    for (int e = threadIdx.x; e < NL * QB; e += blockDim.x) {
        const int l = e / QB, c = e - l * QB;
        if (c < nq) Y[(size_t)l * a.NQ + c] = buf[e];
    }
}

template <typename R, int N1, int N2, int NZA>
__global__ void __launch_bounds__((64 * Dct2Geo<R, N1, N2>::W)) k_yfwd_dct2(DctArgs a) {
    using D = Dct2Geo<R, N1, N2>;
    constexpr int L = D::L, NL = D::NL, G = D::G, NG = D::NG, LB = D::LB, QB = D::QB, MX = D::MX;
    __shared__ cpx<R> buf[D::BUF];
    const cpx<R> *tw = reinterpret_cast<const cpx<R> *>(a.tw);
    const int Lb = xcd_block((int)gridDim.x);
    const int f = Lb % a.nf, cg = Lb / a.nf;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane / G, t = lane - g * G;
    const int line = wv * NG + g;
    const int q0 = cg * QB, nq = min(QB, a.NQ - q0);
    {
        const cpx<R> *Y = reinterpret_cast<const cpx<R> *>(a.y) + (size_t)f * a.YS + q0;
        for (int e = threadIdx.x; e < NL * QB; e += blockDim.x) {
            const int l = e / QB, c = e - l * QB;
            buf[e] = c < nq ? Y[(size_t)l * a.NQ + c] : cpx<R>{R(0), R(0)};
        }
    }
    __syncthreads();
    // pass-A inputs of this line, weighted (1/2 at both end points)
    cpx<R> x[MX];
    sfor<0, MX>([&](auto m) { x[m] = cpx<R>{R(0), R(0)}; });
    if (t < N1)
        sfor<0, NZA>([&](auto m) {
            const int n = t + N1 * decltype(m)::value;
            const cpx<R> v = buf[(n < NL ? n : 0) * QB + line];
            const R wgt = n < NL ? ((n == 0 || n == L / 2) ? R(0.5) : R(1)) : R(0);
            x[decltype(m)::value] = {wgt * v.x, wgt * v.y};
        });
    __syncthreads(); // the tile is read: the lines may overwrite it
    cpx<R> *B = buf + line * LB;
    two_pass_pruned<R, N1, N2, -1, NZA>(x, t, tw, B, lsync);
    if (t < N2) sfor<0, N1>([&](auto k) { B[t + N2 * decltype(k)::value] = x[k]; });
    lsync();
    const int q = q0 + line;
    if (q < a.NQ) {
        const int nj = a.ny + 1;
        const double inv = 1.0 / (double)(L / 2);
        zd *out = a.spec_out[f] + (size_t)q * a.JS;
        cpx<float> *out32 = reinterpret_cast<cpx<float> *>(a.spec_out[f]) + (size_t)q * a.JS;
        const zd *add = a.add[f] ? a.add[f] + (size_t)q * a.JS : nullptr;
        for (int j = t; j < a.JS; j += G) {
            zd v = {0.0, 0.0};
            if (j < nj) {
                const cpx<R> pv = B[j], mv = B[j == 0 ? 0 : L - j];
                const double sc = j == 0 ? 0.5 * inv : inv;
                v = {sc * ((double)pv.x + (double)mv.x), sc * ((double)pv.y + (double)mv.y)};
                if (add) v = v + add[j];
            }
            if (a.out32) out32[j] = {float(v.x), float(v.y)};
            else out[j] = v;
        }
    }
}

// pass-A pruning bound: lines whose nonzero prefix fits NZ1 = ceil((L/3 + 2) / N1) entries per row
// (dealiased meshes, NYS <= 2 NY2 / 3 + 1) take the pruned kernels, others the full ones
template <int L, int N1> constexpr int nza_small() { return (L / 3 + 2 + N1 - 1) / N1; }

template <typename R, int N1, int N2> static bool go_ybwd_dct2(const DctArgs &a, hipStream_t st) {
    using D = Dct2Geo<R, N1, N2>;
    constexpr int L = N1 * N2, NZS = nza_small<L, N1>() < N2 ? nza_small<L, N1>() : N2;
    constexpr int NZF = (L / 2 + 1 + N1 - 1) / N1 < N2 ? (L / 2 + 1 + N1 - 1) / N1 : N2;
    const int ncg = (a.NQ + D::QB - 1) / D::QB;
    const dim3 grid(ncg * a.nf), block(64 * D::W);
    const bool small = a.ny + 1 <= N1 * NZS;
    if (a.CH == 1) {
        if (small) hipLaunchKernelGGL((k_ybwd_dct2<R, N1, N2, 1, NZS>), grid, block, 0, st, a);
        else hipLaunchKernelGGL((k_ybwd_dct2<R, N1, N2, 1, NZF>), grid, block, 0, st, a);
    } else {
        if (small) hipLaunchKernelGGL((k_ybwd_dct2<R, N1, N2, 2, NZS>), grid, block, 0, st, a);
        else hipLaunchKernelGGL((k_ybwd_dct2<R, N1, N2, 2, NZF>), grid, block, 0, st, a);
    }
    return true;
}
template <typename R, int N1, int N2> static bool go_yfwd_dct2(const DctArgs &a, hipStream_t st) {
    using D = Dct2Geo<R, N1, N2>;
    constexpr int L = N1 * N2, NZF = (L / 2 + 1 + N1 - 1) / N1 < N2 ? (L / 2 + 1 + N1 - 1) / N1 : N2;
    const int ncg = (a.NQ + D::QB - 1) / D::QB;
    hipLaunchKernelGGL((k_yfwd_dct2<R, N1, N2, NZF>), dim3(ncg * a.nf), dim3(64 * D::W), 0, st, a);
    return true;
}

// Line lengths L = 2 NY2 with an instantiation (even factors up to 24).
#define CDNS_DCT_SIZES(X) X(64, 8, 8) X(96, 8, 12) X(128, 8, 16) X(192, 12, 16) X(256, 16, 16) X(384, 16, 24) \
    X(480, 20, 24) X(576, 24, 24)

bool dct_supported(int L) {
#define X(NN, A, B) if (L == NN) return true;
    CDNS_DCT_SIZES(X)
#undef X
    return false;
}

bool launch_ybwd_dct(const DctArgs &a, bool fp32, hipStream_t st) {
#define X(NN, A, B) case NN: return fp32 ? go_ybwd_dct2<float, A, B>(a, st) : go_ybwd_dct2<double, A, B>(a, st);
    switch (a.L) { CDNS_DCT_SIZES(X) default: return false; }
#undef X
}
bool launch_yfwd_dct(const DctArgs &a, bool fp32, hipStream_t st) {
#define X(NN, A, B) case NN: return fp32 ? go_yfwd_dct2<float, A, B>(a, st) : go_yfwd_dct2<double, A, B>(a, st);
    switch (a.L) { CDNS_DCT_SIZES(X) default: return false; }
#undef X
}

} // namespace cdns
