// y stages as a DCT-I computed through complex FFTs of length L = 2 NY2 (register-resident
// two-pass FFTs on lane pairs), the alternative to the matrix-core Chebyshev apply of
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
// A workgroup owns Q consecutive columns q of one field (FP64 8, FP32 16, so that each row of the
// Y field is written or read as a 128-byte segment); each FFT line sits on a lane group of the
// wavefront, with each FFT row on a lane pair.  The Y side goes through an LDS tile [l][column]
// that reuses the FFT buffers; the spectral side is read and written column-contiguous.  The
// fields of one column group are neighbouring blocks on one XCD, so the columns they share stay
// in that XCD's L2.
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

constexpr int kDctWaves = 8;

template <typename R, int N1, int N2> struct DctGeo {
    static constexpr int L = N1 * N2, NL = L / 2 + 1;
    static constexpr int MX = N1 > N2 ? N1 : N2, MH = MX / 2, H1 = N1 / 2, H2 = N2 / 2;
    static constexpr int G = 2 * MX <= 16 ? 16 : (2 * MX <= 32 ? 32 : 64); // lanes per line
    static constexpr int NG = 64 / G;                                       // lines per wavefront
    static constexpr int LB0 = TwoPass<N1, N2>::LDS;
    static constexpr int LB = LB0 > L ? LB0 : L; // complex entries per line buffer
    static constexpr int LINES = kDctWaves * NG;  // lines in flight per block
    static constexpr int Q = sizeof(R) == 8 ? 8 : 16; // columns per block
    static constexpr int CPL = Q / LINES > 0 ? Q / LINES : 1; // columns per line buffer (rounds)
    static constexpr int QB = CPL * LINES;                      // columns per block (>= Q)
    static constexpr int NLG = (NL + G - 1) / G;                // output rows per lane and column
    static constexpr int BUF = LINES * LB > NL * QB ? LINES * LB : NL * QB;
};

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

// grid: (column groups x fields), XCD-ordered; block: kDctWaves wavefronts.
template <typename R, int N1, int N2, int CH>
__global__ void __launch_bounds__(64 * kDctWaves) k_ybwd_dct(DctArgs a) {
    using D = DctGeo<R, N1, N2>;
    constexpr int L = D::L, NL = D::NL, G = D::G, NG = D::NG, LB = D::LB, QB = D::QB, CPL = D::CPL;
    constexpr int H1 = D::H1, H2 = D::H2, MH = D::MH, NLG = D::NLG;
    __shared__ cpx<R> tw[L];
    __shared__ cpx<R> buf[D::BUF];
    const int Lb = xcd_block((int)gridDim.x);
    const int f = Lb % a.nf, cg = Lb / a.nf;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane / G, r = lane - g * G;
    const int t = r >> 1, h = r & 1;
    const int line = wv * NG + g;
    cpx<R> *B = buf + line * LB;
    for (int i = threadIdx.x; i < L; i += blockDim.x) tw[i] = reinterpret_cast<const cpx<R> *>(a.tw)[i];
    const int nj = a.ny + 1;
    const R half_end = (a.ny == a.NY2) ? R(0.5) : R(1); // REDFT00 end point, present when NYS = NY2
    cpx<R> yv[CPL][NLG];
    for (int cr = 0; cr < CPL; ++cr) {
        const int cl = line * CPL + cr, q = cg * QB + cl;
        const bool valid = q < a.NQ;
        // the spectral column, FP64, in the (lane, parity) layout of its own wavefront: the
        // wavefront's NG lines are formed one after the other (every lane takes part)
        for (int gg = 0; gg < NG; ++gg) {
            const int lq = (wv * NG + gg) * CPL + cr, qq = cg * QB + lq;
            const bool vq = qq < a.NQ;
            zd s[2][CH];
            if (a.curl) {
                curl_field<CH>(a, f, qq, vq, s, lane);
            } else {
                load_col<CH>(a.spec[f] + (size_t)(vq ? qq : 0) * a.JS, lane, a.ny, s);
            }
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
        if (cr == 0) __syncthreads(); // the root table
        else lsync();
        cpx<R> x[MH];
        sfor<0, MH>([&](auto m) { x[m] = cpx<R>{R(0), R(0)}; });
        if (t < N1)
            sfor<0, H2>([&](auto m) {
                const int n = t + N1 * (2 * decltype(m)::value + h);
                const cpx<R> v = B[n < nj ? n : 0];
                x[m] = n < nj ? v : cpx<R>{R(0), R(0)};
            });
        lsync();
        two_pass_pair<R, N1, N2, +1, false>(x, t, h, tw, B, lsync);
        if (t < N2) sfor<0, H1>([&](auto k) { B[t + N2 * (H1 * h + decltype(k)::value)] = x[k]; });
        lsync();
#pragma unroll
        for (int i = 0; i < NLG; ++i) {
            const int l = r + G * i;
            const int lc = l <= L / 2 ? l : 0;
            const cpx<R> p = B[lc], m = B[lc == 0 ? 0 : L - lc];
            yv[cr][i] = {R(0.5) * (p.x + m.x), R(0.5) * (p.y + m.y)};
        }
        lsync();
    }
    __syncthreads(); // every line is done with its buffer: the tile may overwrite them
    for (int cr = 0; cr < CPL; ++cr) {
        const int cl = line * CPL + cr;
#pragma unroll
        for (int i = 0; i < NLG; ++i) {
            const int l = r + G * i;
            if (l < NL) buf[l * QB + cl] = yv[cr][i];
        }
    }
    __syncthreads();
    const int q0 = cg * QB, nq = min(QB, a.NQ - q0);
    cpx<R> *Y = reinterpret_cast<cpx<R> *>(a.y) + (size_t)f * a.YS + q0;
    for (int e = threadIdx.x; e < NL * QB; e += blockDim.x) {
        const int l = e / QB, c = e - l * QB;
        if (c < nq) Y[(size_t)l * a.NQ + c] = buf[e];
    }
}

// grid: (column groups x fields), XCD-ordered.
template <typename R, int N1, int N2>
__global__ void __launch_bounds__(64 * kDctWaves) k_yfwd_dct(DctArgs a) {
    using D = DctGeo<R, N1, N2>;
    constexpr int L = D::L, NL = D::NL, G = D::G, NG = D::NG, LB = D::LB, QB = D::QB, CPL = D::CPL;
    constexpr int H1 = D::H1, H2 = D::H2, MH = D::MH;
    __shared__ cpx<R> tw[L];
    __shared__ cpx<R> buf[D::BUF];
    const int Lb = xcd_block((int)gridDim.x);
    const int f = Lb % a.nf, cg = Lb / a.nf;
    const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, g = lane / G, r = lane - g * G;
    const int t = r >> 1, h = r & 1;
    const int line = wv * NG + g;
    const int q0 = cg * QB, nq = min(QB, a.NQ - q0);
    for (int i = threadIdx.x; i < L; i += blockDim.x) tw[i] = reinterpret_cast<const cpx<R> *>(a.tw)[i];
    {
        const cpx<R> *Y = reinterpret_cast<const cpx<R> *>(a.y) + (size_t)f * a.YS + q0;
        for (int e = threadIdx.x; e < NL * QB; e += blockDim.x) {
            const int l = e / QB, c = e - l * QB;
            buf[e] = c < nq ? Y[(size_t)l * a.NQ + c] : cpx<R>{R(0), R(0)};
        }
    }
    __syncthreads();
    // pass-A inputs of every column of this line, weighted (1/2 at both end points)
    cpx<R> xin[CPL][MH];
    for (int cr = 0; cr < CPL; ++cr) {
        const int cl = line * CPL + cr;
        sfor<0, MH>([&](auto m) { xin[cr][m] = cpx<R>{R(0), R(0)}; });
        if (t < N1)
            sfor<0, H2>([&](auto m) {
                const int n = t + N1 * (2 * decltype(m)::value + h);
                const cpx<R> v = buf[(n < NL ? n : 0) * QB + cl];
                const R wgt = n < NL ? ((n == 0 || n == L / 2) ? R(0.5) : R(1)) : R(0);
                xin[cr][decltype(m)::value] = {wgt * v.x, wgt * v.y};
            });
    }
    __syncthreads(); // the tile is read: the lines may overwrite it
    cpx<R> *B = buf + line * LB;
    const int nj = a.ny + 1;
    const double inv = 1.0 / (double)(L / 2);
    for (int cr = 0; cr < CPL; ++cr) {
        const int q = q0 + line * CPL + cr;
        const bool valid = q < a.NQ;
        cpx<R> x[MH];
        sfor<0, MH>([&](auto m) { x[m] = xin[cr][m]; });
        two_pass_pair<R, N1, N2, -1, false>(x, t, h, tw, B, lsync);
        if (t < N2) sfor<0, H1>([&](auto k) { B[t + N2 * (H1 * h + decltype(k)::value)] = x[k]; });
        lsync();
        if (valid) {
            zd *out = a.spec_out[f] + (size_t)q * a.JS;
            const zd *add = a.add[f] ? a.add[f] + (size_t)q * a.JS : nullptr;
            for (int j = r; j < a.JS; j += G) {
                zd v = {0.0, 0.0};
                if (j < nj) {
                    const cpx<R> p = B[j], m = B[j == 0 ? 0 : L - j];
                    const double sc = j == 0 ? 0.5 * inv : inv;
                    v = {sc * ((double)p.x + (double)m.x), sc * ((double)p.y + (double)m.y)};
                    if (add) v = v + add[j];
                }
                out[j] = v;
            }
        }
        lsync();
    }
}

template <typename R, int N1, int N2> static void go_ybwd_dct(const DctArgs &a, hipStream_t st) {
    using D = DctGeo<R, N1, N2>;
    const int ncg = (a.NQ + D::QB - 1) / D::QB;
    if (a.CH == 1) hipLaunchKernelGGL((k_ybwd_dct<R, N1, N2, 1>), dim3(ncg * a.nf), dim3(64 * kDctWaves), 0, st, a);
    else hipLaunchKernelGGL((k_ybwd_dct<R, N1, N2, 2>), dim3(ncg * a.nf), dim3(64 * kDctWaves), 0, st, a);
}
template <typename R, int N1, int N2> static void go_yfwd_dct(const DctArgs &a, hipStream_t st) {
    using D = DctGeo<R, N1, N2>;
    const int ncg = (a.NQ + D::QB - 1) / D::QB;
    hipLaunchKernelGGL((k_yfwd_dct<R, N1, N2>), dim3(ncg * a.nf), dim3(64 * kDctWaves), 0, st, a);
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
#define X(NN, A, B) case NN: if (fp32) go_ybwd_dct<float, A, B>(a, st); else go_ybwd_dct<double, A, B>(a, st); return true;
    switch (a.L) { CDNS_DCT_SIZES(X) default: return false; }
#undef X
}
bool launch_yfwd_dct(const DctArgs &a, bool fp32, hipStream_t st) {
#define X(NN, A, B) case NN: if (fp32) go_yfwd_dct<float, A, B>(a, st); else go_yfwd_dct<double, A, B>(a, st); return true;
    switch (a.L) { CDNS_DCT_SIZES(X) default: return false; }
#undef X
}

} // namespace cdns
