// Product-grid (3/2-dealiased) transforms of the channel-dns step on gfx950.
//
// The reference runs y REDFT00 -> identity transpose -> z c2c -> x c2r per component
// (src/fft/transforms.cpp:687-740) and the inverse chain (641-685), with the point-wise
// cross product as a separate pass (src/solver/vector_ops.cpp:250-272).  Here:
//   y stage : dense Chebyshev matrix apply split by parity (T_j(-y) = (-1)^j T_j(y)),
//             LDS-tiled, one complex column per lane;
//   z stage : mixed-radix Stockham FFT over kz in LDS, kx lines batched per workgroup;
//   x stage : c2r of (u,v,w,wx,wy,wz) packed two real lines per complex FFT, the cross
//             product in LDS, and r2c of the three products, in one kernel, so that the
//             nine physical fields never reach HBM.
// Normalisation follows transforms.cpp exactly: backward synthesis is the exact value of
// the truncated series; forward applies 1/NX2 (no Nyquist halving), 1/NZ2 and the DCT-I
// weights 1/(2 NY2) (iy=0), 1/NY2 (iy>=1).
#include "internal.hpp"

namespace cdns {

// ============================================================ y backward ====
// Y[f][l][q] = sum_j Bm[j][l'] * S_f[q][j], l' = min(l, NY2-l), sign (-1)^j for l > NY2/2.
template <typename R, int LW>
__global__ void __launch_bounds__(512) k_ybwd(YArgs a) {
    extern __shared__ unsigned char smem_raw[];
    constexpr int JC = 32, XP = 65;
    cpx<R> *X = reinterpret_cast<cpx<R> *>(smem_raw);
    R *Bc = reinterpret_cast<R *>(X + JC * XP);
    const int f = blockIdx.y;
    const int q0 = blockIdx.x * 64;
    const int wv = threadIdx.x >> 6, t = threadIdx.x & 63;
    const int l0 = wv * LW;
    const zd *__restrict__ in = a.spec[f];
    const int nj = a.nj, NLH = a.NLH;
    cpx<R> accE[LW], accO[LW];
#pragma unroll
    for (int k = 0; k < LW; ++k) { accE[k] = {R(0), R(0)}; accO[k] = {R(0), R(0)}; }
    // zero the coefficient padding once
    for (int e = threadIdx.x; e < LW + 2; e += blockDim.x) Bc[JC * NLH + e] = R(0);
    for (int j0 = 0; j0 < nj; j0 += JC) {
        for (int e = threadIdx.x; e < 64 * JC; e += blockDim.x) {
            const int ql = e / JC, jj = e - ql * JC;
            const int q = q0 + ql, j = j0 + jj;
            zd v = {0.0, 0.0};
            if (q < a.NQ && j < nj) v = in[(size_t)q * a.JS + j];
            X[jj * XP + ql] = {R(v.x), R(v.y)};
        }
        for (int e = threadIdx.x; e < JC * NLH; e += blockDim.x) {
            const int jj = e / NLH, lp = e - jj * NLH;
            const int j = j0 + jj;
            Bc[e] = j < nj ? R(a.coef[(size_t)j * NLH + lp]) : R(0);
        }
        __syncthreads();
        const int jmax = min(JC, nj - j0);
        for (int jj = 0; jj + 1 < jmax; jj += 2) {
            const cpx<R> xe = X[jj * XP + t], xo = X[(jj + 1) * XP + t];
            const R *be = Bc + jj * NLH + l0, *bo = Bc + (jj + 1) * NLH + l0;
#pragma unroll
            for (int k = 0; k < LW; ++k) {
                accE[k].x += be[k] * xe.x;
                accE[k].y += be[k] * xe.y;
                accO[k].x += bo[k] * xo.x;
                accO[k].y += bo[k] * xo.y;
            }
        }
        if (jmax & 1) {
            const int jj = jmax - 1; // even j (j0 and JC are even)
            const cpx<R> xe = X[jj * XP + t];
            const R *be = Bc + jj * NLH + l0;
#pragma unroll
            for (int k = 0; k < LW; ++k) {
                accE[k].x += be[k] * xe.x;
                accE[k].y += be[k] * xe.y;
            }
        }
        __syncthreads();
    }
    const int q = q0 + t;
    if (q >= a.NQ) return;
    cpx<R> *out = reinterpret_cast<cpx<R> *>(a.y) + (size_t)f * a.YS;
#pragma unroll
    for (int k = 0; k < LW; ++k) {
        const int lp = l0 + k;
        if (lp < NLH) {
            out[(size_t)lp * a.NQ + q] = accE[k] + accO[k];
            const int lm = a.NY2 - lp;
            if (lm != lp) out[(size_t)lm * a.NQ + q] = accE[k] - accO[k];
        }
    }
}

// ============================================================= y forward ====
// S_f[q][j] = sum_{l'} F_p[l'][jj] * (Y[l'][q] +- Y[NY2-l'][q]), j = 2 jj + p.
template <typename R, int C>
__global__ void __launch_bounds__(512) k_yfwd(YArgs a) {
    extern __shared__ unsigned char smem_raw[];
    constexpr int LC = 16;
    const int NWJ = a.nwj, NG = a.ng, TQ = NG * C, JL = 64 * NWJ;
    cpx<R> *SD = reinterpret_cast<cpx<R> *>(smem_raw); // [2][LC][TQ]
    R *F = reinterpret_cast<R *>(SD + 2 * LC * TQ);    // [2][LC][JL]
    const int f = blockIdx.y;
    const int q0 = blockIdx.x * TQ;
    const int wv = threadIdx.x >> 6, t = threadIdx.x & 63;
    const int p = wv & 1, js = (wv >> 1) % NWJ, g = wv / (2 * NWJ);
    const int jj = js * 64 + t, j = 2 * jj + p;
    const cpx<R> *__restrict__ Y = reinterpret_cast<const cpx<R> *>(a.y) + (size_t)f * a.YS;
    const int NLH = a.NLH, NY2 = a.NY2;
    cpx<R> acc[C];
#pragma unroll
    for (int c = 0; c < C; ++c) acc[c] = {R(0), R(0)};
    for (int l0 = 0; l0 < NLH; l0 += LC) {
        for (int e = threadIdx.x; e < LC * TQ; e += blockDim.x) {
            const int lc = e / TQ, ql = e - lc * TQ;
            const int lp = l0 + lc, q = q0 + ql;
            cpx<R> s = {R(0), R(0)}, d = {R(0), R(0)};
            if (lp < NLH && q < a.NQ) {
                const cpx<R> y1 = Y[(size_t)lp * a.NQ + q];
                const int lm = NY2 - lp;
                if (lm != lp) {
                    const cpx<R> y2 = Y[(size_t)lm * a.NQ + q];
                    s = y1 + y2;
                    d = y1 - y2;
                } else {
                    s = y1;
                }
            }
            SD[lc * TQ + ql] = s;
            SD[LC * TQ + lc * TQ + ql] = d;
        }
        for (int e = threadIdx.x; e < 2 * LC * JL; e += blockDim.x) {
            const int pp = e / (LC * JL), r = e - pp * LC * JL, lc = r / JL, jl = r - lc * JL;
            const int lp = l0 + lc;
            const int nje = pp ? a.njo : a.nje;
            R v = R(0);
            if (lp < NLH && jl < nje) v = R((pp ? a.coef2 : a.coef)[(size_t)lp * nje + jl]);
            F[e] = v;
        }
        __syncthreads();
        const int lmax = min(LC, NLH - l0);
        const cpx<R> *sd = SD + p * LC * TQ + g * C;
        const R *fp = F + p * LC * JL + jj;
        for (int lc = 0; lc < lmax; ++lc) {
            const R fv = fp[lc * JL];
            const cpx<R> *row = sd + lc * TQ;
#pragma unroll
            for (int c = 0; c < C; ++c) {
                const cpx<R> sv = row[c];
                acc[c].x += fv * sv.x;
                acc[c].y += fv * sv.y;
            }
        }
        __syncthreads();
    }
    if (j >= a.nj) return;
    zd *out = a.spec_out[f];
    cpx<float> *out32 = reinterpret_cast<cpx<float> *>(a.spec_out[f]);
    const zd *add = a.add[f];
#pragma unroll
    for (int c = 0; c < C; ++c) {
        const int q = q0 + g * C + c;
        if (q < a.NQ) {
            zd v = {double(acc[c].x), double(acc[c].y)};
            if (add) v = v + add[(size_t)q * a.JS + j];
            if (a.out32) out32[(size_t)q * a.JS + j] = {float(v.x), float(v.y)};
            else out[(size_t)q * a.JS + j] = v;
        }
    }
}

// ============================================================== z stages ====
// backward: Y[f][l][kx*nzk + kzi] -> slots, FFT(+), Z[f][l][iz][kx]
template <typename R, int IT>
__global__ void __launch_bounds__(256) k_zbwd(ZArgs a) {
    extern __shared__ unsigned char smem_raw[];
    cpx<R> *buf = reinterpret_cast<cpx<R> *>(smem_raw);
    const int kt0 = blockIdx.x * a.KT, l = blockIdx.y, f = blockIdx.z;
    const int nl = min(a.KT, a.nkx - kt0);
    const int N = a.plan.n, NZS = a.NZS;
    const cpx<R> *__restrict__ Y = reinterpret_cast<const cpx<R> *>(a.in) + (size_t)f * a.in_fs + (size_t)l * a.NQ;
    for (int e = threadIdx.x; e < nl * N; e += blockDim.x) {
        const int ln = e / N, s = e - ln * N;
        const int kx = kt0 + ln;
        // slot s holds kz = s (s <= NZS) or s - N (s >= N - NZS); at an overlap the slot
        // N - NZS takes kz = -NZS (transforms.cpp:717-726)
        int kz = 1 << 30;
        if (s <= NZS) kz = s;
        else if (s >= N - NZS) kz = s - N;
        if (s == N - NZS && s <= NZS) kz = -NZS;
        cpx<R> v = {R(0), R(0)};
        if (kz != (1 << 30)) v = Y[(size_t)kx * a.nzk + kz + NZS];
        buf[e] = v;
    }
    __syncthreads();
    fft_block<R, IT>(buf, nl, a.plan, reinterpret_cast<const cpx<R> *>(a.tw), +1);
    cpx<R> *Z = reinterpret_cast<cpx<R> *>(a.out) + (size_t)f * a.out_fs + (size_t)l * N * a.nkx;
    for (int e = threadIdx.x; e < nl * N; e += blockDim.x) {
        const int iz = e / nl, ln = e - iz * nl;
        Z[(size_t)iz * a.nkx + kt0 + ln] = buf[ln * N + iz];
    }
}

// forward: Z[f][l][kx][iz] (line-major) -> FFT(-) -> Y[f][l][kx*nzk + kzi] = X(slot)/NZ2
template <typename R, int IT>
__global__ void __launch_bounds__(256) k_zfwd(ZArgs a) {
    extern __shared__ unsigned char smem_raw[];
    cpx<R> *buf = reinterpret_cast<cpx<R> *>(smem_raw);
    const int kt0 = blockIdx.x * a.KT, l = blockIdx.y, f = blockIdx.z;
    const int nl = min(a.KT, a.nkx - kt0);
    const int N = a.plan.n, NZS = a.NZS;
    // line-major input [l][kx][iz]: the block's lines are one contiguous run
    const cpx<R> *__restrict__ Z = reinterpret_cast<const cpx<R> *>(a.in) + (size_t)f * a.in_fs +
                                   ((size_t)l * a.nkx + kt0) * N;
    for (int e = threadIdx.x; e < nl * N; e += blockDim.x) buf[e] = Z[e];
    __syncthreads();
    fft_block<R, IT>(buf, nl, a.plan, reinterpret_cast<const cpx<R> *>(a.tw), -1);
    const R inv = R(1.0 / N);
    cpx<R> *Y = reinterpret_cast<cpx<R> *>(a.out) + (size_t)f * a.out_fs + (size_t)l * a.NQ;
    for (int e = threadIdx.x; e < nl * a.nzk; e += blockDim.x) {
        const int ln = e / a.nzk, kzi = e - ln * a.nzk;
        const int kz = kzi - NZS;
        const int s = kz >= 0 ? kz : N + kz;
        Y[(size_t)(kt0 + ln) * a.nzk + kzi] = inv * buf[ln * N + s];
    }
}

// ===================================================== x stage + product ====
// Hermitian line k -> complex slot value for packing two real syntheses A + iB.
template <typename R>
__device__ inline cpx<R> pack_c2r(const cpx<R> *A, const cpx<R> *B, int k, int N, int K) {
    // FFTW c2r ignores the imaginary parts of the DC and (even N) Nyquist inputs.
    cpx<R> a = {R(0), R(0)}, b = {R(0), R(0)};
    bool hit = false, mirror = false;
    int src = k;
    if (k <= K) hit = true;
    else if (N - k <= K) { hit = true; mirror = true; src = N - k; }
    if (!hit) return {R(0), R(0)};
    a = A[src];
    b = B[src];
    if (src == 0 || 2 * src == N) { a.y = R(0); b.y = R(0); }
    if (mirror) { a.y = -a.y; b.y = -b.y; }
    return {a.x - b.y, a.y + b.x};
}

template <typename R>
__device__ inline cpx<R> pack_c2r_single(const cpx<R> *A, int k, int N, int K) {
    int src = k;
    bool mirror = false;
    if (k > K) {
        if (N - k > K) return {R(0), R(0)};
        mirror = true;
        src = N - k;
    }
    cpx<R> a = A[src];
    if (src == 0 || 2 * src == N) a.y = R(0);
    if (mirror) a.y = -a.y;
    return a;
}

// One block: P points (l, iz); input fields fin[0..5] = a (3 components), b (3 components).
template <typename R, int IT>
__global__ void __launch_bounds__(256) k_xprod(XArgs a) {
    extern __shared__ unsigned char smem_raw[];
    cpx<R> *buf = reinterpret_cast<cpx<R> *>(smem_raw);
    const int P = a.P, N = a.plan.n, K = a.nkx - 1;
    const int np = a.NL * a.NZ2;
    const int p0 = blockIdx.x * P;
    // 1. pack (u,v), (w,wx), (wy,wz) per point
    for (int e = threadIdx.x; e < 3 * P * N; e += blockDim.x) {
        const int line = e / N, k = e - line * N;
        const int pp = line / 3, gsel = line - pp * 3;
        const int pt = p0 + pp;
        cpx<R> v = {R(0), R(0)};
        if (pt < np) {
            const cpx<R> *A = reinterpret_cast<const cpx<R> *>(a.fin[2 * gsel]) + (size_t)pt * a.nkx;
            const cpx<R> *B = reinterpret_cast<const cpx<R> *>(a.fin[2 * gsel + 1]) + (size_t)pt * a.nkx;
            v = pack_c2r(A, B, k, N, K);
        }
        buf[e] = v;
    }
    __syncthreads();
    fft_block<R, IT>(buf, 3 * P, a.plan, reinterpret_cast<const cpx<R> *>(a.tw), +1);
    // 2. products; point pair (2h, 2h+1) -> forward lines 3h+0 (Nx,Ny of 2h), 3h+1 (Nx,Ny of 2h+1),
    //    3h+2 (Nz of 2h, Nz of 2h+1)
    constexpr int MAXI = 8;
    cpx<R> o0[MAXI], o1[MAXI], o2[MAXI];
    const int items = (P / 2) * N;
#pragma unroll
    for (int it = 0; it < MAXI; ++it) {
        const int e = threadIdx.x + it * blockDim.x;
        if (e < items) {
            const int h = e / N, n = e - h * N;
            R nx[2], ny[2], nz[2];
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const int pp = 2 * h + s;
                const cpx<R> c0 = buf[(3 * pp + 0) * N + n], c1 = buf[(3 * pp + 1) * N + n], c2 = buf[(3 * pp + 2) * N + n];
                const R u = c0.x, v = c0.y, w = c1.x, ox = c1.y, oy = c2.x, oz = c2.y;
                nx[s] = v * oz - w * oy;
                ny[s] = w * ox - u * oz;
                nz[s] = u * oy - v * ox;
            }
            o0[it] = {nx[0], ny[0]};
            o1[it] = {nx[1], ny[1]};
            o2[it] = {nz[0], nz[1]};
        }
    }
    __syncthreads();
#pragma unroll
    for (int it = 0; it < MAXI; ++it) {
        const int e = threadIdx.x + it * blockDim.x;
        if (e < items) {
            const int h = e / N, n = e - h * N;
            buf[(3 * h + 0) * N + n] = o0[it];
            buf[(3 * h + 1) * N + n] = o1[it];
            buf[(3 * h + 2) * N + n] = o2[it];
        }
    }
    __syncthreads();
    fft_block<R, IT>(buf, 3 * (P / 2), a.plan, reinterpret_cast<const cpx<R> *>(a.tw), -1);
    // 3. unpack two real spectra per line, x 1/NX2, keep kx <= NXS
    const R half = R(0.5) / R(N);
    for (int e = threadIdx.x; e < 3 * (P / 2) * (K + 1); e += blockDim.x) {
        const int line = e / (K + 1), k = e - line * (K + 1);
        const int h = line / 3, sel = line - h * 3;
        const cpx<R> F = buf[line * N + k], G = conjc(buf[line * N + (k == 0 ? 0 : N - k)]);
        const cpx<R> s = F + G, d = F - G;
        const cpx<R> first = {half * s.x, half * s.y};   // (F + conj F_{N-k}) / 2
        const cpx<R> second = {half * d.y, -half * d.x}; // (F - conj F_{N-k}) / (2i)
        int pa, ca, pb, cb;
        if (sel < 2) { pa = pb = 2 * h + sel; ca = 0; cb = 1; }
        else { pa = 2 * h; pb = 2 * h + 1; ca = cb = 2; }
        const int pta = p0 + pa, ptb = p0 + pb;
        if (pta < np) reinterpret_cast<cpx<R> *>(a.fout[ca])[zline_index(pta, k, a.nkx, a.NZ2)] = first;
        if (ptb < np) reinterpret_cast<cpx<R> *>(a.fout[cb])[zline_index(ptb, k, a.nkx, a.NZ2)] = second;
    }
}

// Single-field x stages against the reference physical layout [iz][iy][ix] (for the
// channel_xyzfft2 entry point).  Two points share one complex FFT.
template <typename R, int IT>
__global__ void __launch_bounds__(256) k_xbwd_phys(XArgs a) {
    extern __shared__ unsigned char smem_raw[];
    cpx<R> *buf = reinterpret_cast<cpx<R> *>(smem_raw);
    const int P = a.P, N = a.plan.n, K = a.nkx - 1, np = a.NL * a.NZ2;
    const int p0 = blockIdx.x * P;
    const cpx<R> *__restrict__ Z = reinterpret_cast<const cpx<R> *>(a.in);
    const cpx<R> zero = {R(0), R(0)};
    for (int e = threadIdx.x; e < (P / 2) * N; e += blockDim.x) {
        const int h = e / N, k = e - h * N;
        const int pa = p0 + 2 * h, pb = pa + 1;
        cpx<R> v = zero;
        if (pb < np) {
            v = pack_c2r(Z + (size_t)pa * a.nkx, Z + (size_t)pb * a.nkx, k, N, K);
        } else if (pa < np) {
            // odd tail: the second line of the pair is zero
            const cpx<R> *A = Z + (size_t)pa * a.nkx;
            v = pack_c2r_single(A, k, N, K);
        }
        buf[e] = v;
    }
    __syncthreads();
    fft_block<R, IT>(buf, P / 2, a.plan, reinterpret_cast<const cpx<R> *>(a.tw), +1);
    double *phys = a.phys;
    for (int e = threadIdx.x; e < P * N; e += blockDim.x) {
        const int pp = e / N, n = e - pp * N;
        const int pt = p0 + pp;
        if (pt >= np) continue;
        const int l = pt / a.NZ2, iz = pt - l * a.NZ2;
        const cpx<R> c = buf[(pp >> 1) * N + n];
        phys[((size_t)iz * a.NL + l) * N + n] = double((pp & 1) ? c.y : c.x);
    }
}

template <typename R, int IT>
__global__ void __launch_bounds__(256) k_xfwd_phys(XArgs a) {
    extern __shared__ unsigned char smem_raw[];
    cpx<R> *buf = reinterpret_cast<cpx<R> *>(smem_raw);
    const int P = a.P, N = a.plan.n, K = a.nkx - 1, np = a.NL * a.NZ2;
    const int p0 = blockIdx.x * P;
    const double *phys = a.phys;
    for (int e = threadIdx.x; e < (P / 2) * N; e += blockDim.x) {
        const int h = e / N, n = e - h * N;
        R re = R(0), im = R(0);
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const int pt = p0 + 2 * h + s;
            if (pt < np) {
                const int l = pt / a.NZ2, iz = pt - l * a.NZ2;
                const R v = R(phys[((size_t)iz * a.NL + l) * N + n]);
                if (s == 0) re = v; else im = v;
            }
        }
        buf[e] = {re, im};
    }
    __syncthreads();
    fft_block<R, IT>(buf, P / 2, a.plan, reinterpret_cast<const cpx<R> *>(a.tw), -1);
    const R half = R(0.5) / R(N);
    cpx<R> *ZN = reinterpret_cast<cpx<R> *>(a.out);
    for (int e = threadIdx.x; e < (P / 2) * (K + 1); e += blockDim.x) {
        const int h = e / (K + 1), k = e - h * (K + 1);
        const cpx<R> F = buf[h * N + k], G = conjc(buf[h * N + (k == 0 ? 0 : N - k)]);
        const cpx<R> s = F + G, d = F - G;
        const int pa = p0 + 2 * h, pb = pa + 1;
        if (pa < np) ZN[zline_index(pa, k, a.nkx, a.NZ2)] = {half * s.x, half * s.y};
        if (pb < np) ZN[zline_index(pb, k, a.nkx, a.NZ2)] = {half * d.y, -half * d.x};
    }
}

// ================================================================ launch ====
template <typename R> static void launch_ybwd_t(const YArgs &a, int nfields, hipStream_t st) {
    constexpr int LW = 26;
    const int nw = (a.NLH + LW - 1) / LW;
    const size_t sm = 32 * 65 * sizeof(cpx<R>) + (32 * a.NLH + LW + 2) * sizeof(R);
    dim3 grid((a.NQ + 63) / 64, nfields), block(64 * nw);
    hipLaunchKernelGGL((k_ybwd<R, LW>), grid, block, sm, st, a);
}
void launch_ybwd(const YArgs &a, int nfields, bool fp32, hipStream_t st) {
    if (a.mfma && launch_ybwd_mfma(a, nfields, fp32, st)) return;
    if (fp32) launch_ybwd_t<float>(a, nfields, st);
    else launch_ybwd_t<double>(a, nfields, st);
}

template <typename R> static void launch_yfwd_t(YArgs a, int nfields, hipStream_t st) {
    constexpr int C = 32;
    const int nje = (a.nj + 1) / 2;
    a.nwj = (nje + 63) / 64;
    a.ng = 2;
    const int TQ = a.ng * C;
    const size_t sm = 2 * 16 * TQ * sizeof(cpx<R>) + 2 * 16 * 64 * a.nwj * sizeof(R);
    dim3 grid((a.NQ + TQ - 1) / TQ, nfields), block(64 * 2 * a.nwj * a.ng);
    hipLaunchKernelGGL((k_yfwd<R, C>), grid, block, sm, st, a);
}
void launch_yfwd(const YArgs &a, int nfields, bool fp32, hipStream_t st) {
    if (a.mfma && launch_yfwd_mfma(a, nfields, fp32, st)) return;
    if (fp32) launch_yfwd_t<float>(a, nfields, st);
    else launch_yfwd_t<double>(a, nfields, st);
}

#define CDNS_DISPATCH_IT(KERNEL, R, IT, GRID, SM, ST, ARGS)                                                 \
    switch (IT) {                                                                                            \
    case 1: hipLaunchKernelGGL((KERNEL<R, 1>), GRID, dim3(256), SM, ST, ARGS); break;                        \
    case 2: hipLaunchKernelGGL((KERNEL<R, 2>), GRID, dim3(256), SM, ST, ARGS); break;                        \
    case 3: hipLaunchKernelGGL((KERNEL<R, 3>), GRID, dim3(256), SM, ST, ARGS); break;                        \
    case 4: hipLaunchKernelGGL((KERNEL<R, 4>), GRID, dim3(256), SM, ST, ARGS); break;                        \
    default: throw std::runtime_error("FFT batch too large for one workgroup");                             \
    }

static int items_per_thread(int nlines, const FftPlan &p) {
    return (nlines * (p.n / plan_min_radix(p)) + 255) / 256;
}

void launch_zbwd(const ZArgs &a, int nfields, bool fp32, hipStream_t st) {
    if (a.rf && launch_zbwd_rf(a, nfields, fp32, st)) return;
    const int it = items_per_thread(a.KT, a.plan);
    dim3 grid((a.nkx + a.KT - 1) / a.KT, a.NL, nfields);
    if (fp32) {
        const size_t sm = (size_t)a.KT * a.plan.n * sizeof(cpx<float>);
        CDNS_DISPATCH_IT(k_zbwd, float, it, grid, sm, st, a);
    } else {
        const size_t sm = (size_t)a.KT * a.plan.n * sizeof(cpx<double>);
        CDNS_DISPATCH_IT(k_zbwd, double, it, grid, sm, st, a);
    }
}
void launch_zfwd(const ZArgs &a, int nfields, bool fp32, hipStream_t st) {
    if (a.rf && launch_zfwd_rf(a, nfields, fp32, st)) return;
    const int it = items_per_thread(a.KT, a.plan);
    dim3 grid((a.nkx + a.KT - 1) / a.KT, a.NL, nfields);
    if (fp32) {
        const size_t sm = (size_t)a.KT * a.plan.n * sizeof(cpx<float>);
        CDNS_DISPATCH_IT(k_zfwd, float, it, grid, sm, st, a);
    } else {
        const size_t sm = (size_t)a.KT * a.plan.n * sizeof(cpx<double>);
        CDNS_DISPATCH_IT(k_zfwd, double, it, grid, sm, st, a);
    }
}
void launch_xprod(const XArgs &a, bool fp32, hipStream_t st) {
    if (a.rf && launch_xprod_rf(a, fp32, st)) return;
    const int it = items_per_thread(3 * a.P, a.plan);
    if ((a.P / 2) * a.plan.n > 8 * 256) throw std::runtime_error("x-product batch too large");
    dim3 grid((a.NL * a.NZ2 + a.P - 1) / a.P);
    if (fp32) {
        const size_t sm = (size_t)3 * a.P * a.plan.n * sizeof(cpx<float>);
        CDNS_DISPATCH_IT(k_xprod, float, it, grid, sm, st, a);
    } else {
        const size_t sm = (size_t)3 * a.P * a.plan.n * sizeof(cpx<double>);
        CDNS_DISPATCH_IT(k_xprod, double, it, grid, sm, st, a);
    }
}
void launch_xbwd_phys(const XArgs &a, bool fp32, hipStream_t st) {
    if (a.rf && launch_xbwd_phys_rf(a, fp32, st)) return;
    const int it = items_per_thread(a.P / 2, a.plan);
    dim3 grid((a.NL * a.NZ2 + a.P - 1) / a.P);
    if (fp32) {
        CDNS_DISPATCH_IT(k_xbwd_phys, float, it, grid, (size_t)(a.P / 2) * a.plan.n * sizeof(cpx<float>), st, a);
    } else {
        CDNS_DISPATCH_IT(k_xbwd_phys, double, it, grid, (size_t)(a.P / 2) * a.plan.n * sizeof(cpx<double>), st, a);
    }
}
void launch_xfwd_phys(const XArgs &a, bool fp32, hipStream_t st) {
    if (a.rf && launch_xfwd_phys_rf(a, fp32, st)) return;
    const int it = items_per_thread(a.P / 2, a.plan);
    dim3 grid((a.NL * a.NZ2 + a.P - 1) / a.P);
    if (fp32) {
        CDNS_DISPATCH_IT(k_xfwd_phys, float, it, grid, (size_t)(a.P / 2) * a.plan.n * sizeof(cpx<float>), st, a);
    } else {
        CDNS_DISPATCH_IT(k_xfwd_phys, double, it, grid, (size_t)(a.P / 2) * a.plan.n * sizeof(cpx<double>), st, a);
    }
}

} // namespace cdns
