// Diagnostics and statistics of the channel-dns state on the coarse physical grid (NX x (NY+1) x
// NZ): the spectral divergence, finiteness of spectral planes, and the per-plane moment sums of
// the physical velocity and pressure (Statistics::reduce, Diagnostics::evaluate).
#include "internal.hpp"
#include "split.hpp"

namespace cdns {

// vector_ops::divergence (vector_ops.cpp:90-105): i kx u + D v + i kz w; one wavefront per column.
template <int NE>
__global__ void __launch_bounds__(256) k_divergence(DiagArgs a) {
    const int lane = lane_id();
    const int q = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (q >= a.NQ) return;
    const int kxi = q / a.nzk, kzi = q - kxi * a.nzk;
    const double kx = a.WAVX[kxi], kz = a.WAVZ[kzi];
    const size_t off = (size_t)q * a.JS;
    zd U[NE], V[NE], W[NE], D[NE];
    sp::load_col<NE>(a.u + off, lane, a.ny, U);
    sp::load_col<NE>(a.v + off, lane, a.ny, V);
    sp::load_col<NE>(a.w + off, lane, a.ny, W);
    sp::cheb_dy<NE>(V, D, lane, a.ny);
#pragma unroll
    for (int e = 0; e < NE; ++e) D[e] = ((zd{0.0, 0.0} + ikmul(kx, U[e])) + D[e]) + ikmul(kz, W[e]);
    sp::store_col<NE>(a.out + off, lane, a.ny, D);
}
void launch_divergence(const DiagArgs &a, int E, hipStream_t st) {
    dim3 grid((a.NQ + 3) / 4), block(256);
    if (E == 2) hipLaunchKernelGGL((k_divergence<2>), grid, block, 0, st, a);
    else if (E == 4) hipLaunchKernelGGL((k_divergence<4>), grid, block, 0, st, a);
    else throw std::runtime_error("unsupported column length");
}

// flags[b] = 1 when every coefficient j <= ny of the planes handled by block b is finite
__global__ void __launch_bounds__(256) k_planes_finite(const double *planes, size_t nplanes, int NQ, int JS, int ny,
                                                       int *flags) {
    const size_t n = nplanes * NQ * JS;
    bool ok = true;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int j = (int)(i % JS);
        if (j <= ny) {
            const double2 v = reinterpret_cast<const double2 *>(planes)[i];
            ok = ok && isfinite(v.x) && isfinite(v.y);
        }
    }
    const int all = __syncthreads_and(ok ? 1 : 0);
    if (threadIdx.x == 0) flags[blockIdx.x] = all;
}
void launch_planes_finite(const zd *planes, size_t nplanes, int NQ, int JS, int ny, int *flags, int nblocks,
                          hipStream_t st) {
    hipLaunchKernelGGL(k_planes_finite, dim3(nblocks), dim3(256), 0, st, reinterpret_cast<const double *>(planes),
                       nplanes, NQ, JS, ny, flags);
}

// Plane sums of Statistics::reduce (statistics.cpp:54-90) on plane iy of the physical layout
// [iz][iy][ix]: u v w p, uu vv ww, uv vw wu, pu pv pw pp, u^3 v^3 w^3, u^4 v^4 w^4 with
// p = raw pressure - |u|^2/2, and in slot 20 the plane sum of div^2 when a divergence field is
// given.  One workgroup per plane; lanes sum strided points, then fixed-order wavefront and
// workgroup reductions.
constexpr int kSums = kPlaneSums;
__global__ void __launch_bounds__(256) k_plane_moments(DiagArgs a) {
    const int iy = blockIdx.x;
    double acc[kSums];
#pragma unroll
    for (int m = 0; m < kSums; ++m) acc[m] = 0.0;
    bool ok = true;
    const int n = a.NX * a.NZ;
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
        const int iz = i / a.NX, ix = i - iz * a.NX;
        const size_t g = ((size_t)iz * a.NL + iy) * a.NX + ix;
        const double u = a.pu[g], v = a.pv[g], w = a.pw[g], raw = a.pp[g];
        const double uu = u * u, vv = v * v, ww = w * w;
        const double p = raw - 0.5 * (uu + vv + ww);
        ok = ok && isfinite(u) && isfinite(v) && isfinite(w) && isfinite(raw) && isfinite(p);
        acc[0] += u;
        acc[1] += v;
        acc[2] += w;
        acc[3] += p;
        acc[4] += uu;
        acc[5] += vv;
        acc[6] += ww;
        if (a.higher) {
            acc[7] += u * v;
            acc[8] += v * w;
            acc[9] += w * u;
            acc[10] += p * u;
            acc[11] += p * v;
            acc[12] += p * w;
            acc[13] += p * p;
            acc[14] += uu * u;
            acc[15] += vv * v;
            acc[16] += ww * w;
            acc[17] += uu * uu;
            acc[18] += vv * vv;
            acc[19] += ww * ww;
        }
        if (a.pdiv) {
            const double d = a.pdiv[g];
            ok = ok && isfinite(d);
            acc[20] += d * d;
        }
    }
    wave_sum<kSums>(acc);
    __shared__ double part[4][kSums];
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int m = 0; m < kSums; ++m) part[wv][m] = acc[m];
    const int all = __syncthreads_and(ok ? 1 : 0);
    if (threadIdx.x < kSums) {
        const int m = threadIdx.x;
        a.sums[(size_t)iy * kSums + m] = (part[0][m] + part[1][m]) + (part[2][m] + part[3][m]);
    }
    if (threadIdx.x == 0) a.finite[iy] = all;
}
void launch_plane_moments(const DiagArgs &a, hipStream_t st) {
    hipLaunchKernelGGL(k_plane_moments, dim3(a.NL), dim3(256), 0, st, a);
}

} // namespace cdns
