// Spectral-space kernels: curl with the curl part of the pressure wall condition, and
// conversions between the reference's interleaved layout and the backend's column planes.
#include "internal.hpp"
#include "wave.hpp"

namespace cdns {

// vector_ops::enqueue_curl (src/solver/vector_ops.cpp:112-151) plus the curl terms of
// TimeStepper::enqueue_pressure_boundary (src/solver/timestepper.cpp:26-43):
//   wx = D w - i kz v,  wy = i kz u - i kx w,  wz = i kx v - D u,
//   bcw(wall 0) = sum_j t_j,  bcw(wall 1) = sum_j (-1)^(j+1) t_j,  t_j = (i kx/Re) wz_j - (i kz/Re) wx_j.
// One wavefront per (kx,kz) column.
template <int CH>
__global__ void __launch_bounds__(256) k_curl(CurlArgs a) {
    const int lane = lane_id();
    const int q = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (q >= a.NQ) return;
    const int kxi = q / a.nzk, kzi = q - kxi * a.nzk;
    const double kx = a.WAVX[kxi], kz = a.WAVZ[kzi];
    const size_t off = (size_t)q * a.JS;
    zd u[2][CH], v[2][CH], w[2][CH], du[2][CH], dw[2][CH];
    load_col<CH>(a.u[0] + off, lane, a.ny, u);
    load_col<CH>(a.u[1] + off, lane, a.ny, v);
    load_col<CH>(a.u[2] + off, lane, a.ny, w);
    cheb_dy<CH>(u, du, lane, a.ny);
    cheb_dy<CH>(w, dw, lane, a.ny);
    zd ox[2][CH], oy[2][CH], oz[2][CH];
    zd s0 = {0.0, 0.0}, s1 = {0.0, 0.0};
    const zd cxre = {0.0 / a.Re, kx / a.Re}, czre = {0.0, kz / a.Re};
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            ox[p][c] = dw[p][c] - ikmul(kz, v[p][c]);
            oy[p][c] = ikmul(kz, u[p][c]) - ikmul(kx, w[p][c]);
            oz[p][c] = ikmul(kx, v[p][c]) - du[p][c];
            const int j = jof<CH>(lane, c, p);
            if (j <= a.ny) {
                const zd t = cmul(cxre, oz[p][c]) - cmul(czre, ox[p][c]);
                s0 = s0 + t;
                s1 = (p == 0) ? s1 - t : s1 + t;
            }
        }
    store_col<CH>(a.w[0] + off, lane, a.ny, ox);
    store_col<CH>(a.w[1] + off, lane, a.ny, oy);
    store_col<CH>(a.w[2] + off, lane, a.ny, oz);
    if (a.bcw) {
        s0 = wave_sum(s0);
        s1 = wave_sum(s1);
        if (lane == 0) {
            a.bcw[(size_t)q * 2 + 0] = s0;
            a.bcw[(size_t)q * 2 + 1] = s1;
        }
    }
}

void launch_curl(const CurlArgs &a, int E, hipStream_t st) {
    dim3 grid((a.NQ + 3) / 4), block(256);
    if (E == 2) hipLaunchKernelGGL((k_curl<1>), grid, block, 0, st, a);
    else if (E == 4) hipLaunchKernelGGL((k_curl<2>), grid, block, 0, st, a);
    else throw std::runtime_error("unsupported column length");
}

template <int CH>
__global__ void __launch_bounds__(256) k_wall_sum(const zd *nly, const zd *bcw, zd *pbc, int ny, int JS, int NQ) {
    const int lane = lane_id();
    const int q = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (q >= NQ) return;
    zd n[2][CH];
    load_col<CH>(nly + (size_t)q * JS, lane, ny, n);
    zd s0 = {0.0, 0.0}, s1 = {0.0, 0.0};
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            s0 = s0 + n[p][c];
            s1 = p == 0 ? s1 - n[p][c] : s1 + n[p][c];
        }
    s0 = wave_sum(s0);
    s1 = wave_sum(s1);
    if (lane == 0) {
        pbc[(size_t)q * 2 + 0] = bcw[(size_t)q * 2 + 0] + s0;
        pbc[(size_t)q * 2 + 1] = bcw[(size_t)q * 2 + 1] + s1;
    }
}
void launch_wall_sum(const zd *nly, const zd *bcw, zd *pbc, int ny, int JS, int NQ, int E, hipStream_t st) {
    dim3 grid((NQ + 3) / 4), block(256);
    if (E == 2) hipLaunchKernelGGL((k_wall_sum<1>), grid, block, 0, st, nly, bcw, pbc, ny, JS, NQ);
    else hipLaunchKernelGGL((k_wall_sum<2>), grid, block, 0, st, nly, bcw, pbc, ny, JS, NQ);
}

// reference (izi, row, kx, c) -> plane c, column q = kx*nzk + izi, row (padding rows zeroed)
__global__ void k_ref_to_planes(LayoutArgs a) {
    const size_t total = (size_t)a.nkx * a.nzk * a.JS;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int row = (int)(e % a.JS);
        const size_t q = e / a.JS;
        const int kx = (int)(q / a.nzk), izi = (int)(q % a.nzk);
        for (int c = 0; c < a.ncomp; ++c) {
            zd v = {0.0, 0.0};
            if (row < a.nrow) v = a.ref_in[(((size_t)izi * a.nrow + row) * a.nkx + kx) * a.stride + c];
            a.plane_out[c][e] = v;
        }
    }
}
__global__ void k_planes_to_ref(LayoutArgs a) {
    const size_t total = (size_t)a.nzk * a.nrow * a.nkx;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int kx = (int)(e % a.nkx);
        const size_t r = e / a.nkx;
        const int row = (int)(r % a.nrow), izi = (int)(r / a.nrow);
        const size_t src = ((size_t)kx * a.nzk + izi) * a.JS + row;
        for (int c = 0; c < a.ncomp; ++c) a.ref_out[e * a.stride + c] = a.plane_in[c][src];
    }
}
__global__ void k_real_planes_to_ref(const double *plane, double *ref, int nrow, int nkx, int nzk, int JS) {
    const size_t total = (size_t)nzk * nrow * nkx;
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < total; e += (size_t)gridDim.x * blockDim.x) {
        const int kx = (int)(e % nkx);
        const size_t r = e / nkx;
        const int row = (int)(r % nrow), izi = (int)(r / nrow);
        ref[e] = plane[((size_t)kx * nzk + izi) * JS + row];
    }
}

static dim3 grid_for(size_t n) {
    size_t b = (n + 255) / 256;
    if (b > 65535 * 4) b = 65535 * 4;
    return dim3((unsigned)(b ? b : 1));
}

// n complex entries between an FP64 plane and a single-precision plane (the mixed step's N)
__global__ void k_widen(const cpx<float> *src, zd *dst, size_t n) {
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        dst[e] = {double(src[e].x), double(src[e].y)};
}
__global__ void k_narrow(const zd *src, cpx<float> *dst, size_t n) {
    for (size_t e = blockIdx.x * (size_t)blockDim.x + threadIdx.x; e < n; e += (size_t)gridDim.x * blockDim.x)
        dst[e] = {float(src[e].x), float(src[e].y)};
}
void launch_widen(const void *src, zd *dst, size_t n, hipStream_t st) {
    hipLaunchKernelGGL(k_widen, grid_for(n), dim3(256), 0, st, reinterpret_cast<const cpx<float> *>(src), dst, n);
}
void launch_narrow(const zd *src, void *dst, size_t n, hipStream_t st) {
    hipLaunchKernelGGL(k_narrow, grid_for(n), dim3(256), 0, st, src, reinterpret_cast<cpx<float> *>(dst), n);
}
void launch_ref_to_planes(const LayoutArgs &a, hipStream_t st) {
    hipLaunchKernelGGL(k_ref_to_planes, grid_for((size_t)a.nkx * a.nzk * a.JS), dim3(256), 0, st, a);
}
void launch_planes_to_ref(const LayoutArgs &a, hipStream_t st) {
    hipLaunchKernelGGL(k_planes_to_ref, grid_for((size_t)a.nzk * a.nrow * a.nkx), dim3(256), 0, st, a);
}
void launch_real_planes_to_ref(const double *plane, double *ref, int nrow, int nkx, int nzk, int JS, hipStream_t st) {
    hipLaunchKernelGGL(k_real_planes_to_ref, grid_for((size_t)nzk * nrow * nkx), dim3(256), 0, st, plane, ref, nrow,
                       nkx, nzk, JS);
}

} // namespace cdns
