// C ABI of the gfx950 backend (include/cdns.h): context, device state, setup and the step.
#include "../../include/cdns.h"
#include "internal.hpp"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <exception>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <random>

using namespace cdns;

// Horizon storage of cdns_adjoint_run_noalloc: one array per GPU, held until the process ends.
// Host code never uses the array's name as an address: cdns_create looks its device address up
// once per device (scratch_address) and every use goes through that address.
constexpr size_t kNoallocScratchBytes = size_t(64) << 20;
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

// The device address of `scratch` on the current device (looked up on the first call per device).
void *scratch_address(int dev) {
    if (dev < 0 || dev >= kMaxDevices) throw std::runtime_error("device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            throw std::runtime_error(std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = ptr;
    }
    return g_scratch_ptr[dev];
}

// ---------------------------------------------------------------- tables --
FftPlan make_plan(int n) {
    FftPlan p{};
    p.n = n;
    int m = n, a = 0;
    while (m % 2 == 0) { m /= 2; ++a; }
    auto push = [&](int r) {
        if (p.nst >= MAXST) throw std::runtime_error("FFT plan too long");
        p.rad[p.nst++] = r;
    };
    // powers of two as radix-8/4 stages (a lone 2 only when a == 1)
    if (a == 1) push(2);
    else if (a == 2) push(4);
    else if (a >= 3) {
        int r = a % 3, k8 = a / 3;
        if (r == 1) { k8 -= 1; push(4); push(4); }
        else if (r == 2) push(4);
        for (int i = 0; i < k8; ++i) push(8);
    }
    for (int f : {3, 5, 7})
        while (m % f == 0) { m /= f; push(f); }
    if (m != 1) throw std::runtime_error("FFT length " + std::to_string(n) + " has a prime factor above 7");
    return p;
}

std::vector<long double> twiddles(int n) {
    std::vector<long double> t(2 * (size_t)n);
    const long double twopi = 6.283185307179586476925286766559005768L;
    for (int k = 0; k < n; ++k) {
        const long double a = twopi * (long double)k / (long double)n;
        t[2 * k] = cosl(a);
        t[2 * k + 1] = -sinl(a);
    }
    return t;
}

// cos(pi m / N) for m in [0, 2N), with the exact zeros at pi/2 and 3pi/2
std::vector<long double> cospi_table(int N) {
    std::vector<long double> t(2 * (size_t)N);
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int m = 0; m < 2 * N; ++m) {
        long double v = cosl(pi * (long double)m / (long double)N);
        if (2 * m == N || 2 * m == 3 * N) v = 0.0L;
        t[m] = v;
    }
    return t;
}

// y-stage matrices for N + 1 Gauss-Lobatto points from exact-angle cosines: synthesis B[j][l]
// (l <= N/2) and the DCT-I analysis weights split by parity, Fe[l][j/2] (j even), Fo[l][j/2] (odd).
void y_tables(int N, int nj, int NLH, std::vector<double> &B, std::vector<double> &Fe, std::vector<double> &Fo) {
    const auto cs = cospi_table(N);
    B.assign((size_t)nj * NLH, 0.0);
    for (int j = 0; j < nj; ++j)
        for (int l = 0; l < NLH; ++l) {
            const long double bw = (j == N) ? 0.5L : 1.0L; // REDFT00 endpoint weight x the 0.5 z factor
            B[(size_t)j * NLH + l] = (double)(bw * cs[((long long)j * l) % (2 * N)]);
        }
    const int nje = (nj + 1) / 2, njo = nj / 2;
    Fe.assign((size_t)NLH * nje, 0.0);
    Fo.assign((size_t)NLH * std::max(njo, 1), 0.0);
    for (int l = 0; l < NLH; ++l) {
        const long double wl = (l == 0) ? 1.0L : 2.0L;
        for (int j = 0; j < nj; ++j) {
            const long double sc = j == 0 ? 1.0L / (2.0L * N) : 1.0L / N;
            const double v = (double)(sc * wl * cs[((long long)j * l) % (2 * N)]);
            if (j % 2 == 0) Fe[(size_t)l * nje + j / 2] = v;
            else Fo[(size_t)l * njo + j / 2] = v;
        }
    }
}

// T'_j(y_l) at y_l = cos(pi l / N), l < NLH: j^2 at the wall, j sin(pi j l / N) / sin(pi l / N)
// elsewhere (the derivative of T_j = cos(j theta)), from exact-angle sines.
void y_dtable(int N, int nj, int NLH, std::vector<double> &D) {
    const long double pi = acosl(-1.0L);
    auto sinpi = [&](long long m) { // sin(pi m / N)
        m %= 2 * N;
        if (m % N == 0) return 0.0L;
        if (2 * m == N) return 1.0L;
        if (2 * m == 3 * N) return -1.0L;
        return sinl(pi * (long double)m / (long double)N);
    };
    D.assign((size_t)nj * NLH, 0.0);
    for (int j = 0; j < nj; ++j)
        for (int l = 0; l < NLH; ++l)
            D[(size_t)j * NLH + l] = l == 0 ? (double)j * (double)j
                                            : (double)((long double)j * sinpi((long long)j * l) / sinpi(l));
}

struct Event2 {
    hipEvent_t a, b;
    int name;
};

} // namespace

struct cdns_ctx {
    cdns_config cfg{};
    Geometry g{};
    bool fp32 = false;
    bool nl32 = false; // the step's nonlinear term and its history in single precision (mixed)
    hipStream_t st = nullptr;
    std::string err;
    size_t esz = 16; // bytes per physical complex element

    // tables
    double *WAVX = nullptr, *WAVZ = nullptr, *WL = nullptr, *WD = nullptr, *WR = nullptr;
    double *ycoef_b = nullptr, *ycoef_fe = nullptr, *ycoef_fo = nullptr, *ycoef_d = nullptr;
    void *twz = nullptr, *twx = nullptr;
    FftPlan pz{}, px{};
    int KT = 13, P = 4;
    int rf = 1; // register-resident FFTs (CDNS_FFT=stockham selects the LDS Stockham kernels)
    int ymfma = 1; // y stages on the matrix cores (CDNS_YSTAGE=vector selects the vector-ALU kernels)
    // y stages as a DCT-I through FFTs of length 2 NY2 (CDNS_YSTAGE=dct), with its root table
    bool ydct = false;
    void *ydct_tw = nullptr;
    int ycurl = 1; // curl fused into the matrix-core y backward of the step (CDNS_YCURL=0: separate)
    int xpair = 1; // x-stage product FFT rows on lane pairs (CDNS_XPRODUCT=single keeps one lane per row)
    int zpair = 1; // z-stage FFT rows on lane pairs (CDNS_ZSTAGE=single keeps one lane per row)
    std::vector<double> hWAVX, hWAVZ;

    // solver tables
    double g_v = 0.0, pa = 0.0, pb = 0.0;
    double *ad_p = nullptr, *ad_v = nullptr, *wn_p = nullptr, *wn_v = nullptr;
    zd *PI[2] = {}, *VI[2][3] = {}, *ID[2] = {}, *PB = nullptr, *VB[3] = {}, *SIG0 = nullptr;
    // compact real tables (PIr0 PIr1 VIv0 VIv1 VIf0 VIf1 PBr VBv VBf) when u and w share wall
    // conditions (CDNS_SOLVER_TABLES=complex keeps the complex ones in the step)
    bool compact = false;
    double *ctab[9] = {};
    double *u_unit = nullptr, *w_unit = nullptr, *q_unit = nullptr;
    bool setup_done = false;

    // state
    zd *u[3] = {}, *up[3] = {}, *nl[3] = {}, *nlp[3] = {}, *P_ = nullptr, *om[3] = {};
    zd *bc_prev = nullptr, *bc_ext = nullptr, *bcw = nullptr, *vbc = nullptr;
    double *flow = nullptr;
    double time = 0.0;

    // constant body force added to the nonlinear term (3 planes), when set
    zd *body[3] = {};
    bool body_on = false;

    // adjoint workspaces (allocated on first use; FP64 only)
    bool adj_ready = false;
    zd *a_lu[3] = {}, *a_lup[3] = {}, *a_lnp[3] = {}, *a_lbp = nullptr, *a_lf[3] = {};
    zd *a_lNs[3] = {}, *a_lw[3] = {}, *a_lUs[3] = {}, *a_lWs[3] = {};
    void *ZL = nullptr, *ZNa = nullptr;
    double *Jpart = nullptr;
    zd *body_save[3] = {}; // the caller's body force while cdns_adjoint_run uses its own
    // horizon storage of cdns_adjoint_run (checkpoints, one segment's synthesised fields), kept
    // between calls and grown when a call needs more
    void *hz_ck = nullptr, *hz_seg = nullptr;
    size_t hz_ck_bytes = 0, hz_seg_bytes = 0;
    std::vector<double> Jhost; // host copy of the objective's per-column partial sums
    // the device the context was created on and its address of the no-allocation scratch array
    int device = 0;
    void *noalloc_scratch = nullptr;

    // coarse grid NX x (NY+1) x NZ (Transforms::channel_xyzfft): FP64 tables and plans, built on
    // first use; the transforms run in the product-grid workspaces, which are larger
    bool coarse_ready = false;
    double *cy_b = nullptr, *cy_fe = nullptr, *cy_fo = nullptr, *cphys = nullptr, *csums = nullptr;
    void *ctwz = nullptr, *ctwx = nullptr;
    FftPlan cpz{}, cpx{};
    int *cflags = nullptr;
    std::vector<double> cc_weights; // Clenshaw-Curtis weights / 2 (diagnostics.cpp:17-33)

    // workspaces
    void *Ybuf = nullptr, *Zbuf = nullptr, *ZNbuf = nullptr;
    zd *scr[8] = {};       // spectral planes for C-ABI staging
    zd *ref_stage = nullptr; // reference-layout spectral staging (3 components)
    zd *bnd_stage = nullptr; // boundary staging
    double *phys_stage = nullptr;

    std::vector<void *> allocs;

    // profiling
    bool prof = false;
    std::vector<std::string> kname;
    std::map<std::string, int> kidx;
    std::vector<double> ktot;
    std::vector<long long> kcnt;
    std::vector<Event2> pending;
    std::vector<hipEvent_t> pool;

    template <typename T> T *alloc(size_t n) {
        void *p = nullptr;
        CDNS_HIP(hipMalloc(&p, std::max<size_t>(n, 1) * sizeof(T)));
        CDNS_HIP(hipMemsetAsync(p, 0, std::max<size_t>(n, 1) * sizeof(T), st));
        allocs.push_back(p);
        return static_cast<T *>(p);
    }
    hipEvent_t ev() {
        if (!pool.empty()) {
            hipEvent_t e = pool.back();
            pool.pop_back();
            return e;
        }
        hipEvent_t e;
        CDNS_HIP(hipEventCreate(&e));
        return e;
    }
};

namespace {

// Records a HIP event pair around each named launch when profiling is on.
struct Launch {
    cdns_ctx *c;
    int id = -1;
    hipEvent_t a{}, b{};
    Launch(cdns_ctx *ctx, const char *name) : c(ctx) {
        if (!c->prof) return;
        auto it = c->kidx.find(name);
        if (it == c->kidx.end()) {
            id = (int)c->kname.size();
            c->kidx[name] = id;
            c->kname.push_back(name);
            c->ktot.push_back(0.0);
            c->kcnt.push_back(0);
        } else {
            id = it->second;
        }
        a = c->ev();
        b = c->ev();
        CDNS_HIP(hipEventRecord(a, c->st));
    }
    ~Launch() noexcept(false) {
        if (std::uncaught_exceptions() > 0) return;
        CDNS_HIP(hipGetLastError());
        if (id < 0) return;
        CDNS_HIP(hipEventRecord(b, c->st));
        c->pending.push_back({a, b, id});
    }
};

void drain_profile(cdns_ctx *c) {
    for (auto &e : c->pending) {
        CDNS_HIP(hipEventSynchronize(e.b));
        float ms = 0.f;
        CDNS_HIP(hipEventElapsedTime(&ms, e.a, e.b));
        c->ktot[e.name] += ms;
        c->kcnt[e.name] += 1;
        c->pool.push_back(e.a);
        c->pool.push_back(e.b);
    }
    c->pending.clear();
}

// ---------------------------------------------------------- transforms ----
YArgs yargs(cdns_ctx *c) {
    YArgs a{};
    const Geometry &g = c->g;
    a.coef = c->ycoef_b;
    a.nj = g.nj;
    a.nje = (g.nj + 1) / 2;
    a.njo = g.nj / 2;
    a.JS = g.JS;
    a.NQ = g.NQ;
    a.NLH = g.NLH;
    a.NY2 = g.NY2;
    a.YS = g.YS;
    a.y = c->Ybuf;
    a.mfma = c->ymfma;
    return a;
}
DctArgs dct_args(cdns_ctx *c) {
    DctArgs a{};
    const Geometry &g = c->g;
    a.y = c->Ybuf;
    a.YS = g.YS;
    a.tw = c->ydct_tw;
    a.wavx = c->WAVX;
    a.wavz = c->WAVZ;
    a.Re = c->cfg.Re;
    a.ny = g.NYS;
    a.JS = g.JS;
    a.NQ = g.NQ;
    a.nzk = g.nzk;
    a.NY2 = g.NY2;
    a.L = 2 * g.NY2;
    a.CH = g.E / 2;
    return a;
}
ZArgs zargs(cdns_ctx *c, const void *in, void *out, size_t in_fs, size_t out_fs) {
    ZArgs a{};
    a.in = in;
    a.out = out;
    a.in_fs = in_fs;
    a.out_fs = out_fs;
    a.tw = c->twz;
    a.plan = c->pz;
    a.nkx = c->g.nkx;
    a.nzk = c->g.nzk;
    a.NZS = c->g.NZS;
    a.NQ = c->g.NQ;
    a.KT = c->KT;
    a.NL = c->g.NL;
    a.rf = c->rf;
    a.pair = c->zpair;
    return a;
}
XArgs xargs(cdns_ctx *c, const void *in, void *out, size_t in_fs, size_t out_fs, int P) {
    XArgs a{};
    a.in = in;
    a.out = out;
    a.in_fs = in_fs;
    a.out_fs = out_fs;
    a.tw = c->twx;
    a.plan = c->px;
    a.nkx = c->g.nkx;
    a.NL = c->g.NL;
    a.NZ2 = c->g.NZ2;
    a.P = P;
    a.rf = c->rf;
    a.pair = c->xpair;
    for (int i = 0; i < 6; ++i) a.fin[i] = in ? (const char *)in + i * in_fs * c->esz : nullptr;
    for (int i = 0; i < 3; ++i) a.fout[i] = out ? (char *)out + i * out_fs * c->esz : nullptr;
    return a;
}

// spectral planes -> Y fields (n fields) -> Z fields (the context's Z buffer, or zout)
// Y fields (n) -> Z fields (the context's Z buffer, or zout)
void backward_z(cdns_ctx *c, int n, void *zout = nullptr) {
    Launch L(c, "z_backward");
    launch_zbwd(zargs(c, c->Ybuf, zout ? zout : c->Zbuf, c->g.YS, c->g.ZS), n, c->fp32, c->st);
}
void backward_yz(cdns_ctx *c, const zd *const *planes, int n, void *zout = nullptr) {
    if (c->ydct) {
        DctArgs a = dct_args(c);
        for (int f = 0; f < n; ++f) a.spec[f] = planes[f];
        a.nf = n;
        Launch L(c, "y_backward");
        if (!launch_ybwd_dct(a, c->fp32, c->st)) throw std::logic_error("no DCT y stage for this NY2");
    } else {
        YArgs a = yargs(c);
        for (int f = 0; f < n; ++f) a.spec[f] = planes[f];
        Launch L(c, "y_backward");
        launch_ybwd(a, n, c->fp32, c->st);
    }
    backward_z(c, n, zout);
}
// ZN fields (n, the context's ZN buffer or zin) -> Y fields -> spectral planes (+ add planes)
// (out32: the planes hold single-precision complex values)
void forward_zy(cdns_ctx *c, zd *const *planes, int n, const void *zin = nullptr, zd *const *add = nullptr,
                bool out32 = false) {
    {
        Launch L(c, "z_forward");
        launch_zfwd(zargs(c, zin ? zin : c->ZNbuf, c->Ybuf, c->g.ZS, c->g.YS), n, c->fp32, c->st);
    }
    if (c->ydct) {
        DctArgs a = dct_args(c);
        for (int f = 0; f < n; ++f) {
            a.spec_out[f] = planes[f];
            a.add[f] = add ? add[f] : nullptr;
        }
        a.nf = n;
        a.out32 = out32;
        Launch L(c, "y_forward");
        if (!launch_yfwd_dct(a, c->fp32, c->st)) throw std::logic_error("no DCT y stage for this NY2");
        return;
    }
    YArgs a = yargs(c);
    a.coef = c->ycoef_fe;
    a.coef2 = c->ycoef_fo;
    a.out32 = out32;
    for (int f = 0; f < n; ++f) {
        a.spec_out[f] = planes[f];
        a.add[f] = add ? add[f] : nullptr;
    }
    {
        Launch L(c, "y_forward");
        launch_yfwd(a, n, c->fp32, c->st);
    }
}

// N = a x b (+ add) for spectral vector planes (a: 3 planes, b: 3 planes)
void cross_planes(cdns_ctx *c, zd *const *a, zd *const *b, zd *const *out, zd *const *add = nullptr,
                  bool out32 = false) {
    const zd *f6[6] = {a[0], a[1], a[2], b[0], b[1], b[2]};
    backward_yz(c, f6, 6);
    {
        Launch L(c, "x_product");
        launch_xprod(xargs(c, c->Zbuf, c->ZNbuf, c->g.ZS, c->g.ZS, c->P), c->fp32, c->st);
    }
    forward_zy(c, out, 3, nullptr, add, out32);
}

// ------------------------------------------------------------- coarse grid ----
// Transforms::channel_xyzfft (transforms.cpp:452-582): the base grid NX x (NY+1) x NZ of
// to_physical, statistics and diagnostics, always in FP64, in the product-grid workspaces.
void ensure_coarse(cdns_ctx *c) {
    if (c->coarse_ready) return;
    const Geometry &g = c->g;
    if (!rf_supported(g.NZ) || !rf_supported(g.NX))
        throw std::invalid_argument("coarse transform: NX and NZ must be register-FFT lengths");
    const int NLc = g.NY + 1, NLHc = g.NY / 2 + 1;
    auto upload = [&](const std::vector<double> &h) {
        double *d = c->alloc<double>(h.size());
        CDNS_HIP(hipMemcpyAsync(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, c->st));
        return d;
    };
    auto up_tw = [&](int n) {
        const auto t = twiddles(n);
        std::vector<double> dv(t.begin(), t.end());
        return (void *)upload(dv);
    };
    std::vector<double> B, Fe, Fo;
    y_tables(g.NY, g.nj, NLHc, B, Fe, Fo);
    c->cy_b = upload(B);
    c->cy_fe = upload(Fe);
    c->cy_fo = upload(Fo);
    c->cpz = make_plan(g.NZ);
    c->cpx = make_plan(g.NX);
    c->ctwz = up_tw(g.NZ);
    c->ctwx = up_tw(g.NX);
    c->cphys = c->alloc<double>(5 * (size_t)g.NZ * NLc * g.NX);
    c->csums = c->alloc<double>((size_t)NLc * kPlaneSums);
    c->cflags = c->alloc<int>(std::max(NLc, 256));
    // Clenshaw-Curtis weights on [-1, 1] over height 2 (diagnostics.cpp:17-33)
    const int n = g.NY;
    const double pi = std::acos(-1.0);
    c->cc_weights.assign(NLc, 0.0);
    for (int j = 0; j <= n; ++j) {
        const double theta = pi * j / n;
        double weight;
        if (j == 0 || j == n) weight = 1.0 / (n * n - (n % 2 == 0 ? 1.0 : 0.0));
        else {
            double value = 1.0;
            for (int k = 1; k <= (n - 1) / 2; ++k) value -= 2.0 * std::cos(2 * k * theta) / (4.0 * k * k - 1.0);
            if (n % 2 == 0) value -= std::cos(n * theta) / (n * n - 1.0);
            weight = 2.0 * value / n;
        }
        c->cc_weights[j] = weight / 2.0;
    }
    CDNS_HIP(hipStreamSynchronize(c->st));
    c->coarse_ready = true;
}
YArgs yargs_coarse(cdns_ctx *c) {
    YArgs a = yargs(c);
    a.coef = c->cy_b;
    a.NLH = c->g.NY / 2 + 1;
    a.NY2 = c->g.NY;
    a.YS = (size_t)(c->g.NY + 1) * c->g.NQ;
    return a;
}
size_t coarse_zs(const cdns_ctx *c) { return (size_t)(c->g.NY + 1) * c->g.NZ * c->g.nkx; }
ZArgs zargs_coarse(cdns_ctx *c, const void *in, void *out) {
    const size_t ys = (size_t)(c->g.NY + 1) * c->g.NQ;
    ZArgs a = zargs(c, in, out, ys, coarse_zs(c));
    a.tw = c->ctwz;
    a.plan = c->cpz;
    a.NL = c->g.NY + 1;
    a.rf = 1;
    a.coarse = 1;
    return a;
}
XArgs xargs_coarse(cdns_ctx *c, const void *in, void *out, double *phys) {
    XArgs a = xargs(c, in, out, coarse_zs(c), coarse_zs(c), 2 * c->P);
    a.tw = c->ctwx;
    a.plan = c->cpx;
    a.NL = c->g.NY + 1;
    a.NZ2 = c->g.NZ;
    a.rf = 1;
    a.coarse = 1;
    a.phys = phys;
    return a;
}
// spectral plane -> coarse physical field [iz][iy][ix] (device)
void coarse_backward(cdns_ctx *c, const zd *plane, double *phys) {
    ensure_coarse(c);
    YArgs a = yargs_coarse(c);
    a.spec[0] = plane;
    {
        Launch L(c, "coarse_y_backward");
        launch_ybwd(a, 1, false, c->st);
    }
    {
        Launch L(c, "coarse_z_backward");
        launch_zbwd(zargs_coarse(c, c->Ybuf, c->Zbuf), 1, false, c->st);
    }
    {
        Launch L(c, "coarse_x_backward");
        launch_xbwd_phys(xargs_coarse(c, c->Zbuf, nullptr, phys), false, c->st);
    }
}
// coarse physical field (device) -> spectral plane
void coarse_forward(cdns_ctx *c, const double *phys, zd *plane) {
    ensure_coarse(c);
    {
        Launch L(c, "coarse_x_forward");
        launch_xfwd_phys(xargs_coarse(c, nullptr, c->ZNbuf, const_cast<double *>(phys)), false, c->st);
    }
    {
        Launch L(c, "coarse_z_forward");
        launch_zfwd(zargs_coarse(c, c->ZNbuf, c->Ybuf), 1, false, c->st);
    }
    YArgs a = yargs_coarse(c);
    a.coef = c->cy_fe;
    a.coef2 = c->cy_fo;
    a.spec_out[0] = plane;
    a.add[0] = nullptr;
    {
        Launch L(c, "coarse_y_forward");
        launch_yfwd(a, 1, false, c->st);
    }
}
size_t coarse_points(const cdns_ctx *c) { return (size_t)c->g.NZ * (c->g.NY + 1) * c->g.NX; }
// FieldIO::to_physical (field_io.cpp:199-220) of the state into cphys[0..3] (u, v, w, pressure);
// returns whether every retained spectral coefficient is finite
bool state_to_physical(cdns_ctx *c) {
    ensure_coarse(c);
    const zd *src[4] = {c->u[0], c->u[1], c->u[2], c->P_};
    for (int k = 0; k < 4; ++k) coarse_backward(c, src[k], c->cphys + k * coarse_points(c));
    int ok = 1;
    const int nb = 256;
    std::vector<int> flags(nb);
    for (int k = 0; k < 4; ++k) {
        launch_planes_finite(src[k], 1, c->g.NQ, c->g.JS, c->g.NYS, c->cflags, nb, c->st);
        CDNS_HIP(hipMemcpyAsync(flags.data(), c->cflags, nb * sizeof(int), hipMemcpyDeviceToHost, c->st));
        CDNS_HIP(hipStreamSynchronize(c->st));
        for (int f : flags) ok &= f;
    }
    return ok != 0;
}
// Statistics::reduce over cphys (and the plane sum of div^2 when with_div) into host sums
void plane_moments(cdns_ctx *c, bool higher, bool with_div, std::vector<double> &sums, std::vector<int> &finite) {
    const Geometry &g = c->g;
    const int NLc = g.NY + 1;
    const size_t np = coarse_points(c);
    DiagArgs a{};
    a.pu = c->cphys;
    a.pv = c->cphys + np;
    a.pw = c->cphys + 2 * np;
    a.pp = c->cphys + 3 * np;
    a.pdiv = with_div ? c->cphys + 4 * np : nullptr;
    a.NX = g.NX;
    a.NZ = g.NZ;
    a.NL = NLc;
    a.higher = higher ? 1 : 0;
    a.sums = c->csums;
    a.finite = c->cflags;
    {
        Launch L(c, "plane_moments");
        launch_plane_moments(a, c->st);
    }
    sums.resize((size_t)NLc * kPlaneSums);
    finite.resize(NLc);
    CDNS_HIP(hipMemcpyAsync(sums.data(), c->csums, sums.size() * sizeof(double), hipMemcpyDeviceToHost, c->st));
    CDNS_HIP(hipMemcpyAsync(finite.data(), c->cflags, NLc * sizeof(int), hipMemcpyDeviceToHost, c->st));
    CDNS_HIP(hipStreamSynchronize(c->st));
}

void curl_planes(cdns_ctx *c, zd *const *u, zd *const *w, zd *bcw) {
    CurlArgs a{};
    for (int k = 0; k < 3; ++k) {
        a.u[k] = u[k];
        a.w[k] = w[k];
    }
    a.bcw = bcw;
    a.WAVX = c->WAVX;
    a.WAVZ = c->WAVZ;
    a.Re = c->cfg.Re;
    a.ny = c->g.NYS;
    a.JS = c->g.JS;
    a.NQ = c->g.NQ;
    a.nzk = c->g.nzk;
    Launch L(c, "curl");
    launch_curl(a, c->g.E, c->st);
}

// TimeStepper::enqueue_nonlinear_term on the context's velocity planes
void nonlinear_planes(cdns_ctx *c, zd *const *u, zd *const *nl, bool out32 = false) {
    zd *const *add = c->body_on ? c->body : nullptr;
    if (c->ydct && c->ycurl) {
        // DCT-I y backward of u with the curl formed per column: the six Y fields and the wall
        // condition
        DctArgs a = dct_args(c);
        for (int k = 0; k < 3; ++k) a.spec[k] = u[k];
        a.nf = 6;
        a.curl = 1;
        a.bcw = c->bcw;
        {
            Launch L(c, "y_backward_curl");
            if (!launch_ybwd_dct(a, c->fp32, c->st)) throw std::logic_error("no DCT y stage for this NY2");
        }
        backward_z(c, 6);
        {
            Launch L(c, "x_product");
            launch_xprod(xargs(c, c->Zbuf, c->ZNbuf, c->g.ZS, c->g.ZS, c->P), c->fp32, c->st);
        }
        forward_zy(c, nl, 3, nullptr, add, out32);
        return;
    }
    if (c->ymfma && c->ycurl) {
        // y backward of u with the curl fused in: the six Y fields and the wall condition
        YArgs a = yargs(c);
        for (int k = 0; k < 3; ++k) a.spec[k] = u[k];
        a.coefd = c->ycoef_d;
        a.wavx = c->WAVX;
        a.wavz = c->WAVZ;
        a.Re = c->cfg.Re;
        a.nzk = c->g.nzk;
        a.bcw = c->bcw;
        bool done;
        {
            Launch L(c, "y_backward_curl");
            done = launch_ybwd_curl(a, c->fp32, c->st);
        }
        if (done) {
            backward_z(c, 6);
            {
                Launch L(c, "x_product");
                launch_xprod(xargs(c, c->Zbuf, c->ZNbuf, c->g.ZS, c->g.ZS, c->P), c->fp32, c->st);
            }
            forward_zy(c, nl, 3, nullptr, add, out32);
            return;
        }
    }
    curl_planes(c, u, c->om, c->bcw);
    cross_planes(c, u, c->om, nl, add, out32);
}

// ---------------------------------------------------------------- solver --
SolveArgs solve_args(cdns_ctx *c, int mode) {
    SolveArgs a{};
    const Geometry &g = c->g;
    a.mode = mode;
    a.ny = g.NYS;
    a.JS = g.JS;
    a.NQ = g.NQ;
    a.nzk = g.nzk;
    a.nkx = g.nkx;
    a.WAVX = c->WAVX;
    a.WAVZ = c->WAVZ;
    a.WL = c->WL;
    a.WD = c->WD;
    a.WR = c->WR;
    a.Re = c->cfg.Re;
    a.G_pre = c->cfg.G_pre;
    a.g_v = c->g_v;
    a.pa = c->pa;
    a.pb = c->pb;
    for (int k = 0; k < 3; ++k) {
        a.aV[k] = c->cfg.a_V[k];
        a.bV[k] = c->cfg.b_V[k];
    }
    a.ad_p = c->ad_p;
    a.ad_v = c->ad_v;
    for (int i = 0; i < 2; ++i) {
        a.PI[i] = c->PI[i];
        a.ID[i] = c->ID[i];
        for (int k = 0; k < 3; ++k) a.VI[i][k] = c->VI[i][k];
    }
    a.PB = c->PB;
    for (int k = 0; k < 3; ++k) a.VB[k] = c->VB[k];
    a.SIG0 = c->SIG0;
    a.u_unit = c->u_unit;
    a.w_unit = c->w_unit;
    a.q_unit = c->q_unit;
    const double th = c->cfg.THETA * std::acos(-1.0) / 180.0;
    a.grad_pre = c->cfg.INDEX_GRAD_PRE;
    if (a.grad_pre == 1) {
        a.grad_x = -std::cos(th);
        a.grad_z = -std::sin(th);
    } else {
        a.qx_mean = std::cos(th);
        a.qz_mean = std::sin(th);
    }
    a.vbc = c->vbc;
    a.flow = c->flow;
    a.compact = c->compact ? 1 : 0;
    if (c->compact) {
        a.PIr[0] = c->ctab[0];
        a.PIr[1] = c->ctab[1];
        a.VIv[0] = c->ctab[2];
        a.VIv[1] = c->ctab[3];
        a.VIf[0] = c->ctab[4];
        a.VIf[1] = c->ctab[5];
        a.PBr = c->ctab[6];
        a.VBv = c->ctab[7];
        a.VBf = c->ctab[8];
    }
    return a;
}

// EqSolvePV::setup (eqsolve_pv.cpp:495-509), entirely on the device.
void setup(cdns_ctx *c, double g_v) {
    const Geometry &g = c->g;
    c->g_v = g_v;
    {
        Launch L(c, "setup_ad");
        launch_build_ad(c->WAVX, c->WAVZ, c->WL, c->WD, c->WR, c->cfg.G_pre, g.NYS, g.JS, g.NQ, g.nzk, c->ad_p,
                        c->wn_p, c->st);
        launch_build_ad(c->WAVX, c->WAVZ, c->WL, c->WD, c->WR, g_v, g.NYS, g.JS, g.NQ, g.nzk, c->ad_v, c->wn_v, c->st);
    }
    c->pa = 1.0;
    c->pb = 0.0;
    for (int i = 0; i < 2; ++i) {
        SolveArgs a = solve_args(c, MODE_INFLUENCE);
        a.influence_index = i;
        a.P = c->PI[i];
        for (int k = 0; k < 3; ++k) a.V[k] = c->VI[i][k];
        a.div_out = c->ID[i];
        Launch L(c, "setup_influence");
        launch_solve(a, g.E, c->st);
    }
    {
        Launch L(c, "setup_normalize");
        launch_normalize_influence(c->ID[0], c->ID[1], g.NQ, g.NZS, c->cfg.a_V[1], c->st);
    }
    {
        SolveArgs a = solve_args(c, MODE_BPROBLEM);
        a.correct = 1;
        a.P = c->PB;
        for (int k = 0; k < 3; ++k) a.V[k] = c->VB[k];
        a.sig_out = c->SIG0;
        Launch L(c, "setup_bproblem");
        launch_solve(a, g.E, c->st);
    }
    if (c->compact) {
        SolveArgs a = solve_args(c, MODE_STEP);
        Launch L(c, "setup_compact");
        launch_compact_tables(a, c->ctab[0], c->ctab[1], c->ctab[2], c->ctab[3], c->ctab[4], c->ctab[5], c->ctab[6],
                              c->ctab[7], c->ctab[8], c->st);
    }
    c->pa = c->cfg.a_pre0;
    c->pb = c->cfg.b_pre0;
    {
        Launch L(c, "setup_pg_response");
        launch_pg_response(c->WAVX, c->WAVZ, c->WL, c->WD, c->WR, c->ad_v, g_v, c->cfg.Re, c->cfg.a_V, c->cfg.b_V,
                           g.NYS, g.JS, g.NZS, c->u_unit, c->w_unit, c->q_unit, g.E, c->st);
    }
    c->setup_done = true;
}

void step_coefficients(int it, double dt, double Re, double co[5]) {
    if (it < 1 || dt <= 0.0) throw std::invalid_argument("invalid time-step coefficients");
    const int eff = it < 3 ? it : 2;
    const double a0 = (double)eff / dt, a1 = (double)(1 - eff) / (2.0 * dt);
    co[0] = a0;
    co[1] = a1;
    co[2] = (double)eff;
    co[3] = (double)(1 - eff);
    co[4] = (a0 + a1) * Re;
}

// One iteration of Simulation::run (simulation.cpp:173-191) with enqueue_serial_step (142-171).
void step(cdns_ctx *c, int it) {
    double co[5];
    step_coefficients(it, c->cfg.DT, c->cfg.Re, co);
    if (it < 3) {
        setup(c, co[4]);
    } else {
        if (!c->setup_done) throw std::logic_error("solver setup has not run (the first steps must be 1 and 2)");
        // the tables may have been set up for another g since (an adjoint step of iteration 1)
        if (c->g_v != co[4]) setup(c, co[4]);
    }
    c->time += c->cfg.DT;
    nonlinear_planes(c, c->u, c->nl, c->nl32);
    SolveArgs a = solve_args(c, MODE_STEP);
    a.nl32 = c->nl32;
    for (int k = 0; k < 3; ++k) {
        a.u[k] = c->u[k];
        a.up[k] = c->up[k];
        a.nl[k] = c->nl[k];
        a.nlp[k] = c->nlp[k];
        a.V[k] = c->up[k]; // u^{n+1} replaces u^{n-1} in place
    }
    a.a0 = co[0];
    a.a1 = co[1];
    a.b0 = co[2];
    a.b1 = co[3];
    a.bcw = c->bcw;
    a.bc_prev = c->bc_prev;
    a.bc_ext = c->bc_ext;
    a.P = c->P_;
    a.correct = 1;
    a.sigma = 1;
    {
        Launch L(c, "solve_step");
        launch_solve(a, c->g.E, c->st);
    }
    for (int k = 0; k < 3; ++k) {
        std::swap(c->u[k], c->up[k]);
        std::swap(c->nl[k], c->nlp[k]);
    }
}

// ------------------------------------------------ discrete adjoint (new) --
// The reference has no adjoint; this transposes the step above (FP64 contexts only).
double g_for_iteration(const cdns_ctx *c, int it) {
    double co[5];
    step_coefficients(it < 2 ? it : 2, c->cfg.DT, c->cfg.Re, co);
    return co[4];
}

// Make the solver tables those of g (setup runs only when they are not).
void ensure_tables(cdns_ctx *c, double g) {
    if (!c->setup_done || c->g_v != g) setup(c, g);
}

void ensure_adjoint(cdns_ctx *c) {
    if (c->fp32) throw std::logic_error("the discrete adjoint runs on FP64 contexts");
    if (c->adj_ready) return;
    const Geometry &g = c->g;
    for (int k = 0; k < 3; ++k) {
        c->a_lu[k] = c->alloc<zd>(g.S);
        c->a_lup[k] = c->alloc<zd>(g.S);
        c->a_lnp[k] = c->alloc<zd>(g.S);
        c->a_lf[k] = c->alloc<zd>(g.S);
        c->a_lNs[k] = c->alloc<zd>(g.S);
        c->a_lw[k] = c->alloc<zd>(g.S);
        c->a_lUs[k] = c->alloc<zd>(g.S);
        c->a_lWs[k] = c->alloc<zd>(g.S);
    }
    c->a_lbp = c->alloc<zd>((size_t)g.NQ * 2);
    c->ZL = c->alloc<unsigned char>(3 * g.ZS * c->esz);
    c->ZNa = c->alloc<unsigned char>(6 * g.ZS * c->esz);
    c->Jpart = c->alloc<double>(g.NQ);
    c->Jhost.assign(g.NQ, 0.0);
    for (int k = 0; k < 3; ++k) c->body_save[k] = c->alloc<zd>(g.S);
    c->adj_ready = true;
}

AdjArgs adj_args(cdns_ctx *c) {
    AdjArgs a{};
    const Geometry &g = c->g;
    a.ny = g.NYS;
    a.JS = g.JS;
    a.NQ = g.NQ;
    a.nzk = g.nzk;
    a.nkx = g.nkx;
    a.NX2 = g.NX2;
    a.NY2 = g.NY2;
    a.collide = (g.NZ2 - g.NZS <= g.NZS) ? 1 : 0;
    a.WAVX = c->WAVX;
    a.WAVZ = c->WAVZ;
    a.WL = c->WL;
    a.WD = c->WD;
    a.WR = c->WR;
    a.Re = c->cfg.Re;
    a.G_pre = c->cfg.G_pre;
    a.g_v = c->g_v;
    a.pa = c->pa;
    a.pb = c->pb;
    a.XL = c->cfg.XL;
    a.ZL = c->cfg.ZL;
    for (int k = 0; k < 3; ++k) {
        a.aV[k] = c->cfg.a_V[k];
        a.bV[k] = c->cfg.b_V[k];
    }
    a.grad_pre = c->cfg.INDEX_GRAD_PRE;
    a.ad_p = c->ad_p;
    a.ad_v = c->ad_v;
    for (int i = 0; i < 2; ++i) {
        a.PI[i] = c->PI[i];
        a.ID[i] = c->ID[i];
        for (int k = 0; k < 3; ++k) a.VI[i][k] = c->VI[i][k];
    }
    for (int k = 0; k < 3; ++k) a.VB[k] = c->VB[k];
    a.SIG0 = c->SIG0;
    a.u_unit = c->u_unit;
    a.w_unit = c->w_unit;
    a.q_unit = c->q_unit;
    for (int k = 0; k < 3; ++k) {
        a.lu[k] = c->a_lu[k];
        a.lup[k] = c->a_lup[k];
        a.lnp[k] = c->a_lnp[k];
        a.lf[k] = c->a_lf[k];
        a.lNs[k] = c->a_lNs[k];
        a.lw[k] = c->a_lw[k];
        a.lUs[k] = c->a_lUs[k];
        a.lWs[k] = c->a_lWs[k];
    }
    a.lbp = c->a_lbp;
    a.Jpart = c->Jpart;
    return a;
}

// Adjoint of step `it` on the adjoint planes (a_lu, a_lup, a_lnp, a_lbp), accumulating the
// body-force gradient into a_lf when with_force.  Zuw holds the step's synthesised velocity and
// curl in the Z layout (six fields: u, v, w, wx, wy, wz), as the forward step left them.
void adjoint_step(cdns_ctx *c, int it, const void *Zuw, bool with_force) {
    const Geometry &g = c->g;
    double co[5];
    step_coefficients(it, c->cfg.DT, c->cfg.Re, co);
    ensure_tables(c, g_for_iteration(c, it));
    AdjArgs a = adj_args(c);
    a.a0 = co[0];
    a.a1 = co[1];
    a.b0 = co[2];
    a.b1 = co[3];
    if (!with_force)
        for (int k = 0; k < 3; ++k) a.lf[k] = nullptr;
    {
        Launch L(c, "adj_solve");
        launch_adj_solve(a, g.E, c->st);
    }
    if (a.collide) {
        Launch L(c, "adj_fold");
        launch_adj_fold(a, c->st);
    }
    // transposed nonlinear term: Lambda = synthesis of the weighted adjoint, then
    // lU = W x Lambda and lW = Lambda x U, analysed back
    backward_yz(c, c->a_lNs, 3, c->ZL);
    const size_t fs = g.ZS * c->esz;
    auto fld = [&](const void *base, int i) { return (const void *)((const char *)base + i * fs); };
    {
        XArgs x = xargs(c, nullptr, nullptr, 0, 0, c->P);
        for (int i = 0; i < 3; ++i) {
            x.fin[i] = fld(Zuw, 3 + i);
            x.fin[3 + i] = fld(c->ZL, i);
            x.fout[i] = (void *)fld(c->ZNa, i);
        }
        Launch L(c, "adj_x_product");
        launch_xprod(x, c->fp32, c->st);
    }
    {
        XArgs x = xargs(c, nullptr, nullptr, 0, 0, c->P);
        for (int i = 0; i < 3; ++i) {
            x.fin[i] = fld(c->ZL, i);
            x.fin[3 + i] = fld(Zuw, i);
            x.fout[i] = (void *)fld(c->ZNa, 3 + i);
        }
        Launch L(c, "adj_x_product");
        launch_xprod(x, c->fp32, c->st);
    }
    zd *outs[6] = {c->a_lUs[0], c->a_lUs[1], c->a_lUs[2], c->a_lWs[0], c->a_lWs[1], c->a_lWs[2]};
    forward_zy(c, outs, 6, c->ZNa);
    {
        Launch L(c, "adj_combine");
        launch_adj_combine(a, g.E, c->st);
    }
}

// Terminal objective of the velocity planes u: gradient into grad planes, J returned.
double objective(cdns_ctx *c, int which, zd *const *u, zd *const *grad) {
    if (which != 0 && which != 1) throw std::invalid_argument("objective: 0 kinetic energy, 1 mean wall shear");
    AdjArgs a = adj_args(c);
    a.which = which;
    for (int k = 0; k < 3; ++k) {
        a.lu[k] = u[k];
        a.grad[k] = grad[k];
    }
    {
        Launch L(c, "objective");
        launch_objective(a, c->g.E, c->st);
    }
    std::vector<double> &h = c->Jhost;
    CDNS_HIP(hipMemcpyAsync(h.data(), c->Jpart, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->st));
    CDNS_HIP(hipStreamSynchronize(c->st));
    double J = 0.0;
    for (double v : h) J += v;
    return J;
}

void zero_planes(cdns_ctx *c, zd *const *pl, int n, size_t count) {
    for (int k = 0; k < n; ++k) CDNS_HIP(hipMemsetAsync(pl[k], 0, count * sizeof(zd), c->st));
}

// ---------------------------------------------------- layout conversion --
// Reference-layout spectral array (host or device) -> internal planes.
void ref_to_planes(cdns_ctx *c, const double *src, int stride, int ncomp, int nrow, zd *const *planes, int where,
                   int JS) {
    const Geometry &g = c->g;
    const size_t n = (size_t)g.nzk * nrow * g.nkx;
    const zd *dev = reinterpret_cast<const zd *>(src);
    int dstride = stride;
    if (where == CDNS_HOST) {
        // gather the (possibly strided) components into the staging buffer
        std::vector<zd> tmp(n * ncomp);
        const zd *h = reinterpret_cast<const zd *>(src);
        for (size_t e = 0; e < n; ++e)
            for (int k = 0; k < ncomp; ++k) tmp[e * ncomp + k] = h[e * stride + k];
        zd *stage = nrow == 2 ? c->bnd_stage : c->ref_stage;
        CDNS_HIP(hipMemcpyAsync(stage, tmp.data(), tmp.size() * sizeof(zd), hipMemcpyHostToDevice, c->st));
        CDNS_HIP(hipStreamSynchronize(c->st));
        dev = stage;
        dstride = ncomp;
    }
    LayoutArgs a{};
    a.ref_in = dev;
    a.stride = dstride;
    a.ncomp = ncomp;
    a.nrow = nrow;
    for (int k = 0; k < ncomp; ++k) a.plane_out[k] = planes[k];
    a.nkx = g.nkx;
    a.nzk = g.nzk;
    a.JS = JS;
    Launch L(c, "layout_in");
    launch_ref_to_planes(a, c->st);
}

void planes_to_ref(cdns_ctx *c, const zd *const *planes, int ncomp, int nrow, double *dst, int stride, int where,
                   int JS) {
    const Geometry &g = c->g;
    const size_t n = (size_t)g.nzk * nrow * g.nkx;
    LayoutArgs a{};
    a.nrow = nrow;
    a.ncomp = ncomp;
    for (int k = 0; k < ncomp; ++k) a.plane_in[k] = planes[k];
    a.nkx = g.nkx;
    a.nzk = g.nzk;
    a.JS = JS;
    if (where == CDNS_DEVICE) {
        a.ref_out = reinterpret_cast<zd *>(dst);
        a.stride = stride;
        Launch L(c, "layout_out");
        launch_planes_to_ref(a, c->st);
        return;
    }
    zd *stage = nrow == 2 ? c->bnd_stage : c->ref_stage;
    a.ref_out = stage;
    a.stride = ncomp;
    {
        Launch L(c, "layout_out");
        launch_planes_to_ref(a, c->st);
    }
    std::vector<zd> tmp(n * ncomp);
    CDNS_HIP(hipMemcpyAsync(tmp.data(), stage, tmp.size() * sizeof(zd), hipMemcpyDeviceToHost, c->st));
    CDNS_HIP(hipStreamSynchronize(c->st));
    zd *h = reinterpret_cast<zd *>(dst);
    for (size_t e = 0; e < n; ++e)
        for (int k = 0; k < ncomp; ++k) h[e * stride + k] = tmp[e * ncomp + k];
}

int fail(cdns_ctx *c, const std::exception &e) {
    if (c) c->err = e.what();
    return -1;
}

// Storage of one adjoint horizon: a checkpoint (u, u_prev, N_prev: 9 planes, and B_prev) every
// k_ckpt steps, and one segment's synthesised (u, curl u) in the Z layout per step.
struct HorizonSize {
    int kc = 0, nck = 0;
    size_t ckn = 0;       // complex entries per checkpoint
    size_t ck_bytes = 0;  // all checkpoints
    size_t seg_bytes = 0; // one segment
};

HorizonSize horizon_size(const cdns_ctx *c, int nsteps, int k_ckpt) {
    if (nsteps < 1 || k_ckpt < 1) throw std::invalid_argument("nsteps and the checkpoint interval must be >= 1");
    const Geometry &g = c->g;
    HorizonSize h;
    h.kc = std::min(k_ckpt, nsteps);
    h.nck = (nsteps + h.kc - 1) / h.kc;
    h.ckn = (size_t)9 * g.S + (size_t)g.NQ * 2;
    h.ck_bytes = (size_t)h.nck * h.ckn * sizeof(zd);
    h.seg_bytes = (size_t)h.kc * 6 * g.ZS * c->esz;
    return h;
}

// Gradient of objective(u(T)) over iterations 1..nsteps from u0 with the body force `force` (or
// none), with the horizon storage at ck and seg (device memory of horizon_size's sizes).  The
// context's state is overwritten; its time and body force are what they were before.
void adjoint_horizon(cdns_ctx *c, const double *u0, const double *force, int nsteps, int k_ckpt, int objective_id,
                     double *grad_u0, double *grad_force, double *J, int where, void *ck, void *seg) {
    const Geometry &g = c->g;
    const HorizonSize hs = horizon_size(c, nsteps, k_ckpt);
    ensure_adjoint(c);
    const double time0 = c->time;
    const bool had_body = c->body_on;
    void *const zkeep = c->Zbuf;
    const size_t bplane = g.S * sizeof(zd);
    auto restore_context = [&]() {
        c->Zbuf = zkeep;
        c->time = time0;
        if (force && had_body)
            for (int k = 0; k < 3; ++k)
                CDNS_HIP(hipMemcpyAsync(c->body[k], c->body_save[k], bplane, hipMemcpyDeviceToDevice, c->st));
        c->body_on = had_body;
    };
    try {
        if (force) {
            if (had_body)
                for (int k = 0; k < 3; ++k)
                    CDNS_HIP(hipMemcpyAsync(c->body_save[k], c->body[k], bplane, hipMemcpyDeviceToDevice, c->st));
            if (cdns_set_body_force(c, force, where) != 0) throw std::runtime_error(c->err);
        } else {
            c->body_on = false;
        }
        // initial state: u0, no history (steps 1 and 2 weight it by zero)
        ref_to_planes(c, u0, 3, 3, g.nj, c->u, where, g.JS);
        zero_planes(c, c->up, 3, g.S);
        zero_planes(c, c->nlp, 3, g.S);
        CDNS_HIP(hipMemsetAsync(c->bc_prev, 0, (size_t)g.NQ * 2 * sizeof(zd), c->st));
        const int kc = hs.kc, nck = hs.nck;
        const size_t ckn = hs.ckn;
        auto save = [&](int k) {
            zd *base = reinterpret_cast<zd *>(ck) + (size_t)k * ckn;
            for (int i = 0; i < 3; ++i) {
                CDNS_HIP(hipMemcpyAsync(base + (size_t)i * g.S, c->u[i], bplane, hipMemcpyDeviceToDevice, c->st));
                CDNS_HIP(hipMemcpyAsync(base + (size_t)(3 + i) * g.S, c->up[i], bplane, hipMemcpyDeviceToDevice, c->st));
                CDNS_HIP(hipMemcpyAsync(base + (size_t)(6 + i) * g.S, c->nlp[i], bplane, hipMemcpyDeviceToDevice, c->st));
            }
            CDNS_HIP(hipMemcpyAsync(base + (size_t)9 * g.S, c->bc_prev, (size_t)g.NQ * 2 * sizeof(zd),
                                    hipMemcpyDeviceToDevice, c->st));
        };
        auto restore = [&](int k) {
            const zd *base = reinterpret_cast<const zd *>(ck) + (size_t)k * ckn;
            for (int i = 0; i < 3; ++i) {
                CDNS_HIP(hipMemcpyAsync(c->u[i], base + (size_t)i * g.S, bplane, hipMemcpyDeviceToDevice, c->st));
                CDNS_HIP(hipMemcpyAsync(c->up[i], base + (size_t)(3 + i) * g.S, bplane, hipMemcpyDeviceToDevice, c->st));
                CDNS_HIP(hipMemcpyAsync(c->nlp[i], base + (size_t)(6 + i) * g.S, bplane, hipMemcpyDeviceToDevice, c->st));
            }
            CDNS_HIP(hipMemcpyAsync(c->bc_prev, base + (size_t)9 * g.S, (size_t)g.NQ * 2 * sizeof(zd),
                                    hipMemcpyDeviceToDevice, c->st));
        };
        for (int it = 1; it <= nsteps; ++it) {
            if ((it - 1) % kc == 0) save((it - 1) / kc);
            step(c, it);
        }
        // terminal objective and its gradient start the adjoint
        const double Jv = objective(c, objective_id, c->u, c->a_lu);
        if (J) *J = Jv;
        zero_planes(c, c->a_lup, 3, g.S);
        zero_planes(c, c->a_lnp, 3, g.S);
        zero_planes(c, c->a_lf, 3, g.S);
        CDNS_HIP(hipMemsetAsync(c->a_lbp, 0, (size_t)g.NQ * 2 * sizeof(zd), c->st));
        const size_t zstep = (size_t)6 * g.ZS * c->esz;
        for (int k = nck - 1; k >= 0; --k) {
            const int it0 = k * kc + 1, it1 = std::min((k + 1) * kc, nsteps);
            restore(k);
            // recompute the segment, keeping each step's synthesised (u, curl u)
            for (int it = it0; it <= it1; ++it) {
                c->Zbuf = (char *)seg + (size_t)(it - it0) * zstep;
                step(c, it);
            }
            c->Zbuf = zkeep;
            for (int it = it1; it >= it0; --it)
                adjoint_step(c, it, (const char *)seg + (size_t)(it - it0) * zstep, force != nullptr);
        }
        restore_context();
        // leave the tables of the steady iterations in place for later forward steps
        if (nsteps >= 2) ensure_tables(c, g_for_iteration(c, 2));
        const zd *o1[3] = {c->a_lu[0], c->a_lu[1], c->a_lu[2]};
        planes_to_ref(c, o1, 3, g.nj, grad_u0, 3, where, g.JS);
        if (grad_force) {
            const zd *o2[3] = {c->a_lf[0], c->a_lf[1], c->a_lf[2]};
            planes_to_ref(c, o2, 3, g.nj, grad_force, 3, where, g.JS);
        }
        CDNS_HIP(hipStreamSynchronize(c->st));
    } catch (...) {
        (void)hipStreamSynchronize(c->st);
        c->Zbuf = zkeep;
        c->time = time0;
        c->body_on = had_body && !force; // a body force half-replaced is not restored
        throw;
    }
}

} // namespace

// ======================================================================= ABI

extern "C" {

int cdns_create(const cdns_config *cfg, cdns_ctx **out) {
    if (out) *out = nullptr;
    if (!cfg || !out) return -1;
    auto c = std::make_unique<cdns_ctx>();
    try {
        c->cfg = *cfg;
        Geometry &g = c->g;
        g.NX = cfg->NX; g.NY = cfg->NY; g.NZ = cfg->NZ;
        g.NXS = cfg->NXS; g.NYS = cfg->NYS; g.NZS = cfg->NZS;
        g.NX2 = cfg->NX2; g.NY2 = cfg->NY2; g.NZ2 = cfg->NZ2;
        if (g.NYS < 4 || g.NXS < 0 || g.NZS < 0 || 2 * g.NXS > g.NX2 || g.NY2 < g.NYS || 2 * g.NZS > g.NZ2 ||
            g.NZ2 < 1 || g.NYS > 255)
            throw std::invalid_argument("unsupported mesh sizes (need 2 NXS <= NX2, NYS <= NY2, 2 NZS <= NZ2, 4 <= NYS < 256)");
        if (cfg->precision != CDNS_FP64 && cfg->precision != CDNS_MIXED)
            throw std::invalid_argument("precision must be CDNS_FP64 or CDNS_MIXED");
        c->fp32 = cfg->precision == CDNS_MIXED;
        // the mixed step keeps its nonlinear term (the FP32 transforms' output) and that term's
        // history in single precision; CDNS_MIXED_N=f64 keeps them in FP64
        c->nl32 = c->fp32;
        if (const char *e = getenv("CDNS_MIXED_N")) c->nl32 = c->fp32 && std::string(e) != "f64";
        if (const char *e = getenv("CDNS_FFT")) c->rf = std::string(e) != "stockham";
        if (const char *e = getenv("CDNS_YSTAGE")) {
            c->ymfma = std::string(e) != "vector";
            c->ydct = std::string(e) == "dct";
        }
        if (const char *e = getenv("CDNS_YCURL")) c->ycurl = std::string(e) != "0";
        if (const char *e = getenv("CDNS_XPRODUCT")) c->xpair = std::string(e) != "single";
        if (const char *e = getenv("CDNS_ZSTAGE")) c->zpair = std::string(e) != "single";
        c->esz = c->fp32 ? 8 : 16;
        g.nkx = g.NXS + 1;
        g.nzk = 2 * g.NZS + 1;
        g.nj = g.NYS + 1;
        g.E = g.nj <= 128 ? 2 : 4;
        g.JS = (g.nj + 7) / 8 * 8;
        g.NQ = g.nkx * g.nzk;
        g.NL = g.NY2 + 1;
        g.NLH = g.NY2 / 2 + 1;
        g.S = (size_t)g.NQ * g.JS;
        g.YS = (size_t)g.NL * g.NQ;
        g.ZS = (size_t)g.NL * g.NZ2 * g.nkx;
        CDNS_HIP(hipStreamCreateWithFlags(&c->st, hipStreamNonBlocking));
        CDNS_HIP(hipGetDevice(&c->device));
        c->noalloc_scratch = scratch_address(c->device);
        cdns_ctx *C = c.get();

        // wavenumbers and Chebyshev weights (transforms.cpp:71-108)
        const double PI = std::acos(-1.0);
        const double ALPHAX = 2.0 * PI / cfg->XL, ALPHAZ = 2.0 * PI / cfg->ZL;
        C->hWAVX.resize(g.nkx);
        C->hWAVZ.resize(g.nzk);
        for (int k = 0; k < g.nkx; ++k) C->hWAVX[k] = (0 + k) * ALPHAX;
        for (int k = -g.NZS; k <= g.NZS; ++k) C->hWAVZ[k + g.NZS] = k * ALPHAZ;
        std::vector<double> WL(g.nj, 0.0), WD(g.nj, 0.0), WR(g.nj, 0.0);
        for (int iy = g.NYS; iy >= 2; --iy) {
            WL[iy] = 1.0 / (4.0 * iy * (iy - 1));
            WD[iy] = -1.0 / (2.0 * (iy * iy - 1));
            WR[iy] = 1.0 / (4.0 * iy * (iy + 1));
        }
        WL[2] = 2.0 * WL[2];
        for (int iy = g.NYS - 1; iy <= g.NYS; ++iy) WD[iy] = 0.0;
        for (int iy = g.NYS - 3; iy <= g.NYS; ++iy) WR[iy] = 0.0;
        auto upload = [&](const std::vector<double> &h) {
            double *d = C->alloc<double>(h.size());
            CDNS_HIP(hipMemcpyAsync(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice, C->st));
            return d;
        };
        C->WAVX = upload(C->hWAVX);
        C->WAVZ = upload(C->hWAVZ);
        C->WL = upload(WL);
        C->WD = upload(WD);
        C->WR = upload(WR);

        // y-stage matrices from exact-angle cosines
        std::vector<double> B, Fe, Fo;
        y_tables(g.NY2, g.nj, g.NLH, B, Fe, Fo);
        C->ycoef_b = upload(B);
        C->ycoef_fe = upload(Fe);
        C->ycoef_fo = upload(Fo);
        std::vector<double> D;
        y_dtable(g.NY2, g.nj, g.NLH, D);
        C->ycoef_d = upload(D);

        // FFT plans and twiddles
        C->pz = make_plan(g.NZ2);
        C->px = make_plan(g.NX2);
        auto up_tw = [&](int n) -> void * {
            const auto t = twiddles(n);
            if (C->fp32) {
                std::vector<float> f(t.size());
                for (size_t i = 0; i < t.size(); ++i) f[i] = (float)t[i];
                float *d = C->alloc<float>(f.size());
                CDNS_HIP(hipMemcpyAsync(d, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice, C->st));
                CDNS_HIP(hipStreamSynchronize(C->st));
                return d;
            }
            std::vector<double> dv(t.size());
            for (size_t i = 0; i < t.size(); ++i) dv[i] = (double)t[i];
            double *d = upload(dv);
            CDNS_HIP(hipStreamSynchronize(C->st));
            return d;
        };
        C->twz = up_tw(g.NZ2);
        C->twx = up_tw(g.NX2);
        C->ydct = C->ydct && g.nj <= 256 && dct_supported(2 * g.NY2);
        if (C->ydct) C->ydct_tw = up_tw(2 * g.NY2);
        // z lines per workgroup: at most 4 butterflies per thread per stage
        auto ipt = [](int lines, const FftPlan &p) { return (lines * (p.n / plan_min_radix(p)) + 255) / 256; };
        int T = (g.nkx + 15) / 16;
        C->KT = (g.nkx + T - 1) / T;
        while (ipt(C->KT, C->pz) > 4 && C->KT > 1) {
            ++T;
            C->KT = (g.nkx + T - 1) / T;
        }
        // x-product points per workgroup (even), LDS <= 64 KB and <= 4 butterflies per thread
        C->P = 2;
        for (int P = 8; P >= 2; P -= 2) {
            if (ipt(3 * P, C->px) <= 4 && (size_t)3 * P * g.NX2 * C->esz <= 65536 && (P / 2) * g.NX2 <= 8 * 256) {
                C->P = P;
                break;
            }
        }
        if (ipt(3 * C->P, C->px) > 4 || ipt(C->KT, C->pz) > 4)
            throw std::invalid_argument("product grid too large for one workgroup per line batch");

        // solver tables
        C->ad_p = C->alloc<double>(g.S);
        C->ad_v = C->alloc<double>(g.S);
        C->wn_p = C->alloc<double>(g.NQ);
        C->wn_v = C->alloc<double>(g.NQ);
        for (int i = 0; i < 2; ++i) {
            C->PI[i] = C->alloc<zd>(g.S);
            for (int k = 0; k < 3; ++k) C->VI[i][k] = C->alloc<zd>(g.S);
            C->ID[i] = C->alloc<zd>((size_t)g.NQ * 2);
        }
        C->PB = C->alloc<zd>(g.S);
        for (int k = 0; k < 3; ++k) C->VB[k] = C->alloc<zd>(g.S);
        C->SIG0 = C->alloc<zd>((size_t)g.NQ * 2);
        C->compact = cfg->a_V[0] == cfg->a_V[2] && cfg->b_V[0] == cfg->b_V[2];
        if (const char *e = getenv("CDNS_SOLVER_TABLES")) C->compact = C->compact && std::string(e) != "complex";
        if (C->compact)
            for (int i = 0; i < 9; ++i) C->ctab[i] = C->alloc<double>(g.S);
        C->u_unit = C->alloc<double>(g.JS);
        C->w_unit = C->alloc<double>(g.JS);
        C->q_unit = C->alloc<double>(2);
        C->pa = cfg->a_pre0;
        C->pb = cfg->b_pre0;

        // state
        for (int k = 0; k < 3; ++k) {
            C->u[k] = C->alloc<zd>(g.S);
            C->up[k] = C->alloc<zd>(g.S);
            if (C->nl32) {
                C->nl[k] = reinterpret_cast<zd *>(C->alloc<cpx<float>>(g.S));
                C->nlp[k] = reinterpret_cast<zd *>(C->alloc<cpx<float>>(g.S));
            } else {
                C->nl[k] = C->alloc<zd>(g.S);
                C->nlp[k] = C->alloc<zd>(g.S);
            }
            C->om[k] = C->alloc<zd>(g.S);
        }
        C->P_ = C->alloc<zd>(g.S);
        C->bc_prev = C->alloc<zd>((size_t)g.NQ * 2);
        C->bc_ext = C->alloc<zd>((size_t)g.NQ * 2);
        C->bcw = C->alloc<zd>((size_t)g.NQ * 2);
        C->vbc = C->alloc<zd>((size_t)g.NQ * 2 * 3);
        C->flow = C->alloc<double>(2);
        // velocity_boundary (initial_bc.cpp:95-101): wall-normal velocity on the mean mode
        {
            std::vector<zd> h((size_t)g.NQ * 2 * 3, zd{0.0, 0.0});
            const size_t root = (size_t)g.NZS; // q of (kx=0, kz=0)
            h[(size_t)1 * g.NQ * 2 + root * 2 + 0] = {cfg->V_Microblow0, 0.0};
            h[(size_t)1 * g.NQ * 2 + root * 2 + 1] = {cfg->V_Microblow0, 0.0};
            CDNS_HIP(hipMemcpyAsync(C->vbc, h.data(), h.size() * sizeof(zd), hipMemcpyHostToDevice, C->st));
            CDNS_HIP(hipStreamSynchronize(C->st));
        }

        // workspaces
        C->Ybuf = C->alloc<unsigned char>(6 * g.YS * C->esz);
        C->Zbuf = C->alloc<unsigned char>(6 * g.ZS * C->esz);
        C->ZNbuf = C->alloc<unsigned char>(3 * g.ZS * C->esz);
        for (int i = 0; i < 8; ++i) C->scr[i] = C->alloc<zd>(g.S);
        C->ref_stage = C->alloc<zd>((size_t)3 * g.nzk * g.nj * g.nkx);
        C->bnd_stage = C->alloc<zd>((size_t)3 * g.nzk * 2 * g.nkx);
        C->phys_stage = C->alloc<double>((size_t)g.NZ2 * g.NL * g.NX2);
        CDNS_HIP(hipStreamSynchronize(C->st));
    } catch (const std::exception &e) {
        static std::string last;
        last = e.what();
        fprintf(stderr, "cdns_create: %s\n", e.what());
        cdns_destroy(c.release());
        return -1;
    }
    *out = c.release();
    return 0;
}

void cdns_destroy(cdns_ctx *c) {
    if (!c) return;
    if (c->st) (void)hipStreamSynchronize(c->st);
    for (auto &e : c->pending) {
        (void)hipEventDestroy(e.a);
        (void)hipEventDestroy(e.b);
    }
    for (auto e : c->pool) (void)hipEventDestroy(e);
    for (void *p : c->allocs) (void)hipFree(p);
    if (c->hz_ck) (void)hipFree(c->hz_ck);
    if (c->hz_seg) (void)hipFree(c->hz_seg);
    if (c->st) (void)hipStreamDestroy(c->st);
    delete c;
}

const char *cdns_last_error(const cdns_ctx *c) { return c ? c->err.c_str() : "null context"; }

struct cdns_engine {
    std::mt19937_64 g;
};

int cdns_engine_create(unsigned long long seed, cdns_engine **out) {
    if (!out) return 1;
    *out = new (std::nothrow) cdns_engine{std::mt19937_64(seed)};
    return *out ? 0 : 1;
}

int cdns_engine_draw(cdns_engine *e, int kind, void *out, unsigned long long n) {
    if (!e || (!out && n)) return 1;
    if (kind == 0) {
        auto *o = static_cast<unsigned long long *>(out);
        for (unsigned long long i = 0; i < n; ++i) o[i] = e->g();
    } else if (kind == 1) {
        auto *o = static_cast<double *>(out);
        for (unsigned long long i = 0; i < n; ++i) o[i] = static_cast<double>(e->g() >> 11) * 0x1.0p-53 - 0.5;
    } else {
        return 1;
    }
    return 0;
}

void cdns_engine_destroy(cdns_engine *e) { delete e; }

void *cdns_stream(cdns_ctx *c) { return c ? (void *)c->st : nullptr; }

int cdns_sync(cdns_ctx *c) {
    try {
        CDNS_HIP(hipStreamSynchronize(c->st));
        if (c->prof) drain_profile(c);
        return 0;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

int cdns_xyzfft2(cdns_ctx *c, char dir, double *phys, double *spec, int stride, int where) {
    try {
        if (stride < 1) throw std::invalid_argument("spectral stride must be positive");
        if (dir != 'F' && dir != 'B') throw std::invalid_argument("direction must be 'F' or 'B'");
        const Geometry &g = c->g;
        const size_t np = (size_t)g.NZ2 * g.NL * g.NX2;
        double *dphys = where == CDNS_DEVICE ? phys : c->phys_stage;
        if (dir == 'B') {
            zd *pl[1] = {c->scr[0]};
            ref_to_planes(c, spec, stride, 1, g.nj, pl, where, g.JS);
            const zd *cp[1] = {c->scr[0]};
            backward_yz(c, cp, 1);
            XArgs a = xargs(c, c->Zbuf, nullptr, g.ZS, g.ZS, 2 * c->P);
            a.phys = dphys;
            {
                Launch L(c, "x_backward");
                launch_xbwd_phys(a, c->fp32, c->st);
            }
            if (where == CDNS_HOST) {
                CDNS_HIP(hipMemcpyAsync(phys, dphys, np * sizeof(double), hipMemcpyDeviceToHost, c->st));
                CDNS_HIP(hipStreamSynchronize(c->st));
            }
        } else {
            if (where == CDNS_HOST)
                CDNS_HIP(hipMemcpyAsync(dphys, phys, np * sizeof(double), hipMemcpyHostToDevice, c->st));
            XArgs a = xargs(c, nullptr, c->ZNbuf, g.ZS, g.ZS, 2 * c->P);
            a.phys = dphys;
            {
                Launch L(c, "x_forward");
                launch_xfwd_phys(a, c->fp32, c->st);
            }
            zd *pl[1] = {c->scr[0]};
            forward_zy(c, pl, 1);
            const zd *cp[1] = {c->scr[0]};
            planes_to_ref(c, cp, 1, g.nj, spec, stride, where, g.JS);
        }
        if (where == CDNS_HOST) CDNS_HIP(hipStreamSynchronize(c->st));
        return 0;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

int cdns_xyzfft(cdns_ctx *c, char dir, double *phys, double *spec, int stride, int where) {
    try {
        if (stride < 1) throw std::invalid_argument("spectral stride must be positive");
        if (dir != 'F' && dir != 'B') throw std::invalid_argument("direction must be 'F' or 'B'");
        ensure_coarse(c);
        const Geometry &g = c->g;
        const size_t np = coarse_points(c);
        double *dphys = where == CDNS_DEVICE ? phys : c->cphys + 4 * np;
        if (dir == 'B') {
            zd *pl[1] = {c->scr[0]};
            ref_to_planes(c, spec, stride, 1, g.nj, pl, where, g.JS);
            coarse_backward(c, c->scr[0], dphys);
            if (where == CDNS_HOST)
                CDNS_HIP(hipMemcpyAsync(phys, dphys, np * sizeof(double), hipMemcpyDeviceToHost, c->st));
        } else {
            if (where == CDNS_HOST)
                CDNS_HIP(hipMemcpyAsync(dphys, phys, np * sizeof(double), hipMemcpyHostToDevice, c->st));
            coarse_forward(c, dphys, c->scr[0]);
            const zd *cp[1] = {c->scr[0]};
            planes_to_ref(c, cp, 1, g.nj, spec, stride, where, g.JS);
        }
        if (where == CDNS_HOST) CDNS_HIP(hipStreamSynchronize(c->st));
        return 0;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

int cdns_to_physical(cdns_ctx *c, double *velocity, double *pressure, int *finite, int where) {
    try {
        const bool ok = state_to_physical(c);
        const size_t np = coarse_points(c);
        const hipMemcpyKind kind = where == CDNS_HOST ? hipMemcpyDeviceToHost : hipMemcpyDeviceToDevice;
        if (velocity) CDNS_HIP(hipMemcpyAsync(velocity, c->cphys, 3 * np * sizeof(double), kind, c->st));
        if (pressure) CDNS_HIP(hipMemcpyAsync(pressure, c->cphys + 3 * np, np * sizeof(double), kind, c->st));
        CDNS_HIP(hipStreamSynchronize(c->st));
        if (finite) *finite = ok ? 1 : 0;
        return 0;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

int cdns_plane_moments(cdns_ctx *c, int higher, double *sums, int *finite) {
    try {
        state_to_physical(c);
        std::vector<double> s;
        std::vector<int> f;
        plane_moments(c, higher != 0, false, s, f);
        const int NLc = c->g.NY + 1;
        for (int iy = 0; iy < NLc; ++iy) {
            for (int m = 0; m < 20; ++m) sums[(size_t)iy * 20 + m] = s[(size_t)iy * kPlaneSums + m];
            if (finite) finite[iy] = f[iy];
        }
        return 0;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

int cdns_diagnostics(cdns_ctx *c, double out[3], int *finite) {
    try {
        const Geometry &g = c->g;
        bool ok = state_to_physical(c);
        DiagArgs d{};
        d.u = c->u[0];
        d.v = c->u[1];
        d.w = c->u[2];
        d.out = c->scr[0];
        d.WAVX = c->WAVX;
        d.WAVZ = c->WAVZ;
        d.NQ = g.NQ;
        d.nzk = g.nzk;
        d.JS = g.JS;
        d.ny = g.NYS;
        {
            Launch L(c, "divergence");
            launch_divergence(d, g.E, c->st);
        }
        coarse_backward(c, c->scr[0], c->cphys + 4 * coarse_points(c));
        std::vector<double> s;
        std::vector<int> f;
        plane_moments(c, false, true, s, f);
        // Diagnostics::evaluate (diagnostics.cpp:39-75)
        const double inv_plane = 1.0 / ((double)g.NX * g.NZ);
        double ekt = 0.0, tke = 0.0, div2 = 0.0;
        for (int iy = 0; iy <= g.NY; ++iy) {
            const double *p = &s[(size_t)iy * kPlaneSums];
            ok = ok && f[iy];
            ekt += p[5];
            double variance = 0.0;
            for (int k = 0; k < 3; ++k) {
                const double mean = p[k] * inv_plane;
                const double raw = p[4 + k] * inv_plane - mean * mean;
                ok = ok && std::isfinite(raw);
                variance += std::max(0.0, raw);
            }
            tke += 0.5 * c->cc_weights[iy] * variance;
            div2 += c->cc_weights[iy] * inv_plane * p[20];
        }
        out[0] = ekt;
        out[1] = tke;
        out[2] = std::sqrt(div2);
        ok = ok && std::isfinite(out[0]) && std::isfinite(out[1]) && std::isfinite(out[2]);
        if (finite) *finite = ok ? 1 : 0;
        return 0;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

int cdns_curl(cdns_ctx *c, double *out, const double *in, int where) {
    try {
        const Geometry &g = c->g;
        zd *a[3] = {c->scr[0], c->scr[1], c->scr[2]}, *b[3] = {c->scr[3], c->scr[4], c->scr[5]};
        ref_to_planes(c, in, 3, 3, g.nj, a, where, g.JS);
        curl_planes(c, a, b, nullptr);
        const zd *cb[3] = {b[0], b[1], b[2]};
        planes_to_ref(c, cb, 3, g.nj, out, 3, where, g.JS);
        if (where == CDNS_HOST) CDNS_HIP(hipStreamSynchronize(c->st));
        return 0;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

int cdns_cross_product(cdns_ctx *c, double *out, const double *av, const double *bv, int where) {
    try {
        const Geometry &g = c->g;
        zd *a[3] = {c->scr[0], c->scr[1], c->scr[2]}, *b[3] = {c->scr[3], c->scr[4], c->scr[5]};
        ref_to_planes(c, av, 3, 3, g.nj, a, where, g.JS);
        ref_to_planes(c, bv, 3, 3, g.nj, b, where, g.JS);
        zd *o[3] = {c->scr[0], c->scr[1], c->scr[2]};
        cross_planes(c, a, b, o);
        const zd *co[3] = {o[0], o[1], o[2]};
        planes_to_ref(c, co, 3, g.nj, out, 3, where, g.JS);
        if (where == CDNS_HOST) CDNS_HIP(hipStreamSynchronize(c->st));
        return 0;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

int cdns_nonlinear_term(cdns_ctx *c, double *nl, const double *velocity, double *pressure_bc, int where) {
    try {
        const Geometry &g = c->g;
        zd *u[3] = {c->scr[0], c->scr[1], c->scr[2]}, *n[3] = {c->scr[3], c->scr[4], c->scr[5]};
        ref_to_planes(c, velocity, 3, 3, g.nj, u, where, g.JS);
        nonlinear_planes(c, u, n);
        {
            Launch L(c, "wall_sum");
            launch_wall_sum(n[1], c->bcw, c->scr[6], g.NYS, g.JS, g.NQ, g.E, c->st);
        }
        const zd *cn[3] = {n[0], n[1], n[2]};
        planes_to_ref(c, cn, 3, g.nj, nl, 3, where, g.JS);
        const zd *pb[1] = {c->scr[6]};
        planes_to_ref(c, pb, 1, 2, pressure_bc, 1, where, 2);
        if (where == CDNS_HOST) CDNS_HIP(hipStreamSynchronize(c->st));
        return 0;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

int cdns_setup(cdns_ctx *c, double g_v) {
    try {
        setup(c, g_v);
        return 0;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

int cdns_solve_modes(cdns_ctx *c, double *pressure, const double *rhs_p, const double *pressure_bc, double *velocity,
                     const double *rhs_v, const double *velocity_bc, int correct_sigma, int where) {
    try {
        if (!c->setup_done) throw std::logic_error("cdns_setup must run before cdns_solve_modes");
        const Geometry &g = c->g;
        zd *rv[3] = {c->scr[0], c->scr[1], c->scr[2]};
        ref_to_planes(c, rhs_v, 3, 3, g.nj, rv, where, g.JS);
        zd *rp = nullptr;
        if (rhs_p) {
            zd *pl[1] = {c->scr[3]};
            ref_to_planes(c, rhs_p, 1, 1, g.nj, pl, where, g.JS);
            rp = c->scr[3];
        }
        // boundaries: pressure [q][2] in scr[6]; velocity [c][q][2] in a temporary set of planes
        zd *pbc[1] = {c->scr[6]};
        ref_to_planes(c, pressure_bc, 1, 1, 2, pbc, where, 2);
        zd *vb = c->scr[7];
        zd *vbp[3] = {vb, vb + (size_t)g.NQ * 2, vb + (size_t)g.NQ * 4};
        ref_to_planes(c, velocity_bc, 3, 3, 2, vbp, where, 2);
        SolveArgs a = solve_args(c, MODE_GIVEN);
        for (int k = 0; k < 3; ++k) a.rv[k] = rv[k];
        a.rp = rp;
        a.pbc = c->scr[6];
        a.vbc = vb;
        a.correct = 1;
        a.sigma = correct_sigma ? 1 : 0;
        a.P = c->scr[4];
        zd *V[3] = {c->scr[0], c->scr[1], c->scr[5]};
        for (int k = 0; k < 3; ++k) a.V[k] = V[k];
        // V[0], V[1] alias the rhs planes of the same column: each wavefront reads its
        // column before it writes it.
        {
            Launch L(c, "solve_modes");
            launch_solve(a, g.E, c->st);
        }
        const zd *cp[1] = {c->scr[4]};
        planes_to_ref(c, cp, 1, g.nj, pressure, 1, where, g.JS);
        const zd *cv[3] = {V[0], V[1], V[2]};
        planes_to_ref(c, cv, 3, g.nj, velocity, 3, where, g.JS);
        if (where == CDNS_HOST) CDNS_HIP(hipStreamSynchronize(c->st));
        return 0;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

int cdns_get_table(cdns_ctx *c, int which, double *out) {
    try {
        const Geometry &g = c->g;
        CDNS_HIP(hipStreamSynchronize(c->st));
        auto real_plane = [&](const double *plane, int nrow, int JS) {
            double *d = reinterpret_cast<double *>(c->ref_stage);
            launch_real_planes_to_ref(plane, d, nrow, g.nkx, g.nzk, JS, c->st);
            CDNS_HIP(hipMemcpyAsync(out, d, (size_t)g.nzk * nrow * g.nkx * sizeof(double), hipMemcpyDeviceToHost, c->st));
        };
        switch (which) {
        case 0: real_plane(c->wn_p, 1, 1); break;
        case 1: real_plane(c->wn_v, 1, 1); break;
        case 2: real_plane(c->ad_p, g.nj, g.JS); break;
        case 3: real_plane(c->ad_v, g.nj, g.JS); break;
        case 4: case 5: { const zd *p[1] = {c->PI[which - 4]}; planes_to_ref(c, p, 1, g.nj, out, 1, CDNS_HOST, g.JS); break; }
        case 6: case 7: { const zd *p[3] = {c->VI[which - 6][0], c->VI[which - 6][1], c->VI[which - 6][2]}; planes_to_ref(c, p, 3, g.nj, out, 3, CDNS_HOST, g.JS); break; }
        case 8: case 9: { const zd *p[1] = {c->ID[which - 8]}; planes_to_ref(c, p, 1, 2, out, 1, CDNS_HOST, 2); break; }
        case 10: { const zd *p[1] = {c->PB}; planes_to_ref(c, p, 1, g.nj, out, 1, CDNS_HOST, g.JS); break; }
        case 11: { const zd *p[3] = {c->VB[0], c->VB[1], c->VB[2]}; planes_to_ref(c, p, 3, g.nj, out, 3, CDNS_HOST, g.JS); break; }
        case 12: { const zd *p[1] = {c->SIG0}; planes_to_ref(c, p, 1, 2, out, 1, CDNS_HOST, 2); break; }
        case 13: CDNS_HIP(hipMemcpyAsync(out, c->u_unit, g.nj * sizeof(double), hipMemcpyDeviceToHost, c->st)); break;
        case 14: CDNS_HIP(hipMemcpyAsync(out, c->w_unit, g.nj * sizeof(double), hipMemcpyDeviceToHost, c->st)); break;
        case 15: CDNS_HIP(hipMemcpyAsync(out, c->q_unit, 2 * sizeof(double), hipMemcpyDeviceToHost, c->st)); break;
        default: throw std::invalid_argument("unknown table");
        }
        CDNS_HIP(hipStreamSynchronize(c->st));
        return 0;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

static void state_planes(cdns_ctx *c, int field, zd **pl, int &ncomp, int &nrow, int &JS) {
    const Geometry &g = c->g;
    nrow = g.nj;
    JS = g.JS;
    ncomp = 3;
    switch (field) {
    case 0: for (int k = 0; k < 3; ++k) pl[k] = c->u[k]; break;
    case 1: for (int k = 0; k < 3; ++k) pl[k] = c->up[k]; break;
    case 2: for (int k = 0; k < 3; ++k) pl[k] = c->nl[k]; break;
    case 3: for (int k = 0; k < 3; ++k) pl[k] = c->nlp[k]; break;
    case 4: ncomp = 1; pl[0] = c->P_; break;
    case 5: ncomp = 1; nrow = 2; JS = 2; pl[0] = c->bc_ext; break;
    case 6: ncomp = 1; nrow = 2; JS = 2; pl[0] = c->bc_prev; break;
    case 7: nrow = 2; JS = 2; for (int k = 0; k < 3; ++k) pl[k] = c->vbc + (size_t)k * g.NQ * 2; break;
    default: throw std::invalid_argument("unknown state field");
    }
}

int cdns_set_state(cdns_ctx *c, int field, const double *data, int where) {
    try {
        zd *pl[3] = {};
        int ncomp, nrow, JS;
        state_planes(c, field, pl, ncomp, nrow, JS);
        if (c->nl32 && (field == 2 || field == 3)) {
            // single-precision planes: staged through FP64 planes
            zd *stage[3] = {c->scr[0], c->scr[1], c->scr[2]};
            ref_to_planes(c, data, ncomp, ncomp, nrow, stage, where, JS);
            for (int k = 0; k < 3; ++k) launch_narrow(stage[k], pl[k], c->g.S, c->st);
        } else {
            ref_to_planes(c, data, ncomp, ncomp, nrow, pl, where, JS);
        }
        if (where == CDNS_HOST) CDNS_HIP(hipStreamSynchronize(c->st));
        return 0;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

int cdns_get_state(cdns_ctx *c, int field, double *data, int where) {
    try {
        zd *pl[3] = {};
        int ncomp, nrow, JS;
        state_planes(c, field, pl, ncomp, nrow, JS);
        if (c->nl32 && (field == 2 || field == 3)) {
            for (int k = 0; k < 3; ++k) launch_widen(pl[k], c->scr[k], c->g.S, c->st);
            for (int k = 0; k < 3; ++k) pl[k] = c->scr[k];
        }
        const zd *cp[3] = {pl[0], pl[1], pl[2]};
        planes_to_ref(c, cp, ncomp, nrow, data, ncomp, where, JS);
        if (where == CDNS_HOST) CDNS_HIP(hipStreamSynchronize(c->st));
        return 0;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

int cdns_step(cdns_ctx *c, int it) {
    try {
        step(c, it);
        return 0;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

int cdns_run(cdns_ctx *c, int it_begin, int nsteps) {
    try {
        for (int i = 0; i < nsteps; ++i) step(c, it_begin + i);
        return 0;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

int cdns_flow_gradients(cdns_ctx *c, double out[2]) {
    try {
        CDNS_HIP(hipMemcpyAsync(out, c->flow, 2 * sizeof(double), hipMemcpyDeviceToHost, c->st));
        CDNS_HIP(hipStreamSynchronize(c->st));
        return 0;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

double cdns_time(const cdns_ctx *c) { return c ? c->time : 0.0; }

int cdns_profile(cdns_ctx *c, int enable) {
    try {
        CDNS_HIP(hipStreamSynchronize(c->st));
        drain_profile(c);
        c->prof = enable != 0;
        return 0;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

int cdns_profile_read(cdns_ctx *c, char *names, long long *counts, double *total_ms, int cap) {
    try {
        CDNS_HIP(hipStreamSynchronize(c->st));
        drain_profile(c);
        const int n = (int)c->kname.size();
        for (int i = 0; i < n && i < cap; ++i) {
            if (names) {
                std::memset(names + 64 * i, 0, 64);
                std::strncpy(names + 64 * i, c->kname[i].c_str(), 63);
            }
            if (counts) counts[i] = c->kcnt[i];
            if (total_ms) total_ms[i] = c->ktot[i];
        }
        return n;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

int cdns_profile_reset(cdns_ctx *c) {
    try {
        CDNS_HIP(hipStreamSynchronize(c->st));
        drain_profile(c);
        std::fill(c->ktot.begin(), c->ktot.end(), 0.0);
        std::fill(c->kcnt.begin(), c->kcnt.end(), 0LL);
        return 0;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

// ------------------------------------------ discrete adjoint (no reference counterpart) --
int cdns_set_body_force(cdns_ctx *c, const double *force, int where) {
    try {
        const Geometry &g = c->g;
        if (!force) {
            c->body_on = false;
            return 0;
        }
        if (!c->body[0])
            for (int k = 0; k < 3; ++k) c->body[k] = c->alloc<zd>(g.S);
        ref_to_planes(c, force, 3, 3, g.nj, c->body, where, g.JS);
        c->body_on = true;
        if (where == CDNS_HOST) CDNS_HIP(hipStreamSynchronize(c->st));
        return 0;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

int cdns_objective(cdns_ctx *c, int which, const double *velocity, double *grad, double *J, int where) {
    try {
        const Geometry &g = c->g;
        ensure_adjoint(c);
        zd *u[3] = {c->scr[0], c->scr[1], c->scr[2]}, *gp[3] = {c->scr[3], c->scr[4], c->scr[5]};
        ref_to_planes(c, velocity, 3, 3, g.nj, u, where, g.JS);
        const double v = objective(c, which, u, gp);
        if (J) *J = v;
        if (grad) {
            const zd *cg[3] = {gp[0], gp[1], gp[2]};
            planes_to_ref(c, cg, 3, g.nj, grad, 3, where, g.JS);
        }
        if (where == CDNS_HOST) CDNS_HIP(hipStreamSynchronize(c->st));
        return 0;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

int cdns_adjoint_step(cdns_ctx *c, int it, const double *velocity, double *adj_velocity, double *adj_velocity_prev,
                      double *adj_nonlinear_prev, double *adj_pressure_bc_prev, double *adj_force, int where) {
    try {
        const Geometry &g = c->g;
        ensure_adjoint(c);
        if (it < 1) throw std::invalid_argument("iteration numbers start at 1");
        // the step's synthesised velocity and curl, as the forward step computes them
        zd *u[3] = {c->scr[0], c->scr[1], c->scr[2]};
        ref_to_planes(c, velocity, 3, 3, g.nj, u, where, g.JS);
        curl_planes(c, u, c->om, nullptr);
        const zd *f6[6] = {u[0], u[1], u[2], c->om[0], c->om[1], c->om[2]};
        backward_yz(c, f6, 6);
        ref_to_planes(c, adj_velocity, 3, 3, g.nj, c->a_lu, where, g.JS);
        ref_to_planes(c, adj_velocity_prev, 3, 3, g.nj, c->a_lup, where, g.JS);
        ref_to_planes(c, adj_nonlinear_prev, 3, 3, g.nj, c->a_lnp, where, g.JS);
        ref_to_planes(c, adj_pressure_bc_prev, 1, 1, 2, &c->a_lbp, where, 2);
        if (adj_force) ref_to_planes(c, adj_force, 3, 3, g.nj, c->a_lf, where, g.JS);
        adjoint_step(c, it, c->Zbuf, adj_force != nullptr);
        const zd *o1[3] = {c->a_lu[0], c->a_lu[1], c->a_lu[2]};
        const zd *o2[3] = {c->a_lup[0], c->a_lup[1], c->a_lup[2]};
        const zd *o3[3] = {c->a_lnp[0], c->a_lnp[1], c->a_lnp[2]};
        const zd *o4[1] = {c->a_lbp};
        planes_to_ref(c, o1, 3, g.nj, adj_velocity, 3, where, g.JS);
        planes_to_ref(c, o2, 3, g.nj, adj_velocity_prev, 3, where, g.JS);
        planes_to_ref(c, o3, 3, g.nj, adj_nonlinear_prev, 3, where, g.JS);
        planes_to_ref(c, o4, 1, 2, adj_pressure_bc_prev, 1, where, 2);
        if (adj_force) {
            const zd *o5[3] = {c->a_lf[0], c->a_lf[1], c->a_lf[2]};
            planes_to_ref(c, o5, 3, g.nj, adj_force, 3, where, g.JS);
        }
        if (where == CDNS_HOST) CDNS_HIP(hipStreamSynchronize(c->st));
        return 0;
    } catch (const std::exception &e) {
        return fail(c, e);
    }
}

int cdns_adjoint_run(cdns_ctx *c, const double *u0, const double *force, int nsteps, int k_ckpt, int objective_id,
                     double *grad_u0, double *grad_force, double *J, int where) {
    try {
        if (!c) throw std::invalid_argument("null context");
        ensure_adjoint(c);
        const HorizonSize hs = horizon_size(c, nsteps, k_ckpt);
        // the horizon storage stays with the context and grows when a call needs more
        if (hs.ck_bytes > c->hz_ck_bytes) {
            CDNS_HIP(hipStreamSynchronize(c->st));
            if (c->hz_ck) CDNS_HIP(hipFree(c->hz_ck));
            c->hz_ck = nullptr;
            c->hz_ck_bytes = 0;
            CDNS_HIP(hipMalloc(&c->hz_ck, hs.ck_bytes));
            c->hz_ck_bytes = hs.ck_bytes;
        }
        if (hs.seg_bytes > c->hz_seg_bytes) {
            CDNS_HIP(hipStreamSynchronize(c->st));
            if (c->hz_seg) CDNS_HIP(hipFree(c->hz_seg));
            c->hz_seg = nullptr;
            c->hz_seg_bytes = 0;
            CDNS_HIP(hipMalloc(&c->hz_seg, hs.seg_bytes));
            c->hz_seg_bytes = hs.seg_bytes;
        }
        adjoint_horizon(c, u0, force, nsteps, k_ckpt, objective_id, grad_u0, grad_force, J, where, c->hz_ck,
                        c->hz_seg);
        return 0;
    } catch (const std::exception &e) {
        if (c && c->st) (void)hipStreamSynchronize(c->st);
        return fail(c, e);
    }
}

int cdns_adjoint_run_noalloc(cdns_ctx *c, const double *u0, const double *force, int nsteps, int k_ckpt,
                             int objective_id, double *grad_u0, double *grad_force, double *J, int where) {
    try {
        if (!c) throw std::invalid_argument("null context");
        if (where != CDNS_DEVICE)
            throw std::invalid_argument("cdns_adjoint_run_noalloc takes device pointers only (CDNS_DEVICE)");
        if (c->fp32) throw std::logic_error("the discrete adjoint runs on FP64 contexts");
        if (!c->adj_ready)
            throw std::logic_error("cdns_adjoint_run_noalloc: the context's adjoint workspaces are not set up (an "
                                   "earlier cdns_adjoint_run or cdns_adjoint_step on this context creates them)");
        if (force && !c->body[0])
            throw std::logic_error("cdns_adjoint_run_noalloc: no body-force buffer on this context (an earlier "
                                   "cdns_set_body_force or cdns_adjoint_run with a force creates it)");
        const HorizonSize hs = horizon_size(c, nsteps, k_ckpt);
        const size_t ck_bytes = (hs.ck_bytes + 255) / 256 * 256;
        if (ck_bytes + hs.seg_bytes > kNoallocScratchBytes)
            throw std::length_error("cdns_adjoint_run_noalloc: the horizon needs " +
                                    std::to_string(ck_bytes + hs.seg_bytes) + " bytes of storage, more than the " +
                                    std::to_string(kNoallocScratchBytes) + " of its scratch array");
        int dev = -1;
        CDNS_HIP(hipGetDevice(&dev));
        if (dev != c->device) throw std::logic_error("cdns_adjoint_run_noalloc: call it with the context's device current");
        char *base = static_cast<char *>(c->noalloc_scratch);
        adjoint_horizon(c, u0, force, nsteps, k_ckpt, objective_id, grad_u0, grad_force, J, where, base,
                        base + ck_bytes);
        return 0;
    } catch (const std::exception &e) {
        if (c && c->st) (void)hipStreamSynchronize(c->st);
        return fail(c, e);
    }
}

} // extern "C"
