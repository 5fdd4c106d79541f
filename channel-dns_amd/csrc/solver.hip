// Per-(kx,kz) Chebyshev-tau solver of the channel-dns step on gfx950: one 64-lane wavefront
// per wavenumber pair, the even Chebyshev chain in lanes 0-31 and the odd chain in lanes 32-63.
//
// The reference solves each column with sequential y recurrences over 8-lane kx tiles
// (EqSolvePV::solve_modes / solve_tau_tile, src/solver/eqsolve_pv.cpp:64-128, 311-450).
// The recurrences are linear in the data and couple one parity only, so each becomes one affine
// scan over the two 32-lane halves at once; the influence-matrix and sigma corrections are
// lane-parallel over y with wavefront reductions for the wall divergence.
#include "internal.hpp"
#include "split.hpp"

namespace cdns {

__device__ inline zd cdivr(zd z, double d) { return {z.x / d, z.y / d}; }

// solve_tau_column (eqsolve_pv.cpp:16-60) for NR right-hand sides that share the operator
// (wavenumber w, ad, WL/WD/WR) and differ in wall data and closure weights (a, b).  The two
// sweeps and the closure accumulators are affine scans whose multipliers depend on the operator
// only, so the NR right-hand sides ride one scan each; the closure rows j = 0 and j = 1 sit in
// entry 0 of lanes 0 and 32.
template <int NE, int NR>
__device__ void tau_solve(const zd (&r)[NR][NE], zd (&out)[NR][NE], const double (&ad)[NE], const double (&WL)[NE],
                          const double (&WD)[NE], const double (&WR)[NE], int lane, int ny, double w, double g,
                          const double (&a)[NR], const double (&b)[NR], bool zero_mode, const zd (&lower)[NR],
                          const zd (&upper)[NR]) {
    const int p = sp::parity(lane);
    const bool head = (lane & 31) == 0; // entry 0 of this lane is j = p
    zd R[NR][NE];
#pragma unroll
    for (int k = 0; k < NR; ++k)
#pragma unroll
        for (int e = 0; e < NE; ++e) R[k][e] = r[k][e];
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        if (g == 0.0 && a[k] == 0.0 && zero_mode) {
            // compatibility correction of the pure-Neumann mean pressure (eqsolve_pv.cpp:19-27)
            zd corr = {0.0, 0.0};
#pragma unroll
            for (int e = 0; e < NE; ++e) {
                const int j = sp::jof<NE>(lane, e);
                if (p == 0 && j >= 2 && j <= ny) corr = corr + cdivr(r[k][e], (double)(j * j - 1));
            }
            corr = wave_sum(corr);
            if (b[k] != 0.0) corr = corr + cdivr(lower[k] + upper[k], 2.0 * b[k]);
            if (lane == 0) R[k][0] = corr;
        }
    }
    // tau right-hand side and the two closure rows (eqsolve_pv.cpp:28-33)
    double adn[NE], WLn[NE];
    zd o[NR][NE];
#pragma unroll
    for (int e = 0; e < NE; ++e) {
        adn[e] = sp::chain_next<NE>(ad, e, lane);
        WLn[e] = sp::chain_next<NE>(WL, e, lane);
        const int j = sp::jof<NE>(lane, e);
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            const zd rm = sp::chain_prev<NE>(R[k], e, lane), rn = sp::chain_next<NE>(R[k], e, lane);
            const zd closure = p ? cdivr(lower[k] - upper[k], 2.0) : cdivr(lower[k] + upper[k], 2.0);
            if (e == 0 && head) o[k][e] = closure;
            else if (j <= ny) o[k][e] = (WL[e] * rm + WD[e] * R[k][e]) + WR[e] * rn;
            else o[k][e] = {0.0, 0.0};
        }
    }
    // backward sweep out(j) -= WR(j) w ad(j+2) out(j+2), 2 <= j <= ny-2 (eqsolve_pv.cpp:34-35), and
    // the closure accumulators ad01 (eqsolve_pv.cpp:37-50): A(j) = f(j) - WL(j+2) w ad(j+2) A(j+2)
    zd x[NR][NE];
    double A[NR][NE];
    {
        double m[NE], beta[NE], f[NR][NE];
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            const int j = sp::jof<NE>(lane, e);
            m[e] = (j >= 2 && j <= ny - 2) ? -(WR[e] * w * adn[e]) : 0.0;
            beta[e] = (j <= ny - 2) ? -(WLn[e] * w * adn[e]) : 0.0;
#pragma unroll
            for (int k = 0; k < NR; ++k) {
                const double fj = (a[k] == 0.0 && b[k] == 0.0) ? (double)((1 - p) * j * j + p) : a[k] + b[k] * j * j;
                f[k][e] = j <= ny ? fj : 0.0;
            }
        }
        sp::suffix_affine_c<NE, NR>(m, o, x, lane);
        sp::suffix_affine_n<NE, NR>(beta, f, A, lane);
    }
    // closure sums over j >= 2 of this chain
    double S[NR][2];
#pragma unroll
    for (int k = 0; k < NR; ++k) {
        zd s = {0.0, 0.0};
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            const int j = sp::jof<NE>(lane, e);
            if (j >= 2 && j <= ny) s = s + (A[k][e] * ad[e]) * x[k][e];
        }
        S[k][0] = s.x;
        S[k][1] = s.y;
    }
    sp::half_sum<2 * NR>(reinterpret_cast<double(&)[2 * NR]>(S));
    // forward sweep out(j) = (out(j) - WL(j) w out(j-2)) ad(j) (eqsolve_pv.cpp:58-59) from the
    // closure values out(0), out(1) (eqsolve_pv.cpp:51-57)
    double aa[NE];
    zd bb[NR][NE];
#pragma unroll
    for (int e = 0; e < NE; ++e) {
        const int j = sp::jof<NE>(lane, e);
        aa[e] = (j >= 2 && j <= ny) ? -(WL[e] * w * ad[e]) : 0.0;
#pragma unroll
        for (int k = 0; k < NR; ++k) {
            if (e == 0 && head) {
                const bool normalize = p == 0 && g == 0.0 && a[k] == 0.0 && zero_mode;
                const double inv = 1.0 / A[k][0];
                bb[k][e] = normalize ? zd{0.0, 0.0} : zd{(x[k][0].x - S[k][0]) * inv, (x[k][0].y - S[k][1]) * inv};
            } else if (j <= ny) {
                bb[k][e] = ad[e] * x[k][e];
            } else {
                bb[k][e] = {0.0, 0.0};
            }
        }
    }
    sp::prefix_affine_c<NE, NR>(aa, bb, out, lane);
}

// Fused coupled solve of one (kx,kz) pair: EqSolvePV::solve_modes (eqsolve_pv.cpp:311-450),
// with the step's RHS assembly (simulation.cpp:144-160), the nonlinear pressure wall condition
// (timestepper.cpp:26-43), correct_zero_mode (timestepper.cpp:60-68) and correct_flow
// (simulation.cpp:111-140) fused in MODE_STEP; the setup problems of eqsolve_pv.cpp:266-303
// and 452-477 in MODE_INFLUENCE / MODE_BPROBLEM.
template <int NE, int MODE, bool NL32 = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(NE == 2 ? 3 : 1))) k_solve(SolveArgs a) {
    const int lane = lane_id(), p = sp::parity(lane);
    // the column index is wavefront-uniform: made visible as such, the per-column scalars
    // (wavenumbers, wall values, influence and sigma entries) come through scalar loads.
    // Wavefront w -> column: the columns kz = -m and kz = +m of one kx on neighbouring wavefronts
    // of a block, then the kz = 0 columns.  The real tables (ad, and the compact influence and B
    // responses) depend on kx^2 + kz^2 only, bit for bit (the setup's arithmetic is symmetric
    // under kz -> -kz), so both columns of a pair read the pair's kz < 0 column (qt) and the
    // second read of each table line is served from the cache, not from HBM.
    const int w = __builtin_amdgcn_readfirstlane(blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6));
    if (w >= a.NQ) return;
    const int nzk = a.nzk, hz = (nzk - 1) >> 1, npair = (a.NQ / nzk) * 2 * hz;
    int kxi, kzi;
    if (w < npair) {
        const int pr = w >> 1, m = pr % hz + 1;
        kxi = pr / hz;
        kzi = (w & 1) ? hz + m : hz - m;
    } else {
        kxi = w - npair;
        kzi = hz;
    }
    const int q = kxi * nzk + kzi, qt = kxi * nzk + (kzi < hz ? kzi : nzk - 1 - kzi);
    const int ny = a.ny;
    constexpr int mode = MODE;
    const bool zero = kxi == 0 && kzi == hz;
    const double kx = a.WAVX[kxi], kz = a.WAVZ[kzi];
    const size_t off = (size_t)q * a.JS, toff = (size_t)qt * a.JS;
    double WL[NE], WD[NE], WR[NE], adp[NE], adv[NE];
    sp::load_rcol<NE>(a.WL, lane, ny, WL);
    sp::load_rcol<NE>(a.WD, lane, ny, WD);
    sp::load_rcol<NE>(a.WR, lane, ny, WR);
    sp::load_rcol<NE>(a.ad_p + toff, lane, ny, adp);
    sp::load_rcol<NE>(a.ad_v + toff, lane, ny, adv);

    // ---- velocity right-hand side
    zd rv[3][NE];
    zd nly[NE];
    if constexpr (mode == MODE_STEP) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            zd u[NE], up[NE], n[NE], np[NE];
            sp::load_col<NE>(a.u[k] + off, lane, ny, u);
            sp::load_col<NE>(a.up[k] + off, lane, ny, up);
            if constexpr (NL32) {
                sp::load_col<NE>(reinterpret_cast<const cpx<float> *>(a.nl[k]) + off, lane, ny, n);
                sp::load_col<NE>(reinterpret_cast<const cpx<float> *>(a.nlp[k]) + off, lane, ny, np);
            } else {
                sp::load_col<NE>(a.nl[k] + off, lane, ny, n);
                sp::load_col<NE>(a.nlp[k] + off, lane, ny, np);
            }
#pragma unroll
            for (int e = 0; e < NE; ++e) {
                rv[k][e] = ((a.a0 * u[e] + a.a1 * up[e]) + a.b0 * n[e]) + a.b1 * np[e];
                if (k == 1) nly[e] = n[e];
            }
        }
    } else if constexpr (mode == MODE_GIVEN) {
#pragma unroll
        for (int k = 0; k < 3; ++k) sp::load_col<NE>(a.rv[k] + off, lane, ny, rv[k]);
    } else {
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int e = 0; e < NE; ++e) rv[k][e] = {0.0, 0.0};
    }

    // ---- pressure right-hand side
    zd rp[1][NE];
    if (mode == MODE_STEP || (mode == MODE_GIVEN && a.rp == nullptr)) {
        zd d1[NE];
        sp::cheb_dy<NE>(rv[1], d1, lane, ny);
#pragma unroll
        for (int e = 0; e < NE; ++e) rp[0][e] = ((zd{0.0, 0.0} + ikmul(kx, rv[0][e])) + d1[e]) + ikmul(kz, rv[2][e]);
    } else if (mode == MODE_GIVEN) {
        sp::load_col<NE>(a.rp + off, lane, ny, rp[0]);
    } else if (mode == MODE_BPROBLEM) {
        // r_0 = NYS-1, r_j = 2 (NYS - ((j+1) mod 2)) for 1 <= j < NYS (eqsolve_pv.cpp:455-460)
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            const int j = sp::jof<NE>(lane, e);
            double v = 0.0;
            if (j == 0) v = (double)(ny - 1);
            else if (j < ny) v = (double)(2 * (ny - ((j + 1) % 2)));
            rp[0][e] = {v, 0.0};
        }
    } else {
#pragma unroll
        for (int e = 0; e < NE; ++e) rp[0][e] = {0.0, 0.0};
    }

    // ---- boundary data
    zd plo[1] = {{0.0, 0.0}}, pup[1] = {{0.0, 0.0}};
    if constexpr (mode == MODE_STEP) {
        // wall sums of N_y: sum_j N_j at y = +1 and sum_j (-1)^(j+1) N_j at y = -1
        zd sn0 = {0.0, 0.0};
#pragma unroll
        for (int e = 0; e < NE; ++e) sn0 = sn0 + nly[e];
        double t[4] = {sn0.x, sn0.y, p ? sn0.x : -sn0.x, p ? sn0.y : -sn0.y};
        wave_sum<4>(t);
        const zd B0 = a.bcw[(size_t)q * 2 + 0] + zd{t[0], t[1]}, B1 = a.bcw[(size_t)q * 2 + 1] + zd{t[2], t[3]};
        const zd P0 = a.bc_prev[(size_t)q * 2 + 0], P1 = a.bc_prev[(size_t)q * 2 + 1];
        plo[0] = a.b0 * B0 + a.b1 * P0;
        pup[0] = a.b0 * B1 + a.b1 * P1;
        if (lane == 0) {
            a.bc_prev[(size_t)q * 2 + 0] = B0;
            a.bc_prev[(size_t)q * 2 + 1] = B1;
            a.bc_ext[(size_t)q * 2 + 0] = plo[0];
            a.bc_ext[(size_t)q * 2 + 1] = pup[0];
        }
    } else if constexpr (mode == MODE_GIVEN) {
        plo[0] = a.pbc[(size_t)q * 2 + 0];
        pup[0] = a.pbc[(size_t)q * 2 + 1];
    } else if constexpr (mode == MODE_INFLUENCE) {
        plo[0] = {(double)(1 - a.influence_index), 0.0};
        pup[0] = {(double)a.influence_index, 0.0};
    }
    zd vlo[3], vup[3];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        if (mode == MODE_STEP || mode == MODE_GIVEN) {
            vlo[k] = a.vbc[(size_t)k * a.NQ * 2 + (size_t)q * 2 + 0];
            vup[k] = a.vbc[(size_t)k * a.NQ * 2 + (size_t)q * 2 + 1];
        } else {
            vlo[k] = {0.0, 0.0};
            vup[k] = {0.0, 0.0};
        }
    }
    const double wn_p = -(kx * kx + kz * kz + a.G_pre);
    const double wn_v = -(kx * kx + kz * kz + a.g_v);

    // ---- pressure, its y derivative, the three velocity components
    zd pc[1][NE];
    {
        const double pa[1] = {a.pa}, pb[1] = {a.pb};
        tau_solve<NE, 1>(rp, pc, adp, WL, WD, WR, lane, ny, wn_p, a.G_pre, pa, pb, zero, plo, pup);
    }
    const zd tail0 = sp::col_at<NE>(rv[1], ny), tail1 = sp::col_at<NE>(rv[1], ny - 1);
    zd vc[3][NE];
    {
        zd dy[NE], forcing[3][NE];
        sp::cheb_dy<NE>(pc[0], dy, lane, ny);
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int e = 0; e < NE; ++e) {
                const zd grad = k == 0 ? ikmul(kx, pc[0][e]) : (k == 1 ? dy[e] : ikmul(kz, pc[0][e]));
                forcing[k][e] = a.Re * (grad - rv[k][e]);
            }
        const double av[3] = {a.aV[0], a.aV[1], a.aV[2]}, bv[3] = {a.bV[0], a.bV[1], a.bV[2]};
        tau_solve<NE, 3>(forcing, vc, adv, WL, WD, WR, lane, ny, wn_v, a.g_v, av, bv, zero, vlo, vup);
    }

    // ---- per-column data of the step's corrections (influence entries, sigma denominators), read
    // before the wall divergence so that its reductions cover the latency
    const bool compact = a.compact && mode == MODE_STEP;
    zd idc[4] = {}, sgc[2] = {};
    if (compact && a.correct) {
        const zd *I0 = a.ID[0] + (size_t)q * 2, *I1 = a.ID[1] + (size_t)q * 2;
        idc[0] = I0[0];
        idc[1] = I0[1];
        idc[2] = I1[0];
        idc[3] = I1[1];
        if (a.sigma) {
            sgc[0] = a.SIG0[(size_t)q * 2];
            sgc[1] = a.SIG0[(size_t)q * 2 + 1];
        }
    }

    // ---- wall divergence (eqsolve_pv.cpp:397-406)
    zd div0 = {0.0, 0.0}, div1 = {0.0, 0.0};
    if (a.correct || mode == MODE_INFLUENCE) {
        zd tang = {0.0, 0.0}, normal = {0.0, 0.0};
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            const int j = sp::jof<NE>(lane, e);
            tang = tang + (ikmul(kx, vc[0][e]) + ikmul(kz, vc[2][e]));
            normal = normal + (double)(j * j) * vc[1][e];
        }
        // y = +1: tang + normal; y = -1: (-1)^j tang - (-1)^j normal
        const zd d0 = tang + normal, d1 = p ? normal - tang : tang - normal;
        double t[4] = {d0.x, d0.y, d1.x, d1.y};
        wave_sum<4>(t);
        div0 = {t[0], t[1]};
        div1 = {t[2], t[3]};
    }
    if constexpr (mode == MODE_INFLUENCE) {
        if (lane == 0) {
            a.div_out[(size_t)q * 2 + 0] = div0;
            a.div_out[(size_t)q * 2 + 1] = div1;
        }
        sp::store_col<NE>(a.P + off, lane, ny, pc[0]);
#pragma unroll
        for (int k = 0; k < 3; ++k) sp::store_col<NE>(a.V[k] + off, lane, ny, vc[k]);
        return;
    }

    // ---- influence-matrix and sigma corrections (eqsolve_pv.cpp:407-445)
    zd c0 = {0.0, 0.0}, c1 = {0.0, 0.0}, sig[2] = {{0.0, 0.0}, {0.0, 0.0}};
    if (a.correct) {
        const zd *I0 = a.ID[0] + (size_t)q * 2, *I1 = a.ID[1] + (size_t)q * 2;
        const zd i00 = compact ? idc[0] : I0[0], i01 = compact ? idc[1] : I0[1];
        const zd i10 = compact ? idc[2] : I1[0], i11 = compact ? idc[3] : I1[1];
        c0 = cmul(div1, i10) - cmul(div0, i11);
        c1 = cmul(div0, i01) - cmul(div1, i00);
        if (a.sigma) {
            zd ptop, vt0, vt1;
            zd s0, s1;
            if (compact) {
                ptop = sp::col_at<NE>(pc[0], ny) + (a.PIr[0][toff + ny] * c0 + a.PIr[1][toff + ny] * c1);
                vt0 = sp::col_at<NE>(vc[1], ny) + (a.VIv[0][toff + ny] * c0 + a.VIv[1][toff + ny] * c1);
                vt1 = sp::col_at<NE>(vc[1], ny - 1) + (a.VIv[0][toff + ny - 1] * c0 + a.VIv[1][toff + ny - 1] * c1);
                s0 = sgc[0];
                s1 = sgc[1];
            } else {
                ptop = sp::col_at<NE>(pc[0], ny) + (cmul(c0, a.PI[0][off + ny]) + cmul(c1, a.PI[1][off + ny]));
                vt0 = sp::col_at<NE>(vc[1], ny) + (cmul(c0, a.VI[0][1][off + ny]) + cmul(c1, a.VI[1][1][off + ny]));
                vt1 = sp::col_at<NE>(vc[1], ny - 1) +
                      (cmul(c0, a.VI[0][1][off + ny - 1]) + cmul(c1, a.VI[1][1][off + ny - 1]));
                s0 = a.SIG0[(size_t)q * 2 + 0];
                s1 = a.SIG0[(size_t)q * 2 + 1];
            }
            sig[0] = cdiv((-wn_v) * vt0 - a.Re * tail0, s0);
            sig[1] = cdiv((-wn_v) * vt1 + a.Re * ((2.0 * ny) * ptop - tail1), s1);
        }
    }
    // the B response enters with sigma_(1 - p) for P, u, w and sigma_p for v (q = (j+1) mod 2)
    const zd sp_own = p ? sig[1] : sig[0], sp_other = p ? sig[0] : sig[1];
    zd P[NE], V[3][NE];
    if (a.correct && a.sigma && compact) {
        // P += c0 P0 + c1 P1 + s PB;  v += c0 v0 + c1 v1 + s vB;  (u, w) += i (kx, kz) (c0 f0 + c1 f1 + s fB)
        double r0[NE], r1[NE], rb[NE];
        sp::load_rcol<NE>(a.PIr[0] + toff, lane, ny, r0);
        sp::load_rcol<NE>(a.PIr[1] + toff, lane, ny, r1);
        sp::load_rcol<NE>(a.PBr + toff, lane, ny, rb);
#pragma unroll
        for (int e = 0; e < NE; ++e) P[e] = pc[0][e] + ((r0[e] * c0 + r1[e] * c1) + rb[e] * sp_other);
        sp::load_rcol<NE>(a.VIv[0] + toff, lane, ny, r0);
        sp::load_rcol<NE>(a.VIv[1] + toff, lane, ny, r1);
        sp::load_rcol<NE>(a.VBv + toff, lane, ny, rb);
#pragma unroll
        for (int e = 0; e < NE; ++e) V[1][e] = vc[1][e] + ((r0[e] * c0 + r1[e] * c1) + rb[e] * sp_own);
        sp::load_rcol<NE>(a.VIf[0] + toff, lane, ny, r0);
        sp::load_rcol<NE>(a.VIf[1] + toff, lane, ny, r1);
        sp::load_rcol<NE>(a.VBf + toff, lane, ny, rb);
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            const zd phi = (r0[e] * c0 + r1[e] * c1) + rb[e] * sp_other;
            V[0][e] = vc[0][e] + ikmul(kx, phi);
            V[2][e] = vc[2][e] + ikmul(kz, phi);
        }
    } else {
        {
            zd i0[NE], i1[NE], bb[NE];
            if (a.correct) {
                sp::load_col<NE>(a.PI[0] + off, lane, ny, i0);
                sp::load_col<NE>(a.PI[1] + off, lane, ny, i1);
            }
            if (a.sigma) sp::load_col<NE>(a.PB + off, lane, ny, bb);
#pragma unroll
            for (int e = 0; e < NE; ++e) {
                zd v = pc[0][e];
                if (a.correct) v = v + (cmul(c0, i0[e]) + cmul(c1, i1[e]));
                if (a.sigma) v = v + cmul(sp_other, bb[e]);
                P[e] = v;
            }
        }
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            zd i0[NE], i1[NE], bb[NE];
            if (a.correct) {
                sp::load_col<NE>(a.VI[0][k] + off, lane, ny, i0);
                sp::load_col<NE>(a.VI[1][k] + off, lane, ny, i1);
            }
            if (a.sigma) sp::load_col<NE>(a.VB[k] + off, lane, ny, bb);
#pragma unroll
            for (int e = 0; e < NE; ++e) {
                zd v = vc[k][e];
                if (a.correct) v = v + (cmul(c0, i0[e]) + cmul(c1, i1[e]));
                if (a.sigma) v = v + cmul(k == 1 ? sp_own : sp_other, bb[e]);
                V[k][e] = v;
            }
        }
    }

    if (mode == MODE_STEP && zero) {
        // correct_zero_mode, then correct_flow on the mean mode
        const zd vmean = 0.5 * (vlo[1] + vup[1]);
        double u_unit[NE], w_unit[NE];
        sp::load_rcol<NE>(a.u_unit, lane, ny, u_unit);
        sp::load_rcol<NE>(a.w_unit, lane, ny, w_unit);
        double t[2] = {0.0, 0.0};
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            V[1][e] = {0.0, 0.0};
            const int j = sp::jof<NE>(lane, e);
            if (p == 0 && j <= ny) {
                t[0] -= V[0][e].x / (double)(j * j - 1);
                t[1] -= V[2][e].x / (double)(j * j - 1);
            }
        }
        if (lane == 0) V[1][0] = vmean;
        wave_sum<2>(t);
        double dpdx, dpdz;
        if (a.grad_pre == 1) {
            dpdx = a.grad_x;
            dpdz = a.grad_z;
        } else {
            dpdx = (a.qx_mean - t[0]) / a.q_unit[0];
            dpdz = (a.qz_mean - t[1]) / a.q_unit[1];
        }
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            V[0][e].x += dpdx * u_unit[e];
            V[2][e].x += dpdz * w_unit[e];
        }
        if (lane == 0) {
            a.flow[0] = dpdx;
            a.flow[1] = dpdz;
        }
    }
    if constexpr (mode == MODE_BPROBLEM) {
        // sigma denominators (eqsolve_pv.cpp:466-476)
        const zd vn = sp::col_at<NE>(V[1], ny), vm = sp::col_at<NE>(V[1], ny - 1), pn = sp::col_at<NE>(P, ny);
        if (lane == 0) {
            zd s0 = {a.Re + wn_v * vn.x, wn_v * vn.y};
            zd s1 = zd{a.Re + wn_v * vm.x, wn_v * vm.y} - (a.Re * (2.0 * ny)) * pn;
            if (zero) s1 = {1.0, 0.0};
            a.sig_out[(size_t)q * 2 + 0] = s0;
            a.sig_out[(size_t)q * 2 + 1] = s1;
        }
    }
    sp::store_col<NE>(a.P + off, lane, ny, P);
#pragma unroll
    for (int k = 0; k < 3; ++k) sp::store_col<NE>(a.V[k] + off, lane, ny, V[k]);
}

template <int NE> static void launch_solve_mode(const SolveArgs &a, hipStream_t st) {
    dim3 grid((a.NQ + 3) / 4), block(256);
    switch (a.mode) {
    case MODE_STEP:
        if (a.nl32) hipLaunchKernelGGL((k_solve<NE, MODE_STEP, true>), grid, block, 0, st, a);
        else hipLaunchKernelGGL((k_solve<NE, MODE_STEP>), grid, block, 0, st, a);
        break;
    case MODE_GIVEN: hipLaunchKernelGGL((k_solve<NE, MODE_GIVEN>), grid, block, 0, st, a); break;
    case MODE_INFLUENCE: hipLaunchKernelGGL((k_solve<NE, MODE_INFLUENCE>), grid, block, 0, st, a); break;
    case MODE_BPROBLEM: hipLaunchKernelGGL((k_solve<NE, MODE_BPROBLEM>), grid, block, 0, st, a); break;
    default: throw std::runtime_error("unknown solve mode");
    }
}
// E = column stride / 64 complex entries per lane pair: two entries of one chain per lane for
// columns up to 128, four up to 256
void launch_solve(const SolveArgs &a, int E, hipStream_t st) {
    if (E == 2) launch_solve_mode<2>(a, st);
    else if (E == 4) launch_solve_mode<4>(a, st);
    else throw std::runtime_error("unsupported column length");
}

// Compact tables: the influence and B responses come from real wall data and real operators,
// so P and v are real and u, w are i kx phi, i kz phi with one real phi when u and w share their
// wall conditions (eqsolve_pv.cpp:266-303, 452-477).
__global__ void k_compact_tables(SolveArgs a, double *PIr0, double *PIr1, double *VIv0, double *VIv1, double *VIf0,
                                 double *VIf1, double *PBr, double *VBv, double *VBf) {
    const size_t e = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (size_t)a.NQ * a.JS) return;
    const int q = (int)(e / a.JS);
    const int kxi = q / a.nzk, kzi = q - kxi * a.nzk;
    const double kx = a.WAVX[kxi], kz = a.WAVZ[kzi];
    auto phi = [&](const zd *u, const zd *w) {
        if (kx != 0.0) return u[e].y / kx;
        if (kz != 0.0) return w[e].y / kz;
        return 0.0;
    };
    PIr0[e] = a.PI[0][e].x;
    PIr1[e] = a.PI[1][e].x;
    VIv0[e] = a.VI[0][1][e].x;
    VIv1[e] = a.VI[1][1][e].x;
    VIf0[e] = phi(a.VI[0][0], a.VI[0][2]);
    VIf1[e] = phi(a.VI[1][0], a.VI[1][2]);
    PBr[e] = a.PB[e].x;
    VBv[e] = a.VB[1][e].x;
    VBf[e] = phi(a.VB[0], a.VB[2]);
}
void launch_compact_tables(const SolveArgs &a, double *PIr0, double *PIr1, double *VIv0, double *VIv1,
                           double *VIf0, double *VIf1, double *PBr, double *VBv, double *VBf, hipStream_t st) {
    const size_t n = (size_t)a.NQ * a.JS;
    hipLaunchKernelGGL(k_compact_tables, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, a, PIr0, PIr1, VIv0,
                       VIv1, VIf0, VIf1, PBr, VBv, VBf);
}

// build_wn_ad (eqsolve_pv.cpp:148-168): continued-fraction recurrence, one thread per pair.
__global__ void k_build_ad(const double *WAVX, const double *WAVZ, const double *WL, const double *WD,
                           const double *WR, double g, int ny, int JS, int NQ, int nzk, double *ad, double *wn) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= NQ) return;
    const int kxi = q / nzk, kzi = q - kxi * nzk;
    const double w = -(WAVX[kxi] * WAVX[kxi] + WAVZ[kzi] * WAVZ[kzi] + g);
    wn[q] = w;
    double *col = ad + (size_t)q * JS;
    col[0] = 0.0;
    col[1] = 0.0;
    col[ny] = 1.0 / (1.0 + WD[ny] * w);
    col[ny - 1] = 1.0 / (1.0 + WD[ny - 1] * w);
    for (int iy = ny - 2; iy >= 2; --iy) {
        const double fc = WR[iy] * WL[iy + 2];
        const double den = 1.0 + (WD[iy] - fc * w * col[iy + 2]) * w;
        col[iy] = 1.0 / den;
    }
}
void launch_build_ad(const double *WAVX, const double *WAVZ, const double *WL, const double *WD, const double *WR,
                     double g, int ny, int JS, int NQ, int nzk, double *ad, double *wn, hipStream_t st) {
    hipLaunchKernelGGL(k_build_ad, dim3((NQ + 127) / 128), dim3(128), 0, st, WAVX, WAVZ, WL, WD, WR, g, ny, JS, NQ,
                       nzk, ad, wn);
}

// build_influence_coefficients normalisation (eqsolve_pv.cpp:282-302).
__global__ void k_normalize_influence(zd *ID0, zd *ID1, int NQ, int root_q, double aV1) {
    const int q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q >= NQ) return;
    zd *a0 = ID0 + (size_t)q * 2, *a1 = ID1 + (size_t)q * 2;
    zd den = cmul(a0[0], a1[1]) - cmul(a0[1], a1[0]);
    const bool root = q == root_q;
    if (root) den = aV1 == 0.0 ? zd{1.0, 0.0} : cdiv(zd{0.25, 0.0}, a0[0]);
    for (int wall = 0; wall < 2; ++wall) {
        if (root) {
            a0[wall] = wall == 0 ? den : zd{0.0, 0.0};
            a1[wall] = wall == 0 ? den : zd{0.0, 0.0};
        } else {
            a0[wall] = cdiv(a0[wall], den);
            a1[wall] = cdiv(a1[wall], den);
        }
    }
}
void launch_normalize_influence(zd *ID0, zd *ID1, int NQ, int root_q, double aV1, hipStream_t st) {
    hipLaunchKernelGGL(k_normalize_influence, dim3((NQ + 127) / 128), dim3(128), 0, st, ID0, ID1, NQ, root_q, aV1);
}

// build_pressure_gradient_responses (eqsolve_pv.cpp:479-493) on the mean mode: one wavefront.
template <int NE>
__global__ void k_pg_response(const double *WAVX, const double *WAVZ, const double *WL_, const double *WD_,
                              const double *WR_, const double *ad_v, double g_v, double Re, double aV0, double bV0,
                              double aV2, double bV2, int ny, int JS, int root_q, double *u_unit, double *w_unit,
                              double *q_unit) {
    const int lane = lane_id(), p = sp::parity(lane);
    const size_t off = (size_t)root_q * JS;
    double WL[NE], WD[NE], WR[NE], adv[NE];
    sp::load_rcol<NE>(WL_, lane, ny, WL);
    sp::load_rcol<NE>(WD_, lane, ny, WD);
    sp::load_rcol<NE>(WR_, lane, ny, WR);
    sp::load_rcol<NE>(ad_v + off, lane, ny, adv);
    const double w = -(WAVX[0] * WAVX[0] + WAVZ[root_q] * WAVZ[root_q] + g_v);
    zd rhs[2][NE];
#pragma unroll
    for (int e = 0; e < NE; ++e) rhs[0][e] = rhs[1][e] = zd{0.0, 0.0};
    if (lane == 0) rhs[0][0] = rhs[1][0] = zd{Re, 0.0};
    const double av[2] = {aV0, aV2}, bv[2] = {bV0, bV2};
    const zd zw[2] = {{0.0, 0.0}, {0.0, 0.0}};
    zd sol[2][NE];
    tau_solve<NE, 2>(rhs, sol, adv, WL, WD, WR, lane, ny, w, g_v, av, bv, true, zw, zw);
    double qs[2] = {0.0, 0.0};
#pragma unroll
    for (int comp = 0; comp < 2; ++comp) {
        double *unit = comp ? w_unit : u_unit;
#pragma unroll
        for (int e = 0; e < NE; ++e) {
            const int j = sp::jof<NE>(lane, e);
            if (j <= ny) {
                unit[j] = sol[comp][e].x;
                if (p == 0) qs[comp] -= sol[comp][e].x / (double)(j * j - 1);
            }
        }
    }
    wave_sum<2>(qs);
    if (lane == 0) {
        q_unit[0] = qs[0];
        q_unit[1] = qs[1];
    }
}
void launch_pg_response(const double *WAVX, const double *WAVZ, const double *WL, const double *WD, const double *WR,
                        const double *ad_v, double g_v, double Re, const double *aV, const double *bV, int ny, int JS,
                        int root_q, double *u_unit, double *w_unit, double *q_unit, int E, hipStream_t st) {
    if (E == 2)
        hipLaunchKernelGGL((k_pg_response<2>), dim3(1), dim3(64), 0, st, WAVX, WAVZ, WL, WD, WR, ad_v, g_v, Re, aV[0],
                           bV[0], aV[2], bV[2], ny, JS, root_q, u_unit, w_unit, q_unit);
    else
        hipLaunchKernelGGL((k_pg_response<4>), dim3(1), dim3(64), 0, st, WAVX, WAVZ, WL, WD, WR, ad_v, g_v, Re, aV[0],
                           bV[0], aV[2], bV[2], ny, JS, root_q, u_unit, w_unit, q_unit);
}

} // namespace cdns
