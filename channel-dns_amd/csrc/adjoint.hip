// Discrete adjoint of the channel step on gfx950 (FP64).
//
// The reference has no adjoint.  These kernels transpose, operation by operation, the step the
// forward kernels compute (src/sim/simulation.cpp:142-171 as restated by the forward path), so
// that <J dX, L> = <dX, J^T L> holds to round-off with real inner products on (re, im) pairs:
//   k_adj_solve    flow correction^T, zero mode^T, solve_modes^T (influence and sigma rank-2
//                  updates, four transposed tau solves, eqsolve_pv.cpp:16-60, 311-450), the
//                  RHS and history transposes and the pressure wall condition^T
//                  (timestepper.cpp:26-43); one wavefront per (kx,kz) pair
//   k_adj_combine  output weights of the transposed synthesis and curl^T (vector_ops.cpp:112-151)
//   k_objective    terminal kinetic energy or mean wall shear and its gradient
// The transposed nonlinear term reuses the forward transform and product kernels between
// diagonal weights (k_adj_solve writes the input weights, k_adj_combine applies the output ones).
#include "internal.hpp"
#include "wave.hpp"

namespace cdns {

__device__ inline zd ikT(double k, zd v) { return {k * v.y, -k * v.x}; } // (-i k) v
__device__ inline zd conjmul(zd a, zd b) { return cmul(conjc(a), b); }   // conj(a) b
__device__ inline zd crecip(zd z) { return cdiv(zd{1.0, 0.0}, z); }
__device__ inline zd cdivr(zd z, double d) { return {z.x / d, z.y / d}; }

// D^T g for one lane-distributed column, where D is the Chebyshev derivative of
// vector_ops.cpp:17-25: (D^T g)_m = 2m sum_{j < m, m - j odd} c_j g_j, c_0 = 1/2.
template <int CH> __device__ inline void cheb_dyT(const zd (&g)[2][CH], zd (&o)[2][CH], int lane, int ny) {
    zd ge[CH], go[CH], PE[CH], PO[CH];
    double one[CH];
#pragma unroll
    for (int c = 0; c < CH; ++c) {
        const int j0 = jof<CH>(lane, c, 0);
        ge[c] = j0 <= ny ? (j0 == 0 ? 0.5 * g[0][c] : g[0][c]) : zd{0.0, 0.0};
        go[c] = j0 + 1 <= ny ? g[1][c] : zd{0.0, 0.0};
        one[c] = 1.0;
    }
    prefix_affine<CH, zd>(one, ge, PE, lane);
    prefix_affine<CH, zd>(one, go, PO, lane);
#pragma unroll
    for (int c = 0; c < CH; ++c) {
        const int j0 = jof<CH>(lane, c, 0), j1 = j0 + 1;
        const zd prev = chain_prev<CH>(PO, c, lane);
        o[0][c] = j0 <= ny ? (double)(2 * j0) * prev : zd{0.0, 0.0};
        o[1][c] = j1 <= ny ? (double)(2 * j1) * PE[c] : zd{0.0, 0.0};
    }
}

// Transpose of tau_solve (solve_tau_column, eqsolve_pv.cpp:16-60) for one column: l holds the
// adjoint of the solution (clobbered); lr receives the adjoint of the right-hand side, llo/lup
// those of the two wall values.
template <int CH>
__device__ void tau_solve_T(zd (&l)[2][CH], zd (&lr)[2][CH], zd &llo, zd &lup, const double (&ad)[2][CH],
                            const double (&WL)[2][CH], const double (&WD)[2][CH], const double (&WR)[2][CH],
                            int lane, int ny, double w, double g, double a, double b, bool zero_mode) {
    const bool normalize = g == 0.0 && a == 0.0 && zero_mode;
    double adn[2][CH], WLn[2][CH];
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            adn[p][c] = chain_next<CH>(ad[p], c, lane);
            WLn[p][c] = chain_next<CH>(WL[p], c, lane);
        }
    // closure accumulators of the forward: A(j) = f(j) - WL(j+2) w ad(j+2) A(j+2)
    double A[2][CH];
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        double f[CH], beta[CH];
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            const int j = jof<CH>(lane, c, p);
            const double fj = (a == 0.0 && b == 0.0) ? (double)((1 - p) * j * j + p) : a + b * j * j;
            f[c] = j <= ny ? fj : 0.0;
            beta[c] = (j <= ny - 2) ? -(WLn[p][c] * w * adn[p][c]) : 0.0;
        }
        suffix_affine<CH, double>(beta, f, A[p], lane);
    }
    const double A0 = col_at<CH>(A, 0), A1 = col_at<CH>(A, 1);
    // forward sweep^T: mu_j = l_j + a_{j+2} mu_{j+2}, a_j = -WL_j w ad_j
    zd mu[2][CH];
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        double an[CH];
        zd bl[CH];
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            const int j = jof<CH>(lane, c, p);
            an[c] = (j + 2 <= ny) ? -(WLn[p][c] * w * adn[p][c]) : 0.0;
            bl[c] = j <= ny ? l[p][c] : zd{0.0, 0.0};
        }
        suffix_affine<CH, zd>(an, bl, mu[p], lane);
    }
    const zd mu0 = col_at<CH>(mu, 0), mu1 = col_at<CH>(mu, 1);
    // out0 = (x0 - S0) / A0 (0 when normalised), out1 = (x1 - S1) / A1
    const zd lx0 = normalize ? zd{0.0, 0.0} : cdivr(mu0, A0);
    const zd lx1 = cdivr(mu1, A1);
    const zd lS[2] = {-1.0 * lx0, -1.0 * lx1};
    // adjoint of the swept rows x: ad mu (forward sweep) + A ad lS (closure sums)
    zd lx[2][CH];
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            const int j = jof<CH>(lane, c, p);
            if (j < 2) lx[p][c] = p ? lx1 : lx0;
            else if (j <= ny) lx[p][c] = ad[p][c] * mu[p][c] + (A[p][c] * ad[p][c]) * lS[p];
            else lx[p][c] = {0.0, 0.0};
        }
    // backward sweep^T: x_j = o_j + m_j x_{j+2}  ->  nu_j = lx_j + m_{j-2} nu_{j-2}
    zd nu[2][CH];
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        double m[CH], mp[CH];
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            const int j = jof<CH>(lane, c, p);
            m[c] = (j >= 2 && j <= ny - 2) ? -(WR[p][c] * w * adn[p][c]) : 0.0;
        }
#pragma unroll
        for (int c = 0; c < CH; ++c) mp[c] = chain_prev<CH>(m, c, lane);
        prefix_affine<CH, zd>(mp, lx[p], nu[p], lane);
    }
    // tau rows^T: o_j = WL_j R_{j-2} + WD_j R_j + WR_j R_{j+2}, 2 <= j <= ny
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        zd t1[CH], t3[CH];
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            const int j = jof<CH>(lane, c, p);
            const bool row = j >= 2 && j <= ny;
            t1[c] = row ? WL[p][c] * nu[p][c] : zd{0.0, 0.0};
            t3[c] = row ? WR[p][c] * nu[p][c] : zd{0.0, 0.0};
        }
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            const int j = jof<CH>(lane, c, p);
            const bool row = j >= 2 && j <= ny;
            zd v = chain_next<CH>(t1, c, lane) + chain_prev<CH>(t3, c, lane);
            if (row) v = v + WD[p][c] * nu[p][c];
            lr[p][c] = j <= ny ? v : zd{0.0, 0.0};
        }
    }
    const zd nu0 = col_at<CH>(nu, 0), nu1 = col_at<CH>(nu, 1);
    llo = 0.5 * (nu0 + nu1);
    lup = 0.5 * (nu0 - nu1);
    if (normalize) {
        // R_0 = sum_{j even >= 2} r_j / (j^2 - 1) + (lower + upper) / (2b)
        const zd lcorr = col_at<CH>(lr, 0);
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            const int j = jof<CH>(lane, c, 0);
            if (j == 0) lr[0][c] = {0.0, 0.0};
            else if (j >= 2 && j <= ny) lr[0][c] = lr[0][c] + cdivr(lcorr, (double)(j * j - 1));
        }
        if (b != 0.0) {
            llo = llo + cdivr(lcorr, 2.0 * b);
            lup = lup + cdivr(lcorr, 2.0 * b);
        }
    }
}

// Weight of the transposed synthesis's input (Din) and output (Dout = 1 / Din, up to the
// collision rule) for column kx, Chebyshev index j: (1 / w_kx) (s_j / bw_j), with w_kx = 2 for
// the interior kx of the half spectrum, s_j the DCT-I analysis weights (1/(2 NY2) at j = 0,
// 1/NY2 else) and bw_j = 1/2 at j = NY2 (transforms.cpp:641-741).
__device__ inline double din_weight(int kxi, int j, int NX2, int NY2) {
    const double wx = (kxi == 0 || 2 * kxi == NX2) ? 1.0 : 2.0;
    const double s = j == 0 ? 1.0 / (2.0 * NY2) : 1.0 / NY2;
    const double bw = j == NY2 ? 0.5 : 1.0;
    return s / (bw * wx);
}

template <int CH>
__global__ void __launch_bounds__(256) k_adj_solve(AdjArgs a) {
    const int lane = lane_id();
    const int q = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (q >= a.NQ) return;
    const int ny = a.ny;
    const int kxi = q / a.nzk, kzi = q - kxi * a.nzk;
    const bool zero = kxi == 0 && 2 * kzi == a.nzk - 1;
    const double kx = a.WAVX[kxi], kz = a.WAVZ[kzi];
    const size_t off = (size_t)q * a.JS;
    double WL[2][CH], WD[2][CH], WR[2][CH], adp[2][CH], adv[2][CH];
    load_rcol<CH>(a.WL, lane, ny, WL);
    load_rcol<CH>(a.WD, lane, ny, WD);
    load_rcol<CH>(a.WR, lane, ny, WR);
    load_rcol<CH>(a.ad_p + off, lane, ny, adp);
    load_rcol<CH>(a.ad_v + off, lane, ny, adv);

    // adjoint of the new velocity
    zd lv[3][2][CH];
#pragma unroll
    for (int k = 0; k < 3; ++k) load_col<CH>(a.lu[k] + off, lane, ny, lv[k]);
    // correct_flow^T (simulation.cpp:111-140): with a constant flux, dpdx = (Q0 - Q(u*_x)) / Q(u_G),
    // Q(f) = -sum_{j even} Re f_j / (j^2 - 1), added along the unit-gradient response
    if (zero && a.grad_pre != 1) {
        double uu[2][CH], ww[2][CH];
        load_rcol<CH>(a.u_unit, lane, ny, uu);
        load_rcol<CH>(a.w_unit, lane, ny, ww);
        double gx = 0.0, gz = 0.0;
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                gx += lv[0][p][c].x * uu[p][c];
                gz += lv[2][p][c].x * ww[p][c];
            }
        gx = wave_sum(gx);
        gz = wave_sum(gz);
        const double qx = a.q_unit[0], qz = a.q_unit[1];
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            const int j = jof<CH>(lane, c, 0);
            if (j <= ny) {
                const double dq = 1.0 / (double)(j * j - 1);
                lv[0][0][c].x += dq / qx * gx;
                lv[2][0][c].x += dq / qz * gz;
            }
        }
    }
    // correct_zero_mode^T (timestepper.cpp:60-68): the overwritten v of the mean mode
    if (zero) {
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int c = 0; c < CH; ++c) lv[1][p][c] = {0.0, 0.0};
    }

    // ---- solve_modes^T: outputs V_k = vc_k + c0 VI0_k + c1 VI1_k + sigma VB_k
    zd lc0 = {0.0, 0.0}, lc1 = {0.0, 0.0}, ls[2] = {{0.0, 0.0}, {0.0, 0.0}};
    zd vi0ny = {0.0, 0.0}, vi0ny1 = {0.0, 0.0}, vi1ny = {0.0, 0.0}, vi1ny1 = {0.0, 0.0};
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        zd i0[2][CH], i1[2][CH], bb[2][CH];
        load_col<CH>(a.VI[0][k] + off, lane, ny, i0);
        load_col<CH>(a.VI[1][k] + off, lane, ny, i1);
        load_col<CH>(a.VB[k] + off, lane, ny, bb);
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                lc0 = lc0 + conjmul(i0[p][c], lv[k][p][c]);
                lc1 = lc1 + conjmul(i1[p][c], lv[k][p][c]);
                // sigma index (j+1) mod 2 = 1 - p, and p for the wall-normal component
                const int s = k == 1 ? p : 1 - p;
                ls[s] = ls[s] + conjmul(bb[p][c], lv[k][p][c]);
            }
        if (k == 1) {
            vi0ny = col_at<CH>(i0, ny);
            vi0ny1 = col_at<CH>(i0, ny - 1);
            vi1ny = col_at<CH>(i1, ny);
            vi1ny1 = col_at<CH>(i1, ny - 1);
        }
    }
    lc0 = wave_sum(lc0);
    lc1 = wave_sum(lc1);
    ls[0] = wave_sum(ls[0]);
    ls[1] = wave_sum(ls[1]);
    const double Re = a.Re;
    const double wv = -(kx * kx + kz * kz + a.g_v), wp = -(kx * kx + kz * kz + a.G_pre);
    // sigma (eqsolve_pv.cpp:413-428): s0 = (-w VT(ny) - Re r1(ny)) / SIG0_0,
    // s1 = (-w VT(ny-1) + Re (2 ny ptop - r1(ny-1))) / SIG0_1
    const zd ls0 = conjmul(crecip(a.SIG0[(size_t)q * 2 + 0]), ls[0]);
    const zd ls1 = conjmul(crecip(a.SIG0[(size_t)q * 2 + 1]), ls[1]);
    const zd lVTn = (-wv) * ls0, lVTn1 = (-wv) * ls1, lptop = (2.0 * ny * Re) * ls1;
    const zd pi0ny = a.PI[0][off + ny], pi1ny = a.PI[1][off + ny];
    lc0 = lc0 + conjmul(vi0ny, lVTn) + conjmul(vi0ny1, lVTn1) + conjmul(pi0ny, lptop);
    lc1 = lc1 + conjmul(vi1ny, lVTn) + conjmul(vi1ny1, lVTn1) + conjmul(pi1ny, lptop);
    // influence coefficients c0 = div1 I1_0 - div0 I1_1, c1 = div0 I0_1 - div1 I0_0
    const zd *I0 = a.ID[0] + (size_t)q * 2, *I1 = a.ID[1] + (size_t)q * 2;
    const zd ldiv1 = conjmul(I1[0], lc0) - conjmul(I0[0], lc1);
    const zd ldiv0 = conjmul(I0[1], lc1) - conjmul(I1[1], lc0);
    zd lr[3][2][CH], lpc[2][CH], ldy[2][CH];
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            const int j = jof<CH>(lane, c, p);
            // wall divergence (eqsolve_pv.cpp:397-406)
            if (j <= ny) {
                const double sgn = p ? -1.0 : 1.0;
                const zd lt = ldiv0 + sgn * ldiv1, ln = ldiv0 - sgn * ldiv1;
                lv[0][p][c] = lv[0][p][c] + ikT(kx, lt);
                lv[2][p][c] = lv[2][p][c] + ikT(kz, lt);
                lv[1][p][c] = lv[1][p][c] + (double)(j * j) * ln;
            }
            if (j == ny) lv[1][p][c] = lv[1][p][c] + lVTn;
            if (j == ny - 1) lv[1][p][c] = lv[1][p][c] + lVTn1;
            lpc[p][c] = j == ny ? lptop : zd{0.0, 0.0};
            ldy[p][c] = {0.0, 0.0};
#pragma unroll
            for (int k = 0; k < 3; ++k) lr[k][p][c] = {0.0, 0.0};
            if (j == ny) lr[1][p][c] = (-Re) * ls0;
            if (j == ny - 1) lr[1][p][c] = (-Re) * ls1;
        }
    // velocity tau solves^T; forcing_k = Re (grad_k P - r_k)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        zd lf[2][CH], dlo, dup;
        tau_solve_T<CH>(lv[k], lf, dlo, dup, adv, WL, WD, WR, lane, ny, wv, a.g_v, a.aV[k], a.bV[k], zero);
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                const zd lg = Re * lf[p][c];
                lr[k][p][c] = lr[k][p][c] - lg;
                if (k == 0) lpc[p][c] = lpc[p][c] + ikT(kx, lg);
                else if (k == 1) ldy[p][c] = ldy[p][c] + lg;
                else lpc[p][c] = lpc[p][c] + ikT(kz, lg);
            }
    }
    {
        zd t[2][CH];
        cheb_dyT<CH>(ldy, t, lane, ny);
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int c = 0; c < CH; ++c) lpc[p][c] = lpc[p][c] + t[p][c];
    }
    // pressure tau solve^T and the fused divergence RHS^T: rp = i kx r0 + D r1 + i kz r2
    zd lrp[2][CH], lplo, lpup;
    tau_solve_T<CH>(lpc, lrp, lplo, lpup, adp, WL, WD, WR, lane, ny, wp, a.G_pre, a.pa, a.pb, zero);
    {
        zd t[2][CH];
        cheb_dyT<CH>(lrp, t, lane, ny);
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                lr[0][p][c] = lr[0][p][c] + ikT(kx, lrp[p][c]);
                lr[1][p][c] = lr[1][p][c] + t[p][c];
                lr[2][p][c] = lr[2][p][c] + ikT(kz, lrp[p][c]);
            }
    }

    // ---- RHS r = a0 u + a1 up + b0 N + b1 Np, bext = b0 B + b1 Bp, and the history moves
    const zd lbp0 = a.lbp[(size_t)q * 2 + 0], lbp1 = a.lbp[(size_t)q * 2 + 1];
    const zd lB0 = a.b0 * lplo + lbp0, lB1 = a.b0 * lpup + lbp1;
    if (lane == 0) {
        a.lbp[(size_t)q * 2 + 0] = a.b1 * lplo;
        a.lbp[(size_t)q * 2 + 1] = a.b1 * lpup;
    }
    const double Re_inv = 1.0 / Re;
    const int NX2 = a.NX2, NY2 = a.NY2;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        zd up_in[2][CH], np_in[2][CH], lnew[2][CH], lN[2][CH], lw[2][CH];
        load_col<CH>(a.lup[k] + off, lane, ny, up_in);
        load_col<CH>(a.lnp[k] + off, lane, ny, np_in);
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                lnew[p][c] = a.a0 * lr[k][p][c] + up_in[p][c];
                lN[p][c] = a.b0 * lr[k][p][c] + np_in[p][c];
                up_in[p][c] = a.a1 * lr[k][p][c];
                np_in[p][c] = a.b1 * lr[k][p][c];
                // pressure wall condition^T: B0 = sum_j v_j, B1 = sum_j (-1)^(j+1) v_j,
                // v = i kx w_z / Re - i kz w_x / Re + N_y
                const zd lwall = p ? lB0 + lB1 : lB0 - lB1;
                lw[p][c] = {0.0, 0.0};
                if (k == 1) lN[p][c] = lN[p][c] + lwall;
                if (k == 0) lw[p][c] = ikmul(kz * Re_inv, lwall);
                if (k == 2) lw[p][c] = ikT(kx * Re_inv, lwall);
            }
        store_col<CH>(a.lu[k] + off, lane, ny, lnew);
        store_col<CH>(a.lup[k] + off, lane, ny, up_in);
        store_col<CH>(a.lnp[k] + off, lane, ny, np_in);
        store_col<CH>(a.lw[k] + off, lane, ny, lw);
        if (a.lf[k]) {
            zd f[2][CH];
            load_col<CH>(a.lf[k] + off, lane, ny, f);
#pragma unroll
            for (int p = 0; p < 2; ++p)
#pragma unroll
                for (int c = 0; c < CH; ++c) f[p][c] = f[p][c] + lN[p][c];
            store_col<CH>(a.lf[k] + off, lane, ny, f);
        }
        // input weights of the transposed nonlinear term
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int c = 0; c < CH; ++c) lN[p][c] = din_weight(kxi, jof<CH>(lane, c, p), NX2, NY2) * lN[p][c];
        store_col<CH>(a.lNs[k] + off, lane, ny, lN);
    }
}

// lu = lu + Dout lUs + curl^T(Dout lWs + lw): the output weights of the transposed synthesis
// and the curl transpose (w_x = D w - i kz v, w_y = i kz u - i kx w, w_z = i kx v - D u).
template <int CH>
__global__ void __launch_bounds__(256) k_adj_combine(AdjArgs a) {
    const int lane = lane_id();
    const int q = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (q >= a.NQ) return;
    const int ny = a.ny;
    const int kxi = q / a.nzk, kzi = q - kxi * a.nzk;
    const double kx = a.WAVX[kxi], kz = a.WAVZ[kzi];
    const size_t off = (size_t)q * a.JS;
    // the synthesis never reads kz = +NZS when its slot collides with kz = -NZS
    const bool dropped = a.collide && kzi == a.nzk - 1;
    zd lw[3][2][CH], lU[3][2][CH];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        zd ws[2][CH], w0[2][CH];
        load_col<CH>(a.lWs[k] + off, lane, ny, ws);
        load_col<CH>(a.lw[k] + off, lane, ny, w0);
        load_col<CH>(a.lUs[k] + off, lane, ny, lU[k]);
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                const int j = jof<CH>(lane, c, p);
                const double d = dropped ? 0.0 : 1.0 / din_weight(kxi, j, a.NX2, a.NY2);
                lw[k][p][c] = d * ws[p][c] + w0[p][c];
                lU[k][p][c] = d * lU[k][p][c];
            }
    }
    zd dx[2][CH], dz[2][CH];
    cheb_dyT<CH>(lw[0], dx, lane, ny);
    cheb_dyT<CH>(lw[2], dz, lane, ny);
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        zd u[2][CH];
        load_col<CH>(a.lu[k] + off, lane, ny, u);
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                zd ct;
                if (k == 0) ct = ikT(kz, lw[1][p][c]) - dz[p][c];
                else if (k == 1) ct = ikmul(kz, lw[0][p][c]) - ikmul(kx, lw[2][p][c]);
                else ct = dx[p][c] + ikmul(kx, lw[1][p][c]);
                u[p][c] = u[p][c] + lU[k][p][c] + ct;
            }
        store_col<CH>(a.lu[k] + off, lane, ny, u);
    }
}

// Collision rule of the synthesis (NZ2 <= 2 NZS): the slot of kz = -NZS also serves kz = +NZS
// in the analysis, so its transpose folds the +NZS input into -NZS.
__global__ void k_adj_fold(AdjArgs a) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    const int nj = a.ny + 1;
    if (e >= a.nkx * nj) return;
    const int kxi = e / nj, j = e - kxi * nj;
    const size_t lo = (size_t)kxi * a.nzk * a.JS + j, hi = lo + (size_t)(a.nzk - 1) * a.JS;
    for (int k = 0; k < 3; ++k) {
        a.lNs[k][lo] = a.lNs[k][lo] + a.lNs[k][hi];
        a.lNs[k][hi] = {0.0, 0.0};
    }
}

// Terminal objective per column: kinetic energy 1/2 int |u|^2 dV (which = 0) with
// int_{-1}^{1} T_m T_n dy = 1/(1 - (m+n)^2) + 1/(1 - (m-n)^2) for m + n even, or the mean wall
// shear -sum_{j even} j^2 Re u_x(0,0,j) (which = 1).  Writes the gradient planes and the
// column's share of J.
template <int CH>
__global__ void __launch_bounds__(256) k_objective(AdjArgs a) {
    const int lane = lane_id();
    const int q = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (q >= a.NQ) return;
    const int ny = a.ny;
    const int kxi = q / a.nzk, kzi = q - kxi * a.nzk;
    const bool zero = kxi == 0 && 2 * kzi == a.nzk - 1;
    const size_t off = (size_t)q * a.JS;
    double J = 0.0;
    if (a.which == 1) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            zd g[2][CH], u[2][CH];
            load_col<CH>(a.lu[k] + off, lane, ny, u);
#pragma unroll
            for (int p = 0; p < 2; ++p)
#pragma unroll
                for (int c = 0; c < CH; ++c) {
                    const int j = jof<CH>(lane, c, p);
                    const bool on = zero && k == 0 && p == 0 && j <= ny;
                    g[p][c] = on ? zd{-(double)(j * j), 0.0} : zd{0.0, 0.0};
                    if (on) J -= (double)(j * j) * u[p][c].x;
                }
            store_col<CH>(a.grad[k] + off, lane, ny, g);
        }
    } else {
        const double ck = (kxi == 0 ? 1.0 : 2.0) * a.XL * a.ZL;
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            zd u[2][CH], acc[2][CH];
            load_col<CH>(a.lu[k] + off, lane, ny, u);
#pragma unroll
            for (int p = 0; p < 2; ++p)
#pragma unroll
                for (int c = 0; c < CH; ++c) acc[p][c] = {0.0, 0.0};
            for (int m = 0; m <= ny; ++m) {
                const zd um = col_at<CH>(u, m);
#pragma unroll
                for (int c = 0; c < CH; ++c) {
                    const int n = jof<CH>(lane, c, m & 1);
                    if (n <= ny) {
                        const double s = (double)(n + m), d = (double)(n - m);
                        acc[m & 1][c] = acc[m & 1][c] + (1.0 / (1.0 - s * s) + 1.0 / (1.0 - d * d)) * um;
                    }
                }
            }
#pragma unroll
            for (int p = 0; p < 2; ++p)
#pragma unroll
                for (int c = 0; c < CH; ++c) {
                    J += 0.5 * ck * (u[p][c].x * acc[p][c].x + u[p][c].y * acc[p][c].y);
                    acc[p][c] = ck * acc[p][c];
                }
            store_col<CH>(a.grad[k] + off, lane, ny, acc);
        }
    }
    J = wave_sum(J);
    if (lane == 0) a.Jpart[q] = J;
}

void launch_adj_solve(const AdjArgs &a, int E, hipStream_t st) {
    dim3 grid((a.NQ + 3) / 4), block(256);
    if (E == 2) hipLaunchKernelGGL((k_adj_solve<1>), grid, block, 0, st, a);
    else if (E == 4) hipLaunchKernelGGL((k_adj_solve<2>), grid, block, 0, st, a);
    else throw std::runtime_error("unsupported column length");
}
void launch_adj_combine(const AdjArgs &a, int E, hipStream_t st) {
    dim3 grid((a.NQ + 3) / 4), block(256);
    if (E == 2) hipLaunchKernelGGL((k_adj_combine<1>), grid, block, 0, st, a);
    else if (E == 4) hipLaunchKernelGGL((k_adj_combine<2>), grid, block, 0, st, a);
    else throw std::runtime_error("unsupported column length");
}
void launch_adj_fold(const AdjArgs &a, hipStream_t st) {
    const int n = a.nkx * (a.ny + 1);
    hipLaunchKernelGGL(k_adj_fold, dim3((n + 255) / 256), dim3(256), 0, st, a);
}
void launch_objective(const AdjArgs &a, int E, hipStream_t st) {
    dim3 grid((a.NQ + 3) / 4), block(256);
    if (E == 2) hipLaunchKernelGGL((k_objective<1>), grid, block, 0, st, a);
    else if (E == 4) hipLaunchKernelGGL((k_objective<2>), grid, block, 0, st, a);
    else throw std::runtime_error("unsupported column length");
}

} // namespace cdns
