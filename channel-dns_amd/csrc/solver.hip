// Per-(kx,kz) Chebyshev-tau solver of the channel-dns step on gfx950: one 64-lane wavefront
// per wavenumber pair.
//
// The reference solves each column with sequential y recurrences over 8-lane kx tiles
// (EqSolvePV::solve_modes / solve_tau_tile, src/solver/eqsolve_pv.cpp:64-128, 311-450).
// The recurrences are linear in the data, so each becomes an affine scan along the even and
// odd Chebyshev chains (6 shuffle steps), and the influence-matrix and sigma corrections are
// lane-parallel over y with wavefront reductions for the wall divergence.
#include "internal.hpp"
#include "wave.hpp"

namespace cdns {

__device__ inline zd cdivr(zd z, double d) { return {z.x / d, z.y / d}; }

// solve_tau_column (eqsolve_pv.cpp:16-60) for one column held by the wavefront.
template <int CH>
__device__ void tau_solve(const zd (&r)[2][CH], zd (&out)[2][CH], const double (&ad)[2][CH],
                          const double (&WL)[2][CH], const double (&WD)[2][CH], const double (&WR)[2][CH],
                          int lane, int ny, double w, double g, double a, double b, bool zero_mode, zd lower,
                          zd upper) {
    const bool normalize = g == 0.0 && a == 0.0 && zero_mode;
    zd R[2][CH];
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int c = 0; c < CH; ++c) R[p][c] = r[p][c];
    if (normalize) {
        // compatibility correction of the pure-Neumann mean pressure (eqsolve_pv.cpp:19-27)
        zd corr = {0.0, 0.0};
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            const int j = jof<CH>(lane, c, 0);
            if (j >= 2 && j <= ny) corr = corr + cdivr(r[0][c], (double)(j * j - 1));
        }
        corr = wave_sum(corr);
        if (b != 0.0) corr = corr + cdivr(lower + upper, 2.0 * b);
        if (lane == 0) R[0][0] = corr;
    }
    // tau right-hand side and the two closure rows (eqsolve_pv.cpp:28-33)
    zd o[2][CH];
    double adn[2][CH], WLn[2][CH];
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            const int j = jof<CH>(lane, c, p);
            const zd rm = chain_prev<CH>(R[p], c, lane), rn = chain_next<CH>(R[p], c, lane);
            adn[p][c] = chain_next<CH>(ad[p], c, lane);
            WLn[p][c] = chain_next<CH>(WL[p], c, lane);
            if (j == 0) o[p][c] = cdivr(lower + upper, 2.0);
            else if (j == 1) o[p][c] = cdivr(lower - upper, 2.0);
            else if (j <= ny) o[p][c] = (WL[p][c] * rm + WD[p][c] * R[p][c]) + WR[p][c] * rn;
            else o[p][c] = {0.0, 0.0};
        }
    // backward sweep out(j) -= WR(j) w ad(j+2) out(j+2), 2 <= j <= ny-2 (eqsolve_pv.cpp:34-35)
    zd x[2][CH];
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        double m[CH];
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            const int j = jof<CH>(lane, c, p);
            m[c] = (j >= 2 && j <= ny - 2) ? -(WR[p][c] * w * adn[p][c]) : 0.0;
        }
        suffix_affine<CH, zd>(m, o[p], x[p], lane);
    }
    // closure accumulators ad01 (eqsolve_pv.cpp:37-50): A(j) = f(j) - WL(j+2) w ad(j+2) A(j+2)
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
    zd S[2];
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        zd s = {0.0, 0.0};
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            const int j = jof<CH>(lane, c, p);
            if (j >= 2 && j <= ny) s = s + (A[p][c] * ad[p][c]) * x[p][c];
        }
        S[p] = wave_sum(s);
    }
    const double A0 = col_at<CH>(A, 0), A1 = col_at<CH>(A, 1);
    const zd x0 = col_at<CH>(x, 0), x1 = col_at<CH>(x, 1);
    const zd out0 = normalize ? zd{0.0, 0.0} : cdivr(x0 - S[0], A0);
    const zd out1 = cdivr(x1 - S[1], A1);
    // forward sweep out(j) = (out(j) - WL(j) w out(j-2)) ad(j) (eqsolve_pv.cpp:58-59)
#pragma unroll
    for (int p = 0; p < 2; ++p) {
        double aa[CH];
        zd bb[CH];
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            const int j = jof<CH>(lane, c, p);
            if (j < 2) {
                aa[c] = 0.0;
                bb[c] = p ? out1 : out0;
            } else if (j <= ny) {
                aa[c] = -(WL[p][c] * w * ad[p][c]);
                bb[c] = ad[p][c] * x[p][c];
            } else {
                aa[c] = 0.0;
                bb[c] = {0.0, 0.0};
            }
        }
        prefix_affine<CH, zd>(aa, bb, out[p], lane);
    }
}

// Fused coupled solve of one (kx,kz) pair: EqSolvePV::solve_modes (eqsolve_pv.cpp:311-450),
// with the step's RHS assembly (simulation.cpp:204-221), the nonlinear pressure wall condition
// (timestepper.cpp:55-72), correct_zero_mode (timestepper.cpp:89-97) and correct_flow
// (simulation.cpp:173-202) fused in MODE_STEP; the setup problems of eqsolve_pv.cpp:266-303
// and 452-477 in MODE_INFLUENCE / MODE_BPROBLEM.
template <int CH>
__global__ void __launch_bounds__(256) k_solve(SolveArgs a) {
    const int lane = lane_id();
    const int q = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    if (q >= a.NQ) return;
    const int ny = a.ny, mode = a.mode;
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

    // ---- velocity right-hand side
    zd rv[3][2][CH];
    zd nly[2][CH];
    if (mode == MODE_STEP) {
#pragma unroll
        for (int k = 0; k < 3; ++k) {
            zd u[2][CH], up[2][CH], n[2][CH], np[2][CH];
            load_col<CH>(a.u[k] + off, lane, ny, u);
            load_col<CH>(a.up[k] + off, lane, ny, up);
            load_col<CH>(a.nl[k] + off, lane, ny, n);
            load_col<CH>(a.nlp[k] + off, lane, ny, np);
#pragma unroll
            for (int p = 0; p < 2; ++p)
#pragma unroll
                for (int c = 0; c < CH; ++c) {
                    rv[k][p][c] = ((a.a0 * u[p][c] + a.a1 * up[p][c]) + a.b0 * n[p][c]) + a.b1 * np[p][c];
                    if (k == 1) nly[p][c] = n[p][c];
                }
        }
    } else if (mode == MODE_GIVEN) {
#pragma unroll
        for (int k = 0; k < 3; ++k) load_col<CH>(a.rv[k] + off, lane, ny, rv[k]);
    } else {
#pragma unroll
        for (int k = 0; k < 3; ++k)
#pragma unroll
            for (int p = 0; p < 2; ++p)
#pragma unroll
                for (int c = 0; c < CH; ++c) rv[k][p][c] = {0.0, 0.0};
    }

    // ---- pressure right-hand side
    zd rp[2][CH];
    if (mode == MODE_STEP || (mode == MODE_GIVEN && a.rp == nullptr)) {
        zd d1[2][CH];
        cheb_dy<CH>(rv[1], d1, lane, ny);
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int c = 0; c < CH; ++c)
                rp[p][c] = ((zd{0.0, 0.0} + ikmul(kx, rv[0][p][c])) + d1[p][c]) + ikmul(kz, rv[2][p][c]);
    } else if (mode == MODE_GIVEN) {
        load_col<CH>(a.rp + off, lane, ny, rp);
    } else if (mode == MODE_BPROBLEM) {
        // r_0 = NYS-1, r_j = 2 (NYS - ((j+1) mod 2)) for 1 <= j < NYS (eqsolve_pv.cpp:455-460)
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                const int j = jof<CH>(lane, c, p);
                double v = 0.0;
                if (j == 0) v = (double)(ny - 1);
                else if (j < ny) v = (double)(2 * (ny - ((j + 1) % 2)));
                rp[p][c] = {v, 0.0};
            }
    } else {
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int c = 0; c < CH; ++c) rp[p][c] = {0.0, 0.0};
    }

    // ---- boundary data
    zd plo = {0.0, 0.0}, pup = {0.0, 0.0};
    if (mode == MODE_STEP) {
        zd sn0 = {0.0, 0.0}, sn1 = {0.0, 0.0};
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                const int j = jof<CH>(lane, c, p);
                if (j <= ny) {
                    sn0 = sn0 + nly[p][c];
                    sn1 = p == 0 ? sn1 - nly[p][c] : sn1 + nly[p][c];
                }
            }
        sn0 = wave_sum(sn0);
        sn1 = wave_sum(sn1);
        const zd B0 = a.bcw[(size_t)q * 2 + 0] + sn0, B1 = a.bcw[(size_t)q * 2 + 1] + sn1;
        const zd P0 = a.bc_prev[(size_t)q * 2 + 0], P1 = a.bc_prev[(size_t)q * 2 + 1];
        plo = a.b0 * B0 + a.b1 * P0;
        pup = a.b0 * B1 + a.b1 * P1;
        if (lane == 0) {
            a.bc_prev[(size_t)q * 2 + 0] = B0;
            a.bc_prev[(size_t)q * 2 + 1] = B1;
            a.bc_ext[(size_t)q * 2 + 0] = plo;
            a.bc_ext[(size_t)q * 2 + 1] = pup;
        }
    } else if (mode == MODE_GIVEN) {
        plo = a.pbc[(size_t)q * 2 + 0];
        pup = a.pbc[(size_t)q * 2 + 1];
    } else if (mode == MODE_INFLUENCE) {
        plo = {(double)(1 - a.influence_index), 0.0};
        pup = {(double)a.influence_index, 0.0};
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
    zd pc[2][CH], dy[2][CH];
    tau_solve<CH>(rp, pc, adp, WL, WD, WR, lane, ny, wn_p, a.G_pre, a.pa, a.pb, zero, plo, pup);
    cheb_dy<CH>(pc, dy, lane, ny);
    const zd tail0 = col_at<CH>(rv[1], ny), tail1 = col_at<CH>(rv[1], ny - 1);
    zd vc[3][2][CH];
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        zd forcing[2][CH];
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                const zd grad = k == 0 ? ikmul(kx, pc[p][c]) : (k == 1 ? dy[p][c] : ikmul(kz, pc[p][c]));
                forcing[p][c] = a.Re * (grad - rv[k][p][c]);
            }
        tau_solve<CH>(forcing, vc[k], adv, WL, WD, WR, lane, ny, wn_v, a.g_v, a.aV[k], a.bV[k], zero, vlo[k], vup[k]);
    }

    // ---- wall divergence (eqsolve_pv.cpp:397-406)
    zd div0 = {0.0, 0.0}, div1 = {0.0, 0.0};
    if (a.correct || mode == MODE_INFLUENCE) {
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                const int j = jof<CH>(lane, c, p);
                if (j <= ny) {
                    const zd tang = ikmul(kx, vc[0][p][c]) + ikmul(kz, vc[2][p][c]);
                    const zd normal = (double)(j * j) * vc[1][p][c];
                    div0 = div0 + (tang + normal);
                    div1 = div1 + (p == 0 ? tang - normal : normal - tang);
                }
            }
        div0 = wave_sum(div0);
        div1 = wave_sum(div1);
    }
    if (mode == MODE_INFLUENCE) {
        if (lane == 0) {
            a.div_out[(size_t)q * 2 + 0] = div0;
            a.div_out[(size_t)q * 2 + 1] = div1;
        }
        store_col<CH>(a.P + off, lane, ny, pc);
#pragma unroll
        for (int k = 0; k < 3; ++k) store_col<CH>(a.V[k] + off, lane, ny, vc[k]);
        return;
    }

    // ---- influence-matrix and sigma corrections (eqsolve_pv.cpp:407-445)
    zd c0 = {0.0, 0.0}, c1 = {0.0, 0.0}, sig[2] = {{0.0, 0.0}, {0.0, 0.0}};
    if (a.correct && a.compact && mode == MODE_STEP) {
        const zd *I0 = a.ID[0] + (size_t)q * 2, *I1 = a.ID[1] + (size_t)q * 2;
        c0 = cmul(div1, I1[0]) - cmul(div0, I1[1]);
        c1 = cmul(div0, I0[1]) - cmul(div1, I0[0]);
        if (a.sigma) {
            const zd ptop = col_at<CH>(pc, ny) + (a.PIr[0][off + ny] * c0 + a.PIr[1][off + ny] * c1);
            const zd vt0 = col_at<CH>(vc[1], ny) + (a.VIv[0][off + ny] * c0 + a.VIv[1][off + ny] * c1);
            const zd vt1 = col_at<CH>(vc[1], ny - 1) + (a.VIv[0][off + ny - 1] * c0 + a.VIv[1][off + ny - 1] * c1);
            sig[0] = cdiv((-wn_v) * vt0 - a.Re * tail0, a.SIG0[(size_t)q * 2 + 0]);
            sig[1] = cdiv((-wn_v) * vt1 + a.Re * ((2.0 * ny) * ptop - tail1), a.SIG0[(size_t)q * 2 + 1]);
        }
    } else if (a.correct) {
        const zd *I0 = a.ID[0] + (size_t)q * 2, *I1 = a.ID[1] + (size_t)q * 2;
        c0 = cmul(div1, I1[0]) - cmul(div0, I1[1]);
        c1 = cmul(div0, I0[1]) - cmul(div1, I0[0]);
        if (a.sigma) {
            const zd ptop = col_at<CH>(pc, ny) + (cmul(c0, a.PI[0][off + ny]) + cmul(c1, a.PI[1][off + ny]));
            const zd vt0 = col_at<CH>(vc[1], ny) + (cmul(c0, a.VI[0][1][off + ny]) + cmul(c1, a.VI[1][1][off + ny]));
            const zd vt1 = col_at<CH>(vc[1], ny - 1) +
                           (cmul(c0, a.VI[0][1][off + ny - 1]) + cmul(c1, a.VI[1][1][off + ny - 1]));
            sig[0] = cdiv((-wn_v) * vt0 - a.Re * tail0, a.SIG0[(size_t)q * 2 + 0]);
            sig[1] = cdiv((-wn_v) * vt1 + a.Re * ((2.0 * ny) * ptop - tail1), a.SIG0[(size_t)q * 2 + 1]);
        }
    }
    zd P[2][CH], V[3][2][CH];
    if (a.correct && a.sigma && a.compact && mode == MODE_STEP) {
        // P += c0 P0 + c1 P1 + s_(1-p) PB;  v += c0 v0 + c1 v1 + s_p vB;
        // (u, w) += i (kx, kz) (c0 f0 + c1 f1 + s_(1-p) fB)
        double r0[2][CH], r1[2][CH], rb[2][CH];
        load_rcol<CH>(a.PIr[0] + off, lane, ny, r0);
        load_rcol<CH>(a.PIr[1] + off, lane, ny, r1);
        load_rcol<CH>(a.PBr + off, lane, ny, rb);
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int c = 0; c < CH; ++c)
                P[p][c] = pc[p][c] + ((r0[p][c] * c0 + r1[p][c] * c1) + rb[p][c] * sig[1 - p]);
        load_rcol<CH>(a.VIv[0] + off, lane, ny, r0);
        load_rcol<CH>(a.VIv[1] + off, lane, ny, r1);
        load_rcol<CH>(a.VBv + off, lane, ny, rb);
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int c = 0; c < CH; ++c)
                V[1][p][c] = vc[1][p][c] + ((r0[p][c] * c0 + r1[p][c] * c1) + rb[p][c] * sig[p]);
        load_rcol<CH>(a.VIf[0] + off, lane, ny, r0);
        load_rcol<CH>(a.VIf[1] + off, lane, ny, r1);
        load_rcol<CH>(a.VBf + off, lane, ny, rb);
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                const zd phi = (r0[p][c] * c0 + r1[p][c] * c1) + rb[p][c] * sig[1 - p];
                V[0][p][c] = vc[0][p][c] + ikmul(kx, phi);
                V[2][p][c] = vc[2][p][c] + ikmul(kz, phi);
            }
    } else {
    {
        zd i0[2][CH], i1[2][CH], bb[2][CH];
        if (a.correct) {
            load_col<CH>(a.PI[0] + off, lane, ny, i0);
            load_col<CH>(a.PI[1] + off, lane, ny, i1);
        }
        if (a.sigma) load_col<CH>(a.PB + off, lane, ny, bb);
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                zd v = pc[p][c];
                if (a.correct) v = v + (cmul(c0, i0[p][c]) + cmul(c1, i1[p][c]));
                if (a.sigma) v = v + cmul(sig[1 - p], bb[p][c]); // q = (j+1) mod 2 = 1 - p
                P[p][c] = v;
            }
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        zd i0[2][CH], i1[2][CH], bb[2][CH];
        if (a.correct) {
            load_col<CH>(a.VI[0][k] + off, lane, ny, i0);
            load_col<CH>(a.VI[1][k] + off, lane, ny, i1);
        }
        if (a.sigma) load_col<CH>(a.VB[k] + off, lane, ny, bb);
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                zd v = vc[k][p][c];
                if (a.correct) v = v + (cmul(c0, i0[p][c]) + cmul(c1, i1[p][c]));
                if (a.sigma) v = v + cmul(sig[k == 1 ? p : 1 - p], bb[p][c]);
                V[k][p][c] = v;
            }
    }
    }

    if (mode == MODE_STEP && zero) {
        // correct_zero_mode, then correct_flow on the mean mode
        const zd vmean = 0.5 * (vlo[1] + vup[1]);
        double u_unit[2][CH], w_unit[2][CH];
        load_rcol<CH>(a.u_unit, lane, ny, u_unit);
        load_rcol<CH>(a.w_unit, lane, ny, w_unit);
        double qx = 0.0, qz = 0.0;
#pragma unroll
        for (int c = 0; c < CH; ++c) {
            V[1][0][c] = {0.0, 0.0};
            V[1][1][c] = {0.0, 0.0};
            const int j = jof<CH>(lane, c, 0);
            if (j <= ny) {
                qx -= V[0][0][c].x / (double)(j * j - 1);
                qz -= V[2][0][c].x / (double)(j * j - 1);
            }
        }
        if (lane == 0) V[1][0][0] = vmean;
        qx = wave_sum(qx);
        qz = wave_sum(qz);
        double dpdx, dpdz;
        if (a.grad_pre == 1) {
            dpdx = a.grad_x;
            dpdz = a.grad_z;
        } else {
            dpdx = (a.qx_mean - qx) / a.q_unit[0];
            dpdz = (a.qz_mean - qz) / a.q_unit[1];
        }
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                V[0][p][c].x += dpdx * u_unit[p][c];
                V[2][p][c].x += dpdz * w_unit[p][c];
            }
        if (lane == 0) {
            a.flow[0] = dpdx;
            a.flow[1] = dpdz;
        }
    }
    if (mode == MODE_BPROBLEM) {
        // sigma denominators (eqsolve_pv.cpp:466-476)
        const zd vn = col_at<CH>(V[1], ny), vm = col_at<CH>(V[1], ny - 1), pn = col_at<CH>(P, ny);
        if (lane == 0) {
            zd s0 = {a.Re + wn_v * vn.x, wn_v * vn.y};
            zd s1 = zd{a.Re + wn_v * vm.x, wn_v * vm.y} - (a.Re * (2.0 * ny)) * pn;
            if (zero) s1 = {1.0, 0.0};
            a.sig_out[(size_t)q * 2 + 0] = s0;
            a.sig_out[(size_t)q * 2 + 1] = s1;
        }
    }
    store_col<CH>(a.P + off, lane, ny, P);
#pragma unroll
    for (int k = 0; k < 3; ++k) store_col<CH>(a.V[k] + off, lane, ny, V[k]);
}

void launch_solve(const SolveArgs &a, int E, hipStream_t st) {
    dim3 grid((a.NQ + 3) / 4), block(256);
    if (E == 2) hipLaunchKernelGGL((k_solve<1>), grid, block, 0, st, a);
    else if (E == 4) hipLaunchKernelGGL((k_solve<2>), grid, block, 0, st, a);
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
template <int CH>
__global__ void k_pg_response(const double *WAVX, const double *WAVZ, const double *WL_, const double *WD_,
                              const double *WR_, const double *ad_v, double g_v, double Re, double aV0, double bV0,
                              double aV2, double bV2, int ny, int JS, int root_q, double *u_unit, double *w_unit,
                              double *q_unit) {
    const int lane = lane_id();
    const size_t off = (size_t)root_q * JS;
    double WL[2][CH], WD[2][CH], WR[2][CH], adv[2][CH];
    load_rcol<CH>(WL_, lane, ny, WL);
    load_rcol<CH>(WD_, lane, ny, WD);
    load_rcol<CH>(WR_, lane, ny, WR);
    load_rcol<CH>(ad_v + off, lane, ny, adv);
    const double w = -(WAVX[0] * WAVX[0] + WAVZ[root_q] * WAVZ[root_q] + g_v);
    zd rhs[2][CH];
#pragma unroll
    for (int p = 0; p < 2; ++p)
#pragma unroll
        for (int c = 0; c < CH; ++c) rhs[p][c] = {0.0, 0.0};
    if (lane == 0) rhs[0][0] = {Re, 0.0};
    for (int comp = 0; comp < 2; ++comp) {
        zd sol[2][CH];
        tau_solve<CH>(rhs, sol, adv, WL, WD, WR, lane, ny, w, g_v, comp ? aV2 : aV0, comp ? bV2 : bV0, true,
                      zd{0.0, 0.0}, zd{0.0, 0.0});
        double qs = 0.0;
        double *unit = comp ? w_unit : u_unit;
#pragma unroll
        for (int p = 0; p < 2; ++p)
#pragma unroll
            for (int c = 0; c < CH; ++c) {
                const int j = jof<CH>(lane, c, p);
                if (j <= ny) {
                    unit[j] = sol[p][c].x;
                    if (p == 0) qs -= sol[p][c].x / (double)(j * j - 1);
                }
            }
        qs = wave_sum(qs);
        if (lane == 0) q_unit[comp] = qs;
    }
}
void launch_pg_response(const double *WAVX, const double *WAVZ, const double *WL, const double *WD, const double *WR,
                        const double *ad_v, double g_v, double Re, const double *aV, const double *bV, int ny, int JS,
                        int root_q, double *u_unit, double *w_unit, double *q_unit, int E, hipStream_t st) {
    if (E == 2)
        hipLaunchKernelGGL((k_pg_response<1>), dim3(1), dim3(64), 0, st, WAVX, WAVZ, WL, WD, WR, ad_v, g_v, Re, aV[0],
                           bV[0], aV[2], bV[2], ny, JS, root_q, u_unit, w_unit, q_unit);
    else
        hipLaunchKernelGGL((k_pg_response<2>), dim3(1), dim3(64), 0, st, WAVX, WAVZ, WL, WD, WR, ad_v, g_v, Re, aV[0],
                           bV[0], aV[2], bV[2], ny, JS, root_q, u_unit, w_unit, q_unit);
}

} // namespace cdns
