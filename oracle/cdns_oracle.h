/* CPU oracle for the channel-dns per-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library restates the reference solver's
 * algorithm in plain C so that the HIP backend can be checked against it.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it; the product path never links or calls it.
 *
 * All arrays use the reference's host layouts (src/solver/fields.hpp:17-48,
 * src/core/array4d.hpp:13-72, serial single-rank mesh):
 *   spectral scalar  U[(iz+NZS)][iy][kx]          complex, iy<=NYS, kx<=NXS
 *   spectral vector  U[(iz+NZS)][iy][kx][c]       complex, c fastest
 *   boundary         B[(iz+NZS)][wall][kx][c]     complex
 *   fine physical    P[iz][iy][ix]                real, NZ2 x (NY2+1) x NX2
 *   coarse physical  P[iz][iy][ix]                real, NZ  x (NY+1)  x NX
 *   real mode field  wn[(iz+NZS)][kx]; real scalar ad[(iz+NZS)][iy][kx]
 * Complex numbers are two doubles (re, im).
 */
#ifndef CDNS_ORACLE_H
#define CDNS_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int NX, NY, NZ, NXS, NYS, NZS, NX2, NY2, NZ2;
    double XL, ZL, THETA, Re, G_pre, a_pre0, b_pre0;
    double a_V[3], b_V[3];
    int INDEX_GRAD_PRE;
    double DT, V_Microblow0;
} orc_config;

typedef struct orc_ctx orc_ctx;

orc_ctx *orc_create(const orc_config *cfg);
void orc_destroy(orc_ctx *ctx);
const char *orc_last_error(void);
int orc_set_threads(int n);

/* Wavenumber and Chebyshev tables (transforms.cpp:71-108). */
void orc_tables(const orc_ctx *c, double *WAVX, double *WAVZ, double *WL, double *WD, double *WR);

/* Fine (product-grid) transform, channel_xyzfft2 (transforms.cpp:588-741). */
int orc_xyzfft2(orc_ctx *c, char dir, double *phys, double *spec, int stride);
/* Coarse transform, channel_xyzfft (transforms.cpp:452-582). */
int orc_xyzfft(orc_ctx *c, char dir, double *phys, double *spec, int stride);

/* vector_ops (vector_ops.cpp). */
/* Statistics::reduce plane sums (statistics.cpp:54-90) of physical [iz][iy][ix] fields into
 * sums[(NY+1) x 20]; Diagnostics weights (diagnostics.cpp:17-33) and evaluate (:39-75); returns
 * the finite flag. */
void orc_plane_moments(const orc_ctx *c, const double *u, const double *v, const double *w, const double *pr,
                       int higher, double *sums, int *finite);
void orc_cc_weights(const orc_ctx *c, double *wt);
int orc_diagnostics(orc_ctx *c, const double *sums, const int *plane_finite, const double *velocity,
                    int spectral_finite, double out[3]);
void orc_derivative(orc_ctx *c, int axis, double *out, const double *in);
void orc_curl(orc_ctx *c, double *out, const double *in);
void orc_divergence(orc_ctx *c, double *out, const double *in);
void orc_gradient(orc_ctx *c, double *out, const double *in);
void orc_divergence_boundary(orc_ctx *c, double *out, const double *in);
double orc_flow_rate(const orc_ctx *c, const double *scalar_or_vector, int stride, int comp);
void orc_cross_product(orc_ctx *c, double *out, const double *a, const double *b);

/* TimeStepper::nonlinear_term (timestepper.cpp:11-43): writes nl and the pressure BC. */
void orc_nonlinear_term(orc_ctx *c, double *nl, const double *velocity, double *pressure_bc);

/* EqSolvePV (eqsolve_pv.cpp). */
void orc_setup(orc_ctx *c, double g_v);
void orc_solve_modes(orc_ctx *c, double *pressure, const double *rhs_pressure /*NULL=fused div*/,
                     const double *pressure_bc, double *velocity, const double *rhs_velocity,
                     const double *velocity_bc, int correct_sigma);
void orc_solve_staged(orc_ctx *c, double *pressure, const double *rhs_pressure,
                      const double *pressure_bc, double *velocity, const double *rhs_velocity,
                      const double *velocity_bc, int correct_sigma);
/* Single tau solve for one scalar field; which=0 pressure tables, 1 velocity tables. */
void orc_solve_tau(orc_ctx *c, double *out, const double *rhs, double g, double a, double b,
                   const double *bc, int which);
void orc_step_coefficients(int it, double dt, double Re, double out[5]);

/* Setup tables, for comparison with a device setup.  which:
 * 0 wn_p[(NZS2P1)(NXS+1)], 1 wn_v, 2 ad_p [S real], 3 ad_v,
 * 4 P_infl0 [S cplx], 5 P_infl1, 6 V_infl0 [3S cplx], 7 V_infl1, 8 infl_div0 [bnd cplx], 9 infl_div1,
 * 10 P_b, 11 V_b, 12 sigma0 [bnd cplx], 13 u_dpdx [NYS+1], 14 w_dpdz, 15 {qx,qz} */
void orc_get_table(const orc_ctx *c, int which, double *out);
/* g of the current velocity tables (NaN before any setup). */
double orc_setup_g(const orc_ctx *c);

/* Simulation state (simulation.hpp:44-55). field: 0 velocity, 1 velocity_prev,
 * 2 nonlinear, 3 nonlinear_prev, 4 pressure, 5 pressure_bc, 6 pressure_bc_prev, 7 velocity_bc */
void orc_set_state(orc_ctx *c, int field, const double *data);
void orc_get_state(const orc_ctx *c, int field, double *data);
/* One iteration of Simulation::run's serial body (simulation.cpp:173-191); it>=1. */
void orc_step(orc_ctx *c, int it);
void orc_run(orc_ctx *c, int nsteps);
/* Constant body force added to every step's nonlinear term (no reference counterpart). */
void orc_set_body_force(orc_ctx *c, const double *f_or_null);
void orc_flow_gradients(const orc_ctx *c, double out[2]);
double orc_time(const orc_ctx *c);

/* std::mt19937_64 restatement, for seeded-input tests. */
typedef struct { uint64_t mt[312]; int idx; } orc_mt64;
void orc_mt64_seed(orc_mt64 *g, uint64_t seed);
uint64_t orc_mt64_next(orc_mt64 *g);

/* ---- tangent-linear and discrete adjoint of the step (cdns_adjoint.c; no reference
 * counterpart).  A step state X is (u, u_prev, N_prev, B_prev) in reference layouts. */
typedef struct orc_adj orc_adj;
orc_adj *orc_adj_create(orc_ctx *c, const orc_config *cfg);
void orc_adj_destroy(orc_adj *a);
void orc_adj_fine_backward_T(orc_adj *a, const double *phys, double *spec);
void orc_adj_fine_forward_T(orc_adj *a, const double *spec, double *phys);
void orc_adj_curl_T(orc_adj *a, const double *lw, double *lu);
/* transpose of solve_modes (fused RHS, sigma on) on its velocity output, with the tables of
 * the context's current setup */
void orc_adj_solve_T(orc_adj *a, const double *lv, double *lr, double *lbc);
void orc_adj_tl_step(orc_adj *a, int it, const double *const X[4], const double *const dX[4],
                     const double *df, double *const dXout[4]);
void orc_adj_step(orc_adj *a, int it, const double *const X[4], double *const L[4], double *lf);
double orc_adj_objective(orc_adj *a, int which, const double *vel, double *grad);
double orc_adj_run(orc_adj *a, const double *u0, const double *f, int nsteps, int kc, int objective,
                   double *grad_u0, double *grad_f);

#ifdef __cplusplus
}
#endif
#endif
