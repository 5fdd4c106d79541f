/* cdns.h -- C ABI of the MI355X (gfx950) backend for the channel-dns per-step hot path.
 *
 * Plain pointers and sizes only.  A context owns one HIP stream and every device
 * buffer of one serial (single-rank) channel-flow run; calls are enqueued on the
 * stream in order (the FIFO-batch semantics of the reference's CpuDevice,
 * src/core/cpu_device.hpp:30-130) and cdns_sync() waits for them.  Every entry
 * point returns 0 on success and a nonzero code on failure; cdns_last_error()
 * gives the message.  Errors detected by the device are reported at the next
 * cdns_sync(), as CpuDevice::sync rethrows (src/core/cpu_device.cpp:169-177).
 *
 * Layouts are the reference's host layouts (src/solver/fields.hpp:17-48):
 *   spectral scalar  [(iz+NZS)][iy][kx]       complex (re,im), iy<=NYS, kx<=NXS
 *   spectral vector  [(iz+NZS)][iy][kx][c]    c fastest
 *   boundary         [(iz+NZS)][wall][kx][c]
 *   fine physical    [iz][iy][ix]             real, NZ2 x (NY2+1) x NX2
 * `where` selects host (CDNS_HOST) or device (CDNS_DEVICE) pointers.  Device
 * pointers must stay valid until the next cdns_sync().
 */
#ifndef CDNS_H
#define CDNS_H

#ifdef __cplusplus
extern "C" {
#endif

#define CDNS_HOST 0
#define CDNS_DEVICE 1

#define CDNS_FP64 0  /* every stage in binary64 */
#define CDNS_MIXED 1 /* transforms and products in binary32; state, solves, mean mode in binary64 */

/* Params + Mesh of one serial run (src/core/params.hpp:6-51, src/core/mesh.hpp:15-26). */
typedef struct cdns_config {
    int NX, NY, NZ;    /* coarse grid */
    int NXS, NYS, NZS; /* retained modes */
    int NX2, NY2, NZ2; /* product grid */
    double XL, ZL, THETA, Re, G_pre, a_pre0, b_pre0;
    double a_V[3], b_V[3];
    int INDEX_GRAD_PRE;
    double DT, V_Microblow0;
    int precision; /* CDNS_FP64 or CDNS_MIXED */
} cdns_config;

typedef struct cdns_ctx cdns_ctx;

/* Replaces the construction of Transforms + EqSolvePV + TimeStepper + Simulation
 * buffers (src/main.cpp:192-221; src/sim/simulation.cpp:47-67). */
int cdns_create(const cdns_config *cfg, cdns_ctx **out);
void cdns_destroy(cdns_ctx *ctx);
const char *cdns_last_error(const cdns_ctx *ctx);
/* CpuDevice::sync (src/core/cpu_device.cpp:169-177). */
int cdns_sync(cdns_ctx *ctx);
/* The context's HIP stream (hipStream_t), for callers that order their own work with it. */
void *cdns_stream(cdns_ctx *ctx);

/* Transforms::channel_xyzfft2 / enqueue_xyzfft2 (src/fft/transforms.hpp:40-43,
 * transforms.cpp:588-741): dir 'B' synthesises the product-grid field from the
 * spectral component at spec[0], spec[stride], ...; 'F' writes only that component. */
int cdns_xyzfft2(cdns_ctx *ctx, char dir, double *phys, double *spec, int stride, int where);

/* Transforms::channel_xyzfft (src/fft/transforms.hpp:39, transforms.cpp:452-582): the coarse
 * grid NX x (NY+1) x NZ, physical layout [iz][iy][ix], with the Nyquist rules of NXS = NX/2 and
 * NZS = NZ/2; same stride and "write only the selected component" contract as cdns_xyzfft2. */
int cdns_xyzfft(cdns_ctx *ctx, char dir, double *phys, double *spec, int stride, int where);

/* FieldIO::to_physical (src/io/field_io.cpp:199-220): the state's velocity (three consecutive
 * [iz][iy][ix] arrays) and pressure on the coarse grid; finite = every retained spectral
 * coefficient of both is finite.  Either output may be NULL. */
int cdns_to_physical(cdns_ctx *ctx, double *velocity, double *pressure, int *finite, int where);
/* Statistics::reduce (src/sim/statistics.cpp:54-109) on the state: per plane iy the 20 sums
 * u v w p, uu vv ww, uv vw wu, pu pv pw pp, u^3 v^3 w^3, u^4 v^4 w^4 (p = pressure - |u|^2/2;
 * the last 13 only when higher != 0) into sums[(NY+1) x 20], host memory; finite[NY+1] or NULL. */
int cdns_plane_moments(cdns_ctx *ctx, int higher, double *sums, int *finite);
/* Diagnostics::evaluate (src/sim/diagnostics.cpp:39-75) on the state: out = {EKT = sum v^2 over
 * the grid nodes, TKE = Clenshaw-Curtis volume mean of 1/2 sum_c var_xz(u_c), L2 norm of the
 * physical divergence}, finite as the reference combines it. */
int cdns_diagnostics(cdns_ctx *ctx, double out[3], int *finite);

/* vector_ops::curl and vector_ops::cross_product (src/solver/vector_ops.hpp:22,33). */
int cdns_curl(cdns_ctx *ctx, double *out, const double *in, int where);
int cdns_cross_product(cdns_ctx *ctx, double *out, const double *a, const double *b, int where);

/* TimeStepper::nonlinear_term (src/solver/timestepper.hpp:10-11): N = u x curl(u) and
 * the nonlinear pressure wall condition. */
int cdns_nonlinear_term(cdns_ctx *ctx, double *nl, const double *velocity, double *pressure_bc,
                        int where);

/* EqSolvePV::setup (src/solver/eqsolve_pv.hpp:17). */
int cdns_setup(cdns_ctx *ctx, double g_v);
/* EqSolvePV::solve_modes (eqsolve_pv.hpp:48-51); rhs_p == NULL forms div(rhs_v) in place.
 * solve_coupled (eqsolve_pv.hpp:18-20) is rhs_p != NULL with correct_sigma = 0. */
int cdns_solve_modes(cdns_ctx *ctx, double *pressure, const double *rhs_p, const double *pressure_bc,
                     double *velocity, const double *rhs_v, const double *velocity_bc,
                     int correct_sigma, int where);
/* Setup tables in reference layout; `which` numbers as the EqSolvePV members
 * (eqsolve_pv.hpp:61-75): 0 wn_p 1 wn_v 2 ad_p 3 ad_v 4 P_infl0 5 P_infl1 6 V_infl0
 * 7 V_infl1 8 infl_div0 9 infl_div1 10 P_b 11 V_b 12 sigma0 13 u_dpdx 14 w_dpdz 15 {qx,qz}. */
int cdns_get_table(cdns_ctx *ctx, int which, double *out);

/* Simulation state (src/sim/simulation.hpp:44-55).  field: 0 velocity, 1 velocity_prev,
 * 2 nonlinear, 3 nonlinear_prev, 4 pressure, 5 pressure_bc, 6 pressure_bc_prev, 7 velocity_bc.
 * The multistep history (1, 3, 6) is what a restart file does not hold.  A CDNS_MIXED context
 * keeps fields 2 and 3 (the step's nonlinear term and its history) in single precision: values
 * set there are rounded to float, values read are widened. */
int cdns_set_state(cdns_ctx *ctx, int field, const double *data, int where);
int cdns_get_state(cdns_ctx *ctx, int field, double *data, int where);
/* One iteration `it` (>= 1) of Simulation::run's serial body (src/sim/simulation.cpp:173-191):
 * step_coefficients (:36-45), setup on it < 3 (:176-177), enqueue_serial_step (:142-171) with
 * correct_flow (:111-140).  From it = 3 on, the solver tables are those of the steady g; they are
 * set up again if another call (cdns_setup, cdns_adjoint_step of iteration 1) changed them. */
int cdns_step(cdns_ctx *ctx, int it);
/* Iterations it_begin .. it_begin+nsteps-1. */
int cdns_run(cdns_ctx *ctx, int it_begin, int nsteps);
/* dpdx_, dpdz_ after the last step (simulation.cpp:111-140) and the simulated time. */
int cdns_flow_gradients(cdns_ctx *ctx, double out[2]);
double cdns_time(const cdns_ctx *ctx);

/* Per-kernel timing with HIP events on the context stream (no reference counterpart;
 * the reference times only the whole loop, src/main.cpp:223-234). */
int cdns_profile(cdns_ctx *ctx, int enable);
/* Fills up to `cap` entries: kernel name (NUL-separated in `names`, 64 bytes each),
 * launch count and total milliseconds; returns the number of kernels. */
int cdns_profile_read(cdns_ctx *ctx, char *names, long long *counts, double *total_ms, int cap);
int cdns_profile_reset(cdns_ctx *ctx);

/* ---- Discrete adjoint of the step (no reference counterpart: the reference integrates
 * forward only).  FP64 contexts.  Adjoint arrays use the layouts of the state they belong to;
 * gradients are d/d(re) + i d/d(im) of a real objective. */

/* Constant body force f (spectral vector, reference layout) added to every step's nonlinear
 * term, so that it enters the RHS with weight b0 + b1 = 1 and the pressure wall condition
 * through f_y; NULL removes it. */
int cdns_set_body_force(cdns_ctx *ctx, const double *force, int where);
/* Terminal objective of a velocity: which = 0 kinetic energy 1/2 int |u|^2 dV, 1 mean wall
 * shear 1/2 (du/dy(-1) - du/dy(+1)) of the plane mean.  Writes J and (if grad) dJ/du. */
int cdns_objective(cdns_ctx *ctx, int which, const double *velocity, double *grad, double *J,
                   int where);
/* Transpose of step `it` at the state whose velocity is `velocity`: the adjoints of the state
 * after the step (velocity, velocity_prev, nonlinear_prev: vectors; pressure_bc_prev: boundary)
 * are replaced by those of the state before it; adj_force (or NULL) accumulates the body-force
 * gradient.  The solver tables are set up for iteration it if they are not. */
int cdns_adjoint_step(cdns_ctx *ctx, int it, const double *velocity, double *adj_velocity,
                      double *adj_velocity_prev, double *adj_nonlinear_prev,
                      double *adj_pressure_bc_prev, double *adj_force, int where);
/* Gradient of objective(u(T)) after iterations 1..nsteps from u0 (zero history) with body force
 * `force` (or none): forward with a checkpoint every k_ckpt steps, then per segment a forward
 * recomputation that keeps each step's synthesised (u, curl u), and the adjoint steps in reverse.
 * Writes J, grad_u0 and (if non-NULL) grad_force.  The context's state (velocity and history) is
 * overwritten; its time and its body force (cdns_set_body_force) are restored.  The checkpoints
 * and one segment's fields stay allocated with the context between calls (grown when a call needs
 * more): (ceil(nsteps / k_ckpt) (9 S + 2 NQ) + k_ckpt * 6 Z) * 16 bytes, S, Z and NQ the spectral,
 * z-physical and column counts. */
int cdns_adjoint_run(cdns_ctx *ctx, const double *u0, const double *force, int nsteps, int k_ckpt,
                     int objective, double *grad_u0, double *grad_force, double *J, int where);

/* No counterpart in the reference: for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).  cdns_adjoint_run with
 * the checkpoints and the segment's fields in a file-scope device array of 64 MiB per GPU (held
 * until the process ends) instead of memory the context allocates; it calls no hipMalloc or
 * hipFree.  Device pointers only (where = CDNS_DEVICE), the context's device current.  It fails,
 * doing nothing, when the horizon needs more than 64 MiB (see cdns_adjoint_run for the size; the
 * checkpoint part is rounded up to 256 bytes), when the context's adjoint workspaces do not exist
 * yet (an earlier cdns_adjoint_run or cdns_adjoint_step creates them) or, with a force, when no
 * body-force buffer exists yet (cdns_set_body_force creates it).  Every call on one GPU uses the
 * same array (every entry is written before it is read, nothing needs clearing): calls that
 * overlap on one GPU, from any context, race. */
int cdns_adjoint_run_noalloc(cdns_ctx *ctx, const double *u0, const double *force, int nsteps, int k_ckpt,
                             int objective, double *grad_u0, double *grad_force, double *J, int where);

/* ---- Seeded initial fields (host only, no device work). */

/* initial_bc::DisturbanceEngine, which is std::mt19937_64 (src/sim/initial_bc.hpp:13), seeded
 * as DisturbanceEngine(seed).  cdns_engine_draw writes n draws in sequence: kind 0 the raw 64-bit
 * outputs, kind 1 initial_bc::centered_uniform, (engine() >> 11) 2^-53 - 0.5
 * (src/sim/initial_bc.cpp:15-18), as doubles.  The engine continues where the last draw ended. */
typedef struct cdns_engine cdns_engine;
int cdns_engine_create(unsigned long long seed, cdns_engine **out);
int cdns_engine_draw(cdns_engine *engine, int kind, void *out, unsigned long long n);
void cdns_engine_destroy(cdns_engine *engine);

#ifdef __cplusplus
}
#endif
#endif
