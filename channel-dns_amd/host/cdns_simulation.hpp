// C++ host-side mirror of the reference's Simulation serial step loop (src/sim/simulation.hpp:21-62,
// simulation.cpp:47-67, 142-191) over the backend's C ABI (include/cdns.h): the binding a
// maintainer adds to the reference, compiled here as it would be there.  Errors keep the
// reference's behaviour: a failing entry point throws the std::runtime_error that
// CpuDevice::sync would have rethrown (src/core/cpu_device.cpp:169-177).
#pragma once

#include "cdns.h"

#include <stdexcept>
#include <string>
#include <vector>

namespace cdns_host {

// The values of Params/Mesh a serial run hands to the backend (params.hpp:6-51, mesh.hpp:15-26).
struct RunParams {
    int NX, NY, NZ, NXS, NYS, NZS, NX2, NY2, NZ2;
    double XL, ZL, THETA, Re, G_pre, a_pre0, b_pre0;
    double a_V[3], b_V[3];
    int INDEX_GRAD_PRE;
    double DT, V_Microblow0;
    int precision; // CDNS_FP64 or CDNS_MIXED
};

class Simulation {
  public:
    // Simulation::Simulation (simulation.cpp:47-67): the device context holds every buffer.
    explicit Simulation(const RunParams &p) : p_(p) {
        cdns_config c{p.NX, p.NY, p.NZ, p.NXS, p.NYS, p.NZS, p.NX2, p.NY2, p.NZ2,
                      p.XL, p.ZL, p.THETA, p.Re, p.G_pre, p.a_pre0, p.b_pre0,
                      {p.a_V[0], p.a_V[1], p.a_V[2]}, {p.b_V[0], p.b_V[1], p.b_V[2]},
                      p.INDEX_GRAD_PRE, p.DT, p.V_Microblow0, p.precision};
        if (cdns_create(&c, &ctx_) != 0) throw std::runtime_error("cdns_create failed");
    }
    ~Simulation() { cdns_destroy(ctx_); }
    Simulation(const Simulation &) = delete;
    Simulation &operator=(const Simulation &) = delete;

    // Elements of one spectral vector field in the reference layout [iz+NZS][iy][kx][c], as
    // (re, im) pairs of doubles (fields.hpp:17-22).
    size_t vector_doubles() const { return (size_t)(2 * p_.NZS + 1) * (p_.NYS + 1) * (p_.NXS + 1) * 3 * 2; }

    // Simulation::initialize (simulation.cpp:69-91) after read_restart / initialize_velocity:
    // the velocity goes to the device; the multistep history starts empty.
    void initialize(const std::vector<double> &velocity) {
        if (velocity.size() != vector_doubles()) throw std::invalid_argument("velocity size");
        ok(cdns_set_state(ctx_, 0, velocity.data(), CDNS_HOST));
        nt_ = 0;
    }

    // Simulation::run (simulation.cpp:173-191), serial branch: step_coefficients, setup on
    // it < 3, enqueue_serial_step and correct_flow are one cdns_step per iteration.
    void run(int nsteps) {
        for (int it = 1; it <= nsteps; ++it) ok(cdns_step(ctx_, it));
        ok(cdns_sync(ctx_));
        nt_ += nsteps;
    }

    // The state a check step reads (simulation.cpp:193-224): velocity, pressure, gradients.
    std::vector<double> velocity() {
        std::vector<double> v(vector_doubles());
        ok(cdns_get_state(ctx_, 0, v.data(), CDNS_HOST));
        return v;
    }
    void flow_gradients(double &dpdx, double &dpdz) {
        double g[2];
        ok(cdns_flow_gradients(ctx_, g));
        dpdx = g[0];
        dpdz = g[1];
    }
    double time() const { return cdns_time(ctx_); }
    long long steps_done() const { return nt_; }

  private:
    void ok(int rc) const {
        if (rc != 0) throw std::runtime_error(cdns_last_error(ctx_));
    }
    RunParams p_;
    cdns_ctx *ctx_ = nullptr;
    long long nt_ = 0;
};

} // namespace cdns_host
