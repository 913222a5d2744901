// gfx_backend_hip.hpp -- header-only C++ glue over the C ABI for C++ callers
// (the reference's solver tree in particular): status codes become
// std::runtime_error, the way the reference reports errors
// (forward/share/src/io.cpp:45-47; solver.cpp:733-736), and the solver handle
// gets RAII ownership.  No HIP or torch types appear here.
#pragma once

#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "gfx.h"
#include "gfx_io.h"

namespace gfx_hip {

inline void check(gfx_status s, const char* what) {
    if (s != GFX_OK) throw std::runtime_error(std::string(what) + ": " + gfx_last_error());
}

// compute_element_residual<BackendHIP> (forward/share/include/gf/element.hpp:39-42):
// host FP64 arrays, r accumulates -= K u.
inline void compute_element_residual(int n_elem, const double* dxi_dx, const double* jacobian,
                                     const double* lambda_, const double* mu_, const double* D,
                                     const double* weights, int NGLL, const double* u, double* r) {
    check(gfx_compute_element_residual(n_elem, dxi_dx, jacobian, lambda_, mu_, D, weights, NGLL, u, r),
          "HIP element residual");
}

// Owner of a device solver: the role of CudaDeviceState + cuda_allocate_state /
// cuda_free_state (forward/share/include/gf/cuda_step.hpp:20-78).
class DeviceState {
public:
    explicit DeviceState(const gfx_solver_desc& d) {
        gfx_solver* s = nullptr;
        check(gfx_solver_create(&d, &s), "HIP allocate state");
        h_.reset(s);
    }
    gfx_solver* get() const { return h_.get(); }

    // Source of one field (cuda_source_injection's data, set once, not per step).
    void set_source(int field, int direction, const std::vector<int32_t>& cells,
                    const std::vector<double>& weights, const std::vector<double>& stf) {
        check(gfx_solver_set_source(get(), field, direction, int(cells.size()), cells.data(), weights.data(),
                                    int(stf.size()), stf.data()),
              "HIP set source");
    }
    // n whole steps: predictor, element residual, source, assembly, C-PML,
    // corrector and recording (solver.cpp:510-622), on the solver's stream.
    void step(int n = 1) { check(gfx_solver_run(get(), n), "HIP step"); }
    // cuda_copy_state_to_host (cuda_step.cu:260-279), one field.
    void copy_state_to_host(int field, std::vector<double>& u, std::vector<double>& v,
                            std::vector<double>& a, int n_rank_node) const {
        u.resize(size_t(n_rank_node) * 3);
        v.resize(u.size());
        a.resize(u.size());
        check(gfx_solver_get_state(get(), field, u.data(), v.data(), a.data()), "HIP copy state");
    }
    void record(int field, int snap, double* strain, double* disp, double* vel, double* acc) const {
        check(gfx_solver_get_record(get(), field, snap, strain, disp, vel, acc), "HIP record");
    }
    int snapshots() const { return gfx_solver_n_snapshots(get()); }

private:
    struct Del {
        void operator()(gfx_solver* s) const { gfx_solver_destroy(s); }
    };
    std::unique_ptr<gfx_solver, Del> h_;
};

// Per-step stage wrappers, one for each of the reference's
// (forward/share/include/gf/cuda_step.hpp:81-127), same arguments and order of
// use (forward/share/src/solver.cpp:395-443); they act on every field of the
// state except where a field is named.
inline void newmark_predict(DeviceState& s, double dt, double beta) {
    check(gfx_step_newmark_predict(s.get(), dt, beta), "HIP newmark predict");
}
inline void gather_predicted(DeviceState& s) { check(gfx_step_gather_predicted(s.get()), "HIP gather predicted"); }
inline void gather_from_rank(DeviceState& s) { check(gfx_step_gather_from_rank(s.get()), "HIP gather"); }
inline void zero_residual(DeviceState& s) { check(gfx_step_zero_residual(s.get()), "HIP zero residual"); }
inline void launch_element_residual(DeviceState& s, int ngll, int n_elem) {
    check(gfx_step_element_residual(s.get(), ngll, n_elem), "HIP element residual");
}
inline void pml_damping(DeviceState& s) { check(gfx_step_pml_damping(s.get()), "HIP pml damping"); }
inline void source_injection(DeviceState& s, int direction, double stf_val, const double* h_src_weights,
                             int n_src_cell, int field = 0) {
    check(gfx_step_source_injection(s.get(), field, direction, stf_val, h_src_weights, n_src_cell),
          "HIP source injection");
}
inline void scatter_to_rank(DeviceState& s) { check(gfx_step_scatter_to_rank(s.get()), "HIP scatter"); }
inline void newmark_correct(DeviceState& s, double dt, double beta, double gamma) {
    check(gfx_step_newmark_correct(s.get(), dt, beta, gamma), "HIP newmark correct");
}
inline void copy_utilde_to_host(const DeviceState& s, double* host_buf, int field = 0) {
    check(gfx_step_copy_utilde_to_host(s.get(), field, host_buf), "HIP copy u~ to host");
}
inline void copy_utilde_from_host(DeviceState& s, const double* host_buf, int field = 0) {
    check(gfx_step_copy_utilde_from_host(s.get(), field, host_buf), "HIP copy u~ from host");
}
inline void copy_residual_to_host(const DeviceState& s, double* host_buf, int field = 0) {
    check(gfx_step_copy_residual_to_host(s.get(), field, host_buf), "HIP copy residual to host");
}
inline void copy_residual_from_host(DeviceState& s, const double* host_buf, int field = 0) {
    check(gfx_step_copy_residual_from_host(s.get(), field, host_buf), "HIP copy residual from host");
}

// gf::run_forward (forward/share/include/gf/solver.hpp:21) on the GPU.
inline int run_forward(const std::string& direction, bool resume_mode = false, int effective_nprocs = 0) {
    return gfx_run_forward(direction.c_str(), resume_mode ? 1 : 0, effective_nprocs);
}

}  // namespace gfx_hip
