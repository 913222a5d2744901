/*
 * gfx_oracle.h -- CPU FP64 restatement of the reference's elastic CG-SEM
 * forward step (CPU branch of forward/share/src/solver.cpp).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * solver in gf-only-sem_amd/.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product never links it and
 * never falls back to it.
 *
 * Layouts follow the reference exactly (forward/share/include/gf/types.hpp:27-95):
 *   per element point arrays are element-major, node index n = (i*NGLL+j)*NGLL+k,
 *   dxi_dx[e][n][9] with [m*3+p] = d(xi_p)/d(x_m), vectors are xyz-interleaved.
 *
 * Pinning (what checks this restatement, by function; tests/test_oracle.py):
 *   gfo_gll_*             reference tests/test_gll.cpp known answers, sum of
 *                         weights, D exact on polynomials, Lagrange cardinality
 *   gfo_element_residual  tests/test_element.cpp invariants (rigid translation,
 *                         rotation, uniaxial force balance), an independent
 *                         numpy Voigt-matrix stiffness, the six rigid-body null
 *                         modes of K, the mu <= 0 skip
 *   gfo_scatter/gather    tests/test_assembly.cpp exact sums, round trip
 *   gfo_newmark_*         tests/test_newmark.cpp closed forms; the exact
 *                         discrete eigenmode cos(n theta) of the whole step
 *   source injection      momentum balance sum(m a) = F(t_n) at every step
 *   gfo_apply_pml_damping the four cases of tests/test_pml.cpp:11-101; the
 *                         legacy branch of the whole step (damping between
 *                         predictor and corrector) against the closed form of
 *                         a damped rigid translation
 *   gfo_cpml_update_displ_fields  three hand-worked calls with v, a != 0 (the
 *                         swap, c1, c2, interior elements untouched)
 *   gfo_cpml_update_displ_memory  hand-worked update with distinct
 *                         coefficients per direction and component
 *   gfo_cpml_accel_contribution   weighted-sum closed form; steady state u/b
 *   whole C-PML step      three hand-worked steps of one element (call order:
 *                         fields, then the term from the old memory, corrector,
 *                         memory update)
 *   gfo_full_strain       exact strain sym(A) of affine fields u = A x + b on a
 *                         jittered mesh, and the recording-cell subset
 *   gfo_run_from          resume from (u, v, a) equals the uninterrupted run
 *                         without C-PML
 * No reference test covers C-PML; the hand-worked sequences pin the
 * restatement of forward/share/src/pml.cpp:55-216 and its call order.
 */
#ifndef GFX_ORACLE_H
#define GFX_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- GLL quadrature (forward/share/include/gf/gll.hpp:14-147) ---- */
int gfo_gll_nodes(int N, double* nodes);
int gfo_gll_weights(int N, const double* nodes, double* w);
int gfo_gll_derivative_matrix(int N, const double* nodes, double* D);
void gfo_lagrange_basis(double xi, int ngll, const double* nodes, double* ell);

/* ---- element internal force (forward/elastic/src/element_cpu.cpp:19-162) ---- */
void gfo_element_residual(int n_elem, const double* dxi_dx, const double* jacobian,
                          const double* lambda_, const double* mu_, const double* D,
                          const double* weights, int ngll, const double* u, double* r);

/* ---- assembly (forward/share/src/assembly.cpp:10-41) ---- */
void gfo_scatter_to_rank(const double* loc, const int32_t* ibool, int n_elem, int n_node,
                         int n_rank_node, double* glob);
void gfo_gather_from_rank(const double* glob, const int32_t* ibool, int n_elem, int n_node,
                          double* loc);

/* ---- Newmark (forward/share/src/solver.cpp:40-69) ---- */
void gfo_newmark_predict(int n_dof, double dt, double beta, const double* u, const double* v,
                         const double* a, double* ut);
void gfo_newmark_correct(int n_dof, double dt, double beta, double gamma, const double* mass,
                         double* u, double* v, double* a, const double* r);

/* ---- legacy linear-ramp damping (forward/share/src/pml.cpp:13-29) ----
 * v[i] -= damping_node[i/3] * v[i] wherever the node's damping is > 0. */
void gfo_apply_pml_damping(const double* damping_node, double* v, long n_dof);

/* ---- C-PML (forward/share/src/pml.cpp:55-216) ---- */
typedef struct {
    int n_elem, ngll;
    const int32_t* pml_region;   /* [n_elem]            */
    const double* coef_alpha;    /* [n_elem*n_node*9]   */
    const double* coef_abar;     /* [n_elem*n_node*5]   */
    const double* density;       /* [n_elem*n_node]     */
    const double* jacobian;      /* [n_elem*n_node]     */
    const int32_t* ibool;        /* [n_elem*n_node]     */
    double* displ_old;           /* [n_elem*n_node*3]   state */
    double* displ_new;           /* [n_elem*n_node*3]   state */
    double* rmemory;             /* [n_elem*n_node*9]   state */
} gfo_cpml;
void gfo_cpml_update_displ_fields(gfo_cpml* p, const double* u, const double* v, const double* a,
                                  double dt);
void gfo_cpml_update_displ_memory(gfo_cpml* p);
void gfo_cpml_accel_contribution(const gfo_cpml* p, const double* u, const double* v,
                                 const double* gll_w, double* r_loc);

/* ---- strain (forward/share/src/solver.cpp:282-335) ---- */
void gfo_full_strain(int n_cell, const int32_t* cells, const double* dxi_dx, const double* D,
                     int ngll, const double* u_loc_all, double* strain);

/* ---- whole forward run (CPU branch of solver.cpp:389-688) ---- */
typedef struct {
    int n_elem, ngll, n_rank_node;
    const double* dxi_dx;       /* [E*n_node*9] */
    const double* jacobian;     /* [E*n_node]   */
    const double* lambda_;      /* [E*n_node]   */
    const double* mu_;          /* [E*n_node]   */
    const double* mass;         /* [E*n_node]  element-local lumped mass (already * rho) */
    const double* density;      /* [E*n_node]   */
    const double* damping;      /* [E*n_node]  legacy ramp (used only without C-PML) */
    const int32_t* ibool;       /* [E*n_node]  local_cell2global_node */
    int has_cpml;
    const int32_t* pml_region;  /* [E]            */
    const double* coef_alpha;   /* [E*n_node*9]   */
    const double* coef_abar;    /* [E*n_node*5]   */
    int n_src_cell;
    const int32_t* src_elem;    /* [n_src_cell] local element index, -1 = absent */
    const double* src_weights;  /* [n_src_cell*n_node] */
    int direction;              /* 0,1,2 */
    int nsteps;
    double dt;
    int n_stf;
    const double* stf;          /* [n_stf] f(step*dt) */
    int snapshot_stride;
    int n_rec_cell;
    const int32_t* rec_cell;    /* [n_rec_cell] local element index */
    /* shear attenuation by standard linear solids (0 = elastic) */
    int n_sls;
    const double* tau_sigma;    /* [E*n_node*n_sls] stress relaxation times (s) */
    const double* tau_epsilon;  /* [E*n_node*n_sls] strain relaxation times (s) */
} gfo_problem;

/* ---- shear attenuation (standard linear solids) ----
 * Not in the reference (its forward/ solver is elastic; the planned inputs are
 * per-point tau_sigma, tau_epsilon per mechanism).  mu_ is the relaxed shear
 * modulus mu_R.  Per mechanism l:
 *   Y_l = 2 mu_R (tau_eps_l / tau_sig_l - 1),  mu_U = mu_R (1 + sum_l (tau_eps_l/tau_sig_l - 1)),
 *   lambda_U = lambda_R + 2 (mu_R - mu_U) / 3   (bulk modulus unchanged),
 *   dz_l/dt = (Y_l dev(eps) - z_l) / tau_sig_l   (z_l: traceless symmetric memory tensor),
 *   sigma = lambda_U tr(eps) I + 2 mu_U eps - sum_l z_l.
 * Over a step the ODE is integrated exactly for strain linear in time between
 * the two levels: z^{n+1} = a z^n + Y (b dev^n + c dev^{n+1}),
 *   a = exp(-dt/tau), b = tau (1 - a)/dt - a, c = 1 - tau (1 - a)/dt.
 * Memory per point: z [n_sls][5] (xx yy xy xz yz; zz = -xx - yy) and the
 * previous deviatoric strain [5]. */
void gfo_sls_coefficients(double tau_sigma, double tau_epsilon, double mu_relaxed, double dt, double* abcY);
void gfo_element_residual_att(int n_elem, const double* dxi_dx, const double* jacobian, const double* lambda_,
                              const double* mu_, int n_sls, const double* tau_sigma, const double* tau_epsilon,
                              double dt, const double* D, const double* weights, int ngll, const double* u,
                              double* zeta, double* dev_prev, double* r);
/* Stress response of one point to a deviatoric strain history dev[n][5]
 * (dev[0] is the level before the first step): sig[n][5] deviatoric stress
 * after each of the n-1 steps, by the same update as the element residual. */
void gfo_sls_point_history(int n_sls, const double* tau_sigma, const double* tau_epsilon, double mu_relaxed,
                           double dt, int n, const double* dev, double* sig);

/* Runs nsteps steps.  Outputs (any may be NULL):
 *   rec_strain [n_snap][n_rec_cell][n_node][6], rec_disp/vel/acc [n_snap][n_rec_cell][n_node][3]
 *   with n_snap = number of steps with step % snapshot_stride == 0.
 *   u_out/v_out/a_out [n_rank_node*3]: final state.
 * Returns the number of snapshots written, or -1 on bad input. */
int gfo_run(const gfo_problem* pb, double* rec_strain, double* rec_disp, double* rec_vel,
            double* rec_acc, double* u_out, double* v_out, double* a_out);

/* Reference resume (solver.cpp:369-388): runs steps [start_step, nsteps) from
 * the state u, v, a (in/out) with the C-PML state at zero, recording the
 * snapshot steps among them (outputs as gfo_run, first snapshot first).
 * Returns the number of snapshots written, or -1 on bad input. */
int gfo_run_from(const gfo_problem* pb, int start_step, double* u, double* v, double* a,
                 double* rec_strain, double* rec_disp, double* rec_vel, double* rec_acc);

/* Runs nsteps starting from a given state (u,v,a in/out, C-PML state zero at
 * start); no recording.  Used by step-level parity tests. */
int gfo_run_state(const gfo_problem* pb, double* u, double* v, double* a);

/* glibc rand()-based fill in [-1,1], as in the reference's
 * tests/test_element_cuda.cu:27-32 (random_fill). */
void gfo_random_fill(double* v, int n, unsigned seed);

/* Threads for the element-local and nodal loops (OpenMP; the scatter stays
 * serial); returns the count in effect.  Results do not depend on it: every
 * parallel loop writes disjoint entries. */
int gfo_set_threads(int n);

#ifdef __cplusplus
}
#endif
#endif
