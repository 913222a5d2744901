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
 *   complete C-PML        (tests/test_cpml_complete.py) gfo_cpml_recursion
 *                         exact on inputs linear in time (closed-form
 *                         integrals); gfo_cpml_gradient / gfo_cpml_inertia
 *                         reproduce s_x s_y s_z / (s_j s_l) and
 *                         (i w)^2 (s_x s_y s_z - K_x K_y K_z) in the periodic
 *                         steady state to second order in dt (coincident
 *                         poles, undamped axes, alpha = 0, K != 1);
 *                         gfo_element_residual_cpml with d = 0 equals the
 *                         elastic residual of the stretched geometry; a
 *                         d = 0, K = 1 run equals the elastic run bit for bit
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
    /* complete C-PML (1): replaces the reference's acceleration-only C-PML and
     * the legacy ramp; PML elements are those with pml_region != 0 (elastic,
     * relaxed moduli, when n_sls > 0) */
    int cpml_complete;
    const double* pml_K;        /* [E*n_node*3] stretch profiles per axis */
    const double* pml_d;
    const double* pml_alpha;
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

/* ---- complete C-PML ----
 * Not in the reference: its solver applies only the acceleration term of the
 * C-PML (pml.cpp) and defers the strain term (docs/deferred.md:46-80).
 * Stretches s_q(w) = K_q + d_q / (alpha_q + i w), beta_q = alpha_q + d_q / K_q.
 * In a PML element stress column j sees the gradient
 *   G^(j)_kl = F^-1[s_x s_y s_z / (s_j s_l)] * du_k/dx_l,
 *   S_ij = lambda delta_ij sum_k G^(j)_kk + mu (G^(j)_ij + G^(j)_ji),  r_i -= int S_ij dphi/dx_j,
 * and the inertia F^-1[(i w)^2 s_x s_y s_z] * rho u = rho K_x K_y K_z u'' + rho L:
 * r -= rho J w L per element point.  L is linear in v with the coefficient
 * c_v = K_z (K_y d_x + K_x d_y) + K_x K_y d_z; its velocity terms are taken at
 * v_{n+1} = v_n + dt/2 (a_n + a_{n+1}) (trapezoidal, unconditionally stable
 * for the damping): L is evaluated with v = v_n + dt/2 a_n and the rest,
 * dt/2 c_v a_{n+1}, moves to the left, so the lumped mass of a PML element
 * point is rho J w (K_x K_y K_z + dt/2 c_v).
 * The factors are chains of first-order filters (no partial fractions, so
 * coincident poles need no special case):
 *   s_q g = K_q g + d_q conv(g, alpha_q),  g / s_q = (g - (d_q / K_q) conv(g, beta_q)) / K_q,
 *   conv(g, p)(t) = int_0^t exp(-p (t - s)) g(s) ds,
 * advanced exactly for g linear in time over a step:
 *   M^{n+1} = Q + c g^{n+1},  Q <- a M^{n+1} + b g^{n+1},
 *   a = exp(-p dt), b = I0 - I1/dt, c = I1/dt, I0 = (1 - a)/p, I1 = (dt - I0)/p.
 * Inertia: C1 = conv(u, alpha_x), z1 = s_x u, C2 = conv(z1, alpha_y), z2 = s_y z1,
 * C3 = conv(z2, alpha_z); with D = d/dt and D conv(g, p) = g - p conv(g, p):
 *   L = K_z L2 + d_z (D z2 - alpha_z z2 + alpha_z^2 C3),
 *   L2 = K_y L1 + d_y (D z1 - alpha_y z1 + alpha_y^2 C2),  L1 = d_x (v - alpha_x u + alpha_x^2 C1).
 * In the step u is the predicted displacement (u_{n+1}).  With n_sls > 0 the
 * interior elements take the attenuating residual and the PML elements stay
 * elastic with the relaxed moduli.
 * Coefficients per point (27): axis q at [9q]: K, d, alpha, (a, b, c) of
 * alpha, (a, b, c) of beta (zero where d == 0).  Memories per point (45):
 * [(k*3+l)*4 + {conv(g, alpha_m), conv(g, alpha_n), conv(conv(g, alpha_m),
 * alpha_n), conv(s_m s_n g, beta_l)}] with m < n the axes other than l, then
 * [36 + c*3 + {Q1, Q2, Q3}] of the inertia. */
void gfo_cpml_recursion(double p, double dt, double* abc);
void gfo_cpml_point_coefficients(const double* K, const double* d, const double* alpha, double dt, double* coef27);
/* g9[k*3+l] = du_k/dx_l -> G27[j*9 + k*3 + l]; carry36 updated */
void gfo_cpml_gradient(const double* coef27, const double* g9, double* carry36, double* G27);
/* u3, v3 -> L3; carry9 updated */
void gfo_cpml_inertia(const double* coef27, const double* u3, const double* v3, double* carry9, double* L3);
void gfo_element_residual_cpml(int n_elem, const double* dxi_dx, const double* jacobian, const double* lambda_,
                               const double* mu_, const int32_t* pml_region, const double* coef27,
                               const double* D, const double* weights, int ngll, const double* u,
                               double* carry45, double* r);
/* Histories of one point from zero memories: g[n][9] -> G27[n][27] and
 * (u, v)[n][3] -> L3[n][3] (either pair may be NULL). */
void gfo_cpml_filter_history(const double* coef27, int n, const double* g, const double* u, const double* v,
                             double* G27, double* L3);

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
