/*
 * gfx.h -- C ABI of the MI355X-native elastic CG-SEM forward solver.
 *
 * Plain C: pointers and sizes only, no C++ or torch types.  Every entry point
 * returns a gfx_status; on failure gfx_last_error() gives a message (thread
 * local).  Nothing in this library aborts the process.
 *
 * Array layouts on the host side are the reference's own
 * (forward/share/include/gf/types.hpp:27-95): element-major, per element
 * point index n = (i*NGLL + j)*NGLL + k, dxi_dx[e][n][m*3+p] = d(xi_p)/d(x_m),
 * vectors interleaved xyz, node ids 0-based (local_cell2global_node).
 */
#ifndef GFX_H
#define GFX_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    GFX_OK = 0,
    GFX_ERR_INVALID = 1,     /* bad argument or shape */
    GFX_ERR_DEVICE = 2,      /* HIP runtime error */
    GFX_ERR_UNSUPPORTED = 3, /* valid request this build does not implement */
    GFX_ERR_NOMEM = 4,       /* host or device allocation failed */
    GFX_ERR_IO = 5           /* file input/output (HDF5) failed */
} gfx_status;

/* Message of the last failed call on this thread ("" if none). */
const char* gfx_last_error(void);
/* Library version string. */
const char* gfx_version(void);
/* Number of HIP devices visible (0 when no GPU); never fails. */
int gfx_device_count(void);

/* -------------------------------------------------------------------------
 * Element plugin.
 * Replaces gf::compute_element_residual<Backend>
 * (forward/share/include/gf/element.hpp:39-42; CUDA specialisation
 * forward/elastic/src/element_cuda.cu:195-252).  Host FP64 pointers,
 * element-major AoS, r ACCUMULATES the internal force (r -= K u), as the CPU
 * specialisation does (forward/elastic/src/element_cpu.cpp:116).  The device
 * computes in FP64, so results match the CPU backend to rounding.
 * Any NGLL in [2, 8] is accepted.
 * ---------------------------------------------------------------------- */
gfx_status gfx_compute_element_residual(int n_elem, const double* dxi_dx, const double* jacobian,
                                        const double* lambda_, const double* mu_,
                                        const double* D, const double* weights, int ngll,
                                        const double* u, double* r);

/* Same contract, FP32 device arithmetic (the solver's arithmetic); inputs are
 * still the reference's FP64 host arrays. NGLL must be 5. */
gfx_status gfx_compute_element_residual_f32(int n_elem, const double* dxi_dx,
                                            const double* jacobian, const double* lambda_,
                                            const double* mu_, const double* D,
                                            const double* weights, int ngll, const double* u,
                                            double* r);

/* Same contract and FP64 arithmetic as gfx_compute_element_residual, with no
 * device allocation: no counterpart in the reference; it is for callers that
 * cannot allocate device memory (graph capture, serving loops, harnesses that
 * forbid hipMalloc).  The device copies of the inputs and of r are staged in
 * a fixed array of GFX_NOALLOC_SCRATCH_BYTES per GPU, which the library holds
 * from its first use on that GPU until the process ends; each call overwrites
 * every byte it reads, so no contents carry over from one call to the next.
 * Calls needing more (about 3,500 elements at NGLL 5) fail with
 * GFX_ERR_INVALID.  Calls in one process are serialised; calls that overlap on
 * one GPU from several threads would otherwise race on the shared array. */
#define GFX_NOALLOC_SCRATCH_BYTES (64u << 20)
gfx_status gfx_compute_element_residual_noalloc(int n_elem, const double* dxi_dx,
                                                const double* jacobian, const double* lambda_,
                                                const double* mu_, const double* D,
                                                const double* weights, int ngll,
                                                const double* u, double* r);

/* -------------------------------------------------------------------------
 * Device-resident solver state.
 * Replaces CudaDeviceState + cuda_allocate_state/cuda_free_state
 * (forward/share/include/gf/cuda_step.hpp:20-79); the per-step host wrappers
 * (cuda_step.hpp:81-128) are the gfx_step_* entry points below, and
 * gfx_solver_run is their whole step fused.  Adds what the CUDA path lacks: the C-PML memory update of the CPU path
 * (forward/share/src/pml.cpp:55-216), device strain recording
 * (forward/share/src/solver.cpp:282-335,640-675) and many right-hand sides
 * ("fields") advanced by one launch.
 * ---------------------------------------------------------------------- */
typedef struct gfx_solver gfx_solver;

typedef enum {
    GFX_ASSEMBLY_COLORED = 0, /* deterministic: element colouring, plain read-modify-write */
    GFX_ASSEMBLY_ATOMIC = 1,  /* float atomics; order not deterministic */
    GFX_ASSEMBLY_GATHER = 2   /* deterministic: element-local residual, per-node
                                 fixed-order sum fused into the node update */
} gfx_assembly;

typedef enum {
    GFX_ELEMENT_AUTO = 0,  /* wave kernel when applicable, else block kernel */
    GFX_ELEMENT_BLOCK = 1, /* one lane per GLL point, workgroup barriers between stages */
    GFX_ELEMENT_WAVE = 2,  /* one wavefront per element, two points per lane, no barriers */
    GFX_ELEMENT_PAIR = 3,  /* as BLOCK, two fields per pass (nodal C-PML only) */
    GFX_ELEMENT_PENCIL = 4 /* lane per 5-point pencil along k, ten elements per workgroup
                              (fixed; elems_per_block must be 0), k contractions in
                              registers (nodal C-PML only) */
} gfx_element_kernel;

typedef struct {
    int n_elem;              /* local cells */
    int ngll;                /* must be 5 (polynomial order 4) */
    int n_rank_node;         /* unique nodes, ids in [0, n_rank_node) */
    /* [n_elem*ngll^3*...] reference layout, FP64 */
    const double* dxi_dx;    /* x9 */
    const double* jacobian;
    const double* lambda_;
    const double* mu_;
    const double* mass;      /* element-local lumped mass, already times density */
    const double* density;   /* used by C-PML only */
    const double* damping;   /* legacy ramp; used only when has_cpml == 0 */
    const int32_t* ibool;    /* local_cell2global_node */
    const double* elem_centroid; /* [n_elem*3] optional; elements are then processed
                                    along a Morton curve of their centroids */
    /* C-PML (pml_coef_abar non-empty in the partition file => has_cpml) */
    int has_cpml;
    const int32_t* pml_region; /* [n_elem] 0 = interior */
    const double* coef_alpha;  /* x9 */
    const double* coef_abar;   /* x5 */
    /* time stepping */
    double dt;
    int snapshot_stride;
    /* recording cells (local element indices) */
    int n_rec_cell;
    const int32_t* rec_cell;
    /* batching */
    int n_fields;            /* right-hand sides advanced together (>= 1) */
    int assembly;            /* gfx_assembly */
    int elems_per_block;     /* elements per workgroup: 1, 2, 4 or 8; 0 = library default */
    int element_kernel;      /* gfx_element_kernel */
    /* device record ring: snapshots kept (0 = no recording); u,v,a copies on/off */
    int record_capacity;
    int record_uva;
    /* node kernels: field groups per workgroup (1, 2, 4 or 8; 0 = 1): the
       fields of one node block are spread over that many lane groups */
    int node_field_groups;
    /* 0: elements whose dxi_dx is diagonal at every point (off-diagonal
       entries at most 1e-12 of the largest diagonal one, i.e. axis-aligned up
       to round-off) run the element kernels' diagonal form, which leaves out
       the products with the zero entries; 1: every element runs the general
       form */
    int general_geometry;
    /* shear attenuation by n_sls standard linear solids (0 = elastic; at most
       8): per point and mechanism the stress and strain relaxation times
       [n_elem*ngll^3*n_sls] (caller order); mu_ is the relaxed modulus.  The
       elements then run the attenuating block kernel (general geometry; the
       C-PML must take the nodal form).  Not in the reference, whose solver is
       elastic (its planned inputs: /field/cell/tau_sigma, tau_epsilon). */
    int n_sls;
    const double* tau_sigma;
    const double* tau_epsilon;
    /* Complete C-PML (1; needs has_cpml).  Not in the reference,
       whose solver applies only the acceleration part of its C-PML
       (forward/share/src/pml.cpp) and defers the strain part
       (docs/deferred.md:46-80).  From the stretch profiles
       s_q = K_q + d_q / (alpha_q + i w) of every element point
       ([n_elem*ngll^3*3], the partition datasets pml_K, pml_d, pml_alpha), the
       PML elements (pml_region != 0) take the stretched weak form: stress
       column j sees the gradient filtered by s_x s_y s_z / (s_j s_l), the mass
       gains K_x K_y K_z and the inertia terms of (i w)^2 s_x s_y s_z enter the
       residual.  coef_alpha and coef_abar are then not used, and the node
       kernels apply no C-PML term.  The memories (45 per PML point and field)
       are advanced exactly for inputs linear in time over a step.  With
       n_sls > 0 the interior elements attenuate and the PML elements stay
       elastic with the relaxed moduli. */
    int cpml_complete;
    const double* pml_K;
    const double* pml_d;
    const double* pml_alpha;
} gfx_solver_desc;

gfx_status gfx_solver_create(const gfx_solver_desc* desc, gfx_solver** out);

/* The nodes the complete C-PML holds rigid (mask[n_rank_node], 1 = rigid;
 * host only): its outer boundary, the faces of PML elements that no other
 * element shares and whose normal (largest component of grad xi at the face
 * centre) is an axis the element's region damps.  They get zero mass, so the
 * massless-node rule keeps their u and v (zero from rest).  Needs n_elem,
 * ngll, n_rank_node, ibool, dxi_dx and pml_region of the descriptor.  No
 * counterpart in the reference, whose C-PML has free outer faces. */
gfx_status gfx_cpml_rigid_nodes(const gfx_solver_desc* desc, uint8_t* mask);
gfx_status gfx_solver_destroy(gfx_solver* s);

/* Point-force source of one field: the cells containing the point with their
 * normalised Lagrange weights [n_src_cell*ngll^3] (config.h5 /source/cells),
 * the force direction (0=x,1=y,2=z) and stf[n_stf] = f(step*dt) (as
 * forward/share/src/solver.cpp:576-605). */
gfx_status gfx_solver_set_source(gfx_solver* s, int field, int direction, int n_src_cell,
                                 const int32_t* src_elem, const double* src_weights, int n_stf,
                                 const double* stf);

/* Zero u, v, a, C-PML memory, step counter and record buffers. */
gfx_status gfx_solver_reset(gfx_solver* s);

/* Advance n_steps steps.  At every step with step % snapshot_stride == 0 the
 * strain (and u, v, a) of the recording cells is written to the device
 * record buffer for that snapshot. */
gfx_status gfx_solver_run(gfx_solver* s, int n_steps);

/* Current step counter (number of steps taken). */
int gfx_solver_step_index(const gfx_solver* s);
/* Set the step counter (steps already taken, e.g. after gfx_solver_set_state
 * on restart) and the snapshot counter that follows from it. */
gfx_status gfx_solver_set_step(gfx_solver* s, int step);

/* Copy a field's nodal state to host FP64 [n_rank_node*3]; NULL skips
 * (cuda_copy_state_to_host, forward/share/src/cuda_step.cu:260-279). */
gfx_status gfx_solver_get_state(const gfx_solver* s, int field, double* u, double* v, double* a);
/* Load a field's nodal state from host FP64 [n_rank_node*3] (for restart). */
gfx_status gfx_solver_set_state(gfx_solver* s, int field, const double* u, const double* v,
                                const double* a);

/* -------------------------------------------------------------------------
 * Per-stage entry points: the reference's per-step device wrappers
 * (forward/share/include/gf/cuda_step.hpp:81-127), for a driver that keeps the
 * reference's device branch (forward/share/src/solver.cpp:390-444) and its
 * multi-rank exchange hooks.  One step, in that order:
 *   gfx_step_newmark_predict, [gfx_step_copy_utilde_to_host / _from_host],
 *   gfx_step_gather_predicted, gfx_step_zero_residual,
 *   gfx_step_element_residual, gfx_step_pml_damping,
 *   gfx_step_source_injection (per field with a source this step),
 *   gfx_step_scatter_to_rank, [gfx_step_copy_residual_to_host / _from_host],
 *   gfx_step_newmark_correct.
 * Calls out of this order fail with GFX_ERR_INVALID; gfx_solver_run refuses
 * a solver in the middle of such a step.  The state between steps is the one
 * gfx_solver_run keeps (the displacement array holds the next step's
 * predicted displacement), so with GATHER assembly a step driven stage by
 * stage equals a gfx_solver_run step bit for bit (sources injected with the
 * values gfx_solver_set_source gave).  Work runs on the solver's stream; the
 * copies to the host wait for it.  Arrays are caller-numbered FP64
 * [n_rank_node*3].
 * ---------------------------------------------------------------------- */
/* cuda_newmark_predict (cuda_step.hpp:84): checks beta = 0 and dt; the
 * predicted displacement u + dt v + dt^2/2 a was formed by the previous
 * step's corrector (or by creation / gfx_solver_set_state), so no device work. */
gfx_status gfx_step_newmark_predict(gfx_solver* s, double dt, double beta);
/* cuda_gather_predicted / cuda_gather_from_rank (cuda_step.hpp:113-116): the
 * element kernels read the nodal arrays through the node ids, so there is no
 * element-local copy to fill (no device work). */
gfx_status gfx_step_gather_predicted(gfx_solver* s);
gfx_status gfx_step_gather_from_rank(gfx_solver* s);
/* cuda_zero_residual (cuda_step.hpp:87): clears the assembled residual for
 * ATOMIC assembly (a fill kernel); COLORED and GATHER need no clearing. */
gfx_status gfx_step_zero_residual(gfx_solver* s);
/* cuda_launch_element_residual (cuda_step.hpp:107): element internal force of
 * every field (plus element-local C-PML or attenuation terms where active),
 * without sources; ngll must be 5 and n_elem the solver's element count. */
gfx_status gfx_step_element_residual(gfx_solver* s, int ngll, int n_elem);
/* cuda_pml_damping (cuda_step.hpp:90): the legacy ramp v -= d v (solvers
 * without C-PML) acts on v_n inside the corrector, where the reference's CPU
 * step applies it (solver.cpp:565-573); no device work here. */
gfx_status gfx_step_pml_damping(gfx_solver* s);
/* cuda_source_injection (cuda_step.hpp:93-94): r[cell point][direction] +=
 * stf_val * src_weights[c][point] for the n_src_cell cells given to
 * gfx_solver_set_source for this field (n_src_cell must match; negative cell
 * ids are skipped, as off-rank cells are). */
gfx_status gfx_step_source_injection(gfx_solver* s, int field, int direction, double stf_val,
                                     const double* src_weights, int n_src_cell);
/* cuda_scatter_to_rank (cuda_step.hpp:110): GATHER assembly sums each node's
 * element-point residuals in fixed order into the nodal residual; COLORED and
 * ATOMIC assembled it in the element stage. */
gfx_status gfx_step_scatter_to_rank(gfx_solver* s);
/* cuda_newmark_correct (cuda_step.hpp:97): a = r/m (0 where m <= 0, u and v
 * kept), v, C-PML memory and terms, legacy damping, then the next step's
 * predicted displacement; records the snapshot when step % snapshot_stride
 * == 0 and closes the step (beta must be 0, gamma 1/2). */
gfx_status gfx_step_newmark_correct(gfx_solver* s, double dt, double beta, double gamma);
/* cuda_copy_state_to_host (cuda_step.hpp:100-101): gfx_solver_get_state
 * between steps. */
gfx_status gfx_step_copy_state_to_host(const gfx_solver* s, int field, double* u, double* v, double* a);
/* cuda_copy_utilde_{to,from}_host (cuda_step.hpp:119-122): the predicted
 * displacement; "from" only between predict and the element stage. */
gfx_status gfx_step_copy_utilde_to_host(const gfx_solver* s, int field, double* buf);
gfx_status gfx_step_copy_utilde_from_host(gfx_solver* s, int field, const double* buf);
/* cuda_copy_residual_{to,from}_host (cuda_step.hpp:125-128): the assembled
 * residual, between scatter and correct. */
gfx_status gfx_step_copy_residual_to_host(const gfx_solver* s, int field, double* buf);
gfx_status gfx_step_copy_residual_from_host(gfx_solver* s, int field, const double* buf);

/* Number of snapshots held so far. */
int gfx_solver_n_snapshots(const gfx_solver* s);
/* Copy snapshot `snap` of `field`: strain [n_rec_cell][ngll^3][6], and u, v, a
 * [n_rec_cell][ngll^3][3] (any may be NULL), FP64 like the reference's write
 * buffers (forward/share/src/solver.cpp:651-675). */
gfx_status gfx_solver_get_record(const gfx_solver* s, int field, int snap, double* strain,
                                 double* disp, double* vel, double* acc);

/* Complete time-stepping state as an opaque host image (predicted
 * displacement, velocity, acceleration where kept, C-PML memory, the last
 * assembled residual, step and snapshot counters).  Loading it into a solver
 * built from the same inputs, field count and assembly continues the run bit
 * for bit; any other solver rejects it (GFX_ERR_INVALID).  This goes beyond
 * the reference's restart file (forward/share/src/restart.cpp:120-152), which
 * keeps neither the C-PML memory nor the exact step state. */
gfx_status gfx_solver_checkpoint_size(const gfx_solver* s, size_t* bytes);
gfx_status gfx_solver_checkpoint_save(const gfx_solver* s, void* host, size_t bytes);
gfx_status gfx_solver_checkpoint_load(gfx_solver* s, const void* host, size_t bytes);

/* Per-kernel device time of the last gfx_solver_run (HIP events on the
 * solver's stream).  name is "element", "node" (both node kernels),
 * "node_special", "node_regular", "record" or "total"; returns milliseconds
 * summed over the run and the launch count. */
gfx_status gfx_solver_kernel_time(const gfx_solver* s, const char* name, double* ms,
                                  int* launches);
/* Enable/disable per-kernel event timing (default off). */
gfx_status gfx_solver_set_timing(gfx_solver* s, int enable);

/* Switch the element kernel form and elements per workgroup of an existing
 * solver (same contract as the descriptor fields; results do not depend on it). */
gfx_status gfx_solver_set_launch(gfx_solver* s, int element_kernel, int elems_per_block);
/* Capture one whole step into a HIP graph and replay it for every step of
 * gfx_solver_run (1), or launch kernel by kernel (0, default).  The step and
 * snapshot counters live on the device, so one captured step serves every
 * step; the graph is recaptured after any launch or source change.  Per-kernel
 * timing (gfx_solver_set_timing) uses the kernel-by-kernel path. */
gfx_status gfx_solver_set_graph(gfx_solver* s, int enable);
/* Switch the node kernels' field groups per workgroup (1, 2, 4 or 8). */
gfx_status gfx_solver_set_node_groups(gfx_solver* s, int node_field_groups);

/* Autotuner: times trial_steps steps of every launch configuration
 * (assembly COLORED / GATHER x element kernel BLOCK / WAVE x elements per
 * workgroup, then node-kernel field groups for the best of those) for the
 * problem in `base` and writes the fastest into *best (a copy of base with
 * assembly, element_kernel, elems_per_block and node_field_groups set);
 * *best_ms receives its device time per step.  Sources and recording are
 * not part of the trial. */
gfx_status gfx_autotune(const gfx_solver_desc* base, int trial_steps, gfx_solver_desc* best,
                        double* best_ms);

/* Introspection of the solver's setup.  key: "colors", "pml_nodal" (1 when the
 * C-PML memory is kept per node, 0 when element-local), "pml_elements",
 * "special_nodes", "pml_nodes", "fields", "nodes", "elements", "assembly",
 * "element_kernel", "elems_per_block", "node_field_groups", "diag_elements"
 * (elements on the diagonal-geometry form), "n_sls"; with the complete C-PML
 * "cpml_complete", "cpml_points" (PML element points), "cpml_memories"
 * (memories in use per field, summed over those points) and "cpml_groups"
 * (launches of the PML element kernel per step: one per set of damped axes,
 * and per colour under COLORED assembly). */
gfx_status gfx_solver_query(const gfx_solver* s, const char* key, double* value);

/* Algorithmic HBM bytes per step for the current problem (each array counted
 * once per step).  which: 0 = whole step, 1 = element stage, 2 = node stage
 * (both node kernels), 3 = special-node kernel, 4 = regular-node kernel. */
double gfx_solver_algorithmic_bytes(const gfx_solver* s, int which);

/* -------------------------------------------------------------------------
 * Batches of independent solvers (uneven meshes, each with its own dt,
 * C-PML, record ring and run length).  No counterpart in the reference, whose
 * pipeline runs meshes and directions one process at a time.
 * ---------------------------------------------------------------------------- */
typedef struct gfx_batch gfx_batch;
/* The solvers stay owned by the caller and must outlive the batch; while they
 * belong to a batch they must not be run or changed concurrently with it. */
gfx_status gfx_batch_create(int n, gfx_solver* const* solvers, gfx_batch** out);
/* Advance every instance by up to n_steps steps, instance i only while its
 * step counter is below until[i] (until may be NULL).  Each step of all active
 * instances is one replay of a captured graph, cached per active set.  With
 * launch groups enabled (gfx_batch_set_groups) and every active instance
 * using GATHER assembly, nodal C-PML (or none), no attenuation and the same
 * launch shape (element kernel form, elements per workgroup, node field
 * groups), the instances are split into launch groups of about equal work and
 * each group shares one launch per stage (element stage, special nodes,
 * regular nodes, counter update; results bitwise equal to running each
 * instance alone); the groups are independent branches of the graph.
 * Otherwise the graph holds one branch of launches per instance.  Returns
 * when all work has finished. */
gfx_status gfx_batch_run(gfx_batch* b, int n_steps, const int* until);
/* Number of launch groups where launches can be shared (default 0: one
 * branch per instance, whose launches the device overlaps across instances). */
gfx_status gfx_batch_set_groups(gfx_batch* b, int n_groups);
/* Launch groups the last step gfx_batch_run took used (0: one branch per
 * instance). */
int gfx_batch_groups(const gfx_batch* b);
gfx_status gfx_batch_destroy(gfx_batch* b);

/* -------------------------------------------------------------------------
 * Structured-box inputs (restates preprocess/ geometry, material, CFL,
 * C-PML, source-location and recording-map rules for regular hex boxes;
 * the reference preprocessor itself is Python and out of scope).
 * ---------------------------------------------------------------------- */
#define GFX_MAX_LAYERS 8

typedef struct {
    int nx, ny, nz;          /* elements per axis */
    double lx, ly, lz;       /* box size (m); x,y in [0,l], z in [0,lz] downward, free surface z=0 */
    int order;               /* polynomial order N (ngll = N+1) */
    /* material: n_layers >= 1 layers; layer l holds z < ztop[l+1] (last is halfspace) */
    int n_layers;
    double layer_zbot[GFX_MAX_LAYERS]; /* bottom depth of layer l (ignored for the last) */
    double vp[GFX_MAX_LAYERS], vs[GFX_MAX_LAYERS], rho[GFX_MAX_LAYERS];
    /* random velocity perturbation: vp,vs *= (1 + rf_amp * f), f ~ Gaussian field */
    double rf_amp, rf_corr_len;
    int rf_modes;
    uint64_t rf_seed;
    /* C-PML thickness in elements (zmin is the free surface and has none) */
    int pml_xmin, pml_xmax, pml_ymin, pml_ymax, pml_zmax;
    double f0_pml;           /* Hz */
    /* interior vertex jitter as a fraction of the element size (0 = regular box) */
    double jitter;
    uint64_t jitter_seed;
    /* time stepping (cfl_validator.py): dt = output_dt/stride <= cfl*h_min/vp_max */
    double cfl_safety, output_dt, duration;
    double record_depth_max;  /* m; 0 = no recording */
    int want_coords;          /* produce coords/vp/vs (x3 memory) */
    int want_beta;            /* produce pml_coef_beta */
    int want_strain;          /* produce pml_coef_strain (A6..A23 of the strain-based C-PML) */
    /* shear attenuation: n_sls standard linear solids fitted to a constant
       q_mu[layer] over [att_fmin, att_fmax] Hz (gfx_sls_fit); 0 = elastic;
       q_mu <= 0 leaves that layer elastic */
    int n_sls;
    double q_mu[GFX_MAX_LAYERS];
    double att_fmin, att_fmax;
    /* C-PML stretch profiles K, d, alpha per axis at every point
       (pml_K, pml_d, pml_alpha [E,n,3], the inputs of the complete C-PML) */
    int want_profiles;
    /* 0: the profiles as preprocess/pml_cpml.py:63-155 computes them (the
       normalised distance measured from the outer boundary face, and for an
       axis with PML on both faces the later face's values, so d is largest at
       the interior interface); 1: graded into the layer, the distance measured
       from the interior interface as pml_cpml.py:44-60 defines it (d = 0 and
       alpha = alpha_max at the interface, d largest and alpha = 0 at the
       boundary), each face inside its own layer only; same d0, exponent and
       alpha_max */
    int pml_profile;
    /* target reflection R of the damping d0 = 3 vp ln(1/R) / (2 w) of both
       profiles; 0 = the reference's 1e-5 (pml_cpml.py:23) */
    double pml_reflection;
} gfx_box_spec;

typedef struct {
    int n_elem, ngll, n_node_glob;
    int nx, ny, nz;
    double xmin, xmax, ymin, ymax, zmin, zmax;
    double* coords;     /* [E,n,3] or NULL */
    double* jacobian;   /* [E,n]   */
    double* dxi_dx;     /* [E,n,9] */
    double* mass;       /* [E,n]   */
    double* vp;         /* [E,n] or NULL */
    double* vs;         /* [E,n] or NULL */
    double* density;    /* [E,n]   */
    double* lambda_;    /* [E,n]   */
    double* mu_;        /* [E,n]   */
    double* damping;    /* [E,n]   */
    int32_t* ibool;     /* [E*n]   */
    double* node_xyz;   /* [G,3] global node coordinates */
    int32_t* pml_region;  /* [E] */
    double* coef_alpha;   /* [E,n,9] */
    double* coef_beta;    /* [E,n,9] or NULL */
    double* coef_strain;  /* [E,n,18] or NULL: per point A0 A6 A7 A8 | A0 A10 A11 A12 | A0 A14 A15 A16 |
                             A18 A19 | A20 A21 | A22 A23 */
    double* coef_abar;    /* [E,n,5] */
    int n_pml_elem;
    double h_min, vp_max, cfl_dt, dt;
    int snapshot_stride, nsteps;
    double record_depth_actual;
    int n_rec_cell;
    int32_t* rec_cell;  /* [n_rec_cell] */
    int n_sls;            /* shear attenuation mechanisms (0 = elastic) */
    double* tau_sigma;    /* [E,n,n_sls] or NULL */
    double* tau_epsilon;  /* [E,n,n_sls] or NULL */
    double* pml_K;        /* [E,n,3] or NULL (want_profiles) */
    double* pml_d;
    double* pml_alpha;
} gfx_mesh;

gfx_status gfx_box_spec_default(gfx_box_spec* spec);
gfx_status gfx_mesh_build_box(const gfx_box_spec* spec, gfx_mesh** out);
gfx_status gfx_mesh_free(gfx_mesh* m);

/* Locate a point force (preprocess/source_locator.py:244-375): surface mode
 * when z == zmin (free-surface cells), else buried mode (non-PML cells).
 * Fills up to max_cells cells and weights [n_src*ngll^3], normalised to sum 1
 * over all containing cells.  *n_src receives the count. */
gfx_status gfx_mesh_locate_source(const gfx_mesh* m, double x, double y, double z, int max_cells,
                                  int32_t* cells, double* weights, int* n_src);

/* Standard linear solids for a constant shear quality factor q over
 * [f_min, f_max] Hz: relaxation frequencies log-spaced over the band
 * (tau_sigma_l = 1 / (2 pi f_l)) and strengths from the linear least-squares
 * fit of Emmerich and Korn (1987) to 1/Q at 10 n_sls log-spaced frequencies,
 * returned as tau_epsilon_l.  The modulus of the solids is
 * M(w) = M_R (1 + sum_l (tau_eps_l / tau_sig_l - 1) i w tau_sig_l / (1 + i w tau_sig_l)). */
gfx_status gfx_sls_fit(double q, double f_min, double f_max, int n_sls, double* tau_sigma, double* tau_epsilon);

/* Strain-based C-PML coefficients of one point (preprocess/pml_cpml.py:523-734,
 * compute_strain_coefficients): K, d, alpha per direction (x, y, z) and the
 * region code (1 X, 2 Y, 3 Z, 4 XY, 5 XZ, 6 YZ, 7 XYZ; 0 gives zeros) to the
 * 18 values of pml_coef_strain.  With s_a = K_a + d_a / (alpha_a + i w) and
 * beta_a = alpha_a + d_a / max(K_a, 1), each group of four is the partial
 * fraction A0 + A1/(alpha_1 + i w) + A2/(alpha_2 + i w) + A3/(beta_3 + i w) of
 * the region's form G(s_1, s_2, s_3) -- XYZ: s_1 s_2 / s_3, X: s_1, Y: s_2,
 * Z: 1 / s_3, XY: s_1 s_2, XZ: s_1 / s_3, YZ: s_2 / s_3 -- with directions
 * (1, 2, 3) = (z, y, x) for entries 0-3, (x, z, y) for 4-7 and (x, y, z) for
 * 8-11.  (The reference keeps the region code when it permutes the
 * directions, so outside XYZ the first two groups are not s_z s_y / s_x and
 * s_x s_z / s_y; the values are kept as it writes them.)  Entries 12-17 are
 * (K_a, K_a d_a / max(K_a, 1)) of s_x, s_y, s_z in the directions the region
 * damps, (1, 0) in the others.  Poles closer than 1e-6 are separated as the
 * reference does. */
gfx_status gfx_pml_strain_coefficients(int region, const double* K, const double* d, const double* alpha,
                                       double* out18);

/* Ricker wavelet A (1 - 2a^2) exp(-a^2), a = pi f0 (t - t0), sampled at
 * t = i*dt for i < n (preprocess/stf_evaluator.py:19-45). */
gfx_status gfx_ricker(double f0, double t0, double amplitude, double dt, int n, double* out);

#ifdef __cplusplus
}
#endif
#endif
