/* gfx_surrogate.h -- strain Green's tensor tiles on the device, the GLL
 * interpolation judge, and the Fourier-feature MLP that stands in for the
 * tile + interpolation query (BF16 MFMA, FP32 accumulation).
 */
#ifndef GFX_SURROGATE_H
#define GFX_SURROGATE_H

#include "gfx.h"

#ifdef __cplusplus
extern "C" {
#endif

/* ---- tiles ------------------------------------------------------------
 * A tile holds Green's tensor values at the unique GLL nodes of a set of
 * cells: [n_time][n_node][6 strain components][3 force directions] FP32 on
 * the device, with the cells' node lists [n_cell][125] (tensor order
 * (i*5+j)*5+k, i along x), as postprocess/cpp/main.cpp:554-579 assembles
 * greens[nt, n_gll, 6, 3] and writer.hh stores it per tile. */
typedef struct gfx_green gfx_green;
gfx_status gfx_green_create(int n_time, int n_node, const double* node_xyz, int n_cell,
                            const int32_t* cell_node_index, gfx_green** out);
gfx_status gfx_green_destroy(gfx_green* g);
gfx_status gfx_green_set_values(gfx_green* g, int time_index, const float* values);
gfx_status gfx_green_get_values(const gfx_green* g, int time_index, float* values);
/* Map forming node values from a solver's records: node n takes
 * sum_e weight_e * strain[rec_point_e] / sum_e weight_e over entries
 * [off[n], off[n+1]); rec_point = position of the cell in the solver's
 * recording-cell list * 125 + local point; weight = the point's lumped mass
 * (the mass-weighted merge of postprocess/cpp/main.cpp:354-426). */
gfx_status gfx_green_set_capture_map(gfx_green* g, const int64_t* off, const int32_t* rec_point,
                                     const double* weight);
/* Fill time index t from snapshot `snap` of solver s (still in its record
 * ring): field_of_dir[d] is the solver field driven by the force in
 * direction d (d = x, y, z), or -1 to leave that direction at zero. */
gfx_status gfx_green_capture(gfx_green* g, const gfx_solver* s, int snap, const int32_t* field_of_dir,
                             int time_index);

/* ---- Exact tile slabs ----------------------------------------------------
 * The post-processor's node merge and tensor layout for one snapshot, on the
 * device (postprocess/cpp/main.cpp:306-426 merge; 554-575, 581-598 layout;
 * 800-811 per-tile copy).  n_node merged nodes; node n is formed from the
 * entries [off[n], off[n+1]) of (rec_point, weight), rec_point = position of
 * the cell in the solver's recording-cell list * 125 + local point, in the
 * order the post-processor visits them (record cell, then point):
 *   strain  = (sum_e weight_e * strain_e) * (1 / sum_e weight_e)
 *   displ.  = (sum_e displacement_e)      * (1 / sum_e weight_e)
 * (no scaling where the weight sum is not positive), each product and sum
 * rounded in FP64 as the host loop rounds it, then stored as FP32 or FP64.
 * The slab lists n_slab node indices slab_node[] (the tiles' local node
 * lists, concatenated); a capture writes greens[n_slab*18] with the 18 values
 * of an entry force-direction major (index d*6 + c, the byte order of the
 * reference's /field/greens_tensor [nt, n, 6, 3]) and, when made with
 * displacement, displacement[n_slab*9] at index c*3 + d. */
typedef struct gfx_green_slab gfx_green_slab;
gfx_status gfx_green_slab_create(int n_node, const int64_t* off, const int32_t* rec_point, const double* weight,
                                 int64_t n_slab, const int32_t* slab_node, int precision_f64,
                                 int with_displacement, gfx_green_slab** out);
gfx_status gfx_green_slab_destroy(gfx_green_slab* t);
/* Snapshot `snap` of solver s (still in its record ring; field_of_dir[d] =
 * the solver field driven in direction d, or -1 for zeros) into the host
 * buffers greens (float or double per precision_f64) and displacement (NULL
 * unless made with_displacement, which needs a solver recording u, v, a). */
gfx_status gfx_green_slab_capture(gfx_green_slab* t, const gfx_solver* s, int snap, const int32_t* field_of_dir,
                                  void* greens, void* displacement);

/* Judge: GLLInterpolator.interpolate (greenfun/gll_interpolator.py:117-227)
 * on the device for n points xyz[n*3] at tile time indices tix[n]; out[n*18]
 * (NaN outside all cells); cell[n] receives the cell used (-1 outside), may
 * be NULL. */
gfx_status gfx_green_interpolate(const gfx_green* g, int n, const double* xyz, const int32_t* tix, float* out,
                                 int32_t* cell);
/* The same rule on the host in FP64 for one point over values[n_node*nv];
 * GFX_ERR_INVALID when the point is outside all cells (the reference raises). */
gfx_status gfx_gll_interpolate_host(int n_node, const double* node_xyz, int n_cell, const int32_t* cell_node_index,
                                    int nv, const double* values, const double* point, double* out);

/* ---- MLP ----------------------------------------------------------------
 * (x, y, z, t) scaled to [-1, 1] by [in_lo, in_hi] -> [sin, cos](2 pi B x),
 * B ~ N(0, sigma^2) [n_freq][4] -> depth x (Linear(width) + GELU) ->
 * Linear(18), outputs multiplied by out_scale[18] and by the envelope e(r)
 * (gfx_mlp_set_envelope; 1 until set).  Parameters are one FP32 array (per
 * linear layer: W [in][out] row-major, then b [out]); products use BF16
 * operands on MFMA with FP32 accumulation.  Training runs Adam with the step
 * size decaying from lr to lr_final on a cosine over each gfx_mlp_train call. */
typedef struct {
    int n_freq, width, depth, n_out;
    float sigma;
    uint64_t seed;
    float lr, beta1, beta2, eps;
    int batch;
    float lr_final;
} gfx_mlp_config;
typedef struct gfx_mlp gfx_mlp;
gfx_status gfx_mlp_config_default(gfx_mlp_config* c);
gfx_status gfx_mlp_create(const gfx_mlp_config* c, const double* in_lo, const double* in_hi,
                          const double* out_scale, gfx_mlp** out);
gfx_status gfx_mlp_destroy(gfx_mlp* m);
gfx_status gfx_mlp_n_params(const gfx_mlp* m, size_t* n);
/* fourier (may be NULL) receives B [n_freq][4]. */
gfx_status gfx_mlp_get_params(const gfx_mlp* m, float* params, size_t n, float* fourier);
gfx_status gfx_mlp_set_params(gfx_mlp* m, const float* params, size_t n);
/* Source frame of the network, r = horizontal distance from (cx, cy):
 *  - moveout: the network's time input is t - slowness * r (0 = plain t), so
 *    arrivals travelling at about 1/slowness line up;
 *  - amplitude envelope: e(r) piecewise linear through env[i] at
 *    r_i = (i + 0.5) * r_max / n_bins, constant beyond the first and last
 *    knots, env[i] > 0; the network is trained on (value / out_scale) / e(r)
 *    and its output is multiplied back, so it fits the Green's functions with
 *    their geometric decay taken out.  n_bins = 0: no envelope.
 * lo/hi of the time input (gfx_mlp_create) refer to the network's time input. */
gfx_status gfx_mlp_set_envelope(gfx_mlp* m, double cx, double cy, double slowness, double r_max, int n_bins,
                                const float* env);
/* x raw [n][4] (x, y, z, t); y [n][18]. */
gfx_status gfx_mlp_forward(gfx_mlp* m, int n, const double* x, float* y);
/* Loss = mean over n x 18 of ((out - y) / (out_scale e(r)))^2, out the
 * network's scaled output, and its gradient [n_params] (for checking against
 * a CPU model). */
gfx_status gfx_mlp_gradient(gfx_mlp* m, int n, const double* x, const float* y, float* grad, double* loss);
/* Adam on random (node, time) samples of the tile (time index k -> t0 + k dt);
 * *loss = mean loss of the last 10 % of iterations. */
gfx_status gfx_mlp_train(gfx_mlp* m, const gfx_green* g, double t0, double dt, int iters, uint64_t seed,
                         double* loss);
/* Query: per region fallback[cell * n_window + window] = 0: the MLP answers;
 * 1: the judge answers (also everywhere outside the tile); 2: the answer is
 * zero (a region whose values are within the error bound of 0).  Regions are
 * (cell, time window): the tile's n_time samples split into n_window windows,
 * time index k in window k * n_window / n_time.  *n_fallback counts judge
 * answers. */
gfx_status gfx_surrogate_query(gfx_mlp* m, const gfx_green* g, const uint8_t* fallback, int n_window, double t0,
                               double dt, int n, const double* xyz, const int32_t* tix, float* out,
                               int* n_fallback);

#ifdef __cplusplus
}
#endif
#endif
