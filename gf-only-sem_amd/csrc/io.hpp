// io.hpp -- the reference's on-disk formats: partition_{r}.h5 and config.h5
// readers (with the single-process merge), the per-snapshot record writer,
// the restart file, and writers of the same input files for generated meshes.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

#include "gfx.h"

namespace gfx::io {

// Recording map of one partition (forward/share/include/gf/types.hpp:85-101).
struct RecordingMap {
    bool has_recording = false;
    std::vector<int64_t> gll_node_ids;       // [n_unique]
    std::vector<double> gll_node_coords;     // [n_unique*3]
    std::vector<int32_t> rec_cell_local;     // [n_rec_cell]
    std::vector<int32_t> cell_gll_node_index;  // [n_rec_cell*ngll^3] into gll_node_ids
    std::vector<int64_t> rec_cell_global;    // [n_rec_cell]
    // attributes of /recording kept for the writer
    double record_depth_max_m = 0.0, record_depth_actual_m = 0.0;
};

// Per-rank data (RankData, forward/share/include/gf/types.hpp:24-103) without
// the MPI ghost/exchange lists, which a single-process run clears anyway
// (forward/share/src/io.cpp:357-360).
struct Partition {
    int n_local_cell = 0;
    int ngll = 0;
    std::vector<int64_t> local_cell_ids;  // 0-based global element ids
    std::vector<double> coords, jacobian, dxi_dx, mass, vp, vs, density, lambda_, mu_, damping;
    std::vector<int32_t> pml_region;
    std::vector<double> coef_alpha, coef_beta, coef_abar, coef_strain;
    // C-PML stretch profiles /field/cell/pml_K, pml_d, pml_alpha [n_cell, n, 3]
    std::vector<double> pml_K, pml_d, pml_alpha;
    // shear attenuation (planned in the reference, not read by its solver):
    // /field/cell/tau_sigma, tau_epsilon [n_cell, ngll, ngll, ngll, n_sls]
    int n_sls = 0;
    std::vector<double> tau_sigma, tau_epsilon;
    bool has_cpml = false;
    std::vector<int32_t> local_cell2rank_node;    // after the merge: global node ids
    std::vector<int32_t> local_cell2global_node;  // cleared by the merge (moved into the above)
    int n_rank_node = 0;
    int n_exchange_neighbors = 0;  // exchange groups present in the file (unused here)
    RecordingMap recording;
};

// Which optional arrays to load.  The solver needs neither coordinates nor
// velocities nor the unused beta / strain C-PML coefficients
// (forward/share/src/io.cpp:234,236; pml.cpp:52); full = true reads every
// array the reference reads.
struct ReadOpts {
    bool full = true;
    bool profiles = false;  // the C-PML stretch profiles even when !full
};

Partition read_partition(const std::string& path, const ReadOpts& o = {});
// Concatenates every partition_{r}.h5 (r = 0, 1, ... while the file exists):
// element arrays appended, ibool := local_cell2global_node, recording nodes
// de-duplicated by global id, rec_cell_local offset by the elements before
// it, n_rank_node = max id + 1 (forward/share/src/io.cpp:319-450).
Partition read_partition_all(const std::string& dir, const ReadOpts& o = {});

// config.h5 (ConfigData, forward/share/include/gf/types.hpp:113-140).
struct Config {
    std::string title = "untitled";
    int polynomial_order = 0;
    double solver_dt = 0.0, output_dt_s = 0.0;
    int snapshot_stride = 1, nsteps = 0;
    double cfl_safety = 1.0;
    std::string snapshot_precision = "float64";
    double xmin = 0, xmax = 0, ymin = 0, ymax = 0, zmin = 0, zmax = 0;
    double record_depth_max_m = 0.0, record_depth_actual_m = 0.0;
    int nx_elements = 0, ny_elements = 0, nz_elements = 0;
    int pml_xmin = 0, pml_xmax = 0, pml_ymin = 0, pml_ymax = 0, pml_zmin = 0, pml_zmax = 0;
    std::vector<int> tilex_elements, tiley_elements;
    int log_stride = 1;
    double restart_dt_s = 0.0;
    int restart_stride = 0;
    std::vector<double> stf_t, stf_values;
    double source_x = 0, source_y = 0, source_z = 0;
    std::vector<int64_t> src_cell_ids;  // 0-based global element ids
    std::vector<double> src_weights;    // [n_src_cell*ngll^3]
    int n_src_cell = 0;
};
Config read_config(const std::string& path);

// One file per (rank, step): wavefields/<dir>/record_<rank>_<step>.h5 with the
// schema of forward/share/src/record.cpp:175-257 (root attributes, the
// recording map, and [1, n_rec_cell, ngll^3, {6,3}] fields).
class RecordWriter {
public:
    RecordWriter(std::string out_dir, std::string direction, int rank, const RecordingMap& map, int ngll,
                 bool use_float32, double depth_max, double depth_actual);
    // Any field pointer may be null (not written).  No-op without recording cells.
    void write_step(int step, const double* strain, const double* disp, const double* vel,
                    const double* acc) const;
    std::string path_of(int step) const;
    int n_rec_cell() const { return int(map_.rec_cell_local.size()); }

private:
    std::string dir_, direction_;
    int rank_, ngll_;
    bool f32_;
    double depth_max_, depth_actual_;
    RecordingMap map_;
};

// Restart file <dir>/restart_<rank>.h5 (forward/share/src/restart.cpp:82-152:
// attributes rank, source_direction, step, time_s, ngll, use_global_dof,
// n_rank_node; FP64 displacement / velocity / acceleration [n_rank_node*3],
// pml_damping [n_rank_node]).  This build adds /gfx_checkpoint (the solver's
// complete device state, C-PML memory included, as opaque bytes) so that a
// resumed run continues bit for bit.
struct RestartState {
    int step = -1;
    double time_s = 0.0;
    int ngll = 0, n_rank_node = 0, use_global_dof = 1;
    std::string direction;
    std::vector<double> displacement, velocity, acceleration, pml_damping;
    std::vector<uint8_t> checkpoint;  // empty when absent
};
void write_restart(const std::string& dir, int rank, const RestartState& st);
// Throws gfx::Error(GFX_ERR_IO) when the file is missing or unreadable.
RestartState read_restart(const std::string& dir, int rank);

// ---- input writers (the files the reference preprocessor produces) ----
struct SourceSpec {
    double x = 0, y = 0, z = 0;
    std::vector<int64_t> cell_ids;  // global element ids
    std::vector<double> weights;    // [n*ngll^3]
};
// Splits the mesh into n_parts contiguous element ranges and writes
// <dir>/partitions/partition_<r>.h5 (gzip-4 field datasets), plus
// <dir>/config.h5 with the source, STF and time stepping.
void write_inputs(const gfx_mesh& m, const std::string& dir, int n_parts, const SourceSpec& src,
                  const std::vector<double>& stf, const std::string& precision, int restart_stride,
                  int log_stride);
// Stations file for batched runs: /stations attrs n_stations; per station
// /stations/s<k> attrs x, y, z with cells/cell_ids and cells/weights.
void write_stations(const std::string& path, const std::vector<SourceSpec>& st, int ngll);
std::vector<SourceSpec> read_stations(const std::string& path);

}  // namespace gfx::io
