/* gfx_io.h -- file-level entry points: the forward driver over the
 * reference's on-disk formats, and input / inspection helpers.
 *
 * HDF5 is bound at run time (libhdf5.so.103, or the library named by
 * $GFX_HDF5_LIB); without it every function here fails with GFX_ERR_IO and
 * the in-memory API of gfx.h is unaffected.
 */
#ifndef GFX_IO_H
#define GFX_IO_H

#include "gfx.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Drop-in for int gf::run_forward(const std::string& direction,
 * bool resume_mode = false, int effective_nprocs = 0)
 * (forward/share/include/gf/solver.hpp:21; forward/share/src/solver.cpp:73-739).
 * Reads config.h5 and every partitions/partition_{r}.h5 of the current
 * directory (merged as read_partition_all does), advances the elastic step on
 * the GPU and writes wavefields/{direction}/record_0_{step}.h5 every snapshot
 * and wavefields/{direction}/restart_0.h5 every restart stride; log lines go
 * to log/forward_{direction}_0.log.  Returns 0 on success and 1 on any error,
 * as the reference does (solver.cpp:733-736); the message is also in
 * gfx_last_error().  effective_nprocs is accepted for signature compatibility
 * and ignored: one process drives one GPU with the whole mesh. */
int gfx_run_forward(const char* direction, int resume_mode, int effective_nprocs);

typedef struct {
    const char* workdir;    /* NULL or "" = current directory */
    const char* directions; /* comma list of x, y, z ("z", "x,y,z"); one field per direction */
    const char* stations;   /* NULL = the config.h5 source; else a stations file: fields =
                               stations x directions, records under wavefields/s{k}/{dir}/ */
    int resume;             /* continue from the restart files (--resume) */
    int record_uva;         /* also write displacement, velocity, acceleration (default 1) */
    int assembly;           /* -1 = autotune (default), else gfx_assembly */
    int element_kernel;     /* gfx_element_kernel (ignored when autotuning) */
    int elems_per_block;    /* 0 = default (ignored when autotuning) */
    int ring;               /* snapshots the device holds between host drains (0 = 8) */
    int stop_after_step;    /* < 0: run to nsteps; else stop after this step index */
    int restart_stride;     /* < 0: from config.h5; 0: no restart files; > 0: this stride */
    int verbose;            /* echo log lines to stdout */
    int cpml_complete;      /* the complete C-PML (gfx_solver_desc.cpml_complete) from the
                               partitions' /field/cell/pml_K, pml_d, pml_alpha (--cpml-complete) */
} gfx_forward_options;

gfx_status gfx_forward_options_default(gfx_forward_options* o);
/* Batched / configurable forward run; same files as gfx_run_forward. */
gfx_status gfx_run_forward_ex(const gfx_forward_options* o);

/* Green's function tiles straight from the device: the reference's three
 * forward runs (x, y, z) plus its post-processor
 * (postprocess/cpp/main.cpp:446-875, writer.hh:105-310) in one pass.  The
 * three force directions run as three fields of one solver over config.h5 +
 * partitions/ in workdir; every snapshot is merged to the recorded GLL nodes
 * (mass-weighted strain, displacement scaled the same way) and laid out per
 * tile on the GPU (gfx_green_slab_capture), then appended to
 * output_dir/tile_x{tx:03d}_y{ty:03d}.h5, whose attributes, datasets, types,
 * chunking and filters follow writer.hh.  Each tile holds the values the
 * post-processor computes from the record files a gfx_run_forward over the
 * same inputs writes.  No record or restart file is written. */
typedef struct {
    const char* workdir;    /* config.h5 + partitions/; NULL or "" = current directory */
    const char* output_dir; /* NULL or "" = workdir/greenfun (the post-processor's default -o) */
    const char* model_path; /* NULL: bounds from config.h5 /domain, mass from the partitions;
                               else model.h5 /domain and /field/cell/mass (main.cpp:458-483) */
    int with_uva;           /* displacement_tensor and the velocity / acceleration tensors, which the
                               post-processor leaves at zero (main.cpp:401); default 1, as the
                               forward run records u, v, a by default */
    int assembly;           /* -1 = autotune (default), else gfx_assembly */
    int element_kernel;     /* gfx_element_kernel (ignored when autotuning) */
    int elems_per_block;    /* 0 = default (ignored when autotuning) */
    int ring;               /* snapshots the device holds between drains (0 = 8) */
    int verbose;            /* echo log lines (log/forward_greens_0.log) to stdout */
    int cpml_complete;      /* the complete C-PML, as gfx_forward_options.cpml_complete */
} gfx_green_tile_options;

gfx_status gfx_green_tile_options_default(gfx_green_tile_options* o);
/* *n_tiles (may be NULL) receives the number of tile files written. */
gfx_status gfx_run_green_tiles(const gfx_green_tile_options* o, int* n_tiles);

/* ---- General hexahedral meshes ------------------------------------------
 * Drop-in for tools/gmsh_to_hdf5.py (`gmsh_to_hdf5.py input.msh -o model.h5
 * [--aux aux.h5]`): reads the 8-node hexahedra of an ASCII GMSH file (format
 * 4.1 or 2.2; points in file order, as meshio keeps them) and writes
 * model.h5 /topology (vertex_to_coord, edge_to_vertex, surface_to_edge,
 * cell_to_surface with signed 1-based ids; gmsh_to_hdf5.py:45-170) and,
 * when aux_path is given, the CSR adjacency file (gmsh_to_hdf5.py:190-268). */
gfx_status gfx_gmsh_to_topology(const char* msh_path, const char* model_path, const char* aux_path);
/* Inputs for a general hexahedral mesh: the reference preprocessor's pipeline
 * (preprocess/cli.py:396-619) on model.h5 /topology, with the spec's material
 * layers and random field, C-PML thicknesses (in elements), f0, CFL safety,
 * output dt, duration and recording depth.  Cell corners come from the
 * topology (gll_geometry.py:70-129), the domain bounds from the vertices, the
 * node numbering from the quantised coordinate sort of partition.py:192-254,
 * and the C-PML flags from the cell index with nx = spec->nx, ny = spec->ny and
 * nz = n_cell / (nx ny) (cli.py:241-264: cells in x-fastest structured order);
 * the spec's box lengths, nz and jitter are not used.  Write the result with
 * gfx_mesh_write_inputs. */
gfx_status gfx_mesh_build_topology(const char* model_path, const gfx_box_spec* spec, gfx_mesh** out);
/* The same from points [n_point*3] and hexahedra [n_cell*8] (0-based point
 * indices, GMSH corner order), through the same topology extraction. */
gfx_status gfx_mesh_build_hex(int64_t n_point, const double* xyz, int64_t n_cell, const int64_t* hex,
                              const gfx_box_spec* spec, gfx_mesh** out);
/* counts[4] = n_vertex, n_edge, n_surface, n_cell of model.h5 /topology; when
 * cell_corners is given, [n_cell*8] 1-based vertex ids of each cell in GMSH
 * corner order. */
gfx_status gfx_topology_summary(const char* model_path, int64_t* counts, int64_t* cell_corners,
                                int64_t n_cell_corners);

/* Write the inputs the reference preprocessor would produce for a generated
 * box (preprocess/model_writer.py:304-446, preprocess/config_writer.py:57-143
 * schemas): dir/partitions/partition_{r}.h5 for n_parts contiguous element
 * ranges and dir/config.h5 with a point force located at (x, y, z), the STF
 * samples stf[n_stf], the snapshot precision ("float32" / "float64") and a
 * restart stride (0 = none). */
gfx_status gfx_mesh_write_inputs(const gfx_mesh* m, const char* dir, int n_parts, double x, double y,
                                 double z, int n_stf, const double* stf, const char* precision,
                                 int restart_stride);
/* Load inputs in the reference's format (dir/config.h5 and every
 * dir/partitions/partition_{r}.h5, merged as read_partition_all does) into a
 * mesh handle for the in-memory API; element index = position in the merged
 * partitions, ibool = local_cell2global_node.  Free with gfx_mesh_free. */
gfx_status gfx_mesh_read_inputs(const char* dir, gfx_mesh** out);
/* Write one snapshot record file out_dir/direction/record_0_{step}.h5 for the
 * recording map and snapshot precision of the inputs in inputs_dir
 * (RecordWriter::write_step, forward/share/src/record.cpp:175-257); arrays as
 * gfx_solver_get_record returns them, any may be NULL. */
gfx_status gfx_record_write_step(const char* inputs_dir, const char* out_dir, const char* direction, int step,
                                 const double* strain, const double* disp, const double* vel, const double* acc);
/* Stations file for batched runs: each of the n points xyz[n*3] is located
 * in the mesh as a point force (cells + normalised weights). */
gfx_status gfx_write_stations(const char* path, const gfx_mesh* m, int n, const double* xyz);

/* Inspection: 1 when an HDF5 library could be bound. */
int gfx_h5_available(void);
/* Text listing of a file's datasets and attributes ("path type [dims]" lines);
 * *needed receives the length + 1; buf may be NULL. */
gfx_status gfx_h5_describe(const char* path, char* buf, size_t len, size_t* needed);
/* Read a whole dataset converted to FP64; out may be NULL to query *count. */
gfx_status gfx_h5_read_f64(const char* path, const char* dataset, double* out, size_t n, size_t* count);
/* Read a numeric or string attribute of a group or dataset ("/" = root). */
gfx_status gfx_h5_read_attr_f64(const char* path, const char* object, const char* name, double* out);
/* Every element of a numeric attribute; out may be NULL to query *count. */
gfx_status gfx_h5_read_attr_f64_array(const char* path, const char* object, const char* name, double* out,
                                      size_t n, size_t* count);
gfx_status gfx_h5_read_attr_str(const char* path, const char* object, const char* name, char* buf,
                                size_t len);

/* Generic writers (test fixtures, tools): a dataset of type "f64", "f32",
 * "i32" or "i64" with parent groups created (append = 0 truncates the file),
 * and a scalar attribute ("f64", "i32", "i64" or "str" with a C string value)
 * on a group or dataset, the group created when missing. */
gfx_status gfx_h5_write(const char* path, int append, const char* dataset, const char* type, int ndims,
                        const uint64_t* dims, const void* data);
gfx_status gfx_h5_write_attr(const char* path, const char* object, const char* name, const char* type,
                             const void* value);

/* What the config reader (io.cpp:577-709 rules, defaults and legacy layout)
 * makes of a file: out[24] = polynomial_order, solver_dt, output_dt_s,
 * snapshot_stride, nsteps, cfl_safety, xmin, xmax, ymin, ymax, zmin, zmax,
 * len(stf_t), len(stf_values), source x, y, z, n_src_cell, record_depth_max_m,
 * record_depth_actual_m, restart_stride, restart_dt_s, log_stride,
 * len(tilex_elements); title and precision strings into buffers of len. */
gfx_status gfx_config_summary(const char* path, double* out, int n, char* title, char* precision, size_t len);
/* What the partition reader makes of one file (merge_dir = 0) or of a
 * directory merged as read_partition_all (merge_dir = 1): out[12] =
 * n_local_cell, ngll, n_rank_node, has_cpml, has_recording, n_rec_cell,
 * n_unique_gll, len(local_cell_ids), exchange neighbours, len(jacobian),
 * len(local_cell2rank_node), len(local_cell2global_node). */
gfx_status gfx_partition_summary(const char* path, int merge_dir, double* out, int n);

#ifdef __cplusplus
}
#endif
#endif
