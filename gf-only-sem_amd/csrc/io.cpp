// io.cpp -- reference file formats (see io.hpp), on the run-time HDF5 binding.
#include "io.hpp"
#include "mesh.hpp"

#include <algorithm>
#include <filesystem>
#include <unordered_map>

#include "common.hpp"
#include "h5.hpp"

namespace gfx::io {

using h5::hsize_t;
using h5::T;

namespace {

constexpr int DEFLATE = 4;  // partition field datasets (preprocess/model_writer.py:323,462-464)

int ngll_from(h5::hid_t f) {
    // NGLL = shape[1] of /field/cell/coords [E, n, n, n, 3] (io.cpp:201-216);
    // other per-point arrays carry the same axis
    for (const char* name : {"/field/cell/coords", "/field/cell/jacobian", "/field/cell/mass"}) {
        if (!h5::exists(f, name)) continue;
        const auto d = h5::shape(f, name);
        if (d.size() >= 4) return int(d[1]);
    }
    return 5;
}

}  // namespace

Partition read_partition(const std::string& path, const ReadOpts& o) {
    h5::Obj f = h5::open_read(path);
    const h5::hid_t id = f.id();
    Partition p;
    p.local_cell_ids = h5::read<int64_t>(id, "/partition/local_cell_ids");
    p.n_local_cell = int(p.local_cell_ids.size());
    p.ngll = ngll_from(id);
    p.coords = h5::try_read<double>(id, "/field/cell/coords");
    p.jacobian = h5::try_read<double>(id, "/field/cell/jacobian");
    p.dxi_dx = h5::try_read<double>(id, "/field/cell/dxi_dx");
    p.mass = h5::try_read<double>(id, "/field/cell/mass");
    if (o.full) {
        p.vp = h5::try_read<double>(id, "/field/cell/vp");
        p.vs = h5::try_read<double>(id, "/field/cell/vs");
    }
    p.density = h5::try_read<double>(id, "/field/cell/density");
    p.lambda_ = h5::try_read<double>(id, "/field/cell/lambda");
    p.mu_ = h5::try_read<double>(id, "/field/cell/mu");
    p.damping = h5::try_read<double>(id, "/field/cell/damping");
    p.pml_region = h5::try_read<int32_t>(id, "/field/cell/pml_region");
    p.coef_alpha = h5::try_read<double>(id, "/field/cell/pml_coef_alpha");
    if (o.full) {
        p.coef_beta = h5::try_read<double>(id, "/field/cell/pml_coef_beta");
        p.coef_strain = h5::try_read<double>(id, "/field/cell/pml_coef_strain");
    }
    p.coef_abar = h5::try_read<double>(id, "/field/cell/pml_coef_abar");
    if (o.full || o.profiles) {
        p.pml_K = h5::try_read<double>(id, "/field/cell/pml_K");
        p.pml_d = h5::try_read<double>(id, "/field/cell/pml_d");
        p.pml_alpha = h5::try_read<double>(id, "/field/cell/pml_alpha");
    }
    p.has_cpml = !p.coef_abar.empty();  // io.cpp:237
    if (h5::exists(id, "/field/cell/tau_sigma")) {
        std::vector<h5::hsize_t> dims;
        p.tau_sigma = h5::read<double>(id, "/field/cell/tau_sigma", &dims);
        if (dims.size() != 5) throw Error(GFX_ERR_IO, path + ": /field/cell/tau_sigma must be [n_cell, n, n, n, n_sls]");
        p.n_sls = int(dims[4]);
        p.tau_epsilon = h5::read<double>(id, "/field/cell/tau_epsilon");
        if (p.tau_epsilon.size() != p.tau_sigma.size())
            throw Error(GFX_ERR_IO, path + ": tau_epsilon differs in size from tau_sigma");
    }
    p.local_cell2rank_node = h5::try_read<int32_t>(id, "/field/cell/local_cell2rank_node");
    p.local_cell2global_node = h5::try_read<int32_t>(id, "/field/cell/local_cell2global_node");
    if (h5::exists(id, "/field/cell")) {
        h5::Obj g = h5::open_group(id, "/field/cell");
        h5::read_attr_int(g.id(), "n_rank_node", p.n_rank_node);
    }
    if (h5::exists(id, "/partition/exchange")) p.n_exchange_neighbors = int(h5::list(id, "/partition/exchange").size());
    if (h5::exists(id, "/recording")) {
        RecordingMap& r = p.recording;
        r.has_recording = true;
        r.gll_node_ids = h5::read<int64_t>(id, "/recording/gll_node_ids");
        r.gll_node_coords = h5::read<double>(id, "/recording/gll_node_coords");
        r.rec_cell_local = h5::read<int32_t>(id, "/recording/rec_cell_local");
        r.cell_gll_node_index = h5::read<int32_t>(id, "/recording/cell_gll_node_index");
        r.rec_cell_global = h5::read<int64_t>(id, "/recording/rec_cell_global_ids");
        h5::Obj g = h5::open_group(id, "/recording");
        h5::read_attr_double(g.id(), "record_depth_max_m", r.record_depth_max_m);
        h5::read_attr_double(g.id(), "record_depth_actual_m", r.record_depth_actual_m);
    }
    return p;
}

Partition read_partition_all(const std::string& dir, const ReadOpts& o) {
    int n_parts = 0;
    while (std::filesystem::exists(dir + "/partition_" + std::to_string(n_parts) + ".h5")) ++n_parts;
    if (n_parts == 0) throw Error(GFX_ERR_IO, "read_partition_all: no partition files found in " + dir);

    auto cat = [](auto& dst, const auto& src) { dst.insert(dst.end(), src.begin(), src.end()); };
    Partition m;
    int cumulative = 0;
    std::unordered_map<int64_t, int32_t> rec_index;  // global node id -> merged recording index
    for (int r = 0; r < n_parts; ++r) {
        Partition part = read_partition(dir + "/partition_" + std::to_string(r) + ".h5", o);
        if (r > 0 && part.ngll != m.ngll)
            throw Error(GFX_ERR_IO, "read_partition_all: partition " + std::to_string(r) + " has NGLL=" +
                                        std::to_string(part.ngll) + " but expected " + std::to_string(m.ngll));
        if (r == 0) {
            std::vector<int32_t> gib = std::move(part.local_cell2global_node);
            m = std::move(part);
            m.local_cell2rank_node = std::move(gib);
            m.local_cell2global_node.clear();
            cumulative = m.n_local_cell;
            if (m.recording.has_recording)
                for (size_t i = 0; i < m.recording.gll_node_ids.size(); ++i)
                    rec_index[m.recording.gll_node_ids[i]] = int32_t(i);
            continue;
        }
        cat(m.local_cell_ids, part.local_cell_ids);
        cat(m.coords, part.coords);
        cat(m.jacobian, part.jacobian);
        cat(m.dxi_dx, part.dxi_dx);
        cat(m.mass, part.mass);
        cat(m.vp, part.vp);
        cat(m.vs, part.vs);
        cat(m.density, part.density);
        cat(m.lambda_, part.lambda_);
        cat(m.mu_, part.mu_);
        cat(m.damping, part.damping);
        cat(m.pml_region, part.pml_region);
        cat(m.coef_alpha, part.coef_alpha);
        cat(m.coef_beta, part.coef_beta);
        cat(m.coef_abar, part.coef_abar);
        cat(m.coef_strain, part.coef_strain);
        cat(m.pml_K, part.pml_K);
        cat(m.pml_d, part.pml_d);
        cat(m.pml_alpha, part.pml_alpha);
        if (part.n_sls != m.n_sls) throw Error(GFX_ERR_IO, "partitions differ in their number of attenuation solids");
        cat(m.tau_sigma, part.tau_sigma);
        cat(m.tau_epsilon, part.tau_epsilon);
        m.has_cpml = m.has_cpml || part.has_cpml;
        cat(m.local_cell2rank_node, part.local_cell2global_node);
        if (part.recording.has_recording) {
            RecordingMap& mr = m.recording;
            const RecordingMap& pr = part.recording;
            mr.has_recording = true;
            std::vector<int32_t> to_merged(pr.gll_node_ids.size());
            for (size_t i = 0; i < pr.gll_node_ids.size(); ++i) {
                const int64_t gid = pr.gll_node_ids[i];
                auto it = rec_index.find(gid);
                if (it == rec_index.end()) {
                    const int32_t k = int32_t(mr.gll_node_ids.size());
                    rec_index[gid] = k;
                    mr.gll_node_ids.push_back(gid);
                    for (int d = 0; d < 3; ++d) mr.gll_node_coords.push_back(pr.gll_node_coords[i * 3 + d]);
                    to_merged[i] = k;
                } else {
                    to_merged[i] = it->second;
                }
            }
            cat(mr.rec_cell_global, pr.rec_cell_global);
            for (int32_t c : pr.rec_cell_local) mr.rec_cell_local.push_back(c + cumulative);
            for (int32_t k : pr.cell_gll_node_index) mr.cell_gll_node_index.push_back(to_merged[k]);
            if (mr.record_depth_actual_m == 0.0) {
                mr.record_depth_max_m = pr.record_depth_max_m;
                mr.record_depth_actual_m = pr.record_depth_actual_m;
            }
        }
        cumulative += part.n_local_cell;
    }
    m.n_local_cell = cumulative;
    m.n_exchange_neighbors = 0;
    if (!m.local_cell2rank_node.empty())
        m.n_rank_node = *std::max_element(m.local_cell2rank_node.begin(), m.local_cell2rank_node.end()) + 1;
    return m;
}

Config read_config(const std::string& path) {
    h5::Obj f = h5::open_read(path);
    const h5::hid_t id = f.id();
    Config c;
    if (h5::exists(id, "/simulation")) {
        h5::Obj g = h5::open_group(id, "/simulation");
        const h5::hid_t s = g.id();
        h5::read_attr_string(s, "title", c.title);
        int po = 0;
        c.polynomial_order = (h5::read_attr_int(s, "polynomial_order", po) && po > 0) ? po : 3;
        if (!h5::read_attr_double(s, "solver_dt", c.solver_dt))
            throw Error(GFX_ERR_IO, "Missing required /simulation attribute: solver_dt");
        if (!h5::read_attr_double(s, "output_dt_s", c.output_dt_s)) c.output_dt_s = c.solver_dt;
        if (!h5::read_attr_int(s, "snapshot_stride", c.snapshot_stride)) c.snapshot_stride = 1;
        if (!h5::read_attr_int(s, "nsteps", c.nsteps))
            throw Error(GFX_ERR_IO, "Missing required /simulation attribute: nsteps");
        h5::read_attr_double(s, "cfl_safety", c.cfl_safety);
        h5::read_attr_string(s, "snapshot_precision", c.snapshot_precision);
        h5::read_attr_double(s, "record_depth_max_m", c.record_depth_max_m);
        h5::read_attr_double(s, "record_depth_actual_m", c.record_depth_actual_m);
        h5::read_attr_int(s, "nx_elements", c.nx_elements);
        h5::read_attr_int(s, "ny_elements", c.ny_elements);
        h5::read_attr_int(s, "nz_elements", c.nz_elements);
        h5::read_attr_int(s, "pml_xmin", c.pml_xmin);
        h5::read_attr_int(s, "pml_xmax", c.pml_xmax);
        h5::read_attr_int(s, "pml_ymin", c.pml_ymin);
        h5::read_attr_int(s, "pml_ymax", c.pml_ymax);
        h5::read_attr_int(s, "pml_zmin", c.pml_zmin);
        h5::read_attr_int(s, "pml_zmax", c.pml_zmax);
        for (auto v : h5::try_read<int64_t>(s, "tilex_elements")) c.tilex_elements.push_back(int(v));
        for (auto v : h5::try_read<int64_t>(s, "tiley_elements")) c.tiley_elements.push_back(int(v));
        h5::read_attr_double(s, "restart_dt_s", c.restart_dt_s);
        h5::read_attr_int(s, "restart_stride", c.restart_stride);
        h5::read_attr_int(s, "log_stride", c.log_stride);
    } else {
        // legacy flat datasets (io.cpp:631-654)
        auto po = h5::try_read<double>(id, "polynomial_order");
        c.polynomial_order = po.empty() ? 3 : int(po[0]);
        auto dt = h5::try_read<double>(id, "dt");
        c.solver_dt = dt.empty() ? 0.005 : dt[0];
        c.output_dt_s = c.solver_dt;
        auto ns = h5::try_read<double>(id, "nsteps");
        c.nsteps = ns.empty() ? 1000 : int(ns[0]);
        auto cfl = h5::try_read<double>(id, "cfl_safety");
        c.cfl_safety = cfl.empty() ? 1.0 : cfl[0];
        c.snapshot_stride = 1;
        auto f32 = h5::try_read<int64_t>(id, "use_float32");
        if (!f32.empty() && f32[0] == 1) c.snapshot_precision = "float32";
    }
    if (h5::exists(id, "/domain")) {
        h5::Obj g = h5::open_group(id, "/domain");
        h5::read_attr_double(g.id(), "xmin", c.xmin);
        h5::read_attr_double(g.id(), "xmax", c.xmax);
        h5::read_attr_double(g.id(), "ymin", c.ymin);
        h5::read_attr_double(g.id(), "ymax", c.ymax);
        h5::read_attr_double(g.id(), "zmin", c.zmin);
        h5::read_attr_double(g.id(), "zmax", c.zmax);
    } else {
        const char* names[6] = {"xmin", "xmax", "ymin", "ymax", "zmin", "zmax"};
        double* dst[6] = {&c.xmin, &c.xmax, &c.ymin, &c.ymax, &c.zmin, &c.zmax};
        for (int i = 0; i < 6; ++i) {
            auto v = h5::try_read<double>(id, names[i]);
            if (!v.empty()) *dst[i] = v[0];
        }
    }
    if (h5::exists(id, "/source")) {
        h5::Obj g = h5::open_group(id, "/source");
        c.stf_t = h5::try_read<double>(id, "/source/stf_t");
        c.stf_values = h5::try_read<double>(id, "/source/stf_values");
        h5::read_attr_double(g.id(), "x", c.source_x);
        h5::read_attr_double(g.id(), "y", c.source_y);
        h5::read_attr_double(g.id(), "z", c.source_z);
    } else {
        c.stf_t = h5::try_read<double>(id, "stf_t");
        c.stf_values = h5::try_read<double>(id, "stf_values");
        const char* names[3] = {"source_x", "source_y", "source_z"};
        double* dst[3] = {&c.source_x, &c.source_y, &c.source_z};
        for (int i = 0; i < 3; ++i) {
            auto v = h5::try_read<double>(id, names[i]);
            if (!v.empty()) *dst[i] = v[0];
        }
    }
    c.src_cell_ids = h5::try_read<int64_t>(id, "/source/cells/cell_ids");
    c.src_weights = h5::try_read<double>(id, "/source/cells/weights");
    c.n_src_cell = int(c.src_cell_ids.size());
    return c;
}

RecordWriter::RecordWriter(std::string out_dir, std::string direction, int rank, const RecordingMap& map,
                           int ngll, bool use_float32, double depth_max, double depth_actual)
    : dir_(std::move(out_dir)),
      direction_(std::move(direction)),
      rank_(rank),
      ngll_(ngll),
      f32_(use_float32),
      depth_max_(depth_max),
      depth_actual_(depth_actual),
      map_(map) {}

std::string RecordWriter::path_of(int step) const {
    return dir_ + "/" + direction_ + "/record_" + std::to_string(rank_) + "_" + std::to_string(step) + ".h5";
}

void RecordWriter::write_step(int step, const double* strain, const double* disp, const double* vel,
                              const double* acc) const {
    const hsize_t n_rec = map_.rec_cell_local.size();
    if (n_rec == 0) return;  // record.cpp:178-179
    std::filesystem::create_directories(dir_ + "/" + direction_);
    h5::Obj f = h5::create_trunc(path_of(step));
    const h5::hid_t id = f.id();
    const int n_node = ngll_ * ngll_ * ngll_;
    const int n_rec_i = int(n_rec), n_unique_i = int(map_.gll_node_ids.size());
    const bool excludes_pml = map_.has_recording;
    h5::write_attr_string(id, "source_direction", direction_);
    h5::write_attr(id, "rank", T::I32, &rank_);
    h5::write_attr(id, "ngll", T::I32, &ngll_);
    h5::write_attr_string(id, "basis", "gll");
    h5::write_attr(id, "n_rec_cell", T::I32, &n_rec_i);
    h5::write_attr(id, "n_unique_gll", T::I32, &n_unique_i);
    h5::write_attr(id, "excludes_pml", T::HBOOL, &excludes_pml);
    h5::write_attr(id, "record_depth_max_m", T::F64, &depth_max_);
    h5::write_attr(id, "record_depth_actual_m", T::F64, &depth_actual_);
    if (!map_.gll_node_ids.empty()) {
        h5::write_vec(id, "gll_node_ids", map_.gll_node_ids);
        h5::write_vec(id, "gll_node_coords", map_.gll_node_coords, {hsize_t(map_.gll_node_ids.size()), 3});
    }
    h5::write_vec(id, "cell_gll_node_index", map_.cell_gll_node_index, {n_rec, hsize_t(n_node)});
    if (!map_.rec_cell_global.empty()) h5::write_vec(id, "recording_cell_model_index", map_.rec_cell_global);
    h5::WriteOpts whole;
    whole.chunk_whole = true;
    const T ft = f32_ ? T::F32 : T::F64;
    struct Fld {
        const char* name;
        int nc;
        const double* p;
    } flds[4] = {{"strain", 6, strain}, {"displacement", 3, disp}, {"velocity", 3, vel}, {"acceleration", 3, acc}};
    for (const Fld& fl : flds)
        if (fl.p) h5::write(id, fl.name, ft, T::F64, fl.p, {1, n_rec, hsize_t(n_node), hsize_t(fl.nc)}, whole);
}

void write_restart(const std::string& dir, int rank, const RestartState& st) {
    std::filesystem::create_directories(dir);
    const std::string path = dir + "/restart_" + std::to_string(rank) + ".h5";
    const std::string tmp = path + ".tmp";
    {
        // written beside the old file and renamed, so a run killed while
        // writing still leaves the previous restart whole
        h5::Obj f = h5::create_trunc(tmp);
        const h5::hid_t id = f.id();
        h5::write_attr_string(id, "source_direction", st.direction);
        h5::write_attr(id, "rank", T::I32, &rank);
        h5::write_attr(id, "ngll", T::I32, &st.ngll);
        h5::write_attr(id, "use_global_dof", T::I32, &st.use_global_dof);
        h5::write_attr(id, "n_rank_node", T::I32, &st.n_rank_node);
        h5::write_attr(id, "step", T::I32, &st.step);
        h5::write_attr(id, "time_s", T::F64, &st.time_s);
        h5::write_vec(id, "displacement", st.displacement);
        h5::write_vec(id, "velocity", st.velocity);
        h5::write_vec(id, "acceleration", st.acceleration);
        h5::write_vec(id, "pml_damping", st.pml_damping);
        if (!st.checkpoint.empty()) {
            h5::WriteOpts o;
            o.deflate = 1;
            h5::write_vec(id, "gfx_checkpoint", st.checkpoint, {}, o);
        }
    }
    std::filesystem::rename(tmp, path);
}

RestartState read_restart(const std::string& dir, int rank) {
    const std::string path = dir + "/restart_" + std::to_string(rank) + ".h5";
    if (!std::filesystem::exists(path)) throw Error(GFX_ERR_IO, "no restart file: " + path);
    h5::Obj f = h5::open_read(path);
    const h5::hid_t id = f.id();
    RestartState st;
    h5::read_attr_string(id, "source_direction", st.direction);
    h5::read_attr_int(id, "ngll", st.ngll);
    h5::read_attr_int(id, "use_global_dof", st.use_global_dof);
    h5::read_attr_int(id, "n_rank_node", st.n_rank_node);
    if (!h5::read_attr_int(id, "step", st.step)) throw Error(GFX_ERR_IO, "restart file without step: " + path);
    h5::read_attr_double(id, "time_s", st.time_s);
    st.displacement = h5::try_read<double>(id, "displacement");
    st.velocity = h5::try_read<double>(id, "velocity");
    st.acceleration = h5::try_read<double>(id, "acceleration");
    st.pml_damping = h5::try_read<double>(id, "pml_damping");
    st.checkpoint = h5::try_read<uint8_t>(id, "gfx_checkpoint");
    return st;
}

void write_inputs(const gfx_mesh& m, const std::string& dir, int n_parts, const SourceSpec& src,
                  const std::vector<double>& stf, const std::string& precision, int restart_stride,
                  int log_stride) {
    require(n_parts >= 1 && n_parts <= m.n_elem, "n_parts must be in [1, n_elem]");
    require(precision == "float32" || precision == "float64", "precision is float32 or float64");
    const int E = m.n_elem, n = m.ngll, NP = n * n * n;
    std::filesystem::create_directories(dir + "/partitions");
    h5::WriteOpts gz;
    gz.deflate = DEFLATE;

    std::vector<char> is_rec(E, 0);
    for (int i = 0; i < m.n_rec_cell; ++i) is_rec[m.rec_cell[i]] = 1;

    // cells of each rank: the reference's partitioner without METIS
    // (preprocess/partition.py:75-110, 257-301): cell centroids (mean of the
    // GLL points) sorted along the longest axis of their spread, cut into
    // chunks of ceil(E / n_parts); each rank keeps its cells in increasing
    // global order.  Equal centroids keep their cell order (stable sort).
    std::vector<int> owner(E, 0);
    if (n_parts > 1) {
        std::vector<double> cen(size_t(E) * 3, 0.0);
        for (int c = 0; c < E; ++c)
            for (int q = 0; q < NP; ++q)
                for (int d = 0; d < 3; ++d) cen[size_t(c) * 3 + d] += m.node_xyz[int64_t(m.ibool[int64_t(c) * NP + q]) * 3 + d];
        for (double& v : cen) v /= NP;
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int c = 0; c < E; ++c)
            for (int d = 0; d < 3; ++d) {
                lo[d] = std::min(lo[d], cen[size_t(c) * 3 + d]);
                hi[d] = std::max(hi[d], cen[size_t(c) * 3 + d]);
            }
        int ax = 0;
        for (int d = 1; d < 3; ++d)
            if (hi[d] - lo[d] > hi[ax] - lo[ax]) ax = d;
        std::vector<int> order(E);
        for (int c = 0; c < E; ++c) order[c] = c;
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int b) { return cen[size_t(a) * 3 + ax] < cen[size_t(b) * 3 + ax]; });
        const int chunk = (E + n_parts - 1) / n_parts;
        for (int i = 0; i < E; ++i) owner[order[i]] = std::min(i / chunk, n_parts - 1);
    }

    for (int r = 0; r < n_parts; ++r) {
        std::vector<int64_t> ids;
        for (int c = 0; c < E; ++c)
            if (owner[c] == r) ids.push_back(c);
        require(!ids.empty(), "n_parts = " + std::to_string(n_parts) + " leaves partition " + std::to_string(r) +
                                  " without cells");
        const int ne = int(ids.size());
        const hsize_t hn = hsize_t(ne), hg = hsize_t(n);
        const std::string path = dir + "/partitions/partition_" + std::to_string(r) + ".h5";
        h5::Obj f = h5::create_trunc(path);
        const h5::hid_t id = f.id();
        auto slice = [&](const double* a, int per) {
            std::vector<double> out;
            out.reserve(size_t(ne) * NP * per);
            for (int64_t c : ids) out.insert(out.end(), a + c * NP * per, a + (c + 1) * NP * per);
            return out;
        };
        h5::write_vec(id, "/partition/local_cell_ids", ids);
        {
            h5::Obj g = h5::open_group(id, "/partition");
            h5::write_attr(g.id(), "n_ranks", T::I32, &n_parts);
        }
        if (m.coords) h5::write_vec(id, "/field/cell/coords", slice(m.coords, 3), {hn, hg, hg, hg, 3}, gz);
        h5::write_vec(id, "/field/cell/jacobian", slice(m.jacobian, 1), {hn, hg, hg, hg}, gz);
        h5::write_vec(id, "/field/cell/dxi_dx", slice(m.dxi_dx, 9), {hn, hg, hg, hg, 9}, gz);
        h5::write_vec(id, "/field/cell/mass", slice(m.mass, 1), {hn, hg, hg, hg}, gz);
        if (m.vp) h5::write_vec(id, "/field/cell/vp", slice(m.vp, 1), {hn, hg, hg, hg}, gz);
        if (m.vs) h5::write_vec(id, "/field/cell/vs", slice(m.vs, 1), {hn, hg, hg, hg}, gz);
        h5::write_vec(id, "/field/cell/density", slice(m.density, 1), {hn, hg, hg, hg}, gz);
        h5::write_vec(id, "/field/cell/lambda", slice(m.lambda_, 1), {hn, hg, hg, hg}, gz);
        h5::write_vec(id, "/field/cell/mu", slice(m.mu_, 1), {hn, hg, hg, hg}, gz);
        h5::write_vec(id, "/field/cell/damping", slice(m.damping, 1), {hn, hg, hg, hg}, gz);
        if (m.pml_region) {
            std::vector<int32_t> reg;
            for (int64_t c : ids) reg.push_back(m.pml_region[c]);
            h5::write_vec(id, "/field/cell/pml_region", reg, {}, gz);
            h5::write_vec(id, "/field/cell/pml_coef_alpha", slice(m.coef_alpha, 9), {hn, hsize_t(NP), 9}, gz);
            if (m.coef_beta)
                h5::write_vec(id, "/field/cell/pml_coef_beta", slice(m.coef_beta, 9), {hn, hsize_t(NP), 9}, gz);
            if (m.coef_strain)
                h5::write_vec(id, "/field/cell/pml_coef_strain", slice(m.coef_strain, 18), {hn, hsize_t(NP), 18},
                              gz);
            h5::write_vec(id, "/field/cell/pml_coef_abar", slice(m.coef_abar, 5), {hn, hsize_t(NP), 5}, gz);
        }
        if (m.pml_K && m.pml_d && m.pml_alpha) {  // pml_cpml.py compute_pml_profiles output
            h5::write_vec(id, "/field/cell/pml_K", slice(m.pml_K, 3), {hn, hsize_t(NP), 3}, gz);
            h5::write_vec(id, "/field/cell/pml_d", slice(m.pml_d, 3), {hn, hsize_t(NP), 3}, gz);
            h5::write_vec(id, "/field/cell/pml_alpha", slice(m.pml_alpha, 3), {hn, hsize_t(NP), 3}, gz);
        }
        if (m.n_sls > 0 && m.tau_sigma && m.tau_epsilon) {
            const hsize_t L = hsize_t(m.n_sls);
            h5::write_vec(id, "/field/cell/tau_sigma", slice(m.tau_sigma, m.n_sls), {hn, hg, hg, hg, L}, gz);
            h5::write_vec(id, "/field/cell/tau_epsilon", slice(m.tau_epsilon, m.n_sls), {hn, hg, hg, hg, L}, gz);
        }
        // node numbering: global ids, and a compact per-partition numbering
        std::vector<int32_t> gib;
        gib.reserve(size_t(ne) * NP);
        for (int64_t c : ids) gib.insert(gib.end(), m.ibool + c * NP, m.ibool + (c + 1) * NP);
        std::vector<int32_t> rib(gib.size());
        std::unordered_map<int32_t, int32_t> compact;
        for (size_t q = 0; q < gib.size(); ++q) {
            auto it = compact.emplace(gib[q], int32_t(compact.size())).first;
            rib[q] = it->second;
        }
        h5::write_vec(id, "/field/cell/local_cell2global_node", gib, {}, gz);
        h5::write_vec(id, "/field/cell/local_cell2rank_node", rib, {}, gz);
        {
            h5::Obj g = h5::open_group(id, "/field/cell");
            const int nrn = int(compact.size());
            h5::write_attr(g.id(), "n_rank_node", T::I32, &nrn);
        }
        // recording map of the cells in this partition (recording_map.py:52-90 rules)
        std::vector<int32_t> rloc;
        std::vector<int64_t> rglob;
        for (int i = 0; i < ne; ++i)
            if (is_rec[ids[i]]) {
                rloc.push_back(i);
                rglob.push_back(ids[i]);
            }
        if (!rloc.empty()) {
            std::vector<int64_t> nodes;
            for (int64_t c : rglob)
                for (int q = 0; q < NP; ++q) nodes.push_back(m.ibool[c * NP + q]);
            std::sort(nodes.begin(), nodes.end());
            nodes.erase(std::unique(nodes.begin(), nodes.end()), nodes.end());
            std::unordered_map<int64_t, int32_t> at;
            std::vector<double> xyz(nodes.size() * 3);
            for (size_t k = 0; k < nodes.size(); ++k) {
                at[nodes[k]] = int32_t(k);
                for (int d = 0; d < 3; ++d) xyz[k * 3 + d] = m.node_xyz[nodes[k] * 3 + d];
            }
            std::vector<int32_t> cidx;
            for (int64_t c : rglob)
                for (int q = 0; q < NP; ++q) cidx.push_back(at[m.ibool[c * NP + q]]);
            h5::write_vec(id, "/recording/gll_node_ids", nodes);
            h5::write_vec(id, "/recording/gll_node_coords", xyz, {hsize_t(nodes.size()), 3});
            h5::write_vec(id, "/recording/rec_cell_local", rloc);
            h5::write_vec(id, "/recording/rec_cell_global_ids", rglob);
            h5::write_vec(id, "/recording/cell_gll_node_index", cidx);
            h5::Obj g = h5::open_group(id, "/recording");
            h5::write_attr_string(g.id(), "basis", "gll");
            const double dmax = m.record_depth_actual, dact = m.record_depth_actual;
            h5::write_attr(g.id(), "record_depth_max_m", T::F64, &dmax);
            h5::write_attr(g.id(), "record_depth_actual_m", T::F64, &dact);
            const bool excl = true;
            h5::write_attr(g.id(), "excludes_pml", T::HBOOL, &excl);
            const int nu = int(nodes.size()), nr = int(rloc.size());
            h5::write_attr(g.id(), "n_unique_gll", T::I32, &nu);
            h5::write_attr(g.id(), "n_rec_cell", T::I32, &nr);
        }
    }

    // config.h5 (preprocess/config_writer.py:57-143 schema)
    h5::Obj f = h5::create_trunc(dir + "/config.h5");
    const h5::hid_t id = f.id();
    {
        h5::Obj g = h5::require_group(id, "/simulation");
        const h5::hid_t s = g.id();
        const int order = m.ngll - 1;
        const double out_dt = m.dt * m.snapshot_stride;
        h5::write_attr_string(s, "title", "gfx structured box");
        h5::write_attr(s, "polynomial_order", T::I32, &order);
        h5::write_attr(s, "solver_dt", T::F64, &m.dt);
        h5::write_attr(s, "output_dt_s", T::F64, &out_dt);
        h5::write_attr(s, "snapshot_stride", T::I32, &m.snapshot_stride);
        h5::write_attr(s, "nsteps", T::I32, &m.nsteps);
        h5::write_attr_string(s, "snapshot_precision", precision);
        h5::write_attr(s, "record_depth_max_m", T::F64, &m.record_depth_actual);
        h5::write_attr(s, "record_depth_actual_m", T::F64, &m.record_depth_actual);
        h5::write_attr(s, "nx_elements", T::I32, &m.nx);
        h5::write_attr(s, "ny_elements", T::I32, &m.ny);
        h5::write_attr(s, "nz_elements", T::I32, &m.nz);
        h5::write_attr(s, "log_stride", T::I32, &log_stride);
        h5::write_attr(s, "restart_stride", T::I32, &restart_stride);
        // C-PML thickness and one tile over the interior (config_loader.py:153-186:
        // the tile sizes sum to the element count between the x / y C-PML layers)
        const gfx_box_spec& sp = reinterpret_cast<const MeshHolder&>(m).spec;
        const int pml[6] = {sp.pml_xmin, sp.pml_xmax, sp.pml_ymin, sp.pml_ymax, 0, sp.pml_zmax};
        const char* pname[6] = {"pml_xmin", "pml_xmax", "pml_ymin", "pml_ymax", "pml_zmin", "pml_zmax"};
        for (int k = 0; k < 6; ++k) h5::write_attr(s, pname[k], T::I32, &pml[k]);
        h5::write_vec<int64_t>(s, "tilex_elements", {int64_t(m.nx - sp.pml_xmin - sp.pml_xmax)});
        h5::write_vec<int64_t>(s, "tiley_elements", {int64_t(m.ny - sp.pml_ymin - sp.pml_ymax)});
    }
    {
        h5::Obj g = h5::require_group(id, "/domain");
        h5::write_attr(g.id(), "xmin", T::F64, &m.xmin);
        h5::write_attr(g.id(), "xmax", T::F64, &m.xmax);
        h5::write_attr(g.id(), "ymin", T::F64, &m.ymin);
        h5::write_attr(g.id(), "ymax", T::F64, &m.ymax);
        h5::write_attr(g.id(), "zmin", T::F64, &m.zmin);
        h5::write_attr(g.id(), "zmax", T::F64, &m.zmax);
    }
    {
        std::vector<double> t(stf.size());
        for (size_t i = 0; i < t.size(); ++i) t[i] = double(i) * m.dt;
        h5::write_vec(id, "/source/stf_t", t);
        h5::write_vec(id, "/source/stf_values", stf);
        h5::Obj g = h5::open_group(id, "/source");
        h5::write_attr(g.id(), "x", T::F64, &src.x);
        h5::write_attr(g.id(), "y", T::F64, &src.y);
        h5::write_attr(g.id(), "z", T::F64, &src.z);
        const int ns = int(src.cell_ids.size());
        h5::write_attr(g.id(), "n_src_cell", T::I32, &ns);
        if (ns > 0) {
            h5::write_vec(id, "/source/cells/cell_ids", src.cell_ids);
            h5::write_vec(id, "/source/cells/weights", src.weights,
                          {hsize_t(ns), hsize_t(n), hsize_t(n), hsize_t(n)});
        }
    }
}

void write_stations(const std::string& path, const std::vector<SourceSpec>& st, int ngll) {
    h5::Obj f = h5::create_trunc(path);
    const h5::hid_t id = f.id();
    h5::Obj root = h5::require_group(id, "/stations");
    const int ns = int(st.size());
    h5::write_attr(root.id(), "n_stations", T::I32, &ns);
    const hsize_t n = hsize_t(ngll);
    for (int k = 0; k < ns; ++k) {
        const std::string base = "/stations/s" + std::to_string(k);
        h5::write_vec(id, base + "/cells/cell_ids", st[k].cell_ids);
        h5::write_vec(id, base + "/cells/weights", st[k].weights, {hsize_t(st[k].cell_ids.size()), n, n, n});
        h5::Obj g = h5::open_group(id, base);
        h5::write_attr(g.id(), "x", T::F64, &st[k].x);
        h5::write_attr(g.id(), "y", T::F64, &st[k].y);
        h5::write_attr(g.id(), "z", T::F64, &st[k].z);
    }
}

std::vector<SourceSpec> read_stations(const std::string& path) {
    h5::Obj f = h5::open_read(path);
    const h5::hid_t id = f.id();
    h5::Obj root = h5::open_group(id, "/stations");
    int ns = 0;
    if (!h5::read_attr_int(root.id(), "n_stations", ns) || ns < 1)
        throw Error(GFX_ERR_IO, "stations file without n_stations: " + path);
    std::vector<SourceSpec> out(ns);
    for (int k = 0; k < ns; ++k) {
        const std::string base = "/stations/s" + std::to_string(k);
        out[k].cell_ids = h5::read<int64_t>(id, base + "/cells/cell_ids");
        out[k].weights = h5::read<double>(id, base + "/cells/weights");
        h5::Obj g = h5::open_group(id, base);
        h5::read_attr_double(g.id(), "x", out[k].x);
        h5::read_attr_double(g.id(), "y", out[k].y);
        h5::read_attr_double(g.id(), "z", out[k].z);
    }
    return out;
}

}  // namespace gfx::io
