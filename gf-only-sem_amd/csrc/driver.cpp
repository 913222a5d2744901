// driver.cpp -- the forward run over the reference's files (gfx_io.h):
// config + partitions in, per-snapshot records and restart files out.
//
// Structure of one run (forward/share/src/solver.cpp:73-739 is the reference
// driver): read config.h5 and merge the partitions, build the device solver
// (one field per direction, or per station and direction), then advance in
// segments of at most `ring` snapshots.  After each segment the new snapshots
// are copied to host FP64 buffers and handed to a writer thread, which
// creates the HDF5 record files while the GPU runs the next segment.  Restart
// files are written at the restart stride with the reference's datasets and
// a complete solver checkpoint beside them.
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <ctime>
#include <deque>
#include <filesystem>
#include <fstream>
#include <functional>
#include <iomanip>
#include <iostream>
#include <mutex>
#include <sstream>
#include <thread>
#include <unordered_map>

#include "common.hpp"
#include "gfx.h"
#include "gfx_io.h"
#include "h5.hpp"
#include "io.hpp"
#include "mesh.hpp"

namespace fs = std::filesystem;

namespace gfx {
namespace {

// log/forward_<dir>_<rank>.log, echoed to stdout when verbose
// (forward/share/include/gf/logger.hpp:4-31).
class Logger {
public:
    Logger(const fs::path& wd, const std::string& tag, bool echo) : echo_(echo) {
        fs::create_directories(wd / "log");
        file_.open(wd / "log" / ("forward_" + tag + "_0.log"), std::ios::out | std::ios::trunc);
    }
    void info(const std::string& m) { line("INFO", m); }
    void error(const std::string& m) { line("ERROR", m); }

private:
    void line(const char* lvl, const std::string& m) {
        const std::time_t t = std::time(nullptr);
        char ts[32];
        std::strftime(ts, sizeof(ts), "%Y-%m-%d %H:%M:%S", std::localtime(&t));
        std::ostringstream os;
        os << "[" << ts << "] [" << lvl << "] " << m;
        if (file_.is_open()) file_ << os.str() << std::endl;
        if (echo_) std::cout << os.str() << std::endl;
    }
    std::ofstream file_;
    bool echo_;
};

struct Snapshot {
    int field = 0, step = 0;
    std::vector<double> strain, disp, vel, acc;
};

// Single consumer thread that writes record files in arrival order; a bounded
// queue keeps host memory to `cap` snapshots.
class WriterThread {
public:
    WriterThread(std::vector<io::RecordWriter>& w, size_t cap) : w_(w), cap_(cap), th_([this] { loop(); }) {}
    ~WriterThread() {
        try {
            finish();
        } catch (...) {
            // an error is already propagating; the write error was logged by finish() callers
        }
    }
    void push(Snapshot&& s) {
        std::unique_lock<std::mutex> lk(mu_);
        space_.wait(lk, [&] { return q_.size() < cap_ || !err_.empty(); });
        if (!err_.empty()) throw Error(GFX_ERR_IO, err_);
        q_.push_back(std::move(s));
        ready_.notify_one();
    }
    // Waits for the queue to drain; rethrows the first write error.
    void finish() {
        {
            std::unique_lock<std::mutex> lk(mu_);
            if (!done_) {
                done_ = true;
                ready_.notify_one();
            }
        }
        if (th_.joinable()) th_.join();
        if (!err_.empty()) {
            const std::string e = err_;
            err_.clear();
            throw Error(GFX_ERR_IO, e);
        }
    }
    size_t written() const { return written_; }

private:
    void loop() {
        for (;;) {
            Snapshot s;
            {
                std::unique_lock<std::mutex> lk(mu_);
                ready_.wait(lk, [&] { return !q_.empty() || done_; });
                if (q_.empty()) return;
                s = std::move(q_.front());
                q_.pop_front();
                space_.notify_one();
            }
            try {
                auto nz = [](const std::vector<double>& v) { return v.empty() ? nullptr : v.data(); };
                w_[s.field].write_step(s.step, nz(s.strain), nz(s.disp), nz(s.vel), nz(s.acc));
                ++written_;
            } catch (const std::exception& e) {
                std::lock_guard<std::mutex> lk(mu_);
                if (err_.empty()) err_ = e.what();
                space_.notify_all();
            }
        }
    }
    std::vector<io::RecordWriter>& w_;
    size_t cap_;
    std::mutex mu_;
    std::condition_variable ready_, space_;
    std::deque<Snapshot> q_;
    bool done_ = false;
    std::string err_;
    size_t written_ = 0;
    std::thread th_;
};

void check(gfx_status st, const char* what) {
    if (st != GFX_OK) throw Error(st, std::string(what) + ": " + gfx_last_error());
}

std::vector<std::string> split_dirs(const std::string& s) {
    std::vector<std::string> out;
    std::string cur;
    for (char c : s + ",") {
        if (c == ',') {
            if (!cur.empty()) out.push_back(cur);
            cur.clear();
        } else if (c != ' ') {
            cur += c;
        }
    }
    for (const auto& d : out)
        if (d != "x" && d != "y" && d != "z")
            throw Error(GFX_ERR_INVALID, "--direction must be x, y, or z, got '" + d + "'");
    if (out.empty()) throw Error(GFX_ERR_INVALID, "no direction given");
    return out;
}

// Maps global element ids to local indices: first match in partition order,
// -1 when the cell is not in this run (solver.cpp:245-256).
std::vector<int32_t> local_cells(const std::vector<int64_t>& local_ids, const std::vector<int64_t>& want) {
    std::unordered_map<int64_t, int32_t> at;
    at.reserve(local_ids.size());
    for (size_t e = 0; e < local_ids.size(); ++e) at.emplace(local_ids[e], int32_t(e));
    std::vector<int32_t> out(want.size(), -1);
    for (size_t i = 0; i < want.size(); ++i) {
        auto it = at.find(want[i]);
        if (it != at.end()) out[i] = it->second;
    }
    return out;
}

struct FieldSpec {
    int direction;       // 0, 1, 2
    std::string subdir;  // "" or "s<k>"
    std::string dname;   // "x", "y", "z"
    int source;          // index into sources
};

int run_forward_impl(const gfx_forward_options& o) {
    const fs::path wd = (o.workdir && o.workdir[0]) ? fs::path(o.workdir) : fs::current_path();
    const std::vector<std::string> dirs = split_dirs(o.directions ? o.directions : "");
    std::string tag;
    for (const auto& d : dirs) tag += d;
    Logger log(wd, o.stations ? "batched_" + tag : tag, o.verbose != 0);
    try {
        log.info("single process, direction=" + tag + (o.stations ? " stations=" + std::string(o.stations) : ""));
        const auto t_io = std::chrono::steady_clock::now();
        const io::Config cfg = io::read_config((wd / "config.h5").string());
        const int ngll = cfg.polynomial_order + 1;
        require(ngll == 5, "the device solver is built for polynomial order 4 (ngll 5), config has " +
                               std::to_string(ngll));
        io::ReadOpts ro;
        ro.full = false;
        io::Partition part = io::read_partition_all((wd / "partitions").string(), ro);
        const int E = part.n_local_cell, NP = ngll * ngll * ngll;
        require(part.ngll == ngll, "partition NGLL differs from config polynomial_order");
        require(part.n_rank_node > 0 && !part.local_cell2rank_node.empty(),
                "partitions lack /field/cell/local_cell2global_node: the single-process run has no node "
                "numbering (solver.cpp:145,511)");
        const size_t P = size_t(E) * NP;
        require(part.jacobian.size() == P && part.dxi_dx.size() == P * 9 && part.mass.size() == P &&
                    part.lambda_.size() == P && part.mu_.size() == P && part.damping.size() == P &&
                    part.local_cell2rank_node.size() == P,
                "partition arrays do not match n_elem x ngll^3");
        require(!part.has_cpml || (part.coef_alpha.size() == P * 9 && part.coef_abar.size() == P * 5 &&
                                   part.density.size() == P && part.pml_region.size() == size_t(E)),
                "C-PML arrays do not match n_elem x ngll^3");
        {
            std::ostringstream os;
            os << "  n_local_cell=" << E << " n_gll_per_elem=" << NP << " n_rank_node=" << part.n_rank_node
               << " c-pml=" << (part.has_cpml ? "yes" : "no (legacy damping)") << " read "
               << std::chrono::duration<double>(std::chrono::steady_clock::now() - t_io).count() << " s";
            log.info(os.str());
        }

        // fields
        std::vector<io::SourceSpec> srcs;
        std::vector<FieldSpec> fields;
        if (o.stations) {
            srcs = io::read_stations(o.stations);
            for (size_t k = 0; k < srcs.size(); ++k)
                for (const auto& d : dirs)
                    fields.push_back({d == "x" ? 0 : (d == "y" ? 1 : 2), "s" + std::to_string(k), d, int(k)});
        } else {
            io::SourceSpec s;
            s.x = cfg.source_x;
            s.y = cfg.source_y;
            s.z = cfg.source_z;
            s.cell_ids = cfg.src_cell_ids;
            s.weights = cfg.src_weights;
            srcs.push_back(std::move(s));
            for (const auto& d : dirs) fields.push_back({d == "x" ? 0 : (d == "y" ? 1 : 2), "", d, 0});
        }
        const int F = int(fields.size());

        // centroids for the Morton element order
        std::vector<double> centroid;
        if (part.coords.size() == P * 3) {
            centroid.assign(size_t(E) * 3, 0.0);
            for (int e = 0; e < E; ++e)
                for (int q = 0; q < NP; ++q)
                    for (int d = 0; d < 3; ++d) centroid[size_t(e) * 3 + d] += part.coords[(size_t(e) * NP + q) * 3 + d];
            for (auto& v : centroid) v /= NP;
        }
        part.coords.clear();
        part.coords.shrink_to_fit();

        const bool recording = part.recording.has_recording && !part.recording.gll_node_ids.empty();
        const int n_snap_total = cfg.snapshot_stride > 0 ? (cfg.nsteps + cfg.snapshot_stride - 1) / cfg.snapshot_stride : 0;
        const int ring = std::max(1, o.ring > 0 ? o.ring : 8);

        gfx_solver_desc d{};
        d.n_elem = E;
        d.ngll = ngll;
        d.n_rank_node = part.n_rank_node;
        d.dxi_dx = part.dxi_dx.data();
        d.jacobian = part.jacobian.data();
        d.lambda_ = part.lambda_.data();
        d.mu_ = part.mu_.data();
        d.mass = part.mass.data();
        d.density = part.density.empty() ? nullptr : part.density.data();
        d.damping = part.damping.data();
        d.ibool = part.local_cell2rank_node.data();
        d.elem_centroid = centroid.empty() ? nullptr : centroid.data();
        d.has_cpml = part.has_cpml ? 1 : 0;
        d.pml_region = part.has_cpml ? part.pml_region.data() : nullptr;
        d.coef_alpha = part.has_cpml ? part.coef_alpha.data() : nullptr;
        d.coef_abar = part.has_cpml ? part.coef_abar.data() : nullptr;
        d.dt = cfg.solver_dt;
        d.snapshot_stride = cfg.snapshot_stride;
        d.n_rec_cell = recording ? int(part.recording.rec_cell_local.size()) : 0;
        d.rec_cell = recording ? part.recording.rec_cell_local.data() : nullptr;
        d.n_fields = F;
        d.record_capacity = (recording && n_snap_total > 0) ? std::min(ring, n_snap_total) : 0;
        d.record_uva = o.record_uva ? 1 : 0;
        d.assembly = o.assembly < 0 ? GFX_ASSEMBLY_COLORED : o.assembly;
        d.element_kernel = o.element_kernel;
        d.elems_per_block = o.elems_per_block;
        if (o.assembly < 0) {
            gfx_solver_desc best{};
            double ms = 0;
            check(gfx_autotune(&d, 3, &best, &ms), "autotune");
            d.assembly = best.assembly;
            d.element_kernel = best.element_kernel;
            d.elems_per_block = best.elems_per_block;
            d.node_field_groups = best.node_field_groups;
            std::ostringstream os;
            os << "  autotune: assembly=" << d.assembly << " element_kernel=" << d.element_kernel
               << " elems_per_block=" << d.elems_per_block << " (" << ms << " ms/step)";
            log.info(os.str());
        }
        gfx_solver* S = nullptr;
        check(gfx_solver_create(&d, &S), "solver create");
        std::unique_ptr<gfx_solver, gfx_status (*)(gfx_solver*)> guard(S, gfx_solver_destroy);

        // sources: STF f(step*dt) for step < len(stf_t) (solver.cpp:578-582)
        const size_t n_stf = std::min(cfg.stf_t.size(), cfg.stf_values.size());
        for (int f = 0; f < F; ++f) {
            const io::SourceSpec& s = srcs[fields[f].source];
            require(s.weights.size() == s.cell_ids.size() * size_t(NP), "source weights do not match the cells");
            const std::vector<int32_t> loc = local_cells(part.local_cell_ids, s.cell_ids);
            check(gfx_solver_set_source(S, f, fields[f].direction, int(loc.size()), loc.data(), s.weights.data(),
                                        int(n_stf), cfg.stf_values.data()),
                  "set source");
        }

        // record writers, one per field
        const bool f32 = cfg.snapshot_precision == "float32";
        std::vector<io::RecordWriter> writers;
        std::vector<fs::path> field_dir;
        for (int f = 0; f < F; ++f) {
            const fs::path base = fields[f].subdir.empty() ? wd / "wavefields" : wd / "wavefields" / fields[f].subdir;
            writers.emplace_back(base.string(), fields[f].dname, 0, part.recording, ngll, f32, cfg.record_depth_max_m,
                                 cfg.record_depth_actual_m);
            field_dir.push_back(base / fields[f].dname);
        }

        // restart stride (solver.cpp:363-371)
        int restart_stride = 0;
        if (o.restart_stride >= 0) {
            restart_stride = o.restart_stride;
        } else if (cfg.restart_dt_s > 0.0 && cfg.solver_dt > 0.0) {
            restart_stride = std::max(1, int(std::lround(cfg.restart_dt_s / cfg.solver_dt)));
        } else if (cfg.restart_stride > 0) {
            restart_stride = cfg.restart_stride;
        }

        // nodal damping for the restart file (solver.cpp:193-202: last writer wins)
        std::vector<double> node_damp(part.n_rank_node, 0.0);
        for (size_t q = 0; q < P; ++q) node_damp[part.local_cell2rank_node[q]] = part.damping[q];

        int start_step = 0;
        if (o.resume) {
            // every field's restart file must hold the same step; the complete
            // checkpoint sits in the first field's file
            try {
                std::vector<io::RestartState> rs;
                for (int f = 0; f < F; ++f) rs.push_back(io::read_restart(field_dir[f].string(), 0));
                for (int f = 1; f < F; ++f)
                    require(rs[f].step == rs[0].step, "restart files hold different steps");
                const int st = rs[0].step;
                if (st > 0 && st < cfg.nsteps) {
                    bool exact = false;
                    if (!rs[0].checkpoint.empty()) {
                        exact = gfx_solver_checkpoint_load(S, rs[0].checkpoint.data(), rs[0].checkpoint.size()) == GFX_OK;
                        if (!exact) log.info(std::string("  checkpoint not usable (") + gfx_last_error() +
                                             "), resuming from u, v, a");
                    }
                    if (!exact) {
                        for (int f = 0; f < F; ++f) {
                            const size_t n3 = size_t(part.n_rank_node) * 3;
                            require(rs[f].displacement.size() == n3 && rs[f].velocity.size() == n3 &&
                                        rs[f].acceleration.size() == n3,
                                    "restart state size differs from the mesh");
                            check(gfx_solver_set_state(S, f, rs[f].displacement.data(), rs[f].velocity.data(),
                                                       rs[f].acceleration.data()),
                                  "set state");
                        }
                        check(gfx_solver_set_step(S, st + 1), "set step");
                    }
                    start_step = st + 1;
                    std::ostringstream os;
                    os << "  resumed at step " << start_step << " (time_s=" << rs[0].time_s << ", "
                       << (exact ? "exact checkpoint" : "u, v, a only") << ")";
                    log.info(os.str());
                }
            } catch (const std::exception& ex) {
                log.error(std::string("  resume failed: ") + ex.what() + " - starting from scratch");
                check(gfx_solver_reset(S), "reset");
                start_step = 0;
            }
        }

        const int end_step = (o.stop_after_step >= 0) ? std::min(cfg.nsteps, o.stop_after_step + 1) : cfg.nsteps;
        {
            std::ostringstream os;
            os << "  fields=" << F << " snapshots=" << n_snap_total << " (stride=" << cfg.snapshot_stride
               << ") steps " << start_step << ".." << end_step << " of " << cfg.nsteps
               << " restart stride=" << restart_stride << " total sim time: " << cfg.nsteps * cfg.solver_dt << " s";
            log.info(os.str());
        }

        const size_t n_rp = size_t(d.n_rec_cell) * NP;
        WriterThread wt(writers, size_t(2 * ring * F + 1));
        const auto t0 = std::chrono::steady_clock::now();
        int drained = gfx_solver_n_snapshots(S);
        int step = start_step;  // updates done == next update index
        const int stride = cfg.snapshot_stride;
        int next_log = cfg.log_stride > 0 ? ((step / cfg.log_stride) + 1) * cfg.log_stride : end_step;
        while (step < end_step) {
            // segment end: at most `ring` snapshot steps, and stop right after a restart step
            int seg_end = end_step;
            if (stride > 0 && d.record_capacity > 0) seg_end = std::min(seg_end, step + stride * d.record_capacity);
            if (restart_stride > 0) {
                int t = (step / restart_stride) * restart_stride;
                if (t < step || t == 0) t += restart_stride;
                seg_end = std::min(seg_end, t + 1);
            }
            check(gfx_solver_run(S, seg_end - step), "run");
            step = seg_end;
            // drain new snapshots (snapshot j was taken after update j*stride)
            const int have = gfx_solver_n_snapshots(S);
            if (d.record_capacity > 0) {
                for (int j = drained; j < have; ++j)
                    for (int f = 0; f < F; ++f) {
                        Snapshot s;
                        s.field = f;
                        s.step = j * stride;
                        s.strain.resize(n_rp * 6);
                        if (o.record_uva) {
                            s.disp.resize(n_rp * 3);
                            s.vel.resize(n_rp * 3);
                            s.acc.resize(n_rp * 3);
                        }
                        check(gfx_solver_get_record(S, f, j, s.strain.data(), o.record_uva ? s.disp.data() : nullptr,
                                                    o.record_uva ? s.vel.data() : nullptr,
                                                    o.record_uva ? s.acc.data() : nullptr),
                              "get record");
                        wt.push(std::move(s));
                    }
            }
            drained = have;
            // restart after update index t = step - 1 (solver.cpp:624-627)
            const int t = step - 1;
            if (restart_stride > 0 && t > 0 && t % restart_stride == 0) {
                size_t nb = 0;
                check(gfx_solver_checkpoint_size(S, &nb), "checkpoint size");
                std::vector<uint8_t> image(nb);
                check(gfx_solver_checkpoint_save(S, image.data(), nb), "checkpoint");
                for (int f = 0; f < F; ++f) {
                    io::RestartState rs;
                    rs.step = t;
                    rs.time_s = t * cfg.solver_dt;
                    rs.ngll = ngll;
                    rs.n_rank_node = part.n_rank_node;
                    rs.direction = fields[f].dname;
                    const size_t n3 = size_t(part.n_rank_node) * 3;
                    rs.displacement.resize(n3);
                    rs.velocity.resize(n3);
                    rs.acceleration.resize(n3);
                    check(gfx_solver_get_state(S, f, rs.displacement.data(), rs.velocity.data(), rs.acceleration.data()),
                          "get state");
                    rs.pml_damping = node_damp;
                    if (f == 0) rs.checkpoint = std::move(image);
                    io::write_restart(field_dir[f].string(), 0, rs);
                }
            }
            if (step >= next_log || step == end_step) {
                const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
                const int done = step - start_step, left = end_step - step;
                std::ostringstream os;
                os << step << "/" << cfg.nsteps << " " << (100 * step / std::max(1, cfg.nsteps)) << "% elapsed="
                   << std::fixed << std::setprecision(1) << el << "s eta=" << (done > 0 ? el * left / done : 0.0) << "s";
                log.info(os.str());
                while (next_log <= step) next_log += std::max(1, cfg.log_stride);
            }
        }
        wt.finish();
        const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::ostringstream os;
        os << "simulation complete, " << (end_step - start_step) << " steps x " << F << " fields in " << el << "s, "
           << wt.written() << " record files";
        log.info(os.str());
        return 0;
    } catch (const std::exception& ex) {
        log.error(std::string("Error: ") + ex.what());
        set_error(ex.what());
        return 1;
    }
}

}  // namespace
}  // namespace gfx

using namespace gfx;

extern "C" gfx_status gfx_forward_options_default(gfx_forward_options* o) {
    return guarded([&] {
        require(o != nullptr, "null argument");
        *o = gfx_forward_options{};
        o->directions = "z";
        o->record_uva = 1;
        o->assembly = -1;
        o->element_kernel = GFX_ELEMENT_AUTO;
        o->ring = 8;
        o->stop_after_step = -1;
        o->restart_stride = -1;
        o->verbose = 0;
    });
}

extern "C" gfx_status gfx_run_forward_ex(const gfx_forward_options* o) {
    if (!o) {
        set_error("null options");
        return GFX_ERR_INVALID;
    }
    const int rc = run_forward_impl(*o);
    if (rc == 0) {
        clear_error();
        return GFX_OK;
    }
    return GFX_ERR_IO;
}

extern "C" int gfx_run_forward(const char* direction, int resume_mode, int /*effective_nprocs*/) {
    gfx_forward_options o;
    gfx_forward_options_default(&o);
    o.directions = direction ? direction : "";
    o.resume = resume_mode;
    o.verbose = 1;
    // one direction per call, as the reference CLI takes it
    if (!direction || (std::string(direction) != "x" && std::string(direction) != "y" &&
                       std::string(direction) != "z")) {
        set_error(std::string("--direction must be x, y, or z, got '") + (direction ? direction : "") + "'");
        return 1;
    }
    return run_forward_impl(o);
}

extern "C" gfx_status gfx_mesh_write_inputs(const gfx_mesh* m, const char* dir, int n_parts, double x, double y,
                                            double z, int n_stf, const double* stf, const char* precision,
                                            int restart_stride) {
    return guarded([&] {
        require(m && dir && (n_stf == 0 || stf), "null argument");
        io::SourceSpec src;
        src.x = x;
        src.y = y;
        src.z = z;
        std::vector<int32_t> cells(64);
        std::vector<double> w(size_t(64) * m->ngll * m->ngll * m->ngll);
        int n = 0;
        check(gfx_mesh_locate_source(m, x, y, z, 64, cells.data(), w.data(), &n), "locate source");
        src.cell_ids.assign(cells.begin(), cells.begin() + n);
        src.weights.assign(w.begin(), w.begin() + size_t(n) * m->ngll * m->ngll * m->ngll);
        io::write_inputs(*m, dir, n_parts, src, std::vector<double>(stf, stf + n_stf),
                         precision ? precision : "float32", restart_stride, 100);
    });
}

extern "C" gfx_status gfx_write_stations(const char* path, const gfx_mesh* m, int n, const double* xyz) {
    return guarded([&] {
        require(path && m && xyz && n > 0, "bad argument");
        const int NP = m->ngll * m->ngll * m->ngll;
        std::vector<io::SourceSpec> st(n);
        for (int k = 0; k < n; ++k) {
            std::vector<int32_t> cells(64);
            std::vector<double> w(size_t(64) * NP);
            int c = 0;
            check(gfx_mesh_locate_source(m, xyz[3 * k], xyz[3 * k + 1], xyz[3 * k + 2], 64, cells.data(), w.data(), &c),
                  "locate station");
            st[k].x = xyz[3 * k];
            st[k].y = xyz[3 * k + 1];
            st[k].z = xyz[3 * k + 2];
            st[k].cell_ids.assign(cells.begin(), cells.begin() + c);
            st[k].weights.assign(w.begin(), w.begin() + size_t(c) * NP);
        }
        io::write_stations(path, st, m->ngll);
    });
}

extern "C" int gfx_h5_available(void) { return h5::available() ? 1 : 0; }

extern "C" gfx_status gfx_h5_describe(const char* path, char* buf, size_t len, size_t* needed) {
    return guarded([&] {
        require(path != nullptr, "null path");
        const std::string s = h5::describe(path);
        if (needed) *needed = s.size() + 1;
        if (buf && len > 0) {
            const size_t n = std::min(len - 1, s.size());
            std::memcpy(buf, s.data(), n);
            buf[n] = '\0';
        }
    });
}

extern "C" gfx_status gfx_h5_read_f64(const char* path, const char* dataset, double* out, size_t n, size_t* count) {
    return guarded([&] {
        require(path && dataset, "null argument");
        h5::Obj f = h5::open_read(path);
        if (!out) {
            size_t total = 1;
            for (auto v : h5::shape(f.id(), dataset)) total *= size_t(v);
            if (count) *count = total;
            return;
        }
        const std::vector<double> v = h5::read<double>(f.id(), dataset);
        require(n >= v.size(), "output buffer too small");
        std::copy(v.begin(), v.end(), out);
        if (count) *count = v.size();
    });
}

extern "C" gfx_status gfx_h5_read_attr_f64(const char* path, const char* object, const char* name, double* out) {
    return guarded([&] {
        require(path && object && name && out, "null argument");
        h5::Obj f = h5::open_read(path);
        h5::Obj o = h5::open_object(f.id(), object);
        if (!h5::read_attr_double(o.id(), name, *out)) throw Error(GFX_ERR_IO, std::string("no attribute ") + name);
    });
}

extern "C" gfx_status gfx_h5_read_attr_str(const char* path, const char* object, const char* name, char* buf,
                                           size_t len) {
    return guarded([&] {
        require(path && object && name && buf && len > 0, "null argument");
        h5::Obj f = h5::open_read(path);
        h5::Obj o = h5::open_object(f.id(), object);
        std::string s;
        if (!h5::read_attr_string(o.id(), name, s)) throw Error(GFX_ERR_IO, std::string("no string attribute ") + name);
        const size_t n = std::min(len - 1, s.size());
        std::memcpy(buf, s.data(), n);
        buf[n] = '\0';
    });
}

extern "C" gfx_status gfx_mesh_read_inputs(const char* dir, gfx_mesh** out) {
    return guarded([&] {
        require(dir && out, "null argument");
        *out = nullptr;
        const fs::path wd(dir);
        const io::Config cfg = io::read_config((wd / "config.h5").string());
        io::Partition p = io::read_partition_all((wd / "partitions").string());
        require(!p.local_cell2rank_node.empty() && p.n_rank_node > 0,
                "partitions lack /field/cell/local_cell2global_node");
        const int E = p.n_local_cell, ng = p.ngll, nn = ng * ng * ng;
        const size_t P = size_t(E) * nn;
        require(p.coords.size() == P * 3, "partitions lack /field/cell/coords");
        auto H = std::make_unique<MeshHolder>();
        H->coords = std::move(p.coords);
        H->jacobian = std::move(p.jacobian);
        H->dxi_dx = std::move(p.dxi_dx);
        H->mass = std::move(p.mass);
        H->vp = std::move(p.vp);
        H->vs = std::move(p.vs);
        H->density = std::move(p.density);
        H->lambda_ = std::move(p.lambda_);
        H->mu_ = std::move(p.mu_);
        H->damping = std::move(p.damping);
        H->ibool = std::move(p.local_cell2rank_node);
        if (p.has_cpml) {
            H->pml_region = std::move(p.pml_region);
            H->coef_alpha = std::move(p.coef_alpha);
            H->coef_beta = std::move(p.coef_beta);
            H->coef_abar = std::move(p.coef_abar);
        }
        H->rec_cell = std::move(p.recording.rec_cell_local);
        H->node_xyz.assign(size_t(p.n_rank_node) * 3, 0.0);
        for (size_t q = 0; q < P; ++q)
            for (int d = 0; d < 3; ++d) H->node_xyz[size_t(H->ibool[q]) * 3 + d] = H->coords[q * 3 + d];
        H->spec.order = ng - 1;
        gfx_mesh& m = H->m;
        m.n_elem = E;
        m.ngll = ng;
        m.n_node_glob = p.n_rank_node;
        m.nx = cfg.nx_elements;
        m.ny = cfg.ny_elements;
        m.nz = cfg.nz_elements;
        m.xmin = cfg.xmin;
        m.xmax = cfg.xmax;
        m.ymin = cfg.ymin;
        m.ymax = cfg.ymax;
        m.zmin = cfg.zmin;
        m.zmax = cfg.zmax;
        m.dt = cfg.solver_dt;
        m.nsteps = cfg.nsteps;
        m.snapshot_stride = cfg.snapshot_stride;
        m.record_depth_actual = cfg.record_depth_actual_m;
        m.n_pml_elem = 0;
        for (int32_t r : H->pml_region) m.n_pml_elem += r != 0;
        wire_mesh(*H);
        *out = &H.release()->m;
    });
}

extern "C" gfx_status gfx_record_write_step(const char* inputs_dir, const char* out_dir, const char* direction,
                                            int step, const double* strain, const double* disp, const double* vel,
                                            const double* acc) {
    return guarded([&] {
        require(inputs_dir && out_dir && direction, "null argument");
        const fs::path wd(inputs_dir);
        const io::Config cfg = io::read_config((wd / "config.h5").string());
        io::ReadOpts ro;
        ro.full = false;
        const io::Partition p = io::read_partition_all((wd / "partitions").string(), ro);
        io::RecordWriter w(out_dir, direction, 0, p.recording, p.ngll, cfg.snapshot_precision == "float32",
                           cfg.record_depth_max_m, cfg.record_depth_actual_m);
        w.write_step(step, strain, disp, vel, acc);
    });
}

namespace {
h5::T type_from(const char* t) {
    const std::string s(t ? t : "");
    if (s == "f64") return h5::T::F64;
    if (s == "f32") return h5::T::F32;
    if (s == "i32") return h5::T::I32;
    if (s == "i64") return h5::T::I64;
    throw Error(GFX_ERR_INVALID, "type must be f64, f32, i32 or i64");
}
}  // namespace

extern "C" gfx_status gfx_h5_write(const char* path, int append, const char* dataset, const char* type, int ndims,
                                   const uint64_t* dims, const void* data) {
    return guarded([&] {
        require(path && dataset && ndims >= 1 && ndims <= 8 && dims, "bad argument");
        h5::Obj f = (append && fs::exists(path)) ? h5::open_rw(path) : h5::create_trunc(path);
        std::vector<h5::hsize_t> d(dims, dims + ndims);
        size_t total = 1;
        for (auto v : d) total *= size_t(v);
        require(total == 0 || data, "null data");
        const h5::T t = type_from(type);
        h5::WriteOpts o;
        o.overwrite = true;
        h5::write(f.id(), dataset, t, t, data, d, o);
    });
}

extern "C" gfx_status gfx_h5_write_attr(const char* path, const char* object, const char* name, const char* type,
                                        const void* value) {
    return guarded([&] {
        require(path && object && name && type && value, "bad argument");
        h5::Obj f = fs::exists(path) ? h5::open_rw(path) : h5::create_trunc(path);
        const std::string obj(object);
        h5::Obj o = (obj == "/" || h5::exists(f.id(), obj)) ? h5::open_object(f.id(), obj)
                                                             : h5::require_group(f.id(), obj);
        if (std::string(type) == "str")
            h5::write_attr_string(o.id(), name, static_cast<const char*>(value), true);
        else
            h5::write_attr(o.id(), name, type_from(type), value, true);
    });
}

extern "C" gfx_status gfx_config_summary(const char* path, double* out, int n, char* title, char* precision,
                                         size_t len) {
    return guarded([&] {
        require(path && out && n >= 24, "need 24 output slots");
        const io::Config c = io::read_config(path);
        const double v[24] = {double(c.polynomial_order), c.solver_dt, c.output_dt_s, double(c.snapshot_stride),
                              double(c.nsteps), c.cfl_safety, c.xmin, c.xmax, c.ymin, c.ymax, c.zmin, c.zmax,
                              double(c.stf_t.size()), double(c.stf_values.size()), c.source_x, c.source_y,
                              c.source_z, double(c.n_src_cell), c.record_depth_max_m, c.record_depth_actual_m,
                              double(c.restart_stride), c.restart_dt_s, double(c.log_stride),
                              double(c.tilex_elements.size())};
        std::copy(v, v + 24, out);
        auto put = [&](char* b, const std::string& s) {
            if (!b || len == 0) return;
            const size_t k = std::min(len - 1, s.size());
            std::memcpy(b, s.data(), k);
            b[k] = '\0';
        };
        put(title, c.title);
        put(precision, c.snapshot_precision);
    });
}

extern "C" gfx_status gfx_partition_summary(const char* path, int merge_dir, double* out, int n) {
    return guarded([&] {
        require(path && out && n >= 12, "need 12 output slots");
        const io::Partition p = merge_dir ? io::read_partition_all(path) : io::read_partition(path);
        const double v[12] = {double(p.n_local_cell), double(p.ngll), double(p.n_rank_node), p.has_cpml ? 1.0 : 0.0,
                              p.recording.has_recording ? 1.0 : 0.0, double(p.recording.rec_cell_local.size()),
                              double(p.recording.gll_node_ids.size()), double(p.local_cell_ids.size()),
                              double(p.n_exchange_neighbors), double(p.jacobian.size()),
                              double(p.local_cell2rank_node.size()), double(p.local_cell2global_node.size())};
        std::copy(v, v + 12, out);
    });
}
