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
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <ctime>
#include <deque>
#include <filesystem>
#include <fstream>
#include <functional>
#include <iomanip>
#include <iostream>
#include <map>
#include <mutex>
#include <sstream>
#include <thread>
#include <unordered_map>

#include "common.hpp"
#include "gfx.h"
#include "gfx_io.h"
#include "gfx_surrogate.h"
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

// Inputs of a run: config.h5 and every partition merged (solver.cpp:73-220),
// with the element centroids for the Morton element order.
struct RunInputs {
    io::Config cfg;
    io::Partition part;
    std::vector<double> centroid;
    int ngll = 0, E = 0, NP = 0;
    size_t P = 0;
};

void load_inputs(const fs::path& wd, Logger& log, RunInputs& in, bool profiles = false) {
    const auto t_io = std::chrono::steady_clock::now();
    in.cfg = io::read_config((wd / "config.h5").string());
    in.ngll = in.cfg.polynomial_order + 1;
    require(in.ngll == 5, "the device solver is built for polynomial order 4 (ngll 5), config has " +
                              std::to_string(in.ngll));
    io::ReadOpts ro;
    ro.full = false;
    ro.profiles = profiles;
    in.part = io::read_partition_all((wd / "partitions").string(), ro);
    io::Partition& part = in.part;
    in.E = part.n_local_cell;
    in.NP = in.ngll * in.ngll * in.ngll;
    const int E = in.E, NP = in.NP;
    require(part.ngll == in.ngll, "partition NGLL differs from config polynomial_order");
    require(part.n_rank_node > 0 && !part.local_cell2rank_node.empty(),
            "partitions lack /field/cell/local_cell2global_node: the single-process run has no node "
            "numbering (solver.cpp:145,511)");
    const size_t P = size_t(E) * NP;
    in.P = P;
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
    // centroids for the Morton element order
    if (part.coords.size() == P * 3) {
        in.centroid.assign(size_t(E) * 3, 0.0);
        for (int e = 0; e < E; ++e)
            for (int q = 0; q < NP; ++q)
                for (int d = 0; d < 3; ++d)
                    in.centroid[size_t(e) * 3 + d] += part.coords[(size_t(e) * NP + q) * 3 + d];
        for (auto& v : in.centroid) v /= NP;
    }
    part.coords.clear();
    part.coords.shrink_to_fit();
}

// Solver description over the loaded arrays; assembly < 0 autotunes the launch.
gfx_solver_desc solver_desc(RunInputs& in, int F, int record_capacity, bool record_uva, int assembly,
                            int element_kernel, int elems_per_block, Logger& log, bool cpml_complete = false) {
    io::Partition& part = in.part;
    const bool recording = part.recording.has_recording && !part.recording.gll_node_ids.empty();
    gfx_solver_desc d{};
    d.n_elem = in.E;
    d.ngll = in.ngll;
    d.n_rank_node = part.n_rank_node;
    d.dxi_dx = part.dxi_dx.data();
    d.jacobian = part.jacobian.data();
    d.lambda_ = part.lambda_.data();
    d.mu_ = part.mu_.data();
    d.mass = part.mass.data();
    d.density = part.density.empty() ? nullptr : part.density.data();
    d.damping = part.damping.data();
    d.ibool = part.local_cell2rank_node.data();
    d.elem_centroid = in.centroid.empty() ? nullptr : in.centroid.data();
    d.has_cpml = part.has_cpml ? 1 : 0;
    d.pml_region = part.has_cpml ? part.pml_region.data() : nullptr;
    d.coef_alpha = part.has_cpml ? part.coef_alpha.data() : nullptr;
    d.coef_abar = part.has_cpml ? part.coef_abar.data() : nullptr;
    d.dt = in.cfg.solver_dt;
    d.snapshot_stride = in.cfg.snapshot_stride;
    d.n_rec_cell = recording ? int(part.recording.rec_cell_local.size()) : 0;
    d.rec_cell = recording ? part.recording.rec_cell_local.data() : nullptr;
    d.n_fields = F;
    d.record_capacity = recording ? record_capacity : 0;
    d.record_uva = record_uva ? 1 : 0;
    d.assembly = assembly < 0 ? GFX_ASSEMBLY_COLORED : assembly;
    d.element_kernel = element_kernel;
    d.elems_per_block = elems_per_block;
    if (part.n_sls > 0) {  // shear attenuation solids in the partitions
        require(part.tau_sigma.size() == in.P * size_t(part.n_sls) && part.tau_epsilon.size() == part.tau_sigma.size(),
                "attenuation arrays do not match n_elem x ngll^3 x n_sls");
        d.n_sls = part.n_sls;
        d.tau_sigma = part.tau_sigma.data();
        d.tau_epsilon = part.tau_epsilon.data();
    }
    if (cpml_complete) {
        require(part.has_cpml && part.pml_K.size() == in.P * 3 && part.pml_d.size() == in.P * 3 &&
                    part.pml_alpha.size() == in.P * 3,
                "the complete C-PML needs /field/cell/pml_K, pml_d and pml_alpha [n_cell, ngll^3, 3] in the "
                "partitions");
        d.cpml_complete = 1;
        d.pml_K = part.pml_K.data();
        d.pml_d = part.pml_d.data();
        d.pml_alpha = part.pml_alpha.data();
        log.info("  complete C-PML (stretched stresses, rigid outer boundary)");
    }
    if (assembly < 0) {
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
    return d;
}

// Sources: STF f(step*dt) for step < len(stf_t) (solver.cpp:578-582).
void set_sources(gfx_solver* S, const RunInputs& in, const std::vector<io::SourceSpec>& srcs,
                 const std::vector<FieldSpec>& fields) {
    const size_t n_stf = std::min(in.cfg.stf_t.size(), in.cfg.stf_values.size());
    for (size_t f = 0; f < fields.size(); ++f) {
        const io::SourceSpec& s = srcs[fields[f].source];
        require(s.weights.size() == s.cell_ids.size() * size_t(in.NP), "source weights do not match the cells");
        const std::vector<int32_t> loc = local_cells(in.part.local_cell_ids, s.cell_ids);
        check(gfx_solver_set_source(S, int(f), fields[f].direction, int(loc.size()), loc.data(), s.weights.data(),
                                    int(n_stf), in.cfg.stf_values.data()),
              "set source");
    }
}

io::SourceSpec config_source(const io::Config& cfg) {
    io::SourceSpec s;
    s.x = cfg.source_x;
    s.y = cfg.source_y;
    s.z = cfg.source_z;
    s.cell_ids = cfg.src_cell_ids;
    s.weights = cfg.src_weights;
    return s;
}

int run_forward_impl(const gfx_forward_options& o) {
    const fs::path wd = (o.workdir && o.workdir[0]) ? fs::path(o.workdir) : fs::current_path();
    const std::vector<std::string> dirs = split_dirs(o.directions ? o.directions : "");
    std::string tag;
    for (const auto& d : dirs) tag += d;
    Logger log(wd, o.stations ? "batched_" + tag : tag, o.verbose != 0);
    try {
        log.info("single process, direction=" + tag + (o.stations ? " stations=" + std::string(o.stations) : ""));
        RunInputs in;
        load_inputs(wd, log, in, o.cpml_complete != 0);
        const io::Config& cfg = in.cfg;
        io::Partition& part = in.part;
        const int ngll = in.ngll, NP = in.NP;
        const size_t P = in.P;

        // fields
        std::vector<io::SourceSpec> srcs;
        std::vector<FieldSpec> fields;
        if (o.stations) {
            srcs = io::read_stations(o.stations);
            for (size_t k = 0; k < srcs.size(); ++k)
                for (const auto& d : dirs)
                    fields.push_back({d == "x" ? 0 : (d == "y" ? 1 : 2), "s" + std::to_string(k), d, int(k)});
        } else {
            srcs.push_back(config_source(cfg));
            for (const auto& d : dirs) fields.push_back({d == "x" ? 0 : (d == "y" ? 1 : 2), "", d, 0});
        }
        const int F = int(fields.size());

        const int n_snap_total = cfg.snapshot_stride > 0 ? (cfg.nsteps + cfg.snapshot_stride - 1) / cfg.snapshot_stride : 0;
        const int ring = std::max(1, o.ring > 0 ? o.ring : 8);
        gfx_solver_desc d = solver_desc(in, F, n_snap_total > 0 ? std::min(ring, n_snap_total) : 0, o.record_uva != 0,
                                        o.assembly, o.element_kernel, o.elems_per_block, log, o.cpml_complete != 0);
        gfx_solver* S = nullptr;
        check(gfx_solver_create(&d, &S), "solver create");
        std::unique_ptr<gfx_solver, gfx_status (*)(gfx_solver*)> guard(S, gfx_solver_destroy);
        set_sources(S, in, srcs, fields);

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


// ---------------------------------------------------------------------------
// Green's function tiles straight from the device (postprocess/cpp/main.cpp,
// writer.hh): the x, y and z forces run as three fields of one solver, every
// snapshot is merged to the recorded GLL nodes and laid out per tile on the
// GPU, and a writer thread appends it to the open tile files.  No record
// file and no [nt, n_gll, 6, 3] host tensor exist at any point.
// ---------------------------------------------------------------------------

// Tile of the interior grid (writer.hh:28-60; main.cpp:667-745).
struct Tile {
    int tx = 0, ty = 0;
    std::vector<int32_t> cells;       // merged recording-cell indices, ascending
    std::vector<int32_t> nodes;       // tile-local -> merged node, first appearance
    std::vector<int32_t> cell_index;  // [cells * 125] tile-local node or -1
    double x0 = 0, x1 = 0, y0 = 0, y1 = 0;
};

struct TilePlan {
    int n_node = 0;                 // merged recorded GLL nodes
    std::vector<int64_t> off;       // merge entries of node n: [off[n], off[n+1])
    std::vector<int32_t> pt;        // recording point c*125 + p
    std::vector<double> w;          // mass weight (1 where the cell has none)
    std::vector<Tile> tiles;        // sorted by (tx, ty)
    std::vector<int64_t> slab_off;  // tile k owns slab entries [slab_off[k], slab_off[k+1])
    std::vector<int32_t> slab_node;
};

// writer.hh:28-36: the tile holding interior element index i; past the last
// boundary, the last tile.
int tile_of(int64_t i, const std::vector<int>& sizes) {
    int64_t cum = 0;
    for (size_t t = 0; t < sizes.size(); ++t) {
        cum += sizes[t];
        if (i < cum) return int(t);
    }
    return int(sizes.size()) - 1;
}

// The node merge of one rank file (main.cpp:133-266 with one file: merged
// node i = recorded node i) and its mass weights (main.cpp:306-381); then the
// whole-cell tiling (main.cpp:667-745) and tile-local node lists
// (main.cpp:757-790).  mass(c, p) returns the weight of point p of recording
// cell c, or a negative value for "no mass" (weight 1).
TilePlan plan_tiles(const io::Config& cfg, const io::RecordingMap& rm, const double bounds[6],
                    const std::function<double(int, int)>& mass) {
    constexpr int NPC = 125;
    TilePlan T;
    const int nu = int(rm.gll_node_ids.size());
    const int nrc = int(rm.rec_cell_local.size());
    require(nu > 0 && nrc > 0, "the partitions record nothing: no tiles to write");
    require(rm.cell_gll_node_index.size() == size_t(nrc) * NPC, "recording map: cell_gll_node_index size");
    require(rm.gll_node_coords.size() == size_t(nu) * 3, "recording map: gll_node_coords size");
    const std::vector<int32_t>& cgi = rm.cell_gll_node_index;
    T.n_node = nu;
    // entries per node in the order cell c, point p
    T.off.assign(size_t(nu) + 1, 0);
    for (int32_t i : cgi)
        if (i >= 0 && i < nu) ++T.off[size_t(i) + 1];
    for (int n = 0; n < nu; ++n) T.off[n + 1] += T.off[n];
    T.pt.resize(size_t(T.off[nu]));
    T.w.resize(size_t(T.off[nu]));
    std::vector<int64_t> fill(T.off.begin(), T.off.end() - 1);
    for (int c = 0; c < nrc; ++c)
        for (int p = 0; p < NPC; ++p) {
            const int32_t i = cgi[size_t(c) * NPC + p];
            if (i < 0 || i >= nu) continue;
            const int64_t e = fill[i]++;
            T.pt[e] = c * NPC + p;
            const double m = mass(c, p);
            T.w[e] = m < 0.0 ? 1.0 : m;
        }

    // whole-cell tiling by cell centre (nodes p = 0 and p = 124)
    const double xmin = bounds[0], xmax = bounds[1], ymin = bounds[2], ymax = bounds[3];
    const double dx = (xmax - xmin) / double(cfg.nx_elements);
    const double dy = (ymax - ymin) / double(cfg.ny_elements);
    int64_t nix = 0, niy = 0;
    for (int v : cfg.tilex_elements) nix += v;
    for (int v : cfg.tiley_elements) niy += v;
    require(!cfg.tilex_elements.empty() && !cfg.tiley_elements.empty(),
            "config.h5 has no /simulation/tilex_elements or tiley_elements");
    std::map<std::pair<int, int>, std::vector<int32_t>> bins;
    const std::vector<double>& X = rm.gll_node_coords;
    for (int c = 0; c < nrc; ++c) {
        const int32_t i0 = cgi[size_t(c) * NPC], i1 = cgi[size_t(c) * NPC + 124];
        if (i0 < 0 || i0 >= nu || i1 < 0 || i1 >= nu) continue;
        const double cx = 0.5 * (X[size_t(i0) * 3] + X[size_t(i1) * 3]);
        const double cy = 0.5 * (X[size_t(i0) * 3 + 1] + X[size_t(i1) * 3 + 1]);
        int64_t ei = dx > 0 ? int64_t(std::floor((cx - xmin) / dx)) : 0;
        int64_t ej = dy > 0 ? int64_t(std::floor((cy - ymin) / dy)) : 0;
        if (ei < 0) ei = 0;
        if (ej < 0) ej = 0;
        if (cfg.nx_elements > 0 && ei >= cfg.nx_elements) ei = cfg.nx_elements - 1;
        if (cfg.ny_elements > 0 && ej >= cfg.ny_elements) ej = cfg.ny_elements - 1;
        const int64_t ii = ei - cfg.pml_xmin, ij = ej - cfg.pml_ymin;
        if (ii < 0 || ii >= nix || ij < 0 || ij >= niy) continue;
        bins[{tile_of(ii, cfg.tilex_elements), tile_of(ij, cfg.tiley_elements)}].push_back(c);
    }
    require(!bins.empty(), "no recording cell lies in the interior tile grid");

    T.slab_off.push_back(0);
    for (auto& kv : bins) {
        Tile t;
        t.tx = kv.first.first;
        t.ty = kv.first.second;
        t.cells = std::move(kv.second);
        std::unordered_map<int32_t, int32_t> local;
        for (int32_t c : t.cells)
            for (int p = 0; p < NPC; ++p) {
                const int32_t g = cgi[size_t(c) * NPC + p];
                if (g >= 0 && local.emplace(g, int32_t(t.nodes.size())).second) t.nodes.push_back(g);
            }
        t.cell_index.resize(t.cells.size() * NPC);
        for (size_t k = 0; k < t.cells.size(); ++k)
            for (int p = 0; p < NPC; ++p) {
                const int32_t g = cgi[size_t(t.cells[k]) * NPC + p];
                t.cell_index[k * NPC + p] = g >= 0 ? local.at(g) : -1;
            }
        // tile bounds on the element grid (main.cpp:724-741)
        int64_t cx = 0, cy = 0;
        for (int k = 0; k < t.tx; ++k) cx += cfg.tilex_elements[k];
        for (int k = 0; k < t.ty; ++k) cy += cfg.tiley_elements[k];
        t.x0 = xmin + double(cfg.pml_xmin + cx) * dx;
        t.x1 = xmin + double(cfg.pml_xmin + cx + cfg.tilex_elements[t.tx]) * dx;
        t.y0 = ymin + double(cfg.pml_ymin + cy) * dy;
        t.y1 = ymin + double(cfg.pml_ymin + cy + cfg.tiley_elements[t.ty]) * dy;
        T.slab_node.insert(T.slab_node.end(), t.nodes.begin(), t.nodes.end());
        T.slab_off.push_back(int64_t(T.slab_node.size()));
        T.tiles.push_back(std::move(t));
    }
    return T;
}

// One open tile file with its tensor datasets (writer.hh:105-310 schema).
struct TileFile {
    h5::Obj file, greens, disp, vel, acc;
    std::string path;
};

struct TileSlab {
    int snap = 0;
    std::vector<uint8_t> greens, disp;
};

// Writes every tile's slab of a snapshot, on one thread (the HDF5 library is
// not thread safe); a bounded queue keeps host memory to `cap` snapshots.
class TileWriter {
public:
    TileWriter(std::vector<TileFile>& files, const TilePlan& plan, h5::T type, size_t vb, size_t cap)
        : files_(files), plan_(plan), type_(type), vb_(vb), cap_(cap), th_([this] { loop(); }) {}
    ~TileWriter() {
        try {
            finish();
        } catch (...) {
            // an error is already propagating
        }
    }
    void push(TileSlab&& s) {
        std::unique_lock<std::mutex> lk(mu_);
        space_.wait(lk, [&] { return q_.size() < cap_ || !err_.empty(); });
        if (!err_.empty()) throw Error(GFX_ERR_IO, err_);
        q_.push_back(std::move(s));
        ready_.notify_one();
    }
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

private:
    void loop() {
        for (;;) {
            TileSlab s;
            {
                std::unique_lock<std::mutex> lk(mu_);
                ready_.wait(lk, [&] { return !q_.empty() || done_; });
                if (q_.empty()) return;
                s = std::move(q_.front());
                q_.pop_front();
                space_.notify_one();
            }
            try {
                write(s);
            } catch (const std::exception& e) {
                std::lock_guard<std::mutex> lk(mu_);
                if (err_.empty()) err_ = e.what();
                space_.notify_all();
            }
        }
    }
    void write(const TileSlab& s) {
        for (size_t k = 0; k < files_.size(); ++k) {
            const int64_t j0 = plan_.slab_off[k];
            const h5::hsize_t nl = h5::hsize_t(plan_.slab_off[k + 1] - j0);
            const h5::hsize_t t = h5::hsize_t(s.snap);
            h5::write_slab(files_[k].greens.id(), type_, s.greens.data() + size_t(j0) * 18 * vb_, {t, 0, 0, 0},
                           {1, nl, 6, 3});
            if (files_[k].disp.valid()) {
                h5::write_slab(files_[k].disp.id(), type_, s.disp.data() + size_t(j0) * 9 * vb_, {t, 0, 0, 0},
                               {1, nl, 3, 3});
                // velocity and acceleration are left at zero by the merge (main.cpp:401)
                if (zeros_.size() < size_t(nl) * 9 * vb_) zeros_.assign(size_t(nl) * 9 * vb_, 0);
                h5::write_slab(files_[k].vel.id(), type_, zeros_.data(), {t, 0, 0, 0}, {1, nl, 3, 3});
                h5::write_slab(files_[k].acc.id(), type_, zeros_.data(), {t, 0, 0, 0}, {1, nl, 3, 3});
            }
        }
    }
    std::vector<TileFile>& files_;
    const TilePlan& plan_;
    h5::T type_;
    size_t vb_, cap_;
    std::vector<uint8_t> zeros_;
    std::mutex mu_;
    std::condition_variable ready_, space_;
    std::deque<TileSlab> q_;
    bool done_ = false;
    std::string err_;
    std::thread th_;
};

// Creates a tile file with its attributes, /time, /mesh and /source, and the
// empty /field tensors, chunked one time step per chunk with shuffle + gzip 4
// (writer.hh:68-100).
TileFile open_tile(const std::string& path, const Tile& t, const TilePlan&, const io::Config& cfg,
                   const io::RecordingMap& rm, const double bounds[6], int nt, const std::vector<double>& time,
                   const std::vector<double>& stf_t, const std::vector<double>& stf_v, h5::T type, bool uva) {
    TileFile F;
    F.path = path;
    F.file = h5::create_trunc(path);
    const h5::hid_t f = F.file.id();
    auto str = [&](const char* name, const std::string& v) { h5::write_attr_string(f, name, v, false, 64); };
    auto dbl = [&](const char* name, double v) { h5::write_attr(f, name, h5::T::F64, &v); };
    auto i32 = [&](h5::hid_t loc, const char* name, int v) { h5::write_attr(loc, name, h5::T::I32, &v); };
    str("version", "1.0.0");
    str("basis", "gll");
    i32(f, "tile_x_index", t.tx);
    i32(f, "tile_y_index", t.ty);
    dbl("x_min_m", t.x0);
    dbl("x_max_m", t.x1);
    dbl("y_min_m", t.y0);
    dbl("y_max_m", t.y1);
    dbl("z_min_m", bounds[4]);
    dbl("z_max_m", bounds[5]);
    dbl("record_depth_max_m", cfg.record_depth_max_m);
    dbl("record_depth_actual_m", cfg.record_depth_actual_m);
    const double src[3] = {cfg.source_x, cfg.source_y, cfg.source_z};
    h5::write_attr_array(f, "source_xyz_m", h5::T::F64, src, 3);
    str("source_directions", "x,y,z");
    str("greens_quantities", uva ? "strain,displacement,velocity,acceleration" : "strain");
    i32(f, "excludes_pml", 1);
    {
        h5::Obj g = h5::require_group(f, "/time");
        h5::write_vec(f, "/time/t", time);
        h5::write_attr(g.id(), "dt", h5::T::F64, &cfg.solver_dt);
        i32(g.id(), "nsteps", nt);
    }
    const size_t nl = t.nodes.size();
    {
        std::vector<int64_t> ids(nl);
        std::vector<double> xyz(nl * 3);
        for (size_t i = 0; i < nl; ++i) {
            ids[i] = rm.gll_node_ids[t.nodes[i]];
            for (int d = 0; d < 3; ++d) xyz[i * 3 + d] = rm.gll_node_coords[size_t(t.nodes[i]) * 3 + d];
        }
        h5::write_vec(f, "/mesh/gll_node_ids", ids);
        h5::write_vec(f, "/mesh/gll_node_coords", xyz, {h5::hsize_t(nl), 3});
        h5::write_vec(f, "/mesh/cell_gll_node_index", t.cell_index, {h5::hsize_t(t.cells.size()), 125});
    }
    if (!stf_t.empty()) {
        h5::write_vec(f, "/source/stf_t", stf_t);
        h5::write_vec(f, "/source/stf_values", stf_v);
    }
    const h5::hsize_t c1 = nl == 0 ? 1 : h5::hsize_t(nl);
    F.greens = h5::create_dataset(f, "/field/greens_tensor", type, {h5::hsize_t(nt), h5::hsize_t(nl), 6, 3},
                                  {1, c1, 6, 3}, true, 4);
    if (uva) {
        const std::vector<h5::hsize_t> dims = {h5::hsize_t(nt), h5::hsize_t(nl), 3, 3}, ch = {1, c1, 3, 3};
        F.disp = h5::create_dataset(f, "/field/displacement_tensor", type, dims, ch, true, 4);
        F.vel = h5::create_dataset(f, "/field/velocity_tensor", type, dims, ch, true, 4);
        F.acc = h5::create_dataset(f, "/field/acceleration_tensor", type, dims, ch, true, 4);
    }
    return F;
}

int run_green_tiles_impl(const gfx_green_tile_options& o, int* n_tiles_out) {
    const fs::path wd = (o.workdir && o.workdir[0]) ? fs::path(o.workdir) : fs::current_path();
    const fs::path out = (o.output_dir && o.output_dir[0]) ? fs::path(o.output_dir) : wd / "greenfun";
    Logger log(wd, "greens", o.verbose != 0);
    try {
        log.info("single process, Green's function tiles, directions x,y,z -> " + out.string());
        RunInputs in;
        load_inputs(wd, log, in, o.cpml_complete != 0);
        const io::Config& cfg = in.cfg;
        const io::RecordingMap& rm = in.part.recording;
        require(rm.has_recording && !rm.gll_node_ids.empty(), "the partitions carry no recording map");
        require(cfg.snapshot_stride > 0 && cfg.nsteps > 0, "no snapshots: snapshot_stride and nsteps must be > 0");
        const int nt = (cfg.nsteps + cfg.snapshot_stride - 1) / cfg.snapshot_stride;

        // domain bounds and cell mass: model.h5 when given (main.cpp:458-483), else config.h5 /domain and the
        // partitions' mass (the same per-cell arrays the preprocessor writes to both)
        double bounds[6] = {cfg.xmin, cfg.xmax, cfg.ymin, cfg.ymax, cfg.zmin, cfg.zmax};
        std::vector<double> model_mass;
        int64_t n_model_cell = 0;
        if (o.model_path && o.model_path[0]) {
            h5::Obj mf = h5::open_read(o.model_path);
            if (h5::exists(mf.id(), "/domain")) {
                h5::Obj g = h5::open_group(mf.id(), "/domain");
                const char* names[6] = {"xmin", "xmax", "ymin", "ymax", "zmin", "zmax"};
                for (int k = 0; k < 6; ++k) h5::read_attr_double(g.id(), names[k], bounds[k]);
            }
            if (h5::exists(mf.id(), "/field/cell/mass")) {
                std::vector<h5::hsize_t> dims;
                model_mass = h5::read<double>(mf.id(), "/field/cell/mass", &dims);
                require(dims.size() == 4 && dims[1] == 5 && dims[2] == 5 && dims[3] == 5,
                        "model /field/cell/mass must be [n_cell, 5, 5, 5]");
                n_model_cell = int64_t(dims[0]);
            }
        }
        const io::Partition& part = in.part;
        std::function<double(int, int)> mass;
        if (!model_mass.empty()) {
            mass = [&](int c, int p) {
                const int64_t g = c < int(rm.rec_cell_global.size()) ? rm.rec_cell_global[c] : -1;
                return (g >= 0 && g < n_model_cell) ? model_mass[size_t(g) * 125 + p] : -1.0;
            };
        } else {
            n_model_cell = in.E;
            mass = [&](int c, int p) {
                const int64_t g = c < int(rm.rec_cell_global.size()) ? rm.rec_cell_global[c] : -1;
                return (g >= 0 && g < n_model_cell) ? part.mass[size_t(rm.rec_cell_local[c]) * 125 + p] : -1.0;
            };
        }
        const TilePlan plan = plan_tiles(cfg, rm, bounds, mass);
        {
            std::ostringstream os;
            os << "  " << plan.n_node << " recorded GLL nodes, " << rm.rec_cell_local.size() << " recording cells, "
               << plan.tiles.size() << " tiles, " << plan.slab_node.size() << " tile-node entries, " << nt
               << " snapshots";
            log.info(os.str());
        }

        // time axis at output_dt_s and the STF down-sampled to it (main.cpp:518-547)
        std::vector<double> time(nt), stf_t, stf_v;
        for (int s = 0; s < nt; ++s) time[s] = double(s) * cfg.output_dt_s;
        if (!cfg.stf_t.empty()) {
            const int64_t nstf = int64_t(cfg.stf_t.size());
            const int64_t stride = std::max<int64_t>(1, nstf / nt);
            stf_t.resize(nt);
            stf_v.resize(nt);
            for (int s = 0; s < nt; ++s) {
                const int64_t k = std::min(int64_t(s) * stride, nstf - 1);
                stf_t[s] = cfg.stf_t[k];
                stf_v[s] = k < int64_t(cfg.stf_values.size()) ? cfg.stf_values[k] : 0.0;
            }
        }

        // solver: fields x, y, z of the config source
        const std::vector<io::SourceSpec> srcs = {config_source(cfg)};
        const std::vector<FieldSpec> fields = {{0, "", "x", 0}, {1, "", "y", 0}, {2, "", "z", 0}};
        const int ring = std::max(1, o.ring > 0 ? o.ring : 8);
        const bool uva = o.with_uva != 0;
        gfx_solver_desc d = solver_desc(in, 3, std::min(ring, nt), uva, o.assembly, o.element_kernel,
                                        o.elems_per_block, log, o.cpml_complete != 0);
        gfx_solver* S = nullptr;
        check(gfx_solver_create(&d, &S), "solver create");
        std::unique_ptr<gfx_solver, gfx_status (*)(gfx_solver*)> guard(S, gfx_solver_destroy);
        set_sources(S, in, srcs, fields);

        const bool f64 = cfg.snapshot_precision != "float32";
        const size_t vb = f64 ? 8 : 4;
        gfx_green_slab* slab = nullptr;
        check(gfx_green_slab_create(plan.n_node, plan.off.data(), plan.pt.data(), plan.w.data(),
                                    int64_t(plan.slab_node.size()), plan.slab_node.data(), f64 ? 1 : 0, uva ? 1 : 0,
                                    &slab),
              "slab create");
        std::unique_ptr<gfx_green_slab, gfx_status (*)(gfx_green_slab*)> slab_guard(slab, gfx_green_slab_destroy);

        fs::create_directories(out);
        std::vector<TileFile> files;
        const h5::T type = f64 ? h5::T::F64 : h5::T::F32;
        for (const Tile& t : plan.tiles) {
            char name[64];
            std::snprintf(name, sizeof(name), "tile_x%03d_y%03d.h5", t.tx, t.ty);
            files.push_back(open_tile((out / name).string(), t, plan, cfg, rm, bounds, nt, time, stf_t, stf_v, type,
                                      uva));
        }

        const auto t0 = std::chrono::steady_clock::now();
        const int32_t fod[3] = {0, 1, 2};
        const size_t ns = plan.slab_node.size();
        TileWriter tw(files, plan, type, vb, size_t(2 * ring + 1));
        int drained = 0, step = 0;
        while (step < cfg.nsteps) {
            const int seg_end = std::min(cfg.nsteps, step + cfg.snapshot_stride * d.record_capacity);
            check(gfx_solver_run(S, seg_end - step), "run");
            step = seg_end;
            const int have = gfx_solver_n_snapshots(S);
            for (int j = drained; j < have; ++j) {
                TileSlab s;
                s.snap = j;
                s.greens.resize(ns * 18 * vb);
                if (uva) s.disp.resize(ns * 9 * vb);
                check(gfx_green_slab_capture(slab, S, j, fod, s.greens.data(), uva ? s.disp.data() : nullptr),
                      "slab capture");
                tw.push(std::move(s));
            }
            drained = have;
        }
        tw.finish();
        require(drained == nt, "snapshot count differs from the time axis");
        files.clear();
        const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::ostringstream os;
        os << "tiles complete: " << plan.tiles.size() << " files, " << nt << " snapshots, " << cfg.nsteps
           << " steps x 3 fields in " << el << "s";
        log.info(os.str());
        if (n_tiles_out) *n_tiles_out = int(plan.tiles.size());
        return 0;
    } catch (const std::exception& ex) {
        log.error(std::string("Error: ") + ex.what());
        throw;
    }
}

}  // namespace
}  // namespace gfx

using namespace gfx;

extern "C" gfx_status gfx_green_tile_options_default(gfx_green_tile_options* o) {
    return guarded([&] {
        require(o, "null options");
        std::memset(o, 0, sizeof(*o));
        o->with_uva = 1;
        o->assembly = -1;
        o->ring = 8;
    });
}

extern "C" gfx_status gfx_run_green_tiles(const gfx_green_tile_options* o, int* n_tiles) {
    return guarded([&] {
        require(o, "null options");
        run_green_tiles_impl(*o, n_tiles);
    });
}

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

extern "C" gfx_status gfx_h5_read_attr_f64_array(const char* path, const char* object, const char* name,
                                                 double* out, size_t n, size_t* count) {
    return guarded([&] {
        require(path && object && name && count, "null argument");
        h5::Obj f = h5::open_read(path);
        h5::Obj o = h5::open_object(f.id(), object);
        std::vector<double> v;
        if (!h5::read_attr_doubles(o.id(), name, v)) throw Error(GFX_ERR_IO, std::string("no attribute ") + name);
        *count = v.size();
        if (out) {
            require(n >= v.size(), "output buffer too small");
            std::copy(v.begin(), v.end(), out);
        }
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
            H->coef_strain = std::move(p.coef_strain);
            H->coef_abar = std::move(p.coef_abar);
        }
        if (p.pml_K.size() == P * 3 && p.pml_d.size() == P * 3 && p.pml_alpha.size() == P * 3) {
            H->pml_K = std::move(p.pml_K);
            H->pml_d = std::move(p.pml_d);
            H->pml_alpha = std::move(p.pml_alpha);
        }
        H->rec_cell = std::move(p.recording.rec_cell_local);
        H->tau_sigma = std::move(p.tau_sigma);
        H->tau_epsilon = std::move(p.tau_epsilon);
        H->node_xyz.assign(size_t(p.n_rank_node) * 3, 0.0);
        for (size_t q = 0; q < P; ++q)
            for (int d = 0; d < 3; ++d) H->node_xyz[size_t(H->ibool[q]) * 3 + d] = H->coords[q * 3 + d];
        H->spec.order = ng - 1;
        H->spec.pml_xmin = cfg.pml_xmin;
        H->spec.pml_xmax = cfg.pml_xmax;
        H->spec.pml_ymin = cfg.pml_ymin;
        H->spec.pml_ymax = cfg.pml_ymax;
        H->spec.pml_zmax = cfg.pml_zmax;
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
        m.n_sls = H->tau_sigma.empty() ? 0 : p.n_sls;
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
