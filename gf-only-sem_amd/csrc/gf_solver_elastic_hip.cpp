// gf_solver_elastic_hip -- command-line forward solver on the GPU.
//
//   gf_solver_elastic_hip --direction {x,y,z} [--resume]
//       the reference CLI (forward/elastic/src/main.cpp:1-7,129-147): reads
//       config.h5 + partitions/partition_{r}.h5 from the current directory,
//       writes wavefields/{dir}/record_0_{step}.h5 (+ restart_0.h5).
//   gf_solver_elastic_hip --directions x,y,z [--stations stations.h5] [--resume]
//       batched: one field per (station, direction), advanced together;
//       records under wavefields/s{k}/{dir}/ (or wavefields/{dir}/ without stations).
//   Options: --assembly {auto,colored,gather,atomic} --ring N --stop-after STEP
//            --restart-stride N --quiet --workdir DIR --cpml-complete
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>

#include "gfx.h"
#include "gfx_io.h"

static void usage(const char* prog) {
    std::cerr << "Usage: " << prog << " --direction {x,y,z} [--resume]\n"
              << "       " << prog << " --directions x,y,z [--stations FILE] [--resume]\n"
              << "  options: --assembly {auto,colored,gather,atomic} --ring N --stop-after STEP\n"
              << "           --restart-stride N --workdir DIR --quiet --cpml-complete\n"
              << "  All I/O paths are relative to the working directory:\n"
              << "    Input:  config.h5, partitions/partition_{r}.h5\n"
              << "    Output: wavefields/{direction}/record_0_{step}.h5\n"
              << "    Restart: wavefields/{direction}/restart_0.h5\n";
}

int main(int argc, char** argv) {
    gfx_forward_options o;
    gfx_forward_options_default(&o);
    o.verbose = 1;
    std::string directions, stations, workdir;
    bool single = false;
    for (int i = 1; i < argc; ++i) {
        const char* a = argv[i];
        auto next = [&](const char* name) -> const char* {
            if (i + 1 >= argc) {
                std::cerr << "Error: " << name << " needs a value\n";
                usage(argv[0]);
                std::exit(1);
            }
            return argv[++i];
        };
        if (!std::strcmp(a, "--direction")) {
            directions = next(a);
            single = true;
        } else if (!std::strcmp(a, "--directions")) {
            directions = next(a);
        } else if (!std::strcmp(a, "--stations")) {
            stations = next(a);
        } else if (!std::strcmp(a, "--resume")) {
            o.resume = 1;
        } else if (!std::strcmp(a, "--assembly")) {
            const std::string v = next(a);
            o.assembly = v == "colored" ? GFX_ASSEMBLY_COLORED
                         : v == "gather" ? GFX_ASSEMBLY_GATHER
                         : v == "atomic" ? GFX_ASSEMBLY_ATOMIC
                                         : -1;
        } else if (!std::strcmp(a, "--ring")) {
            o.ring = std::atoi(next(a));
        } else if (!std::strcmp(a, "--stop-after")) {
            o.stop_after_step = std::atoi(next(a));
        } else if (!std::strcmp(a, "--restart-stride")) {
            o.restart_stride = std::atoi(next(a));
        } else if (!std::strcmp(a, "--workdir")) {
            workdir = next(a);
        } else if (!std::strcmp(a, "--quiet")) {
            o.verbose = 0;
        } else if (!std::strcmp(a, "--cpml-complete")) {
            o.cpml_complete = 1;
        } else if (!std::strcmp(a, "--help") || !std::strcmp(a, "-h")) {
            usage(argv[0]);
            return 0;
        }
    }
    if (single && directions != "x" && directions != "y" && directions != "z") {
        std::cerr << "Error: --direction must be x, y, or z, got '" << directions << "'\n";
        usage(argv[0]);
        return 1;
    }
    if (directions.empty()) {
        std::cerr << "Error: --direction must be x, y, or z, got ''\n";
        usage(argv[0]);
        return 1;
    }
    if (gfx_device_count() < 1) {
        std::cerr << "Error: no HIP GPU found.\n";
        return 1;
    }
    o.directions = directions.c_str();
    o.stations = stations.empty() ? nullptr : stations.c_str();
    o.workdir = workdir.empty() ? nullptr : workdir.c_str();
    if (o.verbose) {
        std::cout << "gf_solver: direction=" << directions << (o.resume ? " (resume mode)" : "") << std::endl;
        std::cout << "  input:  config.h5 + partitions/partition_{r}.h5\n"
                  << "  output: wavefields/" << (stations.empty() ? "" : "s{k}/") << "{direction}/record_0_{step}.h5"
                  << std::endl;
    }
    const gfx_status st = gfx_run_forward_ex(&o);
    if (st != GFX_OK) {
        std::cerr << "Fatal error: " << gfx_last_error() << std::endl;
        return 1;
    }
    return 0;
}
