// gf_greens_hip -- Green's function tiles and GMSH conversion from the command line.
//
//   gf_greens_hip [model.h5] [-o DIR] [--assembly {auto,colored,gather,atomic}]
//                 [--no-uva] [--ring N] [--workdir DIR] [--quiet] [--cpml-complete]
//       the reference's three forward runs plus its post-processor
//       (postprocess/cpp/main.cpp:39-89 command line: model.h5, output
//       directory -o, default greenfun) as one device run over config.h5 +
//       partitions/ of the working directory; writes DIR/tile_x{tx}_y{ty}.h5.
//       model.h5 (optional) supplies the domain bounds and cell mass.
//   gf_greens_hip --gmsh input.msh [-o model.h5] [--aux aux.h5]
//       tools/gmsh_to_hdf5.py: GMSH hexahedra to model.h5 /topology.
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>

#include "gfx.h"
#include "gfx_io.h"

static void usage(const char* prog) {
    std::cerr << "Usage: " << prog << " [model.h5] [-o DIR] [--assembly {auto,colored,gather,atomic}] [--no-uva]\n"
              << "                 [--ring N] [--workdir DIR] [--quiet] [--cpml-complete]\n"
              << "       " << prog << " --gmsh input.msh [-o model.h5] [--aux aux.h5]\n"
              << "  Input:  config.h5, partitions/partition_{r}.h5 (working directory)\n"
              << "  Output: DIR/tile_x{tx:03d}_y{ty:03d}.h5 (default DIR: greenfun)\n";
}

int main(int argc, char** argv) {
    gfx_green_tile_options o;
    gfx_green_tile_options_default(&o);
    o.verbose = 1;
    std::string model, out, workdir, gmsh, aux;
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
        if (!std::strcmp(a, "-o") || !std::strcmp(a, "--output")) {
            out = next(a);
        } else if (!std::strcmp(a, "--gmsh")) {
            gmsh = next(a);
        } else if (!std::strcmp(a, "--aux")) {
            aux = next(a);
        } else if (!std::strcmp(a, "--assembly")) {
            const std::string v = next(a);
            if (v == "auto") {
                o.assembly = -1;
            } else if (v == "colored" || v == "gather" || v == "atomic") {
                o.assembly = v == "colored" ? GFX_ASSEMBLY_COLORED
                             : v == "gather" ? GFX_ASSEMBLY_GATHER
                                             : GFX_ASSEMBLY_ATOMIC;
            } else {
                std::cerr << "Error: --assembly must be auto, colored, gather or atomic\n";
                return 1;
            }
        } else if (!std::strcmp(a, "--no-uva")) {
            o.with_uva = 0;
        } else if (!std::strcmp(a, "--ring")) {
            o.ring = std::atoi(next(a));
        } else if (!std::strcmp(a, "--workdir")) {
            workdir = next(a);
        } else if (!std::strcmp(a, "--quiet")) {
            o.verbose = 0;
        } else if (!std::strcmp(a, "--cpml-complete")) {
            o.cpml_complete = 1;
        } else if (!std::strcmp(a, "-h") || !std::strcmp(a, "--help")) {
            usage(argv[0]);
            return 0;
        } else if (a[0] != '-' && model.empty()) {
            model = a;
        } else {
            std::cerr << "Error: unknown argument " << a << "\n";
            usage(argv[0]);
            return 1;
        }
    }
    if (!gmsh.empty()) {
        const std::string dst = out.empty() ? "model.h5" : out;
        if (gfx_gmsh_to_topology(gmsh.c_str(), dst.c_str(), aux.empty() ? nullptr : aux.c_str()) != GFX_OK) {
            std::cerr << "Error: " << gfx_last_error() << "\n";
            return 1;
        }
        return 0;
    }
    o.model_path = model.empty() ? nullptr : model.c_str();
    o.output_dir = out.empty() ? nullptr : out.c_str();
    o.workdir = workdir.empty() ? nullptr : workdir.c_str();
    int n_tiles = 0;
    if (gfx_run_green_tiles(&o, &n_tiles) != GFX_OK) {
        std::cerr << "Error: " << gfx_last_error() << "\n";
        return 1;
    }
    std::printf("STAT_NTILES=%d\n", n_tiles);
    return 0;
}
