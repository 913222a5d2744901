// green.hpp -- device-resident strain Green's tensor tile: values at the
// unique GLL nodes of a set of cells, [time][node][6 strain x 3 force dirs],
// the cells' node lists, and the map that forms each node value from the
// solver's element-local strain records (mass-weighted average over the
// element points at that node, postprocess/cpp/main.cpp:354-426).
#pragma once

#include "device.hpp"
#include "gfx.h"

struct gfx_green {
    int n_time = 0, n_node = 0, n_cell = 0;
    gfx::DevBuf<float> val;         // [n_time][n_node][18], index c*3 + d
    gfx::DevBuf<int32_t> cell_idx;  // [n_cell][125]
    gfx::DevBuf<double> xyz;        // [n_node][3]
    gfx::DevBuf<double> box;        // [n_cell][6]: xmin xmax ymin ymax zmin zmax (corner nodes)
    // cell search: uniform bins over the bounding box, each listing the cells
    // whose (slack-widened) box overlaps it
    int nb[3] = {1, 1, 1};
    double b0[3] = {0, 0, 0}, bh[3] = {1, 1, 1};
    gfx::DevBuf<int32_t> bin_off, bin_cell;
    // capture map: node n <- sum_e w_e * strain[rec_point_e] / sum_e w_e
    gfx::DevBuf<int64_t> map_off;   // [n_node + 1]
    gfx::DevBuf<int32_t> map_pt;    // recording point = rec cell position * 125 + local point
    gfx::DevBuf<float> map_w;       // normalised weights
    int64_t map_max = -1;           // largest recording point the map reads
    std::vector<double> h_xyz, h_box;
    std::vector<int32_t> h_cell;
    hipStream_t stream = nullptr;
    ~gfx_green() {
        if (stream) (void)hipStreamDestroy(stream);
    }
};
