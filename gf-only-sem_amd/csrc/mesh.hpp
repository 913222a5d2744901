// mesh.hpp -- owner of the arrays behind a gfx_mesh handle (generated box or
// inputs read back from partition files).
#pragma once

#include <cstdint>
#include <vector>

#include "gfx.h"

namespace gfx {

struct MeshHolder {
    gfx_mesh m{};  // first member: a gfx_mesh* is a MeshHolder*
    std::vector<double> coords, jacobian, dxi_dx, mass, vp, vs, density, lambda_, mu_, damping, node_xyz,
        coef_alpha, coef_beta, coef_abar, coef_strain, tau_sigma, tau_epsilon, pml_K, pml_d, pml_alpha;
    std::vector<int32_t> ibool, pml_region, rec_cell;
    gfx_box_spec spec{};
};

// Points the gfx_mesh fields at the holder's vectors (null for empty ones).
void wire_mesh(MeshHolder& H);

}  // namespace gfx
