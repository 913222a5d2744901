// solver.hip -- device-resident elastic CG-SEM solver and its C ABI.
//
// Replaces the reference's CudaDeviceState and per-step wrappers
// (forward/share/include/gf/cuda_step.hpp:20-129, forward/share/src/cuda_step.cu)
// with one handle that owns a HIP stream, FP32 SoA device arrays and the
// launch schedule of one step:
//   [atomic mode: zero R] -> element kernel per colour -> special-node kernel
//   -> regular-node kernel -> (snapshot steps) record kernel.
// Host-side setup (colouring, node renumbering, C-PML assembly, FP64 -> FP32
// packing) happens once in gfx_solver_create.
#include <array>
#include <atomic>
#include <cmath>
#include <cstring>
#include <functional>
#include <memory>
#include <numeric>

#include "common.hpp"
#include "kernels.hpp"

#include "device.hpp"
#include "green.hpp"
#include "gfx_surrogate.h"

namespace gfx {

namespace {
// Solver elements [b, e) launched together: one colour (COLORED) or all
// elements, split into the axis-aligned part (diag) and the rest, and with the
// complete C-PML into PML elements (pml) and the others.
struct ElemRange {
    int b, e;
    bool diag;
    bool pml;
    int ax;  // complete C-PML: the axes (bit 0 x, 1 y, 2 z) with d != 0 at some point of the range's elements
};

struct FieldSource {
    int direction = -1;
    std::vector<int32_t> cells;  // caller element indices
    std::vector<float> weights;  // [n*125]
    std::vector<float> stf;
};
}  // namespace

}  // namespace gfx

struct gfx_solver {
    int E = 0, G = 0, nf = 1, n_rec = 0, n_pml_elem = 0;
    int64_t Ga = 0, Gpml = 0, G4 = 0, Ppml = 0;
    int64_t fstride = 0, astride = 0;
    int assembly = GFX_ASSEMBLY_COLORED, epb = 2, kernel = GFX_ELEMENT_WAVE;
    int node_fg = 1;  // field groups per node-kernel workgroup (1, 2, 4, 8)
    gfx::DevBuf<int> ctr;  // device step / snapshot counters
    bool use_graph = false;
    hipGraph_t graph = nullptr;
    hipGraphExec_t graph_exec = nullptr;
    int gen = 0;  // bumped whenever captured launches become stale
    double dt = 0;
    int stride = 1;
    bool has_cpml = false, pml_nodal = true;
    // complete C-PML: stretched stresses and inertia terms inside the PML
    // elements' own kernel (coefficients pcoef [28][Ppml], memories in pmem
    // [nf][45][Ppml]); the node kernels then apply no C-PML term
    bool cpml_complete = false;
    gfx::DevBuf<float> pcoef;
    int64_t cpml_mem_active = 0;  // memories the kernel reads and writes, over all PML points
    hipStream_t stream = nullptr;
    std::vector<int32_t> int_of_elem;        // caller element -> solver element
    std::vector<int32_t> node_new_of_old;    // caller node -> solver node
    std::vector<double> mass_node;           // assembled mass, solver numbering
    std::vector<gfx::ElemRange> ranges;      // solver element ranges launched together
    int n_colors = 1;
    int64_t n_diag = 0;                       // elements on the axis-aligned (diagonal) path
    gfx::DevBuf<float> G9, lam, mu, minv, damp;
    int n_sls = 0;                                     // shear attenuation mechanisms
    gfx::DevBuf<float> att_a, att_b, att_c, zeta, devp;
    gfx::DevBuf<uint32_t> ibool;
    // C-PML: nodal (pc, palpha, pmem, pnew over [0,Gpml)) or element-local (Ppml points)
    gfx::DevBuf<float> pc, palpha, pabar, pmem, pnew;
    gfx::DevBuf<uint8_t> pmask;  // nodal C-PML: memory directions in use (bits x, y, z)
    gfx::DevBuf<int32_t> pml_slot, src_off;
    gfx::DevBuf<float> U, W, R, Aacc, Rl;
    // GATHER assembly: node n owns slots [goff_host[n], goff_host[n+1]); slot q
    // sums the entry gpts_host[q] of Rl (increasing in solver order per node;
    // the first is the node's own index).  On the device goff / gpts list
    // every slot but the first, which the node kernels read at the node id.
    gfx::DevBuf<int32_t> goff, gpts;
    std::vector<int32_t> goff_host, gpts_host;
    // entry of each element point in Rl ([P]) and entries per field (G4 + P - G)
    gfx::DevBuf<int32_t> rdst;
    int64_t PR = 0;
    std::vector<gfx::FieldSource> sources;
    gfx::DevBuf<int4> src_ent;
    gfx::DevBuf<float> src_w, stf;
    int nstf = 0;
    // recording
    std::vector<int32_t> rec_cells;
    gfx::DevBuf<int32_t> rec_ibool;
    gfx::DevBuf<float> rec_dd, rec_strain, rec_uva;
    int rec_cap = 0, n_snap = 0;
    bool rec_uva_on = false;
    float D[25];
    int step = 0;
    // per-stage entry points: position inside the current step (0 = between
    // steps) and whether the ATOMIC residual was cleared for it; the source
    // cells (solver numbering) and weights of the last injection
    int stage = 0;
    bool stage_zeroed = false;
    // GATHER: R (not the slot sums of Rl) holds the last step's assembled
    // residual, as after a step driven stage by stage (the exchange may have
    // changed it after the scatter)
    bool r_in_R = false;
    gfx::DevBuf<int32_t> inj_cells;
    gfx::DevBuf<float> inj_w;
    // timing
    bool timing = false;
    double t_elem = 0, t_elem_pml = 0, t_spec = 0, t_reg = 0, t_rec = 0, t_total = 0;
    int n_elem_launch = 0, n_elem_pml_launch = 0, n_spec_launch = 0, n_reg_launch = 0, n_rec_launch = 0,
        n_total_launch = 0;
    // complete C-PML: index of the first range when the PML ranges close the
    // element launches (one colour), else -1 (their time is then not split out)
    int pml_tail = -1;

    ~gfx_solver() {
        if (graph_exec) (void)hipGraphExecDestroy(graph_exec);
        if (graph) (void)hipGraphDestroy(graph);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace gfx {
namespace {

constexpr int NP = dev::NP;

void build_sources(gfx_solver& S) {
    std::vector<std::vector<int4>> per_elem(S.E);
    std::vector<float> w;
    int nstf = 0;
    for (int f = 0; f < S.nf; ++f) {
        const FieldSource& fs = S.sources[f];
        if (fs.direction < 0) continue;
        nstf = std::max(nstf, int(fs.stf.size()));
        for (size_t c = 0; c < fs.cells.size(); ++c) {
            const int e = fs.cells[c];
            if (e < 0) continue;  // source cell not on this rank (solver.cpp:585-587)
            const int ie = S.int_of_elem[e];
            const int off = int(w.size());
            w.insert(w.end(), fs.weights.begin() + c * NP, fs.weights.begin() + (c + 1) * NP);
            per_elem[ie].push_back(make_int4(f, fs.direction, off, 0));
        }
    }
    std::vector<int32_t> off(S.E + 1, 0);
    std::vector<int4> ent;
    for (int e = 0; e < S.E; ++e) {
        off[e] = int32_t(ent.size());
        for (const int4& q : per_elem[e]) ent.push_back(q);
    }
    off[S.E] = int32_t(ent.size());
    const int ns = std::max(nstf, 1);
    std::vector<float> stf(size_t(ns) * S.nf, 0.f);
    for (int f = 0; f < S.nf; ++f)
        for (size_t t = 0; t < S.sources[f].stf.size(); ++t) stf[size_t(f) * ns + t] = S.sources[f].stf[t];
    if (ent.empty()) ent.push_back(make_int4(-1, 0, 0, 0));
    if (w.empty()) w.push_back(0.f);
    S.src_off.upload(off);
    S.src_ent.upload(ent);
    S.src_w.upload(w);
    S.stf.upload(stf);
    S.nstf = nstf;
}

// conv(g, p)(t) = int_0^t exp(-p (t - s)) g(s) ds over one step, for g linear
// in time between the step's two levels: M = Q + c g, then Q <- a M + b g, with
// a = exp(-p dt), b = I0 - I1/dt, c = I1/dt, I0 = (1 - a)/p, I1 = (dt - I0)/p.
void cpml_recursion(double p, double dt, double* abc) {
    const double x = p * dt;
    double i0, c;  // I0 / dt and I1 / dt^2
    if (std::abs(x) < 1e-2) {  // sum_k (-x)^k / (k+1)! and sum_k (-x)^k / (k+2)!
        double term = 1.0, f1 = 1.0, f2 = 2.0;
        i0 = 0.0;
        c = 0.0;
        for (int k = 0; k < 8; ++k) {
            i0 += term / f1;
            c += term / f2;
            term *= -x;
            f1 *= k + 2;
            f2 *= k + 3;
        }
    } else {
        const double em = std::expm1(-x);
        i0 = -em / x;
        c = (x + em) / (x * x);
    }
    abc[0] = std::exp(-x);
    abc[1] = dt * (i0 - c);
    abc[2] = dt * c;
}

// The 27 complete C-PML coefficients of one point: per axis K, d, alpha, the
// recursion of alpha and that of beta = alpha + d / K (both zero on an axis
// with d = 0, whose memories then stay zero).  False on an invalid profile.
bool cpml_point_coefficients(const double* K, const double* d, const double* al, double dt, double* c) {
    bool ok = true;
    for (int q = 0; q < 3; ++q) {
        double* cq = c + q * 9;
        const bool good = std::isfinite(K[q]) && K[q] > 0.0 && std::isfinite(d[q]) && d[q] >= 0.0 &&
                          std::isfinite(al[q]) && al[q] >= 0.0;
        ok = ok && good;
        cq[0] = K[q];
        cq[1] = d[q];
        cq[2] = al[q];
        for (int t = 3; t < 9; ++t) cq[t] = 0.0;
        if (good && d[q] != 0.0) {
            cpml_recursion(al[q], dt, cq + 3);
            cpml_recursion(al[q] + d[q] / K[q], dt, cq + 6);
        }
    }
    return ok;
}

// Nodes of the complete C-PML's outer boundary, held rigid: the faces of PML
// elements that no other element shares and whose normal (the largest
// component of grad xi_a at the face centre) is an axis the element's region
// damps (region codes of preprocess/pml_cpml.py:27-34).  The free surface and
// the PML's interior interface are therefore not included.
std::vector<char> cpml_rigid_nodes(const gfx_solver_desc& d, int E, int G) {
    auto damps = [](int region, int axis) {
        static const int mask[8] = {0, 1, 2, 4, 3, 5, 6, 7};  // bit 0 x, 1 y, 2 z
        return region >= 1 && region <= 7 && (mask[region] >> axis & 1);
    };
    auto pidx = [](int a, int side, int s, int t) {  // point of face (a, side) at in-face indices (s, t)
        const int f = side ? 4 : 0;
        return a == 0 ? (f * 5 + s) * 5 + t : a == 1 ? (s * 5 + f) * 5 + t : (s * 5 + t) * 5 + f;
    };
    struct Face {
        std::array<int32_t, 4> key;
        int32_t e;
        int8_t f;
    };
    std::vector<Face> faces;
    faces.reserve(size_t(E) * 6);
    for (int e = 0; e < E; ++e)
        for (int f = 0; f < 6; ++f) {
            Face q{};
            for (int c = 0; c < 4; ++c)
                q.key[c] = d.ibool[int64_t(e) * NP + pidx(f / 2, f % 2, (c & 1) * 4, (c >> 1) * 4)];
            std::sort(q.key.begin(), q.key.end());
            q.e = e;
            q.f = int8_t(f);
            faces.push_back(q);
        }
    std::sort(faces.begin(), faces.end(), [](const Face& a, const Face& b) { return a.key < b.key; });
    std::vector<char> rigid(G, 0);
    for (size_t i = 0; i < faces.size();) {
        size_t j = i + 1;
        while (j < faces.size() && faces[j].key == faces[i].key) ++j;
        const Face& q = faces[i];
        const int region = d.pml_region[q.e];
        if (j == i + 1 && region != 0) {
            const int a = q.f / 2, side = q.f % 2;
            const double* g = d.dxi_dx + (int64_t(q.e) * NP + pidx(a, side, 2, 2)) * 9;
            int axis = 0;
            for (int m = 1; m < 3; ++m)
                if (std::abs(g[m * 3 + a]) > std::abs(g[axis * 3 + a])) axis = m;
            if (damps(region, axis))
                for (int s2 = 0; s2 < 5; ++s2)
                    for (int t = 0; t < 5; ++t) rigid[d.ibool[int64_t(q.e) * NP + pidx(a, side, s2, t)]] = 1;
        }
        i = j;
    }
    return rigid;
}

void init_state(gfx_solver& S) {
    S.U.zero(S.stream);
    S.W.zero(S.stream);
    S.R.zero(S.stream);
    S.Aacc.zero(S.stream);
    S.Rl.zero(S.stream);
    S.pmem.zero(S.stream);
    S.pnew.zero(S.stream);
    S.zeta.zero(S.stream);
    S.devp.zero(S.stream);
    S.ctr.zero(S.stream);
    GFX_HIP(hipStreamSynchronize(S.stream));
    S.step = 0;
    S.n_snap = 0;
    S.stage = 0;
    S.stage_zeroed = false;
    S.r_in_R = false;
}

// Launchers.  DIAG selects the axis-aligned geometry form for a range whose
// elements all have diagonal dxi/dx.
template <int EPB, bool PE, bool DIAG>
void launch_element_t(gfx_solver& S, const dev::ElemArgs& a, int b, int n) {
    constexpr int NT = ((EPB * NP + 63) / 64) * 64;
    const int grid = (n + EPB - 1) / EPB;
    if (grid == 0) return;
    switch (S.assembly) {
        case GFX_ASSEMBLY_ATOMIC:
            dev::k_element<EPB, dev::ASM_ATOMIC, PE, DIAG><<<grid, NT, 0, S.stream>>>(a, b, n);
            break;
        case GFX_ASSEMBLY_GATHER:
            dev::k_element<EPB, dev::ASM_GATHER, PE, DIAG><<<grid, NT, 0, S.stream>>>(a, b, n);
            break;
        default:
            dev::k_element<EPB, dev::ASM_COLORED, PE, DIAG><<<grid, NT, 0, S.stream>>>(a, b, n);
    }
}

template <bool PE, bool DIAG>
void launch_element_p(gfx_solver& S, const dev::ElemArgs& a, int b, int n) {
    switch (S.epb) {
        case 1: launch_element_t<1, PE, DIAG>(S, a, b, n); break;
        case 2: launch_element_t<2, PE, DIAG>(S, a, b, n); break;
        case 4: launch_element_t<4, PE, DIAG>(S, a, b, n); break;
        default: launch_element_t<8, PE, DIAG>(S, a, b, n); break;
    }
}

template <int WPB, bool DIAG>
void launch_wave_t(gfx_solver& S, const dev::ElemArgs& a, int b, int n) {
    const int grid = (n + WPB - 1) / WPB;
    if (grid == 0) return;
    switch (S.assembly) {
        case GFX_ASSEMBLY_ATOMIC:
            dev::k_element_wave<WPB, dev::ASM_ATOMIC, DIAG><<<grid, 64 * WPB, 0, S.stream>>>(a, b, n);
            break;
        case GFX_ASSEMBLY_GATHER:
            dev::k_element_wave<WPB, dev::ASM_GATHER, DIAG><<<grid, 64 * WPB, 0, S.stream>>>(a, b, n);
            break;
        default:
            dev::k_element_wave<WPB, dev::ASM_COLORED, DIAG><<<grid, 64 * WPB, 0, S.stream>>>(a, b, n);
    }
}

template <int EPB, bool DIAG>
void launch_pair_t(gfx_solver& S, const dev::ElemArgs& a, int b, int n) {
    constexpr int NT = ((EPB * NP + 63) / 64) * 64;
    const int grid = (n + EPB - 1) / EPB;
    if (grid == 0) return;
    switch (S.assembly) {
        case GFX_ASSEMBLY_ATOMIC:
            dev::k_element_pair<EPB, dev::ASM_ATOMIC, DIAG><<<grid, NT, 0, S.stream>>>(a, b, n);
            break;
        case GFX_ASSEMBLY_GATHER:
            dev::k_element_pair<EPB, dev::ASM_GATHER, DIAG><<<grid, NT, 0, S.stream>>>(a, b, n);
            break;
        default:
            dev::k_element_pair<EPB, dev::ASM_COLORED, DIAG><<<grid, NT, 0, S.stream>>>(a, b, n);
    }
}

template <bool DIAG>
void launch_pencil(gfx_solver& S, const dev::ElemArgs& a, int b, int n) {
    const int grid = (n + dev::PENCIL_EPB - 1) / dev::PENCIL_EPB;
    if (grid == 0) return;
    switch (S.assembly) {
        case GFX_ASSEMBLY_ATOMIC:
            dev::k_element_pencil<dev::ASM_ATOMIC, DIAG><<<grid, 256, 0, S.stream>>>(a, b, n);
            break;
        case GFX_ASSEMBLY_GATHER:
            dev::k_element_pencil<dev::ASM_GATHER, DIAG><<<grid, 256, 0, S.stream>>>(a, b, n);
            break;
        default:
            dev::k_element_pencil<dev::ASM_COLORED, DIAG><<<grid, 256, 0, S.stream>>>(a, b, n);
    }
}

// Element kernel form for a request: AUTO = wave when the nodal C-PML form
// applies, else block; WAVE, PAIR and PENCIL need the nodal form.
int resolve_kernel(int requested, bool nodal_ok) {
    require(requested >= GFX_ELEMENT_AUTO && requested <= GFX_ELEMENT_PENCIL, "unknown element kernel");
    if (requested != GFX_ELEMENT_AUTO && requested != GFX_ELEMENT_BLOCK)
        require(nodal_ok, "the wave, pair and pencil element kernels need the nodal C-PML form");
    if (requested == GFX_ELEMENT_AUTO) return nodal_ok ? GFX_ELEMENT_WAVE : GFX_ELEMENT_BLOCK;
    return requested;
}

void check_epb(int kernel, int epb) {
    require(epb == 0 || epb == 1 || epb == 2 || epb == 4 || epb == 8, "elems_per_block must be 1, 2, 4 or 8");
    require(!(kernel == GFX_ELEMENT_PAIR && epb == 8), "the pair element kernel takes 1, 2 or 4 elements per block");
    require(!(kernel == GFX_ELEMENT_PENCIL && epb != 0),
            "the pencil element kernel has a fixed ten elements per block (elems_per_block must be 0)");
}

// Elements per workgroup reported and used for a kernel form and request.
int resolve_epb(int kernel, int epb) {
    if (kernel == GFX_ELEMENT_PENCIL) return dev::PENCIL_EPB;
    return epb ? epb : (kernel == GFX_ELEMENT_WAVE ? 4 : 2);
}

template <bool DIAG>
void launch_element_d(gfx_solver& S, const dev::ElemArgs& a, int b, int n) {
    if (S.kernel == GFX_ELEMENT_PENCIL) {
        launch_pencil<DIAG>(S, a, b, n);
    } else if (S.kernel == GFX_ELEMENT_PAIR) {
        switch (S.epb) {
            case 1: launch_pair_t<1, DIAG>(S, a, b, n); break;
            case 4: launch_pair_t<4, DIAG>(S, a, b, n); break;
            default: launch_pair_t<2, DIAG>(S, a, b, n); break;
        }
    } else if (S.kernel == GFX_ELEMENT_WAVE) {
        switch (S.epb) {
            case 1: launch_wave_t<1, DIAG>(S, a, b, n); break;
            case 2: launch_wave_t<2, DIAG>(S, a, b, n); break;
            case 4: launch_wave_t<4, DIAG>(S, a, b, n); break;
            default: launch_wave_t<8, DIAG>(S, a, b, n); break;
        }
    } else {
        launch_element_p<false, DIAG>(S, a, b, n);
    }
}

// The reference's C-PML in its element-local form (the nodal form does not apply).
bool ref_elem_pml(const gfx_solver& S) { return S.has_cpml && !S.pml_nodal && !S.cpml_complete; }

template <int MODE>
void launch_att(gfx_solver& S, const dev::ElemArgs& a, int b, int n) {
    switch (S.n_sls) {
        case 1: dev::k_element_att<MODE, 1><<<n, 128, 0, S.stream>>>(a, b, n); break;
        case 2: dev::k_element_att<MODE, 2><<<n, 128, 0, S.stream>>>(a, b, n); break;
        case 3: dev::k_element_att<MODE, 3><<<n, 128, 0, S.stream>>>(a, b, n); break;
        case 4: dev::k_element_att<MODE, 4><<<n, 128, 0, S.stream>>>(a, b, n); break;
        default: dev::k_element_att<MODE, 0><<<n, 128, 0, S.stream>>>(a, b, n); break;
    }
}

template <int MODE>
void launch_cpml(gfx_solver& S, const dev::ElemArgs& a, int ax, int b, int n) {
    switch (ax) {
        case 0: dev::k_element_cpml<MODE, 0><<<n, 128, 0, S.stream>>>(a, b, n); break;
        case 1: dev::k_element_cpml<MODE, 1><<<n, 128, 0, S.stream>>>(a, b, n); break;
        case 2: dev::k_element_cpml<MODE, 2><<<n, 128, 0, S.stream>>>(a, b, n); break;
        case 3: dev::k_element_cpml<MODE, 3><<<n, 128, 0, S.stream>>>(a, b, n); break;
        case 4: dev::k_element_cpml<MODE, 4><<<n, 128, 0, S.stream>>>(a, b, n); break;
        case 5: dev::k_element_cpml<MODE, 5><<<n, 128, 0, S.stream>>>(a, b, n); break;
        case 6: dev::k_element_cpml<MODE, 6><<<n, 128, 0, S.stream>>>(a, b, n); break;
        default: dev::k_element_cpml<MODE, 7><<<n, 128, 0, S.stream>>>(a, b, n); break;
    }
}

void launch_element(gfx_solver& S, const dev::ElemArgs& a, const ElemRange& rg) {
    const int b = rg.b, n = rg.e - rg.b;
    if (S.n_sls > 0 && !rg.pml) {  // shear attenuation: one kernel form for every range
        if (n > 0) {
            switch (S.assembly) {
                case GFX_ASSEMBLY_GATHER: launch_att<dev::ASM_GATHER>(S, a, b, n); break;
                case GFX_ASSEMBLY_ATOMIC: launch_att<dev::ASM_ATOMIC>(S, a, b, n); break;
                default: launch_att<dev::ASM_COLORED>(S, a, b, n); break;
            }
        }
        return;
    }
    if (rg.pml) {  // complete C-PML: one element per workgroup, general form (elastic)
        if (n > 0) {
            switch (S.assembly) {
                case GFX_ASSEMBLY_GATHER: launch_cpml<dev::ASM_GATHER>(S, a, rg.ax, b, n); break;
                case GFX_ASSEMBLY_ATOMIC: launch_cpml<dev::ASM_ATOMIC>(S, a, rg.ax, b, n); break;
                default: launch_cpml<dev::ASM_COLORED>(S, a, rg.ax, b, n); break;
            }
        }
    } else if (ref_elem_pml(S)) {
        launch_element_p<true, false>(S, a, b, n);  // element-local C-PML: block kernel, general form
    } else if (rg.diag) {
        launch_element_d<true>(S, a, b, n);
    } else {
        launch_element_d<false>(S, a, b, n);
    }
}

void launch_node_special(gfx_solver& S, const dev::NodeArgs& na, int grid) {
    switch (S.node_fg) {
        case 2: dev::k_node_special<2><<<grid, 256, 0, S.stream>>>(na); break;
        case 4: dev::k_node_special<4><<<grid, 256, 0, S.stream>>>(na); break;
        case 8: dev::k_node_special<8><<<grid, 256, 0, S.stream>>>(na); break;
        default: dev::k_node_special<1><<<grid, 256, 0, S.stream>>>(na); break;
    }
}

template <bool GATHER>
void launch_node_regular_t(gfx_solver& S, const dev::NodeArgs& na, int grid) {
    switch (S.node_fg) {
        case 2: dev::k_node_regular<GATHER, 2><<<grid, 256, 0, S.stream>>>(na); break;
        case 4: dev::k_node_regular<GATHER, 4><<<grid, 256, 0, S.stream>>>(na); break;
        case 8: dev::k_node_regular<GATHER, 8><<<grid, 256, 0, S.stream>>>(na); break;
        default: dev::k_node_regular<GATHER, 1><<<grid, 256, 0, S.stream>>>(na); break;
    }
}

void launch_node_regular(gfx_solver& S, const dev::NodeArgs& na, int grid) {
    if (na.gather)
        launch_node_regular_t<true>(S, na, grid);
    else
        launch_node_regular_t<false>(S, na, grid);
}

struct StepArgs {
    dev::ElemArgs ea{};
    dev::NodeArgs na{};
    dev::RecArgs ra{};
    int spec_grid = 0, reg_grid = 0, zero_grid = 0, rec_grid = 0, rec_threads = 0;
    int64_t r4 = 0;
    bool recording = false;
};
constexpr int REPB = 2;  // recording cells per workgroup

StepArgs make_step_args(gfx_solver& S) {
    StepArgs a;
    const double theta = 1.0 / 8.0;  // pml.cpp:35, coefficients pml.cpp:63-64
    const float c1 = float((1.0 - 2.0 * theta) * 0.5 * S.dt);
    const float c2 = float((1.0 - theta) * 0.5 * S.dt * S.dt);

    dev::ElemArgs& ea = a.ea;
    ea.G = S.G9.p;
    ea.lam = S.lam.p;
    ea.mu = S.mu.p;
    ea.ibool = S.ibool.p;
    ea.P = int64_t(S.E) * NP;
    ea.nf = S.nf;
    ea.fstride = S.fstride;
    ea.U = S.U.p;
    ea.R = S.R.p;
    ea.Rl = S.Rl.p;
    ea.rdst = S.rdst.p;
    ea.PR = S.PR;
    ea.W = S.W.p;
    ea.Aacc = S.Aacc.p;
    ea.astride = S.astride;
    ea.pml_slot = S.pml_slot.p;
    ea.Ppml = S.Ppml;
    ea.palpha = S.palpha.p;
    ea.pabar = S.pabar.p;
    ea.pmem = S.pmem.p;
    ea.pnew = S.pnew.p;
    ea.pcoef = S.pcoef.p;
    ea.c1 = c1;
    ea.c2 = c2;
    ea.dt = float(S.dt);
    ea.src_off = S.src_off.p;
    ea.src_ent = S.src_ent.p;
    ea.src_w = S.src_w.p;
    ea.stf = S.stf.p;
    ea.nstf = S.nstf;
    ea.ctr = S.ctr.p;
    ea.n_sls = S.n_sls;
    ea.att_a = S.att_a.p;
    ea.att_b = S.att_b.p;
    ea.att_c = S.att_c.p;
    ea.zeta = S.zeta.p;
    ea.devp = S.devp.p;
    std::memcpy(ea.D, S.D, sizeof(S.D));

    dev::NodeArgs& na = a.na;
    na.G = S.G;
    na.Ga = S.Ga;
    na.Gpml = S.Gpml;
    na.G4 = S.G4;
    na.P = int64_t(S.E) * NP;
    na.nf = S.nf;
    na.fstride = S.fstride;
    na.astride = S.astride;
    na.gather = (S.assembly == GFX_ASSEMBLY_GATHER) ? 1 : 0;
    na.Rl = S.Rl.p;
    na.PR = S.PR;
    na.off = S.goff.p;
    na.pts = S.gpts.p;
    na.minv = S.minv.p;
    na.damp = S.has_cpml ? nullptr : S.damp.p;
    na.U = S.U.p;
    na.W = S.W.p;
    na.R = S.R.p;
    na.Aacc = S.Aacc.p;
    na.pml_nodal = (S.has_cpml && S.pml_nodal) ? 1 : 0;
    na.pc = S.pc.p;
    na.palpha = S.palpha.p;
    na.pmask = S.pmask.p;
    na.pmem = S.pmem.p;
    na.pnew = S.pnew.p;
    na.c1 = c1;
    na.c2 = c2;
    na.dt = float(S.dt);
    const int fg = S.node_fg;
    a.spec_grid = int((S.Ga + 256 / fg - 1) / (256 / fg));
    a.reg_grid = int(((S.G4 - S.Ga) / 4 + 256 / fg - 1) / (256 / fg));
    a.r4 = int64_t(S.nf) * S.fstride / 4;
    a.zero_grid = int((a.r4 + 255) / 256);

    dev::RecArgs& ra = a.ra;
    ra.ctr = S.ctr.p;
    ra.stride = S.stride;
    ra.cap = S.rec_cap;
    ra.n_rec = S.n_rec;
    ra.ibool = S.rec_ibool.p;
    ra.dd = S.rec_dd.p;
    ra.nf = S.nf;
    ra.fstride = S.fstride;
    ra.astride = S.astride;
    ra.U = S.U.p;
    ra.W = S.W.p;
    ra.Aacc = S.Aacc.p;
    ra.dt = float(S.dt);
    ra.strain = S.rec_strain.p;
    ra.uva = S.rec_uva_on ? S.rec_uva.p : nullptr;
    std::memcpy(ra.D, S.D, sizeof(S.D));
    a.recording = S.stride > 0 && S.rec_cap > 0 && S.n_rec > 0;
    a.rec_grid = (S.n_rec + REPB - 1) / REPB;
    a.rec_threads = ((REPB * NP + 63) / 64) * 64;
    return a;
}

// Enqueues one whole step on S.stream.  The step index and snapshot counter
// live on the device (S.ctr), so the same enqueue can be captured once into a
// graph and replayed.  `record` launches the record kernel (which itself does
// nothing off snapshot steps); ev, when given, receives 5 events: before the
// complete C-PML element ranges when they close the element launches (else
// with the next), after the elements, the special nodes, the regular nodes
// and the end of the step.
void enqueue_step(gfx_solver& S, const StepArgs& a, bool record, hipEvent_t* ev) {
    if (S.assembly == GFX_ASSEMBLY_ATOMIC)
        dev::k_zero<<<a.zero_grid, 256, 0, S.stream>>>(reinterpret_cast<float4*>(S.R.p), a.r4);
    const int nr = int(S.ranges.size());
    const int pml_tail = S.pml_tail >= 0 ? S.pml_tail : nr;
    for (int i = 0; i < nr; ++i) {
        if (ev && i == pml_tail) GFX_HIP(hipEventRecord(ev[0], S.stream));
        launch_element(S, a.ea, S.ranges[i]);
    }
    if (ev && pml_tail == nr) GFX_HIP(hipEventRecord(ev[0], S.stream));
    GFX_HIP(hipGetLastError());
    if (ev) GFX_HIP(hipEventRecord(ev[1], S.stream));
    if (a.spec_grid > 0) launch_node_special(S, a.na, a.spec_grid);
    if (ev) GFX_HIP(hipEventRecord(ev[2], S.stream));
    if (a.reg_grid > 0) launch_node_regular(S, a.na, a.reg_grid);
    GFX_HIP(hipGetLastError());
    if (ev) GFX_HIP(hipEventRecord(ev[3], S.stream));
    if (record && a.recording && a.rec_grid > 0)
        dev::k_record<REPB><<<a.rec_grid, a.rec_threads, 0, S.stream>>>(a.ra);
    dev::k_advance<<<1, 64, 0, S.stream>>>(S.ctr.p, S.stride);
    GFX_HIP(hipGetLastError());
    if (ev) GFX_HIP(hipEventRecord(ev[4], S.stream));
}

// Host mirror of k_advance.
void advance_host(gfx_solver& S) {
    if (S.stride > 0 && S.step % S.stride == 0) ++S.n_snap;
    ++S.step;
}

void drop_graph(gfx_solver& S) {
    ++S.gen;
    if (S.graph_exec) (void)hipGraphExecDestroy(S.graph_exec);
    if (S.graph) (void)hipGraphDestroy(S.graph);
    S.graph_exec = nullptr;
    S.graph = nullptr;
}

// One step captured into a graph (record kernel included; it checks the
// snapshot condition on the device).  Recaptured after any launch change.
void ensure_graph(gfx_solver& S) {
    if (S.graph_exec) return;
    const StepArgs a = make_step_args(S);
    GFX_HIP(hipStreamBeginCapture(S.stream, hipStreamCaptureModeThreadLocal));
    enqueue_step(S, a, true, nullptr);
    GFX_HIP(hipStreamEndCapture(S.stream, &S.graph));
    GFX_HIP(hipGraphInstantiate(&S.graph_exec, S.graph, nullptr, nullptr, 0));
}

void run_steps(gfx_solver& S, int nsteps) {
    if (nsteps > 0) S.r_in_R = false;
    if (S.use_graph && !S.timing) {
        ensure_graph(S);
        for (int t = 0; t < nsteps; ++t) {
            GFX_HIP(hipGraphLaunch(S.graph_exec, S.stream));
            advance_host(S);
        }
        GFX_HIP(hipStreamSynchronize(S.stream));
        return;
    }
    const StepArgs a = make_step_args(S);
    std::vector<hipEvent_t> ev;
    constexpr size_t EV = 5;
    if (S.timing) {
        // ev[0] = start; step t: ev[5t+1] before the complete C-PML ranges (when
        // they close the element launches), ev[5t+2] after elements, ev[5t+3]
        // after the special-node kernel, ev[5t+4] after the regular-node kernel,
        // ev[5t+5] after recording (= start of step t+1)
        ev.resize(size_t(nsteps) * EV + 1);
        for (auto& e : ev) GFX_HIP(hipEventCreate(&e));
        GFX_HIP(hipEventRecord(ev[0], S.stream));
    }
    std::vector<char> did_rec(nsteps, 0);
    for (int t = 0; t < nsteps; ++t) {
        const bool snap = a.recording && S.step % S.stride == 0;
        enqueue_step(S, a, snap, S.timing ? &ev[size_t(t) * EV + 1] : nullptr);
        did_rec[t] = snap ? 1 : 0;
        advance_host(S);
    }
    GFX_HIP(hipStreamSynchronize(S.stream));
    if (S.timing) {
        S.t_elem = S.t_elem_pml = S.t_spec = S.t_reg = S.t_rec = S.t_total = 0;
        S.n_elem_launch = S.n_elem_pml_launch = S.n_spec_launch = S.n_reg_launch = S.n_rec_launch = 0;
        const int n_pml_ranges = S.pml_tail >= 0 ? int(S.ranges.size()) - S.pml_tail : 0;
        for (int t = 0; t < nsteps; ++t) {
            float e = 0, ep = 0, b = 0, c = 0, d = 0;
            const size_t o = size_t(t) * EV;
            GFX_HIP(hipEventElapsedTime(&e, ev[o], ev[o + 2]));
            GFX_HIP(hipEventElapsedTime(&ep, ev[o + 1], ev[o + 2]));
            GFX_HIP(hipEventElapsedTime(&b, ev[o + 2], ev[o + 3]));
            GFX_HIP(hipEventElapsedTime(&c, ev[o + 3], ev[o + 4]));
            GFX_HIP(hipEventElapsedTime(&d, ev[o + 4], ev[o + 5]));
            S.t_elem += e;
            S.t_elem_pml += ep;
            S.t_spec += b;
            S.t_reg += c;
            S.t_rec += d;
            S.n_elem_launch += int(S.ranges.size());
            S.n_elem_pml_launch += n_pml_ranges;
            S.n_spec_launch += a.spec_grid > 0;
            S.n_reg_launch += a.reg_grid > 0;
            S.n_rec_launch += did_rec[t];
        }
        float tot = 0;
        GFX_HIP(hipEventElapsedTime(&tot, ev[0], ev[size_t(nsteps) * EV]));
        S.t_total = tot;
        S.n_total_launch = nsteps;
        for (auto& e : ev) (void)hipEventDestroy(e);
    }
}

// Writes the host step / snapshot counters to the device.
void upload_counters(gfx_solver& S) {
    const int h[2] = {S.step, S.n_snap};
    GFX_HIP(hipStreamSynchronize(S.stream));
    GFX_HIP(hipMemcpy(S.ctr.p, h, sizeof(h), hipMemcpyHostToDevice));
}

// Greedy element colouring: no two elements of one colour share a node.
// Elements are visited in caller order, so the result is deterministic.
std::vector<int> colour_elements(int E, const int32_t* ibool, int G) {
    std::vector<uint64_t> used(G, 0);
    std::vector<int> col(E, 0);
    for (int e = 0; e < E; ++e) {
        uint64_t m = 0;
        for (int p = 0; p < NP; ++p) m |= used[ibool[int64_t(e) * NP + p]];
        int c = 0;
        while (c < 64 && ((m >> c) & 1ULL)) ++c;
        require(c < 64, "element colouring needs more than 64 colours");
        col[e] = c;
        for (int p = 0; p < NP; ++p) used[ibool[int64_t(e) * NP + p]] |= (1ULL << c);
    }
    return col;
}

// Device allocations and copies of create(), queued while the host plan is
// built and run only once it is complete: every host-side check runs before
// the first device call (also on a machine without a device).
struct DeviceWork {
    std::vector<std::function<void()>> ops;
    template <typename T>
    void upload(DevBuf<T>& b, std::vector<T> h) {
        ops.emplace_back([&b, h = std::move(h)] { b.upload(h); });
    }
    template <typename T>
    void alloc(DevBuf<T>& b, size_t n) {
        ops.emplace_back([&b, n] { b.alloc(n); });
    }
    void run() {
        for (auto& f : ops) f();
        ops.clear();
    }
};

void create(const gfx_solver_desc& d, gfx_solver& S) {
    require(d.ngll == 5, "the device solver supports NGLL = 5 (polynomial order 4) only");
    require(d.n_elem > 0 && d.n_rank_node > 0, "empty mesh");
    require(d.dxi_dx && d.jacobian && d.lambda_ && d.mu_ && d.mass && d.ibool,
            "missing geometry/material arrays");
    require(d.n_fields >= 1 && d.n_fields <= 4096, "n_fields must be in [1, 4096]");
    require(d.dt > 0, "dt must be > 0");
    require(d.assembly == GFX_ASSEMBLY_COLORED || d.assembly == GFX_ASSEMBLY_ATOMIC ||
                d.assembly == GFX_ASSEMBLY_GATHER,
            "unknown assembly mode");
    require(d.element_kernel >= GFX_ELEMENT_AUTO && d.element_kernel <= GFX_ELEMENT_PENCIL,
            "unknown element kernel");
    int epb = d.elems_per_block;
    require(epb == 0 || epb == 1 || epb == 2 || epb == 4 || epb == 8,
            "elems_per_block must be 1, 2, 4 or 8");
    const int nfg = d.node_field_groups;
    require(nfg == 0 || nfg == 1 || nfg == 2 || nfg == 4 || nfg == 8, "node_field_groups must be 1, 2, 4 or 8");
    S.node_fg = nfg ? nfg : 1;
    require(d.n_rank_node < int(0x7fffffff) - 8, "too many nodes");
    DeviceWork dv;
    const int E = d.n_elem, G = d.n_rank_node;
    const int64_t P = int64_t(E) * NP;
    for (int64_t p = 0; p < P; ++p)
        require(d.ibool[p] >= 0 && d.ibool[p] < G, "ibool entry out of range");
    S.E = E;
    S.G = G;
    S.nf = d.n_fields;
    S.assembly = d.assembly;
    S.dt = d.dt;
    S.stride = d.snapshot_stride;
    S.has_cpml = d.has_cpml != 0;
    S.cpml_complete = d.cpml_complete != 0;
    if (S.cpml_complete) {
        require(S.has_cpml, "cpml_complete needs has_cpml");
        require(d.pml_region && d.pml_K && d.pml_d && d.pml_alpha,
                "cpml_complete needs pml_region, pml_K, pml_d and pml_alpha");
    } else if (S.has_cpml) {
        require(d.pml_region && d.coef_alpha && d.coef_abar && d.density,
                "has_cpml needs pml_region, coef_alpha, coef_abar and density");
    }

    const GLL gll(4);
    for (int q = 0; q < 25; ++q) S.D[q] = float(gll.D[q]);

    // ---- element order: Morton curve of the centroids (when given), colour-major
    //      for coloured assembly; caller order otherwise ----
    std::vector<uint64_t> key(E, 0);
    if (d.elem_centroid) {
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (int e = 0; e < E; ++e)
            for (int c = 0; c < 3; ++c) {
                lo[c] = std::min(lo[c], d.elem_centroid[e * 3 + c]);
                hi[c] = std::max(hi[c], d.elem_centroid[e * 3 + c]);
            }
        for (int e = 0; e < E; ++e) {
            uint64_t q[3];
            for (int c = 0; c < 3; ++c) {
                const double span = hi[c] > lo[c] ? hi[c] - lo[c] : 1.0;
                const double t = (d.elem_centroid[e * 3 + c] - lo[c]) / span;
                q[c] = uint64_t(std::min(std::max(t, 0.0), 1.0) * 2097151.0);
            }
            uint64_t m = 0;
            for (int b = 20; b >= 0; --b)
                for (int c = 2; c >= 0; --c) m = (m << 1) | ((q[c] >> b) & 1ULL);
            key[e] = m;
        }
    }
    // axis-aligned elements: dxi_a/dx_m = 0 for a != m at every point, up to FP64
    // round-off of the input (|off-diagonal| <= 1e-12 x the largest diagonal
    // entry); they run the diagonal geometry form of the element kernels
    std::vector<char> diag_el(E, 0);
    if (!d.general_geometry) {
        for (int e = 0; e < E; ++e) {
            bool ok = true;
            for (int p = 0; p < NP && ok; ++p) {
                const double* x = d.dxi_dx + (int64_t(e) * NP + p) * 9;
                const double dmax = std::max(std::fabs(x[0]), std::max(std::fabs(x[4]), std::fabs(x[8])));
                for (int c : {1, 2, 3, 5, 6, 7}) ok = ok && std::fabs(x[c]) <= 1e-12 * dmax;
            }
            diag_el[e] = ok;
        }
    }
    // complete C-PML: PML elements run their own kernel, after the others
    // and are grouped by the axes their stretch profiles damp (d != 0 at some
    // point), so that each launch carries only the memories those axes use
    std::vector<char> pml_el(E, 0), pml_ax(E, 0);
    if (S.cpml_complete)
        for (int e = 0; e < E; ++e) {
            pml_el[e] = d.pml_region[e] != 0;
            if (!pml_el[e] || !d.pml_d) continue;
            int ax = 0;
            for (int p = 0; p < NP; ++p)
                for (int q = 0; q < 3; ++q)
                    if (float(d.pml_d[(int64_t(e) * NP + p) * 3 + q]) != 0.f) ax |= 1 << q;
            pml_ax[e] = char(ax);
        }
    std::vector<int32_t> elem_of_int(E);
    std::iota(elem_of_int.begin(), elem_of_int.end(), 0);
    std::vector<int> col(E, 0);
    if (S.assembly == GFX_ASSEMBLY_COLORED) col = colour_elements(E, d.ibool, G);
    std::stable_sort(elem_of_int.begin(), elem_of_int.end(), [&](int a, int b) {
        if (col[a] != col[b]) return col[a] < col[b];
        if (pml_el[a] != pml_el[b]) return pml_el[a] < pml_el[b];
        if (pml_ax[a] != pml_ax[b]) return pml_ax[a] < pml_ax[b];
        if (diag_el[a] != diag_el[b]) return diag_el[a] > diag_el[b];
        return key[a] < key[b];
    });
    {
        int b = 0;
        for (int i = 1; i <= E; ++i) {
            const int e0 = elem_of_int[b];
            if (i == E || col[elem_of_int[i]] != col[e0] || pml_el[elem_of_int[i]] != pml_el[e0] ||
                pml_ax[elem_of_int[i]] != pml_ax[e0] || diag_el[elem_of_int[i]] != diag_el[e0]) {
                const bool pml = pml_el[e0] != 0;
                S.ranges.push_back({b, i, diag_el[e0] != 0 && !pml, pml, int(pml_ax[e0])});
                b = i;
            }
        }
        S.n_colors = S.assembly == GFX_ASSEMBLY_COLORED ? (*std::max_element(col.begin(), col.end()) + 1) : 1;
        S.n_diag = 0;
        for (const ElemRange& rg : S.ranges)
            if (rg.diag) S.n_diag += rg.e - rg.b;
        if (S.cpml_complete && S.assembly != GFX_ASSEMBLY_COLORED) {
            for (int i = 0; i < int(S.ranges.size()); ++i)
                if (S.ranges[i].pml) {
                    S.pml_tail = i;
                    break;
                }
        }
    }
    S.int_of_elem.assign(E, 0);
    for (int i = 0; i < E; ++i) S.int_of_elem[elem_of_int[i]] = i;

    // ---- nodal mass and legacy damping (solver.cpp:193-202), caller numbering ----
    // The complete C-PML weights the mass of a PML element point by
    // K_x K_y K_z + dt/2 c_v: its inertia term L has velocity coefficient
    // c_v = K_z (K_y d_x + K_x d_y) + K_x K_y d_z and takes the velocity at
    // v_{n+1} = w_n + dt/2 a_{n+1} (trapezoidal, stable for any damping); the
    // element kernel evaluates L with v = w_n and the dt/2 c_v a_{n+1} part
    // moves into the mass.
    std::vector<double> mg(G, 0.0), dg(G, 0.0);
    for (int64_t p = 0; p < P; ++p) {
        double m = d.mass[p];
        if (S.cpml_complete && d.pml_region[p / NP] != 0) {
            const double* K = d.pml_K + p * 3;
            const double* dd = d.pml_d + p * 3;
            const double cv = K[2] * (K[1] * dd[0] + K[0] * dd[1]) + K[0] * K[1] * dd[2];
            m *= K[0] * K[1] * K[2] + 0.5 * S.dt * cv;
        }
        mg[d.ibool[p]] += m;
        if (d.damping) dg[d.ibool[p]] = d.damping[p];  // last writer wins
    }
    // the complete C-PML ends in a rigid outer boundary: zero mass there, so
    // the massless-node rule keeps those nodes' u and v (zero from rest)
    if (S.cpml_complete) {
        const std::vector<char> rigid = cpml_rigid_nodes(d, E, G);
        for (int g = 0; g < G; ++g)
            if (rigid[g]) mg[g] = 0.0;
    }

    // ---- node classes: 2 = C-PML node, 1 = other special, 0 = regular ----
    // (the complete C-PML keeps its terms in the element kernel: its PML nodes
    // need no node-kernel memory and stay regular unless rigid or recorded)
    std::vector<char> cls(G, 0);
    if (S.has_cpml && !S.cpml_complete) {
        for (int e = 0; e < E; ++e)
            if (d.pml_region[e] != 0)
                for (int p = 0; p < NP; ++p) cls[d.ibool[int64_t(e) * NP + p]] = 2;
    }
    for (int g = 0; g < G; ++g) {
        if (cls[g]) continue;
        if (mg[g] <= 0.0 || (!S.has_cpml && dg[g] > 0.0)) cls[g] = 1;
    }
    S.rec_cells.assign(d.rec_cell, d.rec_cell + std::max(d.n_rec_cell, 0));
    for (int32_t e : S.rec_cells) {
        require(e >= 0 && e < E, "recording cell out of range");
        for (int p = 0; p < NP; ++p) {
            char& c = cls[d.ibool[int64_t(e) * NP + p]];
            if (!c) c = 1;
        }
    }
    // C-PML memory directions used at each node: direction d when some PML
    // element point on the node has A_{3+d} != 0 (the others keep zero memory)
    std::vector<uint8_t> pm_old(G, 0);  // stays 0 with the complete C-PML, which has no such memories
    if (S.has_cpml && !S.cpml_complete) {
        for (int64_t e = 0; e < E; ++e) {
            if (d.pml_region[e] == 0) continue;
            for (int p = 0; p < NP; ++p) {
                const int64_t src = e * NP + p;
                for (int dir = 0; dir < 3; ++dir)
                    if (d.coef_abar[src * 5 + 2 + dir] != 0.0) pm_old[d.ibool[src]] |= uint8_t(1u << dir);
            }
        }
    }
    // numbering: C-PML nodes, other special nodes, regular nodes; first touch
    // in solver element order inside each class; C-PML nodes grouped by their
    // direction mask (first touch inside a group), so wavefronts share a mask
    // (-1 = not yet seen, -2 = queued in this class, >= 0 = numbered)
    S.node_new_of_old.assign(G, -1);
    int next = 0;
    for (int pass = 2; pass >= 0; --pass) {
        std::vector<int> order;
        for (int i = 0; i < E; ++i) {
            const int64_t e = elem_of_int[i];
            for (int p = 0; p < NP; ++p) {
                const int g = d.ibool[e * NP + p];
                if (cls[g] == pass && S.node_new_of_old[g] == -1) {
                    S.node_new_of_old[g] = -2;
                    order.push_back(g);
                }
            }
        }
        for (int g = 0; g < G; ++g)  // nodes in no element
            if (cls[g] == pass && S.node_new_of_old[g] == -1) {
                S.node_new_of_old[g] = -2;
                order.push_back(g);
            }
        if (pass == 2)
            std::stable_sort(order.begin(), order.end(),
                             [&](int a, int b) { return pm_old[a] < pm_old[b]; });
        for (int g : order) S.node_new_of_old[g] = next++;
        if (pass == 2) S.Gpml = next;
    }
    // every node numbered exactly once: the numbering is a permutation of [0, G)
    require(next == G, "internal: node numbering is not a permutation");
    {
        std::vector<char> hit(G, 0);
        for (int g = 0; g < G; ++g) {
            const int n = S.node_new_of_old[g];
            require(n >= 0 && n < G && !hit[n], "internal: node numbering is not a permutation");
            hit[n] = 1;
        }
    }
    int64_t n_special = 0;
    for (int g = 0; g < G; ++g) n_special += cls[g] ? 1 : 0;
    S.Ga = std::min<int64_t>(((n_special + 3) / 4) * 4, ((int64_t(G) + 3) / 4) * 4);
    S.G4 = ((int64_t(G) + 3) / 4) * 4;

    // ---- geometry / material, FP64 -> FP32 SoA, first-touch flags ----
    std::vector<float> hG(size_t(9) * P), hlam(P), hmu(P);
    std::vector<uint32_t> hib(P);
    std::atomic<bool> neg_jac{false};
    parallel_for(E, [&](int64_t ie) {
        const int64_t e = elem_of_int[ie];
        for (int p = 0; p < NP; ++p) {
            const int64_t src = e * NP + p, dst = ie * NP + p;
            const int pi = p / 25, pj = (p / 5) % 5, pk = p % 5;
            const double jw = d.jacobian[src] * gll.w[pi] * gll.w[pj] * gll.w[pk];
            if (jw < 0) neg_jac = true;
            const double sq = std::sqrt(std::max(jw, 0.0));
            for (int c = 0; c < 9; ++c) hG[c * P + dst] = float(d.dxi_dx[src * 9 + c] * sq);
            const bool solid = d.mu_[src] > 0.0;  // element_cpu.cpp:42-43
            hlam[dst] = solid ? float(d.lambda_[src]) : 0.f;
            hmu[dst] = solid ? float(d.mu_[src]) : 0.f;
            hib[dst] = uint32_t(S.node_new_of_old[d.ibool[src]]);
        }
    });
    if (neg_jac) throw Error(GFX_ERR_UNSUPPORTED, "negative jacobian*weight (inverted element)");
    if (S.assembly == GFX_ASSEMBLY_COLORED) {
        std::vector<char> seen(G, 0);
        for (int64_t q = 0; q < P; ++q) {
            const uint32_t n = hib[q];
            if (!seen[n]) {
                seen[n] = 1;
                hib[q] |= dev::FIRST_TOUCH;
            }
        }
    }
    // ---- shear attenuation: unrelaxed moduli and per-mechanism step coefficients ----
    S.n_sls = d.n_sls;
    require(d.n_sls >= 0 && d.n_sls <= 8, "n_sls must be in [0, 8]");
    if (S.n_sls > 0) {
        require(d.tau_sigma && d.tau_epsilon, "n_sls > 0 needs tau_sigma and tau_epsilon");
        const int L = S.n_sls;
        std::vector<float> ha(size_t(L) * P), hb(size_t(L) * P), hc(size_t(L) * P);
        std::atomic<bool> bad_tau{false};
        parallel_for(E, [&](int64_t ie) {
            const int64_t e = elem_of_int[ie];
            for (int p = 0; p < NP; ++p) {
                const int64_t src = e * NP + p, dst = ie * NP + p;
                const double mu_r = d.mu_[src];
                double s_u = 1.0;
                for (int l = 0; l < L; ++l) {
                    const double ts = d.tau_sigma[src * L + l], te = d.tau_epsilon[src * L + l];
                    if (!(ts > 0) || !(te > 0) || !std::isfinite(ts) || !std::isfinite(te)) bad_tau = true;
                    s_u += te / ts - 1.0;
                    // exact step for strain linear in time: z <- a z + Y (b dev_old + c dev_new)
                    const double a = std::exp(-S.dt / ts), q = ts * (1.0 - a) / S.dt;
                    const double Y = 2.0 * mu_r * (te / ts - 1.0);
                    const bool solid = mu_r > 0.0;
                    ha[size_t(l) * P + dst] = solid ? float(a) : 0.f;
                    hb[size_t(l) * P + dst] = solid ? float(Y * (q - a)) : 0.f;
                    hc[size_t(l) * P + dst] = solid ? float(Y * (1.0 - q)) : 0.f;
                }
                // the complete C-PML's elements stay elastic with the relaxed moduli
                const bool pml_elastic = S.cpml_complete && d.pml_region[e] != 0;
                if (mu_r > 0.0 && !pml_elastic) {
                    const double mu_u = mu_r * s_u;
                    hmu[dst] = float(mu_u);
                    hlam[dst] = float(d.lambda_[src] + 2.0 * (mu_r - mu_u) / 3.0);
                }
            }
        });
        require(!bad_tau, "tau_sigma and tau_epsilon must be positive and finite");
        dv.upload(S.att_a, std::move(ha));
        dv.upload(S.att_b, std::move(hb));
        dv.upload(S.att_c, std::move(hc));
        dv.alloc(S.zeta, size_t(S.nf) * L * 5 * P);
        dv.alloc(S.devp, size_t(S.nf) * 5 * P);
    }
    dv.upload(S.G9, std::move(hG));
    dv.upload(S.lam, std::move(hlam));
    dv.upload(S.mu, std::move(hmu));
    dv.upload(S.ibool, hib);  // hib is read again for the gather slots

    S.mass_node.assign(G, 0.0);
    std::vector<float> hminv(S.G4, 0.f), hdamp(std::max<int64_t>(S.Ga, 1), 0.f);
    for (int g = 0; g < G; ++g) {
        const int n = S.node_new_of_old[g];
        S.mass_node[n] = mg[g];
        hminv[n] = mg[g] > 0.0 ? float(1.0 / mg[g]) : 0.f;
        if (n < S.Ga) hdamp[n] = float(dg[g]);
    }
    dv.upload(S.minv, std::move(hminv));
    dv.upload(S.damp, std::move(hdamp));

    // ---- C-PML ----
    std::vector<int32_t> hslot(E, -1);
    S.pml_nodal = true;
    if (S.cpml_complete) {
        // per PML element point: K, d, alpha and the one-step recursions of
        // alpha and beta = alpha + d / K for each axis, and rho J w (the mass)
        S.pml_nodal = false;
        int npml = 0;
        for (int ie = 0; ie < E; ++ie)
            if (d.pml_region[elem_of_int[ie]] != 0) hslot[ie] = npml++;
        S.Ppml = int64_t(npml) * NP;
        const int64_t Q = std::max<int64_t>(S.Ppml, 1);
        std::vector<float> hc(size_t(dev::CPML_NC) * Q, 0.f);
        std::atomic<bool> bad{false};
        parallel_for(E, [&](int64_t ie) {
            if (hslot[ie] < 0) return;
            const int64_t e = elem_of_int[ie];
            for (int p = 0; p < NP; ++p) {
                const int64_t src = e * NP + p, q = int64_t(hslot[ie]) * NP + p;
                double c[27];
                if (!cpml_point_coefficients(d.pml_K + src * 3, d.pml_d + src * 3, d.pml_alpha + src * 3, S.dt, c))
                    bad = true;
                for (int t = 0; t < 27; ++t) hc[t * Q + q] = float(c[t]);
                hc[27 * Q + q] = float(d.mass[src]);
            }
        });
        require(!bad, "pml_K must be > 0, pml_d and pml_alpha >= 0 (all finite) at PML element points");
        // memories in use per point: per derivative axis l (x3 components) the
        // ones of axes m, n with d != 0, their product, and beta_l's; inertia 3 per damped axis
        S.cpml_mem_active = 0;
        for (int64_t q = 0; q < S.Ppml; ++q) {
            bool on[3];
            for (int ax = 0; ax < 3; ++ax) on[ax] = hc[size_t(ax * 9 + 1) * Q + q] != 0.f;
            for (int l = 0; l < 3; ++l) {
                const int m = l == 0 ? 1 : 0, n = l == 2 ? 1 : 2;
                S.cpml_mem_active += 3 * (int(on[m]) + int(on[n]) + int(on[m] && on[n]) + int(on[l]));
            }
            S.cpml_mem_active += 3 * (int(on[0]) + int(on[1]) + int(on[2]));
        }
        dv.upload(S.pcoef, std::move(hc));
        dv.alloc(S.pmem, size_t(S.nf) * dev::CPML_NM * Q);
    } else if (S.has_cpml) {
        // Nodal form: the acceleration term sum_e w J/rho (A1 v + A2 u + A3 mx + A4 my + A5 mz)
        // assembles to per-node coefficients, and every element's memory variable at a node
        // evolves identically when the coef_alpha of all elements that use it (A_{3+d} != 0)
        // agree.  Checked here; otherwise the element-local form of pml.cpp is used.
        const int64_t Gp = S.Gpml;
        std::vector<double> C(size_t(5) * Gp, 0.0), al(size_t(9) * Gp, 0.0);
        std::vector<char> have(size_t(3) * Gp, 0);
        bool consistent = true;
        for (int64_t e = 0; e < E && consistent; ++e) {
            if (d.pml_region[e] == 0) continue;
            for (int p = 0; p < NP; ++p) {
                const int64_t src = e * NP + p;
                const int64_t n = S.node_new_of_old[d.ibool[src]];
                const int pi = p / 25, pj = (p / 5) % 5, pk = p % 5;
                const double wg = gll.w[pi] * gll.w[pj] * gll.w[pk];
                const double rho = d.density[src];
                const double scale = wg * (rho > 0.0 ? 1.0 / rho : 0.0) * d.jacobian[src];
                for (int q = 0; q < 5; ++q) C[q * Gp + n] += scale * d.coef_abar[src * 5 + q];
                for (int dir = 0; dir < 3; ++dir) {
                    if (d.coef_abar[src * 5 + 2 + dir] == 0.0) continue;  // m_dir unused here
                    const double* a3 = &d.coef_alpha[src * 9 + dir * 3];
                    if (!have[dir * Gp + n]) {
                        have[dir * Gp + n] = 1;
                        for (int q = 0; q < 3; ++q) al[(dir * 3 + q) * Gp + n] = a3[q];
                    } else {
                        for (int q = 0; q < 3; ++q)
                            if (al[(dir * 3 + q) * Gp + n] != a3[q]) consistent = false;
                    }
                }
            }
        }
        if (consistent) {
            std::vector<float> hc(size_t(5) * std::max<int64_t>(Gp, 1), 0.f),
                hal(size_t(9) * std::max<int64_t>(Gp, 1), 0.f);
            for (int64_t n = 0; n < Gp; ++n) {
                for (int q = 0; q < 5; ++q) hc[q * Gp + n] = float(C[q * Gp + n]);
                for (int dir = 0; dir < 3; ++dir)
                    for (int q = 0; q < 3; ++q)
                        hal[(dir * 3 + q) * Gp + n] =
                            have[dir * Gp + n] ? float(al[(dir * 3 + q) * Gp + n]) : (q == 0 ? 1.f : 0.f);
            }
            dv.upload(S.pc, std::move(hc));
            dv.upload(S.palpha, std::move(hal));
            std::vector<uint8_t> hm(std::max<int64_t>(Gp, 1), 0);
            for (int64_t n = 0; n < Gp; ++n)
                hm[n] = uint8_t((have[n] ? 1 : 0) | (have[Gp + n] ? 2 : 0) | (have[2 * Gp + n] ? 4 : 0));
            dv.upload(S.pmask, std::move(hm));
            dv.alloc(S.pmem, size_t(S.nf) * 9 * std::max<int64_t>(Gp, 1));
            dv.alloc(S.pnew, size_t(S.nf) * 3 * std::max<int64_t>(Gp, 1));
        } else {
            S.pml_nodal = false;
            require(S.n_sls == 0,
                    "shear attenuation needs the nodal C-PML form (every element using a memory variable at a "
                    "node must carry the same alpha for that direction)");
            int npml = 0;
            for (int ie = 0; ie < E; ++ie)
                if (d.pml_region[elem_of_int[ie]] != 0) hslot[ie] = npml++;
            S.n_pml_elem = npml;
            S.Ppml = int64_t(npml) * NP;
            const int64_t Q = std::max<int64_t>(S.Ppml, 1);
            std::vector<float> hal(size_t(9) * Q, 0.f), hab(size_t(5) * Q, 0.f);
            for (int ie = 0; ie < E; ++ie) {
                if (hslot[ie] < 0) continue;
                const int64_t e = elem_of_int[ie];
                for (int p = 0; p < NP; ++p) {
                    const int64_t src = e * NP + p, q = int64_t(hslot[ie]) * NP + p;
                    const int pi = p / 25, pj = (p / 5) % 5, pk = p % 5;
                    const double wg = gll.w[pi] * gll.w[pj] * gll.w[pk];
                    const double rho = d.density[src];
                    const double scale = wg * (rho > 0.0 ? 1.0 / rho : 0.0) * d.jacobian[src];
                    for (int c = 0; c < 9; ++c) hal[c * Q + q] = float(d.coef_alpha[src * 9 + c]);
                    for (int c = 0; c < 5; ++c) hab[c * Q + q] = float(scale * d.coef_abar[src * 5 + c]);
                }
            }
            dv.upload(S.palpha, std::move(hal));
            dv.upload(S.pabar, std::move(hab));
            dv.alloc(S.pmem, size_t(S.nf) * 9 * Q);
            dv.alloc(S.pnew, size_t(S.nf) * 3 * Q);
        }
    }
    if (S.has_cpml) {
        int npe = 0;
        for (int e = 0; e < E; ++e) npe += d.pml_region[e] != 0;
        S.n_pml_elem = npe;
    }
    dv.upload(S.pml_slot, std::move(hslot));

    // element-local C-PML runs the block kernel's general form only
    if (ref_elem_pml(S)) {
        for (auto& rg : S.ranges) rg.diag = false;
        S.n_diag = 0;
    }
    // element kernel: the wavefront-per-element form needs no element-local C-PML state
    const bool wave_ok = !ref_elem_pml(S);
    S.kernel = resolve_kernel(d.element_kernel, wave_ok);
    check_epb(S.kernel, epb);
    S.epb = resolve_epb(S.kernel, epb);

    // ---- nodal state: U, W, R [nf][fstride]; a at special nodes [nf][astride] ----
    S.fstride = ((S.G4 * 3 + 63) / 64) * 64;
    S.astride = ((std::max<int64_t>(S.Ga, 1) * 3 + 63) / 64) * 64;
    const size_t fs = size_t(S.fstride) * S.nf;
    dv.alloc(S.ctr, 2);
    dv.alloc(S.U, fs);
    dv.alloc(S.W, fs);
    dv.alloc(S.R, fs);
    dv.alloc(S.Aacc, size_t(S.astride) * S.nf);
    if (S.assembly == GFX_ASSEMBLY_GATHER) {
        // each node owns the contiguous slots [off[n], off[n+1]), filled with its
        // element points in increasing solver order (the fixed summation order);
        // the element kernels store each point's residual at the point itself
        std::vector<int32_t> off(S.G4 + 1, 0), pts(P);
        for (int64_t q = 0; q < P; ++q) ++off[hib[q] + 1];
        for (int64_t n = 0; n < S.G4; ++n) off[n + 1] += off[n];
        require(off[S.G4] == P, "internal: gather slots do not cover every element point");
        std::vector<int32_t> fill(off.begin(), off.end() - 1);
        for (int64_t q = 0; q < P; ++q) pts[fill[hib[q]]++] = int32_t(q);
        // residual entry of each point: a node's first point (lowest solver order) at
        // the node's own index, the other points after the nodes, element by element,
        // ordered by node id inside each element (the points a node-kernel workgroup
        // gathers from one element then sit together); a node's slots list entries,
        // the first one being the node itself.  Entry positions do not change the
        // sums: each node adds its slots in increasing solver point order.
        std::vector<int32_t> rdst(P, -1);
        for (int64_t n = 0; n < S.G4; ++n)
            if (off[n + 1] > off[n]) rdst[pts[off[n]]] = int32_t(n);
        int64_t k = S.G4;
        for (int64_t e = 0; e < E; ++e) {
            int64_t rest[NP];
            int nr = 0;
            for (int p = 0; p < NP; ++p)
                if (rdst[e * NP + p] < 0) rest[nr++] = e * NP + p;
            std::sort(rest, rest + nr, [&](int64_t a, int64_t b) { return hib[a] < hib[b] || (hib[a] == hib[b] && a < b); });
            for (int q = 0; q < nr; ++q) rdst[rest[q]] = int32_t(k++);
        }
        require(k == S.G4 + P - S.G, "internal: residual entries do not match the point count");
        S.PR = (k + 3) / 4 * 4;  // 16-byte aligned field stride
        for (int64_t q = 0; q < P; ++q) pts[q] = rdst[pts[q]];
        // device slot lists: every point but the node's first (read at the node id)
        std::vector<int32_t> xoff(S.G4 + 1, 0), xpts;
        xpts.reserve(size_t(P - S.G));
        for (int64_t n = 0; n < S.G4; ++n) {
            for (int32_t q = off[n] + 1; q < off[n + 1]; ++q) xpts.push_back(pts[q]);
            xoff[n + 1] = int32_t(xpts.size());
        }
        S.goff_host = off;
        S.gpts_host = pts;
        dv.upload(S.goff, std::move(xoff));
        dv.upload(S.gpts, std::move(xpts));
        dv.upload(S.rdst, std::move(rdst));
        dv.alloc(S.Rl, size_t(S.nf) * S.PR * 3);
    }
    std::vector<uint32_t>().swap(hib);

    // ---- recording cells ----
    S.n_rec = int(S.rec_cells.size());
    S.rec_cap = std::max(d.record_capacity, 0);
    S.rec_uva_on = d.record_uva != 0;
    if (S.n_rec > 0) {
        const int64_t nrp = int64_t(S.n_rec) * NP;
        std::vector<int32_t> rib(nrp);
        std::vector<float> rdd(size_t(9) * nrp);
        for (int c = 0; c < S.n_rec; ++c) {
            const int64_t e = S.rec_cells[c];
            for (int p = 0; p < NP; ++p) {
                rib[int64_t(c) * NP + p] = S.node_new_of_old[d.ibool[e * NP + p]];
                for (int q = 0; q < 9; ++q)
                    rdd[q * nrp + int64_t(c) * NP + p] = float(d.dxi_dx[(e * NP + p) * 9 + q]);
            }
        }
        dv.upload(S.rec_ibool, std::move(rib));
        dv.upload(S.rec_dd, std::move(rdd));
        if (S.rec_cap > 0) {
            dv.alloc(S.rec_strain, size_t(S.rec_cap) * S.nf * nrp * 6);
            if (S.rec_uva_on) dv.alloc(S.rec_uva, size_t(S.rec_cap) * S.nf * nrp * 9);
        }
    }

    // ---- the host plan is complete: device buffers, sources, zero state ----
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        throw Error(GFX_ERR_DEVICE, "no HIP device available");
    }
    GFX_HIP(hipStreamCreateWithFlags(&S.stream, hipStreamNonBlocking));
    dv.run();
    S.sources.assign(S.nf, FieldSource{});
    build_sources(S);
    init_state(S);
}

void check_field(const gfx_solver* s, int field) {
    require(s != nullptr, "null solver");
    require(field >= 0 && field < s->nf, "field index out of range");
}

std::vector<float> fetch(const gfx_solver* s, const float* p, size_t n) {
    std::vector<float> h(n);
    if (n) GFX_HIP(hipMemcpy(h.data(), p, n * sizeof(float), hipMemcpyDeviceToHost));
    return h;
}

}  // namespace
}  // namespace gfx

using namespace gfx;

extern "C" int gfx_device_count(void) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) {
        (void)hipGetLastError();
        return 0;
    }
    return n;
}

extern "C" gfx_status gfx_cpml_rigid_nodes(const gfx_solver_desc* d, uint8_t* mask) {
    return guarded([&] {
        require(d && mask, "null argument");
        require(d->ngll == 5, "NGLL = 5 (polynomial order 4) only");
        require(d->n_elem > 0 && d->n_rank_node > 0 && d->ibool && d->dxi_dx && d->pml_region,
                "needs n_elem, n_rank_node, ibool, dxi_dx and pml_region");
        const int64_t P = int64_t(d->n_elem) * NP;
        for (int64_t p = 0; p < P; ++p)
            require(d->ibool[p] >= 0 && d->ibool[p] < d->n_rank_node, "ibool entry out of range");
        const std::vector<char> r = cpml_rigid_nodes(*d, d->n_elem, d->n_rank_node);
        for (int g = 0; g < d->n_rank_node; ++g) mask[g] = uint8_t(r[g]);
    });
}

extern "C" gfx_status gfx_solver_create(const gfx_solver_desc* desc, gfx_solver** out) {
    return guarded([&] {
        require(desc && out, "null argument");
        *out = nullptr;
        auto S = std::make_unique<gfx_solver>();
        create(*desc, *S);
        *out = S.release();
    });
}

extern "C" gfx_status gfx_solver_destroy(gfx_solver* s) {
    return guarded([&] {
        if (s) {
            if (s->stream) (void)hipStreamSynchronize(s->stream);
            delete s;
        }
    });
}

extern "C" gfx_status gfx_solver_set_source(gfx_solver* s, int field, int direction,
                                            int n_src_cell, const int32_t* src_elem,
                                            const double* src_weights, int n_stf,
                                            const double* stf) {
    return guarded([&] {
        check_field(s, field);
        require(direction >= 0 && direction <= 2, "direction must be 0, 1 or 2");
        require(n_src_cell >= 0 && n_stf >= 0, "negative count");
        require(n_src_cell == 0 || (src_elem && src_weights), "missing source arrays");
        require(n_stf == 0 || stf, "missing stf");
        FieldSource fs;
        fs.direction = direction;
        for (int c = 0; c < n_src_cell; ++c) {
            require(src_elem[c] < s->E, "source cell out of range");
            fs.cells.push_back(src_elem[c]);
            for (int p = 0; p < NP; ++p) fs.weights.push_back(float(src_weights[int64_t(c) * NP + p]));
        }
        for (int t = 0; t < n_stf; ++t) fs.stf.push_back(float(stf[t]));
        GFX_HIP(hipStreamSynchronize(s->stream));
        s->sources[field] = std::move(fs);
        drop_graph(*s);
        build_sources(*s);
    });
}

extern "C" gfx_status gfx_solver_reset(gfx_solver* s) {
    return guarded([&] {
        require(s != nullptr, "null solver");
        init_state(*s);
    });
}

extern "C" gfx_status gfx_solver_run(gfx_solver* s, int n_steps) {
    return guarded([&] {
        require(s != nullptr, "null solver");
        require(n_steps >= 0, "n_steps must be >= 0");
        require(s->stage == 0, "a step driven stage by stage is not finished (gfx_step_newmark_correct)");
        if (n_steps > 0) run_steps(*s, n_steps);
    });
}

extern "C" int gfx_solver_step_index(const gfx_solver* s) { return s ? s->step : -1; }

extern "C" gfx_status gfx_solver_set_step(gfx_solver* s, int step) {
    return guarded([&] {
        require(s != nullptr, "null solver");
        require(step >= 0, "step must be >= 0");
        s->step = step;
        // snapshots are taken after the updates with index % stride == 0
        s->n_snap = s->stride > 0 ? (step + s->stride - 1) / s->stride : 0;
        upload_counters(*s);
    });
}
extern "C" int gfx_solver_n_snapshots(const gfx_solver* s) { return s ? s->n_snap : -1; }

extern "C" gfx_status gfx_solver_get_state(const gfx_solver* s, int field, double* u, double* v,
                                           double* a) {
    return guarded([&] {
        check_field(s, field);
        GFX_HIP(hipStreamSynchronize(s->stream));
        const size_t n3 = size_t(s->G) * 3;
        const std::vector<float> hU = fetch(s, s->U.p + int64_t(field) * s->fstride, n3);
        const std::vector<float> hW = fetch(s, s->W.p + int64_t(field) * s->fstride, n3);
        std::vector<float> hR;
        if (s->assembly == GFX_ASSEMBLY_GATHER && !s->r_in_R) {
            // assembled r of the last step: fixed-order sum of the node's slots, as the
            // node kernels form it
            const std::vector<float> hl = fetch(s, s->Rl.p + int64_t(field) * s->PR * 3, size_t(s->PR) * 3);
            const std::vector<int32_t>& off = s->goff_host;
            const std::vector<int32_t>& pts = s->gpts_host;
            hR.assign(size_t(s->G4) * 3, 0.f);
            for (int64_t n = 0; n < s->G; ++n)
                for (int q = off[n]; q < off[n + 1]; ++q)
                    for (int c = 0; c < 3; ++c) hR[n * 3 + c] += hl[int64_t(pts[q]) * 3 + c];
        } else {
            hR = fetch(s, s->R.p + int64_t(field) * s->fstride, n3);
        }
        const std::vector<float> hA = fetch(s, s->Aacc.p + int64_t(field) * s->astride, size_t(s->Ga) * 3);
        const float dt = float(s->dt);
        for (int g = 0; g < s->G; ++g) {
            const int64_t n = s->node_new_of_old[g];
            for (int c = 0; c < 3; ++c) {
                const float w = hW[n * 3 + c];
                // regular nodes keep the assembled r = m a_n in R
                float an;
                if (n < s->Ga) {
                    an = hA[n * 3 + c];
                } else {
                    const double m = s->mass_node[n];
                    an = (m > 0.0) ? hR[n * 3 + c] * float(1.0 / m) : 0.f;
                }
                if (u) u[int64_t(g) * 3 + c] = hU[n * 3 + c] - dt * w;
                if (v) v[int64_t(g) * 3 + c] = w - 0.5f * dt * an;
                if (a) a[int64_t(g) * 3 + c] = an;
            }
        }
    });
}

extern "C" gfx_status gfx_solver_set_state(gfx_solver* s, int field, const double* u,
                                           const double* v, const double* a) {
    return guarded([&] {
        check_field(s, field);
        require(u && v && a, "u, v and a are required");
        const size_t n3 = size_t(s->G) * 3;
        std::vector<float> hU(n3), hW(n3), hR(n3), hA(size_t(s->Ga) * 3, 0.f);
        const double dt = s->dt;
        for (int g = 0; g < s->G; ++g) {
            const int64_t n = s->node_new_of_old[g];
            const double m = s->mass_node[n];
            for (int c = 0; c < 3; ++c) {
                const int64_t o = int64_t(g) * 3 + c;
                // a massless node keeps the given a until its first corrector sets it to 0,
                // so the first predictor uses it (solver.cpp:40-69); massless nodes are
                // special nodes, whose a is stored
                const double ag = a[o];
                hU[n * 3 + c] = float(u[o] + dt * v[o] + 0.5 * dt * dt * ag);
                hW[n * 3 + c] = float(v[o] + 0.5 * dt * ag);
                hR[n * 3 + c] = float(ag * m);
                if (n < s->Ga) hA[n * 3 + c] = float(ag);
            }
        }
        GFX_HIP(hipStreamSynchronize(s->stream));
        const int64_t fo = int64_t(field) * s->fstride;
        GFX_HIP(hipMemcpy(s->U.p + fo, hU.data(), n3 * 4, hipMemcpyHostToDevice));
        GFX_HIP(hipMemcpy(s->W.p + fo, hW.data(), n3 * 4, hipMemcpyHostToDevice));
        GFX_HIP(hipMemcpy(s->R.p + fo, hR.data(), n3 * 4, hipMemcpyHostToDevice));
        if (!hA.empty())
            GFX_HIP(hipMemcpy(s->Aacc.p + int64_t(field) * s->astride, hA.data(), hA.size() * 4,
                              hipMemcpyHostToDevice));
        if (s->assembly == GFX_ASSEMBLY_GATHER) {  // r = m a at the first entry of each node
            std::vector<float> hl(size_t(s->PR) * 3, 0.f);
            for (int64_t n = 0; n < s->G; ++n)
                if (s->goff_host[n + 1] > s->goff_host[n])
                    for (int c = 0; c < 3; ++c)
                        hl[int64_t(s->gpts_host[s->goff_host[n]]) * 3 + c] = hR[n * 3 + c];
            GFX_HIP(hipMemcpy(s->Rl.p + int64_t(field) * s->PR * 3, hl.data(), hl.size() * 4,
                              hipMemcpyHostToDevice));
        }
    });
}

// ---------------------------------------------------------------------------
// Per-stage entry points: the reference's device branch
// (forward/share/src/solver.cpp:390-444) driven one stage at a time on the
// solver's stream, with the multi-rank exchange copies between them.  The
// stages are the fused step cut at the reference's boundaries; the state
// between steps stays the leapfrog form (U = predicted displacement), so a
// step driven stage by stage in the reference's order is the same step as
// gfx_solver_run's.
namespace gfx {
namespace {

enum { STAGE_IDLE = 0, STAGE_PREDICTED = 1, STAGE_ELEMENT = 2, STAGE_ASSEMBLED = 3 };

void stage_require(const gfx_solver* s, int want, const char* fn, const char* after) {
    require(s != nullptr, "null solver");
    if (s->stage != want)
        throw Error(GFX_ERR_INVALID, std::string(fn) + " called out of order: it follows " + after);
}

void check_newmark(const gfx_solver* s, double dt, double beta) {
    require(beta == 0.0, "the solver's Newmark scheme is explicit: beta must be 0");
    require(std::fabs(dt - s->dt) <= 1e-12 * std::fabs(s->dt), "dt differs from the solver's dt");
}

// Caller-numbered FP64 [G*3] <-> solver-numbered FP32 field of a nodal array.
std::vector<double> nodal_to_caller(const gfx_solver* s, const float* dev_field) {
    GFX_HIP(hipStreamSynchronize(s->stream));
    const std::vector<float> h = fetch(s, dev_field, size_t(s->G) * 3);
    std::vector<double> out(size_t(s->G) * 3);
    for (int g = 0; g < s->G; ++g) {
        const int64_t n = s->node_new_of_old[g];
        for (int c = 0; c < 3; ++c) out[size_t(g) * 3 + c] = h[n * 3 + c];
    }
    return out;
}

void nodal_from_caller(gfx_solver* s, float* dev_field, const double* buf) {
    std::vector<float> h(size_t(s->G) * 3);
    for (int g = 0; g < s->G; ++g) {
        const int64_t n = s->node_new_of_old[g];
        for (int c = 0; c < 3; ++c) h[n * 3 + c] = float(buf[size_t(g) * 3 + c]);
    }
    GFX_HIP(hipStreamSynchronize(s->stream));
    GFX_HIP(hipMemcpy(dev_field, h.data(), h.size() * sizeof(float), hipMemcpyHostToDevice));
}

}  // namespace
}  // namespace gfx

extern "C" gfx_status gfx_step_newmark_predict(gfx_solver* s, double dt, double beta) {
    return guarded([&] {
        stage_require(s, STAGE_IDLE, "gfx_step_newmark_predict", "gfx_step_newmark_correct (or creation)");
        check_newmark(s, dt, beta);
        // U already holds u~_{n+1} = u_n + dt (v_n + dt/2 a_n): the corrector of
        // the previous step wrote it (and set_state / creation for the first)
        s->stage = STAGE_PREDICTED;
        s->stage_zeroed = false;
    });
}

extern "C" gfx_status gfx_step_gather_predicted(gfx_solver* s) {
    return guarded([&] {
        stage_require(s, STAGE_PREDICTED, "gfx_step_gather_predicted", "gfx_step_newmark_predict");
        // the element kernels read u~ through the node ids themselves
    });
}

extern "C" gfx_status gfx_step_gather_from_rank(gfx_solver* s) {
    return guarded([&] {
        require(s != nullptr, "null solver");
        // no element-local displacement copy exists: every kernel reads the nodal arrays
    });
}

extern "C" gfx_status gfx_step_zero_residual(gfx_solver* s) {
    return guarded([&] {
        stage_require(s, STAGE_PREDICTED, "gfx_step_zero_residual", "gfx_step_newmark_predict");
        if (s->assembly == GFX_ASSEMBLY_ATOMIC) {
            const int64_t r4 = int64_t(s->nf) * s->fstride / 4;
            dev::k_zero<<<int((r4 + 255) / 256), 256, 0, s->stream>>>(reinterpret_cast<float4*>(s->R.p), r4);
            GFX_HIP(hipGetLastError());
        }
        // COLORED stores at the first touch and GATHER stores per point: nothing to clear
        s->stage_zeroed = true;
    });
}

extern "C" gfx_status gfx_step_element_residual(gfx_solver* s, int ngll, int n_elem) {
    return guarded([&] {
        stage_require(s, STAGE_PREDICTED, "gfx_step_element_residual", "gfx_step_newmark_predict");
        require(ngll == 5, "ngll must be 5");
        require(n_elem == s->E, "n_elem differs from the solver's element count");
        require(s->stage_zeroed || s->assembly != GFX_ASSEMBLY_ATOMIC,
                "ATOMIC assembly adds into the residual: call gfx_step_zero_residual first");
        StepArgs a = make_step_args(*s);
        a.ea.nstf = 0;  // sources come from gfx_step_source_injection
        for (const auto& rg : s->ranges) launch_element(*s, a.ea, rg);
        GFX_HIP(hipGetLastError());
        s->stage = STAGE_ELEMENT;
    });
}

extern "C" gfx_status gfx_step_pml_damping(gfx_solver* s) {
    return guarded([&] {
        stage_require(s, STAGE_ELEMENT, "gfx_step_pml_damping", "gfx_step_element_residual");
        // the legacy ramp v -= d v (without C-PML) acts on v_n inside the
        // corrector, before the update, where the reference applies it
    });
}

extern "C" gfx_status gfx_step_source_injection(gfx_solver* s, int field, int direction, double stf_val,
                                                const double* src_weights, int n_src_cell) {
    return guarded([&] {
        stage_require(s, STAGE_ELEMENT, "gfx_step_source_injection", "gfx_step_element_residual");
        check_field(s, field);
        require(direction >= 0 && direction <= 2, "direction must be 0, 1 or 2");
        const gfx::FieldSource& fs = s->sources[field];
        require(n_src_cell == int(fs.cells.size()),
                "n_src_cell differs from the cells given to gfx_solver_set_source for this field");
        require(n_src_cell == 0 || src_weights, "missing source weights");
        std::vector<int32_t> cells;
        std::vector<float> w;
        for (int c = 0; c < n_src_cell; ++c) {
            if (fs.cells[c] < 0) continue;  // source cell not on this rank (solver.cpp:585-587)
            cells.push_back(s->int_of_elem[fs.cells[c]]);
            for (int p = 0; p < NP; ++p) w.push_back(float(src_weights[int64_t(c) * NP + p]));
        }
        if (cells.empty()) return;
        GFX_HIP(hipStreamSynchronize(s->stream));  // the previous injection's tables are no longer read
        if (s->inj_cells.n < cells.size()) s->inj_cells.alloc(cells.size());
        if (s->inj_w.n < w.size()) s->inj_w.alloc(w.size());
        GFX_HIP(hipMemcpy(s->inj_cells.p, cells.data(), cells.size() * 4, hipMemcpyHostToDevice));
        GFX_HIP(hipMemcpy(s->inj_w.p, w.data(), w.size() * 4, hipMemcpyHostToDevice));
        dev::k_inject<<<1, 128, 0, s->stream>>>(s->Rl.p, s->rdst.p, s->R.p, s->ibool.p, s->inj_cells.p, s->inj_w.p,
                                               int(cells.size()), direction, float(stf_val),
                                               s->PR, s->fstride, field,
                                               s->assembly == GFX_ASSEMBLY_GATHER ? 1 : 0);
        GFX_HIP(hipGetLastError());
    });
}

extern "C" gfx_status gfx_step_scatter_to_rank(gfx_solver* s) {
    return guarded([&] {
        stage_require(s, STAGE_ELEMENT, "gfx_step_scatter_to_rank", "gfx_step_element_residual");
        if (s->assembly == GFX_ASSEMBLY_GATHER) {
            const StepArgs a = make_step_args(*s);
            dim3 grid(unsigned((s->G4 + 255) / 256), unsigned(s->nf));
            dev::k_scatter_slots<<<grid, 256, 0, s->stream>>>(a.na);
            GFX_HIP(hipGetLastError());
        }
        // COLORED and ATOMIC assemble into R inside the element stage
        s->stage = STAGE_ASSEMBLED;
    });
}

extern "C" gfx_status gfx_step_newmark_correct(gfx_solver* s, double dt, double beta, double gamma) {
    return guarded([&] {
        stage_require(s, STAGE_ASSEMBLED, "gfx_step_newmark_correct", "gfx_step_scatter_to_rank");
        check_newmark(s, dt, beta);
        require(gamma == 0.5, "the solver's Newmark scheme has gamma = 1/2");
        StepArgs a = make_step_args(*s);
        a.na.gather = 0;  // R holds the assembled residual (k_scatter_slots, or the exchange)
        if (a.spec_grid > 0) launch_node_special(*s, a.na, a.spec_grid);
        if (a.reg_grid > 0) launch_node_regular(*s, a.na, a.reg_grid);
        if (a.recording && s->step % s->stride == 0 && a.rec_grid > 0)
            dev::k_record<REPB><<<a.rec_grid, a.rec_threads, 0, s->stream>>>(a.ra);
        dev::k_advance<<<1, 64, 0, s->stream>>>(s->ctr.p, s->stride);
        GFX_HIP(hipGetLastError());
        advance_host(*s);
        s->stage = STAGE_IDLE;
        s->r_in_R = true;
    });
}

extern "C" gfx_status gfx_step_copy_utilde_to_host(const gfx_solver* s, int field, double* buf) {
    return guarded([&] {
        check_field(s, field);
        require(buf != nullptr, "null buffer");
        const std::vector<double> h = nodal_to_caller(s, s->U.p + int64_t(field) * s->fstride);
        std::copy(h.begin(), h.end(), buf);
    });
}

extern "C" gfx_status gfx_step_copy_utilde_from_host(gfx_solver* s, int field, const double* buf) {
    return guarded([&] {
        stage_require(s, STAGE_PREDICTED, "gfx_step_copy_utilde_from_host", "gfx_step_newmark_predict");
        check_field(s, field);
        require(buf != nullptr, "null buffer");
        nodal_from_caller(s, s->U.p + int64_t(field) * s->fstride, buf);
    });
}

extern "C" gfx_status gfx_step_copy_residual_to_host(const gfx_solver* s, int field, double* buf) {
    return guarded([&] {
        stage_require(s, STAGE_ASSEMBLED, "gfx_step_copy_residual_to_host", "gfx_step_scatter_to_rank");
        check_field(s, field);
        require(buf != nullptr, "null buffer");
        const std::vector<double> h = nodal_to_caller(s, s->R.p + int64_t(field) * s->fstride);
        std::copy(h.begin(), h.end(), buf);
    });
}

extern "C" gfx_status gfx_step_copy_residual_from_host(gfx_solver* s, int field, const double* buf) {
    return guarded([&] {
        stage_require(s, STAGE_ASSEMBLED, "gfx_step_copy_residual_from_host", "gfx_step_scatter_to_rank");
        check_field(s, field);
        require(buf != nullptr, "null buffer");
        nodal_from_caller(s, s->R.p + int64_t(field) * s->fstride, buf);
    });
}

extern "C" gfx_status gfx_step_copy_state_to_host(const gfx_solver* s, int field, double* u, double* v,
                                                  double* a) {
    return guarded([&] {
        stage_require(s, STAGE_IDLE, "gfx_step_copy_state_to_host", "gfx_step_newmark_correct");
    }) == GFX_OK ? gfx_solver_get_state(s, field, u, v, a) : GFX_ERR_INVALID;
}

extern "C" gfx_status gfx_solver_get_record(const gfx_solver* s, int field, int snap,
                                            double* strain, double* disp, double* vel,
                                            double* acc) {
    return guarded([&] {
        check_field(s, field);
        require(s->rec_cap > 0 && s->n_rec > 0, "recording is off");
        require(snap >= 0 && snap < s->n_snap && snap >= s->n_snap - s->rec_cap,
                "snapshot not held in the record ring");
        require(!(disp || vel || acc) || s->rec_uva_on, "u, v, a records are off");
        const int slot = snap % s->rec_cap;
        const int64_t nrp = int64_t(s->n_rec) * NP;
        GFX_HIP(hipStreamSynchronize(s->stream));
        if (strain) {
            const std::vector<float> h =
                fetch(s, s->rec_strain.p + (int64_t(slot) * s->nf + field) * nrp * 6, size_t(nrp) * 6);
            // FP32 -> FP64 widening on the host threads, one recording point per index
            parallel_for(nrp, [&](int64_t q) {
                for (int c = 0; c < 6; ++c) strain[q * 6 + c] = h[q * 6 + c];
            });
        }
        if (disp || vel || acc) {
            const std::vector<float> h =
                fetch(s, s->rec_uva.p + (int64_t(slot) * s->nf + field) * nrp * 9, size_t(nrp) * 9);
            parallel_for(nrp, [&](int64_t q) {
                for (int c = 0; c < 3; ++c) {
                    if (disp) disp[q * 3 + c] = h[q * 9 + c];
                    if (vel) vel[q * 3 + c] = h[q * 9 + 3 + c];
                    if (acc) acc[q * 3 + c] = h[q * 9 + 6 + c];
                }
            });
        }
    });
}

namespace gfx {
namespace {

// Checkpoint image: a header of int64 words, then the device arrays that
// carry state from one step to the next, as raw FP32.
constexpr int64_t CKPT_MAGIC = 0x31544b5043584647LL;  // "GFXCPKT1"
constexpr int CKPT_WORDS = 16;

uint64_t layout_signature(const gfx_solver& S) {
    uint64_t h = 1469598103934665603ULL;  // FNV-1a over the solver's numbering
    auto mix = [&](const void* p, size_t n) {
        const unsigned char* c = static_cast<const unsigned char*>(p);
        for (size_t i = 0; i < n; ++i) h = (h ^ c[i]) * 1099511628211ULL;
    };
    mix(S.node_new_of_old.data(), S.node_new_of_old.size() * sizeof(int32_t));
    mix(S.int_of_elem.data(), S.int_of_elem.size() * sizeof(int32_t));
    const int64_t k[7] = {S.E, S.G, S.nf, S.Ga, S.Gpml, int64_t(S.assembly), int64_t(S.cpml_complete)};
    mix(k, sizeof(k));
    return h;
}

size_t checkpoint_bytes(const gfx_solver& S) {
    const size_t fl = size_t(S.nf) * S.fstride;
    return CKPT_WORDS * 8 + 4 * (3 * fl + size_t(S.nf) * S.astride + S.pmem.n + S.pnew.n + S.zeta.n + S.devp.n);
}

}  // namespace
}  // namespace gfx

extern "C" gfx_status gfx_solver_checkpoint_size(const gfx_solver* s, size_t* bytes) {
    return guarded([&] {
        require(s && bytes, "null argument");
        *bytes = checkpoint_bytes(*s);
    });
}

extern "C" gfx_status gfx_solver_checkpoint_save(const gfx_solver* s, void* host, size_t bytes) {
    return guarded([&] {
        require(s && host, "null argument");
        require(bytes >= checkpoint_bytes(*s), "checkpoint buffer too small");
        GFX_HIP(hipStreamSynchronize(s->stream));
        int64_t* hdr = static_cast<int64_t*>(host);
        std::memset(hdr, 0, CKPT_WORDS * 8);
        hdr[0] = CKPT_MAGIC;
        hdr[1] = s->E;
        hdr[2] = s->G;
        hdr[3] = s->nf;
        hdr[4] = s->assembly;
        hdr[5] = s->step;
        hdr[6] = s->n_snap;
        hdr[7] = int64_t(layout_signature(*s));
        hdr[8] = int64_t(s->pmem.n);
        hdr[9] = int64_t(s->pnew.n);
        hdr[10] = int64_t(s->zeta.n);
        hdr[11] = int64_t(s->devp.n);
        float* p = reinterpret_cast<float*>(static_cast<char*>(host) + CKPT_WORDS * 8);
        const size_t fl = size_t(s->nf) * s->fstride, al = size_t(s->nf) * s->astride;
        auto get = [&](const float* d, size_t n) {
            if (n) GFX_HIP(hipMemcpy(p, d, n * 4, hipMemcpyDeviceToHost));
            p += n;
        };
        get(s->U.p, fl);
        get(s->W.p, fl);
        get(s->Aacc.p, al);
        get(s->pmem.p, s->pmem.n);
        get(s->pnew.p, s->pnew.n);
        get(s->zeta.p, s->zeta.n);
        get(s->devp.p, s->devp.n);
        // assembled residual of the last step (what get_state reads for a at
        // regular nodes): R itself, or the fixed-order slot sums
        if (s->assembly != GFX_ASSEMBLY_GATHER || s->r_in_R) {
            get(s->R.p, fl);
        } else {
            std::fill(p, p + fl, 0.f);
            for (int f = 0; f < s->nf; ++f) {
                const std::vector<float> hl = fetch(s, s->Rl.p + int64_t(f) * s->PR * 3, size_t(s->PR) * 3);
                float* r = p + int64_t(f) * s->fstride;
                for (int64_t n = 0; n < s->G; ++n)
                    for (int q = s->goff_host[n]; q < s->goff_host[n + 1]; ++q)
                        for (int c = 0; c < 3; ++c) r[n * 3 + c] += hl[int64_t(s->gpts_host[q]) * 3 + c];
            }
        }
    });
}

extern "C" gfx_status gfx_solver_checkpoint_load(gfx_solver* s, const void* host, size_t bytes) {
    return guarded([&] {
        require(s && host, "null argument");
        require(bytes >= size_t(CKPT_WORDS) * 8, "checkpoint image too small");
        const int64_t* hdr = static_cast<const int64_t*>(host);
        require(hdr[0] == CKPT_MAGIC, "not a solver checkpoint image");
        require(hdr[1] == s->E && hdr[2] == s->G && hdr[3] == s->nf && hdr[4] == s->assembly &&
                    hdr[7] == int64_t(layout_signature(*s)) && hdr[8] == int64_t(s->pmem.n) &&
                    hdr[9] == int64_t(s->pnew.n) && hdr[10] == int64_t(s->zeta.n) && hdr[11] == int64_t(s->devp.n),
                "checkpoint was taken from a solver with another mesh, field count, assembly or attenuation");
        require(bytes >= checkpoint_bytes(*s), "checkpoint image truncated");
        const float* p = reinterpret_cast<const float*>(static_cast<const char*>(host) + CKPT_WORDS * 8);
        const size_t fl = size_t(s->nf) * s->fstride, al = size_t(s->nf) * s->astride;
        GFX_HIP(hipStreamSynchronize(s->stream));
        auto put = [&](float* d, size_t n) {
            if (n) GFX_HIP(hipMemcpy(d, p, n * 4, hipMemcpyHostToDevice));
            p += n;
        };
        put(s->U.p, fl);
        put(s->W.p, fl);
        put(s->Aacc.p, al);
        put(s->pmem.p, s->pmem.n);
        put(s->pnew.p, s->pnew.n);
        put(s->zeta.p, s->zeta.n);
        put(s->devp.p, s->devp.n);
        if (s->assembly != GFX_ASSEMBLY_GATHER) {
            put(s->R.p, fl);
        } else {
            // the sum at the first point of each node, zeros elsewhere: get_state
            // then forms the same r
            std::vector<float> hl(size_t(s->PR) * 3);
            for (int f = 0; f < s->nf; ++f) {
                std::fill(hl.begin(), hl.end(), 0.f);
                const float* r = p + int64_t(f) * s->fstride;
                for (int64_t n = 0; n < s->G; ++n)
                    if (s->goff_host[n + 1] > s->goff_host[n])
                        for (int c = 0; c < 3; ++c) hl[int64_t(s->gpts_host[s->goff_host[n]]) * 3 + c] = r[n * 3 + c];
                GFX_HIP(hipMemcpy(s->Rl.p + int64_t(f) * s->PR * 3, hl.data(), hl.size() * 4, hipMemcpyHostToDevice));
            }
        }
        s->step = int(hdr[5]);
        s->n_snap = int(hdr[6]);
        s->r_in_R = false;
        s->stage = 0;
        upload_counters(*s);
    });
}

extern "C" gfx_status gfx_solver_set_timing(gfx_solver* s, int enable) {
    return guarded([&] {
        require(s != nullptr, "null solver");
        s->timing = enable != 0;
    });
}

extern "C" gfx_status gfx_solver_kernel_time(const gfx_solver* s, const char* name, double* ms,
                                             int* launches) {
    return guarded([&] {
        require(s && name && ms && launches, "null argument");
        const std::string n(name);
        if (n == "element") {
            *ms = s->t_elem;
            *launches = s->n_elem_launch;
        } else if (n == "element_pml") {  // the complete C-PML's element launches (within "element")
            *ms = s->t_elem_pml;
            *launches = s->n_elem_pml_launch;
        } else if (n == "node") {
            *ms = s->t_spec + s->t_reg;
            *launches = s->n_spec_launch + s->n_reg_launch;
        } else if (n == "node_special") {
            *ms = s->t_spec;
            *launches = s->n_spec_launch;
        } else if (n == "node_regular") {
            *ms = s->t_reg;
            *launches = s->n_reg_launch;
        } else if (n == "record") {
            *ms = s->t_rec;
            *launches = s->n_rec_launch;
        } else if (n == "total") {
            *ms = s->t_total;
            *launches = s->n_total_launch;
        } else {
            throw Error(GFX_ERR_INVALID, "unknown kernel name: " + n);
        }
    });
}

extern "C" gfx_status gfx_solver_set_launch(gfx_solver* s, int element_kernel, int elems_per_block) {
    return guarded([&] {
        require(s != nullptr, "null solver");
        const bool wave_ok = !ref_elem_pml(*s);
        const int kern = resolve_kernel(element_kernel, wave_ok);
        check_epb(kern, elems_per_block);
        s->kernel = kern;
        s->epb = resolve_epb(s->kernel, elems_per_block);
        drop_graph(*s);
    });
}

extern "C" gfx_status gfx_solver_set_graph(gfx_solver* s, int enable) {
    return guarded([&] {
        require(s != nullptr, "null solver");
        GFX_HIP(hipStreamSynchronize(s->stream));
        if (!enable) drop_graph(*s);
        s->use_graph = enable != 0;
    });
}

extern "C" gfx_status gfx_solver_set_node_groups(gfx_solver* s, int node_field_groups) {
    return guarded([&] {
        require(s != nullptr, "null solver");
        const int g = node_field_groups;
        require(g == 1 || g == 2 || g == 4 || g == 8, "node_field_groups must be 1, 2, 4 or 8");
        s->node_fg = g;
        drop_graph(*s);
    });
}

extern "C" gfx_status gfx_autotune(const gfx_solver_desc* base, int trial_steps,
                                   gfx_solver_desc* best, double* best_ms) {
    return guarded([&] {
        require(base && best && best_ms, "null argument");
        require(trial_steps >= 1, "trial_steps must be >= 1");
        double best_t = 1e300;
        gfx_solver_desc pick = *base;
        const int asms[2] = {GFX_ASSEMBLY_COLORED, GFX_ASSEMBLY_GATHER};
        constexpr int NL = 9;
        const int launch[NL][2] = {{GFX_ELEMENT_BLOCK, 1}, {GFX_ELEMENT_BLOCK, 2}, {GFX_ELEMENT_BLOCK, 4},
                                   {GFX_ELEMENT_WAVE, 2},  {GFX_ELEMENT_WAVE, 4},  {GFX_ELEMENT_WAVE, 8},
                                   {GFX_ELEMENT_PAIR, 1},  {GFX_ELEMENT_PAIR, 2},  {GFX_ELEMENT_PENCIL, 0}};
        auto time_steps = [&](gfx_solver& S) {
            run_steps(S, 1);  // warm-up
            hipEvent_t e0, e1;
            GFX_HIP(hipEventCreate(&e0));
            GFX_HIP(hipEventCreate(&e1));
            GFX_HIP(hipEventRecord(e0, S.stream));
            const bool keep_timing = S.timing;
            S.timing = false;
            run_steps(S, trial_steps);
            S.timing = keep_timing;
            GFX_HIP(hipEventRecord(e1, S.stream));
            GFX_HIP(hipEventSynchronize(e1));
            float ms = 0;
            GFX_HIP(hipEventElapsedTime(&ms, e0, e1));
            (void)hipEventDestroy(e0);
            (void)hipEventDestroy(e1);
            return double(ms) / trial_steps;
        };
        const int fg0 = base->node_field_groups ? base->node_field_groups : 1;
        for (int asm_mode : asms) {
            gfx_solver_desc d = *base;
            d.assembly = asm_mode;
            d.record_capacity = 0;
            d.element_kernel = GFX_ELEMENT_AUTO;
            d.elems_per_block = 0;
            d.node_field_groups = fg0;
            auto S = std::make_unique<gfx_solver>();
            create(d, *S);
            double best_here = 1e300;
            int best_lc = -1;
            for (int li = 0; li < NL; ++li) {
                const auto& lc = launch[li];
                if (lc[0] != GFX_ELEMENT_BLOCK && ref_elem_pml(*S)) continue;
                if (lc[0] == GFX_ELEMENT_PAIR && S->nf < 2) continue;
                S->kernel = lc[0];
                S->epb = resolve_epb(lc[0], lc[1]);
                const double per = time_steps(*S);
                if (per < best_here) {
                    best_here = per;
                    best_lc = li;
                }
            }
            // node-kernel field groups for this assembly's best element launch
            S->kernel = launch[best_lc][0];
            S->epb = resolve_epb(launch[best_lc][0], launch[best_lc][1]);
            int best_fg = fg0;
            for (int fg : {1, 2, 4, 8}) {
                if (fg == fg0 || fg > 2 * S->nf) continue;
                S->node_fg = fg;
                const double per = time_steps(*S);
                if (per < best_here) {
                    best_here = per;
                    best_fg = fg;
                }
            }
            if (best_here < best_t) {
                best_t = best_here;
                pick = *base;
                pick.assembly = asm_mode;
                pick.element_kernel = launch[best_lc][0];
                pick.elems_per_block = launch[best_lc][1];
                pick.node_field_groups = best_fg;
            }
            GFX_HIP(hipStreamSynchronize(S->stream));
        }
        *best = pick;
        *best_ms = best_t;
    });
}

extern "C" gfx_status gfx_solver_query(const gfx_solver* s, const char* key, double* value) {
    return guarded([&] {
        require(s && key && value, "null argument");
        const std::string k(key);
        if (k == "colors") *value = double(s->n_colors);
        else if (k == "diag_elements") *value = double(s->n_diag);
        else if (k == "element_kernel") *value = s->kernel;
        else if (k == "elems_per_block") *value = s->epb;
        else if (k == "node_field_groups") *value = s->node_fg;
        else if (k == "n_sls") *value = s->n_sls;
        else if (k == "pml_nodal") *value = (s->has_cpml && s->pml_nodal) ? 1.0 : 0.0;
        else if (k == "pml_elements") *value = s->n_pml_elem;
        else if (k == "cpml_complete") *value = s->cpml_complete ? 1.0 : 0.0;
        else if (k == "cpml_points") *value = double(s->Ppml);
        else if (k == "cpml_memories") *value = double(s->cpml_mem_active);
        else if (k == "cpml_groups")
            *value = double(std::count_if(s->ranges.begin(), s->ranges.end(), [](const ElemRange& r) { return r.pml; }));
        else if (k == "special_nodes") *value = double(s->Ga);
        else if (k == "pml_nodes") *value = double(s->Gpml);
        else if (k == "fields") *value = s->nf;
        else if (k == "nodes") *value = s->G;
        else if (k == "elements") *value = s->E;
        else if (k == "assembly") *value = s->assembly;
        else throw Error(GFX_ERR_INVALID, "unknown query key: " + k);
    });
}

extern "C" double gfx_solver_algorithmic_bytes(const gfx_solver* s, int which) {
    if (!s) return 0.0;
    const double P = double(s->E) * NP, G = s->G, F = s->nf, Ga = double(s->Ga),
                 Gp = double(s->Gpml), Q = double(s->Ppml);
    // element stage: geometry (9 scaled dxi_dx + lambda + mu + node id) per point,
    // 3 + 3 values at points of axis-aligned elements on the diagonal form; per
    // field the predicted displacement read and the residual written once per node.
    const double Pd = double(s->n_diag) * NP;
    double elem = 48.0 * (P - Pd) + 24.0 * Pd + F * 24.0 * G;
    if (ref_elem_pml(*s)) elem += 56.0 * Q + F * (96.0 * Q + 36.0 * Gp);
    // complete C-PML: 28 coefficients per PML point, and per field the
    // memories in use (of 45) read and written and w read
    if (s->cpml_complete) elem += 112.0 * Q + F * (8.0 * double(s->cpml_mem_active) + 12.0 * Q);
    // shear attenuation: the general form (48 B) at every point, 3 step
    // coefficients per solid per point, and per field the memory (5 per solid)
    // and previous deviatoric strain (5) read and written at every point
    if (s->n_sls > 0) {
        const double L = s->n_sls;
        elem = 48.0 * P + F * 24.0 * G + 12.0 * L * P + F * 8.0 * (5.0 * L + 5.0) * P;
    }
    // node stage: 1/m per node; per field r, u~, w read and u~, w written;
    // special nodes also read and write a; C-PML nodes read 5 + 9 coefficients
    // and per field read and write the memory (9) and "new" (3) values.
    double node = 4.0 * G + F * 60.0 * G + F * 24.0 * Ga;
    if (s->has_cpml && s->pml_nodal) node += 56.0 * Gp + F * 96.0 * Gp;
    // the node stage split by kernel: special nodes [0, Ga) carry a and the
    // C-PML terms, regular nodes [Ga, G) only r, u~ and w
    const double node_reg = (4.0 + F * 60.0) * (G - Ga);
    if (which == 1) return elem;
    if (which == 2) return node;
    if (which == 3) return node - node_reg;
    if (which == 4) return node_reg;
    return elem + node;
}

// ---------------------------------------------------------------------------
// Batches of independent solvers (uneven meshes): one step of every active
// instance is captured as parallel branches of one graph on the batch stream.
// ---------------------------------------------------------------------------
struct gfx_batch {
    std::vector<gfx_solver*> s;
    hipStream_t stream = nullptr;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    // launch groups when the solvers allow shared launches: 0 = one branch of
    // launches per solver (default: the branches overlap one mesh's element
    // stage with another's node stages), k > 0 = k branches of launches shared
    // by their solvers (fewer graph nodes for many small meshes)
    int want_groups = 0;
    int last_groups = 0;
    // One branch of shared launches: argument tables and entries its kernels read.
    struct Group {
        std::vector<int> members;  // solver indices
        gfx::DevBuf<gfx::dev::ElemArgs> etab;
        gfx::DevBuf<gfx::dev::NodeArgs> ntab;
        // per-workgroup work of each launch (general and diagonal element
        // forms, special and regular nodes)
        gfx::DevBuf<int4> map_gen, map_diag, map_spec, map_reg;
        gfx::DevBuf<int*> ctrs;
        gfx::DevBuf<int> strides;
        // launch shape: workgroups per launch, kernel form
        int bg = 0, bd = 0, bs = 0, br = 0, n_ctr = 0;
        int fg = 1, kernel = 0, epb = 1;
    };
    struct Cached {
        std::vector<char> active;
        std::vector<int> gens;
        int groups = 0;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
        std::vector<std::unique_ptr<Group>> grp;
    };
    std::vector<std::unique_ptr<Cached>> cache;
    ~gfx_batch() {
        if (stream) (void)hipStreamSynchronize(stream);
        for (auto& c : cache) {
            if (c->exec) (void)hipGraphExecDestroy(c->exec);
            if (c->graph) (void)hipGraphDestroy(c->graph);
        }
        for (auto e : join)
            if (e) (void)hipEventDestroy(e);
        if (fork) (void)hipEventDestroy(fork);
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace gfx {
namespace {

// Can the active solvers share one launch per stage?  GATHER assembly (no
// colours, no zero fill), nodal C-PML or none, no attenuation, and one launch
// shape: the same element kernel form (block or wavefront), elements per
// workgroup and node field grouping, so that every solver's arithmetic is the
// one of its own launches.
bool groupable(const gfx_batch& B, const std::vector<char>& active) {
    const gfx_solver* first = nullptr;
    for (size_t i = 0; i < B.s.size(); ++i) {
        if (!active[i]) continue;
        const gfx_solver& S = *B.s[i];
        if (S.assembly != GFX_ASSEMBLY_GATHER || (S.has_cpml && !S.pml_nodal) || S.n_sls > 0) return false;
        if (S.kernel != GFX_ELEMENT_BLOCK && S.kernel != GFX_ELEMENT_WAVE) return false;
        if (first && (S.kernel != first->kernel || S.epb != first->epb || S.node_fg != first->node_fg)) return false;
        if (!first) first = &S;
    }
    return first != nullptr;
}

template <bool DIAG>
void launch_element_grp(int kernel, int epb, gfx_batch::Group& c, const gfx::DevBuf<int4>& map, int grid,
                        hipStream_t st) {
    constexpr int M = dev::ASM_GATHER;
    if (kernel == GFX_ELEMENT_WAVE) {
        switch (epb) {
            case 1: dev::k_element_wave_grp<1, M, DIAG><<<grid, 64, 0, st>>>(c.etab.p, map.p); break;
            case 2: dev::k_element_wave_grp<2, M, DIAG><<<grid, 128, 0, st>>>(c.etab.p, map.p); break;
            case 4: dev::k_element_wave_grp<4, M, DIAG><<<grid, 256, 0, st>>>(c.etab.p, map.p); break;
            default: dev::k_element_wave_grp<8, M, DIAG><<<grid, 512, 0, st>>>(c.etab.p, map.p); break;
        }
    } else {
        constexpr auto T = [](int e) { return ((e * NP + 63) / 64) * 64; };
        switch (epb) {
            case 1: dev::k_element_grp<1, M, DIAG><<<grid, T(1), 0, st>>>(c.etab.p, map.p); break;
            case 2: dev::k_element_grp<2, M, DIAG><<<grid, T(2), 0, st>>>(c.etab.p, map.p); break;
            case 4: dev::k_element_grp<4, M, DIAG><<<grid, T(4), 0, st>>>(c.etab.p, map.p); break;
            default: dev::k_element_grp<8, M, DIAG><<<grid, T(8), 0, st>>>(c.etab.p, map.p); break;
        }
    }
}

template <int FG>
void launch_nodes_grp(gfx_batch::Group& c, hipStream_t st) {
    if (c.bs > 0) dev::k_node_special_grp<FG><<<c.bs, 256, 0, st>>>(c.ntab.p, c.map_spec.p);
    if (c.br > 0) dev::k_node_regular_grp<true, FG><<<c.br, 256, 0, st>>>(c.ntab.p, c.map_reg.p);
}

// Argument tables and entries of a group of solvers, uploaded before the
// capture that records the group's launches.
void build_group(gfx_batch& B, gfx_batch::Group& c, const std::vector<StepArgs>& args) {
    std::vector<dev::ElemArgs> et;
    std::vector<dev::NodeArgs> nt;
    std::vector<int4> eg, ed, es, er;  // one entry per workgroup
    std::vector<int*> ctr;
    std::vector<int> strides;
    int fg = 1, kernel = GFX_ELEMENT_BLOCK, epb = 1;
    for (const int i : c.members) {
        gfx_solver& S = *B.s[i];
        const int t = int(et.size());
        et.push_back(args[i].ea);
        nt.push_back(args[i].na);
        fg = S.node_fg;
        kernel = S.kernel;
        epb = S.epb;
        for (const auto& rg : S.ranges) {
            const int n = rg.e - rg.b;
            auto& m = rg.diag ? ed : eg;
            for (int w = 0; w < (n + S.epb - 1) / S.epb; ++w) m.push_back(make_int4(t, rg.b, n, w));
        }
        for (int w = 0; w < args[i].spec_grid; ++w) es.push_back(make_int4(t, 0, 0, w));
        for (int w = 0; w < args[i].reg_grid; ++w) er.push_back(make_int4(t, 0, 0, w));
        ctr.push_back(S.ctr.p);
        strides.push_back(S.stride);
    }
    auto put = [](auto& buf, const auto& v) {
        buf.alloc(std::max<size_t>(v.size(), 1));
        if (!v.empty()) GFX_HIP(hipMemcpy(buf.p, v.data(), v.size() * sizeof(v[0]), hipMemcpyHostToDevice));
    };
    put(c.etab, et);
    put(c.ntab, nt);
    put(c.map_gen, eg);
    put(c.map_diag, ed);
    put(c.map_spec, es);
    put(c.map_reg, er);
    put(c.ctrs, ctr);
    put(c.strides, strides);
    c.bg = int(eg.size());
    c.bd = int(ed.size());
    c.bs = int(es.size());
    c.br = int(er.size());
    c.n_ctr = int(ctr.size());
    c.fg = fg;
    c.kernel = kernel;
    c.epb = epb;
}

// One step of a group of solvers as shared launches on stream st: the element
// ranges of every member (general form, then diagonal form), all special nodes,
// all regular nodes, then each member's recording and one counter update for all.
void enqueue_group(gfx_batch::Group& c, const std::vector<StepArgs>& args, hipStream_t st) {
    if (c.bg > 0) launch_element_grp<false>(c.kernel, c.epb, c, c.map_gen, c.bg, st);
    if (c.bd > 0) launch_element_grp<true>(c.kernel, c.epb, c, c.map_diag, c.bd, st);
    switch (c.fg) {
        case 2: launch_nodes_grp<2>(c, st); break;
        case 4: launch_nodes_grp<4>(c, st); break;
        case 8: launch_nodes_grp<8>(c, st); break;
        default: launch_nodes_grp<1>(c, st); break;
    }
    for (const int i : c.members)
        if (args[i].recording && args[i].rec_grid > 0)
            dev::k_record<REPB><<<args[i].rec_grid, args[i].rec_threads, 0, st>>>(args[i].ra);
    dev::k_advance_grp<<<(c.n_ctr + 63) / 64, 64, 0, st>>>(c.ctrs.p, c.strides.p, c.n_ctr);
    GFX_HIP(hipGetLastError());
}

// Splits the active solvers into k groups of about equal work (element-field
// updates per step): largest first, each to the least loaded group.
std::vector<std::vector<int>> split_groups(const gfx_batch& B, const std::vector<char>& active, int k) {
    std::vector<int> idx;
    for (size_t i = 0; i < B.s.size(); ++i)
        if (active[i]) idx.push_back(int(i));
    auto work = [&](int i) { return double(B.s[i]->E) * B.s[i]->nf; };
    std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return work(a) > work(b); });
    k = std::max(1, std::min<int>(k, int(idx.size())));
    std::vector<std::vector<int>> g(k);
    std::vector<double> load(k, 0.0);
    for (const int i : idx) {
        const int j = int(std::min_element(load.begin(), load.end()) - load.begin());
        g[j].push_back(i);
        load[j] += work(i);
    }
    for (auto& m : g) std::sort(m.begin(), m.end());
    return g;
}

hipGraphExec_t batch_graph(gfx_batch& B, const std::vector<char>& active) {
    std::vector<int> gens(B.s.size());
    for (size_t i = 0; i < B.s.size(); ++i) gens[i] = B.s[i]->gen;
    const int groups = (B.want_groups > 0 && groupable(B, active)) ? B.want_groups : 0;
    for (auto& c : B.cache)
        if (c->active == active && c->gens == gens && c->groups == groups) {
            B.last_groups = int(c->grp.size());
            return c->exec;
        }
    auto cp = std::make_unique<gfx_batch::Cached>();
    gfx_batch::Cached& c = *cp;
    c.active = active;
    c.gens = gens;
    c.groups = groups;
    std::vector<StepArgs> args(B.s.size());
    for (size_t i = 0; i < B.s.size(); ++i)
        if (active[i]) args[i] = make_step_args(*B.s[i]);
    if (groups > 0) {
        for (auto& m : split_groups(B, active, groups)) {
            c.grp.push_back(std::make_unique<gfx_batch::Group>());
            c.grp.back()->members = std::move(m);
            build_group(B, *c.grp.back(), args);
        }
        // one branch per group, on the stream of the group's first solver
        GFX_HIP(hipStreamBeginCapture(B.stream, hipStreamCaptureModeThreadLocal));
        GFX_HIP(hipEventRecord(B.fork, B.stream));
        for (auto& g : c.grp) {
            const int lead = g->members.front();
            hipStream_t st = B.s[lead]->stream;
            GFX_HIP(hipStreamWaitEvent(st, B.fork, 0));
            enqueue_group(*g, args, st);
            GFX_HIP(hipEventRecord(B.join[lead], st));
            GFX_HIP(hipStreamWaitEvent(B.stream, B.join[lead], 0));
        }
        GFX_HIP(hipStreamEndCapture(B.stream, &c.graph));
        GFX_HIP(hipGraphInstantiate(&c.exec, c.graph, nullptr, nullptr, 0));
        B.last_groups = int(c.grp.size());
        B.cache.push_back(std::move(cp));
        return c.exec;
    }
    GFX_HIP(hipStreamBeginCapture(B.stream, hipStreamCaptureModeThreadLocal));
    GFX_HIP(hipEventRecord(B.fork, B.stream));
    for (size_t i = 0; i < B.s.size(); ++i) {
        if (!active[i]) continue;
        gfx_solver& S = *B.s[i];
        GFX_HIP(hipStreamWaitEvent(S.stream, B.fork, 0));
        enqueue_step(S, args[i], true, nullptr);
        GFX_HIP(hipEventRecord(B.join[i], S.stream));
        GFX_HIP(hipStreamWaitEvent(B.stream, B.join[i], 0));
    }
    GFX_HIP(hipStreamEndCapture(B.stream, &c.graph));
    GFX_HIP(hipGraphInstantiate(&c.exec, c.graph, nullptr, nullptr, 0));
    B.cache.push_back(std::move(cp));
    B.last_groups = 0;
    return c.exec;
}

}  // namespace
}  // namespace gfx

extern "C" gfx_status gfx_batch_create(int n, gfx_solver* const* solvers, gfx_batch** out) {
    return guarded([&] {
        require(n >= 1 && solvers && out, "bad argument");
        *out = nullptr;
        auto B = std::make_unique<gfx_batch>();
        for (int i = 0; i < n; ++i) {
            require(solvers[i] != nullptr, "null solver in batch");
            for (int k = 0; k < i; ++k) require(solvers[k] != solvers[i], "a solver appears twice in the batch");
            B->s.push_back(solvers[i]);
        }
        GFX_HIP(hipStreamCreateWithFlags(&B->stream, hipStreamNonBlocking));
        GFX_HIP(hipEventCreateWithFlags(&B->fork, hipEventDisableTiming));
        B->join.resize(n, nullptr);
        for (auto& e : B->join) GFX_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        *out = B.release();
    });
}

extern "C" gfx_status gfx_batch_run(gfx_batch* b, int n_steps, const int* until) {
    return guarded([&] {
        require(b != nullptr, "null batch");
        require(n_steps >= 0, "n_steps must be >= 0");
        for (auto* S : b->s) GFX_HIP(hipStreamSynchronize(S->stream));
        std::vector<char> active(b->s.size());
        for (int t = 0; t < n_steps; ++t) {
            bool any = false;
            for (size_t i = 0; i < b->s.size(); ++i) {
                active[i] = (!until || b->s[i]->step < until[i]) ? 1 : 0;
                any = any || active[i];
            }
            if (!any) break;
            GFX_HIP(hipGraphLaunch(batch_graph(*b, active), b->stream));
            for (size_t i = 0; i < b->s.size(); ++i)
                if (active[i]) advance_host(*b->s[i]);
        }
        GFX_HIP(hipStreamSynchronize(b->stream));
    });
}

extern "C" gfx_status gfx_batch_set_groups(gfx_batch* b, int n_groups) {
    return guarded([&] {
        require(b != nullptr, "null batch");
        require(n_groups >= 0, "n_groups must be >= 0");
        b->want_groups = n_groups;
    });
}

extern "C" int gfx_batch_groups(const gfx_batch* b) { return b ? b->last_groups : 0; }

extern "C" gfx_status gfx_batch_destroy(gfx_batch* b) {
    return guarded([&] { delete b; });
}

// ---------------------------------------------------------------------------
// Green's tensor capture: node values of a tile from one snapshot of the
// solver's record ring (mass-weighted merge, postprocess/cpp/main.cpp:354-426).
// ---------------------------------------------------------------------------
namespace gfx {
namespace dev {
__global__ void k_green_capture(const float* __restrict__ ring, int64_t nrp, const int64_t* __restrict__ off,
                                const int32_t* __restrict__ pt, const float* __restrict__ w, int n_node, int f0,
                                int f1, int f2, float* __restrict__ out) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= n_node) return;
    const int fd[3] = {f0, f1, f2};
    float acc[18];
#pragma unroll
    for (int k = 0; k < 18; ++k) acc[k] = 0.f;
    for (int64_t e = off[n]; e < off[n + 1]; ++e) {
        const float we = w[e];
        const int64_t q = pt[e];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            if (fd[d] < 0) continue;
            const float* s6 = ring + (int64_t(fd[d]) * nrp + q) * 6;
#pragma unroll
            for (int c = 0; c < 6; ++c) acc[c * 3 + d] += we * s6[c];
        }
    }
    float* o = out + int64_t(n) * 18;
#pragma unroll
    for (int k = 0; k < 18; ++k) o[k] = acc[k];
}
}  // namespace dev
}  // namespace gfx

extern "C" gfx_status gfx_green_set_capture_map(gfx_green* g, const int64_t* off, const int32_t* rec_point,
                                                const double* weight) {
    return guarded([&] {
        require(g && off && rec_point && weight, "null argument");
        require(off[0] == 0, "off[0] must be 0");
        const int64_t ne = off[g->n_node];
        std::vector<int64_t> ho(off, off + g->n_node + 1);
        std::vector<int32_t> hp(rec_point, rec_point + ne);
        std::vector<float> hw(ne);
        for (int n = 0; n < g->n_node; ++n) {
            require(ho[n + 1] >= ho[n], "off must be non-decreasing");
            double sw = 0;
            for (int64_t e = ho[n]; e < ho[n + 1]; ++e) sw += weight[e];
            for (int64_t e = ho[n]; e < ho[n + 1]; ++e) hw[e] = sw > 0 ? float(weight[e] / sw) : 0.f;
        }
        int64_t mx = -1;
        for (int32_t q : hp) {
            require(q >= 0, "negative recording point");
            mx = std::max<int64_t>(mx, q);
        }
        g->map_max = mx;
        g->map_off.upload(ho);
        g->map_pt.upload(hp.empty() ? std::vector<int32_t>{0} : hp);
        g->map_w.upload(hw.empty() ? std::vector<float>{0.f} : hw);
    });
}

extern "C" gfx_status gfx_green_capture(gfx_green* g, const gfx_solver* s, int snap, const int32_t* field_of_dir,
                                        int time_index) {
    return guarded([&] {
        require(g && s && field_of_dir, "null argument");
        require(g->map_off.p != nullptr, "capture map not set");
        require(time_index >= 0 && time_index < g->n_time, "time index out of range");
        require(s->rec_cap > 0 && s->n_rec > 0, "the solver records nothing");
        require(snap >= 0 && snap < s->n_snap && snap >= s->n_snap - s->rec_cap, "snapshot not held in the record ring");
        for (int d = 0; d < 3; ++d) require(field_of_dir[d] >= -1 && field_of_dir[d] < s->nf, "bad field index");
        const int64_t nrp = int64_t(s->n_rec) * NP;
        // every map entry must address the solver's recording points
        require(g->map_max < nrp, "capture map addresses recording points the solver does not have");
        const int slot = snap % s->rec_cap;
        const float* ring = s->rec_strain.p + int64_t(slot) * s->nf * nrp * 6;
        GFX_HIP(hipStreamSynchronize(s->stream));
        GFX_HIP(hipStreamSynchronize(g->stream));
        dev::k_green_capture<<<(g->n_node + 255) / 256, 256, 0, g->stream>>>(
            ring, nrp, g->map_off.p, g->map_pt.p, g->map_w.p, g->n_node, field_of_dir[0], field_of_dir[1],
            field_of_dir[2], g->val.p + int64_t(time_index) * g->n_node * 18);
        GFX_HIP(hipGetLastError());
        GFX_HIP(hipStreamSynchronize(g->stream));
    });
}


// ---------------------------------------------------------------------------
// Exact tile slabs: the post-processor's merge (postprocess/cpp/main.cpp:
// 306-426) and Green's tensor layout (main.cpp:554-575, 607-618, 800-811) on
// the device, from one snapshot of the solver's record ring.
// ---------------------------------------------------------------------------
struct gfx_green_slab {
    int n_node = 0;
    int64_t n_slab = 0;
    bool f64 = false, disp = false;
    int64_t map_max = -1;
    gfx::DevBuf<int64_t> off;
    gfx::DevBuf<int32_t> pt, slab_node;
    gfx::DevBuf<double> w, inv;
    gfx::DevBuf<uint8_t> out_g, out_d;
    hipStream_t stream = nullptr;
    ~gfx_green_slab() {
        if (stream) (void)hipStreamDestroy(stream);
    }
};

namespace gfx {
namespace dev {
// One thread per slab entry j (a node of one tile; nodes on tile borders
// appear once per tile).  The sums run over the node's entries in the order
// given, each term rounded as the FP64 host loop rounds it (product, then
// sum; no fused multiply-add), then scale by inv = 1 / sum of weights.
// Strain: dst[d*6 + c] (force direction d major, main.cpp:556-575);
// displacement: plain sums of the records scaled by the same inv,
// dst[c*3 + d] (main.cpp:382-399, 414-423, 581-598).
template <typename V>
__global__ void k_green_slab(const float* __restrict__ strain, const float* __restrict__ uva, int64_t nrp, int f0,
                             int f1, int f2, const int64_t* __restrict__ off, const int32_t* __restrict__ pt,
                             const double* __restrict__ w, const double* __restrict__ inv,
                             const int32_t* __restrict__ slab_node, int64_t n_slab, V* __restrict__ out_g,
                             V* __restrict__ out_d) {
    // products and sums are rounded one by one, as the host loop rounds them:
    // no fused multiply-add
#pragma clang fp contract(off)
    const int64_t j = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (j >= n_slab) return;
    const int n = slab_node[j];
    const int fd[3] = {f0, f1, f2};
    double g[18], u[9];
#pragma unroll
    for (int k = 0; k < 18; ++k) g[k] = 0.0;
#pragma unroll
    for (int k = 0; k < 9; ++k) u[k] = 0.0;
    for (int64_t e = off[n]; e < off[n + 1]; ++e) {
        const double we = w[e];
        const int64_t q = pt[e];
#pragma unroll
        for (int d = 0; d < 3; ++d) {
            if (fd[d] < 0) continue;
            const float* s6 = strain + (int64_t(fd[d]) * nrp + q) * 6;
#pragma unroll
            for (int c = 0; c < 6; ++c) g[d * 6 + c] += double(s6[c]) * we;
            if (out_d) {
                const float* u3 = uva + (int64_t(fd[d]) * nrp + q) * 9;
#pragma unroll
                for (int c = 0; c < 3; ++c) u[c * 3 + d] += double(u3[c]);
            }
        }
    }
    const double s = inv[n];
    V* og = out_g + j * 18;
#pragma unroll
    for (int k = 0; k < 18; ++k) og[k] = V(g[k] * s);
    if (out_d) {
        V* od = out_d + j * 9;
#pragma unroll
        for (int k = 0; k < 9; ++k) od[k] = V(u[k] * s);
    }
}
}  // namespace dev
}  // namespace gfx

extern "C" gfx_status gfx_green_slab_create(int n_node, const int64_t* off, const int32_t* rec_point,
                                            const double* weight, int64_t n_slab, const int32_t* slab_node,
                                            int precision_f64, int with_displacement, gfx_green_slab** out) {
    return guarded([&] {
        require(out && off && slab_node && (off[n_node] == 0 || (rec_point && weight)), "null argument");
        *out = nullptr;
        require(n_node >= 1 && n_slab >= 1, "empty slab");
        require(off[0] == 0, "off[0] must be 0");
        auto T = std::make_unique<gfx_green_slab>();
        T->n_node = n_node;
        T->n_slab = n_slab;
        T->f64 = precision_f64 != 0;
        T->disp = with_displacement != 0;
        const int64_t ne = off[n_node];
        std::vector<int64_t> ho(off, off + n_node + 1);
        std::vector<double> inv(n_node, 1.0);
        for (int n = 0; n < n_node; ++n) {
            require(ho[n + 1] >= ho[n], "off must be non-decreasing");
            // node_weight_sum in entry order; no scaling when it is not positive (main.cpp:414-416)
            double sw = 0.0;
            for (int64_t e = ho[n]; e < ho[n + 1]; ++e) {
                require(std::isfinite(weight[e]), "weights must be finite");
                sw += weight[e];
            }
            if (sw > 0.0) inv[n] = 1.0 / sw;
        }
        std::vector<int32_t> hp(rec_point, rec_point + ne);
        for (int32_t q : hp) {
            require(q >= 0, "negative recording point");
            T->map_max = std::max<int64_t>(T->map_max, q);
        }
        std::vector<int32_t> hs(slab_node, slab_node + n_slab);
        for (int32_t n : hs) require(n >= 0 && n < n_node, "slab node out of range");
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
            (void)hipGetLastError();
            throw gfx::Error(GFX_ERR_DEVICE, "no HIP device available");
        }
        GFX_HIP(hipStreamCreateWithFlags(&T->stream, hipStreamNonBlocking));
        T->off.upload(ho);
        T->pt.upload(hp.empty() ? std::vector<int32_t>{0} : hp);
        T->w.upload(ne ? std::vector<double>(weight, weight + ne) : std::vector<double>{0.0});
        T->inv.upload(inv);
        T->slab_node.upload(hs);
        const size_t vb = T->f64 ? 8 : 4;
        T->out_g.alloc(size_t(n_slab) * 18 * vb);
        if (T->disp) T->out_d.alloc(size_t(n_slab) * 9 * vb);
        *out = T.release();
    });
}

extern "C" gfx_status gfx_green_slab_destroy(gfx_green_slab* t) {
    return guarded([&] { delete t; });
}

extern "C" gfx_status gfx_green_slab_capture(gfx_green_slab* t, const gfx_solver* s, int snap,
                                             const int32_t* field_of_dir, void* greens, void* displacement) {
    return guarded([&] {
        require(t && s && field_of_dir && greens, "null argument");
        require(!t->disp || displacement, "the slab was made with displacement: pass its buffer");
        require(t->disp || !displacement, "the slab was made without displacement");
        require(s->rec_cap > 0 && s->n_rec > 0, "the solver records nothing");
        require(!t->disp || s->rec_uva_on, "displacement needs the solver's u, v, a records");
        require(snap >= 0 && snap < s->n_snap && snap >= s->n_snap - s->rec_cap, "snapshot not held in the record ring");
        for (int d = 0; d < 3; ++d) require(field_of_dir[d] >= -1 && field_of_dir[d] < s->nf, "bad field index");
        const int64_t nrp = int64_t(s->n_rec) * NP;
        require(t->map_max < nrp, "slab map addresses recording points the solver does not have");
        const int slot = snap % s->rec_cap;
        const float* strain = s->rec_strain.p + int64_t(slot) * s->nf * nrp * 6;
        const float* uva = t->disp ? s->rec_uva.p + int64_t(slot) * s->nf * nrp * 9 : nullptr;
        GFX_HIP(hipStreamSynchronize(s->stream));
        const int64_t blocks = (t->n_slab + 255) / 256;
        const size_t vb = t->f64 ? 8 : 4;
        if (t->f64)
            dev::k_green_slab<double><<<dim3(unsigned(blocks)), 256, 0, t->stream>>>(
                strain, uva, nrp, field_of_dir[0], field_of_dir[1], field_of_dir[2], t->off.p, t->pt.p, t->w.p,
                t->inv.p, t->slab_node.p, t->n_slab, reinterpret_cast<double*>(t->out_g.p),
                t->disp ? reinterpret_cast<double*>(t->out_d.p) : nullptr);
        else
            dev::k_green_slab<float><<<dim3(unsigned(blocks)), 256, 0, t->stream>>>(
                strain, uva, nrp, field_of_dir[0], field_of_dir[1], field_of_dir[2], t->off.p, t->pt.p, t->w.p,
                t->inv.p, t->slab_node.p, t->n_slab, reinterpret_cast<float*>(t->out_g.p),
                t->disp ? reinterpret_cast<float*>(t->out_d.p) : nullptr);
        GFX_HIP(hipGetLastError());
        GFX_HIP(hipMemcpyAsync(greens, t->out_g.p, size_t(t->n_slab) * 18 * vb, hipMemcpyDeviceToHost, t->stream));
        if (t->disp)
            GFX_HIP(hipMemcpyAsync(displacement, t->out_d.p, size_t(t->n_slab) * 9 * vb, hipMemcpyDeviceToHost,
                                   t->stream));
        GFX_HIP(hipStreamSynchronize(t->stream));
    });
}
