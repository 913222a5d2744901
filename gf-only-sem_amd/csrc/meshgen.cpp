// meshgen.cpp -- structured hex-box inputs in the reference's partition layout.
//
// The reference builds its inputs with a Python preprocessor (out of scope;
// its numerics are the specification restated here for regular boxes):
//   geometry + lumped mass    preprocess/gll_geometry.py:132-186
//   material, lambda/mu       preprocess/cli.py:283,369-381,448 (mass *= rho)
//   CFL solver dt / stride    preprocess/cfl_validator.py:21-115
//   PML element flags         preprocess/cli.py:236-264
//   legacy damping ramp       preprocess/pml.py:23-124
//   C-PML K/d/alpha, region,  preprocess/pml_cpml.py:42-155,182-248,
//   convolution and A1..A5      257-349,357-513 (SPECFEM3D-style formulas,
//                               including the face-overwrite order and the
//                               safe-denominator clamps)
//   source cells + weights    preprocess/source_locator.py:244-375
//   recording cells           preprocess/recording_map.py:18-110
// Node numbering is an analytic lattice numbering (the reference's quantised
// lexsort in preprocess/partition.py:192-254 yields some other permutation;
// solver results do not depend on it).
#include <cstring>
#include <functional>
#include <limits>
#include <memory>
#include <numeric>
#include <set>

#include "common.hpp"
#include "gfx_io.h"
#include "mesh.hpp"
#include "topology.hpp"

namespace gfx {
namespace {

// --- deterministic RNG (splitmix64) -------------------------------------
struct Rng {
    uint64_t s;
    explicit Rng(uint64_t seed) : s(seed * 0x9E3779B97F4A7C15ULL + 0x1234567ULL) {}
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    double uniform() { return (next() >> 11) * (1.0 / 9007199254740992.0); }
    double normal() {
        double u1 = uniform(), u2 = uniform();
        if (u1 < 1e-300) u1 = 1e-300;
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(2.0 * M_PI * u2);
    }
};

// Linear hex shape functions, GMSH corner order (preprocess/gll_geometry.py:53-68).
const double kCorner[8][3] = {{-1, -1, -1}, {1, -1, -1}, {1, 1, -1}, {-1, 1, -1},
                              {-1, -1, 1},  {1, -1, 1},  {1, 1, 1},  {-1, 1, 1}};

void shape(double xi, double eta, double zeta, double S[8], double dS[8][3]) {
    for (int a = 0; a < 8; ++a) {
        const double ca = kCorner[a][0], cb = kCorner[a][1], cc = kCorner[a][2];
        S[a] = 0.125 * (1 + ca * xi) * (1 + cb * eta) * (1 + cc * zeta);
        dS[a][0] = 0.125 * ca * (1 + cb * eta) * (1 + cc * zeta);
        dS[a][1] = 0.125 * (1 + ca * xi) * cb * (1 + cc * zeta);
        dS[a][2] = 0.125 * (1 + ca * xi) * (1 + cb * eta) * cc;
    }
}

// PML region codes (preprocess/pml_cpml.py:27-34).
enum { R_X = 1, R_Y = 2, R_Z = 3, R_XY = 4, R_XZ = 5, R_YZ = 6, R_XYZ = 7 };

bool axis_active(int region, int axis) {  // pml_cpml.py:159-175
    switch (region) {
        case R_XYZ: return true;
        case R_X: return axis == 0;
        case R_Y: return axis == 1;
        case R_Z: return axis == 2;
        case R_XY: return axis == 0 || axis == 1;
        case R_XZ: return axis == 0 || axis == 2;
        case R_YZ: return axis == 1 || axis == 2;
        default: return false;
    }
}

constexpr double kMinDist = 1e-6;  // pml_cpml.py:24
inline double safe_den(double v) { return std::abs(v) < kMinDist ? kMinDist : v; }

// Recursive-convolution coefficients (pml_cpml.py:257-304).
void conv_coef(double b, double dt, double c[3]) {
    const double temp = std::exp(-0.5 * b * dt);
    c[0] = temp * temp;
    if (std::abs(b) >= kMinDist) {
        c[1] = (1.0 - temp) / b;
        c[2] = c[1] * temp;
    } else {
        const double dt2 = dt * dt, dt3 = dt2 * dt, dt4 = dt3 * dt, b2 = b * b, b3 = b2 * b;
        c[1] = dt * 0.5 + (-0.125 * dt2 * b + (1.0 / 48.0) * dt3 * b2 - (1.0 / 384.0) * dt4 * b3);
        c[2] = dt * 0.5 + (-0.375 * dt2 * b + (7.0 / 48.0) * dt3 * b2 - (5.0 / 128.0) * dt4 * b3);
    }
}

// Partial fractions of s_1 s_2 / s_3 (pml_cpml.py:600-692, _lijk_parameter,
// always in its 1-2-3 form; callers permute the axes): A0 and the residues at
// alpha_1, alpha_2 and beta_3.  Directions the region does not damp drop out.
void lijk(int region, const double k1, const double d1, const double a1, const double k2, const double d2,
          const double a2, const double k3, const double d3, const double a3, double* A) {
    const double b1 = a1 + d1 / std::max(k1, 1.0), b2 = a2 + d2 / std::max(k2, 1.0),
                 b3 = a3 + d3 / std::max(k3, 1.0);
    A[0] = A[1] = A[2] = A[3] = 0.0;
    switch (region) {
        case R_XYZ: {
            A[0] = k1 * k2 / std::max(k3, 1.0);
            A[1] = -A[0] * (a1 - a3) * (a1 - b1) * (a1 - b2) / (safe_den(a1 - a2) * safe_den(a1 - b3));
            A[2] = -A[0] * (a2 - a3) * (a2 - b1) * (a2 - b2) / (safe_den(a2 - a1) * safe_den(a2 - b3));
            A[3] = -A[0] * (b3 - a3) * (b3 - b1) * (b3 - b2) / (safe_den(b3 - a1) * safe_den(b3 - a2));
            break;
        }
        case R_X:
            A[0] = k1;
            A[1] = -A[0] * (a1 - b1);
            break;
        case R_Y:
            A[0] = k2;
            A[2] = -A[0] * (a2 - b2);
            break;
        case R_Z:
            A[0] = 1.0 / std::max(k3, 1.0);
            A[3] = -A[0] * (b3 - a3);
            break;
        case R_XY:
            A[0] = k1 * k2;
            A[1] = -A[0] * (a1 - b1) * (a1 - b2) / safe_den(a1 - a2);
            A[2] = -A[0] * (a2 - b1) * (a2 - b2) / safe_den(a2 - a1);
            break;
        case R_XZ:
            A[0] = k1 / std::max(k3, 1.0);
            A[1] = -A[0] * (a1 - a3) * (a1 - b1) / safe_den(a1 - b3);
            A[3] = -A[0] * (b3 - a3) * (b3 - b1) / safe_den(b3 - a1);
            break;
        case R_YZ:
            A[0] = k2 / std::max(k3, 1.0);
            A[2] = -A[0] * (a2 - a3) * (a2 - b2) / safe_den(a2 - b3);
            A[3] = -A[0] * (b3 - a3) * (b3 - b2) / safe_den(b3 - a2);
            break;
        default:
            break;
    }
}

// (K, d) of the stretch along one axis where the region damps it, (1, 0)
// elsewhere (pml_cpml.py:695-734, _lx/_ly/_lz_parameter).
void lone(int region, int axis, double k, double d, double a, double* A) {
    if (region != 0 && axis_active(region, axis)) {
        A[0] = k;
        A[1] = -k * (a - (a + d / std::max(k, 1.0)));
    } else {
        A[0] = 1.0;
        A[1] = 0.0;
    }
}

// The 18 strain coefficients of one point (pml_cpml.py:523-597).  The
// region's index permutations 231, 132, 123 map (z, y, x), (x, z, y), (x, y, z)
// onto lijk's (1, 2, 3); the region code itself is passed unchanged, as the
// reference does.
void strain_coef(int region, const double K[3], const double d[3], const double al[3], double S[18]) {
    for (int q = 0; q < 18; ++q) S[q] = 0.0;
    if (region == 0) return;
    lijk(region, K[2], d[2], al[2], K[1], d[1], al[1], K[0], d[0], al[0], S + 0);
    lijk(region, K[0], d[0], al[0], K[2], d[2], al[2], K[1], d[1], al[1], S + 4);
    lijk(region, K[0], d[0], al[0], K[1], d[1], al[1], K[2], d[2], al[2], S + 8);
    for (int ax = 0; ax < 3; ++ax) lone(region, ax, K[ax], d[ax], al[ax], S + 12 + 2 * ax);
}

// A1..A5 by region (pml_cpml.py:388-513, _l_parameter).
void abar(int region, const double K[3], const double d[3], const double al[3], double A[5]) {
    const double kx = K[0], ky = K[1], kz = K[2];
    const double ax = al[0], ay = al[1], az = al[2];
    const double bx = ax + d[0] / std::max(kx, 1.0);
    const double by = ay + d[1] / std::max(ky, 1.0);
    const double bz = az + d[2] / std::max(kz, 1.0);
    for (int i = 0; i < 5; ++i) A[i] = 0.0;
    switch (region) {
        case R_XYZ: {
            const double A0 = kx * ky * kz;
            A[0] = A0 * (bx + by + bz - ax - ay - az);
            A[1] = A0 * (bx - ax) * (by - ay - ax) + A0 * (by - ay) * (bz - az - ay) +
                   A0 * (bz - az) * (bx - ax - az);
            const double dxy = safe_den(ax - ay), dyz = safe_den(ay - az);
            const double dyx = safe_den(ay - ax), dzx = safe_den(az - ax), dzy = safe_den(az - ay);
            A[2] = A0 * ax * ax * (bx - ax) * (by - ax) * (bz - ax) / (dyx * dzx);
            A[3] = A0 * ay * ay * (bx - ay) * (by - ay) * (bz - ay) / (dxy * dzy);
            // (alpha_y - alpha_z)(alpha_z - alpha_x) as pml_cpml.py:433 divides: the
            // opposite sign of the residue at alpha_z, kept as the reference writes it
            A[4] = A0 * az * az * (bx - az) * (by - az) * (bz - az) / (dyz * dzx);
            break;
        }
        case R_XY: {
            const double A0 = kx * ky;
            A[0] = A0 * (bx + by - ax - ay);
            A[1] = A0 * (bx - ax) * (by - ay - ax) - A0 * (by - ay) * ay;
            A[2] = A0 * ax * ax * (bx - ax) * (by - ax) / safe_den(ay - ax);
            A[3] = A0 * ay * ay * (bx - ay) * (by - ay) / safe_den(ax - ay);
            break;
        }
        case R_XZ: {
            const double A0 = kx * kz;
            A[0] = A0 * (bx + bz - ax - az);
            A[1] = A0 * (bx - ax) * (-ax) + A0 * (bz - az) * (bx - ax - az);
            A[2] = A0 * ax * ax * (bx - ax) * (bz - ax) / safe_den(az - ax);
            A[4] = A0 * az * az * (bx - az) * (bz - az) / safe_den(ax - az);
            break;
        }
        case R_YZ: {
            const double A0 = ky * kz;
            A[0] = A0 * (by + bz - ay - az);
            A[1] = A0 * (by - ay) * (bz - az - ay) - A0 * (bz - az) * az;
            A[3] = A0 * ay * ay * (by - ay) * (bz - ay) / safe_den(az - ay);
            A[4] = A0 * az * az * (by - az) * (bz - az) / safe_den(ay - az);
            break;
        }
        case R_X: {
            const double diff = bx - ax;
            A[0] = kx * diff;
            A[1] = -kx * ax * diff;
            A[2] = kx * ax * ax * diff;
            break;
        }
        case R_Y: {
            const double diff = by - ay;
            A[0] = ky * diff;
            A[1] = -ky * ay * diff;
            A[3] = ky * ay * ay * diff;
            break;
        }
        case R_Z: {
            const double diff = bz - az;
            A[0] = kz * diff;
            A[1] = -kz * az * diff;
            A[4] = kz * az * az * diff;
            break;
        }
        default: break;
    }
}

// Constant-Q standard linear solids (Emmerich and Korn 1987, linearised least
// squares on 1/Q; see gfx_sls_fit).
void sls_fit(double q, double f_min, double f_max, int L, double* ts, double* te) {
    require(L >= 1 && L <= 8, "n_sls must be in [1, 8]");
    require(f_min > 0 && f_max >= f_min, "attenuation band must have 0 < f_min <= f_max");
    for (int l = 0; l < L; ++l) {
        const double f = L == 1 ? std::sqrt(f_min * f_max) : f_min * std::pow(f_max / f_min, double(l) / (L - 1));
        ts[l] = 1.0 / (2.0 * M_PI * f);
    }
    if (!(q > 0) || std::isinf(q)) {
        for (int l = 0; l < L; ++l) te[l] = ts[l];
        return;
    }
    const double qi = 1.0 / q;
    const int K = 10 * L;
    std::vector<double> ata(size_t(L) * L, 0.0), atb(L, 0.0);
    for (int k = 0; k < K; ++k) {
        const double f = K == 1 ? f_min : f_min * std::pow(f_max / f_min, double(k) / (K - 1));
        const double w = 2.0 * M_PI * f;
        std::vector<double> row(L);
        for (int l = 0; l < L; ++l) row[l] = (w * ts[l] + qi) / (1.0 + w * w * ts[l] * ts[l]);
        for (int a = 0; a < L; ++a) {
            atb[a] += row[a] * qi;
            for (int b = 0; b < L; ++b) ata[size_t(a) * L + b] += row[a] * row[b];
        }
    }
    // Gaussian elimination with partial pivoting
    std::vector<double> y = atb;
    for (int c = 0; c < L; ++c) {
        int piv = c;
        for (int r = c + 1; r < L; ++r)
            if (std::abs(ata[size_t(r) * L + c]) > std::abs(ata[size_t(piv) * L + c])) piv = r;
        require(std::abs(ata[size_t(piv) * L + c]) > 0, "attenuation fit is singular");
        if (piv != c) {
            for (int k = 0; k < L; ++k) std::swap(ata[size_t(c) * L + k], ata[size_t(piv) * L + k]);
            std::swap(y[c], y[piv]);
        }
        for (int r = c + 1; r < L; ++r) {
            const double m = ata[size_t(r) * L + c] / ata[size_t(c) * L + c];
            for (int k = c; k < L; ++k) ata[size_t(r) * L + k] -= m * ata[size_t(c) * L + k];
            y[r] -= m * y[c];
        }
    }
    for (int c = L - 1; c >= 0; --c) {
        for (int k = c + 1; k < L; ++k) y[c] -= ata[size_t(c) * L + k] * y[k];
        y[c] /= ata[size_t(c) * L + c];
    }
    double sy = 0;
    for (int l = 0; l < L; ++l) sy += y[l];
    require(sy < 1.0, "quality factor too low for the standard linear solids");
    for (int l = 0; l < L; ++l) te[l] = ts[l] * (1.0 + y[l] / (1.0 - sy));
}

// General cells: corner coordinates per cell in GMSH order and the domain
// bounds (the vertex extent, preprocess/cli.py:409-417).
struct GeneralCells {
    std::vector<double> cv;  // [E][8][3]
    double bnd[6];
};

// Boundary detection on general hexahedra (preprocess/boundary_detector.py:18-86):
// a face whose centre (mean of its four corners) lies on z = zmin is free
// surface; one on x = xmin/xmax, y = ymin/ymax or z = zmax (tolerance 1e-6 of
// the largest box extent) is absorbing; a cell with an absorbing face is a PML
// cell.  Faces by GMSH corner order: 0-3 and 4-7 are opposite quads.
void face_flags(const GeneralCells& g, std::vector<char>& is_pml) {
    static const int F[6][4] = {{0, 1, 2, 3}, {4, 5, 6, 7}, {0, 1, 5, 4},
                                {1, 2, 6, 5}, {2, 3, 7, 6}, {3, 0, 4, 7}};
    const double* b = g.bnd;
    const double dmax = std::max({b[1] - b[0], b[3] - b[2], b[5] - b[4]});
    const double tol = dmax > 0 ? 1e-6 * dmax : 1e-9;
    // np.isclose(a, v, atol=tol) with its default rtol = 1e-5
    auto close = [&](double a, double v) { return std::abs(a - v) <= tol + 1e-5 * std::abs(v); };
    const int64_t E = int64_t(g.cv.size() / 24);
    for (int64_t e = 0; e < E; ++e) {
        const double* cv = &g.cv[size_t(e) * 24];
        for (int f = 0; f < 6 && !is_pml[e]; ++f) {
            double c[3] = {0, 0, 0};
            for (int a = 0; a < 4; ++a)
                for (int q = 0; q < 3; ++q) c[q] += 0.25 * cv[F[f][a] * 3 + q];
            if (close(c[2], b[4])) continue;  // free surface
            if (close(c[0], b[0]) || close(c[0], b[1]) || close(c[1], b[2]) || close(c[1], b[3]) ||
                close(c[2], b[5]))
                is_pml[e] = 1;
        }
    }
}

void validate(const gfx_box_spec& s, bool general) {
    require(s.nx >= 1 && s.ny >= 1 && (general || s.nz >= 1), "element counts must be >= 1");
    require(general || (s.lx > 0 && s.ly > 0 && s.lz > 0), "box lengths must be > 0");
    require(s.order >= 1 && s.order <= 7, "polynomial order must be in [1, 7]");
    require(s.n_layers >= 1 && s.n_layers <= GFX_MAX_LAYERS, "n_layers must be in [1, 8]");
    for (int l = 0; l < s.n_layers; ++l)
        require(s.vp[l] > 0 && s.vs[l] >= 0 && s.rho[l] > 0, "layer material must be positive");
    require(s.pml_xmin >= 0 && s.pml_xmax >= 0 && s.pml_ymin >= 0 && s.pml_ymax >= 0 &&
                s.pml_zmax >= 0,
            "PML thickness must be >= 0");
    require(s.pml_xmin + s.pml_xmax <= s.nx && s.pml_ymin + s.pml_ymax <= s.ny &&
                (general || s.pml_zmax <= s.nz),
            "PML thicker than the box");
    require(s.cfl_safety > 0 && s.output_dt > 0 && s.duration > 0,
            "cfl_safety, output_dt and duration must be > 0");
    require(general || (s.jitter >= 0 && s.jitter < 0.25), "jitter must be in [0, 0.25)");
    require(s.pml_profile == 0 || s.pml_profile == 1, "pml_profile must be 0 or 1");
    require(s.pml_reflection >= 0.0 && s.pml_reflection < 1.0, "pml_reflection must be in [0, 1)");
    if (!general) {
        const double n_nodes = double(s.nx * s.order + 1) * (s.ny * s.order + 1) * (s.nz * s.order + 1);
        require(n_nodes < 2.0e9, "too many global nodes for int32 ids");
    }
}

// Node numbering of general cells (preprocess/partition.py:192-254): every
// element point's coordinates quantised to 1e-12 of the domain extent, points
// sorted by (x, y, z) of the quantised values, and a new node wherever any of
// them changes.  ibool[p] receives the node; node_xyz the coordinates of the
// node's first point in that order.
int64_t number_by_coordinates(const std::vector<double>& xyz, std::vector<int32_t>& ibool,
                              std::vector<double>& node_xyz) {
    const int64_t P = int64_t(xyz.size() / 3);
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (int64_t p = 0; p < P; ++p)
        for (int c = 0; c < 3; ++c) {
            lo[c] = std::min(lo[c], xyz[size_t(p) * 3 + c]);
            hi[c] = std::max(hi[c], xyz[size_t(p) * 3 + c]);
        }
    double extent = std::numeric_limits<double>::epsilon();
    for (int c = 0; c < 3; ++c) extent = std::max(extent, hi[c] - lo[c]);
    const double scale = 1.0 / (1e-12 * extent);
    std::vector<int64_t> q(size_t(P) * 3);
    for (int64_t p = 0; p < P * 3; ++p) q[size_t(p)] = int64_t(std::nearbyint(xyz[size_t(p)] * scale));
    std::vector<int64_t> order(static_cast<size_t>(P));
    std::iota(order.begin(), order.end(), int64_t(0));
    // lexsort is stable: equal keys keep element-point order
    std::stable_sort(order.begin(), order.end(), [&](int64_t a, int64_t b) {
        for (int c = 0; c < 3; ++c)
            if (q[size_t(a) * 3 + c] != q[size_t(b) * 3 + c]) return q[size_t(a) * 3 + c] < q[size_t(b) * 3 + c];
        return false;
    });
    ibool.assign(size_t(P), 0);
    node_xyz.clear();
    int64_t id = -1;
    for (int64_t r = 0; r < P; ++r) {
        const int64_t p = order[size_t(r)];
        bool fresh = r == 0;
        if (!fresh) {
            const int64_t pp = order[size_t(r - 1)];
            for (int c = 0; c < 3; ++c) fresh = fresh || q[size_t(p) * 3 + c] != q[size_t(pp) * 3 + c];
        }
        if (fresh) {
            ++id;
            node_xyz.insert(node_xyz.end(), &xyz[size_t(p) * 3], &xyz[size_t(p) * 3] + 3);
        }
        ibool[size_t(p)] = int32_t(id);
    }
    return id + 1;
}

void build(MeshHolder& H, const GeneralCells* gen = nullptr) {
    const gfx_box_spec& s = H.spec;
    const bool general = gen != nullptr;
    validate(s, general);
    const GLL gll(s.order);
    const int N = s.order, ng = N + 1, nn = ng * ng * ng;
    const int nx = s.nx, ny = s.ny;
    int64_t E = 0;
    int nz = s.nz;
    // C-PML flags: by cell index when the cells are an x-fastest nx * ny * nz
    // lattice, else from the cells' faces on the domain bounds
    // (preprocess/cli.py:239-264, boundary_detector.py:18-86)
    bool index_flags = true;
    if (general) {
        E = int64_t(gen->cv.size() / 24);
        require(E >= 1, "no cells");
        nz = int(E / (int64_t(nx) * ny));  // cli.py:462: n_cell // (nx * ny)
        index_flags = int64_t(nx) * ny * nz == E;
        if (index_flags) require(s.pml_zmax <= nz, "PML thicker than the box");
        nz = std::max(nz, 1);  // element height for the C-PML widths (cli.py:466)
    } else {
        E = int64_t(nx) * ny * nz;
    }
    require(E * nn < (int64_t(1) << 31), "too many element points for int32 indexing");
    const int64_t P = E * nn;
    const int NX = N * nx + 1, NY = N * ny + 1, NZ = N * nz + 1;
    const double bnd[6] = {general ? gen->bnd[0] : 0.0, general ? gen->bnd[1] : s.lx,
                           general ? gen->bnd[2] : 0.0, general ? gen->bnd[3] : s.ly,
                           general ? gen->bnd[4] : 0.0, general ? gen->bnd[5] : s.lz};
    // element sizes from the bounds (preprocess/cli.py:464-469)
    const double hx = (bnd[1] - bnd[0]) / nx, hy = (bnd[3] - bnd[2]) / ny, hz = (bnd[5] - bnd[4]) / nz;

    gfx_mesh& m = H.m;
    m.n_elem = int(E);
    m.ngll = ng;
    m.nx = nx;
    m.ny = ny;
    m.nz = nz;
    m.xmin = bnd[0];
    m.xmax = bnd[1];
    m.ymin = bnd[2];
    m.ymax = bnd[3];
    m.zmin = bnd[4];
    m.zmax = bnd[5];

    // per element corners; global nodes and connectivity
    std::vector<double> vert;
    std::function<void(int64_t, double (*)[3])> corners;
    int64_t G = 0;
    if (general) {
        corners = [&](int64_t e, double cv[8][3]) {
            for (int a = 0; a < 8; ++a)
                for (int c = 0; c < 3; ++c) cv[a][c] = gen->cv[(size_t(e) * 8 + a) * 3 + c];
        };
        // every element point's coordinates (preprocess/gll_geometry.py:170-175), then the
        // reference's coordinate-sort numbering
        std::vector<double> pxyz(size_t(P) * 3);
        parallel_for(E, [&](int64_t e) {
            double cv[8][3];
            corners(e, cv);
            for (int i = 0; i < ng; ++i)
                for (int j = 0; j < ng; ++j)
                    for (int k = 0; k < ng; ++k) {
                        double S[8], dS[8][3];
                        shape(gll.x[i], gll.x[j], gll.x[k], S, dS);
                        double* X = &pxyz[(size_t(e) * nn + (i * ng + j) * ng + k) * 3];
                        for (int c = 0; c < 3; ++c) {
                            double v = 0;
                            for (int a = 0; a < 8; ++a) v += S[a] * cv[a][c];
                            X[c] = v;
                        }
                    }
        });
        G = number_by_coordinates(pxyz, H.ibool, H.node_xyz);
        require(G < (int64_t(1) << 31) - 8, "too many global nodes for int32 ids");
    } else {
        // vertex lattice, interior vertices optionally jittered
        const int VX = nx + 1, VY = ny + 1, VZ = nz + 1;
        G = int64_t(NX) * NY * NZ;
        vert.resize(size_t(VX) * VY * VZ * 3);
        {
            Rng rng(s.jitter_seed ? s.jitter_seed : 7);
            for (int vz = 0; vz < VZ; ++vz)
                for (int vy = 0; vy < VY; ++vy)
                    for (int vx = 0; vx < VX; ++vx) {
                        double* p = &vert[((size_t(vz) * VY + vy) * VX + vx) * 3];
                        p[0] = (vx == nx) ? s.lx : vx * hx;
                        p[1] = (vy == ny) ? s.ly : vy * hy;
                        p[2] = (vz == nz) ? s.lz : vz * hz;
                        const bool interior =
                            vx > 0 && vx < nx && vy > 0 && vy < ny && vz > 0 && vz < nz;
                        if (s.jitter > 0 && interior) {
                            p[0] += (2 * rng.uniform() - 1) * s.jitter * hx;
                            p[1] += (2 * rng.uniform() - 1) * s.jitter * hy;
                            p[2] += (2 * rng.uniform() - 1) * s.jitter * hz;
                        }
                    }
        }
        auto lattice_corners = [&vert, VX, VY](int ex, int ey, int ez, double cv[8][3]) {
            for (int a = 0; a < 8; ++a) {
                const int vx = ex + (kCorner[a][0] > 0), vy = ey + (kCorner[a][1] > 0),
                          vz = ez + (kCorner[a][2] > 0);
                const double* p = &vert[((size_t(vz) * VY + vy) * VX + vx) * 3];
                cv[a][0] = p[0];
                cv[a][1] = p[1];
                cv[a][2] = p[2];
            }
        };
        corners = [lattice_corners, nx, ny](int64_t e, double cv[8][3]) {
            lattice_corners(int(e % nx), int((e / nx) % ny), int(e / (int64_t(nx) * ny)), cv);
        };

        // global node coordinates from one canonical element per node
        H.node_xyz.resize(G * 3);
        parallel_for(int64_t(NZ) * NY, [&](int64_t row) {
            const int gz = int(row / NY), gy = int(row % NY);
            const int ez = std::min(gz / N, nz - 1), k = gz - ez * N;
            const int ey = std::min(gy / N, ny - 1), j = gy - ey * N;
            for (int gx = 0; gx < NX; ++gx) {
                const int ex = std::min(gx / N, nx - 1), i = gx - ex * N;
                double cv[8][3], S[8], dS[8][3];
                lattice_corners(ex, ey, ez, cv);
                shape(gll.x[i], gll.x[j], gll.x[k], S, dS);
                double* X = &H.node_xyz[((size_t(gz) * NY + gy) * NX + gx) * 3];
                for (int c = 0; c < 3; ++c) {
                    double v = 0;
                    for (int a = 0; a < 8; ++a) v += S[a] * cv[a][c];
                    X[c] = v;
                }
            }
        });

        // connectivity (element x fastest, as preprocess/cli.py:252-254 assumes)
        H.ibool.resize(P);
        parallel_for(E, [&](int64_t e) {
            const int ex = int(e % nx), ey = int((e / nx) % ny), ez = int(e / (int64_t(nx) * ny));
            for (int i = 0; i < ng; ++i)
                for (int j = 0; j < ng; ++j)
                    for (int k = 0; k < ng; ++k) {
                        const int64_t gx = ex * N + i, gy = ey * N + j, gz = ez * N + k;
                        H.ibool[e * nn + (i * ng + j) * ng + k] = int32_t((gz * NY + gy) * NX + gx);
                    }
        });
    }
    m.n_node_glob = int(G);

    // material at global nodes
    std::vector<double> nvp(G), nvs(G), nrho(G);
    {
        std::vector<double> kx, ky, kz, ph;
        const int K = (s.rf_amp != 0.0) ? std::max(1, s.rf_modes) : 0;
        if (K > 0) {
            require(s.rf_corr_len > 0, "rf_corr_len must be > 0");
            Rng rng(s.rf_seed ? s.rf_seed : 2);
            kx.resize(K);
            ky.resize(K);
            kz.resize(K);
            ph.resize(K);
            for (int q = 0; q < K; ++q) {
                kx[q] = rng.normal() / s.rf_corr_len;
                ky[q] = rng.normal() / s.rf_corr_len;
                kz[q] = rng.normal() / s.rf_corr_len;
                ph[q] = 2.0 * M_PI * rng.uniform();
            }
        }
        const double amp_norm = (K > 0) ? std::sqrt(2.0 / K) : 0.0;
        // Regular boxes: x depends on gx only (likewise y, z), so
        // cos(kx x + ky y + kz z + ph) = Re(e^{i kx x} * e^{i(ky y + kz z + ph)}).
        const bool separable = (K > 0) && !general && s.jitter == 0.0;
        std::vector<double> cxr, cxi;
        if (separable) {
            cxr.resize(size_t(K) * NX);
            cxi.resize(size_t(K) * NX);
            for (int q = 0; q < K; ++q)
                for (int gx = 0; gx < NX; ++gx) {
                    const double a = kx[q] * H.node_xyz[size_t(gx) * 3 + 0];
                    cxr[size_t(q) * NX + gx] = std::cos(a);
                    cxi[size_t(q) * NX + gx] = std::sin(a);
                }
        }
        // rows of the node lattice, or runs of 256 nodes of a general mesh
        const int64_t RW = general ? 256 : NX;
        parallel_for((G + RW - 1) / RW, [&](int64_t row) {
            const int64_t g0 = row * RW;
            const int nrow = int(std::min<int64_t>(RW, G - g0));
            std::vector<double> f(nrow, 0.0);
            if (separable) {
                const double y = H.node_xyz[g0 * 3 + 1], z = H.node_xyz[g0 * 3 + 2];
                for (int q = 0; q < K; ++q) {
                    const double b = ky[q] * y + kz[q] * z + ph[q];
                    const double wr = std::cos(b), wi = std::sin(b);
                    const double* xr = &cxr[size_t(q) * NX];
                    const double* xi = &cxi[size_t(q) * NX];
                    for (int gx = 0; gx < NX; ++gx) f[gx] += wr * xr[gx] - wi * xi[gx];
                }
            } else if (K > 0) {
                for (int gx = 0; gx < nrow; ++gx) {
                    const double* X = &H.node_xyz[(g0 + gx) * 3];
                    double acc = 0;
                    for (int q = 0; q < K; ++q)
                        acc += std::cos(kx[q] * X[0] + ky[q] * X[1] + kz[q] * X[2] + ph[q]);
                    f[gx] = acc;
                }
            }
            for (int gx = 0; gx < nrow; ++gx) {
                const int64_t g = g0 + gx;
                const double* X = &H.node_xyz[g * 3];
                int l = 0;  // layer l holds layer_zbot[l-1] <= z < layer_zbot[l]; last is halfspace
                while (l < s.n_layers - 1 && X[2] >= s.layer_zbot[l]) ++l;
                double vp = s.vp[l], vs = s.vs[l];
                if (K > 0) {
                    vp *= 1.0 + s.rf_amp * amp_norm * f[gx];
                    vs *= 1.0 + s.rf_amp * amp_norm * f[gx];
                }
                nvp[g] = vp;
                nvs[g] = vs;
                nrho[g] = s.rho[l];
            }
        });
    }

    // shear attenuation per layer, at the element points' nodes
    m.n_sls = s.n_sls;
    if (s.n_sls > 0) {
        require(s.n_sls <= 8, "n_sls must be in [0, 8]");
        const int L = s.n_sls;
        std::vector<double> lts(size_t(s.n_layers) * L), lte(size_t(s.n_layers) * L);
        for (int l = 0; l < s.n_layers; ++l) sls_fit(s.q_mu[l], s.att_fmin, s.att_fmax, L, &lts[l * L], &lte[l * L]);
        H.tau_sigma.resize(size_t(P) * L);
        H.tau_epsilon.resize(size_t(P) * L);
        parallel_for(E, [&](int64_t e) {
            for (int n = 0; n < nn; ++n) {
                const double z = H.node_xyz[int64_t(H.ibool[e * nn + n]) * 3 + 2];
                int l = 0;
                while (l < s.n_layers - 1 && z >= s.layer_zbot[l]) ++l;
                for (int q = 0; q < L; ++q) {
                    H.tau_sigma[(size_t(e) * nn + n) * L + q] = lts[size_t(l) * L + q];
                    H.tau_epsilon[(size_t(e) * nn + n) * L + q] = lte[size_t(l) * L + q];
                }
            }
        });
    }

    // per element point geometry and material
    H.jacobian.resize(P);
    H.dxi_dx.resize(P * 9);
    H.mass.resize(P);
    H.density.resize(P);
    H.lambda_.resize(P);
    H.mu_.resize(P);
    if (s.want_coords) {
        H.coords.resize(P * 3);
        H.vp.resize(P);
        H.vs.resize(P);
    }
    std::vector<double> hmin_e(E, 1e300), vpmax_e(E, 0.0);
    parallel_for(E, [&](int64_t e) {
        double cv[8][3];
        corners(e, cv);
        double hmin = 1e300, vpm = 0;
        for (int i = 0; i < ng; ++i)
            for (int j = 0; j < ng; ++j)
                for (int k = 0; k < ng; ++k) {
                    const int n = (i * ng + j) * ng + k;
                    const int64_t p = e * nn + n;
                    double S[8], dS[8][3];
                    shape(gll.x[i], gll.x[j], gll.x[k], S, dS);
                    double J[3][3];  // J[r][c] = dx_c / dxi_r (gll_geometry.py:176)
                    for (int r = 0; r < 3; ++r)
                        for (int c = 0; c < 3; ++c) {
                            double v = 0;
                            for (int a = 0; a < 8; ++a) v += dS[a][r] * cv[a][c];
                            J[r][c] = v;
                        }
                    const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                                       J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                                       J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
                    H.jacobian[p] = det;
                    double* dd = &H.dxi_dx[p * 9];
                    if (det > 0) {  // inv(J).ravel(): dd[m*3+q] = dxi_q/dx_m
                        const double id = 1.0 / det;
                        dd[0] = (J[1][1] * J[2][2] - J[1][2] * J[2][1]) * id;
                        dd[1] = (J[0][2] * J[2][1] - J[0][1] * J[2][2]) * id;
                        dd[2] = (J[0][1] * J[1][2] - J[0][2] * J[1][1]) * id;
                        dd[3] = (J[1][2] * J[2][0] - J[1][0] * J[2][2]) * id;
                        dd[4] = (J[0][0] * J[2][2] - J[0][2] * J[2][0]) * id;
                        dd[5] = (J[0][2] * J[1][0] - J[0][0] * J[1][2]) * id;
                        dd[6] = (J[1][0] * J[2][1] - J[1][1] * J[2][0]) * id;
                        dd[7] = (J[0][1] * J[2][0] - J[0][0] * J[2][1]) * id;
                        dd[8] = (J[0][0] * J[1][1] - J[0][1] * J[1][0]) * id;
                    } else {
                        for (int q = 0; q < 9; ++q) dd[q] = 0.0;
                    }
                    const int64_t g = H.ibool[p];
                    const double rho = nrho[g], vp = nvp[g], vs = nvs[g];
                    H.density[p] = rho;
                    H.mu_[p] = rho * vs * vs;
                    H.lambda_[p] = rho * (vp * vp - 2.0 * vs * vs);
                    H.mass[p] = det * gll.w[i] * gll.w[j] * gll.w[k] * rho;
                    if (s.want_coords) {
                        for (int c = 0; c < 3; ++c) H.coords[p * 3 + c] = H.node_xyz[g * 3 + c];
                        H.vp[p] = vp;
                        H.vs[p] = vs;
                    }
                    vpm = std::max(vpm, vp);
                    // CFL spacing to +i/+j/+k neighbours (cfl_validator.py:45-64)
                    const double* X0 = &H.node_xyz[g * 3];
                    const int64_t nb[3] = {i + 1 < ng ? H.ibool[e * nn + ((i + 1) * ng + j) * ng + k] : -1,
                                           j + 1 < ng ? H.ibool[e * nn + (i * ng + j + 1) * ng + k] : -1,
                                           k + 1 < ng ? H.ibool[e * nn + (i * ng + j) * ng + k + 1] : -1};
                    for (int q = 0; q < 3; ++q) {
                        if (nb[q] < 0) continue;
                        const double* X1 = &H.node_xyz[nb[q] * 3];
                        const double dx = X0[0] - X1[0], dy = X0[1] - X1[1], dz = X0[2] - X1[2];
                        hmin = std::min(hmin, std::sqrt(dx * dx + dy * dy + dz * dz));
                    }
                }
        hmin_e[e] = hmin;
        vpmax_e[e] = vpm;
    });
    m.h_min = *std::min_element(hmin_e.begin(), hmin_e.end());
    m.vp_max = *std::max_element(vpmax_e.begin(), vpmax_e.end());
    require(m.h_min > 0 && m.vp_max > 0, "degenerate mesh or material");
    m.cfl_dt = s.cfl_safety * m.h_min / m.vp_max;
    m.dt = 0;
    for (int stride = 1; stride <= 100; ++stride) {  // cfl_validator.py:101-104
        const double dt = s.output_dt / stride;
        if (dt <= m.cfl_dt) {
            m.dt = dt;
            m.snapshot_stride = stride;
            break;
        }
    }
    require(m.dt > 0, "output_dt too large for the CFL limit (no stride in 1..100)");
    m.nsteps = int(std::ceil(s.duration / m.dt));

    // PML flags (cli.py:245-264) and widths (cli.py:464-474)
    std::vector<char> is_pml(E, 0);
    if (index_flags) {
        for (int64_t e = 0; e < E; ++e) {
            const int i = int(e % nx), j = int((e / nx) % ny), k = int(e / (int64_t(nx) * ny));
            if (i < s.pml_xmin || i >= nx - s.pml_xmax || j < s.pml_ymin || j >= ny - s.pml_ymax ||
                k >= nz - s.pml_zmax)
                is_pml[e] = 1;
        }
    } else {
        face_flags(*gen, is_pml);
    }
    const double wid[6] = {s.pml_xmin * hx, s.pml_xmax * hx, s.pml_ymin * hy,
                           s.pml_ymax * hy, 0.0,             s.pml_zmax * hz};
    H.pml_region.assign(E, 0);
    H.coef_alpha.resize(P * 9);
    H.coef_abar.assign(P * 5, 0.0);
    if (s.want_beta) H.coef_beta.resize(P * 9);
    if (s.want_strain) H.coef_strain.assign(P * 18, 0.0);
    if (s.want_profiles) {
        H.pml_K.assign(P * 3, 1.0);
        H.pml_d.assign(P * 3, 0.0);
        H.pml_alpha.assign(P * 3, 0.0);
    }
    H.damping.assign(P, 0.0);
    const double alpha_max[3] = {M_PI * s.f0_pml * 0.9, M_PI * s.f0_pml * 1.0,
                                 M_PI * s.f0_pml * 1.1};
    const double rcoef = s.pml_reflection > 0.0 ? s.pml_reflection : 1e-5;  // R_COEF, pml_cpml.py:23
    const int64_t ncell_approx = std::max<int64_t>(1, (int64_t)std::ceil(std::pow(double(E), 1.0 / 3.0)));
    parallel_for(E, [&](int64_t e) {
        int region = 0;
        if (is_pml[e]) {  // classify_pml_regions, pml_cpml.py:182-248
            double c[3] = {0, 0, 0};
            for (int n = 0; n < nn; ++n) {
                const double* X = &H.node_xyz[int64_t(H.ibool[e * nn + n]) * 3];
                for (int q = 0; q < 3; ++q) c[q] += X[q];
            }
            for (int q = 0; q < 3; ++q) c[q] /= nn;
            const double tol = 1e-6;
            bool act[3] = {false, false, false};
            for (int ax = 0; ax < 3; ++ax) {
                const double wmin = wid[2 * ax], wmax = wid[2 * ax + 1];
                if (wmin > 0 && c[ax] < bnd[2 * ax] + wmin + tol * wmin) act[ax] = true;
                if (wmax > 0 && c[ax] > bnd[2 * ax + 1] - wmax - tol * wmax) act[ax] = true;
            }
            if (act[0] && act[1] && act[2]) region = R_XYZ;
            else if (act[0] && act[1]) region = R_XY;
            else if (act[0] && act[2]) region = R_XZ;
            else if (act[1] && act[2]) region = R_YZ;
            else if (act[0]) region = R_X;
            else if (act[1]) region = R_Y;
            else if (act[2]) region = R_Z;
            else region = R_XYZ;
        }
        H.pml_region[e] = region;
        for (int n = 0; n < nn; ++n) {
            const int64_t p = e * nn + n;
            const double* X = &H.node_xyz[int64_t(H.ibool[p]) * 3];
            const double vp = nvp[H.ibool[p]];
            double K[3] = {1, 1, 1}, d[3] = {0, 0, 0}, al[3] = {0, 0, 0};
            if (is_pml[e] && s.pml_profile == 1) {
                // graded into the layer: depth from the interior interface of each
                // face's own layer, the deeper of an axis's two faces
                for (int ax = 0; ax < 3; ++ax) {
                    if (!axis_active(region, ax)) continue;
                    double depth = -1e300, w = 0.0;
                    for (int side = 0; side < 2; ++side) {
                        const int f = 2 * ax + side;
                        const double wf = wid[f];
                        if (wf <= 0) continue;
                        const double dep = side == 0 ? (bnd[f] + wf - X[ax]) / wf : (X[ax] - (bnd[f] - wf)) / wf;
                        if (dep > depth) {
                            depth = dep;
                            w = wf;
                        }
                    }
                    if (w <= 0.0) continue;  // no layer on this axis
                    depth = std::min(1.0, std::max(0.0, depth));  // points on the interface: depth 0
                    K[ax] = 1.0;  // K_MIN_PML = K_MAX_PML = 1
                    d[ax] = -((2.0 + 1.0) * vp * std::log(rcoef) / (2.0 * w)) * std::pow(depth, 1.2 * 2.0);
                    al[ax] = alpha_max[ax] * (1.0 - depth);
                }
            } else if (is_pml[e]) {  // compute_pml_profiles, pml_cpml.py:69-155 (later faces overwrite)
                for (int f = 0; f < 6; ++f) {
                    const int ax = f / 2;
                    const double w = wid[f];
                    if (w <= 0 || !axis_active(region, ax)) continue;
                    double dist = std::abs(X[ax] - bnd[f]) / w;
                    dist = std::min(1.0, std::max(0.0, dist));
                    double Kv = 1.0 + (1.0 - 1.0) * dist;  // K_MIN_PML=K_MAX_PML=1
                    double dv = -((2.0 + 1.0) * vp * std::log(rcoef) / (2.0 * w)) *
                                std::pow(dist, 1.2 * 2.0);
                    double av = alpha_max[ax] * (1.0 - dist);
                    if (Kv < 1.0 || dv < 0.0) {
                        Kv = 1.0;
                        dv = 0.0;
                    }
                    av = std::max(av, 0.0);
                    K[ax] = Kv;
                    d[ax] = dv;
                    al[ax] = av;
                }
            }
            for (int ax = 0; ax < 3; ++ax) {  // compute_coef_alpha_beta, pml_cpml.py:307-349
                double c3[3];
                conv_coef(al[ax], m.dt, c3);
                for (int q = 0; q < 3; ++q) H.coef_alpha[p * 9 + ax * 3 + q] = c3[q];
                if (s.want_beta) {
                    conv_coef(al[ax] + d[ax] / K[ax], m.dt, c3);
                    for (int q = 0; q < 3; ++q) H.coef_beta[p * 9 + ax * 3 + q] = c3[q];
                }
            }
            if (s.want_profiles)
                for (int ax = 0; ax < 3; ++ax) {
                    H.pml_K[p * 3 + ax] = K[ax];
                    H.pml_d[p * 3 + ax] = d[ax];
                    H.pml_alpha[p * 3 + ax] = al[ax];
                }
            if (region != 0) abar(region, K, d, al, &H.coef_abar[p * 5]);
            if (region != 0 && s.want_strain) strain_coef(region, K, d, al, &H.coef_strain[p * 18]);
        }
        // legacy linear ramp (preprocess/pml.py:55-122)
        if (is_pml[e]) {
            const int thick[6] = {s.pml_xmin, s.pml_xmax, s.pml_ymin, s.pml_ymax, 0, s.pml_zmax};
            for (int f = 0; f < 6; ++f) {
                if (thick[f] <= 0) continue;
                const int ax = f / 2;
                const double extent = bnd[2 * ax + 1] - bnd[2 * ax];
                const double cell = extent / double(ncell_approx);
                const double w = thick[f] * cell;
                if (w <= 0) continue;
                const bool neg = (f % 2) == 0;
                const double bval = bnd[f];
                const double start = neg ? bval + w : bval - w;
                double emin = 1e300, emax = -1e300;
                for (int n = 0; n < nn; ++n) {
                    const double v = H.node_xyz[int64_t(H.ibool[e * nn + n]) * 3 + ax];
                    emin = std::min(emin, v);
                    emax = std::max(emax, v);
                }
                if (neg) {
                    if (emax > start + 1e-12 && emax > bval + w + 1e-12) continue;
                } else {
                    if (emin < start - 1e-12) continue;
                }
                for (int n = 0; n < nn; ++n) {
                    const double v = H.node_xyz[int64_t(H.ibool[e * nn + n]) * 3 + ax];
                    double ramp = neg ? (start - v) / w : (v - start) / w;
                    ramp = std::min(1.0, std::max(0.0, ramp));
                    double& dmp = H.damping[e * nn + n];
                    dmp = std::max(dmp, ramp);
                }
            }
        }
    });
    m.n_pml_elem = 0;
    for (int64_t e = 0; e < E; ++e) m.n_pml_elem += is_pml[e] ? 1 : 0;

    // recording cells (recording_map.py:52-88)
    m.record_depth_actual = 0;
    if (s.record_depth_max > 0) {
        std::set<double> levels;  // every cell's min/max z inside [zmin, zmax]
        for (int64_t e = 0; e < E; ++e) {
            double lo = 1e300, hi = -1e300;
            for (int n = 0; n < nn; ++n) {
                const double zc = H.node_xyz[int64_t(H.ibool[e * nn + n]) * 3 + 2];
                lo = std::min(lo, zc);
                hi = std::max(hi, zc);
            }
            if (m.zmin <= lo && lo <= m.zmax) levels.insert(lo);
            if (m.zmin <= hi && hi <= m.zmax) levels.insert(hi);
        }
        const double target = m.zmin + s.record_depth_max;
        double actual = m.zmax - m.zmin;
        for (double z : levels)
            if (z >= target) {
                actual = z - m.zmin;
                break;
            }
        m.record_depth_actual = actual;
        const double ztarget = m.zmin + actual;
        for (int64_t e = 0; e < E; ++e) {
            if (is_pml[e]) continue;
            double cz = 0;
            for (int n = 0; n < nn; ++n) cz += H.node_xyz[int64_t(H.ibool[e * nn + n]) * 3 + 2];
            cz /= nn;
            if (cz <= ztarget + 1e-12) H.rec_cell.push_back(int32_t(e));
        }
    }

    wire_mesh(H);
}

// Newton solve for natural coordinates of x in a trilinear hex
// (source_locator.py:69-117).  Returns false when not converged or outside.
bool newton_xi(const double x[3], const double cv[8][3], double xi[3]) {
    xi[0] = xi[1] = xi[2] = 0.0;
    bool conv = false;
    for (int it = 0; it < 50; ++it) {
        double S[8], dS[8][3];
        shape(xi[0], xi[1], xi[2], S, dS);
        double xc[3] = {0, 0, 0}, J[3][3];
        for (int c = 0; c < 3; ++c)
            for (int a = 0; a < 8; ++a) xc[c] += S[a] * cv[a][c];
        for (int r = 0; r < 3; ++r)  // J = dN^T @ corners: J[r][c] = dx_c/dxi_r
            for (int c = 0; c < 3; ++c) {
                double v = 0;
                for (int a = 0; a < 8; ++a) v += dS[a][r] * cv[a][c];
                J[r][c] = v;
            }
        const double res[3] = {x[0] - xc[0], x[1] - xc[1], x[2] - xc[2]};
        // solve J dxi = res (numpy solve of the [3,3] J as stored)
        const double det = J[0][0] * (J[1][1] * J[2][2] - J[1][2] * J[2][1]) -
                           J[0][1] * (J[1][0] * J[2][2] - J[1][2] * J[2][0]) +
                           J[0][2] * (J[1][0] * J[2][1] - J[1][1] * J[2][0]);
        if (std::abs(det) < 1e-300) return false;
        double dxi[3];
        for (int q = 0; q < 3; ++q) {
            double M[3][3];
            for (int r = 0; r < 3; ++r)
                for (int c = 0; c < 3; ++c) M[r][c] = (c == q) ? res[r] : J[r][c];
            dxi[q] = (M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) -
                      M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
                      M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0])) /
                     det;
        }
        for (int q = 0; q < 3; ++q) xi[q] += dxi[q];
        if (std::sqrt(dxi[0] * dxi[0] + dxi[1] * dxi[1] + dxi[2] * dxi[2]) < 1e-12) {
            conv = true;
            break;
        }
    }
    if (!conv) return false;
    for (int q = 0; q < 3; ++q)
        if (xi[q] < -(1.0 + 1e-8) || xi[q] > 1.0 + 1e-8) return false;
    return true;
}

}  // namespace
}  // namespace gfx

using namespace gfx;

extern "C" gfx_status gfx_box_spec_default(gfx_box_spec* s) {
    return guarded([&] {
        require(s != nullptr, "null spec");
        std::memset(s, 0, sizeof(*s));
        // homogeneous elastic halfspace box, 4000 hexes of 1 km, order 4
        s->nx = 20;
        s->ny = 20;
        s->nz = 10;
        s->lx = 20000.0;
        s->ly = 20000.0;
        s->lz = 10000.0;
        s->order = 4;
        s->n_layers = 1;
        s->vp[0] = 6000.0;
        s->vs[0] = 3464.0;
        s->rho[0] = 2700.0;
        s->rf_modes = 256;
        s->rf_corr_len = 2000.0;
        s->pml_xmin = s->pml_xmax = s->pml_ymin = s->pml_ymax = s->pml_zmax = 3;
        s->f0_pml = 1.0;
        s->cfl_safety = 0.5;
        s->output_dt = 0.05;
        s->duration = 10.0;
        s->record_depth_max = 2000.0;
        s->want_coords = 1;
        s->want_beta = 0;
        s->want_strain = 0;
    });
}

extern "C" gfx_status gfx_mesh_build_box(const gfx_box_spec* spec, gfx_mesh** out) {
    return guarded([&] {
        require(spec != nullptr && out != nullptr, "null argument");
        *out = nullptr;
        auto H = std::make_unique<MeshHolder>();
        H->spec = *spec;
        build(*H);
        *out = &H.release()->m;
    });
}

namespace {
// A mesh over a topology: corners of every cell in GMSH order
// (preprocess/gll_geometry.py:70-129), bounds from the vertices
// (preprocess/cli.py:409-417), then the same pipeline as a box.
gfx_mesh* build_from_topology(const topo::Topology& T, const gfx_box_spec& spec) {
    GeneralCells gc;
    gc.cv.resize(size_t(T.n_cell) * 24);
    for (int64_t c = 0; c < T.n_cell; ++c) {
        const auto v = topo::cell_vertices(T, c);
        for (int a = 0; a < 8; ++a) {
            require(v[a] >= 1 && v[a] <= T.n_vertex, "vertex id out of range");
            for (int q = 0; q < 3; ++q) gc.cv[(size_t(c) * 8 + a) * 3 + q] = T.vertex_to_coord[size_t(v[a] - 1) * 3 + q];
        }
    }
    for (int q = 0; q < 3; ++q) {
        gc.bnd[2 * q] = 1e300;
        gc.bnd[2 * q + 1] = -1e300;
    }
    for (int64_t v = 0; v < T.n_vertex; ++v)
        for (int q = 0; q < 3; ++q) {
            gc.bnd[2 * q] = std::min(gc.bnd[2 * q], T.vertex_to_coord[size_t(v) * 3 + q]);
            gc.bnd[2 * q + 1] = std::max(gc.bnd[2 * q + 1], T.vertex_to_coord[size_t(v) * 3 + q]);
        }
    auto H = std::make_unique<MeshHolder>();
    H->spec = spec;
    build(*H, &gc);
    return &H.release()->m;
}
}  // namespace

extern "C" gfx_status gfx_gmsh_to_topology(const char* msh_path, const char* model_path, const char* aux_path) {
    return guarded([&] {
        require(msh_path && model_path, "null argument");
        const topo::Topology T = topo::extract(topo::read_gmsh(msh_path));
        topo::write_topology(model_path, T);
        if (aux_path && aux_path[0]) topo::write_auxiliary(aux_path, T);
    });
}

extern "C" gfx_status gfx_mesh_build_topology(const char* model_path, const gfx_box_spec* spec, gfx_mesh** out) {
    return guarded([&] {
        require(model_path && spec && out, "null argument");
        *out = nullptr;
        *out = build_from_topology(topo::read_topology(model_path), *spec);
    });
}

extern "C" gfx_status gfx_mesh_build_hex(int64_t n_point, const double* xyz, int64_t n_cell, const int64_t* hex,
                                         const gfx_box_spec* spec, gfx_mesh** out) {
    return guarded([&] {
        require(xyz && hex && spec && out && n_point > 0 && n_cell > 0, "null or empty argument");
        *out = nullptr;
        topo::HexMesh M;
        M.xyz.assign(xyz, xyz + n_point * 3);
        M.hex.assign(hex, hex + n_cell * 8);
        *out = build_from_topology(topo::extract(M), *spec);
    });
}

extern "C" gfx_status gfx_topology_summary(const char* model_path, int64_t* counts, int64_t* cell_corners,
                                           int64_t n_cell_corners) {
    return guarded([&] {
        require(model_path && counts, "null argument");
        const topo::Topology T = topo::read_topology(model_path);
        counts[0] = T.n_vertex;
        counts[1] = T.n_edge;
        counts[2] = T.n_surface;
        counts[3] = T.n_cell;
        if (cell_corners) {
            require(n_cell_corners >= T.n_cell, "cell_corners holds fewer cells than the topology");
            for (int64_t c = 0; c < T.n_cell; ++c) {
                const auto v = topo::cell_vertices(T, c);
                for (int a = 0; a < 8; ++a) cell_corners[c * 8 + a] = v[a];
            }
        }
    });
}

extern "C" gfx_status gfx_mesh_free(gfx_mesh* m) {
    return guarded([&] {
        if (!m) return;
        // gfx_mesh is the first member of MeshHolder
        delete reinterpret_cast<MeshHolder*>(m);
    });
}

extern "C" gfx_status gfx_mesh_locate_source(const gfx_mesh* m, double x, double y, double z,
                                             int max_cells, int32_t* cells, double* weights,
                                             int* n_src) {
    return guarded([&] {
        require(m && cells && weights && n_src, "null argument");
        const GLL gll(m->ngll - 1);
        const int ng = m->ngll, nn = ng * ng * ng;
        const double xs[3] = {x, y, z};
        // surface mode when z is the free surface (source_locator.py:273-275)
        const bool buried = std::abs(z - m->zmin) > 1e-8 * 1.0;
        const int idx = ng - 1;
        const int corner[8] = {0, idx, idx * ng, idx * ng + idx, idx * ng * ng, idx * ng * ng + idx,
                               (idx * ng + idx) * ng, (idx * ng + idx) * ng + idx};
        const double ztol = 1e-6 * std::max(1.0, std::abs(m->zmax - m->zmin));
        std::vector<int64_t> cand;
        for (int64_t e = 0; e < m->n_elem; ++e) {
            if (!buried) {
                // cells with a face on z = zmin: four corners on the surface
                int on = 0;
                for (int a = 0; a < 8; ++a)
                    on += std::abs(m->node_xyz[int64_t(m->ibool[e * nn + corner[a]]) * 3 + 2] - m->zmin) <= ztol;
                if (on >= 4) cand.push_back(e);
            } else {
                if (m->pml_region && m->pml_region[e] != 0) continue;
                double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
                for (int n = 0; n < nn; ++n) {
                    const double* X = &m->node_xyz[int64_t(m->ibool[e * nn + n]) * 3];
                    for (int q = 0; q < 3; ++q) {
                        lo[q] = std::min(lo[q], X[q]);
                        hi[q] = std::max(hi[q], X[q]);
                    }
                }
                bool in = true;
                for (int q = 0; q < 3; ++q)
                    in = in && (lo[q] - 1e-10 <= xs[q]) && (xs[q] <= hi[q] + 1e-10);
                if (in) cand.push_back(e);
            }
        }
        std::vector<int32_t> found;
        std::vector<double> w;
        double total = 0;
        for (int64_t e : cand) {
            double cv[8][3];
            const int cidx[8][3] = {{0, 0, 0}, {idx, 0, 0}, {idx, idx, 0}, {0, idx, 0},
                                    {0, 0, idx}, {idx, 0, idx}, {idx, idx, idx}, {0, idx, idx}};
            for (int a = 0; a < 8; ++a) {
                const int n = (cidx[a][0] * ng + cidx[a][1]) * ng + cidx[a][2];
                const double* X = &m->node_xyz[int64_t(m->ibool[e * nn + n]) * 3];
                for (int q = 0; q < 3; ++q) cv[a][q] = X[q];
            }
            double xi[3];
            if (!newton_xi(xs, cv, xi)) continue;
            double lx[16], ly[16], lz[16];
            gll.lagrange(xi[0], lx);
            gll.lagrange(xi[1], ly);
            gll.lagrange(xi[2], lz);
            double sum = 0;
            std::vector<double> wc(nn);
            for (int i = 0; i < ng; ++i)
                for (int j = 0; j < ng; ++j)
                    for (int k = 0; k < ng; ++k) {
                        const double v = lx[i] * ly[j] * lz[k];
                        wc[(i * ng + j) * ng + k] = v;
                        sum += v;
                    }
            if (sum < 1e-12) continue;
            found.push_back(int32_t(e));
            w.insert(w.end(), wc.begin(), wc.end());
            total += sum;
        }
        if (found.empty()) throw Error(GFX_ERR_INVALID, "source not contained in any cell");
        require(int(found.size()) <= max_cells, "more source cells than max_cells");
        for (size_t c = 0; c < found.size(); ++c) {
            cells[c] = found[c];
            for (int n = 0; n < nn; ++n) weights[c * nn + n] = w[c * nn + n] / total;
        }
        *n_src = int(found.size());
    });
}

extern "C" gfx_status gfx_pml_strain_coefficients(int region, const double* K, const double* d,
                                                   const double* alpha, double* out18) {
    return guarded([&] {
        require(K && d && alpha && out18, "K, d, alpha and out18 are required");
        require(region >= 0 && region <= R_XYZ, "region must be 0..7");
        strain_coef(region, K, d, alpha, out18);
    });
}

extern "C" gfx_status gfx_sls_fit(double q, double f_min, double f_max, int n_sls, double* tau_sigma,
                                   double* tau_epsilon) {
    return guarded([&] {
        require(tau_sigma && tau_epsilon, "null argument");
        sls_fit(q, f_min, f_max, n_sls, tau_sigma, tau_epsilon);
    });
}

extern "C" gfx_status gfx_ricker(double f0, double t0, double amplitude, double dt, int n,
                                 double* out) {
    return guarded([&] {
        require(out != nullptr && n >= 0, "bad argument");
        for (int i = 0; i < n; ++i) {
            const double t = i * dt;
            const double a = M_PI * f0 * (t - t0);
            out[i] = amplitude * (1.0 - 2.0 * a * a) * std::exp(-(a * a));
        }
    });
}

namespace gfx {

void wire_mesh(MeshHolder& H) {
    gfx_mesh& m = H.m;
    m.coords = H.coords.empty() ? nullptr : H.coords.data();
    m.jacobian = H.jacobian.data();
    m.dxi_dx = H.dxi_dx.data();
    m.mass = H.mass.data();
    m.vp = H.vp.empty() ? nullptr : H.vp.data();
    m.vs = H.vs.empty() ? nullptr : H.vs.data();
    m.density = H.density.data();
    m.lambda_ = H.lambda_.data();
    m.mu_ = H.mu_.data();
    m.damping = H.damping.data();
    m.ibool = H.ibool.data();
    m.node_xyz = H.node_xyz.data();
    m.pml_region = H.pml_region.empty() ? nullptr : H.pml_region.data();
    m.coef_alpha = H.coef_alpha.empty() ? nullptr : H.coef_alpha.data();
    m.coef_beta = H.coef_beta.empty() ? nullptr : H.coef_beta.data();
    m.coef_strain = H.coef_strain.empty() ? nullptr : H.coef_strain.data();
    m.coef_abar = H.coef_abar.empty() ? nullptr : H.coef_abar.data();
    m.n_rec_cell = int(H.rec_cell.size());
    m.rec_cell = H.rec_cell.empty() ? nullptr : H.rec_cell.data();
    m.tau_sigma = H.tau_sigma.empty() ? nullptr : H.tau_sigma.data();
    m.tau_epsilon = H.tau_epsilon.empty() ? nullptr : H.tau_epsilon.data();
    m.pml_K = H.pml_K.empty() ? nullptr : H.pml_K.data();
    m.pml_d = H.pml_d.empty() ? nullptr : H.pml_d.data();
    m.pml_alpha = H.pml_alpha.empty() ? nullptr : H.pml_alpha.data();
}

}  // namespace gfx
