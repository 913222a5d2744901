// topology.cpp -- general hexahedral meshes: GMSH files in, the reference's
// /topology datasets out and back, and each cell's corners in GMSH order.
//
//   GMSH reader           the hexahedron cells and points meshio.read keeps
//                         (tools/gmsh_to_hdf5.py:281-283), ASCII formats 4.1 and 2.2
//   topology extraction   tools/gmsh_to_hdf5.py:45-146 (edges keyed by their sorted
//                         vertex pair, numbered by first appearance; faces keyed by
//                         their edge set, signed by orientation against the first
//                         appearance)
//   /topology writer      tools/gmsh_to_hdf5.py:149-170
//   auxiliary CSR         tools/gmsh_to_hdf5.py:173-268
//   /topology reader      preprocess/topology_reader.py:22-47
//   cell corners          preprocess/gll_geometry.py:70-129 (each corner is the vertex
//                         on exactly its three faces)
#include "topology.hpp"

#include <algorithm>
#include <fstream>
#include <map>
#include <set>
#include <sstream>
#include <unordered_map>

#include "h5.hpp"

namespace gfx::topo {
namespace {

// Hex faces by local vertex (CCW seen from outside), in the order -z, +z, -y, +y, -x, +x.
constexpr int kFace[6][4] = {{0, 3, 2, 1}, {4, 5, 6, 7}, {0, 1, 5, 4}, {3, 7, 6, 2}, {0, 4, 7, 3}, {1, 2, 6, 5}};
// The 12 hex edges by local vertex.
constexpr int kEdge[12][2] = {{0, 1}, {1, 2}, {2, 3}, {3, 0}, {4, 5}, {5, 6},
                              {6, 7}, {7, 4}, {0, 4}, {1, 5}, {2, 6}, {3, 7}};

struct Reader {
    std::ifstream in;
    std::string path;
    int64_t line_no = 0;
    explicit Reader(const std::string& p) : in(p), path(p) {
        if (!in) throw Error(GFX_ERR_IO, "cannot open " + p);
    }
    bool line(std::string& s) {
        if (!std::getline(in, s)) return false;
        ++line_no;
        if (!s.empty() && s.back() == '\r') s.pop_back();
        return true;
    }
    std::string need() {
        std::string s;
        if (!line(s)) fail("unexpected end of file");
        return s;
    }
    [[noreturn]] void fail(const std::string& why) const {
        throw Error(GFX_ERR_IO, path + ":" + std::to_string(line_no) + ": " + why);
    }
    void skip_to(const std::string& end) {
        std::string s;
        while (line(s))
            if (s == end) return;
        fail("missing " + end);
    }
};

template <typename T>
std::vector<T> numbers(const std::string& s) {
    std::istringstream is(s);
    std::vector<T> v;
    T x;
    while (is >> x) v.push_back(x);
    return v;
}

// Points in file order, plus the tag -> point index map (meshio's order).
struct Points {
    std::vector<double> xyz;
    std::unordered_map<int64_t, int64_t> at;
    void add(int64_t tag, double x, double y, double z, Reader& r) {
        if (!at.emplace(tag, int64_t(xyz.size() / 3)).second) r.fail("duplicate node tag " + std::to_string(tag));
        xyz.insert(xyz.end(), {x, y, z});
    }
    int64_t index(int64_t tag, Reader& r) const {
        auto it = at.find(tag);
        if (it == at.end()) r.fail("element refers to unknown node tag " + std::to_string(tag));
        return it->second;
    }
};

void read_nodes_41(Reader& r, Points& P) {
    const auto head = numbers<int64_t>(r.need());
    if (head.size() < 2) r.fail("bad $Nodes header");
    const int64_t blocks = head[0];
    for (int64_t b = 0; b < blocks; ++b) {
        const auto bh = numbers<int64_t>(r.need());
        if (bh.size() < 4) r.fail("bad node block header");
        const int64_t n = bh[3];
        std::vector<int64_t> tags(n);
        for (int64_t i = 0; i < n; ++i) {
            const auto t = numbers<int64_t>(r.need());
            if (t.empty()) r.fail("bad node tag");
            tags[i] = t[0];
        }
        for (int64_t i = 0; i < n; ++i) {
            const auto c = numbers<double>(r.need());  // x y z [parametric coordinates]
            if (c.size() < 3) r.fail("bad node coordinates");
            P.add(tags[i], c[0], c[1], c[2], r);
        }
    }
}

void read_elements_41(Reader& r, const Points& P, std::vector<int64_t>& hex) {
    const auto head = numbers<int64_t>(r.need());
    if (head.size() < 2) r.fail("bad $Elements header");
    const int64_t blocks = head[0];
    for (int64_t b = 0; b < blocks; ++b) {
        const auto bh = numbers<int64_t>(r.need());
        if (bh.size() < 4) r.fail("bad element block header");
        const int64_t type = bh[2], n = bh[3];
        for (int64_t i = 0; i < n; ++i) {
            const std::string s = r.need();
            if (type != 5) continue;  // 8-node hexahedron
            const auto v = numbers<int64_t>(s);
            if (v.size() < 9) r.fail("hexahedron with fewer than 8 nodes");
            for (int a = 0; a < 8; ++a) hex.push_back(P.index(v[1 + a], r));
        }
    }
}

void read_nodes_22(Reader& r, Points& P) {
    const auto head = numbers<int64_t>(r.need());
    if (head.empty()) r.fail("bad $Nodes header");
    for (int64_t i = 0; i < head[0]; ++i) {
        const auto v = numbers<double>(r.need());
        if (v.size() < 4) r.fail("bad node line");
        P.add(int64_t(v[0]), v[1], v[2], v[3], r);
    }
}

void read_elements_22(Reader& r, const Points& P, std::vector<int64_t>& hex) {
    const auto head = numbers<int64_t>(r.need());
    if (head.empty()) r.fail("bad $Elements header");
    for (int64_t i = 0; i < head[0]; ++i) {
        const auto v = numbers<int64_t>(r.need());  // tag type ntags tags... nodes...
        if (v.size() < 3) r.fail("bad element line");
        if (v[1] != 5) continue;
        const size_t first = 3 + size_t(v[2]);
        if (v.size() < first + 8) r.fail("hexahedron with fewer than 8 nodes");
        for (int a = 0; a < 8; ++a) hex.push_back(P.index(v[first + a], r));
    }
}

// Same cyclic orientation of two signed edge loops (tools/gmsh_to_hdf5.py:30-41).
bool same_orientation(const int64_t* a, const int64_t* b) {
    for (int off = 0; off < 4; ++off)
        if (std::llabs(b[off]) == std::llabs(a[0])) {
            for (int i = 0; i < 4; ++i)
                if (a[i] != b[(off + i) % 4]) return false;
            return true;
        }
    return false;
}

}  // namespace

HexMesh read_gmsh(const std::string& path) {
    Reader r(path);
    std::string s;
    double version = 0;
    Points P;
    HexMesh M;
    bool nodes = false, elems = false;
    while (r.line(s)) {
        if (s == "$MeshFormat") {
            const auto v = numbers<double>(r.need());
            if (v.size() < 2) r.fail("bad $MeshFormat");
            version = v[0];
            if (v[1] != 0) r.fail("binary GMSH files are not supported (write ASCII)");
            if (!(version >= 4.0 && version < 5.0) && !(version >= 2.0 && version < 3.0))
                r.fail("GMSH format " + std::to_string(version) + " is not supported (4.1 or 2.2)");
            r.skip_to("$EndMeshFormat");
        } else if (s == "$Nodes") {
            if (version == 0) r.fail("$Nodes before $MeshFormat");
            if (version >= 4.0)
                read_nodes_41(r, P);
            else
                read_nodes_22(r, P);
            r.skip_to("$EndNodes");
            nodes = true;
        } else if (s == "$Elements") {
            if (!nodes) r.fail("$Elements before $Nodes");
            if (version >= 4.0)
                read_elements_41(r, P, M.hex);
            else
                read_elements_22(r, P, M.hex);
            r.skip_to("$EndElements");
            elems = true;
        } else if (!s.empty() && s[0] == '$' && s.rfind("$End", 0) != 0) {
            r.skip_to("$End" + s.substr(1));
        }
    }
    if (!nodes || !elems) r.fail("no $Nodes or $Elements section");
    if (M.hex.empty()) throw Error(GFX_ERR_INVALID, "No hexahedron cells found in mesh");
    M.xyz = std::move(P.xyz);
    return M;
}

Topology extract(const HexMesh& M) {
    const int64_t nc = int64_t(M.hex.size() / 8);
    require(nc > 0, "No hexahedron cells found in mesh");
    Topology T;
    T.n_vertex = int64_t(M.xyz.size() / 3);
    T.vertex_to_coord = M.xyz;
    for (int64_t v : M.hex) require(v >= 0 && v < T.n_vertex, "cell vertex out of range");
    // edges: (low, high) -> 1-based id in order of first appearance
    std::map<std::pair<int64_t, int64_t>, int64_t> edge_id;
    std::vector<int64_t> cell_edges(size_t(nc) * 12);
    for (int64_t c = 0; c < nc; ++c) {
        const int64_t* v = &M.hex[size_t(c) * 8];
        for (int k = 0; k < 12; ++k) {
            int64_t a = v[kEdge[k][0]], b = v[kEdge[k][1]];
            if (a > b) std::swap(a, b);
            auto it = edge_id.emplace(std::make_pair(a, b), int64_t(T.edge_to_vertex.size() / 2) + 1);
            if (it.second) T.edge_to_vertex.insert(T.edge_to_vertex.end(), {a + 1, b + 1});
            cell_edges[size_t(c) * 12 + k] = it.first->second;
        }
    }
    T.n_edge = int64_t(T.edge_to_vertex.size() / 2);
    // faces: unsigned edge set -> 1-based id; the first appearance is canonical
    std::map<std::array<int64_t, 4>, int64_t> face_id;
    T.cell_to_surface.resize(size_t(nc) * 6);
    for (int64_t c = 0; c < nc; ++c) {
        const int64_t* v = &M.hex[size_t(c) * 8];
        for (int f = 0; f < 6; ++f) {
            int64_t se[4];
            std::array<int64_t, 4> key;
            for (int k = 0; k < 4; ++k) {
                const int64_t a = v[kFace[f][k]], b = v[kFace[f][(k + 1) % 4]];
                const int64_t id = edge_id.at(a < b ? std::make_pair(a, b) : std::make_pair(b, a));
                se[k] = a < b ? id : -id;
                key[k] = id;
            }
            std::sort(key.begin(), key.end());
            auto it = face_id.emplace(key, int64_t(T.surface_to_edge.size() / 4) + 1);
            if (it.second) T.surface_to_edge.insert(T.surface_to_edge.end(), se, se + 4);
            const int64_t sid = it.first->second;
            T.cell_to_surface[size_t(c) * 6 + f] =
                same_orientation(se, &T.surface_to_edge[size_t(sid - 1) * 4]) ? sid : -sid;
        }
    }
    T.n_surface = int64_t(T.surface_to_edge.size() / 4);
    T.n_cell = nc;
    return T;
}

void write_topology(const std::string& path, const Topology& T) {
    using h5::hsize_t;
    h5::Obj f = h5::create_trunc(path);
    h5::Obj g = h5::require_group(f.id(), "/topology");
    h5::write_vec(f.id(), "/topology/vertex_to_coord", T.vertex_to_coord, {hsize_t(T.n_vertex), 3});
    h5::write_vec(f.id(), "/topology/edge_to_vertex", T.edge_to_vertex, {hsize_t(T.n_edge), 2});
    h5::write_vec(f.id(), "/topology/surface_to_edge", T.surface_to_edge, {hsize_t(T.n_surface), 4});
    h5::write_vec(f.id(), "/topology/cell_to_surface", T.cell_to_surface, {hsize_t(T.n_cell), 6});
    const std::pair<const char*, int64_t> counts[4] = {
        {"n_vertex", T.n_vertex}, {"n_edge", T.n_edge}, {"n_surface", T.n_surface}, {"n_cell", T.n_cell}};
    for (const auto& kv : counts) h5::write_attr(g.id(), kv.first, h5::T::I64, &kv.second);
}

namespace {
// CSR of (row, 1-based column) pairs, columns sorted and unique per row
// (tools/gmsh_to_hdf5.py:173-187).
void write_csr(h5::hid_t f, const std::string& name, const std::vector<std::pair<int64_t, int64_t>>& pairs,
               int64_t n_rows) {
    std::vector<std::set<int64_t>> rows(static_cast<size_t>(n_rows));
    for (const auto& p : pairs) rows[size_t(p.first)].insert(p.second);
    std::vector<int64_t> indptr(1, 0), indices;
    for (const auto& r : rows) {
        indices.insert(indices.end(), r.begin(), r.end());
        indptr.push_back(int64_t(indices.size()));
    }
    h5::write_vec(f, "/auxiliary/" + name + "/indptr", indptr);
    h5::write_vec(f, "/auxiliary/" + name + "/indices", indices);
}
}  // namespace

void write_auxiliary(const std::string& path, const Topology& T) {
    using h5::hsize_t;
    const int64_t nv = T.n_vertex, ne = T.n_edge, ns = T.n_surface, nc = T.n_cell;
    std::vector<int64_t> s2c(size_t(ns) * 2, 0);
    for (int64_t c = 0; c < nc; ++c)
        for (int f = 0; f < 6; ++f) {
            const int64_t sid = std::llabs(T.cell_to_surface[size_t(c) * 6 + f]);
            int64_t* row = &s2c[size_t(sid - 1) * 2];
            (row[0] == 0 ? row[0] : row[1]) = c + 1;
        }
    std::vector<std::pair<int64_t, int64_t>> ve, vs, vc, es, ec;
    for (int64_t e = 0; e < ne; ++e) {
        ve.push_back({T.edge_to_vertex[size_t(e) * 2] - 1, e + 1});
        ve.push_back({T.edge_to_vertex[size_t(e) * 2 + 1] - 1, e + 1});
    }
    for (int64_t s = 0; s < ns; ++s)
        for (int k = 0; k < 4; ++k) {
            const int64_t e = std::llabs(T.surface_to_edge[size_t(s) * 4 + k]) - 1;
            vs.push_back({T.edge_to_vertex[size_t(e) * 2] - 1, s + 1});
            vs.push_back({T.edge_to_vertex[size_t(e) * 2 + 1] - 1, s + 1});
            es.push_back({e, s + 1});
        }
    for (int64_t c = 0; c < nc; ++c)
        for (int f = 0; f < 6; ++f) {
            const int64_t s = std::llabs(T.cell_to_surface[size_t(c) * 6 + f]) - 1;
            for (int k = 0; k < 4; ++k) {
                const int64_t e = std::llabs(T.surface_to_edge[size_t(s) * 4 + k]) - 1;
                vc.push_back({T.edge_to_vertex[size_t(e) * 2] - 1, c + 1});
                vc.push_back({T.edge_to_vertex[size_t(e) * 2 + 1] - 1, c + 1});
                ec.push_back({e, c + 1});
            }
        }
    h5::Obj f = h5::create_trunc(path);
    h5::write_vec(f.id(), "/auxiliary/surface_to_cell", s2c, {hsize_t(ns), 2});
    write_csr(f.id(), "vertex_to_edge", ve, nv);
    write_csr(f.id(), "vertex_to_surface", vs, nv);
    write_csr(f.id(), "vertex_to_cell", vc, nv);
    write_csr(f.id(), "edge_to_surface", es, ne);
    write_csr(f.id(), "edge_to_cell", ec, ne);
}

Topology read_topology(const std::string& path) {
    h5::Obj f = h5::open_read(path);
    if (!h5::exists(f.id(), "/topology")) throw Error(GFX_ERR_IO, "File '" + path + "' missing /topology/ group");
    h5::Obj g = h5::open_group(f.id(), "/topology");
    Topology T;
    std::vector<h5::hsize_t> d;
    T.vertex_to_coord = h5::read<double>(f.id(), "/topology/vertex_to_coord", &d);
    require(d.size() == 2 && d[1] == 3, "/topology/vertex_to_coord must be [n, 3]");
    T.edge_to_vertex = h5::read<int64_t>(f.id(), "/topology/edge_to_vertex", &d);
    require(d.size() == 2 && d[1] == 2, "/topology/edge_to_vertex must be [n, 2]");
    T.surface_to_edge = h5::read<int64_t>(f.id(), "/topology/surface_to_edge", &d);
    require(d.size() == 2 && d[1] == 4, "/topology/surface_to_edge must be [n, 4]");
    T.cell_to_surface = h5::read<int64_t>(f.id(), "/topology/cell_to_surface", &d);
    require(d.size() == 2 && d[1] == 6, "/topology/cell_to_surface must be [n, 6]");
    T.n_vertex = int64_t(T.vertex_to_coord.size() / 3);
    T.n_edge = int64_t(T.edge_to_vertex.size() / 2);
    T.n_surface = int64_t(T.surface_to_edge.size() / 4);
    T.n_cell = int64_t(T.cell_to_surface.size() / 6);
    const char* names[4] = {"n_vertex", "n_edge", "n_surface", "n_cell"};
    const int64_t have[4] = {T.n_vertex, T.n_edge, T.n_surface, T.n_cell};
    for (int k = 0; k < 4; ++k) {
        double v = -1;
        if (h5::read_attr_double(g.id(), names[k], v))
            require(int64_t(v) == have[k], std::string("/topology attribute ") + names[k] + " differs from its dataset");
    }
    return T;
}

std::array<int64_t, 8> cell_vertices(const Topology& T, int64_t c) {
    require(c >= 0 && c < T.n_cell, "cell out of range");
    // vertices of each face, then each vertex's set of faces (as a bit mask)
    std::map<int64_t, unsigned> faces_of;
    for (int f = 0; f < 6; ++f) {
        const int64_t s = std::llabs(T.cell_to_surface[size_t(c) * 6 + f]) - 1;
        require(s >= 0 && s < T.n_surface, "cell_to_surface entry out of range");
        for (int k = 0; k < 4; ++k) {
            const int64_t e = std::llabs(T.surface_to_edge[size_t(s) * 4 + k]) - 1;
            require(e >= 0 && e < T.n_edge, "surface_to_edge entry out of range");
            for (int q = 0; q < 2; ++q) faces_of[T.edge_to_vertex[size_t(e) * 2 + q]] |= 1u << f;
        }
    }
    // GMSH corner -> its faces (0 -z, 1 +z, 2 -y, 3 +y, 4 -x, 5 +x)
    static const unsigned want[8] = {
        (1u << 0) | (1u << 2) | (1u << 4), (1u << 0) | (1u << 2) | (1u << 5), (1u << 0) | (1u << 3) | (1u << 5),
        (1u << 0) | (1u << 3) | (1u << 4), (1u << 1) | (1u << 2) | (1u << 4), (1u << 1) | (1u << 2) | (1u << 5),
        (1u << 1) | (1u << 3) | (1u << 5), (1u << 1) | (1u << 3) | (1u << 4)};
    std::array<int64_t, 8> out{};
    for (int a = 0; a < 8; ++a) {
        bool found = false;
        for (const auto& kv : faces_of)
            if (kv.second == want[a]) {
                out[a] = kv.first;
                found = true;
                break;
            }
        if (!found)
            throw Error(GFX_ERR_INVALID,
                        "Cannot identify GMSH corner " + std::to_string(a) + " for cell " + std::to_string(c));
    }
    return out;
}

}  // namespace gfx::topo
