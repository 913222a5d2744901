"""General hexahedral meshes: GMSH reader, /topology datasets, cell corners,
coordinate-sort node numbering, and the preprocessing pipeline on them.

Pinned by the assertions of the reference's tests/tools/test_gmsh_to_hdf5.py
and tests/preprocess/test_gll_geometry.py (restated on meshes built here),
by an independent Gauss-Legendre volume of jittered hexahedra, and by
equivalence with the structured-box generator: the same box given as general
cells must produce the same geometry, material, C-PML coefficients, time step
and recording cells, with a node numbering that is the same partition of the
element points into nodes, ordered by (x, y, z)."""
import os
import subprocess

import numpy as np
import pytest

import pygfx

pytestmark = pytest.mark.skipif(not pygfx.h5_available(), reason="no HDF5 library to bind")

# GMSH corner order: v0(-,-,-) v1(+,-,-) v2(+,+,-) v3(-,+,-) v4..v7 the same at +z
CORNER = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]


def write_msh41(path, xyz, hexes, tags=None, split=1):
    """ASCII GMSH 4.1 with the points in `split` node blocks and the hexahedra
    in one element block, plus a line element block that the reader skips."""
    n = len(xyz)
    tags = np.arange(1, n + 1) if tags is None else np.asarray(tags)
    blocks = np.array_split(np.arange(n), split)
    lines = ["$MeshFormat", "4.1 0 8", "$EndMeshFormat", "$Nodes",
             f"{len(blocks)} {n} {tags.min()} {tags.max()}"]
    for b in blocks:
        lines.append(f"3 1 0 {len(b)}")
        lines += [str(int(tags[i])) for i in b]
        lines += [" ".join(repr(float(v)) for v in xyz[i]) for i in b]
    lines += ["$EndNodes", "$Elements", f"2 {len(hexes) + 1} 1 {len(hexes) + 1}", "1 1 1 1",
              f"{len(hexes) + 1} {int(tags[0])} {int(tags[1])}", f"3 1 5 {len(hexes)}"]
    for c, h in enumerate(hexes):
        lines.append(" ".join([str(c + 1)] + [str(int(tags[v])) for v in h]))
    lines += ["$EndElements"]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def write_msh22(path, xyz, hexes):
    lines = ["$MeshFormat", "2.2 0 8", "$EndMeshFormat", "$Nodes", str(len(xyz))]
    lines += [" ".join([str(i + 1)] + [repr(float(v)) for v in p]) for i, p in enumerate(xyz)]
    lines += ["$EndNodes", "$Elements", str(len(hexes))]
    lines += [" ".join([str(c + 1), "5", "2", "0", "1"] + [str(v + 1) for v in h]) for c, h in enumerate(hexes)]
    lines += ["$EndElements"]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def box_cells(nx, ny, nz, h=(1.0, 1.0, 1.0), jitter=0.0, seed=0):
    """Vertices of an nx x ny x nz lattice (interior ones optionally jittered)
    and its hexahedra in x-fastest order, GMSH corner order."""
    rng = np.random.default_rng(seed)
    vid = np.arange((nx + 1) * (ny + 1) * (nz + 1)).reshape(nz + 1, ny + 1, nx + 1)
    zz, yy, xx = np.meshgrid(np.arange(nz + 1), np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    xyz = np.stack([xx * h[0], yy * h[1], zz * h[2]], axis=-1).reshape(-1, 3).astype(np.float64)
    if jitter:
        inner = ((xx > 0) & (xx < nx) & (yy > 0) & (yy < ny) & (zz > 0) & (zz < nz)).ravel()
        xyz[inner] += rng.uniform(-jitter, jitter, (inner.sum(), 3)) * np.asarray(h)
    hexes = []
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                hexes.append([vid[k + c[2], j + c[1], i + c[0]] for c in CORNER])
    return xyz, np.asarray(hexes, np.int64)


def topo_arrays(path):
    r = lambda k: pygfx.h5_read(path, "/topology/" + k)  # noqa: E731
    return {"edge_to_vertex": r("edge_to_vertex").astype(np.int64).reshape(-1, 2),
            "surface_to_edge": r("surface_to_edge").astype(np.int64).reshape(-1, 4),
            "cell_to_surface": r("cell_to_surface").astype(np.int64).reshape(-1, 6),
            "vertex_to_coord": r("vertex_to_coord").reshape(-1, 3)}


def convert(tmp_path, xyz, hexes, fmt="41", aux=False, **kw):
    msh = str(tmp_path / "in.msh")
    (write_msh41 if fmt == "41" else write_msh22)(msh, xyz, hexes, **kw)
    model = str(tmp_path / "model.h5")
    pygfx.gmsh_to_topology(msh, model, str(tmp_path / "aux.h5") if aux else None)
    return model, topo_arrays(model)


TWO_STACKED = [[0, 1, 2, 3, 4, 5, 6, 7], [4, 5, 6, 7, 8, 9, 10, 11]]


def _points(n, seed=42):
    return np.random.default_rng(seed).random((n, 3))


def test_edges_are_deduplicated_canonical_and_one_based(tmp_path):
    _, t = convert(tmp_path, _points(8), [list(range(8))])
    assert t["edge_to_vertex"].shape == (12, 2)
    _, t = convert(tmp_path, _points(12), TWO_STACKED)
    e2v = t["edge_to_vertex"]
    assert e2v.shape == (20, 2)  # 12 + 12 - 4 shared
    assert np.all(e2v[:, 0] < e2v[:, 1]) and np.all(e2v > 0)


def test_surfaces_and_their_signs(tmp_path):
    _, t = convert(tmp_path, _points(8), [list(range(8))])
    assert t["surface_to_edge"].shape == (6, 4)
    assert np.all(t["cell_to_surface"] > 0)
    model, t = convert(tmp_path, _points(12), TWO_STACKED)
    s2e, c2s = t["surface_to_edge"], t["cell_to_surface"]
    assert s2e.shape == (11, 4)
    assert np.all(np.abs(s2e) >= 1) and np.all(np.abs(s2e) <= len(t["edge_to_vertex"]))
    assert c2s.shape == (2, 6)
    # the shared face is +z of the first cell and -z of the second, seen in opposite senses
    assert abs(c2s[0, 1]) == abs(c2s[1, 0]) and c2s[0, 1] != c2s[1, 0]
    counts, _ = pygfx.topology_summary(model)
    assert counts == (12, 20, 11, 2)


def test_topology_file_layout(tmp_path):
    model, _ = convert(tmp_path, _points(8), [list(range(8))], aux=True)
    lay = pygfx.h5_layout(model)
    assert lay["/topology/vertex_to_coord"] == ("f64", (8, 3))
    assert lay["/topology/edge_to_vertex"] == ("i64", (12, 2))
    assert lay["/topology/surface_to_edge"] == ("i64", (6, 4))
    assert lay["/topology/cell_to_surface"] == ("i64", (1, 6))
    for k, v in (("n_vertex", 8), ("n_edge", 12), ("n_surface", 6), ("n_cell", 1)):
        assert lay["/topology@" + k] == ("i64", ())
        assert pygfx.h5_attr(model, "/topology", k) == v
    aux = str(tmp_path / "aux.h5")
    assert pygfx.h5_read(aux, "/auxiliary/vertex_to_edge/indptr").size == 9
    assert pygfx.h5_read(aux, "/auxiliary/vertex_to_edge/indices").size == 24
    _, t = convert(tmp_path, _points(12), TWO_STACKED, aux=True)
    s2c = pygfx.h5_read(aux, "/auxiliary/surface_to_cell").reshape(-1, 2)
    assert int((s2c[:, 1] != 0).sum()) == 1  # one interior surface, on two cells
    assert np.all(s2c[:, 0] != 0)


def test_gmsh_41_and_22_read_alike_with_tags_and_blocks(tmp_path):
    xyz, hexes = box_cells(3, 2, 2, jitter=0.2, seed=4)
    _, a = convert(tmp_path, xyz, hexes, fmt="41", tags=np.arange(len(xyz)) * 7 + 100, split=3)
    _, b = convert(tmp_path, xyz, hexes, fmt="22")
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    assert np.array_equal(a["vertex_to_coord"], xyz)


def test_bad_gmsh_files_are_refused(tmp_path):
    p = tmp_path / "bin.msh"
    p.write_text("$MeshFormat\n4.1 1 8\n$EndMeshFormat\n")
    with pytest.raises(pygfx.GfxError, match="binary"):
        pygfx.gmsh_to_topology(str(p), str(tmp_path / "m.h5"))
    write_msh22(str(tmp_path / "nohex.msh"), _points(4), [])
    with pytest.raises(pygfx.GfxError, match="No hexahedron"):
        pygfx.gmsh_to_topology(str(tmp_path / "nohex.msh"), str(tmp_path / "m.h5"))


def test_corners_come_back_in_gmsh_order(tmp_path):
    """gll_geometry.py's face-membership rule, applied to the stored faces,
    returns every cell's vertices in the GMSH order the cell listed them."""
    xyz, hexes = box_cells(2, 1, 1)
    model, _ = convert(tmp_path, xyz, hexes)
    _, corners = pygfx.topology_summary(model)
    assert np.array_equal(corners - 1, hexes)


def _flat_spec(**kw):
    base = dict(nx=2, ny=1, order=1, layers=[(1e9, 6000.0, 3464.0, 1.0)], rf_amp=0.0, pml_xmin=0, pml_xmax=0,
                pml_ymin=0, pml_ymax=0, pml_zmax=0, record_depth_max=0.0, cfl_safety=0.5, output_dt=1e-4,
                duration=1e-3, want_coords=1)
    base.update(kw)
    return base


def test_unit_cube_geometry(tmp_path):
    """tests/preprocess/test_gll_geometry.py on the two-element mesh: corner
    coordinates in GMSH order, J = 1/8 on a unit cube, sum of w J = 1."""
    xyz, hexes = box_cells(2, 1, 1)
    model, _ = convert(tmp_path, xyz, hexes)
    m = pygfx.build_topology(model, **_flat_spec())
    c = m.coords.reshape(2, 2, 2, 2, 3)
    for e, x0 in ((0, 0.0), (1, 1.0)):
        for a, (i, j, k) in enumerate(CORNER):
            np.testing.assert_allclose(c[e, i, j, k], [x0 + i, j, k], atol=1e-12)
    np.testing.assert_allclose(m.jacobian, 0.125, atol=1e-12)
    mass = (m.mass / m.density).reshape(2, -1).sum(axis=1)  # density 1 layer: mass = w J
    np.testing.assert_allclose(mass, 1.0, atol=1e-12)


def _gauss_volume(cv, n=3):
    """Volume of a trilinear hexahedron by tensor Gauss-Legendre quadrature
    (exact: det J has degree 2 per variable)."""
    g, w = np.polynomial.legendre.leggauss(n)
    vol = 0.0
    sgn = np.array(CORNER) * 2 - 1
    for a, xa in zip(w, g):
        for b, xb in zip(w, g):
            for c, xc in zip(w, g):
                dN = np.stack([sgn[:, 0] * (1 + sgn[:, 1] * xb) * (1 + sgn[:, 2] * xc),
                               sgn[:, 1] * (1 + sgn[:, 0] * xa) * (1 + sgn[:, 2] * xc),
                               sgn[:, 2] * (1 + sgn[:, 0] * xa) * (1 + sgn[:, 1] * xb)], axis=1) / 8
                vol += a * b * c * np.linalg.det(dN.T @ cv)
    return vol


def test_jittered_cells_integrate_to_their_volume():
    xyz, hexes = box_cells(3, 3, 2, h=(500.0, 400.0, 300.0), jitter=0.2, seed=9)
    m = pygfx.build_hex(xyz, hexes, **_flat_spec(nx=3, ny=3, order=4, output_dt=1e-3, duration=1e-2))
    w = (m.mass / m.density).reshape(len(hexes), -1).sum(axis=1)
    want = np.array([_gauss_volume(xyz[h]) for h in hexes])
    np.testing.assert_allclose(w, want, rtol=1e-12)


def test_nodes_are_numbered_in_coordinate_order():
    xyz, hexes = box_cells(3, 2, 2, h=(800.0, 700.0, 600.0), jitter=0.15, seed=2)
    m = pygfx.build_hex(xyz, hexes, **_flat_spec(nx=3, ny=2, order=4, output_dt=1e-3, duration=1e-2))
    nx = m.node_xyz.reshape(-1, 3)
    assert m.n_node == (3 * 4 + 1) * (2 * 4 + 1) * (2 * 4 + 1)
    key = np.round(nx / 1e-9).astype(np.int64)
    order = np.lexsort((key[:, 2], key[:, 1], key[:, 0]))
    assert np.array_equal(order, np.arange(len(nx)))
    # every point of a node sits at the node
    pts = m.coords.reshape(-1, 3)
    np.testing.assert_allclose(pts, nx[m.ibool], atol=1e-9)


def _same_partition(ib_a, ib_b):
    """Both numberings group the element points into the same nodes."""
    fwd = {}
    for a, b in zip(ib_a.tolist(), ib_b.tolist()):
        if fwd.setdefault(a, b) != b:
            return False
    return len(set(fwd.values())) == len(fwd)


@pytest.mark.parametrize("jitter", [0.0, 0.12])
def test_box_as_general_cells_gives_the_box_inputs(jitter):
    spec = dict(nx=6, ny=5, nz=4, lx=6000.0, ly=5000.0, lz=4000.0, order=4, jitter=jitter, jitter_seed=3,
                layers=[(1500.0, 5000.0, 2900.0, 2500.0), (1e9, 6200.0, 3600.0, 2800.0)], rf_amp=0.05,
                rf_corr_len=2000.0, rf_modes=16, rf_seed=5, pml_xmin=1, pml_xmax=2, pml_ymin=1, pml_ymax=1,
                pml_zmax=1, f0_pml=1.0, output_dt=0.05, duration=1.0, record_depth_max=1200.0)
    box = pygfx.build_box(**spec)
    E = box.n_elem
    ib = box.ibool.reshape(E, 125)
    node = box.node_xyz.reshape(-1, 3)
    # the box's own (jittered) corners as general cells
    corner_pts = np.array([(i * 4 * 25 + j * 4 * 5 + k * 4) for (i, j, k) in CORNER])
    vids, hexes = np.unique(ib[:, corner_pts], return_inverse=True)
    gen = pygfx.build_hex(node[vids], hexes.reshape(E, 8), **spec)
    assert gen.n_elem == E and gen.n_node == box.n_node
    assert gen.bounds == box.bounds
    assert (gen.dt, gen.nsteps, gen.snapshot_stride) == (box.dt, box.nsteps, box.snapshot_stride)
    assert np.array_equal(gen.rec_cell, box.rec_cell) and len(box.rec_cell) > 0
    assert np.array_equal(gen.pml_region, box.pml_region)
    assert _same_partition(box.ibool, gen.ibool)
    for k in ("jacobian", "dxi_dx"):
        assert np.array_equal(getattr(gen, k), getattr(box, k)), k
    for k in ("mass", "density", "lambda_", "mu_", "damping", "coef_alpha", "coef_abar", "coords"):
        np.testing.assert_allclose(getattr(gen, k), getattr(box, k), rtol=1e-12, atol=1e-9, err_msg=k)


def test_general_mesh_round_trips_through_the_input_files(tmp_path):
    xyz, hexes = box_cells(5, 4, 3, h=(1000.0, 1000.0, 1000.0), jitter=0.1, seed=1)
    msh, model = str(tmp_path / "m.msh"), str(tmp_path / "model.h5")
    write_msh41(msh, xyz, hexes)
    pygfx.gmsh_to_topology(msh, model)
    m = pygfx.build_topology(model, nx=5, ny=4, order=4, pml_xmin=1, pml_xmax=1, pml_ymin=1, pml_ymax=1,
                             pml_zmax=1, output_dt=0.05, duration=0.5, record_depth_max=1000.0)
    stf = pygfx.ricker(1.0, 0.5, 1e20, m.dt, m.nsteps)
    d = str(tmp_path / "inputs")
    pygfx.write_inputs(m, d, (2500.0, 2000.0, 0.0), stf, n_parts=2)
    back = pygfx.read_inputs(d)
    assert back.n_elem == m.n_elem and back.n_node == m.n_node
    ids = np.concatenate([pygfx.h5_read(os.path.join(d, "partitions", f"partition_{r}.h5"),
                                        "/partition/local_cell_ids") for r in range(2)]).astype(np.int64)
    assert np.array_equal(np.sort(ids), np.arange(m.n_elem))
    for k in ("jacobian", "mass", "lambda_", "mu_"):
        got = getattr(back, k).reshape(m.n_elem, -1)
        np.testing.assert_allclose(got, getattr(m, k).reshape(m.n_elem, -1)[ids], rtol=1e-15, err_msg=k)
    cfg = pygfx.read_config(os.path.join(d, "config.h5"))
    assert cfg["xmax"] == 5000.0 and cfg["n_tilex"] == 1


def test_cells_off_the_lattice_take_pml_flags_from_boundary_faces():
    """A cell count that is not nx_elements * ny_elements * nz takes the
    reference's face test (preprocess/cli.py:258-261, boundary_detector.py):
    the PML cells are those with a face on x = xmin/xmax, y = ymin/ymax or
    z = zmax; the free surface z = 0 does not count.  On a 4 x 3 x 3 jittered
    lattice (36 cells, given with nx_elements 5, ny_elements 3) those are the
    cells with i in {0, 3}, j in {0, 2} or k = 2."""
    xyz, hexes = box_cells(4, 3, 3, h=(1000.0, 1000.0, 1000.0), jitter=0.15, seed=4)
    spec = _flat_spec(nx=5, ny=3, order=4, output_dt=1e-3, duration=1e-2)
    spec.update(pml_xmin=1, pml_xmax=1, pml_ymin=1, pml_ymax=1, pml_zmax=1)
    m = pygfx.build_hex(xyz, hexes, **spec)
    k, j, i = np.meshgrid(np.arange(3), np.arange(3), np.arange(4), indexing="ij")
    want = ((i == 0) | (i == 3) | (j == 0) | (j == 2) | (k == 2)).ravel()
    assert np.array_equal(m.pml_region != 0, want)
    abar = np.abs(m.coef_abar.reshape(36, -1)).max(axis=1)
    assert np.all(abar[~want] == 0.0) and np.all(abar[want] > 0.0)
    # on the lattice count the index rule stays (here: the same cells)
    lat = pygfx.build_hex(xyz, hexes, **dict(spec, nx=4, ny=3))
    assert np.array_equal(lat.pml_region != 0, want)


def test_cli_converts_gmsh(tmp_path):
    cli = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "gf-only-sem_amd", "bin",
                       "gf_greens_hip")
    xyz, hexes = box_cells(2, 2, 1)
    write_msh41(str(tmp_path / "in.msh"), xyz, hexes)
    r = subprocess.run([cli, "--gmsh", "in.msh", "--aux", "aux.h5"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    counts, corners = pygfx.topology_summary(str(tmp_path / "model.h5"))
    assert counts[3] == 4 and np.array_equal(corners - 1, hexes)
    assert (tmp_path / "aux.h5").exists()
    bad = subprocess.run([cli, "--gmsh", "missing.msh"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert bad.returncode == 1 and "cannot open" in bad.stderr
