"""File formats of the forward path: the partition / config readers (with the
single-process merge), the per-snapshot record writer and the input writer.

The first cases restate the reference's own I/O tests on synthetic files
(tests/test_io.cpp:16-207, tests/test_record.cpp:11-114); the rest pin the
merge rules of forward/share/src/io.cpp:319-450 and the on-disk schema
(partition / config / record layouts) on generated boxes.  HDF5 only moves
bytes, so equality here is exact."""
import os

import numpy as np
import pytest

import pygfx

pytestmark = pytest.mark.skipif(not pygfx.h5_available(), reason="no HDF5 library to bind")


def small_box(**kw):
    args = dict(nx=6, ny=5, nz=4, lx=6e3, ly=5e3, lz=4e3, pml_xmin=1, pml_xmax=1, pml_ymin=1, pml_ymax=1,
                pml_zmax=1, output_dt=0.05, duration=1.0, record_depth_max=1000.0)
    args.update(kw)
    return pygfx.build_box(**args)


def synth_partition(path, rank, ngll, n):
    """The synthetic partition of tests/test_io.cpp:16-103."""
    nn = n * ngll ** 3
    pygfx.h5_write(path, "/field/cell/coords", np.zeros((n, ngll, ngll, ngll, 3)), append=False)
    pygfx.h5_write(path, "/field/cell/jacobian", np.ones((n, ngll, ngll, ngll)))
    for name in ("dxi_dx", "mass", "vp", "vs", "density", "damping"):
        pygfx.h5_write(path, f"/field/cell/{name}", np.ones((n, ngll, ngll, ngll)))
    pygfx.h5_write(path, "/partition/local_cell_ids", np.arange(n, dtype=np.int64) + rank * n + 1)
    pygfx.h5_write(path, "/partition/ghost_cell_ids", np.zeros(0, np.int64))
    pygfx.h5_write(path, "/partition/ghost_owners", np.zeros(0, np.int32))
    pygfx.h5_write_attr(path, "/partition/exchange", "note", "single rank")  # empty exchange group
    assert nn == pygfx.h5_read(path, "/field/cell/jacobian").size


def test_read_partition_roundtrip_like_reference(tmp_path):
    # tests/test_io.cpp:105-122
    p = str(tmp_path / "partition_0.h5")
    synth_partition(p, rank=0, ngll=4, n=2)
    d = pygfx.read_partition(p)
    assert d["n_local_cell"] == 2 and d["ngll"] == 4 and d["n_local_cell_ids"] == 2
    assert d["n_exchange_neighbors"] == 0 and d["has_cpml"] == 0 and d["has_recording"] == 0
    assert d["n_jacobian"] == 2 * 64


def test_read_config_roundtrip_like_reference(tmp_path):
    # tests/test_io.cpp:124-207
    p = str(tmp_path / "config.h5")
    stf = np.linspace(0.0, 1.0, 10)
    pygfx.h5_write(p, "/source/stf_values", stf, append=False)
    pygfx.h5_write(p, "/source/stf_t", stf)
    for k, v in (("title", "test"), ("polynomial_order", 3), ("solver_dt", 0.001), ("output_dt_s", 0.01),
                 ("snapshot_stride", 10), ("nsteps", 500), ("cfl_safety", 0.5), ("snapshot_precision", "float32")):
        pygfx.h5_write_attr(p, "/simulation", k, v)
    for k, v in (("xmin", 0.0), ("xmax", 1.0), ("ymin", 0.0), ("ymax", 2.0), ("zmin", 0.0), ("zmax", 3.0)):
        pygfx.h5_write_attr(p, "/domain", k, v)
    for k, v in (("x", 0.5), ("y", 0.6), ("z", 0.0)):
        pygfx.h5_write_attr(p, "/source", k, v)
    c = pygfx.read_config(p)
    assert c["title"] == "test" and c["polynomial_order"] == 3
    assert c["solver_dt"] == 0.001 and c["output_dt_s"] == 0.01
    assert c["snapshot_stride"] == 10 and c["nsteps"] == 500
    assert c["snapshot_precision"] == "float32" and c["n_stf_values"] == 10
    assert c["source_x"] == 0.5 and c["source_y"] == 0.6 and c["ymax"] == 2.0 and c["zmax"] == 3.0


def test_read_config_defaults_required_and_legacy(tmp_path):
    # io.cpp:585-613: polynomial_order defaults to 3, output_dt to solver_dt,
    # stride to 1; solver_dt and nsteps are required
    p = str(tmp_path / "c.h5")
    pygfx.h5_write_attr(p, "/simulation", "solver_dt", 0.002)
    pygfx.h5_write_attr(p, "/simulation", "nsteps", 7)
    c = pygfx.read_config(p)
    assert c["polynomial_order"] == 3 and c["output_dt_s"] == 0.002 and c["snapshot_stride"] == 1
    assert c["title"] == "untitled" and c["snapshot_precision"] == "float64"
    q = str(tmp_path / "bad.h5")
    pygfx.h5_write_attr(q, "/simulation", "nsteps", 7)
    with pytest.raises(pygfx.GfxError, match="solver_dt"):
        pygfx.read_config(q)
    # legacy flat datasets (io.cpp:631-654)
    r = str(tmp_path / "legacy.h5")
    pygfx.h5_write(r, "dt", np.array([0.004]), append=False)
    pygfx.h5_write(r, "nsteps", np.array([12.0]))
    pygfx.h5_write(r, "use_float32", np.array([1], np.int64))
    c = pygfx.read_config(r)
    assert c["solver_dt"] == 0.004 and c["nsteps"] == 12 and c["polynomial_order"] == 3
    assert c["snapshot_precision"] == "float32"


# Appendix of the partition schema (preprocess/model_writer.py:304-446 as read by io.cpp:182-317)
PARTITION_SCHEMA = {
    "/field/cell/coords": ("f64", lambda n: [n, 5, 5, 5, 3]),
    "/field/cell/jacobian": ("f64", lambda n: [n, 5, 5, 5]),
    "/field/cell/dxi_dx": ("f64", lambda n: [n, 5, 5, 5, 9]),
    "/field/cell/mass": ("f64", lambda n: [n, 5, 5, 5]),
    "/field/cell/vp": ("f64", lambda n: [n, 5, 5, 5]),
    "/field/cell/vs": ("f64", lambda n: [n, 5, 5, 5]),
    "/field/cell/density": ("f64", lambda n: [n, 5, 5, 5]),
    "/field/cell/lambda": ("f64", lambda n: [n, 5, 5, 5]),
    "/field/cell/mu": ("f64", lambda n: [n, 5, 5, 5]),
    "/field/cell/damping": ("f64", lambda n: [n, 5, 5, 5]),
    "/field/cell/pml_region": ("i32", lambda n: [n]),
    "/field/cell/pml_coef_alpha": ("f64", lambda n: [n, 125, 9]),
    "/field/cell/pml_coef_abar": ("f64", lambda n: [n, 125, 5]),
    "/field/cell/local_cell2global_node": ("i32", lambda n: [n * 125]),
    "/field/cell/local_cell2rank_node": ("i32", lambda n: [n * 125]),
    "/partition/local_cell_ids": ("i64", lambda n: [n]),
}


def parse_describe(text):
    ds, attrs = {}, {}
    for line in text.strip().splitlines():
        name, typ, dims = line.split()
        dims = [int(x) for x in dims.strip("[]").split(",") if x]
        if "@" in name:
            attrs[name] = (typ, dims)
        else:
            ds[name] = (typ, dims)
    return ds, attrs


def test_written_inputs_follow_the_partition_and_config_schema(tmp_path):
    m = small_box()
    stf = pygfx.ricker(1.0, 0.5, 1e20, m.dt, m.nsteps)
    pygfx.write_inputs(m, tmp_path, (3000.0, 2000.0, 0.0), stf, n_parts=2)  # a surface vertex
    sizes = []
    for r in range(2):
        ds, attrs = parse_describe(pygfx.h5_describe(tmp_path / "partitions" / f"partition_{r}.h5"))
        n = ds["/partition/local_cell_ids"][1][0]
        sizes.append(n)
        for name, (typ, dims) in PARTITION_SCHEMA.items():
            assert ds[name] == (typ, dims(n)), name
        assert "/field/cell@n_rank_node" in attrs and "/partition@n_ranks" in attrs
    assert sum(sizes) == m.n_elem
    ds, attrs = parse_describe(pygfx.h5_describe(tmp_path / "config.h5"))
    for a in ("title", "polynomial_order", "solver_dt", "output_dt_s", "snapshot_stride", "nsteps",
              "snapshot_precision", "record_depth_max_m", "record_depth_actual_m"):
        assert f"/simulation@{a}" in attrs, a
    for a in ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax"):
        assert f"/domain@{a}" in attrs
    assert ds["/source/stf_values"] == ("f64", [m.nsteps]) and ds["/source/stf_t"] == ("f64", [m.nsteps])
    assert ds["/source/cells/cell_ids"] == ("i64", [4]) and ds["/source/cells/weights"] == ("f64", [4, 5, 5, 5])
    c = pygfx.read_config(str(tmp_path / "config.h5"))
    assert c["solver_dt"] == m.dt and c["nsteps"] == m.nsteps and c["snapshot_stride"] == m.snapshot_stride
    assert c["n_src_cell"] == 4 and c["snapshot_precision"] == "float32"
    assert np.array_equal(pygfx.h5_read(tmp_path / "config.h5", "/source/stf_values"), stf)
    # the located weights are normalised over the containing cells (source_locator.py:361-365)
    assert abs(pygfx.h5_read(tmp_path / "config.h5", "/source/cells/weights").sum() - 1.0) < 1e-12


@pytest.mark.parametrize("n_parts", [1, 3])
def test_inputs_roundtrip_through_the_merge(tmp_path, n_parts):
    m = small_box(jitter=0.1, jitter_seed=5)
    stf = pygfx.ricker(1.0, 0.5, 1e20, m.dt, m.nsteps)
    pygfx.write_inputs(m, tmp_path, (3000.0, 2500.0, 0.0), stf, n_parts=n_parts)
    m2 = pygfx.read_inputs(tmp_path)
    # the merge concatenates the partitions' cells (io.cpp:361-447): cell c of the
    # merged mesh is global cell ids[c]
    ids = np.concatenate([pygfx.h5_read(tmp_path / "partitions" / f"partition_{r}.h5", "/partition/local_cell_ids")
                          for r in range(n_parts)]).astype(np.int64)
    assert np.array_equal(np.sort(ids), np.arange(m.n_elem))
    E = m.n_elem
    for k in ("coords", "jacobian", "dxi_dx", "mass", "density", "lambda_", "mu_", "damping", "ibool",
              "pml_region", "coef_alpha", "coef_abar"):
        assert np.array_equal(m.arrays[k].reshape(E, -1)[ids], m2.arrays[k].reshape(E, -1)), k
    assert np.array_equal(m.arrays["node_xyz"], m2.arrays["node_xyz"])
    assert np.array_equal(np.sort(ids[m2.rec_cell]), np.sort(m.rec_cell))
    assert (m2.n_node, m2.dt, m2.nsteps, m2.snapshot_stride) == (m.n_node, m.dt, m.nsteps, m.snapshot_stride)
    merged = pygfx.read_partition(str(tmp_path / "partitions"), merge=True)
    # n_rank_node = max global id + 1; ibool := local_cell2global_node (io.cpp:361-368,442-447)
    assert merged["n_rank_node"] == m.ibool.max() + 1 == m.n_node
    assert merged["n_local_cell"] == m.n_elem and merged["n_local_cell2global_node"] == 0
    assert merged["n_rec_cell"] == len(m.rec_cell)
    uniq = np.unique(m.ibool.reshape(-1, 125)[m.rec_cell])
    assert merged["n_unique_gll"] == uniq.size  # recording nodes de-duplicated across partitions
    # per partition: compact rank numbering with its own n_rank_node
    for r in range(n_parts):
        d = pygfx.read_partition(str(tmp_path / "partitions" / f"partition_{r}.h5"))
        rib = pygfx.h5_read(tmp_path / "partitions" / f"partition_{r}.h5", "/field/cell/local_cell2rank_node")
        gib = pygfx.h5_read(tmp_path / "partitions" / f"partition_{r}.h5", "/field/cell/local_cell2global_node")
        assert d["n_rank_node"] == rib.max() + 1 == np.unique(gib).size


def test_stretch_profiles_roundtrip(tmp_path):
    """pml_K, pml_d, pml_alpha [n_cell, 125, 3] (the datasets pml_cpml.py
    computes) are written when generated and merged back by cell id."""
    m = small_box(want_profiles=1, pml_profile=1)
    stf = pygfx.ricker(1.0, 0.5, 1e20, m.dt, m.nsteps)
    pygfx.write_inputs(m, tmp_path, (3000.0, 2500.0, 0.0), stf, n_parts=2)
    assert "/field/cell/pml_K f64 [" in pygfx.h5_describe(tmp_path / "partitions" / "partition_0.h5")
    m2 = pygfx.read_inputs(tmp_path)
    ids = np.concatenate([pygfx.h5_read(tmp_path / "partitions" / f"partition_{r}.h5", "/partition/local_cell_ids")
                          for r in range(2)]).astype(np.int64)
    for k in ("pml_K", "pml_d", "pml_alpha"):
        assert np.array_equal(m.arrays[k].reshape(m.n_elem, -1)[ids], m2.arrays[k].reshape(m.n_elem, -1)), k
    bare = tmp_path / "bare"
    pygfx.write_inputs(small_box(), bare, (3000.0, 2500.0, 0.0), stf, n_parts=1)
    assert "pml_K" not in pygfx.h5_describe(bare / "partitions" / "partition_0.h5")
    assert pygfx.read_inputs(bare).arrays["pml_K"] is None


def test_partitions_are_centroid_slabs_along_the_longest_axis(tmp_path):
    """Without METIS the reference sorts cell centroids along the axis of
    largest spread and cuts the order into chunks of ceil(E / n)
    (preprocess/partition.py:75-110); each partition lists its cells in
    increasing global order."""
    m = small_box(jitter=0.1, jitter_seed=5)
    E, n = m.n_elem, 4
    stf = pygfx.ricker(1.0, 0.5, 1e20, m.dt, m.nsteps)
    pygfx.write_inputs(m, tmp_path, (3000.0, 2500.0, 0.0), stf, n_parts=n)
    cen = m.coords.reshape(E, 125, 3).mean(axis=1)
    ax = int(np.argmax(np.ptp(cen, axis=0)))
    chunk = -(-E // n)
    prev_max = -np.inf
    for r in range(n):
        ids = pygfx.h5_read(tmp_path / "partitions" / f"partition_{r}.h5", "/partition/local_cell_ids").astype(np.int64)
        assert np.all(np.diff(ids) > 0)
        assert len(ids) == (chunk if r < n - 1 else E - chunk * (n - 1))
        c = cen[ids, ax]
        assert c.min() >= prev_max
        prev_max = c.max()


@pytest.mark.parametrize("precision", ["float32", "float64"])
def test_record_file_schema_and_values(tmp_path, precision):
    """RecordWriter (record.cpp:175-257 / tests/test_record.cpp): one file per
    snapshot with the recording map and [1, n_rec, ngll^3, {6,3}] fields; the
    map of merged partitions points each cell node at its global node."""
    m = small_box()
    stf = pygfx.ricker(1.0, 0.5, 1e20, m.dt, m.nsteps)
    inp = tmp_path / "in"
    pygfx.write_inputs(m, inp, (3000.0, 2500.0, 0.0), stf, n_parts=3, precision=precision)
    nrec = len(m.rec_cell)
    rng = np.random.default_rng(0)
    strain = rng.standard_normal(nrec * 125 * 6) * 1e-7
    disp, vel, acc = (rng.standard_normal(nrec * 125 * 3) for _ in range(3))
    pygfx.write_record(inp, tmp_path / "wavefields", "y", 40, strain, disp, vel, acc)
    path = tmp_path / "wavefields" / "y" / "record_0_40.h5"
    assert path.exists()
    ds, attrs = parse_describe(pygfx.h5_describe(path))
    ft = "f32" if precision == "float32" else "f64"
    assert ds["/strain"] == (ft, [1, nrec, 125, 6])
    for name in ("/displacement", "/velocity", "/acceleration"):
        assert ds[name] == (ft, [1, nrec, 125, 3])
    n_u = ds["/gll_node_ids"][1][0]
    assert ds["/gll_node_ids"][0] == "i64" and ds["/gll_node_coords"] == ("f64", [n_u, 3])
    assert ds["/cell_gll_node_index"] == ("i32", [nrec, 125])
    assert ds["/recording_cell_model_index"] == ("i64", [nrec])
    for a in ("source_direction", "rank", "ngll", "basis", "n_rec_cell", "n_unique_gll", "excludes_pml",
              "record_depth_max_m", "record_depth_actual_m"):
        assert f"/@{a}" in attrs, a
    assert pygfx.h5_attr(path, "/", "source_direction", string=True) == "y"
    assert pygfx.h5_attr(path, "/", "basis", string=True) == "gll"
    assert pygfx.h5_attr(path, "/", "n_rec_cell") == nrec and pygfx.h5_attr(path, "/", "ngll") == 5
    cast = (lambda x: x.astype(np.float32).astype(np.float64)) if precision == "float32" else (lambda x: x)
    assert np.array_equal(pygfx.h5_read(path, "/strain"), cast(strain))
    assert np.array_equal(pygfx.h5_read(path, "/velocity"), cast(vel))
    ids = pygfx.h5_read(path, "/gll_node_ids").astype(np.int64)
    cidx = pygfx.h5_read(path, "/cell_gll_node_index").astype(np.int64)
    assert np.unique(ids).size == ids.size
    want = m.ibool.reshape(-1, 125)[pygfx.h5_read(path, "/recording_cell_model_index").astype(np.int64)]
    assert np.array_equal(ids[cidx].reshape(nrec, 125), want)
    xyz = pygfx.h5_read(path, "/gll_node_coords").reshape(-1, 3)
    assert np.array_equal(xyz, m.node_xyz.reshape(-1, 3)[ids])


def test_stations_file(tmp_path):
    m = small_box()
    pts = np.array([[2500.0, 2000.0, 0.0], [3000.0, 2000.0, 0.0], [3300.0, 2100.0, 700.0]])
    p = tmp_path / "stations.h5"
    pygfx.write_stations(p, m, pts)
    assert pygfx.h5_attr(p, "/stations", "n_stations") == 3
    for k in range(3):
        w = pygfx.h5_read(p, f"/stations/s{k}/cells/weights")
        assert abs(w.sum() - 1.0) < 1e-12
        assert pygfx.h5_attr(p, f"/stations/s{k}", "z") == pts[k, 2]
    assert pygfx.h5_read(p, "/stations/s1/cells/cell_ids").size == 4  # surface vertex: 4 cells


def test_missing_inputs_raise(tmp_path):
    with pytest.raises(pygfx.GfxError):
        pygfx.read_inputs(tmp_path)
    os.makedirs(tmp_path / "partitions")
    with pytest.raises(pygfx.GfxError, match="no partition files"):
        pygfx.read_partition(str(tmp_path / "partitions"), merge=True)
    # a partition without local_cell2global_node cannot drive a single-process run
    synth_partition(str(tmp_path / "partitions" / "partition_0.h5"), 0, 5, 2)
    pygfx.h5_write_attr(str(tmp_path / "config.h5"), "/simulation", "solver_dt", 0.01)
    pygfx.h5_write_attr(str(tmp_path / "config.h5"), "/simulation", "nsteps", 3)
    with pytest.raises(pygfx.GfxError, match="local_cell2global_node"):
        pygfx.read_inputs(tmp_path)
