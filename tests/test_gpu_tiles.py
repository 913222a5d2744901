"""Green's function tiles straight from the device against the post-processor.

The reference makes tiles in two steps: three forward runs (x, y, z) write
record files, then postprocess/cpp/main.cpp merges them and writes
tile_x*_y*.h5 (writer.hh).  Here the forward driver writes those record
files, the numpy restatement of the post-processor (oracle/postprocess.py)
turns them into expected tiles, and gfx_run_green_tiles, which never writes a
record file, must produce the same tile files: same names, same datasets,
attributes, types and shapes, and the same values bit for bit (both runs use
one fixed deterministic launch, so the solver's records are identical and
the device merge rounds as the FP64 host loop does)."""
import os
import subprocess

import numpy as np
import pytest

import meshcases as mc
import postprocess as pp
import pygfx

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not pygfx.h5_available(), reason="no HDF5 library to bind")]

LAUNCH = dict(assembly=pygfx.GFX_ASSEMBLY_GATHER, element_kernel=pygfx.GFX_ELEMENT_BLOCK, elems_per_block=1)


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if pygfx.device_count() < 1:
        pytest.fail("no HIP device visible; GPU tests must run on an MI355X box")


def make_inputs(d, precision, tilex=(1, 3), tiley=(2, 2)):
    """8 x 8 x 5 halfspace with 2-element C-PML: a 4 x 4 interior cut into tiles."""
    m = mc.small_halfspace(nx=8, ny=8, nz=5, duration=1.0)
    stf = pygfx.ricker(1.0, 0.6, 1e20, m.dt, m.nsteps)
    xyz = (0.5 * (m.bounds[0] + m.bounds[1]) + 300.0, 0.5 * (m.bounds[2] + m.bounds[3]) - 200.0, 0.0)
    pygfx.write_inputs(m, d, xyz, stf, precision=precision)
    cfg_path = os.path.join(d, "config.h5")
    pygfx.h5_write(cfg_path, "/simulation/tilex_elements", np.asarray(tilex, np.int64))
    pygfx.h5_write(cfg_path, "/simulation/tiley_elements", np.asarray(tiley, np.int64))
    return m


def read_cfg(d):
    p = os.path.join(d, "config.h5")
    a = lambda obj, k: pygfx.h5_attr(p, obj, k)  # noqa: E731
    cfg = {k: int(a("/simulation", k)) for k in ("nx_elements", "ny_elements", "pml_xmin", "pml_ymin")}
    cfg.update({k: a("/simulation", k) for k in ("output_dt_s", "solver_dt", "record_depth_max_m",
                                                  "record_depth_actual_m")})
    cfg["tilex_elements"] = pygfx.h5_read(p, "/simulation/tilex_elements").astype(int).tolist()
    cfg["tiley_elements"] = pygfx.h5_read(p, "/simulation/tiley_elements").astype(int).tolist()
    bounds = tuple(a("/domain", k) for k in ("xmin", "xmax", "ymin", "ymax", "zmin", "zmax"))
    src = tuple(a("/source", k) for k in ("x", "y", "z"))
    return cfg, bounds, src, pygfx.h5_read(p, "/source/stf_t"), pygfx.h5_read(p, "/source/stf_values")


def expected_tiles(d, m, uva):
    """The post-processor over the forward driver's record files."""
    cfg, bounds, src, stf_t, stf_v = read_cfg(d)
    mass = pygfx.read_inputs(d).mass.reshape(-1, 125)
    steps = [s for s in range(m.nsteps) if s % m.snapshot_stride == 0]
    merged = []
    for dname in "xyz":
        base = os.path.join(d, "wavefields", dname)
        first = os.path.join(base, f"record_0_{steps[0]}.h5")
        nrc = int(pygfx.h5_attr(first, "/", "n_rec_cell"))
        f = {"gll_node_ids": pygfx.h5_read(first, "/gll_node_ids").astype(np.int64),
             "gll_node_coords": pygfx.h5_read(first, "/gll_node_coords").reshape(-1, 3),
             "cell_gll_node_index": pygfx.h5_read(first, "/cell_gll_node_index").astype(np.int64).reshape(nrc, 125),
             "recording_cell_model_index": pygfx.h5_read(first, "/recording_cell_model_index").astype(np.int64)}
        group = []
        for s in steps:
            p = os.path.join(base, f"record_0_{s}.h5")
            rec = {"strain": pygfx.h5_read(p, "/strain").reshape(nrc, 125, 6)}
            if uva:
                rec["displacement"] = pygfx.h5_read(p, "/displacement").reshape(nrc, 125, 3)
            group.append([rec])
        merged.append(pp.merge_direction([f], group, cell_mass=mass, n_model_cell=mass.shape[0]))
    return pp.tiles(cfg, bounds, *merged, stf_t, stf_v, src)


STR_ATTRS = ("version", "basis", "source_directions", "greens_quantities")
I32_ATTRS = ("tile_x_index", "tile_y_index", "excludes_pml")


def check_tile(path, want, precision):
    ftype = "f32" if precision == "float32" else "f64"
    lay = pygfx.h5_layout(path)
    names = {k for k in want if not k.startswith("@") and "@" not in k}
    attrs = {k for k in want if k.startswith("@")}
    assert {k for k in lay if "@" not in k} == names, path
    assert {k for k in lay if k.startswith("/@")} == {"/" + k for k in attrs}, path
    assert {k for k in lay if k.startswith("/time@")} == {"/time@dt", "/time@nsteps"}
    for k in STR_ATTRS:
        assert lay["/@" + k][0] == "str"
        assert pygfx.h5_attr(path, "/", k, string=True) == want["@" + k], k
    for k in I32_ATTRS:
        assert lay["/@" + k] == ("i32", ())
        assert int(pygfx.h5_attr(path, "/", k)) == want["@" + k], k
    for k in attrs - {"@" + a for a in STR_ATTRS + I32_ATTRS}:
        got = pygfx.h5_attr_array(path, "/", k[1:])
        assert lay["/" + k][0] == "f64", k
        assert np.array_equal(got, np.atleast_1d(np.asarray(want[k], np.float64))), k
    assert lay["/time@nsteps"] == ("i32", ()) and int(pygfx.h5_attr(path, "/time", "nsteps")) == want["/time@nsteps"]
    assert pygfx.h5_attr(path, "/time", "dt") == want["/time@dt"]
    types = {"/mesh/gll_node_ids": "i64", "/mesh/cell_gll_node_index": "i32"}
    for k in names:
        w = np.asarray(want[k])
        typ = types.get(k, ftype if k.startswith("/field/") else "f64")
        assert lay[k] == (typ, w.shape), (k, lay[k], w.shape)
        got = pygfx.h5_read(path, k)
        if typ == "f32":
            w = w.astype(np.float32)
        assert np.array_equal(got, w.astype(np.float64).ravel()), (path, k)


@pytest.mark.parametrize("precision,uva", [("float32", True), ("float64", True), ("float32", False)])
def test_device_tiles_equal_the_post_processed_records(tmp_path, precision, uva):
    d = str(tmp_path)
    m = make_inputs(d, precision)
    pygfx.run_forward(d, directions="x,y,z", record_uva=uva, restart_stride=0, **LAUNCH)
    want = expected_tiles(d, m, uva)
    assert len(want) == 4
    out = os.path.join(d, "tiles")
    n = pygfx.run_green_tiles(d, output_dir=out, with_uva=uva, **LAUNCH)
    assert n == len(want)
    assert sorted(os.listdir(out)) == sorted(want)
    for name, w in want.items():
        check_tile(os.path.join(out, name), w, precision)
    # the strain is not trivially zero: the source reached every tile
    for name in want:
        g = pygfx.h5_read(os.path.join(out, name), "/field/greens_tensor")
        assert np.abs(g).max() > 0, name


def test_reference_tile_schema_checks(tmp_path):
    """tests/test_postprocess_tile.cpp's assertions on a device tile."""
    d = str(tmp_path)
    make_inputs(d, "float32", tilex=(4,), tiley=(4,))
    out = os.path.join(d, "greenfun")
    assert pygfx.run_green_tiles(d, **LAUNCH) == 1
    path = os.path.join(out, "tile_x000_y000.h5")
    lay = pygfx.h5_layout(path)
    assert "/@source_xyz_m" in lay and lay["/@source_xyz_m"][1] == (3,)
    assert "strain" in pygfx.h5_attr(path, "/", "greens_quantities", string=True)
    assert len(lay["/mesh/gll_node_coords"][1]) == 2 and lay["/mesh/gll_node_coords"][1][1] == 3
    assert len(lay["/mesh/gll_node_ids"][1]) == 1
    g = lay["/field/greens_tensor"][1]
    assert len(g) == 4 and g[2] == 6 and g[3] == 3


def test_slab_capture_rejects_what_it_cannot_serve():
    m = mc.small_halfspace(nx=6, ny=6, nz=4, pml=1, duration=0.3)
    s = pygfx.Solver(m, n_fields=3, record_capacity=2, record_uva=False, **LAUNCH)
    try:
        off = np.array([0, 1], np.int64)
        with pytest.raises(pygfx.GfxError, match="slab node out of range"):
            pygfx.GreenSlab(off, [0], [1.0], [1])
        slab = pygfx.GreenSlab(off, [0], [1.0], [0], displacement=True)
        s.run(2 * m.snapshot_stride)
        with pytest.raises(pygfx.GfxError, match="u, v, a records"):
            slab.capture(s, 0)
        far = pygfx.GreenSlab(off, [10 ** 8], [1.0], [0])
        with pytest.raises(pygfx.GfxError, match="recording points"):
            far.capture(s, 0)
    finally:
        s.close()


def test_cli_writes_the_tiles(tmp_path):
    d = str(tmp_path)
    make_inputs(d, "float32", tilex=(2, 2), tiley=(4,))
    cli = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "gf-only-sem_amd", "bin",
                       "gf_greens_hip")
    r = subprocess.run([cli, "-o", "out", "--assembly", "gather", "--quiet"], cwd=d, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    assert "STAT_NTILES=2" in r.stdout
    assert sorted(os.listdir(os.path.join(d, "out"))) == ["tile_x000_y000.h5", "tile_x001_y000.h5"]


def test_device_tiles_with_complete_cpml(tmp_path):
    """--cpml-complete on both paths: the tiles equal the post-processed
    records of the forward driver's complete-C-PML run bit for bit, and differ
    from the reference-C-PML tiles once reflections reach the tiles."""
    d = str(tmp_path)
    m = mc.small_halfspace(nx=8, ny=8, nz=5, duration=3.0, want_profiles=1, pml_profile=1)
    stf = pygfx.ricker(1.0, 0.6, 1e20, m.dt, m.nsteps)
    xyz = (0.5 * (m.bounds[0] + m.bounds[1]) + 300.0, 0.5 * (m.bounds[2] + m.bounds[3]) - 200.0, 0.0)
    pygfx.write_inputs(m, d, xyz, stf, precision="float32")
    cfg_path = os.path.join(d, "config.h5")
    pygfx.h5_write(cfg_path, "/simulation/tilex_elements", np.asarray((4,), np.int64))
    pygfx.h5_write(cfg_path, "/simulation/tiley_elements", np.asarray((4,), np.int64))
    pygfx.run_forward(d, directions="x,y,z", record_uva=True, restart_stride=0, cpml_complete=True, **LAUNCH)
    want = expected_tiles(d, m, True)
    out = os.path.join(d, "tiles_c")
    assert pygfx.run_green_tiles(d, output_dir=out, with_uva=True, cpml_complete=True, **LAUNCH) == len(want)
    for name, w in want.items():
        check_tile(os.path.join(out, name), w, "float32")
    ref = os.path.join(d, "tiles_r")
    pygfx.run_green_tiles(d, output_dir=ref, with_uva=True, **LAUNCH)
    for name in want:
        a = pygfx.h5_read(os.path.join(out, name), "/field/greens_tensor")
        b = pygfx.h5_read(os.path.join(ref, name), "/field/greens_tensor")
        diff = np.linalg.norm(a - b) / np.linalg.norm(b)
        print(f"{name}: complete vs reference C-PML tile rel L2 {diff:.3e}")
        assert diff > 1e-3
