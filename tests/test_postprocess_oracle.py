"""Pins the numpy restatement of the tile post-processor (oracle/postprocess.py).

The reference holds no golden tiles (tests/test_postprocess_tile.cpp checks
the schema of a tile it writes itself), so the restatement is held to:
- hand-computed answers on tiny record sets whose arithmetic is exact
  (mass-weighted node values, the displacement scaled by the same weight sum,
  the direction-major order of the 18 values, first-appearance node order
  across rank files);
- a plain sequential loop written from postprocess/cpp/main.cpp:306-426
  (file, cell, point order; product then sum; one multiply by the inverse
  weight sum) on random records, bit for bit;
- the invariant that a field constant over every record point stays constant;
- hand-worked tile keys, tile bounds and STF down-sampling."""
import math

import numpy as np
import pytest

import postprocess as pp

NPC = 125


def _file(ids, coords, cgi, rcm=None):
    return {"gll_node_ids": np.asarray(ids, np.int64), "gll_node_coords": np.asarray(coords, np.float64),
            "cell_gll_node_index": np.asarray(cgi, np.int64).reshape(-1, NPC),
            "recording_cell_model_index": None if rcm is None else np.asarray(rcm, np.int64)}


def _two_cells():
    """Cells 0 and 1 share node 0 (point 124 of cell 0, point 0 of cell 1);
    every other point has a node of its own."""
    c0 = np.arange(NPC) + 1
    c1 = np.arange(NPC) + NPC + 1
    c0[124] = 0
    c1[0] = 0
    n = 2 * NPC + 1
    ids = np.arange(n) + 1
    coords = np.zeros((n, 3))
    return _file(ids, coords, np.stack([c0, c1]), rcm=[0, 1]), n


def test_shared_node_is_the_mass_weighted_mean():
    f, n = _two_cells()
    strain = np.zeros((2, NPC, 6))
    strain[0, 124] = 2.0
    strain[1, 0] = 6.0
    mass = np.ones((2, NPC))
    mass[0, 124] = 1.0
    mass[1, 0] = 3.0
    m = pp.merge_direction([f], [[{"strain": strain}]], cell_mass=mass, n_model_cell=2)
    # (1*2 + 3*6) / (1 + 3) = 5, exact in binary
    assert np.array_equal(m["strain"][0, 0], np.full(6, 5.0))
    # no mass file: the plain mean (2 + 6) / 2
    m1 = pp.merge_direction([f], [[{"strain": strain}]])
    assert np.array_equal(m1["strain"][0, 0], np.full(6, 4.0))
    # a cell whose model index is out of range weighs 1
    f2 = dict(f, recording_cell_model_index=np.array([0, 7]))
    m2 = pp.merge_direction([f2], [[{"strain": strain}]], cell_mass=mass, n_model_cell=2)
    assert np.array_equal(m2["strain"][0, 0], np.full(6, 4.0))


def test_displacement_is_summed_then_scaled_by_the_weight_sum():
    f, n = _two_cells()
    strain = np.zeros((2, NPC, 6))
    disp = np.zeros((2, NPC, 3))
    disp[0, 124] = 8.0
    disp[1, 0] = 8.0
    disp[0, 5] = 3.0
    mass = np.ones((2, NPC))
    mass[0, 124], mass[1, 0], mass[0, 5] = 1.0, 3.0, 0.5
    m = pp.merge_direction([f], [[{"strain": strain, "displacement": disp}]], cell_mass=mass, n_model_cell=2)
    # main.cpp:394-423: (8 + 8) * (1 / (1 + 3)) = 4, and 3 * (1 / 0.5) = 6 on a single point
    assert np.array_equal(m["displacement"][0, 0], np.full(3, 4.0))
    assert np.array_equal(m["displacement"][0, 6], np.full(3, 6.0))  # point 5 of cell 0 is node 6


def test_tensor_layout_is_force_direction_major():
    f, n = _two_cells()
    merged = []
    for d in range(3):
        strain = np.zeros((2, NPC, 6))
        strain[:, :, :] = 100.0 * d + np.arange(6)
        merged.append(pp.merge_direction([f], [[{"strain": strain}]]))
    g = pp.greens_tensor(*merged)
    assert g.shape == (1, n, 6, 3)
    flat = g.reshape(1, n, 18)
    for d in range(3):
        for c in range(6):
            k = d * 6 + c
            assert flat[0, 3, k] == 100.0 * d + c
            # what a reader indexing [.., 6, 3] finds there
            assert g[0, 3, k // 3, k % 3] == 100.0 * d + c
    # node 125 belongs to no cell: it merges to zero
    assert not flat[0, NPC + 1].any()


def test_nodes_merge_in_first_appearance_order_across_files():
    a = _file([10, 20, 30], np.arange(9).reshape(3, 3), np.zeros((1, NPC), np.int64))
    b = _file([30, 40, 10], np.arange(9).reshape(3, 3) + 100, np.full((1, NPC), 1))
    s = [[{"strain": np.ones((1, NPC, 6))}, {"strain": 3 * np.ones((1, NPC, 6))}]]
    m = pp.merge_direction([a, b], s)
    assert m["gll_node_ids"].tolist() == [10, 20, 30, 40]
    # coordinates of the first file that names a node
    assert m["gll_node_coords"][2].tolist() == [6.0, 7.0, 8.0]
    assert m["gll_node_coords"][3].tolist() == [103.0, 104.0, 105.0]
    # file b's local node 1 (id 40) is merged node 3; its cell points there
    assert np.all(m["cell_gll_node_index"][1] == 3)
    assert np.array_equal(m["strain"][0, 0], np.ones(6))
    assert np.array_equal(m["strain"][0, 3], 3 * np.ones(6))


def _loop_merge(files, group, mass, n_model_cell):
    """main.cpp:306-426 as the C++ loop runs it, one node value at a time."""
    ids = {}
    l2g = []
    for f in files:
        m = []
        for gid in f["gll_node_ids"].tolist():
            m.append(ids.setdefault(gid, len(ids)))
        l2g.append(m)
    ng = len(ids)
    acc = [[0.0] * 6 for _ in range(ng)]
    wsum = [0.0] * ng
    for fi, f in enumerate(files):
        cgi = f["cell_gll_node_index"]
        rcm = f["recording_cell_model_index"]
        st = group[fi]["strain"]
        for c in range(cgi.shape[0]):
            for p in range(NPC):
                loc = int(cgi[c, p])
                if loc < 0 or loc >= len(l2g[fi]):
                    continue
                g = l2g[fi][loc]
                w = 1.0
                gc = int(rcm[c])
                if 0 <= gc < n_model_cell:
                    w = float(mass[gc, p])
                for k in range(6):
                    acc[g][k] = acc[g][k] + float(st[c, p, k]) * w
                wsum[g] += w
    out = np.zeros((ng, 6))
    for g in range(ng):
        if wsum[g] > 0.0:
            inv = 1.0 / wsum[g]
            out[g] = [a * inv for a in acc[g]]
        else:
            out[g] = acc[g]
    return out


def test_vector_merge_equals_the_sequential_loop_bitwise():
    rng = np.random.default_rng(3)
    files = []
    for k in range(2):
        n_cell = 3
        ids = rng.choice(400, size=200, replace=False) + 1
        cgi = rng.integers(-1, 200, size=(n_cell, NPC))
        files.append(_file(ids, rng.normal(size=(200, 3)), cgi, rcm=rng.integers(-1, 6, size=n_cell)))
    group = [{"strain": rng.normal(size=(3, NPC, 6)).astype(np.float32).astype(np.float64)} for _ in files]
    mass = rng.uniform(0.1, 5.0, size=(5, NPC))
    m = pp.merge_direction(files, [group], cell_mass=mass, n_model_cell=5)
    ref = _loop_merge(files, group, mass, 5)
    assert np.array_equal(m["strain"][0], ref)


def test_constant_field_stays_constant():
    rng = np.random.default_rng(5)
    cgi = rng.integers(0, 60, size=(4, NPC))
    f = _file(np.arange(60) + 1, np.zeros((60, 3)), cgi, rcm=[0, 1, 2, 3])
    val = np.array([1.5, -2.25, 3.0, 0.1, 7.0, -0.3])
    strain = np.broadcast_to(val, (4, NPC, 6)).copy()
    m = pp.merge_direction([f], [[{"strain": strain}]], cell_mass=rng.uniform(0.5, 2.0, (4, NPC)),
                           n_model_cell=4)
    seen = np.unique(cgi)
    np.testing.assert_allclose(m["strain"][0, seen], np.broadcast_to(val, (len(seen), 6)), rtol=4e-16 * 8,
                               atol=0)


def _grid_merged(nx, ny, h=1000.0):
    """One recording cell per element of an nx x ny surface grid; each cell's
    point 0 sits at its lower corner and point 124 at its upper corner."""
    ids, coords, cgi = [], [], []
    for j in range(ny):
        for i in range(nx):
            base = len(ids)
            row = np.arange(NPC) + base
            pts = np.zeros((NPC, 3))
            pts[0] = (i * h, j * h, 0.0)
            pts[124] = ((i + 1) * h, (j + 1) * h, 1000.0)
            ids.extend(range(base + 1, base + NPC + 1))
            coords.append(pts)
            cgi.append(row)
    return {"gll_node_ids": np.asarray(ids), "gll_node_coords": np.concatenate(coords),
            "cell_gll_node_index": np.asarray(cgi)}


def test_tiles_bin_cells_by_centre_over_the_interior_grid():
    cfg = {"nx_elements": 6, "ny_elements": 5, "pml_xmin": 1, "pml_ymin": 2,
           "tilex_elements": [1, 3], "tiley_elements": [2, 1]}
    merged = _grid_merged(6, 5)
    bins = pp.tile_cells(cfg, (0.0, 6000.0, 0.0, 5000.0), merged)
    # interior columns 1..4 -> tile x 0 (column 1), 1 (columns 2..4); column 5 is outside;
    # interior rows 2..4 -> tile y 0 (rows 2, 3), 1 (row 4); rows 0, 1 are outside
    assert list(bins) == [(0, 0), (0, 1), (1, 0), (1, 1)]
    cell = lambda i, j: j * 6 + i  # noqa: E731
    assert bins[(0, 0)] == [cell(1, 2), cell(1, 3)]
    assert bins[(0, 1)] == [cell(1, 4)]
    assert bins[(1, 0)] == [cell(2, 2), cell(3, 2), cell(4, 2), cell(2, 3), cell(3, 3), cell(4, 3)]
    assert bins[(1, 1)] == [cell(2, 4), cell(3, 4), cell(4, 4)]
    assert pp.tile_bounds(cfg, (0.0, 6000.0, 0.0, 5000.0), 1, 0) == (2000.0, 5000.0, 2000.0, 4000.0)
    assert pp.tile_bounds(cfg, (0.0, 6000.0, 0.0, 5000.0), 0, 1) == (1000.0, 2000.0, 4000.0, 5000.0)
    assert pp.find_tile_index(99, [1, 3]) == 1


def test_stf_is_downsampled_by_integer_stride_and_clamped():
    t = np.arange(10) * 0.5
    v = np.arange(10) * 2.0
    st, sv = pp.downsample_stf(t, v, 4)
    assert st.tolist() == [0.0, 1.0, 2.0, 3.0] and sv.tolist() == [0.0, 4.0, 8.0, 12.0]
    st, _ = pp.downsample_stf(t, v, 3)
    assert st.tolist() == [0.0, 1.5, 3.0]
    st, _ = pp.downsample_stf(t, v, 12)
    assert st.tolist() == [0.0, 0.5, 1.0, 1.5, 2.0, 2.5, 3.0, 3.5, 4.0, 4.5, 4.5, 4.5]


def test_tile_dict_carries_the_writer_schema():
    """The fields tests/test_postprocess_tile.cpp asserts, on a tile built here."""
    merged = _grid_merged(2, 2)
    n = len(merged["gll_node_ids"])
    for f in ("strain",):
        merged[f] = np.zeros((3, n, 6))
    merged["displacement"] = None
    merged["n_steps"] = 3
    cfg = {"nx_elements": 2, "ny_elements": 2, "pml_xmin": 0, "pml_ymin": 0, "tilex_elements": [2],
           "tiley_elements": [2], "output_dt_s": 0.25, "solver_dt": 0.05, "record_depth_max_m": 1000.0,
           "record_depth_actual_m": 1000.0}
    out = pp.tiles(cfg, (0.0, 2000.0, 0.0, 2000.0, 0.0, 1000.0), merged, merged, merged,
                   np.arange(15) * 0.05, np.ones(15), (1.0, 2.0, 3.0))
    t = out["tile_x000_y000.h5"]
    assert t["@source_xyz_m"].tolist() == [1.0, 2.0, 3.0]
    assert "strain" in t["@greens_quantities"]
    assert t["/mesh/gll_node_coords"].shape == (n, 3)
    assert t["/mesh/gll_node_ids"].ndim == 1
    g = t["/field/greens_tensor"]
    assert g.ndim == 4 and g.shape[2] == 6 and g.shape[3] == 3
    assert t["/time/t"].tolist() == [0.0, 0.25, 0.5]
    assert t["/source/stf_t"].tolist() == [0.0, 0.25, 0.5]
    assert math.isclose(t["/time@dt"], 0.05)
