"""Structured-box input generator (restatement of the reference preprocessor's
rules for regular boxes) against the known answers of the reference's Python
tests and closed forms."""
import numpy as np
import pytest

import pygfx


def test_unit_cube_jacobian_and_mass():
    # tests/preprocess/test_gll_geometry.py:89-97: unit cube, N=1: J = 1/8, mass sum = 1
    m = pygfx.build_box(nx=1, ny=1, nz=1, lx=1.0, ly=1.0, lz=1.0, order=1,
                        layers=[(0.0, 2.0, 1.0, 1.0)], pml_xmin=0, pml_xmax=0, pml_ymin=0,
                        pml_ymax=0, pml_zmax=0, record_depth_max=0.0, output_dt=0.01, duration=0.1)
    assert np.allclose(m.jacobian, 0.125, atol=1e-12)
    assert abs(m.mass.sum() - 1.0) < 1e-12
    assert np.allclose(m.coords.reshape(-1, 3)[0], 0.0)
    assert np.allclose(m.coords.reshape(-1, 3)[-1], 1.0)
    assert np.allclose(m.dxi_dx.reshape(-1, 9)[0], [2, 0, 0, 0, 2, 0, 0, 0, 2])


def test_default_box_sizes_cfl_and_counts():
    m = pygfx.build_box()  # 20 x 20 x 10 hexes of 1 km, order 4, C-PML 3
    assert m.n_elem == 4000
    assert m.n_node == 81 * 81 * 41  # unique lattice nodes (test_partition.py:472-488)
    assert len(np.unique(m.ibool)) == m.n_node
    # h_min = (1 - sqrt(3/7)) / 2 km; dt = 0.05 / stride <= 0.5 h_min / vp (cfl_validator.py)
    h = 1000.0 * (1 - np.sqrt(3 / 7)) / 2
    assert abs(m.h_min - h) < 1e-9
    assert m.snapshot_stride == 4 and abs(m.dt - 0.0125) < 1e-15 and m.nsteps == 800
    assert m.n_pml_elem == 4000 - 14 * 14 * 7
    assert len(m.rec_cell) == 14 * 14 * 2 and m.record_depth_actual == 2000.0
    assert abs(m.mass.sum() - 20e3 * 20e3 * 10e3 * 2700) < 1e-6 * 1.08e16
    # lambda / mu from vp, vs, rho (cli.py:369-381)
    assert np.allclose(m.mu_, 2700 * 3464.0 ** 2)
    assert np.allclose(m.lambda_, 2700 * (6000.0 ** 2 - 2 * 3464.0 ** 2))


def test_surface_vertex_source_splits_over_four_cells():
    # tests/preprocess/test_source_locator.py:197-205
    m = pygfx.build_box()
    cells, w = m.locate_source(10000.0, 10000.0, 0.0)
    assert len(cells) == 4
    w = w.reshape(4, -1)
    assert abs(w.sum() - 1.0) < 1e-12
    for c in range(4):
        nz = np.flatnonzero(np.abs(w[c]) > 1e-14)
        assert len(nz) == 1 and abs(w[c, nz[0]] - 0.25) < 1e-14
        # that corner sits at the source position
        g = m.ibool[cells[c] * 125 + nz[0]]
        assert np.allclose(m.node_xyz[3 * g: 3 * g + 3], [10000.0, 10000.0, 0.0])


def test_buried_interior_source_weights_interpolate_linear_fields():
    m = pygfx.build_box(jitter=0.1, jitter_seed=3)
    p = np.array([9321.0, 10777.0, 1234.0])
    cells, w = m.locate_source(*p)
    assert len(cells) == 1
    xyz = m.node_xyz.reshape(-1, 3)[m.ibool[cells[0] * 125: cells[0] * 125 + 125]]
    assert abs(w.sum() - 1.0) < 1e-12
    # trilinear elements: Lagrange weights reproduce the coordinates of the point
    assert np.allclose(w @ xyz, p, atol=1e-6)


def test_ricker_values():
    # tests/preprocess/test_stf_evaluator.py:49-66: peak 1 at t0, samples at i*dt
    v = pygfx.ricker(5.0, 0.3, 1.0, 1e-4, 3001)
    assert abs(v[3000] - 1.0) < 1e-12
    t = np.arange(3001) * 1e-4
    a = np.pi * 5.0 * (t - 0.3)
    assert np.allclose(v, (1 - 2 * a * a) * np.exp(-a * a), atol=1e-15)


def test_pml_regions_and_coefficients():
    m = pygfx.build_box()
    reg = m.pml_region.reshape(10, 20, 20)  # [ez][ey][ex]
    assert reg[0, 10, 10] == 0  # interior
    assert reg[0, 10, 0] == 1 and reg[0, 10, 19] == 1  # x only
    assert reg[0, 0, 10] == 2  # y only
    assert reg[9, 10, 10] == 3  # z only (bottom)
    assert reg[0, 0, 0] == 4 and reg[9, 10, 0] == 5 and reg[9, 0, 10] == 6 and reg[9, 0, 0] == 7
    # interior cells: alpha = d = 0 -> coef (1, dt/2, dt/2), abar = 0
    al = m.coef_alpha.reshape(4000, 125, 9)
    ab = m.coef_abar.reshape(4000, 125, 5)
    e = (0 * 20 + 10) * 20 + 10
    assert np.allclose(al[e], np.tile([1.0, m.dt / 2, m.dt / 2], 3))
    assert np.abs(ab[e]).max() == 0.0
    # recursive-convolution identity a1 + a2 = (1 - a0)/b for b = alpha > 0
    e = (0 * 20 + 10) * 20 + 19  # x-only PML cell on the xmax side
    a0, a1, a2 = al[e, :, 0], al[e, :, 1], al[e, :, 2]
    b = -np.log(a0) / m.dt
    ok = b > 1e-3
    assert np.allclose((a1 + a2)[ok], ((1 - a0) / b)[ok], rtol=1e-10)


def test_pml_alpha_consistent_at_shared_nodes():
    """Where a memory variable is used (A_{3+d} != 0), every element sharing the
    node carries the same coef_alpha for direction d -- the property that lets
    the device solver keep C-PML memory per node."""
    m = pygfx.build_box(nx=12, ny=10, nz=8, lx=12e3, ly=10e3, lz=8e3)
    al = m.coef_alpha.reshape(-1, 9)
    ab = m.coef_abar.reshape(-1, 5)
    reg = np.repeat(m.pml_region, 125)
    for d in range(3):
        use = (reg != 0) & (ab[:, 2 + d] != 0)
        nodes = m.ibool[use]
        vals = al[use, 3 * d: 3 * d + 3]
        order = np.argsort(nodes, kind="stable")
        n_s, v_s = nodes[order], vals[order]
        first = np.r_[True, n_s[1:] != n_s[:-1]]
        idx = np.cumsum(first) - 1
        ref = v_s[first][idx]
        assert np.array_equal(v_s, ref)


def test_recording_depth_snaps_to_cell_faces():
    # recording_map.py:52-88: 2.0 km request on 0.6 km cells snaps to 2.4 km
    m = pygfx.build_box(nx=12, ny=12, nz=8, lx=7200.0, ly=7200.0, lz=4800.0, pml_xmin=2,
                        pml_xmax=2, pml_ymin=2, pml_ymax=2, pml_zmax=2)
    assert abs(m.record_depth_actual - 2400.0) < 1e-9
    assert len(m.rec_cell) == 8 * 8 * 4


def test_layered_material_and_random_field():
    m = pygfx.build_box(nx=10, ny=10, nz=10, lx=6000.0, ly=6000.0, lz=6000.0,
                        layers=[(3000.0, 5000.0, 2900.0, 2500.0), (0.0, 6200.0, 3600.0, 2800.0)],
                        rf_amp=0.05, rf_corr_len=2000.0, rf_modes=64, rf_seed=2)
    z = m.coords.reshape(-1, 3)[:, 2]
    rho = m.density
    assert np.all(rho[z < 3000.0] == 2500.0) and np.all(rho[z >= 3000.0] == 2800.0)
    ratio = m.vp / np.where(z < 3000.0, 5000.0, 6200.0)
    assert 0.75 < ratio.min() < 1.0 < ratio.max() < 1.25
    assert np.allclose(m.vs / np.where(z < 3000.0, 2900.0, 3600.0), ratio)
    # shared nodes carry identical material in every element
    g = m.ibool
    order = np.argsort(g, kind="stable")
    gs, vs = g[order], m.vp[order]
    first = np.r_[True, gs[1:] != gs[:-1]]
    assert np.array_equal(vs, vs[first][np.cumsum(first) - 1])


def test_invalid_specs_raise():
    with pytest.raises(pygfx.GfxError):
        pygfx.build_box(nx=0)
    with pytest.raises(pygfx.GfxError):
        pygfx.build_box(pml_xmin=15, pml_xmax=15)
    with pytest.raises(pygfx.GfxError):
        pygfx.build_box(output_dt=10.0)  # no CFL stride in 1..100
