"""Surrogate pieces that run without a GPU: the FP64 host restatement of the
GLL interpolation judge (greenfun/gll_interpolator.py:117-227), pinned by
what the mathematics fixes (exactness on tensor polynomials of degree <= 4
per axis, cardinality at nodes, partition of unity, clipping, the outside
error), and the CPU model of the MLP (oracle/mlp.py), pinned by central
finite differences."""
import os
import sys

import numpy as np
import pytest

import pygfx

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp as mlp_oracle  # noqa: E402


def tile_from_mesh(mesh, cells):
    """Unique nodes of the given cells and each cell's 125 indices into them."""
    ib = mesh.ibool.reshape(-1, 125)[np.asarray(cells)]
    nodes, inv = np.unique(ib.ravel(), return_inverse=True)
    xyz = mesh.node_xyz.reshape(-1, 3)[nodes]
    return nodes, xyz, inv.reshape(-1, 125).astype(np.int32)


@pytest.fixture(scope="module")
def box():
    m = pygfx.build_box(nx=4, ny=3, nz=3, lx=2400.0, ly=1800.0, lz=1800.0, pml_xmin=0, pml_xmax=0, pml_ymin=0,
                        pml_ymax=0, pml_zmax=0, output_dt=0.05, duration=0.2, record_depth_max=0.0)
    cells = np.arange(m.n_elem)
    nodes, xyz, cidx = tile_from_mesh(m, cells)
    return m, xyz, cidx


def poly(p):
    x, y, z = p[..., 0] / 1000.0, p[..., 1] / 1000.0, p[..., 2] / 1000.0
    return 1.0 + 0.3 * x ** 4 * y ** 3 - 2.0 * x * z ** 4 + 0.7 * y ** 4 * z ** 4 + x ** 2 * y * z ** 3


def test_judge_is_exact_on_tensor_polynomials_of_degree_four(box):
    m, xyz, cidx = box
    vals = np.stack([poly(xyz), 3.0 * np.ones(len(xyz))], axis=1)
    rng = np.random.default_rng(7)
    lo, hi = xyz.min(axis=0), xyz.max(axis=0)
    for _ in range(40):
        p = lo + rng.random(3) * (hi - lo)
        out = pygfx.gll_interpolate_host(xyz, cidx, vals, p)
        assert abs(out[0] - poly(p)) <= 1e-9 * max(1.0, abs(poly(p)))
        assert abs(out[1] - 3.0) <= 1e-12  # partition of unity (weights renormalised)


def test_judge_returns_node_values_at_nodes_and_rejects_outside(box):
    m, xyz, cidx = box
    rng = np.random.default_rng(1)
    vals = rng.standard_normal((len(xyz), 3))
    for k in rng.integers(0, len(xyz), 20):
        assert np.array_equal(pygfx.gll_interpolate_host(xyz, cidx, vals, xyz[k]), vals[k])
        # within the 1e-6 m exact-node tolerance
        assert np.array_equal(pygfx.gll_interpolate_host(xyz, cidx, vals, xyz[k] + 4e-7), vals[k])
    hi = xyz.max(axis=0)
    with pytest.raises(pygfx.GfxError, match="outside"):
        pygfx.gll_interpolate_host(xyz, cidx, vals, hi + np.array([1.0, 0.0, 0.0]))


def test_judge_clips_reference_coordinates_inside_the_slack(box):
    m, xyz, cidx = box
    vals = poly(xyz)[:, None]
    hi = xyz.max(axis=0)
    lo = xyz.min(axis=0)
    p = np.array([hi[0] + 5e-7, 0.5 * (lo[1] + hi[1]) + 37.0, 0.5 * (lo[2] + hi[2]) - 11.0])
    q = p.copy()
    q[0] = hi[0]  # clipped to the face
    assert np.allclose(pygfx.gll_interpolate_host(xyz, cidx, vals, p), pygfx.gll_interpolate_host(xyz, cidx, vals, q),
                       rtol=1e-12, atol=0)


def _tiny_model(emulate, seed=0):
    rng = np.random.default_rng(seed)
    nf, wd, dp = 3, 8, 2
    iw = 2 * nf
    n = 0
    for l in range(dp + 1):
        ow = 18 if l == dp else wd
        n += iw * ow + ow
        iw = ow
    w = rng.standard_normal(n) * 0.4
    B = rng.standard_normal((nf, 4))
    lo, hi = np.zeros(4), np.array([10.0, 10.0, 5.0, 2.0])
    scale = rng.uniform(0.5, 2.0, 18)
    return mlp_oracle.Model(w, B, nf, wd, dp, lo, hi, scale, emulate_bf16=emulate)


def test_cpu_mlp_gradient_matches_finite_differences():
    m = _tiny_model(False)
    rng = np.random.default_rng(3)
    x = rng.random((5, 4)) * np.array([10.0, 10.0, 5.0, 2.0])
    y = rng.standard_normal((5, 18))
    _, g = m.loss_grad(x, y)
    h = 1e-6
    for i in rng.choice(m.n_params, 60, replace=False):
        wp, wm = m.w.copy(), m.w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (m.loss_grad(x, y, wp)[0] - m.loss_grad(x, y, wm)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])


def test_bf16_rounding_is_nearest_even():
    # BF16 keeps 7 fraction bits: spacing 2^-7 at 1.0
    v = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -9, 1.0 + 2 ** -9, 1.0 + 3 * 2 ** -8, -2.5])
    r = mlp_oracle.bf16(v)
    assert r[0] == 1.0
    assert r[1] == 1.0  # exact tie -> even (fraction 0)
    assert r[2] == 1.0 + 2 ** -7  # above the tie
    assert r[3] == 1.0  # below the tie
    assert r[4] == 1.0 + 2 ** -6  # tie between 1+2^-7 (odd) and 1+2^-6 (even) -> even
    assert r[5] == -2.5
