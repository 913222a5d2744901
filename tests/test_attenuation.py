"""Shear attenuation by standard linear solids: the fit and the oracle's update.

The reference's solver is elastic (its attenuation is a planned extension with
per-point tau_sigma / tau_epsilon inputs), so nothing here compares with
reference output.  The oracle's restatement is pinned by the mathematics:
- the fitted solids' quality factor, evaluated from the complex modulus
  M(w) = M_R (1 + sum_l (te_l/ts_l - 1) i w ts_l / (1 + i w ts_l)), stays
  within a few percent of the target over the band;
- the oracle's stress response to a strain ramp (one step) then hold equals
  the analytic hereditary integral sigma = 2 int M(t - s) de(s) of the
  solids' relaxation function M(t) = M_R (1 + sum_l (te_l/ts_l - 1)
  exp(-t/ts_l)), at every step, to round-off (the update is exact for strain
  linear in time);
- solids with te = ts leave the elastic oracle run unchanged."""
import numpy as np
import pytest

import meshcases as mc
import pygfx
import pyoracle


def q_of(ts, te, f):
    w = 2 * np.pi * np.asarray(f)[:, None]
    m = 1 + np.sum((te / ts - 1) * 1j * w * ts / (1 + 1j * w * ts), axis=1)
    return m.real / m.imag


@pytest.mark.parametrize("q,band,n", [(20.0, (0.1, 5.0), 3), (50.0, (0.05, 2.0), 3), (200.0, (0.5, 10.0), 4),
                                      (80.0, (0.2, 1.0), 2)])
def test_fitted_solids_hold_q_over_the_band(q, band, n):
    ts, te = pygfx.sls_fit(q, band[0], band[1], n)
    assert np.all(te > ts) and np.all(ts > 0)
    np.testing.assert_allclose(ts, 1 / (2 * np.pi * np.geomspace(band[0], band[1], n)), rtol=1e-12)
    qf = q_of(ts, te, np.geomspace(band[0], band[1], 400))
    assert np.max(np.abs(qf / q - 1)) < 0.06, (qf.min(), qf.max())


def test_infinite_q_fits_elastic_solids():
    ts, te = pygfx.sls_fit(0.0, 0.1, 5.0, 3)
    assert np.array_equal(ts, te)
    with pytest.raises(pygfx.GfxError, match="band"):
        pygfx.sls_fit(50.0, 5.0, 0.1, 3)


def test_oracle_update_is_the_hereditary_integral():
    ts, te = pygfx.sls_fit(15.0, 0.2, 8.0, 3)
    mu_r, dt, n = 3.0e10, 0.004, 5000
    eps1 = np.array([1e-6, -2e-6, 3e-7, 5e-7, -4e-7])
    dev = np.zeros((n, 5))
    dev[1:] = eps1  # ramp from 0 to eps1 over the first step, then held
    sig = pyoracle.sls_point_history(ts, te, mu_r, dt, dev)
    t = dt * np.arange(1, n)
    r = te / ts - 1
    # 2 (eps1/dt) int_0^dt M(t - s) ds
    integ = mu_r * dt + mu_r * np.sum(r * ts * (np.exp(-(t[:, None] - dt) / ts) - np.exp(-t[:, None] / ts)), axis=1)
    want = 2 * (integ / dt)[:, None] * eps1[None, :]
    np.testing.assert_allclose(sig, want, rtol=1e-10, atol=0)
    # relaxed at the end: 2 mu_R eps1
    np.testing.assert_allclose(sig[-1], 2 * mu_r * eps1, rtol=1e-4)


def test_step_coefficients():
    a, b, c, y = pyoracle.sls_coefficients(0.05, 0.06, 2.0, 0.01)
    assert np.isclose(a, np.exp(-0.2))
    assert np.isclose(b + c, 1 - a)  # constant strain relaxes to Y eps
    assert np.isclose(y, 2 * 2.0 * (0.06 / 0.05 - 1))


def test_oracle_elastic_solids_equal_the_elastic_oracle():
    mesh = mc.small_halfspace(nx=6, ny=6, nz=4, pml=1, duration=0.6)
    cells, w, stf = mc.surface_centre_source(mesh)
    P = mesh.n_elem * 125
    ts = np.tile(pygfx.sls_fit(0.0, 0.1, 5.0, 3)[0], P)
    args = (mesh.arrays, mesh.n_elem, 5, mesh.n_node, mesh.dt, mesh.nsteps, mesh.snapshot_stride, 2, cells, w, stf,
            mesh.rec_cell)
    a = pyoracle.Run(*args).run()
    b = pyoracle.Run(*args, tau_sigma=ts, tau_epsilon=ts).run()
    for k in ("strain", "u", "v", "a"):
        np.testing.assert_allclose(b[k], a[k], rtol=0, atol=1e-12 * np.abs(a[k]).max(), err_msg=k)


def test_box_generator_assigns_solids_per_layer():
    m = pygfx.build_box(nx=4, ny=4, nz=4, lx=4000.0, ly=4000.0, lz=4000.0, order=4,
                        layers=[(2000.0, 5000.0, 2900.0, 2500.0), (1e9, 6200.0, 3600.0, 2800.0)],
                        n_sls=3, q_mu=[30.0, 0.0], att_fmin=0.1, att_fmax=5.0, pml_xmin=1, pml_xmax=1, pml_ymin=1,
                        pml_ymax=1, pml_zmax=1, duration=0.2)
    assert m.n_sls == 3
    ts = m.tau_sigma.reshape(-1, 3)
    te = m.tau_epsilon.reshape(-1, 3)
    z = m.node_xyz.reshape(-1, 3)[m.ibool, 2]
    top, deep = z < 2000.0, z >= 2000.0
    fit30 = pygfx.sls_fit(30.0, 0.1, 5.0, 3)
    assert np.allclose(ts[top], fit30[0]) and np.allclose(te[top], fit30[1])
    assert np.array_equal(ts[deep], te[deep])


@pytest.mark.skipif(not pygfx.h5_available(), reason="no HDF5 library to bind")
def test_solids_round_trip_through_the_partitions(tmp_path):
    m = mc.small_halfspace(nx=6, ny=6, nz=4, pml=1, duration=0.5, n_sls=3, q_mu=[30.0], att_fmin=0.1,
                           att_fmax=4.0)
    stf = pygfx.ricker(1.0, 0.5, 1e20, m.dt, m.nsteps)
    d = str(tmp_path)
    pygfx.write_inputs(m, d, (3000.0, 3000.0, 0.0), stf, n_parts=2)
    lay = pygfx.h5_layout(str(tmp_path / "partitions" / "partition_0.h5"))
    assert lay["/field/cell/tau_sigma"] == ("f64", (72, 5, 5, 5, 3))
    back = pygfx.read_inputs(d)
    assert back.n_sls == 3
    assert np.array_equal(back.tau_sigma, m.tau_sigma) and np.array_equal(back.tau_epsilon, m.tau_epsilon)
