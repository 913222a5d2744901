"""GPU parity on the reference's edge cases, for every assembly mode:

  * GLL points with mu <= 0 contribute nothing to the element force
    (forward/elastic/src/element_cpu.cpp:42-43), whatever their lambda;
  * nodes with assembled mass <= 0 get a = 0 and keep u and v
    (forward/share/src/solver.cpp:57-61).

Each case is built so that breaking the rule moves the oracle's records far past
the gate (checked in the test), and the GPU must match the oracle within 1e-4."""
import numpy as np
import pytest

import meshcases as mc
import pygfx
import pyoracle

pytestmark = pytest.mark.gpu

TOL, TOL_ACC = 1e-4, 1e-3
ASSEMBLIES = [pygfx.GFX_ASSEMBLY_GATHER, pygfx.GFX_ASSEMBLY_COLORED, pygfx.GFX_ASSEMBLY_ATOMIC]


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if pygfx.device_count() < 1:
        pytest.fail("no HIP device visible; GPU tests must run on an MI355X box")


def _oracle(mesh, src, direction=2):
    cells, w, stf = src
    return pyoracle.Run(mesh.arrays, mesh.n_elem, 5, mesh.n_node, mesh.dt, mesh.nsteps, mesh.snapshot_stride,
                        direction, cells, w, stf, mesh.rec_cell).run(uva=True)


def _gpu(mesh, src, assembly, direction=2):
    cells, w, stf = src
    s = pygfx.Solver(mesh, n_fields=1, assembly=assembly, record_uva=True)
    s.set_source(0, direction, cells, w, stf)
    s.run(mesh.nsteps)
    n = s.n_snapshots
    return s, np.stack([s.get_record(0, k) for k in range(n)]).reshape(n, s.n_rec, 125, 6)


_CASES = {}


def _mu_case():
    """5 % of the points of a jittered layered mesh get mu = 0 or mu < 0 (lambda kept)."""
    if "mu" not in _CASES:
        base = mc.small_layered(duration=1.5)
        mu = base.mu_.copy()
        pts = np.random.default_rng(21).choice(mu.size, size=mu.size // 20, replace=False)
        half = len(pts) // 2
        mu[pts[:half]] = 0.0
        mu[pts[half:]] *= -1.0
        mesh = base.derive(mu_=mu)
        src = mc.buried_source(mesh)
        # not skipping them (mu -> tiny positive: lambda tr(eps) acts) moves the records
        unskipped = base.derive(mu_=np.where(mu <= 0, 1e-30, mu))
        sens = mc.rel_l2(_oracle(unskipped, src)["strain"], _oracle(mesh, src)["strain"])
        _CASES["mu"] = (mesh, src, _oracle(mesh, src), sens)
    return _CASES["mu"]


def _massless_case():
    """2 % of the nodes (seeded, plus some nodes of the source cells other than the
    forced ones) get zero mass."""
    if "mass" not in _CASES:
        base = mc.small_halfspace(duration=1.5)
        ib = base.ibool
        rng = np.random.default_rng(22)
        nodes = rng.choice(base.n_node, size=base.n_node // 50, replace=False)
        src = mc.surface_centre_source(base)
        cell_nodes = ib.reshape(-1, 125)[src[0]]
        forced = np.unique(cell_nodes[src[1].reshape(-1, 125) != 0])
        nodes = np.setdiff1d(np.union1d(nodes, cell_nodes.ravel()[::7]), forced)
        mass = np.where(np.isin(ib, nodes), 0.0, base.mass)
        mesh = base.derive(mass=mass)
        sens = mc.rel_l2(_oracle(base, src)["strain"], _oracle(mesh, src)["strain"])
        _CASES["mass"] = (mesh, src, _oracle(mesh, src), sens, nodes)
    return _CASES["mass"]


@pytest.mark.parametrize("assembly", ASSEMBLIES)
def test_mu_nonpositive_points_are_skipped(assembly):
    mesh, src, ref, sens = _mu_case()
    print(f"case: mu <= 0 at 5% of points, assembly={assembly}; rule worth {sens:.3e}", flush=True)
    assert sens > 100 * TOL
    s, st = _gpu(mesh, src, assembly)
    err = mc.rel_l2(st, ref["strain"])
    u, v, a = s.get_state(0)
    print(f"strain rel L2 {err:.3e}")
    assert err <= TOL
    assert mc.rel_l2(u, ref["u"]) <= TOL and mc.rel_l2(v, ref["v"]) <= TOL


@pytest.mark.parametrize("assembly", ASSEMBLIES)
def test_massless_nodes_keep_u_and_v(assembly):
    mesh, src, ref, sens, nodes = _massless_case()
    print(f"case: {len(nodes)} massless nodes, assembly={assembly}; rule worth {sens:.3e}", flush=True)
    assert sens > 100 * TOL
    s, st = _gpu(mesh, src, assembly)
    err = mc.rel_l2(st, ref["strain"])
    u, v, a = s.get_state(0)
    print(f"strain rel L2 {err:.3e}")
    assert err <= TOL
    assert mc.rel_l2(u, ref["u"]) <= TOL and mc.rel_l2(v, ref["v"]) <= TOL
    assert mc.rel_l2(a, ref["a"]) <= TOL_ACC
    assert np.all(a.reshape(-1, 3)[nodes] == 0.0)


@pytest.mark.parametrize("assembly", ASSEMBLIES)
def test_massless_nodes_from_a_moving_state(assembly):
    """From a random state, massless nodes keep their u and v (a = 0) step after step,
    the rest follows the oracle (gfo_run_state)."""
    mesh, src, _, _, nodes = _massless_case()
    n = 3 * mesh.n_node
    rng = np.random.default_rng(5)
    u0, v0, a0 = (rng.standard_normal(n) * sc for sc in (1e-3, 1e-2, 1.0))
    nsteps = 40
    run = pyoracle.Run(mesh.arrays, mesh.n_elem, 5, mesh.n_node, mesh.dt, nsteps, 0, 2, None, None, None, None)
    ru, rv, ra = run.run_state(u0, v0, a0)
    s = pygfx.Solver(mesh, n_fields=1, assembly=assembly, rec_cell=np.zeros(0, np.int32))
    s.set_state(0, u0, v0, a0)
    s.run(nsteps)
    u, v, a = s.get_state(0)
    print(f"moving state, assembly={assembly}: u {mc.rel_l2(u, ru):.3e} v {mc.rel_l2(v, rv):.3e}")
    assert mc.rel_l2(u, ru) <= TOL and mc.rel_l2(v, rv) <= TOL and mc.rel_l2(a, ra) <= TOL_ACC
    k = (3 * nodes[:, None] + np.arange(3)).ravel()
    assert np.allclose(u[k], u0[k], rtol=1e-6, atol=1e-6 * np.abs(u0).max())
    assert np.allclose(v[k], v0[k], rtol=1e-6, atol=1e-6 * np.abs(v0).max())
    assert np.all(a[k] == 0.0)
