"""The per-stage entry points driven in the order of the reference's device
branch (forward/share/src/solver.cpp:390-444):

    predict, [copy u~ out / back], gather_predicted, zero_residual,
    element_residual, pml_damping, source_injection (when stf != 0),
    scatter_to_rank, [copy r out / back], correct, [copy state out]

Gates:
  * GATHER assembly: the stage-driven run equals gfx_solver_run's bit for bit
    (state and every record snapshot), with and without C-PML, two fields;
  * COLORED / ATOMIC: the source is added after the element sum instead of
    inside it (one rounding apart; ATOMIC sums in no fixed order), so the two
    runs agree to 1e-5 rel L2 in strain, u and v, 1e-3 in a (the parity bars'
    ratio for the acceleration);
  * the stage-driven run against the FP64 oracle at the parity bar (1e-4);
  * exchange copies (u~ and r out to the host and back) leave the run
    unchanged, a residual changed between scatter and correct moves it, and
    out-of-order calls are refused without touching the state."""
import numpy as np
import pytest

import meshcases as mc
import pygfx
import pyoracle

pytestmark = pytest.mark.gpu

TOL = 1e-4
GFX_ERR_INVALID = 1


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if pygfx.device_count() < 1:
        pytest.fail("no HIP device visible; GPU tests must run on an MI355X box")


def _mesh(has_cpml):
    return mc.small_halfspace(nx=6, ny=6, nz=5, pml=2 if has_cpml else 1, duration=1.0)


def _sources(mesh):
    zc, zw, stf = mc.surface_centre_source(mesh)
    bc, bw, bstf = mc.buried_source(mesh)
    return [(zc, zw, stf, 2), (bc, bw, bstf, 0)]


def _solver(mesh, srcs, assembly, has_cpml):
    s = pygfx.Solver(mesh, n_fields=len(srcs), assembly=assembly, record_uva=True, has_cpml=has_cpml)
    for f, (c, w, stf, d) in enumerate(srcs):
        s.set_source(f, d, c, w, stf)
    return s


def _drive(s, mesh, srcs, nsteps, exchange_every=0):
    """The reference's device loop, stage by stage."""
    dt, beta, gamma = mesh.dt, 0.0, 0.5
    for step in range(nsteps):
        s.stage("newmark_predict", dt, beta)
        if exchange_every and step % exchange_every == 0:  # cuda_copy_utilde_to/from_host
            for f in range(len(srcs)):
                s.stage("copy_utilde_from_host", f, s.nodal_copy("utilde", f))
        s.stage("gather_predicted")
        s.stage("zero_residual")
        s.stage("element_residual", 5, mesh.n_elem)
        s.stage("pml_damping")
        for f, (c, w, stf, d) in enumerate(srcs):
            val = float(stf[step]) if step < len(stf) else 0.0
            if val != 0.0:  # solver.cpp:422-431
                s.stage("source_injection", f, d, val, np.asarray(w, np.float64), len(c))
        s.stage("scatter_to_rank")
        if exchange_every and step % exchange_every == 0:  # cuda_copy_residual_to/from_host
            for f in range(len(srcs)):
                s.stage("copy_residual_from_host", f, s.nodal_copy("residual", f))
        s.stage("newmark_correct", dt, beta, gamma)


def _snapshot(s, nf):
    st = [s.get_state(f) for f in range(nf)]
    rec = [np.stack([np.concatenate([x.ravel() for x in s.get_record(f, k, uva=True)])
                     for k in range(s.n_snapshots)]) for f in range(nf)]
    return st, rec


@pytest.mark.parametrize("has_cpml", [True, False])
def test_gather_stages_equal_fused_run_bitwise(has_cpml):
    mesh = _mesh(has_cpml)
    srcs = _sources(mesh)
    n = mesh.nsteps
    print(f"case: stages vs run, GATHER, cpml={has_cpml}, {mesh.n_elem} elements, {n} steps", flush=True)
    a = _solver(mesh, srcs, pygfx.GFX_ASSEMBLY_GATHER, has_cpml)
    a.run(n)
    b = _solver(mesh, srcs, pygfx.GFX_ASSEMBLY_GATHER, has_cpml)
    _drive(b, mesh, srcs, n, exchange_every=7)
    assert a.step == b.step == n and a.n_snapshots == b.n_snapshots > 0
    sa, ra = _snapshot(a, 2)
    sb, rb = _snapshot(b, 2)
    for f in range(2):
        for x, y in zip(sa[f], sb[f]):
            assert np.array_equal(x, y)
        assert np.array_equal(ra[f], rb[f])
    # the fused run continues from a stage-driven state
    a.run(5)
    b.run(5)
    assert all(np.array_equal(x, y) for x, y in zip(a.get_state(1), b.get_state(1)))


@pytest.mark.parametrize("assembly", [pygfx.GFX_ASSEMBLY_COLORED, pygfx.GFX_ASSEMBLY_ATOMIC])
def test_colored_atomic_stages_match_fused_run(assembly):
    mesh = _mesh(True)
    srcs = _sources(mesh)
    n = mesh.nsteps
    print(f"case: stages vs run, assembly={assembly}", flush=True)
    a = _solver(mesh, srcs, assembly, True)
    a.run(n)
    b = _solver(mesh, srcs, assembly, True)
    _drive(b, mesh, srcs, n)
    sa, ra = _snapshot(a, 2)
    sb, rb = _snapshot(b, 2)
    for f in range(2):
        n_strain = a.n_rec * 125 * 6
        assert mc.rel_l2(rb[f][:, :n_strain], ra[f][:, :n_strain]) < 1e-5  # strain records
        for x, y in zip(sb[f][:2], sa[f][:2]):  # u, v
            assert mc.rel_l2(x, y) < 1e-5
        assert mc.rel_l2(sb[f][2], sa[f][2]) < 1e-3  # a, the most cancellation-prone


def test_stage_driven_run_matches_oracle():
    mesh = _mesh(True)
    c, w, stf = mc.surface_centre_source(mesh)
    s = _solver(mesh, [(c, w, stf, 2)], pygfx.GFX_ASSEMBLY_GATHER, True)
    _drive(s, mesh, [(c, w, stf, 2)], mesh.nsteps)
    got = np.stack([s.get_record(0, k) for k in range(s.n_snapshots)])
    ref = pyoracle.Run(mesh.arrays, mesh.n_elem, 5, mesh.n_node, mesh.dt, mesh.nsteps, mesh.snapshot_stride, 2,
                       c, w, stf, mesh.rec_cell).run()
    err = mc.rel_l2(got, ref["strain"])
    print(f"stage-driven strain rel L2 vs oracle {err:.3e}")
    assert err <= TOL
    u, v, a = np.zeros(3 * mesh.n_node), np.zeros(3 * mesh.n_node), np.zeros(3 * mesh.n_node)
    s.stage("copy_state_to_host", 0, u, v, a)  # no-op check of the call itself
    uu, vv, aa = s.get_state(0)
    assert mc.rel_l2(uu, ref["u"]) <= TOL and mc.rel_l2(vv, ref["v"]) <= TOL


def test_residual_exchange_moves_the_run_and_order_is_enforced():
    mesh = _mesh(True)
    srcs = _sources(mesh)[:1]
    a = _solver(mesh, srcs, pygfx.GFX_ASSEMBLY_GATHER, True)
    b = _solver(mesh, srcs, pygfx.GFX_ASSEMBLY_GATHER, True)
    _drive(a, mesh, srcs, 30)
    _drive(b, mesh, srcs, 29)
    # step 29 of b with a neighbour's contribution added to r between scatter and correct
    dt = mesh.dt
    b.stage("newmark_predict", dt, 0.0)
    b.stage("element_residual", 5, mesh.n_elem)
    c, w, stf, d = srcs[0]
    b.stage("source_injection", 0, d, float(stf[29]), np.asarray(w, np.float64), len(c))
    b.stage("scatter_to_rank")
    r = b.nodal_copy("residual", 0)
    r_mod = r.copy()
    r_mod[3 * (mesh.n_node // 2) + 2] += 1e12
    b.stage("copy_residual_from_host", 0, r_mod)
    assert np.array_equal(b.nodal_copy("residual", 0), r_mod.astype(np.float32).astype(np.float64))
    b.stage("newmark_correct", dt, 0.0, 0.5)
    assert not np.array_equal(a.get_state(0)[2], b.get_state(0)[2])
    # out of order: correct / scatter / source before the element stage; run mid-step; bad arguments
    s = _solver(mesh, srcs, pygfx.GFX_ASSEMBLY_ATOMIC, True)
    u0 = s.get_state(0)
    assert s.stage_status("newmark_correct", dt, 0.0, 0.5) == GFX_ERR_INVALID
    assert s.stage_status("scatter_to_rank") == GFX_ERR_INVALID
    assert s.stage_status("newmark_predict", dt, 0.25) == GFX_ERR_INVALID  # implicit beta refused
    assert s.stage_status("newmark_predict", 2 * dt, 0.0) == GFX_ERR_INVALID  # another dt
    s.stage("newmark_predict", dt, 0.0)
    assert s.stage_status("element_residual", 5, mesh.n_elem) == GFX_ERR_INVALID  # ATOMIC: zero first
    assert s.stage_status("element_residual", 4, mesh.n_elem) == GFX_ERR_INVALID
    with pytest.raises(pygfx.GfxError):
        s.run(1)
    s.stage("zero_residual")
    s.stage("element_residual", 5, mesh.n_elem)
    assert s.stage_status("newmark_predict", dt, 0.0) == GFX_ERR_INVALID
    s.stage("scatter_to_rank")
    s.stage("newmark_correct", dt, 0.0, 0.5)
    s.run(1)
    assert s.step == 2
    assert all(np.all(np.isfinite(x)) for x in s.get_state(0)) and u0 is not None
