"""GPU parity: the HIP solver (through the C ABI) against the FP64 CPU oracle.

Tolerances:
  * element plugin, FP64 device arithmetic: max rel diff < 1e-12, as the
    reference's own CUDA-vs-CPU test (tests/test_element_cuda.cu:119-135);
  * FP32 solver: aggregate rel L2 of recorded strain <= 1e-4 against the FP64
    oracle on the same inputs (the parity bar), final u, v <= 1e-4;
  * acceleration (a = M^-1 (f - K u), the most cancellation-prone output)
    <= 1e-3.
"""
import numpy as np
import pytest

import meshcases as mc
import pygfx
import pyoracle

pytestmark = pytest.mark.gpu

TOL = 1e-4
TOL_ACC = 1e-3


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if pygfx.device_count() < 1:
        pytest.fail("no HIP device visible; GPU tests must run on an MI355X box")


def _random_element(N, seed):
    """tests/test_element_cuda.cu:35-95 RandomElement, same glibc rand() seeds."""
    ng = N + 1
    nn = ng ** 3
    x, w, D = pyoracle.gll(N)
    dd = pyoracle.random_fill(nn * 9, seed)
    dd[0::9] += 3.0
    dd[4::9] += 3.0
    dd[8::9] += 3.0
    jac = np.abs(pyoracle.random_fill(nn, seed + 1)) + 0.01
    vp = np.abs(pyoracle.random_fill(nn, seed + 2)) + 1000.0
    vs = np.abs(pyoracle.random_fill(nn, seed + 3)) + 500.0
    rho = np.abs(pyoracle.random_fill(nn, seed + 4)) + 1000.0
    mu = rho * vs * vs
    lam = rho * (vp * vp - 2.0 * vs * vs)
    return dict(dd=dd, jac=jac, lam=lam, mu=mu, D=D, w=w, ng=ng, nn=nn)


@pytest.mark.parametrize("N,seed,useed", [(3, 42, 99), (5, 123, 456)])
def test_element_plugin_fp64_matches_cpu(N, seed, useed):
    el = _random_element(N, seed)
    u = pyoracle.random_fill(el["nn"] * 3, useed)
    r_cpu = np.zeros(el["nn"] * 3)
    pyoracle.element_residual(1, el["dd"], el["jac"], el["lam"], el["mu"], el["D"], el["w"], el["ng"], u, r_cpu)
    r_gpu = np.zeros(el["nn"] * 3)
    pygfx.element_residual(el["dd"], el["jac"], el["lam"], el["mu"], el["D"], el["w"], el["ng"], u, r_gpu)
    rel = np.abs(r_cpu - r_gpu) / np.maximum(np.abs(r_cpu), 1e-14)
    assert rel.max() < 1e-12


def test_element_plugin_accumulates_and_rigid_translation():
    el = _random_element(3, 77)
    u = np.tile([1.0, 2.0, 3.0], el["nn"])
    r = np.full(el["nn"] * 3, 5.0)
    pygfx.element_residual(el["dd"], el["jac"], el["lam"], el["mu"], el["D"], el["w"], el["ng"], u, r)
    assert np.abs(r - 5.0).max() < 1e-5  # tests/test_element_cuda.cu:168-187, r accumulates


def test_element_plugin_fp32_many_elements():
    mesh = mc.small_layered(nx=4, ny=4, nz=3, pml=1)
    E, nn = mesh.n_elem, 125
    rng = np.random.default_rng(3)
    u = rng.standard_normal(E * nn * 3)
    x, w, D = pyoracle.gll(4)
    ref = np.zeros(E * nn * 3)
    pyoracle.element_residual(E, mesh.dxi_dx, mesh.jacobian, mesh.lambda_, mesh.mu_, D, w, 5, u, ref)
    out = np.zeros(E * nn * 3)
    pygfx.element_residual(mesh.dxi_dx, mesh.jacobian, mesh.lambda_, mesh.mu_, D, w, 5, u, out, fp32=True)
    assert mc.rel_l2(out, ref) < 1e-5
    out64 = np.zeros(E * nn * 3)
    pygfx.element_residual(mesh.dxi_dx, mesh.jacobian, mesh.lambda_, mesh.mu_, D, w, 5, u, out64)
    assert mc.rel_l2(out64, ref) < 1e-13


def _oracle(mesh, cells, w, stf, direction, nsteps=None, has_cpml=True):
    nsteps = mesh.nsteps if nsteps is None else nsteps
    run = pyoracle.Run(mesh.arrays, mesh.n_elem, mesh.ngll, mesh.n_node, mesh.dt, nsteps,
                       mesh.snapshot_stride, direction, cells, w, stf, mesh.rec_cell,
                       has_cpml=has_cpml)
    return run.run(uva=True)


def _gpu(mesh, sources, nsteps=None, assembly=pygfx.GFX_ASSEMBLY_COLORED, epb=0, has_cpml=True,
         kernel=pygfx.GFX_ELEMENT_AUTO, node_groups=0, general_geometry=False):
    nsteps = mesh.nsteps if nsteps is None else nsteps
    s = pygfx.Solver(mesh, n_fields=len(sources), assembly=assembly, elems_per_block=epb,
                     record_uva=True, has_cpml=has_cpml, element_kernel=kernel,
                     node_field_groups=node_groups, general_geometry=general_geometry)
    for f, (cells, w, stf, d) in enumerate(sources):
        s.set_source(f, d, cells, w, stf)
    s.run(nsteps)
    return s


def _records(s, field, n_snap):
    nn = 125
    st = np.stack([s.get_record(field, k) for k in range(n_snap)]).reshape(n_snap, s.n_rec, nn, 6)
    return st


@pytest.mark.parametrize("assembly,kernel", [
    (pygfx.GFX_ASSEMBLY_GATHER, pygfx.GFX_ELEMENT_WAVE),
    (pygfx.GFX_ASSEMBLY_GATHER, pygfx.GFX_ELEMENT_BLOCK),
    (pygfx.GFX_ASSEMBLY_COLORED, pygfx.GFX_ELEMENT_WAVE),
    (pygfx.GFX_ASSEMBLY_COLORED, pygfx.GFX_ELEMENT_BLOCK),
    (pygfx.GFX_ASSEMBLY_ATOMIC, pygfx.GFX_ELEMENT_WAVE),
    (pygfx.GFX_ASSEMBLY_ATOMIC, pygfx.GFX_ELEMENT_BLOCK),
    (pygfx.GFX_ASSEMBLY_GATHER, pygfx.GFX_ELEMENT_PAIR),
    (pygfx.GFX_ASSEMBLY_COLORED, pygfx.GFX_ELEMENT_PAIR)])
def test_solver_matches_oracle_halfspace(assembly, kernel):
    mesh = mc.small_halfspace()
    cells, w, stf = mc.surface_centre_source(mesh)
    ref = _oracle(mesh, cells, w, stf, 2)
    s = _gpu(mesh, [(cells, w, stf, 2)], assembly=assembly, kernel=kernel)
    assert s.query("element_kernel") == kernel
    n_snap = ref["strain"].shape[0]
    assert s.n_snapshots == n_snap
    st = _records(s, 0, n_snap)
    err = mc.rel_l2(st, ref["strain"])
    print(f"halfspace assembly={assembly} kernel={kernel} strain rel L2 = {err:.3e}")
    assert err <= TOL
    u, v, a = s.get_state(0)
    assert mc.rel_l2(u, ref["u"]) <= TOL
    assert mc.rel_l2(v, ref["v"]) <= TOL
    assert mc.rel_l2(a, ref["a"]) <= TOL_ACC
    _, d, vel, acc = s.get_record(0, n_snap - 1, uva=True)
    assert mc.rel_l2(d, ref["disp"][-1]) <= TOL
    assert mc.rel_l2(vel, ref["vel"][-1]) <= TOL
    assert mc.rel_l2(acc, ref["acc"][-1]) <= TOL_ACC


@pytest.mark.parametrize("direction,kernel", [(0, pygfx.GFX_ELEMENT_WAVE), (1, pygfx.GFX_ELEMENT_BLOCK)])
def test_solver_matches_oracle_layered_jittered(direction, kernel):
    mesh = mc.small_layered()
    cells, w, stf = mc.buried_source(mesh)
    ref = _oracle(mesh, cells, w, stf, direction)
    s = _gpu(mesh, [(cells, w, stf, direction)], kernel=kernel)
    st = _records(s, 0, ref["strain"].shape[0])
    err = mc.rel_l2(st, ref["strain"])
    print(f"layered dir={direction} strain rel L2 = {err:.3e}")
    assert err <= TOL


def test_element_local_cpml_fallback_matches_oracle():
    """Inputs whose C-PML memory coefficients differ between elements at a shared
    node cannot use the nodal form; the solver falls back to element-local memory."""
    mesh = mc.small_halfspace(duration=1.5)
    arrays = mesh.arrays
    al = arrays["coef_alpha"].reshape(-1, 125, 9)
    ab = arrays["coef_abar"].reshape(-1, 125, 5)
    # a node shared by two elements that both use its x-direction memory (A3 != 0)
    use = (mesh.pml_region[:, None] != 0) & (ab[:, :, 2] != 0)
    nodes = mesh.ibool.reshape(-1, 125)
    ids, counts = np.unique(nodes[use], return_counts=True)
    shared = ids[counts >= 2][0]
    e, p = (int(x[0]) for x in np.nonzero(use & (nodes == shared)))
    al[e, p, 0] *= 0.999  # its x-direction alpha in one of the elements
    cells, w, stf = mc.surface_centre_source(mesh)
    ref = _oracle(mesh, cells, w, stf, 2)
    s = pygfx.Solver(mesh, n_fields=1, record_uva=True)
    assert s.query("pml_nodal") == 0.0
    s.set_source(0, 2, cells, w, stf)
    s.run(mesh.nsteps)
    st = _records(s, 0, ref["strain"].shape[0])
    err = mc.rel_l2(st, ref["strain"])
    print(f"element-local C-PML strain rel L2 = {err:.3e}")
    assert err <= TOL
    s2 = pygfx.Solver(mc.small_halfspace(duration=1.5), n_fields=1)
    assert s2.query("pml_nodal") == 1.0 and s2.query("colors") == 8


def test_solver_legacy_damping_path():
    mesh = mc.small_halfspace(duration=1.5)
    cells, w, stf = mc.surface_centre_source(mesh)
    ref = _oracle(mesh, cells, w, stf, 2, has_cpml=False)
    s = _gpu(mesh, [(cells, w, stf, 2)], has_cpml=False)
    st = _records(s, 0, ref["strain"].shape[0])
    assert mc.rel_l2(st, ref["strain"]) <= TOL
    u, v, a = s.get_state(0)
    assert mc.rel_l2(u, ref["u"]) <= TOL


@pytest.mark.parametrize("assembly,kernel", [
    (pygfx.GFX_ASSEMBLY_COLORED, pygfx.GFX_ELEMENT_WAVE),
    (pygfx.GFX_ASSEMBLY_COLORED, pygfx.GFX_ELEMENT_BLOCK),
    (pygfx.GFX_ASSEMBLY_GATHER, pygfx.GFX_ELEMENT_WAVE),
    (pygfx.GFX_ASSEMBLY_GATHER, pygfx.GFX_ELEMENT_BLOCK)])
def test_deterministic_assembly_is_bitwise_reproducible_and_epb_invariant(assembly, kernel):
    mesh = mc.small_layered(duration=0.6)
    cells, w, stf = mc.buried_source(mesh)
    runs = []
    for epb in (1, 2, 4, 8, 2):
        s = _gpu(mesh, [(cells, w, stf, 2)], epb=epb, kernel=kernel, assembly=assembly)
        runs.append(s.get_state(0))
    for r in runs[1:]:
        for x, y in zip(r, runs[0]):
            assert np.array_equal(x, y)


@pytest.mark.parametrize("assembly", [pygfx.GFX_ASSEMBLY_COLORED, pygfx.GFX_ASSEMBLY_GATHER])
def test_batched_fields_equal_single_runs(assembly):
    mesh = mc.small_halfspace(duration=1.0)
    cells, w, stf = mc.surface_centre_source(mesh)
    c2, w2, stf2 = mc.buried_source(mesh, frac=(0.45, 0.4, 0.0), depth=1500.0, f0=1.2)
    sources = [(cells, w, stf, 0), (cells, w, stf, 1), (cells, w, stf, 2), (c2, w2, stf2, 2)]
    sb = _gpu(mesh, sources, assembly=assembly)
    for f, src in enumerate(sources):
        s1 = _gpu(mesh, [src], assembly=assembly)
        for x, y in zip(sb.get_state(f), s1.get_state(0)):
            assert np.array_equal(x, y)
        n_snap = s1.n_snapshots
        assert np.array_equal(_records(sb, f, n_snap), _records(s1, 0, n_snap))


@pytest.mark.parametrize("assembly,nf", [(pygfx.GFX_ASSEMBLY_COLORED, 3), (pygfx.GFX_ASSEMBLY_GATHER, 4)])
def test_pair_kernel_matches_block_kernel_and_oracle(assembly, nf):
    """Two fields per pass (odd and even field counts) against one per pass,
    and field 0 against the oracle."""
    mesh = mc.small_layered(duration=0.8)
    cells, w, stf = mc.buried_source(mesh)
    c2, w2, stf2 = mc.buried_source(mesh, frac=(0.4, 0.6, 0.0), depth=900.0)
    srcs = [(cells, w, stf, 2), (c2, w2, stf2, 0), (cells, w, stf, 1), (c2, w2, stf2, 2)][:nf]
    a = _gpu(mesh, srcs, assembly=assembly, kernel=pygfx.GFX_ELEMENT_PAIR, epb=2)
    b = _gpu(mesh, srcs, assembly=assembly, kernel=pygfx.GFX_ELEMENT_BLOCK, epb=2)
    assert int(a.query("element_kernel")) == pygfx.GFX_ELEMENT_PAIR
    for f in range(nf):
        for x, y in zip(a.get_state(f), b.get_state(f)):
            assert mc.rel_l2(x, y) <= 1e-6
    ref = _oracle(mesh, cells, w, stf, 2)
    got = np.stack([a.get_record(0, k) for k in range(a.n_snapshots)])
    assert mc.rel_l2(got, ref["strain"][:a.n_snapshots]) <= TOL


@pytest.mark.parametrize("assembly,nf", [(pygfx.GFX_ASSEMBLY_GATHER, 3), (pygfx.GFX_ASSEMBLY_COLORED, 2),
                                         (pygfx.GFX_ASSEMBLY_ATOMIC, 1)])
def test_pencil_kernel_matches_block_kernel_and_oracle(assembly, nf):
    """Lane-per-pencil element kernel (k contractions in registers, ten
    elements per workgroup, last workgroup partly empty) against the
    lane-per-point block kernel, and field 0 against the oracle.  Both run the
    same per-point operation sequence, so with deterministic assembly the
    states agree to the last bit; atomic assembly (float atomics, order
    varies from run to run) only to rounding: 1e-5 rel L2, the acceleration
    being the most cancellation-prone of u, v, a."""
    mesh = mc.small_layered(duration=0.8)
    assert mesh.n_elem % 10 != 0  # a partial last workgroup is exercised
    cells, w, stf = mc.buried_source(mesh)
    c2, w2, stf2 = mc.buried_source(mesh, frac=(0.4, 0.6, 0.0), depth=900.0)
    srcs = [(cells, w, stf, 2), (c2, w2, stf2, 0), (cells, w, stf, 1)][:nf]
    a = _gpu(mesh, srcs, assembly=assembly, kernel=pygfx.GFX_ELEMENT_PENCIL)
    b = _gpu(mesh, srcs, assembly=assembly, kernel=pygfx.GFX_ELEMENT_BLOCK, epb=2)
    assert int(a.query("element_kernel")) == pygfx.GFX_ELEMENT_PENCIL
    assert int(a.query("elems_per_block")) == 10
    for f in range(nf):
        for x, y in zip(a.get_state(f), b.get_state(f)):
            if assembly == pygfx.GFX_ASSEMBLY_ATOMIC:
                assert mc.rel_l2(x, y) <= 1e-5
            else:
                assert np.array_equal(x, y)
    ref = _oracle(mesh, cells, w, stf, 2)
    got = np.stack([a.get_record(0, k) for k in range(a.n_snapshots)])
    err = mc.rel_l2(got, ref["strain"][:a.n_snapshots])
    print(f"pencil assembly={assembly} strain rel L2 = {err:.3e}")
    assert err <= TOL


@pytest.mark.parametrize("assembly,kernel", [(pygfx.GFX_ASSEMBLY_COLORED, pygfx.GFX_ELEMENT_BLOCK),
                                             (pygfx.GFX_ASSEMBLY_GATHER, pygfx.GFX_ELEMENT_WAVE),
                                             (pygfx.GFX_ASSEMBLY_GATHER, pygfx.GFX_ELEMENT_PENCIL)])
def test_axis_aligned_elements_take_the_diagonal_form(assembly, kernel):
    """The halfspace mesh is axis-aligned (dxi_dx diagonal up to FP64 round-off);
    a fifth of its elements get a real off-diagonal term, so both the diagonal
    and the general form run in one step (split ranges, per colour too).  The
    oracle gets the same inputs.  Leaving out the round-off-sized off-diagonal
    products changes FP32 rounding, which the explicit time loop carries on, so
    the two forms are not bitwise equal: each meets the parity bars against the
    oracle, and they agree with each other to the same bars."""
    mesh = mc.small_halfspace(duration=0.8)
    dd = mesh.arrays["dxi_dx"].reshape(-1, 125, 9)
    rng = np.random.default_rng(12)
    pick = rng.choice(mesh.n_elem, mesh.n_elem // 5, replace=False)
    dd[pick, :, 1] = 1e-6 * dd[pick, :, 0]
    cells, w, stf = mc.surface_centre_source(mesh)
    ref = _oracle(mesh, cells, w, stf, 2)
    s = _gpu(mesh, [(cells, w, stf, 2)], assembly=assembly, kernel=kernel)
    g = _gpu(mesh, [(cells, w, stf, 2)], assembly=assembly, kernel=kernel, general_geometry=True)
    assert int(s.query("diag_elements")) == mesh.n_elem - len(pick)
    assert int(g.query("diag_elements")) == 0
    n_snap = ref["strain"].shape[0]
    err = mc.rel_l2(_records(s, 0, n_snap), ref["strain"])
    same = all(np.array_equal(x, y) for x, y in zip(s.get_state(0), g.get_state(0)))
    print(f"diagonal form assembly={assembly} kernel={kernel}: strain rel L2 = {err:.3e}, "
          f"identical to the general form: {same}")
    assert err <= TOL
    assert mc.rel_l2(_records(g, 0, n_snap), ref["strain"]) <= TOL
    for run in (s, g):
        u, v, a = run.get_state(0)
        assert mc.rel_l2(u, ref["u"]) <= TOL and mc.rel_l2(v, ref["v"]) <= TOL
        assert mc.rel_l2(a, ref["a"]) <= TOL_ACC
    for x, y, bar in zip(s.get_state(0), g.get_state(0), (TOL, TOL, TOL_ACC)):
        assert mc.rel_l2(x, y) <= bar


@pytest.mark.parametrize("assembly", [pygfx.GFX_ASSEMBLY_COLORED, pygfx.GFX_ASSEMBLY_GATHER])
def test_node_field_groups_do_not_change_results(assembly):
    """Spreading the fields of a node block over 1, 2, 4 or 8 lane groups only
    changes which lane does the arithmetic, never the arithmetic."""
    mesh = mc.small_halfspace(duration=0.8)
    cells, w, stf = mc.surface_centre_source(mesh)
    c2, w2, stf2 = mc.buried_source(mesh, frac=(0.45, 0.4, 0.0), depth=1500.0, f0=1.2)
    sources = [(cells, w, stf, 0), (cells, w, stf, 1), (cells, w, stf, 2), (c2, w2, stf2, 2), (c2, w2, stf2, 0)]
    base = _gpu(mesh, sources, assembly=assembly, node_groups=1)
    n_snap = base.n_snapshots
    for fg in (2, 4, 8):
        s = _gpu(mesh, sources, assembly=assembly, node_groups=fg)
        assert int(s.query("node_field_groups")) == fg
        for f in range(len(sources)):
            for x, y in zip(s.get_state(f), base.get_state(f)):
                assert np.array_equal(x, y)
            assert np.array_equal(_records(s, f, n_snap), _records(base, f, n_snap))


@pytest.mark.parametrize("assembly", [pygfx.GFX_ASSEMBLY_COLORED, pygfx.GFX_ASSEMBLY_GATHER])
def test_set_state_get_state_roundtrip_and_continuation(assembly):
    """Loading a state and continuing equals running through (restart semantics)."""
    mesh = mc.small_halfspace(duration=1.0)
    cells, w, stf = mc.surface_centre_source(mesh)
    ref = _oracle(mesh, cells, w, stf, 2)
    s = _gpu(mesh, [(cells, w, stf, 2)], nsteps=40, assembly=assembly)
    u, v, a = s.get_state(0)
    s2 = pygfx.Solver(mesh, n_fields=1, assembly=assembly)
    s2.set_state(0, u, v, a)
    u2, v2, a2 = s2.get_state(0)
    for x, y in ((u, u2), (v, v2), (a, a2)):
        assert mc.rel_l2(y, x) < 1e-6
    # continue from step 40 with the source shifted in time
    s2.set_source(0, 2, cells, w, stf[40:])
    s2.run(mesh.nsteps - 40)
    uf = s2.get_state(0)[0]
    assert mc.rel_l2(uf, ref["u"]) <= TOL


@pytest.mark.parametrize("assembly", [pygfx.GFX_ASSEMBLY_COLORED, pygfx.GFX_ASSEMBLY_GATHER])
def test_reset_and_restart_state_roundtrip(assembly):
    mesh = mc.small_halfspace(duration=1.0)
    cells, w, stf = mc.surface_centre_source(mesh)
    s = _gpu(mesh, [(cells, w, stf, 2)], nsteps=30, assembly=assembly)
    u1, v1, a1 = s.get_state(0)
    s.reset()
    assert s.step == 0 and s.n_snapshots == 0
    assert np.abs(s.get_state(0)[0]).max() == 0.0
    s.run(30)
    for x, y in zip(s.get_state(0), (u1, v1, a1)):
        assert np.array_equal(x, y)


def test_autotuned_configuration_matches_oracle():
    mesh = mc.small_layered(duration=1.0)
    cells, w, stf = mc.buried_source(mesh)
    pick = pygfx.autotune(mesh, n_fields=2, trial_steps=2)
    assert pick["assembly"] in (pygfx.GFX_ASSEMBLY_COLORED, pygfx.GFX_ASSEMBLY_GATHER)
    assert pick["element_kernel"] in (pygfx.GFX_ELEMENT_BLOCK, pygfx.GFX_ELEMENT_WAVE, pygfx.GFX_ELEMENT_PAIR,
                                      pygfx.GFX_ELEMENT_PENCIL)
    assert pick["elems_per_block"] in (1, 2, 4, 8) and pick["ms_per_step"] > 0
    ref = _oracle(mesh, cells, w, stf, 1)
    s = pygfx.Solver(mesh, n_fields=2, assembly=pick["assembly"], element_kernel=pick["element_kernel"],
                     elems_per_block=pick["elems_per_block"])
    s.set_source(0, 1, cells, w, stf)
    s.set_source(1, 1, cells, w, stf)
    s.run(mesh.nsteps)
    st = _records(s, 1, ref["strain"].shape[0])
    assert mc.rel_l2(st, ref["strain"]) <= TOL
