"""Pins the FP64 CPU oracle (oracle/gfx_oracle.c) before it is trusted as the
parity checker.

Sources of truth, per oracle function:
  * the reference's own C++ unit tests (known answers and invariants):
    tests/test_gll.cpp, tests/test_element.cpp, tests/test_assembly.cpp,
    tests/test_newmark.cpp;
  * closed forms: an element stiffness matrix built independently in Voigt
    B-matrix form with numpy, the exact discrete eigenmode cos(n theta) of the
    central-difference step, momentum balance of the point source, and the
    steady state u/b of the C-PML memory recurrence.
"""
import numpy as np
import pytest

import pyoracle

# ---------------------------------------------------------------- GLL (tests/test_gll.cpp)


def test_gll_endpoints_and_known_values():
    x, w, D = pyoracle.gll(1)
    assert abs(x[0] + 1) < 1e-15 and abs(x[1] - 1) < 1e-15  # test_gll.cpp:11-16
    x, w, D = pyoracle.gll(2)
    assert np.allclose(x, [-1, 0, 1], atol=1e-15)  # test_gll.cpp:18-24
    assert np.allclose(w, [1 / 3, 4 / 3, 1 / 3], atol=1e-15)  # test_gll.cpp:50-58
    x, w, D = pyoracle.gll(3)
    assert abs(x[0] + x[3]) < 1e-15 and abs(x[1] + x[2]) < 1e-15  # test_gll.cpp:26-36
    assert np.all(np.diff(x) > 0)
    x, w, D = pyoracle.gll(5)
    for i in range(3):
        assert abs(x[i] + x[5 - i]) < 1e-14  # test_gll.cpp:38-48
    # N = 4 closed form: 0, +-sqrt(3/7); weights 1/10, 49/90, 32/45
    x, w, D = pyoracle.gll(4)
    assert np.allclose(x, [-1, -np.sqrt(3 / 7), 0, np.sqrt(3 / 7), 1], atol=1e-15)
    assert np.allclose(w, [0.1, 49 / 90, 32 / 45, 49 / 90, 0.1], atol=1e-15)


@pytest.mark.parametrize("N", [1, 2, 3, 4, 5])
def test_gll_weights_sum_to_two(N):
    assert abs(pyoracle.gll(N)[1].sum() - 2.0) < 1e-14  # test_gll.cpp:60-69


@pytest.mark.parametrize("N", [2, 3, 4])
def test_derivative_matrix_rows_and_polynomials(N):
    x, w, D = pyoracle.gll(N)
    D = D.reshape(N + 1, N + 1)
    assert np.abs(D.sum(axis=1)).max() < 1e-13  # test_gll.cpp:71-85
    # exact for every polynomial of degree <= N (test_gll.cpp:87-104 does x^2 at N=3)
    for deg in range(N + 1):
        assert np.allclose(D @ x ** deg, deg * x ** max(deg - 1, 0) if deg else 0 * x, atol=1e-12)


@pytest.mark.parametrize("N", [2, 3, 4])
def test_lagrange_partition_of_unity_and_cardinality(N):
    x, _, _ = pyoracle.gll(N)
    for xi in np.linspace(-1, 1, 10):
        assert abs(pyoracle.lagrange(xi, x).sum() - 1.0) < 1e-14  # test_gll.cpp:106-120
    for i in range(N + 1):
        assert np.allclose(pyoracle.lagrange(x[i], x), np.eye(N + 1)[i], atol=1e-15)  # :122-137


# ------------------------------------------------ element residual (tests/test_element.cpp)


def unit_cube(N):
    """tests/test_element.cpp:15-76: [0,1]^3, dxi_dx = diag(2), J = 1/8."""
    ng = N + 1
    nn = ng ** 3
    x, w, D = pyoracle.gll(N)
    coords = np.zeros((nn, 3))
    for i in range(ng):
        for j in range(ng):
            for k in range(ng):
                coords[(i * ng + j) * ng + k] = 0.5 * (np.array([x[i], x[j], x[k]]) + 1)
    dd = np.tile([2.0, 0, 0, 0, 2.0, 0, 0, 0, 2.0], nn)
    jac = np.full(nn, 0.125)
    vp, vs, rho = 3000.0, 1500.0, 2500.0
    mu = np.full(nn, rho * vs * vs)
    lam = np.full(nn, rho * (vp * vp - 2 * vs * vs))
    return dict(ng=ng, nn=nn, coords=coords, dd=dd, jac=jac, lam=lam, mu=mu, D=D, w=w, x=x)


def residual(el, u, n_elem=1):
    r = np.zeros(u.size)
    pyoracle.element_residual(n_elem, el["dd"], el["jac"], el["lam"], el["mu"], el["D"], el["w"],
                              el["ng"], np.ascontiguousarray(u), r)
    return r


def test_rigid_translation_and_rotation_zero():
    el = unit_cube(3)
    u = np.tile([1.0, 2.0, 3.0], el["nn"])
    assert np.abs(residual(el, u)).max() < 1e-5  # test_element.cpp:78-100
    th = 1e-6
    u = np.zeros((el["nn"], 3))
    u[:, 0] = -th * el["coords"][:, 1]
    u[:, 1] = th * el["coords"][:, 0]
    assert np.abs(residual(el, u.ravel())).max() < 1e-6  # test_element.cpp:102-125


def test_uniaxial_strain_force_balance():
    el = unit_cube(3)
    u = np.zeros((el["nn"], 3))
    u[:, 0] = 1e-3 * el["coords"][:, 0]
    r = residual(el, u.ravel()).reshape(-1, 3)
    assert abs(r[:, 0].sum()) < 1e-8  # test_element.cpp:127-156


def _lagrange_deriv_matrix(nodes):
    """D[i][j] = l_j'(x_i) from numpy polynomials (independent of gll.hpp's formula)."""
    n = len(nodes)
    D = np.zeros((n, n))
    for j in range(n):
        others = np.delete(nodes, j)
        poly = np.poly1d(others, r=True) / np.prod(nodes[j] - others)
        D[:, j] = np.polyder(poly)(nodes)
    return D


def _voigt_stiffness(el):
    """Element stiffness in Voigt form: K = sum_q w_q J_q B_q^T C_q B_q."""
    ng, nn = el["ng"], el["nn"]
    x, wq = el["x"], el["w"]
    Dl = _lagrange_deriv_matrix(x)
    K = np.zeros((3 * nn, 3 * nn))
    for qi in range(ng):
        for qj in range(ng):
            for qk in range(ng):
                q = (qi * ng + qj) * ng + qk
                dd = el["dd"][9 * q: 9 * q + 9].reshape(3, 3)  # dd[m][p] = dxi_p/dx_m
                if el["mu"][q] <= 0:
                    continue
                # dN_a/dxi_p at quadrature point q for a = (i,j,k)
                dN = np.zeros((nn, 3))
                for a in range(nn):
                    i, j, k = a // (ng * ng), (a // ng) % ng, a % ng
                    dN[a, 0] = Dl[qi, i] * (j == qj) * (k == qk)
                    dN[a, 1] = Dl[qj, j] * (i == qi) * (k == qk)
                    dN[a, 2] = Dl[qk, k] * (i == qi) * (j == qj)
                gx = dN @ dd.T  # gx[a, m] = dN_a/dx_m
                B = np.zeros((6, 3 * nn))
                B[0, 0::3] = gx[:, 0]
                B[1, 1::3] = gx[:, 1]
                B[2, 2::3] = gx[:, 2]
                B[3, 0::3] = gx[:, 1]
                B[3, 1::3] = gx[:, 0]  # engineering shear xy
                B[4, 0::3] = gx[:, 2]
                B[4, 2::3] = gx[:, 0]
                B[5, 1::3] = gx[:, 2]
                B[5, 2::3] = gx[:, 1]
                lam, mu = el["lam"][q], el["mu"][q]
                Cm = np.zeros((6, 6))
                Cm[:3, :3] = lam
                Cm[np.arange(3), np.arange(3)] += 2 * mu
                Cm[np.arange(3, 6), np.arange(3, 6)] = mu
                K += wq[qi] * wq[qj] * wq[qk] * el["jac"][q] * (B.T @ Cm @ B)
    return K


def _random_element(N, seed):
    """tests/test_element_cuda.cu:35-95 (same glibc rand() seeds)."""
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
    return dict(ng=ng, nn=nn, dd=dd, jac=jac, lam=rho * (vp * vp - 2 * vs * vs), mu=rho * vs * vs,
                D=D, w=w, x=x)


@pytest.mark.parametrize("N,seed,useed", [(3, 42, 99), (4, 7, 8)])
def test_element_residual_equals_independent_voigt_stiffness(N, seed, useed):
    el = _random_element(N, seed)
    K = _voigt_stiffness(el)
    assert np.allclose(K, K.T, rtol=0, atol=1e-9 * np.abs(K).max())  # symmetric
    u = pyoracle.random_fill(el["nn"] * 3, useed)
    r = residual(el, u)
    ref = -K @ u
    assert np.abs(r - ref).max() < 1e-11 * np.abs(ref).max()


def test_element_stiffness_has_six_rigid_modes():
    el = unit_cube(2)
    ng, nn = el["ng"], el["nn"]
    K = np.zeros((3 * nn, 3 * nn))
    for col in range(3 * nn):
        e = np.zeros(3 * nn)
        e[col] = 1.0
        K[:, col] = -residual(el, e)
    ev = np.linalg.eigvalsh(0.5 * (K + K.T))
    assert np.abs(ev[:6]).max() < 1e-8 * ev[-1]
    assert ev[6] > 1e-6 * ev[-1]


def test_mu_nonpositive_point_is_skipped():
    el = unit_cube(2)
    el["mu"] = el["mu"].copy()
    el["mu"][:] = 0.0
    u = pyoracle.random_fill(el["nn"] * 3, 5)
    assert np.abs(residual(el, u)).max() == 0.0  # element_cpu.cpp:42-43


# ------------------------------------------------ assembly (tests/test_assembly.cpp:144-251)


def test_scatter_accumulates_shared_nodes():
    ib = np.array([0, 1, 1, 2], np.int32)
    loc = np.array([10, 20, 30, 40, 50, 60, 100, 200, 300, 400, 500, 600], float)
    glob = pyoracle.scatter(loc, ib, 2, 2, 3)
    assert np.array_equal(glob, [10, 20, 30, 140, 250, 360, 400, 500, 600])


def test_gather_copies_and_round_trip():
    ib = np.array([0, 1, 1, 2], np.int32)
    glob = np.arange(1.0, 10.0)
    loc = pyoracle.gather(glob, ib, 2, 2)
    assert np.array_equal(loc, [1, 2, 3, 4, 5, 6, 4, 5, 6, 7, 8, 9])
    ib3 = np.array([0, 1, 0], np.int32)
    loc = np.array([100, 101, 102, 200, 201, 202, 300, 301, 302], float)
    assert np.array_equal(pyoracle.scatter(loc, ib3, 3, 1, 2), [400, 402, 404, 200, 201, 202])


# ------------------------------------------------ Newmark (tests/test_newmark.cpp)


def test_newmark_predictor_and_corrector_closed_forms():
    import ctypes as C
    dt = 1e-3
    u, v, a = np.zeros(9), np.zeros(9), np.full(9, 2.0)
    ut = np.zeros(9)
    pyoracle.lib.gfo_newmark_predict(9, dt, 0.0, pyoracle.P(u), pyoracle.P(v), pyoracle.P(a), pyoracle.P(ut))
    assert np.allclose(ut, 0.5 * dt * dt * 2.0, atol=1e-15)  # test_newmark.cpp:14-45
    mass = np.full(2, 10.0)
    u, v, a, r = np.full(6, 0.5), np.full(6, 0.1), np.zeros(6), np.full(6, 2.0)
    pyoracle.lib.gfo_newmark_correct(6, dt, 0.0, 0.5, pyoracle.P(mass), pyoracle.P(u), pyoracle.P(v),
                                     pyoracle.P(a), pyoracle.P(r))
    assert np.allclose(a, 0.2, atol=1e-15)  # test_newmark.cpp:47-77
    assert np.allclose(v, 0.1 + 0.5 * dt * 0.2, atol=1e-15)
    assert np.allclose(u, 0.5 + dt * 0.1, atol=1e-15)  # u += dt v_old + dt^2/2 a_old (a_old = 0)
    # massless node: a = 0, u and v unchanged (solver.cpp:57-61)
    mass = np.array([0.0])
    u, v, a, r = np.ones(3), np.ones(3), np.ones(3), np.ones(3)
    pyoracle.lib.gfo_newmark_correct(3, dt, 0.0, 0.5, pyoracle.P(mass), pyoracle.P(u), pyoracle.P(v),
                                     pyoracle.P(a), pyoracle.P(r))
    assert np.array_equal(u, np.ones(3)) and np.array_equal(v, np.ones(3)) and np.array_equal(a, np.zeros(3))


# ------------------------------------------------ whole step: exact discrete eigenmode


def _free_mesh(nx=2, ny=1, nz=1):
    import pygfx
    return pygfx.build_box(nx=nx, ny=ny, nz=nz, lx=1000.0 * nx, ly=1000.0 * ny, lz=1000.0 * nz,
                           order=2, layers=[(0.0, 3000.0, 1700.0, 2400.0)], pml_xmin=0, pml_xmax=0,
                           pml_ymin=0, pml_ymax=0, pml_zmax=0, jitter=0.0, record_depth_max=0.0,
                           duration=1.0)


def test_step_reproduces_exact_discrete_eigenmode():
    """u_n = cos(n theta) phi with cos(theta) = 1 - w^2 dt^2 / 2 for Newmark
    beta = 0, gamma = 1/2 started from (phi, 0, -w^2 phi)."""
    import scipy.linalg as sl
    m = _free_mesh()
    ng = m.ngll
    nn = ng ** 3
    x, w, D = pyoracle.gll(ng - 1)
    G = m.n_node
    K = np.zeros((3 * G, 3 * G))
    for col in range(3 * G):
        ug = np.zeros(3 * G)
        ug[col] = 1.0
        loc = pyoracle.gather(ug, m.ibool, m.n_elem, nn)
        rl = np.zeros_like(loc)
        pyoracle.element_residual(m.n_elem, m.dxi_dx, m.jacobian, m.lambda_, m.mu_, D, w, ng, loc, rl)
        K[:, col] = -pyoracle.scatter(rl, m.ibool, m.n_elem, nn, G)
    mass = np.zeros(G)
    np.add.at(mass, m.ibool, m.mass)
    M = np.repeat(mass, 3)
    ev, vec = sl.eigh(0.5 * (K + K.T), np.diag(M))
    mode = 10  # first non-rigid modes are 6..; take a higher one
    om2, phi = ev[mode], vec[:, mode]
    dt = 0.2 * 2.0 / np.sqrt(ev[-1])  # well inside the stability limit
    nsteps = 37
    arrays = dict(m.arrays)
    arrays["damping"] = np.zeros_like(m.damping)
    run = pyoracle.Run(arrays, m.n_elem, ng, G, dt, nsteps, 0, 2, None, None, None, None, has_cpml=False)
    u, v, a = run.run_state(phi, np.zeros_like(phi), -om2 * phi)
    theta = np.arccos(1.0 - 0.5 * om2 * dt * dt)
    assert np.allclose(u, np.cos(nsteps * theta) * phi, atol=1e-9 * np.abs(phi).max())
    # v_n = (u_{n+1} - u_{n-1}) / (2 dt) = -sin(n theta) sin(theta) / dt * phi
    assert np.allclose(v, -np.sin(nsteps * theta) * np.sin(theta) / dt * phi,
                       atol=1e-8 * np.abs(phi).max() / dt)


def test_source_momentum_balance():
    """Free body, no damping: sum_g m_g a_g = F(t_n) in the force direction at every step
    (solver.cpp:576-605: stf_values[step] = f(step dt), weights normalised to 1)."""
    import pygfx
    m = _free_mesh(2, 2, 1)
    cells, wts = m.locate_source(1000.0, 1000.0, 0.0)
    assert len(cells) == 4
    arrays = dict(m.arrays)
    arrays["damping"] = np.zeros_like(m.damping)
    stf = np.array([0.0, 3.0, -1.5, 7.0, 2.0])
    mass = np.zeros(m.n_node)
    np.add.at(mass, m.ibool, m.mass)
    for direction in range(3):
        for n in range(1, len(stf) + 1):
            run = pyoracle.Run(arrays, m.n_elem, m.ngll, m.n_node, 1e-4, n, 0, direction, cells, wts,
                               stf, None, has_cpml=False)
            z = np.zeros(3 * m.n_node)
            u, v, a = run.run_state(z, z, z)
            mom = (mass[:, None] * a.reshape(-1, 3)).sum(axis=0)
            want = np.zeros(3)
            want[direction] = stf[n - 1]
            assert np.allclose(mom, want, atol=1e-9 * max(1.0, np.abs(stf).max()))


# ------------------------------------------------ C-PML (forward/share/src/pml.cpp)


def test_cpml_memory_steady_state_is_u_over_b():
    """For constant u (v = a = 0) the recurrence m = a0 m + a1 new + a2 old, with the
    convolution coefficients of pml_cpml.py:257-304 for rate b, converges to u / b:
    the exponential-kernel convolution of a constant.  Distinct b per direction and
    distinct u per component pin the [comp*3 + dir] and [dir*3 + k] layouts."""
    import ctypes as C
    ng = 2
    nn = ng ** 3
    dt = 0.01
    b = np.array([3.0, 11.0, 0.7])
    alpha = np.zeros((nn, 9))
    for d in range(3):
        t = np.exp(-0.5 * b[d] * dt)
        c1 = (1 - t) / b[d]
        alpha[:, d * 3: d * 3 + 3] = [t * t, c1, c1 * t]
    ubar = np.array([1.0, -2.0, 5.0])
    u = np.tile(ubar, nn)
    zero = np.zeros(3 * nn)
    old, new, mem = np.zeros(3 * nn), np.zeros(3 * nn), np.zeros(9 * nn)
    region = np.array([7], np.int32)
    ib = np.arange(nn, dtype=np.int32)
    abar = np.zeros(5 * nn)
    ones = np.ones(nn)
    p = pyoracle.Cpml(1, ng, pyoracle.P(region, C.c_int32), pyoracle.P(np.ascontiguousarray(alpha.ravel())),
                      pyoracle.P(abar), pyoracle.P(ones), pyoracle.P(ones), pyoracle.P(ib, C.c_int32),
                      pyoracle.P(old), pyoracle.P(new), pyoracle.P(mem))
    keep = (alpha, abar, ones, ib, region)  # noqa: F841
    for _ in range(4000):
        pyoracle.lib.gfo_cpml_update_displ_fields(C.byref(p), pyoracle.P(u), pyoracle.P(zero), pyoracle.P(zero), dt)
        pyoracle.lib.gfo_cpml_update_displ_memory(C.byref(p))
    got = np.ctypeslib.as_array(p.rmemory, shape=(9 * nn,)).reshape(nn, 3, 3)  # [node][comp][dir]
    want = ubar[:, None] / b[None, :]
    assert np.allclose(got, want[None], rtol=1e-9)


def test_cpml_accel_term_is_weighted_sum():
    """r += w_i w_j w_k J / rho (A1 v + A2 u + A3 mx + A4 my + A5 mz), added per component
    (pml.cpp:151-216); checked with one non-zero coefficient at a time."""
    import ctypes as C
    ng = 2
    nn = ng ** 3
    x, w, D = pyoracle.gll(1)
    region = np.array([1], np.int32)
    ib = np.arange(nn, dtype=np.int32)
    rho = np.full(nn, 2.0)
    jac = np.full(nn, 3.0)
    u = np.tile([1.0, 2.0, 3.0], nn)
    v = np.tile([10.0, 20.0, 30.0], nn)
    mem = np.tile([[100.0, 200.0, 300.0], [400.0, 500.0, 600.0], [700.0, 800.0, 900.0]], (nn, 1, 1)).ravel()
    alpha = np.zeros(9 * nn)
    for kcoef, expect in ((0, v[:3]), (1, u[:3]), (2, mem[0:9:3]), (3, mem[1:9:3]), (4, mem[2:9:3])):
        abar = np.zeros((nn, 5))
        abar[:, kcoef] = 1.0
        abar = abar.ravel()
        old, new = np.zeros(3 * nn), np.zeros(3 * nn)
        memc = mem.copy()
        p = pyoracle.Cpml(1, ng, pyoracle.P(region, C.c_int32), pyoracle.P(alpha), pyoracle.P(abar),
                          pyoracle.P(rho), pyoracle.P(jac), pyoracle.P(ib, C.c_int32), pyoracle.P(old),
                          pyoracle.P(new), pyoracle.P(memc))
        r = np.zeros(3 * nn)
        pyoracle.lib.gfo_cpml_accel_contribution(C.byref(p), pyoracle.P(u), pyoracle.P(v), pyoracle.P(w), pyoracle.P(r))
        scale = 1.0 * 1.0 * 1.0 * 3.0 / 2.0  # N = 1: w = 1
        assert np.allclose(r.reshape(nn, 3), scale * np.asarray(expect)[None, :])


# ------------------------------------- legacy ramp damping (forward/share/src/pml.cpp:13-29)


def _damp(d, v):
    v = np.ascontiguousarray(v, np.float64).copy()
    pyoracle.lib.gfo_apply_pml_damping(pyoracle.P(np.ascontiguousarray(d, np.float64)), pyoracle.P(v), v.size)
    return v


def test_legacy_damping_reference_cases():
    """The four cases of the reference's tests/test_pml.cpp:11-101 (NGLL 4, one or two
    elements, damping per node shared by its 3 DOFs)."""
    n_node = 2 * 4 ** 3  # test_pml.cpp:11-28: zero damping leaves v unchanged
    assert np.all(_damp(np.zeros(n_node), np.ones(3 * n_node)) == 1.0)
    n_node = 4 ** 3  # :30-56: d = 0.5 on the first half halves v there, the rest stays 2
    d = np.zeros(n_node)
    d[: n_node // 2] = 0.5
    v = _damp(d, np.full(3 * n_node, 2.0))
    assert np.all(v[: 3 * (n_node // 2)] == 1.0) and np.all(v[3 * (n_node // 2):] == 2.0)
    assert np.all(_damp(np.ones(n_node), np.full(3 * n_node, 3.0)) == 0.0)  # :58-75
    d = np.zeros(n_node)  # :77-101: damping is per node, not per DOF
    d[0] = 0.8
    v = _damp(d, np.ones(3 * n_node))
    assert np.allclose(v[:3], 0.2, atol=1e-15) and np.all(v[3:] == 1.0)


def test_legacy_damping_whole_run_closed_form():
    """Without C-PML, solver.cpp:563-574 damps v after the predictor and before the
    corrector, whose u update then uses the damped v.  A rigid translation with v0 and a
    uniform d has K u = 0, so v_n = (1-d)^n v0 and u_n = dt v0 sum_{k=1..n} (1-d)^k.
    (Damping after the corrector would give u_1 = dt v0; damping before the predictor
    would give a u that leads v by one factor.)"""
    m = _free_mesh(2, 1, 1)
    arrays = dict(m.arrays)
    d = 0.05
    arrays["damping"] = np.full_like(m.damping, d)
    dt, n = 1e-3, 7
    run = pyoracle.Run(arrays, m.n_elem, m.ngll, m.n_node, dt, n, 0, 2, None, None, None, None,
                       has_cpml=False)
    v0 = np.array([0.3, -1.2, 2.0])
    u, v, a = run.run_state(np.zeros(3 * m.n_node), np.tile(v0, m.n_node), np.zeros(3 * m.n_node))
    q = 1.0 - d
    assert np.allclose(v.reshape(-1, 3), q ** n * v0, rtol=1e-12, atol=1e-12)
    want_u = dt * v0 * sum(q ** k for k in range(1, n + 1))
    assert np.allclose(u.reshape(-1, 3), want_u, rtol=1e-11, atol=1e-15)
    assert np.abs(a).max() < 1e-9 * np.abs(v0).max() / dt


# ------------------------------------- C-PML step by hand (forward/share/src/pml.cpp:55-216)


def _cpml_state(ng, n_elem, region, alpha, abar, rho=None, jac=None):
    import ctypes as C
    nn = ng ** 3
    keep = dict(region=np.asarray(region, np.int32), alpha=np.ascontiguousarray(alpha, np.float64).ravel(),
                abar=np.ascontiguousarray(abar, np.float64).ravel(),
                rho=np.ones(n_elem * nn) if rho is None else rho, jac=np.ones(n_elem * nn) if jac is None else jac,
                ib=np.arange(n_elem * nn, dtype=np.int32), old=np.zeros(3 * n_elem * nn),
                new=np.zeros(3 * n_elem * nn), mem=np.zeros(9 * n_elem * nn))
    p = pyoracle.Cpml(n_elem, ng, pyoracle.P(keep["region"], C.c_int32), pyoracle.P(keep["alpha"]),
                      pyoracle.P(keep["abar"]), pyoracle.P(keep["rho"]), pyoracle.P(keep["jac"]),
                      pyoracle.P(keep["ib"], C.c_int32), pyoracle.P(keep["old"]), pyoracle.P(keep["new"]),
                      pyoracle.P(keep["mem"]))
    return p, keep


def test_cpml_displ_fields_three_steps_known_answer():
    """pml.cpp:55-110 over three calls with v, a != 0: old <- previous new (the swap),
    new = u + c1 v, old += c2 a, with c1 = (1 - 2 theta) dt / 2 and c2 = (1 - theta) dt^2 / 2,
    theta = 1/8.  dt = 0.1 gives c1 = 0.0375, c2 = 0.004375.  Expected values worked by hand:
      call 1  u (1, 2, 3)     v (10, 20, 30)  a (100, 200, 300)
              new = (1.375, 2.75, 4.125), old = 0 + c2 a = (0.4375, 0.875, 1.3125)
      call 2  u (-1, .5, 2)   v (4, -8, 0)    a (-40, 16, 8)
              old = (1.375, 2.75, 4.125) + c2 a = (1.2, 2.82, 4.16), new = (-0.85, 0.2, 2)
      call 3  u (2, -2, 1)    v (-16, 16, 8)  a (80, -160, 0)
              old = (-0.85, 0.2, 2) + c2 a = (-0.5, -0.5, 2), new = (1.4, -1.4, 1.3)
    Element 1 is interior (region 0) and is never touched."""
    import ctypes as C
    ng, nn = 2, 8
    p, keep = _cpml_state(ng, 2, [3, 0], np.zeros(2 * nn * 9), np.zeros(2 * nn * 5))
    calls = [((1, 2, 3), (10, 20, 30), (100, 200, 300), (1.375, 2.75, 4.125), (0.4375, 0.875, 1.3125)),
             ((-1, .5, 2), (4, -8, 0), (-40, 16, 8), (-0.85, 0.2, 2), (1.2, 2.82, 4.16)),
             ((2, -2, 1), (-16, 16, 8), (80, -160, 0), (1.4, -1.4, 1.3), (-0.5, -0.5, 2))]
    for u, v, a, want_new, want_old in calls:
        uu, vv, aa = (np.tile(np.asarray(x, np.float64), 2 * nn) for x in (u, v, a))
        pyoracle.lib.gfo_cpml_update_displ_fields(C.byref(p), pyoracle.P(uu), pyoracle.P(vv), pyoracle.P(aa), 0.1)
        new = np.ctypeslib.as_array(p.displ_new, shape=(2 * nn * 3,)).reshape(2, nn, 3)
        old = np.ctypeslib.as_array(p.displ_old, shape=(2 * nn * 3,)).reshape(2, nn, 3)
        assert np.allclose(new[0], want_new, rtol=0, atol=1e-14)
        assert np.allclose(old[0], want_old, rtol=0, atol=1e-14)
        assert np.all(new[1] == 0.0) and np.all(old[1] == 0.0)


def test_cpml_memory_update_known_answer():
    """pml.cpp:112-149: m[c*3+d] = alpha[d*3+0] m + alpha[d*3+1] new_c + alpha[d*3+2] old_c,
    with distinct coefficients per direction, non-zero memory and distinct components.
    alpha_x = (.5, .25, .125), alpha_y = (.9, .1, .05), alpha_z = (0, 1, -1);
    m[c][d] = 10 c + d + 1; new = (1.4, -1.4, 1.3); old = (-0.5, -0.5, 2).  By hand:
      c = 0: x .5*1 + .25*1.4 - .125*.5 = 0.7875,  y .9*2 + .1*1.4 - .05*.5 = 1.915,   z 1.4 + .5 = 1.9
      c = 1: x 5.5 - .35 - .0625 = 5.0875,         y 10.8 - .14 - .025 = 10.635,      z -1.4 + .5 = -0.9
      c = 2: x 10.5 + .325 + .25 = 11.075,         y 19.8 + .13 + .1 = 20.03,         z 1.3 - 2 = -0.7"""
    import ctypes as C
    ng, nn = 2, 8
    alpha = np.tile([.5, .25, .125, .9, .1, .05, 0.0, 1.0, -1.0], nn)
    p, keep = _cpml_state(ng, 1, [7], alpha, np.zeros(nn * 5))
    keep["mem"][:] = np.tile([1, 2, 3, 11, 12, 13, 21, 22, 23], nn)
    keep["new"][:] = np.tile([1.4, -1.4, 1.3], nn)
    keep["old"][:] = np.tile([-0.5, -0.5, 2.0], nn)
    pyoracle.lib.gfo_cpml_update_displ_memory(C.byref(p))
    want = [0.7875, 1.915, 1.9, 5.0875, 10.635, -0.9, 11.075, 20.03, -0.7]
    assert np.allclose(keep["mem"].reshape(nn, 9), want, rtol=0, atol=1e-13)


def _one_element_n1(abar, alpha, region=7):
    """One N = 1 element on [-1, 1]^3 with rho = J = 1 (GLL weights 1): every point is its
    own node with mass 1, so a rigid translation has K u = 0 exactly and each node's C-PML
    acceleration is abar . (v, u, m_x, m_y, m_z)."""
    nn = 8
    eye = np.tile(np.eye(3).ravel(), nn)
    return dict(dxi_dx=eye, jacobian=np.ones(nn), lambda_=np.full(nn, 2.0), mu_=np.ones(nn),
                mass=np.ones(nn), density=np.ones(nn), damping=np.zeros(nn),
                ibool=np.arange(nn, dtype=np.int32), pml_region=np.array([region], np.int32),
                coef_alpha=np.tile(alpha, nn), coef_abar=np.tile(abar, nn))


def test_cpml_whole_step_three_steps_known_answer():
    """Three steps of the CPU branch of solver.cpp:510-621 with C-PML on a rigid
    translation, worked by hand from its call order: fields from (u_n, v_n, a_n) before the
    element kernel (pml.cpp:55-110), a_{n+1} = A1 v_n + A2 u_n + A3 m_x + A4 m_y + A5 m_z with
    the memory of the previous step, Newmark corrector, then the memory update from the new
    and old fields (pml.cpp:112-149).  dt = 0.1, A = (1, -2, 4, 8, 16), alpha_x = (.5, .25,
    .125), alpha_y = (.9, .1, .05), alpha_z = (0, 1, -1).  x starts from (u, v, a) = (0, 1, 0):
      step 0: new 0.0375, old 0, a 1, u 0.1, v 1.05, m (0.009375, 0.00375, 0.0375)
      step 1: new 0.139375, old 0.041875, a 1.05 - 0.2 + 0.0375 + 0.03 + 0.6 = 1.5175,
              u 0.21, v 1.175875, m (0.044765625, 0.01940625, 0.0975)
      step 2: a 1.175875 - 0.42 + 0.1790625 + 0.15525 + 1.56 = 2.6501875,
              u 0.21 + 0.1175875 + 0.0075875 = 0.335175, v 1.175875 + 0.05 (1.5175 + 2.6501875)
    y from (1, 0, 0) and z from (0, 0, 2) are worked the same way.  Using u~ instead of u_n
    in the C-PML term, updating the memory before it is used, or taking 'old' from the
    current 'new' each fails."""
    alpha = [.5, .25, .125, .9, .1, .05, 0.0, 1.0, -1.0]
    arrays = _one_element_n1([1.0, -2.0, 4.0, 8.0, 16.0], alpha)
    run = pyoracle.Run(arrays, 1, 2, 8, 0.1, 3, 0, 2, None, None, None, None, has_cpml=True)
    u0 = np.tile([0.0, 1.0, 0.0], 8)
    v0 = np.tile([1.0, 0.0, 0.0], 8)
    a0 = np.tile([0.0, 0.0, 2.0], 8)
    u, v, a = run.run_state(u0, v0, a0)
    want_u = [0.335175, 1.117, 0.02947875]
    want_v = [1.384259375, 1.50051875, 0.1101615625]
    want_a = [2.6501875, 2.610375, 0.30748125]
    for got, want in ((u, want_u), (v, want_v), (a, want_a)):
        assert np.allclose(got.reshape(8, 3), want, rtol=1e-13, atol=1e-15)
    # an interior element (region 0) takes no C-PML term: plain rigid motion
    arrays = _one_element_n1([1.0, -2.0, 4.0, 8.0, 16.0], alpha, region=0)
    run = pyoracle.Run(arrays, 1, 2, 8, 0.1, 3, 0, 2, None, None, None, None, has_cpml=True)
    u, v, a = run.run_state(u0, v0, a0)
    assert np.allclose(u.reshape(8, 3), [0.3, 1.0, 0.03], rtol=0, atol=1e-14)
    assert np.allclose(v.reshape(8, 3), [1.0, 0.0, 0.1], rtol=0, atol=1e-14)
    assert np.all(a == 0.0)


# ------------------------------------------ recorded strain (forward/share/src/solver.cpp:282-335)


def test_full_strain_is_exact_for_affine_displacement():
    """u = A x + b at the nodes of a jittered (non-affine) order-4 mesh: the element maps
    are trilinear, so u is a polynomial of degree <= 1 per reference axis, the GLL
    derivative is exact, and the strain at every point of every cell is sym(A) in the
    reference's Voigt order (xx, yy, zz, (u_x,y + u_y,x)/2, (u_x,z + u_z,x)/2,
    (u_y,z + u_z,y)/2).  A transposed dxi_dx, a wrong D index or a dropped chain-rule
    term fails it (the jitter makes dxi_dx full and non-symmetric)."""
    import ctypes as C

    import pygfx
    m = pygfx.build_box(nx=3, ny=3, nz=2, lx=2700.0, ly=2700.0, lz=1800.0, order=4,
                        layers=[(0.0, 3000.0, 1700.0, 2400.0)], jitter=0.15, jitter_seed=9,
                        pml_xmin=0, pml_xmax=0, pml_ymin=0, pml_ymax=0, pml_zmax=0,
                        record_depth_max=0.0, duration=0.1)
    dd = m.dxi_dx.reshape(-1, 3, 3)
    assert np.abs(dd - dd.transpose(0, 2, 1)).max() > 1e-6 * np.abs(dd).max()  # non-symmetric
    A = np.array([[0.3, -1.1, 0.7], [2.0, -0.4, 0.9], [-0.6, 1.3, 0.25]]) * 1e-3
    xyz = m.node_xyz.reshape(-1, 3)
    ug = (xyz @ A.T + np.array([5.0, -2.0, 1.0])).ravel()
    nn = 125
    uloc = pyoracle.gather(ug, m.ibool, m.n_elem, nn)
    x, w, D = pyoracle.gll(4)
    strain = np.zeros(m.n_elem * nn * 6)
    pyoracle.lib.gfo_full_strain(m.n_elem, pyoracle.P(None, C.c_int32), pyoracle.P(m.dxi_dx),
                                 pyoracle.P(D), 5, pyoracle.P(uloc), pyoracle.P(strain))
    want = np.array([A[0, 0], A[1, 1], A[2, 2], 0.5 * (A[0, 1] + A[1, 0]), 0.5 * (A[0, 2] + A[2, 0]),
                     0.5 * (A[1, 2] + A[2, 1])])
    assert np.allclose(strain.reshape(-1, 6), want, rtol=0, atol=1e-12 * np.abs(A).max())
    # the recording subset: the same values for a list of cells, in list order
    cells = np.array([5, 0, 11], np.int32)
    sub = np.zeros(len(cells) * nn * 6)
    pyoracle.lib.gfo_full_strain(len(cells), pyoracle.P(cells, C.c_int32), pyoracle.P(m.dxi_dx),
                                 pyoracle.P(D), 5, pyoracle.P(uloc), pyoracle.P(sub))
    assert np.array_equal(sub.reshape(3, nn, 6), strain.reshape(-1, nn, 6)[cells])


def test_run_from_is_the_reference_resume():
    """gfo_run_from(k + 1, state after step k) is solver.cpp:369-388's resume.  Without
    C-PML (whose memory the reference's restart does not keep) it continues the
    uninterrupted run exactly: same records for the later snapshot steps, same final
    state; started at step 0 from zeros it is gfo_run."""
    import meshcases as mc
    m = mc.small_halfspace(nx=4, ny=4, nz=3, pml=1, duration=1.0)
    cells, w, stf = mc.surface_centre_source(m)
    full = pyoracle.Run(m.arrays, m.n_elem, 5, m.n_node, m.dt, m.nsteps, m.snapshot_stride, 2, cells, w, stf,
                        m.rec_cell, has_cpml=False).run(uva=True)
    k = 37
    head = pyoracle.Run(m.arrays, m.n_elem, 5, m.n_node, m.dt, k + 1, m.snapshot_stride, 2, cells, w, stf,
                        m.rec_cell, has_cpml=False)
    z = np.zeros(3 * m.n_node)
    u, v, a = head.run_state(z, z, z)
    tail = pyoracle.Run(m.arrays, m.n_elem, 5, m.n_node, m.dt, m.nsteps, m.snapshot_stride, 2, cells, w, stf,
                        m.rec_cell, has_cpml=False).run_from(k + 1, u, v, a, uva=True)
    later = [i for i, s in enumerate(range(0, m.nsteps, m.snapshot_stride)) if s > k]
    assert tail["strain"].shape[0] == len(later)
    for key in ("strain", "disp", "vel", "acc"):
        assert np.array_equal(tail[key], full[key][later]), key
    for key in ("u", "v", "a"):
        assert np.array_equal(tail[key], full[key]), key
    again = pyoracle.Run(m.arrays, m.n_elem, 5, m.n_node, m.dt, m.nsteps, m.snapshot_stride, 2, cells, w, stf,
                         m.rec_cell).run_from(0, z, z, z)
    ref = pyoracle.Run(m.arrays, m.n_elem, 5, m.n_node, m.dt, m.nsteps, m.snapshot_stride, 2, cells, w, stf,
                       m.rec_cell).run()
    assert np.array_equal(again["strain"], ref["strain"]) and np.array_equal(again["u"], ref["u"])
