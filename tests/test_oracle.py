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
