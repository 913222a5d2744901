"""Complete C-PML in the FP64 oracle: the strain (stretched-stress) term the
reference defers (docs/deferred.md:46-80) and the inertia term written from the
stretched weak form.  Pinned by the mathematics, not by retyping the update:

  * the one-step recursion of conv(g, p) is exact for inputs linear in time
    (closed-form integrals, including p = 0 and the series branch);
  * the gradient filters reproduce s_x s_y s_z / (s_j s_l) and the inertia
    filter (i w)^2 (s_x s_y s_z - K_x K_y K_z) in the periodic steady state of a
    sampled cosine, to second order in dt (the error falls 4x when dt halves),
    with coincident poles and undamped axes among the cases;
  * with d = 0 the stretches are real coordinate stretches: a PML element's
    residual equals the elastic residual of the stretched geometry
    (dxi/dx / K, J K_x K_y K_z), which fixes which stress column sees which
    derivative;
  * with d = 0 and K = 1 a whole run equals the elastic run bit for bit;
  * the solver's rigid outer boundary (found from unshared element faces and
    the regions' damped axes) is the set of nodes on the box faces that carry
    a layer (found from the coordinates)."""
import numpy as np
import pytest

import meshcases as mc
import pygfx
import pyoracle as po


def conv_exact_ramp(p, T):
    """int_0^T exp(-p (T - s)) s ds."""
    x = p * T
    if x < 1e-3:
        return T * T * (0.5 - x / 6 + x * x / 24 - x ** 3 / 120)
    return (x + np.expm1(-x)) / (p * p)


@pytest.mark.parametrize("p", [0.0, 1e-9, 1e-4, 0.3, 5.0, 80.0, 2000.0])
def test_recursion_is_exact_for_inputs_linear_in_time(p):
    dt = 1e-3
    a, b, c = po.cpml_recursion(p, dt)
    assert a == pytest.approx(np.exp(-p * dt), rel=1e-15)
    n = 700
    q, out = 0.0, []
    for s in range(n):  # g(t) = t sampled at t_s = (s + 1) dt, g = 0 at the level before
        g = (s + 1) * dt
        m = q + c * g
        q = a * m + b * g
        out.append(m)
    want = np.array([conv_exact_ramp(p, (s + 1) * dt) for s in range(n)])
    np.testing.assert_allclose(out, want, rtol=1e-12, atol=1e-18)
    # a kink on a grid level: g = (t - t0)_+ stays exact
    t0 = 0.2
    q, got = 0.0, None
    for s in range(n):
        t = (s + 1) * dt
        g = max(t - t0, 0.0)
        got = q + c * g
        q = a * got + b * g
    assert got == pytest.approx(conv_exact_ramp(p, n * dt - t0), rel=1e-11)


def _stretch(K, d, al, w):
    return K + d / (al + 1j * w)


def _coef_cases():
    """(K, d, alpha) per axis: generic, coincident alphas, beta_l equal to an
    alpha, an undamped axis, alpha = 0 on a damped axis, K != 1."""
    return [
        ((1.0, 1.0, 1.0), (30.0, 12.0, 45.0), (3.0, 5.0, 2.0)),
        ((1.0, 1.0, 1.0), (30.0, 12.0, 45.0), (4.0, 4.0, 4.0)),
        ((1.0, 1.0, 1.0), (6.0, 20.0, 0.0), (3.0, 9.0, 2.5)),   # beta_x = alpha_y = 9
        ((1.0, 1.0, 1.0), (25.0, 0.0, 0.0), (3.0, 7.0, 1.0)),   # x only
        ((1.0, 1.0, 1.0), (0.0, 18.0, 33.0), (5.0, 0.0, 6.0)),  # alpha_y = 0 where d_y > 0
        ((1.4, 1.0, 2.2), (20.0, 35.0, 10.0), (2.0, 6.0, 3.5)),
    ]


def _dc(x, zero_pole):
    """A damped axis with alpha = 0 integrates without decay, so the start-up
    (the input jumps from 0 at the level before the first step) leaves a
    constant offset: such cases are compared up to a constant, over whole
    periods."""
    return x - x.mean(axis=0) if zero_pole else x


def _steady_gradient_error(K, d, al, dt, w, T):
    K, d, al = (np.array(x, float) for x in (K, d, al))
    zp = bool(np.any((d > 0) & (al == 0)))
    coef = po.cpml_point_coefficients(K, d, al, dt)
    n = int(round(T / dt))
    t = (np.arange(n) + 1) * dt
    s = _stretch(K, d, al, w)
    err = 0.0
    for kl in (1, 5, 6):  # (k, l) = (0, 1), (1, 2), (2, 0)
        k, l = divmod(kl, 3)
        g = np.zeros((n, 9))
        g[:, kl] = np.cos(w * t)
        G, _ = po.cpml_filter_history(coef, g=g)
        tail = slice(n - n // 10, n)
        for j in range(3):
            H = s.prod() / (s[j] * s[l])
            want = _dc(np.real(H * np.exp(1j * w * t[tail])), zp)
            got = _dc(G[tail, j, kl], zp)
            err = max(err, np.max(np.abs(got - want)) / abs(H))
            others = np.delete(G[tail, j, :], kl, axis=1)
            assert np.all(others == 0.0)
    return err


@pytest.mark.parametrize("case", range(6))
def test_gradient_filters_are_the_stretch_ratios(case):
    K, d, al = _coef_cases()[case]
    w = 2 * np.pi * 2.5  # the tail (T/10) holds whole periods at both steps
    e1 = _steady_gradient_error(K, d, al, 2e-3, w, 12.0)
    e2 = _steady_gradient_error(K, d, al, 1e-3, w, 12.0)
    assert e2 < 2e-4, e2
    assert 3.0 < e1 / e2 < 5.0, (e1, e2)  # second order in dt


def _steady_inertia_error(K, d, al, dt, w, T):
    K, d, al = (np.array(x, float) for x in (K, d, al))
    zp = bool(np.any((d > 0) & (al == 0)))
    coef = po.cpml_point_coefficients(K, d, al, dt)
    n = int(round(T / dt))
    t = (np.arange(n) + 1) * dt
    u = np.zeros((n, 3))
    v = np.zeros((n, 3))
    amp = np.array([1.0, -0.6, 0.3])
    u[:] = amp * np.cos(w * t)[:, None]
    v[:] = -w * amp * np.sin(w * t)[:, None]
    _, L = po.cpml_filter_history(coef, u=u, v=v)
    s = _stretch(K, d, al, w)
    H = (1j * w) ** 2 * (s.prod() - K.prod())
    tail = slice(n - n // 10, n)
    want = _dc(np.real(H * np.exp(1j * w * t[tail]))[:, None] * amp, zp)
    return np.max(np.abs(_dc(L[tail], zp) - want)) / abs(H)


@pytest.mark.parametrize("case", range(6))
def test_inertia_filter_is_the_stretched_mass_term(case):
    K, d, al = _coef_cases()[case]
    w = 2 * np.pi * 2.5
    e1 = _steady_inertia_error(K, d, al, 2e-3, w, 12.0)
    e2 = _steady_inertia_error(K, d, al, 1e-3, w, 12.0)
    assert e2 < 2e-4, e2
    if e1 > 1e-12:
        assert 3.0 < e1 / e2 < 5.0, (e1, e2)


def test_undamped_axes_keep_zero_memories():
    coef = po.cpml_point_coefficients([1.0, 1.3, 1.0], [0.0, 0.0, 0.0], [3.0, 4.0, 5.0], 1e-3).reshape(3, 9)
    assert np.all(coef[:, 3:] == 0.0)
    G, L = po.cpml_filter_history(coef.ravel(), g=np.ones((5, 9)), u=np.ones((5, 3)), v=np.ones((5, 3)))
    assert np.all(L == 0.0)
    # pure real stretch: G^(j)_kl = K_x K_y K_z / (K_j K_l) g
    Kv = np.array([1.0, 1.3, 1.0])
    for j in range(3):
        for kl in range(9):
            assert G[-1, j, kl] == pytest.approx(Kv.prod() / (Kv[j] * Kv[kl % 3]), rel=1e-15)


def test_real_stretch_equals_elastic_residual_of_stretched_geometry():
    rng = np.random.default_rng(5)
    x, w, D = po.gll(4)
    E, nn = 3, 125
    dd = rng.normal(size=(E, nn, 9)) * 1e-3
    jac = rng.uniform(0.5, 2.0, (E, nn)) * 1e8
    lam = rng.uniform(1e10, 3e10, (E, nn))
    mu = rng.uniform(0.5e10, 2e10, (E, nn))
    mu[1, 7] = 0.0  # the mu <= 0 skip holds inside PML elements too
    u = rng.normal(size=(E, nn, 3))
    Kv = np.array([1.3, 1.7, 2.1])
    coef = np.zeros((E, nn, 27))
    for e in range(E):
        for n in range(nn):
            coef[e, n] = po.cpml_point_coefficients(Kv, [0, 0, 0], [1, 2, 3], 1e-3)
    region = np.full(E, 7, np.int32)
    r = np.zeros(E * nn * 3)
    carry = np.zeros(E * nn * 45)
    po.element_residual_cpml(E, dd.ravel(), jac.ravel(), lam.ravel(), mu.ravel(), region, coef.ravel(), D, w, 5,
                             u.ravel(), carry, r)
    # stretched geometry x~_m = K_m x_m: d xi / d x~_m = (d xi / d x_m) / K_m, J~ = J K_x K_y K_z
    dds = (dd.reshape(E, nn, 3, 3) / Kv[None, None, :, None]).ravel()
    ref = np.zeros_like(r)
    po.element_residual(E, dds, (jac * Kv.prod()).ravel(), lam.ravel(), mu.ravel(), D, w, 5, u.ravel(), ref)
    np.testing.assert_allclose(r, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    # interior elements (region 0) are the plain elastic residual
    r0 = np.zeros_like(r)
    po.element_residual_cpml(E, dd.ravel(), jac.ravel(), lam.ravel(), mu.ravel(), np.zeros(E, np.int32),
                             coef.ravel(), D, w, 5, u.ravel(), carry, r0)
    ref0 = np.zeros_like(r)
    po.element_residual(E, dd.ravel(), jac.ravel(), lam.ravel(), mu.ravel(), D, w, 5, u.ravel(), ref0)
    assert np.array_equal(r0, ref0)


def _profiles(mesh, K=1.0, d=0.0, alpha=2.0):
    P = mesh.n_elem * 125
    return dict(pml_K=np.full(P * 3, K), pml_d=np.full(P * 3, d), pml_alpha=np.full(P * 3, alpha))


def test_unit_undamped_complete_run_is_the_elastic_run():
    mesh = mc.small_halfspace(nx=5, ny=5, nz=4, pml=1, duration=0.4)
    c, w, stf = mc.surface_centre_source(mesh)
    arr = dict(mesh.arrays)
    arr.update(_profiles(mesh))
    arr["damping"] = np.zeros_like(arr["damping"])
    args = (arr, mesh.n_elem, 5, mesh.n_node, mesh.dt, mesh.nsteps, mesh.snapshot_stride, 2, c, w, stf,
            mesh.rec_cell)
    a = po.Run(*args, has_cpml=False).run()
    b = po.Run(*args, has_cpml=True, cpml_complete=True).run()
    for k in ("strain", "u", "v", "a"):
        assert np.array_equal(a[k], b[k]), k


def test_rigid_boundary_is_the_layered_faces():
    for mesh, faces in ((mc.small_halfspace(nx=6, ny=6, nz=4, pml=2, want_profiles=1), None),
                        (mc.small_layered(nx=7, ny=6, nz=5, pml=2, want_profiles=1), None),
                        (mc.small_halfspace(nx=6, ny=5, nz=4, pml=1, pml_xmin=0, pml_zmax=2, want_profiles=1),
                         ("xmax", "ymin", "ymax", "zmax"))):
        got = pygfx.cpml_rigid_nodes(mesh)
        want = mc.rigid_outer_nodes(mesh) if faces is None else mc.rigid_outer_nodes(mesh, faces)
        assert np.array_equal(got, want)


def test_generator_profiles():
    """pml_profile 0 is what preprocess/pml_cpml.py:63-155 computes: per damped
    axis the normalised distance from a boundary face, the later face (max)
    overwriting the earlier (min), so the min-side layer gets dist = 1
    (d = d0, alpha = 0) throughout and the max-side layer d = d0 at its
    interior interface, 0 at the boundary.  pml_profile 1 measures the depth
    from the interior interface of each face's own layer.  Halfspace, vp 6000,
    d0 = 3 vp ln(1/R) / (2 w), exponent 2.4, alpha_max = pi f0 (0.9, 1, 1.1)."""
    kw = dict(nx=8, ny=7, nz=5, pml=2, want_profiles=1, want_coords=1, f0_pml=1.2)
    ref = mc.small_halfspace(pml_profile=0, **kw)
    grad = mc.small_halfspace(pml_profile=1, pml_reflection=1e-3, **kw)
    P = ref.n_elem * 125
    xyz = ref.coords.reshape(P, 3)
    reg = np.repeat(ref.pml_region, 125)
    b = ref.bounds
    w = 2000.0
    amax = np.pi * 1.2 * np.array([0.9, 1.0, 1.1])
    damps = {1: (0,), 2: (1,), 3: (2,), 4: (0, 1), 5: (0, 2), 6: (1, 2), 7: (0, 1, 2)}
    for m, R, graded in ((ref, 1e-5, False), (grad, 1e-3, True)):
        d0 = 3.0 * 6000.0 * np.log(1.0 / R) / (2.0 * w)
        K, d, al = (m.arrays[k].reshape(P, 3) for k in ("pml_K", "pml_d", "pml_alpha"))
        assert np.all(K == 1.0)
        for ax, (lo, hi) in enumerate(((b[0], b[1]), (b[2], b[3]), (b[4], b[5]))):
            on = np.array([r in damps and ax in damps[r] for r in reg])
            assert np.all(d[~on, ax] == 0.0) and np.all(al[~on, ax] == 0.0)
            x = xyz[on, ax]
            if graded:
                depth = (x - (hi - w)) / w
                if ax < 2:
                    depth = np.maximum(depth, (lo + w - x) / w)
                dist = np.clip(depth, 0.0, 1.0)
            else:
                dist = np.clip(np.abs(x - hi) / w, 0.0, 1.0)  # the max face written last
            np.testing.assert_allclose(d[on, ax], d0 * dist ** 2.4, rtol=1e-12, atol=1e-9)
            np.testing.assert_allclose(al[on, ax], amax[ax] * (1.0 - dist), rtol=1e-12, atol=1e-12)
    # the reference's acceleration coefficients follow the profiles it is given
    assert not np.array_equal(ref.coef_abar, grad.coef_abar)


def test_attenuation_with_no_pml_elements_is_the_attenuating_run():
    """Complete C-PML with attenuation: the interior elements take the
    attenuating residual (mu zeroed in the PML) and the PML elements the
    stretched elastic one (mu zeroed inside).  With no PML element the run is
    the plain attenuating run bit for bit."""
    mesh = mc.small_layered(nx=5, ny=5, nz=4, pml=1, duration=0.4, n_sls=3, q_mu=[20.0, 20.0, 20.0],
                            att_fmin=0.1, att_fmax=4.0, want_profiles=1)
    c, w, stf = mc.buried_source(mesh)
    arr = dict(mesh.arrays)
    arr["pml_region"] = np.zeros_like(arr["pml_region"])
    arr["damping"] = np.zeros_like(arr["damping"])
    args = (arr, mesh.n_elem, 5, mesh.n_node, mesh.dt, mesh.nsteps, mesh.snapshot_stride, 2, c, w, stf,
            mesh.rec_cell)
    att = dict(tau_sigma=mesh.tau_sigma, tau_epsilon=mesh.tau_epsilon)
    a = po.Run(*args, has_cpml=False, **att).run()
    b = po.Run(*args, has_cpml=True, cpml_complete=True, **att).run()
    for k in ("strain", "u", "v", "a"):
        assert np.array_equal(a[k], b[k]), k
