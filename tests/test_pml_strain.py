"""Strain-based C-PML coefficients (pml_coef_strain, A6..A23).

The reference generator writes them (preprocess/pml_cpml.py:523-734) and its
solver does not use them yet; the element-kernel correction that would is not
built here either.  The coefficients are pinned by the mathematics, not by
retyping the formulas: with s_a(w) = K_a + d_a / (alpha_a + i w), every group
of four must reproduce its region's rational function G(s_1, s_2, s_3) at any
frequency as A0 + A1/(alpha_1 + i w) + A2/(alpha_2 + i w) + A3/(beta_3 + i w),
and the last six are the (K, d) pairs of the single stretches."""
import numpy as np
import pytest

import meshcases as mc
import pygfx

X, Y, Z, XY, XZ, YZ, XYZ = 1, 2, 3, 4, 5, 6, 7
ACTIVE = {X: (0,), Y: (1,), Z: (2,), XY: (0, 1), XZ: (0, 2), YZ: (1, 2), XYZ: (0, 1, 2)}
FORM = {XYZ: lambda s1, s2, s3: s1 * s2 / s3, X: lambda s1, s2, s3: s1, Y: lambda s1, s2, s3: s2,
        Z: lambda s1, s2, s3: 1 / s3, XY: lambda s1, s2, s3: s1 * s2, XZ: lambda s1, s2, s3: s1 / s3,
        YZ: lambda s1, s2, s3: s2 / s3}
PERM = [(2, 1, 0), (0, 2, 1), (0, 1, 2)]  # directions (1, 2, 3) of the three groups


def point(rng, region):
    """K >= 1, d > 0 and distinct alphas in the damped directions; K = 1, d = 0 elsewhere."""
    K, d, a = np.ones(3), np.zeros(3), rng.uniform(0.5, 6.0, 3)
    for ax in ACTIVE[region]:
        K[ax] = rng.uniform(1.0, 3.0)
        d[ax] = rng.uniform(5.0, 400.0)
    return K, d, a


@pytest.mark.parametrize("region", [X, Y, Z, XY, XZ, YZ, XYZ])
def test_groups_are_the_partial_fractions_of_the_region_form(region):
    rng = np.random.default_rng(region)
    w = np.geomspace(1e-2, 1e3, 40)
    for _ in range(20):
        K, d, a = point(rng, region)
        S = pygfx.pml_strain_coefficients(region, K, d, a)
        beta = a + d / np.maximum(K, 1.0)
        s = [K[q] + d[q] / (a[q] + 1j * w) for q in range(3)]
        for g, (p1, p2, p3) in enumerate(PERM):
            A0, A1, A2, A3 = S[4 * g:4 * g + 4]
            want = FORM[region](s[p1], s[p2], s[p3])
            got = A0 + A1 / (a[p1] + 1j * w) + A2 / (a[p2] + 1j * w) + A3 / (beta[p3] + 1j * w)
            np.testing.assert_allclose(got, want, rtol=1e-9, atol=1e-12, err_msg=f"group {g}")
        for ax in range(3):
            pair = S[12 + 2 * ax:14 + 2 * ax]
            want = (K[ax], K[ax] * d[ax] / max(K[ax], 1.0)) if ax in ACTIVE[region] else (1.0, 0.0)
            np.testing.assert_allclose(pair, want, rtol=1e-12)


def test_xyz_groups_are_the_full_stretch_ratios():
    """In the corner region all three groups are the modified-gradient factors
    s_y s_z / s_x, s_x s_z / s_y, s_x s_y / s_z of the stretched weak form."""
    rng = np.random.default_rng(11)
    K, d, a = point(rng, XYZ)
    S = pygfx.pml_strain_coefficients(XYZ, K, d, a)
    w = 2.0 * np.pi * 0.7
    s = K + d / (a + 1j * w)
    beta = a + d / K
    for g, (num1, num2, den) in enumerate([(2, 1, 0), (0, 2, 1), (0, 1, 2)]):
        A0, A1, A2, A3 = S[4 * g:4 * g + 4]
        val = A0 + A1 / (a[num1] + 1j * w) + A2 / (a[num2] + 1j * w) + A3 / (beta[den] + 1j * w)
        assert np.isclose(val, s[num1] * s[num2] / s[den], rtol=1e-10)


def test_interior_points_and_bad_input():
    assert np.array_equal(pygfx.pml_strain_coefficients(0, [1, 1, 1], [1, 2, 3], [1, 2, 3]), np.zeros(18))
    with pytest.raises(pygfx.GfxError, match="region"):
        pygfx.pml_strain_coefficients(8, [1, 1, 1], [0, 0, 0], [1, 1, 1])


def test_generator_writes_the_coefficients_per_region():
    m = pygfx.build_box(nx=6, ny=6, nz=4, lx=6000.0, ly=6000.0, lz=4000.0, layers=[(0.0, 6000.0, 3464.0, 2700.0)],
                        pml_xmin=1, pml_xmax=1, pml_ymin=1, pml_ymax=1, pml_zmax=1, duration=0.2, want_strain=1)
    S = m.coef_strain.reshape(m.n_elem, 125, 18)
    reg = m.pml_region
    assert np.all(S[reg == 0] == 0)
    assert np.all(np.isfinite(S))
    # x-only cells damp x alone: (K, d) of s_y and s_z stay (1, 0), s_x has d > 0 inside
    xo = S[reg == X]
    assert xo.size and np.all(xo[..., 14] == 1) and np.all(xo[..., 15] == 0) and np.all(xo[..., 17] == 0)
    assert np.all(xo[..., 12] >= 1) and np.any(xo[..., 13] > 0)
    # the box has corner cells, which carry all three groups
    assert np.any(reg == XYZ)


@pytest.mark.skipif(not pygfx.h5_available(), reason="no HDF5 library to bind")
def test_coefficients_round_trip_through_the_partitions(tmp_path):
    m = mc.small_halfspace(nx=6, ny=6, nz=4, pml=1, duration=0.3, want_strain=1)
    stf = pygfx.ricker(1.0, 0.5, 1e20, m.dt, m.nsteps)
    d = str(tmp_path)
    pygfx.write_inputs(m, d, (3000.0, 3000.0, 0.0), stf, n_parts=2)
    lay = pygfx.h5_layout(str(tmp_path / "partitions" / "partition_0.h5"))
    assert lay["/field/cell/pml_coef_strain"][1][1:] == (125, 18)  # [E, 125, 18] as pml_coef_beta
    back = pygfx.read_inputs(d)
    ids = np.concatenate([pygfx.h5_read(str(tmp_path / "partitions" / f"partition_{r}.h5"),
                                        "/partition/local_cell_ids") for r in range(2)]).astype(np.int64)
    E = m.n_elem
    assert np.array_equal(back.coef_strain.reshape(E, -1), m.coef_strain.reshape(E, -1)[ids])


@pytest.mark.parametrize("profile", [0, 1])
def test_abar_are_the_partial_fractions_of_the_stretched_inertia(profile):
    """pml_coef_abar (preprocess/pml_cpml.py:388-513) at every C-PML point with
    distinct poles, checked against the point's own K, d, alpha (pml_K, pml_d,
    pml_alpha): (i w)^2 (s_x s_y s_z - K_x K_y K_z) =
    A1 i w + A2 + A3/(alpha_x + i w) + A4/(alpha_y + i w) + A5/(alpha_z + i w)
    at 30 frequencies.  Every region matches to 1e-8 except one term: in corner
    (XYZ) cells the reference's A5 has the opposite sign (its denominator is
    (alpha_y - alpha_z)(alpha_z - alpha_x) where the residue needs
    (alpha_x - alpha_z)(alpha_y - alpha_z)); the generator keeps the reference's
    value, and the test pins that sign."""
    m = mc.small_halfspace(nx=8, ny=7, nz=5, pml=2, want_profiles=1, pml_profile=profile)
    P = m.n_elem * 125
    K, d, al = (m.arrays[k].reshape(P, 3) for k in ("pml_K", "pml_d", "pml_alpha"))
    ab = m.coef_abar.reshape(P, 5)
    reg = np.repeat(m.pml_region, 125)
    w = np.geomspace(1e-2, 1e3, 30)
    checked = {r: 0 for r in range(1, 8)}
    for p in np.flatnonzero(reg):
        act = d[p] > 1e-3
        a = al[p][act]
        if not act.any() or (len(a) > 1 and np.min(np.abs(np.diff(np.sort(a)))) < 1e-2):
            continue
        s = K[p] + d[p] / (al[p][None, :] + 1j * w[:, None])
        lhs = (1j * w) ** 2 * (s.prod(axis=1) - K[p].prod())
        A = ab[p].copy()
        if reg[p] == XYZ:
            A[4] = -A[4]
        rhs = A[0] * 1j * w + A[1] + sum(A[2 + q] / (al[p, q] + 1j * w) for q in range(3))
        err = np.max(np.abs(lhs - rhs) / (np.abs(lhs) + 1e-12 * (w ** 2 + 1)))
        assert err < 1e-8, (reg[p], K[p], d[p], al[p], ab[p], err)
        checked[reg[p]] += 1
    assert all(n > 50 for n in checked.values()), checked
