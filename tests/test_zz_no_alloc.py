"""The allocation-free element plugin (gfx_compute_element_residual_noalloc)
against the allocating one (gfx_compute_element_residual), through the C ABI.

Both run the same FP64 device kernel on the same inputs, so the bar is bit
equality (integer-exact work is compared bit for bit; this is the same
arithmetic on the same bytes).  Two calls in a row, the second with a smaller
input, so the second call runs on the scratch array the first one filled.
Sizes past the scratch maximum are refused before any device work."""
import ctypes as C

import numpy as np
import pytest

import pygfx
import pyoracle

pytestmark = pytest.mark.gpu

_dp = C.POINTER(C.c_double)
SCRATCH_BYTES = 64 << 20
GFX_ERR_INVALID = 1


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if pygfx.device_count() < 1:
        pytest.fail("no HIP device visible; GPU tests must run on an MI355X box")


def _fn():
    fn = pygfx.lib.gfx_compute_element_residual_noalloc
    fn.restype = C.c_int
    fn.argtypes = [C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, C.c_int, _dp, _dp]
    return fn


def _err():
    return pygfx.lib.gfx_last_error().decode()


def _p(a):
    return a.ctypes.data_as(_dp)


def _inputs(n_elem, ngll, seed):
    rng = np.random.default_rng(seed)
    nn = ngll ** 3
    P = n_elem * nn
    dd = rng.standard_normal(P * 9) * 0.2
    dd[0::9] += 3.0
    dd[4::9] += 3.0
    dd[8::9] += 3.0
    jac = np.abs(rng.standard_normal(P)) + 0.01
    rho = 2000.0 + 1000.0 * rng.random(P)
    vs = 2000.0 + 1500.0 * rng.random(P)
    mu = rho * vs * vs
    mu[rng.random(P) < 0.02] = 0.0  # the mu <= 0 skip (element_cpu.cpp:42-43) on some points
    lam = rho * (3.0 * vs * vs)
    _, w, D = pyoracle.gll(ngll - 1)
    u = rng.standard_normal(P * 3)
    r0 = rng.standard_normal(P * 3)  # r accumulates
    return [np.ascontiguousarray(x, np.float64) for x in (dd, jac, lam, mu, D, w, u)], r0


def _call(fn, n_elem, ngll, arrs, r):
    return fn(n_elem, *[_p(a) for a in arrs[:6]], ngll, _p(arrs[6]), _p(r))


def test_noalloc_equals_allocating_entry_point_twice_shrinking():
    fn = _fn()
    ref_fn = pygfx.lib.gfx_compute_element_residual
    for n_elem, ngll, seed in ((3000, 5, 11), (257, 5, 12), (40, 4, 13)):
        print(f"case: noalloc plugin n_elem {n_elem} ngll {ngll} seed {seed}", flush=True)
        arrs, r0 = _inputs(n_elem, ngll, seed)
        want = r0.copy()
        assert ref_fn(n_elem, *[_p(a) for a in arrs[:6]], ngll, _p(arrs[6]), _p(want)) == 0, _err()
        got = r0.copy()
        assert _call(fn, n_elem, ngll, arrs, got) == 0, _err()
        assert np.array_equal(got, want), (n_elem, ngll, np.abs(got - want).max())
        assert not np.array_equal(got, r0)


def test_noalloc_matches_oracle_and_rejects_oversize():
    fn = _fn()
    n_elem, ngll = 64, 5
    arrs, r0 = _inputs(n_elem, ngll, 21)
    got = r0.copy()
    assert _call(fn, n_elem, ngll, arrs, got) == 0, _err()
    ref = r0.copy()
    pyoracle.element_residual(n_elem, *arrs[:6], ngll, arrs[6], ref)
    rel = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-14)
    assert rel.max() < 1e-12
    # per element at NGLL 5: 125 points x (9 + 3 + 3 + 3) doubles = 18,000 B; 64 MiB holds < 3,800
    big = 4000
    bad = np.zeros(1)
    rc = fn(big, _p(bad), _p(bad), _p(bad), _p(bad), _p(bad), _p(bad), ngll, _p(bad), _p(bad))
    assert rc == GFX_ERR_INVALID
    assert "scratch" in _err()
    assert fn(0, None, None, None, None, None, None, ngll, None, None) == 0
