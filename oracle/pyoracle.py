"""ctypes binding of oracle/liboracle.so -- TEST INFRASTRUCTURE ONLY.

The FP64 CPU restatement of the reference solver is the parity checker.  Only
tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class Problem(C.Structure):
    _fields_ = [
        ("n_elem", C.c_int), ("ngll", C.c_int), ("n_rank_node", C.c_int),
        ("dxi_dx", _dp), ("jacobian", _dp), ("lambda_", _dp), ("mu_", _dp), ("mass", _dp),
        ("density", _dp), ("damping", _dp), ("ibool", _ip),
        ("has_cpml", C.c_int), ("pml_region", _ip), ("coef_alpha", _dp), ("coef_abar", _dp),
        ("n_src_cell", C.c_int), ("src_elem", _ip), ("src_weights", _dp),
        ("direction", C.c_int), ("nsteps", C.c_int), ("dt", C.c_double),
        ("n_stf", C.c_int), ("stf", _dp), ("snapshot_stride", C.c_int),
        ("n_rec_cell", C.c_int), ("rec_cell", _ip),
        ("n_sls", C.c_int), ("tau_sigma", _dp), ("tau_epsilon", _dp),
        ("cpml_complete", C.c_int), ("pml_K", _dp), ("pml_d", _dp), ("pml_alpha", _dp),
    ]


class Cpml(C.Structure):
    _fields_ = [
        ("n_elem", C.c_int), ("ngll", C.c_int),
        ("pml_region", _ip), ("coef_alpha", _dp), ("coef_abar", _dp), ("density", _dp),
        ("jacobian", _dp), ("ibool", _ip),
        ("displ_old", _dp), ("displ_new", _dp), ("rmemory", _dp),
    ]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is not built; run `make -C oracle`")
    lib = C.CDLL(LIB_PATH)
    sig = {
        "gfo_gll_nodes": (C.c_int, [C.c_int, _dp]),
        "gfo_gll_weights": (C.c_int, [C.c_int, _dp, _dp]),
        "gfo_gll_derivative_matrix": (C.c_int, [C.c_int, _dp, _dp]),
        "gfo_lagrange_basis": (None, [C.c_double, C.c_int, _dp, _dp]),
        "gfo_element_residual": (None, [C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, C.c_int, _dp, _dp]),
        "gfo_scatter_to_rank": (None, [_dp, _ip, C.c_int, C.c_int, C.c_int, _dp]),
        "gfo_gather_from_rank": (None, [_dp, _ip, C.c_int, C.c_int, _dp]),
        "gfo_newmark_predict": (None, [C.c_int, C.c_double, C.c_double, _dp, _dp, _dp, _dp]),
        "gfo_newmark_correct": (None, [C.c_int, C.c_double, C.c_double, C.c_double, _dp, _dp, _dp, _dp, _dp]),
        "gfo_apply_pml_damping": (None, [_dp, _dp, C.c_long]),
        "gfo_cpml_update_displ_fields": (None, [C.POINTER(Cpml), _dp, _dp, _dp, C.c_double]),
        "gfo_cpml_update_displ_memory": (None, [C.POINTER(Cpml)]),
        "gfo_cpml_accel_contribution": (None, [C.POINTER(Cpml), _dp, _dp, _dp, _dp]),
        "gfo_full_strain": (None, [C.c_int, _ip, _dp, _dp, C.c_int, _dp, _dp]),
        "gfo_run": (C.c_int, [C.POINTER(Problem), _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
        "gfo_run_from": (C.c_int, [C.POINTER(Problem), C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, _dp]),
        "gfo_run_state": (C.c_int, [C.POINTER(Problem), _dp, _dp, _dp]),
        "gfo_sls_coefficients": (None, [C.c_double, C.c_double, C.c_double, C.c_double, _dp]),
        "gfo_sls_point_history": (None, [C.c_int, _dp, _dp, C.c_double, C.c_double, C.c_int, _dp, _dp]),
        "gfo_cpml_recursion": (None, [C.c_double, C.c_double, _dp]),
        "gfo_cpml_point_coefficients": (None, [_dp, _dp, _dp, C.c_double, _dp]),
        "gfo_cpml_gradient": (None, [_dp, _dp, _dp, _dp]),
        "gfo_cpml_inertia": (None, [_dp, _dp, _dp, _dp, _dp]),
        "gfo_element_residual_cpml": (None, [C.c_int, _dp, _dp, _dp, _dp, _ip, _dp, _dp, _dp, C.c_int, _dp, _dp,
                                             _dp]),
        "gfo_cpml_filter_history": (None, [_dp, C.c_int, _dp, _dp, _dp, _dp, _dp]),
        "gfo_random_fill": (None, [_dp, C.c_int, C.c_uint]),
        "gfo_set_threads": (C.c_int, [C.c_int]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def P(a, ct=C.c_double):
    if a is None:
        return C.cast(None, C.POINTER(ct))
    assert a.flags.c_contiguous
    return a.ctypes.data_as(C.POINTER(ct))


def gll(N):
    x = np.zeros(N + 1)
    w = np.zeros(N + 1)
    D = np.zeros((N + 1) * (N + 1))
    assert lib.gfo_gll_nodes(N, P(x)) == 0
    assert lib.gfo_gll_weights(N, P(x), P(w)) == 0
    assert lib.gfo_gll_derivative_matrix(N, P(x), P(D)) == 0
    return x, w, D


def lagrange(xi, nodes):
    ell = np.zeros(len(nodes))
    lib.gfo_lagrange_basis(xi, len(nodes), P(np.ascontiguousarray(nodes, np.float64)), P(ell))
    return ell


def element_residual(n_elem, dxi_dx, jac, lam, mu, D, w, ngll, u, r):
    lib.gfo_element_residual(n_elem, P(dxi_dx), P(jac), P(lam), P(mu), P(D), P(w), ngll, P(u), P(r))


def scatter(loc, ibool, n_elem, n_node, n_rank_node):
    glob = np.zeros(n_rank_node * 3)
    lib.gfo_scatter_to_rank(P(loc), P(ibool, C.c_int32), n_elem, n_node, n_rank_node, P(glob))
    return glob


def gather(glob, ibool, n_elem, n_node):
    loc = np.zeros(n_elem * n_node * 3)
    lib.gfo_gather_from_rank(P(glob), P(ibool, C.c_int32), n_elem, n_node, P(loc))
    return loc


def random_fill(n, seed):
    v = np.zeros(n)
    lib.gfo_random_fill(P(v), n, seed)
    return v


class Run:
    """One forward run of the oracle on mesh arrays (reference layout)."""

    def __init__(self, arrays, n_elem, ngll, n_node, dt, nsteps, snapshot_stride, direction,
                 src_cells, src_weights, stf, rec_cell, has_cpml=True, tau_sigma=None, tau_epsilon=None,
                 cpml_complete=False):
        k = self._k = {}

        def keep(name, x, dt_=np.float64):
            if x is None:
                k[name] = None
                return None
            k[name] = np.ascontiguousarray(x, dtype=dt_)
            return k[name]

        pb = Problem()
        pb.n_elem, pb.ngll, pb.n_rank_node = n_elem, ngll, n_node
        for name in ("dxi_dx", "jacobian", "lambda_", "mu_", "mass", "density", "damping",
                     "coef_alpha", "coef_abar"):
            setattr(pb, name, P(keep(name, arrays[name])))
        pb.ibool = P(keep("ibool", arrays["ibool"], np.int32), C.c_int32)
        pb.has_cpml = 1 if has_cpml else 0
        pb.pml_region = P(keep("pml_region", arrays["pml_region"], np.int32), C.c_int32)
        sc = keep("src_cells", src_cells, np.int32)
        pb.n_src_cell = 0 if sc is None else sc.size
        pb.src_elem = P(sc, C.c_int32)
        pb.src_weights = P(keep("src_weights", src_weights))
        pb.direction = direction
        pb.nsteps = nsteps
        pb.dt = dt
        st = keep("stf", stf)
        pb.n_stf = 0 if st is None else st.size
        pb.stf = P(st)
        pb.snapshot_stride = snapshot_stride
        rc = keep("rec", rec_cell, np.int32)
        pb.n_rec_cell = 0 if rc is None else rc.size
        pb.rec_cell = P(rc, C.c_int32)
        if tau_sigma is not None:
            ts = keep("tau_sigma", tau_sigma)
            pb.n_sls = ts.size // (n_elem * ngll ** 3)
            pb.tau_sigma = P(ts)
            pb.tau_epsilon = P(keep("tau_epsilon", tau_epsilon))
        if cpml_complete:
            pb.cpml_complete = 1
            for name in ("pml_K", "pml_d", "pml_alpha"):
                setattr(pb, name, P(keep(name, arrays[name])))
        self.pb = pb
        self.nn = ngll ** 3
        self.n_node = n_node
        self.n_snap = sum(1 for s in range(nsteps) if snapshot_stride > 0 and s % snapshot_stride == 0)

    def run(self, uva=False):
        nrec = self.pb.n_rec_cell
        strain = np.zeros(self.n_snap * nrec * self.nn * 6)
        d = v = a = None
        if uva:
            d, v, a = (np.zeros(self.n_snap * nrec * self.nn * 3) for _ in range(3))
        uo, vo, ao = (np.zeros(self.n_node * 3) for _ in range(3))
        ns = lib.gfo_run(C.byref(self.pb), P(strain), P(d), P(v), P(a), P(uo), P(vo), P(ao))
        assert ns == self.n_snap, (ns, self.n_snap)
        out = {"strain": strain.reshape(self.n_snap, nrec, self.nn, 6), "u": uo, "v": vo, "a": ao}
        if uva:
            for name, x in (("disp", d), ("vel", v), ("acc", a)):
                out[name] = x.reshape(self.n_snap, nrec, self.nn, 3)
        return out

    def run_from(self, start_step, u, v, a, uva=False):
        """The reference's resume: steps [start_step, nsteps) from (u, v, a), C-PML
        state zero; records of the snapshot steps among them."""
        u, v, a = (np.ascontiguousarray(x, np.float64).copy() for x in (u, v, a))
        nrec = self.pb.n_rec_cell
        st = self.pb.snapshot_stride
        n_snap = sum(1 for s in range(start_step, self.pb.nsteps) if st > 0 and s % st == 0)
        strain = np.zeros(n_snap * nrec * self.nn * 6)
        d = vel = acc = None
        if uva:
            d, vel, acc = (np.zeros(n_snap * nrec * self.nn * 3) for _ in range(3))
        ns = lib.gfo_run_from(C.byref(self.pb), int(start_step), P(u), P(v), P(a), P(strain), P(d), P(vel), P(acc))
        assert ns == n_snap, (ns, n_snap)
        out = {"strain": strain.reshape(n_snap, nrec, self.nn, 6), "u": u, "v": v, "a": a}
        if uva:
            for name, x in (("disp", d), ("vel", vel), ("acc", acc)):
                out[name] = x.reshape(n_snap, nrec, self.nn, 3)
        return out

    def run_state(self, u, v, a):
        u, v, a = (np.ascontiguousarray(x, np.float64).copy() for x in (u, v, a))
        assert lib.gfo_run_state(C.byref(self.pb), P(u), P(v), P(a)) == 0
        return u, v, a


def set_threads(n):
    """Threads of the oracle's element loop; returns the count in effect."""
    return int(lib.gfo_set_threads(int(n)))



def sls_coefficients(tau_sigma, tau_epsilon, mu_relaxed, dt):
    """(a, b, c, Y) of one mechanism's exact step for strain linear in time."""
    out = np.zeros(4)
    lib.gfo_sls_coefficients(tau_sigma, tau_epsilon, mu_relaxed, dt, P(out))
    return out


def sls_point_history(tau_sigma, tau_epsilon, mu_relaxed, dt, dev):
    """Deviatoric stress [n-1, 5] of one point for the strain history dev [n, 5]."""
    ts = np.ascontiguousarray(tau_sigma, np.float64)
    te = np.ascontiguousarray(tau_epsilon, np.float64)
    dev = np.ascontiguousarray(dev, np.float64)
    sig = np.zeros((len(dev) - 1, 5))
    lib.gfo_sls_point_history(ts.size, P(ts), P(te), mu_relaxed, dt, len(dev), P(dev), P(sig))
    return sig


def cpml_recursion(p, dt):
    """(a, b, c) of conv(g, p) over one step for g linear in time."""
    out = np.zeros(3)
    lib.gfo_cpml_recursion(float(p), float(dt), P(out))
    return out


def cpml_point_coefficients(K, d, alpha, dt):
    """The 27 complete C-PML coefficients of one point."""
    K, d, alpha = (np.ascontiguousarray(x, np.float64) for x in (K, d, alpha))
    out = np.zeros(27)
    lib.gfo_cpml_point_coefficients(P(K), P(d), P(alpha), float(dt), P(out))
    return out


def cpml_filter_history(coef, g=None, u=None, v=None):
    """One point from zero memories: gradient history g [n,9] -> G [n,3,9]
    (column j, entry k*3+l) and/or (u, v) [n,3] -> inertia L [n,3]."""
    coef = np.ascontiguousarray(coef, np.float64)
    n = len(g) if g is not None else len(u)
    G = np.zeros((n, 27)) if g is not None else None
    L = np.zeros((n, 3)) if u is not None else None
    g = None if g is None else np.ascontiguousarray(g, np.float64)
    u = None if u is None else np.ascontiguousarray(u, np.float64)
    v = None if v is None else np.ascontiguousarray(v, np.float64)
    lib.gfo_cpml_filter_history(P(coef), n, P(g), P(u), P(v), P(G), P(L))
    return (None if G is None else G.reshape(n, 3, 9)), L


def element_residual_cpml(n_elem, dxi_dx, jac, lam, mu, pml_region, coef, D, w, ngll, u, carry, r):
    lib.gfo_element_residual_cpml(n_elem, P(dxi_dx), P(jac), P(lam), P(mu), P(pml_region, C.c_int32), P(coef), P(D),
                                  P(w), ngll, P(u), P(carry), P(r))
