"""ctypes bindings of libgfx.so (the MI355X elastic CG-SEM solver C ABI).

The library is built in-tree by ``make -C gf-only-sem_amd``; importing this
module fails loudly when it is missing (there is no CPU fallback).
"""
from __future__ import annotations

import ctypes as C
import dataclasses
import os
from dataclasses import dataclass, field

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# GFX_LIB selects another build of the library (the host-sanitizer build, for CPU checks)
LIB_PATH = os.environ.get("GFX_LIB") or os.path.join(os.path.dirname(_HERE), "libgfx.so")

GFX_ASSEMBLY_COLORED = 0
GFX_ASSEMBLY_ATOMIC = 1
GFX_ASSEMBLY_GATHER = 2
GFX_ELEMENT_AUTO = 0
GFX_ELEMENT_BLOCK = 1
GFX_ELEMENT_WAVE = 2
GFX_ELEMENT_PAIR = 3
GFX_ELEMENT_PENCIL = 4
MAX_LAYERS = 8

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class GfxError(RuntimeError):
    def __init__(self, status: int, msg: str):
        super().__init__(f"gfx status {status}: {msg}")
        self.status = status


class BoxSpec(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("lx", C.c_double), ("ly", C.c_double), ("lz", C.c_double),
        ("order", C.c_int),
        ("n_layers", C.c_int),
        ("layer_zbot", C.c_double * MAX_LAYERS),
        ("vp", C.c_double * MAX_LAYERS),
        ("vs", C.c_double * MAX_LAYERS),
        ("rho", C.c_double * MAX_LAYERS),
        ("rf_amp", C.c_double), ("rf_corr_len", C.c_double),
        ("rf_modes", C.c_int),
        ("rf_seed", C.c_uint64),
        ("pml_xmin", C.c_int), ("pml_xmax", C.c_int), ("pml_ymin", C.c_int),
        ("pml_ymax", C.c_int), ("pml_zmax", C.c_int),
        ("f0_pml", C.c_double),
        ("jitter", C.c_double),
        ("jitter_seed", C.c_uint64),
        ("cfl_safety", C.c_double), ("output_dt", C.c_double), ("duration", C.c_double),
        ("record_depth_max", C.c_double),
        ("want_coords", C.c_int),
        ("want_beta", C.c_int),
        ("want_strain", C.c_int),
        ("n_sls", C.c_int),
        ("q_mu", C.c_double * MAX_LAYERS),
        ("att_fmin", C.c_double), ("att_fmax", C.c_double),
        ("want_profiles", C.c_int),
        ("pml_profile", C.c_int),
        ("pml_reflection", C.c_double),
    ]


class _Mesh(C.Structure):
    _fields_ = [
        ("n_elem", C.c_int), ("ngll", C.c_int), ("n_node_glob", C.c_int),
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("xmin", C.c_double), ("xmax", C.c_double), ("ymin", C.c_double),
        ("ymax", C.c_double), ("zmin", C.c_double), ("zmax", C.c_double),
        ("coords", _dp), ("jacobian", _dp), ("dxi_dx", _dp), ("mass", _dp),
        ("vp", _dp), ("vs", _dp), ("density", _dp), ("lambda_", _dp), ("mu_", _dp),
        ("damping", _dp), ("ibool", _ip), ("node_xyz", _dp),
        ("pml_region", _ip), ("coef_alpha", _dp), ("coef_beta", _dp), ("coef_strain", _dp), ("coef_abar", _dp),
        ("n_pml_elem", C.c_int),
        ("h_min", C.c_double), ("vp_max", C.c_double), ("cfl_dt", C.c_double), ("dt", C.c_double),
        ("snapshot_stride", C.c_int), ("nsteps", C.c_int),
        ("record_depth_actual", C.c_double),
        ("n_rec_cell", C.c_int),
        ("rec_cell", _ip),
        ("n_sls", C.c_int),
        ("tau_sigma", _dp),
        ("tau_epsilon", _dp),
        ("pml_K", _dp), ("pml_d", _dp), ("pml_alpha", _dp),
    ]


class SolverDesc(C.Structure):
    _fields_ = [
        ("n_elem", C.c_int), ("ngll", C.c_int), ("n_rank_node", C.c_int),
        ("dxi_dx", _dp), ("jacobian", _dp), ("lambda_", _dp), ("mu_", _dp),
        ("mass", _dp), ("density", _dp), ("damping", _dp), ("ibool", _ip),
        ("elem_centroid", _dp),
        ("has_cpml", C.c_int), ("pml_region", _ip), ("coef_alpha", _dp), ("coef_abar", _dp),
        ("dt", C.c_double), ("snapshot_stride", C.c_int),
        ("n_rec_cell", C.c_int), ("rec_cell", _ip),
        ("n_fields", C.c_int), ("assembly", C.c_int), ("elems_per_block", C.c_int),
        ("element_kernel", C.c_int),
        ("record_capacity", C.c_int), ("record_uva", C.c_int),
        ("node_field_groups", C.c_int),
        ("general_geometry", C.c_int),
        ("n_sls", C.c_int), ("tau_sigma", _dp), ("tau_epsilon", _dp),
        ("cpml_complete", C.c_int), ("pml_K", _dp), ("pml_d", _dp), ("pml_alpha", _dp),
    ]


class ForwardOptions(C.Structure):
    _fields_ = [
        ("workdir", C.c_char_p), ("directions", C.c_char_p), ("stations", C.c_char_p),
        ("resume", C.c_int), ("record_uva", C.c_int), ("assembly", C.c_int),
        ("element_kernel", C.c_int), ("elems_per_block", C.c_int), ("ring", C.c_int),
        ("stop_after_step", C.c_int), ("restart_stride", C.c_int), ("verbose", C.c_int),
        ("cpml_complete", C.c_int),
    ]


class GreenTileOptions(C.Structure):
    _fields_ = [
        ("workdir", C.c_char_p), ("output_dir", C.c_char_p), ("model_path", C.c_char_p),
        ("with_uva", C.c_int), ("assembly", C.c_int), ("element_kernel", C.c_int),
        ("elems_per_block", C.c_int), ("ring", C.c_int), ("verbose", C.c_int),
        ("cpml_complete", C.c_int),
    ]


class MlpConfig(C.Structure):
    _fields_ = [("n_freq", C.c_int), ("width", C.c_int), ("depth", C.c_int), ("n_out", C.c_int),
                ("sigma", C.c_float), ("seed", C.c_uint64), ("lr", C.c_float), ("beta1", C.c_float),
                ("beta2", C.c_float), ("eps", C.c_float), ("batch", C.c_int), ("lr_final", C.c_float)]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is not built; run `make -C gf-only-sem_amd` (or __graft_entry__.build())")
    lib = C.CDLL(LIB_PATH)
    st = C.c_int
    vp = C.c_void_p
    sig = {
        "gfx_last_error": (C.c_char_p, []),
        "gfx_version": (C.c_char_p, []),
        "gfx_device_count": (C.c_int, []),
        "gfx_compute_element_residual": (st, [C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, C.c_int, _dp, _dp]),
        "gfx_compute_element_residual_f32": (st, [C.c_int, _dp, _dp, _dp, _dp, _dp, _dp, C.c_int, _dp, _dp]),
        "gfx_solver_create": (st, [C.POINTER(SolverDesc), C.POINTER(vp)]),
        "gfx_cpml_rigid_nodes": (st, [C.POINTER(SolverDesc), C.POINTER(C.c_uint8)]),
        "gfx_solver_destroy": (st, [vp]),
        "gfx_solver_set_source": (st, [vp, C.c_int, C.c_int, C.c_int, _ip, _dp, C.c_int, _dp]),
        "gfx_solver_reset": (st, [vp]),
        "gfx_solver_run": (st, [vp, C.c_int]),
        "gfx_solver_step_index": (C.c_int, [vp]),
        "gfx_solver_get_state": (st, [vp, C.c_int, _dp, _dp, _dp]),
        "gfx_solver_set_state": (st, [vp, C.c_int, _dp, _dp, _dp]),
        "gfx_solver_n_snapshots": (C.c_int, [vp]),
        "gfx_step_newmark_predict": (st, [vp, C.c_double, C.c_double]),
        "gfx_step_gather_predicted": (st, [vp]),
        "gfx_step_gather_from_rank": (st, [vp]),
        "gfx_step_zero_residual": (st, [vp]),
        "gfx_step_element_residual": (st, [vp, C.c_int, C.c_int]),
        "gfx_step_pml_damping": (st, [vp]),
        "gfx_step_source_injection": (st, [vp, C.c_int, C.c_int, C.c_double, _dp, C.c_int]),
        "gfx_step_scatter_to_rank": (st, [vp]),
        "gfx_step_newmark_correct": (st, [vp, C.c_double, C.c_double, C.c_double]),
        "gfx_step_copy_state_to_host": (st, [vp, C.c_int, _dp, _dp, _dp]),
        "gfx_step_copy_utilde_to_host": (st, [vp, C.c_int, _dp]),
        "gfx_step_copy_utilde_from_host": (st, [vp, C.c_int, _dp]),
        "gfx_step_copy_residual_to_host": (st, [vp, C.c_int, _dp]),
        "gfx_step_copy_residual_from_host": (st, [vp, C.c_int, _dp]),
        "gfx_solver_get_record": (st, [vp, C.c_int, C.c_int, _dp, _dp, _dp, _dp]),
        "gfx_solver_kernel_time": (st, [vp, C.c_char_p, C.POINTER(C.c_double), C.POINTER(C.c_int)]),
        "gfx_solver_set_timing": (st, [vp, C.c_int]),
        "gfx_solver_algorithmic_bytes": (C.c_double, [vp, C.c_int]),
        "gfx_solver_query": (st, [vp, C.c_char_p, C.POINTER(C.c_double)]),
        "gfx_solver_set_launch": (st, [vp, C.c_int, C.c_int]),
        "gfx_solver_set_node_groups": (st, [vp, C.c_int]),
        "gfx_solver_set_graph": (st, [vp, C.c_int]),
        "gfx_batch_create": (st, [C.c_int, C.POINTER(vp), C.POINTER(vp)]),
        "gfx_batch_run": (st, [vp, C.c_int, _ip]),
        "gfx_batch_set_groups": (st, [vp, C.c_int]),
        "gfx_batch_groups": (C.c_int, [vp]),
        "gfx_batch_destroy": (st, [vp]),
        "gfx_green_create": (st, [C.c_int, C.c_int, _dp, C.c_int, _ip, C.POINTER(vp)]),
        "gfx_green_destroy": (st, [vp]),
        "gfx_green_set_values": (st, [vp, C.c_int, C.POINTER(C.c_float)]),
        "gfx_green_get_values": (st, [vp, C.c_int, C.POINTER(C.c_float)]),
        "gfx_green_set_capture_map": (st, [vp, C.POINTER(C.c_int64), _ip, _dp]),
        "gfx_green_capture": (st, [vp, vp, C.c_int, _ip, C.c_int]),
        "gfx_green_interpolate": (st, [vp, C.c_int, _dp, _ip, C.POINTER(C.c_float), _ip]),
        "gfx_gll_interpolate_host": (st, [C.c_int, _dp, C.c_int, _ip, C.c_int, _dp, _dp, _dp]),
        "gfx_mlp_config_default": (st, [C.POINTER(MlpConfig)]),
        "gfx_mlp_create": (st, [C.POINTER(MlpConfig), _dp, _dp, _dp, C.POINTER(vp)]),
        "gfx_mlp_destroy": (st, [vp]),
        "gfx_mlp_n_params": (st, [vp, C.POINTER(C.c_size_t)]),
        "gfx_mlp_get_params": (st, [vp, C.POINTER(C.c_float), C.c_size_t, C.POINTER(C.c_float)]),
        "gfx_mlp_set_params": (st, [vp, C.POINTER(C.c_float), C.c_size_t]),
        "gfx_mlp_forward": (st, [vp, C.c_int, _dp, C.POINTER(C.c_float)]),
        "gfx_mlp_gradient": (st, [vp, C.c_int, _dp, C.POINTER(C.c_float), C.POINTER(C.c_float),
                                  C.POINTER(C.c_double)]),
        "gfx_mlp_train": (st, [vp, vp, C.c_double, C.c_double, C.c_int, C.c_uint64, C.POINTER(C.c_double)]),
        "gfx_mlp_set_envelope": (st, [vp, C.c_double, C.c_double, C.c_double, C.c_double, C.c_int,
                                      C.POINTER(C.c_float)]),
        "gfx_surrogate_query": (st, [vp, vp, C.POINTER(C.c_uint8), C.c_int, C.c_double, C.c_double, C.c_int, _dp,
                                     _ip, C.POINTER(C.c_float), C.POINTER(C.c_int)]),
        "gfx_autotune": (st, [C.POINTER(SolverDesc), C.c_int, C.POINTER(SolverDesc), C.POINTER(C.c_double)]),
        "gfx_box_spec_default": (st, [C.POINTER(BoxSpec)]),
        "gfx_mesh_build_box": (st, [C.POINTER(BoxSpec), C.POINTER(C.POINTER(_Mesh))]),
        "gfx_mesh_free": (st, [C.POINTER(_Mesh)]),
        "gfx_mesh_locate_source": (st, [C.POINTER(_Mesh), C.c_double, C.c_double, C.c_double,
                                        C.c_int, _ip, _dp, C.POINTER(C.c_int)]),
        "gfx_ricker": (st, [C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, _dp]),
        "gfx_sls_fit": (st, [C.c_double, C.c_double, C.c_double, C.c_int, _dp, _dp]),
        "gfx_pml_strain_coefficients": (st, [C.c_int, _dp, _dp, _dp, _dp]),
        "gfx_solver_set_step": (st, [vp, C.c_int]),
        "gfx_solver_checkpoint_size": (st, [vp, C.POINTER(C.c_size_t)]),
        "gfx_solver_checkpoint_save": (st, [vp, vp, C.c_size_t]),
        "gfx_solver_checkpoint_load": (st, [vp, vp, C.c_size_t]),
        "gfx_run_forward": (C.c_int, [C.c_char_p, C.c_int, C.c_int]),
        "gfx_forward_options_default": (st, [C.POINTER(ForwardOptions)]),
        "gfx_run_forward_ex": (st, [C.POINTER(ForwardOptions)]),
        "gfx_mesh_write_inputs": (st, [C.POINTER(_Mesh), C.c_char_p, C.c_int, C.c_double, C.c_double,
                                       C.c_double, C.c_int, _dp, C.c_char_p, C.c_int]),
        "gfx_mesh_read_inputs": (st, [C.c_char_p, C.POINTER(C.POINTER(_Mesh))]),
        "gfx_record_write_step": (st, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_int, _dp, _dp, _dp, _dp]),
        "gfx_write_stations": (st, [C.c_char_p, C.POINTER(_Mesh), C.c_int, _dp]),
        "gfx_h5_available": (C.c_int, []),
        "gfx_h5_describe": (st, [C.c_char_p, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]),
        "gfx_h5_read_f64": (st, [C.c_char_p, C.c_char_p, _dp, C.c_size_t, C.POINTER(C.c_size_t)]),
        "gfx_h5_read_attr_f64": (st, [C.c_char_p, C.c_char_p, C.c_char_p, _dp]),
        "gfx_h5_read_attr_str": (st, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, C.c_size_t]),
        "gfx_h5_write": (st, [C.c_char_p, C.c_int, C.c_char_p, C.c_char_p, C.c_int, C.POINTER(C.c_uint64), vp]),
        "gfx_h5_write_attr": (st, [C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, vp]),
        "gfx_config_summary": (st, [C.c_char_p, _dp, C.c_int, C.c_char_p, C.c_char_p, C.c_size_t]),
        "gfx_partition_summary": (st, [C.c_char_p, C.c_int, _dp, C.c_int]),
        "gfx_h5_read_attr_f64_array": (st, [C.c_char_p, C.c_char_p, C.c_char_p, _dp, C.c_size_t,
                                            C.POINTER(C.c_size_t)]),
        "gfx_green_tile_options_default": (st, [C.POINTER(GreenTileOptions)]),
        "gfx_run_green_tiles": (st, [C.POINTER(GreenTileOptions), C.POINTER(C.c_int)]),
        "gfx_green_slab_create": (st, [C.c_int, C.POINTER(C.c_int64), _ip, _dp, C.c_int64, _ip, C.c_int, C.c_int,
                                       C.POINTER(vp)]),
        "gfx_green_slab_destroy": (st, [vp]),
        "gfx_green_slab_capture": (st, [vp, vp, C.c_int, _ip, vp, vp]),
        "gfx_gmsh_to_topology": (st, [C.c_char_p, C.c_char_p, C.c_char_p]),
        "gfx_mesh_build_topology": (st, [C.c_char_p, C.POINTER(BoxSpec), C.POINTER(C.POINTER(_Mesh))]),
        "gfx_mesh_build_hex": (st, [C.c_int64, _dp, C.c_int64, C.POINTER(C.c_int64), C.POINTER(BoxSpec),
                                    C.POINTER(C.POINTER(_Mesh))]),
        "gfx_topology_summary": (st, [C.c_char_p, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.c_int64]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def _check(status: int):
    if status != 0:
        raise GfxError(status, lib.gfx_last_error().decode())


def _ptr(a: np.ndarray | None, ct=C.c_double):
    if a is None:
        return C.cast(None, C.POINTER(ct))
    return a.ctypes.data_as(C.POINTER(ct))


def device_count() -> int:
    return int(lib.gfx_device_count())


def element_residual(dxi_dx, jacobian, lambda_, mu_, D, weights, ngll, u, r, fp32=False):
    """gf::compute_element_residual<BackendHIP>: r (float64, in/out) -= K u."""
    arrs = [np.ascontiguousarray(x, dtype=np.float64) for x in (dxi_dx, jacobian, lambda_, mu_, D, weights, u)]
    assert r.dtype == np.float64 and r.flags.c_contiguous
    n_elem = arrs[1].size // (ngll ** 3)
    fn = lib.gfx_compute_element_residual_f32 if fp32 else lib.gfx_compute_element_residual
    _check(fn(n_elem, *[_ptr(a) for a in arrs[:6]], ngll, _ptr(arrs[6]), _ptr(r)))


def sls_fit(q, f_min, f_max, n_sls):
    """(tau_sigma, tau_epsilon) of n_sls standard linear solids for a constant q."""
    ts, te = np.zeros(n_sls), np.zeros(n_sls)
    _check(lib.gfx_sls_fit(q, f_min, f_max, n_sls, _ptr(ts), _ptr(te)))
    return ts, te


def pml_strain_coefficients(region, K, d, alpha):
    """The 18 strain-based C-PML coefficients of one point (gfx_pml_strain_coefficients)."""
    k, dd, a = (np.ascontiguousarray(x, np.float64) for x in (K, d, alpha))
    out = np.zeros(18)
    _check(lib.gfx_pml_strain_coefficients(int(region), _ptr(k), _ptr(dd), _ptr(a), _ptr(out)))
    return out


def ricker(f0, t0, amplitude, dt, n):
    out = np.zeros(n, dtype=np.float64)
    _check(lib.gfx_ricker(f0, t0, amplitude, dt, n, _ptr(out)))
    return out


@dataclass
class Mesh:
    """Structured-box inputs in the reference partition layout (numpy copies)."""
    n_elem: int
    ngll: int
    n_node: int
    nx: int
    ny: int
    nz: int
    bounds: tuple
    arrays: dict
    dt: float
    snapshot_stride: int
    nsteps: int
    h_min: float
    vp_max: float
    cfl_dt: float
    n_pml_elem: int
    record_depth_actual: float
    rec_cell: np.ndarray
    n_sls: int = 0
    _handle: object = field(default=None, repr=False)
    _parent: object = field(default=None, repr=False)

    def __getattr__(self, k):
        arrays = self.__dict__.get("arrays", {})
        if k in arrays:
            return arrays[k]
        raise AttributeError(k)

    def centroids(self):
        """Element centroids [E,3] from the 8 corner nodes."""
        corners = np.array([0, 4, 20, 24, 100, 104, 120, 124])
        ib = self.arrays["ibool"].reshape(self.n_elem, 125)[:, corners]
        xyz = self.arrays["node_xyz"].reshape(-1, 3)
        return np.ascontiguousarray(xyz[ib].mean(axis=1))

    def derive(self, rec_cell=None, **arrays):
        """Same mesh with some arrays (same shapes) or the recording cells replaced.
        Source location uses the geometry, which is unchanged, through this mesh."""
        a = dict(self.arrays)
        for k, v in arrays.items():
            if k not in a or a[k] is None or np.asarray(v).size != a[k].size:
                raise KeyError(f"{k}: not an array of this mesh, or a different size")
            a[k] = np.ascontiguousarray(v, dtype=a[k].dtype).reshape(a[k].shape)
        rec = self.rec_cell if rec_cell is None else np.ascontiguousarray(rec_cell, np.int32)
        return dataclasses.replace(self, arrays=a, rec_cell=rec, _handle=None, _parent=self)

    def locate_source(self, x, y, z, max_cells=8):
        if self._handle is None and self._parent is not None:
            return self._parent.locate_source(x, y, z, max_cells)
        nn = self.ngll ** 3
        cells = np.zeros(max_cells, dtype=np.int32)
        w = np.zeros(max_cells * nn, dtype=np.float64)
        n = C.c_int(0)
        _check(lib.gfx_mesh_locate_source(self._handle, x, y, z, max_cells, _ptr(cells, C.c_int32),
                                          _ptr(w), C.byref(n)))
        return cells[: n.value].copy(), w[: n.value * nn].copy()

    def close(self):
        if self._handle is not None:
            lib.gfx_mesh_free(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def default_spec() -> BoxSpec:
    s = BoxSpec()
    _check(lib.gfx_box_spec_default(C.byref(s)))
    return s


def make_spec(**kw) -> BoxSpec:
    """Box spec from defaults; `layers` = [(zbot, vp, vs, rho), ...] (last zbot ignored)."""
    s = default_spec()
    layers = kw.pop("layers", None)
    q_mu = kw.pop("q_mu", None)
    if q_mu is not None:
        for l, q in enumerate(q_mu):
            s.q_mu[l] = q
    for k, v in kw.items():
        if not hasattr(s, k):
            raise KeyError(k)
        setattr(s, k, v)
    if layers is not None:
        s.n_layers = len(layers)
        for l, (zb, vp, vs, rho) in enumerate(layers):
            s.layer_zbot[l] = zb
            s.vp[l] = vp
            s.vs[l] = vs
            s.rho[l] = rho
    return s


def build_box(spec: BoxSpec | None = None, keep_handle=True, copy=True, **kw) -> Mesh:
    """Build a box mesh.  copy=False returns zero-copy views that are valid only
    while the returned Mesh is alive (use for very large meshes)."""
    if spec is None:
        spec = make_spec(**kw)
    h = C.POINTER(_Mesh)()
    _check(lib.gfx_mesh_build_box(C.byref(spec), C.byref(h)))
    return _wrap_mesh(h, keep_handle, copy)


def gmsh_to_topology(msh_path, model_path, aux_path=None):
    """tools/gmsh_to_hdf5.py: an ASCII GMSH hexahedral mesh to model.h5 /topology."""
    _check(lib.gfx_gmsh_to_topology(os.fsencode(msh_path), os.fsencode(model_path),
                                    None if aux_path is None else os.fsencode(aux_path)))


def build_topology(model_path, spec: BoxSpec | None = None, keep_handle=True, copy=True, **kw) -> Mesh:
    """The preprocessor's pipeline on model.h5 /topology (general hexahedra in
    x-fastest structured order; spec nx, ny give the C-PML flag grid)."""
    if spec is None:
        spec = make_spec(**kw)
    h = C.POINTER(_Mesh)()
    _check(lib.gfx_mesh_build_topology(os.fsencode(model_path), C.byref(spec), C.byref(h)))
    return _wrap_mesh(h, keep_handle, copy)


def build_hex(xyz, hexes, spec: BoxSpec | None = None, keep_handle=True, copy=True, **kw) -> Mesh:
    """The same from points [n, 3] and hexahedra [n_cell, 8] (0-based, GMSH corner order)."""
    if spec is None:
        spec = make_spec(**kw)
    xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
    hexes = np.ascontiguousarray(hexes, np.int64).reshape(-1, 8)
    h = C.POINTER(_Mesh)()
    _check(lib.gfx_mesh_build_hex(len(xyz), _ptr(xyz), len(hexes), hexes.ctypes.data_as(C.POINTER(C.c_int64)),
                                  C.byref(spec), C.byref(h)))
    return _wrap_mesh(h, keep_handle, copy)


def topology_summary(model_path):
    """(n_vertex, n_edge, n_surface, n_cell) of model.h5 /topology and each cell's
    1-based corner vertex ids [n_cell, 8] in GMSH order."""
    counts = np.zeros(4, np.int64)
    _check(lib.gfx_topology_summary(os.fsencode(model_path), counts.ctypes.data_as(C.POINTER(C.c_int64)),
                                    None, 0))
    corners = np.zeros((int(counts[3]), 8), np.int64)
    _check(lib.gfx_topology_summary(os.fsencode(model_path), counts.ctypes.data_as(C.POINTER(C.c_int64)),
                                    corners.ctypes.data_as(C.POINTER(C.c_int64)), len(corners)))
    return tuple(int(c) for c in counts), corners


def read_inputs(directory, keep_handle=True, copy=True) -> Mesh:
    """Load config.h5 + partitions/partition_{r}.h5 (the reference's input
    files, merged as read_partition_all does) as a Mesh."""
    h = C.POINTER(_Mesh)()
    _check(lib.gfx_mesh_read_inputs(os.fsencode(directory), C.byref(h)))
    return _wrap_mesh(h, keep_handle, copy)


def _wrap_mesh(h, keep_handle, copy) -> Mesh:
    m = h.contents
    E, ng = m.n_elem, m.ngll
    nn = ng ** 3
    P = E * nn
    G = m.n_node_glob

    def arr(p, n, dt=np.float64):
        # views over the library's buffers (valid until Mesh.close())
        if not p or n == 0:
            return None
        a = np.ctypeslib.as_array(p, shape=(n,))
        return a.copy() if copy else a

    arrays = {
        "coords": arr(m.coords, P * 3),
        "jacobian": arr(m.jacobian, P),
        "dxi_dx": arr(m.dxi_dx, P * 9),
        "mass": arr(m.mass, P),
        "vp": arr(m.vp, P),
        "vs": arr(m.vs, P),
        "density": arr(m.density, P),
        "lambda_": arr(m.lambda_, P),
        "mu_": arr(m.mu_, P),
        "damping": arr(m.damping, P),
        "ibool": arr(m.ibool, P, np.int32),
        "node_xyz": arr(m.node_xyz, G * 3),
        "pml_region": arr(m.pml_region, E, np.int32),
        "coef_alpha": arr(m.coef_alpha, P * 9),
        "coef_beta": arr(m.coef_beta, P * 9),
        "coef_strain": arr(m.coef_strain, P * 18),
        "coef_abar": arr(m.coef_abar, P * 5),
        "tau_sigma": arr(m.tau_sigma, P * m.n_sls),
        "tau_epsilon": arr(m.tau_epsilon, P * m.n_sls),
        "pml_K": arr(m.pml_K, P * 3),
        "pml_d": arr(m.pml_d, P * 3),
        "pml_alpha": arr(m.pml_alpha, P * 3),
    }
    rec = arr(m.rec_cell, m.n_rec_cell, np.int32)
    rec = None if rec is None else rec.copy()
    mesh = Mesh(n_elem=E, ngll=ng, n_node=G, nx=m.nx, ny=m.ny, nz=m.nz,
                bounds=(m.xmin, m.xmax, m.ymin, m.ymax, m.zmin, m.zmax), arrays=arrays,
                dt=m.dt, snapshot_stride=m.snapshot_stride, nsteps=m.nsteps, h_min=m.h_min,
                vp_max=m.vp_max, cfl_dt=m.cfl_dt, n_pml_elem=m.n_pml_elem,
                record_depth_actual=m.record_depth_actual, n_sls=m.n_sls,
                rec_cell=rec if rec is not None else np.zeros(0, np.int32), _handle=h)
    if not keep_handle:
        if not copy:
            mesh.arrays = {k: (None if v is None else v.copy()) for k, v in arrays.items()}
        mesh.close()
    return mesh


def write_inputs(mesh: Mesh, directory, src_xyz, stf, n_parts=1, precision="float32", restart_stride=0):
    """Write config.h5 + partitions/partition_{r}.h5 for this mesh and a point
    force at src_xyz (the files the reference solver reads)."""
    stf = np.ascontiguousarray(stf, np.float64)
    _check(lib.gfx_mesh_write_inputs(mesh._handle, os.fsencode(directory), n_parts, *map(float, src_xyz),
                                     stf.size, _ptr(stf), precision.encode(), restart_stride))


def write_stations(path, mesh: Mesh, xyz):
    xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
    _check(lib.gfx_write_stations(os.fsencode(path), mesh._handle, xyz.shape[0], _ptr(xyz)))


def write_record(inputs_dir, out_dir, direction, step, strain, disp=None, vel=None, acc=None):
    """One record file in the reference schema for the recording map of inputs_dir."""
    arrs = [None if x is None else np.ascontiguousarray(x, np.float64) for x in (strain, disp, vel, acc)]
    _check(lib.gfx_record_write_step(os.fsencode(inputs_dir), os.fsencode(out_dir), direction.encode(), step,
                                     *[_ptr(a) for a in arrs]))


def run_forward(workdir=None, directions="z", stations=None, resume=False, record_uva=True, assembly=-1,
                element_kernel=GFX_ELEMENT_AUTO, elems_per_block=0, ring=8, stop_after_step=-1,
                restart_stride=-1, verbose=False, cpml_complete=False):
    """The forward driver over the reference's files (gfx_run_forward_ex)."""
    o = ForwardOptions()
    _check(lib.gfx_forward_options_default(C.byref(o)))
    keep = [None if workdir is None else os.fsencode(workdir), directions.encode(),
            None if stations is None else os.fsencode(stations)]
    o.workdir, o.directions, o.stations = keep
    o.resume = 1 if resume else 0
    o.record_uva = 1 if record_uva else 0
    o.assembly = assembly
    o.element_kernel = element_kernel
    o.elems_per_block = elems_per_block
    o.ring = ring
    o.stop_after_step = stop_after_step
    o.restart_stride = restart_stride
    o.verbose = 1 if verbose else 0
    o.cpml_complete = 1 if cpml_complete else 0
    _check(lib.gfx_run_forward_ex(C.byref(o)))


def run_green_tiles(workdir=None, output_dir=None, model_path=None, with_uva=True, assembly=-1,
                    element_kernel=GFX_ELEMENT_AUTO, elems_per_block=0, ring=8, verbose=False,
                    cpml_complete=False) -> int:
    """Green's function tiles straight from the device (gfx_run_green_tiles); returns the tile count."""
    o = GreenTileOptions()
    _check(lib.gfx_green_tile_options_default(C.byref(o)))
    keep = [None if x is None else os.fsencode(x) for x in (workdir, output_dir, model_path)]
    o.workdir, o.output_dir, o.model_path = keep
    o.with_uva = 1 if with_uva else 0
    o.assembly = assembly
    o.element_kernel = element_kernel
    o.elems_per_block = elems_per_block
    o.ring = ring
    o.verbose = 1 if verbose else 0
    o.cpml_complete = 1 if cpml_complete else 0
    n = C.c_int(0)
    _check(lib.gfx_run_green_tiles(C.byref(o), C.byref(n)))
    return n.value


def h5_available() -> bool:
    return bool(lib.gfx_h5_available())


def h5_describe(path) -> str:
    need = C.c_size_t(0)
    _check(lib.gfx_h5_describe(os.fsencode(path), None, 0, C.byref(need)))
    buf = C.create_string_buffer(need.value + 1)
    _check(lib.gfx_h5_describe(os.fsencode(path), buf, need.value + 1, C.byref(need)))
    return buf.value.decode()


def h5_read(path, dataset) -> np.ndarray:
    n = C.c_size_t(0)
    _check(lib.gfx_h5_read_f64(os.fsencode(path), dataset.encode(), _ptr(None), 0, C.byref(n)))
    out = np.zeros(n.value)
    _check(lib.gfx_h5_read_f64(os.fsencode(path), dataset.encode(), _ptr(out), out.size, C.byref(n)))
    return out


_H5T = {np.dtype(np.float64): "f64", np.dtype(np.float32): "f32", np.dtype(np.int32): "i32",
        np.dtype(np.int64): "i64"}


def h5_write(path, dataset, data, append=True):
    """Write a numpy array as a dataset (parent groups created)."""
    a = np.ascontiguousarray(data)
    dims = (C.c_uint64 * max(a.ndim, 1))(*(a.shape if a.ndim else (1,)))
    _check(lib.gfx_h5_write(os.fsencode(path), 1 if append else 0, dataset.encode(), _H5T[a.dtype].encode(),
                            max(a.ndim, 1), dims, a.ctypes.data_as(C.c_void_p)))


def h5_write_attr(path, obj, name, value):
    """Scalar attribute: str, int (int32) or float (float64)."""
    if isinstance(value, str):
        b = C.create_string_buffer(value.encode())
        _check(lib.gfx_h5_write_attr(os.fsencode(path), obj.encode(), name.encode(), b"str", C.cast(b, C.c_void_p)))
        return
    if isinstance(value, (int, np.integer)) and not isinstance(value, bool):
        v, t = C.c_int32(int(value)), b"i32"
    else:
        v, t = C.c_double(float(value)), b"f64"
    _check(lib.gfx_h5_write_attr(os.fsencode(path), obj.encode(), name.encode(), t, C.cast(C.byref(v), C.c_void_p)))


_CONFIG_KEYS = ("polynomial_order", "solver_dt", "output_dt_s", "snapshot_stride", "nsteps", "cfl_safety",
                "xmin", "xmax", "ymin", "ymax", "zmin", "zmax", "n_stf_t", "n_stf_values", "source_x",
                "source_y", "source_z", "n_src_cell", "record_depth_max_m", "record_depth_actual_m",
                "restart_stride", "restart_dt_s", "log_stride", "n_tilex")
_PART_KEYS = ("n_local_cell", "ngll", "n_rank_node", "has_cpml", "has_recording", "n_rec_cell", "n_unique_gll",
              "n_local_cell_ids", "n_exchange_neighbors", "n_jacobian", "n_local_cell2rank_node",
              "n_local_cell2global_node")


def read_config(path) -> dict:
    """What the config reader makes of config.h5 (defaults and legacy layout included)."""
    out = np.zeros(24)
    t, p = C.create_string_buffer(1024), C.create_string_buffer(64)
    _check(lib.gfx_config_summary(os.fsencode(path), _ptr(out), 24, t, p, 64))
    d = dict(zip(_CONFIG_KEYS, out.tolist()))
    d["title"], d["snapshot_precision"] = t.value.decode(), p.value.decode()
    return d


def read_partition(path, merge=False) -> dict:
    """What the partition reader makes of one file, or of a directory merged."""
    out = np.zeros(12)
    _check(lib.gfx_partition_summary(os.fsencode(path), 1 if merge else 0, _ptr(out), 12))
    return dict(zip(_PART_KEYS, out.tolist()))


def h5_attr(path, obj, name, string=False):
    if string:
        buf = C.create_string_buffer(4096)
        _check(lib.gfx_h5_read_attr_str(os.fsencode(path), obj.encode(), name.encode(), buf, 4096))
        return buf.value.decode()
    v = C.c_double(0)
    _check(lib.gfx_h5_read_attr_f64(os.fsencode(path), obj.encode(), name.encode(), C.byref(v)))
    return v.value


def h5_attr_array(path, obj, name) -> np.ndarray:
    """Every element of a numeric attribute as float64."""
    n = C.c_size_t(0)
    _check(lib.gfx_h5_read_attr_f64_array(os.fsencode(path), obj.encode(), name.encode(), _ptr(None), 0,
                                          C.byref(n)))
    out = np.zeros(n.value)
    _check(lib.gfx_h5_read_attr_f64_array(os.fsencode(path), obj.encode(), name.encode(), _ptr(out), out.size,
                                          C.byref(n)))
    return out


def h5_layout(path) -> dict:
    """{path: (type, shape)} for every dataset and {path@attr: (type, shape)} for every attribute."""
    out = {}
    for line in h5_describe(path).splitlines():
        name, typ, dims = line.rsplit(" ", 2)
        shape = tuple(int(v) for v in dims.strip("[]").split(",") if v)
        out[name] = (typ, shape)
    return out


def make_desc(mesh: Mesh, keep: list, n_fields=1, assembly=GFX_ASSEMBLY_COLORED, elems_per_block=0,
              record_capacity=None, record_uva=False, has_cpml=True, dt=None,
              snapshot_stride=None, rec_cell=None, element_kernel=GFX_ELEMENT_AUTO, morton=True,
              node_field_groups=0, general_geometry=False, attenuation=True, cpml_complete=False):
    """Solver descriptor over mesh arrays; arrays it points to are appended to keep.
    cpml_complete: the complete C-PML from the mesh's pml_K, pml_d, pml_alpha."""
    a = mesh.arrays

    def kp(x, dt_=np.float64):
        if x is None:
            return None
        x = np.ascontiguousarray(x, dtype=dt_)
        keep.append(x)
        return x

    rec = mesh.rec_cell if rec_cell is None else np.asarray(rec_cell, np.int32)
    d = SolverDesc()
    d.n_elem = mesh.n_elem
    d.ngll = mesh.ngll
    d.n_rank_node = mesh.n_node
    for name in ("dxi_dx", "jacobian", "lambda_", "mu_", "mass", "density", "damping",
                 "coef_alpha", "coef_abar"):
        setattr(d, name, _ptr(kp(a[name])))
    d.ibool = _ptr(kp(a["ibool"], np.int32), C.c_int32)
    if morton:
        d.elem_centroid = _ptr(kp(mesh.centroids()))
    d.has_cpml = 1 if has_cpml else 0
    d.pml_region = _ptr(kp(a["pml_region"], np.int32), C.c_int32)
    d.dt = mesh.dt if dt is None else dt
    d.snapshot_stride = mesh.snapshot_stride if snapshot_stride is None else snapshot_stride
    rk = kp(rec, np.int32)
    d.n_rec_cell = rk.size
    d.rec_cell = _ptr(rk, C.c_int32)
    d.n_fields = n_fields
    d.assembly = assembly
    d.elems_per_block = elems_per_block
    d.element_kernel = element_kernel
    if record_capacity is None:
        record_capacity = 0 if rk.size == 0 else (mesh.nsteps // max(d.snapshot_stride, 1) + 1)
    d.record_capacity = record_capacity
    d.record_uva = 1 if record_uva else 0
    d.node_field_groups = node_field_groups
    d.general_geometry = 1 if general_geometry else 0
    if attenuation and getattr(mesh, "n_sls", 0) > 0:
        d.n_sls = mesh.n_sls
        d.tau_sigma = _ptr(kp(a["tau_sigma"]))
        d.tau_epsilon = _ptr(kp(a["tau_epsilon"]))
    if cpml_complete:
        if a.get("pml_K") is None:
            raise ValueError("cpml_complete needs the mesh's pml_K, pml_d, pml_alpha (want_profiles=1)")
        d.cpml_complete = 1
        d.pml_K = _ptr(kp(a["pml_K"]))
        d.pml_d = _ptr(kp(a["pml_d"]))
        d.pml_alpha = _ptr(kp(a["pml_alpha"]))
    return d


def cpml_rigid_nodes(mesh: Mesh):
    """Node mask [n_node] of the complete C-PML's rigid outer boundary."""
    keep = []
    d = make_desc(mesh, keep)
    mask = np.zeros(mesh.n_node, np.uint8)
    _check(lib.gfx_cpml_rigid_nodes(C.byref(d), mask.ctypes.data_as(C.POINTER(C.c_uint8))))
    return mask.astype(bool)


def autotune(mesh: Mesh, n_fields=1, trial_steps=3, **kw):
    """Fastest (assembly, element_kernel, elems_per_block) for this mesh and field count."""
    keep = []
    base = make_desc(mesh, keep, n_fields=n_fields, **kw)
    best = SolverDesc()
    ms = C.c_double(0)
    _check(lib.gfx_autotune(C.byref(base), trial_steps, C.byref(best), C.byref(ms)))
    return dict(assembly=best.assembly, element_kernel=best.element_kernel,
                elems_per_block=best.elems_per_block, node_field_groups=best.node_field_groups,
                ms_per_step=ms.value)


class Solver:
    """Device-resident solver over one mesh and n_fields right-hand sides."""

    def __init__(self, mesh: Mesh, n_fields=1, assembly=GFX_ASSEMBLY_COLORED, elems_per_block=0,
                 record_capacity=None, record_uva=False, has_cpml=True, dt=None,
                 snapshot_stride=None, rec_cell=None, element_kernel=GFX_ELEMENT_AUTO,
                 morton=True, node_field_groups=0, general_geometry=False, attenuation=True,
                 cpml_complete=False):
        self._keep = []
        d = make_desc(mesh, self._keep, n_fields=n_fields, assembly=assembly,
                      elems_per_block=elems_per_block, record_capacity=record_capacity,
                      record_uva=record_uva, has_cpml=has_cpml, dt=dt,
                      snapshot_stride=snapshot_stride, rec_cell=rec_cell,
                      element_kernel=element_kernel, morton=morton,
                      node_field_groups=node_field_groups, general_geometry=general_geometry,
                      attenuation=attenuation, cpml_complete=cpml_complete)
        rk_size = d.n_rec_cell
        h = C.c_void_p()
        _check(lib.gfx_solver_create(C.byref(d), C.byref(h)))
        self._keep = []
        self.h = h
        self.mesh = mesh
        self.n_fields = n_fields
        self.n_rec = int(rk_size)
        self.ngll = mesh.ngll

    def close(self):
        if getattr(self, "h", None):
            lib.gfx_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_source(self, field, direction, cells, weights, stf):
        cells = np.ascontiguousarray(cells, np.int32)
        weights = np.ascontiguousarray(weights, np.float64)
        stf = np.ascontiguousarray(stf, np.float64)
        _check(lib.gfx_solver_set_source(self.h, field, direction, cells.size, _ptr(cells, C.c_int32),
                                         _ptr(weights), stf.size, _ptr(stf)))

    def reset(self):
        _check(lib.gfx_solver_reset(self.h))

    def run(self, n_steps):
        _check(lib.gfx_solver_run(self.h, int(n_steps)))

    @property
    def step(self):
        return int(lib.gfx_solver_step_index(self.h))

    @property
    def n_snapshots(self):
        return int(lib.gfx_solver_n_snapshots(self.h))

    def get_state(self, field=0):
        n = self.mesh.n_node * 3
        u, v, a = (np.zeros(n) for _ in range(3))
        _check(lib.gfx_solver_get_state(self.h, field, _ptr(u), _ptr(v), _ptr(a)))
        return u, v, a

    def set_state(self, field, u, v, a):
        u, v, a = (np.ascontiguousarray(x, np.float64) for x in (u, v, a))
        _check(lib.gfx_solver_set_state(self.h, field, _ptr(u), _ptr(v), _ptr(a)))

    def get_record(self, field, snap, uva=False):
        nn = self.ngll ** 3
        s = np.zeros(self.n_rec * nn * 6)
        if uva:
            u, v, a = (np.zeros(self.n_rec * nn * 3) for _ in range(3))
            _check(lib.gfx_solver_get_record(self.h, field, snap, _ptr(s), _ptr(u), _ptr(v), _ptr(a)))
            return s, u, v, a
        _check(lib.gfx_solver_get_record(self.h, field, snap, _ptr(s), _ptr(None), _ptr(None), _ptr(None)))
        return s

    # Per-stage entry points (the reference's cuda_* step wrappers,
    # forward/share/include/gf/cuda_step.hpp:81-127).
    def stage(self, name, *args):
        """Call gfx_step_<name>(handle, *args); arrays pass as float64 numpy."""
        fn = getattr(lib, "gfx_step_" + name)
        keep = [np.ascontiguousarray(a, np.float64) if isinstance(a, np.ndarray) else None for a in args]
        conv = [_ptr(k) if k is not None else a for k, a in zip(keep, args)]
        _check(fn(self.h, *conv))

    def stage_status(self, name, *args):
        """As stage(), returning the status code instead of raising."""
        return int(getattr(lib, "gfx_step_" + name)(self.h, *args))

    def nodal_copy(self, what, field):
        """gfx_step_copy_{utilde,residual}_to_host: caller-numbered [n_node*3] float64."""
        out = np.zeros(3 * self.mesh.n_node)
        _check(getattr(lib, f"gfx_step_copy_{what}_to_host")(self.h, field, _ptr(out)))
        return out

    def set_timing(self, on=True):
        _check(lib.gfx_solver_set_timing(self.h, 1 if on else 0))

    def kernel_time(self, name):
        ms = C.c_double(0)
        n = C.c_int(0)
        _check(lib.gfx_solver_kernel_time(self.h, name.encode(), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def set_launch(self, element_kernel, elems_per_block=0, node_field_groups=None):
        _check(lib.gfx_solver_set_launch(self.h, element_kernel, elems_per_block))
        if node_field_groups is not None:
            _check(lib.gfx_solver_set_node_groups(self.h, node_field_groups))

    def query(self, key):
        v = C.c_double(0)
        _check(lib.gfx_solver_query(self.h, key.encode(), C.byref(v)))
        return v.value

    def algorithmic_bytes(self, which=0):
        return float(lib.gfx_solver_algorithmic_bytes(self.h, which))

    def set_graph(self, on=True):
        _check(lib.gfx_solver_set_graph(self.h, 1 if on else 0))

    def set_step(self, step):
        _check(lib.gfx_solver_set_step(self.h, int(step)))

    def checkpoint(self) -> bytes:
        """Complete time-stepping state as an opaque image (bit-exact resume)."""
        n = C.c_size_t(0)
        _check(lib.gfx_solver_checkpoint_size(self.h, C.byref(n)))
        buf = C.create_string_buffer(n.value)
        _check(lib.gfx_solver_checkpoint_save(self.h, buf, n.value))
        return buf.raw

    def restore(self, image: bytes):
        buf = C.create_string_buffer(image, len(image))
        _check(lib.gfx_solver_checkpoint_load(self.h, buf, len(image)))


class Batch:
    """Independent solvers (uneven meshes) advanced together, one captured
    graph replay per step.  groups=k > 0 splits the solvers into k launch
    groups, each sharing one launch per stage, where every active instance
    allows it; groups=0 (the library default) or solvers that cannot share
    keep one branch of launches per instance."""

    def __init__(self, solvers, groups=None):
        self.solvers = list(solvers)
        arr = (C.c_void_p * len(self.solvers))(*[s.h.value for s in self.solvers])
        h = C.c_void_p()
        _check(lib.gfx_batch_create(len(self.solvers), arr, C.byref(h)))
        self.h = h
        if groups is not None:
            _check(lib.gfx_batch_set_groups(self.h, int(groups)))

    @property
    def last_groups(self):
        """Launch groups of the last step run (0: one branch per instance)."""
        return int(lib.gfx_batch_groups(self.h))

    def run(self, n_steps, until=None):
        u = None if until is None else np.ascontiguousarray(until, np.int32)
        _check(lib.gfx_batch_run(self.h, int(n_steps), _ptr(u, C.c_int32)))

    def close(self):
        if getattr(self, "h", None):
            lib.gfx_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_fp = C.POINTER(C.c_float)


def gll_interpolate_host(node_xyz, cell_node_index, values, point):
    """FP64 host restatement of GLLInterpolator.interpolate for one point."""
    xyz = np.ascontiguousarray(node_xyz, np.float64).reshape(-1, 3)
    cidx = np.ascontiguousarray(cell_node_index, np.int32).reshape(-1, 125)
    vals = np.ascontiguousarray(values, np.float64).reshape(xyz.shape[0], -1)
    p = np.ascontiguousarray(point, np.float64)
    out = np.zeros(vals.shape[1])
    _check(lib.gfx_gll_interpolate_host(xyz.shape[0], _ptr(xyz), cidx.shape[0], _ptr(cidx, C.c_int32), vals.shape[1],
                                        _ptr(vals), _ptr(p), _ptr(out)))
    return out


class GreenTile:
    """Device-resident Green's tensor tile [n_time][n_node][6][3] with its cells."""

    def __init__(self, n_time, node_xyz, cell_node_index):
        self.xyz = np.ascontiguousarray(node_xyz, np.float64).reshape(-1, 3)
        self.cells = np.ascontiguousarray(cell_node_index, np.int32).reshape(-1, 125)
        self.n_time, self.n_node, self.n_cell = int(n_time), self.xyz.shape[0], self.cells.shape[0]
        h = C.c_void_p()
        _check(lib.gfx_green_create(self.n_time, self.n_node, _ptr(self.xyz), self.n_cell,
                                    _ptr(self.cells, C.c_int32), C.byref(h)))
        self.h = h

    def set_values(self, t, v):
        v = np.ascontiguousarray(v, np.float32).reshape(self.n_node, 18)
        _check(lib.gfx_green_set_values(self.h, int(t), v.ctypes.data_as(_fp)))

    def get_values(self, t):
        v = np.zeros((self.n_node, 18), np.float32)
        _check(lib.gfx_green_get_values(self.h, int(t), v.ctypes.data_as(_fp)))
        return v

    def set_capture_map(self, off, rec_point, weight):
        off = np.ascontiguousarray(off, np.int64)
        rp = np.ascontiguousarray(rec_point, np.int32)
        w = np.ascontiguousarray(weight, np.float64)
        _check(lib.gfx_green_set_capture_map(self.h, off.ctypes.data_as(C.POINTER(C.c_int64)), _ptr(rp, C.c_int32),
                                             _ptr(w)))

    def capture(self, solver, snap, field_of_dir, t):
        fd = np.ascontiguousarray(field_of_dir, np.int32)
        _check(lib.gfx_green_capture(self.h, solver.h, int(snap), _ptr(fd, C.c_int32), int(t)))

    def interpolate(self, xyz, tix):
        xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
        tix = np.ascontiguousarray(np.broadcast_to(tix, (xyz.shape[0],)), np.int32)
        out = np.zeros((xyz.shape[0], 18), np.float32)
        cell = np.zeros(xyz.shape[0], np.int32)
        _check(lib.gfx_green_interpolate(self.h, xyz.shape[0], _ptr(xyz), _ptr(tix, C.c_int32),
                                         out.ctypes.data_as(_fp), _ptr(cell, C.c_int32)))
        return out, cell

    def close(self):
        if getattr(self, "h", None):
            lib.gfx_green_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def green_capture_map(mesh, rec_cell, tile_nodes):
    """CSR map tile node -> (recording point, lumped mass) over every point of
    the recording cells that sits on the node (the mass-weighted merge)."""
    rec_cell = np.asarray(rec_cell, np.int64)
    ib = mesh.ibool.reshape(-1, 125)[rec_cell].ravel()
    mass = mesh.mass.reshape(-1, 125)[rec_cell].ravel()
    pos = np.full(mesh.n_node, -1, np.int64)
    pos[np.asarray(tile_nodes, np.int64)] = np.arange(len(tile_nodes))
    sel = pos[ib] >= 0
    tn = pos[ib][sel]
    rp = np.flatnonzero(sel).astype(np.int64)
    order = np.argsort(tn, kind="stable")
    tn, rp, w = tn[order], rp[order], mass[sel][order]
    off = np.zeros(len(tile_nodes) + 1, np.int64)
    np.add.at(off, tn + 1, 1)
    return np.cumsum(off), rp.astype(np.int32), w


class GreenSlab:
    """Exact per-tile node merge of one snapshot on the device (gfx_green_slab_*)."""

    def __init__(self, off, rec_point, weight, slab_node, f64=False, displacement=False):
        self._keep = [np.ascontiguousarray(off, np.int64), np.ascontiguousarray(rec_point, np.int32),
                      np.ascontiguousarray(weight, np.float64), np.ascontiguousarray(slab_node, np.int32)]
        o, r, w, sn = self._keep
        self.n_slab = len(sn)
        self.dtype = np.float64 if f64 else np.float32
        self.displacement = displacement
        h = C.c_void_p()
        _check(lib.gfx_green_slab_create(len(o) - 1, o.ctypes.data_as(C.POINTER(C.c_int64)), _ptr(r, C.c_int32),
                                         _ptr(w), self.n_slab, _ptr(sn, C.c_int32), 1 if f64 else 0,
                                         1 if displacement else 0, C.byref(h)))
        self.h = h

    def capture(self, solver, snap, field_of_dir=(0, 1, 2)):
        g = np.zeros((self.n_slab, 18), self.dtype)
        d = np.zeros((self.n_slab, 9), self.dtype) if self.displacement else None
        fd = np.asarray(field_of_dir, np.int32)
        _check(lib.gfx_green_slab_capture(self.h, solver.h, snap, _ptr(fd, C.c_int32),
                                          g.ctypes.data_as(C.c_void_p),
                                          None if d is None else d.ctypes.data_as(C.c_void_p)))
        return g, d

    def close(self):
        if getattr(self, "h", None):
            lib.gfx_green_slab_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class MLP:
    """Fourier-feature coordinate MLP on the device (BF16 MFMA, FP32 master weights)."""

    def __init__(self, in_lo, in_hi, out_scale, **cfg):
        c = MlpConfig()
        _check(lib.gfx_mlp_config_default(C.byref(c)))
        for k, v in cfg.items():
            setattr(c, k, v)
        self.cfg = c
        lo, hi, sc = (np.ascontiguousarray(a, np.float64) for a in (in_lo, in_hi, out_scale))
        h = C.c_void_p()
        _check(lib.gfx_mlp_create(C.byref(c), _ptr(lo), _ptr(hi), _ptr(sc), C.byref(h)))
        self.h = h
        n = C.c_size_t(0)
        _check(lib.gfx_mlp_n_params(self.h, C.byref(n)))
        self.n_params = n.value

    def params(self):
        w = np.zeros(self.n_params, np.float32)
        b = np.zeros((self.cfg.n_freq, 4), np.float32)
        _check(lib.gfx_mlp_get_params(self.h, w.ctypes.data_as(_fp), self.n_params, b.ctypes.data_as(_fp)))
        return w, b

    def set_params(self, w):
        w = np.ascontiguousarray(w, np.float32)
        _check(lib.gfx_mlp_set_params(self.h, w.ctypes.data_as(_fp), w.size))

    def set_envelope(self, cx, cy, r_max, env, slowness=0.0):
        """Source frame: output envelope e(r) over the horizontal distance r
        from (cx, cy), piecewise linear through env[i] at (i + 0.5) r_max /
        len(env) (empty: none), and time input t - slowness * r."""
        env = np.ascontiguousarray(env, np.float32).ravel()
        _check(lib.gfx_mlp_set_envelope(self.h, float(cx), float(cy), float(slowness), float(r_max), env.size,
                                        env.ctypes.data_as(_fp) if env.size else None))

    def forward(self, x):
        x = np.ascontiguousarray(x, np.float64).reshape(-1, 4)
        y = np.zeros((x.shape[0], 18), np.float32)
        _check(lib.gfx_mlp_forward(self.h, x.shape[0], _ptr(x), y.ctypes.data_as(_fp)))
        return y

    def gradient(self, x, y):
        x = np.ascontiguousarray(x, np.float64).reshape(-1, 4)
        y = np.ascontiguousarray(y, np.float32).reshape(-1, 18)
        g = np.zeros(self.n_params, np.float32)
        loss = C.c_double(0)
        _check(lib.gfx_mlp_gradient(self.h, x.shape[0], _ptr(x), y.ctypes.data_as(_fp), g.ctypes.data_as(_fp),
                                    C.byref(loss)))
        return g, loss.value

    def train(self, tile, t0, dt, iters, seed=1):
        loss = C.c_double(0)
        _check(lib.gfx_mlp_train(self.h, tile.h, t0, dt, int(iters), int(seed), C.byref(loss)))
        return loss.value

    def query(self, tile, fallback, t0, dt, xyz, tix):
        """fallback: [n_cell] or [n_cell][n_window] per region: 0 the network answers,
        1 the judge, 2 zero."""
        xyz = np.ascontiguousarray(xyz, np.float64).reshape(-1, 3)
        tix = np.ascontiguousarray(np.broadcast_to(tix, (xyz.shape[0],)), np.int32)
        fb = np.ascontiguousarray(fallback, np.uint8)
        fb = fb.reshape(tile.n_cell, -1)
        out = np.zeros((xyz.shape[0], 18), np.float32)
        nf = C.c_int(0)
        _check(lib.gfx_surrogate_query(self.h, tile.h, fb.ctypes.data_as(C.POINTER(C.c_uint8)), fb.shape[1], t0,
                                       dt, xyz.shape[0], _ptr(xyz), _ptr(tix, C.c_int32), out.ctypes.data_as(_fp),
                                       C.byref(nf)))
        return out, nf.value

    def close(self):
        if getattr(self, "h", None):
            lib.gfx_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

