"""GPU parity of the C-PML step (forward/share/src/pml.cpp:55-216, called at
forward/share/src/solver.cpp:545-548,558-562,619-621) on inputs where it dominates
the PML (meshcases.amplified_cpml; tests/test_cpml_sensitivity.py shows that there
each of its parts moves the records by 1e-3 to 1 on the oracle).

Gates: recorded strain of the surface-layer PML cells and u, v of every PML node
within 1e-4 rel L2 of the FP64 oracle, for the nodal form (memory kept per node)
and the element-local form (pml.cpp's own layout), under GATHER, COLORED and
ATOMIC assembly.  Negative controls: the same GPU run with the C-PML term removed
(A = 0) or with the legacy ramp instead misses the gate by more than 100x."""
import numpy as np
import pytest

import meshcases as mc
import pygfx
import pyoracle

pytestmark = pytest.mark.gpu

TOL = 1e-4


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if pygfx.device_count() < 1:
        pytest.fail("no HIP device visible; GPU tests must run on an MI355X box")


def _element_local(mesh):
    """Perturb the x-direction alpha of one element at one node that another element
    shares, so that the nodal form no longer applies (the solver then keeps the
    memory per element point, as pml.cpp does)."""
    al = mesh.coef_alpha.reshape(-1, 125, 9).copy()
    ab = mesh.coef_abar.reshape(-1, 125, 5)
    use = (mesh.pml_region[:, None] != 0) & (ab[:, :, 2] != 0)
    nodes = mesh.ibool.reshape(-1, 125)
    ids, counts = np.unique(nodes[use], return_counts=True)
    shared = ids[counts >= 2][0]
    e, p = (int(x[0]) for x in np.nonzero(use & (nodes == shared)))
    al[e, p, 0] *= 0.999
    return mesh.derive(coef_alpha=al.reshape(-1))


_CASES = {}


def _case(form):
    if form not in _CASES:
        mesh = mc.amplified_cpml(mc.small_halfspace(duration=4.0))
        if form == "element":
            mesh = _element_local(mesh)
        cells, w, stf = mc.surface_centre_source(mesh)
        ref = pyoracle.Run(mesh.arrays, mesh.n_elem, 5, mesh.n_node, mesh.dt, mesh.nsteps, mesh.snapshot_stride,
                           2, cells, w, stf, mesh.rec_cell).run(uva=True)
        _CASES[form] = (mesh, (cells, w, stf), ref)
    return _CASES[form]


def _gpu(mesh, src, assembly, has_cpml=True):
    cells, w, stf = src
    s = pygfx.Solver(mesh, n_fields=1, assembly=assembly, record_uva=True, has_cpml=has_cpml)
    s.set_source(0, 2, cells, w, stf)
    s.run(mesh.nsteps)
    n_snap = s.n_snapshots
    st = np.stack([s.get_record(0, k) for k in range(n_snap)]).reshape(n_snap, s.n_rec, 125, 6)
    return s, st


@pytest.mark.parametrize("form", ["nodal", "element"])
@pytest.mark.parametrize("assembly", [pygfx.GFX_ASSEMBLY_GATHER, pygfx.GFX_ASSEMBLY_COLORED,
                                      pygfx.GFX_ASSEMBLY_ATOMIC])
def test_cpml_step_matches_oracle_where_it_dominates(form, assembly):
    mesh, src, ref = _case(form)
    print(f"case: amplified C-PML, form={form}, assembly={assembly}, {mesh.n_elem} elements, "
          f"{mesh.nsteps} steps", flush=True)
    s, st = _gpu(mesh, src, assembly)
    assert s.query("pml_nodal") == (1.0 if form == "nodal" else 0.0)
    assert st.shape == ref["strain"].shape
    pn = mc.pml_nodes(mesh)
    u, v, a = s.get_state(0)
    errs = (mc.rel_l2(st, ref["strain"]), mc.rel_l2(u.reshape(-1, 3)[pn], ref["u"].reshape(-1, 3)[pn]),
            mc.rel_l2(v.reshape(-1, 3)[pn], ref["v"].reshape(-1, 3)[pn]))
    print(f"strain {errs[0]:.3e}, PML-node u {errs[1]:.3e}, v {errs[2]:.3e}")
    assert max(errs) <= TOL


def test_cpml_negative_controls_miss_the_gate():
    """Without its C-PML term (A = 0), or with the legacy ramp in its place, the GPU
    run misses the oracle's C-PML run by more than 100x the gate: the parity test
    above cannot pass with the C-PML step missing."""
    mesh, src, ref = _case("nodal")
    zero = mesh.derive(coef_abar=np.zeros_like(mesh.coef_abar))
    _, st0 = _gpu(zero, src, pygfx.GFX_ASSEMBLY_GATHER)
    _, st1 = _gpu(mesh, src, pygfx.GFX_ASSEMBLY_GATHER, has_cpml=False)
    e0, e1 = mc.rel_l2(st0, ref["strain"]), mc.rel_l2(st1, ref["strain"])
    print(f"strain vs the C-PML oracle: A = 0 {e0:.3e}, legacy ramp {e1:.3e}")
    assert e0 > 100 * TOL and e1 > 100 * TOL
