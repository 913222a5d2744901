"""The amplified C-PML case (meshcases.amplified_cpml) is a sensitive parity case: on
it the oracle's C-PML step (forward/share/src/pml.cpp:55-216) moves the records far
beyond the 1e-4 parity gate, part by part, and the run stays stable.  The GPU parity
test on the same case (tests/test_gpu_cpml.py) therefore checks every part of the
C-PML step; on the generated coefficients the whole term is worth ~1e-8."""
import numpy as np
import pytest

import meshcases as mc
import pyoracle


def _run(mesh, **over):
    arrays = dict(mesh.arrays)
    arrays.update(over)
    cells, w, stf = mc.surface_centre_source(mesh)
    return pyoracle.Run(arrays, mesh.n_elem, 5, mesh.n_node, mesh.dt, mesh.nsteps, mesh.snapshot_stride, 2,
                        cells, w, stf, mesh.rec_cell).run(uva=True)


@pytest.fixture(scope="module")
def case():
    mesh = mc.amplified_cpml(mc.small_halfspace(duration=4.0))
    return mesh, _run(mesh)


def _diff(mesh, full, other):
    pn = mc.pml_nodes(mesh)
    us = lambda x: x.reshape(-1, 3)[pn]  # noqa: E731
    return mc.rel_l2(other["strain"], full["strain"]), mc.rel_l2(us(other["u"]), us(full["u"]))


def test_amplified_case_records_pml_cells_and_is_stable(case):
    mesh, full = case
    assert mesh.rec_cell.size > 0 and np.all(mesh.pml_region[mesh.rec_cell] != 0)
    peak = np.abs(full["disp"]).max(axis=(1, 2, 3))
    assert np.all(np.isfinite(peak))
    assert peak[-1] < 0.6 * peak.max()  # the wave field decays in the PML


def test_every_part_of_the_cpml_step_moves_the_records(case):
    mesh, full = case
    ab = mesh.coef_abar.reshape(-1, 5)
    no_cpml = _diff(mesh, full, _run(mesh, coef_abar=np.zeros_like(mesh.coef_abar)))
    mem0 = ab.copy()
    mem0[:, 2:] = 0.0
    no_memory = _diff(mesh, full, _run(mesh, coef_abar=mem0.reshape(-1)))
    al = mesh.coef_alpha.reshape(-1, 3, 3)
    dirs = _diff(mesh, full, _run(mesh, coef_alpha=np.ascontiguousarray(al[:, ::-1, :]).reshape(-1)))
    c12 = _diff(mesh, full, _run(mesh, coef_alpha=np.ascontiguousarray(al[:, :, [0, 2, 1]]).reshape(-1)))
    print(f"strain / PML-node u changes: no C-PML {no_cpml}, no memory {no_memory}, x<->z poles {dirs}, "
          f"new/old weights swapped {c12}")
    assert min(no_cpml) > 1e-1
    assert min(no_memory) > 1e-2
    assert min(dirs) > 1e-2
    assert min(c12) > 5e-4  # 5x the parity gate: the old/new weights are told apart
