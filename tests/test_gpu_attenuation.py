"""Shear attenuation on the GPU against the FP64 oracle of the same solids.

Bars as for the elastic solver (strain, u, v within 1e-4 aggregate rel L2);
the attenuated run must differ from the elastic one by far more than that,
solids with te = ts must reproduce the elastic kernel, lower Q must damp the
recorded strain more, and a checkpoint must carry the solids' memory."""
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


def att_mesh(q, n_sls=3, duration=1.5, **kw):
    return mc.small_layered(duration=duration, n_sls=n_sls, q_mu=[q, q, q], att_fmin=0.1, att_fmax=4.0, **kw)


def gpu_run(mesh, src, assembly=pygfx.GFX_ASSEMBLY_GATHER, attenuation=True, nsteps=None):
    cells, w, stf, d = src
    s = pygfx.Solver(mesh, n_fields=1, assembly=assembly, record_uva=True, attenuation=attenuation)
    s.set_source(0, d, cells, w, stf)
    s.run(mesh.nsteps if nsteps is None else nsteps)
    return s


def records(s):
    return np.stack([s.get_record(0, k) for k in range(s.n_snapshots)])


@pytest.mark.parametrize("assembly,n_sls", [(pygfx.GFX_ASSEMBLY_GATHER, 3), (pygfx.GFX_ASSEMBLY_COLORED, 3),
                                            (pygfx.GFX_ASSEMBLY_ATOMIC, 3), (pygfx.GFX_ASSEMBLY_GATHER, 1),
                                            (pygfx.GFX_ASSEMBLY_GATHER, 5)])
def test_attenuated_solver_matches_oracle(assembly, n_sls):
    """n_sls 1-4 hold the solids' coefficients in registers, 5-8 read them per field."""
    mesh = att_mesh(20.0, n_sls=n_sls)
    cells, w, stf = mc.buried_source(mesh)
    s = gpu_run(mesh, (cells, w, stf, 2), assembly=assembly)
    assert int(s.query("n_sls")) == n_sls
    ref = pyoracle.Run(mesh.arrays, mesh.n_elem, 5, mesh.n_node, mesh.dt, mesh.nsteps, mesh.snapshot_stride, 2,
                       cells, w, stf, mesh.rec_cell, tau_sigma=mesh.tau_sigma,
                       tau_epsilon=mesh.tau_epsilon).run()
    ela = pyoracle.Run(mesh.arrays, mesh.n_elem, 5, mesh.n_node, mesh.dt, mesh.nsteps, mesh.snapshot_stride, 2,
                       cells, w, stf, mesh.rec_cell).run()
    got = records(s)
    err = mc.rel_l2(got, ref["strain"])
    gap = mc.rel_l2(ela["strain"], ref["strain"])
    print(f"attenuated strain rel L2 = {err:.3e} (elastic vs attenuated {gap:.3e})")
    assert err <= TOL and gap > 100 * TOL
    u, v, a = s.get_state(0)
    assert mc.rel_l2(u, ref["u"]) <= TOL and mc.rel_l2(v, ref["v"]) <= TOL


def test_elastic_solids_reproduce_the_elastic_kernel():
    mesh = att_mesh(0.0)  # te = ts everywhere
    cells, w, stf = mc.buried_source(mesh)
    a = gpu_run(mesh, (cells, w, stf, 0))
    b = gpu_run(mesh, (cells, w, stf, 0), attenuation=False)
    assert int(a.query("n_sls")) == 3 and int(b.query("n_sls")) == 0
    # FP32 rounding of another operation order only (each run is ~1e-6 from the oracle)
    assert mc.rel_l2(records(a), records(b)) <= 1e-5


def test_lower_q_damps_more():
    peaks = []
    for q in (15.0, 60.0, 0.0):
        mesh = att_mesh(q)
        cells, w, stf = mc.surface_centre_source(mesh)
        peaks.append(np.abs(records(gpu_run(mesh, (cells, w, stf, 2)))[-40:]).max())
    assert peaks[0] < peaks[1] < peaks[2], peaks


def test_checkpoint_carries_the_memory():
    mesh = att_mesh(25.0, duration=0.8)
    cells, w, stf = mc.buried_source(mesh)
    s = gpu_run(mesh, (cells, w, stf, 1), nsteps=60)
    image = s.checkpoint()
    s.run(40)
    full = s.get_state(0)
    t = gpu_run(mesh, (cells, w, stf, 1), nsteps=0)
    t.restore(image)
    t.run(40)
    for x, y in zip(t.get_state(0), full):
        assert np.array_equal(x, y)
    e = gpu_run(mesh, (cells, w, stf, 1), nsteps=0, attenuation=False)
    with pytest.raises(pygfx.GfxError, match="attenuation"):
        e.restore(image)


@pytest.mark.skipif(not pygfx.h5_available(), reason="no HDF5 library to bind")
def test_forward_driver_reads_the_solids(tmp_path):
    mesh = mc.small_halfspace(nx=8, ny=8, nz=5, duration=1.0, n_sls=3, q_mu=[25.0], att_fmin=0.1, att_fmax=4.0)
    xyz = (0.5 * (mesh.bounds[0] + mesh.bounds[1]), 0.5 * (mesh.bounds[2] + mesh.bounds[3]), 0.0)
    stf = pygfx.ricker(1.0, 0.6, 1e20, mesh.dt, mesh.nsteps)
    d = str(tmp_path)
    pygfx.write_inputs(mesh, d, xyz, stf, precision="float64")
    pygfx.run_forward(d, directions="z", assembly=pygfx.GFX_ASSEMBLY_GATHER, restart_stride=0)
    back = pygfx.read_inputs(d)
    cells, w = back.locate_source(*xyz)
    s = gpu_run(back, (cells, w, stf, 2))
    steps = [k for k in range(mesh.nsteps) if k % mesh.snapshot_stride == 0]
    rec = np.stack([pygfx.h5_read(str(tmp_path / "wavefields" / "z" / f"record_0_{k}.h5"), "/strain")
                    for k in steps])
    mine = records(s).reshape(rec.shape)
    assert mc.rel_l2(rec, mine) <= 1e-6
    ela = gpu_run(back, (cells, w, stf, 2), attenuation=False)
    assert mc.rel_l2(records(ela).reshape(rec.shape), rec) > 1e-3
