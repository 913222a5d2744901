"""The solver on a general hexahedral mesh (GMSH -> /topology -> inputs).

A jittered lattice written as a GMSH file goes through gfx_gmsh_to_topology
and gfx_mesh_build_topology; the GPU solver on it matches the FP64 oracle on
the same arrays (strain within 1e-4 aggregate rel L2), and a generated box
given as general cells gives the box's own records to 1e-6."""
import numpy as np
import pytest

import meshcases as mc
import pygfx
import pyoracle
from test_topology import box_cells, write_msh41

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not pygfx.h5_available(), reason="no HDF5 library to bind")]

CORNER_PTS = np.array([(i * 4 * 25 + j * 4 * 5 + k * 4) for (i, j, k) in
                       [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]])


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if pygfx.device_count() < 1:
        pytest.fail("no HIP device visible; GPU tests must run on an MI355X box")


def _run(mesh, cells, w, stf, nf=3):
    s = pygfx.Solver(mesh, n_fields=nf, assembly=pygfx.GFX_ASSEMBLY_GATHER, record_uva=True)
    for f in range(nf):
        s.set_source(f, f, cells, w, stf)
    s.run(mesh.nsteps)
    return s


def test_gmsh_mesh_matches_oracle(tmp_path):
    xyz, hexes = box_cells(8, 7, 5, h=(900.0, 900.0, 900.0), jitter=0.12, seed=6)
    msh, model = str(tmp_path / "m.msh"), str(tmp_path / "model.h5")
    write_msh41(msh, xyz, hexes, split=2)
    pygfx.gmsh_to_topology(msh, model)
    mesh = pygfx.build_topology(model, nx=8, ny=7, order=4, layers=mc.LAYERED, rf_amp=0.05, rf_corr_len=2000.0,
                                rf_modes=64, rf_seed=11, pml_xmin=2, pml_xmax=2, pml_ymin=2, pml_ymax=2,
                                pml_zmax=2, f0_pml=1.5, output_dt=0.05, duration=1.5, record_depth_max=1800.0)
    assert mesh.n_pml_elem > 0 and len(mesh.rec_cell) > 0
    cells, w, stf = mc.buried_source(mesh)
    s = _run(mesh, cells, w, stf, nf=1)
    ref = pyoracle.Run(mesh.arrays, mesh.n_elem, 5, mesh.n_node, mesh.dt, mesh.nsteps, mesh.snapshot_stride, 0,
                       cells, w, stf, mesh.rec_cell).run()
    n = ref["strain"].shape[0]
    got = np.stack([s.get_record(0, k) for k in range(n)])
    err = mc.rel_l2(got, ref["strain"])
    print(f"general mesh strain rel L2 = {err:.3e}")
    assert err <= 1e-4


def test_cells_off_the_lattice_run_with_face_pml_flags():
    """A 9 x 7 x 5 jittered lattice given with nx_elements 8 (315 cells, not a
    multiple of 8 x 7): the C-PML cells come from the boundary faces
    (preprocess/boundary_detector.py), the widths from extent / nx; the GPU
    solver on these inputs matches the oracle."""
    xyz, hexes = box_cells(9, 7, 5, h=(900.0, 900.0, 900.0), jitter=0.12, seed=7)
    mesh = pygfx.build_hex(xyz, hexes, nx=8, ny=7, order=4, layers=mc.LAYERED, rf_amp=0.05, rf_corr_len=2000.0,
                           rf_modes=64, rf_seed=11, pml_xmin=1, pml_xmax=1, pml_ymin=1, pml_ymax=1, pml_zmax=1,
                           f0_pml=1.5, output_dt=0.05, duration=1.5, record_depth_max=1800.0)
    k, j, i = np.meshgrid(np.arange(5), np.arange(7), np.arange(9), indexing="ij")
    boundary = ((i == 0) | (i == 8) | (j == 0) | (j == 6) | (k == 4)).ravel()
    assert np.array_equal(mesh.pml_region != 0, boundary)
    cells, w, stf = mc.buried_source(mesh)
    s = _run(mesh, cells, w, stf, nf=1)
    ref = pyoracle.Run(mesh.arrays, mesh.n_elem, 5, mesh.n_node, mesh.dt, mesh.nsteps, mesh.snapshot_stride, 0,
                       cells, w, stf, mesh.rec_cell).run()
    n = ref["strain"].shape[0]
    got = np.stack([s.get_record(0, kk) for kk in range(n)])
    err = mc.rel_l2(got, ref["strain"])
    print(f"off-lattice general mesh strain rel L2 = {err:.3e}")
    assert err <= 1e-4


def test_box_as_general_cells_runs_as_the_box():
    spec = dict(nx=8, ny=8, nz=5, lx=8000.0, ly=8000.0, lz=5000.0, layers=mc.HALFSPACE, jitter=0.1, jitter_seed=4,
                pml_xmin=2, pml_xmax=2, pml_ymin=2, pml_ymax=2, pml_zmax=2, f0_pml=1.0, duration=1.0,
                record_depth_max=2000.0)
    box = pygfx.build_box(**spec)
    ib = box.ibool.reshape(box.n_elem, 125)
    vids, hexes = np.unique(ib[:, CORNER_PTS], return_inverse=True)
    gen = pygfx.build_hex(box.node_xyz.reshape(-1, 3)[vids], hexes.reshape(box.n_elem, 8), **spec)
    cells, w, stf = mc.surface_centre_source(box)
    a = _run(box, cells, w, stf)
    cells_g, w_g, _ = mc.surface_centre_source(gen)
    assert np.array_equal(cells_g, cells)
    b = _run(gen, cells_g, w_g, stf)
    for f in range(3):
        ra = np.stack([a.get_record(f, k) for k in range(a.n_snapshots)])
        rb = np.stack([b.get_record(f, k) for k in range(b.n_snapshots)])
        assert mc.rel_l2(rb, ra) <= 1e-6, f
