"""Graph-captured steps and batches of uneven meshes, on the GPU.

A captured step is replayed many times with the program's own eager work
(state reads, record reads, synchronisations) between the replays; every
replay must equal the kernel-by-kernel run bit for bit.  A batch of meshes of
different size, material, C-PML thickness, time step and run length, advanced
as one graph (one branch per mesh, or launches shared by all meshes), must
reproduce each mesh's single run bit for bit, and one of them is checked
against the FP64 oracle."""
import numpy as np
import pytest

import meshcases as mc
import pygfx
import pyoracle

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if pygfx.device_count() < 1:
        pytest.fail("no HIP device visible; GPU tests must run on an MI355X box")


def _solver(mesh, sources, assembly, graph=False, **kw):
    s = pygfx.Solver(mesh, n_fields=len(sources), assembly=assembly, record_uva=True, **kw)
    for f, (cells, w, stf, d) in enumerate(sources):
        s.set_source(f, d, cells, w, stf)
    if graph:
        s.set_graph(True)
    return s


def _all_records(s, n_snap):
    return [s.get_record(f, k, uva=True) for f in range(s.n_fields) for k in range(n_snap)]


@pytest.mark.parametrize("assembly", [pygfx.GFX_ASSEMBLY_COLORED, pygfx.GFX_ASSEMBLY_GATHER])
def test_graph_replays_equal_eager_steps(assembly):
    mesh = mc.small_layered(duration=1.0)
    cells, w, stf = mc.buried_source(mesh)
    c2, w2, stf2 = mc.buried_source(mesh, frac=(0.4, 0.6, 0.0), depth=900.0)
    src = [(cells, w, stf, 2), (c2, w2, stf2, 0)]
    eager = _solver(mesh, src, assembly)
    graph = _solver(mesh, src, assembly, graph=True)
    done = 0
    for chunk in (1, 5, 3, 17, mesh.nsteps):
        chunk = min(chunk, mesh.nsteps - done)
        if chunk <= 0:
            break
        eager.run(chunk)
        graph.run(chunk)
        done += chunk
        assert graph.step == eager.step == done
        for f in range(2):  # eager reads between replays
            for x, y in zip(graph.get_state(f), eager.get_state(f)):
                assert np.array_equal(x, y)
    assert graph.n_snapshots == eager.n_snapshots
    for a, b in zip(_all_records(graph, graph.n_snapshots), _all_records(eager, eager.n_snapshots)):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


def test_graph_is_recaptured_after_launch_and_source_changes():
    mesh = mc.small_halfspace(duration=0.6)
    cells, w, stf = mc.surface_centre_source(mesh)
    eager = _solver(mesh, [(cells, w, stf, 2)], pygfx.GFX_ASSEMBLY_GATHER)
    graph = _solver(mesh, [(cells, w, stf, 2)], pygfx.GFX_ASSEMBLY_GATHER, graph=True)
    for s in (eager, graph):
        s.run(10)
        s.set_launch(pygfx.GFX_ELEMENT_BLOCK, 4, node_field_groups=2)
        s.run(10)
        s.set_source(0, 1, cells, w, 0.5 * stf)  # new source arrays on the device
        s.run(mesh.nsteps - 20)
    for x, y in zip(graph.get_state(0), eager.get_state(0)):
        assert np.array_equal(x, y)


def _uneven_meshes():
    """Three boxes of different size, layering, C-PML thickness, dt and length."""
    m0 = mc.small_halfspace(nx=6, ny=6, nz=4, pml=1, duration=0.8)
    m1 = mc.small_layered(nx=9, ny=7, nz=5, pml=2, duration=1.2)
    m2 = pygfx.build_box(nx=7, ny=5, nz=5, lx=4200.0, ly=3000.0, lz=3000.0,
                         layers=[(1200.0, 3000.0, 1700.0, 2300.0), (0.0, 5200.0, 3000.0, 2900.0)],
                         pml_xmin=3, pml_xmax=1, pml_ymin=1, pml_ymax=1, pml_zmax=2, f0_pml=1.5,
                         duration=0.5, record_depth_max=1200.0)
    return [m0, m1, m2]


@pytest.mark.parametrize("assembly,groups", [(pygfx.GFX_ASSEMBLY_COLORED, 2), (pygfx.GFX_ASSEMBLY_GATHER, 0),
                                             (pygfx.GFX_ASSEMBLY_GATHER, 1), (pygfx.GFX_ASSEMBLY_GATHER, 2)])
def test_batch_of_uneven_meshes_equals_single_runs(assembly, groups):
    """GATHER batches share one launch per stage among the meshes of a launch
    group; COLORED ones, and GATHER with groups=0, keep one branch per mesh."""
    meshes = _uneven_meshes()
    assert len({m.dt for m in meshes}) > 1 and len({m.nsteps for m in meshes}) == 3
    srcs = []
    for m in meshes:
        cells, w, stf = mc.surface_centre_source(m, f0=1.5, t0=0.3)
        srcs.append([(cells, w, stf, 2), (cells, w, stf, 0)])
    batched = [_solver(m, s, assembly) for m, s in zip(meshes, srcs)]
    single = [_solver(m, s, assembly) for m, s in zip(meshes, srcs)]
    batch = pygfx.Batch(batched, groups=groups)
    until = [m.nsteps for m in meshes]
    # in pieces, with eager reads in between
    batch.run(7, until)
    assert batch.last_groups == (groups if assembly == pygfx.GFX_ASSEMBLY_GATHER else 0)
    for s in batched:
        s.get_state(0)
    batch.run(10 ** 6, until)
    for s, m in zip(batched, meshes):
        assert s.step == m.nsteps
    for s, m in zip(single, meshes):
        s.run(m.nsteps)
    for b, s in zip(batched, single):
        assert b.n_snapshots == s.n_snapshots
        for f in range(2):
            for x, y in zip(b.get_state(f), s.get_state(f)):
                assert np.array_equal(x, y)
        for x, y in zip(_all_records(b, b.n_snapshots), _all_records(s, s.n_snapshots)):
            for p, q in zip(x, y):
                assert np.array_equal(p, q)
    # one instance against the FP64 oracle
    m = meshes[1]
    cells, w, stf, d = srcs[1][0]
    ref = pyoracle.Run(m.arrays, m.n_elem, 5, m.n_node, m.dt, m.nsteps, m.snapshot_stride, d, cells, w, stf,
                       m.rec_cell).run()["strain"]
    got = np.stack([batched[1].get_record(0, k) for k in range(batched[1].n_snapshots)])
    err = mc.rel_l2(got, ref)
    print(f"batched instance strain rel L2 vs oracle = {err:.3e}")
    assert err <= 1e-4
    batch.close()


def test_launch_groups_fall_back_to_branches_when_a_mesh_cannot_share():
    """An attenuating mesh runs its own element kernel form, so a batch holding
    one keeps one branch per mesh even with launch groups requested; dropping it
    from the active set (its run is over) lets the rest share launches."""
    m0 = mc.small_halfspace(nx=6, ny=6, nz=4, pml=1, duration=0.3)
    m1 = mc.small_layered(duration=0.8, n_sls=2, q_mu=[30.0, 30.0, 30.0], att_fmin=0.1, att_fmax=4.0)
    m2 = mc.small_halfspace(nx=7, ny=6, nz=4, pml=1, duration=0.6)
    meshes = [m0, m1, m2]
    srcs = [[mc.surface_centre_source(m) + (2,)] for m in meshes]
    batched = [_solver(m, s, pygfx.GFX_ASSEMBLY_GATHER) for m, s in zip(meshes, srcs)]
    single = [_solver(m, s, pygfx.GFX_ASSEMBLY_GATHER) for m, s in zip(meshes, srcs)]
    assert int(batched[1].query("n_sls")) == 2
    until = [40, 20, 40]
    batch = pygfx.Batch(batched, groups=1)
    batch.run(10, until)
    assert batch.last_groups == 0  # the attenuating mesh is active
    batch.run(10 ** 6, until)
    assert batch.last_groups == 1  # meshes 0 and 2 alone
    for b, s, n in zip(batched, single, until):
        s.run(n)
        for x, y in zip(b.get_state(0), s.get_state(0)):
            assert np.array_equal(x, y)
    batch.close()
