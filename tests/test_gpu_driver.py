"""The forward driver over the reference's files, on the GPU.

A generated halfspace is written as config.h5 + partitions (the reference's
input files), run through gfx_run_forward_ex / the CLI, and the record files
are compared with the FP64 oracle run on the same arrays (strain, u, v within
1e-4 aggregate rel L2, a within 1e-3: the bars of the in-memory parity tests).
Partition count, interruption + resume and batching must not change a single
bit of the records."""
import os
import subprocess
import sys

import numpy as np
import pytest

import meshcases as mc
import pygfx
import pyoracle

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not pygfx.h5_available(), reason="no HDF5 library to bind")]

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CLI = os.path.join(ROOT, "gf-only-sem_amd", "bin", "gf_solver_elastic_hip")
TOL, TOL_ACC = 1e-4, 1e-3


@pytest.fixture(scope="module", autouse=True)
def _need_gpu():
    if pygfx.device_count() < 1:
        pytest.fail("no HIP device visible; GPU tests must run on an MI355X box")


def make_inputs(d, n_parts=1, precision="float64", restart_stride=0, duration=2.0):
    m = mc.small_halfspace(nx=8, ny=8, nz=5, duration=duration)
    stf = pygfx.ricker(1.0, 1.5, 1e20, m.dt, m.nsteps)
    xyz = (0.5 * (m.bounds[0] + m.bounds[1]), 0.5 * (m.bounds[2] + m.bounds[3]), 0.0)
    pygfx.write_inputs(m, d, xyz, stf, n_parts=n_parts, precision=precision, restart_stride=restart_stride)
    return m, stf, xyz


def read_records(d, direction, steps, sub=""):
    """Record fields per snapshot step, cells in increasing global id: a run over
    several partitions lists them in merged order, each record file naming its
    cells' global ids (recording_cell_model_index)."""
    out = {k: [] for k in ("strain", "displacement", "velocity", "acceleration")}
    for s in steps:
        p = os.path.join(d, "wavefields", sub, direction, f"record_0_{s}.h5")
        order = np.argsort(pygfx.h5_read(p, "/recording_cell_model_index"), kind="stable")
        for k in out:
            out[k].append(pygfx.h5_read(p, "/" + k).reshape(len(order), -1)[order].ravel())
    return {k: np.stack(v) for k, v in out.items()}


def snapshot_steps(m, upto=None):
    n = m.nsteps if upto is None else upto
    return [s for s in range(n) if s % m.snapshot_stride == 0]


def oracle_records(m, stf, xyz, direction):
    assert np.all(np.diff(m.rec_cell) > 0)  # the oracle records cells in increasing global id
    cells, w = m.locate_source(*xyz)
    run = pyoracle.Run(m.arrays, m.n_elem, 5, m.n_node, m.dt, m.nsteps, m.snapshot_stride, direction,
                       cells, w, stf, m.rec_cell)
    return run.run(uva=True)


def test_forward_run_matches_oracle(tmp_path):
    m, stf, xyz = make_inputs(tmp_path, n_parts=2)
    print(f"forward driver: {m.n_elem} elements, {m.nsteps} steps", flush=True)
    pygfx.run_forward(str(tmp_path), directions="z")
    steps = snapshot_steps(m)
    rec = read_records(tmp_path, "z", steps)
    ref = oracle_records(m, stf, xyz, 2)
    n = len(steps)
    es = mc.rel_l2(rec["strain"], ref["strain"][:n])
    print(f"driver strain rel L2 = {es:.3e}")
    assert es <= TOL
    assert mc.rel_l2(rec["displacement"], ref["disp"][:n]) <= TOL
    assert mc.rel_l2(rec["velocity"], ref["vel"][:n]) <= TOL
    assert mc.rel_l2(rec["acceleration"], ref["acc"][:n]) <= TOL_ACC
    assert os.path.exists(tmp_path / "log" / "forward_z_0.log")


def test_partition_count_does_not_change_records(tmp_path):
    a, b = tmp_path / "one", tmp_path / "three"
    m, _, _ = make_inputs(a, n_parts=1, duration=1.0)
    make_inputs(b, n_parts=3, duration=1.0)
    for d in (a, b):
        pygfx.run_forward(str(d), directions="y", assembly=pygfx.GFX_ASSEMBLY_COLORED,
                          element_kernel=pygfx.GFX_ELEMENT_BLOCK, elems_per_block=2)
    steps = snapshot_steps(m)
    ra, rb = read_records(a, "y", steps), read_records(b, "y", steps)
    for k in ra:
        assert np.array_equal(ra[k], rb[k]), k


@pytest.mark.parametrize("assembly", [pygfx.GFX_ASSEMBLY_COLORED, pygfx.GFX_ASSEMBLY_GATHER])
def test_interrupt_and_resume_is_bit_exact(tmp_path, assembly):
    """Kill after step k (stop_after_step), resume from the restart file, and
    compare every later record with an uninterrupted run."""
    full, cut = tmp_path / "full", tmp_path / "cut"
    m, _, _ = make_inputs(full, restart_stride=6, duration=1.5)
    make_inputs(cut, restart_stride=6, duration=1.5)
    kw = dict(directions="x", assembly=assembly, element_kernel=pygfx.GFX_ELEMENT_BLOCK, elems_per_block=2)
    pygfx.run_forward(str(full), **kw)
    k = 40  # stop after step 40; last restart at step 36
    pygfx.run_forward(str(cut), stop_after_step=k, **kw)
    rs = cut / "wavefields" / "x" / "restart_0.h5"
    assert pygfx.h5_attr(rs, "/", "step") == 36
    assert pygfx.h5_read(rs, "/displacement").size == 3 * m.n_node
    # records written after the restart step are overwritten by the resumed run
    pygfx.run_forward(str(cut), resume=True, **kw)
    steps = snapshot_steps(m)
    ra, rb = read_records(full, "x", steps), read_records(cut, "x", steps)
    for key in ra:
        assert np.array_equal(ra[key], rb[key]), key


def test_resume_from_displacement_velocity_acceleration_only(tmp_path):
    """A restart file with only the reference's datasets (no checkpoint image)
    resumes from u, v, a with the C-PML memory at zero, as the reference does
    (solver.cpp:369-388).  The resumed records equal the oracle's resume from the
    same file (gfo_run_from) within the parity gates, and differ from the
    uninterrupted run (the dropped memory)."""
    full, cut = tmp_path / "full", tmp_path / "cut"
    m, _, _ = make_inputs(full, restart_stride=10, duration=1.5)
    make_inputs(cut, restart_stride=10, duration=1.5)
    kw = dict(directions="z", assembly=pygfx.GFX_ASSEMBLY_COLORED)
    pygfx.run_forward(str(full), **kw)
    pygfx.run_forward(str(cut), stop_after_step=50, **kw)
    rs = str(cut / "wavefields" / "z" / "restart_0.h5")
    step = int(pygfx.h5_attr(rs, "/", "step"))
    keep = {k: pygfx.h5_read(rs, "/" + k) for k in ("displacement", "velocity", "acceleration", "pml_damping")}
    pygfx.h5_write(rs, "/displacement", keep["displacement"], append=False)
    for k in ("velocity", "acceleration", "pml_damping"):
        pygfx.h5_write(rs, "/" + k, keep[k])
    for k, v in (("step", step), ("ngll", 5), ("rank", 0), ("use_global_dof", 1), ("n_rank_node", m.n_node)):
        pygfx.h5_write_attr(rs, "/", k, v)
    pygfx.h5_write_attr(rs, "/", "time_s", step * m.dt)
    pygfx.run_forward(str(cut), resume=True, **kw)
    steps = [s for s in snapshot_steps(m) if s > step]
    ra, rb = read_records(full, "z", steps), read_records(cut, "z", steps)
    stf = pygfx.ricker(1.0, 1.5, 1e20, m.dt, m.nsteps)
    cells, w = m.locate_source(0.5 * (m.bounds[0] + m.bounds[1]), 0.5 * (m.bounds[2] + m.bounds[3]), 0.0)
    ref = pyoracle.Run(m.arrays, m.n_elem, 5, m.n_node, m.dt, m.nsteps, m.snapshot_stride, 2, cells, w, stf,
                       m.rec_cell).run_from(step + 1, keep["displacement"], keep["velocity"],
                                            keep["acceleration"], uva=True)
    es = mc.rel_l2(rb["strain"], ref["strain"])
    print(f"resume at step {step + 1} from u, v, a: strain rel L2 vs oracle resume {es:.3e}, "
          f"vs uninterrupted {mc.rel_l2(rb['strain'], ra['strain']):.3e}")
    assert es <= TOL
    assert mc.rel_l2(rb["displacement"], ref["disp"]) <= TOL
    assert mc.rel_l2(rb["velocity"], ref["vel"]) <= TOL
    assert mc.rel_l2(rb["acceleration"], ref["acc"]) <= TOL_ACC
    assert not np.array_equal(rb["strain"], ra["strain"])


def test_batched_stations_equal_single_runs(tmp_path):
    batch, single = tmp_path / "batch", tmp_path / "single"
    m, _, _ = make_inputs(batch, duration=1.0)
    pts = np.array([[3000.0, 4000.0, 0.0], [5000.0, 3000.0, 0.0]])
    pygfx.write_stations(batch / "stations.h5", m, pts)
    kw = dict(assembly=pygfx.GFX_ASSEMBLY_COLORED, element_kernel=pygfx.GFX_ELEMENT_BLOCK, elems_per_block=2)
    pygfx.run_forward(str(batch), directions="x,y,z", stations=str(batch / "stations.h5"), **kw)
    steps = snapshot_steps(m)
    for k in range(2):
        make_inputs(single, duration=1.0)
        pygfx.write_stations(single / "one.h5", m, pts[k:k + 1])
        for d in "xyz":
            pygfx.run_forward(str(single), directions=d, stations=str(single / "one.h5"), **kw)
            ra = read_records(batch, d, steps, sub=f"s{k}")
            rb = read_records(single, d, steps, sub="s0")
            for key in ra:
                assert np.array_equal(ra[key], rb[key]), (k, d, key)


def test_cli_runs_the_reference_command_line(tmp_path):
    m, stf, xyz = make_inputs(tmp_path, duration=0.5, precision="float32")
    r = subprocess.run([CLI, "--direction", "z", "--quiet"], cwd=tmp_path, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stderr
    steps = snapshot_steps(m)
    for s in steps:
        assert (tmp_path / "wavefields" / "z" / f"record_0_{s}.h5").exists()
    p = tmp_path / "wavefields" / "z" / f"record_0_{steps[-1]}.h5"
    assert "/strain f32 [1," in pygfx.h5_describe(p)
    bad = subprocess.run([CLI, "--direction", "w"], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert bad.returncode == 1 and "must be x, y, or z" in bad.stderr
    empty = tmp_path / "empty"
    empty.mkdir()
    miss = subprocess.run([CLI, "--direction", "x", "--quiet"], cwd=empty, capture_output=True, text=True,
                          timeout=60)
    assert miss.returncode == 1 and "config.h5" in miss.stderr


def test_forward_run_with_complete_cpml(tmp_path):
    """--cpml-complete: the driver reads /field/cell/pml_K, pml_d, pml_alpha
    from the partitions and runs the complete C-PML; records against the
    oracle's complete C-PML (rigid outer boundary as zero mass), and the CLI
    flag gives the same files bit for bit."""
    m = mc.small_halfspace(nx=8, ny=8, nz=5, duration=2.0, want_profiles=1, pml_profile=1)
    stf = pygfx.ricker(1.0, 1.5, 1e20, m.dt, m.nsteps)
    xyz = (0.5 * (m.bounds[0] + m.bounds[1]), 0.5 * (m.bounds[2] + m.bounds[3]), 0.0)
    pygfx.write_inputs(m, tmp_path, xyz, stf, n_parts=2)
    print(f"forward driver, complete C-PML: {m.n_elem} elements, {m.nsteps} steps", flush=True)
    pygfx.run_forward(str(tmp_path), directions="z", cpml_complete=True)
    steps = snapshot_steps(m)
    rec = read_records(tmp_path, "z", steps)
    cells, w = m.locate_source(*xyz)
    ref = pyoracle.Run(mc.cpml_complete_arrays(m), m.n_elem, 5, m.n_node, m.dt, m.nsteps, m.snapshot_stride, 2,
                       cells, w, stf, m.rec_cell, cpml_complete=True).run(uva=True)
    old = oracle_records(m, stf, xyz, 2)
    n = len(steps)
    for k, r in (("strain", "strain"), ("displacement", "disp"), ("velocity", "vel")):
        e = mc.rel_l2(rec[k], ref[r][:n])
        print(f"driver {k} rel L2 vs the complete C-PML oracle = {e:.3e}, vs the reference C-PML's "
              f"{mc.rel_l2(rec[k], old[r][:n]):.3e}")
        assert e <= TOL
    cli_dir = tmp_path / "cli"
    pygfx.write_inputs(m, cli_dir, xyz, stf, n_parts=1)
    r = subprocess.run([CLI, "--direction", "z", "--quiet", "--cpml-complete"], cwd=cli_dir, capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    rec_cli = read_records(cli_dir, "z", steps)
    for k in rec:
        assert np.array_equal(rec_cli[k], rec[k]), k
    # inputs without the profiles are refused
    bare = tmp_path / "bare"
    pygfx.write_inputs(mc.small_halfspace(nx=8, ny=8, nz=5, duration=0.2), bare, xyz, stf[:16], n_parts=1)
    with pytest.raises(pygfx.GfxError, match="pml_K"):
        pygfx.run_forward(str(bare), directions="z", cpml_complete=True)
