#!/usr/bin/env python3
"""Throughput benchmark of the MI355X elastic CG-SEM time step.

Workload: the 500k-element layered mesh (100 x 100 x 50 hexes of 0.6 km,
polynomial order 4, 3 layers + 5 % Gaussian random velocity field, C-PML 5
elements thick on the sides and bottom, free surface on top, Ricker f0 = 1.5 Hz,
strain recorded at the cells down to 2.4 km every snapshot stride).
  * plain:   one z-direction point force at the surface centre (1 field);
  * batched: x, y and z point forces at 8 seeded surface stations in the
             central 40 x 40 km (24 fields) advanced by one launch per colour.
Metric: GLL-point updates/s = elements * 125 * fields * steps / time-loop wall
time (inputs resident in HBM, device recording included).  `value` is the
batched aggregate; the plain figure and the batched/plain ratio sit beside it.

A step is one pass of the whole hot path (element kernels for all colours,
C-PML memory, sources, Newmark node kernel, and the recording kernel on
snapshot steps).  The timed region is bracketed by device synchronisations.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "gf-only-sem_amd"))

import numpy as np  # noqa: E402

METRIC = "GLL-point updates/s, batched 3×N sources vs plain, at ≤1e-4 rel L2 strain"
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cfg2_spec(pygfx, nx=100, ny=100, nz=50, h=600.0, duration=30.0, n_sls=0, profiles=False):
    from pygfx import workloads
    spec = workloads.layered_spec(nx, ny, nz, h=h, duration=duration, n_sls=n_sls)
    spec.want_profiles = 1 if profiles else 0  # pml_K, pml_d, pml_alpha for the complete C-PML leg
    return spec


def stations(mesh, n=8, seed=3):
    """Seeded surface stations in the central 40 x 40 km (80 % of the extent on a
    smaller mesh)."""
    from pygfx import workloads
    ext = min(mesh.bounds[1] - mesh.bounds[0], mesh.bounds[3] - mesh.bounds[2])
    return workloads.stations(mesh, n=n, seed=seed, half=min(20000.0, 0.4 * ext))


def time_run(pygfx, solver, steps, warmup):
    solver.run(warmup)
    solver.set_timing(True)
    t0 = time.perf_counter()
    solver.run(steps)  # returns after hipStreamSynchronize on the solver stream
    t1 = time.perf_counter()
    solver.set_timing(False)
    out = dict(wall_s=t1 - t0)
    for name in ("element", "element_pml", "node", "node_special", "node_regular", "record", "total"):
        ms, n = solver.kernel_time(name)
        out[name] = (ms, n)
    return out


def copy_calibration(n_copies=3, nbytes=1 << 30):
    """Device-to-device copies of a known size through the HIP runtime, run
    first under a rocprofv3 --pmc pass: what FETCH_SIZE / WRITE_SIZE read for
    them calibrates the counters against bytes actually moved."""
    import ctypes as C
    hip = C.CDLL("libamdhip64.so.7")
    a, b = C.c_void_p(), C.c_void_p()
    if hip.hipMalloc(C.byref(a), C.c_size_t(nbytes)) or hip.hipMalloc(C.byref(b), C.c_size_t(nbytes)):
        raise RuntimeError("hipMalloc failed")
    try:
        if hip.hipMemset(a, 1, C.c_size_t(nbytes)):
            raise RuntimeError("hipMemset failed")
        for _ in range(n_copies):
            if hip.hipMemcpy(b, a, C.c_size_t(nbytes), 3):  # hipMemcpyDeviceToDevice
                raise RuntimeError("hipMemcpy failed")
        hip.hipDeviceSynchronize()
    finally:
        hip.hipFree(a)
        hip.hipFree(b)
    log(f"copy calibration: {n_copies} x {nbytes} B device-to-device")


def pmc_summary(fetch_csv, write_csv, n_copies=3, nbytes=1 << 30):
    """Per-kernel HBM traffic from two rocprofv3 --pmc passes (FETCH_SIZE, then
    WRITE_SIZE; counter values in KiB per dispatch), each scaled by the factor
    that makes the calibration copies read as the bytes they moved."""
    import csv

    def load(path, counter):
        per = {}
        for r in csv.DictReader(open(path)):
            if r.get("Counter_Name", counter) != counter:
                continue
            name = r["Kernel_Name"].split("(")[0].replace("void ", "").strip()
            per.setdefault(name, []).append(float(r["Counter_Value"]) * 1024.0)
        return per

    fetch = load(fetch_csv, "FETCH_SIZE")
    write = load(write_csv, "WRITE_SIZE")

    def factor(per, expect):
        copies = [v for k, v in per.items() if "copybuffer" in k.lower() or "copy_buffer" in k.lower()]
        tot = sum(sum(v) for v in copies)
        return (expect * n_copies / tot) if tot > 0 else None

    ff = factor(fetch, nbytes)
    fw = factor(write, nbytes)
    lines = [f"calibration: {n_copies} device-to-device copies of {nbytes} B; "
             f"FETCH_SIZE factor {ff if ff is None else round(ff, 4)}, "
             f"WRITE_SIZE factor {fw if fw is None else round(fw, 4)}",
             f"{'kernel':72s} {'dispatches':>10s} {'fetch GB':>10s} {'write GB':>10s} {'total GB':>10s}"]
    for k in sorted(set(fetch) | set(write)):
        f = fetch.get(k, [])
        w = write.get(k, [])
        fm = (sum(f) / len(f) * (ff or 1.0)) if f else 0.0
        wm = (sum(w) / len(w) * (fw or 1.0)) if w else 0.0
        lines.append(f"{k[:72]:72s} {max(len(f), len(w)):10d} {fm / 1e9:10.4f} {wm / 1e9:10.4f} "
                     f"{(fm + wm) / 1e9:10.4f}")
    return "\n".join(lines)


def traffic_from_profiles(kernel_prefix, launch, fields, dims):
    """HBM bytes per dispatch of the element kernel as committed from the PMC
    passes (profiles/*.txt lines 'traffic <kernel> <launch> <fields> <mesh>
    <bytes>'), with the launch the pass ran.  The pass of the same launch is
    used when there is one; otherwise a pass of the same kernel family, field
    count and workload at another launch (its traffic is set by the geometry,
    displacement and residual arrays, not by the workgroup shape), named in the
    result.  (None, None) when no pass covers the kernel family and workload."""
    tag = f"{launch['assembly']}/{launch['element_kernel']}/{launch['elems_per_block']}"
    mesh = "x".join(str(d) for d in dims)
    pdir = os.path.join(ROOT, "profiles")
    if not os.path.isdir(pdir):
        return None, None
    same, other = None, None
    for fn in sorted(os.listdir(pdir)):
        if not fn.endswith(".txt"):
            continue
        for line in open(os.path.join(pdir, fn)):
            p = line.split()
            if len(p) != 6 or p[0] != "traffic" or p[3] != str(fields) or p[4] != mesh:
                continue
            if p[1] == kernel_prefix and p[2] == tag:
                same = (float(p[5]), p[2])
            elif p[1].startswith("k_element") and kernel_prefix.startswith("k_element") and \
                    "k_element_cpml" not in (p[1], kernel_prefix) and \
                    p[2].split("/")[0] == launch["assembly"] and other is None:
                other = (float(p[5]), f"{p[1]} {p[2]}")
    return same or other or (None, None)


# LDS reads per element point and field of the block element kernel's field
# loop (k_element<EPB, GATHER, no element-local C-PML, diagonal or general form),
# counted in its gfx950 code: 39 ds_read2_b32 + 12 ds_read_b32 = 90 dwords
# (45 displacement values for the gradient, 45 flux values for the divergence).
# The measured rate of 32-bit and read2 LDS reads on this part is about 104
# bytes per clock per CU: 256 CUs x 2.4 GHz x 104 B = 63.9 TB/s.
BLOCK_LDS_DWORDS = 90
LDS_MEASURED_TBS = 63.9


def element_lds(launch, n_elem, n_fields, ms_elem):
    """LDS read rate of the block element kernel's field loop (its binding limit)."""
    if launch["element_kernel"] != "block" or ms_elem <= 0:
        return None
    tbs = n_elem * 125 * n_fields * BLOCK_LDS_DWORDS * 4 / (ms_elem * 1e-3) / 1e12
    return {"dwords_per_point_field": BLOCK_LDS_DWORDS, "achieved_tbs": tbs, "measured_rate_tbs": LDS_MEASURED_TBS,
            "frac_of_measured_rate": tbs / LDS_MEASURED_TBS}


def uneven_specs(pygfx, n=16, seed=4, lo=8e3, hi=2e5, pml=(3, 6)):
    from pygfx import workloads
    return workloads.uneven_specs(n=n, seed=seed, lo=lo, hi=hi, pml=pml)


def uneven_workload(pygfx, steps, warmup, group_counts=(0, 1, 2, 4), n=16, lo=8e3, hi=2e5, pml=(3, 6)):
    """x, y and z forces at the surface centre of each of n uneven meshes:
    the meshes advanced together as one captured graph per step, for each
    number of launch groups in group_counts (0: one branch of launches per
    mesh; k: k branches, each sharing one launch per stage among its meshes),
    against the same solvers advanced one after another, kernel by kernel.
    value is the fastest batched form."""
    t0 = time.perf_counter()
    meshes = [pygfx.build_box(sp, copy=False) for sp in uneven_specs(pygfx, n=n, lo=lo, hi=hi, pml=pml)]
    solvers = []
    for m in meshes:
        s = pygfx.Solver(m, n_fields=3, assembly=pygfx.GFX_ASSEMBLY_GATHER, record_capacity=2)
        cx = 0.5 * (m.bounds[0] + m.bounds[1])
        cy = 0.5 * (m.bounds[2] + m.bounds[3])
        cells, w = m.locate_source(cx, cy, 0.0)
        stf = pygfx.ricker(1.5, 1.0, 1e20, m.dt, m.nsteps)
        for f in range(3):
            s.set_source(f, f, cells, w, stf)
        solvers.append(s)
    E = sum(m.n_elem for m in meshes)
    log(f"uneven: {len(meshes)} meshes, {E} elements ({min(m.n_elem for m in meshes)}..."
        f"{max(m.n_elem for m in meshes)}), dt {min(m.dt for m in meshes):.4g}..{max(m.dt for m in meshes):.4g} s "
        f"({time.perf_counter() - t0:.1f} s setup)")
    for s in solvers:
        s.run(warmup)
    t0 = time.perf_counter()
    for s in solvers:
        s.run(steps)
    t_seq = time.perf_counter() - t0
    times = {}
    for groups in group_counts:
        batch = pygfx.Batch(solvers, groups=groups)
        batch.run(warmup)
        t0 = time.perf_counter()
        batch.run(steps)
        times[groups] = time.perf_counter() - t0
        assert batch.last_groups == groups
        batch.close()
    best = min(times, key=times.get)
    t_bat = times[best]
    pts = E * 125 * 3 * steps
    out = {"meshes": len(meshes), "elements": E, "fields_per_mesh": 3, "steps": steps,
           "elements_min": min(m.n_elem for m in meshes), "elements_max": max(m.n_elem for m in meshes),
           "value": pts / t_bat, "unit": "GLL-point updates/s", "launch_groups": best,
           "ms_per_step_batched": 1e3 * t_bat / steps,
           "ms_per_step_by_launch_groups": {str(g): 1e3 * t / steps for g, t in times.items()},
           "ms_per_step_sequential": 1e3 * t_seq / steps, "speedup_vs_sequential": t_seq / t_bat}
    for s in solvers:
        s.close()
    for m in meshes:
        m.close()
    return out


def surrogate_run(pygfx, args):
    """Learned Green's-function surrogate on the layered mesh: device tile of
    strain Green's tensors (6 x 3) at the GLL nodes of the recording cells to
    1.8 km, one sample per snapshot; MLP (64 Fourier frequencies, 6 x 256
    GELU, BF16 MFMA) trained on it; 1M seeded queries scored against GLL
    interpolation of the tile, with fallback in (cell, time window) regions
    whose validated error is above 1 % of the tile's RMS.  One result per
    configuration in --surrogate-configs (a JSON list of experiment keyword
    overrides), all on the same tile."""
    from pygfx import greens
    dims = (40, 40, 20) if args.small else (100, 100, 50)
    t0 = time.perf_counter()
    mesh = pygfx.build_box(cfg2_spec(pygfx, *dims), copy=False)
    log(f"mesh: {mesh.n_elem} elements ({time.perf_counter() - t0:.1f} s)")
    stf = pygfx.ricker(1.5, 1.0, 1e20, mesh.dt, mesh.nsteps)
    src = (0.5 * (mesh.bounds[0] + mesh.bounds[1]), 0.5 * (mesh.bounds[2] + mesh.bounds[3]), 0.0)
    tile, info = greens.build_tile(mesh, src, stf, depth_max=1800.0, log=log,
                                   assembly=pygfx.GFX_ASSEMBLY_GATHER)
    del mesh
    # default: envelope on, moveout at the top layer's Rayleigh speed (about
    # 2.7 km/s), Fourier bandwidths 1/20 km (x, y), 1/5 km (z), 0.7 Hz (time)
    default = {"slowness": 1 / 2700.0, "freq": [5e-5, 2e-4, 0.7]}
    configs = json.loads(args.surrogate_configs) if args.surrogate_configs else [default]
    outs = []
    for cfg in configs:
        kw = dict(iters=args.surrogate_iters, n_query=1_000_000, bound=0.01)
        kw.update(cfg)
        if "freq" in kw:
            kw["freq"] = tuple(kw["freq"])
        res = greens.experiment(tile, info["dt_snap"], src[:2], log=log, **kw)
        out = {"metric": "surrogate rel L2 vs GLL interpolation of the tile (1M seeded queries)",
               "workload": f"layered {dims[0]}x{dims[1]}x{dims[2]}, x/y/z forces at the surface centre, "
                           f"tile = GLL nodes of recording cells to 1.8 km",
               "tile": {k: v for k, v in info.items() if k not in ("cell_ids", "source")},
               "config": cfg}
        out.update(res)
        log(json.dumps(out))
        outs.append(out)
    tile.close()
    return outs


def host_threads():
    """Threads this job may use on the host: the CPUs in its affinity set, capped by
    OMP_NUM_THREADS when the environment sets it (a GPU job's share of a shared host)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    env = os.environ.get("OMP_NUM_THREADS", "")
    if env.isdigit() and int(env) > 0:
        n = min(n, int(env))
    return max(1, n)


def cpu_baseline(seconds=15.0):
    """FP64 CPU restatement of the reference solver (oracle/) on a 20 x 20 x 10
    sub-box of the same 0.6 km layered discretisation: one thread (the
    reference's single-process solver) and all host threads (its element-local and nodal loops
    parallel; results identical).  The all-threads figure is the baseline."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import pygfx
    import pyoracle

    mesh = pygfx.build_box(cfg2_spec(pygfx, nx=20, ny=20, nz=10, duration=30.0))
    cells, w = mesh.locate_source(6000.0, 6000.0, 0.0)
    stf = pygfx.ricker(1.5, 1.0, 1e20, mesh.dt, mesh.nsteps)

    def timed(threads, budget):
        got = pyoracle.set_threads(threads)
        probe = pyoracle.Run(mesh.arrays, mesh.n_elem, 5, mesh.n_node, mesh.dt, 5, mesh.snapshot_stride,
                             2, cells, w, stf, mesh.rec_cell)
        t0 = time.perf_counter()
        probe.run()
        per = (time.perf_counter() - t0) / 5
        n = max(5, min(mesh.nsteps, int(budget / per)))
        run = pyoracle.Run(mesh.arrays, mesh.n_elem, 5, mesh.n_node, mesh.dt, n, mesh.snapshot_stride,
                           2, cells, w, stf, mesh.rec_cell)
        t0 = time.perf_counter()
        run.run()
        dt = time.perf_counter() - t0
        return got, n, dt, mesh.n_elem * 125 * n / dt

    one = timed(1, seconds * 0.5)
    many = timed(host_threads(), seconds * 0.5)
    return {"value": many[3], "unit": "GLL-point updates/s", "cores": many[0], "kind": "port",
            "sample": f"oracle (FP64 C restatement of the reference CPU step, element-local and nodal loops on {many[0]} "
                      f"threads: every host thread this job may use, os.sched_getaffinity capped by OMP_NUM_THREADS "
                      f"({os.environ.get('OMP_NUM_THREADS', 'unset')}), of {os.cpu_count()} CPUs on the host), "
                      f"{mesh.n_elem}-element sub-box of the 0.6 km layered mesh, z force, "
                      f"{many[1]} steps with C-PML and recording, {many[2]:.1f} s",
            "value_1_core": one[3],
            "sample_1_core": f"same, 1 thread, {one[1]} steps, {one[2]:.1f} s"}


def cpml_complete_leg(pygfx, mesh, F, lc, st, stf, args, value_ref, launch, dims):
    """The batched workload again with the complete C-PML (stretched stresses
    and inertia terms in the PML elements' own kernel, the reference's profiles
    of this mesh): whole-step throughput and the PML kernel's time, bytes and
    HBM rate (events around its launch)."""
    t0 = time.perf_counter()
    s = pygfx.Solver(mesh, n_fields=F, record_capacity=2, cpml_complete=True, **lc)
    for f in range(F):
        x, y = st[f // 3]
        cells, w = mesh.locate_source(x, y, 0.0)
        s.set_source(f, f % 3, cells, w, stf)
    log(f"complete C-PML solver ready ({time.perf_counter() - t0:.1f} s)")
    t = time_run(pygfx, s, args.steps, args.warmup)
    value = mesh.n_elem * 125 * F * args.steps / t["wall_s"]
    ms_pml, n_pml = t["element_pml"]
    Q = s.query("cpml_points")
    mem = s.query("cpml_memories")
    # per PML point: scaled geometry (9), lambda, mu, node id, residual entry and
    # 28 coefficients; per field and point u and w read, the residual written,
    # and the memories in use read and written
    pml_bytes = (Q * 4.0 * (9 + 4 + 28) + F * (Q * 36.0 + 8.0 * mem)) if n_pml else None
    out = {"value": value, "unit": "GLL-point updates/s", "ms_per_step": 1e3 * t["wall_s"] / args.steps,
           "vs_reference_cpml": value / value_ref,
           "ms_element_per_step": t["element"][0] / args.steps,
           "ms_pml_kernel_per_step": (ms_pml / args.steps) if n_pml else None,
           "pml_elements": mesh.n_pml_elem, "pml_points": int(Q),
           "memories_in_use_per_point": mem / Q if Q else 0.0,
           "pml_kernel_bytes_per_step": pml_bytes,
           "pml_kernel_gbs": (pml_bytes / (ms_pml / args.steps * 1e-3) / 1e9) if n_pml and ms_pml > 0 else None,
           "pml_kernel_traffic": traffic_from_profiles("k_element_cpml", launch, F, dims)[0],
           "note": "the reference's own profiles of this mesh (pml_profile 0); PML kernel bytes: per PML "
                   "point the geometry, moduli, node id, residual entry and 28 coefficients, and per field "
                   "u and w read, the residual written and the memories in use read and written"}
    s.close()
    return out


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--fields", type=int, default=24)
    ap.add_argument("--assembly", choices=["auto", "colored", "gather", "atomic"], default="auto",
                    help="auto = the library autotuner picks assembly, element kernel and block size")
    ap.add_argument("--epb", type=int, default=0)
    ap.add_argument("--kernel", choices=["auto", "block", "wave", "pair", "pencil"], default="auto")
    ap.add_argument("--node-groups", type=int, default=0, help="node-kernel field groups (1, 2, 4, 8)")
    ap.add_argument("--small", action="store_true", help="40x40x20 mesh instead of 100x100x50 (smoke)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-uneven", action="store_true", help="skip the 16-mesh uneven batch line")
    ap.add_argument("--general-geometry", action="store_true",
                    help="run every element on the general (full dxi/dx) form, no axis-aligned path")
    ap.add_argument("--attenuation", type=int, default=0, metavar="N_SLS",
                    help="shear attenuation with this many standard linear solids (step benchmark only)")
    ap.add_argument("--surrogate", action="store_true",
                    help="instead of the step benchmark: Green tile of the layered mesh (x, y, z forces at the "
                         "surface centre, cells to 1.8 km, every snapshot), MLP training, 1M scored queries")
    ap.add_argument("--surrogate-iters", type=int, default=16000)
    ap.add_argument("--surrogate-configs", default="",
                    help='JSON list of experiment overrides, e.g. \'[{"sigma": 1.0}, {"envelope": false}]\'')
    ap.add_argument("--no-cpml-complete", action="store_true",
                    help="skip the complete C-PML line (the batched fields again with the stretched-stress C-PML)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run 3 device-to-device copies of 1 GiB (PMC counter calibration)")
    ap.add_argument("--pmc-summary", nargs=2, metavar=("FETCH_CSV", "WRITE_CSV"),
                    help="print per-kernel HBM traffic from two rocprofv3 --pmc passes and exit")
    args = ap.parse_args()
    if args.pmc_summary:
        print(pmc_summary(*args.pmc_summary))
        return
    if args.gpus != 1:
        print("bench.py measures one GPU: this build has no multi-GPU path (each GPU would run "
              "its own fields or meshes with no communication); run with --gpus 1.", file=sys.stderr)
        sys.exit(2)

    import pygfx

    if pygfx.device_count() < 1:
        print("no HIP device visible", file=sys.stderr)
        sys.exit(1)
    if args.calibrate_copy:
        copy_calibration()
    if args.surrogate:
        for out in surrogate_run(pygfx, args):
            print(json.dumps(out), flush=True)
        return
    asm_of = {"colored": pygfx.GFX_ASSEMBLY_COLORED, "gather": pygfx.GFX_ASSEMBLY_GATHER,
              "atomic": pygfx.GFX_ASSEMBLY_ATOMIC}
    kern_of = {"auto": pygfx.GFX_ELEMENT_AUTO, "block": pygfx.GFX_ELEMENT_BLOCK,
               "wave": pygfx.GFX_ELEMENT_WAVE, "pair": pygfx.GFX_ELEMENT_PAIR, "pencil": pygfx.GFX_ELEMENT_PENCIL}

    def launch_config(n_fields):
        if args.assembly == "auto":
            t0 = time.perf_counter()
            pick = pygfx.autotune(mesh, n_fields=n_fields, trial_steps=3, record_capacity=0,
                                  general_geometry=args.general_geometry)
            log(f"autotune F={n_fields}: {pick} ({time.perf_counter() - t0:.1f} s)")
            return dict(assembly=pick["assembly"], element_kernel=pick["element_kernel"],
                        elems_per_block=pick["elems_per_block"], node_field_groups=pick["node_field_groups"])
        return dict(assembly=asm_of[args.assembly], element_kernel=kern_of[args.kernel],
                    elems_per_block=args.epb, node_field_groups=args.node_groups)

    names = {pygfx.GFX_ASSEMBLY_COLORED: "colored", pygfx.GFX_ASSEMBLY_GATHER: "gather",
             pygfx.GFX_ASSEMBLY_ATOMIC: "atomic"}
    knames = {pygfx.GFX_ELEMENT_BLOCK: "block", pygfx.GFX_ELEMENT_WAVE: "wave", pygfx.GFX_ELEMENT_PAIR: "pair",
              pygfx.GFX_ELEMENT_PENCIL: "pencil", 0: "auto"}

    t0 = time.perf_counter()
    dims = (40, 40, 20) if args.small else (100, 100, 50)
    want_cpml = not args.no_cpml_complete
    mesh = pygfx.build_box(cfg2_spec(pygfx, *dims, n_sls=args.attenuation, profiles=want_cpml), copy=False)
    log(f"mesh: {mesh.n_elem} elements, {mesh.n_node} nodes, {mesh.n_pml_elem} PML elements, "
        f"{len(mesh.rec_cell)} recording cells, dt={mesh.dt:.5g} s, stride={mesh.snapshot_stride}, "
        f"nsteps={mesh.nsteps} ({time.perf_counter() - t0:.1f} s)")
    E = mesh.n_elem
    G = mesh.n_node
    cx = 0.5 * (mesh.bounds[0] + mesh.bounds[1])
    cy = 0.5 * (mesh.bounds[2] + mesh.bounds[3])
    stf = pygfx.ricker(1.5, 1.0, 1e20, mesh.dt, mesh.nsteps)

    result = {}
    plain = None
    if not args.no_plain:
        lc1 = launch_config(1)
        t0 = time.perf_counter()
        s1 = pygfx.Solver(mesh, n_fields=1, record_capacity=2, general_geometry=args.general_geometry, **lc1)
        cells, w = mesh.locate_source(cx, cy, 0.0)
        s1.set_source(0, 2, cells, w, stf)
        log(f"plain solver ready ({time.perf_counter() - t0:.1f} s)")
        tp = time_run(pygfx, s1, args.steps, args.warmup)
        plain = dict(value=E * 125 * 1 * args.steps / tp["wall_s"],
                     ms_per_step=1e3 * tp["wall_s"] / args.steps,
                     ms_element_per_step=tp["element"][0] / args.steps,
                     ms_node_per_step=tp["node"][0] / args.steps,
                     node_updates_per_s=G * args.steps / tp["wall_s"],
                     bytes_per_step=s1.algorithmic_bytes(0),
                     launch={"assembly": names[int(s1.query("assembly"))],
                             "element_kernel": knames[int(s1.query("element_kernel"))],
                             "elems_per_block": int(s1.query("elems_per_block")),
                             "node_field_groups": int(s1.query("node_field_groups"))})
        plain["hbm_gbs_whole_step"] = plain["bytes_per_step"] / (plain["ms_per_step"] * 1e-3) / 1e9
        log("plain:", json.dumps(plain))
        s1.close()
        del s1

    F = args.fields
    lcb = launch_config(F)
    t0 = time.perf_counter()
    sb = pygfx.Solver(mesh, n_fields=F, record_capacity=2, general_geometry=args.general_geometry, **lcb)
    st = stations(mesh, n=max(1, (F + 2) // 3))
    for f in range(F):
        x, y = st[f // 3]
        cells, w = mesh.locate_source(x, y, 0.0)
        sb.set_source(f, f % 3, cells, w, stf)
    log(f"batched solver ready, {F} fields ({time.perf_counter() - t0:.1f} s)")
    tb = time_run(pygfx, sb, args.steps, args.warmup)
    value = E * 125 * F * args.steps / tb["wall_s"]
    ms_step = 1e3 * tb["wall_s"] / args.steps

    # roofline of the dominant kernel: the one with the most device time per
    # step (HIP events on the solver stream around each kernel)
    launch = {"assembly": names[int(sb.query("assembly"))],
              "element_kernel": knames[int(sb.query("element_kernel"))],
              "elems_per_block": int(sb.query("elems_per_block")),
              "node_field_groups": int(sb.query("node_field_groups"))}
    kern = {"element": ({"wave": "k_element_wave", "pair": "k_element_pair",
                         "pencil": "k_element_pencil"}.get(launch["element_kernel"], "k_element"), 1),
            "node_special": ("k_node_special", 3), "node_regular": ("k_node_regular", 4)}
    dom = max(kern, key=lambda k: tb[k][0])
    kname, which = kern[dom]
    ms_dom, n_dom = tb[dom]
    launches_per_step = n_dom / args.steps
    alg_launch = sb.algorithmic_bytes(which) / launches_per_step
    ms_launch = ms_dom / n_dom
    achieved = alg_launch / (ms_launch * 1e-3) / 1e9
    traffic, traffic_launch = traffic_from_profiles(kname, launch, F, dims)
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "GLL-point updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"layered {dims[0]}x{dims[1]}x{dims[2]} hexes of 0.6 km, order 4, "
                               f"C-PML 5, x/y/z forces at {len(st)} surface stations = {F} fields",
                   "elements": E, "nodes": G, "fields": F, "pml_elements": mesh.n_pml_elem,
                   "recording_cells": int(len(mesh.rec_cell)), "snapshot_stride": mesh.snapshot_stride,
                   "dt_s": mesh.dt, "parallelism": "none (1 GPU)",
                   "launch": launch,
                   "axis_aligned_elements_on_diagonal_form": int(sb.query("diag_elements")),
                   "attenuation_solids": int(sb.query("n_sls"))},
        "node_updates_per_s": G * F * args.steps / tb["wall_s"],
        "plain": plain,
        "batched_vs_plain": (value / plain["value"]) if plain else None,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_pass_launch": traffic_launch,
                     "kernel": kname, "algorithmic_bytes_per_launch": alg_launch,
                     "ms_per_launch": ms_launch, "launches_per_step": launches_per_step},
        "element_lds": element_lds(launch, E, F, tb["element"][0] / args.steps),
        "whole_step_hbm_gbs": sb.algorithmic_bytes(0) / (ms_step * 1e-3) / 1e9,
        "ms_per_step_by_kernel": {k: tb[k][0] / args.steps for k in ("element", "node_special",
                                                                      "node_regular", "record")},
        "gbs_by_kernel": {k: sb.algorithmic_bytes(w) / (tb[k][0] / args.steps * 1e-3) / 1e9
                          for k, (_, w) in kern.items() if tb[k][0] > 0},
        "ms_record_per_snapshot": (tb["record"][0] / tb["record"][1]) if tb["record"][1] else None,
    }
    # records leave the GPU once per snapshot: time copying one snapshot of every
    # field to host FP64 buffers (the forward driver's drain), and the
    # throughput with that copy added for every snapshot step (not overlapped)
    if sb.n_snapshots > 0:
        t0 = time.perf_counter()
        for f in range(F):
            sb.get_record(f, sb.n_snapshots - 1)
        drain = time.perf_counter() - t0
        snaps = args.steps / mesh.snapshot_stride
        result["record_drain"] = {
            "s_per_snapshot": drain, "bytes_per_snapshot": int(F * len(mesh.rec_cell) * 125 * 6 * 4),
            "value_with_drain": E * 125 * F * args.steps / (tb["wall_s"] + snaps * drain),
            "note": "one snapshot of every field (strain at the recording cells) copied to host "
                    "FP64 arrays after the timed steps; value_with_drain adds that copy at every snapshot "
                    "step without overlap (the forward driver overlaps it with the next segment)"}
    sb.close()
    del sb
    if want_cpml:
        result["cpml_complete"] = cpml_complete_leg(pygfx, mesh, F, lcb, st, stf, args, value, launch, dims)
        log("complete C-PML:", json.dumps(result["cpml_complete"]))
    if not args.no_uneven and not args.small:
        result["uneven_batch"] = uneven_workload(pygfx, args.steps, args.warmup)
        log("uneven:", json.dumps(result["uneven_batch"]))
        # many small meshes: per-launch cost dominates each mesh's step
        result["uneven_batch_small"] = uneven_workload(pygfx, args.steps, args.warmup, n=64, lo=500.0, hi=4000.0,
                                                       pml=(1, 2))
        log("uneven small:", json.dumps(result["uneven_batch_small"]))
    if not args.no_cpu_baseline:
        log("cpu baseline ...")
        result["cpu_baseline"] = cpu_baseline()
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
