"""Green's-function tiles on the device and the learned surrogate for them.

Pipeline (everything stays in HBM):
  1. a solver with three fields (x, y, z point forces at one source) records
     strain at the recording cells every snapshot;
  2. each snapshot is merged into a tile of Green's tensors at the unique GLL
     nodes of the chosen cells (mass-weighted average of the element-local
     strain, postprocess/cpp/main.cpp:354-426), [time][node][6][3];
  3. a Fourier-feature MLP is trained on (node, time) samples of the tile;
  4. queries (x, y, z, time index) are answered by the MLP, except in cells
     whose validation error exceeds the bound, which fall back to GLL
     interpolation of the tile (the judge, gll_interpolator.py:117-227).
"""
from __future__ import annotations

import time

import numpy as np

from . import GreenTile, MLP, Solver, green_capture_map


def tile_cells(mesh, depth_max):
    """Recording cells whose centre lies above depth_max (a shallow tile)."""
    ib = mesh.ibool.reshape(-1, 125)[mesh.rec_cell]
    zc = mesh.node_xyz.reshape(-1, 3)[ib][:, :, 2].mean(axis=1)
    return mesh.rec_cell[zc <= depth_max]


def tile_nodes(mesh, cells):
    ib = mesh.ibool.reshape(-1, 125)[np.asarray(cells)]
    nodes, inv = np.unique(ib.ravel(), return_inverse=True)
    return nodes, mesh.node_xyz.reshape(-1, 3)[nodes], inv.reshape(-1, 125).astype(np.int32)


def build_tile(mesh, source_xyz, stf, depth_max, nsteps=None, ring=8, log=print, **solver_kw):
    """Run x, y, z forces at source_xyz and capture every snapshot into a tile.
    Returns (tile, info)."""
    nsteps = mesh.nsteps if nsteps is None else nsteps
    stride = mesh.snapshot_stride
    n_snap = (nsteps + stride - 1) // stride
    cells = tile_cells(mesh, depth_max)
    nodes, xyz, cidx = tile_nodes(mesh, cells)
    t0 = time.perf_counter()
    s = Solver(mesh, n_fields=3, record_capacity=ring, record_uva=False, **solver_kw)
    sc, sw = mesh.locate_source(*source_xyz)
    for d in range(3):
        s.set_source(d, d, sc, sw, stf)
    tile = GreenTile(n_snap, xyz, cidx)
    # recording points of the tile cells: their positions in the solver's recording list
    off, rp, w = green_capture_map(mesh, mesh.rec_cell, nodes)
    tile.set_capture_map(off, rp, w)
    log(f"tile: {len(cells)} cells, {len(nodes)} nodes, {n_snap} samples "
        f"({n_snap * len(nodes) * 72 / 1e9:.1f} GB on the device); setup {time.perf_counter() - t0:.1f} s")
    t0 = time.perf_counter()
    done = 0
    while s.step < nsteps:
        s.run(min(stride * ring, nsteps - s.step))
        for k in range(done, s.n_snapshots):
            tile.capture(s, k, [0, 1, 2], k)
        done = s.n_snapshots
    t_run = time.perf_counter() - t0
    log(f"tile: {nsteps} steps x 3 fields + capture in {t_run:.1f} s")
    s.close()
    info = dict(cells=len(cells), nodes=len(nodes), samples=n_snap, tile_bytes=n_snap * len(nodes) * 72,
                solve_s=t_run, dt_snap=mesh.dt * stride, cell_ids=cells)
    return tile, info


def random_queries(tile, n, seed, t_count):
    rng = np.random.default_rng(seed)
    lo, hi = tile.xyz.min(axis=0), tile.xyz.max(axis=0)
    pts = lo + rng.random((n, 3)) * (hi - lo)
    tix = rng.integers(0, t_count, n).astype(np.int32)
    return pts, tix


def experiment(tile, dt_snap, iters=4000, n_query=1_000_000, bound=0.01, log=print, **mlp_kw):
    """Train the MLP on the tile and score it against the judge.

    Cells are flagged for fallback on a validation set (seed 6, n_query/4
    points); the scores are on an independent test set (seed 5, n_query
    points).  Rel L2 is over all 18 outputs of all points of a set."""
    lo = np.r_[tile.xyz.min(axis=0), 0.0]
    hi = np.r_[tile.xyz.max(axis=0), (tile.n_time - 1) * dt_snap]
    # per-output scale: RMS of the tile values over a sample of times
    ts = np.linspace(0, tile.n_time - 1, min(tile.n_time, 16)).astype(int)
    sq = np.zeros(18)
    for t in ts:
        v = tile.get_values(int(t)).astype(np.float64)
        sq += (v * v).mean(axis=0)
    scale = np.sqrt(sq / len(ts))
    scale = np.where(scale > 0, scale, 1.0)
    net = MLP(lo, hi, scale, **mlp_kw)
    t0 = time.perf_counter()
    loss = net.train(tile, 0.0, dt_snap, iters, seed=1)
    t_train = time.perf_counter() - t0
    log(f"surrogate: {net.n_params} parameters, {iters} Adam iterations in {t_train:.1f} s, loss {loss:.4g}")

    def score(pts, tix):
        t0 = time.perf_counter()
        jv, cell = tile.interpolate(pts, tix)
        tj = time.perf_counter() - t0
        t0 = time.perf_counter()
        mv = net.forward(np.c_[pts, tix * dt_snap])
        tm = time.perf_counter() - t0
        return jv, cell, mv, tj, tm

    vp, vt = random_queries(tile, max(1, n_query // 4), 6, tile.n_time)
    jv, cell, mv, _, _ = score(vp, vt)
    ok = cell >= 0
    err2 = np.bincount(cell[ok], weights=((mv[ok] - jv[ok]).astype(np.float64) ** 2).sum(axis=1),
                       minlength=tile.n_cell)
    ref2 = np.bincount(cell[ok], weights=(jv[ok].astype(np.float64) ** 2).sum(axis=1), minlength=tile.n_cell)
    cell_rel = np.sqrt(err2 / np.maximum(ref2, 1e-300))
    fallback = (cell_rel > bound) | (ref2 == 0)
    qp, qt = random_queries(tile, n_query, 5, tile.n_time)
    jv, cell, mv, tj, tm = score(qp, qt)
    rel = float(np.linalg.norm(mv - jv) / np.linalg.norm(jv))
    out, n_fb = net.query(tile, fallback.astype(np.uint8), 0.0, dt_snap, qp, qt)
    rel_q = float(np.linalg.norm(out - jv) / np.linalg.norm(jv))
    res = dict(params=int(net.n_params), iterations=iters, train_s=t_train, final_loss=loss,
               rel_l2_mlp=rel, cells_flagged=int(fallback.sum()), cells=int(tile.n_cell),
               queries=n_query, fallback_fraction=n_fb / n_query, rel_l2_with_fallback=rel_q,
               mlp_bytes_bf16=int(net.n_params) * 2, tile_bytes=int(tile.n_time) * tile.n_node * 72,
               judge_query_s=tj, mlp_query_s=tm)
    net.close()
    return res
