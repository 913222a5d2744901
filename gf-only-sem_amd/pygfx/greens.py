"""Green's-function tiles on the device and the learned surrogate for them.

Pipeline (everything stays in HBM):
  1. a solver with three fields (x, y, z point forces at one source) records
     strain at the recording cells every snapshot;
  2. each snapshot is merged into a tile of Green's tensors at the unique GLL
     nodes of the chosen cells (mass-weighted average of the element-local
     strain, postprocess/cpp/main.cpp:354-426), [time][node][6][3];
  3. a Fourier-feature MLP is trained on (node, time) samples of the tile,
     with the geometric decay from the source divided out (a radial envelope
     fitted to the tile);
  4. queries (x, y, z, time index) are answered by the MLP, except in
     (cell, time window) regions whose validated error is above the bound,
     which fall back to GLL interpolation of the tile (the judge,
     gll_interpolator.py:117-227).

The gate: a region is served by the network when its RMS error on the
validation queries is at most `bound` times the RMS of the judge's values
over all validation queries.  Every served region then contributes at most
bound^2 of the total squared norm per query, so the aggregate rel L2 of the
combined answer is at most `bound` (on the validation distribution).  The
same gate applied to a zero predictor shows how much of that comes from the
network rather than from regions whose values are small anyway.
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
                solve_s=t_run, dt_snap=mesh.dt * stride, cell_ids=cells, source=tuple(source_xyz))
    return tile, info


def random_queries(tile, n, seed, t_count):
    rng = np.random.default_rng(seed)
    lo, hi = tile.xyz.min(axis=0), tile.xyz.max(axis=0)
    pts = lo + rng.random((n, 3)) * (hi - lo)
    tix = rng.integers(0, t_count, n).astype(np.int32)
    return pts, tix


def output_scale(tile, n_times=16):
    """Per-output RMS of the tile over a sample of times (1 where zero)."""
    ts = np.linspace(0, tile.n_time - 1, min(tile.n_time, n_times)).astype(int)
    sq = np.zeros(18)
    for t in ts:
        v = tile.get_values(int(t)).astype(np.float64)
        sq += (v * v).mean(axis=0)
    scale = np.sqrt(sq / len(ts))
    return np.where(scale > 0, scale, 1.0)


def radial_envelope(tile, scale, src_xy, n_bins=64, n_times=32, floor=1e-3):
    """RMS of the scaled tile values per bin of horizontal distance from the
    source (over a sample of times and all 18 outputs), empty bins filled
    from their neighbours, floored at floor x the largest bin, smoothed over
    three bins in log space.  Returns (r_max, knots)."""
    r = np.hypot(tile.xyz[:, 0] - src_xy[0], tile.xyz[:, 1] - src_xy[1])
    r_max = float(r.max()) * (1 + 1e-9) + 1e-9
    b = np.minimum((r / r_max * n_bins).astype(int), n_bins - 1)
    ts = np.linspace(0, tile.n_time - 1, min(tile.n_time, n_times)).astype(int)
    s2 = np.zeros(n_bins)
    for t in ts:
        v = tile.get_values(int(t)).astype(np.float64) / scale
        s2 += np.bincount(b, weights=(v * v).sum(axis=1), minlength=n_bins)
    cnt = np.bincount(b, minlength=n_bins) * len(ts) * 18
    have = cnt > 0
    rms = np.zeros(n_bins)
    rms[have] = np.sqrt(s2[have] / cnt[have])
    idx = np.arange(n_bins)
    rms = np.interp(idx, idx[have], rms[have])
    rms = np.maximum(rms, floor * rms.max())
    lg = np.log(rms)
    lg = np.convolve(np.r_[lg[0], lg, lg[-1]], np.ones(3) / 3, mode="valid")
    return r_max, np.exp(lg).astype(np.float32)


def frequency_box(tile, t_lo, t_hi, freq_xy, freq_z, freq_t, sigma):
    """Input range [lo, hi] per dimension such that Fourier frequencies drawn
    from N(0, sigma^2) in scaled units have standard deviation freq_* in
    physical units (cycles per metre, per second); [t_lo, t_hi] is the range
    of the network's time input."""
    c = np.r_[0.5 * (tile.xyz.min(axis=0) + tile.xyz.max(axis=0)), 0.5 * (t_lo + t_hi)]
    f = np.array([freq_xy, freq_xy, freq_z, freq_t])
    half = sigma / f  # physical half-width of one scaled unit
    return c - half, c + half


def region_sums(tile, cell, tix, pred, ref, n_window):
    """Per (cell, time window) region: query count, squared error, squared value."""
    ok = cell >= 0
    nreg = tile.n_cell * n_window
    reg = cell[ok].astype(np.int64) * n_window + (tix[ok].astype(np.int64) * n_window) // tile.n_time
    e2 = np.bincount(reg, weights=((pred[ok] - ref[ok]).astype(np.float64) ** 2).sum(axis=1), minlength=nreg)
    r2 = np.bincount(reg, weights=(ref[ok].astype(np.float64) ** 2).sum(axis=1), minlength=nreg)
    cnt = np.bincount(reg, minlength=nreg)
    return cnt, e2, r2


def fallback_bytes(tile, flags, n_window):
    """Tile bytes the judge still needs: for each window, the unique nodes of
    the cells flagged in it, times the window's samples, 72 B each."""
    f = flags.reshape(tile.n_cell, n_window)
    total = 0
    for w in range(n_window):
        t_b = w * tile.n_time // n_window
        t_e = (w + 1) * tile.n_time // n_window
        cells = np.nonzero(f[:, w])[0]
        if cells.size:
            total += np.unique(tile.cells[cells]).size * (t_e - t_b) * 72
    return total


def experiment(tile, dt_snap, source_xy, iters=4000, n_query=1_000_000, n_valid=8_000_000, n_window=10,
               bound=0.01, freq=(1 / 3000.0, 1 / 4000.0, 1.0), envelope=True, slowness=0.0, log=print,
               **mlp_kw):
    """Train the MLP on the tile and score it against the judge.

    Regions are flagged for fallback on a validation set (seed 6, n_valid
    points); the scores are on an independent test set (seed 5, n_query
    points).  Rel L2 is over all 18 outputs of all points of a set.
    slowness (s/m) > 0 feeds the network the retarded time t - slowness * r."""
    t_span = (tile.n_time - 1) * dt_snap
    scale = output_scale(tile)
    sigma = float(mlp_kw.pop("sigma", 1.0))
    r_all = np.hypot(tile.xyz[:, 0] - source_xy[0], tile.xyz[:, 1] - source_xy[1])
    lo, hi = frequency_box(tile, -slowness * float(r_all.max()), t_span, *freq, sigma)
    net = MLP(lo, hi, scale, sigma=sigma, **mlp_kw)
    r_max, env = radial_envelope(tile, scale, source_xy) if envelope else (1.0, np.zeros(0, np.float32))
    net.set_envelope(source_xy[0], source_xy[1], r_max, env, slowness=slowness)
    t0 = time.perf_counter()
    loss = net.train(tile, 0.0, dt_snap, iters, seed=1)
    t_train = time.perf_counter() - t0
    log(f"surrogate: {net.n_params} parameters, {iters} Adam iterations in {t_train:.1f} s, loss {loss:.4g}")

    # validation: region error map from n_valid points, in chunks
    nreg = tile.n_cell * n_window
    cnt = np.zeros(nreg, np.int64)
    e2 = np.zeros(nreg)
    r2 = np.zeros(nreg)
    rng = np.random.default_rng(6)
    lo3, hi3 = tile.xyz.min(axis=0), tile.xyz.max(axis=0)
    chunk = 1_000_000
    for b in range(0, n_valid, chunk):
        m = min(chunk, n_valid - b)
        pts = lo3 + rng.random((m, 3)) * (hi3 - lo3)
        tix = rng.integers(0, tile.n_time, m).astype(np.int32)
        jv, cell = tile.interpolate(pts, tix)
        mv = net.forward(np.c_[pts, tix * dt_snap])
        c, e, r = region_sums(tile, cell, tix, mv, jv, n_window)
        cnt += c
        e2 += e
        r2 += r
    rms_all = np.sqrt(r2.sum() / max(cnt.sum(), 1))
    seen = cnt > 0
    limit = bound * rms_all
    rms_err = np.where(seen, np.sqrt(e2 / np.maximum(cnt, 1)), np.inf)
    rms_ref = np.where(seen, np.sqrt(r2 / np.maximum(cnt, 1)), np.inf)
    serve = rms_err <= limit
    serve_zero = rms_ref <= limit
    fallback = ~serve

    # test set
    qp, qt = random_queries(tile, n_query, 5, tile.n_time)
    t0 = time.perf_counter()
    jv, cell = tile.interpolate(qp, qt)
    t_judge = time.perf_counter() - t0
    t0 = time.perf_counter()
    mv = net.forward(np.c_[qp, qt * dt_snap])
    t_mlp = time.perf_counter() - t0
    nj = np.linalg.norm(jv.astype(np.float64))
    rel = float(np.linalg.norm((mv - jv).astype(np.float64)) / nj)
    out, n_fb = net.query(tile, fallback.astype(np.uint8), 0.0, dt_snap, qp, qt)
    rel_q = float(np.linalg.norm((out - jv).astype(np.float64)) / nj)
    # the zero predictor under the same gate
    reg = np.where(cell >= 0, cell.astype(np.int64) * n_window + (qt.astype(np.int64) * n_window) // tile.n_time, -1)
    served_zero = (reg >= 0) & serve_zero[np.maximum(reg, 0)]
    zero_out = np.where(served_zero[:, None], 0.0, jv)
    rel_zero = float(np.linalg.norm((zero_out - jv).astype(np.float64)) / nj)
    # combined gate: per region the cheapest answer within the bound -- zero
    # where the values themselves are, else the network where its error is,
    # else the judge.  Without the network the same gate has zero or judge.
    policy = np.where(serve_zero, 2, np.where(serve, 0, 1)).astype(np.uint8)
    out_c, n_judge_c = net.query(tile, policy, 0.0, dt_snap, qp, qt)
    rel_c = float(np.linalg.norm((out_c - jv).astype(np.float64)) / nj)
    pol_q = np.where(reg >= 0, policy[np.maximum(reg, 0)], 1)
    net_only = serve & ~serve_zero
    tb = tile.n_time * tile.n_node * 72
    res = dict(params=int(net.n_params), iterations=iters, train_s=t_train, final_loss=loss,
               sigma=sigma, freq_xy=freq[0], freq_z=freq[1], freq_t=freq[2], envelope=bool(envelope),
               slowness=slowness,
               rel_l2_mlp=rel, n_window=n_window, regions=int(nreg), regions_served=int(serve.sum()),
               regions_served_by_zero=int(serve_zero.sum()), valid_queries=int(n_valid), queries=n_query,
               fallback_fraction=n_fb / n_query, rel_l2_with_fallback=rel_q,
               zero_predictor_served_fraction=float(served_zero.mean()), rel_l2_zero_gate=rel_zero,
               mlp_bytes_bf16=int(net.n_params) * 2, tile_bytes=tb,
               fallback_tile_bytes=int(fallback_bytes(tile, fallback, n_window)),
               zero_gate_tile_bytes=int(fallback_bytes(tile, ~serve_zero, n_window)),
               combined_rel_l2=rel_c, combined_judge_fraction=n_judge_c / n_query,
               combined_zero_fraction=float((pol_q == 2).mean()),
               combined_network_fraction=float((pol_q == 0).mean()),
               network_only_regions=int(net_only.sum()),
               combined_tile_bytes=int(fallback_bytes(tile, policy == 1, n_window)),
               judge_query_s=t_judge, mlp_query_s=t_mlp)
    net.close()
    return res
