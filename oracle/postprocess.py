"""Green's function tile post-processor, restated in numpy -- TEST INFRASTRUCTURE ONLY.

The checker for the device tile writer.  It restates postprocess/cpp/main.cpp
(record merge, Green's tensor layout, whole-cell tiling) and
postprocess/cpp/writer.hh (tile file contents) over arrays already read from
record files, so that a test can hold the device's tiles against what the
reference's post-processor makes of the same records.  Only tests/ may import
it.

FP64 throughout, with the reference loop's order of operations: per node, the
products strain*weight are added one at a time in the order record file, cell,
point (numpy's ufunc.at applies unbuffered updates in index order), then
multiplied by 1 / (sum of weights).
"""
from __future__ import annotations

import math

import numpy as np

NPC = 125


def merge_direction(files, steps, cell_mass=None, n_model_cell=0, ngll=5):
    """main.cpp:121-429 for one force direction.

    files: per rank file, in the order the post-processor lists them, a dict
        with gll_node_ids [n], gll_node_coords [n, 3], cell_gll_node_index
        [n_rec_cell, 125] and recording_cell_model_index [n_rec_cell] (or None).
    steps: per snapshot, a list over the same files of dicts with "strain"
        [n_rec_cell, 125, 6] and optionally "displacement" [n_rec_cell, 125, 3].
    cell_mass: model /field/cell/mass as [n_model_cell, 125] (None = no mass).
    """
    # first pass: union of node ids in first-appearance order (main.cpp:133-248)
    where = {}
    ids, coords, maps = [], [], []
    for f in files:
        l2g = np.empty(len(f["gll_node_ids"]), np.int64)
        for li, gid in enumerate(np.asarray(f["gll_node_ids"], np.int64).tolist()):
            if gid not in where:
                where[gid] = len(ids)
                ids.append(gid)
                coords.append(np.asarray(f["gll_node_coords"], np.float64)[li])
            l2g[li] = where[gid]
        maps.append(l2g)
    ng = len(ids)
    cgi, rcm = [], []
    for f, l2g in zip(files, maps):
        loc = np.asarray(f["cell_gll_node_index"], np.int64).reshape(-1, NPC)
        ok = (loc >= 0) & (loc < len(l2g))
        cgi.append(np.where(ok, l2g[np.clip(loc, 0, max(len(l2g) - 1, 0))], -1))
        r = f.get("recording_cell_model_index")
        rcm.append(np.full(len(loc), -1, np.int64) if r is None else np.asarray(r, np.int64))
    cell_index = np.concatenate(cgi) if cgi else np.zeros((0, NPC), np.int64)

    use_mass = cell_mass is not None and ngll > 0 and n_model_cell > 0
    has_disp = bool(steps) and all("displacement" in s for s in steps[0])
    strain = np.zeros((len(steps), ng, 6))
    disp = np.zeros((len(steps), ng, 3))
    for k, group in enumerate(steps):
        wsum = np.zeros(ng)
        for fi, rec in enumerate(group):
            g = cgi[fi]
            valid = (g >= 0) & (g < ng)
            w = np.ones(g.shape)
            if use_mass:
                gc = rcm[fi]
                okc = (gc >= 0) & (gc < n_model_cell)
                w[okc] = np.asarray(cell_mass, np.float64).reshape(-1, NPC)[gc[okc]]
            node = g[valid]
            s6 = np.asarray(rec["strain"], np.float64).reshape(-1, NPC, 6)[valid]
            ww = w[valid]
            # main.cpp:374-378: dst += src*weight in cell, point order
            for c in range(6):
                np.add.at(strain[k, :, c], node, s6[:, c] * ww)
            np.add.at(wsum, node, ww)
            if has_disp:
                u3 = np.asarray(rec["displacement"], np.float64).reshape(-1, NPC, 3)[valid]
                for c in range(3):
                    np.add.at(disp[k, :, c], node, u3[:, c])
        # main.cpp:412-424: both scaled by 1 / (sum of weights) where positive
        pos = wsum > 0.0
        inv = np.ones(ng)
        inv[pos] = 1.0 / wsum[pos]
        strain[k] *= inv[:, None]
        if has_disp:
            disp[k] *= inv[:, None]
    return {
        "gll_node_ids": np.asarray(ids, np.int64),
        "gll_node_coords": np.asarray(coords, np.float64).reshape(-1, 3),
        "cell_gll_node_index": cell_index,
        "strain": strain,
        "displacement": disp if has_disp else None,
        "n_steps": len(steps),
    }


def greens_tensor(fx, fy, fz):
    """main.cpp:554-575: per node the 18 values direction-major (d*6 + c),
    stored under the shape [nt, n, 6, 3]."""
    g = np.stack([fx["strain"], fy["strain"], fz["strain"]], axis=2)  # [nt, n, 3, 6]
    return g.reshape(g.shape[0], g.shape[1], 6, 3)


def displacement_tensor(fx, fy, fz):
    """main.cpp:581-598: [nt, n, 3 (component), 3 (direction)]."""
    return np.stack([fx["displacement"], fy["displacement"], fz["displacement"]], axis=3)


def find_tile_index(i, sizes):
    """writer.hh:28-36."""
    cum = 0
    for t, s in enumerate(sizes):
        cum += s
        if i < cum:
            return t
    return len(sizes) - 1


def tile_cells(cfg, bounds, merged):
    """main.cpp:667-714: {(tx, ty): [merged cell indices]} by the centre of
    points 0 and 124 of each cell, over the interior element grid."""
    xmin, xmax, ymin, ymax = bounds[:4]
    nx, ny = cfg["nx_elements"], cfg["ny_elements"]
    dx = (xmax - xmin) / nx if nx else math.inf
    dy = (ymax - ymin) / ny if ny else math.inf
    tot_x, tot_y = sum(cfg["tilex_elements"]), sum(cfg["tiley_elements"])
    X = merged["gll_node_coords"]
    n_rec = len(merged["gll_node_ids"])
    bins = {}
    for ci, row in enumerate(merged["cell_gll_node_index"]):
        i0, i1 = int(row[0]), int(row[124])
        if not (0 <= i0 < n_rec and 0 <= i1 < n_rec):
            continue
        cx = 0.5 * (X[i0, 0] + X[i1, 0])
        cy = 0.5 * (X[i0, 1] + X[i1, 1])
        ei = math.floor((cx - xmin) / dx) if dx > 0 else 0
        ej = math.floor((cy - ymin) / dy) if dy > 0 else 0
        ei, ej = max(ei, 0), max(ej, 0)
        if nx > 0 and ei >= nx:
            ei = nx - 1
        if ny > 0 and ej >= ny:
            ej = ny - 1
        ii, ij = ei - cfg["pml_xmin"], ej - cfg["pml_ymin"]
        if not (0 <= ii < tot_x and 0 <= ij < tot_y):
            continue
        key = (find_tile_index(ii, cfg["tilex_elements"]), find_tile_index(ij, cfg["tiley_elements"]))
        bins.setdefault(key, []).append(ci)
    return dict(sorted(bins.items()))


def tile_bounds(cfg, bounds, tx, ty):
    """main.cpp:724-741."""
    xmin, xmax, ymin, ymax = bounds[:4]
    dx = (xmax - xmin) / cfg["nx_elements"]
    dy = (ymax - ymin) / cfg["ny_elements"]
    cx = sum(cfg["tilex_elements"][:tx])
    cy = sum(cfg["tiley_elements"][:ty])
    i0 = cfg["pml_xmin"] + cx
    j0 = cfg["pml_ymin"] + cy
    return (xmin + i0 * dx, xmin + (i0 + cfg["tilex_elements"][tx]) * dx,
            ymin + j0 * dy, ymin + (j0 + cfg["tiley_elements"][ty]) * dy)


def downsample_stf(stf_t, stf_values, nt):
    """main.cpp:518-540: every (nstf // nt)-th sample (stride at least 1),
    clamped to the last."""
    nstf = len(stf_t)
    if nstf == 0 or nt == 0:
        return np.zeros(0), np.zeros(0)
    stride = max(1, nstf // nt)
    idx = np.minimum(np.arange(nt) * stride, nstf - 1)
    return np.asarray(stf_t, np.float64)[idx], np.asarray(stf_values, np.float64)[idx]


def tiles(cfg, bounds, fx, fy, fz, stf_t, stf_values, source_xyz):
    """main.cpp:616-875 and writer.hh:105-310: every tile's contents as a dict
    (attributes under "@name", datasets under their paths, tensors in FP64
    before the file's precision is applied)."""
    nt = fx["n_steps"]
    time = np.arange(nt, dtype=np.float64) * cfg["output_dt_s"]
    st, sv = downsample_stf(stf_t, stf_values, nt)
    G = greens_tensor(fx, fy, fz)
    has_d = all(f["displacement"] is not None for f in (fx, fy, fz))
    D = displacement_tensor(fx, fy, fz) if has_d else None
    out = {}
    for (tx, ty), cells in tile_cells(cfg, bounds, fx).items():
        local, nodes = {}, []
        for ci in cells:
            for g in fx["cell_gll_node_index"][ci].tolist():
                if g >= 0 and g not in local:
                    local[g] = len(nodes)
                    nodes.append(g)
        cidx = np.array([[local[g] if g >= 0 else -1 for g in fx["cell_gll_node_index"][ci].tolist()]
                         for ci in cells], np.int32)
        nodes = np.asarray(nodes, np.int64)
        x0, x1, y0, y1 = tile_bounds(cfg, bounds, tx, ty)
        quant = "strain" + (",displacement,velocity,acceleration" if has_d else "")
        t = {
            "@version": "1.0.0", "@basis": "gll", "@tile_x_index": tx, "@tile_y_index": ty,
            "@x_min_m": x0, "@x_max_m": x1, "@y_min_m": y0, "@y_max_m": y1,
            "@z_min_m": bounds[4], "@z_max_m": bounds[5],
            "@record_depth_max_m": cfg["record_depth_max_m"],
            "@record_depth_actual_m": cfg["record_depth_actual_m"],
            "@source_xyz_m": np.asarray(source_xyz, np.float64), "@source_directions": "x,y,z",
            "@greens_quantities": quant, "@excludes_pml": 1,
            "/time/t": time, "/time@dt": cfg["solver_dt"], "/time@nsteps": nt,
            "/mesh/gll_node_ids": fx["gll_node_ids"][nodes],
            "/mesh/gll_node_coords": fx["gll_node_coords"][nodes],
            "/mesh/cell_gll_node_index": cidx,
            "/field/greens_tensor": G[:, nodes],
        }
        if len(st):
            t["/source/stf_t"] = st
            t["/source/stf_values"] = sv
        if has_d:
            t["/field/displacement_tensor"] = D[:, nodes]
            # velocity and acceleration are never accumulated (main.cpp:401)
            t["/field/velocity_tensor"] = np.zeros_like(D[:, nodes])
            t["/field/acceleration_tensor"] = np.zeros_like(D[:, nodes])
        out[f"tile_x{tx:03d}_y{ty:03d}.h5"] = t
    return out
