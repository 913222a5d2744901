"""Small, seeded box problems shared by the CPU (oracle) and GPU parity tests."""
from __future__ import annotations

import numpy as np

import pygfx

HALFSPACE = [(0.0, 6000.0, 3464.0, 2700.0)]
# three-layer model (layer bottoms scaled to the small boxes)
LAYERED = [(1800.0, 5000.0, 2900.0, 2500.0), (3600.0, 6200.0, 3600.0, 2800.0),
           (0.0, 8000.0, 4600.0, 3300.0)]


def small_halfspace(nx=8, ny=8, nz=6, pml=2, duration=2.5, **kw):
    """Homogeneous halfspace of 1 km hexes, C-PML on five sides, order 4."""
    spec = dict(nx=nx, ny=ny, nz=nz, lx=1000.0 * nx, ly=1000.0 * ny, lz=1000.0 * nz,
                layers=HALFSPACE, pml_xmin=pml, pml_xmax=pml, pml_ymin=pml, pml_ymax=pml,
                pml_zmax=pml, f0_pml=1.0, duration=duration, record_depth_max=2000.0)
    spec.update(kw)
    return pygfx.build_box(**spec)


def small_layered(nx=8, ny=7, nz=6, pml=2, duration=2.0, jitter=0.12, **kw):
    """Three layers + 5 % random velocity field on a jittered (non-affine) mesh."""
    spec = dict(nx=nx, ny=ny, nz=nz, lx=900.0 * nx, ly=900.0 * ny, lz=900.0 * nz,
                layers=LAYERED, rf_amp=0.05, rf_corr_len=2000.0, rf_modes=64, rf_seed=11,
                jitter=jitter, jitter_seed=5, pml_xmin=pml, pml_xmax=pml, pml_ymin=pml,
                pml_ymax=pml, pml_zmax=pml, f0_pml=1.5, duration=duration,
                record_depth_max=1800.0)
    spec.update(kw)
    return pygfx.build_box(**spec)


def surface_centre_source(mesh, f0=1.0, t0=1.5, amp=1e20):
    x = 0.5 * (mesh.bounds[0] + mesh.bounds[1])
    y = 0.5 * (mesh.bounds[2] + mesh.bounds[3])
    cells, w = mesh.locate_source(x, y, 0.0)
    stf = pygfx.ricker(f0, t0, amp, mesh.dt, mesh.nsteps)
    return cells, w, stf


def buried_source(mesh, frac=(0.47, 0.53, 0.0), depth=None, f0=1.5, t0=1.0, amp=1e20):
    x = mesh.bounds[0] + frac[0] * (mesh.bounds[1] - mesh.bounds[0])
    y = mesh.bounds[2] + frac[1] * (mesh.bounds[3] - mesh.bounds[2])
    z = depth if depth is not None else 0.3 * (mesh.bounds[5] - mesh.bounds[4])
    cells, w = mesh.locate_source(x, y, z)
    stf = pygfx.ricker(f0, t0, amp, mesh.dt, mesh.nsteps)
    return cells, w, stf


def rigid_outer_nodes(mesh, faces=("xmin", "xmax", "ymin", "ymax", "zmax")):
    """Nodes on the named box faces (those that carry a C-PML layer), found
    from the coordinates."""
    xyz = mesh.node_xyz.reshape(-1, 3)
    b = mesh.bounds
    tol = 1e-6 * max(b[1] - b[0], b[3] - b[2], b[5] - b[4])
    where = {"xmin": (0, b[0]), "xmax": (0, b[1]), "ymin": (1, b[2]), "ymax": (1, b[3]), "zmax": (2, b[5])}
    on = np.zeros(len(xyz), bool)
    for name in faces:
        ax, v = where[name]
        on |= np.abs(xyz[:, ax] - v) <= tol
    return on


def cpml_complete_arrays(mesh):
    """Oracle inputs of the complete C-PML: the mesh arrays with zero mass at
    the element points on the rigid outer boundary (the oracle's massless-node
    rule then holds them, as the solver's rigid nodes)."""
    arr = dict(mesh.arrays)
    mass = np.array(mesh.mass, copy=True)
    mass[rigid_outer_nodes(mesh)[mesh.ibool]] = 0.0
    arr["mass"] = mass
    return arr


def rel_l2(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / nb) if nb > 0 else float(np.linalg.norm(a))


def per_trace_rel_l2(s, ref, floor=1e-3):
    """Per (cell, point, component) trace rel L2 over time; traces whose norm is
    below floor * max trace norm are skipped (symmetry zeros)."""
    s = np.asarray(s).reshape(ref.shape[0], -1)
    r = np.asarray(ref).reshape(ref.shape[0], -1)
    nr = np.linalg.norm(r, axis=0)
    keep = nr >= floor * nr.max()
    return np.linalg.norm(s[:, keep] - r[:, keep], axis=0) / nr[keep]


def amplified_cpml(mesh, damp=0.1, spring=0.01, memory=20.0):
    """Copy of `mesh` whose C-PML inputs make the reference's C-PML step dominate the
    PML, recording the surface-layer PML cells.

    The reference adds w J / rho (A1 v + A2 u + A3 m_x + A4 m_y + A5 m_z) to the residual
    (forward/share/src/pml.cpp:197-212) and divides by the mass rho J w, so the term
    reaches the acceleration as A / rho^2.  With the generated coefficients and rho = 2700
    it moves the records by ~1e-8, below any parity gate.  Here A' = rho^2 (...):
      A1' = -damp rho^2 A1          velocity damping over the generated ramp A1;
      A2' = -spring rho^2 |A2|      a spring;
      A3+d' = -memory rho^2 b_d A1 / max A1   memory coupling (m_d ~ u / b_d for slowly
                                    varying u, so a lagged spring of strength `memory`);
    and coef_alpha from poles b_d (1/s) that are linear in the node position, different
    per direction, through the reference's convolution coefficients
    (preprocess/pml_cpml.py:261-304): (exp(-b dt), (1 - exp(-b dt/2)) / b, c1 exp(-b dt/2)).
    Being functions of the node position, the poles agree between the elements at a node,
    so the nodal C-PML form applies.  The oracle runs every part of the C-PML step on
    these inputs; on the 8x8x6 halfspace over 4 s each part moves the recorded strain by
    1e-3 to 1 (tests/test_cpml_sensitivity.py), and the wave field decays."""
    a = mesh.arrays
    rho = a["density"]
    ab0 = a["coef_abar"].reshape(-1, 5)
    xyz = a["node_xyz"].reshape(-1, 3)[a["ibool"]]
    lx, ly, lz = mesh.bounds[1], mesh.bounds[3], mesh.bounds[5]
    b = np.stack([6.0 + 10.0 * xyz[:, 0] / lx, 12.0 + 8.0 * xyz[:, 1] / ly,
                  20.0 + 15.0 * xyz[:, 2] / lz], axis=1)
    t = np.exp(-0.5 * b * mesh.dt)
    c1 = (1.0 - t) / b
    alpha = np.stack([t * t, c1, c1 * t], axis=2).reshape(-1)
    r2 = rho * rho
    prof = ab0[:, 0] / ab0[:, 0].max()
    ab = np.zeros_like(ab0)
    ab[:, 0] = -damp * r2 * ab0[:, 0]
    ab[:, 1] = -spring * r2 * np.abs(ab0[:, 1])
    ab[:, 2:] = -memory * (r2 * prof)[:, None] * b
    top = mesh.bounds[4] + (mesh.bounds[5] - mesh.bounds[4]) / mesh.nz
    rec = np.nonzero((mesh.centroids()[:, 2] < top) & (mesh.pml_region != 0))[0]
    return mesh.derive(rec_cell=rec, coef_alpha=alpha, coef_abar=ab.reshape(-1))


def pml_nodes(mesh):
    """Global ids of the nodes of the C-PML elements."""
    return np.unique(mesh.ibool.reshape(mesh.n_elem, -1)[mesh.pml_region != 0])
