"""The benchmark workloads as box specs (seeded, reproducible).

  halfspace_spec  20 x 20 x 10 hexes of 1 km (4,000 elements), homogeneous
                  halfspace Vp 6000 / Vs 3464 / rho 2700, order 4, C-PML 3 on the
                  sides and bottom, CFL 0.5, 10 s, output dt 0.05 s, strain
                  recorded to 2 km (the reference's CPU-runnable case);
  layered_spec    100 x 100 x 50 hexes of 0.6 km (500,000 elements), three layers
                  (0-3 km 5000/2900/2500, 3-20 km 6200/3600/2800, halfspace
                  8000/4600/3300) with a 5 % Gaussian random velocity field of
                  2 km correlation length, C-PML 5, 30 s, recorded to 2 km;
  stations        seeded surface stations for the batched x/y/z fields;
  uneven_specs    seeded boxes of uneven size for batches of independent meshes.
"""
from __future__ import annotations

import numpy as np

from . import make_spec

HALFSPACE = [(0.0, 6000.0, 3464.0, 2700.0)]
LAYERED = [(3000.0, 5000.0, 2900.0, 2500.0), (20000.0, 6200.0, 3600.0, 2800.0), (0.0, 8000.0, 4600.0, 3300.0)]


def halfspace_spec(nx=20, ny=20, nz=10, h=1000.0, pml=3, duration=10.0):
    return make_spec(nx=nx, ny=ny, nz=nz, lx=h * nx, ly=h * ny, lz=h * nz, order=4, layers=HALFSPACE,
                     pml_xmin=pml, pml_xmax=pml, pml_ymin=pml, pml_ymax=pml, pml_zmax=pml, f0_pml=1.0,
                     cfl_safety=0.5, output_dt=0.05, duration=duration, record_depth_max=2000.0)


def layered_spec(nx=100, ny=100, nz=50, h=600.0, duration=30.0, n_sls=0, pml=5):
    """n_sls > 0 adds shear attenuation (Q_mu 60 / 120 / 250 by layer over 0.05-2 Hz)."""
    return make_spec(
        n_sls=n_sls, q_mu=[60.0, 120.0, 250.0], att_fmin=0.05, att_fmax=2.0,
        nx=nx, ny=ny, nz=nz, lx=h * nx, ly=h * ny, lz=h * nz, layers=LAYERED,
        rf_amp=0.05, rf_corr_len=2000.0, rf_modes=256, rf_seed=2,
        pml_xmin=pml, pml_xmax=pml, pml_ymin=pml, pml_ymax=pml, pml_zmax=pml, f0_pml=1.5,
        cfl_safety=0.5, output_dt=0.05, duration=duration, record_depth_max=2000.0,
        want_coords=0, want_beta=0)


def stations(mesh, n=8, seed=3, half=20000.0):
    """n seeded surface points, uniform in the central 2 half x 2 half square."""
    rng = np.random.default_rng(seed)
    cx = 0.5 * (mesh.bounds[0] + mesh.bounds[1])
    cy = 0.5 * (mesh.bounds[2] + mesh.bounds[3])
    return [(cx + rng.uniform(-half, half), cy + rng.uniform(-half, half)) for _ in range(n)]


def uneven_specs(n=16, seed=4, h=600.0, lo=8e3, hi=2e5, pml=(3, 6)):
    """n independent boxes (seeded): n_target log-uniform in [lo, hi] elements,
    nx = ny = round((2n)^(1/3)), nz = round(nx/2), 0.6 km hexes; 2-4 layers with
    Vs ~ U(1500, 4500) increasing with depth, Vp = 1.73 Vs, rho ~ U(2300, 3300);
    C-PML U{pml[0]..pml[1]}; run length U(5, 30) s; Ricker f0 = 1.5 Hz."""
    rng = np.random.default_rng(seed)
    specs = []
    for _ in range(n):
        nt = float(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        nx = int(round((2 * nt) ** (1 / 3)))
        nz = max(4, int(round(nx / 2)))
        nl = int(rng.integers(2, 5))
        vs = np.sort(rng.uniform(1500.0, 4500.0, nl))
        rho = rng.uniform(2300.0, 3300.0, nl)
        lz = h * nz
        tops = np.sort(rng.uniform(0.1, 0.7, nl - 1)) * lz
        layers = [(float(tops[i]), 1.73 * vs[i], vs[i], rho[i]) for i in range(nl - 1)]
        layers.append((0.0, 1.73 * vs[-1], vs[-1], rho[-1]))
        npml = int(rng.integers(pml[0], pml[1] + 1))
        dur = float(rng.uniform(5.0, 30.0))
        specs.append(make_spec(nx=nx, ny=nx, nz=nz, lx=h * nx, ly=h * nx, lz=lz, layers=layers,
                               pml_xmin=npml, pml_xmax=npml, pml_ymin=npml, pml_ymax=npml, pml_zmax=npml,
                               f0_pml=1.5, cfl_safety=0.5, output_dt=0.05, duration=dur,
                               record_depth_max=2000.0, want_coords=0, want_beta=0))
    return specs
