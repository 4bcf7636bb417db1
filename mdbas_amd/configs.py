"""The five synthetic water workloads, with their SIMU-deck settings.

Each entry gives what a reference SIMU deck would set (cutoff, cuton,
delr, ewald alpha / nbsp / grid, timestep, list period, link ratio,
temperature) for a seeded rigid TIP3P box built by :mod:`mdbas_amd.water`.
Grid sizes are the FFT mesh the reference would use: ``ewald recvec m``
gives a ``2*m`` mesh (spme.c:54-73), so 32^3 / 64^3 / 128^3 come from
recvec 16 / 32 / 64.  The replica workload asks for ~1 A spacing rounded
up to a multiple of 4, which the reference's power-of-two rounding cannot
express; those meshes are used as given.
"""
from __future__ import annotations

import math

from . import water

COMMON = dict(cutOff=12.0, cutOn=10.0, delr=2.0, alpha=0.32, nbsp=4,
              tolShake=1e-8, maxCycle=150, update=20, linkRatio=1, temp=300.0)

CONFIGS = {
    "c0": dict(COMMON, name="tip3p_1k_fp64", n_waters=1000, L=31.04, seed=0, grid=32,
               timeStep=0.001, nsteps=1000, precision="fp64"),
    "c1": dict(COMMON, name="tip3p_10k_mixed", n_waters=10000, L=67.0, seed=1, grid=64,
               timeStep=0.002, nsteps=5000, precision="mixed"),
    "c2": dict(COMMON, name="tip3p_64k_mixed", n_waters=64000, L=124.1, seed=2, grid=128,
               timeStep=0.002, nsteps=2000, precision="mixed"),
    "c3": dict(COMMON, name="tip3p_1k_adjoint", n_waters=1000, L=31.04, seed=0, grid=32,
               timeStep=0.001, nsteps=200, checkpoint=20, precision="fp64"),
}


def replica_configs(n_replicas: int = 256, seed: int = 3):
    """The uneven replica batch: W ~ U{216..512}, L = (W/0.0334)^(1/3),
    T spaced geometrically over 280-400 K, 9 A cutoff / 7.5 A switch,
    alpha 0.35, mesh ~1 A rounded up to a multiple of 4."""
    u = water.uniform(seed, 31, range(n_replicas))
    out = []
    for k in range(n_replicas):
        w = 216 + int(u[k] * 297)
        L = (w / 0.0334) ** (1.0 / 3.0)
        m = 4 * int(math.ceil(L / 4.0))
        t = 280.0 * (400.0 / 280.0) ** (k / max(1, n_replicas - 1))
        out.append(dict(COMMON, name=f"replica{k}", n_waters=w, L=L, seed=seed * 100000 + k,
                        grid=m, cutOff=9.0, cutOn=7.5, alpha=0.35, timeStep=0.002,
                        nsteps=1000, temp=t, precision="mixed"))
    return out


def build(cfg: dict) -> water.WaterBox:
    box = water.make_water_box(cfg["n_waters"], cfg["L"], cfg["seed"], temp=cfg["temp"],
                               tol_shake=cfg["tolShake"], max_cycle=cfg["maxCycle"])
    box.params.update(cfg)
    return box
