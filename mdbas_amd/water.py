"""Seeded synthetic rigid TIP3P water boxes in the reference's array layout.

The reference ships no input decks and has no water builder, so the
workloads are generated here.  Layout follows what ``init_system``
(init.c:133-157) hands the hot path: SoA float64 arrays of length N for
coordinates, velocities, charges, LJ parameters and masses, plus the
constraint list that ``setup`` (init.c:322-343) derives from FORF bonds
under ``consth``: per molecule O-H1, O-H2, H1-H2, with rc2 = r0^2.

Parameter conventions (io.c:1174-1175, tools/prepare.c:1999-2000):
``eps[i] = sqrt(eps_kcal * 418.4)`` so the pair uses ``eps_i * eps_j``, and
``sig[i] = (Rmin/2) / 2^(1/6)`` so the pair uses ``sig_i + sig_j``.

Random numbers come from a counter-based SplitMix64 hash, so a box depends
only on (seed, stream, counter), not on numpy's generator versions.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

KCAL_TO_IU = 418.4                # global.h:41 (10 J/mol per kcal/mol)
CHARGE_CONST = 332.0716 * 418.4   # chgcharmm * kcaltoiu, init.c:102-103
RBOLTZ_IU = 0.83144621            # global.h:47

R_OH = 0.9572
THETA_HOH = 104.52
Q_O, Q_H = -0.834, 0.417
M_O, M_H = 15.9994, 1.008
EPS_O_KCAL = 0.1521
RMIN2_O = 1.7682
R_HH = 2.0 * R_OH * math.sin(math.radians(THETA_HOH) / 2.0)

_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _splitmix64(z: np.ndarray) -> np.ndarray:
    """SplitMix64 finaliser (Steele, Lea, Flood 2014), vectorised, wrapping."""
    with np.errstate(over="ignore"):
        z = (z + np.uint64(0x9E3779B97F4A7C15)) & _M64
        z = ((z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)) & _M64
        z = ((z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)) & _M64
        return z ^ (z >> np.uint64(31))


def uniform(seed: int, stream: int, counter) -> np.ndarray:
    """Uniform doubles in (0, 1) for counters (array), fixed seed/stream."""
    c = np.asarray(counter, dtype=np.uint64)
    with np.errstate(over="ignore"):
        key = _splitmix64(np.uint64(seed) * np.uint64(0x100000001B3) ^ np.uint64(stream))
        u = _splitmix64(key ^ (c * np.uint64(0xD6E8FEB86659FD93)))
    return ((u >> np.uint64(11)).astype(np.float64) + 0.5) * (1.0 / 9007199254740992.0)


def normal(seed: int, stream: int, counter) -> np.ndarray:
    """Standard normals by Box-Muller from two uniforms per counter."""
    c = np.asarray(counter, dtype=np.uint64)
    u1 = uniform(seed, stream, c * np.uint64(2))
    u2 = uniform(seed, stream, c * np.uint64(2) + np.uint64(1))
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * math.pi * u2)


def _rotations(seed: int, ids: np.ndarray, attempt: np.ndarray) -> np.ndarray:
    """Uniform random rotation matrices from normalised 4-Gaussians."""
    base = (ids.astype(np.uint64) << np.uint64(8)) | attempt.astype(np.uint64)
    qv = np.stack([normal(seed, 11, base * np.uint64(4) + np.uint64(k)) for k in range(4)], 1)
    qv /= np.linalg.norm(qv, axis=1, keepdims=True)
    w, x, y, z = qv.T
    R = np.empty((len(ids), 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z)
    R[:, 0, 1] = 2 * (x * y - z * w)
    R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w)
    R[:, 1, 1] = 1 - 2 * (x * x + z * z)
    R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w)
    R[:, 2, 1] = 2 * (y * z + x * w)
    R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def _body_frame() -> np.ndarray:
    h = math.radians(THETA_HOH) / 2.0
    return np.array([[0.0, 0.0, 0.0],
                     [R_OH * math.sin(h), 0.0, R_OH * math.cos(h)],
                     [-R_OH * math.sin(h), 0.0, R_OH * math.cos(h)]])


@dataclass
class WaterBox:
    """One synthetic system in the reference's SoA layout."""
    n_waters: int
    L: float
    x: np.ndarray
    y: np.ndarray
    z: np.ndarray
    vx: np.ndarray
    vy: np.ndarray
    vz: np.ndarray
    q: np.ndarray
    eps: np.ndarray
    sig: np.ndarray
    mass: np.ndarray
    rmass: np.ndarray
    frozen: np.ndarray
    nAtConst: np.ndarray
    ca: np.ndarray
    cb: np.ndarray
    rc2: np.ndarray
    params: dict = field(default_factory=dict)

    @property
    def n(self) -> int:
        return len(self.x)

    @property
    def box9(self) -> np.ndarray:
        return np.array([self.L, 0, 0, 0, self.L, 0, 0, 0, self.L], dtype=np.float64)

    def copy(self) -> "WaterBox":
        d = {k: (v.copy() if isinstance(v, np.ndarray) else v) for k, v in self.__dict__.items()}
        d["params"] = dict(self.params)
        return WaterBox(**d)


def _lattice_n(w: int) -> int:
    n = 1
    while n * n * n < w:
        n += 1
    return n


def build_positions(n_waters: int, L: float, seed: int, min_hh: float = 1.6,
                    max_rounds: int = 200) -> np.ndarray:
    """(W,3,3) O,H1,H2 coordinates on a cubic lattice, random orientations,
    re-drawing any molecule whose H is closer than ``min_hh`` to an H of a
    molecule on a neighbouring site."""
    n = _lattice_n(n_waters)
    a = L / n
    nsite = n ** 3
    if n_waters < nsite:
        key = uniform(seed, 7, np.arange(nsite))
        sites = np.sort(np.argsort(key, kind="stable")[:n_waters])
    else:
        sites = np.arange(nsite)
    ix, iy, iz = sites // (n * n), (sites // n) % n, sites % n
    centre = np.stack([(ix + 0.5) * a - L / 2, (iy + 0.5) * a - L / 2, (iz + 0.5) * a - L / 2], 1)
    site_mol = -np.ones(nsite, dtype=np.int64)
    site_mol[sites] = np.arange(n_waters)
    body = _body_frame()
    ids = np.arange(n_waters)
    attempt = np.zeros(n_waters, dtype=np.int64)
    pos = centre[:, None, :] + np.einsum("mij,aj->mai", _rotations(seed, ids, attempt), body)
    offsets = [(dx, dy, dz) for dx in (-1, 0, 1) for dy in (-1, 0, 1) for dz in (-1, 0, 1)
               if (dx, dy, dz) > (0, 0, 0)]
    for _ in range(max_rounds):
        bad = np.zeros(n_waters, dtype=bool)
        for dx, dy, dz in offsets:
            nb = site_mol[(((ix + dx) % n) * n + (iy + dy) % n) * n + (iz + dz) % n]
            ok = nb >= 0
            i_m, j_m = ids[ok], nb[ok]
            hi = pos[i_m][:, 1:3, None, :]
            hj = pos[j_m][:, None, 1:3, :]
            d = hj - hi
            d -= L * np.round(d / L)
            dmin = np.sqrt((d * d).sum(-1)).reshape(len(i_m), -1).min(1)
            viol = dmin < min_hh
            bad[np.maximum(i_m[viol], j_m[viol])] = True
        if not bad.any():
            break
        attempt[bad] += 1
        sel = ids[bad]
        pos[sel] = centre[sel][:, None, :] + np.einsum(
            "mij,aj->mai", _rotations(seed, sel, attempt[sel]), body)
    return pos


def _init_constvel(x, y, z, vx, vy, vz, mass, nat, ca, cb, L, tol, max_cycle):
    """Constraint projection of initial velocities, init.c:724-833
    (Jacobi sweep over all constraints until max |dv| < tol)."""
    d = np.stack([x[cb] - x[ca], y[cb] - y[ca], z[cb] - z[ca]], 1)
    d -= L * np.trunc(d / L + np.where(d >= 0, 0.5, -0.5))
    d /= np.sqrt((d * d).sum(1))[:, None]
    v = np.stack([vx, vy, vz], 1)
    n = len(x)
    for _ in range(max_cycle):
        dv = (d * (v[cb] - v[ca])).sum(1)
        if np.abs(dv).max() < tol:
            break
        msum = mass[ca] + mass[cb]
        w1 = mass[cb] * dv / msum
        w2 = mass[ca] * dv / msum
        acc = np.zeros((n, 3))
        np.add.at(acc, ca, w1[:, None] * d)
        np.add.at(acc, cb, -w2[:, None] * d)
        # each atom receives acc/nAtConst once per constraint it is in
        v += acc
    else:
        raise RuntimeError("initial-velocity constraint projection did not converge (code 310)")
    return v[:, 0].copy(), v[:, 1].copy(), v[:, 2].copy()


def make_water_box(n_waters: int, L: float, seed: int, temp: float = 300.0,
                   tol_shake: float = 1e-8, max_cycle: int = 150) -> WaterBox:
    """Rigid TIP3P box: positions, Maxwell-Boltzmann velocities at ``temp``
    following init.c:627-713 (Gaussian * sqrt(1/m), constraint projection,
    centre-of-mass removal, exact rescale to T * N_df * R / 2)."""
    pos = build_positions(n_waters, L, seed)
    N = 3 * n_waters
    xyz = pos.reshape(N, 3)
    q = np.tile([Q_O, Q_H, Q_H], n_waters)
    mass = np.tile([M_O, M_H, M_H], n_waters)
    rmass = 1.0 / mass
    eps = np.tile([math.sqrt(EPS_O_KCAL * KCAL_TO_IU), 0.0, 0.0], n_waters)
    sig = np.tile([RMIN2_O / 2.0 ** (1.0 / 6.0), 0.0, 0.0], n_waters)
    o = 3 * np.arange(n_waters)
    ca = np.stack([o, o, o + 1], 1).reshape(-1).astype(np.int32)
    cb = np.stack([o + 1, o + 2, o + 2], 1).reshape(-1).astype(np.int32)
    rc2 = np.tile([R_OH * R_OH, R_OH * R_OH, R_HH * R_HH], n_waters)
    nat = np.full(N, 2, dtype=np.int32)
    frozen = np.zeros(N, dtype=np.int32)

    cnt = np.arange(N, dtype=np.uint64)
    vx = normal(seed, 21, cnt) * np.sqrt(rmass)
    vy = normal(seed, 22, cnt) * np.sqrt(rmass)
    vz = normal(seed, 23, cnt) * np.sqrt(rmass)
    x, y, z = xyz[:, 0].copy(), xyz[:, 1].copy(), xyz[:, 2].copy()
    vx, vy, vz = _init_constvel(x, y, z, vx, vy, vz, mass, nat, ca, cb, L, tol_shake, max_cycle)
    cm = np.array([(vx * mass).sum(), (vy * mass).sum(), (vz * mass).sum()]) / mass.sum()
    vx, vy, vz = vx - cm[0], vy - cm[1], vz - cm[2]
    ndf = 3 * N - 3 - len(ca)                      # utils.c:569-582
    kin0 = 0.5 * temp * ndf * RBOLTZ_IU            # utils.c:551-561
    kin = 0.5 * (mass * (vx * vx + vy * vy + vz * vz)).sum()
    f = math.sqrt(kin0 / kin)
    return WaterBox(n_waters, float(L), x, y, z, vx * f, vy * f, vz * f, q, eps, sig,
                    mass, rmass, frozen, nat, ca, cb, rc2,
                    params={"temp": temp, "ndf": ndf, "seed": seed})
