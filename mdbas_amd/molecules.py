"""Seeded synthetic boxes of small flexible molecules with every bonded term.

The water workloads have constraints and nonbonded terms only.  This
system exercises the rest of ``energy()`` (energy.c:103-205): bonds
(harmonic and Morse), Urey-Bradley terms, angles, dihedrals (cosine and
harmonic), impropers and the 1-4 pairs that the dihedrals create
(list.c:334-549), in the post-``setup`` layout that the hot path receives
(init.c:248-616: bonds to hydrogen became constraints under ``consth``).

One molecule, five atoms::

        E(H)
        |
    A - B - C - D

bonds A-B, B-C (harmonic), C-D (Morse), constraint B-E; angles A-B-C,
B-C-D, A-B-E, C-B-E; Urey-Bradley A...C; dihedrals A-B-C-D (cosine, n = 3)
and E-B-C-D (harmonic); improper B-A-C-E (harmonic).  The 1-4 pairs are
(A, D) and (E, D).  Charges sum to zero per molecule.  Force constants are
in internal units (kcal/mol x 418.4, global.h:41); eps/sig follow the
reference's stored forms (eps = sqrt(eps_kcal * 418.4), sig = Rmin/2 /
2^(1/6), io.c:1174-1175), and eps14 / sig14 are the separate 1-4 sets that
FORF's ``vdw`` lines may carry (io.c:1176-1185).
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field

import numpy as np

from .water import KCAL_TO_IU, RBOLTZ_IU, WaterBox, _init_constvel, normal, uniform, _rotations

SIXTH_ROOT_2 = 2.0 ** (1.0 / 6.0)
K = KCAL_TO_IU

#            A        B        C        D        E
MASS = [12.011, 12.011, 12.011, 15.999, 1.008]
CHARGE = [-0.27, 0.09, 0.05, -0.34, 0.47]
EPS_KCAL = [0.080, 0.055, 0.055, 0.152, 0.022]
RMIN2 = [2.060, 2.175, 2.175, 1.770, 1.320]
EPS14_KCAL = [0.010, 0.010, 0.010, 0.120, 0.022]
RMIN2_14 = [1.900, 1.900, 1.900, 1.400, 1.320]
LABELS = ["CA", "CB", "CC", "OD", "HE"]

R_AB, R_BC, R_CD, R_BE = 1.53, 1.53, 1.43, 1.09
TH_ABC, TH_BCD, TH_ABE, TH_CBE = 111.0, 109.5, 110.0, 110.0


class MdgBond(C.Structure):      # BOND, global.h:247-252
    _fields_ = [("type", C.c_int), ("a", C.c_int), ("b", C.c_int), ("k", C.c_double),
                ("r0", C.c_double), ("beta", C.c_double)]


class MdgAngle(C.Structure):     # ANGLE, global.h:254-259
    _fields_ = [("type", C.c_int), ("a", C.c_int), ("b", C.c_int), ("c", C.c_int),
                ("k", C.c_double), ("theta0", C.c_double)]


class MdgDihe(C.Structure):      # DIHE, global.h:261-267
    _fields_ = [("type", C.c_int), ("order", C.c_int), ("a", C.c_int), ("b", C.c_int),
                ("c", C.c_int), ("d", C.c_int), ("k", C.c_double), ("phi0", C.c_double),
                ("mult", C.c_double)]


def _place(a, b, c, r, theta, phi):
    """Natural extension reference frame: the atom at distance r from c,
    angle theta (b, c, new) and dihedral phi (a, b, c, new), degrees."""
    th, ph = math.radians(theta), math.radians(phi)
    bc = c - b
    bc /= np.linalg.norm(bc)
    n = np.cross(b - a, bc)
    n /= np.linalg.norm(n)
    m = np.cross(n, bc)
    d2 = np.array([-r * math.cos(th), r * math.sin(th) * math.cos(ph), r * math.sin(th) * math.sin(ph)])
    return c + d2[0] * bc + d2[1] * m + d2[2] * n


def body_frame() -> np.ndarray:
    """Rest geometry of one molecule (rows A, B, C, D, E), centred."""
    A = np.array([0.0, 0.0, 0.0])
    B = np.array([R_AB, 0.0, 0.0])
    Cc = B + R_BC * np.array([-math.cos(math.radians(TH_ABC)), math.sin(math.radians(TH_ABC)), 0.0])
    D = _place(A, B, Cc, R_CD, TH_BCD, 60.0)
    E = _place(Cc, A, B, R_BE, TH_ABE, 120.0)
    X = np.stack([A, B, Cc, D, E])
    return X - X.mean(0)


def _dihedral(p0, p1, p2, p3):
    """phi as internal.c:278-331 defines it (atan2 of the normals)."""
    dab, dbc, dcd = p0 - p1, p1 - p2, p2 - p3
    pb, pc = np.cross(dab, dbc), np.cross(dbc, dcd)
    cosp = pb @ pc / (np.linalg.norm(pb) * np.linalg.norm(pc))
    sinp = dbc @ np.cross(pc, pb) / (np.linalg.norm(pb) * np.linalg.norm(pc) * np.linalg.norm(dbc))
    return math.atan2(sinp, cosp)


@dataclass
class MolBox(WaterBox):
    """A water-box-like system with bonded topology (post-setup)."""
    bonds: list = field(default_factory=list)       # (type, a, b, k, r0, beta)
    ubs: list = field(default_factory=list)         # (type, a, b, k, r0, beta)
    angles: list = field(default_factory=list)      # (type, a, b, c, k, theta0)
    dihes: list = field(default_factory=list)       # (type, order, a, b, c, d, k, phi0, mult)
    imprs: list = field(default_factory=list)
    eps14: np.ndarray = None
    sig14: np.ndarray = None
    scal14: float = 0.5
    labels: list = field(default_factory=list)

    def copy(self) -> "MolBox":
        d = {k: (v.copy() if isinstance(v, (np.ndarray, list)) else v)
             for k, v in self.__dict__.items()}
        d["params"] = dict(self.params)
        return MolBox(**d)

    # ctypes arrays in the reference layouts
    def c_bonds(self):
        return _carr(MdgBond, self.bonds)

    def c_ubs(self):
        return _carr(MdgBond, self.ubs)

    def c_angles(self):
        return _carr(MdgAngle, self.angles)

    def c_dihes(self):
        return _carr(MdgDihe, self.dihes)

    def c_imprs(self):
        return _carr(MdgDihe, self.imprs)


def _carr(T, rows):
    a = (T * max(1, len(rows)))()
    for k, r in enumerate(rows):
        for (name, _), v in zip(T._fields_, r):
            setattr(a[k], name, v)
    return a


def make_mol_box(n_mol: int, L: float, seed: int, temp: float = 300.0, tol_shake: float = 1e-8,
                 max_cycle: int = 150, min_dist: float = 2.2, max_rounds: int = 300) -> MolBox:
    """n_mol molecules on a cubic lattice in an L box with uniform random
    orientations (re-drawn while any atom comes within min_dist of an atom
    of a neighbouring site), Maxwell-Boltzmann velocities at temp with the
    constraint projection, centre-of-mass removal and exact rescale of
    init.c:627-713."""
    n = 1
    while n ** 3 < n_mol:
        n += 1
    a = L / n
    nsite = n ** 3
    if n_mol < nsite:
        key = uniform(seed, 7, np.arange(nsite))
        sites = np.sort(np.argsort(key, kind="stable")[:n_mol])
    else:
        sites = np.arange(nsite)
    ix, iy, iz = sites // (n * n), (sites // n) % n, sites % n
    centre = np.stack([(ix + 0.5) * a - L / 2, (iy + 0.5) * a - L / 2, (iz + 0.5) * a - L / 2], 1)
    site_mol = -np.ones(nsite, dtype=np.int64)
    site_mol[sites] = np.arange(n_mol)
    body = body_frame()
    ids = np.arange(n_mol)
    attempt = np.zeros(n_mol, dtype=np.int64)
    pos = centre[:, None, :] + np.einsum("mij,aj->mai", _rotations(seed, ids, attempt), body)
    offsets = [(dx, dy, dz) for dx in (-1, 0, 1) for dy in (-1, 0, 1) for dz in (-1, 0, 1)
               if (dx, dy, dz) > (0, 0, 0)]
    for _ in range(max_rounds):
        bad = np.zeros(n_mol, dtype=bool)
        for dx, dy, dz in offsets:
            nb = site_mol[(((ix + dx) % n) * n + (iy + dy) % n) * n + (iz + dz) % n]
            ok = nb >= 0
            i_m, j_m = ids[ok], nb[ok]
            d = pos[j_m][:, None, :, :] - pos[i_m][:, :, None, :]
            d -= L * np.round(d / L)
            dmin = np.sqrt((d * d).sum(-1)).reshape(len(i_m), -1).min(1)
            viol = dmin < min_dist
            bad[np.maximum(i_m[viol], j_m[viol])] = True
        if not bad.any():
            break
        attempt[bad] += 1
        sel = ids[bad]
        pos[sel] = centre[sel][:, None, :] + np.einsum(
            "mij,aj->mai", _rotations(seed, sel, attempt[sel]), body)
    else:
        raise RuntimeError("could not place the molecules")

    # thermal distortion of the heavy atoms A, C, D (B and E keep the
    # constraint B-E at its length), so every bonded term is off its rest
    # value
    jit = np.stack([normal(seed, 31 + k, np.arange(n_mol * 5, dtype=np.uint64)) for k in range(3)], 1)
    jit = 0.06 * jit.reshape(n_mol, 5, 3)
    jit[:, [1, 4], :] = 0.0
    pos = pos + jit
    N = 5 * n_mol
    xyz = pos.reshape(N, 3)
    mass = np.tile(MASS, n_mol)
    rmass = 1.0 / mass
    q = np.tile(CHARGE, n_mol)
    eps = np.sqrt(np.tile(EPS_KCAL, n_mol) * K)
    sig = np.tile(RMIN2, n_mol) / SIXTH_ROOT_2
    eps14 = np.sqrt(np.tile(EPS14_KCAL, n_mol) * K)
    sig14 = np.tile(RMIN2_14, n_mol) / SIXTH_ROOT_2
    phi_abcd = _dihedral(*body[[0, 1, 2, 3]])
    phi_ebcd = _dihedral(*body[[4, 1, 2, 3]])
    phi_imp = _dihedral(*body[[1, 0, 2, 4]])
    bonds, ubs, angles, dihes, imprs = [], [], [], [], []
    ca, cb, rc2 = [], [], []
    for m in range(n_mol):
        A, B, Cc, D, E = (5 * m + k for k in range(5))
        bonds += [(0, A, B, 600.0 * K, R_AB, 0.0), (0, B, Cc, 600.0 * K, R_BC, 0.0),
                  (1, Cc, D, 90.0 * K, R_CD, 2.0)]
        ca.append(B)
        cb.append(E)
        rc2.append(R_BE * R_BE)
        angles += [(0, A, B, Cc, 100.0 * K, math.radians(TH_ABC)),
                   (0, B, Cc, D, 100.0 * K, math.radians(TH_BCD)),
                   (0, A, B, E, 70.0 * K, math.radians(TH_ABE)),
                   (0, Cc, B, E, 70.0 * K, math.radians(TH_CBE))]
        ubs.append((0, A, Cc, 20.0 * K, float(np.linalg.norm(body[2] - body[0])), 0.0))
        dihes += [(1, 0, A, B, Cc, D, 1.4 * K, 0.0, 3.0),
                  (2, 0, E, B, Cc, D, 2.0 * K, phi_ebcd + 0.3, 1.0)]
        imprs.append((2, 0, B, A, Cc, E, 20.0 * K, phi_imp, 1.0))
    del phi_abcd
    ca = np.array(ca, np.int32)
    cb = np.array(cb, np.int32)
    rc2 = np.array(rc2)
    nat = np.zeros(N, np.int32)
    np.add.at(nat, ca, 1)
    np.add.at(nat, cb, 1)
    frozen = np.zeros(N, np.int32)
    cnt = np.arange(N, dtype=np.uint64)
    vx = normal(seed, 21, cnt) * np.sqrt(rmass)
    vy = normal(seed, 22, cnt) * np.sqrt(rmass)
    vz = normal(seed, 23, cnt) * np.sqrt(rmass)
    x, y, z = xyz[:, 0].copy(), xyz[:, 1].copy(), xyz[:, 2].copy()
    vx, vy, vz = _init_constvel(x, y, z, vx, vy, vz, mass, nat, ca, cb, L, tol_shake, max_cycle)
    cm = np.array([(vx * mass).sum(), (vy * mass).sum(), (vz * mass).sum()]) / mass.sum()
    vx, vy, vz = vx - cm[0], vy - cm[1], vz - cm[2]
    ndf = 3 * N - 3 - len(ca)
    kin0 = 0.5 * temp * ndf * RBOLTZ_IU
    kin = 0.5 * (mass * (vx * vx + vy * vy + vz * vz)).sum()
    f = math.sqrt(kin0 / kin)
    return MolBox(n_waters=n_mol, L=float(L), x=x, y=y, z=z, vx=vx * f, vy=vy * f, vz=vz * f,
                  q=q, eps=eps, sig=sig, mass=mass, rmass=rmass, frozen=frozen, nAtConst=nat,
                  ca=ca, cb=cb, rc2=rc2, params={"temp": temp, "ndf": ndf, "seed": seed},
                  bonds=bonds, ubs=ubs, angles=angles, dihes=dihes, imprs=imprs,
                  eps14=eps14, sig14=sig14, scal14=0.5, labels=LABELS * n_mol)


# settings shared with the replica workload (9 A cutoff, switch from 7.5 A)
MOL_COMMON = dict(cutOff=9.0, cutOn=7.5, delr=2.0, alpha=0.35, nbsp=4, tolShake=1e-8,
                  maxCycle=150, update=20, linkRatio=1, temp=300.0, timeStep=0.001)

MOL_CONFIGS = {
    # 216 molecules (1,080 atoms) in a 30 A box: O(N^2) list (2 cells per edge)
    "flex216": dict(MOL_COMMON, name="flex_216", n_mol=216, L=30.0, seed=41, grid=32),
    # 729 molecules (3,645 atoms) in a 45 A box: link cells (4 per edge)
    "flex729": dict(MOL_COMMON, name="flex_729", n_mol=729, L=45.0, seed=42, grid=48),
}


def build(cfg: dict) -> MolBox:
    box = make_mol_box(cfg["n_mol"], cfg["L"], cfg["seed"], temp=cfg["temp"],
                       tol_shake=cfg["tolShake"], max_cycle=cfg["maxCycle"])
    box.params.update(cfg)
    return box
