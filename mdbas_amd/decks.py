"""Reference-format input decks for the synthetic boxes.

Writes a box built by :mod:`mdbas_amd.water` or :mod:`mdbas_amd.molecules`
as the SIMU / CONF / FORF files a reference run reads (io.c:29-1199), and
its velocities as a REST file (io.c:785-870, the binary layout
``mdg_write_rest`` writes), so the same system can be read back through
``mdg_read_*`` (or the reference's own readers) and run by the C host
driver ``mdbas_gpu``.

Numbers are written so that the reference's parsing gives back the box's
values bit for bit: coordinates go through the binary REST file, and every
text field is the shortest decimal whose parse (``atof``, then the
reader's own scaling: ``x 418.4`` for energies, ``sqrt(eps * 418.4)`` for
LJ well depths, ``r0 * r0`` for constraint lengths, degrees to radians)
reproduces the stored double.

SIMU layout (io.c:113-711)::

    charmm / timestep / nsteps / cutoff / cuton / delr / elec spme /
    ewald prec|alpha|nbsp|recvec / vdw switch / list / link / consth /
    integrator velocity / ensemble nve / temperature / seed / print /
    prop / traj / resconf / [restart] / pbc 1 + 3 lines / end
"""
from __future__ import annotations

import math
import os
import struct

import numpy as np

KCAL = 418.4
PI = 3.141592653589793
ATOM_BYTES = 28      # char label[5], segn[5], resn[5]; 1 pad byte; int type, resi, ires
ENERGY_DOUBLES = 19  # ENERGY, global.h:282-290


def _exact(f, target: float, guess: float) -> str:
    """The text of a double v near guess with f(v) == target exactly (a
    few ulps searched on each side), else repr(guess)."""
    v = guess
    cands = [v]
    lo = hi = v
    for _ in range(8):
        lo = math.nextafter(lo, -math.inf)
        hi = math.nextafter(hi, math.inf)
        cands += [lo, hi]
    for c in cands:
        if f(float(repr(c))) == target:
            return repr(c)
    return repr(guess)


def _kcal(k_iu: float) -> str:
    return _exact(lambda v: v * KCAL, k_iu, k_iu / KCAL)


def _eps(eps_stored: float) -> str:
    if eps_stored == 0.0:
        return "0.0"
    return _exact(lambda v: math.sqrt(v * KCAL), eps_stored, eps_stored * eps_stored / KCAL)


def _deg(rad: float) -> str:
    return _exact(lambda v: v * PI / 180.0, rad, rad * 180.0 / PI)


def _r0(rc2: float) -> str:
    return _exact(lambda v: v * v, rc2, math.sqrt(rc2))


def labels_of(box) -> list:
    """CONF atom labels: hydrogens (first letter H, init.c:324) as the
    reference identifies them."""
    if hasattr(box, "labels") and box.labels:
        return list(box.labels)
    return ["OH2", "H1", "H2"] * box.n_waters


def write_simu(path: str, box, cfg: dict, nsteps: int, *, restart: bool = True,
               prop: int = 1, traj: int = 0, resconf: int = 0, printout: int = 1,
               seed: int | None = None, nb14: float | None = None, elec: str = "spme",
               mesh: int | None = None, ensemble: str = "nve", taut: float | None = None,
               taup: float | None = None, pressure: float | None = None):
    """The SIMU deck for box with the workload settings of cfg; ensemble is
    the SIMU value ("nve", "nvt beren", "npt hoover", ...), with the bath's
    taut / taup (ps) and the target pressure (bar) when given."""
    m = mesh if mesh is not None else cfg["grid"]
    rec = m // 2          # recvec r gives a 2r mesh (spme.c:54-73) for power-of-two r
    lines = ["charmm",
             f"timestep {float(cfg['timeStep'])!r}",
             f"nsteps {nsteps}",
             f"cutoff {float(cfg['cutOff'])!r}",
             f"cuton {float(cfg['cutOn'])!r}",
             f"delr {float(cfg['delr'])!r}",
             f"elec {elec}",
             "ewald prec 1e-6",
             f"ewald alpha {float(cfg['alpha'])!r}",
             f"ewald nbsp {cfg['nbsp']}",
             f"ewald recvec {rec} {rec} {rec}",
             "vdw switch",
             f"list {cfg['update']}",
             f"link {cfg['linkRatio']}",
             "consth",
             f"shake tol {float(cfg['tolShake'])!r}",
             f"shake maxcycle {cfg['maxCycle']}",
             "integrator velocity",
             f"ensemble {ensemble}",
             f"temperature {float(cfg.get('temp', 300.0))!r}",
             f"print {printout}"]
    for key, val in (("taut", taut), ("taup", taup), ("pressure", pressure)):
        if val is not None:
            lines.append(f"{key} {float(val)!r}")
    if nb14 is not None:
        lines.append(f"nb14 {float(nb14)!r}")
    if seed is not None:
        lines.append(f"seed {seed}")
    if prop:
        lines.append(f"prop {prop}")
    if traj:
        lines.append(f"traj {traj}")
    if resconf:
        lines.append(f"resconf {resconf}")
    if restart:
        lines.append("restart")
    L = repr(float(box.L))
    lines += ["pbc 1", f"{L} 0.0 0.0", f"0.0 {L} 0.0", f"0.0 0.0 {L}", "end"]
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")


def write_conf(path: str, box):
    """CONF (io.c:744-779): '*' title lines, the count, then per atom
    serial, residue index, residue name, label, x, y, z, segment, residue
    id, weight."""
    lab = labels_of(box)
    per = box.n // box.n_waters
    with open(path, "w") as f:
        f.write("* synthetic box\n*\n")
        f.write(f"{box.n:5d}\n")
        for i in range(box.n):
            res = i // per + 1
            f.write(f"{i + 1:5d}{res:5d} MOL  {lab[i]:<4s} {float(box.x[i])!r} {float(box.y[i])!r} "
                    f"{float(box.z[i])!r} SYS  {res:<4d} 0.00000\n")


def write_forf(path: str, box):
    """FORF (io.c:890-1195).  Water: the three O-H, O-H and H-H bonds per
    molecule, which consth turns into constraints (init.c:314-343).
    Molecule boxes: their bonds, with each constraint written as a bond
    (its partner is a hydrogen, so consth turns it back into that
    constraint, in the same order), Urey-Bradley terms, angles, dihedrals,
    impropers."""
    n = box.n
    per = n // box.n_waters
    out = ["# synthetic box", f"atoms {n}"]
    types = {}
    lab = labels_of(box)
    for i in range(n):
        t = types.setdefault(lab[i], len(types) + 1)
        out.append(f"{i + 1} {lab[i]} {t} {float(box.q[i])!r} {float(box.mass[i])!r} "
                   f"{int(box.frozen[i])}")
    bonds = []
    if hasattr(box, "bonds"):
        # per molecule: its bonds, then its constraints as bonds
        cons_of = {}
        for a, b, r in zip(box.ca, box.cb, box.rc2):
            cons_of.setdefault(int(a) // per, []).append((int(a), int(b), float(r)))
        by_mol = {}
        for t in box.bonds:
            by_mol.setdefault(t[1] // per, []).append(t)
        for m in range(box.n_waters):
            for (ty, a, b, k, r0, beta) in by_mol.get(m, []):
                bonds.append(f"{a + 1} {b + 1} {ty} {_kcal(float(k))} {float(r0)!r} "
                             f"{float(beta)!r}")
            for (a, b, r) in cons_of.get(m, []):
                bonds.append(f"{a + 1} {b + 1} 0 0.0 {_r0(r)} 0.0")
    else:
        for a, b, r in zip(box.ca, box.cb, box.rc2):
            bonds.append(f"{int(a) + 1} {int(b) + 1} 0 450.0 {_r0(float(r))} 0.0")
    out.append(f"bonds {len(bonds)}")
    out += bonds
    if hasattr(box, "bonds"):
        out.append(f"urey-bradley {len(box.ubs)}")
        out += [f"{a + 1} {b + 1} {_kcal(float(k))} {float(r0)!r} {float(beta)!r}"
                for (_, a, b, k, r0, beta) in box.ubs]
        out.append(f"angles {len(box.angles)}")
        out += [f"{a + 1} {b + 1} {c + 1} {ty} {_kcal(float(k))} {_deg(float(th))}"
                for (ty, a, b, c, k, th) in box.angles]
        for key, terms in (("dihedrals", box.dihes), ("impropers", box.imprs)):
            out.append(f"{key} {len(terms)}")
            out += [f"{a + 1} {b + 1} {c + 1} {d + 1} {ty} {order} {_kcal(float(k))} "
                    f"{_deg(float(phi))} {float(mult)!r}"
                    for (ty, order, a, b, c, d, k, phi, mult) in terms]
    eps14 = getattr(box, "eps14", None)
    sig14 = getattr(box, "sig14", None)
    out.append(f"vdw {n}")
    for i in range(n):
        e14 = box.eps[i] if eps14 is None else eps14[i]
        s14 = box.sig[i] if sig14 is None else sig14[i]
        out.append(f"{i + 1} {types[lab[i]]} {_eps(float(box.eps[i]))} {float(box.sig[i])!r} 0.0 "
                   f"{_eps(float(e14))} {float(s14)!r} 0.0")
    out.append("end")
    with open(path, "w") as f:
        f.write("\n".join(out) + "\n")


def write_rest(path: str, box, step: int = 0):
    """REST (io.c:1276-1369): step, atom count, ATOM records, positions,
    velocities, zero forces, a zero ENERGY record, chiT = chiP = 0."""
    n = box.n
    lab = labels_of(box)
    per = n // box.n_waters
    with open(path, "wb") as f:
        f.write(struct.pack("<ii", step, n))
        for i in range(n):
            res = i // per + 1
            f.write(struct.pack("<5s5s5sxiii", lab[i].encode()[:4], b"SYS", b"MOL", 0, res, res))
        for a in (box.x, box.y, box.z, box.vx, box.vy, box.vz):
            f.write(np.ascontiguousarray(a, dtype="<f8").tobytes())
        f.write(np.zeros(3 * n, dtype="<f8").tobytes())
        f.write(np.zeros(ENERGY_DOUBLES + 2, dtype="<f8").tobytes())


def write_deck(dirpath: str, box, cfg: dict, nsteps: int, **simu):
    """SIMU, CONF, FORF and (with restart, the default) REST in dirpath."""
    os.makedirs(dirpath, exist_ok=True)
    write_simu(os.path.join(dirpath, "SIMU"), box, cfg, nsteps, **simu)
    write_conf(os.path.join(dirpath, "CONF"), box)
    write_forf(os.path.join(dirpath, "FORF"), box)
    if simu.get("restart", True):
        write_rest(os.path.join(dirpath, "REST"), box)


def read_prop(path: str) -> np.ndarray:
    """PROP records (io.c:1230-1274) as an (nrec, 29) array: step, time,
    temperature, pressure, volume, then tot kin pot elec vdw bond ang ub
    dihe impr virtot virpot virelec virvdw virbond virub virshake consv in
    kcal/mol, then a b c alpha beta gamma."""
    return np.fromfile(path, dtype="<f8").reshape(-1, 29)


def read_rest(path: str) -> dict:
    """The REST file's step, count, positions, velocities and forces."""
    raw = open(path, "rb").read()
    step, n = struct.unpack_from("<ii", raw, 0)
    off = 8 + ATOM_BYTES * n
    arr = np.frombuffer(raw, dtype="<f8", count=9 * n, offset=off).reshape(9, n)
    return {"step": step, "n": n, "x": arr[0], "y": arr[1], "z": arr[2], "vx": arr[3],
            "vy": arr[4], "vz": arr[5], "fx": arr[6], "fy": arr[7], "fz": arr[8]}
