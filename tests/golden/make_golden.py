"""Writes tests/golden/*.npz from the compiled reference (oracle/_ref).

The reference ships no tests, fixtures or example inputs, so its own
compiled code (oracle/Makefile builds list.c, ewald.c, vdw.c, elec.c,
utils.c, shake.c, integrate.c, init.c, energy.c and internal.c from the
reference tree)
is driven here on synthetic inputs and its outputs are stored as small
fixtures.  The test suite checks the FP64 oracle against them, which pins
the oracle on machines where the reference tree is absent.

Run from the repository root:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from mdbas_amd import configs  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def digest(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def functors():
    import ctypes as C
    R = O.ref()
    wb = configs.build(configs.CONFIGS["c0"])
    rs = O.RefSystem(wb)
    r = np.concatenate([np.linspace(0.8, 12.0, 400), [10.0, 10.0 + 1e-12, 12.0]])
    alpha = 0.32
    qel = 138935.4835 * 0.834 * 0.417
    out = {"r": r}
    for name, fn in (("dir", R.ref_ewald_dir), ("corr", R.ref_ewald_corr)):
        e = np.zeros_like(r)
        d = np.zeros_like(r)
        for k, rr in enumerate(r):
            dE = C.c_double(0)
            fn.argtypes = [C.c_double] * 4 + [C.POINTER(C.c_double)]
            e[k] = fn(alpha, qel, rr, 1.0 / rr, C.byref(dE))
            d[k] = dE.value
        out[f"{name}_E"], out[f"{name}_dE"] = e, d
    veps, vsig = wb.eps[0] * wb.eps[0], wb.sig[0] + wb.sig[0]
    e = np.zeros_like(r)
    d = np.zeros_like(r)
    R.ref_vdw_switch.argtypes = [C.c_void_p] + [C.c_double] * 4 + [C.POINTER(C.c_double)]
    for k, rr in enumerate(r):
        dE = C.c_double(0)
        e[k] = R.ref_vdw_switch(rs.h, veps, vsig, rr * rr, 1.0 / rr, C.byref(dE))
        d[k] = dE.value
    out["vdw_E"], out["vdw_dE"] = e, d
    out["vdw_par"] = np.array([veps, vsig])
    R.ref_coulomb.argtypes = [C.c_void_p, C.c_int] + [C.c_double] * 3 + [C.POINTER(C.c_double)]
    for t in (2, 3, 4):
        e = np.zeros_like(r)
        d = np.zeros_like(r)
        for k, rr in enumerate(r):
            dE = C.c_double(0)
            e[k] = R.ref_coulomb(rs.h, t, -0.834 * 0.417, rr * rr, 1.0 / rr, C.byref(dE))
            d[k] = dE.value
        out[f"coul{t}_E"], out[f"coul{t}_dE"] = e, d
    out["alpha"], out["qel"] = np.array(alpha), np.array(qel)
    # nint and init_box
    xs = np.array([-2.5, -1.5, -0.5, -0.49999999999999994, 0.0, 0.49999999999999994, 0.5, 1.5,
                   2.5, 1e-300, -1e-300])
    R.ref_nint.argtypes = [C.c_double]
    out["nint_x"] = xs
    out["nint_v"] = np.array([R.ref_nint(v) for v in xs], np.int32)
    boxes = np.array([[31.04, 0, 0, 0, 31.04, 0, 0, 0, 31.04],
                      [67.0, 0, 0, 0, 67.0, 0, 0, 0, 67.0],
                      [30.0, 0, 0, 0, 32.5, 0, 0, 0, 35.1],
                      [30.0, 0, 0, 5.0, 31.0, 0, -3.0, 4.0, 33.0]])
    bo = np.zeros((len(boxes), 24))
    for k, b9 in enumerate(boxes):
        R.ref_init_box(1, O.P(np.ascontiguousarray(b9)), O.P(bo[k]))
    out["boxes"], out["box_out"] = boxes, bo
    np.savez_compressed(os.path.join(OUT, "functors.npz"), **out)


def lists():
    out = {}
    cases = [("c0", configs.CONFIGS["c0"]),
             ("link3k", dict(configs.COMMON, n_waters=3000, L=(3000 / 0.0334) ** (1 / 3), seed=5,
                             grid=32, cutOff=9.0, cutOn=7.5, timeStep=0.002, alpha=0.35)),
             ("w64", dict(configs.COMMON, n_waters=64, L=(64 / 0.0334) ** (1 / 3), seed=9,
                          grid=16, cutOff=9.0, cutOn=7.5, timeStep=0.002, alpha=0.35)),
             # the two benchmark boxes: 30,000 atoms (4^3 link cells) and
             # 192,000 atoms (8^3 link cells), 12 A cutoff, 2 A skin
             ("c1", configs.CONFIGS["c1"]),
             ("c2", configs.CONFIGS["c2"])]
    for name, cfg in cases:
        wb = configs.build(cfg)
        rs = O.RefSystem(wb, cfg)
        x, y, z = wb.x.copy(), wb.y.copy(), wb.z.copy()
        kl, nl, np_, no = rs.makelist(x, y, z, 0)
        off, lst = rs.exclusions()
        out[f"{name}_keylink"] = np.array(kl)
        out[f"{name}_count"] = np.array(len(nl))
        out[f"{name}_digest"] = np.array(digest(nl, np_, no))
        out[f"{name}_wrapped_digest"] = np.array(digest(x, y, z))
        out[f"{name}_excl_digest"] = np.array(digest(off, lst))
        if name == "w64":
            out["w64_nl"], out["w64_np"], out["w64_no"] = nl, np_, no
        # update build after a seeded displacement of the wrapped coordinates
        rng = np.random.default_rng(1234)
        for a in (x, y, z):
            a += rng.normal(0.0, 0.05, len(a))
        xp, yp, zp = x.copy(), y.copy(), z.copy()
        kl, nl, np_, no = rs.makelist(x, y, z, cfg["update"])
        out[f"{name}_upd_count"] = np.array(len(nl))
        out[f"{name}_upd_digest"] = np.array(digest(nl, np_, no))
        out[f"{name}_upd_pos_digest"] = np.array(digest(xp, yp, zp))
    np.savez_compressed(os.path.join(OUT, "lists.npz"), **out)


def dynamics():
    """VV + SHAKE with the reference's own integrator, on forces from the
    reference's nonbond_energy (SWITCH Coulomb + switched LJ), 10 steps of
    the 64-water box; plus classic-Ewald energies for the SPME limit."""
    cfg = dict(configs.COMMON, n_waters=64, L=(64 / 0.0334) ** (1 / 3), seed=9, grid=16,
               cutOff=9.0, cutOn=7.5, timeStep=0.002, alpha=0.35, update=5)
    wb = configs.build(cfg)
    rs = O.RefSystem(wb, cfg)
    st = [wb.x.copy(), wb.y.copy(), wb.z.copy(), wb.vx.copy(), wb.vy.copy(), wb.vz.copy()]
    out = {"x0": np.stack(st)}
    rs.makelist(st[0], st[1], st[2], 0)
    f, e = rs.nonbond(4, st[0], st[1], st[2])
    kin, vshake, pots = [], [], []
    for step in range(1, 11):
        o1 = rs.vv(1, *st, f)
        rs.makelist(st[0], st[1], st[2], step)
        f, e = rs.nonbond(4, st[0], st[1], st[2])
        o2 = rs.vv(2, *st, f)
        kin.append(o2[0])
        vshake.append(o1[1])
        pots.append(e[0] + e[1])
    out["x10"] = np.stack(st)
    out["kin"], out["virshake"], out["pot"] = np.array(kin), np.array(vshake), np.array(pots)
    # classic Ewald reciprocal energy / forces at several kmax (64 waters)
    xs = wb.x.copy(), wb.y.copy(), wb.z.copy()
    ek = []
    for kmax in (4, 6, 8, 10):
        er, fr, stres, vir = rs.ewald_rec(kmax, *xs)
        ek.append(er)
        if kmax == 10:
            out["ewald_f10"] = fr
            out["ewald_vir10"] = np.array(vir)
    out["ewald_kmax"] = np.array([4, 6, 8, 10])
    out["ewald_E"] = np.array(ek)
    out["ewald_pos"] = np.stack(xs)
    # init_constvel on a raw Gaussian draw
    v = np.stack([wb.vx.copy(), wb.vy.copy(), wb.vz.copy()])
    rng = np.random.default_rng(7)
    v += rng.normal(0, 1.0, v.shape) * np.sqrt(wb.rmass)
    out["cv_in"] = v.copy()
    vx, vy, vz = v[0].copy(), v[1].copy(), v[2].copy()
    rs.constvel(wb.x.copy(), wb.y.copy(), wb.z.copy(), vx, vy, vz)
    out["cv_out"] = np.stack([vx, vy, vz])
    np.savez_compressed(os.path.join(OUT, "dynamics.npz"), **out)


def bonded():
    """The molecule boxes of mdbas_amd.molecules through the compiled
    reference: exclusions and 1-4 list of makelist (list.c:108-894,
    1125-1791), the pair-list digest, the bonded energies and forces of
    internal.c (bond, angle, Urey-Bradley, dihedral, improper) and the 1-4
    pair energies and forces of ewald14_energy / nonbond14_energy
    (energy.c:356-439, 654-742) for the Ewald path and SHIFT1 / SHIFT2 /
    SWITCH."""
    from mdbas_amd import molecules as Mo
    out = {}
    for name in ("flex216", "flex729"):
        cfg = Mo.MOL_CONFIGS[name]
        wb = Mo.build(cfg)
        rs = O.RefSystem(wb, cfg)
        x, y, z = wb.x.copy(), wb.y.copy(), wb.z.copy()
        kl, nl, np_, no = rs.makelist(x, y, z, 0)
        off, lst = rs.exclusions()
        out[f"{name}_keylink"] = np.array(kl)
        out[f"{name}_digest"] = np.array(digest(nl, np_, no))
        out[f"{name}_excl_digest"] = np.array(digest(off, lst))
        out[f"{name}_list14"] = rs.list14()
        f, e = rs.bonded(x, y, z)
        out[f"{name}_bonded_E"] = e
        if name == "flex216":
            out[f"{name}_bonded_f"] = f
            out[f"{name}_pos"] = np.stack([x, y, z])
        for ew, et in ((1, 0), (0, 2), (0, 3), (0, 4)):
            f, e = rs.nb14(ew, et, x, y, z)
            out[f"{name}_nb14_{ew}{et}_E"] = e
            if name == "flex216":
                out[f"{name}_nb14_{ew}{et}_f"] = f
    np.savez_compressed(os.path.join(OUT, "bonded.npz"), **out)


if __name__ == "__main__":
    if not O.ref_available():
        O.build(ref=True)
    functors()
    lists()
    dynamics()
    bonded()
    for f in sorted(os.listdir(OUT)):
        if f.endswith(".npz"):
            print(f, os.path.getsize(os.path.join(OUT, f)), "bytes")
