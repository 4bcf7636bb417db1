"""ctypes wrapper of the CPU FP64 oracle (test infrastructure only).

Only tests, the smoke check and the benchmark's CPU-baseline leg import
this package; it is the checker, never the product path.  It loads
``oracle/liboracle.so`` (the restatement in mdoracle.c) and, where it was
built, ``oracle/_ref/libmdbas_ref.so`` (the reference's own sources).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libmdbas_ref.so")

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class OBox(C.Structure):
    _fields_ = [("type", C.c_int)] + [(k, C.c_double) for k in (
        "a1 a2 a3 b1 b2 b3 c1 c2 c3 u1 u2 u3 v1 v2 v3 w1 w2 w3 pa pb pc u v w det vol".split())]


class OParam(C.Structure):
    _fields_ = [("nAtom", C.c_int), ("cutOff", C.c_double), ("cutOn", C.c_double),
                ("delr", C.c_double), ("chargeConst", C.c_double), ("alpha", C.c_double),
                ("m1", C.c_int), ("m2", C.c_int), ("m3", C.c_int), ("nbsp", C.c_int),
                ("timeStep", C.c_double), ("tolShake", C.c_double),
                ("maxCycle", C.c_int), ("update", C.c_int), ("linkRatio", C.c_int),
                ("cutOff2", C.c_double), ("cutOn2", C.c_double), ("switch2", C.c_double),
                ("rcutOff", C.c_double), ("rcutOff2", C.c_double), ("rTimeStep", C.c_double)]


class OEnergy(C.Structure):
    _fields_ = [(k, C.c_double) for k in
                "eDir eCorr eRec eVdw virDir virCorr virRec virVdw".split()] + [
                ("stressRec", C.c_double * 6), ("nInCut", C.c_int), ("nListed", C.c_int)]


class OBath(C.Structure):
    _fields_ = [("ens", C.c_int)] + [(k, C.c_double) for k in (
        "tauT tauP chiT chiP kinTemp0 temp0 press0 virpot conint consv kin virshake".split())] + [
        ("stress", C.c_double * 9)]


# Thermostat/barostat state handed between VV stages: BATH, the PARAM
# targets, the ENERGY members the integrators touch, the box vectors
# (a1..c3) and the nine box stresses.
ENS_SCALARS = "tauT tauP chiT chiP kinTemp0 temp0 press0 virpot conint consv kin virshake".split()


def ens_state(ens, box9, **kw):
    st = {k: float(kw.get(k, 0.)) for k in ENS_SCALARS}
    st["ens"] = int(ens)
    st["box9"] = np.array(box9, np.float64)
    st["stress"] = np.zeros(9)
    return st


def box_from9(box9, btype: int = 1) -> OBox:
    b = OBox()
    b.type = btype
    (b.a1, b.a2, b.a3, b.b1, b.b2, b.b3, b.c1, b.c2, b.c3) = (float(v) for v in box9)
    lib().oc_init_box(C.byref(b))
    return b


def build(ref: bool = True) -> None:
    """Compile liboracle.so (and the reference library when its sources are
    present) with the committed Makefile."""
    target = "all" if ref else "oracle"
    subprocess.run(["make", "-s", "-C", HERE, target], check=True)


_LIB = None
_REF = None


def lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            build(ref=False)
        L = C.CDLL(LIB_PATH)
        L.oc_spme.restype = C.c_double
        L.oc_ewald_classic.restype = C.c_double
        L.oc_dist.restype = C.c_double
        L.oc_ewald_dir.restype = C.c_double
        L.oc_ewald_corr.restype = C.c_double
        L.oc_vdw_switch.restype = C.c_double
        L.oc_verlet_list.restype = C.c_long
        L.oc_link_cell_list.restype = C.c_long
        _LIB = L
    return _LIB


def ref_available() -> bool:
    return os.path.exists(REF_PATH)


def ref():
    global _REF
    if _REF is None:
        R = C.CDLL(REF_PATH)
        for f in ("ref_ewald_dir", "ref_ewald_corr", "ref_vdw_switch", "ref_coulomb",
                  "ref_ewald_rec"):
            getattr(R, f).restype = C.c_double
        R.ref_create.restype = C.c_void_p
        R.ref_list_count.restype = C.c_long
        for f in ("ref_makelist", "ref_list_count", "ref_list_copy", "ref_excl_copy",
                  "ref_vdw_switch", "ref_coulomb", "ref_nonbond", "ref_ewald_rec", "ref_vv",
                  "ref_constvel", "ref_set_bonded", "ref_n14", "ref_list14_copy", "ref_bonded",
                  "ref_nb14", "ref_vv_ens"):
            getattr(R, f).argtypes = None
        _REF = R
    return _REF


def P(a):
    """ctypes pointer to a contiguous numpy array (double or int32)."""
    if a is None:
        return None
    if a.dtype == np.float64:
        return a.ctypes.data_as(_dp)
    if a.dtype == np.int32:
        return a.ctypes.data_as(_ip)
    raise TypeError(a.dtype)


def make_box(L: float) -> OBox:
    b = OBox()
    b.type = 1
    b.a1 = b.b2 = b.c3 = float(L)
    lib().oc_init_box(C.byref(b))
    return b


def make_param(cfg: dict, n: int) -> OParam:
    p = OParam()
    p.nAtom = n
    p.cutOff, p.cutOn, p.delr = cfg["cutOff"], cfg["cutOn"], cfg["delr"]
    p.chargeConst = 332.0716 * 418.4
    p.alpha = cfg["alpha"]
    p.m1 = p.m2 = p.m3 = cfg["grid"]
    p.nbsp = cfg["nbsp"]
    p.timeStep = cfg["timeStep"]
    p.tolShake = cfg["tolShake"]
    p.maxCycle, p.update, p.linkRatio = cfg["maxCycle"], cfg["update"], cfg["linkRatio"]
    lib().oc_param_derive(C.byref(p))
    return p


class OracleSystem:
    """One water box with oracle-side exclusions and list state."""

    def __init__(self, wb, cfg: dict | None = None):
        self.wb = wb
        self.cfg = dict(wb.params) if cfg is None else cfg
        self.p = make_param(self.cfg, wb.n)
        self.b = make_box(wb.L)
        self.keyLink = lib().oc_choose_link(C.byref(self.p), C.byref(self.b))
        self.exOff = np.zeros(wb.n + 1, np.int32)
        self.bonded_topology = hasattr(wb, "bonds")
        if self.bonded_topology:
            t = topology_arrays(wb)
            cap = 16 * (len(wb.ca) + len(wb.bonds) + len(wb.angles) + len(wb.dihes)
                        + len(wb.imprs)) + 16
            self.exLst = np.zeros(cap, np.int32)
            l14 = np.zeros(2 * (5 * len(wb.dihes) + 1), np.int32)
            n14 = C.c_int(0)
            ne = lib().oc_exclusions14(wb.n, len(wb.bonds), P(t["ba"]), P(t["bb"]), len(wb.ca),
                                       P(wb.ca), P(wb.cb), len(wb.angles), P(t["aa"]),
                                       P(t["ab"]), P(t["ac"]), len(wb.dihes), P(t["d4"]),
                                       len(wb.imprs), P(t["i4"]), int(self.keyLink),
                                       P(self.exOff), P(self.exLst), cap, P(l14), C.byref(n14))
            self.list14 = l14[:2 * n14.value].copy()
        else:
            cap = 8 * len(wb.ca) + 16
            self.exLst = np.zeros(cap, np.int32)
            ne = lib().oc_exclusions(wb.n, 0, None, None, len(wb.ca), P(wb.ca), P(wb.cb), 0, None,
                                     None, None, int(self.keyLink), P(self.exOff), P(self.exLst),
                                     cap)
            self.list14 = np.zeros(0, np.int32)
        if ne < 0:
            raise RuntimeError(f"oracle exclusions failed: {ne}")
        self.exLst = self.exLst[:ne].copy()
        self.neighList = None
        self.q, self.eps, self.sig = wb.q, wb.eps, wb.sig

    def set_params(self, q=None, eps=None, sig=None):
        """Per-atom nonbonded parameters used by energy() and energy_grad()
        (the water box's own when None)."""
        self.q = self.wb.q if q is None else np.ascontiguousarray(q, np.float64)
        self.eps = self.wb.eps if eps is None else np.ascontiguousarray(eps, np.float64)
        self.sig = self.wb.sig if sig is None else np.ascontiguousarray(sig, np.float64)

    def build_list(self, x, y, z, build: bool = True):
        """Reference-format list; x,y,z may be wrapped in place (link cells,
        first build) exactly as list.c does."""
        n = self.wb.n
        np_ = np.zeros(n, np.int32)
        no = np.zeros(n, np.int32)
        # room for the expected pair count (half the atoms within cutOff +
        # delr of each atom, with headroom), so one pass normally suffices
        rl = self.cfg["cutOff"] + self.cfg["delr"]
        cap = max(getattr(self, "_cap", 0),
                  int(0.5 * n * 4.18879 * rl ** 3 * n / self.wb.L ** 3 * 1.3) + 1024)
        nl = np.zeros(cap, np.int32)
        for _ in range(2):
            if self.keyLink:
                cnt = lib().oc_link_cell_list(C.byref(self.p), C.byref(self.b), int(build),
                                              P(x), P(y), P(z), P(self.wb.frozen), P(self.exOff),
                                              P(self.exLst), P(nl), C.c_long(cap), P(np_), P(no))
            else:
                cnt = lib().oc_verlet_list(C.byref(self.p), C.byref(self.b), P(x), P(y), P(z),
                                           P(self.wb.frozen), P(self.exOff), P(self.exLst),
                                           P(nl), C.c_long(cap), P(np_), P(no))
            if cnt < 0:
                raise RuntimeError(f"oracle list failed: {cnt}")
            if cnt <= cap:
                break
            cap = int(cnt)
            nl = np.zeros(cap, np.int32)
            build = False
        self._cap = cap
        self.neighList, self.neighPair, self.neighOrder = nl[:cnt], np_, no
        return self.neighList, self.neighPair, self.neighOrder

    def ewald_real(self, x, y, z, f=None):
        n = self.wb.n
        f = np.zeros((3, n)) if f is None else f
        e = OEnergy()
        lib().oc_ewald_real(C.byref(self.p), C.byref(self.b), P(x), P(y), P(z),
                            P(f[0]), P(f[1]), P(f[2]), P(self.q), P(self.eps),
                            P(self.sig), P(self.neighList), P(self.neighPair),
                            P(self.neighOrder), P(self.exOff), P(self.exLst), C.byref(e))
        return f, e

    def spme(self, x, y, z, f=None):
        n = self.wb.n
        f = np.zeros((3, n)) if f is None else f
        st = np.zeros(6)
        vir = C.c_double(0)
        e = lib().oc_spme(C.byref(self.p), C.byref(self.b), P(x), P(y), P(z),
                          P(f[0]), P(f[1]), P(f[2]), P(self.q), P(st), C.byref(vir))
        return e, f, st, vir.value

    def energy_grad(self, x, y, z):
        """dE_pot/d(q, eps, sig) per atom at x,y,z over the current list."""
        n = self.wb.n
        g = np.zeros((3, n))
        lib().oc_energy_grad(C.byref(self.p), C.byref(self.b), P(x), P(y), P(z), P(self.q),
                             P(self.eps), P(self.sig), P(self.neighList), P(self.neighPair),
                             P(self.neighOrder), P(self.exOff), P(self.exLst), P(g[0]),
                             P(g[1]), P(g[2]))
        return g

    def energy(self, x, y, z):
        """energy() (energy.c:103-205) for the Ewald path: zero forces, SPME,
        real-space pairs, correction.  Returns forces (3,N) and a dict."""
        eRec, f, st, vr = self.spme(x, y, z)
        f, e = self.ewald_real(x, y, z, f)
        elec = e.eDir + eRec + e.eCorr
        return f, dict(elec=elec, vdw=e.eVdw, pot=elec + e.eVdw, eDir=e.eDir, eRec=eRec,
                       eCorr=e.eCorr, virelec=e.virDir + vr + e.virCorr, virvdw=e.virVdw,
                       nInCut=e.nInCut, nListed=e.nListed, stressRec=st)

    def nonbond_energy(self, elec_type, x, y, z):
        """energy() on the plain cutoff path (keyEwald 0, energy.c:151-161):
        zeroed forces, nonbond_energy with the coulomb functor of elec_type
        (SHIFT1 2, SHIFT2 3, SWITCH 4) and the LJ switch over the current
        list.  Returns forces (3,N) and a dict."""
        n = self.wb.n
        f = np.zeros((3, n))
        e = [C.c_double(0) for _ in range(4)]
        lib().oc_nonbond(C.byref(self.p), C.byref(self.b), int(elec_type), P(x), P(y), P(z),
                         P(f[0]), P(f[1]), P(f[2]), P(self.q), P(self.eps), P(self.sig),
                         P(self.neighList), P(self.neighPair), P(self.neighOrder),
                         *[C.byref(v) for v in e])
        elec, vdw, ve, vv = (v.value for v in e)
        return f, dict(elec=elec, vdw=vdw, pot=elec + vdw, virelec=ve, virvdw=vv)

    def md_run_plain(self, elec_type, nsteps, x, y, z, vx, vy, vz):
        """The MD loop of mdbas.c:244-363 on the plain cutoff path: list
        build, step-0 energy, then per step VV stage 1, the list update
        every `update` steps, nonbond_energy, VV stage 2.  Returns (err, f,
        etot, epot) with per-step energies (length nsteps + 1)."""
        upd = self.cfg["update"]
        self.build_list(x, y, z, build=True)
        f, e = self.nonbond_energy(elec_type, x, y, z)
        m = self.wb.mass
        ek = 0.5 * float((m * (vx * vx + vy * vy + vz * vz)).sum())
        et, ep = [e["pot"] + ek], [e["pot"]]
        for step in range(1, nsteps + 1):
            err, _, _, _ = self.vv(1, x, y, z, vx, vy, vz, f)
            if err:
                return err, f, np.array(et), np.array(ep)
            if step % upd == 0:
                self.build_list(x, y, z, build=False)
            f, e = self.nonbond_energy(elec_type, x, y, z)
            err, ek, _, _ = self.vv(2, x, y, z, vx, vy, vz, f)
            if err:
                return err, f, np.array(et), np.array(ep)
            et.append(e["pot"] + ek)
            ep.append(e["pot"])
        return 0, f, np.array(et), np.array(ep)

    def bonded(self, x, y, z, f=None):
        """Bonded energies (internal.c:21-604) in energy()'s order; forces
        accumulate into f.  Returns (f, [bond, ub, ang, dihe, impr, virbond,
        virub])."""
        wb = self.wb
        f = np.zeros((3, wb.n)) if f is None else f
        out = np.zeros(7)
        lib().oc_bonded(C.byref(self.b), len(wb.bonds), wb.c_bonds(), len(wb.ubs), wb.c_ubs(),
                        len(wb.angles), wb.c_angles(), len(wb.dihes), wb.c_dihes(),
                        len(wb.imprs), wb.c_imprs(), P(x), P(y), P(z), P(f[0]), P(f[1]),
                        P(f[2]), P(out))
        return f, out

    def nb14(self, ewald, elec_type, x, y, z, f=None, vdw_on=1):
        """1-4 pairs (ewald14_energy or nonbond14_energy); returns (f,
        [elec, vdw, virelec, virvdw])."""
        wb = self.wb
        f = np.zeros((3, wb.n)) if f is None else f
        out = np.zeros(4)
        lib().oc_nb14(C.byref(self.p), C.byref(self.b), int(ewald), int(elec_type), int(vdw_on),
                      C.c_double(wb.scal14), len(self.list14) // 2, P(self.list14), P(x), P(y),
                      P(z), P(f[0]), P(f[1]), P(f[2]), P(self.q), P(wb.eps14), P(wb.sig14),
                      P(out))
        return f, out

    def energy_full(self, x, y, z):
        """energy() (energy.c:103-205) on the Ewald path with 1-4 pairs and
        every bonded term: forces (3,N) and the ENERGY fields."""
        f, e = self.energy(x, y, z)
        f, e14 = self.nb14(1, 0, x, y, z, f)
        f, eb = self.bonded(x, y, z, f)
        elec = e["elec"] + e14[0]
        vdw = e["vdw"] + e14[1]
        d = dict(e, elec=elec, vdw=vdw, bond=eb[0], ub=eb[1], ang=eb[2], dihe=eb[3], impr=eb[4],
                 virbond=eb[5], virub=eb[6], virelec=e["virelec"] + e14[2],
                 virvdw=e["virvdw"] + e14[3])
        d["pot"] = elec + vdw + eb[0] + eb[2] + eb[1] + eb[3] + eb[4]
        return f, d

    def md_run_full(self, nsteps, x, y, z, vx, vy, vz):
        """The MD loop of mdbas.c:244-363 (as oc_md_run) with energy_full:
        list build, step-0 energy, then per step VV stage 1, the list update
        every `update` steps, energy_full, VV stage 2.  Returns (err, f,
        etot, epot) with per-step energies (length nsteps + 1)."""
        upd = self.cfg["update"]
        self.build_list(x, y, z, build=True)
        f, e = self.energy_full(x, y, z)
        m = self.wb.mass
        ek = 0.5 * float((m * (vx * vx + vy * vy + vz * vz)).sum())
        et, ep = [e["pot"] + ek], [e["pot"]]
        for step in range(1, nsteps + 1):
            err, _, _, _ = self.vv(1, x, y, z, vx, vy, vz, f)
            if err:
                return err, f, np.array(et), np.array(ep)
            if step % upd == 0:
                self.build_list(x, y, z, build=False)
            f, e = self.energy_full(x, y, z)
            err, ek, _, _ = self.vv(2, x, y, z, vx, vy, vz, f)
            if err:
                return err, f, np.array(et), np.array(ep)
            et.append(e["pot"] + ek)
            ep.append(e["pot"])
        return 0, f, np.array(et), np.array(ep)

    def vv(self, stage, x, y, z, vx, vy, vz, f):
        ek = C.c_double(0)
        vs = C.c_double(0)
        nc = C.c_int(0)
        wb = self.wb
        err = lib().oc_vv_nve(C.byref(self.p), C.byref(self.b), int(stage), len(wb.ca), P(wb.ca),
                              P(wb.cb), P(wb.rc2), P(x), P(y), P(z), P(vx), P(vy), P(vz),
                              P(f[0]), P(f[1]), P(f[2]), P(wb.mass), P(wb.rmass),
                              P(wb.nAtConst), C.byref(ek), C.byref(vs), C.byref(nc))
        return err, ek.value, vs.value, nc.value

    def vv_conv(self, stage, x, y, z, vx, vy, vz, f, cap=None):
        """vv with each constraint's convergence test at every SHAKE sweep:
        returns (err, conv) with conv a bool array [sweeps run, nConst]."""
        wb = self.wb
        cap = self.cfg["maxCycle"] if cap is None else cap
        conv = np.zeros((cap, len(wb.ca)), np.uint8)
        nc = C.c_int(0)
        err = lib().oc_vv_nve_conv(C.byref(self.p), C.byref(self.b), int(stage), len(wb.ca),
                                   P(wb.ca), P(wb.cb), P(wb.rc2), P(x), P(y), P(z), P(vx), P(vy),
                                   P(vz), P(f[0]), P(f[1]), P(f[2]), P(wb.mass), P(wb.rmass),
                                   P(wb.nAtConst), conv.ctypes.data_as(C.c_void_p), int(cap),
                                   C.byref(nc))
        return err, conv[:min(nc.value, cap)].astype(bool)

    def vv_ens(self, st, stage, x, y, z, vx, vy, vz, f):
        """oc_vv_ens: one stage of a thermostat/barostat integrator; st
        (ens_state) is updated in place, box included."""
        wb = self.wb
        b = box_from9(st["box9"], self.b.type)
        s = OBath()
        s.ens = st["ens"]
        for k in ENS_SCALARS:
            setattr(s, k, st[k])
        for k in range(9):
            s.stress[k] = st["stress"][k]
        nc = C.c_int(0)
        err = lib().oc_vv_ens(C.byref(self.p), C.byref(b), C.byref(s), int(stage), len(wb.ca),
                              P(wb.ca), P(wb.cb), P(wb.rc2), P(x), P(y), P(z), P(vx), P(vy),
                              P(vz), P(f[0]), P(f[1]), P(f[2]), P(wb.mass), P(wb.rmass),
                              P(wb.nAtConst), C.byref(nc))
        for k in ENS_SCALARS:
            st[k] = getattr(s, k)
        st["stress"] = np.array(s.stress[:])
        st["box9"] = np.array([b.a1, b.a2, b.a3, b.b1, b.b2, b.b3, b.c1, b.c2, b.c3])
        return err

    def md_run_ens(self, st, nsteps, x, y, z, vx, vy, vz):
        """The MD loop of mdbas.c:244-363 on the Ewald path with the
        thermostat/barostat of st (ens_state, updated in place): consv
        cleared at each step's start, VV stage 1, the list update every
        `update` steps, energy() with virpot = virelec + virvdw
        (energy.c:196), VV stage 2; the box follows st between calls.
        Returns (err, f, records) with one record per step from step 0:
        tot, pot, kin, consv, volume."""
        upd = self.cfg["update"]
        btype = self.b.type
        self.build_list(x, y, z, build=True)
        f, e = self.energy(x, y, z)
        st["virpot"] = e["virelec"] + e["virvdw"]
        m = self.wb.mass
        ek = 0.5 * float((m * (vx * vx + vy * vy + vz * vz)).sum())
        st["kin"] = ek
        rec = [(e["pot"] + ek, e["pot"], ek, st["consv"], self.b.vol)]
        for step in range(1, nsteps + 1):
            st["consv"] = 0.
            err = self.vv_ens(st, 1, x, y, z, vx, vy, vz, f)
            if err:
                return err, f, np.array(rec)
            self.b = box_from9(st["box9"], btype)
            if step % upd == 0:
                self.build_list(x, y, z, build=False)
            f, e = self.energy(x, y, z)
            st["virpot"] = e["virelec"] + e["virvdw"]
            err = self.vv_ens(st, 2, x, y, z, vx, vy, vz, f)
            if err:
                return err, f, np.array(rec)
            self.b = box_from9(st["box9"], btype)
            rec.append((e["pot"] + st["kin"], e["pot"], st["kin"], st["consv"], self.b.vol))
        return 0, f, np.array(rec)

    def sym_exclusions(self):
        """The exclusion table with each pair at both atoms (sym = 1)."""
        if not hasattr(self, "_sym"):
            wb = self.wb
            off = np.zeros(wb.n + 1, np.int32)
            cap = 8 * len(wb.ca) + 16
            lst = np.zeros(cap, np.int32)
            ne = lib().oc_exclusions(wb.n, 0, None, None, len(wb.ca), P(wb.ca), P(wb.cb), 0, None,
                                     None, None, 1, P(off), P(lst), cap)
            if ne < 0:
                raise RuntimeError(f"oracle exclusions failed: {ne}")
            self._sym = (off, lst[:ne].copy())
        return self._sym

    def hvp(self, x, y, z, u):
        """oc_hvp over the current list: H u (3, n) and d/d(q, eps, sig)
        (grad E . u)."""
        n = self.wb.n
        u = np.ascontiguousarray(u, np.float64).reshape(3, n)
        hu = np.zeros((3, n))
        mq, me, ms = np.zeros(n), np.zeros(n), np.zeros(n)
        off, lst = self.sym_exclusions()
        lib().oc_hvp(C.byref(self.p), C.byref(self.b), P(x), P(y), P(z), P(self.q), P(self.eps),
                     P(self.sig), P(self.neighList), P(self.neighPair), P(self.neighOrder), P(off),
                     P(lst), P(u), P(hu), P(mq), P(me), P(ms))
        return hu, mq, me, ms

    def spme_hvp(self, x, y, z, u=None):
        """oc_spme_hvp: H_rec u, d/dq (grad E_rec . u) and grad E_rec."""
        n = self.wb.n
        hu, mq, g = np.zeros((3, n)), np.zeros(n), np.zeros((3, n))
        uu = None if u is None else np.ascontiguousarray(u, np.float64).reshape(3, n)
        lib().oc_spme_hvp(C.byref(self.p), C.byref(self.b), P(x), P(y), P(z), P(self.q), P(uu),
                          P(hu), P(mq), P(g))
        return hu, mq, g

    def traj_adjoint(self, nsteps, x, y, z, vx, vy, vz):
        """oc_traj_adjoint from (x, v) (not modified): J and dJ/d(q, eps,
        sig) of the time-averaged potential over steps 1..nsteps."""
        wb, n = self.wb, self.wb.n
        J = C.c_double(0)
        dq, de, ds = np.zeros(n), np.zeros(n), np.zeros(n)
        err = lib().oc_traj_adjoint(C.byref(self.p), C.byref(self.b), int(self.keyLink),
                                    int(nsteps), P(x), P(y), P(z), P(vx), P(vy), P(vz), P(self.q),
                                    P(self.eps), P(self.sig), P(wb.mass), P(wb.rmass),
                                    P(wb.frozen), P(wb.nAtConst), len(wb.ca), P(wb.ca), P(wb.cb),
                                    P(wb.rc2), C.byref(J), P(dq), P(de), P(ds))
        if err:
            raise RuntimeError(f"oc_traj_adjoint failed: {err}")
        return J.value, dq, de, ds

    def md_run(self, nsteps, x, y, z, vx, vy, vz):
        n = self.wb.n
        f = np.zeros((3, n))
        et, ep, ek = (np.zeros(nsteps + 1) for _ in range(3))
        sec = C.c_double(0)
        wb = self.wb
        err = lib().oc_md_run(C.byref(self.p), C.byref(self.b), int(self.keyLink), int(nsteps),
                              P(x), P(y), P(z), P(vx), P(vy), P(vz), P(f[0]), P(f[1]), P(f[2]),
                              P(self.q), P(self.eps), P(self.sig), P(wb.mass), P(wb.rmass),
                              P(wb.frozen), P(wb.nAtConst), len(wb.ca), P(wb.ca), P(wb.cb),
                              P(wb.rc2), P(et), P(ep), P(ek), C.byref(sec))
        return err, f, et, ep, ek, sec.value


def topology_arrays(wb):
    """Flat int32 index arrays of a bonded topology for the C oracle."""
    i32 = lambda v: np.ascontiguousarray(np.array(v, np.int32).reshape(-1))  # noqa: E731
    return dict(ba=i32([b[1] for b in wb.bonds]), bb=i32([b[2] for b in wb.bonds]),
                aa=i32([a[1] for a in wb.angles]), ab=i32([a[2] for a in wb.angles]),
                ac=i32([a[3] for a in wb.angles]),
                d4=i32([d[2:6] for d in wb.dihes]), i4=i32([d[2:6] for d in wb.imprs]))


class RefSystem:
    """The compiled reference (oracle/_ref) driven on the same arrays."""

    def __init__(self, wb, cfg: dict | None = None, noLink: int = 0):
        cfg = dict(wb.params) if cfg is None else cfg
        self.wb, self.cfg = wb, cfg
        R = ref()
        R.ref_create.argtypes = [C.c_int, C.c_int, _dp, C.c_double, C.c_double, C.c_double,
                                 C.c_double, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double,
                                 C.c_int, C.c_int, _ip, _ip, _dp, _ip, _ip]
        self.box9 = wb.box9
        self.h = C.c_void_p(R.ref_create(wb.n, 1, P(self.box9), cfg["cutOff"], cfg["cutOn"],
                                         cfg["delr"], cfg["alpha"], cfg["linkRatio"],
                                         cfg["update"], noLink, cfg["timeStep"],
                                         cfg["tolShake"], cfg["maxCycle"], len(wb.ca), P(wb.ca),
                                         P(wb.cb), P(wb.rc2), P(wb.frozen), P(wb.nAtConst)))
        if hasattr(wb, "bonds"):
            self.set_bonded()

    def set_bonded(self):
        """Hands the molecule topology to the reference structs."""
        wb = self.wb
        ref().ref_set_bonded(self.h, len(wb.bonds), wb.c_bonds(), len(wb.ubs), wb.c_ubs(),
                             len(wb.angles), wb.c_angles(), len(wb.dihes), wb.c_dihes(),
                             len(wb.imprs), wb.c_imprs(), C.c_double(wb.scal14))

    def list14(self):
        R = ref()
        n = R.ref_n14(self.h)
        out = np.zeros(max(1, 2 * n), np.int32)
        R.ref_list14_copy(self.h, P(out))
        return out[:2 * n]

    def bonded(self, x, y, z, f=None):
        n = self.wb.n
        f = np.zeros((3, n)) if f is None else f
        out = np.zeros(7)
        ref().ref_bonded(self.h, P(x), P(y), P(z), P(f[0]), P(f[1]), P(f[2]), P(out))
        return f, out

    def nb14(self, ewald, elec_type, x, y, z, f=None, vdw_type=2):
        n = self.wb.n
        f = np.zeros((3, n)) if f is None else f
        out = np.zeros(4)
        ref().ref_nb14(self.h, C.c_int(ewald), C.c_int(elec_type), C.c_int(vdw_type), P(x), P(y),
                       P(z), P(f[0]), P(f[1]), P(f[2]), P(self.wb.q), P(self.wb.eps14),
                       P(self.wb.sig14), P(out))
        return f, out

    def makelist(self, x, y, z, step=0):
        R = ref()
        keylink = R.ref_makelist(self.h, P(x), P(y), P(z), C.c_int(step))
        cnt = R.ref_list_count(self.h)
        n = self.wb.n
        nl = np.zeros(max(cnt, 1), np.int32)
        np_ = np.zeros(n, np.int32)
        no = np.zeros(n, np.int32)
        R.ref_list_copy(self.h, P(nl), P(np_), P(no))
        return keylink, nl[:cnt], np_, no

    def exclusions(self):
        n = self.wb.n
        off = np.zeros(n + 1, np.int32)
        lst = np.zeros(1, np.int32)
        cap = ref().ref_excl_copy(self.h, P(off), P(lst), C.c_int(0))
        lst = np.zeros(max(cap, 1), np.int32)
        ne = ref().ref_excl_copy(self.h, P(off), P(lst), C.c_int(cap))
        return off, lst[:ne]

    def nonbond(self, elecType, x, y, z):
        n = self.wb.n
        f = np.zeros((3, n))
        out = np.zeros(4)
        ref().ref_nonbond(self.h, C.c_int(elecType), P(x), P(y), P(z), P(f[0]), P(f[1]),
                          P(f[2]), P(self.wb.q), P(self.wb.eps), P(self.wb.sig), P(out))
        return f, out

    def ewald_rec(self, kmax, x, y, z):
        n = self.wb.n
        f = np.zeros((3, n))
        st = np.zeros(6)
        vir = C.c_double(0)
        e = ref().ref_ewald_rec(self.h, C.c_int(kmax), P(x), P(y), P(z), P(f[0]), P(f[1]),
                                P(f[2]), P(self.wb.q), P(st), C.byref(vir))
        return e, f, st, vir.value

    def vv(self, stage, x, y, z, vx, vy, vz, f):
        out = np.zeros(2)
        ref().ref_vv(self.h, C.c_int(stage), P(x), P(y), P(z), P(vx), P(vy), P(vz),
                     P(f[0]), P(f[1]), P(f[2]), P(self.wb.mass), P(self.wb.rmass), P(out))
        return out

    def vv_ens(self, st, stage, x, y, z, vx, vy, vz, f):
        """vv_integrate with st["ens"] through the compiled reference; st
        (ens_state) is updated in place."""
        a = np.zeros(31)
        a[0] = st["ens"]
        for i, k in enumerate(ENS_SCALARS):
            a[1 + i] = st[k]
        a[13:22] = st["box9"]
        a[22:31] = st["stress"]
        ref().ref_vv_ens(self.h, C.c_int(stage), P(a), P(x), P(y), P(z), P(vx), P(vy), P(vz),
                         P(f[0]), P(f[1]), P(f[2]), P(self.wb.mass), P(self.wb.rmass))
        for i, k in enumerate(ENS_SCALARS):
            st[k] = float(a[1 + i])
        st["box9"] = a[13:22].copy()
        st["stress"] = a[22:31].copy()

    def constvel(self, x, y, z, vx, vy, vz):
        ref().ref_constvel(self.h, P(x), P(y), P(z), P(vx), P(vy), P(vz), P(self.wb.mass))
