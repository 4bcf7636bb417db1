"""mdbas_amd -- MI355X (gfx950) nonbonded force path with MDBas semantics.

The product is ``libmdgpu.so`` (HIP kernels behind the C ABI declared in
``include/mdgpu.h``).  This module is the thin Python side of that ABI:
ctypes mirrors of the reference structs (global.h) and a ``Context`` that
forwards to the ``mdg_*`` entry points.  There is no CPU fallback: loading
fails loudly if the library is missing, and every compute call runs on
the GPU.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmdgpu.so")

MDG_FP64, MDG_MIXED = 0, 1
ERRORS = {1006: "input larger than the fixed scratch", 1001: "HIP/rocFFT error", 1002: "unsupported setting", 1003: "list capacity",
          1004: "bad argument", 1005: "SHAKE group re-diverged", 311: "SHAKE not converged",
          110: "1-4 pair without dihedral", 111: "exclusion list last atom", 412: "link cells"}

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)


class MDG_PBC(C.Structure):
    _fields_ = [("type", C.c_int)] + [(k, C.c_double) for k in (
        "a a1 a2 a3 b b1 b2 b3 c c1 c2 c3 u u1 u2 u3 v v1 v2 v3 w w1 w2 w3 "
        "pa pb pc det vol vol0 stress1 stress2 stress3 stress4 stress5 stress6 "
        "stress7 stress8 stress9").split()]


class MDG_CTRL(C.Structure):
    _fields_ = [(k, C.c_int) for k in (
        "newjob keyRand keyTraj keyProp keyForF keyRest printOut printProp printTraj printRest "
        "keyMd mdType seed keyMinim keyLink noLink keyConstH keyNb14 keyNumForce keyEwald "
        "keyAlpha keyMmax elecType vdwType integrator ens").split()]


class MDG_PARAM(C.Structure):
    _fields_ = [("step", C.c_int), ("nSteps", C.c_int), ("timeStep", C.c_double),
                ("rTimeStep", C.c_double), ("nAtom", C.c_int), ("nDegFree", C.c_int),
                ("nFrozen", C.c_int), ("nBond", C.c_int), ("nAngle", C.c_int),
                ("nDihedral", C.c_int), ("nImproper", C.c_int), ("nUb", C.c_int),
                ("nConst", C.c_int), ("maxCycle", C.c_int), ("tolShake", C.c_double)] + [
                (k, C.c_double) for k in (
                    "chargeConst delr scal14 cutOff rcutOff cutOff2 rcutOff2 cutOn cutOn2 "
                    "switch2 temp0 press0 kinTemp0 tolMinim maxminsiz maxminst").split()]


class MDG_BATH(C.Structure):
    _fields_ = [(k, C.c_double) for k in "tauT chiT tauP chiP compress".split()]


class MDG_NEIGH(C.Structure):
    _fields_ = [(k, C.c_int) for k in "update linkRatio nCells sizeList nPair14".split()]


class MDG_ENERGY(C.Structure):
    _fields_ = [(k, C.c_double) for k in (
        "tot pot kin elec vdw bond ang ub dihe impr conint consv virelec virvdw virbond virub "
        "virshake virpot virtot").split()]


class MDG_CONSTRAINT(C.Structure):
    _fields_ = [("a", C.c_int), ("b", C.c_int), ("rc2", C.c_double)]


class MDG_EWALD(C.Structure):
    _fields_ = [("nbsp", C.c_int), ("mmax", C.c_int), ("m1max", C.c_int), ("m2max", C.c_int),
                ("m3max", C.c_int), ("prec", C.c_double), ("tol", C.c_double),
                ("tol1", C.c_double), ("alpha", C.c_double)]


class MDG_ATOM(C.Structure):
    _fields_ = [("label", C.c_char * 5), ("segn", C.c_char * 5), ("resn", C.c_char * 5),
                ("type", C.c_int), ("resi", C.c_int), ("ires", C.c_int)]


class MDG_IO(C.Structure):
    _fields_ = ([(k, C.c_void_p) for k in
                 "confFile forfFile propFile restFile rconFile trajFile simuFile".split()] +
                [(k, C.c_char * 512) for k in
                 "confName forfName propName restName rconName trajName simuName".split()])


class MDG_STEP(C.Structure):
    _fields_ = [("step", C.c_int), ("listRebuilt", C.c_int)] + [(k, C.c_double) for k in (
        "pot kin tot elec vdw eDir eRec eCorr virelec virvdw virshake nInCut "
        "ebond eub eang edihe eimpr virbond virub conint consv chiT").split()]


# Every entry point include/mdgpu.h declares (checked by the ABI test).
EXPORTS = ("mdg_create mdg_destroy mdg_last_error mdg_precision mdg_key_link mdg_set_topology "
           "mdg_makelist mdg_ewald_energy mdg_spme_energy mdg_nonbond_energy mdg_energy "
           "mdg_vv_integrate mdg_upload_state mdg_download_state mdg_initial_energy "
           "mdg_run_steps mdg_set_profiling mdg_phase_times mdg_list_stats "
           "mdg_list_pairs mdg_set_nonbond_params mdg_backward mdg_backward_traj "
           "mdg_batch_create mdg_batch_size mdg_autotune mdg_set_prune mdg_prune_stats mdg_set_pair_path "
           "mdg_prune_skin mdg_get_launch mdg_set_launch mdg_set_bonded "
           "mdg_read_simu mdg_read_conf mdg_read_rest mdg_read_forf mdg_setup "
           "mdg_charge_const mdg_init_box mdg_get_kinfromtemp mdg_image_update mdg_kinetic "
           "mdg_spme_mesh mdg_init_vel mdg_write_conf mdg_write_prop mdg_write_rest "
           "mdg_write_dcd_header mdg_write_dcd_traj mdg_nonbond_energy_noalloc mdg_set_bath "
           "mdg_get_bath").split()

_LIB = None


class MdgError(RuntimeError):
    pass


def lib_path() -> str:
    return LIB_PATH


def load():
    """Load libmdgpu.so; raises if it was not built (no fallback path)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise MdgError(f"{LIB_PATH} is missing: run __graft_entry__.build() (or "
                       "python mdbas_amd/build.py) first")
    L = C.CDLL(LIB_PATH)
    L.mdg_create.restype = C.c_void_p
    L.mdg_create.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                             C.POINTER(C.c_int)]
    L.mdg_batch_create.restype = C.c_void_p
    L.mdg_batch_create.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int,
                                   C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_int)]
    L.mdg_destroy.argtypes = [C.c_void_p]
    L.mdg_prune_skin.restype = C.c_double
    L.mdg_prune_skin.argtypes = [C.c_void_p]
    L.mdg_last_error.restype = C.c_char_p
    L.mdg_last_error.argtypes = [C.c_void_p]
    for f in EXPORTS:
        fn = getattr(L, f)
        if f not in ("mdg_create", "mdg_batch_create", "mdg_destroy", "mdg_last_error",
                     "mdg_prune_skin", "mdg_charge_const", "mdg_kinetic", "mdg_init_box",
                     "mdg_get_kinfromtemp", "mdg_image_update"):
            fn.restype = C.c_int
    L.mdg_charge_const.restype = C.c_double
    L.mdg_charge_const.argtypes = [C.c_int]
    L.mdg_kinetic.restype = C.c_double
    for f in ("mdg_init_box", "mdg_get_kinfromtemp", "mdg_image_update"):
        getattr(L, f).restype = None
    dp = C.POINTER(C.c_double)
    L.mdg_backward.argtypes = [C.c_void_p, C.c_double, dp, dp, dp]
    L.mdg_backward_traj.argtypes = [C.c_void_p, C.c_int, C.c_int, dp, dp, dp, dp]
    _LIB = L
    return L


def P(a):
    if a is None:
        return None
    if a.dtype == np.float64:
        return a.ctypes.data_as(_dp)
    if a.dtype == np.int32:
        return a.ctypes.data_as(_ip)
    raise TypeError(a.dtype)


def reference_structs(wb, cfg: dict):
    """The CTRL/PARAM/PBC/EWALD/NEIGH values a reference SIMU deck for this
    workload produces (io.c:29-723, init.c:102-109, setup's consth)."""
    ctrl = MDG_CTRL()
    ctrl.newjob = 1
    ctrl.keyMd = 1
    ctrl.mdType = 1
    ctrl.keyConstH = 1
    ctrl.keyEwald = cfg.get("keyEwald", 2)   # 0: plain cutoff Coulomb of elecType
    ctrl.keyAlpha = 1
    ctrl.keyMmax = 1
    ctrl.elecType = cfg.get("elecType", 0)
    ctrl.vdwType = 2
    ctrl.integrator = 1
    ctrl.ens = 0
    p = MDG_PARAM()
    p.nAtom = wb.n
    p.timeStep = cfg["timeStep"]
    p.rTimeStep = 1.0 / cfg["timeStep"]
    p.nConst = len(wb.ca)
    p.maxCycle = cfg["maxCycle"]
    p.tolShake = cfg["tolShake"]
    p.chargeConst = 332.0716 * 418.4
    p.delr = cfg["delr"]
    p.cutOff = cfg["cutOff"]
    p.cutOn = cfg["cutOn"]
    p.rcutOff = 1.0 / p.cutOff
    p.cutOff2 = p.cutOff * p.cutOff
    p.rcutOff2 = 1.0 / p.cutOff2
    p.cutOn2 = p.cutOn * p.cutOn
    p.switch2 = 1.0 / (p.cutOff2 - p.cutOn2) ** 3
    p.temp0 = cfg.get("temp", 300.0)
    p.nDegFree = 3 * wb.n - 3 - len(wb.ca)
    b = MDG_PBC()
    b.type = 1
    b.a1 = b.b2 = b.c3 = float(wb.L)
    e = MDG_EWALD()
    e.nbsp = cfg["nbsp"]
    e.m1max = e.m2max = e.m3max = cfg["grid"]
    e.mmax = cfg["grid"] ** 3
    e.alpha = cfg["alpha"]
    e.prec = 1e-6
    ng = MDG_NEIGH()
    ng.update = cfg["update"]
    ng.linkRatio = cfg["linkRatio"]
    if hasattr(wb, "bonds"):          # a molecule box with bonded terms (molecules.py)
        ctrl.keyNb14 = 1
        p.nBond, p.nUb, p.nAngle = len(wb.bonds), len(wb.ubs), len(wb.angles)
        p.nDihedral, p.nImproper = len(wb.dihes), len(wb.imprs)
        p.scal14 = wb.scal14
    return ctrl, p, b, e, ng


class Context:
    """One device context (one HIP stream) for one water box."""

    def __init__(self, wb, cfg: dict | None = None, precision: int = MDG_FP64):
        self.L = load()
        self.wb = wb
        self.cfg = dict(wb.params) if cfg is None else cfg
        self.ctrl, self.param, self.pbc, self.ewald, self.neigh = reference_structs(wb, self.cfg)
        st = C.c_int(0)
        h = self.L.mdg_create(C.byref(self.ctrl), C.byref(self.param), C.byref(self.pbc),
                              C.byref(self.ewald), C.byref(self.neigh), int(precision),
                              C.byref(st))
        if not h:
            raise MdgError(f"mdg_create failed: {st.value} {ERRORS.get(st.value, '')}")
        self.h = C.c_void_p(h)
        self.precision = precision
        n = wb.n
        self.cons = (MDG_CONSTRAINT * max(1, len(wb.ca)))()
        for k in range(len(wb.ca)):
            self.cons[k].a, self.cons[k].b, self.cons[k].rc2 = int(wb.ca[k]), int(wb.cb[k]), float(wb.rc2[k])
        self._chk(self.L.mdg_set_topology(self.h, P(wb.q), P(wb.eps), P(wb.sig), P(wb.mass),
                                          P(wb.rmass), P(wb.frozen), P(wb.nAtConst), self.cons,
                                          len(wb.ca)), "mdg_set_topology")
        self.n = n
        self._lists = None
        if hasattr(wb, "bonds"):
            t = self._terms()
            self._chk(self.L.mdg_set_bonded(self.h, C.byref(self.ctrl), C.byref(self.param), *t,
                                            P(wb.eps14), P(wb.sig14)), "mdg_set_bonded")

    def _chk(self, rc, what):
        if rc != 0:
            msg = self.L.mdg_last_error(self.h).decode() if getattr(self, "h", None) else ""
            raise MdgError(f"{what} -> {rc} ({ERRORS.get(rc, '')}): {msg}")

    def close(self):
        if getattr(self, "h", None):
            self.L.mdg_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def key_link(self) -> int:
        return self.L.mdg_key_link(self.h)

    # ---- drop-in mirrors -------------------------------------------------
    def makelist(self, x, y, z, step: int = 0):
        """mdg_makelist; returns (neighList, neighPair, neighOrder, exclOff,
        exclLst) as numpy arrays (copies of the malloc'd reference-format
        arrays, which are then freed)."""
        if self._lists is None:
            self._lists = dict(nl=C.POINTER(C.c_int)(), np=C.POINTER(C.c_int)(),
                               no=C.POINTER(C.c_int)(), n14=C.POINTER(C.c_int)(),
                               el=C.POINTER(C.POINTER(C.c_int))(), ep=C.POINTER(C.c_int)())
        d = self._lists
        self.param.step = step
        t = self._terms()
        self._chk(self.L.mdg_makelist(self.h, C.byref(self.ctrl), C.byref(self.param),
                                      C.byref(self.pbc), C.byref(self.neigh), self.cons, t[0],
                                      t[2], t[3], t[4], P(x), P(y), P(z), P(self.wb.frozen),
                                      C.byref(d["nl"]), C.byref(d["np"]), C.byref(d["no"]),
                                      C.byref(d["n14"]), C.byref(d["el"]), C.byref(d["ep"])),
                  "mdg_makelist")
        n = self.n
        np_ = np.ctypeslib.as_array(d["np"], shape=(n,)).copy()
        no = np.ctypeslib.as_array(d["no"], shape=(n,)).copy()
        tot = int(np_.sum())
        nl = np.ctypeslib.as_array(d["nl"], shape=(max(tot, 1),))[:tot].copy()
        ep = np.ctypeslib.as_array(d["ep"], shape=(n,)).copy()
        rows = n if self.key_link else n - 1
        off = np.zeros(n + 1, np.int32)
        lst = []
        for i in range(n):
            off[i + 1] = off[i] + ep[i]
            if i < rows and ep[i]:
                lst.extend(d["el"][i][k] for k in range(ep[i]))
        self._excl = (d["el"], d["ep"])
        n14 = self.neigh.nPair14
        self.list14 = np.ctypeslib.as_array(d["n14"], shape=(max(1, 2 * n14),))[:2 * n14].copy()
        return nl, np_, no, off, np.array(lst, np.int32)

    def _terms(self):
        """The bonded term tables (bond, ub, angle, dihe, impr) as ctypes
        arrays, or Nones for a water box."""
        if not hasattr(self.wb, "bonds"):
            return (None,) * 5
        if not hasattr(self, "_term_arrays"):
            wb = self.wb
            self._term_arrays = (wb.c_bonds(), wb.c_ubs(), wb.c_angles(), wb.c_dihes(),
                                 wb.c_imprs())
        return self._term_arrays

    def energy(self, x, y, z, nl, np_, no, elec_type=None, keyEwald=None):
        """The drop-in energy() (mdg_energy, energy.c:103-205) over a
        reference-format list: zeroed forces, Ewald (or plain cutoff) pairs
        with their 1-4 pairs, every bonded term.  Returns forces (3,N) and
        the ENERGY struct."""
        n = self.n
        f = np.zeros((3, n))
        v = np.zeros((3, n))
        e = MDG_ENERGY()
        el, ep = self._excl
        t = self._terms()
        wb = self.wb
        saved = (self.ctrl.elecType, self.ctrl.keyEwald)
        if elec_type is not None:
            self.ctrl.elecType = elec_type
        if keyEwald is not None:
            self.ctrl.keyEwald = keyEwald
        eps14 = getattr(wb, "eps14", None)
        sig14 = getattr(wb, "sig14", None)
        l14 = getattr(self, "list14", np.zeros(0, np.int32))
        try:
            self._chk(self.L.mdg_energy(self.h, C.byref(self.ctrl), C.byref(self.param), C.byref(e),
                                        C.byref(self.ewald), C.byref(self.pbc), C.byref(self.neigh),
                                        *t, P(x), P(y), P(z), P(v[0]), P(v[1]), P(v[2]), P(f[0]),
                                        P(f[1]), P(f[2]), P(wb.q), P(wb.eps), P(wb.sig),
                                        P(eps14), P(sig14), P(wb.frozen), P(nl), P(np_), P(no),
                                        P(l14) if len(l14) else None, el, ep), "mdg_energy")
        finally:
            self.ctrl.elecType, self.ctrl.keyEwald = saved
        return f, e

    def ewald_energy(self, x, y, z, nl, np_, no, f=None):
        """mdg_ewald_energy over a reference-format list; returns forces (3,N)
        and the accumulated ENERGY struct."""
        n = self.n
        f = np.zeros((3, n)) if f is None else f
        e = MDG_ENERGY()
        el, ep = self._excl
        self._chk(self.L.mdg_ewald_energy(self.h, C.byref(self.ctrl), C.byref(self.param),
                                          C.byref(e), C.byref(self.ewald), C.byref(self.pbc),
                                          P(x), P(y), P(z), P(f[0]), P(f[1]), P(f[2]),
                                          P(self.wb.q), P(self.wb.eps), P(self.wb.sig), P(nl),
                                          P(np_), P(no), el, ep), "mdg_ewald_energy")
        return f, e

    def nonbond_energy(self, elec_type, x, y, z, nl, np_, no):
        n = self.n
        f = np.zeros((3, n))
        e = MDG_ENERGY()
        saved = self.ctrl.elecType
        self.ctrl.elecType = elec_type
        try:
            self._chk(self.L.mdg_nonbond_energy(self.h, C.byref(self.ctrl), C.byref(self.param),
                                                C.byref(e), C.byref(self.pbc), P(x), P(y), P(z),
                                                P(f[0]), P(f[1]), P(f[2]), P(self.wb.q),
                                                P(self.wb.eps), P(self.wb.sig), P(nl), P(np_),
                                                P(no)), "mdg_nonbond_energy")
        finally:
            self.ctrl.elecType = saved
        return f, e

    def spme_energy(self, x, y, z, f=None):
        n = self.n
        f = np.zeros((3, n)) if f is None else f
        st = np.zeros(6)
        vir = C.c_double(0)
        en = C.c_double(0)
        self._chk(self.L.mdg_spme_energy(self.h, C.byref(self.param), C.byref(self.ewald),
                                         C.byref(self.pbc), P(x), P(y), P(z), P(f[0]), P(f[1]),
                                         P(f[2]), P(self.wb.q), P(st), C.byref(vir),
                                         C.byref(en)), "mdg_spme_energy")
        return en.value, f, st, vir.value

    def vv_integrate(self, stage, x, y, z, vx, vy, vz, f):
        e = MDG_ENERGY()
        nat = self.wb.nAtConst
        self._chk(self.L.mdg_vv_integrate(self.h, C.byref(self.ctrl), C.byref(self.param),
                                          C.byref(e), C.byref(self.pbc), None, self.cons, P(x),
                                          P(y), P(z), P(vx), P(vy), P(vz), P(f[0]), P(f[1]),
                                          P(f[2]), P(self.wb.mass), P(self.wb.rmass), P(nat),
                                          int(stage)), "mdg_vv_integrate")
        return e

    def vv_integrate_ens(self, st, stage, x, y, z, vx, vy, vz, f):
        """vv_integrate with a thermostat/barostat (st["ens"] 1..4): st is
        the state dict of oracle.ens_state (bath, targets, the ENERGY
        members, box vectors, nine stresses), updated in place."""
        b9 = [float(v) for v in st["box9"]]
        pbc = self.pbc
        (pbc.a1, pbc.a2, pbc.a3, pbc.b1, pbc.b2, pbc.b3, pbc.c1, pbc.c2, pbc.c3) = b9
        self.L.mdg_init_box(C.byref(pbc))
        for k in range(9):
            setattr(pbc, f"stress{k + 1}", float(st["stress"][k]))
        bath = MDG_BATH(tauT=st["tauT"], chiT=st["chiT"], tauP=st["tauP"], chiP=st["chiP"])
        e = MDG_ENERGY(virpot=st["virpot"], conint=st["conint"], consv=st["consv"], kin=st["kin"],
                       virshake=st["virshake"])
        self.param.kinTemp0, self.param.temp0 = st["kinTemp0"], st["temp0"]
        self.param.press0 = st["press0"]
        self.ctrl.ens = int(st["ens"])
        try:
            self._chk(self.L.mdg_vv_integrate(self.h, C.byref(self.ctrl), C.byref(self.param),
                                              C.byref(e), C.byref(pbc), C.byref(bath), self.cons,
                                              P(x), P(y), P(z), P(vx), P(vy), P(vz), P(f[0]),
                                              P(f[1]), P(f[2]), P(self.wb.mass), P(self.wb.rmass),
                                              P(self.wb.nAtConst), int(stage)),
                      "mdg_vv_integrate")
        finally:
            self.ctrl.ens = 0
        st["chiT"], st["chiP"] = bath.chiT, bath.chiP
        for k in ("conint", "consv", "kin", "virshake"):
            st[k] = getattr(e, k)
        st["box9"] = np.array([pbc.a1, pbc.a2, pbc.a3, pbc.b1, pbc.b2, pbc.b3, pbc.c1, pbc.c2,
                               pbc.c3])
        st["stress"] = np.array([getattr(pbc, f"stress{k + 1}") for k in range(9)])
        return e

    # ---- device-resident ------------------------------------------------
    def upload(self, x, y, z, vx, vy, vz):
        self._chk(self.L.mdg_upload_state(self.h, P(x), P(y), P(z), P(vx), P(vy), P(vz)),
                  "mdg_upload_state")

    def download(self):
        n = self.n
        a = [np.zeros(n) for _ in range(9)]
        self._chk(self.L.mdg_download_state(self.h, *[P(v) for v in a]), "mdg_download_state")
        return a

    def initial_energy(self) -> MDG_STEP:
        r = MDG_STEP()
        self._chk(self.L.mdg_initial_energy(self.h, C.byref(r)), "mdg_initial_energy")
        return r

    def set_bath(self, ens: int, tauT: float = 0.1, chiT: float = 0.0, kinTemp0: float = 0.0,
                 conint: float = 0.0, consv: float = 0.0):
        """mdg_set_bath: the device step's thermostat (0 NVE, 1 NVT_B, 3
        NVT_H)."""
        ctrl = MDG_CTRL()
        C.pointer(ctrl)[0] = self.ctrl
        ctrl.ens = int(ens)
        param = MDG_PARAM()
        C.pointer(param)[0] = self.param
        param.kinTemp0 = float(kinTemp0)
        bath = MDG_BATH(tauT=float(tauT), chiT=float(chiT))
        e = MDG_ENERGY(conint=float(conint), consv=float(consv))
        self._chk(self.L.mdg_set_bath(self.h, C.byref(ctrl), C.byref(param), C.byref(bath),
                                      C.byref(e)), "mdg_set_bath")

    def get_bath(self):
        bath, e = MDG_BATH(), MDG_ENERGY()
        self._chk(self.L.mdg_get_bath(self.h, C.byref(bath), C.byref(e)), "mdg_get_bath")
        return dict(chiT=bath.chiT, conint=e.conint, consv=e.consv)

    def run(self, nsteps: int, use_graph: bool = True):
        recs = (MDG_STEP * max(1, nsteps))()
        self._chk(self.L.mdg_run_steps(self.h, int(nsteps), int(use_graph), recs),
                  "mdg_run_steps")
        return recs

    def autotune(self, nsteps: int = 60):
        """Times candidate launch shapes on this system and keeps the best;
        returns them (the uploaded state is restored)."""
        out = (C.c_int * 2)()
        self._chk(self.L.mdg_autotune(self.h, int(nsteps), out), "mdg_autotune")
        return dict(self.get_launch(), prune_skin=float(self.L.mdg_prune_skin(self.h)))

    # ---- adjoint --------------------------------------------------------
    def set_params(self, q, eps, sig):
        """New per-atom charges and LJ eps / sig (FP64 arrays of length n)."""
        q, eps, sig = (np.ascontiguousarray(a, np.float64) for a in (q, eps, sig))
        self._chk(self.L.mdg_set_nonbond_params(self.h, P(q), P(eps), P(sig)),
                  "mdg_set_nonbond_params")

    def backward(self, dL_dE: float = 1.0):
        """dL/d(q, eps, sig) per atom at the current device state."""
        g = np.zeros((3, self.n))
        self._chk(self.L.mdg_backward(self.h, float(dL_dE), P(g[0]), P(g[1]), P(g[2])),
                  "mdg_backward")
        return g

    def backward_traj(self, nsteps: int, ckpt: int):
        """(J, dJ/d(q, eps, sig)) for J = mean potential energy over the next
        nsteps steps (the device state advances by nsteps)."""
        g = np.zeros((3, self.n))
        J = C.c_double(0)
        self._chk(self.L.mdg_backward_traj(self.h, int(nsteps), int(ckpt), C.byref(J), P(g[0]),
                                           P(g[1]), P(g[2])), "mdg_backward_traj")
        return J.value, g

    def set_profiling(self, on: bool):
        self._chk(self.L.mdg_set_profiling(self.h, int(bool(on))), "mdg_set_profiling")

    def phase_times(self):
        out = np.zeros(5)
        self._chk(self.L.mdg_phase_times(self.h, P(out)), "mdg_phase_times")
        return dict(pair_ms=out[0], spme_ms=out[1], integrate_ms=out[2], rebuild_ms=out[3],
                    step_ms=out[4])

    def get_launch(self):
        out = (C.c_int * 4)()
        self._chk(self.L.mdg_get_launch(self.h, out), "mdg_get_launch")
        return dict(shake_spread=out[0], pair_blocks=out[1], spme_first=out[2], spme_prio=out[3])

    def set_launch(self, shake_spread: int, pair_blocks: int, spme_first: int = 0,
                   spme_prio: int = 0):
        """Launch shapes: SHAKE spread (0: a molecule per lane; 1, 2, 4, 8, 16: a
        molecule per quad of lanes on every spread-th quad), the pair kernel's
        grid cap, SPME branch first at the fork (0/1), wave priority (0: none, 1: the
        binned SPME kernels raise theirs, 2: the mixed pair kernel raises its own)."""
        v = (C.c_int * 4)(shake_spread, pair_blocks, spme_first, spme_prio)
        self._chk(self.L.mdg_set_launch(self.h, v), "mdg_set_launch")

    def set_prune(self, skin: float):
        """Pruning distance beyond the cutoff for the mixed pair kernel's
        shorter per-step list (0 switches it off)."""
        self._chk(self.L.mdg_set_prune(self.h, C.c_double(skin)), "mdg_set_prune")

    def set_pair_path(self, path: int):
        """Pair displacements of the mixed kernel: 0 relative coordinates
        where the box allows (default), 1 FP64 from the positions, 2 relative
        with every entry's cutoff test in FP64."""
        self._chk(self.L.mdg_set_pair_path(self.h, int(path)), "mdg_set_pair_path")

    def prune_stats(self):
        out = (C.c_long * 2)()
        self._chk(self.L.mdg_prune_stats(self.h, out), "mdg_prune_stats")
        return dict(pruned_launches=out[0], pruned_entries=out[1],
                    skin=float(self.L.mdg_prune_skin(self.h)))

    def list_stats(self):
        out = (C.c_long * 3)()
        self._chk(self.L.mdg_list_stats(self.h, out), "mdg_list_stats")
        return dict(cluster_pairs=out[0], atom_pairs=out[1], clusters=out[2])

    def list_pairs(self):
        cnt = C.c_long(0)
        self._chk(self.L.mdg_list_pairs(self.h, None, None, C.c_long(0), C.byref(cnt)),
                  "mdg_list_pairs")
        k = cnt.value
        pi = np.zeros(max(k, 1), np.int32)
        pj = np.zeros(max(k, 1), np.int32)
        self._chk(self.L.mdg_list_pairs(self.h, P(pi), P(pj), C.c_long(k), C.byref(cnt)),
                  "mdg_list_pairs")
        return pi[:k], pj[:k]


class BatchContext(Context):
    """Many independent water boxes (each small enough for the O(N^2)
    list) advanced together: one context, every kernel covering all
    replicas.  Arrays are the replicas' concatenated; records come per step
    and replica."""

    def __init__(self, boxes, precision: int = MDG_MIXED):
        self.L = load()
        self.boxes = boxes
        self.R = len(boxes)
        cfg0 = dict(boxes[0].params)
        self.cfg = cfg0
        self.ctrl, self.param, _, self.ewald, self.neigh = reference_structs(boxes[0], cfg0)
        pbcs = (MDG_PBC * self.R)()
        for k, b in enumerate(boxes):
            pbcs[k].type = 1
            pbcs[k].a1 = pbcs[k].b2 = pbcs[k].c3 = float(b.L)
        nat = np.array([b.n for b in boxes], np.int32)
        mesh = np.array([b.params["grid"] for b in boxes], np.int32)
        st = C.c_int(0)
        h = self.L.mdg_batch_create(C.byref(self.ctrl), C.byref(self.param), C.byref(self.ewald),
                                    C.byref(self.neigh), self.R, nat.ctypes.data, pbcs,
                                    mesh.ctypes.data, int(precision), C.byref(st))
        if not h:
            raise MdgError(f"mdg_batch_create failed: {st.value} {ERRORS.get(st.value, '')}")
        self.h = C.c_void_p(h)
        self.precision = precision
        self.aoff = np.concatenate([[0], np.cumsum(nat)]).astype(np.int64)
        self.n = int(self.aoff[-1])
        cat = lambda k: np.ascontiguousarray(np.concatenate([getattr(b, k) for b in boxes]))  # noqa: E731
        self.q, self.eps, self.sig = cat("q"), cat("eps"), cat("sig")
        self.mass, self.rmass = cat("mass"), cat("rmass")
        self.frozen, self.nAtConst = cat("frozen"), cat("nAtConst")
        ca = np.concatenate([b.ca + o for b, o in zip(boxes, self.aoff[:-1])])
        cb = np.concatenate([b.cb + o for b, o in zip(boxes, self.aoff[:-1])])
        rc2 = np.concatenate([b.rc2 for b in boxes])
        self.cons = (MDG_CONSTRAINT * max(1, len(ca)))()
        for k in range(len(ca)):
            self.cons[k].a, self.cons[k].b, self.cons[k].rc2 = int(ca[k]), int(cb[k]), float(rc2[k])
        self._chk(self.L.mdg_set_topology(self.h, P(self.q), P(self.eps), P(self.sig),
                                          P(self.mass), P(self.rmass), P(self.frozen),
                                          P(self.nAtConst), self.cons, len(ca)),
                  "mdg_set_topology")
        self.wb = None
        self._lists = None

    def upload_boxes(self):
        cat = lambda k: np.ascontiguousarray(np.concatenate([getattr(b, k) for b in self.boxes]))  # noqa: E731
        self.upload(*[cat(k) for k in ("x", "y", "z", "vx", "vy", "vz")])

    def split(self, a):
        """Per-replica views of a concatenated per-atom array."""
        return [a[self.aoff[r]:self.aoff[r + 1]] for r in range(self.R)]

    def initial_energy(self):
        recs = (MDG_STEP * self.R)()
        self._chk(self.L.mdg_initial_energy(self.h, recs), "mdg_initial_energy")
        return recs

    def run(self, nsteps: int, use_graph: bool = True):
        """Records as an (nsteps, R) array of MDG_STEP."""
        recs = (MDG_STEP * max(1, nsteps * self.R))()
        self._chk(self.L.mdg_run_steps(self.h, int(nsteps), int(use_graph), recs), "mdg_run_steps")
        return recs


def batch_records(recs, R: int):
    """(steps, R) arrays of every MDG_STEP field."""
    d = records_to_dict(recs)
    return {k: v.reshape(-1, R) for k, v in d.items()}


def records_to_dict(recs):
    keys = [f[0] for f in MDG_STEP._fields_]
    return {k: np.array([getattr(r, k) for r in recs]) for k in keys}


# ---- reference decks (host-side readers of libmdgpu) ------------------------

class _InDir:
    def __init__(self, d):
        self.d = d

    def __enter__(self):
        self.old = os.getcwd()
        os.chdir(self.d)

    def __exit__(self, *a):
        os.chdir(self.old)


def read_deck(dirpath: str, simu: str = "SIMU") -> dict:
    """init_system's reading chain (init.c:30-172) through libmdgpu's host
    readers, in dirpath: mdg_read_simu, the Coulomb constant,
    mdg_read_rest (under "restart") or mdg_read_conf, mdg_read_forf,
    mdg_setup, mdg_get_kinfromtemp, mdg_init_box, mdg_image_update.
    Returns the structs and numpy copies of every array."""
    L = load()
    io, ctrl, param, bath = MDG_IO(), MDG_CTRL(), MDG_PARAM(), MDG_BATH()
    neigh, ewald, box, ener = MDG_NEIGH(), MDG_EWALD(), MDG_PBC(), MDG_ENERGY()
    io.simuName = simu.encode()
    dpp = C.POINTER(C.c_double)
    with _InDir(dirpath):
        def chk(rc, what):
            if rc:
                raise MdgError(f"{what} -> {rc}")
        chk(L.mdg_read_simu(C.byref(io), C.byref(ctrl), C.byref(param), C.byref(bath),
                            C.byref(neigh), C.byref(ewald), C.byref(box)), "mdg_read_simu")
        param.chargeConst = L.mdg_charge_const(ctrl.mdType)
        atom = C.POINTER(MDG_ATOM)()
        xyz = [dpp() for _ in range(9)]
        if ctrl.keyRest:
            chk(L.mdg_read_rest(C.byref(io), C.byref(param), C.byref(ener), C.byref(bath),
                                C.byref(atom), *[C.byref(a) for a in xyz]), "mdg_read_rest")
        else:
            chk(L.mdg_read_conf(C.byref(io), C.byref(param), C.byref(atom),
                                *[C.byref(a) for a in xyz[:3]]), "mdg_read_conf")
        n = param.nAtom
        arr = {k: np.zeros(n) for k in ("q", "mass", "rmass", "eps", "sig", "eps14", "sig14")}
        frozen = np.zeros(n, np.int32)
        nat = np.zeros(n, np.int32)
        cons = C.POINTER(MDG_CONSTRAINT)()
        from .molecules import MdgAngle, MdgBond, MdgDihe
        bond, ub = C.POINTER(MdgBond)(), C.POINTER(MdgBond)()
        angle = C.POINTER(MdgAngle)()
        dihe, impr = C.POINTER(MdgDihe)(), C.POINTER(MdgDihe)()
        chk(L.mdg_read_forf(C.byref(io), C.byref(param), atom, C.byref(cons), C.byref(bond),
                            C.byref(angle), C.byref(dihe), C.byref(impr), C.byref(ub),
                            P(arr["eps"]), P(arr["sig"]), P(arr["eps14"]), P(arr["sig14"]),
                            P(arr["mass"]), P(arr["q"]), P(frozen), P(nat)), "mdg_read_forf")
        chk(L.mdg_setup(C.byref(ctrl), C.byref(param), atom, C.byref(cons), C.byref(bond),
                        C.byref(angle), C.byref(dihe), C.byref(impr), C.byref(ub),
                        P(arr["mass"]), P(arr["rmass"]), P(frozen), P(nat)), "mdg_setup")
        L.mdg_get_kinfromtemp(C.byref(param), C.byref(box))
        L.mdg_init_box(C.byref(box))
        pos = [np.ctypeslib.as_array(a, shape=(n,)).copy() if a else np.zeros(n) for a in xyz]
        L.mdg_image_update(C.byref(param), C.byref(box), P(pos[0]), P(pos[1]), P(pos[2]))
    atoms = [MDG_ATOM.from_buffer_copy(atom[i]) for i in range(n)]
    atom_bytes = C.string_at(atom, n * C.sizeof(MDG_ATOM))

    def rows(ptr, count, T):
        return [T.from_buffer_copy(ptr[k]) for k in range(count)]
    out = dict(io=io, ctrl=ctrl, param=param, bath=bath, neigh=neigh, ewald=ewald, box=box,
               ener=ener, atoms=atoms, atom_bytes=atom_bytes, frozen=frozen, nAtConst=nat,
               cons=rows(cons, param.nConst, MDG_CONSTRAINT), bond=rows(bond, param.nBond, MdgBond),
               ub=rows(ub, param.nUb, MdgBond), angle=rows(angle, param.nAngle, MdgAngle),
               dihe=rows(dihe, param.nDihedral, MdgDihe), impr=rows(impr, param.nImproper, MdgDihe))
    for k, a in zip(("x", "y", "z", "vx", "vy", "vz", "fx", "fy", "fz"), pos):
        out[k] = a
    out.update(arr)
    libc = C.CDLL(None)
    libc.free.argtypes = [C.c_void_p]
    for p_ in [atom, cons, bond, ub, angle, dihe, impr] + xyz:
        if p_:
            libc.free(C.cast(p_, C.c_void_p))
    return out
