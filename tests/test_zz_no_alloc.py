"""mdg_nonbond_energy_noalloc (no counterpart in the reference: the
computation of mdg_nonbond_energy for callers that cannot allocate device
memory) against mdg_nonbond_energy on the same inputs, through the C ABI:
two calls in a row, the second on a smaller system, each to the FP64 bar
(1e-12).  The test allocates the device arrays itself through the HIP
runtime libmdgpu links (the one process-wide runtime; no second one is
loaded); the no-alloc call allocates nothing."""
import ctypes as C
import os

import numpy as np
import pytest

import mdbas_amd as M
import oracle as O
from mdbas_amd import configs

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def systems():
    c0 = configs.CONFIGS["c0"]
    L = (216 / 0.0334) ** (1 / 3)
    small = dict(configs.COMMON, name="tip3p_216", n_waters=216, L=L, seed=11, grid=20,
                 cutOff=9.0, cutOn=7.5, alpha=0.35, timeStep=0.002, precision="fp64")
    return [(c0, 4), (small, 2)]          # the second call is smaller


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime libmdgpu.so links."""

    def __init__(self):
        M.load()
        self.h = C.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib",
                                     "libamdhip64.so.7"))
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipFree.argtypes = [C.c_void_p]
        self.h.hipDeviceSynchronize.argtypes = []
        self.live = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), a.nbytes) == 0
        assert self.h.hipMemcpy(p, a.ctypes.data, a.nbytes, 1) == 0     # host to device
        self.live.append(p)
        return p

    def get(self, p, like):
        out = np.empty_like(like)
        assert self.h.hipMemcpy(out.ctypes.data, p, out.nbytes, 2) == 0  # device to host
        return out

    def sync(self):
        assert self.h.hipDeviceSynchronize() == 0

    def free(self):
        for p in self.live:
            self.h.hipFree(p)
        self.live = []


def test_noalloc_matches_nonbond_energy():
    hip = Hip()
    L = M.load()
    for cfg, elec in systems():
        wb = configs.build(cfg)
        print("no-alloc", cfg["name"], wb.n, "elecType", elec, flush=True)
        o = O.OracleSystem(wb)
        x, y, z = wb.x.copy(), wb.y.copy(), wb.z.copy()
        nl, np_, no = o.build_list(x, y, z)
        ctx = M.Context(wb)
        f_ref, e_ref = ctx.nonbond_energy(elec, x, y, z, nl, np_, no)
        ctrl, param, pbc = ctx.ctrl, ctx.param, ctx.pbc
        ctrl.elecType = elec
        ctx.close()
        xd, yd, zd, qd, ed, sd = (hip.put(a) for a in (x, y, z, wb.q, wb.eps, wb.sig))
        zero = np.zeros(wb.n)
        fxd, fyd, fzd = hip.put(zero), hip.put(zero), hip.put(zero)
        nld, npd, nod = hip.put(nl), hip.put(np_), hip.put(no)
        en = hip.put(np.zeros(4))
        hip.sync()
        rc = L.mdg_nonbond_energy_noalloc(C.byref(ctrl), C.byref(param), C.byref(pbc), xd, yd,
                                          zd, fxd, fyd, fzd, qd, ed, sd, nld, npd, nod, en, None)
        hip.sync()
        assert rc == 0
        f = np.stack([hip.get(p, zero) for p in (fxd, fyd, fzd)])
        e = hip.get(en, np.zeros(4))
        hip.free()
        print("  elec", e[0], e_ref.elec, "vdw", e[1], e_ref.vdw, "force", relerr(f, f_ref),
              flush=True)
        assert e[0] == pytest.approx(e_ref.elec, rel=1e-12)
        assert e[1] == pytest.approx(e_ref.vdw, rel=1e-12)
        assert e[2] == pytest.approx(e_ref.virelec, rel=1e-12)
        assert e[3] == pytest.approx(e_ref.virvdw, rel=1e-12)
        assert relerr(f, f_ref) < 1e-12


def test_noalloc_rejects_oversized_input():
    L = M.load()
    ctrl, param, pbc = M.MDG_CTRL(), M.MDG_PARAM(), M.MDG_PBC()
    param.nAtom = 262144 + 1
    pbc.type = 1
    d = C.c_void_p(1)     # never dereferenced: the size check comes first
    rc = L.mdg_nonbond_energy_noalloc(C.byref(ctrl), C.byref(param), C.byref(pbc), *([d] * 12),
                                      d, None)
    assert rc == 1006
