"""Times the FP64 oracle against the compiled reference (oracle/_ref) on the
same 3,000-atom water deck, for the parts both have: the O(N^2) Verlet list
build, the SWITCH-Coulomb + switched-LJ pair loop of nonbond_energy, and
the two velocity Verlet + SHAKE stages.  One pinned host core each, best of
a few repetitions.  The ratio shows the oracle that bench.py times as the
CPU baseline is no slower than the reference itself.

Needs the reference build (python -c "import oracle; oracle.build()"), so
it runs where the reference tree is present.

  python oracle/calibrate.py
"""
from __future__ import annotations

import ctypes as C
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from mdbas_amd import configs  # noqa: E402


def best(fn, reps=3):
    t = []
    for _ in range(reps):
        t0 = time.perf_counter()
        fn()
        t.append(time.perf_counter() - t0)
    return min(t)


def main():
    try:
        os.sched_setaffinity(0, {sorted(os.sched_getaffinity(0))[0]})
    except (AttributeError, OSError):
        pass
    if not O.ref_available():
        print("reference build not available (oracle/_ref)")
        return 1
    wb = configs.build(configs.CONFIGS["c0"])
    o = O.OracleSystem(wb)
    r = O.RefSystem(wb)
    x0, y0, z0 = wb.x.copy(), wb.y.copy(), wb.z.copy()
    rows = []

    # list build (O(N^2) Verlet list at c0)
    rows.append(("list build", best(lambda: o.build_list(x0.copy(), y0.copy(), z0.copy())),
                 best(lambda: r.makelist(x0.copy(), y0.copy(), z0.copy()))))

    # pair loop: nonbond_energy with SWITCH Coulomb (elecType 4) over the same list
    o.build_list(x0.copy(), y0.copy(), z0.copy())
    r.makelist(x0.copy(), y0.copy(), z0.copy())
    L = O.lib()
    eE, eV, vE, vV = (C.c_double(0) for _ in range(4))

    def oracle_pairs():
        f = np.zeros((3, wb.n))
        L.oc_nonbond(C.byref(o.p), C.byref(o.b), 4, O.P(x0), O.P(y0), O.P(z0), O.P(f[0]),
                     O.P(f[1]), O.P(f[2]), O.P(wb.q), O.P(wb.eps), O.P(wb.sig),
                     O.P(o.neighList), O.P(o.neighPair), O.P(o.neighOrder), C.byref(eE),
                     C.byref(eV), C.byref(vE), C.byref(vV))

    rows.append(("pair loop (nonbond_energy, SWITCH)", best(oracle_pairs),
                 best(lambda: r.nonbond(4, x0, y0, z0))))

    # VV stages 1 + 2 with SHAKE, on the same forces
    f = np.zeros((3, wb.n))
    L.oc_nonbond(C.byref(o.p), C.byref(o.b), 4, O.P(x0), O.P(y0), O.P(z0), O.P(f[0]), O.P(f[1]),
                 O.P(f[2]), O.P(wb.q), O.P(wb.eps), O.P(wb.sig), O.P(o.neighList),
                 O.P(o.neighPair), O.P(o.neighOrder), C.byref(eE), C.byref(eV), C.byref(vE),
                 C.byref(vV))

    def vv_pair(obj):
        def run():
            s = [a.copy() for a in (x0, y0, z0, wb.vx, wb.vy, wb.vz)]
            obj.vv(1, *s, f)
            obj.vv(2, *s, f)
        return run

    rows.append(("VV stages 1 + 2 with SHAKE", best(vv_pair(o)), best(vv_pair(r))))

    print(f"c0 deck: {wb.n} atoms, one pinned core ({os.uname().machine})")
    print(f"{'part':40s} {'oracle s':>10s} {'reference s':>12s} {'ratio':>7s}")
    to = tr = 0.
    for name, a, b in rows:
        to += a
        tr += b
        print(f"{name:40s} {a:10.4f} {b:12.4f} {a / b:7.3f}")
    print(f"{'total':40s} {to:10.4f} {tr:12.4f} {to / tr:7.3f}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
