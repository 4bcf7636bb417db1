"""Thermostats and barostats (vv_nvt_b, vv_npt_b, vv_nvt_h, vv_npt_h,
integrate.c:1636-2618).

CPU: the oracle's restatement (oc_vv_ens) against the compiled reference
(oracle/_ref, integrate.c built with zero-initialised locals, the value the
restatement gives the NPT integrators' virshake where they read it before
SHAKE has set it), with and without constraints, and against closed forms:
Berendsen at tauT = dt lands the kinetic energy on its target in one step,
and the Nose-Hoover thermostat with a zero friction and a kinetic energy on
target is plain velocity Verlet.

GPU: the drop-in mdg_vv_integrate for each ensemble against the oracle,
stage by stage over several steps with the same forces."""
import copy

import numpy as np
import pytest

import oracle as O
from mdbas_amd import configs

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="reference tree not built here")
RBOLTZUI = 0.83144621   # global.h:47
ENSEMBLES = {1: "NVT_B", 2: "NPT_B", 3: "NVT_H", 4: "NPT_H"}


def unconstrained(wb):
    """The same atoms with no constraints (nAtConst zero)."""
    w = copy.copy(wb)
    w.ca = np.zeros(0, np.int32)
    w.cb = np.zeros(0, np.int32)
    w.rc2 = np.zeros(0)
    w.nAtConst = np.zeros(wb.n, np.int32)
    return w


def start_state(wb, ens, f_scale=1.0):
    """A bath a little off its targets so every term moves: temperature
    target 10% above the box's, chiT/chiP non-zero, virpot of a liquid."""
    m = wb.mass
    kin = 0.5 * float((m * (wb.vx ** 2 + wb.vy ** 2 + wb.vz ** 2)).sum())
    return O.ens_state(ens, wb.box9, tauT=0.1, tauP=0.5, chiT=0.3, chiP=-0.2,
                       kinTemp0=1.1 * kin, temp0=300., press0=0.0163,
                       virpot=-0.8 * kin * f_scale)


@pytest.fixture(scope="module")
def c0():
    wb = configs.build(configs.CONFIGS["c0"])
    o = O.OracleSystem(wb)
    x, y, z = wb.x.copy(), wb.y.copy(), wb.z.copy()
    o.build_list(x, y, z)
    f, _ = o.energy(x, y, z)
    return wb, f


def run_pair(sys_a, sys_b, wb, f, ens, steps):
    """The same ensemble through two implementations, stage by stage;
    yields after every stage."""
    sa = [wb.x.copy(), wb.y.copy(), wb.z.copy(), wb.vx.copy(), wb.vy.copy(), wb.vz.copy()]
    sb = [a.copy() for a in sa]
    ta, tb = start_state(wb, ens), start_state(wb, ens)
    for step in range(steps):
        for stage in (1, 2):
            ea = sys_a(ta, stage, sa, f)
            eb = sys_b(tb, stage, sb, f)
            yield step, stage, ea, eb, sa, sb, ta, tb


@needs_ref
@pytest.mark.parametrize("constrained", [True, False], ids=["shake", "free"])
@pytest.mark.parametrize("ens", [1, 2, 3, 4], ids=lambda e: ENSEMBLES[e])
def test_oracle_ensemble_matches_compiled_reference(c0, ens, constrained):
    """Three steps of each integrator (both stages) from the 3,000-atom
    water box with fixed forces: positions, velocities, the box, the bath
    (chiT, chiP), conint, consv, kin, virshake and the kinetic stress equal
    the compiled reference's bit for bit."""
    wb, f = c0
    if not constrained:
        wb = unconstrained(wb)
    o = O.OracleSystem(wb)
    r = O.RefSystem(wb)

    def sys_o(st, stage, s, f):
        return o.vv_ens(st, stage, *s, f)

    def sys_r(st, stage, s, f):
        r.vv_ens(st, stage, *s, f.copy())
        return 0

    for step, stage, ea, eb, sa, sb, ta, tb in run_pair(sys_o, sys_r, wb, f, ens, 3):
        assert ea == 0
        for a, b in zip(sa, sb):
            assert np.array_equal(a, b), (step, stage)
        for k in O.ENS_SCALARS:
            assert ta[k] == tb[k], (step, stage, k, ta[k], tb[k])
        assert np.array_equal(ta["box9"], tb["box9"])
        assert np.array_equal(ta["stress"], tb["stress"])
    if ens in (2, 4):
        assert not np.array_equal(ta["box9"], wb.box9)   # the barostat moved the box
    if ens in (3, 4):
        assert ta["conint"] != 0. and ta["consv"] != 0.


def test_berendsen_reaches_target_in_one_step_at_tau_dt(c0):
    """vv_nvt_b's lambda^2 = 1 + dt/tauT (K0/K - 1): with tauT = dt and no
    constraints the stage-2 kinetic energy is K0 (to rounding)."""
    wb, f = c0
    wb = unconstrained(wb)
    o = O.OracleSystem(wb)
    s = [wb.x.copy(), wb.y.copy(), wb.z.copy(), wb.vx.copy(), wb.vy.copy(), wb.vz.copy()]
    st = start_state(wb, 1)
    st["tauT"] = o.p.timeStep
    assert o.vv_ens(st, 1, *s, f) == 0
    assert o.vv_ens(st, 2, *s, f) == 0
    ek = 0.5 * float((wb.mass * (s[3] ** 2 + s[4] ** 2 + s[5] ** 2)).sum())
    assert abs(st["kin"] - st["kinTemp0"]) < 1e-13 * st["kinTemp0"]
    assert abs(ek - st["kinTemp0"]) < 1e-12 * st["kinTemp0"]


def test_hoover_at_rest_is_velocity_verlet(c0):
    """vv_nvt_h with chiT = 0 and the target set to the kinetic energy
    leaves chiT at rounding level, so lambda = exp(-dt chiT / 2) rounds to
    1: the positions and velocities are those of plain VV (oc_vv_nve)."""
    wb, f = c0
    o = O.OracleSystem(wb)
    s = [wb.x.copy(), wb.y.copy(), wb.z.copy(), wb.vx.copy(), wb.vy.copy(), wb.vz.copy()]
    t = [a.copy() for a in s]
    st = start_state(wb, 3)
    st["chiT"] = 0.
    st["kinTemp0"] = 0.5 * float((wb.mass * (s[3] ** 2 + s[4] ** 2 + s[5] ** 2)).sum())
    assert o.vv_ens(st, 1, *s, f) == 0
    assert abs(st["chiT"]) < 1e-15
    assert o.vv(1, *t, f)[0] == 0
    for a, b in zip(s, t):
        assert np.array_equal(a, b)


def test_npt_b_box_follows_pressure(c0):
    """vv_npt_b: gamma = 1 + beta dt (P - P0)/tauP, the box scales by
    gamma^(1/3); a pressure above target grows the box, below shrinks it."""
    wb, f = c0
    wb = unconstrained(wb)
    o = O.OracleSystem(wb)
    vol0 = float(np.prod(wb.box9[[0, 4, 8]]))
    ek = 0.5 * float((wb.mass * (wb.vx ** 2 + wb.vy ** 2 + wb.vz ** 2)).sum())
    pp = (2. * ek - start_state(wb, 2)["virpot"]) / (3. * vol0)
    for press0, grow in ((pp - 10., True), (pp + 10., False)):
        s = [wb.x.copy(), wb.y.copy(), wb.z.copy(), wb.vx.copy(), wb.vy.copy(), wb.vz.copy()]
        st = start_state(wb, 2)
        st["press0"] = press0
        assert o.vv_ens(st, 1, *s, f) == 0
        g = 1. + 0.007372 * o.p.timeStep * (pp - press0) / st["tauP"]
        assert np.isclose(st["box9"][0], wb.box9[0] * g ** (1 / 3), rtol=1e-14, atol=0)
        assert (st["box9"][0] > wb.box9[0]) == grow


@pytest.mark.gpu
@pytest.mark.parametrize("constrained", [True, False], ids=["shake", "free"])
@pytest.mark.parametrize("ens", [1, 2, 3, 4], ids=lambda e: ENSEMBLES[e])
def test_dropin_ensemble_against_oracle(c0, ens, constrained):
    """mdg_vv_integrate with ctrl->ens = NVT_B / NPT_B / NVT_H / NPT_H,
    three steps of both stages on the 3,000-atom water box with fixed
    forces, against oc_vv_ens: positions to 1e-10 A, velocities, the box,
    chiT, chiP, conint, consv, kin and virshake to 1e-10 relative (the
    device sums the kinetic energy in a different order; everything else
    is the same FP64 expression), and the kinetic stress."""
    import mdbas_amd as M
    wb, f = c0
    if not constrained:
        wb = unconstrained(wb)
    o = O.OracleSystem(wb)
    g = M.Context(wb)
    print("ensemble", ENSEMBLES[ens], "constrained", constrained, flush=True)

    def sys_o(st, stage, s, f):
        return o.vv_ens(st, stage, *s, f)

    def sys_g(st, stage, s, f):
        g.vv_integrate_ens(st, stage, *s, f)
        return 0

    def close(a, b, rel):
        return abs(a - b) <= rel * max(abs(a), abs(b), 1e-300)

    for step, stage, ea, eb, sa, sb, ta, tb in run_pair(sys_o, sys_g, wb, f, ens, 3):
        assert ea == 0
        for a, b in zip(sa[:3], sb[:3]):
            assert np.abs(a - b).max() < 1e-10, (step, stage)
        for a, b in zip(sa[3:], sb[3:]):
            assert np.abs(a - b).max() < 1e-10 * np.abs(a).max(), (step, stage)
        for k in ("chiT", "chiP", "conint", "consv", "kin", "virshake"):
            assert close(ta[k], tb[k], 1e-10), (step, stage, k, ta[k], tb[k])
        assert np.allclose(ta["box9"], tb["box9"], rtol=1e-13, atol=0)
        assert np.allclose(ta["stress"], tb["stress"], rtol=1e-10, atol=1e-10 * abs(ta["kin"]))
    g.close()


@pytest.mark.gpu
@pytest.mark.parametrize("prec", ["fp64", "mixed"])
@pytest.mark.parametrize("ens", [1, 3], ids=lambda e: ENSEMBLES[e])
def test_device_thermostat_against_oracle(c0, ens, prec):
    """mdg_set_bath + mdg_run_steps with NVT_B / NVT_H: 25 device-resident
    steps (crossing the list rebuild at step 20) against the oracle's MD
    loop with oc_vv_ens from the same bath (chiT 0.3, target 10% above the
    box's kinetic energy): total and potential energy, kinetic energy,
    consv and the final positions.  FP64 to 1e-10 of |E_tot(0)| and 1e-9 A;
    mixed to 1e-5 and 1e-4 A."""
    import mdbas_amd as M
    wb = c0[0]
    o = O.OracleSystem(wb)
    s = [wb.x.copy(), wb.y.copy(), wb.z.copy(), wb.vx.copy(), wb.vy.copy(), wb.vz.copy()]
    st = start_state(wb, ens)
    kT0, chiT0 = st["kinTemp0"], st["chiT"]
    err, f, rec = o.md_run_ens(st, 25, *s)
    assert err == 0
    g = M.Context(wb, precision=M.MDG_FP64 if prec == "fp64" else M.MDG_MIXED)
    g.set_bath(ens, tauT=0.1, chiT=chiT0, kinTemp0=kT0)
    g.upload(wb.x, wb.y, wb.z, wb.vx, wb.vy, wb.vz)
    g.initial_energy()
    R = M.records_to_dict(g.run(25, use_graph=True))
    x = g.download()
    bath = g.get_bath()
    g.close()
    scale = abs(rec[0, 0])
    bar, pos = (1e-10, 1e-9) if prec == "fp64" else (1e-5, 1e-4)
    for k, col in (("tot", 0), ("pot", 1), ("kin", 2), ("consv", 3)):
        e = np.abs(R[k] - rec[1:, col]).max() / scale
        print(ENSEMBLES[ens], prec, k, e, flush=True)
        assert e < bar, k
    if ens == 3:
        assert np.abs(R["consv"]).min() > 0.
        assert bath["chiT"] == pytest.approx(st["chiT"], rel=1e-6 if prec == "mixed" else 1e-10)
    dx = max(np.abs(x[i] - s[i]).max() for i in range(3))
    print("pos", dx, flush=True)
    assert dx < pos


@pytest.mark.gpu
def test_device_thermostat_graph_modes_agree(c0):
    """The same NVT_H run replayed by step graphs, by one graph for the
    whole run and by plain launches gives the same records bit for bit
    (mixed: fixed-point forces, fixed-order sums)."""
    import mdbas_amd as M
    wb = c0[0]
    out = []
    for mode in (0, 1, 2):
        g = M.Context(wb, precision=M.MDG_MIXED)
        st = start_state(wb, 3)
        g.set_bath(3, tauT=0.1, chiT=st["chiT"], kinTemp0=st["kinTemp0"])
        g.upload(wb.x, wb.y, wb.z, wb.vx, wb.vy, wb.vz)
        g.initial_energy()
        out.append(M.records_to_dict(g.run(22, use_graph=mode)))
        g.close()
    for k in ("tot", "kin", "consv", "chiT"):
        assert np.array_equal(out[0][k], out[1][k]), k
        assert np.array_equal(out[0][k], out[2][k]), k
