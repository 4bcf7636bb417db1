"""Plain cutoff electrostatics (keyEwald 0: the SHIFT1 / SHIFT2 / SWITCH
functors of elec.c:137-263 through nonbond_energy, energy.c:216-345) on
the device-resident path: the cluster pair kernels (FP64 and mixed) with
no SPME, no excluded-pair correction and no self or net-charge terms,
against the FP64 oracle on the same inputs (the 3,000-atom water box)."""
import numpy as np
import pytest

import mdbas_amd as M
import oracle as O
from mdbas_amd import configs

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


@pytest.fixture(scope="module")
def c0():
    cfg = configs.CONFIGS["c0"]
    return configs.build(cfg), cfg


def plain(cfg, elec):
    return dict(cfg, keyEwald=0, elecType=elec)


@pytest.mark.parametrize("elec", [2, 3, 4])
@pytest.mark.parametrize("prec", [M.MDG_FP64, M.MDG_MIXED])
def test_device_plain_energy_against_oracle(elec, prec, c0):
    """Step-0 energy and forces of the device step: FP64 to 1e-12, mixed
    RMS relative force error <= 1e-5 and the potential within 5e-6."""
    wb, cfg = c0
    o = O.OracleSystem(wb)
    x, y, z = wb.x.copy(), wb.y.copy(), wb.z.copy()
    o.build_list(x, y, z)
    fo, eo = o.nonbond_energy(elec, x, y, z)
    ctx = M.Context(wb, plain(cfg, elec), precision=prec)
    ctx.upload(wb.x, wb.y, wb.z, wb.vx, wb.vy, wb.vz)
    r = ctx.initial_energy()
    f = np.stack(ctx.download()[6:9])
    ctx.close()
    print("plain", elec, prec, r.elec, eo["elec"], r.vdw, eo["vdw"], relerr(f, fo), flush=True)
    assert r.eRec == 0.0 and r.eCorr == 0.0
    if prec == M.MDG_FP64:
        assert r.elec == pytest.approx(eo["elec"], rel=1e-12)
        assert r.vdw == pytest.approx(eo["vdw"], rel=1e-11)
        assert r.virelec == pytest.approx(eo["virelec"], rel=1e-11)
        assert relerr(f, fo) < 1e-12
    else:
        assert relerr(f, fo) <= 1e-5
        assert r.pot == pytest.approx(eo["pot"], rel=5e-6)


@pytest.mark.parametrize("elec", [2, 4])
def test_device_plain_trajectory_against_oracle(elec, c0):
    """25 FP64 device steps (dt 1 fs, SHAKE, list rebuilt at step 20)
    against the oracle's plain-path MD loop: energies to 1e-10, positions
    to 1e-9 A."""
    wb, cfg = c0
    nsteps = 25
    ctx = M.Context(wb, plain(cfg, elec), precision=M.MDG_FP64)
    ctx.upload(wb.x, wb.y, wb.z, wb.vx, wb.vy, wb.vz)
    ctx.initial_energy()
    recs = M.records_to_dict(ctx.run(nsteps, use_graph=True))
    g = ctx.download()
    ctx.close()
    o = O.OracleSystem(wb)
    s = [wb.x.copy(), wb.y.copy(), wb.z.copy(), wb.vx.copy(), wb.vy.copy(), wb.vz.copy()]
    err, fo, et, ep = o.md_run_plain(elec, nsteps, *s)
    assert err == 0
    dx = max(np.abs(g[k] - s[k]).max() for k in range(3))
    print("plain traj", elec, recs["tot"][-1], et[-1], "pos", dx, flush=True)
    assert np.abs(recs["tot"] - et[1:]).max() < 1e-10 * abs(et[0])
    assert np.abs(recs["pot"] - ep[1:]).max() < 1e-10 * abs(ep[0])
    assert dx < 1e-9


def test_mixed_plain_trajectory_tracks_fp64(c0):
    """100 mixed steps (SHIFT2) follow the FP64 device run: the total
    energy and its change since step 0 agree to 1e-5 and 5e-6 of the
    energy at every step (FP32 pair forces, RMS error ~6e-7, let the two
    trajectories separate slowly)."""
    wb, cfg = c0
    out = {}
    for prec in (M.MDG_FP64, M.MDG_MIXED):
        ctx = M.Context(wb, plain(cfg, 3), precision=prec)
        ctx.upload(wb.x, wb.y, wb.z, wb.vx, wb.vy, wb.vz)
        ctx.initial_energy()
        out[prec] = M.records_to_dict(ctx.run(100, use_graph=True))["tot"]
        ctx.close()
    m, f = out[M.MDG_MIXED], out[M.MDG_FP64]
    d = np.abs(m - f).max() / abs(f[0])
    dd = np.abs((m - m[0]) - (f - f[0])).max() / abs(f[0])
    print("mixed vs fp64 plain", d, "energy change", dd, flush=True)
    assert d < 1e-5
    assert dd < 5e-6
