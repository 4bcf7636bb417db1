"""GPU parity tests: the HIP library (through its C ABI) against the FP64
oracle on the same seeded inputs.

Bars: the neighbour pair set and list order are bit-exact; FP64 energies
and forces agree to 1e-12 relative (summation order differs, nothing
else); the mixed-precision path's RMS relative force error vs FP64 is at
most 1e-5 (the target the mixed path is built to)."""
import numpy as np
import pytest

import mdbas_amd as M
import oracle as O

pytestmark = pytest.mark.gpu


def relerr(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def pairset(nl, np_, no, n):
    a = np.repeat(no, np_[no]).astype(np.int64)
    return np.sort(np.minimum(a, nl) * n + np.maximum(a, nl))


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def c0(c0_box):
    o = O.OracleSystem(c0_box)
    x, y, z = c0_box.x.copy(), c0_box.y.copy(), c0_box.z.copy()
    o.build_list(x, y, z)
    f, e = o.energy(x, y, z)
    return dict(wb=c0_box, o=o, x=(x, y, z), f=f, e=e)


# ------------------------------------------------------------------ drop-in list

@pytest.mark.parametrize("which", ["c0", "link"])
def test_makelist_bit_exact(which, c0_box, link_box):
    wb = c0_box if which == "c0" else link_box
    say("makelist", which, wb.n)
    o = O.OracleSystem(wb)
    x, y, z = wb.x.copy(), wb.y.copy(), wb.z.copy()
    nl, np_, no = o.build_list(x, y, z)
    ctx = M.Context(wb)
    assert ctx.key_link == o.keyLink == (1 if which == "link" else 0)
    gx, gy, gz = wb.x.copy(), wb.y.copy(), wb.z.copy()
    gnl, gnp, gno, eoff, elst = ctx.makelist(gx, gy, gz)
    assert np.array_equal(gnl, nl) and np.array_equal(gnp, np_) and np.array_equal(gno, no)
    assert np.array_equal(gx, x) and np.array_equal(gy, y) and np.array_equal(gz, z)
    assert np.array_equal(eoff, o.exOff) and np.array_equal(elst, o.exLst)
    # update build (no wrap) after a seeded displacement
    rng = np.random.default_rng(77)
    for a, b in ((x, gx), (y, gy), (z, gz)):
        d = rng.normal(0.0, 0.05, len(a))
        a += d
        b += d
    nl, np_, no = o.build_list(x, y, z, build=False)
    gnl, gnp, gno, _, _ = ctx.makelist(gx, gy, gz, step=wb.params["update"])
    assert np.array_equal(gnl, nl) and np.array_equal(gnp, np_) and np.array_equal(gno, no)
    ctx.close()


# ------------------------------------------------------------------ drop-in energies

def test_ewald_energy_parity(c0):
    wb = c0["wb"]
    ctx = M.Context(wb)
    gx, gy, gz = wb.x.copy(), wb.y.copy(), wb.z.copy()
    gnl, gnp, gno, _, _ = ctx.makelist(gx, gy, gz)
    f, e = ctx.ewald_energy(gx, gy, gz, gnl, gnp, gno)
    eo = c0["e"]
    say("ewald", e.elec, eo["elec"], e.vdw, eo["vdw"])
    assert e.elec == pytest.approx(eo["elec"], rel=1e-12)
    assert e.vdw == pytest.approx(eo["vdw"], rel=1e-11)
    assert e.virelec == pytest.approx(eo["virelec"], rel=1e-11)
    assert e.virvdw == pytest.approx(eo["virvdw"], rel=1e-11)
    assert relerr(f, c0["f"]) < 1e-12
    ctx.close()


@pytest.mark.parametrize("which", ["c0", "small"])
def test_spme_energy_parity(which, c0_box, small_box):
    wb = c0_box if which == "c0" else small_box   # 32^3 and a 20^3 (not a power of two) mesh
    o = O.OracleSystem(wb)
    er, fr, st, vr = o.spme(wb.x.copy(), wb.y.copy(), wb.z.copy())
    ctx = M.Context(wb)
    eg, fg, sg, vg = ctx.spme_energy(wb.x.copy(), wb.y.copy(), wb.z.copy())
    say("spme", which, eg, er)
    assert eg == pytest.approx(er, rel=1e-12)
    assert vg == pytest.approx(vr, rel=1e-10)
    assert np.allclose(sg, st, rtol=1e-10, atol=1e-8 * np.abs(st).max())
    assert relerr(fg, fr) < 1e-12
    ctx.close()


@pytest.mark.parametrize("elec", [2, 3, 4])
def test_nonbond_energy_parity(elec, c0):
    import ctypes as C
    wb, o = c0["wb"], c0["o"]
    x, y, z = c0["x"]
    fo = np.zeros((3, wb.n))
    vals = [C.c_double(0) for _ in range(4)]
    O.lib().oc_nonbond(C.byref(o.p), C.byref(o.b), elec, O.P(x), O.P(y), O.P(z), O.P(fo[0]),
                       O.P(fo[1]), O.P(fo[2]), O.P(wb.q), O.P(wb.eps), O.P(wb.sig),
                       O.P(o.neighList), O.P(o.neighPair), O.P(o.neighOrder),
                       *[C.byref(v) for v in vals])
    ctx = M.Context(wb)
    fg, eg = ctx.nonbond_energy(elec, x, y, z, o.neighList, o.neighPair, o.neighOrder)
    assert eg.elec == pytest.approx(vals[0].value, rel=1e-12)
    assert eg.vdw == pytest.approx(vals[1].value, rel=1e-11)
    assert relerr(fg, fo) < 1e-12
    ctx.close()


def test_vv_integrate_parity(c0):
    wb, o = c0["wb"], c0["o"]
    s_o = [wb.x.copy(), wb.y.copy(), wb.z.copy(), wb.vx.copy(), wb.vy.copy(), wb.vz.copy()]
    s_g = [a.copy() for a in s_o]
    ctx = M.Context(wb)
    f = c0["f"]
    err, _, vs, ncyc = o.vv(1, *s_o, f)
    eg = ctx.vv_integrate(1, *s_g, f.copy())
    say("vv1 shake cycles", ncyc)
    assert err == 0
    assert max(np.abs(a - b).max() for a, b in zip(s_o, s_g)) < 1e-12
    assert eg.virshake == pytest.approx(vs, rel=1e-10)
    err, ek, _, _ = o.vv(2, *s_o, f)
    eg = ctx.vv_integrate(2, *s_g, f.copy())
    assert max(np.abs(a - b).max() for a, b in zip(s_o, s_g)) < 1e-12
    assert eg.kin == pytest.approx(ek, rel=1e-12)
    ctx.close()


# ------------------------------------------------------------------ device-resident

@pytest.mark.parametrize("which", ["c0", "link"])
def test_cluster_pair_set_bit_exact(which, c0_box, link_box):
    wb = c0_box if which == "c0" else link_box
    o = O.OracleSystem(wb)
    x, y, z = wb.x.copy(), wb.y.copy(), wb.z.copy()
    nl, np_, no = o.build_list(x, y, z)
    ctx = M.Context(wb)
    ctx.upload(x, y, z, wb.vx, wb.vy, wb.vz)
    ctx.initial_energy()
    pi, pj = ctx.list_pairs()
    got = np.sort(pi.astype(np.int64) * wb.n + pj)
    want = pairset(nl, np_, no, wb.n)
    say("cluster list", which, ctx.list_stats(), len(want))
    assert np.array_equal(got, want)
    ctx.close()


def test_device_fp64_energy_and_trajectory(c0):
    wb = c0["wb"]
    ctx = M.Context(wb, precision=M.MDG_FP64)
    ctx.upload(wb.x, wb.y, wb.z, wb.vx, wb.vy, wb.vz)
    r = ctx.initial_energy()
    eo = c0["e"]
    assert r.pot == pytest.approx(eo["pot"], rel=1e-12)
    assert r.nInCut == eo["nInCut"]
    a = ctx.download()
    assert relerr(np.stack(a[6:9]), c0["f"]) < 1e-12
    nsteps = 25          # crosses the list rebuild at step 20
    recs = M.records_to_dict(ctx.run(nsteps, use_graph=True))
    o = O.OracleSystem(wb)
    s = [wb.x.copy(), wb.y.copy(), wb.z.copy(), wb.vx.copy(), wb.vy.copy(), wb.vz.copy()]
    err, fo, et, ep, ek, _ = o.md_run(nsteps, *s)
    assert err == 0
    g = ctx.download()
    say("traj tot", recs["tot"][-1], et[-1], "pos", max(np.abs(g[k] - s[k]).max() for k in range(3)))
    assert np.abs(recs["tot"] - et[1:]).max() < 1e-10 * abs(et[0])
    assert np.abs(recs["pot"] - ep[1:]).max() < 1e-10 * abs(ep[0])
    assert max(np.abs(g[k] - s[k]).max() for k in range(3)) < 1e-9
    assert relerr(np.stack(g[6:9]), fo) < 1e-9
    ctx.close()


@pytest.mark.parametrize("which", ["c0", "c1"])
def test_mixed_force_error(which, c0_box, c1_box):
    wb = c0_box if which == "c0" else c1_box
    out = {}
    for prec in (M.MDG_FP64, M.MDG_MIXED):
        ctx = M.Context(wb, precision=prec)
        ctx.upload(wb.x, wb.y, wb.z, wb.vx, wb.vy, wb.vz)
        r = ctx.initial_energy()
        out[prec] = (r, np.stack(ctx.download()[6:9]))
        ctx.close()
    f64, fm = out[M.MDG_FP64][1], out[M.MDG_MIXED][1]
    rms = relerr(fm, f64)
    mx = float((np.linalg.norm(fm - f64, axis=0) / np.linalg.norm(f64, axis=0)).max())
    say("mixed", which, "rms rel", rms, "max rel", mx, "pot rel",
        abs(out[M.MDG_MIXED][0].pot / out[M.MDG_FP64][0].pot - 1.0))
    assert rms <= 1e-5
    assert out[M.MDG_MIXED][0].pot == pytest.approx(out[M.MDG_FP64][0].pot, rel=5e-6)


def test_mixed_spme_outlier_atoms(c0_box):
    """Atoms more than one mesh spacing outside the box (whole molecules
    shifted by a box vector) take the binned spread's point-by-point path,
    with the reference's mesh-index rule (spme.c:453-459); FP64 spreads
    them with atomics under the same rule.  Mixed vs FP64 to the mixed bar,
    and the shifted system's energy differs from the unshifted one (the
    rule is not periodic for such atoms, so the path is exercised)."""
    wb = c0_box.copy()
    L = wb.L
    for w in range(0, wb.n_waters, 37):          # every 37th molecule, O H H
        s = slice(3 * w, 3 * w + 3)
        wb.x[s] += L
        wb.y[s] -= 2.0 * L
    pots, forces = {}, {}
    for prec in (M.MDG_FP64, M.MDG_MIXED):
        ctx = M.Context(wb, precision=prec)
        ctx.upload(wb.x, wb.y, wb.z, wb.vx, wb.vy, wb.vz)
        pots[prec] = ctx.initial_energy().pot
        forces[prec] = np.stack(ctx.download()[6:9])
        ctx.close()
    rms = relerr(forces[M.MDG_MIXED], forces[M.MDG_FP64])
    say("outliers rms rel", rms, "pot", pots[M.MDG_MIXED], pots[M.MDG_FP64])
    assert rms <= 1e-5
    assert pots[M.MDG_MIXED] == pytest.approx(pots[M.MDG_FP64], rel=5e-6)
    c = M.Context(c0_box, precision=M.MDG_FP64)
    c.upload(c0_box.x, c0_box.y, c0_box.z, c0_box.vx, c0_box.vy, c0_box.vz)
    p0 = c.initial_energy().pot
    c.close()
    assert abs(p0 - pots[M.MDG_FP64]) > 1e-6 * abs(p0)


def test_mixed_graph_replay_is_bitwise_deterministic(c0_box):
    """The mixed path has no floating-point atomics: graph-replayed runs
    (per-step graphs, and the whole run as one graph) and a run of plain
    launches from the same state agree bit for bit, replay after replay."""
    wb = c0_box
    res = []
    for graph in (1, 0, 2):
        ctx = M.Context(wb, precision=M.MDG_MIXED)
        ctx.upload(wb.x, wb.y, wb.z, wb.vx, wb.vy, wb.vz)
        ctx.initial_energy()
        r1 = M.records_to_dict(ctx.run(12, use_graph=graph))
        _ = ctx.download()               # host work between replays
        r2 = M.records_to_dict(ctx.run(12, use_graph=graph))
        res.append((r1, r2, ctx.download()))
        ctx.close()
    (a1, a2, sa) = res[0]
    for (b1, b2, sb) in res[1:]:
        for k in ("tot", "pot", "kin", "eRec"):
            assert np.array_equal(a1[k], b1[k]) and np.array_equal(a2[k], b2[k]), k
        for u, v in zip(sa, sb):
            assert np.array_equal(u, v)


def _same_records(a, b):
    for k in a:
        assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("skin", [0.5, 0.02])
def test_mixed_pruned_list_is_exact(c0_box, skin):
    """The pruned list leaves out only entries whose pairs stay beyond the
    cutoff while it is in use, so with it the mixed trajectory is the one
    of the full list bit for bit: through two list rebuilds, in step-block
    graphs, per-step graphs and plain launches.  With a 0.5 A skin most
    steps run on the pruned list; with 0.02 A the move guard sends most
    steps back to the full one."""
    wb = c0_box
    out = []
    for s in (0.0, skin):
        ctx = M.Context(wb, precision=M.MDG_MIXED)
        ctx.set_prune(s)
        ctx.upload(wb.x, wb.y, wb.z, wb.vx, wb.vy, wb.vz)
        r0 = ctx.initial_energy()
        r1 = M.records_to_dict(ctx.run(45, use_graph=True))
        r2 = M.records_to_dict(ctx.run(7, use_graph=False))
        out.append((r0, r1, r2, ctx.download(), ctx.prune_stats(), ctx.list_stats()))
        ctx.close()
    (a0, a1, a2, sa, pa, la), (b0, b1, b2, sb, pb, lb) = out
    say("prune", skin, pa, pb, la)
    assert a0.pot == b0.pot
    _same_records(a1, b1)
    _same_records(a2, b2)
    for u, v in zip(sa, sb):
        assert np.array_equal(u, v)
    assert pa["pruned_launches"] == 0
    if skin == 0.5:
        assert pb["pruned_launches"] >= 40
        assert 0 < pb["pruned_entries"] < lb["cluster_pairs"]


def test_batch_pruned_list_is_exact(replicas):
    """The replica batch with and without the pruned list: the same
    trajectories bit for bit over 25 steps (one rebuild)."""
    out = []
    for s in (0.0, 1.0):
        b = M.BatchContext(replicas, precision=M.MDG_MIXED)
        b.set_prune(s)
        b.upload_boxes()
        b.initial_energy()
        r = M.batch_records(b.run(25, use_graph=True), b.R)
        out.append((r, b.download(), b.prune_stats()))
        b.close()
    (ra, sa, pa), (rb, sb, pb) = out
    say("batch prune", pa, pb)
    _same_records(ra, rb)
    for u, v in zip(sa, sb):
        assert np.array_equal(u, v)
    assert pa["pruned_launches"] == 0 and pb["pruned_launches"] >= 15


# ------------------------------------------------------------------ adjoint

@pytest.mark.parametrize("prec", [M.MDG_FP64, M.MDG_MIXED])
def test_backward_matches_oracle_gradient(prec, c0):
    """mdg_backward (FP64 in either context precision) against the oracle's
    dE/d(q, eps, sig) on the same list and coordinates."""
    wb, o = c0["wb"], c0["o"]
    g_o = o.energy_grad(*c0["x"])
    ctx = M.Context(wb, precision=prec)
    ctx.upload(wb.x, wb.y, wb.z, wb.vx, wb.vy, wb.vz)
    ctx.initial_energy()
    g = ctx.backward(2.5)
    ctx.close()
    for k in range(3):
        say("backward", prec, k, relerr(g[k], 2.5 * g_o[k]))
        assert relerr(g[k], 2.5 * g_o[k]) < 1e-11


def _c3_water(wb):
    import torch
    from mdbas_amd import autograd as AG
    water = AG.Water(wb)
    q = torch.tensor(wb.q, dtype=torch.float64, requires_grad=True)
    eps_o = torch.tensor([0.1521], dtype=torch.float64, requires_grad=True)
    rmin_o = torch.tensor([1.7682], dtype=torch.float64, requires_grad=True)
    zero = torch.zeros(1, dtype=torch.float64)
    return torch, AG, water, q, eps_o, rmin_o, zero


def test_backward_torch_harness_against_finite_differences(c0_box):
    """Instantaneous potential: per-type LJ (eps kcal/mol, Rmin/2) and
    per-atom charges as FP64 leaves through torch.autograd, against the
    4-point stencil of numderiv.c:102 (h = 1e-6) on GPU FP64 energies."""
    from mdbas_amd.numderiv import derivative
    torch, AG, water, q, eps_o, rmin_o, zero = _c3_water(c0_box)
    E = AG.potential_energy(water, q, torch.cat([eps_o, zero]), torch.cat([rmin_o, zero]))
    E.backward()
    with torch.no_grad():
        def fe(qv, ev, rv):
            return float(AG.potential_energy(water, qv, torch.cat([ev, zero]),
                                             torch.cat([rv, zero])))
        fd_e = derivative(lambda d: fe(q, eps_o + d, rmin_o), 1e-6, 4)
        fd_r = derivative(lambda d: fe(q, eps_o, rmin_o + d), 1e-6, 4)
        say("adjoint eps", float(eps_o.grad), fd_e, "rmin", float(rmin_o.grad), fd_r)
        assert float(eps_o.grad) == pytest.approx(fd_e, rel=1e-7)
        assert float(rmin_o.grad) == pytest.approx(fd_r, rel=1e-7)
        # E is quadratic in each charge, so the 4-point stencil is exact for
        # any h; h = 1e-4 keeps the rounding of E (1e-16 of the sum of
        # |pair terms|, about 1e-8) small against h
        for i in (0, 1, 2, 1500):
            dq = torch.zeros_like(q)
            dq[i] = 1.0
            fd = derivative(lambda d: fe(q + d * dq, eps_o, rmin_o), 1e-4, 4)
            assert float(q.grad[i]) == pytest.approx(fd, rel=1e-8, abs=1e-4), i
    water.close()


def test_backward_traj_checkpointing_is_exact(c0_box):
    """The same 30-step trajectory gradient with one segment and with
    segments of 20 + 10 steps (recomputed from checkpoints)."""
    wb = c0_box
    out = []
    for ckpt in (30, 20, 7):
        ctx = M.Context(wb, precision=M.MDG_FP64)
        ctx.upload(wb.x, wb.y, wb.z, wb.vx, wb.vy, wb.vz)
        ctx.initial_energy()
        out.append(ctx.backward_traj(30, ckpt))
        ctx.close()
    (J0, g0) = out[0]
    for J, g in out[1:]:
        assert J == pytest.approx(J0, rel=1e-13)
        for k in range(3):
            assert relerr(g[k], g0[k]) < 1e-9


def _traj_fd_checks(c0_box, T, K, h):
    from mdbas_amd.numderiv import derivative
    torch, AG, water, q, eps_o, rmin_o, zero = _c3_water(c0_box)
    J = AG.time_averaged_potential(water, q, torch.cat([eps_o, zero]), torch.cat([rmin_o, zero]),
                                   T, K)
    J.backward()
    rng = np.random.default_rng(8)
    dq = torch.tensor(rng.normal(size=len(q)), dtype=torch.float64)
    with torch.no_grad():
        def fj(qv, ev, rv):
            water._set(qv, torch.sqrt(torch.cat([ev, zero])[water.types] * 418.4),
                       torch.cat([rv, zero])[water.types] / AG.SIXTH_ROOT_2)
            water.ctx.initial_energy()
            R = M.records_to_dict(water.ctx.run(T, use_graph=True))
            return float(R["pot"].mean())
        assert fj(q, eps_o, rmin_o) == pytest.approx(float(J), rel=1e-12)
        checks = [("eps_O", float(eps_o.grad), derivative(lambda d: fj(q, eps_o + d, rmin_o), h, 4)),
                  ("rmin_O", float(rmin_o.grad), derivative(lambda d: fj(q, eps_o, rmin_o + d), h, 4)),
                  ("q.dir", float(q.grad @ dq), derivative(lambda d: fj(q + d * dq, eps_o, rmin_o), h, 4))]
        # a directional derivative is judged on its own scale |g| |dq| (the
        # random direction's terms largely cancel)
        scale = {"eps_O": None, "rmin_O": None, "q.dir": float(q.grad.norm() * dq.norm())}
    water.close()
    for name, g, fd in checks:
        say("traj adjoint T", T, name, g, fd, (g - fd) / abs(fd), "scale", scale[name])
    return [(name, g, fd, scale[name]) for name, g, fd in checks]


def test_backward_traj_against_finite_differences(c0_box):
    """Mean potential over 50 steps (dt = 1 fs, checkpoints every 20):
    the reverse-sweep gradient against the 4-point stencil of
    numderiv.c:102 (h = 1e-5) on whole FP64 GPU trajectories, for the O
    eps and Rmin/2 leaves and a random direction in the charges."""
    for name, g, fd, scale in _traj_fd_checks(c0_box, 50, 20, 1e-5):
        assert abs(g - fd) <= 1e-5 * (scale or abs(fd)), name


def test_backward_traj_200_steps_against_finite_differences(c0_box):
    """The 200-step average (checkpoints every 20) with h = 1e-6.  Over
    200 fs the trajectory's sensitivity grows (the stencils at h = 1e-4,
    1e-5 and 1e-6 themselves spread by about 3e-4 of the derivative: SHAKE
    stops on a tolerance, so J(theta) has small steps that the dynamics
    amplifies), so the bar here is 1e-3; the 50-step test holds 1e-5."""
    for name, g, fd, scale in _traj_fd_checks(c0_box, 200, 20, 1e-6):
        assert abs(g - fd) <= 1e-3 * (scale or abs(fd)), name


@pytest.mark.parametrize("T,K", [(1, 1), (5, 5), (50, 20), (200, 20)])
def test_backward_traj_against_oracle_adjoint(c0_box, T, K):
    """The device trajectory adjoint (analytic Hessian-vector products,
    checkpoints every K steps) against the oracle's FP64 adjoint of the same
    T-step objective (oc_traj_adjoint, every step stored, pinned against
    differences of J on the CPU): J to 1e-12, each of dJ/dq, dJ/deps and
    dJ/dsig to 1e-8 relative (norm over atoms).  (200, 20) is the c3
    configuration."""
    wb = c0_box
    o = O.OracleSystem(wb)
    say("oracle adjoint T", T)
    Jo, dq, de, ds = o.traj_adjoint(T, wb.x, wb.y, wb.z, wb.vx, wb.vy, wb.vz)
    ctx = M.Context(wb, precision=M.MDG_FP64)
    ctx.upload(wb.x, wb.y, wb.z, wb.vx, wb.vy, wb.vz)
    ctx.initial_energy()
    J, g = ctx.backward_traj(T, K)
    ctx.close()
    errs = [relerr(g[0], dq), relerr(g[1], de), relerr(g[2], ds)]
    say("traj adjoint vs oracle T", T, "J", J, Jo, "rel", errs)
    assert J == pytest.approx(Jo, rel=1e-12)
    for e in errs:
        assert e < 1e-8


# ------------------------------------------------------------------ replica batch

@pytest.fixture(scope="module")
def replicas():
    from mdbas_amd import configs
    cfgs = configs.replica_configs(6)
    return [configs.build(c) for c in cfgs]


def test_batch_pair_set_bit_exact(replicas):
    """The batch's cluster list holds, replica by replica, exactly the
    reference O(N^2) list (global indices offset by each replica's start)."""
    b = M.BatchContext(replicas, precision=M.MDG_MIXED)
    b.upload_boxes()
    b.initial_energy()
    pi, pj = b.list_pairs()
    got = np.sort(pi.astype(np.int64) * b.n + pj)
    want = []
    for r, wb in enumerate(replicas):
        o = O.OracleSystem(wb)
        nl, np_, no = o.build_list(wb.x.copy(), wb.y.copy(), wb.z.copy())
        off = int(b.aoff[r])
        a = np.repeat(no, np_[no]).astype(np.int64) + off
        c = nl.astype(np.int64) + off
        want.append(np.minimum(a, c) * b.n + np.maximum(a, c))
    want = np.sort(np.concatenate(want))
    say("batch list", b.list_stats(), len(want))
    b.close()
    assert np.array_equal(got, want)


def test_batch_fp64_matches_single_contexts(replicas):
    """Every replica of an FP64 batch follows its own single-system FP64
    context: step-0 energies and forces to 1e-12, 25 steps (through the
    rebuild at step 20) to 1e-10 in energy and 1e-9 A in position."""
    nsteps = 25
    b = M.BatchContext(replicas, precision=M.MDG_FP64)
    b.upload_boxes()
    r0 = b.initial_energy()
    fb = [b.split(a) for a in b.download()[6:9]]
    rb = M.batch_records(b.run(nsteps, use_graph=True), b.R)
    xb = [b.split(a) for a in b.download()[:3]]
    b.close()
    for r, wb in enumerate(replicas):
        c = M.Context(wb, precision=M.MDG_FP64)
        c.upload(wb.x, wb.y, wb.z, wb.vx, wb.vy, wb.vz)
        s0 = c.initial_energy()
        f = np.stack(c.download()[6:9])
        rs = M.records_to_dict(c.run(nsteps, use_graph=True))
        xs = c.download()[:3]
        c.close()
        say("batch fp64 replica", r, wb.n, r0[r].pot, s0.pot, rb["tot"][-1, r], rs["tot"][-1])
        assert r0[r].pot == pytest.approx(s0.pot, rel=1e-12)
        assert r0[r].nInCut == s0.nInCut
        assert relerr(np.stack([fb[k][r] for k in range(3)]), f) < 1e-12
        assert np.abs(rb["tot"][:, r] - rs["tot"]).max() < 1e-10 * abs(rs["tot"][0])
        assert np.abs(rb["pot"][:, r] - rs["pot"]).max() < 1e-10 * abs(rs["pot"][0])
        assert max(np.abs(xb[k][r] - xs[k]).max() for k in range(3)) < 1e-9


def test_batch_mixed_matches_single_contexts(replicas):
    """Mixed batch against mixed single contexts (the same FP32 pair math,
    another summation order): energies to 2e-6 over 25 steps, step-0
    forces to 2e-6 RMS."""
    nsteps = 25
    b = M.BatchContext(replicas, precision=M.MDG_MIXED)
    b.upload_boxes()
    r0 = b.initial_energy()
    fb = [b.split(a) for a in b.download()[6:9]]
    rb = M.batch_records(b.run(nsteps, use_graph=True), b.R)
    b.close()
    for r, wb in enumerate(replicas):
        c = M.Context(wb, precision=M.MDG_MIXED)
        c.upload(wb.x, wb.y, wb.z, wb.vx, wb.vy, wb.vz)
        s0 = c.initial_energy()
        f = np.stack(c.download()[6:9])
        rs = M.records_to_dict(c.run(nsteps, use_graph=True))
        c.close()
        assert r0[r].pot == pytest.approx(s0.pot, rel=2e-6)
        assert relerr(np.stack([fb[k][r] for k in range(3)]), f) < 2e-6
        assert np.abs(rb["tot"][:, r] - rs["tot"]).max() < 2e-6 * abs(rs["tot"][0])


def test_autotune_keeps_results(c0_box):
    """Launch shapes change no arithmetic: after mdg_autotune the mixed
    trajectory's positions are bit for bit those of the default shapes,
    and the uploaded state is restored."""
    wb = c0_box
    out = []
    for tune in (False, True):
        ctx = M.Context(wb, precision=M.MDG_MIXED)
        ctx.upload(wb.x, wb.y, wb.z, wb.vx, wb.vy, wb.vz)
        ctx.initial_energy()
        if tune:
            say("autotune", ctx.autotune(40))
            a = ctx.download()
            assert all(np.array_equal(a[k], v) for k, v in enumerate((wb.x, wb.y, wb.z)))
        r = M.records_to_dict(ctx.run(25, use_graph=True))
        out.append((r, ctx.download()))
        ctx.close()
    (ra, sa), (rb, sb) = out
    for u, v in zip(sa[:6], sb[:6]):
        assert np.array_equal(u, v)
    assert np.allclose(ra["tot"], rb["tot"], rtol=1e-13, atol=0)


@pytest.mark.parametrize("prec", ["mixed", "fp64"])
def test_launch_shapes_keep_results(c0_box, prec):
    """mdg_set_launch round-trips, and the branch order at the SPME fork,
    the SHAKE spread (a molecule per quad of lanes, or per lane at spread
    0) and the pair grid change no per-atom arithmetic: the trajectories'
    positions and velocities agree bit for bit (the step records' FP64
    sums of per-block partials are grouped by the launch shape, so they
    agree to 1e-13).  FP64 forces are summed with FP64 atomics, whose order
    varies from run to run, so FP64 trajectories agree to 1e-9 A."""
    wb = c0_box
    out = []
    for shape in ((1, 1024, 0, 0), (2, 512, 1, 1), (0, 768, 0, 1), (1, 1024, 0, 2)):
        ctx = M.Context(wb, precision=M.MDG_MIXED if prec == "mixed" else M.MDG_FP64)
        ctx.set_launch(*shape)
        assert tuple(ctx.get_launch().values()) == shape
        ctx.upload(wb.x, wb.y, wb.z, wb.vx, wb.vy, wb.vz)
        ctx.initial_energy()
        r = M.records_to_dict(ctx.run(25, use_graph=True))
        out.append((r, ctx.download()))
        ctx.close()
    ra, sa = out[0]
    for rb, sb in out[1:]:
        if prec == "mixed":
            for u, v in zip(sa[:6], sb[:6]):
                assert np.array_equal(u, v)
        else:
            assert max(np.abs(u - v).max() for u, v in zip(sa[:3], sb[:3])) < 1e-9
        for k in ("tot", "pot", "kin", "eRec", "virshake"):
            assert np.allclose(ra[k], rb[k], rtol=1e-13 if prec == "mixed" else 1e-10, atol=0), k
        assert np.array_equal(ra["nInCut"], rb["nInCut"])


def test_batch_shake_per_lane_matches_per_quad(replicas):
    """A replica batch with one molecule per lane in the SHAKE kernels
    (spread 0) follows the batch with one per quad bit for bit over 25
    steps (through the rebuild at step 20); the per-replica records, sums
    of per-molecule terms in a fixed order, agree exactly too."""
    out = []
    for spread in (1, 0):
        b = M.BatchContext(replicas, precision=M.MDG_MIXED)
        b.set_launch(spread, 1024, 0, 0)
        b.upload_boxes()
        b.initial_energy()
        rb = M.batch_records(b.run(25, use_graph=True), b.R)
        out.append((rb, b.download()))
        b.close()
    (ra, sa), (rb, sb) = out
    for u, v in zip(sa[:6], sb[:6]):
        assert np.array_equal(u, v)
    for k in ("tot", "kin", "virshake"):
        assert np.array_equal(ra[k], rb[k]), k
