"""GPU parity at the benchmark sizes, through the C ABI.

The 30,000-atom box (c1: 67.0 A, 4^3 link cells, 64^3 mesh), the
192,000-atom box (c2: 124.1 A, 8^3 link cells, 128^3 mesh) and the full
256-replica batch (c4), each against the FP64 oracle on the same inputs:

* the neighbour list of the drop-in makelist, element for element (first
  build with its in-place wrap, and the update build after a seeded
  displacement), against the oracle and against the digests of the
  compiled reference's own lists (tests/golden/lists.npz, written by
  tests/golden/make_golden.py from list.c:1793-2619);
* the device cluster list's pair set, bit for bit;
* FP64 device energies and forces to 1e-12 (oracle: real space over the
  reference list + SPME), and a 25-step FP64 trajectory at c1 crossing
  the list rebuild at step 20;
* mixed-precision forces against the oracle's FP64 forces, RMS relative
  error <= 1e-5 (the mixed path's stated bar)."""
import hashlib
import os

import numpy as np
import pytest

import mdbas_amd as M
import oracle as O

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def say(*a):
    print(*a, flush=True)


def digest(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def relerr(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def pairset(nl, np_, no, n):
    a = np.repeat(no, np_[no]).astype(np.int64)
    return np.sort(np.minimum(a, nl) * n + np.maximum(a, nl))


@pytest.fixture(scope="module")
def gl():
    return np.load(os.path.join(GOLD, "lists.npz"))


def _box(which, c1_box, c2_box):
    return c1_box if which == "c1" else c2_box


@pytest.mark.parametrize("which", ["c1", "c2"])
def test_makelist_bit_exact_at_scale(which, c1_box, c2_box, gl):
    """Drop-in makelist (link cells) against the oracle and the compiled
    reference's digests: neighList, neighPair, neighOrder and the wrapped
    coordinates of the first build, then the update build."""
    wb = _box(which, c1_box, c2_box)
    say("makelist", which, wb.n)
    ctx = M.Context(wb)
    assert ctx.key_link == 1 == int(gl[f"{which}_keylink"])
    gx, gy, gz = wb.x.copy(), wb.y.copy(), wb.z.copy()
    gnl, gnp, gno, eoff, elst = ctx.makelist(gx, gy, gz)
    assert len(gnl) == int(gl[f"{which}_count"])
    assert digest(gnl, gnp, gno) == str(gl[f"{which}_digest"])
    assert digest(gx, gy, gz) == str(gl[f"{which}_wrapped_digest"])
    o = O.OracleSystem(wb)
    x, y, z = wb.x.copy(), wb.y.copy(), wb.z.copy()
    nl, np_, no = o.build_list(x, y, z)
    assert np.array_equal(gnl, nl) and np.array_equal(gnp, np_) and np.array_equal(gno, no)
    assert np.array_equal(eoff, o.exOff) and np.array_equal(elst, o.exLst)
    del nl, gnl
    rng = np.random.default_rng(1234)          # make_golden.py's displacement
    for a in (gx, gy, gz):
        a += rng.normal(0.0, 0.05, len(a))
    assert digest(gx, gy, gz) == str(gl[f"{which}_upd_pos_digest"])
    gnl, gnp, gno, _, _ = ctx.makelist(gx, gy, gz, step=wb.params["update"])
    say("update", which, len(gnl))
    assert len(gnl) == int(gl[f"{which}_upd_count"])
    assert digest(gnl, gnp, gno) == str(gl[f"{which}_upd_digest"])
    ctx.close()


@pytest.mark.parametrize("which", ["c1", "c2"])
def test_cluster_pair_set_bit_exact_at_scale(which, c1_box, c2_box):
    """The device-resident cluster-pair list (the list the benchmark runs)
    holds exactly the reference list's atom pairs."""
    wb = _box(which, c1_box, c2_box)
    o = O.OracleSystem(wb)
    nl, np_, no = o.build_list(wb.x.copy(), wb.y.copy(), wb.z.copy())
    want = pairset(nl, np_, no, wb.n)
    del nl
    for prec in (M.MDG_MIXED, M.MDG_FP64):
        ctx = M.Context(wb, precision=prec)
        ctx.upload(wb.x, wb.y, wb.z, wb.vx, wb.vy, wb.vz)
        ctx.initial_energy()
        pi, pj = ctx.list_pairs()
        got = np.sort(pi.astype(np.int64) * wb.n + pj)
        say("cluster list", which, prec, ctx.list_stats(), len(want))
        ctx.close()
        assert np.array_equal(got, want)


@pytest.fixture(scope="module")
def oracle_energy():
    cache = {}

    def get(wb, key):
        if key not in cache:
            o = O.OracleSystem(wb)
            x, y, z = wb.x.copy(), wb.y.copy(), wb.z.copy()
            o.build_list(x, y, z)
            f, e = o.energy(x, y, z)
            cache[key] = (f, e)
        return cache[key]
    return get


@pytest.mark.parametrize("which", ["c1", "c2"])
def test_device_fp64_energy_at_scale(which, c1_box, c2_box, oracle_energy):
    """FP64 device energy and forces of the uploaded state (list build,
    SPME, real space, excluded-pair correction) against the oracle.  The
    potential is a sum of large terms of both signs (at 30k atoms the
    reciprocal part alone is -2.6e8 against a total of -5.8e6, and the
    direct part sums 1e7 pair terms of both signs) summed in another order
    than the oracle's, so the total and each part are held to 1e-12 of the
    summed magnitudes of the parts; forces to 1e-12 relative."""
    wb = _box(which, c1_box, c2_box)
    fo, eo = oracle_energy(wb, which)
    ctx = M.Context(wb, precision=M.MDG_FP64)
    ctx.upload(wb.x, wb.y, wb.z, wb.vx, wb.vy, wb.vz)
    r = ctx.initial_energy()
    f = np.stack(ctx.download()[6:9])
    ctx.close()
    say("fp64", which, r.pot, eo["pot"], r.eRec, eo["eRec"], relerr(f, fo))
    assert r.nInCut == eo["nInCut"]
    scale = abs(eo["eDir"]) + abs(eo["eRec"]) + abs(eo["eCorr"]) + abs(eo["vdw"])
    say("pot difference / scale", abs(r.pot - eo["pot"]) / scale)
    assert abs(r.pot - eo["pot"]) <= 1e-12 * scale
    for k in ("eRec", "eDir", "eCorr", "vdw"):
        assert abs(getattr(r, k) - eo[k]) <= 1e-12 * scale, k
    assert relerr(f, fo) < 1e-12


@pytest.mark.parametrize("which", ["c1", "c2"])
def test_mixed_forces_against_oracle_at_scale(which, c1_box, c2_box, oracle_energy):
    """Mixed-precision forces (FP32 pair and spline math, int64 fixed-point
    accumulation) against the oracle's FP64 forces: RMS relative error
    <= 1e-5; the potential within 5e-6."""
    wb = _box(which, c1_box, c2_box)
    fo, eo = oracle_energy(wb, which)
    ctx = M.Context(wb, precision=M.MDG_MIXED)
    ctx.upload(wb.x, wb.y, wb.z, wb.vx, wb.vy, wb.vz)
    r = ctx.initial_energy()
    f = np.stack(ctx.download()[6:9])
    ctx.close()
    rms = relerr(f, fo)
    mx = float((np.linalg.norm(f - fo, axis=0) / np.linalg.norm(fo, axis=0)).max())
    say("mixed vs oracle", which, "rms", rms, "max", mx, "pot", r.pot / eo["pot"] - 1)
    assert r.nInCut == eo["nInCut"]
    assert rms <= 1e-5
    assert r.pot == pytest.approx(eo["pot"], rel=5e-6)


def _mixed_run(wb, path, skin=0.0, nsteps=25):
    ctx = M.Context(wb, precision=M.MDG_MIXED)
    ctx.set_pair_path(path)
    ctx.set_prune(skin)
    ctx.upload(wb.x, wb.y, wb.z, wb.vx, wb.vy, wb.vz)
    r0 = ctx.initial_energy()
    f0 = np.stack(ctx.download()[6:9])
    recs = M.records_to_dict(ctx.run(nsteps, use_graph=True))
    out = (r0, f0, recs, ctx.download(), ctx.prune_stats())
    ctx.close()
    return out


@pytest.mark.parametrize("which", ["c1", "c2"])
def test_mixed_relative_coordinates(which, c1_box, c2_box, oracle_energy):
    """The mixed pair kernel in relative coordinates (path 0: FP32
    displacements from cluster origins, the FP64 cutoff test only for
    entries with a listed pair near the cutoff) against the same with the
    FP64 test on every entry (path 2): bit-identical step-0 forces and a
    bit-identical 25-step trajectory through the rebuild at step 20, so the
    rounding bound never lets the FP32 test decide otherwise than the
    reference's FP64 one.  Against FP64 displacements (path 1): the paths
    differ (path 0 does run in relative coordinates), the in-cutoff counts
    agree at step 0 (within 1e-6 over the 25 steps, as the trajectories part
    at the FP32 level), step-0 forces agree to 1e-6, and both meet the mixed
    bar against the oracle (RMS relative force error <= 1e-5, nInCut
    exact).  With the pruned list (0.8 A) path 0 is the full list's bit
    for bit (c1).  c2's 124 A box has origin offsets up to 62 A, so its
    rounding bound is the larger one."""
    wb = _box(which, c1_box, c2_box)
    fo, eo = oracle_energy(wb, which)
    n = 25 if which == "c1" else 12
    a = _mixed_run(wb, 0, nsteps=n)
    b = _mixed_run(wb, 2, nsteps=n)
    c = _mixed_run(wb, 1, nsteps=n)
    say("rel vs exact", a[0].pot, b[0].pot, "rel vs fp64-disp pot", a[0].pot / c[0].pot - 1,
        "forces", relerr(a[1], c[1]), "vs oracle", relerr(a[1], fo), relerr(c[1], fo))
    assert np.array_equal(a[1], b[1])
    assert a[0].pot == b[0].pot and a[0].nInCut == b[0].nInCut
    for k in ("tot", "pot", "kin", "eDir", "virelec", "virvdw", "virshake", "nInCut"):
        assert np.array_equal(a[2][k], b[2][k]), k
    for u, v in zip(a[3][:6], b[3][:6]):
        assert np.array_equal(u, v)
    assert not np.array_equal(a[1], c[1])
    # the two trajectories part at the FP32 level, and within a few steps
    # a pair or two on the cutoff sphere falls on different sides
    assert np.abs(a[2]["nInCut"] - c[2]["nInCut"]).max() <= 1e-6 * eo["nInCut"]
    assert relerr(a[1], c[1]) < 1e-6
    for r, f in ((a[0], a[1]), (c[0], c[1])):
        assert r.nInCut == eo["nInCut"]
        assert relerr(f, fo) <= 1e-5
    if which != "c1":
        return
    p = _mixed_run(wb, 0, skin=0.8, nsteps=45)
    q = _mixed_run(wb, 0, skin=0.0, nsteps=45)
    say("pruned launches", p[4])
    assert p[4]["pruned_launches"] >= 30
    for k in ("tot", "pot", "eDir", "virelec", "virvdw", "nInCut"):
        assert np.array_equal(p[2][k], q[2][k]), k
    for u, v in zip(p[3][:6], q[3][:6]):
        assert np.array_equal(u, v)


def test_device_fp64_trajectory_c1(c1_box):
    """25 FP64 device steps of the 30,000-atom box (dt 2 fs, SHAKE, list
    rebuilt at step 20) against the oracle's MD loop."""
    wb = c1_box
    nsteps = 25
    ctx = M.Context(wb, precision=M.MDG_FP64)
    ctx.upload(wb.x, wb.y, wb.z, wb.vx, wb.vy, wb.vz)
    ctx.initial_energy()
    recs = M.records_to_dict(ctx.run(nsteps, use_graph=True))
    g = ctx.download()
    ctx.close()
    o = O.OracleSystem(wb)
    s = [wb.x.copy(), wb.y.copy(), wb.z.copy(), wb.vx.copy(), wb.vy.copy(), wb.vz.copy()]
    err, fo, et, ep, ek, _ = o.md_run(nsteps, *s)
    assert err == 0
    dx = max(np.abs(g[k] - s[k]).max() for k in range(3))
    say("c1 traj tot", recs["tot"][-1], et[-1], "pos", dx)
    assert np.abs(recs["tot"] - et[1:]).max() < 1e-10 * abs(et[0])
    assert np.abs(recs["pot"] - ep[1:]).max() < 1e-10 * abs(ep[0])
    assert dx < 1e-9
    assert relerr(np.stack(g[6:9]), fo) < 1e-9


# ------------------------------------------------------------------ 256 replicas

@pytest.fixture(scope="module")
def replicas256():
    from mdbas_amd import configs
    return [configs.build(c) for c in configs.replica_configs(256)]


def test_batch256_pair_set_bit_exact(replicas256):
    """All 256 replicas in one batch: the cluster list holds, replica by
    replica, exactly the reference O(N^2) list (verlet_list,
    list.c:896-1123), offsets added."""
    reps = replicas256
    b = M.BatchContext(reps, precision=M.MDG_MIXED)
    b.upload_boxes()
    b.initial_energy()
    pi, pj = b.list_pairs()
    got = np.sort(pi.astype(np.int64) * b.n + pj)
    aoff = b.aoff.copy()
    n = b.n
    b.close()
    want = []
    for r, wb in enumerate(reps):
        o = O.OracleSystem(wb)
        nl, np_, no = o.build_list(wb.x.copy(), wb.y.copy(), wb.z.copy())
        a = np.repeat(no, np_[no]).astype(np.int64) + aoff[r]
        c = nl.astype(np.int64) + aoff[r]
        want.append(np.minimum(a, c) * n + np.maximum(a, c))
    want = np.sort(np.concatenate(want))
    say("batch256 list", len(got), len(want))
    assert np.array_equal(got, want)


def test_batch256_energies_against_oracle(replicas256):
    """Step-0 energies and forces of every one of the 256 replicas: FP64
    batch against the oracle to 1e-12 (energy: of the summed magnitudes of
    its parts), mixed batch forces against the oracle's to RMS <= 1e-5 per
    replica and the mixed potential to 1e-6 of the summed magnitudes (in
    these small boxes the FP32 mesh energy is large beside the total)."""
    reps = replicas256
    out = {}
    for prec in (M.MDG_FP64, M.MDG_MIXED):
        b = M.BatchContext(reps, precision=prec)
        b.upload_boxes()
        r0 = b.initial_energy()
        f = [b.split(a) for a in b.download()[6:9]]
        out[prec] = ([r0[r].pot for r in range(b.R)], [r0[r].nInCut for r in range(b.R)],
                     [np.stack([f[k][r] for k in range(3)]) for r in range(b.R)])
        b.close()
    worst = [0.0, 0.0, 0.0, 0.0]
    for r, wb in enumerate(reps):
        o = O.OracleSystem(wb)
        x, y, z = wb.x.copy(), wb.y.copy(), wb.z.copy()
        o.build_list(x, y, z)
        fo, eo = o.energy(x, y, z)
        p64, n64, f64 = (v[r] for v in out[M.MDG_FP64])
        pm, nm, fm = (v[r] for v in out[M.MDG_MIXED])
        assert n64 == eo["nInCut"] == nm, r
        scale = abs(eo["eDir"]) + abs(eo["eRec"]) + abs(eo["eCorr"]) + abs(eo["vdw"])
        assert abs(p64 - eo["pot"]) <= 1e-12 * scale, r
        assert relerr(f64, fo) < 1e-12, r
        worst[0] = max(worst[0], abs(p64 / eo["pot"] - 1))
        worst[1] = max(worst[1], relerr(f64, fo))
        worst[2] = max(worst[2], relerr(fm, fo))
        worst[3] = max(worst[3], abs(pm - eo["pot"]) / scale)
        assert relerr(fm, fo) <= 1e-5, r
        assert abs(pm - eo["pot"]) <= 1e-6 * scale, r
    say("batch256 worst: fp64 pot", worst[0], "fp64 f", worst[1], "mixed f rms", worst[2],
        "mixed pot / scale", worst[3])


def test_batch256_fp64_trajectories_against_oracle(replicas256):
    """The FP64 batch of all 256 replicas runs 25 steps (dt 2 fs, list
    rebuilt at step 20); eight replicas spread over the size range follow
    the oracle's MD loop to 1e-10 in energy and 1e-9 A in position."""
    reps = replicas256
    nsteps = 25
    b = M.BatchContext(reps, precision=M.MDG_FP64)
    b.upload_boxes()
    b.initial_energy()
    rb = M.batch_records(b.run(nsteps, use_graph=True), b.R)
    xb = [b.split(a) for a in b.download()[:3]]
    b.close()
    order = np.argsort([wb.n for wb in reps])
    sample = [int(order[k]) for k in np.linspace(0, len(reps) - 1, 8).astype(int)]
    for r in sample:
        wb = reps[r]
        o = O.OracleSystem(wb)
        s = [wb.x.copy(), wb.y.copy(), wb.z.copy(), wb.vx.copy(), wb.vy.copy(), wb.vz.copy()]
        err, fo, et, ep, ek, _ = o.md_run(nsteps, *s)
        assert err == 0
        dx = max(np.abs(xb[k][r] - s[k]).max() for k in range(3))
        say("batch256 replica", r, wb.n, rb["tot"][-1, r], et[-1], "pos", dx)
        assert np.abs(rb["tot"][:, r] - et[1:]).max() < 1e-10 * abs(et[0]), r
        assert np.abs(rb["pot"][:, r] - ep[1:]).max() < 1e-10 * abs(ep[0]), r
        assert dx < 1e-9, r
