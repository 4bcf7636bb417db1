"""Bonded terms and 1-4 pairs: the oracle against the compiled reference's
golden vectors and the mathematics, and the GPU drop-ins (makelist,
energy) against the oracle.

System: mdbas_amd.molecules (five-atom flexible molecules with harmonic
and Morse bonds, angles, Urey-Bradley, cosine and harmonic dihedrals, a
harmonic improper and a hydrogen constraint), 216 molecules in a 30 A box
(O(N^2) list) and 729 in a 45 A box (link cells)."""
import ctypes as C
import hashlib
import os

import numpy as np
import pytest

import mdbas_amd as M
import oracle as O
from mdbas_amd import molecules as Mo

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def digest(*arrays) -> str:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def relerr(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


@pytest.fixture(scope="module")
def gb():
    return np.load(os.path.join(GOLD, "bonded.npz"))


@pytest.fixture(scope="module")
def boxes():
    return {k: Mo.build(c) for k, c in Mo.MOL_CONFIGS.items()}


# ------------------------------------------------------------------ oracle (CPU)

@pytest.mark.parametrize("name", ["flex216", "flex729"])
def test_oracle_bonded_matches_reference_golden(name, gb, boxes):
    """Exclusions, 1-4 list, pair list, bonded and 1-4 energies of the
    oracle against the compiled reference's (tests/golden/bonded.npz)."""
    wb = boxes[name]
    o = O.OracleSystem(wb)
    assert o.keyLink == int(gb[f"{name}_keylink"])
    assert digest(o.exOff, o.exLst) == str(gb[f"{name}_excl_digest"])
    assert np.array_equal(o.list14, gb[f"{name}_list14"])
    x, y, z = wb.x.copy(), wb.y.copy(), wb.z.copy()
    nl, np_, no = o.build_list(x, y, z)
    assert digest(nl, np_, no) == str(gb[f"{name}_digest"])
    f, e = o.bonded(x, y, z)
    assert np.allclose(e, gb[f"{name}_bonded_E"], rtol=1e-13, atol=1e-9)
    if name == "flex216":
        assert np.array_equal(np.stack([x, y, z]), gb[f"{name}_pos"])
        assert relerr(f, gb[f"{name}_bonded_f"]) < 1e-14
    for ew, et in ((1, 0), (0, 2), (0, 3), (0, 4)):
        f, e = o.nb14(ew, et, x, y, z)
        assert np.allclose(e, gb[f"{name}_nb14_{ew}{et}_E"], rtol=1e-13, atol=1e-9), (ew, et)
        if name == "flex216":
            assert relerr(f, gb[f"{name}_nb14_{ew}{et}_f"]) < 1e-13, (ew, et)


def _fd_check(energy, wb, atoms, h=1e-5):
    """-dE/dx by central differences at a few atoms against the forces."""
    x, y, z = wb.x.copy(), wb.y.copy(), wb.z.copy()
    f, _ = energy(x, y, z)
    worst = 0.0
    for i in atoms:
        for c, arr in enumerate((x, y, z)):
            v = arr[i]
            arr[i] = v + h
            ep = energy(x, y, z)[1]
            arr[i] = v - h
            em = energy(x, y, z)[1]
            arr[i] = v
            num = -(ep - em) / (2 * h)
            worst = max(worst, abs(num - f[c, i]) / (abs(f[c, i]) + 1e-3 * np.abs(f).max()))
    return worst


def test_oracle_bonded_forces_are_gradients(boxes):
    """Every bonded term's forces are minus the gradient of its energy
    (central differences, h = 1e-5 A) -- the mathematics behind
    internal.c's hand-written force expressions."""
    wb = boxes["flex216"]
    o = O.OracleSystem(wb)
    atoms = [0, 1, 2, 3, 4, 5 * 100 + 2, 5 * 215 + 3]
    err = _fd_check(lambda x, y, z: (o.bonded(x, y, z)[0], o.bonded(x, y, z)[1][:5].sum()),
                    wb, atoms)
    assert err < 1e-6


@pytest.mark.parametrize("mode", [(1, 0), (0, 2), (0, 3), (0, 4)])
def test_oracle_nb14_forces_are_gradients(mode, boxes):
    wb = boxes["flex216"]
    o = O.OracleSystem(wb)
    ew, et = mode
    atoms = [0, 3, 4, 5 * 37 + 3, 5 * 37 + 4]
    err = _fd_check(lambda x, y, z: (o.nb14(ew, et, x, y, z)[0], o.nb14(ew, et, x, y, z)[1][:2].sum()),
                    wb, atoms)
    assert err < 1e-6


def test_oracle_exclusions_are_the_three_bond_graph(boxes):
    """Brute force: the exclusion set is every pair within three bonds
    (bonds, constraints, angles, dihedrals, impropers all lie inside that
    set for this molecule), and the 1-4 list is the pairs exactly three
    bonds apart, each once."""
    wb = boxes["flex216"]
    o = O.OracleSystem(wb)
    n = wb.n
    adj = [set() for _ in range(n)]
    for b in wb.bonds:
        adj[b[1]].add(b[2])
        adj[b[2]].add(b[1])
    for a, b in zip(wb.ca, wb.cb):
        adj[a].add(int(b))
        adj[b].add(int(a))
    dist = {}
    for s in range(n):
        seen = {s: 0}
        frontier = [s]
        for d in (1, 2, 3):
            nxt = []
            for u in frontier:
                for v in adj[u]:
                    if v not in seen:
                        seen[v] = d
                        nxt.append(v)
            frontier = nxt
        for v, d in seen.items():
            if v > s:
                dist[(s, v)] = d
    excl = set()
    for i in range(n):
        for k in range(o.exOff[i], o.exOff[i + 1]):
            j = int(o.exLst[k])
            excl.add((min(i, j), max(i, j)))
    assert excl == set(dist)
    p14 = {(min(a, b), max(a, b)) for a, b in o.list14.reshape(-1, 2)}
    assert len(p14) == len(o.list14) // 2
    assert p14 == {p for p, d in dist.items() if d == 3}


def test_oracle_one_four_without_dihedral_is_error_110():
    """A 1-4 path with no dihedral parameters is DIHE_NONPARAM_ERROR
    (list.c:835-836): a four-atom chain of bonds."""
    ba = np.array([0, 1, 2], np.int32)
    bb = np.array([1, 2, 3], np.int32)
    off = np.zeros(5, np.int32)
    lst = np.zeros(64, np.int32)
    l14 = np.zeros(8, np.int32)
    n14 = C.c_int(0)
    e = np.zeros(0, np.int32)
    r = O.lib().oc_exclusions14(4, 3, O.P(ba), O.P(bb), 0, None, None, 0, None, None, None, 0,
                                None, 0, None, 0, O.P(off), O.P(lst), 64, O.P(l14), C.byref(n14))
    assert r == -110
    d4 = np.array([0, 1, 2, 3], np.int32)
    r = O.lib().oc_exclusions14(4, 3, O.P(ba), O.P(bb), 0, None, None, 0, None, None, None, 1,
                                O.P(d4), 0, None, 0, O.P(off), O.P(lst), 64, O.P(l14), C.byref(n14))
    assert r == 6 and n14.value == 1 and list(l14[:2]) == [0, 3]
    del e


# ------------------------------------------------------------------ GPU drop-ins

@pytest.mark.gpu
@pytest.mark.parametrize("name", ["flex216", "flex729"])
def test_makelist_bonded_bit_exact(name, boxes, gb):
    """Drop-in makelist with the full bonded topology: the pair list,
    the exclusion table and the 1-4 list equal the oracle's and the
    compiled reference's."""
    wb = boxes[name]
    o = O.OracleSystem(wb)
    x, y, z = wb.x.copy(), wb.y.copy(), wb.z.copy()
    nl, np_, no = o.build_list(x, y, z)
    ctx = M.Context(wb)
    gx, gy, gz = wb.x.copy(), wb.y.copy(), wb.z.copy()
    gnl, gnp, gno, eoff, elst = ctx.makelist(gx, gy, gz)
    print("bonded makelist", name, len(gnl), ctx.neigh.nPair14, flush=True)
    assert np.array_equal(gnl, nl) and np.array_equal(gnp, np_) and np.array_equal(gno, no)
    assert np.array_equal(eoff, o.exOff) and np.array_equal(elst, o.exLst)
    assert np.array_equal(ctx.list14, o.list14)
    assert np.array_equal(ctx.list14, gb[f"{name}_list14"])
    ctx.close()


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["flex216", "flex729"])
def test_energy_dropin_bonded_ewald(name, boxes):
    """mdg_energy (energy.c:103-205) on the Ewald path with 1-4 pairs and
    every bonded term, FP64 on the GPU, against the oracle: each ENERGY
    field to 1e-12 relative, forces to 1e-12."""
    wb = boxes[name]
    o = O.OracleSystem(wb)
    x, y, z = wb.x.copy(), wb.y.copy(), wb.z.copy()
    o.build_list(x, y, z)
    fo, eo = o.energy_full(x, y, z)
    ctx = M.Context(wb)
    gx, gy, gz = wb.x.copy(), wb.y.copy(), wb.z.copy()
    gnl, gnp, gno, _, _ = ctx.makelist(gx, gy, gz)
    f, e = ctx.energy(gx, gy, gz, gnl, gnp, gno)
    ctx.close()
    print("energy", name, e.pot, eo["pot"], relerr(f, fo), flush=True)
    for k in ("elec", "vdw", "bond", "ang", "ub", "dihe", "impr", "virbond", "virub", "virelec",
              "virvdw", "pot"):
        assert getattr(e, k) == pytest.approx(eo[k], rel=1e-12, abs=1e-9), k
    assert relerr(f, fo) < 1e-12


@pytest.mark.gpu
@pytest.mark.parametrize("elec", [2, 3, 4])
def test_energy_dropin_bonded_plain(elec, boxes):
    """mdg_energy on the plain cutoff path (keyEwald 0) with SHIFT1 /
    SHIFT2 / SWITCH Coulomb, its 1-4 pairs (nonbond14_energy) and the
    bonded terms, against the oracle."""
    wb = boxes["flex216"]
    o = O.OracleSystem(wb)
    x, y, z = wb.x.copy(), wb.y.copy(), wb.z.copy()
    nl, np_, no = o.build_list(x, y, z)
    fo = np.zeros((3, wb.n))
    vals = [C.c_double(0) for _ in range(4)]
    O.lib().oc_nonbond(C.byref(o.p), C.byref(o.b), elec, O.P(x), O.P(y), O.P(z), O.P(fo[0]),
                       O.P(fo[1]), O.P(fo[2]), O.P(wb.q), O.P(wb.eps), O.P(wb.sig),
                       O.P(nl), O.P(np_), O.P(no), *[C.byref(v) for v in vals])
    fo, e14 = o.nb14(0, elec, x, y, z, fo)
    fo, eb = o.bonded(x, y, z, fo)
    ctx = M.Context(wb)
    gx, gy, gz = wb.x.copy(), wb.y.copy(), wb.z.copy()
    gnl, gnp, gno, _, _ = ctx.makelist(gx, gy, gz)
    f, e = ctx.energy(gx, gy, gz, gnl, gnp, gno, elec_type=elec, keyEwald=0)
    ctx.close()
    assert e.elec == pytest.approx(vals[0].value + e14[0], rel=1e-12)
    assert e.vdw == pytest.approx(vals[1].value + e14[1], rel=1e-12)
    assert e.bond == pytest.approx(eb[0], rel=1e-12)
    assert e.dihe == pytest.approx(eb[3], rel=1e-12)
    assert relerr(f, fo) < 1e-12


# ------------------------------------------------------------------ device-resident

@pytest.mark.gpu
@pytest.mark.parametrize("name", ["flex216", "flex729"])
def test_device_bonded_energy(name, boxes):
    """The device-resident step with bonded terms and 1-4 pairs
    (mdg_set_bonded): step-0 record and forces against the oracle's
    energy() -- FP64 to 1e-12, mixed forces to RMS 1e-5."""
    wb = boxes[name]
    o = O.OracleSystem(wb)
    x, y, z = wb.x.copy(), wb.y.copy(), wb.z.copy()
    o.build_list(x, y, z)
    fo, eo = o.energy_full(x, y, z)
    for prec in (M.MDG_FP64, M.MDG_MIXED):
        ctx = M.Context(wb, precision=prec)
        ctx.upload(wb.x, wb.y, wb.z, wb.vx, wb.vy, wb.vz)
        r = ctx.initial_energy()
        f = np.stack(ctx.download()[6:9])
        ctx.close()
        print("device bonded", name, prec, r.pot, eo["pot"], relerr(f, fo), flush=True)
        assert r.nInCut == eo["nInCut"]
        if prec == M.MDG_FP64:
            assert r.pot == pytest.approx(eo["pot"], rel=1e-12)
            for k, kk in (("ebond", "bond"), ("eub", "ub"), ("eang", "ang"), ("edihe", "dihe"),
                          ("eimpr", "impr"), ("elec", "elec"), ("vdw", "vdw")):
                assert getattr(r, k) == pytest.approx(eo[kk], rel=1e-12, abs=1e-9), k
            assert relerr(f, fo) < 1e-12
        else:
            assert relerr(f, fo) <= 1e-5
            assert r.pot == pytest.approx(eo["pot"], rel=5e-6)


@pytest.mark.gpu
def test_device_bonded_trajectory(boxes):
    """25 device steps (FP64, dt 1 fs, SHAKE on the B-E bonds, list rebuilt
    at step 20) of the O(N^2) molecule box against the oracle's MD loop
    with every bonded term: energies to 1e-10, positions to 1e-9 A."""
    wb = boxes["flex216"]
    nsteps = 25
    ctx = M.Context(wb, precision=M.MDG_FP64)
    ctx.upload(wb.x, wb.y, wb.z, wb.vx, wb.vy, wb.vz)
    ctx.initial_energy()
    recs = M.records_to_dict(ctx.run(nsteps, use_graph=True))
    g = ctx.download()
    ctx.close()
    o = O.OracleSystem(wb)
    s = [wb.x.copy(), wb.y.copy(), wb.z.copy(), wb.vx.copy(), wb.vy.copy(), wb.vz.copy()]
    err, fo, et, ep = o.md_run_full(nsteps, *s)
    assert err == 0
    dx = max(np.abs(g[k] - s[k]).max() for k in range(3))
    print("bonded traj", recs["tot"][-1], et[-1], "pos", dx, flush=True)
    assert np.abs(recs["tot"] - et[1:]).max() < 1e-10 * abs(et[0])
    assert np.abs(recs["pot"] - ep[1:]).max() < 1e-10 * abs(ep[0])
    assert dx < 1e-9
