"""Reference-format input and output (host side of libmdgpu, no GPU).

* Decks written by mdbas_amd.decks read back through mdg_read_simu /
  mdg_read_rest / mdg_read_forf / mdg_setup bit for bit equal to the box
  they were written from (water and the flexible-molecule box).
* The same decks read by the compiled reference's own read_SIMU, read_rest,
  read_CONF, read_FORF, setup, get_kinfromtemp, init_box and image_update
  (oracle/_ref, io.c and init.c built where they lie) give the same
  structs, per-atom arrays and term tables, field for field.
* The writers (RCON, PROP, REST, DCD header and frames) write the same
  bytes as the reference's write_CONF, write_prop, write_rest,
  write_DCD_header and write_DCD_traj on the same data (the DCD title
  lines hold the date and user, and REST's ATOM records a padding byte, so
  those bytes are left out).
* SIMU error numbers, the Ewald mesh of spme.c:32-73 in closed form, and
  init_vel's procedure with given deviates (constraint projection against
  the reference's own init_constvel)."""
import ctypes as C
import math
import os

import numpy as np
import pytest

import mdbas_amd as M
import oracle as O
from mdbas_amd import configs, decks
from mdbas_amd import molecules as Mo

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="reference library not built")


def small_water(n=64, seed=5, frozen_waters=0):
    cfg = dict(configs.COMMON, name="t", n_waters=n, L=(n / 0.0334) ** (1 / 3), seed=seed,
               grid=16, cutOff=5.0, cutOn=4.0, timeStep=0.001, precision="fp64")
    wb = configs.build(cfg)
    if frozen_waters:
        wb.frozen[:3 * frozen_waters] = 1
    return wb, cfg


@pytest.fixture(scope="module")
def water_deck(tmp_path_factory):
    wb, cfg = small_water()
    d = str(tmp_path_factory.mktemp("water"))
    decks.write_deck(d, wb, cfg, 10, traj=2, resconf=5)
    return d, wb, cfg


@pytest.fixture(scope="module")
def mol_deck(tmp_path_factory):
    cfg = Mo.MOL_CONFIGS["flex216"]
    mb = Mo.build(cfg)
    d = str(tmp_path_factory.mktemp("mol"))
    decks.write_deck(d, mb, cfg, 10, nb14=0.5)
    return d, mb, cfg


@pytest.fixture(scope="module")
def frozen_conf_deck(tmp_path_factory):
    """Read from CONF (no restart), five frozen waters."""
    wb, cfg = small_water(n=27, seed=9, frozen_waters=5)
    d = str(tmp_path_factory.mktemp("frozen"))
    decks.write_deck(d, wb, cfg, 4, restart=False, traj=1, seed=77)
    return d, wb, cfg


def fields(s):
    return {f[0]: getattr(s, f[0]) for f in s._fields_}


def same_struct(a, b, skip=()):
    for f in a._fields_:
        k = f[0]
        if k in skip:
            continue
        va, vb = getattr(a, k), getattr(b, k)
        if isinstance(va, float) and math.isnan(va) and math.isnan(vb):
            continue
        assert va == vb, k


# ------------------------------------------------------------------ round trip

def test_water_deck_reads_back_bit_exact(water_deck):
    d, wb, cfg = water_deck
    D = M.read_deck(d)
    p = D["param"]
    assert p.nAtom == wb.n and p.nConst == len(wb.ca) and p.nBond == 0
    assert p.nDegFree == 3 * wb.n - 3 - len(wb.ca)              # utils.c:569-582
    assert p.chargeConst == 332.0716 * 418.4                     # init.c:102-103
    for k in ("x", "y", "z", "vx", "vy", "vz", "q", "eps", "sig", "mass", "rmass"):
        assert np.array_equal(D[k], getattr(wb, k)), k
    assert np.array_equal(D["nAtConst"], wb.nAtConst)
    assert [(c.a, c.b, c.rc2) for c in D["cons"]] == list(zip(wb.ca, wb.cb, wb.rc2))
    assert D["ctrl"].keyConstH == 1 and D["ctrl"].keyEwald == 2 and D["ctrl"].keyRest == 1
    assert D["box"].a1 == wb.L and D["box"].u1 == 1.0 / wb.L


def test_molecule_deck_reads_back_bit_exact(mol_deck):
    d, mb, cfg = mol_deck
    D = M.read_deck(d)
    for k in ("x", "y", "z", "vx", "vy", "vz", "q", "eps", "sig", "eps14", "sig14", "mass",
              "rmass"):
        assert np.array_equal(D[k], getattr(mb, k)), k

    def rows(rs):
        return [tuple(getattr(r, f[0]) for f in r._fields_) for r in rs]
    assert rows(D["bond"]) == [tuple(t) for t in mb.bonds]
    assert rows(D["ub"]) == [tuple(t) for t in mb.ubs]
    assert rows(D["angle"]) == [tuple(t) for t in mb.angles]
    assert rows(D["dihe"]) == [tuple(t) for t in mb.dihes]
    assert rows(D["impr"]) == [tuple(t) for t in mb.imprs]
    assert [(c.a, c.b, c.rc2) for c in D["cons"]] == list(zip(mb.ca, mb.cb, mb.rc2))
    assert np.array_equal(D["nAtConst"], mb.nAtConst)
    assert D["ctrl"].keyNb14 == 1 and D["param"].scal14 == 0.5


# ------------------------------------------------------------------ against the reference

def ref_read(d):
    R = O.ref()
    R.ref_read_deck.restype = C.c_void_p
    R.ref_read_deck.argtypes = [C.c_char_p]
    old = os.getcwd()
    os.chdir(d)
    try:
        h = C.c_void_p(R.ref_read_deck(b"SIMU"))
    finally:
        os.chdir(old)
    S = dict(ctrl=M.MDG_CTRL(), param=M.MDG_PARAM(), bath=M.MDG_BATH(), neigh=M.MDG_NEIGH(),
             ewald=M.MDG_EWALD(), box=M.MDG_PBC(), ener=M.MDG_ENERGY(), io=M.MDG_IO())
    R.ref_deck_structs(h, *[C.byref(S[k]) for k in ("ctrl", "param", "bath", "neigh", "ewald",
                                                    "box", "ener", "io")])
    p = S["param"]
    n = p.nAtom
    atoms = (M.MDG_ATOM * n)()
    xyz = np.zeros(3 * n)
    par = np.zeros(7 * n)
    fr = np.zeros(n, np.int32)
    nat = np.zeros(n, np.int32)
    R.ref_deck_atoms(h, atoms, M.P(xyz), M.P(par), M.P(fr), M.P(nat))
    S.update(atoms=list(atoms), x=xyz[:n], y=xyz[n:2 * n], z=xyz[2 * n:], frozen=fr, nAtConst=nat)
    for k, key in enumerate(("q", "eps", "sig", "eps14", "sig14", "mass", "rmass")):
        S[key] = par[k * n:(k + 1) * n]
    T = dict(cons=(M.MDG_CONSTRAINT * max(1, p.nConst))(), bond=(Mo.MdgBond * max(1, p.nBond))(),
             ub=(Mo.MdgBond * max(1, p.nUb))(), angle=(Mo.MdgAngle * max(1, p.nAngle))(),
             dihe=(Mo.MdgDihe * max(1, p.nDihedral))(),
             impr=(Mo.MdgDihe * max(1, p.nImproper))())
    R.ref_deck_terms(h, *[T[k] for k in ("cons", "bond", "ub", "angle", "dihe", "impr")])
    cnt = dict(cons=p.nConst, bond=p.nBond, ub=p.nUb, angle=p.nAngle, dihe=p.nDihedral,
               impr=p.nImproper)
    for k in T:
        S[k] = list(T[k])[:cnt[k]]
    return S, h


@needs_ref
@pytest.mark.parametrize("which", ["water", "mol", "frozen"])
def test_readers_match_reference(which, water_deck, mol_deck, frozen_conf_deck):
    d = {"water": water_deck, "mol": mol_deck, "frozen": frozen_conf_deck}[which][0]
    mine = M.read_deck(d)
    ref, _ = ref_read(d)
    for k in ("ctrl", "param", "bath", "neigh", "ewald", "box"):
        same_struct(mine[k], ref[k])
    same_struct(mine["ener"], ref["ener"])
    for k in ("confName", "forfName", "propName", "restName", "rconName", "trajName"):
        assert getattr(mine["io"], k) == getattr(ref["io"], k)
    for a, b in zip(mine["atoms"], ref["atoms"]):
        assert (a.label, a.segn, a.resn, a.type, a.resi, a.ires) == \
               (b.label, b.segn, b.resn, b.type, b.resi, b.ires)
    for k in ("x", "y", "z", "q", "eps", "sig", "eps14", "sig14", "mass", "rmass", "frozen",
              "nAtConst"):
        assert np.array_equal(mine[k], ref[k]), k
    for k in ("cons", "bond", "ub", "angle", "dihe", "impr"):
        assert len(mine[k]) == len(ref[k]), k
        for a, b in zip(mine[k], ref[k]):
            same_struct(a, b)


@needs_ref
def test_frozen_deck_setup_drops_frozen_constraints(frozen_conf_deck):
    """setup (init.c:248-616) with frozen atoms: masses zeroed, constraints
    between two frozen atoms dropped, nFrozen and the degrees of freedom."""
    d, wb, _ = frozen_conf_deck
    D = M.read_deck(d)
    p = D["param"]
    assert p.nFrozen == 15
    assert p.nConst == len(wb.ca) - 15
    assert (D["mass"][:15] == 0).all() and (D["rmass"][:15] == 0).all()
    assert p.nDegFree == 3 * (wb.n - 15) - 3 - p.nConst


def _lib():
    return M.load()


def _write_mine(d, D, ener, xyz, vxyz, fxyz):
    """My writers on the deck D into directory d (same file names)."""
    L = _lib()
    from mdbas_amd import _InDir
    n = D["param"].nAtom
    atoms = (M.MDG_ATOM * n)(*D["atoms"])
    with _InDir(d):
        io = D["io"]
        bath = D["bath"]
        x, y, z = xyz
        vx, vy, vz = vxyz
        fx, fy, fz = fxyz
        assert L.mdg_write_conf(C.byref(io), C.byref(D["param"]), atoms, M.P(x), M.P(y),
                                M.P(z)) == 0
        assert L.mdg_write_prop(C.byref(io), C.byref(D["param"]), C.byref(ener),
                                C.byref(D["box"])) == 0
        assert L.mdg_write_rest(C.byref(io), C.byref(D["param"]), C.byref(ener), C.byref(bath),
                                atoms, *[M.P(a) for a in (x, y, z, vx, vy, vz, fx, fy, fz)]) == 0
        assert L.mdg_write_dcd_header(C.byref(io), C.byref(D["ctrl"]), C.byref(D["param"]),
                                      C.byref(D["box"]), M.P(D["frozen"])) == 0
        for a in (xyz, vxyz):
            assert L.mdg_write_dcd_traj(C.byref(io), C.byref(D["param"]), C.byref(D["box"]),
                                        *[M.P(v) for v in a], M.P(D["frozen"])) == 0


@needs_ref
def test_writers_match_reference(frozen_conf_deck, tmp_path):
    """The reference's static first-frame flag (io.c:27) makes this the only
    test that calls its DCD writer in the process."""
    d, wb, _ = frozen_conf_deck
    D = M.read_deck(d)
    n = wb.n
    rng = np.random.default_rng(3)
    xyz = [D["x"] + 0, D["y"] + 0, D["z"] + 0]
    vxyz = [rng.normal(size=n) for _ in range(3)]
    fxyz = [rng.normal(size=n) * 100 for _ in range(3)]
    ener = M.MDG_ENERGY()
    for k, f in enumerate(ener._fields_):
        setattr(ener, f[0], 1000.0 * (k + 1) + 0.125 * k)
    D["param"].step = 40
    D["bath"].chiT, D["bath"].chiP = 0.25, -0.5
    mine_dir, ref_dir = tmp_path / "mine", tmp_path / "ref"
    mine_dir.mkdir()
    ref_dir.mkdir()
    _write_mine(str(mine_dir), D, ener, xyz, vxyz, fxyz)
    _, h = ref_read(d)
    R = O.ref()
    R.ref_deck_write.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.c_void_p, C.c_double, C.c_double]
    # the reference's deck state carries step 0; PROP and REST need step 40
    old = os.getcwd()
    os.chdir(str(ref_dir))
    try:
        cat = lambda a: np.ascontiguousarray(np.concatenate(a))  # noqa: E731
        X, V, F = cat(xyz), cat(vxyz), cat(fxyz)
        R.ref_deck_set_step.argtypes = [C.c_void_p, C.c_int]
        R.ref_deck_set_step(h, 40)
        R.ref_deck_write(h, C.byref(D["io"]), C.byref(ener), X.ctypes.data, V.ctypes.data,
                         F.ctypes.data, 0.25, -0.5)
    finally:
        os.chdir(old)
    rd = lambda p, name: open(os.path.join(str(p), name), "rb").read()  # noqa: E731
    assert rd(mine_dir, "RCON") == rd(ref_dir, "RCON")
    assert rd(mine_dir, "PROP") == rd(ref_dir, "PROP")
    a, b = bytearray(rd(mine_dir, "REST")), bytearray(rd(ref_dir, "REST"))
    assert len(a) == len(b)
    for i in range(n):   # ATOM padding byte (offset 15 of each 28-byte record)
        a[8 + 28 * i + 15] = b[8 + 28 * i + 15] = 0
    assert a == b
    a, b = bytearray(rd(mine_dir, "TRAJ")), bytearray(rd(ref_dir, "TRAJ"))
    assert len(a) == len(b)
    t0 = 4 + 4 + 80 + 4 + 4 + 4   # start of the title lines
    a[t0:t0 + 240] = b[t0:t0 + 240] = bytes(240)
    assert a == b
    # the first frame holds every atom, the second only the free ones
    hdr = 92 + 252 + 12 + 8 + 4 * (n - 15)
    first = 56 + 3 * (8 + 4 * n)
    assert len(a) == hdr + first + 56 + 3 * (8 + 4 * (n - 15))


# ------------------------------------------------------------------ SIMU errors, mesh, velocities

@pytest.mark.parametrize("text,code", [
    ("charmm\nbogus 1\nend\n", 61),
    ("timestep\nend\n", 63),
    ("elec coulomb\nend\n", 62),
    ("ensemble nvt langevin\nend\n", 62),
    ("pbc 1\n10.0 0.0 0.0\n0.0 -10.0 0.0\n0.0 0.0 10.0\nend\n", 62),
    ("pbc 1\n10.0 0.0\n", 63),
])
def test_simu_errors(text, code, tmp_path):
    (tmp_path / "SIMU").write_text(text)
    with pytest.raises(M.MdgError, match=f"-> {code}"):
        M.read_deck(str(tmp_path))


def test_simu_missing_file(tmp_path):
    with pytest.raises(M.MdgError, match="-> 60"):
        M.read_deck(str(tmp_path))


def test_spme_mesh_closed_form():
    """spme.c:32-73: each edge is twice the smallest power of two >= the
    requested count (recvec r), capped at 512; without recvec the count is
    nint(0.25 + p*alpha*tol1/pi); without alpha, alpha =
    sqrt(|ln(prec*rc*tol)|)/rc with tol = sqrt(|ln(prec*rc)|)."""
    L = M.load()
    ctrl, p, e, b = M.MDG_CTRL(), M.MDG_PARAM(), M.MDG_EWALD(), M.MDG_PBC()
    p.cutOff = 12.0
    b.type, b.a1, b.b2, b.c3 = 1, 31.04, 31.04, 31.04
    L.mdg_init_box(C.byref(b))
    for req, want in ((16, 32), (17, 64), (1, 2), (3, 8), (300, 512), (1000, 512)):
        ctrl.keyMmax, ctrl.keyAlpha = 1, 1
        e.prec, e.alpha = 1e-6, 0.32
        e.m1max, e.m2max, e.m3max = req, 16, 5
        assert L.mdg_spme_mesh(C.byref(ctrl), C.byref(p), C.byref(e), C.byref(b)) == 0
        assert (e.m1max, e.m2max, e.m3max) == (want, 32, 16)
        assert e.mmax == want * 32 * 16
    ctrl.keyMmax, ctrl.keyAlpha = 0, 0
    e.prec = 1e-6
    assert L.mdg_spme_mesh(C.byref(ctrl), C.byref(p), C.byref(e), C.byref(b)) == 0
    tol = math.sqrt(abs(math.log(1e-6 * 12.0)))
    alpha = math.sqrt(abs(math.log(1e-6 * 12.0 * tol))) / 12.0
    tol1 = math.sqrt(-math.log(1e-6 * 12.0 * (2.0 * tol * alpha) ** 2))
    assert e.alpha == pytest.approx(alpha, rel=1e-15)
    m = int(0.25 + 31.04 * alpha * tol1 / math.pi + 0.5)
    p2 = 1
    while m > p2 and p2 < 256:
        p2 *= 2
    assert e.m1max == 2 * p2


@needs_ref
def test_init_vel_with_given_deviates(water_deck):
    """mdg_init_vel: velocities g*sqrt(1/m), projected onto the
    constraints as the reference's own init_constvel does (compiled, same
    input), zero total momentum, kinetic energy exactly kinTemp0."""
    d, wb, cfg = water_deck
    D = M.read_deck(d)
    L = M.load()
    n = wb.n
    g = np.random.default_rng(11).normal(size=3 * n)
    v = [np.zeros(n) for _ in range(3)]
    cons = (M.MDG_CONSTRAINT * len(D["cons"]))(*D["cons"])
    assert L.mdg_init_vel(C.byref(D["param"]), C.byref(D["box"]), cons, M.P(D["x"]),
                          M.P(D["y"]), M.P(D["z"]), *[M.P(a) for a in v], M.P(D["mass"]),
                          M.P(D["rmass"]), M.P(D["frozen"]), M.P(D["nAtConst"]), M.P(g)) == 0
    mass = D["mass"]
    kin = 0.5 * (mass * (v[0] ** 2 + v[1] ** 2 + v[2] ** 2)).sum()
    assert kin == pytest.approx(D["param"].kinTemp0, rel=1e-13)
    for a in v:
        assert abs((a * mass).sum()) < 1e-12 * np.abs(a * mass).sum()
    # the projection step against the compiled init_constvel
    s = g.reshape(3, n) * np.sqrt(D["rmass"])
    rs = O.RefSystem(wb)
    pre = [s[0].copy(), s[1].copy(), s[2].copy()]
    rs.constvel(D["x"], D["y"], D["z"], *pre)
    cm = [(a * mass).sum() / mass.sum() for a in pre]
    pre = [a - c for a, c in zip(pre, cm)]
    f = math.sqrt(D["param"].kinTemp0 / (0.5 * (mass * (pre[0] ** 2 + pre[1] ** 2 +
                                                         pre[2] ** 2)).sum()))
    for a, b in zip(v, pre):
        assert np.abs(a - b * f).max() < 1e-12 * np.abs(a).max()
    # constrained bonds carry no relative velocity along them: below the
    # projection's tolerance (tolShake) before the final scaling by f
    for c in D["cons"]:
        dv = np.array([v[k][c.b] - v[k][c.a] for k in range(3)])
        dr = np.array([D[k][c.b] - D[k][c.a] for k in "xyz"])
        assert abs(dv @ dr) / np.linalg.norm(dr) < D["param"].tolShake * f
