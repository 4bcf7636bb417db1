"""The C host driver (mdbas_amd/mdbas_gpu, built from host/mdbas_gpu.c
with gcc against libmdgpu.so): a reference deck in, the reference's
output files out.

It runs mdbas.c's order (mdbas.c:244-363) through the C ABI, either with
the drop-in calls each step (mdg_vv_integrate stage 1 -> mdg_makelist ->
mdg_energy -> mdg_vv_integrate stage 2) or device-resident
(mdg_run_steps with snapshot writers).  The GPU tests compare its PROP
records and final REST state with the FP64 oracle's MD loop on the same
system (the deck reads back bit for bit, tests/test_io.py)."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
from mdbas_amd import configs, decks
from mdbas_amd import molecules as Mo

HERE = os.path.dirname(os.path.abspath(__file__))
DRIVER = os.path.join(os.path.dirname(HERE), "mdbas_amd", "mdbas_gpu")
KCAL = 418.4


def test_driver_built_and_prints_usage():
    assert os.path.exists(DRIVER), "build first: python -c 'import __graft_entry__ as g; g.build()'"
    r = subprocess.run([DRIVER, "--help"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0 and "-m dropin|device" in r.stdout


def test_driver_reports_simu_errors(tmp_path):
    (tmp_path / "SIMU").write_text("charmm\nbogus\nend\n")
    r = subprocess.run([DRIVER], cwd=str(tmp_path), capture_output=True, text=True, timeout=60)
    assert r.returncode == 61, r.stderr


def run(deck, *args):
    print("driver", deck, args, flush=True)
    r = subprocess.run([DRIVER, *args], cwd=deck, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, (r.returncode, r.stderr[-2000:])
    assert "Normal Termination" in open(os.path.join(deck, "OUTPUT")).read()
    return decks.read_prop(os.path.join(deck, "PROP")), decks.read_rest(os.path.join(deck, "REST"))


@pytest.fixture(scope="module")
def c0_oracle():
    wb = configs.build(configs.CONFIGS["c0"])
    o = O.OracleSystem(wb)
    s = [wb.x.copy(), wb.y.copy(), wb.z.copy(), wb.vx.copy(), wb.vy.copy(), wb.vz.copy()]
    err, fo, et, ep, ek, _ = o.md_run(25, *s)
    assert err == 0
    return wb, s, fo, et, ep


def check(prop, rest, s, et, ep, rel_e, pos):
    assert prop.shape[0] == 26 and list(prop[:, 0]) == list(range(26))
    e_scale = abs(et[0])
    assert np.abs(prop[:, 5] * KCAL - et).max() < rel_e * e_scale
    assert np.abs(prop[:, 7] * KCAL - ep).max() < rel_e * e_scale
    assert rest["step"] == 25
    dx = max(np.abs(rest[k] - s[i]).max() for i, k in enumerate("xyz"))
    print("energy", np.abs(prop[:, 5] * KCAL - et).max() / e_scale, "pos", dx, flush=True)
    assert dx < pos


@pytest.mark.gpu
@pytest.mark.parametrize("mode", [("dropin", "fp64"), ("device", "fp64"), ("device", "mixed")])
def test_driver_c0_deck_against_oracle(mode, c0_oracle, tmp_path):
    """The 3,000-atom water deck (SPME 32^3, 12 A cutoff, SHAKE, dt 1 fs),
    25 steps from a REST file, list rebuilt at step 20: PROP energies and
    the final REST positions against the oracle (FP64: 1e-10 relative
    energy, 1e-9 A; mixed: 1e-5 relative energy, 1e-4 A), one DCD frame
    every 5 steps, RCON/REST every 10."""
    wb, s, fo, et, ep = c0_oracle
    d = str(tmp_path)
    decks.write_deck(d, wb, configs.CONFIGS["c0"], 25, traj=5, resconf=10)
    prop, rest = run(d, "-m", mode[0], "-p", mode[1])
    if mode[1] == "fp64":
        check(prop, rest, s, et, ep, 1e-10, 1e-9)
        f = np.stack([rest["fx"], rest["fy"], rest["fz"]])
        assert np.sqrt(((f - fo) ** 2).sum() / (fo ** 2).sum()) < 1e-9
    else:
        check(prop, rest, s, et, ep, 1e-5, 1e-4)
    traj = open(os.path.join(d, "TRAJ"), "rb").read()
    frame = 56 + 3 * (8 + 4 * wb.n)
    assert len(traj) == 92 + 252 + 12 + 5 * frame
    # the last frame holds the step-25 positions (as float)
    last = np.frombuffer(traj[-frame + 56:], dtype="<u4")
    xs = np.frombuffer(traj[-frame + 56 + 4:-frame + 56 + 4 + 4 * wb.n], dtype="<f4")
    assert last[0] == 4 * wb.n
    assert np.array_equal(xs, rest["x"].astype(np.float32))


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["dropin", "device"])
def test_driver_molecule_deck_against_oracle(mode, tmp_path):
    """The flexible-molecule deck (bonds, Morse, Urey-Bradley, angles,
    dihedrals, impropers, 1-4 pairs at scal14 0.5, consth constraints), 25
    FP64 steps against the oracle's MD loop with every bonded term."""
    cfg = Mo.MOL_CONFIGS["flex216"]
    mb = Mo.build(cfg)
    o = O.OracleSystem(mb)
    s = [mb.x.copy(), mb.y.copy(), mb.z.copy(), mb.vx.copy(), mb.vy.copy(), mb.vz.copy()]
    err, fo, et, ep = o.md_run_full(25, *s)
    assert err == 0
    d = str(tmp_path)
    decks.write_deck(d, mb, cfg, 25, nb14=0.5)
    prop, rest = run(d, "-m", mode, "-p", "fp64")
    check(prop, rest, s, et, ep, 1e-10, 1e-9)
    # bonded energies are in the PROP record (bond, ang, ub, dihe, impr)
    assert (prop[:, 10] != 0).all() and (prop[:, 13] != 0).all()


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["dropin", "device"])
def test_driver_plain_cutoff_deck_against_oracle(mode, c0_oracle, tmp_path):
    """The same water deck with "elec switch" (keyEwald 0: the cutoff
    SWITCH functor through nonbond_energy, no SPME): 25 FP64 steps against
    the oracle's plain-path MD loop."""
    wb = c0_oracle[0]
    o = O.OracleSystem(wb)
    s = [wb.x.copy(), wb.y.copy(), wb.z.copy(), wb.vx.copy(), wb.vy.copy(), wb.vz.copy()]
    err, fo, et, ep = o.md_run_plain(4, 25, *s)
    assert err == 0
    d = str(tmp_path)
    decks.write_deck(d, wb, configs.CONFIGS["c0"], 25, elec="switch")
    prop, rest = run(d, "-m", mode, "-p", "fp64")
    check(prop, rest, s, et, ep, 1e-10, 1e-9)


ENSEMBLE_DECKS = [("nvt beren", 1, "dropin"), ("npt beren", 2, "dropin"),
                  ("nvt hoover", 3, "dropin"), ("npt hoover", 4, "dropin"),
                  ("nvt beren", 1, "device"), ("nvt hoover", 3, "device")]


@pytest.mark.gpu
@pytest.mark.parametrize("ensemble,ens,mode", ENSEMBLE_DECKS,
                         ids=[e[0].replace(" ", "_") + "_" + e[2] for e in ENSEMBLE_DECKS])
def test_driver_ensemble_deck_against_oracle(ensemble, ens, mode, c0_oracle, tmp_path):
    """The water deck with a thermostat or barostat ("ensemble nvt beren",
    "npt hoover", ..., taut 0.1 ps, taup 0.5 ps, pressure 1 bar), 15 FP64
    steps through the drop-in calls, and for the thermostats also
    device-resident (list rebuilt every 20 steps, so once at the start)
    against the oracle's MD loop with oc_vv_ens: PROP total and potential
    energy, consv and volume to 1e-10 relative, final positions to 1e-8 A
    (the device sums the kinetic energy in another order, and the bath
    feeds it back into every velocity)."""
    wb = c0_oracle[0]
    cfg = configs.CONFIGS["c0"]
    o = O.OracleSystem(wb)
    s = [wb.x.copy(), wb.y.copy(), wb.z.copy(), wb.vx.copy(), wb.vy.copy(), wb.vz.copy()]
    ndeg = 3 * wb.n - 3 - len(wb.ca)                   # utils.c:569-582
    temp0 = float(cfg["temp"])
    st = O.ens_state(ens, wb.box9, tauT=0.1, tauP=0.5, temp0=temp0,
                     kinTemp0=0.5 * temp0 * float(ndeg) * 0.83144621,
                     press0=1.0 * 6.02214129e-3)
    err, f, rec = o.md_run_ens(st, 15, *s)
    assert err == 0
    d = str(tmp_path)
    decks.write_deck(d, wb, cfg, 15, ensemble=ensemble, taut=0.1, taup=0.5, pressure=1.0)
    prop, rest = run(d, "-m", mode, "-p", "fp64")
    assert prop.shape[0] == 16
    scale = abs(rec[0, 0])
    for col, k, name in ((5, 0, "tot"), (7, 1, "pot"), (22, 3, "consv")):
        err_k = np.abs(prop[:, col] * KCAL - rec[:, k]).max() / scale
        print(name, err_k, flush=True)
        assert err_k < 1e-10, name
    assert np.abs(prop[:, 4] - rec[:, 4]).max() < 1e-10 * rec[0, 4]
    if ens in (2, 4):
        assert rec[-1, 4] != rec[0, 4]
    if ens in (3, 4):
        assert np.abs(rec[1:, 3]).min() > 0.
    dx = max(np.abs(rest[k] - s[i]).max() for i, k in enumerate("xyz"))
    print("pos", dx, flush=True)
    assert dx < 1e-8
