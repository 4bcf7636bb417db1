"""CPU pins of the oracle's second derivatives and trajectory adjoint.

oc_hvp (H u and d/dtheta (grad E . u) of the Ewald-path potential) and
oc_traj_adjoint (dJ/d(q, eps, sig) of the time-averaged potential through
VV + SHAKE) are checked against what the mathematics fixes, on a
216-water box small enough for many oracle runs:

* H u against central differences of the oracle's own forces along u
  (with the mean-force removal of spme.c:648-670 accounted for), and the
  symmetry v . H u = u . H v;
* the SPME part alone against differences of the reciprocal gradient;
* the mixed terms against differences of grad E . u over the parameters;
* the trajectory adjoint against central differences of J itself, along
  random parameter directions, with SHAKE converged to 1e-12 so that J is
  smooth at the step sizes used.
"""
import numpy as np
import pytest

import oracle as O
from mdbas_amd import configs

SMALL = dict(configs.COMMON, name="tip3p_216", n_waters=216, L=18.64, seed=7, grid=20,
             cutOff=8.0, cutOn=7.0, delr=1.0, alpha=0.35, timeStep=0.001, tolShake=1e-12,
             maxCycle=1000, precision="fp64")


@pytest.fixture(scope="module")
def small():
    wb = configs.build(SMALL)
    o = O.OracleSystem(wb)
    x, y, z = wb.x.copy(), wb.y.copy(), wb.z.copy()
    o.build_list(x, y, z)
    return wb, o, (x, y, z)


def raw_grad(o, x, y, z):
    """grad E = -(F + m), m the mean raw reciprocal force the reference
    removes."""
    f, _ = o.energy(x, y, z)
    _, _, g = o.spme_hvp(x, y, z)
    m = -g.mean(axis=1, keepdims=True)
    return -(f + m)


def shifted(pos, u, h):
    return [pos[c] + h * u[c] for c in range(3)]


def test_hvp_matches_force_differences(small):
    wb, o, pos = small
    rng = np.random.default_rng(3)
    u = rng.normal(size=(3, wb.n))
    hu = o.hvp(*pos, u)[0]
    h = 1e-6
    gp = raw_grad(o, *shifted(pos, u, h))
    gm = raw_grad(o, *shifted(pos, u, -h))
    fd = (gp - gm) / (2 * h)
    err = np.sqrt(((fd - hu) ** 2).sum() / (hu ** 2).sum())
    print("H u vs differences of grad E:", err)
    assert err < 1e-6


def test_hvp_is_symmetric(small):
    wb, o, pos = small
    rng = np.random.default_rng(4)
    u, v = rng.normal(size=(3, wb.n)), rng.normal(size=(3, wb.n))
    hu, hv = o.hvp(*pos, u)[0], o.hvp(*pos, v)[0]
    a, b = (v * hu).sum(), (u * hv).sum()
    assert abs(a - b) < 1e-11 * max(abs(a), np.abs(v * hu).sum())


def test_spme_hvp_matches_reciprocal_differences(small):
    wb, o, pos = small
    rng = np.random.default_rng(5)
    u = rng.normal(size=(3, wb.n))
    hu, mq, _ = o.spme_hvp(*pos, u)
    h = 1e-6
    gp = o.spme_hvp(*shifted(pos, u, h))[2]
    gm = o.spme_hvp(*shifted(pos, u, -h))[2]
    fd = (gp - gm) / (2 * h)
    err = np.sqrt(((fd - hu) ** 2).sum() / (hu ** 2).sum())
    print("H_rec u vs differences:", err)
    assert err < 1e-6


@pytest.mark.parametrize("which", ["q", "eps", "sig"])
def test_mixed_terms_match_parameter_differences(small, which):
    wb, o, pos = small
    rng = np.random.default_rng(6)
    u = rng.normal(size=(3, wb.n))
    hu, mq, me, ms = o.hvp(*pos, u)
    base = {"q": wb.q, "eps": wb.eps, "sig": wb.sig}
    d = rng.normal(size=wb.n) * (base[which] != 0)
    mixed = {"q": mq, "eps": me, "sig": ms}[which]
    h = 1e-6 * np.abs(base[which]).max()

    def gdotu(s):
        kw = {which: base[which] + s * d}
        o.set_params(**kw)
        return (raw_grad(o, *pos) * u).sum()

    try:
        fd = (gdotu(h) - gdotu(-h)) / (2 * h)
    finally:
        o.set_params()
    an = (mixed * d).sum()
    print(which, an, fd)
    assert abs(an - fd) < 1e-6 * abs(an)


def J_of(o, wb, T):
    s = [wb.x.copy(), wb.y.copy(), wb.z.copy(), wb.vx.copy(), wb.vy.copy(), wb.vz.copy()]
    err, f, et, ep, ek, _ = o.md_run(T, *s)
    assert err == 0
    return ep[1:].mean()


@pytest.mark.parametrize("which", ["q", "eps", "sig"])
def test_traj_adjoint_matches_differences_of_J(small, which):
    """Four steps, 1 fs: dJ . d from the adjoint against (J(+h) - J(-h))/2h
    along a random direction d of one parameter set, to 1e-7."""
    wb, _, _ = small
    T = 4
    o = O.OracleSystem(wb)
    J, dq, de, ds = o.traj_adjoint(T, wb.x, wb.y, wb.z, wb.vx, wb.vy, wb.vz)
    assert J == pytest.approx(J_of(o, wb, T), rel=1e-13)
    rng = np.random.default_rng({"q": 11, "eps": 12, "sig": 13}[which])
    base = {"q": wb.q, "eps": wb.eps, "sig": wb.sig}[which]
    d = rng.normal(size=wb.n) * (base != 0)
    grad = {"q": dq, "eps": de, "sig": ds}[which]
    # steps at the plateau of each difference's convergence (measured over
    # 1e-7..1e-3 of the largest parameter): below it J's rounding (a sum
    # of ~1e5 terms of both signs) dominates, above it the curvature
    h = {"q": 1e-5, "eps": 1e-4, "sig": 1e-6}[which] * np.abs(base).max()
    vals = []
    for s in (h, -h):
        o.set_params(**{which: base + s * d})
        vals.append(J_of(o, wb, T))
    o.set_params()
    fd = (vals[0] - vals[1]) / (2 * h)
    an = (grad * d).sum()
    print(which, an, fd, abs(an - fd) / abs(an))
    assert abs(an - fd) < 1e-7 * abs(an)
