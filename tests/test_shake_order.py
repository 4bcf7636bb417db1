"""The device's SHAKE stopping rule against the reference's, on the oracle.

The reference sweeps every constraint until all of them pass the test in
the same sweep C (shake.c:243-338 for positions, 370-457 for velocities).
The device iterates each constraint group (a molecule) on its own and stops
it once it has passed at two consecutive sweeps; sweeps after that count as
passed without being checked.  Its stop sweep is then the first sweep at
which every group passed or had stopped, and the second kernel raises
MDG_ESHAKE_ORDER when some group that had stopped fails the test at that
sweep.  For one system of three-site molecules the stop is per wavefront of
16 molecules (all of them at two consecutive passing sweeps at once); in a
replica batch it is per molecule.

Each constraint group's iterates depend on that group alone, so the
oracle's global sweeps give every group's pass/fail history up to C.  The
tests replay the device's rules on that history and check that they find
the same C and raise no order error over MD trajectories of the c0 (1 fs)
and c1-density (2 fs) boxes, and that the replay does raise it on a history
where a stopped group fails again.
"""
import numpy as np
import pytest

import oracle as O
from mdbas_amd import configs

WAVE = 16  # three-site molecules per wavefront (four lanes each)


def groups_of(wb):
    """Constraint -> group (connected component of the constraint graph)."""
    parent = np.arange(wb.n)

    def root(a):
        while parent[a] != a:
            parent[a] = parent[parent[a]]
            a = parent[a]
        return a

    for a, b in zip(wb.ca, wb.cb):
        ra, rb = root(int(a)), root(int(b))
        if ra != rb:
            parent[max(ra, rb)] = min(ra, rb)
    roots = np.array([root(int(a)) for a in wb.ca])
    _, g = np.unique(roots, return_inverse=True)
    return g


def per_group(conv, g):
    """[sweeps, constraints] -> [sweeps, groups]: a group passes when all
    its constraints do."""
    out = np.ones((conv.shape[0], g.max() + 1), bool)
    np.logical_and.at(out, (slice(None), g), conv)
    return out


def device_stop(cg, width=1):
    """Replays the device rule on a group history cg [sweeps, groups].

    Groups are taken `width` at a time (1: each group on its own; WAVE: a
    wavefront, whose groups stop together at the first sweep where all of
    them passed at that sweep and the one before).  Sweeps after a stop
    count as passed.  Returns (stop sweep, order error)."""
    K, G = cg.shape
    bits = cg.copy()
    for lo in range(0, G, width):
        blk = cg[:, lo:lo + width]
        both = blk[1:].all(1) & blk[:-1].all(1)
        hit = np.flatnonzero(both)
        if hit.size:
            bits[hit[0] + 2:, lo:lo + width] = True
    stop = int(np.flatnonzero(bits.all(1))[0])
    return stop, not cg[stop].all()


def trajectory_histories(cfg, nsteps):
    wb = configs.build(cfg)
    o = O.OracleSystem(wb)
    s = [wb.x.copy(), wb.y.copy(), wb.z.copy(), wb.vx.copy(), wb.vy.copy(), wb.vz.copy()]
    o.build_list(*s[:3])
    f, _ = o.energy(*s[:3])
    g = groups_of(wb)
    out = []
    for step in range(1, nsteps + 1):
        for stage in (1, 2):
            err, conv = o.vv_conv(stage, *s, f)
            assert err == 0
            out.append((step, stage, per_group(conv, g)))
            if stage == 1:
                if step % cfg["update"] == 0:
                    o.build_list(*s[:3], build=False)
                f, _ = o.energy(*s[:3])
    return out


@pytest.fixture(scope="module", params=["c0", "c1"])
def histories(request):
    cfg = configs.CONFIGS[request.param]
    return request.param, trajectory_histories(cfg, 12 if request.param == "c0" else 3)


def test_history_follows_the_global_rule(histories):
    """The recorded history stops at the first sweep where every group
    passes (the reference's rule), so its last row is that sweep."""
    _, hs = histories
    for step, stage, cg in hs:
        rows = cg.all(1)
        assert rows[-1] and not rows[:-1].any(), (step, stage)


@pytest.mark.parametrize("width", [1, WAVE])
def test_device_rule_finds_the_reference_stop(histories, width):
    name, hs = histories
    redivergent = 0
    for step, stage, cg in hs:
        stop, order = device_stop(cg, width)
        assert not order, (name, step, stage)
        assert stop == cg.shape[0] - 1, (name, step, stage, stop)
        # groups that passed twice in a row and later failed within the window
        both = cg[1:] & cg[:-1]
        for j in range(cg.shape[1]):
            h = np.flatnonzero(both[:, j])
            if h.size and not cg[h[0] + 1:, j].all():
                redivergent += 1
    print(name, "width", width, "steps x stages", len(hs), "sweeps",
          sorted({cg.shape[0] for _, _, cg in hs}), "re-diverged groups", redivergent)
    assert redivergent == 0


def test_replay_raises_on_a_redivergent_group():
    """Group 0 passes at sweeps 1 and 2, fails at 3 and 4, passes from 5;
    group 1 passes from sweep 3.  The device stops group 0 after sweep 2 and
    takes sweep 3 as the stop, where group 0 fails: the order error.  The
    reference's stop is sweep 5."""
    cg = np.array([[0, 0], [1, 0], [1, 0], [0, 1], [0, 1], [1, 1]], bool)
    assert device_stop(cg, 1) == (3, True)
    assert np.flatnonzero(cg.all(1))[0] == 5
    # as one wavefront the pair never passes at two consecutive sweeps
    # before sweep 5, so the checked bits stand and sweep 5 is taken, as
    # the reference takes it
    assert device_stop(cg, 2) == (5, False)


def test_replay_keeps_a_late_group():
    """A group that never passes twice in a row before the stop keeps its
    checked bits: the stop is the reference's."""
    cg = np.array([[0, 0], [1, 0], [0, 1], [1, 0], [1, 1]], bool)
    assert device_stop(cg, 1) == (4, False)
