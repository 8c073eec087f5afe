"""Recovery paths of the step engine, through the C ABI: a neighbour-capacity
overflow is replayed from the call's start with larger capacities and ends
bitwise where a run without it ends; a call that fails (SHAKE) leaves every
member of the batch as it was, and the survivors then continue bitwise as if
the failing call had never happened; the autotuner gives the state back
unchanged; the device-side rebuild decision fires in a trajectory that the
oracle follows.

The reference exits on overflow (Check_Static_Parameters,
src/dynamics.cu:52-234) and on SHAKE failure; its rebuild trigger is the
half-skin test of CoordsUpdate (src/mdcuda.cu:453-458) read back by
src/dynamics.cu:727-737."""
import numpy as np
import pytest

import fenzi
import oracle
import systems

pytestmark = pytest.mark.gpu


def _state(s):
    x, v = s.get_state()
    return s.get_fixed(), v, s.energies()["step"]


def _same(a, b):
    return np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2]


def test_capacity_overflow_replay_is_bitwise_equal_to_a_run_without_it(ctx):
    """System B holds exactly the pair-list entries and clusters of its current
    list; a later build needs more, overflows inside an fz_step call, and the
    call restores, grows and replays.  A (default capacities) and B stay
    bitwise equal after every call, with the same list-build count."""
    topo, xyz, params = systems.slab(seed=31)
    sa = fenzi.System(ctx, topo, params, fenzi.MIXED, xyz=xyz)
    sb = fenzi.System(ctx, topo, params, fenzi.MIXED, xyz=xyz)
    for s in (sa, sb):
        s.reassign_velocities(400.0, 5, 0)
        s.step(10)
    pairs, _ = sb.cluster_pairs()
    _, _, slots = sb.cluster_layout()
    bytes0 = sb.info()["device_bytes"]
    sb.set_neighbor_capacity(len(pairs), len(slots) // 8)
    bytes1 = sb.info()["device_bytes"]
    assert bytes1 < bytes0
    assert _same(_state(sa), _state(sb))
    assert np.array_equal(sb.cluster_pairs()[0], pairs)
    grown = False
    for chunk in range(30):
        sa.step(20)
        sb.step(20)
        ea, eb = sa.energies(), sb.energies()
        assert _same(_state(sa), _state(sb)), chunk
        assert ea["list_builds"] == eb["list_builds"] and ea["potential"] == eb["potential"]
        grown = grown or sb.info()["device_bytes"] > bytes1
        if grown and chunk >= 3:
            break
    assert grown, "no build outgrew the reduced capacity"


def test_failed_batch_call_leaves_every_member_as_it_was(ctx):
    """A batch whose third member cannot satisfy SHAKE (tolerance 1e-30; the
    solver gives up after its iteration cap, a host-reported error) fails; the
    other members are back at their state before the call, and from there they
    continue bitwise like copies that never joined the batch."""
    cfgs = [systems.slab(seed=41), systems.water_box(seed=42)]
    members, copies = [], []
    for topo, xyz, params in cfgs:
        for group in (members, copies):
            s = fenzi.System(ctx, topo, params, fenzi.MIXED, xyz=xyz)
            s.reassign_velocities(300.0, 7, 0)
            s.step(12)
            group.append(s)
    topo, xyz, params = systems.water_box(n_waters=200, L=19.0, seed=43)
    bad = fenzi.System(ctx, topo, dict(params, shake_tol=1e-30), fenzi.MIXED, xyz=xyz)
    before = [_state(s) for s in members]
    bad_before = _state(bad)
    b = fenzi.Batch(ctx, members + [bad])
    with pytest.raises(fenzi.FenziError) as err:
        b.step(30)
    assert err.value.code == fenzi.ERR_SHAKE
    for s, st in zip(members, before):
        assert _same(_state(s), st)
    assert _same(_state(bad), bad_before)
    b.close()
    for s, c in zip(members, copies):
        s.step(30)
        c.step(30)
        assert _same(_state(s), _state(c))
        assert s.energies()["list_builds"] == c.energies()["list_builds"]


@pytest.mark.parametrize("explicit", [False, True])
def test_autotune_gives_the_state_back(ctx, explicit):
    topo, xyz, params = systems.slab(seed=46)
    if explicit:
        params = dict(params, column_grid=(5, 4))
    s = fenzi.System(ctx, topo, params, fenzi.MIXED, xyz=xyz)
    s.reassign_velocities(300.0, 8, 0)
    s.step(15)
    s.minimize(3, 0.01)
    s.reassign_velocities(300.0, 8, 1)
    st = _state(s)
    grid0 = s.cluster_layout()[0]
    tuned = s.autotune(6)
    assert _same(_state(s), st)
    assert tuned["ms_per_step"] < 1e29
    if explicit:
        assert s.cluster_layout()[0] == grid0 == (5, 4)
    s.step(5)
    assert s.energies()["step"] == st[2] + 5


@pytest.mark.parametrize("prec", [fenzi.FP64, fenzi.MIXED])
def test_rebuild_trigger_fires_in_a_trajectory_the_oracle_follows(ctx, prec):
    """Pair-list distance only 0.2 A beyond the cutoff (half-skin 0.1 A): the
    list is rebuilt many times in 40 steps at 2 fs and the trajectory still
    tracks the oracle, which evaluates every pair within the cutoff directly."""
    topo, xyz, params = systems.slab(seed=47)
    params = dict(params, pairlist=params["cutoff"] + 0.2, dt_ps=0.002)
    orc = oracle.Oracle(topo, params)
    u = orc.to_fixed(xyz)
    v0 = orc.maxwell(u, 300.0, seed=3, counter=0)
    s = fenzi.System(ctx, topo, params, prec)
    s.set_fixed(u)
    s.reassign_velocities(300.0, 3, 0)
    b0 = s.energies()["list_builds"]
    u_ref, v_ref, rc = orc.step(u, v0, 40)
    assert rc == 0
    s.step(40)
    builds = s.energies()["list_builds"] - b0
    assert builds >= 3, builds
    x_dev, v_dev = s.get_state()
    d = x_dev - orc.from_fixed(u_ref)
    d -= np.array(params["box"]) * np.round(d / np.array(params["box"]))
    assert np.abs(d).max() < (1e-8 if prec == fenzi.FP64 else 1e-3)
    assert systems.rel_rms(v_dev, v_ref) < (1e-7 if prec == fenzi.FP64 else 2e-3)
    f_ref, e_ref, _ = orc.forces(s.get_fixed())
    f, e = s.forces()
    assert systems.pair_count_ok(e["pairs_within_cutoff"], orc, s.get_fixed(), params["cutoff"], prec == fenzi.MIXED)
    assert systems.rel_rms(f, f_ref) < (1e-10 if prec == fenzi.FP64 else 1e-5)
