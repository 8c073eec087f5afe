"""Andersen barostat (fz_set_barostat, molecular scaling of the constraint
clusters) on the GPU against the oracle's or_step_npt, through the C ABI: trajectories (positions, velocities, volume,
piston momentum) in both precision forms with list rebuilds under a changing
box; a run split into several calls equals one call bit for bit (mixed form);
a capacity overflow replayed from the call's start ends where a run without
it ends; a box collapse is refused with the state restored; the extended
energy KE + U + p_ref V + Pi^2 / 2M is kept as well as the fixed-box step keeps
KE + U.  Reference: the piston of reduce_virial
(src/cell_based_kernels.cu:403-600), whose NPT step is a stub in the
reference (src/npt_kernels.cu:26-33)."""
import numpy as np
import pytest

import fenzi
import oracle
import systems

pytestmark = pytest.mark.gpu


_relaxed = {}


def _setup(ctx, prec, seed=61, pairlist_extra=None, dt_ps=0.001):
    """A small slab relaxed by the oracle's steepest descent (computed once per
    seed: the oracle's threaded sums differ in the last bit between calls)."""
    topo, xyz, params = systems.slab(seed=seed)
    params = dict(params, dt_ps=dt_ps)
    if pairlist_extra is not None:
        params["pairlist"] = params["cutoff"] + pairlist_extra
    orc = oracle.Oracle(topo, params)
    if seed not in _relaxed:
        _relaxed[seed], _ = orc.minimize_sd(orc.to_fixed(xyz), 60, 0.05)
    u = _relaxed[seed].copy()
    s = fenzi.System(ctx, topo, params, prec)
    s.set_fixed(u)
    return topo, params, orc, u, s


def _state(s):
    b = s.barostat()
    return s.get_fixed(), s.get_state()[1], b["volume"], b["Pi"], s.energies()["step"]


def _same(a, b):
    return (np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2] and a[3] == b[3]
            and a[4] == b[4])


@pytest.mark.parametrize("prec", [fenzi.FP64, fenzi.MIXED])
def test_npt_trajectory_vs_oracle(ctx, prec):
    """40 steps with a light piston (the box moves by about 1 %) and a thin
    skin (rebuilds under the changing box), against the oracle."""
    topo, params, orc, u, s = _setup(ctx, prec, pairlist_extra=0.4)
    pref, M = 1.0, 1e-5
    v0 = orc.maxwell(u, 300.0, seed=4, counter=0)
    pc = orc.piston(u, pref, M)
    s.reassign_velocities(300.0, 4, 0)
    s.set_barostat(pref, M)
    b0 = s.energies()["list_builds"]
    u_ref, v_ref, e_ref, w_ref, rc = orc.step_npt(u, v0, 40, pc)
    assert rc == 0
    s.step(40)
    b = s.barostat()
    assert s.energies()["list_builds"] - b0 >= 2
    assert abs(pc.V - np.prod(params["box"])) > 1e-3 * pc.V           # the box moved
    tolV = 1e-10 if prec == fenzi.FP64 else 1e-6
    assert abs(b["volume"] - pc.V) < tolV * pc.V, (b["volume"], pc.V)
    assert np.allclose(b["box"], orc.box, rtol=tolV, atol=0)
    assert abs(b["Pi"] - pc.Pi) < (1e-7 if prec == fenzi.FP64 else 1e-3) * max(abs(pc.Pi), 1e-3)
    x_dev, v_dev = s.get_state()
    d = x_dev - orc.from_fixed(u_ref)
    d -= orc.box * np.round(d / orc.box)
    assert np.abs(d).max() < (1e-8 if prec == fenzi.FP64 else 1e-3)
    assert systems.rel_rms(v_dev, v_ref) < (1e-7 if prec == fenzi.FP64 else 2e-3)
    e = s.energies()
    assert e["volume"] == pytest.approx(b["volume"], rel=1e-15)
    P_ref = (2.0 * orc.kinetic(v_ref) + np.trace(w_ref)) / (3.0 * pc.V) / oracle.PU
    assert e["pressure"] == pytest.approx(P_ref, rel=1e-6 if prec == fenzi.FP64 else 1e-3, abs=1.0)
    orc.set_box(b["box"])
    res, vres = orc.constraint_residual(s.get_fixed(), v_dev)
    assert res < 1e-8
    s.close()


def test_npt_split_run_equals_one_run(ctx):
    runs = []
    for split in ([30], [7, 11, 12]):
        _, _, _, _, s = _setup(ctx, fenzi.MIXED, seed=62)
        s.reassign_velocities(300.0, 6, 0)
        s.set_barostat(1.0, 1e-5)
        for n in split:
            s.step(n)
        runs.append(_state(s))
        s.close()
    assert _same(runs[0], runs[1])


def test_npt_capacity_overflow_replay_is_bitwise(ctx):
    """A barostat run whose list outgrows a reduced capacity inside a call ends
    bitwise where the run with default capacities ends."""
    out = []
    for reduce in (False, True):
        _, _, _, _, s = _setup(ctx, fenzi.MIXED, seed=63)
        s.reassign_velocities(400.0, 9, 0)
        s.set_barostat(-500.0, 1e-5)   # tension: the box grows, then shrinks back
        s.step(10)
        if reduce:
            pairs, _ = s.cluster_pairs()
            _, _, slots = s.cluster_layout()
            bytes1 = s.info()["device_bytes"]
            s.set_neighbor_capacity(len(pairs), len(slots) // 8)
            assert s.info()["device_bytes"] < bytes1
        for _ in range(8):
            s.step(15)
        out.append((_state(s), s.energies()["list_builds"], s.info()["device_bytes"]))
        s.close()
    assert _same(out[0][0], out[1][0]) and out[0][1] == out[1][1]


def test_npt_box_collapse_is_refused_and_restored(ctx):
    _, _, _, _, s = _setup(ctx, fenzi.MIXED, seed=64)
    s.reassign_velocities(300.0, 2, 0)
    s.set_barostat(1.0, 1e-5)
    s.step(5)
    # 4e5 atm on this piston compresses the box past 0.85 of its size within
    # about 30 steps, smoothly (the box guard refuses the step that would go below)
    s.set_barostat(4e5, 1e-4)
    before = _state(s)
    with pytest.raises(fenzi.FenziError) as err:
        s.step(60)
    assert err.value.code == fenzi.ERR_STATE
    after = _state(s)
    assert np.array_equal(after[0], before[0]) and np.array_equal(after[1], before[1]) and after[2] == before[2]
    s.close()


def test_npt_extended_energy_kept_as_the_fixed_box_energy(ctx):
    """FP64 form, 300 steps at 1 fs: the extended energy's range stays within a
    small multiple of the range of KE + U in the same run without the piston."""
    ranges = []
    for npt in (False, True):
        _, _, orc, _, s = _setup(ctx, fenzi.FP64, seed=65)
        s.reassign_velocities(300.0, 3, 0)
        if npt:
            s.set_barostat(1.0, 1e-4)
        h = []
        for _ in range(30):
            s.step(10)
            e = s.energies()
            val = e["kinetic"] + e["potential"]
            if npt:
                b = s.barostat()
                val += 1.0 * oracle.PU * b["volume"] + b["Pi"] ** 2 / (2.0 * 1e-4)
            h.append(val)
        ranges.append(np.ptp(h))
        s.close()
    assert ranges[1] < 3.0 * ranges[0] + 0.05, ranges


def test_npt_refuses_fixed_atoms(ctx):
    _, _, _, _, s = _setup(ctx, fenzi.MIXED, seed=66)
    fixed = np.zeros(s.n, np.uint8)
    fixed[:5] = 1
    s.set_fixed_atoms(fixed)
    with pytest.raises(fenzi.FenziError) as err:
        s.set_barostat(1.0, 1e-4)
    assert err.value.code == fenzi.ERR_ARG
    s.set_fixed_atoms(None)
    s.set_barostat(1.0, 1e-4)
    with pytest.raises(fenzi.FenziError):
        s.set_fixed_atoms(fixed)
    s.close()
