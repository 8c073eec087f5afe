"""Serving form: continuous batching of MD jobs (fz_server_*), and the
chunking invariances it relies on (a run split into several calls equals one
call).  Mixed precision is bitwise deterministic (fixed-point sums), so these
comparisons are exact."""
import numpy as np
import pytest

import fenzi
import systems

pytestmark = pytest.mark.gpu


def _jobs():
    # uneven sizes, minimisation lengths, run lengths and reassignment periods,
    # so that the batch's rounds cut every job at points it never sees alone
    return [
        (systems.slab(seed=1), 30, 95, 40, 101),
        (systems.water_box(seed=2), 0, 120, 0, 102),
        (systems.slab(n_chains=3, n_waters=300, L=24.0, seed=3, K=(24, 24, 24)), 45, 70, 25, 103),
        (systems.water_box(n_waters=220, L=19.0, seed=4), 12, 150, 60, 104),
    ]


def _same_energies(a, b):
    skip = {"list_builds"}
    return all(np.array_equal(a[k], b[k]) for k in a if k not in skip)


def test_minimize_split_matches_one_call(ctx):
    topo, xyz, params = systems.slab(seed=7)
    a = fenzi.System(ctx, topo, params, fenzi.MIXED, xyz=xyz)
    a.minimize(40, 0.05)
    b = fenzi.System(ctx, topo, params, fenzi.MIXED, xyz=xyz)
    b.minimize(13, 0.05)
    b.minimize(27, 0.05)
    assert np.array_equal(a.get_fixed(), b.get_fixed())
    assert a.energies()["potential"] == b.energies()["potential"]


def test_steps_split_matches_one_call(ctx):
    topo, xyz, params = systems.slab(seed=8)
    runs = []
    for split in ([24], [5, 7, 12]):
        s = fenzi.System(ctx, topo, params, fenzi.MIXED, xyz=xyz)
        s.reassign_velocities(300.0, 3, 0)
        for n in split:
            s.step(n)
        runs.append((s.get_fixed(), s.get_state()[1], s.energies()))
        s.close()
    assert np.array_equal(runs[0][0], runs[1][0])
    assert np.array_equal(runs[0][1], runs[1][1])
    assert _same_energies(runs[0][2], runs[1][2])


@pytest.mark.parametrize("prec", [fenzi.FP64, fenzi.MIXED])
def test_dynamics_after_minimize_starts_from_its_forces(ctx, prec):
    """The first half kick after a minimisation uses the forces at the
    minimised positions (same trajectory as a fresh system started there)."""
    topo, xyz, params = systems.slab(seed=9)
    a = fenzi.System(ctx, topo, params, prec, xyz=xyz)
    a.minimize(20, 0.05)
    u = a.get_fixed()
    a.reassign_velocities(300.0, 5, 0)
    a.step(3)
    b = fenzi.System(ctx, topo, params, prec)
    b.set_fixed(u)
    b.reassign_velocities(300.0, 5, 0)
    b.step(3)
    va, vb = a.get_state()[1], b.get_state()[1]
    assert systems.rel_rms(va, vb) < 1e-6


def _run_server(ctx, jobs, **kw):
    srv = fenzi.Server(ctx, kw.pop("budget", 8e9), precision=fenzi.MIXED, **kw)
    ids = [srv.submit(t, p, x, m, n, 300.0, r, seed) for (t, x, p), m, n, r, seed in jobs]
    srv.run_until_done()
    out = []
    for jid in ids:
        info = srv.info(jid)
        assert info["state"] == fenzi.JOB_DONE, info
        out.append(srv.result(jid))
    srv.close()
    return out


def test_server_result_identical_alone_and_in_mix(ctx):
    jobs = _jobs()
    alone = [_run_server(ctx, [j], chunk_steps=25)[0] for j in jobs]
    mix = _run_server(ctx, jobs, chunk_steps=25)
    for a, m, j in zip(alone, mix, jobs):
        assert np.array_equal(a["fixed"], m["fixed"])
        assert np.array_equal(a["vel"], m["vel"])
        assert _same_energies(a["energies"], m["energies"])
        assert m["energies"]["step"] == j[2]


def test_server_admission_respects_budget_and_slots(ctx):
    jobs = _jobs() + _jobs()[:2]
    est = [fenzi.estimate_bytes(t, p, fenzi.MIXED) for (t, x, p), *_ in jobs]
    budget = 2.5 * max(est)
    srv = fenzi.Server(ctx, budget, max_active=3, precision=fenzi.MIXED, chunk_steps=40)
    ids = [srv.submit(t, p, x, m, n, 300.0, r, s) for (t, x, p), m, n, r, s in jobs]
    peak = 0
    for _ in range(10000):
        a, q, d = srv.poll()
        infos = [srv.info(i) for i in ids]
        active = [k for k, i in enumerate(infos) if i["state"] in (fenzi.JOB_MINIMIZING, fenzi.JOB_DYNAMICS)]
        assert len(active) <= 3
        assert sum(est[k] for k in active) <= budget
        peak = max(peak, len(active))
        if a == 0 and q == 0:
            break
    assert peak >= 2
    for i in ids:
        info = srv.info(i)
        assert info["state"] == fenzi.JOB_DONE
        assert info["t_submit"] <= info["t_admit"] <= info["t_done"]
    srv.close()


def test_server_isolates_a_failed_job(ctx):
    jobs = _jobs()[:2]
    (t, x, p) = systems.water_box(seed=5)
    bad = dict(p, cutoff=12.0, cuton=10.0, pairlist=13.0)  # cutoff above half the box
    srv = fenzi.Server(ctx, 8e9, precision=fenzi.MIXED, chunk_steps=50)
    good = [srv.submit(tt, pp, xx, m, n, 300.0, r, s) for (tt, xx, pp), m, n, r, s in jobs]
    badid = srv.submit(t, bad, x, 10, 10, 300.0, 0, 1)
    srv.run_until_done()
    assert srv.info(badid)["state"] == fenzi.JOB_FAILED
    for g in good:
        assert srv.info(g)["state"] == fenzi.JOB_DONE
    with pytest.raises(fenzi.FenziError):
        srv.result(badid)
    srv.release(badid)
    srv.close()


def test_server_isolates_a_job_failing_mid_run(ctx):
    """A job whose SHAKE cannot converge (tolerance 1e-30) fails in its first
    dynamics round, inside the batch of every dynamics job: the batch call
    restores every member to the round's start, the members are retried one by
    one, the failing job is marked failed, and every other job ends bitwise as
    it does alone, with its full step count."""
    jobs = _jobs()[:3]
    alone = [_run_server(ctx, [j], chunk_steps=25)[0] for j in jobs]
    t, x, p = systems.water_box(n_waters=200, L=19.0, seed=6)
    srv = fenzi.Server(ctx, 8e9, precision=fenzi.MIXED, chunk_steps=25)
    ids = [srv.submit(tt, pp, xx, m, n, 300.0, r, s) for (tt, xx, pp), m, n, r, s in jobs]
    bad = srv.submit(t, dict(p, shake_tol=1e-30), x, 10, 60, 300.0, 0, 1)
    srv.run_until_done()
    assert srv.info(bad)["state"] == fenzi.JOB_FAILED
    for jid, a, j in zip(ids, alone, jobs):
        info = srv.info(jid)
        assert info["state"] == fenzi.JOB_DONE and info["steps_done"] == j[2]
        r = srv.result(jid)
        assert np.array_equal(r["fixed"], a["fixed"])
        assert np.array_equal(r["vel"], a["vel"])
        assert r["energies"]["step"] == j[2]
    srv.close()


def test_admission_estimate_bounds_the_device_footprint(ctx):
    """fz_server_estimate_bytes (the admission figure) is at least what a
    system of that job actually holds on the device, from small to large."""
    import fz_synth
    cases = [systems.water_box(seed=2), systems.slab(seed=1)]
    for n, f, seed in ((1000, 0.0, 1), (6000, 0.3, 2), (24000, 0.15, 3), (60000, 0.1, 4)):
        topo, xyz, params = fz_synth.ensemble_member(n, f, seed)
        cases.append((topo, xyz, params))
    for topo, xyz, params in cases:
        est = fenzi.estimate_bytes(topo, params, fenzi.MIXED)
        s = fenzi.System(ctx, topo, params, fenzi.MIXED, xyz=xyz)
        s.minimize(50, 0.05)  # the generated structures overlap: relax before dynamics
        s.reassign_velocities(300.0, 1, 0)
        s.step(20)
        used = s.info()["device_bytes"]
        print("admission estimate: %d atoms, %d bytes used, %d estimated" % (topo["n_atoms"], used, est))
        assert used <= est, (topo["n_atoms"], used, est)
        assert est < 3 * used, (topo["n_atoms"], used, est)
        s.close()
