"""The five benchmark configurations at their own sizes under the GPU tests,
against the FP64 oracle, through the C ABI:

* the 2,400-water box (7,200 atoms, 30 A, cutoff 10, PME 32^3, beta 0.34),
  FP64, with both PME and the direct Ewald sum on the oracle side (the
  PME-vs-Ewald error reported apart from the GPU-vs-oracle error);
* neighbour layouts and (I, J, mask) lists bit for bit at 17K atoms (the fused
  column scan and the four-part search) and at 230K atoms (the separate column
  scan and the one-part search);
* the 230K-atom slab in FP64 (the FP64 brick spread);
* a batch spanning the ensemble's size range (1,500 to 24,000 atoms, brick
  spread), every member's forces after minimisation and dynamics against the
  oracle, and two members alone against their batch trajectories;
* serving jobs spanning 1,000 to 60,000 atoms whose final states' forces agree
  with the oracle;
* the adaptive steepest descent against its oracle restatement move for move.

Tolerances: FP64 form 1e-10 relative RMS on forces, mixed form 1e-5."""
import numpy as np
import pytest

import fenzi
import fz_synth
import oracle
import systems

pytestmark = pytest.mark.gpu


def _gpu_forces(ctx, topo, params, prec, u):
    s = fenzi.System(ctx, topo, params, prec)
    s.set_fixed(u)
    f, e = s.forces()
    return s, f, e


def test_water2400_vs_oracle_with_pme_and_direct_ewald(ctx):
    topo, xyz, params = fz_synth.config(0)
    orc = oracle.Oracle(topo, params)
    u = orc.to_fixed(xyz)
    s, f, e = _gpu_forces(ctx, topo, params, fenzi.FP64, u)
    f_pme, e_pme, w_pme = orc.forces(u)
    # the GPU against the same-form oracle (smooth PME, order 4, 32^3)
    assert systems.rel_rms(f, f_pme) < 1e-10
    assert abs(e["pme"] - e_pme["pme"]) < 1e-9 * abs(e_pme["pme"])
    assert np.abs(e["virial"] - w_pme).max() < 1e-8 * np.abs(w_pme).max()
    assert e["pairs_within_cutoff"] == orc.count_pairs_within(u, params["cutoff"])
    # the method against the direct Ewald sum (reciprocal parts alone, and totals)
    fr_pme, er_pme, _ = orc.forces(u, oracle.PME)
    fr_ew, er_ew, _ = orc.forces(u, oracle.EWALD)
    # the method's own error (order-4 B-splines on a 0.94 A grid at beta 0.34):
    # 3.2e-3 of the reciprocal force and 9.7e-4 of the reciprocal energy, 1.4e-5
    # of the total force on this configuration
    pme_vs_ewald = systems.rel_rms(fr_pme, fr_ew)
    assert pme_vs_ewald < 5e-3, pme_vs_ewald
    assert abs(er_pme["pme"] - er_ew["pme"]) < 2e-3 * abs(er_ew["pme"])
    f_ew_total, _, _ = orc.forces(u, oracle.ALL & ~oracle.PME | oracle.EWALD)
    total_vs_ewald = systems.rel_rms(f, f_ew_total)
    assert total_vs_ewald < 1e-4, total_vs_ewald
    print("water2400: GPU vs PME oracle %.2e, PME vs direct Ewald (reciprocal) %.2e, GPU vs Ewald oracle %.2e"
          % (systems.rel_rms(f, f_pme), pme_vs_ewald, total_vs_ewald))
    # 20 NVE steps at 1 fs against the oracle, from the box relaxed by 50
    # steepest-descent moves (the lattice start overlaps molecules: 2.66 g/cm^3)
    s.minimize(50, 0.05)
    u = s.get_fixed()
    v0 = orc.maxwell(u, 300.0, seed=0, counter=0)
    s.reassign_velocities(300.0, 0, 0)
    u_ref, v_ref, rc = orc.step(u, v0, 20)
    assert rc == 0
    s.step(20)
    # (SHAKE stops at its 1e-8 tolerance on either side: ~1e-8 A differences)
    d = (s.get_fixed() - u_ref).view(np.int64).astype(np.float64) * (params["box"][0] / 2.0 ** 64)
    assert np.abs(d).max() < 1e-6
    s.close()


@pytest.mark.parametrize("cid", [0, 1, 2])
def test_full_size_layout_and_pairs_bitexact(ctx, cid):
    """Column binning, cluster slots and the cluster-pair list with its lane masks
    equal the oracle's restatement bit for bit at the configurations' own sizes
    (C0/C1: fused column scan, search in four parts; C2: separate column scan,
    one part); the pairs inside the cutoff are counted alike."""
    topo, xyz, params = fz_synth.config(cid)
    orc = oracle.Oracle(topo, params)
    u = orc.to_fixed(xyz)
    s, _, e = _gpu_forces(ctx, topo, params, fenzi.MIXED, u)
    (ncx, ncy), col, slots = s.cluster_layout()
    pairs, masks = s.cluster_pairs()
    ocol, oslots, opairs, omasks = orc.cluster_pairs(u, ncx, ncy)
    assert np.array_equal(col, ocol)
    assert np.array_equal(slots, oslots)
    assert len(pairs) == len(opairs)
    a = np.lexsort((masks, pairs[:, 1], pairs[:, 0]))
    b = np.lexsort((omasks, opairs[:, 1], opairs[:, 0]))
    assert np.array_equal(pairs[a], opairs[b]) and np.array_equal(masks[a], omasks[b])
    assert systems.pair_count_ok(e["pairs_within_cutoff"], orc, u, params["cutoff"], True)
    s.close()


def test_slab230k_fp64_vs_oracle(ctx):
    """The denominator of the mixed-over-FP64 speed ratio: the 230K slab in the
    FP64 form (brick spread with FP64 accumulators) against the oracle."""
    topo, xyz, params = fz_synth.config(2)
    orc = oracle.Oracle(topo, params)
    u = orc.to_fixed(xyz)
    f_ref, e_ref, w_ref = orc.forces(u)
    s, f, e = _gpu_forces(ctx, topo, params, fenzi.FP64, u)
    assert systems.rel_rms(f, f_ref) < 1e-10
    for k, name in (("vdw", "vdw"), ("elec", "elec"), ("pme", "pme"), ("ewexcl", "ewald_excl"), ("bond", "bond"),
                    ("angle", "angle"), ("dihed", "dihedral")):
        assert abs(e[name] - e_ref[k]) < 1e-8 * max(1.0, abs(e_ref[k])), k
    assert np.abs(e["virial"] - w_ref).max() < 1e-8 * np.abs(w_ref).max()
    assert systems.pair_count_ok(e["pairs_within_cutoff"], orc, u, params["cutoff"], False)
    s.close()


def test_batch_over_the_ensemble_size_range_vs_oracle(ctx):
    """32 members drawn as the ensemble's (log-uniform 1,500-24,000 atoms, chain
    fraction 0-0.3, cutoff 10): one batch minimises and steps them (PME plane
    kernels with 256-thread workgroups, and the pair kernel split between the
    members with and without a rebuild); every member's forces at its final
    positions match the oracle, and the smallest and largest members run alone
    (1,024-thread plane workgroups, one pair launch) follow their batch
    trajectories bitwise."""
    n, frac = fz_synth.batch_sizes(32)
    members = [fz_synth.ensemble_member(int(n[k]), float(frac[k]), 1000 + k) for k in range(32)]
    assert sum(int(m[0]["n_atoms"]) for m in members) >= 65536
    systems_ = [fenzi.System(ctx, t, p, fenzi.MIXED, xyz=x) for t, x, p in members]
    b = fenzi.Batch(ctx, systems_)
    b.minimize(100, 0.05)
    for k, s in enumerate(systems_):
        s.reassign_velocities(300.0, 1000 + k, 0)
    b.step(40)
    b.close()
    for (t, x, p), s in zip(members, systems_):
        orc = oracle.Oracle(t, p)
        u = s.get_fixed()
        f_ref, e_ref, _ = orc.forces(u)
        f, e = s.forces()
        assert systems.rel_rms(f, f_ref) < 1e-5, t["n_atoms"]
        assert abs(e["potential"] - sum(e_ref.values())) < 1e-5 * sum(abs(v) for v in e_ref.values())
    for k in (int(np.argmin(n)), int(np.argmax(n))):
        t, x, p = members[k]
        a = fenzi.System(ctx, t, p, fenzi.MIXED, xyz=x)
        a.minimize(100, 0.05)
        a.reassign_velocities(300.0, 1000 + k, 0)
        a.step(40)
        assert np.array_equal(a.get_fixed(), systems_[k].get_fixed()), k
        a.close()
    for s in systems_:
        s.close()


def test_serving_jobs_over_the_stream_size_range_vs_oracle(ctx):
    """Jobs of 1,000 to 60,000 atoms (the serving stream's range) through the
    continuous-batching server (steepest descent, then NVT with reassignment):
    every job finishes with its full step count and its final positions carry
    forces that match the oracle."""
    sizes = [(1000, 0.0), (3500, 0.25), (12000, 0.1), (30000, 0.2), (60000, 0.05)]
    jobs = [fz_synth.ensemble_member(nn, ff, 7000 + k) for k, (nn, ff) in enumerate(sizes)]
    srv = fenzi.Server(ctx, 200e9, precision=fenzi.MIXED, chunk_steps=50)
    ids = [srv.submit(t, p, x, 120, 60 + 20 * k, 300.0, 50, 5000 + k) for k, (t, x, p) in enumerate(jobs)]
    srv.run_until_done()
    for k, (jid, (t, x, p)) in enumerate(zip(ids, jobs)):
        info = srv.info(jid)
        assert info["state"] == fenzi.JOB_DONE and info["steps_done"] == 60 + 20 * k
        r = srv.result(jid)
        orc = oracle.Oracle(t, p)
        f_ref, _, _ = orc.forces(r["fixed"])
        s, f, _ = _gpu_forces(ctx, t, p, fenzi.MIXED, r["fixed"])
        assert systems.rel_rms(f, f_ref) < 1e-5, t["n_atoms"]
        res, _ = orc.constraint_residual(r["fixed"])
        assert res < 1e-6
        s.close()
    srv.close()


def test_adaptive_minimiser_vs_oracle_move_for_move(ctx):
    """fz_minimize (the server's minimiser) against Oracle.minimize_sd, FP64
    form: the same accepted and rejected moves, the same final positions and
    step size carried to a second call."""
    topo, xyz, params = systems.slab(n_chains=3, n_waters=250, L=22.0, seed=12, K=(24, 24, 24))
    orc = oracle.Oracle(topo, params)
    u0 = orc.to_fixed(xyz)
    s = fenzi.System(ctx, topo, params, fenzi.FP64)
    s.set_fixed(u0)
    u_ref, step = orc.minimize_sd(u0, 25, 0.05)
    s.minimize(25, 0.05)
    d = (s.get_fixed() - u_ref).view(np.int64).astype(np.float64) * (params["box"][0] / 2.0 ** 64)
    assert np.abs(d).max() < 1e-8
    u_ref, step = orc.minimize_sd(u_ref, 15, 0.05, step=step)
    s.minimize(15, 0.05)
    d = (s.get_fixed() - u_ref).view(np.int64).astype(np.float64) * (params["box"][0] / 2.0 ** 64)
    assert np.abs(d).max() < 1e-8
    s.close()
