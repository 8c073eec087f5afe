"""GPU parity: libfenzi's HIP path against the FP64 oracle, through the C ABI."""
import numpy as np
import pytest

import os

import fenzi
import fz_synth
import oracle
import systems

pytestmark = pytest.mark.gpu


def _pme_numpy(q, box, beta):
    """Reference PME convolution with numpy.fft (a library routine)."""
    K2, K1, K0 = q.shape
    V = box[0] * box[1] * box[2]
    Q = np.fft.fftn(q)
    m = []
    bsp = []
    for K, L in ((K2, box[2]), (K1, box[1]), (K0, box[0])):
        j = np.arange(K)
        jj = np.where(j <= K // 2, j, j - K)
        m.append(jj / L)
        bsp.append(9.0 / (2.0 + np.cos(2 * np.pi * j / K)) ** 2)
    mz, my, mx = np.meshgrid(m[0], m[1], m[2], indexing="ij")
    m2 = mx ** 2 + my ** 2 + mz ** 2
    B = bsp[0][:, None, None] * bsp[1][None, :, None] * bsp[2][None, None, :]
    with np.errstate(divide="ignore", invalid="ignore"):
        G = 332.0716 / (np.pi * V) * np.exp(-np.pi ** 2 * m2 / beta ** 2) / m2 * B
    G[0, 0, 0] = 0.0
    E = 0.5 * (G * np.abs(Q) ** 2).sum()
    phi = np.real(np.fft.ifftn(G * Q)) * q.size
    return phi, E


@pytest.mark.parametrize("K", [(16, 16, 16), (12, 20, 30), (48, 48, 72), (32, 27, 25), (128, 40, 36)])
@pytest.mark.parametrize("prec", [fenzi.FP64, fenzi.MIXED])
def test_pme_fft_identity_roundtrip(ctx, K, prec):
    rng = np.random.default_rng(sum(K))
    q = rng.standard_normal((K[2], K[1], K[0]))
    phi, _ = ctx.pme_convolve(q, (30.0, 31.0, 32.0), 0.3, prec, mode=1)
    tol = 1e-12 if prec == fenzi.FP64 else 2e-6
    err = np.abs(phi - q * q.size).max() / (np.abs(q).max() * q.size)
    assert err < tol, err


@pytest.mark.parametrize("K", [(16, 16, 16), (24, 20, 30), (48, 48, 72)])
@pytest.mark.parametrize("prec", [fenzi.FP64, fenzi.MIXED])
def test_pme_convolution_vs_numpy(ctx, K, prec):
    rng = np.random.default_rng(7 + K[0])
    q = rng.standard_normal((K[2], K[1], K[0]))
    box = (25.0, 27.0, 33.0)
    phi, E = ctx.pme_convolve(q, box, 0.32, prec, mode=0)
    phi_ref, E_ref = _pme_numpy(q, box, 0.32)
    tol = 1e-11 if prec == fenzi.FP64 else 2e-6
    assert np.abs(phi - phi_ref).max() / np.abs(phi_ref).max() < tol
    assert abs(E - E_ref) / abs(E_ref) < tol


def _system(topo, params, prec, ctx, xyz, orc):
    s = fenzi.System(ctx, topo, params, prec)
    s.set_fixed(orc.to_fixed(xyz))
    return s


@pytest.mark.parametrize("make", [systems.water_box, systems.slab])
def test_cluster_layout_and_pairs_bitexact(ctx, make):
    topo, xyz, params = make()
    orc = oracle.Oracle(topo, params)
    u = orc.to_fixed(xyz)
    s = _system(topo, params, fenzi.FP64, ctx, xyz, orc)
    (ncx, ncy), col, slots = s.cluster_layout()
    pairs, masks = s.cluster_pairs()
    ocol, oslots, opairs, omasks = orc.cluster_pairs(u, ncx, ncy)
    assert np.array_equal(col, ocol)
    assert np.array_equal(slots, oslots)
    key = lambda p, m: sorted(zip(p[:, 0].tolist(), p[:, 1].tolist(), m.tolist()))
    assert key(pairs, masks) == key(opairs, omasks)


@pytest.mark.parametrize("make", [systems.water_box, systems.slab])
def test_pair_list_covers_every_pair_within_cutoff(ctx, make):
    """Every non-excluded pair within the cutoff appears in exactly one tile lane."""
    topo, xyz, params = make()
    orc = oracle.Oracle(topo, params)
    u = orc.to_fixed(xyz)
    s = _system(topo, params, fenzi.MIXED, ctx, xyz, orc)
    _, e = s.forces()
    assert e["pairs_within_cutoff"] == orc.count_pairs_within(u, params["cutoff"])


@pytest.mark.parametrize("make", [systems.water_box, systems.slab])
@pytest.mark.parametrize("prec", [fenzi.FP64, fenzi.MIXED])
def test_forces_energies_virial_vs_oracle(ctx, make, prec):
    topo, xyz, params = make()
    orc = oracle.Oracle(topo, params)
    u = orc.to_fixed(xyz)
    f_ref, e_ref, vir_ref = orc.forces(u)
    s = _system(topo, params, prec, ctx, xyz, orc)
    f, e = s.forces()
    # tolerances: FP64 form forces 1e-10 relative RMS, energy terms 1e-9 relative
    # (absolute floor 1e-6 kcal/mol: energies are summed in 2^-28 fixed point);
    # mixed form 1e-5 for both
    ftol, etol = (1e-10, 1e-9) if prec == fenzi.FP64 else (1e-5, 1e-5)
    assert systems.rel_rms(f, f_ref) < ftol
    names = dict(bond="bond", angle="angle", ub="urey_bradley", impr="improper", dihed="dihedral",
                 vdw="vdw", elec="elec", pme="pme", ewexcl="ewald_excl", ewself="ewald_self")
    scale = sum(abs(v) for v in e_ref.values())
    for k, v in e_ref.items():
        assert abs(e[names[k]] - v) <= max(1e-6, etol * max(abs(v), 1e-3 * scale)), (k, e[names[k]], v)
    assert np.abs(e["virial"] - vir_ref).max() <= (1e-8 if prec == fenzi.FP64 else 1e-5) * np.abs(vir_ref).max()


@pytest.mark.parametrize("prec", [fenzi.FP64, fenzi.MIXED])
def test_md_steps_track_oracle(ctx, prec):
    topo, xyz, params = systems.slab()
    orc = oracle.Oracle(topo, params)
    u = orc.to_fixed(xyz)
    v0 = orc.maxwell(u, 300.0, seed=3, counter=0)
    s = _system(topo, params, prec, ctx, xyz, orc)
    s.reassign_velocities(300.0, 3, 0)
    _, v_dev = s.get_state()
    # same Philox draw; differences are libm ulps and the 2^-28 fixed-point momentum sum
    assert systems.rel_rms(v_dev, v0) < 1e-10
    u_ref, v_ref, rc = orc.step(u, v0, 10)
    assert rc == 0
    s.step(10)
    x_dev, v_dev = s.get_state()
    x_ref = orc.from_fixed(u_ref)
    d = x_dev - x_ref
    d -= np.array(params["box"]) * np.round(d / np.array(params["box"]))
    tol = 1e-9 if prec == fenzi.FP64 else 2e-5
    assert np.abs(d).max() < tol
    assert systems.rel_rms(v_dev, v_ref) < (1e-8 if prec == fenzi.FP64 else 1e-4)
    res, vres = orc.constraint_residual(s.get_fixed(), v_dev)
    assert res < 1e-9 and vres < 1e-6


def test_mixed_is_bitwise_reproducible(ctx):
    topo, xyz, params = systems.slab()
    runs = []
    for _ in range(2):
        s = fenzi.System(ctx, topo, params, fenzi.MIXED, xyz=xyz)
        s.reassign_velocities(300.0, 9, 0)
        s.step(25)
        runs.append((s.get_fixed(), s.get_state()[1]))
        s.close()
    assert np.array_equal(runs[0][0], runs[1][0])
    assert np.array_equal(runs[0][1], runs[1][1])


def test_batch_matches_alone_bitwise(ctx):
    cfgs = [systems.slab(seed=1), systems.water_box(seed=2), systems.slab(n_chains=3, n_waters=300, L=24.0, seed=3,
                                                                        K=(24, 24, 24))]
    alone = []
    for topo, xyz, params in cfgs:
        s = fenzi.System(ctx, topo, params, fenzi.MIXED, xyz=xyz)
        s.reassign_velocities(300.0, 4, 0)
        s.step(20)
        alone.append(s.get_fixed())
        s.close()
    mix = []
    for topo, xyz, params in cfgs:
        s = fenzi.System(ctx, topo, params, fenzi.MIXED, xyz=xyz)
        s.reassign_velocities(300.0, 4, 0)
        mix.append(s)
    b = fenzi.Batch(ctx, mix)
    b.step(20)
    for s, ref in zip(mix, alone):
        assert np.array_equal(s.get_fixed(), ref)
    # per-phase timing of the batch (eager launches) advances every member
    ph = b.phase_times(2)
    assert ph["pair"] > 0.0 and all(v >= 0.0 for v in ph.values())
    assert all(s.energies()["step"] == 22 for s in mix)
    b.close()


def test_second_engine_over_a_built_system_steps_identically(ctx):
    """A batch of one over a system whose list another engine built starts its
    dynamics without a rebuild; it lays out its own pair-kernel work first and
    continues the trajectory bitwise."""
    topo, xyz, params = systems.slab(seed=1)
    a = fenzi.System(ctx, topo, params, fenzi.MIXED, xyz=xyz)
    a.reassign_velocities(300.0, 4, 0)
    a.step(10)
    ref = a.get_fixed()
    a.close()
    b = fenzi.System(ctx, topo, params, fenzi.MIXED, xyz=xyz)
    b.reassign_velocities(300.0, 4, 0)
    b.step(5)
    one = fenzi.Batch(ctx, [b])
    one.step(5)
    assert np.array_equal(b.get_fixed(), ref)
    one.close()
    b.close()


def test_brick_atomic_and_plane_spread_give_the_same_trajectory(ctx):
    """Three PME paths: a launch over 64K atoms or more spreads charges by grid
    bricks, a smaller one by global atomics, and systems up to 64K atoms by
    default by the plane kernels (spread and x, y transforms per z plane in
    LDS).  Three ~23K-atom systems run alone (planes, then atomics) and as one
    batch (bricks) follow bitwise identical mixed-form trajectories."""
    # PME grids that are not multiples of the brick edge: blocks wrap across the boundary
    cfgs = [systems.slab(n_chains=6, n_waters=7700, L=62.0, seed=20 + k, K=(60, 54, 50)) for k in range(3)]
    assert sum(int(c[0]["n_atoms"]) for c in cfgs) >= 65536
    assert max(int(c[0]["n_atoms"]) for c in cfgs) < 65536
    runs = {}
    for path in (fenzi.PME_PLANES, fenzi.PME_GRID):
        runs[path] = []
        for topo, xyz, params in cfgs:
            s = fenzi.System(ctx, topo, params, fenzi.MIXED, xyz=xyz)
            eff = s.set_launch(pme_path=fenzi.PME_AUTO)
            assert eff["pme_path"] == fenzi.PME_PLANES
            assert s.set_launch(pme_path=path)["pme_path"] == path
            s.reassign_velocities(300.0, 6, 0)
            s.step(10)
            runs[path].append(s.get_fixed())
            s.close()
    mix = []
    for topo, xyz, params in cfgs:
        s = fenzi.System(ctx, topo, params, fenzi.MIXED, xyz=xyz)
        s.reassign_velocities(300.0, 6, 0)
        mix.append(s)
    b = fenzi.Batch(ctx, mix)
    assert b.set_launch(pme_path=fenzi.PME_GRID)["pme_path"] == fenzi.PME_GRID
    b.step(10)
    for k, s in enumerate(mix):
        assert np.array_equal(runs[fenzi.PME_PLANES][k], runs[fenzi.PME_GRID][k])
        assert np.array_equal(s.get_fixed(), runs[fenzi.PME_GRID][k])
    b.close()


@pytest.mark.parametrize("prec", [fenzi.FP64, fenzi.MIXED])
@pytest.mark.parametrize("make", ["water_box", "slab"])
def test_plane_kernels_match_the_grid_passes(ctx, prec, make):
    """The PME plane kernels against the spread + x/y passes on the same
    positions: the mixed form bitwise (the same integer grid and line
    transforms), the FP64 form to rounding (its LDS sums run in another order);
    then a few steps on each path, bitwise in the mixed form."""
    topo, xyz, params = getattr(systems, make)()
    out = {}
    for path in (fenzi.PME_GRID, fenzi.PME_PLANES):
        s = fenzi.System(ctx, topo, params, prec, xyz=xyz)
        assert s.set_launch(pme_path=path)["pme_path"] == path
        f, e = s.forces()
        s.reassign_velocities(300.0, 4, 0)
        s.step(6)
        out[path] = (f, e["pme"], e["virial"], s.get_fixed())
        s.close()
    (fa, ea, va, ua), (fb, eb, vb, ub) = out[fenzi.PME_GRID], out[fenzi.PME_PLANES]
    if prec == fenzi.MIXED:
        assert np.array_equal(fa, fb)
        assert ea == eb and np.array_equal(va, vb)
        assert np.array_equal(ua, ub)
    else:
        assert systems.rel_rms(fb, fa) < 1e-13
        assert abs(eb - ea) <= 1e-12 * abs(ea)
        assert np.max(np.abs(vb - va)) <= 1e-11 * np.max(np.abs(va))


def test_plane_path_refused_where_a_plane_does_not_fit(ctx):
    """FZ_PME_PLANES on a grid whose plane needs more than 64 KiB of LDS is an
    argument error; the automatic choice then keeps the grid passes."""
    topo, xyz, params = systems.water_box()
    big = dict(params, K=(96, 96, 24))
    s = fenzi.System(ctx, topo, big, fenzi.MIXED, xyz=xyz)
    with pytest.raises(fenzi.FenziError):
        s.set_launch(pme_path=fenzi.PME_PLANES)
    assert s.set_launch()["pme_path"] == fenzi.PME_GRID
    s.close()


@pytest.mark.parametrize("prec", [fenzi.FP64, fenzi.MIXED])
def test_minimize_lowers_energy_and_keeps_constraints(ctx, prec):
    topo, xyz, params = systems.slab()
    orc = oracle.Oracle(topo, params)
    s = fenzi.System(ctx, topo, params, prec, xyz=xyz)
    _, e0 = s.forces()
    s.minimize(100, 0.05)
    f1, e1 = s.forces()
    assert e1["potential"] < e0["potential"]
    res, _ = orc.constraint_residual(s.get_fixed())
    assert res < 1e-9
    # energies after minimisation agree with the oracle at the new positions
    f_ref, e_ref, _ = orc.forces(s.get_fixed())
    assert systems.rel_rms(f1, f_ref) < (1e-10 if prec == fenzi.FP64 else 1e-5)


@pytest.mark.parametrize("prec,tol", [(fenzi.FP64, 1e-7), (fenzi.MIXED, 1e-4)])
@pytest.mark.parametrize("stages", [
    [(1, 2e-6, 1e-6, 40)],
    [(2, 1e-10, 3e-6, 40)],
    # a second call whose conjugate-gradient factor overshoots: one move is undone
    # and the schedule continues by steepest descent
    [(1, 5e-5, 2e-6, 30), (1, 1.5e-3, 2e-4, 40)],
])
def test_reference_minimiser_schedule_vs_oracle(ctx, prec, tol, stages):
    """fz_minimize_ex (conjugate gradient -> steepest descent schedule) follows
    the oracle's restatement move for move: same final positions."""
    topo, xyz, params = systems.slab(n_chains=3, n_waters=250, L=22.0, seed=12, K=(24, 24, 24))
    orc = oracle.Oracle(topo, params)
    u_ref = orc.to_fixed(xyz)
    s = fenzi.System(ctx, topo, params, prec)
    s.set_fixed(u_ref)
    _, e0 = s.forces()
    for method, cgfac, sdfac, n in stages:
        s.minimize_ex(n, fenzi.min_params(method, cgfac=cgfac, sdfac=sdfac, cgrate=0.1, sdrate=0.01))
        u_ref, _, done = orc.minimize_ref(u_ref, n, method, cgfac, sdfac, 0.1, 0.01)
        assert done > 0
        d = (s.get_fixed() - u_ref).view(np.int64).astype(np.float64) * (params["box"][0] / 2.0 ** 64)
        assert np.abs(d).max() < tol
    _, e1 = s.forces()
    assert e1["potential"] < e0["potential"]
    assert np.abs(s.get_state()[1]).max() == 0.0
    s.close()


@pytest.mark.parametrize("prec", [fenzi.FP64, fenzi.MIXED])
def test_reference_deck_forces_vs_oracle(ctx, prec, tmp_path):
    """The reference's own dmpc_small input (CHARMM lipid + SPC/Fw water, with
    Urey-Bradley, impropers, NBFIX and multi-term dihedrals) through the parser,
    GPU vs oracle."""
    deck = fenzi.Deck(systems.dmpc_small_deck(str(tmp_path)))
    topo, params, xyz = deck.topology(), deck.params(), deck.coordinates()
    orc = oracle.Oracle(topo, params)
    u = orc.to_fixed(xyz)
    f_ref, e_ref, vir_ref = orc.forces(u)
    s = fenzi.System(ctx, topo, params, prec)
    s.set_fixed(u)
    f, e = s.forces()
    assert systems.rel_rms(f, f_ref) < (1e-10 if prec == fenzi.FP64 else 1e-5)
    epot_ref = sum(e_ref.values())
    assert abs(e["potential"] - epot_ref) < (1e-8 if prec == fenzi.FP64 else 1e-5) * sum(abs(v) for v in e_ref.values())
    assert systems.pair_count_ok(e["pairs_within_cutoff"], orc, u, params["cutoff"], prec == fenzi.MIXED)


@pytest.mark.parametrize("cid,prec", [(1, fenzi.FP64), (1, fenzi.MIXED), (2, fenzi.MIXED)])
def test_full_size_workloads_vs_oracle(ctx, cid, prec):
    """The 17K and 230K slabs at their full size (generated coordinates)."""
    topo, xyz, params = fz_synth.config(cid)
    orc = oracle.Oracle(topo, params)
    u = orc.to_fixed(xyz)
    f_ref, e_ref, vir_ref = orc.forces(u)
    s = fenzi.System(ctx, topo, params, prec)
    s.set_fixed(u)
    f, e = s.forces()
    assert systems.rel_rms(f, f_ref) < (1e-10 if prec == fenzi.FP64 else 1e-5)
    assert np.abs(e["virial"] - vir_ref).max() < (1e-8 if prec == fenzi.FP64 else 1e-5) * np.abs(vir_ref).max()
    for k, name in (("vdw", "vdw"), ("elec", "elec"), ("pme", "pme"), ("ewexcl", "ewald_excl")):
        assert abs(e[name] - e_ref[k]) < (1e-8 if prec == fenzi.FP64 else 1e-5) * max(1.0, abs(e_ref[k])), k


def test_fenzi_driver_runs_a_deck(tmp_path):
    import subprocess
    topo, xyz, params = systems.slab()
    path = fz_synth.write_deck(topo, xyz, params, str(tmp_path), "drv", steps=40, output_every=20, reassign_every=20)
    exe = os.path.join(os.path.dirname(fenzi.LIB_PATH), "fenzi")
    out = subprocess.run([exe, "-d", "0", "-f", path, "-p", "mixed"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.splitlines()
    assert lines[0] == fenzi.energy_header().rstrip("\n")
    rows = [l.split() for l in lines[1:] if l and l[0].isdigit()]
    assert [int(r[0]) for r in rows] == [0, 20, 40]
    assert all(len(r) == 16 for r in rows)
    assert "Ns per day" in out.stdout


def test_fenzi_driver_checkpoints_restarts_and_writes_a_trajectory(tmp_path):
    import subprocess
    topo, xyz, params = systems.water_box(seed=21)
    path = fz_synth.write_deck(topo, xyz, params, str(tmp_path), "ck", steps=40, output_every=20, reassign_every=0)
    prefix = str(tmp_path / "ck_run")
    with open(path, "a") as f:
        f.write("outputtrajectory %s.dcd\nprintlevel 2\ncheckpointTiming 20\ntrajfrequencyTiming 20\n"
                "checkpointformat binary\n" % prefix)
    exe = os.path.join(os.path.dirname(fenzi.LIB_PATH), "fenzi")
    out = subprocess.run([exe, "-d", "0", "-f", path, "-p", "mixed"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    n = topo["n_atoms"]
    r40 = fenzi.restart_read(prefix + ".rst", n, binary=True)
    assert r40["step"] == 40
    raw = open(prefix + "_dev0.dcd", "rb").read()
    assert len(raw) == (8 + 84) + (8 + 84) + (8 + 4) + 2 * ((8 + 48) + 3 * (8 + 4 * n))
    # a second run with a later end step resumes from the checkpoint
    txt = open(path).read().replace("dynasteps 40", "dynasteps 60")
    open(path, "w").write(txt)
    out = subprocess.run([exe, "-d", "0", "-f", path, "-p", "mixed"], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    rows = [l.split() for l in out.stdout.splitlines()[1:] if l and l[0].isdigit()]
    assert [int(r[0]) for r in rows] == [40, 60]
    r60 = fenzi.restart_read(prefix + ".rst", n, binary=True)
    assert r60["step"] == 60 and not np.array_equal(r60["xyz"], r40["xyz"])


@pytest.mark.parametrize("prec", [fenzi.FP64, fenzi.MIXED])
def test_step_virial_with_constraint_part_vs_oracle(ctx, prec):
    """The virial reported after fz_step: force virial at the final positions
    plus the constraint virial of the last RATTLE projection, against the
    oracle's or_step_ex (rigid waters and SHAKE'd chain-free slab)."""
    topo, xyz, params = systems.slab(seed=14)
    orc = oracle.Oracle(topo, params)
    u = orc.to_fixed(xyz)
    v0 = orc.maxwell(u, 300.0, seed=5, counter=0)
    u_ref, _, e_ref, w_ref, rc = orc.step_ex(u, v0, 6)
    assert rc == 0
    s = _system(topo, params, prec, ctx, xyz, orc)
    s.reassign_velocities(300.0, 5, 0)
    s.step(6)
    e = s.energies()
    _, _, w_forces = orc.forces(u_ref)
    wc = w_ref - w_forces
    assert np.abs(wc).max() > 1e-2 * np.abs(w_ref).max()   # the constraint part matters here
    tol = 1e-7 if prec == fenzi.FP64 else 2e-5
    assert np.abs(e["virial"] - w_ref).max() < tol * np.abs(w_ref).max()
    # the step's energies (mixed form: pair terms in FP32, summed in FP64)
    scale = sum(abs(v) for v in e_ref.values())
    assert abs(e["potential"] - sum(e_ref.values())) < (1e-9 if prec == fenzi.FP64 else 1e-5) * scale
    for k, name in (("vdw", "vdw"), ("elec", "elec"), ("pme", "pme")):
        assert abs(e[name] - e_ref[k]) < (1e-9 if prec == fenzi.FP64 else 1e-5) * scale, k


def test_launch_shapes_and_pair_timing_leave_the_forces_unchanged(ctx):
    """fz_set_launch reports the shapes in effect and keeps them across calls;
    the mixed-form forces do not depend on them (fixed-point sums), the
    reciprocal energy's per-workgroup FP64 partial sums do, at rounding level;
    fz_pair_kernel_time returns a positive time and leaves the system's forces
    and energies as they were (it evaluates them again after its launches);
    the energies count one pair-kernel tile per list entry."""
    topo, xyz, params = systems.slab(seed=12)
    s = fenzi.System(ctx, topo, params, fenzi.MIXED, xyz=xyz)
    f0, e0 = s.forces()
    eff = s.set_launch(fft_lines=4, pair_grid=512)
    assert eff["fft_lines"] == 4 and eff["pair_grid"] == 512
    assert s.set_launch() == eff  # -1: keep
    f1, e1 = s.forces()
    assert np.array_equal(f0, f1)
    assert abs(e1["potential"] - e0["potential"]) <= 1e-10 * abs(e0["potential"])
    ms = s.pair_kernel_time(5)
    assert ms > 0.0
    f2, e2 = s.forces()
    assert np.array_equal(f1, f2) and e2["potential"] == e1["potential"]
    # one evaluated tile per list entry
    assert e2["pair_tiles"] == len(s.cluster_pairs()[0])
    with pytest.raises(fenzi.FenziError):
        s.set_launch(fft_lines=1000)
    s.close()


def _same_energies(a, b):
    for k, v in a.items():
        if isinstance(v, np.ndarray):
            assert np.array_equal(v, b[k]), k
        elif k != "pair_tiles":
            assert v == b[k], k


def test_large_system_alone_matches_a_batch_member(ctx):
    """A 66,000-atom system (brick spread, grid FFTs) through several list
    builds: alone, one pair launch after the list step; in a batch, the pair
    kernel over its kept list beside the build and over its rebuilt list after
    it.  The trajectories and energies are bitwise the same."""
    topo, xyz, params = systems.water_box(n_waters=22000, L=88.0, seed=31)
    params = dict(params, K=(90, 90, 90), pairlist=params["cutoff"] + 0.8)
    a = fenzi.System(ctx, topo, params, fenzi.MIXED, xyz=xyz)
    assert a.info()["n_atoms"] >= 65536
    a.reassign_velocities(300.0, 4, 0)
    a.step(30)
    ua, ea = a.get_fixed(), a.energies()
    a.close()
    assert ea["list_builds"] >= 3
    b0 = fenzi.System(ctx, topo, params, fenzi.MIXED, xyz=xyz)
    t2, x2, p2 = systems.water_box(seed=2)
    b1 = fenzi.System(ctx, t2, p2, fenzi.MIXED, xyz=x2)
    for s in (b0, b1):
        s.reassign_velocities(300.0, 4, 0)
    b = fenzi.Batch(ctx, [b0, b1])
    b.step(30)
    assert np.array_equal(b0.get_fixed(), ua)
    _same_energies(b0.energies(), ea)
    b.close()
