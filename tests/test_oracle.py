"""Pins for the FP64 oracle (CPU only): known-answer vectors, closed forms,
library routines and finite differences, each chosen so that a plausible
mistake in the oracle (a dropped term, a wrong sign or index, a transposed
operand) fails at least one of them."""
import math
import os

import numpy as np
import pytest

import fz_synth
import oracle
import systems

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


# ---- Philox4x32-10: Random123 known-answer vectors ---------------------------
@pytest.mark.parametrize("ctr,key,out", [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
])
def test_philox_known_answers(ctr, key, out):
    import ctypes as C
    L = oracle.lib()
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    L.or_philox(c, k, o)
    assert list(o) == out


# ---- FFT: committed numpy.fft vectors, then live numpy for more shapes -------
def test_fft_golden_vectors():
    g = np.load(os.path.join(GOLDEN, "fft_vectors.npz"))
    for name in ("a", "b"):
        x = g["in_" + name]
        assert np.abs(oracle.fft3d(x, -1) - g["fwd_" + name]).max() < 1e-12 * np.abs(g["fwd_" + name]).max()
        assert np.abs(oracle.fft3d(x, +1) - g["inv_" + name]).max() < 1e-12 * np.abs(g["inv_" + name]).max()


@pytest.mark.parametrize("shape", [(5, 6, 7), (16, 9, 10), (3, 25, 8), (2, 3, 49)])
def test_fft_matches_numpy(shape):
    rng = np.random.default_rng(1)
    x = rng.standard_normal(shape) + 1j * rng.standard_normal(shape)
    assert np.allclose(oracle.fft3d(x, -1), np.fft.fftn(x), rtol=0, atol=1e-11 * np.abs(x).sum())


# ---- Ewald: NaCl Madelung constant --------------------------------------------
def _ions(charges, xyz, box, cutoff, beta, K=(16, 16, 16)):
    n = len(charges)
    topo = dict(n_atoms=n, n_types=1, charge=np.asarray(charges, float), mass=np.full(n, 10.0),
                type=np.zeros(n, np.int32), lj_eps=np.zeros(1), lj_sig2=np.ones(1), lj14_eps=np.zeros(1),
                lj14_sig2=np.ones(1), n_bonds=0, bond_atoms=np.zeros(0, np.int32), bond_k=np.zeros(0),
                bond_r0=np.zeros(0), n_angles=0, angle_atoms=np.zeros(0, np.int32), angle_k=np.zeros(0),
                angle_theta0=np.zeros(0), ub_k=np.zeros(0), ub_s0=np.zeros(0), n_dihedrals=0,
                dihedral_atoms=np.zeros(0, np.int32), dihedral_k=np.zeros(0), dihedral_n=np.zeros(0, np.int32),
                dihedral_delta=np.zeros(0), n_impropers=0, improper_atoms=np.zeros(0, np.int32),
                improper_k=np.zeros(0), improper_psi0=np.zeros(0), n_constraints=0,
                constraint_atoms=np.zeros(0, np.int32), constraint_length=np.zeros(0))
    params = dict(box=box, cutoff=cutoff, cuton=cutoff - 1.0, pairlist=cutoff + 1.0, beta=beta, K=K)
    o = oracle.Oracle(topo, params)
    return o, o.to_fixed(np.asarray(xyz, float))


def test_nacl_madelung_constant():
    r0, m = 2.8, 3                       # nearest-neighbour distance, 3x3x3 conventional cells
    pts, q = [], []
    for i in range(2 * m):
        for j in range(2 * m):
            for k in range(2 * m):
                pts.append((i * r0, j * r0, k * r0))
                q.append(1.0 if (i + j + k) % 2 == 0 else -1.0)
    L = 2 * m * r0
    o, u = _ions(q, pts, (L, L, L), cutoff=8.0, beta=0.6)
    _, e, _ = o.forces(u, oracle.REAL | oracle.EWALD | oracle.SELF)
    total = e["elec"] + e["pme"] + e["ewself"]
    madelung = -total / (len(q) / 2) * r0 / 332.0716
    assert abs(madelung - 1.747564594633182) < 1e-8


def test_ewald_total_is_beta_independent():
    rng = np.random.default_rng(4)
    x = rng.uniform(0, 18.0, (40, 3))
    q = np.tile([1.0, -1.0], 20)
    e = []
    for beta in (0.45, 0.55):
        o, u = _ions(q, x, (18.0, 18.0, 18.0), cutoff=8.9, beta=beta)
        _, ee, _ = o.forces(u, oracle.REAL | oracle.EWALD | oracle.SELF)
        e.append(ee["elec"] + ee["pme"] + ee["ewself"])
    assert abs(e[0] - e[1]) < 1e-7 * abs(e[0])


def test_pme_converges_to_direct_ewald():
    topo, xyz, p = systems.water_box(100, 15.0, seed=3)
    errs = []
    fe = None
    for K in (12, 16, 24, 32):
        params = dict(p, K=(K, K, K), cutoff=7.0, cuton=5.0, pairlist=8.0)
        o = oracle.Oracle(topo, params)
        u = o.to_fixed(xyz)
        if fe is None:
            fe, ee, _ = o.forces(u, oracle.EWALD)
        fp, ep, _ = o.forces(u, oracle.PME)
        errs.append(systems.rel_rms(fp, fe))
    # order-4 splines: force error falls roughly as h^3 (12 -> 32: ~95x here)
    assert errs[-1] < 3e-4 and errs[0] / errs[-1] > 30
    assert all(a > b for a, b in zip(errs, errs[1:]))


# ---- forces are the negative gradient of the energy ---------------------------
def _fd_system():
    topo, xyz = fz_synth.make_system(3, 60, (22.0, 22.0, 22.0), seed=2)
    # add Urey-Bradley terms and impropers so that every term type is exercised
    na = topo["n_angles"]
    topo["ub_k"] = np.where(np.arange(na) % 3 == 0, 15.0, 0.0)
    topo["ub_s0"] = np.full(na, 2.4)
    imp = np.array([[0, 1, 2, 3], [40, 41, 42, 43], [80, 81, 82, 83]], np.int32)
    topo["improper_atoms"] = imp.ravel()
    topo["improper_k"] = np.array([20.0, 35.0, 10.0])
    topo["improper_psi0"] = np.zeros(3)
    topo["n_impropers"] = 3
    topo["dihedral_delta"] = np.where(np.arange(topo["n_dihedrals"]) % 2 == 0, 0.0, math.radians(35.0))
    params = dict(box=(22.0, 22.0, 22.0), cutoff=9.0, cuton=7.0, pairlist=10.0, beta=0.35, K=(24, 24, 24))
    return topo, xyz, params


@pytest.mark.parametrize("flags", [oracle.REAL, oracle.EXCL, oracle.BONDED, oracle.PME, oracle.EWALD])
def test_force_is_minus_energy_gradient(flags):
    topo, xyz, params = _fd_system()
    o = oracle.Oracle(topo, params)
    u = o.to_fixed(xyz)
    f, _, _ = o.forces(u, flags)
    L = params["box"][0]
    rng = np.random.default_rng(flags)
    atoms = rng.choice(topo["n_atoms"], 12, replace=False).tolist() + [0, 1, 2, 41]
    for i in atoms:
        for d in range(3):
            du = np.uint64(int(1e-5 / L * 2 ** 64))
            up, um = u.copy(), u.copy()
            up[i, d] += du
            um[i, d] -= du
            h = int(du) * L / 2 ** 64
            ep = sum(o.forces(up, flags)[1].values())
            em = sum(o.forces(um, flags)[1].values())
            fd = -(ep - em) / (2 * h)
            assert abs(fd - f[i, d]) < 2e-5 * max(1.0, abs(f[i, d])), (flags, i, d, fd, f[i, d])


def test_virial_matches_box_scaling_derivative():
    """W_aa = -L_a dE/dL_a at fixed fractional coordinates, for every term together."""
    topo, xyz, params = _fd_system()
    o = oracle.Oracle(topo, params)
    u = o.to_fixed(xyz)
    flags = oracle.REAL | oracle.EXCL | oracle.BONDED | oracle.EWALD
    _, _, vir = o.forces(u, flags)
    for a in range(3):
        e = []
        for s in (1 + 1e-6, 1 - 1e-6):
            box = list(params["box"])
            box[a] *= s
            ob = oracle.Oracle(topo, dict(params, box=tuple(box)))
            e.append(sum(ob.forces(u, flags)[1].values()))
        dEdL = (e[0] - e[1]) / (2e-6 * params["box"][a])
        assert abs(-params["box"][a] * dEdL - vir[a, a]) < 1e-4 * np.abs(vir).max()


def test_pair_forces_obey_newton_third_law():
    topo, xyz, params = _fd_system()
    o = oracle.Oracle(topo, params)
    f, _, _ = o.forces(o.to_fixed(xyz), oracle.REAL | oracle.EXCL | oracle.BONDED)
    assert np.abs(f.sum(axis=0)).max() < 1e-9 * np.abs(f).max()


def test_dihedral_sign_convention():
    """phi = +60 deg in the IUPAC convention: E = K(1 + cos(n phi - delta)) for
    delta = 90 deg, n = 1 is K(1 + sin 60)."""
    topo, xyz, params = _fd_system()
    th = math.radians(60.0)
    pos = np.array([[1.0, 0.0, 0.0], [0.0, 0.0, 0.0], [0.0, 0.0, 1.5], [math.cos(th), math.sin(th), 1.5]]) + 5.0
    t = dict(topo)
    t["n_atoms"] = 4
    for k in ("charge", "mass", "type"):
        t[k] = topo[k][:4]
    for k in ("bonds", "angles", "impropers", "constraints"):
        t["n_" + k] = 0
    t["n_dihedrals"] = 1
    t["dihedral_atoms"] = np.array([0, 1, 2, 3], np.int32)
    t["dihedral_k"], t["dihedral_n"], t["dihedral_delta"] = np.array([2.0]), np.array([1], np.int32), np.array([math.pi / 2])
    o = oracle.Oracle(t, params)
    _, e, _ = o.forces(o.to_fixed(pos), oracle.BONDED)
    assert abs(e["dihed"] - 2.0 * (1.0 + math.sin(th))) < 1e-12


def test_angle_and_bond_closed_forms():
    topo, xyz, params = _fd_system()
    t = dict(topo)
    t["n_atoms"] = 3
    for k in ("charge", "mass", "type"):
        t[k] = topo[k][:3]
    t["n_dihedrals"] = t["n_impropers"] = t["n_constraints"] = 0
    t["n_bonds"], t["bond_atoms"], t["bond_k"], t["bond_r0"] = 1, np.array([0, 1], np.int32), np.array([300.0]), np.array([1.5])
    t["n_angles"], t["angle_atoms"] = 1, np.array([0, 1, 2], np.int32)
    t["angle_k"], t["angle_theta0"] = np.array([50.0]), np.array([math.radians(100.0)])
    t["ub_k"], t["ub_s0"] = np.zeros(1), np.zeros(1)
    pos = np.array([[1.7, 0, 0], [0, 0, 0], [0, 2.0, 0]]) + 4.0    # right angle, bond 1.7
    o = oracle.Oracle(t, params)
    _, e, _ = o.forces(o.to_fixed(pos), oracle.BONDED)
    assert abs(e["bond"] - 300.0 * 0.2 ** 2) < 1e-10
    assert abs(e["angle"] - 50.0 * math.radians(10.0) ** 2) < 1e-10


# ---- exclusions, pair sets, cluster pairs ----------------------------------------
def test_exclusion_sets_follow_bonds_angles_dihedrals():
    topo, xyz = fz_synth.make_system(2, 5, (30.0, 30.0, 30.0), seed=1)
    params = dict(box=(30.0,) * 3, cutoff=9.0, cuton=7.0, pairlist=10.0, beta=0.3, K=(24, 24, 24))
    o = oracle.Oracle(topo, params)
    ew = {tuple(p) for p in o.pairs(1).tolist()}
    p14 = {tuple(p) for p in o.pairs(2).tolist()}
    # chain of 38: 37 bonded + 36 1-3 pairs, 35 1-4 pairs; water: 3 pairs (O-H, O-H, H-H)
    assert len(ew) == 2 * (37 + 36) + 5 * 3
    assert len(p14) == 2 * 35
    assert (0, 3) in p14 and (0, 2) in ew and (0, 1) in ew and (0, 4) not in p14 | ew


def test_cluster_pairs_cover_exactly_the_pairs_within_list_distance():
    topo, xyz, params = systems.water_box(200, 19.0, seed=8)
    o = oracle.Oracle(topo, params)
    u = o.to_fixed(xyz)
    col, slots, pairs, masks = o.cluster_pairs(u, 4, 4)
    covered = set()
    for (I, J), m in zip(pairs.tolist(), masks.tolist()):
        assert I <= J and m != 0
        for l in range(64):
            if m >> l & 1:
                a, b = slots[8 * I + (l & 7)], slots[8 * J + (l >> 3)]
                key = (min(a, b), max(a, b))
                assert key not in covered
                covered.add(key)
    # brute force in FP64 (water only: a pair is excluded iff both atoms are in one
    # molecule): every pair within the list distance is covered, and nothing beyond
    # the FP32 test's 1e-6 relative margin is
    n = topo["n_atoms"]
    ia, ib = np.triu_indices(n, 1)
    sc = np.array(params["box"]) / 2.0 ** 64
    dd = ((u[ib] - u[ia]).view(np.int64)).astype(np.float64) * sc
    r2 = (dd ** 2).sum(1)
    keep = (ia // 3) != (ib // 3)
    rl2 = params["pairlist"] ** 2
    inner = set(zip(ia[keep & (r2 < rl2)].tolist(), ib[keep & (r2 < rl2)].tolist()))
    outer = set(zip(ia[keep & (r2 < rl2 * (1 + 4e-6))].tolist(), ib[keep & (r2 < rl2 * (1 + 4e-6))].tolist()))
    assert inner <= covered <= outer
    assert len(inner) == o.count_pairs_within(u, params["pairlist"])
    # every atom sits in exactly one slot, in the column its x, y fractions name
    real = slots[slots >= 0]
    assert sorted(real.tolist()) == list(range(topo["n_atoms"]))


# ---- integrator, constraints, velocities ---------------------------------------------
def test_shake_keeps_rigid_water_geometry():
    topo, xyz, params = systems.water_box(150, 17.0, seed=4)
    params = dict(params, cutoff=8.0, cuton=6.0, pairlist=9.0, K=(20, 20, 20), dt_ps=0.001)
    o = oracle.Oracle(topo, params)
    u = o.to_fixed(xyz)
    v = o.maxwell(u, 300.0, 1, 0)
    u, v, rc = o.step(u, v, 5)
    assert rc == 0
    res, vres = o.constraint_residual(u, v)
    assert res < params["shake_tol"] and vres < 1e-9
    x = o.from_fixed(u)
    box = np.array(params["box"])
    for w in range(0, 30, 3):
        d1 = x[w + 1] - x[w]; d1 -= box * np.round(d1 / box)
        d2 = x[w + 2] - x[w + 1]; d2 -= box * np.round(d2 / box)
        assert abs(np.linalg.norm(d1) - fz_synth.R_OH) < 1e-8
        assert abs(np.linalg.norm(d2) - fz_synth.R_HH) < 1e-8


def _two_atom_bond(k, r0, m=12.0):
    return dict(n_atoms=2, n_types=1, charge=np.zeros(2), mass=np.full(2, m), type=np.zeros(2, np.int32),
                lj_eps=np.zeros(1), lj_sig2=np.ones(1), lj14_eps=np.zeros(1), lj14_sig2=np.ones(1),
                n_bonds=1, bond_atoms=np.array([0, 1], np.int32), bond_k=np.array([k]), bond_r0=np.array([r0]),
                n_angles=0, angle_atoms=np.zeros(0, np.int32), angle_k=np.zeros(0), angle_theta0=np.zeros(0),
                ub_k=np.zeros(0), ub_s0=np.zeros(0), n_dihedrals=0, dihedral_atoms=np.zeros(0, np.int32),
                dihedral_k=np.zeros(0), dihedral_n=np.zeros(0, np.int32), dihedral_delta=np.zeros(0),
                n_impropers=0, improper_atoms=np.zeros(0, np.int32), improper_k=np.zeros(0),
                improper_psi0=np.zeros(0), n_constraints=0, constraint_atoms=np.zeros(0, np.int32),
                constraint_length=np.zeros(0))


def _bond_length_after_schedule(k, x0, nsteps, method, cgfac, sdfac, cgrate, sdrate):
    """The minimiser schedule on one harmonic bond, as a scalar recurrence on the
    stretch x = d - r0 (E = k x^2; the force on the far atom is f = -2 k x along
    the bond, the near atom gets -f, so each move changes d by twice the far
    atom's displacement)."""
    x, hprev, fprev = x0, 0.0, 0.0
    e_prev, have, done = 0.0, False, 0
    for _ in range(nsteps):
        if method == 0:
            break
        f = -2.0 * k * x
        if method == 1:
            p2 = fprev * fprev
            if p2 < 1e-8:
                p2 = f * f
            g = f * f / p2 if p2 > 0 else 0.0
            h = (f + g * hprev) * cgfac
            hprev, fprev = h, f
        else:
            h = f * sdfac
        xn = x + 2.0 * h
        e = k * xn * xn
        if have and e > e_prev:
            method = 2 if method == 1 else 0
            continue
        x, e_prev, have = xn, e, True
        idx = done
        done += 1
        if idx > 0 and idx % 100 == 0:
            if method == 1:
                cgfac *= cgrate * 100.0
            else:
                sdfac *= sdrate * 100.0
    return x, method, done


@pytest.mark.parametrize("method,cgfac,sdfac,nsteps", [
    (2, 1e-10, 1e-3, 30),      # steepest descent: x_k = x_0 (1 - 4 k sdfac)^k
    (2, 1e-10, 2e-5, 250),     # with two factor ramps (x10 after moves 100 and 200)
    (1, 6e-3, 1e-3, 40),       # conjugate gradient overshoots, switches to steepest descent
    (1, 1e-4, 1e-3, 30),       # conjugate gradient with gamma = 1 on a single bond
])
def test_reference_minimiser_schedule_on_one_bond(method, cgfac, sdfac, nsteps):
    k, r0, x0 = 100.0, 1.5, 0.2
    o = oracle.Oracle(_two_atom_bond(k, r0), dict(box=(30.0,) * 3, cutoff=9.0, cuton=8.0, pairlist=10.0,
                                                  beta=0.3, K=(16, 16, 16), dt_ps=0.001))
    u = o.to_fixed(np.array([[0.0, 0.0, 0.0], [r0 + x0, 0.0, 0.0]]))
    u2, m2, done = o.minimize_ref(u, nsteps, method, cgfac, sdfac, 0.1, 0.01)
    x2 = o.from_fixed(u2)
    x_ref, m_ref, done_ref = _bond_length_after_schedule(k, x0, nsteps, method, cgfac, sdfac, 0.1, 0.01)
    assert m2 == m_ref and done == done_ref
    assert abs((x2[1, 0] - x2[0, 0] - r0) - x_ref) < 1e-9
    if method == 2 and nsteps == 30:
        assert abs(x_ref - x0 * (1 - 4 * k * sdfac) ** nsteps) < 1e-12
    if method == 1 and cgfac == 4e-4:
        assert m_ref == 2  # the overshoot did switch methods


def test_harmonic_bond_oscillates_at_the_closed_form_frequency():
    k, r0, m = 100.0, 1.0, 12.0
    t = dict(n_atoms=2, n_types=1, charge=np.zeros(2), mass=np.full(2, m), type=np.zeros(2, np.int32),
             lj_eps=np.zeros(1), lj_sig2=np.ones(1), lj14_eps=np.zeros(1), lj14_sig2=np.ones(1),
             n_bonds=1, bond_atoms=np.array([0, 1], np.int32), bond_k=np.array([k]), bond_r0=np.array([r0]),
             n_angles=0, angle_atoms=np.zeros(0, np.int32), angle_k=np.zeros(0), angle_theta0=np.zeros(0),
             ub_k=np.zeros(0), ub_s0=np.zeros(0), n_dihedrals=0, dihedral_atoms=np.zeros(0, np.int32),
             dihedral_k=np.zeros(0), dihedral_n=np.zeros(0, np.int32), dihedral_delta=np.zeros(0),
             n_impropers=0, improper_atoms=np.zeros(0, np.int32), improper_k=np.zeros(0), improper_psi0=np.zeros(0),
             n_constraints=0, constraint_atoms=np.zeros(0, np.int32), constraint_length=np.zeros(0))
    params = dict(box=(30.0,) * 3, cutoff=9.0, cuton=8.0, pairlist=10.0, beta=0.3, K=(16, 16, 16), dt_ps=0.0005)
    o = oracle.Oracle(t, params)
    x = np.array([[0.0, 0.0, 0.0], [r0 + 0.1, 0.0, 0.0]])
    u = o.to_fixed(x)
    v = np.zeros((2, 3))
    mu = m / 2.0
    omega = math.sqrt(2.0 * k / mu)       # E = k (r - r0)^2
    nsteps = 400
    u2, v2, _ = o.step(u, v, nsteps)
    x2 = o.from_fixed(u2)
    h = params["dt_ps"] * 20.4548
    # velocity Verlet on a harmonic oscillator is exactly a rotation with
    # frequency (2/h) asin(omega h / 2)
    omega_h = 2.0 / h * math.asin(omega * h / 2.0)
    assert abs((x2[1, 0] - x2[0, 0] - r0) - 0.1 * math.cos(omega_h * nsteps * h)) < 1e-9


def test_maxwell_velocities_statistics():
    topo, xyz, params = systems.water_box(3000, 31.0, seed=1)
    o = oracle.Oracle(topo, params)
    u = o.to_fixed(xyz)
    v = o.maxwell(u, 300.0, 12, 0)
    p = (topo["mass"][:, None] * v).sum(axis=0)
    assert np.abs(p).max() < 1e-9
    assert abs(o.temperature(v) - 300.0) < 10.0
    v2 = o.maxwell(u, 300.0, 12, 1)
    assert not np.allclose(v, v2)


# ---- FP32 erfc fit --------------------------------------------------------------------
def test_erfc_fit_relative_error():
    from scipy.special import erfc
    hdr = open(os.path.join(os.path.dirname(GOLDEN), "..", "fen-zi_amd", "csrc", "fz_erfc_fit.h")).read()
    vals = {}
    for line in hdr.splitlines():
        if line.startswith("#define FZ_ERFC_"):
            _, k, v = line.split()
            vals[k] = float(v.rstrip("f"))
    c = [vals["FZ_ERFC_C%d" % i] for i in range(int(vals["FZ_ERFC_DEG"]) + 1)]
    x = np.linspace(0, 4.0, 20001)
    t = 1 / (1 + vals["FZ_ERFC_P"] * x)
    y = np.exp(-x * x) * t * np.polyval(c[::-1], t)
    assert np.abs(y / erfc(x) - 1).max() < 2e-7


def _free_atoms(masses):
    """Non-interacting atoms of the given masses (no charges, no LJ, no terms)."""
    n = len(masses)
    t = _two_atom_bond(0.0, 1.0)
    t.update(n_atoms=n, charge=np.zeros(n), mass=np.asarray(masses, float), type=np.zeros(n, np.int32),
             n_bonds=0, bond_atoms=np.zeros(0, np.int32), bond_k=np.zeros(0), bond_r0=np.zeros(0))
    return t


def test_com_restraint_is_the_gradient_of_a_harmonic_well():
    """With a one-atom group 1 the restraint force is -grad of fc/2 |mi(x1 - c0 - d)|^2
    (minimum image), c0 the mass-weighted centre of group 0; group 0 and
    ungrouped atoms feel nothing."""
    L = 20.0
    params = dict(box=(L, L, L), cutoff=5.0, cuton=4.0, pairlist=6.0, beta=0.3, K=(16, 16, 16))
    masses = [12.0, 16.0, 1.0, 14.0]
    o = oracle.Oracle(_free_atoms(masses), params)
    group = [0, 0, 1, -1]
    dist, fc = np.array([3.0, -2.0, 6.5]), 0.7
    o.set_com_restraint(group, dist, fc)
    # atom 2 sits across the boundary from group 0's centre, so the minimum image matters
    x = np.array([[1.0, 2.0, -3.0], [2.5, 1.0, -2.0], [-8.5, 4.0, 7.0], [0.0, 0.0, 0.0]])

    def well(xx):
        c0 = (masses[0] * xx[0] + masses[1] * xx[1]) / (masses[0] + masses[1])
        r = xx[2] - c0 - dist
        r -= L * np.round(r / L)
        return 0.5 * fc * (r ** 2).sum()

    f, _, _ = o.forces(o.to_fixed(x), flags=oracle.RESTRAINT)
    h = 1e-5
    for k in range(3):
        xp, xm = x.copy(), x.copy()
        xp[2, k] += h
        xm[2, k] -= h
        assert abs(f[2, k] + (well(xp) - well(xm)) / (2 * h)) < 1e-7
    assert np.abs(f[[0, 1, 3]]).max() == 0.0
    assert np.abs(f[2]).max() > 1.0


def test_com_restraint_uses_mass_weighted_centres():
    """Moving two group-0 atoms so that their mass-weighted centre stays put leaves
    the force unchanged; translating group 0 by t changes it by fc t; every atom
    of group 1 receives the same force."""
    L = 30.0
    params = dict(box=(L, L, L), cutoff=5.0, cuton=4.0, pairlist=6.0, beta=0.3, K=(16, 16, 16))
    masses = [12.0, 16.0, 1.0, 14.0, 2.0]
    o = oracle.Oracle(_free_atoms(masses), params)
    o.set_com_restraint([0, 0, 1, 1, -1], (1.0, 2.0, 3.0), 0.25)
    x = np.array([[1.0, 2.0, -3.0], [2.5, 1.0, -2.0], [4.0, 4.0, 4.0], [5.0, 3.0, 6.0], [0.0, 0.0, 0.0]])
    f0, _, _ = o.forces(o.to_fixed(x), flags=oracle.RESTRAINT)
    assert np.array_equal(f0[2], f0[3])
    y = x.copy()
    step = np.array([0.4, -0.3, 0.2])
    y[0] += step * masses[1]
    y[1] -= step * masses[0]
    f1, _, _ = o.forces(o.to_fixed(y), flags=oracle.RESTRAINT)
    assert np.abs(f1 - f0).max() < 1e-9
    t = np.array([0.5, -0.25, 1.0])
    z = x.copy()
    z[:2] += t
    f2, _, _ = o.forces(o.to_fixed(z), flags=oracle.RESTRAINT)
    assert np.abs(f2[2] - f0[2] - 0.25 * t).max() < 1e-9


def test_fixed_atoms_stay_put_and_constraints_between_them_are_dropped():
    topo, xyz, params = systems.water_box(150, 17.0, seed=4)
    params = dict(params, cutoff=8.0, cuton=6.0, pairlist=9.0, K=(20, 20, 20), dt_ps=0.001)
    o = oracle.Oracle(topo, params)
    fixed = np.zeros(topo["n_atoms"], np.uint8)
    fixed[:30] = 1          # waters 0-9 entirely
    o.set_fixed(fixed)
    u0 = o.to_fixed(xyz)
    v = o.maxwell(u0, 300.0, 2, 0)
    assert np.abs(v[fixed == 1]).max() == 0.0
    m = np.asarray(topo["mass"])
    assert np.abs((m[:, None] * v).sum(0)).max() < 1e-9  # whole molecules fixed: momentum-free draw
    fixed[30] = 1           # and the oxygen of water 10 (its constraints still act on its hydrogens)
    o.set_fixed(fixed)
    v = o.maxwell(u0, 300.0, 2, 0)
    assert np.abs(v[fixed == 1]).max() == 0.0
    u, v, rc = o.step(u0, v, 5)
    assert rc == 0
    assert np.array_equal(u[fixed == 1], u0[fixed == 1])
    assert np.abs(v[fixed == 1]).max() == 0.0
    assert not np.array_equal(u[31], u0[31])  # water 10's hydrogens move
    res, vres = o.constraint_residual(u, v)
    assert res < params["shake_tol"] and vres < 1e-9
    # 3 N_mobile - 6 - N_constraints: the 10 fixed waters' 30 constraints are gone
    dof = 3 * (topo["n_atoms"] - 31) - 6 - (topo["n_constraints"] - 30)
    assert abs(o.temperature(v) - 2 * o.kinetic(v) / (oracle.RC * dof)) < 1e-12 * o.temperature(v)


def test_adaptive_minimiser_lowers_energy_and_keeps_constraints():
    """The restated adaptive steepest descent: the largest force moves max_step
    on the first move, the energy falls, SHAKE holds, and a split run (step
    size carried over) follows one run."""
    topo, xyz, params = systems.slab(n_chains=2, n_waters=120, L=20.0, seed=12, K=(20, 20, 20))
    o = oracle.Oracle(topo, params)
    u0 = o.to_fixed(xyz)
    f0, e0, _ = o.forces(u0)
    u1, s1 = o.minimize_sd(u0, 1, 0.05)
    d = (u1 - u0).view(np.int64).astype(np.float64) * (params["box"][0] / 2.0 ** 64)
    imax = int(np.argmax((f0 * f0).sum(1)))
    if s1 > 0.05 - 1e-12:     # accepted: the largest-force atom moved by the step (before SHAKE of its cluster)
        assert np.linalg.norm(d[imax]) > 0.0
    ua, sa = o.minimize_sd(u0, 30, 0.05)
    ub, sb = o.minimize_sd(u0, 12, 0.05)
    ub, sb = o.minimize_sd(ub, 18, 0.05, step=sb)
    # the oracle's threaded force sums differ from run to run in the last bit
    assert np.abs((ua - ub).view(np.int64)).max() * params["box"][0] / 2.0 ** 64 < 1e-9 and sa == sb
    _, e1, _ = o.forces(ua)
    assert sum(e1.values()) < sum(e0.values())
    res, _ = o.constraint_residual(ua)
    assert res < params["shake_tol"] * 10


def test_adaptive_minimiser_on_one_bond_follows_the_scalar_recurrence():
    """On one harmonic bond (E = k x^2, both atoms feel |F| = 2 k |x|) every
    accepted move shifts each atom by the step toward the other, so the stretch
    follows x -> x - 2 s sign(x) when |x - 2 s sign(x)| < |x| (then s -> min(1.2 s,
    s_max)), else x stays and s -> s / 2."""
    k, r0, x0, smax = 100.0, 1.5, 0.37, 0.05
    o = oracle.Oracle(_two_atom_bond(k, r0), dict(box=(30.0,) * 3, cutoff=9.0, cuton=8.0, pairlist=10.0,
                                                  beta=0.3, K=(16, 16, 16), dt_ps=0.001))
    u = o.to_fixed(np.array([[0.0, 0.0, 0.0], [r0 + x0, 0.0, 0.0]]))
    u2, s2 = o.minimize_sd(u, 25, smax)
    x2 = o.from_fixed(u2)
    x, s = x0, smax
    for _ in range(25):
        xn = x - 2.0 * s * np.sign(x)
        if abs(xn) < abs(x):
            x, s = xn, min(1.2 * s, smax)
        else:
            s *= 0.5
    assert abs((x2[1, 0] - x2[0, 0] - r0) - x) < 1e-12 and abs(s2 - s) < 1e-15
