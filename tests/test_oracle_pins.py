"""Pins for the parts of the FP64 oracle that finite differences of its own
energy cannot pin (a sign or factor error made in both an energy and its force
passes those): the Ewald exclusion correction and the 1-4 Coulomb term against
closed forms for isolated molecules, the total electrostatic energy's
independence of the Ewald splitting with exclusions present, every component
of the virial (reciprocal part against an independent triclinic Ewald sum
under shear, pair and bonded parts against homogeneous deformations), the
CHARMM switch and the Lennard-Jones form against their closed-form values, and
the PRM-to-table conversions against the reference's own parameter file.

Reference code pinned here: Ewald exclusion correction src/mdcuda.cu:520-553;
1-4 pairs src/bonded_forces.cu:1155-1268; self energy src/prmread.cu:1734;
reciprocal virial src/pme_kernels.cu:569-572; switch src/mem.cu:557-560 and
src/nonbonded_forces.cu:146-233; LJ conversions src/prmread.cu:2460-2514,
2540-2565, 2720-2736."""
import math
import os

import numpy as np
import pytest

import fenzi
import fz_synth
import oracle
import systems

CC = 332.0716


def _topo(charges, bonds=(), angles=(), diheds=(), types=None, lj=None, lj14=None):
    """Topology of the given atoms with zero-force-constant bonded terms (they
    only define exclusions: 1-2, 1-3 and 1-4 pairs)."""
    n = len(charges)
    nt = 1 if types is None else int(max(types)) + 1
    eps, sig2 = (np.zeros(nt * nt), np.ones(nt * nt)) if lj is None else lj
    eps14, sig214 = (eps, sig2) if lj14 is None else lj14
    b = np.asarray(bonds, np.int32).reshape(-1, 2)
    a = np.asarray(angles, np.int32).reshape(-1, 3)
    d = np.asarray(diheds, np.int32).reshape(-1, 4)
    t = dict(n_atoms=n, n_types=nt, charge=np.asarray(charges, float), mass=np.full(n, 12.0),
             type=np.zeros(n, np.int32) if types is None else np.asarray(types, np.int32),
             lj_eps=np.asarray(eps, float), lj_sig2=np.asarray(sig2, float),
             lj14_eps=np.asarray(eps14, float), lj14_sig2=np.asarray(sig214, float),
             n_bonds=len(b), bond_atoms=b.ravel(), bond_k=np.zeros(len(b)), bond_r0=np.ones(len(b)),
             n_angles=len(a), angle_atoms=a.ravel(), angle_k=np.zeros(len(a)), angle_theta0=np.zeros(len(a)),
             ub_k=np.zeros(len(a)), ub_s0=np.zeros(len(a)),
             n_dihedrals=len(d), dihedral_atoms=d.ravel(), dihedral_k=np.zeros(len(d)),
             dihedral_n=np.ones(len(d), np.int32), dihedral_delta=np.zeros(len(d)),
             n_impropers=0, improper_atoms=np.zeros(0, np.int32), improper_k=np.zeros(0),
             improper_psi0=np.zeros(0), n_constraints=0, constraint_atoms=np.zeros(0, np.int32),
             constraint_length=np.zeros(0))
    return t


def _molecule(charges, xyz, L, beta=0.4, bonds=(), angles=(), diheds=(), types=None, lj=None, lj14=None,
              cutoff=None, cuton=None, K=16):
    """Oracle and fixed-point positions of _topo(...) in a cubic box of edge L."""
    t = _topo(charges, bonds, angles, diheds, types, lj, lj14)
    cutoff = cutoff or min(12.0, 0.45 * L)
    cuton = cutoff - 2.0 if cuton is None else cuton
    params = dict(box=(L, L, L), cutoff=cutoff, cuton=cuton, pairlist=cutoff + 1.0, beta=beta, K=(K, K, K))
    o = oracle.Oracle(t, params)
    return o, o.to_fixed(np.asarray(xyz, float))


ELEC = oracle.REAL | oracle.EXCL | oracle.EWALD | oracle.SELF


def _elec_total(o, u):
    _, e, _ = o.forces(u, ELEC)
    return e["elec"] + e["ewexcl"] + e["pme"] + e["ewself"]


def _lattice_limit(energy_of_L, L1=40.0, L2=60.0):
    """a of E(L) = a / L^3 + b / L^5 from two box sizes (Richardson step)."""
    E1, E2 = energy_of_L(L1), energy_of_L(L2)
    return (E1 * L1 ** 5 - E2 * L2 ** 5) / (L1 ** 2 - L2 ** 2)


# ---- exclusion correction, self energy, 1-4 Coulomb: closed forms -----------------
def test_isolated_water_ewald_energy_is_the_tinfoil_dipole_term():
    """One rigid water (all three pairs excluded) per cubic cell.  With every
    intramolecular pair excluded the Ewald energy is the image interaction alone:
    for a cubic lattice of point dipoles under tin-foil boundary conditions that
    is -(2 pi / 3 V) |p|^2 (the spherical lattice sum of a cubic lattice
    vanishes), plus quadrupole terms of order L^-5.  A sign or factor error in
    the exclusion correction or the self energy leaves O(100) kcal/mol here."""
    w = fz_synth.water_local() + 1.3
    q = np.array([fz_synth.Q_O, fz_synth.Q_H, fz_synth.Q_H])
    p = (q[:, None] * w).sum(0)

    def energy(L):
        o, u = _molecule(q, w, L, bonds=[[0, 1], [0, 2]], angles=[[1, 0, 2]])
        return _elec_total(o, u)

    a = _lattice_limit(energy)
    assert abs(a / (-2.0 * math.pi / 3.0 * (p @ p) * CC) - 1.0) < 5e-6
    # and the splitting parameter drops out
    o, u = _molecule(q, w, 40.0, beta=0.3, bonds=[[0, 1], [0, 2]], angles=[[1, 0, 2]])
    o2, u2 = _molecule(q, w, 40.0, beta=0.5, bonds=[[0, 1], [0, 2]], angles=[[1, 0, 2]])
    assert abs(_elec_total(o, u) - _elec_total(o2, u2)) < 1e-10


def test_isolated_chain_has_full_coulomb_between_its_1_4_ends():
    """Four charged atoms bonded 0-1-2-3: the 1-2 and 1-3 pairs carry no
    Coulomb energy (exclusion correction), the 1-4 pair carries the full
    CC q0 q3 / r (its real-space erfc term plus the reciprocal erf part), and
    the lattice adds -(2 pi / 3 V) |p|^2."""
    x = np.array([[0.0, 0.0, 0.0], [1.53, 0.0, 0.0], [2.1, 1.42, 0.0], [3.4, 1.9, 1.1]]) + 2.0
    q = np.array([0.5, -0.3, 0.2, -0.4])
    p = (q[:, None] * x).sum(0)
    e14 = CC * q[0] * q[3] / np.linalg.norm(x[3] - x[0])
    top = dict(bonds=[[0, 1], [1, 2], [2, 3]], angles=[[0, 1, 2], [1, 2, 3]], diheds=[[0, 1, 2, 3]])

    def energy(L):
        o, u = _molecule(q, x, L, **top)
        return _elec_total(o, u) - e14

    a = _lattice_limit(energy)
    assert abs(a / (-2.0 * math.pi / 3.0 * (p @ p) * CC) - 1.0) < 1e-5
    # the 1-4 pair's real-space term is erfc-screened, not bare: per term
    o, u = _molecule(q, x, 40.0, **top)
    _, e, _ = o.forces(u, oracle.EXCL)
    r14 = np.linalg.norm(x[3] - x[0])
    from scipy.special import erf, erfc
    assert abs(e["elec"] - e14 * erfc(0.4 * r14)) < 1e-9 * abs(e14)
    d = [np.linalg.norm(x[i] - x[j]) for i, j in ((0, 1), (1, 2), (2, 3), (0, 2), (1, 3))]
    qq = [q[0] * q[1], q[1] * q[2], q[2] * q[3], q[0] * q[2], q[1] * q[3]]
    ew = -CC * sum(c * erf(0.4 * r) / r for c, r in zip(qq, d))
    assert abs(e["ewexcl"] - ew) < 1e-9 * abs(ew)


def test_electrostatics_with_exclusions_is_beta_independent():
    """Waters and a charged chain (1-2, 1-3 and 1-4 exclusions): the sum of the
    real-space, 1-4, exclusion-correction, reciprocal and self terms is the
    Coulomb energy of the non-excluded pairs plus the 1-4 pairs, whatever the
    splitting parameter."""
    topo, xyz = fz_synth.make_system(1, 90, (24.0, 24.0, 24.0), seed=6)
    rng = np.random.default_rng(6)
    q = np.asarray(topo["charge"], float).copy()
    chain = np.arange(fz_synth.CHAIN_LEN)
    q[chain] = rng.uniform(-0.4, 0.4, len(chain))
    q[chain] -= q[chain].mean()
    topo["charge"] = q
    e = []
    for beta in (0.42, 0.5):
        params = dict(box=(24.0,) * 3, cutoff=11.5, cuton=9.5, pairlist=12.0, beta=beta, K=(24, 24, 24))
        o = oracle.Oracle(topo, params)
        u = o.to_fixed(xyz)
        e.append(_elec_total(o, u))
        assert o.pairs(2).shape[0] == fz_synth.CHAIN_LEN - 3
    assert abs(e[0] - e[1]) < 2e-8 * abs(e[0])


# ---- virial: every component -------------------------------------------------------
def _triclinic_ewald(s, q, L, beta, eps, nmax):
    """Reciprocal Ewald energy of fractional coordinates s in the cell
    (I + eps) diag(L): E = CC / (2 pi V) sum_{n != 0} exp(-pi^2 m^2 / beta^2) / m^2
    |sum_j q_j exp(2 pi i n . s_j)|^2, m = (I + eps)^-T diag(1/L) n."""
    H = (np.eye(3) + eps) @ np.diag(L)
    V = abs(np.linalg.det(H))
    r = np.arange(-nmax, nmax + 1)
    n = np.stack(np.meshgrid(r, r, r, indexing="ij"), -1).reshape(-1, 3)
    n = n[np.any(n != 0, axis=1)]
    m = n @ np.linalg.inv(H)          # rows: n^T H^-1 = (H^-T n)^T
    m2 = (m * m).sum(1)
    S = np.exp(2j * np.pi * (n @ s.T)) @ q
    return CC / (2.0 * np.pi * V) * (np.exp(-np.pi ** 2 * m2 / beta ** 2) / m2 * np.abs(S) ** 2).sum()


def test_reciprocal_virial_matches_triclinic_ewald_under_strain():
    """W_ab = -dE/d eps_ab (symmetric strain, fractional coordinates fixed) for
    every component, against an Ewald sum written for a general (sheared) cell;
    the PME virial agrees with it to the mesh accuracy."""
    rng = np.random.default_rng(9)
    L = np.array([17.0, 19.0, 21.0])
    n = 36
    x = rng.uniform(0.0, 1.0, (n, 3)) * L
    q = np.tile([0.7, -0.7], n // 2)
    beta = 0.42
    t = _topo(q)
    params = dict(box=tuple(L), cutoff=8.0, cuton=7.0, pairlist=9.0, beta=beta, K=(48, 48, 48))
    o = oracle.Oracle(t, params)
    u = o.to_fixed(x)
    _, e_ew, w_ew = o.forces(u, oracle.EWALD)
    _, e_pme, w_pme = o.forces(u, oracle.PME)
    s = u.astype(np.float64) / 2.0 ** 64
    nmax = 12
    assert abs(_triclinic_ewald(s, q, L, beta, np.zeros((3, 3)), nmax) - e_ew["pme"]) < 1e-9 * abs(e_ew["pme"])
    h = 1e-5
    for a, b in ((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)):
        eps = np.zeros((3, 3))
        eps[a, b] = eps[b, a] = h
        dE = (_triclinic_ewald(s, q, L, beta, eps, nmax) - _triclinic_ewald(s, q, L, beta, -eps, nmax)) / (2 * h)
        w = w_ew[a, a] if a == b else w_ew[a, b] + w_ew[b, a]
        assert abs(-dE - w) < 1e-6 * np.abs(w_ew).max(), (a, b, -dE, w)
    assert np.allclose(w_ew, w_ew.T, atol=1e-12 * np.abs(w_ew).max())
    assert np.abs(w_pme - w_ew).max() < 2e-4 * np.abs(w_ew).max()
    assert np.abs(w_ew - np.diag(np.diag(w_ew))).max() > 1e-3 * np.abs(w_ew).max()  # off-diagonals not trivial


def test_pme_virial_matches_box_scaling_derivative():
    """W_aa = -L_a dE/dL_a of the PME energy at fixed fractional coordinates (the
    mesh charges do not change, only the influence function does), together
    with every other term."""
    topo, xyz = fz_synth.make_system(3, 60, (22.0, 22.0, 22.0), seed=2)
    params = dict(box=(22.0, 23.0, 24.0), cutoff=9.0, cuton=7.0, pairlist=10.0, beta=0.35, K=(24, 25, 27))
    o = oracle.Oracle(topo, params)
    u = o.to_fixed(xyz)
    flags = oracle.REAL | oracle.EXCL | oracle.BONDED | oracle.PME
    _, _, vir = o.forces(u, flags)
    for a in range(3):
        e = []
        for sgn in (1, -1):
            box = list(params["box"])
            box[a] *= 1 + sgn * 1e-6
            ob = oracle.Oracle(topo, dict(params, box=tuple(box)))
            e.append(sum(ob.forces(u, flags)[1].values()))
        dEdL = (e[0] - e[1]) / (2e-6 * params["box"][a])
        assert abs(-params["box"][a] * dEdL - vir[a, a]) < 2e-5 * np.abs(vir).max()


def test_pair_and_bonded_virial_off_diagonals_match_a_homogeneous_deformation():
    """An isolated cluster (a chain and waters, far from its periodic images):
    deforming every position by x -> (I + eps) x with a symmetric eps changes
    the energy by -sum_ab W_ab eps_ab, which pins the off-diagonal components
    of the pair, exclusion-correction, 1-4 and bonded virials."""
    topo, xyz = fz_synth.make_system(1, 12, (24.0, 24.0, 24.0), seed=3)
    topo["ub_k"] = np.where(np.arange(topo["n_angles"]) % 4 == 0, 12.0, 0.0)
    topo["ub_s0"] = np.full(topo["n_angles"], 2.5)
    L0 = 24.0
    # unwrap: chain beads in sequence, each water around its oxygen
    x = xyz.copy()
    for i in range(1, fz_synth.CHAIN_LEN):
        d = x[i] - x[i - 1]
        x[i] = x[i - 1] + d - L0 * np.round(d / L0)
    for o_ in range(fz_synth.CHAIN_LEN, topo["n_atoms"], 3):
        for h in (o_ + 1, o_ + 2):
            d = x[h] - x[o_]
            x[h] = x[o_] + d - L0 * np.round(d / L0)
        c = x[o_] - x[:fz_synth.CHAIN_LEN].mean(0)
        sh = L0 * np.round(c / L0)
        x[o_:o_ + 3] -= sh
    x -= x.mean(0)
    ext = np.ptp(x, axis=0).max()
    L = 3.0 * ext + 40.0
    params = dict(box=(L, L, L), cutoff=12.0, cuton=10.0, pairlist=13.0, beta=0.3, K=(16, 16, 16))
    o = oracle.Oracle(topo, params)
    flags = oracle.REAL | oracle.EXCL | oracle.BONDED
    _, _, vir = o.forces(o.to_fixed(x), flags)
    h = 1e-6
    for a, b in ((0, 1), (0, 2), (1, 2), (0, 0), (2, 2)):
        eps = np.zeros((3, 3))
        eps[a, b] = eps[b, a] = h
        ep = sum(o.forces(o.to_fixed(x @ (np.eye(3) + eps).T), flags)[1].values())
        em = sum(o.forces(o.to_fixed(x @ (np.eye(3) - eps).T), flags)[1].values())
        w = vir[a, a] if a == b else vir[a, b] + vir[b, a]
        assert abs(-(ep - em) / (2 * h) - w) < 1e-5 * np.abs(vir).max(), (a, b)
    assert np.abs(vir - vir.T).max() < 1e-9 * np.abs(vir).max()


# ---- switch and Lennard-Jones closed forms ---------------------------------------------
def _lj_pair(eps_ij, sig2_ij, r, cutoff=10.0, cuton=8.0):
    o, u = _molecule([0.0, 0.0], [[5.0, 5.0, 5.0], [5.0 + r, 5.0, 5.0]], 40.0,
                     lj=(np.array([eps_ij]), np.array([sig2_ij])), cutoff=cutoff, cuton=cuton)
    f, e, _ = o.forces(u, oracle.REAL)
    return e["vdw"], f[1, 0]


def test_lennard_jones_minimum_and_switch_end_values():
    """E = eps_ij ((s/r)^12 - (s/r)^6) (eps_ij = 4 eps): minimum -eps at
    r = 2^(1/6) s with zero force; the switch S(r) is 1 up to r_on, 1/2 at
    r^2 = (r_on^2 + r_c^2) / 2, 0 from r_c, with a continuous energy and force."""
    eps, rmin = 0.1521, 2.0 * 1.7682
    s = rmin / 2.0 ** (1.0 / 6.0)
    e, f = _lj_pair(4 * eps, s * s, rmin)
    assert abs(e + eps) < 1e-13 and abs(f) < 1e-12
    # switch midpoint (r_on = 8, r_c = 10): S = 1/2 of the bare energy
    rm = math.sqrt(0.5 * (64.0 + 100.0))
    bare = 4 * eps * ((s / rm) ** 12 - (s / rm) ** 6)
    assert abs(_lj_pair(4 * eps, s * s, rm)[0] - 0.5 * bare) < 1e-15
    # S(r_on) = 1 and S'(r_on) = 0: just inside the switching range the energy
    # and force are the bare ones to O(delta^2) and O(delta)
    def bare_ef(r):
        return 4 * eps * ((s / r) ** 12 - (s / r) ** 6), 4 * eps * (12 * (s / r) ** 12 - 6 * (s / r) ** 6) / r

    d = 1e-6
    e, f = _lj_pair(4 * eps, s * s, 8.0 + d)
    eb, fb = bare_ef(8.0 + d)
    assert abs(e - eb) < 1e-13 and abs(f - fb) < 1e-8 and abs(fb) > 1e-3
    assert _lj_pair(4 * eps, s * s, 8.0 - d) == pytest.approx(bare_ef(8.0 - d), rel=1e-13)
    # S(r_c) = 0 and S'(r_c) = 0: energy and force vanish continuously
    e, f = _lj_pair(4 * eps, s * s, 10.0 - d)
    assert abs(e) < 1e-14 and abs(f) < 1e-8
    assert _lj_pair(4 * eps, s * s, 10.0 + d) == (0.0, 0.0)


def test_1_4_pairs_use_the_1_4_table():
    """A 1-4 pair takes the 1-4 LJ parameters (NBXMOD 5), not the general ones."""
    x = np.array([[0.0, 0.0, 0.0], [1.53, 0.0, 0.0], [2.1, 1.42, 0.0], [3.4, 1.9, 1.1]]) + 2.0
    r = np.linalg.norm(x[3] - x[0])
    lj = (np.array([4 * 0.056]), np.array([(2 * 2.01 / 2 ** (1 / 6)) ** 2]))
    lj14 = (np.array([4 * 0.01]), np.array([(2 * 1.9 / 2 ** (1 / 6)) ** 2]))
    o, u = _molecule(np.zeros(4), x, 40.0, bonds=[[0, 1], [1, 2], [2, 3]], angles=[[0, 1, 2], [1, 2, 3]],
                     diheds=[[0, 1, 2, 3]], lj=lj, lj14=lj14)
    _, e, _ = o.forces(u, oracle.EXCL | oracle.REAL)
    s2 = lj14[1][0] / (r * r)
    assert abs(e["vdw"] - lj14[0][0] * (s2 ** 6 - s2 ** 3)) < 1e-14


# ---- PRM -> pair tables on the reference's own parameter file ----------------------------
def _psf_type_names(psf_path, rtf_path):
    """Type name of every atom: the PSF's numeric type through the RTF's MASS records."""
    code = {}
    with open(rtf_path) as f:
        for l in f:
            w = l.split()
            if len(w) >= 3 and w[0] == "MASS":
                code[w[1]] = w[2]
    with open(psf_path) as f:
        lines = f.read().splitlines()
    k = next(i for i, l in enumerate(lines) if "!NATOM" in l)
    n = int(lines[k].split()[0])
    return [code[l.split()[5]] for l in lines[k + 1:k + 1 + n]]


def test_dmpc_small_pair_tables_follow_the_prm_conversions(tmp_path):
    """OT (eps -0.1554253, Rmin/2 1.7765722), CTL2 (eps -0.056, Rmin/2 2.010,
    1-4: -0.01, 1.9) and the HT-OT NBFIX (Emin 0, Rmin 1.9927) of
    spcfw_lipid_forpme.prm give eps_ij = 4 sqrt(eps_i eps_j),
    sigma_ij = (Rmin_i/2 + Rmin_j/2) / 2^(1/6) and, for NBFIX,
    sigma_ij = Rmin_ij / 2^(1/6); the OT-OT well then sits at Rmin with depth
    -eps."""
    path = systems.dmpc_small_deck(str(tmp_path))
    deck = fenzi.Deck(path)
    t = deck.topology()
    names = _psf_type_names(os.path.join(str(tmp_path), "dmpc_small.psf"),
                            os.path.join(str(tmp_path), "spcfw_lipid_forpme.rtf"))
    T = t["n_types"]
    typ = {}
    for i, nm in enumerate(names):
        typ.setdefault(nm, int(t["type"][i]))
    s6 = 2.0 ** (1.0 / 6.0)
    eps = t["lj_eps"].reshape(T, T)
    sig2 = t["lj_sig2"].reshape(T, T)
    eps14 = t["lj14_eps"].reshape(T, T)
    sig214 = t["lj14_sig2"].reshape(T, T)
    ot, ht, c2 = typ["OT"], typ["HT"], typ["CTL2"]
    assert abs(eps[ot, ot] - 4 * 0.1554253) < 1e-12
    assert abs(sig2[ot, ot] - (2 * 1.7765722 / s6) ** 2) < 1e-10
    assert abs(eps[c2, c2] - 4 * 0.056) < 1e-12 and abs(sig2[c2, c2] - (2 * 2.010 / s6) ** 2) < 1e-10
    assert abs(eps[ot, c2] - 4 * math.sqrt(0.1554253 * 0.056)) < 1e-12
    assert abs(sig2[ot, c2] - ((1.7765722 + 2.010) / s6) ** 2) < 1e-10
    assert abs(eps14[c2, c2] - 4 * 0.01) < 1e-12 and abs(sig214[c2, c2] - (2 * 1.9 / s6) ** 2) < 1e-10
    assert eps[ht, ot] == 0.0 and abs(sig2[ht, ot] - (1.9927 / s6) ** 2) < 1e-10
    assert sig2[ot, ht] == sig2[ht, ot]
    # the OT-OT pair through the oracle: minimum at Rmin = 2 (Rmin/2), depth -eps, no force
    e, f = _lj_pair(eps[ot, ot], sig2[ot, ot], 2 * 1.7765722)
    assert abs(e + 0.1554253) < 1e-12 and abs(f) < 1e-11
    deck.close()


def test_constraint_virial_satisfies_the_rigid_molecule_identity():
    """For a rigid molecule, d^2/dt^2 sum_i m_i |r_i - R|^2 = 0 gives, at every
    instant, sum_i (r_i - R) . (F_i + G_i) = -sum_i m_i |v_i - V|^2 (G the
    constraint forces).  The trace of the constraint virial that the final
    RATTLE projection yields (the step's virial minus the force virial) meets
    it to first order in the time step: RATTLE's multipliers see the relative
    velocities half a step earlier (src/shake.cu:300-306 computes the same
    quantity from SHAKE)."""
    topo, xyz, params = systems.water_box(150, 17.0, seed=4)
    p0 = dict(params, cutoff=8.0, cuton=6.0, pairlist=9.0, K=(20, 20, 20))
    o = oracle.Oracle(topo, dict(p0, dt_ps=0.0005))
    u = o.to_fixed(xyz)
    v = o.maxwell(u, 300.0, 1, 0)
    u, v, _ = o.step(u, v, 40)
    m = np.asarray(topo["mass"])
    L = np.array(p0["box"])
    errs = []
    for dt, n in ((0.0005, 4), (0.00025, 8), (0.000125, 16)):
        od = oracle.Oracle(topo, dict(p0, dt_ps=dt))
        u2, v2, _, W, rc = od.step_ex(u, v, n)
        assert rc == 0
        f, _, Wf = od.forces(u2)
        Wc = W - Wf
        assert np.abs(Wc - Wc.T).max() < 1e-9 * np.abs(Wc).max()
        x = od.from_fixed(u2)
        rhs = 0.0
        for w in range(0, topo["n_atoms"], 3):
            idx = [w, w + 1, w + 2]
            xr = x[idx] - x[w]
            xr -= L * np.round(xr / L)
            M = m[idx].sum()
            R = (m[idx, None] * xr).sum(0) / M
            V = (m[idx, None] * v2[idx]).sum(0) / M
            rhs += -(m[idx] * ((v2[idx] - V) ** 2).sum(1)).sum() - ((xr - R) * f[idx]).sum()
        errs.append(abs(np.trace(Wc) / rhs - 1.0))
    assert errs[-1] < 1.5e-3
    assert 1.6 < errs[0] / errs[1] < 2.6 and 1.6 < errs[1] / errs[2] < 2.6, errs
