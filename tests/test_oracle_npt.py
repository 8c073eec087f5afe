"""Pins of the oracle's Andersen barostat (or_step_npt): the volume V follows
V'' = (P - p_ref) / M, P = (2 KE + tr W) / 3V, the piston of the reference's
reduce_virial (src/cell_based_kernels.cu:403-600; its NPT step is a stub,
src/npt_kernels.cu:26-33).

* An infinitely heavy piston never moves the box, and the step is then the
  fixed-box step (or_step_ex) bit for bit.
* A free gas (no charges, no LJ, no bonds) has U = W = 0: every particle
  momentum p = m L v is conserved, so KE V^(2/3) is invariant, and V follows the
  one-dimensional Hamiltonian K0 V^(-2/3) + p_ref V + Pi^2 / 2M; the oracle's V(t)
  matches a fine Runge-Kutta solution of that ODE.
* With interactions and no constraints, every piece of the splitting is an
  exact flow of the extended Hamiltonian KE + U + p_ref V + Pi^2 / 2M: it is
  conserved as well as the fixed-box step conserves KE + U, and the step is time
  reversible.  A sign or factor error in the piston's forces (pressure, kinetic
  term) breaks the conservation."""
import os

import numpy as np

import fz_synth
import oracle

PU = oracle.PU


def _free_gas(n=200, L=30.0, seed=3):
    rng = np.random.default_rng(seed)
    z1 = lambda k: np.zeros(k, np.int32)
    topo = dict(n_atoms=n, n_types=1, charge=np.zeros(n), mass=np.full(n, 14.0), type=z1(n),
                lj_eps=np.zeros(1), lj_sig2=np.ones(1), lj14_eps=np.zeros(1), lj14_sig2=np.ones(1),
                bond_atoms=z1(0), bond_k=np.zeros(0), bond_r0=np.zeros(0), angle_atoms=z1(0), angle_k=np.zeros(0),
                angle_theta0=np.zeros(0), ub_k=np.zeros(0), ub_s0=np.zeros(0), dihedral_atoms=z1(0),
                dihedral_k=np.zeros(0), dihedral_n=z1(0), dihedral_delta=np.zeros(0), improper_atoms=z1(0),
                improper_k=np.zeros(0), improper_psi0=np.zeros(0), constraint_atoms=z1(0),
                constraint_length=np.zeros(0))
    for k in ("bonds", "angles", "dihedrals", "impropers", "constraints"):
        topo["n_" + k] = 0
    xyz = rng.uniform(-L / 2, L / 2, (n, 3))
    params = dict(box=(L, L, L), cutoff=9.0, cuton=7.0, pairlist=10.0, beta=0.3, K=(16, 16, 16), dt_ps=0.002,
                  shake_tol=1e-10)
    return topo, xyz, params


def _hext(orc, v, e, pc):
    return orc.kinetic(v) + sum(e.values()) + pc.p_ref * pc.V + pc.Pi ** 2 / (2.0 * pc.mass)


def test_heavy_piston_is_the_fixed_box_step():
    # one thread: the oracle's threaded force sums are not bitwise reproducible
    oracle.set_threads(1)
    try:
        _heavy_piston()
    finally:
        oracle.set_threads(os.cpu_count() or 1)


def _heavy_piston():
    topo, xyz = fz_synth.make_system(3, 150, (24.0, 24.0, 24.0), seed=21)
    params = dict(box=(24.0, 24.0, 24.0), cutoff=9.0, cuton=7.0, pairlist=10.5, beta=0.34, K=(24, 24, 24),
                  dt_ps=0.001, shake_tol=1e-10)
    orc = oracle.Oracle(topo, params)
    u, _ = orc.minimize_sd(orc.to_fixed(xyz), 60, 0.05)
    v = orc.maxwell(u, 300.0, seed=2, counter=0)
    u_nve, v_nve, e_nve, w_nve, rc = orc.step_ex(u, v, 12)
    assert rc == 0
    pc = orc.piston(u, 1.0, 1e300)
    u_npt, v_npt, e_npt, w_npt, rc = orc.step_npt(u, v, 12, pc)
    assert rc == 0
    assert pc.V == 24.0 ** 3
    assert np.array_equal(u_npt, u_nve) and np.array_equal(v_npt, v_nve)
    assert np.array_equal(w_npt, w_nve)


def _gas_volumes(dt_ps, t_total, n_out, pref_atm, M):
    topo, xyz, params = _free_gas()
    params = dict(params, dt_ps=dt_ps)
    orc = oracle.Oracle(topo, params)
    u = orc.to_fixed(xyz)
    v = orc.maxwell(u, 300.0, seed=5, counter=0)
    pc = orc.piston(u, pref_atm, M)
    assert pc.Wtr == 0.0
    K0 = orc.kinetic(v) * pc.V ** (2.0 / 3.0)
    steps = int(round(t_total / dt_ps))
    chunk = steps // n_out
    Vs = [pc.V]
    for _ in range(n_out):
        u, v, e, w, rc = orc.step_npt(u, v, chunk, pc)
        assert rc == 0
        # the exact invariant of the discrete step (rounding only)
        assert abs(orc.kinetic(v) * pc.V ** (2.0 / 3.0) - K0) < 1e-12 * K0
        Vs.append(pc.V)
    return np.array(Vs), K0


def test_free_gas_volume_follows_the_reduced_hamiltonian():
    """V(t) against a Runge-Kutta solution of V' = Pi / M,
    Pi' = 2 K0 V^(-5/3) / 3 - p_ref, at steps of 2 fs and 1 fs: the error is
    small and falls about fourfold when the step halves (second order)."""
    pref_atm, M, t_total, n_out = 400.0, 2e-7, 0.8, 20
    V2, K0 = _gas_volumes(0.002, t_total, n_out, pref_atm, M)
    V1, _ = _gas_volumes(0.001, t_total, n_out, pref_atm, M)
    p = pref_atm * PU

    def f(y):
        return np.array([y[1] / M, 2.0 * K0 * y[0] ** (-5.0 / 3.0) / 3.0 - p])

    y = np.array([V2[0], 0.0])
    nsub = 20000
    h = 20.4548 * t_total / nsub
    ref = [y[0]]
    for k in range(nsub):
        k1 = f(y); k2 = f(y + 0.5 * h * k1); k3 = f(y + 0.5 * h * k2); k4 = f(y + h * k3)
        y = y + h / 6.0 * (k1 + 2 * k2 + 2 * k3 + k4)
        if (k + 1) % (nsub // n_out) == 0:
            ref.append(y[0])
    ref = np.array(ref)
    span = np.ptp(ref)
    assert span > 1e-2 * ref[0]                     # the piston moves
    e2, e1 = np.abs(V2 - ref).max(), np.abs(V1 - ref).max()
    assert e2 < 2e-2 * span, (e2, span)
    assert 3.0 < e2 / e1 < 5.0, (e2, e1)


def test_unconstrained_extended_energy_and_reversibility():
    box = (26.0, 26.0, 26.0)
    topo, xyz = fz_synth.make_system(10, 0, box, seed=5)
    params = dict(box=box, cutoff=10.0, cuton=8.0, pairlist=11.5, beta=0.32, K=(27, 27, 27), dt_ps=0.001,
                  shake_tol=1e-12)
    orc = oracle.Oracle(topo, params)
    u0, _ = orc.minimize_sd(orc.to_fixed(xyz), 150, 0.05)
    v0 = orc.maxwell(u0, 300.0, seed=1, counter=0)
    # fixed box: range of KE + U over 100 steps
    u, v = u0.copy(), v0.copy()
    h_nve = []
    for _ in range(10):
        u, v, e, w, rc = orc.step_ex(u, v, 10)
        h_nve.append(orc.kinetic(v) + sum(e.values()))
    # barostat: range of the extended energy over the same steps
    pc = orc.piston(u0, 1.0, 1e-4)
    u, v = u0.copy(), v0.copy()
    h_npt = []
    for _ in range(10):
        u, v, e, w, rc = orc.step_npt(u, v, 10, pc)
        h_npt.append(_hext(orc, v, e, pc))
    assert abs(pc.V - box[0] ** 3) > 1.0          # the box moved
    assert np.ptp(h_npt) < 1.5 * np.ptp(h_nve) + 1e-3, (np.ptp(h_npt), np.ptp(h_nve))
    # reversibility: 20 steps, reverse velocities and piston momentum, 20 steps
    orc.set_box(box)
    pc = orc.piston(u0, 1.0, 1e-4)
    u1, v1, _, _, _ = orc.step_npt(u0, v0, 20, pc)
    V1 = pc.V
    pc.Pi = -pc.Pi
    u2, v2, _, _, _ = orc.step_npt(u1, -v1, 20, pc)
    d = (u2 - u0).view(np.int64).astype(np.float64) * box[0] / 2.0 ** 64
    assert abs(V1 - box[0] ** 3) > 0.1
    assert np.abs(d).max() < 1e-9 and np.abs(v2 + v0).max() < 1e-9
    assert abs(pc.V - box[0] ** 3) < 1e-9 * box[0] ** 3
