"""GPU parity of the COM restraint (consharm) and fixed atoms (consfix) against
the FP64 oracle, through the C ABI."""
import os
import subprocess

import numpy as np
import pytest

import fenzi
import fz_synth
import oracle
import systems

pytestmark = pytest.mark.gpu


def _slab():
    return systems.slab(n_chains=4, n_waters=300, L=24.0, seed=7, K=(24, 24, 24))


def _groups(n):
    g = np.full(n, -1, np.int32)
    g[:38] = 0        # chain 0
    g[38:76] = 1      # chain 1
    return g


def _fixed_mask(topo):
    n = topo["n_atoms"]
    m = np.zeros(n, np.uint8)
    m[76:114] = 1                 # chain 2
    w0 = 4 * 38                   # first water
    m[w0:w0 + 30] = 1             # ten whole waters
    m[w0 + 30] = 1                # the oxygen of the eleventh
    return m


@pytest.mark.parametrize("prec", [fenzi.FP64, fenzi.MIXED])
def test_com_restraint_forces_vs_oracle(ctx, prec):
    topo, xyz, params = _slab()
    orc = oracle.Oracle(topo, params)
    u = orc.to_fixed(xyz)
    g, dist, fc = _groups(topo["n_atoms"]), (2.0, -1.0, 9.0), 0.8
    s = fenzi.System(ctx, topo, params, prec)
    s.set_fixed(u)
    f_plain, e_plain = s.forces()
    s.set_com_restraint(g, dist, fc)
    f, e = s.forces()
    orc.set_com_restraint(g, dist, fc)
    f_ref, _, _ = orc.forces(u)
    f_res, _, _ = orc.forces(u, flags=oracle.RESTRAINT)
    assert np.abs(f_res[38:76]).max() > 1.0  # the restraint acts
    assert systems.rel_rms(f, f_ref) < (1e-10 if prec == fenzi.FP64 else 1e-5)
    # the restraint's own contribution, isolated on the device by the difference
    assert np.abs((f - f_plain) - f_res).max() < 1e-8
    assert e["potential"] == pytest.approx(e_plain["potential"], rel=1e-12, abs=1e-9)  # no energy term
    s.set_com_restraint(None)
    f_off, _ = s.forces()
    assert np.abs(f_off - f_plain).max() < 1e-9
    s.close()


@pytest.mark.parametrize("prec", [fenzi.FP64, fenzi.MIXED])
def test_fixed_atoms_md_vs_oracle(ctx, prec):
    topo, xyz, params = _slab()
    orc = oracle.Oracle(topo, params)
    u = orc.to_fixed(xyz)
    fixed = _fixed_mask(topo)
    orc.set_fixed(fixed)
    s = fenzi.System(ctx, topo, params, prec)
    s.set_fixed(u)
    s.set_fixed_atoms(fixed)
    v0 = orc.maxwell(u, 300.0, seed=5, counter=0)
    s.reassign_velocities(300.0, 5, 0)
    _, v_dev = s.get_state()
    assert np.abs(v_dev[fixed == 1]).max() == 0.0
    assert systems.rel_rms(v_dev, v0) < 1e-10
    u_ref, v_ref, rc = orc.step(u, v0, 10)
    assert rc == 0
    s.step(10)
    u_dev = s.get_fixed()
    assert np.array_equal(u_dev[fixed == 1], u[fixed == 1])  # fixed atoms did not move at all
    x_dev, v_dev = s.get_state()
    d = x_dev - orc.from_fixed(u_ref)
    d -= np.array(params["box"]) * np.round(d / np.array(params["box"]))
    assert np.abs(d).max() < (1e-9 if prec == fenzi.FP64 else 2e-5)
    assert systems.rel_rms(v_dev, v_ref) < (1e-8 if prec == fenzi.FP64 else 1e-4)
    res, vres = orc.constraint_residual(u_dev, v_dev)
    assert res < 1e-9 and vres < 1e-6
    e = s.energies()
    assert e["temperature"] == pytest.approx(orc.temperature(v_dev), rel=1e-6)
    s.close()


@pytest.mark.parametrize("prec,tol", [(fenzi.FP64, 1e-7), (fenzi.MIXED, 1e-4)])
def test_minimiser_with_fixed_atoms_and_restraint_vs_oracle(ctx, prec, tol):
    topo, xyz, params = systems.slab(n_chains=3, n_waters=250, L=22.0, seed=12, K=(24, 24, 24))
    orc = oracle.Oracle(topo, params)
    u = orc.to_fixed(xyz)
    fixed = np.zeros(topo["n_atoms"], np.uint8)
    fixed[76:114] = 1
    g = _groups(topo["n_atoms"])
    s = fenzi.System(ctx, topo, params, prec)
    s.set_fixed(u)
    s.set_fixed_atoms(fixed)
    s.set_com_restraint(g, (1.0, 1.0, 4.0), 0.05)
    orc.set_fixed(fixed)
    orc.set_com_restraint(g, (1.0, 1.0, 4.0), 0.05)
    s.minimize_ex(30, fenzi.min_params(1, cgfac=2e-6, sdfac=1e-6, cgrate=0.1, sdrate=0.01))
    u_ref, _, done = orc.minimize_ref(u, 30, 1, 2e-6, 1e-6, 0.1, 0.01)
    assert done > 0
    u_dev = s.get_fixed()
    assert np.array_equal(u_dev[fixed == 1], u[fixed == 1])
    d = (u_dev - u_ref).view(np.int64).astype(np.float64) * (params["box"][0] / 2.0 ** 64)
    assert np.abs(d).max() < tol
    s.close()


def test_fenzi_driver_with_restraint_and_fixed_segments(tmp_path):
    topo, xyz, params = _slab()
    n = topo["n_atoms"]
    crd = np.array(["BULK"] * n, dtype=object)
    crd[:38], crd[38:76] = "A", "B"
    psf = np.array(["SYS"] * n, dtype=object)
    psf[76:114] = "FIXD"
    prefix = str(tmp_path / "rst_run")
    path = fz_synth.write_deck(topo, xyz, dict(params, dt_ps=0.001), str(tmp_path), "rst", steps=40,
                               output_every=20, psf_seg=psf, crd_seg=crd,
                               extra=["consharm A B", "consharmdist 1 1 5", "consharmfc 0.2", "consfix 1 FIXD",
                                      "outputtrajectory %s.dcd" % prefix, "printlevel 1", "checkpointTiming 40"])
    exe = os.path.join(os.path.dirname(fenzi.LIB_PATH), "fenzi")
    r = subprocess.run([exe, "-d", "0", "-f", path], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    rows = [l.split() for l in r.stdout.splitlines()[1:] if l and l[0].isdigit()]
    assert [int(w[0]) for w in rows] == [0, 20, 40]
    rst = fenzi.restart_read(prefix + ".rst", n, binary=True)
    assert rst["step"] == 40
    box = np.array(params["box"])
    # fixed atoms end where they started (the restart file stores float32)
    d = rst["xyz"][76:114] - xyz[76:114]
    d -= box * np.round(d / box)
    assert np.abs(d).max() < 1e-4
    assert np.abs(rst["vel"][76:114]).max() == 0.0
    moved = rst["xyz"][:76] - xyz[:76]
    moved -= box * np.round(moved / box)
    assert np.abs(moved).max() > 1e-3
