"""Host-side tests that need no GPU: the C ABI library loads and exports every
declared symbol, the input-deck parsers read the reference's own deck, the
energy table has the reference's format, and a written deck reads back into
the same physics."""
import ctypes
import os
import re

import numpy as np
import pytest

import fenzi
import fz_synth
import oracle
import systems

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "fenzi.h")).read()
    declared = set(re.findall(r"\b(fz_[a-z0-9_]+)\s*\(", hdr))
    L = ctypes.CDLL(fenzi.LIB_PATH)
    missing = [s for s in sorted(declared) if not hasattr(L, s)]
    assert not missing, missing
    assert declared <= set(fenzi.EXPORTS) | {"fz_last_error"}


def test_reference_deck_dmpc_small(tmp_path):
    deck = fenzi.Deck(systems.dmpc_small_deck(str(tmp_path)))
    # counts stated in the deck's own header comments and PSF section headers
    assert deck.psf_counts() == dict(bonds=14096, angles=19108, dihedrals=22536, impropers=144)
    p = deck.params()
    assert p["box"] == (46.8, 46.8, 76.0) and p["K"] == (64, 64, 64)
    assert (p["cutoff"], p["cuton"], p["pairlist"], p["beta"], p["dt_ps"]) == (8.0, 7.0, 9.5, 0.32, 0.001)
    assert p["shake_tol"] == 1e-10
    s = deck.settings()
    assert (s["temperature"], s["seed"], s["tempreset_every"], s["output_every"], s["shake"]) == (298.0, 2345, 20000, 100, 1)
    t = deck.topology()
    assert t["n_atoms"] == 17004
    x = deck.coordinates()
    assert x.shape == (17004, 3) and np.allclose(x[0], [13.904, -17.451, 17.063])
    assert abs(t["charge"].sum()) < 1e-6
    # every constraint holds a hydrogen (mass <= 1.5) and has a bond length
    ca = t["constraint_atoms"].reshape(-1, 2)
    assert len(ca) > 0 and np.all(np.minimum(t["mass"][ca[:, 0]], t["mass"][ca[:, 1]]) <= 1.5)
    assert np.all(t["constraint_length"] > 0.9)
    assert t["n_dihedrals"] >= 22536 and t["n_impropers"] == 144
    assert np.all(t["lj_eps"] >= 0) and np.allclose(t["lj_eps"].reshape(16, 16), t["lj_eps"].reshape(16, 16).T)


def test_energy_table_format():
    hdr = fenzi.energy_header()
    ref = "%-9s %-10s %-10s %-11s %-11s " % ("Step", "Time (ps)", "Temp (K)", "bond", "angle")
    ref += "%-11s " % "urey-b" + "%-8s " % "impr"
    ref += "%-11s %-11s %-13s " % ("dihed", "VDW", "elec")
    ref += "%-11s %-13s %-13s " % ("PME", "Ewexcl", "Ewself")
    ref += "%-13s %-13s %-13s\n" % ("sum PE", "Kinetic", "Potential")
    assert hdr == ref
    e = fenzi.Energies()
    e.step, e.temperature, e.bond, e.pme, e.potential, e.kinetic = 100, 298.5, 12.25, -5.5, -1234.5, 567.25
    row = fenzi.energy_row(e, 0.001)
    assert row == ("%-9d %-10.4f %-10.4f %-11.4f %-11.4f " % (100, 0.1, 298.5, 12.25, 0.0) +
                   "%-11.4f " % 0.0 + "%-8.4f " % 0.0 + "%-11.4f %-11.4f %-13.4f " % (0.0, 0.0, 0.0) +
                   "%-11.4f %-13.4f %-13.4f " % (-5.5, 0.0, 0.0) + "%-13.4f %-13.4f %-13.4f\n" % (-1234.5, 567.25, -1234.5))


def test_written_deck_reads_back_into_the_same_physics(tmp_path):
    topo, xyz = fz_synth.make_system(3, 150, (24.0, 24.0, 24.0), seed=6)
    params = dict(box=(24.0, 24.0, 24.0), cutoff=9.0, cuton=7.0, pairlist=10.5, beta=0.33, K=(24, 24, 24),
                  dt_ps=0.002, shake_tol=1e-8)
    path = fz_synth.write_deck(topo, xyz, params, str(tmp_path), "rt")
    deck = fenzi.Deck(path)
    t2 = deck.topology()
    p2 = deck.params()
    assert t2["n_atoms"] == topo["n_atoms"] and t2["n_constraints"] == topo["n_constraints"]
    assert p2["K"] == params["K"] and p2["box"] == params["box"]
    o1 = oracle.Oracle(topo, params)
    o2 = oracle.Oracle(t2, dict(params, **{k: p2[k] for k in ("cutoff", "cuton", "pairlist", "beta")}))
    f1, e1, _ = o1.forces(o1.to_fixed(xyz))
    f2, e2, _ = o2.forces(o2.to_fixed(deck.coordinates()))
    assert systems.rel_rms(f2, f1) < 1e-7
    for k in e1:
        assert abs(e1[k] - e2[k]) <= 1e-7 * max(1.0, abs(e1[k])), k


def test_invalid_inputs_are_rejected_without_a_device(tmp_path):
    """Parse errors come back as codes and messages; nothing exits."""
    bad = tmp_path / "bad.in"
    bad.write_text("coordinates nothere.crd\nstructure nothere.psf\nparameters nothere.prm\n"
                   "cellBasisVector1 10 0 0\ncellBasisVector2 0 10 0\ncellBasisVector3 0 0 10\n")
    try:
        fenzi.Deck(str(bad))
    except fenzi.FenziError as e:
        assert "not found" in str(e)
    else:
        raise AssertionError("missing files accepted")


# ---- checkpoint (.rst) and trajectory (.dcd) files -------------------------------------
def _state(n=37, seed=3):
    rng = np.random.default_rng(seed)
    return rng.uniform(-20, 20, (n, 3)), rng.normal(0, 0.02, (n, 3)), np.array([40.5, 41.25, 60.0])


def test_binary_restart_layout_and_round_trip(tmp_path):
    import struct
    x, v, box = _state()
    p = str(tmp_path / "a.rst")
    fenzi.restart_write(p, 1200, 2345.0, box, x, v, binary=True)
    raw = open(p, "rb").read()
    # SaveCheckpointBinary: int natom, int step, double seed, float3 box, float3 zero, 6 floats per atom
    assert len(raw) == 4 + 4 + 8 + 12 + 12 + 24 * len(x)
    n, step, seed = struct.unpack_from("<iid", raw, 0)
    assert (n, step, seed) == (len(x), 1200, 2345.0)
    assert np.allclose(struct.unpack_from("<3f", raw, 16), box.astype(np.float32), rtol=0, atol=0)
    assert struct.unpack_from("<3f", raw, 28) == (0.0, 0.0, 0.0)
    rec = np.frombuffer(raw, dtype="<f4", offset=40).reshape(-1, 6)
    assert np.array_equal(rec[:, :3], x.astype(np.float32)) and np.array_equal(rec[:, 3:], v.astype(np.float32))
    r = fenzi.restart_read(p, len(x), binary=True)
    assert r["step"] == 1200 and r["seed"] == 2345.0
    assert np.array_equal(r["xyz"], x.astype(np.float32).astype(np.float64))
    assert np.array_equal(r["vel"], v.astype(np.float32).astype(np.float64))
    with pytest.raises(fenzi.FenziError):
        fenzi.restart_read(p, len(x) + 1, binary=True)
    open(p, "wb").write(raw[:-10])
    with pytest.raises(fenzi.FenziError):
        fenzi.restart_read(p, len(x), binary=True)


def test_ascii_restart_layout_and_round_trip(tmp_path):
    x, v, box = _state(5, seed=4)
    p = str(tmp_path / "a.rst")
    fenzi.restart_write(p, 77, 12.5, box, x, v, binary=False)
    lines = open(p).read().split("\n")
    # SaveCheckpointAscii: "%12d \n", "%12d \n", "%le \n", box + zeros, "%10f" x 6 per atom
    assert lines[0] == "%12d " % 5 and lines[1] == "%12d " % 77 and lines[2] == "%e " % 12.5
    assert lines[3] == "%12f %12f %12f %12f %12f %12f" % (40.5, 41.25, 60.0, 0, 0, 0)
    a = [float(t) for t in lines[4].split()]
    assert np.allclose(a, np.concatenate([x[0], v[0]]), atol=6e-7)
    r = fenzi.restart_read(p, 5, binary=False)
    assert r["step"] == 77 and r["seed"] == 12.5
    assert np.allclose(r["xyz"], x, atol=6e-7) and np.allclose(r["vel"], v, atol=6e-7)


def _fortran_records(raw):
    import struct
    out, k = [], 0
    while k < len(raw):
        (n,) = struct.unpack_from("<I", raw, k)
        body = raw[k + 4:k + 4 + n]
        (m,) = struct.unpack_from("<I", raw, k + 4 + n)
        assert m == n
        out.append(body)
        k += 8 + n
    return out


def test_dcd_layout(tmp_path):
    import struct
    x, _, box = _state(11, seed=5)
    p = str(tmp_path / "t.dcd")
    d = fenzi.Dcd(p, len(x), nframes=2, first_step=50, every=50, nsteps=100, dt_ps=0.002)
    d.write(x, box)
    d.write(x + 1.0, box)
    d.close()
    recs = _fortran_records(open(p, "rb").read())
    assert len(recs) == 3 + 2 * 4
    hdr = recs[0]
    assert hdr[:4] == b"CORD" and len(hdr) == 84
    ints = struct.unpack_from("<20i", hdr, 4)
    assert ints[0:4] == (2, 50, 50, 100) and ints[7] == 1 and ints[8] == 0 and ints[10] == 1 and ints[19] == 34
    assert abs(struct.unpack_from("<f", hdr, 4 + 9 * 4)[0] - 0.002 / 4.888821e-2) < 1e-7
    assert struct.unpack_from("<i", recs[1], 0)[0] == 1 and len(recs[1]) == 84
    assert struct.unpack_from("<i", recs[2], 0)[0] == 11
    for f, shift in ((0, 0.0), (1, 1.0)):
        cell = struct.unpack("<6d", recs[3 + 4 * f])
        assert cell == (40.5, 0.0, 41.25, 0.0, 0.0, 60.0)
        for k in range(3):
            col = np.frombuffer(recs[4 + 4 * f + k], dtype="<f4")
            assert np.array_equal(col, (x[:, k] + shift).astype(np.float32))


def test_deck_restraint_and_fixed_segments(tmp_path):
    """consharm / consharmdist / consharmfc select groups by the CRD segment name,
    consfix fixes atoms by the PSF segment name (both case-insensitive, as the
    reference's cicompare)."""
    topo, xyz, params = systems.slab(n_chains=3, n_waters=40, L=22.0, seed=2, K=(24, 24, 24))
    n = topo["n_atoms"]
    crd = np.array(["BULK"] * n, dtype=object)
    crd[:38] = "MEMA"
    crd[38:76] = "memb"
    psf = np.array(["SYS"] * n, dtype=object)
    psf[76:114] = "FIX1"
    psf[-6:] = "fix2"
    path = fz_synth.write_deck(topo, xyz, dict(params, dt_ps=0.001), str(tmp_path), "seg", psf_seg=psf, crd_seg=crd,
                               extra=["consharm mema MEMB", "consharmdist 1.5 -2 3", "consharmfc 0.25",
                                      "consfix 2 fix1 FIX2"])
    deck = fenzi.Deck(path)
    group, dist, fc = deck.com_restraint()
    assert list(np.flatnonzero(group == 0)) == list(range(38))
    assert list(np.flatnonzero(group == 1)) == list(range(38, 76))
    assert (group[76:] == -1).all()
    assert np.allclose(dist, [1.5, -2.0, 3.0]) and fc == 0.25
    fixed = deck.fixed_atoms()
    assert list(np.flatnonzero(fixed)) == list(range(76, 114)) + list(range(n - 6, n))
    deck.close()
    # without the keywords: no restraint, nothing fixed
    path = fz_synth.write_deck(topo, xyz, dict(params, dt_ps=0.001), str(tmp_path), "plain")
    deck = fenzi.Deck(path)
    assert deck.com_restraint() is None
    assert deck.fixed_atoms().sum() == 0
    deck.close()
