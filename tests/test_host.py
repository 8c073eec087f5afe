"""Host-side tests that need no GPU: the C ABI library loads and exports every
declared symbol, the input-deck parsers read the reference's own deck, the
energy table has the reference's format, and a written deck reads back into
the same physics."""
import ctypes
import os
import re

import numpy as np

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
