"""Small seeded test systems shared by the CPU and GPU tests."""
import numpy as np

import fz_synth


def water_box(n_waters=300, L=21.0, seed=11):
    topo, xyz = fz_synth.make_system(0, n_waters, (L, L, L), seed=seed)
    params = dict(box=(L, L, L), cutoff=9.0, cuton=7.0, pairlist=10.5, beta=0.34, K=(24, 24, 24),
                  dt_ps=0.001, shake_tol=1e-10)
    return topo, xyz, params


def slab(n_chains=6, n_waters=400, L=26.0, seed=5, K=(27, 27, 27)):
    topo, xyz = fz_synth.make_system(n_chains, n_waters, (L, L, L), seed=seed)
    params = dict(box=(L, L, L), cutoff=10.0, cuton=8.0, pairlist=11.5, beta=0.32, K=K,
                  dt_ps=0.001, shake_tol=1e-10)
    return topo, xyz, params


def pair_count_ok(count, orc, u, rc, mixed, delta=4e-5):
    """The device's count of non-excluded pairs inside the cutoff against the
    oracle's FP64 count: equal in the FP64 form; in the mixed form, which tests
    r^2 < rc^2 in FP32 arithmetic (relative rounding of a few 1e-7 at the
    cutoff, well inside delta / rc), between the FP64 counts within rc - delta
    and rc + delta."""
    if not mixed:
        return count == orc.count_pairs_within(u, rc)
    return orc.count_pairs_within(u, rc - delta) <= count <= orc.count_pairs_within(u, rc + delta)


def rel_rms(a, b):
    return float(np.sqrt(((a - b) ** 2).sum() / (b ** 2).sum()))


def dmpc_small_deck(tmpdir):
    """Unpack the reference's dmpc_small input deck (tests/examples_fenzi/dmpc_small,
    kept here as a gzip'd fixture) and return the path of its .in file."""
    import gzip
    import os
    import shutil
    src = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "dmpc_small")
    for f in os.listdir(src):
        if f.endswith(".gz"):
            with gzip.open(os.path.join(src, f), "rb") as a, open(os.path.join(tmpdir, f[:-3]), "wb") as b:
                shutil.copyfileobj(a, b)
        else:
            shutil.copy(os.path.join(src, f), os.path.join(tmpdir, f))
    return os.path.join(tmpdir, "fenzi.in")
