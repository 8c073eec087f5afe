#!/usr/bin/env python3
"""Benchmark of the per-step force-and-integrate path on one MI355X.

Headline line (value): ns/day of the ~17K-atom membrane-like slab (128 chains
x 38 beads + 4,046 rigid waters, 48x48x72 A, cutoff 12 A, PME 48x48x72,
dt 2 fs, NVT) in the mixed-precision form.  Beside it: the same system in
the plain FP64 form, the ~230K-atom slab (1,536 chains + 57,216 waters,
132^3 A, PME 128^3) in both forms and their ratio, mixed-vs-FP64 force /
energy / virial errors, per-phase device times, the pair kernel's roofline
and the FP64 CPU oracle timed on the host.

Systems are synthetic (seeded), relaxed with libfenzi's own steepest descent
before timing (untimed), then given Maxwell velocities at 300 K.  The timed
region is K steps of fz_step (graph-replayed steps, device-side neighbour
rebuilds) bracketed by device synchronisation; ns/day = 86400 / t_step[s] *
dt[fs] * 1e-6.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "fen-zi_amd"))

import numpy as np  # noqa: E402

import fenzi  # noqa: E402
import fz_synth  # noqa: E402

METRIC = "ns/day at 17K & 230K atoms; batched agg ns/day + p99 latency, mixed vs FP64"
FP32_PEAK = 157.3e12   # MI355X FP32 vector, published
FP64_PEAK = 78.6e12    # MI355X FP64 vector, published
HBM_PEAK = 8.0e12
PAIR_FLOP = 66         # algorithmic flop per non-excluded half-pair inside the cutoff


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def ns_day(ms, dt_ps):
    return 86400.0 / (ms * 1e-3) * dt_ps * 1e-3


def prepare(ctx, cid, min_steps):
    t0 = time.perf_counter()
    topo, xyz, p = fz_synth.config(cid)
    s = fenzi.System(ctx, topo, p, fenzi.FP64, xyz=xyz)
    s.minimize(min_steps, 0.05)
    x, _ = s.get_state()
    e = s.energies()
    s.close()
    log("workload %d: %d atoms generated+relaxed in %.1fs (Epot %.1f)" % (cid, topo["n_atoms"], time.perf_counter() - t0,
                                                                        e["potential"]))
    return topo, x, p


def md(ctx, topo, x, p, prec, steps, warmup, seed=1):
    s = fenzi.System(ctx, topo, p, prec, xyz=x)
    s.reassign_velocities(p["temperature"], seed, 0)
    s.step(warmup)
    t0 = time.perf_counter()
    s.step(steps)
    dt = time.perf_counter() - t0
    ms = dt * 1e3 / steps
    e = s.energies()
    ph = s.phase_times(10)
    out = dict(ms_per_step=ms, ns_per_day=ns_day(ms, p["dt_ps"]), temperature=e["temperature"],
               pairs_within_cutoff=int(e["pairs_within_cutoff"]), list_builds=int(e["list_builds"]), phases_ms=ph,
               device_bytes=s.info()["device_bytes"])
    s.close()
    return out


def compare_forms(ctx, topo, x, p):
    """Mixed-precision form against the FP64 form at the same positions."""
    res = {}
    a = fenzi.System(ctx, topo, p, fenzi.FP64, xyz=x)
    f64, e64 = a.forces()
    u = a.get_fixed()
    a.close()
    b = fenzi.System(ctx, topo, p, fenzi.MIXED)
    b.set_fixed(u)
    f32, e32 = b.forces()
    b.close()
    res["force_rel_rms"] = float(np.sqrt(((f32 - f64) ** 2).sum() / (f64 ** 2).sum()))
    res["energy_rel_err"] = {k: float(abs(e32[k] - e64[k]) / max(abs(e64[k]), 1e-12))
                             for k in ("vdw", "elec", "pme", "ewald_excl", "bond", "angle", "dihedral", "potential")}
    res["virial_rel_err"] = float(np.abs(e32["virial"] - e64["virial"]).max() / np.abs(e64["virial"]).max())
    return res


def roofline(run, prec):
    t = run["phases_ms"]["pair"] * 1e-3
    flops = PAIR_FLOP * run["pairs_within_cutoff"]
    ach = flops / t
    peak = FP32_PEAK if prec == fenzi.MIXED else FP64_PEAK
    return {"kernel": "pair (LJ + real-space Ewald, cluster-pair tiles)", "bound": "valu",
            "achieved": ach / 1e12, "peak": peak / 1e12, "unit": "TFLOP/s", "frac": ach / peak, "traffic": None,
            "flop_per_launch": flops, "launch_ms": run["phases_ms"]["pair"]}


def cpu_baseline(topo, x, p, budget_s=20.0):
    """FP64 oracle (C + OpenMP) on the host cores: velocity Verlet + SHAKE/RATTLE + PME."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle
    threads = oracle.set_threads(min(16, os.cpu_count() or 1))
    o = oracle.Oracle(topo, p)
    u = o.to_fixed(x)
    v = o.maxwell(u, p["temperature"], 1, 0)
    t0 = time.perf_counter()
    u, v, _ = o.step(u, v, 1)
    t1 = max(time.perf_counter() - t0, 1e-3)
    n = max(2, min(200, int(budget_s / t1)))
    t0 = time.perf_counter()
    o.step(u, v, n)
    ms = (time.perf_counter() - t0) * 1e3 / n
    return {"value": ns_day(ms, p["dt_ps"]), "unit": "ns/day", "cores": threads, "kind": "port",
            "sample": "%d steps of the %d-atom slab (FP64 oracle, C/OpenMP: switched LJ + erfc real space, "
                      "order-4 PME with own FFT, bonded, SHAKE/RATTLE); %.1f ms/step" % (n, topo["n_atoms"], ms)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=300)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--no-large", action="store_true", help="skip the 230K-atom system")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline")
    ap.add_argument("--min-steps", type=int, default=300)
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py measures one GPU: this build runs one process on one MI355X; "
              "multi-GPU runs are not part of it (--gpus %d requested)" % a.gpus, file=sys.stderr)
        sys.exit(2)
    ctx = fenzi.Context(0)
    topo1, x1, p1 = prepare(ctx, 1, a.min_steps)
    r1m = md(ctx, topo1, x1, p1, fenzi.MIXED, a.steps, a.warmup)
    log("17K mixed: %.3f ms/step %.1f ns/day" % (r1m["ms_per_step"], r1m["ns_per_day"]))
    r1d = md(ctx, topo1, x1, p1, fenzi.FP64, a.steps, a.warmup)
    log("17K fp64 : %.3f ms/step %.1f ns/day" % (r1d["ms_per_step"], r1d["ns_per_day"]))
    err1 = compare_forms(ctx, topo1, x1, p1)
    extra = {"slab17k_mixed": r1m, "slab17k_fp64": r1d, "slab17k_mixed_vs_fp64": err1,
             "slab17k_mixed_over_fp64": r1d["ms_per_step"] / r1m["ms_per_step"]}
    if not a.no_large:
        topo2, x2, p2 = prepare(ctx, 2, a.min_steps)
        s2 = max(50, a.steps // 3)
        r2m = md(ctx, topo2, x2, p2, fenzi.MIXED, s2, max(10, a.warmup // 2))
        log("230K mixed: %.3f ms/step %.1f ns/day" % (r2m["ms_per_step"], r2m["ns_per_day"]))
        r2d = md(ctx, topo2, x2, p2, fenzi.FP64, s2, max(10, a.warmup // 2))
        log("230K fp64 : %.3f ms/step %.1f ns/day" % (r2d["ms_per_step"], r2d["ns_per_day"]))
        extra.update({"slab230k_mixed": r2m, "slab230k_fp64": r2d,
                      "slab230k_mixed_vs_fp64": compare_forms(ctx, topo2, x2, p2),
                      "slab230k_mixed_over_fp64": r2d["ms_per_step"] / r2m["ms_per_step"]})
    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(topo1, x1, p1)
        log("cpu oracle: %.3f ns/day on %d threads" % (cpu["value"], cpu["cores"]))
    line = {
        "metric": METRIC, "value": r1m["ns_per_day"], "unit": "ns/day", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": r1m["ms_per_step"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "17,002-atom membrane-like slab (128 x 38-bead chains + 4,046 rigid waters), "
                               "48x48x72 A, cutoff 12 A (switch 10, list 13.5), PME 48x48x72 order 4, "
                               "beta 0.26, dt 2 fs, NVT 300 K (velocity reassignment every 500 steps), "
                               "mixed precision (FP32 arithmetic, int64 fixed-point sums, FP64 SHAKE)",
                   "atoms": int(topo1["n_atoms"]), "precision": "mixed"},
        "roofline": roofline(r1m, fenzi.MIXED),
        "cpu_baseline": cpu,
        "extra": extra,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
