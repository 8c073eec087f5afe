"""ctypes wrapper of liboracle.so -- the FP64 CPU checker.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")

E_NAMES = ["bond", "angle", "ub", "impr", "dihed", "vdw", "elec", "pme", "ewexcl", "ewself"]
REAL, EXCL, BONDED, PME, EWALD, SELF, RESTRAINT = 1, 2, 4, 8, 16, 32, 64
ALL = REAL | EXCL | BONDED | PME | SELF | RESTRAINT
RC = 0.001987  # kcal/(mol K), src/defs.in.h:156

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_up = C.POINTER(C.c_uint64)


class Topo(C.Structure):
    _fields_ = [("n_atoms", C.c_int32), ("n_types", C.c_int32),
                ("charge", _dp), ("mass", _dp), ("type", _ip),
                ("lj_eps", _dp), ("lj_sig2", _dp), ("lj14_eps", _dp), ("lj14_sig2", _dp),
                ("n_bonds", C.c_int32), ("bond_atoms", _ip), ("bond_k", _dp), ("bond_r0", _dp),
                ("n_angles", C.c_int32), ("angle_atoms", _ip), ("angle_k", _dp), ("angle_theta0", _dp),
                ("ub_k", _dp), ("ub_s0", _dp),
                ("n_dihedrals", C.c_int32), ("dihedral_atoms", _ip), ("dihedral_k", _dp),
                ("dihedral_n", _ip), ("dihedral_delta", _dp),
                ("n_impropers", C.c_int32), ("improper_atoms", _ip), ("improper_k", _dp),
                ("improper_psi0", _dp),
                ("n_constraints", C.c_int32), ("constraint_atoms", _ip), ("constraint_length", _dp)]


class Params(C.Structure):
    _fields_ = [("box", C.c_double * 3), ("cutoff", C.c_double), ("cuton", C.c_double),
                ("pairlist", C.c_double), ("beta", C.c_double), ("K", C.c_int32 * 3),
                ("dt_ps", C.c_double), ("shake_tol", C.c_double), ("ewald_kmax_tol", C.c_double)]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.or_create.restype = C.c_void_p
        L.or_create.argtypes = [C.POINTER(Topo), C.POINTER(Params)]
        L.or_destroy.argtypes = [C.c_void_p]
        L.or_set_threads.argtypes = [C.c_int]
        L.or_set_threads.restype = C.c_int
        L.or_to_fixed.argtypes = [_dp, _dp, _up, C.c_int64]
        L.or_from_fixed.argtypes = [_dp, _up, _dp, C.c_int64]
        L.or_forces.argtypes = [C.c_void_p, _up, _dp, _dp, _dp, C.c_int]
        L.or_n_pairs.argtypes = [C.c_void_p, C.c_int]
        L.or_n_pairs.restype = C.c_int64
        L.or_get_pairs.argtypes = [C.c_void_p, C.c_int, _ip]
        L.or_count_pairs_within.argtypes = [C.c_void_p, _up, C.c_double]
        L.or_count_pairs_within.restype = C.c_int64
        L.or_cluster_pairs.argtypes = [C.c_void_p, _up, C.c_int32, C.c_int32, _ip, _ip, C.c_int64,
                                       C.POINTER(C.c_int64), _ip, _up, C.c_int64]
        L.or_cluster_pairs.restype = C.c_int64
        L.or_step.argtypes = [C.c_void_p, _up, _dp, C.c_int]
        L.or_step.restype = C.c_int
        L.or_step_ex.argtypes = [C.c_void_p, _up, _dp, C.c_int, _dp, _dp]
        L.or_step_ex.restype = C.c_int
        L.or_step_npt.argtypes = [C.c_void_p, C.POINTER(Piston), _up, _dp, C.c_int, _dp, _dp]
        L.or_step_npt.restype = C.c_int
        L.or_virial_trace.argtypes = [C.c_void_p, _up]
        L.or_virial_trace.restype = C.c_double
        L.or_set_box.argtypes = [C.c_void_p, _dp]
        L.or_get_box.argtypes = [C.c_void_p, _dp]
        L.or_shake_positions.argtypes = [C.c_void_p, _up, _up]
        L.or_shake_positions.restype = C.c_int
        L.or_maxwell.argtypes = [C.c_void_p, _up, C.c_double, C.c_uint64, C.c_uint64, _dp]
        L.or_kinetic.argtypes = [C.c_void_p, _dp]
        L.or_kinetic.restype = C.c_double
        L.or_temperature.argtypes = [C.c_void_p, _dp]
        L.or_temperature.restype = C.c_double
        L.or_constraint_residual.argtypes = [C.c_void_p, _up, _dp, _dp]
        L.or_constraint_residual.restype = C.c_double
        L.or_fft3d.argtypes = [_ip, _dp, _dp, C.c_int]
        L.or_set_fixed.argtypes = [C.c_void_p, C.POINTER(C.c_uint8)]
        L.or_set_fixed.restype = C.c_int
        L.or_set_com_restraint.argtypes = [C.c_void_p, _ip, _dp, C.c_double]
        L.or_set_com_restraint.restype = C.c_int
        L.or_philox.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        _lib = L
    return _lib


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


# atm -> kcal/mol/A^3 (PU, src/defs.in.h:105)
PU = 1.4586e-5


class Piston(C.Structure):
    """Andersen piston (or_piston): reference box L0, p_ref [kcal/mol/A^3],
    mass [amu/A^4], volume V, momentum Pi = M dV/dt, virial trace Wtr."""
    _fields_ = [("L0", C.c_double * 3), ("p_ref", C.c_double), ("mass", C.c_double), ("V", C.c_double),
                ("Pi", C.c_double), ("Wtr", C.c_double)]


class Oracle:
    """FP64 oracle for one system.  topo: dict of numpy arrays (fz_synth layout)."""

    def __init__(self, topo, params):
        self._keep = {}

        def arr(name, dt):
            a = np.ascontiguousarray(topo[name], dtype=dt)
            if a.size == 0:
                a = np.zeros(1, dtype=dt)
            self._keep[name] = a
            return _p(a, C.c_int32 if dt == np.int32 else C.c_double)

        t = Topo()
        t.n_atoms, t.n_types = int(topo["n_atoms"]), int(topo["n_types"])
        for f in ("charge", "mass", "lj_eps", "lj_sig2", "lj14_eps", "lj14_sig2", "bond_k", "bond_r0",
                  "angle_k", "angle_theta0", "ub_k", "ub_s0", "dihedral_k", "dihedral_delta",
                  "improper_k", "improper_psi0", "constraint_length"):
            setattr(t, f, arr(f, np.float64))
        for f in ("type", "bond_atoms", "angle_atoms", "dihedral_atoms", "dihedral_n", "improper_atoms",
                  "constraint_atoms"):
            setattr(t, f, arr(f, np.int32))
        for k in ("bonds", "angles", "dihedrals", "impropers", "constraints"):
            setattr(t, "n_" + k, int(topo["n_" + k]))
        p = Params()
        p.box[:] = [float(x) for x in params["box"]]
        p.cutoff, p.cuton, p.pairlist, p.beta = (float(params[k]) for k in ("cutoff", "cuton", "pairlist", "beta"))
        p.K[:] = [int(k) for k in params["K"]]
        p.dt_ps = float(params.get("dt_ps", 0.001))
        p.shake_tol = float(params.get("shake_tol", 1e-8))
        p.ewald_kmax_tol = float(params.get("ewald_kmax_tol", 1e-18))
        self.t, self.p = t, p
        self.n = t.n_atoms
        self.box = np.array(params["box"], dtype=np.float64)
        self.fixed = np.zeros(self.n, bool)
        self.h = C.c_void_p(lib().or_create(C.byref(t), C.byref(p)))

    def set_fixed(self, fixed):
        """consfix mask (None releases every atom)."""
        if fixed is None:
            self.fixed = np.zeros(self.n, bool)
            assert lib().or_set_fixed(self.h, None) == 0
            return
        m = np.ascontiguousarray(np.asarray(fixed).astype(np.uint8))
        self._keep["fixed"] = m
        assert lib().or_set_fixed(self.h, _p(m, C.c_uint8)) == 0
        self.fixed = m.astype(bool)

    def set_com_restraint(self, group, dist=(10.0, 10.0, 20.0), fc=0.01):
        """consharm: group[i] in {0, 1, -1}; group-1 atoms receive fc * (dist + c0 - c1)."""
        if group is None:
            assert lib().or_set_com_restraint(self.h, None, None, 0.0) == 0
            return
        g = np.ascontiguousarray(np.asarray(group, dtype=np.int32))
        dd = np.ascontiguousarray(np.asarray(dist, dtype=np.float64))
        self._keep["com_grp"], self._keep["com_d"] = g, dd
        assert lib().or_set_com_restraint(self.h, _p(g, C.c_int32), _p(dd, C.c_double), float(fc)) == 0

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.or_destroy(self.h)
            self.h = None

    def to_fixed(self, xyz):
        x = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        u = np.empty(x.shape, dtype=np.uint64)
        lib().or_to_fixed(_p(self.box, C.c_double), _p(x, C.c_double), _p(u, C.c_uint64), len(x))
        return u

    def from_fixed(self, u):
        u = np.ascontiguousarray(u, dtype=np.uint64).reshape(-1, 3)
        x = np.empty(u.shape)
        lib().or_from_fixed(_p(self.box, C.c_double), _p(u, C.c_uint64), _p(x, C.c_double), len(u))
        return x

    def forces(self, u, flags=ALL):
        u = np.ascontiguousarray(u, dtype=np.uint64)
        f = np.zeros((self.n, 3))
        e = np.zeros(10)
        v = np.zeros(9)
        lib().or_forces(self.h, _p(u, C.c_uint64), _p(f, C.c_double), _p(e, C.c_double), _p(v, C.c_double), flags)
        return f, dict(zip(E_NAMES, e)), v.reshape(3, 3)

    def pairs(self, which):
        n = lib().or_n_pairs(self.h, which)
        ij = np.zeros((max(n, 1), 2), np.int32)
        lib().or_get_pairs(self.h, which, _p(ij, C.c_int32))
        return ij[:n]

    def count_pairs_within(self, u, r):
        u = np.ascontiguousarray(u, dtype=np.uint64)
        return lib().or_count_pairs_within(self.h, _p(u, C.c_uint64), float(r))

    def cluster_pairs(self, u, nx, ny, cap=None):
        u = np.ascontiguousarray(u, dtype=np.uint64)
        col = np.zeros(self.n, np.int32)
        scap = self.n + 8 * nx * ny + 8
        slots = np.zeros(scap, np.int32)
        ns = C.c_int64(0)
        cap = cap or 4096
        while True:
            pairs = np.zeros((cap, 2), np.int32)
            masks = np.zeros(cap, np.uint64)
            r = lib().or_cluster_pairs(self.h, _p(u, C.c_uint64), nx, ny, _p(col, C.c_int32), _p(slots, C.c_int32),
                                       scap, C.byref(ns), _p(pairs, C.c_int32), _p(masks, C.c_uint64), cap)
            if r >= 0:
                return col, slots[: ns.value], pairs[:r], masks[:r]
            cap = -r

    def step(self, u, v, nsteps):
        u = np.ascontiguousarray(u, dtype=np.uint64).copy()
        v = np.ascontiguousarray(v, dtype=np.float64).copy()
        rc = lib().or_step(self.h, _p(u, C.c_uint64), _p(v, C.c_double), int(nsteps))
        return u, v, rc

    def step_ex(self, u, v, nsteps):
        """step() that also returns the last step's energies and virial (forces
        at the final positions plus the final RATTLE's constraint virial)."""
        u = np.ascontiguousarray(u, dtype=np.uint64).copy()
        v = np.ascontiguousarray(v, dtype=np.float64).copy()
        e = np.zeros(10)
        w = np.zeros(9)
        rc = lib().or_step_ex(self.h, _p(u, C.c_uint64), _p(v, C.c_double), int(nsteps), _p(e, C.c_double),
                              _p(w, C.c_double))
        return u, v, dict(zip(E_NAMES, e)), w.reshape(3, 3), rc

    def set_box(self, box):
        self.box = np.array(box, dtype=np.float64)
        lib().or_set_box(self.h, _p(self.box, C.c_double))

    def piston(self, u, pref_atm, mass):
        """A piston at rest at the current box, Wtr the virial trace at u."""
        pc = Piston()
        pc.L0[:] = [float(x) for x in self.box]
        pc.p_ref, pc.mass = float(pref_atm) * PU, float(mass)
        pc.V = float(np.prod(self.box))
        pc.Pi = 0.0
        u = np.ascontiguousarray(u, dtype=np.uint64)
        pc.Wtr = lib().or_virial_trace(self.h, _p(u, C.c_uint64))
        return pc

    def step_npt(self, u, v, nsteps, pc):
        """nsteps Andersen-barostat steps; pc (Piston) is advanced in place and
        the oracle's box follows it.  Returns u, v, energies, virial, rc."""
        u = np.ascontiguousarray(u, dtype=np.uint64).copy()
        v = np.ascontiguousarray(v, dtype=np.float64).copy()
        e = np.zeros(10)
        w = np.zeros(9)
        rc = lib().or_step_npt(self.h, C.byref(pc), _p(u, C.c_uint64), _p(v, C.c_double), int(nsteps),
                               _p(e, C.c_double), _p(w, C.c_double))
        self.box = np.zeros(3)
        lib().or_get_box(self.h, _p(self.box, C.c_double))
        return u, v, dict(zip(E_NAMES, e)), w.reshape(3, 3), rc

    def shake_positions(self, uold, u):
        u = np.ascontiguousarray(u, dtype=np.uint64).copy()
        uold = np.ascontiguousarray(uold, dtype=np.uint64)
        rc = lib().or_shake_positions(self.h, _p(uold, C.c_uint64), _p(u, C.c_uint64))
        return u, rc

    def minimize_ref(self, u, nsteps, method=1, cgfac=1e-10, sdfac=1e-4, cgrate=0.1, sdrate=0.01):
        """The reference's minimisation schedule (MinimizeStep, src/dynamics.cu:491-620;
        ConjugatedGradient / SteepestDescent, src/mdcuda.cu:126-330; factor ramp,
        src/fenzi.cu:402-415) in FP64, with this build's SHAKE after every move:
        method 1 moves r += H, H = (F + gamma H_prev) cgfac, gamma = |F|^2/|F_prev|^2
        per atom (F_prev taken as F when |F_prev|^2 < 1e-8); method 2 moves
        r += F sdfac.  An energy above the last accepted one (the first move is
        always accepted) undoes the move and demotes 1 -> 2 -> 0 (stop).  After
        the accepted moves with 0-based index 100, 200, ... the active factor is
        multiplied by rate * 100.  Returns (u, method, accepted moves)."""
        sc = np.array(self.box) / 2.0 ** 64
        u = np.ascontiguousarray(u, dtype=np.uint64).copy()
        H = np.zeros((self.n, 3))
        Fp = np.zeros((self.n, 3))
        f, _, _ = self.forces(u)
        e_prev, have, done = 0.0, False, 0
        for _ in range(nsteps):
            if method == 0:
                break
            uold = u.copy()
            if method == 1:
                f2 = (f * f).sum(1)
                p2 = (Fp * Fp).sum(1)
                p2 = np.where(p2 < 1e-8, f2, p2)
                g = np.where(p2 > 0, f2 / np.where(p2 > 0, p2, 1.0), 0.0)
                Hn = (f + g[:, None] * H) * cgfac
                H = np.where(self.fixed[:, None], H, Hn)  # fixed atoms keep no direction
                Fp = np.where(self.fixed[:, None], Fp, f)
                h = np.where(self.fixed[:, None], 0.0, H)
            else:
                h = np.where(self.fixed[:, None], 0.0, f * sdfac)
            u = u + np.rint(h / sc).astype(np.int64).view(np.uint64)
            u, _ = self.shake_positions(uold, u)
            fn, e, _ = self.forces(u)
            pot = sum(e.values())
            if have and pot > e_prev:
                u = uold
                method = 2 if method == 1 else 0
                f, _, _ = self.forces(u)
                continue
            f = fn
            e_prev, have = pot, True
            idx = done
            done += 1
            if idx > 0 and idx % 100 == 0:
                if method == 1:
                    cgfac *= cgrate * 100.0
                else:
                    sdfac *= sdrate * 100.0
        return u, method, done

    def minimize_sd(self, u, nsteps, max_step, step=0.0):
        """The library's adaptive steepest descent (fz_minimize; the reference's
        MinimizeStep SD branch, src/dynamics.cu:491-620 and
        src/mdcuda.cu:245-330, with its own step rule): the mobile atom with the
        largest force moves `step` A along it and every mobile atom moves by
        the same multiple of its force, then SHAKE against the previous
        positions; an energy below the last accepted one is accepted and the
        step grows by 1.2 (capped at max_step), otherwise the move is undone and
        the step halves.  step <= 0 starts at max_step.  Returns (u, step)."""
        sc = np.array(self.box) / 2.0 ** 64
        u = np.ascontiguousarray(u, dtype=np.uint64).copy()
        step = max_step if step <= 0.0 else min(step, max_step)
        f, e, _ = self.forces(u)
        e_old = sum(e.values())
        mobile = ~self.fixed
        for _ in range(nsteps):
            f2 = (f * f).sum(1)
            f2max = f2[mobile].max() if mobile.any() else 0.0
            a = step / np.sqrt(f2max) if f2max > 0 else 0.0
            uold = u.copy()
            h = np.where(mobile[:, None], a * f, 0.0)
            u = u + np.rint(h / sc).astype(np.int64).view(np.uint64)
            u, _ = self.shake_positions(uold, u)
            fn, e, _ = self.forces(u)
            pot = sum(e.values())
            if pot < e_old:
                e_old = pot
                step = min(step * 1.2, max_step)
                f = fn
            else:
                u = uold
                step *= 0.5
        return u, step

    def maxwell(self, u, T, seed, counter=0):
        u = np.ascontiguousarray(u, dtype=np.uint64)
        v = np.zeros((self.n, 3))
        lib().or_maxwell(self.h, _p(u, C.c_uint64), float(T), int(seed), int(counter), _p(v, C.c_double))
        return v

    def kinetic(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        return lib().or_kinetic(self.h, _p(v, C.c_double))

    def temperature(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        return lib().or_temperature(self.h, _p(v, C.c_double))

    def constraint_residual(self, u, v=None):
        u = np.ascontiguousarray(u, dtype=np.uint64)
        vr = C.c_double(0)
        vp = _p(np.ascontiguousarray(v, dtype=np.float64), C.c_double) if v is not None else None
        r = lib().or_constraint_residual(self.h, _p(u, C.c_uint64), C.byref(vr), vp)
        return r, vr.value


def fft3d(data, sign=-1):
    """Unnormalised complex 3-D DFT of data[K2, K1, K0] (complex128)."""
    a = np.ascontiguousarray(data, dtype=np.complex128)
    K = np.array([a.shape[2], a.shape[1], a.shape[0]], np.int32)
    out = np.empty_like(a)
    lib().or_fft3d(_p(K, C.c_int32), a.ctypes.data_as(_dp), out.ctypes.data_as(_dp), int(sign))
    return out


def set_threads(n):
    return lib().or_set_threads(int(n))
