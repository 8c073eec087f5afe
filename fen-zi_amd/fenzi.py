"""ctypes binding of libfenzi (include/fenzi.h) for tests, bench and scripts.

The library is the product; this module only marshals numpy arrays across the
C ABI.  It raises if libfenzi.so is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfenzi.so")

FP64, MIXED = 0, 1
E_FIELDS = ["bond", "angle", "urey_bradley", "improper", "dihedral", "vdw", "elec", "pme",
            "ewald_excl", "ewald_self", "potential", "kinetic"]

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)
_up = C.POINTER(C.c_uint64)


class Topology(C.Structure):
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
                ("pairlist", C.c_double), ("beta", C.c_double), ("pme_grid", C.c_int32 * 3),
                ("pme_order", C.c_int32), ("dt_ps", C.c_double), ("shake_tol", C.c_double),
                ("column_grid", C.c_int32 * 2)]


class Energies(C.Structure):
    _fields_ = [("step", C.c_int64), ("temperature", C.c_double)] + \
               [(f, C.c_double) for f in E_FIELDS] + \
               [("virial", C.c_double * 9), ("pairs_within_cutoff", C.c_int64), ("list_builds", C.c_int64),
                ("volume", C.c_double), ("pressure", C.c_double), ("pair_tiles", C.c_int64)]

    def as_dict(self):
        d = {f: getattr(self, f) for f in E_FIELDS}
        d["temperature"] = self.temperature
        d["step"] = self.step
        d["virial"] = np.array(self.virial[:]).reshape(3, 3)
        d["pairs_within_cutoff"] = self.pairs_within_cutoff
        d["list_builds"] = self.list_builds
        d["volume"] = self.volume
        d["pressure"] = self.pressure
        d["pair_tiles"] = self.pair_tiles
        return d


class RunSettings(C.Structure):
    _fields_ = [("temperature", C.c_double), ("dynasteps", C.c_int64), ("minimizesteps", C.c_int64),
                ("startstep", C.c_int64), ("tempreset_every", C.c_int64), ("output_every", C.c_int64),
                ("seed", C.c_uint64), ("shake", C.c_int32), ("print_to_stdout", C.c_int32),
                ("output_prefix", C.c_char * 256), ("cgfac", C.c_double), ("sdfac", C.c_double),
                ("cgrate", C.c_double), ("sdrate", C.c_double), ("checkpoint_every", C.c_int64),
                ("traj_every", C.c_int64), ("print_level", C.c_int32), ("checkpoint_binary", C.c_int32),
                ("file_prefix", C.c_char * 256), ("pmass", C.c_double), ("pref", C.c_double)]


# every symbol include/fenzi.h declares
EXPORTS = ["fz_last_error", "fz_ctx_create", "fz_ctx_destroy", "fz_system_create", "fz_system_destroy",
           "fz_set_state", "fz_get_state", "fz_set_fixed_positions", "fz_get_fixed_positions",
           "fz_reassign_velocities", "fz_compute_forces", "fz_step", "fz_get_energies", "fz_minimize",
           "fz_system_info", "fz_phase_times", "fz_autotune", "fz_batch_autotune", "fz_get_cluster_layout", "fz_get_cluster_pairs", "fz_pme_convolve",
           "fz_batch_create", "fz_batch_step", "fz_batch_minimize", "fz_batch_destroy", "fz_deck_load", "fz_deck_destroy",
           "fz_deck_topology", "fz_deck_params", "fz_deck_coordinates", "fz_deck_settings",
           "fz_deck_psf_counts", "fz_format_energy_header", "fz_format_energy_row",
           "fz_server_create", "fz_server_submit", "fz_server_poll", "fz_server_job_info", "fz_server_job_result",
           "fz_server_release", "fz_server_estimate_bytes", "fz_server_destroy",
           "fz_minimize_ex", "fz_batch_minimize_ex", "fz_restart_write", "fz_restart_read", "fz_dcd_open",
           "fz_dcd_write_frame", "fz_dcd_close", "fz_set_fixed_atoms", "fz_set_com_restraint",
           "fz_deck_com_restraint", "fz_deck_fixed_atoms", "fz_set_neighbor_capacity",
           "fz_pme_convolve_noalloc", "fz_pme_convolve_noalloc_bytes", "fz_set_barostat", "fz_get_barostat",
           "fz_set_launch", "fz_batch_set_launch", "fz_pair_kernel_time", "fz_batch_phase_times"]

MIN_ADAPTIVE_SD, MIN_REF_CG, MIN_REF_SD = 0, 1, 2
PME_AUTO, PME_GRID, PME_PLANES = 0, 1, 2


class MinParams(C.Structure):
    _fields_ = [("method", C.c_int32), ("max_step", C.c_double), ("cgfac", C.c_double), ("sdfac", C.c_double),
                ("cgrate", C.c_double), ("sdrate", C.c_double)]


def min_params(method=MIN_REF_CG, max_step=0.05, cgfac=1e-10, sdfac=1e-4, cgrate=0.1, sdrate=0.01):
    """fz_min_params; defaults are the reference's (src/globals.cu:487-497)."""
    return MinParams(int(method), float(max_step), float(cgfac), float(sdfac), float(cgrate), float(sdrate))

JOB_QUEUED, JOB_MINIMIZING, JOB_DYNAMICS, JOB_DONE, JOB_FAILED = 0, 1, 2, 3, -1


class ServerConfig(C.Structure):
    _fields_ = [("hbm_budget_bytes", C.c_double), ("max_active", C.c_int32), ("precision", C.c_int32),
                ("chunk_steps", C.c_int32), ("min_max_step", C.c_double)]


class JobInfo(C.Structure):
    _fields_ = [("id", C.c_int64), ("state", C.c_int32), ("n_atoms", C.c_int64), ("steps_done", C.c_int64),
                ("min_done", C.c_int64), ("device_bytes", C.c_int64), ("t_submit", C.c_double),
                ("t_admit", C.c_double), ("t_done", C.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}

_lib = None


class FenziError(RuntimeError):
    """A negative status of the C ABI; .code is the FZ_ERR_* value."""
    code = 0


# status codes of include/fenzi.h
ERR_ARG, ERR_HIP, ERR_CAPACITY, ERR_PARSE, ERR_NOMEM, ERR_SHAKE, ERR_STATE, ERR_BUDGET = -1, -2, -3, -4, -5, -6, -7, -8


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FenziError("libfenzi.so is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        vp = C.c_void_p
        L.fz_last_error.restype = C.c_char_p
        L.fz_ctx_create.argtypes = [C.c_int, C.POINTER(vp)]
        L.fz_ctx_destroy.argtypes = [vp]
        L.fz_system_create.argtypes = [vp, C.POINTER(Topology), C.POINTER(Params), C.c_int, C.POINTER(vp)]
        L.fz_system_destroy.argtypes = [vp]
        L.fz_set_state.argtypes = [vp, _dp, _dp]
        L.fz_get_state.argtypes = [vp, _dp, _dp]
        L.fz_set_fixed_positions.argtypes = [vp, _up]
        L.fz_get_fixed_positions.argtypes = [vp, _up]
        L.fz_reassign_velocities.argtypes = [vp, C.c_double, C.c_uint64, C.c_uint64]
        L.fz_compute_forces.argtypes = [vp, _dp, C.POINTER(Energies)]
        L.fz_step.argtypes = [vp, C.c_int]
        L.fz_set_barostat.argtypes = [vp, C.c_double, C.c_double]
        L.fz_get_barostat.argtypes = [vp, _dp, _dp]
        L.fz_get_energies.argtypes = [vp, C.POINTER(Energies)]
        L.fz_minimize.argtypes = [vp, C.c_int, C.c_double]
        L.fz_system_info.argtypes = [vp, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.POINTER(C.c_int32)]
        L.fz_set_neighbor_capacity.argtypes = [vp, C.c_int64, C.c_int32]
        L.fz_phase_times.argtypes = [vp, C.c_int, _dp]
        L.fz_autotune.argtypes = [vp, C.c_int, _dp]
        L.fz_batch_autotune.argtypes = [vp, C.c_int, _dp]
        L.fz_batch_phase_times.argtypes = [vp, C.c_int, _dp]
        L.fz_set_launch.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int32)]
        L.fz_pair_kernel_time.argtypes = [vp, C.c_int, _dp]
        L.fz_batch_set_launch.argtypes = [vp, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_int32)]
        L.fz_get_cluster_layout.argtypes = [vp, _ip, _ip, _ip, C.c_int64, C.POINTER(C.c_int64)]
        L.fz_get_cluster_pairs.argtypes = [vp, _ip, _up, C.c_int64]
        L.fz_get_cluster_pairs.restype = C.c_int64
        L.fz_pme_convolve.argtypes = [vp, _ip, _dp, C.c_double, C.c_int, C.c_int, _dp, _dp, _dp]
        L.fz_pme_convolve_noalloc.argtypes = [vp, _ip, _dp, C.c_double, C.c_int, C.c_int, vp, vp, vp, vp]
        L.fz_pme_convolve_noalloc_bytes.restype = C.c_int64
        L.fz_batch_create.argtypes = [vp, C.POINTER(vp), C.c_int, C.POINTER(vp)]
        L.fz_batch_step.argtypes = [vp, C.c_int]
        L.fz_batch_minimize.argtypes = [vp, C.c_int, C.c_double]
        L.fz_minimize_ex.argtypes = [vp, C.c_int, C.POINTER(MinParams)]
        L.fz_restart_write.argtypes = [C.c_char_p, C.c_int32, C.c_int64, C.c_int64, C.c_double, _dp, _dp, _dp]
        L.fz_restart_read.argtypes = [C.c_char_p, C.c_int32, C.c_int64, C.POINTER(C.c_int64),
                                      C.POINTER(C.c_double), _dp, _dp, _dp]
        L.fz_dcd_open.argtypes = [C.c_char_p, C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_int64, C.c_double,
                                  C.POINTER(vp)]
        L.fz_dcd_write_frame.argtypes = [vp, _dp, _dp]
        L.fz_dcd_close.argtypes = [vp]
        L.fz_batch_minimize_ex.argtypes = [vp, C.c_int, C.POINTER(MinParams)]
        L.fz_batch_destroy.argtypes = [vp]
        L.fz_deck_load.argtypes = [C.c_char_p, C.POINTER(vp)]
        L.fz_deck_destroy.argtypes = [vp]
        L.fz_deck_topology.argtypes = [vp]
        L.fz_deck_topology.restype = C.POINTER(Topology)
        L.fz_deck_params.argtypes = [vp]
        L.fz_deck_params.restype = C.POINTER(Params)
        L.fz_deck_coordinates.argtypes = [vp]
        L.fz_deck_coordinates.restype = _dp
        L.fz_deck_settings.argtypes = [vp, C.POINTER(RunSettings)]
        L.fz_deck_psf_counts.argtypes = [vp, C.POINTER(C.c_int64)]
        L.fz_deck_com_restraint.argtypes = [vp, C.POINTER(C.c_int32), _dp, _dp]
        L.fz_deck_fixed_atoms.argtypes = [vp, C.POINTER(C.c_uint8)]
        L.fz_set_fixed_atoms.argtypes = [vp, C.POINTER(C.c_uint8)]
        L.fz_set_com_restraint.argtypes = [vp, C.POINTER(C.c_int32), _dp, C.c_double]
        L.fz_format_energy_header.argtypes = [C.c_char_p, C.c_size_t]
        L.fz_format_energy_row.argtypes = [C.POINTER(Energies), C.c_double, C.c_char_p, C.c_size_t]
        L.fz_server_create.argtypes = [vp, C.POINTER(ServerConfig), C.POINTER(vp)]
        L.fz_server_submit.argtypes = [vp, C.POINTER(Topology), C.POINTER(Params), _dp, C.c_int32, C.c_int32,
                                       C.c_double, C.c_int32, C.c_uint64, C.POINTER(C.c_int64)]
        L.fz_server_poll.argtypes = [vp, C.POINTER(C.c_int32), C.POINTER(C.c_int32), C.POINTER(C.c_int32)]
        L.fz_server_job_info.argtypes = [vp, C.c_int64, C.POINTER(JobInfo)]
        L.fz_server_job_result.argtypes = [vp, C.c_int64, _dp, _dp, _up, C.POINTER(Energies)]
        L.fz_server_release.argtypes = [vp, C.c_int64]
        L.fz_server_estimate_bytes.argtypes = [C.POINTER(Topology), C.POINTER(Params), C.c_int32]
        L.fz_server_estimate_bytes.restype = C.c_int64
        L.fz_server_destroy.argtypes = [vp]
        _lib = L
    return _lib


def _check(rc, what):
    if rc < 0:
        err = FenziError("%s: %s (code %d)" % (what, lib().fz_last_error().decode(), rc))
        err.code = int(rc)
        raise err
    return rc


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def make_topology(topo, keep):
    """fz_topology over numpy arrays of a topology dict (fz_synth layout)."""
    def arr(name, dt):
        a = np.ascontiguousarray(topo[name], dtype=dt)
        if a.size == 0:
            a = np.zeros(1, dtype=dt)
        keep.append(a)
        return _p(a, C.c_int32 if dt == np.int32 else C.c_double)

    t = Topology()
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
    return t


def make_params(params):
    p = Params()
    p.box[:] = [float(x) for x in params["box"]]
    p.cutoff, p.cuton, p.pairlist, p.beta = (float(params[k]) for k in ("cutoff", "cuton", "pairlist", "beta"))
    p.pme_grid[:] = [int(k) for k in params["K"]]
    p.pme_order = 4
    p.dt_ps = float(params.get("dt_ps", 0.001))
    p.shake_tol = float(params.get("shake_tol", 1e-8))
    cg = params.get("column_grid", (0, 0))
    p.column_grid[:] = [int(cg[0]), int(cg[1])]
    return p


class Context:
    def __init__(self, device=0):
        self.h = C.c_void_p()
        _check(lib().fz_ctx_create(device, C.byref(self.h)), "fz_ctx_create")

    def close(self):
        if self.h:
            lib().fz_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def pme_convolve(self, q, box, beta, precision=FP64, mode=0):
        q = np.ascontiguousarray(q, dtype=np.float64)
        K = np.array([q.shape[2], q.shape[1], q.shape[0]], np.int32)
        phi = np.empty_like(q)
        e = C.c_double(0)
        b = np.asarray(box, np.float64)
        _check(lib().fz_pme_convolve(self.h, _p(K, C.c_int32), _p(b, C.c_double), float(beta), precision, mode,
                                     _p(q, C.c_double), _p(phi, C.c_double), C.byref(e)), "fz_pme_convolve")
        return phi, e.value

    def pme_convolve_noalloc(self, K, box, beta, q_dev, phi_dev, energy_dev=None, precision=FP64, mode=0,
                             stream=None):
        """fz_pme_convolve_noalloc on device addresses (ints): q_dev, phi_dev
        hold K0*K1*K2 doubles, energy_dev one double (or None); enqueued on
        stream (None = the null stream), not synchronised."""
        Kc = np.asarray(K, np.int32)
        b = np.asarray(box, np.float64)
        _check(lib().fz_pme_convolve_noalloc(self.h, _p(Kc, C.c_int32), _p(b, C.c_double), float(beta), precision,
                                             mode, C.c_void_p(q_dev), C.c_void_p(phi_dev),
                                             C.c_void_p(energy_dev) if energy_dev else None,
                                             C.c_void_p(stream) if stream else None), "fz_pme_convolve_noalloc")


def noalloc_scratch_bytes():
    """Bytes of the static device array behind fz_pme_convolve_noalloc (per GPU)."""
    return int(lib().fz_pme_convolve_noalloc_bytes())


class System:
    def __init__(self, ctx, topo, params, precision=FP64, xyz=None, vel=None):
        self.ctx = ctx
        self._keep = []
        self.n = int(topo["n_atoms"])
        t = make_topology(topo, self._keep)
        p = make_params(params)
        self.h = C.c_void_p()
        _check(lib().fz_system_create(ctx.h, C.byref(t), C.byref(p), precision, C.byref(self.h)), "fz_system_create")
        self.precision = precision
        self.dt_ps = p.dt_ps
        if xyz is not None:
            self.set_state(xyz, vel)

    def close(self):
        if self.h:
            lib().fz_system_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, xyz, vel=None):
        x = np.ascontiguousarray(xyz, dtype=np.float64).reshape(self.n, 3)
        v = None if vel is None else np.ascontiguousarray(vel, dtype=np.float64).reshape(self.n, 3)
        _check(lib().fz_set_state(self.h, _p(x, C.c_double), None if v is None else _p(v, C.c_double)),
               "fz_set_state")

    def get_state(self):
        x = np.empty((self.n, 3))
        v = np.empty((self.n, 3))
        _check(lib().fz_get_state(self.h, _p(x, C.c_double), _p(v, C.c_double)), "fz_get_state")
        return x, v

    def set_fixed(self, u):
        u = np.ascontiguousarray(u, dtype=np.uint64).reshape(self.n, 3)
        _check(lib().fz_set_fixed_positions(self.h, _p(u, C.c_uint64)), "fz_set_fixed_positions")

    def get_fixed(self):
        u = np.empty((self.n, 3), np.uint64)
        _check(lib().fz_get_fixed_positions(self.h, _p(u, C.c_uint64)), "fz_get_fixed_positions")
        return u

    def reassign_velocities(self, T, seed, counter=0):
        _check(lib().fz_reassign_velocities(self.h, float(T), int(seed), int(counter)), "fz_reassign_velocities")

    def set_fixed_atoms(self, fixed):
        """consfix: fixed[i] true holds atom i in place (None releases all)."""
        if fixed is None:
            _check(lib().fz_set_fixed_atoms(self.h, None), "fz_set_fixed_atoms")
            return
        m = np.ascontiguousarray(np.asarray(fixed).astype(np.uint8))
        assert m.shape == (self.n,)
        _check(lib().fz_set_fixed_atoms(self.h, _p(m, C.c_uint8)), "fz_set_fixed_atoms")

    def set_com_restraint(self, group, dist=(10.0, 10.0, 20.0), fc=0.01):
        """consharm: group[i] in {0, 1, -1}; group-1 atoms receive fc * (dist + c0 - c1)."""
        if group is None:
            _check(lib().fz_set_com_restraint(self.h, None, None, 0.0), "fz_set_com_restraint")
            return
        g = np.ascontiguousarray(np.asarray(group, dtype=np.int32))
        assert g.shape == (self.n,)
        dd = np.ascontiguousarray(np.asarray(dist, dtype=np.float64))
        _check(lib().fz_set_com_restraint(self.h, _p(g, C.c_int32), _p(dd, C.c_double), float(fc)),
               "fz_set_com_restraint")

    def forces(self):
        f = np.empty((self.n, 3))
        e = Energies()
        _check(lib().fz_compute_forces(self.h, _p(f, C.c_double), C.byref(e)), "fz_compute_forces")
        return f, e.as_dict()

    def step(self, n):
        _check(lib().fz_step(self.h, int(n)), "fz_step")

    def minimize(self, n, max_step=0.05):
        _check(lib().fz_minimize(self.h, int(n), float(max_step)), "fz_minimize")

    def minimize_ex(self, n, params):
        _check(lib().fz_minimize_ex(self.h, int(n), C.byref(params)), "fz_minimize_ex")

    def energies(self):
        e = Energies()
        _check(lib().fz_get_energies(self.h, C.byref(e)), "fz_get_energies")
        return e.as_dict()

    PHASES = ["kick_drift_shake", "nblist", "gather_sorted", "pair", "bonded", "pme_spread", "pme_fft_conv",
              "pme_gather", "kick_rattle"]

    def phase_times(self, nsteps):
        ms = np.zeros(len(self.PHASES))
        _check(lib().fz_phase_times(self.h, int(nsteps), _p(ms, C.c_double)), "fz_phase_times")
        return dict(zip(self.PHASES, ms.tolist()))

    def pair_kernel_time(self, reps=20):
        """fz_pair_kernel_time: mean ms per force-only pair-kernel launch."""
        ms = C.c_double()
        _check(lib().fz_pair_kernel_time(self.h, int(reps), C.byref(ms)), "fz_pair_kernel_time")
        return ms.value

    def autotune(self, nsteps=20):
        out = np.zeros(4)
        _check(lib().fz_autotune(self.h, int(nsteps), _p(out, C.c_double)), "fz_autotune")
        return dict(column_width=out[0], fft_lines=int(out[1]), pair_grid=int(out[2]), ms_per_step=out[3])

    def set_launch(self, fft_lines=-1, pair_grid=-1, pme_path=-1):
        """fz_set_launch; returns the shapes in effect (fft_lines, pair_grid, pme_path)."""
        out = np.zeros(3, np.int32)
        _check(lib().fz_set_launch(self.h, int(fft_lines), int(pair_grid), int(pme_path), _p(out, C.c_int32)),
               "fz_set_launch")
        return dict(fft_lines=int(out[0]), pair_grid=int(out[1]), pme_path=int(out[2]))

    def set_neighbor_capacity(self, pairs_cap=0, clusters_cap=0):
        """Neighbour capacities (0 keeps one); the list held stays the same."""
        _check(lib().fz_set_neighbor_capacity(self.h, int(pairs_cap), int(clusters_cap)), "fz_set_neighbor_capacity")

    def set_barostat(self, pref_atm, piston_mass):
        """Andersen barostat (fz_set_barostat): pref in atm, piston mass in
        amu/A^4; piston_mass <= 0 turns it off."""
        _check(lib().fz_set_barostat(self.h, float(pref_atm), float(piston_mass)), "fz_set_barostat")

    def barostat(self):
        """Current box and piston state: dict(box, volume, Pi, Wtr)."""
        box, st = np.zeros(3), np.zeros(3)
        _check(lib().fz_get_barostat(self.h, _p(box, C.c_double), _p(st, C.c_double)), "fz_get_barostat")
        return dict(box=box, volume=st[0], Pi=st[1], Wtr=st[2])

    def info(self):
        n, b, p = C.c_int64(), C.c_int64(), C.c_int32()
        _check(lib().fz_system_info(self.h, C.byref(n), C.byref(b), C.byref(p)), "fz_system_info")
        return dict(n_atoms=n.value, device_bytes=b.value, precision=p.value)

    def cluster_layout(self):
        g = np.zeros(2, np.int32)
        col = np.zeros(self.n, np.int32)
        ns = C.c_int64()
        _check(lib().fz_get_cluster_layout(self.h, _p(g, C.c_int32), _p(col, C.c_int32), None, 0, C.byref(ns)),
               "fz_get_cluster_layout")
        slots = np.zeros(max(1, ns.value), np.int32)
        _check(lib().fz_get_cluster_layout(self.h, None, None, _p(slots, C.c_int32), ns.value, C.byref(ns)),
               "fz_get_cluster_layout")
        return (int(g[0]), int(g[1])), col, slots[: ns.value]

    def cluster_pairs(self):
        n = _check(lib().fz_get_cluster_pairs(self.h, None, None, 0), "fz_get_cluster_pairs")
        pairs = np.zeros((max(1, n), 2), np.int32)
        masks = np.zeros(max(1, n), np.uint64)
        n2 = _check(lib().fz_get_cluster_pairs(self.h, _p(pairs, C.c_int32), _p(masks, C.c_uint64), n),
                    "fz_get_cluster_pairs")
        return pairs[:n2], masks[:n2]


class Batch:
    def __init__(self, ctx, systems):
        self.ctx = ctx
        self.systems = list(systems)
        arr = (C.c_void_p * len(self.systems))(*[s.h.value for s in self.systems])
        self.h = C.c_void_p()
        _check(lib().fz_batch_create(ctx.h, arr, len(self.systems), C.byref(self.h)), "fz_batch_create")

    def step(self, n):
        _check(lib().fz_batch_step(self.h, int(n)), "fz_batch_step")

    def minimize(self, n, max_step=0.05):
        _check(lib().fz_batch_minimize(self.h, int(n), float(max_step)), "fz_batch_minimize")

    def minimize_ex(self, n, params):
        _check(lib().fz_batch_minimize_ex(self.h, int(n), C.byref(params)), "fz_batch_minimize_ex")

    def autotune(self, nsteps=10):
        out = np.zeros(4)
        _check(lib().fz_batch_autotune(self.h, int(nsteps), _p(out, C.c_double)), "fz_batch_autotune")
        return dict(column_width=out[0], fft_lines=int(out[1]), pair_grid=int(out[2]), ms_per_step=out[3])

    def phase_times(self, nsteps):
        """fz_batch_phase_times: mean ms per phase of a batch step (eager launches)."""
        ms = np.zeros(len(System.PHASES))
        _check(lib().fz_batch_phase_times(self.h, int(nsteps), _p(ms, C.c_double)), "fz_batch_phase_times")
        return dict(zip(System.PHASES, ms.tolist()))

    def set_launch(self, fft_lines=-1, pair_grid=-1, pme_path=-1):
        """fz_batch_set_launch; returns the shapes in effect."""
        out = np.zeros(3, np.int32)
        _check(lib().fz_batch_set_launch(self.h, int(fft_lines), int(pair_grid), int(pme_path), _p(out, C.c_int32)),
               "fz_batch_set_launch")
        return dict(fft_lines=int(out[0]), pair_grid=int(out[1]), pme_path=int(out[2]))

    def close(self):
        if self.h:
            lib().fz_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def restart_write(path, step, seed, box, xyz, vel, binary=True):
    x = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
    v = np.ascontiguousarray(vel, dtype=np.float64).reshape(-1, 3)
    b = np.ascontiguousarray(box, dtype=np.float64)
    _check(lib().fz_restart_write(path.encode(), int(binary), len(x), int(step), float(seed), _p(b, C.c_double),
                                  _p(x, C.c_double), _p(v, C.c_double)), "fz_restart_write")


def restart_read(path, n_atoms, binary=True):
    x = np.empty((n_atoms, 3))
    v = np.empty((n_atoms, 3))
    b = np.empty(3)
    st = C.c_int64()
    sd = C.c_double()
    _check(lib().fz_restart_read(path.encode(), int(binary), int(n_atoms), C.byref(st), C.byref(sd),
                                 _p(b, C.c_double), _p(x, C.c_double), _p(v, C.c_double)), "fz_restart_read")
    return dict(step=st.value, seed=sd.value, box=b, xyz=x, vel=v)


class Dcd:
    """CHARMM DCD trajectory writer (fz_dcd_* in include/fenzi.h)."""

    def __init__(self, path, n_atoms, nframes, first_step, every, nsteps, dt_ps):
        self.h = C.c_void_p()
        _check(lib().fz_dcd_open(path.encode(), int(n_atoms), int(nframes), int(first_step), int(every), int(nsteps),
                                 float(dt_ps), C.byref(self.h)), "fz_dcd_open")

    def write(self, xyz, box):
        x = np.ascontiguousarray(xyz, dtype=np.float64)
        b = np.ascontiguousarray(box, dtype=np.float64)
        _check(lib().fz_dcd_write_frame(self.h, _p(x, C.c_double), _p(b, C.c_double)), "fz_dcd_write_frame")

    def close(self):
        if self.h:
            _check(lib().fz_dcd_close(self.h), "fz_dcd_close")
            self.h = C.c_void_p()


def estimate_bytes(topo, params, precision=MIXED):
    """Device bytes the server budgets for one job of this topology."""
    keep = []
    t = make_topology(topo, keep)
    p = make_params(params)
    return _check(lib().fz_server_estimate_bytes(C.byref(t), C.byref(p), precision), "fz_server_estimate_bytes")


class Server:
    """Continuous-batching job server (fz_server_* in include/fenzi.h)."""

    def __init__(self, ctx, hbm_budget_bytes, max_active=64, precision=MIXED, chunk_steps=100, min_max_step=0.05):
        self.ctx = ctx
        cfg = ServerConfig(float(hbm_budget_bytes), int(max_active), int(precision), int(chunk_steps),
                           float(min_max_step))
        self.h = C.c_void_p()
        _check(lib().fz_server_create(ctx.h, C.byref(cfg), C.byref(self.h)), "fz_server_create")
        self.n_atoms = {}

    def submit(self, topo, params, xyz, min_steps, md_steps, temperature=300.0, reassign_every=0, seed=0):
        keep = []
        t = make_topology(topo, keep)
        p = make_params(params)
        x = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        jid = C.c_int64()
        _check(lib().fz_server_submit(self.h, C.byref(t), C.byref(p), _p(x, C.c_double), int(min_steps),
                                      int(md_steps), float(temperature), int(reassign_every), int(seed),
                                      C.byref(jid)), "fz_server_submit")
        self.n_atoms[jid.value] = x.shape[0]
        return jid.value

    def poll(self):
        a, q, d = C.c_int32(), C.c_int32(), C.c_int32()
        _check(lib().fz_server_poll(self.h, C.byref(a), C.byref(q), C.byref(d)), "fz_server_poll")
        return a.value, q.value, d.value

    def info(self, jid):
        i = JobInfo()
        _check(lib().fz_server_job_info(self.h, int(jid), C.byref(i)), "fz_server_job_info")
        return i.as_dict()

    def result(self, jid):
        n = self.n_atoms[jid]
        x = np.empty((n, 3))
        v = np.empty((n, 3))
        u = np.empty((n, 3), np.uint64)
        e = Energies()
        _check(lib().fz_server_job_result(self.h, int(jid), _p(x, C.c_double), _p(v, C.c_double),
                                          _p(u, C.c_uint64), C.byref(e)), "fz_server_job_result")
        return dict(xyz=x, vel=v, fixed=u, energies=e.as_dict())

    def release(self, jid):
        _check(lib().fz_server_release(self.h, int(jid)), "fz_server_release")
        self.n_atoms.pop(jid, None)

    def run_until_done(self, max_rounds=1000000):
        for _ in range(max_rounds):
            a, q, _ = self.poll()
            if a == 0 and q == 0:
                return
        raise FenziError("server did not drain")

    def close(self):
        if self.h:
            lib().fz_server_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Deck:
    def __init__(self, path):
        self.h = C.c_void_p()
        _check(lib().fz_deck_load(path.encode(), C.byref(self.h)), "fz_deck_load")

    def close(self):
        if self.h:
            lib().fz_deck_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def topology(self):
        t = lib().fz_deck_topology(self.h).contents
        n = t.n_atoms

        def a(ptr, cnt, dt=np.float64):
            if cnt == 0:
                return np.zeros(0, dt)
            return np.ctypeslib.as_array(ptr, shape=(cnt,)).astype(dt).copy()
        T = t.n_types
        return dict(
            n_atoms=n, n_types=T, charge=a(t.charge, n), mass=a(t.mass, n), type=a(t.type, n, np.int32),
            lj_eps=a(t.lj_eps, T * T), lj_sig2=a(t.lj_sig2, T * T), lj14_eps=a(t.lj14_eps, T * T),
            lj14_sig2=a(t.lj14_sig2, T * T),
            n_bonds=t.n_bonds, bond_atoms=a(t.bond_atoms, 2 * t.n_bonds, np.int32), bond_k=a(t.bond_k, t.n_bonds),
            bond_r0=a(t.bond_r0, t.n_bonds),
            n_angles=t.n_angles, angle_atoms=a(t.angle_atoms, 3 * t.n_angles, np.int32),
            angle_k=a(t.angle_k, t.n_angles), angle_theta0=a(t.angle_theta0, t.n_angles),
            ub_k=a(t.ub_k, t.n_angles), ub_s0=a(t.ub_s0, t.n_angles),
            n_dihedrals=t.n_dihedrals, dihedral_atoms=a(t.dihedral_atoms, 4 * t.n_dihedrals, np.int32),
            dihedral_k=a(t.dihedral_k, t.n_dihedrals), dihedral_n=a(t.dihedral_n, t.n_dihedrals, np.int32),
            dihedral_delta=a(t.dihedral_delta, t.n_dihedrals),
            n_impropers=t.n_impropers, improper_atoms=a(t.improper_atoms, 4 * t.n_impropers, np.int32),
            improper_k=a(t.improper_k, t.n_impropers), improper_psi0=a(t.improper_psi0, t.n_impropers),
            n_constraints=t.n_constraints,
            constraint_atoms=a(t.constraint_atoms, 2 * t.n_constraints, np.int32),
            constraint_length=a(t.constraint_length, t.n_constraints))

    def params(self):
        p = lib().fz_deck_params(self.h).contents
        return dict(box=tuple(p.box[:]), cutoff=p.cutoff, cuton=p.cuton, pairlist=p.pairlist, beta=p.beta,
                    K=tuple(p.pme_grid[:]), dt_ps=p.dt_ps, shake_tol=p.shake_tol)

    def coordinates(self):
        n = lib().fz_deck_topology(self.h).contents.n_atoms
        return np.ctypeslib.as_array(lib().fz_deck_coordinates(self.h), shape=(n * 3,)).reshape(n, 3).copy()

    def settings(self):
        s = RunSettings()
        _check(lib().fz_deck_settings(self.h, C.byref(s)), "fz_deck_settings")
        return {f: getattr(s, f) for f, _ in RunSettings._fields_}

    def psf_counts(self):
        c = np.zeros(4, np.int64)
        _check(lib().fz_deck_psf_counts(self.h, _p(c, C.c_int64)), "fz_deck_psf_counts")
        return dict(zip(["bonds", "angles", "dihedrals", "impropers"], c.tolist()))

    def com_restraint(self):
        """(group, dist, fc) of the consharm keywords, or None."""
        n = self.topology()["n_atoms"]
        g = np.zeros(n, np.int32)
        dd = np.zeros(3)
        fc = C.c_double(0.0)
        r = lib().fz_deck_com_restraint(self.h, _p(g, C.c_int32), _p(dd, C.c_double), C.byref(fc))
        _check(min(r, 0), "fz_deck_com_restraint")
        return (g, dd, fc.value) if r == 1 else None

    def fixed_atoms(self):
        """consfix mask (uint8 per atom)."""
        n = self.topology()["n_atoms"]
        m = np.zeros(n, np.uint8)
        r = lib().fz_deck_fixed_atoms(self.h, _p(m, C.c_uint8))
        _check(min(r, 0), "fz_deck_fixed_atoms")
        return m


def energy_header():
    b = C.create_string_buffer(1024)
    lib().fz_format_energy_header(b, 1024)
    return b.value.decode()


def energy_row(e: Energies, dt_ps):
    b = C.create_string_buffer(1024)
    lib().fz_format_energy_row(C.byref(e), float(dt_ps), b, 1024)
    return b.value.decode()


def ns_per_day(ms_per_step, dt_ps):
    return 86400.0 / (ms_per_step * 1e-3) * dt_ps * 1e-3
