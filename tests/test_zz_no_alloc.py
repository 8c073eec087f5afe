"""fz_pme_convolve_noalloc (every buffer beyond the caller's grids in one static
device array, no allocation) against fz_pme_convolve, through the C ABI, in
both precision forms: two calls in a row on the same GPU, the second on a
smaller grid, each equal to the allocating entry point's result (FP64: 1e-12
of the largest potential value, mixed: 1e-6; the two differ only in where
the per-axis tables are computed, host or device libm), and a grid beyond
the scratch array's 2^18 points rejected before anything is enqueued.

The caller's device grids are allocated here through the HIP runtime
libfenzi links (one runtime per process)."""
import ctypes as C

import numpy as np
import pytest

import fenzi

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class _Hip:
    def __init__(self):
        self.h = C.CDLL("libamdhip64.so.7")
        self.bufs = []

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), C.c_size_t(nbytes)) == 0
        self.bufs.append(p)
        return p.value

    def put(self, dst, a):
        a = np.ascontiguousarray(a)
        assert self.h.hipMemcpy(C.c_void_p(dst), a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), H2D) == 0

    def get(self, src, n, dtype=np.float64):
        a = np.empty(n, dtype)
        assert self.h.hipMemcpy(a.ctypes.data_as(C.c_void_p), C.c_void_p(src), C.c_size_t(a.nbytes), D2H) == 0
        return a

    def sync(self):
        assert self.h.hipDeviceSynchronize() == 0

    def close(self):
        for p in self.bufs:
            self.h.hipFree(p)
        self.bufs = []


@pytest.mark.parametrize("prec", [fenzi.FP64, fenzi.MIXED])
def test_noalloc_convolution_matches_the_allocating_entry_point(ctx, prec):
    hip = _Hip()
    try:
        n_max = 40 * 36 * 48
        q_dev, phi_dev, e_dev = hip.alloc(8 * n_max), hip.alloc(8 * n_max), hip.alloc(8)
        box, beta = (31.0, 29.0, 37.0), 0.31
        rng = np.random.default_rng(11 + prec)
        for K in ((40, 36, 48), (16, 20, 18)):  # the second call on a smaller grid
            n = K[0] * K[1] * K[2]
            q = rng.standard_normal((K[2], K[1], K[0]))
            print("noalloc prec=%d K=%s" % (prec, K), flush=True)
            hip.put(q_dev, q)
            hip.put(phi_dev, np.full(n, np.nan))
            ctx.pme_convolve_noalloc(K, box, beta, q_dev, phi_dev, e_dev, precision=prec, mode=0)
            hip.sync()
            phi = hip.get(phi_dev, n).reshape(q.shape)
            e = float(hip.get(e_dev, 1)[0])
            phi_ref, e_ref = ctx.pme_convolve(q, box, beta, prec, mode=0)
            tol = 1e-12 if prec == fenzi.FP64 else 1e-6
            assert np.isfinite(phi).all()
            assert np.abs(phi - phi_ref).max() <= tol * np.abs(phi_ref).max()
            assert abs(e - e_ref) <= tol * abs(e_ref)
        # an oversize grid is refused before anything is enqueued
        with pytest.raises(fenzi.FenziError) as err:
            ctx.pme_convolve_noalloc((64, 64, 72), box, beta, q_dev, phi_dev, e_dev, precision=prec)
        assert err.value.code == fenzi.ERR_ARG
    finally:
        hip.sync()
        hip.close()
