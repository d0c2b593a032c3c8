"""The allocation-free spectrum entry point (dj_field_power_spectrum_noalloc) against the allocating one
(dj_field_power_spectrum), both through the C ABI.  The device buffers of the inputs and outputs are
the caller's: this test allocates them with the HIP runtime the library itself links (libamdhip64),
since the entry point takes device pointers.  Two calls in a row on one device, the second with a
smaller field and fewer bins, so that what the first left in the shared scratch array must not leak
into the second; and a bin count beyond the scratch array is refused."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

D = pytest.importorskip("discodj_amd")

H2D, D2H = 1, 2


def _hip():
    """The HIP runtime the library was loaded with (the copy mapped into this process for it)."""
    D.lib()
    paths = [ln.split()[-1] for ln in open("/proc/self/maps") if "libamdhip64.so" in ln]
    hip = C.CDLL(paths[0])
    hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    hip.hipFree.argtypes = [C.c_void_p]
    hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return hip


class _Dev:
    def __init__(self, hip, nbytes):
        self.hip, self.p = hip, C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), nbytes) == 0

    def put(self, a):
        assert self.hip.hipMemcpy(self.p, a.ctypes.data, a.nbytes, H2D) == 0

    def get(self, a):
        assert self.hip.hipMemcpy(a.ctypes.data, self.p, a.nbytes, D2H) == 0
        return a

    def free(self):
        self.hip.hipFree(self.p)


@pytest.mark.parametrize("prec", [64, 32])
def test_noalloc_power_spectrum_matches_the_allocating_entry_point(prec):
    hip = _hip()
    L = D.lib()
    dt = np.float64 if prec == 64 else np.float32
    tol = 1e-12 if prec == 64 else 1e-6  # same kernels; only the order of the double atomics differs
    for N, bins, log, seed in ((64, 40, 1, 1), (32, 12, 0, 2)):
        field = np.random.default_rng(seed).normal(size=(N, N, N)).astype(dt)
        with D.PMSolverHIP(D.PMConfig(res_pm=N, n_part=16, boxsize=2.5 * N), precision=prec) as S:
            k_ref, pk_ref, nm_ref = S.field_power_spectrum(field, bins=bins, logarithmic=bool(log))
            d_f, d_k, d_pk, d_nm = _Dev(hip, field.nbytes), _Dev(hip, 8 * bins), _Dev(hip, 8 * bins), _Dev(hip, 8 * bins)
            try:
                d_f.put(field)
                rc = L.dj_field_power_spectrum_noalloc(S._ctx, d_f.p, bins, log, d_k.p, d_pk.p, d_nm.p)
                assert rc == 0, D.lib().dj_last_error(S._ctx)
                assert L.dj_synchronize(S._ctx) == 0
                k, pk, nm = d_k.get(np.empty(bins)), d_pk.get(np.empty(bins)), d_nm.get(np.empty(bins, np.int64))
                assert L.dj_field_power_spectrum_noalloc(S._ctx, d_f.p, 257, log, d_k.p, d_pk.p, d_nm.p) == -2
            finally:
                for d in (d_f, d_k, d_pk, d_nm):
                    d.free()
        assert np.array_equal(nm, nm_ref)
        assert np.allclose(k, k_ref, rtol=1e-13, atol=0)
        ok = nm_ref > 0
        assert np.max(np.abs(pk[ok] / pk_ref[ok] - 1)) < tol and np.all(pk[~ok] == 0)
