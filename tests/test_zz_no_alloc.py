"""The allocation-free MLP loss/gradient entry point (librpa_amd_mlp_loss_grad_noalloc) against
the allocating one (librpa_amd_mlp_loss_grad): the same kernels in the same order, so the gradient
and the loss must agree bit for bit, on two calls in a row (the second with fewer samples, so the
static work array still holds the first call's larger contents), and a sample count above the
array's maximum must be refused before any kernel runs.

Device memory comes from the HIP runtime the library is bound to (loaded by its soname)."""
import ctypes as C

import numpy as np
import pytest

import librpa_amd as L

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2  # hipMemcpyHostToDevice, hipMemcpyDeviceToHost


def _hip():
    L.load()
    h = C.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipMalloc.restype = C.c_int
    h.hipFree.argtypes = [C.c_void_p]
    h.hipFree.restype = C.c_int
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipMemcpy.restype = C.c_int
    h.hipDeviceSynchronize.restype = C.c_int
    return h


def test_noalloc_loss_grad_matches_allocating_entry(lib):
    print("no-alloc MLP loss/gradient vs the allocating entry", flush=True)
    hip = _hip()
    rng = np.random.default_rng(12)
    npar = L.SCREEN_NPARAM
    p = (0.1 * rng.standard_normal(npar)).astype(np.float32)
    nmax = 20000
    sizes = {"p": 4 * npar, "X": 4 * 8 * nmax, "t": 4 * nmax, "g": 4 * npar, "l": 8}
    d = {}
    try:
        for k, b in sizes.items():
            ptr = C.c_void_p()
            assert hip.hipMalloc(C.byref(ptr), b) == 0
            d[k] = ptr.value
        assert hip.hipMemcpy(d["p"], p.ctypes.data, p.nbytes, H2D) == 0
        for n in (nmax, 3001):
            X = rng.standard_normal((n, 8)).astype(np.float32)
            t = rng.standard_normal(n).astype(np.float32)
            loss_ref, g_ref = L.mlp_loss_grad(p, X, t)
            assert hip.hipMemcpy(d["X"], X.ctypes.data, X.nbytes, H2D) == 0
            assert hip.hipMemcpy(d["t"], t.ctypes.data, t.nbytes, H2D) == 0
            L.mlp_loss_grad_noalloc(d["p"], d["X"], d["t"], n, d["g"], d["l"], None)
            assert hip.hipDeviceSynchronize() == 0
            g = np.empty(npar, dtype=np.float32)
            loss = np.empty(1)
            assert hip.hipMemcpy(g.ctypes.data, d["g"], g.nbytes, D2H) == 0
            assert hip.hipMemcpy(loss.ctypes.data, d["l"], 8, D2H) == 0
            print(f"  n={n}: loss {loss[0]!r} (allocating entry {loss_ref!r}), "
                  f"max |grad diff| {np.abs(g - g_ref).max():.1e}", flush=True)
            assert np.array_equal(g, g_ref)
            assert loss[0] == loss_ref
        with pytest.raises(L.LibrpaError, match="exceed"):
            L.mlp_loss_grad_noalloc(d["p"], d["X"], d["t"], L.MLP_NOALLOC_MAX + 1, d["g"], d["l"], None)
    finally:
        for ptr in d.values():
            hip.hipFree(ptr)
