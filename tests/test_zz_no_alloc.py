"""GPU test of fmm3d_lfmm3dparttarg_dev_noalloc (the entry point that
allocates no device memory: its buffers are slices of a static device
array): through the C ABI, its results equal those of
fmm3d_lfmm3dparttarg_dev on the same inputs bitwise, on two calls in a row,
the second with a smaller input, a lower iprec and another precision mode."""
import ctypes
import os

import numpy as np
import pytest

import fmmlib3d_amd as F
from fmm_inputs import points, strengths

pytestmark = pytest.mark.gpu


def _hip():
    rt = ctypes.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libamdhip64.so"))
    rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    rt.hipFree.argtypes = [ctypes.c_void_p]
    rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    rt.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    rt.hipDeviceSynchronize.argtypes = []
    return rt


def _run_both(rt, lib, prec, iprec, src, q, ds, dv, trg):
    """The same call through _dev and _dev_noalloc; returns both results."""
    ns, nt = src.shape[0], trg.shape[0]
    bufs = []

    def dev(a):
        p = ctypes.c_void_p()
        assert rt.hipMalloc(ctypes.byref(p), max(a.nbytes, 8)) == 0
        assert rt.hipMemcpy(p, a.ctypes.data_as(ctypes.c_void_p), a.nbytes, 1) == 0  # H2D
        bufs.append(p)
        return p

    def zeros_dev(nbytes):
        p = ctypes.c_void_p()
        assert rt.hipMalloc(ctypes.byref(p), nbytes) == 0
        assert rt.hipMemset(p, 0, nbytes) == 0
        bufs.append(p)
        return p

    shapes = [(ns,), (ns, 3), (nt,), (nt, 3)]
    res = {}
    try:
        ins = [dev(np.ascontiguousarray(a)) for a in (src, q, ds, dv, trg)]
        for name in ("fmm3d_lfmm3dparttarg_dev", "fmm3d_lfmm3dparttarg_dev_noalloc"):
            outs = [zeros_dev(int(np.prod(s)) * 16) for s in shapes]
            print("case", name, "prec", prec, "iprec", iprec, "ns", ns, "nt", nt, flush=True)
            ier = getattr(lib, name)(prec, iprec, ns, ins[0], 1, ins[1], 1, ins[2], ins[3], 1, outs[0], 1,
                                     outs[1], nt, ins[4], 1, outs[2], 1, outs[3], None)
            assert ier == 0, (name, ier)
            assert rt.hipDeviceSynchronize() == 0
            hs = []
            for s, d in zip(shapes, outs):
                h = np.zeros(s, np.complex128)
                assert rt.hipMemcpy(h.ctypes.data_as(ctypes.c_void_p), d, h.nbytes, 2) == 0  # D2H
                hs.append(h)
            res[name] = hs
    finally:
        for p in bufs:
            rt.hipFree(p)
    return res["fmm3d_lfmm3dparttarg_dev"], res["fmm3d_lfmm3dparttarg_dev_noalloc"]


def test_noalloc_equals_dev_on_two_calls():
    rt = _hip()
    lib = F.load()
    rng = np.random.default_rng(77)
    # first call: 60,000 clustered sources + 30,000 targets, iprec 2, FP64
    # second call (smaller, reusing the array): 20,000 + 5,000 uniform, iprec 1, FP32 near field
    for prec, iprec, ns, nt, dist in ((F.PREC_FP64, 2, 60000, 30000, "clusters"),
                                      (F.PREC_MIXED, 1, 20000, 5000, "cube")):
        src = points(rng, ns, dist)
        trg = points(rng, nt, dist)
        q, ds, dv = strengths(rng, ns, True, True)
        a, b = _run_both(rt, lib, prec, iprec, src, q, ds, dv, trg)
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
        assert np.linalg.norm(a[2]) > 0
