"""The MATLAB MEX gateway (fmmlib3d_amd/matlab/fmm3d_mi355x_mex.c) compiled
against a test double of MATLAB's mx API (tests/matlab/: real and complex
column-major matrices, mxMalloc temporaries, mexErrMsgIdAndTxt returning by
longjmp) and linked to libfmm3d_mi355x.so: every routine it offers gives
bitwise the results of the Python module on the same inputs (the gateway's
argument order, shapes and flag handling), and malformed calls end in
MATLAB-style errors before any GPU work.  The .m files beside the gateway
need a MATLAB interpreter and are not run here."""
import ctypes
import os

import numpy as np
import pytest

import fmmlib3d_amd as F
from fmm_inputs import icosphere

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "matlab", "libmex_standin.so")

pytestmark = pytest.mark.skipif(not os.path.exists(SO), reason="gateway test library not built (make -C tests/matlab)")


class Mex:
    def __init__(self):
        L = ctypes.CDLL(SO)
        vp = ctypes.c_void_p
        for name, res, args in [
                ("standin_real", vp, [ctypes.c_size_t, ctypes.c_size_t, vp]),
                ("standin_complex", vp, [ctypes.c_size_t, ctypes.c_size_t, vp]),
                ("standin_string", vp, [ctypes.c_char_p]),
                ("standin_data", vp, [vp]),
                ("standin_is_complex", ctypes.c_int, [vp]),
                ("standin_call", ctypes.c_int, [ctypes.c_int, vp, ctypes.c_int, vp, ctypes.c_char_p, ctypes.c_size_t]),
                ("mxGetM", ctypes.c_size_t, [vp]), ("mxGetN", ctypes.c_size_t, [vp]),
                ("mxGetScalar", ctypes.c_double, [vp]), ("mxDestroyArray", None, [vp])]:
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        self.L = L

    def arg(self, v):
        """A MATLAB value from a Python one: str, scalar, or a numpy array
        given in MATLAB's column-major shape as (n, m) C order (m x n)."""
        L = self.L
        if isinstance(v, str):
            return L.standin_string(v.encode())
        if np.isscalar(v):
            if isinstance(v, complex):
                a = np.array([v], np.complex128)
                return L.standin_complex(1, 1, a.ctypes.data)
            return L.standin_real(1, 1, np.array([float(v)]).ctypes.data)
        a = np.asarray(v)
        if a.size == 0:
            return L.standin_real(0, 0, None)
        n = a.shape[0]
        m = a.size // n if a.ndim > 1 else 1
        if a.ndim == 1:
            m, n = 1, a.shape[0]
        if np.iscomplexobj(a):
            a = np.ascontiguousarray(a, np.complex128)
            return L.standin_complex(m, n, a.ctypes.data)
        a = np.ascontiguousarray(a, np.float64)
        return L.standin_real(m, n, a.ctypes.data)

    def call(self, nlhs, *args):
        L = self.L
        prhs = (ctypes.c_void_p * len(args))(*[self.arg(a) for a in args])
        plhs = (ctypes.c_void_p * max(nlhs, 1))()
        err = ctypes.create_string_buffer(512)
        rc = L.standin_call(nlhs, plhs, len(args), prhs, err, 512)
        for p in prhs:
            L.mxDestroyArray(p)
        if rc:
            raise RuntimeError(err.value.decode())
        out = []
        for k in range(nlhs):
            p = plhs[k]
            m, n = L.mxGetM(p), L.mxGetN(p)
            if L.standin_is_complex(p):
                d = np.ctypeslib.as_array(ctypes.cast(L.standin_data(p), ctypes.POINTER(ctypes.c_double)),
                                          shape=(2 * m * n,)).copy().view(np.complex128)
                val = d.reshape(n, m) if m > 1 else d
            else:
                val = L.mxGetScalar(p)
            L.mxDestroyArray(p)
            out.append(val)
        return out


@pytest.fixture(scope="module")
def mex():
    return Mex()


def test_gateway_rejects_malformed_calls(mex):
    with pytest.raises(RuntimeError, match="unknown routine"):
        mex.call(1, "lfmm3dpartfoo", 1)
    with pytest.raises(RuntimeError, match="10 arguments"):
        mex.call(5, "lfmm3dparttarg", 1, np.zeros((4, 3)))
    bad = np.zeros((4, 2))   # a 2 x 4 source
    with pytest.raises(RuntimeError, match="source must be a real 3 x n"):
        mex.call(5, "lfmm3dparttarg", 1, bad, np.ones(4, complex), np.zeros(0), np.zeros(0), 1, 1,
                 np.zeros(0), 0, 0)
    with pytest.raises(RuntimeError, match="charge must have 4"):
        mex.call(5, "lfmm3dparttarg", 1, np.zeros((4, 3)), np.ones(3, complex), np.zeros(0), np.zeros(0), 1, 1,
                 np.zeros(0), 0, 0)
    with pytest.raises(RuntimeError, match="triaflat must be a real 9 x n"):
        mex.call(5, "lfmm3dtriatarg", 1, np.zeros((4, 3)), np.zeros((4, 3)), 1, np.ones(4, complex), 0,
                 np.zeros(0), 1, 1, np.zeros(0), 0, 0)


@pytest.mark.gpu
def test_gateway_particle_routines_match_module(mex):
    rng = np.random.default_rng(11)
    n, nt = 3000, 500
    src = rng.random((n, 3))
    trg = rng.random((nt, 3))
    q = rng.normal(size=n) + 1j * rng.normal(size=n)
    ds = rng.normal(size=n) + 1j * rng.normal(size=n)
    dv = rng.normal(size=(n, 3))
    ier, pot, fld, pt, ft = mex.call(5, "lfmm3dparttarg", 2, src, q, ds, dv, 1, 1, trg, 1, 1)
    r = F.lfmm3dparttarg(2, src, q, ds, dv, trg)
    assert ier == 0 and r[0] == 0
    for a, b in zip((pot, fld, pt, ft), r[1:]):
        assert np.array_equal(a, b)
    zk = 3.0 + 0.5j
    ier, pot, fld, pt, ft = mex.call(5, "hfmm3dparttarg", 2, zk, src, q, np.zeros(0), np.zeros(0), 1, 1, trg, 1, 1)
    r = F.hfmm3dparttarg(2, zk, src, q, None, None, trg)
    assert ier == 0 and r[0] == 0
    for a, b in zip((pot, fld, pt, ft), r[1:]):
        assert np.array_equal(a, b)
    pot, fld, pt, ft = mex.call(4, "l3dpartdirect", src, q, ds, dv, 1, 1, trg, 1, 1)
    r = F.l3dpartdirect(src, q, ds, dv, trg)
    for a, b in zip((pot, fld, pt, ft), r):
        assert np.array_equal(a, b)
    pot, fld, pt, ft = mex.call(4, "h3dpartdirect", zk, src, q, ds, dv, 1, 1, trg, 1, 1)
    r = F.h3dpartdirect(zk, src, q, ds, dv, trg)
    for a, b in zip((pot, fld, pt, ft), r):
        assert np.array_equal(a, b)


@pytest.mark.gpu
def test_gateway_triangle_routines_match_module(mex):
    rng = np.random.default_rng(12)
    T = icosphere(3)
    n = T.shape[0]
    cen = T.mean(axis=1)
    trg = rng.normal(size=(300, 3))
    q = rng.normal(size=n) + 1j * rng.normal(size=n)
    mu = rng.normal(size=n) + 1j * rng.normal(size=n)
    tri = T.reshape(n, 9)
    ier, pot, fld, pt, ft = mex.call(5, "lfmm3dtriatarg", 2, tri, cen, 1, q, 1, mu, 1, 1, trg, 1, 1)
    r = F.lfmm3dtriatarg(2, T, q, mu, trg)
    assert ier == 0 and r[0] == 0
    for a, b in zip((pot, fld, pt, ft), r[1:]):
        assert np.array_equal(a, b)
    zk = 2.0 + 0.1j
    ier, pot, fld, pt, ft = mex.call(5, "hfmm3dtriatarg", 2, zk, tri, cen, 2, q, 1, mu, 1, 1, trg, 1, 1)
    r = F.hfmm3dtriatarg(2, zk, T, q, mu, trg, ifcharge=2, ifdipole=1)
    assert ier == 0 and r[0] == 0
    for a, b in zip((pot, fld, pt, ft), r[1:]):
        assert np.array_equal(a, b)
    pot, fld, pt, ft = mex.call(4, "l3dtriadirect", tri, cen, 1, q, 1, mu, 1, 1, trg, 1, 1)
    r = F.l3dtriadirect(T, q, mu, trg)
    for a, b in zip((pot, fld, pt, ft), r):
        assert np.array_equal(a, b)
    pot, fld, pt, ft = mex.call(4, "h3dtriadirect", 4, zk, tri, cen, 1, q, 2, mu, 1, 1, trg, 1, 1)
    r = F.h3dtriadirect(4, zk, T, q, mu, trg, ifcharge=1, ifdipole=2)
    for a, b in zip((pot, fld, pt, ft), r):
        assert np.array_equal(a, b)
