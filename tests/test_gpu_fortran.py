"""The Fortran drop-in: a Fortran program calling lfmm3dparttarg and
l3dpartdirect with the reference's calling sequences, linked against
libfmm3d_mi355x_f.so (ISO_C_BINDING over the C ABI)."""
import os
import subprocess

import pytest

import fmmlib3d_amd.build as B

pytestmark = pytest.mark.gpu


@pytest.mark.skipif(not os.path.exists(B.FDRIVER), reason="Fortran driver not built (no Fortran compiler)")
def test_fortran_driver_runs_on_gpu():
    out = subprocess.run([B.FDRIVER], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert out.returncode == 0, out.stdout
    line = [l for l in out.stdout.splitlines() if l.startswith("relerr")][-1]
    errs = [float(x) for x in line.split()[-4:]]
    # iprec = 1: tolerance 0.5e-3 for the potential, a digit less for the field
    assert errs[0] < 5e-4 and errs[2] < 5e-4
    assert errs[1] < 5e-3 and errs[3] < 5e-3
    line = [l for l in out.stdout.splitlines() if l.startswith("hrelerr")][-1]
    errs = [float(x) for x in line.split()[-2:]]
    assert errs[0] < 5e-4 and errs[1] < 5e-3


@pytest.mark.skipif(not os.path.exists(B.FLIB), reason="Fortran library not built (no Fortran compiler)")
def test_fortran_triangle_entries_match_c_abi():
    """The triangle entries of the Fortran drop-in (lfmm3dtriatarg_,
    hfmm3dtriatarg_, h3dtriadirect_: every argument by reference, the
    reference's order) give bitwise the results of the C ABI."""
    import ctypes

    import numpy as np

    import fmmlib3d_amd as F
    from fmm_inputs import icosphere

    lib = ctypes.CDLL(B.FLIB)
    rng = np.random.default_rng(5)
    T = icosphere(3)
    n = T.shape[0]
    q = rng.normal(size=n) + 1j * rng.normal(size=n)
    mu = rng.normal(size=n) + 1j * rng.normal(size=n)
    trg = rng.normal(size=(200, 3))
    tri = np.ascontiguousarray(T.reshape(n, 9))
    nrm = np.cross(T[:, 1] - T[:, 0], T[:, 2] - T[:, 0])
    nrm = np.ascontiguousarray(nrm / np.linalg.norm(nrm, axis=1)[:, None])
    cen = np.ascontiguousarray(T.mean(axis=1))
    I = lambda v: ctypes.byref(ctypes.c_int(v))  # noqa: E731
    P = lambda a: a.ctypes.data_as(ctypes.c_void_p)  # noqa: E731
    zk = np.array([2.0 + 0.5j])

    def outs():
        return (np.zeros(n, np.complex128), np.zeros((n, 3), np.complex128), np.zeros(200, np.complex128),
                np.zeros((200, 3), np.complex128))

    pot, fld, pt, ft = outs()
    ier = ctypes.c_int(-1)
    lib.hfmm3dtriatarg_(ctypes.byref(ier), I(2), P(zk), I(n), P(tri), P(nrm), P(cen), I(1), P(q), I(1), P(mu),
                        P(nrm), I(1), P(pot), I(1), P(fld), I(200), P(trg), I(1), P(pt), I(1), P(ft))
    r = F.hfmm3dtriatarg(2, zk[0], T, q, mu, trg)
    assert ier.value == 0 and r[0] == 0
    for a, b in zip((pot, fld, pt, ft), r[1:]):
        assert np.array_equal(a, b)
    pot, fld, pt, ft = outs()
    lib.lfmm3dtriatarg_(ctypes.byref(ier), I(2), I(n), P(tri), P(nrm), P(cen), I(1), P(q), I(1), P(mu),
                        P(nrm), I(1), P(pot), I(1), P(fld), I(200), P(trg), I(1), P(pt), I(1), P(ft))
    r = F.lfmm3dtriatarg(2, T, q, mu, trg)
    assert ier.value == 0 and r[0] == 0
    for a, b in zip((pot, fld, pt, ft), r[1:]):
        assert np.array_equal(a, b)
    pot, fld, pt, ft = outs()
    lib.h3dtriadirect_(I(4), P(zk), I(n), P(tri), P(nrm), P(cen), I(1), P(q), I(2), P(mu), P(nrm), I(1), P(pot),
                       I(1), P(fld), I(200), P(trg), I(1), P(pt), I(1), P(ft))
    r = F.h3dtriadirect(4, zk[0], T, q, mu, trg, ifcharge=1, ifdipole=2)
    for a, b in zip((pot, fld, pt, ft), r):
        assert np.array_equal(a, b)
