"""Checks that pin the oracle through mathematics alone (no reference
outputs): closed forms, exact identities and brute force on tiny inputs,
chosen so that a dropped term, a wrong sign or index or a transposed
operand fails one of them."""
import ctypes
import math

import numpy as np
import pytest
from scipy.special import lpmv

import oracle_bindings as ob
from oracle_bindings import _vp
from fmm_inputs import IPREC_EPS, points, rel, strengths


def idx(n, m):
    return n * n + n + m


def yhat(n, m, x):
    """(-1)^m sqrt((n-|m|)!/(n+|m|)!) P_n^|m|(cos t) e^{i m phi}, P without
    Condon-Shortley phase (scipy's lpmv includes it: multiply by (-1)^|m|)."""
    r = np.linalg.norm(x)
    ct = x[2] / r
    ph = math.atan2(x[1], x[0])
    am = abs(m)
    P = lpmv(am, n, ct) * (-1) ** am
    return (-1) ** m * math.sqrt(math.factorial(n - am) / math.factorial(n + am)) * P * np.exp(1j * m * ph)


def test_solid_harmonics_closed_form(rng):
    o = ob.oracle()
    p = 12
    for _ in range(5):
        x = rng.normal(size=3)
        R = np.zeros((p + 1) ** 2, np.complex128)
        S = np.zeros((p + 1) ** 2, np.complex128)
        o.orc_regular(p, _vp(x), _vp(R))
        o.orc_irregular(p, _vp(x), _vp(S))
        r = np.linalg.norm(x)
        for n in range(p + 1):
            for m in range(-n, n + 1):
                y = yhat(n, m, x)
                assert abs(R[idx(n, m)] - r ** n * y) < 1e-12 * max(1, r ** n)
                assert abs(S[idx(n, m)] - y / r ** (n + 1)) < 1e-12 * max(1, r ** -(n + 1))


def test_addition_theorem(rng):
    """1/|x-y| = sum_nm conj(R_n^m(y)) S_n^m(x) for |y| < |x|."""
    o = ob.oracle()
    p = 40
    x = np.array([0.9, -1.1, 0.7])
    d = rng.normal(size=3)
    y = 0.4 * d / np.linalg.norm(d)      # |y|/|x| = 0.25: the series has converged at p = 40
    R = np.zeros((p + 1) ** 2, np.complex128)
    S = np.zeros((p + 1) ** 2, np.complex128)
    o.orc_regular(p, _vp(y), _vp(R))
    o.orc_irregular(p, _vp(x), _vp(S))
    assert abs(np.sum(np.conj(R) * S) - 1 / np.linalg.norm(x - y)) < 1e-13


def _direct_numpy(src, q, trg):
    d = trg[:, None, :] - src[None, :, :]
    r = np.linalg.norm(d, axis=2)
    return (q[None, :] / r).sum(1)


def test_direct_field_is_minus_gradient(rng):
    """fld = -grad pot (central differences) and the dipole potential is the
    limit of a charge pair: checks signs and the dipole formula."""
    ns = 40
    src = rng.random((ns, 3))
    q, ds, dv = strengths(rng, ns, True, True, unit_dipvec=False)
    trg = rng.random((6, 3)) + 2.0
    pot, fld, pt, ft = ob.orc_l3dpartdirect(src, q, ds, dv, trg)
    h = 1e-5
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        pp = ob.orc_l3dpartdirect(src, q, ds, dv, trg + e)[2]
        pm = ob.orc_l3dpartdirect(src, q, ds, dv, trg - e)[2]
        assert rel(-(pp - pm) / (2 * h), ft[:, k]) < 1e-8
    # charges only: brute-force numpy sum
    pot2 = ob.orc_l3dpartdirect(src, q, None, None, trg)[2]
    assert rel(pot2, _direct_numpy(src, q, trg)) < 1e-14
    # dipole = d/ds of a charge displaced along dipvec: (q(y+hv) - q(y-hv))/(2h)
    pd = ob.orc_l3dpartdirect(src, None, ds, dv, trg)[2]
    acc = np.zeros(trg.shape[0], np.complex128)
    for i in range(ns):
        a = _direct_numpy(src[i:i + 1] + h * dv[i], ds[i:i + 1], trg)
        b = _direct_numpy(src[i:i + 1] - h * dv[i], ds[i:i + 1], trg)
        acc += (a - b) / (2 * h)
    assert rel(pd, acc) < 1e-8


def test_direct_skips_only_self(rng):
    src = rng.random((30, 3))
    q = rng.normal(size=30) + 0j
    pot = ob.orc_l3dpartdirect(src, q)[0]
    ref = np.array([sum(q[i] / np.linalg.norm(src[j] - src[i]) for i in range(30) if i != j) for j in range(30)])
    assert rel(pot, ref) < 1e-14


def test_rotation_is_orthogonal(rng):
    o = ob.oracle()
    p = 30
    K = (p + 1) ** 2
    a = rng.normal(size=K) + 1j * rng.normal(size=K)
    b = np.zeros(K, np.complex128)
    c = np.zeros(K, np.complex128)
    o.orc_rotate_y(p, ctypes.c_double(0.83), _vp(a), _vp(b))
    o.orc_rotate_y(p, ctypes.c_double(-0.83), _vp(b), _vp(c))
    assert rel(c, a) < 1e-13
    for n in range(p + 1):
        sl = slice(n * n, (n + 1) ** 2)
        assert abs(np.linalg.norm(b[sl]) - np.linalg.norm(a[sl])) < 1e-12 * np.linalg.norm(a[sl])


def _p2m(src, q, c, p, sc=1.0):
    o = ob.oracle()
    M = np.zeros((p + 1) ** 2, np.complex128)
    o.orc_l3dformmp(ctypes.c_double(sc), _vp(np.ascontiguousarray(src)), _vp(q), src.shape[0], _vp(c), p, _vp(M))
    return M


def test_m2m_is_exact(rng):
    """Shifting a degree-p multipole gives the degree-p multipole of the same
    sources about the new centre exactly."""
    o = ob.oracle()
    p = 14
    c1 = np.array([0.1, 0.2, 0.3])
    c2 = c1 + np.array([-0.13, 0.07, 0.11])
    src = c1 + 0.05 * rng.normal(size=(7, 3))
    q = rng.normal(size=7) + 1j * rng.normal(size=7)
    M1 = _p2m(src, q, c1, p, 2.0)
    M2 = np.zeros_like(M1)
    o.orc_l3dmpmp(ctypes.c_double(2.0), _vp(c1), _vp(M1), p, ctypes.c_double(1.5), _vp(c2), _vp(M2), p)
    assert rel(M2, _p2m(src, q, c2, p, 1.5)) < 1e-13


def test_l2l_is_exact(rng):
    """A local expansion is a polynomial: re-centring it changes no value."""
    o = ob.oracle()
    p = 12
    c0 = np.array([0.0, 0.0, 0.0])
    c1 = np.array([0.05, -0.04, 0.03])
    src = np.array([[2.0, 1.0, -1.5], [-1.7, 2.2, 0.4]])
    q = np.array([1.0 + 0.5j, -0.7 + 0.2j])
    L = np.zeros((p + 1) ** 2, np.complex128)
    o.orc_l3dformta(ctypes.c_double(1.0), _vp(src), _vp(q), 2, _vp(c0), p, _vp(L))
    L1 = np.zeros_like(L)
    o.orc_l3dlocloc(ctypes.c_double(1.0), _vp(c0), _vp(L), p, ctypes.c_double(1.0), _vp(c1), _vp(L1), p)
    x = c1 + 0.02 * rng.normal(size=(5, 3))
    a, fa = np.zeros(5, np.complex128), np.zeros((5, 3), np.complex128)
    b, fb = np.zeros(5, np.complex128), np.zeros((5, 3), np.complex128)
    o.orc_l3dtaeval(ctypes.c_double(1.0), _vp(c0), _vp(L), p, _vp(x), 5, 1, _vp(a), 1, _vp(fa))
    o.orc_l3dtaeval(ctypes.c_double(1.0), _vp(c1), _vp(L1), p, _vp(x), 5, 1, _vp(b), 1, _vp(fb))
    assert rel(b, a) < 1e-13 and rel(fb, fa) < 1e-13


def test_m2l_converges_to_direct(rng):
    o = ob.oracle()
    c1 = np.array([0.0, 0.0, 0.0])
    c0 = np.array([2.0, 1.0, 0.0])
    src = c1 + 0.2 * (rng.random((5, 3)) - 0.5)
    q = rng.normal(size=5) + 0j
    x = c0 + 0.1 * (rng.random((3, 3)) - 0.5)
    direct = ob.orc_l3dpartdirect(src, q, None, None, x)[2]
    errs = []
    for p in (4, 8, 16):
        M = _p2m(src, q, c1, p)
        L = np.zeros_like(M)
        o.orc_l3dmploc(ctypes.c_double(1.0), _vp(c1), _vp(M), p, p, ctypes.c_double(1.0), _vp(c0), _vp(L), p)
        pot, fld = np.zeros(3, np.complex128), np.zeros((3, 3), np.complex128)
        o.orc_l3dtaeval(ctypes.c_double(1.0), _vp(c0), _vp(L), p, _vp(x), 3, 1, _vp(pot), 0, _vp(fld))
        errs.append(rel(pot, direct))
    assert errs[2] < 1e-10 and errs[0] > errs[1] > errs[2]


@pytest.mark.parametrize("iprec,dist", [(-1, "clusters"), (0, "cube"), (1, "sphere"), (2, "cube"), (3, "clusters")])
def test_fmm_within_iprec_tolerance(rng, iprec, dist):
    ns, nt = 3000, 1500
    src = points(rng, ns, dist)
    trg = points(rng, nt, dist)
    q, ds, dv = strengths(rng, ns, True, True)
    ier, pot, fld, pt, ft = ob.orc_lfmm3dparttarg(iprec, src, q, ds, dv, trg)
    d = ob.orc_l3dpartdirect(src, q, ds, dv, trg)
    assert ier == 0
    # iprec names the tolerance of the potential; the field, a derivative of
    # the truncated expansions, carries about one digit less
    tol = IPREC_EPS[iprec]
    for a, b, t in zip((pot, fld, pt, ft), d, (tol, 10 * tol, tol, 10 * tol)):
        assert rel(a, b) < t
