"""The Helmholtz half of the CPU oracle, pinned three ways: against the
golden fixtures written by the reference Fortran (tests/golden/make_golden.py),
against closed forms (scipy's spherical Bessel functions, the direct sum) and
against invariants the translations must satisfy (M2M, L2L and M2L reproduce
what forming the expansion at the new centre gives)."""
import ctypes
import glob
import os

import numpy as np
import pytest
import scipy.special as sp

import oracle_bindings as ob
from fmm_inputs import IPREC_EPS, points, rel, strengths
from oracle_bindings import ZC, _vp, zc

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
HELM_FIXTURES = sorted(glob.glob(os.path.join(GOLD, "helm_targ_*.npz")) + glob.glob(os.path.join(GOLD, "helm_self_*.npz")))
V, I_, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double


def load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


def _o():
    o = ob.oracle()
    o.orc_h3dall.argtypes = [I_, ZC, D, V]
    o.orc_jfuns3d.argtypes = [I_, ZC, D, V]
    o.orc_jfuns3d.restype = I_
    o.orc_h3dterms.argtypes = [D, ZC, D, V]
    o.orc_h3dterms.restype = I_
    o.orc_h3dterms_list2.argtypes = [D, ZC, D, V, V]
    for f in ("orc_h3dformmp", "orc_h3dformta"):
        getattr(o, f).argtypes = [ZC, D, V, V, I_, V, I_, V]
    for f in ("orc_h3dformmp_dp", "orc_h3dformta_dp"):
        getattr(o, f).argtypes = [ZC, D, V, V, V, I_, V, I_, V]
    for f in ("orc_h3dmpeval", "orc_h3dtaeval"):
        getattr(o, f).argtypes = [ZC, D, V, V, I_, V, I_, I_, V, I_, V]
    o.orc_h3dmpmp.argtypes = [ZC, D, V, V, I_, D, V, V, I_]
    o.orc_h3dlocloc.argtypes = [ZC, D, V, V, I_, D, V, V, I_]
    o.orc_h3dmploc.argtypes = [ZC, D, V, V, I_, I_, D, V, V, I_]
    return o


def _zeros(p):
    return np.zeros((p + 1) ** 2, np.complex128)


def form(kind, zk, sc, src, q, c, p, ds=None, dv=None):
    o = _o()
    out = _zeros(p)
    src = np.ascontiguousarray(src)
    c = np.ascontiguousarray(c, dtype=np.float64)
    if q is not None:
        getattr(o, "orc_h3dform" + kind)(zc(zk), sc, _vp(src), _vp(np.ascontiguousarray(q)), src.shape[0], _vp(c), p, _vp(out))
    if ds is not None:
        getattr(o, "orc_h3dform" + kind + "_dp")(zc(zk), sc, _vp(src), _vp(np.ascontiguousarray(ds)),
                                                 _vp(np.ascontiguousarray(dv)), src.shape[0], _vp(c), p, _vp(out))
    return out


def evaluate(kind, zk, sc, c, ex, p, x):
    o = _o()
    x = np.ascontiguousarray(x)
    pot = np.zeros(x.shape[0], np.complex128)
    fld = np.zeros((x.shape[0], 3), np.complex128)
    getattr(o, "orc_h3d" + kind + "eval")(zc(zk), sc, _vp(np.ascontiguousarray(c, dtype=np.float64)),
                                          _vp(ex), p, _vp(x), x.shape[0], 1, _vp(pot), 1, _vp(fld))
    return pot, fld


# ----------------------------------------------------------------------------
# golden fixtures
# ----------------------------------------------------------------------------
@pytest.mark.parametrize("path", HELM_FIXTURES, ids=[os.path.basename(p)[:-4] for p in HELM_FIXTURES])
def test_helm_fmm_matches_reference(path):
    """Same tree, orders and truncation as the reference; the translations
    differ only by the reference's quadrature error, so the two FMMs agree
    far inside the iprec tolerance (a tenth of it is required)."""
    f = np.load(path, allow_pickle=False)
    g = lambda k: f[k] if k in f.files else None  # noqa: E731
    iprec = int(f["iprec"])
    ier, pot, fld, pt, ft = ob.orc_hfmm3dparttarg(iprec, complex(f["zk"]), f["source"], g("charge"),
                                                  g("dipstr"), g("dipvec"), g("target"))
    assert ier == 0
    tol = 0.1 * IPREC_EPS[iprec]
    assert rel(pot, f["pot"]) < tol
    assert rel(fld, f["fld"]) < tol
    if "target" in f.files:
        assert rel(pt, f["pottarg"]) < tol
        assert rel(ft, f["fldtarg"]) < tol


def test_helm_direct_matches_reference():
    f = load("helm_direct.npz")
    pot, fld, pt, ft = ob.orc_h3dpartdirect(complex(f["zk"]), f["source"], f["charge"], f["dipstr"],
                                            f["dipvec"], f["target"])
    for a, k in ((pot, "pot"), (fld, "fld"), (pt, "pottarg"), (ft, "fldtarg")):
        assert rel(a, f[k]) < 1e-13


def test_helm_orders_match_reference():
    o = _o()
    f = load("helm_orders.npz")
    for row, tab in zip(f["rows"], f["itables"]):
        size, zr, zi, eps, n = row
        ier = ctypes.c_int(0)
        assert o.orc_h3dterms(size, zc(complex(zr, zi)), eps, ctypes.byref(ier)) == int(n), row
        it = np.zeros(343, np.int32)
        o.orc_h3dterms_list2(size, zc(complex(zr, zi)), eps, _vp(it), ctypes.byref(ier))
        assert np.array_equal(it.reshape(7, 7, 7), tab), row


def test_helm_operators_match_reference():
    """Formation and evaluation are the same formulas as the reference's
    (roundoff only); the translations are the same operators computed in
    closed form instead of by quadrature on a sphere."""
    f = load("helm_operators.npz")
    p, zk = int(f["p"]), complex(f["zk"])
    sc1, sc2, sc3 = float(f["sc1"]), float(f["sc2"]), float(f["sc3"])
    src, q, ds, dv, trg = f["source"], f["charge"], f["dipstr"], f["dipvec"], f["target"]
    mp = form("mp", zk, sc1, src, q, f["c1"], p)
    assert rel(mp, f["p2m"]) < 1e-13
    assert rel(form("mp", zk, sc1, src, None, f["c1"], p, ds, dv), f["p2m_dp"]) < 1e-13
    lo = form("ta", zk, sc1, src, q, f["c0"], p)
    assert rel(lo, f["p2l"]) < 1e-13
    assert rel(form("ta", zk, sc1, src, None, f["c0"], p, ds, dv), f["p2l_dp"]) < 1e-13
    pot, fld = evaluate("mp", zk, sc1, f["c1"], f["p2m"], p, trg)
    assert rel(pot, f["m2p_pot"]) < 1e-13 and rel(fld, f["m2p_fld"]) < 1e-13
    pot, fld = evaluate("ta", zk, sc1, f["c0"], f["p2l"], p, trg)
    assert rel(pot, f["l2p_pot"]) < 1e-13 and rel(fld, f["l2p_fld"]) < 1e-13
    # Translations: the reference's quadrature leaves errors in the top
    # coefficients (up to 5% at order 9 for this M2L), invisible where the
    # expansion is evaluated.  Compare there, and require the oracle to be
    # at least as close as the reference to the exact answer.
    o = _o()
    far = trg
    m2m = _zeros(p)
    o.orc_h3dmpmp(zc(zk), sc1, _vp(f["c1"]), _vp(f["p2m"]), p, sc2, _vp(f["c2"]), _vp(m2m), p)
    a, _ = evaluate("mp", zk, sc2, f["c2"], m2m, p, far)
    b, _ = evaluate("mp", zk, sc2, f["c2"], f["m2m"], p, far)
    assert rel(a, b) < 1e-6
    assert rel(m2m, form("mp", zk, sc2, src, q, f["c2"], p)) < 1e-12      # exact for a truncated input
    d, fd = ob.orc_direct_sample(src, far, -np.ones(far.shape[0], np.int32), q, zk=zk)
    for pt, key in ((p, "m2l"), (6, "m2l_trunc6")):
        m2l = _zeros(p)
        o.orc_h3dmploc(zc(zk), sc1, _vp(f["c1"]), _vp(f["p2m"]), p, pt, sc1, _vp(f["c0"]), _vp(m2l), p)
        a, fa = evaluate("ta", zk, sc1, f["c0"], m2l, p, far)
        b, fb = evaluate("ta", zk, sc1, f["c0"], f[key], p, far)
        assert rel(a, b) < 1e-6 and rel(fa, fb) < 1e-6, key
        assert rel(a, d) < 1.2 * rel(b, d), key
    l2l = _zeros(p)
    o.orc_h3dlocloc(zc(zk), sc1, _vp(f["c0"]), _vp(f["p2l"]), p, sc3, _vp(f["c3"]), _vp(l2l), p)
    x = f["c3"] + 0.02 * np.array([[0.3, -0.2, 0.5], [0.1, 0.4, -0.3]])
    a, fa = evaluate("ta", zk, sc3, f["c3"], l2l, p, x)
    b, fb = evaluate("ta", zk, sc3, f["c3"], f["l2l"], p, x)
    assert rel(a, b) < 1e-12 and rel(fa, fb) < 1e-10


# ----------------------------------------------------------------------------
# closed forms and invariants
# ----------------------------------------------------------------------------
@pytest.mark.parametrize("z", [0.3, 2.5, 17.0, 1.0 + 0.7j, 40.0 + 2.0j, 1e-3])
def test_helm_bessel_functions_closed_form(z):
    o = _o()
    n, sc = 30, 0.7
    h = np.zeros(n + 1, np.complex128)
    j = np.zeros(n + 1, np.complex128)
    o.orc_h3dall(n, zc(z), sc, _vp(h))
    assert o.orc_jfuns3d(n, zc(z), sc, _vp(j)) == 0
    k = np.arange(n + 1)
    jr = sp.spherical_jn(k, z)
    hr = jr + 1j * sp.spherical_yn(k, z)
    ok_h = np.isfinite(hr) & (np.abs(hr * sc ** k) < 1e250)
    assert np.allclose(h[ok_h], (hr * sc ** k)[ok_h], rtol=1e-12, atol=0)
    ok_j = np.abs(jr) > 1e-280
    assert np.allclose(j[ok_j], (jr / sc ** k)[ok_j], rtol=1e-11, atol=0)


@pytest.mark.parametrize("zk", [5.0, 2.0 + 1.0j])
def test_helm_expansions_match_direct(rng, zk):
    """P2M then M2P and P2L then L2P (charges and dipoles) reproduce the
    direct sum; the fields also agree with -grad of the potential."""
    src = 0.5 + 0.1 * (rng.random((7, 3)) - 0.5)
    q, ds, dv = strengths(rng, 7, True, True, unit_dipvec=False)
    c = np.array([0.5, 0.5, 0.5])
    far = c + np.array([0.9, -0.7, 0.8]) + 0.05 * rng.random((4, 3))
    p = 30
    mp = form("mp", zk, 0.4, src, q, c, p, ds, dv)
    pot, fld = evaluate("mp", zk, 0.4, c, mp, p, far)
    pd, fd = ob.orc_direct_sample(src, far, -np.ones(4, np.int32), q, ds, dv, zk=zk)
    assert rel(pot, pd) < 1e-12 and rel(fld, fd) < 1e-12
    lc = far.mean(0)
    lo = form("ta", zk, 0.4, src, q, lc, p, ds, dv)
    near = lc + 0.05 * (rng.random((4, 3)) - 0.5)
    pot, fld = evaluate("ta", zk, 0.4, lc, lo, p, near)
    pd, fd = ob.orc_direct_sample(src, near, -np.ones(4, np.int32), q, ds, dv, zk=zk)
    assert rel(pot, pd) < 1e-12 and rel(fld, fd) < 1e-12


def test_helm_translations_are_consistent(rng):
    """M2M equals forming the multipole at the parent, L2L equals forming the
    local at the child (both exact up to the truncation of the input), M2L
    converges to forming the local directly."""
    o = _o()
    zk = 6.0 + 0.2j
    c1 = np.array([0.1, 0.2, 0.3])
    src = c1 + 0.05 * (rng.random((5, 3)) - 0.5)
    q = rng.normal(size=5) + 1j * rng.normal(size=5)
    p = 22
    mp = form("mp", zk, 0.5, src, q, c1, p)
    c2 = c1 + np.array([0.0625, -0.0625, 0.0625])
    m2m = _zeros(p)
    o.orc_h3dmpmp(zc(zk), 0.5, _vp(c1), _vp(mp), p, 0.9, _vp(c2), _vp(m2m), p)
    direct = form("mp", zk, 0.9, src, q, c2, p)
    k = 12 ** 2     # low orders are exact; the top ones feel the input truncation
    assert rel(m2m[:k], direct[:k]) < 1e-12
    c0 = c1 + np.array([0.5, 0.375, -0.375])
    for pt, tol in ((22, 1e-7), (30, 1e-9)):
        mpp = form("mp", zk, 0.5, src, q, c1, pt)
        loc = _zeros(pt)
        o.orc_h3dmploc(zc(zk), 0.5, _vp(c1), _vp(mpp), pt, pt, 0.5, _vp(c0), _vp(loc), pt)
        dl = form("ta", zk, 0.5, src, q, c0, pt)
        assert rel(loc[:64], dl[:64]) < tol
    lo = form("ta", zk, 0.5, src, q, c0, p)
    c3 = c0 + np.array([-0.03, 0.03, 0.03])
    l2l = _zeros(p)
    o.orc_h3dlocloc(zc(zk), 0.5, _vp(c0), _vp(lo), p, 0.3, _vp(c3), _vp(l2l), p)
    x = c3 + 0.02 * (rng.random((3, 3)) - 0.5)
    a, _ = evaluate("ta", zk, 0.3, c3, l2l, p, x)
    pd, _ = ob.orc_direct_sample(src, x, -np.ones(3, np.int32), q, zk=zk)
    assert rel(a, pd) < 1e-11


@pytest.mark.parametrize("iprec,dist,zk", [(1, "cube", 8.0), (2, "sphere", 12.0), (2, "clusters", 2.0 + 0.3j),
                                           (0, "plummer", 20.0), (3, "cube", 1.0)])
def test_helm_fmm_within_iprec_tolerance(rng, iprec, dist, zk):
    src = points(rng, 3000, dist)
    trg = points(rng, 800, dist)
    q, ds, dv = strengths(rng, 3000, True, True)
    ier, pot, fld, pt, ft = ob.orc_hfmm3dparttarg(iprec, zk, src, q, ds, dv, trg)
    assert ier == 0
    d = ob.orc_h3dpartdirect(zk, src, q, ds, dv, trg)
    eps = IPREC_EPS[iprec]
    assert rel(pot, d[0]) < eps and rel(pt, d[2]) < eps
    assert rel(fld, d[1]) < 10 * eps and rel(ft, d[3]) < 10 * eps


def test_helm_error_codes():
    src = np.random.default_rng(0).random((10, 3))
    q = np.ones(10, np.complex128)
    assert ob.orc_hfmm3dparttarg(9, 1.0, src, q)[0] == 32
