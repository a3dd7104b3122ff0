"""Oracle for quadrupole sources (oracle/fmm_oracle_qp.c) pinned against the
reference's own routines (laprouts3d.f:8691 lpotfld3dhess_qp, :7914
l3dformmp_qp_trunc, :8261 l3dformta_qp_trunc) where the reference library
builds, and independently by the mathematics: fld = -grad pot by central
differences, the expansions reproducing the direct sum when evaluated, and a
pure-trace quadrupole (the Laplacian of 1/r) producing nothing."""
import ctypes

import numpy as np
import pytest

import oracle_bindings as ob

needs_ref = pytest.mark.skipif(not ob.ref_available(), reason="reference library not built")


def _qp(rng, n):
    return (rng.normal(size=n) + 1j * rng.normal(size=n), rng.normal(size=(n, 6)))


@needs_ref
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_qp_direct_matches_reference(seed):
    rng = np.random.default_rng(seed)
    qs, qv = _qp(rng, 1)
    s = rng.uniform(-1, 1, 3)
    t = rng.uniform(-1, 1, 3)
    po, fo = ob.orc_direct_qp_sample(s[None], qs, qv, t[None], np.array([-1]))
    pr, fr, _ = ob.ref_lpotfld3dhess_qp(s, qs[0], qv[0], t)
    assert abs(po[0] - pr) <= 1e-13 * abs(pr)
    assert np.abs(fo[0] - fr).max() <= 1e-13 * np.abs(fr).max()


@needs_ref
@pytest.mark.parametrize("irregular", [False, True])
@pytest.mark.parametrize("p", [1, 2, 5, 12])
def test_qp_expansions_match_reference(irregular, p):
    rng = np.random.default_rng(10 + p)
    src = rng.uniform(-0.25, 0.25, (7, 3))
    qs, qv = _qp(rng, 7)
    c = np.array([2.0, -1.0, 1.5]) if irregular else np.array([0.02, 0.01, -0.03])
    a = ob.orc_l3dform_qp(irregular, 0.6, src, qs, qv, c, p)
    b = ob.ref_l3dform_qp(irregular, 0.6, src, qs, qv, c, p)
    assert np.abs(a - b).max() <= 1e-14 * np.abs(b).max()


def test_qp_field_is_minus_gradient():
    rng = np.random.default_rng(3)
    src = rng.uniform(-0.5, 0.5, (4, 3))
    qs, qv = _qp(rng, 4)
    x = np.array([1.3, -0.9, 1.1])
    h = 1e-5
    _, f = ob.orc_direct_qp_sample(src, qs, qv, x[None], np.array([-1]))
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        pp, _ = ob.orc_direct_qp_sample(src, qs, qv, (x + e)[None], np.array([-1]))
        pm, _ = ob.orc_direct_qp_sample(src, qs, qv, (x - e)[None], np.array([-1]))
        assert abs(-(pp[0] - pm[0]) / (2 * h) - f[0, k]) <= 1e-7 * np.abs(f).max()


def test_qp_potential_is_second_derivative_of_charge():
    """pot = quadstr sum_k quadvec_k d_k(1/r) with d_xx, d_xy ... taken at the
    target: second differences of 1/r."""
    s = np.array([0.1, -0.2, 0.3])
    x = np.array([0.9, 0.4, -0.6])
    h = 1e-4

    def g(y):
        return 1.0 / np.linalg.norm(y - s)

    pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    for k, (a, b) in enumerate(pairs):
        qv = np.zeros(6)
        qv[k] = 1.0
        ea, eb = np.eye(3)[a] * h, np.eye(3)[b] * h
        d2 = (g(x + ea + eb) - g(x + ea - eb) - g(x - ea + eb) + g(x - ea - eb)) / (4 * h * h)
        p, _ = ob.orc_direct_qp_sample(s[None], np.array([1.0 + 0j]), qv[None], x[None], np.array([-1]))
        assert abs(p[0] - d2) <= 1e-6 * abs(d2)


def test_qp_pure_trace_is_silent():
    rng = np.random.default_rng(4)
    src = rng.uniform(-0.2, 0.2, (3, 3))
    qs = np.ones(3, np.complex128)
    qv = np.tile([1.0, 1.0, 1.0, 0.0, 0.0, 0.0], (3, 1))
    p, f = ob.orc_direct_qp_sample(src, qs, qv, np.array([[1.0, 2.0, -1.0]]), np.array([-1]))
    assert abs(p[0]) < 1e-14 and np.abs(f).max() < 1e-14
    mp = ob.orc_l3dform_qp(False, 1.0, src, qs, qv, np.zeros(3), 8)
    loc = ob.orc_l3dform_qp(True, 1.0, src, qs, qv, np.array([3.0, 0.0, 0.0]), 8)
    assert np.abs(mp).max() < 1e-14 and np.abs(loc).max() < 1e-14


def _eval(irregular, rscale, center, coef, p, pts):
    o = ob.oracle()
    c = np.ascontiguousarray(center, np.float64)
    pts = np.ascontiguousarray(pts, np.float64)
    pot = np.zeros(pts.shape[0], np.complex128)
    fld = np.zeros((pts.shape[0], 3), np.complex128)
    fn = o.orc_l3dtaeval if irregular else o.orc_l3dmpeval
    fn(ctypes.c_double(rscale), ob._vp(c), ob._vp(coef), int(p), ob._vp(pts), pts.shape[0], 1, ob._vp(pot), 1,
       ob._vp(fld))
    return pot, fld


@pytest.mark.parametrize("irregular", [False, True])
def test_qp_expansions_reproduce_direct_sum(irregular):
    """Expansions evaluated by the charge evaluators (l3dmpeval / l3dtaeval)
    converge to the direct sum geometrically in p."""
    rng = np.random.default_rng(5)
    src = rng.uniform(-0.2, 0.2, (5, 3))
    qs, qv = _qp(rng, 5)
    if irregular:
        center = np.array([1.0, 0.8, -0.9])
        pts = center + rng.uniform(-0.2, 0.2, (6, 3))
    else:
        center = np.zeros(3)
        pts = rng.uniform(-1, 1, (6, 3))
        pts = 1.3 * pts / np.linalg.norm(pts, axis=1)[:, None]
    pd, fd = ob.orc_direct_qp_sample(src, qs, qv, pts, -np.ones(6, np.int32))
    errs = []
    for p in (6, 12, 24):
        coef = ob.orc_l3dform_qp(irregular, 0.5, src, qs, qv, center, p)
        pe, fe = _eval(irregular, 0.5, center, coef, p, pts)
        errs.append(max(np.abs(pe - pd).max() / np.abs(pd).max(), np.abs(fe - fd).max() / np.abs(fd).max()))
    assert errs[2] < 1e-9 and errs[1] < errs[0] * 0.1
