"""The oracle's flat-triangle layer potentials (oracle/fmm_oracle_tria.c,
restating the reference's analytic quadratures direct3dtritarglaps2 /
direct3dtritarglapd2 / direct3dtrilaps2 / direct3dtrilapd2,
src/triquadflatlib.f:147-818, and l3dtriadirect, src/lfmm3dtria.f:1385)
pinned four ways: against the reference's own routines built from source
(oracle/_ref), against Gauss quadrature of the integrands off the plane,
against the jump relations of the single- and dipole-layer potentials
across the triangle, and by the field being minus the gradient of the
potential (central differences).  CPU only."""
import numpy as np
import pytest

import oracle_bindings as ob
from fmm_inputs import icosphere, rel


def _close(a, b, tol):
    a, b = np.atleast_1d(np.asarray(a, float)), np.atleast_1d(np.asarray(b, float))
    return np.max(np.abs(a - b)) <= tol * max(1.0, np.max(np.abs(b)))


@pytest.mark.skipif(not ob.ref_available(), reason="reference library not built here")
def test_oracle_tri_matches_reference_routines():
    rng = np.random.default_rng(51)
    for it in range(300):
        tri = rng.normal(size=(3, 3))
        c = tri.mean(0)
        if it % 3 == 0:
            x = c + 0.5 * rng.normal(size=3)            # near
        elif it % 3 == 1:
            x = 3.0 * rng.normal(size=3)                # anywhere
        else:
            x = tri[0] + 1.7 * (tri[1] - tri[0]) + 0.3 * (tri[2] - tri[0])   # in the plane, outside
        a = ob.orc_tri_unit(tri, x)
        b = ob.ref_tri_unit(tri, x)
        for u, v in zip(a, b):
            assert _close(u, v, 1e-12), (it, u, v)
    for it in range(100):
        tri = rng.normal(size=(3, 3))
        a = ob.orc_tri_unit(tri, tri.mean(0), True)
        b = ob.ref_tri_unit(tri, tri.mean(0), True)
        for u, v in zip(a, b):
            assert _close(u, v, 1e-12), (it, u, v)


def _gauss_tri(tri, x, q=60):
    """Collapsed-square Gauss quadrature of the four integrands at x."""
    g, w = np.polynomial.legendre.leggauss(q)
    u, wu = (g + 1) / 2, w / 2
    U, V = np.meshgrid(u, u, indexing="ij")
    W = np.outer(wu, wu) * U
    xi, eta = (U * (1 - V)).ravel(), (U * V).ravel()
    e1, e2 = tri[1] - tri[0], tri[2] - tri[0]
    n = np.cross(e1, e2)
    area2 = np.linalg.norm(n)
    n = n / area2
    y = tri[0] + xi[:, None] * e1 + eta[:, None] * e2
    wt = W.ravel() * area2
    d = x - y
    r = np.linalg.norm(d, axis=1)
    slp = np.sum(wt / r)
    slpf = np.sum(wt[:, None] * d / r[:, None] ** 3, axis=0)
    nd = d @ n
    dlp = np.sum(wt * nd / r ** 3)
    dlpf = np.sum(wt[:, None] * (3 * nd[:, None] * d / r[:, None] ** 5 - n / r[:, None] ** 3), axis=0)
    return slp, slpf, dlp, dlpf


def test_oracle_tri_matches_quadrature_off_plane():
    rng = np.random.default_rng(52)
    for it in range(40):
        tri = rng.normal(size=(3, 3))
        e = tri.max(0) - tri.min(0)
        x = tri.mean(0) + 2.0 * np.max(e) * rng.normal(size=3)   # a few triangle sizes away
        a = ob.orc_tri_unit(tri, x)
        b = _gauss_tri(tri, x)
        for u, v in zip(a, b):
            assert _close(u, v, 1e-10), (it, u, v)


def test_oracle_tri_jump_relations():
    """Across the triangle at an interior point: the single layer is
    continuous and its normal field jumps by 4 pi; the dipole-layer
    potential jumps by 4 pi (Green's function 1/r)."""
    rng = np.random.default_rng(53)
    for it in range(30):
        tri = rng.normal(size=(3, 3))
        n = np.cross(tri[1] - tri[0], tri[2] - tri[0])
        n /= np.linalg.norm(n)
        b = rng.dirichlet([1, 1, 1])
        x0 = b @ tri
        eps = 1e-11
        up = ob.orc_tri_unit(tri, x0 + eps * n)
        dn = ob.orc_tri_unit(tri, x0 - eps * n)
        assert abs(up[0] - dn[0]) < 1e-6
        assert abs((up[1] - dn[1]) @ n - 4 * np.pi) < 1e-6
        assert abs(up[2] - dn[2] - 4 * np.pi) < 1e-6


def test_oracle_tri_field_is_minus_gradient():
    rng = np.random.default_rng(54)
    h = 1e-5
    for it in range(30):
        tri = rng.normal(size=(3, 3))
        x = tri.mean(0) + rng.normal(size=3)
        a = ob.orc_tri_unit(tri, x)
        for k in range(3):
            e = np.zeros(3)
            e[k] = h
            p, m = ob.orc_tri_unit(tri, x + e), ob.orc_tri_unit(tri, x - e)
            assert abs(-(p[0] - m[0]) / (2 * h) - a[1][k]) < 1e-6 * max(1, abs(a[1][k]))
            assert abs(-(p[2] - m[2]) / (2 * h) - a[3][k]) < 1e-6 * max(1, abs(a[3][k]))


@pytest.mark.skipif(not ob.ref_available(), reason="reference library not built here")
def test_oracle_tria_direct_matches_reference():
    rng = np.random.default_rng(55)
    T = icosphere(2)
    n = T.shape[0]
    cen = T.mean(axis=1)
    q = rng.normal(size=n) + 1j * rng.normal(size=n)
    mu = rng.normal(size=n) + 1j * rng.normal(size=n)
    trg = rng.normal(size=(50, 3)) * 0.7
    self = np.arange(n, dtype=np.int32)
    po, fo = ob.orc_tria_direct_sample(T, cen, self, q, mu)
    pt, ft = ob.orc_tria_direct_sample(T, trg, None, q, mu)
    rp, rf, rpt, rft = ob.ref_l3dtriadirect(T, cen, q, mu, trg)
    assert rel(po, rp) < 1e-13 and rel(fo, rf) < 1e-13
    assert rel(pt, rpt) < 1e-13 and rel(ft, rft) < 1e-13
