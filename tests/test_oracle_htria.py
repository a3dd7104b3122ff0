"""The oracle's Helmholtz flat-triangle layer potentials
(oracle/fmm_oracle_htria.c, restating direct3dtritarghelms3 /
direct3dtritarghelmd3 / direct3dtrihelms2 / direct3dtrihelmd2,
src/triquadflatlib.f:819-1358 with triquadhelm2 / triquadhelmd2 /
triquadselfhelm / triquadselfhelmd, and h3dtriadirect, src/hfmm3dtria.f:1456)
pinned against the reference's own routines built from source (oracle/_ref),
its symmetric rules (triasymq) against the reference's and against the
moment equations, the k -> 0 limit against the Laplace closed forms, a point
far from the triangle against high-order product Gauss quadrature of
e^{ikr}/r, and the field against central differences of the potential.
CPU only."""
from math import factorial

import numpy as np
import pytest

import oracle_bindings as ob
from fmm_inputs import icosphere, rel


def _close(a, b, tol):
    a, b = np.atleast_1d(np.asarray(a)), np.atleast_1d(np.asarray(b))
    return np.max(np.abs(a - b)) <= tol * max(1.0, np.max(np.abs(b)))


@pytest.mark.parametrize("degree,nodes", [(1, 1), (2, 3), (3, 6), (4, 6), (5, 7), (6, 12)])
def test_symrule_moments(degree, nodes):
    """Each rule integrates every monomial x^i y^j, i + j <= degree, exactly
    on the reference triangle, has the reference's node count, positive
    weights and nodes inside the triangle."""
    xi, eta, w = ob.orc_tri_symrule(degree)
    assert len(w) == nodes
    assert np.all(w > 0) and np.all(xi > 0) and np.all(eta > 0) and np.all(xi + eta < 1)
    for i in range(degree + 1):
        for j in range(degree + 1 - i):
            exact = factorial(i) * factorial(j) / factorial(i + j + 2)
            assert abs(np.sum(w * xi ** i * eta ** j) - exact) < 1e-15, (i, j)
    assert len(ob.orc_tri_symrule(7)[2]) == 0


@pytest.mark.skipif(not ob.ref_available(), reason="reference library not built here")
@pytest.mark.parametrize("degree", [1, 2, 3, 4, 5, 6])
def test_symrule_matches_reference_triasymq(degree):
    a = ob.orc_tri_symrule(degree)
    b = ob.ref_triasymq(degree)
    assert len(a[2]) == len(b[2])
    key = lambda r: np.lexsort((np.round(r[1], 9), np.round(r[0], 9)))
    ia, ib = key(a), key(b)
    for u, v in zip(a, b):
        assert np.max(np.abs(u[ia] - v[ib])) < 1e-14


def _cases(rng, n):
    for it in range(n):
        tri = rng.normal(size=(3, 3))
        zk = complex(rng.uniform(0.2, 4.0), rng.uniform(0.0, 0.5) if it % 2 else 0.0)
        deg = [2, 4, 6, 5][it % 4]
        kind = it % 5
        if kind == 0:
            x, self = tri.mean(0), True
        elif kind == 1:
            x, self = tri[0] + 0.3 * (tri[1] - tri[0]) + 1.3 * (tri[2] - tri[0]), False   # in the plane, outside
        elif kind == 2:
            x, self = tri.mean(0) + 1e-3 * rng.normal(size=3), False                   # very close
        elif kind == 3:
            x, self = tri.mean(0) + rng.normal(size=3), False
        else:
            x, self = tri.mean(0) + 1e-3 * rng.normal(size=3), False
            zk = 1e-4 * zk                                                           # Taylor branch
        yield tri, x, zk, deg, self


@pytest.mark.skipif(not ob.ref_available(), reason="reference library not built here")
def test_oracle_htri_matches_reference_routines():
    rng = np.random.default_rng(81)
    for tri, x, zk, deg, self in _cases(rng, 200):
        for ifl in (0, 1):
            a = ob.orc_htri_unit(tri, x, zk, deg, self, lap_c=not ifl, lap_d=not ifl)
            b = ob.ref_htri_unit(tri, x, zk, deg, self, ifl, ifl)
            for u, v in zip(a, b):
                assert _close(u, v, 1e-12), (tri, x, zk, deg, self, ifl, u, v)


def test_oracle_htri_small_k_is_laplace():
    """As k -> 0 the Helmholtz integrals tend to the Laplace closed forms
    (the difference part is O(k))."""
    rng = np.random.default_rng(82)
    for it in range(40):
        tri = rng.normal(size=(3, 3))
        self = it % 4 == 0
        x = tri.mean(0) if self else tri.mean(0) + rng.normal(size=3)
        a = ob.orc_htri_unit(tri, x, 1e-9, 6, self)
        b = ob.orc_tri_unit(tri, x, self)
        for u, v in zip(a, b):
            assert _close(u, v, 1e-8)
        d = ob.orc_htri_unit(tri, x, 1e-9, 6, self, lap_c=False, lap_d=False)   # the difference kernel
        for u in d:
            assert np.max(np.abs(u)) < 1e-7


def _gauss_tri_helm(tri, x, zk, q=60):
    """Collapsed-square Gauss quadrature of e^{ikr}/r, its gradient, the
    dipole-layer kernel n.grad_y and its gradient at x (far from tri)."""
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
    e = np.exp(1j * zk * r)
    G = e / r
    dG = (1j * zk * r - 1) * e / r ** 3                  # dG/dr / r
    slp = np.sum(wt * G)
    slpf = -np.sum(wt[:, None] * d * dG[:, None], axis=0)
    nd = d @ n
    dlp = -np.sum(wt * nd * dG)                         # n . grad_y G = -n . grad_x G
    # d2G: second derivative factor, grad_x (n.d dG) = n dG + n.d d (dG'/r)
    ddG = ((-zk ** 2 * r ** 2 - 3j * zk * r + 3) * e / r ** 5)
    dlpf = np.sum(wt[:, None] * (n[None, :] * dG[:, None] + (nd * ddG)[:, None] * d), axis=0)
    return slp, slpf, dlp, dlpf


def test_oracle_htri_far_matches_quadrature():
    """At a point many triangle sizes away the symmetric rule on the smooth
    part and the closed forms together equal a high-order quadrature of the
    full kernel."""
    rng = np.random.default_rng(83)
    for it in range(30):
        tri = 0.1 * rng.normal(size=(3, 3))
        x = tri.mean(0) + rng.normal(size=3) * 1.5
        zk = complex(rng.uniform(0.5, 3.0), 0.1 * (it % 2))
        a = ob.orc_htri_unit(tri, x, zk, 6)
        b = _gauss_tri_helm(tri, x, zk)
        for u, v in zip(a, b):
            assert _close(u, v, 1e-9), (it, u, v)


def test_oracle_htri_field_is_minus_gradient():
    rng = np.random.default_rng(84)
    h = 1e-5
    for it in range(30):
        tri = rng.normal(size=(3, 3))
        x = tri.mean(0) + rng.normal(size=3)
        zk = [2.0, 3.0 + 0.4j, 0.004][it % 3]                 # the last one on the Taylor branch
        a = ob.orc_htri_unit(tri, x, zk, 6)
        for k in range(3):
            e = np.zeros(3)
            e[k] = h
            p, m = ob.orc_htri_unit(tri, x + e, zk, 6), ob.orc_htri_unit(tri, x - e, zk, 6)
            assert abs(-(p[0] - m[0]) / (2 * h) - a[1][k]) < 1e-6 * max(1, abs(a[1][k]))
            assert abs(-(p[2] - m[2]) / (2 * h) - a[3][k]) < 1e-6 * max(1, abs(a[3][k]))


@pytest.mark.skipif(not ob.ref_available(), reason="reference library not built here")
@pytest.mark.parametrize("zk,nq,fc,fd", [(1.3, 6, 1, 1), (2 + 0.3j, 4, 1, 1), (0.7, 2, 2, 2), (1.1, 6, 2, 1),
                                         (5.0, 5, 1, 2), (1.0, 1, 1, 1)])
def test_oracle_htria_direct_matches_reference(zk, nq, fc, fd):
    rng = np.random.default_rng(85)
    T = icosphere(2)
    n = T.shape[0]
    cen = T.mean(axis=1)
    q = rng.normal(size=n) + 1j * rng.normal(size=n)
    mu = rng.normal(size=n) + 1j * rng.normal(size=n)
    trg = rng.normal(size=(40, 3)) * 0.8
    rp, rf, rpt, rft = ob.ref_h3dtriadirect(nq, zk, T, cen, q, mu, trg, fc, fd)
    po, fo = ob.orc_htria_direct_sample(T, zk, nq, cen, np.arange(n, dtype=np.int32), q, mu, fc, fd)
    pt, ft = ob.orc_htria_direct_sample(T, zk, nq, trg, None, q, mu, fc, fd)
    assert rel(po, rp) < 1e-13 and rel(fo, rf) < 1e-13
    assert rel(pt, rpt) < 1e-13 and rel(ft, rft) < 1e-13


def _green_data(level, zk, x0=(0.0, 0.2, 0.0)):
    """The reference drivers' check (examples/hfmm3dtria_driver.f, and its
    Laplace twin): u = G(x - x0) for x0 inside a triangulated unit sphere,
    single-layer density du/dn / (4 pi) (as fld . n), dipole-layer density
    u / (4 pi); Green's identity gives pot = u/2 on the surface and
    fld . n / (2 pi) = the single-layer density, to discretisation
    accuracy."""
    T = icosphere(level)
    cen = T.mean(axis=1)
    nrm = np.cross(T[:, 1] - T[:, 0], T[:, 2] - T[:, 0])
    nrm /= np.linalg.norm(nrm, axis=1)[:, None]
    d = cen - np.asarray(x0)
    r = np.linalg.norm(d, axis=1)
    if zk is None:
        u, grad = 1 / r, -d / r[:, None] ** 3
    else:
        e = np.exp(1j * zk * r)
        u, grad = e / r, ((1j * zk * r - 1) * e / r ** 3)[:, None] * d
    charge = np.sum(-grad * nrm, 1) / (4 * np.pi)
    dipstr = u / (4 * np.pi) + 0j
    return T, cen, nrm, u, charge, dipstr


def _green_errors(pot, fld, nrm, u, charge):
    e_pot = np.linalg.norm(pot - u / 2) / np.linalg.norm(u / 2)
    e_dn = np.linalg.norm(np.sum(fld * nrm, 1) / (2 * np.pi) - charge) / np.linalg.norm(charge)
    return e_pot, e_dn


@pytest.mark.parametrize("zk", [None, 1.0 + 0.1j])
def test_oracle_tria_green_identity_converges(zk):
    """Green's identity on the sphere with the oracle's direct sums: the
    errors are discretisation errors and fall with refinement (about 4x for
    the potential, 2x for the normal field per level); a wrong sign or a
    missing self term leaves O(1) errors."""
    errs = []
    for level in (2, 3):
        T, cen, nrm, u, charge, dipstr = _green_data(level, zk)
        idx = np.arange(T.shape[0], dtype=np.int32)
        if zk is None:
            pot, fld = ob.orc_tria_direct_sample(T, cen, idx, charge, dipstr)
        else:
            pot, fld = ob.orc_htria_direct_sample(T, zk, 6, cen, idx, charge, dipstr)
        errs.append(_green_errors(pot, fld, nrm, u, charge))
    (p2, n2), (p3, n3) = errs
    assert p3 < 6e-3 and p3 < p2 / 3
    assert n3 < 6e-2 and n3 < 0.7 * n2
