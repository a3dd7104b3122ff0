"""Helmholtz triangle FMM (fmm3d_hfmm3dtriatarg / fmm3d_h3dtriadirect) on
the GPU against the oracle's direct sum (which the CPU tests pin to the
reference's own routines): the device direct sum to roundoff for every rule
degree and density flag, the FMM within the iprec tolerance
(hfmm3dtria.f:262-270) for the potential and one digit less for the field,
on a triangulated sphere (uniform octree) and on a soup of small triangles
(adaptive octree, lists 3 and 4), with charges, dipoles, both, and the
difference kernel (flag 2)."""
import numpy as np
import pytest

import fmmlib3d_amd as F
import oracle_bindings as ob
from fmm_inputs import IPREC_EPS, icosphere, rel

pytestmark = pytest.mark.gpu


def _densities(rng, n, ifc, ifd):
    q = rng.normal(size=n) + 1j * rng.normal(size=n) if ifc else None
    mu = rng.normal(size=n) + 1j * rng.normal(size=n) if ifd else None
    return q, mu


def _targets(rng, m, radius=1.0):
    x = rng.normal(size=(m, 3))
    x /= np.linalg.norm(x, axis=1)[:, None]
    return x * rng.uniform(0.9 * radius, 1.1 * radius, m)[:, None]


@pytest.mark.parametrize("nq", [1, 2, 4, 5, 6])
@pytest.mark.parametrize("fc,fd", [(1, 1), (1, 0), (0, 1), (2, 2)])
def test_gpu_htria_direct_equals_oracle(nq, fc, fd):
    rng = np.random.default_rng(91)
    T = icosphere(3)
    n = T.shape[0]
    q, mu = _densities(rng, n, fc, fd)
    trg = _targets(rng, 300)
    zk = 2.5 + 0.2j
    pot, fld, pt, ft = F.h3dtriadirect(nq, zk, T, q, mu, trg, ifcharge=fc or 1, ifdipole=fd or 1)
    cen = T.mean(axis=1)
    po, fo = ob.orc_htria_direct_sample(T, zk, nq, cen, np.arange(n, dtype=np.int32), q, mu, fc, fd)
    pto, fto = ob.orc_htria_direct_sample(T, zk, nq, trg, None, q, mu, fc, fd)
    assert rel(pot, po) < 1e-12 and rel(fld, fo) < 1e-12
    assert rel(pt, pto) < 1e-12 and rel(ft, fto) < 1e-12


def _check(T, zk, iprec, q, mu, trg, out, rng, fc=1, fd=1, tag=""):
    ier, pot, fld, pt, ft = out
    assert ier == 0
    n = T.shape[0]
    nq = 2 if iprec < 0 else 4 if iprec == 0 else 6
    idx = rng.choice(n, 300, replace=False)
    it = rng.choice(len(trg), min(300, len(trg)), replace=False)
    cen = T.mean(axis=1)
    po, fo = ob.orc_htria_direct_sample(T, zk, nq, cen[idx], idx.astype(np.int32), q, mu, fc, fd)
    pto, fto = ob.orc_htria_direct_sample(T, zk, nq, trg[it], None, q, mu, fc, fd)
    e = [rel(pot[idx], po), rel(fld[idx], fo), rel(pt[it], pto), rel(ft[it], fto)]
    print("%s iprec %d: pot %.2e fld %.2e pottarg %.2e fldtarg %.2e" % (tag, iprec, *e))
    return e


@pytest.mark.parametrize("iprec", [0, 1, 2, 3])
@pytest.mark.parametrize("ifc,ifd", [(1, 1), (1, 0), (0, 1)])
def test_gpu_htria_fmm_sphere(iprec, ifc, ifd):
    rng = np.random.default_rng(92 + iprec)
    T = icosphere(5)                                  # 20,480 triangles
    n = T.shape[0]
    q, mu = _densities(rng, n, ifc, ifd)
    trg = _targets(rng, 2000)
    zk = 3.0 + 0.1j
    out = F.hfmm3dtriatarg(iprec, zk, T, q, mu, trg)
    e = _check(T, zk, iprec, q, mu, trg, out, rng, ifc, ifd, "sphere")
    bar = IPREC_EPS[iprec]
    # iprec 3: a leaf of 80 points (triangles and targets) is a few triangles
    # wide, so triangles reach well outside the box their expansion is
    # centred on; the potential at the centroids then comes to 5.6e-10
    # (the reference's own FMM gives 3.0e-8 at iprec 3 on a sphere with
    # targets, hfmm3dtriatarg).  Held to twice the tolerance there.
    pbar = 2 * bar if iprec == 3 else bar
    assert e[0] < pbar and e[2] < bar
    assert e[1] < 10 * bar and e[3] < 10 * bar


@pytest.mark.parametrize("fc,fd", [(2, 2), (2, 1), (1, 2)])
def test_gpu_htria_fmm_difference_kernel(fc, fd):
    """Flag 2: the near field with e^{ikr}/r - 1/r, the far field as the
    Helmholtz far field minus the Laplace one.  The difference kernel's
    field is O(k^2 h) smaller than either kernel's while the two far fields
    carry their own truncation errors, so relative to it the errors are
    larger: held to 5x / 50x the tolerance (the reference's own FMM gives
    2.6e-6 and 1.1e-5 for the field at centroids and targets on a sphere
    with both flags 2 at iprec 2)."""
    rng = np.random.default_rng(96)
    T = icosphere(5)
    n = T.shape[0]
    q, mu = _densities(rng, n, 1, 1)
    trg = _targets(rng, 2000)
    zk = 2.0
    out = F.hfmm3dtriatarg(2, zk, T, q, mu, trg, ifcharge=fc, ifdipole=fd)
    e = _check(T, zk, 2, q, mu, trg, out, rng, fc, fd, "flag2 %d%d" % (fc, fd))
    assert e[0] < 5 * IPREC_EPS[2] and e[2] < 5 * IPREC_EPS[2]
    assert e[1] < 50 * IPREC_EPS[2] and e[3] < 50 * IPREC_EPS[2]


@pytest.mark.parametrize("iprec", [1, 2])
def test_gpu_htria_fmm_soup_adaptive(iprec):
    """Clustered small triangles: leaves on many levels (lists 3 and 4)."""
    rng = np.random.default_rng(97 + iprec)
    c = rng.random((6, 3)) * 0.8 + 0.1
    T = c[rng.integers(0, 6, 30000)][:, None, :] + 0.03 * rng.normal(size=(30000, 1, 3))
    T = T + 0.0005 * rng.normal(size=(30000, 3, 3))
    n = T.shape[0]
    q, mu = _densities(rng, n, 1, 1)
    trg = rng.random((3000, 3))
    zk = 4.0
    out = F.hfmm3dtriatarg(iprec, zk, T, q, mu, trg)
    tree = F.d3tstrcr(T.mean(axis=1), trg, {1: 20, 2: 40}[iprec])
    assert any(len(l) for l in tree[4][2])      # list 3 is not empty
    e = _check(T, zk, iprec, q, mu, trg, out, rng, 1, 1, "soup")
    bar = IPREC_EPS[iprec]
    assert e[0] < bar and e[2] < bar
    assert e[1] < 10 * bar and e[3] < 10 * bar


def test_gpu_htria_self_one_leaf_and_entries():
    """iprec 6 puts everything in one leaf: the FMM is the direct sum with
    the degree-6 rule; the self and targets-only entries agree with it."""
    rng = np.random.default_rng(99)
    T = icosphere(2)
    n = T.shape[0]
    q, mu = _densities(rng, n, 1, 1)
    trg = _targets(rng, 100)
    zk = 1.5 + 0.3j
    ier, pot, fld, pt, ft = F.hfmm3dtriatarg(6, zk, T, q, mu, trg)
    assert ier == 0
    pd, fd, ptd, ftd = F.h3dtriadirect(6, zk, T, q, mu, trg)
    assert rel(pot, pd) < 1e-12 and rel(fld, fd) < 1e-12
    assert rel(pt, ptd) < 1e-12 and rel(ft, ftd) < 1e-12
    ier2, pot2, fld2 = F.hfmm3dtriaself(6, zk, T, q, mu)
    assert ier2 == 0 and np.array_equal(pot2, pot) and np.array_equal(fld2, fld)
    ier3, pt3, ft3 = F.hfmm3dtriampftarg(6, zk, T, q, mu, trg)
    assert ier3 == 0 and rel(pt3, ptd) < 1e-12 and rel(ft3, ftd) < 1e-12


def test_gpu_htria_invalid_flag():
    T = icosphere(1)
    q = np.ones(T.shape[0], np.complex128)
    ier, *_ = F.hfmm3dtriatarg(1, 1.0, T, q, None, None, ifcharge=3)
    assert ier == 128


@pytest.mark.skipif(not ob.ref_available(), reason="reference library not built here")
def test_gpu_htria_fmm_vs_reference_fmm():
    """Both FMMs against each other on the same sphere (iprec 2): within the
    sum of their tolerances."""
    rng = np.random.default_rng(100)
    T = icosphere(4)
    n = T.shape[0]
    q, mu = _densities(rng, n, 1, 1)
    trg = _targets(rng, 500)
    zk = 2.0
    ier, pot, fld, pt, ft = F.hfmm3dtriatarg(2, zk, T, q, mu, trg)
    rier, rpot, rfld, rpt, rft = ob.ref_hfmm3dtriatarg(2, zk, T, T.mean(axis=1), q, mu, trg)
    assert ier == 0 and rier == 0
    e = [rel(pot, rpot), rel(fld, rfld), rel(pt, rpt), rel(ft, rft)]
    print("vs reference FMM: pot %.2e fld %.2e pottarg %.2e fldtarg %.2e" % tuple(e))
    assert e[0] < 2 * IPREC_EPS[2] and e[2] < 2 * IPREC_EPS[2]
    assert e[1] < 20 * IPREC_EPS[2] and e[3] < 20 * IPREC_EPS[2]


@pytest.mark.parametrize("iprec", [4, 5])
def test_gpu_htria_fmm_high_iprec(iprec):
    """iprec 4 and 5: within the tolerance at iprec 4; at iprec 5 the FP64
    rounding of the sums, about 4e-14 here, is the floor."""
    rng = np.random.default_rng(101 + iprec)
    T = icosphere(4)
    n = T.shape[0]
    q, mu = _densities(rng, n, 1, 1)
    trg = _targets(rng, 1000)
    zk = 2.0 + 0.1j
    out = F.hfmm3dtriatarg(iprec, zk, T, q, mu, trg)
    e = _check(T, zk, iprec, q, mu, trg, out, rng, 1, 1, "sphere")
    assert max(e) < max(IPREC_EPS[iprec], 2e-13)


@pytest.mark.parametrize("m", [1, 2])
def test_gpu_htria_fmm_one_or_two_triangles(m):
    rng = np.random.default_rng(105 + m)
    T = icosphere(1)[:m]
    q, mu = _densities(rng, m, 1, 1)
    trg = _targets(rng, 50)
    zk = 3.0
    ier, pot, fld, pt, ft = F.hfmm3dtriatarg(2, zk, T, q, mu, trg)
    pd, fd, ptd, ftd = F.h3dtriadirect(6, zk, T, q, mu, trg)
    assert ier == 0
    for a, b in zip((pot, fld, pt, ft), (pd, fd, ptd, ftd)):
        assert rel(a, b) < 1e-14


@pytest.mark.parametrize("zk", [None, 1.0 + 0.1j])
def test_gpu_tria_fmms_green_identity(zk):
    """The reference drivers' own check (examples/hfmm3dtria_driver.f and
    lfmm3dtria_driver.f: iprec 1, a point source at (0, 0.2, 0) inside the
    sphere, densities from its potential and normal derivative) through the
    FMMs on a 20,480-triangle sphere: pot = u/2 and fld . n / (2 pi) = the
    single-layer density to discretisation accuracy (measured on the
    oracle: 1.2e-3 and 2.3e-2 at 5,120 triangles, falling 4x and 2x per
    refinement)."""
    from test_oracle_htria import _green_data, _green_errors
    T, cen, nrm, u, charge, dipstr = _green_data(5, zk)
    trg = cen - np.array([2.0, 0.0, 0.0])           # the drivers' targets, unused by the check
    if zk is None:
        ier, pot, fld, pt, ft = F.lfmm3dtriatarg(1, T, charge, dipstr, trg)
    else:
        ier, pot, fld, pt, ft = F.hfmm3dtriatarg(1, zk, T, charge, dipstr, trg)
    assert ier == 0
    e_pot, e_dn = _green_errors(pot, fld, nrm, u, charge)
    print("green identity zk %s: ext %.2e dudn %.2e" % (zk, e_pot, e_dn))
    assert e_pot < 6e-4 and e_dn < 1.5e-2
