"""GPU parity of the Helmholtz particle FMM (hfmm3dparttarg and friends
through the C ABI).  Bars: against the CPU oracle, which applies the same
operators in the same order (same tree, orders, list-2 truncation and
closed-form coaxial shifts), 1e-10 relative L2; against the reference's
golden fixtures (whose translations use quadrature on a sphere) a tenth of
the iprec tolerance; against the FP64 direct sum the iprec tolerance."""
import glob
import os

import numpy as np
import pytest

import fmmlib3d_amd as F
import oracle_bindings as ob
from fmm_inputs import IPREC_EPS, points, rel, strengths

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
HELM_FIXTURES = sorted(glob.glob(os.path.join(GOLD, "helm_targ_*.npz")) + glob.glob(os.path.join(GOLD, "helm_self_*.npz")))
ORACLE_TOL = 1e-10


def _check(a, b, tol):
    if b is None:
        assert a is None
        return
    nb = np.linalg.norm(b)
    if nb == 0:
        assert np.linalg.norm(a) == 0
    else:
        assert rel(a, b) < tol, rel(a, b)


@pytest.mark.parametrize("path", HELM_FIXTURES, ids=[os.path.basename(p)[:-4] for p in HELM_FIXTURES])
def test_gpu_helm_matches_reference_fixture(path):
    f = np.load(path, allow_pickle=False)
    g = lambda k: f[k] if k in f.files else None  # noqa: E731
    iprec = int(f["iprec"])
    ier, pot, fld, pt, ft = F.hfmm3dparttarg(iprec, complex(f["zk"]), f["source"], g("charge"), g("dipstr"),
                                             g("dipvec"), g("target"))
    assert ier == 0
    tol = 0.1 * IPREC_EPS[iprec]
    _check(pot, f["pot"], tol)
    _check(fld, f["fld"], tol)
    if "target" in f.files:
        _check(pt, f["pottarg"], tol)
        _check(ft, f["fldtarg"], tol)


def test_gpu_helm_direct_matches_reference_fixture():
    f = np.load(os.path.join(GOLD, "helm_direct.npz"), allow_pickle=False)
    pot, fld, pt, ft = F.h3dpartdirect(complex(f["zk"]), f["source"], f["charge"], f["dipstr"], f["dipvec"],
                                       f["target"])
    for a, k in ((pot, "pot"), (fld, "fld"), (pt, "pottarg"), (ft, "fldtarg")):
        _check(a, f[k], 1e-13)


CASES = [
    # ns, nt, iprec, dist, zk, ifcharge, ifdipole
    (5000, 2000, 1, "cube", 8.0, 1, 1),
    (6000, 0, 2, "sphere", 20.0, 1, 0),
    (5000, 1500, 2, "clusters", 2.0 + 0.5j, 1, 1),
    (4000, 2000, 0, "plummer", 25.0, 1, 1),
    (3000, 1000, 3, "cube", 1.5, 0, 1),
    (3000, 800, -1, "shell", 12.0, 1, 0),
    (2000, 500, -2, "plane", 4.0, 1, 1),
    (1, 3, 1, "cube", 3.0, 1, 1),        # a single source
    (150, 40, 1, "cube", 5.0, 1, 1),     # N <= nbox: root leaf, no far field
    (300, 100, 6, "cube", 2.0, 1, 1),    # iprec = 6: the FMM is a direct sum
    (8000, 2000, 2, "cube", 0.05, 1, 1), # low frequency: rscale < 1 on every level
]


@pytest.mark.parametrize("ns,nt,iprec,dist,zk,ifc,ifd", CASES)
def test_gpu_helm_vs_oracle(rng, ns, nt, iprec, dist, zk, ifc, ifd):
    print("case", ns, nt, iprec, dist, zk, ifc, ifd, flush=True)
    src = points(rng, ns, dist)
    trg = points(rng, nt, dist)
    q, ds, dv = strengths(rng, ns, ifc, ifd, unit_dipvec=False)
    g = F.hfmm3dparttarg(iprec, zk, src, q, ds, dv, trg)
    o = ob.orc_hfmm3dparttarg(iprec, zk, src, q, ds, dv, trg)
    assert g[0] == o[0] == 0
    for k in range(1, 5 if nt else 3):
        _check(g[k], o[k], ORACLE_TOL)


@pytest.mark.parametrize("flags", [(1, 0, 0, 0), (0, 1, 0, 0), (0, 0, 1, 0), (0, 0, 0, 1), (1, 1, 0, 1)])
def test_gpu_helm_output_flags(rng, flags):
    """Outputs are written only when their flag is 1 (src/hfmm3dpart.f:521-530)."""
    src = points(rng, 4000, "cube")
    trg = points(rng, 1500, "cube")
    q, ds, dv = strengths(rng, 4000, True, True)
    ifpot, iffld, ifpt, ifft = flags
    g = F.hfmm3dparttarg(1, 7.0, src, q, ds, dv, trg, ifpot, iffld, ifpt, ifft)
    o = ob.orc_hfmm3dparttarg(1, 7.0, src, q, ds, dv, trg, ifpot, iffld, ifpt, ifft)
    assert g[0] == 0
    for k, flag in zip(range(1, 5), flags):
        if flag:
            _check(g[k], o[k], ORACLE_TOL)
        else:
            assert g[k] is None


def test_gpu_helm_selfcall_and_direct(rng):
    src = points(rng, 5000, "sphere")
    q, ds, dv = strengths(rng, 5000, True, True)
    ier, pot, fld = F.hfmm3dpartself(2, 10.0, src, q, ds, dv)
    o = ob.orc_hfmm3dparttarg(2, 10.0, src, q, ds, dv, None)
    assert ier == 0
    _check(pot, o[1], ORACLE_TOL)
    _check(fld, o[2], ORACLE_TOL)
    trg = points(rng, 700, "cube")
    g = F.h3dpartdirect(3.0 + 0.2j, src[:2000], q[:2000], ds[:2000], dv[:2000], trg)
    d = ob.orc_h3dpartdirect(3.0 + 0.2j, src[:2000], q[:2000], ds[:2000], dv[:2000], trg)
    for a, b in zip(g, d):
        _check(a, b, 1e-13)


def test_gpu_helm_error_codes():
    src = np.random.default_rng(0).random((10, 3))
    q = np.ones(10, np.complex128)
    assert F.hfmm3dparttarg(1, 0.0, src, q)[0] == 128       # zk = 0
    assert F.hfmm3dparttarg(9, 1.0, src, q)[0] == 32        # invalid iprec
    assert F.hfmm3dparttarg(1, 1.0, src[:0], q[:0])[0] == 4  # no sources


def test_gpu_helm_device_tree_equals_host_tree(rng, monkeypatch):
    """Device and host octree builders give bitwise equal Helmholtz results."""
    src = points(rng, 30000, "clusters")
    trg = points(rng, 6000, "clusters")
    q, ds, dv = strengths(rng, 30000, True, True)
    monkeypatch.setenv("FMM3D_TREE", "host")
    a = F.hfmm3dparttarg(2, 6.0 + 0.2j, src, q, ds, dv, trg)
    monkeypatch.delenv("FMM3D_TREE")
    b = F.hfmm3dparttarg(2, 6.0 + 0.2j, src, q, ds, dv, trg)
    assert a[0] == b[0] == 0
    for x, y in zip(a[1:], b[1:]):
        assert np.array_equal(x, y)


@pytest.mark.parametrize("mode", [F.PREC_FP64, F.PREC_MIXED])
def test_gpu_helm_near_field_stream_does_not_change_results(rng, monkeypatch, mode):
    """Near field on the second stream or on the caller's stream
    (FMM3D_NEAR_STREAM=0): bitwise equal results."""
    src = points(rng, 25000, "sphere")
    trg = points(rng, 5000, "cube")
    q, ds, dv = strengths(rng, 25000, True, True)
    monkeypatch.setenv("FMM3D_NEAR_STREAM", "0")
    a = F.hfmm3dparttarg(2, 8.0, src, q, ds, dv, trg, precision=mode)
    monkeypatch.delenv("FMM3D_NEAR_STREAM")
    b = F.hfmm3dparttarg(2, 8.0, src, q, ds, dv, trg, precision=mode)
    assert a[0] == b[0] == 0
    for x, y in zip(a[1:], b[1:]):
        assert np.array_equal(x, y)


def test_gpu_helm_deterministic(rng):
    src = points(rng, 20000, "sphere")
    q, _, _ = strengths(rng, 20000, True, False)
    a = F.hfmm3dpartself(2, 20.0, src, q)
    b = F.hfmm3dpartself(2, 20.0, src, q)
    assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])


@pytest.mark.slow
def test_gpu_helm_sphere_500k_accuracy():
    """500,000 sources on the unit sphere, zk = 20, iprec = 2, charges with
    standard normal parts: error against the FP64 direct sum at 200 sampled
    sources, and linearity in the charges."""
    rng = np.random.default_rng(3)
    n = 500_000
    src = points(rng, n, "sphere")
    q = rng.normal(size=n) + 1j * rng.normal(size=n)
    ier, pot, fld = F.hfmm3dpartself(2, 20.0, src, q)
    assert ier == 0
    idx = rng.choice(n, 200, replace=False)
    pd, fd = ob.orc_direct_sample(src, src[idx], idx.astype(np.int32), q, zk=20.0)
    assert rel(pot[idx], pd) < IPREC_EPS[2]
    assert rel(fld[idx], fd) < 10 * IPREC_EPS[2]
    q2 = rng.normal(size=n) + 1j * rng.normal(size=n)
    a = F.hfmm3dpartself(2, 20.0, src, q2)
    b = F.hfmm3dpartself(2, 20.0, src, q + q2)
    assert rel(pot + a[1], b[1]) < 1e-12


HPREC_CASES = [(1, "cube", 6.0), (1, "plummer", 25.0), (2, "sphere", 20.0), (2, "clusters", 2.0 + 0.5j),
               (2, "plummer", 30.0), (2, "plummer", 4.0), (3, "cube", 10.0), (3, "plummer", 25.0)]


@pytest.mark.parametrize("iprec,dist,zk", HPREC_CASES)
def test_gpu_helm_fp32_near_field_within_tolerance(rng, iprec, dist, zk):
    """FMM3D_PREC_MIXED: FP32 pair arithmetic in the near field (phase k r
    reduced in FP64), FP64 translations; against the FP64 direct sum at the
    tolerance iprec names for the potential and a digit less for the field.
    At iprec 3 the library runs FP64 throughout."""
    src = points(rng, 20000, dist)
    trg = points(rng, 4000, dist)
    q, ds, dv = strengths(rng, 20000, True, True)
    g = F.hfmm3dparttarg(iprec, zk, src, q, ds, dv, trg, precision=F.PREC_MIXED)
    assert g[0] == 0
    assert F.last_precision() == (F.PREC_FP64 if iprec >= 3 else F.PREC_MIXED)
    idx = rng.choice(4000, 300, replace=False)
    pd, fd = ob.orc_direct_sample(src, trg[idx], -np.ones(300, np.int32), q, ds, dv, zk=zk)
    ep, ef = rel(g[3][idx], pd), rel(g[4][idx], fd)
    print("helm iprec %d %s zk %s: pot %.2e fld %.2e" % (iprec, dist, zk, ep, ef))
    assert ep < IPREC_EPS[iprec]
    assert ef < 10 * IPREC_EPS[iprec]
