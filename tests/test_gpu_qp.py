"""Laplace FMM with quadrupole sources (fmm3d_lfmm3dparttarg_qp; FP64)
against the oracle's direct sums (orc_l3ddirect_qp_sample restates the
reference's lpotfld3dhess_qp, src/laprouts3d.f:8691; the charge and dipole
parts orc_l3ddirect_sample): to roundoff when the whole problem is one leaf
(near field only), within a bar tied to iprec otherwise, on inputs that fill
lists 3 and 4 (the quadrupoles' P2M and P2L); and, at a size no direct sum
reaches, that a pure-trace quadrupole (the Laplacian of 1/r) produces
nothing and that the call is linear in its sources."""
import numpy as np
import pytest

import fmmlib3d_amd as F
import oracle_bindings as ob
from fmm_inputs import IPREC_EPS, points, rel, strengths

pytestmark = pytest.mark.gpu


def quads(rng, n):
    return rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n), rng.uniform(-1, 1, (n, 6))


def direct(src, pts, self_index, q, ds, dv, qs, qv):
    pot, fld = ob.orc_direct_qp_sample(src, qs, qv, pts, self_index)
    if q is not None or ds is not None:
        p2, f2 = ob.orc_direct_sample(src, pts, self_index, q, ds, dv)
        pot, fld = pot + p2, fld + f2
    return pot, fld


@pytest.mark.parametrize("ifc,ifd", [(0, 0), (1, 1)])
def test_gpu_qp_one_leaf_equals_direct(rng, ifc, ifd):
    """iprec 6 puts every point in the root box: the near field alone."""
    ns, nt = 2500, 600
    src = points(rng, ns, "clusters")
    trg = points(rng, nt, "cube")
    q, ds, dv = strengths(rng, ns, bool(ifc), bool(ifd))
    qs, qv = quads(rng, ns)
    ier, pot, fld, pt, ft = F.lfmm3dparttarg_qp(6, src, q, ds, dv, qs, qv, trg)
    assert ier == 0
    ps, fs = direct(src, src, np.arange(ns, dtype=np.int32), q, ds, dv, qs, qv)
    pd, fd = direct(src, trg, -np.ones(nt, np.int32), q, ds, dv, qs, qv)
    assert rel(pot, ps) < 1e-12 and rel(fld, fs) < 1e-12
    assert rel(pt, pd) < 1e-12 and rel(ft, fd) < 1e-12


CASES = [(0, "cube", 0), (1, "cube", 0), (2, "cube", 1), (2, "clusters", 0), (2, "plummer", 1), (3, "clusters", 1),
         (3, "sphere", 0), (1, "clusters/cube", 0), (3, "clusters/cube", 0), (-2, "clusters", 1), (-1, "plummer", 0)]


@pytest.mark.parametrize("iprec,dist,ifc", CASES)
def test_gpu_qp_vs_direct(rng, iprec, dist, ifc):
    """Clustered, Plummer and sphere inputs put boxes in lists 3 and 4 (the
    quadrupoles' P2L and the list-3 evaluation of their multipoles); the
    cube exercises P2M, M2L and the locals; clustered sources with targets
    spread over the cube ("clusters/cube") leave most targets with far-field
    values alone; at iprec -2 the expansions have order 2, the lowest whose
    multipoles carry a quadrupole.  Bar: 10 times the iprec tolerance (the expansion orders
    are those chosen for charges; a quadrupole's coefficients grow with the
    degree by two more factors of n)."""
    ns, nt = 40000, 6000
    sd, td = (dist.split("/") + [dist])[:2]
    src = points(rng, ns, sd)
    trg = points(rng, nt, td)
    q, ds, dv = strengths(rng, ns, bool(ifc), bool(ifc))
    qs, qv = quads(rng, ns)
    ier, pot, fld, pt, ft = F.lfmm3dparttarg_qp(iprec, src, q, ds, dv, qs, qv, trg)
    assert ier == 0
    i_s = rng.choice(ns, 300, replace=False)
    i_t = rng.choice(nt, 300, replace=False)
    ps, fs = direct(src, src[i_s], i_s.astype(np.int32), q, ds, dv, qs, qv)
    pd, fd = direct(src, trg[i_t], -np.ones(300, np.int32), q, ds, dv, qs, qv)
    e = [rel(pot[i_s], ps), rel(fld[i_s], fs), rel(pt[i_t], pd), rel(ft[i_t], fd)]
    print("iprec %d %s ifc %d: rel err pot %.2e fld %.2e pottarg %.2e fldtarg %.2e" % ((iprec, dist, ifc) + tuple(e)))
    assert max(e) < 10 * IPREC_EPS[iprec]


def test_gpu_qp_without_quadrupoles_is_the_plain_call(rng):
    src = points(rng, 20000, "plummer")
    trg = points(rng, 3000, "cube")
    q, ds, dv = strengths(rng, 20000, True, True)
    a = F.lfmm3dparttarg_qp(2, src, q, ds, dv, None, None, trg)
    b = F.lfmm3dparttarg(2, src, q, ds, dv, trg)
    assert a[0] == 0 and b[0] == 0
    for x, y in zip(a[1:], b[1:]):
        assert np.array_equal(x, y)


def test_gpu_qp_pure_trace_and_linearity_at_size():
    """300,000 sources: quadrupoles with quadvec = (1, 1, 1, 0, 0, 0) are the
    Laplacian of 1/r and produce nothing (to roundoff against the size of a
    traceless set's output), and the call is linear: charges and quadrupoles
    together equal the sum of the two calls apart."""
    rng = np.random.default_rng(43)
    n = 300_000
    src = points(rng, n, "clusters")
    trg = points(rng, 50_000, "cube")
    qs, qv = quads(rng, n)
    tr = np.tile([1.0, 1.0, 1.0, 0.0, 0.0, 0.0], (n, 1))
    a = F.lfmm3dparttarg_qp(2, src, quadstr=qs, quadvec=tr, target=trg)
    b = F.lfmm3dparttarg_qp(2, src, quadstr=qs, quadvec=qv, target=trg)
    assert a[0] == 0 and b[0] == 0
    for x, y in zip(a[1:], b[1:]):
        assert np.linalg.norm(x) < 1e-9 * np.linalg.norm(y)
    q, _, _ = strengths(rng, n, True, False)
    c = F.lfmm3dparttarg(2, src, q, target=trg)
    d = F.lfmm3dparttarg_qp(2, src, q, quadstr=qs, quadvec=qv, target=trg)
    assert c[0] == 0 and d[0] == 0
    for x, y, z in zip(b[1:], c[1:], d[1:]):
        assert rel(z, x + y) < 1e-12


def test_gpu_qp_sources_only(rng):
    """No targets (the self call): pot and fld at the sources, targets' outputs None."""
    ns = 30000
    src = points(rng, ns, "plummer")
    qs, qv = quads(rng, ns)
    ier, pot, fld, pt, ft = F.lfmm3dparttarg_qp(1, src, quadstr=qs, quadvec=qv)
    assert ier == 0 and pt is None and ft is None
    i_s = rng.choice(ns, 200, replace=False)
    ps, fs = direct(src, src[i_s], i_s.astype(np.int32), None, None, None, qs, qv)
    assert rel(pot[i_s], ps) < IPREC_EPS[1] and rel(fld[i_s], fs) < IPREC_EPS[1]
