"""FMM with Hessians (fmm3d_lfmm3dparttarg_hess; FP64): against the oracle's
direct-sum Hessians (restating the reference's lpotfld3dhess /
lpotfld3dhess_dp, src/laprouts3d.f:5442, :5619), exactly when the whole
problem is one leaf (near field only), within a bar tied to iprec
otherwise; pot and fld of the same call equal those of the plain call
bitwise; and the trace pot_xx + pot_yy + pot_zz vanishes (to roundoff) on a
large problem, where it is checked without a direct sum."""
import numpy as np
import pytest

import fmmlib3d_amd as F
import oracle_bindings as ob
from fmm_inputs import IPREC_EPS, points, rel, strengths

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ifc,ifd", [(1, 0), (0, 1), (1, 1)])
def test_gpu_hess_one_leaf_equals_direct(rng, ifc, ifd):
    """iprec 6 puts every point in the root box: the Hessians are the near
    field alone and must equal the direct sum to roundoff."""
    src = points(rng, 3000, "clusters")
    trg = points(rng, 700, "cube")
    q, ds, dv = strengths(rng, 3000, bool(ifc), bool(ifd))
    g = F.lfmm3dparttarg_hess(6, src, q, ds, dv, trg)
    assert g[0] == 0
    hs = ob.orc_direct_hess_sample(src, src, np.arange(3000, dtype=np.int32), q, ds, dv)
    ht = ob.orc_direct_hess_sample(src, trg, -np.ones(700, np.int32), q, ds, dv)
    assert rel(g[3], hs) < 1e-12
    assert rel(g[6], ht) < 1e-12


CASES = [(1, "cube", 1, 1), (2, "cube", 1, 1), (2, "clusters", 1, 1), (2, "plummer", 1, 0), (3, "clusters", 0, 1),
         (0, "plummer", 1, 1), (2, "sphere", 1, 1)]


@pytest.mark.parametrize("iprec,dist,ifc,ifd", CASES)
def test_gpu_hess_vs_direct(rng, iprec, dist, ifc, ifd):
    """Clustered, Plummer and sphere inputs put boxes in lists 3 and 4 (the
    multipole Hessians of list 3); the uniform cube exercises the local
    Hessians.  Bar: 30 times the iprec tolerance (two derivatives of the
    truncated expansions)."""
    ns, nt = 40000, 6000
    src = points(rng, ns, dist)
    trg = points(rng, nt, dist)
    q, ds, dv = strengths(rng, ns, bool(ifc), bool(ifd))
    g = F.lfmm3dparttarg_hess(iprec, src, q, ds, dv, trg)
    assert g[0] == 0
    plain = F.lfmm3dparttarg(iprec, src, q, ds, dv, trg)
    for a, b in zip((g[1], g[2], g[4], g[5]), plain[1:]):
        assert np.array_equal(a, b)
    i_s = rng.choice(ns, 300, replace=False)
    i_t = rng.choice(nt, 300, replace=False)
    hs = ob.orc_direct_hess_sample(src, src[i_s], i_s.astype(np.int32), q, ds, dv)
    ht = ob.orc_direct_hess_sample(src, trg[i_t], -np.ones(300, np.int32), q, ds, dv)
    es, et = rel(g[3][i_s], hs), rel(g[6][i_t], ht)
    print("iprec %d %s: hess rel err sources %.2e targets %.2e" % (iprec, dist, es, et))
    bar = 30 * IPREC_EPS[iprec]
    assert es < bar and et < bar


@pytest.mark.slow
def test_gpu_hess_traceless_at_size():
    rng = np.random.default_rng(41)
    n = 300_000
    src = points(rng, n, "cube")
    trg = points(rng, n, "cube")
    q, ds, dv = strengths(rng, n, True, True)
    g = F.lfmm3dparttarg_hess(2, src, q, ds, dv, trg)
    assert g[0] == 0
    for h in (g[3], g[6]):
        tr = h[:, 0] + h[:, 1] + h[:, 2]
        assert np.linalg.norm(tr) < 1e-9 * np.linalg.norm(h)
