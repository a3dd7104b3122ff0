"""The oracle against the live reference library (oracle/_ref), where it
has been built from the reference sources: random cases beyond the fixtures."""
import numpy as np
import pytest

import oracle_bindings as ob
from fmm_inputs import points, rel, strengths

pytestmark = pytest.mark.skipif(not ob.ref_available(), reason="oracle/_ref not built")

CASES = [(3000, 2000, 1, "cube", 1, 1), (4000, 0, 2, "sphere", 1, 0), (3000, 3000, 0, "clusters", 1, 1),
         (2000, 1000, 3, "cube", 0, 1), (1500, 500, -1, "plummer", 1, 1), (2500, 700, -2, "shell", 1, 1),
         (1, 5, 1, "cube", 1, 1), (200, 0, 6, "cube", 1, 1),
         # p = 51 and 64, with leaves at level 2 (a far field)
         (12000, 2000, 4, "cube", 1, 1), (15000, 1000, 5, "clusters", 1, 1)]


@pytest.mark.parametrize("ns,nt,iprec,dist,ifc,ifd", CASES)
def test_fmm_vs_reference(rng, ns, nt, iprec, dist, ifc, ifd):
    src = points(rng, ns, dist)
    trg = points(rng, nt, dist)
    q, ds, dv = strengths(rng, ns, ifc, ifd, unit_dipvec=False)
    r = ob.ref_lfmm3dparttarg(iprec, src, q, ds, dv, trg)
    o = ob.orc_lfmm3dparttarg(iprec, src, q, ds, dv, trg)
    assert r[0] == o[0] == 0
    for k in range(1, 5 if nt else 3):
        if np.linalg.norm(r[k]) == 0:
            assert np.linalg.norm(o[k]) == 0
        else:
            assert rel(o[k], r[k]) < 1e-12


@pytest.mark.parametrize("ns,nt,nbox,dist", [(3000, 2000, 40, "cube"), (5000, 0, 50, "sphere"),
                                             (4000, 3000, 80, "clusters"), (3000, 100, 30, "plummer")])
def test_tree_vs_reference(rng, ns, nt, nbox, dist):
    src = points(rng, ns, dist)
    trg = points(rng, nt, dist)
    b1, i1, t1, l1, n1 = ob.ref_tree(src, trg, nbox)
    b2, i2, t2, l2, n2 = ob.orc_tree(src, trg, nbox)
    assert n1 == n2 and np.array_equal(b1[:, :17], b2[:, :17])
    assert np.array_equal(i1, i2) and np.array_equal(t1, t2)
    for k in range(5):
        assert l1[k] == l2[k]


def test_nsource_zero_gives_ier4():
    """d3tstrcr returns 128 for n < 1 and lfmm3dparttarg turns that into
    ier = 4 after its retries (src/lfmm3dpart.f:309-330)."""
    src = np.zeros((1, 3))
    import ctypes
    from oracle_bindings import _vp, _i
    r = ob.ref()
    ier = ctypes.c_int(0)
    z = np.zeros(8, np.complex128)
    r.lfmm3dparttarg_(ctypes.byref(ier), _i(1), _i(0), _vp(src), _i(1), _vp(z), _i(0), _vp(z), _vp(src),
                      _i(1), _vp(z), _i(1), _vp(z), _i(0), _vp(src), _i(0), _vp(z), _i(0), _vp(z))
    assert ier.value == 4
    assert ob.oracle().orc_lfmm3dparttarg(1, 0, _vp(src), 1, _vp(z), 0, None, None, 1, _vp(z), 1, _vp(z),
                                          0, None, 0, None, 0, None) == 4


HCASES = [(3000, 1000, 1, "cube", 8.0, 1, 1), (4000, 0, 2, "sphere", 20.0, 1, 0),
          (3000, 800, 2, "clusters", 2.0 + 0.5j, 1, 1), (2500, 500, 0, "plummer", 25.0, 0, 1),
          (2000, 0, 3, "cube", 1.5, 1, 1), (1, 4, 1, "cube", 3.0, 1, 1)]


@pytest.mark.parametrize("ns,nt,iprec,dist,zk,ifc,ifd", HCASES)
def test_helm_fmm_vs_reference(rng, ns, nt, iprec, dist, zk, ifc, ifd):
    """The Helmholtz oracle and the reference agree to a tenth of the iprec
    tolerance (the translations differ by the reference's quadrature only)."""
    from fmm_inputs import IPREC_EPS
    src = points(rng, ns, dist)
    trg = points(rng, nt, dist)
    q, ds, dv = strengths(rng, ns, ifc, ifd, unit_dipvec=False)
    r = ob.ref_hfmm3dparttarg(iprec, zk, src, q, ds, dv, trg)
    o = ob.orc_hfmm3dparttarg(iprec, zk, src, q, ds, dv, trg)
    assert r[0] == o[0] == 0
    for k in range(1, 5 if nt else 3):
        if np.linalg.norm(r[k]) == 0:
            assert np.linalg.norm(o[k]) == 0
        else:
            assert rel(o[k], r[k]) < 0.1 * IPREC_EPS[iprec]
