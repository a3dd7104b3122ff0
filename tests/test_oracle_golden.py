"""The CPU oracle against the golden fixtures produced by the reference
Fortran (tests/golden/make_golden.py)."""
import glob
import os

import numpy as np
import pytest

import oracle_bindings as ob
from fmm_inputs import rel

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
FMM_FIXTURES = sorted(glob.glob(os.path.join(GOLD, "lap_targ_*.npz")) + glob.glob(os.path.join(GOLD, "lap_self_*.npz")))


def load(name):
    return np.load(os.path.join(GOLD, name), allow_pickle=False)


@pytest.mark.parametrize("path", FMM_FIXTURES, ids=[os.path.basename(p)[:-4] for p in FMM_FIXTURES])
def test_fmm_matches_reference(path):
    f = np.load(path, allow_pickle=False)
    g = lambda k: f[k] if k in f.files else None  # noqa: E731
    ier, pot, fld, pt, ft = ob.orc_lfmm3dparttarg(int(f["iprec"]), f["source"], g("charge"), g("dipstr"),
                                                  g("dipvec"), g("target"))
    assert ier == 0
    assert rel(pot, f["pot"]) < 1e-12
    assert rel(fld, f["fld"]) < 1e-12
    if "target" in f.files:
        assert rel(pt, f["pottarg"]) < 1e-12
        assert rel(ft, f["fldtarg"]) < 1e-12


def test_direct_matches_reference():
    f = load("lap_direct.npz")
    pot, fld, pt, ft = ob.orc_l3dpartdirect(f["source"], f["charge"], f["dipstr"], f["dipvec"], f["target"])
    for a, k in ((pot, "pot"), (fld, "fld"), (pt, "pottarg"), (ft, "fldtarg")):
        assert rel(a, f[k]) < 1e-13


def test_tree_matches_reference():
    f = load("tree_clusters.npz")
    boxes, isrc, itrg, lists, nlev = ob.orc_tree(f["source"], f["target"], int(f["nbox"]))
    assert nlev == int(f["nlev"])
    assert np.array_equal(boxes[:, :17], f["boxes"][:, :17])
    assert np.array_equal(isrc, f["isrc"]) and np.array_equal(itrg, f["itrg"])
    counts, flat = f["list_counts"], f["lists"]
    o = 0
    for k in range(5):
        for b in range(boxes.shape[0]):
            c = counts[k, b]
            assert lists[k][b] == tuple(flat[o:o + c].tolist()), (k, b)
            o += c


def test_operators_match_reference():
    import ctypes
    from oracle_bindings import _vp
    f = load("lap_operators.npz")
    o = ob.oracle()
    p = int(f["p"])
    K = (p + 1) ** 2
    sc, sc2 = ctypes.c_double(float(f["rscale"])), ctypes.c_double(float(f["rscale2"]))
    src, q, ds, dv, trg = f["source"], f["charge"], f["dipstr"], f["dipvec"], f["target"]
    c0, c1, c2, c3 = (np.ascontiguousarray(f[k]) for k in ("c0", "c1", "c2", "c3"))
    ns, nt = src.shape[0], trg.shape[0]
    z = lambda: np.zeros(K, np.complex128)  # noqa: E731
    mp, mpd, lo, lod = z(), z(), z(), z()
    o.orc_l3dformmp(sc, _vp(src), _vp(q), ns, _vp(c1), p, _vp(mp))
    o.orc_l3dformmp_dp(sc, _vp(src), _vp(ds), _vp(dv), ns, _vp(c1), p, _vp(mpd))
    o.orc_l3dformta(sc, _vp(src), _vp(q), ns, _vp(c0), p, _vp(lo))
    o.orc_l3dformta_dp(sc, _vp(src), _vp(ds), _vp(dv), ns, _vp(c0), p, _vp(lod))
    assert rel(mp, f["p2m"]) < 1e-13 and rel(mpd, f["p2m_dp"]) < 1e-13
    assert rel(lo, f["p2l"]) < 1e-13 and rel(lod, f["p2l_dp"]) < 1e-13
    pot, fld = np.zeros(nt, np.complex128), np.zeros((nt, 3), np.complex128)
    o.orc_l3dmpeval(sc, _vp(c1), _vp(mp), p, _vp(trg), nt, 1, _vp(pot), 1, _vp(fld))
    assert rel(pot, f["m2p_pot"]) < 1e-13 and rel(fld, f["m2p_fld"]) < 1e-13
    pot, fld = np.zeros(nt, np.complex128), np.zeros((nt, 3), np.complex128)
    o.orc_l3dtaeval(sc, _vp(c0), _vp(lo), p, _vp(trg), nt, 1, _vp(pot), 1, _vp(fld))
    assert rel(pot, f["l2p_pot"]) < 1e-13 and rel(fld, f["l2p_fld"]) < 1e-13
    m2m, m2l, m2l5, l2l = z(), z(), z(), z()
    o.orc_l3dmpmp(sc, _vp(c1), _vp(mp), p, sc2, _vp(c2), _vp(m2m), p)
    o.orc_l3dmploc(sc, _vp(c1), _vp(mp), p, p, sc2, _vp(c0), _vp(m2l), p)
    o.orc_l3dmploc(sc, _vp(c1), _vp(mp), p, 5, sc2, _vp(c0), _vp(m2l5), p)
    o.orc_l3dlocloc(sc, _vp(c0), _vp(lo), p, sc2, _vp(c3), _vp(l2l), p)
    assert rel(m2m, f["m2m"]) < 1e-13
    assert rel(m2l, f["m2l"]) < 1e-13
    assert rel(m2l5, f["m2l_trunc5"]) < 1e-13
    assert rel(l2l, f["l2l"]) < 1e-13


def test_orders_match_reference():
    import ctypes
    from oracle_bindings import _vp
    f = load("lap_orders.npz")
    o = ob.oracle()
    for e, n, it in zip(f["eps"], f["nterms"], f["itables"]):
        ier = ctypes.c_int(0)
        assert o.orc_l3dterms(ctypes.c_double(float(e)), ctypes.byref(ier)) == n
        tab = np.zeros(343, np.int32)
        o.orc_l3dterms_list2(ctypes.c_double(float(e)), _vp(tab), ctypes.byref(ier))
        assert np.array_equal(tab.reshape(7, 7, 7), it)
