"""ctypes bindings for the CPU oracle (oracle/liboracle.so) and, where it has
been built, the reference Fortran library (oracle/_ref/libfmmlib3d_ref.so).

Test infrastructure: used by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg only.  Arrays follow the reference's Fortran layouts:
source(3,n) is a C-contiguous (n,3) float64 array, charges complex128,
fld(3,n) a C-contiguous (n,3) complex128 array.
"""
import ctypes
import os

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_SO = os.path.join(ROOT, "oracle", "liboracle.so")
REF_SO = os.path.join(ROOT, "oracle", "_ref", "libfmmlib3d_ref.so")

_oracle = None
_ref = None


def _vp(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def oracle():
    global _oracle
    if _oracle is None:
        if not os.path.exists(ORACLE_SO):
            raise RuntimeError("oracle not built: run `make -C oracle`")
        _oracle = ctypes.CDLL(ORACLE_SO)
        _oracle.orc_l3dterms.restype = ctypes.c_int
        _oracle.orc_lfmm3dparttarg.restype = ctypes.c_int
    return _oracle


def ref_available():
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        if not ref_available():
            raise RuntimeError("reference library not built: run `make -C oracle ref`")
        _ref = ctypes.CDLL(REF_SO)
    return _ref


def _i(v):
    return ctypes.byref(ctypes.c_int(int(v)))


def _d(v):
    return ctypes.byref(ctypes.c_double(float(v)))


def _z(v):
    a = np.array([complex(v)], dtype=np.complex128)
    return a, _vp(a)


# --------------------------------------------------------------------------
# packed (n*n+n+m) <-> reference mpole(0:p,-p:p) column-major
# --------------------------------------------------------------------------
def packed_to_ref(M, p):
    A = np.zeros((2 * p + 1, p + 1), dtype=np.complex128)  # [m+p, n] -> Fortran (n,m)
    for n in range(p + 1):
        for m in range(-n, n + 1):
            A[m + p, n] = M[n * n + n + m]
    return np.ascontiguousarray(A).reshape(-1)


def ref_to_packed(A, p):
    A = A.reshape(2 * p + 1, p + 1)
    M = np.zeros((p + 1) ** 2, dtype=np.complex128)
    for n in range(p + 1):
        for m in range(-n, n + 1):
            M[n * n + n + m] = A[m + p, n]
    return M


# --------------------------------------------------------------------------
# oracle wrappers
# --------------------------------------------------------------------------
def _c(a, dt):
    return None if a is None else np.ascontiguousarray(a, dtype=dt)


def orc_lfmm3dparttarg(iprec, source, charge=None, dipstr=None, dipvec=None, target=None,
                       ifpot=1, iffld=1, ifpottarg=1, iffldtarg=1):
    o = oracle()
    src = _c(source, np.float64)
    ns = src.shape[0]
    trg = _c(target, np.float64) if target is not None else np.zeros((1, 3))
    nt = 0 if target is None else trg.shape[0]
    q = _c(charge, np.complex128)
    ds = _c(dipstr, np.complex128)
    dv = _c(dipvec, np.float64)
    pot = np.zeros(ns, np.complex128)
    fld = np.zeros((ns, 3), np.complex128)
    pt = np.zeros(max(nt, 1), np.complex128)
    ft = np.zeros((max(nt, 1), 3), np.complex128)
    ier = o.orc_lfmm3dparttarg(int(iprec), ns, _vp(src), int(q is not None), _vp(q),
                               int(ds is not None), _vp(ds), _vp(dv), ifpot, _vp(pot), iffld,
                               _vp(fld), nt, _vp(trg), ifpottarg if nt else 0, _vp(pt),
                               iffldtarg if nt else 0, _vp(ft))
    return ier, pot, fld, pt[:nt], ft[:nt]


def orc_l3dpartdirect(source, charge=None, dipstr=None, dipvec=None, target=None):
    o = oracle()
    src = _c(source, np.float64)
    ns = src.shape[0]
    trg = _c(target, np.float64) if target is not None else np.zeros((1, 3))
    nt = 0 if target is None else trg.shape[0]
    q = _c(charge, np.complex128)
    ds = _c(dipstr, np.complex128)
    dv = _c(dipvec, np.float64)
    pot = np.zeros(ns, np.complex128)
    fld = np.zeros((ns, 3), np.complex128)
    pt = np.zeros(max(nt, 1), np.complex128)
    ft = np.zeros((max(nt, 1), 3), np.complex128)
    o.orc_l3dpartdirect(ns, _vp(src), int(q is not None), _vp(q), int(ds is not None), _vp(ds),
                        _vp(dv), 1, _vp(pot), 1, _vp(fld), nt, _vp(trg), 1 if nt else 0, _vp(pt),
                        1 if nt else 0, _vp(ft))
    return pot, fld, pt[:nt], ft[:nt]


def orc_direct_sample(source, points, self_index, charge=None, dipstr=None, dipvec=None, zk=None):
    """FP64 direct sum at `points`; self_index[j] >= 0 skips that source."""
    o = oracle()
    src = _c(source, np.float64)
    pts = _c(points, np.float64)
    si = _c(self_index, np.int32)
    q = _c(charge, np.complex128)
    ds = _c(dipstr, np.complex128)
    dv = _c(dipvec, np.float64)
    pot = np.zeros(pts.shape[0], np.complex128)
    fld = np.zeros((pts.shape[0], 3), np.complex128)
    if zk is None:
        o.orc_l3ddirect_sample(src.shape[0], _vp(src), int(q is not None), _vp(q),
                               int(ds is not None), _vp(ds), _vp(dv), pts.shape[0], _vp(pts),
                               _vp(si), _vp(pot), _vp(fld))
    else:
        o.orc_h3ddirect_sample.argtypes = [ZC] + [ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
                                                 ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                                 ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
                                                 ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        o.orc_h3ddirect_sample(zc(zk), src.shape[0], _vp(src), int(q is not None), _vp(q),
                               int(ds is not None), _vp(ds), _vp(dv), pts.shape[0], _vp(pts),
                               _vp(si), _vp(pot), _vp(fld))
    return pot, fld


def orc_direct_hess_sample(source, points, self_index, charge=None, dipstr=None, dipvec=None):
    """FP64 direct-sum Hessians (n, 6) at `points` (order pot_xx, pot_yy,
    pot_zz, pot_xy, pot_xz, pot_yz); self_index[j] >= 0 skips that source."""
    o = oracle()
    src = _c(source, np.float64)
    pts = _c(points, np.float64)
    si = _c(self_index, np.int32)
    q = _c(charge, np.complex128)
    ds = _c(dipstr, np.complex128)
    dv = _c(dipvec, np.float64)
    hs = np.zeros((pts.shape[0], 6), np.complex128)
    o.orc_l3ddirect_hess_sample(src.shape[0], _vp(src), int(q is not None), _vp(q), int(ds is not None), _vp(ds),
                                _vp(dv), pts.shape[0], _vp(pts), _vp(si), _vp(hs))
    return hs


def ref_lpotfld3dhess(source, charge, target):
    """The reference's lpotfld3dhess (one charge, one target): pot, fld, hess."""
    r = ref()
    s, t = _c(source, np.float64), _c(target, np.float64)
    q = np.array([charge], np.complex128)
    pot = np.zeros(1, np.complex128)
    fld = np.zeros(3, np.complex128)
    hs = np.zeros(6, np.complex128)
    one = ctypes.c_int(1)
    r.lpotfld3dhess_(ctypes.byref(one), ctypes.byref(one), _vp(s), _vp(q), _vp(t), _vp(pot), _vp(fld), _vp(hs))
    return pot[0], fld, hs


def ref_lpotfld3dhess_dp(source, dipstr, dipvec, target):
    """The reference's lpotfld3dhess_dp (one dipole, one target): pot, fld, hess."""
    r = ref()
    s, t, dv = _c(source, np.float64), _c(target, np.float64), _c(dipvec, np.float64)
    ds = np.array([dipstr], np.complex128)
    pot = np.zeros(1, np.complex128)
    fld = np.zeros(3, np.complex128)
    hs = np.zeros(6, np.complex128)
    one = ctypes.c_int(1)
    r.lpotfld3dhess_dp_(ctypes.byref(one), ctypes.byref(one), _vp(s), _vp(ds), _vp(dv), _vp(t), _vp(pot), _vp(fld),
                        _vp(hs))
    return pot[0], fld, hs


class ZC(ctypes.Structure):
    """C `double _Complex` passed by value (two doubles, SysV classes SSE,SSE)."""
    _fields_ = [("re", ctypes.c_double), ("im", ctypes.c_double)]


def zc(v):
    v = complex(v)
    return ZC(v.real, v.imag)


def orc_hfmm3dparttarg(iprec, zk, source, charge=None, dipstr=None, dipvec=None, target=None,
                       ifpot=1, iffld=1, ifpottarg=1, iffldtarg=1):
    o = oracle()
    src = _c(source, np.float64)
    ns = src.shape[0]
    trg = _c(target, np.float64) if target is not None else np.zeros((1, 3))
    nt = 0 if target is None else trg.shape[0]
    q = _c(charge, np.complex128)
    ds = _c(dipstr, np.complex128)
    dv = _c(dipvec, np.float64)
    pot = np.zeros(ns, np.complex128)
    fld = np.zeros((ns, 3), np.complex128)
    pt = np.zeros(max(nt, 1), np.complex128)
    ft = np.zeros((max(nt, 1), 3), np.complex128)
    V, I_ = ctypes.c_void_p, ctypes.c_int
    o.orc_hfmm3dparttarg.restype = I_
    o.orc_hfmm3dparttarg.argtypes = [I_, ZC, I_, V, I_, V, I_, V, V, I_, V, I_, V, I_, V, I_, V, I_, V]
    ier = o.orc_hfmm3dparttarg(int(iprec), zc(zk), ns, _vp(src), int(q is not None), _vp(q),
                               int(ds is not None), _vp(ds), _vp(dv), ifpot, _vp(pot), iffld,
                               _vp(fld), nt, _vp(trg), ifpottarg if nt else 0, _vp(pt),
                               iffldtarg if nt else 0, _vp(ft))
    return ier, pot, fld, pt[:nt], ft[:nt]


def orc_h3dpartdirect(zk, source, charge=None, dipstr=None, dipvec=None, target=None):
    o = oracle()
    src = _c(source, np.float64)
    ns = src.shape[0]
    trg = _c(target, np.float64) if target is not None else np.zeros((1, 3))
    nt = 0 if target is None else trg.shape[0]
    q = _c(charge, np.complex128)
    ds = _c(dipstr, np.complex128)
    dv = _c(dipvec, np.float64)
    pot = np.zeros(ns, np.complex128)
    fld = np.zeros((ns, 3), np.complex128)
    pt = np.zeros(max(nt, 1), np.complex128)
    ft = np.zeros((max(nt, 1), 3), np.complex128)
    V, I_ = ctypes.c_void_p, ctypes.c_int
    o.orc_h3dpartdirect.argtypes = [ZC, I_, V, I_, V, I_, V, V, I_, V, I_, V, I_, V, I_, V, I_, V]
    o.orc_h3dpartdirect(zc(zk), ns, _vp(src), int(q is not None), _vp(q), int(ds is not None), _vp(ds),
                        _vp(dv), 1, _vp(pot), 1, _vp(fld), nt, _vp(trg), 1 if nt else 0, _vp(pt),
                        1 if nt else 0, _vp(ft))
    return pot, fld, pt[:nt], ft[:nt]


# --------------------------------------------------------------------------
# reference wrappers (only when oracle/_ref is built)
# --------------------------------------------------------------------------
def ref_lfmm3dparttarg(iprec, source, charge=None, dipstr=None, dipvec=None, target=None,
                       ifpot=1, iffld=1, ifpottarg=1, iffldtarg=1):
    r = ref()
    src = _c(source, np.float64)
    ns = src.shape[0]
    trg = _c(target, np.float64) if target is not None else np.zeros((1, 3))
    nt = 0 if target is None else trg.shape[0]
    q = _c(charge, np.complex128) if charge is not None else np.zeros(1, np.complex128)
    ds = _c(dipstr, np.complex128) if dipstr is not None else np.zeros(1, np.complex128)
    dv = _c(dipvec, np.float64) if dipvec is not None else np.zeros(3)
    pot = np.zeros(ns, np.complex128)
    fld = np.zeros((ns, 3), np.complex128)
    pt = np.zeros(max(nt, 1), np.complex128)
    ft = np.zeros((max(nt, 1), 3), np.complex128)
    ier = ctypes.c_int(0)
    r.lfmm3dparttarg_(ctypes.byref(ier), _i(iprec), _i(ns), _vp(src), _i(charge is not None),
                      _vp(q), _i(dipstr is not None), _vp(ds), _vp(dv), _i(ifpot), _vp(pot),
                      _i(iffld), _vp(fld), _i(nt), _vp(trg), _i(ifpottarg if nt else 0), _vp(pt),
                      _i(iffldtarg if nt else 0), _vp(ft))
    return ier.value, pot, fld, pt[:nt], ft[:nt]


def ref_l3dpartdirect(source, charge=None, dipstr=None, dipvec=None, target=None):
    r = ref()
    src = _c(source, np.float64)
    ns = src.shape[0]
    trg = _c(target, np.float64) if target is not None else np.zeros((1, 3))
    nt = 0 if target is None else trg.shape[0]
    q = _c(charge, np.complex128) if charge is not None else np.zeros(1, np.complex128)
    ds = _c(dipstr, np.complex128) if dipstr is not None else np.zeros(1, np.complex128)
    dv = _c(dipvec, np.float64) if dipvec is not None else np.zeros(3)
    pot = np.zeros(ns, np.complex128)
    fld = np.zeros((ns, 3), np.complex128)
    pt = np.zeros(max(nt, 1), np.complex128)
    ft = np.zeros((max(nt, 1), 3), np.complex128)
    r.l3dpartdirect_(_i(ns), _vp(src), _i(charge is not None), _vp(q), _i(dipstr is not None),
                     _vp(ds), _vp(dv), _i(1), _vp(pot), _i(1), _vp(fld), _i(nt), _vp(trg),
                     _i(1 if nt else 0), _vp(pt), _i(1 if nt else 0), _vp(ft))
    return pot, fld, pt[:nt], ft[:nt]


def ref_hfmm3dparttarg(iprec, zk, source, charge=None, dipstr=None, dipvec=None, target=None,
                       ifpot=1, iffld=1, ifpottarg=1, iffldtarg=1):
    r = ref()
    src = _c(source, np.float64)
    ns = src.shape[0]
    trg = _c(target, np.float64) if target is not None else np.zeros((1, 3))
    nt = 0 if target is None else trg.shape[0]
    q = _c(charge, np.complex128) if charge is not None else np.zeros(1, np.complex128)
    ds = _c(dipstr, np.complex128) if dipstr is not None else np.zeros(1, np.complex128)
    dv = _c(dipvec, np.float64) if dipvec is not None else np.zeros(3)
    pot = np.zeros(ns, np.complex128)
    fld = np.zeros((ns, 3), np.complex128)
    pt = np.zeros(max(nt, 1), np.complex128)
    ft = np.zeros((max(nt, 1), 3), np.complex128)
    ier = ctypes.c_int(0)
    zka, zkp = _z(zk)
    r.hfmm3dparttarg_(ctypes.byref(ier), _i(iprec), zkp, _i(ns), _vp(src), _i(charge is not None),
                      _vp(q), _i(dipstr is not None), _vp(ds), _vp(dv), _i(ifpot), _vp(pot),
                      _i(iffld), _vp(fld), _i(nt), _vp(trg), _i(ifpottarg if nt else 0), _vp(pt),
                      _i(iffldtarg if nt else 0), _vp(ft))
    return ier.value, pot, fld, pt[:nt], ft[:nt]


def ref_h3dpartdirect(zk, source, charge=None, dipstr=None, dipvec=None, target=None):
    r = ref()
    src = _c(source, np.float64)
    ns = src.shape[0]
    trg = _c(target, np.float64) if target is not None else np.zeros((1, 3))
    nt = 0 if target is None else trg.shape[0]
    q = _c(charge, np.complex128) if charge is not None else np.zeros(1, np.complex128)
    ds = _c(dipstr, np.complex128) if dipstr is not None else np.zeros(1, np.complex128)
    dv = _c(dipvec, np.float64) if dipvec is not None else np.zeros(3)
    pot = np.zeros(ns, np.complex128)
    fld = np.zeros((ns, 3), np.complex128)
    pt = np.zeros(max(nt, 1), np.complex128)
    ft = np.zeros((max(nt, 1), 3), np.complex128)
    zka, zkp = _z(zk)
    r.h3dpartdirect_(zkp, _i(ns), _vp(src), _i(charge is not None), _vp(q), _i(dipstr is not None),
                     _vp(ds), _vp(dv), _i(1), _vp(pot), _i(1), _vp(fld), _i(nt), _vp(trg),
                     _i(1 if nt else 0), _vp(pt), _i(1 if nt else 0), _vp(ft))
    return pot, fld, pt[:nt], ft[:nt]


def ref_tree(source, target, nbox):
    """Run the reference d3tstrcr and return (boxes20, isrc, itrg, lists) with
    lists[k][b] a sorted tuple of 1-based box ids (k = 0..4 for lists 1..5)."""
    r = ref()
    src = _c(source, np.float64)
    trg = _c(target, np.float64) if target is not None else np.zeros((1, 3))
    ns = src.shape[0]
    nt = 0 if target is None else trg.shape[0]
    lw = 200 * (ns + nt) + 100000
    w = np.zeros(lw, np.int32)
    iz = np.zeros(ns, np.int32)
    izt = np.zeros(max(nt, 1), np.int32)
    laddr = np.zeros(400, np.int32)
    center = np.zeros(3)
    ier = ctypes.c_int(0)
    nboxes = ctypes.c_int(0)
    nlev = ctypes.c_int(0)
    size = ctypes.c_double(0)
    lused = ctypes.c_int(0)
    r.d3tstrcr_(ctypes.byref(ier), _vp(src), _i(ns), _i(nbox), ctypes.byref(nboxes), _vp(iz),
                _vp(laddr), ctypes.byref(nlev), _vp(center), ctypes.byref(size), _vp(trg), _i(nt),
                _vp(izt), _vp(w), _i(lw), ctypes.byref(lused))
    if ier.value:
        raise RuntimeError("d3tstrcr ier=%d" % ier.value)
    nb = nboxes.value
    boxes = np.zeros((nb, 20), np.int32)
    box = np.zeros(20, np.int32)
    c0 = np.zeros(3)
    corners = np.zeros(24)
    lst = np.zeros(20000, np.int32)
    nl = ctypes.c_int(0)
    lists = [[() for _ in range(nb)] for _ in range(5)]
    for b in range(1, nb + 1):
        r.d3tgetb_(ctypes.byref(ier), _i(b), _vp(box), _vp(c0), _vp(corners), _vp(w))
        boxes[b - 1] = box
        for k in range(1, 6):
            nl.value = 0
            r.d3tgetl_(ctypes.byref(ier), _i(b), _i(k), _vp(lst), ctypes.byref(nl), _vp(w))
            if ier.value == 0 and nl.value > 0:
                lists[k - 1][b - 1] = tuple(sorted(lst[: nl.value].tolist()))
    return boxes, iz.copy(), izt[:nt].copy(), lists, nlev.value


def orc_tree(source, target, nbox, maxboxes=200000, maxlist=20000000):
    o = oracle()
    src = _c(source, np.float64)
    trg = _c(target, np.float64) if target is not None else np.zeros((1, 3))
    ns = src.shape[0]
    nt = 0 if target is None else trg.shape[0]
    nboxes = ctypes.c_int(0)
    nlev = ctypes.c_int(0)
    boxes = np.zeros((maxboxes, 20), np.int32)
    isrc = np.zeros(ns, np.int32)
    itrg = np.zeros(max(nt, 1), np.int32)
    counts = np.zeros(5 * maxboxes, np.int32)
    lists = np.zeros(maxlist, np.int32)
    ier = o.orc_tree_summary(_vp(src), ns, _vp(trg), nt, int(nbox), ctypes.byref(nboxes),
                             ctypes.byref(nlev), _vp(boxes), maxboxes, _vp(isrc), _vp(itrg),
                             _vp(counts), _vp(lists), maxlist)
    if ier:
        raise RuntimeError("orc_tree_summary ier=%d" % ier)
    nb = nboxes.value
    out = [[() for _ in range(nb)] for _ in range(5)]
    o_ = 0
    for k in range(5):
        for b in range(nb):
            c = counts[k * maxboxes + b]
            out[k][b] = tuple(lists[o_: o_ + c].tolist())
            o_ += c
    return boxes[:nb], isrc, itrg[:nt], out, nlev.value


# ---- Laplace layer potentials of flat triangles ----

def orc_tri_unit(tri, x, self=False):
    """The oracle's unit-density integrals of one triangle (3x3 vertices) at x:
    (SLP pot, SLP fld[3], DLP pot, DLP fld[3])."""
    o = oracle()
    t = _c(tri, np.float64).reshape(9)
    p = _c(x, np.float64).reshape(3)
    out = np.zeros(8)
    o.orc_tri_unit(_vp(t), _vp(p), ctypes.c_int(1 if self else 0), _vp(out))
    return out[0], out[1:4].copy(), out[4], out[5:8].copy()


def orc_tria_direct_sample(tri, points, self_index, charge=None, dipstr=None):
    """The oracle's direct triangle sum at `points`; self_index[j] >= 0 names
    the triangle whose centroid point j is."""
    o = oracle()
    t = _c(tri, np.float64).reshape(-1, 9)
    pts = _c(points, np.float64).reshape(-1, 3)
    n, npts = t.shape[0], pts.shape[0]
    q = _c(charge, np.complex128) if charge is not None else np.zeros(1, np.complex128)
    ds = _c(dipstr, np.complex128) if dipstr is not None else np.zeros(1, np.complex128)
    si = _c(self_index, np.int32) if self_index is not None else None
    pot = np.zeros(npts, np.complex128)
    fld = np.zeros((npts, 3), np.complex128)
    o.orc_l3dtriadirect_sample(ctypes.c_int(n), _vp(t), ctypes.c_int(charge is not None), _vp(q),
                               ctypes.c_int(dipstr is not None), _vp(ds), ctypes.c_int(npts), _vp(pts),
                               _vp(si), _vp(pot), _vp(fld))
    return pot, fld


def ref_tri_unit(tri, x, self=False):
    """The reference's analytic triangle quadratures for one triangle at x
    (direct3dtritarglaps2 / direct3dtritarglapd2, or at the triangle's own
    centroid direct3dtrilaps2 / direct3dtrilapd2): unit-density (SLP pot,
    SLP fld, DLP pot, DLP fld)."""
    r = ref()
    t = np.asfortranarray(_c(tri, np.float64).reshape(3, 3).T)   # triang(3,3): column j = vertex j
    p = _c(x, np.float64).reshape(3)
    e1, e2 = t[:, 1] - t[:, 0], t[:, 2] - t[:, 0]
    nrm = np.cross(e1, e2)
    nrm = nrm / np.linalg.norm(nrm)
    one = np.array([1.0 + 0j])
    pot = np.zeros(1, np.complex128)
    fld = np.zeros(3, np.complex128)
    out = []
    if self:
        r.direct3dtrilaps2_(_i(1), _i(1), _vp(p), _vp(one), _vp(t), _i(1), _vp(pot), _i(1), _vp(fld))
    else:
        r.direct3dtritarglaps2_(_i(1), _vp(p), _vp(one), _vp(t), _i(1), _vp(pot), _i(1), _vp(fld))
    out += [pot[0].real, fld.real.copy()]
    pot[:] = 0
    fld[:] = 0
    if self:
        r.direct3dtrilapd2_(_i(1), _i(1), _vp(p), _vp(one), _vp(t), _vp(nrm), _i(1), _vp(pot), _i(1), _vp(fld))
    else:
        r.direct3dtritarglapd2_(_i(1), _vp(p), _vp(one), _vp(t), _vp(nrm), _i(1), _vp(pot), _i(1), _vp(fld))
    out += [pot[0].real, fld.real.copy()]
    return tuple(out)


def _tri_args(tri, cent, charge, dipstr, target):
    t = np.asfortranarray(np.transpose(_c(tri, np.float64).reshape(-1, 3, 3), (0, 2, 1)))  # (n, 3xyz, 3vert) C
    t = np.ascontiguousarray(_c(tri, np.float64).reshape(-1, 9))     # vertex-major = Fortran triaflat(3,3,n)
    n = t.shape[0]
    v = t.reshape(n, 3, 3)
    nrm = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
    nrm = np.ascontiguousarray(nrm / np.linalg.norm(nrm, axis=1)[:, None])
    c = _c(cent, np.float64)
    q = _c(charge, np.complex128) if charge is not None else np.zeros(1, np.complex128)
    ds = _c(dipstr, np.complex128) if dipstr is not None else np.zeros(1, np.complex128)
    trg = _c(target, np.float64) if target is not None else np.zeros((1, 3))
    nt = 0 if target is None else trg.shape[0]
    return t, nrm, c, q, ds, trg, nt, n


def ref_l3dtriadirect(tri, cent, charge=None, dipstr=None, target=None):
    """The reference's l3dtriadirect (src/lfmm3dtria.f:1385)."""
    r = ref()
    t, nrm, c, q, ds, trg, nt, n = _tri_args(tri, cent, charge, dipstr, target)
    pot = np.zeros(n, np.complex128)
    fld = np.zeros((n, 3), np.complex128)
    pt = np.zeros(max(nt, 1), np.complex128)
    ft = np.zeros((max(nt, 1), 3), np.complex128)
    r.l3dtriadirect_(_i(n), _vp(t), _vp(nrm), _vp(c), _i(charge is not None), _vp(q),
                     _i(dipstr is not None), _vp(ds), _vp(nrm), _i(1), _vp(pot), _i(1), _vp(fld),
                     _i(nt), _vp(trg), _i(1 if nt else 0), _vp(pt), _i(1 if nt else 0), _vp(ft))
    return pot, fld, pt[:nt], ft[:nt]


def ref_lfmm3dtriatarg(iprec, tri, cent, charge=None, dipstr=None, target=None):
    """The reference's triangle FMM lfmm3dtriatarg (src/lfmm3dtria.f:121)."""
    r = ref()
    t, nrm, c, q, ds, trg, nt, n = _tri_args(tri, cent, charge, dipstr, target)
    pot = np.zeros(n, np.complex128)
    fld = np.zeros((n, 3), np.complex128)
    pt = np.zeros(max(nt, 1), np.complex128)
    ft = np.zeros((max(nt, 1), 3), np.complex128)
    ier = ctypes.c_int(0)
    r.lfmm3dtriatarg_(ctypes.byref(ier), _i(iprec), _i(n), _vp(t), _vp(nrm), _vp(c), _i(charge is not None),
                      _vp(q), _i(dipstr is not None), _vp(ds), _vp(nrm), _i(1), _vp(pot), _i(1), _vp(fld),
                      _i(nt), _vp(trg), _i(1 if nt else 0), _vp(pt), _i(1 if nt else 0), _vp(ft))
    return ier.value, pot, fld, pt[:nt], ft[:nt]


# ---- Helmholtz layer potentials of flat triangles ----

def orc_tri_symrule(degree):
    """The oracle's D3-symmetric rule of the given degree on the reference
    triangle: (xi, eta, w), weights summing to 1/2."""
    o = oracle()
    xi, eta, w = np.zeros(16), np.zeros(16), np.zeros(16)
    n = o.orc_tri_symrule(ctypes.c_int(degree), _vp(xi), _vp(eta), _vp(w))
    return xi[:n].copy(), eta[:n].copy(), w[:n].copy()


def orc_htri_unit(tri, x, zk, degree, self=False, lap_c=True, lap_d=True):
    """The oracle's unit-density Helmholtz integrals of one triangle at x:
    (SLP pot, SLP fld[3], DLP pot, DLP fld[3]), complex."""
    o = oracle()
    xi, eta, w = orc_tri_symrule(degree)
    t = _c(tri, np.float64).reshape(9)
    p = _c(x, np.float64).reshape(3)
    out = np.zeros(8, np.complex128)
    o.orc_htri_unit(_vp(t), _vp(p), ctypes.c_int(1 if self else 0), ctypes.c_double(complex(zk).real),
                    ctypes.c_double(complex(zk).imag), ctypes.c_int(len(w)), _vp(xi), _vp(eta), _vp(w),
                    ctypes.c_int(1 if lap_c else 0), ctypes.c_int(1 if lap_d else 0), _vp(out))
    return out[0], out[1:4].copy(), out[4], out[5:8].copy()


def orc_htria_direct_sample(tri, zk, nqtri, points, self_index, charge=None, dipstr=None, ifcharge=1,
                            ifdipole=1):
    """The oracle's direct Helmholtz triangle sum (h3dtriadirect) at
    `points`; flag 2 subtracts the Laplace part of that density."""
    o = oracle()
    t = _c(tri, np.float64).reshape(-1, 9)
    pts = _c(points, np.float64).reshape(-1, 3)
    n, npts = t.shape[0], pts.shape[0]
    q = _c(charge, np.complex128) if charge is not None else np.zeros(1, np.complex128)
    ds = _c(dipstr, np.complex128) if dipstr is not None else np.zeros(1, np.complex128)
    si = _c(self_index, np.int32) if self_index is not None else None
    pot = np.zeros(npts, np.complex128)
    fld = np.zeros((npts, 3), np.complex128)
    z = np.array([complex(zk)], np.complex128)
    o.orc_h3dtriadirect_sample(ctypes.c_int(n), _vp(t), ctypes.c_double(z[0].real), ctypes.c_double(z[0].imag),
                               ctypes.c_int(nqtri), ctypes.c_int(ifcharge if charge is not None else 0), _vp(q),
                               ctypes.c_int(ifdipole if dipstr is not None else 0), _vp(ds), ctypes.c_int(npts),
                               _vp(pts), _vp(si), _vp(pot), _vp(fld))
    return pot, fld


def ref_triasymq(degree):
    """The reference's triasymq rule on the triangle (0,0), (1,0), (0,1)."""
    r = ref()
    v1, v2, v3 = np.array([0.0, 0.0]), np.array([1.0, 0.0]), np.array([0.0, 1.0])
    rn = np.zeros((500, 2))
    w = np.zeros(500)
    nn = ctypes.c_int(0)
    r.triasymq_(_i(degree), _vp(v1), _vp(v2), _vp(v3), _vp(rn), _vp(w), ctypes.byref(nn))
    return rn[:nn.value, 0].copy(), rn[:nn.value, 1].copy(), w[:nn.value].copy()


def ref_htri_unit(tri, x, zk, nqtri, self=False, ifl_c=0, ifl_d=0):
    """The reference's Helmholtz triangle quadratures for one triangle at x
    (direct3dtritarghelms3 / direct3dtritarghelmd3, or at the triangle's own
    centroid direct3dtrihelms2 / direct3dtrihelmd2), unit densities; ifl 1
    for the difference kernel."""
    r = ref()
    t = np.asfortranarray(_c(tri, np.float64).reshape(3, 3).T)
    p = _c(x, np.float64).reshape(3)
    e1, e2 = t[:, 1] - t[:, 0], t[:, 2] - t[:, 0]
    nrm = np.cross(e1, e2)
    nrm = nrm / np.linalg.norm(nrm)
    z = np.array([complex(zk)], np.complex128)
    one = np.array([1.0 + 0j])
    pot = np.zeros(1, np.complex128)
    fld = np.zeros(3, np.complex128)
    if self:
        r.direct3dtrihelms2_(_i(ifl_c), _i(1), _i(1), _vp(z), _i(nqtri), _vp(p), _vp(one), _vp(t), _vp(pot), _vp(fld))
    else:
        r.direct3dtritarghelms3_(_i(ifl_c), _i(1), _vp(p), _vp(z), _i(nqtri), _vp(one), _vp(t), _i(1), _vp(pot),
                                 _i(1), _vp(fld))
    out = [pot[0], fld.copy()]
    pot[:] = 0
    fld[:] = 0
    if self:
        r.direct3dtrihelmd2_(_i(ifl_d), _i(1), _i(1), _vp(z), _i(nqtri), _vp(p), _vp(one), _vp(t), _vp(nrm),
                             _vp(pot), _vp(fld))
    else:
        r.direct3dtritarghelmd3_(_i(ifl_d), _i(1), _vp(p), _vp(z), _i(nqtri), _vp(one), _vp(t), _vp(nrm), _i(1),
                                 _vp(pot), _i(1), _vp(fld))
    out += [pot[0], fld.copy()]
    return tuple(out)


def ref_h3dtriadirect(nqtri, zk, tri, cent, charge=None, dipstr=None, target=None, ifcharge=1, ifdipole=1):
    """The reference's h3dtriadirect (src/hfmm3dtria.f:1456)."""
    r = ref()
    t, nrm, c, q, ds, trg, nt, n = _tri_args(tri, cent, charge, dipstr, target)
    z = np.array([complex(zk)], np.complex128)
    pot = np.zeros(n, np.complex128)
    fld = np.zeros((n, 3), np.complex128)
    pt = np.zeros(max(nt, 1), np.complex128)
    ft = np.zeros((max(nt, 1), 3), np.complex128)
    r.h3dtriadirect_(_i(nqtri), _vp(z), _i(n), _vp(t), _vp(nrm), _vp(c),
                     _i(ifcharge if charge is not None else 0), _vp(q),
                     _i(ifdipole if dipstr is not None else 0), _vp(ds), _vp(nrm), _i(1), _vp(pot), _i(1),
                     _vp(fld), _i(nt), _vp(trg), _i(1 if nt else 0), _vp(pt), _i(1 if nt else 0), _vp(ft))
    return pot, fld, pt[:nt], ft[:nt]


def ref_hfmm3dtriatarg(iprec, zk, tri, cent, charge=None, dipstr=None, target=None, ifcharge=1, ifdipole=1):
    """The reference's Helmholtz triangle FMM hfmm3dtriatarg
    (src/hfmm3dtria.f:133)."""
    r = ref()
    t, nrm, c, q, ds, trg, nt, n = _tri_args(tri, cent, charge, dipstr, target)
    z = np.array([complex(zk)], np.complex128)
    pot = np.zeros(n, np.complex128)
    fld = np.zeros((n, 3), np.complex128)
    pt = np.zeros(max(nt, 1), np.complex128)
    ft = np.zeros((max(nt, 1), 3), np.complex128)
    ier = ctypes.c_int(0)
    r.hfmm3dtriatarg_(ctypes.byref(ier), _i(iprec), _vp(z), _i(n), _vp(t), _vp(nrm), _vp(c),
                      _i(ifcharge if charge is not None else 0), _vp(q),
                      _i(ifdipole if dipstr is not None else 0), _vp(ds), _vp(nrm), _i(1), _vp(pot), _i(1),
                      _vp(fld), _i(nt), _vp(trg), _i(1 if nt else 0), _vp(pt), _i(1 if nt else 0), _vp(ft))
    return ier.value, pot, fld, pt[:nt], ft[:nt]


# --------------------------------------------------------------------------
# quadrupole sources (oracle/fmm_oracle_qp.c; laprouts3d.f:7914, :8261, :8691)
# --------------------------------------------------------------------------
def orc_qp_ladder(n, m, irregular):
    o = oracle()
    L = np.zeros(5)
    (o.orc_qp_ladder_irregular if irregular else o.orc_qp_ladder_regular)(int(n), int(m), _vp(L))
    return L


def orc_l3dform_qp(irregular, rscale, source, quadstr, quadvec, center, p):
    """Packed multipole (irregular=False) or local (True) expansion of order p
    of quadrupoles; quadvec (n, 6) ordered xx, yy, zz, xy, xz, yz."""
    o = oracle()
    src = _c(np.atleast_2d(source), np.float64)
    qs = _c(np.atleast_1d(quadstr), np.complex128)
    qv = _c(np.atleast_2d(quadvec), np.float64)
    c = _c(center, np.float64)
    out = np.zeros((p + 1) ** 2, np.complex128)
    o.orc_l3dform_qp(int(bool(irregular)), ctypes.c_double(rscale), _vp(src), _vp(qs), _vp(qv), src.shape[0],
                     _vp(c), int(p), _vp(out))
    return out


def orc_direct_qp_sample(source, quadstr, quadvec, points, self_index):
    """FP64 pot and fld at `points` of quadrupole sources; self_index[j] >= 0
    skips that source."""
    o = oracle()
    src = _c(source, np.float64)
    qs = _c(quadstr, np.complex128)
    qv = _c(quadvec, np.float64)
    pts = _c(points, np.float64)
    si = _c(self_index, np.int32)
    pot = np.zeros(pts.shape[0], np.complex128)
    fld = np.zeros((pts.shape[0], 3), np.complex128)
    o.orc_l3ddirect_qp_sample(src.shape[0], _vp(src), _vp(qs), _vp(qv), pts.shape[0], _vp(pts), _vp(si), _vp(pot),
                              _vp(fld))
    return pot, fld


def ref_lpotfld3dhess_qp(source, quadstr, quadvec, target):
    """The reference's lpotfld3dhess_qp (one quadrupole, one target): pot, fld, hess."""
    r = ref()
    s, t, qv = _c(source, np.float64), _c(target, np.float64), _c(quadvec, np.float64)
    qs = np.array([quadstr], np.complex128)
    pot = np.zeros(1, np.complex128)
    fld = np.zeros(3, np.complex128)
    hs = np.zeros(6, np.complex128)
    one = ctypes.c_int(1)
    r.lpotfld3dhess_qp_(ctypes.byref(one), ctypes.byref(one), _vp(s), _vp(qs), _vp(qv), _vp(t), _vp(pot), _vp(fld),
                        _vp(hs))
    return pot[0], fld, hs


def ref_l3dform_qp(irregular, rscale, source, quadstr, quadvec, center, p):
    """The reference's l3dformmp_qp_trunc / l3dformta_qp_trunc, packed."""
    r = ref()
    src = _c(np.atleast_2d(source), np.float64)
    qs = _c(np.atleast_1d(quadstr), np.complex128)
    qv = _c(np.atleast_2d(quadvec), np.float64)
    c = _c(center, np.float64)
    nlege = p + 10
    lw = 4 * (nlege + 1) ** 2 + 100
    w = np.zeros(lw)
    lused = ctypes.c_int(0)
    r.ylgndrfwini_(_i(nlege), _vp(w), _i(lw), ctypes.byref(lused))
    A = np.zeros((2 * p + 1) * (p + 1), np.complex128)
    ier = ctypes.c_int(0)
    fn = r.l3dformta_qp_trunc_ if irregular else r.l3dformmp_qp_trunc_
    fn(ctypes.byref(ier), _d(rscale), _vp(src), _vp(qs), _vp(qv), _i(src.shape[0]), _vp(c), _i(p), _vp(A), _vp(w),
       _i(nlege))
    return ref_to_packed(A, p)
