"""fmmlib3d_amd: MI355X (gfx950) particle FMM with fmmlib3d's calling sequences.

Python mirror of the reference routines (src/lfmm3dpart.f, src/hfmm3dpart.f)
over the C ABI of libfmm3d_mi355x.so (include/fmm3d_mi355x.h).  Arguments and
flags mean what they mean in the reference; like the reference's f2py
intents (lfmm3dpart.f:205-226) the outputs are returned:

    ier, pot, fld, pottarg, fldtarg = lfmm3dparttarg(iprec, source, charge=q,
                                                     target=t)

Arrays: source/target (n,3) float64, charge/dipstr (n,) complex128,
dipvec (n,3) float64; pot (n,) complex128, fld (n,3) complex128.
There is no CPU fallback: when the HIP library is missing or no GPU is
visible the calls raise.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libfmm3d_mi355x.so")

PREC_FP64 = 0     # every stage in FP64
PREC_MIXED = 1    # FP32 near-field pair arithmetic, FP64 far field
PREC_FAR32 = 2    # FP32 expansions/translations, FP64 near field
PREC_FP32 = 3     # FP32 near and far field

# every symbol include/fmm3d_mi355x.h declares
EXPORTED = (
    "fmm3d_lfmm3dparttarg", "fmm3d_lfmm3dpartself", "fmm3d_l3dpartdirect",
    "fmm3d_lfmm3dparttarg_ex", "fmm3d_lfmm3dparttarg_dev", "fmm3d_lfmm3dparttarg_dev_noalloc",
    "fmm3d_lfmm3dparttarg_hess", "fmm3d_lfmm3dparttarg_qp",
    "fmm3d_last_timings",
    "fmm3d_last_precision", "fmm3d_version", "fmm3d_d3tstrcr",
    "fmm3d_hfmm3dparttarg", "fmm3d_hfmm3dpartself", "fmm3d_h3dpartdirect",
    "fmm3d_hfmm3dparttarg_ex", "fmm3d_hfmm3dparttarg_dev",
    "fmm3d_server_create", "fmm3d_server_submit", "fmm3d_server_wait", "fmm3d_server_stats",
    "fmm3d_server_destroy",
    "fmm3d_lfmm3dtriatarg", "fmm3d_lfmm3dtriaself", "fmm3d_lfmm3dtriampftarg", "fmm3d_l3dtriadirect",
    "fmm3d_hfmm3dtriatarg", "fmm3d_hfmm3dtriaself", "fmm3d_hfmm3dtriampftarg", "fmm3d_h3dtriadirect",
)

_lib = None


class FMMError(RuntimeError):
    def __init__(self, ier, msg=""):
        super().__init__("ier=%d %s" % (ier, msg))
        self.ier = ier


def load():
    """Load the native library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libfmm3d_mi355x.so not built: run `python -m fmmlib3d_amd.build`")
        lib = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        i = ctypes.c_int
        lib.fmm3d_version.restype = ctypes.c_char_p
        targ = [i, i, vp, i, vp, i, vp, vp, i, vp, i, vp, i, vp, i, vp, i, vp]
        lib.fmm3d_lfmm3dparttarg.argtypes = targ
        lib.fmm3d_lfmm3dparttarg.restype = i
        lib.fmm3d_lfmm3dparttarg_ex.argtypes = [i] + targ
        lib.fmm3d_lfmm3dparttarg_ex.restype = i
        lib.fmm3d_lfmm3dparttarg_dev.argtypes = [i] + targ + [vp]
        lib.fmm3d_lfmm3dparttarg_dev.restype = i
        lib.fmm3d_lfmm3dparttarg_dev_noalloc.argtypes = [i] + targ + [vp]
        lib.fmm3d_lfmm3dparttarg_dev_noalloc.restype = i
        lib.fmm3d_lfmm3dparttarg_hess.argtypes = [i, i, vp, i, vp, i, vp, vp, i, vp, i, vp, i, vp, i, vp, i, vp,
                                                  i, vp, i, vp]
        lib.fmm3d_lfmm3dparttarg_hess.restype = i
        lib.fmm3d_lfmm3dparttarg_qp.argtypes = [i, i, vp, i, vp, i, vp, vp, i, vp, vp, i, vp, i, vp, i, vp, i, vp,
                                                i, vp]
        lib.fmm3d_lfmm3dparttarg_qp.restype = i
        lib.fmm3d_lfmm3dpartself.argtypes = [i, i, vp, i, vp, i, vp, vp, i, vp, i, vp]
        lib.fmm3d_lfmm3dpartself.restype = i
        lib.fmm3d_l3dpartdirect.argtypes = [i, vp, i, vp, i, vp, vp, i, vp, i, vp, i, vp, i, vp, i, vp]
        lib.fmm3d_l3dpartdirect.restype = None
        lib.fmm3d_d3tstrcr.argtypes = [i, vp, i, vp, i, vp, vp, vp, vp, vp, i, vp, vp, vp, vp, i, vp]
        lib.fmm3d_d3tstrcr.restype = i
        htarg = [i, vp] + targ[1:]
        lib.fmm3d_hfmm3dparttarg.argtypes = htarg
        lib.fmm3d_hfmm3dparttarg.restype = i
        lib.fmm3d_hfmm3dparttarg_ex.argtypes = [i] + htarg
        lib.fmm3d_hfmm3dparttarg_ex.restype = i
        lib.fmm3d_hfmm3dparttarg_dev.argtypes = [i] + htarg + [vp]
        lib.fmm3d_hfmm3dparttarg_dev.restype = i
        lib.fmm3d_hfmm3dpartself.argtypes = [i, vp, i, vp, i, vp, i, vp, vp, i, vp, i, vp]
        lib.fmm3d_hfmm3dpartself.restype = i
        lib.fmm3d_h3dpartdirect.argtypes = [vp, i, vp, i, vp, i, vp, vp, i, vp, i, vp, i, vp, i, vp, i, vp]
        lib.fmm3d_h3dpartdirect.restype = None
        ll = ctypes.c_longlong
        lib.fmm3d_server_create.argtypes = [i, ll]
        lib.fmm3d_server_create.restype = vp
        lib.fmm3d_server_submit.argtypes = [vp, i, i, i, vp, i, vp, i, vp, i, vp, vp, i, i, i, vp, i, i]
        lib.fmm3d_server_submit.restype = ll
        lib.fmm3d_server_wait.argtypes = [vp, ll, i, vp, vp, vp, vp, vp, vp]
        lib.fmm3d_server_wait.restype = i
        lib.fmm3d_server_stats.argtypes = [vp, vp, vp, vp, vp]
        lib.fmm3d_server_stats.restype = None
        lib.fmm3d_server_destroy.argtypes = [vp]
        lib.fmm3d_server_destroy.restype = None
        lib.fmm3d_last_timings.argtypes = [vp, i]
        lib.fmm3d_last_timings.restype = None
        lib.fmm3d_last_precision.argtypes = []
        lib.fmm3d_last_precision.restype = i
        ttarg = [i, i, vp, vp, vp, i, vp, i, vp, vp, i, vp, i, vp, i, vp, i, vp, i, vp]
        lib.fmm3d_lfmm3dtriatarg.argtypes = ttarg
        lib.fmm3d_lfmm3dtriatarg.restype = i
        lib.fmm3d_lfmm3dtriaself.argtypes = ttarg[:14]
        lib.fmm3d_lfmm3dtriaself.restype = i
        lib.fmm3d_lfmm3dtriampftarg.argtypes = ttarg[:10] + ttarg[14:]
        lib.fmm3d_lfmm3dtriampftarg.restype = i
        lib.fmm3d_l3dtriadirect.argtypes = ttarg[1:]
        lib.fmm3d_l3dtriadirect.restype = None
        htt = [i, vp] + ttarg[1:]
        lib.fmm3d_hfmm3dtriatarg.argtypes = htt
        lib.fmm3d_hfmm3dtriatarg.restype = i
        lib.fmm3d_hfmm3dtriaself.argtypes = htt[:15]
        lib.fmm3d_hfmm3dtriaself.restype = i
        lib.fmm3d_hfmm3dtriampftarg.argtypes = htt[:11] + htt[15:]
        lib.fmm3d_hfmm3dtriampftarg.restype = i
        lib.fmm3d_h3dtriadirect.argtypes = htt
        lib.fmm3d_h3dtriadirect.restype = None
        _lib = lib
    return _lib


def version():
    return load().fmm3d_version().decode()


def last_precision():
    """The PREC_* mode the last FMM call of this thread ran in (PREC_FP64
    when the requested mode fell back to FP64 at iprec >= 3)."""
    return int(load().fmm3d_last_precision())


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _in(a, dt, shape_last=None):
    if a is None:
        return None
    a = np.ascontiguousarray(a, dtype=dt)
    return a


def _prep(source, charge, dipstr, dipvec, target):
    src = _in(source, np.float64)
    if src.ndim != 2 or src.shape[1] != 3:
        raise ValueError("source must be (n,3)")
    ns = src.shape[0]
    q = _in(charge, np.complex128)
    ds = _in(dipstr, np.complex128)
    dv = _in(dipvec, np.float64)
    if q is not None and q.shape != (ns,):
        raise ValueError("charge must be (nsource,)")
    if (ds is None) != (dv is None):
        raise ValueError("dipstr and dipvec go together")
    if ds is not None and (ds.shape != (ns,) or dv.shape != (ns, 3)):
        raise ValueError("dipstr (n,) and dipvec (n,3)")
    trg = _in(target, np.float64)
    if trg is not None and (trg.ndim != 2 or trg.shape[1] != 3):
        raise ValueError("target must be (n,3)")
    return src, q, ds, dv, trg


def lfmm3dparttarg(iprec, source, charge=None, dipstr=None, dipvec=None, target=None,
                   ifpot=1, iffld=1, ifpottarg=1, iffldtarg=1, precision=PREC_FP64):
    """Laplace FMM (lfmm3dparttarg, src/lfmm3dpart.f:138).  Returns
    (ier, pot, fld, pottarg, fldtarg); outputs whose flag is 0 are None."""
    lib = load()
    src, q, ds, dv, trg = _prep(source, charge, dipstr, dipvec, target)
    ns = src.shape[0]
    nt = 0 if trg is None else trg.shape[0]
    pot = np.zeros(ns, np.complex128) if ifpot else None
    fld = np.zeros((ns, 3), np.complex128) if iffld else None
    pt = np.zeros(nt, np.complex128) if (ifpottarg and nt) else None
    ft = np.zeros((nt, 3), np.complex128) if (iffldtarg and nt) else None
    ier = lib.fmm3d_lfmm3dparttarg_ex(
        int(precision), int(iprec), ns, _ptr(src), int(q is not None), _ptr(q),
        int(ds is not None), _ptr(ds), _ptr(dv), int(bool(ifpot)), _ptr(pot), int(bool(iffld)),
        _ptr(fld), nt, _ptr(trg), int(pt is not None), _ptr(pt), int(ft is not None), _ptr(ft))
    return ier, pot, fld, pt, ft


def lfmm3dparttarg_hess(iprec, source, charge=None, dipstr=None, dipvec=None, target=None,
                        ifpot=1, iffld=1, ifhess=1, ifpottarg=1, iffldtarg=1, ifhesstarg=1):
    """Laplace FMM with Hessians (FP64).  Returns (ier, pot, fld, hess,
    pottarg, fldtarg, hesstarg); hess is (n, 6) complex, per point (pot_xx,
    pot_yy, pot_zz, pot_xy, pot_xz, pot_yz), the order of the reference's
    lpotfld3dhess (src/laprouts3d.f:5442); outputs whose flag is 0 are None."""
    lib = load()
    src, q, ds, dv, trg = _prep(source, charge, dipstr, dipvec, target)
    ns = src.shape[0]
    nt = 0 if trg is None else trg.shape[0]
    pot = np.zeros(ns, np.complex128) if ifpot else None
    fld = np.zeros((ns, 3), np.complex128) if iffld else None
    hs = np.zeros((ns, 6), np.complex128) if ifhess else None
    pt = np.zeros(nt, np.complex128) if (ifpottarg and nt) else None
    ft = np.zeros((nt, 3), np.complex128) if (iffldtarg and nt) else None
    ht = np.zeros((nt, 6), np.complex128) if (ifhesstarg and nt) else None
    ier = lib.fmm3d_lfmm3dparttarg_hess(
        int(iprec), ns, _ptr(src), int(q is not None), _ptr(q), int(ds is not None), _ptr(ds), _ptr(dv),
        int(bool(ifpot)), _ptr(pot), int(bool(iffld)), _ptr(fld), int(hs is not None), _ptr(hs),
        nt, _ptr(trg), int(pt is not None), _ptr(pt), int(ft is not None), _ptr(ft), int(ht is not None), _ptr(ht))
    return ier, pot, fld, hs, pt, ft, ht


def lfmm3dparttarg_qp(iprec, source, charge=None, dipstr=None, dipvec=None, quadstr=None, quadvec=None,
                      target=None, ifpot=1, iffld=1, ifpottarg=1, iffldtarg=1):
    """Laplace FMM with quadrupole sources as well (FP64).  quadstr (n,)
    complex and quadvec (n, 6) real, ordered (xx, yy, zz, xy, xz, yz), the
    source of the reference's lpotfld3dhess_qp (src/laprouts3d.f:8691).
    Returns (ier, pot, fld, pottarg, fldtarg); outputs whose flag is 0 are None."""
    lib = load()
    src, q, ds, dv, trg = _prep(source, charge, dipstr, dipvec, target)
    ns = src.shape[0]
    nt = 0 if trg is None else trg.shape[0]
    qs = _in(quadstr, np.complex128)
    qv = _in(quadvec, np.float64)
    if (qs is None) != (qv is None):
        raise ValueError("quadstr and quadvec go together")
    if qs is not None and (qs.shape != (ns,) or qv.shape != (ns, 6)):
        raise ValueError("quadstr (n,) and quadvec (n,6)")
    pot = np.zeros(ns, np.complex128) if ifpot else None
    fld = np.zeros((ns, 3), np.complex128) if iffld else None
    pt = np.zeros(nt, np.complex128) if (ifpottarg and nt) else None
    ft = np.zeros((nt, 3), np.complex128) if (iffldtarg and nt) else None
    ier = lib.fmm3d_lfmm3dparttarg_qp(
        int(iprec), ns, _ptr(src), int(q is not None), _ptr(q), int(ds is not None), _ptr(ds), _ptr(dv),
        int(qs is not None), _ptr(qs), _ptr(qv), int(bool(ifpot)), _ptr(pot), int(bool(iffld)), _ptr(fld),
        nt, _ptr(trg), int(pt is not None), _ptr(pt), int(ft is not None), _ptr(ft))
    return ier, pot, fld, pt, ft


def _tri_prep(triaflat, charge, dipstr, target, source):
    tri = np.ascontiguousarray(triaflat, np.float64).reshape(-1, 3, 3)   # (n, vertex, xyz)
    n = tri.shape[0]
    v0, v1, v2 = tri[:, 0], tri[:, 1], tri[:, 2]
    # centroids ((v0 + v1) + v2) / 3 and unit normals (v1 - v0) x (v2 - v0), by columns
    # (numpy's mean over the vertex axis and cross take 10x longer on these shapes)
    cen = (v0 + v1 + v2) / 3.0 if source is None else np.ascontiguousarray(source, np.float64).reshape(n, 3)
    a, b = v1 - v0, v2 - v0
    nrm = np.empty((n, 3))
    nrm[:, 0] = a[:, 1] * b[:, 2] - a[:, 2] * b[:, 1]
    nrm[:, 1] = a[:, 2] * b[:, 0] - a[:, 0] * b[:, 2]
    nrm[:, 2] = a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]
    nrm /= np.sqrt(nrm[:, 0] ** 2 + nrm[:, 1] ** 2 + nrm[:, 2] ** 2)[:, None]
    q = None if charge is None else np.ascontiguousarray(charge, np.complex128).reshape(n)
    ds = None if dipstr is None else np.ascontiguousarray(dipstr, np.complex128).reshape(n)
    trg = None if target is None else np.ascontiguousarray(target, np.float64).reshape(-1, 3)
    return np.ascontiguousarray(tri), cen, nrm, q, ds, trg


def lfmm3dtriatarg(iprec, triaflat, charge=None, dipstr=None, target=None, source=None,
                   ifpot=1, iffld=1, ifpottarg=1, iffldtarg=1):
    """Triangle FMM (lfmm3dtriatarg, src/lfmm3dtria.f:121): triaflat is
    (n, 3, 3) = per triangle its three vertices; charge / dipstr the
    single- and dipole-layer densities (the dipole layer along the normal
    (v2-v1)x(v3-v1)); pot/fld at `source` (default: the centroids) and at
    target.  Returns (ier, pot, fld, pottarg, fldtarg)."""
    lib = load()
    tri, cen, nrm, q, ds, trg = _tri_prep(triaflat, charge, dipstr, target, source)
    n = tri.shape[0]
    nt = 0 if trg is None else trg.shape[0]
    pot = np.zeros(n, np.complex128)
    fld = np.zeros((n, 3), np.complex128)
    pt = np.zeros(max(nt, 1), np.complex128)
    ft = np.zeros((max(nt, 1), 3), np.complex128)
    ier = lib.fmm3d_lfmm3dtriatarg(
        int(iprec), n, _ptr(tri), _ptr(nrm), _ptr(cen), int(q is not None), _ptr(q), int(ds is not None), _ptr(ds),
        _ptr(nrm), int(bool(ifpot)), _ptr(pot), int(bool(iffld)), _ptr(fld), nt, _ptr(trg),
        int(bool(ifpottarg) and nt > 0), _ptr(pt), int(bool(iffldtarg) and nt > 0), _ptr(ft))
    return ier, pot, fld, pt[:nt], ft[:nt]


def lfmm3dtriaself(iprec, triaflat, charge=None, dipstr=None, source=None, ifpot=1, iffld=1):
    """Triangle FMM without targets (lfmm3dtriaself): (ier, pot, fld)."""
    g = lfmm3dtriatarg(iprec, triaflat, charge, dipstr, None, source, ifpot, iffld, 0, 0)
    return g[0], g[1], g[2]


def lfmm3dtriampftarg(iprec, triaflat, charge=None, dipstr=None, target=None, source=None):
    """Triangle FMM with outputs at the targets only (lfmm3dtriampftarg):
    (ier, pottarg, fldtarg)."""
    lib = load()
    tri, cen, nrm, q, ds, trg = _tri_prep(triaflat, charge, dipstr, target, source)
    n = tri.shape[0]
    nt = 0 if trg is None else trg.shape[0]
    pt = np.zeros(max(nt, 1), np.complex128)
    ft = np.zeros((max(nt, 1), 3), np.complex128)
    ier = lib.fmm3d_lfmm3dtriampftarg(int(iprec), n, _ptr(tri), _ptr(nrm), _ptr(cen), int(q is not None), _ptr(q),
                                      int(ds is not None), _ptr(ds), _ptr(nrm), nt, _ptr(trg), int(nt > 0), _ptr(pt),
                                      int(nt > 0), _ptr(ft))
    return ier, pt[:nt], ft[:nt]


def l3dtriadirect(triaflat, charge=None, dipstr=None, target=None, source=None):
    """Direct triangle sum on the device (l3dtriadirect, src/lfmm3dtria.f:1385):
    (pot, fld, pottarg, fldtarg)."""
    lib = load()
    tri, cen, nrm, q, ds, trg = _tri_prep(triaflat, charge, dipstr, target, source)
    n = tri.shape[0]
    nt = 0 if trg is None else trg.shape[0]
    pot = np.zeros(n, np.complex128)
    fld = np.zeros((n, 3), np.complex128)
    pt = np.zeros(max(nt, 1), np.complex128)
    ft = np.zeros((max(nt, 1), 3), np.complex128)
    lib.fmm3d_l3dtriadirect(n, _ptr(tri), _ptr(nrm), _ptr(cen), int(q is not None), _ptr(q), int(ds is not None),
                            _ptr(ds), _ptr(nrm), 1, _ptr(pot), 1, _ptr(fld), nt, _ptr(trg), int(nt > 0), _ptr(pt),
                            int(nt > 0), _ptr(ft))
    return pot, fld, pt[:nt], ft[:nt]


def hfmm3dtriatarg(iprec, zk, triaflat, charge=None, dipstr=None, target=None, source=None,
                   ifpot=1, iffld=1, ifpottarg=1, iffldtarg=1, ifcharge=1, ifdipole=1):
    """Helmholtz triangle FMM (hfmm3dtriatarg, src/hfmm3dtria.f:133), the
    arguments as lfmm3dtriatarg plus the wavenumber zk; ifcharge / ifdipole
    = 2 subtract that density's Laplace part (the difference kernel).
    Returns (ier, pot, fld, pottarg, fldtarg)."""
    lib = load()
    tri, cen, nrm, q, ds, trg = _tri_prep(triaflat, charge, dipstr, target, source)
    n = tri.shape[0]
    nt = 0 if trg is None else trg.shape[0]
    pot = np.zeros(n, np.complex128)
    fld = np.zeros((n, 3), np.complex128)
    pt = np.zeros(max(nt, 1), np.complex128)
    ft = np.zeros((max(nt, 1), 3), np.complex128)
    z = _zk(zk)
    ier = lib.fmm3d_hfmm3dtriatarg(
        int(iprec), _ptr(z), n, _ptr(tri), _ptr(nrm), _ptr(cen), int(ifcharge) if q is not None else 0, _ptr(q),
        int(ifdipole) if ds is not None else 0, _ptr(ds), _ptr(nrm), int(bool(ifpot)), _ptr(pot),
        int(bool(iffld)), _ptr(fld), nt, _ptr(trg), int(bool(ifpottarg) and nt > 0), _ptr(pt),
        int(bool(iffldtarg) and nt > 0), _ptr(ft))
    return ier, pot, fld, pt[:nt], ft[:nt]


def hfmm3dtriaself(iprec, zk, triaflat, charge=None, dipstr=None, source=None, ifpot=1, iffld=1, ifcharge=1,
                   ifdipole=1):
    """Helmholtz triangle FMM without targets (hfmm3dtriaself): (ier, pot, fld)."""
    g = hfmm3dtriatarg(iprec, zk, triaflat, charge, dipstr, None, source, ifpot, iffld, 0, 0, ifcharge, ifdipole)
    return g[0], g[1], g[2]


def hfmm3dtriampftarg(iprec, zk, triaflat, charge=None, dipstr=None, target=None, source=None, ifcharge=1,
                      ifdipole=1):
    """Helmholtz triangle FMM with outputs at the targets only
    (hfmm3dtriampftarg): (ier, pottarg, fldtarg)."""
    lib = load()
    tri, cen, nrm, q, ds, trg = _tri_prep(triaflat, charge, dipstr, target, source)
    n = tri.shape[0]
    nt = 0 if trg is None else trg.shape[0]
    pt = np.zeros(max(nt, 1), np.complex128)
    ft = np.zeros((max(nt, 1), 3), np.complex128)
    z = _zk(zk)
    ier = lib.fmm3d_hfmm3dtriampftarg(int(iprec), _ptr(z), n, _ptr(tri), _ptr(nrm), _ptr(cen),
                                      int(ifcharge) if q is not None else 0, _ptr(q),
                                      int(ifdipole) if ds is not None else 0, _ptr(ds), _ptr(nrm), nt, _ptr(trg),
                                      int(nt > 0), _ptr(pt), int(nt > 0), _ptr(ft))
    return ier, pt[:nt], ft[:nt]


def h3dtriadirect(nqtri, zk, triaflat, charge=None, dipstr=None, target=None, source=None, ifcharge=1,
                  ifdipole=1):
    """Direct Helmholtz triangle sum on the device (h3dtriadirect,
    src/hfmm3dtria.f:1456) with the symmetric rule of degree nqtri:
    (pot, fld, pottarg, fldtarg)."""
    lib = load()
    tri, cen, nrm, q, ds, trg = _tri_prep(triaflat, charge, dipstr, target, source)
    n = tri.shape[0]
    nt = 0 if trg is None else trg.shape[0]
    pot = np.zeros(n, np.complex128)
    fld = np.zeros((n, 3), np.complex128)
    pt = np.zeros(max(nt, 1), np.complex128)
    ft = np.zeros((max(nt, 1), 3), np.complex128)
    z = _zk(zk)
    lib.fmm3d_h3dtriadirect(int(nqtri), _ptr(z), n, _ptr(tri), _ptr(nrm), _ptr(cen),
                            int(ifcharge) if q is not None else 0, _ptr(q), int(ifdipole) if ds is not None else 0,
                            _ptr(ds), _ptr(nrm), 1, _ptr(pot), 1, _ptr(fld), nt, _ptr(trg), int(nt > 0), _ptr(pt),
                            int(nt > 0), _ptr(ft))
    return pot, fld, pt[:nt], ft[:nt]


def lfmm3dpartself(iprec, source, charge=None, dipstr=None, dipvec=None, ifpot=1, iffld=1,
                   precision=PREC_FP64):
    """Laplace FMM at the sources only (lfmm3dpartself, src/lfmm3dpart.f:87)."""
    ier, pot, fld, _, _ = lfmm3dparttarg(iprec, source, charge, dipstr, dipvec, None, ifpot, iffld,
                                         0, 0, precision)
    return ier, pot, fld


def l3dpartdirect(source, charge=None, dipstr=None, dipvec=None, target=None,
                  ifpot=1, iffld=1, ifpottarg=1, iffldtarg=1):
    """O(N^2) direct sum on the GPU (l3dpartdirect, src/lfmm3dpart.f:1391)."""
    lib = load()
    src, q, ds, dv, trg = _prep(source, charge, dipstr, dipvec, target)
    ns = src.shape[0]
    nt = 0 if trg is None else trg.shape[0]
    pot = np.zeros(ns, np.complex128) if ifpot else None
    fld = np.zeros((ns, 3), np.complex128) if iffld else None
    pt = np.zeros(nt, np.complex128) if (ifpottarg and nt) else None
    ft = np.zeros((nt, 3), np.complex128) if (iffldtarg and nt) else None
    lib.fmm3d_l3dpartdirect(ns, _ptr(src), int(q is not None), _ptr(q), int(ds is not None),
                            _ptr(ds), _ptr(dv), int(bool(ifpot)), _ptr(pot), int(bool(iffld)),
                            _ptr(fld), nt, _ptr(trg), int(pt is not None), _ptr(pt),
                            int(ft is not None), _ptr(ft))
    return pot, fld, pt, ft


def _zk(zk):
    return np.array([complex(zk).real, complex(zk).imag], dtype=np.float64)


def hfmm3dparttarg(iprec, zk, source, charge=None, dipstr=None, dipvec=None, target=None,
                   ifpot=1, iffld=1, ifpottarg=1, iffldtarg=1, precision=PREC_FP64):
    """Helmholtz FMM, kernel e^{ik r}/r (hfmm3dparttarg, src/hfmm3dpart.f:141).
    Returns (ier, pot, fld, pottarg, fldtarg); outputs whose flag is 0 are None."""
    lib = load()
    src, q, ds, dv, trg = _prep(source, charge, dipstr, dipvec, target)
    ns = src.shape[0]
    nt = 0 if trg is None else trg.shape[0]
    pot = np.zeros(ns, np.complex128) if ifpot else None
    fld = np.zeros((ns, 3), np.complex128) if iffld else None
    pt = np.zeros(nt, np.complex128) if (ifpottarg and nt) else None
    ft = np.zeros((nt, 3), np.complex128) if (iffldtarg and nt) else None
    z = _zk(zk)
    ier = lib.fmm3d_hfmm3dparttarg_ex(
        int(precision), int(iprec), _ptr(z), ns, _ptr(src), int(q is not None), _ptr(q),
        int(ds is not None), _ptr(ds), _ptr(dv), int(bool(ifpot)), _ptr(pot), int(bool(iffld)),
        _ptr(fld), nt, _ptr(trg), int(pt is not None), _ptr(pt), int(ft is not None), _ptr(ft))
    return ier, pot, fld, pt, ft


def hfmm3dpartself(iprec, zk, source, charge=None, dipstr=None, dipvec=None, ifpot=1, iffld=1,
                   precision=PREC_FP64):
    """Helmholtz FMM at the sources only (hfmm3dpartself, src/hfmm3dpart.f:90)."""
    ier, pot, fld, _, _ = hfmm3dparttarg(iprec, zk, source, charge, dipstr, dipvec, None, ifpot,
                                         iffld, 0, 0, precision)
    return ier, pot, fld


def h3dpartdirect(zk, source, charge=None, dipstr=None, dipvec=None, target=None,
                  ifpot=1, iffld=1, ifpottarg=1, iffldtarg=1):
    """O(N^2) Helmholtz direct sum on the GPU (h3dpartdirect, src/hfmm3dpart.f:1268)."""
    lib = load()
    src, q, ds, dv, trg = _prep(source, charge, dipstr, dipvec, target)
    ns = src.shape[0]
    nt = 0 if trg is None else trg.shape[0]
    pot = np.zeros(ns, np.complex128) if ifpot else None
    fld = np.zeros((ns, 3), np.complex128) if iffld else None
    pt = np.zeros(nt, np.complex128) if (ifpottarg and nt) else None
    ft = np.zeros((nt, 3), np.complex128) if (iffldtarg and nt) else None
    z = _zk(zk)
    lib.fmm3d_h3dpartdirect(_ptr(z), ns, _ptr(src), int(q is not None), _ptr(q), int(ds is not None),
                            _ptr(ds), _ptr(dv), int(bool(ifpot)), _ptr(pot), int(bool(iffld)),
                            _ptr(fld), nt, _ptr(trg), int(pt is not None), _ptr(pt),
                            int(ft is not None), _ptr(ft))
    return pot, fld, pt, ft


def d3tstrcr(source, target, nbox):
    """The host octree (d3tstrcr, src/d3tstrcr.f:24): returns
    (ier, boxes20, isource, itarget, lists, nlev, center, size) with lists[k][b]
    a tuple of 1-based box ids (k = 0..3 for lists 1..4)."""
    lib = load()
    src = np.ascontiguousarray(source, np.float64)
    trg = None if target is None else np.ascontiguousarray(target, np.float64)
    ns = src.shape[0]
    nt = 0 if trg is None else trg.shape[0]
    nb = ctypes.c_int(0)
    nlev = ctypes.c_int(0)
    nlist = ctypes.c_int(0)
    center = np.zeros(3)
    size = ctypes.c_double(0)
    ier = lib.fmm3d_d3tstrcr(ns, _ptr(src), nt, _ptr(trg), int(nbox), ctypes.byref(nb),
                             ctypes.byref(nlev), _ptr(center), ctypes.byref(size),
                             ctypes.byref(nlist), 0, None, None, None, None, 0, None)
    if ier:
        return ier, None, None, None, None, 0, center, 0.0
    M = nb.value
    boxes = np.zeros((M, 20), np.int32)
    isrc = np.zeros(ns, np.int32)
    itrg = np.zeros(max(nt, 1), np.int32)
    counts = np.zeros(4 * M, np.int32)
    lists = np.zeros(max(nlist.value, 1), np.int32)
    ier = lib.fmm3d_d3tstrcr(ns, _ptr(src), nt, _ptr(trg), int(nbox), ctypes.byref(nb),
                             ctypes.byref(nlev), _ptr(center), ctypes.byref(size),
                             ctypes.byref(nlist), M, _ptr(boxes), _ptr(isrc), _ptr(itrg),
                             _ptr(counts), lists.size, _ptr(lists))
    out = [[() for _ in range(M)] for _ in range(4)]
    o = 0
    for k in range(4):
        for b in range(M):
            c = counts[k * M + b]
            out[k][b] = tuple(lists[o:o + c].tolist())
            o += c
    return ier, boxes, isrc, itrg[:nt], out, nlev.value, center, size.value


def last_timings():
    """Per-stage milliseconds of the last call on this thread (see
    fmm3d_last_timings in include/fmm3d_mi355x.h)."""
    out = np.zeros(13)
    load().fmm3d_last_timings(_ptr(out), 13)
    keys = ("tree_plan", "upload", "p2m", "p2l", "m2m", "m2l", "l2l", "l2p", "m2p", "p2p", "total", "tree",
            "lists")
    return dict(zip(keys, out.tolist()))


class Server:
    """Request server over the C ABI (fmm3d_server_* in include/fmm3d_mi355x.h):
    several evaluations run at once on the GPU, each worker with its own
    stream and resident device memory, admitted against a memory budget.

        with Server(nworkers=4, mem_budget=64 << 30) as srv:
            rid = srv.submit(0, 2, source, charge=q)          # Laplace, iprec 2
            ier, pot, fld, pottarg, fldtarg, latency_ms = srv.wait(rid)
    """

    def __init__(self, nworkers=4, mem_budget=64 << 30):
        self._lib = load()
        self._h = self._lib.fmm3d_server_create(int(nworkers), int(mem_budget))
        if not self._h:
            raise RuntimeError("fmm3d_server_create failed (no GPU, or bad arguments)")
        self._meta = {}

    def submit(self, kernel, iprec, source, charge=None, dipstr=None, dipvec=None, target=None,
               zk=None, ifpot=1, iffld=1, ifpottarg=1, iffldtarg=1, precision=PREC_FP64):
        """kernel 0 = Laplace, 1 = Helmholtz (zk required).  Returns a request id."""
        src, q, ds, dv, trg = _prep(source, charge, dipstr, dipvec, target)
        ns = src.shape[0]
        nt = 0 if trg is None else trg.shape[0]
        z = _zk(zk) if zk is not None else None
        rid = self._lib.fmm3d_server_submit(
            self._h, int(kernel), int(precision), int(iprec), _ptr(z), ns, _ptr(src), int(q is not None),
            _ptr(q), int(ds is not None), _ptr(ds), _ptr(dv), int(bool(ifpot)), int(bool(iffld)), nt,
            _ptr(trg), int(bool(ifpottarg) and nt > 0), int(bool(iffldtarg) and nt > 0))
        if rid < 0:
            raise FMMError(-rid, "submission rejected")
        self._meta[rid] = (ns, nt, bool(ifpot), bool(iffld), bool(ifpottarg) and nt > 0, bool(iffldtarg) and nt > 0)
        return rid

    def wait(self, rid, block=True):
        """(ier, pot, fld, pottarg, fldtarg, latency_ms), or None if not done
        and block is False."""
        ns, nt, wp, wf, wpt, wft = self._meta[rid]
        pot = np.zeros(ns, np.complex128) if wp else None
        fld = np.zeros((ns, 3), np.complex128) if wf else None
        pt = np.zeros(nt, np.complex128) if wpt else None
        ft = np.zeros((nt, 3), np.complex128) if wft else None
        ier = ctypes.c_int(0)
        lat = ctypes.c_double(0)
        r = self._lib.fmm3d_server_wait(self._h, rid, int(bool(block)), _ptr(pot), _ptr(fld), _ptr(pt), _ptr(ft),
                                        ctypes.byref(ier), ctypes.byref(lat))
        if r == 0:
            return None
        if r < 0:
            raise KeyError(rid)
        del self._meta[rid]
        return ier.value, pot, fld, pt, ft, lat.value

    def stats(self):
        q = ctypes.c_longlong(0)
        r = ctypes.c_int(0)
        b = ctypes.c_longlong(0)
        fz = ctypes.c_longlong(0)
        self._lib.fmm3d_server_stats(self._h, ctypes.byref(q), ctypes.byref(r), ctypes.byref(b), ctypes.byref(fz))
        return {"queued": q.value, "running": r.value, "bytes_in_use": b.value, "fused_requests": fz.value}

    def close(self):
        if self._h:
            self._lib.fmm3d_server_destroy(self._h)
            self._h = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
