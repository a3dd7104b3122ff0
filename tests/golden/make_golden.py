"""Generate the golden fixtures in tests/golden/ from the reference library.

The reference Fortran (fmmlib3d, compiled from its own sources by
`make -C oracle ref` into oracle/_ref/libfmmlib3d_ref.so) is run on small
seeded inputs; inputs and outputs are stored as .npz (no pickles).  Run in a
container where the reference sources are present:

    make -C oracle ref && python tests/golden/make_golden.py [case names]
"""
import ctypes
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
import oracle_bindings as ob  # noqa: E402
from oracle_bindings import _vp, _i, _d  # noqa: E402


def points(rng, n, dist):
    if dist == "cube":
        return rng.random((n, 3))
    if dist == "sphere":
        x = rng.normal(size=(n, 3))
        return x / np.linalg.norm(x, axis=1)[:, None]
    if dist == "clusters":
        c = rng.random((8, 3)) * 0.8 + 0.1
        return c[rng.integers(0, 8, n)] + 0.02 * rng.normal(size=(n, 3))
    if dist == "plane":
        x = rng.random((n, 3))
        x[:, 2] = 0.25
        return x
    raise ValueError(dist)


FMM_CASES = [
    # name, seed, iprec, ns, nt, dist, ifcharge, ifdipole
    ("lap_targ_iprec1_cube", 101, 1, 1500, 700, "cube", 1, 1),
    ("lap_self_iprec2_sphere", 102, 2, 2000, 0, "sphere", 1, 0),
    ("lap_targ_iprec0_clusters_dip", 103, 0, 1500, 800, "clusters", 0, 1),
    ("lap_self_iprec3_cube", 104, 3, 1200, 0, "cube", 1, 1),
    ("lap_targ_iprecm1_clusters", 105, -1, 1000, 1000, "clusters", 1, 1),
    ("lap_targ_rootleaf", 106, 1, 120, 30, "cube", 1, 1),
    ("lap_targ_iprecm2_plane", 107, -2, 900, 300, "plane", 1, 1),
    # p = 51 and 64 (nbox 1200, 1400): enough points for leaves at level 2,
    # so list 2 and the far field exist
    ("lap_targ_iprec4_cube", 108, 4, 12000, 2000, "cube", 1, 1),
    ("lap_self_iprec5_cube", 109, 5, 14000, 0, "cube", 1, 0),
]


def make_fmm(name, seed, iprec, ns, nt, dist, ifc, ifd):
    rng = np.random.default_rng(seed)
    src = points(rng, ns, dist)
    trg = points(rng, nt, dist) if nt else None
    q = (rng.uniform(-1, 1, ns) + 1j * rng.uniform(-1, 1, ns)) if ifc else None
    ds = (rng.uniform(-1, 1, ns) + 1j * rng.uniform(-1, 1, ns)) if ifd else None
    dv = rng.normal(size=(ns, 3)) if ifd else None
    ier, pot, fld, pt, ft = ob.ref_lfmm3dparttarg(iprec, src, q, ds, dv, trg)
    assert ier == 0, (name, ier)
    out = dict(iprec=np.int32(iprec), source=src, pot=pot, fld=fld)
    if trg is not None:
        out.update(target=trg, pottarg=pt, fldtarg=ft)
    if q is not None:
        out["charge"] = q
    if ds is not None:
        out.update(dipstr=ds, dipvec=dv)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)


def make_direct():
    rng = np.random.default_rng(201)
    src = rng.random((600, 3))
    trg = rng.random((200, 3))
    q = rng.normal(size=600) + 1j * rng.normal(size=600)
    ds = rng.normal(size=600) + 1j * rng.normal(size=600)
    dv = rng.normal(size=(600, 3))
    pot, fld, pt, ft = ob.ref_l3dpartdirect(src, q, ds, dv, trg)
    np.savez_compressed(os.path.join(HERE, "lap_direct.npz"), source=src, target=trg, charge=q,
                        dipstr=ds, dipvec=dv, pot=pot, fld=fld, pottarg=pt, fldtarg=ft)


def make_tree():
    rng = np.random.default_rng(301)
    src = points(rng, 3000, "clusters")
    trg = points(rng, 1500, "clusters")
    boxes, isrc, itrg, lists, nlev = ob.ref_tree(src, trg, 40)
    flat = []
    counts = np.zeros((5, boxes.shape[0]), np.int32)
    for k in range(5):
        for b in range(boxes.shape[0]):
            counts[k, b] = len(lists[k][b])
            flat.extend(lists[k][b])
    np.savez_compressed(os.path.join(HERE, "tree_clusters.npz"), source=src, target=trg,
                        nbox=np.int32(40), boxes=boxes, isrc=isrc, itrg=itrg, nlev=np.int32(nlev),
                        list_counts=counts, lists=np.array(flat, np.int32))


def make_operators():
    """Coefficient-level fixtures of the reference's expansion routines
    (reference layout mpole(0:p,-p:p) converted to packed n*n+n+m)."""
    r = ob.ref()
    rng = np.random.default_rng(401)
    nlege = 100
    wlege = np.zeros(40000)
    lused = ctypes.c_int(0)
    r.ylgndrfwini_(_i(nlege), _vp(wlege), _i(40000), ctypes.byref(lused))
    ier = ctypes.c_int(0)
    p = 9
    c1 = np.array([0.3, -0.2, 0.5])
    c0 = np.array([2.0, 1.1, -0.7])
    c2 = c1 + np.array([0.11, -0.07, 0.05])
    c3 = c0 + np.array([0.05, 0.02, -0.04])
    ns = 6
    src = c1 + 0.1 * (rng.random((ns, 3)) - 0.5)
    q = rng.normal(size=ns) + 1j * rng.normal(size=ns)
    ds = rng.normal(size=ns) + 1j * rng.normal(size=ns)
    dv = rng.normal(size=(ns, 3))
    trg = c0 + 0.2 * (rng.random((5, 3)) - 0.5)
    sc = 3.7
    sc2 = sc * 0.5
    L = (p + 1) * (2 * p + 1)
    z = lambda: np.zeros(L, np.complex128)  # noqa: E731
    mp, mpd, lo, lod = z(), z(), z(), z()
    r.l3dformmp_add_trunc_(ctypes.byref(ier), _d(sc), _vp(src), _vp(q), _i(ns), _vp(c1), _i(p), _i(p), _vp(mp), _vp(wlege), _i(nlege))
    r.l3dformmp_dp_add_trunc_(ctypes.byref(ier), _d(sc), _vp(src), _vp(ds), _vp(dv), _i(ns), _vp(c1), _i(p), _i(p), _vp(mpd), _vp(wlege), _i(nlege))
    r.l3dformta_add_trunc_(ctypes.byref(ier), _d(sc), _vp(src), _vp(q), _i(ns), _vp(c0), _i(p), _i(p), _vp(lo), _vp(wlege), _i(nlege))
    r.l3dformta_dp_add_trunc_(ctypes.byref(ier), _d(sc), _vp(src), _vp(ds), _vp(dv), _i(ns), _vp(c0), _i(p), _i(p), _vp(lod), _vp(wlege), _i(nlege))
    nt = trg.shape[0]
    mpot, mfld = np.zeros(nt, np.complex128), np.zeros((nt, 3), np.complex128)
    r.l3dmpevalall_trunc_(_d(sc), _vp(c1), _vp(mp), _i(p), _i(p), _vp(trg), _i(nt), _i(1), _vp(mpot), _i(1), _vp(mfld), _vp(wlege), _i(nlege), ctypes.byref(ier))
    lpot, lfld = np.zeros(nt, np.complex128), np.zeros((nt, 3), np.complex128)
    r.l3dtaevalall_trunc_(_d(sc), _vp(c0), _vp(lo), _i(p), _i(p), _vp(trg), _i(nt), _i(1), _vp(lpot), _i(1), _vp(lfld), _vp(wlege), _i(nlege), ctypes.byref(ier))
    m2m = z()
    r.l3dmpmpquadu_add_(_d(sc), _vp(c1), _vp(mp), _i(p), _d(sc2), _vp(c2), _vp(m2m), _i(p), _i(p), ctypes.byref(ier))
    m2l = z()
    r.l3dmplocquadu_add_trunc_(_d(sc), _vp(c1), _vp(mp), _i(p), _i(p), _d(sc2), _vp(c0), _vp(m2l), _i(p), _i(p), ctypes.byref(ier))
    m2l5 = z()
    r.l3dmplocquadu_add_trunc_(_d(sc), _vp(c1), _vp(mp), _i(p), _i(5), _d(sc2), _vp(c0), _vp(m2l5), _i(p), _i(5), ctypes.byref(ier))
    l2l = z()
    r.l3dloclocquadu_add_(_d(sc), _vp(c0), _vp(lo), _i(p), _d(sc2), _vp(c3), _vp(l2l), _i(p), _i(p), ctypes.byref(ier))
    pk = lambda a: ob.ref_to_packed(a, p)  # noqa: E731
    np.savez_compressed(os.path.join(HERE, "lap_operators.npz"), p=np.int32(p), rscale=sc, rscale2=sc2,
                        c0=c0, c1=c1, c2=c2, c3=c3, source=src, charge=q, dipstr=ds, dipvec=dv,
                        target=trg, p2m=pk(mp), p2m_dp=pk(mpd), p2l=pk(lo), p2l_dp=pk(lod),
                        m2p_pot=mpot, m2p_fld=mfld, l2p_pot=lpot, l2p_fld=lfld, m2m=pk(m2m),
                        m2l=pk(m2l), m2l_trunc5=pk(m2l5), l2l=pk(l2l))


def make_orders():
    r = ob.ref()
    eps = [.5e0, .5e-1, .5e-2, .5e-3, .5e-6, .5e-9, .5e-12, .5e-15]
    nterms = []
    for e in eps:
        n = ctypes.c_int(0)
        ier = ctypes.c_int(0)
        r.l3dterms_(_d(e), ctypes.byref(n), ctypes.byref(ier))
        nterms.append(n.value)
    itables = []
    for e in eps:
        it = np.zeros(343, np.int32)
        ier = ctypes.c_int(0)
        r.l3dterms_list2_(_d(e), _vp(it), ctypes.byref(ier))
        # Fortran itable(-3:3,-3:3,-3:3): first index fastest -> [k][j][i]
        itables.append(it.reshape(7, 7, 7).transpose(2, 1, 0).copy())
    np.savez_compressed(os.path.join(HERE, "lap_orders.npz"), eps=np.array(eps),
                        nterms=np.array(nterms, np.int32), itables=np.array(itables))


HELM_CASES = [
    # name, seed, iprec, zk, ns, nt, dist, ifcharge, ifdipole
    ("helm_targ_iprec1_cube", 501, 1, 6.0, 2000, 800, "cube", 1, 1),
    ("helm_self_iprec2_sphere", 502, 2, 20.0, 3000, 0, "sphere", 1, 0),
    ("helm_targ_iprec2_clusters_cplx", 503, 2, 3.0 + 0.5j, 2500, 600, "clusters", 1, 1),
    ("helm_targ_iprec0_cube_dip", 504, 0, 12.0, 1500, 700, "cube", 0, 1),
    ("helm_self_iprec3_cube", 505, 3, 2.0, 1800, 0, "cube", 1, 1),
    ("helm_targ_rootleaf", 506, 1, 1.0, 100, 40, "cube", 1, 1),
    # zk = 30 at iprec 3: orders of 30 and more at levels 2-3, where the
    # reference switches to rotations by projection (h3dtrans.f:318)
    ("helm_targ_iprec3_zk30", 507, 3, 30.0, 8000, 1000, "cube", 1, 1),
]


def make_helm(name, seed, iprec, zk, ns, nt, dist, ifc, ifd):
    rng = np.random.default_rng(seed)
    src = points(rng, ns, dist)
    trg = points(rng, nt, dist) if nt else None
    q = (rng.normal(size=ns) + 1j * rng.normal(size=ns)) if ifc else None
    ds = (rng.uniform(-1, 1, ns) + 1j * rng.uniform(-1, 1, ns)) if ifd else None
    dv = rng.normal(size=(ns, 3)) if ifd else None
    ier, pot, fld, pt, ft = ob.ref_hfmm3dparttarg(iprec, zk, src, q, ds, dv, trg)
    assert ier == 0, (name, ier)
    out = dict(iprec=np.int32(iprec), zk=np.complex128(zk), source=src, pot=pot, fld=fld)
    if trg is not None:
        out.update(target=trg, pottarg=pt, fldtarg=ft)
    if q is not None:
        out["charge"] = q
    if ds is not None:
        out.update(dipstr=ds, dipvec=dv)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), **out)


def make_helm_direct():
    rng = np.random.default_rng(601)
    src = rng.random((500, 3))
    trg = rng.random((150, 3))
    q = rng.normal(size=500) + 1j * rng.normal(size=500)
    ds = rng.normal(size=500) + 1j * rng.normal(size=500)
    dv = rng.normal(size=(500, 3))
    zk = 7.5 + 0.2j
    pot, fld, pt, ft = ob.ref_h3dpartdirect(zk, src, q, ds, dv, trg)
    np.savez_compressed(os.path.join(HERE, "helm_direct.npz"), zk=np.complex128(zk), source=src,
                        target=trg, charge=q, dipstr=ds, dipvec=dv, pot=pot, fld=fld, pottarg=pt,
                        fldtarg=ft)


def make_helm_operators():
    """Coefficient-level fixtures of the reference's Helmholtz expansion
    routines (packed n*n+n+m), for one geometry: child box of side 0.25 at c1,
    its parent at c2, a well-separated box at c0 with its child at c3."""
    r = ob.ref()
    rng = np.random.default_rng(701)
    nlege = 200
    wlege = np.zeros(100000)
    lused = ctypes.c_int(0)
    r.ylgndrfwini_(_i(nlege), _vp(wlege), _i(100000), ctypes.byref(lused))
    ier = ctypes.c_int(0)
    zk = 4.0 + 0.3j
    zka, zkp = ob._z(zk)
    p = 9
    side = 0.25
    c1 = np.array([0.3, -0.2, 0.5])
    c2 = c1 + side / 2 * np.array([1.0, -1.0, 1.0])      # parent centre (side 0.5)
    c0 = c1 + side * np.array([3.0, 2.0, -2.0])          # list-2 offset (3, 2, -2)
    c3 = c0 + side / 4 * np.array([-1.0, 1.0, 1.0])      # child of c0 (side 0.125)
    ns = 6
    src = c1 + side * (rng.random((ns, 3)) - 0.5)
    q = rng.normal(size=ns) + 1j * rng.normal(size=ns)
    ds = rng.normal(size=ns) + 1j * rng.normal(size=ns)
    dv = rng.normal(size=(ns, 3))
    trg = c0 + side * (rng.random((5, 3)) - 0.5)
    sc1 = min(abs(zk) * side, 1.0)
    sc2 = min(abs(zk) * 2 * side, 1.0)
    sc3 = min(abs(zk) * side / 2, 1.0)
    L = (p + 1) * (2 * p + 1)
    z = lambda: np.zeros(L, np.complex128)  # noqa: E731
    mp, mpd, lo, lod = z(), z(), z(), z()
    r.h3dformmp_add_trunc_(ctypes.byref(ier), zkp, _d(sc1), _vp(src), _vp(q), _i(ns), _vp(c1), _i(p), _i(p), _vp(mp), _vp(wlege), _i(nlege))
    r.h3dformmp_dp_add_trunc_(ctypes.byref(ier), zkp, _d(sc1), _vp(src), _vp(ds), _vp(dv), _i(ns), _vp(c1), _i(p), _i(p), _vp(mpd), _vp(wlege), _i(nlege))
    r.h3dformta_add_trunc_(ctypes.byref(ier), zkp, _d(sc1), _vp(src), _vp(q), _i(ns), _vp(c0), _i(p), _i(p), _vp(lo), _vp(wlege), _i(nlege))
    r.h3dformta_dp_add_trunc_(ctypes.byref(ier), zkp, _d(sc1), _vp(src), _vp(ds), _vp(dv), _i(ns), _vp(c0), _i(p), _i(p), _vp(lod), _vp(wlege), _i(nlege))
    nt = trg.shape[0]
    mpot, mfld = np.zeros(nt, np.complex128), np.zeros((nt, 3), np.complex128)
    r.h3dmpevalall_trunc_(zkp, _d(sc1), _vp(c1), _vp(mp), _i(p), _i(p), _vp(trg), _i(nt), _i(1), _vp(mpot), _i(1), _vp(mfld), _vp(wlege), _i(nlege), ctypes.byref(ier))
    lpot, lfld = np.zeros(nt, np.complex128), np.zeros((nt, 3), np.complex128)
    r.h3dtaevalall_trunc_(zkp, _d(sc1), _vp(c0), _vp(lo), _i(p), _i(p), _vp(trg), _i(nt), _i(1), _vp(lpot), _i(1), _vp(lfld), _vp(wlege), _i(nlege), ctypes.byref(ier))

    def quad(nq):
        x, w = np.zeros(nq), np.zeros(nq)
        r.legewhts_(_i(nq), _vp(x), _vp(w), _i(1))
        return x, w
    # M2M into the parent: radius = parent corner distance, nquad = max(6, 2.5 p)
    x, w = quad(max(6, int(p * 2.5)))
    m2m = z()
    r.h3dmpmpquadu_add_(zkp, _d(sc1), _vp(c1), _vp(mp), _i(p), _d(sc2), _vp(c2), _vp(m2m), _i(p), _i(p),
                        _d(side * np.sqrt(3.0)), _vp(x), _vp(w), _i(len(x)), ctypes.byref(ier))
    # M2L: radius = source box corner distance, nquad = max(6, 1.2 p)
    x, w = quad(max(6, int(p * 1.2)))
    m2l = z()
    r.h3dmplocquadu_add_trunc_(zkp, _d(sc1), _vp(c1), _vp(mp), _i(p), _i(p), _d(sc1), _vp(c0), _vp(m2l), _i(p), _i(p),
                               _d(side * np.sqrt(3.0) / 2), _vp(x), _vp(w), _i(len(x)), ctypes.byref(ier))
    m2l6 = z()
    r.h3dmplocquadu_add_trunc_(zkp, _d(sc1), _vp(c1), _vp(mp), _i(p), _i(6), _d(sc1), _vp(c0), _vp(m2l6), _i(p), _i(6),
                               _d(side * np.sqrt(3.0) / 2), _vp(x), _vp(w), _i(len(x)), ctypes.byref(ier))
    # L2L into the child: radius = child corner distance, nquad = max(6, 2 p)
    x, w = quad(max(6, 2 * p))
    l2l = z()
    r.h3dloclocquadu_add_(zkp, _d(sc1), _vp(c0), _vp(lo), _i(p), _d(sc3), _vp(c3), _vp(l2l), _i(p), _i(p),
                          _d(side / 2 * np.sqrt(3.0) / 2), _vp(x), _vp(w), _i(len(x)), ctypes.byref(ier))
    pk = lambda a: ob.ref_to_packed(a, p)  # noqa: E731
    np.savez_compressed(os.path.join(HERE, "helm_operators.npz"), p=np.int32(p), zk=np.complex128(zk),
                        sc1=sc1, sc2=sc2, sc3=sc3, c0=c0, c1=c1, c2=c2, c3=c3, source=src, charge=q,
                        dipstr=ds, dipvec=dv, target=trg, p2m=pk(mp), p2m_dp=pk(mpd), p2l=pk(lo),
                        p2l_dp=pk(lod), m2p_pot=mpot, m2p_fld=mfld, l2p_pot=lpot, l2p_fld=lfld,
                        m2m=pk(m2m), m2l=pk(m2l), m2l_trunc6=pk(m2l6), l2l=pk(l2l))


def make_helm_orders():
    r = ob.ref()
    eps = [.5e-1, .5e-3, .5e-6, .5e-9]
    cases = [(2.0, 1.0), (0.5, 20.0), (0.125, 20.0), (0.03125, 20.0), (1.0, 3.0 + 0.5j), (0.25, 30.0),
             (0.001, 1.0), (2.0, 40.0)]
    rows, tabs = [], []
    for size, zk in cases:
        zka, zkp = ob._z(zk)
        for e in eps:
            n = ctypes.c_int(0)
            ier = ctypes.c_int(0)
            r.h3dterms_(_d(size), zkp, _d(e), ctypes.byref(n), ctypes.byref(ier))
            it = np.zeros(343, np.int32)
            r.h3dterms_list2_(_d(size), zkp, _d(e), _vp(it), ctypes.byref(ier))
            rows.append((size, complex(zk).real, complex(zk).imag, e, n.value))
            tabs.append(it.reshape(7, 7, 7).transpose(2, 1, 0).copy())
    np.savez_compressed(os.path.join(HERE, "helm_orders.npz"), rows=np.array(rows), itables=np.array(tabs))


def main():
    only = set(sys.argv[1:])   # names of the FMM cases to (re)write; none = everything
    if only:
        for c in FMM_CASES:
            if c[0] in only:
                make_fmm(*c)
        for c in HELM_CASES:
            if c[0] in only:
                make_helm(*c)
        print("fixtures written to", HERE)
        return
    for c in FMM_CASES:
        make_fmm(*c)
    make_direct()
    make_tree()
    make_operators()
    make_orders()
    for c in HELM_CASES:
        make_helm(*c)
    make_helm_direct()
    make_helm_operators()
    make_helm_orders()
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
