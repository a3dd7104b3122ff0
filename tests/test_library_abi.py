"""The C-ABI library: it loads, exports what include/*.h declares, rejects
bad arguments with the reference's ier values before touching the GPU, and
its host octree is the reference's tree.  No device compute here."""
import glob
import os
import re

import numpy as np
import pytest

import fmmlib3d_amd as F
import oracle_bindings as ob
from fmm_inputs import points

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    names = set()
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        txt = re.sub(r"/\*.*?\*/", "", open(h).read(), flags=re.S)
        names |= set(re.findall(r"\b(fmm3d_\w+)\s*\(", txt))
    return names


def test_library_exports_every_declared_symbol():
    lib = F.load()
    names = declared()
    assert names, "no declarations found"
    for n in sorted(names):
        assert hasattr(lib, n), n
    assert set(F.EXPORTED) == names


def test_null_arrays_rejected_before_device_use():
    """An array the flags ask for that is NULL returns 128 (host and device
    entry points; the direct sums compute nothing) before any device call."""
    lib = F.load()
    src = np.zeros((4, 3))
    q = np.ones(4, np.complex128)
    out = np.zeros(4, np.complex128)
    fo = np.zeros((4, 3), np.complex128)
    p = F._ptr
    # source NULL, charge NULL with ifcharge=1, dipoles NULL with ifdipole=1
    assert lib.fmm3d_lfmm3dparttarg(1, 4, None, 1, p(q), 0, None, None, 1, p(out), 0, None,
                                    0, None, 0, None, 0, None) == 128
    assert lib.fmm3d_lfmm3dparttarg(1, 4, p(src), 1, None, 0, None, None, 1, p(out), 0, None,
                                    0, None, 0, None, 0, None) == 128
    assert lib.fmm3d_lfmm3dparttarg(1, 4, p(src), 0, None, 1, p(q), None, 1, p(out), 0, None,
                                    0, None, 0, None, 0, None) == 128
    # outputs NULL with their flag set, target NULL with ntarget > 0
    assert lib.fmm3d_lfmm3dparttarg(1, 4, p(src), 1, p(q), 0, None, None, 1, None, 0, None,
                                    0, None, 0, None, 0, None) == 128
    assert lib.fmm3d_lfmm3dparttarg(1, 4, p(src), 1, p(q), 0, None, None, 0, None, 1, None,
                                    0, None, 0, None, 0, None) == 128
    assert lib.fmm3d_lfmm3dparttarg(1, 4, p(src), 1, p(q), 0, None, None, 1, p(out), 1, p(fo),
                                    4, None, 0, None, 0, None) == 128
    assert lib.fmm3d_lfmm3dparttarg(1, 4, p(src), 1, p(q), 0, None, None, 1, p(out), 1, p(fo),
                                    4, p(src), 1, None, 0, None) == 128
    # the device entry points check the same before touching the device
    assert lib.fmm3d_lfmm3dparttarg_dev(0, 1, 4, p(src), 1, p(q), 0, None, None, 1, None, 0, None,
                                        0, None, 0, None, 0, None, None) == 128
    zk = np.array([1.0, 0.0])
    assert lib.fmm3d_hfmm3dparttarg_dev(0, 1, p(zk), 4, p(src), 1, None, 0, None, None, 1, p(out), 0, None,
                                        0, None, 0, None, 0, None, None) == 128
    assert lib.fmm3d_hfmm3dparttarg(1, p(zk), 4, p(src), 0, None, 1, p(q), None, 1, p(out), 0, None,
                                    0, None, 0, None, 0, None) == 128
    # quadrupoles: quadstr or quadvec NULL with ifquad = 1, and the iprec check
    qv = np.zeros((4, 6))
    assert lib.fmm3d_lfmm3dparttarg_qp(1, 4, p(src), 0, None, 0, None, None, 1, None, p(qv), 1, p(out), 0, None,
                                       0, None, 0, None, 0, None) == 128
    assert lib.fmm3d_lfmm3dparttarg_qp(1, 4, p(src), 0, None, 0, None, None, 1, p(q), None, 1, p(out), 0, None,
                                       0, None, 0, None, 0, None) == 128
    assert lib.fmm3d_lfmm3dparttarg_qp(7, 4, p(src), 0, None, 0, None, None, 1, p(q), p(qv), 1, p(out), 0, None,
                                       0, None, 0, None, 0, None) == 32
    with pytest.raises(ValueError):
        F.lfmm3dparttarg_qp(1, src, quadstr=q)
    # the direct sum has no error channel: it leaves the outputs untouched
    out[:] = 7
    lib.fmm3d_l3dpartdirect(4, p(src), 1, None, 0, None, None, 1, p(out), 0, None, 0, None, 0, None, 0, None)
    assert np.all(out == 7)


def test_version_string():
    assert "gfx950" in F.version()


def test_argument_errors_before_device_use():
    src = np.zeros((4, 3))
    q = np.ones(4, np.complex128)
    assert F.lfmm3dparttarg(1, np.zeros((0, 3)), np.zeros(0, np.complex128))[0] == 4
    assert F.lfmm3dparttarg(7, src, q)[0] == 32
    assert F.lfmm3dparttarg(-3, src, q)[0] == 32
    lib = F.load()
    # negative ntarget
    assert lib.fmm3d_lfmm3dparttarg(1, 4, F._ptr(src), 1, F._ptr(q), 0, None, None, 1, None, 0, None,
                                    -1, None, 0, None, 0, None) == 128
    # unknown precision mode
    assert F.lfmm3dparttarg(1, src, q, precision=9)[0] == 128
    # Helmholtz: the same checks, plus zk = 0 and modes other than FP64
    assert F.hfmm3dparttarg(1, 2.0, np.zeros((0, 3)), np.zeros(0, np.complex128))[0] == 4
    assert F.hfmm3dparttarg(7, 2.0, src, q)[0] == 32
    assert F.hfmm3dparttarg(1, 0.0, src, q)[0] == 128
    assert F.hfmm3dparttarg(1, 2.0, src, q, precision=F.PREC_FP32)[0] == 128
    assert F.hfmm3dparttarg(1, 2.0, src, q, precision=F.PREC_FAR32)[0] == 128
    assert lib.fmm3d_hfmm3dparttarg(1, None, 4, F._ptr(src), 1, F._ptr(q), 0, None, None, 1, None, 0,
                                    None, 0, None, 0, None, 0, None) == 128


@pytest.mark.parametrize("ns,nt,nbox,dist", [(3000, 2000, 40, "cube"), (5000, 0, 50, "sphere"),
                                             (4000, 3000, 80, "clusters"), (3000, 100, 30, "plummer"),
                                             (900, 300, 40, "plane"), (50, 10, 160, "cube"),
                                             (20000, 5000, 10, "plummer"), (30000, 0, 20, "clusters"),
                                             (40000, 40000, 400, "cube")])
def test_host_octree_equals_oracle(rng, ns, nt, nbox, dist):
    src = points(rng, ns, dist)
    trg = points(rng, nt, dist)
    ier, b1, i1, t1, l1, n1, c1, s1 = F.d3tstrcr(src, trg, nbox)
    b2, i2, t2, l2, n2 = ob.orc_tree(src, trg, nbox)
    assert ier == 0 and n1 == n2
    assert np.array_equal(b1[:, :17], b2[:, :17])
    assert np.array_equal(i1, i2) and np.array_equal(t1, t2)
    for k in range(4):
        assert l1[k] == l2[k], "list %d" % (k + 1)


def test_host_octree_equals_golden_tree():
    f = np.load(os.path.join(ROOT, "tests", "golden", "tree_clusters.npz"), allow_pickle=False)
    ier, boxes, isrc, itrg, lists, nlev, _, _ = F.d3tstrcr(f["source"], f["target"], int(f["nbox"]))
    assert ier == 0 and nlev == int(f["nlev"])
    assert np.array_equal(boxes[:, :17], f["boxes"][:, :17])
    assert np.array_equal(isrc, f["isrc"]) and np.array_equal(itrg, f["itrg"])
    counts, flat = f["list_counts"], f["lists"]
    o = 0
    for k in range(5):
        for b in range(boxes.shape[0]):
            c = counts[k, b]
            if k < 4:
                assert lists[k][b] == tuple(flat[o:o + c].tolist())
            o += c


def test_host_octree_rejects_empty():
    assert F.d3tstrcr(np.zeros((0, 3)), None, 10)[0] == 128


def test_noalloc_rejects_problem_beyond_its_array():
    """fmm3d_lfmm3dparttarg_dev_noalloc returns 256 for a problem larger
    than its static device array, before any device call (the pointers are
    never dereferenced)."""
    import ctypes
    lib = F.load()
    dummy = ctypes.c_void_p(256)
    n = 700_000
    ier = lib.fmm3d_lfmm3dparttarg_dev_noalloc(0, 2, n, dummy, 1, dummy, 0, None, None, 1, dummy, 1, dummy,
                                              n, dummy, 1, dummy, 1, dummy, None)
    assert ier == 256
    # argument errors come first, as in the other entry points
    assert lib.fmm3d_lfmm3dparttarg_dev_noalloc(0, 9, n, dummy, 1, dummy, 0, None, None, 1, dummy, 1, dummy,
                                               n, dummy, 1, dummy, 1, dummy, None) == 32


def test_triangle_entries_reject_bad_arguments_before_device_use():
    """The triangle FMMs check their arguments on the host: NULL triangles or
    densities the flags ask for, a density flag outside 0..2 (Helmholtz) or
    outside 0..1 (Laplace), iprec outside -2..6, zk = 0; the direct sums,
    with no error channel, leave their outputs untouched."""
    from fmm_inputs import icosphere
    lib = F.load()
    p = F._ptr
    T = np.ascontiguousarray(icosphere(0).reshape(-1, 9))
    n = T.shape[0]
    cen = np.ascontiguousarray(T.reshape(n, 3, 3).mean(axis=1))
    q = np.ones(n, np.complex128)
    pot = np.zeros(n, np.complex128)
    zk = np.array([2.0, 0.0])
    # Helmholtz
    assert lib.fmm3d_hfmm3dtriatarg(1, p(zk), n, None, None, p(cen), 1, p(q), 0, None, None, 1, p(pot), 0, None,
                                    0, None, 0, None, 0, None) == 128
    assert lib.fmm3d_hfmm3dtriatarg(1, p(zk), n, p(T), None, p(cen), 2, None, 0, None, None, 1, p(pot), 0, None,
                                    0, None, 0, None, 0, None) == 128
    assert lib.fmm3d_hfmm3dtriatarg(1, p(zk), n, p(T), None, p(cen), 0, None, 1, None, None, 1, p(pot), 0, None,
                                    0, None, 0, None, 0, None) == 128
    assert lib.fmm3d_hfmm3dtriatarg(1, p(zk), n, p(T), None, p(cen), 3, p(q), 0, None, None, 1, p(pot), 0, None,
                                    0, None, 0, None, 0, None) == 128
    assert lib.fmm3d_hfmm3dtriatarg(9, p(zk), n, p(T), None, p(cen), 1, p(q), 0, None, None, 1, p(pot), 0, None,
                                    0, None, 0, None, 0, None) == 32
    z0 = np.zeros(2)
    assert lib.fmm3d_hfmm3dtriatarg(1, p(z0), n, p(T), None, p(cen), 1, p(q), 0, None, None, 1, p(pot), 0, None,
                                    0, None, 0, None, 0, None) == 128
    assert lib.fmm3d_hfmm3dtriatarg(1, None, n, p(T), None, p(cen), 1, p(q), 0, None, None, 1, p(pot), 0, None,
                                    0, None, 0, None, 0, None) == 128
    assert F.hfmm3dtriatarg(1, 2.0, np.zeros((0, 3, 3)), np.zeros(0, np.complex128))[0] == 4
    # Laplace
    assert lib.fmm3d_lfmm3dtriatarg(1, n, None, None, p(cen), 1, p(q), 0, None, None, 1, p(pot), 0, None,
                                    0, None, 0, None, 0, None) == 128
    assert lib.fmm3d_lfmm3dtriatarg(-3, n, p(T), None, p(cen), 1, p(q), 0, None, None, 1, p(pot), 0, None,
                                    0, None, 0, None, 0, None) == 32
    # the direct sums compute nothing on bad input
    pot[:] = 7
    lib.fmm3d_h3dtriadirect(7, p(zk), n, p(T), None, p(cen), 1, p(q), 0, None, None, 1, p(pot), 0, None, 0, None,
                            0, None, 0, None)                                   # no rule of degree 7
    lib.fmm3d_h3dtriadirect(6, p(zk), n, p(T), None, p(cen), 1, None, 0, None, None, 1, p(pot), 0, None, 0, None,
                            0, None, 0, None)                                   # charge NULL
    lib.fmm3d_l3dtriadirect(n, None, None, p(cen), 1, p(q), 0, None, None, 1, p(pot), 0, None, 0, None, 0, None,
                            0, None)                                            # triangles NULL
    assert np.all(pot == 7)
