"""The request server: requests of different kernels, sizes, distributions
and tolerances run together on the GPU (several workers, one stream each),
and every result equals the result of the same call made alone."""
import numpy as np
import pytest

import fmmlib3d_amd as F
from fmm_inputs import points, strengths

pytestmark = pytest.mark.gpu


def _mix(rng):
    reqs = []
    for k, (ns, nt, iprec, dist, zk, ifd) in enumerate([
            (20000, 5000, 2, "cube", None, 1), (8000, 0, 1, "shell", None, 0), (30000, 2000, 3, "clusters", None, 0),
            (12000, 3000, 2, "sphere", 8.0, 1), (5000, 5000, 1, "plummer", None, 1), (15000, 0, 2, "cube", 2.0 + 0.5j, 0),
            (3000, 1000, 0, "cube", None, 1), (40000, 0, 1, "clusters", None, 0)]):
        src = points(rng, ns, dist)
        trg = points(rng, nt, dist)
        q, ds, dv = strengths(rng, ns, True, bool(ifd))
        reqs.append((iprec, zk, src, q, ds, dv, trg))
    return reqs


def _alone(iprec, zk, src, q, ds, dv, trg):
    if zk is None:
        return F.lfmm3dparttarg(iprec, src, q, ds, dv, trg)
    return F.hfmm3dparttarg(iprec, zk, src, q, ds, dv, trg)


def _same(a, b):
    for x, y in zip(a, b):
        if y is None:
            assert x is None
        else:
            assert np.array_equal(x, y)


@pytest.mark.parametrize("nworkers,budget", [(3, 16 << 30), (4, 1 << 20)])
def test_server_results_equal_standalone(rng, nworkers, budget):
    """(3 workers, 16 GiB): requests overlap on the device.  (4 workers, 1 MiB):
    every estimate exceeds the budget, so admission runs them one at a time."""
    reqs = _mix(rng)
    alone = [_alone(*r) for r in reqs]
    with F.Server(nworkers=nworkers, mem_budget=budget) as srv:
        ids = [srv.submit(1 if zk is not None else 0, iprec, src, q, ds, dv, trg, zk=zk)
               for (iprec, zk, src, q, ds, dv, trg) in reqs]
        outs = [srv.wait(i) for i in reversed(ids)][::-1]
        assert srv.stats()["running"] == 0
    for o, a in zip(outs, alone):
        assert o[0] == 0 and a[0] == 0
        assert o[5] > 0.0
        _same(o[1:5], a[1:5])


def test_server_rejects_bad_requests(rng):
    src = points(rng, 100, "cube")
    q, _, _ = strengths(rng, 100, True, False)
    with F.Server(nworkers=1, mem_budget=1 << 30) as srv:
        with pytest.raises(F.FMMError):
            srv.submit(0, 9, src, q)
        with pytest.raises(F.FMMError):
            srv.submit(1, 1, src, q)            # Helmholtz without zk
        rid = srv.submit(0, 1, src, q)
        assert srv.wait(rid)[0] == 0


@pytest.mark.parametrize("iprec,ifd,with_targets", [(1, 0, False), (2, 1, True)])
def test_server_fuses_small_requests(rng, iprec, ifd, with_targets):
    """Small Laplace requests queued together with the same parameters run as
    one forest in one graph-captured launch sequence; each result is bitwise
    the result of the same call made alone.  Two rounds on the same server,
    so the workers reuse their forest storage."""
    with F.Server(nworkers=2, mem_budget=16 << 30) as srv:
        for rnd in range(2):
            reqs = []
            for k in range(24):
                n = int(rng.integers(1000, 20000))
                src = points(rng, n, "clusters" if k % 2 else "cube")
                trg = points(rng, int(rng.integers(100, 3000)), "cube") if with_targets else None
                q, ds, dv = strengths(rng, n, True, bool(ifd))
                reqs.append((iprec, None, src, q, ds, dv, trg))
            alone = [_alone(*r) for r in reqs]
            ids = [srv.submit(0, iprec, src, q, ds, dv, trg) for (iprec, zk, src, q, ds, dv, trg) in reqs]
            outs = [srv.wait(i) for i in ids]
            st = srv.stats()
            print("round", rnd, "fused requests so far:", st["fused_requests"])
            assert st["fused_requests"] >= 4 * (rnd + 1)
            for o, a in zip(outs, alone):
                assert o[0] == 0 and a[0] == 0
                _same(o[1:5], a[1:5])
