"""The two serving streams at their full size, through the request server
(4 workers, requests overlapping on the device): every result equals the
same call made alone bitwise, and is within the tolerance its iprec names of
the FP64 direct sum (oracle) at sampled points (src/lfmm3dpart.f:277-285):
potential at that tolerance, field at ten times it.

- the Laplace stream: 512 lfmm3dpartself requests, N log-uniform in
  [1e4, 2e6], cube/shell/clusters, iprec 1-3, charges, FP64;
- the mixed stream: 256 Plummer requests, Laplace or Helmholtz (zk in
  [1, 30]), iprec 2-3, charges + dipoles, mixed precision (the library runs
  FP64 at iprec 3)."""
import numpy as np
import pytest

import fmmlib3d_amd as F
import oracle_bindings as ob
from fmm_inputs import IPREC_EPS, rel, stream4_requests, stream5_requests

pytestmark = [pytest.mark.gpu, pytest.mark.slow]


def _check_stream(reqs, submit, alone, direct, nsample):
    rng = np.random.default_rng(99)
    worst = {}
    with F.Server(nworkers=4, mem_budget=96 << 30) as srv:
        ids = [submit(srv, r) for r in reqs]
        for k, (r, i) in enumerate(zip(reqs, ids)):
            o = srv.wait(i)
            assert o[0] == 0, (k, o[0])
            a = alone(r)
            assert a[0] == 0
            assert np.array_equal(o[1], a[1]) and np.array_equal(o[2], a[2]), k
            iprec, x = r[-2], r[-1]
            idx = rng.choice(x.shape[0], min(nsample, x.shape[0]), replace=False)
            pd, fd = direct(r, idx)
            ep, ef = rel(o[1][idx], pd), rel(o[2][idx], fd)
            worst[iprec] = max(worst.get(iprec, (0, 0))[0], ep), max(worst.get(iprec, (0, 0))[1], ef)
            assert ep < IPREC_EPS[iprec], (k, iprec, ep)
            assert ef < 10 * IPREC_EPS[iprec], (k, iprec, ef)
            if k % 64 == 0:
                print("request %d (N=%d, iprec %d): pot %.1e fld %.1e" % (k, x.shape[0], iprec, ep, ef), flush=True)
    print("worst per iprec (pot, fld):", worst)


def test_laplace_stream_512_requests():
    reqs = [(q, it, x) for (it, x, q) in stream4_requests(4, 512)]

    def submit(srv, r):
        q, it, x = r
        return srv.submit(0, it, x, q)

    def alone(r):
        q, it, x = r
        return F.lfmm3dpartself(it, x, q)

    def direct(r, idx):
        q, it, x = r
        return ob.orc_direct_sample(x, x[idx], idx.astype(np.int32), q)

    _check_stream(reqs, submit, alone, direct, 50)


def test_mixed_stream_256_requests():
    reqs = [(zk, q, ds, dv, iprec, x) for (zk, iprec, x, q, ds, dv) in stream5_requests(5, 256)]

    def submit(srv, r):
        zk, q, ds, dv, iprec, x = r
        return srv.submit(0 if zk is None else 1, iprec, x, charge=q, dipstr=ds, dipvec=dv, zk=zk,
                          precision=F.PREC_MIXED)

    def alone(r):
        zk, q, ds, dv, iprec, x = r
        if zk is None:
            return F.lfmm3dpartself(iprec, x, q, ds, dv, precision=F.PREC_MIXED)
        return F.hfmm3dpartself(iprec, zk, x, q, ds, dv, precision=F.PREC_MIXED)

    def direct(r, idx):
        zk, q, ds, dv, iprec, x = r
        return ob.orc_direct_sample(x, x[idx], idx.astype(np.int32), q, ds, dv, zk=zk)

    _check_stream(reqs, submit, alone, direct, 100)
