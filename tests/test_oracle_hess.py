"""The oracle's direct-sum Hessians (oracle/fmm_oracle.c lap_pair_hess,
restating lpotfld3dhess / lpotfld3dhess_dp, src/laprouts3d.f:5442, :5619)
pinned three ways: against the reference's own compiled routines (live,
where oracle/_ref is built), against central differences of the oracle's
field (fld = -grad pot, so d fld_a / d x_b = -pot_ab), and by the Laplace
trace pot_xx + pot_yy + pot_zz = 0.  CPU only."""
import numpy as np
import pytest

import oracle_bindings as ob
from fmm_inputs import points, rel, strengths

IA = [0, 1, 2, 0, 0, 1]
IB = [0, 1, 2, 1, 2, 2]


@pytest.mark.skipif(not ob.ref_available(), reason="reference library not built here")
def test_oracle_hess_matches_reference_routines():
    rng = np.random.default_rng(31)
    for _ in range(20):
        s, t = rng.random(3), rng.random(3) + 0.3
        q = complex(rng.normal(), rng.normal())
        _, _, href = ob.ref_lpotfld3dhess(s, q, t)
        h = ob.orc_direct_hess_sample(s[None], t[None], np.array([-1], np.int32), charge=np.array([q]))[0]
        assert rel(h, href) < 1e-14
        ds = complex(rng.normal(), rng.normal())
        dv = rng.normal(size=3)
        _, _, href = ob.ref_lpotfld3dhess_dp(s, ds, dv, t)
        h = ob.orc_direct_hess_sample(s[None], t[None], np.array([-1], np.int32), dipstr=np.array([ds]),
                                      dipvec=dv[None])[0]
        assert rel(h, href) < 1e-14


@pytest.mark.parametrize("ifc,ifd", [(1, 0), (0, 1), (1, 1)])
def test_oracle_hess_is_minus_field_derivative_and_traceless(ifc, ifd):
    rng = np.random.default_rng(32)
    src = points(rng, 300, "cube")
    trg = points(rng, 40, "cube") + 1.5   # away from the sources
    q, ds, dv = strengths(rng, 300, bool(ifc), bool(ifd), unit_dipvec=False)
    none = -np.ones(40, np.int32)
    h = ob.orc_direct_hess_sample(src, trg, none, q, ds, dv)
    eps = 1e-5
    fd = np.zeros_like(h)
    for b in range(3):
        e = np.zeros(3)
        e[b] = eps
        _, fp = ob.orc_direct_sample(src, trg + e, none, q, ds, dv)
        _, fm = ob.orc_direct_sample(src, trg - e, none, q, ds, dv)
        dfl = -(fp - fm) / (2 * eps)     # d pot_a / d x_b
        for k in range(6):
            if IB[k] == b:
                fd[:, k] = dfl[:, IA[k]]
    assert rel(h, fd) < 1e-8
    assert np.max(np.abs(h[:, 0] + h[:, 1] + h[:, 2])) < 1e-10 * np.max(np.abs(h))
