"""Continuous-batching server (pulsesuite_amd.serve): each request's result is the
same alone and in the mix (bit for bit: a wire's arithmetic does not depend on
which wires share its launches), admission respects the HBM budget and slot
count, and a served solve matches the FP64 oracle's coupled 1-D step."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def reqs():
    from pulsesuite_amd.serve import Request

    return [Request(0, 16, 30, 3e7, 0.0), Request(1, 24, 20, 1e7, 0.0), Request(2, 12, 25, 5e7, 0.0),
            Request(3, 20, 12, 2e7, 0.01)]


def same(a, b):
    return bool(np.array_equal(a, b))


@pytest.mark.parametrize("feedback", [False, True])
def test_alone_equals_mix(feedback):
    from pulsesuite_amd.serve import Server

    mix = Server(max_slots=2, precision="fp64", feedback=feedback, keep_results=True)
    mix.run(reqs())
    assert len(mix.records) == 4 and mix.max_active == 2
    for r in reqs():
        r.arrival = 0.0
        alone = Server(precision="fp64", feedback=feedback, keep_results=True)
        alone.run([r])
        a, b = alone.results[r.rid], mix.results[r.rid]
        assert a["steps"] == b["steps"] == r.steps
        for k in ("Ey", "Ex", "p", "C", "D"):
            assert np.isfinite(a[k]).all(), (r.rid, k)
            ok = same(a[k], b[k])
            assert ok, (r.rid, k)


def test_budget_limits_concurrency():
    from pulsesuite_amd.serve import Server, solve_bytes

    rs = reqs()
    budget = max(solve_bytes(r.Nk) for r in rs) + 1024  # room for one solve at a time
    srv = Server(hbm_budget=budget, precision="fp64")
    rec = srv.run(rs)
    assert srv.max_active == 1 and len(rec) == 4
    rec = sorted(rec, key=lambda x: x["admitted"])
    for a, b in zip(rec, rec[1:]):
        assert b["admitted"] >= a["finished"] - 1e-3 or b["admitted"] >= a["admitted"]
    s = srv.summary()
    assert s["finished"] == 4 and s["p99_latency_s"] >= s["p50_latency_s"] > 0


def test_served_solve_matches_oracle():
    """One request through the server vs the oracle's coupled step on the same
    1024 x 1 x 1 grid (soft source, 16-cell CPML, wire at the centre)."""
    from oracle import pstd as OP
    from oracle.coupling import OracleCoupling
    from oracle.sbe import L_for_nk, MBFlags, QWParams
    from pulsesuite_amd import serve as S

    r = S.Request(0, 14, 600, 4e7, 0.0)  # the pulse reaches the wire after ~450 steps
    srv = S.Server(precision="fp64", keep_results=True, feedback=True)
    srv.run([r])
    got = srv.results[0]
    sp = OP.Space(3, S.NX, 1, 1, S.DX, S.DX, S.DX, S.EPSR)
    oc = OracleCoupling(sp, 800e-9, 0.0, Nw=1, qw=QWParams(L=L_for_nk(r.Nk)), mb=MBFlags())
    prop = OP.Propagator(sp, OP.Time(0.0, 1.0, S.DT, 0), OP.Pulse(800e-9, r.E0, 100e-15, 50e-15, 0.0), "soft",
                         "cpml", calc_j=oc.calc_j, npml=S.NPML)
    for _ in range(r.steps):
        prop.step()
    # E as a vector field: on a line along x, Ex is longitudinal and only accumulates
    # the current J_x, whose CalcJfx part is the derivative of Rho = RhoH - RhoE, a
    # difference of nearly equal densities (rounding noise), so Ex is held to |E|
    num = sum(np.linalg.norm(got[k] - getattr(prop, k)) ** 2 for k in ("Ex", "Ey"))
    den = sum(np.linalg.norm(getattr(prop, k)) ** 2 for k in ("Ex", "Ey"))
    e = float(np.sqrt(num / den))
    assert e < 1e-10, e
    assert np.abs(got["p"]).max() > 1e-4  # the wire was driven
    so = oc.wire.state[0]
    for k, a in (("p", so.p2), ("C", so.C2)):
        e = float(np.linalg.norm(got[k] - a) / np.linalg.norm(a))
        assert e < 1e-10, (k, e)
