"""Batched wires of uneven Nk in one launch (pulsesuite_amd.batch) against the
FP64 oracle run wire by wire, and invariance of the results under the launch
plan (bucket count, streams) and under HIP-graph capture and replay."""

import numpy as np
import pytest

from oracle import sbe as O

pytestmark = pytest.mark.gpu
HBAR = 1.0545718176461565e-34
E = 1.602176634e-19


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / max(np.linalg.norm(b), 1e-300))


def specs():
    from pulsesuite_amd.batch import WireSpec

    nks = [12, 20, 16, 28, 14, 24]
    e0 = [3e7, 1e8, 5e6, 2e7, 6e7, 1e7]
    det = [0.0, 0.03, -0.04, 0.01, -0.02, 0.05]
    return [WireSpec(k, e, d, pol=(1.0, 0.5, 0.2), tw=4e-15, tp=5e-15) for k, e, d in zip(nks, e0, det)]


def drive(sp, w0, t):
    """docs/source/examples/sbe_wire_1d.py:45-54"""
    u = w0 * (t - sp.tp)
    w0tw = w0 * sp.tw
    return np.exp(-(u**2) / w0tw**2) * np.cos(u) * np.exp(-(u**20) / (2 * w0tw) ** 20)


def test_batch_matches_oracle_wire_by_wire():
    import torch

    from pulsesuite_amd.batch import WireBatch

    sp = specs()
    dt = 4e-17
    b = WireBatch(sp, dt=dt, nrr=16)
    nsteps = 4
    b.run(nsteps, graph=False)
    torch.cuda.synchronize()
    for i, s in enumerate(sp):
        L = O.L_for_nk(s.Nk)
        rr = np.linspace(-1.2 * L, 1.2 * L, 16)
        ref = O.OracleWire(rr, 0.0, 0.0, 800e-9, 1, O.QWParams(L=L), O.MBFlags())
        w0 = (ref.gap + s.detuning_eV * E) / HBAR
        for n in range(nsteps):
            f = drive(s, w0, n * dt)
            Ex, Ey, Ez = (np.full(16, s.E0 * c * f, dtype=complex) for c in s.pol)
            r = ref.QWCalculator(Ex, Ey, Ez, rr, dt, 1)
        st, so = b.get_state(i), ref.state[0]
        assert b.steps_done(i) == nsteps
        for k in ("p2", "C2", "D2", "p1"):
            assert rel(st[k], getattr(so, k)) < 1e-10, (i, k, rel(st[k], getattr(so, k)))
        P = b.polarization(i).cpu().numpy()
        for c in range(3):
            assert rel(P[c], r[c]) < 1e-10, (i, c)
        assert np.abs(so.p2).max() > 0


def test_plans_and_graph_replays_identical():
    """Eager single-launch steps vs a bucketed, two-stream plan captured in a graph:
    every replay, with eager work and synchronisations between replays, must give
    the eager states bit for bit."""
    import torch

    from pulsesuite_amd.batch import Plan, WireBatch

    sp = specs()
    b = WireBatch(sp, dt=4e-17, nrr=16)
    snap = b.snapshot()
    ref = []
    for _ in range(4):
        b.step_eager(Plan(1, 1))
        torch.cuda.synchronize()
        ref.append([b.get_state(i) for i in range(len(sp))])
    for plan in (Plan(3, 1), Plan(3, 2)):
        b.restore(snap)
        g = b.capture(plan)  # runs step 1 eagerly
        torch.cuda.synchronize()
        for n in range(4):
            if n > 0:
                g.replay()
                x = torch.randn(256, 256, device="cuda", dtype=torch.float64)
                (x @ x).sum().item()  # eager work between replays
            torch.cuda.synchronize()
            for i in range(len(sp)):
                st = b.get_state(i)
                for k in ("p2", "C2", "D2"):
                    ok = bool(np.array_equal(st[k], ref[n][i][k]))
                    assert ok, (plan, n, i, k)


@pytest.mark.parametrize("graph", [False, True])
def test_autotune_restores_state_and_picks_a_plan(graph):
    import torch

    from pulsesuite_amd.batch import WireBatch

    b = WireBatch(specs(), dt=4e-17, nrr=16)
    b.run(2, graph=False)
    torch.cuda.synchronize()
    before = [b.get_state(i) for i in range(b.N)]
    ctl = b.dev.ctl.clone()
    best, times = b.autotune(reps=2, graph=graph)
    assert best in times and len(times) >= 3
    assert torch.equal(ctl, b.dev.ctl)
    for i in range(b.N):
        st = b.get_state(i)
        for k in st:
            ok = bool(np.array_equal(st[k], before[i][k]))
            assert ok, (i, k)
    b.run(3)  # tuned plan, captured
    torch.cuda.synchronize()
    assert all(b.steps_done(i) == 5 for i in range(b.N))
