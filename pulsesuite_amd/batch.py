"""Batched quantum wires of uneven k-grid size, stepped together on one GPU.

Every wire is an independent SBE solve (its own Nk, pump amplitude, detuning;
the per-wire step of PSTD3D/SBEs.py:5838-5997 driven by the uniform Gaussian
pulse of docs/source/examples/sbe_wire_1d.py:45-54,99-100).  One batch step is
one drive launch for all wires plus the SBE step kernels launched once per
bucket of wires: wires are sorted by Nk and cut into contiguous buckets of
roughly equal work (sum of Nk^3), each launch sized for its bucket's largest
wire, buckets spread over a few streams.  The step is captured once in a HIP
graph (torch.cuda.CUDAGraph) and replayed; the drive time and the state
rotation live in device counters, so replays need no host work.

The plan (bucket count, stream count) is chosen by :meth:`WireBatch.autotune`,
which times each candidate's step on the batch itself and restores the state
afterwards.  A wire's arithmetic does not depend on the plan (every kernel
indexes a wire's work by its own Nk), so results are identical under every plan.
"""

import dataclasses
import math
import time

import numpy as np

from . import _lib
from .SBEs import DeviceWires
from .wire_model import L_of_nk, MBFlags, QWParams, build_model

_HBAR = 1.0545718176461565e-34
_E = 1.602176634e-19
_C0 = 2.99792458e8


@dataclasses.dataclass
class WireSpec:
    """One wire of a batch: k-grid size, pump and its timing.

    E0: peak field (V/m); pol: weights of (Ex, Ey, Ez); detuning_eV: photon energy
    minus the band gap; tw, tp: pulse width and centre (s); t0: start time (s)."""
    Nk: int
    E0: float = 1e7
    detuning_eV: float = 0.0
    pol: tuple = (1.0, 0.0, 0.0)
    tw: float = 10e-15
    tp: float = 50e-15
    t0: float = 0.0


def sample_specs(n, seed=0, nk_lo=64, nk_hi=1024, e_lo=1e6, e_hi=1e9, det_eV=0.05):
    """Seeded batch: Nk uniform over the even values in [nk_lo, nk_hi], E0
    log-uniform over [e_lo, e_hi], detuning uniform over +-det_eV."""
    rng = np.random.default_rng(seed)
    nks = 2 * rng.integers(nk_lo // 2, nk_hi // 2 + 1, size=n)
    e0 = np.exp(rng.uniform(np.log(e_lo), np.log(e_hi), size=n))
    det = rng.uniform(-det_eV, det_eV, size=n)
    return [WireSpec(int(k), float(e), float(d)) for k, e, d in zip(nks, e0, det)]


@dataclasses.dataclass(frozen=True)
class Plan:
    buckets: int = 1
    streams: int = 1


class WireBatch:
    """Device-resident batch of independent wires.

    specs: list of WireSpec.  qw_params: base material parameters (L is set per
    wire from its Nk).  The wires are stored sorted by Nk (largest first);
    ``slot[i]`` is the storage position of spec i."""

    def __init__(self, specs, precision="fp64", dt=10e-18, mb_flags=None, qw_params=None, nrr=16, device=None):
        import torch

        if not specs:
            raise ValueError("empty batch")
        self.specs = list(specs)
        self.precision = precision
        self.dt = float(dt)
        mb = mb_flags or MBFlags()
        base = qw_params or QWParams()
        order = sorted(range(len(specs)), key=lambda i: -specs[i].Nk)
        self.order = order
        self.slot = [0] * len(specs)
        for s, i in enumerate(order):
            self.slot[i] = s
        models = {}
        for nk in sorted({s.Nk for s in specs}):
            L = L_of_nk(nk)
            rr = np.linspace(-1.2 * L, 1.2 * L, nrr)
            m = build_model(dataclasses.replace(base, L=L), mb, rr, 0.0)
            if m.Nk != nk:
                raise ValueError(f"Nk={nk} is not reachable (gives {m.Nk}); use even Nk")
            models[nk] = m
        self.models = models
        sorted_models = [models[specs[i].Nk] for i in order]
        n = len(specs)
        self.N = n
        self.nk = np.array([specs[i].Nk for i in order])
        # independent wires: own wireoff latch each, gate disabled (Emax0 = 0)
        self.dev = DeviceWires(sorted_models, n, 0.0, self.dt, precision, device, w_index=[1] * n,
                               groups=list(range(n)))
        self.device = self.dev.device
        pulses = np.zeros((n, 8))
        for s, i in enumerate(order):
            sp = specs[i]
            gap = sorted_models[s].scalars["gap"]
            w0 = (gap + sp.detuning_eV * _E) / _HBAR
            pulses[s, :3] = sp.E0 * np.asarray(sp.pol, dtype=float)
            pulses[s, 3:7] = (w0, sp.tw, sp.tp, sp.t0)
        self.pulses = torch.as_tensor(pulses, dtype=torch.float64, device=self.device)
        self.plan = Plan()
        self._graph = None
        self._graph_plan = None
        self._streams = []
        self._events = []
        self.tuning = None

    # ------------------------------------------------------------ planning
    def ranges(self, nb):
        """Contiguous slot ranges of roughly equal work (sum Nk^3), at most nb."""
        return balanced_ranges(self.nk, nb)

    def _ensure_streams(self, ns):
        import torch

        while len(self._streams) < ns:
            self._streams.append(torch.cuda.Stream(device=self.device))  # default priority
            self._events.append((torch.cuda.Event(), torch.cuda.Event()))

    def _launch(self, plan):
        """Enqueue one batch step on the current stream (forking to side streams)."""
        import torch

        cur = torch.cuda.current_stream(self.device)
        _lib.check(_lib.lib.pss_drive_uniform(_lib.ptr(self.dev.d_descs), _lib.ptr(self.pulses), self.N,
                                              _lib.stream_ptr(cur)), "pss_drive_uniform")
        rng = self.ranges(plan.buckets)
        ns = min(plan.streams, len(rng))
        if ns <= 1:
            for r in rng:
                self.dev.step(wires=r, stream=cur)
            return
        self._ensure_streams(ns)
        fork = self._events[0][0]
        fork.record(cur)
        for j in range(ns):
            self._streams[j].wait_event(fork)
        for b, r in enumerate(rng):
            self.dev.step(wires=r, stream=self._streams[b % ns])
        for j in range(ns):
            ev = self._events[j][1]
            ev.record(self._streams[j])
            cur.wait_event(ev)

    def step_eager(self, plan=None):
        self._launch(plan or self.plan)

    def capture(self, plan=None):
        """Capture one batch step under ``plan`` in a CUDA/HIP graph."""
        import torch

        plan = plan or self.plan
        self._launch(plan)  # warm-up outside the graph (library state, LDS limits)
        torch.cuda.synchronize(self.device)
        g = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream(device=self.device)
        s.wait_stream(torch.cuda.current_stream(self.device))
        with _lib.gc_paused(), torch.cuda.stream(s):
            with torch.cuda.graph(g, stream=s):
                self._launch(plan)
        torch.cuda.current_stream(self.device).wait_stream(s)
        self._graph, self._graph_plan = g, plan
        return g

    def run(self, nsteps, graph=True):
        """Advance every wire by nsteps (graph replays, or eager launches)."""
        if not graph:
            for _ in range(nsteps):
                self._launch(self.plan)
            return
        if nsteps <= 0:
            return
        if self._graph is None or self._graph_plan != self.plan:
            self.capture(self.plan)  # executes one eager step; the captured launches do not run
            nsteps -= 1
        for _ in range(nsteps):
            self._graph.replay()

    # ------------------------------------------------------------ autotuner
    def snapshot(self):
        d = self.dev
        return ([{k: v.clone() for k, v in st.items()} for st in d.state], d.ctl.clone(), d.on.clone())

    def restore(self, snap):
        d = self.dev
        states, ctl, on = snap
        for st, sv in zip(d.state, states):
            for k in st:
                st[k].copy_(sv[k])
        d.ctl.copy_(ctl)
        d.on.copy_(on)

    def candidates(self):
        out = []
        for nb in (1, 2, 4, 8, 16):
            if nb > 1 and len(self.ranges(nb)) < nb:
                continue
            for ns in (1, 2, 4):
                if ns <= nb:
                    out.append(Plan(nb, ns))
        return out

    def autotune(self, reps=3, candidates=None, graph=False):
        """Time each plan's step on this batch (eager launches, or captured graphs with
        graph=True), keep the fastest, restore the state."""
        import torch

        snap = self.snapshot()
        times = {}
        for plan in candidates or self.candidates():
            if graph:
                g = self.capture(plan)
                run = g.replay
            else:
                self._launch(plan)  # warm-up
                run = lambda p=plan: self._launch(p)  # noqa: E731
            torch.cuda.synchronize(self.device)
            t0 = time.perf_counter()
            for _ in range(reps):
                run()
            torch.cuda.synchronize(self.device)
            times[plan] = (time.perf_counter() - t0) / reps
            self.restore(snap)
        best = min(times, key=times.get)
        self.plan = best
        self._graph = None
        self.tuning = {f"b{p.buckets}s{p.streams}": t for p, t in times.items()}
        torch.cuda.synchronize(self.device)
        return best, times

    # ------------------------------------------------------------ results
    def kpoint_steps(self, nsteps):
        return int(self.nk.sum()) * nsteps

    def get_state(self, i):
        """State of spec i (numpy), as DeviceWires.get_state."""
        return self.dev.get_state(self.slot[i])

    def polarization(self, i):
        """(Pxx, Pyy, Pzz) of spec i on its propagation grid (torch, complex128)."""
        return self.dev.P_out[self.slot[i]]

    def steps_done(self, i):
        return int(self.dev.ctl[self.slot[i], 0].item())

    def memory_bytes(self):
        return batch_bytes([s.Nk for s in self.specs], nrr=self.dev.nrr[0])


def balanced_ranges(nks, nb):
    """Cut a list of wires (sorted by Nk) into at most nb contiguous ranges of roughly
    equal work, sum of Nk^3."""
    n = len(nks)
    if n == 0:
        return []
    cum = np.cumsum(np.asarray(nks, dtype=np.float64) ** 3)
    cuts = [0]
    for b in range(1, nb):
        c = int(np.searchsorted(cum, cum[-1] * b / nb, side="left")) + 1
        if cuts[-1] < c < n:
            cuts.append(c)
    cuts.append(n)
    return [(a, b) for a, b in zip(cuts[:-1], cuts[1:]) if b > a]


def wire_bytes(nk, nrr=16, precision="mixed"):
    """Device bytes one wire holds: 9 Nk x Nk complex state matrices plus its workspace."""
    nr = 2 * nk + 2
    code = _lib.FP64 if precision == "fp64" else _lib.MIXED
    return 9 * 16 * nk * nk + int(_lib.lib.pss_sbe_workspace_bytes(nk, nr, nrr, code))


def batch_bytes(nks, nrr=16):
    return sum(wire_bytes(k, nrr) for k in nks)
