"""Continuous-batching solve server for wire + 1-D PSTD requests on one GPU.

A request is one quantum wire (its own Nk) coupled to its own 1-D PSTD line
(1024 cells, 16-cell CPML, soft source; the coupled step of PSTD3D/PSTD3D.py:523-627
with the wire glue of PSTD3D/rhoPJ.py:1060-1329 on an Nx x 1 x 1 grid), run for a
requested number of steps.  The server keeps every admitted solve resident on the
device and advances all of them together, one tick per step:

  1. per solve: E half step (update, CPML, source, point sample) and wire drive line;
  2. one batched SBE step over every in-flight wire (wires of different Nk share the
     launches: sorted by Nk, cut into work-balanced buckets, as in batch.WireBatch);
  3. per solve: current line, its FFT, B half step.
Steps 1 and 3 of a solve are captured as two HIP graphs at its first tick (the
source time lives in a device counter), so a tick costs two graph launches per
solve plus the batched SBE launches.

Requests that arrive while others run are admitted at the next tick when a slot is
free and their device memory fits in the HBM budget (first fit in arrival order);
finished solves leave the batch at once, and their memory returns to the budget when
the GPU has finished their last step (an event per solve).  A wire's
arithmetic does not depend on which other wires share its launches, so a result
is the same alone and in the mix.

feedback=False (the default) runs the wires with the reference's ignorewire switch
(PSTD3D/SBEs.py:5993-5997): every kernel runs and the wire's P and rho are zeroed
before they form the current, which keeps long solves finite; with feedback=True
the point-current loop is closed.

Latency = completion (the GPU finished the solve's last step, seen through an event)
minus arrival; throughput = sum of Nk x steps over finished requests / wall time.
"""

import ctypes as C
import dataclasses
import math
import time
from collections import deque

import numpy as np

from . import _lib
from .batch import balanced_ranges, wire_bytes
from .SBEs import feature_bits, precision_mask
from .wire_model import L_of_nk, MBFlags, QWParams

# 1-D line: 1024 cells of 10 nm with 16-cell CPML.  On an Nx x 1 x 1 grid the
# reference takes eps_r = 1 (GetEpsr returns 1 when Nz == 1), so the leapfrog
# spectral step limit is 2 dx / (pi c) = 21.2 as; 0.9 of it is used.
NX, NPML, DX, EPSR = 1024, 16, 10e-9, 9.1
DT = 0.9 * 2.0 * DX / (math.pi * 2.99792458e8)


@dataclasses.dataclass
class Request:
    """One solve: wire of Nk points, steps, source amplitude E0 (V/m), arrival time (s after
    the server starts) and the wire's sample point x0 on the line (m; 0 = line centre)."""
    rid: int
    Nk: int
    steps: int
    E0: float = 1e7
    arrival: float = 0.0
    x0: float = 0.0


def source_plane_x():
    """x of the soft source's plane on the served line (cell round(0.25 NX), PSTD3D/tfsf.py:26-60)."""
    x = np.linspace(-DX * (NX - 1) / 2.0, DX * (NX - 1) / 2.0, NX)
    return float(x[int(round(0.25 * NX))])


def poisson_requests(n, rate, seed=0, nk_lo=64, nk_hi=2048, st_lo=500, st_hi=20000, e_lo=1e6, e_hi=1e8):
    """Seeded Poisson stream: exponential gaps at `rate` per second, Nk log-uniform over
    [nk_lo, nk_hi] rounded to even, steps log-uniform over [st_lo, st_hi], source
    amplitude log-uniform over [e_lo, e_hi]."""
    rng = np.random.default_rng(seed)
    gaps = rng.exponential(1.0 / rate, n)
    arr = np.cumsum(gaps) - gaps[0]
    nk = 2 * np.maximum(np.round(np.exp(rng.uniform(math.log(nk_lo), math.log(nk_hi), n)) / 2), 2).astype(int)
    st = np.round(np.exp(rng.uniform(math.log(st_lo), math.log(st_hi), n))).astype(int)
    e0 = np.exp(rng.uniform(math.log(e_lo), math.log(e_hi), n))
    return [Request(i, int(k), int(s), float(e), float(a)) for i, (k, s, e, a) in enumerate(zip(nk, st, e0, arr))]


def solve_bytes(nk, nx=NX):
    """Device bytes of one solve: the wire (state + workspace on an nx-point line) and
    the 1-D PSTD fields (E, B, J, CPML psi and scratch, coupling lines)."""
    return wire_bytes(nk, nrr=nx) + 48 * 16 * nx


class Solve:
    """One admitted request, resident on the device."""

    def __init__(self, req, precision, mb, device, feedback=False):
        from .PSTD3D import PSTD3DPropagator
        from .rhoPJ import QWArray
        from .typepulse import ps
        from .typespace import ss
        from .typetime import ts

        self.req = req
        sp = ss(3, NX, 1, 1, DX, DX, DX, EPSR)
        mbf = dataclasses.replace(mb, ignorewire=not feedback)
        self.qa = QWArray(QWParams(L=L_of_nk(req.Nk)), mbf, Nw=1, Emax0=0.0, precision=precision, x0=req.x0)
        self.prop = PSTD3DPropagator(sp, ts(0.0, 1.0, DT, 0), ps(800e-9, req.E0, 100e-15, 50e-15, 0.0),
                                     source_type="soft", boundary_type="cpml", calc_j=self.qa, npml=NPML,
                                     check_every=0, device=device, precision=precision, device_time=True)
        if self.qa.model.Nk != req.Nk:
            raise ValueError(f"request {req.rid}: Nk={req.Nk} not reachable (gives {self.qa.model.Nk})")
        self.done = 0
        self.nbytes = solve_bytes(req.Nk)
        self.admitted = None
        self.event = None
        self.g_e = self.g_b = None

    def desc(self):
        return self.qa.wires.descs[0]

    def _launch_e(self):
        self.prop.step_e_half()
        self.qa.drive(self.prop.space, self.prop.time)

    def _launch_b(self):
        self.qa.deposit(self.prop.space, self.prop.time)
        self.prop.step_b_half()

    def capture(self):
        """Record the two half steps (E side with the wire drive line, B side with the
        current line) as HIP graphs; the source time is a device counter, so replays
        advance it.  Host bookkeeping done by the launches is rolled back."""
        import torch

        tm, prop = self.prop.time, self.prop
        saved = (tm.t, tm.n, prop._nstep)
        s = torch.cuda.Stream(device=prop.device)
        s.wait_stream(torch.cuda.current_stream(prop.device))
        self.g_e, self.g_b = torch.cuda.CUDAGraph(), torch.cuda.CUDAGraph()
        with _lib.gc_paused(), torch.cuda.stream(s):
            with torch.cuda.graph(self.g_e, stream=s):
                self._launch_e()
            with torch.cuda.graph(self.g_b, stream=s):
                self._launch_b()
        torch.cuda.current_stream(prop.device).wait_stream(s)
        tm.t, tm.n, prop._nstep = saved

    def half_e(self, graphs=False):
        if graphs:
            if self.g_e is None:
                self.capture()
            self.g_e.replay()
            self.prop.time.t += self.prop.time.dt / 2.0
        else:
            self._launch_e()

    def half_b(self, graphs=False):
        if graphs:
            self.g_b.replay()
            tm = self.prop.time
            tm.t += tm.dt / 2.0
            tm.n += 1
            self.prop._nstep += 1
        else:
            self._launch_b()
        self.done += 1

    def result(self):
        """Host copies of what a caller keeps: the field line and the wire state."""
        st = self.qa.wires.get_state(0)
        return {"Ey": self.prop.Ey.cpu().numpy().copy(), "Ex": self.prop.Ex.cpu().numpy().copy(),
                "p": st["p2"], "C": st["C2"], "D": st["D2"], "steps": self.done}


class Server:
    """Admission by HBM budget and slot count, one batched tick per step."""

    def __init__(self, hbm_budget=64 << 30, max_slots=64, precision="mixed", mb_flags=None, buckets=4,
                 device=None, feedback=False, keep_results=False, graphs=True, streams=8):
        import torch

        self.budget = int(hbm_budget)
        self.max_slots = max_slots
        self.precision = precision
        self._pmask = precision_mask(_lib.MIXED if precision == "mixed" else _lib.FP64)
        self.mb = mb_flags or MBFlags()
        self.buckets = buckets
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.feedback = feedback
        self.keep = keep_results
        self.graphs = graphs
        # the per-solve half steps are small (1-D lines): spread them over a few streams
        # (default priority) so they run side by side; joined before and after the SBE step
        self.nstreams = max(1, streams)
        self._pool = [torch.cuda.Stream(device=self.device) for _ in range(self.nstreams)] if streams > 1 else []
        self._ev_fork = torch.cuda.Event()
        self._ev_join = [torch.cuda.Event() for _ in range(self.nstreams)]
        self.used = 0
        self.active = []
        self._dirty = True
        self._d_descs = None
        self._ranges = []
        self.results = {}
        self.records = []
        self.ticks = 0
        self.max_active = 0

    # ---------------------------------------------------------------- batching
    def _rebuild(self):
        """Descriptor array of the in-flight wires, sorted by Nk (largest first)."""
        import torch

        self.active.sort(key=lambda s: -s.req.Nk)
        n = len(self.active)
        arr = (_lib.WireDesc * n)()
        for i, s in enumerate(self.active):
            arr[i] = s.desc()
        raw = bytes(memoryview(arr).cast("B"))
        host = torch.frombuffer(bytearray(raw), dtype=torch.uint8)
        self._d_descs = host.to(self.device)
        nk = np.array([s.req.Nk for s in self.active])
        # few wires: one launch sequence over all of them fills the chip best; more wires:
        # up to `buckets` sequences side by side (one per 8 wires)
        nb = max(1, min(self.buckets, n // 8))
        self._ranges = balanced_ranges(nk, nb)
        self._dirty = False

    def _sbe_step(self):
        """One batched SBE step of every in-flight wire: one launch sequence per bucket,
        buckets spread over the stream pool."""
        import torch

        sz = C.sizeof(_lib.WireDesc)
        base = self._d_descs.data_ptr()
        main = torch.cuda.current_stream(self.device)
        nb = len(self._ranges)
        ns = min(len(self._pool), nb) if self._pool else 0
        if ns > 1:
            self._ev_fork.record(main)
            for j in range(ns):
                self._pool[j].wait_event(self._ev_fork)
        for b, (lo, hi) in enumerate(self._ranges):
            sub = self.active[lo:hi]
            m = [s.qa.model for s in sub]
            st = _lib.stream_ptr(self._pool[b % ns] if ns > 1 else main)
            _lib.check(_lib.lib.pss_sbe_step_p(C.c_void_p(base + lo * sz), hi - lo, min(x.Nk for x in m),
                                               max(x.Nk for x in m), max(x.Nr for x in m), NX,
                                               self._pmask | feature_bits(m), 0, st),
                       "pss_sbe_step_p")
        if ns > 1:
            for j in range(ns):
                self._ev_join[j].record(self._pool[j])
                main.wait_event(self._ev_join[j])

    def tick(self):
        if self._dirty:
            self._rebuild()
        self.max_active = max(self.max_active, len(self.active))
        self._each(lambda s: s.half_e(self.graphs))
        self._sbe_step()
        self._each(lambda s: s.half_b(self.graphs))
        self.ticks += 1

    def _each(self, fn):
        """fn(solve) for every in-flight solve, forked over the stream pool and joined
        back into the current stream."""
        import torch

        if not self._pool:
            for s in self.active:
                fn(s)
            return
        main = torch.cuda.current_stream(self.device)
        self._ev_fork.record(main)
        ns = min(self.nstreams, len(self.active))
        for j in range(ns):
            self._pool[j].wait_event(self._ev_fork)
        for i, s in enumerate(self.active):
            with torch.cuda.stream(self._pool[i % ns]):
                fn(s)
        for j in range(ns):
            self._ev_join[j].record(self._pool[j])
            main.wait_event(self._ev_join[j])

    # ---------------------------------------------------------------- loop
    def admit(self, req):
        need = solve_bytes(req.Nk)
        if len(self.active) >= self.max_slots or self.used + need > self.budget:
            return None
        s = Solve(req, self.precision, self.mb, self.device, self.feedback)
        self.used += s.nbytes
        self.active.append(s)
        self._dirty = True
        return s

    def run(self, requests, time_limit=None):
        """Serve `requests` (arrival times relative to the start); returns the records."""
        import torch

        pending = deque(sorted(requests, key=lambda r: r.arrival))
        for r in pending:
            if solve_bytes(r.Nk) > self.budget:
                raise ValueError(f"request {r.rid} (Nk={r.Nk}) can never fit the HBM budget")
        waiting = []
        draining = []
        t0 = time.perf_counter()
        arrived = {}
        while pending or waiting or self.active or draining:
            now = time.perf_counter() - t0
            if time_limit is not None and now > time_limit:
                break
            while pending and pending[0].arrival <= now:
                r = pending.popleft()
                arrived[r.rid] = r.arrival
                waiting.append(r)
            still = []
            for r in waiting:  # first fit, arrival order
                s = self.admit(r)
                if s is None:
                    still.append(r)
                else:
                    s.admitted = now
            waiting = still
            if self.active:
                self.tick()
                left = []
                for s in self.active:
                    if s.done >= s.req.steps:
                        ev = torch.cuda.Event()
                        ev.record()
                        s.event = ev
                        draining.append(s)
                        self._dirty = True
                    else:
                        left.append(s)
                self.active = left
            elif draining:
                torch.cuda.synchronize(self.device)
            elif pending:
                time.sleep(max(0.0, min(pending[0].arrival - now, 0.05)))
            keep = []
            for s in draining:
                if s.event.query():
                    # its device memory goes back to the budget only now: until its last
                    # step has run, the solve's buffers are still in use
                    self.used -= s.nbytes
                    done_t = time.perf_counter() - t0
                    r = s.req
                    self.records.append(dict(rid=r.rid, Nk=r.Nk, steps=r.steps, arrival=r.arrival,
                                             admitted=s.admitted, finished=done_t, latency=done_t - r.arrival))
                    if self.keep:
                        self.results[r.rid] = s.result()
                else:
                    keep.append(s)
            draining = keep
        torch.cuda.synchronize(self.device)
        self.wall = time.perf_counter() - t0
        return self.records

    def summary(self):
        rec = self.records
        if not rec:
            return {"finished": 0}
        lat = np.array([r["latency"] for r in rec])
        kps = sum(r["Nk"] * r["steps"] for r in rec)
        return {"finished": len(rec), "wall_s": self.wall, "ticks": self.ticks,
                "kpoint_steps_per_s": kps / self.wall, "p50_latency_s": float(np.percentile(lat, 50)),
                "p99_latency_s": float(np.percentile(lat, 99)), "mean_latency_s": float(lat.mean())}
