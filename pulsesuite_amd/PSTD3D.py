"""3-D pseudo-spectral Maxwell propagator on the GPU (drop-in for PSTD3D/PSTD3D.py).

Fields are contiguous complex128 CUDA tensors of shape (Nx, Ny, Nz), kept in
k-space between half steps as in the reference.  ``UpdateE3D`` / ``UpdateB3D``
(PSTD3D.py:29-113) update in place; ``PSTD3DPropagator`` (PSTD3D.py:233-858)
runs the leapfrog loop: E curl update, CPML-E (with the soft source and the
wire point-sample fused into its real-space pass), t += dt/2, ``calc_j``,
B curl update, CPML-B (+ source), t += dt/2.

Boundaries: "mask" (the reference's default, PSTD3D/absorber.py), "cpml" and "none"
(periodic).  ``polarization`` (keyword extra, default "y" as the reference) picks the soft
source's and the IC seed's field: "y" injects Ey and Bz = Ey/v, "z" injects Ez and
By = -Ez/v (a plane wave along +x; E along x would be longitudinal there).

Differences from the reference, on purpose: no per-step diagnostics IFFTs, Emax.dat,
snapshots or summary files; the NaN/overflow guard checks the k-space fields every
``check_every`` steps and raises RuntimeError as the reference does (PSTD3D.py:510-512).
"""

import numpy as np
import torch
from scipy.constants import c as _c0
from scipy.constants import mu_0 as _mu0

from . import _lib
from ._grid import GridContext
from .typespace import GetDx, GetEpsr, GetXArray, GetYArray, GetZArray

_twopi = 2.0 * np.pi
_ctx = {}


def _context(space, dt, device):
    key = (space.Nx, space.Ny, space.Nz, space.dx, space.dy, space.dz, GetEpsr(space), dt, str(device))
    if key not in _ctx:
        _ctx[key] = GridContext(space, dt, with_cpml=False, device=device)
    return _ctx[key]


def UpdateE3D(space, time, Bx, By, Bz, Jx, Jy, Jz, Ex, Ey, Ez):
    """E += i v2 dt (k x B) - mu0 v2 dt J, in place (PSTD3D/PSTD3D.py:29-76)."""
    ctx = _context(space, time.dt, Ex.device)
    P = _lib.ptr
    _lib.check(_lib.lib.pss_update_e3d(ctx.gp, P(Bx), P(By), P(Bz), P(Jx), P(Jy), P(Jz), None, None, None,
                                       P(Ex), P(Ey), P(Ez), _lib.stream_ptr()), "pss_update_e3d")


def UpdateB3D(space, time, Ex, Ey, Ez, Bx, By, Bz):
    """B -= i dt (k x E), in place (PSTD3D/PSTD3D.py:79-113)."""
    ctx = _context(space, time.dt, Ex.device)
    P = _lib.ptr
    _lib.check(_lib.lib.pss_update_b3d(ctx.gp, P(Ex), P(Ey), P(Ez), P(Bx), P(By), P(Bz), _lib.stream_ptr()),
               "pss_update_b3d")


def SeedInitialCondition(space, time, pulse, Ey, Bz, npml_x=0, b_sign=1.0):
    """Seed Ey, Bz with the pulse and FFT to k-space (PSTD3D/PSTD3D.py:134-226).

    The seed is evaluated on the host once at set-up; the transform runs on the GPU.
    b_sign = -1 seeds (Ez, By) of a z-polarised pulse through the same arguments."""
    from .fftw import fft_3D

    Nx = space.Nx
    x, yy, zz = GetXArray(space), GetYArray(space), GetZArray(space)
    omega0 = pulse.CalcOmega0()
    tau_G = pulse.CalcTau()
    v = _c0 / np.sqrt(GetEpsr(space))
    sigma_x = v * tau_G
    k_med = omega0 / v
    xi = (x - x[Nx // 2])[:, None, None]
    gauss = np.exp(-(yy[None, :, None] ** 2 + zz[None, None, :] ** 2) / pulse.w0**2)
    E = pulse.Amp * gauss * np.exp(-(xi**2) / sigma_x**2) * np.exp(1j * (k_med * xi + (pulse.chirp / v**2) * xi**2))
    if npml_x > 0:
        taper = np.ones(Nx)
        for i in range(npml_x):
            taper[i] = 0.5 * (1.0 - np.cos(np.pi * i / npml_x))
        for i in range(Nx - npml_x, Nx):
            taper[i] = 0.5 * (1.0 - np.cos(np.pi * (Nx - 1 - i) / npml_x))
        E = E * taper[:, None, None]
    Ey.copy_(torch.as_tensor(E, dtype=torch.complex128))
    Bz.copy_(torch.as_tensor(b_sign * E / v, dtype=torch.complex128))
    fft_3D(Ey)
    fft_3D(Bz)


def InitializeTFSF(space, pulse):
    """Normalised Gaussian source profile, sigma = 5 dx at 25 % of x (PSTD3D/tfsf.py:26-60)."""
    x = GetXArray(space)
    dx = GetDx(space)
    xp = x[int(round(space.Nx * 0.25))]
    s = 5.0 * dx
    return np.exp(-0.5 * ((x - xp) / s) ** 2) / (s * np.sqrt(_twopi))


class PSTD3DPropagator:
    """PSTD3D/PSTD3D.py:233-629 on the GPU.  Keyword-only extras: ``npml`` (CPML
    thickness override), ``check_every`` (NaN guard period), ``device``."""

    def __init__(self, space, time, pulse, *, source_type="ic", boundary_type="mask", calc_j=None,
                 output_dir="output", snapshot_interval=500, npml=None, check_every=50, device=None,
                 precision="fp64", device_time=False, overlap=None, polarization="y"):
        if boundary_type not in ("mask", "cpml", "none"):
            raise ValueError(f"unknown boundary_type {boundary_type!r}")
        if polarization not in ("y", "z"):
            raise ValueError("polarization must be 'y' or 'z' (the source plane wave travels along +x)")
        self.polarization = polarization
        # soft-source / seed components: E and B indices and the sign of B = (x^ x E)/v
        self._e_comp, self._b_comp, self._b_sign = (1, 2, 1.0) if polarization == "y" else (2, 1, -1.0)
        if source_type not in ("ic", "soft"):
            raise ValueError(f"unknown source_type {source_type!r}")
        self.space, self.time, self.pulse = space, time, pulse
        self.source_type, self.boundary_type = source_type, boundary_type
        self.calc_j = calc_j
        self.check_every = check_every
        self.device = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        from .cpml import CPML

        self.cpml = CPML(space.Nx, space.Ny, space.Nz, space.dx, space.dy, space.dz, time.dt, GetEpsr(space),
                         npml=npml if boundary_type == "cpml" else 0, device=self.device)
        if boundary_type != "cpml":
            self.cpml.ctx.g.npml_x = self.cpml.ctx.g.npml_y = self.cpml.ctx.g.npml_z = 0
        self.ctx = self.cpml.ctx
        self.absorber = None
        if boundary_type == "mask":
            from .absorber import Absorber

            self.absorber = Absorber(space.Nx, space.Ny, space.Nz, time.dt, ctx=self.ctx)
            self._partial = torch.zeros(6144, dtype=torch.float64, device=self.device)
        if precision not in ("fp64", "mixed"):
            raise ValueError(f"unknown precision {precision!r}")
        # mixed: the CPML's spectral derivatives (real parts, PML cells only) in FP32;
        # the fields, their transforms and the curl updates stay FP64
        self.precision = precision
        self._curl_early = precision == "mixed"  # pipelined runs: E curl term with the correction
        self.ctx.g.precision = _lib.MIXED if precision == "mixed" else _lib.FP64
        # device_time: the source reads the time from a device scalar advanced by a kernel
        # after each half step, so a half step can be captured once and replayed
        self._tdev = None
        if device_time:
            self._tdev = torch.full((1,), float(time.t), dtype=torch.float64, device=self.device)
            self.ctx.g.tdev = self._tdev.data_ptr()
        self.v = _c0 / np.sqrt(GetEpsr(space))
        self.v2 = _c0**2 / GetEpsr(space)
        f = self.ctx.field
        self.Ex, self.Ey, self.Ez, self.Bx, self.By, self.Bz = (f() for _ in range(6))
        self.Jx, self.Jy, self.Jz = (f() for _ in range(3))
        dev = self.device
        self._rx = torch.as_tensor(GetXArray(space), device=dev)
        self._ry = torch.as_tensor(GetYArray(space), device=dev)
        self._rz = torch.as_tensor(GetZArray(space), device=dev)
        self._tfsf = None
        if source_type == "soft":
            self._tfsf = torch.as_tensor(InitializeTFSF(space, pulse), device=dev)
        if source_type == "ic":
            npx = self.absorber.npml_x if self.absorber is not None else self.cpml.npml_x
            Es, Bs = (self.Ey, self.Bz) if polarization == "y" else (self.Ez, self.By)
            SeedInitialCondition(space, time, pulse, Es, Bs, npx, b_sign=self._b_sign)
        if calc_j is not None and hasattr(calc_j, "bind"):
            calc_j.bind(self)
        self._nstep = 0
        # overlap: the calc_j hook (reads E, writes J) runs on a side stream beside the
        # B half step (reads E, writes B); both join before the next E update, which
        # reads J and B.  Same launches and operands as the serial order, so identical
        # results.  The hook never reads the device clock, so this holds with device time.
        if overlap is None:
            overlap = calc_j is not None and getattr(calc_j, "separable", False)
        self.overlap = bool(overlap) and calc_j is not None
        self._side = None
        self._graph = None
        self._pgraphs = None
        # pipelined steps recorded per graph: each replay boundary (the join, then the next
        # graph's start) leaves the GPU idle for ~20 us, paid once per this many steps
        # (128^3 pipelined step, 1/2/4/8 per graph: mixed 0.585/0.581/0.579/0.579 ms,
        # FP64 0.809/0.806/0.805/0.805 ms)
        self.steps_per_graph = 4

    def _src(self, comp, amp, t):
        if self._tfsf is None:
            return None
        p = self.pulse
        w0inv2 = 0.0 if not np.isfinite(p.w0) else 1.0 / p.w0**2
        t_arg = 0.0 if self._tdev is not None else t  # device time: offset 0 from *tdev
        par = (_lib.C.c_double * 9)(t_arg, amp, p.CalcOmega0(), p.CalcTau(), p.chirp, self.v, p.Tp, w0inv2,
                                    GetDx(self.space))
        return (comp, par, self._tfsf, self._rx, self._ry, self._rz)

    def step(self, stream=None):
        """One leapfrog step (PSTD3D/PSTD3D.py:523-627)."""
        self.step_e_half(stream)
        if self.calc_j is None:
            self.step_b_half(stream)
        elif not self.overlap:
            self.calc_j(self.space, self.time, self.Ex, self.Ey, self.Ez, self.Jx, self.Jy, self.Jz)
            self.step_b_half(stream)
        else:
            cur = stream if stream is not None else torch.cuda.current_stream(self.device)
            self._ensure_side()
            self._ev_e.record(cur)
            self._side.wait_event(self._ev_e)
            with torch.cuda.stream(self._side):
                self.calc_j(self.space, self.time, self.Ex, self.Ey, self.Ez, self.Jx, self.Jy, self.Jz)
            self._ev_j.record(self._side)
            self.step_b_half(stream)
            cur.wait_event(self._ev_j)

    def _ensure_side(self):
        if self._side is None:
            self._side = torch.cuda.Stream(device=self.device)  # default priority
            self._ev_e, self._ev_j = torch.cuda.Event(), torch.cuda.Event()

    def capture_step(self):
        """Record one step (both streams when overlapped) as a HIP graph.  Needs
        device_time=True: the soft source reads the time from the device counter that
        the recorded step advances, so every replay is the next step.  The recording
        executes nothing; the host bookkeeping of the recorded launches is rolled back."""
        if self._tdev is None:
            raise RuntimeError("capture_step needs a propagator built with device_time=True")
        if self.overlap:
            self._ensure_side()
        tm = self.time
        saved = (tm.t, tm.n, self._nstep, self.check_every)
        self.check_every = 0
        s = torch.cuda.Stream(device=self.device)
        s.wait_stream(torch.cuda.current_stream(self.device))
        g = torch.cuda.CUDAGraph()
        try:
            with _lib.gc_paused(), torch.cuda.stream(s):
                with torch.cuda.graph(g, stream=s):
                    self.step()
        finally:
            torch.cuda.current_stream(self.device).wait_stream(s)
            tm.t, tm.n, self._nstep, self.check_every = saved
        self._graph = g
        return g

    def run_graph(self, nsteps):
        """Advance nsteps by replaying the recorded step (recorded on first use)."""
        if self._graph is None:
            self.capture_step()
        tm = self.time
        for _ in range(nsteps):
            self._graph.replay()
            tm.t += tm.dt / 2.0  # the two half steps, as step() advances the host clock
            tm.t += tm.dt / 2.0
            tm.n += 1
            self._nstep += 1
            if self.check_every and self._nstep % self.check_every == 0:
                self.check_finite()

    def step_e_half(self, stream=None):
        """E update + CPML-E half step (with the soft source on Ey and the wire's point
        sample), then t += dt/2: the part of a step before the calc_j hook."""
        tm = self.time
        self._update_e(stream)
        sample = self.calc_j.sample_args() if getattr(self.calc_j, "separable", False) else None
        self._cpml_e(stream, sample)
        self._advance(tm.dt / 2.0, _lib.stream_ptr(stream))
        tm.t += tm.dt / 2.0

    def _update_e(self, stream):
        """E += i v2 dt (k x B) - mu0 v2 dt J (J from the last calc_j call)."""
        tm = self.time
        g = self.ctx.g
        g.dt = tm.dt
        g.v2 = self.v2
        P = _lib.ptr
        sep = getattr(self.calc_j, "separable", False)
        jl, gy, gz = self.calc_j.jsep() if sep else (None, None, None)
        Jd = (None, None, None) if sep else (self.Jx, self.Jy, self.Jz)
        _lib.check(_lib.lib.pss_update_e3d(self.ctx.gp, P(self.Bx), P(self.By), P(self.Bz), *(P(t) for t in Jd),
                                           P(jl), P(gy), P(gz), P(self.Ex), P(self.Ey), P(self.Ez),
                                           _lib.stream_ptr(stream)), "pss_update_e3d")

    def _cpml_e(self, stream, sample=None):
        """CPML correction of E from B and psi_E (or the masking absorber), with the soft
        source on its E component at the current (host or device) time; optionally the
        wire's point sample of the corrected E."""
        tm = self.time
        src = self._src(self._e_comp, self.pulse.Amp, tm.t)
        if self.absorber is not None:
            self.absorber.apply(self.Ex, self.Ey, self.Ez, src=src, stream=stream)
            if sample is not None:
                sidx, sw, sout = sample
                P = _lib.ptr
                _lib.check(_lib.lib.pss_ksample(self.ctx.gp, P(self.Ex), P(self.Ey), P(self.Ez), P(sidx), P(sw),
                                                P(self._partial), P(sout), _lib.stream_ptr(stream)), "pss_ksample")
            return
        self.cpml.ApplyCPML_E(self.Ex, self.Ey, self.Ez, self.Bx, self.By, self.Bz, dt=tm.dt, v2=self.v2,
                              src=src, sample=sample, stream=stream)

    def _cpml_e_next(self, stream):
        """The next step's CPML correction of E in a pipelined run, soft source on Ey at the
        current (host or device) time; with the curl term of its E update (both read only
        B) in mixed precision.  In FP64 the field branch is the longer one and the curl term
        stays in the E update: on the correction's FP64 last pass it lengthened that branch
        by more than it took off the E update (128^3: 0.813 -> 0.851 ms per step), while in
        mixed precision the step got shorter (0.590 -> 0.586 ms)."""
        tm = self.time
        if self._curl_early:
            self.cpml.CurlE_ApplyCPML_E(self.Ex, self.Ey, self.Ez, self.Bx, self.By, self.Bz, tm.dt, self.v2,
                                        src=self._src(self._e_comp, self.pulse.Amp, tm.t), stream=stream)
        else:
            self._cpml_e(stream)

    # ---- pipelined steps (graph replays of several steps)
    # The CPML correction of an E half step reads only B, psi_E and the clock, not J, and so
    # does the curl term of the E update; both add to E like the current term does, so the
    # three commute.  Applied right after the B half step of the step before (in mixed
    # precision with the curl term in the correction's last forward pass), they run while
    # the wire step (calc_j) still works on the side stream; what waits for the wire's
    # current is the rest of the E update and the point sample.  Same operations on the same operands in another order: results
    # agree with eager steps to rounding.  The state between replays holds the next step's
    # correction and curl term, so the sequence is closed: a prologue applies the first ones
    # and the last step applies none.
    def _pipelined_step(self, pre_next, stream=None):
        tm = self.time
        P = _lib.ptr
        st = _lib.stream_ptr(stream)
        g = self.ctx.g
        g.dt = tm.dt
        g.v2 = self.v2
        jl, gy, gz = self.calc_j.jsep()
        sidx, sw, sout = self.calc_j.sample_args()
        # the E update (only its current term when the curl term came with the correction,
        # before) and the point sample of the new E in one pass over E; the sample's last
        # kernel also advances the device clock by dt/2
        Bp = (None, None, None) if self._curl_early else (P(self.Bx), P(self.By), P(self.Bz))
        # the wires' input lines filled with the sample by the same kernels, when the hook
        # keeps them in one block
        la = self.calc_j.line_args() if hasattr(self.calc_j, "line_args") else None
        lines, line_n, nlines = (P(la[0]), la[1], la[2]) if la is not None and la[0] is not None else (None, 0, 0)
        _lib.check(_lib.lib.pss_update_e3d_sample(self.ctx.gp, *Bp, None, None, None,
                                                  P(jl), P(gy), P(gz), P(self.Ex), P(self.Ey), P(self.Ez), P(sidx),
                                                  P(sw), P(self.cpml.scratch), P(sout), lines, line_n, nlines,
                                                  P(self._tdev), tm.dt / 2.0, 0 if self._curl_early else 1, st),
                   "pss_update_e3d_sample")
        tm.t += tm.dt / 2.0
        cur = stream if stream is not None else torch.cuda.current_stream(self.device)
        self._ensure_side()
        self._ev_e.record(cur)
        # recorded first, the field branch is the one the graph starts at once (the other
        # branch's queue starts later); it is the longer one here
        # FP64: the B curl term came with the E update (same pass over the cells)
        self.step_b_half(stream, curl=self._curl_early)
        if pre_next:
            self._cpml_e_next(stream)
        self._side.wait_event(self._ev_e)
        with torch.cuda.stream(self._side):
            if lines is not None:
                self.calc_j.step_filled(self.space, self.time)
            else:
                self.calc_j(self.space, self.time, self.Ex, self.Ey, self.Ez, self.Jx, self.Jy, self.Jz)
        self._ev_j.record(self._side)
        cur.wait_event(self._ev_j)

    def pipelined(self):
        """True when run_pipelined applies: the CPML boundary, device time and an overlapped,
        separable calc_j."""
        return (self.boundary_type == "cpml" and self._tdev is not None and self.overlap
                and getattr(self.calc_j, "separable", False))

    def capture_pipelined(self, steps_per_graph=None):
        """Record the pipelined step (with and without the next step's CPML-E correction)
        as HIP graphs: steps_per_graph (default self.steps_per_graph) steps with the
        correction in one graph, one such step, and the closing step; like capture_step,
        the recording executes nothing."""
        if not self.pipelined():
            raise RuntimeError("pipelined steps need device_time=True and an overlapped separable calc_j")
        self._ensure_side()
        tm = self.time
        saved = (tm.t, tm.n, self._nstep, self.check_every)
        self.check_every = 0
        s = torch.cuda.Stream(device=self.device)
        s.wait_stream(torch.cuda.current_stream(self.device))
        K = int(self.steps_per_graph if steps_per_graph is None else steps_per_graph)
        if K < 1:
            raise ValueError("steps_per_graph >= 1")
        graphs = []
        try:
            for pres in ((True,) * K, (True,), (False,)) if K > 1 else ((True,), (False,)):
                g = torch.cuda.CUDAGraph()
                with _lib.gc_paused(), torch.cuda.stream(s):
                    with torch.cuda.graph(g, stream=s):
                        for pre in pres:
                            self._pipelined_step(pre)
                graphs.append(g)
                tm.t, tm.n, self._nstep = saved[:3]
        finally:
            torch.cuda.current_stream(self.device).wait_stream(s)
            tm.t, tm.n, self._nstep, self.check_every = saved
        self._pgraphs = (graphs[0], K, graphs[-2], graphs[-1])
        return self._pgraphs

    def run_pipelined(self, nsteps):
        """Advance nsteps: the first step's CPML-E correction, then nsteps - 1 replays of
        the pipelined step and one of the closing step.  Afterwards the state is that of
        nsteps eager steps (to rounding)."""
        if nsteps <= 0:
            return
        if self._pgraphs is None:
            self.capture_pipelined()
        gk, K, g1, gl = self._pgraphs
        self._cpml_e_next(None)
        rest = nsteps - 1
        while rest > 0:
            k = K if rest >= K else 1
            (gk if k == K else g1).replay()
            self._book_steps(k)
            rest -= k
        gl.replay()
        self._book_steps(1)

    def _book_steps(self, k):
        """Host bookkeeping of k replayed steps (time, counters, the periodic check)."""
        tm = self.time
        before = self._nstep
        for _ in range(k):
            tm.t += tm.dt / 2.0
            tm.t += tm.dt / 2.0
            tm.n += 1
        self._nstep += k
        if self.check_every and self._nstep // self.check_every > before // self.check_every:
            self.check_finite()

    def step_b_half(self, stream=None, curl=True):
        """B update + CPML-B half step (soft source on Bz), then t += dt/2: the part of a
        step after the calc_j hook.  curl=False: the CPML-B correction only (the curl term
        of the B update was applied already)."""
        tm = self.time
        st = _lib.stream_ptr(stream)
        src = self._src(self._b_comp, self._b_sign * self.pulse.Amp / self.v, tm.t)
        if self.absorber is not None:
            if curl:
                self.ctx.g.dt = tm.dt
                P = _lib.ptr
                _lib.check(_lib.lib.pss_update_b3d(self.ctx.gp, P(self.Ex), P(self.Ey), P(self.Ez), P(self.Bx),
                                                   P(self.By), P(self.Bz), st), "pss_update_b3d")
            self.absorber.apply(self.Bx, self.By, self.Bz, src=src, stream=stream)
        elif curl:
            self.cpml.UpdateB_ApplyCPML_B(self.Bx, self.By, self.Bz, self.Ex, self.Ey, self.Ez, tm.dt, src=src,
                                          stream=stream)
        else:
            self.cpml.ApplyCPML_B(self.Bx, self.By, self.Bz, self.Ex, self.Ey, self.Ez, dt=tm.dt, src=src,
                                  stream=stream)
        self._advance(tm.dt / 2.0, st)
        tm.t += tm.dt / 2.0
        tm.n += 1
        self._nstep += 1
        if self.check_every and self._nstep % self.check_every == 0:
            self.check_finite()

    def _advance(self, dt2, st):
        if self._tdev is not None:
            _lib.check(_lib.lib.pss_add_scalar(_lib.ptr(self._tdev), dt2, st), "pss_add_scalar")

    def check_finite(self):
        for F in (self.Ex, self.Ey, self.Ez):
            if not bool(torch.isfinite(torch.view_as_real(F)).all()):
                raise RuntimeError(f"NaN or overflow at step {self.time.n}")

    def run(self, nsteps=None):
        n = self.time.CalcNt() + 1 - self.time.n if nsteps is None else nsteps
        for _ in range(n):
            self.step()


def InitializeFields(Nx, Ny, Nz, device=None):
    """Nine zero (Nx, Ny, Nz) complex128 device tensors: Ex, Ey, Ez, Bx, By, Bz, Jx, Jy, Jz
    (PSTD3D/PSTD3D.py:114-131)."""
    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    return tuple(torch.zeros((Nx, Ny, Nz), dtype=torch.complex128, device=dev) for _ in range(9))


_default = None  # the last propagator run by PSTD_3D_Propagator (not thread-safe, as the reference)


def PSTD_3D_Propagator(space, time, pulse, *, source_type="ic", boundary_type="mask", calc_j=None,
                       output_dir="output", snapshot_interval=500, **kw):
    """Run a whole simulation (PSTD3D/PSTD3D.py:864-909): build the propagator, keep it as
    the module's default and run it to the end of the time grid."""
    global _default
    _default = PSTD3DPropagator(space, time, pulse, source_type=source_type, boundary_type=boundary_type,
                                calc_j=calc_j, output_dir=output_dir, snapshot_interval=snapshot_interval, **kw)
    _default.run()
    return _default
