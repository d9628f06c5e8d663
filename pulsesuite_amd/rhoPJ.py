"""Wire <-> grid coupling on the GPU: the ``calc_j`` hook of the coupled step.

Follows QWArray.QuantumWire (PSTD3D/rhoPJ.py:1060-1329) for the current path:
the wire sees E sampled at one point (EAtXYZ, PSTD3D/usefulsubs.py:2016-2085),
each sub-band wire w = 1..Nw runs its SBE step on the x line, the current
J(x) = sum_w dP_w/dt / Nw + CalcJfx (rhoPJ.py:561-592, 1183-1194) is deposited
as J(x) g_y(y) g_z(z) (QWPlacement, rhoPJ.py:736-781) and enters the E update
in k-space.  Because the deposit is separable, its 3-D FFT is the product of
three 1-D FFTs: J_hat(kx,ky,kz) = FFT_x(J)(kx) FFT_y(g_y)(ky) FFT_z(g_z)(kz); the
E-update kernel forms it on the fly and no J grid is stored or transformed.

Two current semantics (``current=``):
  * "reference" (default): what the reference's code does.  QuantumWire re-creates
    PxxOld/PyyOld/PzzOld/RhoOld as zero arrays on every call (rhoPJ.py:1138-1147)
    before it forms J = sum_w (P_w - P_old_w)/dt/Nw (:1183-1194), so J = sum_w P_w/dt/Nw
    and CalcJfx sees RhoOld = 0.
  * "delta": the previous call's P and rho are kept, so J = (P - P_old)/dt, the time
    derivative that the reference's comment (:1182) describes.
The longitudinal-field diagnostics of QuantumWire (ElongfromRho, RhoBSpec,
RhoBAnalytic, RhoBSmooth, RhoBAnSmooth; rhoPJ.py:784-969, 1206-1285) are not on the J
path: ``QWArray(longitudinal=True)`` computes them on the device after every wire step,
as the reference does, into device grids of the array; ``ElongfromRho`` and
``QWPlacement`` are also module functions with the reference's signatures.
"""

import ctypes as C
import weakref

import numpy as np
import torch
from scipy.constants import epsilon_0 as eps0

from . import _lib
from .SBEs import DeviceWires
from .typespace import GetDx, GetDy, GetDz, GetEpsr, GetKxArray, GetXArray, GetYArray, GetZArray, ss
from .wire_model import build_model

_CDT = torch.complex128


def _dev_c(a, device):
    """A contiguous complex128 device tensor of a (numpy or torch) array."""
    if isinstance(a, torch.Tensor):
        return a.to(device=device, dtype=_CDT).contiguous()
    return torch.as_tensor(np.ascontiguousarray(a, dtype=np.complex128), device=device)


def _check_grid(t, shape, name):
    if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == _CDT and t.is_contiguous()
            and tuple(t.shape) == tuple(shape)):
        raise TypeError(f"{name}: a contiguous complex128 CUDA tensor of shape {tuple(shape)} is required")


_space_ctx_cache = {}


def _space_ctx(space, device):
    """Grid context (FFT-order wavenumbers and twiddles) of a space, cached."""
    from ._grid import GridContext

    key = (space.Nx, space.Ny, space.Nz, GetDx(space), GetDy(space), GetDz(space), GetEpsr(space), str(device))
    if key not in _space_ctx_cache:
        _space_ctx_cache[key] = GridContext(space, 1.0, with_cpml=False, device=device)
    return _space_ctx_cache[key]


def _place(out, terms, nw, dx, stream=None):
    """pss_qw_place: terms are (lines, lstride, gx, gy, gz, sign, mode) with device tensors
    (gx may be None); lines[w * lstride + i] of complex entries."""
    n = len(terms)
    P = _lib.ptr
    lines = (C.c_void_p * n)(*(C.c_void_p(t[0].data_ptr()) for t in terms))
    stride = (C.c_int64 * n)(*(t[1] for t in terms))
    gx = (C.c_void_p * n)(*(P(t[2]) for t in terms))
    gy = (C.c_void_p * n)(*(P(t[3]) for t in terms))
    gz = (C.c_void_p * n)(*(P(t[4]) for t in terms))
    sign = (C.c_double * n)(*(t[5] for t in terms))
    mode = (C.c_int32 * n)(*(t[6] for t in terms))
    Nx, Ny, Nz = out.shape
    _lib.check(_lib.lib.pss_qw_place(Nx, Ny, Nz, nw, n, lines, stride, gx, gy, gz, sign, mode, float(dx), P(out),
                                     _lib.stream_ptr(stream)), "pss_qw_place")


def QWPlacement(Fwire, gx, gy, gz, Fgrid):
    """Fgrid[i, j, k] = gx[i] gy[j] gz[k] Fwire[i] on the device (PSTD3D/rhoPJ.py:736-781);
    Fgrid a complex128 CUDA grid (written in place), the lines numpy or torch arrays."""
    _check_grid(Fgrid, Fgrid.shape, "Fgrid")
    dev = Fgrid.device
    f, a, b, c = (_dev_c(v, dev) for v in (Fwire, gx, gy, gz))
    _place(Fgrid, [(f, 0, a, b, c, 1.0, 0)], 1, 1.0)


def ElongfromRho(space, Rho, Px, Py, Pz, ExlfromPRho, EylfromPRho, EzlfromPRho, ExlfromP, EylfromP, EzlfromP,
                 ExlfromRho, EylfromRho, EzlfromRho, RhoBSpec, _ctx=None, _scratch=None):
    """Longitudinal fields from P and the free charge (PSTD3D/rhoPJ.py:784-969) on the device.

    All arguments are complex128 CUDA grids [Nx][Ny][Nz].  As in the reference, Px, Py and
    Pz are transformed to k-space in place; Rho is not changed; the nine fields and
    RhoBSpec are written in real space.  Runs the library's 3-D FFTs and the k-space loop
    of pss_elong_k (rhoPJ.py:596-700)."""
    shape = (space.Nx, space.Ny, space.Nz)
    names = ("Rho", "Px", "Py", "Pz", "ExlfromPRho", "EylfromPRho", "EzlfromPRho", "ExlfromP", "EylfromP",
             "EzlfromP", "ExlfromRho", "EylfromRho", "EzlfromRho", "RhoBSpec")
    grids = (Rho, Px, Py, Pz, ExlfromPRho, EylfromPRho, EzlfromPRho, ExlfromP, EylfromP, EzlfromP, ExlfromRho,
             EylfromRho, EzlfromRho, RhoBSpec)
    for g, nm in zip(grids, names):
        _check_grid(g, shape, nm)
    if space.Nx < 2:
        raise ValueError("ElongfromRho: Nx >= 2 (the reference's small = (qx[1] - qx[0]) / 1e4)")
    ctx = _ctx if _ctx is not None else _space_ctx(space, Rho.device)
    st = _lib.stream_ptr()
    Rk = _scratch if _scratch is not None else torch.empty_like(Rho)
    Rk.copy_(Rho)
    for F in (Rk, Px, Py, Pz):
        _lib.check(_lib.lib.pss_fft3d(ctx.gp, _lib.ptr(F), -1, st), "pss_fft3d")
    outs = grids[4:]
    arr = (C.c_void_p * 10)(*(_lib.ptr(o) for o in outs))
    qx = GetKxArray(space)
    small = (qx[1] - qx[0]) / 1e4
    inv_eps = 1.0 / (eps0 * GetEpsr(space))
    _lib.check(_lib.lib.pss_elong_k(ctx.gp, _lib.ptr(Rk), _lib.ptr(Px), _lib.ptr(Py), _lib.ptr(Pz), arr, small,
                                    inv_eps, st), "pss_elong_k")
    for o in outs:
        _lib.check(_lib.lib.pss_fft3d(ctx.gp, _lib.ptr(o), +1, st), "pss_fft3d")


LONGITUDINAL = ("Rho", "Px", "Py", "Pz", "ExlfromPRho", "EylfromPRho", "EzlfromPRho", "ExlfromP", "EylfromP",
                "EzlfromP", "ExlfromRho", "EylfromRho", "EzlfromRho", "RhoBSpec", "RhoBAnalytic", "RhoBSmooth",
                "RhoBAnSmooth")


def _locator(x, x0):
    i = np.searchsorted(x, x0, side="right") - 1
    return int(max(0, min(i, len(x) - 2)))


def trilinear(shape, rx, ry, rz, x0, y0, z0):
    """Flat indices and weights of the 8 corners of EAtXYZ (usefulsubs.py:2016-2085)."""
    axes = []
    for a, a0 in ((rx, x0), (ry, y0), (rz, z0)):
        if len(a) == 1:
            axes.append((0, np.array([1.0, 0.0])))
        else:
            i = _locator(a, a0)
            al = (a0 - a[i]) / (a[i + 1] - a[i])
            axes.append((i, np.array([1.0 - al, al])))
    Nx, Ny, Nz = shape
    idx, w = [], []
    for r in range(2):
        for q in range(2):
            for p in range(2):
                ii, jj, kk = axes[0][0] + p, axes[1][0] + q, axes[2][0] + r
                ww = axes[0][1][p] * axes[1][1][q] * axes[2][1][r]
                if ii < Nx and jj < Ny and kk < Nz and ww != 0.0:
                    idx.append((ii * Ny + jj) * Nz + kk)
                    w.append(ww)
                else:
                    idx.append(-1)
                    w.append(0.0)
    return np.array(idx, dtype=np.int32), np.array(w)


class QWArray:
    """Nw sub-band wires along x through (x0, y0, z0), coupled to a PSTD3DPropagator.

    qw_params / mb_flags configure the wires (see SBEs.SBESolver); Emax0 sets
    the wireoff gate (1e-3 Emax0); current selects the current semantics, "reference"
    (J = P/dt, the reference's code) or "delta" (J = dP/dt); integrator "leapfrog" (the
    reference's wire step) or "rk4" (DeviceWires.step_rk4: eager, not capturable);
    longitudinal=True: after every wire step, QuantumWire's diagnostics (PSTD3D/rhoPJ.py:
    1206-1285) into ``self.longitudinal_grids`` (names in LONGITUDINAL): rho and P placed
    on the grid (P left in k-space by ElongfromRho, as in the reference), the fields of
    ElongfromRho, RhoBSpec, RhoBAnalytic, RhoBSmooth, RhoBAnSmooth.
    """

    separable = True

    def __init__(self, qw_params, mb_flags, Nw=1, Emax0=1.0, y0=0.0, z0=0.0, x0=0.0, ay=5e-9, az=5e-9,
                 precision="fp64", current="reference", integrator="leapfrog", longitudinal=False):
        if current not in ("reference", "delta"):
            raise ValueError(f"current must be 'reference' or 'delta', not {current!r}")
        if integrator not in ("leapfrog", "rk4"):
            raise ValueError(f"integrator must be 'leapfrog' or 'rk4', not {integrator!r}")
        self.integrator = integrator
        self.current = current
        self.qw, self.mb = qw_params, mb_flags
        self.Nw = Nw
        self.Emax0 = Emax0
        self.x0, self.y0, self.z0, self.ay, self.az = x0, y0, z0, ay, az
        self.precision = precision
        self.longitudinal = bool(longitudinal)
        self.longitudinal_grids = None
        self._prop = None

    @property
    def prop(self):
        """The propagator this array is bound to (held weakly: the propagator holds the
        array as its calc_j hook, and a strong reference back would make a cycle that only
        the cyclic collector frees, at an arbitrary later moment, with its device buffers
        and captured graphs)."""
        return self._prop() if self._prop is not None else None

    def bind(self, prop):
        """Allocate device state for propagator ``prop`` (called by PSTD3DPropagator)."""
        self._prop = weakref.ref(prop)
        sp = prop.space
        dev = prop.device
        self.rx, self.ry, self.rz = GetXArray(sp), GetYArray(sp), GetZArray(sp)
        Nx = sp.Nx
        self.model = build_model(self.qw, self.mb, self.rx)
        self.wires = DeviceWires(self.model, self.Nw, self.Emax0, prop.time.dt, self.precision, dev)
        idx, w = trilinear((sp.Nx, sp.Ny, sp.Nz), self.rx, self.ry, self.rz, self.x0, self.y0, self.z0)
        self.sidx = torch.as_tensor(idx, device=dev)
        self.sw = torch.as_tensor(w, device=dev)
        self.sample = torch.zeros(3, dtype=torch.complex128, device=dev)
        self.P = self.wires.P_all  # (Nw, 3, Nx)
        self.Rho = self.wires.Rho_all  # (Nw, Nx)
        self.Pold = torch.zeros_like(self.P)
        self.RhoOld = torch.zeros_like(self.Rho)
        self.jl = torch.zeros((3, Nx), dtype=torch.complex128, device=dev)
        gy = np.exp(-((self.ry - self.y0) ** 2) / (2.0 * self.ay**2)).astype(complex)
        gz = np.exp(-((self.rz - self.z0) ** 2) / (2.0 * self.az**2)).astype(complex)
        self.gy = torch.as_tensor(gy, device=dev)
        self.gz = torch.as_tensor(gz, device=dev)
        ctx = prop.ctx
        st = _lib.stream_ptr()
        # FFT of the transverse gates, once: passes along y and z of (1, Ny, 1) / (1, 1, Nz) views
        self._gline_y = _line_ctx(1, sp.Ny, 1, dev)
        self._gline_z = _line_ctx(1, 1, sp.Nz, dev)
        # the current line's FFT along x: in the current-line kernel (one workgroup) up to
        # 256 cells, where that is faster than a separate launch (Nx = 128: 10.6 us for both
        # against 6.3 + 8.1 us); longer lines take the multi-workgroup line FFT
        self._fft_in_jline = sp.Nx <= 256
        self._gline_x = None if self._fft_in_jline else _line_ctx(Nx, 1, 1, dev)
        if sp.Ny > 1:
            _lib.check(_lib.lib.pss_fft_line(self._gline_y.gp, _lib.ptr(self.gy), _lib.ptr(self.gy), 1, -1, st), "fft gy")
        if sp.Nz > 1:
            _lib.check(_lib.lib.pss_fft_line(self._gline_z.gp, _lib.ptr(self.gz), _lib.ptr(self.gz), 2, -1, st), "fft gz")
        self.dx = GetDx(sp)
        if self.longitudinal:
            if Nx < 2:
                raise ValueError("longitudinal diagnostics need Nx >= 2")
            shp = (sp.Nx, sp.Ny, sp.Nz)
            self.longitudinal_grids = {n: torch.zeros(shp, dtype=_CDT, device=dev) for n in LONGITUDINAL}
            self._lscratch = torch.zeros(shp, dtype=_CDT, device=dev)
            gyr = np.exp(-((self.ry - self.y0) ** 2) / (2.0 * self.ay**2))
            gzr = np.exp(-((self.rz - self.z0) ** 2) / (2.0 * self.az**2))
            self._gates = {k: torch.as_tensor(v.astype(complex), device=dev) for k, v in (
                ("gy", gyr), ("gz", gzr), ("dgy", -(self.ry - self.y0) / self.ay**2 * gyr),
                ("dgz", -(self.rz - self.z0) / self.az**2 * gzr))}
            self._lctx = ctx

    def sample_args(self):
        return (self.sidx, self.sw, self.sample)

    def jsep(self):
        return (self.jl, self.gy, self.gz)

    def line_args(self):
        """The wires' input lines as one block [Nw][3][Nx] (for a sample kernel that fills
        them itself), with Nx and Nw."""
        return (self.wires.E_all, self.wires.nrr[0], self.Nw)

    def step_filled(self, space, time):
        """calc_j hook when the input lines already hold the sample (filled by the sample
        kernel): the wires step and the J line is deposited."""
        self.wires.set_dt(time.dt)
        self._step_wires()
        self.deposit(space, time)

    def __call__(self, space, time, Ex, Ey, Ez, Jx, Jy, Jz):
        """calc_j hook: wires step on the sampled field, J line -> k-space factors."""
        self.drive(space, time)
        self._step_wires()
        self.deposit(space, time)

    def _step_wires(self):
        # the hook runs beside the field half steps: a split-K of at most 4 (mixed) or 8
        # (FP64) keeps the contractions' workgroups from crowding the field passes (measured
        # at Nk = 256, 128^3: caps 1/2/4/8 gave 0.634/0.598/0.591/0.596 ms mixed and
        # 1.137/0.865/0.813/0.807 ms FP64 per pipelined step)
        if self.integrator == "rk4":
            self.wires.step_rk4()
            return
        self.wires.step(ksplit_cap=4 if self.precision == "mixed" else 8)

    def drive(self, space, time):
        """Wire lines <- the point sample taken during the E half step."""
        st = _lib.stream_ptr()
        W = self.wires
        for i in range(self.Nw):
            _lib.check(_lib.lib.pss_fill_line(_lib.ptr(self.sample), _lib.ptr(W.E_in[i]), space.Nx, st), "fill")
        W.set_dt(time.dt)

    def deposit(self, space, time):
        """J line from the wires' new P and rho, then its FFT along x."""
        st = _lib.stream_ptr()
        ink = self._fft_in_jline and space.Nx > 1
        _lib.check(_lib.lib.pss_jline(_lib.ptr(self.P), _lib.ptr(self.Rho), _lib.ptr(self.Pold), _lib.ptr(self.RhoOld),
                                      self.Nw, space.Nx, time.dt, self.dx, _lib.ptr(self.jl), 1 if ink else 0,
                                      1 if self.current == "delta" else 0, st),
                   "jline")
        if space.Nx > 1 and not ink:
            lines = (_lib.C.c_void_p * 3)(*(_lib.ptr(self.jl[c]) for c in range(3)))
            _lib.check(_lib.lib.pss_fft_line_batch(self._gline_x.gp, lines, lines, 3, 0, -1, st), "fft jline")
        if self.longitudinal:
            self._diagnostics(space)

    def _diagnostics(self, space):
        """QuantumWire's longitudinal fields and bound-charge diagnostics (PSTD3D/rhoPJ.py:
        1206-1285) from this step's wire outputs: rho and P (means over the sub-band wires)
        placed with the gates, ElongfromRho, RhoBAnalytic (the product rule: -dPx/dx gy gz -
        Py dgy gz - Pz gy dgz, central difference in x), and RhoBSmooth / RhoBAnSmooth (Gaussian
        k-space filter of widths 2 dx, 2 dy, 2 dz)."""
        G, g = self.longitudinal_grids, self._gates
        Nx, Nw = space.Nx, self.Nw
        P, R = self.P, self.Rho  # (Nw, 3, Nx), (Nw, Nx)
        _place(G["Rho"], [(R, Nx, None, g["gy"], g["gz"], 1.0, 0)], Nw, self.dx)
        for c, nm in enumerate(("Px", "Py", "Pz")):
            _place(G[nm], [(P[:, c], 3 * Nx, None, g["gy"], g["gz"], 1.0, 0)], Nw, self.dx)
        _place(G["RhoBAnalytic"], [(P[:, 0], 3 * Nx, None, g["gy"], g["gz"], -1.0, 1),
                                    (P[:, 1], 3 * Nx, None, g["dgy"], g["gz"], -1.0, 0),
                                    (P[:, 2], 3 * Nx, None, g["gy"], g["dgz"], -1.0, 0)], Nw, self.dx)
        ElongfromRho(space, *(G[n] for n in LONGITUDINAL[:14]), _ctx=self._lctx, _scratch=self._lscratch)
        st = _lib.stream_ptr()
        gp = self._lctx.gp
        sx, sy, sz = 2.0 * GetDx(space), 2.0 * GetDy(space), 2.0 * GetDz(space)
        for src, dst in (("RhoBSpec", "RhoBSmooth"), ("RhoBAnalytic", "RhoBAnSmooth")):
            G[dst].copy_(G[src])
            _lib.check(_lib.lib.pss_fft3d(gp, _lib.ptr(G[dst]), -1, st), "pss_fft3d")
            _lib.check(_lib.lib.pss_kfilter(gp, _lib.ptr(G[dst]), sx, sy, sz, st), "pss_kfilter")
            _lib.check(_lib.lib.pss_fft3d(gp, _lib.ptr(G[dst]), +1, st), "pss_fft3d")


_line_cache = {}


def _line_ctx(Nx, Ny, Nz, dev):
    from ._grid import GridContext

    key = (Nx, Ny, Nz, str(dev))
    if key not in _line_cache:
        _line_cache[key] = GridContext(ss(3, Nx, Ny, Nz, 1.0, 1.0, 1.0, 1.0), 1.0, with_cpml=False, device=dev)
    return _line_cache[key]
