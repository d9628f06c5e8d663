"""Quantum-wire SBE solver on MI355X: drop-in for PulseSuite's ``SBEs`` module.

Same call surface as PSTD3D/SBEs.py: ``SBESolver(q, rr, r0, Emaxxx, lam, Nw, QW,
backend)`` (:4635), the module shims ``InitializeSBE`` / ``QWCalculator`` /
``chiqw`` / ``getqc`` / ``QWArea`` / ``ShutOffOptics`` (:6020-6078), and
``QWCalculator(Exx, Eyy, Ezz, Vrr, rr, q, dt, w, Pxx, Pyy, Pzz, Rho, DoQWP,
DoQWDl)`` with 1-based ``w`` and in-place outputs.  The wire state lives on the
GPU; every step runs in libpss (there is no host fallback).  With
``write_fields=True`` the reference's files of the jmax cadence are written (the state,
energies, dephasing and scattering rates, the fields and sources on both grids; see
``pulsesuite_amd.output``) on a host thread; the per-step info and PdotE files are not.

Extra, GPU-native entry points: ``step_all`` advances all Nw wires in one
batched launch from device-resident field lines; ``get_state`` / ``set_state``
read and write (p1, p2, C1, C2, D1, D2) for checkpointing and tests.
"""

import ctypes as C
import os

import numpy as np
import torch

from . import _lib
from .wire_model import MBFlags, QWParams, build_model, read_mb_params, read_qw_params

__all__ = ["SBESolver", "InitializeSBE", "QWCalculator", "chiqw", "getqc", "QWArea", "ShutOffOptics",
           "WriteSBEsData", "ReadSBEsData", "dpdt", "dCdt", "dDdt", "DeviceWires", "QWParams", "MBFlags"]

_CDT = torch.complex128

# stage bits of pss_sbe_stages (include/pss.h)
ST_FRONT, ST_TRACES, ST_OFFDIAG, ST_CALCH, ST_RHS, ST_COLL, ST_TAIL = 1, 2, 4, 8, 16, 32, 64
ST_DERIV = ST_FRONT | ST_TRACES | ST_OFFDIAG | ST_CALCH | ST_RHS


def _dev(device):
    if device is None:
        device = torch.device("cuda", torch.cuda.current_device())
    return torch.device(device)


# read-only device tables per (model, name, device), shared by every DeviceWires
_DEVICE_TABLES = {}


class DeviceWires:
    """Device-resident state, tables, workspace and descriptors of Nw wires sharing one model.

    ``models`` may be one WireModel (the SBESolver case, every wire identical up
    to its index w) or a list with one model per wire (batched wires of
    different Nk).  ``nrr`` is the propagation-grid length of every wire.
    """

    def __init__(self, models, Nw, Emax0, dt=0.0, precision="fp64", device=None, w_index=None, groups=None):
        self.device = _dev(device)
        if not isinstance(models, (list, tuple)):
            models = [models] * Nw
        if len(models) != Nw:
            raise ValueError("need one model per wire")
        self.models = list(models)
        self.Nw = Nw
        self.precision = precision
        self.prec_code = _lib.FP64 if precision == "fp64" else _lib.MIXED
        self.dt = float(dt)
        dev = self.device
        self._keep = []
        self._tables = {}
        self.state = []
        self.ctl = torch.zeros((Nw, 4), dtype=torch.int32, device=dev)
        # wireoff latch: one per solver group (group id per wire; default all wires share one)
        self.group = list(groups) if groups is not None else [0] * Nw
        self.on = torch.zeros(max(self.group) + 1, dtype=torch.int32, device=dev)
        nrr = [len(m.arrays["sa_x"]) for m in self.models]
        self.nrr = nrr
        if len(set(nrr)) == 1:
            # one block per quantity so the coupling kernels see [Nw][3][n] / [Nw][n]
            self.E_all = torch.zeros((Nw, 3, nrr[0]), dtype=_CDT, device=dev)
            self.P_all = torch.zeros((Nw, 3, nrr[0]), dtype=_CDT, device=dev)
            self.Rho_all = torch.zeros((Nw, nrr[0]), dtype=_CDT, device=dev)
            self.E_in = [self.E_all[i] for i in range(Nw)]
            self.P_out = [self.P_all[i] for i in range(Nw)]
            self.Rho_out = [self.Rho_all[i] for i in range(Nw)]
        else:
            self.E_all = self.P_all = self.Rho_all = None
            self.E_in = [torch.zeros((3, n), dtype=_CDT, device=dev) for n in nrr]
            self.P_out = [torch.zeros((3, n), dtype=_CDT, device=dev) for n in nrr]
            self.Rho_out = [torch.zeros(n, dtype=_CDT, device=dev) for n in nrr]
        self.ws = []
        self.descs = (_lib.WireDesc * Nw)()
        w_index = w_index or list(range(1, Nw + 1))
        self.w_index = w_index
        self.gate = 1e-3 * Emax0
        for i, m in enumerate(self.models):
            self._setup_wire(i, m, w_index[i])
        self.d_descs = torch.empty(C.sizeof(self.descs), dtype=torch.uint8, device=dev)
        self.upload_descs()
        self.max_Nk = max(m.Nk for m in self.models)
        self.max_Nr = max(m.Nr for m in self.models)
        self.max_nrr = max(nrr)

    # ------------------------------------------------------------ set-up
    def _cached(self, m, name):
        e = _DEVICE_TABLES.get((id(m), name, str(self.device)))
        return e[1] if e is not None and e[0] is m else None

    def _store(self, m, name, t):
        if len(_DEVICE_TABLES) > 65536:
            _DEVICE_TABLES.clear()
        _DEVICE_TABLES[(id(m), name, str(self.device))] = (m, t)
        return t

    def _table(self, m, name):
        t = self._cached(m, name)
        if t is None:
            a = m.arrays[name]
            if np.iscomplexobj(a):
                t = torch.as_tensor(np.ascontiguousarray(a), dtype=_CDT).view(torch.float64)
            elif a.dtype.kind in "iu":
                t = torch.as_tensor(np.ascontiguousarray(a, dtype=np.int32))
            else:
                t = torch.as_tensor(np.ascontiguousarray(a, dtype=np.float64))
            t = self._store(m, name, t.to(self.device))
        return t

    def _v0(self, m):
        """V(q) rows (eh, ee, hh) of the model on the device, computed once per model."""
        V0 = self._cached(m, "V0")
        if V0 is None:
            s = m.scalars
            y = self._table(m, "r")
            q = torch.as_tensor(m.arrays["Vq"], dtype=torch.float64, device=self.device)
            V0 = torch.empty((3, m.Nk), dtype=torch.float64, device=self.device)
            st = _lib.stream_ptr()
            for n, (a1, a2) in enumerate(((s["ae"], s["ah"]), (s["ae"], s["ae"]), (s["ah"], s["ah"]))):
                _lib.check(_lib.lib.pss_vint(_lib.ptr(y), m.Nr, _lib.ptr(q), m.Nk, a1, a2, s["Delta0"], s["vpref"],
                                             _lib.ptr(V0[n]), st), "pss_vint")
            self._store(m, "V0", V0)
        return V0

    def _setup_wire(self, i, m, w):
        dev = self.device
        Nk, Nr = m.Nk, m.Nr
        # the dephasing table and the collision-line recurrence need parabolic bands on a
        # uniform grid: E[k] = c kr[k]^2 with kr[k] = dkr (k - k0)
        kr = m.arrays["kr"]
        for name in ("Ee", "Eh"):
            E = m.arrays[name]
            c = E[0] / kr[0] ** 2
            if not np.allclose(E, c * kr**2, rtol=1e-12, atol=0.0):
                raise ValueError(f"{name} is not parabolic in kr; the device kernels need E = c kr^2")
        if Nk > 1 and not np.allclose(np.diff(kr), kr[1] - kr[0], rtol=1e-12, atol=0.0):
            raise ValueError("kr is not uniform; the device kernels need a uniform k grid")
        nrr = self.nrr[i]
        s = m.scalars
        st = {}
        for name in ("p", "C", "D"):
            st[name] = torch.zeros((3, Nk, Nk), dtype=_CDT, device=dev)
        fe = torch.as_tensor(m.arrays["fe0"], dtype=_CDT, device=dev)
        fh = torch.as_tensor(m.arrays["fh0"], dtype=_CDT, device=dev)
        for b in range(3):
            st["C"][b].diagonal().copy_(fe)
            st["D"][b].diagonal().copy_(fh)
        self.state.append(st)
        wsb = _lib.lib.pss_sbe_workspace_bytes(Nk, Nr, nrr, self.prec_code)
        ws = torch.zeros(wsb // 8, dtype=torch.float64, device=dev)
        self.ws.append(ws)
        d = self.descs[i]
        d.Nk, d.Nr, d.nrr, d.w = Nk, Nr, nrr, w
        d.flags, d.precision = m.flags, self.prec_code
        d.dt = self.dt
        for k in ("L", "gap", "ehint", "area", "dcv", "dkr", "dr_r", "drr", "gam_e", "gam_h", "gam_eh", "Vscale",
                  "N0", "hOph", "hGph", "Gph", "density_max"):
            setattr(d, k, float(s[k]))
        d.Edc = float(s.get("Edc", 0.0))
        for j, a in enumerate(s.get("rsp", (0.0,) * 4)):
            d.rsp[j] = float(a)
        d.yw = float((-1) ** int(np.floor((w - 1) / 2.0)))
        d.zsign = float((-1) ** w)
        d.gate = self.gate
        for k in ("kr", "Ee", "Eh", "win", "r", "chie", "chih", "qe", "qh", "ktab", "xe", "xh", "iqv", "twid",
                  "sa_x", "sa_t", "sa_b", "sa_h", "sa_end", "sa_idx", "sa_dx", "sb_t", "sb_b", "sb_h", "sb_end",
                  "sb_idx", "sb_dx"):
            setattr(d, k, self._table(m, k).data_ptr())
        d.V0 = self._v0(m).data_ptr()
        for b in range(3):
            d.p[b] = st["p"][b].data_ptr()
            d.C[b] = st["C"][b].data_ptr()
            d.D[b] = st["D"][b].data_ptr()
        d.ctl = self.ctl[i].data_ptr()
        d.on = self.on[self.group[i]:self.group[i] + 1].data_ptr()
        d.Exx, d.Eyy, d.Ezz = (self.E_in[i][c].data_ptr() for c in range(3))
        d.Pxx, d.Pyy, d.Pzz = (self.P_out[i][c].data_ptr() for c in range(3))
        d.Rho = self.Rho_out[i].data_ptr()
        d.ws = ws.data_ptr()

    def upload_descs(self):
        raw = bytes(memoryview(self.descs).cast("B"))
        host = torch.frombuffer(bytearray(raw), dtype=torch.uint8)
        self.d_descs.copy_(host)

    def set_desc(self, i, **fields):
        """Change fields of wire i's descriptor (flags, gam_e, gam_h, gam_eh, ...) and upload
        the descriptors when anything changed."""
        d = self.descs[i]
        changed = False
        for k, v in fields.items():
            if getattr(d, k) != v:
                setattr(d, k, v)
                changed = True
        if changed:
            self.upload_descs()

    def set_dt(self, dt):
        dt = float(dt)
        if dt != self.dt:
            self.dt = dt
            for i in range(self.Nw):
                self.descs[i].dt = dt
            self.upload_descs()

    # ------------------------------------------------------------ stepping
    def step(self, wires=None, stream=None, ksplit_cap=0):
        """One step of the wires in the contiguous range ``wires`` (default: all);
        ksplit_cap > 0 caps the contractions' split-K factor (a step that shares the GPU
        with other work does better with fewer workgroups)."""
        lo, hi = (0, self.Nw) if wires is None else wires
        sz = C.sizeof(_lib.WireDesc)
        ptr = C.c_void_p(self.d_descs.data_ptr() + lo * sz)
        sub = self.models[lo:hi]
        _lib.check(_lib.lib.pss_sbe_step_p(ptr, hi - lo, min(m.Nk for m in sub), max(m.Nk for m in sub),
                                           max(m.Nr for m in sub),
                                           max(self.nrr[lo:hi]), precision_mask(self.prec_code, sub), ksplit_cap,
                                           _lib.stream_ptr(stream)), "pss_sbe_step_p")

    def step_rk4(self, wires=None, stream=None):
        """One step of the wires in ``wires`` with the classic RK4 integrator in place of the
        leapfrog: the coherent right-hand side (dpdt, dCdt, dDdt with the Hamiltonian,
        screening and dephasing of each stage's state, pss_sbe_stages with dX/dt only) in
        the stage order of _rk4 (libpulsesuite/integrator.py:143-157), the combines on
        pss_rk4_stage; then the relaxation, transport, renormalisation and polarisation of
        the step (its collision, phonon and emission rates from the updated state).  The
        prev buffer holds y during the step.  Eager only: it reads the wires' step counters
        and the wireoff latch on the host.  A wire group that is gated off takes the plain
        step (zero outputs, nothing advanced)."""
        if torch.cuda.is_current_stream_capturing():
            raise RuntimeError("the RK4 wire step reads step counters on the host and cannot be captured")
        lo, hi = (0, self.Nw) if wires is None else wires
        sub = self.models[lo:hi]
        pm = precision_mask(self.prec_code, sub)
        s = stream if stream is not None else torch.cuda.current_stream(self.device)
        sp = _lib.stream_ptr(s)
        self.kernel("pss_sbe_stages", pm, ST_FRONT, 0, wires=(lo, hi), stream=s)  # the wireoff gate
        on = self.on.cpu().numpy()
        if not all(on[self.group[i]] for i in range(lo, hi)):
            if any(on[self.group[i]] for i in range(lo, hi)):
                raise RuntimeError("RK4 step: wires of one launch must be all on or all off")
            self.step((lo, hi), s)
            return
        par = [int(v) & 1 for v in self.ctl[lo:hi, 0].cpu().numpy()]
        h = self.dt
        work = []
        with torch.cuda.stream(s):
            for i, pb in zip(range(lo, hi), par):
                st = self.state[i]
                cur, nxt = 1 + pb, 2 - pb
                scr = self._rk_scratch(i)
                for n in ("p", "C", "D"):
                    st[n][0].copy_(st[n][cur])  # y
                work.append((st, cur, nxt, scr))

            def rk(stage, n, y, dydt=None, dyt=None, dym=None, ytmp=None):
                _lib.check(_lib.lib.pss_rk4_stage(stage, 2 * y.numel(), float(h), _lib.ptr(y), _lib.ptr(dydt),
                                                  _lib.ptr(dyt), _lib.ptr(dym), _lib.ptr(ytmp), sp), "pss_rk4_stage")

            def deriv():
                self.kernel("pss_sbe_stages", pm, ST_DERIV, 1, wires=(lo, hi), stream=s)

            deriv()  # dydt -> next
            for st, cur, nxt, scr in work:
                for j, n in enumerate(("p", "C", "D")):
                    scr[0][j].copy_(st[n][nxt])
                    rk(0, n, st[n][0], dydt=st[n][nxt], ytmp=st[n][cur])
            deriv()  # dyt -> next
            for st, cur, nxt, scr in work:
                for j, n in enumerate(("p", "C", "D")):
                    scr[1][j].copy_(st[n][nxt])
                    rk(1, n, st[n][0], dyt=st[n][nxt], ytmp=st[n][cur])
            deriv()  # dym -> next
            for st, cur, nxt, scr in work:
                for j, n in enumerate(("p", "C", "D")):
                    rk(2, n, st[n][0], dyt=scr[1][j], dym=st[n][nxt], ytmp=st[n][cur])
                    scr[2][j].copy_(st[n][nxt])
            deriv()  # dyt -> next
            for st, cur, nxt, scr in work:
                for j, n in enumerate(("p", "C", "D")):
                    rk(3, n, st[n][0], dydt=scr[0][j], dyt=st[n][nxt], dym=scr[2][j])
                    st[n][cur].copy_(st[n][0])
                    st[n][nxt].copy_(st[n][0])
        # rates of the updated state, then relaxation ... polarisation on the next buffers
        self.kernel("pss_sbe_stages", pm, ST_FRONT, 0, wires=(lo, hi), stream=s)
        self.kernel("pss_sbe_stages", pm, ST_COLL | ST_TAIL, 0, wires=(lo, hi), stream=s)

    def _rk_scratch(self, i):
        """dydt, dyt, dym of wire i ([3][3][Nk][Nk]: stage value, then p, C, D), allocated on
        the first RK4 step."""
        sc = getattr(self, "_rk", None)
        if sc is None:
            sc = self._rk = {}
        if i not in sc:
            Nk = self.models[i].Nk
            sc[i] = torch.zeros((3, 3, Nk, Nk), dtype=_CDT, device=self.device)
        return sc[i]

    def kernel(self, name, *extra, wires=None, stream=None):
        """Run a kernel-level entry (pss_calc_h, pss_sbe_rhs, pss_coll_rates, pss_qwpol) on the
        wires in the contiguous range ``wires`` (default: all); ``extra`` are the entry's
        arguments after the size bounds (precision mask or rate selection)."""
        lo, hi = (0, self.Nw) if wires is None else wires
        sz = C.sizeof(_lib.WireDesc)
        ptr = C.c_void_p(self.d_descs.data_ptr() + lo * sz)
        sub = self.models[lo:hi]
        fn = getattr(_lib.lib, name)
        _lib.check(fn(ptr, hi - lo, min(m.Nk for m in sub), max(m.Nk for m in sub), max(m.Nr for m in sub),
                      max(self.nrr[lo:hi]), *extra, _lib.stream_ptr(stream)), name)

    def buffer(self, i, name, which):
        """Device view of wire i's state buffer ``which`` in ("prev", "cur", "next") of p, C or D
        (the rotation follows the wire's step counter)."""
        s = int(self.ctl[i, 0].item()) & 1
        b = {"prev": 0, "cur": 1 + s, "next": 2 - s}[which]
        return self.state[i][name][b]

    def packed_state(self, i, stream=None):
        """(p1, p2, C1, C2, D1, D2) of wire i as one device tensor [6][Nk][Nk] (pss_sbe_get_state)."""
        Nk = self.models[i].Nk
        out = torch.empty((6, Nk, Nk), dtype=_CDT, device=self.device)
        sz = C.sizeof(_lib.WireDesc)
        _lib.check(_lib.lib.pss_sbe_get_state(C.c_void_p(self.d_descs.data_ptr() + i * sz), Nk, _lib.ptr(out),
                                              _lib.stream_ptr(stream)), "pss_sbe_get_state")
        return out

    def load_packed_state(self, i, packed, stream=None):
        """Write wire i's (p1, p2, C1, C2, D1, D2) from a device tensor [6][Nk][Nk] (pss_sbe_set_state)."""
        Nk = self.models[i].Nk
        packed = packed.to(self.device, _CDT).contiguous()
        if tuple(packed.shape) != (6, Nk, Nk):
            raise ValueError(f"packed state must be [6][{Nk}][{Nk}]")
        sz = C.sizeof(_lib.WireDesc)
        _lib.check(_lib.lib.pss_sbe_set_state(C.c_void_p(self.d_descs.data_ptr() + i * sz), Nk, _lib.ptr(packed),
                                              _lib.stream_ptr(stream)), "pss_sbe_set_state")

    # ------------------------------------------------------------ state access
    def _cur(self, i):
        return 1 + (int(self.ctl[i, 0].item()) & 1)

    def get_state(self, i):
        st = self.state[i]
        c = self._cur(i)
        return {k: st[n][b].cpu().numpy().copy() for k, n, b in
                (("p1", "p", 0), ("p2", "p", c), ("C1", "C", 0), ("C2", "C", c), ("D1", "D", 0), ("D2", "D", c))}

    def set_state(self, i, p1, p2, C1, C2, D1, D2):
        st = self.state[i]
        c = self._cur(i)
        for n, b, a in (("p", 0, p1), ("p", c, p2), ("C", 0, C1), ("C", c, C2), ("D", 0, D1), ("D", c, D2)):
            st[n][b].copy_(torch.as_tensor(np.ascontiguousarray(a), dtype=_CDT))

    def force_on(self):
        self.on.fill_(1)

    def is_on(self, i):
        return bool(self.on[self.group[i]].item())

    def workspace_view(self, i, name):
        """Read-only view of a named workspace segment (diagnostics and tests)."""
        m = self.models[i]
        Nk, Nr, nrr = m.Nk, m.Nr, self.nrr[i]
        lay = ws_layout(Nk, Nr, nrr)
        sizes = {"vc": 3 * Nk, "e1d": Nk, "gam": 2 * Nk, "De": (2 * Nk + 1) * Nk, "Dh": (2 * Nk + 1) * Nk,
                 "H": 3 * Nk * Nk * 2, "W": 4 * Nk, "n3": 4 * Nk, "eq": 3 * (2 * Nk - 1) * 2, "pr": 3 * Nr,
                 "rho": 2 * Nr * 2, "er": 3 * Nr, "rsp": Nk, "dcc": 8 * Nk, "outr": 5 * nrr}
        o = lay[name]
        return self.ws[i][o:o + sizes[name]]


def precision_mask(prec_code, models=()):
    """pss_sbe_step_p's mask for a batch of wires of one precision: 1 FP64 or 2 mixed, plus 4
    when a wire has DCTrans and 8 when one has Recomb (those terms' kernels run only then)."""
    return (2 if prec_code == _lib.MIXED else 1) | feature_bits(models)


def feature_bits(models):
    f = 0
    for m in models:
        if m.mb.DCTrans:
            f |= 4
        if m.mb.Recomb:
            f |= 8
    return f


def ws_layout(Nk, Nr, nrr):
    """Offsets (in doubles) of the workspace segments; mirrors ws_layout in csrc/sbe.hip."""
    def al(n):
        return (n + 31) // 32 * 32
    nd = 2 * Nk - 1
    segs = [("er", 3 * Nr), ("spa", 3 * nrr), ("eq", 3 * nd * 2), ("vc", 3 * Nk), ("e1d", Nk), ("ne2", 2 * Nk),
            ("nh2", 2 * Nk), ("tr", 4 * nd * 2), ("gam", 2 * Nk), ("De", (2 * Nk + 1) * Nk),
            ("Dh", (2 * Nk + 1) * Nk), ("H", 3 * Nk * Nk * 2), ("W", 4 * Nk), ("n3", 4 * Nk), ("red", 32),
            ("ds", 4 * nd * 2), ("pr", 3 * Nr), ("rho", 2 * Nr * 2), ("spb", 5 * Nr), ("outr", 5 * nrr),
            ("cp", 2 * Nk * Nk), ("dcc", 8 * Nk), ("rsp", Nk)]
    np64 = (Nk + 63) // 64 * 64
    ksm = 16 if np64 <= 128 else 8 if np64 <= 256 else 4 if np64 <= 512 else 1
    out = {}
    o = 0
    for n, sz in segs:
        out[n] = o
        o += al(sz)
    out["pk"] = o  # split-K partial tiles
    o += al(6 * ksm * Nk * Nk if ksm > 1 else 0)
    out["total_fp64"] = o
    out["sp"] = o  # mixed precision only: FP16 operand planes
    o += al(16 * np64 * np64)
    out["total"] = o
    return out


class SBESolver:
    """Class-based SBE solver, GPU resident (PSTD3D/SBEs.py:4631-6013).

    Keyword-only extras: ``qw_params`` / ``mb_flags`` replace the files
    params/qw.params and params/mb.params read from the CWD; ``precision``
    selects "fp64" (default) or "mixed" contractions; ``device`` the GPU;
    ``integrator`` "leapfrog" (the reference's step) or "rk4" (DeviceWires.step_rk4);
    ``write_fields`` the reference's output files every jmax steps (dataQW/Wire/*, dataQW/
    Prop/*, PSTD3D/SBEs.py:5724-5727, 5931-5934, PSTD3D/qwoptics.py:798-799) written on a
    host thread (``flush_output`` waits for them).
    """

    def __init__(self, q, rr, r0, Emaxxx, lam, Nw, QW, backend="auto", *, qw_params=None, mb_flags=None,
                 precision="fp64", device=None, integrator="leapfrog", write_fields=False):
        self._backend = backend  # stored and not read, as the reference does (PSTD3D/SBEs.py:4750)
        if integrator not in ("leapfrog", "rk4"):
            raise ValueError(f"integrator must be 'leapfrog' or 'rk4', not {integrator!r}")
        self.integrator = integrator
        self.Nw = Nw
        self.QW = bool(QW)
        if not QW:
            return
        qw = qw_params if qw_params is not None else read_qw_params("params/qw.params")
        mb = mb_flags if mb_flags is not None else read_mb_params("params/mb.params")
        self.model = build_model(qw, mb, rr, r0)
        self.mb = self.model.mb
        s = self.model.scalars
        self.Nk, self.Nr = self.model.Nk, self.model.Nr
        self.L, self.area, self.qc, self.dcv, self.gap = s["L"], s["area"], s["qc"], s["dcv"], s["gap"]
        self.kr = self.model.arrays["kr"]
        self.r = self.model.arrays["r"]
        self.Emax0 = Emaxxx
        self.lam = lam
        self.rr = np.asarray(rr, dtype=np.float64)
        self.dev = DeviceWires(self.model, Nw, Emaxxx, 0.0, precision, device)
        self.t = 0.0
        self.xxx = 1
        self.jjj = 1
        self.jmax = qw.jmax
        self._on_host = [False] * Nw
        self.write_fields = bool(write_fields)
        self._writer = None
        if self.write_fields:
            from .output import AsyncWriter
            self._writer = AsyncWriter()

    def flush_output(self):
        """Wait until every submitted output file is written."""
        if self._writer is not None:
            self._writer.flush()

    # ---------------------------------------------------------- public step
    def QWCalculator(self, Exx, Eyy, Ezz, Vrr, rr, q, dt, w, Pxx, Pyy, Pzz, Rho, DoQWP, DoQWDl):
        """PSTD3D/SBEs.py:5838-5997.  Arrays may be numpy or CUDA tensors; outputs in place."""
        if w == 1:  # PSTD3D/SBEs.py:5849-5856
            if self.jjj == self.jmax:
                self.jjj = 0
            self.xxx += 1
            self.jjj += 1
            self.t += dt
        i = w - 1
        dev = self.dev
        dev.set_dt(dt)
        for c, F in enumerate((Exx, Eyy, Ezz)):
            dev.E_in[i][c].copy_(torch.as_tensor(F).to(dev.device, _CDT), non_blocking=True)
        write = self.write_fields and self.jjj == self.jmax and self.integrator == "leapfrog"
        Hd = None
        if self.integrator == "rk4":
            dev.step_rk4((i, i + 1))
        elif write:
            # the same stages in two launches: the Hamiltonian's diagonals (the Ee/Eh files)
            # are read before the tail, where a DCTrans shift reuses their memory
            pm = precision_mask(dev.prec_code, dev.models[i:i + 1])
            dev.kernel("pss_sbe_stages", pm, ST_DERIV, 0, wires=(i, i + 1))
            H = dev.workspace_view(i, "H").view(_CDT).reshape(3, self.Nk, self.Nk)
            Hd = (H[1].diagonal().cpu().numpy(), H[2].diagonal().cpu().numpy())
            dev.kernel("pss_sbe_stages", pm, ST_COLL | ST_TAIL, 0, wires=(i, i + 1))
        else:
            dev.step((i, i + 1))
        if write and dev.is_on(i):  # the reference returns before any output while the wire is off
            from .output import snapshot, write_snapshot
            lines = [np.zeros(len(rr), complex) if F is None else
                     np.asarray(F.cpu().numpy() if isinstance(F, torch.Tensor) else F) for F in (Exx, Eyy, Ezz, Vrr)]
            snap = snapshot(dev, i, H=Hd, flags=self.mb)
            snap["prop"] = lines + [snap["outr"][c].astype(complex) for c in range(3)]
            self._writer.submit(write_snapshot, snap, w, self.xxx, np.asarray(rr, dtype=np.float64))
        for c, F in enumerate((Pxx, Pyy, Pzz)):
            _copy_out(F, dev.P_out[i][c])
        _copy_out(Rho, dev.Rho_out[i])
        if not self._on_host[i]:
            self._on_host[i] = dev.is_on(i)
        if self._on_host[i]:
            DoQWP[0] = self.mb.Optics
            DoQWDl[0] = self.mb.LF

    def step_all(self, dt, stream=None):
        """Advance all Nw wires from the device lines ``dev.E_in`` (one launch)."""
        self.t += dt
        self.xxx += 1
        self.dev.set_dt(dt)
        if self.integrator == "rk4":
            self.dev.step_rk4(None, stream)
        else:
            self.dev.step(None, stream)

    # ---------------------------------------------------------- accessors
    def chiqw(self):
        from .wire_model import e0, eps0, hbar
        ww = 2 * np.pi * 2.99792458e8 / self.lam
        Ee = self.model.arrays["Ee"] + self.gap
        Eh = self.model.arrays["Eh"]
        geh = self.model.scalars["gam_eh"]
        den1 = Ee + Eh - 1j * hbar * geh - hbar * ww
        den2 = Ee + Eh + 1j * hbar * geh + hbar * ww
        return 4 * self.dcv**2 / eps0 / self.area * self.model.scalars["dkr"] / (2 * np.pi) * np.sum(
            (Ee + Eh) / (den1 * den2))

    def getqc(self):
        return self.qc

    def QWArea(self):
        return self.area

    def ShutOffOptics(self):
        self.mb.Optics = False
        for i in range(self.Nw):
            self.dev.descs[i].flags &= ~_lib.OPTICS
        self.dev.upload_descs()

    def write_sbes_data(self, n, backup_dir="dataQW/backup/"):
        """Checkpoint in the reference's text format (PSTD3D/SBEs.py:5536-5558).  The
        reference's YY2/CC2/DD2 (the averaged previous step) and YY3/CC3/DD3 (the newest
        step) are the device's prev and cur buffers; its YY1/CC1/DD1 (one step older, never
        read back by the step) are not kept on the device and are written as YY2/CC2/DD2."""
        Nk, Nw = self.Nk, self.Nw
        arrs = {k: np.zeros((Nk, Nk, Nw), complex) for k in BACKUP_KEYS}
        for w in range(Nw):
            st = self.dev.get_state(w)
            for X, p1, p2 in (("YY", "p1", "p2"), ("CC", "C1", "C2"), ("DD", "D1", "D2")):
                arrs[X + "1"][:, :, w] = st[p1]
                arrs[X + "2"][:, :, w] = st[p1]
                arrs[X + "3"][:, :, w] = st[p2]
        write_backup(arrs, n, backup_dir)

    def read_sbes_data(self, Nt, backup_dir="dataQW/backup/"):
        """Resume from a checkpoint (PSTD3D/SBEs.py:5560-5581): prev <- YY2/CC2/DD2,
        cur <- YY3/CC3/DD3, as the next step's _checkout reads them (:5073-5080)."""
        a = read_backup(self.Nk, self.Nw, Nt, backup_dir)
        for w in range(self.Nw):
            self.dev.set_state(w, p1=a["YY2"][:, :, w], p2=a["YY3"][:, :, w], C1=a["CC2"][:, :, w],
                               C2=a["CC3"][:, :, w], D1=a["DD2"][:, :, w], D2=a["DD3"][:, :, w])

    def get_state(self, w):
        return self.dev.get_state(w - 1)

    def set_state(self, w, **kw):
        self.dev.set_state(w - 1, **kw)


def _rhs(which, C, D, p, Heh, Hee, Hhh, GamE=None, GamH=None, Off=None):
    """One of dpdt / dCdt / dDdt on the device (pss_rhs_dense); NumPy in, NumPy out (copies
    to and from the GPU), CUDA tensors in, a CUDA tensor out."""
    host = not isinstance(p, torch.Tensor)
    dev = p.device if not host else torch.device("cuda", torch.cuda.current_device())

    def t(a, dt=_CDT):
        return None if a is None else torch.as_tensor(np.asarray(a) if not isinstance(a, torch.Tensor) else a,
                                                      dtype=dt, device=dev).contiguous()

    args = [t(C), t(D), t(p), t(Heh), t(Hee), t(Hhh), t(GamE, torch.float64), t(GamH, torch.float64), t(Off)]
    Nk = args[2].shape[0]
    out = torch.empty((Nk, Nk), dtype=_CDT, device=dev)
    _lib.check(_lib.lib.pss_rhs_dense(Nk, which, *(_lib.ptr(a) for a in args), _lib.ptr(out), _lib.stream_ptr()),
               "pss_rhs_dense")
    return out.cpu().numpy() if host else out


def dpdt(C, D, p, Heh, Hee, Hhh, GamE, GamH, OffP):
    """dp/dt of the SBEs (PSTD3D/SBEs.py:1850-1976), output indexed [kh, ke]."""
    return _rhs(0, C, D, p, Heh, Hee, Hhh, GamE, GamH, OffP)


def dCdt(Cee, Dhh, Phe, Heh, Hee, Hhh, GamE=None, GamH=None, OffE=None):
    """dC/dt (PSTD3D/SBEs.py:2221-2338); GamE, GamH, OffE are not used, as in the reference."""
    return _rhs(1, Cee, Dhh, Phe, Heh, Hee, Hhh)


def dDdt(Cee, Dhh, Phe, Heh, Hee, Hhh, GamE=None, GamH=None, OffH=None):
    """dD/dt (PSTD3D/SBEs.py:2513-2636); GamE, GamH, OffH are not used, as in the reference."""
    return _rhs(2, Cee, Dhh, Phe, Heh, Hee, Hhh)


BACKUP_KEYS = ("CC1", "CC2", "CC3", "DD1", "DD2", "DD3", "YY1", "YY2", "YY3")


def write_backup(arrays, n, backup_dir="dataQW/backup/"):
    """The reference's checkpoint files (PSTD3D/SBEs.py:5536-5558): for every key of
    BACKUP_KEYS one text file ``{key}.{n}.dat`` holding arrays[key] (complex [Nk][Nk][Nw]) as
    "re im" lines, wire by wire, k2 outer, k1 inner (shortest round-trip float text)."""
    os.makedirs(backup_dir, exist_ok=True)
    for key in BACKUP_KEYS:
        a = np.asarray(arrays[key])
        with open(os.path.join(backup_dir, f"{key}.{n}.dat"), "w", encoding="utf-8") as f:
            for w in range(a.shape[2]):
                col = a[:, :, w].T.reshape(-1)  # k1 fastest
                f.write("".join(f"{x!r} {y!r}\n" for x, y in zip(col.real.tolist(), col.imag.tolist())))


def read_backup(Nk, Nw, Nt, backup_dir="dataQW/backup/"):
    """Inverse of write_backup (PSTD3D/SBEs.py:5560-5581): {key: complex [Nk][Nk][Nw]}."""
    out = {}
    for key in BACKUP_KEYS:
        with open(os.path.join(backup_dir, f"{key}.{Nt}.dat"), "r", encoding="utf-8") as f:
            vals = np.array([complex(float(a), float(b)) for a, b in (line.split()[:2] for line in f)])
        if vals.size != Nk * Nk * Nw:
            raise ValueError(f"{key}.{Nt}.dat holds {vals.size} values, expected {Nk * Nk * Nw}")
        out[key] = vals.reshape(Nw, Nk, Nk).transpose(2, 1, 0).copy()
    return out


def _copy_out(dst, src):
    if isinstance(dst, torch.Tensor):
        dst.copy_(src.to(dst.device, dst.dtype))
    else:
        dst[:] = src.cpu().numpy()


_default_solver = None


def InitializeSBE(q, rr, r0, Emaxxx, lam, Nw, QW, **kw):
    """Module shim (PSTD3D/SBEs.py:6023-6026)."""
    global _default_solver
    _default_solver = SBESolver(q, rr, r0, Emaxxx, lam, Nw, QW, **kw)


def QWCalculator(Exx, Eyy, Ezz, Vrr, rr, q, dt, w, Pxx, Pyy, Pzz, Rho, DoQWP, DoQWDl):
    """Module shim (PSTD3D/SBEs.py:6029-6048)."""
    _default_solver.QWCalculator(Exx, Eyy, Ezz, Vrr, rr, q, dt, w, Pxx, Pyy, Pzz, Rho, DoQWP, DoQWDl)


def chiqw():
    return _default_solver.chiqw()


def getqc():
    return _default_solver.getqc()


def QWArea():
    return _default_solver.QWArea()


def ShutOffOptics():
    _default_solver.ShutOffOptics()


def WriteSBEsData(n):
    """Module shim (PSTD3D/SBEs.py:6071-6073)."""
    _default_solver.write_sbes_data(n)


def ReadSBEsData(Nt):
    """Module shim (PSTD3D/SBEs.py:6076-6078)."""
    _default_solver.read_sbes_data(Nt)
