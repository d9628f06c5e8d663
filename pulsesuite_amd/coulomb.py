"""Coulomb module on the GPU: drop-in for PulseSuite's PSTD3D/coulomb.py module surface.

``InitializeCoulomb(y, ky, L, Delta0, me, mh, Ee, Eh, ge, gh, alphae, alphah, er, Qy, kkp,
screened)`` (PSTD3D/coulomb.py:1351-1374) builds the module's singleton ``CoulombModule``:
a one-wire device context whose unscreened V(q) rows are integrated on the GPU (the
reference's Vint with its series K0, PSTD3D/coulomb.py:39-104) and whose 1-D
susceptibility rows follow CalcChi1D (:580-630).  Every interaction quantity of this path is
Toeplitz in k1 - k2 (V(|q|), E1D(|q|)); the device keeps rows, and the module functions
expand them into the reference's Nk x Nk (x 3) arrays.

  CalcScreenedArrays(screened, L, ne, nh, VC, E1D)   fills VC, E1D in place (:1386-1391)
  MBCE / MBCH(ne0, nh0, ky, Ee, Eh, VC, geh, g, Win, Wout)  accumulate into Win, Wout
                                                          (:1418-1423; cores :957-1071)
  SetLorentzDelta(flag)                                Lorentzian instead of Gaussian lines
  module attributes _qe, _qh, _Veh0, _Vee0, _Vhh0, _Chi1De, _Chi1Dh, _LorentzDelta, and
  _k3, _UnDel, _Ceh, _Cee, _Chh (:1427-1447): the device never stores the (Nk+1)^3
  collision tables (its kernels generate each entry where it is read); reading these five
  attributes builds them on the host, for inspection only.

Arrays may be NumPy arrays or CUDA tensors; outputs are written in place.  The rates run on
the step's own kernels (k_coll, k_wsum through pss_coll_rates; the screening through
pss_sbe_stages).  VC passed to MBCE / MBCH must be Toeplitz in k1 - k2, as the reference's
always is (its rows VC[q, 0, :] are what the device reads).
"""

import numpy as np
import torch

from . import _lib
from .SBEs import DeviceWires
from .wire_model import MBFlags, explicit_model

ST_FRONT = 1  # stage bit of the screening (pss_sbe_stages)

_instance = None
_LorentzDelta = False


def _np(a):
    return a.detach().cpu().numpy() if isinstance(a, torch.Tensor) else np.asarray(a)


def _write(dst, val):
    """dst[...] = val for a NumPy array or a tensor (in place)."""
    if isinstance(dst, torch.Tensor):
        dst.copy_(torch.as_tensor(val, dtype=dst.dtype))
    else:
        dst[...] = val


def _add(dst, val):
    if isinstance(dst, torch.Tensor):
        dst += torch.as_tensor(val, dtype=dst.dtype, device=dst.device)
    else:
        dst += val


def toeplitz(rows):
    """[..., Nk] rows of |k1 - k2| -> [Nk, Nk, ...] matrices."""
    rows = np.asarray(rows)
    Nk = rows.shape[-1]
    d = np.abs(np.arange(Nk)[:, None] - np.arange(Nk)[None, :])
    return np.moveaxis(rows[..., d], 0, -1) if rows.ndim == 2 else rows[d]


class CoulombModule:
    """PSTD3D/coulomb.py:1079-1164 with the device context of its functions."""

    def __init__(self, y, ky, L, Delta0, me, mh, Ee, Eh, ge, gh, alphae, alphah, er, Qy, kkp, screened,
                 LorentzDelta=False, device=None):
        self.ky = np.asarray(_np(ky), dtype=np.float64)
        self.Nk = len(self.ky)
        self.L, self.me, self.mh, self.ge, self.gh = float(L), float(me), float(mh), float(ge), float(gh)
        self.screened = bool(screened)
        self.LorentzDelta = bool(LorentzDelta)
        mb = MBFlags(Optics=False, Excitons=False, EHs=True, Screened=self.screened, Phonon=False, LF=False,
                     DiagDph=False, OffDiagDph=False, LorentzDelta=self.LorentzDelta)
        self.model = explicit_model(self.ky, _np(Ee), _np(Eh), mb, y=_np(y), L=L, Delta0=Delta0, me=me, mh=mh,
                                    ge=ge, gh=gh, alphae=alphae, alphah=alphah, er=er, Qy=_np(Qy), kkp=_np(kkp))
        self.dev = DeviceWires(self.model, 1, 0.0, 0.0, "fp64", device)
        self.dev.force_on()
        self._ee, self._eh = self.model.arrays["Ee"], self.model.arrays["Eh"]
        V0 = self.dev._v0(self.model).cpu().numpy()
        self.Veh0, self.Vee0, self.Vhh0 = (toeplitz(V0[n]) for n in range(3))
        self.qe, self.qh = toeplitz(self.model.arrays["qe"]), toeplitz(self.model.arrays["qh"])
        self.Chi1De, self.Chi1Dh = toeplitz(self.model.arrays["chie"]), toeplitz(self.model.arrays["chih"])

    def SetLorentzDelta(self, boolean):
        self.LorentzDelta = bool(boolean)
        f = self.dev.descs[0].flags
        f = (f | _lib.LORENTZDELTA) if self.LorentzDelta else (f & ~_lib.LORENTZDELTA)
        self.dev.set_desc(0, flags=f)

    # ---------------------------------------------------------------- screening
    def CalcScreenedArrays(self, screened, L, ne, nh, VC, E1D):
        """PSTD3D/coulomb.py:1195-1230: VC[:, :, n] = V0_n / E1D, E1D = Eps1D(n1D) when
        screened (else 1), with n1D = sum(Re ne + Re nh) / 2L clipped to the density limit."""
        d = self.dev
        f = d.descs[0].flags
        f = (f | _lib.SCREENED) if screened else (f & ~_lib.SCREENED)
        d.set_desc(0, flags=f, L=float(L))
        cur_c, cur_d = d.buffer(0, "C", "cur"), d.buffer(0, "D", "cur")
        cur_c.diagonal().copy_(torch.as_tensor(_np(ne), dtype=cur_c.dtype))
        cur_d.diagonal().copy_(torch.as_tensor(_np(nh), dtype=cur_d.dtype))
        d.kernel("pss_sbe_stages", 1, ST_FRONT, 0)
        vc = d.workspace_view(0, "vc").cpu().numpy().reshape(3, self.Nk)
        e1d = d.workspace_view(0, "e1d").cpu().numpy()
        _write(VC, toeplitz(vc))
        _write(E1D, toeplitz(e1d))

    def Eps1D(self, n1D, Nk=None):
        """PSTD3D/coulomb.py:1172-1193 (a closed form of the stored rows)."""
        return (np.ones_like(self.Chi1De) - self.Chi1De * 2 * np.log(np.abs((self.qe - np.pi * n1D) / (self.qe + n1D)))
                - self.Chi1Dh * 2 * np.log(np.abs((self.qh - np.pi * n1D) / (self.qh + n1D))))

    # ---------------------------------------------------------------- collision rates
    def _rates(self, ne0, nh0, VC, geh, g, holes):
        d = self.dev
        kw = dict(gam_eh=float(geh))
        kw["gam_h" if holes else "gam_e"] = float(g)
        d.set_desc(0, **kw)
        vc = _np(VC)[:, 0, :].T.astype(np.float64)  # Toeplitz rows V_n(|q|)
        d.workspace_view(0, "vc").copy_(torch.as_tensor(vc.reshape(-1)))
        for name, n in (("C", ne0), ("D", nh0)):
            nxt = d.buffer(0, name, "next")
            nxt.diagonal().copy_(torch.as_tensor(np.real(_np(n)).astype(complex)))
        d.kernel("pss_coll_rates", 1)
        W = d.workspace_view(0, "W").cpu().numpy().reshape(4, self.Nk)
        return (W[2], W[3]) if holes else (W[0], W[1])

    def MBCE(self, ne0, nh0, ky, Ee, Eh, VC, geh, ge, Win, Wout):
        """Electron in / out rates of carrier-carrier scattering, added to Win, Wout
        (PSTD3D/coulomb.py:957-1012, 1298-1309)."""
        wi, wo = self._rates(ne0, nh0, VC, geh, ge, False)
        _add(Win, wi)
        _add(Wout, wo)

    def MBCH(self, ne0, nh0, ky, Ee, Eh, VC, geh, gh, Win, Wout):
        """Hole in / out rates (PSTD3D/coulomb.py:1016-1071, 1311-1322)."""
        wi, wo = self._rates(ne0, nh0, VC, geh, gh, True)
        _add(Win, wi)
        _add(Wout, wo)

    # ---------------------------------------------------------------- host tables on request
    def _undel(self):
        """UnDel (PSTD3D/coulomb.py:336-364): (Nk+1)^2 ones, zero row and column 0 and diagonal."""
        u = np.ones((self.Nk + 1, self.Nk + 1))
        u[0, :] = 0.0
        u[:, 0] = 0.0
        np.fill_diagonal(u, 0.0)
        return u

    def _k3(self):
        """k3[k1, k2, k4] = k1 + k2 - k4 (1-based, 0 outside 1..Nk), PSTD3D/coulomb.py:271-299."""
        n = self.Nk
        a = np.arange(1, n + 1)
        k3 = a[:, None, None] + a[None, :, None] - a[None, None, :]
        return np.where((k3 >= 1) & (k3 <= n), k3, 0).astype(np.int32)

    def _mb_tables(self):
        """Ceh, Cee, Chh (PSTD3D/coulomb.py:367-499) on the host, (Nk+1)^3 each: slot
        (k1+1, k2+1, k4+1) holds the line shape of E1[k1] + E2[k2] - E2[k30] - E1[k4] with
        k30 = k4 + k2 - k1 (e-h) or k1 + k2 - k4 (e-e, h-h), zero when k1 = k4 or k2 = k30."""
        from scipy.constants import hbar

        n = self.Nk
        Ee, Eh = self._ee, self._eh
        A = np.arange(n)[:, None, None]
        B = np.arange(n)[None, :, None]
        Cc = np.arange(n)[None, None, :]
        out = []
        for E1, E2, g, eh in ((Ee, Eh, (self.ge + self.gh) / 2.0, True), (Ee, Ee, self.ge, False),
                              (Eh, Eh, self.gh, False)):
            k30 = Cc + B - A if eh else A + B - Cc
            ok = (k30 >= 0) & (k30 < n)
            kc = np.clip(k30, 0, n - 1)
            Ed = E1[A] + E2[B] - E2[kc] - E1[Cc]
            und = (A != Cc) & (B != kc)
            hw = hbar * g
            if self.LorentzDelta:
                val = 2.0 * g / (Ed**2 + hw**2)
            else:
                val = 2.0 * np.pi / hbar * np.exp(-((Ed / hw) ** 2)) / (np.sqrt(np.pi) * hw)
            T = np.zeros((n + 1, n + 1, n + 1))
            T[1:, 1:, 1:] = np.where(ok & und, val, 0.0)
            out.append(T)
        return out


def SetLorentzDelta(boolean):
    """PSTD3D/coulomb.py:1343-1348"""
    global _LorentzDelta
    _LorentzDelta = bool(boolean)
    if _instance is not None:
        _instance.SetLorentzDelta(boolean)


def InitializeCoulomb(y, ky, L, Delta0, me, mh, Ee, Eh, ge, gh, alphae, alphah, er, Qy, kkp, screened):
    """PSTD3D/coulomb.py:1351-1374"""
    global _instance
    _instance = CoulombModule(y, ky, L, Delta0, me, mh, Ee, Eh, ge, gh, alphae, alphah, er, Qy, kkp, screened,
                              LorentzDelta=_LorentzDelta)


def _require_instance():
    if _instance is None:
        raise ValueError("CoulombModule not initialized. Call InitializeCoulomb first.")
    return _instance


def CalcScreenedArrays(screened, L, ne, nh, VC, E1D):
    _require_instance().CalcScreenedArrays(screened, L, ne, nh, VC, E1D)


def Eps1D(n1D, Nk):
    return _require_instance().Eps1D(n1D, Nk)


def MBCE(ne0, nh0, ky, Ee, Eh, VC, geh, ge, Win, Wout, k3=None, Ceh=None, Cee=None):
    _require_instance().MBCE(ne0, nh0, ky, Ee, Eh, VC, geh, ge, Win, Wout)


def MBCH(ne0, nh0, ky, Ee, Eh, VC, geh, gh, Win, Wout, k3=None, Ceh=None, Chh=None):
    _require_instance().MBCH(ne0, nh0, ky, Ee, Eh, VC, geh, gh, Win, Wout)


def __getattr__(name):
    """Reads of coulomb._qe, ._Veh0, ... as the reference allows (PSTD3D/coulomb.py:1427-1447)."""
    direct = {"_qe": "qe", "_qh": "qh", "_Veh0": "Veh0", "_Vee0": "Vee0", "_Vhh0": "Vhh0", "_Chi1De": "Chi1De",
              "_Chi1Dh": "Chi1Dh", "_LorentzDelta": "LorentzDelta"}
    built = {"_k3", "_UnDel", "_Ceh", "_Cee", "_Chh"}
    if name in direct or name in built:
        if _instance is None:
            return None
        if name in direct:
            return getattr(_instance, direct[name])
        if name == "_k3":
            return _instance._k3()
        if name == "_UnDel":
            return _instance._undel()
        return dict(zip(("_Ceh", "_Cee", "_Chh"), _instance._mb_tables()))[name]
    raise AttributeError(f"module 'coulomb' has no attribute {name!r}")
