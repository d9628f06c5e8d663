"""Oracle: one quantum wire's SBE time step (PulseSuite ``SBESolver`` class path).

Test infrastructure only (see ``oracle/__init__.py``).  FP64 NumPy restatement of
PSTD3D/SBEs.py:4631-6013 and the physics modules it calls.  O(Nk^3) loops of
the reference are written in matrix form (GEMMs, diagonal-wise Toeplitz
products); ``tests/test_oracle_sbe.py`` checks every matrix form against a
literal loop transcription of the reference at small Nk.

Pinning (details in the tests):
  grid / kkp / Nk(L) / Fermi / phonon N0, Vscale, EP / GaussDelta / UnDel /
  k3 / dephasing index tables / Xcv0 / QWWindow       pinned (closed forms)
  K03 (accurate, cik01)                               pinned (scipy.special.k0)
  _K03_jit and Vint                                   pinned by limits (x->0 leading
                                                      term, asymptotic branch, cutoffs);
                                                      series branch value unpinned
  CalcH                                               pinned (closed forms: one-entry
                                                      exchange and direct terms)
  dpdt                                                pinned (free carriers: the band
                                                      frequency), + loop transcription
  QWPolarization3 / QWRho5 / Prop2QW                  pinned (one entry = one plane
                                                      wave; diagonal coherences = the
                                                      window profile; a cubic's windowed
                                                      real part)
  MBCE / MBCH / MBPE / MBPH / CalcGammaE/H /
  OffDiagDephasing2                                   pinned (each term vanishes at zero
                                                      and full occupation), + loop
                                                      transcription, W >= 0
  dCdt / dDdt                                         loop transcription + Hermiticity
                                                      (no closed form beyond the shared
                                                      free-carrier case)
  whole step (_sbe_calculator)                        free-carrier convergence to the
                                                      closed form, OBE special case,
                                                      RK4 order; otherwise build-defined
                                                      composition of the pinned parts
"""

from dataclasses import dataclass, field

import numpy as np
from scipy.constants import c as c0
from scipy.constants import e as e0
from scipy.constants import epsilon_0 as eps0
from scipy.constants import hbar
from scipy.constants import k as kB
from scipy.constants import m_e as me0
from scipy.special import k0 as _scipy_k0

from .helpers import (
    FFTG,
    GetArray0Index,
    GetKArray,
    GetSpaceArray,
    iFFTG,
    locator,
    rescale_1D_dpc,
)

eV = 1.602176634e-19  # PSTD3D/SBEs.py:79
pi = np.pi
twopi = 2.0 * np.pi
ii = 1j
TEMP = 77.0  # PSTD3D/phonons.py:22
EPSR0_HOST = 10.0  # PSTD3D/phonons.py:23
EPSRINF_HOST = 8.2  # PSTD3D/phonons.py:24


# ---------------------------------------------------------------------------
# Parameters (qw.params / mb.params)   PSTD3D/SBEs.py:4990-5034
# ---------------------------------------------------------------------------
@dataclass
class QWParams:
    L: float = 100e-9
    Delta0: float = 5e-9
    gap_eV: float = 1.5
    me_rel: float = 0.07
    mh_rel: float = 0.45
    HO_eV: float = 100e-3
    gam_e: float = 1e12
    gam_h: float = 1e12
    gam_eh: float = 1e12
    epsr: float = 9.1
    Oph_eV: float = 36e-3
    Gph_eV: float = 3e-3
    Edc: float = 0.0
    jmax: int = 500
    ntmax: int = 20000


@dataclass
class MBFlags:
    Optics: bool = True
    Excitons: bool = True
    EHs: bool = True
    Screened: bool = True
    Phonon: bool = True
    DCTrans: bool = False
    LF: bool = True
    FreePot: bool = False
    DiagDph: bool = True
    OffDiagDph: bool = True
    Recomb: bool = False
    PLSpec: bool = False
    ignorewire: bool = False
    Xqwparams: bool = False
    LorentzDelta: bool = False
    OBE: bool = False


def L_for_nk(nk):
    """Midpoint wire length giving an even Nk (inverse of PSTD3D/SBEs.py:4766-4770)."""
    gap = 1.5 * e0
    me = 0.07 * me0
    kmax = np.sqrt(1.2 * gap * 2.0 * me / hbar**2)
    m = nk // 2
    lo = m * pi / kmax
    hi = (m + 1) * pi / kmax
    return 0.5 * (lo + hi)


# ---------------------------------------------------------------------------
# Special functions
# ---------------------------------------------------------------------------
def K03_jit(x):
    """Series/asymptotic 'K0' used by the Numba Coulomb path.

    PSTD3D/coulomb.py:39-66, vectorised with the same per-element early exit.
    (Not the modified Bessel K0 for 0.5 < x < 2; copied as is.)
    """
    x = np.asarray(x, dtype=np.float64)
    out = np.zeros_like(x)
    big = x > 100.0
    tiny = (~big) & (x < 1e-10)
    small = (~big) & (~tiny) & (x < 2.0)
    large = (~big) & (~tiny) & (x >= 2.0)
    out[tiny] = 23.0
    xs = x[small]
    x2 = xs * xs
    res = -np.log(0.5 * xs) - 0.5772156649015329
    term = np.ones_like(xs)
    active = np.ones(xs.shape, dtype=bool)
    for k in range(1, 20):
        term = np.where(active, term * 0.25 * x2 / (k * k), term)
        res = np.where(active, res + term * (1.0 / k - 0.5772156649015329), res)
        active = active & ~(np.abs(term) < 1e-15)
        if not active.any():
            break
    out[small] = res
    xl = x[large]
    out[large] = np.sqrt(np.pi / (2.0 * xl)) * np.exp(-xl)
    return out


def K03(x):
    """Accurate K0 (cik01 path), 0 beyond x>100.  PSTD3D/usefulsubs.py:1794-1820"""
    x = np.asarray(x, dtype=np.float64)
    return np.where(x > 1e2, 0.0, _scipy_k0(np.minimum(x, 1e2)))


def FermiDistr(En):
    """PSTD3D/phonons.py:317-340"""
    return 1.0 / (np.exp(En / kB / TEMP) + 1.0)


def GaussDelta(a, b):
    """PSTD3D/coulomb.py:245-263 (vectorised)."""
    if abs(b) < 1e-300:
        return np.zeros_like(a)
    return 1.0 / (np.sqrt(pi) * b) * np.exp(-((a / b) ** 2))


def Vint(Qyk, y, alphae, alphah, Delta0):
    """Gaussian-overlap Coulomb integral with the Numba K0.  PSTD3D/coulomb.py:70-104,106-176"""
    Ny = len(y)
    N1 = Ny // 4
    N2 = 3 * Ny // 4
    aey2 = (alphae * y) ** 2
    ahy2 = (alphah * y) ** 2
    kmin = (alphae + alphah) / 4.0
    dk = max(abs(Qyk), kmin)
    multconst = alphae * alphah / np.pi * (y[1] - y[0]) ** 2
    yi = y[N1 : N2 + 1]
    r = np.sqrt((yi[:, None] - yi[None, :]) ** 2 + Delta0**2)
    ex = np.exp(-aey2[N1 : N2 + 1][:, None] - ahy2[N1 : N2 + 1][None, :])
    return float(np.sum(ex * multconst * K03_jit(dk * r)))


# ---------------------------------------------------------------------------
# The wire
# ---------------------------------------------------------------------------
@dataclass
class WireState:
    p1: np.ndarray
    p2: np.ndarray
    C1: np.ndarray
    C2: np.ndarray
    D1: np.ndarray
    D2: np.ndarray


class OracleWire:
    """One quantum wire (or Nw sub-band wires sharing tables), SBESolver semantics.

    Construction follows SBESolver.__init__ (PSTD3D/SBEs.py:4635-4984) with the
    parameter files replaced by dataclasses; file output is omitted.
    """

    def __init__(self, rr, r0, Emaxxx, lam, Nw, qw=None, mb=None, integrator="leapfrog"):
        self.qwp = qw or QWParams()
        self.mb = mb or MBFlags()
        q = self.qwp
        self.L = q.L
        self.Delta0 = q.Delta0
        self.gap = q.gap_eV * e0
        self.me = q.me_rel * me0
        self.mh = q.mh_rel * me0
        self.HO = q.HO_eV * e0
        self.gam_e = q.gam_e
        self.gam_h = q.gam_h
        self.epsr = q.epsr
        self.Oph = q.Oph_eV * e0 / hbar
        self.Gph = q.Gph_eV * e0 / hbar
        self.Emax0 = Emaxxx
        self.t = 0.0
        self.small = 1e-200
        self.Nw = Nw

        kmax = np.sqrt(1.2 * self.gap * 2.0 * self.me / (hbar**2))
        self.dkr = twopi / (2.0 * self.L)
        Nk = int(np.floor(kmax / self.dkr) * 2 + 1)
        self.Nr = Nk * 2
        self.Nk = Nk - 1
        Nk = self.Nk
        Nr = self.Nr

        # _get_arrays  PSTD3D/SBEs.py:5040-5052
        dnk = (Nk - 1) // 2
        self.NK0 = dnk + 1
        self.kr = np.array([self.dkr * (-dnk + k - 0.5) for k in range(Nk)])
        self.r = GetSpaceArray(Nr, 2 * self.L)
        Qr = GetKArray(Nr, 2 * self.L)
        Qr = np.roll(Qr, Nr // 2)
        Qr[0] = -Qr[0]
        self.Qr = Qr
        self.NQ0 = GetArray0Index(Qr)
        self.dr = (self.r[2] - self.r[1]) * (Nr - 1) / float(Nr)

        # material constants  PSTD3D/SBEs.py:4806-4835
        self.dcv = np.sqrt((e0 * hbar) ** 2 / (6.0 * me0 * self.gap) * (me0 / self.me - 1.0))
        self.alphae = np.sqrt(self.me * self.HO) / hbar
        self.alphah = np.sqrt(self.mh * self.HO) / hbar
        self.ehint = np.sqrt(2.0 * self.alphae * self.alphah / (self.alphae**2 + self.alphah**2))
        self.gam_eh = (self.gam_e + self.gam_h) / 2.0
        self.qc = 2.0 * self.alphae * self.alphah / (self.alphae + self.alphah)
        self.area = np.sqrt(2.0 * pi) / np.sqrt(self.alphae**2 + self.alphah**2) * self.Delta0

        self.Ee = hbar**2 * self.kr**2 / (2.0 * self.me)
        self.Eh = hbar**2 * self.kr**2 / (2.0 * self.mh)

        # initial state  PSTD3D/SBEs.py:4837-4854
        fermi_e = FermiDistr(self.Ee + self.gap / 2.0)
        fermi_h = FermiDistr(self.Ee + self.gap / 2.0)
        self.state = []
        for _ in range(Nw):
            C = np.diag(fermi_e).astype(complex)
            D = np.diag(fermi_h).astype(complex)
            P = np.zeros((Nk, Nk), dtype=complex)
            self.state.append(WireState(P.copy(), P.copy(), C.copy(), C.copy(), D.copy(), D.copy()))

        if self.mb.OBE:  # PSTD3D/SBEs.py:4884-4891
            m = self.mb
            m.Optics, m.Excitons, m.EHs, m.Phonon = True, False, False, False
            m.Recomb, m.LF, m.DCTrans = False, False, False

        # kkp  PSTD3D/SBEs.py:5054-5063, 3729-3739
        q = self.kr[:, None] - self.kr[None, :]
        self.kkp = (np.round(q / self.dkr)).astype(int) + self.NQ0

        self._init_optics()
        self._init_coulomb()
        self._init_phonons()
        self._init_dephasing()
        self._init_dc()
        if self.mb.Recomb:
            self._init_emission()

        rr_offset = np.asarray(rr) - r0
        self.Nr1 = locator(rr_offset, self.r[0])
        self.Nr2 = locator(rr_offset, self.r[Nr - 1])
        self.wph = (self.gap + hbar * self.Oph) / hbar
        self.wireoff = True
        if integrator not in ("leapfrog", "rk4"):
            raise ValueError(f"integrator must be 'leapfrog' or 'rk4', not {integrator!r}")
        self.integrator = integrator
        self.xxx = 1
        self.jjj = 1
        self.jmax = self.qwp.jmax

    # --- DC transport  PSTD3D/dcfield.py:782-793, 1169-1247 ------------------
    def _init_dc(self):
        """The conjugate grid of the momentum shift: FFT-order wavenumbers of an Nk-point
        grid of length (Nk - 1) dky, dky = kr[2] - kr[1] (dcfield.py:782-793, 39-68)."""
        kr = self.kr
        dky = kr[2] - kr[1] if self.Nk > 2 else kr[1] - kr[0]
        self.dcY = GetKArray(self.Nk, (self.Nk - 1) * dky)

    def Transport(self, X, Eac, dt, k1nek2):
        """In place: shift X in momentum by dk = -e (Edc + Eac) dt / hbar through the FFT
        (DCFieldModule.Transport, dcfield.py:1213-1247): both indices with a 2-D FFT when
        k1nek2 (ShiftN2D, :1191-1211), else the diagonal only (ShiftN1D, :1169-1189)."""
        dk = -e0 * (self.qwp.Edc + Eac) / hbar * dt
        Y = self.dcY
        if k1nek2:
            F = np.fft.fft2(X)
            F = F * np.exp(-ii * (Y[:, None] + Y[None, :]) * dk)
            X[:, :] = np.fft.ifft2(F)
        else:
            n = np.fft.fft(np.diag(X).copy())
            n = n * np.exp(-ii * Y * dk)
            np.fill_diagonal(X, np.fft.ifft(n))

    # --- spontaneous emission  PSTD3D/emission.py:36-200, 230-260 ---------------
    def _init_emission(self):
        """InitializeEmission (emission.py:36-91) with CalcHOmega (:230-260): the
        photon-energy midpoint grid HOmega = (n + 1/2) dhw, n < ceil(4 (kB T + hbar geh) / dhw),
        dhw = min(kB T, hbar geh) / 20, and the line weights `square`."""
        Nk = self.Nk
        self.idel = np.ones((Nk, Nk))
        np.fill_diagonal(self.idel, 0.0)
        kBT, hg = kB * TEMP, hbar * self.gam_eh
        hwmax = (kBT + hg) * 4.0
        dhw = min(kBT, hg) / 20.0
        Nw = int(np.ceil(hwmax / dhw))
        if Nw < 10:
            raise ValueError("Error: temperature is too low in emission.py")
        self.HOmega = (np.arange(Nw) + 0.5) * dhw
        dcv = abs(self.dcv)
        lrtz = (hg / pi) / (self.HOmega**2 + hg**2)
        self.square = (3.0 * dcv**2 * self.ehint / eps0 / np.sqrt(self.epsr) * self.ehint / hbar * lrtz
                       * np.exp(-self.HOmega / kB / TEMP))

    def Ec(self, ne, nh, VC):
        """Coulomb energy correction (emission.py:134-175):
        Ec[k] = sum_i ne_i (Vee[k,k] - Vee[i,k]) + nh_i (Vhh[k,k] - Vhh[i,k])
                + sum_{i != k} (ne_i - nh_i) Veh[k,k] - Veh[k,k]."""
        Veh, Vee, Vhh = VC[:, :, 0], VC[:, :, 1], VC[:, :, 2]
        vd = np.diag(Veh)
        return (ne @ (np.diag(Vee)[None, :] - Vee) + nh @ (np.diag(Vhh)[None, :] - Vhh)
                + (ne @ self.idel) * vd - (nh @ self.idel) * vd - vd)

    def SpontIntegral(self, Ek):
        """sum_n (HOmega_n + Ek) rho0(HOmega_n + Ek) square_n dhw, rho0(x) = x^2 / (c^3 pi^2 hbar^3)
        (emission.py:178-227), for an array of Ek."""
        dhw = self.HOmega[1] - self.HOmega[0]
        x = self.HOmega[None, :] + np.asarray(Ek, dtype=np.float64)[:, None]
        return np.sum(x * (x**2 / (c0**3 * pi**2 * hbar**3)) * self.square[None, :], axis=1) * dhw

    def SpontEmission(self, ne, nh, VC):
        """Rsp[k] = SpontIntegral(gap + Ee + Eh + Ec(Re ne, Re nh, VC)) (emission.py:94-131)."""
        Ek = self.gap + self.Ee + self.Eh + self.Ec(np.real(ne), np.real(nh), VC)
        return self.SpontIntegral(Ek)

    # --- QWOptics  PSTD3D/qwoptics.py:581-669 ----------------------------
    def _init_optics(self):
        Nk = self.Nk
        self.QWWindow = np.exp(-((self.r / (self.L / 2.0)) ** 150))
        self.Expikr = np.exp(1j * np.outer(self.kr, self.r))
        self.Expikrc = np.conj(self.Expikr)
        kh = np.arange(Nk)
        self.Xcv0 = np.tile((self.dcv * ((-1.0) ** (kh + 1)))[None, :], (Nk, 1)).astype(complex)
        self.Ycv0 = np.full((Nk, Nk), self.dcv, dtype=complex)
        self.Zcv0 = np.full((Nk, Nk), -self.dcv, dtype=complex)
        self.Xvc0 = np.conj(self.Xcv0.T)
        self.Yvc0 = np.conj(self.Ycv0.T)
        self.Zvc0 = np.conj(self.Zcv0.T)

    # --- CoulombModule.__init__  PSTD3D/coulomb.py:1124-1164 ----------------
    def _init_coulomb(self):
        Nk = self.Nk
        ky = self.kr
        ae, ah = self.alphae, self.alphah
        # MakeQs  PSTD3D/coulomb.py:302-333
        dky = ky[None, :] - ky[:, None]
        self.qe = np.maximum(np.abs(dky), ae / 2.0)
        self.qh = np.maximum(np.abs(dky), ah / 2.0)
        # CalcCoulombArrays  PSTD3D/coulomb.py:502-577 -- only the indices kkp reaches
        prefactor = e0**2 / (twopi * eps0 * self.epsr * self.L)
        used = np.unique(self.kkp)
        eh = np.zeros(len(self.Qr))
        ee = np.zeros(len(self.Qr))
        hh = np.zeros(len(self.Qr))
        for k in used:
            eh[k] = prefactor * Vint(self.Qr[k], self.r, ae, ah, self.Delta0)
            ee[k] = prefactor * Vint(self.Qr[k], self.r, ae, ae, self.Delta0)
            hh[k] = prefactor * Vint(self.Qr[k], self.r, ah, ah, self.Delta0)
        self.Veh0 = eh[self.kkp]
        self.Vee0 = ee[self.kkp]
        self.Vhh0 = hh[self.kkp]
        # CalcChi1D  PSTD3D/coulomb.py:580-630
        Re = np.sqrt((2.0 / ae) ** 2 + self.Delta0**2)
        Rh = np.sqrt((2.0 / ah) ** 2 + self.Delta0**2)
        scale = e0**2 / (twopi * eps0 * self.epsr * hbar**2)
        self.Chi1De = self.me * K03(self.qe * Re) / self.qe * scale
        self.Chi1Dh = self.mh * K03(self.qh * Rh) / self.qh * scale

    def Eps1D(self, n1D):
        """PSTD3D/coulomb.py:1172-1193"""
        return (
            np.ones_like(self.Chi1De)
            - self.Chi1De * 2 * np.log(np.abs((self.qe - pi * n1D) / (self.qe + n1D)))
            - self.Chi1Dh * 2 * np.log(np.abs((self.qh - pi * n1D) / (self.qh + n1D)))
        )

    def CalcScreenedArrays(self, ne, nh):
        """Returns VC (Nk,Nk,3), E1D.  PSTD3D/coulomb.py:1195-1230"""
        VC = np.stack([self.Veh0, self.Vee0, self.Vhh0], axis=2).astype(float)
        E1D = np.ones((self.Nk, self.Nk))
        if self.mb.Screened:
            density_1D = np.sum(np.real(ne) + np.real(nh)) / 2.0 / self.L
            density_max = min(self.qe[1, 1], self.qh[1, 1]) / pi * 0.99
            density_1D = min(density_1D, density_max)
            E1D = self.Eps1D(density_1D)
            VC = VC / E1D[:, :, None]
        return VC, E1D

    # --- phonons  PSTD3D/phonons.py:55-136 ----------------------------------
    def _init_phonons(self):
        Nk = self.Nk
        self.N0 = 1.0 / (np.exp(hbar * self.Oph / kB / TEMP) - 1.0)
        Ee, Eh = self.Ee, self.Eh
        hO, hG = hbar * self.Oph, hbar * self.Gph
        dE = Ee[:, None] - Ee[None, :]  # [k, k1]
        EP = self.N0 / ((dE - hO) ** 2 + hG**2) + (self.N0 + 1.0) / ((dE + hO) ** 2 + hG**2)
        dH = Eh[:, None] - Eh[None, :]
        HP = self.N0 / ((dH - hO) ** 2 + hG**2) + (self.N0 + 1.0) / ((dH + hO) ** 2 + hG**2)
        idel = 1.0 - np.eye(Nk)
        self.EP = EP * 2.0 * self.Gph * idel
        self.HP = HP * 2.0 * self.Gph * idel
        self.EPT = self.EP.T
        self.HPT = self.HP.T
        self.Vscale = hbar * self.Oph * self.epsr * (1.0 / EPSRINF_HOST - 1.0 / EPSR0_HOST)

    def MBPE(self, ne, VC, E1D, Win, Wout):
        """Accumulating phonon rates for electrons.  PSTD3D/phonons.py:140-205"""
        Vep = VC[:, :, 1] / E1D * self.Vscale
        Win += np.sum(Vep * ne[None, :] * self.EPT.T, axis=1)
        Wout += np.sum(Vep * (1.0 - ne)[None, :] * self.EP.T, axis=1)

    def MBPH(self, nh, VC, E1D, Win, Wout):
        """Accumulating phonon rates for holes.  PSTD3D/phonons.py:208-268"""
        Vhp = VC[:, :, 2] / E1D * self.Vscale
        Win += np.sum(Vhp * nh[:, None] * self.HPT, axis=0)
        Wout += np.sum(Vhp * (1.0 - nh)[:, None] * self.HP, axis=0)

    # --- dephasing tables  PSTD3D/dephasing.py:471-538 ----------------------
    def _init_dephasing(self):
        ky = self.kr
        Nk = self.Nk
        me, mh = self.me, self.mh
        dk = ky[1] - ky[0]
        kmax = ky[Nk - 1] + dk
        kmin = ky[0] - dk
        Nk0 = (Nk - 1) // 2 + 1

        def idx(v):
            return (np.round(np.clip(v, kmin, kmax) / dk)).astype(np.int64) + Nk0 - 1

        K = ky[:, None]  # [k, q]
        Q = ky[None, :]
        self.k_p_q = idx(K + Q)
        self.k_m_q = idx(K - Q)
        k1p_0 = ((me + mh) * Q - 2 * mh * K) / 2.0 / me
        self.k1p = idx(k1p_0)
        self.k1p_m_q = idx(k1p_0 - Q)
        k1_0 = ((me + mh) * Q - 2 * me * K) / 2.0 / mh
        self.k1 = idx(k1_0)
        self.k1_m_q = idx(k1_0 - Q)
        self.xe = me / hbar**2 * np.abs(ky) / (np.abs(ky) + dk * 1e-5) ** 2 / dk
        self.xh = mh / hbar**2 * np.abs(ky) / (np.abs(ky) + dk * 1e-5) ** 2 / dk
        dq = ky[1] - ky[0]
        self.iq_vxx2 = np.round(np.abs(ky / dq)).astype(np.int32)  # Vxx2  dephasing.py:362-378

    def CalcGammaE(self, ne0, nh0, VC):
        """PSTD3D/dephasing.py:29-66, 540-580"""
        Nk = self.Nk
        Veh2 = VC[self.iq_vxx2, 0, 0] ** 2
        Vee2 = VC[self.iq_vxx2, 0, 1] ** 2
        ne = np.real(ne0)
        nh = np.real(nh0)
        se, sh = 1.0 - ne, 1.0 - nh
        G = np.zeros(Nk)
        kpq = self.k_p_q
        ok = (kpq >= 0) & (kpq < Nk)
        kc = np.clip(kpq, 0, Nk - 1)
        term = pi / hbar * Vee2[None, :] * ne[kc] * se[kc] * np.abs(self.xe)[None, :]
        G += np.sum(np.where(ok, term, 0.0), axis=1)
        a, b, c = self.k1p_m_q, self.k1p, self.k_m_q
        ok = (a >= 0) & (a < Nk) & (b >= 0) & (b < Nk) & (c >= 0) & (c < Nk)
        a, b, c = np.clip(a, 0, Nk - 1), np.clip(b, 0, Nk - 1), np.clip(c, 0, Nk - 1)
        term = pi / hbar * Veh2[None, :] * (nh[a] * sh[b] * ne[c] + sh[a] * nh[b] * se[c]) * np.abs(self.xh)[None, :]
        G += np.sum(np.where(ok, term, 0.0), axis=1)
        return G

    def CalcGammaH(self, ne0, nh0, VC):
        """PSTD3D/dephasing.py:69-103, 582-624"""
        Nk = self.Nk
        Veh2 = VC[self.iq_vxx2, 0, 0] ** 2
        Vhh2 = VC[self.iq_vxx2, 0, 2] ** 2
        ne = np.real(ne0)
        nh = np.real(nh0)
        se, sh = 1.0 - ne, 1.0 - nh
        G = np.zeros(Nk)
        kpq = self.k_p_q
        ok = (kpq >= 0) & (kpq < Nk)
        kc = np.clip(kpq, 0, Nk - 1)
        term = pi / hbar * Vhh2[None, :] * nh[kc] * sh[kc] * np.abs(self.xh)[None, :]
        G += np.sum(np.where(ok, term, 0.0), axis=1)
        a, b, c = self.k1_m_q, self.k1, self.k_m_q
        ok = (a >= 0) & (a < Nk) & (b >= 0) & (b < Nk) & (c >= 0) & (c < Nk)
        a, b, c = np.clip(a, 0, Nk - 1), np.clip(b, 0, Nk - 1), np.clip(c, 0, Nk - 1)
        term = pi / hbar * Veh2[None, :] * (ne[a] * se[b] * nh[c] + se[a] * ne[b] * sh[c]) * np.abs(self.xe)[None, :]
        G += np.sum(np.where(ok, term, 0.0), axis=1)
        return G

    # --- off-diagonal dephasing  PSTD3D/dephasing.py:227-356, 381-414, 713-742 --
    def _vsq(self, VC):
        Nk = self.Nk
        Vsq = np.zeros((2 * Nk + 1, 3))
        for q in range(1, Nk):
            Vsq[q + Nk, :] = VC[q, 0, :] ** 2
            Vsq[-q + Nk, :] = VC[q, 0, :] ** 2
        return Vsq

    @staticmethod
    def _lrtz(a, b):
        return (b / pi) / (a**2 + b**2)

    def _offdiag_D(self, nA, nB, EA, EB, VsqAA, VsqAB, gAA, gAB):
        """Shared body of CalcOffDiagDeph_E2/H2 for carrier A with partner B."""
        Nk = self.Nk
        D = np.zeros((2 * Nk + 1, Nk))
        kk = np.arange(Nk)
        qq = np.arange(-Nk, Nk + 1)
        for k in range(Nk):
            # same-species term: q in [max(p-Nk,-k), min(p-1, Nk-k-1)], p-q in [0,Nk)
            P = kk[:, None]
            Q = qq[None, :]
            ok = (Q >= np.maximum(P - Nk, -k)) & (Q <= np.minimum(P - 1, Nk - k - 1))
            kpq = k + Q
            pmq = P - Q
            ok &= (kpq >= 0) & (kpq < Nk) & (pmq >= 0) & (pmq < Nk)
            kpqc = np.clip(kpq, 0, Nk - 1)
            pmqc = np.clip(pmq, 0, Nk - 1)
            Pc = np.broadcast_to(P, ok.shape)
            Ed = EA[kpqc] + EA[pmqc] - EA[Pc] - EA[k]
            val = VsqAA[Q + Nk] * self._lrtz(Ed, hbar * gAA) * (
                nA[pmqc] * (1.0 - nA[Pc]) * (1.0 - nA[k]) + (1.0 - nA[pmqc]) * nA[Pc] * nA[k]
            )
            D[:, k] += np.sum(np.where(ok, val, 0.0), axis=0)
        for k in range(Nk):
            P = kk[:, None]
            Q = qq[None, :]
            ok = (Q >= np.maximum(-P, -k)) & (Q <= np.minimum(Nk - P - 1, Nk - k - 1))
            kpq = k + Q
            ppq = P + Q
            ok &= (kpq >= 0) & (kpq < Nk) & (ppq >= 0) & (ppq < Nk)
            kpqc = np.clip(kpq, 0, Nk - 1)
            ppqc = np.clip(ppq, 0, Nk - 1)
            Pc = np.broadcast_to(P, ok.shape)
            Ed = EA[kpqc] + EB[ppqc] - EB[Pc] - EA[k]
            val = VsqAB[Q + Nk] * self._lrtz(Ed, hbar * gAB) * (
                nB[ppqc] * (1.0 - nB[Pc]) * (1.0 - nA[k]) + (1.0 - nB[ppqc]) * nB[Pc] * nA[k]
            )
            D[:, k] += np.sum(np.where(ok, val, 0.0), axis=0)
        return D * pi / hbar

    def CalcOffDiagDeph_E2(self, ne, nh, VC, gee, geh):
        Vsq = self._vsq(VC)
        return self._offdiag_D(np.abs(ne), np.abs(nh), self.Ee, self.Eh, Vsq[:, 1], Vsq[:, 0], gee, geh)

    def CalcOffDiagDeph_H2(self, ne, nh, VC, ghh, geh):
        Vsq = self._vsq(VC)
        return self._offdiag_D(np.abs(nh), np.abs(ne), self.Eh, self.Ee, Vsq[:, 2], Vsq[:, 0], ghh, geh)

    def OffDiagDephasing2(self, ne, nh, p, VC):
        """x[kh,ke]; PSTD3D/dephasing.py:321-356, 713-742 (as matrix products)."""
        Nk = self.Nk
        g = np.array([self.gam_eh, self.gam_e, self.gam_h])
        De = self.CalcOffDiagDeph_E2(ne, nh, VC, g[0], g[2])
        Dh = self.CalcOffDiagDeph_H2(ne, nh, VC, g[1], g[2])
        # Mh[kh, k'] = Dh[k'-kh+Nk, kh] (k' != kh); Me[ke, k'] = De[k'-ke+Nk, ke]
        a = np.arange(Nk)
        qidx = a[None, :] - a[:, None] + Nk
        Mh = Dh[qidx, a[:, None]] * (qidx != Nk)
        Me = De[qidx, a[:, None]] * (qidx != Nk)
        x = Mh @ p + p @ Me.T
        return x * ii * hbar, De, Dh

    # --- collision sums  PSTD3D/coulomb.py:367-499, 957-1071 ----------------
    def _ceh_read(self, a, b, c, kind):
        """Value the cores read as Ceh/Cee/Chh[a,b,c] with 0-based a,b,c.

        The tables are filled at 1-based slots (PSTD3D/coulomb.py:410-425) but the
        cores index them 0-based (:975,986,1000), so slot (a,b,c) holds the entry
        built for 0-based (a-1, b-1, c-1) and row/column 0 is zero.
        """
        Nk = self.Nk
        A, B, C = a - 1, b - 1, c - 1
        ok = (A >= 0) & (B >= 0) & (C >= 0)
        Ac, Bc, Cc = np.maximum(A, 0), np.maximum(B, 0), np.maximum(C, 0)
        if kind == "eh":
            k30 = Cc + Bc - Ac  # k3[c, b, a] 0-based = c+b-a
            E1, E2 = self.Ee, self.Eh
            g = hbar * self.gam_eh
            gam = self.gam_eh
        elif kind == "ee":
            k30 = Ac + Bc - Cc
            E1, E2 = self.Ee, self.Ee
            g = hbar * self.gam_e
            gam = self.gam_e
        else:
            k30 = Ac + Bc - Cc
            E1, E2 = self.Eh, self.Eh
            g = hbar * self.gam_h
            gam = self.gam_h
        ok = ok & (k30 >= 0) & (k30 < Nk)
        k30c = np.clip(k30, 0, Nk - 1)
        Ed = E1[Ac] + E2[Bc] - E2[k30c] - E1[Cc]
        und = (Ac != Cc) & (Bc != k30c)
        if self.mb.LorentzDelta:
            val = 2.0 * gam / (Ed**2 + g**2)
        else:
            val = twopi / hbar * GaussDelta(Ed, g)
        return np.where(ok & und, val, 0.0)

    def MBCE(self, ne0, nh0, VC, Win, Wout):
        """Accumulating carrier-carrier rates for electrons.  PSTD3D/coulomb.py:957-1012,1298-1309"""
        Nk = self.Nk
        Veh2 = VC[:, :, 0] ** 2
        Vee2 = VC[:, :, 1] ** 2
        ne = np.zeros(Nk + 1)
        nh = np.zeros(Nk + 1)
        ne[1:] = np.abs(ne0)
        nh[1:] = np.abs(nh0)
        q1 = np.arange(1, Nk + 1)[:, None]
        q2 = np.arange(1, Nk + 1)[None, :]
        for k in range(1, Nk + 1):
            kp, k1 = q1, q2
            k1p = kp + k1 - k  # k3[kp-1,k1-1,k-1]
            ok = (k1p >= 1) & (k1p <= Nk)
            k1pc = np.clip(k1p, 0, Nk)
            Win[k - 1] += np.sum(np.where(ok, Veh2[k - 1, k1 - 1] * (1.0 - nh[kp]) * nh[k1pc] * ne[k1]
                                          * self._ceh_read(k - 1, kp - 1, k1 - 1, "eh"), 0.0))
            k1p = kp + k - k1  # k3[kp-1,k-1,k1-1]
            ok = (k1p >= 1) & (k1p <= Nk)
            k1pc = np.clip(k1p, 0, Nk)
            Wout[k - 1] += np.sum(np.where(ok, Veh2[k1 - 1, k - 1] * (1.0 - ne[k1]) * (1.0 - nh[kp]) * nh[k1pc]
                                           * self._ceh_read(k1 - 1, kp - 1, k - 1, "eh"), 0.0))
            k2, k4 = q1, q2
            k30 = k + k2 - k4  # k3[k-1,k2-1,k4-1]
            ok = (k30 >= 1) & (k30 <= Nk)
            k30c = np.clip(k30, 0, Nk)
            Win[k - 1] += np.sum(np.where(ok, Vee2[k - 1, k4 - 1] * (1.0 - ne[k2]) * ne[k30c] * ne[k4]
                                          * self._ceh_read(k - 1, k2 - 1, k4 - 1, "ee"), 0.0))
            k30 = k4 + k2 - k  # k3[k4-1,k2-1,k-1]
            ok = (k30 >= 1) & (k30 <= Nk)
            k30c = np.clip(k30, 0, Nk)
            Wout[k - 1] += np.sum(np.where(ok, Vee2[k4 - 1, k - 1] * (1.0 - ne[k4]) * (1.0 - ne[k2]) * ne[k30c]
                                           * self._ceh_read(k4 - 1, k2 - 1, k - 1, "ee"), 0.0))

    def MBCH(self, ne0, nh0, VC, Win, Wout):
        """Accumulating carrier-carrier rates for holes.  PSTD3D/coulomb.py:1016-1071,1311-1322"""
        Nk = self.Nk
        Veh2 = VC[:, :, 0] ** 2
        Vhh2 = VC[:, :, 2] ** 2
        ne = np.zeros(Nk + 1)
        nh = np.zeros(Nk + 1)
        ne[1:] = np.abs(ne0)
        nh[1:] = np.abs(nh0)
        q1 = np.arange(1, Nk + 1)[:, None]
        q2 = np.arange(1, Nk + 1)[None, :]
        for kp in range(1, Nk + 1):
            k, k1 = q1, q2
            k1p = kp + k1 - k  # k3[kp-1,k1-1,k-1]
            ok = (k1p >= 1) & (k1p <= Nk)
            k1pc = np.clip(k1p, 0, Nk)
            Win[kp - 1] += np.sum(np.where(ok, Veh2[k - 1, k1 - 1] * (1.0 - ne[k]) * nh[k1pc] * ne[k1]
                                           * self._ceh_read(k - 1, kp - 1, k1 - 1, "eh"), 0.0))
            k1p = kp + k - k1  # k3[kp-1,k-1,k1-1]
            ok = (k1p >= 1) & (k1p <= Nk)
            k1pc = np.clip(k1p, 0, Nk)
            Wout[kp - 1] += np.sum(np.where(ok, Veh2[k1 - 1, k - 1] * (1.0 - ne[k]) * (1.0 - nh[k1pc]) * ne[k1]
                                            * self._ceh_read(k - 1, k1pc - 1, k1 - 1, "eh"), 0.0))
            k2p, k4p = q1, q2
            k3p = kp + k2p - k4p  # k3[kp-1,k2p-1,k4p-1]
            ok = (k3p >= 1) & (k3p <= Nk)
            k3pc = np.clip(k3p, 0, Nk)
            Win[kp - 1] += np.sum(np.where(ok, Vhh2[kp - 1, k4p - 1] * (1.0 - nh[k2p]) * nh[k3pc] * nh[k4p]
                                           * self._ceh_read(kp - 1, k2p - 1, k4p - 1, "hh"), 0.0))
            k3p = k4p + k2p - kp  # k3[k4p-1,k2p-1,kp-1]
            ok = (k3p >= 1) & (k3p <= Nk)
            k3pc = np.clip(k3p, 0, Nk)
            Wout[kp - 1] += np.sum(np.where(ok, Vhh2[k4p - 1, kp - 1] * (1.0 - nh[k4p]) * (1.0 - nh[k2p]) * nh[k3pc]
                                            * self._ceh_read(k4p - 1, k2p - 1, kp - 1, "hh"), 0.0))

    # --- Hamiltonians  PSTD3D/SBEs.py:820-935 -------------------------------
    def CalcH(self, Meh, C, D, p, VC):
        Nk = self.Nk
        Vq = np.zeros((Nk, 3))  # Vq[|q|, n], Vq[0]=0
        Vq[1:, :] = VC[1:, 0, :]
        Heh = Meh.astype(complex).copy()
        Hee = np.diag(self.Ee + self.gap).astype(complex)
        Hhh = np.diag(self.Eh).astype(complex)
        mb = self.mb
        if not mb.Excitons:
            return Heh, Hee, Hhh
        idx = np.arange(Nk)
        T = [Vq[np.abs(idx[:, None] - idx[None, :]), n] for n in range(3)]  # Toeplitz, zero diagonal

        def exch(X, Tn, full=True):
            # out[k1,k2] = sum_{q != 0} V(|q|) X[k2+q, k1+q]  (X^T along diagonals)
            Xt = X.T
            out = np.zeros((Nk, Nk), dtype=complex)
            for d in range(-(Nk - 1), Nk):
                if not full and d != 0:
                    continue
                i = np.arange(max(0, -d), min(Nk, Nk - d))
                m = len(i)
                x = Xt[i, i + d]
                out[i, i + d] = Tn[:m, :m] @ x
            return out

        if mb.LF:
            Heh += exch(p, T[0])
            Hee -= exch(C, T[1])
            Hhh -= exch(D, T[2])
            # direct terms: trace along diagonal q_d = k1-k2
            tC = np.zeros(2 * Nk - 1, dtype=complex)  # sum_k C[k, k+qd]
            tD = np.zeros(2 * Nk - 1, dtype=complex)  # sum_k D[k+qd, k]
            tD2 = np.zeros(2 * Nk - 1, dtype=complex)  # sum_k D[k, k+qd]
            tC2 = np.zeros(2 * Nk - 1, dtype=complex)  # sum_k C[k+qd, k]
            for qd in range(-(Nk - 1), Nk):
                tC[qd + Nk - 1] = np.trace(C, offset=qd)
                tD[qd + Nk - 1] = np.trace(D, offset=-qd)
                tD2[qd + Nk - 1] = np.trace(D, offset=qd)
                tC2[qd + Nk - 1] = np.trace(C, offset=-qd)
            QD = idx[:, None] - idx[None, :]
            Vd = [Vq[np.abs(QD), n] for n in range(3)]
            nz = QD != 0
            Hee += np.where(nz, Vd[1] * tC[QD + Nk - 1] - Vd[0] * tD[QD + Nk - 1], 0.0)
            Hhh += np.where(nz, Vd[2] * tD2[QD + Nk - 1] - Vd[0] * tC2[QD + Nk - 1], 0.0)
        else:
            Heh += exch(p, T[0])
            Hee -= exch(C, T[1], full=False)
            Hhh -= exch(D, T[2], full=False)
        return Heh, Hee, Hhh

    def calc_Meh(self, Ex, Ey, Ez, w):
        """PSTD3D/SBEs.py:5165-5175 with qwoptics.yw (PSTD3D/qwoptics.py:1099-1112)."""
        q = self.kkp
        ywv = (-1) ** int(np.floor((w - 1) / 2.0))
        return (-self.ehint * self.Xcv0 * Ex[q] - self.ehint * self.Ycv0 * Ey[q] * ywv
                - self.ehint * self.Zcv0 * Ez[q] * (-1) ** w)

    # --- RHS  PSTD3D/SBEs.py:1953-1976, 2320-2338, 2618-2636 ---------------
    @staticmethod
    def dpdt(C, D, p, Heh, Hee, Hhh, GamE, GamH, OffP):
        r = (Hhh @ p + p @ Hee.T - Heh.T @ C - (Heh @ D).T + Heh.T
             - ii * hbar * (GamE[None, :] + GamH[:, None]) * p + OffP)
        return r / (ii * hbar)

    @staticmethod
    def dCdt(C, D, p, Heh, Hee, Hhh):
        r = C @ Hee.T - Hee.T @ C + np.conj(p.T) @ Heh.T - np.conj(Heh) @ p
        return r / (ii * hbar)

    @staticmethod
    def dDdt(C, D, p, Heh, Hee, Hhh):
        r = D @ Hhh.T - Hhh.T @ D + np.conj(p) @ Heh - np.conj(Heh.T) @ p.T
        return r / (ii * hbar)

    # --- optics  PSTD3D/qwoptics.py:673-721, 815-946, 948-1077, 723-811 -----
    def Prop2QW(self, RR, Exx, Eyy, Ezz):
        out = []
        for F in (Exx, Eyy, Ezz):
            G = rescale_1D_dpc(RR, F, self.r)
            G = G * self.QWWindow
            G = np.real(G).astype(complex)
            FFTG(G)
            out.append(G)
        return out

    def QWPolarization3(self, p, w):
        Nk = self.Nk
        ywv = (-1) ** int(np.floor((w - 1) / 2.0))
        E, Ec = self.Expikr, self.Expikrc
        res = []
        for X, fac in ((self.Xvc0, 1.0), (self.Yvc0, ywv), (self.Zvc0, (-1) ** w)):
            S = np.sum(Ec * ((p * X) @ E), axis=0)  # sum_kh conj(e^{i kh r}) sum_ke (pX)[kh,ke] e^{i ke r}
            P = np.real(S) * self.QWWindow * fac
            P = P * 2 * self.ehint / self.area / self.L
            P = P.astype(complex)
            FFTG(P)
            res.append(P)
        return res

    def QWRho5(self, C, D, ne, nh):
        Nr = self.Nr
        dr = self.r[1] - self.r[0]
        NeTotal = np.sum(np.abs(ne)) + 1e-50
        NhTotal = np.sum(np.abs(nh)) + 1e-50
        E, Ec = self.Expikr, self.Expikrc
        re = np.sum(Ec * (C @ E), axis=0) * self.QWWindow / (2.0 * self.L)
        rh = np.sum(Ec * (np.conj(D) @ E), axis=0) * self.QWWindow / (2.0 * self.L)
        re = re - np.real(np.sum(re[0:10]) + np.sum(re[Nr - 10:Nr])) / 20.0
        rh = rh - np.real(np.sum(rh[0:10]) + np.sum(rh[Nr - 10:Nr])) / 20.0
        re = re * NeTotal / (np.sum(np.abs(re)) * dr + 1e-100)
        rh = rh * NhTotal / (np.sum(np.abs(rh)) * dr + 1e-100)
        FFTG(re)
        FFTG(rh)
        return re, rh

    def QW2Prop(self, Px, Py, Pz, re, rh, RR):
        dr = self.r[1] - self.r[0]
        dRR = RR[1] - RR[0]
        for F in (Px, Py, Pz, re, rh):
            iFFTG(F)
        total = None
        if self.mb.LF:
            re[:] = np.abs(re)
            rh[:] = np.abs(rh)
            total = (np.sum(re) * dr + np.sum(rh) * dr + 1e-100) / 2.0
            re[:] = re * total / (np.sum(re) * dr + 1e-100)
            rh[:] = rh * total / (np.sum(rh) * dr + 1e-100)
        Pxx = rescale_1D_dpc(self.r, Px, RR)
        Pyy = rescale_1D_dpc(self.r, Py, RR)
        Pzz = rescale_1D_dpc(self.r, Pz, RR)
        RhoE = rescale_1D_dpc(self.r, re, RR)
        RhoH = rescale_1D_dpc(self.r, rh, RR)
        if self.mb.LF:
            RhoE = np.abs(RhoE).astype(complex)
            RhoH = np.abs(RhoH).astype(complex)
            RhoE = RhoE * total / (np.sum(RhoE) * dRR + 1e-100)
            RhoH = RhoH * total / (np.sum(RhoH) * dRR + 1e-100)
        return Pxx, Pyy, Pzz, RhoE, RhoH

    # --- one SBE step  PSTD3D/SBEs.py:5624-5832 -----------------------------
    def _prep(self, C2, D2, p2, Ex, Ey, Ez, w):
        """_preparation (PSTD3D/SBEs.py:5185-5238) of the state (p2, C2, D2): screened V, E1D,
        the Hamiltonians, dephasing rates, off-diagonal dephasing and (Recomb) Rsp."""
        mb = self.mb
        Nk = self.Nk
        ne = np.diag(C2).copy()
        nh = np.diag(D2).copy()
        Meh = self.calc_Meh(Ex, Ey, Ez, w) if mb.Optics else np.zeros((Nk, Nk), complex)
        VC, E1D = self.CalcScreenedArrays(ne, nh)
        Heh, Hee, Hhh = self.CalcH(Meh, C2, D2, p2, VC)
        GamE = np.full(Nk, self.gam_e)
        GamH = np.full(Nk, self.gam_h)
        if mb.DiagDph:
            GamE = self.CalcGammaE(ne, nh, VC)
            GamH = self.CalcGammaH(ne, nh, VC)
        OffP = np.zeros((Nk, Nk), complex)
        if mb.OffDiagDph:
            OffP, _, _ = self.OffDiagDephasing2(ne, nh, p2, VC)
        Rsp = self.SpontEmission(ne, nh, VC) if mb.Recomb else np.zeros(Nk)
        return VC, E1D, (Heh, Hee, Hhh, GamE, GamH, OffP), Rsp

    def _rhs(self, p, C, D, Ex, Ey, Ez, w):
        """(dp/dt, dC/dt, dD/dt) of a state, with the Hamiltonian and dephasing of that state."""
        _, _, (Heh, Hee, Hhh, GamE, GamH, OffP), _ = self._prep(C, D, p, Ex, Ey, Ez, w)
        return (self.dpdt(C, D, p, Heh, Hee, Hhh, GamE, GamH, OffP), self.dCdt(C, D, p, Heh, Hee, Hhh),
                self.dDdt(C, D, p, Heh, Hee, Hhh))

    def _rk4_coherent(self, st, Ex, Ey, Ez, dt, w):
        """One classic RK4 step of dX/dt = RHS(X), X = (p, C, D), from the cur state, the field
        held over the step, with the stage order of _rk4 (libpulsesuite/integrator.py:143-157):
          ytmp = y + (h/2) dydt ; dyt = D(ytmp) ; ytmp = y + (h/2) dyt ; dym = D(ytmp)
          ytmp = y + h dym ; dym = dym + dyt ; dyt = D(ytmp) ; y + (h/6) (dydt + dyt + 2 dym)."""
        h = dt
        y = (st.p2, st.C2, st.D2)
        f = lambda x: self._rhs(x[0], x[1], x[2], Ex, Ey, Ez, w)  # noqa: E731
        dydt = f(y)
        ytmp = tuple(a + (h / 2.0) * b for a, b in zip(y, dydt))
        dyt = f(ytmp)
        ytmp = tuple(a + (h / 2.0) * b for a, b in zip(y, dyt))
        dym = f(ytmp)
        ytmp = tuple(a + h * b for a, b in zip(y, dym))
        dym = tuple(a + b for a, b in zip(dym, dyt))
        dyt = f(ytmp)
        return tuple(a + (h / 6.0) * ((b + c) + 2.0 * d) for a, b, c, d in zip(y, dydt, dyt, dym))

    def sbe_calculator(self, Ex, Ey, Ez, dt, w):
        mb = self.mb
        Nk = self.Nk
        st = self.state[w - 1]
        p1, p2, C1, C2, D1, D2 = st.p1, st.p2, st.C1, st.C2, st.D1, st.D2
        if self.integrator == "rk4":
            # the coherent RK4 update, then the relaxation, transport and renormalisation of
            # the leapfrog step with the rates of the updated state (an operator split)
            p3, C3, D3 = self._rk4_coherent(st, Ex, Ey, Ez, dt, w)
            VC, E1D, _, Rsp = self._prep(C3, D3, p3, Ex, Ey, Ez, w)
            p1, C1, D1 = p3, C3, D3
        else:
            VC, E1D, (Heh, Hee, Hhh, GamE, GamH, OffP), Rsp = self._prep(C2, D2, p2, Ex, Ey, Ez, w)
            dp = self.dpdt(C2, D2, p2, Heh, Hee, Hhh, GamE, GamH, OffP)
            dC = self.dCdt(C2, D2, p2, Heh, Hee, Hhh)
            dD = self.dDdt(C2, D2, p2, Heh, Hee, Hhh)
            C3 = C1 + dC * dt * 2.0
            D3 = D1 + dD * dt * 2.0
            p3 = p1 + dp * dt * 2.0
        self.last_Rsp = Rsp  # emission rates of the step, for tests
        C3, D3, p3 = C3.copy(), D3.copy(), p3.copy()
        if mb.EHs or mb.Phonon:  # _relaxation  PSTD3D/SBEs.py:5288-5337
            ne3 = np.diag(C3).copy()
            nh3 = np.diag(D3).copy()
            WinE, WoutE = np.zeros(Nk), np.zeros(Nk)
            WinH, WoutH = np.zeros(Nk), np.zeros(Nk)
            if mb.EHs:
                self.MBCE(np.real(ne3), np.real(nh3), VC, WinE, WoutE)
            if mb.Phonon:
                self.MBPE(np.real(ne3), VC, E1D, WinE, WoutE)
            if mb.EHs:
                self.MBCH(np.real(ne3), np.real(nh3), VC, WinH, WoutH)
            if mb.Phonon:
                self.MBPH(np.real(nh3), VC, E1D, WinH, WoutH)
            self.last_W = (WinE, WoutE, WinH, WoutH)  # collision rates, for tests
            We = WinE + WoutE
            Wh = WinH + WoutH
            ne3 = np.exp(-We * dt) * ((-1 + np.exp(We * dt) + ne3) * WinE + ne3 * WoutE) / (We + self.small)
            nh3 = np.exp(-Wh * dt) * ((-1 + np.exp(Wh * dt) + nh3) * WinH + nh3 * WoutH) / (Wh + self.small)
            if mb.Recomb:  # PSTD3D/SBEs.py:5335-5337 (nh decays with the updated ne)
                ne3 = ne3 * np.exp(-Rsp * nh3 * dt)
                nh3 = nh3 * np.exp(-Rsp * ne3 * dt)
            C3[np.arange(Nk), np.arange(Nk)] = ne3
            D3[np.arange(Nk), np.arange(Nk)] = nh3
        if mb.DCTrans:  # PSTD3D/SBEs.py:5698-5702: Eac = Re Ex(q = 0) of the wire's field
            Eac = np.real(Ex[self.NQ0])
            for X in (C3, D3, p3):
                self.Transport(X, Eac, dt, mb.LF)
        ne3 = np.abs(np.diag(C3)).astype(complex)
        nh3 = np.abs(np.diag(D3)).astype(complex)
        total = (np.sum(np.abs(ne3)) + np.sum(np.abs(nh3))) / 2.0
        ne3 = ne3 * total / (np.sum(np.abs(ne3)) + self.small)
        nh3 = nh3 * total / (np.sum(np.abs(nh3)) + self.small)
        C3[np.arange(Nk), np.arange(Nk)] = ne3
        D3[np.arange(Nk), np.arange(Nk)] = nh3
        st.p1, st.p2 = (p1 + p3) / 2.0, p3
        st.C1, st.C2 = (C1 + C3) / 2.0, C3
        st.D1, st.D2 = (D1 + D3) / 2.0, D3
        Nr = self.Nr
        if mb.Optics:
            Px, Py, Pz = self.QWPolarization3(p3, w)
        else:
            Px = Py = Pz = np.zeros(Nr, complex)
        if mb.LF:
            Re, Rh = self.QWRho5(C3, D3, ne3, nh3)
            Re = 2.0 * e0 * Re / self.area * self.ehint
            Rh = 2.0 * e0 * Rh / self.area * self.ehint
        else:
            Re = np.zeros(Nr, complex)
            Rh = np.zeros(Nr, complex)
        return Px, Py, Pz, Re, Rh

    # --- public entry  PSTD3D/SBEs.py:5838-5997 ------------------------------
    def QWCalculator(self, Exx, Eyy, Ezz, rr, dt, w):
        """Returns (Pxx, Pyy, Pzz, Rho) on rr; None outputs while the wire is gated off."""
        n = len(rr)
        zero = np.zeros(n, complex)
        if w == 1:
            if self.jjj == self.jmax:
                self.jjj = 0
            self.xxx += 1
            self.jjj += 1
            self.t += dt
        fm = np.max(np.sqrt(np.abs(Exx) ** 2 + np.abs(Eyy) ** 2 + np.abs(Ezz) ** 2))
        if self.wireoff and fm < 1e-3 * self.Emax0:
            return zero, zero.copy(), zero.copy(), zero.copy()
        self.wireoff = False
        Ex, Ey, Ez = self.Prop2QW(rr, Exx, Eyy, Ezz)
        Px, Py, Pz, re, rh = self.sbe_calculator(Ex, Ey, Ez, dt, w)
        Pxx, Pyy, Pzz, RhoE, RhoH = self.QW2Prop(Px, Py, Pz, re, rh, rr)
        Rho = RhoH - RhoE
        self.last_rho = (RhoE, RhoH)  # scale of the difference, for tolerance checks
        if self.mb.ignorewire:
            return zero, zero.copy(), zero.copy(), zero.copy()
        return Pxx, Pyy, Pzz, Rho
