"""Host-side set-up of one quantum wire: parameters, grids, constants and the
init tables the device step reads.

Mirrors SBESolver.__init__ (PSTD3D/SBEs.py:4635-4984) and the table builders it
calls, laid out for the GPU: every Coulomb/screening quantity is a Toeplitz row
indexed by |k1-k2| (the reference's Nk x Nk matrices are Toeplitz by
construction, PSTD3D/coulomb.py:302-333, 502-577), the carrier-carrier
collision tables (Nk+1)^3 are not stored (the device regenerates them), and
V(q) is integrated on the GPU.
"""

from dataclasses import dataclass, field

import numpy as np
from scipy.constants import e as e0
from scipy.constants import epsilon_0 as eps0
from scipy.constants import hbar
from scipy.constants import k as kB
from scipy.constants import m_e as me0
from scipy.special import k0 as _k0

pi = np.pi
twopi = 2.0 * np.pi
TEMP = 77.0          # PSTD3D/phonons.py:22
EPSR0_HOST = 10.0    # PSTD3D/phonons.py:23
EPSRINF_HOST = 8.2   # PSTD3D/phonons.py:24


@dataclass
class QWParams:
    """qw.params, in file order (PSTD3D/SBEs.py:4990-5014)."""
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


MB_NAMES = ("Optics", "Excitons", "EHs", "Screened", "Phonon", "DCTrans", "LF", "FreePot", "DiagDph",
            "OffDiagDph", "Recomb", "PLSpec", "ignorewire", "Xqwparams", "LorentzDelta")


@dataclass
class MBFlags:
    """mb.params, in file order (PSTD3D/SBEs.py:5016-5034), plus OBE (:4884-4891)."""
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


def _first_token(line):
    return line.split()[0]


def _num(tok):
    """A number token; Fortran exponents (1d-9, 2D3) read as e (PSTD3D/typetime.py:31-50)."""
    return float(tok.replace("d", "e").replace("D", "E"))


def _flag(tok):
    """0 / 1 as the reference reads them; Fortran .true. / .false. accepted too."""
    t = tok.strip().lower()
    if t in (".true.", "t", "true"):
        return True
    if t in (".false.", "f", "false"):
        return False
    return bool(int(t))


def read_qw_params(path="params/qw.params"):
    """PSTD3D/SBEs.py:4990-5014: the first token of each of 15 lines (values kept in the
    file's units)."""
    with open(path, "r", encoding="utf-8") as f:
        vals = [_first_token(f.readline()) for _ in range(15)]
    p = QWParams(*[_num(v) for v in vals[:13]], int(_num(vals[13])), int(_num(vals[14])))
    return p


def read_mb_params(path="params/mb.params"):
    """PSTD3D/SBEs.py:5016-5034: the first token of each of 15 lines."""
    with open(path, "r", encoding="utf-8") as f:
        vals = [_flag(_first_token(f.readline())) for _ in range(15)]
    return MBFlags(*vals)


def nk_of_L(L, gap_eV=1.5, me_rel=0.07):
    """Nk for a wire length (PSTD3D/SBEs.py:4766-4770)."""
    gap = gap_eV * e0
    me = me_rel * me0
    kmax = np.sqrt(1.2 * gap * 2.0 * me / (hbar**2))
    dkr = twopi / (2.0 * L)
    return int(np.floor(kmax / dkr) * 2 + 1) - 1


def L_of_nk(nk, gap_eV=1.5, me_rel=0.07):
    """Midpoint of the L interval that yields an even Nk."""
    gap = gap_eV * e0
    me = me_rel * me0
    kmax = np.sqrt(1.2 * gap * 2.0 * me / hbar**2)
    m = nk // 2
    return 0.5 * (m * pi / kmax + (m + 1) * pi / kmax)


def _k03(x):
    """Accurate K0 with the reference's cut-off (PSTD3D/usefulsubs.py:1794-1820)."""
    x = np.asarray(x, dtype=np.float64)
    return np.where(x > 1e2, 0.0, _k0(np.minimum(x, 1e2)))


def fermi(En):
    """PSTD3D/phonons.py:317-340"""
    return 1.0 / (np.exp(En / kB / TEMP) + 1.0)


def spline_factors(x):
    """y-independent part of the FMM spline (libpulsesuite/spliner.py:1287-1363):
    spacings h, elimination factors t, pivots b, end-condition constants."""
    x = np.asarray(x, dtype=np.float64)
    n = len(x)
    if n < 4:
        raise ValueError("spline grids need at least 4 knots")
    h = np.zeros(n)
    h[: n - 1] = np.diff(x)
    b = np.zeros(n)
    b[0] = -h[0]
    for i in range(1, n - 1):
        b[i] = 2.0 * (h[i - 1] + h[i])
    b[n - 1] = -h[n - 2]
    t = np.zeros(n)
    for i in range(1, n):
        t[i] = h[i - 1] / b[i - 1]
        b[i] = b[i] - t[i] * h[i - 1]
    end = np.array([x[3] - x[1], x[2] - x[0], h[0] ** 2, x[3] - x[0],
                    x[n - 1] - x[n - 3], x[n - 2] - x[n - 4], h[n - 2] ** 2, x[n - 1] - x[n - 4]])
    return h, t, b, end


def spline_eval_index(x, u):
    """Interval index (locate, spliner.py:1137-1170) and offset of targets u; -1 outside."""
    x = np.asarray(x, dtype=np.float64)
    u = np.asarray(u, dtype=np.float64)
    n = len(x)
    j = np.searchsorted(x, u, side="right") - 1
    j = np.where(u < x[0], 0, j)
    j = np.where(u >= x[n - 1], n - 1, j)
    inside = (u >= np.min(x)) & (u <= np.max(x))
    dx = u - x[np.clip(j, 0, n - 1)]
    return np.where(inside, j, -1).astype(np.int32), np.where(inside, dx, 0.0)


@dataclass
class WireModel:
    """Everything the device step needs for one SBESolver (shared by its Nw wires)."""
    qw: QWParams
    mb: MBFlags
    Nk: int = 0
    Nr: int = 0
    arrays: dict = field(default_factory=dict)
    scalars: dict = field(default_factory=dict)

    @property
    def flags(self):
        from . import _lib as L
        m = self.mb
        f = 0
        for name, bit in (("Optics", L.OPTICS), ("Excitons", L.EXCITONS), ("EHs", L.EHS), ("Screened", L.SCREENED),
                          ("Phonon", L.PHONON), ("DCTrans", L.DCTRANS), ("LF", L.LF), ("DiagDph", L.DIAGDPH),
                          ("OffDiagDph", L.OFFDIAGDPH), ("Recomb", L.RECOMB), ("ignorewire", L.IGNOREWIRE),
                          ("LorentzDelta", L.LORENTZDELTA)):
            if getattr(m, name):
                f |= bit
        return f


_MODEL_CACHE = {}


def build_model(qw, mb, rr, r0=0.0):
    """Grids, constants and host tables (PSTD3D/SBEs.py:4756-4984).

    Models are cached per (material, flags, propagation grid, r0): wires of the same
    kind (batches, serving admission) share one model and its device tables."""
    rr = np.asarray(rr, dtype=np.float64)
    key = (tuple(sorted(vars(qw).items())), tuple(sorted(vars(mb).items())), rr.shape, rr.tobytes(), float(r0))
    m = _MODEL_CACHE.get(key)
    if m is None:
        m = _build_model(qw, mb, rr, r0)
        if len(_MODEL_CACHE) > 4096:
            _MODEL_CACHE.clear()
        _MODEL_CACHE[key] = m
    return m


def _build_model(qw, mb, rr, r0=0.0):
    M = WireModel(qw=qw, mb=MBFlags(**vars(mb)))
    mb = M.mb
    # FreePot adds the monopole couplings Wee, Whh to Hee, Hhh (PSTD3D/SBEs.py:848-853), which the
    # reference's _preparation leaves at zero (their CalcWnn is not called, :5207-5209): no term.
    L = qw.L
    gap = qw.gap_eV * e0
    me, mh = qw.me_rel * me0, qw.mh_rel * me0
    HO = qw.HO_eV * e0
    Oph = qw.Oph_eV * e0 / hbar
    Gph = qw.Gph_eV * e0 / hbar
    kmax = np.sqrt(1.2 * gap * 2.0 * me / (hbar**2))
    dkr = twopi / (2.0 * L)
    nk1 = int(np.floor(kmax / dkr) * 2 + 1)
    Nr = nk1 * 2
    Nk = nk1 - 1
    M.Nk, M.Nr = Nk, Nr
    dnk = (Nk - 1) // 2
    kr = np.array([dkr * (-dnk + k - 0.5) for k in range(Nk)])
    r = np.linspace(-L, L, Nr)
    dcv = np.sqrt((e0 * hbar) ** 2 / (6.0 * me0 * gap) * (me0 / me - 1.0))
    ae = np.sqrt(me * HO) / hbar
    ah = np.sqrt(mh * HO) / hbar
    ehint = np.sqrt(2.0 * ae * ah / (ae**2 + ah**2))
    gam_e, gam_h = qw.gam_e, qw.gam_h
    gam_eh = (gam_e + gam_h) / 2.0
    area = np.sqrt(2.0 * pi) / np.sqrt(ae**2 + ah**2) * qw.Delta0
    Ee = hbar**2 * kr**2 / (2.0 * me)
    Eh = hbar**2 * kr**2 / (2.0 * mh)
    if mb.OBE:
        mb.Optics, mb.Excitons, mb.EHs, mb.Phonon, mb.Recomb, mb.LF, mb.DCTrans = (
            True, False, False, False, False, False, False)

    # Toeplitz rows (|k1 - k2| = d)
    d = np.arange(Nk)
    dky = np.abs(kr[d] - kr[0])
    qe = np.maximum(dky, ae / 2.0)
    qh = np.maximum(dky, ah / 2.0)
    Re = np.sqrt((2.0 / ae) ** 2 + qw.Delta0**2)
    Rh = np.sqrt((2.0 / ah) ** 2 + qw.Delta0**2)
    scale = e0**2 / (twopi * eps0 * qw.epsr * hbar**2)
    chie = me * _k03(qe * Re) / qe * scale
    chih = mh * _k03(qh * Rh) / qh * scale

    # dephasing index tables  PSTD3D/dephasing.py:471-538
    dk = kr[1] - kr[0]
    kmx = kr[Nk - 1] + dk
    kmn = kr[0] - dk
    Nk0 = (Nk - 1) // 2 + 1
    K = kr[:, None]
    Q = kr[None, :]

    def idx(v):
        return (np.round(np.clip(v, kmn, kmx) / dk)).astype(np.int64) + Nk0 - 1

    k1p0 = ((me + mh) * Q - 2 * mh * K) / 2.0 / me
    k10 = ((me + mh) * Q - 2 * me * K) / 2.0 / mh
    ktab = np.stack([idx(K + Q), idx(K - Q), idx(k1p0), idx(k1p0 - Q), idx(k10), idx(k10 - Q)]).astype(np.int32)
    xe = me / hbar**2 * np.abs(kr) / (np.abs(kr) + dk * 1e-5) ** 2 / dk
    xh = mh / hbar**2 * np.abs(kr) / (np.abs(kr) + dk * 1e-5) ** 2 / dk
    iqv = np.round(np.abs(kr / dk)).astype(np.int32)

    # optics  PSTD3D/qwoptics.py:636-656
    win = np.exp(-((r / (L / 2.0)) ** 150))
    m = np.arange(Nr)
    twid = np.exp(-2j * np.pi * m / Nr)

    # splines rr -> r (Prop2QW) and r -> rr (QW2Prop)
    rr = np.asarray(rr, dtype=np.float64)
    sa_h, sa_t, sa_b, sa_end = spline_factors(rr)
    sa_idx, sa_dx = spline_eval_index(rr, r)
    sb_h, sb_t, sb_b, sb_end = spline_factors(r)
    sb_idx, sb_dx = spline_eval_index(r, rr)

    N0 = 1.0 / (np.exp(hbar * Oph / kB / TEMP) - 1.0)
    Vscale = hbar * Oph * qw.epsr * (1.0 / EPSRINF_HOST - 1.0 / EPSR0_HOST)
    fe = fermi(Ee + gap / 2.0)
    fh = fermi(Ee + gap / 2.0)  # the reference seeds holes with Ee too (PSTD3D/SBEs.py:4840-4841)

    M.arrays = dict(kr=kr, Ee=Ee, Eh=Eh, win=win, r=r, chie=chie, chih=chih, qe=qe, qh=qh, ktab=ktab,
                    xe=xe, xh=xh, iqv=iqv, twid=twid, sa_x=rr, sa_t=sa_t, sa_b=sa_b, sa_h=sa_h, sa_end=sa_end,
                    sa_idx=sa_idx, sa_dx=sa_dx, sb_t=sb_t, sb_b=sb_b, sb_h=sb_h, sb_end=sb_end, sb_idx=sb_idx,
                    sb_dx=sb_dx, fe0=fe, fh0=fh, Vq=d * dkr)
    M.scalars = dict(L=L, gap=gap, me=me, mh=mh, ehint=ehint, area=area, dcv=dcv, dkr=dkr, ae=ae, ah=ah,
                     dr_r=r[1] - r[0], drr=rr[1] - rr[0], gam_e=gam_e, gam_h=gam_h, gam_eh=gam_eh,
                     Vscale=Vscale, N0=N0, hOph=hbar * Oph, hGph=hbar * Gph, Gph=Gph,
                     density_max=min(ae / 2.0, ah / 2.0) / pi * 0.99, Delta0=qw.Delta0, epsr=qw.epsr,
                     vpref=e0**2 / (twopi * eps0 * qw.epsr * L), Oph=Oph,
                     dr=(r[2] - r[1]) * (Nr - 1) / float(Nr), qc=2.0 * ae * ah / (ae + ah),
                     Edc=float(qw.Edc), rsp=rsp_poly(gam_eh, dcv, ehint, qw.epsr) if mb.Recomb else (0.0,) * 4)
    return M


def rsp_poly(gam_eh, dcv, ehint, epsr):
    """Coefficients a of SpontIntegral(E) = a0 + a1 E + a2 E^2 + a3 E^3 (PSTD3D/emission.py:178-227):
    the sum over the photon-energy grid h_n = (n + 1/2) dhw of CalcHOmega (:230-260) of
    (h_n + E)^3 / (c^3 pi^2 hbar^3) square_n dhw, square of InitializeEmission (:36-91),
    expanded in E (its moments; every term is positive, so the cubic and the direct sum agree
    to rounding)."""
    from scipy.constants import c as c0

    kBT, hg = kB * TEMP, hbar * gam_eh
    hwmax = (kBT + hg) * 4.0
    dhw = min(kBT, hg) / 20.0
    Nw = int(np.ceil(hwmax / dhw))
    if Nw < 10:
        raise ValueError("Error: temperature is too low in emission.py")
    H = (np.arange(Nw) + 0.5) * dhw
    dcv = abs(dcv)
    square = (3.0 * dcv**2 * ehint / eps0 / np.sqrt(epsr) * ehint / hbar * ((hg / pi) / (H**2 + hg**2))
              * np.exp(-H / kB / TEMP))
    w = square * dhw / (c0**3 * pi**2 * hbar**3)
    return (float(np.sum(H**3 * w)), float(3.0 * np.sum(H**2 * w)), float(3.0 * np.sum(H * w)), float(np.sum(w)))


def explicit_model(ky, Ee, Eh, mb, *, y=None, L=None, Delta0=5e-9, me=0.07 * me0, mh=0.45 * me0, ge=1e12, gh=1e12,
                   alphae=None, alphah=None, er=9.1, Qy=None, kkp=None, Oph=0.0, Gph=0.0, epsr_ph=None):
    """A WireModel built from explicitly given grids and constants, as the reference's module
    initialisers receive them (InitializeCoulomb PSTD3D/coulomb.py:1351-1374,
    InitializePhonons PSTD3D/phonons.py:55-136), for the kernel-level entries those modules'
    functions run on.  Quantities those functions do not use (the optical tables, the
    propagation-grid splines) hold neutral placeholders of the right sizes.

    y: the wire's real-space grid (the V(q) integral runs over it); Qy, kkp: the momentum
    difference grid and the index map of CalcCoulombArrays (PSTD3D/coulomb.py:502-577),
    V(|Qy[kkp[k1, k2]]|) depends on k1 - k2 only."""
    kr = np.asarray(ky, dtype=np.float64)
    Ee = np.asarray(Ee, dtype=np.float64)
    Eh = np.asarray(Eh, dtype=np.float64)
    Nk = len(kr)
    M = WireModel(qw=QWParams(), mb=MBFlags(**vars(mb)))
    dk = kr[1] - kr[0]
    L = float(L) if L is not None else twopi / (2.0 * dk)
    y = np.asarray(y, dtype=np.float64) if y is not None else np.linspace(-L, L, 2 * Nk + 2)
    Nr = len(y)
    M.Nk, M.Nr = Nk, Nr
    ae = alphae if alphae is not None else 1.0 / Delta0
    ah = alphah if alphah is not None else 1.0 / Delta0
    d = np.arange(Nk)
    dky = np.abs(kr[d] - kr[0])
    qe = np.maximum(dky, ae / 2.0)
    qh = np.maximum(dky, ah / 2.0)
    Re = np.sqrt((2.0 / ae) ** 2 + Delta0**2)
    Rh = np.sqrt((2.0 / ah) ** 2 + Delta0**2)
    scale = e0**2 / (twopi * eps0 * er * hbar**2)
    chie = me * _k03(qe * Re) / qe * scale
    chih = mh * _k03(qh * Rh) / qh * scale
    if Qy is not None and kkp is not None:
        Vq = np.abs(np.asarray(Qy, dtype=np.float64)[np.asarray(kkp)[d, 0]])
    else:
        Vq = d * dk
    kmx, kmn, Nk0 = kr[Nk - 1] + dk, kr[0] - dk, (Nk - 1) // 2 + 1
    K, Q = kr[:, None], kr[None, :]

    def idx(v):
        return (np.round(np.clip(v, kmn, kmx) / dk)).astype(np.int64) + Nk0 - 1

    k1p0 = ((me + mh) * Q - 2 * mh * K) / 2.0 / me
    k10 = ((me + mh) * Q - 2 * me * K) / 2.0 / mh
    ktab = np.stack([idx(K + Q), idx(K - Q), idx(k1p0), idx(k1p0 - Q), idx(k10), idx(k10 - Q)]).astype(np.int32)
    xe = me / hbar**2 * np.abs(kr) / (np.abs(kr) + dk * 1e-5) ** 2 / dk
    xh = mh / hbar**2 * np.abs(kr) / (np.abs(kr) + dk * 1e-5) ** 2 / dk
    iqv = np.round(np.abs(kr / dk)).astype(np.int32)
    rr = np.linspace(y[0], y[-1], 8)
    sa_h, sa_t, sa_b, sa_end = spline_factors(rr)
    sa_idx, sa_dx = spline_eval_index(rr, y)
    sb_h, sb_t, sb_b, sb_end = spline_factors(y)
    sb_idx, sb_dx = spline_eval_index(y, rr)
    epsr_ph = er if epsr_ph is None else epsr_ph
    N0 = 1.0 / (np.exp(hbar * Oph / kB / TEMP) - 1.0) if Oph > 0 else 0.0
    Vscale = hbar * Oph * epsr_ph * (1.0 / EPSRINF_HOST - 1.0 / EPSR0_HOST)
    z = np.zeros(Nk)
    M.arrays = dict(kr=kr, Ee=Ee, Eh=Eh, win=np.ones(Nr), r=y, chie=chie, chih=chih, qe=qe, qh=qh, ktab=ktab, xe=xe,
                    xh=xh, iqv=iqv, twid=np.exp(-2j * np.pi * np.arange(Nr) / Nr), sa_x=rr, sa_t=sa_t, sa_b=sa_b,
                    sa_h=sa_h, sa_end=sa_end, sa_idx=sa_idx, sa_dx=sa_dx, sb_t=sb_t, sb_b=sb_b, sb_h=sb_h,
                    sb_end=sb_end, sb_idx=sb_idx, sb_dx=sb_dx, fe0=z, fh0=z, Vq=Vq)
    M.scalars = dict(L=L, gap=0.0, me=me, mh=mh, ehint=1.0, area=1.0, dcv=0.0, dkr=dk, ae=ae, ah=ah,
                     dr_r=y[1] - y[0], drr=rr[1] - rr[0], gam_e=ge, gam_h=gh, gam_eh=(ge + gh) / 2.0,
                     Vscale=Vscale, N0=N0, hOph=hbar * Oph, hGph=hbar * Gph, Gph=Gph,
                     density_max=min(ae / 2.0, ah / 2.0) / pi * 0.99, Delta0=Delta0, epsr=er,
                     vpref=e0**2 / (twopi * eps0 * er * L), Oph=Oph, dr=(y[2] - y[1]) * (Nr - 1) / float(Nr),
                     qc=2.0 * ae * ah / (ae + ah))
    return M
