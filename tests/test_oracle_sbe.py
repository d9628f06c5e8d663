"""Pins for the SBE oracle: literal loop transcriptions of the reference at small
Nk (brute force), closed forms, library routines and invariants.

Each ``loop_*`` function below is a second, independent transcription of the
named reference loop (statement order kept), so a transposed operand, a
dropped term or an index shift in the oracle's matrix form fails here.
"""

import numpy as np
import pytest
from scipy.constants import e as e0
from scipy.constants import hbar
from scipy.constants import k as kB
from scipy.special import k0 as sk0

from oracle import sbe as O
from oracle.helpers import FFTG, iFFTG, rescale_1D_dpc, spline_dpc

RNG = np.random.default_rng(42)


def hermitian_pop(n, rng, scale=1.0):
    """Hermitian C with eigenvalues in [0, 1] (occupation-like)."""
    X = rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))
    Q, _ = np.linalg.qr(X)
    ev = rng.uniform(0, 1, n) * scale
    return (Q * ev) @ Q.conj().T


def make_wire(nk=10, mb=None, Nw=1):
    L = O.L_for_nk(nk)
    rr = np.linspace(-2 * L, 2 * L, 40)
    w = O.OracleWire(rr, 0.0, 1e7, 800e-9, Nw, O.QWParams(L=L), mb or O.MBFlags())
    assert w.Nk == nk
    return w, rr


# ---------------------------------------------------------------- loop forms
def loop_calcH(Meh, C, D, p, VC, Nk, Ee, Eh, gap, excitons, lf):
    """PSTD3D/SBEs.py:820-935 literally."""
    Ct, Dt, pt = C.T.copy(), D.T.copy(), p.T.copy()
    V = np.zeros((2 * Nk + 1, 3))
    for q in range(1, Nk):
        for n in range(3):
            V[Nk + q, n] = VC[q, 0, n]
            V[Nk - q, n] = VC[q, 0, n]
    Heh = Meh.astype(complex).copy()
    Hee = np.zeros((Nk, Nk), complex)
    Hhh = np.zeros((Nk, Nk), complex)
    for k in range(Nk):
        Hee[k, k] = Ee[k] + gap
        Hhh[k, k] = Eh[k]
    if excitons and lf:
        for k2 in range(Nk):
            for k1 in range(Nk):
                for q in range(max(-k1, -k2), min(Nk - 1 - k1, Nk - 1 - k2) + 1):
                    if q != 0:
                        Heh[k1, k2] += V[Nk + q, 0] * pt[k1 + q, k2 + q]
        for k2 in range(Nk):
            for k1 in range(Nk):
                for q in range(max(-k1, -k2), min(Nk - 1 - k1, Nk - 1 - k2) + 1):
                    if q != 0:
                        Hee[k1, k2] -= V[Nk + q, 1] * Ct[k1 + q, k2 + q]
                q_d = k1 - k2
                if q_d != 0:
                    for k in range(max(0, -q_d), min(Nk - 1, Nk - 1 - q_d) + 1):
                        Hee[k1, k2] += V[Nk + q_d, 1] * C[k, k + q_d] - V[Nk + q_d, 0] * D[k + q_d, k]
        for k2 in range(Nk):
            for k1 in range(Nk):
                for q in range(max(-k1, -k2), min(Nk - 1 - k1, Nk - 1 - k2) + 1):
                    if q != 0:
                        Hhh[k1, k2] -= V[Nk + q, 2] * Dt[k1 + q, k2 + q]
                q_d = k1 - k2
                if q_d != 0:
                    for k in range(max(0, -q_d), min(Nk - 1, Nk - 1 - q_d) + 1):
                        Hhh[k1, k2] += V[Nk + q_d, 2] * D[k, k + q_d] - V[Nk + q_d, 0] * C[k + q_d, k]
    elif excitons:
        for k2 in range(Nk):
            for k1 in range(Nk):
                for q in range(max(-k1, -k2), min(Nk - 1 - k1, Nk - 1 - k2) + 1):
                    if q != 0:
                        Heh[k1, k2] += V[Nk + q, 0] * pt[k1 + q, k2 + q]
        for k2 in range(Nk):
            for q in range(-k2, Nk - k2):
                if q != 0:
                    Hee[k2, k2] -= V[Nk + q, 1] * Ct[k2 + q, k2 + q]
                    Hhh[k2, k2] -= V[Nk + q, 2] * Dt[k2 + q, k2 + q]
    return Heh, Hee, Hhh


def loop_dpdt(C, D, p, Heh, Hee, Hhh, GamE, GamH, OffP):
    """PSTD3D/SBEs.py:1953-1976"""
    Nk = p.shape[0]
    r = np.zeros((Nk, Nk), complex)
    for ke in range(Nk):
        for kh in range(Nk):
            t1 = np.sum(Hhh[kh, :] * p[:, ke])
            t2 = np.sum(Hee[ke, :] * p[kh, :])
            t3 = np.sum(Heh[:, kh] * C[:, ke])
            t4 = np.sum(Heh[ke, :] * D[:, kh])
            r[kh, ke] = (t1 + t2 - t3 - t4 + Heh[ke, kh] - 1j * hbar * (GamE[ke] + GamH[kh]) * p[kh, ke]
                         + OffP[kh, ke])
    return r / (1j * hbar)


def loop_dCdt(C, D, p, Heh, Hee, Hhh):
    """PSTD3D/SBEs.py:2320-2338"""
    Nk = C.shape[0]
    r = np.zeros((Nk, Nk), complex)
    Hhe = np.conj(Heh.T)
    Peh = np.conj(p.T)
    for k2 in range(Nk):
        for k1 in range(Nk):
            r[k1, k2] = (np.sum(Hee[k2, :] * C[k1, :]) - np.sum(Hee[:, k1] * C[:, k2])
                         + np.sum(Heh[k2, :] * Peh[k1, :]) - np.sum(Hhe[:, k1] * p[:, k2]))
    return r / (1j * hbar)


def loop_dDdt(C, D, p, Heh, Hee, Hhh):
    """PSTD3D/SBEs.py:2618-2636"""
    Nk = D.shape[0]
    r = np.zeros((Nk, Nk), complex)
    Peh = np.conj(p.T)
    Hhe = np.conj(Heh.T)
    for k2 in range(Nk):
        for k1 in range(Nk):
            r[k1, k2] = (np.sum(Hhh[k2, :] * D[k1, :]) - np.sum(Hhh[:, k1] * D[:, k2])
                         + np.sum(Heh[:, k2] * Peh[:, k1]) - np.sum(Hhe[k1, :] * p[k2, :]))
    return r / (1j * hbar)


def loop_tables(ky, Ee, Eh, ge, gh, lorentz=False):
    """MakeK3, MakeUnDel, GaussDelta, CalcMBArrays: PSTD3D/coulomb.py:245-499."""
    N = len(ky)
    k3 = np.zeros((N, N, N), dtype=np.int32)
    for k4 in range(N):
        for k2 in range(N):
            for k1 in range(N):
                v = (k1 + 1) + (k2 + 1) - (k4 + 1)
                k3[k1, k2, k4] = v if 1 <= v <= N else 0
    UnDel = np.ones((N + 1, N + 1))
    UnDel[0, :] = 0.0
    UnDel[:, 0] = 0.0
    for i in range(1, N + 1):
        UnDel[i, i] = 0.0

    def gd(a, b):
        return 0.0 if abs(b) < 1e-300 else 1.0 / (np.sqrt(np.pi) * b) * np.exp(-((a / b) ** 2))

    geh = (ge + gh) / 2.0
    Ceh = np.zeros((N + 1, N + 1, N + 1))
    Cee = np.zeros_like(Ceh)
    Chh = np.zeros_like(Ceh)
    for k1 in range(N):
        for k2 in range(N):
            for k4 in range(N):
                a, b, c = k1 + 1, k2 + 1, k4 + 1
                k30 = k3[k4, k2, k1]
                if k30 > 0:
                    E = Ee[k1] + Eh[k2] - Eh[k30 - 1] - Ee[k4]
                    f = 2.0 * geh / (E**2 + (hbar * geh) ** 2) if lorentz else 2 * np.pi / hbar * gd(E, hbar * geh)
                    Ceh[a, b, c] = UnDel[a, c] * UnDel[b, k30] * f
                k30 = k3[k1, k2, k4]
                if k30 > 0:
                    E = Ee[k1] + Ee[k2] - Ee[k30 - 1] - Ee[k4]
                    f = 2.0 * ge / (E**2 + (hbar * ge) ** 2) if lorentz else 2 * np.pi / hbar * gd(E, hbar * ge)
                    Cee[a, b, c] = UnDel[a, c] * UnDel[b, k30] * f
                    E = Eh[k1] + Eh[k2] - Eh[k30 - 1] - Eh[k4]
                    f = 2.0 * gh / (E**2 + (hbar * gh) ** 2) if lorentz else 2 * np.pi / hbar * gd(E, hbar * gh)
                    Chh[a, b, c] = UnDel[a, c] * UnDel[b, k30] * f
    return k3, UnDel, Ceh, Cee, Chh


def loop_MBCE(ne, nh, Veh2, Vee2, Win, Wout, k3, Ceh, Cee):
    """PSTD3D/coulomb.py:957-1012 (1-based padded ne, nh)."""
    Nk = len(ne) - 1
    for k in range(1, Nk + 1):
        for q1 in range(1, Nk + 1):
            for q2 in range(1, Nk + 1):
                kp, k1 = q1, q2
                v = k3[kp - 1, k1 - 1, k - 1]
                if v > 0:
                    Win[k - 1] += Veh2[k - 1, k1 - 1] * (1 - nh[kp]) * nh[v] * ne[k1] * Ceh[k - 1, kp - 1, k1 - 1]
                v = k3[kp - 1, k - 1, k1 - 1]
                if v > 0:
                    Wout[k - 1] += Veh2[k1 - 1, k - 1] * (1 - ne[k1]) * (1 - nh[kp]) * nh[v] * Ceh[k1 - 1, kp - 1, k - 1]
                k2, k4 = q1, q2
                v = k3[k - 1, k2 - 1, k4 - 1]
                if v > 0:
                    Win[k - 1] += Vee2[k - 1, k4 - 1] * (1 - ne[k2]) * ne[v] * ne[k4] * Cee[k - 1, k2 - 1, k4 - 1]
                v = k3[k4 - 1, k2 - 1, k - 1]
                if v > 0:
                    Wout[k - 1] += Vee2[k4 - 1, k - 1] * (1 - ne[k4]) * (1 - ne[k2]) * ne[v] * Cee[k4 - 1, k2 - 1, k - 1]


def loop_MBCH(ne, nh, Veh2, Vhh2, Win, Wout, k3, Ceh, Chh):
    """PSTD3D/coulomb.py:1016-1071"""
    Nk = len(ne) - 1
    for kp in range(1, Nk + 1):
        for q1 in range(1, Nk + 1):
            for q2 in range(1, Nk + 1):
                k, k1 = q1, q2
                v = k3[kp - 1, k1 - 1, k - 1]
                if v > 0:
                    Win[kp - 1] += Veh2[k - 1, k1 - 1] * (1 - ne[k]) * nh[v] * ne[k1] * Ceh[k - 1, kp - 1, k1 - 1]
                v = k3[kp - 1, k - 1, k1 - 1]
                if v > 0:
                    Wout[kp - 1] += Veh2[k1 - 1, k - 1] * (1 - ne[k]) * (1 - nh[v]) * ne[k1] * Ceh[k - 1, v - 1, k1 - 1]
                k2p, k4p = q1, q2
                v = k3[kp - 1, k2p - 1, k4p - 1]
                if v > 0:
                    Win[kp - 1] += Vhh2[kp - 1, k4p - 1] * (1 - nh[k2p]) * nh[v] * nh[k4p] * Chh[kp - 1, k2p - 1, k4p - 1]
                v = k3[k4p - 1, k2p - 1, kp - 1]
                if v > 0:
                    Wout[kp - 1] += Vhh2[k4p - 1, kp - 1] * (1 - nh[k4p]) * (1 - nh[k2p]) * nh[v] * Chh[k4p - 1, k2p - 1, kp - 1]


def loop_offdiag_E2(ne, nh, Vsq, Ee, Eh, gee, geh):
    """PSTD3D/dephasing.py:227-268"""
    Nk = len(ne)
    lr = lambda a, b: (b / np.pi) / (a**2 + b**2)  # noqa: E731
    D = np.zeros((2 * Nk + 1, Nk))
    for k in range(Nk):
        for p in range(Nk):
            for q in range(max(p - Nk, -k), min(p - 1, Nk - k - 1) + 1):
                if 0 <= k + q < Nk and 0 <= p - q < Nk:
                    E = Ee[k + q] + Ee[p - q] - Ee[p] - Ee[k]
                    D[q + Nk, k] += Vsq[q + Nk, 1] * lr(E, hbar * gee) * (
                        ne[p - q] * (1 - ne[p]) * (1 - ne[k]) + (1 - ne[p - q]) * ne[p] * ne[k])
    for k in range(Nk):
        for p in range(Nk):
            for q in range(max(-p, -k), min(Nk - p - 1, Nk - k - 1) + 1):
                E = Ee[k + q] + Eh[p + q] - Eh[p] - Ee[k]
                D[q + Nk, k] += Vsq[q + Nk, 0] * lr(E, hbar * geh) * (
                    nh[p + q] * (1 - nh[p]) * (1 - ne[k]) + (1 - nh[p + q]) * nh[p] * ne[k])
    return D * np.pi / hbar


def loop_offdiag_H2(ne, nh, Vsq, Ee, Eh, ghh, geh):
    """PSTD3D/dephasing.py:271-317"""
    Nk = len(ne)
    lr = lambda a, b: (b / np.pi) / (a**2 + b**2)  # noqa: E731
    D = np.zeros((2 * Nk + 1, Nk))
    for k in range(Nk):
        for p in range(Nk):
            for q in range(max(p - Nk, -k), min(p - 1, Nk - k - 1) + 1):
                if 0 <= k + q < Nk and 0 <= p - q < Nk:
                    E = Eh[k + q] + Eh[p - q] - Eh[p] - Eh[k]
                    D[q + Nk, k] += Vsq[q + Nk, 2] * lr(E, hbar * ghh) * (
                        nh[p - q] * (1 - nh[p]) * (1 - nh[k]) + (1 - nh[p - q]) * nh[p] * nh[k])
    for k in range(Nk):
        for p in range(Nk):
            for q in range(max(-p, -k), min(Nk - p - 1, Nk - k - 1) + 1):
                E = Eh[k + q] + Ee[p + q] - Ee[p] - Eh[k]
                D[q + Nk, k] += Vsq[q + Nk, 0] * lr(E, hbar * geh) * (
                    ne[p + q] * (1 - ne[p]) * (1 - nh[k]) + (1 - ne[p + q]) * ne[p] * nh[k])
    return D * np.pi / hbar


def loop_offG(Dh, De, p):
    """PSTD3D/dephasing.py:321-356"""
    Nk = p.shape[0]
    pt = p.T
    undel = np.ones(2 * Nk + 1)
    undel[Nk] = 0.0
    x = np.zeros((Nk, Nk), complex)
    for ke in range(Nk):
        for kh in range(Nk):
            for q in range(-kh, Nk - kh):
                x[kh, ke] += Dh[q + Nk, kh] * p[kh + q, ke] * undel[q + Nk]
    for ke in range(Nk):
        for kh in range(Nk):
            for q in range(-ke, Nk - ke):
                x[kh, ke] += De[q + Nk, ke] * pt[ke + q, kh] * undel[q + Nk]
    return x * 1j * hbar


def loop_gamma(w, ne0, nh0, VC):
    """PSTD3D/dephasing.py:29-103 with the tables of :471-538 and Vxx2 :362-378."""
    ky = w.kr
    Nk = len(ky)
    dk = ky[1] - ky[0]
    kmax, kmin = ky[Nk - 1] + dk, ky[0] - dk
    Nk0 = (Nk - 1) // 2 + 1
    me, mh = w.me, w.mh
    t = {n: np.zeros((Nk, Nk), dtype=np.int32) for n in ("kpq", "kmq", "k1mq", "k1pmq", "k1", "k1p")}
    for q in range(Nk):
        for k in range(Nk):
            t["kpq"][k, q] = int(np.round(np.clip(ky[k] + ky[q], kmin, kmax) / dk)) + Nk0 - 1
            t["kmq"][k, q] = int(np.round(np.clip(ky[k] - ky[q], kmin, kmax) / dk)) + Nk0 - 1
            k1p0 = ((me + mh) * ky[q] - 2 * mh * ky[k]) / 2.0 / me
            t["k1p"][k, q] = int(np.round(np.clip(k1p0, kmin, kmax) / dk)) + Nk0 - 1
            t["k1pmq"][k, q] = int(np.round(np.clip(k1p0 - ky[q], kmin, kmax) / dk)) + Nk0 - 1
            k10 = ((me + mh) * ky[q] - 2 * me * ky[k]) / 2.0 / mh
            t["k1"][k, q] = int(np.round(np.clip(k10, kmin, kmax) / dk)) + Nk0 - 1
            t["k1mq"][k, q] = int(np.round(np.clip(k10 - ky[q], kmin, kmax) / dk)) + Nk0 - 1
    xe = me / hbar**2 * np.abs(ky) / (np.abs(ky) + dk * 1e-5) ** 2 / dk
    xh = mh / hbar**2 * np.abs(ky) / (np.abs(ky) + dk * 1e-5) ** 2 / dk
    iq = np.round(np.abs(ky / dk)).astype(np.int32)
    V2 = [np.array([VC[iq[i], 0, n] ** 2 for i in range(Nk)]) for n in range(3)]
    ne = np.zeros(Nk + 2)
    nh = np.zeros(Nk + 2)
    ne[1:Nk + 1] = np.real(ne0)
    nh[1:Nk + 1] = np.real(nh0)
    se = np.zeros(Nk + 2)
    sh = np.zeros(Nk + 2)
    se[1:Nk + 1] = 1 - ne[1:Nk + 1]
    sh[1:Nk + 1] = 1 - nh[1:Nk + 1]
    GE = np.zeros(Nk)
    GH = np.zeros(Nk)
    ok = lambda i: 1 <= i < Nk + 1  # noqa: E731
    for q in range(Nk):
        for k in range(Nk):
            a = t["kpq"][k, q] + 1
            if ok(a):
                GE[k] += np.pi / hbar * V2[1][q] * ne[a] * se[a] * abs(xe[q])
                GH[k] += np.pi / hbar * V2[2][q] * nh[a] * sh[a] * abs(xh[q])
    for q in range(Nk):
        for k in range(Nk):
            a, b, c = t["k1pmq"][k, q] + 1, t["k1p"][k, q] + 1, t["kmq"][k, q] + 1
            if ok(a) and ok(b) and ok(c):
                GE[k] += np.pi / hbar * V2[0][q] * (nh[a] * sh[b] * ne[c] + sh[a] * nh[b] * se[c]) * abs(xh[q])
            a, b = t["k1mq"][k, q] + 1, t["k1"][k, q] + 1
            if ok(a) and ok(b) and ok(c):
                GH[k] += np.pi / hbar * V2[0][q] * (ne[a] * se[b] * nh[c] + se[a] * ne[b] * sh[c]) * abs(xe[q])
    return GE, GH


def loop_pol(w, p, wi):
    """PSTD3D/qwoptics.py:56-82 then x 2 ehint/area/L and FFTG (:933-946)."""
    Nr, Nk = w.Nr, w.Nk
    yw = (-1) ** int(np.floor((wi - 1) / 2.0))
    P = [np.zeros(Nr, complex) for _ in range(3)]
    for r in range(Nr):
        for ke in range(Nk):
            for kh in range(Nk):
                base = p[kh, ke] * w.Expikr[ke, r] * w.Expikrc[kh, r]
                P[0][r] += (base * w.Xvc0[kh, ke]).real * w.QWWindow[r]
                P[1][r] += (base * w.Yvc0[kh, ke]).real * w.QWWindow[r] * yw
                P[2][r] += (base * w.Zvc0[kh, ke]).real * w.QWWindow[r] * (-1) ** wi
    for A in P:
        A *= 2 * w.ehint / w.area / w.L
        FFTG(A)
    return P


# ---------------------------------------------------------------- tests
@pytest.mark.parametrize("excitons,lf", [(True, True), (True, False)])
def test_calcH_matches_loop(excitons, lf):
    mb = O.MBFlags(Excitons=excitons, LF=lf)
    w, _ = make_wire(10, mb)
    Nk = w.Nk
    C = hermitian_pop(Nk, RNG)
    D = hermitian_pop(Nk, RNG)
    p = 0.5 * (RNG.uniform(-1, 1, (Nk, Nk)) + 1j * RNG.uniform(-1, 1, (Nk, Nk)))
    Meh = RNG.standard_normal((Nk, Nk)) * 1e-21 + 0j
    VC, _ = w.CalcScreenedArrays(np.diag(C), np.diag(D))
    ref = loop_calcH(Meh, C, D, p, VC, Nk, w.Ee, w.Eh, w.gap, excitons, lf)
    got = w.CalcH(Meh, C, D, p, VC)
    for a, b in zip(ref, got):
        assert np.allclose(a, b, rtol=1e-12, atol=1e-12 * np.abs(a).max())


def test_rhs_matches_loops():
    w, _ = make_wire(10)
    Nk = w.Nk
    C = hermitian_pop(Nk, RNG)
    D = hermitian_pop(Nk, RNG)
    p = 0.3 * (RNG.standard_normal((Nk, Nk)) + 1j * RNG.standard_normal((Nk, Nk)))
    H = [(RNG.standard_normal((Nk, Nk)) + 1j * RNG.standard_normal((Nk, Nk))) * 1e-20 for _ in range(3)]
    GE, GH = RNG.uniform(1e11, 1e12, Nk), RNG.uniform(1e11, 1e12, Nk)
    Off = (RNG.standard_normal((Nk, Nk)) + 1j * RNG.standard_normal((Nk, Nk))) * 1e-22
    a = loop_dpdt(C, D, p, *H, GE, GH, Off)
    b = w.dpdt(C, D, p, *H, GE, GH, Off)
    assert np.allclose(a, b, rtol=1e-12, atol=1e-12 * np.abs(a).max())
    assert np.allclose(loop_dCdt(C, D, p, *H), w.dCdt(C, D, p, *H), rtol=1e-12, atol=1e-3)
    a = loop_dDdt(C, D, p, *H)
    assert np.allclose(a, w.dDdt(C, D, p, *H), rtol=1e-12, atol=1e-12 * np.abs(a).max())


@pytest.mark.parametrize("lorentz", [False, True])
def test_mbce_mbch_match_table_loops(lorentz):
    """On-the-fly index-shifted tables == literal CalcMBArrays + shifted cores."""
    mb = O.MBFlags(LorentzDelta=lorentz)
    w, _ = make_wire(8, mb)
    Nk = w.Nk
    ne0 = RNG.uniform(0, 0.6, Nk)
    nh0 = RNG.uniform(0, 0.6, Nk)
    VC, _ = w.CalcScreenedArrays(ne0, nh0)
    k3, UnDel, Ceh, Cee, Chh = loop_tables(w.kr, w.Ee, w.Eh, w.gam_e, w.gam_h, lorentz)
    ne = np.zeros(Nk + 1)
    nh = np.zeros(Nk + 1)
    ne[1:] = np.abs(ne0)
    nh[1:] = np.abs(nh0)
    Wi, Wo = np.zeros(Nk), np.zeros(Nk)
    loop_MBCE(ne, nh, VC[:, :, 0] ** 2, VC[:, :, 1] ** 2, Wi, Wo, k3, Ceh, Cee)
    Gi, Go = np.zeros(Nk), np.zeros(Nk)
    w.MBCE(ne0, nh0, VC, Gi, Go)
    assert np.allclose(Gi, Wi, rtol=1e-12, atol=1e-30) and np.allclose(Go, Wo, rtol=1e-12, atol=1e-30)
    assert (Wi >= 0).all() and (Wo >= 0).all()
    Wi, Wo = np.zeros(Nk), np.zeros(Nk)
    loop_MBCH(ne, nh, VC[:, :, 0] ** 2, VC[:, :, 2] ** 2, Wi, Wo, k3, Ceh, Chh)
    Gi, Go = np.zeros(Nk), np.zeros(Nk)
    w.MBCH(ne0, nh0, VC, Gi, Go)
    assert np.allclose(Gi, Wi, rtol=1e-12, atol=1e-30) and np.allclose(Go, Wo, rtol=1e-12, atol=1e-30)
    assert Wi.max() > 0 and Wo.max() > 0


def test_mbce_index_shift_is_real():
    """The shifted read differs from an unshifted one: the quirk is live."""
    w, _ = make_wire(8)
    Nk = w.Nk
    a = w._ceh_read(np.arange(1, Nk)[:, None, None], np.arange(1, Nk)[None, :, None],
                    np.arange(1, Nk)[None, None, :], "ee")
    assert (w._ceh_read(np.array(0), np.array(3), np.array(4), "ee") == 0)
    assert a.max() > 0


def test_gamma_matches_loop():
    w, _ = make_wire(12)
    ne0 = RNG.uniform(0, 0.7, w.Nk) + 0j
    nh0 = RNG.uniform(0, 0.7, w.Nk) + 0j
    VC, _ = w.CalcScreenedArrays(ne0, nh0)
    GE, GH = loop_gamma(w, ne0, nh0, VC)
    assert np.allclose(w.CalcGammaE(ne0, nh0, VC), GE, rtol=1e-12)
    assert np.allclose(w.CalcGammaH(ne0, nh0, VC), GH, rtol=1e-12)
    assert (GE > 0).all() and (GH > 0).all()


def test_offdiag_matches_loops():
    w, _ = make_wire(10)
    Nk = w.Nk
    ne0 = RNG.uniform(0, 0.7, Nk) + 0.01j
    nh0 = RNG.uniform(0, 0.7, Nk) - 0.01j
    p = 0.3 * (RNG.standard_normal((Nk, Nk)) + 1j * RNG.standard_normal((Nk, Nk)))
    VC, _ = w.CalcScreenedArrays(ne0, nh0)
    Vsq = w._vsq(VC)
    De = loop_offdiag_E2(np.abs(ne0), np.abs(nh0), Vsq, w.Ee, w.Eh, w.gam_eh, w.gam_h)
    Dh = loop_offdiag_H2(np.abs(ne0), np.abs(nh0), Vsq, w.Ee, w.Eh, w.gam_e, w.gam_h)
    x, De2, Dh2 = w.OffDiagDephasing2(ne0, nh0, p, VC)
    assert np.allclose(De2, De, rtol=1e-12, atol=1e-30)
    assert np.allclose(Dh2, Dh, rtol=1e-12, atol=1e-30)
    xr = loop_offG(Dh, De, p)
    assert np.allclose(x, xr, rtol=1e-12, atol=1e-12 * np.abs(xr).max())


def test_polarization_matches_loop():
    w, _ = make_wire(8)
    Nk = w.Nk
    p = 0.3 * (RNG.standard_normal((Nk, Nk)) + 1j * RNG.standard_normal((Nk, Nk)))
    for wi in (1, 2, 3):
        ref = loop_pol(w, p, wi)
        got = w.QWPolarization3(p, wi)
        for a, b in zip(ref, got):
            assert np.allclose(a, b, rtol=1e-11, atol=1e-11 * np.abs(a).max())


def test_rho_matches_loop():
    w, _ = make_wire(8)
    Nk, Nr = w.Nk, w.Nr
    C = hermitian_pop(Nk, RNG)
    D = hermitian_pop(Nk, RNG)
    ne, nh = np.diag(C).copy(), np.diag(D).copy()
    re = np.zeros(Nr, complex)
    rh = np.zeros(Nr, complex)
    for ri in range(Nr):  # PSTD3D/qwoptics.py:367-380
        for k2 in range(Nk):
            for k1 in range(Nk):
                f = w.Expikrc[k1, ri] * w.Expikr[k2, ri] * w.QWWindow[ri] / (2.0 * w.L)
                re[ri] += C[k1, k2] * f
                rh[ri] += np.conj(D[k1, k2]) * f
    dr = w.r[1] - w.r[0]
    for A, n in ((re, ne), (rh, nh)):  # :1065-1077
        A -= np.real(np.sum(A[0:10]) + np.sum(A[Nr - 10:Nr])) / 20.0
        A *= (np.sum(np.abs(n)) + 1e-50) / (np.sum(np.abs(A)) * dr + 1e-100)
        FFTG(A)
    a, b = w.QWRho5(C, D, ne, nh)
    assert np.allclose(a, re, rtol=1e-11, atol=1e-11 * np.abs(re).max())
    assert np.allclose(b, rh, rtol=1e-11, atol=1e-11 * np.abs(rh).max())


def test_phonons_closed_forms_and_limits():
    """N0, Vscale closed forms (T/test_phonons.py:146-182); zero/full population."""
    w, _ = make_wire(10)
    assert np.isclose(w.N0, 1.0 / (np.exp(hbar * w.Oph / kB / 77.0) - 1.0), rtol=1e-12)
    assert np.isclose(w.Vscale, hbar * w.Oph * w.epsr * (1 / 8.2 - 1 / 10.0), rtol=1e-12)
    assert np.isclose(O.FermiDistr(0.0), 0.5)
    Nk = w.Nk
    VC, E1D = w.CalcScreenedArrays(np.zeros(Nk), np.zeros(Nk))
    Wi, Wo = np.zeros(Nk), np.zeros(Nk)
    w.MBPE(np.zeros(Nk), VC, E1D, Wi, Wo)
    assert np.all(Wi == 0)
    Wi, Wo = np.zeros(Nk), np.zeros(Nk)
    w.MBPE(np.ones(Nk), VC, E1D, Wi, Wo)
    assert np.all(Wo == 0)
    # literal jit loops  PSTD3D/phonons.py:140-155, 208-220
    ne = RNG.uniform(0, 1, Nk)
    Vep = VC[:, :, 1] / E1D * w.Vscale
    Wi2 = np.array([sum(Vep[k, i] * ne[i] * w.EPT[i, k] for i in range(Nk)) for k in range(Nk)])
    Wo2 = np.array([sum(Vep[k, i] * (1 - ne[i]) * w.EP[i, k] for i in range(Nk)) for k in range(Nk)])
    Wi, Wo = np.zeros(Nk), np.zeros(Nk)
    w.MBPE(ne, VC, E1D, Wi, Wo)
    assert np.allclose(Wi, Wi2, rtol=1e-12) and np.allclose(Wo, Wo2, rtol=1e-12)
    Vhp = VC[:, :, 2] / E1D * w.Vscale
    Wi2 = np.array([sum(Vhp[i, k] * ne[i] * w.HPT[i, k] for i in range(Nk)) for k in range(Nk)])
    Wo2 = np.array([sum(Vhp[i, k] * (1 - ne[i]) * w.HP[i, k] for i in range(Nk)) for k in range(Nk)])
    Wi, Wo = np.zeros(Nk), np.zeros(Nk)
    w.MBPH(ne, VC, E1D, Wi, Wo)
    assert np.allclose(Wi, Wi2, rtol=1e-12) and np.allclose(Wo, Wo2, rtol=1e-12)


def test_grid_closed_forms():
    w, _ = make_wire(16)
    Nk = w.Nk
    assert w.Nr == 2 * Nk + 2 and w.NQ0 == Nk + 1
    k = np.arange(Nk)
    assert np.allclose(w.kr, w.dkr * (k - Nk / 2 + 0.5), rtol=1e-14, atol=0)
    assert np.array_equal(w.kkp, k[:, None] - k[None, :] + Nk + 1)
    assert np.allclose(w.Qr, (np.arange(w.Nr) - w.Nr // 2) * w.dkr, rtol=1e-14, atol=1e-6)
    # Toeplitz V and the symmetric q = +-d values of Vint
    assert np.allclose(w.Veh0, w.Veh0.T) and np.allclose(w.Veh0[1:, 1:], w.Veh0[:-1, :-1])


def test_k03_pins():
    """Accurate K03 == scipy k0; _K03_jit: cutoffs, x->0 leading term, asymptote."""
    x = np.logspace(-3, 1.9, 50)
    assert np.allclose(O.K03(x), sk0(x), rtol=1e-14)
    assert O.K03(np.array([150.0]))[0] == 0.0
    j = O.K03_jit(np.array([1e-12, 150.0, 3.0, 50.0]))
    assert j[0] == 23.0 and j[1] == 0.0
    assert np.isclose(j[2], np.sqrt(np.pi / 6.0) * np.exp(-3.0))
    xs = np.array([1e-6])
    assert np.isclose(O.K03_jit(xs)[0], -np.log(xs[0] / 2) - 0.5772156649015329, rtol=1e-9)
    # series branch: independent summation of sum_k (x^2/4)^k/(k!)^2 (1/k - gamma)
    for xv in (0.3, 1.0, 1.9):
        s = -np.log(0.5 * xv) - 0.5772156649015329
        t = 1.0
        for kk in range(1, 20):
            t *= 0.25 * xv * xv / kk**2
            s += t * (1.0 / kk - 0.5772156649015329)
            if abs(t) < 1e-15:
                break
        assert np.isclose(O.K03_jit(np.array([xv]))[0], s, rtol=1e-14)


def test_vint_literal_double_loop():
    w, _ = make_wire(8)
    y = w.r
    Ny = len(y)
    N1, N2 = Ny // 4, 3 * Ny // 4
    for Q in (0.0, 3 * w.dkr):
        kmin = (w.alphae + w.alphah) / 4.0
        dk = max(abs(Q), kmin)
        mc = w.alphae * w.alphah / np.pi * (y[1] - y[0]) ** 2
        s = 0.0
        for i in range(N1, N2 + 1):
            for j in range(N1, N2 + 1):
                r = np.sqrt((y[i] - y[j]) ** 2 + w.Delta0**2)
                s += np.exp(-(w.alphae * y[i]) ** 2 - (w.alphah * y[j]) ** 2) * mc * O.K03_jit(np.array([dk * r]))[0]
        assert np.isclose(O.Vint(Q, y, w.alphae, w.alphah, w.Delta0), s, rtol=1e-12)


def test_fftg_is_centred_dft():
    N = 14
    F = RNG.standard_normal(N) + 1j * RNG.standard_normal(N)
    G = F.copy()
    FFTG(G)
    n = np.arange(N)
    s = (-1.0) ** n
    ref = -s * (np.exp(-2j * np.pi * np.outer(n, n) / N) @ (s * F)) / N
    assert np.allclose(G, ref, rtol=1e-13, atol=1e-14)
    iFFTG(G)
    assert np.allclose(G, F, rtol=1e-13, atol=1e-14)


def test_spline_reproduces_cubics_and_dense_solve():
    x = np.sort(RNG.uniform(-1, 1, 17))
    y = (0.3 - 1.2 * x + 0.7 * x**2 - 2.1 * x**3) * (1 + 0.5j)
    u = np.linspace(x[0], x[-1], 101)
    got = rescale_1D_dpc(x, y, u)
    assert np.allclose(got, (0.3 - 1.2 * u + 0.7 * u**2 - 2.1 * u**3) * (1 + 0.5j), rtol=1e-10, atol=1e-11)
    assert np.all(rescale_1D_dpc(x, y, np.array([x[0] - 0.1, x[-1] + 0.1])) == 0)
    # interpolation, C2 continuity and the FMM end conditions fix the spline uniquely
    y2 = np.sin(3 * x) + 1j * np.cos(2 * x)
    b, c, d = spline_dpc(x, y2)
    n = len(x)
    h = np.diff(x)
    s = lambda i, t: y2[i] + t * (b[i] + t * (c[i] + t * d[i]))  # noqa: E731
    for i in range(n - 1):
        assert np.isclose(s(i, h[i]), y2[i + 1], rtol=1e-10, atol=1e-12)
        assert np.isclose(b[i] + 2 * c[i] * h[i] + 3 * d[i] * h[i] ** 2, b[i + 1], rtol=1e-9, atol=1e-9)
        assert np.isclose(c[i] + 3 * d[i] * h[i], c[i + 1], rtol=1e-9, atol=1e-8)

    def dd3(xs, ys):
        f1 = np.diff(ys) / np.diff(xs)
        f2 = np.diff(f1) / (xs[2:] - xs[:-2])
        return (f2[1] - f2[0]) / (xs[3] - xs[0])
    assert np.isclose(d[0], dd3(x[:4], y2[:4]), rtol=1e-8)
    assert np.isclose(d[n - 2], dd3(x[-4:], y2[-4:]), rtol=1e-8)


def test_step_invariants_and_hermiticity():
    w, rr = make_wire(12)
    E = np.full(len(rr), 3e7, complex)
    for _ in range(6):
        out = w.QWCalculator(E, 0.5 * E, 0.1 * E, rr, 2e-17, 1)
    st = w.state[0]
    for M in (st.C2, st.D2, st.C1, st.D1):
        assert np.abs(M - M.conj().T).max() <= 1e-12 * np.abs(M).max()
    assert np.isclose(np.abs(np.diag(st.C2)).sum(), np.abs(np.diag(st.D2)).sum(), rtol=1e-12)
    assert all(np.isfinite(o).all() for o in out)
    assert np.abs(out[0]).max() > 0


def test_no_field_is_stationary():
    """p = 0 and no field: p stays 0 (the SBEs have no source)."""
    w, rr = make_wire(10)
    w.wireoff = False
    Z = np.zeros(len(rr), complex)
    for _ in range(3):
        w.QWCalculator(Z, Z, Z, rr, 1e-17, 1)
    assert np.abs(w.state[0].p2).max() == 0.0


def _obe_rel_err(dt, n, k):
    mb = O.MBFlags(OBE=True, DiagDph=False, OffDiagDph=False, Screened=False)
    w, rr = make_wire(8, mb)
    E = np.full(len(rr), 1e6, complex)
    for _ in range(n):
        w.QWCalculator(E, 0 * E, 0 * E, rr, dt, 1)
    Ex, Ey, Ez = w.Prop2QW(rr, E, 0 * E, 0 * E)
    Meh = w.calc_Meh(Ex, Ey, Ez, 1)
    # exact linear response  p(T) = M/(h Om) (exp(-i Om T) - 1); after n steps the
    # newest p sits at T = (n+1) dt
    T = (n + 1) * dt
    Om = (w.Ee[k] + w.gap + w.Eh[k]) / hbar - 1j * (w.gam_e + w.gam_h)
    pe = Meh[k, k] / (hbar * Om) * (np.exp(-1j * Om * T) - 1.0)
    return abs(w.state[0].p2[k, k] - pe) / abs(pe)


def test_obe_mode_is_two_level_bloch():
    """OBE mode (PSTD3D/SBEs.py:4884-4891) at weak field: every k is a damped
    two-level system with a closed-form solution.  The leapfrog+averaging step
    (PSTD3D/SBEs.py:5685-5718) converges to it at first order in dt."""
    for k in (0, 4, 7):
        e1 = _obe_rel_err(1e-17, 40, k)
        e2 = _obe_rel_err(0.5e-17, 80, k)
        assert e1 < 0.03
        assert 1.7 < e1 / e2 < 2.3


@pytest.mark.parametrize("lorentz", [False, True])
def test_collision_regrouping_matches_mbce_mbch(lorentz):
    """The device evaluates each carrier-carrier line shape once per (species, kind)
    and spreads it over Win[o] and Wout[t] (o = output momentum, t = partner):
    S(o, t) = sum_j C[o,j,t] (1-n[j]) n[j+t-o]  etc.  Check that regrouping against
    the loop-form MBCE/MBCH (PSTD3D/coulomb.py:957-1071) on random populations."""
    w, _ = make_wire(nk=12, mb=O.MBFlags(LorentzDelta=lorentz))
    Nk = w.Nk
    rng = np.random.default_rng(3)
    ne0, nh0 = rng.uniform(0, 1, Nk), rng.uniform(0, 1, Nk)
    v = rng.uniform(0.5, 2, (3, Nk))
    idx = np.abs(np.arange(Nk)[:, None] - np.arange(Nk)[None, :])
    VC = np.stack([v[c][idx] for c in range(3)], axis=2)
    Wi, Wo, Hi, Ho = (np.zeros(Nk) for _ in range(4))
    w.MBCE(ne0, nh0, VC, Wi, Wo)
    w.MBCH(ne0, nh0, VC, Hi, Ho)
    ne, nh = np.r_[0, ne0], np.r_[0, nh0]

    def C(kind, a, b, c):  # read C_kind[a-1, b-1, c-1] for 1-based a, b, c
        return float(w._ceh_read(np.array(a - 1), np.array(b - 1), np.array(c - 1), kind))

    R = np.zeros((4, Nk))
    for o in range(1, Nk + 1):
        for t in range(1, Nk + 1):
            Sx = Ss = Hx = Hs = 0.0
            for j in range(1, Nk + 1):
                if 1 <= j + t - o <= Nk:
                    Sx += C("eh", o, j, t) * (1 - nh[j]) * nh[j + t - o]
                    Hx += C("eh", j, o, j + t - o) * (1 - ne[j]) * ne[j + t - o]
                if 1 <= o + j - t <= Nk:
                    Ss += C("ee", o, j, t) * (1 - ne[j]) * ne[o + j - t]
                    Hs += C("hh", o, j, t) * (1 - nh[j]) * nh[o + j - t]
            q = abs(o - t)
            S = v[0][q] ** 2 * Sx + v[1][q] ** 2 * Ss
            H = v[0][q] ** 2 * Hx + v[2][q] ** 2 * Hs
            R[0, o - 1] += ne[t] * S
            R[1, t - 1] += (1 - ne[o]) * S
            R[2, o - 1] += nh[t] * H
            R[3, t - 1] += (1 - nh[o]) * H
    for a, b in zip(R, (Wi, Wo, Hi, Ho)):
        assert np.abs(a - b).max() <= 1e-14 * np.abs(b).max()
