"""Closed-form pins of the SBE oracle: answers the mathematics of the reference's loops fixes
for special inputs, derived from the reference's text (cited per test), not from the
oracle's matrix forms.

Each input is chosen so that a plausible transcription error changes the answer: a single
nonzero matrix entry moves to a different place when an operand is transposed or an index
shifted; electron and hole energies differ (me != mh), so swapping kh and ke changes a
frequency; Pauli factors vanish at zero and full occupation one term at a time.
"""

import numpy as np
import pytest
from scipy.constants import hbar

from oracle import sbe as O
from oracle.helpers import FFTG, iFFTG


def make_wire(nk=12, mb=None):
    L = O.L_for_nk(nk)
    rr = np.linspace(-2 * L, 2 * L, 48)
    w = O.OracleWire(rr, 0.0, 1e7, 800e-9, 1, O.QWParams(L=L), mb or O.MBFlags())
    assert w.Nk == nk
    return w, rr


def toeplitz_vc(nk, rng):
    """VC[k1, k2, n] = row_n[|k1 - k2|] with random positive rows (the reference's V is
    Toeplitz in k1 - k2, PSTD3D/coulomb.py:502-577)."""
    rows = rng.uniform(0.5, 2.0, (3, nk)) * 1e-20
    d = np.abs(np.arange(nk)[:, None] - np.arange(nk)[None, :])
    return np.stack([rows[n][d] for n in range(3)], axis=2), rows


FREE = O.MBFlags(Excitons=False, EHs=False, Screened=False, Phonon=False, LF=False, DiagDph=False,
                 OffDiagDph=False)


# ------------------------------------------------------------------ CalcH
@pytest.mark.parametrize("a,b", [(2, 7), (9, 1), (5, 5)])
def test_calcH_single_entry_exchange_and_direct_terms(a, b):
    """_CalcH_jit (PSTD3D/SBEs.py:820-935) for one nonzero entry.

    p[a, b] = z: Heh[k1, k2] += V_eh(|q|) p^T[k1 + q, k2 + q] puts V_eh(|q|) z at
    (b - q, a - q) for every q != 0 that stays on the grid, nothing elsewhere.
    C[a, b] = z: exchange -V_ee(|q|) z at (b - q, a - q) in Hee; direct term
    +V_ee(|b - a|) z on every entry of the diagonal k1 - k2 = b - a (q_d = b - a != 0)
    in Hee, and -V_eh(|a - b|) z on the diagonal k1 - k2 = a - b in Hhh."""
    nk = 12
    w, _ = make_wire(nk)
    rng = np.random.default_rng(a * 31 + b)
    VC, rows = toeplitz_vc(nk, rng)
    z = 0.3 - 0.7j
    Z = np.zeros((nk, nk), complex)
    E = np.diag(w.Ee + w.gap).astype(complex)
    H = np.diag(w.Eh).astype(complex)

    p = Z.copy()
    p[a, b] = z
    Heh, Hee, Hhh = w.CalcH(Z, Z, Z, p, VC)
    want = Z.copy()
    for q in range(-nk, nk + 1):
        if q != 0 and 0 <= b - q < nk and 0 <= a - q < nk:
            want[b - q, a - q] += rows[0][abs(q)] * z
    assert np.array_equal(Heh, want)
    assert np.array_equal(Hee, E) and np.array_equal(Hhh, H)

    C = Z.copy()
    C[a, b] = z
    Heh, Hee, Hhh = w.CalcH(Z, C, Z, Z, VC)
    want_e, want_h = E.copy(), H.copy()
    for q in range(-nk, nk + 1):
        if q != 0 and 0 <= b - q < nk and 0 <= a - q < nk:
            want_e[b - q, a - q] -= rows[1][abs(q)] * z
    for k1 in range(nk):
        for k2 in range(nk):
            if k1 - k2 == b - a and b != a:
                want_e[k1, k2] += rows[1][abs(b - a)] * z
            if k1 - k2 == a - b and a != b:
                want_h[k1, k2] -= rows[0][abs(a - b)] * z
    np.testing.assert_allclose(Hee, want_e, rtol=1e-15, atol=1e-15 * np.abs(want_e).max())
    np.testing.assert_allclose(Hhh, want_h, rtol=1e-15, atol=1e-15 * np.abs(want_h).max())
    assert np.array_equal(Heh, Z)


# ------------------------------------------------------------------ dpdt / time step
def test_dpdt_free_carriers_is_the_band_frequency():
    """_dpdt_jit (PSTD3D/SBEs.py:1953-1976) with Heh = 0, Hee = diag(Ee + gap),
    Hhh = diag(Eh), no off-diagonal dephasing: dp[kh, ke]/dt =
    (-i (Eh[kh] + Ee[ke] + gap)/hbar - (GamE[ke] + GamH[kh])) p[kh, ke].  Ee != Eh
    (me != mh), so a swapped kh/ke changes the answer."""
    nk = 12
    w, _ = make_wire(nk, FREE)
    rng = np.random.default_rng(3)
    p = rng.standard_normal((nk, nk)) + 1j * rng.standard_normal((nk, nk))
    C = np.diag(rng.uniform(0, 1, nk)).astype(complex)
    D = np.diag(rng.uniform(0, 1, nk)).astype(complex)
    ge, gh = rng.uniform(1e11, 1e12, nk), rng.uniform(1e11, 1e12, nk)
    Z = np.zeros((nk, nk), complex)
    dp = w.dpdt(C, D, p, Z, np.diag(w.Ee + w.gap).astype(complex), np.diag(w.Eh).astype(complex), ge, gh, Z)
    lam = -1j * (w.Eh[:, None] + w.Ee[None, :] + w.gap) / hbar - (ge[None, :] + gh[:, None])
    np.testing.assert_allclose(dp, lam * p, rtol=1e-13)
    assert not np.allclose(dp, lam.T * p, rtol=1e-6)
    # dC/dt and dD/dt: with Heh = 0 a diagonal C, D does not move (PSTD3D/SBEs.py:2320-2338, 2618-2636)
    assert np.abs(w.dCdt(C, D, p, Z, np.diag(w.Ee + w.gap).astype(complex), np.diag(w.Eh).astype(complex))).max() == 0
    assert np.abs(w.dDdt(C, D, p, Z, np.diag(w.Ee + w.gap).astype(complex), np.diag(w.Eh).astype(complex))).max() == 0


def _free_step_err(dt, T):
    nk = 10
    w, rr = make_wire(nk, FREE)
    w.wireoff = False
    rng = np.random.default_rng(5)
    p0 = (rng.standard_normal((nk, nk)) + 1j * rng.standard_normal((nk, nk))) * 1e-3
    st = w.state[0]
    st.p1, st.p2 = p0.copy(), p0.copy()
    Zl = np.zeros(len(rr), complex)
    n = int(round(T / dt))
    for _ in range(n):
        w.QWCalculator(Zl, Zl, Zl, rr, dt, 1)
    lam = -1j * (w.Eh[:, None] + w.Ee[None, :] + w.gap) / hbar - (w.gam_e + w.gam_h)
    # the newest p (p2) is the leapfrog value one step ahead of p1: at (n + 1) dt
    exact = p0 * np.exp(lam * (n + 1) * dt)
    return np.linalg.norm(st.p2 - exact) / np.linalg.norm(exact)


def test_step_free_carriers_converges_to_closed_form():
    """The whole step (_sbe_calculator, PSTD3D/SBEs.py:5624-5832) with every many-body flag
    off and no field: the leapfrog with averaging (:5685-5718) integrates
    p(t) = p0 exp(-i (Eh + Ee + gap) t / hbar - (gam_e + gam_h) t), converging at first
    order in dt."""
    e1 = _free_step_err(2e-18, 1e-15)
    e2 = _free_step_err(1e-18, 1e-15)
    assert e1 < 0.05
    assert 1.7 < e1 / e2 < 2.3


# ------------------------------------------------------------------ optics
@pytest.mark.parametrize("a,b,wi", [(3, 8, 1), (7, 2, 2), (5, 5, 3)])
def test_polarization_single_entry_is_one_plane_wave(a, b, wi):
    """QWPolarization3 (PSTD3D/qwoptics.py:56-82, 815-946) of p with one entry p[kh = a,
    ke = b] = z: Px(r) = Re(z Xvc0[a, b] e^{i (kr[b] - kr[a]) r}) W(r) 2 ehint/(area L), with
    Xvc0[kh, ke] = conj(Xcv0[ke, kh]) = dcv (-1)^(kh + 1) (:620-632); Py with dcv and yw(w),
    Pz with -dcv and (-1)^w; each then FFTG'd."""
    nk = 12
    w, _ = make_wire(nk)
    p = np.zeros((nk, nk), complex)
    z = 0.8 + 0.25j
    p[a, b] = z
    Px, Py, Pz = w.QWPolarization3(p, wi)
    phase = np.exp(1j * (w.kr[b] - w.kr[a]) * w.r)
    pref = 2 * w.ehint / w.area / w.L
    ywv = (-1) ** int(np.floor((wi - 1) / 2.0))
    for got, dip, fac in ((Px, w.dcv * (-1) ** (a + 1), 1.0), (Py, w.dcv, ywv), (Pz, -w.dcv, (-1) ** wi)):
        want = (np.real(z * dip * phase) * w.QWWindow * fac * pref).astype(complex)
        FFTG(want)
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-13 * np.abs(want).max())


def test_rho_of_diagonal_coherences_is_the_window_profile():
    """QWRho5 (PSTD3D/qwoptics.py:367-380, 948-1077) of diagonal C, D: the k sums give
    sum_k n_k |e^{i k r}|^2 W(r)/(2L), a uniform density times the window; after the edge
    subtraction (mean of the 10 first and 10 last points) and the normalisation to
    sum|n| (+1e-50) over sum|rho| dr (+1e-100) the profile is u = W - edge mean,
    rho = NeTotal u/(sum|u| dr), FFTG'd."""
    nk = 12
    w, _ = make_wire(nk)
    rng = np.random.default_rng(8)
    ne, nh = rng.uniform(0, 0.5, nk), rng.uniform(0, 0.3, nk)
    re, rh = w.QWRho5(np.diag(ne).astype(complex), np.diag(nh).astype(complex), ne, nh)
    W = w.QWWindow
    Nr = len(W)
    dr = w.r[1] - w.r[0]
    for got, n in ((re, ne), (rh, nh)):
        u = n.sum() * W / (2 * w.L)
        u = u - (u[:10].sum() + u[Nr - 10:].sum()) / 20.0
        want = (u * (n.sum() + 1e-50) / (np.abs(u).sum() * dr + 1e-100)).astype(complex)
        FFTG(want)
        np.testing.assert_allclose(got, want, rtol=0, atol=1e-13 * np.abs(want).max())


def test_prop2qw_of_a_cubic_is_its_windowed_real_part():
    """Prop2QW (PSTD3D/qwoptics.py:673-721): the spline of a cubic reproduces it exactly
    (libpulsesuite/spliner.py:1287-1396), so Ex on the wire grid is FFTG(Re(f(r)) W(r));
    QW2Prop (:723-811) of FFTG(g) for a cubic g on r returns g on rr (inside the wire's
    span; 0 outside)."""
    w, rr = make_wire(12, O.MBFlags(LF=False))
    f = lambda x: (1.0 + 2.0j) + (0.5 - 1j) * x / w.L + 0.25 * (x / w.L) ** 2 - 0.125j * (x / w.L) ** 3
    Ex, Ey, Ez = w.Prop2QW(rr, f(rr), 2 * f(rr), 0 * rr.astype(complex))
    want = (np.real(f(w.r)) * w.QWWindow).astype(complex)
    FFTG(want)
    np.testing.assert_allclose(Ex, want, rtol=0, atol=1e-12 * np.abs(want).max())
    np.testing.assert_allclose(Ey, 2 * want, rtol=0, atol=1e-12 * np.abs(want).max())
    g = f(w.r)
    lines = [g.copy() for _ in range(5)]
    for x in lines:
        FFTG(x)
    Pxx, Pyy, Pzz, Re, Rh = w.QW2Prop(*lines, rr)
    inside = (rr >= w.r[0]) & (rr <= w.r[-1])
    np.testing.assert_allclose(Pxx[inside], f(rr[inside]), rtol=0, atol=1e-12 * np.abs(g).max())
    assert np.all(Pxx[~inside] == 0.0)


# ------------------------------------------------------------------ rates at zero / full occupation
@pytest.mark.parametrize("fill", [0.0, 1.0])
def test_rates_vanish_at_zero_and_full_occupation_term_by_term(fill):
    """Every Pauli product of the scattering rates has a factor n and a factor 1 - n of
    the same carrier, so each term vanishes when that carrier's occupation is all 0 or
    all 1, whatever the other carrier holds:
      CalcGammaE/H (PSTD3D/dephasing.py:29-103): ne(1-ne) and nh(1-nh)ne + (1-nh)nh(1-ne);
      CalcOffDiagDeph_E2/H2 (:227-317): same-species n(1-n)(1-n) + (1-n)n n, cross-species
      nB(1-nB)(1-nA) + (1-nB)nB nA;
      MBCE (PSTD3D/coulomb.py:957-1012): Win_eh (1-nh)nh ne, Wout_eh (1-ne)(1-nh)nh,
      Win_ee (1-ne)ne ne, Wout_ee (1-ne)(1-ne)ne; MBCH (:1016-1071) likewise.
    So with ne at 0 or 1 and nh random, the electron-electron parts vanish and only the
    electron-hole parts remain (and the converse); with both at 0 or 1, everything is 0."""
    nk = 12
    w, _ = make_wire(nk)
    rng = np.random.default_rng(11)
    VC, _ = toeplitz_vc(nk, rng)
    rnd = rng.uniform(0.05, 0.95, nk)
    full = np.full(nk, fill)
    z = lambda x: np.zeros(nk)

    def mbce(ne, nh, only):
        """Electron rates with one carrier-carrier term kept (the other's V set to 0)."""
        V = VC.copy()
        V[:, :, 1 if only == "eh" else 0] = 0.0
        Win, Wout = np.zeros(nk), np.zeros(nk)
        w.MBCE(ne, nh, V, Win, Wout)
        return Win, Wout

    def mbch(ne, nh, only):
        V = VC.copy()
        V[:, :, 2 if only == "eh" else 0] = 0.0
        Win, Wout = np.zeros(nk), np.zeros(nk)
        w.MBCH(ne, nh, V, Win, Wout)
        return Win, Wout

    # electron-electron terms vanish with ne at 0/1 whatever nh; hole-hole with nh at 0/1
    for W in mbce(full, rnd, "ee") + mbch(rnd, full, "hh"):
        assert np.abs(W).max() == 0.0
    # electron-hole terms: the hole factor (1 - nh) nh (electrons) / (1 - ne) ne (holes)
    for W in mbce(rnd, full, "eh") + mbch(full, rnd, "eh"):
        assert np.abs(W).max() == 0.0
    # and they do not vanish for partial occupations (the pins are not vacuous)
    assert np.abs(mbce(rnd, rnd, "ee")[0]).max() > 0 and np.abs(mbce(rnd, rnd, "eh")[1]).max() > 0
    assert np.abs(mbch(rnd, rnd, "hh")[0]).max() > 0 and np.abs(mbch(rnd, rnd, "eh")[1]).max() > 0

    # diagonal dephasing: both rates vanish with both carriers at 0/1
    assert np.abs(w.CalcGammaE(full, full, VC)).max() == 0.0
    assert np.abs(w.CalcGammaH(full, full, VC)).max() == 0.0
    # the same-species part of Gamma_E alone (hole coupling V_eh = 0): vanishes with ne at 0/1
    V = VC.copy()
    V[:, :, 0] = 0.0
    assert np.abs(w.CalcGammaE(full, rnd, V)).max() == 0.0
    assert np.abs(w.CalcGammaE(rnd, rnd, V)).max() > 0
    V = VC.copy()
    V[:, :, 0] = 0.0
    assert np.abs(w.CalcGammaH(rnd, full, V)).max() == 0.0

    # off-diagonal dephasing: with both at 0/1 every D vanishes; same-species parts alone
    # (V_eh = 0) vanish with their own carrier at 0/1
    p = rng.standard_normal((nk, nk)) + 1j * rng.standard_normal((nk, nk))
    x, De, Dh = w.OffDiagDephasing2(full, full, p, VC)
    assert np.abs(De).max() == 0.0 and np.abs(Dh).max() == 0.0 and np.abs(x).max() == 0.0
    V = VC.copy()
    V[:, :, 0] = 0.0
    _, De, Dh = w.OffDiagDephasing2(full, rnd, p, V)
    assert np.abs(De).max() == 0.0 and np.abs(Dh).max() > 0
    _, De, Dh = w.OffDiagDephasing2(rnd, full, p, V)
    assert np.abs(Dh).max() == 0.0 and np.abs(De).max() > 0


# ------------------------------------------------------------------ DC transport
def _dc_wire(nk=12, Edc=0.0, **mb):
    L = O.L_for_nk(nk)
    rr = np.linspace(-2 * L, 2 * L, 48)
    w = O.OracleWire(rr, 0.0, 1e7, 800e-9, 1, O.QWParams(L=L, Edc=Edc), O.MBFlags(**mb))
    return w


def _field_for_shift(w, m, dt):
    """Eac that moves a distribution by exactly m grid points in one call: the phase of
    Transport (PSTD3D/dcfield.py:1213-1247) is exp(-i Y_j dk), Y_j = s_j 2 pi / ((Nk-1) dky)
    with s_j the FFT-order integer of index j, so dk = m (Nk - 1) dky / Nk gives the DFT
    shift theorem's exp(-2 pi i j m / Nk), a circular roll by m."""
    dky = w.kr[2] - w.kr[1]
    dk = 2.0 * np.pi * m / w.Nk * ((w.Nk - 1) * dky) / (2.0 * np.pi)
    return -dk * hbar / (O.e0 * dt) - w.qwp.Edc


@pytest.mark.parametrize("full", [True, False])
@pytest.mark.parametrize("m", [1, 3, -2])
def test_transport_integer_shift_is_a_circular_roll(full, m):
    w = _dc_wire(Edc=2e5)
    rng = np.random.default_rng(11)
    nk = w.Nk
    X = rng.standard_normal((nk, nk)) + 1j * rng.standard_normal((nk, nk))
    dt = 1e-16
    Y = X.copy()
    w.Transport(Y, _field_for_shift(w, m, dt), dt, full)
    if full:
        want = np.roll(np.roll(X, m, axis=0), m, axis=1)
    else:
        want = X.copy()
        np.fill_diagonal(want, np.roll(np.diag(X), m))
    assert np.abs(Y - want).max() < 1e-12 * np.abs(X).max()


def test_transport_zero_field_identity_unitarity_and_composition():
    w = _dc_wire(Edc=3e6)
    rng = np.random.default_rng(12)
    nk = w.Nk
    X = rng.standard_normal((nk, nk)) + 1j * rng.standard_normal((nk, nk))
    dt = 2e-17
    Y = X.copy()
    w.Transport(Y, -3e6, dt, True)  # Edc + Eac = 0
    assert np.abs(Y - X).max() < 1e-14 * np.abs(X).max()
    A = X.copy()
    w.Transport(A, 1.3e7, dt, True)
    assert abs(np.linalg.norm(A) - np.linalg.norm(X)) < 1e-13 * np.linalg.norm(X)  # a phase in Fourier space
    assert np.abs(A - X).max() > 1e-4 * np.abs(X).max()  # a fractional shift moved it
    # two shifts add: (Edc + E1) + (Edc + E2) = Edc + (E1 + E2 + Edc)
    E1, E2 = 0.3e7, 0.7e7
    B = X.copy()
    w.Transport(B, E1, dt, True)
    w.Transport(B, E2, dt, True)
    C = X.copy()
    w.Transport(C, E1 + E2 + 3e6, dt, True)
    assert np.abs(B - C).max() < 1e-13 * np.abs(X).max()


# ------------------------------------------------------------------ spontaneous emission
def test_ec_special_populations():
    """emission.py:134-175 at empty bands (Ec = -Veh(0)) and one occupied electron state j:
    Ec[k] = Vee(0) - Vee(|j-k|) + Veh(0) for k != j, and -Veh(0) at k = j (the idel factor);
    holes the same with Vhh and the opposite sign of the Veh term."""
    w = _dc_wire(nk=10, Recomb=True)
    nk = w.Nk
    VC, rows = toeplitz_vc(nk, np.random.default_rng(5))
    z = np.zeros(nk)
    assert np.allclose(w.Ec(z, z, VC), -rows[0][0], rtol=1e-15, atol=0)
    j = 3
    e = np.zeros(nk)
    e[j] = 1.0
    k = np.arange(nk)
    want_e = rows[1][0] - rows[1][np.abs(j - k)] + np.where(k != j, rows[0][0], 0.0) - rows[0][0]
    want_h = rows[2][0] - rows[2][np.abs(j - k)] - np.where(k != j, rows[0][0], 0.0) - rows[0][0]
    assert np.allclose(w.Ec(e, z, VC), want_e, rtol=1e-14, atol=1e-36)
    assert np.allclose(w.Ec(z, e, VC), want_h, rtol=1e-14, atol=1e-36)


def test_spont_integral_is_the_midpoint_rule_of_its_integral():
    """SpontIntegral (emission.py:178-227) is the midpoint rule on [0, Nw dhw] of
    (h + E)^3 / (c^3 pi^2 hbar^3) x 3 dcv^2 ehint^2 / (eps0 sqrt(epsr) hbar) L(h, hbar geh) exp(-h / kB T),
    L(a, b) = (b / pi) / (a^2 + b^2) (InitializeEmission, :36-91): scipy's adaptive quadrature
    of that integral agrees to the midpoint rule's O(dhw^2) (a left-point grid would be off
    by O(dhw), ~1e-3 here)."""
    from scipy.constants import c, epsilon_0, k as kB
    from scipy.integrate import quad

    w = _dc_wire(nk=10, Recomb=True)
    H = w.HOmega
    dhw = H[1] - H[0]
    top = len(H) * dhw
    hg = hbar * w.gam_eh
    pref = 3.0 * abs(w.dcv) ** 2 * w.ehint**2 / (epsilon_0 * np.sqrt(w.epsr) * hbar)

    def f(h, E):
        return (h + E) ** 3 / (c**3 * np.pi**2 * hbar**3) * pref * (hg / np.pi) / (h**2 + hg**2) * np.exp(-h / (kB * O.TEMP))

    for E in (1.4 * O.e0, 1.6 * O.e0):
        want, _ = quad(f, 0.0, top, args=(E,), points=[hg, 5 * hg], limit=400, epsrel=1e-12)
        got = w.SpontIntegral(np.array([E]))[0]
        assert abs(got / want - 1.0) < 2e-5
    # Rsp grows with the transition energy as (h + E)^3
    r = w.SpontIntegral(np.array([1.0, 2.0]) * O.e0)
    assert 7.0 < r[1] / r[0] < 8.5


# ------------------------------------------------------------------ RK4 mode
def test_rk4_mode_is_fourth_order_and_leapfrog_is_not():
    """The oracle's RK4 mode (the stage order of _rk4, libpulsesuite/integrator.py:143-157,
    on the wire's coherent right-hand side) on the free two-level wire (OBE flags): halving
    dt shrinks the error of p against a dt/8 run ~16-fold; a wrong stage weight or node
    drops the order.  The reference's leapfrog with its prev averaging converges far more
    slowly on the same problem."""
    nk = 16
    L = O.L_for_nk(nk)
    rr = np.linspace(-3 * L, 3 * L, 48)
    x = rr / L
    E = (1e6 * np.exp(-((x / 1.5) ** 2)) * np.cos(2.0 * x)).astype(complex)
    T = 8e-16

    def run(dt, integ):
        w = O.OracleWire(rr, 0.0, 1e5, 800e-9, 1, O.QWParams(L=L), O.MBFlags(OBE=True), integrator=integ)
        for _ in range(int(round(T / dt))):
            w.QWCalculator(E, 0.5 * E, 0.3 * E, rr, dt, 1)
        return w.state[0].p2

    ref = run(T / 64, "rk4")
    e1 = np.linalg.norm(run(T / 8, "rk4") - ref)
    e2 = np.linalg.norm(run(T / 16, "rk4") - ref)
    assert 14.0 < e1 / e2 < 18.5, (e1, e2)
    lref = run(T / 64, "leapfrog")
    l1 = np.linalg.norm(run(T / 8, "leapfrog") - lref)
    l2 = np.linalg.norm(run(T / 16, "leapfrog") - lref)
    assert l1 / l2 < 4.5 and l1 > 100 * e1


# ---------------------------------------------------------------- longitudinal fields
def _elong_space():
    from oracle import pstd as OP

    return OP.Space(3, 8, 4, 16, 20e-9, 30e-9, 10e-9, 9.1)


def _mode(space, a, b, c):
    i, j, k = np.meshgrid(np.arange(space.Nx), np.arange(space.Ny), np.arange(space.Nz), indexing="ij")
    return np.exp(2j * np.pi * (a * i / space.Nx + b * j / space.Ny + c * k / space.Nz))


@pytest.mark.parametrize("mode", [(1, 0, 0), (0, 3, 0), (0, 0, 13), (3, 1, 5)])
def test_elong_of_one_fourier_mode(mode):
    """ElongfromRho (PSTD3D/rhoPJ.py:784-969) of one Fourier mode exp(i q.r): from rho,
    E = -i q rho / (eps (|q|^2 + s)); from P = x_hat exp(i q.r), E = -q q_x / (eps (|q|^2 + s));
    RhoB = i eps q.E - rho (Gauss's law), so -i q_x |q|^2 / (|q|^2 + s) for that P (the bound
    charge -div P), and -s / (|q|^2 + s) rho for rho alone; s = dq_x / 1e4.  Every output is
    longitudinal (parallel to q), and P is returned transformed in place."""
    from scipy.constants import epsilon_0

    from oracle import coupling as OC
    from oracle import pstd as OP

    sp = _elong_space()
    a, b, c = mode
    q = np.array([OP.GetKxArray(sp)[a], OP.GetKyArray(sp)[b], OP.GetKzArray(sp)[c]])
    s = (OP.GetKxArray(sp)[1] - OP.GetKxArray(sp)[0]) / 1e4
    eps = epsilon_0 * OP.GetEpsr(sp)
    q2 = float(q @ q) + s
    f = _mode(sp, a, b, c)
    zero = np.zeros_like(f)
    EPR, EP, ER, RhoB = OC.ElongfromRho(sp, f, zero.copy(), zero.copy(), zero.copy())
    for n in range(3):
        assert np.allclose(ER[n], -1j * q[n] / (eps * q2) * f, rtol=1e-12, atol=0)
        assert np.allclose(EPR[n], ER[n], rtol=1e-15, atol=0)
        assert np.abs(EP[n]).max() == 0.0
    # q^2 / (q^2 + s) rho - rho cancels to ~7e-14 of rho: rounding leaves ~1e-16 absolute
    assert np.allclose(RhoB, -s / q2 * f, rtol=0, atol=1e-15)
    Px = f.copy()
    EPR, EP, ER, RhoB = OC.ElongfromRho(sp, zero, Px, zero.copy(), zero.copy())
    assert np.allclose(Px, np.fft.fftn(f), rtol=0, atol=1e-9)
    for n in range(3):
        assert np.allclose(EP[n], -q[n] * q[0] / (eps * q2) * f, rtol=1e-12, atol=1e-300)
        assert np.abs(ER[n]).max() == 0.0
    assert np.allclose(RhoB, -1j * q[0] * (q2 - s) / q2 * f, rtol=1e-12, atol=1e-300)


def test_elong_uniform_charge():
    """The q = 0 mode: a uniform rho gives no field and RhoB = -rho."""
    from oracle import coupling as OC

    sp = _elong_space()
    f = np.full((sp.Nx, sp.Ny, sp.Nz), 2.5 + 1.0j)
    zero = np.zeros_like(f)
    EPR, EP, ER, RhoB = OC.ElongfromRho(sp, f, zero.copy(), zero.copy(), zero.copy())
    for F in EPR + EP + ER:
        assert np.abs(F).max() < 1e-30
    assert np.allclose(RhoB, -f, rtol=1e-14)


def test_rhob_analytic_product_rule_and_filter():
    """RhoBAnalytic (PSTD3D/rhoPJ.py:1235-1261) of P placed with Gaussian gates is
    -dPx/dx gy gz - Py dgy gz - Pz gy dgz, with the central difference exact for a quadratic
    line away from the periodic ends; RhoBSmooth of one Fourier mode is the mode times
    exp(-(sx^2 qx^2 + sy^2 qy^2 + sz^2 qz^2) / 2), s = 2 (dx, dy, dz)  (:1263-1285)."""
    from oracle import coupling as OC
    from oracle import pstd as OP

    sp = _elong_space()
    i = np.arange(sp.Nx)
    dx = OP.GetDx(sp)
    Pavg = np.array([0.3 + 0.7 * i + 0.2 * i**2, 1.0 - 0.5 * i, 0.25 * i**2], dtype=complex)
    ry, rz = OP.GetYArray(sp), OP.GetZArray(sp)
    gy, gz = np.exp(-ry**2 / (2 * 25e-18)), np.exp(-rz**2 / (2 * 36e-18))
    dgy, dgz = -ry / 25e-18 * gy, -rz / 36e-18 * gz
    f = _mode(sp, 1, 2, 3)
    An, smooth, _ = OC.rhob_diagnostics(sp, Pavg, gy, gz, dgy, dgz, f)
    dP = (0.7 + 0.4 * i) / dx
    want = (-dP[:, None, None] * gy[None, :, None] * gz[None, None, :]
            - Pavg[1][:, None, None] * dgy[None, :, None] * gz[None, None, :]
            - Pavg[2][:, None, None] * gy[None, :, None] * dgz[None, None, :])
    assert np.allclose(An[1:-1], want[1:-1], rtol=1e-12, atol=0)
    q = (OP.GetKxArray(sp)[1], OP.GetKyArray(sp)[2], OP.GetKzArray(sp)[3])
    s = (2 * dx, 2 * OP.GetDy(sp), 2 * OP.GetDz(sp))
    damp = np.exp(-0.5 * sum((s[n] * q[n]) ** 2 for n in range(3)))
    assert np.allclose(smooth, damp * f, rtol=1e-12, atol=1e-15)
