"""Oracle helpers: grids, centred FFTs, the FMM cubic spline, locators.

Test infrastructure only (see ``oracle/__init__.py``).

Pinning:
  GetSpaceArray / GetKArray / GetArray0Index / locator  pinned (closed forms).
  FFTG / iFFTG          pinned (explicit DFT definition + round trip).
  spline_dpc / seval_dpc / rescale_1D_dpc
                        pinned (exact reproduction of cubics, dense-solve
                        transcription) and golden (reference spliner fixture).
"""

import numpy as np

twopi = 2.0 * np.pi


def GetSpaceArray(N, L):
    """N points evenly spaced over [-L/2, L/2].  PSTD3D/typespace.py:352-359"""
    if N == 1:
        return np.array([0.0])
    return np.linspace(-L / 2.0, L / 2.0, N)


def GetKArray(N, L):
    """FFT-order wavenumbers, +N/2 at index N/2.  PSTD3D/typespace.py:362-381"""
    if N == 1:
        return np.array([0.0])
    dl = twopi / L if L > 0 else 1.0
    i = np.arange(N)
    return np.where(i <= N // 2, i, i - N).astype(np.float64) * dl


def GetArray0Index(x):
    """First index whose value rounds to zero.  PSTD3D/usefulsubs.py:1028-1058"""
    dx = x[2] - x[1]
    xi = np.round((x + 0.1 * dx) / dx).astype(np.int32)
    hits = np.nonzero(xi == 0)[0]
    if len(hits) == 0:
        raise ValueError("Zero index not found in array")
    return int(hits[0])


def locator(x, x0):
    """Clamped searchsorted index.  PSTD3D/usefulsubs.py:1823-1850"""
    i = np.searchsorted(x, x0, side="right") - 1
    return int(max(0, min(i, len(x) - 2)))


def locate(xx, x):
    """Bracketing index, n-1 at/after the last knot.  libpulsesuite/spliner.py:1137-1170

    Vectorised over ``x``; equals the reference's bisection for sorted ``xx``.
    """
    xx = np.asarray(xx)
    x = np.asarray(x, dtype=np.float64)
    n = len(xx)
    j = np.searchsorted(xx, x, side="right") - 1
    j = np.where(x < xx[0], 0, j)
    j = np.where(x >= xx[n - 1], n - 1, j)
    return j


def nyquist_1D(F):
    """Flip the sign of odd 0-based entries.  PSTD3D/usefulsubs.py:2265-2275"""
    F[1::2] = -F[1::2]


def FFTG(F):
    """Centred forward FFT scaled by -1/N, in place.  PSTD3D/usefulsubs.py:2278-2304"""
    N = len(F)
    nyquist_1D(F)
    F[:] = np.fft.fft(F)
    nyquist_1D(F)
    F[:] = -F / (N * 1.0)


def iFFTG(F):
    """Centred inverse FFT scaled by -N, in place.  PSTD3D/usefulsubs.py:2307-2333"""
    N = len(F)
    nyquist_1D(F)
    F[:] = np.fft.ifft(F)
    nyquist_1D(F)
    F[:] = -F * N


def spline_dpc(x, y):
    """FMM cubic-spline coefficients (b, c, d) for complex y.

    Follows libpulsesuite/spliner.py:1287-1363 statement by statement
    (third-derivative end conditions from cubics through the end 4 points).
    """
    n = len(x)
    y = np.asarray(y, dtype=np.complex128)
    b = np.zeros(n, dtype=complex)
    c = np.zeros(n, dtype=complex)
    d = np.zeros(n, dtype=complex)
    if n < 2:
        raise ValueError("Not enough points to spline.")
    if n < 3:
        b[0] = (y[1] - y[0]) / (x[1] - x[0])
        b[1] = b[0]
        return b, c, d
    d[0] = x[1] - x[0]
    c[1] = (y[1] - y[0]) / d[0]
    for i in range(1, n - 1):
        d[i] = x[i + 1] - x[i]
        b[i] = 2.0 * (d[i - 1] + d[i])
        c[i + 1] = (y[i + 1] - y[i]) / d[i]
        c[i] = c[i + 1] - c[i]
    b[0] = -d[0]
    b[n - 1] = -d[n - 2]
    c[0] = 0.0
    c[n - 1] = 0.0
    if n != 3:
        c[0] = c[2] / (x[3] - x[1]) - c[1] / (x[2] - x[0])
        c[n - 1] = c[n - 2] / (x[n - 1] - x[n - 3]) - c[n - 3] / (x[n - 2] - x[n - 4])
        c[0] = c[0] * d[0] ** 2 / (x[3] - x[0])
        c[n - 1] = -c[n - 1] * d[n - 2] ** 2 / (x[n - 1] - x[n - 4])
    for i in range(1, n):
        t = d[i - 1] / b[i - 1]
        b[i] = b[i] - t * d[i - 1]
        c[i] = c[i] - t * c[i - 1]
    c[n - 1] = c[n - 1] / b[n - 1]
    for i in range(n - 2, -1, -1):
        c[i] = (c[i] - d[i] * c[i + 1]) / b[i]
    b[n - 1] = (y[n - 1] - y[n - 2]) / d[n - 2] + d[n - 2] * (c[n - 2] + 2.0 * c[n - 1])
    for i in range(n - 1):
        b[i] = (y[i + 1] - y[i]) / d[i] - d[i] * (c[i + 1] + 2.0 * c[i])
        d[i] = (c[i + 1] - c[i]) / d[i]
        c[i] = 3.0 * c[i]
    c[n - 1] = 3.0 * c[n - 1]
    d[n - 1] = d[n - 2]
    return b, c, d


def seval_dpc(u, x, y, b, c, d):
    """Evaluate the spline at points u.  libpulsesuite/spliner.py:1367-1396"""
    i = locate(x, u)
    dx = u - x[i]
    return y[i] + dx * (b[i] + dx * (c[i] + dx * d[i]))


def rescale_1D_dpc(x0, z0, x1):
    """Spline z0(x0) onto x1; zero outside [min x0, max x0].

    libpulsesuite/spliner.py:153-188
    """
    if len(x0) != len(z0):
        raise ValueError("Bad sizes in rescale_1D_dpc")
    z0 = np.asarray(z0, dtype=np.complex128)
    b, c, d = spline_dpc(x0, z0)
    x1 = np.asarray(x1, dtype=np.float64)
    inside = (x1 >= np.min(x0)) & (x1 <= np.max(x0))
    out = np.zeros(len(x1), dtype=complex)
    if np.any(inside):
        out[inside] = seval_dpc(x1[inside], x0, z0, b, c, d)
    return out


def rk4(y, dydt, t, h, D):
    """Numerical-Recipes RK4 step, stage order as libpulsesuite/integrator.py:143-157"""
    ytmp = y + (h / 2.0) * dydt
    dyt = D(t + h / 2.0, ytmp)
    ytmp = y + (h / 2.0) * dyt
    dym = D(t + h / 2.0, ytmp)
    ytmp = y + h * dym
    dym = dym + dyt
    dyt = D(t + h, ytmp)
    y[:] = y + (h / 6.0) * (dydt + dyt + 2.0 * dym)
    return t + h
