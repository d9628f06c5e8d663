"""Output files of the wire step at the jmax cadence, in the reference's text formats.

The writers keep the reference's names, paths and line formats (PSTD3D/qwoptics.py:
1195-1392, PSTD3D/usefulsubs.py:2088-2192, PSTD3D/dephasing.py:417-438, the rates of
PSTD3D/SBEs.py:5288-5320): files under dataQW/ of the working directory, one value or one
'z re im' line per point as str(np.float32).  ``snapshot`` copies what they need of one
wire from the device right after its step (the state, the Hamiltonian's diagonals, the
dephasing and scattering rates, the fields and sources on the wire grid and on the
propagation line); ``AsyncWriter`` formats and writes the files on a host thread, in
submission order, so the step is not held up by file output."""

import os
import queue
import threading

import numpy as np
import torch

_E0 = 1.602176634e-19  # scipy.constants.e (the reference's eV)
_state = {"firsttime": True}  # printITReal2's first-call flag (PSTD3D/qwoptics.py:47, 1209-1221)


def _path(file, n, width):
    filename = f"dataQW/{file}{n:0{width}d}.dat"
    os.makedirs(os.path.dirname(filename), exist_ok=True)
    return filename


def printIT(Dx, z, n, file):
    """'z re im' per point (PSTD3D/usefulsubs.py:2088-2120)."""
    with open(_path(file, n, 6), "w", encoding="utf-8") as f:
        for i in range(len(z)):
            f.write(f"{np.float32(z[i])} {np.float32(np.real(Dx[i]))} {np.float32(np.imag(Dx[i]))}\n")


def printITR(Dx, z, n, file):
    """'z value' per point of a real array (PSTD3D/usefulsubs.py:2123-2153)."""
    with open(_path(file, n, 6), "w", encoding="utf-8") as f:
        for i in range(len(z)):
            f.write(f"{np.float32(z[i])} {np.float32(Dx[i])}\n")


def printIT2D(Dx, z, n, file):
    """|Dx[i, j]| per line, j outer (PSTD3D/usefulsubs.py:2156-2182); z is not written."""
    with open(_path(file, n, 7), "w", encoding="utf-8") as f:
        for j in range(Dx.shape[1]):
            for i in range(Dx.shape[0]):
                f.write(f"{np.float32(np.abs(Dx[i, j]))}\n")


def printITReal2(Dx, z, n, file):
    """'re z' per point on the process's first call, 're' afterwards (PSTD3D/qwoptics.py:1195-1221)."""
    with open(_path(file, n, 6), "w", encoding="utf-8") as f:
        if _state["firsttime"]:
            for i in range(len(z)):
                f.write(f"{np.float32(np.real(Dx[i]))} {np.float32(z[i])}\n")
            _state["firsttime"] = False
        else:
            for i in range(len(z)):
                f.write(f"{np.float32(np.real(Dx[i]))}\n")


def printGam(Dx, z, n, file):
    """'z value' per point, five-digit index (PSTD3D/dephasing.py:417-426)."""
    filename = f"dataQW/{file}{n:05d}.dat"
    os.makedirs(os.path.dirname(filename), exist_ok=True)
    with open(filename, "w", encoding="utf-8") as f:
        for i in range(len(z)):
            f.write(f"{np.float32(z[i])} {np.float32(Dx[i])}\n")


def WriteSBESolns(ky, ne, nh, C, D, P, Ee, Eh, w, xxx):
    """PSTD3D/qwoptics.py:1269-1298"""
    wire = f"{w:02d}"
    printIT2D(C, ky, xxx, f"Wire/C/C.{wire}.k.kp.")
    printIT2D(D, ky, xxx, f"Wire/D/D.{wire}.k.kp.")
    printIT2D(P, ky, xxx, f"Wire/P/P.{wire}.k.kp.")
    printIT(ne, ky, xxx, f"Wire/ne/ne.{wire}.k.")
    printIT(nh, ky, xxx, f"Wire/nh/nh.{wire}.k.")
    printIT(Ee, ky, xxx, f"Wire/Ee/Ee.{wire}.k.")
    printIT(Eh, ky, xxx, f"Wire/Eh/Eh.{wire}.k.")


def WriteDephasing(ky, gamE, gamH, w, xxx):
    """PSTD3D/dephasing.py:429-438"""
    wire = f"{w:02d}"
    printGam(gamE, ky, xxx, f"Wire/Ge/Ge.{wire}.k.")
    printGam(gamH, ky, xxx, f"Wire/Gh/Gh.{wire}.k.")


def WriteRates(ky, WinE, WinH, WoutE, WoutH, ne, nh, w, xxx):
    """The in / out scattering rates times the blocking factors (PSTD3D/SBEs.py:5294-5320)."""
    wire = f"{w:02d}"
    printITR(WinE * (1 - np.real(ne)), ky, xxx, f"Wire/Win/Win.e.{wire}.k.")
    printITR(WinH * (1 - np.real(nh)), ky, xxx, f"Wire/Win/Win.h.{wire}.k.")
    printITR(WoutE * np.real(ne), ky, xxx, f"Wire/Wout/Wout.e.{wire}.k.")
    printITR(WoutH * np.real(nh), ky, xxx, f"Wire/Wout/Wout.h.{wire}.k.")


def WriteQWFields(QY, Ex, Ey, Ez, Vr, Px, Py, Pz, Re, Rh, sp, w, xxx):
    """PSTD3D/qwoptics.py:1322-1355"""
    wire = f"{w:02d}"
    for name, a in (("Ex", Ex), ("Ey", Ey), ("Ez", Ez), ("Vr", Vr), ("Px", Px), ("Py", Py), ("Pz", Pz),
                    ("Re", Re), ("Rh", Rh), ("Rho", Rh - Re)):
        printIT(a, QY, xxx, f"Wire/{name}/{name}.{wire}.{sp}.")


def WritePropFields(y, Ex, Ey, Ez, Vr, Px, Py, Pz, Re, Rh, sp, w, xxx):
    """PSTD3D/qwoptics.py:1358-1392"""
    wire = f"{w:02d}"
    for name, a in (("Ex", Ex), ("Ey", Ey), ("Ez", Ez), ("Vr", Vr), ("Px", Px), ("Py", Py), ("Pz", Pz),
                    ("Re", Re), ("Rh", Rh)):
        printITReal2(a, y, xxx, f"Prop/{name}/{name}.{wire}.{sp}.")
    printIT(Rh - Re, y, xxx, f"Prop/Rho/Rho.{wire}.{sp}.")


def snapshot(dev, i, H=None, prop_lines=None, flags=None):
    """Host copies of what the writers need of wire i after its step: the new state X3
    (the cur buffers), populations, the Hamiltonian's diagonals H (taken between the RHS and
    the tail of the step, before a DCTrans shift reuses that memory; given by the caller),
    the dephasing and scattering rates, the populations before relaxation, the fields,
    polarisation and densities on the wire grid, and the propagation-line outputs."""
    m = dev.models[i]
    Nk, Nr, n = m.Nk, m.Nr, dev.nrr[i]
    snap = {"kr": m.arrays["kr"], "r": m.arrays["r"], "Nk": Nk}
    for X in ("C", "D", "p"):
        snap[X] = dev.buffer(i, X, "cur").cpu().numpy()
    if H is not None:
        snap["Ee"], snap["Eh"] = H[0] / _E0, H[1] / _E0
    snap["gam"] = dev.workspace_view(i, "gam").cpu().numpy()
    snap["W"] = dev.workspace_view(i, "W").cpu().numpy()
    snap["n3"] = dev.workspace_view(i, "n3").cpu().numpy()
    snap["er"] = dev.workspace_view(i, "er").cpu().numpy().reshape(3, Nr)
    snap["pr"] = dev.workspace_view(i, "pr").cpu().numpy().reshape(3, Nr)
    snap["rho"] = dev.workspace_view(i, "rho").cpu().numpy().view(complex).reshape(2, Nr)
    snap["outr"] = dev.workspace_view(i, "outr").cpu().numpy().reshape(5, n)
    if prop_lines is not None:
        snap["prop"] = [np.asarray(a.cpu().numpy() if isinstance(a, torch.Tensor) else a) for a in prop_lines]
    snap["flags"] = flags
    return snap


def write_snapshot(snap, w, xxx, rr):
    """The reference's files of one wire at one output step (PSTD3D/SBEs.py:5294-5320,
    5724-5727, PSTD3D/qwoptics.py:798-799, PSTD3D/SBEs.py:5931-5934), from ``snapshot``."""
    fl = snap["flags"]
    kr, r, Nk = snap["kr"], snap["r"], snap["Nk"]
    C, D, p = snap["C"], snap["D"], snap["p"]
    ne3, nh3 = np.diag(C).copy(), np.diag(D).copy()
    if fl.EHs or fl.Phonon:
        W, n3 = snap["W"], snap["n3"]
        WriteRates(kr, W[:Nk], W[2 * Nk:3 * Nk], W[Nk:2 * Nk], W[3 * Nk:], n3[2 * Nk:3 * Nk], n3[3 * Nk:], w, xxx)
    if "Ee" in snap:
        WriteSBESolns(kr, ne3, nh3, C, D, p, snap["Ee"].astype(complex), snap["Eh"].astype(complex), w, xxx)
    if fl.DiagDph:
        WriteDephasing(kr, snap["gam"][:Nk], snap["gam"][Nk:], w, xxx)
    er, pr, rho = snap["er"], snap["pr"], snap["rho"]
    zero = np.zeros(len(r), complex)
    re, rh = (rho[0], rho[1]) if fl.LF else (zero, zero)
    WriteQWFields(r, er[0].astype(complex), er[1].astype(complex), er[2].astype(complex), zero,
                  pr[0].astype(complex), pr[1].astype(complex), pr[2].astype(complex), re, rh, "r", w, xxx)
    if "prop" in snap:
        Exx, Eyy, Ezz, Vrr, Pxx, Pyy, Pzz = snap["prop"]
        o = snap["outr"]
        WritePropFields(rr, Exx, Eyy, Ezz, Vrr, Pxx, Pyy, Pzz, o[3].astype(complex), o[4].astype(complex), "r", w,
                        xxx)


class AsyncWriter:
    """A host thread that runs submitted writes in order; ``flush`` waits for all of them."""

    def __init__(self):
        self._q = queue.Queue()
        self._err = None
        self._t = threading.Thread(target=self._run, daemon=True)
        self._t.start()

    def _run(self):
        while True:
            job = self._q.get()
            if job is None:
                self._q.task_done()
                return
            try:
                job()
            except Exception as e:  # reported by flush
                self._err = e
            self._q.task_done()

    def submit(self, fn, *args):
        self._q.put(lambda: fn(*args))

    def flush(self):
        self._q.join()
        if self._err is not None:
            e, self._err = self._err, None
            raise RuntimeError(f"output writer failed: {e}") from e

    def close(self):
        self.flush()
        self._q.put(None)
        self._t.join()
