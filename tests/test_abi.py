"""C ABI checks that need no GPU: the library loads, exports every function that
include/pss.h declares, and the ctypes mirrors of the descriptor structs have the
C layout (offsets checked with a tiny gcc program built from the header)."""

import ctypes
import os
import re
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "pss.h")


def declared():
    txt = open(HDR).read()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|const char\*)\s+(pss_\w+)\s*\(", txt, re.M)))


def test_library_exports_every_declared_symbol():
    from pulsesuite_amd import _lib

    names = declared()
    assert len(names) >= 10
    for n in names:
        assert hasattr(_lib.lib, n), n
    assert set(_lib.EXPORTS) == set(names)
    assert b"gfx950" in _lib.lib.pss_version()


def test_workspace_layout_matches_python_mirror():
    from pulsesuite_amd import _lib
    from pulsesuite_amd.SBEs import ws_layout

    for nk, nr, nrr in ((16, 34, 48), (128, 258, 1024), (1024, 2050, 128)):
        lay = ws_layout(nk, nr, nrr)
        assert _lib.lib.pss_sbe_workspace_bytes(nk, nr, nrr, _lib.FP64) == 8 * lay["total_fp64"]
        assert _lib.lib.pss_sbe_workspace_bytes(nk, nr, nrr, _lib.MIXED) == 8 * lay["total"]


@pytest.mark.parametrize("struct", ["pss_wire_desc", "pss_grid"])
def test_struct_layout_matches_ctypes(struct):
    from pulsesuite_amd import _lib

    cls = _lib.WireDesc if struct == "pss_wire_desc" else _lib.Grid
    fields = [f[0] for f in cls._fields_ if not f[0].startswith("pad")]
    prog = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HDR}"', 'int main(void){',
            f'printf("%zu\\n", sizeof({struct}));']
    for f in fields:
        prog.append(f'printf("%zu\\n", offsetof({struct}, {f}));')
    prog.append("return 0;}")
    with tempfile.TemporaryDirectory() as td:
        src = os.path.join(td, "t.c")
        exe = os.path.join(td, "t")
        open(src, "w").write("\n".join(prog))
        subprocess.run(["gcc", src, "-o", exe], check=True)
        out = [int(x) for x in subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()]
    assert out[0] == ctypes.sizeof(cls)
    for f, off in zip(fields, out[1:]):
        assert getattr(cls, f).offset == off, f


@pytest.mark.parametrize("shape,pml", [((64, 1, 1), (6, 3, 0)), ((32, 32, 32), (17, 6, 6)), ((32, 32, 32), (6, -1, 6))])
def test_cpml_half_rejects_pml_that_does_not_fit(shape, pml):
    """A PML needs an axis longer than 1 and 2 npml <= N (the slab arrays of the
    derivatives and psi assume it); the check runs before any device work, so a bad
    grid is refused on a machine without a GPU too (null buffers are never touched)."""
    from pulsesuite_amd import _lib

    g = _lib.Grid()
    g.Nx, g.Ny, g.Nz = shape
    g.npml_x, g.npml_y, g.npml_z = pml
    rc = _lib.lib.pss_cpml_half(ctypes.byref(g), 0, None, None, None, None, None, None, None, None, -1, None, None,
                                None, None, None, None, None, None, None)
    assert rc != 0
    assert b"npml" in _lib.lib.pss_last_error()


def test_gc_paused_holds_the_collector_off_and_restores_it():
    """Graph captures run under gc_paused: the collector is off inside (no destructor of
    a dead graph or event can run mid-capture) and back to its prior state afterwards."""
    import gc

    from pulsesuite_amd import _lib

    assert gc.isenabled()
    with _lib.gc_paused():
        assert not gc.isenabled()
    assert gc.isenabled()
    gc.disable()
    try:
        with _lib.gc_paused():
            pass
        assert not gc.isenabled()
    finally:
        gc.enable()


def test_jline_fft_needs_power_of_two():
    """The in-kernel FFT of the current line is radix-2: other lengths are refused before
    any device work (null buffers are never touched)."""
    from pulsesuite_amd import _lib

    assert _lib.lib.pss_jline(None, None, None, None, 1, 96, 1e-17, 2e-8, None, 1, 0, None) != 0
    assert b"power-of-two" in _lib.lib.pss_last_error()
    assert _lib.lib.pss_jline(None, None, None, None, 1, 4096, 1e-17, 2e-8, None, 0, 1, None) != 0


def test_pipelined_bookkeeping_over_multi_step_replays():
    """run_pipelined replays graphs of several steps: the host clock, the step counters and
    the periodic finite check must advance as over single steps (a check whenever a
    multiple of check_every is crossed, once per replay).  Host logic only."""
    import types

    from pulsesuite_amd.PSTD3D import PSTD3DPropagator

    checks = []
    fake = types.SimpleNamespace(time=types.SimpleNamespace(t=0.0, dt=0.5, n=0), _nstep=0, check_every=3,
                                 check_finite=lambda: checks.append(fake._nstep))
    for k in (4, 4, 1, 1, 4, 1):
        PSTD3DPropagator._book_steps(fake, k)
    assert fake._nstep == 15 and fake.time.n == 15 and fake.time.t == 7.5
    assert checks == [4, 8, 9, 14, 15]  # replays that crossed 3, 6, 9, 12, 15


def test_fft_planes_acc_refuses_bad_shapes():
    """pss_fft_planes_acc needs Ny = Nz = 64 or 128 and a power-of-two Nx (its x-planes come
    in conjugate pairs g, Nx - g): other shapes are refused before any device work."""
    import ctypes as C

    from pulsesuite_amd import _lib

    nul = (C.c_void_p * 3)()
    for nx, ny, nz, msg in ((96, 128, 128, b"power of two"), (128, 64, 128, b"Ny = Nz"), (128, 256, 256, b"Ny = Nz")):
        g = _lib.Grid()
        g.Nx, g.Ny, g.Nz = nx, ny, nz
        assert _lib.lib.pss_fft_planes_acc(C.byref(g), nul, nul, None, None, None, 1.0, None) != 0
        assert msg in _lib.lib.pss_last_error()
