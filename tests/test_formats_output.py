"""Text formats of the output files (PSTD3D/usefulsubs.py:2088-2182, PSTD3D/qwoptics.py:
1195-1221, PSTD3D/dephasing.py:417-426): known answers for small arrays.  The reference
formats each value as f"{np.float32(x)}", which prints the float64 representation of the
float32-rounded value (3e-9 -> 3.000000026176508e-09, 1e12 -> 999999995904.0)."""

import numpy as np


def test_print_formats(tmp_path, monkeypatch):
    from pulsesuite_amd import output as O

    monkeypatch.chdir(tmp_path)
    monkeypatch.setitem(O._state, "firsttime", True)
    z = np.array([0.5, 1.25])
    O.printIT(np.array([1 + 2j, 3e-9 - 1j]), z, 7, "Wire/x/x.01.k.")
    assert (tmp_path / "dataQW/Wire/x/x.01.k.000007.dat").read_text() == "0.5 1.0 2.0\n1.25 3.000000026176508e-09 -1.0\n"
    O.printITR(np.array([2.0, -0.25]), z, 12, "Wire/y/y.")
    assert (tmp_path / "dataQW/Wire/y/y.000012.dat").read_text() == "0.5 2.0\n1.25 -0.25\n"
    O.printIT2D(np.array([[3 + 4j, 1], [0, -2]]), z, 5, "Wire/C/C.01.k.kp.")
    # |D[i, j]| with j outer: (0,0), (1,0), (0,1), (1,1)
    assert (tmp_path / "dataQW/Wire/C/C.01.k.kp.0000005.dat").read_text() == "5.0\n0.0\n1.0\n2.0\n"
    O.printGam(np.array([1e12, 2e12]), z, 3, "Wire/Ge/Ge.01.k.")
    assert (tmp_path / "dataQW/Wire/Ge/Ge.01.k.00003.dat").read_text() == "0.5 999999995904.0\n1.25 1999999991808.0\n"
    O.printITReal2(np.array([1.5 + 9j, -2.0]), z, 4, "Prop/Ex/Ex.01.r.")
    O.printITReal2(np.array([1.5 + 9j, -2.0]), z, 5, "Prop/Ex/Ex.01.r.")
    assert (tmp_path / "dataQW/Prop/Ex/Ex.01.r.000004.dat").read_text() == "1.5 0.5\n-2.0 1.25\n"
    assert (tmp_path / "dataQW/Prop/Ex/Ex.01.r.000005.dat").read_text() == "1.5\n-2.0\n"


def test_async_writer_order_and_errors():
    import pytest

    from pulsesuite_amd.output import AsyncWriter

    out = []
    w = AsyncWriter()
    for i in range(50):
        w.submit(out.append, i)
    w.flush()
    assert out == list(range(50))
    w.submit(lambda: 1 / 0)
    with pytest.raises(RuntimeError):
        w.flush()
    w.close()
