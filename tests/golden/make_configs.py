"""Write the oracle fixtures of the five benchmark workloads (tests/golden/config_*.npz).

Runs the FP64 oracle (oracle/, NumPy on the CPU) on the inputs that tests/config_cases.py
defines, and stores what tests/test_gpu_configs.py compares against: per checkpoint the
sketches and norms of the six field components (the fields themselves on 1-D lines), the
sketches and norms of the wire's p, C, D, its populations, and its polarisation and charge
lines.  C4 stores the eight oracle wires of the batch, C5 one served request.

Usage: python tests/golden/make_configs.py C1 C2 C3 C4 C5   (one process per case; C4 runs
its eight wires in parallel processes).  Single-threaded BLAS per process is advised
(OMP_NUM_THREADS=1); the largest cases (C3, the Nk ~ 1024 wire of C4) take tens of minutes.
"""

import os
import sys
import time
from concurrent.futures import ProcessPoolExecutor

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))  # tests/
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))  # repo root

import config_cases as CC  # noqa: E402


def log(*a):
    print(f"[{time.strftime('%H:%M:%S')}]", *a, flush=True)


def run_coupled(name):
    prop, oc = CC.oracle_coupled(name)
    c = CC.coupled_params(name)
    out = {"dt": c["dt"], "amp": c["amp"], "x0": c["x0"], "steps": c["steps"],
           "checkpoints": np.array(c["checkpoints"])}
    t0 = time.time()
    for n in range(1, c["steps"] + 1):
        prop.step()
        if n % 100 == 0 or n in c["checkpoints"]:
            log(name, "step", n, f"{time.time() - t0:.0f}s", "max|E|",
                max(float(np.abs(getattr(prop, f)).max()) for f in ("Ex", "Ey", "Ez")))
        if n in c["checkpoints"]:
            fields = {f: getattr(prop, f) for f in CC.FIELDS}
            st = oc.wire.state[0]
            state = {"p2": st.p2, "C2": st.C2, "D2": st.D2}
            rec = CC.coupled_record(fields, state, oc.Pold[:, :, 0].copy(), oc.RhoOld[:, 0].copy(), seed=7 + n)
            for k, v in rec.items():
                out[f"s{n}_{k}"] = v
    np.savez_compressed(os.path.join(HERE, f"config_{name}.npz"), **out)
    log(name, "written")


def _c4_one(i):
    specs = CC.c4_specs()
    sp = specs[i]
    t0 = time.time()
    st, r = CC.c4_oracle_wire(sp)
    sk = CC.sketch([st.p2, st.C2, st.D2], seed=100 + i)
    log("C4 wire", i, "Nk", sp.Nk, f"{time.time() - t0:.0f}s")
    return i, {"sk_p2": sk[0], "sk_C2": sk[1], "sk_D2": sk[2], "norm_p2": np.linalg.norm(st.p2),
               "norm_C2": np.linalg.norm(st.C2), "norm_D2": np.linalg.norm(st.D2),
               "ne": np.diag(st.C2).real.copy(), "nh": np.diag(st.D2).real.copy(),
               "P": np.stack(r[:3]), "Nk": sp.Nk}


def run_c4():
    specs = CC.c4_specs()
    picks = CC.c4_oracle_picks(specs)
    out = {"picks": np.array(picks)}
    # largest wires first so the pool finishes together
    order = sorted(picks, key=lambda i: -specs[i].Nk)
    with ProcessPoolExecutor(max_workers=min(8, len(order))) as ex:
        for i, rec in ex.map(_c4_one, order):
            for k, v in rec.items():
                out[f"w{i}_{k}"] = v
    np.savez_compressed(os.path.join(HERE, "config_C4.npz"), **out)
    log("C4 written")


def run_c5():
    from oracle import pstd as OP
    from oracle.coupling import OracleCoupling
    from oracle.sbe import L_for_nk, MBFlags, QWParams
    from pulsesuite_amd import serve as S

    c = CC.C5_ORACLE
    sp = OP.Space(3, S.NX, 1, 1, S.DX, S.DX, S.DX, S.EPSR)
    oc = OracleCoupling(sp, 800e-9, 0.0, Nw=1, qw=QWParams(L=L_for_nk(c["nk"])), mb=MBFlags(), current="reference")
    prop = OP.Propagator(sp, OP.Time(0.0, 1.0, S.DT, 0), OP.Pulse(800e-9, c["E0"], 100e-15, 50e-15, 0.0), "soft",
                         "cpml", calc_j=oc.calc_j, npml=S.NPML)
    for n in range(c["steps"]):
        prop.step()
    st = oc.wire.state[0]
    np.savez_compressed(os.path.join(HERE, "config_C5.npz"), Ex=prop.Ex.reshape(-1), Ey=prop.Ey.reshape(-1),
                        p=st.p2, C=st.C2, D=st.D2)
    log("C5 written")


if __name__ == "__main__":
    names = sys.argv[1:] or ["C1", "C2", "C3", "C4", "C5"]
    for nm in names:
        if nm in CC.COUPLED:
            run_coupled(nm)
        elif nm == "C4":
            run_c4()
        elif nm == "C5":
            run_c5()
        else:
            raise SystemExit(f"unknown case {nm}")
