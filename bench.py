"""Benchmark of the coupled SBE-PSTD time step on one MI355X.

Workload (default): 3-D PSTD on a 128^3 complex grid with a 12-cell CPML on all
faces, additive soft plane source, spectral curl via 3-D FFTs, and one
embedded quantum wire (Nk = 256, GaAs parameters, screened = 0 "bare") driven
by the point field, its current line formed, transformed and added to the E
update every step -- 1000 leapfrog steps.

The wire runs with the reference's ignorewire switch (PSTD3D/SBEs.py:5993-5997):
every kernel of the step runs, and the wire's P and rho outputs are zeroed before
they form the current.  With the back-action on, the point-current loop of this
configuration grows exponentially (about 1.5e16 1/s, the same at dt/2 and dt/10,
the same in the FP64 oracle) and overflows after ~45 steps; --feedback runs it so.

Metric: k-point.steps/s = (sum over wires of Nk) x executed steps / wall time,
inputs resident in HBM, timed region bracketed by synchronize on both sides.
The FP64 form and the mixed form run the same algorithm; the mixed line is
reported only with its error against the FP64 run.

Prints ONE JSON line (rank 0).  --gpus N > 1 is refused: this build runs on one GPU.
"""

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)
# the serving and batched runs spread small launches over up to 8 streams; give the
# process as many hardware queues (HIP reads this at its first GPU call)
os.environ.setdefault("GPU_MAX_HW_QUEUES", "8")

METRIC = "k-point·steps/s mixed vs FP64 GPU at ≤1e-6 rel err; p50/p99 solve latency"
HBM_PEAK = 8.0e12


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--grid", type=int, default=128)
    ap.add_argument("--nk", type=int, default=256)
    ap.add_argument("--npml", type=int, default=12)
    ap.add_argument("--precision", default="both", choices=["fp64", "mixed", "both"])
    ap.add_argument("--cpu-steps", type=int, default=2, help="steps of the CPU oracle sample (0 = skip)")
    ap.add_argument("--no-kernel-timing", action="store_true")
    ap.add_argument("--eager", action="store_true", help="coupled steps as eager launches instead of graph replays")
    ap.add_argument("--traffic-reps", type=int, default=3, help="copies and 3-D FFTs of the traffic workload")
    ap.add_argument("--feedback", action="store_true", help="keep the wire's P/rho in the current (diverges)")
    ap.add_argument("--batched-wires", type=int, default=64, help="wires of the batched run (0 = skip)")
    ap.add_argument("--batched-steps", type=int, default=5)
    ap.add_argument("--batched-nk-max", type=int, default=1024)
    ap.add_argument("--batched-eager", action="store_true", help="batched run without autotune or graph capture")
    ap.add_argument("--workloads", default="coupled,batched,serving,precision256",
                    help="comma list of coupled, batched, serving, precision256 (the JSON line needs coupled); "
                         "alone: traffic (kernels for the HBM-counter runs), ceiling (one FFT pass against "
                         "plain in-place / copy kernels of the same size)")
    ap.add_argument("--p256-steps", type=int, default=1000, help="timed steps of the 256^3 / Nk=1024 precision form")
    ap.add_argument("--serve-requests", type=int, default=48, help="requests of the serving run (0 = skip)")
    ap.add_argument("--serve-rate", type=float, default=2.0,
                    help="Poisson arrival rate (1/s) of the latency run; a burst run (all at t=0) follows")
    return ap.parse_args()


AMP = 1e7  # soft-source amplitude (V/m)


def setup(grid, nk, npml, precision, feedback=False, device_time=True):
    from pulsesuite_amd.PSTD3D import PSTD3DPropagator
    from pulsesuite_amd.rhoPJ import QWArray
    from pulsesuite_amd.typepulse import ps
    from pulsesuite_amd.typespace import ss
    from pulsesuite_amd.typetime import ts
    from pulsesuite_amd.wire_model import L_of_nk, MBFlags, QWParams

    dx = 20e-9
    epsr = 9.1
    dt = 0.9 * dx / (2.99792458e8 / np.sqrt(epsr) * np.sqrt(3.0))
    sp = ss(3, grid, grid, grid, dx, dx, dx, epsr)
    pulse = ps(800e-9, AMP, 100e-15, 50e-15, 0.0)
    mb = MBFlags(Screened=False, ignorewire=not feedback)  # bare Coulomb for this config
    qa = QWArray(QWParams(L=L_of_nk(nk)), mb, Nw=1, Emax0=0.0, precision=precision)
    prop = PSTD3DPropagator(sp, ts(0.0, 1.0, dt, 0), pulse, source_type="soft", boundary_type="cpml",
                            calc_j=qa, npml=npml, check_every=0, precision=precision, device_time=device_time)
    return prop, qa


def time_steps(prop, steps, warmup, graph=True):
    """Wall time of `steps` steps after `warmup` eager ones: graph replays (default) of
    the pipelined step (the next step's CPML-E correction beside the wire step) when the
    propagator supports it, else of the plain step; or eager launches."""
    import torch

    for _ in range(warmup):
        prop.step()
    pipe = graph and prop.pipelined()
    if pipe:
        prop.capture_pipelined()
    elif graph:
        prop.capture_step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    if pipe:
        prop.run_pipelined(steps)
    elif graph:
        prop.run_graph(steps)
    else:
        for _ in range(steps):
            prop.step()
    torch.cuda.synchronize()
    return time.perf_counter() - t0


# HBM bytes per k_fft2<8,16,0,double> launch (one pass over a 128^3 complex128 field),
# from rocprofv3 --pmc FETCH_SIZE and --pmc WRITE_SIZE runs of `bench.py --workloads
# traffic`, each counter scaled by what it reads for a 512 MiB device copy
# (FETCH_SIZE x 2.00, WRITE_SIZE x 1.00): 16471 KiB x 2 + 32768 KiB per pass;
# profiles/fft_pass_128cube_traffic.md.
FFT_PASS_TRAFFIC_128 = (16471.0 * 2.0 + 32768.0) * 1024


def fft_pass_timing(prop, reps=50):
    """Average duration of one FP64 3-D FFT (3 two-level register/LDS passes) on this grid,
    HIP events on the launch stream; algorithmic bytes = 3 passes x (read + write)."""
    import torch

    from pulsesuite_amd import _lib

    F = prop.Ex.clone()
    s = torch.cuda.current_stream()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    for _ in range(3):
        _lib.lib.pss_fft3d(prop.ctx.gp, _lib.ptr(F), -1, _lib.stream_ptr())
    ev0.record(s)
    for _ in range(reps):
        _lib.check(_lib.lib.pss_fft3d(prop.ctx.gp, _lib.ptr(F), -1, _lib.stream_ptr()), "fft")
    ev1.record(s)
    torch.cuda.synchronize()
    ms = ev0.elapsed_time(ev1) / reps
    n = prop.ctx.n
    passes = sum(1 for N in prop.ctx.shape if N > 1)
    bytes_ = passes * 2 * 16 * n
    return ms, bytes_


def batched(nw, steps, nk_max=1024, seed=0, eager=False):
    """Batched wires of uneven Nk (uniform over even values in [64, nk_max]), pump
    log-uniform over 3 decades, detuning +-50 meV; one captured launch per step with
    the autotuned plan; FP64 and mixed on the same batch, errors of mixed vs FP64."""
    import torch

    from pulsesuite_amd.batch import WireBatch, sample_specs
    from pulsesuite_amd.wire_model import MBFlags

    specs = sample_specs(nw, seed=seed, nk_lo=64, nk_hi=nk_max, e_lo=1e6, e_hi=1e9)
    for sp in specs:  # start 2 fs before the pulse peak so the timed steps are driven
        sp.t0 = sp.tp - 2e-15
    out = {"workload": f"{nw} wires, Nk uniform even [64,{nk_max}], E0 log-uniform [1e6,1e9] V/m, detuning "
                       f"+-50 meV, screened, dt 10 as, from 2 fs before the pulse peak, seed {seed}",
           "wires": nw, "steps": steps,
           "kpoints": int(sum(s.Nk for s in specs))}
    states = {}
    for prec in ("fp64", "mixed"):
        b = WireBatch(specs, precision=prec, dt=10e-18, mb_flags=MBFlags(Screened=True), nrr=16)
        if eager:
            b.run(1, graph=False)
        else:
            print(f"[bench] batched {prec}: autotune", file=sys.stderr, flush=True)
            b.autotune(reps=1)
            b.run(1)  # capture with the tuned plan (one eager step)
        best = b.plan
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        b.run(steps, graph=not eager)
        torch.cuda.synchronize()
        el = time.perf_counter() - t0
        out[prec] = {"value": b.kpoint_steps(steps) / el, "ms_per_step": 1e3 * el / steps,
                     "plan": f"buckets={best.buckets} streams={best.streams}" + (" eager" if eager else " graph")}
        states[prec] = [b.get_state(i) for i in range(0, nw, max(1, nw // 8))]
        del b
        torch.cuda.empty_cache()
    errs = {}
    for k in ("p2", "C2", "D2"):
        num = sum(np.linalg.norm(m[k] - f[k]) ** 2 for m, f in zip(states["mixed"], states["fp64"]))
        den = sum(np.linalg.norm(f[k]) ** 2 for f in states["fp64"])
        errs[k[0]] = float(np.sqrt(num / max(den, 1e-300)))
    out["mixed_over_fp64"] = out["mixed"]["value"] / out["fp64"]["value"]
    out["mixed_rel_err_vs_fp64"] = errs
    return out


SERVE = dict(nk_lo=64, nk_hi=512, st_lo=200, st_hi=2000, e_lo=1e6, e_hi=1e8, seed=0)


def serving(nreq, rate, precision="mixed"):
    """Continuous-batching runs of a seeded stream of wire + 1-D PSTD requests (1024
    cells, 16-cell CPML), admitted by HBM budget: latency at a Poisson rate below the
    server's capacity, then the same requests all arriving at once (peak throughput)."""
    from pulsesuite_amd.serve import Server, poisson_requests

    out = {}
    for tag, r in (("poisson", rate), ("burst", float("inf"))):
        print(f"[bench] serving {tag}", file=sys.stderr, flush=True)
        reqs = poisson_requests(nreq, r, **SERVE)
        srv = Server(hbm_budget=64 << 30, max_slots=64, precision=precision)
        srv.run(reqs)
        o = srv.summary()
        o.update(requests=nreq, max_in_flight=srv.max_active,
                 arrivals=f"Poisson {r}/s" if tag == "poisson" else "all at t=0")
        out[tag] = o
        del srv
    out.update(precision=precision,
               workload=f"{nreq} requests, Nk log-uniform [{SERVE['nk_lo']},{SERVE['nk_hi']}], steps log-uniform "
                        f"[{SERVE['st_lo']},{SERVE['st_hi']}], wire + 1-D PSTD 1024 cells npml 16, seed {SERVE['seed']}")
    return out


def precision_form(grid=256, nk=1024, npml=12, steps=100, warmup=2, feedback=False, graph=True):
    """The larger precision form: the coupled step on a grid^3 field grid with one
    Nk-point wire, run in FP64 and in the mixed form for the same steps; throughput
    of both and the mixed run's relative error per output against the FP64 run."""
    res = {}
    for prec in ("fp64", "mixed"):
        print(f"[bench] precision form {grid}^3 Nk={nk} {prec}", file=sys.stderr, flush=True)
        prop, qa = setup(grid, nk, npml, prec, feedback)
        el = time_steps(prop, steps, warmup, graph=graph)
        res[prec] = dict(prop=prop, qa=qa, value=nk * steps / el, ms=1e3 * el / steps)
    errs = precision_errors(res["fp64"]["prop"], res["fp64"]["qa"], res["mixed"]["prop"], res["mixed"]["qa"])
    out = {"workload": f"3-D PSTD {grid}^3 + {npml}-cell CPML + soft source, 1 wire Nk={nk}, "
                       f"{warmup} + {steps} leapfrog steps",
           "steps": steps,
           "fp64": {"value": res["fp64"]["value"], "ms_per_step": res["fp64"]["ms"]},
           "mixed": {"value": res["mixed"]["value"], "ms_per_step": res["mixed"]["ms"]},
           "mixed_over_fp64": res["mixed"]["value"] / res["fp64"]["value"],
           "mixed_rel_err_vs_fp64": errs,
           "mixed_within_1e-6": all(v <= 1e-6 for v in errs.values())}
    del res
    import torch

    torch.cuda.empty_cache()
    return out


def rel_err(a, b):
    import torch

    return float(torch.linalg.vector_norm(a - b) / torch.clamp(torch.linalg.vector_norm(b), min=1e-300))


def precision_errors(p64, q64, pmx, qmx):
    """Relative L2 of the mixed run against the FP64 run after the same steps:
    fields (E, B as vector fields), wire polarisation state p, populations (diagonals
    of C, D), and the coherences C, D."""
    import torch

    out = {}
    for tag, names in (("E", ("Ex", "Ey", "Ez")), ("B", ("Bx", "By", "Bz"))):
        num = sum(float(torch.linalg.vector_norm(getattr(pmx, n) - getattr(p64, n))) ** 2 for n in names)
        den = sum(float(torch.linalg.vector_norm(getattr(p64, n))) ** 2 for n in names)
        out[tag] = (num / max(den, 1e-300)) ** 0.5
    a, b = q64.wires.get_state(0), qmx.wires.get_state(0)
    out["p"] = float(np.linalg.norm(b["p2"] - a["p2"]) / max(np.linalg.norm(a["p2"]), 1e-300))
    for k in ("C2", "D2"):
        da, db = np.diag(a[k]).real, np.diag(b[k]).real
        out["n_" + k[0].lower()] = float(np.linalg.norm(db - da) / max(np.linalg.norm(da), 1e-300))
        out[k[0]] = float(np.linalg.norm(b[k] - a[k]) / max(np.linalg.norm(a[k]), 1e-300))
    return out


def traffic_kernels(grid, reps=3):
    """For `rocprofv3 --pmc FETCH_SIZE` / `--pmc WRITE_SIZE` runs: `reps` plain device
    copies of 512 MiB (the counters' calibration: known bytes read and written), then
    `reps` FP64 3-D FFTs of one grid^3 field (3 k_fft2 passes each)."""
    import torch

    from pulsesuite_amd import _lib
    from pulsesuite_amd._grid import GridContext
    from pulsesuite_amd.typespace import ss

    src = torch.ones(64 << 20, dtype=torch.float64, device="cuda")  # 512 MiB
    dst = torch.empty_like(src)
    for _ in range(reps):
        dst.copy_(src)
    ctx = GridContext(ss(3, grid, grid, grid, 20e-9, 20e-9, 20e-9, 9.1), 1e-17, with_cpml=False, device="cuda")
    F = torch.ones((grid, grid, grid), dtype=torch.complex128, device="cuda")
    for _ in range(reps):
        _lib.check(_lib.lib.pss_fft3d(ctx.gp, _lib.ptr(F), -1, _lib.stream_ptr()), "fft")
    torch.cuda.synchronize()
    print(json.dumps({"traffic_kernels": {"copy_bytes": src.numel() * 8, "fft_grid": grid, "reps": reps}}))


def pass_ceiling(grids=(128, 256), reps=100):
    """What bounds one FFT pass at a given size: per-axis FP64 passes (in place and out of
    place, each axis repeated) against torch's in-place elementwise scale and out-of-place
    copy of the same complex128 field, HIP events on the launch stream.  Bytes per launch
    = read + write of the field."""
    import torch

    from pulsesuite_amd import _lib
    from pulsesuite_amd._grid import GridContext
    from pulsesuite_amd.typespace import ss

    def timed(fn):
        s = torch.cuda.current_stream()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        for _ in range(5):
            fn()
        e0.record(s)
        for _ in range(reps):
            fn()
        e1.record(s)
        torch.cuda.synchronize()
        return 1e3 * e0.elapsed_time(e1) / reps

    out = {}
    for grid in grids:
        ctx = GridContext(ss(3, grid, grid, grid, 20e-9, 20e-9, 20e-9, 9.1), 1e-17, with_cpml=False, device="cuda")
        a = torch.ones((grid, grid, grid), dtype=torch.complex128, device="cuda")
        b = torch.empty_like(a)
        nbytes = 2 * a.numel() * 16
        r = {"bytes_per_launch": nbytes}
        r["torch_inplace_scale_us"] = timed(lambda: a.mul_(1.0))
        r["torch_copy_us"] = timed(lambda: b.copy_(a))
        for ax in range(3):
            r[f"pass{ax}_inplace_us"] = timed(lambda: _lib.check(_lib.lib.pss_fft_line(
                ctx.gp, _lib.ptr(a), _lib.ptr(a), ax, -1, _lib.stream_ptr()), "pass"))
            r[f"pass{ax}_out_of_place_us"] = timed(lambda: _lib.check(_lib.lib.pss_fft_line(
                ctx.gp, _lib.ptr(a), _lib.ptr(b), ax, -1, _lib.stream_ptr()), "pass"))
        r["fft3d_pass_avg_us"] = timed(lambda: _lib.check(_lib.lib.pss_fft3d(
            ctx.gp, _lib.ptr(a), -1, _lib.stream_ptr()), "fft")) / 3
        for k in [k for k in r if k.endswith("_us")]:
            r[k.replace("_us", "_TBps")] = round(nbytes / (r[k] * 1e-6) / 1e12, 3)
            r[k] = round(r[k], 2)
        out[f"{grid}^3"] = r
        del a, b, ctx
        torch.cuda.empty_cache()
    print(json.dumps({"pass_ceiling": out}))


def cpu_baseline(grid, nk, npml, steps, feedback=False):
    """The FP64 NumPy oracle on the same workload for `steps` steps (bounded sample)."""
    from oracle import pstd as OP
    from oracle.coupling import OracleCoupling
    from oracle.sbe import L_for_nk, MBFlags, QWParams

    dx = 20e-9
    epsr = 9.1
    dt = 0.9 * dx / (2.99792458e8 / np.sqrt(epsr) * np.sqrt(3.0))
    sp = OP.Space(3, grid, grid, grid, dx, dx, dx, epsr)
    oc = OracleCoupling(sp, 800e-9, 0.0, Nw=1, qw=QWParams(L=L_for_nk(nk)),
                        mb=MBFlags(Screened=False, ignorewire=not feedback))
    prop = OP.Propagator(sp, OP.Time(0.0, 1.0, dt, 0), OP.Pulse(800e-9, AMP, 100e-15, 50e-15, 0.0), "soft", "cpml",
                         calc_j=oc.calc_j, npml=npml)
    prop.step()  # warm-up
    t0 = time.perf_counter()
    for _ in range(steps):
        prop.step()
    el = time.perf_counter() - t0
    return nk * steps / el, el


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is a design note only",
                          "n_gpus": a.gpus}))
        sys.exit(2)
    import torch

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    wl = set(a.workloads.split(","))
    if "traffic" in wl:  # kernels for the HBM-counter runs: a plain copy of known size, then 3-D FFTs
        traffic_kernels(a.grid, a.traffic_reps)
        return
    if "ceiling" in wl:  # one FFT pass against plain in-place / copy kernels of the same size
        pass_ceiling()
        return
    if "coupled" not in wl:  # partial runs (profiling): no headline value
        line = {"metric": METRIC, "value": None, "note": "coupled workload skipped"}
        if "batched" in wl and a.batched_wires > 0:
            line["batched"] = batched(a.batched_wires, a.batched_steps, a.batched_nk_max, eager=a.batched_eager)
        if "serving" in wl and a.serve_requests > 0:
            line["serving"] = serving(a.serve_requests, a.serve_rate)
        if "precision256" in wl:
            line["precision_form_256"] = precision_form(steps=a.p256_steps, feedback=a.feedback, graph=not a.eager)
        print(json.dumps(line))
        return
    results = {}
    precs = ["fp64", "mixed"] if a.precision == "both" else [a.precision]
    for prec in precs:
        print(f"[bench] coupled step {prec}", file=sys.stderr, flush=True)
        prop, qa = setup(a.grid, a.nk, a.npml, prec, a.feedback)
        el = time_steps(prop, a.steps, a.warmup, graph=not a.eager)
        results[prec] = dict(seconds=el, value=a.nk * a.steps / el, ms=1e3 * el / a.steps, prop=prop, qa=qa)
    main_prec = precs[-1]
    r = results[main_prec]
    errs = None
    if len(precs) == 2:
        errs = precision_errors(results["fp64"]["prop"], results["fp64"]["qa"], results["mixed"]["prop"],
                                results["mixed"]["qa"])
    line = {
        "metric": METRIC,
        "value": r["value"],
        "unit": "k-point*steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": r["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128/f64" if main_prec == "fp64" else
                 "mixed (SBE products fp16 hi/lo MFMA, fp32 accumulate; CPML derivative transforms fp32; fields, "
                 "populations, time step fp64)",
        "data": "synthetic (seeded soft-source pulse, reference initial wire state)",
        "config": {"workload": f"3-D PSTD {a.grid}^3 + {a.npml}-cell CPML + soft source, 1 wire Nk={a.nk}, "
                               f"{a.steps} leapfrog steps", "grid": a.grid, "nk": a.nk, "npml": a.npml,
                   "wires": 1, "precision": main_prec,
                   "launch": "eager" if a.eager else ("HIP graph replays of pipelined steps, 4 per graph (wire step on "
                                                        "a second stream beside the B half and the next CPML-E "
                                                        "correction)")},
    }
    if errs is not None:
        line["fp64_value"] = results["fp64"]["value"]
        line["fp64_ms_per_step"] = results["fp64"]["ms"]
        line["mixed_over_fp64"] = results["mixed"]["value"] / results["fp64"]["value"]
        line["mixed_rel_err_vs_fp64"] = errs
        line["mixed_within_1e-6"] = all(v <= 1e-6 for v in errs.values())
    if not a.no_kernel_timing:
        # one pass launch: algorithmic bytes = read + write of every complex128 cell
        ms, b = fft_pass_timing(r["prop"])
        npass = sum(1 for N in r["prop"].ctx.shape if N > 1)
        pass_ms, pass_bytes = ms / npass, b / npass
        ach = pass_bytes / (pass_ms * 1e-3)
        line["roofline"] = {"bound": "hbm", "kernel": f"k_fft2 pass of the {a.grid}^3 complex128 3-D FFT",
                            "achieved": ach / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": ach / HBM_PEAK,
                            "traffic": FFT_PASS_TRAFFIC_128 if a.grid == 128 else None,
                            "algorithmic_bytes": pass_bytes, "avg_us": 1e3 * pass_ms}
    if "batched" in wl and a.batched_wires > 0:
        line["batched"] = batched(a.batched_wires, a.batched_steps, a.batched_nk_max, eager=a.batched_eager)
    if "serving" in wl and a.serve_requests > 0:
        line["serving"] = serving(a.serve_requests, a.serve_rate)
    if "precision256" in wl:
        line["precision_form_256"] = precision_form(steps=a.p256_steps, feedback=a.feedback, graph=not a.eager)
    if a.cpu_steps > 0:
        v, el = cpu_baseline(a.grid, a.nk, a.npml, a.cpu_steps, a.feedback)
        line["cpu_baseline"] = {"value": v, "unit": "k-point*steps/s", "cores": len(os.sched_getaffinity(0)),
                                "kind": "port", "sample": f"{a.cpu_steps} steps of the same workload, "
                                f"{el:.1f} s, NumPy/SciPy FP64 oracle"}
    print(json.dumps(line))


if __name__ == "__main__":
    main()
