"""Benchmark: nonlinear-term evaluations per second on one MI355X.

Workload: the u x omega nonlinear term on a spherical shell at lmax = 255
(nlat = 384, nphi = 768), nr = 128 physical shells, FC-Gram C = 25, d = 5
(radial FFT length 153), shell ri = 0.35, ro = 1.0.  Velocity potentials W, Z
are seeded complex normal with spectrum (l+1)^-2 and radial profile
sin(pi (r - ri)/(ro - ri)).  One step = one full evaluation
(W, Z) -> (NQ, NS, NT) with all inputs resident in HBM.

Prints one JSON line.  --gpus above 1 is refused: the path is single-GPU.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

LMAX, NR, C, D = 255, 128, 25, 5
RI, RO = 0.35, 1.0
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFLOPS = 78.6


def workload():
    nlat = 4 * ((3 * (LMAX + 1) + 7) // 8)
    nphi = 2 * nlat
    n = NR + C
    dz = (RO - RI) / (NR - 1)
    return nlat, nphi, n, dz


def seeded_fields(nlm, n):
    from tests import helpers as H
    return H.shell_fields(LMAX, NR, n, RI, RO, seed=20261203, slope=2.0)


def cpu_baseline(W, Z, nlat, nphi, n, dz, gpu_out, budget_s=15.0):
    """The oracle (CPU FP64 restatement, OpenMP) on a bounded sample of shells of
    the same workload; also checks the GPU result on those shells."""
    from oracle import oracle as O
    from tests import helpers as H
    import redondito_amd.sphp as S
    A, Q = O.load_tables(S.TABLE_DIR, C, D)
    sht = O.SHT(LMAX, nlat, nphi)
    fc = O.FC(n, C, D, O.fc_dir(A, Q))
    cores = os.cpu_count() or 1
    # the radial part is done in full on every call; time it separately
    t0 = time.perf_counter()
    O.nonlinear_uxw(sht, fc, RI, dz, W, Z, 0, 0)
    t_rad = time.perf_counter() - t0
    nsample = 4
    while True:
        t0 = time.perf_counter()
        ref = O.nonlinear_uxw(sht, fc, RI, dz, W, Z, 0, nsample)
        t = time.perf_counter() - t0
        if t > budget_s / 4 or nsample >= NR:
            break
        nsample = min(NR, nsample * 4)
    t_ang = max(t - t_rad, 1e-9) / nsample
    t_eval = t_rad + t_ang * NR
    errs = [H.rell2_joint([g[:nsample] for g in gpu_out], [r[:nsample] for r in ref])]
    errs += [H.rell2(g[:nsample], r[:nsample]) for g, r in zip(gpu_out, ref)]
    return {
        "value": 1.0 / t_eval,
        "unit": "nonlinear-term evals/s",
        "cores": cores,
        "kind": "port",
        "sample": (f"oracle u x omega on shells 0..{nsample - 1} of {NR} (radial step on all shells) "
                   f"= {t:.2f} s, extrapolated to all {NR} shells; OpenMP threads = {cores}"),
    }, {"joint": errs[0], "NQ": errs[1], "NS": errs[2], "NT": errs[3], "shells": nsample}


def copy_calibration(nbytes=1 << 30, reps=3):
    """A plain device copy of known size, for scaling the FETCH_SIZE and
    WRITE_SIZE counters of a profiler run (each copy reads and writes nbytes)."""
    import torch
    a = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(reps):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes": nbytes, "reps": reps}))


def measured_traffic(kernel):
    """HBM bytes per launch of a kernel from the committed counter summary
    (profiles/hbm_traffic_uxw_lmax255_nr128.json), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_uxw_lmax255_nr128.json")
    try:
        with open(path) as fh:
            d = json.load(fh)
        return d["kernels"][kernel]["bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 1 GiB device copy (counter calibration) and exit")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this path runs on one GPU only; --gpus above 1 is not supported", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        copy_calibration()
        return

    import torch
    import redondito_amd.sphp as S

    nlat, nphi, n, dz = workload()
    dev = torch.device("cuda:0")
    plan = S.FCPlan(LMAX, nlat, nphi, n, C, D, dz, RI, device=dev)
    W, Z = seeded_fields(plan.nlm, n)
    Wd = torch.from_numpy(W).to(dev)
    Zd = torch.from_numpy(Z).to(dev)
    out = [torch.zeros_like(Wd) for _ in range(3)]

    for _ in range(args.warmup):
        S.nonlinear_uxw(plan, Wd, Zd, *out)
    torch.cuda.synchronize()

    # timed region, with per-stage HIP events on the plan's stream
    plan.profile(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        S.nonlinear_uxw(plan, Wd, Zd, *out)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    stages = plan.stage_times()
    plan.profile(False)
    ms_step = (t1 - t0) * 1e3 / args.steps
    evals = 1e3 / ms_step

    L, nlm = LMAX + 1, plan.nlm
    Sb = 16.0 * nlm * NR
    Fm = 16.0 * nlat * L * NR
    G = 8.0 * nlat * nphi * NR
    leg_flops = 10.0 * nlat * nlm * NR       # one vector synthesis or analysis
    per_stage = {k: v[0] / max(v[1], 1) for k, v in stages.items()}
    alg = {
        "radial": ("hbm", 8 * Sb),
        "synthesis_u": ("mfma", leg_flops),
        "synthesis_w": ("mfma", leg_flops),
        "phi_cross": ("hbm", 9 * Fm),
        "analysis": ("mfma", leg_flops),
    }
    dom = max(alg, key=lambda k: per_stage.get(k, 0.0))
    bound, work = alg[dom]
    t_dom = per_stage[dom] * 1e-3
    if bound == "hbm":
        achieved, peak, unit = work / t_dom / 1e9, HBM_PEAK_GBS, "GB/s"
    else:
        achieved, peak, unit = work / t_dom / 1e12, FP64_PEAK_TFLOPS, "TFLOP/s"
    # stage-wise roofline time of the whole evaluation (FP64)
    t_roof = (5 * Sb + 11 * Sb + 6 * Sb + 6 * Fm + 6 * Fm + 6 * G + 9 * G + 3 * G + 3 * Fm + 3 * Fm + 3 * Sb) / (HBM_PEAK_GBS * 1e9)
    t_roof_leg = 3 * leg_flops / (FP64_PEAK_TFLOPS * 1e12)

    host_out = [o.cpu().numpy() for o in out]
    cpu = None
    parity = None
    if not args.no_cpu:
        cpu, parity = cpu_baseline(W, Z, nlat, nphi, n, dz, host_out)

    line = {
        "metric": "nonlinear-term evals/s @lmax=255,nr=128: mixed-prec vs FP64, rel err <=1e-9",
        "value": round(evals, 3),
        "unit": "nonlinear-term evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic seeded W,Z: complex normal, (l+1)^-2, sin radial profile",
        "config": {"workload": "u x omega nonlinear term, plain FP64 form", "lmax": LMAX, "nlat": nlat,
                   "nphi": nphi, "nr": NR, "fcgram_C": C, "fcgram_d": D, "radial_fft_len": n,
                   "form": "fp64"},
        "roofline": {"bound": bound, "kernel": dom, "achieved": round(achieved, 3), "peak": peak, "unit": unit,
                     "frac": round(achieved / peak, 4), "traffic": measured_traffic(dom)},
        "stage_ms": {k: round(v, 4) for k, v in per_stage.items()},
        "eval_roofline_frac": round((t_roof + t_roof_leg) / (ms_step * 1e-3), 4),
        "rel_err_vs_fp64": 0.0,
        "rel_l2_vs_oracle": parity,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    plan.close()


if __name__ == "__main__":
    main()
