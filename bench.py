"""Benchmark: nonlinear-term evaluations per second on one MI355X.

Workload: the u x omega nonlinear term on a spherical shell at lmax = 255
(nlat = 384, nphi = 768), nr = 128 physical shells, FC-Gram C = 25, d = 5
(radial FFT length 153), shell ri = 0.35, ro = 1.0.  Velocity potentials W, Z
are seeded complex normal with spectrum (l+1)^-2 and radial profile
sin(pi (r - ri)/(ro - ri)).  One step = one full evaluation
(W, Z) -> (NQ, NS, NT) with all inputs resident in HBM.

Four forms run on the same inputs: FP64, and the precision forms with the
Legendre contractions on FP32 MFMA, as BF16x3 splits, and as int8 digit slices
of the FP64 operands with exact int32 accumulation (the Ozaki form).  Each is timed over
the same number of steps and its error against the FP64 output is reported
per field and per degree band.  The headline value is the fastest form whose
joint relative L2 error against FP64 is at most 1e-9.

Also reported:
  * each form's evaluation replayed from a hipGraph;
  * absolute scalar transforms/s, the wall solves at the same size and the
    HBM bandwidth of a plain device copy;
  * the batching form: 48 uneven problems, persistent launches, hipGraph,
    for the scalar 3D round trip and for the nonlinear term of every problem;
  * the large coupled step at lmax = 511, nr = 256 (u x omega, u x B and
    u . grad T) in all forms, with per-field and per-band errors;
  * the CPU oracle timed on all host threads and on one, with its 3D
    transform at the size the reference's CPU timing is quoted on.

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
SEED = 20261203
HBM_PEAK_GBS = 8000.0
# dense MFMA peaks; the int8 digit form spends 28 int8 products per FP64 product,
# so its FP64-equivalent peak is the 5.0 POP/s int8 peak over 28
MFMA_PEAK_TFLOPS = {"fp64": 78.6, "fp32": 157.3, "bf16x3": 2500.0, "ozaki": 5000.0 / 28, "ozaki6": 5000.0 / 21}
# (name, precision form, Ozaki digits per operand)
FORMS = (("fp64", 0, 7), ("fp32", 1, 7), ("bf16x3", 2, 7), ("ozaki", 3, 7), ("ozaki6", 3, 6))
ERR_BAR = 1e-9


def workload():
    nlat = 4 * ((3 * (LMAX + 1) + 7) // 8)
    nphi = 2 * nlat
    n = NR + C
    dz = (RO - RI) / (NR - 1)
    return nlat, nphi, n, dz


def seeded_fields(n):
    from tests import helpers as H
    return H.shell_fields(LMAX, NR, n, RI, RO, seed=SEED, slope=2.0)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def cpu_calibration(O, S):
    """The oracle's 3D transforms at the size LEEME.md:36 quotes for the
    reference's CPU path (lmax 63, nlat = nphi = 128, nz = 128, C = 25, d = 5:
    11 ms per transform on one MPI rank), on one thread and on all threads;
    ms per transform (the mean of the forward and the inverse)."""
    lmax, nlat, nphi, nz = 63, 128, 128, 128
    A, Q = O.load_tables(S.TABLE_DIR, C, D)
    sht = O.SHT(lmax, nlat, nphi)
    fc = O.FC(nz, C, D, O.fc_dir(A, Q))
    rng = np.random.default_rng(SEED + 3)
    nlm = (lmax + 1) * (lmax + 2) // 2
    a = rng.standard_normal((nz, nlm)) + 1j * rng.standard_normal((nz, nlm))
    out = {}
    for label, threads in (("one_thread_ms", 1), ("all_threads_ms", 0)):
        O.set_threads(threads)
        best = 1e30
        for _ in range(3):
            t0 = time.perf_counter()
            g = sht.synth(fc.inverse(a.copy()))
            t1 = time.perf_counter()
            fc.forward(sht.anal(g))
            t2 = time.perf_counter()
            best = min(best, 0.5 * (t2 - t0))
        out[label] = round(best * 1e3, 2)
    O.set_threads(0)
    out["reference_ms_one_rank"] = 11.0
    out["config"] = "3D transform, lmax 63, nlat = nphi = 128, nz = 128, C 25, d 5 (LEEME.md:36)"
    return out


def cpu_baseline(W, Z, nlat, nphi, n, dz, gpu_out, budget_s=15.0):
    """The oracle (CPU FP64 restatement, OpenMP) on a bounded sample of shells of
    the same workload, on all host threads and on one; also checks the GPU
    FP64 result on those shells."""
    from oracle import oracle as O
    from tests import helpers as H
    import redondito_amd.sphp as S
    A, Q = O.load_tables(S.TABLE_DIR, C, D)
    sht = O.SHT(LMAX, nlat, nphi)
    fc = O.FC(n, C, D, O.fc_dir(A, Q))
    cores = O.set_threads(0)
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
    # one thread: the radial step on every shell plus two shells
    O.set_threads(1)
    t0 = time.perf_counter()
    O.nonlinear_uxw(sht, fc, RI, dz, W, Z, 0, 0)
    t1_rad = time.perf_counter() - t0
    t0 = time.perf_counter()
    O.nonlinear_uxw(sht, fc, RI, dz, W, Z, 0, 2)
    t1 = time.perf_counter() - t0
    O.set_threads(0)
    t1_eval = t1_rad + max(t1 - t1_rad, 1e-9) / 2 * NR
    return {
        "value": 1.0 / t_eval,
        "unit": "nonlinear-term evals/s",
        "cores": cores,
        "kind": "port",
        "sample": (f"oracle u x omega on shells 0..{nsample - 1} of {NR} (radial step on all shells) "
                   f"= {t:.2f} s, extrapolated to all {NR} shells; OpenMP threads = {cores}"),
        "cpu_model": cpu_model(),
        "one_thread": {"value": 1.0 / t1_eval, "sample": (f"oracle u x omega on shells 0..1 (radial step on all "
                                                          f"shells) = {t1:.2f} s on 1 thread, extrapolated")},
        "calibration": cpu_calibration(O, S),
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


def copy_bandwidth(nbytes=1 << 30, reps=5):
    """HBM bandwidth of a plain device copy (read + write bytes per second),
    to set beside the 8 TB/s peak."""
    import torch
    a = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda:0")
    b = torch.empty_like(a)
    b.copy_(a)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        b.copy_(a)
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a, b
    return round(2 * nbytes / (ms * 1e-3) / 1e9, 1)


def measured_traffic(kernel):
    """HBM bytes per launch of a stage's kernel from the committed counter
    summary (profiles/hbm_traffic_uxw_lmax255_nr128.json), or None."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_uxw_lmax255_nr128.json")
    try:
        with open(path) as fh:
            d = json.load(fh)
        return d["kernels"][kernel]["bytes_per_launch"]
    except (OSError, KeyError, ValueError):
        return None


def time_form(S, plan, prec, Wd, Zd, out, steps, warmup):
    """Warm up, then time exactly `steps` evaluations bracketed by device
    synchronisations; returns (ms per step, per-stage ms)."""
    import torch
    for _ in range(warmup):
        S.nonlinear_uxw(plan, Wd, Zd, *out, prec=prec)
    torch.cuda.synchronize()
    plan.profile(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        S.nonlinear_uxw(plan, Wd, Zd, *out, prec=prec)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    stages = plan.stage_times()
    plan.profile(False)
    return (t1 - t0) * 1e3 / steps, {k: v[0] / max(v[1], 1) for k, v in stages.items()}


def time_graph(S, plan, prec, Wd, Zd, out, steps):
    """One evaluation captured into a hipGraph (after rd_plan_prepare), then
    `steps` replays timed like the eager steps; returns ms per step."""
    import torch
    plan.prepare(prec)
    plan.capture_uxw(Wd, Zd, *out, prec=prec)
    plan.replay()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        plan.replay()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / steps


def time_calls(fn, reps=10):
    import torch
    fn()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) * 1e3 / reps


def aux_timings(S, plan, Wd):
    """Absolute scalar transforms (one SHT over all nr shells) and the FP64
    wall solves at the same size."""
    import torch
    nlat, nphi, nlm = plan.nlat, plan.nphi, plan.nlm
    f = torch.zeros((NR, nphi, nlat), dtype=torch.float64, device=Wd.device)
    a = Wd[:NR].contiguous()
    b = torch.zeros_like(a)
    syn = time_calls(lambda: S.fftp2d_complex_to_real_xy(plan, a, f))
    ana = time_calls(lambda: S.fftp2d_real_to_complex_xy(plan, f, b))
    bcp = S.BCPlan()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    S.setup_bc(bcp, plan, ("noslip", "noslip"), "v")
    torch.cuda.synchronize()
    setup = (time.perf_counter() - t0) * 1e3
    tor, phi, pol = Wd.clone(), Wd.clone(), torch.zeros_like(Wd)
    inv = time_calls(lambda: S.invDl(plan, phi, pol))
    full = time_calls(lambda: S.v_imposebc(tor, phi, pol, plan))
    fast = time_calls(lambda: S.v_imposebc_fast(tor, phi, pol, plan))
    return {
        "scalar_synthesis_ms": round(syn, 4), "scalar_analysis_ms": round(ana, 4),
        "transforms_per_s": round(2e3 / (syn + ana), 2),
        "bc_setup_ms": round(setup, 2), "invDl_ms": round(inv, 4),
        "v_imposebc_ms": round(full, 4), "v_imposebc_fast_ms": round(fast, 4),
    }


def batch_problems(seed=SEED, n=48):
    """48 shell problems: lmax from {31, 63, 95, 127, 191, 255}, nlat = 3(lmax+1)/2
    (rounded to the grid rule), nphi = 2 nlat, nr from {32, 64, 128}, one seed."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        lmax = int(rng.choice([31, 63, 95, 127, 191, 255]))
        nr = int(rng.choice([32, 64, 128]))
        nlat = 4 * ((3 * (lmax + 1) + 7) // 8)
        out.append((lmax, nlat, 2 * nlat, nr + C, (RO - RI) / (nr - 1), RI))
    return out


def batch_bench(reps):
    """The batching form: 48 uneven scalar problems, FP32 MFMA Legendre, the
    round trip spec -> grid -> spec of all of them, eager and hipGraph-replayed,
    with tuned and fixed tiles."""
    import torch
    import redondito_amd.sphp as S
    from tests import helpers as H
    probs = batch_problems()
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(SEED + 1)
    spec, grid, out = [], [], []
    for lmax, nlat, nphi, nz, dz, ri in probs:
        ell = H.ell_of(lmax)
        c = (rng.standard_normal((nz, ell.size)) + 1j * rng.standard_normal((nz, ell.size))) * (ell + 1.0) ** -1.5
        c[:, : lmax + 1] = c[:, : lmax + 1].real
        spec.append(torch.from_numpy(c).to(dev))
        grid.append(torch.empty((nz, nphi, nlat), dtype=torch.float64, device=dev))
        out.append(torch.empty_like(spec[-1]))
    b = S.Batch(probs, prec=S.RD_PREC_FP32)
    ms_fixed, ms_tuned, choice = b.tune(spec, grid, out, reps=2)
    split = b.get_split()

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / reps

    eager = timed(lambda: (b.inverse(spec, grid), b.forward(grid, out)))
    b.capture(spec, grid, out)
    graph = timed(b.replay)
    # accuracy of the FP32 round trip on band-limited, FC-Gram-consistent data
    b.forward(grid, spec)
    b.inverse(spec, grid)
    b.forward(grid, out)
    torch.cuda.synchronize()
    err = max(H.rell2(o.cpu().numpy(), s.cpu().numpy()) for o, s in zip(out, spec))
    nsyn = sum(p[3] for p in probs)     # planes transformed per direction
    info = b.info()
    b.close()
    return {
        "problems": len(probs), "prec": "fp32 Legendre", "planes_per_direction": nsyn,
        "round_trip_ms_eager": round(eager, 4), "round_trip_ms_graph": round(graph, 4),
        "round_trip_ms_fixed_tiles": round(ms_fixed, 4), "round_trip_ms_tuned_tiles": round(ms_tuned, 4),
        "tuned_choice": {"plane_tile_blocks": choice[0], "workgroups_per_cu": choice[1],
                         "one_problem_lmax_min": split[0], "one_problem_count": split[1]},
        "transforms_per_s": round(2 * len(probs) * 1e3 / graph, 1),
        "round_trip_rel_err": float(f"{err:.3e}"),
        "work_items": {"synthesis": info["syn_items"], "analysis": info["ana_items"]},
    }


def batch_uxw_bench(reps):
    """The nonlinear-term form of the batch on the same 48-problem draw: u x
    omega of every problem (velocity potentials seeded as in the headline
    workload), the Legendre stage of all problems as one persistent launch per
    direction, FP64 and FP32 Legendre, eager and hipGraph-replayed, tuned
    against fixed tiles; one step = one evaluation of all 48 problems."""
    import torch
    import redondito_amd.sphp as S
    from tests import helpers as H
    probs = batch_problems()
    dev = torch.device("cuda:0")
    W, Z = [], []
    for i, (lmax, nlat, nphi, nz, dz, ri) in enumerate(probs):
        w, z = H.shell_fields(lmax, nz - C, nz, ri, RO, seed=SEED + 100 + i, slope=2.0)
        W.append(torch.from_numpy(w).to(dev))
        Z.append(torch.from_numpy(z).to(dev))
    res, outs = {}, {}

    def timed(fn):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1e3 / reps

    # the baseline the batch replaces: one plan per problem, the problems'
    # evaluations one after another on one stream
    single = {}
    plans = [S.FCPlan(lmax, nlat, nphi, nz, C, D, dz, ri, device=dev) for lmax, nlat, nphi, nz, dz, ri in probs]
    for name, prec in (("fp64", S.RD_PREC_FP64), ("fp32", S.RD_PREC_FP32)):
        out1 = [[torch.zeros_like(w) for w in W] for _ in range(3)]

        def run_all():
            for i, pl in enumerate(plans):
                S.nonlinear_uxw(pl, W[i], Z[i], out1[0][i], out1[1][i], out1[2][i], prec=prec)

        single[name] = timed(run_all)
        del out1
    for pl in plans:
        pl.close()
    del plans
    torch.cuda.empty_cache()

    for name, prec in (("fp64", S.RD_PREC_FP64), ("fp32", S.RD_PREC_FP32)):
        b = S.Batch(probs, prec=prec, form="uxw")
        out = [[torch.zeros_like(w) for w in W] for _ in range(3)]
        fixed, tuned, choice = b.tune_uxw(W, Z, *out, reps=2)
        split = b.get_split()
        eager = timed(lambda: b.nonlinear_uxw(W, Z, *out))
        b.capture_uxw(W, Z, *out)
        graph = timed(b.replay)
        torch.cuda.synchronize()
        outs[name] = [[x.cpu().numpy() for x in o] for o in out]
        res[name] = {"ms_eager": round(eager, 3), "ms_graph": round(graph, 3), "ms_fixed_tiles": round(fixed, 3),
                     "ms_tuned_tiles": round(tuned, 3),
                     "tuned_choice": {"plane_tile_blocks": choice[0], "workgroups_per_cu": choice[1],
                                      "one_problem_lmax_min": split[0], "one_problem_count": split[1]},
                     "evals_per_s": round(len(probs) * 1e3 / min(graph, eager), 1),
                     "ms_single_plans": round(single[name], 3),
                     "speedup_vs_single_plans": round(single[name] / min(graph, eager), 3)}
        del out
        b.close()
        torch.cuda.empty_cache()
    num = sum(float(np.linalg.norm(a - r)) ** 2 for k in range(3) for a, r in zip(outs["fp32"][k], outs["fp64"][k]))
    den = sum(float(np.linalg.norm(r)) ** 2 for k in range(3) for r in outs["fp64"][k])
    res["fp32"]["rel_err_vs_fp64"] = float(f"{(num / den) ** 0.5:.3e}")
    return {"problems": len(probs), "step": "u x omega of every problem", **res}


CP_LMAX, CP_NLAT, CP_NPHI, CP_NR = 511, 768, 1536, 256
CP_NAMES = ("NQ", "NS", "NT", "MQ", "MS", "MT", "NH")


def coupled_bench(reps):
    """The large coupled state: lmax = 511, nlat = 768, nphi = 1536, nr = 256
    (radial length 281), velocity W, Z ~ (l+1)^-2, magnetic Pb, Tb ~ (l+1)^-2.5,
    temperature T ~ (l+1)^-2.  One step = u x omega, u x B and u . grad T.
    The three Legendre forms are timed on the same inputs and their error
    against the FP64 step reported per output field and degree band; the FP64
    wall solves are timed at the same size."""
    import torch
    import redondito_amd.sphp as S
    from tests import helpers as H
    dev = torch.device("cuda:0")
    n = CP_NR + C
    dz = (RO - RI) / (CP_NR - 1)
    plan = S.FCPlan(CP_LMAX, CP_NLAT, CP_NPHI, n, C, D, dz, RI, device=dev)
    ins = [torch.from_numpy(x).to(dev) for x in H.coupled_fields(CP_LMAX, CP_NR, n, RI, RO, seed=SEED + 7)]
    forms, host = {}, {}
    for name, prec, digits in FORMS:
        plan.set_ozaki_digits(digits)
        out = [torch.zeros_like(ins[0]) for _ in range(7)]
        ms = time_calls(lambda: S.nonlinear_coupled(plan, *ins, *out, prec=prec), reps)
        host[name] = [o.cpu().numpy() for o in out]
        del out
        forms[name] = {"ms_per_step": round(ms, 3), "steps_per_s": round(1e3 / ms, 2)}
    for name in forms:
        rep = H.band_errors(host[name], host["fp64"], CP_LMAX, CP_NR, names=CP_NAMES)
        grp = [H.rell2_joint([x[:CP_NR] for x in host[name][a:b]], [x[:CP_NR] for x in host["fp64"][a:b]])
               for a, b in ((0, 3), (3, 6), (6, 7))]
        forms[name]["rel_err_vs_fp64"] = {
            "u_x_omega": float(f"{grp[0]:.3e}"), "u_x_B": float(f"{grp[1]:.3e}"), "u_grad_T": float(f"{grp[2]:.3e}"),
            **{f: {b: float(f"{e:.3e}") for b, e in rep[f].items()} for f in CP_NAMES}}
        forms[name]["speedup_vs_fp64"] = round(forms["fp64"]["ms_per_step"] / forms[name]["ms_per_step"], 3)
    host.clear()
    bcp = S.BCPlan()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    S.setup_bc(bcp, plan, ("noslip", "noslip"), "v")
    torch.cuda.synchronize()
    setup = (time.perf_counter() - t0) * 1e3
    tor, phi, pol = ins[1].clone(), ins[0].clone(), torch.zeros_like(ins[0])
    inv = time_calls(lambda: S.invDl(plan, phi, pol), reps)
    fast = time_calls(lambda: S.v_imposebc_fast(tor, phi, pol, plan), reps)
    plan.close()
    return {
        "lmax": CP_LMAX, "nlat": CP_NLAT, "nphi": CP_NPHI, "nr": CP_NR, "radial_fft_len": n,
        "step": "u x omega, u x B, u . grad T", "forms": forms,
        "bc_setup_ms": round(setup, 2), "invDl_ms": round(inv, 4), "v_imposebc_fast_ms": round(fast, 4),
    }


def roofline(form, per_stage, nlat, nlm, kernel=None):
    """A kernel of a form against its bound (the dominant one unless named),
    algorithmic work per launch."""
    L = LMAX + 1
    Sb = 16.0 * nlm * NR                    # one spectral field on nr shells
    Fm = 16.0 * nlat * L * NR               # one Fourier field (m <= lmax)
    leg_flops = 10.0 * nlat * nlm * NR      # one vector synthesis or analysis
    alg = {
        # reads W, Z, writes six (Q,S,T) arrays, clears the outputs' rows past nr
        "radial": ("hbm", 8 * Sb + 3 * 16.0 * nlm * C),
        "synthesis": ("mfma", 2 * leg_flops),   # u and omega in one launch
        "phi_cross": ("hbm", 9 * Fm),
        "analysis": ("mfma", leg_flops),
    }
    dom = kernel or max(alg, key=lambda k: per_stage.get(k, 0.0))
    bound, work = alg[dom]
    t_dom = per_stage[dom] * 1e-3
    if bound == "hbm":
        achieved, peak, unit = work / t_dom / 1e9, HBM_PEAK_GBS, "GB/s"
    else:
        achieved, peak, unit = work / t_dom / 1e12, MFMA_PEAK_TFLOPS[form], "TFLOP/s"
    return {"bound": bound, "kernel": dom, "achieved": round(achieved, 3), "peak": peak, "unit": unit,
            "frac": round(achieved / peak, 4), "traffic": measured_traffic(dom) if form == "fp64" else None}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--fp64-only", action="store_true", help="time only the FP64 form (for profiler runs)")
    ap.add_argument("--no-batch", action="store_true", help="skip the batching-form measurement")
    ap.add_argument("--no-coupled", action="store_true", help="skip the lmax = 511 coupled-step measurement")
    ap.add_argument("--coupled-only", action="store_true",
                    help="only run the lmax = 511 coupled-step measurement (for profiler runs) and print it")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run a 1 GiB device copy (counter calibration) and exit")
    ap.add_argument("--fp64-ceilings", action="store_true",
                    help="only measure the FP64 MFMA and VALU ceilings of the device and exit")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this path runs on one GPU only; --gpus above 1 is not supported", file=sys.stderr)
        sys.exit(2)
    if args.calibrate_copy:
        copy_calibration()
        return
    if args.fp64_ceilings:
        from redondito_amd import sphp as S
        mf, vf = S.fp64_ceilings()
        m4 = S.fp64_mfma4x4_ceiling()
        print(json.dumps({"fp64_mfma_ceiling_tflops": round(mf, 2), "fp64_mfma4x4_ceiling_tflops": round(m4, 2),
                          "fp64_valu_ceiling_tflops": round(vf, 2), "published_fp64_peak_tflops": 78.6}))
        return
    if args.coupled_only:
        print(json.dumps({"coupled": coupled_bench(max(2, args.steps))}))
        return

    import torch
    import redondito_amd.sphp as S
    from tests import helpers as H

    nlat, nphi, n, dz = workload()
    dev = torch.device("cuda:0")
    plan = S.FCPlan(LMAX, nlat, nphi, n, C, D, dz, RI, device=dev)
    W, Z = seeded_fields(n)
    Wd = torch.from_numpy(W).to(dev)
    Zd = torch.from_numpy(Z).to(dev)

    forms = {}
    host = {}
    for name, prec, digits in FORMS:
        if args.fp64_only and prec != 0:
            continue
        plan.set_ozaki_digits(digits)
        out = [torch.zeros_like(Wd) for _ in range(3)]
        ms, stages = time_form(S, plan, prec, Wd, Zd, out, args.steps, args.warmup)
        host[name] = [o.cpu().numpy() for o in out]
        gms = time_graph(S, plan, prec, Wd, Zd, [torch.zeros_like(Wd) for _ in range(3)], args.steps)
        forms[name] = {"evals_per_s": round(1e3 / ms, 3), "ms_per_step": round(ms, 4),
                       "graph_ms_per_step": round(gms, 4),
                       "stage_ms": {k: round(v, 4) for k, v in stages.items()}, "_stages": stages}
    for name in forms:
        rep = H.band_errors(host[name], host["fp64"], LMAX, NR)
        forms[name]["rel_err_vs_fp64"] = {
            "joint": float(f"{rep['joint']:.3e}"),
            **{f: {b: float(f"{e:.3e}") for b, e in rep[f].items()} for f in ("NQ", "NS", "NT")}}
        forms[name]["speedup_vs_fp64"] = round(forms["fp64"]["ms_per_step"] / forms[name]["ms_per_step"], 3)

    ok = [k for k in forms if forms[k]["rel_err_vs_fp64"]["joint"] <= ERR_BAR]
    head = min(ok, key=lambda k: forms[k]["ms_per_step"])
    nlm = plan.nlm
    roof = roofline(head, forms[head]["_stages"], nlat, nlm)
    # every kernel of the evaluation against its own bound
    roof_all = {k: roofline(head, forms[head]["_stages"], nlat, nlm, kernel=k)
                for k in ("radial", "synthesis", "phi_cross", "analysis")}
    # whole-evaluation bound, stage by stage: sum of max(flops / MFMA peak,
    # algorithmic bytes / HBM peak) over the Legendre contractions (two
    # syntheses, one analysis), the radial step and the phi step
    L = LMAX + 1
    Sb, Fm = 16.0 * nlm * NR, 16.0 * nlat * L * NR
    bw, leg = HBM_PEAK_GBS * 1e9, 10.0 * nlat * nlm * NR / (MFMA_PEAK_TFLOPS[head] * 1e12)
    t_roof = ((8 * Sb + 48.0 * nlm * C) / bw + 2 * max(leg, (3 * Sb + 3 * Fm) / bw) + 9 * Fm / bw
              + max(leg, (3 * Fm + 3 * Sb) / bw))
    aux = aux_timings(S, plan, Wd)
    aux["hbm_copy_gbs"] = copy_bandwidth()
    mf, vf = S.fp64_ceilings()
    aux["fp64_mfma_ceiling_tflops"] = round(mf, 2)   # measured, beside the 78.6 published
    aux["fp64_mfma4x4_ceiling_tflops"] = round(S.fp64_mfma4x4_ceiling(), 2)
    aux["fp64_valu_ceiling_tflops"] = round(vf, 2)
    batch = None if (args.no_batch or args.fp64_only) else batch_bench(max(2, min(args.steps, 10)))
    batch_uxw = None if (args.no_batch or args.fp64_only) else batch_uxw_bench(max(2, min(args.steps, 10)))
    coupled = None if (args.no_coupled or args.fp64_only) else coupled_bench(max(2, min(args.steps, 5)))

    cpu = parity = None
    if not args.no_cpu:
        cpu, parity = cpu_baseline(W, Z, nlat, nphi, n, dz, host["fp64"])

    for f in forms.values():
        f.pop("_stages")
    line = {
        "metric": "nonlinear-term evals/s @lmax=255,nr=128: mixed-prec vs FP64, rel err <=1e-9",
        "value": forms[head]["evals_per_s"],
        "unit": "nonlinear-term evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": forms[head]["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"fp64": "f64", "fp32": "f32 Legendre, f64 elsewhere",
                  "bf16x3": "bf16x3 Legendre, f64 elsewhere",
                  "ozaki": "int8 digit-sliced f64 Legendre (exact int32 accumulation), f64 elsewhere",
                  "ozaki6": "int8 digit-sliced f64 Legendre, 6 digits (exact int32 accumulation), f64 elsewhere"}[head],
        "data": "synthetic seeded W,Z: complex normal, (l+1)^-2, sin radial profile",
        "config": {"workload": "u x omega nonlinear term", "form": head, "lmax": LMAX, "nlat": nlat,
                   "nphi": nphi, "nr": NR, "fcgram_C": C, "fcgram_d": D, "radial_fft_len": n},
        "roofline": roof,
        "rooflines": roof_all,
        "eval_roofline_frac": round(t_roof / (forms[head]["ms_per_step"] * 1e-3), 4),
        "forms": forms,
        "aux": aux,
        "batch": batch,
        "batch_uxw": batch_uxw,
        "coupled": coupled,
        "rel_l2_vs_oracle": parity,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line))
    plan.close()


if __name__ == "__main__":
    main()
