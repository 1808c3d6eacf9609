"""Batching form: uneven scalar shell problems through persistent Legendre
launches, a captured hipGraph and the tile tuner.

* FP64 batch == the single-plan 3D transforms (the same tile arithmetic), and
  the single-plan path is pinned to the oracle by the 3D and SHT tests.
* The round trip forward(inverse(forward(g))) returns forward(g) (band-limited
  in l, FC-Gram-consistent in r) to 1e-12.
* Graph replays equal the eager round trip on every replay, with eager work
  between replays.
* FP32 Legendre stays within 1e-5 of FP64 (the reported precision form).
* Tuning changes the schedule, not the result.
"""
import numpy as np
import pytest
import torch

from tests import helpers as H

pytestmark = pytest.mark.gpu

RI, RO = 0.35, 1.0
PROBLEMS = [(31, 32), (63, 64), (95, 32), (31, 128), (127, 64), (63, 32)]   # (lmax, nr)


def dev():
    return torch.device("cuda:0")


def problem_list(spec=PROBLEMS):
    out = []
    for lmax, nr in spec:
        nlat = 4 * ((3 * (lmax + 1) + 7) // 8)
        out.append((lmax, nlat, 2 * nlat, nr + 25, (RO - RI) / (nr - 1), RI))
    return out


def seeded_spectra(probs, seed):
    rng = np.random.default_rng(seed)
    out = []
    for lmax, nlat, nphi, nz, dz, ri in probs:
        ell = H.ell_of(lmax)
        c = (rng.standard_normal((nz, ell.size)) + 1j * rng.standard_normal((nz, ell.size)))
        c *= (ell + 1.0) ** -1.5
        c[:, : lmax + 1] = c[:, : lmax + 1].real
        out.append(torch.from_numpy(c).to(dev()))
    return out


def grids(probs):
    return [torch.zeros((p[3], p[2], p[1]), dtype=torch.float64, device=dev()) for p in probs]


def specs(probs):
    return [torch.zeros((p[3], (p[0] + 1) * (p[0] + 2) // 2), dtype=torch.complex128, device=dev())
            for p in probs]


def test_batch_matches_single_plans():
    """The round trip of every problem equals the single-plan 3D transforms,
    with every problem on its own chain (the default), every problem in the
    persistent launches, and split between the two."""
    from redondito_amd import sphp as S
    probs = problem_list()
    spec = seeded_spectra(probs, 1)
    refs = []
    for i, p in enumerate(probs):
        plan = S.FCPlan(p[0], p[1], p[2], p[3], 25, 5, p[4], p[5])
        sc = spec[i].clone()
        gr = torch.zeros((p[3], p[2], p[1]), dtype=torch.float64, device=dev())
        S.fftp3d_complex_to_real(plan, sc, gr)
        so = torch.zeros_like(spec[i])
        S.fftp3d_real_to_complex(plan, gr.clone(), so)
        torch.cuda.synchronize()
        refs.append((gr, so))
        plan.close()
    b = S.Batch(probs, prec=S.RD_PREC_FP64)
    print("batch info", b.info())
    assert b.info()["nprob"] == len(probs) and b.get_split() == (0, len(probs))
    mid = sorted(p[0] for p in probs)[len(probs) // 2]
    n_mid = sum(p[0] >= mid for p in probs)
    for split, n_one in ((0, len(probs)), (2**31 - 1, 0), (mid, n_mid)):
        b.split(split)
        assert b.get_split() == (split, n_one)
        assert (b.info()["syn_items"] > 0) == (n_one < len(probs))
        g = grids(probs)
        b.inverse(spec, g)
        out = specs(probs)
        b.forward(g, out)
        torch.cuda.synchronize()
        for i, p in enumerate(probs):
            gr, so = refs[i]
            eg = (g[i] - gr).abs().max().item() / gr.abs().max().item()
            es = (out[i] - so).abs().max().item() / so.abs().max().item()
            print(f"split {split} problem {p[:4]}: grid {eg:.1e} spec {es:.1e}")
            assert eg <= 1e-15 and es <= 1e-15
    b.close()


def test_batch_round_trip_and_graph():
    """Round trip and its graph with the problems split between one-problem
    chains and the persistent launches."""
    from redondito_amd import sphp as S
    probs = problem_list()
    b = S.Batch(probs, prec=S.RD_PREC_FP64)
    b.split(sorted(p[0] for p in probs)[len(probs) // 2])
    assert 0 < b.get_split()[1] < len(probs)
    rng = np.random.default_rng(3)
    g0 = [torch.from_numpy(rng.standard_normal((p[3], p[2], p[1]))).to(dev()) for p in probs]
    s0 = specs(probs)
    b.forward(g0, s0)                      # band-limited, FC-Gram-consistent spectra
    g1, s1 = grids(probs), specs(probs)
    b.inverse(s0, g1)
    b.forward(g1, s1)
    torch.cuda.synchronize()
    for a, r in zip(s1, s0):
        assert H.rell2(a.cpu().numpy(), r.cpu().numpy()) < 1e-12
    # capture the round trip s0 -> g2 -> s2 and replay it with eager work in between
    g2, s2 = grids(probs), specs(probs)
    b.capture(s0, g2, s2)
    assert b.info()["captured"] == 1
    other = seeded_spectra(probs, 9)
    for rep in range(3):
        for t in s2:
            t.zero_()
        b.replay()
        torch.cuda.synchronize()
        for a, r in zip(s2, s1):
            assert torch.equal(a, r), rep
        b.inverse(other, grids(probs))     # eager work on the same batch between replays
        torch.cuda.synchronize()
    b.close()


def test_batch_fp32_form():
    from redondito_amd import sphp as S
    probs = problem_list()
    spec = seeded_spectra(probs, 5)
    res = {}
    for prec in (S.RD_PREC_FP64, S.RD_PREC_FP32):
        b = S.Batch(probs, prec=prec)
        g, out = grids(probs), specs(probs)
        b.inverse(spec, g)
        b.forward(g, out)
        torch.cuda.synchronize()
        res[prec] = ([x.cpu().numpy() for x in g], [x.cpu().numpy() for x in out])
        b.close()
    eg = max(H.rell2(a, r) for a, r in zip(res[1][0], res[0][0]))
    es = max(H.rell2(a, r) for a, r in zip(res[1][1], res[0][1]))
    print(f"fp32 vs fp64: grid {eg:.2e} spectra {es:.2e}")
    assert eg < 1e-5 and es < 1e-5 and eg > 0


def test_batch_tune_keeps_results():
    from redondito_amd import sphp as S
    probs = problem_list()
    b = S.Batch(probs, prec=S.RD_PREC_FP64)
    spec = seeded_spectra(probs, 11)
    g, out = grids(probs), specs(probs)
    b.inverse(spec, g)
    b.forward(g, out)
    torch.cuda.synchronize()
    ref = [x.clone() for x in out]
    ms_fixed, ms_tuned, choice = b.tune(spec, grids(probs), specs(probs), reps=2)
    print(f"tune: fixed {ms_fixed:.3f} ms, tuned {ms_tuned:.3f} ms, choice {choice}, split {b.get_split()}")
    # two separate timings (the threshold pass, then the tile pass): within noise
    assert ms_tuned <= 1.1 * ms_fixed and choice[0] in (2, 4) and 1 <= choice[1] <= 8
    assert b.info()["jb"] == choice[0]
    out2 = specs(probs)
    b.inverse(spec, g)
    b.forward(g, out2)
    torch.cuda.synchronize()
    for a, r in zip(out2, ref):
        assert torch.equal(a, r)
    b.close()


def test_batch_bad_arguments():
    from redondito_amd import sphp as S
    probs = problem_list([(31, 32)])
    with pytest.raises(S.RdError):
        S.Batch(probs, prec=S.RD_PREC_BF16X3)
    b = S.Batch(probs, prec=S.RD_PREC_FP64)
    with pytest.raises(S.RdError):
        b.replay()                         # nothing captured
    with pytest.raises(S.RdError):
        b.config(3, 2)
    b.close()


def test_batch_benchmark_draw_48_problems():
    """The batching workload itself: the 48-problem draw of the benchmark
    (lmax from {31, 63, 95, 127, 191, 255}, nlat = 3(lmax+1)/2, nphi = 2 nlat,
    nr from {32, 64, 128}).  This reaches the batch's nphi = 576 and 768 phi
    paths and the lmax 191 / 255 persistent work items.

    * FP64 batch == single plans, bit for bit, for every problem (grid and
      spectra, sphp_sca.F90:237-270, 400-433 per problem).
    * The single plan of each distinct (lmax, nr) against the oracle: the 3D
      inverse (z inverse, then synthesis per plane) and forward (analysis per
      plane, then continuation and z forward) to 1e-12.
    * The FP32-Legendre batch within 1e-5 of FP64, and its round trip on
      band-limited, FC-Gram-consistent spectra within 1e-5."""
    from bench import batch_problems
    from oracle import oracle as O
    from redondito_amd import sphp as S
    probs = batch_problems()
    assert len(probs) == 48 and max(p[0] for p in probs) == 255
    assert {p[2] for p in probs} >= {576, 768}
    spec = seeded_spectra(probs, 20261204)
    b = S.Batch(probs, prec=S.RD_PREC_FP64)
    g, out = grids(probs), specs(probs)
    b.inverse(spec, g)
    b.forward(g, out)
    torch.cuda.synchronize()
    A, Q = O.load_tables(S.TABLE_DIR, 25, 5)
    dirc = O.fc_dir(A, Q)
    seen = set()
    for i, p in enumerate(probs):
        plan = S.FCPlan(p[0], p[1], p[2], p[3], 25, 5, p[4], p[5])
        gr = torch.zeros_like(g[i])
        S.fftp3d_complex_to_real(plan, spec[i].clone(), gr)
        so = torch.zeros_like(out[i])
        S.fftp3d_real_to_complex(plan, g[i].clone(), so)
        torch.cuda.synchronize()
        assert torch.equal(g[i], gr), f"problem {i} {p[:4]}: grid"
        assert torch.equal(out[i], so), f"problem {i} {p[:4]}: spectra"
        key = (p[0], p[3])
        if key not in seen:
            seen.add(key)
            sht = O.SHT(p[0], p[1], p[2])
            fc = O.FC(p[3], 25, 5, dirc)
            ref_g = sht.synth(fc.inverse(spec[i].cpu().numpy().copy()))
            gh = g[i].cpu().numpy()
            ref_s = fc.forward(sht.anal(gh))
            eg, es = H.rell2(gh, ref_g), H.rell2(out[i].cpu().numpy(), ref_s)
            print(f"lmax={p[0]} nphi={p[2]} nz={p[3]}: inverse {eg:.2e} forward {es:.2e} vs oracle", flush=True)
            assert eg < 1e-12 and es < 1e-12
        plan.close()
    b.close()
    del g
    b32 = S.Batch(probs, prec=S.RD_PREC_FP32)
    g32, o32 = grids(probs), specs(probs)
    b32.inverse(spec, g32)
    b32.forward(g32, o32)
    torch.cuda.synchronize()
    e32 = max(H.rell2(a.cpu().numpy(), r.cpu().numpy()) for a, r in zip(o32, out))
    # round trip of band-limited, FC-Gram-consistent spectra (the FP64 forward output)
    b32.inverse(out, g32)
    b32.forward(g32, o32)
    torch.cuda.synchronize()
    rt = max(H.rell2(a.cpu().numpy(), r.cpu().numpy()) for a, r in zip(o32, out))
    print(f"fp32 batch vs fp64 {e32:.2e}; fp32 round trip {rt:.2e}")
    assert 0 < e32 < 1e-5 and 0 < rt < 1e-5
    b32.close()


# ------------------------------------------------ nonlinear-term form
UXW_PROBLEMS = [(31, 24), (63, 40), (47, 128), (95, 32), (31, 64), (127, 96)]   # (lmax, nr)


def _uxw_inputs(probs, seed):
    out = []
    for i, (lmax, nlat, nphi, nz, dz, ri) in enumerate(probs):
        W, Z = H.shell_fields(lmax, nz - 25, nz, ri, RO, seed=seed + i, slope=2.0)
        out.append((torch.from_numpy(W).to(dev()), torch.from_numpy(Z).to(dev())))
    return out


def test_batch_nonlinear_uxw_matches_single_plans():
    """The nonlinear-term form of the batch (u x omega of every problem: the
    problems below the degree threshold share one persistent Legendre launch
    per direction, those at or above it run their one-problem evaluation)
    equals rd_nonlinear_uxw of each problem's own plan bit for bit (same tile
    arithmetic, same accumulation order), in FP64 and in the FP32 form, with
    every problem persistent, every problem on its own, and split between
    the two; its hipGraph replays equal the eager result, with eager work in
    between; the tuner changes the schedule, not the result."""
    from redondito_amd import sphp as S
    probs = problem_list(UXW_PROBLEMS)
    ins = _uxw_inputs(probs, 20261205)
    W = [w for w, _ in ins]
    Z = [z for _, z in ins]
    for prec in (S.RD_PREC_FP64, S.RD_PREC_FP32):
        refs = []
        for i, p in enumerate(probs):
            plan = S.FCPlan(p[0], p[1], p[2], p[3], 25, 5, p[4], p[5])
            ref = [torch.full_like(W[i], 3.0) for _ in range(3)]
            S.nonlinear_uxw(plan, W[i], Z[i], *ref, prec=prec)
            torch.cuda.synchronize()
            refs.append(ref)
            plan.close()
        b = S.Batch(probs, prec=prec, form="uxw")
        assert b.info()["jb"] == 1 and b.get_split() == (0, len(probs))
        # all on their own (the default), none, and lmax >= 63 (3 of 6)
        for split, n_one in ((0, len(probs)), (2**31 - 1, 0), (63, 3)):
            b.split(split)
            assert b.get_split() == (split, n_one)
            # the persistent launches hold work items only for the others
            assert (b.info()["syn_items"] > 0) == (n_one < len(probs))
            out = [specs(probs) for _ in range(3)]
            b.nonlinear_uxw(W, Z, *out)
            torch.cuda.synchronize()
            for i, p in enumerate(probs):
                d = max((a[i] - r).abs().max().item() for a, r in zip(out, refs[i]))
                print(f"prec {prec} split {split} problem {p[:4]}: max |batch - single| {d:.1e}")
                assert all(torch.equal(a[i], r) for a, r in zip(out, refs[i])), (prec, split, p[:4], d)
        # graph of the split schedule: replays equal the eager result, with
        # eager work between them
        g = [specs(probs) for _ in range(3)]
        b.capture_uxw(W, Z, *g)
        assert b.info()["captured"] == 1
        other = [specs(probs) for _ in range(3)]
        for rep in range(3):
            for t in g:
                for x in t:
                    x.zero_()
            b.replay()
            torch.cuda.synchronize()
            assert all(torch.equal(x[i], y) for i in range(len(probs)) for x, y in zip(g, refs[i])), rep
            b.nonlinear_uxw(Z, W, *other)          # eager work on other arrays
            torch.cuda.synchronize()
        fixed, tuned, choice = b.tune_uxw(W, Z, *other, reps=2)
        split = b.get_split()
        print(f"prec {prec}: tune fixed {fixed:.3f} ms tuned {tuned:.3f} ms choice {choice} split {split}")
        assert choice[0] in (1, 2) and split[0] in (2**31 - 1, 0) + tuple(p[0] for p in probs)
        t2 = [specs(probs) for _ in range(3)]
        b.nonlinear_uxw(W, Z, *t2)
        torch.cuda.synchronize()
        assert all(torch.equal(x[i], y) for i in range(len(probs)) for x, y in zip(t2, refs[i]))
        with pytest.raises(S.RdError):
            b.inverse(W, grids(probs))             # the round trip is not this form's
        b.close()
