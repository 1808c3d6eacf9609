"""rd_sht_synthesis_noalloc: the scalar synthesis with its Fourier scratch in
a fixed device array (no allocation on the call), through the C-ABI, against
rd_sht_synthesis (the same kernels and tables: equal bit for bit).  Two calls
in a row, the second with a smaller input on a smaller plan, so that the
second result cannot lean on what the first left in the shared scratch; a
request beyond the fixed scratch is refused before any launch and leaves the
output untouched."""
import ctypes

import numpy as np
import pytest
import torch

from tests import helpers as H

pytestmark = pytest.mark.gpu


def _spectra(lmax, nplanes, seed):
    rng = np.random.default_rng(seed)
    ell = H.ell_of(lmax)
    c = (rng.standard_normal((nplanes, ell.size)) + 1j * rng.standard_normal((nplanes, ell.size))) * (ell + 1.0) ** -1.5
    c[:, : lmax + 1] = c[:, : lmax + 1].real
    return torch.from_numpy(c).to("cuda:0")


def _both(S, plan, a):
    L = S.lib()
    n = a.shape[0]
    ref = torch.full((n, plan.nphi, plan.nlat), 7.0, dtype=torch.float64, device="cuda:0")
    got = torch.full_like(ref, -3.0)
    h = plan.h
    assert L.rd_sht_synthesis(h, n, ctypes.c_void_p(a.data_ptr()), ctypes.c_void_p(ref.data_ptr())) == 0
    assert L.rd_sht_synthesis_noalloc(h, n, ctypes.c_void_p(a.data_ptr()), ctypes.c_void_p(got.data_ptr())) == 0
    torch.cuda.synchronize()
    return got, ref


def test_noalloc_synthesis_matches_the_allocating_entry_point():
    from redondito_amd import sphp as S
    big = S.FCPlan(127, 192, 384, 40)
    a1 = _spectra(127, 40, 1)
    got1, ref1 = _both(S, big, a1)
    print(f"lmax 127, 40 planes: max |diff| {(got1 - ref1).abs().max().item():.1e}")
    assert torch.equal(got1, ref1)
    small = S.FCPlan(31, 48, 96, 9)
    a2 = _spectra(31, 7, 2)
    got2, ref2 = _both(S, small, a2)
    print(f"lmax 31, 7 planes: max |diff| {(got2 - ref2).abs().max().item():.1e}")
    assert torch.equal(got2, ref2)
    # the FP64 synthesis itself against the oracle, so that equality means parity
    from oracle import oracle as O
    o = O.SHT(31, 48, 96).synth(a2.cpu().numpy())
    assert H.rell2(got2.cpu().numpy(), o) < 1e-13
    big.close()
    small.close()


def test_noalloc_synthesis_rejects_sizes_beyond_its_scratch():
    from redondito_amd import sphp as S
    L = S.lib()
    plan = S.FCPlan(255, 384, 768, 48)
    per_plane = 16 * 256 * 384
    n_ok = S.RD_NOALLOC_SCRATCH_BYTES // per_plane
    assert n_ok < 48
    a = _spectra(255, 48, 3)
    out = torch.full((48, 768, 384), 5.0, dtype=torch.float64, device="cuda:0")
    rc = L.rd_sht_synthesis_noalloc(plan.h, n_ok + 1, ctypes.c_void_p(a.data_ptr()), ctypes.c_void_p(out.data_ptr()))
    torch.cuda.synchronize()
    assert rc == -1 and bool((out == 5.0).all())
    # the largest plane count that fits works and matches
    got, ref = _both(S, plan, a[:n_ok].contiguous())
    assert torch.equal(got, ref)
    plan.close()
