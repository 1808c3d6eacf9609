"""GPU parity: the HIP path through the C-ABI against the CPU oracle.

Tolerances: the FP64 path must match the oracle to 1e-12 relative L2 (the
bound set for the plain form); known-answer fields of the reference's tests
are held to the levels LEEME.md records for its CPU path.
"""
import ctypes

import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu

TOL = 1e-12
TABLES = None


def dev():
    return torch.device("cuda:0")


def sphp():
    from redondito_amd import sphp as S
    return S


def tables_dir():
    import redondito_amd.sphp as S
    return S.TABLE_DIR


def to_dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev())


def to_host(x):
    torch.cuda.synchronize()
    return x.cpu().numpy()


def fc_dir():
    A, Q = O.load_tables(tables_dir(), 25, 5)
    return O.fc_dir(A, Q)


# ------------------------------------------------------------------ grid
@pytest.mark.parametrize("lmax,nlat,nphi", [(15, 24, 48), (31, 48, 96), (127, 192, 384), (255, 384, 768)])
def test_grid_matches_oracle(lmax, nlat, nphi):
    S = sphp()
    plan = S.FCPlan(lmax, nlat, nphi, 4)
    ct, st, wt, z, kz, ell, emm = plan.grid()
    oc = O.SHT(lmax, nlat, nphi)
    c2, s2, w2 = oc.grid()
    assert np.max(np.abs(ct - c2)) < 1e-15
    assert np.max(np.abs(st - s2) / s2) < 1e-15
    assert np.max(np.abs(wt - w2) / w2) < 1e-13
    assert np.array_equal(ell, H.ell_of(lmax)) and np.array_equal(emm, H.emm_of(lmax))


# ---------------------------------------------------------- scalar SHT
def test_prueba_sphp_known_fields():
    """The nine closed-form fields of src/prueba_sphp.f90:121-184, plane k scaled by k."""
    S = sphp()
    lmax, nz = 15, 6
    nlat, nphi = 3 * (lmax + 1) // 2, 3 * (lmax + 1)
    plan = S.FCPlan(lmax, nlat, nphi, nz)
    ct, st = plan.grid()[:2]
    cth, sth, ph = H.grid_arrays(ct, st, nphi)
    k = np.arange(1, nz + 1)[:, None, None]
    for ic in range(1, 10):
        f, a = H.scalar_case(ic, lmax, cth, sth, ph)
        fin = to_dev(k * f[None])
        b = torch.zeros((nz, plan.nlm), dtype=torch.complex128, device=dev())
        S.fftp2d_real_to_complex_xy(plan, fin, b)
        ea = np.max(np.abs(to_host(b) - k[:, :, 0] * a[None]))
        g = torch.zeros_like(fin)
        S.fftp2d_complex_to_real_xy(plan, to_dev(k[:, :, 0] * a[None]), g)
        es = np.max(np.abs(to_host(g) - k * f[None]))
        print(f"case {ic}: analysis {ea:.2e} synthesis {es:.2e}")
        assert ea < 1e-12 and es < 1e-12


@pytest.mark.parametrize("lmax,nlat,nphi,npl", [(15, 24, 48, 6), (31, 48, 96, 24), (127, 192, 384, 5),
                                                (255, 384, 768, 3),
                                                # nphi = 768 kernels with m > lmax zero and a last
                                                # tile of two theta rows
                                                (200, 386, 768, 3)])
def test_scalar_sht_vs_oracle(lmax, nlat, nphi, npl):
    S = sphp()
    plan = S.FCPlan(lmax, nlat, nphi, npl)
    oc = O.SHT(lmax, nlat, nphi)
    a = H.random_spectrum(lmax, npl, seed=11 + lmax, slope=1.5)
    f_ref = oc.synth(a)
    g = torch.zeros((npl, nphi, nlat), dtype=torch.float64, device=dev())
    S.fftp2d_complex_to_real_xy(plan, to_dev(a), g)
    e1 = H.rell2(to_host(g), f_ref)
    b = torch.zeros((npl, plan.nlm), dtype=torch.complex128, device=dev())
    S.fftp2d_real_to_complex_xy(plan, to_dev(f_ref), b)
    b_ref = oc.anal(f_ref)
    e2 = H.rell2(to_host(b), b_ref)
    e3 = H.rell2(to_host(b), a)
    print(f"synthesis {e1:.2e} analysis {e2:.2e} round trip {e3:.2e}")
    assert e1 < TOL and e2 < TOL and e3 < TOL


# ---------------------------------------------------------- vector SHT
def test_prueba_vec_known_fields():
    """The eight vector fields of src/prueba_vec.f90:139-209."""
    S = sphp()
    lmax, nz = 15, 6
    nlat, nphi = 3 * (lmax + 1) // 2, 3 * (lmax + 1)
    plan = S.FCPlan(lmax, nlat, nphi, nz)
    ct, st = plan.grid()[:2]
    cth, sth, ph = H.grid_arrays(ct, st, nphi)
    k = np.arange(1, nz + 1)[:, None, None]
    for ic in range(1, 9):
        fs, cs = H.vector_case(ic, lmax, cth, sth, ph)
        vin = [to_dev(k * f[None]) for f in fs]
        qst = [torch.zeros((nz, plan.nlm), dtype=torch.complex128, device=dev()) for _ in range(3)]
        S.fftp2d_vec_to_qst(plan, *vin, *qst)
        ea = max(np.max(np.abs(to_host(x) - k[:, :, 0] * c[None])) for x, c in zip(qst, cs))
        vout = [torch.zeros_like(vin[0]) for _ in range(3)]
        S.fftp2d_qst_to_vec(plan, *[to_dev(k[:, :, 0] * c[None]) for c in cs], *vout)
        es = max(np.max(np.abs(to_host(v) - k * f[None])) for v, f in zip(vout, fs))
        print(f"case {ic}: analysis {ea:.2e} synthesis {es:.2e}")
        assert ea < 1e-12 and es < 5e-12


@pytest.mark.parametrize("lmax,nlat,nphi,npl", [(15, 24, 48, 6), (31, 48, 96, 8), (127, 192, 384, 4),
                                                (255, 384, 768, 2)])
def test_vector_sht_vs_oracle(lmax, nlat, nphi, npl):
    S = sphp()
    plan = S.FCPlan(lmax, nlat, nphi, npl)
    oc = O.SHT(lmax, nlat, nphi)
    q = H.random_spectrum(lmax, npl, seed=3, slope=1.0)
    s = H.random_spectrum(lmax, npl, seed=4, slope=1.0, zero_l0=True)
    t = H.random_spectrum(lmax, npl, seed=5, slope=1.0, zero_l0=True)
    v_ref = oc.vsynth(q, s, t)
    vout = [torch.zeros((npl, nphi, nlat), dtype=torch.float64, device=dev()) for _ in range(3)]
    S.fftp2d_qst_to_vec(plan, to_dev(q), to_dev(s), to_dev(t), *vout)
    e1 = max(H.rell2(to_host(v), r) for v, r in zip(vout, v_ref))
    qst = [torch.zeros((npl, plan.nlm), dtype=torch.complex128, device=dev()) for _ in range(3)]
    S.fftp2d_vec_to_qst(plan, *[to_dev(v) for v in v_ref], *qst)
    a_ref = oc.vanal(*v_ref)
    e2 = max(H.rell2(to_host(x), r) for x, r in zip(qst, a_ref))
    e3 = max(H.rell2(to_host(x), r) for x, r in zip(qst, (q, s, t)))
    print(f"synthesis {e1:.2e} analysis {e2:.2e} round trip {e3:.2e}")
    assert e1 < TOL and e2 < TOL and e3 < TOL


# --------------------------------------------------------------- radial
def test_prueba_z_periodic():
    """src/prueba_z.f90:45-75: c cos(2 pi q z/L) + d sin(2 pi p z/L), nz = 16."""
    S = sphp()
    lmax, nz, q, p = 7, 16, 3, 5
    nlat, nphi = 12, 24
    plan = S.FCPlan(lmax, nlat, nphi, nz)
    nlm = plan.nlm
    lm = np.arange(1, nlm + 1)
    cl = lm / nlm - 0.5j
    dl = 0.3 + 1j * (1.0 - lm / nlm)
    j = np.arange(nz)[:, None]
    f0 = cl[None] * np.cos(2 * np.pi * q * j / nz) + dl[None] * np.sin(2 * np.pi * p * j / nz)
    e = np.zeros((nz, nlm), complex)
    e[q] = cl / 2; e[nz - q] = cl / 2; e[p] = -1j * dl / 2; e[nz - p] = 1j * dl / 2
    f = to_dev(f0)
    S.fftp1d_real_to_complex_z(plan, f)
    e1 = np.max(np.abs(to_host(f) - e))
    S.fftp1d_complex_to_real_z(plan, f)
    e2 = np.max(np.abs(to_host(f) - f0))
    print(f"coefficients {e1:.2e} round trip {e2:.2e}")
    assert e1 < 1e-15 and e2 < 1e-14


# radial lengths 49, 64, 121, 153, 256; the primes 89 and 281 (Rader form:
# 88 = 8 x 11 in LDS Stockham stages, 280 = 14 x 20 in register passes) and 47
# (46 = 2 x 23, direct prime stage)
@pytest.mark.parametrize("nr,lmax", [(24, 31), (39, 7), (96, 15), (128, 15), (231, 7), (64, 15), (256, 7),
                                     (22, 15)])
def test_z_fcgram_vs_oracle(nr, lmax):
    S = sphp()
    C, d = 25, 5
    n = nr + C
    dz = 1.0 / (nr - 1)
    nlat = S.grid_nlat(lmax)
    plan = S.FCPlan(lmax, nlat, 2 * nlat, n, C, d, dz, 0.0)
    fc = O.FC(n, C, d, fc_dir())
    rng = np.random.default_rng(nr)
    a = rng.standard_normal((n, plan.nlm)) + 1j * rng.standard_normal((n, plan.nlm))
    ref = fc.forward(a.copy())
    x = to_dev(a)
    S.fftp1d_real_to_complex_z(plan, x)
    e1 = H.rell2(to_host(x), ref)
    ref2 = fc.inverse(ref.copy())
    S.fftp1d_complex_to_real_z(plan, x)
    e2 = H.rell2(to_host(x), ref2)
    d1 = torch.zeros_like(x)
    S.derivr(plan, to_dev(a), d1, 1, "r")
    kz = O.kz(n, dz)
    e3 = H.rell2(to_host(d1)[:nr], fc.derivr_r(a.copy(), kz, 1)[:nr])
    d2 = torch.zeros_like(x)
    S.derivr(plan, to_dev(a), d2, 2, "r")
    e4 = H.rell2(to_host(d2)[:nr], fc.derivr_r(a.copy(), kz, 2)[:nr])
    print(f"n={n}: forward {e1:.2e} inverse {e2:.2e} derivr {e3:.2e} derivr2 {e4:.2e}")
    assert e1 < TOL and e2 < TOL and e3 < TOL and e4 < TOL


@pytest.mark.parametrize("nz", [64, 128, 256])
def test_prueba_z_fcgram_convergence(nz):
    """src/prueba_z.f90:79-125: d/dz of sin(2.3z+0.4)+z^2 on [0,1], Nyquist zeroed
    by the test's own multiplier; LEEME.md:34 reports 2.5e-6, 4.9e-8, 2.2e-9."""
    S = sphp()
    C, d, lmax = 25, 5, 7
    dz = 1.0 / (nz - C - 1)
    plan = S.FCPlan(lmax, 12, 24, nz, C, d, dz, 0.0)
    nlm = plan.nlm
    lm = np.arange(1, nlm + 1)
    cl = lm / nlm - 0.5j
    zz = np.arange(nz - C) * dz
    f = np.zeros((nz, nlm), complex)
    f[: nz - C] = cl[None] * (np.sin(2.3 * zz + 0.4) + zz ** 2)[:, None]
    e = cl[None] * (2.3 * np.cos(2.3 * zz + 0.4) + 2 * zz)[:, None]
    x = to_dev(f)
    S.fftp1d_real_to_complex_z(plan, x)
    kn = np.arange(nz, dtype=float)
    kn[kn > nz / 2] -= nz
    kn[nz // 2] = 0
    mult = to_dev((1j * 2 * np.pi * kn / (nz * dz))[:, None] * np.ones((1, nlm)))
    x.mul_(mult)
    S.fftp1d_complex_to_real_z(plan, x)
    err = np.max(np.abs(to_host(x)[: nz - C] - e)) / np.max(np.abs(e))
    published = {64: 2.5e-6, 128: 4.9e-8, 256: 2.2e-9}[nz]
    print(f"nz={nz}: {err:.2e} (published {published:.1e})")
    assert err < 2 * published


# ------------------------------------------------------------ operators
def _pseudo_setup(lmax=7, nz=128, rint=0.5, rext=1.5):
    S = sphp()
    C, d = 25, 5
    dz = (rext - rint) / (nz - C - 1)
    nlat = 4 * ((3 * (lmax + 1) + 7) // 8)
    plan = S.FCPlan(lmax, nlat, 2 * nlat, nz, C, d, dz, rint)
    return S, plan, dz


def test_prueba_pseudo_operators_vs_oracle():
    """src/prueba_pseudo.f90:42-115 on f = c_lm e^r sin 3r, r in [0.5, 1.5]."""
    S, plan, dz = _pseudo_setup()
    nz, nlm, lmax = plan.nz, plan.nlm, plan.lmax
    n_phys = nz - plan.C
    z = plan.grid()[3]
    ell = H.ell_of(lmax)
    lm = np.arange(1, nlm + 1)
    cl = (1.0 + 0.1 * lm) + 1j * (0.5 - 0.05 * lm)
    g = lambda r, p: np.exp(r) * (np.sin(3 * r) if p == 0 else (np.sin(3 * r) + 3 * np.cos(3 * r)) if p == 1
                                  else (-8 * np.sin(3 * r) + 6 * np.cos(3 * r)))
    a = cl[None] * g(z, 0)[:, None]
    fc = O.FC(nz, 25, 5, fc_dir())
    kz = O.kz(nz, dz)
    A = to_dev(a)
    B = torch.zeros_like(A)
    rel = lambda x, y: np.max(np.abs(x[:n_phys] - y[:n_phys])) / np.max(np.abs(y[:n_phys]))
    out = {}
    S.derivr(plan, A, B, 1, "r")
    d1 = to_host(B)
    out["derivr1_vs_oracle"] = H.rell2(d1, fc.derivr_r(a, kz, 1))
    out["derivr1_exact"] = rel(d1, cl[None] * g(z, 1)[:, None])
    S.derivr(plan, A, B, 2, "r")
    d2 = to_host(B)
    out["derivr2_vs_oracle"] = H.rell2(d2, fc.derivr_r(a, kz, 2))
    out["derivr2_exact"] = rel(d2, cl[None] * g(z, 2)[:, None])
    Cc = A.clone()
    S.fftp1d_real_to_complex_z(plan, Cc)
    E = torch.zeros_like(A)
    S.derivr(plan, Cc, E, 1, "k")
    S.fftp1d_complex_to_real_z(plan, E)
    out["derivr_k_vs_r"] = rel(to_host(E), d1)
    S.laplang(plan, A, B)
    out["laplang"] = rel(to_host(B), -(ell * (ell + 1.0))[None] * a)
    S.multr(plan, A, B, -2)
    out["multr"] = rel(to_host(B), a / (z ** 2)[:, None])
    S.dlr(plan, A, B)
    dl = to_host(B)
    out["dlr_vs_oracle"] = H.rell2(dl, fc.dlr(a, kz, z, ell))
    out["dlr_exact"] = rel(dl, cl[None] * g(z, 2)[:, None] - (ell * (ell + 1.0))[None] * a / (z ** 2)[:, None])
    Cz = torch.zeros_like(A)
    S.saxpby_c(plan, Cz, A, 2.0, B, -3.0)
    out["saxpby"] = rel(to_host(Cz), 2 * a - 3 * dl)
    S.copy(plan, B, Cz)
    out["copy"] = rel(to_host(Cz), dl)
    Cc = A.clone()
    S.fc_filter(plan, Cc, "r")
    out["fc_filter_change"] = rel(to_host(Cc), a)
    # FP64 conditioning of the FC-Gram operators at this input (the continuation
    # matrix has entries ~4e3): GPU and oracle may differ by a few times this
    floor = H.noise_floor(lambda x: [fc.derivr_r(x, kz, 1), fc.derivr_r(x, kz, 2), fc.dlr(x, kz, z, ell)], [a])
    for k, v in out.items():
        print(f"{k:22s} {v:.2e}")
    print("noise floor (d1, d2, dlr):", ", ".join(f"{f:.2e}" for f in floor))
    tol = [max(TOL, 30 * f) for f in floor]
    assert out["derivr1_vs_oracle"] < tol[0] and out["derivr2_vs_oracle"] < tol[1] and out["dlr_vs_oracle"] < tol[2]
    assert out["derivr1_exact"] < 2 * 2.3e-7 and out["derivr2_exact"] < 2 * 3.6e-5 and out["dlr_exact"] < 2 * 2.8e-6
    assert out["derivr_k_vs_r"] < 1e-13 and out["laplang"] < 1e-15 and out["multr"] < 1e-15
    assert out["saxpby"] < 1e-15 and out["copy"] == 0.0
    # LEEME.md:45 publishes a change of 1.7e-11 for this field; a no-op filter
    # (a round trip, ~5e-16) fails the lower bound
    assert 0.3 * 1.7e-11 < out["fc_filter_change"] < 3 * 1.7e-11


@pytest.mark.parametrize("nz", [128, 153])
def test_fc_filter_vs_oracle(nz):
    """fc_filter (pseudospec_hd.f90:369-467) in both spaces against the oracle,
    and its multiplier against the 40-digit closed form: exactly 1 at kz = 0,
    1e-16 at the even-n Nyquist bin."""
    S, plan, dz = _pseudo_setup(nz=nz)
    nlm = plan.nlm
    rng = np.random.default_rng(nz)
    a = rng.standard_normal((nz, nlm)) + 1j * rng.standard_normal((nz, nlm))
    kzv = O.kz(nz, dz)
    assert np.max(np.abs(plan.grid()[4] - kzv)) <= 1e-15 * np.max(np.abs(kzv))
    dkz = 2 * np.pi / (dz * nz)
    A = to_dev(a)
    S.fc_filter(plan, A, "k")
    ref = O.fc_filter_k(a.copy(), kzv, dkz)
    e_k = H.rell2(to_host(A), ref)
    ones = to_dev(np.ones((nz, nlm), dtype=np.complex128))
    S.fc_filter(plan, ones, "k")
    mult = to_host(ones)
    assert np.all(mult == mult[:, :1]) and np.all(mult.imag == 0)
    assert H.filter_check(mult[:, 0].real, nz, dz)
    if nz % 2 == 0:
        assert abs(mult[nz // 2, 0].real - 1e-16) < 1e-29
    # physical space: forward FC-Gram transform, filter, inverse
    n_phys = nz - plan.C
    fc = O.FC(nz, 25, 5, fc_dir())
    B = to_dev(a)
    S.fc_filter(plan, B, "r")
    filt_r = lambda x: fc.inverse(O.fc_filter_k(fc.forward(x.copy()), kzv, dkz))[:n_phys]
    ref_r = filt_r(a)
    e_r = H.rell2(to_host(B)[:n_phys], ref_r)
    # random columns through the FC-Gram continuation (entries ~4e3): the
    # r-space result is held to its FP64 conditioning, as derivr is
    floor = H.noise_floor(filt_r, [a])
    print(f"nz={nz}: k-space {e_k:.2e}, r-space {e_r:.2e} (noise floor {floor:.2e})")
    assert e_k < 1e-14 and e_r < max(TOL, 30 * floor)


@pytest.mark.parametrize("boun", [5, 6])
def test_neumann_and_robin_reconstruct_vs_oracle(boun):
    """neumann_reconstruct (orders 1, 2) and robin_reconstruct along z
    (src/sphp/fcgram_mod.f90:374-650) on the device against the oracle, and
    the wall value of a degree-4 radial polynomial recovered from its outward
    normal derivative."""
    S, plan, dz = _pseudo_setup(lmax=15, nz=65)
    nz, nlm, C, d = plan.nz, plan.nlm, 25, 5
    nr = nz - C
    z = plan.grid()[3]
    rng = np.random.default_rng(boun)
    c = rng.standard_normal((5, nlm)) + 1j * rng.standard_normal((5, nlm))
    f = sum(c[i][None] * (z[:, None] ** i) for i in range(5))
    rb = 0 if boun == 5 else nr - 1
    for ord_ in (1, 2):
        neu = S.load_neumann_tables(tables_dir(), d, ord_, dz)
        dc = [c[i] * np.prod(np.arange(i - ord_ + 1, i + 1)) if i >= ord_ else 0 * c[i] for i in range(5)]
        g = sum(dc[i] * z[rb] ** (i - ord_) for i in range(ord_, 5))
        if ord_ == 1 and boun == 5:
            g = -g
        a = f.copy()
        a[rb] = g
        A = to_dev(a)
        S.neumann_reconstruct(plan, A, boun, ord_, neu)
        got = to_host(A)
        ref = O.fc_reconstruct(a, C, d, boun, neu)
        assert np.array_equal(np.delete(got, rb, 0), np.delete(a, rb, 0))
        e_orc = np.max(np.abs(got[rb] - ref[rb])) / np.max(np.abs(ref[rb]))
        e_poly = np.max(np.abs(got[rb] - f[rb])) / np.max(np.abs(f[rb]))
        print(f"boun {boun} ord {ord_}: vs oracle {e_orc:.1e}, vs polynomial {e_poly:.1e}")
        assert e_orc < 1e-15 and e_poly < 1e-12
    neu = S.load_neumann_tables(tables_dir(), d, 1, dz)
    coef = rng.uniform(0.5, 2.0, nlm)
    fp = sum(c[i] * i * z[rb] ** (i - 1) for i in range(1, 5)) * (-1.0 if boun == 5 else 1.0)
    a = f.copy()
    a[rb] = fp + coef * f[rb]
    A = to_dev(a)
    S.robin_reconstruct(plan, A, boun, neu, to_dev(coef))
    got = to_host(A)
    ref = O.fc_reconstruct(a, C, d, boun, neu, coef)
    assert np.max(np.abs(got[rb] - ref[rb])) / np.max(np.abs(ref[rb])) < 1e-15
    assert np.max(np.abs(got[rb] - f[rb])) / np.max(np.abs(f[rb])) < 1e-12
    with pytest.raises(S.RdError):
        S.neumann_reconstruct(plan, A, 3, 1, neu)


def test_redist_is_the_one_rank_copy():
    """fftp3d_redist_k_to_lm / _lm_to_k (src/sphp/sphp_sca.F90:629-760): with one
    task the exchange is a copy of the (nlm, planes) block, both directions,
    for a full and a partial plane count; out == in is a no-op; a plane count
    beyond nz is an error."""
    S, plan, dz = _pseudo_setup(lmax=15, nz=40)
    nz, nlm = plan.nz, plan.nlm
    rng = np.random.default_rng(4)
    a = rng.standard_normal((nz, nlm)) + 1j * rng.standard_normal((nz, nlm))
    A = to_dev(a)
    for fn in (S.fftp3d_redist_k_to_lm, S.fftp3d_redist_lm_to_k):
        B = torch.full_like(A, 5.0)
        fn(plan, A, B)
        assert np.array_equal(to_host(B), a)
        part = A[:7].clone()
        Bp = torch.full_like(part, 5.0)
        fn(plan, part, Bp)
        assert np.array_equal(to_host(Bp), a[:7])
        fn(plan, A, A)
        assert np.array_equal(to_host(A), a)
    L = S.lib()
    assert L.rd_redist_k_to_lm(plan.h, nz + 1, ctypes.c_void_p(A.data_ptr()), ctypes.c_void_p(A.data_ptr())) != 0


def test_saxpby_r_vs_numpy():
    """saxpby_r (pseudospec_hd.f90:329-366): z = a x + b y on the grid, z may
    alias x or y (the update is pointwise)."""
    S, plan, dz = _pseudo_setup(lmax=15, nz=40)
    shp = (plan.nz, plan.nphi, plan.nlat)
    rng = np.random.default_rng(6)
    x, y = rng.standard_normal(shp), rng.standard_normal(shp)
    X, Y = to_dev(x), to_dev(y)
    Z = torch.zeros_like(X)
    S.saxpby_r(plan, Z, X, 1.25, Y, -0.5)
    ref = 1.25 * x - 0.5 * y
    assert np.max(np.abs(to_host(Z) - ref)) <= 2.3e-16 * np.max(np.abs(ref))
    S.saxpby_r(plan, X, X, 2.0, Y, 3.0)          # z = x
    assert np.max(np.abs(to_host(X) - (2 * x + 3 * y))) <= 4.5e-16 * np.max(np.abs(2 * x + 3 * y))
    Xp, Yp = to_dev(x[:5]), to_dev(y[:5])
    Zp = torch.zeros_like(Xp)
    S.saxpby_r(plan, Zp, Xp, 1.0, Yp, 1.0)       # a partial plane range
    assert np.max(np.abs(to_host(Zp) - (x[:5] + y[:5]))) <= 4.5e-16 * np.max(np.abs(x[:5] + y[:5]))


def test_prodvec_vs_hand_cross():
    """prodvec against the hand-written cross product (src/prueba_pseudo.f90:98-115);
    planes past nz - C are left untouched."""
    S, plan, dz = _pseudo_setup(lmax=7, nz=40)
    nlat, nphi, nz = plan.nlat, plan.nphi, plan.nz
    ct, st, wt, z = plan.grid()[:4]
    th = np.arccos(ct)
    phi = 2 * np.pi * np.arange(nphi) / nphi
    T = np.broadcast_to(th[None, None, :], (nz, nphi, nlat))
    Ph = np.broadcast_to(phi[None, :, None], (nz, nphi, nlat))
    Z = np.broadcast_to(z[:, None, None], (nz, nphi, nlat))
    u = [np.sin(T) * np.cos(Ph), Z + 0 * T, np.cos(T)]
    v = [Z ** 2 + 0 * T, np.sin(Ph) + 0 * T, np.sin(T) * np.cos(T)]
    U = [to_dev(x.copy()) for x in u]
    V = [to_dev(x.copy()) for x in v]
    Wt = [torch.full((nz, nphi, nlat), 7.0, dtype=torch.float64, device=dev()) for _ in range(3)]
    S.prodvec(plan, *U, *V, *Wt)
    ref = [u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]]
    npk = nz - plan.C
    err = max(np.max(np.abs(to_host(w)[:npk] - r[:npk])) for w, r in zip(Wt, ref))
    untouched = all(np.all(to_host(w)[npk:] == 7.0) for w in Wt)
    print(f"prodvec {err:.2e}")
    assert err < 1e-15 and untouched


# ----------------------------------------------------------------- 3D
def test_prueba_3d_periodic():
    """src/prueba_3d.f90:45-78: five exact (lm, kz) entries, nz = 8."""
    S = sphp()
    lmax, nz, q, p = 15, 8, 3, 1
    nlat, nphi = 24, 48
    plan = S.FCPlan(lmax, nlat, nphi, nz)
    ct, st = plan.grid()[:2]
    cth, sth, ph = H.grid_arrays(ct, st, nphi)
    k = np.arange(nz)[:, None, None]
    f0 = sth[None] * np.cos(ph[None]) * np.cos(2 * np.pi * q * k / nz) + cth[None] * np.sin(2 * np.pi * p * k / nz) + 1.0
    L = lambda l, m: H.lm_index(lmax, l, m)
    e = np.zeros((nz, plan.nlm), complex)
    e[q, L(1, 1)] = np.sqrt(2 * np.pi / 3) / 2
    e[nz - q, L(1, 1)] = np.sqrt(2 * np.pi / 3) / 2
    e[p, L(1, 0)] = -1j * np.sqrt(4 * np.pi / 3) / 2
    e[nz - p, L(1, 0)] = 1j * np.sqrt(4 * np.pi / 3) / 2
    e[0, L(0, 0)] = np.sqrt(4 * np.pi)
    f = to_dev(f0)
    b = torch.zeros((nz, plan.nlm), dtype=torch.complex128, device=dev())
    S.fftp3d_real_to_complex(plan, f, b)
    e1 = np.max(np.abs(to_host(b) - e))
    g = torch.zeros_like(f)
    S.fftp3d_complex_to_real(plan, b, g)
    e2 = np.max(np.abs(to_host(g) - f0))
    print(f"coefficients {e1:.2e} round trip {e2:.2e}")
    assert e1 < 1e-14 and e2 < 1e-13


# -------------------------------------------------------- nonlinear term
@pytest.mark.parametrize("lmax,nr,grid", [(15, 24, None), (31, 24, None), (63, 40, None), (127, 96, None),
                                          (255, 12, None),
                                          # radial length 153 = 17 x 9 (the nr = 128 workload)
                                          (47, 128, None),
                                          # nphi = 768 kernel with m > lmax zero, rows not a multiple of 8
                                          (200, 11, (386, 768)),
                                          # nphi = 1536 kernel: second 256-block of m and its mirrors
                                          (300, 12, (452, 1536))])
def test_nonlinear_uxw_vs_oracle(lmax, nr, grid):
    S = sphp()
    C, d, rint, rext = 25, 5, 0.35, 1.0
    n = nr + C
    dz = (rext - rint) / (nr - 1)
    nlat = 4 * ((3 * (lmax + 1) + 7) // 8)
    nphi = 2 * nlat
    if grid is not None:
        nlat, nphi = grid
    plan = S.FCPlan(lmax, nlat, nphi, n, C, d, dz, rint)
    W, Z = H.shell_fields(lmax, nr, n, rint, rext, seed=20261203 + lmax, slope=2.0)
    oc = O.SHT(lmax, nlat, nphi)
    fc = O.FC(n, C, d, fc_dir())
    ref = O.nonlinear_uxw(oc, fc, rint, dz, W, Z)
    Wd, Zd = to_dev(W), to_dev(Z)
    out = [torch.full_like(Wd, 3.0) for _ in range(3)]
    S.nonlinear_uxw(plan, Wd, Zd, *out)
    got = [to_host(o) for o in out]
    errs = [H.rell2(g[:nr], r[:nr]) for g, r in zip(got, ref)]
    joint = H.rell2_joint([g[:nr] for g in got], [r[:nr] for r in ref])
    floor = H.noise_floor(lambda w, zz: [x[:nr] for x in O.nonlinear_uxw(oc, fc, rint, dz, w, zz)], [W, Z])
    tail = max(float(np.max(np.abs(g[nr:]))) for g in got)
    assert np.array_equal(to_host(Wd), W) and np.array_equal(to_host(Zd), Z)
    print(f"lmax={lmax} nr={nr}: joint {joint:.2e}; NQ {errs[0]:.2e} NS {errs[1]:.2e} NT {errs[2]:.2e}; "
          f"noise floor NQ {floor[0]:.2e} NS {floor[1]:.2e} NT {floor[2]:.2e}")
    assert joint < TOL and tail == 0.0
    bounds = H.field_bounds([r[:nr] for r in ref], floor, TOL)
    assert all(e < bd for e, bd in zip(errs, bounds)), bounds


def test_nonlinear_solid_body_rotation():
    """u = Omega z x r (Z_10 = Omega r^2 sqrt(4 pi/3)), omega = 2 Omega z:
    u x omega = grad(Omega^2 r^2 sin^2 theta), i.e. NQ_00, NQ_20, NS_20 only."""
    S = sphp()
    lmax, nr, C, d, rint, rext = 15, 64, 25, 5, 0.35, 1.0
    n = nr + C
    dz = (rext - rint) / (nr - 1)
    nlat = 4 * ((3 * (lmax + 1) + 7) // 8)
    plan = S.FCPlan(lmax, nlat, 2 * nlat, n, C, d, dz, rint)
    r = rint + dz * np.arange(nr)
    Om = 0.7
    W = np.zeros((n, plan.nlm), complex)
    Z = np.zeros_like(W)
    Z[:nr, H.lm_index(lmax, 1, 0)] = Om * r ** 2 * np.sqrt(4 * np.pi / 3)
    NQ, NS, NT = (torch.zeros((n, plan.nlm), dtype=torch.complex128, device=dev()) for _ in range(3))
    S.nonlinear_uxw(plan, to_dev(W), to_dev(Z), NQ, NS, NT)
    eQ = np.zeros((nr, plan.nlm), complex)
    eS = np.zeros_like(eQ)
    eQ[:, H.lm_index(lmax, 0, 0)] = 2 * Om ** 2 * r * (2 / 3) * np.sqrt(4 * np.pi)
    eQ[:, H.lm_index(lmax, 2, 0)] = -2 * Om ** 2 * r * (2 / 3) * np.sqrt(4 * np.pi / 5)
    eS[:, H.lm_index(lmax, 2, 0)] = -Om ** 2 * r * (2 / 3) * np.sqrt(4 * np.pi / 5)
    e = [H.rell2(to_host(NQ)[:nr], eQ), H.rell2(to_host(NS)[:nr], eS),
         float(np.max(np.abs(to_host(NT)[:nr])))]
    print(f"solid body: NQ {e[0]:.2e} NS {e[1]:.2e} |NT| {e[2]:.2e}")
    assert e[0] < 1e-9 and e[1] < 1e-9 and e[2] < 1e-10


def test_nonlinear_uxw_prime_281_vs_extended_truth():
    """u x omega at nr = 256: radial length 281, a prime, which takes the Rader
    register-pass kernels (lmax 13: the lm count 105 is not a multiple of their
    8-column tile), against the same composition in x87 extended precision.
    At this radius the second radial derivative multiplies FP64 rounding by
    kz^2 ~ 1.5e6: the FP64 oracle is itself ~8e-12 from the truth in NQ and
    ~1.4e-11 in NT, so the GPU is held to the truth: 1e-12 jointly, and per
    field at least as close as the oracle (within 1.5x, plus 1e-13).  Rows past
    nr are zero and the inputs are left as they were."""
    S = sphp()
    lmax, nr, C, d, rint, rext = 13, 256, 25, 5, 0.35, 1.0
    n = nr + C
    dz = (rext - rint) / (nr - 1)
    nlat = 4 * ((3 * (lmax + 1) + 7) // 8)
    nphi = 2 * nlat
    plan = S.FCPlan(lmax, nlat, nphi, n, C, d, dz, rint)
    W, Z = H.shell_fields(lmax, nr, n, rint, rext, seed=20261203 + lmax, slope=2.0)
    Wd, Zd = to_dev(W), to_dev(Z)
    out = [torch.full_like(Wd, 3.0) for _ in range(3)]
    S.nonlinear_uxw(plan, Wd, Zd, *out)
    full = [to_host(o) for o in out]
    assert np.array_equal(to_host(Wd), W) and np.array_equal(to_host(Zd), Z)
    assert all(float(np.max(np.abs(g[nr:]))) == 0.0 for g in full)
    got = [g[:nr] for g in full]
    A, Q = O.load_tables(tables_dir(), C, d)
    ref = [r[:nr] for r in O.nonlinear_uxw(O.SHT(lmax, nlat, nphi), O.FC(n, C, d, O.fc_dir(A, Q)), rint, dz, W, Z)]
    tru = [t[:nr].astype(np.complex128) for t in O.nonlinear_uxw_ld(lmax, nlat, nphi, A, Q, rint, dz, W, Z)]
    e_gpu = [H.rell2(g, t) for g, t in zip(got, tru)]
    e_orc = [H.rell2(r, t) for r, t in zip(ref, tru)]
    j_gpu = H.rell2_joint(got, tru)
    for nm, a, b_ in zip(("NQ", "NS", "NT"), e_gpu, e_orc):
        print(f"{nm}: gpu vs truth {a:.3e}, oracle vs truth {b_:.3e}")
    print(f"joint: gpu vs truth {j_gpu:.3e}, oracle vs truth {H.rell2_joint(ref, tru):.3e}")
    assert j_gpu < TOL
    assert all(a <= 1.5 * b_ + 1e-13 for a, b_ in zip(e_gpu, e_orc)), (e_gpu, e_orc)
    plan.close()


def test_nonlinear_uxw_headline_shells_vs_extended_truth():
    """u x omega at the benchmark size (lmax 255, nlat 384, nphi 768, nr 128,
    the benchmark's seeded input) on shells 62..65, against the same
    composition computed in x87 extended precision (liboracle_ld.so, 64-bit
    significand, same FP64 inputs and tables).  NT is ~1700x smaller than NQ
    here, so FP64 rounding of the large fields dominates its relative error;
    the truth shows how much of it each FP64 implementation carries.  The GPU
    must be at least as close to the truth as the FP64 oracle, field by
    field (within 1.5x, plus 1e-13), and within 1e-12 jointly."""
    S = sphp()
    lmax, nr, C, d, rint, rext = 255, 128, 25, 5, 0.35, 1.0
    n = nr + C
    dz = (rext - rint) / (nr - 1)
    nlat = 4 * ((3 * (lmax + 1) + 7) // 8)
    nphi = 2 * nlat
    k0, nk = 62, 4
    plan = S.FCPlan(lmax, nlat, nphi, n, C, d, dz, rint)
    W, Z = H.shell_fields(lmax, nr, n, rint, rext, seed=20261203, slope=2.0)
    Wd, Zd = to_dev(W), to_dev(Z)
    out = [torch.zeros_like(Wd) for _ in range(3)]
    S.nonlinear_uxw(plan, Wd, Zd, *out)
    sl = slice(k0, k0 + nk)
    got = [to_host(o)[sl] for o in out]
    A, Q = O.load_tables(tables_dir(), C, d)
    ref = [r[sl] for r in O.nonlinear_uxw(O.SHT(lmax, nlat, nphi), O.FC(n, C, d, O.fc_dir(A, Q)),
                                          rint, dz, W, Z, k0, nk)]
    tru = [t[sl].astype(np.complex128) for t in O.nonlinear_uxw_ld(lmax, nlat, nphi, A, Q, rint, dz, W, Z, k0, nk)]
    e_gpu = [H.rell2(g, t) for g, t in zip(got, tru)]
    e_orc = [H.rell2(r, t) for r, t in zip(ref, tru)]
    j_gpu = H.rell2_joint(got, tru)
    j_orc = H.rell2_joint(ref, tru)
    j_go = H.rell2_joint(got, ref)
    for nm, a, b_ in zip(("NQ", "NS", "NT"), e_gpu, e_orc):
        print(f"{nm}: gpu vs truth {a:.3e}, oracle vs truth {b_:.3e}")
    print(f"joint: gpu vs truth {j_gpu:.3e}, oracle vs truth {j_orc:.3e}, gpu vs oracle {j_go:.3e}")
    assert j_gpu < TOL and j_go < TOL
    assert all(a <= 1.5 * b_ + 1e-13 for a, b_ in zip(e_gpu, e_orc)), (e_gpu, e_orc)
    plan.close()
