"""GPU parity of the no-slip wall solves (src/boundary/) against the oracle.

* derivr_mat: the device FC-Gram matrices against the oracle's, within the
  FP64 conditioning of the continuation (a few 1e-12 of the largest entry).
* invDl_setup: M_l against the oracle's; both come from a row-scaled LU of
  B_l, so they agree to cond(B_l) * eps.  The solve they define is pinned by
  the residual B_l W = P phi.
* invDl: the device GEMM against the oracle's loop over the same M_l, to 1e-13.
* v_imposebc / v_imposebc_fast: the ten checks of src/prueba_boundary.f90 at
  the levels LEEME.md:45 records, and the outputs against the oracle's
  compositions.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from tests import helpers as H

pytestmark = pytest.mark.gpu


def dev():
    return torch.device("cuda:0")


def to_dev(x):
    return torch.from_numpy(np.ascontiguousarray(x)).to(dev())


def to_host(x):
    torch.cuda.synchronize()
    return x.cpu().numpy()


def _setup(lmax, nz, rint=0.5, rext=1.5):
    from redondito_amd import sphp as S
    C, d = 25, 5
    dz = (rext - rint) / (nz - C - 1)
    nlat = 4 * ((3 * (lmax + 1) + 7) // 8)
    plan = S.FCPlan(lmax, nlat, 2 * nlat, nz, C, d, dz, rint)
    A, Q = O.load_tables(S.TABLE_DIR, C, d)
    fc = O.FC(nz, C, d, O.fc_dir(A, Q))
    z, kz, ell = plan.grid()[3], plan.grid()[4], H.ell_of(lmax)
    return S, plan, fc, z, kz, ell


@pytest.mark.parametrize("lmax,nz", [(7, 128), (31, 89), (127, 153)])
def test_derivr_mat_and_binv_vs_oracle(lmax, nz):
    S, plan, fc, z, kz, ell = _setup(lmax, nz)
    N = nz - 25
    d1 = to_host(S.derivr_mat(plan, 1))
    d2 = to_host(S.derivr_mat(plan, 2))
    o1, o2 = fc.derivr_mat(kz, 1), fc.derivr_mat(kz, 2)
    e1 = np.abs(d1 - o1).max() / np.abs(o1).max()
    e2 = np.abs(d2 - o2).max() / np.abs(o2).max()
    print(f"derivr_mat d1 {e1:.2e} d2 {e2:.2e}")
    assert e1 < 1e-12 and e2 < 1e-12
    bcp = S.BCPlan()
    S.setup_bc(bcp, plan, ("noslip", "noslip"), "v")
    binv, d2b = (to_host(x) for x in S.bc_tables(plan))
    ob = O.invdl_setup(N, lmax, o1, o2, z[:N])
    eb = np.abs(binv - ob).max() / np.abs(ob).max()
    assert np.array_equal(d2b.T, O.d2bor(d2))
    # the solve itself: B_l M_l = P S^-1 columns, checked through the residual
    worst = 0.0
    for l in sorted({1, 2, lmax // 2, lmax}):
        M = binv[l - 1].T                     # N x (N-4)
        B = d2.copy()
        B[np.arange(N), np.arange(N)] -= l * (l + 1) / z[:N] ** 2
        B[0] = 0
        B[0, 0] = 1
        B[1] = d1[0]
        B[N - 2] = d1[N - 1]
        B[N - 1] = 0
        B[N - 1, N - 1] = 1
        R = B @ M
        P = np.zeros((N, N - 4))
        P[np.arange(2, N - 2), np.arange(N - 4)] = 1.0
        scale = np.abs(B).max(axis=1)[:, None] * np.abs(M).max()
        worst = max(worst, (np.abs(R - P) / scale).max())
    print(f"binv vs oracle {eb:.2e}, scaled residual {worst:.2e}")
    assert worst < 1e-13
    assert eb < 1e-8


@pytest.mark.parametrize("lmax,nz", [(7, 128), (63, 121), (255, 153)])
def test_invdl_vs_oracle(lmax, nz):
    S, plan, fc, z, kz, ell = _setup(lmax, nz)
    N = nz - 25
    bcp = S.BCPlan()
    S.setup_bc(bcp, plan, ("noslip", "noslip"), "v")
    binv = to_host(S.bc_tables(plan)[0])
    rng = np.random.default_rng(lmax)
    phi = rng.standard_normal((nz, ell.size)) + 1j * rng.standard_normal((nz, ell.size))
    Wd = torch.full((nz, ell.size), 3.0 + 0j, dtype=torch.complex128, device=dev())
    S.invDl(plan, to_dev(phi), Wd)
    W = to_host(Wd)
    ref = O.invdl(N, nz, lmax, binv, phi)
    err = H.rell2(W, ref)
    print(f"invDl lmax={lmax} nz={nz}: rel L2 {err:.2e}")
    assert err < 1e-13
    assert np.all(W[N:] == 0) and np.all(W[:, 0] == 0)


def test_prueba_boundary_levels_on_device():
    """src/prueba_boundary.f90 (lmax = 7, nz = 128, rint = 0.5, rext = 1.5)
    through the C-ABI, and the outputs against the oracle's compositions."""
    lmax, nz = 7, 128
    S, plan, fc, z, kz, ell = _setup(lmax, nz)
    N = nz - 25
    bcp = S.BCPlan()
    S.setup_bc(bcp, plan, ("noslip", "noslip"), "v")
    d1 = to_host(S.derivr_mat(plan, 1))
    z0, w0 = H.boundary_inputs(z, ell, N)
    W0, F0 = to_dev(w0), torch.empty((nz, ell.size), dtype=torch.complex128, device=dev())
    S.dlr(plan, W0, F0)
    f0 = to_host(F0)
    fs = H.boundary_stage_phi(f0, ell, N)
    tor, fi, pol = to_dev(z0), to_dev(fs), torch.zeros_like(W0)
    S.v_imposebc(tor, fi, pol, plan)
    pol2 = torch.zeros_like(W0)
    tor2, fi2 = tor.clone(), fi.clone()
    S.v_imposebc(tor2, fi2, pol2, plan)
    tor3, fi3, pol3 = to_dev(z0), to_dev(fs), torch.zeros_like(W0)
    S.v_imposebc_fast(tor3, fi3, pol3, plan)
    h = [to_host(x) for x in (tor, fi, pol, pol2, tor3, fi3, pol3)]
    err = H.boundary_errs(z, ell, N, d1, z0, w0, f0, *h)
    print("prueba_boundary errs", " ".join(f"{e:.2e}" for e in err))
    for k, (e, lim) in enumerate(zip(err, H.BOUNDARY_LIMITS)):
        assert e <= lim, (k + 1, e, lim)
    # against the oracle's compositions with the device tables
    binv, d2b = (to_host(x) for x in S.bc_tables(plan))
    ot, of, op = O.v_imposebc(fc, kz, z, ell, binv, lmax, z0, fs)
    ot3, of3, op3 = O.v_imposebc_fast(fc, kz, z, ell, binv, d2b.T.copy(), lmax, z0, fs)
    assert np.array_equal(h[0], ot)
    assert H.rell2(h[2], op) < 1e-13 and H.rell2(h[6], op3) < 1e-13
    assert H.rell2(h[1], of) < 1e-11 and H.rell2(h[5], of3) < 1e-12


def test_bc_errors():
    from redondito_amd import sphp as S
    S_, plan, fc, z, kz, ell = _setup(7, 60)
    W = torch.zeros((60, ell.size), dtype=torch.complex128, device=dev())
    with pytest.raises(S.RdError):
        S.invDl(plan, W, W.clone())          # no setup yet
    with pytest.raises(S.RdError):
        S.setup_bc(S.BCPlan(), plan, ("noslip", "freeslip"), "v")
