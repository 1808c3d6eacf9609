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


def _stage_setup(lmax, nr):
    from redondito_amd import sphp as S
    C, d, rint, rext = 25, 5, 0.35, 1.0
    nz = nr + C
    dz = (rext - rint) / (nr - 1)
    nlat = 4 * ((3 * (lmax + 1) + 7) // 8)
    plan = S.FCPlan(lmax, nlat, 2 * nlat, nz, C, d, dz, rint)
    bcp = S.BCPlan()
    S.setup_bc(bcp, plan, ("noslip", "noslip"), "v")
    binv, d2b = (to_host(x) for x in S.bc_tables(plan))
    d2b = d2b.T.copy()
    A, Q = O.load_tables(S.TABLE_DIR, C, d)
    fc = O.FC(nz, C, d, O.fc_dir(A, Q))
    sht = O.SHT(lmax, nlat, 2 * nlat)
    z, kz = plan.grid()[3], plan.grid()[4]
    return S, plan, fc, sht, z, kz, H.ell_of(lmax), binv, d2b, dz, rint


@pytest.mark.parametrize("lmax,nr,filt", [(15, 24, True), (31, 40, False)])
def test_stage_uxw_vs_oracle(lmax, nr, filt):
    """One Runge-Kutta stage (nonlinear term, explicit update with viscosity,
    radial filter, no-slip walls) against the oracle's composition of the same
    reference operators, from a state that already satisfies the walls."""
    S, plan, fc, sht, z, kz, ell, binv, d2b, dz, rint = _stage_setup(lmax, nr)
    nz = nr + 25
    W, Zt = H.shell_fields(lmax, nr, nz, rint, 1.0, seed=lmax, slope=2.0)
    # a wall-consistent start: phi = D_l W, then the wall step
    Zs, phis, Ws = O.v_imposebc_fast(fc, kz, z, ell, binv, d2b, lmax, Zt, fc.dlr(W, kz, z, ell))
    o, dt, nu = 2, 1e-3, 0.05
    Z0, phi0 = Zs * 1.01, phis * 0.99
    ref = O.stage_uxw(sht, fc, rint, dz, kz, z, ell, binv, d2b, lmax, o, dt, nu, filt, Z0, phi0, Zs, phis, Ws)
    Zd, phid, Wd = to_dev(Zs), to_dev(phis), to_dev(Ws)
    S.stage_uxw(plan, to_dev(Z0), to_dev(phi0), Zd, phid, Wd, o, dt, nu, filter=filt)
    got = [to_host(x) for x in (Zd, phid, Wd)]
    errs = [H.rell2(g[:nr], r[:nr]) for g, r in zip(got, ref)]
    floor = H.noise_floor(lambda zz, pp: [x[:nr] for x in O.stage_uxw(sht, fc, rint, dz, kz, z, ell, binv, d2b, lmax,
                                                                       o, dt, nu, filt, Z0, phi0, zz, pp, Ws)],
                          [Zs, phis])
    print(f"lmax={lmax} nr={nr} filter={filt}: Z {errs[0]:.2e} phi {errs[1]:.2e} W {errs[2]:.2e}; "
          f"floor {floor[0]:.2e} {floor[1]:.2e} {floor[2]:.2e}")
    assert all(e < max(1e-12, 30 * f) for e, f in zip(errs, floor))
    # the walls hold after the stage: Z = W = dW/dr = 0
    N = nr
    assert np.max(np.abs(got[0][[0, N - 1]])) == 0 and np.max(np.abs(got[2][[0, N - 1]])) < 1e-13 * np.abs(got[2]).max()


def test_stage_uxw_solid_body_rotation_is_steady():
    """u = Omega z x r: u x omega is a gradient, so its curl vanishes and the
    inviscid stage leaves Z unchanged off the walls (the wall step zeroes it on
    them) and phi = W = 0."""
    S, plan, fc, sht, z, kz, ell, binv, d2b, dz, rint = _stage_setup(15, 40)
    nz, nlm = plan.nz, plan.nlm
    nr = nz - 25
    Zs = np.zeros((nz, nlm), complex)
    Zs[:nr, H.lm_index(15, 1, 0)] = 0.7 * z[:nr] ** 2 * np.sqrt(4 * np.pi / 3)
    zero = np.zeros_like(Zs)
    Zd, phid, Wd = to_dev(Zs), to_dev(zero), to_dev(zero)
    S.stage_uxw(plan, to_dev(Zs), to_dev(zero), Zd, phid, Wd, 1, 0.1, 0.0, filter=False)
    Zg, pg, Wg = (to_host(x) for x in (Zd, phid, Wd))
    scale = np.abs(Zs).max()
    dZ = np.abs(Zg[1:nr - 1] - Zs[1:nr - 1]).max() / scale
    print(f"solid body: interior Z change {dZ:.2e}, |phi| {np.abs(pg[:nr]).max():.2e}, |W| {np.abs(Wg[:nr]).max():.2e}")
    assert dZ < 1e-11 and np.abs(pg[:nr]).max() < 1e-9 and np.abs(Wg[:nr]).max() < 1e-10
    assert np.all(Zg[[0, nr - 1]] == 0)


def test_stage_uxw_toroidal_viscous_term():
    """An axisymmetric azimuthal flow (Z only in l = 1, m = 0, Z = a f(r) with
    f = sin(pi (r - ri)/(ro - ri))): u x omega lies in the meridional planes,
    so its toroidal part N_T vanishes and the stage changes Z by exactly
    dt/o nu D_1 Z = dt/o nu a (f'' - 2 f / r^2), to the FC-Gram derivative
    accuracy; a sign or operator slip in the viscous term fails."""
    S, plan, fc, sht, z, kz, ell, binv, d2b, dz, rint = _stage_setup(15, 64)
    nz, nlm = plan.nz, plan.nlm
    nr = nz - 25
    r = z[:nr]
    k = np.pi / (1.0 - rint)
    f = np.sin(k * (r - rint))
    fpp = -k * k * f
    i10 = H.lm_index(15, 1, 0)
    Zs = np.zeros((nz, nlm), complex)
    Zs[:nr, i10] = 0.3 * f
    zero = np.zeros_like(Zs)
    o, dt, nu = 2, 0.2, 0.05
    Zd, phid, Wd = to_dev(Zs), to_dev(zero), to_dev(zero)
    S.stage_uxw(plan, to_dev(Zs), to_dev(zero), Zd, phid, Wd, o, dt, nu, filter=False)
    Zg = to_host(Zd)
    got = (Zg[1:nr - 1, i10] - Zs[1:nr - 1, i10]) / (dt / o * nu)
    exact = 0.3 * (fpp - 2 * f / r ** 2)[1:nr - 1]
    e = np.max(np.abs(got - exact)) / np.max(np.abs(exact))
    other = np.max(np.abs(np.delete(Zg[:nr], i10, 1)))
    print(f"toroidal viscous term vs D_1 f: {e:.2e}; other modes {other:.1e}")
    assert e < 1e-4 and other < 1e-14
