"""Host-side mirror of REDONDITO's sphp / pseudo interface over the C-ABI.

Routine names, argument order and meaning follow the reference's Fortran
(src/sphp/sphp_sca.F90, src/sphp/sphp_vec.F90, src/sphp/fcgram_mod.f90,
src/pseudo/pseudospec_hd.f90, src/boundary/boundary_mod.f90); the arrays are
torch tensors resident on the GPU:

  spectral a(lm, k)          -> complex128 tensor of shape (nz, nlm)
  physical f(theta, phi, k)  -> float64 tensor of shape (nplanes, nphi, nlat)

The work is queued on torch's current stream (the plan is bound to it at
every call), so ordering with surrounding torch code is the stream's.  There
is no CPU fallback: loading fails loudly when the HIP library is absent.
"""
import ctypes
import os

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_lib", "librdsphp.so")
TABLE_DIR = os.path.join(_HERE, "tables")

RD_PREC_FP64, RD_PREC_FP32, RD_PREC_BF16X3, RD_PREC_OZAKI = 0, 1, 2, 3
RD_NOALLOC_SCRATCH_BYTES = 64 << 20   # fixed scratch of rd_sht_synthesis_noalloc (include/redondito.h)

_lib = None


class RdError(RuntimeError):
    pass


def lib():
    """The loaded C-ABI library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RdError(f"HIP library not built: {LIB_PATH} (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        vp, i, d = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
        L.rd_error_string.restype = ctypes.c_char_p
        L.rd_error_string.argtypes = [i]
        L.rd_load_dirichlet_tables.argtypes = [ctypes.c_char_p, i, i, ctypes.POINTER(d)]
        L.rd_plan_create.argtypes = [ctypes.POINTER(vp), i, i, i, i, i, i, d, d, ctypes.POINTER(d)]
        L.rd_plan_destroy.argtypes = [vp]
        L.rd_plan_set_stream.argtypes = [vp, vp]
        L.rd_plan_set_ozaki_digits.argtypes = [vp, i]
        L.rd_plan_sync.argtypes = [vp]
        L.rd_plan_info.argtypes = [vp, ctypes.POINTER(i)]
        L.rd_plan_grid.argtypes = [vp] + [ctypes.POINTER(d)] * 5 + [ctypes.POINTER(i)] * 2
        for nm in ("rd_sht_analysis", "rd_sht_synthesis", "rd_sht_synthesis_noalloc", "rd_redist_k_to_lm",
                   "rd_redist_lm_to_k"):
            getattr(L, nm).argtypes = [vp, i, vp, vp]
        for nm in ("rd_vsht_analysis", "rd_vsht_synthesis"):
            getattr(L, nm).argtypes = [vp, i] + [vp] * 6
        for nm in ("rd_z_forward", "rd_z_inverse", "rd_dirichlet"):
            getattr(L, nm).argtypes = [vp, vp]
        for nm in ("rd_3d_forward", "rd_3d_inverse", "rd_laplang", "rd_dlr", "rd_copy"):
            getattr(L, nm).argtypes = [vp, vp, vp]
        L.rd_derivr.argtypes = [vp, i, ctypes.c_char, vp, vp]
        L.rd_multr.argtypes = [vp, i, vp, vp]
        L.rd_saxpby_c.argtypes = [vp, vp, vp, d, vp, d]
        L.rd_saxpby_r.argtypes = [vp, i, vp, vp, d, vp, d]
        L.rd_prodvec.argtypes = [vp, i] + [vp] * 9
        L.rd_fc_filter.argtypes = [vp, ctypes.c_char, vp]
        L.rd_load_neumann_tables.argtypes = [ctypes.c_char_p, i, i, d, ctypes.POINTER(d)]
        L.rd_neumann_reconstruct.argtypes = [vp, vp, i, i, ctypes.POINTER(d)]
        L.rd_robin_reconstruct.argtypes = [vp, vp, i, ctypes.POINTER(d), vp]
        L.rd_nonlinear_uxw.argtypes = [vp, i, vp, vp, vp, vp, vp]
        L.rd_nonlinear_uxw_capture.argtypes = [vp, i, vp, vp, vp, vp, vp]
        L.rd_plan_prepare.argtypes = [vp, i]
        L.rd_plan_replay.argtypes = [vp]
        L.rd_stage_uxw.argtypes = [vp, i, i, d, d, i, vp, vp, vp, vp, vp]
        L.rd_fp64_ceilings.argtypes = [ctypes.POINTER(d), ctypes.POINTER(d)]
        L.rd_fp64_mfma4x4_ceiling.argtypes = [ctypes.POINTER(d)]
        L.rd_nonlinear_coupled.argtypes = [vp, i] + [vp] * 12
        L.rd_plan_profile.argtypes = [vp, i]
        L.rd_plan_stage_times.argtypes = [vp, ctypes.POINTER(d), ctypes.POINTER(ctypes.c_long), i]
        L.rd_derivr_mat.argtypes = [vp, i, vp]
        L.rd_bc_setup.argtypes = [vp, i, i]
        L.rd_bc_tables.argtypes = [vp, vp, vp]
        L.rd_invdl.argtypes = [vp, vp, vp]
        L.rd_v_imposebc.argtypes = [vp, vp, vp, vp]
        L.rd_v_imposebc_fast.argtypes = [vp, vp, vp, vp]
        _lib = L
    return _lib


def _check(rc, what):
    if rc != 0:
        raise RdError(f"{what}: {lib().rd_error_string(rc).decode()} ({rc})")


def load_dirichlet_tables(tdir, C, d):
    """dir = A.Q^T (C x d) from A{C}-{d}.dat and Q{d}.dat
    (load_dirichlet_tables, src/sphp/fcgram_mod.f90:186-263)."""
    import numpy as np
    out = np.zeros(C * d)
    _check(lib().rd_load_dirichlet_tables(str(tdir).encode(), C, d,
                                          out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
           "load_dirichlet_tables")
    return out.reshape(d, C).T.copy()


def load_neumann_tables(tdir, d, ord_, dx):
    """neu (d values) of load_neumann_tables (src/sphp/fcgram_mod.f90:267-371)
    for matching points at spacing dx: Q{d}.dat and Q{ord}n{d}.dat."""
    import numpy as np
    out = np.full(d, np.nan)
    _check(lib().rd_load_neumann_tables(str(tdir).encode(), d, ord_, float(dx),
                                        out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))),
           "load_neumann_tables")
    return out


def grid_nlat(lmax):
    """nlat from the 3/2 rule of grid_init (src/pseudo/pseudospec_mod.f90:66-68)."""
    return 4 * ((3 * (lmax + 1) + 7) // 8)


class FCPlan:
    """The FCPLAN of the reference: SHT grid + radial FC-Gram plan.

    fcgram_create_plan(this, N=(nlat,nphi,nz), C=(0,0,Cz), o=(0,0,oz),
    dx=(1,1,dz), tdir, ...) -> FCPlan(lmax, nlat, nphi, nz, Cz, oz, dz,
    rint, tdir).  Cz = oz = 0 makes z periodic.
    """

    def __init__(self, lmax, nlat, nphi, nz, C=0, d=0, dz=1.0, rint=0.0, tdir=TABLE_DIR, device=None):
        import numpy as np
        self.lmax, self.nlat, self.nphi, self.nz, self.C, self.d = lmax, nlat, nphi, nz, C, d
        self.nlm = (lmax + 1) * (lmax + 2) // 2
        self.nr = nz - C
        self.dz, self.rint = dz, rint
        self.device = torch.device(device if device is not None else "cuda")
        if C > 0 and d > 0:
            self.dir = load_dirichlet_tables(tdir, C, d)
            dv = np.asfortranarray(self.dir).ravel(order="F").copy()
            dptr = dv.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        elif C == 0 and d == 0:
            self.dir, dptr = None, None
        else:
            raise RdError("Mismatch in continuation or matching points in z direction")
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            _check(lib().rd_plan_create(ctypes.byref(h), lmax, nlat, nphi, nz, C, d, dz, rint, dptr),
                   "fcgram_create_plan")
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().rd_plan_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def h(self):
        """Plan handle bound to torch's current stream."""
        lib().rd_plan_set_stream(self._h, ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream))
        return self._h

    # (the u and omega syntheses are one launch; the radial step also clears
    # the output rows past nr)
    STAGES = ("radial", "synthesis", "phi_cross", "analysis")

    def set_ozaki_digits(self, digits):
        """Digits per operand (7, 6 or 5) of this plan's later RD_PREC_OZAKI calls."""
        _check(lib().rd_plan_set_ozaki_digits(self._h, int(digits)), "plan_set_ozaki_digits")

    def prepare(self, prec):
        """Build the tables precision form prec needs now (no allocation on its later calls)."""
        _check(lib().rd_plan_prepare(self._h, prec), "plan_prepare")

    def capture_uxw(self, W, Z, NQ, NS, NT, prec=RD_PREC_FP64):
        """Capture one nonlinear_uxw into the plan's hipGraph (replay())."""
        _check(lib().rd_nonlinear_uxw_capture(self.h, prec, self._spec(W), self._spec(Z), self._spec(NQ),
                                              self._spec(NS), self._spec(NT)), "nonlinear_uxw_capture")

    def replay(self):
        _check(lib().rd_plan_replay(self.h), "plan_replay")

    def profile(self, enable=True):
        """Switch per-stage event timing of nonlinear_uxw on or off (resets it)."""
        _check(lib().rd_plan_profile(self._h, 1 if enable else 0), "profile")

    def stage_times(self):
        """{stage: (total_ms, launches)} accumulated since profile(True)."""
        ms = (ctypes.c_double * 8)()
        cnt = (ctypes.c_long * 8)()
        n = lib().rd_plan_stage_times(self._h, ms, cnt, 8)
        if n < 0:
            _check(n, "stage_times")
        return {self.STAGES[i]: (ms[i], cnt[i]) for i in range(n)}

    def grid(self):
        """(ct, st, wt, z, kz, ell, emm) as numpy arrays (box_init's grids)."""
        import numpy as np
        ct, st, wt = (np.zeros(self.nlat) for _ in range(3))
        z, kz = np.zeros(self.nz), np.zeros(self.nz)
        ell, emm = np.zeros(self.nlm, np.int32), np.zeros(self.nlm, np.int32)
        P = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        Pi = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
        _check(lib().rd_plan_grid(self._h, P(ct), P(st), P(wt), P(z), P(kz), Pi(ell), Pi(emm)), "grid")
        return ct, st, wt, z, kz, ell, emm

    # -- shape checks -------------------------------------------------
    def _spec(self, a, rows=None):
        rows = self.nz if rows is None else rows
        if a.dtype != torch.complex128 or not a.is_cuda or not a.is_contiguous():
            raise RdError("spectral arrays must be contiguous complex128 CUDA tensors")
        if a.shape != (rows, self.nlm):
            raise RdError(f"spectral shape {tuple(a.shape)} != ({rows}, {self.nlm})")
        return ctypes.c_void_p(a.data_ptr())

    def _phys(self, f, nplanes):
        if f.dtype != torch.float64 or not f.is_cuda or not f.is_contiguous():
            raise RdError("physical arrays must be contiguous float64 CUDA tensors")
        if f.shape != (nplanes, self.nphi, self.nlat):
            raise RdError(f"physical shape {tuple(f.shape)} != ({nplanes}, {self.nphi}, {self.nlat})")
        return ctypes.c_void_p(f.data_ptr())


# ---------------------------------------------------------------- sphp
def fftp2d_real_to_complex_xy(plan, inp, out):
    """in(nlat,nphi,planes) -> out(lm,planes) (src/sphp/sphp_sca.F90:273-329)."""
    n = inp.shape[0]
    _check(lib().rd_sht_analysis(plan.h, n, plan._phys(inp, n), plan._spec(out, n)), "fftp2d_real_to_complex_xy")


def fftp2d_complex_to_real_xy(plan, inp, out):
    """in(lm,planes) -> out(nlat,nphi,planes) (src/sphp/sphp_sca.F90:436-492)."""
    n = inp.shape[0]
    _check(lib().rd_sht_synthesis(plan.h, n, plan._spec(inp, n), plan._phys(out, n)), "fftp2d_complex_to_real_xy")


def fftp2d_vec_to_qst(plan, vr, vt, vp, q, s, t):
    """(v_r,v_theta,v_phi) -> (Q,S,T) (src/sphp/sphp_vec.F90:49-106)."""
    n = vr.shape[0]
    _check(lib().rd_vsht_analysis(plan.h, n, plan._phys(vr, n), plan._phys(vt, n), plan._phys(vp, n),
                                  plan._spec(q, n), plan._spec(s, n), plan._spec(t, n)), "fftp2d_vec_to_qst")


def fftp2d_qst_to_vec(plan, q, s, t, vr, vt, vp):
    """(Q,S,T) -> (v_r,v_theta,v_phi) (src/sphp/sphp_vec.F90:109-165)."""
    n = q.shape[0]
    _check(lib().rd_vsht_synthesis(plan.h, n, plan._spec(q, n), plan._spec(s, n), plan._spec(t, n),
                                   plan._phys(vr, n), plan._phys(vt, n), plan._phys(vp, n)), "fftp2d_qst_to_vec")


def fftp1d_real_to_complex_z(plan, inout):
    """In place, continuation + forward FFT + 1/nz (src/sphp/sphp_sca.F90:332-397)."""
    _check(lib().rd_z_forward(plan.h, plan._spec(inout)), "fftp1d_real_to_complex_z")


def fftp1d_complex_to_real_z(plan, inout):
    """In place, backward FFT (src/sphp/sphp_sca.F90:495-530)."""
    _check(lib().rd_z_inverse(plan.h, plan._spec(inout)), "fftp1d_complex_to_real_z")


def fftp3d_real_to_complex(plan, inp, out):
    """(src/sphp/sphp_sca.F90:237-270)."""
    _check(lib().rd_3d_forward(plan.h, plan._phys(inp, plan.nz), plan._spec(out)), "fftp3d_real_to_complex")


def fftp3d_complex_to_real(plan, inp, out):
    """in is destroyed (src/sphp/sphp_sca.F90:400-433)."""
    _check(lib().rd_3d_inverse(plan.h, plan._spec(inp), plan._phys(out, plan.nz)), "fftp3d_complex_to_real")


def fftp3d_redist_k_to_lm(plan, inp, out):
    """(src/sphp/sphp_sca.F90:629-693); a copy on one GPU."""
    n = inp.shape[0]
    _check(lib().rd_redist_k_to_lm(plan.h, n, plan._spec(inp, n), plan._spec(out, n)), "fftp3d_redist_k_to_lm")


def fftp3d_redist_lm_to_k(plan, inp, out):
    """(src/sphp/sphp_sca.F90:696-760); a copy on one GPU."""
    n = inp.shape[0]
    _check(lib().rd_redist_lm_to_k(plan.h, n, plan._spec(inp, n), plan._spec(out, n)), "fftp3d_redist_lm_to_k")


# -------------------------------------------------------------- pseudo
def derivr(plan, a, b, p, esp):
    """b = d^p a / dr^p in (lm,r) ('r') or (lm,kz) ('k') (pseudospec_hd.f90:18-135)."""
    if esp not in ("r", "k"):
        raise RdError(f"Unknown space {esp} in derivr")
    _check(lib().rd_derivr(plan.h, p, esp.encode(), plan._spec(a), plan._spec(b)), "derivr")


def laplang(plan, a, b):
    _check(lib().rd_laplang(plan.h, plan._spec(a), plan._spec(b)), "laplang")


def multr(plan, a, b, p):
    _check(lib().rd_multr(plan.h, p, plan._spec(a), plan._spec(b)), "multr")


def dlr(plan, a, b):
    _check(lib().rd_dlr(plan.h, plan._spec(a), plan._spec(b)), "dlr")


def saxpby_c(plan, z, x, a, y, b):
    _check(lib().rd_saxpby_c(plan.h, plan._spec(z), plan._spec(x), a, plan._spec(y), b), "saxpby_c")


def saxpby_r(plan, z, x, a, y, b):
    n = z.shape[0]
    _check(lib().rd_saxpby_r(plan.h, n, plan._phys(z, n), plan._phys(x, n), a, plan._phys(y, n), b), "saxpby_r")


def prodvec(plan, ar, at, ap, br, bt, bp, cr, ct, cp):
    n = ar.shape[0]
    args = [plan._phys(x, n) for x in (ar, at, ap, br, bt, bp, cr, ct, cp)]
    _check(lib().rd_prodvec(plan.h, n, *args), "prodvec")


def neumann_reconstruct(plan, a, boun, ord_, neu):
    """Wall value of a (nz, nlm) from its normal derivative in row 0 (boun 5)
    or nz - C - 1 (boun 6) (neumann_reconstruct, fcgram_mod.f90:374-517)."""
    import numpy as np
    w = np.ascontiguousarray(neu, np.float64)
    _check(lib().rd_neumann_reconstruct(plan.h, plan._spec(a), boun, ord_,
                                        w.ctypes.data_as(ctypes.POINTER(ctypes.c_double))), "neumann_reconstruct")


def robin_reconstruct(plan, a, boun, neu, coef):
    """As neumann_reconstruct for f' + coef f = g, coef a float64 CUDA tensor
    of nlm values (robin_reconstruct, fcgram_mod.f90:520-650)."""
    import numpy as np
    w = np.ascontiguousarray(neu, np.float64)
    if coef.dtype != torch.float64 or not coef.is_cuda or coef.numel() != plan.nlm or not coef.is_contiguous():
        raise RdError("robin coefficients must be a contiguous float64 CUDA tensor of nlm values")
    _check(lib().rd_robin_reconstruct(plan.h, plan._spec(a), boun, w.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                      ctypes.c_void_p(coef.data_ptr())), "robin_reconstruct")


def fc_filter(plan, a, esp):
    if esp not in ("r", "k"):
        raise RdError(f"Unknown space {esp} in fc_filter")
    _check(lib().rd_fc_filter(plan.h, esp.encode(), plan._spec(a)), "fc_filter")


def copy(plan, a, b):
    _check(lib().rd_copy(plan.h, plan._spec(a), plan._spec(b)), "copy")


def dirichlet(plan, a):
    _check(lib().rd_dirichlet(plan.h, plan._spec(a)), "dirichlet")


def nonlinear_uxw(plan, W, Z, NQ, NS, NT, prec=RD_PREC_FP64):
    """N = u x omega, u = curl curl (W r) + curl (Z r), as (Q,S,T) in (lm, r)."""
    _check(lib().rd_nonlinear_uxw(plan.h, prec, plan._spec(W), plan._spec(Z), plan._spec(NQ),
                                  plan._spec(NS), plan._spec(NT)), "nonlinear_uxw")


def stage_uxw(plan, Z0, phi0, Z, phi, W, o, dt, nu, filter=True, prec=RD_PREC_FP64):
    """One stage Z, phi = (Z0, phi0) + dt/o rhs of the explicit Runge-Kutta
    scheme (rhs from u x omega and viscosity nu), radial filter, no-slip walls
    (v_imposebc_fast); Z, phi, W updated in place.  Needs setup_bc."""
    _check(lib().rd_stage_uxw(plan.h, prec, int(o), float(dt), float(nu), 1 if filter else 0,
                              *[plan._spec(x) for x in (Z0, phi0, Z, phi, W)]), "stage_uxw")


def fp64_ceilings():
    """Measured FP64 throughput of the current device in TFLOP/s:
    (v_mfma_f64_16x16x4 loop, v_fma_f64 loop), register-only."""
    m, v = ctypes.c_double(0.0), ctypes.c_double(0.0)
    _check(lib().rd_fp64_ceilings(ctypes.byref(m), ctypes.byref(v)), "fp64_ceilings")
    return m.value, v.value


def fp64_mfma4x4_ceiling():
    """Measured FP64 throughput in TFLOP/s of a register-only
    v_mfma_f64_4x4x4_4b_f64 loop."""
    t = ctypes.c_double(0.0)
    _check(lib().rd_fp64_mfma4x4_ceiling(ctypes.byref(t)), "fp64_mfma4x4_ceiling")
    return t.value


def nonlinear_coupled(plan, W, Z, Pb, Tb, T, NQ, NS, NT, MQ, MS, MT, NH, prec=RD_PREC_FP64):
    """One coupled step: (NQ, NS, NT) = u x omega, (MQ, MS, MT) = u x B and
    NH = u . grad T, with u from (W, Z) and B from (Pb, Tb) as poloidal /
    toroidal potentials, all (lm, r)."""
    _check(lib().rd_nonlinear_coupled(plan.h, prec, *[plan._spec(x) for x in (W, Z, Pb, Tb, T, NQ, NS, NT, MQ,
                                                                                MS, MT, NH)]),
           "nonlinear_coupled")


# ------------------------------------------------------------ boundary
RD_BC_NOSLIP = 0
_BC_KINDS = {"noslip": RD_BC_NOSLIP}


def v_parsebc(s):
    """Wall type from its name (v_parsebc, src/boundary/vboundary.f90:10-31)."""
    key = s.strip().lower()
    if key not in _BC_KINDS:
        raise RdError(f"Unknown boundary condition type {s}")
    return _BC_KINDS[key]


class BCPlan:
    """BCPLAN (src/boundary/boundary_mod.f90:14-16): the wall types."""

    def __init__(self):
        self.bcint = self.bcext = None


def setup_bc(this, planfc, bckind, field):
    """Wall types and the per-l recovery matrices M_l on the device
    (setup_bc / v_setup / invDl_setup, boundary_mod.f90:35-58,140-241)."""
    if field != "v":
        return
    this.bcint, this.bcext = v_parsebc(bckind[0]), v_parsebc(bckind[1])
    _check(lib().rd_bc_setup(planfc.h, this.bcint, this.bcext), "invDl_setup")


def derivr_mat(plan, p):
    """d^p/dr^p on the nr physical points as an (nr, nr) float64 tensor,
    d[i, j] = d^p e_j / dr^p at r_i (derivr_mat, boundary_mod.f90:96-137)."""
    N = plan.nz - plan.C
    # column-major N x N on the device == the transpose of a row-major tensor
    dt = torch.empty((N, N), dtype=torch.float64, device=plan.device)
    _check(lib().rd_derivr_mat(plan.h, p, ctypes.c_void_p(dt.data_ptr())), "derivr_mat")
    return dt.t()


def bc_tables(plan):
    """Copies of the wall tables: binv as (lmax, N-4, N) with
    binv[l-1, j, i] = M_l[i, j], and d2bor as (N, 4) with d2bor[j, k] = entry
    j of wall row k (rows 1, 2, N-1, N of d2/dr2)."""
    N = plan.nz - plan.C
    binv = torch.empty((plan.lmax, N - 4, N), dtype=torch.float64, device=plan.device)
    d2bor = torch.empty((N, 4), dtype=torch.float64, device=plan.device)
    _check(lib().rd_bc_tables(plan.h, ctypes.c_void_p(binv.data_ptr()), ctypes.c_void_p(d2bor.data_ptr())),
           "bc_tables")
    return binv, d2bor


def invDl(plan, a, b):
    """W = M_l phi(3:N-2) (invDl, boundary_mod.f90:244-292)."""
    _check(lib().rd_invdl(plan.h, plan._spec(a), plan._spec(b)), "invDl")


def v_imposebc(tor, phi, pol, plan):
    """Z = 0 at the walls, W = M_l phi, phi = D_l W (vboundary.f90:59-88)."""
    _check(lib().rd_v_imposebc(plan.h, plan._spec(tor), plan._spec(phi), plan._spec(pol)), "v_imposebc")


def v_imposebc_fast(tor, phi, pol, plan):
    """As v_imposebc, phi recomputed on rows 1, 2, N-1, N only (vboundary.f90:91-130)."""
    _check(lib().rd_v_imposebc_fast(plan.h, plan._spec(tor), plan._spec(phi), plan._spec(pol)),
           "v_imposebc_fast")


# ------------------------------------------------------------ batching form
class Batch:
    """Many independent shell problems transformed together (the reference's
    fftp3d_complex_to_real / fftp3d_real_to_complex per problem, with the
    Legendre stage of all problems in one persistent launch per direction).

    problems: sequence of (lmax, nlat, nphi, nz, dz, rint); one FC-Gram table
    (C, d) for all.  Spectral arrays are complex128 (nz, nlm) tensors, grids
    float64 (nz, nphi, nlat) tensors, one per problem.
    """

    def __init__(self, problems, C=25, d=5, prec=RD_PREC_FP32, tdir=TABLE_DIR, device=None, form="3d"):
        import numpy as np
        L = lib()
        if form not in ("3d", "uxw"):
            raise RdError(f"unknown batch form {form}")
        self.form = form
        i, vp, dd = ctypes.c_int, ctypes.c_void_p, ctypes.c_double
        if not getattr(L, "_batch_bound", False):
            L.rd_batch_create.argtypes = [ctypes.POINTER(vp), i] + [ctypes.POINTER(i)] * 4 + [
                i, i, ctypes.POINTER(dd), ctypes.POINTER(dd), ctypes.POINTER(dd), i]
            for nm in ("rd_batch_destroy", "rd_batch_sync", "rd_batch_replay"):
                getattr(L, nm).argtypes = [vp]
            L.rd_batch_set_stream.argtypes = [vp, vp]
            L.rd_batch_info.argtypes = [vp, ctypes.POINTER(i)]
            for nm in ("rd_batch_3d_inverse", "rd_batch_3d_forward"):
                getattr(L, nm).argtypes = [vp, ctypes.POINTER(vp), ctypes.POINTER(vp)]
            L.rd_batch_capture.argtypes = [vp] + [ctypes.POINTER(vp)] * 3
            L.rd_batch_config.argtypes = [vp, i, i]
            L.rd_batch_tune.argtypes = [vp] + [ctypes.POINTER(vp)] * 3 + [i, ctypes.POINTER(dd),
                                                                           ctypes.POINTER(dd), ctypes.POINTER(i)]
            L.rd_batch_create_uxw.argtypes = L.rd_batch_create.argtypes
            for nm in ("rd_batch_nonlinear_uxw", "rd_batch_capture_uxw"):
                getattr(L, nm).argtypes = [vp] + [ctypes.POINTER(vp)] * 5
            L.rd_batch_split.argtypes = [vp, i]
            L.rd_batch_get_split.argtypes = [vp, ctypes.POINTER(i), ctypes.POINTER(i)]
            L.rd_batch_tune_uxw.argtypes = [vp] + [ctypes.POINTER(vp)] * 5 + [i, ctypes.POINTER(dd),
                                                                               ctypes.POINTER(dd), ctypes.POINTER(i)]
            L._batch_bound = True
        self.problems = [tuple(p) for p in problems]
        self.device = torch.device(device if device is not None else "cuda")
        n = len(self.problems)
        arr = lambda k, t: (t * n)(*[p[k] for p in self.problems])
        dirp = None
        if C > 0:
            dv = np.asfortranarray(load_dirichlet_tables(tdir, C, d)).ravel(order="F").copy()
            self._dir = dv
            dirp = dv.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        h = ctypes.c_void_p()
        with torch.cuda.device(self.device):
            create = L.rd_batch_create_uxw if form == "uxw" else L.rd_batch_create
            _check(create(ctypes.byref(h), n, arr(0, i), arr(1, i), arr(2, i), arr(3, i), C, d,
                          arr(4, dd), arr(5, dd), dirp, prec), "batch_create")
        self._h = h
        self.nlm = [(p[0] + 1) * (p[0] + 2) // 2 for p in self.problems]

    @property
    def h(self):
        lib().rd_batch_set_stream(self._h, ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream))
        return self._h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            lib().rd_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _ptrs(self, ts, kind):
        if len(ts) != len(self.problems):
            raise RdError("one array per problem")
        for t, p, nlm in zip(ts, self.problems, self.nlm):
            if kind == "spec":
                ok = t.dtype == torch.complex128 and tuple(t.shape) == (p[3], nlm)
            else:
                ok = t.dtype == torch.float64 and tuple(t.shape) == (p[3], p[2], p[1])
            if not (ok and t.is_cuda and t.is_contiguous()):
                raise RdError(f"bad {kind} array {tuple(t.shape)} {t.dtype}")
        return (ctypes.c_void_p * len(ts))(*[t.data_ptr() for t in ts])

    def info(self):
        v = (ctypes.c_int * 8)()
        _check(lib().rd_batch_info(self._h, v), "batch_info")
        keys = ("nprob", "prec", "jb", "wg_per_cu", "n_cu", "syn_items", "ana_items", "captured")
        return dict(zip(keys, list(v)))

    def inverse(self, spec, grid):
        _check(lib().rd_batch_3d_inverse(self.h, self._ptrs(spec, "spec"), self._ptrs(grid, "grid")),
               "batch_3d_inverse")

    def forward(self, grid, spec):
        _check(lib().rd_batch_3d_forward(self.h, self._ptrs(grid, "grid"), self._ptrs(spec, "spec")),
               "batch_3d_forward")

    def capture(self, spec, grid, spec_out):
        _check(lib().rd_batch_capture(self.h, self._ptrs(spec, "spec"), self._ptrs(grid, "grid"),
                                      self._ptrs(spec_out, "spec")), "batch_capture")

    def replay(self):
        _check(lib().rd_batch_replay(self.h), "batch_replay")

    def config(self, jb, wg_per_cu):
        _check(lib().rd_batch_config(self.h, jb, wg_per_cu), "batch_config")

    def nonlinear_uxw(self, W, Z, NQ, NS, NT):
        """u x omega of every problem (form "uxw"): lists of (nz, nlm) complex128 tensors."""
        _check(lib().rd_batch_nonlinear_uxw(self.h, *[self._ptrs(x, "spec") for x in (W, Z, NQ, NS, NT)]),
               "batch_nonlinear_uxw")

    def capture_uxw(self, W, Z, NQ, NS, NT):
        _check(lib().rd_batch_capture_uxw(self.h, *[self._ptrs(x, "spec") for x in (W, Z, NQ, NS, NT)]),
               "batch_capture_uxw")

    def split(self, lmax_min):
        """Problems with lmax >= lmax_min run one-problem chains (0: all, a
        value above every lmax: none); the results do not change."""
        _check(lib().rd_batch_split(self.h, int(min(lmax_min, 2**31 - 1))), "batch_split")

    def get_split(self):
        """(lmax threshold, number of problems run as one-problem chains)."""
        a, n = ctypes.c_int(), ctypes.c_int()
        _check(lib().rd_batch_get_split(self._h, ctypes.byref(a), ctypes.byref(n)), "batch_get_split")
        return a.value, n.value

    def tune_uxw(self, W, Z, NQ, NS, NT, reps=3):
        """Returns (ms_fixed, ms_tuned, (jb, wg_per_cu)) per evaluation of every problem."""
        f, t = ctypes.c_double(), ctypes.c_double()
        ch = (ctypes.c_int * 2)()
        _check(lib().rd_batch_tune_uxw(self.h, *[self._ptrs(x, "spec") for x in (W, Z, NQ, NS, NT)], reps,
                                       ctypes.byref(f), ctypes.byref(t), ch), "batch_tune_uxw")
        return f.value, t.value, (ch[0], ch[1])

    def tune(self, spec, grid, spec_out, reps=3):
        """Returns (ms_fixed, ms_tuned, (jb, wg_per_cu)) per round trip."""
        f, t = ctypes.c_double(), ctypes.c_double()
        ch = (ctypes.c_int * 2)()
        _check(lib().rd_batch_tune(self.h, self._ptrs(spec, "spec"), self._ptrs(grid, "grid"),
                                   self._ptrs(spec_out, "spec"), reps, ctypes.byref(f), ctypes.byref(t), ch),
               "batch_tune")
        return f.value, t.value, (ch[0], ch[1])
