"""ctypes front end of the CPU oracle (oracle/sphp_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and
by the CPU-baseline leg of bench.py, always as the checker.  The product
path (redondito_amd) never imports it.

Arrays follow the reference layouts: spectral (lm, k) complex128 with lm
fastest, i.e. numpy shape (nplanes, nlm); physical (theta, phi, k) float64
with theta fastest, i.e. numpy shape (nplanes, nphi, nlat).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "_build", "liboracle.so")
_lib = None

_P = ctypes.c_void_p
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.c_int
_L = ctypes.c_long


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB):
            build()
        L = ctypes.CDLL(_LIB)
        L.orc_sht_create.restype = _P
        L.orc_sht_create.argtypes = [_I, _I, _I]
        L.orc_sht_destroy.argtypes = [_P]
        L.orc_sht_grid.argtypes = [_P, _D, _D, _D]
        L.orc_sht_tables.argtypes = [_P, _D, _D, _D]
        for nm in ("orc_sh_to_spat", "orc_spat_to_sh"):
            getattr(L, nm).argtypes = [_P, _I, _D, _D]
        for nm in ("orc_shqst_to_spat", "orc_spat_to_shqst"):
            getattr(L, nm).argtypes = [_P, _I, _D, _D, _D, _D, _D, _D]
        L.orc_fc_dir.argtypes = [_I, _I, _D, _D, _D]
        L.orc_fc_create.restype = _P
        L.orc_fc_create.argtypes = [_I, _I, _I, _D]
        L.orc_fc_destroy.argtypes = [_P]
        L.orc_z_forward.argtypes = [_P, _I, _L, _D]
        L.orc_z_inverse.argtypes = [_P, _I, _L, _D]
        L.orc_kz.argtypes = [_I, ctypes.c_double, _D]
        L.orc_derivr_k.argtypes = [_I, _I, _L, _D, _I, _D, _D]
        L.orc_derivr_r.argtypes = [_P, _D, _I, _I, _L, _D, _D]
        L.orc_laplang.argtypes = [_I, _I, _L, ctypes.POINTER(ctypes.c_int), _D, _D]
        L.orc_multr.argtypes = [_I, _I, _L, _D, _I, _D, _D]
        L.orc_saxpby_c.argtypes = [_L, _D, _D, ctypes.c_double, _D, ctypes.c_double]
        L.orc_dlr.argtypes = [_P, _D, _D, ctypes.POINTER(ctypes.c_int), _I, _L, _D, _D]
        L.orc_prodvec.argtypes = [_L] + [_D] * 9
        L.orc_fc_filter_k.argtypes = [_I, _I, _L, _D, ctypes.c_double, _D]
        L.orc_set_threads.restype = _I
        L.orc_set_threads.argtypes = [_I]
        L.orc_neumann_table.argtypes = [_I, _I, ctypes.c_double, _D, _D, ctypes.c_double, _D]
        L.orc_fc_reconstruct.restype = _I
        L.orc_fc_reconstruct.argtypes = [_I, _I, _I, _I, _L, _I, _D, _D, _D]
        L.orc_nonlinear_uxw.argtypes = [_P, _P, ctypes.c_double, ctypes.c_double,
                                        _D, _D, _D, _D, _D, _I, _I]
        L.orc_nonlinear_coupled.argtypes = [_P, _P, ctypes.c_double, ctypes.c_double] + [_D] * 12 + [_I, _I]
        L.orc_derivr_mat.argtypes = [_P, _D, _I, _D]
        L.orc_invdl_setup.restype = _I
        L.orc_invdl_setup.argtypes = [_I, _I, _D, _D, _D, _D]
        L.orc_invdl.argtypes = [_I, _I, _I, _D, _D, _D]
        _lib = L
    return _lib


def _d(a):
    assert a.flags["C_CONTIGUOUS"]
    assert a.dtype in (np.float64, np.complex128)
    return a.ctypes.data_as(_D)


def set_threads(n):
    """OpenMP threads of the oracle's loops (0: the default); returns the count."""
    return lib().orc_set_threads(int(n))


def nlm_of(lmax):
    return (lmax + 1) * (lmax + 2) // 2


def lm_index(lmax, l, m):
    """0-based position of (l, m) in the m-major order."""
    return m * (2 * lmax + 1 - m) // 2 + l


def ell_emm(lmax):
    ell = np.empty(nlm_of(lmax), np.int32)
    emm = np.empty(nlm_of(lmax), np.int32)
    for m in range(lmax + 1):
        for l in range(m, lmax + 1):
            ell[lm_index(lmax, l, m)] = l
            emm[lm_index(lmax, l, m)] = m
    return ell, emm


def grid_nlat(lmax):
    """nlat from the 3/2 rule of grid_init (pseudospec_mod.f90:66-68)."""
    return 4 * ((3 * (lmax + 1) + 7) // 8)


class SHT:
    """Oracle spherical-harmonic transform on one (lmax, nlat, nphi) grid."""

    def __init__(self, lmax, nlat, nphi):
        self.lmax, self.nlat, self.nphi = lmax, nlat, nphi
        self.nlm = nlm_of(lmax)
        self.h = lib().orc_sht_create(lmax, nlat, nphi)
        if not self.h:
            raise ValueError("bad SHT sizes")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_sht_destroy(self.h)
            self.h = None

    def grid(self):
        ct, st, wt = (np.empty(self.nlat) for _ in range(3))
        lib().orc_sht_grid(self.h, _d(ct), _d(st), _d(wt))
        return ct, st, wt

    def tables(self):
        P, dP, mPs = (np.empty((self.nlm, self.nlat)) for _ in range(3))
        lib().orc_sht_tables(self.h, _d(P), _d(dP), _d(mPs))
        return P, dP, mPs

    def _spec(self, a):
        a = np.ascontiguousarray(a, np.complex128)
        if a.ndim == 1:
            a = a[None]
        assert a.shape[1] == self.nlm
        return a

    def _phys(self, f):
        f = np.ascontiguousarray(f, np.float64)
        if f.ndim == 2:
            f = f[None]
        assert f.shape[1:] == (self.nphi, self.nlat)
        return f

    def synth(self, a):
        a = self._spec(a)
        f = np.empty((a.shape[0], self.nphi, self.nlat))
        lib().orc_sh_to_spat(self.h, a.shape[0], _d(a), _d(f))
        return f

    def anal(self, f):
        f = self._phys(f)
        a = np.empty((f.shape[0], self.nlm), np.complex128)
        lib().orc_spat_to_sh(self.h, f.shape[0], _d(f), _d(a))
        return a

    def vsynth(self, q, s, t):
        q, s, t = self._spec(q), self._spec(s), self._spec(t)
        out = [np.empty((q.shape[0], self.nphi, self.nlat)) for _ in range(3)]
        lib().orc_shqst_to_spat(self.h, q.shape[0], _d(q), _d(s), _d(t), *[_d(o) for o in out])
        return out

    def vanal(self, vr, vt, vp):
        vr, vt, vp = self._phys(vr), self._phys(vt), self._phys(vp)
        out = [np.empty((vr.shape[0], self.nlm), np.complex128) for _ in range(3)]
        lib().orc_spat_to_shqst(self.h, vr.shape[0], _d(vr), _d(vt), _d(vp), *[_d(o) for o in out])
        return out


def load_tables(tdir, C, d):
    """A (C x d) and Q (d x d) in the Fortran stream layout of
    load_dirichlet_tables (fcgram_mod.f90:227-254)."""
    A = np.fromfile(os.path.join(tdir, f"A{C}-{d}.dat"), "<f8")
    Q = np.fromfile(os.path.join(tdir, f"Q{d}.dat"), "<f8")
    if A.size != C * d or Q.size != d * d:
        raise ValueError("FC-Gram table size mismatch")
    return A.reshape(d, C).T.copy(), Q.reshape(d, d).T.copy()


def fc_dir(A, Q):
    C, d = A.shape
    Af = np.asfortranarray(A).ravel(order="F").copy()
    Qf = np.asfortranarray(Q).ravel(order="F").copy()
    out = np.empty(C * d)
    lib().orc_fc_dir(C, d, _d(Af), _d(Qf), _d(out))
    return out.reshape(d, C).T.copy()


class FC:
    """Oracle radial plan: n total points, C continuation, d matching."""

    def __init__(self, n, C=0, d=0, dir_=None):
        self.n, self.C, self.d = n, C, d
        if C > 0:
            dv = np.asfortranarray(dir_).ravel(order="F").copy()
            self.h = lib().orc_fc_create(n, C, d, _d(dv))
        else:
            self.h = lib().orc_fc_create(n, 0, 0, None)
        if not self.h:
            raise ValueError("bad FC-Gram sizes")

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_fc_destroy(self.h)
            self.h = None

    def forward(self, a):
        """a: (n, ncols) complex, modified in place (continuation rows too)."""
        assert a.shape[0] == self.n and a.dtype == np.complex128 and a.flags["C_CONTIGUOUS"]
        lib().orc_z_forward(self.h, a.shape[1], a.shape[1], _d(a))
        return a

    def inverse(self, a):
        assert a.shape[0] == self.n and a.dtype == np.complex128 and a.flags["C_CONTIGUOUS"]
        lib().orc_z_inverse(self.h, a.shape[1], a.shape[1], _d(a))
        return a

    def derivr_r(self, a, kz, p):
        b = np.empty_like(a)
        lib().orc_derivr_r(self.h, _d(kz), p, a.shape[1], a.shape[1], _d(np.ascontiguousarray(a)), _d(b))
        return b

    def dlr(self, a, kz, z, ell):
        b = np.empty_like(a)
        ell = np.ascontiguousarray(ell, np.int32)
        lib().orc_dlr(self.h, _d(kz), _d(z), ell.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                      a.shape[1], a.shape[1], _d(np.ascontiguousarray(a)), _d(b))
        return b

    def derivr_mat(self, kz, p):
        N = self.n - self.C
        dm = np.empty((N, N), order="F")
        out = np.empty(N * N)
        lib().orc_derivr_mat(self.h, _d(kz), p, _d(out))
        dm[:] = out.reshape(N, N).T
        return np.ascontiguousarray(dm)


def kz(n, dz):
    out = np.empty(n)
    lib().orc_kz(n, dz, _d(out))
    return out


def derivr_k(a, kzv, p):
    b = np.empty_like(a)
    lib().orc_derivr_k(a.shape[0], a.shape[1], a.shape[1], _d(kzv), p, _d(np.ascontiguousarray(a)), _d(b))
    return b


def laplang(a, ell):
    b = np.empty_like(a)
    ell = np.ascontiguousarray(ell, np.int32)
    lib().orc_laplang(a.shape[0], a.shape[1], a.shape[1],
                      ell.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _d(np.ascontiguousarray(a)), _d(b))
    return b


def multr(a, z, p):
    b = np.empty_like(a)
    lib().orc_multr(a.shape[0], a.shape[1], a.shape[1], _d(z), p, _d(np.ascontiguousarray(a)), _d(b))
    return b


def saxpby_c(x, a, y, b):
    z = np.empty_like(x)
    lib().orc_saxpby_c(x.size, _d(z), _d(np.ascontiguousarray(x)), a, _d(np.ascontiguousarray(y)), b)
    return z


def prodvec(a, b):
    """a, b: tuples of 3 float64 arrays of equal shape -> 3 arrays."""
    a = [np.ascontiguousarray(x, np.float64) for x in a]
    b = [np.ascontiguousarray(x, np.float64) for x in b]
    c = [np.empty_like(a[0]) for _ in range(3)]
    lib().orc_prodvec(a[0].size, *[_d(x) for x in a + b + c])
    return c


def fc_filter_k(a, kzv, dkz):
    a = np.ascontiguousarray(a)
    lib().orc_fc_filter_k(a.shape[0], a.shape[1], a.shape[1], _d(kzv), dkz, _d(a))
    return a


def load_neumann_files(tdir, d, ord_):
    """Q (d x d) and (dxp, Qn) as load_neumann_tables reads them
    (fcgram_mod.f90:325-353): Q{d}.dat, then Q1n{d}.dat or Q2n{d}.dat with a
    leading FP64 dxp; stream FP64, column major."""
    Q = np.fromfile(os.path.join(tdir, f"Q{d}.dat"), "<f8")
    b = np.fromfile(os.path.join(tdir, f"Q{ord_}n{d}.dat"), "<f8")
    if Q.size != d * d or b.size != 1 + d * d:
        raise ValueError("Neumann table size mismatch")
    return Q.reshape(d, d).T.copy(), float(b[0]), b[1:].reshape(d, d).T.copy()


def neumann_table(d, ord_, dx, Q, Qn, dxp):
    """neu of load_neumann_tables (fcgram_mod.f90:355-363)."""
    Qf = np.ascontiguousarray(np.asfortranarray(Q).ravel(order="F"))
    Qnf = np.ascontiguousarray(np.asfortranarray(Qn).ravel(order="F"))
    neu = np.empty(d)
    lib().orc_neumann_table(d, ord_, dx, _d(Qf), _d(Qnf), dxp, _d(neu))
    return neu


def fc_reconstruct(a, C, d, boun, neu, coef=None):
    """neumann_reconstruct (coef None) or robin_reconstruct (coef: one real
    per column) along z on a (n, ncols) complex array; returns a copy."""
    a = np.ascontiguousarray(a, np.complex128).copy()
    neu = np.ascontiguousarray(neu, np.float64)
    cp = None if coef is None else _d(np.ascontiguousarray(coef, np.float64))
    rc = lib().orc_fc_reconstruct(a.shape[0], C, d, a.shape[1], a.shape[1], boun, _d(neu), cp, _d(a))
    if rc:
        raise ValueError("unsupported boundary")
    return a


def nonlinear_uxw(sht, fc, rint, dz, W, Z, plane0=0, nplanes=None):
    """N = u x omega as (NQ, NS, NT), each (n, nlm) complex; W, Z (n, nlm)."""
    n = fc.n
    if nplanes is None:
        nplanes = n - fc.C
    W = np.ascontiguousarray(W, np.complex128)
    Z = np.ascontiguousarray(Z, np.complex128)
    assert W.shape == (n, sht.nlm) and Z.shape == W.shape
    out = [np.empty_like(W) for _ in range(3)]
    lib().orc_nonlinear_uxw(sht.h, fc.h, rint, dz, _d(W), _d(Z), *[_d(o) for o in out], plane0, nplanes)
    return out


def nonlinear_coupled(sht, fc, rint, dz, W, Z, Pb, Tb, T, plane0=0, nplanes=None):
    """One coupled step: ((NQ, NS, NT) = u x omega, (MQ, MS, MT) = u x B,
    NH = u . grad T), each (n, nlm) complex; inputs (n, nlm) complex."""
    n = fc.n
    if nplanes is None:
        nplanes = n - fc.C
    ins = [np.ascontiguousarray(x, np.complex128) for x in (W, Z, Pb, Tb, T)]
    for x in ins:
        assert x.shape == (n, sht.nlm)
    out = [np.empty_like(ins[0]) for _ in range(7)]
    lib().orc_nonlinear_coupled(sht.h, fc.h, rint, dz, *[_d(x) for x in ins], *[_d(o) for o in out],
                                plane0, nplanes)
    return out


def invdl_setup(N, lmax, d1, d2, z):
    binv = np.zeros((lmax, N - 4, N))
    d1f = np.asfortranarray(d1).ravel(order="F").copy()
    d2f = np.asfortranarray(d2).ravel(order="F").copy()
    zz = np.ascontiguousarray(z, np.float64)
    rc = lib().orc_invdl_setup(N, lmax, _d(d1f), _d(d2f), _d(zz), _d(binv))
    if rc:
        raise RuntimeError(f"LU of B_l failed for l = {rc}")
    return binv


def invdl(N, n, lmax, binv, phi):
    phi = np.ascontiguousarray(phi, np.complex128)
    out = np.empty_like(phi)
    lib().orc_invdl(N, n, lmax, _d(binv), _d(phi), _d(out))
    return out


def dirichlet(a, N):
    """dirichlet (boundary_mod.f90:62-93): rows 1 and N = nz - C set to zero."""
    a = np.array(a, np.complex128, copy=True)
    a[0] = 0
    a[N - 1] = 0
    return a


def d2bor(d2):
    """Rows 1, 2, N-1, N of d2/dr2 (boundary_mod.f90:230-235), shape (4, N)."""
    N = d2.shape[0]
    return np.ascontiguousarray(d2[[0, 1, N - 2, N - 1]])


def v_imposebc(fc, kz, z, ell, binv, lmax, tor, phi):
    """vboundary.f90:59-88: Z = 0 at the walls, W = M_l phi, phi = D_l W.
    Returns (tor, phi, pol)."""
    N = fc.n - fc.C
    tor = dirichlet(tor, N)
    pol = invdl(N, fc.n, lmax, binv, phi)
    return tor, fc.dlr(pol, kz, z, ell), pol


def v_imposebc_fast(fc, kz, z, ell, binv, d2b, lmax, tor, phi):
    """vboundary.f90:91-130: phi recomputed on rows 1, 2, N-1, N only, l = 0
    mode of phi zero.  Returns (tor, phi, pol)."""
    N = fc.n - fc.C
    tor = dirichlet(tor, N)
    pol = invdl(N, fc.n, lmax, binv, phi)
    phi = np.array(phi, np.complex128, copy=True)
    ll = ell * (ell + 1.0)
    for k, i in enumerate((0, 1, N - 2, N - 1)):
        phi[i] = d2b[k] @ pol[:N] - ll * pol[i] / z[i] ** 2
    phi[:, ell == 0] = 0
    return tor, phi, pol


def fc_filter_r(fc, a, kzv, dkz):
    """fc_filter_r (pseudospec_hd.f90:443-467): forward z, fc_filter_k, inverse z."""
    return fc.inverse(fc_filter_k(fc.forward(np.array(a, np.complex128, copy=True)), kzv, dkz))


def stage_uxw(sht, fc, rint, dz, kzv, z, ell, binv, d2b, lmax, o, dt, nu, filt, Z0, phi0, Z, phi, W):
    """One stage a = a0 + dt/o rhs(a) of the explicit Runge-Kutta scheme for
    the toroidal potential Z and phi = D_l W of u = curl curl (W r) + curl (Z r):
      rhs_Z   = r N_T + nu D_l Z,
      rhs_phi = d(r N_S)/dr - N_Q + nu D_l phi,
    N = u x omega (nonlinear_uxw) from (W, Z); then fc_filter along r (filt)
    and v_imposebc_fast (vboundary.f90:91-130).  Returns (Z, phi, W)."""
    NQ, NS, NT = nonlinear_uxw(sht, fc, rint, dz, W, Z)
    c = dt / o
    zz = z[:, None]
    Zn = Z0 + c * (zz * NT + nu * fc.dlr(Z, kzv, z, ell))
    phin = phi0 + c * (fc.derivr_r(np.ascontiguousarray(zz * NS), kzv, 1) - NQ + nu * fc.dlr(phi, kzv, z, ell))
    if filt:
        dkz = 2 * np.pi / (dz * fc.n)
        Zn = fc_filter_r(fc, Zn, kzv, dkz)
        phin = fc_filter_r(fc, phin, kzv, dkz)
    return v_imposebc_fast(fc, kzv, z, ell, binv, d2b, lmax, Zn, phin)


# ---------------------------------------------------------------------------
# Extended-precision build of the same source (liboracle_ld.so): every step in
# x87 long double (64-bit significand) on the same FP64 inputs and FC-Gram
# tables.  Used as a higher-precision truth when two FP64 implementations
# (this oracle and the GPU path) disagree by more than their rounding.
_LIB_LD = os.path.join(_HERE, "_build", "liboracle_ld.so")
_lib_ld = None
_DL = ctypes.POINTER(ctypes.c_longdouble)
_LDT = ctypes.c_longdouble


def lib_ld():
    global _lib_ld
    if _lib_ld is None:
        if not os.path.exists(_LIB_LD):
            build()
        L = ctypes.CDLL(_LIB_LD)
        L.orc_sht_create.restype = _P
        L.orc_sht_create.argtypes = [_I, _I, _I]
        L.orc_sht_destroy.argtypes = [_P]
        L.orc_sht_grid.argtypes = [_P, _DL, _DL, _DL]
        L.orc_fc_dir.argtypes = [_I, _I, _DL, _DL, _DL]
        L.orc_fc_create.restype = _P
        L.orc_fc_create.argtypes = [_I, _I, _I, _DL]
        L.orc_fc_destroy.argtypes = [_P]
        L.orc_nonlinear_uxw.argtypes = [_P, _P, _LDT, _LDT, _DL, _DL, _DL, _DL, _DL, _I, _I]
        _lib_ld = L
    return _lib_ld


def _dl(a):
    assert a.flags["C_CONTIGUOUS"] and a.dtype in (np.longdouble, np.clongdouble)
    return a.ctypes.data_as(_DL)


def nonlinear_uxw_ld(lmax, nlat, nphi, A, Q, rint, dz, W, Z, plane0=0, nplanes=None):
    """u x omega as in nonlinear_uxw, computed in extended precision from the
    FP64 inputs W, Z (n, nlm) and the FP64 tables A (C x d), Q (d x d).
    Returns (NQ, NS, NT) as clongdouble arrays; rows outside
    plane0 .. plane0 + nplanes - 1 are zero."""
    L = lib_ld()
    C, d = A.shape
    n = W.shape[0]
    if nplanes is None:
        nplanes = n - C
    Af = np.ascontiguousarray(np.asfortranarray(A).ravel(order="F"), np.longdouble)
    Qf = np.ascontiguousarray(np.asfortranarray(Q).ravel(order="F"), np.longdouble)
    dr = np.empty(C * d, np.longdouble)
    L.orc_fc_dir(C, d, _dl(Af), _dl(Qf), _dl(dr))
    # orc_fc_create takes dir column major, as orc_fc_dir writes it
    fc = L.orc_fc_create(n, C, d, _dl(dr))
    sht = L.orc_sht_create(lmax, nlat, nphi)
    if not fc or not sht:
        raise ValueError("bad sizes")
    Wl = np.ascontiguousarray(W, np.clongdouble)
    Zl = np.ascontiguousarray(Z, np.clongdouble)
    out = [np.empty_like(Wl) for _ in range(3)]
    try:
        L.orc_nonlinear_uxw(sht, fc, _LDT(rint), _LDT(dz), _dl(Wl), _dl(Zl), *[_dl(o) for o in out],
                            plane0, nplanes)
    finally:
        L.orc_sht_destroy(sht)
        L.orc_fc_destroy(fc)
    return out
