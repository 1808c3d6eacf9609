"""Shared test inputs: the reference's known-answer fields and seeded spectra."""
import numpy as np

PI = np.pi


def lm_index(lmax, l, m):
    return m * (2 * lmax + 1 - m) // 2 + l


def nlm_of(lmax):
    return (lmax + 1) * (lmax + 2) // 2


def ell_of(lmax):
    ell = np.empty(nlm_of(lmax), np.int32)
    for m in range(lmax + 1):
        for l in range(m, lmax + 1):
            ell[lm_index(lmax, l, m)] = l
    return ell


def emm_of(lmax):
    emm = np.empty(nlm_of(lmax), np.int32)
    for m in range(lmax + 1):
        for l in range(m, lmax + 1):
            emm[lm_index(lmax, l, m)] = m
    return emm


def rell2(x, y):
    x, y = np.asarray(x), np.asarray(y)
    return float(np.linalg.norm((x - y).ravel()) / max(np.linalg.norm(y.ravel()), 1e-300))


def rell2_joint(xs, ys):
    """Relative L2 of several outputs taken as one vector."""
    num = sum(float(np.linalg.norm((np.asarray(x) - np.asarray(y)).ravel())) ** 2 for x, y in zip(xs, ys))
    den = sum(float(np.linalg.norm(np.asarray(y).ravel())) ** 2 for y in ys)
    return (num / max(den, 1e-300)) ** 0.5


def field_bounds(refs, floors, tol, factor=30.0):
    """Per-field relative-L2 bounds for outputs refs (one group) against an
    FP64 reference: tol, or factor x the field's noise floor, or the relative
    error at which the field's absolute error reaches a tenth of the group's
    joint tolerance, whichever is largest.  The last term covers fields far
    smaller than their group (the toroidal part of a near-gradient u x w is
    1e4 times smaller than the poloidal part at lmax 300): their rounding
    noise is that of the group, which one-ulp input noise does not show."""
    norms = [float(np.linalg.norm(np.asarray(r).ravel())) for r in refs]
    group = sum(x * x for x in norms) ** 0.5
    return [max(tol, factor * f, 0.1 * tol * group / max(nm, 1e-300)) for nm, f in zip(norms, floors)]


def noise_floor(fn, inputs, seed=0, rel=2.0 ** -52):
    """FP64 conditioning of fn at these inputs: the relative L2 change of each
    output when every input gets independent relative noise of one ulp.  Two
    correct FP64 implementations with different rounding differ by about this."""
    rng = np.random.default_rng(seed)
    base = fn(*inputs)
    pert = [x * (1 + rel * rng.standard_normal(x.shape)) if np.iscomplexobj(x) is False
            else x.real * (1 + rel * rng.standard_normal(x.shape)) + 1j * x.imag * (1 + rel * rng.standard_normal(x.shape))
            for x in inputs]
    out = fn(*pert)
    if isinstance(base, (list, tuple)):
        return [rell2(o, b) for o, b in zip(out, base)]
    return rell2(out, base)


def sectoral_c(l):
    """c = sqrt((2l+1)/(4 pi)) (2l-1)!!/sqrt((2l)!) (src/prueba_sphp.f90:165-174)."""
    c = np.sqrt((2 * l + 1) / (4 * PI))
    for k in range(1, l + 1):
        c *= np.sqrt((2 * k - 1.0) / (2 * k))
    return c


def legendre_p(l, x):
    p0, p1 = np.ones_like(x), x
    if l == 0:
        return p0
    for k in range(1, l):
        p0, p1 = p1, ((2 * k + 1) * x * p1 - k * p0) / (k + 1)
    return p1


def scalar_case(ic, lmax, cth, sth, ph):
    """Field f(phi, theta) and exact coefficients a of prueba_sphp case ic
    (src/prueba_sphp.f90:121-184); arrays are (nphi, nlat)."""
    a = np.zeros(nlm_of(lmax), complex)
    L = lambda l, m: lm_index(lmax, l, m)
    if ic == 1:
        f = np.ones_like(cth); a[L(0, 0)] = np.sqrt(4 * PI)
    elif ic == 2:
        f = cth.copy(); a[L(1, 0)] = np.sqrt(4 * PI / 3)
    elif ic == 3:
        f = sth * np.cos(ph); a[L(1, 1)] = np.sqrt(2 * PI / 3)
    elif ic == 4:
        f = sth * np.sin(ph); a[L(1, 1)] = -1j * np.sqrt(2 * PI / 3)
    elif ic == 5:
        f = 3 * cth ** 2 - 1; a[L(2, 0)] = 4 * np.sqrt(PI / 5)
    elif ic == 6:
        f = sth ** 2 * np.cos(2 * ph); a[L(2, 2)] = 2 * np.sqrt(2 * PI / 15)
    elif ic == 7:
        f = legendre_p(lmax, cth); a[L(lmax, 0)] = np.sqrt(4 * PI / (2 * lmax + 1))
    elif ic == 8:
        f = sth ** lmax * np.cos(lmax * ph); a[L(lmax, lmax)] = 1 / (2 * sectoral_c(lmax))
    elif ic == 9:
        f = 2 - 3 * cth + sth * np.sin(ph) + 0.5 * sth ** 2 * np.cos(2 * ph)
        a[L(0, 0)] = 2 * np.sqrt(4 * PI); a[L(1, 0)] = -3 * np.sqrt(4 * PI / 3)
        a[L(1, 1)] = -1j * np.sqrt(2 * PI / 3); a[L(2, 2)] = np.sqrt(2 * PI / 15)
    else:
        raise ValueError(ic)
    return f, a


def vector_case(ic, lmax, cth, sth, ph):
    """(fr, ft, fp) and exact (aq, as, at) of prueba_vec case ic
    (src/prueba_vec.f90:139-209)."""
    n = nlm_of(lmax)
    aq, as_, at = np.zeros(n, complex), np.zeros(n, complex), np.zeros(n, complex)
    z = np.zeros_like(cth)
    fr, ft, fp = z.copy(), z.copy(), z.copy()
    L = lambda l, m: lm_index(lmax, l, m)
    if ic == 1:
        fr = cth.copy(); aq[L(1, 0)] = np.sqrt(4 * PI / 3)
    elif ic == 2:
        ft = -sth; as_[L(1, 0)] = np.sqrt(4 * PI / 3)
    elif ic == 3:
        fp = sth.copy(); at[L(1, 0)] = np.sqrt(4 * PI / 3)
    elif ic == 4:
        ft = cth * np.cos(ph); fp = -np.sin(ph); as_[L(1, 1)] = np.sqrt(2 * PI / 3)
    elif ic == 5:
        ft = -np.sin(ph); fp = -cth * np.cos(ph); at[L(1, 1)] = np.sqrt(2 * PI / 3)
    elif ic == 6:
        fr = legendre_p(lmax, cth); aq[L(lmax, 0)] = np.sqrt(4 * PI / (2 * lmax + 1))
    elif ic == 7:
        ft = -lmax * sth ** (lmax - 1) * np.sin(lmax * ph)
        fp = -lmax * sth ** (lmax - 1) * cth * np.cos(lmax * ph)
        at[L(lmax, lmax)] = 1 / (2 * sectoral_c(lmax))
    elif ic == 8:
        fr = cth.copy(); ft = cth * np.cos(ph) - np.sin(ph); fp = -np.sin(ph) - cth * np.cos(ph)
        aq[L(1, 0)] = np.sqrt(4 * PI / 3); as_[L(1, 1)] = np.sqrt(2 * PI / 3); at[L(1, 1)] = np.sqrt(2 * PI / 3)
    else:
        raise ValueError(ic)
    return (fr, ft, fp), (aq, as_, at)


def grid_arrays(ct, st, nphi):
    """(cth, sth, ph) of shape (nphi, nlat) for the theta-contiguous planes."""
    ph = 2 * PI * np.arange(nphi) / nphi
    cth = np.broadcast_to(ct[None, :], (nphi, ct.size)).copy()
    sth = np.broadcast_to(st[None, :], (nphi, st.size)).copy()
    phg = np.broadcast_to(ph[:, None], (nphi, ct.size)).copy()
    return cth, sth, phg


def random_spectrum(lmax, nplanes, seed, slope=0.0, zero_l0=False):
    """Seeded complex normal coefficients, scaled by (l+1)^-slope, m = 0 real."""
    rng = np.random.default_rng(seed)
    n = nlm_of(lmax)
    a = rng.standard_normal((nplanes, n)) + 1j * rng.standard_normal((nplanes, n))
    a *= (ell_of(lmax) + 1.0) ** (-slope)
    a[:, : lmax + 1] = a[:, : lmax + 1].real
    if zero_l0:
        a[:, 0] = 0
    return a


def shell_fields(lmax, nr, n, rint, rext, seed, slope=2.0):
    """W, Z (n, nlm) in (lm, r): c_lm (l+1)^-slope sin(pi (r - ri)/(ro - ri)) on the
    nr physical shells, l = 0 zero, m = 0 real; continuation rows zero."""
    rng = np.random.default_rng(seed)
    nl = nlm_of(lmax)
    r = rint + (rext - rint) * np.arange(nr) / (nr - 1)
    prof = np.sin(PI * (r - rint) / (rext - rint))
    out = []
    for f in range(2):
        c = rng.standard_normal(nl) + 1j * rng.standard_normal(nl)
        c *= (ell_of(lmax) + 1.0) ** (-slope)
        c[: lmax + 1] = c[: lmax + 1].real
        c[0] = 0
        a = np.zeros((n, nl), complex)
        a[:nr] = prof[:, None] * c[None, :]
        out.append(a)
    return out


def coupled_fields(lmax, nr, n, rint, rext, seed):
    """The coupled state (W, Z, Pb, Tb, T), each (n, nlm) in (lm, r): velocity
    potentials with spectrum (l+1)^-2, magnetic potentials (l+1)^-2.5 and a
    temperature (l+1)^-2 with its l = 0 mode, all with the sin radial profile."""
    W, Z = shell_fields(lmax, nr, n, rint, rext, seed, slope=2.0)
    Pb, Tb = shell_fields(lmax, nr, n, rint, rext, seed + 1, slope=2.5)
    rng = np.random.default_rng(seed + 2)
    r = rint + (rext - rint) * np.arange(nr) / (nr - 1)
    c = (rng.standard_normal(nlm_of(lmax)) + 1j * rng.standard_normal(nlm_of(lmax))) * (ell_of(lmax) + 1.0) ** -2
    c[: lmax + 1] = c[: lmax + 1].real
    T = np.zeros((n, nlm_of(lmax)), complex)
    T[:nr] = (1.0 + r[:, None]) * np.sin(PI * (r - rint) / (rext - rint))[:, None] * c[None, :]
    return W, Z, Pb, Tb, T


# ------------------------------------------------------------ boundary
def boundary_inputs(z, ell, N):
    """The fields of src/prueba_boundary.f90:58-78 on the rows of z: Z0, the
    admissible W0 = c (r - r1)^2 (r - r2)^2 e^r and c_lm per mode."""
    nlm = ell.size
    lm = np.arange(1, nlm + 1)
    cl = (1.0 + 0.1 * lm) + 1j * (0.5 - 0.05 * lm)
    r1, r2 = z[0], z[N - 1]
    z0 = cl[None] * (1.0 + np.exp(z) * np.sin(3 * z))[:, None]
    w0 = cl[None] * ((z - r1) ** 2 * (z - r2) ** 2 * np.exp(z))[:, None]
    return z0, w0


def boundary_stage_phi(f0, ell, N):
    """phi as a Runge-Kutta stage leaves it: wall rows and the l = 0 mode
    disturbed (src/prueba_boundary.f90:80-89)."""
    fi = np.array(f0, np.complex128, copy=True)
    fi[0] += 1.0
    fi[1] -= 2.0
    fi[N - 2] += 3.0
    fi[N - 1] -= 4.0
    fi[:, ell == 0] = 5.0 + 1.0j
    return fi


def boundary_errs(z, ell, N, d1, z0, w0, f0, tor, fi, pol, pol2, tor3, fi3, pol3):
    """The ten checks of src/prueba_boundary.f90:92-128, in order."""
    err = np.zeros(10)
    l0 = ell == 0
    nl = ~l0
    r1, r2 = z[0], z[N - 1]
    err[0] = max(np.abs(tor[0]).max(), np.abs(tor[N - 1]).max())
    err[1] = max(np.abs(tor[1:N - 1] - z0[1:N - 1]).max(), np.abs(tor[N:] - z0[N:]).max())
    err[4] = max(np.abs(fi[:, l0]).max(), np.abs(pol[:, l0]).max())
    wm = np.abs(pol[:N, nl]).max(axis=0)
    err[2] = (np.maximum(np.abs(pol[0, nl]), np.abs(pol[N - 1, nl])) / wm).max()
    dw = d1 @ pol[:N, nl]
    err[3] = (np.maximum(np.abs(dw[0]), np.abs(dw[N - 1])) * (r2 - r1) / wm).max()
    err[5] = (np.abs(fi[2:N - 2, nl] - f0[2:N - 2, nl]).max(axis=0) / np.abs(f0[2:N - 2, nl]).max(axis=0)).max()
    err[7] = (np.abs(pol[:N, nl] - w0[:N, nl]).max(axis=0) / np.abs(w0[:N, nl]).max(axis=0)).max()
    err[6] = np.abs(pol2 - pol).max() / np.abs(pol).max()
    err[8] = max(np.abs(pol3 - pol).max(), np.abs(tor3 - tor).max())
    err[9] = np.abs(fi3[:N] - fi[:N]).max() / np.abs(fi[:N]).max()
    return err


# published levels of prueba_boundary (LEEME.md:45) with margins for the
# FC-Gram tables built here: walls exact, dW/dr 1e-14, interior phi 3e-12,
# twice 2e-13, W vs W0 6e-8 (truncation), fast == full, fast phi 3e-12
BOUNDARY_LIMITS = (0.0, 0.0, 1e-13, 1e-12, 0.0, 3e-11, 2e-12, 6e-7, 0.0, 3e-11)


def l_bands(lmax):
    """Degree bands [lo, hi) for error reports: 0-15, 16-31, then octaves."""
    edges = [0, 16, 32]
    while edges[-1] <= lmax:
        edges.append(edges[-1] * 2)
    edges[-1] = lmax + 1
    return [(a, b) for a, b in zip(edges[:-1], edges[1:]) if a <= lmax]


def band_errors(out, ref, lmax, nr, names=("NQ", "NS", "NT")):
    """Relative L2 error of each output field (named by names) on the nr
    physical shells, over all degrees and per degree band."""
    ell = ell_of(lmax)
    rep = {}
    for name, o, r in zip(names, out, ref):
        o, r = o[:nr], r[:nr]
        d = {"all": rell2(o, r)}
        for a, b in l_bands(lmax):
            sel = (ell >= a) & (ell < b)
            d[f"l{a}-{b - 1}"] = rell2(o[:, sel], r[:, sel])
        rep[name] = d
    rep["joint"] = rell2_joint([o[:nr] for o in out], [r[:nr] for r in ref])
    return rep


# ---------------------------------------------------------- fc_filter
FILTER_ALPHA = 16.0 * np.log(10.0)


def filter_multiplier_mp(k, n, p=50):
    """exp(-alpha (2 k / n)^(2p)) in 40-digit arithmetic, alpha = 16 ln 10
    (pseudospec_hd.f90:425-437, with kz / Dkz = k, the FFT-ordered bin)."""
    import mpmath as mp
    mp.mp.dps = 40
    x = mp.mpf(2 * int(k)) / n
    return float(mp.exp(-16 * mp.log(10) * x ** (2 * p)))


def filter_check(mult, n, dz):
    """True when `mult` (the per-bin factor of a filter) agrees with the
    closed form at every bin and is 1 exactly at kz = 0.  The bound follows
    the formula's conditioning: x = 2 kz/(n Dkz) carries a few ulps (kz/Dkz is
    not exactly k in FP64), and d ln f / d ln x = -100 alpha x^100, so four
    ulps of x give 4 eps 100 alpha x^100 relative in f (up to 3e-13 at the
    Nyquist bin), plus 1e-14."""
    k = np.rint(_kz(n, dz) / (2 * np.pi / (dz * n))).astype(int)
    ref = np.array([filter_multiplier_mp(kk, n) for kk in k])
    x = 2.0 * k / n
    tol = (1e-14 + 4 * 2.2e-16 * 100 * FILTER_ALPHA * x ** 100) * ref + 1e-30
    return mult[0] == 1.0 and bool(np.all(np.abs(mult - ref) <= tol))


def _kz(n, dz):
    from oracle import oracle as O
    return O.kz(n, dz)
