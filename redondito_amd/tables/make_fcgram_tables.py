"""Builds the FC-Gram Dirichlet continuation tables A{C}-{d}.dat and Q{d}.dat.

The reference loads these two tables from an external directory
(src/sphp/fcgram_mod.f90:186-263) and forms the continuation operator
dir = A . Q^T (fcgram_mod.f90:257-258).  The files are Fortran unformatted
stream FP64, column major: A is C x d, Q is d x d.

Construction (FC-Gram, Bruno & Lyon 2010; Albin & Bruno 2011), written here
from the published method and run in 60-digit arithmetic because the
least-squares problem is ill-conditioned in FP64:

  * matching points x_j = j, j = 0..d-1 (unit spacing, x_{d-1} is the wall);
  * Q[:, k] = the k-th discrete Gram polynomial on those points (QR of the
    Vandermonde matrix, so Q is orthogonal and Q^T f are the coefficients
    of the degree-<d interpolant of f);
  * for every Gram polynomial g_k, a trigonometric polynomial T_k of period
    P is fitted in least squares (truncated SVD) to g_k on the oversampled
    matching interval [0, d-1] and to zero on the oversampled interval
    [d+C, 2d+C-1], which is where the opposite wall's matching points sit
    once the sequence is wrapped periodically;
  * A[ii-1, k] = T_k(d-1+ii), ii = 1..C: the blend-to-zero continuation of
    g_k sampled on the C continuation points.

Neumann reconstructors Q1n{d}.dat and Q2n{d}.dat (read by
load_neumann_tables, fcgram_mod.f90:267-371: a leading FP64 dxp, then Qn,
d x d, column major).  The reference reconstructs the wall value as
f_d = Q(d,:) Qn^T v with v = (f_1, ..., f_{d-1}, g), g the ord-th derivative
at the wall x_d: so Qn^T maps v to the Gram coefficients of the degree-<d
polynomial p with p(x_j) = f_j (j < d) and p^(ord)(x_d) = g, i.e.
Qn = M^-T with M[j, k] = g_k(x_j) for j < d and M[d-1, k] = g_k^(ord)(x_{d-1}).
The matching points are at unit spacing, so dxp = 1.

Usage:  python make_fcgram_tables.py [C] [d] [outdir]
        python make_fcgram_tables.py --neumann [d] [outdir]
"""
import os
import struct
import sys

import mpmath as mp

mp.mp.dps = 60

# Oversampling of the fitting intervals and width of the auxiliary period.
N_OS = 20


def gram_basis(d):
    """Orthonormal discrete Gram polynomials on x = 0..d-1.

    Returns (Q, coef) with Q[j][k] = g_k(x_j) and coef[i][k] the monomial
    coefficients of g_k (g_k(t) = sum_i coef[i][k] t^i).
    """
    # Modified Gram-Schmidt on the monomials, carrying the coefficients.
    cols, coefs = [], []
    for k in range(d):
        v = [mp.mpf(j) ** k for j in range(d)]
        c = [mp.mpf(0)] * d
        c[k] = mp.mpf(1)
        for q, qc in zip(cols, coefs):
            r = mp.fsum(a * b for a, b in zip(q, v))
            v = [a - r * b for a, b in zip(v, q)]
            c = [a - r * b for a, b in zip(c, qc)]
        nrm = mp.sqrt(mp.fsum(a * a for a in v))
        cols.append([a / nrm for a in v])
        coefs.append([a / nrm for a in c])
    Q = [[cols[k][j] for k in range(d)] for j in range(d)]
    coef = [[coefs[k][i] for k in range(d)] for i in range(d)]
    return Q, coef


def poly(coefk, t):
    return mp.fsum(c * t ** i for i, c in enumerate(coefk))


def blend_tables(C, d):
    Q, coef = gram_basis(d)
    period = mp.mpf(2 * (d + C))
    nmodes = (d + C) // 2           # highest harmonic, below a quarter of the unit-grid Nyquist
    fit_t, is_match = [], []
    for j in range((d - 1) * N_OS + 1):
        fit_t.append(mp.mpf(j) / N_OS)
        is_match.append(True)
    for j in range((d - 1) * N_OS + 1):
        fit_t.append(mp.mpf(d + C) + mp.mpf(j) / N_OS)
        is_match.append(False)
    ncol = 2 * nmodes + 1
    M = mp.matrix(len(fit_t), ncol)
    w = 2 * mp.pi / period
    for i, t in enumerate(fit_t):
        M[i, 0] = 1
        for q in range(1, nmodes + 1):
            M[i, 2 * q - 1] = mp.cos(q * w * t)
            M[i, 2 * q] = mp.sin(q * w * t)
    U, S, V = mp.svd_r(M)
    smax = max(S[i] for i in range(len(S)))
    A = [[None] * d for _ in range(C)]
    resid = []
    for k in range(d):
        coefk = [coef[i][k] for i in range(d)]
        y = mp.matrix([poly(coefk, t) if m else mp.mpf(0) for t, m in zip(fit_t, is_match)])
        # truncated-SVD least squares
        uty = U.T * y
        x = mp.matrix(ncol, 1)
        for s in range(len(S)):
            if S[s] > smax * mp.mpf('1e-40'):
                coeff = uty[s] / S[s]
                for c in range(ncol):
                    x[c] += V[s, c] * coeff
        r = M * x - y
        resid.append(max(abs(r[i]) for i in range(len(fit_t))))

        def T(t):
            v = x[0]
            for q in range(1, nmodes + 1):
                v += x[2 * q - 1] * mp.cos(q * w * t) + x[2 * q] * mp.sin(q * w * t)
            return v

        for ii in range(1, C + 1):
            A[ii - 1][k] = T(mp.mpf(d - 1 + ii))
    return A, Q, resid


def neumann_tables(d, ord_):
    """Qn (d x d) of the reconstructor of order ord_ (1: first, 2: second
    derivative at the wall x_{d-1})."""
    Q, coef = gram_basis(d)
    t = mp.mpf(d - 1)
    M = mp.matrix(d, d)
    for k in range(d):
        for j in range(d - 1):
            M[j, k] = Q[j][k]
        # ord_-th derivative of g_k(t) = sum_i coef[i][k] t^i at the wall
        M[d - 1, k] = mp.fsum(coef[i][k] * mp.ff(i, ord_) * t ** (i - ord_) for i in range(ord_, d))
    Qn = (M ** -1).T
    return [[Qn[r, c] for c in range(d)] for r in range(d)]


def write_neumann(path, Qn, d, dxp=1.0):
    with open(path, "wb") as fh:
        fh.write(struct.pack("<d", dxp))
        for c in range(d):
            for r in range(d):
                fh.write(struct.pack("<d", float(Qn[r][c])))


def write_stream(path, mat, nrow, ncol):
    """Fortran unformatted stream, column major, FP64 little endian."""
    with open(path, "wb") as fh:
        for c in range(ncol):
            for r in range(nrow):
                fh.write(struct.pack("<d", float(mat[r][c])))


def main():
    if len(sys.argv) > 1 and sys.argv[1] == "--neumann":
        d = int(sys.argv[2]) if len(sys.argv) > 2 else 5
        out = sys.argv[3] if len(sys.argv) > 3 else os.path.dirname(os.path.abspath(__file__))
        for o in (1, 2):
            write_neumann(os.path.join(out, f"Q{o}n{d}.dat"), neumann_tables(d, o), d)
        return
    C = int(sys.argv[1]) if len(sys.argv) > 1 else 25
    d = int(sys.argv[2]) if len(sys.argv) > 2 else 5
    out = sys.argv[3] if len(sys.argv) > 3 else os.path.dirname(os.path.abspath(__file__))
    A, Q, resid = blend_tables(C, d)
    write_stream(os.path.join(out, f"A{C}-{d}.dat"), A, C, d)
    write_stream(os.path.join(out, f"Q{d}.dat"), Q, d, d)
    print("fit residual per Gram polynomial:", [mp.nstr(r, 3) for r in resid])


if __name__ == "__main__":
    main()
