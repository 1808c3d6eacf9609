// Exact-accumulation form of the Legendre contractions (RD_PREC_OZAKI):
// Ozaki-style splitting of FP64 operands into int8 digits, products on the
// int8 MFMA with exact int32 accumulation, recombination in FP64.
//
// Every contraction of one order m is written as real GEMMs C = A B:
//   synthesis  (rows theta, K over l, columns = plane x {re, im})
//     f0: sym  Pe Q_e                 anti Po Q_o
//     f1: sym [Do|Me][S_o; iT_e]      anti [De|Mo][S_e; iT_o]
//    -f2: sym [Do|Me][T_o; -iS_e]     anti [De|Mo][T_e; -iS_o]
//   north = sym + anti, south = sym - anti (the FP64 kernels' symmetric and
//   antisymmetric accumulators);
//   analysis   (rows l, K over the northern thetas, columns as above)
//     Q_e = Pe r+, Q_o = Po r-, S_e = [De|Me][t-; -ip+], S_o = [Do|Mo][t+; -ip-],
//     T_e = [De|Me][-p-; -it+], T_o = [Do|Mo][-p+; -it-]
//   with x+- = x_north +- x_south.
// So each direction has four distinct A matrices per m (tables, sliced once
// per plan) and six B matrices (sliced per call by a memory-bound kernel).
//
// Splitting: a row of A (or a column of B) with largest magnitude M gets the
// exponent e = ilogb(M) + 2, so |x| 2^-e < 1/2, and OZ_S base-128 digits
// d_p = rint(128 r), r <- 128 r - d_p, all in [-64, 64] (negation is safe in
// int8).  The digit products d_p(A) d_q(B) with p + q <= OZ_S - 1 (0-based)
// are accumulated exactly in int32 per level p + q, and
//   C = 2^(eA + eB) sum_L 128^-(L+2) acc_L.
// The truncation of each operand at OZ_S digits bounds the error by
// 2^-(7 OZ_S + 1) of the row or column maximum per element.  The tables are
// sliced to OZ_S = 7 digits once; a call uses the plan's digit count D
// (5, 6 or 7: the first D digit planes of the tables and operands, the
// products with p + q <= D - 1), for an element error of 2^-(7 D + 1).
#include "rd_internal.h"

#include <algorithm>
#include <vector>

namespace rd {

namespace {

constexpr int OZ_S = 7;        // digits per operand
constexpr int OZ_CT = 64;      // columns per workgroup tile

typedef int v4i __attribute__((ext_vector_type(4)));

__device__ __forceinline__ int oz_exponent(double mx)
{
    return mx > 0.0 ? ilogb(mx) + 2 : 0;
}

// The first D digits of x 2^-e (|x 2^-e| < 1/2) into d[0..D).
template <int D>
__device__ __forceinline__ void oz_digits(double x, int e, int8_t (&d)[D])
{
    double r = ldexp(x, -e);
#pragma unroll
    for (int p = 0; p < D; p++) {
        r *= 128.0;
        const double q = rint(r);
        r -= q;
        d[p] = (int8_t)(int)q;
    }
}

// Slices of one matrix: [OZ_S][rows][K] int8 from off, K a multiple of 64,
// split = start of the second part of a concatenated K (or K).
struct OzMat {
    long off;
    int K, split, rows;
};

__host__ __device__ inline int rup(int a, int b) { return (a + b - 1) / b * b; }

// ------------------------------------------------------------ tables
// synthesis A: rows theta (thp), K over l_p
__global__ void syn_tab_oz_kernel(const double *tP, const double *tD, const double *tM, const long *syn_off,
                                  int lmax, int nth, int nthp, const OzMat *am, int8_t *A, int *eA, int ra)
{
    const int th = blockIdx.x * blockDim.x + threadIdx.x;
    const int a = blockIdx.y, m = blockIdx.z;
    const OzMat M = am[m * 4 + a];
    if (th >= M.rows) return;
    const int L = lmax + 1, re = (L - m + 1) / 2, ro = (L - m) / 2;
    auto tab = [&](const double *t, int par, int row) -> double {
        const int nrow = par ? ro : re;
        if (row < 0 || row >= nrow || th >= nth) return 0.0;
        return t[(syn_off[2 * m + par] + row) * nthp + th];
    };
    auto val = [&](int k) -> double {
        switch (a) {
        case 0: return tab(tP, 0, k);
        case 1: return tab(tP, 1, k);
        case 2: return k < M.split ? tab(tD, 1, k) : tab(tM, 0, k - M.split);
        default: return k < M.split ? tab(tD, 0, k) : tab(tM, 1, k - M.split);
        }
    };
    double mx = 0.0;
    for (int k = 0; k < M.K; k++) mx = fmax(mx, fabs(val(k)));
    const int e = oz_exponent(mx);
    eA[(m * 4 + a) * ra + th] = e;
    for (int k = 0; k < M.K; k++) {
        int8_t d[OZ_S];
        oz_digits(val(k), e, d);
#pragma unroll
        for (int p = 0; p < OZ_S; p++) A[M.off + ((long)p * M.rows + th) * M.K + k] = d[p];
    }
}

// analysis A: rows l_p, K over the northern thetas (twice for [D|M])
__global__ void ana_tab_oz_kernel(const double *tP, const double *tD, const double *tM, const long *ana_off,
                                  const int *ana_ncol, int lmax, int nth, int kth, const OzMat *am, int8_t *A,
                                  int *eA, int ra)
{
    const int lp = blockIdx.x * blockDim.x + threadIdx.x;
    const int a = blockIdx.y, m = blockIdx.z;
    const OzMat M = am[m * 4 + a];
    if (lp >= M.rows) return;
    const int L = lmax + 1, par = a & 1, nrow = par ? (L - m) / 2 : (L - m + 1) / 2;
    const int ncol = ana_ncol[m];
    auto tab = [&](const double *t, int th) -> double {
        if (lp >= nrow || th < 0 || th >= nth) return 0.0;
        return t[ana_off[2 * m + par] + (long)th * ncol + lp];
    };
    auto val = [&](int k) -> double {
        if (a < 2) return tab(tP, k);
        return k < kth ? tab(tD, k) : tab(tM, k - kth);
    };
    double mx = 0.0;
    for (int k = 0; k < M.K; k++) mx = fmax(mx, fabs(val(k)));
    const int e = oz_exponent(mx);
    eA[(m * 4 + a) * ra + lp] = e;
    for (int k = 0; k < M.K; k++) {
        int8_t d[OZ_S];
        oz_digits(val(k), e, d);
#pragma unroll
        for (int p = 0; p < OZ_S; p++) A[M.off + ((long)p * M.rows + lp) * M.K + k] = d[p];
    }
}

// ------------------------------------------------------------ operands
struct OzBArgs {
    const double2 *in[3];
    int8_t *B;
    int *eB;
    const OzMat *bm;
    long ld;
    int lmax, nlat, nth, kth, nplanes, ncols, nmat;
};

// The first D digits of four values x[i] 2^-e, packed per digit plane:
// byte i of w[p] is digit p of value i.  The digits are those of oz_digits
// (d_p = rint(128 r), r <- 128 r - d_p, ties to even): t = 128 r + 1.5 2^52
// rounds 128 r to an integer in its low bits, whose low byte is the int8
// digit, and 128 r - d_p = fma(r, 128, 1.5 2^52 - t) is exact.
template <int D>
__device__ __forceinline__ void oz_pack4(const double (&x)[4], int e, unsigned (&w)[D])
{
    constexpr double M = 6755399441055744.0;   // 1.5 2^52
#pragma unroll
    for (int p = 0; p < D; p++) w[p] = 0u;
#pragma unroll
    for (int i = 0; i < 4; i++) {
        double r = ldexp(x[i], -e);
#pragma unroll
        for (int p = 0; p < D; p++) {
            const double t = fma(r, 128.0, M);
            const unsigned lo = (unsigned)__double2loint(t);
            r = fma(r, 128.0, M - t);
            w[p] |= (lo & 0xffu) << (8 * i);
        }
    }
}

// Stores the D digit planes of four consecutive K values (kidx..kidx+3) of
// column col of matrix b.
template <int D>
__device__ __forceinline__ void oz_store4(const OzBArgs &A, const OzMat &M, int col, int kidx, const double (&x)[4],
                                          int e)
{
    unsigned w[D];
    oz_pack4<D>(x, e, w);
    int8_t *dst = A.B + M.off + (long)col * M.K + kidx;
#pragma unroll
    for (int p = 0; p < D; p++) *(unsigned *)(dst + (long)p * A.ncols * M.K) = w[p];
}

// Operand slicing kernels: lane item (plane k, word w) holds K indices
// 4w .. 4w+3 of every operand matrix of one order m and emits all their
// digits from one set of complex loads.  The G lanes of one plane (G = 1 ..
// 64, the smallest power of two holding its words; more words than 64 are
// taken in turns) are consecutive, so a wave covers 64 / G planes and every
// lane has work; the column maxima of a plane are reduced over its G lanes.
// Digits of a column are written only where a matrix has data; the rest of
// the [p][column][K] layout (K padding) stays zero from its allocation.
constexpr int OZ_WPB = 4;   // waves per workgroup

// Group width (lanes per plane) for nw words per column.
__host__ __device__ inline int oz_group_any(int nw)
{
    int g = 1;
    while (g < nw && g < 64) g <<= 1;
    return g;
}
__device__ __forceinline__ int oz_group(int nw) { return oz_group_any(nw); }
inline int oz_group_host(int nw) { return oz_group_any(nw); }

// Maxima of the 12 columns (6 matrices x {re, im}) of a plane, over its G lanes.
__device__ __forceinline__ void oz_group_max12(double (&v)[12], int G)
{
#pragma unroll
    for (int i = 0; i < 12; i++)
        for (int off = 1; off < G; off <<= 1) v[i] = fmax(v[i], __shfl_xor(v[i], off));
}

// Running maxima of column pair (2b, 2b + 1) over four complex values; swap
// for a rotated part (i x or -i x: the re column holds Im x, the im column Re x).
__device__ __forceinline__ void oz_max4c(double *mx, int b, const double2 (&v)[4], bool swap)
{
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const double x = fabs(swap ? v[i].y : v[i].x), y = fabs(swap ? v[i].x : v[i].y);
        mx[2 * b] = fmax(mx[2 * b], x);
        mx[2 * b + 1] = fmax(mx[2 * b + 1], y);
    }
}

// re and im columns of a part of matrix b at kidx: the values v, or i v
// (rot = +1: re -Im, im Re) or -i v (rot = -1: re Im, im -Re), or -v (rot = 2).
template <int D>
__device__ __forceinline__ void oz_put_part(const OzBArgs &A, const OzMat &M, int b, int k, int kidx,
                                            const double2 (&v)[4], int rot, const int *es)
{
    double re[4], im[4];
#pragma unroll
    for (int i = 0; i < 4; i++) {
        re[i] = rot == 0 ? v[i].x : rot == 1 ? -v[i].y : rot == -1 ? v[i].y : -v[i].x;
        im[i] = rot == 0 ? v[i].y : rot == 1 ? v[i].x : rot == -1 ? -v[i].x : -v[i].y;
    }
    oz_store4<D>(A, M, 2 * k, kidx, re, es[2 * b]);
    oz_store4<D>(A, M, 2 * k + 1, kidx, im, es[2 * b + 1]);
}

// synthesis operands: words w over the row index t of the even and odd l of
// order m (l = m + 2t + parity).  Matrix b is [part 1; part 2] over K, part 2
// from K index split on:
//   b  part 1       part 2
//   0  Q_e          -
//   1  Q_o          -
//   2  S_o          i T_e    (split ro16)
//   3  S_e          i T_o    (split re16)
//   4  T_o          -i S_e   (split ro16)
//   5  T_e          -i S_o   (split re16)
// One workgroup per (block of OZ_WPB * 64 / G planes, m).
template <bool VEC, int D>
__global__ __launch_bounds__(64 * OZ_WPB) void syn_b_oz_kernel(OzBArgs A)
{
    const int m = blockIdx.y;
    const int L = A.lmax + 1, re = (L - m + 1) / 2, ro = (L - m) / 2;
    const int re16 = rup(re, 16), ro16 = rup(ro, 16);
    const int nw = (re + 3) / 4, G = oz_group(nw);
    const int lane = threadIdx.x & 63, g = lane % G;
    const int k = (blockIdx.x * OZ_WPB + (threadIdx.x >> 6)) * (64 / G) + lane / G;
    if ((blockIdx.x * OZ_WPB + (threadIdx.x >> 6)) * (64 / G) >= A.nplanes) return;   // whole wave idle
    const bool kin = k < A.nplanes;
    const int kc = kin ? k : 0;
    const long lm0 = (long)m * (2 * A.lmax + 3 - m) / 2 + kc * A.ld;   // (m, m) of plane kc
    // the operand matrices' descriptors of this m, loaded once here, in
    // uniform control flow (scalar loads), instead of once per part written
    OzMat Mb[6];
#pragma unroll
    for (int b = 0; b < (VEC ? 6 : 2); b++) Mb[b] = A.bm[m * 6 + b];
    // four rows t0..t0+3 of field f, parity par (zero past their count or off-plane)
    auto load4 = [&](int f, int par, int t0, double2 (&v)[4]) {
        const int nrow = par ? ro : re;
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int t = t0 + i;
            // (clamped to the last row of the parity, or to (m, m) when it has none)
            double2 x = A.in[f][nrow > 0 ? lm0 + 2 * min(t, nrow - 1) + par : lm0];
            if (m == 0) x.y = 0.0;
            v[i] = (kin && t < nrow) ? x : make_double2(0.0, 0.0);
        }
    };
    double mx[12];
#pragma unroll
    for (int i = 0; i < 12; i++) mx[i] = 0.0;
    for (int w = g; w < nw; w += G) {
        double2 qe[4], qo[4];
        load4(0, 0, 4 * w, qe);
        load4(0, 1, 4 * w, qo);
        oz_max4c(mx, 0, qe, false);
        oz_max4c(mx, 1, qo, false);
        if (VEC) {
            double2 se[4], so[4], te[4], to[4];
            load4(1, 0, 4 * w, se);
            load4(1, 1, 4 * w, so);
            load4(2, 0, 4 * w, te);
            load4(2, 1, 4 * w, to);
            oz_max4c(mx, 2, so, false); oz_max4c(mx, 2, te, true);
            oz_max4c(mx, 3, se, false); oz_max4c(mx, 3, to, true);
            oz_max4c(mx, 4, to, false); oz_max4c(mx, 4, se, true);
            oz_max4c(mx, 5, te, false); oz_max4c(mx, 5, so, true);
        }
    }
    oz_group_max12(mx, G);
    int es[12];
#pragma unroll
    for (int i = 0; i < 12; i++) es[i] = oz_exponent(mx[i]);
    if (!kin) return;
    for (int w = g; w < nw; w += G) {
        const int t0 = 4 * w;
        double2 qe[4], qo[4];
        load4(0, 0, t0, qe);
        load4(0, 1, t0, qo);
        oz_put_part<D>(A, Mb[0], 0, k, t0, qe, 0, es);
        if (t0 < ro) oz_put_part<D>(A, Mb[1], 1, k, t0, qo, 0, es);
        if (VEC) {
            double2 se[4], so[4], te[4], to[4];
            load4(1, 0, t0, se);
            load4(1, 1, t0, so);
            load4(2, 0, t0, te);
            load4(2, 1, t0, to);
            if (t0 < ro) {
                oz_put_part<D>(A, Mb[2], 2, k, t0, so, 0, es);
                oz_put_part<D>(A, Mb[4], 4, k, t0, to, 0, es);
                oz_put_part<D>(A, Mb[3], 3, k, re16 + t0, to, 1, es);
                oz_put_part<D>(A, Mb[5], 5, k, re16 + t0, so, -1, es);
            }
            oz_put_part<D>(A, Mb[2], 2, k, ro16 + t0, te, 1, es);
            oz_put_part<D>(A, Mb[4], 4, k, ro16 + t0, se, -1, es);
            oz_put_part<D>(A, Mb[3], 3, k, t0, se, 0, es);
            oz_put_part<D>(A, Mb[5], 5, k, t0, te, 0, es);
        }
    }
    if (g == 0)
        for (int i = 0; i < 2 * A.nmat; i++) A.eB[(m * 6 + i / 2) * A.ncols + 2 * k + (i & 1)] = es[i];
}

// analysis operands: words w over the northern thetas t; x+- = x_north +-
// x_south of the Fourier coefficients of field x (r, theta, phi):
//   b  part 1   part 2 (from kth)
//   0  r+       -
//   1  r-       -
//   2  t-       -i p+
//   3  t+       -i p-
//   4  -p-      -i t+
//   5  -p+      -i t-
template <bool VEC, int D>
__global__ __launch_bounds__(64 * OZ_WPB) void ana_b_oz_kernel(OzBArgs A)
{
    const int m = blockIdx.y;
    const int L = A.lmax + 1;
    const int nw = (A.nth + 3) / 4, G = oz_group(nw);
    const int lane = threadIdx.x & 63, g = lane % G;
    const int k = (blockIdx.x * OZ_WPB + (threadIdx.x >> 6)) * (64 / G) + lane / G;
    if ((blockIdx.x * OZ_WPB + (threadIdx.x >> 6)) * (64 / G) >= A.nplanes) return;   // whole wave idle
    const bool kin = k < A.nplanes;
    const long base = ((long)(kin ? k : 0) * L + m) * A.nlat;
    OzMat Mb[6];   // the operand matrices' descriptors of this m, loaded once (as in syn_b_oz_kernel)
#pragma unroll
    for (int b = 0; b < (VEC ? 6 : 2); b++) Mb[b] = A.bm[m * 6 + b];
    // x+ and x- of field f at thetas t0..t0+3 (zero past nth or off-plane)
    auto load4 = [&](int f, int t0, double2 (&sp)[4], double2 (&sm)[4]) {
#pragma unroll
        for (int i = 0; i < 4; i++) {
            const int t = t0 + i, tc = min(t, A.nth - 1);
            double2 n = A.in[f][base + tc], s = A.in[f][base + A.nlat - 1 - tc];
            if (m == 0) {
                n.y = 0.0;
                s.y = 0.0;
            }
            const bool ok = kin && t < A.nth;
            sp[i] = ok ? make_double2(n.x + s.x, n.y + s.y) : make_double2(0.0, 0.0);
            sm[i] = ok ? make_double2(n.x - s.x, n.y - s.y) : make_double2(0.0, 0.0);
        }
    };
    double mx[12];
#pragma unroll
    for (int i = 0; i < 12; i++) mx[i] = 0.0;
    for (int w = g; w < nw; w += G) {
        double2 rp[4], rm[4];
        load4(0, 4 * w, rp, rm);
        oz_max4c(mx, 0, rp, false);
        oz_max4c(mx, 1, rm, false);
        if (VEC) {
            double2 tp[4], tm[4], pp[4], pm[4];
            load4(1, 4 * w, tp, tm);
            load4(2, 4 * w, pp, pm);
            oz_max4c(mx, 2, tm, false); oz_max4c(mx, 2, pp, true);
            oz_max4c(mx, 3, tp, false); oz_max4c(mx, 3, pm, true);
            oz_max4c(mx, 4, pm, false); oz_max4c(mx, 4, tp, true);
            oz_max4c(mx, 5, pp, false); oz_max4c(mx, 5, tm, true);
        }
    }
    oz_group_max12(mx, G);
    int es[12];
#pragma unroll
    for (int i = 0; i < 12; i++) es[i] = oz_exponent(mx[i]);
    if (!kin) return;
    for (int w = g; w < nw; w += G) {
        const int t0 = 4 * w;
        double2 rp[4], rm[4];
        load4(0, t0, rp, rm);
        oz_put_part<D>(A, Mb[0], 0, k, t0, rp, 0, es);
        oz_put_part<D>(A, Mb[1], 1, k, t0, rm, 0, es);
        if (VEC) {
            double2 tp[4], tm[4], pp[4], pm[4];
            load4(1, t0, tp, tm);
            load4(2, t0, pp, pm);
            oz_put_part<D>(A, Mb[2], 2, k, t0, tm, 0, es);
            oz_put_part<D>(A, Mb[2], 2, k, A.kth + t0, pp, -1, es);
            oz_put_part<D>(A, Mb[3], 3, k, t0, tp, 0, es);
            oz_put_part<D>(A, Mb[3], 3, k, A.kth + t0, pm, -1, es);
            oz_put_part<D>(A, Mb[4], 4, k, t0, pm, 2, es);
            oz_put_part<D>(A, Mb[4], 4, k, A.kth + t0, tp, -1, es);
            oz_put_part<D>(A, Mb[5], 5, k, t0, pp, 2, es);
            oz_put_part<D>(A, Mb[5], 5, k, A.kth + t0, tm, -1, es);
        }
    }
    if (g == 0)
        for (int i = 0; i < 2 * A.nmat; i++) A.eB[(m * 6 + i / 2) * A.ncols + 2 * k + (i & 1)] = es[i];
}

// ------------------------------------------------------------ GEMM
struct OzGArgs {
    const int8_t *A, *B;
    const OzMat *am, *bm;
    const int *eA, *eB;
    const int2 *items;          // (m, row tile)
    double2 *out[3];
    long ld;
    int ra, ncols, n_jt;
    int lmax, nlat, nth, nplanes;
};

// LDS-staged digit GEMMs of one workgroup item.  The item is a 64-row slab of
// the A matrices of a pair of GEMMs (s = 0, 1: Ma[s] x Mb[s]) times the
// 64-column tiles jt0 <= jt < jt0 + njt of the B matrices.  Per K step of 64
// the D digit planes of the A rows and B columns go to LDS (rows at a pitch
// of 80 bytes: the 16-byte fragment reads of 16 lanes then cover all 64
// banks), and each wave multiplies its 16 rows by the 64 columns:
// acc[p + q][j] += A_p B_q[j] for p + q <= D - 1.  Thread t stages row t / 4,
// 16-byte chunk t % 4 of every plane of both operands.  The GEMMs (jt, s) run
// as one sequence of K steps, each step's staging loads issued during the
// previous step's MFMAs across GEMM boundaries, so that only the item's first
// step waits for them in full; epi(jt, s, acc) runs after each GEMM's last
// step.  A pair with live[s] false is skipped (at least one is live); every
// live pair has K > 0, a multiple of 64.
constexpr int OZ_PITCH = 80;
constexpr int OZ_TILE = OZ_S * 64 * OZ_PITCH;     // bytes of one operand tile

template <int D>
__device__ __forceinline__ void oz_zero(v4i (&acc)[D][4])
{
#pragma unroll
    for (int L = 0; L < D; L++)
#pragma unroll
        for (int j = 0; j < 4; j++) acc[L][j] = (v4i){0, 0, 0, 0};
}

template <int D, typename Epi>
__device__ __forceinline__ void oz_item(const int8_t *Aall, const int8_t *Ball, int ncols, const OzMat (&Ma)[2],
                                        const OzMat (&Mb)[2], const bool (&live)[2], int arow0, int jt0,
                                        int njt, int8_t *As, int8_t *Bs, Epi &&epi)
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, r = lane & 15, g = lane >> 4;
    const int srow = tid >> 2, sq = tid & 3;
    v4i ra[D], rb[D];
    auto gload = [&](int jt, int s, int kb) {
        const int K = Ma[s].K;
        const int8_t *Ab = Aall + Ma[s].off, *Bb = Ball + Mb[s].off;
#pragma unroll
        for (int p = 0; p < D; p++) {
            ra[p] = *(const v4i *)(Ab + ((long)p * Ma[s].rows + arow0 + srow) * K + kb + 16 * sq);
            rb[p] = *(const v4i *)(Bb + ((long)p * ncols + 64 * jt + srow) * K + kb + 16 * sq);
        }
    };
    // the step after (jt, s, kb)
    auto advance = [&](int &jt, int &s, int &kb) {
        kb += 64;
        if (kb < Ma[s].K) return;
        kb = 0;
        if (s == 0 && live[1]) {
            s = 1;
            return;
        }
        s = live[0] ? 0 : 1;
        jt++;
    };
    v4i acc[D][4];
    oz_zero<D>(acc);
    const int jend = jt0 + njt;
    int jt = jt0, s = live[0] ? 0 : 1, kb = 0;
    gload(jt, s, kb);
    while (jt < jend) {
#pragma unroll
        for (int p = 0; p < D; p++) {
            *(v4i *)(As + (p * 64 + srow) * OZ_PITCH + 16 * sq) = ra[p];
            *(v4i *)(Bs + (p * 64 + srow) * OZ_PITCH + 16 * sq) = rb[p];
        }
        __syncthreads();
        int jn = jt, sn = s, kn = kb;
        advance(jn, sn, kn);
        if (jn < jend) gload(jn, sn, kn);
        v4i a[D];
#pragma unroll
        for (int p = 0; p < D; p++) a[p] = *(const v4i *)(As + (p * 64 + wave * 16 + r) * OZ_PITCH + 16 * g);
#pragma unroll
        for (int q = 0; q < D; q++) {
            v4i bq[4];
#pragma unroll
            for (int j = 0; j < 4; j++) bq[j] = *(const v4i *)(Bs + (q * 64 + 16 * j + r) * OZ_PITCH + 16 * g);
#pragma unroll
            for (int p = 0; p < D - q; p++)
#pragma unroll
                for (int j = 0; j < 4; j++)
                    acc[p + q][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[p], bq[j], acc[p + q][j], 0, 0, 0);
        }
        __syncthreads();
        if (kn == 0) {   // (jt, s) has had its last step
            epi(jt, s, acc);
            oz_zero<D>(acc);
        }
        jt = jn;
        s = sn;
        kb = kn;
    }
}

// One digit GEMM of a 64-row x 64-column tile with the staging of oz_item:
// the next K step's chunks are loaded into registers during the current
// step's MFMAs (the analysis tiles, whose K of 3 or 6 steps is pipelined
// within the GEMM; a sequence across both parities measured slower for them).
template <int D>
__device__ __forceinline__ void oz_gemm_lds(const int8_t *Ab, int arows, int arow0, const int8_t *Bb,
                                            int ncols, int col0, int K, int8_t *As, int8_t *Bs,
                                            v4i (&acc)[D][4])
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, r = lane & 15, g = lane >> 4;
    const int srow = tid >> 2, sq = tid & 3;
    v4i ra[D], rb[D];
    auto gload = [&](int kb) {
#pragma unroll
        for (int p = 0; p < D; p++) {
            ra[p] = *(const v4i *)(Ab + ((long)p * arows + arow0 + srow) * K + kb + 16 * sq);
            rb[p] = *(const v4i *)(Bb + ((long)p * ncols + col0 + srow) * K + kb + 16 * sq);
        }
    };
    gload(0);
    for (int kb = 0; kb < K; kb += 64) {
#pragma unroll
        for (int p = 0; p < D; p++) {
            *(v4i *)(As + (p * 64 + srow) * OZ_PITCH + 16 * sq) = ra[p];
            *(v4i *)(Bs + (p * 64 + srow) * OZ_PITCH + 16 * sq) = rb[p];
        }
        __syncthreads();
        if (kb + 64 < K) gload(kb + 64);
        v4i a[D];
#pragma unroll
        for (int p = 0; p < D; p++) a[p] = *(const v4i *)(As + (p * 64 + wave * 16 + r) * OZ_PITCH + 16 * g);
#pragma unroll
        for (int q = 0; q < D; q++) {
            v4i bq[4];
#pragma unroll
            for (int j = 0; j < 4; j++) bq[j] = *(const v4i *)(Bs + (q * 64 + 16 * j + r) * OZ_PITCH + 16 * g);
#pragma unroll
            for (int p = 0; p < D - q; p++)
#pragma unroll
                for (int j = 0; j < 4; j++)
                    acc[p + q][j] = __builtin_amdgcn_mfma_i32_16x16x64_i8(a[p], bq[j], acc[p + q][j], 0, 0, 0);
        }
        __syncthreads();
    }
}

// FP64 value of output element (v, j): the levels from the smallest weight up.
template <int D>
__device__ __forceinline__ double oz_value(const v4i (&acc)[D][4], int j, int v, int e)
{
    double s = 0.0;
#pragma unroll
    for (int L = D - 1; L >= 0; L--) s += ldexp((double)acc[L][j][v], -7 * (L + 2));
    return ldexp(s, e);
}

// synthesis: workgroup = 64 thetas (16 per wave) of one m, one field, every
// column tile, both parities (the symmetric part of a column tile is held as
// FP64 while its antisymmetric GEMM runs); D digits per operand
template <int D>
__global__ __launch_bounds__(256, 2) void syn_oz_kernel(OzGArgs A)
{
    extern __shared__ __align__(16) int8_t oz_sm[];
    int8_t *As = oz_sm, *Bs = oz_sm + OZ_TILE;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, r = lane & 15, g = lane >> 4;
    const int2 it = A.items[blockIdx.x];
    const int m = it.x, arow0 = it.y * 64, row0 = arow0 + wave * 16;
    const int f = blockIdx.y;
    const int L = A.lmax + 1;
    const int am[2] = {f == 0 ? 0 : 2, f == 0 ? 1 : 3};
    const int bmx[2] = {f == 0 ? 0 : (f == 1 ? 2 : 4), f == 0 ? 1 : (f == 1 ? 3 : 5)};
    const OzMat Ma[2] = {A.am[m * 4 + am[0]], A.am[m * 4 + am[1]]};
    const OzMat Mb[2] = {A.bm[m * 6 + bmx[0]], A.bm[m * 6 + bmx[1]]};
    // K = 0 only for the odd-l matrix of m = lmax; the even-l part is never empty
    const bool live[2] = {true, Ma[1].K > 0};
    const double sg = f == 2 ? -1.0 : 1.0;
    double *out = (double *)A.out[f];
    double sym[4][4];
    oz_item<D>(A.A, A.B, A.ncols, Ma, Mb, live, arow0, 0, A.n_jt, As, Bs, [&](int jt, int s, const v4i (&acc)[D][4]) {
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int col = 64 * jt + 16 * j + r;
            const int eb = A.eB[(m * 6 + bmx[s]) * A.ncols + col];
            double res[4];
#pragma unroll
            for (int v = 0; v < 4; v++)
                res[v] = oz_value<D>(acc, j, v, A.eA[(m * 4 + am[s]) * A.ra + row0 + 4 * g + v] + eb);
            if (s == 0 && live[1]) {
#pragma unroll
                for (int v = 0; v < 4; v++) sym[j][v] = res[v];
                continue;
            }
            const int k = col >> 1, c = col & 1;
            if (k >= A.nplanes) continue;
            const long base = ((long)k * L + m) * A.nlat;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int th = row0 + 4 * g + v;
                if (th >= A.nth) continue;
                // s = 1: sym held from s = 0; s = 0 without an odd part: anti = 0
                const double sy = s == 1 ? sym[j][v] : res[v], an = s == 1 ? res[v] : 0.0;
                out[2 * (base + th) + c] = sg * (sy + an);
                out[2 * (base + A.nlat - 1 - th) + c] = sg * (sy - an);
            }
        }
    });
}

// analysis: workgroup = 64 rows l_p (16 per wave) x 64 columns, one field,
// both parities; D digits per operand
template <int D>
__global__ __launch_bounds__(256, 2) void ana_oz_kernel(OzGArgs A)
{
    extern __shared__ __align__(16) int8_t oz_sm[];
    int8_t *As = oz_sm, *Bs = oz_sm + OZ_TILE;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, r = lane & 15, g = lane >> 4;
    const int jt = blockIdx.x % A.n_jt;
    const int2 it = A.items[blockIdx.x / A.n_jt];
    const int m = it.x, arow0 = it.y * 64, row0 = arow0 + wave * 16, col0 = jt * 64;
    const int f = blockIdx.y;
    const long lm0 = (long)m * (2 * A.lmax + 3 - m) / 2;
    double *out = (double *)A.out[f];
#pragma unroll
    for (int par = 0; par < 2; par++) {
        const int a = f == 0 ? par : 2 + par;
        const int b = f == 0 ? par : (f == 1 ? 2 + par : 4 + par);
        const OzMat Ma = A.am[m * 4 + a], Mb = A.bm[m * 6 + b];
        if (arow0 >= Ma.rows) continue;       // uniform over the workgroup
        v4i acc[D][4];
        oz_zero<D>(acc);
        oz_gemm_lds<D>(A.A + Ma.off, Ma.rows, arow0, A.B + Mb.off, A.ncols, col0, Ma.K, As, Bs, acc);
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int col = col0 + 16 * j + r, k = col >> 1, c = col & 1;
            if (k >= A.nplanes) continue;
            const int eb = A.eB[(m * 6 + b) * A.ncols + col];
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int lp = row0 + 4 * g + v;
                const int l = m + par + 2 * lp;
                if (l > A.lmax) continue;
                double val = oz_value<D>(acc, j, v, A.eA[(m * 4 + a) * A.ra + lp] + eb);
                if (m == 0 && c == 1) val = 0.0;
                out[2 * (lm0 + (l - m) + k * A.ld) + c] = val;
            }
        }
    }
}

}  // namespace

// ------------------------------------------------------------ host
struct OzDir {
    OzMat *d_am = nullptr, *d_bm = nullptr;
    int8_t *A = nullptr, *B = nullptr;
    int *eA = nullptr, *eB = nullptr;
    int2 *d_items = nullptr;
    int n_items = 0, ra = 0;
};

struct OzState {
    OzDir syn, ana;
    int ncols = 0;
};

static OzState *oz(Plan *p) { return (OzState *)p->oz; }

static int oz_alloc_dir(Plan *p, OzDir &D, const std::vector<OzMat> &am, const std::vector<OzMat> &bm,
                        long abytes, long bbytes, const std::vector<int2> &items, int ra, int ncols)
{
    const int L = p->L;
    RD_HIP(hipMalloc(&D.d_am, sizeof(OzMat) * am.size()));
    RD_HIP(hipMalloc(&D.d_bm, sizeof(OzMat) * bm.size()));
    RD_HIP(hipMemcpyAsync(D.d_am, am.data(), sizeof(OzMat) * am.size(), hipMemcpyHostToDevice, p->stream));
    RD_HIP(hipMemcpyAsync(D.d_bm, bm.data(), sizeof(OzMat) * bm.size(), hipMemcpyHostToDevice, p->stream));
    RD_HIP(hipMalloc(&D.A, abytes));
    RD_HIP(hipMalloc(&D.B, bbytes));
    RD_HIP(hipMemsetAsync(D.B, 0, bbytes, p->stream));
    RD_HIP(hipMalloc(&D.eA, sizeof(int) * 4L * L * ra));
    RD_HIP(hipMemsetAsync(D.eA, 0, sizeof(int) * 4L * L * ra, p->stream));
    RD_HIP(hipMalloc(&D.eB, sizeof(int) * 6L * L * ncols));
    RD_HIP(hipMemsetAsync(D.eB, 0, sizeof(int) * 6L * L * ncols, p->stream));
    RD_HIP(hipMalloc(&D.d_items, sizeof(int2) * items.size()));
    RD_HIP(hipMemcpyAsync(D.d_items, items.data(), sizeof(int2) * items.size(), hipMemcpyHostToDevice,
                          p->stream));
    D.n_items = (int)items.size();
    D.ra = ra;
    return RD_OK;
}

// Builds the sliced tables and the operand scratch on first use.
int ensure_oz_tables(Plan *p)
{
    if (p->oz) return RD_OK;
    OzState *S = new OzState();
    p->oz = S;
    const int L = p->L, nth = p->nth;
    S->ncols = rup(2 * p->nz, OZ_CT);
    const int thp = rup(nth, 64), kth = rup(nth, 16);
    // synthesis: A rows theta, K over l_p
    {
        std::vector<OzMat> am(4 * L), bm(6 * L);
        std::vector<int2> items;
        long ao = 0, bo = 0;
        for (int m = 0; m < L; m++) {
            const int re = (L - m + 1) / 2, ro = (L - m) / 2, re16 = rup(re, 16), ro16 = rup(ro, 16);
            const int K[4] = {rup(re, 64), rup(ro, 64), rup(ro16 + re16, 64), rup(re16 + ro16, 64)};
            const int sp[4] = {K[0], K[1], ro16, re16};
            for (int a = 0; a < 4; a++) {
                am[4 * m + a] = OzMat{ao, K[a], sp[a], thp};
                ao += (long)OZ_S * thp * K[a];
            }
            const int bk[6] = {0, 1, 2, 3, 2, 3};
            for (int b = 0; b < 6; b++) {
                bm[6 * m + b] = OzMat{bo, K[bk[b]], sp[bk[b]], S->ncols};
                bo += (long)OZ_S * S->ncols * K[bk[b]];
            }
            for (int tt = 0; tt < thp / 64; tt++) items.push_back(make_int2(m, tt));
        }
        int rc = oz_alloc_dir(p, S->syn, am, bm, ao, bo, items, thp, S->ncols);
        if (rc) return rc;
        hipLaunchKernelGGL(syn_tab_oz_kernel, dim3(thp / 64, 4, L), dim3(64), 0, p->stream, p->syn_P, p->syn_D,
                           p->syn_M, p->d_syn_off, p->lmax, nth, p->nthp, S->syn.d_am, S->syn.A, S->syn.eA, thp);
        RD_LAUNCH_CHECK();
    }
    // analysis: A rows l_p, K over the northern thetas
    {
        const int ra = rup((L + 1) / 2, 64);
        std::vector<OzMat> am(4 * L), bm(6 * L);
        std::vector<int2> items;
        long ao = 0, bo = 0;
        const int K1 = rup(kth, 64), K2 = rup(2 * kth, 64);
        for (int m = 0; m < L; m++) {
            const int re = (L - m + 1) / 2, ro = (L - m) / 2;
            const int rows[4] = {rup(re, 64), rup(ro, 64), rup(re, 64), rup(ro, 64)};
            const int K[4] = {K1, K1, K2, K2};
            for (int a = 0; a < 4; a++) {
                am[4 * m + a] = OzMat{ao, K[a], kth, rows[a]};
                ao += (long)OZ_S * rows[a] * K[a];
            }
            const int bk[6] = {K1, K1, K2, K2, K2, K2};
            for (int b = 0; b < 6; b++) {
                bm[6 * m + b] = OzMat{bo, bk[b], kth, S->ncols};
                bo += (long)OZ_S * S->ncols * bk[b];
            }
            for (int rt = 0; rt < rows[0] / 64; rt++) items.push_back(make_int2(m, rt));
        }
        int rc = oz_alloc_dir(p, S->ana, am, bm, ao, bo, items, ra, S->ncols);
        if (rc) return rc;
        hipLaunchKernelGGL(ana_tab_oz_kernel, dim3(ra / 64, 4, L), dim3(64), 0, p->stream, p->ana_P, p->ana_D,
                           p->ana_M, p->d_ana_off, p->d_ana_ncol, p->lmax, nth, kth, S->ana.d_am, S->ana.A,
                           S->ana.eA, ra);
        RD_LAUNCH_CHECK();
    }
    return RD_OK;
}

void free_oz_tables(Plan *p)
{
    OzState *S = oz(p);
    if (!S) return;
    for (OzDir *D : {&S->syn, &S->ana}) {
        void *ptrs[] = {D->d_am, D->d_bm, D->A, D->B, D->eA, D->eB, D->d_items};
        for (void *q : ptrs)
            if (q) (void)hipFree(q);
    }
    delete S;
    p->oz = nullptr;
}

int launch_synthesis_oz(Plan *p, bool vec, int nplanes, const double2 *q, const double2 *s, const double2 *t,
                        long ld, double2 *f0, double2 *f1, double2 *f2)
{
    int rc = ensure_oz_tables(p);
    if (rc) return rc;
    OzState *S = oz(p);
    if (2 * nplanes > S->ncols) return RD_ERR_ARG;
    OzBArgs B;
    B.in[0] = q;
    B.in[1] = vec ? s : q;
    B.in[2] = vec ? t : q;
    B.B = S->syn.B;
    B.eB = S->syn.eB;
    B.bm = S->syn.d_bm;
    B.ld = ld;
    B.lmax = p->lmax;
    B.nlat = p->nlat;
    B.nth = p->nth;
    B.kth = 0;
    B.nplanes = nplanes;
    B.ncols = S->ncols;
    B.nmat = vec ? 6 : 2;
    const int D = p->oz_digits;
    // planes per workgroup for the longest columns (m = 0)
    const int ppw = OZ_WPB * 64 / oz_group_host(((p->L + 1) / 2 + 3) / 4);
    const dim3 bgrid((nplanes + ppw - 1) / ppw, p->L);
#define RD_SYN_B(DD)                                                                                    \
    if (vec)                                                                                            \
        hipLaunchKernelGGL((syn_b_oz_kernel<true, DD>), bgrid, dim3(64 * OZ_WPB), 0, p->stream, B);     \
    else                                                                                                \
        hipLaunchKernelGGL((syn_b_oz_kernel<false, DD>), bgrid, dim3(64 * OZ_WPB), 0, p->stream, B);
    if (D == 5) { RD_SYN_B(5) } else if (D == 6) { RD_SYN_B(6) } else { RD_SYN_B(7) }
#undef RD_SYN_B
    RD_LAUNCH_CHECK();
    OzGArgs G = {};
    G.A = S->syn.A;
    G.B = S->syn.B;
    G.am = S->syn.d_am;
    G.bm = S->syn.d_bm;
    G.eA = S->syn.eA;
    G.eB = S->syn.eB;
    G.items = S->syn.d_items;
    G.out[0] = f0;
    G.out[1] = vec ? f1 : f0;
    G.out[2] = vec ? f2 : f0;
    G.ld = ld;
    G.ra = S->syn.ra;
    G.ncols = S->ncols;
    G.n_jt = (2 * nplanes + OZ_CT - 1) / OZ_CT;
    G.lmax = p->lmax;
    G.nlat = p->nlat;
    G.nth = p->nth;
    G.nplanes = nplanes;
    const dim3 ggrid(S->syn.n_items, vec ? 3 : 1);
    if (D == 5)
        hipLaunchKernelGGL(syn_oz_kernel<5>, ggrid, dim3(256), 2 * OZ_TILE, p->stream, G);
    else if (D == 6)
        hipLaunchKernelGGL(syn_oz_kernel<6>, ggrid, dim3(256), 2 * OZ_TILE, p->stream, G);
    else
        hipLaunchKernelGGL(syn_oz_kernel<7>, ggrid, dim3(256), 2 * OZ_TILE, p->stream, G);
    RD_LAUNCH_CHECK();
    return RD_OK;
}

int launch_analysis_oz(Plan *p, bool vec, int nplanes, const double2 *g0, const double2 *g1, const double2 *g2,
                       double2 *q, double2 *s, double2 *t, long ld)
{
    int rc = ensure_oz_tables(p);
    if (rc) return rc;
    OzState *S = oz(p);
    if (2 * nplanes > S->ncols) return RD_ERR_ARG;
    OzBArgs B;
    B.in[0] = g0;
    B.in[1] = vec ? g1 : g0;
    B.in[2] = vec ? g2 : g0;
    B.B = S->ana.B;
    B.eB = S->ana.eB;
    B.bm = S->ana.d_bm;
    B.ld = ld;
    B.lmax = p->lmax;
    B.nlat = p->nlat;
    B.nth = p->nth;
    B.kth = rup(p->nth, 16);
    B.nplanes = nplanes;
    B.ncols = S->ncols;
    B.nmat = vec ? 6 : 2;
    const int D = p->oz_digits;
    const int ppw = OZ_WPB * 64 / oz_group_host((p->nth + 3) / 4);
    const dim3 bgrid((nplanes + ppw - 1) / ppw, p->L);
#define RD_ANA_B(DD)                                                                                    \
    if (vec)                                                                                            \
        hipLaunchKernelGGL((ana_b_oz_kernel<true, DD>), bgrid, dim3(64 * OZ_WPB), 0, p->stream, B);     \
    else                                                                                                \
        hipLaunchKernelGGL((ana_b_oz_kernel<false, DD>), bgrid, dim3(64 * OZ_WPB), 0, p->stream, B);
    if (D == 5) { RD_ANA_B(5) } else if (D == 6) { RD_ANA_B(6) } else { RD_ANA_B(7) }
#undef RD_ANA_B
    RD_LAUNCH_CHECK();
    OzGArgs G = {};
    G.A = S->ana.A;
    G.B = S->ana.B;
    G.am = S->ana.d_am;
    G.bm = S->ana.d_bm;
    G.eA = S->ana.eA;
    G.eB = S->ana.eB;
    G.items = S->ana.d_items;
    G.out[0] = q;
    G.out[1] = vec ? s : q;
    G.out[2] = vec ? t : q;
    G.ld = ld;
    G.ra = S->ana.ra;
    G.ncols = S->ncols;
    G.n_jt = (2 * nplanes + OZ_CT - 1) / OZ_CT;
    G.lmax = p->lmax;
    G.nlat = p->nlat;
    G.nth = p->nth;
    G.nplanes = nplanes;
    const dim3 ggrid(S->ana.n_items * G.n_jt, vec ? 3 : 1);
    if (D == 5)
        hipLaunchKernelGGL(ana_oz_kernel<5>, ggrid, dim3(256), 2 * OZ_TILE, p->stream, G);
    else if (D == 6)
        hipLaunchKernelGGL(ana_oz_kernel<6>, ggrid, dim3(256), 2 * OZ_TILE, p->stream, G);
    else
        hipLaunchKernelGGL(ana_oz_kernel<7>, ggrid, dim3(256), 2 * OZ_TILE, p->stream, G);
    RD_LAUNCH_CHECK();
    return RD_OK;
}

}  // namespace rd
