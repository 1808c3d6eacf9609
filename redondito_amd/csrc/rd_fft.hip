// Longitude real FFTs, the fused u x w grid step, and the FC-Gram radial
// transforms, as LDS Stockham kernels (FP64 complex).
//
// phi: the Fourier coefficients live in F[k][m][theta] (complex, theta
// fastest).  Two real rows are packed into one complex FFT of length nphi
// (x + i y), the Hermitian half m = 0..lmax is the only part stored, and the
// forward transform is scaled by 2 pi / nphi so that the Legendre analysis
// with the Gauss weights gives the orthonormal coefficients (the SHTns
// convention the reference configures, src/sphp/sphp_sca.F90:117-122).
//
// z: columns of a (nlm, nz) complex array, lm fastest.  A workgroup loads a
// tile of consecutive lm for all nz rows (coalesced row segments), applies the
// FC-Gram continuation (src/sphp/sphp_sca.F90:369-381), the c2c FFT along k
// and the 1/nz factor (:391), and stores in place.
#include "rd_internal.h"

#include <cmath>
#include <vector>

namespace rd {

FftSeq make_fft_seq(int n)
{
    FftSeq s;
    s.n = n;
    s.nst = 0;
    int m = n;
    while (m % 4 == 0 && s.nst < kMaxStages) {
        s.rad[s.nst++] = 4;
        m /= 4;
    }
    for (int f = 2; m > 1 && s.nst < kMaxStages; f++)
        while (m % f == 0) {
            s.rad[s.nst++] = f;
            m /= f;
        }
    return s;
}

__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cscale(double2 a, double s) { return make_double2(a.x * s, a.y * s); }
// sign * i * a
__device__ __forceinline__ double2 cmuli(double2 a, int sign)
{
    return sign > 0 ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x);
}
// exp(sign 2 pi i t / n) from a table of the first nh entries of
// exp(-2 pi i t / n); nh = n/2 (even n) uses exp(-i pi) = -1 for the rest.
__device__ __forceinline__ double2 twid(const double2 *tw, int t, int nh, int sign)
{
    double2 w;
    if (t < nh) {
        w = tw[t];
    } else {
        w = tw[t - nh];
        w.x = -w.x;
        w.y = -w.y;
    }
    return sign > 0 ? make_double2(w.x, -w.y) : w;
}

// Copies the twiddle table into LDS (n/2 entries for even n, else n) and
// returns the number of entries held.
__device__ int load_twiddles(double2 *dst, const double2 *tw, int n)
{
    const int nh = (n % 2 == 0) ? n / 2 : n;
    for (int t = threadIdx.x; t < nh; t += blockDim.x) dst[t] = tw[t];
    return nh;
}

// Workgroup index remapped so that consecutive logical tiles run on the same
// XCD (the hardware deals workgroups round-robin over the 8 XCDs, each with
// its own L2): neighbouring tiles then share cache lines in one L2.
__device__ __forceinline__ int xcd_tile(int bid, int nblocks)
{
    const int q = nblocks / 8, r = nblocks % 8;
    const int x = bid % 8, i = bid / 8;
    return x * q + (x < r ? x : r) + i;
}

// One radix-R Stockham stage over nb sequences of length n (a -> b).
template <int R>
__device__ void fft_stage(const double2 *a, double2 *b, int nb, int n, int Ns, int sign,
                          const double2 *tw, int nh)
{
    const int nbf = n / R, tot = nb * nbf, step = n / (Ns * R);
    for (int bf = threadIdx.x; bf < tot; bf += blockDim.x) {
        const int bb = bf / nbf, j = bf - bb * nbf;
        const int k = j % Ns;
        const double2 *in = a + bb * n;
        double2 *out = b + bb * n;
        double2 v[R];
#pragma unroll
        for (int r = 0; r < R; r++) v[r] = in[j + r * nbf];
        if (Ns > 1) {
#pragma unroll
            for (int r = 1; r < R; r++) v[r] = cmul(v[r], twid(tw, r * k * step, nh, sign));
        }
        if (R == 2) {
            double2 t0 = cadd(v[0], v[1]), t1 = csub(v[0], v[1]);
            v[0] = t0;
            v[1] = t1;
        } else if (R == 3) {
            const double c3 = 0.86602540378443864676;
            double2 t1 = cadd(v[1], v[2]);
            double2 t2 = make_double2(v[0].x - 0.5 * t1.x, v[0].y - 0.5 * t1.y);
            double2 t3 = cmuli(cscale(csub(v[1], v[2]), c3), sign);
            v[0] = cadd(v[0], t1);
            v[1] = cadd(t2, t3);
            v[2] = csub(t2, t3);
        } else if (R == 4) {
            double2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
            double2 s13 = cadd(v[1], v[3]), d13 = cmuli(csub(v[1], v[3]), sign);
            v[0] = cadd(s02, s13);
            v[2] = csub(s02, s13);
            v[1] = cadd(d02, d13);
            v[3] = csub(d02, d13);
        }
        const int idxD = (j / Ns) * Ns * R + k;
        if (R <= 4) {
#pragma unroll
            for (int q = 0; q < R; q++) out[idxD + q * Ns] = v[q];
        } else {
            // direct R-point DFT: the inputs stay in registers, the roots come
            // from the LDS table, one output at a time
            const int e = n / R;
            for (int q = 0; q < R; q++) {
                double2 acc = v[0];
                int idx = 0;
                const int eq = q * e;
#pragma unroll
                for (int r = 1; r < R; r++) {
                    idx += eq;
                    if (idx >= n) idx -= n;
                    acc = cadd(acc, cmul(v[r], twid(tw, idx, nh, sign)));
                }
                out[idxD + q * Ns] = acc;
            }
        }
    }
}

// Generic radix: every thread computes whole output elements,
// out[idxD + q Ns] = sum_r in[j + r nbf] W_{Ns R}^{r (k + q Ns)}.
__device__ void fft_stage_gen(const double2 *a, double2 *b, int nb, int n, int Ns, int R, int sign,
                              const double2 *tw, int nh)
{
    const int nbf = n / R, step = n / (Ns * R);
    for (int o = threadIdx.x; o < nb * n; o += blockDim.x) {
        const int bb = o / n, rem = o - bb * n;
        const int k = rem % Ns, t = rem / Ns;
        const int q = t % R, jg = t / R;
        const int j = jg * Ns + k;
        const double2 *in = a + bb * n;
        const int e = (k + q * Ns) * step;   // < n
        double2 s = make_double2(0.0, 0.0);
        int idx = 0;
        for (int r = 0; r < R; r++) {
            s = cadd(s, cmul(in[j + r * nbf], twid(tw, idx, nh, sign)));
            idx += e;
            if (idx >= n) idx -= n;
        }
        b[o] = s;
    }
}

// Batched in-LDS FFT; returns the buffer that holds the result (a or b).
// Radices 2, 3 and 4 have hand-written butterflies; RB (> 4, chosen per
// length) gets a register DFT; any other factor uses the generic stage.
template <int RB>
__device__ double2 *lds_fft(double2 *a, double2 *b, int nb, const FftSeq &seq, int sign,
                            const double2 *tw, int nh)
{
    const int n = seq.n;
    int Ns = 1;
    for (int s = 0; s < seq.nst; s++) {
        const int R = seq.rad[s];
        if (R == 4)
            fft_stage<4>(a, b, nb, n, Ns, sign, tw, nh);
        else if (R == 2)
            fft_stage<2>(a, b, nb, n, Ns, sign, tw, nh);
        else if (R == 3)
            fft_stage<3>(a, b, nb, n, Ns, sign, tw, nh);
        else {
            if constexpr (RB > 4) {
                if (R == RB) fft_stage<RB>(a, b, nb, n, Ns, sign, tw, nh);
                else fft_stage_gen(a, b, nb, n, Ns, R, sign, tw, nh);
            } else {
                fft_stage_gen(a, b, nb, n, Ns, R, sign, tw, nh);
            }
        }
        __syncthreads();
        double2 *t = a;
        a = b;
        b = t;
        Ns *= R;
    }
    return a;
}

// The register-DFT radix of a length: its largest prime factor if that is
// one of 5, 7, 11, 13, 17, 19, else 0.
int big_radix(const FftSeq &seq)
{
    int rb = 0;
    for (int s = 0; s < seq.nst; s++) {
        const int R = seq.rad[s];
        if (R == 5 || R == 7 || R == 11 || R == 13 || R == 17 || R == 19) rb = R > rb ? R : rb;
    }
    return rb;
}

// Dynamic LDS above 64 KiB has to be allowed per kernel; the request is
// checked against the device limit first so an oversize launch is an error,
// not a fault.
template <typename K>
static int allow_lds(K kernel, size_t bytes)
{
    static int limit = -1;
    if (limit < 0) {
        int devid = 0;
        RD_HIP(hipGetDevice(&devid));
        RD_HIP(hipDeviceGetAttribute(&limit, hipDeviceAttributeMaxSharedMemoryPerBlock, devid));
    }
    if ((long)bytes > (long)limit) {
        fprintf(stderr, "redondito: kernel needs %zu bytes of LDS, device allows %d\n", bytes, limit);
        return RD_ERR_ARG;
    }
    if (bytes > 65536) RD_HIP(hipFuncSetAttribute((const void *)kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
    return RD_OK;
}

// ---------------------------------------------------------------- phi
struct PhiPtrs {
    const double2 *F[6];
    double2 *Fo[6];
    const double *f[6];
    double *fo[6];
};

// 4 rows (two packed complex FFTs) of one plane per workgroup;
// blockIdx.y selects the field.  F[k][m][theta] -> f[k][phi][theta].
__global__ __launch_bounds__(256) void phi_inverse_kernel(PhiPtrs P, int lmax, int nlat, int nphi,
                                                          int nplanes, FftSeq seq, const double2 *tw)
{
    extern __shared__ double2 sm[];
    const int ntt = (nlat + 3) / 4;
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const int k = t / ntt, th0 = (t % ntt) * 4;
    if (k >= nplanes) return;
    const int L = lmax + 1;
    const double2 *F = P.F[blockIdx.y];
    double *f = P.fo[blockIdx.y];
    double2 *A = sm, *B = sm + 2 * nphi, *twl = sm + 4 * nphi;
    const int nh = load_twiddles(twl, tw, nphi);
    for (int e = threadIdx.x; e < 2 * nphi; e += blockDim.x) {
        const int b = e / nphi, m = e - b * nphi;
        const int r1 = th0 + 2 * b, r2 = r1 + 1;
        double2 z = make_double2(0.0, 0.0);
        int mm = -1;
        bool cj = false;
        if (m <= lmax) mm = m;
        else if (m >= nphi - lmax) { mm = nphi - m; cj = true; }
        if (mm >= 0) {
            const long base = ((long)k * L + mm) * nlat;
            double2 f1 = r1 < nlat ? F[base + r1] : make_double2(0.0, 0.0);
            double2 f2 = r2 < nlat ? F[base + r2] : make_double2(0.0, 0.0);
            if (mm == 0) { f1.y = 0.0; f2.y = 0.0; }
            if (cj) { f1.y = -f1.y; f2.y = -f2.y; }
            z = make_double2(f1.x - f2.y, f1.y + f2.x);
        }
        A[e] = z;
    }
    __syncthreads();
    double2 *R = lds_fft<0>(A, B, 2, seq, +1, twl, nh);
    for (int e = threadIdx.x; e < 2 * nphi; e += blockDim.x) {
        const int b = e / nphi, j = e - b * nphi;
        const int r1 = th0 + 2 * b, r2 = r1 + 1;
        const double2 z = R[e];
        const long o = ((long)k * nphi + j) * nlat;
        if (r1 < nlat) f[o + r1] = z.x;
        if (r2 < nlat) f[o + r2] = z.y;
    }
}

// f[k][phi][theta] -> F[k][m][theta] = (2 pi / nphi) sum_j f e^{-i m phi_j}, m <= lmax.
__global__ __launch_bounds__(256) void phi_forward_kernel(PhiPtrs P, int lmax, int nlat, int nphi,
                                                          int nplanes, FftSeq seq, const double2 *tw)
{
    extern __shared__ double2 sm[];
    const int ntt = (nlat + 3) / 4;
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const int k = t / ntt, th0 = (t % ntt) * 4;
    if (k >= nplanes) return;
    const int L = lmax + 1;
    const double *f = P.f[blockIdx.y];
    double2 *F = P.Fo[blockIdx.y];
    double2 *A = sm, *B = sm + 2 * nphi, *twl = sm + 4 * nphi;
    const int nh = load_twiddles(twl, tw, nphi);
    for (int e = threadIdx.x; e < 2 * nphi; e += blockDim.x) {
        const int b = e / nphi, j = e - b * nphi;
        const int r1 = th0 + 2 * b, r2 = r1 + 1;
        const long o = ((long)k * nphi + j) * nlat;
        A[e] = make_double2(r1 < nlat ? f[o + r1] : 0.0, r2 < nlat ? f[o + r2] : 0.0);
    }
    __syncthreads();
    double2 *R = lds_fft<0>(A, B, 2, seq, -1, twl, nh);
    const double sc = 0.5 * (2.0 * M_PI / nphi);
    for (int e = threadIdx.x; e < 2 * L; e += blockDim.x) {
        const int b = e / L, m = e - b * L;
        const int r1 = th0 + 2 * b, r2 = r1 + 1;
        const double2 zp = R[b * nphi + m], zm = R[b * nphi + (m == 0 ? 0 : nphi - m)];
        // X1 = (Z[m] + conj Z[-m]) / 2, X2 = (Z[m] - conj Z[-m]) / (2i)
        double2 x1 = make_double2(sc * (zp.x + zm.x), sc * (zp.y - zm.y));
        double2 x2 = make_double2(sc * (zp.y + zm.y), -sc * (zp.x - zm.x));
        if (m == 0) { x1.y = 0.0; x2.y = 0.0; }
        const long base = ((long)k * L + m) * nlat;
        if (r1 < nlat) F[base + r1] = x1;
        if (r2 < nlat) F[base + r2] = x2;
    }
}

// One (plane, theta) row per workgroup: Z_x = U_x + i W_x for x = r, theta,
// phi; three inverse FFTs give u_x + i w_x on the nphi points; the pointwise
// product N = u x w (prodvec, src/pseudo/pseudospec_hd.f90:235-248) is packed
// as (N_r + i N_theta, N_phi) and transformed forward.  Rows of Fout may alias
// rows of Fin: each workgroup reads its own row completely before writing it.
__global__ __launch_bounds__(512) void phi_cross_kernel(PhiPtrs P, int lmax, int nlat, int nphi,
                                                        int nplanes, FftSeq seq, const double2 *tw)
{
    extern __shared__ double2 sm[];
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const int k = t / nlat, th = t % nlat;
    if (k >= nplanes) return;
    const int L = lmax + 1;
    double2 *A = sm, *B = sm + 3 * nphi, *twl = sm + 6 * nphi;
    const int nh = load_twiddles(twl, tw, nphi);
    for (int e = threadIdx.x; e < 3 * nphi; e += blockDim.x) {
        const int x = e / nphi, m = e - x * nphi;
        double2 z = make_double2(0.0, 0.0);
        int mm = -1;
        bool cj = false;
        if (m <= lmax) mm = m;
        else if (m >= nphi - lmax) { mm = nphi - m; cj = true; }
        if (mm >= 0) {
            const long o = ((long)k * L + mm) * nlat + th;
            double2 u = P.F[x][o], w = P.F[3 + x][o];
            if (mm == 0) { u.y = 0.0; w.y = 0.0; }
            if (cj) { u.y = -u.y; w.y = -w.y; }
            z = make_double2(u.x - w.y, u.y + w.x);
        }
        A[e] = z;
    }
    __syncthreads();
    double2 *R = lds_fft<0>(A, B, 3, seq, +1, twl, nh);
    double2 *O = (R == A) ? B : A;
    for (int j = threadIdx.x; j < nphi; j += blockDim.x) {
        const double2 zr = R[j], zt = R[nphi + j], zp = R[2 * nphi + j];
        // u = (zr.x, zt.x, zp.x), w = (zr.y, zt.y, zp.y)
        const double nr = zt.x * zp.y - zp.x * zt.y;
        const double nt = zp.x * zr.y - zr.x * zp.y;
        const double np = zr.x * zt.y - zt.x * zr.y;
        O[j] = make_double2(nr, nt);
        O[nphi + j] = make_double2(np, 0.0);
    }
    __syncthreads();
    double2 *Y = lds_fft<0>(O, R, 2, seq, -1, twl, nh);
    const double sc = 0.5 * (2.0 * M_PI / nphi);
    for (int m = threadIdx.x; m < L; m += blockDim.x) {
        const double2 yp = Y[m], ym = Y[m == 0 ? 0 : nphi - m];
        double2 xr = make_double2(sc * (yp.x + ym.x), sc * (yp.y - ym.y));
        double2 xt = make_double2(sc * (yp.y + ym.y), -sc * (yp.x - ym.x));
        double2 xp = cscale(Y[nphi + m], 2.0 * sc);
        if (m == 0) { xr.y = 0.0; xt.y = 0.0; xp.y = 0.0; }
        const long o = ((long)k * L + m) * nlat + th;
        P.Fo[0][o] = xr;
        P.Fo[1][o] = xt;
        P.Fo[2][o] = xp;
    }
}

int launch_phi_inverse(Plan *p, int nfields, int nplanes, const double2 *const *F, double *const *f)
{
    PhiPtrs P = {};
    for (int i = 0; i < nfields; i++) {
        P.F[i] = F[i];
        P.fo[i] = f[i];
    }
    const int ntt = (p->nlat + 3) / 4;
    size_t shm = sizeof(double2) * (4 * p->nphi + p->nphi / 2);
    {
        int lrc = allow_lds(phi_inverse_kernel, shm);
        if (lrc) return lrc;
    }
    hipLaunchKernelGGL(phi_inverse_kernel, dim3(nplanes * ntt, nfields), dim3(256), shm, p->stream, P,
                       p->lmax, p->nlat, p->nphi, nplanes, p->seq_phi, p->tw_phi);
    RD_LAUNCH_CHECK();
    return RD_OK;
}

int launch_phi_forward(Plan *p, int nfields, int nplanes, const double *const *f, double2 *const *F)
{
    PhiPtrs P = {};
    for (int i = 0; i < nfields; i++) {
        P.f[i] = f[i];
        P.Fo[i] = F[i];
    }
    const int ntt = (p->nlat + 3) / 4;
    size_t shm = sizeof(double2) * (4 * p->nphi + p->nphi / 2);
    {
        int lrc = allow_lds(phi_forward_kernel, shm);
        if (lrc) return lrc;
    }
    hipLaunchKernelGGL(phi_forward_kernel, dim3(nplanes * ntt, nfields), dim3(256), shm, p->stream, P,
                       p->lmax, p->nlat, p->nphi, nplanes, p->seq_phi, p->tw_phi);
    RD_LAUNCH_CHECK();
    return RD_OK;
}

int launch_phi_cross(Plan *p, int nplanes, const double2 *const *Fin, double2 *const *Fout)
{
    PhiPtrs P = {};
    for (int i = 0; i < 6; i++) P.F[i] = Fin[i];
    for (int i = 0; i < 3; i++) P.Fo[i] = Fout[i];
    size_t shm = sizeof(double2) * (6 * p->nphi + p->nphi / 2);
    {
        int lrc = allow_lds(phi_cross_kernel, shm);
        if (lrc) return lrc;
    }
    hipLaunchKernelGGL(phi_cross_kernel, dim3(nplanes * p->nlat), dim3(512), shm, p->stream, P,
                       p->lmax, p->nlat, p->nphi, nplanes, p->seq_phi, p->tw_phi);
    RD_LAUNCH_CHECK();
    return RD_OK;
}

// ----------------------------------------------------------------- z
constexpr int ZTC = 8;   // lm columns per workgroup

struct ZArgs {
    double2 *a;
    const double2 *ain;
    const double *dir;
    const double *kz;
    long ld;
    int ncols, n, C, d, order;
};

// Continuation of the rows n-C..n-1 of every column held in LDS
// (column c at X + c*n), from rows 0..d-1 and n-C-d..n-C-1.
__device__ void fc_continue(double2 *X, int ncol, int n, int C, int d, const double *dir)
{
    for (int e = threadIdx.x; e < ncol * C; e += blockDim.x) {
        const int c = e / C, ii = e - c * C + 1;
        const double2 *x = X + c * n;
        double2 s = make_double2(0.0, 0.0);
        for (int jj = 1; jj <= d; jj++) {
            const double a1 = dir[(ii - 1) + C * (jj - 1)];
            const double a2 = dir[(C - ii) + C * (jj - 1)];
            const double2 xr = x[n - C - d + jj - 1], xl = x[d - jj];
            s.x += a1 * xr.x + a2 * xl.x;
            s.y += a1 * xr.y + a2 * xl.y;
        }
        X[c * n + n - C + ii - 1] = s;
    }
}

__device__ void z_load(double2 *X, const double2 *a, long ld, int c0, int ncols, int n, int rows)
{
    for (int e = threadIdx.x; e < rows * ZTC; e += blockDim.x) {
        const int r = e / ZTC, c = e - r * ZTC;
        X[c * n + r] = (c0 + c < ncols) ? a[(long)r * ld + c0 + c] : make_double2(0.0, 0.0);
    }
}

__device__ void z_store(const double2 *X, double2 *a, long ld, int c0, int ncols, int n, int rows,
                        double scale)
{
    for (int e = threadIdx.x; e < rows * ZTC; e += blockDim.x) {
        const int r = e / ZTC, c = e - r * ZTC;
        if (c0 + c < ncols) a[(long)r * ld + c0 + c] = cscale(X[c * n + r], scale);
    }
}

// mode 0: forward (continuation, FFT -1, /n); 1: inverse (FFT +1);
// 2: derivative of order `order` in (lm,r), output to a from ain.
template <int RB>
__global__ __launch_bounds__(512) void z_kernel(ZArgs A, int mode, FftSeq seq, const double2 *tw)
{
    extern __shared__ double2 sm[];
    const int n = A.n, c0 = blockIdx.x * ZTC;
    double2 *X = sm, *Y = sm + ZTC * n, *twl = sm + 2 * ZTC * n;
    const int nh = load_twiddles(twl, tw, n);
    const double2 *src = (mode == 2) ? A.ain : A.a;
    const int nr = n - A.C;
    if (mode == 1) {
        z_load(X, src, A.ld, c0, A.ncols, n, n);
        __syncthreads();
        double2 *R = lds_fft<RB>(X, Y, ZTC, seq, +1, twl, nh);
        z_store(R, A.a, A.ld, c0, A.ncols, n, n, 1.0);
        return;
    }
    z_load(X, src, A.ld, c0, A.ncols, n, A.C > 0 ? nr : n);
    __syncthreads();
    if (A.C > 0) {
        fc_continue(X, ZTC, n, A.C, A.d, A.dir);
        __syncthreads();
    }
    double2 *R = lds_fft<RB>(X, Y, ZTC, seq, -1, twl, nh);
    if (mode == 0) {
        z_store(R, A.a, A.ld, c0, A.ncols, n, n, 1.0 / n);
        return;
    }
    // (i kz)^p / n, then the inverse transform
    const int p = A.order;
    for (int e = threadIdx.x; e < ZTC * n; e += blockDim.x) {
        const int r = e % n;
        double kp = 1.0;
        for (int q = 0; q < p; q++) kp *= A.kz[r];
        double2 v = cscale(R[e], kp / n);
        const int pm = p & 3;
        if (pm == 1) v = make_double2(-v.y, v.x);
        else if (pm == 2) v = make_double2(-v.x, -v.y);
        else if (pm == 3) v = make_double2(v.y, -v.x);
        R[e] = v;
    }
    __syncthreads();
    double2 *O = (R == X) ? Y : X;
    double2 *R2 = lds_fft<RB>(R, O, ZTC, seq, +1, twl, nh);
    z_store(R2, A.a, A.ld, c0, A.ncols, n, n, 1.0);
}

static int z_launch(Plan *p, int mode, double2 *a, const double2 *ain, int order)
{
    ZArgs A;
    A.a = a;
    A.ain = ain;
    A.dir = p->d_dir;
    A.kz = p->d_kz;
    A.ld = p->nlm;
    A.ncols = p->nlm;
    A.n = p->nz;
    A.C = p->C;
    A.d = p->d;
    A.order = order;
    size_t shm = sizeof(double2) * (2 * ZTC + 1) * p->nz;
    const dim3 grid((p->nlm + ZTC - 1) / ZTC), block(512);
#define RD_Z_CASE(RBV)                                                                         \
    case RBV: {                                                                                \
        int lrc = allow_lds(z_kernel<RBV>, shm);                                               \
        if (lrc) return lrc;                                                                   \
        hipLaunchKernelGGL(z_kernel<RBV>, grid, block, shm, p->stream, A, mode, p->seq_z, p->tw_z); \
        break;                                                                                 \
    }
    switch (big_radix(p->seq_z)) {
        RD_Z_CASE(0)
        RD_Z_CASE(5)
        RD_Z_CASE(7)
        RD_Z_CASE(11)
        RD_Z_CASE(13)
        RD_Z_CASE(17)
        RD_Z_CASE(19)
    default: return RD_ERR_ARG;
    }
#undef RD_Z_CASE
    RD_LAUNCH_CHECK();
    return RD_OK;
}

int launch_z_forward(Plan *p, double2 *a) { return z_launch(p, 0, a, nullptr, 0); }
int launch_z_inverse(Plan *p, double2 *a) { return z_launch(p, 1, a, nullptr, 0); }
int launch_derivr_r(Plan *p, int order, const double2 *a, double2 *b) { return z_launch(p, 2, b, a, order); }

// Radial step of the nonlinear term for a tile of lm columns:
//   W^ = FFT(continue(W))/n, Z^ likewise;
//   dW = IFFT(i kz W^), d2W = IFFT(-kz^2 W^), dZ = IFFT(i kz Z^)
//   u:     Q = l(l+1) W / r^2,  S = dW / r,  T = Z / r
//   omega: Q = l(l+1) Z / r^2,  S = dZ / r,  T = -(d2W - l(l+1) W / r^2) / r
// on the physical rows k < nr.
struct UxwRadArgs {
    const double2 *W, *Z;
    double2 *out[6];
    const double *dir, *kz, *z;
    const int *ell;
    long ld;
    int ncols, n, C, d;
};

__device__ void spec_mul_ikz(const double2 *S, double2 *D, const double *kz, int n, int p)
{
    for (int e = threadIdx.x; e < ZTC * n; e += blockDim.x) {
        const int r = e % n;
        double kp = 1.0;
        for (int q = 0; q < p; q++) kp *= kz[r];
        double2 v = cscale(S[e], kp);
        if (p == 1) v = make_double2(-v.y, v.x);
        else if (p == 2) v = make_double2(-v.x, -v.y);
        D[e] = v;
    }
}

template <int RB>
__global__ __launch_bounds__(512) void uxw_radial_kernel(UxwRadArgs A, FftSeq seq, const double2 *tw)
{
    extern __shared__ double2 sm[];
    const int n = A.n, c0 = blockIdx.x * ZTC, nr = n - A.C;
    double2 *X = sm, *Y = sm + ZTC * n, *S = sm + 2 * ZTC * n, *twl = sm + 3 * ZTC * n;
    const int nh = load_twiddles(twl, tw, n);
    const double inv_n = 1.0 / n;
    for (int fld = 0; fld < 2; fld++) {
        const double2 *src = fld == 0 ? A.W : A.Z;
        z_load(X, src, A.ld, c0, A.ncols, n, A.C > 0 ? nr : n);
        __syncthreads();
        if (A.C > 0) {
            fc_continue(X, ZTC, n, A.C, A.d, A.dir);
            __syncthreads();
        }
        double2 *R = lds_fft<RB>(X, Y, ZTC, seq, -1, twl, nh);
        for (int e = threadIdx.x; e < ZTC * n; e += blockDim.x) S[e] = cscale(R[e], inv_n);
        __syncthreads();
        const int nder = fld == 0 ? 2 : 1;
        for (int p = 1; p <= nder; p++) {
            spec_mul_ikz(S, X, A.kz, n, p);
            __syncthreads();
            double2 *D = lds_fft<RB>(X, Y, ZTC, seq, +1, twl, nh);
            for (int e = threadIdx.x; e < nr * ZTC; e += blockDim.x) {
                const int r = e / ZTC, c = e - r * ZTC;
                const int lm = c0 + c;
                if (lm >= A.ncols) continue;
                const double rr = A.z[r], r1 = 1.0 / rr, r2 = 1.0 / (rr * rr);
                const long o = (long)r * A.ld + lm;
                const double2 dv = D[c * n + r];
                if (fld == 0 && p == 1) {
                    A.out[1][o] = cscale(dv, r1);               // u: S
                } else if (fld == 0 && p == 2) {
                    const double ll = (double)A.ell[lm] * (A.ell[lm] + 1);
                    const double2 w = A.W[o];
                    const double2 dl = make_double2(dv.x - ll * w.x * r2, dv.y - ll * w.y * r2);
                    A.out[5][o] = cscale(dl, -r1);              // omega: T
                    A.out[0][o] = cscale(w, ll * r2);           // u: Q
                } else {
                    const double ll = (double)A.ell[lm] * (A.ell[lm] + 1);
                    const double2 zz = A.Z[o];
                    A.out[4][o] = cscale(dv, r1);               // omega: S
                    A.out[2][o] = cscale(zz, r1);               // u: T
                    A.out[3][o] = cscale(zz, ll * r2);          // omega: Q
                }
            }
            __syncthreads();
        }
    }
}

int launch_uxw_radial(Plan *p, const double2 *W, const double2 *Z, double2 *const *qst)
{
    UxwRadArgs A;
    A.W = W;
    A.Z = Z;
    for (int i = 0; i < 6; i++) A.out[i] = qst[i];
    A.dir = p->d_dir;
    A.kz = p->d_kz;
    A.z = p->d_z;
    A.ell = p->d_ell;
    A.ld = p->nlm;
    A.ncols = p->nlm;
    A.n = p->nz;
    A.C = p->C;
    A.d = p->d;
    size_t shm = sizeof(double2) * (3 * ZTC + 1) * p->nz;
    const dim3 grid((p->nlm + ZTC - 1) / ZTC), block(512);
#define RD_U_CASE(RBV)                                                                          \
    case RBV: {                                                                                 \
        int lrc = allow_lds(uxw_radial_kernel<RBV>, shm);                                       \
        if (lrc) return lrc;                                                                    \
        hipLaunchKernelGGL(uxw_radial_kernel<RBV>, grid, block, shm, p->stream, A, p->seq_z, p->tw_z); \
        break;                                                                                  \
    }
    switch (big_radix(p->seq_z)) {
        RD_U_CASE(0)
        RD_U_CASE(5)
        RD_U_CASE(7)
        RD_U_CASE(11)
        RD_U_CASE(13)
        RD_U_CASE(17)
        RD_U_CASE(19)
    default: return RD_ERR_ARG;
    }
#undef RD_U_CASE
    RD_LAUNCH_CHECK();
    return RD_OK;
}

}  // namespace rd
