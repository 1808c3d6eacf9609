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
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
// sign * i * a
__device__ __forceinline__ double2 cmuli(double2 a, int sign)
{
    return sign > 0 ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x);
}
// the same in FP32 (the FP32 phi kernels of the precision forms)
__device__ __forceinline__ float2 cmul(float2 a, float2 b)
{
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
__device__ __forceinline__ float2 cmuli(float2 a, int sign)
{
    return sign > 0 ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}
// Streaming (nontemporal) store of one complex value: for outputs read by a
// later kernel, so that they do not displace lines the kernel (or its
// neighbours on the XCD) read again.
__device__ __forceinline__ void st_stream(double2 *p, double2 v)
{
    typedef double vd2 __attribute__((ext_vector_type(2)));
    const vd2 t = {v.x, v.y};
    __builtin_nontemporal_store(t, (vd2 *)p);
}

// scalar type and constructor of a complex type
template <typename V> struct Cx;
template <> struct Cx<double2> {
    using S = double;
    static __device__ __forceinline__ double2 make(double x, double y) { return make_double2(x, y); }
};
template <> struct Cx<float2> {
    using S = float;
    static __device__ __forceinline__ float2 make(double x, double y) { return make_float2((float)x, (float)y); }
};
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
#pragma unroll 1
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

// The register-DFT kernel variant for the radial FFT: in Rader form the
// inner length is short and only radices 5 and 7 pay for their registers.
int big_radix(const FftSeq &seq);
static int z_radix(const ZFft &Z)
{
    const int rb = big_radix(Z.seq);
    return (Z.rader && rb > 7) ? 0 : rb;
}

// The register-DFT radix of a length: its largest prime factor if that is
// one of 5, 7, 11, 13, else 0 (the generic stage is faster for 17 and 19,
// whose register DFTs need too many VGPRs).
int big_radix(const FftSeq &seq)
{
    int rb = 0;
    for (int s = 0; s < seq.nst; s++) {
        const int R = seq.rad[s];
        if (R == 5 || R == 7 || R == 11 || R == 13) rb = R > rb ? R : rb;
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

// Workgroup sums of NV values per thread.  Lane 0 of every wave posts its
// wave's sums to red, one barrier, then every thread adds the NW posts in the
// same order, so all threads hold bit-identical totals.  The barrier also
// publishes any LDS writes made before the call.
template <int NV, int NW>
__device__ __forceinline__ void block_sums(double (&v)[NV], double (*red)[NV])
{
#pragma unroll
    for (int i = 0; i < NV; i++)
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) v[i] += __shfl_xor(v[i], off);
    if ((threadIdx.x & 63) == 0)
#pragma unroll
        for (int i = 0; i < NV; i++) red[threadIdx.x >> 6][i] = v[i];
    __syncthreads();
#pragma unroll
    for (int i = 0; i < NV; i++) {
        double s = 0.0;
        for (int q = 0; q < NW; q++) s += red[q][i];
        v[i] = s;
    }
}

// Two real rows u, w packed as one complex sequence u + i a w share the FFT's
// rounding: each comes back with an error of about eps (|u| + a |w|).  With a
// a power of two within a factor 2 of |u| / |w| (su, sw sums of squares) the
// packing costs nothing beyond the rounding of the rows themselves, and
// dividing by a afterwards is exact.
// A value every lane holds with the same bits, moved to scalar registers.
__device__ __forceinline__ double wave_uniform(double x)
{
    const int lo = __builtin_amdgcn_readfirstlane(__double2loint(x));
    const int hi = __builtin_amdgcn_readfirstlane(__double2hiint(x));
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double pack_scale(double su, double sw)
{
    if (!(su > 0.0) || !(sw > 0.0)) return 1.0;
    int e;
    (void)frexp(sqrt(su / sw), &e);
    return ldexp(1.0, e);
}

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

// Row (k, th) of the six Fourier fields (U_r, U_t, U_p, W_r, W_t, W_p), as the
// Hermitian-extended sequences packed Z_x = U_x + i a W_x in A[x nphi + m]
// (B is scratch), a = pack_scale(|U|^2, |W|^2); returns 1/a.  Ends with a
// barrier.
constexpr int PHI_GEN_THREADS = 512;

__device__ double phi_load_pairs(const PhiPtrs &P, int k, int th, int lmax, int nlat, int nphi, double2 *A,
                                 double2 *B, double (*red)[2])
{
    const int L = lmax + 1;
    double s[2] = {0.0, 0.0};
    for (int e = threadIdx.x; e < 3 * nphi; e += blockDim.x) {
        const int x = e / nphi, m = e - x * nphi;
        double2 u = make_double2(0.0, 0.0), w = u;
        int mm = -1;
        bool cj = false;
        if (m <= lmax) mm = m;
        else if (m >= nphi - lmax) { mm = nphi - m; cj = true; }
        if (mm >= 0) {
            const long o = ((long)k * L + mm) * nlat + th;
            u = P.F[x][o];
            w = P.F[3 + x][o];
            if (mm == 0) { u.y = 0.0; w.y = 0.0; }
            if (cj) { u.y = -u.y; w.y = -w.y; }
        }
        s[0] += u.x * u.x + u.y * u.y;
        s[1] += w.x * w.x + w.y * w.y;
        A[e] = u;
        B[e] = w;
    }
    block_sums<2, PHI_GEN_THREADS / 64>(s, red);
    const double a = pack_scale(s[0], s[1]);
    for (int e = threadIdx.x; e < 3 * nphi; e += blockDim.x) {
        const double2 u = A[e], w = B[e];
        A[e] = make_double2(u.x - a * w.y, u.y + a * w.x);
    }
    __syncthreads();
    return 1.0 / a;
}

// One (plane, theta) row per workgroup: three inverse FFTs of the balanced
// packs give u_x + i a w_x on the nphi points; the pointwise product
// N = u x w (prodvec, src/pseudo/pseudospec_hd.f90:235-248) is packed as
// (N_theta + i N_phi, N_r) and transformed forward (the horizontal pair,
// from which S and T come, apart from the radial part).  Rows of Fout may
// alias rows of Fin: each workgroup reads its own row completely before
// writing it.
__global__ __launch_bounds__(PHI_GEN_THREADS) void phi_cross_kernel(PhiPtrs P, int lmax, int nlat, int nphi,
                                                                    int nplanes, FftSeq seq, const double2 *tw)
{
    extern __shared__ double2 sm[];
    __shared__ double red[PHI_GEN_THREADS / 64][2];
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const int k = t / nlat, th = t % nlat;
    if (k >= nplanes) return;
    const int L = lmax + 1;
    double2 *A = sm, *B = sm + 3 * nphi, *twl = sm + 6 * nphi;
    const int nh = load_twiddles(twl, tw, nphi);
    const double ia = phi_load_pairs(P, k, th, lmax, nlat, nphi, A, B, red);
    double2 *R = lds_fft<0>(A, B, 3, seq, +1, twl, nh);
    double2 *O = (R == A) ? B : A;
    for (int j = threadIdx.x; j < nphi; j += blockDim.x) {
        const double2 zr = R[j], zt = R[nphi + j], zp = R[2 * nphi + j];
        // u = (zr.x, zt.x, zp.x), w = (zr.y, zt.y, zp.y) / a
        const double wr = zr.y * ia, wt = zt.y * ia, wp = zp.y * ia;
        const double nr = zt.x * wp - zp.x * wt;
        const double nt = zp.x * wr - zr.x * wp;
        const double np = zr.x * wt - zt.x * wr;
        O[j] = make_double2(nt, np);
        O[nphi + j] = make_double2(nr, 0.0);
    }
    __syncthreads();
    double2 *Y = lds_fft<0>(O, R, 2, seq, -1, twl, nh);
    const double sc = 0.5 * (2.0 * M_PI / nphi);
    for (int m = threadIdx.x; m < L; m += blockDim.x) {
        const double2 yp = Y[m], ym = Y[m == 0 ? 0 : nphi - m];
        double2 xt = make_double2(sc * (yp.x + ym.x), sc * (yp.y - ym.y));
        double2 xp = make_double2(sc * (yp.y + ym.y), -sc * (yp.x - ym.x));
        double2 xr = cscale(Y[nphi + m], 2.0 * sc);
        if (m == 0) { xr.y = 0.0; xt.y = 0.0; xp.y = 0.0; }
        const long o = ((long)k * L + m) * nlat + th;
        P.Fo[0][o] = xr;
        P.Fo[1][o] = xt;
        P.Fo[2][o] = xp;
    }
}

// One (plane, theta) row per workgroup: three inverse FFTs of the balanced
// packs give u_x + i a g_x on the nphi points; the pointwise scalar product
// D = u . g is transformed forward into Fo[0].  Fo[0] may alias a row of F:
// each workgroup reads its own row completely before writing it.
__global__ __launch_bounds__(PHI_GEN_THREADS) void phi_dot_kernel(PhiPtrs P, int lmax, int nlat, int nphi,
                                                                  int nplanes, FftSeq seq, const double2 *tw)
{
    extern __shared__ double2 sm[];
    __shared__ double red[PHI_GEN_THREADS / 64][2];
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const int k = t / nlat, th = t % nlat;
    if (k >= nplanes) return;
    const int L = lmax + 1;
    double2 *A = sm, *B = sm + 3 * nphi, *twl = sm + 6 * nphi;
    const int nh = load_twiddles(twl, tw, nphi);
    const double ia = phi_load_pairs(P, k, th, lmax, nlat, nphi, A, B, red);
    double2 *R = lds_fft<0>(A, B, 3, seq, +1, twl, nh);
    double2 *O = (R == A) ? B : A;
    for (int j = threadIdx.x; j < nphi; j += blockDim.x) {
        const double2 zr = R[j], zt = R[nphi + j], zp = R[2 * nphi + j];
        O[j] = make_double2((zr.x * zr.y + zt.x * zt.y + zp.x * zp.y) * ia, 0.0);
    }
    __syncthreads();
    double2 *Y = lds_fft<0>(O, R, 1, seq, -1, twl, nh);
    const double sc = 2.0 * M_PI / nphi;
    for (int m = threadIdx.x; m < L; m += blockDim.x) {
        double2 x = cscale(Y[m], sc);
        if (m == 0) x.y = 0.0;
        P.Fo[0][((long)k * L + m) * nlat + th] = x;
    }
}


// ------------------------------------------------ phi / cross, nphi = 768
// The same step for nphi = 3 x 16 x 16 with at most 256 stored m:
//   inverse: m = 256 a + b, phi index j = 3 q + r.  The middle block a = 1 of
//   the padded spectrum is zero, so the radix-3 step is one complex FMA,
//   fused with the loads:  Y_r[b] = (Z[b] + w3^{2r} Z[512 + b]) w768^{b r};
//   then z[3 q + r] = IDFT_256(Y_r)[q], run as 16 x 16 with both 16-point
//   DFTs in registers and one LDS transpose between them.
//   forward: decimation in time j = 3 n1 + n2, the two 256-point DFTs again
//   as 16 x 16, and only the outputs m <= lmax (and their mirrors) are
//   formed.
// Seven barriers per row instead of one per Stockham pass.

// Length-16 DFT in registers, out[k] = sum_j v[j] exp(sign 2 pi i j k / 16),
// as 4 x 4 with the w16 twiddles as constants.
template <typename V>
__device__ __forceinline__ void dft4_inplace(V &a, V &b, V &c, V &d, int sign)
{
    const V s02 = cadd(a, c), d02 = csub(a, c);
    const V s13 = cadd(b, d), d13 = cmuli(csub(b, d), sign);
    a = cadd(s02, s13);
    c = csub(s02, s13);
    b = cadd(d02, d13);
    d = csub(d02, d13);
}

template <typename V>
__device__ __forceinline__ void dft16(V (&v)[16], int sign)
{
    // step 1: 4-point DFTs over j1 for each j0 (elements v[4 j1 + j0])
#pragma unroll
    for (int j0 = 0; j0 < 4; j0++) dft4_inplace(v[j0], v[4 + j0], v[8 + j0], v[12 + j0], sign);
    // after step 1, v[4 k0 + j0] holds A[j0][k0]; twiddle by w16^(j0 k0)
    const double c1 = 0.92387953251128675613, s1 = 0.38268343236508977173, c2 = 0.70710678118654752440;
    const double sg = sign > 0 ? 1.0 : -1.0;
    const V w1 = Cx<V>::make(c1, sg * s1), w2 = Cx<V>::make(c2, sg * c2), w3 = Cx<V>::make(s1, sg * c1);
    const V w6 = Cx<V>::make(-c2, sg * c2), w9 = Cx<V>::make(-c1, -sg * s1);
    v[4 + 1] = cmul(v[4 + 1], w1);   // j0 = 1, k0 = 1
    v[8 + 1] = cmul(v[8 + 1], w2);   // j0 = 1, k0 = 2
    v[12 + 1] = cmul(v[12 + 1], w3); // j0 = 1, k0 = 3
    v[4 + 2] = cmul(v[4 + 2], w2);   // j0 = 2, k0 = 1
    v[8 + 2] = cmuli(v[8 + 2], sign); // j0 = 2, k0 = 2: w16^4 = sign i
    v[12 + 2] = cmul(v[12 + 2], w6); // j0 = 2, k0 = 3
    v[4 + 3] = cmul(v[4 + 3], w3);   // j0 = 3, k0 = 1
    v[8 + 3] = cmul(v[8 + 3], w6);   // j0 = 3, k0 = 2
    v[12 + 3] = cmul(v[12 + 3], w9); // j0 = 3, k0 = 3
    // step 3: 4-point DFTs over j0 for each k0 -> out[k0 + 4 k1]
#pragma unroll
    for (int k0 = 0; k0 < 4; k0++) dft4_inplace(v[4 * k0], v[4 * k0 + 1], v[4 * k0 + 2], v[4 * k0 + 3], sign);
    // now v[4 k0 + k1] = out[k0 + 4 k1]; reorder to natural order
    V o[16];
#pragma unroll
    for (int k0 = 0; k0 < 4; k0++)
#pragma unroll
        for (int k1 = 0; k1 < 4; k1++) o[k0 + 4 * k1] = v[4 * k0 + k1];
#pragma unroll
    for (int i = 0; i < 16; i++) v[i] = o[i];
}

// T16[(q-1) 16 + b] = e^{-2 pi i b q/256}, q = 1..15, b = 0..15, from the
// global table exp(-2 pi i t/n) of the n = 256 st roots of unity: the
// twiddles of a 16 x 16 pass, read at a constant offset per q.
template <typename V>
__device__ __forceinline__ void fill_t16(V *T16, const double2 *tw, int st, int tid)
{
    if (tid < 15 * 16) {
        const double2 w = tw[st * (tid & 15) * ((tid >> 4) + 1)];
        T16[tid] = Cx<V>::make(w.x, w.y);
    }
}

// The twiddles e^{-2 pi i b q/256}, q = 1..15, of lane b of a 16 x 16 pass:
// T16 holds them all, but only q = 1, 2, 3, 4, 8, 12 are read, and the
// others are formed as one complex product w_{4a} w_r (q = 4a + r), one
// rounding beyond the table's: six LDS reads per lane instead of fifteen.
template <typename V>
struct Tw16 {
    V t1, t2, t3, t4, t8, t12;
    __device__ __forceinline__ Tw16(const V *T16, int b)
        : t1(T16[b]), t2(T16[16 + b]), t3(T16[32 + b]), t4(T16[48 + b]), t8(T16[112 + b]), t12(T16[176 + b])
    {
    }
    __device__ __forceinline__ V get(int q) const
    {
        const int a = q >> 2, r = q & 3;
        const V hi = a == 1 ? t4 : (a == 2 ? t8 : t12);
        const V lo = r == 1 ? t1 : (r == 2 ? t2 : t3);
        if (r == 0) return hi;
        if (a == 0) return lo;
        return cmul(hi, lo);
    }
};

// One (plane, theta) row per workgroup, in place in a single LDS buffer so
// that four rows run per CU.  Every phase reads its inputs into registers,
// waits at a barrier, then writes its outputs over the same buffer:
//   R3  radix-3 step of the inverse transform of the packed sequences
//       z_x = u_x + i a w_x (x = r, theta, phi; a a power of two that
//       balances u and w): the spectrum block 256..511 is
//       zero, so Y[3x+r][b] = (Z[b] + e^{2 pi i 512 r/768} Z[512+b]) e^{2 pi i b r/768}
//   P1  first 16-point pass of the 256-point inverse FFTs over b = 16 b2 + b1
//   P2  second inverse pass: grid values z_x at j = 3 (q1 + 16 q2) + r
//   CR  N = u x w packed as (N_theta + i N_phi, N_r)
//   F1  first forward pass over n1 = 16 a2 + a1 of j = 3 n1 + n2
//   F2  second forward pass: Yf[3s+n2][k] = 256-point spectra
//   FIN X[m] = sum_n2 e^{-2 pi i n2 m/768} Yf[n2][m mod 256], unpacking of the
//       real spectra, scaling and store.
// The 16 x 16 transposes use an XOR swizzle (column q ^ row) instead of
// padding.  The 16-point passes take their twiddles from a [q][b] table of
// the 256th roots in LDS (one read at a constant offset per value, lanes of
// consecutive b on consecutive slots); R3 and FIN read e^{-2 pi i b/768} and
// e^{-2 pi i 2b/768} straight from the global table of the 768 roots.
// The phases are written for V = double2 (the instance launched) and for
// V = float2 (the row in FP32, loads, pack scale and stores in FP64); the FP32
// instance measured slower (0.745 against 0.661 ms at lmax 255, 128 planes)
// and is not launched.
constexpr int PHI_BUF = 9 * 256;

// Thread index with the wave index permuted per workgroup (XOR with a hash of
// blockIdx.x, LW = log2 of the waves per workgroup).  The 16-point passes give
// their work to the first few waves; rotating the roles from row to row keeps
// any one SIMD of the CU from always receiving the busiest wave.
template <int LW>
__device__ __forceinline__ int rotated_tid(int tid)
{
    return tid ^ (int)((((unsigned)blockIdx.x * 2654435761u) >> (32 - LW)) << 6);
}

// Orders LDS accesses among the lanes of one wave: a wave's LDS instructions
// complete in issue order, so a compiler fence at wavefront scope is enough
// when every lane that reads a slot is in the wave that wrote it.
__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Coefficient m = min(tid, lmax) of the six fields of row (k, th), loaded
// unconditionally: the zeros above lmax and the dropped imaginary part of
// m = 0 are applied where the values are used, so that nothing here (no
// branch, no select) makes the compiler wait for the loads.
__device__ __forceinline__ void cross768_load(const PhiPtrs &P, int lmax, int nlat, int k, int th,
                                              double2 (&v)[6])
{
    const int L = lmax + 1, m = min((int)threadIdx.x, lmax);
#pragma unroll
    for (int f = 0; f < 6; f++) v[f] = P.F[f][((long)k * L + m) * nlat + th];
}

// The phases of one row from its loaded coefficients v (see above); ends
// with the stores, the LDS buffer still read by FIN (a barrier is needed
// before the buffer is written again).
template <typename V>
__device__ __forceinline__ void cross768_row(const PhiPtrs &P, int lmax, int nlat, int k, int th,
                                             const double2 (&vin)[6], const V *T16, V *B, double (*red)[2],
                                             double *ia_sp, double2 g1, double2 g2, int rtid)
{
    using S = typename Cx<V>::S;
    constexpr int N = 768;
    double &ia_s = *ia_sp;
    const int L = lmax + 1;
    const int tid = threadIdx.x;
    // the row's six coefficient vectors [f][m], m = 0..255 (zero above lmax),
    // and the sums of squares of u and of w
    // (the thread's own coefficients stay in registers for R3, which reads
    // only the mirrored ones back)
    double s[2] = {0.0, 0.0};
    const bool live = tid <= lmax;
    V own[6];
#pragma unroll
    for (int f = 0; f < 6; f++) {
        const double2 v = make_double2(live ? vin[f].x : 0.0, (live && tid > 0) ? vin[f].y : 0.0);
        own[f] = Cx<V>::make(v.x, v.y);
        B[f * 256 + tid] = own[f];
        s[f / 3] += v.x * v.x + v.y * v.y;
    }
    block_sums<2, 4>(s, red);
    // balanced packs z_x = u_x + i a w_x; 1/a kept in LDS for CR (read there
    // after several barriers)
    const double a = wave_uniform(pack_scale(s[0], s[1]));
    const S as = (S)a;   // a power of two: exact in FP32 as well
    if (tid == 0) ia_s = 1.0 / a;

    // ---- R3 -> Y[3x+r][b]
    {
        const int b = tid, mm = (256 - b) & 255;
        V zb[3], zc[3];
#pragma unroll
        for (int x = 0; x < 3; x++) {
            const V u = own[x], w = cscale(own[3 + x], as);   // b = tid
            const V uc = B[x * 256 + mm], wc = cscale(B[(3 + x) * 256 + mm], as);
            zb[x] = V{u.x - w.y, u.y + w.x};
            zc[x] = b == 0 ? V{S(0), S(0)} : V{uc.x + wc.y, -uc.y + wc.x};
        }
        __syncthreads();
        // e^{2 pi i 512/768} and e^{2 pi i 256/768}
        const S h = S(0.86602540378443864676);
        const V w3a = V{S(-0.5), -h}, w3b = V{S(-0.5), h};
        const V t1 = Cx<V>::make(g1.x, -g1.y), t2 = Cx<V>::make(g2.x, -g2.y);
#pragma unroll
        for (int x = 0; x < 3; x++) {
            B[(3 * x) * 256 + b] = cadd(zb[x], zc[x]);
            B[(3 * x + 1) * 256 + b] = cmul(cadd(zb[x], cmul(zc[x], w3a)), t1);
            B[(3 * x + 2) * 256 + b] = cmul(cadd(zb[x], cmul(zc[x], w3b)), t2);
        }
    }
    __syncthreads();
    // ---- P1: T[xr][b1][q1] = e^{2 pi i b1 q1/256} sum_b2 Y[xr][16 b2 + b1] e^{2 pi i b2 q1/16}
    // The three sequences xr = 3x + r of field x are worked by lanes 0..47 of
    // wave x, and P1 and P2 of them touch only slots x N .. x N + 767
    // (P2 writes the grid order j = 3 (q1 + 16 q2) + r): wave-local ordering
    // suffices until CR.
    {
        const int x = rtid >> 6, r = (rtid & 63) >> 4, xr = 3 * x + r, b1 = rtid & 15;
        const bool act = rtid < 192 && r < 3;
        V v[16];
        if (act) {
#pragma unroll
            for (int b2 = 0; b2 < 16; b2++) v[b2] = B[xr * 256 + 16 * b2 + b1];
        }
        wave_lds_sync();
        if (act) {
            dft16(v, +1);
            const Tw16<V> tw(T16, b1);
            B[(xr * 16 + b1) * 16 + b1] = v[0];
#pragma unroll
            for (int q1 = 1; q1 < 16; q1++) B[(xr * 16 + b1) * 16 + (q1 ^ b1)] = cmul(v[q1], cconj(tw.get(q1)));
        }
    }
    wave_lds_sync();
    // ---- P2: grid values z_x at j = 3 (q1 + 16 q2) + r, stored at x N + j
    {
        int t1 = rtid;   // opaque copy, as in F2
        asm volatile("" : "+v"(t1));
        const int x = t1 >> 6, r = (t1 & 63) >> 4, xr = 3 * x + r, q1 = t1 & 15;
        const bool act = t1 < 192 && r < 3;
        V v[16];
        if (act) {
#pragma unroll
            for (int b1 = 0; b1 < 16; b1++) v[b1] = B[(xr * 16 + b1) * 16 + (q1 ^ b1)];
        }
        wave_lds_sync();
        if (act) {
            dft16(v, +1);
#pragma unroll
            for (int q2 = 0; q2 < 16; q2++) B[x * N + 3 * (q1 + 16 * q2) + r] = v[q2];
        }
    }
    __syncthreads();
    // ---- CR: (N_theta + i N_phi) at j, N_r at N + j
    {
        V o0[3], o1[3];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const int j = tid + 256 * i;
            const V zr = B[j], zt = B[N + j], zp = B[2 * N + j];
            const S ia = (S)ia_s, wr = zr.y * ia, wt = zt.y * ia, wp = zp.y * ia;
            o0[i] = V{zp.x * wr - zr.x * wp, zr.x * wt - zt.x * wr};
            o1[i] = V{zt.x * wp - zp.x * wt, S(0)};
        }
        // each thread overwrites only the slots it has just read
#pragma unroll
        for (int i = 0; i < 3; i++) {
            B[tid + 256 * i] = o0[i];
            B[N + tid + 256 * i] = o1[i];
        }
    }
    __syncthreads();
    // ---- F1: F[3s+n2][a1][k1] = e^{-2 pi i a1 k1/256} sum_a2 P_s[3 (16 a2 + a1) + n2] e^{-2 pi i a2 k1/16}
    // Packed sequence s is worked by lanes 0..47 of wave s, and F1 and F2 of
    // it touch only slots s N .. s N + 767: wave-local ordering suffices
    // until FIN.
    {
        const int s = rtid >> 6, rem = rtid & 63, n2 = rem >> 4, a1 = rem & 15;
        const bool act = rtid < 128 && rem < 48;
        V v[16];
        if (act) {
#pragma unroll
            for (int a2 = 0; a2 < 16; a2++) v[a2] = B[s * N + 3 * (16 * a2 + a1) + n2];
        }
        wave_lds_sync();
        if (act) {
            dft16(v, -1);
            const Tw16<V> tw(T16, a1);
            B[((s * 3 + n2) * 16 + a1) * 16 + a1] = v[0];
#pragma unroll
            for (int k1 = 1; k1 < 16; k1++) B[((s * 3 + n2) * 16 + a1) * 16 + (k1 ^ a1)] = cmul(v[k1], tw.get(k1));
        }
    }
    wave_lds_sync();
    // ---- F2: Yf[3s+n2][k1 + 16 k2]
    {
        // indices recomputed from an opaque copy of tid: shared with F1 they
        // stay live across the barrier and cost a spill
        int t2 = rtid;
        asm volatile("" : "+v"(t2));
        const int s = t2 >> 6, rem = t2 & 63, n2 = rem >> 4, k1 = rem & 15;
        const bool act = t2 < 128 && rem < 48;
        V v[16];
        if (act) {
#pragma unroll
            for (int a1 = 0; a1 < 16; a1++) v[a1] = B[((s * 3 + n2) * 16 + a1) * 16 + (k1 ^ a1)];
        }
        wave_lds_sync();
        if (act) {
            dft16(v, -1);
#pragma unroll
            for (int k2 = 0; k2 < 16; k2++) B[(s * 3 + n2) * 256 + k1 + 16 * k2] = v[k2];
        }
    }
    __syncthreads();
    // ---- FIN
    const double sc = 0.5 * (2.0 * M_PI / N);
    const int m = tid;
    if (m < L) {
        const int mneg = (256 - m) & 255;
        const V c1 = Cx<V>::make(g1.x, g1.y), c2 = Cx<V>::make(g2.x, g2.y);
        const V xpv = cadd(B[m], cadd(cmul(B[256 + m], c1), cmul(B[512 + m], c2)));
        const V x1v = cadd(B[768 + m], cadd(cmul(B[1024 + m], c1), cmul(B[1280 + m], c2)));
        const V xmv = cadd(B[mneg], cadd(cmul(B[256 + mneg], cconj(c1)), cmul(B[512 + mneg], cconj(c2))));
        const double2 xp = make_double2(xpv.x, xpv.y), x1 = make_double2(x1v.x, x1v.y);
        double2 xm = make_double2(xmv.x, xmv.y);
        if (m == 0) xm = xp;
        double2 gt = make_double2(sc * (xp.x + xm.x), sc * (xp.y - xm.y));
        double2 gp = make_double2(sc * (xp.y + xm.y), -sc * (xp.x - xm.x));
        double2 gr = cscale(x1, 2.0 * sc);
        if (m == 0) {
            gr.y = 0.0;
            gt.y = 0.0;
            gp.y = 0.0;
        }
        const long o = ((long)k * L + m) * nlat + th;
        P.Fo[0][o] = gr;
        P.Fo[1][o] = gt;
        P.Fo[2][o] = gp;
    }
}

// One row per workgroup (the grid covers every row).
template <typename V>
__global__ __launch_bounds__(256, 4) void phi_cross768_kernel(PhiPtrs P, int lmax, int nlat, int nplanes,
                                                              const double2 *tw)
{
    __shared__ V T16[15 * 16];   // T16[(q-1) 16 + b] = e^{-2 pi i b q/256}, q = 1..15
    __shared__ V B[PHI_BUF];
    __shared__ double red[4][2];
    __shared__ double ia_s;
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const int k = t / nlat, th = t % nlat;
    if (k >= nplanes) return;
    const int tid = threadIdx.x;
    const int rtid = rotated_tid<2>(tid);   // wave 0 works all four passes, wave 3 none
    fill_t16(T16, tw, 3, tid);
    // e^{-2 pi i b/768} and e^{-2 pi i 2b/768} for b = tid: R3 takes their
    // conjugates, FIN (m = tid) the values themselves
    const double2 g1 = tw[tid], g2 = tw[2 * tid];
    double2 v[6];
    cross768_load(P, lmax, nlat, k, th, v);
    cross768_row<V>(P, lmax, nlat, k, th, v, T16, B, red, &ia_s, g1, g2, rtid);
}

// --------------------------------------- scalar phi transforms, nphi = 768
// Four theta rows of one plane per workgroup (the tiling of phi_inverse_kernel
// and phi_forward_kernel), packed in pairs (row 2p + i row 2p+1) into two
// complex 768-point transforms that follow the steps of phi_cross768_kernel:
// the radix-3 step on the loads (block 256..511 of the padded spectrum is
// zero) and 16 x 16 register passes.  Packed sequence p runs on the lanes of
// wave p, so only the phases that read across packed sequences need
// workgroup barriers.  Loads and stores cover four adjacent theta.
// F[k][m][theta] -> f[k][phi][theta]; blockIdx.y selects the field.
__global__ __launch_bounds__(256) void phi_inverse768_kernel(PhiPtrs P, int lmax, int nlat, int nplanes,
                                                             const double2 *tw)
{
    constexpr int N = 768;
    __shared__ double2 T16[15 * 16];
    __shared__ double2 B[6 * 256];
    const int ntt = (nlat + 3) / 4;
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const int k = t / ntt, th0 = (t % ntt) * 4;
    if (k >= nplanes) return;
    const int L = lmax + 1;
    const double2 *F = P.F[blockIdx.y];
    double *f = P.fo[blockIdx.y];
    const int tid = threadIdx.x;
    const int rtid = rotated_tid<2>(tid);
    fill_t16(T16, tw, 3, tid);
    const double2 g1 = tw[tid], g2 = tw[2 * tid];
    // coefficient m = tid of rows th0 .. th0+3 at B[row 256 + m]; loaded from
    // clamped addresses without branches, the zeros applied afterwards, so
    // that the four loads are in flight together
    {
        const int mc = min(tid, lmax);
        double2 v[4];
#pragma unroll
        for (int r = 0; r < 4; r++) v[r] = F[((long)k * L + mc) * nlat + min(th0 + r, nlat - 1)];
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const bool live = tid <= lmax && th0 + r < nlat;
            B[r * 256 + tid] = make_double2(live ? v[r].x : 0.0, (live && tid > 0) ? v[r].y : 0.0);
        }
    }
    __syncthreads();
    // ---- R3 -> Y[3p+r][b]
    {
        const int b = tid, mm = (256 - b) & 255;
        double2 zb[2], zc[2];
#pragma unroll
        for (int p = 0; p < 2; p++) {
            const double2 u = B[(2 * p) * 256 + b], w = B[(2 * p + 1) * 256 + b];
            const double2 uc = B[(2 * p) * 256 + mm], wc = B[(2 * p + 1) * 256 + mm];
            zb[p] = make_double2(u.x - w.y, u.y + w.x);
            zc[p] = b == 0 ? make_double2(0.0, 0.0) : make_double2(uc.x + wc.y, -uc.y + wc.x);
        }
        __syncthreads();
        const double h = 0.86602540378443864676;
        const double2 w3a = make_double2(-0.5, -h), w3b = make_double2(-0.5, h);
        const double2 t1 = cconj(g1), t2 = cconj(g2);
#pragma unroll
        for (int p = 0; p < 2; p++) {
            B[(3 * p) * 256 + b] = cadd(zb[p], zc[p]);
            B[(3 * p + 1) * 256 + b] = cmul(cadd(zb[p], cmul(zc[p], w3a)), t1);
            B[(3 * p + 2) * 256 + b] = cmul(cadd(zb[p], cmul(zc[p], w3b)), t2);
        }
    }
    __syncthreads();
    // ---- P1 (packed sequence x on lanes 0..47 of wave x)
    {
        const int x = rtid >> 6, r = (rtid & 63) >> 4, xr = 3 * x + r, b1 = rtid & 15;
        const bool act = rtid < 128 && r < 3;
        double2 v[16];
        if (act) {
#pragma unroll
            for (int b2 = 0; b2 < 16; b2++) v[b2] = B[xr * 256 + 16 * b2 + b1];
        }
        wave_lds_sync();
        if (act) {
            dft16(v, +1);
            const Tw16<double2> tw(T16, b1);
            B[(xr * 16 + b1) * 16 + b1] = v[0];
#pragma unroll
            for (int q1 = 1; q1 < 16; q1++) B[(xr * 16 + b1) * 16 + (q1 ^ b1)] = cmul(v[q1], cconj(tw.get(q1)));
        }
    }
    wave_lds_sync();
    // ---- P2: z_x at j = 3 (q1 + 16 q2) + r, stored at x N + j
    {
        int t1 = rtid;   // opaque copy: shared with P1 the indices stay live
        asm volatile("" : "+v"(t1));
        const int x = t1 >> 6, r = (t1 & 63) >> 4, xr = 3 * x + r, q1 = t1 & 15;
        const bool act = t1 < 128 && r < 3;
        double2 v[16];
        if (act) {
#pragma unroll
            for (int b1 = 0; b1 < 16; b1++) v[b1] = B[(xr * 16 + b1) * 16 + (q1 ^ b1)];
        }
        wave_lds_sync();
        if (act) {
            dft16(v, +1);
#pragma unroll
            for (int q2 = 0; q2 < 16; q2++) B[x * N + 3 * (q1 + 16 * q2) + r] = v[q2];
        }
    }
    __syncthreads();
    // ---- rows th0 + 2p (real part) and th0 + 2p + 1 (imaginary part)
#pragma unroll
    for (int i = 0; i < 3; i++) {
        const int j = tid + 256 * i;
        const double2 z0 = B[j], z1 = B[N + j];
        const long o = ((long)k * N + j) * nlat + th0;
        if (th0 < nlat) f[o] = z0.x;
        if (th0 + 1 < nlat) f[o + 1] = z0.y;
        if (th0 + 2 < nlat) f[o + 2] = z1.x;
        if (th0 + 3 < nlat) f[o + 3] = z1.y;
    }
}

// f[k][phi][theta] -> F[k][m][theta] = (2 pi / 768) sum_j f e^{-i m phi_j}, m <= lmax,
// four theta rows per workgroup as in phi_inverse768_kernel.
__global__ __launch_bounds__(256) void phi_forward768_kernel(PhiPtrs P, int lmax, int nlat, int nplanes,
                                                             const double2 *tw)
{
    constexpr int N = 768;
    __shared__ double2 T16[15 * 16];
    __shared__ double2 B[6 * 256];
    const int ntt = (nlat + 3) / 4;
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const int k = t / ntt, th0 = (t % ntt) * 4;
    if (k >= nplanes) return;
    const int L = lmax + 1;
    const double *f = P.f[blockIdx.y];
    double2 *F = P.Fo[blockIdx.y];
    const int tid = threadIdx.x;
    const int rtid = rotated_tid<2>(tid);
    fill_t16(T16, tw, 3, tid);
    const double2 c1 = tw[tid], c2 = tw[2 * tid];   // e^{-2 pi i m/768}, e^{-2 pi i 2m/768}, m = tid
    // packed rows z_p[j] = row th0+2p + i row th0+2p+1 at B[p N + j]
    // (clamped loads without branches, the zeros applied afterwards)
    {
        double v[3][4];
#pragma unroll
        for (int i = 0; i < 3; i++)
#pragma unroll
            for (int r = 0; r < 4; r++)
                v[i][r] = f[((long)k * N + tid + 256 * i) * nlat + min(th0 + r, nlat - 1)];
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const int j = tid + 256 * i;
            double w[4];
#pragma unroll
            for (int r = 0; r < 4; r++) w[r] = th0 + r < nlat ? v[i][r] : 0.0;
            B[j] = make_double2(w[0], w[1]);
            B[N + j] = make_double2(w[2], w[3]);
        }
    }
    __syncthreads();
    // ---- F1 (packed sequence s on lanes 0..47 of wave s)
    {
        const int s = rtid >> 6, rem = rtid & 63, n2 = rem >> 4, a1 = rem & 15;
        const bool act = rtid < 128 && rem < 48;
        double2 v[16];
        if (act) {
#pragma unroll
            for (int a2 = 0; a2 < 16; a2++) v[a2] = B[s * N + 3 * (16 * a2 + a1) + n2];
        }
        wave_lds_sync();
        if (act) {
            dft16(v, -1);
            const Tw16<double2> tw(T16, a1);
            B[((s * 3 + n2) * 16 + a1) * 16 + a1] = v[0];
#pragma unroll
            for (int k1 = 1; k1 < 16; k1++) B[((s * 3 + n2) * 16 + a1) * 16 + (k1 ^ a1)] = cmul(v[k1], tw.get(k1));
        }
    }
    wave_lds_sync();
    // ---- F2: Yf[3s+n2][k1 + 16 k2]
    {
        int t2 = rtid;   // opaque copy, as in phi_cross768_kernel
        asm volatile("" : "+v"(t2));
        const int s = t2 >> 6, rem = t2 & 63, n2 = rem >> 4, k1 = rem & 15;
        const bool act = t2 < 128 && rem < 48;
        double2 v[16];
        if (act) {
#pragma unroll
            for (int a1 = 0; a1 < 16; a1++) v[a1] = B[((s * 3 + n2) * 16 + a1) * 16 + (k1 ^ a1)];
        }
        wave_lds_sync();
        if (act) {
            dft16(v, -1);
#pragma unroll
            for (int k2 = 0; k2 < 16; k2++) B[(s * 3 + n2) * 256 + k1 + 16 * k2] = v[k2];
        }
    }
    __syncthreads();
    // ---- X[m] = sum_n2 e^{-2 pi i n2 m/768} Yf[n2][m], unpacked into the two rows
    const int m = tid;
    if (m < L) {
        const int mneg = (256 - m) & 255;
        const double sc = 0.5 * (2.0 * M_PI / N);
        const long base = ((long)k * L + m) * nlat + th0;
#pragma unroll
        for (int p = 0; p < 2; p++) {
            const double2 *Y = B + 3 * p * 256;
            const double2 zp = cadd(Y[m], cadd(cmul(Y[256 + m], c1), cmul(Y[512 + m], c2)));
            double2 zm = cadd(Y[mneg], cadd(cmul(Y[256 + mneg], cconj(c1)), cmul(Y[512 + mneg], cconj(c2))));
            if (m == 0) zm = zp;
            // X1 = (Z[m] + conj Z[-m]) / 2, X2 = (Z[m] - conj Z[-m]) / (2i)
            double2 x1 = make_double2(sc * (zp.x + zm.x), sc * (zp.y - zm.y));
            double2 x2 = make_double2(sc * (zp.y + zm.y), -sc * (zp.x - zm.x));
            if (m == 0) {
                x1.y = 0.0;
                x2.y = 0.0;
            }
            if (th0 + 2 * p < nlat) F[base + 2 * p] = x1;
            if (th0 + 2 * p + 1 < nlat) F[base + 2 * p + 1] = x2;
        }
    }
}

// ----------------------------------------- phi / cross and dot, nphi = 1536
// The same step for nphi = 6 x 16 x 16 with at most 512 stored m:
//   inverse: m = 256 a + b, phi index j = 6 q + r.  Blocks a = 2, 3 of the
//   padded spectrum are zero, so the radix-6 step takes four terms, fused
//   with the loads:
//     Y_r[b] = (Z_0 + w6^r Z_1 + w6^{4r} Z_4 + w6^{5r} Z_5) e^{2 pi i b r/1536},
//   Z_a = Z[256 a + b]; then z[6 q + r] = IDFT_256(Y_r)[q] as 16 x 16.
//   forward: j = 6 n1 + n2, six 256-point DFTs per packed sequence, and only
//   the outputs m <= lmax (and their mirrors) are formed from them.
// One row per workgroup of 512 threads, in place in one 72 KB LDS buffer,
// two rows per CU.  DOT: the forward part transforms the one real row u . g
// (u + i a g packed as in the cross product) instead of the packed pair
// (N_theta + i N_phi, N_r).  A sequence's 16 lanes lie in one wave, so the
// exchanges inside one 256-point sequence need only wave-local ordering.
// V = float2: the row is transformed in FP32 (36 KB of LDS, three rows per CU)
// for the BF16x3 form, whose Legendre error dominates what the FP32
// transforms add (u x omega 5.8e-6 -> 6.2e-6 at lmax 511); loads, the pack
// scale and the stores stay FP64.
// The FP32 form keeps FP64 rows: FP32 transforms raised its u x omega error at
// lmax 511 from 4e-7 to 2e-6.
constexpr int PHI6_BUF = 18 * 256;

// p[r-1] = w^r, r = 1..5, in FP64 (a few ulps from the correctly rounded roots)
__device__ __forceinline__ void pow5(double2 w, double2 (&p)[5])
{
    p[0] = w;
    p[1] = cmul(w, w);
    p[2] = cmul(p[1], w);
    p[3] = cmul(p[1], p[1]);
    p[4] = cmul(p[3], w);
}

// e^{2 pi i k/6} for a compile-time k (any integer)
template <typename V>
__device__ __forceinline__ V root6(int k)
{
    using S = typename Cx<V>::S;
    const S h = S(0.86602540378443864676);
    switch (((k % 6) + 6) % 6) {
    case 0: return V{S(1), S(0)};
    case 1: return V{S(0.5), h};
    case 2: return V{S(-0.5), h};
    case 3: return V{S(-1), S(0)};
    case 4: return V{S(-0.5), -h};
    default: return V{S(0.5), -h};
    }
}

template <typename V, bool DOT>
__global__ __launch_bounds__(512, 2) void phi1536_kernel(PhiPtrs P, int lmax, int nlat, int nplanes,
                                                         const double2 *tw)
{
    using S = typename Cx<V>::S;
    constexpr int N = 1536, NS = DOT ? 1 : 2;
    __shared__ V T16[15 * 16];
    __shared__ V B[PHI6_BUF];
    __shared__ double red[8][2];
    __shared__ double ia_s;
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    const int k = t / nlat, th = t % nlat;
    if (k >= nplanes) return;
    const int L = lmax + 1;
    const int tid = threadIdx.x;
    const int rtid = rotated_tid<3>(tid);
    fill_t16(T16, tw, 6, tid);
    const double2 g1 = tw[tid & 255];   // e^{-2 pi i b/1536}, b = tid mod 256
    // the row's six coefficient vectors at B[f 512 + m], m = 0..511 (zero
    // above lmax), and the sums of squares of u and of the second factor
    // (clamped loads without branches, the zeros applied afterwards, as in
    // phi_cross768_kernel)
    double s[2] = {0.0, 0.0};
    {
        const int mc = min(tid, lmax);
        const bool live = tid <= lmax;
        double2 vin[6];
#pragma unroll
        for (int f = 0; f < 6; f++) vin[f] = P.F[f][((long)k * L + mc) * nlat + th];
#pragma unroll
        for (int f = 0; f < 6; f++) {
            const double2 v = make_double2(live ? vin[f].x : 0.0, (live && tid > 0) ? vin[f].y : 0.0);
            B[f * 512 + tid] = Cx<V>::make(v.x, v.y);
            s[f / 3] += v.x * v.x + v.y * v.y;
        }
    }
    block_sums<2, 8>(s, red);
    const double a = wave_uniform(pack_scale(s[0], s[1]));
    const S as = (S)a;   // a power of two: exact in FP32 as well
    if (tid == 0) ia_s = 1.0 / a;

    // ---- R6 -> Y[6x+r][b] for the items (x, b) = tid and tid + 512
    {
        V z[2][4];
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int i = tid + 512 * h, x = i >> 8, b = i & 255;
            if (i < 768) {
                // packs U + i a W at m = b, 256 + b and at the mirrors of
                // 512 - b (zero for b = 0: m = 512 is above every lmax here)
                // and 256 - b
                const int ms[4] = {b, 256 + b, 512 - b, 256 - b};
#pragma unroll
                for (int q = 0; q < 4; q++) {
                    V zz = {S(0), S(0)};
                    if (q != 2 || b != 0) {
                        const V u = B[x * 512 + ms[q]], w = cscale(B[(3 + x) * 512 + ms[q]], as);
                        zz = q < 2 ? V{u.x - w.y, u.y + w.x} : V{u.x + w.y, -u.y + w.x};
                    }
                    z[h][q] = zz;
                }
            }
        }
        // e^{-2 pi i b r/1536}, r = 1..5, as FP64 powers of the table value
        // (b = tid mod 256 for both items)
        double2 g[5];
        pow5(g1, g);
        __syncthreads();
#pragma unroll
        for (int h = 0; h < 2; h++) {
            const int i = tid + 512 * h, x = i >> 8, b = i & 255;
            if (i < 768) {
#pragma unroll
                for (int r = 0; r < 6; r++) {
                    // e^{2 pi i 256 r/1536}, e^{2 pi i 1024 r/1536}, e^{2 pi i 1280 r/1536}: sixth roots
                    const V w1 = root6<V>(r), w4 = root6<V>(4 * r), w5 = root6<V>(5 * r);
                    const V y = cadd(cadd(z[h][0], cmul(z[h][1], w1)), cadd(cmul(z[h][2], w4), cmul(z[h][3], w5)));
                    B[(6 * x + r) * 256 + b] = r == 0 ? y : cmul(y, Cx<V>::make(g[r - 1].x, -g[r - 1].y));
                }
            }
        }
    }
    __syncthreads();
    // ---- P1: T[xr][b1][q1] = e^{2 pi i b1 q1/256} sum_b2 Y[xr][16 b2 + b1] e^{2 pi i b2 q1/16}
    {
        const int xr = rtid >> 4, b1 = rtid & 15;
        const bool act = rtid < 288;
        V v[16];
        if (act) {
#pragma unroll
            for (int b2 = 0; b2 < 16; b2++) v[b2] = B[xr * 256 + 16 * b2 + b1];
        }
        wave_lds_sync();
        if (act) {
            dft16(v, +1);
            const Tw16<V> tw(T16, b1);
            B[(xr * 16 + b1) * 16 + b1] = v[0];
#pragma unroll
            for (int q1 = 1; q1 < 16; q1++) B[(xr * 16 + b1) * 16 + (q1 ^ b1)] = cmul(v[q1], cconj(tw.get(q1)));
        }
    }
    wave_lds_sync();
    // ---- P2: grid values z_x at j = 6 (q1 + 16 q2) + r, stored at x N + j
    {
        int t1 = rtid;   // opaque copy: shared with P1 the indices stay live and spill
        asm volatile("" : "+v"(t1));
        const int xr = t1 >> 4, q1 = t1 & 15, x = xr / 6, r = xr - 6 * x;
        const bool act = t1 < 288;
        V v[16];
        if (act) {
#pragma unroll
            for (int b1 = 0; b1 < 16; b1++) v[b1] = B[(xr * 16 + b1) * 16 + (q1 ^ b1)];
        }
        __syncthreads();
        if (act) {
            dft16(v, +1);
#pragma unroll
            for (int q2 = 0; q2 < 16; q2++) B[x * N + 6 * (q1 + 16 * q2) + r] = v[q2];
        }
    }
    __syncthreads();
    // ---- CR: each thread overwrites only the slots it reads
    {
        const S ia = (S)ia_s;
#pragma unroll
        for (int i = 0; i < 3; i++) {
            const int j = tid + 512 * i;
            const V zr = B[j], zt = B[N + j], zp = B[2 * N + j];
            const S wr = zr.y * ia, wt = zt.y * ia, wp = zp.y * ia;
            if (DOT) {
                B[j] = V{zr.x * wr + zt.x * wt + zp.x * wp, S(0)};
            } else {
                B[j] = V{zp.x * wr - zr.x * wp, zr.x * wt - zt.x * wr};
                B[N + j] = V{zt.x * wp - zp.x * wt, S(0)};
            }
        }
    }
    __syncthreads();
    // ---- F1: F[6s+n2][a1][k1] = e^{-2 pi i a1 k1/256} sum_a2 P_s[6 (16 a2 + a1) + n2] e^{-2 pi i a2 k1/16}
    {
        const int s = rtid / 96, rem = rtid % 96, n2 = rem >> 4, a1 = rem & 15;
        const bool act = rtid < 96 * NS;
        V v[16];
        if (act) {
#pragma unroll
            for (int a2 = 0; a2 < 16; a2++) v[a2] = B[s * N + 6 * (16 * a2 + a1) + n2];
        }
        __syncthreads();
        if (act) {
            dft16(v, -1);
            const Tw16<V> tw(T16, a1);
            B[((s * 6 + n2) * 16 + a1) * 16 + a1] = v[0];
#pragma unroll
            for (int k1 = 1; k1 < 16; k1++) B[((s * 6 + n2) * 16 + a1) * 16 + (k1 ^ a1)] = cmul(v[k1], tw.get(k1));
        }
    }
    wave_lds_sync();
    // ---- F2: Yf[6s+n2][k1 + 16 k2]
    {
        int t2 = rtid;
        asm volatile("" : "+v"(t2));
        const int s = t2 / 96, rem = t2 % 96, n2 = rem >> 4, k1 = rem & 15;
        const bool act = t2 < 96 * NS;
        V v[16];
        if (act) {
#pragma unroll
            for (int a1 = 0; a1 < 16; a1++) v[a1] = B[((s * 6 + n2) * 16 + a1) * 16 + (k1 ^ a1)];
        }
        wave_lds_sync();
        if (act) {
            dft16(v, -1);
#pragma unroll
            for (int k2 = 0; k2 < 16; k2++) B[(s * 6 + n2) * 256 + k1 + 16 * k2] = v[k2];
        }
    }
    // FIN's twiddles e^{-2 pi i n2 m/1536} (m = tid): FP64 powers of the
    // table value, loaded before the barrier
    const double2 f1 = tw[tid];
    __syncthreads();
    // ---- FIN: X[m] = sum_n2 e^{-2 pi i n2 m/1536} Yf[n2][m mod 256]
    const int m = tid;
    if (m < L) {
        const int mm = m & 255, mneg = (256 - mm) & 255;
        double2 cg[5];
        pow5(f1, cg);
        V c[6];
        c[0] = V{S(1), S(0)};
#pragma unroll
        for (int n2 = 1; n2 < 6; n2++) c[n2] = Cx<V>::make(cg[n2 - 1].x, cg[n2 - 1].y);
        V xpv = {S(0), S(0)};
#pragma unroll
        for (int n2 = 0; n2 < 6; n2++) xpv = cadd(xpv, cmul(B[n2 * 256 + mm], c[n2]));
        const double2 xp = make_double2(xpv.x, xpv.y);
        const long o = ((long)k * L + m) * nlat + th;
        if (DOT) {
            double2 x = cscale(xp, 2.0 * M_PI / N);
            if (m == 0) x.y = 0.0;
            P.Fo[0][o] = x;
        } else {
            const double sc = 0.5 * (2.0 * M_PI / N);
            V x1v = {S(0), S(0)}, xmv = {S(0), S(0)};
#pragma unroll
            for (int n2 = 0; n2 < 6; n2++) {
                x1v = cadd(x1v, cmul(B[(6 + n2) * 256 + mm], c[n2]));
                xmv = cadd(xmv, cmul(B[n2 * 256 + mneg], cconj(c[n2])));
            }
            const double2 x1 = make_double2(x1v.x, x1v.y);
            double2 xm = make_double2(xmv.x, xmv.y);
            if (m == 0) xm = xp;
            double2 gt = make_double2(sc * (xp.x + xm.x), sc * (xp.y - xm.y));
            double2 gp = make_double2(sc * (xp.y + xm.y), -sc * (xp.x - xm.x));
            double2 gr = cscale(x1, 2.0 * sc);
            if (m == 0) {
                gr.y = 0.0;
                gt.y = 0.0;
                gp.y = 0.0;
            }
            P.Fo[0][o] = gr;
            P.Fo[1][o] = gt;
            P.Fo[2][o] = gp;
        }
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
    if (p->nphi == 768 && p->lmax <= 255) {
        hipLaunchKernelGGL(phi_inverse768_kernel, dim3(nplanes * ntt, nfields), dim3(256), 0, p->stream, P,
                           p->lmax, p->nlat, nplanes, p->tw_phi);
        RD_LAUNCH_CHECK();
        return RD_OK;
    }
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
    if (p->nphi == 768 && p->lmax <= 255) {
        hipLaunchKernelGGL(phi_forward768_kernel, dim3(nplanes * ntt, nfields), dim3(256), 0, p->stream, P,
                           p->lmax, p->nlat, nplanes, p->tw_phi);
        RD_LAUNCH_CHECK();
        return RD_OK;
    }
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

// Fin: (u_r, u_t, u_p, g_r, g_t, g_p) Fourier fields -> Fout = u . g.
int launch_phi_dot(Plan *p, int nplanes, const double2 *const *Fin, double2 *Fout, int prec)
{
    PhiPtrs P = {};
    for (int i = 0; i < 6; i++) P.F[i] = Fin[i];
    P.Fo[0] = Fout;
    if (p->nphi == 1536 && p->lmax <= 511) {
        if (prec == RD_PREC_BF16X3)
            hipLaunchKernelGGL((phi1536_kernel<float2, true>), dim3(nplanes * p->nlat), dim3(512), 0, p->stream, P,
                               p->lmax, p->nlat, nplanes, p->tw_phi);
        else
            hipLaunchKernelGGL((phi1536_kernel<double2, true>), dim3(nplanes * p->nlat), dim3(512), 0, p->stream,
                               P, p->lmax, p->nlat, nplanes, p->tw_phi);
        RD_LAUNCH_CHECK();
        return RD_OK;
    }
    size_t shm = sizeof(double2) * (6 * p->nphi + p->nphi / 2);
    {
        int lrc = allow_lds(phi_dot_kernel, shm);
        if (lrc) return lrc;
    }
    hipLaunchKernelGGL(phi_dot_kernel, dim3(nplanes * p->nlat), dim3(PHI_GEN_THREADS), shm, p->stream, P, p->lmax,
                       p->nlat, p->nphi, nplanes, p->seq_phi, p->tw_phi);
    RD_LAUNCH_CHECK();
    return RD_OK;
}

int launch_phi_cross(Plan *p, int nplanes, const double2 *const *Fin, double2 *const *Fout, int prec)
{
    PhiPtrs P = {};
    for (int i = 0; i < 6; i++) P.F[i] = Fin[i];
    for (int i = 0; i < 3; i++) P.Fo[i] = Fout[i];
    if (p->nphi == 768 && p->lmax <= 255) {
        // every form transforms these rows in FP64: the FP32 instance of this
        // kernel (BF16x3 form) took 0.745 against 0.661 ms at lmax 255, 128
        // planes, and lost accuracy besides
        hipLaunchKernelGGL(phi_cross768_kernel<double2>, dim3(nplanes * p->nlat), dim3(256), 0, p->stream, P,
                           p->lmax, p->nlat, nplanes, p->tw_phi);
        RD_LAUNCH_CHECK();
        return RD_OK;
    }
    if (p->nphi == 1536 && p->lmax <= 511) {
        if (prec == RD_PREC_BF16X3)
            hipLaunchKernelGGL((phi1536_kernel<float2, false>), dim3(nplanes * p->nlat), dim3(512), 0, p->stream, P,
                               p->lmax, p->nlat, nplanes, p->tw_phi);
        else
            hipLaunchKernelGGL((phi1536_kernel<double2, false>), dim3(nplanes * p->nlat), dim3(512), 0, p->stream,
                               P, p->lmax, p->nlat, nplanes, p->tw_phi);
        RD_LAUNCH_CHECK();
        return RD_OK;
    }
    size_t shm = sizeof(double2) * (6 * p->nphi + p->nphi / 2);
    {
        int lrc = allow_lds(phi_cross_kernel, shm);
        if (lrc) return lrc;
    }
    hipLaunchKernelGGL(phi_cross_kernel, dim3(nplanes * p->nlat), dim3(PHI_GEN_THREADS), shm, p->stream, P,
                       p->lmax, p->nlat, p->nphi, nplanes, p->seq_phi, p->tw_phi);
    RD_LAUNCH_CHECK();
    return RD_OK;
}

// ----------------------------------------------------------------- z
// Radial FFT of nb sequences of length Z.n held at a + s n (b is scratch of
// the same size, xs kZScalars double2 of LDS scratch); returns the buffer
// holding the result, in the same layout.  The Rader form gathers
// a_p = x[g^p] into packed sequences of length n - 1, convolves them with c by
// two FFTs and the spectrum cf[s], and scatters X[g^-q] = x_0 + conv[q],
// X[0] = x_0 + sum_p a_p (the first forward coefficient).
template <int RB>
__device__ double2 *z_fft(double2 *a, double2 *b, int nb, const ZFft &Z, int sign, const double2 *tw, int nh,
                          double2 *xs)
{
    if (!Z.rader) return lds_fft<RB>(a, b, nb, Z.seq, sign, tw, nh);
    const int n = Z.n, n1 = n - 1;
    for (int e = threadIdx.x; e < nb * n1; e += blockDim.x) {
        const int q = e / n1, p = e - q * n1;
        b[e] = a[q * n + Z.pin[p]];
    }
    for (int q = threadIdx.x; q < nb; q += blockDim.x) xs[q] = a[q * n];
    __syncthreads();
    double2 *F = lds_fft<RB>(b, a, nb, Z.seq, -1, tw, nh);
    const double2 *cf = Z.cf[sign > 0 ? 1 : 0];
    for (int e = threadIdx.x; e < nb * n1; e += blockDim.x) {
        const int q = e / n1, k = e - q * n1;
        const double2 v = F[e];
        if (k == 0) xs[nb + q] = v;
        F[e] = cmul(v, cf[k]);
    }
    __syncthreads();
    double2 *G = lds_fft<RB>(F, F == a ? b : a, nb, Z.seq, +1, tw, nh);
    double2 *X = G == a ? b : a;
    for (int e = threadIdx.x; e < nb * n1; e += blockDim.x) {
        const int q = e / n1, r = e - q * n1;
        X[q * n + Z.pout[r]] = cadd(xs[q], G[e]);
    }
    for (int q = threadIdx.x; q < nb; q += blockDim.x) X[q * n] = cadd(xs[q], xs[nb + q]);
    __syncthreads();
    return X;
}

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
__global__ __launch_bounds__(512) void z_kernel(ZArgs A, int mode, ZFft Z, const double2 *tw)
{
    extern __shared__ double2 sm[];
    const int n = A.n, c0 = blockIdx.x * ZTC;
    double2 *X = sm, *Y = sm + ZTC * n, *twl = sm + 2 * ZTC * n, *xs = sm + (2 * ZTC + 1) * n;
    const int nh = load_twiddles(twl, tw, Z.seq.n);
    const double2 *src = (mode == 2) ? A.ain : A.a;
    const int nr = n - A.C;
    if (mode == 1) {
        z_load(X, src, A.ld, c0, A.ncols, n, n);
        __syncthreads();
        double2 *R = z_fft<RB>(X, Y, ZTC, Z, +1, twl, nh, xs);
        z_store(R, A.a, A.ld, c0, A.ncols, n, n, 1.0);
        return;
    }
    z_load(X, src, A.ld, c0, A.ncols, n, A.C > 0 ? nr : n);
    __syncthreads();
    if (A.C > 0) {
        fc_continue(X, ZTC, n, A.C, A.d, A.dir);
        __syncthreads();
    }
    double2 *R = z_fft<RB>(X, Y, ZTC, Z, -1, twl, nh, xs);
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
    double2 *R2 = z_fft<RB>(R, O, ZTC, Z, +1, twl, nh, xs);
    z_store(R2, A.a, A.ld, c0, A.ncols, n, n, 1.0);
}

int z_fact_launch(Plan *p, const ZArgs &A, int mode);

static int z_launch(Plan *p, int mode, double2 *a, const double2 *ain, int order, int ncols, long ld)
{
    ZArgs A;
    A.a = a;
    A.ain = ain;
    A.dir = p->d_dir;
    A.kz = p->d_kz;
    A.ld = ld;
    A.ncols = ncols;
    A.n = p->nz;
    A.C = p->C;
    A.d = p->d;
    A.order = order;
    {
        const int frc = z_fact_launch(p, A, mode);
        if (frc != 1) return frc;
    }
    size_t shm = sizeof(double2) * ((2 * ZTC + 1) * p->nz + kZScalars);
    const dim3 grid((ncols + ZTC - 1) / ZTC), block(512);
#define RD_Z_CASE(RBV)                                                                         \
    case RBV: {                                                                                \
        int lrc = allow_lds(z_kernel<RBV>, shm);                                               \
        if (lrc) return lrc;                                                                   \
        hipLaunchKernelGGL(z_kernel<RBV>, grid, block, shm, p->stream, A, mode, p->zfft, p->tw_z); \
        break;                                                                                 \
    }
    switch (z_radix(p->zfft)) {
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

int launch_z_forward(Plan *p, double2 *a) { return z_launch(p, 0, a, nullptr, 0, p->nlm, p->nlm); }
int launch_z_inverse(Plan *p, double2 *a) { return z_launch(p, 1, a, nullptr, 0, p->nlm, p->nlm); }
int launch_derivr_r(Plan *p, int order, const double2 *a, double2 *b)
{
    return z_launch(p, 2, b, a, order, p->nlm, p->nlm);
}
int launch_derivr_cols(Plan *p, int order, const double2 *a, double2 *b, int ncols, long ld)
{
    return z_launch(p, 2, b, a, order, ncols, ld);
}

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
    int nout;   // 6: u and omega; 3: u only (Q, S, T of W, Z; no curl)
    // arrays whose rows nr..n-1 are cleared (the outputs of the nonlinear
    // term), or null; the factorised kernel's W workgroups clear their columns
    double2 *zero[3];
};

// lm columns per workgroup in the fused radial step.  One column keeps the
// workgroup's LDS at 7 n double2 (31 KB at n = 281), so five workgroups fit
// on a CU; four columns (112 KB) allowed one, and the lmax = 511 coupled step
// took 92 ms instead of 74.
constexpr int UTC = 1;

// Columns c of a batch held in LDS at X + c*n: rows nr..n-1 from the FC-Gram
// continuation of rows 0..d-1 and nr-d..nr-1.
__device__ void z_load_cols(double2 *X, const double2 *a, long ld, int c0, int ncols, int n, int rows,
                            int nc)
{
    for (int e = threadIdx.x; e < rows * nc; e += blockDim.x) {
        const int r = e / nc, c = e - r * nc;
        X[c * n + r] = (c0 + c < ncols) ? a[(long)r * ld + c0 + c] : make_double2(0.0, 0.0);
    }
}

// Both forward transforms (W and Z columns) run as one batch, the three
// derivative spectra (i kz W^, -kz^2 W^, i kz Z^) as one inverse batch.
template <int RB>
__global__ __launch_bounds__(256) void uxw_radial_kernel(UxwRadArgs A, ZFft Z, const double2 *tw)
{
    extern __shared__ double2 sm[];
    const int n = A.n, nr = n - A.C;
    const int c0 = xcd_tile(blockIdx.x, gridDim.x) * UTC;
    double2 *P = sm, *Q = sm + 3 * UTC * n, *twl = sm + 6 * UTC * n, *xs = sm + (6 * UTC + 1) * n;
    const int nh = load_twiddles(twl, tw, Z.seq.n);
    const int rows = A.C > 0 ? nr : n;
    // the spectrum of Z only feeds the curl outputs (nout = 6)
    const int nfwd = A.nout == 6 ? 2 : 1;
    z_load_cols(P, A.W, A.ld, c0, A.ncols, n, rows, UTC);
    if (nfwd == 2) z_load_cols(P + UTC * n, A.Z, A.ld, c0, A.ncols, n, rows, UTC);
    __syncthreads();
    if (A.C > 0) {
        fc_continue(P, nfwd * UTC, n, A.C, A.d, A.dir);
        __syncthreads();
    }
    double2 *R = z_fft<RB>(P, Q, nfwd * UTC, Z, -1, twl, nh, xs);
    double2 *I = (R == P) ? Q : P;
    // spectra / n times (i kz)^p: sequences [0,UTC) dW, [UTC,2UTC) d2W, [2UTC,3UTC) dZ
    // (only dW when nout = 3)
    const double inv_n = 1.0 / n;
    const int ninv = A.nout == 6 ? 3 : 1;
    for (int e = threadIdx.x; e < ninv * UTC * n; e += blockDim.x) {
        const int sq = e / n, r = e - sq * n;
        const int kind = sq / UTC, c = sq - kind * UTC;
        const double k = A.kz[r];
        double2 v = R[(kind == 2 ? UTC + c : c) * n + r];
        if (kind == 1) {
            v = cscale(v, -k * k * inv_n);
        } else {
            v = cscale(v, k * inv_n);
            v = make_double2(-v.y, v.x);
        }
        I[e] = v;
    }
    __syncthreads();
    double2 *D = z_fft<RB>(I, R, ninv * UTC, Z, +1, twl, nh, xs);
    // u:     Q = l(l+1) W / r^2, S = dW / r, T = Z / r
    // omega: Q = l(l+1) Z / r^2, S = dZ / r, T = -(d2W - l(l+1) W / r^2) / r
    for (int e = threadIdx.x; e < nr * UTC; e += blockDim.x) {
        const int r = e / UTC, c = e - r * UTC;
        const int lm = c0 + c;
        if (lm >= A.ncols) continue;
        const double rr = A.z[r], r1 = 1.0 / rr, r2 = 1.0 / (rr * rr);
        const double ll = (double)A.ell[lm] * (A.ell[lm] + 1);
        const long o = (long)r * A.ld + lm;
        const double2 w = A.W[o], zz = A.Z[o];
        const double2 dw = D[c * n + r];
        A.out[0][o] = cscale(w, ll * r2);
        A.out[1][o] = cscale(dw, r1);
        A.out[2][o] = cscale(zz, r1);
        if (ninv == 3) {
            const double2 d2w = D[(UTC + c) * n + r], dz = D[(2 * UTC + c) * n + r];
            A.out[3][o] = cscale(zz, ll * r2);
            A.out[4][o] = cscale(dz, r1);
            A.out[5][o] = cscale(make_double2(d2w.x - ll * w.x * r2, d2w.y - ll * w.y * r2), -r1);
        }
    }
}

// ------------------------------------------- factorised radial step, n = N1 N2
// cos and sin of 2 pi m / R, correctly rounded.
template <int R> struct DftConst;
template <> struct DftConst<7> {
    static constexpr double c[7] = {1.0, 0.6234898018587335, -0.2225209339563144, -0.9009688679024191,
                                    -0.9009688679024191, -0.2225209339563144, 0.6234898018587335};
    static constexpr double s[7] = {0.0, 0.7818314824680298, 0.9749279121818236, 0.4338837391175581,
                                    -0.4338837391175581, -0.9749279121818236, -0.7818314824680298};
};
template <> struct DftConst<9> {
    static constexpr double c[9] = {1.0, 0.766044443118978, 0.17364817766693036, -0.5, -0.9396926207859084,
                                    -0.9396926207859084, -0.5, 0.17364817766693036, 0.766044443118978};
    static constexpr double s[9] = {0.0, 0.6427876096865394, 0.984807753012208, 0.8660254037844386,
                                    0.3420201433256687, -0.3420201433256687, -0.8660254037844386,
                                    -0.984807753012208, -0.6427876096865394};
};
template <> struct DftConst<11> {
    static constexpr double c[11] = {1.0, 0.8412535328311812, 0.41541501300188644, -0.14231483827328514,
                                     -0.6548607339452851, -0.9594929736144974, -0.9594929736144974,
                                     -0.6548607339452851, -0.14231483827328514, 0.41541501300188644,
                                     0.8412535328311812};
    static constexpr double s[11] = {0.0, 0.5406408174555976, 0.9096319953545183, 0.9898214418809327,
                                     0.7557495743542583, 0.28173255684142967, -0.28173255684142967,
                                     -0.7557495743542583, -0.9898214418809327, -0.9096319953545183,
                                     -0.5406408174555976};
};
template <> struct DftConst<17> {
    static constexpr double c[17] = {1.0, 0.9324722294043558, 0.7390089172206591, 0.44573835577653825,
                                     0.092268359463302, -0.2736629900720829, -0.6026346363792564,
                                     -0.8502171357296141, -0.9829730996839018, -0.9829730996839018,
                                     -0.8502171357296141, -0.6026346363792564, -0.2736629900720829,
                                     0.092268359463302, 0.44573835577653825, 0.7390089172206591,
                                     0.9324722294043558};
    static constexpr double s[17] = {0.0, 0.3612416661871529, 0.6736956436465572, 0.8951632913550623,
                                     0.9957341762950345, 0.961825643172819, 0.7980172272802395,
                                     0.5264321628773558, 0.18374951781657034, -0.18374951781657034,
                                     -0.5264321628773558, -0.7980172272802395, -0.961825643172819,
                                     -0.9957341762950345, -0.8951632913550623, -0.6736956436465572,
                                     -0.3612416661871529};
};

template <> struct DftConst<5> {
    static constexpr double c[5] = {1.0, 0.30901699437494745, -0.8090169943749475, -0.8090169943749475,
                                    0.30901699437494745};
    static constexpr double s[5] = {0.0, 0.9510565162951535, 0.5877852522924731, -0.5877852522924731,
                                    -0.9510565162951535};
};

// In-register DFT of odd length R, X[k] = sum_n x[n] e^{sign 2 pi i n k / R},
// pairing n with R - n (a_j = x_j + x_{R-j}, b_j = x_j - x_{R-j}):
//   X[k]   = x_0 + sum_j a_j cos(2 pi j k / R) + sign i sum_j b_j sin(2 pi j k / R)
//   X[R-k] = the same with -sign.
template <int R>
__device__ __forceinline__ void dft_odd(double2 (&x)[R], int sign)
{
    constexpr int H = (R - 1) / 2;
    double2 a[H], b[H];
    double2 s0 = x[0];
#pragma unroll
    for (int j = 1; j <= H; j++) {
        a[j - 1] = cadd(x[j], x[R - j]);
        b[j - 1] = csub(x[j], x[R - j]);
        s0 = cadd(s0, a[j - 1]);
    }
    const double sg = sign > 0 ? 1.0 : -1.0;
    double2 y[R];
    y[0] = s0;
#pragma unroll
    for (int k = 1; k <= H; k++) {
        double2 C = x[0], S = make_double2(0.0, 0.0);
#pragma unroll
        for (int j = 1; j <= H; j++) {
            const int m = (j * k) % R;
            C.x += a[j - 1].x * DftConst<R>::c[m];
            C.y += a[j - 1].y * DftConst<R>::c[m];
            S.x += b[j - 1].x * DftConst<R>::s[m];
            S.y += b[j - 1].y * DftConst<R>::s[m];
        }
        y[k] = make_double2(C.x - sg * S.y, C.y + sg * S.x);
        y[R - k] = make_double2(C.x + sg * S.y, C.y - sg * S.x);
    }
#pragma unroll
    for (int k = 0; k < R; k++) x[k] = y[k];
}

// The same DFT with each output handed to emit(k, X[k]) as soon as it is
// formed (k = 0, then the pairs k, R - k), so that only the sums a_j and
// differences b_j stay live, not an output array beside them: the form of
// the passes whose outputs go straight to LDS or to global memory.
template <int R, typename E>
__device__ __forceinline__ void dft_odd_emit(const double2 (&x)[R], int sign, E &&emit)
{
    constexpr int H = (R - 1) / 2;
    double2 a[H], b[H];
    const double2 x0 = x[0];
    double2 s0 = x0;
#pragma unroll
    for (int j = 1; j <= H; j++) {
        a[j - 1] = cadd(x[j], x[R - j]);
        b[j - 1] = csub(x[j], x[R - j]);
        s0 = cadd(s0, a[j - 1]);
    }
    const double sg = sign > 0 ? 1.0 : -1.0;
    emit(0, s0);
#pragma unroll
    for (int k = 1; k <= H; k++) {
        double2 C = x0, S = make_double2(0.0, 0.0);
#pragma unroll
        for (int j = 1; j <= H; j++) {
            const int m = (j * k) % R;
            C.x += a[j - 1].x * DftConst<R>::c[m];
            C.y += a[j - 1].y * DftConst<R>::c[m];
            S.x += b[j - 1].x * DftConst<R>::s[m];
            S.y += b[j - 1].y * DftConst<R>::s[m];
        }
        emit(k, make_double2(C.x - sg * S.y, C.y + sg * S.x));
        emit(R - k, make_double2(C.x + sg * S.y, C.y - sg * S.x));
    }
}

// The radial step of the nonlinear term for n = N1 N2 (n = 153 = 17 x 9 at
// nr = 128, 121 = 11 x 11, 49 = 7 x 7), CB lm columns per workgroup.  With
// r = n2 + N2 n1 and k = k1 + N1 k2:
//   X[k] = sum_n2 e^{-2 pi i n2 k2/N2} e^{-2 pi i n2 k1/n} sum_n1 x[r] e^{-2 pi i n1 k1/N1}
// Pass A (thread = column, n2) holds the N1 rows n2 + N2 n1 of a column: it
// loads them straight from HBM (coalesced over the columns), applies the
// FC-Gram continuation for the rows past nr in registers and does the N1-point
// DFTs; pass B (thread = column, k1) does the N2-point DFTs and keeps the W and
// Z spectra in registers, from which it starts the three inverse transforms
// (i kz W^, -kz^2 W^, i kz Z^, each / n).  One LDS buffer of CB n values
// carries every transpose.
template <int N1, int N2>
struct RadFact {
    static constexpr int N = N1 * N2;
    static constexpr int CB = 256 / (N1 > N2 ? N1 : N2);
};

// Pass-A rows n2 + N2 n1 of column col of src, for the factorised radial
// kernels: every row (clamped to the physical ones) and the 2 D wall rows are
// loaded first, without branches, then the continuation rows past nr are
// formed from the walls with the FC-Gram operator dirl (sphp_sca.F90:369-381);
// ok = false gives zeros.
template <int N1, int N2, int D>
__device__ __forceinline__ void fact_load_rows(const double2 *src, long ld, int col, bool ok, int n2, int nr,
                                               int C, const double *dirl, double2 (&x)[N1])
{
    double2 wr[D > 0 ? D : 1], wl[D > 0 ? D : 1];
#pragma unroll
    for (int n1 = 0; n1 < N1; n1++) x[n1] = src[(long)min(n2 + N2 * n1, nr - 1) * ld + col];
#pragma unroll
    for (int jj = 0; jj < D; jj++) {
        wr[jj] = src[(long)(nr - D + jj) * ld + col];   // row nr - d + jj - 1, jj 1-based
        wl[jj] = src[(long)(D - 1 - jj) * ld + col];    // row d - jj
    }
#pragma unroll
    for (int n1 = 0; n1 < N1; n1++) {
        const int r = n2 + N2 * n1;
        double2 v = x[n1];
        if (D > 0 && r >= nr) {
            const int ii = r - nr + 1;
            v = make_double2(0.0, 0.0);
#pragma unroll
            for (int jj = 0; jj < D; jj++) {
                const double a1 = dirl[(ii - 1) + C * jj];
                const double a2 = dirl[(C - ii) + C * jj];
                v.x += a1 * wr[jj].x + a2 * wl[jj].x;
                v.y += a1 * wr[jj].y + a2 * wl[jj].y;
            }
        }
        x[n1] = ok ? v : make_double2(0.0, 0.0);
    }
}

// D: the FC-Gram matching points (the continuation rows are formed from
// D rows at each wall, D a compile-time count so that every load of a column
// is issued at once, without branches); D = 0 for a periodic radius (no
// continuation).  z, kz and dir are staged in LDS.
constexpr int kRadMaxC = 32;

// Tiling of uxw_radial_fact_kernel: CB lm columns per workgroup, 16 (two whole
// 128-byte lines of each row) or 32 for the short lengths, one thread per
// (column, n2) in pass A and per (column, k1) in pass B on NT threads; the
// pass-B items past NT (k1 = 16 at n = 153) are second items of the first
// threads, with their spectra held in LDS.
template <int N1, int N2>
struct RadLine {
    static constexpr int NM = N1 > N2 ? N1 : N2;
    static constexpr int CB = NM > 8 ? 16 : 32;
    static constexpr int NT = CB * NM > 256 ? 256 : CB * NM;
    static constexpr int NB2 = CB * N1 > NT ? CB * N1 - NT : 0;   // second pass-B items
};

template <int N1, int N2, int D>
__global__ __launch_bounds__((RadLine<N1, N2>::NT), 3) void uxw_radial_fact_kernel(UxwRadArgs A, const double2 *tw)
{
    constexpr int N = N1 * N2, CB = RadLine<N1, N2>::CB, NT = RadLine<N1, N2>::NT;
    constexpr int NB2 = RadLine<N1, N2>::NB2;
    constexpr int DW = D > 0 ? D : 1;
    __shared__ double2 buf[CB * N];
    __shared__ double2 twl[N];   // e^{-2 pi i t / n}
    __shared__ double zl[N], kzl[N];
    __shared__ double dirl[D > 0 ? kRadMaxC * D : 1];
    // wall rows of the workgroup's columns, [j][column]: j < D row nr - D + j,
    // j >= D row 2 D - 1 - j (the D matching rows at each wall)
    __shared__ double2 wall[2 * DW][CB];
    // spectra of the second pass-B items, [item][k2]
    __shared__ double2 xs2[NB2 > 0 ? NB2 : 1][N2];
    const int tid = threadIdx.x;
    const int nr = N - A.C;
    // blockIdx.y: the field, 0 for W (u_Q, u_S, w_T), 1 for Z (u_T, w_Q, w_S)
    const int fld = blockIdx.y;
    const double2 *src = fld ? A.Z : A.W;
    const int c0 = xcd_tile(blockIdx.x, gridDim.x) * CB;
    // byte offsets of (row, column) in the (nlm, n) arrays: below 2^32 (the
    // launch takes the generic kernel otherwise), so that the loads and
    // stores take a uniform base and a 32-bit offset
    const unsigned ldb = (unsigned)A.ld * 16u;
    for (int t = tid; t < N; t += NT) {
        twl[t] = tw[t];
        zl[t] = A.z[t];
        kzl[t] = A.kz[t];
    }
    if (D > 0) {
        for (int t = tid; t < A.C * D; t += NT) dirl[t] = A.dir[t];
        for (int t = tid; t < 2 * D * CB; t += NT) {
            const int c = t % CB, j = t / CB;
            const int row = j < D ? nr - D + j : 2 * D - 1 - j;
            wall[j][c] = src[(long)row * A.ld + min(c0 + c, A.ncols - 1)];
        }
    }
    // every thread has a pass-B item when CB N1 >= NT; threads t < NB2 have a
    // second one, (column t % CB, k1 = (NT + t) / CB)
    if (fld == 0 && A.zero[0])
#pragma unroll
        for (int a = 0; a < 3; a++)
            for (int t = tid; t < A.C * CB; t += NT) {
                const int col = c0 + t % CB;
                if (col < A.ncols) st_stream(&A.zero[a][(long)(nr + t / CB) * A.ld + col], make_double2(0.0, 0.0));
            }
    const bool inA = tid < CB * N2, inB = CB * N1 >= NT || tid < CB * N1;
    const int cb = tid % CB, k1 = tid / CB;   // pass B
    const bool inB2 = tid < NB2;
    const int cb2 = (NT + tid) % CB, k12 = (NT + tid) / CB;
    const double inv_n = 1.0 / N;
    // l(l+1) of the thread's pass-A column (0 for idle lanes)
    double ll;
    {
        const int lma = c0 + tid % CB;
        const bool okA = inA && lma < A.ncols;
        const int ell = A.ell[okA ? lma : 0];
        ll = okA ? (double)ell * (ell + 1) : 0.0;
    }

    // Pass-A indices of the thread, recomputed in every phase from an opaque
    // copy of tid: shared between phases they (and the row offsets formed
    // from them) would stay live across the barriers and crowd out the
    // spectrum held by pass B.
    struct PA {
        int ca, n2;
        bool ok;           // a real column
        unsigned o0;       // byte offset of row n2 of the column (a valid column for idle lanes)
    };
    auto pa = [&]() {
        int t = tid;
        asm volatile("" : "+v"(t));
        PA p;
        p.ca = t % CB;
        p.n2 = t / CB;
        const int lma = c0 + p.ca;
        p.ok = t < CB * N2 && lma < A.ncols;
        p.o0 = (unsigned)p.n2 * ldb + (unsigned)(p.ok ? lma : 0) * 16u;
        return p;
    };
    auto at = [](const double2 *base, unsigned off) { return *(const double2 *)((const char *)base + off); };
    auto put = [](double2 *base, unsigned off, double2 v) { st_stream((double2 *)((char *)base + off), v); };

    // Rows n2 + N2 n1 of the thread's column of the field: the continuation
    // rows past nr are formed first from the staged walls with the FC-Gram
    // operator (sphp_sca.F90:369-381), then every row clamped to the physical
    // ones is loaded, without branches, so that the wall reads and the loads
    // in flight do not hold registers together.  The continuation rows lie in
    // the last NCR values of n1, since C <= kRadMaxC.
    constexpr int NCR = (kRadMaxC + N2 - 1) / N2 + 1 < N1 ? (kRadMaxC + N2 - 1) / N2 + 1 : N1;
    auto load_rows = [&](const PA &p, double2 (&x)[N1]) {
        double2 cv[NCR];
#pragma unroll
        for (int i = 0; i < NCR; i++) {
            const int r = p.n2 + N2 * (N1 - NCR + i);
            const int ii = min(max(r - nr + 1, 1), A.C > 0 ? A.C : 1);
            double2 v = make_double2(0.0, 0.0);
#pragma unroll
            for (int jj = 0; jj < D; jj++) {
                const double a1 = dirl[(ii - 1) + A.C * jj];
                const double a2 = dirl[(A.C - ii) + A.C * jj];
                const double2 wr = wall[jj][p.ca], wl = wall[D + jj][p.ca];
                v.x += a1 * wr.x + a2 * wl.x;
                v.y += a1 * wr.y + a2 * wl.y;
            }
            cv[i] = v;
        }
        const unsigned olast = p.o0 + (unsigned)(nr - 1 - p.n2) * ldb;   // row nr - 1
#pragma unroll
        for (int n1 = 0; n1 < N1; n1++) {
            const int r = p.n2 + N2 * n1;
            x[n1] = at(src, r < nr ? p.o0 + (unsigned)(N2 * n1) * ldb : olast);
        }
#pragma unroll
        for (int n1 = 0; n1 < N1; n1++) {
            const int r = p.n2 + N2 * n1;
            double2 v = x[n1];
            if (D > 0 && n1 >= N1 - NCR && r >= nr) v = cv[n1 - (N1 - NCR)];
            x[n1] = p.ok ? v : make_double2(0.0, 0.0);
        }
    };
    // pass A forward: N1-point DFTs, twiddles, to LDS as [column][k1][n2]
    auto fwd_a = [&](const PA &p, const double2 (&x)[N1]) {
        dft_odd_emit<N1>(x, -1, [&](int q, double2 v) { buf[p.ca * N + q * N2 + p.n2] = cmul(v, twl[(p.n2 * q) % N]); });
    };
    // pass B forward of item (column c, kk): N2-point DFTs -> X[kk + N1 k2]
    auto fwd_b = [&](int c, int kk, double2 (&y)[N2]) {
#pragma unroll
        for (int q = 0; q < N2; q++) y[q] = buf[c * N + kk * N2 + q];
        dft_odd<N2>(y, -1);
    };
    // pass B inverse: N2-point DFTs of a spectrum, twiddles, to LDS as [column][n2][k1]
    auto inv_b = [&](int c, int kk, const double2 (&y)[N2]) {
        dft_odd_emit<N2>(y, +1, [&](int q, double2 v) { buf[c * N + q * N1 + kk] = cmul(v, cconj(twl[(q * kk) % N])); });
    };
    // the forward pass B of the second items, their spectra kept in xs2
    auto fwd_b2 = [&]() {
        if (inB2) {
            double2 z[N2];
            fwd_b(cb2, k12, z);
#pragma unroll
            for (int q = 0; q < N2; q++) xs2[tid][q] = z[q];
        }
    };
    // the inverse pass B of the second items, of the spectrum times f(k, X)
    auto inv_b2 = [&](auto &&f) {
        if (inB2) {
            double2 z[N2];
#pragma unroll
            for (int q = 0; q < N2; q++) z[q] = f(kzl[k12 + N1 * q], xs2[tid][q]);
            inv_b(cb2, k12, z);
        }
    };
    // pass A inverse: N1-point DFTs; emit(n1, value at row n2 + N2 n1)
    auto inv_a = [&](const PA &p, double2 (&x)[N1], auto &&emit) {
#pragma unroll
        for (int q = 0; q < N1; q++) x[q] = buf[p.ca * N + p.n2 * N1 + q];
        dft_odd_emit<N1>(x, +1, emit);
    };

    // i kz / n times a spectrum value (the first radial derivative)
    auto ddr = [&](double kz, double2 v) {
        const double k = kz * inv_n;
        return make_double2(-k * v.y, k * v.x);
    };
    double2 x[N1], y[N2], X[N2];
    __syncthreads();
    if (fld == 0) {
        // ---- W: u_Q = l(l+1) W / r^2, forward transform
        if (inA) {
            const PA p = pa();
            load_rows(p, x);
            if (p.ok)
#pragma unroll
                for (int n1 = 0; n1 < N1; n1++) {
                    const int r = p.n2 + N2 * n1;
                    if (r < nr) {
                        const double rr = zl[r];
                        put(A.out[0], p.o0 + (unsigned)(N2 * n1) * ldb, cscale(x[n1], ll / (rr * rr)));
                    }
                }
            fwd_a(p, x);
        }
        __syncthreads();
        if (inB) fwd_b(cb, k1, X);
        fwd_b2();
        __syncthreads();
        // ---- u_S = dW / r
        if (inB) {
#pragma unroll
            for (int q = 0; q < N2; q++) y[q] = ddr(kzl[k1 + N1 * q], X[q]);
            inv_b(cb, k1, y);
        }
        inv_b2(ddr);
        __syncthreads();
        if (inA) {
            const PA p = pa();
            inv_a(p, x, [&](int n1, double2 v) {
                const int r = p.n2 + N2 * n1;
                if (p.ok && r < nr) put(A.out[1], p.o0 + (unsigned)(N2 * n1) * ldb, cscale(v, 1.0 / zl[r]));
            });
        }
        if (A.nout == 6) {
            __syncthreads();
            // ---- w_T = -(d2W - l(l+1) W / r^2) / r
            auto d2dr2 = [&](double k, double2 v) { return cscale(v, -k * k * inv_n); };
            if (inB) {
#pragma unroll
                for (int q = 0; q < N2; q++) y[q] = d2dr2(kzl[k1 + N1 * q], X[q]);
                inv_b(cb, k1, y);
            }
            inv_b2(d2dr2);
            __syncthreads();
            if (inA) {
                const PA p = pa();
                // W at this thread's rows again, each row loaded where its
                // output is formed (clamped: no branch before the load)
                inv_a(p, x, [&](int n1, double2 v) {
                    const int r = p.n2 + N2 * n1;
                    const unsigned o = p.o0 + (unsigned)(N2 * min(n1, (nr - 1 - p.n2) / N2)) * ldb;
                    const double2 w = at(A.W, o);
                    if (p.ok && r < nr) {
                        const double rr = zl[r], r2 = 1.0 / (rr * rr);
                        put(A.out[5], o, cscale(make_double2(v.x - ll * w.x * r2, v.y - ll * w.y * r2), -1.0 / rr));
                    }
                });
            }
        }
    } else {
        // ---- Z: u_T = Z / r, w_Q = l(l+1) Z / r^2, forward transform
        if (inA) {
            const PA p = pa();
            load_rows(p, x);
            if (p.ok)
#pragma unroll
                for (int n1 = 0; n1 < N1; n1++) {
                    const int r = p.n2 + N2 * n1;
                    if (r < nr) {
                        const double rr = zl[r];
                        const unsigned o = p.o0 + (unsigned)(N2 * n1) * ldb;
                        put(A.out[2], o, cscale(x[n1], 1.0 / rr));
                        if (A.nout == 6) put(A.out[3], o, cscale(x[n1], ll / (rr * rr)));
                    }
                }
            if (A.nout == 6) fwd_a(p, x);
        }
        if (A.nout == 6) {   // the spectrum of Z only feeds the curl outputs
            __syncthreads();
            if (inB) fwd_b(cb, k1, X);
            fwd_b2();
            __syncthreads();
            // ---- w_S = dZ / r
            if (inB) {
#pragma unroll
                for (int q = 0; q < N2; q++) y[q] = ddr(kzl[k1 + N1 * q], X[q]);
                inv_b(cb, k1, y);
            }
            inv_b2(ddr);
            __syncthreads();
            if (inA) {
                const PA p = pa();
                inv_a(p, x, [&](int n1, double2 v) {
                    const int r = p.n2 + N2 * n1;
                    if (p.ok && r < nr) put(A.out[4], p.o0 + (unsigned)(N2 * n1) * ldb, cscale(v, 1.0 / zl[r]));
                });
            }
        }
    }
}

// Factor pair of n handled by uxw_radial_fact_kernel, or 0.
static int radial_factorised(int n)
{
    return (n == 153 || n == 121 || n == 49) ? n : 0;
}

// Single-field radial transforms for n = N1 N2, with the register DFTs and
// the two passes of uxw_radial_fact_kernel; the modes and stored rows are
// those of z_kernel: 0 forward (continuation, FFT -1, / n), 1 inverse
// (FFT +1), 2 derivative of order A.order (forward, (i kz)^p / n, inverse)
// from A.ain into A.a.  Every one of the n rows is stored.  A workgroup
// reads all rows of its CB columns before it writes any, so a and ain may be
// the same array.
template <int N1, int N2, int D>
__global__ __launch_bounds__(256) void z_fact_kernel(ZArgs A, int mode, const double2 *tw)
{
    constexpr int N = N1 * N2, CB = RadFact<N1, N2>::CB;
    __shared__ double2 buf[CB * N];
    __shared__ double2 twl[N];   // e^{-2 pi i t / n}
    __shared__ double kzl[N];
    __shared__ double dirl[D > 0 ? kRadMaxC * D : 1];
    const int tid = threadIdx.x;
    for (int t = tid; t < N; t += 256) {
        twl[t] = tw[t];
        kzl[t] = A.kz[t];
    }
    if (D > 0)
        for (int t = tid; t < A.C * D; t += 256) dirl[t] = A.dir[t];
    const int nr = N - A.C;
    const int c0 = xcd_tile(blockIdx.x, gridDim.x) * CB;
    const bool inA = tid < CB * N2, inB = tid < CB * N1;
    const int ca = tid % CB, n2 = tid / CB;   // pass A
    const int cb = tid % CB, k1 = tid / CB;   // pass B
    const bool okA = inA && c0 + ca < A.ncols, okB = inB && c0 + cb < A.ncols;
    const int cola = okA ? c0 + ca : 0, colb = okB ? c0 + cb : 0;   // valid columns for idle lanes' loads
    const double2 *src = (mode == 2) ? A.ain : A.a;
    __syncthreads();
    double2 x[N1], y[N2];
    if (mode == 1) {
        // pass-B thread (column, k1) holds the spectrum rows k1 + N1 q
        if (inB) {
#pragma unroll
            for (int q = 0; q < N2; q++) y[q] = src[(long)(k1 + N1 * q) * A.ld + colb];
#pragma unroll
            for (int q = 0; q < N2; q++) y[q] = okB ? y[q] : make_double2(0.0, 0.0);
            dft_odd<N2>(y, +1);
#pragma unroll
            for (int q = 0; q < N2; q++) buf[cb * N + q * N1 + k1] = cmul(y[q], cconj(twl[(q * k1) % N]));
        }
    } else {
        // pass A: rows n2 + N2 n1, continuation rows from the wall rows
        if (inA) {
            fact_load_rows<N1, N2, D>(src, A.ld, cola, okA, n2, nr, A.C, dirl, x);
            dft_odd<N1>(x, -1);
#pragma unroll
            for (int q = 0; q < N1; q++) buf[ca * N + q * N2 + n2] = cmul(x[q], twl[(n2 * q) % N]);
        }
        __syncthreads();
        // pass B: y[q] = X[k1 + N1 q]
        if (inB) {
#pragma unroll
            for (int q = 0; q < N2; q++) y[q] = buf[cb * N + k1 * N2 + q];
            dft_odd<N2>(y, -1);
        }
        if (mode == 0) {
            if (okB)
#pragma unroll
                for (int q = 0; q < N2; q++) A.a[(long)(k1 + N1 * q) * A.ld + c0 + cb] = cscale(y[q], 1.0 / N);
            return;
        }
        __syncthreads();
        // (i kz)^p / n, then the inverse pass B
        if (inB) {
            const int pm = A.order & 3;
#pragma unroll
            for (int q = 0; q < N2; q++) {
                double kp = 1.0;
                for (int e = 0; e < A.order; e++) kp *= kzl[k1 + N1 * q];
                double2 v = cscale(y[q], kp / N);
                if (pm == 1) v = make_double2(-v.y, v.x);
                else if (pm == 2) v = make_double2(-v.x, -v.y);
                else if (pm == 3) v = make_double2(v.y, -v.x);
                y[q] = v;
            }
            dft_odd<N2>(y, +1);
#pragma unroll
            for (int q = 0; q < N2; q++) buf[cb * N + q * N1 + k1] = cmul(y[q], cconj(twl[(q * k1) % N]));
        }
    }
    __syncthreads();
    // inverse pass A: rows n2 + N2 n1
    if (inA) {
#pragma unroll
        for (int q = 0; q < N1; q++) x[q] = buf[ca * N + n2 * N1 + q];
        dft_odd<N1>(x, +1);
        if (okA)
#pragma unroll
            for (int n1 = 0; n1 < N1; n1++) A.a[(long)(n2 + N2 * n1) * A.ld + c0 + ca] = x[n1];
    }
}

// ------------------------------ Rader radial step, prime n = RA RB + 1 (281)
// For a prime radial length the transform is Rader's: with g a primitive root
// of n, X[g^-q] = x_0 + conv[q], conv the cyclic convolution of length n - 1
// of a_p = x[g^p] with c_m = e^{s 2 pi i g^-m / n} (the generic kernels do
// the same in LDS Stockham stages).  Here every convolution runs as register
// passes over n - 1 = RA RB, and the Rader orderings are folded into the HBM
// loads and stores, so that LDS carries only the two transposes of each
// convolution:
//   forward (s = -1): a thread holds a_p = x[g^p] for p = n2 + RB n1 and ends
//     with X at the rows g^-p;
//   inverse (s = +1): from b_p = S[g^-p], the same thread ends with
//     Y[g^p] = S_0 + conv'[p], conv' = b * h, h_d = e^{+2 pi i g^d / n}, whose
//     kernel spectrum is cf_+ read backwards;
// so the inverse of a spectrum formed in place lands on the rows the thread
// loaded, and row 0 (X[0] = x_0 + sum a, Y[0] = S_0 + sum b) comes from the
// DC term of the convolution's own transform.

// Prime-factor (Good-Thomas) register DFTs of R = P Q points, P and Q
// coprime: input x[(Q a + P b) mod R], output X[(Q AL c + P BE d) mod R] with
// AL = Q^-1 mod P and BE = P^-1 mod Q, so the Q-point DFTs over b and the
// P-point DFTs over a need no twiddles.  X[k] = sum_n x[n] e^{sign 2 pi i n k / R}.
template <int R> struct Pfa;
template <> struct Pfa<14> {
    static constexpr int P = 2, Q = 7, AL = 1, BE = 4;
};
template <> struct Pfa<20> {
    static constexpr int P = 4, Q = 5, AL = 1, BE = 4;
};

template <int R, typename E>
__device__ __forceinline__ void dft_pfa_emit(const double2 (&x)[R], int sign, E &&emit)
{
    constexpr int P = Pfa<R>::P, Q = Pfa<R>::Q, AL = Pfa<R>::AL, BE = Pfa<R>::BE;
    static_assert(P * Q == R && (Q * AL) % P == 1 && (P * BE) % Q == 1, "prime-factor map");
    double2 u[P][Q];
#pragma unroll
    for (int a = 0; a < P; a++) {
#pragma unroll
        for (int b = 0; b < Q; b++) u[a][b] = x[(Q * a + P * b) % R];
        dft_odd<Q>(u[a], sign);
    }
    // each output handed to emit(k, X[k]) as soon as its P-point DFT is done
#pragma unroll
    for (int d = 0; d < Q; d++) {
        if constexpr (P == 2) {
            const double2 s0 = cadd(u[0][d], u[1][d]), s1 = csub(u[0][d], u[1][d]);
            u[0][d] = s0;
            u[1][d] = s1;
        } else {
            dft4_inplace(u[0][d], u[1][d], u[2][d], u[3][d], sign);
        }
#pragma unroll
        for (int c = 0; c < P; c++) emit((Q * AL * c + P * BE * d) % R, u[c][d]);
    }
}

template <int R>
__device__ __forceinline__ void dft_pfa(double2 (&x)[R], int sign)
{
    dft_pfa_emit<R>(x, sign, [&](int k, double2 y) { x[k] = y; });
}

// An index that the compiler cannot compute before the value y exists: LDS
// reads addressed by it are issued where y is used, not hoisted above the
// DFT that forms y (where they would hold their registers beside its values).
__device__ __forceinline__ int after(int i, double2 y)
{
    asm volatile("" : "+v"(i) : "v"(y.x));
    return i;
}

constexpr int rd_powmod(int b, int e, int m)
{
    long r = 1;
    for (int i = 0; i < e; i++) r = r * b % m;
    return (int)r;
}
// smallest primitive root of a prime n (the rule of the host's tables)
constexpr int rd_primroot(int n)
{
    for (int g = 2; g < n; g++) {
        bool ok = true;
        int m = n - 1;
        for (int f = 2; f <= m; f++)
            if (m % f == 0) {
                if (rd_powmod(g, (n - 1) / f, n) == 1) ok = false;
                while (m % f == 0) m /= f;
            }
        if (ok) return g;
    }
    return 0;
}

// Tiling of the Rader kernels: CB lm columns per workgroup (one 128-byte line
// of each row), one pass-A item (column, n2) per thread, pass-B items
// (column, k1) on the first CB RA threads; 53 KB of LDS, three workgroups per
// CU.
template <int RA, int RB>
struct RaderLine {
    static constexpr int N1 = RA * RB, N = N1 + 1;
    static constexpr int CB = 8;
    static constexpr int NT = CB * RB;
    static constexpr int G = rd_primroot(N), GI = rd_powmod(G, N - 2, N);
    static constexpr int GR = rd_powmod(G, RB, N), GIR = rd_powmod(GI, RB, N);
    static constexpr int NTW = (RA - 1) * (RB - 1) + 1;   // twiddles e^{-2 pi i t / N1}, t = n2 k1
    static_assert(RB <= 32, "rows() takes n2 < 32");
};

// LDS of the Rader kernels: the transpose buffer [RA][RB][CB], the two kernel
// spectra, the twiddles, the wall rows [2D][CB], the radii and the FC-Gram
// operator.
template <int RA, int RB, int D>
static size_t rader_lds(int C)
{
    using L = RaderLine<RA, RB>;
    const int DW = D > 0 ? D : 1;
    return sizeof(double2) * (L::N1 * L::CB + 2 * L::N1 + L::NTW + 2 * DW * L::CB) + sizeof(double) * (L::N + C * DW);
}

struct RaderTabs {
    const double2 *kf;   // forward: cf_-[k] = DFT of c_m (s = -1) / (n - 1)
    const double2 *kp;   // inverse: cf_+[k], read at (n - 1 - k) mod (n - 1)
    const double2 *tw;   // e^{-2 pi i t / (n - 1)}
};

// One cyclic convolution of length N1 = RA RB, in place on the values
// v[n1] = x[n2 + RB n1] of the thread's pass-A item (column tid % CB,
// n2 = tid / CB): conv = IDFT(DFT(x) K), K (in LDS) holding the 1 / N1.
//   pass A: RA-point DFTs over n1, twiddles e^{-2 pi i n2 k1 / N1}, to
//     T[k1][n2][column];
//   pass B (threads < CB RA, item (column, k1 = tid / CB)): RB-point DFTs over
//     n2 -> k2, times K[k1 + RA k2], RB-point inverse DFTs over k2 -> n2,
//     twiddles e^{+2 pi i n2 k1 / N1}, back to the slots it read;
//   pass A': RA-point inverse DFTs over k1 -> conv[n2 + RB n1].
// Pass A of a following convolution writes only the slots pass A' of the same
// thread read: two barriers per convolution.  dc: on pass-B items with k1 = 0,
// the sum of the inputs.
template <int RA, int RB, int CB>
__device__ __forceinline__ void rader_conv(double2 (&v)[RA], const double2 *K, double2 *T, const double2 *twl,
                                           int tid, double2 &dc)
{
    const int c = tid % CB, j = tid / CB;
    dft_pfa_emit<RA>(v, -1, [&](int k1, double2 y) {
        T[(k1 * RB + j) * CB + c] = k1 == 0 ? y : cmul(y, twl[after(j, y) * k1]);
    });
    __syncthreads();
    if (tid < CB * RA) {
        double2 w[RB];
#pragma unroll
        for (int q = 0; q < RB; q++) w[q] = T[(j * RB + q) * CB + c];
        dft_pfa_emit<RB>(w, -1, [&](int k2, double2 y) {
            if (k2 == 0) dc = y;
            w[k2] = cmul(y, K[after(j, y) + RA * k2]);
        });
        dft_pfa_emit<RB>(w, +1, [&](int q, double2 y) {
            T[(j * RB + q) * CB + c] = q == 0 ? y : cmul(y, cconj(twl[q * after(j, y)]));
        });
    }
    __syncthreads();
#pragma unroll
    for (int k1 = 0; k1 < RA; k1++) v[k1] = T[(k1 * RB + j) * CB + c];
    dft_pfa<RA>(v, +1);
}

// State and helpers shared by the Rader kernels: LDS staging, the rows of a
// thread's pass-A item, the FC-Gram continuation and the wavenumbers.
template <int RA, int RB, int D>
struct RaderCtx {
    using L = RaderLine<RA, RB>;
    static constexpr int N = L::N, N1 = L::N1, CB = L::CB, NT = L::NT, DW = D > 0 ? D : 1;
    double2 *T, *Kf, *Ki, *twl, *wall;   // Kf forward, Ki inverse kernel spectrum
    double *zl, *dirl;
    int nr, C, c0, ncols;
    unsigned ldb, oc;   // bytes per row; byte offset of the thread's column (a valid one for idle lanes)
    bool ok;            // the thread's column is a real one
    double dk;          // kz[k] = k' dk, k' = k or k - n

    __device__ RaderCtx(double2 *sm, int tile, const RaderTabs &R, const double2 *src, const double *z,
                        const double *kz, const double *dir, long ld, int ncols_, int C_)
    {
        T = sm;
        Kf = T + N1 * CB;
        Ki = Kf + N1;
        twl = Ki + N1;
        wall = twl + L::NTW;
        zl = (double *)(wall + 2 * DW * CB);
        dirl = zl + N;
        C = C_;
        nr = N - C;
        ncols = ncols_;
        c0 = tile * CB;
        ldb = (unsigned)ld * 16u;
        const int tid = threadIdx.x;
        for (int t = tid; t < L::NTW; t += NT) twl[t] = R.tw[t];
        for (int t = tid; t < N1; t += NT) {
            Kf[t] = R.kf[t];
            Ki[t] = R.kp[t == 0 ? 0 : N1 - t];
        }
        if (z)
            for (int t = tid; t < N; t += NT) zl[t] = z[t];
        if (D > 0 && src) {
            for (int t = tid; t < C * D; t += NT) dirl[t] = dir[t];
            // [j][column]: j < D row nr - D + j, j >= D row 2 D - 1 - j
            for (int t = tid; t < 2 * D * CB; t += NT) {
                const int c = t % CB, j = t / CB;
                const int row = j < D ? nr - D + j : 2 * D - 1 - j;
                wall[j * CB + c] = src[(long)row * ld + min(c0 + c, ncols - 1)];
            }
        }
        const int lm = c0 + tid % CB;
        ok = lm < ncols;
        oc = (unsigned)(ok ? lm : ncols - 1) * 16u;
        dk = kz[1];
    }
    // r[n1] = base^(n2 + RB n1) mod n for the thread's n2, from an opaque copy
    // of the thread index (recomputed where used rather than held)
    template <int BASE, int STEP>
    __device__ __forceinline__ void rows(int (&r)[RA]) const
    {
        int t = threadIdx.x;
        asm volatile("" : "+v"(t));
        const int j = t / CB;
        int x = 1, b = BASE;
#pragma unroll
        for (int bit = 0; bit < 5; bit++) {
            if ((j >> bit) & 1) x = x * b % N;
            b = b * b % N;
        }
#pragma unroll
        for (int n1 = 0; n1 < RA; n1++) {
            r[n1] = x;
            x = x * STEP % N;
        }
    }
    __device__ __forceinline__ double kzof(int k) const { return (double)(k <= (N - 1) / 2 ? k : k - N) * dk; }
    __device__ __forceinline__ unsigned off(int row) const { return (unsigned)row * ldb + oc; }
    static __device__ __forceinline__ double2 at(const double2 *base, unsigned o)
    {
        return *(const double2 *)((const char *)base + o);
    }
    static __device__ __forceinline__ void put(double2 *base, unsigned o, double2 v)
    {
        st_stream((double2 *)((char *)base + o), v);
    }
    // x[g^p] at p = n2 + RB n1 and x_0 of the thread's column: every load from
    // a row clamped to the physical ones, issued at once, then the rows past
    // nr replaced by the FC-Gram continuation of the staged walls
    // (sphp_sca.F90:369-381); zeros for idle lanes
    __device__ __forceinline__ void load(const double2 *src, double2 (&v)[RA], double2 &x0) const
    {
        int r[RA];
        rows<L::G, L::GR>(r);
        x0 = at(src, oc);
#pragma unroll
        for (int n1 = 0; n1 < RA; n1++) v[n1] = at(src, off(min(r[n1], nr - 1)));
        const int ca = threadIdx.x % CB;
#pragma unroll
        for (int n1 = 0; n1 < RA; n1++) {
            if (D > 0 && r[n1] >= nr) {
                const int ii = r[n1] - nr + 1;
                double2 s = make_double2(0.0, 0.0);
#pragma unroll
                for (int jj = 0; jj < D; jj++) {
                    const double a1 = dirl[(ii - 1) + C * jj], a2 = dirl[(C - ii) + C * jj];
                    const double2 wr = wall[jj * CB + ca], wl = wall[(D + jj) * CB + ca];
                    s.x += a1 * wr.x + a2 * wl.x;
                    s.y += a1 * wr.y + a2 * wl.y;
                }
                v[n1] = s;
            }
            if (!ok) v[n1] = make_double2(0.0, 0.0);
        }
        if (!ok) x0 = make_double2(0.0, 0.0);
    }
};

// i kz / n times a spectrum value (the first radial derivative)
__device__ __forceinline__ double2 rader_ddr(double kz, double inv_n, double2 v)
{
    const double k = kz * inv_n;
    return make_double2(-k * v.y, k * v.x);
}

// The radial step of the nonlinear term (the outputs of uxw_radial_kernel) for
// a prime n = RA RB + 1, as three kinds of workgroup, each with one forward and
// one inverse transform of its column tile:
//   0: W -> u_Q = l(l+1) W / r^2, u_S = dW / r
//   1: Z -> u_T = Z / r, w_Q = l(l+1) Z / r^2, w_S = dZ / r
//   2: W -> w_T = -(d2W - l(l+1) W / r^2) / r
// (with nout = 3 only kinds 0 and 1, and kind 1 without transforms).  Holding
// W's spectrum for both of its inverse transforms would spill; the kinds of a
// tile are consecutive workgroups on one XCD, so kind 2 reads W from L2.
// kz[0] = 0, so the derivative spectra have no row-0 term.
template <int RA, int RB, int D>
__global__ __launch_bounds__((RaderLine<RA, RB>::NT), 3) void uxw_radial_rader_kernel(UxwRadArgs A, RaderTabs R)
{
    using X = RaderCtx<RA, RB, D>;
    constexpr int N = X::N, CB = X::CB, NT = X::NT;
    constexpr int G = X::L::G, GR = X::L::GR, GI = X::L::GI, GIR = X::L::GIR;
    extern __shared__ double2 sm[];
    const int nk = A.nout == 6 ? 3 : 2;
    const int t = xcd_tile(blockIdx.x, gridDim.x), kind = t % nk;
    const double2 *src = kind == 1 ? A.Z : A.W;
    const X c(sm, t / nk, R, src, A.z, A.kz, A.dir, A.ld, A.ncols, A.C);
    const int tid = threadIdx.x;
    const int nr = c.nr;
    if (kind == 0 && A.zero[0])
#pragma unroll
        for (int a = 0; a < 3; a++)
            for (int e = tid; e < A.C * CB; e += NT) {
                const int col = c.c0 + e % CB;
                if (col < A.ncols) st_stream(&A.zero[a][(long)(nr + e / CB) * A.ld + col], make_double2(0.0, 0.0));
            }
    double ll;
    {
        const int lm = c.c0 + tid % CB;
        const int ell = A.ell[c.ok ? lm : 0];
        ll = c.ok ? (double)ell * (ell + 1) : 0.0;
    }
    double2 v[RA], x0, dc;
    int r[RA];
    __syncthreads();
    c.load(src, v, x0);
    // ---- the outputs formed from the loads, at the rows read and at row 0
    if (kind != 2 && c.ok) {
        c.template rows<G, GR>(r);
        auto emit = [&](unsigned o, double2 x, double rr) {
            if (kind == 0) {
                X::put(A.out[0], o, cscale(x, ll / (rr * rr)));
            } else {
                X::put(A.out[2], o, cscale(x, 1.0 / rr));
                if (A.nout == 6) X::put(A.out[3], o, cscale(x, ll / (rr * rr)));
            }
        };
#pragma unroll
        for (int n1 = 0; n1 < RA; n1++)
            if (r[n1] < nr) emit(c.off(r[n1]), v[n1], c.zl[r[n1]]);
        if (tid < CB) emit(c.oc, x0, c.zl[0]);
    }
    if (kind == 1 && A.nout != 6) return;
    // ---- forward transform; the spectrum at the rows g^-p times i kz / n
    //      (kinds 0, 1) or -kz^2 / n (kind 2); inverse transform
    rader_conv<RA, RB, CB>(v, c.Kf, c.T, c.twl, tid, dc);
    c.template rows<GI, GIR>(r);
    const double inv_n = 1.0 / N;
#pragma unroll
    for (int n1 = 0; n1 < RA; n1++) {
        const double k = c.kzof(r[n1]);
        const double2 x = cadd(x0, v[n1]);
        v[n1] = kind == 2 ? cscale(x, -k * k * inv_n) : rader_ddr(k, inv_n, x);
    }
    rader_conv<RA, RB, CB>(v, c.Ki, c.T, c.twl, tid, dc);
    // ---- out = (y - s l(l+1) W / r^2) f: s = 0, f = 1 / r for u_S and w_S;
    //      s = 1, f = -1 / r for w_T, with W read again at each row
    double2 *dst = kind == 0 ? A.out[1] : kind == 1 ? A.out[4] : A.out[5];
    const double sw = kind == 2 ? ll : 0.0, sf = kind == 2 ? -1.0 : 1.0;
    auto fin = [&](unsigned o, double2 y, double2 w, double rr) {
        const double r2 = 1.0 / (rr * rr);
        X::put(dst, o, cscale(make_double2(y.x - sw * w.x * r2, y.y - sw * w.y * r2), sf / rr));
    };
    if (tid < CB && c.ok) fin(c.oc, dc, x0, c.zl[0]);
    c.template rows<G, GR>(r);
    if (kind == 2) {
#pragma unroll
        for (int n1 = 0; n1 < RA; n1++) {
            const unsigned o = c.off(min(r[n1], nr - 1));
            const double2 w = X::at(A.W, o);
            if (c.ok && r[n1] < nr) fin(o, v[n1], w, c.zl[r[n1]]);
        }
    } else if (c.ok) {
#pragma unroll
        for (int n1 = 0; n1 < RA; n1++)
            if (r[n1] < nr) fin(c.off(r[n1]), v[n1], make_double2(0.0, 0.0), c.zl[r[n1]]);
    }
}

// Single-field radial transforms of prime length n = RA RB + 1, with the modes
// and stored rows of z_kernel: 0 forward (continuation, FFT -1, / n), 1
// inverse (FFT +1), 2 derivative of order A.order >= 1 (forward, (i kz)^p / n,
// inverse) from A.ain into A.a.  Every one of the n rows is stored; a
// workgroup reads all rows of its columns before it writes any, so a and ain
// may be the same array.
template <int RA, int RB, int D, int MODE>
__global__ __launch_bounds__((RaderLine<RA, RB>::NT), 3) void z_rader_kernel(ZArgs A, RaderTabs R)
{
    constexpr int mode = MODE;
    using X = RaderCtx<RA, RB, D>;
    constexpr int N = X::N, CB = X::CB;
    constexpr int G = X::L::G, GR = X::L::GR, GI = X::L::GI, GIR = X::L::GIR;
    extern __shared__ double2 sm[];
    const double2 *src = (mode == 2) ? A.ain : A.a;
    const X c(sm, xcd_tile(blockIdx.x, gridDim.x), R, mode == 1 ? nullptr : src, nullptr, A.kz, A.dir, A.ld, A.ncols,
              A.C);
    const int tid = threadIdx.x;
    double2 v[RA], x0, dc;
    int r[RA];
    __syncthreads();
    if (mode == 1) {
        // the spectrum at the rows g^-p, and S_0
        c.template rows<GI, GIR>(r);
        x0 = X::at(src, c.oc);
#pragma unroll
        for (int n1 = 0; n1 < RA; n1++) v[n1] = X::at(src, c.off(r[n1]));
#pragma unroll
        for (int n1 = 0; n1 < RA; n1++) v[n1] = c.ok ? v[n1] : make_double2(0.0, 0.0);
        rader_conv<RA, RB, CB>(v, c.Ki, c.T, c.twl, tid, dc);
        c.template rows<G, GR>(r);
        if (c.ok) {
#pragma unroll
            for (int n1 = 0; n1 < RA; n1++) X::put(A.a, c.off(r[n1]), cadd(x0, v[n1]));
            if (tid < CB) X::put(A.a, c.oc, cadd(x0, dc));
        }
        return;
    }
    c.load(src, v, x0);
    rader_conv<RA, RB, CB>(v, c.Kf, c.T, c.twl, tid, dc);
    const double inv_n = 1.0 / N;
    c.template rows<GI, GIR>(r);
    if (mode == 0) {
        if (c.ok) {
#pragma unroll
            for (int n1 = 0; n1 < RA; n1++) X::put(A.a, c.off(r[n1]), cscale(cadd(x0, v[n1]), inv_n));
            if (tid < CB) X::put(A.a, c.oc, cscale(cadd(x0, dc), inv_n));
        }
        return;
    }
    // (i kz)^p / n at the rows g^-p
    const int pm = A.order & 3;
#pragma unroll
    for (int n1 = 0; n1 < RA; n1++) {
        const double k = c.kzof(r[n1]);
        double kp = inv_n;
        for (int e = 0; e < A.order; e++) kp *= k;
        double2 w = cscale(cadd(x0, v[n1]), kp);
        if (pm == 1) w = make_double2(-w.y, w.x);
        else if (pm == 2) w = make_double2(-w.x, -w.y);
        else if (pm == 3) w = make_double2(w.y, -w.x);
        v[n1] = w;
    }
    rader_conv<RA, RB, CB>(v, c.Ki, c.T, c.twl, tid, dc);
    c.template rows<G, GR>(r);
    if (c.ok) {
#pragma unroll
        for (int n1 = 0; n1 < RA; n1++) X::put(A.a, c.off(r[n1]), v[n1]);
        if (tid < CB) X::put(A.a, c.oc, dc);
    }
}

// The Rader register-pass kernels take n = 281 (n - 1 = 14 x 20) with the
// host's primitive root and Rader tables, the standard FC-Gram matching
// (d = 5) or the periodic radius, and byte offsets below 2^32.  Returns the
// D to launch with, or -1.
static int rader_fits(const Plan *p, long ncols, long ld)
{
    using L = RaderLine<14, 20>;
    if (p->nz != L::N || !p->zfft.rader || !p->zfft.cf[0] || primitive_root(p->nz) != L::G) return -1;
    if ((long)p->nz * (ld > ncols ? ld : ncols) * (long)sizeof(double2) >= (1L << 32)) return -1;
    return p->C == 0 ? 0 : (p->d == 5 && p->C <= kRadMaxC ? 5 : -1);
}

static RaderTabs rader_tabs(const Plan *p)
{
    RaderTabs R;
    R.kf = p->zfft.cf[0];
    R.kp = p->zfft.cf[1];
    R.tw = p->tw_z;
    return R;
}

template <int D, int MODE>
static int launch_z_rader_m(Plan *p, const ZArgs &A)
{
    using L = RaderLine<14, 20>;
    const size_t shm = rader_lds<14, 20, D>(p->C);
    const int lrc = allow_lds(z_rader_kernel<14, 20, D, MODE>, shm);
    if (lrc) return lrc;
    hipLaunchKernelGGL((z_rader_kernel<14, 20, D, MODE>), dim3((A.ncols + L::CB - 1) / L::CB), dim3(L::NT), shm,
                       p->stream, A, rader_tabs(p));
    RD_LAUNCH_CHECK();
    return RD_OK;
}

template <int D>
static int launch_z_rader(Plan *p, const ZArgs &A, int mode)
{
    return mode == 0 ? launch_z_rader_m<D, 0>(p, A) : mode == 1 ? launch_z_rader_m<D, 1>(p, A)
                                                                : launch_z_rader_m<D, 2>(p, A);
}

template <int D>
static int launch_uxw_rader(Plan *p, const UxwRadArgs &A)
{
    using L = RaderLine<14, 20>;
    const size_t shm = rader_lds<14, 20, D>(p->C);
    const int lrc = allow_lds(uxw_radial_rader_kernel<14, 20, D>, shm);
    if (lrc) return lrc;
    const int nk = A.nout == 6 ? 3 : 2;
    hipLaunchKernelGGL((uxw_radial_rader_kernel<14, 20, D>), dim3(nk * ((A.ncols + L::CB - 1) / L::CB)), dim3(L::NT),
                       shm, p->stream, A, rader_tabs(p));
    RD_LAUNCH_CHECK();
    return RD_OK;
}

// The factorised single-field radial transform when n is one of the factor
// pairs above; returns 1 (nothing launched) otherwise.
int z_fact_launch(Plan *p, const ZArgs &A, int mode)
{
    // prime n = 281: the Rader register passes (a derivative of order 0 keeps
    // a row-0 term they do not form)
    {
        const int rd = rader_fits(p, A.ncols, A.ld);
        if (rd >= 0 && !(mode == 2 && A.order < 1)) return rd == 5 ? launch_z_rader<5>(p, A, mode) : launch_z_rader<0>(p, A, mode);
    }
    // the standard FC-Gram matching (d = 5) and the periodic radius; other
    // (C, d) take the Stockham kernel
    const int fd = p->C == 0 ? 0 : (p->d == 5 && p->C <= kRadMaxC ? 5 : -1);
    switch (fd < 0 ? 0 : radial_factorised(p->nz)) {
#define RD_ZF_CASE(NV, F1, F2)                                                                  \
    case NV: {                                                                                  \
        constexpr int CB = RadFact<F1, F2>::CB;                                                 \
        if (fd == 5)                                                                            \
            hipLaunchKernelGGL((z_fact_kernel<F1, F2, 5>), dim3((A.ncols + CB - 1) / CB), dim3(256), 0,      \
                               p->stream, A, mode, p->tw_z);                                    \
        else                                                                                    \
            hipLaunchKernelGGL((z_fact_kernel<F1, F2, 0>), dim3((A.ncols + CB - 1) / CB), dim3(256), 0,      \
                               p->stream, A, mode, p->tw_z);                                    \
        RD_LAUNCH_CHECK();                                                                      \
        return RD_OK;                                                                           \
    }
        RD_ZF_CASE(153, 17, 9)
        RD_ZF_CASE(121, 11, 11)
        RD_ZF_CASE(49, 7, 7)
#undef RD_ZF_CASE
    default: return 1;
    }
}

int launch_uxw_radial(Plan *p, const double2 *W, const double2 *Z, double2 *const *qst, int nout,
                      double2 *const *zero3)
{
    if (nout != 6 && nout != 3) return RD_ERR_ARG;
    UxwRadArgs A;
    for (int i = 0; i < 3; i++) A.zero[i] = zero3 ? zero3[i] : nullptr;
    A.W = W;
    A.Z = Z;
    for (int i = 0; i < 6; i++) A.out[i] = i < nout ? qst[i] : nullptr;
    A.nout = nout;
    A.dir = p->d_dir;
    A.kz = p->d_kz;
    A.z = p->d_z;
    A.ell = p->d_ell;
    A.ld = p->nlm;
    A.ncols = p->nlm;
    A.n = p->nz;
    A.C = p->C;
    A.d = p->d;
    // the factorised kernel for the standard FC-Gram matching (d = 5) and the
    // periodic radius, when byte offsets into the (nlm, nz) arrays fit in 32
    // bits; other plans take the generic kernel
    const bool off32 = (long)p->nz * p->nlm * (long)sizeof(double2) < (1L << 32);
    const int fd = !off32 ? -1 : p->C == 0 ? 0 : (p->d == 5 && p->C <= kRadMaxC ? 5 : -1);
    switch (fd < 0 ? 0 : radial_factorised(p->nz)) {
#define RD_F_CASE(NV, F1, F2)                                                                   \
    case NV: {                                                                                  \
        constexpr int CB = RadLine<F1, F2>::CB, NT = RadLine<F1, F2>::NT;                       \
        if (fd == 5)                                                                            \
            hipLaunchKernelGGL((uxw_radial_fact_kernel<F1, F2, 5>), dim3((p->nlm + CB - 1) / CB, 2), dim3(NT), \
                               0, p->stream, A, p->tw_z);                                       \
        else                                                                                    \
            hipLaunchKernelGGL((uxw_radial_fact_kernel<F1, F2, 0>), dim3((p->nlm + CB - 1) / CB, 2), dim3(NT), \
                               0, p->stream, A, p->tw_z);                                       \
        RD_LAUNCH_CHECK();                                                                      \
        return RD_OK;                                                                           \
    }
        RD_F_CASE(153, 17, 9)
        RD_F_CASE(121, 11, 11)
        RD_F_CASE(49, 7, 7)
#undef RD_F_CASE
    default: break;
    }
    {
        const int rd = off32 ? rader_fits(p, p->nlm, p->nlm) : -1;
        if (rd >= 0) return rd == 5 ? launch_uxw_rader<5>(p, A) : launch_uxw_rader<0>(p, A);
    }
    size_t shm = sizeof(double2) * ((6 * UTC + 1) * p->nz + kZScalars);
    const dim3 grid((p->nlm + UTC - 1) / UTC), block(256);
#define RD_U_CASE(RBV)                                                                        \
    case RBV: {                                                                                 \
        int lrc = allow_lds(uxw_radial_kernel<RBV>, shm);                                       \
        if (lrc) return lrc;                                                                    \
        hipLaunchKernelGGL(uxw_radial_kernel<RBV>, grid, block, shm, p->stream, A, p->zfft, p->tw_z); \
        break;                                                                                  \
    }
    switch (z_radix(p->zfft)) {
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
    for (int i = 0; zero3 && i < 3; i++) {
        const int rc = launch_zero_rows(p, zero3[i], p->nz - p->C, p->C);
        if (rc) return rc;
    }
    return RD_OK;
}

}  // namespace rd
