// No-slip boundary conditions of the velocity on the shell walls, on the
// device in FP64: the FC-Gram derivative matrices, the per-l W recovery
// matrices M_l (src/boundary/boundary_mod.f90:96-241), W = M_l phi(3:N-2)
// (:244-292) and the two forms of v_imposebc (src/boundary/vboundary.f90:59-130).
//
// Layouts (N = nr physical points, all column major like the Fortran arrays):
//   d1, d2  : N x N,            d[i + N*j] = d^p e_j / dr^p at r_i
//   binv    : N x (N-4) per l,  binv[i + N*j + N*(N-4)*(l-1)], l = 1..lmax
//   d2bor   : 4 x N,            rows 1, 2, N-1, N of d2
#include "rd_internal.h"

#include <algorithm>
#include <vector>

namespace rd {

using v4d = __attribute__((__vector_size__(4 * sizeof(double)))) double;

static __device__ __forceinline__ v4d mfma64(double a, double b, v4d c)
{
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

static __device__ __forceinline__ long lm_of(int lmax, int l, int m)
{
    return (long)m * (2 * lmax + 1 - m) / 2 + l;
}

// Columns e_0..e_{N-1} of an (N columns) x (n rows) complex array, plane stride N.
__global__ void identity_cols_kernel(double2 *a, int N, int n)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long)N * n) return;
    const int j = (int)(e % N), k = (int)(e / N);
    a[e] = make_double2(j == k ? 1.0 : 0.0, 0.0);
}

// d[i + N*j] = Re b[j + N*i] for i, j < N.
__global__ void extract_dmat_kernel(const double2 *b, double *d, int N)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= N * N) return;
    const int i = e % N, j = e / N;
    d[e] = b[j + (long)N * i].x;
}

// B_l (boundary_mod.f90:177-191): rows 1 and N impose W = 0, rows 2 and N-1
// dW/dr = 0 (rows 1 and N of d1), rows 3..N-2 are D_l = d2 - l(l+1)/r^2;
// each row is divided by its largest absolute entry (esc).
__global__ __launch_bounds__(256) void bl_build_kernel(const double *d1, const double *d2,
                                                       const double *z, double *B, double *esc,
                                                       int N)
{
    const int l = blockIdx.x + 1;
    double *b = B + (long)blockIdx.x * N * N;
    double *s = esc + (long)blockIdx.x * N;
    const double ll = (double)l * (l + 1);
    for (int i = threadIdx.x; i < N; i += blockDim.x) {
        double mx = 0.0;
        for (int j = 0; j < N; j++) {
            double v;
            if (i == 0) v = (j == 0) ? 1.0 : 0.0;
            else if (i == 1) v = d1[0 + (long)N * j];
            else if (i == N - 2) v = d1[N - 1 + (long)N * j];
            else if (i == N - 1) v = (j == N - 1) ? 1.0 : 0.0;
            else {
                v = d2[i + (long)N * j];
                if (j == i) v -= ll / (z[i] * z[i]);
            }
            b[i + (long)N * j] = v;
            mx = fmax(mx, fabs(v));
        }
        s[i] = mx;
        for (int j = 0; j < N; j++) b[i + (long)N * j] = b[i + (long)N * j] / mx;
    }
}

// In-place LU with partial pivoting of one scaled B_l per workgroup
// (right-looking, the pivot is the first row of largest magnitude, as
// LAPACK's DGETRF picks it).  fail receives the smallest l whose matrix is
// singular.
__global__ __launch_bounds__(256) void bl_lu_kernel(double *B, int *piv, int N, int *fail)
{
    __shared__ double rv[4];
    __shared__ int ri[4];
    __shared__ int pk_s;
    const int l = blockIdx.x + 1;
    double *a = B + (long)blockIdx.x * N * N;
    int *pv = piv + (long)blockIdx.x * N;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    for (int k = 0; k < N; k++) {
        // pivot search over rows k..N-1 of column k
        double bv = -1.0;
        int bi = N;
        for (int i = k + tid; i < N; i += 256) {
            const double v = fabs(a[i + (long)N * k]);
            if (v > bv) { bv = v; bi = i; }
        }
        for (int o = 32; o > 0; o >>= 1) {
            const double ov = __shfl_xor(bv, o);
            const int oi = __shfl_xor(bi, o);
            if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
        }
        if (lane == 0) { rv[wave] = bv; ri[wave] = bi; }
        __syncthreads();
        if (tid == 0) {
            double v = rv[0];
            int i0 = ri[0];
            for (int w = 1; w < 4; w++)
                if (rv[w] > v || (rv[w] == v && ri[w] < i0)) { v = rv[w]; i0 = ri[w]; }
            pk_s = (v > 0.0) ? i0 : -1;
            pv[k] = i0;
        }
        __syncthreads();
        const int pk = pk_s;
        if (pk < 0) {
            if (tid == 0) atomicMin(fail, l);
            return;
        }
        if (pk != k)
            for (int j = tid; j < N; j += 256) {
                const double t = a[k + (long)N * j];
                a[k + (long)N * j] = a[pk + (long)N * j];
                a[pk + (long)N * j] = t;
            }
        __syncthreads();
        const double inv = 1.0 / a[k + (long)N * k];
        for (int i = k + 1 + tid; i < N; i += 256) a[i + (long)N * k] *= inv;
        __syncthreads();
        const int c = N - k - 1;
        for (int e = tid; e < c * c; e += 256) {
            const int i = k + 1 + e % c, j = k + 1 + e / c;
            a[i + (long)N * j] -= a[i + (long)N * k] * a[k + (long)N * j];
        }
        __syncthreads();
    }
}

// M_l = columns 3..N-2 of B_l^-1 (boundary_mod.f90:201-210): solve
// (S^-1 B_l) x = e_{j+2} / esc_{j+2}, one right-hand side per thread, in a
// scratch laid out [i][j] so that the threads' accesses coalesce.
__global__ __launch_bounds__(256) void bl_solve_kernel(const double *B, const int *piv,
                                                       const double *esc, double *X, double *binv,
                                                       int N)
{
    const int K = N - 4;
    const double *a = B + (long)blockIdx.x * N * N;
    const int *pv = piv + (long)blockIdx.x * N;
    const double *s = esc + (long)blockIdx.x * N;
    double *x = X + (long)blockIdx.x * N * K;
    double *out = binv + (long)blockIdx.x * N * K;
    for (int j = threadIdx.x; j < K; j += blockDim.x) {
        for (int i = 0; i < N; i++) x[(long)i * K + j] = (i == j + 2) ? 1.0 / s[j + 2] : 0.0;
        for (int k = 0; k < N; k++) {
            const int pk = pv[k];
            if (pk != k) {
                const double t = x[(long)k * K + j];
                x[(long)k * K + j] = x[(long)pk * K + j];
                x[(long)pk * K + j] = t;
            }
        }
        for (int c = 0; c < N; c++) {
            const double xc = x[(long)c * K + j];
            for (int i = c + 1; i < N; i++) x[(long)i * K + j] -= a[i + (long)N * c] * xc;
        }
        for (int c = N - 1; c >= 0; c--) {
            const double xc = x[(long)c * K + j] / a[c + (long)N * c];
            x[(long)c * K + j] = xc;
            for (int i = 0; i < c; i++) x[(long)i * K + j] -= a[i + (long)N * c] * xc;
        }
        for (int i = 0; i < N; i++) out[i + (long)N * j] = x[(long)i * K + j];
    }
}

// Rows 1, 2, N-1, N of d2 (boundary_mod.f90:230-235).
__global__ void d2bor_kernel(const double *d2, double *d2bor, int N)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= N) return;
    d2bor[0 + 4 * j] = d2[0 + (long)N * j];
    d2bor[1 + 4 * j] = d2[1 + (long)N * j];
    d2bor[2 + 4 * j] = d2[N - 2 + (long)N * j];
    d2bor[3 + 4 * j] = d2[N - 1 + (long)N * j];
}

// W = M_l phi(3:N-2) for every lm with l >= 1 (boundary_mod.f90:260-289), as
// one real GEMM per l whose columns are the real and imaginary parts of the
// modes m of that l.  Workgroup tile: one l, 32 modes (64 columns), 64 rows
// r (one 16-row MFMA tile per wave); the phi rows are staged through LDS in
// chunks of 16.
constexpr int IV_KC = 16;

struct InvdlArgs {
    const double *binv;
    const int *queue;
    const double2 *phi;
    double2 *W;
    long nlm;
    int N, lmax;
};

__global__ __launch_bounds__(256) void invdl_kernel(InvdlArgs A)
{
    __shared__ double Bs[IV_KC][64 + 2];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int ent = A.queue[blockIdx.x];
    const int l = ent & 0xffff, m0 = ent >> 16;
    const int nm = min(32, l + 1 - m0);
    const int N = A.N, K = N - 4;
    const int row = blockIdx.y * 64 + wave * 16 + (lane & 15);
    const int kk = lane >> 4, col = lane & 15;
    const double *M = A.binv + (long)N * K * (l - 1);

    v4d acc[4];
#pragma unroll
    for (int c = 0; c < 4; c++) acc[c] = (v4d){0.0, 0.0, 0.0, 0.0};

    for (int kc = 0; kc < K; kc += IV_KC) {
        for (int e = tid; e < IV_KC * 32; e += 256) {
            const int r = e / 32, mi = e % 32, k = kc + r;
            double2 v = make_double2(0.0, 0.0);
            if (mi < nm && k < K) v = A.phi[lm_of(A.lmax, l, m0 + mi) + A.nlm * (k + 2)];
            Bs[r][2 * mi] = v.x;
            Bs[r][2 * mi + 1] = v.y;
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < IV_KC / 4; s++) {
            const int k = kc + 4 * s + kk;
            const double a = (row < N && k < K) ? M[row + (long)N * k] : 0.0;
#pragma unroll
            for (int c = 0; c < 4; c++) acc[c] = mfma64(a, Bs[4 * s + kk][c * 16 + col], acc[c]);
        }
        __syncthreads();
    }
    double *Wd = (double *)A.W;
#pragma unroll
    for (int c = 0; c < 4; c++)
#pragma unroll
        for (int v = 0; v < 4; v++) {
            const int r = blockIdx.y * 64 + wave * 16 + kk + 4 * v;
            const int cc = c * 16 + col, mi = cc >> 1;
            if (r < N && mi < nm) Wd[2 * (lm_of(A.lmax, l, m0 + mi) + A.nlm * r) + (cc & 1)] = acc[c][v];
        }
}

// W of the l = 0 mode (lm = 0) on the physical rows is zero: the gauge freedom.
__global__ void zero_l0_kernel(double2 *W, long nlm, int N)
{
    const int r = blockIdx.x * blockDim.x + threadIdx.x;
    if (r < N) W[nlm * r] = make_double2(0.0, 0.0);
}

// v_imposebc_fast, step 3 (vboundary.f90:117-128): phi on rows 1, 2, N-1, N
// from d2bor and W; the l = 0 mode of phi is zero on every row.
__global__ void imposebc_rows_kernel(const double *d2bor, const double *z, const int *ell,
                                     const double2 *pol, double2 *phi, long nlm, int N, int nz)
{
    const long lm = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (lm >= nlm) return;
    const int l = ell[lm];
    if (l == 0) {
        for (int k = 0; k < nz; k++) phi[lm + nlm * k] = make_double2(0.0, 0.0);
        return;
    }
    const int fila[4] = {0, 1, N - 2, N - 1};
    double2 s[4];
#pragma unroll
    for (int q = 0; q < 4; q++) s[q] = make_double2(0.0, 0.0);
    for (int j = 0; j < N; j++) {
        const double2 w = pol[lm + nlm * j];
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const double d = d2bor[q + 4 * j];
            s[q].x += d * w.x;
            s[q].y += d * w.y;
        }
    }
    const double ll = (double)l * (l + 1);
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int i = fila[q];
        const double2 w = pol[lm + nlm * i];
        const double f = ll / (z[i] * z[i]);
        phi[lm + nlm * i] = make_double2(s[q].x - f * w.x, s[q].y - f * w.y);
    }
}

// ------------------------------------------------------------------ host

int build_derivr_mat(Plan *p, int order, double *d)
{
    const int N = p->nr, n = p->nz;
    double2 *a = nullptr, *b = nullptr;
    RD_HIP(hipMalloc(&a, sizeof(double2) * (size_t)N * n));
    RD_HIP(hipMalloc(&b, sizeof(double2) * (size_t)N * n));
    const long tot = (long)N * n;
    hipLaunchKernelGGL(identity_cols_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, p->stream, a,
                       N, n);
    RD_LAUNCH_CHECK();
    int rc = launch_derivr_cols(p, order, a, b, N, N);
    if (!rc) {
        hipLaunchKernelGGL(extract_dmat_kernel, dim3((N * N + 255) / 256), dim3(256), 0, p->stream, b, d, N);
        RD_LAUNCH_CHECK();
    }
    RD_HIP(hipStreamSynchronize(p->stream));
    (void)hipFree(a);
    (void)hipFree(b);
    return rc;
}

void free_bc(Plan *p)
{
    void *dev[] = {p->bc_d1, p->bc_d2, p->bc_binv, p->bc_d2bor, p->bc_queue};
    for (void *d : dev)
        if (d) (void)hipFree(d);
    p->bc_d1 = p->bc_d2 = p->bc_binv = p->bc_d2bor = nullptr;
    p->bc_queue = nullptr;
    p->bc_ready = 0;
}

int build_bc(Plan *p, int *failed_l)
{
    const int N = p->nr, K = N - 4, lmax = p->lmax;
    *failed_l = 0;
    if (N < 6 || lmax < 1) return RD_ERR_STATE;
    free_bc(p);
    RD_HIP(hipMalloc(&p->bc_d1, sizeof(double) * N * N));
    RD_HIP(hipMalloc(&p->bc_d2, sizeof(double) * N * N));
    RD_HIP(hipMalloc(&p->bc_binv, sizeof(double) * (size_t)N * K * lmax));
    RD_HIP(hipMalloc(&p->bc_d2bor, sizeof(double) * 4 * N));
    int rc = build_derivr_mat(p, 1, p->bc_d1);
    if (!rc) rc = build_derivr_mat(p, 2, p->bc_d2);
    if (rc) return rc;

    double *B = nullptr, *esc = nullptr, *X = nullptr;
    int *piv = nullptr, *fail = nullptr;
    RD_HIP(hipMalloc(&B, sizeof(double) * (size_t)N * N * lmax));
    RD_HIP(hipMalloc(&X, sizeof(double) * (size_t)N * K * lmax));
    RD_HIP(hipMalloc(&esc, sizeof(double) * (size_t)N * lmax));
    RD_HIP(hipMalloc(&piv, sizeof(int) * (size_t)N * lmax));
    RD_HIP(hipMalloc(&fail, sizeof(int)));
    const int big = 1 << 30;
    RD_HIP(hipMemcpyAsync(fail, &big, sizeof(int), hipMemcpyHostToDevice, p->stream));
    hipLaunchKernelGGL(bl_build_kernel, dim3(lmax), dim3(256), 0, p->stream, p->bc_d1, p->bc_d2, p->d_z, B,
                       esc, N);
    RD_LAUNCH_CHECK();
    hipLaunchKernelGGL(bl_lu_kernel, dim3(lmax), dim3(256), 0, p->stream, B, piv, N, fail);
    RD_LAUNCH_CHECK();
    hipLaunchKernelGGL(bl_solve_kernel, dim3(lmax), dim3(128), 0, p->stream, B, piv, esc, X, p->bc_binv, N);
    RD_LAUNCH_CHECK();
    hipLaunchKernelGGL(d2bor_kernel, dim3((N + 63) / 64), dim3(64), 0, p->stream, p->bc_d2, p->bc_d2bor, N);
    RD_LAUNCH_CHECK();
    int hfail = big;
    RD_HIP(hipMemcpyAsync(&hfail, fail, sizeof(int), hipMemcpyDeviceToHost, p->stream));
    RD_HIP(hipStreamSynchronize(p->stream));
    (void)hipFree(B);
    (void)hipFree(X);
    (void)hipFree(esc);
    (void)hipFree(piv);
    (void)hipFree(fail);
    if (hfail != big) {
        *failed_l = hfail;
        return RD_ERR_LU;
    }

    // (l, m chunk) tiles of invDl, heaviest l first
    std::vector<int> q;
    for (int l = lmax; l >= 1; l--)
        for (int m0 = 0; m0 <= l; m0 += 32) q.push_back(l | (m0 << 16));
    p->bc_qlen = (int)q.size();
    RD_HIP(hipMalloc(&p->bc_queue, sizeof(int) * q.size()));
    RD_HIP(hipMemcpyAsync(p->bc_queue, q.data(), sizeof(int) * q.size(), hipMemcpyHostToDevice, p->stream));
    RD_HIP(hipStreamSynchronize(p->stream));
    p->bc_ready = 1;
    return RD_OK;
}

int launch_invdl(Plan *p, const double2 *phi, double2 *W)
{
    const int N = p->nr;
    InvdlArgs A;
    A.binv = p->bc_binv;
    A.queue = p->bc_queue;
    A.phi = phi;
    A.W = W;
    A.nlm = p->nlm;
    A.N = N;
    A.lmax = p->lmax;
    hipLaunchKernelGGL(invdl_kernel, dim3((unsigned)p->bc_qlen, (unsigned)((N + 63) / 64)), dim3(256), 0,
                       p->stream, A);
    RD_LAUNCH_CHECK();
    hipLaunchKernelGGL(zero_l0_kernel, dim3((N + 63) / 64), dim3(64), 0, p->stream, W, (long)p->nlm, N);
    RD_LAUNCH_CHECK();
    return launch_zero_rows(p, W, N, p->nz - N);
}

int launch_imposebc_rows(Plan *p, const double2 *pol, double2 *phi)
{
    hipLaunchKernelGGL(imposebc_rows_kernel, dim3((unsigned)((p->nlm + 255) / 256)), dim3(256), 0, p->stream,
                       p->bc_d2bor, p->d_z, p->d_ell, pol, phi, (long)p->nlm, p->nr, p->nz);
    RD_LAUNCH_CHECK();
    return RD_OK;
}

}  // namespace rd
