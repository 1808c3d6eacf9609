// Pointwise pseudospectral operators on (nlm, nz) complex and
// (nlat, nphi, planes) real arrays (src/pseudo/pseudospec_hd.f90).
// Grid-stride loops over the flat index; lm is the fastest index.
#include "rd_internal.h"

#include <cmath>

namespace rd {

static inline dim3 grid_for(long n)
{
    long b = (n + 255) / 256;
    if (b > 65536L * 8) b = 65536L * 8;
    return dim3((unsigned)b);
}

// b = i^p kz^p a  (derivr_k, pseudospec_hd.f90:86-95)
__global__ void derivr_k_kernel(const double2 *a, double2 *b, const double *kz, long nlm, int nz, int p)
{
    const long n = nlm * nz;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int k = (int)(i / nlm);
        double kp = 1.0;
        for (int q = 0; q < p; q++) kp *= kz[k];
        double2 v = a[i];
        v.x *= kp;
        v.y *= kp;
        const int pm = p & 3;
        if (pm == 1) v = make_double2(-v.y, v.x);
        else if (pm == 2) v = make_double2(-v.x, -v.y);
        else if (pm == 3) v = make_double2(v.y, -v.x);
        b[i] = v;
    }
}

// b = -l(l+1) a  (laplang, :160-169)
__global__ void laplang_kernel(const double2 *a, double2 *b, const int *ell, long nlm, int nz)
{
    const long n = nlm * nz;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int l = ell[i % nlm];
        const double f = -(double)l * (l + 1);
        const double2 v = a[i];
        b[i] = make_double2(f * v.x, f * v.y);
    }
}

// b = z_k^p a  (multr, :196-205)
__global__ void multr_kernel(const double2 *a, double2 *b, const double *z, long nlm, int nz, int p)
{
    const long n = nlm * nz;
    const int ap = p < 0 ? -p : p;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int k = (int)(i / nlm);
        double zp = 1.0;
        for (int q = 0; q < ap; q++) zp *= z[k];
        if (p < 0) zp = 1.0 / zp;
        const double2 v = a[i];
        b[i] = make_double2(zp * v.x, zp * v.y);
    }
}

// z = a x + b y on n doubles  (saxpby_c / saxpby_r, :314-323, :352-363)
__global__ void saxpby_kernel(long n, double *z, const double *x, double a, const double *y, double b)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        z[i] = a * x[i] + b * y[i];
}

// C = A x B in (r, theta, phi)  (prodvec, :235-248)
__global__ void prodvec_kernel(long n, const double *ar, const double *at, const double *ap,
                               const double *br, const double *bt, const double *bp, double *cr,
                               double *ct, double *cp)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const double xr = ar[i], xt = at[i], xp = ap[i];
        const double yr = br[i], yt = bt[i], yp = bp[i];
        cr[i] = xt * yp - xp * yt;
        ct[i] = xp * yr - xr * yp;
        cp[i] = xr * yt - xt * yr;
    }
}

// a *= exp(-alpha (2 kz / (nz Dkz))^(2p)), alpha = 16 ln 10, p = 50  (fc_filter_k, :425-437)
__global__ void fc_filter_kernel(double2 *a, const double *kz, double dkz, long nlm, int nz)
{
    const long n = nlm * nz;
    const double alpha = 16.0 * log(10.0);
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int k = (int)(i / nlm);
        const double x = 2.0 * kz[k] / nz / dkz;
        const double f = exp(-alpha * pow(x, 100.0));
        double2 v = a[i];
        a[i] = make_double2(f * v.x, f * v.y);
    }
}

// a(lm, 1) = a(lm, nz - C) = 0  (dirichlet, boundary_mod.f90:82-90)
__global__ void dirichlet_kernel(double2 *a, long nlm, int last)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < nlm; i += (long)gridDim.x * blockDim.x) {
        a[i] = make_double2(0.0, 0.0);
        a[(long)last * nlm + i] = make_double2(0.0, 0.0);
    }
}

// Wall value from its normal derivative along z (neumann_reconstruct /
// robin_reconstruct, fcgram_mod.f90:466-517, 612-650, on the (lm, nz) layout):
// row rb holds g on entry; f_b = w[d-1] g + sum_{k<d} w[k-1] a(rows[k-1]),
// divided by coef[lm] w[d-1] + 1 when coef is given (Robin).
struct FcRecon {
    double w[8];
    int rows[8];
    int d, rb;
};

__global__ void fc_reconstruct_kernel(double2 *a, long nlm, FcRecon R, const double *coef)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < nlm; i += (long)gridDim.x * blockDim.x) {
        const double2 g = a[(long)R.rb * nlm + i];
        double2 v = make_double2(R.w[R.d - 1] * g.x, R.w[R.d - 1] * g.y);
        for (int k = 0; k < R.d - 1; k++) {
            const double2 x = a[(long)R.rows[k] * nlm + i];
            v.x += R.w[k] * x.x;
            v.y += R.w[k] * x.y;
        }
        if (coef) {
            const double den = coef[i] * R.w[R.d - 1] + 1.0;
            v = make_double2(v.x / den, v.y / den);
        }
        a[(long)R.rb * nlm + i] = v;
    }
}

__global__ void zero_kernel(double2 *a, long n)
{
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        a[i] = make_double2(0.0, 0.0);
}

int launch_derivr_k(Plan *p, int order, const double2 *a, double2 *b)
{
    hipLaunchKernelGGL(derivr_k_kernel, grid_for((long)p->nlm * p->nz), dim3(256), 0, p->stream, a, b,
                       p->d_kz, (long)p->nlm, p->nz, order);
    RD_LAUNCH_CHECK();
    return RD_OK;
}

int launch_laplang(Plan *p, const double2 *a, double2 *b)
{
    hipLaunchKernelGGL(laplang_kernel, grid_for((long)p->nlm * p->nz), dim3(256), 0, p->stream, a, b,
                       p->d_ell, (long)p->nlm, p->nz);
    RD_LAUNCH_CHECK();
    return RD_OK;
}

int launch_multr(Plan *p, int pw, const double2 *a, double2 *b)
{
    hipLaunchKernelGGL(multr_kernel, grid_for((long)p->nlm * p->nz), dim3(256), 0, p->stream, a, b,
                       p->d_z, (long)p->nlm, p->nz, pw);
    RD_LAUNCH_CHECK();
    return RD_OK;
}

int launch_saxpby(Plan *p, long n, double *z, const double *x, double a, const double *y, double b)
{
    hipLaunchKernelGGL(saxpby_kernel, grid_for(n), dim3(256), 0, p->stream, n, z, x, a, y, b);
    RD_LAUNCH_CHECK();
    return RD_OK;
}

int launch_prodvec(Plan *p, long npts, const double *ar, const double *at, const double *ap,
                   const double *br, const double *bt, const double *bp, double *cr, double *ct,
                   double *cp)
{
    if (npts <= 0) return RD_OK;
    hipLaunchKernelGGL(prodvec_kernel, grid_for(npts), dim3(256), 0, p->stream, npts, ar, at, ap, br, bt,
                       bp, cr, ct, cp);
    RD_LAUNCH_CHECK();
    return RD_OK;
}

int launch_fc_filter_k(Plan *p, double2 *a)
{
    const double dkz = 2.0 * M_PI / (p->dz * p->nz);
    hipLaunchKernelGGL(fc_filter_kernel, grid_for((long)p->nlm * p->nz), dim3(256), 0, p->stream, a,
                       p->d_kz, dkz, (long)p->nlm, p->nz);
    RD_LAUNCH_CHECK();
    return RD_OK;
}

int launch_dirichlet(Plan *p, double2 *a)
{
    hipLaunchKernelGGL(dirichlet_kernel, grid_for(p->nlm), dim3(256), 0, p->stream, a, (long)p->nlm,
                       p->nz - p->C - 1);
    RD_LAUNCH_CHECK();
    return RD_OK;
}

int launch_fc_reconstruct(Plan *p, double2 *a, int boun, const double *neu, const double *coef)
{
    FcRecon R = {};
    R.d = p->d;
    if (R.d < 1 || R.d > 8) return RD_ERR_ARG;
    R.rb = (boun == 5) ? 0 : p->nr - 1;
    for (int k = 1; k <= R.d; k++) {
        R.w[k - 1] = neu[k - 1];
        if (k < R.d) R.rows[k - 1] = (boun == 5) ? R.d - k : p->nr - 1 - R.d + k;
    }
    hipLaunchKernelGGL(fc_reconstruct_kernel, grid_for(p->nlm), dim3(256), 0, p->stream, a, (long)p->nlm, R,
                       coef);
    RD_LAUNCH_CHECK();
    return RD_OK;
}

int launch_zero_rows(Plan *p, double2 *a, int row0, int nrows)
{
    if (nrows <= 0) return RD_OK;
    const long n = (long)nrows * p->nlm;
    hipLaunchKernelGGL(zero_kernel, grid_for(n), dim3(256), 0, p->stream, a + (long)row0 * p->nlm, n);
    RD_LAUNCH_CHECK();
    return RD_OK;
}

}  // namespace rd
