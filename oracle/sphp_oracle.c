/*
 * CPU FP64 restatement of REDONDITO's pseudospectral transform path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle: the tests, the smoke
 * check and the CPU-baseline leg of the benchmark use it as the checker.
 * Nothing in the product library links or calls it.
 *
 * It restates, in plain C99 + OpenMP, the semantics of:
 *   src/sphp/sphp_sca.F90   2D/1D/3D scalar transforms (stage order,
 *                           continuation indexing :369-381, 1/n :391)
 *   src/sphp/sphp_vec.F90   vector (Q,S,T) transforms :49-165
 *   src/sphp/fcgram_mod.f90 dir = A.Q^T :257-258
 *   src/pseudo/pseudospec_mod.f90  grid rule :66-68, z grid :196-200,
 *                           kz :223-227
 *   src/pseudo/pseudospec_hd.f90   derivr/laplang/multr/prodvec/dlr/
 *                           saxpby/fc_filter/copy
 *   src/boundary/boundary_mod.f90  derivr_mat :96-137, invDl_setup
 *                           :140-241, invDl :244-292
 * and the SHTns conventions the reference configures
 * (src/sphp/sphp_sca.F90:117-122): Gauss grid, orthonormal harmonics
 * without the Condon-Shortley phase, theta-contiguous planes, m-major lm
 * order LM(l,m) = m(2 lmax + 1 - m)/2 + l.  The real-field rule is
 * f = sum a_l0 Y_l0 + sum_{m>0} 2 Re(a_lm Y_lm) (src/prueba_sphp.f90:5-6),
 * and the vector convention v = Q Y r + S r grad Y - T r x grad Y
 * (src/prueba_vec.f90:5-7).
 *
 * Layouts follow the reference: spectral a(lm, k) complex, lm fastest;
 * physical f(theta, phi, k) real, theta fastest.
 */
#include <omp.h>
#include <stdlib.h>
#include <string.h>
#include <tgmath.h>

/* The arithmetic type.  The default build is FP64 (liboracle.so).  The same
 * source built with -DORC_LONG_DOUBLE (liboracle_ld.so) runs every step in
 * x87 extended precision (64-bit significand) on the same FP64 inputs and
 * tables: that build serves as a higher-precision truth against which the
 * rounding of two FP64 implementations (this one and the GPU's) is measured.
 * Through <tgmath.h> each math call takes the precision of its argument. */
#ifdef ORC_LONG_DOUBLE
typedef long double real;
typedef long double _Complex cplx;
#else
typedef double real;
typedef double _Complex cplx;
#endif
#define ORC_PI ((real)3.14159265358979323846264338327950288L)

/* OpenMP threads of the parallel loops (n < 1: the runtime's default);
 * returns the count now in effect. */
int orc_set_threads(int n)
{
    static int deflt = 0;
    if (!deflt) deflt = omp_get_max_threads();
    omp_set_num_threads(n > 0 ? n : deflt);
    return omp_get_max_threads();
}

/* ------------------------------------------------------------------ */
/* Complex FFT: recursive mixed radix, generic radix-p butterflies.     */
/* ------------------------------------------------------------------ */
typedef struct {
    int n;
    int nfac;
    int fac[64];
    cplx *twf; /* exp(-2 pi i t / n) */
    cplx *twb; /* exp(+2 pi i t / n) */
} orc_fft;

static orc_fft *fft_create(int n)
{
    orc_fft *p = calloc(1, sizeof(*p));
    p->n = n;
    int m = n;
    for (int f = 4; m > 1 && f == 4;) { /* radix 4 first, then primes */
        if (m % 4 == 0) { p->fac[p->nfac++] = 4; m /= 4; } else break;
    }
    for (int f = 2; m > 1; f++)
        while (m % f == 0) { p->fac[p->nfac++] = f; m /= f; }
    p->twf = malloc(sizeof(cplx) * n);
    p->twb = malloc(sizeof(cplx) * n);
    for (int t = 0; t < n; t++) {
        real a = 2.0 * ORC_PI * (real)t / (real)n;
        p->twf[t] = cos(a) - I * sin(a);
        p->twb[t] = cos(a) + I * sin(a);
    }
    return p;
}

static void fft_destroy(orc_fft *p)
{
    if (!p) return;
    free(p->twf);
    free(p->twb);
    free(p);
}

static void fft_rec(int n, const int *fac, const cplx *in, long is, cplx *out,
                    const cplx *tw, int N0, cplx *tmp)
{
    if (n == 1) { out[0] = in[0]; return; }
    int p = fac[0], m = n / p;
    long step = N0 / n;
    for (int j = 0; j < p; j++)
        fft_rec(m, fac + 1, in + j * is, is * p, out + (long)j * m, tw, N0, tmp);
    for (int k1 = 0; k1 < m; k1++) {
        for (int j = 0; j < p; j++)
            tmp[j] = out[(long)j * m + k1] * tw[((long)j * k1 * step) % N0];
        for (int k2 = 0; k2 < p; k2++) {
            cplx s = 0;
            for (int j = 0; j < p; j++)
                s += tmp[j] * tw[((long)j * k2 * m * step) % N0];
            out[k1 + (long)m * k2] = s;
        }
    }
}

/* out = sum_j in[j*is] exp(sign 2 pi i j k / n); sign -1 forward. */
static void fft_exec(const orc_fft *p, int sign, const cplx *in, long is, cplx *out, cplx *tmp)
{
    fft_rec(p->n, p->fac, in, is, out, sign < 0 ? p->twf : p->twb, p->n, tmp);
}

/* ------------------------------------------------------------------ */
/* Grid and associated-Legendre tables                                  */
/* ------------------------------------------------------------------ */

/* Coefficients of P_n(cos t) = sum_{k=0..n} h_k cos((n - 2k) t),
 * h_k = g_k g_{n-k}, g_k = prod_{j<=k} (2j-1)/(2j). */
void orc_gauss_coefs(int n, real *h)
{
    real *g = malloc(sizeof(real) * (n + 1));
    g[0] = 1.0;
    for (int k = 1; k <= n; k++) g[k] = g[k - 1] * (2.0 * k - 1.0) / (2.0 * k);
    for (int k = 0; k <= n; k++) h[k] = g[k] * g[n - k];
    free(g);
}

/* P_n(cos t) and dP_n/dt from the cosine series, the derivative summed with
 * Neumaier compensation. */
static void legendre_cos_series(int n, const real *h, real t, real *P, real *dP)
{
    real p = (n % 2 == 0) ? h[n / 2] : 0.0;
    real s = 0.0, c = 0.0;
    for (int k = 0; k < (n + 1) / 2; k++) {
        int m = n - 2 * k;
        p += 2.0 * h[k] * cos(m * t);
        real term = -2.0 * h[k] * m * sin(m * t);
        real u = s + term;
        c += (fabs(s) >= fabs(term)) ? (s - u) + term : (term - u) + s;
        s = u;
    }
    *P = p;
    *dP = s + c;
}

/* Gauss-Legendre nodes from north to south.  Newton runs on theta with
 * P_n(cos theta) evaluated from its cosine series, which keeps theta (and so
 * sin theta) accurate to an ulp at the poles, where cos theta alone cannot;
 * w = 2 / (dP_n/dtheta)^2. */
void orc_gauss(int nlat, real *ct, real *st, real *wt)
{
    real *h = malloc(sizeof(real) * (nlat + 1));
    orc_gauss_coefs(nlat, h);
    for (int i = 0; i < (nlat + 1) / 2; i++) {
        real th = ORC_PI * (i + 0.75) / (nlat + 0.5);
        real P, dP;
        int extra = 0;
        for (int it = 0; it < 60; it++) {
            legendre_cos_series(nlat, h, th, &P, &dP);
            real dth = P / dP;
            th -= dth;
            if (fabs(dth) < 1e-15 * th && ++extra > 2) break;
        }
        legendre_cos_series(nlat, h, th, &P, &dP);
        real w = 2.0 / (dP * dP);
        ct[i] = cos(th);
        st[i] = sin(th);
        wt[i] = w;
        ct[nlat - 1 - i] = -cos(th);
        st[nlat - 1 - i] = sin(th);
        wt[nlat - 1 - i] = w;
    }
    if (nlat % 2 == 1) { ct[nlat / 2] = 0.0; st[nlat / 2] = 1.0; }
    free(h);
}

static inline long lm_index(int lmax, int l, int m)
{
    return (long)m * (2 * lmax + 1 - m) / 2 + l;
}

/* Orthonormal, no-CS P~_lm(theta_i), dP~_lm/dtheta and m P~_lm / sin(theta),
 * each stored [lm][i] (m-major lm).  P~_lm = s^m Q_lm(x), the recurrence is
 * run on P~ and on D = s^m dQ/dx with a power-of-two scale that keeps the
 * sectoral start s^m representable. */
void orc_legendre(int lmax, int nlat, const real *ct, const real *st,
                  real *P, real *dP, real *mPs)
{
    #pragma omp parallel for schedule(dynamic)
    for (int m = 0; m <= lmax; m++) {
        /* c_m = sqrt((2m+1)/(4 pi)) prod_{k=1..m} sqrt((2k-1)/(2k)) */
        real cm = sqrt((2.0 * m + 1.0) / (4.0 * ORC_PI));
        for (int k = 1; k <= m; k++) cm *= sqrt((real)(2 * k - 1) / (2 * k));
        for (int i = 0; i < nlat; i++) {
            real x = ct[i], s = st[i];
            real v = cm;
            int e2 = 0;
            for (int k = 0; k < m; k++) {
                v *= s;
                if (v < 0x1p-400) { v *= 0x1p400; e2 -= 400; }
            }
            /* l = m */
            real p_lm2 = 0.0, p_lm1 = v;     /* scaled P~ */
            real d_lm2 = 0.0, d_lm1 = 0.0;   /* scaled s^m dQ/dx */
            for (int l = m; l <= lmax; l++) {
                real p, d;
                if (l == m) {
                    p = p_lm1;
                    d = 0.0;
                } else if (l == m + 1) {
                    real a = sqrt((real)(2 * m + 3));
                    p = a * x * p_lm1;
                    d = a * p_lm1;
                } else {
                    real a = sqrt((4.0 * l * l - 1.0) / ((real)l * l - (real)m * m));
                    real b = sqrt(((l - 1.0) * (l - 1.0) - (real)m * m) /
                                    (4.0 * (l - 1.0) * (l - 1.0) - 1.0));
                    p = a * (x * p_lm1 - b * p_lm2);
                    d = a * (p_lm1 + x * d_lm1 - b * d_lm2);
                }
                if (l > m) {
                    p_lm2 = p_lm1; p_lm1 = p;
                    d_lm2 = d_lm1; d_lm1 = d;
                } else {
                    d_lm1 = 0.0;
                }
                long idx = lm_index(lmax, l, m) * nlat + i;
                real pv = ldexp(p, e2), dv = ldexp(d, e2);
                P[idx] = pv;
                dP[idx] = (m * x * pv - s * s * dv) / s;
                mPs[idx] = m * pv / s;
            }
        }
    }
}

/* ------------------------------------------------------------------ */
/* SHT handle                                                           */
/* ------------------------------------------------------------------ */
typedef struct {
    int lmax, nlat, nphi;
    long nlm;
    real *ct, *st, *wt;
    real *P, *dP, *mPs;
    orc_fft *fphi;
} orc_sht;

void *orc_sht_create(int lmax, int nlat, int nphi)
{
    if (lmax < 0 || nlat < 2 || nphi < 2 * lmax + 1) return NULL;
    orc_sht *h = calloc(1, sizeof(*h));
    h->lmax = lmax;
    h->nlat = nlat;
    h->nphi = nphi;
    h->nlm = (long)(lmax + 1) * (lmax + 2) / 2;
    h->ct = malloc(sizeof(real) * nlat);
    h->st = malloc(sizeof(real) * nlat);
    h->wt = malloc(sizeof(real) * nlat);
    orc_gauss(nlat, h->ct, h->st, h->wt);
    h->P = malloc(sizeof(real) * h->nlm * nlat);
    h->dP = malloc(sizeof(real) * h->nlm * nlat);
    h->mPs = malloc(sizeof(real) * h->nlm * nlat);
    orc_legendre(lmax, nlat, h->ct, h->st, h->P, h->dP, h->mPs);
    h->fphi = fft_create(nphi);
    return h;
}

void orc_sht_destroy(void *vh)
{
    orc_sht *h = vh;
    if (!h) return;
    free(h->ct); free(h->st); free(h->wt);
    free(h->P); free(h->dP); free(h->mPs);
    fft_destroy(h->fphi);
    free(h);
}

void orc_sht_grid(void *vh, real *ct, real *st, real *wt)
{
    orc_sht *h = vh;
    memcpy(ct, h->ct, sizeof(real) * h->nlat);
    memcpy(st, h->st, sizeof(real) * h->nlat);
    memcpy(wt, h->wt, sizeof(real) * h->nlat);
}

void orc_sht_tables(void *vh, real *P, real *dP, real *mPs)
{
    orc_sht *h = vh;
    size_t nb = sizeof(real) * h->nlm * h->nlat;
    memcpy(P, h->P, nb);
    memcpy(dP, h->dP, nb);
    memcpy(mPs, h->mPs, nb);
}

/* F[m*nlat + i] -> f[i + nlat*j]: f = Re F_0 + 2 Re sum_{m>0} F_m e^{i m phi_j} */
static void phi_synth(const orc_sht *h, const cplx *F, real *f, cplx *X, cplx *Y, cplx *tmp)
{
    int nlat = h->nlat, nphi = h->nphi, lmax = h->lmax;
    for (int i = 0; i < nlat; i++) {
        for (int k = 0; k < nphi; k++) X[k] = 0;
        X[0] = creal(F[i]);
        for (int m = 1; m <= lmax; m++) {
            X[m] = F[(long)m * nlat + i];
            X[nphi - m] = conj(F[(long)m * nlat + i]);
        }
        fft_exec(h->fphi, +1, X, 1, Y, tmp);
        for (int j = 0; j < nphi; j++) f[i + (long)nlat * j] = creal(Y[j]);
    }
}

/* f -> G[m*nlat + i] = (2 pi / nphi) sum_j f(i,j) e^{-i m phi_j} */
static void phi_anal(const orc_sht *h, const real *f, cplx *G, cplx *X, cplx *Y, cplx *tmp)
{
    int nlat = h->nlat, nphi = h->nphi, lmax = h->lmax;
    real sc = 2.0 * ORC_PI / nphi;
    for (int i = 0; i < nlat; i++) {
        for (int j = 0; j < nphi; j++) X[j] = f[i + (long)nlat * j];
        fft_exec(h->fphi, -1, X, 1, Y, tmp);
        for (int m = 0; m <= lmax; m++) G[(long)m * nlat + i] = sc * Y[m];
    }
}

typedef struct {
    cplx *F0, *F1, *F2, *X, *Y, *tmp;
} sht_work;

static sht_work work_alloc(const orc_sht *h)
{
    sht_work w;
    size_t nF = (size_t)(h->lmax + 1) * h->nlat;
    w.F0 = malloc(sizeof(cplx) * nF);
    w.F1 = malloc(sizeof(cplx) * nF);
    w.F2 = malloc(sizeof(cplx) * nF);
    w.X = malloc(sizeof(cplx) * h->nphi);
    w.Y = malloc(sizeof(cplx) * h->nphi);
    w.tmp = malloc(sizeof(cplx) * h->nphi);
    return w;
}

static void work_free(sht_work *w)
{
    free(w->F0); free(w->F1); free(w->F2);
    free(w->X); free(w->Y); free(w->tmp);
}

/* One plane of SH_to_spat: a(nlm) -> f(nlat, nphi). */
static void sh_to_spat_plane(const orc_sht *h, const cplx *a, real *f, sht_work *w)
{
    int nlat = h->nlat, lmax = h->lmax;
    for (int m = 0; m <= lmax; m++) {
        for (int i = 0; i < nlat; i++) {
            cplx s = 0;
            for (int l = m; l <= lmax; l++) {
                long lm = lm_index(lmax, l, m);
                s += h->P[lm * nlat + i] * a[lm];
            }
            w->F0[(long)m * nlat + i] = s;
        }
    }
    phi_synth(h, w->F0, f, w->X, w->Y, w->tmp);
}

/* One plane of spat_to_SH: f -> a, a_lm = sum_i w_i P~_lm G_m(i). */
static void spat_to_sh_plane(const orc_sht *h, const real *f, cplx *a, sht_work *w)
{
    int nlat = h->nlat, lmax = h->lmax;
    phi_anal(h, f, w->F0, w->X, w->Y, w->tmp);
    for (int m = 0; m <= lmax; m++) {
        for (int l = m; l <= lmax; l++) {
            long lm = lm_index(lmax, l, m);
            cplx s = 0;
            for (int i = 0; i < nlat; i++)
                s += h->wt[i] * h->P[lm * nlat + i] * w->F0[(long)m * nlat + i];
            a[lm] = (m == 0) ? creal(s) : s;
        }
    }
}

/* One plane of SHqst_to_spat:
 *   v_r = sum Q Y;  v_th = dS/dth + (1/s) dT/dph;  v_ph = (1/s) dS/dph - dT/dth */
static void shqst_to_spat_plane(const orc_sht *h, const cplx *q, const cplx *s_, const cplx *t,
                                real *vr, real *vt, real *vp, sht_work *w)
{
    int nlat = h->nlat, lmax = h->lmax;
    for (int m = 0; m <= lmax; m++) {
        for (int i = 0; i < nlat; i++) {
            cplx r = 0, th = 0, ph = 0;
            for (int l = m; l <= lmax; l++) {
                long lm = lm_index(lmax, l, m);
                real P = h->P[lm * nlat + i], dP = h->dP[lm * nlat + i];
                real mP = h->mPs[lm * nlat + i];
                r += P * q[lm];
                th += dP * s_[lm] + I * mP * t[lm];
                ph += I * mP * s_[lm] - dP * t[lm];
            }
            w->F0[(long)m * nlat + i] = r;
            w->F1[(long)m * nlat + i] = th;
            w->F2[(long)m * nlat + i] = ph;
        }
    }
    phi_synth(h, w->F0, vr, w->X, w->Y, w->tmp);
    phi_synth(h, w->F1, vt, w->X, w->Y, w->tmp);
    phi_synth(h, w->F2, vp, w->X, w->Y, w->tmp);
}

/* One plane of spat_to_SHqst (S, T of l = 0 are zero). */
static void spat_to_shqst_plane(const orc_sht *h, const real *vr, const real *vt,
                                const real *vp, cplx *q, cplx *s_, cplx *t, sht_work *w)
{
    int nlat = h->nlat, lmax = h->lmax;
    phi_anal(h, vr, w->F0, w->X, w->Y, w->tmp);
    phi_anal(h, vt, w->F1, w->X, w->Y, w->tmp);
    phi_anal(h, vp, w->F2, w->X, w->Y, w->tmp);
    for (int m = 0; m <= lmax; m++) {
        for (int l = m; l <= lmax; l++) {
            long lm = lm_index(lmax, l, m);
            cplx aq = 0, as = 0, at = 0;
            for (int i = 0; i < nlat; i++) {
                long k = (long)m * nlat + i;
                real wi = h->wt[i];
                real P = h->P[lm * nlat + i], dP = h->dP[lm * nlat + i];
                real mP = h->mPs[lm * nlat + i];
                aq += wi * P * w->F0[k];
                as += wi * (dP * w->F1[k] - I * mP * w->F2[k]);
                at += wi * (-I * mP * w->F1[k] - dP * w->F2[k]);
            }
            real ll = (real)l * (l + 1);
            q[lm] = (m == 0) ? creal(aq) : aq;
            if (l == 0) {
                s_[lm] = 0;
                t[lm] = 0;
            } else {
                s_[lm] = (m == 0) ? creal(as) / ll : as / ll;
                t[lm] = (m == 0) ? creal(at) / ll : at / ll;
            }
        }
    }
}

/* Plane loops.  Spectral a + k*nlm, physical f + k*nlat*nphi. */
void orc_sh_to_spat(void *vh, int nplanes, const real *a_, real *f)
{
    orc_sht *h = vh;
    const cplx *a = (const cplx *)a_;
    long np = (long)h->nlat * h->nphi;
    #pragma omp parallel
    {
        sht_work w = work_alloc(h);
        #pragma omp for schedule(dynamic)
        for (int k = 0; k < nplanes; k++) sh_to_spat_plane(h, a + k * h->nlm, f + k * np, &w);
        work_free(&w);
    }
}

void orc_spat_to_sh(void *vh, int nplanes, const real *f, real *a_)
{
    orc_sht *h = vh;
    cplx *a = (cplx *)a_;
    long np = (long)h->nlat * h->nphi;
    #pragma omp parallel
    {
        sht_work w = work_alloc(h);
        #pragma omp for schedule(dynamic)
        for (int k = 0; k < nplanes; k++) spat_to_sh_plane(h, f + k * np, a + k * h->nlm, &w);
        work_free(&w);
    }
}

void orc_shqst_to_spat(void *vh, int nplanes, const real *q, const real *s, const real *t,
                       real *vr, real *vt, real *vp)
{
    orc_sht *h = vh;
    long np = (long)h->nlat * h->nphi;
    #pragma omp parallel
    {
        sht_work w = work_alloc(h);
        #pragma omp for schedule(dynamic)
        for (int k = 0; k < nplanes; k++)
            shqst_to_spat_plane(h, (const cplx *)q + k * h->nlm, (const cplx *)s + k * h->nlm,
                                (const cplx *)t + k * h->nlm, vr + k * np, vt + k * np,
                                vp + k * np, &w);
        work_free(&w);
    }
}

void orc_spat_to_shqst(void *vh, int nplanes, const real *vr, const real *vt, const real *vp,
                       real *q, real *s, real *t)
{
    orc_sht *h = vh;
    long np = (long)h->nlat * h->nphi;
    #pragma omp parallel
    {
        sht_work w = work_alloc(h);
        #pragma omp for schedule(dynamic)
        for (int k = 0; k < nplanes; k++)
            spat_to_shqst_plane(h, vr + k * np, vt + k * np, vp + k * np,
                                (cplx *)q + k * h->nlm, (cplx *)s + k * h->nlm,
                                (cplx *)t + k * h->nlm, &w);
        work_free(&w);
    }
}

/* ------------------------------------------------------------------ */
/* Radial (z) transforms with FC-Gram continuation                      */
/* ------------------------------------------------------------------ */
typedef struct {
    int n, C, d;
    real *dir; /* C x d, column major: dir[(ii-1) + C*(jj-1)] */
    orc_fft *fz;
} orc_fc;

/* dir = A . Q^T with A (C x d) and Q (d x d) column major (fcgram_mod.f90:257-258). */
void orc_fc_dir(int C, int d, const real *A, const real *Q, real *dir)
{
    for (int i = 0; i < C; i++)
        for (int j = 0; j < d; j++) {
            real s = 0.0;
            for (int k = 0; k < d; k++) s += A[i + (long)C * k] * Q[j + (long)d * k];
            dir[i + (long)C * j] = s;
        }
}

void *orc_fc_create(int n, int C, int d, const real *dir)
{
    if (n < 1 || C < 0 || (C > 0 && (d < 1 || n < C + 2 * d))) return NULL;
    orc_fc *p = calloc(1, sizeof(*p));
    p->n = n;
    p->C = C;
    p->d = d;
    if (C > 0) {
        p->dir = malloc(sizeof(real) * C * d);
        memcpy(p->dir, dir, sizeof(real) * C * d);
    }
    p->fz = fft_create(n);
    return p;
}

void orc_fc_destroy(void *vp)
{
    orc_fc *p = vp;
    if (!p) return;
    free(p->dir);
    fft_destroy(p->fz);
    free(p);
}

/* Continuation rows n-C+1..n (1-based), sphp_sca.F90:369-381. */
static void fc_continue_col(const orc_fc *p, cplx *col, long ld)
{
    int n = p->n, C = p->C, d = p->d;
    for (int ii = 1; ii <= C; ii++) {
        cplx s = p->dir[(ii - 1) + (long)C * 0] * col[(long)(n - C - d) * ld] +
                 p->dir[(C - ii) + (long)C * 0] * col[(long)(d - 1) * ld];
        for (int jj = 2; jj <= d; jj++)
            s += p->dir[(ii - 1) + (long)C * (jj - 1)] * col[(long)(n - C - d + jj - 1) * ld] +
                 p->dir[(C - ii) + (long)C * (jj - 1)] * col[(long)(d - jj) * ld];
        col[(long)(n - C + ii - 1) * ld] = s;
    }
}

/* fftp1d_real_to_complex_z: continuation, forward c2c (sign -1), /n; in place. */
void orc_z_forward(void *vp, int ncols, long ld, real *a_)
{
    orc_fc *p = vp;
    cplx *a = (cplx *)a_;
    int n = p->n;
    #pragma omp parallel
    {
        cplx *x = malloc(sizeof(cplx) * n), *y = malloc(sizeof(cplx) * n);
        cplx *tmp = malloc(sizeof(cplx) * n);
        #pragma omp for schedule(static)
        for (int c = 0; c < ncols; c++) {
            if (p->C > 0) fc_continue_col(p, a + c, ld);
            for (int k = 0; k < n; k++) x[k] = a[c + k * ld];
            fft_exec(p->fz, -1, x, 1, y, tmp);
            for (int k = 0; k < n; k++) a[c + k * ld] = y[k] / (real)n;
        }
        free(x); free(y); free(tmp);
    }
}

/* fftp1d_complex_to_real_z: backward c2c (sign +1), unnormalized, in place. */
void orc_z_inverse(void *vp, int ncols, long ld, real *a_)
{
    orc_fc *p = vp;
    cplx *a = (cplx *)a_;
    int n = p->n;
    #pragma omp parallel
    {
        cplx *x = malloc(sizeof(cplx) * n), *y = malloc(sizeof(cplx) * n);
        cplx *tmp = malloc(sizeof(cplx) * n);
        #pragma omp for schedule(static)
        for (int c = 0; c < ncols; c++) {
            for (int k = 0; k < n; k++) x[k] = a[c + k * ld];
            fft_exec(p->fz, +1, x, 1, y, tmp);
            for (int k = 0; k < n; k++) a[c + k * ld] = y[k];
        }
        free(x); free(y); free(tmp);
    }
}

/* kz in FFT order times Dkz = 2 pi / (dz n) (pseudospec_mod.f90:196,223-227).
 * Even n keeps the Nyquist bin at -n/2 as the reference does; for odd n the
 * reference loop leaves index n unset, here the correct ordering
 * 0..(n-1)/2, -(n-1)/2..-1 is used. */
void orc_kz(int n, real dz, real *kz)
{
    real dk = 2.0 * ORC_PI / (dz * n);
    if (n % 2 == 0) {
        for (int k = 0; k < n / 2; k++) {
            kz[k] = k * dk;
            kz[k + n / 2] = (k - n / 2) * dk;
        }
    } else {
        for (int k = 0; k < n; k++) kz[k] = (k <= (n - 1) / 2 ? k : k - n) * dk;
    }
}

/* derivr_k: b = i^p kz^p a (pseudospec_hd.f90:60-98). */
void orc_derivr_k(int n, int ncols, long ld, const real *kz, int p, const real *a_, real *b_)
{
    const cplx *a = (const cplx *)a_;
    cplx *b = (cplx *)b_;
    cplx ip = cpow(I, p);
    if (p == 0) ip = 1; else if (p == 1) ip = I; else if (p == 2) ip = -1; else if (p == 3) ip = -I;
    else if (p == 4) ip = 1;
    #pragma omp parallel for schedule(static)
    for (int k = 0; k < n; k++) {
        real kp = 1.0;
        for (int e = 0; e < p; e++) kp *= kz[k];
        for (int c = 0; c < ncols; c++) b[c + k * ld] = ip * kp * a[c + k * ld];
    }
}

/* derivr_r: copy, forward z, derivr_k, inverse z (pseudospec_hd.f90:101-135). */
void orc_derivr_r(void *vp, const real *kz, int p, int ncols, long ld, const real *a, real *b)
{
    orc_fc *fp = vp;
    int n = fp->n;
    size_t nb = sizeof(cplx) * (size_t)ld * n;
    real *c1 = malloc(nb);
    memcpy(c1, a, nb);
    orc_z_forward(vp, ncols, ld, c1);
    orc_derivr_k(n, ncols, ld, kz, p, c1, b);
    orc_z_inverse(vp, ncols, ld, b);
    free(c1);
}

/* laplang: b = -l(l+1) a (pseudospec_hd.f90:138-172). */
void orc_laplang(int n, int ncols, long ld, const int *ell, const real *a_, real *b_)
{
    const cplx *a = (const cplx *)a_;
    cplx *b = (cplx *)b_;
    for (int k = 0; k < n; k++)
        for (int c = 0; c < ncols; c++)
            b[c + k * ld] = -(real)ell[c] * (ell[c] + 1) * a[c + k * ld];
}

/* multr: b = z_k^p a (pseudospec_hd.f90:175-208). */
void orc_multr(int n, int ncols, long ld, const real *z, int p, const real *a_, real *b_)
{
    const cplx *a = (const cplx *)a_;
    cplx *b = (cplx *)b_;
    for (int k = 0; k < n; k++) {
        real zp = 1.0;
        for (int e = 0; e < abs(p); e++) zp *= z[k];
        if (p < 0) zp = 1.0 / zp;
        for (int c = 0; c < ncols; c++) b[c + k * ld] = zp * a[c + k * ld];
    }
}

/* saxpby_c: zz = a x + b y, complex (pseudospec_hd.f90:291-326). */
void orc_saxpby_c(long npts, real *zz, const real *x, real a, const real *y, real b)
{
    for (long i = 0; i < 2 * npts; i++) zz[i] = a * x[i] + b * y[i];
}

/* dlr: b = d2a/dr2 - l(l+1) a / r^2 (pseudospec_hd.f90:254-288). */
void orc_dlr(void *vp, const real *kz, const real *z, const int *ell, int ncols, long ld,
             const real *a, real *b)
{
    orc_fc *fp = vp;
    int n = fp->n;
    size_t nb = sizeof(cplx) * (size_t)ld * n;
    real *c1 = malloc(nb), *c2 = malloc(nb);
    memset(c1, 0, nb);
    memset(c2, 0, nb);
    orc_derivr_r(vp, kz, 2, ncols, ld, a, b);
    orc_laplang(n, ncols, ld, ell, a, c1);
    orc_multr(n, ncols, ld, z, -2, c1, c2);
    orc_saxpby_c((long)ld * n, b, b, 1.0, c2, 1.0);
    free(c1);
    free(c2);
}

/* prodvec: C = A x B in (r,theta,phi) on npts points (pseudospec_hd.f90:211-251). */
void orc_prodvec(long npts, const real *ar, const real *at, const real *ap,
                 const real *br, const real *bt, const real *bp,
                 real *cr, real *ct, real *cp)
{
    #pragma omp parallel for schedule(static)
    for (long i = 0; i < npts; i++) {
        real r = at[i] * bp[i] - ap[i] * bt[i];
        real t = ap[i] * br[i] - ar[i] * bp[i];
        real p = ar[i] * bt[i] - at[i] * br[i];
        cr[i] = r;
        ct[i] = t;
        cp[i] = p;
    }
}

/* fc_filter_k: a *= exp(-alpha (2 kz/(n Dkz))^(2p)), alpha = 16 ln 10, p = 50
 * (pseudospec_hd.f90:405-437). */
void orc_fc_filter_k(int n, int ncols, long ld, const real *kz, real dkz, real *a_)
{
    cplx *a = (cplx *)a_;
    real alpha = 16.0 * log(10.0);
    for (int k = 0; k < n; k++) {
        real x = 2.0 * kz[k] / n / dkz;
        real f = exp(-alpha * pow(x, 100.0));
        for (int c = 0; c < ncols; c++) a[c + k * ld] *= f;
    }
}

/* Neumann reconstructor of order ord (1 or 2) from Q (d x d) and Qn (d x d),
 * both column major, read with the table spacing dxp:
 * neu = Q(d,:) Qn^T, neu(d) *= (dx / dxp)^ord (fcgram_mod.f90:355-363). */
void orc_neumann_table(int d, int ord, real dx, const real *Q, const real *Qn, real dxp, real *neu)
{
    for (int j = 0; j < d; j++) {
        real s = 0.0;
        for (int k = 0; k < d; k++) s += Q[(d - 1) + (long)d * k] * Qn[j + (long)d * k];
        neu[j] = s;
    }
    const real f = dx / dxp;
    neu[d - 1] *= (ord == 2) ? f * f : f;
}

/* neumann_reconstruct / robin_reconstruct along z (boundaries 5, z(1), and
 * 6, z(nz - Cz); fcgram_mod.f90:466-517, 612-650) on (ncols, n) complex
 * arrays with row stride ld, z being the row index as in the spectral (lm, nz)
 * arrays of the path.  The boundary row holds the normal derivative g on
 * entry and the reconstructed value on exit:
 *   f_b = neu(d) g + sum_{k=1}^{d-1} neu(k) f(row_k),
 * rows d+1-k (boundary 5) or nz-Cz-d+k (boundary 6), 1-based; with coef
 * (one real per column, Robin f' + a f = g) f_b is further divided by
 * a neu(d) + 1.  Returns 0, or -1 for another boundary. */
int orc_fc_reconstruct(int n, int C, int d, int ncols, long ld, int boun, const real *neu, const real *coef,
                       real *a_)
{
    cplx *a = (cplx *)a_;
    if (boun != 5 && boun != 6) return -1;
    const int rb = (boun == 5) ? 0 : n - C - 1;
    for (int c = 0; c < ncols; c++) {
        cplx v = neu[d - 1] * a[c + (long)rb * ld];
        for (int k = 1; k <= d - 1; k++) {
            const int row = (boun == 5) ? d - k : n - C - 1 - d + k;
            v += neu[k - 1] * a[c + (long)row * ld];
        }
        if (coef) v = v / (coef[c] * neu[d - 1] + 1);
        a[c + (long)rb * ld] = v;
    }
    return 0;
}

/* ------------------------------------------------------------------ */
/* Nonlinear term u x omega, composed from the reference operators      */
/* ------------------------------------------------------------------ */
/*
 * u = curl curl (W r) + curl (Z r), r the unit radial vector.  In SHTns
 * (Q,S,T):  u:     Q = l(l+1) W / r^2,  S = dW/dr / r,  T = Z / r
 *           omega: Q = l(l+1) Z / r^2,  S = dZ/dr / r,  T = -D_l W / r
 * The radial derivatives are derivr (p = 1) and dlr, the syntheses are
 * fftp2d_qst_to_vec, the product is prodvec on the physical planes and the
 * analysis is fftp2d_vec_to_qst.
 *
 * W, Z, NQ, NS, NT are (nlm, n) complex; rows 0..nr-1 (nr = n - C) are the
 * physical shells.  W and Z are not modified (the radial transforms work on
 * copies).  Rows nr..n-1 of the outputs are set to zero.
 */
void orc_nonlinear_uxw(void *vsht, void *vfc, real rint, real dz,
                       const real *W, const real *Z, real *NQ, real *NS, real *NT,
                       int plane0, int nplanes_eval)
{
    orc_sht *h = vsht;
    orc_fc *fp = vfc;
    int n = fp->n, nr = n - fp->C, lmax = h->lmax;
    long nlm = h->nlm;
    size_t nb = sizeof(cplx) * (size_t)nlm * n;
    real *kz = malloc(sizeof(real) * n), *z = malloc(sizeof(real) * n);
    int *ell = malloc(sizeof(int) * nlm);
    orc_kz(n, dz, kz);
    for (int k = 0; k < n; k++) z[k] = rint + dz * k;
    for (int m = 0; m <= lmax; m++)
        for (int l = m; l <= lmax; l++) ell[lm_index(lmax, l, m)] = l;

    real *dW = malloc(nb), *dlW = malloc(nb), *dZ = malloc(nb);
    orc_derivr_r(fp, kz, 1, (int)nlm, nlm, W, dW);
    orc_dlr(fp, kz, z, ell, (int)nlm, nlm, W, dlW);
    orc_derivr_r(fp, kz, 1, (int)nlm, nlm, Z, dZ);

    cplx *uq = malloc(nb), *us = malloc(nb), *ut = malloc(nb);
    cplx *wq = malloc(nb), *ws = malloc(nb), *wt = malloc(nb);
    const cplx *cW = (const cplx *)W, *cZ = (const cplx *)Z;
    for (int k = 0; k < nr; k++) {
        real r1 = 1.0 / z[k], r2 = 1.0 / (z[k] * z[k]);
        for (long c = 0; c < nlm; c++) {
            long i = c + k * nlm;
            real ll = (real)ell[c] * (ell[c] + 1);
            uq[i] = ll * cW[i] * r2;
            us[i] = ((cplx *)dW)[i] * r1;
            ut[i] = cZ[i] * r1;
            wq[i] = ll * cZ[i] * r2;
            ws[i] = ((cplx *)dZ)[i] * r1;
            wt[i] = -((cplx *)dlW)[i] * r1;
        }
    }
    memset(NQ, 0, nb);
    memset(NS, 0, nb);
    memset(NT, 0, nb);
    int k0 = plane0, k1 = plane0 + nplanes_eval;
    if (k1 > nr) k1 = nr;
    long np = (long)h->nlat * h->nphi;
    #pragma omp parallel
    {
        sht_work w = work_alloc(h);
        real *g = malloc(sizeof(real) * np * 9);
        #pragma omp for schedule(dynamic)
        for (int k = k0; k < k1; k++) {
            long o = k * nlm;
            shqst_to_spat_plane(h, uq + o, us + o, ut + o, g, g + np, g + 2 * np, &w);
            shqst_to_spat_plane(h, wq + o, ws + o, wt + o, g + 3 * np, g + 4 * np, g + 5 * np, &w);
            orc_prodvec(np, g, g + np, g + 2 * np, g + 3 * np, g + 4 * np, g + 5 * np,
                        g + 6 * np, g + 7 * np, g + 8 * np);
            spat_to_shqst_plane(h, g + 6 * np, g + 7 * np, g + 8 * np, (cplx *)NQ + o,
                                (cplx *)NS + o, (cplx *)NT + o, &w);
        }
        free(g);
        work_free(&w);
    }
    free(uq); free(us); free(ut); free(wq); free(ws); free(wt);
    free(dW); free(dlW); free(dZ);
    free(kz); free(z); free(ell);
}

/* (Q,S,T) of v = curl curl (P r) + curl (Q r) and of curl v, as in
 * orc_nonlinear_uxw, on rows 0..nr-1 of freshly allocated (nlm, n) arrays
 * (rows nr..n-1 zero): vq, vs, vt, cq, cs, ct. */
static void curl_qst(orc_fc *fp, const real *kz, const real *z, const int *ell, long nlm,
                     const real *P, const real *Q, cplx *out[6])
{
    int n = fp->n, nr = n - fp->C;
    size_t nb = sizeof(cplx) * (size_t)nlm * n;
    real *dP = malloc(nb), *dlP = malloc(nb), *dQ = malloc(nb);
    orc_derivr_r(fp, kz, 1, (int)nlm, nlm, P, dP);
    orc_dlr(fp, kz, z, ell, (int)nlm, nlm, P, dlP);
    orc_derivr_r(fp, kz, 1, (int)nlm, nlm, Q, dQ);
    for (int i = 0; i < 6; i++) out[i] = calloc(1, nb);
    const cplx *cP = (const cplx *)P, *cQ = (const cplx *)Q;
    for (int k = 0; k < nr; k++) {
        real r1 = 1.0 / z[k], r2 = 1.0 / (z[k] * z[k]);
        for (long c = 0; c < nlm; c++) {
            long i = c + k * nlm;
            real ll = (real)ell[c] * (ell[c] + 1);
            out[0][i] = ll * cP[i] * r2;
            out[1][i] = ((cplx *)dP)[i] * r1;
            out[2][i] = cQ[i] * r1;
            out[3][i] = ll * cQ[i] * r2;
            out[4][i] = ((cplx *)dQ)[i] * r1;
            out[5][i] = -((cplx *)dlP)[i] * r1;
        }
    }
    free(dP); free(dlP); free(dQ);
}

/*
 * One nonlinear step of the coupled state (velocity W, Z; magnetic field
 * Pb, Tb with B = curl curl (Pb r) + curl (Tb r); temperature T):
 *   (NQ,NS,NT) = u x omega, (MQ,MS,MT) = u x B, NH = u . grad T,
 * with grad T = (dT/dr, (1/r) dT/dtheta, (1/(r sin)) dT/dphi), i.e. the
 * (Q,S,T) vector (derivr(T), T/r, 0).  Same composition and layouts as
 * orc_nonlinear_uxw; the scalar analysis is fftp2d_real_to_complex_xy.
 */
void orc_nonlinear_coupled(void *vsht, void *vfc, real rint, real dz,
                           const real *W, const real *Z, const real *Pb, const real *Tb,
                           const real *T, real *NQ, real *NS, real *NT, real *MQ, real *MS,
                           real *MT, real *NH, int plane0, int nplanes_eval)
{
    orc_sht *h = vsht;
    orc_fc *fp = vfc;
    int n = fp->n, nr = n - fp->C, lmax = h->lmax;
    long nlm = h->nlm;
    size_t nb = sizeof(cplx) * (size_t)nlm * n;
    real *kz = malloc(sizeof(real) * n), *z = malloc(sizeof(real) * n);
    int *ell = malloc(sizeof(int) * nlm);
    orc_kz(n, dz, kz);
    for (int k = 0; k < n; k++) z[k] = rint + dz * k;
    for (int m = 0; m <= lmax; m++)
        for (int l = m; l <= lmax; l++) ell[lm_index(lmax, l, m)] = l;

    cplx *uw[6], *bj[6];
    curl_qst(fp, kz, z, ell, nlm, W, Z, uw);
    curl_qst(fp, kz, z, ell, nlm, Pb, Tb, bj);
    cplx *gq = malloc(nb), *gs = calloc(1, nb), *gt = calloc(1, nb);
    orc_derivr_r(fp, kz, 1, (int)nlm, nlm, T, (real *)gq);
    for (int k = 0; k < nr; k++)
        for (long c = 0; c < nlm; c++) gs[c + k * nlm] = ((const cplx *)T)[c + k * nlm] / z[k];

    real *outs[7] = {NQ, NS, NT, MQ, MS, MT, NH};
    for (int i = 0; i < 7; i++) memset(outs[i], 0, nb);
    int k0 = plane0, k1 = plane0 + nplanes_eval;
    if (k1 > nr) k1 = nr;
    long np = (long)h->nlat * h->nphi;
    #pragma omp parallel
    {
        sht_work w = work_alloc(h);
        real *g = malloc(sizeof(real) * np * 16);
        real *u = g, *om = g + 3 * np, *b = g + 6 * np, *gr = g + 9 * np, *c = g + 12 * np,
               *d = g + 15 * np;
        #pragma omp for schedule(dynamic)
        for (int k = k0; k < k1; k++) {
            long o = k * nlm;
            shqst_to_spat_plane(h, uw[0] + o, uw[1] + o, uw[2] + o, u, u + np, u + 2 * np, &w);
            shqst_to_spat_plane(h, uw[3] + o, uw[4] + o, uw[5] + o, om, om + np, om + 2 * np, &w);
            shqst_to_spat_plane(h, bj[0] + o, bj[1] + o, bj[2] + o, b, b + np, b + 2 * np, &w);
            shqst_to_spat_plane(h, gq + o, gs + o, gt + o, gr, gr + np, gr + 2 * np, &w);
            orc_prodvec(np, u, u + np, u + 2 * np, om, om + np, om + 2 * np, c, c + np, c + 2 * np);
            spat_to_shqst_plane(h, c, c + np, c + 2 * np, (cplx *)NQ + o, (cplx *)NS + o, (cplx *)NT + o, &w);
            orc_prodvec(np, u, u + np, u + 2 * np, b, b + np, b + 2 * np, c, c + np, c + 2 * np);
            spat_to_shqst_plane(h, c, c + np, c + 2 * np, (cplx *)MQ + o, (cplx *)MS + o, (cplx *)MT + o, &w);
            for (long i = 0; i < np; i++)
                d[i] = u[i] * gr[i] + u[np + i] * gr[np + i] + u[2 * np + i] * gr[2 * np + i];
            spat_to_sh_plane(h, d, (cplx *)NH + o, &w);
        }
        free(g);
        work_free(&w);
    }
    for (int i = 0; i < 6; i++) { free(uw[i]); free(bj[i]); }
    free(gq); free(gs); free(gt);
    free(kz); free(z); free(ell);
}

/* ------------------------------------------------------------------ */
/* Boundary: FC-Gram derivative matrices and the per-l W recovery       */
/* ------------------------------------------------------------------ */

/* derivr_mat (boundary_mod.f90:96-137): d[i + N*j] = d^p/dr^p of e_j at r_i,
 * on the N = n - C physical points. */
void orc_derivr_mat(void *vp, const real *kz, int p, real *dmat)
{
    orc_fc *fp = vp;
    int n = fp->n, N = n - fp->C;
    size_t nb = sizeof(cplx) * (size_t)N * n;
    real *a = calloc(1, nb), *b = calloc(1, nb);
    for (int j = 0; j < N; j++) a[2 * (j + (long)N * j)] = 1.0;
    orc_derivr_r(vp, kz, p, N, N, a, b);
    for (int j = 0; j < N; j++)
        for (int i = 0; i < N; i++) dmat[i + (long)N * j] = b[2 * (j + (long)N * i)];
    free(a);
    free(b);
}

/* Dense LU with partial pivoting, column major; returns 0 on success. */
static int lu_factor(int n, real *A, int *piv)
{
    for (int k = 0; k < n; k++) {
        int pk = k;
        real mx = fabs(A[k + (long)n * k]);
        for (int i = k + 1; i < n; i++)
            if (fabs(A[i + (long)n * k]) > mx) { mx = fabs(A[i + (long)n * k]); pk = i; }
        piv[k] = pk;
        if (mx == 0.0) return k + 1;
        if (pk != k)
            for (int j = 0; j < n; j++) {
                real t = A[k + (long)n * j];
                A[k + (long)n * j] = A[pk + (long)n * j];
                A[pk + (long)n * j] = t;
            }
        real inv = 1.0 / A[k + (long)n * k];
        for (int i = k + 1; i < n; i++) A[i + (long)n * k] *= inv;
        for (int j = k + 1; j < n; j++) {
            real akj = A[k + (long)n * j];
            if (akj != 0.0)
                for (int i = k + 1; i < n; i++) A[i + (long)n * j] -= A[i + (long)n * k] * akj;
        }
    }
    return 0;
}

static void lu_solve(int n, const real *LU, const int *piv, real *b)
{
    for (int k = 0; k < n; k++) {
        int pk = piv[k];
        if (pk != k) { real t = b[k]; b[k] = b[pk]; b[pk] = t; }
    }
    for (int j = 0; j < n; j++)
        for (int i = j + 1; i < n; i++) b[i] -= LU[i + (long)n * j] * b[j];
    for (int j = n - 1; j >= 0; j--) {
        b[j] /= LU[j + (long)n * j];
        for (int i = 0; i < j; i++) b[i] -= LU[i + (long)n * j] * b[j];
    }
}

/* invDl_setup (boundary_mod.f90:140-241): binv[(i) + N*(j) + N*(N-4)*(l-1)],
 * M_l = columns 3..N-2 of B_l^{-1}; returns 0 or the failing l. */
int orc_invdl_setup(int N, int lmax, const real *d1, const real *d2, const real *z,
                    real *binv)
{
    real *B = malloc(sizeof(real) * N * N), *esc = malloc(sizeof(real) * N);
    real *col = malloc(sizeof(real) * N);
    int *piv = malloc(sizeof(int) * N);
    int fail = 0;
    for (int l = 1; l <= lmax && !fail; l++) {
        memset(B, 0, sizeof(real) * N * N);
        B[0] = 1.0;
        for (int j = 0; j < N; j++) B[1 + (long)N * j] = d1[0 + (long)N * j];
        for (int i = 2; i < N - 2; i++) {
            for (int j = 0; j < N; j++) B[i + (long)N * j] = d2[i + (long)N * j];
            B[i + (long)N * i] -= (real)l * (l + 1) / (z[i] * z[i]);
        }
        for (int j = 0; j < N; j++) B[N - 2 + (long)N * j] = d1[N - 1 + (long)N * j];
        B[N - 1 + (long)N * (N - 1)] = 1.0;
        for (int i = 0; i < N; i++) {
            real mx = 0.0;
            for (int j = 0; j < N; j++) mx = fmax(mx, fabs(B[i + (long)N * j]));
            esc[i] = mx;
            for (int j = 0; j < N; j++) B[i + (long)N * j] /= mx;
        }
        if (lu_factor(N, B, piv)) { fail = l; break; }
        for (int j = 0; j < N - 4; j++) {
            memset(col, 0, sizeof(real) * N);
            col[j + 2] = 1.0 / esc[j + 2];
            lu_solve(N, B, piv, col);
            memcpy(binv + (long)N * j + (long)N * (N - 4) * (l - 1), col, sizeof(real) * N);
        }
    }
    free(B); free(esc); free(col); free(piv);
    return fail;
}

/* invDl (boundary_mod.f90:244-292): W = M_l phi(3:N-2) per lm, l = 0 -> 0;
 * rows N..n-1 of b are zero. */
void orc_invdl(int N, int n, int lmax, const real *binv, const real *a_, real *b_)
{
    const cplx *a = (const cplx *)a_;
    cplx *b = (cplx *)b_;
    long nlm = (long)(lmax + 1) * (lmax + 2) / 2;
    memset(b, 0, sizeof(cplx) * nlm * n);
    #pragma omp parallel for schedule(dynamic)
    for (int m = 0; m <= lmax; m++)
        for (int l = (m > 1 ? m : 1); l <= lmax; l++) {
            long lm = lm_index(lmax, l, m);
            const real *M = binv + (long)N * (N - 4) * (l - 1);
            for (int i = 0; i < N; i++) {
                cplx s = 0;
                for (int j = 0; j < N - 4; j++) s += M[i + (long)N * j] * a[lm + nlm * (j + 2)];
                b[lm + nlm * i] = s;
            }
        }
}
