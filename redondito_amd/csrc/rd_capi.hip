// C-ABI entry points (include/redondito.h): plan life cycle and the
// reference's sphp / pseudo / boundary routines on device pointers.
#include "rd_internal.h"

#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

using namespace rd;

namespace rd {
// Smallest primitive root of a prime n: g^((n-1)/f) != 1 for every prime f | n-1.
int primitive_root(int n)
{
    std::vector<int> fs;
    for (int m = n - 1, f = 2; m > 1; f++)
        if (m % f == 0) {
            fs.push_back(f);
            while (m % f == 0) m /= f;
        }
    auto powmod = [n](long b, long e) {
        long r = 1;
        b %= n;
        for (; e; e >>= 1, b = b * b % n)
            if (e & 1) r = r * b % n;
        return r;
    };
    for (int g = 2; g < n; g++) {
        bool ok = true;
        for (int f : fs) ok = ok && powmod(g, (n - 1) / f) != 1;
        if (ok) return g;
    }
    return 0;
}

// One evaluation of u x omega on the plan's stream: the radial step (which
// also clears the output rows nr..nz-1), the u and omega syntheses, the phi
// transforms with the cross product, the analysis.  ev: ST_COUNT + 1 events
// recorded around the stages, or null.
int uxw_evaluate(Plan *p, int prec, const double2 *W, const double2 *Z, double2 *NQ, double2 *NS, double2 *NT,
                 hipEvent_t *ev)
{
    const int nr = p->nr;
#define MARK(i) \
    if (ev) RD_HIP(hipEventRecord(ev[i], p->stream))
    double2 *outs[3] = {NQ, NS, NT};
    MARK(0);
    int rc = launch_uxw_radial(p, W, Z, p->S, 6, outs);
    MARK(1);
    if (!rc) {
        const double2 *su[3] = {p->S[0], p->S[1], p->S[2]}, *sw[3] = {p->S[3], p->S[4], p->S[5]};
        double2 *fu[3] = {p->F[0], p->F[1], p->F[2]}, *fw[3] = {p->F[3], p->F[4], p->F[5]};
        rc = launch_synthesis2(p, prec, nr, su, fu, sw, fw, p->nlm);
    }
    MARK(2);
    if (!rc) {
        const double2 *Fin[6] = {p->F[0], p->F[1], p->F[2], p->F[3], p->F[4], p->F[5]};
        double2 *Fout[3] = {p->F[0], p->F[1], p->F[2]};
        rc = launch_phi_cross(p, nr, Fin, Fout, prec);
    }
    MARK(3);
    if (!rc) rc = launch_analysis(p, prec, true, nr, p->F[0], p->F[1], p->F[2], NQ, NS, NT, p->nlm);
    MARK(4);
#undef MARK
    return rc;
}
}  // namespace rd

namespace {

inline bool bad_plan(const rd_plan *p) { return p == nullptr; }

Plan *P(rd_plan *p) { return reinterpret_cast<Plan *>(p); }
const Plan *P(const rd_plan *p) { return reinterpret_cast<const Plan *>(p); }

int upload_twiddles(Plan *p, int n, double2 **dst)
{
    std::vector<double2> h(n);
    for (int t = 0; t < n; t++) {
        const double a = 2.0 * M_PI * (double)t / (double)n;
        h[t] = make_double2(cos(a), -sin(a));
    }
    RD_HIP(hipMalloc(dst, sizeof(double2) * n));
    RD_HIP(hipMemcpyAsync(*dst, h.data(), sizeof(double2) * n, hipMemcpyHostToDevice, p->stream));
    return RD_OK;
}

bool is_prime(int n)
{
    if (n < 2) return false;
    for (int f = 2; (long)f * f <= n; f++)
        if (n % f == 0) return false;
    return true;
}

int largest_prime_factor(int n)
{
    int lp = 1;
    for (int f = 2; n > 1; f++)
        while (n % f == 0) {
            lp = f;
            n /= f;
        }
    return lp;
}

// The radial FFT of length nz: direct Stockham stages, or the Rader form for
// a prime nz > 19 whose nz - 1 is 13-smooth (the direct stage of a large
// prime radix is O(n^2) per sequence).  The Rader spectra are computed in
// long double by a direct DFT and rounded once.
int setup_zfft(Plan *p)
{
    const int n = p->nz;
    ZFft &Z = p->zfft;
    Z = ZFft{};
    Z.n = n;
    Z.rader = is_prime(n) && n > 19 && largest_prime_factor(n - 1) <= 13;
    if (!Z.rader) {
        Z.seq = make_fft_seq(n);
        return upload_twiddles(p, n, &p->tw_z);
    }
    const int n1 = n - 1, g = primitive_root(n);
    std::vector<int> perm(2 * n1);
    long gp = 1;
    for (int q = 0; q < n1; q++, gp = gp * g % n) perm[q] = (int)gp;            // g^q
    for (int q = 0; q < n1; q++) perm[n1 + q] = perm[(n1 - q) % n1];              // g^-q
    std::vector<double2> cf(2 * n1);
    const long double tp = 2.0L * 3.14159265358979323846264338327950288L;
    for (int si = 0; si < 2; si++) {
        const long double sg = si == 0 ? -1.0L : 1.0L;
        for (int k = 0; k < n1; k++) {
            long double re = 0.0L, im = 0.0L;
            for (int m = 0; m < n1; m++) {
                // c_m e^{-2 pi i m k / n1}, c_m = e^{s 2 pi i g^-m / n}
                const long double a = sg * tp * (long double)perm[n1 + m] / n
                                      - tp * (long double)(((long)m * k) % n1) / n1;
                re += cosl(a);
                im += sinl(a);
            }
            cf[si * n1 + k] = make_double2((double)(re / n1), (double)(im / n1));
        }
    }
    RD_HIP(hipMalloc(&p->z_perm, sizeof(int) * 2 * n1));
    RD_HIP(hipMalloc(&p->z_cf, sizeof(double2) * 2 * n1));
    RD_HIP(hipMemcpyAsync(p->z_perm, perm.data(), sizeof(int) * 2 * n1, hipMemcpyHostToDevice, p->stream));
    RD_HIP(hipMemcpyAsync(p->z_cf, cf.data(), sizeof(double2) * 2 * n1, hipMemcpyHostToDevice, p->stream));
    RD_HIP(hipStreamSynchronize(p->stream));
    Z.seq = make_fft_seq(n1);
    Z.pin = p->z_perm;
    Z.pout = p->z_perm + n1;
    Z.cf[0] = p->z_cf;
    Z.cf[1] = p->z_cf + n1;
    return upload_twiddles(p, n1, &p->tw_z);
}

void drop_plan_graph(Plan *p)
{
    if (p->gexec) (void)hipGraphExecDestroy(p->gexec);
    if (p->graph) (void)hipGraphDestroy(p->graph);
    p->gexec = nullptr;
    p->graph = nullptr;
}

void free_plan(Plan *p)
{
    if (!p) return;
    drop_plan_graph(p);
    free_bc(p);
    free_precision_tables(p);
    free_oz_tables(p);
    void *dev[] = {p->d_ct, p->d_st, p->d_wt, p->d_z, p->d_kz, p->d_dir, p->d_ell,
                   p->syn_P, p->syn_D, p->syn_M, p->ana_P, p->ana_D, p->ana_M, p->d_syn_off,
                   p->d_ana_off, p->d_ana_ncol, p->d_syn_q, p->d_ana_q, p->tw_phi, p->tw_z,
                   p->z_perm, p->z_cf};
    for (void *d : dev)
        if (d) (void)hipFree(d);
    for (int i = 0; i < 6; i++) {
        if (p->F[i]) (void)hipFree(p->F[i]);
        if (p->S[i]) (void)hipFree(p->S[i]);
    }
    if (p->ev) {
        for (int i = 0; i < kProfSlots * (ST_COUNT + 1); i++) (void)hipEventDestroy(p->ev[i]);
        free(p->ev);
        free(p->slot_used);
    }
    if (p->own_stream) (void)hipStreamDestroy(p->own_stream);
    free(p->h_ct);
    free(p->h_st);
    free(p->h_wt);
    free(p->h_z);
    free(p->h_kz);
    free(p->h_ell);
    free(p->h_emm);
    delete p;
}

int synth_scalar(Plan *p, int nplanes, const double *in, double *out)
{
    int rc = launch_synthesis(p, RD_PREC_FP64, false, nplanes, (const double2 *)in, nullptr, nullptr, p->nlm, p->F[0],
                              nullptr, nullptr);
    if (rc) return rc;
    const double2 *F[1] = {p->F[0]};
    double *f[1] = {out};
    return launch_phi_inverse(p, 1, nplanes, F, f);
}

int anal_scalar(Plan *p, int nplanes, const double *in, double *out)
{
    const double *f[1] = {in};
    double2 *F[1] = {p->F[0]};
    int rc = launch_phi_forward(p, 1, nplanes, f, F);
    if (rc) return rc;
    return launch_analysis(p, RD_PREC_FP64, false, nplanes, p->F[0], nullptr, nullptr, (double2 *)out, nullptr,
                           nullptr, p->nlm);
}

}  // namespace

extern "C" {

const char *rd_error_string(int code)
{
    switch (code) {
    case RD_OK: return "ok";
    case RD_ERR_ARG: return "bad argument or size";
    case RD_ERR_HIP: return "HIP runtime error";
    case RD_ERR_TABLE: return "FC-Gram table missing or of the wrong size";
    case RD_ERR_NOMEM: return "device allocation failed";
    case RD_ERR_LU: return "singular boundary matrix";
    case RD_ERR_STATE: return "required setup not done";
    default: return "unknown error";
    }
}

int rd_load_dirichlet_tables(const char *tdir, int C, int d, double *dir)
{
    if (!tdir || C <= 0 || d <= 0 || !dir) return RD_ERR_ARG;
    char pa[1024], pq[1024];
    snprintf(pa, sizeof pa, "%s/A%d-%d.dat", tdir, C, d);
    snprintf(pq, sizeof pq, "%s/Q%d.dat", tdir, d);
    std::vector<double> A((size_t)C * d), Q((size_t)d * d);
    FILE *fa = fopen(pa, "rb");
    if (!fa) {
        fprintf(stderr, "redondito: could not find table %s\n", pa);
        return RD_ERR_TABLE;
    }
    size_t na = fread(A.data(), sizeof(double), A.size(), fa);
    int extra = fgetc(fa);
    fclose(fa);
    FILE *fq = fopen(pq, "rb");
    if (!fq) {
        fprintf(stderr, "redondito: could not find table %s\n", pq);
        return RD_ERR_TABLE;
    }
    size_t nq = fread(Q.data(), sizeof(double), Q.size(), fq);
    int extraq = fgetc(fq);
    fclose(fq);
    if (na != A.size() || nq != Q.size() || extra != EOF || extraq != EOF) {
        fprintf(stderr, "redondito: FC-Gram table of the wrong size (%s, %s)\n", pa, pq);
        return RD_ERR_TABLE;
    }
    // dir = A . Q^T, all column major
    for (int i = 0; i < C; i++)
        for (int j = 0; j < d; j++) {
            double s = 0.0;
            for (int k = 0; k < d; k++) s += A[i + (size_t)C * k] * Q[j + (size_t)d * k];
            dir[i + (size_t)C * j] = s;
        }
    return RD_OK;
}

// Reads a stream FP64 file of exactly n values (after `skip` leading ones
// returned in head).
static int read_table(const char *path, double *head, int skip, double *v, size_t n)
{
    FILE *f = fopen(path, "rb");
    if (!f) {
        fprintf(stderr, "redondito: could not find table %s\n", path);
        return RD_ERR_TABLE;
    }
    size_t nh = skip ? fread(head, sizeof(double), (size_t)skip, f) : 0;
    size_t nv = fread(v, sizeof(double), n, f);
    int extra = fgetc(f);
    fclose(f);
    if (nh != (size_t)skip || nv != n || extra != EOF) {
        fprintf(stderr, "redondito: FC-Gram table of the wrong size (%s)\n", path);
        return RD_ERR_TABLE;
    }
    return RD_OK;
}

int rd_load_neumann_tables(const char *tdir, int d, int ord, double dx, double *neu)
{
    if (!tdir || d <= 0 || !neu || !(dx > 0.0)) return RD_ERR_ARG;
    if (ord != 1 && ord != 2) {
        fprintf(stderr, "[WARNING] load_neu_tables called with undefined order. Skipping...\n");
        return RD_OK;
    }
    char pq[1024], pn[1024];
    snprintf(pq, sizeof pq, "%s/Q%d.dat", tdir, d);
    snprintf(pn, sizeof pn, "%s/Q%dn%d.dat", tdir, ord, d);
    std::vector<double> Q((size_t)d * d), Qn((size_t)d * d);
    double dxp = 0.0;
    int rc = read_table(pq, nullptr, 0, Q.data(), Q.size());
    if (!rc) rc = read_table(pn, &dxp, 1, Qn.data(), Qn.size());
    if (rc) return rc;
    if (!(dxp > 0.0)) {
        fprintf(stderr, "redondito: table spacing %g in %s\n", dxp, pn);
        return RD_ERR_TABLE;
    }
    // neu = Q(d,:) Qn^T, the last weight scaled by (dx/dxp)^ord (column major)
    for (int j = 0; j < d; j++) {
        double s = 0.0;
        for (int k = 0; k < d; k++) s += Q[(d - 1) + (size_t)d * k] * Qn[j + (size_t)d * k];
        neu[j] = s;
    }
    const double f = dx / dxp;
    neu[d - 1] *= (ord == 2) ? f * f : f;
    return RD_OK;
}

// The plan constructor; nwork Fourier and spectral scratch arrays (6 for the
// nonlinear term, 1 for scalar transforms).
static int plan_create_impl(Plan **out, int lmax, int nlat, int nphi, int nz, int C, int d, double dz,
                            double rint, const double *dir, int nwork)
{
    if (!out || nwork < 1 || nwork > 6) return RD_ERR_ARG;
    *out = nullptr;
    if (lmax < 0 || nlat < 2 || (nlat & 1) || nlat <= lmax || nphi <= 2 * lmax || nphi < 2) return RD_ERR_ARG;
    if (nz < 1 || C < 0) return RD_ERR_ARG;
    if (C > 0 && (d < 1 || nz - C < 2 * d || !dir)) return RD_ERR_ARG;
    if (C == 0 && d != 0) return RD_ERR_ARG;
    if (!(dz > 0.0)) return RD_ERR_ARG;

    Plan *p = new Plan();
    memset(p, 0, sizeof(Plan));
    p->oz_digits = 7;
    p->lmax = lmax;
    p->L = lmax + 1;
    p->nlat = nlat;
    p->nphi = nphi;
    p->nlm = (lmax + 1) * (lmax + 2) / 2;
    p->nth = nlat / 2;
    p->nthp = (p->nth + 63) / 64 * 64;
    p->nthp16 = (p->nth + 15) / 16 * 16;
    p->nz = nz;
    p->C = C;
    p->d = d;
    p->nr = nz - C;
    p->dz = dz;
    p->rint = rint;
    p->seq_phi = make_fft_seq(nphi);

    int rc = RD_OK;
#define TRY(x)                  \
    do {                        \
        rc = (x);               \
        if (rc) {               \
            free_plan(p);       \
            return rc;          \
        }                       \
    } while (0)
#define TRYH(x)                      \
    do {                             \
        if ((x) != hipSuccess) {     \
            free_plan(p);            \
            return RD_ERR_HIP;       \
        }                            \
    } while (0)

    TRYH(hipGetDevice(&p->device));
    if (hipDeviceGetAttribute(&p->n_cu, hipDeviceAttributeMultiprocessorCount, p->device) != hipSuccess ||
        p->n_cu < 1)
        p->n_cu = 256;
    TRYH(hipStreamCreateWithFlags(&p->own_stream, hipStreamNonBlocking));
    p->stream = p->own_stream;

    // host grid: z, kz (pseudospec_mod.f90:196-200,223-227), l and m per lm
    p->h_ct = (double *)malloc(sizeof(double) * nlat);
    p->h_st = (double *)malloc(sizeof(double) * nlat);
    p->h_wt = (double *)malloc(sizeof(double) * nlat);
    p->h_z = (double *)malloc(sizeof(double) * nz);
    p->h_kz = (double *)malloc(sizeof(double) * nz);
    p->h_ell = (int *)malloc(sizeof(int) * p->nlm);
    p->h_emm = (int *)malloc(sizeof(int) * p->nlm);
    for (int k = 0; k < nz; k++) p->h_z[k] = rint + dz * k;
    const double dk = 2.0 * M_PI / (dz * nz);
    if (nz % 2 == 0) {
        for (int k = 0; k < nz / 2; k++) {
            p->h_kz[k] = k * dk;
            p->h_kz[k + nz / 2] = (k - nz / 2) * dk;
        }
    } else {
        for (int k = 0; k < nz; k++) p->h_kz[k] = (k <= (nz - 1) / 2 ? k : k - nz) * dk;
    }
    for (int m = 0; m <= lmax; m++)
        for (int l = m; l <= lmax; l++) {
            const long i = (long)m * (2 * lmax + 1 - m) / 2 + l;
            p->h_ell[i] = l;
            p->h_emm[i] = m;
        }

    TRYH(hipMalloc(&p->d_ct, sizeof(double) * nlat));
    TRYH(hipMalloc(&p->d_st, sizeof(double) * nlat));
    TRYH(hipMalloc(&p->d_wt, sizeof(double) * nlat));
    TRYH(hipMalloc(&p->d_z, sizeof(double) * nz));
    TRYH(hipMalloc(&p->d_kz, sizeof(double) * nz));
    TRYH(hipMalloc(&p->d_ell, sizeof(int) * p->nlm));
    TRYH(hipMemcpyAsync(p->d_z, p->h_z, sizeof(double) * nz, hipMemcpyHostToDevice, p->stream));
    TRYH(hipMemcpyAsync(p->d_kz, p->h_kz, sizeof(double) * nz, hipMemcpyHostToDevice, p->stream));
    TRYH(hipMemcpyAsync(p->d_ell, p->h_ell, sizeof(int) * p->nlm, hipMemcpyHostToDevice, p->stream));
    if (C > 0) {
        TRYH(hipMalloc(&p->d_dir, sizeof(double) * C * d));
        TRYH(hipMemcpyAsync(p->d_dir, dir, sizeof(double) * C * d, hipMemcpyHostToDevice, p->stream));
    }
    TRY(upload_twiddles(p, nphi, &p->tw_phi));
    TRY(setup_zfft(p));
    TRY(build_legendre_tables(p));

    p->fsize = (long)nz * p->L * nlat;
    p->ssize = (long)p->nlm * nz;
    p->nwork = nwork;
    for (int i = 0; i < nwork; i++) {
        if (hipMalloc(&p->F[i], sizeof(double2) * p->fsize) != hipSuccess) {
            free_plan(p);
            return RD_ERR_NOMEM;
        }
        if (hipMalloc(&p->S[i], sizeof(double2) * p->ssize) != hipSuccess) {
            free_plan(p);
            return RD_ERR_NOMEM;
        }
    }
    TRYH(hipStreamSynchronize(p->stream));
#undef TRY
#undef TRYH
    *out = p;
    return RD_OK;
}

int rd_plan_create(rd_plan **out, int lmax, int nlat, int nphi, int nz, int C, int d, double dz,
                   double rint, const double *dir)
{
    if (!out) return RD_ERR_ARG;
    Plan *p = nullptr;
    const int rc = plan_create_impl(&p, lmax, nlat, nphi, nz, C, d, dz, rint, dir, 6);
    *out = reinterpret_cast<rd_plan *>(p);
    return rc;
}

int rd_plan_destroy(rd_plan *plan)
{
    if (bad_plan(plan)) return RD_ERR_ARG;
    Plan *p = P(plan);
    (void)hipStreamSynchronize(p->stream);
    free_plan(p);
    return RD_OK;
}

int rd_plan_set_stream(rd_plan *plan, void *s)
{
    if (bad_plan(plan)) return RD_ERR_ARG;
    Plan *p = P(plan);
    p->stream = (hipStream_t)s;   // NULL: the default stream
    return RD_OK;
}

int rd_plan_set_ozaki_digits(rd_plan *plan, int digits)
{
    if (bad_plan(plan)) return RD_ERR_ARG;
    if (digits < 5 || digits > 7) return RD_ERR_ARG;
    P(plan)->oz_digits = digits;
    return RD_OK;
}

int rd_plan_sync(rd_plan *plan)
{
    if (bad_plan(plan)) return RD_ERR_ARG;
    RD_HIP(hipStreamSynchronize(P(plan)->stream));
    return RD_OK;
}

int rd_plan_info(const rd_plan *plan, int *info)
{
    if (bad_plan(plan) || !info) return RD_ERR_ARG;
    const Plan *p = P(plan);
    int v[9] = {p->lmax, p->nlat, p->nphi, p->nlm, p->nz, p->C, p->d, p->nr, 0};
    memcpy(info, v, sizeof v);
    return RD_OK;
}

int rd_plan_grid(const rd_plan *plan, double *ct, double *st, double *wt, double *z, double *kz,
                 int *ell, int *emm)
{
    if (bad_plan(plan)) return RD_ERR_ARG;
    const Plan *p = P(plan);
    if (ct) memcpy(ct, p->h_ct, sizeof(double) * p->nlat);
    if (st) memcpy(st, p->h_st, sizeof(double) * p->nlat);
    if (wt) memcpy(wt, p->h_wt, sizeof(double) * p->nlat);
    if (z) memcpy(z, p->h_z, sizeof(double) * p->nz);
    if (kz) memcpy(kz, p->h_kz, sizeof(double) * p->nz);
    if (ell) memcpy(ell, p->h_ell, sizeof(int) * p->nlm);
    if (emm) memcpy(emm, p->h_emm, sizeof(int) * p->nlm);
    return RD_OK;
}

int rd_sht_analysis(rd_plan *plan, int nplanes, const double *in, double *out)
{
    if (bad_plan(plan) || !in || !out) return RD_ERR_ARG;
    Plan *p = P(plan);
    if (nplanes < 0 || nplanes > p->nz) return RD_ERR_ARG;
    if (nplanes == 0) return RD_OK;
    return anal_scalar(p, nplanes, in, out);
}

int rd_sht_synthesis(rd_plan *plan, int nplanes, const double *in, double *out)
{
    if (bad_plan(plan) || !in || !out) return RD_ERR_ARG;
    Plan *p = P(plan);
    if (nplanes < 0 || nplanes > p->nz) return RD_ERR_ARG;
    if (nplanes == 0) return RD_OK;
    return synth_scalar(p, nplanes, in, out);
}

int rd_vsht_analysis(rd_plan *plan, int nplanes, const double *vr, const double *vt, const double *vp,
                     double *q, double *s, double *t)
{
    if (bad_plan(plan) || !vr || !vt || !vp || !q || !s || !t) return RD_ERR_ARG;
    Plan *p = P(plan);
    if (nplanes < 0 || nplanes > p->nz) return RD_ERR_ARG;
    if (nplanes == 0) return RD_OK;
    const double *f[3] = {vr, vt, vp};
    double2 *F[3] = {p->F[0], p->F[1], p->F[2]};
    int rc = launch_phi_forward(p, 3, nplanes, f, F);
    if (rc) return rc;
    return launch_analysis(p, RD_PREC_FP64, true, nplanes, p->F[0], p->F[1], p->F[2], (double2 *)q, (double2 *)s,
                           (double2 *)t, p->nlm);
}

int rd_vsht_synthesis(rd_plan *plan, int nplanes, const double *q, const double *s, const double *t,
                      double *vr, double *vt, double *vp)
{
    if (bad_plan(plan) || !vr || !vt || !vp || !q || !s || !t) return RD_ERR_ARG;
    Plan *p = P(plan);
    if (nplanes < 0 || nplanes > p->nz) return RD_ERR_ARG;
    if (nplanes == 0) return RD_OK;
    int rc = launch_synthesis(p, RD_PREC_FP64, true, nplanes, (const double2 *)q, (const double2 *)s,
                              (const double2 *)t, p->nlm, p->F[0], p->F[1], p->F[2]);
    if (rc) return rc;
    const double2 *F[3] = {p->F[0], p->F[1], p->F[2]};
    double *f[3] = {vr, vt, vp};
    return launch_phi_inverse(p, 3, nplanes, F, f);
}

int rd_z_forward(rd_plan *plan, double *inout)
{
    if (bad_plan(plan) || !inout) return RD_ERR_ARG;
    return launch_z_forward(P(plan), (double2 *)inout);
}

int rd_z_inverse(rd_plan *plan, double *inout)
{
    if (bad_plan(plan) || !inout) return RD_ERR_ARG;
    return launch_z_inverse(P(plan), (double2 *)inout);
}

int rd_3d_forward(rd_plan *plan, const double *in, double *out)
{
    if (bad_plan(plan) || !in || !out) return RD_ERR_ARG;
    Plan *p = P(plan);
    int rc = anal_scalar(p, p->nz, in, out);
    if (rc) return rc;
    return launch_z_forward(p, (double2 *)out);
}

int rd_3d_inverse(rd_plan *plan, double *in, double *out)
{
    if (bad_plan(plan) || !in || !out) return RD_ERR_ARG;
    Plan *p = P(plan);
    int rc = launch_z_inverse(p, (double2 *)in);
    if (rc) return rc;
    return synth_scalar(p, p->nz, in, out);
}

int rd_redist_k_to_lm(rd_plan *plan, int nplanes, const double *in, double *out)
{
    if (bad_plan(plan) || !in || !out) return RD_ERR_ARG;
    Plan *p = P(plan);
    if (nplanes < 0 || nplanes > p->nz) return RD_ERR_ARG;
    if (in == out || nplanes == 0) return RD_OK;
    RD_HIP(hipMemcpyAsync(out, in, sizeof(double2) * (size_t)p->nlm * nplanes, hipMemcpyDeviceToDevice,
                          p->stream));
    return RD_OK;
}

int rd_redist_lm_to_k(rd_plan *plan, int nplanes, const double *in, double *out)
{
    return rd_redist_k_to_lm(plan, nplanes, in, out);
}

int rd_derivr(rd_plan *plan, int order, char esp, const double *a, double *b)
{
    if (bad_plan(plan) || !a || !b || order < 0) return RD_ERR_ARG;
    Plan *p = P(plan);
    if (esp == 'r') {
        if (a == b) return RD_ERR_ARG;
        return launch_derivr_r(p, order, (const double2 *)a, (double2 *)b);
    }
    if (esp == 'k') return launch_derivr_k(p, order, (const double2 *)a, (double2 *)b);
    fprintf(stderr, "redondito: unknown space %c in derivr\n", esp);
    return RD_ERR_ARG;
}

int rd_laplang(rd_plan *plan, const double *a, double *b)
{
    if (bad_plan(plan) || !a || !b) return RD_ERR_ARG;
    return launch_laplang(P(plan), (const double2 *)a, (double2 *)b);
}

int rd_multr(rd_plan *plan, int pw, const double *a, double *b)
{
    if (bad_plan(plan) || !a || !b) return RD_ERR_ARG;
    return launch_multr(P(plan), pw, (const double2 *)a, (double2 *)b);
}

int rd_dlr(rd_plan *plan, const double *a, double *b)
{
    if (bad_plan(plan) || !a || !b || a == b) return RD_ERR_ARG;
    Plan *p = P(plan);
    if (p->nwork < 2) return RD_ERR_STATE;
    int rc = launch_derivr_r(p, 2, (const double2 *)a, (double2 *)b);
    if (!rc) rc = launch_laplang(p, (const double2 *)a, p->S[0]);
    if (!rc) rc = launch_multr(p, -2, p->S[0], p->S[1]);
    if (!rc) rc = launch_saxpby(p, 2L * p->nlm * p->nz, b, b, 1.0, (const double *)p->S[1], 1.0);
    return rc;
}

int rd_saxpby_c(rd_plan *plan, double *z, const double *x, double a, const double *y, double b)
{
    if (bad_plan(plan) || !z || !x || !y) return RD_ERR_ARG;
    Plan *p = P(plan);
    return launch_saxpby(p, 2L * p->nlm * p->nz, z, x, a, y, b);
}

int rd_saxpby_r(rd_plan *plan, int nplanes, double *z, const double *x, double a, const double *y,
                double b)
{
    if (bad_plan(plan) || !z || !x || !y) return RD_ERR_ARG;
    Plan *p = P(plan);
    if (nplanes < 0 || nplanes > p->nz) return RD_ERR_ARG;
    return launch_saxpby(p, (long)p->nlat * p->nphi * nplanes, z, x, a, y, b);
}

int rd_prodvec(rd_plan *plan, int nplanes, const double *ar, const double *at, const double *ap,
               const double *br, const double *bt, const double *bp, double *cr, double *ct, double *cp)
{
    if (bad_plan(plan) || !ar || !at || !ap || !br || !bt || !bp || !cr || !ct || !cp) return RD_ERR_ARG;
    Plan *p = P(plan);
    if (nplanes < 0 || nplanes > p->nz) return RD_ERR_ARG;
    const int np = nplanes < p->nr ? nplanes : p->nr;
    return launch_prodvec(p, (long)p->nlat * p->nphi * np, ar, at, ap, br, bt, bp, cr, ct, cp);
}

static int fc_reconstruct(rd_plan *plan, double *a, int boun, int ord, const double *neu, const double *coef)
{
    if (bad_plan(plan) || !a || !neu) return RD_ERR_ARG;
    Plan *p = P(plan);
    if ((boun != 5 && boun != 6) || (ord != 1 && ord != 2) || p->C == 0 || p->d < 1 || p->d > 8) {
        fprintf(stderr, "[ERROR] Neumann reconstruction not performed. Unsupported derivative order, "
                        "boundary surface or unallocated table.\n");
        return RD_ERR_ARG;
    }
    return launch_fc_reconstruct(p, (double2 *)a, boun, neu, coef);
}

int rd_neumann_reconstruct(rd_plan *plan, double *a, int boun, int ord, const double *neu)
{
    return fc_reconstruct(plan, a, boun, ord, neu, nullptr);
}

int rd_robin_reconstruct(rd_plan *plan, double *a, int boun, const double *neu, const double *coef)
{
    if (!coef) return RD_ERR_ARG;
    return fc_reconstruct(plan, a, boun, 1, neu, coef);
}

int rd_fc_filter(rd_plan *plan, char esp, double *a)
{
    if (bad_plan(plan) || !a) return RD_ERR_ARG;
    Plan *p = P(plan);
    if (esp == 'k') return launch_fc_filter_k(p, (double2 *)a);
    if (esp == 'r') {
        int rc = launch_z_forward(p, (double2 *)a);
        if (!rc) rc = launch_fc_filter_k(p, (double2 *)a);
        if (!rc) rc = launch_z_inverse(p, (double2 *)a);
        return rc;
    }
    fprintf(stderr, "redondito: unknown space %c in fc_filter\n", esp);
    return RD_ERR_ARG;
}

int rd_copy(rd_plan *plan, const double *a, double *b)
{
    if (bad_plan(plan) || !a || !b) return RD_ERR_ARG;
    Plan *p = P(plan);
    if (a == b) return RD_OK;
    RD_HIP(hipMemcpyAsync(b, a, sizeof(double2) * (size_t)p->nlm * p->nz, hipMemcpyDeviceToDevice,
                          p->stream));
    return RD_OK;
}

static int harvest_slot(Plan *p, int s)
{
    if (!p->slot_used[s]) return RD_OK;
    hipEvent_t *e = p->ev + (long)s * (ST_COUNT + 1);
    RD_HIP(hipEventSynchronize(e[ST_COUNT]));
    for (int i = 0; i < ST_COUNT; i++) {
        float ms = 0.f;
        RD_HIP(hipEventElapsedTime(&ms, e[i], e[i + 1]));
        p->stage_ms[i] += ms;
        p->stage_n[i] += 1;
    }
    p->slot_used[s] = 0;
    return RD_OK;
}

int rd_plan_profile(rd_plan *plan, int enable)
{
    if (bad_plan(plan)) return RD_ERR_ARG;
    Plan *p = P(plan);
    const int nev = kProfSlots * (ST_COUNT + 1);
    if (enable && !p->profiling) {
        p->ev = (hipEvent_t *)calloc(nev, sizeof(hipEvent_t));
        p->slot_used = (int *)calloc(kProfSlots, sizeof(int));
        if (!p->ev || !p->slot_used) return RD_ERR_NOMEM;
        for (int i = 0; i < nev; i++) RD_HIP(hipEventCreate(&p->ev[i]));
        p->next_slot = 0;
    }
    if (!enable && p->profiling) {
        RD_HIP(hipStreamSynchronize(p->stream));
        for (int i = 0; i < nev; i++) (void)hipEventDestroy(p->ev[i]);
        free(p->ev);
        free(p->slot_used);
        p->ev = nullptr;
        p->slot_used = nullptr;
    }
    p->profiling = enable ? 1 : 0;
    for (int i = 0; i < kMaxStages; i++) {
        p->stage_ms[i] = 0.0;
        p->stage_n[i] = 0;
    }
    return RD_OK;
}

int rd_plan_stage_times(rd_plan *plan, double *ms, long *count, int nmax)
{
    if (bad_plan(plan) || !ms || !count) return RD_ERR_ARG;
    Plan *p = P(plan);
    if (p->profiling)
        for (int s = 0; s < kProfSlots; s++) {
            int rc = harvest_slot(p, s);
            if (rc) return rc;
        }
    for (int i = 0; i < nmax && i < ST_COUNT; i++) {
        ms[i] = p->stage_ms[i];
        count[i] = p->stage_n[i];
    }
    return ST_COUNT;
}

int rd_nonlinear_uxw(rd_plan *plan, int prec, const double *W, const double *Z, double *NQ, double *NS,
                     double *NT)
{
    if (bad_plan(plan) || !W || !Z || !NQ || !NS || !NT) return RD_ERR_ARG;
    if (prec < RD_PREC_FP64 || prec > RD_PREC_OZAKI) return RD_ERR_ARG;
    Plan *p = P(plan);
    if (p->nwork < 6) return RD_ERR_STATE;
    hipEvent_t *ev = nullptr;
    if (p->profiling) {
        const int s = p->next_slot;
        p->next_slot = (s + 1) % kProfSlots;
        int hrc = harvest_slot(p, s);
        if (hrc) return hrc;
        ev = p->ev + (long)s * (ST_COUNT + 1);
        p->slot_used[s] = 1;
    }
    return uxw_evaluate(p, prec, (const double2 *)W, (const double2 *)Z, (double2 *)NQ, (double2 *)NS,
                        (double2 *)NT, ev);
}

int rd_plan_prepare(rd_plan *plan, int prec)
{
    if (bad_plan(plan)) return RD_ERR_ARG;
    if (prec < RD_PREC_FP64 || prec > RD_PREC_OZAKI) return RD_ERR_ARG;
    Plan *p = P(plan);
    int rc = RD_OK;
    if (prec == RD_PREC_OZAKI) rc = ensure_oz_tables(p);
    else if (prec != RD_PREC_FP64) rc = ensure_precision_tables(p, prec);
    if (rc) return rc;
    RD_HIP(hipStreamSynchronize(p->stream));
    return RD_OK;
}

int rd_nonlinear_uxw_capture(rd_plan *plan, int prec, const double *W, const double *Z, double *NQ, double *NS,
                             double *NT)
{
    if (bad_plan(plan) || !W || !Z || !NQ || !NS || !NT) return RD_ERR_ARG;
    Plan *p = P(plan);
    if (p->nwork < 6) return RD_ERR_STATE;
    int rc = rd_plan_prepare(plan, prec);
    if (rc) return rc;
    drop_plan_graph(p);
    // recorded on the plan's own stream without the per-stage events, then
    // launched on whichever stream the plan is bound to
    const hipStream_t user = p->stream;
    const int prof = p->profiling;
    p->stream = p->own_stream;
    p->profiling = 0;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(p->own_stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        rc = rd_nonlinear_uxw(plan, prec, W, Z, NQ, NS, NT);
        e = hipStreamEndCapture(p->own_stream, &g);
    }
    p->stream = user;
    p->profiling = prof;
    if (rc || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return rc ? rc : RD_ERR_HIP;
    }
    p->graph = g;
    RD_HIP(hipGraphInstantiate(&p->gexec, g, nullptr, nullptr, 0));
    return RD_OK;
}

int rd_plan_replay(rd_plan *plan)
{
    if (bad_plan(plan)) return RD_ERR_ARG;
    Plan *p = P(plan);
    if (!p->gexec) return RD_ERR_STATE;
    RD_HIP(hipGraphLaunch(p->gexec, p->stream));
    return RD_OK;
}

// One nonlinear step of the coupled state on the plan's six Fourier and six
// spectral work arrays:
//   B:      radial(Pb, Tb) -> S0..S2 (B only: one inverse radial transform), synthesis -> F3..F5
//   u, w:   radial(W, Z) -> S0..S5, synthesis of u -> F0..F2
//   u x B:  phi_cross(F0..F2, F3..F5) -> F3..F5, analysis -> MQ, MS, MT
//   u x w:  synthesis of w -> F3..F5, phi_cross -> F3..F5, analysis -> NQ, NS, NT
//   u.gT:   (dT/dr, T/r, 0) -> S0..S2, synthesis -> F3..F5, phi_dot -> F3,
//           scalar analysis -> NH
int rd_nonlinear_coupled(rd_plan *plan, int prec, const double *W, const double *Z, const double *Pb,
                         const double *Tb, const double *T, double *NQ, double *NS, double *NT, double *MQ,
                         double *MS, double *MT, double *NH)
{
    if (bad_plan(plan) || !W || !Z || !Pb || !Tb || !T || !NQ || !NS || !NT || !MQ || !MS || !MT || !NH)
        return RD_ERR_ARG;
    if (prec < RD_PREC_FP64 || prec > RD_PREC_OZAKI) return RD_ERR_ARG;
    Plan *p = P(plan);
    if (p->nwork < 6 || p->nr < 1) return RD_ERR_STATE;
    const int nr = p->nr;
    double2 **S = p->S, **F = p->F;
    const double2 *u_b[6] = {F[0], F[1], F[2], F[3], F[4], F[5]};
    double2 *o3[3] = {F[3], F[4], F[5]};
    int rc = launch_uxw_radial(p, (const double2 *)Pb, (const double2 *)Tb, S, 3);   // B only
    if (!rc) rc = launch_synthesis(p, prec, true, nr, S[0], S[1], S[2], p->nlm, F[3], F[4], F[5]);
    if (!rc) rc = launch_uxw_radial(p, (const double2 *)W, (const double2 *)Z, S);
    if (!rc) rc = launch_synthesis(p, prec, true, nr, S[0], S[1], S[2], p->nlm, F[0], F[1], F[2]);
    if (!rc) rc = launch_phi_cross(p, nr, u_b, o3, prec);
    if (!rc)
        rc = launch_analysis(p, prec, true, nr, F[3], F[4], F[5], (double2 *)MQ, (double2 *)MS, (double2 *)MT,
                             p->nlm);
    if (!rc) rc = launch_synthesis(p, prec, true, nr, S[3], S[4], S[5], p->nlm, F[3], F[4], F[5]);
    if (!rc) rc = launch_phi_cross(p, nr, u_b, o3, prec);
    if (!rc)
        rc = launch_analysis(p, prec, true, nr, F[3], F[4], F[5], (double2 *)NQ, (double2 *)NS, (double2 *)NT,
                             p->nlm);
    if (!rc) rc = launch_derivr_r(p, 1, (const double2 *)T, S[0]);
    if (!rc) rc = launch_multr(p, -1, (const double2 *)T, S[1]);
    if (!rc) rc = launch_zero_rows(p, S[2], 0, p->nz);
    if (!rc) rc = launch_synthesis(p, prec, true, nr, S[0], S[1], S[2], p->nlm, F[3], F[4], F[5]);
    if (!rc) rc = launch_phi_dot(p, nr, u_b, F[3], prec);
    if (!rc)
        rc = launch_analysis(p, prec, false, nr, F[3], nullptr, nullptr, (double2 *)NH, nullptr, nullptr, p->nlm);
    double *outs[7] = {NQ, NS, NT, MQ, MS, MT, NH};
    for (int i = 0; i < 7 && !rc; i++) rc = launch_zero_rows(p, (double2 *)outs[i], nr, p->nz - nr);
    return rc;
}

int rd_dirichlet(rd_plan *plan, double *a)
{
    if (bad_plan(plan) || !a) return RD_ERR_ARG;
    Plan *p = P(plan);
    if (p->nr < 1) return RD_ERR_STATE;
    return launch_dirichlet(p, (double2 *)a);
}

int rd_derivr_mat(rd_plan *plan, int order, double *d)
{
    if (bad_plan(plan) || !d || order < 0) return RD_ERR_ARG;
    Plan *p = P(plan);
    if (p->nr < 1) return RD_ERR_STATE;
    return build_derivr_mat(p, order, d);
}

int rd_bc_setup(rd_plan *plan, int bcint, int bcext)
{
    if (bad_plan(plan)) return RD_ERR_ARG;
    if (bcint != RD_BC_NOSLIP || bcext != RD_BC_NOSLIP) {
        fprintf(stderr, "redondito: unknown boundary condition type %d/%d\n", bcint, bcext);
        return RD_ERR_ARG;
    }
    Plan *p = P(plan);
    if (p->nr < 6 || p->lmax < 1) return RD_ERR_STATE;
    int failed_l = 0;
    int rc = build_bc(p, &failed_l);
    if (rc == RD_ERR_LU)
        fprintf(stderr, "redondito: LU of B_l failed for l = %d in invDl_setup\n", failed_l);
    if (rc) free_bc(p);
    return rc;
}

int rd_bc_tables(rd_plan *plan, double *binv, double *d2bor)
{
    if (bad_plan(plan)) return RD_ERR_ARG;
    Plan *p = P(plan);
    if (!p->bc_ready) return RD_ERR_STATE;
    const size_t N = (size_t)p->nr;
    if (binv)
        RD_HIP(hipMemcpyAsync(binv, p->bc_binv, sizeof(double) * N * (N - 4) * p->lmax,
                              hipMemcpyDeviceToDevice, p->stream));
    if (d2bor)
        RD_HIP(hipMemcpyAsync(d2bor, p->bc_d2bor, sizeof(double) * 4 * N, hipMemcpyDeviceToDevice, p->stream));
    return RD_OK;
}

int rd_invdl(rd_plan *plan, const double *phi, double *W)
{
    if (bad_plan(plan) || !phi || !W || phi == W) return RD_ERR_ARG;
    Plan *p = P(plan);
    if (!p->bc_ready) return RD_ERR_STATE;
    return launch_invdl(p, (const double2 *)phi, (double2 *)W);
}

int rd_v_imposebc(rd_plan *plan, double *tor, double *phi, double *pol)
{
    if (bad_plan(plan) || !tor || !phi || !pol || phi == pol || tor == pol || tor == phi) return RD_ERR_ARG;
    Plan *p = P(plan);
    if (!p->bc_ready) return RD_ERR_STATE;
    int rc = launch_dirichlet(p, (double2 *)tor);
    if (!rc) rc = launch_invdl(p, (const double2 *)phi, (double2 *)pol);
    if (!rc) rc = rd_dlr(plan, pol, phi);
    return rc;
}

int rd_v_imposebc_fast(rd_plan *plan, double *tor, double *phi, double *pol)
{
    if (bad_plan(plan) || !tor || !phi || !pol || phi == pol || tor == pol || tor == phi) return RD_ERR_ARG;
    Plan *p = P(plan);
    if (!p->bc_ready) return RD_ERR_STATE;
    int rc = launch_dirichlet(p, (double2 *)tor);
    if (!rc) rc = launch_invdl(p, (const double2 *)phi, (double2 *)pol);
    if (!rc) rc = launch_imposebc_rows(p, (const double2 *)pol, (double2 *)phi);
    return rc;
}

// b = D_l a = d2a/dr2 - l(l+1) a / r^2 with the scratch t1, t2 (dlr,
// pseudospec_hd.f90:254-288, on chosen buffers).
static int dl_into(Plan *p, const double2 *a, double2 *b, double2 *t1, double2 *t2)
{
    int rc = launch_derivr_r(p, 2, a, b);
    if (!rc) rc = launch_laplang(p, a, t1);
    if (!rc) rc = launch_multr(p, -2, t1, t2);
    if (!rc) rc = launch_saxpby(p, 2L * p->nlm * p->nz, (double *)b, (double *)b, 1.0, (const double *)t2, 1.0);
    return rc;
}

int rd_stage_uxw(rd_plan *plan, int prec, int o, double dt, double nu, int filter, const double *Z0,
                 const double *phi0, double *Z, double *phi, double *W)
{
    if (bad_plan(plan) || !Z0 || !phi0 || !Z || !phi || !W || o < 1) return RD_ERR_ARG;
    if (Z == phi || Z == W || phi == W || Z0 == phi || phi0 == Z || Z0 == W || phi0 == W) return RD_ERR_ARG;
    if (prec < RD_PREC_FP64 || prec > RD_PREC_OZAKI) return RD_ERR_ARG;
    Plan *p = P(plan);
    if (p->nwork < 6 || !p->bc_ready) return RD_ERR_STATE;
    double2 **S = p->S;
    const long nd = 2L * p->nlm * p->nz;
    const double c = dt / o;
    // N = u x omega into S0..S2 (the evaluation is done with S0..S5 when its
    // analysis writes them)
    int rc = rd_nonlinear_uxw(plan, prec, W, Z, (double *)S[0], (double *)S[1], (double *)S[2]);
    // toroidal: Z = Z0 + c (r NT + nu D_l Z)
    if (!rc) rc = dl_into(p, (const double2 *)Z, S[3], S[4], S[5]);
    if (!rc) rc = launch_multr(p, 1, S[2], S[4]);
    if (!rc) rc = launch_saxpby(p, nd, Z, Z0, 1.0, (const double *)S[4], c);
    if (!rc) rc = launch_saxpby(p, nd, Z, Z, 1.0, (const double *)S[3], c * nu);
    // poloidal: phi = phi0 + c (d(r NS)/dr - NQ + nu D_l phi)
    if (!rc) rc = dl_into(p, (const double2 *)phi, S[3], S[4], S[5]);
    if (!rc) rc = launch_multr(p, 1, S[1], S[4]);
    if (!rc) rc = launch_derivr_r(p, 1, S[4], S[5]);
    if (!rc) rc = launch_saxpby(p, nd, (double *)S[5], (const double *)S[5], 1.0, (const double *)S[0], -1.0);
    if (!rc) rc = launch_saxpby(p, nd, phi, phi0, 1.0, (const double *)S[5], c);
    if (!rc) rc = launch_saxpby(p, nd, phi, phi, 1.0, (const double *)S[3], c * nu);
    if (!rc && filter) rc = rd_fc_filter(plan, 'r', Z);
    if (!rc && filter) rc = rd_fc_filter(plan, 'r', phi);
    if (!rc) rc = rd_v_imposebc_fast(plan, Z, phi, W);
    return rc;
}

}  // extern "C"

namespace rd {

int plan_create(Plan **out, int lmax, int nlat, int nphi, int nz, int C, int d, double dz, double rint,
                const double *dir, int nwork)
{
    return plan_create_impl(out, lmax, nlat, nphi, nz, C, d, dz, rint, dir, nwork);
}

void plan_destroy(Plan *p)
{
    if (!p) return;
    (void)hipStreamSynchronize(p->stream);
    free_plan(p);
}

}  // namespace rd
