// Internal declarations of the MI355X pseudospectral library (gfx950 only).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cstdio>

#include "../../include/redondito.h"

namespace rd {

constexpr int kMaxStages = 24;

// Radix sequence of one FFT length; the Stockham stages run in this order.
struct FftSeq {
    int n;
    int nst;
    int rad[kMaxStages];
};

FftSeq make_fft_seq(int n);
// smallest primitive root of a prime n (the generator of the Rader tables)
int primitive_root(int n);

// The radial FFT of length n as the kernels run it.  A prime n whose n - 1
// has no prime factor above 13 runs in Rader form: X[0] = sum_j x_j and
// X[g^-q] = x_0 + sum_p x[g^p] c[q - p] (g a primitive root mod n, c_m =
// e^{s 2 pi i g^-m / n}), the cyclic convolution by two FFTs of length n - 1
// (seq) and the precomputed spectrum cf[s] of c divided by n - 1.  Otherwise
// seq is of length n itself.
struct ZFft {
    FftSeq seq;
    int n, rader;
    const int *pin, *pout;          // g^p mod n, g^-q mod n
    const double2 *cf[2];           // s = -1, s = +1
};
constexpr int kZScalars = 32;       // LDS scratch (double2) the Rader form needs

// Workgroup index remapped so that consecutive logical tiles run on the same
// XCD (the hardware deals workgroups round-robin over the 8 XCDs, each with
// its own L2): tiles that share operands then share one L2.  Only for grids
// whose tiles cost the same; uneven work goes through per-XCD queues.
constexpr int kXcds = 8;
__device__ __forceinline__ int xcd_tile(int bid, int nblocks)
{
    const int q = nblocks / 8, r = nblocks % 8;
    const int x = bid % 8, i = bid / 8;
    return x * q + (x < r ? x : r) + i;
}

// Sizes and device tables of one plan.
struct Plan {
    int lmax, L, nlat, nphi, nlm, nth, nthp, nthp16;
    int nz, C, d, nr;
    double dz, rint;
    int device;                    // the device current at creation (the plan's tables live there)
    int n_cu;                      // compute units of that device (persistent grids)
    hipStream_t stream;
    hipStream_t own_stream;
    // grid (host copies)
    double *h_ct, *h_st, *h_wt, *h_z, *h_kz;
    int *h_ell, *h_emm;
    // device grid
    double *d_ct, *d_st, *d_wt, *d_z, *d_kz, *d_dir;
    int *d_ell;
    // associated-Legendre tables: synthesis [m,p][row l_p][theta nthp],
    // analysis [m,p][theta nthp16][col l_p] with weights (and 1/l(l+1))
    double *syn_P, *syn_D, *syn_M;
    double *ana_P, *ana_D, *ana_M;
    // precision-form tables, built on first use: FP32 copies in the same
    // layouts, and BF16 hi/lo splits with four consecutive K values packed
    // (synthesis [m,p][row/4][theta][4], analysis [m,p][theta/4][col][4])
    float *syn32_P, *syn32_D, *syn32_M, *ana32_P, *ana32_D, *ana32_M;
    uint16_t *synb_hi[3], *synb_lo[3], *anab_hi[3], *anab_lo[3];
    void *oz;                      // int8 digit tables and operand scratch of the Ozaki form
    int oz_digits;                 // digits per operand of the Ozaki form's calls (5, 6 or 7)
    long *d_syn_off, *d_ana_off;   // per (m,p)
    int *d_ana_ncol;               // per m
    long syn_rows, ana_elems;
    // Per-XCD work queues, stored interleaved (entry e of XCD x at e*8+x, -1 =
    // empty): synthesis entries m | tt<<16, analysis entries m | lt<<16.
    int *d_syn_q, *d_ana_q;
    int syn_qlen, ana_qlen;
    // FFTs; tw_z is of the length of zfft.seq (n - 1 in Rader form)
    FftSeq seq_phi;
    ZFft zfft;
    double2 *tw_phi, *tw_z;        // exp(-2 pi i t / n)
    int *z_perm;                   // Rader index tables (2 (n - 1) ints)
    double2 *z_cf;                 // Rader spectra (2 (n - 1))
    // no-slip wall solves, built by rd_bc_setup (column major, N = nr):
    // d1, d2 N x N; binv N x (N-4) per l = 1..lmax; d2bor 4 x N
    int bc_ready;
    double *bc_d1, *bc_d2, *bc_binv, *bc_d2bor;
    int *bc_queue;                 // (l | m0 << 16) tiles of invDl
    int bc_qlen;
    // workspace (nwork of each; 6 for the nonlinear term)
    int nwork;
    double2 *F[6];                 // Fourier coefficients [k][m][theta], nz planes
    double2 *S[6];                 // spectral scratch (nlm, nz)
    long fsize, ssize;             // in elements of the type
    // one captured evaluation of rd_nonlinear_uxw (rd_nonlinear_uxw_capture)
    hipGraph_t graph;
    hipGraphExec_t gexec;
    // per-stage event timing of rd_nonlinear_uxw: a ring of event slots,
    // harvested lazily so that timing adds no host synchronisation
    int profiling;
    hipEvent_t *ev;                // kProfSlots x (ST_COUNT + 1)
    int *slot_used;
    int next_slot;
    double stage_ms[kMaxStages];
    long stage_n[kMaxStages];
};

constexpr int kProfSlots = 512;

// Stages of rd_nonlinear_uxw, in order.
// (the u and omega syntheses are one launch; the radial step also clears
// the output rows past nr)
enum UxwStage { ST_RADIAL = 0, ST_SYN, ST_PHI_CROSS, ST_ANALYSIS, ST_COUNT };

// ---------------------------------------------------------------- kernels
// Legendre contractions over all planes (Fourier layout F[k][m][theta]);
// prec is RD_PREC_FP64, RD_PREC_FP32, RD_PREC_BF16X3 or RD_PREC_OZAKI.
int launch_synthesis(Plan *p, int prec, bool vec, int nplanes, const double2 *q, const double2 *s,
                     const double2 *t, long ld, double2 *f0, double2 *f1, double2 *f2);
// Two vector syntheses (field sets a and b) in one launch where the form has
// the merged kernel (FP64, FP32), else one after the other.
int launch_synthesis2(Plan *p, int prec, int nplanes, const double2 *const *in_a, double2 *const *out_a,
                      const double2 *const *in_b, double2 *const *out_b, long ld);
int launch_analysis(Plan *p, int prec, bool vec, int nplanes, const double2 *g0, const double2 *g1,
                    const double2 *g2, double2 *q, double2 *s, double2 *t, long ld);
int build_legendre_tables(Plan *p);
int ensure_precision_tables(Plan *p, int prec);
void free_precision_tables(Plan *p);
// in2/out2: a second field set synthesised by the same launch (vector form)
int launch_synthesis_bf3(Plan *p, bool vec, int nplanes, const double2 *q, const double2 *s, const double2 *t,
                         long ld, double2 *f0, double2 *f1, double2 *f2, const double2 *const *in2 = nullptr,
                         double2 *const *out2 = nullptr);
int ensure_oz_tables(Plan *p);
void free_oz_tables(Plan *p);
int launch_synthesis_oz(Plan *p, bool vec, int nplanes, const double2 *q, const double2 *s, const double2 *t,
                        long ld, double2 *f0, double2 *f1, double2 *f2);
int launch_analysis_oz(Plan *p, bool vec, int nplanes, const double2 *g0, const double2 *g1, const double2 *g2,
                       double2 *q, double2 *s, double2 *t, long ld);
int launch_analysis_bf3(Plan *p, bool vec, int nplanes, const double2 *g0, const double2 *g1,
                        const double2 *g2, double2 *q, double2 *s, double2 *t, long ld);

// phi transforms between F[k][m][theta] and f(theta, phi, k).
int launch_phi_inverse(Plan *p, int nfields, int nplanes, const double2 *const *F, double *const *f);
int launch_phi_forward(Plan *p, int nfields, int nplanes, const double *const *f, double2 *const *F);
// Fused: six Fourier fields (u_r,u_t,u_p,w_r,w_t,w_p) -> u x w -> three Fourier fields.
// prec RD_PREC_BF16X3 transforms the rows in FP32 where an FP32 kernel exists
// (nphi = 768 and 1536); every other form in FP64.
int launch_phi_cross(Plan *p, int nplanes, const double2 *const *Fin, double2 *const *Fout, int prec);
// Fused: (u_r,u_t,u_p,g_r,g_t,g_p) -> u . g -> one Fourier field.
int launch_phi_dot(Plan *p, int nplanes, const double2 *const *Fin, double2 *Fout, int prec);

// Radial transforms on (nlm, nz) complex, in place.
int launch_z_forward(Plan *p, double2 *a);
int launch_z_inverse(Plan *p, double2 *a);
// Radial derivative of order p in (lm,r): b = derivr(a, p).
int launch_derivr_r(Plan *p, int order, const double2 *a, double2 *b);
// Fused radial step of the nonlinear term: W, Z -> the six (Q,S,T) arrays of
// u and omega (nout = 6), or the three of u only (nout = 3).
// zero3: three (nlm, nz) arrays whose rows nr..nz-1 the kernel clears
// (the continuation rows of the nonlinear term's outputs), or null.
int launch_uxw_radial(Plan *p, const double2 *W, const double2 *Z, double2 *const *qst, int nout = 6,
                      double2 *const *zero3 = nullptr);
// Radial derivative of order p of ncols columns with plane stride ld.
int launch_derivr_cols(Plan *p, int order, const double2 *a, double2 *b, int ncols, long ld);

// Pointwise operators.
int launch_derivr_k(Plan *p, int order, const double2 *a, double2 *b);
int launch_laplang(Plan *p, const double2 *a, double2 *b);
int launch_multr(Plan *p, int pw, const double2 *a, double2 *b);
int launch_saxpby(Plan *p, long n, double *z, const double *x, double a, const double *y, double b);
int launch_prodvec(Plan *p, long npts, const double *ar, const double *at, const double *ap,
                   const double *br, const double *bt, const double *bp, double *cr, double *ct,
                   double *cp);
int launch_fc_filter_k(Plan *p, double2 *a);
int launch_dirichlet(Plan *p, double2 *a);
int launch_zero_rows(Plan *p, double2 *a, int row0, int nrows);
// One u x omega evaluation on the plan's stream (rd_nonlinear_uxw without its
// argument checks); ev: ST_COUNT + 1 events recorded around the stages, or null.
int uxw_evaluate(Plan *p, int prec, const double2 *W, const double2 *Z, double2 *NQ, double2 *NS, double2 *NT,
                 hipEvent_t *ev);
// Wall value of every lm column from its normal derivative (boun 5: z(1),
// 6: z(nz - C)), Neumann weights neu (d values), Robin coefficients coef
// (nlm reals, device) or NULL.
int launch_fc_reconstruct(Plan *p, double2 *a, int boun, const double *neu, const double *coef);

// Plan life cycle for internal users (the batch): nwork scratch arrays.
int plan_create(Plan **out, int lmax, int nlat, int nphi, int nz, int C, int d, double dz, double rint,
                const double *dir, int nwork);
void plan_destroy(Plan *p);

// Persistent Legendre launches over many problems: work items (problem, m,
// tile, plane tile) pulled from a counter, largest first.  Scalar form: one
// synthesis and one analysis operand set per problem over its nz planes (the
// 3D transforms).  Vector form (vec): nsets (Q,S,T) synthesis sets per
// problem (u and omega of the nonlinear term) and one vector analysis set,
// over the problem's nr physical planes.
struct BatchLeg {
    int nprob, prec, jb, vec, nsets;
    void *d_syn_args, *d_ana_args;       // per-problem kernel arguments
    int4 *d_syn_items, *d_ana_items;
    int n_syn_items, n_ana_items;
    int *d_counter;                      // [0] synthesis, [1] analysis
    size_t arg_bytes_syn, arg_bytes_ana;
    void *h_syn_args, *h_ana_args;       // host copies of the uploaded arguments
};
int batch_leg_init(BatchLeg *b, Plan *const *plans, int nprob, int prec, int jb, int vec = 0, int nsets = 1);
// Binds the per-problem operands (spectral and Fourier arrays); uploads the
// arguments when they changed (synchronous on the stream).
int batch_leg_bind(BatchLeg *b, Plan *const *plans, const double2 *const *syn_in, double2 *const *syn_out,
                   const double2 *const *ana_in, double2 *const *ana_out, hipStream_t s);
// Vector form: syn_in / syn_out hold 3 arrays per (problem, set), set fastest
// ((p nsets + s) 3 + f); ana_in / ana_out 3 per problem.
int batch_vleg_bind(BatchLeg *b, Plan *const *plans, const double2 *const *syn_in, double2 *const *syn_out,
                    const double2 *const *ana_in, double2 *const *ana_out, hipStream_t s);
int launch_batch_synthesis(BatchLeg *b, hipStream_t s, int nwg);
int launch_batch_analysis(BatchLeg *b, hipStream_t s, int nwg);
void batch_leg_free(BatchLeg *b);

// Wall solves.
int build_derivr_mat(Plan *p, int order, double *d);
int build_bc(Plan *p, int *failed_l);
void free_bc(Plan *p);
int launch_invdl(Plan *p, const double2 *phi, double2 *W);
int launch_imposebc_rows(Plan *p, const double2 *pol, double2 *phi);

}  // namespace rd

#define RD_HIP(call)                                                                     \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess) {                                                          \
            fprintf(stderr, "redondito: %s failed: %s (%s:%d)\n", #call,                 \
                    hipGetErrorString(e_), __FILE__, __LINE__);                           \
            return RD_ERR_HIP;                                                           \
        }                                                                                \
    } while (0)

#define RD_LAUNCH_CHECK()                                                                \
    do {                                                                                 \
        hipError_t e_ = hipGetLastError();                                               \
        if (e_ != hipSuccess) {                                                          \
            fprintf(stderr, "redondito: launch failed: %s (%s:%d)\n",                    \
                    hipGetErrorString(e_), __FILE__, __LINE__);                           \
            return RD_ERR_HIP;                                                           \
        }                                                                                \
    } while (0)
