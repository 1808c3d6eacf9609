// Associated-Legendre tables and the per-m Legendre contractions (FP64 MFMA).
//
// Synthesis, for every m and every plane k (SHTns SH_to_spat / SHqst_to_spat
// semantics, called per plane by src/sphp/sphp_sca.F90:484-486 and
// src/sphp/sphp_vec.F90:152-160):
//   F_m(theta, k) = sum_l P~_lm(theta) a_lm(k)
// run as a GEMM  [theta x l] . [l x (k, re/im)]  with the equatorial symmetry
// P~_lm(pi - theta) = (-1)^(l-m) P~_lm(theta): the even and odd l - m halves are
// accumulated separately over the northern nodes and combined into north and
// south rows.  Analysis is the transposed contraction with the Gauss weights
// (spat_to_SH / spat_to_SHqst, src/sphp/sphp_sca.F90:311-313,
// src/sphp/sphp_vec.F90:88-92).
//
// Each workgroup is 4 wave64s; each wave owns a 16-row slab of the output
// and issues v_mfma_f64_16x16x4f64.  The spectral operand is staged through
// LDS with coalesced loads along l; the table operand is read straight from
// HBM/L2 in the layout the MFMA A-operand wants (16 consecutive doubles per
// lane group).
#include "rd_internal.h"
#include <type_traits>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

namespace rd {

typedef double v4d __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

// 16x16x4 MFMA of one operand type.  A lane (i + 16k) holds A[i][k], B lane
// (j + 16k) holds B[k][j] for both types; the accumulator rows differ:
// lane l, register v holds D[row(l/16, v)][l%16] with row = l/16 + 4v for
// F64 and 4(l/16) + v for F32 (measured on gfx950).
template <typename T> struct Mf;
template <> struct Mf<double> {
    using V = v4d;
    static __device__ __forceinline__ V mma(double a, double b, V c)
    {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int kk, int v) { return kk + 4 * v; }
};
template <> struct Mf<float> {
    using V = v4f;
    static __device__ __forceinline__ V mma(float a, float b, V c)
    {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ __forceinline__ int row(int kk, int v) { return 4 * kk + v; }
};

__device__ __forceinline__ long lm_index(int lmax, int l, int m)
{
    return (long)m * (2 * lmax + 1 - m) / 2 + l;
}

// ------------------------------------------------------------------ grid
// P_n(cos t) = sum_k h_k cos((n - 2k) t) and its t-derivative (Neumaier
// compensated), h_k = g_k g_{n-k}, g_k = prod_{j<=k} (2j-1)/(2j).
__device__ void legendre_cos_series(int n, const double *h, double t, double *P, double *dP)
{
    double p = (n % 2 == 0) ? h[n / 2] : 0.0;
    double s = 0.0, c = 0.0;
    for (int k = 0; k < (n + 1) / 2; k++) {
        const int m = n - 2 * k;
        double sn, cs;
        sincos(m * t, &sn, &cs);
        p += 2.0 * h[k] * cs;
        const double term = -2.0 * h[k] * m * sn;
        const double u = s + term;
        c += (fabs(s) >= fabs(term)) ? (s - u) + term : (term - u) + s;
        s = u;
    }
    *P = p;
    *dP = s + c;
}

// Gauss-Legendre nodes, north to south (one thread per northern node):
// Newton on theta with P_n(cos theta) from its cosine series, so theta and
// sin(theta) stay accurate to an ulp at the poles; w = 2 / (dP_n/dtheta)^2.
__global__ void gauss_kernel(int nlat, const double *h, double *ct, double *st, double *wt)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (nlat + 1) / 2) return;
    double th = M_PI * (i + 0.75) / (nlat + 0.5);
    double P, dP;
    int extra = 0;
    for (int it = 0; it < 60; it++) {
        legendre_cos_series(nlat, h, th, &P, &dP);
        const double dth = P / dP;
        th -= dth;
        if (fabs(dth) < 1e-15 * th && ++extra > 2) break;
    }
    legendre_cos_series(nlat, h, th, &P, &dP);
    const double w = 2.0 / (dP * dP);
    double s, x;
    sincos(th, &s, &x);
    ct[i] = x;
    st[i] = s;
    wt[i] = w;
    ct[nlat - 1 - i] = -x;
    st[nlat - 1 - i] = s;
    wt[nlat - 1 - i] = w;
}

// One thread per (m, northern node): the FP64 recurrence
//   P~_mm = c_m s^m,  P~_{m+1,m} = sqrt(2m+3) x P~_mm,
//   P~_lm = a_lm (x P~_{l-1,m} - b_lm P~_{l-2,m})
// run on P~ and on D = s^m dQ/dx (P~ = s^m Q(x)) with a power-of-two scale
// for the sectoral start, then
//   dP~/dtheta = (m x P~ - s^2 D) / s,   m P~ / s.
__global__ void legendre_tables_kernel(int lmax, int nth, int nthp, int nthp16, const double *ct,
                                       const double *st, const double *wt, const long *syn_off,
                                       const long *ana_off, const int *ana_ncol, double *sP,
                                       double *sD, double *sM, double *aP, double *aD, double *aM)
{
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    int m = blockIdx.y;
    if (i >= nth) return;
    double cm = sqrt((2.0 * m + 1.0) / (4.0 * M_PI));
    for (int k = 1; k <= m; k++) cm *= sqrt((2.0 * k - 1.0) / (2.0 * k));
    double x = ct[i], s = st[i], w = wt[i];
    double v = cm;
    int e2 = 0;
    for (int k = 0; k < m; k++) {
        v *= s;
        if (v < 0x1p-400) {
            v *= 0x1p400;
            e2 -= 400;
        }
    }
    double p1 = v, p2 = 0.0, d1 = 0.0, d2 = 0.0;
    int ncol = ana_ncol[m];
    for (int l = m; l <= lmax; l++) {
        double pc, dc;
        if (l == m) {
            pc = p1;
            dc = 0.0;
        } else if (l == m + 1) {
            double a = sqrt(2.0 * m + 3.0);
            pc = a * x * p1;
            dc = a * p1;
        } else {
            double a = sqrt((4.0 * l * l - 1.0) / ((double)l * l - (double)m * m));
            double b = sqrt(((l - 1.0) * (l - 1.0) - (double)m * m) / (4.0 * (l - 1.0) * (l - 1.0) - 1.0));
            pc = a * (x * p1 - b * p2);
            dc = a * (p1 + x * d1 - b * d2);
        }
        if (l > m) {
            p2 = p1;
            p1 = pc;
            d2 = d1;
            d1 = dc;
        }
        double P = ldexp(pc, e2), Dq = ldexp(dc, e2);
        double dP = (m * x * P - s * s * Dq) / s;
        double mP = m * P / s;
        int par = (l - m) & 1, lp = (l - m) >> 1;
        long so = (syn_off[2 * m + par] + lp) * nthp + i;
        sP[so] = P;
        sD[so] = dP;
        sM[so] = mP;
        double ll = (double)l * (l + 1);
        double il = l > 0 ? 1.0 / ll : 0.0;
        long ao = ana_off[2 * m + par] + (long)i * ncol + lp;
        aP[ao] = w * P;
        aD[ao] = w * dP * il;
        aM[ao] = w * mP * il;
    }
}

static inline int rup(int a, int b) { return (a + b - 1) / b * b; }

// Interleave per-XCD queues: entry e of queue x goes to e*kXcds + x, shorter
// queues padded with -1.  Workgroup b then runs on XCD b % kXcds.
static std::vector<int> interleave_queues(const std::vector<std::vector<int>> &q, int *qlen)
{
    size_t len = 0;
    for (const auto &v : q) len = std::max(len, v.size());
    std::vector<int> out(len * kXcds, -1);
    for (int x = 0; x < kXcds; x++)
        for (size_t e = 0; e < q[x].size(); e++) out[e * kXcds + x] = q[x][e];
    *qlen = (int)len;
    return out;
}

// Synthesis: the work of a tile of order m is proportional to L - m, so m and
// L-1-m are paired into chunks of equal cost, dealt round-robin to the XCDs.
static std::vector<int> synthesis_queues(int L, int n_tt, int *qlen)
{
    std::vector<std::vector<int>> q(kXcds);
    int c = 0;
    for (int a = 0, b = L - 1; a <= b; a++, b--, c++) {
        auto &dst = q[c % kXcds];
        for (int tt = 0; tt < n_tt; tt++) dst.push_back(a | (tt << 16));
        if (b != a)
            for (int tt = 0; tt < n_tt; tt++) dst.push_back(b | (tt << 16));
    }
    return interleave_queues(q, qlen);
}

// Analysis: tiles cost the same; all l-tiles of one m share its Fourier
// coefficients, so whole m groups go to the least-loaded XCD (largest first).
static std::vector<int> analysis_queues(const std::vector<int> &ncol, int *qlen)
{
    const int L = (int)ncol.size();
    std::vector<std::vector<int>> q(kXcds);
    for (int m = 0; m < L; m++) {   // ncol is non-increasing in m
        int x = 0;
        for (int y = 1; y < kXcds; y++)
            if (q[y].size() < q[x].size()) x = y;
        for (int lt = 0; lt < ncol[m] / 32; lt++) q[x].push_back(m | (lt << 16));
    }
    return interleave_queues(q, qlen);
}

int build_legendre_tables(Plan *p)
{
    const int L = p->L;
    std::vector<long> syn_off(2 * L), ana_off(2 * L);
    std::vector<int> ana_ncol(L);
    long rows = 0, elems = 0;
    for (int m = 0; m < L; m++) {
        int ne = (L - m + 1) / 2;             // l - m even count (>= odd count)
        int rowsm = rup(ne, 16);
        syn_off[2 * m] = rows;
        syn_off[2 * m + 1] = rows + rowsm;
        rows += 2 * rowsm;
        int ncol = rup(ne, 32);
        ana_ncol[m] = ncol;
        ana_off[2 * m] = elems;
        ana_off[2 * m + 1] = elems + (long)p->nthp16 * ncol;
        elems += 2L * p->nthp16 * ncol;
    }
    // the kernels index the tables with 32-bit element offsets
    if (rows * (long)p->nthp >= (1L << 32) || elems >= (1L << 32)) return RD_ERR_ARG;
    p->syn_rows = rows;
    p->ana_elems = elems;
    std::vector<int> synq = synthesis_queues(L, p->nthp / 64, &p->syn_qlen);
    std::vector<int> anaq = analysis_queues(ana_ncol, &p->ana_qlen);
    size_t sb = sizeof(double) * rows * p->nthp, ab = sizeof(double) * elems;
    RD_HIP(hipMalloc(&p->syn_P, sb));
    RD_HIP(hipMalloc(&p->syn_D, sb));
    RD_HIP(hipMalloc(&p->syn_M, sb));
    RD_HIP(hipMalloc(&p->ana_P, ab));
    RD_HIP(hipMalloc(&p->ana_D, ab));
    RD_HIP(hipMalloc(&p->ana_M, ab));
    RD_HIP(hipMalloc(&p->d_syn_off, sizeof(long) * 2 * L));
    RD_HIP(hipMalloc(&p->d_ana_off, sizeof(long) * 2 * L));
    RD_HIP(hipMalloc(&p->d_ana_ncol, sizeof(int) * L));
    RD_HIP(hipMalloc(&p->d_syn_q, sizeof(int) * synq.size()));
    RD_HIP(hipMalloc(&p->d_ana_q, sizeof(int) * anaq.size()));
    hipStream_t st = p->stream;
    RD_HIP(hipMemsetAsync(p->syn_P, 0, sb, st));
    RD_HIP(hipMemsetAsync(p->syn_D, 0, sb, st));
    RD_HIP(hipMemsetAsync(p->syn_M, 0, sb, st));
    RD_HIP(hipMemsetAsync(p->ana_P, 0, ab, st));
    RD_HIP(hipMemsetAsync(p->ana_D, 0, ab, st));
    RD_HIP(hipMemsetAsync(p->ana_M, 0, ab, st));
    RD_HIP(hipMemcpyAsync(p->d_syn_off, syn_off.data(), sizeof(long) * 2 * L, hipMemcpyHostToDevice, st));
    RD_HIP(hipMemcpyAsync(p->d_ana_off, ana_off.data(), sizeof(long) * 2 * L, hipMemcpyHostToDevice, st));
    RD_HIP(hipMemcpyAsync(p->d_ana_ncol, ana_ncol.data(), sizeof(int) * L, hipMemcpyHostToDevice, st));
    RD_HIP(hipMemcpyAsync(p->d_syn_q, synq.data(), sizeof(int) * synq.size(), hipMemcpyHostToDevice, st));
    RD_HIP(hipMemcpyAsync(p->d_ana_q, anaq.data(), sizeof(int) * anaq.size(), hipMemcpyHostToDevice, st));

    // cosine-series coefficients of P_nlat (host FP64, shared with the nodes kernel)
    std::vector<double> g(p->nlat + 1), h(p->nlat + 1);
    g[0] = 1.0;
    for (int k = 1; k <= p->nlat; k++) g[k] = g[k - 1] * (2.0 * k - 1.0) / (2.0 * k);
    for (int k = 0; k <= p->nlat; k++) h[k] = g[k] * g[p->nlat - k];
    double *d_h = nullptr;
    RD_HIP(hipMalloc(&d_h, sizeof(double) * (p->nlat + 1)));
    RD_HIP(hipMemcpyAsync(d_h, h.data(), sizeof(double) * (p->nlat + 1), hipMemcpyHostToDevice, st));
    hipLaunchKernelGGL(gauss_kernel, dim3((p->nlat / 2 + 63) / 64), dim3(64), 0, st, p->nlat, d_h, p->d_ct,
                       p->d_st, p->d_wt);
    RD_LAUNCH_CHECK();
    hipLaunchKernelGGL(legendre_tables_kernel, dim3((p->nth + 63) / 64, L), dim3(64), 0, st, p->lmax,
                       p->nth, p->nthp, p->nthp16, p->d_ct, p->d_st, p->d_wt, p->d_syn_off,
                       p->d_ana_off, p->d_ana_ncol, p->syn_P, p->syn_D, p->syn_M, p->ana_P,
                       p->ana_D, p->ana_M);
    RD_LAUNCH_CHECK();
    RD_HIP(hipMemcpyAsync(p->h_ct, p->d_ct, sizeof(double) * p->nlat, hipMemcpyDeviceToHost, st));
    RD_HIP(hipMemcpyAsync(p->h_st, p->d_st, sizeof(double) * p->nlat, hipMemcpyDeviceToHost, st));
    RD_HIP(hipMemcpyAsync(p->h_wt, p->d_wt, sizeof(double) * p->nlat, hipMemcpyDeviceToHost, st));
    RD_HIP(hipStreamSynchronize(st));
    RD_HIP(hipFree(d_h));
    return RD_OK;
}

// ------------------------------------------------------------- synthesis
// Workgroup tile: one m, 64 northern nodes (16 per wave), SJT real columns
// (SJT/2 planes).  The l loop runs in chunks of 32 (16 per parity).
constexpr int SKC = 32;

// 16-column blocks per wave of the one-problem kernels.  FP64 vector tiles
// keep one block: the accumulators then fit 108 VGPRs, which buys 4 waves
// per SIMD, and at this occupancy the global-load latency of the staging
// loop hides better than with two blocks at 2 waves and a double-buffered
// stage (0.54 against 0.59 ms at lmax 255, 256 planes).
template <typename T, bool VEC>
constexpr int syn_jb()
{
    return VEC ? 1 : 4;
}

template <bool VEC, int JBV = (VEC ? 2 : 4)>
struct SynCfg {
    static constexpr int NF = VEC ? 3 : 1;
    static constexpr int JB = JBV;               // 16-column blocks per wave
    static constexpr int JT = 16 * JB;           // columns per workgroup
    static constexpr int JTP = JT + 2;           // LDS row pitch (doubles)
};

template <typename T>
struct SynArgs {
    const T *tP, *tD, *tM;
    const long *off;
    const int *queue;
    const double2 *in[3];
    double2 *out[3];
    // a second field set transformed by the same launch (nset = 2): the
    // one-problem kernel runs both sets' plane tiles of one (m, theta tile)
    // entry back to back, so the entry's table rows serve both from L2
    const double2 *in2[3];
    double2 *out2[3];
    long ld;
    int lmax, nlat, nth, nthp, nplanes, n_jt, nset;
};

// One synthesis tile (m, theta tile tt, plane tile jt) of one problem.
// T = double: the FP64 path; T = float: the FP32 precision form (tables and
// operands rounded to FP32, FP32 accumulation, FP64 output).
// PF: prefetch the next spectral chunk and the next step's table values
// into registers (the one-problem kernels).  The persistent batch kernels
// keep PF off: the extra registers cost them more occupancy than the
// prefetch gains (round trip 13.4 against 14.2 ms).
template <typename T, bool VEC, int JBV, bool PF = true>
__device__ __forceinline__ void syn_tile(
    const SynArgs<T> &A, int m, int tt, int jt,
    T (&Bs)[SynCfg<VEC, JBV>::NF][SKC][SynCfg<VEC, JBV>::JTP])
{
    using Cfg = SynCfg<VEC, JBV>;
    using V = typename Mf<T>::V;
    constexpr int NF = Cfg::NF, JB = Cfg::JB, JT = Cfg::JT;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int L = A.lmax + 1;
    const int k0 = jt * (JT / 2);
    const int th0 = tt * 64 + wave * 16;
    const int col = lane & 15, kk = lane >> 4;
    const T sgn_i = (col & 1) ? T(1) : T(-1);   // (i X)_j = sgn_i * X_{j^1}

    V acc[NF][2][JB];
#pragma unroll
    for (int f = 0; f < NF; f++)
#pragma unroll
        for (int s = 0; s < 2; s++)
#pragma unroll
            for (int b = 0; b < JB; b++) acc[f][s][b] = (V){0, 0, 0, 0};

    const long offe = A.off[2 * m] * A.nthp, offo = A.off[2 * m + 1] * A.nthp;
    const long lm0 = lm_index(A.lmax, m, m);

    // a chunk is NF x (JT/2 planes) x SKC l's, l fastest across threads;
    // each thread moves NLD double2 of it: element tid + 256 i, whose field
    // (256 i) / PER is a compile-time constant because PER is a multiple of
    // 256, so the field's base pointer is a scalar kernel argument.  The
    // loads go to clamped addresses, unconditionally, so that they issue
    // back to back with no wait between them; entries beyond lmax or the
    // last plane are zeroed when the chunk is stored to LDS, which happens
    // after the MFMAs that the loads overlap.
    constexpr int PER = (JT / 2) * SKC;
    constexpr int NLD = NF * PER / 256;
    static_assert(PER % 256 == 0, "a thread's chunk elements must each lie in one field");
    double2 r[NLD];
    unsigned rlive = 0;   // bit i: element i lies inside the spectrum
    auto load = [&](int lbase) {
        rlive = 0;
#pragma unroll
        for (int i = 0; i < NLD; i++) {
            const int f = (256 * i) / PER, q = tid + (256 * i) % PER, pl = q / SKC, ll = q % SKC;
            const int l = lbase + ll, k = k0 + pl;
            const int lc = l < A.lmax ? l : A.lmax, kc = k < A.nplanes ? k : A.nplanes - 1;
            r[i] = A.in[f][lm0 + (lc - m) + kc * A.ld];
            rlive |= (unsigned)(l <= A.lmax && k < A.nplanes) << i;
        }
    };
    auto store = [&]() {
#pragma unroll
        for (int i = 0; i < NLD; i++) {
            const int f = (256 * i) / PER, q = tid + (256 * i) % PER, pl = q / SKC, ll = q % SKC;
            const bool ok = (rlive >> i) & 1u;
            Bs[f][ll][2 * pl] = ok ? (T)r[i].x : T(0);
            Bs[f][ll][2 * pl + 1] = ok ? (T)r[i].y : T(0);
        }
    };
    // The table values of step g (rows 4 g + kk of both parities) are
    // loaded one step ahead, across chunk boundaries, so that their latency
    // overlaps the MFMAs of step g - 1.  32-bit element offsets from the
    // uniform table bases (tables stay below 4 GB): one VGPR per address
    // instead of a 64-bit pointer per table and step.
    struct Tab {
        T Pe, Po, De, Do, Me, Mo;
    };
    auto ltab = [&](int row) {
        Tab t;
        const unsigned ie = (unsigned)offe + (unsigned)(row * A.nthp + th0 + col);
        const unsigned io = (unsigned)offo + (unsigned)(row * A.nthp + th0 + col);
        t.Pe = A.tP[ie];
        t.Po = A.tP[io];
        t.De = t.Do = t.Me = t.Mo = T(0);
        if (VEC) {
            t.De = A.tD[ie];
            t.Do = A.tD[io];
            t.Me = A.tM[ie];
            t.Mo = A.tM[io];
        }
        return t;
    };
    const int nsteps = ((A.lmax - m) / SKC + 1) * (SKC / 8);
    // two register sets, selected by the parity of the unrolled step (four
    // steps per chunk), so that no copy of a set in flight is needed: a copy
    // would wait for the loads in the step that issues them
    int g = 0;
    Tab tb[2] = {};
    if constexpr (PF) {
        tb[0] = ltab(kk);
        load(m);
    }
    for (int lbase = m; lbase <= A.lmax; lbase += SKC) {
        if constexpr (PF) {
            store();
            __syncthreads();
            if (lbase + SKC <= A.lmax) load(lbase + SKC);
        } else {
            load(lbase);
            store();
            __syncthreads();
        }
#pragma unroll
        for (int st = 0; st < SKC / 8; st++, g++) {
            Tab cur = {};
            if constexpr (PF) {
                // the last step reloads row kk, which every table holds
                const int gn = (g + 1 < nsteps) ? g + 1 : 0;
                tb[(st + 1) & 1] = ltab(4 * gn + kk);
                cur = tb[st & 1];
            } else {
                const long ie = offe + (long)(4 * g + kk) * A.nthp + th0 + col;
                const long io = offo + (long)(4 * g + kk) * A.nthp + th0 + col;
                cur.Pe = A.tP[ie];
                cur.Po = A.tP[io];
                if (VEC) {
                    cur.De = A.tD[ie];
                    cur.Do = A.tD[io];
                    cur.Me = A.tM[ie];
                    cur.Mo = A.tM[io];
                }
            }
            const T Pe = cur.Pe, Po = cur.Po, De = cur.De, Do = cur.Do, Me = cur.Me, Mo = cur.Mo;
            const int lle = 8 * st + 2 * kk, llo = lle + 1;
            // steps whose eight l all lie above lmax hold only padding
            if (lbase + 8 * st <= A.lmax) {
#pragma unroll
                for (int b = 0; b < JB; b++) {
                    const int j = b * 16 + col;
                    acc[0][0][b] = Mf<T>::mma(Pe, Bs[0][lle][j], acc[0][0][b]);
                    acc[0][1][b] = Mf<T>::mma(Po, Bs[0][llo][j], acc[0][1][b]);
                    if (VEC) {
                        const T Se = Bs[1][lle][j], So = Bs[1][llo][j];
                        const T Te = Bs[2][lle][j], To = Bs[2][llo][j];
                        const T iSe = sgn_i * Bs[1][lle][j ^ 1], iSo = sgn_i * Bs[1][llo][j ^ 1];
                        const T iTe = sgn_i * Bs[2][lle][j ^ 1], iTo = sgn_i * Bs[2][llo][j ^ 1];
                        // v_theta: symmetric = dP_o S_o + mP_e (iT)_e, antisymmetric = dP_e S_e + mP_o (iT)_o
                        acc[1][0][b] = Mf<T>::mma(Do, So, acc[1][0][b]);
                        acc[1][0][b] = Mf<T>::mma(Me, iTe, acc[1][0][b]);
                        acc[1][1][b] = Mf<T>::mma(De, Se, acc[1][1][b]);
                        acc[1][1][b] = Mf<T>::mma(Mo, iTo, acc[1][1][b]);
                        // v_phi: symmetric = mP_e (iS)_e - dP_o T_o, antisymmetric = mP_o (iS)_o - dP_e T_e
                        acc[2][0][b] = Mf<T>::mma(Me, iSe, acc[2][0][b]);
                        acc[2][0][b] = Mf<T>::mma(Do, -To, acc[2][0][b]);
                        acc[2][1][b] = Mf<T>::mma(Mo, iSo, acc[2][1][b]);
                        acc[2][1][b] = Mf<T>::mma(De, -Te, acc[2][1][b]);
                    }
                }
            }
        }
        __syncthreads();
    }

    // north = sym + anti, south = sym - anti
#pragma unroll
    for (int f = 0; f < NF; f++) {
        double *out = (double *)A.out[f];
#pragma unroll
        for (int b = 0; b < JB; b++) {
            const int j = b * 16 + col;
            const int k = k0 + (j >> 1), c = j & 1;
            if (k >= A.nplanes) continue;
            const long base = ((long)k * L + m) * A.nlat;
#pragma unroll
            for (int v = 0; v < 4; v++) {
                const int th = th0 + Mf<T>::row(kk, v);
                if (th >= A.nth) continue;
                const double sy = acc[f][0][b][v], an = acc[f][1][b][v];
                out[2 * (base + th) + c] = sy + an;
                out[2 * (base + A.nlat - 1 - th) + c] = sy - an;
            }
        }
    }
}

template <typename T, bool VEC>
__global__ __launch_bounds__(256, 4) void synthesis_kernel(SynArgs<T> A)
{
    using Cfg = SynCfg<VEC, syn_jb<T, VEC>()>;
    __shared__ T Bs[Cfg::NF][SKC][Cfg::JTP];
    // per-XCD queue entry (m, theta tile); the plane tiles of one entry (of
    // every field set) run back to back on one XCD, so its table rows are
    // read from HBM once
    const int xcd = blockIdx.x % kXcds, i = blockIdx.x / kXcds;
    const int nj = A.n_jt * A.nset, jts = i % nj;
    const int ent = A.queue[(i / nj) * kXcds + xcd];
    if (ent < 0) return;
    SynArgs<T> S = A;
    if (jts >= A.n_jt)
#pragma unroll
        for (int f = 0; f < 3; f++) {
            S.in[f] = A.in2[f];
            S.out[f] = A.out2[f];
        }
    syn_tile<T, VEC, Cfg::JB>(S, ent & 0xffff, ent >> 16, jts % A.n_jt, Bs);
}

// Synthesis of the field set (q, s, t) -> (f0, f1, f2), and of the second
// set (in2 -> out2) in the same launch when in2 is given (vector form only).
template <typename T>
static int launch_synthesis_t(Plan *p, const T *tP, const T *tD, const T *tM, bool vec, int nplanes,
                              const double2 *q, const double2 *s, const double2 *t, long ld, double2 *f0,
                              double2 *f1, double2 *f2, const double2 *const *in2 = nullptr,
                              double2 *const *out2 = nullptr)
{
    SynArgs<T> A = {};
    A.nset = (vec && in2) ? 2 : 1;
    for (int f = 0; f < 3; f++) {
        A.in2[f] = A.nset == 2 ? in2[f] : nullptr;
        A.out2[f] = A.nset == 2 ? out2[f] : nullptr;
    }
    A.tP = tP;
    A.tD = tD;
    A.tM = tM;
    A.off = p->d_syn_off;
    A.in[0] = q;
    A.in[1] = vec ? s : q;
    A.in[2] = vec ? t : q;
    A.out[0] = f0;
    A.out[1] = vec ? f1 : f0;
    A.out[2] = vec ? f2 : f0;
    A.ld = ld;
    A.lmax = p->lmax;
    A.nlat = p->nlat;
    A.nth = p->nth;
    A.nthp = p->nthp;
    A.nplanes = nplanes;
    A.queue = p->d_syn_q;
    const int JT = vec ? SynCfg<true, syn_jb<T, true>()>::JT : SynCfg<false, syn_jb<T, false>()>::JT;
    A.n_jt = (2 * nplanes + JT - 1) / JT;
    dim3 grid((unsigned)(kXcds * p->syn_qlen * A.n_jt * A.nset));
    if (vec)
        hipLaunchKernelGGL((synthesis_kernel<T, true>), grid, dim3(256), 0, p->stream, A);
    else
        hipLaunchKernelGGL((synthesis_kernel<T, false>), grid, dim3(256), 0, p->stream, A);
    RD_LAUNCH_CHECK();
    return RD_OK;
}

int launch_synthesis(Plan *p, int prec, bool vec, int nplanes, const double2 *q, const double2 *s,
                     const double2 *t, long ld, double2 *f0, double2 *f1, double2 *f2)
{
    if (prec == RD_PREC_FP64)
        return launch_synthesis_t<double>(p, p->syn_P, p->syn_D, p->syn_M, vec, nplanes, q, s, t, ld, f0, f1,
                                          f2);
    if (prec == RD_PREC_OZAKI) return launch_synthesis_oz(p, vec, nplanes, q, s, t, ld, f0, f1, f2);
    int rc = ensure_precision_tables(p, prec);
    if (rc) return rc;
    if (prec == RD_PREC_FP32)
        return launch_synthesis_t<float>(p, p->syn32_P, p->syn32_D, p->syn32_M, vec, nplanes, q, s, t, ld, f0,
                                         f1, f2);
    return launch_synthesis_bf3(p, vec, nplanes, q, s, t, ld, f0, f1, f2);
}

int launch_synthesis2(Plan *p, int prec, int nplanes, const double2 *const *in_a, double2 *const *out_a,
                      const double2 *const *in_b, double2 *const *out_b, long ld)
{
    if (prec == RD_PREC_FP64)
        return launch_synthesis_t<double>(p, p->syn_P, p->syn_D, p->syn_M, true, nplanes, in_a[0], in_a[1], in_a[2],
                                          ld, out_a[0], out_a[1], out_a[2], in_b, out_b);
    if (prec == RD_PREC_FP32) {
        int rc = ensure_precision_tables(p, prec);
        if (rc) return rc;
        return launch_synthesis_t<float>(p, p->syn32_P, p->syn32_D, p->syn32_M, true, nplanes, in_a[0], in_a[1],
                                         in_a[2], ld, out_a[0], out_a[1], out_a[2], in_b, out_b);
    }
    if (prec == RD_PREC_BF16X3) {
        int rc = ensure_precision_tables(p, prec);
        if (rc) return rc;
        return launch_synthesis_bf3(p, true, nplanes, in_a[0], in_a[1], in_a[2], ld, out_a[0], out_a[1], out_a[2],
                                    in_b, out_b);
    }
    int rc = launch_synthesis(p, prec, true, nplanes, in_a[0], in_a[1], in_a[2], ld, out_a[0], out_a[1], out_a[2]);
    if (!rc) rc = launch_synthesis(p, prec, true, nplanes, in_b[0], in_b[1], in_b[2], ld, out_b[0], out_b[1], out_b[2]);
    return rc;
}

// -------------------------------------------------------------- analysis
// Workgroup tile: one m, 32 rows l_p per parity (2 waves x 16), AJT real
// columns (2 waves x JB x 16).  The theta loop runs over the northern nodes
// in chunks of 16.
constexpr int ATC = 16;

// 16-column blocks per wave of the one-problem kernels.  Vector tiles keep
// one block: FP64 then fits 128 VGPRs at 4 waves per SIMD (0.54 against
// 0.66 ms with two blocks at 2 waves, lmax 255, 256 planes) and FP32 gains
// in the same way (0.45 against 0.50 ms); the table rows are read twice as
// often from L2, which costs less than the latency hiding it buys.
template <typename T, bool VEC>
constexpr int ana_jb()
{
    return VEC ? 1 : 4;
}

template <bool VEC, int JBV = (VEC ? 2 : 4)>
struct AnaCfg {
    static constexpr int NF = VEC ? 3 : 1;
    static constexpr int JB = JBV;
    static constexpr int JT = 2 * 16 * JB;
    static constexpr int JTP = JT + 2;
};

template <typename T>
struct AnaArgs {
    const T *tP, *tD, *tM;
    const long *off;
    const int *ncol;
    const int *queue;
    const double2 *in[3];
    double2 *out[3];
    long ld;
    int lmax, nlat, nth, nthp16, nplanes, n_jt;
};

// One analysis tile (m, l tile lt, plane tile jt) of one problem, staged
// through registers with in-step table loads: the tile of the persistent
// batch kernels (the one-problem kernels use ana_tile_dma below); T as in
// syn_tile.
template <typename T, bool VEC, int JBV>
__device__ __forceinline__ void ana_tile(const AnaArgs<T> &A, int m, int lt, int jt,
                                         T (&Gn)[AnaCfg<VEC, JBV>::NF][ATC][AnaCfg<VEC, JBV>::JTP],
                                         T (&Gs)[AnaCfg<VEC, JBV>::NF][ATC][AnaCfg<VEC, JBV>::JTP])
{
    using Cfg = AnaCfg<VEC, JBV>;
    using V = typename Mf<T>::V;
    constexpr int NF = Cfg::NF, JB = Cfg::JB, JT = Cfg::JT;

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int L = A.lmax + 1;
    const int wl = wave & 1, wj = wave >> 1;
    const int k0 = jt * (JT / 2);
    const int col = lane & 15, kk = lane >> 4;
    const T sgn_mi = (col & 1) ? T(-1) : T(1);   // (-i X)_j = sgn_mi * X_{j^1}
    const int ncol = A.ncol[m];
    const int lp0 = lt * 32 + wl * 16;
    // a wave whose 16 rows l_p lie above both parities' row counts (the
    // even count (L - m + 1) / 2 is the larger) holds only padding: it stages
    // and waits at the barriers but skips the MFMAs
    const bool live = lp0 < (L - m + 1) / 2;

    V acc[NF][2][JB];
#pragma unroll
    for (int f = 0; f < NF; f++)
#pragma unroll
        for (int s = 0; s < 2; s++)
#pragma unroll
            for (int b = 0; b < JB; b++) acc[f][s][b] = (V){0, 0, 0, 0};

    const long offe = A.off[2 * m], offo = A.off[2 * m + 1];
    struct Tab {
        T Pe, Po, De, Do, Me, Mo;
    };
    // table values at theta row t (32-bit element offsets from the uniform
    // table bases, as in syn_tile)
    auto ltab = [&](int t) {
        Tab v;
        const unsigned rowoff = (unsigned)(t * ncol + lp0 + col);
        const unsigned ie = (unsigned)offe + rowoff, io = (unsigned)offo + rowoff;
        v.Pe = A.tP[ie];
        v.Po = A.tP[io];
        v.De = v.Do = v.Me = v.Mo = T(0);
        if (VEC) {
            v.De = A.tD[ie];
            v.Do = A.tD[io];
            v.Me = A.tM[ie];
            v.Mo = A.tM[io];
        }
        return v;
    };
    for (int tc = 0; tc < A.nth; tc += ATC) {
        constexpr int PER = (JT / 2) * ATC;
        for (int e = tid; e < NF * PER; e += 256) {
            int f = e / PER, r = e % PER;
            int pl = r / ATC, tl = r % ATC;
            int th = tc + tl, k = k0 + pl;
            double2 vn = make_double2(0.0, 0.0), vs = vn;
            if (th < A.nth && k < A.nplanes) {
                const long base = ((long)k * L + m) * A.nlat;
                vn = A.in[f][base + th];
                vs = A.in[f][base + A.nlat - 1 - th];
            }
            Gn[f][tl][2 * pl] = (T)vn.x;
            Gn[f][tl][2 * pl + 1] = (T)vn.y;
            Gs[f][tl][2 * pl] = (T)vs.x;
            Gs[f][tl][2 * pl + 1] = (T)vs.y;
        }
        __syncthreads();
#pragma unroll
        for (int st = 0; st < ATC / 4; st++) {
            const int tl = 4 * st + kk;
            const Tab cur = ltab(tc + tl);
            const T Pe = cur.Pe, Po = cur.Po, De = cur.De, Do = cur.Do, Me = cur.Me, Mo = cur.Mo;
            if (live) {
#pragma unroll
                for (int b = 0; b < JB; b++) {
                    const int j = (wj * JB + b) * 16 + col;
                    const T rn = Gn[0][tl][j], rs = Gs[0][tl][j];
                    acc[0][0][b] = Mf<T>::mma(Pe, rn + rs, acc[0][0][b]);
                    acc[0][1][b] = Mf<T>::mma(Po, rn - rs, acc[0][1][b]);
                    if (VEC) {
                        const T tn = Gn[1][tl][j], ts = Gs[1][tl][j];
                        const T pn = Gn[2][tl][j], ps = Gs[2][tl][j];
                        const T mtn = sgn_mi * Gn[1][tl][j ^ 1], mts = sgn_mi * Gs[1][tl][j ^ 1];
                        const T mpn = sgn_mi * Gn[2][tl][j ^ 1], mps = sgn_mi * Gs[2][tl][j ^ 1];
                        // S_e = dP_e (t_n - t_s) + mP_e (-i)(p_n + p_s)
                        acc[1][0][b] = Mf<T>::mma(De, tn - ts, acc[1][0][b]);
                        acc[1][0][b] = Mf<T>::mma(Me, mpn + mps, acc[1][0][b]);
                        // S_o = dP_o (t_n + t_s) + mP_o (-i)(p_n - p_s)
                        acc[1][1][b] = Mf<T>::mma(Do, tn + ts, acc[1][1][b]);
                        acc[1][1][b] = Mf<T>::mma(Mo, mpn - mps, acc[1][1][b]);
                        // T_e = mP_e (-i)(t_n + t_s) - dP_e (p_n - p_s)
                        acc[2][0][b] = Mf<T>::mma(Me, mtn + mts, acc[2][0][b]);
                        acc[2][0][b] = Mf<T>::mma(De, ps - pn, acc[2][0][b]);
                        // T_o = mP_o (-i)(t_n - t_s) - dP_o (p_n + p_s)
                        acc[2][1][b] = Mf<T>::mma(Mo, mtn - mts, acc[2][1][b]);
                        acc[2][1][b] = Mf<T>::mma(Do, -(pn + ps), acc[2][1][b]);
                    }
                }
            }
        }
        __syncthreads();
    }

    const long lm0 = lm_index(A.lmax, m, m);
#pragma unroll
    for (int f = 0; f < NF; f++) {
        double *out = (double *)A.out[f];
#pragma unroll
        for (int par = 0; par < 2; par++)
#pragma unroll
            for (int b = 0; b < JB; b++) {
                const int j = (wj * JB + b) * 16 + col;
                const int k = k0 + (j >> 1), c = j & 1;
                if (k >= A.nplanes) continue;
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int lp = lp0 + Mf<T>::row(kk, v);
                    const int l = m + par + 2 * lp;
                    if (l > A.lmax) continue;
                    double val = acc[f][par][b][v];
                    if (m == 0 && c == 1) val = 0.0;
                    out[2 * (lm0 + (l - m) + k * A.ld) + c] = val;
                }
            }
    }
}

// Analysis tile of the one-problem kernels, FP64 and FP32 forms, whose grid
// chunks (8 northern rows and their southern mirrors of NF fields and P
// planes, P = 16 for vector tiles, 64 for scalar ones) move from global
// memory straight into LDS (global_load_lds_dwordx4), double-buffered: the
// moves of chunk c + 1 are issued when chunk c starts, so the staging latency
// overlaps the MFMAs and one barrier per chunk suffices.  The moves are
// inline asm (lds_move16) with explicit waits.  Every table load a chunk
// needs (its second step's values and the next chunk's first) is issued
// before that chunk's moves, and the moves are the last vector-memory
// operations of the chunk, so vmcnt(0) at its end waits for exactly what was
// issued before the moves plus the moves themselves: no count of the loads
// the compiler emits enters the wait.  A plain s_barrier then publishes the
// moves (the fence of __syncthreads is not needed).  One move writes 64
// consecutive 16-byte LDS slots (8 planes x 8 rows) from a wave-uniform base;
// lane 8 p + s loads row (s - p) mod 8 of plane p, so that each plane's eight
// lanes read one 128-byte segment and the MFMA operand reads (one row, eight
// planes) fall on eight different bank groups.  Slot of (field f, north/south
// h, plane pl, row t): 8 P (2 f + h) + 8 pl + (t + pl) mod 8.  LDS holds FP64;
// the FP32 form rounds on the operand read.
__device__ double2 ana_zero16 = {0.0, 0.0};

// One 16-byte LDS move per lane: lane i's global element lands at slot i
// after the wave-uniform LDS base.  Inline asm, because the compiler, seeing
// an LDS move in flight, waits for it before the next barrier with a fence
// (vmcnt(0)); the caller orders its moves with explicit vmcnt waits instead.
// The instruction has no register result: its address operand is read at
// issue.  Waits the compiler places for its own loads stay valid with these
// moves in the count: vmcnt retires in order, so extra operations only make
// them stricter.  An LDS move reads M0 one wait state after an SALU write of
// M0 at the earliest (a gfx9 hazard), hence the s_nop between the two; the
// compiler itself uses M0 for nothing else in these kernels (the CPU tests
// check the built assembly for both).
__device__ __forceinline__ void lds_move16(const double2 *src, double2 *lds)
{
    const unsigned base = (unsigned)(size_t)(__attribute__((address_space(3))) void *)lds;
    asm volatile("s_mov_b32 m0, %0\n\ts_nop 0\n\tglobal_load_lds_dwordx4 %1, off" ::"s"(base), "v"(src)
                 : "memory", "m0");
}

// s_waitcnt vmcnt(n) with expcnt and lgkmcnt left free (n < 64)
#define RD_WAIT_VMCNT(n) __builtin_amdgcn_s_waitcnt(((n) & 15) | (((n) >> 4) << 14) | (7 << 4) | (15 << 8))

constexpr int ADC = 8;   // rows per LDS-move chunk

template <bool VEC>
struct AnaDma {
    static constexpr int NF = VEC ? 3 : 1;
    static constexpr int JB = VEC ? 1 : 4;    // 16-column blocks per wave
    static constexpr int P = 16 * JB;         // planes per workgroup
    static constexpr int SLOTS = NF * 16 * P; // double2 per chunk buffer
    static constexpr int NB = SLOTS / 64;     // moves per chunk
    static_assert(NB % 4 == 0, "moves must split evenly over the four waves");
};

template <typename T, bool VEC>
__device__ __forceinline__ void ana_tile_dma(const AnaArgs<T> &A, int m, int lt, int jt,
                                             double2 (&G)[2][AnaDma<VEC>::SLOTS])
{
    using V = typename Mf<T>::V;
    using D = AnaDma<VEC>;
    constexpr int NF = D::NF, JB = D::JB, P = D::P;
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int L = A.lmax + 1;
    const int wl = wave & 1, wj = wave >> 1;
    const int k0 = jt * P;
    const int col = lane & 15, kk = lane >> 4;
    const T sgn_mi = (col & 1) ? T(-1) : T(1);   // (-i X)_j = sgn_mi * X_{j^1}
    const int ncol = A.ncol[m];
    const int lp0 = lt * 32 + wl * 16;
    const bool live = lp0 < (L - m + 1) / 2;

    V acc[NF][2][JB];
#pragma unroll
    for (int f = 0; f < NF; f++)
#pragma unroll
        for (int s = 0; s < 2; s++)
#pragma unroll
            for (int b = 0; b < JB; b++) acc[f][s][b] = (V){0, 0, 0, 0};

    // wave w issues moves NB/4 w .. NB/4 (w + 1) - 1 of a chunk; move
    // ib = (2 f + h) P/8 + pb holds planes 8 pb .. 8 pb + 7 of field f,
    // rows north (h = 0) or south
    auto issue = [&](int tc, int buf) {
        const int p = lane >> 3, t = ((lane & 7) - p) & 7;
#pragma unroll
        for (int i = 0; i < D::NB / 4; i++) {
            const int ib = (D::NB / 4) * wave + i;   // wave-uniform
            const int fh = ib / (P / 8), pb = ib % (P / 8);
            const int f = fh >> 1, h = fh & 1;
            const int th = tc + t, k = k0 + 8 * pb + p;
            const double2 *src = &ana_zero16;
            if (th < A.nth && k < A.nplanes)
                src = A.in[f] + (((long)k * L + m) * A.nlat + (h ? A.nlat - 1 - th : th));
            lds_move16(src, &G[buf][64 * ib]);
        }
    };
    const long offe = A.off[2 * m], offo = A.off[2 * m + 1];
    struct Tab {
        T Pe, Po, De, Do, Me, Mo;
    };
    auto ltab = [&](int t) {
        Tab v;
        const unsigned rowoff = (unsigned)(t * ncol + lp0 + col);
        const unsigned ie = (unsigned)offe + rowoff, io = (unsigned)offo + rowoff;
        v.Pe = A.tP[ie];
        v.Po = A.tP[io];
        v.De = v.Do = v.Me = v.Mo = T(0);
        if (VEC) {
            v.De = A.tD[ie];
            v.Do = A.tD[io];
            v.Me = A.tM[ie];
            v.Mo = A.tM[io];
        }
        return v;
    };
    static_assert(ADC == 8, "two MFMA steps of four rows per chunk");
    const int nsteps = ((A.nth + ADC - 1) / ADC) * (ADC / 4);
    int buf = 0;
    // tb[0]: this chunk's first step (already landed), tb[1]: its second,
    // tb[2]: the next chunk's first (the last chunk reloads row kk)
    Tab tb[3];
    tb[0] = ltab(kk);
    issue(0, 0);
    RD_WAIT_VMCNT(0);
    __builtin_amdgcn_s_barrier();
    for (int tc = 0, g = 0; tc < A.nth; tc += ADC, g += ADC / 4, buf ^= 1) {
        tb[1] = ltab(4 * (g + 1) + kk);
        tb[2] = ltab(4 * ((g + 2 < nsteps) ? g + 2 : 0) + kk);
        __builtin_amdgcn_sched_barrier(0);
        if (tc + ADC < A.nth) issue(tc + ADC, buf ^ 1);
        __builtin_amdgcn_sched_barrier(0);
        const double *Gd = (const double *)G[buf];
#pragma unroll
        for (int st = 0; st < ADC / 4; st++) {
            const int tl = 4 * st + kk;
            const Tab cur = tb[st];
            if (live) {
#pragma unroll
                for (int b = 0; b < JB; b++) {
                    // column j = 16 (wj JB + b) + col: plane j / 2, part j & 1
                    const int pl = (16 * (wj * JB + b) + col) >> 1, c = col & 1;
                    const int sl = 8 * pl + ((tl + pl) & 7);
                    auto at = [&](int f, int h, int part) {
                        return (T)Gd[2 * (8 * P * (2 * f + h) + sl) + part];
                    };
                    const T rn = at(0, 0, c), rs = at(0, 1, c);
                    acc[0][0][b] = Mf<T>::mma(cur.Pe, rn + rs, acc[0][0][b]);
                    acc[0][1][b] = Mf<T>::mma(cur.Po, rn - rs, acc[0][1][b]);
                    if (VEC) {
                        const T tn = at(1, 0, c), ts = at(1, 1, c);
                        const T pn = at(2, 0, c), ps = at(2, 1, c);
                        const T mtn = sgn_mi * at(1, 0, c ^ 1), mts = sgn_mi * at(1, 1, c ^ 1);
                        const T mpn = sgn_mi * at(2, 0, c ^ 1), mps = sgn_mi * at(2, 1, c ^ 1);
                        // S_e = dP_e (t_n - t_s) + mP_e (-i)(p_n + p_s)
                        acc[1][0][b] = Mf<T>::mma(cur.De, tn - ts, acc[1][0][b]);
                        acc[1][0][b] = Mf<T>::mma(cur.Me, mpn + mps, acc[1][0][b]);
                        // S_o = dP_o (t_n + t_s) + mP_o (-i)(p_n - p_s)
                        acc[1][1][b] = Mf<T>::mma(cur.Do, tn + ts, acc[1][1][b]);
                        acc[1][1][b] = Mf<T>::mma(cur.Mo, mpn - mps, acc[1][1][b]);
                        // T_e = mP_e (-i)(t_n + t_s) - dP_e (p_n - p_s)
                        acc[2][0][b] = Mf<T>::mma(cur.Me, mtn + mts, acc[2][0][b]);
                        acc[2][0][b] = Mf<T>::mma(cur.De, ps - pn, acc[2][0][b]);
                        // T_o = mP_o (-i)(t_n - t_s) - dP_o (p_n + p_s)
                        acc[2][1][b] = Mf<T>::mma(cur.Mo, mtn - mts, acc[2][1][b]);
                        acc[2][1][b] = Mf<T>::mma(cur.Do, -(pn + ps), acc[2][1][b]);
                    }
                }
            }
        }
        __builtin_amdgcn_sched_barrier(0);
        RD_WAIT_VMCNT(0);   // this wave's moves of the next chunk (its last memory operations) have landed
        __builtin_amdgcn_s_barrier();
        tb[0] = tb[2];
    }

    const long lm0 = lm_index(A.lmax, m, m);
#pragma unroll
    for (int f = 0; f < NF; f++) {
        double *out = (double *)A.out[f];
#pragma unroll
        for (int par = 0; par < 2; par++)
#pragma unroll
            for (int b = 0; b < JB; b++) {
                const int j = (wj * JB + b) * 16 + col;
                const int k = k0 + (j >> 1), c = j & 1;
                if (k >= A.nplanes) continue;
#pragma unroll
                for (int v = 0; v < 4; v++) {
                    const int lp = lp0 + Mf<T>::row(kk, v);
                    const int l = m + par + 2 * lp;
                    if (l > A.lmax) continue;
                    double val = acc[f][par][b][v];
                    if (m == 0 && c == 1) val = 0.0;
                    out[2 * (lm0 + (l - m) + k * A.ld) + c] = val;
                }
            }
    }
}

template <typename T, bool VEC>
__global__ __launch_bounds__(256, 4) void analysis_kernel(AnaArgs<T> A)
{
    // per-XCD queue entry (m, l tile); the plane tiles of one entry run back
    // to back on one XCD
    const int xcd = blockIdx.x % kXcds, i = blockIdx.x / kXcds;
    const int jt = i % A.n_jt;
    const int ent = A.queue[(i / A.n_jt) * kXcds + xcd];
    if (ent < 0) return;
    static_assert(AnaDma<VEC>::JB == ana_jb<T, VEC>(), "the LDS-move tile and the launch agree on the tiling");
    __shared__ double2 G[2][AnaDma<VEC>::SLOTS];
    ana_tile_dma<T, VEC>(A, ent & 0xffff, ent >> 16, jt, G);
}

template <typename T>
static int launch_analysis_t(Plan *p, const T *tP, const T *tD, const T *tM, bool vec, int nplanes,
                             const double2 *g0, const double2 *g1, const double2 *g2, double2 *q, double2 *s,
                             double2 *t, long ld)
{
    AnaArgs<T> A;
    A.tP = tP;
    A.tD = tD;
    A.tM = tM;
    A.off = p->d_ana_off;
    A.ncol = p->d_ana_ncol;
    A.queue = p->d_ana_q;
    A.in[0] = g0;
    A.in[1] = vec ? g1 : g0;
    A.in[2] = vec ? g2 : g0;
    A.out[0] = q;
    A.out[1] = vec ? s : q;
    A.out[2] = vec ? t : q;
    A.ld = ld;
    A.lmax = p->lmax;
    A.nlat = p->nlat;
    A.nth = p->nth;
    A.nthp16 = p->nthp16;
    A.nplanes = nplanes;
    const int JT = vec ? AnaCfg<true, ana_jb<T, true>()>::JT : AnaCfg<false, ana_jb<T, false>()>::JT;
    A.n_jt = (2 * nplanes + JT - 1) / JT;
    dim3 grid((unsigned)(kXcds * p->ana_qlen * A.n_jt));
    if (vec)
        hipLaunchKernelGGL((analysis_kernel<T, true>), grid, dim3(256), 0, p->stream, A);
    else
        hipLaunchKernelGGL((analysis_kernel<T, false>), grid, dim3(256), 0, p->stream, A);
    RD_LAUNCH_CHECK();
    return RD_OK;
}

int launch_analysis(Plan *p, int prec, bool vec, int nplanes, const double2 *g0, const double2 *g1,
                    const double2 *g2, double2 *q, double2 *s, double2 *t, long ld)
{
    if (prec == RD_PREC_FP64)
        return launch_analysis_t<double>(p, p->ana_P, p->ana_D, p->ana_M, vec, nplanes, g0, g1, g2, q, s, t,
                                         ld);
    if (prec == RD_PREC_OZAKI) return launch_analysis_oz(p, vec, nplanes, g0, g1, g2, q, s, t, ld);
    int rc = ensure_precision_tables(p, prec);
    if (rc) return rc;
    if (prec == RD_PREC_FP32)
        return launch_analysis_t<float>(p, p->ana32_P, p->ana32_D, p->ana32_M, vec, nplanes, g0, g1, g2, q, s,
                                        t, ld);
    return launch_analysis_bf3(p, vec, nplanes, g0, g1, g2, q, s, t, ld);
}

// ------------------------------------------------ batched persistent launches
// Many independent problems through one launch per direction: every
// workgroup pulls work items (problem, m, tile, plane tile) from a counter
// until the list is exhausted; the list is ordered largest first so that the
// uneven m-blocks of problems of different lmax balance across the CUs.
__global__ void zero_counters_kernel(int *c, int n)
{
    if ((int)threadIdx.x < n) c[threadIdx.x] = 0;
}

template <typename T, int JBV>
__global__ __launch_bounds__(256) void batch_synthesis_kernel(const SynArgs<T> *probs, const int4 *items,
                                                              int nitems, int *counter)
{
    using Cfg = SynCfg<false, JBV>;
    __shared__ T Bs[Cfg::NF][SKC][Cfg::JTP];
    __shared__ int next;
    for (;;) {
        if (threadIdx.x == 0) next = atomicAdd(counter, 1);
        __syncthreads();
        const int it = next;
        __syncthreads();
        if (it >= nitems) return;
        const int4 w = items[it];
        const SynArgs<T> A = probs[w.x];
        syn_tile<T, false, JBV, false>(A, w.y, w.z, w.w, Bs);
    }
}

template <typename T, int JBV>
__global__ __launch_bounds__(256) void batch_analysis_kernel(const AnaArgs<T> *probs, const int4 *items,
                                                             int nitems, int *counter)
{
    using Cfg = AnaCfg<false, JBV>;
    __shared__ T Gn[Cfg::NF][ATC][Cfg::JTP];
    __shared__ T Gs[Cfg::NF][ATC][Cfg::JTP];
    __shared__ int next;
    for (;;) {
        if (threadIdx.x == 0) next = atomicAdd(counter, 1);
        __syncthreads();
        const int it = next;
        __syncthreads();
        if (it >= nitems) return;
        const int4 w = items[it];
        const AnaArgs<T> A = probs[w.x];
        ana_tile<T, false, JBV>(A, w.y, w.z, w.w, Gn, Gs);
    }
}

// Vector forms of the persistent launches (the nonlinear term of many
// problems): the same work-item loop over the (Q,S,T) tiles.
template <typename T, int JBV>
__global__ __launch_bounds__(256) void batch_vsynthesis_kernel(const SynArgs<T> *sets, const int4 *items,
                                                               int nitems, int *counter)
{
    using Cfg = SynCfg<true, JBV>;
    __shared__ T Bs[Cfg::NF][SKC][Cfg::JTP];
    __shared__ int next;
    for (;;) {
        if (threadIdx.x == 0) next = atomicAdd(counter, 1);
        __syncthreads();
        const int it = next;
        __syncthreads();
        if (it >= nitems) return;
        const int4 w = items[it];
        const SynArgs<T> A = sets[w.x];
        syn_tile<T, true, JBV, false>(A, w.y, w.z, w.w, Bs);
    }
}

template <typename T, int JBV>
__global__ __launch_bounds__(256) void batch_vanalysis_kernel(const AnaArgs<T> *probs, const int4 *items,
                                                              int nitems, int *counter)
{
    using Cfg = AnaCfg<true, JBV>;
    __shared__ T Gn[Cfg::NF][ATC][Cfg::JTP];
    __shared__ T Gs[Cfg::NF][ATC][Cfg::JTP];
    __shared__ int next;
    for (;;) {
        if (threadIdx.x == 0) next = atomicAdd(counter, 1);
        __syncthreads();
        const int it = next;
        __syncthreads();
        if (it >= nitems) return;
        const int4 w = items[it];
        const AnaArgs<T> A = probs[w.x];
        ana_tile<T, true, JBV>(A, w.y, w.z, w.w, Gn, Gs);
    }
}

int batch_leg_init(BatchLeg *b, Plan *const *plans, int nprob, int prec, int jb, int vec, int nsets)
{
    if (prec != RD_PREC_FP64 && prec != RD_PREC_FP32) return RD_ERR_ARG;
    if (vec ? (jb != 1 && jb != 2) : (jb != 2 && jb != 4)) return RD_ERR_ARG;
    if (nsets < 1 || (!vec && nsets != 1)) return RD_ERR_ARG;
    b->nprob = nprob;
    b->prec = prec;
    b->jb = jb;
    b->vec = vec;
    b->nsets = nsets;
    b->arg_bytes_syn = prec == RD_PREC_FP64 ? sizeof(SynArgs<double>) : sizeof(SynArgs<float>);
    b->arg_bytes_ana = prec == RD_PREC_FP64 ? sizeof(AnaArgs<double>) : sizeof(AnaArgs<float>);
    std::vector<int4> si, ai;
    std::vector<long> sc, ac;
    for (int p = 0; p < nprob; p++) {
        Plan *pl = plans[p];
        if (prec == RD_PREC_FP32) {
            int rc = ensure_precision_tables(pl, prec);
            if (rc) return rc;
        }
        const int L = pl->L;
        const int np = vec ? pl->nr : pl->nz;
        const int sjt = (2 * np + 16 * jb - 1) / (16 * jb), ntt = pl->nthp / 64;
        const int ajt = (2 * np + 32 * jb - 1) / (32 * jb);
        for (int m = 0; m < L; m++) {
            for (int set = 0; set < nsets; set++)
                for (int tt = 0; tt < ntt; tt++)
                    for (int jt = 0; jt < sjt; jt++) {
                        si.push_back(make_int4(p * nsets + set, m, tt, jt));
                        sc.push_back((L - m + SKC - 1) / SKC);
                    }
            const int ncol = rup((L - m + 1) / 2, 32);
            for (int lt = 0; lt < ncol / 32; lt++)
                for (int jt = 0; jt < ajt; jt++) {
                    ai.push_back(make_int4(p, m, lt, jt));
                    ac.push_back((pl->nth + ATC - 1) / ATC);
                }
        }
    }
    auto order = [](std::vector<int4> &v, const std::vector<long> &cost) {
        std::vector<size_t> idx(v.size());
        for (size_t i = 0; i < idx.size(); i++) idx[i] = i;
        std::stable_sort(idx.begin(), idx.end(), [&](size_t a, size_t c) { return cost[a] > cost[c]; });
        std::vector<int4> o(v.size());
        for (size_t i = 0; i < idx.size(); i++) o[i] = v[idx[i]];
        v.swap(o);
    };
    order(si, sc);
    order(ai, ac);
    b->n_syn_items = (int)si.size();
    b->n_ana_items = (int)ai.size();
    RD_HIP(hipMalloc(&b->d_syn_items, sizeof(int4) * si.size()));
    RD_HIP(hipMalloc(&b->d_ana_items, sizeof(int4) * ai.size()));
    RD_HIP(hipMalloc(&b->d_syn_args, b->arg_bytes_syn * nprob * nsets));
    RD_HIP(hipMalloc(&b->d_ana_args, b->arg_bytes_ana * nprob));
    RD_HIP(hipMalloc(&b->d_counter, 2 * sizeof(int)));
    RD_HIP(hipMemcpy(b->d_syn_items, si.data(), sizeof(int4) * si.size(), hipMemcpyHostToDevice));
    RD_HIP(hipMemcpy(b->d_ana_items, ai.data(), sizeof(int4) * ai.size(), hipMemcpyHostToDevice));
    b->h_syn_args = calloc((size_t)nprob * nsets, b->arg_bytes_syn);
    b->h_ana_args = calloc(nprob, b->arg_bytes_ana);
    if (!b->h_syn_args || !b->h_ana_args) return RD_ERR_NOMEM;
    return RD_OK;
}

template <typename T>
static void fill_args(Plan *const *plans, int nprob, const T *const *ts, const T *const *ta,
                      const double2 *const *syn_in, double2 *const *syn_out, const double2 *const *ana_in,
                      double2 *const *ana_out, SynArgs<T> *sa, AnaArgs<T> *aa)
{
    for (int p = 0; p < nprob; p++) {
        const Plan *pl = plans[p];
        SynArgs<T> &S = sa[p];
        S.tP = S.tD = S.tM = ts[p];
        S.off = pl->d_syn_off;
        S.queue = nullptr;
        for (int f = 0; f < 3; f++) {
            S.in[f] = syn_in ? syn_in[p] : nullptr;
            S.out[f] = syn_out ? syn_out[p] : nullptr;
        }
        S.ld = pl->nlm;
        S.lmax = pl->lmax;
        S.nlat = pl->nlat;
        S.nth = pl->nth;
        S.nthp = pl->nthp;
        S.nplanes = pl->nz;
        S.n_jt = 0;
        AnaArgs<T> &A = aa[p];
        A.tP = A.tD = A.tM = ta[p];
        A.off = pl->d_ana_off;
        A.ncol = pl->d_ana_ncol;
        A.queue = nullptr;
        for (int f = 0; f < 3; f++) {
            A.in[f] = ana_in ? ana_in[p] : nullptr;
            A.out[f] = ana_out ? ana_out[p] : nullptr;
        }
        A.ld = pl->nlm;
        A.lmax = pl->lmax;
        A.nlat = pl->nlat;
        A.nth = pl->nth;
        A.nthp16 = pl->nthp16;
        A.nplanes = pl->nz;
        A.n_jt = 0;
    }
}

int batch_leg_bind(BatchLeg *b, Plan *const *plans, const double2 *const *syn_in, double2 *const *syn_out,
                   const double2 *const *ana_in, double2 *const *ana_out, hipStream_t s)
{
    const int np = b->nprob;
    std::vector<char> hs(b->arg_bytes_syn * np), ha(b->arg_bytes_ana * np);
    if (b->prec == RD_PREC_FP64) {
        std::vector<const double *> ts(np), ta(np);
        for (int p = 0; p < np; p++) {
            ts[p] = plans[p]->syn_P;
            ta[p] = plans[p]->ana_P;
        }
        fill_args<double>(plans, np, ts.data(), ta.data(), syn_in, syn_out, ana_in, ana_out,
                          (SynArgs<double> *)hs.data(), (AnaArgs<double> *)ha.data());
    } else {
        std::vector<const float *> ts(np), ta(np);
        for (int p = 0; p < np; p++) {
            ts[p] = plans[p]->syn32_P;
            ta[p] = plans[p]->ana32_P;
        }
        fill_args<float>(plans, np, ts.data(), ta.data(), syn_in, syn_out, ana_in, ana_out,
                         (SynArgs<float> *)hs.data(), (AnaArgs<float> *)ha.data());
    }
    const bool same = memcmp(hs.data(), b->h_syn_args, hs.size()) == 0 &&
                      memcmp(ha.data(), b->h_ana_args, ha.size()) == 0;
    if (same) return RD_OK;
    RD_HIP(hipStreamSynchronize(s));
    RD_HIP(hipMemcpy(b->d_syn_args, hs.data(), hs.size(), hipMemcpyHostToDevice));
    RD_HIP(hipMemcpy(b->d_ana_args, ha.data(), ha.size(), hipMemcpyHostToDevice));
    memcpy(b->h_syn_args, hs.data(), hs.size());
    memcpy(b->h_ana_args, ha.data(), ha.size());
    return RD_OK;
}

template <typename T>
static void fill_vargs(const BatchLeg *b, Plan *const *plans, const double2 *const *syn_in,
                       double2 *const *syn_out, const double2 *const *ana_in, double2 *const *ana_out,
                       SynArgs<T> *sa, AnaArgs<T> *aa)
{
    const bool f32 = b->prec == RD_PREC_FP32;
    for (int p = 0; p < b->nprob; p++) {
        const Plan *pl = plans[p];
        for (int set = 0; set < b->nsets; set++) {
            SynArgs<T> &S = sa[p * b->nsets + set];
            S.tP = (const T *)(f32 ? (const void *)pl->syn32_P : (const void *)pl->syn_P);
            S.tD = (const T *)(f32 ? (const void *)pl->syn32_D : (const void *)pl->syn_D);
            S.tM = (const T *)(f32 ? (const void *)pl->syn32_M : (const void *)pl->syn_M);
            S.off = pl->d_syn_off;
            S.queue = nullptr;
            for (int f = 0; f < 3; f++) {
                S.in[f] = syn_in[(p * b->nsets + set) * 3 + f];
                S.out[f] = syn_out[(p * b->nsets + set) * 3 + f];
            }
            S.ld = pl->nlm;
            S.lmax = pl->lmax;
            S.nlat = pl->nlat;
            S.nth = pl->nth;
            S.nthp = pl->nthp;
            S.nplanes = pl->nr;
            S.n_jt = 0;
        }
        AnaArgs<T> &A = aa[p];
        A.tP = (const T *)(f32 ? (const void *)pl->ana32_P : (const void *)pl->ana_P);
        A.tD = (const T *)(f32 ? (const void *)pl->ana32_D : (const void *)pl->ana_D);
        A.tM = (const T *)(f32 ? (const void *)pl->ana32_M : (const void *)pl->ana_M);
        A.off = pl->d_ana_off;
        A.ncol = pl->d_ana_ncol;
        A.queue = nullptr;
        for (int f = 0; f < 3; f++) {
            A.in[f] = ana_in[p * 3 + f];
            A.out[f] = ana_out[p * 3 + f];
        }
        A.ld = pl->nlm;
        A.lmax = pl->lmax;
        A.nlat = pl->nlat;
        A.nth = pl->nth;
        A.nthp16 = pl->nthp16;
        A.nplanes = pl->nr;
        A.n_jt = 0;
    }
}

int batch_vleg_bind(BatchLeg *b, Plan *const *plans, const double2 *const *syn_in, double2 *const *syn_out,
                    const double2 *const *ana_in, double2 *const *ana_out, hipStream_t s)
{
    if (!b->vec) return RD_ERR_STATE;
    const int np = b->nprob;
    std::vector<char> hs(b->arg_bytes_syn * np * b->nsets), ha(b->arg_bytes_ana * np);
    if (b->prec == RD_PREC_FP64)
        fill_vargs<double>(b, plans, syn_in, syn_out, ana_in, ana_out, (SynArgs<double> *)hs.data(),
                           (AnaArgs<double> *)ha.data());
    else
        fill_vargs<float>(b, plans, syn_in, syn_out, ana_in, ana_out, (SynArgs<float> *)hs.data(),
                          (AnaArgs<float> *)ha.data());
    const bool same = memcmp(hs.data(), b->h_syn_args, hs.size()) == 0 &&
                      memcmp(ha.data(), b->h_ana_args, ha.size()) == 0;
    if (same) return RD_OK;
    RD_HIP(hipStreamSynchronize(s));
    RD_HIP(hipMemcpy(b->d_syn_args, hs.data(), hs.size(), hipMemcpyHostToDevice));
    RD_HIP(hipMemcpy(b->d_ana_args, ha.data(), ha.size(), hipMemcpyHostToDevice));
    memcpy(b->h_syn_args, hs.data(), hs.size());
    memcpy(b->h_ana_args, ha.data(), ha.size());
    return RD_OK;
}

int launch_batch_synthesis(BatchLeg *b, hipStream_t s, int nwg)
{
    hipLaunchKernelGGL(zero_counters_kernel, dim3(1), dim3(64), 0, s, b->d_counter, 1);
    RD_LAUNCH_CHECK();
#define RD_BS(K, T, J)                                                                                     \
    hipLaunchKernelGGL((K<T, J>), dim3(nwg), dim3(256), 0, s, (const SynArgs<T> *)b->d_syn_args,          \
                       b->d_syn_items, b->n_syn_items, b->d_counter)
    if (b->vec) {
        if (b->prec == RD_PREC_FP64) {
            if (b->jb == 1) RD_BS(batch_vsynthesis_kernel, double, 1);
            else RD_BS(batch_vsynthesis_kernel, double, 2);
        } else {
            if (b->jb == 1) RD_BS(batch_vsynthesis_kernel, float, 1);
            else RD_BS(batch_vsynthesis_kernel, float, 2);
        }
    } else if (b->prec == RD_PREC_FP64) {
        if (b->jb == 2) RD_BS(batch_synthesis_kernel, double, 2);
        else RD_BS(batch_synthesis_kernel, double, 4);
    } else {
        if (b->jb == 2) RD_BS(batch_synthesis_kernel, float, 2);
        else RD_BS(batch_synthesis_kernel, float, 4);
    }
#undef RD_BS
    RD_LAUNCH_CHECK();
    return RD_OK;
}

int launch_batch_analysis(BatchLeg *b, hipStream_t s, int nwg)
{
    hipLaunchKernelGGL(zero_counters_kernel, dim3(1), dim3(64), 0, s, b->d_counter + 1, 1);
    RD_LAUNCH_CHECK();
#define RD_BA(K, T, J)                                                                                     \
    hipLaunchKernelGGL((K<T, J>), dim3(nwg), dim3(256), 0, s, (const AnaArgs<T> *)b->d_ana_args,          \
                       b->d_ana_items, b->n_ana_items, b->d_counter + 1)
    if (b->vec) {
        if (b->prec == RD_PREC_FP64) {
            if (b->jb == 1) RD_BA(batch_vanalysis_kernel, double, 1);
            else RD_BA(batch_vanalysis_kernel, double, 2);
        } else {
            if (b->jb == 1) RD_BA(batch_vanalysis_kernel, float, 1);
            else RD_BA(batch_vanalysis_kernel, float, 2);
        }
    } else if (b->prec == RD_PREC_FP64) {
        if (b->jb == 2) RD_BA(batch_analysis_kernel, double, 2);
        else RD_BA(batch_analysis_kernel, double, 4);
    } else {
        if (b->jb == 2) RD_BA(batch_analysis_kernel, float, 2);
        else RD_BA(batch_analysis_kernel, float, 4);
    }
#undef RD_BA
    RD_LAUNCH_CHECK();
    return RD_OK;
}

void batch_leg_free(BatchLeg *b)
{
    void *d[] = {b->d_syn_args, b->d_ana_args, b->d_syn_items, b->d_ana_items, b->d_counter};
    for (void *x : d)
        if (x) (void)hipFree(x);
    free(b->h_syn_args);
    free(b->h_ana_args);
    memset(b, 0, sizeof(*b));
}

}  // namespace rd

