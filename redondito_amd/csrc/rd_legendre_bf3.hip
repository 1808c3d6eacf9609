// Precision forms of the Legendre contractions.
//
// FP32: the FP64 tables rounded to FP32, consumed by the FP32 instances of
// synthesis_kernel / analysis_kernel (rd_legendre.hip).
//
// BF16x3: every operand x is split into two BF16 values, hi = bf16(x) and
// lo = bf16(x - hi), and each product is taken as hi*hi + hi*lo + lo*hi on
// v_mfma_f32_16x16x16_bf16 with FP32 accumulation (the lo*lo term, below
// 2^-16 relative, is dropped).  The table splits come from the FP64 tables;
// the spectral / Fourier operands are split when staged into LDS.
//
// Operand layouts of the 16x16x16 BF16 MFMA (measured on gfx950): lane l
// holds A[l%16][4(l/16) + e] and B[4(l/16) + e][l%16], e = 0..3, as four
// packed BF16; D lane l, register v holds D[4(l/16) + v][l%16].  The tables
// are therefore stored with four consecutive K values (rows l_p for the
// synthesis, nodes theta for the analysis) next to each other:
//   synthesis  [m,p][row/4][theta nthp][4]
//   analysis   [m,p][theta/4][col l_p][4]
// at the same block offsets as the FP64 tables.
#include "rd_internal.h"

#include <cstdint>

namespace rd {

typedef short v4s __attribute__((ext_vector_type(4)));
typedef float v4f __attribute__((ext_vector_type(4)));

static constexpr unsigned long long kSigns = 0x8000800080008000ull;

__device__ __forceinline__ unsigned short bf16_rne(float x)
{
    unsigned u = __float_as_uint(x);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}

__device__ __forceinline__ float bf16_to_float(unsigned short h) { return __uint_as_float((unsigned)h << 16); }

// x = hi + lo + O(2^-17 |x|)
__device__ __forceinline__ void split_bf16(double x, unsigned short &hi, unsigned short &lo)
{
    hi = bf16_rne((float)x);
    lo = bf16_rne((float)(x - (double)bf16_to_float(hi)));
}

__device__ __forceinline__ v4s flip(v4s x, unsigned long long mask)
{
    return __builtin_bit_cast(v4s, __builtin_bit_cast(unsigned long long, x) ^ mask);
}

__device__ __forceinline__ v4f mma_bf(v4s a, v4s b, v4f c)
{
    return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}

// c += a*b with a = ah + al, b = bh + bl (three products, small ones first)
__device__ __forceinline__ v4f mma3(v4s ah, v4s al, v4s bh, v4s bl, v4f c)
{
    c = mma_bf(al, bh, c);
    c = mma_bf(ah, bl, c);
    return mma_bf(ah, bh, c);
}

__device__ __forceinline__ long lm_idx(int lmax, int l, int m)
{
    return (long)m * (2 * lmax + 1 - m) / 2 + l;
}

// ---------------------------------------------------------------- tables
__global__ void to_fp32_kernel(const double *src, float *dst, long n)
{
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        dst[i] = (float)src[i];
}

// synthesis table [row][theta] -> [row/4][theta][row%4] (rows of every block
// start at a multiple of 16, so the global row index can be used)
__global__ void syn_bf3_kernel(const double *src, unsigned short *hi, unsigned short *lo, long rows, int nthp)
{
    const long n = rows * nthp;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const long row = i / nthp;
        const int th = (int)(i % nthp);
        const long d = (row >> 2) * 4L * nthp + 4L * th + (row & 3);
        split_bf16(src[i], hi[d], lo[d]);
    }
}

// analysis table block (m,p) [theta][col] -> [theta/4][col][theta%4]
__global__ void ana_bf3_kernel(const double *src, unsigned short *hi, unsigned short *lo, const long *off,
                               const int *ncol, int nthp16)
{
    const int mp = blockIdx.y, nc = ncol[mp >> 1];
    const long base = off[mp];
    const long n = (long)nthp16 * nc;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const int th = (int)(i / nc), c = (int)(i % nc);
        const long d = base + (long)(th >> 2) * 4 * nc + 4L * c + (th & 3);
        split_bf16(src[base + i], hi[d], lo[d]);
    }
}

int ensure_precision_tables(Plan *p, int prec)
{
    if (prec != RD_PREC_FP32 && prec != RD_PREC_BF16X3) return RD_ERR_ARG;
    const long ns = p->syn_rows * p->nthp, na = p->ana_elems;
    const double *syn[3] = {p->syn_P, p->syn_D, p->syn_M};
    const double *ana[3] = {p->ana_P, p->ana_D, p->ana_M};
    if (prec == RD_PREC_FP32) {
        if (p->syn32_P) return RD_OK;
        float **s32[3] = {&p->syn32_P, &p->syn32_D, &p->syn32_M};
        float **a32[3] = {&p->ana32_P, &p->ana32_D, &p->ana32_M};
        for (int t = 0; t < 3; t++) {
            RD_HIP(hipMalloc(s32[t], sizeof(float) * ns));
            RD_HIP(hipMalloc(a32[t], sizeof(float) * na));
            hipLaunchKernelGGL(to_fp32_kernel, dim3(2048), dim3(256), 0, p->stream, syn[t], *s32[t], ns);
            RD_LAUNCH_CHECK();
            hipLaunchKernelGGL(to_fp32_kernel, dim3(2048), dim3(256), 0, p->stream, ana[t], *a32[t], na);
            RD_LAUNCH_CHECK();
        }
        return RD_OK;
    }
    if (p->synb_hi[0]) return RD_OK;
    for (int t = 0; t < 3; t++) {
        RD_HIP(hipMalloc(&p->synb_hi[t], sizeof(uint16_t) * ns));
        RD_HIP(hipMalloc(&p->synb_lo[t], sizeof(uint16_t) * ns));
        RD_HIP(hipMalloc(&p->anab_hi[t], sizeof(uint16_t) * na));
        RD_HIP(hipMalloc(&p->anab_lo[t], sizeof(uint16_t) * na));
        hipLaunchKernelGGL(syn_bf3_kernel, dim3(2048), dim3(256), 0, p->stream, syn[t], p->synb_hi[t],
                           p->synb_lo[t], p->syn_rows, p->nthp);
        RD_LAUNCH_CHECK();
        hipLaunchKernelGGL(ana_bf3_kernel, dim3(64, 2 * p->L), dim3(256), 0, p->stream, ana[t], p->anab_hi[t],
                           p->anab_lo[t], p->d_ana_off, p->d_ana_ncol, p->nthp16);
        RD_LAUNCH_CHECK();
    }
    return RD_OK;
}

void free_precision_tables(Plan *p)
{
    float *f[6] = {p->syn32_P, p->syn32_D, p->syn32_M, p->ana32_P, p->ana32_D, p->ana32_M};
    for (float *x : f)
        if (x) (void)hipFree(x);
    for (int t = 0; t < 3; t++) {
        uint16_t *b[4] = {p->synb_hi[t], p->synb_lo[t], p->anab_hi[t], p->anab_lo[t]};
        for (uint16_t *x : b)
            if (x) (void)hipFree(x);
    }
}

// ------------------------------------------------------------- synthesis
// Same tiling as synthesis_kernel: one m, 64 northern nodes (16 per wave),
// JT real columns; the l loop in chunks of 32 (one K=16 MFMA per parity).
// LDS holds the chunk as [field][parity][column][l_p], hi and lo.  Vector
// tiles take one 16-column block per wave (0.45 against 0.55 ms with two
// blocks at 2 waves, lmax 255, 256 planes).  With the next chunk's table
// fragments prefetched into a second register set and unconditional chunk
// loads the vector tile takes 112 VGPRs (4 waves per SIMD): u and omega
// syntheses at lmax 255, 128 planes, 0.758 -> 0.656 ms, outputs bit-identical.
constexpr int BP = 20;   // column pitch in BF16 elements (16 + 4, 8-byte aligned)

struct Bf3Args {
    const unsigned short *th_[3], *tl_[3];   // P, dP/dtheta, mP/sin: hi and lo
    const long *off;
    const int *ncol;
    const int *queue;
    const double2 *in[3];
    double2 *out[3];
    // a second field set synthesised by the same launch (nset = 2): the plane
    // tiles of both sets of one (m, theta tile) entry run back to back on one
    // XCD, so the entry's table fragments serve both from L2
    const double2 *in2[3];
    double2 *out2[3];
    long ld;
    int lmax, nlat, nth, nthp, nthp16, nplanes, n_jt, nset;
};

template <bool VEC>
__global__ __launch_bounds__(256, VEC ? 4 : 1) void synthesis_bf3_kernel(Bf3Args A)
{
    constexpr int NF = VEC ? 3 : 1, JB = VEC ? 1 : 4, JT = 16 * JB;
    __shared__ __align__(16) unsigned short Bh[NF][2][JT][BP];
    __shared__ __align__(16) unsigned short Bl[NF][2][JT][BP];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int L = A.lmax + 1;
    const int xcd = blockIdx.x % kXcds, i = blockIdx.x / kXcds;
    const int nj = A.n_jt * A.nset, jts = i % nj, jt = jts % A.n_jt;
    const int ent = A.queue[(i / nj) * kXcds + xcd];
    if (ent < 0) return;
    if (jts >= A.n_jt)
#pragma unroll
        for (int f = 0; f < 3; f++) {
            A.in[f] = A.in2[f];
            A.out[f] = A.out2[f];
        }
    const int m = ent & 0xffff, tt = ent >> 16;
    const int k0 = jt * (JT / 2);
    const int th0 = tt * 64 + wave * 16;
    const int col = lane & 15, kk = lane >> 4;
    // (i X)_j = sgn_i X_{j^1} with sgn_i = -1 on even columns (real parts)
    const unsigned long long mi = (col & 1) ? 0ull : kSigns;

    v4f acc[NF][2][JB];
#pragma unroll
    for (int f = 0; f < NF; f++)
#pragma unroll
        for (int s = 0; s < 2; s++)
#pragma unroll
            for (int b = 0; b < JB; b++) acc[f][s][b] = (v4f){0.f, 0.f, 0.f, 0.f};

    const long offe = A.off[2 * m] * A.nthp, offo = A.off[2 * m + 1] * A.nthp;
    const long lm0 = lm_idx(A.lmax, m, m);

    // a chunk is NF x (JT/2 planes) x 32 l's, l fastest across threads; each
    // thread moves NLD double2 of it: element tid + 256 i, whose field
    // (256 i) / PER is a compile-time constant.  The next chunk is loaded into
    // registers while the current one is multiplied out of LDS, from clamped
    // addresses, unconditionally (a load under a condition makes the loads of
    // a chunk issue one after another); entries beyond lmax or the last plane
    // are zeroed where the chunk is split into LDS.
    constexpr int PER = (JT / 2) * 32, NLD = NF * PER / 256;
    static_assert(PER % 256 == 0, "a thread's chunk elements must each lie in one field");
    double2 rv[NLD];
    unsigned rlive = 0;   // bit i: element i lies inside the spectrum
    auto load = [&](int lbase) {
        rlive = 0;
#pragma unroll
        for (int i = 0; i < NLD; i++) {
            const int f = (256 * i) / PER, r = tid + (256 * i) % PER, pl = r / 32, ll = r % 32;
            const int l = lbase + ll, k = k0 + pl;
            const int lc = l < A.lmax ? l : A.lmax, kc = k < A.nplanes ? k : A.nplanes - 1;
            rv[i] = A.in[f][lm0 + (lc - m) + kc * A.ld];
            rlive |= (unsigned)(l <= A.lmax && k < A.nplanes) << i;
        }
    };
    // table fragments (4 consecutive rows of both parities of the three
    // tables) of the chunk at lbase, 32-bit element offsets from the uniform
    // table bases; the next chunk's are loaded before the current chunk's
    // MFMAs, into the other of two register sets
    struct Tab {
        v4s Peh, Pel, Poh, Pol, Deh, Del, Doh, Dol, Meh, Mel, Moh, Mol;
    };
    auto ltab = [&](int lbase, Tab &t) {
        const int rg = (((lbase - m) >> 1) >> 2) + kk;   // group of 4 rows
        const unsigned ie = (unsigned)offe + (unsigned)(rg * 4 * A.nthp + 4 * (th0 + col));
        const unsigned io = (unsigned)offo + (unsigned)(rg * 4 * A.nthp + 4 * (th0 + col));
        t.Peh = *(const v4s *)(A.th_[0] + ie);
        t.Pel = *(const v4s *)(A.tl_[0] + ie);
        t.Poh = *(const v4s *)(A.th_[0] + io);
        t.Pol = *(const v4s *)(A.tl_[0] + io);
        if (VEC) {
            t.Deh = *(const v4s *)(A.th_[1] + ie);
            t.Del = *(const v4s *)(A.tl_[1] + ie);
            t.Doh = *(const v4s *)(A.th_[1] + io);
            t.Dol = *(const v4s *)(A.tl_[1] + io);
            t.Meh = *(const v4s *)(A.th_[2] + ie);
            t.Mel = *(const v4s *)(A.tl_[2] + ie);
            t.Moh = *(const v4s *)(A.th_[2] + io);
            t.Mol = *(const v4s *)(A.tl_[2] + io);
        }
    };
    int lbase = m;
    auto chunk = [&](const Tab &cur, Tab &nxt) {
#pragma unroll
        for (int i = 0; i < NLD; i++) {
            const int f = (256 * i) / PER, r = tid + (256 * i) % PER, pl = r / 32, ll = r % 32;
            const int par = ll & 1, lp = ll >> 1;
            const bool ok = (rlive >> i) & 1u;
            unsigned short h, lo;
            split_bf16(ok ? rv[i].x : 0.0, h, lo);
            Bh[f][par][2 * pl][lp] = h;
            Bl[f][par][2 * pl][lp] = lo;
            split_bf16(ok ? rv[i].y : 0.0, h, lo);
            Bh[f][par][2 * pl + 1][lp] = h;
            Bl[f][par][2 * pl + 1][lp] = lo;
        }
        __syncthreads();
        const bool more = lbase + 32 <= A.lmax;
        if (more) load(lbase + 32);
        ltab(more ? lbase + 32 : lbase, nxt);   // (the last chunk reloads its own rows)
        const v4s Peh = cur.Peh, Pel = cur.Pel, Poh = cur.Poh, Pol = cur.Pol;
        const v4s Deh = cur.Deh, Del = cur.Del, Doh = cur.Doh, Dol = cur.Dol;
        const v4s Meh = cur.Meh, Mel = cur.Mel, Moh = cur.Moh, Mol = cur.Mol;
#pragma unroll
        for (int b = 0; b < JB; b++) {
            const int j = b * 16 + col;
#define LDB(X, f, par, jj) (*(const v4s *)&X[f][par][jj][4 * kk])
            acc[0][0][b] = mma3(Peh, Pel, LDB(Bh, 0, 0, j), LDB(Bl, 0, 0, j), acc[0][0][b]);
            acc[0][1][b] = mma3(Poh, Pol, LDB(Bh, 0, 1, j), LDB(Bl, 0, 1, j), acc[0][1][b]);
            if (VEC) {
                const v4s Seh = LDB(Bh, 1, 0, j), Sel = LDB(Bl, 1, 0, j);
                const v4s Soh = LDB(Bh, 1, 1, j), Sol = LDB(Bl, 1, 1, j);
                const v4s mTeh = flip(LDB(Bh, 2, 0, j), kSigns), mTel = flip(LDB(Bl, 2, 0, j), kSigns);
                const v4s mToh = flip(LDB(Bh, 2, 1, j), kSigns), mTol = flip(LDB(Bl, 2, 1, j), kSigns);
                const v4s iSeh = flip(LDB(Bh, 1, 0, j ^ 1), mi), iSel = flip(LDB(Bl, 1, 0, j ^ 1), mi);
                const v4s iSoh = flip(LDB(Bh, 1, 1, j ^ 1), mi), iSol = flip(LDB(Bl, 1, 1, j ^ 1), mi);
                const v4s iTeh = flip(LDB(Bh, 2, 0, j ^ 1), mi), iTel = flip(LDB(Bl, 2, 0, j ^ 1), mi);
                const v4s iToh = flip(LDB(Bh, 2, 1, j ^ 1), mi), iTol = flip(LDB(Bl, 2, 1, j ^ 1), mi);
                // v_theta: symmetric = dP_o S_o + mP_e (iT)_e, antisymmetric = dP_e S_e + mP_o (iT)_o
                acc[1][0][b] = mma3(Doh, Dol, Soh, Sol, acc[1][0][b]);
                acc[1][0][b] = mma3(Meh, Mel, iTeh, iTel, acc[1][0][b]);
                acc[1][1][b] = mma3(Deh, Del, Seh, Sel, acc[1][1][b]);
                acc[1][1][b] = mma3(Moh, Mol, iToh, iTol, acc[1][1][b]);
                // v_phi: symmetric = mP_e (iS)_e - dP_o T_o, antisymmetric = mP_o (iS)_o - dP_e T_e
                acc[2][0][b] = mma3(Meh, Mel, iSeh, iSel, acc[2][0][b]);
                acc[2][0][b] = mma3(Doh, Dol, mToh, mTol, acc[2][0][b]);
                acc[2][1][b] = mma3(Moh, Mol, iSoh, iSol, acc[2][1][b]);
                acc[2][1][b] = mma3(Deh, Del, mTeh, mTel, acc[2][1][b]);
            }
#undef LDB
        }
        __syncthreads();
        lbase += 32;
    };
    // two table register sets, alternating with the chunk parity (no copy of
    // a set whose loads are in flight)
    Tab ta = {}, tb = {};
    load(m);
    ltab(m, ta);
    for (;;) {
        chunk(ta, tb);
        if (lbase > A.lmax) break;
        chunk(tb, ta);
        if (lbase > A.lmax) break;
    }

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
                const int th = th0 + 4 * kk + v;
                if (th >= A.nth) continue;
                const double sy = acc[f][0][b][v], an = acc[f][1][b][v];
                out[2 * (base + th) + c] = sy + an;
                out[2 * (base + A.nlat - 1 - th) + c] = sy - an;
            }
        }
    }
}

// -------------------------------------------------------------- analysis
// Same tiling as analysis_kernel: one m, 32 rows l_p (2 waves x 16), JT real
// columns; the theta loop in chunks of 16 (one K=16 MFMA).  LDS holds the
// north +- south combinations as [field][sum/dif][column][theta], hi and lo.
// Vector tiles take one 16-column block per wave, as in synthesis (0.39
// against 0.54 ms with two blocks).
template <bool VEC>
__global__ __launch_bounds__(256, VEC ? 4 : 1) void analysis_bf3_kernel(Bf3Args A)
{
    constexpr int NF = VEC ? 3 : 1, JB = VEC ? 1 : 4, JT = 2 * 16 * JB;
    __shared__ __align__(16) unsigned short Gh[NF][2][JT][BP];
    __shared__ __align__(16) unsigned short Gl[NF][2][JT][BP];

    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int L = A.lmax + 1;
    const int xcd = blockIdx.x % kXcds, i = blockIdx.x / kXcds;
    const int jt = i % A.n_jt;
    const int ent = A.queue[(i / A.n_jt) * kXcds + xcd];
    if (ent < 0) return;
    const int m = ent & 0xffff, lt = ent >> 16;
    const int wl = wave & 1, wj = wave >> 1;
    const int k0 = jt * (JT / 2);
    const int col = lane & 15, kk = lane >> 4;
    // (-i X)_j = sgn X_{j^1} with sgn = -1 on odd columns (imaginary parts)
    const unsigned long long mmi = (col & 1) ? kSigns : 0ull;
    const int ncol = A.ncol[m];
    const int lp0 = lt * 32 + wl * 16;
    // a wave whose 16 rows l_p are all padding stages and waits at the
    // barriers but skips the table loads and MFMAs
    const bool live = lp0 < (L - m + 1) / 2;

    v4f acc[NF][2][JB];
#pragma unroll
    for (int f = 0; f < NF; f++)
#pragma unroll
        for (int s = 0; s < 2; s++)
#pragma unroll
            for (int b = 0; b < JB; b++) acc[f][s][b] = (v4f){0.f, 0.f, 0.f, 0.f};

    const long offe = A.off[2 * m], offo = A.off[2 * m + 1];

    // a chunk is NF x (JT/2 planes) x 16 northern nodes, node fastest across
    // threads; each thread loads element tid + 256 i of it (field (256 i) /
    // PER, a compile-time constant), north and south, from clamped addresses,
    // unconditionally, so that the loads issue back to back (a load under a
    // condition makes them go out one after another); nodes past nth and
    // planes past the last are zeroed where the chunk is split into LDS.
    // (Loading the next chunk during the current one's MFMAs was slower:
    // 0.352 against 0.290 ms at lmax 255, 128 planes; its registers cost 8
    // spills at 4 waves per SIMD.)
    constexpr int PER = (JT / 2) * 16, NLD = NF * PER / 256;
    static_assert(PER % 256 == 0, "a thread's chunk elements must each lie in one field");
    double2 rn[NLD], rs[NLD];
    unsigned rlive = 0;
    auto load = [&](int tc) {
        rlive = 0;
#pragma unroll
        for (int i = 0; i < NLD; i++) {
            const int f = (256 * i) / PER, r = tid + (256 * i) % PER, pl = r / 16, tl = r % 16;
            const int th = tc + tl, k = k0 + pl;
            const int thc = th < A.nth ? th : A.nth - 1, kc = k < A.nplanes ? k : A.nplanes - 1;
            const long base = ((long)kc * L + m) * A.nlat;
            rn[i] = A.in[f][base + thc];
            rs[i] = A.in[f][base + A.nlat - 1 - thc];
            rlive |= (unsigned)(th < A.nth && k < A.nplanes) << i;
        }
    };
    // table fragments (4 consecutive nodes of both parities of the three
    // tables) of the chunk at tc, loaded after the chunk's barrier
    struct Tab {
        v4s Peh, Pel, Poh, Pol, Deh, Del, Doh, Dol, Meh, Mel, Moh, Mol;
    };
    auto ltab = [&](int tc, Tab &t) {
        const long tg = (tc >> 2) + kk;   // group of 4 nodes
        const long ie = offe + tg * 4L * ncol + 4L * (lp0 + col);
        const long io = offo + tg * 4L * ncol + 4L * (lp0 + col);
        t.Peh = *(const v4s *)(A.th_[0] + ie);
        t.Pel = *(const v4s *)(A.tl_[0] + ie);
        t.Poh = *(const v4s *)(A.th_[0] + io);
        t.Pol = *(const v4s *)(A.tl_[0] + io);
        if (VEC) {
            t.Deh = *(const v4s *)(A.th_[1] + ie);
            t.Del = *(const v4s *)(A.tl_[1] + ie);
            t.Doh = *(const v4s *)(A.th_[1] + io);
            t.Dol = *(const v4s *)(A.tl_[1] + io);
            t.Meh = *(const v4s *)(A.th_[2] + ie);
            t.Mel = *(const v4s *)(A.tl_[2] + ie);
            t.Moh = *(const v4s *)(A.th_[2] + io);
            t.Mol = *(const v4s *)(A.tl_[2] + io);
        }
    };
    for (int tc = 0; tc < A.nth; tc += 16) {
        load(tc);
#pragma unroll
        for (int i = 0; i < NLD; i++) {
            const int f = (256 * i) / PER, r = tid + (256 * i) % PER, pl = r / 16, tl = r % 16;
            const bool ok = (rlive >> i) & 1u;
            const double2 vn = ok ? rn[i] : make_double2(0.0, 0.0), vs = ok ? rs[i] : make_double2(0.0, 0.0);
            unsigned short h, lo;
            split_bf16(vn.x + vs.x, h, lo);
            Gh[f][0][2 * pl][tl] = h;
            Gl[f][0][2 * pl][tl] = lo;
            split_bf16(vn.y + vs.y, h, lo);
            Gh[f][0][2 * pl + 1][tl] = h;
            Gl[f][0][2 * pl + 1][tl] = lo;
            split_bf16(vn.x - vs.x, h, lo);
            Gh[f][1][2 * pl][tl] = h;
            Gl[f][1][2 * pl][tl] = lo;
            split_bf16(vn.y - vs.y, h, lo);
            Gh[f][1][2 * pl + 1][tl] = h;
            Gl[f][1][2 * pl + 1][tl] = lo;
        }
        __syncthreads();
        if (live) {
            Tab cur = {};
            ltab(tc, cur);
            const v4s Peh = cur.Peh, Pel = cur.Pel, Poh = cur.Poh, Pol = cur.Pol;
            const v4s Deh = cur.Deh, Del = cur.Del, Doh = cur.Doh, Dol = cur.Dol;
            const v4s Meh = cur.Meh, Mel = cur.Mel, Moh = cur.Moh, Mol = cur.Mol;
#pragma unroll
            for (int b = 0; b < JB; b++) {
                const int j = (wj * JB + b) * 16 + col;
#define LDG(X, f, sd, jj) (*(const v4s *)&X[f][sd][jj][4 * kk])
                acc[0][0][b] = mma3(Peh, Pel, LDG(Gh, 0, 0, j), LDG(Gl, 0, 0, j), acc[0][0][b]);
                acc[0][1][b] = mma3(Poh, Pol, LDG(Gh, 0, 1, j), LDG(Gl, 0, 1, j), acc[0][1][b]);
                if (VEC) {
                    const v4s s1h = LDG(Gh, 1, 0, j), s1l = LDG(Gl, 1, 0, j);
                    const v4s d1h = LDG(Gh, 1, 1, j), d1l = LDG(Gl, 1, 1, j);
                    const v4s ns2h = flip(LDG(Gh, 2, 0, j), kSigns), ns2l = flip(LDG(Gl, 2, 0, j), kSigns);
                    const v4s nd2h = flip(LDG(Gh, 2, 1, j), kSigns), nd2l = flip(LDG(Gl, 2, 1, j), kSigns);
                    const v4s rs1h = flip(LDG(Gh, 1, 0, j ^ 1), mmi), rs1l = flip(LDG(Gl, 1, 0, j ^ 1), mmi);
                    const v4s rd1h = flip(LDG(Gh, 1, 1, j ^ 1), mmi), rd1l = flip(LDG(Gl, 1, 1, j ^ 1), mmi);
                    const v4s rs2h = flip(LDG(Gh, 2, 0, j ^ 1), mmi), rs2l = flip(LDG(Gl, 2, 0, j ^ 1), mmi);
                    const v4s rd2h = flip(LDG(Gh, 2, 1, j ^ 1), mmi), rd2l = flip(LDG(Gl, 2, 1, j ^ 1), mmi);
                    // S_e = dP_e dif_t + mP_e (-i) sum_p ; S_o = dP_o sum_t + mP_o (-i) dif_p
                    acc[1][0][b] = mma3(Deh, Del, d1h, d1l, acc[1][0][b]);
                    acc[1][0][b] = mma3(Meh, Mel, rs2h, rs2l, acc[1][0][b]);
                    acc[1][1][b] = mma3(Doh, Dol, s1h, s1l, acc[1][1][b]);
                    acc[1][1][b] = mma3(Moh, Mol, rd2h, rd2l, acc[1][1][b]);
                    // T_e = mP_e (-i) sum_t - dP_e dif_p ; T_o = mP_o (-i) dif_t - dP_o sum_p
                    acc[2][0][b] = mma3(Meh, Mel, rs1h, rs1l, acc[2][0][b]);
                    acc[2][0][b] = mma3(Deh, Del, nd2h, nd2l, acc[2][0][b]);
                    acc[2][1][b] = mma3(Moh, Mol, rd1h, rd1l, acc[2][1][b]);
                    acc[2][1][b] = mma3(Doh, Dol, ns2h, ns2l, acc[2][1][b]);
                }
#undef LDG
            }
        }
        __syncthreads();
    }

    const long lm0 = lm_idx(A.lmax, m, m);
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
                    const int lp = lp0 + 4 * kk + v;
                    const int l = m + par + 2 * lp;
                    if (l > A.lmax) continue;
                    double val = acc[f][par][b][v];
                    if (m == 0 && c == 1) val = 0.0;
                    out[2 * (lm0 + (l - m) + k * A.ld) + c] = val;
                }
            }
    }
}

// --------------------------------------------------------------- launchers
static Bf3Args bf3_args(Plan *p, bool syn, bool vec, int nplanes, const double2 *i0, const double2 *i1,
                        const double2 *i2, double2 *o0, double2 *o1, double2 *o2, long ld)
{
    Bf3Args A;
    for (int t = 0; t < 3; t++) {
        A.th_[t] = syn ? p->synb_hi[t] : p->anab_hi[t];
        A.tl_[t] = syn ? p->synb_lo[t] : p->anab_lo[t];
    }
    A.off = syn ? p->d_syn_off : p->d_ana_off;
    A.ncol = p->d_ana_ncol;
    A.queue = syn ? p->d_syn_q : p->d_ana_q;
    A.in[0] = i0;
    A.in[1] = vec ? i1 : i0;
    A.in[2] = vec ? i2 : i0;
    A.out[0] = o0;
    A.out[1] = vec ? o1 : o0;
    A.out[2] = vec ? o2 : o0;
    A.ld = ld;
    A.lmax = p->lmax;
    A.nlat = p->nlat;
    A.nth = p->nth;
    A.nthp = p->nthp;
    A.nthp16 = p->nthp16;
    A.nplanes = nplanes;
    A.nset = 1;
    for (int f = 0; f < 3; f++) {
        A.in2[f] = nullptr;
        A.out2[f] = nullptr;
    }
    return A;
}

int launch_synthesis_bf3(Plan *p, bool vec, int nplanes, const double2 *q, const double2 *s, const double2 *t,
                         long ld, double2 *f0, double2 *f1, double2 *f2, const double2 *const *in2,
                         double2 *const *out2)
{
    Bf3Args A = bf3_args(p, true, vec, nplanes, q, s, t, f0, f1, f2, ld);
    if (vec && in2) {
        A.nset = 2;
        for (int f = 0; f < 3; f++) {
            A.in2[f] = in2[f];
            A.out2[f] = out2[f];
        }
    }
    const int JT = vec ? 16 : 64;
    A.n_jt = (2 * nplanes + JT - 1) / JT;
    dim3 grid((unsigned)(kXcds * p->syn_qlen * A.n_jt * A.nset));
    if (vec)
        hipLaunchKernelGGL(synthesis_bf3_kernel<true>, grid, dim3(256), 0, p->stream, A);
    else
        hipLaunchKernelGGL(synthesis_bf3_kernel<false>, grid, dim3(256), 0, p->stream, A);
    RD_LAUNCH_CHECK();
    return RD_OK;
}

int launch_analysis_bf3(Plan *p, bool vec, int nplanes, const double2 *g0, const double2 *g1,
                        const double2 *g2, double2 *q, double2 *s, double2 *t, long ld)
{
    Bf3Args A = bf3_args(p, false, vec, nplanes, g0, g1, g2, q, s, t, ld);
    const int JT = vec ? 32 : 128;
    A.n_jt = (2 * nplanes + JT - 1) / JT;
    dim3 grid((unsigned)(kXcds * p->ana_qlen * A.n_jt));
    if (vec)
        hipLaunchKernelGGL(analysis_bf3_kernel<true>, grid, dim3(256), 0, p->stream, A);
    else
        hipLaunchKernelGGL(analysis_bf3_kernel<false>, grid, dim3(256), 0, p->stream, A);
    RD_LAUNCH_CHECK();
    return RD_OK;
}

}  // namespace rd
