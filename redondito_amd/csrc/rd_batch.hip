// Batching form: many independent shell problems of different sizes (lmax,
// grid, radial length) transformed together.  Each problem has its own plan
// (tables, twiddles, one Fourier and one spectral scratch array); the
// Legendre stage of all problems runs as one persistent launch per direction
// (work items ordered largest first, pulled from a counter), the longitude
// and radial FFTs run per problem, spread over four lanes (the batch stream
// and three side streams forked from it and joined back with events, so that
// the small launches of different problems overlap; a captured graph keeps
// them as parallel branches).  The whole round trip
// can be captured once into a hipGraph and replayed, and a tuner times the
// tile / grid choices of the persistent launches and keeps the fastest.
// The problems at or above a degree threshold run their own one-problem chain
// instead (the kernels of the single-plan routines, spread over the lanes), and
// the persistent launches take the rest: a large problem fills the device on
// its own and its one-problem kernels (prefetching; in the nonlinear-term form
// both syntheses in one launch) are faster than the persistent tiles, while
// many small problems gain from sharing launches.  The tuners pick the
// threshold.
//
// Per problem p (the reference's fftp3d_complex_to_real and
// fftp3d_real_to_complex, src/sphp/sphp_sca.F90:237-270,400-433, applied to
// every plan of the batch):
//   inverse: spec (nz, nlm) -> radial inverse FFT -> synthesis -> phi inverse
//            -> grid (nz, nphi, nlat); spec is left unchanged (the radial
//            transform runs on the plan's copy)
//   forward: grid -> phi forward -> analysis -> FC-Gram continuation and
//            radial forward FFT (1/nz) -> spec
#include "rd_internal.h"

#include <algorithm>
#include <climits>
#include <vector>

using namespace rd;

constexpr int kLanes = 4;   // streams the per-problem stages are spread over
// default degree threshold of the one-problem chains: every problem on its own
// (on the 48-problem benchmark draw, lmax 31 to 255, the fastest threshold of
// the nonlinear-term form: 25.8 ms against 30.8 ms with every problem in the
// persistent launches); the tuners choose among the batch's own degrees
constexpr int kSplitDefault = 0;

struct rd_batch {
    int nprob, prec;
    int uxw;                          // the nonlinear-term form (rd_batch_create_uxw)
    std::vector<Plan *> plans;
    // problems with lmax >= split run one-problem chains (indices `large`),
    // the rest the persistent launches (`small`)
    int split;
    std::vector<int> small, large, all;
    hipStream_t stream, own_stream;
    hipStream_t side[kLanes - 1];   // default priority, like every stream here
    hipEvent_t ev_fork, ev_join[kLanes - 1];
    BatchLeg leg;
    int jb, wg_per_cu, n_cu;
    std::vector<double2 *> ana_out;   // bound output arrays of the analysis
    hipGraph_t graph;
    hipGraphExec_t gexec;
    bool captured;
    // the graph's own copies of the launch arguments and work counters, so
    // that eager calls with other arrays do not change what a replay reads
    void *g_syn_args, *g_ana_args;
    int *g_counter;
};

namespace {

int set_stream(rd_batch *b, hipStream_t s)
{
    b->stream = s;
    for (Plan *p : b->plans) p->stream = s;
    return RD_OK;
}

void drop_graph(rd_batch *b)
{
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    void *d[] = {b->g_syn_args, b->g_ana_args, b->g_counter};
    for (void *x : d)
        if (x) (void)hipFree(x);
    b->gexec = nullptr;
    b->graph = nullptr;
    b->g_syn_args = b->g_ana_args = nullptr;
    b->g_counter = nullptr;
    b->captured = false;
}

int nwg(const rd_batch *b) { return b->n_cu * b->wg_per_cu; }

void partition(rd_batch *b)
{
    b->small.clear();
    b->large.clear();
    b->all.clear();
    for (int i = 0; i < b->nprob; i++) {
        b->all.push_back(i);
        (b->plans[i]->lmax >= b->split ? b->large : b->small).push_back(i);
    }
}

std::vector<Plan *> plans_of(const rd_batch *b, const std::vector<int> &ids)
{
    std::vector<Plan *> v;
    for (int i : ids) v.push_back(b->plans[i]);
    return v;
}

int bind(rd_batch *b, double *const *spec_out)
{
    for (int p = 0; p < b->nprob; p++)
        if (spec_out) b->ana_out[p] = (double2 *)spec_out[p];
    const int np = (int)b->small.size();
    if (np == 0) return RD_OK;
    std::vector<const double2 *> sin(np), ain(np);
    std::vector<double2 *> sout(np), aout(np);
    for (int j = 0; j < np; j++) {
        const int p = b->small[j];
        sin[j] = b->plans[p]->S[0];
        sout[j] = b->plans[p]->F[0];
        ain[j] = b->plans[p]->F[0];
        aout[j] = b->ana_out[p];
    }
    const std::vector<Plan *> pls = plans_of(b, b->small);
    return batch_leg_bind(&b->leg, pls.data(), sin.data(), sout.data(), ain.data(), aout.data(), b->stream);
}

// Enqueues f(i) for every problem i of ids, the j-th of them with its plan
// bound to lane j % kLanes: the side streams wait for the batch stream's work
// so far, and the batch stream waits for all of them afterwards.
template <typename Fn>
int per_problem(rd_batch *b, const std::vector<int> &ids, Fn &&f)
{
    if (ids.empty()) return RD_OK;
    RD_HIP(hipEventRecord(b->ev_fork, b->stream));
    for (int l = 0; l < kLanes - 1; l++) RD_HIP(hipStreamWaitEvent(b->side[l], b->ev_fork, 0));
    int rc = RD_OK;
    for (size_t j = 0; j < ids.size() && !rc; j++) {
        const int i = ids[j], lane = (int)(j % kLanes);
        Plan *p = b->plans[i];
        p->stream = lane == 0 ? b->stream : b->side[lane - 1];
        rc = f(i);
        p->stream = b->stream;
    }
    for (int l = 0; l < kLanes - 1; l++) {
        RD_HIP(hipEventRecord(b->ev_join[l], b->side[l]));
        RD_HIP(hipStreamWaitEvent(b->stream, b->ev_join[l], 0));
    }
    return rc;
}

// the inverse chain, enqueued on the batch stream and its lanes (no host
// synchronisation): the one-problem chains of `large` (rd_3d_inverse's
// kernels, the batch's precision), then `small` through the persistent
// synthesis
int enqueue_inverse(rd_batch *b, const double *const *spec, double *const *grid)
{
    auto copy_z = [&](int i) {
        Plan *p = b->plans[i];
        RD_HIP(hipMemcpyAsync(p->S[0], spec[i], sizeof(double2) * p->ssize, hipMemcpyDeviceToDevice, p->stream));
        return launch_z_inverse(p, p->S[0]);
    };
    auto phi = [&](int i) {
        Plan *p = b->plans[i];
        const double2 *F[1] = {p->F[0]};
        double *f[1] = {grid[i]};
        return launch_phi_inverse(p, 1, p->nz, F, f);
    };
    int rc = per_problem(b, b->large, [&](int i) {
        Plan *p = b->plans[i];
        int r = copy_z(i);
        if (!r) r = launch_synthesis(p, b->prec, false, p->nz, p->S[0], nullptr, nullptr, p->nlm, p->F[0], nullptr,
                                     nullptr);
        return r ? r : phi(i);
    });
    if (rc || b->small.empty()) return rc;
    rc = per_problem(b, b->small, copy_z);
    if (rc) return rc;
    rc = launch_batch_synthesis(&b->leg, b->stream, nwg(b));
    if (rc) return rc;
    return per_problem(b, b->small, phi);
}

int enqueue_forward(rd_batch *b, const double *const *grid)
{
    auto phi = [&](int i) {
        Plan *p = b->plans[i];
        const double *f[1] = {grid[i]};
        double2 *F[1] = {p->F[0]};
        return launch_phi_forward(p, 1, p->nz, f, F);
    };
    int rc = per_problem(b, b->large, [&](int i) {
        Plan *p = b->plans[i];
        int r = phi(i);
        if (!r) r = launch_analysis(p, b->prec, false, p->nz, p->F[0], nullptr, nullptr, b->ana_out[i], nullptr,
                                    nullptr, p->nlm);
        return r ? r : launch_z_forward(p, b->ana_out[i]);
    });
    if (rc || b->small.empty()) return rc;
    rc = per_problem(b, b->small, phi);
    if (rc) return rc;
    rc = launch_batch_analysis(&b->leg, b->stream, nwg(b));
    if (rc) return rc;
    return per_problem(b, b->small, [&](int i) { return launch_z_forward(b->plans[i], b->ana_out[i]); });
}

// The nonlinear term of every problem (rd_nonlinear_uxw per problem).  The
// problems of `large` run their one-problem evaluation over the lanes; those
// of `small` run the radial steps (which also clear the output rows
// nr..nz-1), one persistent vector synthesis over their u and omega sets, the
// fused phi transforms and cross products, and one persistent vector
// analysis.
int bind_uxw(rd_batch *b, double *const *NQ, double *const *NS, double *const *NT)
{
    const int np = (int)b->small.size();
    if (np == 0) return RD_OK;
    std::vector<const double2 *> sin(np * 6), ain(np * 3);
    std::vector<double2 *> sout(np * 6), aout(np * 3);
    for (int j = 0; j < np; j++) {
        const int p = b->small[j];
        Plan *pl = b->plans[p];
        for (int f = 0; f < 6; f++) {
            sin[j * 6 + f] = pl->S[f];
            sout[j * 6 + f] = pl->F[f];
        }
        for (int f = 0; f < 3; f++) ain[j * 3 + f] = pl->F[f];
        aout[j * 3 + 0] = (double2 *)NQ[p];
        aout[j * 3 + 1] = (double2 *)NS[p];
        aout[j * 3 + 2] = (double2 *)NT[p];
    }
    const std::vector<Plan *> pls = plans_of(b, b->small);
    return batch_vleg_bind(&b->leg, pls.data(), sin.data(), sout.data(), ain.data(), aout.data(), b->stream);
}

int enqueue_uxw(rd_batch *b, const double *const *W, const double *const *Z, double *const *NQ,
                double *const *NS, double *const *NT)
{
    int rc = per_problem(b, b->large, [&](int i) {
        return uxw_evaluate(b->plans[i], b->prec, (const double2 *)W[i], (const double2 *)Z[i], (double2 *)NQ[i],
                            (double2 *)NS[i], (double2 *)NT[i], nullptr);
    });
    if (rc || b->small.empty()) return rc;
    rc = per_problem(b, b->small, [&](int i) {
        Plan *p = b->plans[i];
        double2 *outs[3] = {(double2 *)NQ[i], (double2 *)NS[i], (double2 *)NT[i]};
        return launch_uxw_radial(p, (const double2 *)W[i], (const double2 *)Z[i], p->S, 6, outs);
    });
    if (rc) return rc;
    rc = launch_batch_synthesis(&b->leg, b->stream, nwg(b));
    if (rc) return rc;
    rc = per_problem(b, b->small, [&](int i) {
        Plan *p = b->plans[i];
        const double2 *Fin[6] = {p->F[0], p->F[1], p->F[2], p->F[3], p->F[4], p->F[5]};
        double2 *Fout[3] = {p->F[0], p->F[1], p->F[2]};
        return launch_phi_cross(p, p->nr, Fin, Fout, b->prec);
    });
    if (rc) return rc;
    return launch_batch_analysis(&b->leg, b->stream, nwg(b));
}

bool uxw_args_ok(const rd_batch *b, const double *const *W, const double *const *Z, double *const *NQ,
                 double *const *NS, double *const *NT)
{
    if (!b || !b->uxw || !W || !Z || !NQ || !NS || !NT) return false;
    for (int i = 0; i < b->nprob; i++)
        if (!W[i] || !Z[i] || !NQ[i] || !NS[i] || !NT[i]) return false;
    return true;
}

int configure(rd_batch *b, int jb, int wg_per_cu)
{
    const bool okjb = b->uxw ? (jb == 1 || jb == 2) : (jb == 2 || jb == 4);
    if (!okjb || wg_per_cu < 1 || wg_per_cu > 8) return RD_ERR_ARG;
    drop_graph(b);
    RD_HIP(hipStreamSynchronize(b->stream));
    batch_leg_free(&b->leg);
    b->leg = BatchLeg{};
    partition(b);
    // the one-problem chains of a reduced-precision batch read the plans' own
    // precision tables: built here, before any capture
    if (b->prec != RD_PREC_FP64)
        for (int i : b->large) {
            const int rc = ensure_precision_tables(b->plans[i], b->prec);
            if (rc) return rc;
        }
    if (!b->small.empty()) {
        const std::vector<Plan *> pls = plans_of(b, b->small);
        int rc = batch_leg_init(&b->leg, pls.data(), (int)pls.size(), b->prec, jb, b->uxw, b->uxw ? 2 : 1);
        if (rc) return rc;
    }
    b->jb = jb;
    b->wg_per_cu = wg_per_cu;
    return RD_OK;
}

void destroy(rd_batch *b)
{
    if (!b) return;
    (void)hipStreamSynchronize(b->stream);
    drop_graph(b);
    batch_leg_free(&b->leg);
    for (Plan *p : b->plans) {
        if (!p) continue;
        p->stream = p->own_stream;
        plan_destroy(p);
    }
    for (int l = 0; l < kLanes - 1; l++)
        if (b->side[l]) (void)hipStreamDestroy(b->side[l]);
    if (b->ev_fork) (void)hipEventDestroy(b->ev_fork);
    for (int l = 0; l < kLanes - 1; l++)
        if (b->ev_join[l]) (void)hipEventDestroy(b->ev_join[l]);
    if (b->own_stream) (void)hipStreamDestroy(b->own_stream);
    delete b;
}

}  // namespace

extern "C" {

static int batch_create(rd_batch **out, int nprob, const int *lmax, const int *nlat, const int *nphi,
                        const int *nz, int C, int d, const double *dz, const double *rint, const double *dir,
                        int prec, int uxw)
{
    if (!out || nprob < 1 || !lmax || !nlat || !nphi || !nz || !dz || !rint) return RD_ERR_ARG;
    if (prec != RD_PREC_FP64 && prec != RD_PREC_FP32) return RD_ERR_ARG;
    *out = nullptr;
    rd_batch *b = new rd_batch();
    b->nprob = nprob;
    b->prec = prec;
    b->uxw = uxw;
    b->plans.assign(nprob, nullptr);
    b->ana_out.assign(nprob, nullptr);
    bool ok = hipStreamCreateWithFlags(&b->own_stream, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&b->ev_fork, hipEventDisableTiming) == hipSuccess;
    for (int l = 0; l < kLanes - 1 && ok; l++)
        ok = hipStreamCreateWithFlags(&b->side[l], hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&b->ev_join[l], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        destroy(b);
        return RD_ERR_HIP;
    }
    b->stream = b->own_stream;
    int dev = 0;
    (void)hipGetDevice(&dev);
    if (hipDeviceGetAttribute(&b->n_cu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        b->n_cu < 1)
        b->n_cu = 256;
    for (int i = 0; i < nprob; i++) {
        int rc = plan_create(&b->plans[i], lmax[i], nlat[i], nphi[i], nz[i], C, d, dz[i], rint[i], dir,
                             uxw ? 6 : 1);
        if (!rc && uxw && b->plans[i]->nr < 1) rc = RD_ERR_ARG;
        if (rc) {
            destroy(b);
            return rc;
        }
    }
    set_stream(b, b->stream);
    b->split = kSplitDefault;
    int rc = configure(b, uxw ? 1 : 4, 2);
    if (rc) {
        destroy(b);
        return rc;
    }
    *out = b;
    return RD_OK;
}

}  // extern "C"

namespace {

// One timed pass of the tuner over a round trip or a nonlinear-term
// evaluation.
template <typename Fn>
int time_choices(rd_batch *b, int reps, Fn &&run, double *ms_fixed, double *ms_tuned, int *choice)
{
    const int cand_jb_s[2] = {4, 2}, cand_jb_v[2] = {1, 2}, cand_wg[4] = {2, 1, 3, 4};
    const int *cand_jb = b->uxw ? cand_jb_v : cand_jb_s;
    hipEvent_t e0, e1;
    RD_HIP(hipEventCreate(&e0));
    RD_HIP(hipEventCreate(&e1));
    double best = 1e30, fixed = 0.0;
    int bj = cand_jb[0], bw = 2;
    int rc = RD_OK;
    for (int a = 0; a < 2 && !rc; a++)
        for (int c = 0; c < 4 && !rc; c++) {
            rc = configure(b, cand_jb[a], cand_wg[c]);
            // one untimed pass (binds the arguments), then reps timed
            if (!rc) rc = run();
            if (rc) break;
            (void)hipEventRecord(e0, b->stream);
            for (int r = 0; r < reps && !rc; r++) rc = run();
            (void)hipEventRecord(e1, b->stream);
            if (hipEventSynchronize(e1) != hipSuccess) rc = RD_ERR_HIP;
            float ms = 0.f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            const double t = ms / reps;
            if (a == 0 && c == 0) fixed = t;
            if (t < best) {
                best = t;
                bj = cand_jb[a];
                bw = cand_wg[c];
            }
        }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    rc = configure(b, bj, bw);
    if (rc) return rc;
    if (ms_fixed) *ms_fixed = fixed;
    if (ms_tuned) *ms_tuned = best;
    if (choice) {
        choice[0] = bj;
        choice[1] = bw;
    }
    return RD_OK;
}

// The degree threshold of the one-problem chains, at the default tiles (none,
// or every distinct degree of the batch; the default threshold first, its time
// being the untuned one), then the persistent tiles for the threshold kept.
template <typename Fn>
int tune_all(rd_batch *b, int reps, Fn &&run, double *ms_fixed, double *ms_tuned, int *choice)
{
    const int jb0 = b->uxw ? 1 : 4;
    std::vector<int> cand = {kSplitDefault, INT_MAX};
    for (Plan *p : b->plans)
        if (std::find(cand.begin(), cand.end(), p->lmax) == cand.end()) cand.push_back(p->lmax);
    hipEvent_t e0, e1;
    RD_HIP(hipEventCreate(&e0));
    RD_HIP(hipEventCreate(&e1));
    double best = 1e30, fixed = 0.0;
    int bs = kSplitDefault, rc = RD_OK;
    for (size_t c = 0; c < cand.size() && !rc; c++) {
        b->split = cand[c];
        rc = configure(b, jb0, 2);
        if (!rc) rc = run();
        if (rc) break;
        (void)hipEventRecord(e0, b->stream);
        for (int r = 0; r < reps && !rc; r++) rc = run();
        (void)hipEventRecord(e1, b->stream);
        if (hipEventSynchronize(e1) != hipSuccess) rc = RD_ERR_HIP;
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (c == 0) fixed = ms / reps;
        if (ms / reps < best) {
            best = ms / reps;
            bs = cand[c];
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    b->split = bs;
    rc = time_choices(b, reps, run, nullptr, ms_tuned, choice);
    if (ms_fixed) *ms_fixed = fixed;
    return rc;
}

}  // namespace

extern "C" {

int rd_batch_create(rd_batch **out, int nprob, const int *lmax, const int *nlat, const int *nphi, const int *nz,
                    int C, int d, const double *dz, const double *rint, const double *dir, int prec)
{
    return batch_create(out, nprob, lmax, nlat, nphi, nz, C, d, dz, rint, dir, prec, 0);
}

int rd_batch_create_uxw(rd_batch **out, int nprob, const int *lmax, const int *nlat, const int *nphi,
                        const int *nz, int C, int d, const double *dz, const double *rint, const double *dir,
                        int prec)
{
    return batch_create(out, nprob, lmax, nlat, nphi, nz, C, d, dz, rint, dir, prec, 1);
}

int rd_batch_nonlinear_uxw(rd_batch *b, const double *const *W, const double *const *Z, double *const *NQ,
                           double *const *NS, double *const *NT)
{
    if (!uxw_args_ok(b, W, Z, NQ, NS, NT)) return RD_ERR_ARG;
    int rc = bind_uxw(b, NQ, NS, NT);
    if (rc) return rc;
    return enqueue_uxw(b, W, Z, NQ, NS, NT);
}

int rd_batch_capture_uxw(rd_batch *b, const double *const *W, const double *const *Z, double *const *NQ,
                         double *const *NS, double *const *NT)
{
    if (!uxw_args_ok(b, W, Z, NQ, NS, NT)) return RD_ERR_ARG;
    drop_graph(b);
    int rc = bind_uxw(b, NQ, NS, NT);
    if (rc) return rc;
    RD_HIP(hipStreamSynchronize(b->stream));
    const bool leg = !b->small.empty();
    const size_t ns = b->leg.arg_bytes_syn * b->leg.nprob * b->leg.nsets, na = b->leg.arg_bytes_ana * b->leg.nprob;
    if (leg) {
        RD_HIP(hipMalloc(&b->g_syn_args, ns));
        RD_HIP(hipMalloc(&b->g_ana_args, na));
        RD_HIP(hipMalloc(&b->g_counter, 2 * sizeof(int)));
        RD_HIP(hipMemcpy(b->g_syn_args, b->leg.d_syn_args, ns, hipMemcpyDeviceToDevice));
        RD_HIP(hipMemcpy(b->g_ana_args, b->leg.d_ana_args, na, hipMemcpyDeviceToDevice));
    }
    hipStream_t user = b->stream;
    void *sa = b->leg.d_syn_args, *aa = b->leg.d_ana_args;
    int *ct = b->leg.d_counter;
    if (leg) {
        b->leg.d_syn_args = b->g_syn_args;
        b->leg.d_ana_args = b->g_ana_args;
        b->leg.d_counter = b->g_counter;
    }
    set_stream(b, b->own_stream);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(b->own_stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        rc = enqueue_uxw(b, W, Z, NQ, NS, NT);
        e = hipStreamEndCapture(b->own_stream, &g);
    }
    set_stream(b, user);
    b->leg.d_syn_args = sa;
    b->leg.d_ana_args = aa;
    b->leg.d_counter = ct;
    if (rc || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return rc ? rc : RD_ERR_HIP;
    }
    b->graph = g;
    RD_HIP(hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
    b->captured = true;
    return RD_OK;
}

int rd_batch_tune_uxw(rd_batch *b, const double *const *W, const double *const *Z, double *const *NQ,
                      double *const *NS, double *const *NT, int reps, double *ms_fixed, double *ms_tuned,
                      int *choice)
{
    if (!uxw_args_ok(b, W, Z, NQ, NS, NT) || reps < 1) return RD_ERR_ARG;
    return tune_all(
        b, reps,
        [&]() {
            int rc = bind_uxw(b, NQ, NS, NT);
            return rc ? rc : enqueue_uxw(b, W, Z, NQ, NS, NT);
        },
        ms_fixed, ms_tuned, choice);
}

int rd_batch_split(rd_batch *b, int lmax_min)
{
    if (!b || lmax_min < 0) return RD_ERR_ARG;
    b->split = lmax_min;
    return configure(b, b->jb, b->wg_per_cu);
}

int rd_batch_get_split(const rd_batch *b, int *lmax_min, int *n_one)
{
    if (!b || !lmax_min || !n_one) return RD_ERR_ARG;
    *lmax_min = b->split;
    *n_one = (int)b->large.size();
    return RD_OK;
}

int rd_batch_destroy(rd_batch *b)
{
    if (!b) return RD_ERR_ARG;
    destroy(b);
    return RD_OK;
}

int rd_batch_set_stream(rd_batch *b, void *s)
{
    if (!b) return RD_ERR_ARG;
    return set_stream(b, (hipStream_t)s);   // NULL: the default stream
}

int rd_batch_sync(rd_batch *b)
{
    if (!b) return RD_ERR_ARG;
    RD_HIP(hipStreamSynchronize(b->stream));
    return RD_OK;
}

int rd_batch_info(const rd_batch *b, int *info)
{
    if (!b || !info) return RD_ERR_ARG;
    info[0] = b->nprob;
    info[1] = b->prec;
    info[2] = b->jb;
    info[3] = b->wg_per_cu;
    info[4] = b->n_cu;
    info[5] = b->leg.n_syn_items;
    info[6] = b->leg.n_ana_items;
    info[7] = b->captured ? 1 : 0;
    return RD_OK;
}

int rd_batch_3d_inverse(rd_batch *b, const double *const *spec, double *const *grid)
{
    if (!b || !spec || !grid) return RD_ERR_ARG;
    if (b->uxw) return RD_ERR_STATE;
    for (int i = 0; i < b->nprob; i++)
        if (!spec[i] || !grid[i]) return RD_ERR_ARG;
    int rc = bind(b, nullptr);
    if (rc) return rc;
    return enqueue_inverse(b, spec, grid);
}

int rd_batch_3d_forward(rd_batch *b, const double *const *grid, double *const *spec)
{
    if (!b || !grid || !spec) return RD_ERR_ARG;
    if (b->uxw) return RD_ERR_STATE;
    for (int i = 0; i < b->nprob; i++)
        if (!spec[i] || !grid[i]) return RD_ERR_ARG;
    int rc = bind(b, spec);
    if (rc) return rc;
    return enqueue_forward(b, grid);
}

int rd_batch_capture(rd_batch *b, const double *const *spec, double *const *grid, double *const *spec_out)
{
    if (!b || !spec || !grid || !spec_out) return RD_ERR_ARG;
    if (b->uxw) return RD_ERR_STATE;
    for (int i = 0; i < b->nprob; i++)
        if (!spec[i] || !grid[i] || !spec_out[i]) return RD_ERR_ARG;
    drop_graph(b);
    int rc = bind(b, spec_out);
    if (rc) return rc;
    RD_HIP(hipStreamSynchronize(b->stream));
    const bool leg = !b->small.empty();
    const size_t ns = b->leg.arg_bytes_syn * b->leg.nprob, na = b->leg.arg_bytes_ana * b->leg.nprob;
    if (leg) {
        RD_HIP(hipMalloc(&b->g_syn_args, ns));
        RD_HIP(hipMalloc(&b->g_ana_args, na));
        RD_HIP(hipMalloc(&b->g_counter, 2 * sizeof(int)));
        RD_HIP(hipMemcpy(b->g_syn_args, b->leg.d_syn_args, ns, hipMemcpyDeviceToDevice));
        RD_HIP(hipMemcpy(b->g_ana_args, b->leg.d_ana_args, na, hipMemcpyDeviceToDevice));
    }
    // record on the batch's own (non-default) stream, with the graph's copies
    // of the arguments and counters; the graph is launched later on whichever
    // stream the batch is bound to
    hipStream_t user = b->stream;
    void *sa = b->leg.d_syn_args, *aa = b->leg.d_ana_args;
    int *ct = b->leg.d_counter;
    if (leg) {
        b->leg.d_syn_args = b->g_syn_args;
        b->leg.d_ana_args = b->g_ana_args;
        b->leg.d_counter = b->g_counter;
    }
    set_stream(b, b->own_stream);
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamBeginCapture(b->own_stream, hipStreamCaptureModeThreadLocal);
    if (e == hipSuccess) {
        rc = enqueue_inverse(b, spec, grid);
        if (!rc) rc = enqueue_forward(b, grid);
        e = hipStreamEndCapture(b->own_stream, &g);
    }
    set_stream(b, user);
    b->leg.d_syn_args = sa;
    b->leg.d_ana_args = aa;
    b->leg.d_counter = ct;
    if (rc || e != hipSuccess) {
        if (g) (void)hipGraphDestroy(g);
        return rc ? rc : RD_ERR_HIP;
    }
    b->graph = g;
    RD_HIP(hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0));
    b->captured = true;
    return RD_OK;
}

int rd_batch_replay(rd_batch *b)
{
    if (!b) return RD_ERR_ARG;
    if (!b->captured) return RD_ERR_STATE;
    RD_HIP(hipGraphLaunch(b->gexec, b->stream));
    return RD_OK;
}

int rd_batch_config(rd_batch *b, int jb, int wg_per_cu)
{
    if (!b) return RD_ERR_ARG;
    return configure(b, jb, wg_per_cu);
}

int rd_batch_tune(rd_batch *b, const double *const *spec, double *const *grid, double *const *spec_out,
                  int reps, double *ms_fixed, double *ms_tuned, int *choice)
{
    if (!b || !spec || !grid || !spec_out || reps < 1) return RD_ERR_ARG;
    if (b->uxw) return RD_ERR_STATE;
    return tune_all(
        b, reps,
        [&]() {
            int rc = bind(b, spec_out);
            if (!rc) rc = enqueue_inverse(b, spec, grid);
            return rc ? rc : enqueue_forward(b, grid);
        },
        ms_fixed, ms_tuned, choice);
}

}  // extern "C"
