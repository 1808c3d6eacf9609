// The scalar synthesis (fftp2d_complex_to_real_xy) without any allocation:
// its Fourier coefficients F[k][m][theta] go to a fixed device array of this
// file instead of the plan's scratch, so that callers that cannot allocate
// (code under stream capture, serving loops, harnesses that forbid hipMalloc)
// can call it with only the plan (built beforehand) and their own input and
// output.  Same kernels, same tables, same results as rd_sht_synthesis.
#include "rd_internal.h"

#include <mutex>

namespace rd {

namespace {

// Each device the library runs on holds one copy of this array until the
// process ends.  Host code never uses its name as an address: the device
// address comes from hipGetSymbolAddress, once per device, with that device
// current.
constexpr size_t kScratchElems = RD_NOALLOC_SCRATCH_BYTES / sizeof(double2);
__device__ __align__(256) double2 scratch[kScratchElems];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double2 *g_scratch[kMaxDevices];

// The device address of this device's scratch array (looked up on the first
// call on that device).
int scratch_on(int device, double2 **out)
{
    if (device < 0 || device >= kMaxDevices) return RD_ERR_ARG;
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[device]) {
        int cur = -1;
        RD_HIP(hipGetDevice(&cur));
        if (cur != device) RD_HIP(hipSetDevice(device));
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (cur != device) RD_HIP(hipSetDevice(cur));
        if (e != hipSuccess || !ptr) {
            fprintf(stderr, "redondito: hipGetSymbolAddress(scratch) failed: %s\n", hipGetErrorString(e));
            return RD_ERR_HIP;
        }
        g_scratch[device] = (double2 *)ptr;
    }
    *out = g_scratch[device];
    return RD_OK;
}

}  // namespace

}  // namespace rd

using namespace rd;

extern "C" int rd_sht_synthesis_noalloc(rd_plan *plan, int nplanes, const double *in, double *out)
{
    if (!plan || !in || !out) return RD_ERR_ARG;
    Plan *p = reinterpret_cast<Plan *>(plan);
    if (nplanes < 0 || nplanes > p->nz) return RD_ERR_ARG;
    if (nplanes == 0) return RD_OK;
    const size_t need = sizeof(double2) * (size_t)nplanes * p->L * p->nlat;
    if (need > RD_NOALLOC_SCRATCH_BYTES) {
        fprintf(stderr,
                "redondito: rd_sht_synthesis_noalloc needs %zu bytes of Fourier scratch for %d planes "
                "(lmax %d, nlat %d); the fixed scratch holds %zu\n",
                need, nplanes, p->lmax, p->nlat, (size_t)RD_NOALLOC_SCRATCH_BYTES);
        return RD_ERR_ARG;
    }
    double2 *F = nullptr;
    int rc = scratch_on(p->device, &F);
    if (rc) return rc;
    // The synthesis writes every (plane, m <= lmax, theta) entry the phi
    // transform then reads, so nothing in the scratch needs clearing.
    rc = launch_synthesis(p, RD_PREC_FP64, false, nplanes, (const double2 *)in, nullptr, nullptr, p->nlm, F,
                          nullptr, nullptr);
    if (rc) return rc;
    const double2 *Fin[1] = {F};
    double *f[1] = {out};
    return launch_phi_inverse(p, 1, nplanes, Fin, f);
}
