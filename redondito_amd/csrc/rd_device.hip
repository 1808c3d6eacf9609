// Device selection and the host-address -> device-copy registry that the
// Fortran drop-in layer uses to find the device copy of a caller's array.
// Replaces src/memmgt/gdevice_mod.f90 (device_init :112-141, gdev_alloc
// :158-193, gdev_free :196-217, gdev_lookup :220-235, gdev_update_to/from
// :238-279, gdev_copy :282-302, gdev_present :305-317) with plain HIP: no
// OpenMP-offload association, the registry is the only mapping.  An interior
// host address (a plane of a registered array) resolves to the same offset
// in the device copy.
#include "rd_internal.h"

#include <map>
#include <mutex>

namespace {

struct Entry {
    void *dp;
    size_t nb;
};

std::mutex g_mu;
std::map<uintptr_t, Entry> g_reg;   // keyed by host base address

// Entry containing host address h (with byte offset), or null.
const Entry *find(const void *h, size_t *off)
{
    const uintptr_t a = (uintptr_t)h;
    auto it = g_reg.upper_bound(a);
    if (it == g_reg.begin()) return nullptr;
    --it;
    if (a >= it->first + it->second.nb) return nullptr;
    *off = a - it->first;
    return &it->second;
}

}  // namespace

extern "C" {

int rd_device_init(int rank)
{
    int ndev = 0;
    RD_HIP(hipGetDeviceCount(&ndev));
    if (ndev <= 0) {
        fprintf(stderr, "redondito: no GPU device found\n");
        return RD_ERR_STATE;
    }
    RD_HIP(hipSetDevice((rank < 0 ? 0 : rank) % ndev));
    return RD_OK;
}

int rd_device_sync(void)
{
    RD_HIP(hipDeviceSynchronize());
    return RD_OK;
}

int rd_dev_alloc(const void *host, size_t nbytes)
{
    if (!host) return RD_ERR_ARG;
    if (nbytes == 0) return RD_OK;
    std::lock_guard<std::mutex> lk(g_mu);
    size_t off = 0;
    if (find(host, &off)) return RD_ERR_ARG;       // already registered
    void *dp = nullptr;
    if (hipMalloc(&dp, nbytes) != hipSuccess) return RD_ERR_NOMEM;
    g_reg[(uintptr_t)host] = Entry{dp, nbytes};
    return RD_OK;
}

int rd_dev_free(const void *host)
{
    std::lock_guard<std::mutex> lk(g_mu);
    auto it = g_reg.find((uintptr_t)host);
    if (it == g_reg.end()) return RD_OK;
    RD_HIP(hipFree(it->second.dp));
    g_reg.erase(it);
    return RD_OK;
}

void *rd_dev_lookup(const void *host)
{
    std::lock_guard<std::mutex> lk(g_mu);
    size_t off = 0;
    const Entry *e = find(host, &off);
    return e ? (void *)((char *)e->dp + off) : nullptr;
}

int rd_dev_present(const void *host) { return rd_dev_lookup(host) != nullptr; }

int rd_dev_index(const void *host)
{
    std::lock_guard<std::mutex> lk(g_mu);
    int i = 1;
    for (const auto &kv : g_reg) {
        if (kv.first == (uintptr_t)host) return i;
        i++;
    }
    return 0;
}

static int span(const void *host, size_t nbytes, void **dp)
{
    std::lock_guard<std::mutex> lk(g_mu);
    size_t off = 0;
    const Entry *e = find(host, &off);
    if (!e) {
        fprintf(stderr, "redondito: array at %p has no device copy\n", host);
        return RD_ERR_STATE;
    }
    if (off + nbytes > e->nb) return RD_ERR_ARG;
    *dp = (char *)e->dp + off;
    return RD_OK;
}

int rd_dev_update_to(const void *host, size_t nbytes)
{
    void *dp = nullptr;
    int rc = span(host, nbytes, &dp);
    if (rc) return rc;
    RD_HIP(hipMemcpy(dp, host, nbytes, hipMemcpyHostToDevice));
    return RD_OK;
}

int rd_dev_update_from(void *host, size_t nbytes)
{
    void *dp = nullptr;
    int rc = span(host, nbytes, &dp);
    if (rc) return rc;
    RD_HIP(hipMemcpy(host, dp, nbytes, hipMemcpyDeviceToHost));
    return RD_OK;
}

int rd_dev_copy(const void *hdst, const void *hsrc, size_t nbytes)
{
    void *d = nullptr, *s = nullptr;
    int rc = span(hdst, nbytes, &d);
    if (!rc) rc = span(hsrc, nbytes, &s);
    if (rc) return rc;
    RD_HIP(hipMemcpy(d, s, nbytes, hipMemcpyDeviceToDevice));
    return RD_OK;
}

}  // extern "C"
