// msj_noalloc.hip -- msj_self_join_noalloc: the whole join with no device
// allocation, for callers that cannot allocate (serving loops that must not
// call hipMalloc, code that runs between a caller's own allocations).
//
// Every device buffer the join needs beyond the caller's inputs and outputs
// (the column copy of the rows, node-net scratch, FP64 join rows, result
// keys and sort buffers, filter tiles) is carved front to back out of
// `scratch`, one file-scope device array of MSJ_NOALLOC_SCRATCH_BYTES: the
// join pipeline's buffers borrow from it while the call runs (msj::Arena).
// Each device holds its own copy of the array from the first kernel launch
// of this library on it until the process ends.
//
// Set-up, once per device with that device current: the device address of
// `scratch` (hipGetSymbolAddress; in host code the array's name is not a
// device address), and the timing events.  Calls on one device share its
// array: a per-device lock runs them one at a time, and each call clears
// the part of the array the previous call used before it starts.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "msj_internal.h"

#pragma clang fp contract(off)

// Device scratch of the allocation-free join: each device holds one copy.
// Host code reaches it only through the per-device address from
// hipGetSymbolAddress (DeviceSlot::ptr), never through this name.
static __device__ __align__(256) unsigned char scratch[MSJ_NOALLOC_SCRATCH_BYTES];

namespace {

using msj::Error;

constexpr int MAX_DEVICES = 64;

struct DeviceSlot {
    std::once_flag once;
    int status = MSJ_OK;         // result of the set-up
    std::string error;
    unsigned char *ptr = nullptr;  // device address of `scratch` on this device
    hipEvent_t ev[8] = {};
    size_t dirty = 0;            // bytes of the array the last call used
    std::mutex mu;
};

DeviceSlot g_slot[MAX_DEVICES];

void setup(DeviceSlot &sl)
{
    // the caller has made the device current
    hipError_t e = hipGetSymbolAddress((void **)&sl.ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !sl.ptr) {
        (void)hipGetLastError();
        sl.status = MSJ_E_HIP;
        sl.error = std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e);
        sl.ptr = nullptr;
        return;
    }
    for (auto &ev : sl.ev) {
        e = hipEventCreate(&ev);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            sl.status = MSJ_E_HIP;
            sl.error = std::string("hipEventCreate failed: ") + hipGetErrorString(e);
            return;
        }
    }
    // the array starts as the loader left it: clear it once
    e = hipMemset(sl.ptr, 0, MSJ_NOALLOC_SCRATCH_BYTES);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        sl.status = MSJ_E_HIP;
        sl.error = std::string("clearing the scratch failed: ") + hipGetErrorString(e);
    }
}

struct Args {
    const void *rows;
    uint32_t n, d;
    double eps;
    const msj_params *p;
    uint32_t *counts;
    uint64_t *offsets;
    uint32_t *nbrs;
    uint64_t nbrs_cap;
    uint32_t *node_order;
    msj_stats *stats;
};

void run(void *a_)
{
    Args &a = *(Args *)a_;
    const auto t0 = std::chrono::steady_clock::now();
    msj_params prm;
    if (a.p)
        prm = *a.p;
    else
        msj_default_params(&prm);
    msj::validate_params(prm);
    if (a.stats)
        std::memset(a.stats, 0, sizeof(*a.stats));
    if (a.d == 0)
        throw Error(MSJ_E_INVAL, "d must be positive");
    if (!(a.eps > 0.0) || !std::isfinite(a.eps))
        throw Error(MSJ_E_INVAL, "eps must be positive and finite");
    if (a.n > 0 && (!a.rows || !a.counts || !a.nbrs || !a.node_order))
        throw Error(MSJ_E_INVAL, "rows, counts, nbrs and node_order must be device pointers");
    if (!a.offsets)
        throw Error(MSJ_E_INVAL, "offsets must be a device pointer");
    const bool f32 = prm.storage == MSJ_STORAGE_F32;
    // the column copy and the FP64 join rows alone must fit
    const uint64_t data_bytes =
        (uint64_t)a.n * a.d * (f32 ? 4 : 8) + (uint64_t)a.n * msj::padded_dim(a.d) * 8;
    if (data_bytes + (uint64_t)a.n * 64 > MSJ_NOALLOC_SCRATCH_BYTES)
        throw Error(MSJ_E_OVERFLOW,
                    std::to_string(a.n) + " x " + std::to_string(a.d) +
                        " rows do not fit the allocation-free scratch (" +
                        std::to_string(data_bytes) + " bytes for the data copies alone, " +
                        std::to_string((uint64_t)MSJ_NOALLOC_SCRATCH_BYTES) + " available)");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        throw Error(MSJ_E_NODEV, "no HIP device visible");
    }
    if (prm.device < 0 || prm.device >= ndev || prm.device >= MAX_DEVICES)
        throw Error(MSJ_E_NODEV, "device ordinal out of range");
    int prev = -1;
    if (hipGetDevice(&prev) != hipSuccess)
        prev = -1;
    MSJ_HIP(hipSetDevice(prm.device));
    struct Restore {
        int prev;
        ~Restore()
        {
            if (prev >= 0)
                (void)hipSetDevice(prev);
        }
    } restore{prev};
    DeviceSlot &sl = g_slot[prm.device];
    std::call_once(sl.once, [&] { setup(sl); });
    if (sl.status != MSJ_OK)
        throw Error(sl.status, sl.error);

    std::lock_guard<std::mutex> lk(sl.mu);
    hipStream_t s = nullptr;  // the device's null stream
    if (sl.dirty)  // what the previous call left behind starts at zero again
        MSJ_HIP(hipMemsetAsync(sl.ptr, 0, sl.dirty, s));
    msj::Arena arena{(char *)sl.ptr, (size_t)MSJ_NOALLOC_SCRATCH_BYTES, 0};
    struct Dirty {
        DeviceSlot &sl;
        msj::Arena &ar;
        ~Dirty() { sl.dirty = std::max(sl.dirty, ar.used); }
    } dirty{sl, arena};
    msj::ArenaScope scope(&arena);

    msj_stats st;
    std::memset(&st, 0, sizeof(st));
    if (a.n == 0) {
        MSJ_HIP(hipMemsetAsync(a.offsets, 0, sizeof(uint64_t), s));
        MSJ_HIP(hipStreamSynchronize(s));
        if (a.stats)
            *a.stats = st;
        return;
    }
    msj::DevBuf<unsigned char> xs;
    xs.ensure((size_t)a.n * a.d * (f32 ? 4 : 8));
    msj::rows_to_soa_device(a.rows, a.n, a.d, f32, xs.p, s);
    msj::JoinCore core;
    // cap 0: the result-key capacity is taken from what the arena has left
    // once the index is built
    const uint64_t R = core.run(xs.p, a.rows, f32, a.n, a.d, a.eps, prm, 0, s, sl.ev, st);
    st.total_pairs = R;
    if (R > a.nbrs_cap) {
        core.batch_record(st);
        if (a.stats)
            *a.stats = st;  // total_pairs tells the caller how much room to give
        throw Error(MSJ_E_OVERFLOW, std::to_string(R) + " result pairs exceed nbrs_cap " +
                                        std::to_string(a.nbrs_cap));
    }
    core.finish(a.n, R, a.counts, a.offsets, a.nbrs, s, sl.ev, st);
    MSJ_HIP(hipMemcpyAsync(a.node_order, core.ix.perm.p, sizeof(uint32_t) * a.n,
                           hipMemcpyDeviceToDevice, s));
    MSJ_HIP(hipStreamSynchronize(s));
    st.ms_total = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0)
                      .count();
    if (a.stats)
        *a.stats = st;
}

}  // namespace

extern "C" int msj_self_join_noalloc(const void *rows, uint32_t n, uint32_t d, double eps,
                                     const msj_params *p, uint32_t *counts, uint64_t *offsets,
                                     uint32_t *nbrs, uint64_t nbrs_cap, uint32_t *node_order,
                                     msj_stats *stats)
{
    Args a{rows, n, d, eps, p, counts, offsets, nbrs, nbrs_cap, node_order, stats};
    return msj::guarded_call(run, &a);
}
