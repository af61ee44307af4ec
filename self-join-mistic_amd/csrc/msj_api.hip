// msj_api.hip -- the C ABI (include/msj.h): argument checking, device
// context, stage timing and result hand-over.
//
// Replaces nodeLauncher5 (reference launcher.cu:3093-3532) and the host helper
// stddev (utils.cu:22-53) used by main (main.cu:111).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "msj_internal.h"

#pragma clang fp contract(off)

namespace {

thread_local std::string g_last_error;

void set_error(const std::string &m) { g_last_error = m; }

}  // namespace

namespace msj {

thread_local Arena *g_arena = nullptr;

void hip_check(hipError_t e, const char *what, const char *file, int line)
{
    if (e == hipSuccess)
        return;
    (void)hipGetLastError();
    int code = (e == hipErrorOutOfMemory) ? MSJ_E_NOMEM : MSJ_E_HIP;
    throw Error(code, std::string(what) + " failed at " + file + ":" + std::to_string(line) + ": " +
                          hipGetErrorString(e));
}

}  // namespace msj

// Page-locked host buffer (hipHostMalloc); grow-only.
struct PinnedBuf {
    void *p = nullptr;
    size_t cap = 0;
    PinnedBuf() = default;
    PinnedBuf(const PinnedBuf &) = delete;
    PinnedBuf &operator=(const PinnedBuf &) = delete;
    ~PinnedBuf() { release(); }
    void ensure(size_t bytes)
    {
        if (bytes <= cap)
            return;
        release();
        hipError_t e = hipHostMalloc(&p, bytes, hipHostMallocDefault);
        if (e != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            throw msj::Error(MSJ_E_NOMEM, std::string("hipHostMalloc of ") +
                                              std::to_string(bytes) + " bytes failed");
        }
        cap = bytes;
    }
    void release()
    {
        if (p)
            (void)hipHostFree(p);
        p = nullptr;
        cap = 0;
    }
};

struct msj_batch {
    msj_params prm;
    msj::BatchEngine *e = nullptr;
    uint32_t njobs = 0;
};

struct msj_ctx {
    msj_params prm;
    hipStream_t stream = nullptr;
    hipStream_t copy_stream = nullptr;  // result pages to host
    hipEvent_t ev[8] = {};
    hipEvent_t page_ev[2] = {};
    PinnedBuf pin[2];                   // double buffer of result pages
    uint32_t n = 0, d = 0;
    bool f32 = false;
    msj::DevBuf<unsigned char> staging;  // row-major upload (empty after an unordered load)
    msj::DevBuf<unsigned char> xs;       // SoA, row order
    msj::JoinCore core;                  // index, join and their device state
    msj::DevBuf<uint32_t> r_counts;
    msj::DevBuf<uint64_t> r_offsets;
    msj::DevBuf<uint32_t> r_nbrs;
    uint64_t last_R = 0;
};

namespace {

using msj::Error;

struct DeviceGuard {
    int prev = -1;
    explicit DeviceGuard(int dev)
    {
        if (hipGetDevice(&prev) != hipSuccess)
            prev = -1;
        MSJ_HIP(hipSetDevice(dev));
    }
    ~DeviceGuard()
    {
        if (prev >= 0)
            (void)hipSetDevice(prev);
    }
};

}  // namespace

void msj::validate_params(const msj_params &p)
{
    if (p.struct_size != sizeof(msj_params))
        throw Error(MSJ_E_INVAL, "msj_params.struct_size mismatch: call msj_default_params first");
    if (p.maxrp == 0 || p.maxrp > MSJ_MAX_LAYERS)
        throw Error(MSJ_E_INVAL, "maxrp must be in 1..16");
    if (p.rp_per_layer == 0 || p.rp_per_layer > 32)
        throw Error(MSJ_E_INVAL, "rp_per_layer must be in 1..32");
    if (!(p.sample_per >= 0.0 && p.sample_per <= 1.0))
        throw Error(MSJ_E_INVAL, "sample_per must be in [0, 1]");
    if (p.storage != MSJ_STORAGE_F64 && p.storage != MSJ_STORAGE_F32)
        throw Error(MSJ_E_INVAL, "unknown storage");
    if (p.output != MSJ_OUT_HOST && p.output != MSJ_OUT_DEVICE)
        throw Error(MSJ_E_INVAL, "unknown output mode");
    if (p.form != MSJ_FORM_PLAIN && p.form != MSJ_FORM_MFMA_BF16 && p.form != MSJ_FORM_MFMA_SPLIT)
        throw Error(MSJ_E_INVAL, "unknown form");
    if (p.index != MSJ_INDEX_NODENET && p.index != MSJ_INDEX_BRUTE && p.index != MSJ_INDEX_COSS)
        throw Error(MSJ_E_INVAL, "unknown candidate generator (index)");
}

namespace {

template <class F>
int guarded(F &&f)
{
    try {
        f();
        return MSJ_OK;
    } catch (const Error &e) {
        set_error(e.what());
        return e.code;
    } catch (const std::bad_alloc &) {
        set_error("host allocation failed");
        return MSJ_E_NOMEM;
    } catch (const std::exception &e) {
        set_error(e.what());
        return MSJ_E_HIP;
    }
}

__global__ void k_soa(const unsigned char *__restrict__ rm, unsigned char *__restrict__ soa,
                      uint32_t n, uint32_t d, uint32_t esz)
{
    const size_t total = (size_t)n * d, stride = (size_t)gridDim.x * blockDim.x;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += stride) {
        uint32_t j = (uint32_t)(idx / n), i = (uint32_t)(idx - (size_t)j * n);
        if (esz == 8)
            ((double *)soa)[idx] = ((const double *)rm)[(size_t)i * d + j];
        else
            ((float *)soa)[idx] = ((const float *)rm)[(size_t)i * d + j];
    }
}

// Rows [row0, row0 + m) of an n-row table, row-major FP64 in `chunk`, into
// the FP64 columns soa[j * n + row].
__global__ void k_rows_to_soa(const double *__restrict__ chunk, uint32_t m, uint32_t d,
                              uint32_t row0, uint32_t n, double *__restrict__ soa)
{
    const size_t total = (size_t)m * d, stride = (size_t)gridDim.x * blockDim.x;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += stride) {
        const uint32_t j = (uint32_t)(idx / m), i = (uint32_t)(idx - (size_t)j * m);
        soa[(size_t)j * n + row0 + i] = chunk[(size_t)i * d + j];
    }
}

// xs[j * n + i] = soa[order[j] * n + i] at the storage precision (FP32:
// rounded to nearest, as a host (float) cast).
template <class T>
__global__ void k_permute_cols(const double *__restrict__ soa, const uint32_t *__restrict__ order,
                               uint32_t n, uint32_t d, T *__restrict__ xs)
{
    const size_t total = (size_t)n * d, stride = (size_t)gridDim.x * blockDim.x;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += stride) {
        const uint32_t j = (uint32_t)(idx / n), i = (uint32_t)(idx - (size_t)j * n);
        xs[idx] = (T)soa[(size_t)order[j] * n + i];
    }
}

// utils.cu:42-48: dimensions by decreasing sigma; thrust's stable ascending
// radix order of the doubles, then reversed (ties: higher index first).
void order_by_sigma(const double *dev, uint32_t d, uint32_t *order, double *sigma)
{
    auto rkey = [](double x) {
        uint64_t u;
        std::memcpy(&u, &x, 8);
        return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
    };
    std::vector<uint32_t> asc(d);
    for (uint32_t i = 0; i < d; ++i)
        asc[i] = i;
    std::stable_sort(asc.begin(), asc.end(),
                     [&](uint32_t a, uint32_t b) { return rkey(dev[a]) < rkey(dev[b]); });
    for (uint32_t i = 0; i < d; ++i) {
        order[i] = asc[d - 1 - i];
        if (sigma)
            sigma[i] = dev[asc[d - 1 - i]];
    }
}

}  // namespace

void msj::rows_to_soa_device(const void *rows, uint32_t n, uint32_t d, bool f32, void *xs,
                             hipStream_t s)
{
    const size_t total = (size_t)n * d;
    if (total == 0)
        return;
    msj::check_grid(msj::grid_stride(total, 256), 256);
    hipLaunchKernelGGL(k_soa, dim3(msj::grid_stride(total, 256)), dim3(256), 0, s,
                       (const unsigned char *)rows, (unsigned char *)xs, n, d, f32 ? 4u : 8u);
    MSJ_HIP(hipGetLastError());
}

void msj::load_soa(const void *data, uint32_t n, uint32_t d, bool f32,
                   msj::DevBuf<unsigned char> &xs, msj::DevBuf<unsigned char> &staging,
                   hipStream_t s)
{
    const uint32_t esz = f32 ? 4 : 8;
    const size_t bytes = (size_t)n * d * esz;
    if (bytes == 0)
        return;
    staging.ensure(bytes);
    xs.ensure(bytes);
    MSJ_HIP(hipMemcpyAsync(staging.p, data, bytes, hipMemcpyHostToDevice, s));
    msj::check_grid(msj::grid_stride((size_t)n * d, 256), 256);
    hipLaunchKernelGGL(k_soa, dim3(msj::grid_stride((size_t)n * d, 256)), dim3(256), 0, s,
                       staging.p, xs.p, n, d, esz);
    MSJ_HIP(hipGetLastError());
    MSJ_HIP(hipStreamSynchronize(s));
}

namespace {

// launcher.cu:3203-3233: the reference's batch plan, kept for its stderr record.
void reference_batch_plan(const std::vector<uint32_t> &nstart, const std::vector<uint64_t> &ncalcs,
                          uint32_t n, double eps, uint64_t sum_calcs, uint32_t *batches,
                          double *var_calc)
{
    double varCalc = sum_calcs * 1.0 / n / n;
    unsigned batchCount = 0;
    unsigned long long calcCounter = 0;
    unsigned pointCounter = 0;
    double factor1 = (1 + varCalc * CMP) * 2000 / std::sqrt((double)n);
    double factor2 = CALC_PER_BATCH * std::pow(10, 10) * 1.0 * std::sqrt(std::pow(1 + eps, 20));
    size_t M = ncalcs.size();
    for (size_t i = 0; i < M; ++i) {
        uint32_t sz = nstart[i + 1] - nstart[i];
        for (uint32_t j = 0; j < sz; ++j) {
            calcCounter += ncalcs[i] / sz;
            pointCounter++;
            if (calcCounter * factor1 >= factor2 || pointCounter >= KERNEL_BLOCKS * BLOCK_SIZE) {
                batchCount++;
                calcCounter = 0;
                pointCounter = 0;
            }
        }
    }
    *batches = batchCount + 1;
    *var_calc = varCalc;
}

void do_load(msj_ctx *c, const void *data, uint32_t n, uint32_t d)
{
    if (d == 0)
        throw Error(MSJ_E_INVAL, "d must be positive");
    if (n > 0 && !data)
        throw Error(MSJ_E_INVAL, "data is NULL");
    if ((uint64_t)n * d >= (1ull << 40))
        throw Error(MSJ_E_OVERFLOW, "n*d too large");
    DeviceGuard g(c->prm.device);
    c->f32 = c->prm.storage == MSJ_STORAGE_F32;
    c->n = n;
    c->d = d;
    c->staging.release();
    if (n == 0)
        return;
    // the row-major upload stays: joins gather whole rows from it
    msj::load_soa(data, n, d, c->f32, c->xs, c->staging, c->stream);
}

// main.cu:42-50 + 106-122 on the device: n rows of d FP64 values in file
// order stream through the pinned double buffer (the host copies chunk k+1
// while chunk k crosses the link and is scattered into FP64 columns), the
// per-column sigma is summed on the device in the reference's order, and
// the columns are kept in decreasing-sigma order at the storage precision.
// `rows` may be a read-only file mapping: it is read once, sequentially.
void do_load_unordered(msj_ctx *c, const double *rows, uint32_t n, uint32_t d,
                       uint32_t *order_out, double *sigma_out)
{
    if (d == 0)
        throw Error(MSJ_E_INVAL, "d must be positive");
    if (n > 0 && !rows)
        throw Error(MSJ_E_INVAL, "data is NULL");
    if ((uint64_t)n * d >= (1ull << 40))
        throw Error(MSJ_E_OVERFLOW, "n*d too large");
    DeviceGuard g(c->prm.device);
    hipStream_t s = c->stream;
    c->f32 = c->prm.storage == MSJ_STORAGE_F32;
    c->n = 0;
    c->d = d;
    c->staging.release();  // the rows arrive in file order, not in column order: no copy kept
    std::vector<uint32_t> order(d);
    std::vector<double> sig(d, 0.0);
    if (n == 0) {
        for (uint32_t j = 0; j < d; ++j)
            order[j] = j;
    } else {
        msj::DevBuf<double> soa, chunk[2], dsig;
        msj::DevBuf<uint32_t> dorder;
        soa.ensure((size_t)n * d);
        const size_t chunk_bytes = (size_t)16 << 20;
        const uint32_t R = (uint32_t)std::max<size_t>(1, chunk_bytes / (sizeof(double) * d));
        for (auto &b : chunk)
            b.ensure((size_t)R * d);
        c->pin[0].ensure((size_t)R * d * sizeof(double));
        c->pin[1].ensure((size_t)R * d * sizeof(double));
        const uint32_t nchunks = (uint32_t)((n + (uint64_t)R - 1) / R);
        for (uint32_t k = 0; k < nchunks; ++k) {
            const uint32_t row0 = k * R, m = std::min<uint32_t>(R, n - row0);
            const size_t bytes = (size_t)m * d * sizeof(double);
            if (k >= 2)  // the copy out of this pinned buffer (chunk k-2) is done
                MSJ_HIP(hipEventSynchronize(c->page_ev[k & 1]));
            std::memcpy(c->pin[k & 1].p, rows + (size_t)row0 * d, bytes);
            MSJ_HIP(hipMemcpyAsync(chunk[k & 1].p, c->pin[k & 1].p, bytes, hipMemcpyHostToDevice,
                                   s));
            MSJ_HIP(hipEventRecord(c->page_ev[k & 1], s));
            const size_t cnt = (size_t)m * d;
            msj::check_grid(msj::grid_stride(cnt, 256), 256);
            hipLaunchKernelGGL(k_rows_to_soa, dim3(msj::grid_stride(cnt, 256)), dim3(256), 0, s,
                               chunk[k & 1].p, m, d, row0, n, soa.p);
            MSJ_HIP(hipGetLastError());
        }
        dsig.ensure(d);
        msj::column_sigma(soa.p, n, d, 1, n, dsig.p, s);
        std::vector<double> dev(d);
        MSJ_HIP(hipMemcpyAsync(dev.data(), dsig.p, sizeof(double) * d, hipMemcpyDeviceToHost, s));
        MSJ_HIP(hipStreamSynchronize(s));
        order_by_sigma(dev.data(), d, order.data(), sig.data());
        dorder.ensure(d);
        MSJ_HIP(hipMemcpyAsync(dorder.p, order.data(), sizeof(uint32_t) * d,
                               hipMemcpyHostToDevice, s));
        const size_t total = (size_t)n * d;
        c->xs.ensure(total * (c->f32 ? 4 : 8));
        msj::check_grid(msj::grid_stride(total, 256), 256);
        if (c->f32)
            hipLaunchKernelGGL(k_permute_cols<float>, dim3(msj::grid_stride(total, 256)),
                               dim3(256), 0, s, soa.p, dorder.p, n, d, (float *)c->xs.p);
        else
            hipLaunchKernelGGL(k_permute_cols<double>, dim3(msj::grid_stride(total, 256)),
                               dim3(256), 0, s, soa.p, dorder.p, n, d, (double *)c->xs.p);
        MSJ_HIP(hipGetLastError());
        MSJ_HIP(hipStreamSynchronize(s));
    }
    c->n = n;
    if (order_out)
        std::memcpy(order_out, order.data(), sizeof(uint32_t) * d);
    if (sigma_out)
        std::memcpy(sigma_out, sig.data(), sizeof(double) * d);
}

// Device -> host copy in pages through two pinned buffers on the copy
// stream: the DMA of page k+1 runs while the host copies page k out of the
// other buffer (the reference's batch loop pages results the same way,
// launcher.cu:3418-3506).  `ready` is recorded on the compute stream after
// the source is complete.  Returns the number of pages.
uint64_t page_to_host(msj_ctx *c, void *dst, const void *src, size_t bytes, size_t page,
                      hipEvent_t ready)
{
    if (bytes == 0)
        return 0;
    page = std::min(page, bytes);  // a short transfer pins no more than it moves
    c->pin[0].ensure(page);
    c->pin[1].ensure(page);
    MSJ_HIP(hipStreamWaitEvent(c->copy_stream, ready, 0));
    const size_t np = (bytes + page - 1) / page;
    auto issue = [&](size_t k) {
        const size_t off = k * page, len = std::min(page, bytes - off);
        MSJ_HIP(hipMemcpyAsync(c->pin[k & 1].p, (const char *)src + off, len,
                               hipMemcpyDeviceToHost, c->copy_stream));
        MSJ_HIP(hipEventRecord(c->page_ev[k & 1], c->copy_stream));
    };
    issue(0);
    for (size_t k = 0; k < np; ++k) {
        if (k + 1 < np)
            issue(k + 1);  // its buffer was drained by the host at step k-1
        MSJ_HIP(hipEventSynchronize(c->page_ev[k & 1]));
        const size_t off = k * page, len = std::min(page, bytes - off);
        std::memcpy((char *)dst + off, c->pin[k & 1].p, len);
    }
    return np;
}

float elapsed(hipEvent_t a, hipEvent_t b)
{
    float ms = 0.f;
    MSJ_HIP(hipEventElapsedTime(&ms, a, b));
    return ms;
}

void do_join(msj_ctx *c, double eps, msj_result *out)
{
    if (!out)
        throw Error(MSJ_E_INVAL, "out is NULL");
    std::memset(out, 0, sizeof(*out));
    if (!(eps > 0.0) || !std::isfinite(eps))
        throw Error(MSJ_E_INVAL, "eps must be positive and finite");
    if (c->d == 0)
        throw Error(MSJ_E_INVAL, "no data loaded");
    const auto t_start = std::chrono::steady_clock::now();
    DeviceGuard g(c->prm.device);
    const uint32_t n = c->n, d = c->d;
    const bool dev_out = c->prm.output == MSJ_OUT_DEVICE;
    hipStream_t s = c->stream;
    msj_stats &st = out->stats;
    out->n = n;
    out->on_device = dev_out ? 1u : 0u;
    if (n == 0) {
        out->offsets = (uint64_t *)std::calloc(1, sizeof(uint64_t));
        out->on_device = 0;
        if (!out->offsets)
            throw std::bad_alloc();
        return;
    }

    // ---- index, join, ordered results (subsystems a-d)
    uint64_t cap = 0;
    {
        size_t freeb = 0, totalb = 0;
        MSJ_HIP(hipMemGetInfo(&freeb, &totalb));
        uint64_t want = std::max<uint64_t>(std::max<uint64_t>(c->last_R, 128ull * n), 1ull << 22);
        uint64_t budget = (uint64_t)(freeb / 4) / sizeof(uint64_t);
        cap = std::max<uint64_t>(std::min(want, budget), 1);
    }
    const uint64_t R =
        c->core.run(c->xs.p, c->staging.p, c->f32, n, d, eps, c->prm, cap, s, c->ev, st);
    c->last_R = R;
    c->r_counts.ensure(n);
    c->r_offsets.ensure((size_t)n + 1);
    c->r_nbrs.ensure(std::max<uint64_t>(R, 1));
    c->core.finish(n, R, c->r_counts.p, c->r_offsets.p, c->r_nbrs.p, s, c->ev, st);
    out->total_pairs = R;

    if (dev_out) {
        out->counts = c->r_counts.p;
        out->offsets = c->r_offsets.p;
        out->nbrs = c->r_nbrs.p;
        out->node_order = c->core.ix.perm.p;
    } else {
        MSJ_HIP(hipEventRecord(c->ev[6], s));
        out->counts = (uint32_t *)std::malloc(sizeof(uint32_t) * n);
        out->offsets = (uint64_t *)std::malloc(sizeof(uint64_t) * ((size_t)n + 1));
        out->nbrs = (uint32_t *)std::malloc(sizeof(uint32_t) * std::max<uint64_t>(R, 1));
        out->node_order = (uint32_t *)std::malloc(sizeof(uint32_t) * n);
        if (!out->counts || !out->offsets || !out->nbrs || !out->node_order) {
            msj_result_free(out);
            throw std::bad_alloc();
        }
        const auto t_copy = std::chrono::steady_clock::now();
        const size_t page = (size_t)(c->prm.page_pairs ? c->prm.page_pairs : (1ull << 24)) * 4;
        page_to_host(c, out->counts, c->r_counts.p, sizeof(uint32_t) * n, page, c->ev[6]);
        page_to_host(c, out->offsets, c->r_offsets.p, sizeof(uint64_t) * ((size_t)n + 1), page,
                     c->ev[6]);
        st.num_pages = (uint32_t)page_to_host(c, out->nbrs, c->r_nbrs.p, sizeof(uint32_t) * R,
                                              page, c->ev[6]);
        page_to_host(c, out->node_order, c->core.ix.perm.p, sizeof(uint32_t) * n, page, c->ev[6]);
        st.ms_d2h = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() -
                                                              t_copy)
                        .count();
    }
    st.ms_total =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start)
            .count();
}

}  // namespace

uint64_t msj::JoinCore::run(const void *xs, const void *rows, bool f32, uint32_t n, uint32_t d,
                            double eps,
                            const msj_params &prm, uint64_t cap, hipStream_t s, hipEvent_t *ev,
                            msj_stats &st)
{
    // ---- index (subsystems a, b)
    MSJ_HIP(hipEventRecord(ev[0], s));
    LayerStats ls;
    if (prm.index == MSJ_INDEX_BRUTE)
        build_single_node(n, ix, ls, s);
    else
        builder.build(xs, f32, n, d, eps, prm, ix, ls, s);
    join.order(ix, xs, f32, n, d, s);
    xr.ensure((size_t)n * padded_dim(d));
    gather_node_rows(xs, rows, xr.p, f32, join.jperm.p, n, d, s);
    const bool use_mfma = prm.form != MSJ_FORM_PLAIN && mfma_filter_applies(eps, n, d);
    join.plan(ix, n, d, s, !use_mfma);
    MSJ_HIP(hipEventRecord(ev[1], s));
    if (use_mfma)  // part of the join: filter rows, thresholds, work items
        mfma.prepare(ix, join, xr.p, n, d, eps, prm.form == MSJ_FORM_MFMA_SPLIT, s);

    // ---- join (subsystem c)
    if (cap == 0) {
        // under an arena: the result keys get half of what is left after the
        // filter's band buffers, the other half is the pair sort's
        const uint64_t left = g_arena ? (uint64_t)(g_arena->cap - g_arena->used) : 0;
        const uint64_t spare = std::min<uint64_t>(left / 4, 32ull << 20);
        cap = std::max<uint64_t>(1, (left - spare) / (2 * sizeof(uint64_t)));
    }
    float kms = 0.f, mms = 0.f;
    uint64_t launches = 0, dims = 0;
    uint64_t R;
    if (use_mfma) {
        R = mfma.run(ix, join, xr.p, n, d, eps, cap, prm.count_work != 0, s, ev[2], ev[7], ev[3],
                     &launches, &kms, &mms, &st);
    } else {
        R = join.run_plain(ix, xr.p, n, d, eps, cap, s, ev[2], ev[3], &launches, &kms,
                           prm.count_work ? &dims : nullptr, prm.verbose != 0);
        mms = kms;
        st.form_used = MSJ_FORM_PLAIN;
    }
    if (R >= 0xFFFFFFFFull * 16)
        throw Error(MSJ_E_OVERFLOW, "result pair count too large");
    MSJ_HIP(hipEventRecord(ev[4], s));
    st.ms_main_kernel = mms;
    st.tile_rows = join.tile_rows;
    st.dims_evaluated = dims;
    st.ms_join_kernel = kms;
    st.join_launches = launches;
    st.num_nodes = ix.M;
    st.num_layers = ls.layers;
    st.num_candidates = ls.ncand;
    st.num_pages = 0;
    st.num_calcs = ix.num_calcs;
    st.total_pairs = R;
    for (uint32_t l = 0; l < ls.layers; ++l) {
        st.chosen[l] = ls.chosen[l];
        st.nodes_per_layer[l] = ls.nodes[l];
        st.calcs_per_layer[l] = ls.calcs[l];
        st.sse[l] = ls.sse[l];
    }
    // node sizes / numCalcs for the reference's batch record, which the host
    // computes while the pair sort runs (batch_record)
    hstart.resize((size_t)ix.M + 1);
    hcalcs.resize(ix.M);
    MSJ_HIP(hipMemcpyAsync(hstart.data(), ix.nstart.p, hstart.size() * 4, hipMemcpyDeviceToHost,
                           s));
    MSJ_HIP(hipMemcpyAsync(hcalcs.data(), ix.ncalcs.p, hcalcs.size() * 8, hipMemcpyDeviceToHost,
                           s));
    MSJ_HIP(hipStreamSynchronize(s));
    rec_n = n;
    rec_eps = eps;
    rec_calcs = ix.num_calcs;
    rec_pending = true;
    return R;
}

void msj::JoinCore::batch_record(msj_stats &st)
{
    if (!rec_pending)
        return;
    reference_batch_plan(hstart, hcalcs, rec_n, rec_eps, rec_calcs, &st.ref_num_batches,
                         &st.ref_var_calc);
    rec_pending = false;
}

void msj::JoinCore::finish(uint32_t n, uint64_t R, uint32_t *counts, uint64_t *offsets,
                           uint32_t *nbrs, hipStream_t s, hipEvent_t *ev, msj_stats &st)
{
    // ---- ordered results (subsystem d)
    join.finish(n, R, counts, offsets, nbrs, s);
    MSJ_HIP(hipEventRecord(ev[5], s));
    batch_record(st);  // on the host, while the device sorts
    MSJ_HIP(hipStreamSynchronize(s));
    st.ms_index = elapsed(ev[0], ev[1]);
    st.ms_join = elapsed(ev[1], ev[4]);
    st.ms_output = elapsed(ev[4], ev[5]);
}

int msj::guarded_call(void (*fn)(void *), void *arg)
{
    return guarded([&] { fn(arg); });
}

extern "C" {

const char *msj_version(void) { return "msj 0.2.0 (gfx950, HIP)"; }

const char *msj_strerror(int status)
{
    switch (status) {
    case MSJ_OK: return "ok";
    case MSJ_E_INVAL: return "invalid argument";
    case MSJ_E_NOMEM: return "out of memory";
    case MSJ_E_HIP: return "HIP runtime error";
    case MSJ_E_OVERFLOW: return "count overflow";
    case MSJ_E_UNSUPPORTED: return "unsupported";
    case MSJ_E_NODEV: return "no HIP device";
    default: return "unknown status";
    }
}

const char *msj_last_error(void) { return g_last_error.c_str(); }

void msj_default_params(msj_params *p)
{
    if (!p)
        return;
    std::memset(p, 0, sizeof(*p));
    p->struct_size = sizeof(msj_params);
    p->device = CUDA_DEVICE;
    p->maxrp = MAXRP;
    p->rp_per_layer = RPPERLAYER;
    p->numidim = NUMIDIM;
    p->seed = 1;
    p->sample_per = SAMPLE_PER;
    p->min_calcs_per_node = MIN_CALCS_PER_NODE;
    p->form = MSJ_FORM_PLAIN;
    p->storage = MSJ_STORAGE_F64;
    p->output = MSJ_OUT_HOST;
    p->verbose = 0;
    p->count_work = 0;
    p->page_pairs = 0;
}

int msj_ctx_create(const msj_params *p, msj_ctx **out)
{
    return guarded([&] {
        if (!out)
            throw Error(MSJ_E_INVAL, "ctx is NULL");
        *out = nullptr;
        msj_params prm;
        if (p)
            prm = *p;
        else
            msj_default_params(&prm);
        msj::validate_params(prm);
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
            (void)hipGetLastError();
            throw Error(MSJ_E_NODEV, "no HIP device visible");
        }
        if (prm.device < 0 || prm.device >= ndev)
            throw Error(MSJ_E_NODEV, "device ordinal out of range");
        DeviceGuard g(prm.device);
        msj_ctx *c = new msj_ctx();
        c->prm = prm;
        try {
            MSJ_HIP(hipStreamCreate(&c->stream));
            MSJ_HIP(hipStreamCreateWithFlags(&c->copy_stream, hipStreamNonBlocking));
            for (auto &e : c->ev)
                MSJ_HIP(hipEventCreate(&e));
            for (auto &e : c->page_ev)
                MSJ_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
        } catch (...) {
            msj_ctx_destroy(c);
            throw;
        }
        *out = c;
    });
}

void msj_ctx_destroy(msj_ctx *c)
{
    if (!c)
        return;
    {
        int prev = -1;
        (void)hipGetDevice(&prev);
        (void)hipSetDevice(c->prm.device);
        if (c->stream)
            (void)hipStreamSynchronize(c->stream);
        if (c->copy_stream)
            (void)hipStreamSynchronize(c->copy_stream);
        for (auto &e : c->ev)
            if (e)
                (void)hipEventDestroy(e);
        for (auto &e : c->page_ev)
            if (e)
                (void)hipEventDestroy(e);
        if (c->copy_stream)
            (void)hipStreamDestroy(c->copy_stream);
        if (c->stream)
            (void)hipStreamDestroy(c->stream);
        delete c;  // releases device and pinned buffers on this device
        if (prev >= 0)
            (void)hipSetDevice(prev);
        return;
    }
}

int msj_batch_create(const msj_params *p, msj_batch **out)
{
    return guarded([&] {
        if (!out)
            throw Error(MSJ_E_INVAL, "out is NULL");
        *out = nullptr;
        msj_params prm;
        if (p)
            prm = *p;
        else
            msj_default_params(&prm);
        msj::validate_params(prm);
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
            (void)hipGetLastError();
            throw Error(MSJ_E_NODEV, "no HIP device visible");
        }
        if (prm.device < 0 || prm.device >= ndev)
            throw Error(MSJ_E_NODEV, "device ordinal out of range");
        DeviceGuard g(prm.device);
        std::unique_ptr<msj_batch> b(new msj_batch());
        b->prm = prm;
        b->e = msj::batch_create(prm);
        *out = b.release();
    });
}

void msj_batch_destroy(msj_batch *b)
{
    if (!b)
        return;
    int prev = -1;
    (void)hipGetDevice(&prev);
    (void)hipSetDevice(b->prm.device);
    msj::batch_destroy(b->e);
    delete b;
    if (prev >= 0)
        (void)hipSetDevice(prev);
}

int msj_batch_load(msj_batch *b, const msj_job *jobs, uint32_t njobs)
{
    return guarded([&] {
        if (!b || (!jobs && njobs))
            throw Error(MSJ_E_INVAL, "batch or jobs is NULL");
        DeviceGuard g(b->prm.device);
        b->njobs = 0;
        msj::batch_load(b->e, jobs, njobs);
        b->njobs = njobs;
    });
}

int msj_batch_run(msj_batch *b, msj_result *outs, msj_batch_stats *st)
{
    return guarded([&] {
        if (!b || (!outs && b->njobs))
            throw Error(MSJ_E_INVAL, "batch or outs is NULL");
        DeviceGuard g(b->prm.device);
        try {
            msj::batch_run(b->e, outs, st);
        } catch (...) {
            for (uint32_t j = 0; j < b->njobs; ++j)
                msj_result_free(&outs[j]);
            throw;
        }
    });
}

int msj_join_batch(const msj_job *jobs, uint32_t njobs, const msj_params *p, msj_result *outs)
{
    msj_params prm;
    if (p)
        prm = *p;
    else
        msj_default_params(&prm);
    prm.output = MSJ_OUT_HOST;
    msj_batch *b = nullptr;
    int rc = msj_batch_create(&prm, &b);
    if (rc == MSJ_OK)
        rc = msj_batch_load(b, jobs, njobs);
    if (rc == MSJ_OK)
        rc = msj_batch_run(b, outs, nullptr);
    msj_batch_destroy(b);
    return rc;
}

int msj_ctx_load(msj_ctx *c, const void *data, uint32_t n, uint32_t d)
{
    return guarded([&] {
        if (!c)
            throw Error(MSJ_E_INVAL, "ctx is NULL");
        do_load(c, data, n, d);
    });
}

int msj_ctx_set_params(msj_ctx *c, const msj_params *p)
{
    return guarded([&] {
        if (!c || !p)
            throw Error(MSJ_E_INVAL, "ctx or params is NULL");
        msj::validate_params(*p);
        if (p->device != c->prm.device || p->storage != c->prm.storage)
            throw Error(MSJ_E_INVAL, "device and storage are fixed when the context is created");
        c->prm = *p;
    });
}

int msj_ctx_load_unordered(msj_ctx *c, const double *rows, uint32_t n, uint32_t d,
                           uint32_t *order, double *sigma)
{
    return guarded([&] {
        if (!c)
            throw Error(MSJ_E_INVAL, "ctx is NULL");
        do_load_unordered(c, rows, n, d, order, sigma);
    });
}

int msj_ctx_join(msj_ctx *c, double eps, msj_result *out)
{
    return guarded([&] {
        if (!c)
            throw Error(MSJ_E_INVAL, "ctx is NULL");
        do_join(c, eps, out);
    });
}

int msj_self_join(const void *data, uint32_t n, uint32_t d, double eps, const msj_params *p,
                  msj_result *out)
{
    msj_params prm;
    if (p)
        prm = *p;
    else
        msj_default_params(&prm);
    prm.output = MSJ_OUT_HOST;  // the context dies with this call
    if (out)
        std::memset(out, 0, sizeof(*out));
    // argument errors come back before any device is touched
    int rc = guarded([&] {
        msj::validate_params(prm);
        if (!out)
            throw Error(MSJ_E_INVAL, "out is NULL");
        if (d == 0)
            throw Error(MSJ_E_INVAL, "d must be positive");
        if (n > 0 && !data)
            throw Error(MSJ_E_INVAL, "data is NULL");
        if ((uint64_t)n * d >= (1ull << 40))
            throw Error(MSJ_E_OVERFLOW, "n*d too large");
        if (!(eps > 0.0) || !std::isfinite(eps))
            throw Error(MSJ_E_INVAL, "eps must be positive and finite");
    });
    if (rc != MSJ_OK)
        return rc;
    msj_ctx *c = nullptr;
    rc = msj_ctx_create(&prm, &c);
    if (rc != MSJ_OK)
        return rc;
    rc = msj_ctx_load(c, data, n, d);
    if (rc == MSJ_OK)
        rc = msj_ctx_join(c, eps, out);
    msj_ctx_destroy(c);
    return rc;
}

void msj_result_free(msj_result *r)
{
    if (!r)
        return;
    if (!r->on_device) {
        std::free(r->counts);
        std::free(r->offsets);
        std::free(r->nbrs);
        std::free(r->node_order);
    }
    r->counts = nullptr;
    r->offsets = nullptr;
    r->nbrs = nullptr;
    r->node_order = nullptr;
    r->total_pairs = 0;
}

// utils.cu:22-53 (host; sequential sums as the reference computes them)
int msj_dim_order(const double *A, uint32_t n, uint32_t d, uint32_t *order, double *sigma)
{
    return guarded([&] {
        if (!A || !order || d == 0)
            throw Error(MSJ_E_INVAL, "bad arguments to msj_dim_order");
        std::vector<double> dev(d);
        auto column = [&](uint32_t i) {
            double mean = 0.0;
            for (uint32_t j = 0; j < n; ++j)
                mean = mean + A[(size_t)d * j + i];
            mean = mean / n;
            double devmean = 0.0;
            for (uint32_t j = 0; j < n; ++j) {
                double t = A[(size_t)d * j + i] - mean;
                double sq = t * t;
                devmean = devmean + sq;
            }
            devmean = devmean / n;
            dev[i] = std::sqrt(devmean);
        };
        // columns are independent; each column's sums stay sequential
        unsigned nt = std::max(1u, std::min<unsigned>(std::thread::hardware_concurrency(), d));
        std::vector<std::thread> pool;
        for (unsigned t = 0; t < nt; ++t)
            pool.emplace_back([&, t] {
                for (uint32_t i = t; i < d; i += nt)
                    column(i);
            });
        for (auto &th : pool)
            th.join();
        order_by_sigma(dev.data(), d, order, sigma);
    });
}

}  // extern "C"
