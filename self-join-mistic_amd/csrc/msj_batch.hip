// msj_batch.hip -- batching form: many independent joins on one GPU.
//
// A step (batch_run) is a pipeline over four streams (the GPU grants a
// process four hardware queues):
//   builders (two threads, a stream each): every job's node net, join
//     order, FP64 rows and work items, largest jobs first;
//   join stream: a job whose candidate evaluations fill the GPU on their own
//     (MSJ_SOLO_CALCS or more) is joined as soon as its build is done, by the
//     kernels of msj_self_join in the batch's form (the plain kernel, or the
//     MFMA filter and its FP64 band recheck), while the builders go on with the next
//     jobs (the build of job k+1 overlaps the join of job k, as the
//     reference's NUMSTREAMS batches overlap, params.cuh:63); the other jobs
//     wait for one persistent kernel per row-width class whose waves pop
//     work items from the class's part of a device queue (one atomic ticket
//     per item; the part is cut into one segment per XCD) holding all of
//     their items; the reset of the queue and the class kernels are captured
//     once as a hipGraph and replayed (the job table they read is refreshed
//     by a copy before each replay);
//   output (one thread, its own stream): as each job's join completes, its
//     pair sort and CSR arrays (and the copies to host).
// Per-job results are those of msj_self_join (the same per-item code, the
// same node nets).
//
// Autotuner.  The tile shape of the plain form is the work per item (query
// rows x candidate columns, set by the lane evaluations per item), the
// number of resident workgroups per CU and, for rows wider than 32, the
// dimensions per slab (the candidate coordinates a wave holds in registers
// and the query rows it stages in LDS at a time).  For each class of the
// persistent launch not tuned yet, its jobs are planned with 2^16, 2^18 and
// 2^20 evaluations per item and the class kernel is timed alone on a prefix
// of its queue; then, with the best of these, 1, 2 and 4 workgroups per CU;
// then, for the wide class, slabs of 8 and 16 dimensions.  The choices are
// kept by the batch for later runs.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <exception>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "msj_join_items.h"

#pragma clang fp contract(off)

namespace msj {

namespace {

// Persistent kernel of one row-width class (DS = ds <= 32, or 0 for the
// slab-major variant of ds > 32): its waves pop tickets from the class's
// queue; ticket t is item job_off[jb] + t of the table, i.e. an item of one
// of the class's jobs [jb, je).  At most `limit` tickets are served (the
// autotuner times prefixes of the queue).
constexpr uint32_t NSEG = 8;  // queue segments per class: one per XCD

// The XCD (accelerator complex die) this wave runs on; any value in
// [0, NSEG) is correct for the queue logic, the right one keeps a segment
// under one L2.
__device__ inline uint32_t xcd_id()
{
    // s_getreg of HW_REG_XCC_ID (id 20), bits [3:0]
    return (uint32_t)__builtin_amdgcn_s_getreg(20 | (0 << 6) | (3 << 11)) % NSEG;
}

template <int DS, int SL = HSL>
__device__ inline void batch_loop(const JobView *__restrict__ views,
                                  const uint64_t *__restrict__ job_off, uint32_t jb, uint32_t je,
                                  unsigned long long *__restrict__ queue, uint64_t limit)
{
    __shared__ __align__(16) double qsh[WPB][LDS_WAVE_BYTES / 8];
    __shared__ unsigned long long hsh[WPB][HITS];
    const uint32_t lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x / WAVE;
    WaveHits H(hsh[wv], lane);
    uint64_t dims = 0;
    const uint64_t base = job_off[jb];
    const uint64_t total = min(job_off[je] - base, limit);
    // The queue is split into NSEG contiguous segments, one per XCD: a wave
    // drains its own XCD's segment first, so consecutive items (the query
    // groups of one candidate chunk) run at the same time under one L2; then
    // it helps with the others.  Every wave leaves once all are drained.
    const uint32_t home = xcd_id();
    uint32_t k = 0;
    while (k < NSEG) {
        const uint32_t x = (home + k) % NSEG;
        const uint64_t slo = total * x / NSEG, shi = total * (x + 1) / NSEG;
        unsigned long long tv = 0;
        if (lane == 0)
            tv = atomicAdd(&queue[x], 1ull);
        // lane 0's ticket, in SGPRs: the job lookup and the job's view stay
        // scalar
        const uint64_t tk = ((uint64_t)__builtin_amdgcn_readfirstlane((uint32_t)(tv >> 32)) << 32) |
                            __builtin_amdgcn_readfirstlane((uint32_t)tv);
        if (tk >= shi - slo) {
            ++k;  // segment drained
            continue;
        }
        const uint64_t t = slo + tk;
        const uint64_t i = base + t;
        uint32_t lo = jb, hi = je;  // job with job_off[j] <= i < job_off[j+1]
        while (hi - lo > 1) {
            const uint32_t mid = (lo + hi) / 2;
            if (job_off[mid] <= i)
                lo = mid;
            else
                hi = mid;
        }
        const JobView J = views[lo];
        const JoinItem w = J.items[i - job_off[lo]];
        if (DS > 0)
            item_sym<(DS > 0 ? DS : 4), false>(J, w, qsh[wv], H, dims);
        else
            item_hd<false, SL>(J, w, qsh[wv], H, dims);
        if (H.fill)
            H.flush(J);  // the keys belong to this job
    }
}

template <int DS>
__global__ __launch_bounds__(256) void k_join_batch(const JobView *__restrict__ views,
                                                    const uint64_t *__restrict__ job_off,
                                                    uint32_t jb, uint32_t je,
                                                    unsigned long long *__restrict__ queue,
                                                    uint64_t limit)
{
    batch_loop<DS>(views, job_off, jb, je, queue, limit);
}

// ds > 32: four waves per SIMD, as the single-join kernel; SL dimensions
// per slab (tuned: 8 or 16).
template <int SL>
__global__ __launch_bounds__(256, 4) void k_join_batch_hd(const JobView *__restrict__ views,
                                                          const uint64_t *__restrict__ job_off,
                                                          uint32_t jb, uint32_t je,
                                                          unsigned long long *__restrict__ queue,
                                                          uint64_t limit)
{
    batch_loop<0, SL>(views, job_off, jb, je, queue, limit);
}

__global__ void k_batch_reset(const JobView *__restrict__ views, uint32_t njobs,
                              unsigned long long *__restrict__ queues, uint32_t nq)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < njobs)
        *views[j].counter = 0ull;
    if (j < nq)
        queues[j] = 0ull;
}

__global__ void k_batch_counts(const JobView *__restrict__ views, uint32_t njobs,
                               unsigned long long *__restrict__ dst)
{
    const uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j < njobs)
        dst[j] = *views[j].counter;
}

#define LAUNCH(kern, grid, block, ...)                                                   \
    do {                                                                                 \
        ::msj::check_grid((grid), (block));                                              \
        hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, s, __VA_ARGS__);            \
        MSJ_HIP(hipGetLastError());                                                      \
    } while (0)

struct BatchJob {
    uint32_t n = 0, d = 0;
    double eps = 0.0;
    DevBuf<unsigned char> xs;
    DevBuf<double> xr;
    Index ix;
    JoinRunner join;
    LayerStats ls;
    DevBuf<uint32_t> counts;
    DevBuf<uint64_t> offsets;
    DevBuf<uint32_t> nbrs;
    uint64_t R = 0, last_R = 0;
    float ms_index = 0.f;
    bool solo = false;            // joined alone by the single-join kernel
    bool mfma_form = false;       // solo, joined by the MFMA filter (params.form)
    MfmaRunner mfma;
    msj_stats mst = {};           // the filter's statistics of the last join
    hipEvent_t done = nullptr;    // its join is complete (recorded on the join stream)
    ~BatchJob()
    {
        if (done)
            (void)hipEventDestroy(done);
    }
};

// A job whose candidate evaluations reach this fills the GPU with its own
// kernel; smaller ones share the persistent launch.
constexpr uint64_t MSJ_SOLO_CALCS = 1ull << 33;

// A closable queue of job indices between the pipeline's threads.
struct JobQueue {
    std::mutex mu;
    std::condition_variable cv;
    std::deque<uint32_t> q;
    bool closed = false;
    void push(uint32_t j)
    {
        {
            std::lock_guard<std::mutex> lk(mu);
            q.push_back(j);
        }
        cv.notify_one();
    }
    void close()
    {
        {
            std::lock_guard<std::mutex> lk(mu);
            closed = true;
        }
        cv.notify_all();
    }
    // false once closed and drained
    bool pop(uint32_t &j)
    {
        std::unique_lock<std::mutex> lk(mu);
        cv.wait(lk, [&] { return !q.empty() || closed; });
        if (q.empty())
            return false;
        j = q.front();
        q.pop_front();
        return true;
    }
};

struct Worker {
    hipStream_t s = nullptr;
    IndexBuilder builder;
    DevBuf<unsigned char> staging;
};

double now_ms()
{
    return std::chrono::duration<double, std::milli>(
               std::chrono::steady_clock::now().time_since_epoch())
        .count();
}


// One class: the jobs of one padded row width, [jb, je) of the sorted table.
struct Cls {
    uint32_t ds, jb, je;
    uint64_t epi;   // tuned lane evaluations per item
    uint32_t wg;    // tuned workgroups per CU
    uint32_t sl;    // ds > 32: tuned dimensions per slab (8 or 16)
};

// The autotuner's choice for one row-width class.
struct Tuned {
    uint64_t epi;
    uint32_t wg, sl;
};

constexpr uint32_t MAX_CLASSES = 8;  // padded widths 4..32 step 4, and ds > 32

void launch_class(const Cls &c, uint32_t grid, const JobView *views, const uint64_t *job_off,
                  unsigned long long *queue, uint64_t limit, hipStream_t s)
{
    check_grid(grid, WAVE * WPB);
    const dim3 g(grid), blk(WAVE * WPB);
    switch (c.ds) {
#define MSJ_C(v)                                                                            \
    case v:                                                                                 \
        hipLaunchKernelGGL(k_join_batch<v>, g, blk, 0, s, views, job_off, c.jb, c.je, queue, \
                           limit);                                                          \
        break;
        MSJ_C(4)
        MSJ_C(8)
        MSJ_C(12)
        MSJ_C(16)
        MSJ_C(20)
        MSJ_C(24)
        MSJ_C(28)
        MSJ_C(32)
#undef MSJ_C
    default:
        if (c.sl == 8)
            hipLaunchKernelGGL(k_join_batch_hd<8>, g, blk, 0, s, views, job_off, c.jb, c.je,
                               queue, limit);
        else
            hipLaunchKernelGGL(k_join_batch_hd<HSL>, g, blk, 0, s, views, job_off, c.jb, c.je,
                               queue, limit);
    }
    MSJ_HIP(hipGetLastError());
}

}  // namespace

struct BatchEngine {
    msj_params prm;
    std::vector<std::unique_ptr<BatchJob>> jobs;
    std::vector<std::unique_ptr<Worker>> workers;  // builders, a stream each
    hipStream_t stream = nullptr;   // joins: solo kernels and the persistent launch
    hipStream_t ostream = nullptr;  // outputs
    hipEvent_t ev[2] = {}, oev[2] = {}, fev[3] = {};  // fev: a filter join's stages
    DevBuf<JobView> views;
    DevBuf<uint64_t> job_off;
    DevBuf<unsigned long long> queues, rcounts;
    std::vector<uint32_t> order;  // table position -> job (the pooled jobs by padded width)
    std::vector<Cls> cls;
    std::map<uint32_t, Tuned> tuned;  // ds (0: every ds > 32) -> choice
    int cus = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    std::vector<uint64_t> graph_key;

    // Runs f(job index, worker) for all jobs, job j on worker j % W.
    template <class F>
    void parallel(F &&f)
    {
        const uint32_t W = (uint32_t)workers.size();
        std::vector<std::thread> th;
        std::vector<std::exception_ptr> err(W);
        for (uint32_t w = 0; w < W; ++w)
            th.emplace_back([&, w] {
                try {
                    MSJ_HIP(hipSetDevice(prm.device));
                    for (uint32_t j = w; j < jobs.size(); j += W)
                        f(j, *workers[w]);
                    MSJ_HIP(hipStreamSynchronize(workers[w]->s));
                } catch (...) {
                    err[w] = std::current_exception();
                }
            });
        for (auto &t : th)
            t.join();
        for (auto &e : err)
            if (e)
                std::rethrow_exception(e);
    }

    // Classes and the job table of the jobs `sel` (sorted by padded width)
    // from their plans.
    void upload_table(const std::vector<uint32_t> &sel)
    {
        order = sel;
        std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
            return padded_dim(jobs[a]->d) < padded_dim(jobs[b]->d);
        });
        std::vector<Cls> nc;
        std::vector<JobView> hv;
        std::vector<uint64_t> off(1, 0);
        for (uint32_t k = 0; k < order.size(); ++k) {
            BatchJob &B = *jobs[order[k]];
            const uint32_t ds = padded_dim(B.d);
            const uint32_t key = ds <= 32 ? ds : 0u;  // every ds > 32 shares one kernel
            if (nc.empty() || (nc.back().ds <= 32 ? nc.back().ds : 0u) != key)
                nc.push_back(Cls{ds, k, k, 0, 0, (uint32_t)HSL});
            nc.back().je = k + 1;
            hv.push_back(B.join.view(B.ix, B.xr.p, B.n, B.d, B.eps));
            off.push_back(off.back() + B.join.nitems);
        }
        if (nc.size() > MAX_CLASSES)
            throw Error(MSJ_E_UNSUPPORTED, "too many row-width classes");
        for (auto &c : nc) {
            const auto it = tuned.find(c.ds <= 32 ? c.ds : 0u);
            c.epi = it == tuned.end() ? 0 : it->second.epi;
            c.wg = it == tuned.end() ? 2 : it->second.wg;
            c.sl = it == tuned.end() ? (uint32_t)HSL : it->second.sl;
        }
        cls = nc;
        views.ensure(std::max<size_t>(hv.size(), 1));
        job_off.ensure(off.size());
        if (!hv.empty())
            MSJ_HIP(hipMemcpyAsync(views.p, hv.data(), hv.size() * sizeof(JobView),
                                   hipMemcpyHostToDevice, stream));
        MSJ_HIP(hipMemcpyAsync(job_off.p, off.data(), off.size() * sizeof(uint64_t),
                               hipMemcpyHostToDevice, stream));
        MSJ_HIP(hipStreamSynchronize(stream));  // the host vectors go out of scope
    }

    uint32_t grid_for_wg(uint32_t wg) const { return (uint32_t)std::max(1, cus) * wg; }

    // Reset, then one persistent kernel per class, one after the other on
    // the join stream.
    void enqueue_all(uint64_t limit, int only)
    {
        hipStream_t s = stream;
        const uint32_t nj = (uint32_t)order.size();
        LAUNCH(k_batch_reset, grid_for(std::max<uint32_t>(nj, MAX_CLASSES * NSEG), 64), 64, views.p,
               nj, queues.p, MAX_CLASSES * NSEG);
        for (uint32_t c = 0; c < cls.size(); ++c) {
            if (only >= 0 && (int)c != only)
                continue;
            launch_class(cls[c], grid_for_wg(cls[c].wg), views.p, job_off.p,
                         queues.p + c * NSEG, limit, s);
        }
    }

    float timed(uint64_t limit, int only)
    {
        MSJ_HIP(hipEventRecord(ev[0], stream));
        enqueue_all(limit, only);
        MSJ_HIP(hipEventRecord(ev[1], stream));
        MSJ_HIP(hipEventSynchronize(ev[1]));
        float ms = 0.f;
        MSJ_HIP(hipEventElapsedTime(&ms, ev[0], ev[1]));
        return ms;
    }

    // The persistent launch through the captured graph (captured again when
    // the table's address, the jobs, the classes or their grids change).
    void launch_graph()
    {
        std::vector<uint64_t> key = {(uint64_t)(uintptr_t)views.p, (uint64_t)(uintptr_t)job_off.p};
        for (uint32_t j : order)
            key.push_back(j);
        for (auto &c : cls) {
            key.push_back(c.ds);
            key.push_back(c.jb);
            key.push_back(c.je);
            key.push_back(c.wg);
            key.push_back(c.sl);
        }
        if (!gexec || key != graph_key) {
            if (gexec)
                (void)hipGraphExecDestroy(gexec);
            if (graph)
                (void)hipGraphDestroy(graph);
            gexec = nullptr;
            graph = nullptr;
            MSJ_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal));
            enqueue_all(~0ull, -1);
            MSJ_HIP(hipStreamEndCapture(stream, &graph));
            MSJ_HIP(hipGraphInstantiate(&gexec, graph, nullptr, nullptr, 0));
            graph_key = key;
        }
        MSJ_HIP(hipGraphLaunch(gexec, stream));
    }

    // Plans the pooled jobs of width class key_ds with e evaluations per item.
    void replan(const std::vector<uint32_t> &pooled, uint32_t key_ds, uint64_t e)
    {
        for (uint32_t j : pooled) {
            BatchJob &B = *jobs[j];
            const uint32_t ds = padded_dim(B.d);
            if ((ds <= 32 ? ds : 0u) == key_ds)
                B.join.plan(B.ix, B.n, B.d, stream, true, e);
        }
    }

    // Autotuning of the classes of the persistent launch not tuned yet:
    // lane evaluations per item (2^16, 2^18, 2^20, with 2 workgroups per
    // CU), then workgroups per CU (1, 2, 4); each trial runs the class alone
    // over a prefix of its queue (an eighth, at least 4096 items), timed per
    // full queue.
    void tune(const std::vector<uint32_t> &pooled)
    {
        upload_table(pooled);
        for (uint32_t c = 0; c < cls.size(); ++c) {
            const uint32_t key = cls[c].ds <= 32 ? cls[c].ds : 0u;
            if (tuned.count(key))
                continue;
            auto trial = [&](uint32_t wg, uint32_t sl) {
                upload_table(pooled);
                cls[c].wg = wg;
                cls[c].sl = sl;
                uint64_t items = 0;
                for (uint32_t k = cls[c].jb; k < cls[c].je; ++k)
                    items += jobs[order[k]]->join.nitems;
                const uint64_t lim = std::min<uint64_t>(items, std::max<uint64_t>(4096, items / 8));
                const float ms = timed(lim, (int)c);
                return items ? ms * (double)items / (double)lim : 0.0;
            };
            double best = -1.0;
            uint64_t best_e = 1ull << 18;
            for (uint64_t e : {1ull << 16, 1ull << 18, 1ull << 20}) {
                replan(pooled, key, e);
                const double t = trial(2, HSL);
                if (best < 0 || t < best) {
                    best = t;
                    best_e = e;
                }
            }
            replan(pooled, key, best_e);
            best = -1.0;
            uint32_t best_wg = 2;
            for (uint32_t wg : {1u, 2u, 4u}) {
                const double t = trial(wg, HSL);
                if (best < 0 || t < best) {
                    best = t;
                    best_wg = wg;
                }
            }
            uint32_t best_sl = HSL;
            if (key == 0) {  // wide rows: dimensions per slab, 8 or 16
                best = -1.0;
                for (uint32_t sl : {8u, (uint32_t)HSL}) {
                    const double t = trial(best_wg, sl);
                    if (best < 0 || t < best) {
                        best = t;
                        best_sl = sl;
                    }
                }
            }
            tuned[key] = Tuned{best_e, best_wg, best_sl};
        }
    }

    ~BatchEngine()
    {
        (void)hipSetDevice(prm.device);
        if (stream)
            (void)hipStreamSynchronize(stream);
        if (ostream)
            (void)hipStreamSynchronize(ostream);
        jobs.clear();
        if (gexec)
            (void)hipGraphExecDestroy(gexec);
        if (graph)
            (void)hipGraphDestroy(graph);
        for (auto &w : workers)
            if (w->s) {
                (void)hipStreamSynchronize(w->s);
                (void)hipStreamDestroy(w->s);
            }
        for (auto &e : ev)
            if (e)
                (void)hipEventDestroy(e);
        for (auto &e : oev)
            if (e)
                (void)hipEventDestroy(e);
        for (auto &e : fev)
            if (e)
                (void)hipEventDestroy(e);
        if (ostream)
            (void)hipStreamDestroy(ostream);
        if (stream)
            (void)hipStreamDestroy(stream);
    }
};

BatchEngine *batch_create(const msj_params &p)
{
    std::unique_ptr<BatchEngine> b(new BatchEngine());
    b->prm = p;
    MSJ_HIP(hipStreamCreate(&b->stream));
    MSJ_HIP(hipStreamCreate(&b->ostream));
    for (auto &e : b->ev)
        MSJ_HIP(hipEventCreate(&e));
    for (auto &e : b->oev)
        MSJ_HIP(hipEventCreate(&e));
    for (auto &e : b->fev)
        MSJ_HIP(hipEventCreate(&e));
    int cus = 0;
    MSJ_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, p.device));
    b->cus = cus;
    const uint32_t W = 2;  // with the join and output streams: four streams
    for (uint32_t w = 0; w < W; ++w) {
        b->workers.emplace_back(new Worker());
        MSJ_HIP(hipStreamCreate(&b->workers.back()->s));
    }
    b->queues.ensure(MAX_CLASSES * NSEG);
    return b.release();
}

void batch_destroy(BatchEngine *b) { delete b; }

void batch_load(BatchEngine *b, const msj_job *jobs, uint32_t njobs)
{
    const bool f32 = b->prm.storage == MSJ_STORAGE_F32;
    for (uint32_t j = 0; j < njobs; ++j) {
        const msj_job &J = jobs[j];
        if (J.d == 0 || J.n == 0)
            throw Error(MSJ_E_INVAL, "every job needs n >= 1 and d >= 1");
        if (!J.data)
            throw Error(MSJ_E_INVAL, "job data is NULL");
        if (!(J.eps > 0.0) || !std::isfinite(J.eps))
            throw Error(MSJ_E_INVAL, "job eps must be positive and finite");
        if ((uint64_t)J.n * J.d >= (1ull << 32))
            throw Error(MSJ_E_OVERFLOW, "n*d too large");
    }
    b->jobs.clear();
    for (uint32_t j = 0; j < njobs; ++j) {
        b->jobs.emplace_back(new BatchJob());
        b->jobs.back()->n = jobs[j].n;
        b->jobs.back()->d = jobs[j].d;
        b->jobs.back()->eps = jobs[j].eps;
        MSJ_HIP(hipEventCreateWithFlags(&b->jobs.back()->done, hipEventDisableTiming));
    }
    b->parallel([&](uint32_t j, Worker &W) {
        BatchJob &B = *b->jobs[j];
        load_soa(jobs[j].data, B.n, B.d, f32, B.xs, W.staging, W.s);
    });
    for (auto &w : b->workers)
        w->staging.release();
    b->rcounts.ensure(std::max<uint32_t>(njobs, 1));
}

namespace {

// The output stage of one job whose join is complete: its pair count (a job
// whose pair buffer overflowed is joined again alone with an exact buffer),
// the pair sort and CSR arrays, and the result fields.
void finish_job(BatchEngine *b, uint32_t j, msj_result &o, const msj_batch_stats &S)
{
    BatchJob &B = *b->jobs[j];
    hipStream_t s = b->ostream;
    const bool dev_out = b->prm.output == MSJ_OUT_DEVICE;
    if (!B.mfma_form) {  // the filter's run already returned its exact count
        unsigned long long R = 0;
        MSJ_HIP(hipMemcpyAsync(&R, B.join.counter.p, sizeof(R), hipMemcpyDeviceToHost, s));
        MSJ_HIP(hipStreamSynchronize(s));
        B.R = R;
    }
    if (!B.mfma_form && B.R > B.join.pairs.cap) {
        B.join.pairs.ensure(B.R);
        uint64_t launches = 0;
        float kms = 0.f;
        B.R = B.join.run_plain(B.ix, B.xr.p, B.n, B.d, B.eps, B.R, s, b->oev[0], b->oev[1],
                               &launches, &kms, nullptr);
    }
    B.last_R = B.R;
    B.counts.ensure(B.n);
    B.offsets.ensure((size_t)B.n + 1);
    B.nbrs.ensure(std::max<uint64_t>(B.R, 1));
    B.join.finish(B.n, B.R, B.counts.p, B.offsets.p, B.nbrs.p, s);
    o.n = B.n;
    o.total_pairs = B.R;
    o.stats.num_nodes = B.ix.M;
    o.stats.num_layers = B.ls.layers;
    o.stats.num_candidates = B.ls.ncand;
    o.stats.num_calcs = B.ix.num_calcs;
    o.stats.total_pairs = B.R;
    o.stats.tile_rows = B.join.tile_rows;
    for (uint32_t l = 0; l < B.ls.layers; ++l) {
        o.stats.chosen[l] = B.ls.chosen[l];
        o.stats.nodes_per_layer[l] = B.ls.nodes[l];
        o.stats.calcs_per_layer[l] = B.ls.calcs[l];
        o.stats.sse[l] = B.ls.sse[l];
    }
    o.stats.ms_index = B.ms_index;
    o.stats.join_launches = 1;
    o.stats.form_used = MSJ_FORM_PLAIN;
    if (B.mfma_form) {
        o.stats.join_launches = 2;
        o.stats.form_used = B.mst.form_used;
        o.stats.band_pairs = B.mst.band_pairs;
        o.stats.max_score_err = B.mst.max_score_err;
        o.stats.mean_score_err = B.mst.mean_score_err;
        o.stats.filter_checked = B.mst.filter_checked;
        o.stats.filter_violations = B.mst.filter_violations;
        o.stats.filter_checked_bf16 = B.mst.filter_checked_bf16;
        o.stats.mfma_flops = B.mst.mfma_flops;
        o.stats.split_share_ppm = B.mst.split_share_ppm;
    }
    if (dev_out) {
        MSJ_HIP(hipStreamSynchronize(s));
        o.on_device = 1;
        o.counts = B.counts.p;
        o.offsets = B.offsets.p;
        o.nbrs = B.nbrs.p;
        o.node_order = B.ix.perm.p;
    } else {
        o.counts = (uint32_t *)std::malloc(sizeof(uint32_t) * B.n);
        o.offsets = (uint64_t *)std::malloc(sizeof(uint64_t) * ((size_t)B.n + 1));
        o.nbrs = (uint32_t *)std::malloc(sizeof(uint32_t) * std::max<uint64_t>(B.R, 1));
        o.node_order = (uint32_t *)std::malloc(sizeof(uint32_t) * B.n);
        if (!o.counts || !o.offsets || !o.nbrs || !o.node_order)
            throw std::bad_alloc();
        MSJ_HIP(hipMemcpyAsync(o.counts, B.counts.p, sizeof(uint32_t) * B.n,
                               hipMemcpyDeviceToHost, s));
        MSJ_HIP(hipMemcpyAsync(o.offsets, B.offsets.p, sizeof(uint64_t) * ((size_t)B.n + 1),
                               hipMemcpyDeviceToHost, s));
        if (B.R)
            MSJ_HIP(hipMemcpyAsync(o.nbrs, B.nbrs.p, sizeof(uint32_t) * B.R,
                                   hipMemcpyDeviceToHost, s));
        MSJ_HIP(hipMemcpyAsync(o.node_order, B.ix.perm.p, sizeof(uint32_t) * B.n,
                               hipMemcpyDeviceToHost, s));
        MSJ_HIP(hipStreamSynchronize(s));
    }
    (void)S;
}

}  // namespace

void batch_run(BatchEngine *b, msj_result *outs, msj_batch_stats *st)
{
    const double t0 = now_ms();
    msj_batch_stats S;
    std::memset(&S, 0, sizeof(S));
    const uint32_t nj = (uint32_t)b->jobs.size();
    const bool f32 = b->prm.storage == MSJ_STORAGE_F32;
    S.njobs = nj;
    S.workers = (uint32_t)b->workers.size();
    for (uint32_t j = 0; j < nj; ++j)
        std::memset(&outs[j], 0, sizeof(msj_result));

    // build order: largest first (their joins start first)
    std::vector<uint32_t> border(nj);
    for (uint32_t j = 0; j < nj; ++j)
        border[j] = j;
    std::stable_sort(border.begin(), border.end(),
                     [&](uint32_t a, uint32_t c) { return b->jobs[a]->n > b->jobs[c]->n; });

    JobQueue built, finished;
    std::atomic<uint32_t> next{0};
    std::mutex err_mu;
    std::exception_ptr err;
    auto fail = [&](std::exception_ptr e) {
        std::lock_guard<std::mutex> lk(err_mu);
        if (!err)
            err = e;
    };
    std::atomic<bool> failed{false};

    // builders: node nets, join orders, rows, items
    std::vector<std::thread> builders;
    std::atomic<uint32_t> builders_left{(uint32_t)b->workers.size()};
    for (uint32_t w = 0; w < b->workers.size(); ++w)
        builders.emplace_back([&, w] {
            Worker &W = *b->workers[w];
            try {
                MSJ_HIP(hipSetDevice(b->prm.device));
                for (uint32_t k; !failed && (k = next++) < nj;) {
                    const uint32_t j = border[k];
                    BatchJob &B = *b->jobs[j];
                    const double a = now_ms();
                    B.ls = LayerStats();
                    W.builder.build(B.xs.p, f32, B.n, B.d, B.eps, b->prm, B.ix, B.ls, W.s);
                    B.join.order(B.ix, B.xs.p, f32, B.n, B.d, W.s);
                    B.xr.ensure((size_t)B.n * padded_dim(B.d));
                    gather_node_rows(B.xs.p, nullptr, B.xr.p, f32, B.join.jperm.p, B.n, B.d,
                                     W.s);
                    B.solo = B.ix.num_calcs >= MSJ_SOLO_CALCS;
                    B.mfma_form = B.solo && b->prm.form != MSJ_FORM_PLAIN &&
                                  mfma_filter_applies(B.eps, B.n, B.d);
                    const uint32_t ds = padded_dim(B.d);
                    const auto it = b->tuned.find(ds <= 32 ? ds : 0u);
                    const uint64_t epi = B.solo || it == b->tuned.end() ? 0 : it->second.epi;
                    B.join.plan(B.ix, B.n, B.d, W.s, !B.mfma_form, epi);
                    if (B.mfma_form)  // rows, planes, thresholds and tiles of the filter
                        B.mfma.prepare(B.ix, B.join, B.xr.p, B.n, B.d, B.eps,
                                       b->prm.form == MSJ_FORM_MFMA_SPLIT, W.s);
                    B.join.counter.ensure(1);
                    const uint64_t cap = std::max<uint64_t>(B.last_R + B.last_R / 8, 64ull * B.n + 4096);
                    if (B.join.pairs.cap < cap)
                        B.join.pairs.ensure(cap);
                    MSJ_HIP(hipStreamSynchronize(W.s));
                    B.ms_index = (float)(now_ms() - a);
                    built.push(j);
                }
            } catch (...) {
                failed = true;
                fail(std::current_exception());
            }
            if (--builders_left == 0)
                built.close();
        });

    // output stage: every job as its join completes
    double out_ms = 0.0;
    std::thread outputs([&] {
        try {
            MSJ_HIP(hipSetDevice(b->prm.device));
            uint32_t j;
            while (finished.pop(j)) {
                if (failed)
                    continue;
                MSJ_HIP(hipEventSynchronize(b->jobs[j]->done));
                const double a = now_ms();
                finish_job(b, j, outs[j], S);
                out_ms += now_ms() - a;
            }
        } catch (...) {
            failed = true;
            fail(std::current_exception());
        }
    });

    // dispatch: solo jobs as their builds finish, then the persistent launch
    double t_built = t0, t_tune = 0.0;
    bool first = true;
    std::vector<uint32_t> pooled;
    try {
        hipStream_t s = b->stream;
        uint32_t j;
        while (built.pop(j)) {
            if (failed)
                continue;
            BatchJob &B = *b->jobs[j];
            if (!B.solo) {
                pooled.push_back(j);
                continue;
            }
            if (first) {
                MSJ_HIP(hipEventRecord(b->ev[0], s));
                first = false;
            }
            if (B.mfma_form) {
                // the filter waits for its own counters (and reruns on a
                // full result buffer) before the next job's dispatch
                uint64_t launches = 0;
                float kms = 0.f, mms = 0.f;
                B.mst = msj_stats{};
                B.R = B.mfma.run(B.ix, B.join, B.xr.p, B.n, B.d, B.eps, B.join.pairs.cap,
                                 b->prm.count_work != 0, s, b->fev[0], b->fev[1], b->fev[2],
                                 &launches, &kms, &mms, &B.mst);
            } else {
                B.join.launch_plain(B.ix, B.xr.p, B.n, B.d, B.eps, s);
            }
            MSJ_HIP(hipEventRecord(B.done, s));
            ++S.solo_jobs;
            finished.push(j);
        }
        t_built = now_ms();
        if (!failed && !pooled.empty()) {
            // autotuning of the pooled row widths not seen before
            bool need = false;
            for (uint32_t k : pooled) {
                const uint32_t ds = padded_dim(b->jobs[k]->d);
                need = need || !b->tuned.count(ds <= 32 ? ds : 0u);
            }
            if (need) {
                const double a = now_ms();
                MSJ_HIP(hipStreamSynchronize(s));  // the solo joins first: tuning times alone
                b->tune(pooled);
                t_tune = now_ms() - a;
            }
            b->upload_table(pooled);
            if (first) {
                MSJ_HIP(hipEventRecord(b->ev[0], s));
                first = false;
            }
            b->launch_graph();
            S.graph_replays = 1;
            for (uint32_t k : pooled) {
                MSJ_HIP(hipEventRecord(b->jobs[k]->done, s));
                finished.push(k);
            }
        }
        if (!first)
            MSJ_HIP(hipEventRecord(b->ev[1], s));
    } catch (...) {
        failed = true;
        fail(std::current_exception());
    }
    finished.close();
    for (auto &t : builders)
        t.join();
    outputs.join();
    if (err) {
        for (uint32_t k = 0; k < nj; ++k)
            if (!outs[k].on_device)
                msj_result_free(&outs[k]);
        std::rethrow_exception(err);
    }
    if (!first) {
        MSJ_HIP(hipEventSynchronize(b->ev[1]));
        float ms = 0.f;
        MSJ_HIP(hipEventElapsedTime(&ms, b->ev[0], b->ev[1]));
        S.ms_join = ms;
    }
    const double t_end = now_ms();
    S.ms_index = t_built - t0;
    S.ms_tune = t_tune;
    for (auto &c : b->cls)
        S.grid_blocks += b->grid_for_wg(c.wg);
    for (uint32_t j = 0; j < nj; ++j) {
        S.total_pairs += b->jobs[j]->R;
        S.num_calcs += b->jobs[j]->ix.num_calcs;
        S.work_items += b->jobs[j]->join.nitems;
    }
    uint32_t k = 0;
    for (auto &c : b->cls) {
        if (k == 8)
            break;
        S.class_ds[k] = c.ds <= 32 ? c.ds : 0u;
        S.class_epi_log2[k] = (uint32_t)bits_for(c.epi) - 1;
        S.class_slab[k] = c.ds > 32 ? c.sl : 0u;
        ++k;
    }
    S.classes = k;
    S.ms_total = t_end - t0;
    S.ms_output = out_ms;
    for (uint32_t j = 0; j < nj; ++j) {
        outs[j].stats.ms_total = S.ms_total;
        outs[j].stats.ms_join = S.ms_join;
        outs[j].stats.ms_join_kernel = S.ms_join;
    }
    if (st)
        *st = S;
}

}  // namespace msj
