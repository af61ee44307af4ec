// msj_join.hip -- plain FP64 distance-and-compare over the node net
// (subsystem c) and the ordered result compaction (subsystem d).
//
// Reference: kernel.cu:992-1046 nodeByPoint5 and kernel.cu:833-865
// cachedDistanceCheck (ILP = 4); launcher.cu:3418-3506 batch loop.
//
// Work decomposition.  A work item is one wave: up to 64 query points of one
// node X against a contiguous segment of X's candidate list (the points of
// X's neighbour nodes, concatenated in adjacency order).  When X has fewer
// than 64 points the wave's lanes form 64/|X| groups that stride the segment
// together, so small nodes still fill the wave.  Query coordinates of the
// first DR dimensions live in registers; candidate coordinates are read
// column-wise from the node-ordered SoA copy (lanes of a group read the same
// address).
//
// Decision: exactly the reference's sequence per 4-dimension chunk,
//   s = fma(d0, d0, s); s = s + d1*d1; s = s + d2*d2; s = s + d3*d3;
//   reject as soon as s > eps^2,
// with this file compiled without floating-point contraction.
//
// Emission: one atomic per wave per step (ballot + popcount prefix), not per
// pair.  Keys are (row << qbits | neighbour row); one radix sort then yields
// every row's neighbours ascending, the counts and the CSR offsets.
#include <algorithm>

#include "msj_internal.h"

#pragma clang fp contract(off)

namespace msj {

namespace {

constexpr uint32_t QB = 64;                  // queries per work item
constexpr uint64_t EVALS_PER_ITEM = 262144;  // pair evaluations per work item

// `dims` (counting builds only) accumulates the dimensions consumed before
// the decision: min(d, 4 * chunks evaluated).
template <class T, int DR, bool COUNT>
__device__ inline bool a3_accept(const double (&qc)[DR], const T *__restrict__ xn, uint32_t n,
                                 uint32_t d, uint32_t p, uint32_t c, double eps2, uint32_t &dims)
{
    double s = 0.0;
#pragma unroll
    for (int l = 0; l < DR; l += 4) {
        if (l < (int)d) {
            if (COUNT)
                dims += min(4u, d - (uint32_t)l);
            double t0 = qc[l] - (double)xn[(size_t)l * n + c];
            s = fma(t0, t0, s);
#pragma unroll
            for (int j = 1; j < 4; ++j) {
                if (l + j < (int)d) {
                    double t = qc[l + j] - (double)xn[(size_t)(l + j) * n + c];
                    double sq = t * t;
                    s = s + sq;
                }
            }
            if (s > eps2)
                return false;
        }
    }
    for (uint32_t l = DR; l < d; l += 4) {
        if (COUNT)
            dims += min(4u, d - l);
        double t0 = (double)xn[(size_t)l * n + p] - (double)xn[(size_t)l * n + c];
        s = fma(t0, t0, s);
        for (uint32_t j = 1; j < 4 && l + j < d; ++j) {
            double t = (double)xn[(size_t)(l + j) * n + p] - (double)xn[(size_t)(l + j) * n + c];
            double sq = t * t;
            s = s + sq;
        }
        if (s > eps2)
            return false;
    }
    return true;
}

template <class T, int DR, bool COUNT>
__global__ __launch_bounds__(256) void k_join_plain(
    const T *__restrict__ xn, uint32_t n, uint32_t d, double eps2,
    const uint32_t *__restrict__ nstart, const uint32_t *__restrict__ aoff,
    const uint32_t *__restrict__ adj, const uint32_t *__restrict__ perm,
    const JoinItem *__restrict__ items, uint32_t nitems, int qbits,
    unsigned long long *__restrict__ out, unsigned long long *__restrict__ counter,
    unsigned long long cap, unsigned long long *__restrict__ work)
{
    const uint32_t lane = threadIdx.x & (WAVE - 1);
    const uint32_t it = blockIdx.x * (blockDim.x / WAVE) + (threadIdx.x / WAVE);
    if (it >= nitems)
        return;  // wave-uniform
    const JoinItem w = items[it];
    const uint32_t x0 = nstart[w.x];
    const uint32_t xsz = nstart[w.x + 1] - x0;
    const uint32_t qn = min(QB, xsz - w.qbeg);
    const uint32_t G = QB / qn;
    const uint32_t qi = lane % qn, g = lane / qn;
    const bool lane_on = g < G;
    const uint32_t p = x0 + w.qbeg + (lane_on ? qi : 0u);

    double qc[DR];
#pragma unroll
    for (int j = 0; j < DR; ++j)
        qc[j] = (j < (int)d) ? (double)xn[(size_t)j * n + p] : 0.0;
    const unsigned long long pkey = (unsigned long long)perm[p] << qbits;
    const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (WAVE - lane));

    uint32_t a = aoff[w.x];
    uint32_t ybase = 0;
    uint32_t ystart = nstart[adj[a]];
    uint32_t ysz = nstart[adj[a] + 1] - ystart;
    const uint32_t iters = (w.cend - w.cbeg + G - 1) / G;
    uint32_t dims = 0;
    for (uint32_t itr = 0; itr < iters; ++itr) {
        const uint32_t c = w.cbeg + g + itr * G;
        const bool valid = lane_on && c < w.cend;
        bool acc = false;
        uint32_t pos = 0;
        if (valid) {
            while (c >= ybase + ysz) {  // advance to the neighbour node holding c
                ybase += ysz;
                ++a;
                const uint32_t Y = adj[a];
                ystart = nstart[Y];
                ysz = nstart[Y + 1] - ystart;
            }
            pos = ystart + (c - ybase);
            acc = a3_accept<T, DR, COUNT>(qc, xn, n, d, p, pos, eps2, dims);
        }
        const uint64_t m = __ballot(acc);
        if (m) {
            const uint32_t leader = (uint32_t)__ffsll((unsigned long long)m) - 1u;
            unsigned long long base = 0;
            if (lane == leader)
                base = atomicAdd(counter, (unsigned long long)__popcll(m));
            base = __shfl(base, (int)leader, WAVE);
            if (acc) {
                const unsigned long long idx = base + (unsigned long long)__popcll(m & lt);
                if (idx < cap)
                    out[idx] = pkey | (unsigned long long)perm[pos];
            }
        }
    }
    if (COUNT) {
        unsigned long long v = dims;
#pragma unroll
        for (int o = WAVE / 2; o > 0; o >>= 1)
            v += __shfl_down(v, o, WAVE);
        if (lane == 0)
            atomicAdd(work, v);
    }
}

// Work items per node: ceil(|X|/64) query chunks x segments of the
// candidate list sized so one item is about EVALS_PER_ITEM evaluations.
__global__ __launch_bounds__(256) void k_item_count(const uint32_t *__restrict__ nstart,
                                                    const uint64_t *__restrict__ sumsz, uint32_t M,
                                                    uint32_t *__restrict__ cnt,
                                                    uint32_t *__restrict__ seglen)
{
    uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= M)
        return;
    uint32_t sz = nstart[x + 1] - nstart[x];
    uint64_t L = sumsz[x];
    uint32_t qn = min(QB, sz);
    uint64_t sl = EVALS_PER_ITEM / qn;
    uint64_t nseg = (L + sl - 1) / sl;
    if (nseg == 0)
        nseg = 1;
    uint64_t nq = (sz + QB - 1) / QB;
    cnt[x] = (uint32_t)(nq * nseg);
    seglen[x] = (uint32_t)sl;
}

__global__ __launch_bounds__(256) void k_item_fill(const uint32_t *__restrict__ nstart,
                                                   const uint64_t *__restrict__ sumsz,
                                                   const uint32_t *__restrict__ seglen,
                                                   const uint32_t *__restrict__ itoff, uint32_t M,
                                                   uint32_t nitems, JoinItem *__restrict__ items)
{
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nitems)
        return;
    // node x with itoff[x] <= k < itoff[x+1]
    uint32_t lo = 0, hi = M;
    while (hi - lo > 1) {
        uint32_t mid = lo + (hi - lo) / 2;
        if (itoff[mid] <= k)
            lo = mid;
        else
            hi = mid;
    }
    uint32_t x = lo;
    uint32_t local = k - itoff[x];
    uint64_t L = sumsz[x];
    uint32_t sl = seglen[x];
    uint32_t nseg = (uint32_t)((L + sl - 1) / sl);
    if (nseg == 0)
        nseg = 1;
    uint32_t q = local / nseg, sg = local % nseg;
    JoinItem w;
    w.x = x;
    w.qbeg = q * QB;
    w.cbeg = (uint32_t)std::min<uint64_t>((uint64_t)sg * sl, L);
    w.cend = (uint32_t)std::min<uint64_t>((uint64_t)(sg + 1) * sl, L);
    items[k] = w;
}

template <class T>
__global__ __launch_bounds__(256) void k_gather(const T *__restrict__ xs, T *__restrict__ xn,
                                                const uint32_t *__restrict__ perm, uint32_t n,
                                                uint32_t d)
{
    size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)n * d)
        return;
    uint32_t j = (uint32_t)(idx / n), i = (uint32_t)(idx - (size_t)j * n);
    xn[idx] = xs[(size_t)j * n + perm[i]];
}

__global__ __launch_bounds__(256) void k_rows_count(const uint64_t *__restrict__ pairs, size_t R,
                                                    int qbits, uint32_t *__restrict__ counts)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= R)
        return;
    uint32_t row = (uint32_t)(pairs[i] >> qbits);
    // sorted keys: count each row once at the end of its run
    if (i + 1 == R || (uint32_t)(pairs[i + 1] >> qbits) != row) {
        size_t lo = 0, hi = i + 1;  // first index of this row
        while (lo < hi) {
            size_t mid = lo + (hi - lo) / 2;
            if ((uint32_t)(pairs[mid] >> qbits) < row)
                lo = mid + 1;
            else
                hi = mid;
        }
        counts[row] = (uint32_t)(i + 1 - lo);
    }
}

__global__ __launch_bounds__(256) void k_rows_nbrs(const uint64_t *__restrict__ pairs, size_t R,
                                                   uint64_t qmask, uint32_t *__restrict__ nbrs)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < R)
        nbrs[i] = (uint32_t)(pairs[i] & qmask);
}

template <class T>
__global__ void k_zero(T *p, size_t n)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = 0;
}

#define LAUNCH(kern, grid, block, ...)                                                   \
    do {                                                                                 \
        hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, s, __VA_ARGS__);            \
        MSJ_HIP(hipGetLastError());                                                      \
    } while (0)

template <class T, int DR>
void launch_plain(const Index &ix, const void *xn, uint32_t n, uint32_t d, double eps2,
                  const JoinItem *items, uint32_t nitems, int qbits, unsigned long long *out,
                  unsigned long long *counter, unsigned long long cap, unsigned long long *work,
                  hipStream_t s)
{
    const unsigned waves_per_block = 4;
    if (work)
        LAUNCH((k_join_plain<T, DR, true>), grid_for(nitems, waves_per_block),
               WAVE * waves_per_block, (const T *)xn, n, d, eps2, ix.nstart.p, ix.aoff.p, ix.adj.p,
               ix.perm.p, items, nitems, qbits, out, counter, cap, work);
    else
        LAUNCH((k_join_plain<T, DR, false>), grid_for(nitems, waves_per_block),
               WAVE * waves_per_block, (const T *)xn, n, d, eps2, ix.nstart.p, ix.aoff.p, ix.adj.p,
               ix.perm.p, items, nitems, qbits, out, counter, cap, work);
}

template <class T>
void dispatch_plain(const Index &ix, const void *xn, uint32_t n, uint32_t d, double eps2,
                    const JoinItem *items, uint32_t nitems, int qbits, unsigned long long *out,
                    unsigned long long *counter, unsigned long long cap, unsigned long long *work,
                    hipStream_t s)
{
#define MSJ_DR(v)                                                                        \
    if (d <= v) {                                                                        \
        launch_plain<T, v>(ix, xn, n, d, eps2, items, nitems, qbits, out, counter, cap, work, \
                           s);                                                           \
        return;                                                                          \
    }
    MSJ_DR(4)
    MSJ_DR(8)
    MSJ_DR(12)
    MSJ_DR(16)
    MSJ_DR(20)
    MSJ_DR(24)
    MSJ_DR(32)
    MSJ_DR(48)
#undef MSJ_DR
    launch_plain<T, 64>(ix, xn, n, d, eps2, items, nitems, qbits, out, counter, cap, work, s);
}

}  // namespace

void gather_node_order(const void *xs, void *xn, bool f32, const uint32_t *perm, uint32_t n,
                       uint32_t d, hipStream_t s)
{
    size_t m = (size_t)n * d;
    if (f32)
        LAUNCH(k_gather<float>, grid_for(m, 256), 256, (const float *)xs, (float *)xn, perm, n, d);
    else
        LAUNCH(k_gather<double>, grid_for(m, 256), 256, (const double *)xs, (double *)xn, perm, n,
               d);
}

void JoinRunner::plan(const Index &ix, uint32_t n, hipStream_t s)
{
    const uint32_t M = ix.M;
    itcnt.ensure(M);
    seglen.ensure(M);
    itoff.ensure((size_t)M + 1);
    LAUNCH(k_item_count, grid_for(M, 256), 256, ix.nstart.p, ix.sumsz.p, M, itcnt.p, seglen.p);
    DevBuf<uint64_t> tot;
    tot.ensure((size_t)M + 1);
    scan.excl_u32_to_u64(itcnt.p, tot.p, M, s);
    uint64_t total = 0;
    MSJ_HIP(hipMemcpyAsync(&total, tot.p + M, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    MSJ_HIP(hipStreamSynchronize(s));
    if (total >= 0xFFFFFFFFull)
        throw Error(MSJ_E_OVERFLOW, "join work items exceed 2^32");
    nitems = (uint32_t)total;
    scan.excl_u32_to_u32(itcnt.p, itoff.p, M, s);
    items.ensure(std::max<uint32_t>(nitems, 1));
    LAUNCH(k_item_fill, grid_for(nitems, 256), 256, ix.nstart.p, ix.sumsz.p, seglen.p, itoff.p, M,
           nitems, items.p);
}

uint64_t JoinRunner::run_plain(const Index &ix, const void *xn, bool f32, uint32_t n, uint32_t d,
                               double eps, uint64_t cap_hint, hipStream_t s, hipEvent_t k0,
                               hipEvent_t k1, uint64_t *launches, float *kernel_ms,
                               uint64_t *dims_evaluated)
{
    const double eps2 = eps * eps;  // launcher.cu:3200
    const int qbits = std::max(1, bits_for(n ? n - 1 : 0));
    counter.ensure(1);
    if (pairs.cap < cap_hint)
        pairs.ensure(cap_hint);
    *kernel_ms = 0.f;
    unsigned long long *work = nullptr;
    if (dims_evaluated) {
        work_ctr.ensure(1);
        work = work_ctr.p;
    }
    for (int attempt = 0; attempt < 2; ++attempt) {
        LAUNCH(k_zero<unsigned long long>, 1, 64, counter.p, (size_t)1);
        if (work)
            LAUNCH(k_zero<unsigned long long>, 1, 64, work, (size_t)1);
        MSJ_HIP(hipEventRecord(k0, s));
        if (nitems) {
            if (f32)
                dispatch_plain<float>(ix, xn, n, d, eps2, items.p, nitems, qbits,
                                      (unsigned long long *)pairs.p, counter.p, pairs.cap, work, s);
            else
                dispatch_plain<double>(ix, xn, n, d, eps2, items.p, nitems, qbits,
                                       (unsigned long long *)pairs.p, counter.p, pairs.cap, work,
                                       s);
            ++*launches;
        }
        MSJ_HIP(hipEventRecord(k1, s));
        unsigned long long R = 0;
        MSJ_HIP(hipMemcpyAsync(&R, counter.p, sizeof(R), hipMemcpyDeviceToHost, s));
        MSJ_HIP(hipStreamSynchronize(s));
        float ms = 0.f;
        MSJ_HIP(hipEventElapsedTime(&ms, k0, k1));
        *kernel_ms += ms;
        if (R <= pairs.cap) {
            if (work)
                MSJ_HIP(hipMemcpy(dims_evaluated, work, sizeof(uint64_t), hipMemcpyDeviceToHost));
            return R;
        }
        pairs.ensure(R);  // exact size; the rerun cannot overflow
    }
    throw Error(MSJ_E_OVERFLOW, "result count changed between identical join passes");
}

void JoinRunner::finish(uint32_t n, uint64_t R, uint32_t *counts, uint64_t *offsets, uint32_t *nbrs,
                        hipStream_t s)
{
    const int qbits = std::max(1, bits_for(n ? n - 1 : 0));
    const int rbits = bits_for(n ? n - 1 : 0);
    sorter.sort(pairs.p, nullptr, R, qbits + rbits, s);
    LAUNCH(k_zero<uint32_t>, grid_for(n, 256), 256, counts, (size_t)n);
    if (R) {
        LAUNCH(k_rows_count, grid_for(R, 256), 256, pairs.p, (size_t)R, qbits, counts);
        LAUNCH(k_rows_nbrs, grid_for(R, 256), 256, pairs.p, (size_t)R, (1ull << qbits) - 1,
               nbrs);
    }
    scan.excl_u32_to_u64(counts, offsets, n, s);
}

}  // namespace msj
