// msj_index.hip -- node-net construction on the GPU (subsystems a and b).
//
// Reference: nodes.cu:5-356 buildNodeNet, 358-449 initNodes, 452-633
// splitNodes, 677-695 updateNodeCalcs, 719-747 updateNeighbors;
// utils.cu:22-137 stddev/createRPArray; kernel.cu:880-909 binningKernel.
//
// Per layer, everything except the 64*d glibc rand() draws and the 64-entry
// sigma sort runs on the device:
//   1. sample distances to 64 random candidates and their population sigma
//      (the host FP64 operation order: no fused multiply-add), max corner;
//   2. one kernel computes all 32 reference-point keys and min(6,d) axis keys
//      per point (FP64, fma accumulation as nvcc contracts binningKernel);
//   3. the 38 candidate partitions are evaluated in ONE radix sort of
//      (candidate, parent node, key) composites; run lengths in sorted order
//      are the would-be node sizes in the reference's node order, and the
//      sum of squared deviations is accumulated sequentially per candidate,
//      in that order, so it is bit-identical to nodes.cu:194-199;
//   4. the winner is applied with a stable (parent, key) sort carrying the
//      point ids (points stay ascending within a node, as thrust's stable
//      sorts leave them);
//   5. adjacency of the new layer comes from the parent layer's adjacency:
//      for each child and each parent neighbour Q, the children of Q whose
//      key is within +-1 (binary search in Q's sorted child keys), or all of
//      them when either side stopped splitting.  Count pass, prefix scan,
//      write pass; the neighbour lists come out in the reference's order.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "msj_internal.h"

#pragma clang fp contract(off)

namespace msj {

// ------------------------------------------------------- glibc TYPE_3 rand
void GlibcRand::seed(uint32_t s)
{
    if (s == 0)
        s = 1;
    int32_t word = (int32_t)s;
    r[0] = word;
    for (int i = 1; i < 31; i++) {
        long hi = word / 127773;
        long lo = word % 127773;
        long w = 16807 * lo - 2836 * hi;
        if (w < 0)
            w += 2147483647;
        word = (int32_t)w;
        r[i] = word;
    }
    f = 3;
    b = 0;
    for (int i = 0; i < 310; i++)
        (void)next();
}

int32_t GlibcRand::next()
{
    uint32_t val = (uint32_t)r[f] + (uint32_t)r[b];
    r[f] = (int32_t)val;
    int32_t res = (int32_t)(val >> 1);
    if (++f >= 31) {
        f = 0;
        ++b;
    } else if (++b >= 31) {
        b = 0;
    }
    return res;
}

namespace {

// x86-64 host conversions used by the reference's host code.
inline uint32_t host_cvt_u32(double q)  // (unsigned)q: cvttsd2si 64-bit, low half
{
    if (!(q > -9223372036854775808.0 && q < 9223372036854775808.0))
        return 0u;
    return (uint32_t)(uint64_t)(int64_t)q;
}
inline int32_t host_cvt_i32(double q)  // (int)q: cvttsd2si 32-bit
{
    if (!(q > -2147483649.0 && q < 2147483648.0))
        return (int32_t)0x80000000u;
    return (int32_t)q;
}
inline uint64_t radix_key_f64(double x)
{
    uint64_t u;
    std::memcpy(&u, &x, 8);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

constexpr int KEY_RP = 32;  // register accumulators of the key kernel

__global__ __launch_bounds__(256) void k_iota(uint32_t *p, uint32_t n)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = i;
}

template <class T>
__global__ void k_fill(T *p, size_t n, T v)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = v;
}

// utils.cu:84-89: distmat[i*test_rp + j] = sqrt(sum (rp - x)^2), host order.
// The candidates come transposed (trpT[k * ntest + j]), so the lanes of a
// wave, consecutive candidates j of one sample row i, read consecutive words.
template <class T>
__global__ __launch_bounds__(256) void k_rp_distmat(const T *__restrict__ xs, uint32_t n,
                                                    uint32_t d, uint32_t S,
                                                    const double *__restrict__ trpT,
                                                    uint32_t ntest, double *__restrict__ distmat)
{
    size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)S * ntest)
        return;
    uint32_t i = (uint32_t)(idx / ntest), j = (uint32_t)(idx % ntest);
    double dist = 0;
    for (uint32_t k = 0; k < d; ++k) {
        double diff = trpT[(size_t)k * ntest + j] - (double)xs[(size_t)k * n + i];
        double sq = diff * diff;
        dist = dist + sq;
    }
    distmat[idx] = sqrt(dist);
}

// utils.cu:29-41 per column: sequential mean, then sequential sum of squares.
// One block per column; element (r, c) is A[r * rs + c * cs].  The additions
// stay in row order, made by one thread; the block's other threads stream
// the column through a double-buffered LDS tile, so the next tile's global
// loads are in flight while the current one is summed.
constexpr uint32_t SIG_TILE = 2048;  // rows per LDS tile
constexpr uint32_t SIG_BLOCK = 256;

__global__ __launch_bounds__(SIG_BLOCK) void k_col_sigma(const double *__restrict__ A,
                                                          uint32_t rows, size_t rs, size_t cs,
                                                          double *__restrict__ sigma)
{
    constexpr uint32_t PER = SIG_TILE / SIG_BLOCK;
    __shared__ double tile[2][SIG_TILE];
    __shared__ double mean_sh;
    const uint32_t c = blockIdx.x, tid = threadIdx.x;
    const uint32_t ntiles = (rows + SIG_TILE - 1) / SIG_TILE;
    double mean = 0.0;
    for (int pass = 0; pass < 2; ++pass) {
        double acc = 0.0;  // thread 0
        double v[PER];
        auto fetch = [&](uint32_t t) {
#pragma unroll
            for (uint32_t k = 0; k < PER; ++k) {
                const uint32_t r = t * SIG_TILE + k * SIG_BLOCK + tid;
                v[k] = r < rows ? A[(size_t)r * rs + (size_t)c * cs] : 0.0;
            }
        };
        auto put = [&](uint32_t t) {
#pragma unroll
            for (uint32_t k = 0; k < PER; ++k)
                tile[t & 1][k * SIG_BLOCK + tid] = v[k];
        };
        if (ntiles) {
            fetch(0);
            put(0);
        }
        __syncthreads();
        for (uint32_t t = 0; t < ntiles; ++t) {
            if (t + 1 < ntiles)
                fetch(t + 1);
            if (tid == 0) {
                const double *T = tile[t & 1];
                const uint32_t m = min(SIG_TILE, rows - t * SIG_TILE);
                const uint32_t m16 = m & ~15u;
                // the next 16 values are read while the current 16 are added
                double w[16];
#pragma unroll
                for (uint32_t k = 0; k < 16; ++k)
                    w[k] = m16 ? T[k] : 0.0;
                for (uint32_t r = 0; r < m16; r += 16) {
                    const uint32_t rn = r + 16 < m16 ? r + 16 : r;
                    double nx[16];
#pragma unroll
                    for (uint32_t k = 0; k < 16; ++k)
                        nx[k] = T[rn + k];
#pragma unroll
                    for (uint32_t k = 0; k < 16; ++k) {
                        if (pass == 0) {
                            acc = acc + w[k];
                        } else {
                            double d = w[k] - mean;
                            double sq = d * d;
                            acc = acc + sq;
                        }
                    }
#pragma unroll
                    for (uint32_t k = 0; k < 16; ++k)
                        w[k] = nx[k];
                }
                for (uint32_t r = m16; r < m; ++r) {
                    if (pass == 0) {
                        acc = acc + T[r];
                    } else {
                        double d = T[r] - mean;
                        double sq = d * d;
                        acc = acc + sq;
                    }
                }
            }
            if (t + 1 < ntiles)
                put(t + 1);
            __syncthreads();
        }
        if (tid == 0) {
            if (pass == 0)
                mean_sh = acc / (double)rows;
            else
                sigma[c] = sqrt(acc / (double)rows);
        }
        __syncthreads();
        mean = mean_sh;
    }
}

// utils.cu:105-118: max corner over the sample rows, starting from 0.  A
// maximum does not depend on the order it is taken in: one block per column.
template <class T>
__global__ __launch_bounds__(256) void k_col_max(const T *__restrict__ xs, uint32_t n, uint32_t d,
                                                 uint32_t S, double *__restrict__ maxc)
{
    __shared__ double part[256 / WAVE];
    const uint32_t j = blockIdx.x;
    double m = 0;
    for (uint32_t i = threadIdx.x; i < S; i += blockDim.x) {
        double v = (double)xs[(size_t)j * n + i];
        if (m < v)
            m = v;
    }
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) {
        double v = __shfl_down(m, o, WAVE);
        if (m < v)
            m = v;
    }
    if ((threadIdx.x & (WAVE - 1)) == 0)
        part[threadIdx.x / WAVE] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 256 / WAVE; ++w)
            if (m < part[w])
                m = part[w];
        maxc[j] = m;
    }
}

// kernel.cu:880-909 (RP keys) + nodes.cu:128-135 (axis keys).  Thread
// (i, blockIdx.y) keeps the distances of point i to reference points
// [KEY_RPT * y, KEY_RPT * y + KEY_RPT) in registers and reads the point's
// column once; RP coordinates are wave-uniform (scalar loads of rpT[j][..]).
// Half of the 32 accumulators per thread keeps enough waves resident to
// hide the loads.  The y = 0 threads also write the axis keys.  With kmax
// given (zeroed before), the largest key is folded into it: one atomic per
// workgroup at most, none when the workgroup's maximum does not exceed what
// kmax already holds.
// COSS: the tree's axis keys instead (tree.cu:284).
constexpr int KEY_RPT = 16;  // reference points per thread; grid.y = KEY_RP / KEY_RPT

template <class T, bool COSS = false>
__global__ __launch_bounds__(256) void k_keys(const T *__restrict__ xs, uint32_t n, uint32_t d,
                                              const double *__restrict__ rpT, uint32_t nrp,
                                              uint32_t naxis, double eps,
                                              uint32_t *__restrict__ keys,
                                              uint32_t *__restrict__ kmax)
{
    __shared__ uint32_t part[256 / WAVE];
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t t0 = blockIdx.y * KEY_RPT;
    uint32_t m = 0;
    if (i < n && t0 < nrp) {
        double acc[KEY_RPT];
#pragma unroll
        for (int t = 0; t < KEY_RPT; ++t)
            acc[t] = 0.0;
        double xn = (double)xs[i];  // d >= 1; column j + 1 is read during column j
        for (uint32_t j = 0; j < d; ++j) {
            const double x = xn;
            if (j + 1 < d)
                xn = (double)xs[(size_t)(j + 1) * n + i];
            const double *r = rpT + (size_t)j * KEY_RP + t0;
#pragma unroll
            for (int t = 0; t < KEY_RPT; ++t) {
                double df = x - r[t];
                acc[t] = fma(df, df, acc[t]);
            }
        }
#pragma unroll
        for (int t = 0; t < KEY_RPT; ++t)
            if (t0 + t < nrp) {
                const uint32_t k = rp_key(acc[t], eps);
                keys[(size_t)(t0 + t) * n + i] = k;
                m = max(m, k);
            }
    }
    if (i < n && blockIdx.y == 0)
        for (uint32_t a = 0; a < naxis; ++a) {
            const uint32_t k = COSS ? coss_axis_key((double)xs[(size_t)a * n + i], eps)
                                    : axis_key((double)xs[(size_t)a * n + i], eps);
            keys[(size_t)(nrp + a) * n + i] = k;
            m = max(m, k);
        }
    if (!kmax)
        return;
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1)
        m = max(m, (uint32_t)__shfl_down(m, o, WAVE));
    if ((threadIdx.x & (WAVE - 1)) == 0)
        part[threadIdx.x / WAVE] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        for (uint32_t w = 1; w < blockDim.x / WAVE; ++w)
            m = max(m, part[w]);
        if (m > __hip_atomic_load(kmax, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
            atomicMax(kmax, m);
    }
}

// Launch of k_keys over n points and every reference-point group.
template <class T, bool COSS>
void launch_keys(const T *xs, uint32_t n, uint32_t d, const double *rpT, uint32_t nrp,
                 uint32_t naxis, double eps, uint32_t *keys, uint32_t *kmax, hipStream_t s)
{
    const dim3 g(grid_for(n, 256), (nrp + KEY_RPT - 1) / KEY_RPT);  // nrp in 1..KEY_RP
    check_grid((size_t)g.x * g.y, 256);
    hipLaunchKernelGGL((k_keys<T, COSS>), g, dim3(256), 0, s, xs, n, d, rpT, nrp, naxis, eps,
                       keys, kmax);
    MSJ_HIP(hipGetLastError());
}

// Composite (candidate | parent node | key) for candidates [t0, t0+tcount).
__global__ __launch_bounds__(256) void k_cand_keys(const uint32_t *__restrict__ keys,
                                                   const uint32_t *__restrict__ perm,
                                                   const uint32_t *__restrict__ pnode,
                                                   const uint8_t *__restrict__ nsplit, uint32_t n,
                                                   uint32_t t0, uint32_t tcount, int kb, int nb,
                                                   uint64_t *__restrict__ ck)
{
    size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)tcount * n)
        return;
    uint32_t t = (uint32_t)(idx / n), i = (uint32_t)(idx - (size_t)t * n);
    uint32_t node = pnode[i];
    uint64_t k = nsplit[node] ? keys[(size_t)(t0 + t) * n + perm[i]] : 0u;
    ck[idx] = ((uint64_t)t << (nb + kb)) | ((uint64_t)node << kb) | k;
}

__global__ __launch_bounds__(256) void k_run_flags(const uint64_t *__restrict__ ck, size_t m,
                                                   uint32_t *__restrict__ flags)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m)
        flags[i] = (i == 0 || ck[i] != ck[i - 1]) ? 1u : 0u;
}

// starts[r] = first position of run r; starts[#runs] = m.
__global__ __launch_bounds__(256) void k_run_starts(const uint32_t *__restrict__ flags,
                                                    const uint32_t *__restrict__ runidx, size_t m,
                                                    uint32_t *__restrict__ starts)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < m && flags[i])
        starts[runidx[i]] = (uint32_t)i;
    if (i == 0)
        starts[runidx[m]] = (uint32_t)m;
}

__global__ void k_runfirst(const uint32_t *__restrict__ runidx, uint32_t n, uint32_t tcount,
                           uint32_t *__restrict__ runfirst)
{
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t <= tcount)
        runfirst[t] = runidx[(size_t)t * n];
}

// nodes.cu:194-199: sum over nodes, in node order, of (size - N/#nodes)^2.
// Sequential in node order; the next U + 1 run starts are read while the
// current U runs are added.
__global__ void k_cand_sse(const uint32_t *__restrict__ runfirst,
                           const uint32_t *__restrict__ starts, uint32_t tcount, uint32_t n,
                           double *__restrict__ sse, uint32_t *__restrict__ cnt)
{
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= tcount)
        return;
    constexpr uint32_t U = 16;
    const uint32_t r0 = runfirst[t], r1 = runfirst[t + 1];
    const double avg = (double)n * 1.0 / (double)(r1 - r0);
    double s = 0;
    uint32_t r = r0;
    uint32_t st[U + 1];
    if (r + U <= r1) {
#pragma unroll
        for (uint32_t k = 0; k <= U; ++k)
            st[k] = starts[r + k];  // starts holds r1 + 1 entries from r0
    }
    for (; r + U <= r1; r += U) {
        uint32_t nx[U + 1];
        const bool more = r + 2 * U <= r1;
#pragma unroll
        for (uint32_t k = 0; k <= U; ++k)
            nx[k] = more ? starts[r + U + k] : 0u;
#pragma unroll
        for (uint32_t k = 0; k < U; ++k) {
            double df = (double)(st[k + 1] - st[k]) - avg;
            double sq = df * df;
            s = s + sq;
        }
#pragma unroll
        for (uint32_t k = 0; k <= U; ++k)
            st[k] = nx[k];
    }
    for (; r < r1; ++r) {
        double df = (double)(starts[r + 1] - starts[r]) - avg;
        double sq = df * df;
        s = s + sq;
    }
    sse[t] = s;
    cnt[t] = r1 - r0;
}

// Dense form of the candidate node sizes, for layers whose (parent node,
// key) bins fit a workgroup's LDS: bin node * K + key (K = max key + 1; key
// 0 for a parent that stopped splitting) of candidate t counts the points
// with that parent and key.  The non-empty bins in bin order are exactly the
// runs of the composite-key sort, in the same order, so the SSE below adds
// the same terms in the same sequence.
constexpr uint32_t DENSE_BINS = 8192;   // bins per candidate: 32 KB of LDS
constexpr uint32_t DENSE_CHUNK = 8192;  // points per histogram workgroup
constexpr uint32_t STOPPED = 0x80000000u;

// code[p] = first bin of point p's parent, STOPPED set when it stopped splitting.
__global__ __launch_bounds__(256) void k_point_bins(const uint32_t *__restrict__ perm,
                                                    const uint32_t *__restrict__ pnode,
                                                    const uint8_t *__restrict__ nsplit,
                                                    uint32_t n, uint32_t K,
                                                    uint32_t *__restrict__ code)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const uint32_t node = pnode[i];
    code[perm[i]] = node * K | (nsplit[node] ? 0u : STOPPED);
}

// Workgroup (x, t): points [x * DENSE_CHUNK, ...) of candidate t counted in
// LDS (T <= DENSE_BINS bins), then added to tab[t * T + bin].
__global__ __launch_bounds__(256) void k_dense_hist(const uint32_t *__restrict__ keys,
                                                    const uint32_t *__restrict__ code, uint32_t n,
                                                    uint32_t T, uint32_t *__restrict__ tab)
{
    __shared__ uint32_t h[DENSE_BINS];
    const uint32_t t = blockIdx.y, tid = threadIdx.x;
    for (uint32_t b = tid; b < T; b += blockDim.x)
        h[b] = 0;
    __syncthreads();
    const uint32_t p0 = blockIdx.x * DENSE_CHUNK, p1 = min(n, p0 + DENSE_CHUNK);
    const uint32_t *kt = keys + (size_t)t * n;
    for (uint32_t p = p0 + tid; p < p1; p += blockDim.x) {
        const uint32_t c = code[p];
        const uint32_t bin = (c & STOPPED) ? (c & ~STOPPED) : c + kt[p];
        atomicAdd(&h[bin], 1u);
    }
    __syncthreads();
    uint32_t *tt = tab + (size_t)t * T;
    for (uint32_t b = tid; b < T; b += blockDim.x)
        if (h[b])
            atomicAdd(&tt[b], h[b]);
}

// nodes.cu:194-199 over the dense bins of candidate t = blockIdx.x: the
// non-empty bins are counted in parallel, then one thread adds
// (size - N/#nodes)^2 bin by bin in bin order (an empty bin adds +0.0, which
// leaves the non-negative sum unchanged).
__global__ __launch_bounds__(256) void k_dense_sse(const uint32_t *__restrict__ tab, uint32_t T,
                                                   uint32_t n, double *__restrict__ sse,
                                                   uint32_t *__restrict__ cnt)
{
    __shared__ uint32_t h[DENSE_BINS + 32];
    __shared__ uint32_t part[256 / WAVE];
    const uint32_t t = blockIdx.x, tid = threadIdx.x;
    const uint32_t *tt = tab + (size_t)t * T;
    uint32_t c = 0;
    for (uint32_t b = tid; b < T + 32; b += blockDim.x) {
        const uint32_t v = b < T ? tt[b] : 0u;
        h[b] = v;
        c += v ? 1u : 0u;
    }
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1)
        c += (uint32_t)__shfl_down(c, o, WAVE);
    if ((tid & (WAVE - 1)) == 0)
        part[tid / WAVE] = c;
    __syncthreads();
    if (tid != 0)
        return;
    uint32_t nodes = 0;
    for (uint32_t w = 0; w < blockDim.x / WAVE; ++w)
        nodes += part[w];
    const double avg = (double)n * 1.0 / (double)nodes;
    double s = 0;
    uint32_t v[16];
#pragma unroll
    for (uint32_t k = 0; k < 16; ++k)
        v[k] = h[k];
    for (uint32_t b = 0; b < T; b += 16) {
        uint32_t nx[16];
#pragma unroll
        for (uint32_t k = 0; k < 16; ++k)
            nx[k] = h[b + 16 + k];  // b < T, so b + 16 + k < T + 32: zeros past T
#pragma unroll
        for (uint32_t k = 0; k < 16; ++k) {
            double df = (double)v[k] - avg;
            double sq = df * df;
            s = s + (v[k] ? sq : 0.0);
        }
#pragma unroll
        for (uint32_t k = 0; k < 16; ++k)
            v[k] = nx[k];
    }
    sse[t] = s;
    cnt[t] = nodes;
}

// nodes.cu:198,207: the candidate with the smallest SSE, the lowest index on
// a tie (strict '<'); kept per layer for the statistics read after the last
// layer, so the host never waits for it.
__global__ void k_pick_best(const double *__restrict__ sse, uint32_t ncand,
                            uint32_t *__restrict__ best, double *__restrict__ best_sse)
{
    if (threadIdx.x != 0)
        return;
    uint32_t b = 0;
    for (uint32_t t = 1; t < ncand; ++t)
        if (sse[t] < sse[b])
            b = t;
    *best = b;
    *best_sse = sse[b];
}

// Keys of the winner *best (candidate rows of `keys`, n each).
__global__ __launch_bounds__(256) void k_apply_keys(const uint32_t *__restrict__ keys,
                                                    const uint32_t *__restrict__ best,
                                                    const uint32_t *__restrict__ perm,
                                                    const uint32_t *__restrict__ pnode,
                                                    const uint8_t *__restrict__ nsplit, uint32_t n,
                                                    int kb, uint64_t *__restrict__ ck,
                                                    uint32_t *__restrict__ vals)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const uint32_t *key_t = keys + (size_t)(*best) * n;
    uint32_t node = pnode[i];
    uint32_t row = perm[i];
    uint64_t k = nsplit[node] ? key_t[row] : 0u;
    ck[i] = ((uint64_t)node << kb) | k;
    vals[i] = row;
}

// Children of the new layer: start, parent, key (UINT_MAX when the parent
// stopped splitting, as splitNodes stores it, nodes.cu:482).
__global__ __launch_bounds__(256) void k_new_nodes(const uint64_t *__restrict__ ck,
                                                   const uint32_t *__restrict__ flags,
                                                   const uint32_t *__restrict__ runidx, uint32_t n,
                                                   int kb, const uint8_t *__restrict__ nsplit,
                                                   uint32_t *__restrict__ cstart,
                                                   uint32_t *__restrict__ cparent,
                                                   uint32_t *__restrict__ ckey,
                                                   uint32_t *__restrict__ pnode_new)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    uint32_t c = runidx[i] + flags[i] - 1u;
    pnode_new[i] = c;
    if (flags[i]) {
        uint64_t v = ck[i];
        uint32_t parent = (uint32_t)(v >> kb);
        uint64_t kmask = kb ? ((kb >= 64) ? ~0ull : ((1ull << kb) - 1)) : 0ull;
        cstart[c] = i;
        cparent[c] = parent;
        ckey[c] = nsplit[parent] ? (uint32_t)(v & kmask) : 0xFFFFFFFFu;
    }
    if (i == 0)
        cstart[runidx[n]] = n;
}

__global__ __launch_bounds__(256) void k_cfirst(const uint32_t *__restrict__ cparent, uint32_t mnew,
                                                uint32_t mold, uint32_t *__restrict__ cfirst)
{
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < mnew && (c == 0 || cparent[c] != cparent[c - 1]))
        cfirst[cparent[c]] = c;
    if (c == 0)
        cfirst[mold] = mnew;
}

struct AdjView {
    const uint32_t *cparent, *ckey, *cstart, *cfirst;
    const uint8_t *osplit;
    const uint32_t *oaoff, *oadj;
};

__device__ inline uint32_t lower_bound_u32(const uint32_t *a, uint32_t lo, uint32_t hi, uint32_t v)
{
    while (lo < hi) {
        uint32_t mid = lo + (hi - lo) / 2;
        if (a[mid] < v)
            lo = mid + 1;
        else
            hi = mid;
    }
    return lo;
}

// Child c of parent P (key k; s: P split) and one entry Q of P's
// neighbour list: the children of Q that are c's neighbours, in child order
// (nodes.cu:557-578 siblings when Q == P, nodes.cu:719-747 otherwise).
template <class F>
__device__ inline void visit_entry(const AdjView &A, uint32_t c, uint32_t P, uint32_t k, bool s,
                                   uint32_t pf, uint32_t pl, uint32_t Q, F &&emit)
{
    if (Q == P) {
        emit(c);
        if (s) {
            if (c > pf && A.ckey[c - 1] + 1u == k)
                emit(c - 1);
            if (c + 1 < pl && A.ckey[c + 1] - 1u == k)
                emit(c + 1);
        }
        return;
    }
    const uint32_t q0 = A.cfirst[Q], q1 = A.cfirst[Q + 1];
    if (!s || !A.osplit[Q]) {
        for (uint32_t v = q0; v < q1; ++v)
            emit(v);
        return;
    }
    // keys k-1, k, k+1 modulo 2^32; Q's child keys ascend
    uint32_t hit[3];
    int nh = 0;
    const uint32_t want[3] = {k - 1u, k, k + 1u};
    for (int w = 0; w < 3; ++w) {
        uint32_t pos = lower_bound_u32(A.ckey, q0, q1, want[w]);
        if (pos < q1 && A.ckey[pos] == want[w])
            hit[nh++] = pos;
    }
    for (int x = 1; x < nh; ++x)  // emit in child order
        for (int y = x; y > 0 && hit[y - 1] > hit[y]; --y) {
            uint32_t tmp = hit[y];
            hit[y] = hit[y - 1];
            hit[y - 1] = tmp;
        }
    for (int x = 0; x < nh; ++x)
        emit(hit[x]);
}

// One wave per child c (wave-uniform): its lanes take the entries of the
// parent's neighbour list in turn, so the per-entry searches run side by
// side; the neighbours come out in the reference's order, entry by entry in
// list order.
__global__ __launch_bounds__(256) void k_adj_count(AdjView A, uint32_t mnew,
                                                   uint32_t *__restrict__ deg,
                                                   uint64_t *__restrict__ sumsz,
                                                   uint64_t *__restrict__ ncalcs)
{
    const uint32_t c = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    const uint32_t lane = threadIdx.x & (WAVE - 1);
    if (c >= mnew)
        return;
    const uint32_t P = A.cparent[c], k = A.ckey[c];
    const bool s = A.osplit[P] != 0;
    const uint32_t pf = A.cfirst[P], pl = A.cfirst[P + 1];
    uint32_t cnt = 0;
    uint64_t sz = 0;
    for (uint32_t a = A.oaoff[P] + lane; a < A.oaoff[P + 1]; a += WAVE)
        visit_entry(A, c, P, k, s, pf, pl, A.oadj[a], [&](uint32_t v) {
            ++cnt;
            sz += A.cstart[v + 1] - A.cstart[v];
        });
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) {
        cnt += (uint32_t)__shfl_xor(cnt, o, WAVE);
        sz += (uint64_t)__shfl_xor((unsigned long long)sz, o, WAVE);
    }
    if (lane == 0) {
        deg[c] = cnt;
        sumsz[c] = sz;
        ncalcs[c] = sz * (uint64_t)(A.cstart[c + 1] - A.cstart[c]);  // nodes.cu:691
    }
}

__global__ __launch_bounds__(256) void k_adj_write(AdjView A, uint32_t mnew,
                                                   const uint32_t *__restrict__ aoff,
                                                   uint32_t *__restrict__ adj)
{
    const uint32_t c = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE;
    const uint32_t lane = threadIdx.x & (WAVE - 1);
    if (c >= mnew)
        return;
    const uint32_t P = A.cparent[c], k = A.ckey[c];
    const bool s = A.osplit[P] != 0;
    const uint32_t pf = A.cfirst[P], pl = A.cfirst[P + 1];
    const uint32_t a0 = A.oaoff[P], a1 = A.oaoff[P + 1];
    uint32_t w = aoff[c];
    for (uint32_t base = a0; base < a1; base += WAVE) {
        const uint32_t a = base + lane;
        const uint32_t Q = a < a1 ? A.oadj[a] : 0u;
        uint32_t cnt = 0;
        if (a < a1)
            visit_entry(A, c, P, k, s, pf, pl, Q, [&](uint32_t) { ++cnt; });
        uint32_t incl = cnt;  // inclusive prefix over the lanes
#pragma unroll
        for (int o = 1; o < WAVE; o <<= 1) {
            const uint32_t up = (uint32_t)__shfl_up(incl, o, WAVE);
            if (lane >= (uint32_t)o)
                incl += up;
        }
        uint32_t pos = w + incl - cnt;
        if (a < a1)
            visit_entry(A, c, P, k, s, pf, pl, Q, [&](uint32_t v) { adj[pos++] = v; });
        w += (uint32_t)__shfl(incl, WAVE - 1, WAVE);
    }
}

__global__ __launch_bounds__(256) void k_nsplit(const uint64_t *__restrict__ ncalcs, uint32_t m,
                                                uint64_t minc, uint8_t *__restrict__ nsplit)
{
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c < m)
        nsplit[c] = ncalcs[c] >= minc ? 1 : 0;  // nodes.cu:476
}

__global__ void k_root(uint8_t *nsplit, uint32_t *aoff, uint32_t *adj)
{
    if (threadIdx.x == 0) {
        nsplit[0] = 1;
        aoff[0] = 0;
        aoff[1] = 1;
        adj[0] = 0;
    }
}

template <class T>
void swap_buf(DevBuf<T> &a, DevBuf<T> &b)
{
    std::swap(a.p, b.p);
    std::swap(a.cap, b.cap);
    std::swap(a.borrowed, b.borrowed);
}

template <class T>
T d2h_scalar(const T *dptr, hipStream_t s)
{
    T v;
    MSJ_HIP(hipMemcpyAsync(&v, dptr, sizeof(T), hipMemcpyDeviceToHost, s));
    MSJ_HIP(hipStreamSynchronize(s));
    return v;
}

#define LAUNCH(kern, grid, block, ...)                                                   \
    do {                                                                                 \
        ::msj::check_grid((grid), (block));                                              \
        hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, s, __VA_ARGS__);            \
        MSJ_HIP(hipGetLastError());                                                      \
    } while (0)

// Host state of the reference-point generation of one build: the test
// candidates and sigmas of a group of layers, each layer's reference points.
struct RpHost {
    uint32_t nrp, d, l0 = 0, lc = 0;  // layers [l0, l0 + lc) are in trp / sig / rpT
    std::vector<double> trp, trpT, sig, maxc, box, rp, rpT;
    std::vector<uint32_t> order;
    RpHost(uint32_t nrp, uint32_t d)
        : nrp(nrp), d(d), maxc(d), box((size_t)2 * nrp * d), rp((size_t)nrp * d),
          order(2 * nrp)
    {
    }
};

// Layers whose test candidates share one distance matrix (S x lc*2*nrp
// doubles, at most about 256 MiB).
inline uint32_t rp_group_layers(uint32_t S, uint32_t ntest, uint32_t left)
{
    const uint64_t per = (uint64_t)std::max<uint32_t>(S, 1) * ntest * sizeof(double);
    const uint64_t fit = std::max<uint64_t>(1, (256ull << 20) / per);
    return (uint32_t)std::min<uint64_t>(left, fit);
}

// utils.cu:66-137 createRPArray for layers [l0, l0 + lc).  The glibc stream's
// draws do not depend on any layer's result, so the 2*nrp test candidates of
// every layer in the group are drawn in stream order up front, their
// sample-distance sigmas come from one distance matrix and one sigma launch
// (one column per candidate, each summed sequentially as before), and the
// sample's max corner, the same for every layer, from one reduction; one
// round trip to the host then gives every layer's stable sigma order and box
// corners.  Layer l's reference points end up transposed in
// B.rpT + (l - l0) * d * KEY_RP.
template <class T>
void group_rps(IndexBuilder &B, const T *xs, uint32_t n, uint32_t d, uint32_t S, uint32_t nrp,
               uint32_t l0, uint32_t lc, GlibcRand &rng, RpHost &h, hipStream_t s)
{
    const uint32_t ntest = nrp * 2, nt = lc * ntest;
    h.l0 = l0;
    h.lc = lc;
    h.trp.resize((size_t)nt * d);
    h.sig.resize(nt);
    h.rpT.assign((size_t)lc * d * KEY_RP, 0.0);
    for (size_t i = 0; i < h.trp.size(); ++i)
        h.trp[i] = (double)rng.next() / (double)2147483647;
    h.trpT.resize(h.trp.size());
    for (uint32_t j = 0; j < nt; ++j)
        for (uint32_t k = 0; k < d; ++k)
            h.trpT[(size_t)k * nt + j] = h.trp[(size_t)j * d + k];
    B.test_rp.ensure((size_t)nt * d);
    B.distmat.ensure((size_t)std::max<uint32_t>(S, 1) * nt);
    B.sigma.ensure(nt);
    B.rpT.ensure((size_t)lc * d * KEY_RP);
    MSJ_HIP(hipMemcpyAsync(B.test_rp.p, h.trpT.data(), h.trpT.size() * sizeof(double),
                           hipMemcpyHostToDevice, s));
    if (S > 0)
        LAUNCH(k_rp_distmat<T>, grid_for((size_t)S * nt, 256), 256, xs, n, d, S, B.test_rp.p, nt,
               B.distmat.p);
    LAUNCH(k_col_sigma, nt, SIG_BLOCK, B.distmat.p, S, (size_t)nt, (size_t)1, B.sigma.p);
    if (l0 == 0) {
        LAUNCH(k_col_max<T>, d, 256, xs, n, d, S, B.maxc.p);
        MSJ_HIP(hipMemcpyAsync(h.maxc.data(), B.maxc.p, d * sizeof(double),
                               hipMemcpyDeviceToHost, s));
    }
    MSJ_HIP(hipMemcpyAsync(h.sig.data(), B.sigma.p, nt * sizeof(double), hipMemcpyDeviceToHost,
                           s));
    MSJ_HIP(hipStreamSynchronize(s));
    // box points: max corner, then corner patterns (utils.cu:104-136)
    for (uint32_t j = 0; j < d; ++j)
        h.box[j] = h.maxc[j];
    const uint32_t step = host_cvt_u32(std::pow(2.0, (double)d) / ntest);
    for (uint32_t i = 1; i < ntest; ++i)
        for (uint32_t j = 0; j < d; ++j) {
            uint32_t div = (uint32_t)host_cvt_i32(std::pow(2.0, (double)j));
            uint32_t bit = ((i * step) / div) % 2u;
            h.box[(size_t)i * d + j] = (double)bit * h.box[j];
        }
    for (uint32_t g = 0; g < lc; ++g) {
        const double *trp = h.trp.data() + (size_t)g * ntest * d;
        const double *sig = h.sig.data() + (size_t)g * ntest;
        // thrust::sort_by_key (stable, IEEE radix order) ascending, reversed
        for (uint32_t i = 0; i < ntest; ++i)
            h.order[i] = i;
        std::stable_sort(h.order.begin(), h.order.end(), [&](uint32_t a, uint32_t b) {
            return radix_key_f64(sig[a]) < radix_key_f64(sig[b]);
        });
        std::reverse(h.order.begin(), h.order.end());
        for (uint32_t i = 0; i < nrp / 2; ++i)
            for (uint32_t j = 0; j < d; ++j)
                h.rp[(size_t)i * d + j] = trp[(size_t)h.order[i] * d + j];
        for (uint32_t i = 0; i < nrp / 2 + nrp % 2; ++i)
            for (uint32_t j = 0; j < d; ++j)
                h.rp[(size_t)(nrp / 2 + i) * d + j] = h.box[(size_t)i * d + j];
        double *rpT = h.rpT.data() + (size_t)g * d * KEY_RP;
        for (uint32_t t = 0; t < nrp; ++t)
            for (uint32_t j = 0; j < d; ++j)
                rpT[(size_t)j * KEY_RP + t] = h.rp[(size_t)t * d + j];
    }
    MSJ_HIP(hipMemcpyAsync(B.rpT.p, h.rpT.data(), h.rpT.size() * sizeof(double),
                           hipMemcpyHostToDevice, s));
}

// Layer l's transposed reference points (d x KEY_RP), drawing the next group
// of layers first when l is past the current one.
template <class T>
const double *layer_rps(IndexBuilder &B, const T *xs, uint32_t n, uint32_t d, uint32_t S,
                        uint32_t nrp, uint32_t l, uint32_t layers, GlibcRand &rng, RpHost &h,
                        hipStream_t s)
{
    if (h.lc == 0 || l >= h.l0 + h.lc)
        group_rps(B, xs, n, d, S, nrp, l, rp_group_layers(S, 2 * nrp, layers - l), rng, h, s);
    return B.rpT.p + (size_t)(l - h.l0) * d * KEY_RP;
}

template <class T>
void build_impl(IndexBuilder &B, const T *xs, uint32_t n, uint32_t d, double eps,
                const msj_params &p, Index &out, LayerStats &st, hipStream_t s)
{
    const uint32_t nrp = p.rp_per_layer;
    const uint32_t naxis = std::min(p.numidim, d);
    const uint32_t ncand = nrp + naxis;
    const uint32_t layers = std::min<uint32_t>(p.maxrp, MSJ_MAX_LAYERS);
    // (unsigned)(numPoints * SAMPLE_PER), utils.cu:68
    const uint32_t S = host_cvt_u32((double)(unsigned long long)n * p.sample_per);
    st.ncand = ncand;
    st.layers = layers;

    GlibcRand rng;
    rng.seed(p.seed);

    B.keys.ensure((size_t)ncand * n);
    if (layers > 1)
        B.keys2.ensure((size_t)ncand * n);
    B.maxc.ensure(d);
    B.u32scalar.ensure(std::max<uint32_t>(layers, 1));
    B.u64scalar.ensure(std::max<uint32_t>(layers, 1));
    B.best.ensure(std::max<uint32_t>(layers, 1));
    B.best_sse.ensure(std::max<uint32_t>(layers, 1));
    out.perm.ensure(n);
    B.vals.ensure(n);
    B.pnode.ensure(n);
    B.pnode2.ensure(n);
    B.nsplit.ensure(1);
    B.aoff1.ensure(2);
    B.adj1.ensure(1);

    LAUNCH(k_iota, grid_for(n, 256), 256, out.perm.p, n);
    LAUNCH(k_fill<uint32_t>, grid_for(n, 256), 256, B.pnode.p, (size_t)n, 0u);
    LAUNCH(k_root, 1, 64, B.nsplit.p, B.aoff1.p, B.adj1.p);
    uint32_t mold = 1;

    RpHost rh(nrp, d);
    // MSJ_NODE_SSE=sort keeps every layer on the composite-key sort (tests
    // compare the two forms); otherwise layers whose bins fit use the dense one.
    const char *sse_env = std::getenv("MSJ_NODE_SSE");
    const bool sse_by_sort = sse_env && std::strcmp(sse_env, "sort") == 0;

    // ---- (1) reference points, utils.cu:66-137, and (2) the keys of every
    // candidate partition of layer l into `dst`, the largest key into
    // u32scalar[l].  Keys do not depend on the node net: layer l + 1's are
    // computed during layer l, and its largest key comes back with layer l's
    // adjacency size.
    LAUNCH(k_fill<uint32_t>, 1, 64, B.u32scalar.p, (size_t)std::max<uint32_t>(layers, 1), 0u);
    auto layer_keys = [&](uint32_t l, uint32_t *dst) {
        const double *rpT = layer_rps(B, xs, n, d, S, nrp, l, layers, rng, rh, s);
        launch_keys<T, false>(xs, n, d, rpT, nrp, naxis, eps, dst, B.u32scalar.p + l, s);
        if (B.probe && B.probe->keys && B.probe->layer == l) {
            MSJ_HIP(hipMemcpyAsync(B.probe->keys, dst, sizeof(uint32_t) * ncand * n,
                                   hipMemcpyDeviceToHost, s));
            MSJ_HIP(hipStreamSynchronize(s));
        }
    };
    uint32_t kmax_next = 0;
    if (layers) {
        layer_keys(0, B.keys.p);
        kmax_next = d2h_scalar(B.u32scalar.p, s);
    }

    for (uint32_t l = 0; l < layers; ++l) {
        const uint32_t kmax = kmax_next;
        const int kb = bits_for(kmax);
        const int nb = bits_for(mold - 1);

        // ---- (3) candidate node sizes and SSE
        int tb_all = bits_for(ncand - 1);
        uint32_t group = ncand;
        if (tb_all + nb + kb > 64) {
            int room = 64 - nb - kb;  // >= 0 since nb, kb <= 32
            group = room >= 31 ? ncand : std::max<uint32_t>(1u, std::min<uint32_t>(ncand, 1u << room));
        }
        B.sse.ensure(ncand);
        B.cnt_nodes.ensure(ncand);
        B.runfirst.ensure((size_t)ncand + 1);
        const uint64_t K = (uint64_t)kmax + 1, nbins = (uint64_t)mold * K;
        const bool dense = nbins <= DENSE_BINS && !sse_by_sort;
        if (dense) {
            const size_t tabn = (size_t)ncand * nbins;
            B.code.ensure(n);
            B.tab.ensure(tabn);
            LAUNCH(k_point_bins, grid_for(n, 256), 256, out.perm.p, B.pnode.p, B.nsplit.p, n,
                   (uint32_t)K, B.code.p);
            LAUNCH(k_fill<uint32_t>, grid_for(tabn, 256), 256, B.tab.p, tabn, 0u);
            const dim3 hg((n + DENSE_CHUNK - 1) / DENSE_CHUNK, ncand);
            hipLaunchKernelGGL(k_dense_hist, hg, dim3(256), 0, s, B.keys.p, B.code.p, n,
                               (uint32_t)nbins, B.tab.p);
            MSJ_HIP(hipGetLastError());
            LAUNCH(k_dense_sse, ncand, 256, B.tab.p, (uint32_t)nbins, n, B.sse.p, B.cnt_nodes.p);
        }
        for (uint32_t t0 = 0; !dense && t0 < ncand; t0 += group) {
            const uint32_t tc = std::min(group, ncand - t0);
            const size_t m = (size_t)tc * n;
            const int tb = bits_for(tc - 1);
            B.ck.ensure(m);
            B.flags.ensure(m);
            B.runidx.ensure(m + 1);
            B.starts.ensure(m + 1);
            LAUNCH(k_cand_keys, grid_for(m, 256), 256, B.keys.p, out.perm.p, B.pnode.p,
                   B.nsplit.p, n, t0, tc, kb, nb, B.ck.p);
            B.sorter.sort(B.ck.p, nullptr, m, tb + nb + kb, s);
            LAUNCH(k_run_flags, grid_for(m, 256), 256, B.ck.p, m, B.flags.p);
            B.scan.excl_u32_to_u32(B.flags.p, B.runidx.p, m, s);
            LAUNCH(k_run_starts, grid_for(m, 256), 256, B.flags.p, B.runidx.p, m, B.starts.p);
            LAUNCH(k_runfirst, 1, 64 * ((tc + 1 + 63) / 64), B.runidx.p, n, tc, B.runfirst.p);
            LAUNCH(k_cand_sse, 1, 64 * ((tc + 63) / 64), B.runfirst.p, B.starts.p, tc, n,
                   B.sse.p + t0, B.cnt_nodes.p + t0);
        }
        LAUNCH(k_pick_best, 1, 64, B.sse.p, ncand, B.best.p + l, B.best_sse.p + l);

        // ---- (4) apply the winner
        B.ck.ensure(n);
        B.flags.ensure(n);
        B.runidx.ensure((size_t)n + 1);
        LAUNCH(k_apply_keys, grid_for(n, 256), 256, B.keys.p, B.best.p + l, out.perm.p,
               B.pnode.p, B.nsplit.p, n, kb, B.ck.p, B.vals.p);
        B.sorter.sort(B.ck.p, B.vals.p, n, nb + kb, s);
        LAUNCH(k_run_flags, grid_for(n, 256), 256, B.ck.p, (size_t)n, B.flags.p);
        B.scan.excl_u32_to_u32(B.flags.p, B.runidx.p, n, s);
        const uint32_t mnew = d2h_scalar(B.runidx.p + n, s);
        B.cstart.ensure((size_t)mnew + 1);
        B.cparent.ensure(mnew);
        B.ckey.ensure(mnew);
        B.cfirst.ensure((size_t)mold + 1);
        LAUNCH(k_new_nodes, grid_for(n, 256), 256, B.ck.p, B.flags.p, B.runidx.p, n, kb,
               B.nsplit.p, B.cstart.p, B.cparent.p, B.ckey.p, B.pnode2.p);
        LAUNCH(k_cfirst, grid_for(mnew, 256), 256, B.cparent.p, mnew, mold, B.cfirst.p);
        if (B.probe && B.probe->node_keys) {
            B.probe->cparent[l].resize(mnew);
            B.probe->ckey[l].resize(mnew);
            MSJ_HIP(hipMemcpyAsync(B.probe->cparent[l].data(), B.cparent.p, sizeof(uint32_t) * mnew,
                                   hipMemcpyDeviceToHost, s));
            MSJ_HIP(hipMemcpyAsync(B.probe->ckey[l].data(), B.ckey.p, sizeof(uint32_t) * mnew,
                                   hipMemcpyDeviceToHost, s));
            MSJ_HIP(hipStreamSynchronize(s));
        }

        // ---- (5) adjacency of the new layer
        B.deg.ensure(mnew);
        B.sumsz2.ensure(mnew);
        B.ncalcs2.ensure(mnew);
        B.aoff2.ensure((size_t)mnew + 1);
        AdjView A{B.cparent.p, B.ckey.p, B.cstart.p, B.cfirst.p, B.nsplit.p, B.aoff1.p, B.adj1.p};
        LAUNCH(k_adj_count, grid_for((size_t)mnew * WAVE, 256), 256, A, mnew, B.deg.p, B.sumsz2.p,
               B.ncalcs2.p);
        {
            // degree prefix in 64 bits first, to detect > 2^32 adjacency entries
            B.tmp64.ensure((size_t)mnew + 1);
            B.scan.excl_u32_to_u64(B.deg.p, B.tmp64.p, mnew, s);
            if (l + 1 < layers)
                layer_keys(l + 1, B.keys2.p);
            uint64_t edges;
            MSJ_HIP(hipMemcpyAsync(&edges, B.tmp64.p + mnew, sizeof(edges),
                                   hipMemcpyDeviceToHost, s));
            if (l + 1 < layers)
                MSJ_HIP(hipMemcpyAsync(&kmax_next, B.u32scalar.p + l + 1, sizeof(kmax_next),
                                       hipMemcpyDeviceToHost, s));
            MSJ_HIP(hipStreamSynchronize(s));
            if (edges >= 0xFFFFFFFFull)
                throw Error(MSJ_E_OVERFLOW, "node adjacency exceeds 2^32 entries");
            B.scan.excl_u32_to_u32(B.deg.p, B.aoff2.p, mnew, s);
            B.adj2.ensure(std::max<uint64_t>(edges, 1));
            LAUNCH(k_adj_write, grid_for((size_t)mnew * WAVE, 256), 256, A, mnew, B.aoff2.p,
                   B.adj2.p);
            out.edges = (uint32_t)edges;
        }
        B.nsplit2.ensure(mnew);
        LAUNCH(k_nsplit, grid_for(mnew, 256), 256, B.ncalcs2.p, mnew, p.min_calcs_per_node,
               B.nsplit2.p);
        // the layer's candidate evaluations, read back once after the last layer
        reduce_sum_u64(B.ncalcs2.p, mnew, B.u64scalar.p + l, s);

        swap_buf(out.perm, B.vals);
        swap_buf(B.pnode, B.pnode2);
        swap_buf(B.nsplit, B.nsplit2);
        swap_buf(B.aoff1, B.aoff2);
        swap_buf(B.adj1, B.adj2);
        swap_buf(B.keys, B.keys2);
        mold = mnew;
        st.nodes[l] = mnew;
    }
    if (layers) {
        std::vector<unsigned long long> calcs(layers);
        std::vector<uint32_t> best(layers);
        std::vector<double> best_sse(layers);
        MSJ_HIP(hipMemcpyAsync(calcs.data(), B.u64scalar.p, sizeof(unsigned long long) * layers,
                               hipMemcpyDeviceToHost, s));
        MSJ_HIP(hipMemcpyAsync(best.data(), B.best.p, sizeof(uint32_t) * layers,
                               hipMemcpyDeviceToHost, s));
        MSJ_HIP(hipMemcpyAsync(best_sse.data(), B.best_sse.p, sizeof(double) * layers,
                               hipMemcpyDeviceToHost, s));
        MSJ_HIP(hipStreamSynchronize(s));
        for (uint32_t l = 0; l < layers; ++l) {
            st.chosen[l] = best[l];
            st.sse[l] = best_sse[l];
            st.calcs[l] = calcs[l];
            if (p.verbose)
                printf("Layer %u Selecting RP %u with Nodes: %u and calcs: %llu and stdev %f\n",
                       l, st.chosen[l], st.nodes[l], calcs[l], st.sse[l]);
        }
        out.num_calcs = calcs[layers - 1];
    }
    out.M = mold;
    swap_buf(out.nstart, B.cstart);
    swap_buf(out.aoff, B.aoff1);
    swap_buf(out.adj, B.adj1);
    swap_buf(out.sumsz, B.sumsz2);
    swap_buf(out.ncalcs, B.ncalcs2);
}

__global__ void k_single_node(uint32_t n, uint32_t *nstart, uint32_t *aoff, uint32_t *adj,
                              uint64_t *sumsz, uint64_t *ncalcs)
{
    if (threadIdx.x == 0) {
        nstart[0] = 0;
        nstart[1] = n;
        aoff[0] = 0;
        aoff[1] = 1;
        adj[0] = 0;
        sumsz[0] = n;
        ncalcs[0] = (uint64_t)n * n;
    }
}


// ------------------------------------------------------------------------
// Tree/COSS bins: buildTree with MINSQRS (tree.cu:3-593) and the 3^L
// neighbour probes of generateRanges / binarySearch (tree.cu:596-759,
// 886-996).  Every layer scores all candidates (the layer's reference points
// and min(L, d) axes) by the sum of squared deviations of every bin of the
// layer (empty ones included) from N / #non-empty, and keeps the first
// smallest.  A bin address is (rank of the point's previous bin - 1) * nb +
// (key - skip); the ranks of the non-empty bins carry to the next layer.
// Candidates whose layer needs 2^32 bins or more, or whose key span wraps
// the unsigned bin count, are not eligible (the reference overflows there).

struct CossSlot {
    uint32_t t, skip, nb, pad;
};

struct CossWalk {
    uint32_t j, avg;
    uint64_t count;
    double s0;
};

__global__ void k_minmax_init(uint32_t *mm, uint32_t ncand)
{
    uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < ncand) {
        mm[2 * t] = 0xFFFFFFFFu;
        mm[2 * t + 1] = 0u;
    }
}

// tree.cu:181-196 / 281-287: smallest and largest key of every candidate
// (grid.y = candidate).
__global__ __launch_bounds__(256) void k_cand_minmax(const uint32_t *__restrict__ keys, uint32_t n,
                                                     uint32_t *__restrict__ mm)
{
    const uint32_t t = blockIdx.y;
    const uint32_t *k = keys + (size_t)t * n;
    uint32_t lo = 0xFFFFFFFFu, hi = 0u;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x) {
        const uint32_t v = k[i];
        lo = v < lo ? v : lo;
        hi = v > hi ? v : hi;
    }
    for (int o = WAVE / 2; o > 0; o >>= 1) {
        const uint32_t a = __shfl_xor(lo, o, WAVE), b = __shfl_xor(hi, o, WAVE);
        lo = a < lo ? a : lo;
        hi = b > hi ? b : hi;
    }
    if ((threadIdx.x & (WAVE - 1)) == 0) {
        atomicMin(&mm[2 * t], lo);
        atomicMax(&mm[2 * t + 1], hi);
    }
}

// tree.cu:224-253: (slot << ab) | bin address of every point under the
// eligible candidates [j0, j0 + jc).
__global__ __launch_bounds__(256) void k_coss_addr(const uint32_t *__restrict__ keys,
                                                   const uint32_t *__restrict__ prank,
                                                   const CossSlot *__restrict__ slots,
                                                   uint32_t jc, uint32_t n, int first, int ab,
                                                   uint64_t *__restrict__ ck)
{
    const size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= (size_t)jc * n)
        return;
    const uint32_t j = (uint32_t)(idx / n), i = (uint32_t)(idx - (size_t)j * n);
    const CossSlot c = slots[j];
    const uint32_t bn = keys[(size_t)c.t * n + i] - c.skip;
    const uint32_t addr = first ? bn : (prank[i] - 1u) * c.nb + bn;
    ck[idx] = ((uint64_t)j << ab) | addr;
}

__device__ inline unsigned long long sat_add(unsigned long long a, unsigned long long b)
{
    const unsigned long long r = a + b;
    return r < a ? ~0ull : r;
}

// Per slot (one block each): sum over its non-empty bins of
// (count - N / #non-empty)^2 in integers, saturating at 2^64 - 1.
__global__ __launch_bounds__(256) void k_coss_stats(const uint32_t *__restrict__ runfirst,
                                                    const uint32_t *__restrict__ starts, uint32_t n,
                                                    unsigned long long *__restrict__ acc)
{
    __shared__ unsigned long long red[256 / WAVE];
    const uint32_t j = blockIdx.x;
    const uint32_t r0 = runfirst[j], r1 = runfirst[j + 1];
    const uint32_t A = n / (r1 - r0);
    unsigned long long a = 0;
    for (uint32_t r = r0 + threadIdx.x; r < r1; r += blockDim.x) {
        const uint32_t c = starts[r + 1] - starts[r];
        const uint64_t df = c > A ? c - A : A - c;
        a = sat_add(a, df * df);
    }
    for (int o = WAVE / 2; o > 0; o >>= 1)
        a = sat_add(a, __shfl_xor(a, o, WAVE));
    if ((threadIdx.x & (WAVE - 1)) == 0)
        red[threadIdx.x / WAVE] = a;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned long long t = 0;
        for (int w = 0; w < 256 / WAVE; ++w)
            t = sat_add(t, red[w]);
        acc[j] = t;
    }
}

// tree.cu:272-275 / 387-394 in the reference's own order: the statistic
// of one candidate as the sequential FP64 sum over all of its bins (used
// when the exact integer sum reaches 2^53 and rounding makes the order
// matter).  One thread per candidate; runs of empty bins are added by
// repeat_add.
__global__ void k_coss_walk(const CossWalk *__restrict__ w, uint32_t nw,
                            const uint32_t *__restrict__ runfirst,
                            const uint32_t *__restrict__ starts, const uint64_t *__restrict__ ck,
                            uint64_t amask, double *__restrict__ out)
{
    const uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nw)
        return;
    const CossWalk W = w[k];
    const double a = (double)W.avg;
    const double c = a * a;  // (0 - avg)^2 of an empty bin
    double s = W.s0;
    uint64_t pos = 0;
    for (uint32_t r = runfirst[W.j]; r < runfirst[W.j + 1]; ++r) {
        const uint32_t st = starts[r];
        const uint64_t addr = ck[st] & amask;
        s = repeat_add(s, c, addr - pos);
        const double e = (double)(starts[r + 1] - st) - a;
        const double sq = e * e;
        s = s + sq;
        pos = addr + 1;
    }
    out[k] = repeat_add(s, c, W.count - pos);
}

// The chosen candidate: bin address (sort key), row, bin number.
__global__ __launch_bounds__(256) void k_coss_best(const uint32_t *__restrict__ key,
                                                   const uint32_t *__restrict__ prank, uint32_t n,
                                                   int first, uint32_t skip, uint32_t nb,
                                                   uint64_t *__restrict__ ck,
                                                   uint32_t *__restrict__ vals,
                                                   uint32_t *__restrict__ bnum)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const uint32_t bn = key[i] - skip;
    ck[i] = first ? bn : (prank[i] - 1u) * nb + bn;
    vals[i] = i;
    bnum[i] = bn;
}

// tree.cu:416-430: 1-based rank of every point's non-empty bin.
__global__ __launch_bounds__(256) void k_coss_rank(const uint32_t *__restrict__ flags,
                                                   const uint32_t *__restrict__ runidx,
                                                   const uint32_t *__restrict__ vals, uint32_t n,
                                                   uint32_t *__restrict__ prank)
{
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n)
        prank[vals[q]] = runidx[q] + flags[q];
}

__global__ __launch_bounds__(256) void k_u32_to_u64(const uint32_t *__restrict__ a, uint32_t n,
                                                    uint64_t *__restrict__ b)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        b[i] = a[i];
}

// Bin-number tuple of every final bin (node order = lexicographic order).
__global__ __launch_bounds__(256) void k_coss_tuples(const uint32_t *__restrict__ perm,
                                                     const uint32_t *__restrict__ nstart,
                                                     const uint32_t *__restrict__ bnum, uint32_t n,
                                                     uint32_t M, uint32_t L,
                                                     uint32_t *__restrict__ tup)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= M)
        return;
    const uint32_t row = perm[nstart[v]];
    for (uint32_t l = 0; l < L; ++l)
        tup[(size_t)v * L + l] = bnum[(size_t)l * n + row];
}

// tree.cu:886-996: the bins whose bin numbers differ from v's by at most
// one in every layer (the 3^L probes), found by a depth-first walk over
// the sorted tuples that narrows one layer at a time, so the hits come out
// in node order.  WRITE = false counts them (degree, sum of their sizes,
// numCalcs = |v| * that sum, tree.cu:756); WRITE = true lists them.
template <bool WRITE>
__global__ __launch_bounds__(256) void k_coss_adj(const uint32_t *__restrict__ tup, uint32_t M,
                                                  uint32_t L, const uint32_t *__restrict__ nstart,
                                                  uint32_t *__restrict__ deg,
                                                  const uint32_t *__restrict__ aoff,
                                                  uint32_t *__restrict__ adj,
                                                  uint64_t *__restrict__ sumsz,
                                                  uint64_t *__restrict__ ncalcs)
{
    const uint32_t v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= M)
        return;
    const uint32_t *key = tup + (size_t)v * L;
    uint32_t lo[MSJ_MAX_LAYERS], hi[MSJ_MAX_LAYERS];
    uint8_t dig[MSJ_MAX_LAYERS];
    lo[0] = 0;
    hi[0] = M;
    dig[0] = 0;
    uint32_t cnt = 0, w = WRITE ? aoff[v] : 0;
    uint64_t sz = 0;
    int l = 0;
    while (l >= 0) {
        if (dig[l] == 3) {
            if (--l >= 0)
                ++dig[l];
            continue;
        }
        const uint32_t want = key[l] + (uint32_t)dig[l] - 1u;
        // [a, b): tuples of [lo, hi) whose layer-l number is `want`
        uint32_t a = lo[l], e = hi[l];
        while (a < e) {
            const uint32_t m = a + (e - a) / 2;
            if (tup[(size_t)m * L + l] < want)
                a = m + 1;
            else
                e = m;
        }
        uint32_t b = a;
        e = hi[l];
        while (b < e) {
            const uint32_t m = b + (e - b) / 2;
            if (tup[(size_t)m * L + l] <= want)
                b = m + 1;
            else
                e = m;
        }
        if (a == b) {
            ++dig[l];
            continue;
        }
        if ((uint32_t)l + 1 == L) {  // tuples are distinct: one bin
            if (WRITE)
                adj[w++] = a;
            ++cnt;
            sz += nstart[a + 1] - nstart[a];
            ++dig[l];
            continue;
        }
        ++l;
        lo[l] = a;
        hi[l] = b;
        dig[l] = 0;
    }
    if (!WRITE) {
        deg[v] = cnt;
        sumsz[v] = sz;
        ncalcs[v] = sz * (uint64_t)(nstart[v + 1] - nstart[v]);
    }
}

template <class T>
void build_coss_impl(IndexBuilder &B, const T *xs, uint32_t n, uint32_t d, double eps,
                     const msj_params &p, Index &out, LayerStats &st, hipStream_t s)
{
    const uint32_t nrp = p.rp_per_layer;
    const uint32_t L = std::min<uint32_t>(p.maxrp, MSJ_MAX_LAYERS);
    const uint32_t naxis = std::min(L, d);  // maxRP axes (tree.cu:278), j < d
    const uint32_t ncand = nrp + naxis;
    const uint32_t S = host_cvt_u32((double)(unsigned long long)n * p.sample_per);
    st = LayerStats{};
    st.ncand = ncand;
    st.layers = L;

    GlibcRand rng;
    rng.seed(p.seed);
    B.keys.ensure((size_t)ncand * n);
    B.maxc.ensure(d);
    B.cmm.ensure(2 * (size_t)ncand);
    B.cslots.ensure(4 * (size_t)ncand);
    B.cacc.ensure(ncand);
    B.cscore.ensure(ncand);
    B.cwalk.ensure((sizeof(CossWalk) / 4) * (size_t)ncand);
    B.prank.ensure(n);
    B.bnum.ensure((size_t)L * n);
    B.vals.ensure(n);
    B.u64scalar.ensure(1);
    B.runfirst.ensure((size_t)ncand + 1);

    // row 0 seeds the axis key bounds: int maxBin = ceil((x + 0.000001) / eps) + 10,
    // minBin = maxBin - 1 (tree.cu:281-282), compared as unsigned
    std::vector<T> x0(naxis);
    for (uint32_t a = 0; a < naxis; ++a)
        MSJ_HIP(hipMemcpyAsync(&x0[a], xs + (size_t)a * n, sizeof(T), hipMemcpyDeviceToHost, s));
    MSJ_HIP(hipStreamSynchronize(s));
    std::vector<uint32_t> seed_max(naxis);
    for (uint32_t a = 0; a < naxis; ++a)
        seed_max[a] = (uint32_t)host_cvt_i32(std::ceil(((double)x0[a] + 0.000001) / eps) + 10.0);

    LAUNCH(k_fill<uint32_t>, grid_for(n, 256), 256, B.prank.p, (size_t)n, 1u);
    RpHost rh(nrp, d);
    std::vector<double> sumsq(ncand, 0.0), score(ncand);
    std::vector<uint32_t> mm(2 * (size_t)ncand), ne_of(ncand);
    std::vector<uint64_t> count_of(ncand);
    std::vector<CossSlot> slots;
    std::vector<CossWalk> walks;
    std::vector<uint32_t> h_rf;
    std::vector<unsigned long long> h_acc;
    std::vector<double> h_walk;
    uint32_t prev_ne = 0;

    for (uint32_t cl = 0; cl < L; ++cl) {
        // ---- candidates and their keys (createRPArray, binningKernel, tree.cu:284)
        const double *rpT = layer_rps(B, xs, n, d, S, nrp, cl, L, rng, rh, s);
        launch_keys<T, true>(xs, n, d, rpT, nrp, naxis, eps, B.keys.p, nullptr, s);
        if (B.probe && B.probe->keys && B.probe->layer == cl) {
            MSJ_HIP(hipMemcpyAsync(B.probe->keys, B.keys.p, sizeof(uint32_t) * ncand * n,
                                   hipMemcpyDeviceToHost, s));
            MSJ_HIP(hipStreamSynchronize(s));
        }
        LAUNCH(k_minmax_init, 1, 64 * ((ncand + 63) / 64), B.cmm.p, ncand);
        hipLaunchKernelGGL(k_cand_minmax, dim3(std::min<unsigned>(grid_for(n, 256), 256), ncand),
                           dim3(256), 0, s, B.keys.p, n, B.cmm.p);
        MSJ_HIP(hipGetLastError());
        MSJ_HIP(hipMemcpyAsync(mm.data(), B.cmm.p, sizeof(uint32_t) * 2 * ncand,
                               hipMemcpyDeviceToHost, s));
        MSJ_HIP(hipStreamSynchronize(s));

        // ---- eligible candidates: skip = minBin - 2, nb = maxBin + 2 - skip
        for (uint32_t t = 0; t < nrp; ++t)
            sumsq[t] = 0.0;
        slots.clear();
        uint64_t maxcount = 1;
        for (uint32_t t = 0; t < ncand; ++t) {
            uint32_t lo = mm[2 * t], hi = mm[2 * t + 1];
            if (t >= nrp) {
                const uint32_t sm = seed_max[t - nrp];
                hi = std::max(hi, sm);
                lo = std::min(lo, sm - 1u);
            }
            const uint32_t skip = lo - 2u, nb = hi + 2u - skip;
            const uint64_t count = cl == 0 ? nb : (uint64_t)prev_ne * nb;
            score[t] = INFINITY;
            count_of[t] = count;
            if ((uint64_t)(hi - lo) + 4u > 0xFFFFFFFFull || count == 0 || count > 0xFFFFFFFFull)
                continue;
            slots.push_back(CossSlot{t, skip, nb, 0u});
            maxcount = std::max(maxcount, count);
        }
        const uint32_t nel = (uint32_t)slots.size();
        if (nel == 0)
            throw Error(MSJ_E_OVERFLOW, "tree/COSS layer " + std::to_string(cl) +
                                            ": no candidate fits in 2^32 bins");
        MSJ_HIP(hipMemcpyAsync(B.cslots.p, slots.data(), sizeof(CossSlot) * nel,
                               hipMemcpyHostToDevice, s));
        const int ab = bits_for(maxcount - 1);
        const uint64_t amask = ab >= 64 ? ~0ull : ((1ull << ab) - 1);
        // groups of slots whose (slot, address) keys fit the u32 run counters
        const uint32_t G = (uint32_t)std::max<uint64_t>(
            1, std::min<uint64_t>(nel, (uint64_t)0x7FFFFFFF / n));
        for (uint32_t j0 = 0; j0 < nel; j0 += G) {
            const uint32_t jc = std::min(G, nel - j0);
            const size_t m = (size_t)jc * n;
            B.ck.ensure(m);
            B.flags.ensure(m);
            B.runidx.ensure(m + 1);
            B.starts.ensure(m + 1);
            LAUNCH(k_coss_addr, grid_for(m, 256), 256, B.keys.p, B.prank.p,
                   (const CossSlot *)B.cslots.p + j0, jc, n, cl == 0 ? 1 : 0, ab, B.ck.p);
            B.sorter.sort(B.ck.p, nullptr, m, ab + bits_for(jc - 1), s);
            LAUNCH(k_run_flags, grid_for(m, 256), 256, B.ck.p, m, B.flags.p);
            B.scan.excl_u32_to_u32(B.flags.p, B.runidx.p, m, s);
            LAUNCH(k_run_starts, grid_for(m, 256), 256, B.flags.p, B.runidx.p, m, B.starts.p);
            LAUNCH(k_runfirst, 1, 64 * ((jc + 1 + 63) / 64), B.runidx.p, n, jc, B.runfirst.p);
            LAUNCH(k_coss_stats, jc, 256, B.runfirst.p, B.starts.p, n, B.cacc.p);
            h_rf.resize(jc + 1);
            h_acc.resize(jc);
            MSJ_HIP(hipMemcpyAsync(h_rf.data(), B.runfirst.p, sizeof(uint32_t) * (jc + 1),
                                   hipMemcpyDeviceToHost, s));
            MSJ_HIP(hipMemcpyAsync(h_acc.data(), B.cacc.p, sizeof(unsigned long long) * jc,
                                   hipMemcpyDeviceToHost, s));
            MSJ_HIP(hipStreamSynchronize(s));
            walks.clear();
            for (uint32_t j = 0; j < jc; ++j) {
                const uint32_t t = slots[j0 + j].t;
                const uint32_t ne = h_rf[j + 1] - h_rf[j];
                const uint32_t A = n / ne;  // numPoints / layerBinNonEmpty
                double s0 = sumsq[t];
                if (t >= nrp && (double)A >= (double)n / 3.0)
                    s0 = (double)((uint64_t)n * n);
                ne_of[t] = ne;
                // exact in integers below 2^53: then every partial sum of the
                // reference's FP64 loop is exact and its order does not matter
                const unsigned __int128 total =
                    (unsigned __int128)h_acc[j] +
                    (unsigned __int128)(count_of[t] - ne) * ((uint64_t)A * A) +
                    (unsigned __int128)(s0 < 9007199254740992.0 ? (uint64_t)s0 : 0);
                const bool exact = h_acc[j] != ~0ull && s0 < 9007199254740992.0 &&
                                   s0 == std::floor(s0) && total < ((unsigned __int128)1 << 53);
                if (exact)
                    score[t] = (double)(uint64_t)total;
                else
                    walks.push_back(CossWalk{j, A, count_of[t], s0});
            }
            if (!walks.empty()) {
                const uint32_t nw = (uint32_t)walks.size();
                MSJ_HIP(hipMemcpyAsync(B.cwalk.p, walks.data(), sizeof(CossWalk) * nw,
                                       hipMemcpyHostToDevice, s));
                LAUNCH(k_coss_walk, grid_for(nw, 64), 64, (const CossWalk *)B.cwalk.p, nw,
                       B.runfirst.p, B.starts.p, B.ck.p, amask, B.cscore.p);
                h_walk.resize(nw);
                MSJ_HIP(hipMemcpyAsync(h_walk.data(), B.cscore.p, sizeof(double) * nw,
                                       hipMemcpyDeviceToHost, s));
                MSJ_HIP(hipStreamSynchronize(s));
                for (uint32_t k = 0; k < nw; ++k)
                    score[slots[j0 + walks[k].j].t] = h_walk[k];
            }
            for (uint32_t j = 0; j < jc; ++j)
                sumsq[slots[j0 + j].t] = score[slots[j0 + j].t];
        }
        uint32_t best = 0;  // MINSQRS: the first smallest (tree.cu:400-404)
        for (uint32_t t = 1; t < ncand; ++t)
            if (score[best] > score[t])
                best = t;
        const CossSlot *bs = nullptr;
        for (const CossSlot &c : slots)
            if (c.t == best)
                bs = &c;
        if (!bs)
            throw Error(MSJ_E_OVERFLOW, "tree/COSS layer " + std::to_string(cl) +
                                            ": no candidate fits in 2^32 bins");

        // ---- apply the winner: ranks of its non-empty bins, bin numbers
        B.ck.ensure(n);
        B.flags.ensure(n);
        B.runidx.ensure((size_t)n + 1);
        LAUNCH(k_coss_best, grid_for(n, 256), 256, B.keys.p + (size_t)best * n, B.prank.p, n,
               cl == 0 ? 1 : 0, bs->skip, bs->nb, B.ck.p, B.vals.p, B.bnum.p + (size_t)cl * n);
        B.sorter.sort(B.ck.p, B.vals.p, n, bits_for(count_of[best] - 1), s);
        LAUNCH(k_run_flags, grid_for(n, 256), 256, B.ck.p, (size_t)n, B.flags.p);
        B.scan.excl_u32_to_u32(B.flags.p, B.runidx.p, n, s);
        LAUNCH(k_coss_rank, grid_for(n, 256), 256, B.flags.p, B.runidx.p, B.vals.p, n,
               B.prank.p);
        const uint32_t ne = d2h_scalar(B.runidx.p + n, s);
        if (ne != ne_of[best])
            throw Error(MSJ_E_HIP, "tree/COSS: inconsistent bin counts");
        prev_ne = ne;
        st.chosen[cl] = best;
        st.sse[cl] = score[best];
        st.nodes[cl] = ne;
        if (p.verbose)
            printf("Layer %u sumqrs: %f BinCount: %llu nonEmpty: %u (candidate %u)\n", cl,
                   score[best], (unsigned long long)count_of[best], ne, best);
    }

    // ---- points by final bin, stable by row (tree.cu:472-503)
    const uint32_t M = prev_ne;
    B.ck.ensure(n);
    LAUNCH(k_u32_to_u64, grid_for(n, 256), 256, B.prank.p, n, B.ck.p);
    LAUNCH(k_iota, grid_for(n, 256), 256, B.vals.p, n);
    B.sorter.sort(B.ck.p, B.vals.p, n, bits_for(M), s);
    swap_buf(out.perm, B.vals);
    LAUNCH(k_run_flags, grid_for(n, 256), 256, B.ck.p, (size_t)n, B.flags.p);
    B.scan.excl_u32_to_u32(B.flags.p, B.runidx.p, n, s);
    out.nstart.ensure((size_t)M + 1);
    LAUNCH(k_run_starts, grid_for(n, 256), 256, B.flags.p, B.runidx.p, (size_t)n, out.nstart.p);
    B.tup.ensure((size_t)M * L);
    LAUNCH(k_coss_tuples, grid_for(M, 256), 256, out.perm.p, out.nstart.p, B.bnum.p, n, M, L,
           B.tup.p);
    if (B.probe && B.probe->node_keys) {
        B.probe->tuples.resize((size_t)M * L);
        MSJ_HIP(hipMemcpyAsync(B.probe->tuples.data(), B.tup.p, sizeof(uint32_t) * M * L,
                               hipMemcpyDeviceToHost, s));
        MSJ_HIP(hipStreamSynchronize(s));
    }

    // ---- neighbouring bins (tree.cu:886-996)
    B.deg.ensure(M);
    out.sumsz.ensure(M);
    out.ncalcs.ensure(M);
    out.aoff.ensure((size_t)M + 1);
    LAUNCH(k_coss_adj<false>, grid_for(M, 256), 256, B.tup.p, M, L, out.nstart.p, B.deg.p,
           nullptr, nullptr, out.sumsz.p, out.ncalcs.p);
    B.tmp64.ensure((size_t)M + 1);
    B.scan.excl_u32_to_u64(B.deg.p, B.tmp64.p, M, s);
    const uint64_t edges = d2h_scalar(B.tmp64.p + M, s);
    if (edges >= 0xFFFFFFFFull)
        throw Error(MSJ_E_OVERFLOW, "node adjacency exceeds 2^32 entries");
    B.scan.excl_u32_to_u32(B.deg.p, out.aoff.p, M, s);
    out.adj.ensure(std::max<uint64_t>(edges, 1));
    LAUNCH(k_coss_adj<true>, grid_for(M, 256), 256, B.tup.p, M, L, out.nstart.p, nullptr,
           out.aoff.p, out.adj.p, nullptr, nullptr);
    reduce_sum_u64(out.ncalcs.p, M, B.u64scalar.p, s);
    const uint64_t calcs = d2h_scalar(B.u64scalar.p, s);
    out.M = M;
    out.edges = (uint32_t)edges;
    out.num_calcs = calcs;
    st.calcs[L - 1] = calcs;
}

}  // namespace

void column_sigma(const double *A, uint32_t rows, uint32_t cols, size_t rs, size_t cs,
                  double *sigma, hipStream_t s)
{
    if (cols)
        LAUNCH(k_col_sigma, cols, SIG_BLOCK, A, rows, rs, cs, sigma);
}

void build_single_node(uint32_t n, Index &out, LayerStats &st, hipStream_t s)
{
    out.perm.ensure(n);
    out.nstart.ensure(2);
    out.aoff.ensure(2);
    out.adj.ensure(1);
    out.sumsz.ensure(1);
    out.ncalcs.ensure(1);
    LAUNCH(k_iota, grid_for(n, 256), 256, out.perm.p, n);
    LAUNCH(k_single_node, 1, 64, n, out.nstart.p, out.aoff.p, out.adj.p, out.sumsz.p,
           out.ncalcs.p);
    out.M = 1;
    out.edges = 1;
    out.num_calcs = (uint64_t)n * n;
    st = LayerStats{};
}

void IndexBuilder::build(const void *xs, bool f32, uint32_t n, uint32_t d, double eps,
                         const msj_params &p, Index &out, LayerStats &st, hipStream_t s)
{
    if (p.index == MSJ_INDEX_COSS) {
        if (f32)
            build_coss_impl<float>(*this, (const float *)xs, n, d, eps, p, out, st, s);
        else
            build_coss_impl<double>(*this, (const double *)xs, n, d, eps, p, out, st, s);
    } else if (f32)
        build_impl<float>(*this, (const float *)xs, n, d, eps, p, out, st, s);
    else
        build_impl<double>(*this, (const double *)xs, n, d, eps, p, out, st, s);
}

}  // namespace msj
