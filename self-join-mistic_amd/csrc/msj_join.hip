// msj_join.hip -- plain FP64 distance-and-compare over the node net
// (subsystem c) and the ordered result compaction (subsystem d).
//
// Reference: kernel.cu:992-1046 nodeByPoint5 and kernel.cu:833-865
// cachedDistanceCheck (ILP = 4); launcher.cu:3418-3506 batch loop.
//
// Symmetry.  The test is bit-symmetric: q - c = -(c - q) exactly, so both
// orders of a pair give the same sum and the same decision.  Each unordered
// candidate pair is therefore evaluated once -- node X against the nodes
// Y >= X of its neighbour list, and inside X only pairs p <= c -- and an
// accepted pair is emitted in both orders (the self pair once).
//
// Work decomposition.  A work item is one wave: a range of X's points (the
// queries) against a range of X's upper candidate list (the points of the
// neighbour nodes Y >= X, concatenated).  The wave walks that range 64
// candidates at a time, one candidate per lane with its coordinates in VGPRs
// (rows of up to 32 dimensions whole, wider rows 16 dimensions per slab),
// and loops over the queries, whose rows it stages in its own LDS slice: a
// query's coordinates are wave-uniform LDS addresses (broadcast reads).
// Items are ordered candidate-chunk-major, so the query groups that share a
// candidate chunk run side by side in one workgroup (L1/L2 reuse).  Rows in
// join order: node order, and inside a node a space-filling-curve order
// (JoinRunner::order).
//
// Decision: exactly the reference's sequence per 4-dimension chunk,
//   s = fma(d0, d0, s); s = s + d1*d1; s = s + d2*d2; s = s + d3*d3;
//   reject once s > eps^2,
// with this file compiled without floating-point contraction.  The wave
// leaves the chunk loop as soon as no lane is still undecided.
//
// Emission: accepted keys (row << qbits | neighbour row) go to a per-wave
// LDS buffer; the wave reserves global space with one atomic per flush and
// copies the buffer out with coalesced stores.  One radix sort then yields
// every row's neighbours ascending, the counts and the CSR offsets.
#include <algorithm>
#include <cmath>
#include <vector>
#include <cstdio>

#include "msj_join_items.h"

#pragma clang fp contract(off)

namespace msj {

namespace {

constexpr uint64_t EVALS_PER_ITEM = 262144;  // lane evaluations per work item (default)
// ds > 32: HQ (24) queries x 1,024 candidates.  The query groups of a
// candidate chunk are consecutive items, dealt out four to a workgroup, so
// the workgroups on one XCD that share a chunk (1,024 rows, 736 KB at d =
// 90) run side by side and find it in their L2 (515,345 x 90: 945 -> 808 ms
// with chunks of 1,024 instead of 16,384 rows, 788 ms with 24 queries per
// item instead of 16).
constexpr uint64_t EVALS_PER_ITEM_HD = 1024ull * HQ;

__device__ inline double wave_min(double v)
{
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) {
        double t = __shfl_xor(v, o, WAVE);
        v = t < v ? t : v;
    }
    return v;
}
__device__ inline double wave_max(double v)
{
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) {
        double t = __shfl_xor(v, o, WAVE);
        v = t > v ? t : v;
    }
    return v;
}

__device__ inline void add_dims(uint64_t dims, unsigned long long *work)
{
    unsigned long long v = dims;
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1)
        v += __shfl_down(v, o, WAVE);
    if ((threadIdx.x & (WAVE - 1)) == 0 && v)
        atomicAdd(work, v);
}

// One wave per work item, grid-stride over the items (a launch stays under
// 2^32 work-items); DS = ds <= 32.
template <int DS, bool COUNT>
__global__ __launch_bounds__(256) void k_join_sym(JobView J, unsigned long long *__restrict__ work)
{
    __shared__ __align__(16) double qsh[WPB][LDS_WAVE_BYTES / 8];
    __shared__ unsigned long long hsh[WPB][HITS];
    const uint32_t lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x / WAVE;
    WaveHits H(hsh[wv], lane);
    uint64_t dims = 0;
    const uint64_t stride = (uint64_t)gridDim.x * WPB;
    for (uint64_t it = (uint64_t)blockIdx.x * WPB + wv; it < J.nitems; it += stride) {
        // readfirstlane keeps the item and the loop counters in SGPRs
        const JoinItem w = J.items[__builtin_amdgcn_readfirstlane((uint32_t)it)];
        item_sym<DS, COUNT>(J, w, qsh[wv], H, dims);
    }
    if (H.fill)
        H.flush(J);
    if (COUNT)
        add_dims(dims, work);
}

// The same for ds > 32 (slab-major variant); 128 VGPRs, four waves per SIMD.
template <bool COUNT>
__global__ __launch_bounds__(256, 4) void k_join_hd(JobView J,
                                                    unsigned long long *__restrict__ work)
{
    __shared__ __align__(16) double qsh[WPB][HQ * HSL];
    __shared__ unsigned long long hsh[WPB][HITS];
    const uint32_t lane = threadIdx.x & (WAVE - 1), wv = threadIdx.x / WAVE;
    WaveHits H(hsh[wv], lane);
    uint64_t dims = 0;
    const uint64_t stride = (uint64_t)gridDim.x * WPB;
    for (uint64_t it = (uint64_t)blockIdx.x * WPB + wv; it < J.nitems; it += stride) {
        const JoinItem w = J.items[__builtin_amdgcn_readfirstlane((uint32_t)it)];
        item_hd<COUNT>(J, w, qsh[wv], H, dims);
    }
    if (H.fill)
        H.flush(J);
    if (COUNT)
        add_dims(dims, work);
}

// Upper neighbour lists: entries Y >= X of X's adjacency, and their sizes;
// trow[x] = sum over those entries of |X| + |Y| (rows of X's candidate tiles).
__global__ __launch_bounds__(256) void k_upper_count(const uint32_t *__restrict__ nstart,
                                                     const uint32_t *__restrict__ aoff,
                                                     const uint32_t *__restrict__ adj, uint32_t M,
                                                     uint32_t *__restrict__ cnt,
                                                     uint64_t *__restrict__ len,
                                                     uint64_t *__restrict__ trow)
{
    uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= M)
        return;
    uint32_t k = 0;
    uint64_t L = 0;
    for (uint32_t e = aoff[x]; e < aoff[x + 1]; ++e) {
        uint32_t y = adj[e];
        if (y >= x) {
            ++k;
            L += nstart[y + 1] - nstart[y];
        }
    }
    cnt[x] = k;
    len[x] = L;
    trow[x] = (uint64_t)k * (nstart[x + 1] - nstart[x]) + L;
}

// ... and, per entry, the offset of its first point inside the node's
// concatenated upper candidate list (a work item finds its first entry by
// binary search).
__global__ __launch_bounds__(256) void k_upper_write(const uint32_t *__restrict__ nstart,
                                                     const uint32_t *__restrict__ aoff,
                                                     const uint32_t *__restrict__ adj, uint32_t M,
                                                     const uint32_t *__restrict__ uoff,
                                                     uint32_t *__restrict__ uadj,
                                                     uint32_t *__restrict__ ucum)
{
    uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= M)
        return;
    uint32_t o = uoff[x], c = 0;
    for (uint32_t e = aoff[x]; e < aoff[x + 1]; ++e)
        if (adj[e] >= x) {
            ucum[o] = c;
            c += nstart[adj[e] + 1] - nstart[adj[e]];
            uadj[o++] = adj[e];
        }
}

// Items per node: query ranges of up to qmax points x candidate ranges sized
// so one item is about `epi` lane evaluations (multiples of 64).
__device__ inline void item_shape(uint32_t sz, uint64_t L, uint32_t qmax, uint64_t epi,
                                  uint32_t *qper, uint32_t *nq, uint32_t *cper, uint32_t *nc)
{
    *qper = min(qmax, sz);
    *nq = (sz + *qper - 1) / *qper;
    uint64_t cp = epi / *qper;
    cp = (cp + WAVE - 1) / WAVE * WAVE;
    *cper = (uint32_t)cp;
    uint64_t k = (L + cp - 1) / cp;
    *nc = (uint32_t)(k ? k : 1);
}

__global__ __launch_bounds__(256) void k_item_count(const uint32_t *__restrict__ nstart,
                                                    const uint64_t *__restrict__ ulen, uint32_t M,
                                                    uint32_t qmax, uint64_t epi,
                                                    uint32_t *__restrict__ cnt)
{
    uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= M)
        return;
    uint32_t qper, nq, cper, nc;
    item_shape(nstart[x + 1] - nstart[x], ulen[x], qmax, epi, &qper, &nq, &cper, &nc);
    cnt[x] = nq * nc;
}

__global__ __launch_bounds__(256) void k_item_fill(const uint32_t *__restrict__ nstart,
                                                   const uint64_t *__restrict__ ulen,
                                                   const uint32_t *__restrict__ itoff, uint32_t M,
                                                   uint32_t qmax, uint64_t epi,
                                                   uint32_t nitems, JoinItem *__restrict__ items)
{
    uint32_t k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nitems)
        return;
    uint32_t lo = 0, hi = M;  // node x with itoff[x] <= k < itoff[x+1]
    while (hi - lo > 1) {
        uint32_t mid = lo + (hi - lo) / 2;
        if (itoff[mid] <= k)
            lo = mid;
        else
            hi = mid;
    }
    const uint32_t x = lo;
    const uint32_t sz = nstart[x + 1] - nstart[x];
    const uint64_t L = ulen[x];
    uint32_t qper, nq, cper, nc;
    item_shape(sz, L, qmax, epi, &qper, &nq, &cper, &nc);
    const uint32_t local = k - itoff[x];
    // candidate-chunk-major: the query groups of one candidate chunk are
    // consecutive items (run side by side, they share the chunk in cache)
    const uint32_t ck = local / nq, qi = local % nq;
    JoinItem w;
    w.x = x;
    w.qbeg = qi * qper;
    w.qend = min(sz, (qi + 1) * qper);
    w.cbeg = (uint32_t)std::min<uint64_t>((uint64_t)ck * cper, L);
    w.cend = (uint32_t)std::min<uint64_t>((uint64_t)(ck + 1) * cper, L);
    items[k] = w;
}

// xr[i*ds + j] = xs[j*n + perm[i]] widened to FP64 (exact for FP32 inputs),
// zero for the padding columns d <= j < ds.
template <class T>
__global__ __launch_bounds__(256) void k_gather_rows(const T *__restrict__ xs,
                                                     double *__restrict__ xr,
                                                     const uint32_t *__restrict__ perm, uint32_t n,
                                                     uint32_t d, uint32_t ds)
{
    const size_t total = (size_t)n * ds, stride = (size_t)gridDim.x * blockDim.x;
    for (size_t idx = (size_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total; idx += stride) {
        uint32_t i = (uint32_t)(idx / ds), j = (uint32_t)(idx - (size_t)i * ds);
        xr[idx] = j < d ? (double)xs[(size_t)j * n + perm[i]] : 0.0;
    }
}

// The same from rows still in the uploaded row-major layout (rows[r*d + j]):
// one wave copies one row per step, reading it whole.
template <class T>
__global__ __launch_bounds__(256) void k_gather_rows_aos(const T *__restrict__ rows,
                                                         double *__restrict__ xr,
                                                         const uint32_t *__restrict__ perm,
                                                         uint32_t n, uint32_t d, uint32_t ds)
{
    const uint32_t lane = threadIdx.x & (WAVE - 1);
    const uint32_t wstride = gridDim.x * (blockDim.x / WAVE);
    for (uint32_t i = (blockIdx.x * blockDim.x + threadIdx.x) / WAVE; i < n; i += wstride) {
        const T *src = rows + (size_t)perm[i] * d;
        double *dst = xr + (size_t)i * ds;
        for (uint32_t j = lane; j < ds; j += WAVE)
            dst[j] = j < d ? (double)src[j] : 0.0;
    }
}

// Sorted keys (row << qbits | neighbour): every key's neighbour id into
// nbrs, and at the start of each row's run its position into first[row].
__global__ __launch_bounds__(256) void k_rows_first(const uint64_t *__restrict__ pairs, size_t R,
                                                    int qbits, uint64_t qmask,
                                                    uint64_t *__restrict__ first,
                                                    uint32_t *__restrict__ nbrs)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < R; i += stride) {
        const uint64_t v = pairs[i];
        const uint64_t row = v >> qbits;
        nbrs[i] = (uint32_t)(v & qmask);
        if (i == 0 || (pairs[i - 1] >> qbits) != row)
            first[row] = i;
    }
}

// At the end of each row's run: its length.
__global__ __launch_bounds__(256) void k_rows_count(const uint64_t *__restrict__ pairs, size_t R,
                                                    int qbits, const uint64_t *__restrict__ first,
                                                    uint32_t *__restrict__ counts)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < R; i += stride) {
        const uint64_t row = pairs[i] >> qbits;
        if (i + 1 == R || (pairs[i + 1] >> qbits) != row)
            counts[row] = (uint32_t)(i + 1 - first[row]);
    }
}

// Per-column minimum and maximum of the first columns: blockIdx.y is the
// column, the blocks along x reduce row ranges into part[j][b] = (lo, hi);
// k_col_range_final reduces those (a minimum or maximum does not depend on
// the order it is taken in).
constexpr uint32_t CR_BLOCKS = 64;

template <class T>
__global__ __launch_bounds__(256) void k_col_range(const T *__restrict__ xs, uint32_t n,
                                                   double *__restrict__ part)
{
    __shared__ double lo_s[256 / WAVE], hi_s[256 / WAVE];
    const uint32_t j = blockIdx.y, b = blockIdx.x;
    double lo = __builtin_inf(), hi = -__builtin_inf();
    for (uint32_t i = b * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        double v = (double)xs[(size_t)j * n + i];
        lo = v < lo ? v : lo;
        hi = v > hi ? v : hi;
    }
    lo = wave_min(lo);
    hi = wave_max(hi);
    if ((threadIdx.x & (WAVE - 1)) == 0) {
        lo_s[threadIdx.x / WAVE] = lo;
        hi_s[threadIdx.x / WAVE] = hi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int w = 1; w < 256 / WAVE; ++w) {
            lo = lo_s[w] < lo ? lo_s[w] : lo;
            hi = hi_s[w] > hi ? hi_s[w] : hi;
        }
        part[2 * ((size_t)j * gridDim.x + b)] = lo;
        part[2 * ((size_t)j * gridDim.x + b) + 1] = hi;
    }
}

__global__ void k_col_range_final(const double *__restrict__ part, uint32_t nb,
                                  double *__restrict__ mm)
{
    const uint32_t j = blockIdx.x;
    if (threadIdx.x != 0)
        return;
    double lo = __builtin_inf(), hi = -__builtin_inf();
    for (uint32_t b = 0; b < nb; ++b) {
        const double l = part[2 * ((size_t)j * nb + b)], h = part[2 * ((size_t)j * nb + b) + 1];
        lo = l < lo ? l : lo;
        hi = h > hi ? h : hi;
    }
    mm[2 * j] = lo;
    mm[2 * j + 1] = hi;
}

__device__ inline uint32_t spread3(uint32_t v)  // 10 bits -> every third bit
{
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// Key (node, Morton code of the first min(3, d) columns quantised to 10
// bits) for every node-order position: sorting it orders points inside each
// node along a space-filling curve.  Values: the node-order row ids.
template <class T>
__global__ __launch_bounds__(256) void k_morton_keys(const T *__restrict__ xs, uint32_t n,
                                                     uint32_t km, const double *__restrict__ mm,
                                                     const uint32_t *__restrict__ nstart,
                                                     uint32_t M, const uint32_t *__restrict__ perm,
                                                     uint64_t *__restrict__ keys,
                                                     uint32_t *__restrict__ vals)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    uint32_t lo = 0, hi = M;  // node holding position i
    while (hi - lo > 1) {
        uint32_t mid = lo + (hi - lo) / 2;
        if (nstart[mid] <= i)
            lo = mid;
        else
            hi = mid;
    }
    const uint32_t row = perm[i];
    uint32_t code = 0;
    for (uint32_t k = 0; k < km; ++k) {
        const double a = mm[2 * k], b = mm[2 * k + 1];
        const double v = (double)xs[(size_t)k * n + row];
        double q = b > a ? (v - a) / (b - a) * 1023.0 : 0.0;
        q = q > 0.0 ? (q < 1023.0 ? q : 1023.0) : 0.0;  // also maps NaN to 0
        code |= spread3((uint32_t)q) << k;
    }
    keys[i] = ((uint64_t)lo << 30) | code;
    vals[i] = row;
}

// Principal directions of the rows, for the join order of wide rows.
constexpr uint32_t PCA_SAMPLE = 4096;  // rows of the covariance sample
constexpr uint32_t PCA_DIMS = 128;     // leading dimensions it covers
constexpr uint32_t PCA_K = 6;          // directions computed (at most)

// S[j * ns + i] = column j < dp of sample row i = i * n / ns.
template <class T>
__global__ __launch_bounds__(256) void k_pca_sample(const T *__restrict__ xs, uint32_t n,
                                                    uint32_t ns, uint32_t dp,
                                                    double *__restrict__ S)
{
    const uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ns * dp)
        return;
    const uint32_t j = e / ns, i = e - j * ns;
    const uint32_t row = (uint32_t)((uint64_t)i * n / ns);
    S[e] = (double)xs[(size_t)j * n + row];
}

// mu[j] = mean of sample column j (one wave per column).
__global__ __launch_bounds__(256) void k_pca_mean(const double *__restrict__ S, uint32_t ns,
                                                  uint32_t dp, double *__restrict__ mu)
{
    const uint32_t j = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    const uint32_t lane = threadIdx.x & (WAVE - 1);
    if (j >= dp)
        return;
    double a = 0.0;
    for (uint32_t i = lane; i < ns; i += WAVE)
        a += S[(size_t)j * ns + i];
    for (int o = WAVE / 2; o > 0; o >>= 1)
        a += __shfl_xor(a, o, WAVE);
    if (lane == 0)
        mu[j] = a / ns;
}

// C[a * dp + b] (a <= b) = sum over the sample of the centred products (one
// wave per pair).
__global__ __launch_bounds__(256) void k_pca_cov(const double *__restrict__ S, uint32_t ns,
                                                 uint32_t dp, const double *__restrict__ mu,
                                                 double *__restrict__ C)
{
    const uint32_t w = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE;
    const uint32_t lane = threadIdx.x & (WAVE - 1);
    if (w >= dp * dp)
        return;
    const uint32_t a = w / dp, b = w - a * dp;
    if (b < a)
        return;
    double acc = 0.0;
    for (uint32_t i = lane; i < ns; i += WAVE)
        acc += (S[(size_t)a * ns + i] - mu[a]) * (S[(size_t)b * ns + i] - mu[b]);
    for (int o = WAVE / 2; o > 0; o >>= 1)
        acc += __shfl_xor(acc, o, WAVE);
    if (lane == 0)
        C[w] = acc;
}

// P[k * n + row] = (x_row - mu) . V_k over the first dp columns, k < PCA_K.
template <class T>
__global__ __launch_bounds__(256) void k_pca_proj(const T *__restrict__ xs, uint32_t n,
                                                  uint32_t dp, const double *__restrict__ mu,
                                                  const double *__restrict__ V,
                                                  double *__restrict__ P)
{
    const uint32_t row = blockIdx.x * blockDim.x + threadIdx.x;
    if (row >= n)
        return;
    double p[PCA_K] = {};
    for (uint32_t j = 0; j < dp; ++j) {
        const double v = (double)xs[(size_t)j * n + row] - mu[j];
#pragma unroll
        for (uint32_t k = 0; k < PCA_K; ++k)
            p[k] += v * V[k * dp + j];
    }
#pragma unroll
    for (uint32_t k = 0; k < PCA_K; ++k)
        P[(size_t)k * n + row] = p[k];
}

// Top K (<= PCA_K) eigenvectors of the symmetric dp x dp matrix C (upper
// triangle filled) by power iteration with deflation, a fixed 32 steps per
// direction: the directions only order the rows, so they need not be exact,
// only the same from run to run.  False if C is degenerate.
bool top_directions(std::vector<double> &C, uint32_t dp, uint32_t K, std::vector<double> &V)
{
    for (uint32_t a = 0; a < dp; ++a)
        for (uint32_t b = 0; b < a; ++b)
            C[(size_t)a * dp + b] = C[(size_t)b * dp + a];
    V.assign((size_t)PCA_K * dp, 0.0);
    std::vector<double> v(dp), w(dp);
    for (uint32_t k = 0; k < K; ++k) {
        for (uint32_t j = 0; j < dp; ++j)
            v[j] = 1.0 + 0.001 * j + (j == k ? 1.0 : 0.0);
        double lam = 0.0;
        for (int it = 0; it < 32; ++it) {
            double nn = 0.0;
            for (uint32_t a = 0; a < dp; ++a) {
                const double *c = &C[(size_t)a * dp];
                double t0 = 0.0, t1 = 0.0, t2 = 0.0, t3 = 0.0;
                uint32_t b = 0;
                for (; b + 4 <= dp; b += 4) {
                    t0 += c[b] * v[b];
                    t1 += c[b + 1] * v[b + 1];
                    t2 += c[b + 2] * v[b + 2];
                    t3 += c[b + 3] * v[b + 3];
                }
                for (; b < dp; ++b)
                    t0 += c[b] * v[b];
                const double t = (t0 + t1) + (t2 + t3);
                w[a] = t;
                nn += t * t;
            }
            nn = std::sqrt(nn);
            if (!(nn > 0.0) || !std::isfinite(nn))
                return false;
            lam = nn;
            for (uint32_t j = 0; j < dp; ++j)
                v[j] = w[j] / nn;
        }
        for (uint32_t j = 0; j < dp; ++j)
            V[(size_t)k * dp + j] = v[j];
        for (uint32_t a = 0; a < dp; ++a)  // deflate
            for (uint32_t b = 0; b < dp; ++b)
                C[(size_t)a * dp + b] -= lam * v[a] * v[b];
    }
    return true;
}

// As k_morton_keys for K coordinates of B = 30 / K bits each, in K columns
// of xs (double, SoA).
template <uint32_t K, class T>
__global__ __launch_bounds__(256) void k_morton_keys_k(const T *__restrict__ xs, uint32_t n,
                                                       const double *__restrict__ mm,
                                                       const uint32_t *__restrict__ nstart,
                                                       uint32_t M,
                                                       const uint32_t *__restrict__ perm,
                                                       uint64_t *__restrict__ keys,
                                                       uint32_t *__restrict__ vals)
{
    constexpr uint32_t B = 30 / K;
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    uint32_t lo = 0, hi = M;  // node holding position i
    while (hi - lo > 1) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (nstart[mid] <= i)
            lo = mid;
        else
            hi = mid;
    }
    const uint32_t row = perm[i];
    const double top = (double)((1u << B) - 1);
    uint32_t code = 0;
#pragma unroll
    for (uint32_t k = 0; k < K; ++k) {
        const double a = mm[2 * k], b = mm[2 * k + 1];
        const double v = (double)xs[(size_t)k * n + row];
        double t = b > a ? (v - a) / (b - a) * top : 0.0;
        t = t > 0.0 ? (t < top ? t : top) : 0.0;  // also maps NaN to 0
        const uint32_t q = (uint32_t)t;
#pragma unroll
        for (uint32_t bb = 0; bb < B; ++bb)
            code |= ((q >> bb) & 1u) << (bb * K + k);
    }
    keys[i] = ((uint64_t)lo << 30) | code;
    vals[i] = row;
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
        ::msj::check_grid((grid), (block));                                              \
        hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, s, __VA_ARGS__);            \
        MSJ_HIP(hipGetLastError());                                                      \
    } while (0)

template <int DS>
void launch_sym(const JobView &V, unsigned long long *work, hipStream_t s)
{
    if (work)
        LAUNCH((k_join_sym<DS, true>), grid_stride(V.nitems, WPB), WAVE * WPB, V, work);
    else
        LAUNCH((k_join_sym<DS, false>), grid_stride(V.nitems, WPB), WAVE * WPB, V, work);
}

void dispatch_plain(const JobView &V, unsigned long long *work, hipStream_t s)
{
    switch (V.ds) {
    case 4: launch_sym<4>(V, work, s); return;
    case 8: launch_sym<8>(V, work, s); return;
    case 12: launch_sym<12>(V, work, s); return;
    case 16: launch_sym<16>(V, work, s); return;
    case 20: launch_sym<20>(V, work, s); return;
    case 24: launch_sym<24>(V, work, s); return;
    case 28: launch_sym<28>(V, work, s); return;
    case 32: launch_sym<32>(V, work, s); return;
    default:
        if (work)
            LAUNCH(k_join_hd<true>, grid_stride(V.nitems, WPB), WAVE * WPB, V, work);
        else
            LAUNCH(k_join_hd<false>, grid_stride(V.nitems, WPB), WAVE * WPB, V, work);
    }
}

}  // namespace

void gather_node_rows(const void *xs, const void *rows, double *xr, bool f32,
                      const uint32_t *perm, uint32_t n, uint32_t d, hipStream_t s)
{
    const uint32_t ds = padded_dim(d);
    if (rows) {
        const unsigned g = grid_stride((size_t)n * WAVE, 256);
        if (f32)
            LAUNCH(k_gather_rows_aos<float>, g, 256, (const float *)rows, xr, perm, n, d, ds);
        else
            LAUNCH(k_gather_rows_aos<double>, g, 256, (const double *)rows, xr, perm, n, d, ds);
        return;
    }
    size_t m = (size_t)n * ds;
    if (f32)
        LAUNCH(k_gather_rows<float>, grid_stride(m, 256), 256, (const float *)xs, xr, perm, n, d,
               ds);
    else
        LAUNCH(k_gather_rows<double>, grid_stride(m, 256), 256, (const double *)xs, xr, perm, n,
               d, ds);
}

void JoinRunner::order(const Index &ix, const void *xs, bool f32, uint32_t n, uint32_t d,
                       hipStream_t s)
{
    const uint32_t km = d < 3 ? d : 3u;
    mm.ensure(2 * PCA_K + 2 * PCA_K * CR_BLOCKS);  // the result, then the partial ranges
    const uint32_t crb = std::max<uint32_t>(1u, std::min<uint32_t>(CR_BLOCKS, (n + 255) / 256));
    jperm.ensure(n);
    jkeys.ensure(n);
    // Wide rows (padded width > 32): order each node along the Morton curve
    // of the rows' projections on the leading principal directions (of a
    // sample, over the first 128 columns): three up to d = 128, six above
    // (the 256-row tiles of the filter).  Rows that sit together in that
    // order are close in the dominant directions, so the pairs that one wave
    // or one tile evaluates reach their decisions after similar numbers of
    // dimensions.  Only the evaluation order changes.
    if (padded_dim(d) > 32 && n >= 2) {
        const uint32_t ns = std::min<uint32_t>(n, PCA_SAMPLE);
        const uint32_t dp = std::min<uint32_t>(d, PCA_DIMS);
        const size_t oS = 0, oMu = oS + (size_t)ns * dp, oC = oMu + dp, oV = oC + (size_t)dp * dp,
                     oP = oV + (size_t)PCA_K * dp;
        pca.ensure(oP + (size_t)PCA_K * n);
        if (f32)
            LAUNCH(k_pca_sample<float>, grid_for((size_t)ns * dp, 256), 256, (const float *)xs, n,
                   ns, dp, pca.p + oS);
        else
            LAUNCH(k_pca_sample<double>, grid_for((size_t)ns * dp, 256), 256, (const double *)xs,
                   n, ns, dp, pca.p + oS);
        LAUNCH(k_pca_mean, grid_for(dp, 4), 256, pca.p + oS, ns, dp, pca.p + oMu);
        LAUNCH(k_zero<double>, grid_for((size_t)dp * dp, 256), 256, pca.p + oC, (size_t)dp * dp);
        LAUNCH(k_pca_cov, grid_for((size_t)dp * dp, 4), 256, pca.p + oS, ns, dp, pca.p + oMu,
               pca.p + oC);
        std::vector<double> hC((size_t)dp * dp), hV;
        MSJ_HIP(hipMemcpyAsync(hC.data(), pca.p + oC, sizeof(double) * hC.size(),
                               hipMemcpyDeviceToHost, s));
        MSJ_HIP(hipStreamSynchronize(s));
        const uint32_t K = d > 128 ? 6u : 3u;
        if (top_directions(hC, dp, K, hV)) {
            MSJ_HIP(hipMemcpyAsync(pca.p + oV, hV.data(), sizeof(double) * hV.size(),
                                   hipMemcpyHostToDevice, s));
            if (f32)
                LAUNCH(k_pca_proj<float>, grid_for(n, 256), 256, (const float *)xs, n, dp,
                       pca.p + oMu, pca.p + oV, pca.p + oP);
            else
                LAUNCH(k_pca_proj<double>, grid_for(n, 256), 256, (const double *)xs, n, dp,
                       pca.p + oMu, pca.p + oV, pca.p + oP);
            hipLaunchKernelGGL(k_col_range<double>, dim3(crb, K), dim3(256), 0, s,
                               (const double *)(pca.p + oP), n, mm.p + 2 * PCA_K);
            MSJ_HIP(hipGetLastError());
            LAUNCH(k_col_range_final, K, 64, mm.p + 2 * PCA_K, crb, mm.p);
            if (K == 6)
                LAUNCH((k_morton_keys_k<6, double>), grid_for(n, 256), 256, (const double *)(pca.p + oP),
                       n, mm.p, ix.nstart.p, ix.M, ix.perm.p, jkeys.p, jperm.p);
            else
                LAUNCH((k_morton_keys_k<3, double>), grid_for(n, 256), 256, (const double *)(pca.p + oP),
                       n, mm.p, ix.nstart.p, ix.M, ix.perm.p, jkeys.p, jperm.p);
            sorter.sort(jkeys.p, jperm.p, n, bits_for(ix.M ? ix.M - 1 : 0) + 30, s);
            MSJ_HIP(hipStreamSynchronize(s));  // hV is read by the copy above
            return;
        }
    }
    // rows of 9 to 32 dimensions: the first six columns (five bits each)
    if (padded_dim(d) > 8) {
        if (f32) {
            hipLaunchKernelGGL(k_col_range<float>, dim3(crb, 6), dim3(256), 0, s,
                               (const float *)xs, n, mm.p + 2 * PCA_K);
            MSJ_HIP(hipGetLastError());
            LAUNCH(k_col_range_final, 6, 64, mm.p + 2 * PCA_K, crb, mm.p);
            LAUNCH((k_morton_keys_k<6, float>), grid_for(n, 256), 256, (const float *)xs, n,
                   mm.p, ix.nstart.p, ix.M, ix.perm.p, jkeys.p, jperm.p);
        } else {
            hipLaunchKernelGGL(k_col_range<double>, dim3(crb, 6), dim3(256), 0, s,
                               (const double *)xs, n, mm.p + 2 * PCA_K);
            MSJ_HIP(hipGetLastError());
            LAUNCH(k_col_range_final, 6, 64, mm.p + 2 * PCA_K, crb, mm.p);
            LAUNCH((k_morton_keys_k<6, double>), grid_for(n, 256), 256, (const double *)xs, n,
                   mm.p, ix.nstart.p, ix.M, ix.perm.p, jkeys.p, jperm.p);
        }
        sorter.sort(jkeys.p, jperm.p, n, bits_for(ix.M ? ix.M - 1 : 0) + 30, s);
        return;
    }
    if (f32) {
        hipLaunchKernelGGL(k_col_range<float>, dim3(crb, km), dim3(256), 0, s, (const float *)xs,
                           n, mm.p + 6);
        MSJ_HIP(hipGetLastError());
        LAUNCH(k_col_range_final, km, 64, mm.p + 6, crb, mm.p);
        LAUNCH(k_morton_keys<float>, grid_for(n, 256), 256, (const float *)xs, n, km, mm.p,
               ix.nstart.p, ix.M, ix.perm.p, jkeys.p, jperm.p);
    } else {
        hipLaunchKernelGGL(k_col_range<double>, dim3(crb, km), dim3(256), 0, s,
                           (const double *)xs, n, mm.p + 6);
        MSJ_HIP(hipGetLastError());
        LAUNCH(k_col_range_final, km, 64, mm.p + 6, crb, mm.p);
        LAUNCH(k_morton_keys<double>, grid_for(n, 256), 256, (const double *)xs, n, km, mm.p,
               ix.nstart.p, ix.M, ix.perm.p, jkeys.p, jperm.p);
    }
    sorter.sort(jkeys.p, jperm.p, n, bits_for(ix.M ? ix.M - 1 : 0) + 30, s);
}

JobView JoinRunner::view(const Index &ix, const double *xr, uint32_t n, uint32_t d,
                         double eps) const
{
    JobView V;
    V.xr = xr;
    V.d = d;
    V.ds = padded_dim(d);
    V.eps2 = eps * eps;  // launcher.cu:3200
    V.nstart = ix.nstart.p;
    V.uoff = uoff.p;
    V.uadj = uadj.p;
    V.ucum = ucum.p;
    V.perm = jperm.p;
    V.items = items.p;
    V.nitems = nitems;
    V.qbits = std::max(1, bits_for(n ? n - 1 : 0));
    V.out = (unsigned long long *)pairs.p;
    V.counter = counter.p;
    V.cap = pairs.cap;
    return V;
}

void JoinRunner::plan(const Index &ix, uint32_t n, uint32_t d, hipStream_t s, bool plain_items,
                      uint64_t evals_per_item)
{
    const uint32_t M = ix.M;
    const uint32_t ds = padded_dim(d);
    const uint32_t qmax = item_queries(ds);
    const uint64_t epi =
        evals_per_item ? evals_per_item : (ds > 32 ? EVALS_PER_ITEM_HD : EVALS_PER_ITEM);
    itcnt.ensure(M);
    itoff.ensure((size_t)M + 1);
    ulen.ensure(M);
    uoff.ensure((size_t)M + 1);
    tot64.ensure((size_t)M + 1);
    trow.ensure(M);
    tsum.ensure(1);
    LAUNCH(k_upper_count, grid_for(M, 256), 256, ix.nstart.p, ix.aoff.p, ix.adj.p, M, itcnt.p,
           ulen.p, trow.p);
    reduce_sum_u64(trow.p, M, tsum.p, s);
    scan.excl_u32_to_u32(itcnt.p, uoff.p, M, s);
    uint32_t ue = 0;
    unsigned long long trows = 0;
    MSJ_HIP(hipMemcpyAsync(&ue, uoff.p + M, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    MSJ_HIP(hipMemcpyAsync(&trows, tsum.p, sizeof(trows), hipMemcpyDeviceToHost, s));
    MSJ_HIP(hipStreamSynchronize(s));
    tile_rows = trows;
    uadj.ensure(std::max<uint32_t>(ue, 1));
    ucum.ensure(std::max<uint32_t>(ue, 1));
    LAUNCH(k_upper_write, grid_for(M, 256), 256, ix.nstart.p, ix.aoff.p, ix.adj.p, M, uoff.p,
           uadj.p, ucum.p);
    if (!plain_items) {
        nitems = 0;
        return;
    }

    LAUNCH(k_item_count, grid_for(M, 256), 256, ix.nstart.p, ulen.p, M, qmax, epi, itcnt.p);
    scan.excl_u32_to_u64(itcnt.p, tot64.p, M, s);
    uint64_t total = 0;
    MSJ_HIP(hipMemcpyAsync(&total, tot64.p + M, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    MSJ_HIP(hipStreamSynchronize(s));
    if (total >= 0xFFFFFFFFull)
        throw Error(MSJ_E_OVERFLOW, "join work items exceed 2^32");
    nitems = (uint32_t)total;
    scan.excl_u32_to_u32(itcnt.p, itoff.p, M, s);
    items.ensure(std::max<uint32_t>(nitems, 1));
    LAUNCH(k_item_fill, grid_for(nitems, 256), 256, ix.nstart.p, ulen.p, itoff.p, M, qmax, epi,
           nitems, items.p);
}

uint64_t JoinRunner::run_plain(const Index &ix, const double *xr, uint32_t n, uint32_t d,
                               double eps, uint64_t cap_hint, hipStream_t s, hipEvent_t k0,
                               hipEvent_t k1, uint64_t *launches, float *kernel_ms,
                               uint64_t *dims_evaluated, bool progress)
{
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
        if (nitems && !progress) {
            dispatch_plain(view(ix, xr, n, d, eps), work, s);
            ++*launches;
        } else if (nitems) {
            const JobView V = view(ix, xr, n, d, eps);
            const uint32_t parts = 16, step = (nitems + parts - 1) / parts;
            for (uint32_t k = 0; k * step < nitems; ++k) {
                JobView P = V;
                P.items = V.items + (size_t)k * step;
                P.nitems = std::min(step, nitems - k * step);
                dispatch_plain(P, work, s);
                ++*launches;
                MSJ_HIP(hipStreamSynchronize(s));
                std::fprintf(stderr, "[msj] join: %u/%u item ranges done\n", k + 1,
                             (nitems + step - 1) / step);
            }
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

void JoinRunner::launch_plain(const Index &ix, const double *xr, uint32_t n, uint32_t d,
                              double eps, hipStream_t s)
{
    counter.ensure(1);
    LAUNCH(k_zero<unsigned long long>, 1, 64, counter.p, (size_t)1);
    if (nitems)
        dispatch_plain(view(ix, xr, n, d, eps), nullptr, s);
}

void JoinRunner::finish(uint32_t n, uint64_t R, uint32_t *counts, uint64_t *offsets, uint32_t *nbrs,
                        hipStream_t s)
{
    const int qbits = std::max(1, bits_for(n ? n - 1 : 0));
    const int rbits = bits_for(n ? n - 1 : 0);
    sorter.sort(pairs.p, nullptr, R, qbits + rbits, s);
    LAUNCH(k_zero<uint32_t>, grid_for(n, 256), 256, counts, (size_t)n);
    if (R) {
        // offsets[] holds the runs' first positions until the scan below
        LAUNCH(k_rows_first, grid_stride(R, 256), 256, pairs.p, (size_t)R, qbits,
               (1ull << qbits) - 1, offsets, nbrs);
        LAUNCH(k_rows_count, grid_stride(R, 256), 256, pairs.p, (size_t)R, qbits, offsets,
               counts);
    }
    scan.excl_u32_to_u64(counts, offsets, n, s);
}

}  // namespace msj
