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
// candidates at a time, one candidate per lane with its first DR
// coordinates in VGPRs, and loops over the queries; a query's coordinates
// are wave-uniform, so they are scalar loads into SGPRs and reach every lane
// with no broadcast instruction.  The FP64 data is a node-ordered row-major
// copy, so a query row is one contiguous scalar load.
//
// Decision: exactly the reference's sequence per 4-dimension chunk,
//   s = fma(d0, d0, s); s = s + d1*d1; s = s + d2*d2; s = s + d3*d3;
//   reject once s > eps^2,
// with this file compiled without floating-point contraction.  The wave
// leaves the chunk loop as soon as no lane is still undecided.
//
// Emission: one atomic per wave per accepting query (ballot + popcount
// prefix), not per pair.  Keys are (row << qbits | neighbour row); one radix
// sort then yields every row's neighbours ascending, the counts and the
// CSR offsets.
#include <algorithm>

#include "msj_internal.h"

#pragma clang fp contract(off)

namespace msj {

namespace {

constexpr uint64_t EVALS_PER_ITEM = 262144;  // lane evaluations per work item
constexpr uint32_t WPB = 4;                  // waves per block
constexpr uint32_t LDS_WAVE_BYTES = 8192;    // staged query rows per wave
constexpr uint32_t HITS = 256;               // per-wave LDS buffer of result keys
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

// Queries per work item for a row width DS (in-register dimensions): as many
// FP64 rows as fit LDS_WAVE_BYTES, at most 64.
__host__ __device__ constexpr uint32_t qmax_for(uint32_t DS)
{
    return (LDS_WAVE_BYTES / (8u * DS)) < 64u ? (LDS_WAVE_BYTES / (8u * DS)) : 64u;
}

// Wave-wide vote on a lane predicate, kept in SGPRs (no VALU round trip).
__device__ inline uint64_t wave_vote(bool v) { return __builtin_amdgcn_ballot_w64(v); }

// Rows of xr are ds = roundup(d, 4) doubles, zero-padded: a padded
// coordinate contributes s + (0 - 0)^2 = s exactly, so the chunk loop needs
// no per-dimension guard.  DS is ds when ds <= 32 (candidate rows live in
// VGPRs); otherwise DS = 32 and the remaining chunks come from memory.
//
// A work item's query rows are contiguous in xr (node order); the wave
// copies them into its LDS slice once, then every query of the inner loop
// is a wave-uniform LDS address (a broadcast read) instead of a scalar load
// that would miss the small scalar cache.
//
// Result keys go to a per-wave LDS buffer; the wave reserves global space
// with one atomic per flush (at most HITS keys) and copies the buffer out
// with coalesced stores.  A global atomic per accepting step would
// serialise on the single counter.
template <int DS, bool TAIL, bool COUNT>
__global__ __launch_bounds__(256) void k_join_sym(
    const double *__restrict__ xr, uint32_t d, uint32_t ds, double eps2,
    const uint32_t *__restrict__ nstart, const uint32_t *__restrict__ uoff,
    const uint32_t *__restrict__ uadj, const uint32_t *__restrict__ perm,
    const JoinItem *__restrict__ items, uint32_t nitems, int qbits,
    unsigned long long *__restrict__ out, unsigned long long *__restrict__ counter,
    unsigned long long cap, unsigned long long *__restrict__ work)
{
    constexpr uint32_t QM = qmax_for(DS);
    __shared__ double qsh[WPB][QM * DS];
    __shared__ unsigned long long hsh[WPB][HITS];
    const uint32_t lane = threadIdx.x & (WAVE - 1);
    const uint32_t wv = threadIdx.x / WAVE;
    const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (WAVE - lane));
    unsigned long long *hb = hsh[wv];
    uint32_t fill = 0;  // wave-uniform
    uint64_t dims = 0;
    // copies the buffered keys out: one atomic for the whole buffer
    auto flush = [&]() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        unsigned long long base = 0;
        if (lane == 0)
            base = atomicAdd(counter, (unsigned long long)fill);
        base = __shfl(base, 0, WAVE);
        for (uint32_t e = lane; e < fill; e += WAVE)
            if (base + e < cap)
                out[base + e] = hb[e];
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        fill = 0;
    };
    // grid-stride over groups of WPB items (a launch stays under 2^32 lanes)
    for (uint64_t it0 = (uint64_t)blockIdx.x * WPB; it0 < nitems;
         it0 += (uint64_t)gridDim.x * WPB) {
        // wave-uniform by construction; readfirstlane lets the compiler keep the
        // item and the loop counters in SGPRs
        const bool have = it0 + wv < nitems;
        const uint32_t it = __builtin_amdgcn_readfirstlane(have ? (uint32_t)(it0 + wv) : 0u);
        JoinItem w = {0, 0, 0, 0, 0};
        if (have)
            w = items[it];
        const uint32_t X = w.x;
        const uint32_t x0 = have ? nstart[X] : 0u;
        const uint32_t stride = TAIL ? ds : (uint32_t)DS;
        const uint32_t qn = w.qend - w.qbeg;
        // stage the item's query rows (first DS dimensions) in LDS
        double *qs = qsh[wv];
        for (uint32_t e = lane; e < qn * DS; e += WAVE) {
            const uint32_t r = e / DS, k = e - r * DS;
            qs[e] = xr[(size_t)(x0 + w.qbeg + r) * stride + k];
        }
        __syncthreads();
        if (have) {
            const uint64_t qall = qn >= WAVE ? ~0ull : ((1ull << qn) - 1ull);

            // per-lane cursor into X's upper neighbour list
            uint32_t a = uoff[X];
            uint32_t Y = uadj[a];
            uint32_t ybase = 0;
            uint32_t ystart = nstart[Y];
            uint32_t ysz = nstart[Y + 1] - ystart;

            for (uint32_t cs = w.cbeg; cs < w.cend; cs += WAVE) {
                const uint32_t ci = cs + lane;
                const bool cvalid = ci < w.cend;
                uint32_t c = 0;
                if (cvalid) {
                    while (ci >= ybase + ysz) {
                        ybase += ysz;
                        ++a;
                        Y = uadj[a];
                        ystart = nstart[Y];
                        ysz = nstart[Y + 1] - ystart;
                    }
                    c = ystart + (ci - ybase);
                }
                const bool self_node = cvalid && Y == X;
                const double *crow = xr + (size_t)c * stride;  // c = 0 for idle lanes
                double cc[DS];
        #pragma unroll
                for (int k = 0; k < DS; ++k)
                    cc[k] = crow[k];
                const unsigned long long corig = perm[c];

                uint64_t qmask = qall;
                while (qmask) {
                    const uint32_t qi = (uint32_t)__builtin_ctzll(qmask);
                    qmask &= qmask - 1ull;
                    const uint32_t p = x0 + w.qbeg + qi;
                    bool alive = cvalid && (!self_node || p <= c);
                    if (!wave_vote(alive))
                        continue;
                    const double *qrow = qs + qi * DS;  // wave-uniform LDS address
                    double s = 0.0;
        #pragma unroll
                    for (int l = 0; l < DS; l += 4) {
                        if (l > 0 && !wave_vote(alive))
                            break;
                        if (COUNT && alive && (uint32_t)l < d)
                            dims += min(4u, d - (uint32_t)l);
                        const double2 qa = *(const double2 *)(qrow + l);
                        const double2 qb = *(const double2 *)(qrow + l + 2);
                        double t0 = qa.x - cc[l];
                        s = fma(t0, t0, s);
                        double t1 = qa.y - cc[l + 1];
                        double q1 = t1 * t1;
                        s = s + q1;
                        double t2 = qb.x - cc[l + 2];
                        double q2 = t2 * t2;
                        s = s + q2;
                        double t3 = qb.y - cc[l + 3];
                        double q3 = t3 * t3;
                        s = s + q3;
                        alive = alive && !(s > eps2);
                    }
                    if (TAIL) {
                        const double *grow = xr + (size_t)p * stride;
                        for (uint32_t l = DS; l < ds && wave_vote(alive); l += 4) {
                            if (COUNT && alive && l < d)
                                dims += min(4u, d - l);
                            double t0 = grow[l] - crow[l];
                            s = fma(t0, t0, s);
                            double t1 = grow[l + 1] - crow[l + 1];
                            double q1 = t1 * t1;
                            s = s + q1;
                            double t2 = grow[l + 2] - crow[l + 2];
                            double q2 = t2 * t2;
                            s = s + q2;
                            double t3 = grow[l + 3] - crow[l + 3];
                            double q3 = t3 * t3;
                            s = s + q3;
                            alive = alive && !(s > eps2);
                        }
                    }
                    const uint64_t m1 = wave_vote(alive);
                    if (m1) {
                        const bool two = alive && c != p;
                        const uint64_t m2 = wave_vote(two);
                        const uint32_t tot = (uint32_t)(__popcll(m1) + __popcll(m2));
                        if (fill + tot > HITS)
                            flush();
                        if (alive) {
                            const unsigned long long porig = perm[p];
                            const uint32_t idx =
                                fill + (uint32_t)(__popcll(m1 & lt) + __popcll(m2 & lt));
                            hb[idx] = (porig << qbits) | corig;
                            if (two)
                                hb[idx + 1] = (corig << qbits) | porig;
                        }
                        fill += tot;
                    }
                }
            }
        }
        __syncthreads();  // the next group's staging reuses qsh
    }
    if (fill)
        flush();
    if (COUNT) {
        unsigned long long v = dims;
#pragma unroll
        for (int o = WAVE / 2; o > 0; o >>= 1)
            v += __shfl_down(v, o, WAVE);
        if (lane == 0)
            atomicAdd(work, v);
    }
}

// High-dimensional rows (ds > 32).  Slab-major order: a wave holds the
// 64 candidates' next HSL dimensions in VGPRs (each loaded once) and
// advances the partial sums of all of the item's queries (up to 32, one
// FP64 sum per lane and query in VGPRs) through that slab, chunk by chunk in
// the reference's order, with the same early-exit test; a query row slab is
// staged in LDS and read as a broadcast.  Decisions are those of the
// per-pair sequence: the chunks of every pair run in order, each followed by
// the s > eps^2 test.
constexpr int HSL = 16;  // dimensions per slab
constexpr int HQ = 32;   // queries per item

template <bool COUNT>
__global__ __launch_bounds__(256) void k_join_hd(
    const double *__restrict__ xr, uint32_t d, uint32_t ds, double eps2,
    const uint32_t *__restrict__ nstart, const uint32_t *__restrict__ uoff,
    const uint32_t *__restrict__ uadj, const uint32_t *__restrict__ perm,
    const JoinItem *__restrict__ items, uint32_t nitems, int qbits,
    unsigned long long *__restrict__ out, unsigned long long *__restrict__ counter,
    unsigned long long cap, unsigned long long *__restrict__ work)
{
    __shared__ double qsh[WPB][HQ * HSL];
    __shared__ unsigned long long hsh[WPB][HITS];
    const uint32_t lane = threadIdx.x & (WAVE - 1);
    const uint32_t wv = threadIdx.x / WAVE;
    const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (WAVE - lane));
    unsigned long long *hb = hsh[wv];
    double *qs = qsh[wv];
    uint32_t fill = 0;  // wave-uniform
    uint64_t dims = 0;
    auto wave_sync = [&]() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    auto flush = [&]() {
        wave_sync();
        unsigned long long base = 0;
        if (lane == 0)
            base = atomicAdd(counter, (unsigned long long)fill);
        base = __shfl(base, 0, WAVE);
        for (uint32_t e = lane; e < fill; e += WAVE)
            if (base + e < cap)
                out[base + e] = hb[e];
        wave_sync();
        fill = 0;
    };
    // one wave per item, grid-stride (no block-wide synchronisation here)
    for (uint64_t it0 = (uint64_t)blockIdx.x * WPB + wv; it0 < nitems;
         it0 += (uint64_t)gridDim.x * WPB) {
        const uint32_t it = __builtin_amdgcn_readfirstlane((uint32_t)it0);
        const JoinItem w = items[it];
        const uint32_t X = w.x;
        const uint32_t x0 = nstart[X];
        const uint32_t qn = w.qend - w.qbeg;  // <= HQ
        const uint32_t q0 = x0 + w.qbeg;      // first query (node order)
        uint32_t a = uoff[X];
        uint32_t Y = uadj[a];
        uint32_t ybase = 0;
        uint32_t ystart = nstart[Y];
        uint32_t ysz = nstart[Y + 1] - ystart;
        for (uint32_t cs = w.cbeg; cs < w.cend; cs += WAVE) {
            const uint32_t ci = cs + lane;
            const bool cvalid = ci < w.cend;
            uint32_t c = 0;
            if (cvalid) {
                while (ci >= ybase + ysz) {
                    ybase += ysz;
                    ++a;
                    Y = uadj[a];
                    ystart = nstart[Y];
                    ysz = nstart[Y + 1] - ystart;
                }
                c = ystart + (ci - ybase);
            }
            const bool self_node = cvalid && Y == X;
            uint32_t alive = 0;  // bit q: pair (query q, this lane's candidate) undecided
            for (uint32_t q = 0; q < qn; ++q)
                if (cvalid && (!self_node || q0 + q <= c))
                    alive |= 1u << q;
            double sq[HQ];
#pragma unroll
            for (int q = 0; q < HQ; ++q)
                sq[q] = 0.0;
            const double *crow = xr + (size_t)c * ds;
            for (uint32_t l0 = 0; l0 < ds; l0 += HSL) {
                if (!wave_vote(alive != 0))
                    break;
                // query slab -> LDS (qn rows of HSL dims), candidate slab -> VGPRs
                wave_sync();
                for (uint32_t e = lane; e < HQ * HSL; e += WAVE) {
                    const uint32_t q = e / HSL, j = e % HSL;
                    qs[e] = (q < qn && l0 + j < ds) ? xr[(size_t)(q0 + q) * ds + l0 + j] : 0.0;
                }
                double cc[HSL];
#pragma unroll
                for (int j = 0; j < HSL; j += 2) {
                    if (l0 + j < ds) {
                        const double2 v = *(const double2 *)(crow + l0 + j);
                        cc[j] = v.x;
                        cc[j + 1] = v.y;
                    } else {
                        cc[j] = 0.0;
                        cc[j + 1] = 0.0;
                    }
                }
                wave_sync();
                // fully unrolled over q (sq[] stays in VGPRs): no break or
                // continue; a decided query or chunk is skipped by its vote
#pragma unroll
                for (int q = 0; q < HQ; ++q) {
                    bool al = (alive >> q) & 1u;  // 0 for q >= qn
                    if (!wave_vote(al))
                        continue;
                    const double *qrow = qs + q * HSL;  // wave-uniform LDS address
                    double s = sq[q];
#pragma unroll
                    for (int k = 0; k < HSL; k += 4) {
                        if (l0 + k >= ds || !wave_vote(al))
                            continue;
                        if (COUNT && al && l0 + k < d)
                            dims += min(4u, d - (l0 + k));
                        const double2 qa = *(const double2 *)(qrow + k);
                        const double2 qb = *(const double2 *)(qrow + k + 2);
                        double t0 = qa.x - cc[k];
                        s = fma(t0, t0, s);
                        double t1 = qa.y - cc[k + 1];
                        double p1 = t1 * t1;
                        s = s + p1;
                        double t2 = qb.x - cc[k + 2];
                        double p2 = t2 * t2;
                        s = s + p2;
                        double t3 = qb.y - cc[k + 3];
                        double p3 = t3 * t3;
                        s = s + p3;
                        al = al && !(s > eps2);
                    }
                    sq[q] = s;
                    if (!al)
                        alive &= ~(1u << q);
                }
            }
            // emission: the pairs still undecided after every slab are results
            const unsigned long long corig = perm[c];
            for (uint32_t q = 0; q < qn; ++q) {
                const bool acc = (alive >> q) & 1u;
                const uint64_t m1 = wave_vote(acc);
                if (!m1)
                    continue;
                const uint32_t p = q0 + q;
                const bool two = acc && c != p;
                const uint64_t m2 = wave_vote(two);
                const uint32_t tot = (uint32_t)(__popcll(m1) + __popcll(m2));
                if (fill + tot > HITS)
                    flush();
                if (acc) {
                    const unsigned long long porig = perm[p];
                    const uint32_t idx = fill + (uint32_t)(__popcll(m1 & lt) + __popcll(m2 & lt));
                    hb[idx] = (porig << qbits) | corig;
                    if (two)
                        hb[idx + 1] = (corig << qbits) | porig;
                }
                fill += tot;
            }
        }
    }
    if (fill)
        flush();
    if (COUNT) {
        unsigned long long v = dims;
#pragma unroll
        for (int o = WAVE / 2; o > 0; o >>= 1)
            v += __shfl_down(v, o, WAVE);
        if (lane == 0)
            atomicAdd(work, v);
    }
}

// Upper neighbour lists: entries Y >= X of X's adjacency, and their sizes.
__global__ __launch_bounds__(256) void k_upper_count(const uint32_t *__restrict__ nstart,
                                                     const uint32_t *__restrict__ aoff,
                                                     const uint32_t *__restrict__ adj, uint32_t M,
                                                     uint32_t *__restrict__ cnt,
                                                     uint64_t *__restrict__ len)
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
}

__global__ __launch_bounds__(256) void k_upper_write(const uint32_t *__restrict__ aoff,
                                                     const uint32_t *__restrict__ adj, uint32_t M,
                                                     const uint32_t *__restrict__ uoff,
                                                     uint32_t *__restrict__ uadj)
{
    uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= M)
        return;
    uint32_t o = uoff[x];
    for (uint32_t e = aoff[x]; e < aoff[x + 1]; ++e)
        if (adj[e] >= x)
            uadj[o++] = adj[e];
}

// Items per node: query ranges of up to qmax points x candidate ranges sized
// so one item is about EVALS_PER_ITEM lane evaluations (multiples of 64).
__device__ inline void item_shape(uint32_t sz, uint64_t L, uint32_t qmax, uint32_t *qper,
                                  uint32_t *nq, uint32_t *cper, uint32_t *nc)
{
    *qper = min(qmax, sz);
    *nq = (sz + *qper - 1) / *qper;
    uint64_t cp = EVALS_PER_ITEM / *qper;
    cp = (cp + WAVE - 1) / WAVE * WAVE;
    *cper = (uint32_t)cp;
    uint64_t k = (L + cp - 1) / cp;
    *nc = (uint32_t)(k ? k : 1);
}

__global__ __launch_bounds__(256) void k_item_count(const uint32_t *__restrict__ nstart,
                                                    const uint64_t *__restrict__ ulen, uint32_t M,
                                                    uint32_t qmax, uint32_t *__restrict__ cnt)
{
    uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= M)
        return;
    uint32_t qper, nq, cper, nc;
    item_shape(nstart[x + 1] - nstart[x], ulen[x], qmax, &qper, &nq, &cper, &nc);
    cnt[x] = nq * nc;
}

__global__ __launch_bounds__(256) void k_item_fill(const uint32_t *__restrict__ nstart,
                                                   const uint64_t *__restrict__ ulen,
                                                   const uint32_t *__restrict__ itoff, uint32_t M,
                                                   uint32_t qmax, uint32_t nitems,
                                                   JoinItem *__restrict__ items)
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
    item_shape(sz, L, qmax, &qper, &nq, &cper, &nc);
    const uint32_t local = k - itoff[x];
    const uint32_t qi = local / nc, ck = local % nc;
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

__global__ __launch_bounds__(256) void k_rows_count(const uint64_t *__restrict__ pairs, size_t R,
                                                    int qbits, uint32_t *__restrict__ counts)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < R; i += stride) {
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
}

__global__ __launch_bounds__(256) void k_rows_nbrs(const uint64_t *__restrict__ pairs, size_t R,
                                                   uint64_t qmask, uint32_t *__restrict__ nbrs)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < R; i += stride)
        nbrs[i] = (uint32_t)(pairs[i] & qmask);
}

// Per-column minimum and maximum of the first columns (one block per column).
template <class T>
__global__ __launch_bounds__(256) void k_col_range(const T *__restrict__ xs, uint32_t n,
                                                   double *__restrict__ mm)
{
    __shared__ double lo_s[256 / WAVE], hi_s[256 / WAVE];
    const uint32_t j = blockIdx.x;
    double lo = __builtin_inf(), hi = -__builtin_inf();
    for (uint32_t i = threadIdx.x; i < n; i += blockDim.x) {
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
        mm[2 * j] = lo;
        mm[2 * j + 1] = hi;
    }
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

template <int DS, bool TAIL>
void launch_sym(const JoinRunner &J, const Index &ix, const double *xr, uint32_t d, uint32_t ds,
                double eps2, int qbits, unsigned long long *out, unsigned long long cap,
                unsigned long long *work, hipStream_t s)
{
    const unsigned wpb = WPB;
    if (work)
        LAUNCH((k_join_sym<DS, TAIL, true>), grid_stride(J.nitems, wpb), WAVE * wpb, xr, d, ds,
               eps2, ix.nstart.p, J.uoff.p, J.uadj.p, J.jperm.p, J.items.p, J.nitems, qbits, out,
               J.counter.p, cap, work);
    else
        LAUNCH((k_join_sym<DS, TAIL, false>), grid_stride(J.nitems, wpb), WAVE * wpb, xr, d, ds,
               eps2, ix.nstart.p, J.uoff.p, J.uadj.p, J.jperm.p, J.items.p, J.nitems, qbits, out,
               J.counter.p, cap, work);
}

void dispatch_sym(const JoinRunner &J, const Index &ix, const double *xr, uint32_t d, double eps2,
                  int qbits, unsigned long long *out, unsigned long long cap,
                  unsigned long long *work, hipStream_t s)
{
    const uint32_t ds = padded_dim(d);
    switch (ds) {
#define MSJ_DS(v)                                                                   \
    case v:                                                                         \
        launch_sym<v, false>(J, ix, xr, d, ds, eps2, qbits, out, cap, work, s);       \
        return;
        MSJ_DS(4)
        MSJ_DS(8)
        MSJ_DS(12)
        MSJ_DS(16)
        MSJ_DS(20)
        MSJ_DS(24)
        MSJ_DS(28)
        MSJ_DS(32)
#undef MSJ_DS
    default:
        if (work)
            LAUNCH(k_join_hd<true>, grid_stride(J.nitems, WPB), WAVE * WPB, xr, d, ds, eps2,
                   ix.nstart.p, J.uoff.p, J.uadj.p, J.jperm.p, J.items.p, J.nitems, qbits, out,
                   J.counter.p, cap, work);
        else
            LAUNCH(k_join_hd<false>, grid_stride(J.nitems, WPB), WAVE * WPB, xr, d, ds, eps2,
                   ix.nstart.p, J.uoff.p, J.uadj.p, J.jperm.p, J.items.p, J.nitems, qbits, out,
                   J.counter.p, cap, work);
    }
}

}  // namespace

void gather_node_rows(const void *xs, double *xr, bool f32, const uint32_t *perm, uint32_t n,
                      uint32_t d, hipStream_t s)
{
    const uint32_t ds = padded_dim(d);
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
    mm.ensure(2 * 3);
    jperm.ensure(n);
    jkeys.ensure(n);
    if (f32) {
        LAUNCH(k_col_range<float>, km, 256, (const float *)xs, n, mm.p);
        LAUNCH(k_morton_keys<float>, grid_for(n, 256), 256, (const float *)xs, n, km, mm.p,
               ix.nstart.p, ix.M, ix.perm.p, jkeys.p, jperm.p);
    } else {
        LAUNCH(k_col_range<double>, km, 256, (const double *)xs, n, mm.p);
        LAUNCH(k_morton_keys<double>, grid_for(n, 256), 256, (const double *)xs, n, km, mm.p,
               ix.nstart.p, ix.M, ix.perm.p, jkeys.p, jperm.p);
    }
    sorter.sort(jkeys.p, jperm.p, n, bits_for(ix.M ? ix.M - 1 : 0) + 30, s);
}

void JoinRunner::plan(const Index &ix, uint32_t n, uint32_t d, hipStream_t s, bool plain_items)
{
    const uint32_t M = ix.M;
    const uint32_t ds = padded_dim(d);
    const uint32_t qmax = ds <= 32 ? qmax_for(ds) : (uint32_t)HQ;
    itcnt.ensure(M);
    itoff.ensure((size_t)M + 1);
    ulen.ensure(M);
    uoff.ensure((size_t)M + 1);
    tot64.ensure((size_t)M + 1);
    LAUNCH(k_upper_count, grid_for(M, 256), 256, ix.nstart.p, ix.aoff.p, ix.adj.p, M, itcnt.p,
           ulen.p);
    scan.excl_u32_to_u32(itcnt.p, uoff.p, M, s);
    uint32_t ue = 0;
    MSJ_HIP(hipMemcpyAsync(&ue, uoff.p + M, sizeof(uint32_t), hipMemcpyDeviceToHost, s));
    MSJ_HIP(hipStreamSynchronize(s));
    uadj.ensure(std::max<uint32_t>(ue, 1));
    LAUNCH(k_upper_write, grid_for(M, 256), 256, ix.aoff.p, ix.adj.p, M, uoff.p, uadj.p);
    if (!plain_items) {
        nitems = 0;
        return;
    }

    LAUNCH(k_item_count, grid_for(M, 256), 256, ix.nstart.p, ulen.p, M, qmax, itcnt.p);
    scan.excl_u32_to_u64(itcnt.p, tot64.p, M, s);
    uint64_t total = 0;
    MSJ_HIP(hipMemcpyAsync(&total, tot64.p + M, sizeof(uint64_t), hipMemcpyDeviceToHost, s));
    MSJ_HIP(hipStreamSynchronize(s));
    if (total >= 0xFFFFFFFFull)
        throw Error(MSJ_E_OVERFLOW, "join work items exceed 2^32");
    nitems = (uint32_t)total;
    scan.excl_u32_to_u32(itcnt.p, itoff.p, M, s);
    items.ensure(std::max<uint32_t>(nitems, 1));
    LAUNCH(k_item_fill, grid_for(nitems, 256), 256, ix.nstart.p, ulen.p, itoff.p, M, qmax, nitems,
           items.p);
}

uint64_t JoinRunner::run_plain(const Index &ix, const double *xr, uint32_t n, uint32_t d,
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
            dispatch_sym(*this, ix, xr, d, eps2, qbits, (unsigned long long *)pairs.p, pairs.cap,
                         work, s);
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
        LAUNCH(k_rows_count, grid_stride(R, 256), 256, pairs.p, (size_t)R, qbits, counts);
        LAUNCH(k_rows_nbrs, grid_stride(R, 256), 256, pairs.p, (size_t)R, (1ull << qbits) - 1,
               nbrs);
    }
    scan.excl_u32_to_u64(counts, offsets, n, s);
}

}  // namespace msj
