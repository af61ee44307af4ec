// msj_join_items.h -- one work item of the plain FP64 join, as device
// functions shared by the single-join kernels (msj_join.hip) and the
// persistent batch kernel (msj_batch.hip).
//
// Reference: kernel.cu:992-1046 nodeByPoint5 and kernel.cu:833-865
// cachedDistanceCheck (ILP = 4).
//
// Symmetry.  The test is bit-symmetric: q - c = -(c - q) exactly, so both
// orders of a pair give the same sum and the same decision.  Each unordered
// candidate pair is therefore evaluated once -- node X against the nodes
// Y >= X of its neighbour list, and inside X only pairs p <= c -- and an
// accepted pair is emitted in both orders (the self pair once).
//
// Decision: exactly the reference's sequence per 4-dimension chunk,
//   s = fma(d0, d0, s); s = s + d1*d1; s = s + d2*d2; s = s + d3*d3;
//   reject once s > eps^2,
// with every including file compiled without floating-point contraction.
//
// Emission: result keys (row << qbits | neighbour row) go to a per-wave LDS
// buffer; the wave reserves global space with one atomic per flush and
// copies the buffer out with coalesced stores.  One radix sort then yields
// every row's neighbours ascending, the counts and the CSR offsets.
#pragma once

#include "msj_internal.h"

#pragma clang fp contract(off)

namespace msj {

constexpr uint32_t WPB = 4;                // waves per block
constexpr uint32_t LDS_WAVE_BYTES = 8192;  // staged query rows per wave
constexpr uint32_t HITS = 256;             // per-wave LDS buffer of result keys
constexpr int HSL = 16;                    // high-d kernel: dimensions per slab
constexpr int HQ = 24;                     // high-d kernel: queries per item

// Queries per work item for a row width DS (in-register dimensions): as many
// FP64 rows as fit LDS_WAVE_BYTES, at most 64.
__host__ __device__ constexpr uint32_t qmax_for(uint32_t DS)
{
    return (LDS_WAVE_BYTES / (8u * DS)) < 64u ? (LDS_WAVE_BYTES / (8u * DS)) : 64u;
}

// Queries per work item of a join with padded row width ds.
__host__ __device__ constexpr uint32_t item_queries(uint32_t ds)
{
    return ds <= 32 ? qmax_for(ds) : (uint32_t)HQ;
}

// Wave-wide vote on a lane predicate, kept in SGPRs (no VALU round trip).
__device__ inline uint64_t wave_vote(bool v) { return __builtin_amdgcn_ballot_w64(v); }

__device__ inline void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Per-wave LDS buffer of result keys.
struct WaveHits {
    unsigned long long *hb;
    uint32_t fill;  // wave-uniform
    uint32_t lane;
    uint64_t lt;    // lanes below this one
    __device__ WaveHits(unsigned long long *buf, uint32_t l)
        : hb(buf), fill(0), lane(l), lt(l == 0 ? 0ull : ((~0ull) >> (WAVE - l)))
    {
    }
    __device__ void flush(const JobView &J)
    {
        wave_sync();
        unsigned long long base = 0;
        if (lane == 0)
            base = atomicAdd(J.counter, (unsigned long long)fill);
        base = __shfl(base, 0, WAVE);
        for (uint32_t e = lane; e < fill; e += WAVE)
            if (base + e < J.cap)
                J.out[base + e] = hb[e];
        wave_sync();
        fill = 0;
    }
    // Emits (p, c) and, unless p == c, (c, p) for the lanes with `acc`.
    __device__ void emit(const JobView &J, bool acc, uint32_t p, uint32_t c,
                         unsigned long long porig, unsigned long long corig)
    {
        const uint64_t m1 = wave_vote(acc);
        if (!m1)
            return;
        const bool two = acc && c != p;
        const uint64_t m2 = wave_vote(two);
        const uint32_t tot = (uint32_t)(__popcll(m1) + __popcll(m2));
        if (fill + tot > HITS)
            flush(J);
        if (acc) {
            const uint32_t idx = fill + (uint32_t)(__popcll(m1 & lt) + __popcll(m2 & lt));
            hb[idx] = (porig << J.qbits) | corig;
            if (two)
                hb[idx + 1] = (corig << J.qbits) | porig;
        }
        fill += tot;
    }
};

// The entry of X's upper list whose points hold candidate offset c (the
// last entry starting at or before c): binary search, wave-uniform.
__device__ inline uint32_t first_entry(const JobView &J, uint32_t X, uint32_t c)
{
    uint32_t lo = J.uoff[X], hi = J.uoff[X + 1];
    while (hi - lo > 1) {
        const uint32_t mid = lo + (hi - lo) / 2;
        if (J.ucum[mid] <= c)
            lo = mid;
        else
            hi = mid;
    }
    return lo;
}

// Rows of xr are ds = roundup(d, 4) doubles, zero-padded: a padded
// coordinate contributes s + (0 - 0)^2 = s exactly, so the chunk loop needs
// no per-dimension guard.  DS = ds (<= 32): candidate rows live in VGPRs.
//
// One wave runs one item: queries [qbeg, qend) of node X against entries
// [cbeg, cend) of X's upper candidate list, 64 candidates at a time, one
// per lane with its coordinates in VGPRs; the item's query rows are staged
// in the wave's LDS slice `qs`, so every query of the inner loop is a
// wave-uniform LDS address (a broadcast read).
template <int DS, bool COUNT>
__device__ inline void item_sym(const JobView &J, const JoinItem &w, double *qs, WaveHits &H,
                                uint64_t &dims)
{
    const uint32_t lane = H.lane;
    const uint32_t X = w.x;
    const uint32_t x0 = J.nstart[X];
    const uint32_t qn = w.qend - w.qbeg;
    wave_sync();  // the previous item's reads of qs are done
    for (uint32_t e = lane; e < qn * DS; e += WAVE) {
        const uint32_t r = e / DS, k = e - r * DS;
        qs[e] = J.xr[(size_t)(x0 + w.qbeg + r) * DS + k];
    }
    wave_sync();
    const uint64_t qall = qn >= WAVE ? ~0ull : ((1ull << qn) - 1ull);
    // row ids of the item's queries, query r in lane r (qn <= 64): an
    // emitting query reads its id with a lane read, not a memory load
    const uint32_t qperm = lane < qn ? J.perm[x0 + w.qbeg + lane] : 0u;

    // per-lane cursor into X's upper neighbour list, from the entry holding
    // the item's first candidate
    uint32_t a = first_entry(J, X, w.cbeg);
    uint32_t ybase = J.ucum[a];
    uint32_t Y = J.uadj[a];
    uint32_t ystart = J.nstart[Y];
    uint32_t ysz = J.nstart[Y + 1] - ystart;

    for (uint32_t cs = w.cbeg; cs < w.cend; cs += WAVE) {
        const uint32_t ci = cs + lane;
        const bool cvalid = ci < w.cend;
        uint32_t c = 0;
        if (cvalid) {
            while (ci >= ybase + ysz) {
                ybase += ysz;
                ++a;
                Y = J.uadj[a];
                ystart = J.nstart[Y];
                ysz = J.nstart[Y + 1] - ystart;
            }
            c = ystart + (ci - ybase);
        }
        const uint64_t vmask = wave_vote(cvalid);           // lanes with a candidate
        const uint64_t smask = wave_vote(cvalid && Y == X);  // ... inside node X
        const double *crow = J.xr + (size_t)c * DS;  // c = 0 for idle lanes
        double cc[DS];
#pragma unroll
        for (int k = 0; k < DS; ++k)
            cc[k] = crow[k];
        const unsigned long long corig = J.perm[c];

        uint64_t qmask = qall;
        while (qmask) {
            const uint32_t qi = (uint32_t)__builtin_ctzll(qmask);
            qmask &= qmask - 1ull;
            const uint32_t p = x0 + w.qbeg + qi;
            // undecided lanes as a wave mask; inside X only pairs p <= c
            uint64_t am = vmask;
            if (smask)
                am &= ~smask | wave_vote(p <= c);
            if (!am)
                continue;
            const double *qrow = qs + qi * DS;  // wave-uniform LDS address
            double s = 0.0;
#pragma unroll
            for (int l = 0; l < DS; l += 4) {
                if (l > 4 && !am)
                    break;
                if (COUNT && ((am >> lane) & 1ull) && (uint32_t)l < J.d)
                    dims += min(4u, J.d - (uint32_t)l);
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
                // partial sums only grow, so the test after the first chunk
                // can wait for the second (the counting build tests every
                // chunk: its dimension count follows the per-chunk exit)
                if (COUNT || l > 0 || l + 4 >= DS)
                    am &= wave_vote(!(s > J.eps2));
            }
            if (am) {
                const bool acc = (am >> lane) & 1ull;
                H.emit(J, acc, p, c, (unsigned long long)__builtin_amdgcn_readlane(qperm, qi),
                       corig);
            }
        }
    }
}

// High-dimensional rows (ds > 32).  Slab-major order: the wave holds the
// 64 candidates' next SL dimensions (HSL, or the batch autotuner's choice)
// in VGPRs (each loaded once) and advances the partial sums of all of the
// item's queries (up to HQ, one FP64 sum per lane and query in VGPRs)
// through that slab, chunk by chunk in the reference's order; a query row
// slab is staged in the wave's LDS slice and read as a broadcast.  The
// undecided pairs of query q are a wave mask amq[q] (wave-uniform, SGPRs).
//
// Decisions are those of the per-pair sequence (each chunk followed by the
// s > eps^2 test): partial sums never decrease, so a pair is accepted iff
// its full sum passes the test, and any schedule of tests that ends with
// the full sum decides the same.  The join tests after every second chunk;
// chunks past ds (the last slab of a width that is not a multiple of SL)
// add (0 - 0)^2 = 0 exactly, so a slab runs its chunks in pairs with no
// per-chunk guard.  The counting build tests every chunk: its dimension
// count follows the per-chunk exit.
template <bool COUNT, int SL = HSL>
__device__ inline void item_hd(const JobView &J, const JoinItem &w, double *qs, WaveHits &H,
                               uint64_t &dims)
{
    static_assert(SL % 8 == 0, "chunks run in pairs");
    const uint32_t lane = H.lane;
    const uint32_t ds = J.ds;
    const uint32_t X = w.x;
    const uint32_t x0 = J.nstart[X];
    const uint32_t qn = w.qend - w.qbeg;  // <= HQ
    const uint32_t q0 = x0 + w.qbeg;      // first query (node order)
    const uint32_t qperm = lane < qn ? J.perm[q0 + lane] : 0u;  // query q's row id in lane q
    uint32_t a = first_entry(J, X, w.cbeg);
    uint32_t ybase = J.ucum[a];
    uint32_t Y = J.uadj[a];
    uint32_t ystart = J.nstart[Y];
    uint32_t ysz = J.nstart[Y + 1] - ystart;
    for (uint32_t cs = w.cbeg; cs < w.cend; cs += WAVE) {
        const uint32_t ci = cs + lane;
        const bool cvalid = ci < w.cend;
        uint32_t c = 0;
        if (cvalid) {
            while (ci >= ybase + ysz) {
                ybase += ysz;
                ++a;
                Y = J.uadj[a];
                ystart = J.nstart[Y];
                ysz = J.nstart[Y + 1] - ystart;
            }
            c = ystart + (ci - ybase);
        }
        const bool self_node = cvalid && Y == X;
        // bit `lane` of amq[q]: pair (query q, this lane's candidate) undecided
        uint64_t amq[HQ];
        double sq[HQ];
#pragma unroll
        for (int q = 0; q < HQ; ++q) {
            amq[q] = wave_vote(cvalid && (uint32_t)q < qn && (!self_node || q0 + q <= c));
            sq[q] = 0.0;
        }
        const double *crow = J.xr + (size_t)c * ds;
        for (uint32_t l0 = 0; l0 < ds; l0 += SL) {
            uint64_t any = 0;
#pragma unroll
            for (int q = 0; q < HQ; ++q)
                any |= amq[q];
            if (!any)
                break;
            // query slab -> LDS (qn rows of SL dims), candidate slab -> VGPRs
            wave_sync();
            for (uint32_t e = lane; e < HQ * SL; e += WAVE) {
                const uint32_t q = e / SL, j = e % SL;
                qs[e] = (q < qn && l0 + j < ds) ? J.xr[(size_t)(q0 + q) * ds + l0 + j] : 0.0;
            }
            double cc[SL];
#pragma unroll
            for (int j = 0; j < SL; j += 2) {
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
            // fully unrolled over q (sq[] in VGPRs, amq[] in SGPRs): no
            // break or continue out of it; a decided query is skipped by its mask
#pragma unroll
            for (int q = 0; q < HQ; ++q) {
                uint64_t am = amq[q];
                if (!am)
                    continue;
                // the query's slab in one batch of broadcast reads (one LDS
                // round trip per query and slab, not per chunk)
                const double *qrow = qs + q * SL;  // wave-uniform LDS address
                double2 qv[SL / 2];
#pragma unroll
                for (int j = 0; j < SL / 2; ++j)
                    qv[j] = *(const double2 *)(qrow + 2 * j);
                double s = sq[q];
                auto chunk = [&](int k) __attribute__((always_inline)) {
                    const double2 qa = qv[k / 2];
                    const double2 qb = qv[k / 2 + 1];
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
                };
                if constexpr (COUNT) {
#pragma unroll
                    for (int k = 0; k < SL; k += 4) {
                        if (l0 + k >= ds || !am)
                            continue;
                        if (((am >> lane) & 1ull) && l0 + k < J.d)
                            dims += min(4u, J.d - (l0 + k));
                        chunk(k);
                        am &= wave_vote(!(s > J.eps2));
                    }
                } else {
#pragma unroll
                    for (int k = 0; k < SL; k += 8) {
                        // wave-uniform: the pair of chunks starts inside the
                        // row and some pair of query q is undecided
                        if (k == 0 || (am && l0 + k < ds)) {
                            chunk(k);
                            chunk(k + 4);
                            am &= wave_vote(!(s > J.eps2));
                        }
                    }
                }
                sq[q] = s;
                amq[q] = am;
            }
        }
        // the pairs still undecided after every slab are results
        uint32_t alive = 0;  // bit q: pair (query q, this lane's candidate)
#pragma unroll
        for (int q = 0; q < HQ; ++q)
            alive |= (uint32_t)((amq[q] >> lane) & 1ull) << q;
        const unsigned long long corig = J.perm[c];
        for (uint32_t q = 0; q < qn; ++q) {
            const bool acc = (alive >> q) & 1u;
            const uint32_t p = q0 + q;
            H.emit(J, acc, p, c, (unsigned long long)__builtin_amdgcn_readlane(qperm, q), corig);
        }
    }
}

// Runs item w of join J with the variant for its row width; the LDS slice
// qs holds LDS_WAVE_BYTES.
template <bool COUNT>
__device__ inline void item_any(const JobView &J, const JoinItem &w, double *qs, WaveHits &H,
                                uint64_t &dims)
{
    switch (J.ds) {
    case 4: item_sym<4, COUNT>(J, w, qs, H, dims); break;
    case 8: item_sym<8, COUNT>(J, w, qs, H, dims); break;
    case 12: item_sym<12, COUNT>(J, w, qs, H, dims); break;
    case 16: item_sym<16, COUNT>(J, w, qs, H, dims); break;
    case 20: item_sym<20, COUNT>(J, w, qs, H, dims); break;
    case 24: item_sym<24, COUNT>(J, w, qs, H, dims); break;
    case 28: item_sym<28, COUNT>(J, w, qs, H, dims); break;
    case 32: item_sym<32, COUNT>(J, w, qs, H, dims); break;
    default: item_hd<COUNT>(J, w, qs, H, dims); break;
    }
}

}  // namespace msj
