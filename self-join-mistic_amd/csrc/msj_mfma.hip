// msj_mfma.hip -- precision form of the distance-and-compare step: a Gram
// filter on the matrix cores with an FP64 recheck of an error band.
//
// The decision it must reproduce is the plain form's (kernel.cu:833-865
// cachedDistanceCheck, restated in msj_join.hip): with s the FP64 sum of the
// chunked sequence, a pair is a result iff !(s > eps^2).  The partial sums of
// that sequence never decrease (every step adds a non-negative rounded term),
// so the early exit does not change the decision: it is !(s_full > eps^2).
//
// Filter.  Rows are centred on the column means mu (FP64), a' = RN(x - mu),
// and stored as BF16 (h = bf16(a')) or split-BF16 (h and l = bf16(a' - h)).
// For a candidate pair the matrix cores give G ~ a'.b' (FP32 accumulation of
// the exact BF16 products; split: h.h + h.l + l.h), and
//   D~ = |a'|^2 + |b'|^2 - 2 G
// estimates the squared distance.  With u = 2^-8 (BF16 rounding, taken with
// a 2^-15 allowance for rounding through FP32) and m products per pair:
//   |P - a'.b'|  <= e_p |a'||b'|,  e_p = 2u + u^2           (BF16)
//                                    e_p = 3u^2 + 2u^3 + 2u^4 (split-BF16)
//   |G - P|      <= gamma * sum|products|, gamma = 2 m 2^-24 / (1 - 2 m 2^-24)
// (P: the exact sum of the products; gamma allows every FP32 addition of the
// accumulation, in any order, an error of two units in the last place), so
//   |2G - 2a'.b'| <= e_g (|a'|^2 + |b'|^2),  e_g = e_p + gamma (1+u)^2 (1+2u).
// Per row the kernel keeps alpha = RU((1+C)|a'|^2) and alpha' = RD((1-C)|a'|^2)
// in FP32, C = 1.01 (e_g + 2^-24) + 2^-40, and decides
//   certain accept:  fma(-2, G, alpha_p + alpha_c)  <= lo
//   certain reject:  fma(-2, G, alpha'_p + alpha'_c) > hi
//   otherwise:       band -> FP64 sequence,
// lo = RD(eps^2 / ((1+2^-23)(1+g)) (1-2^-30) - 2^-58),
// hi = RU(eps^2 / ((1-2^-24)(1-g)) (1+2^-30) + 2^-58),
// g = (ds + 4) 2^-52 bounding the FP64 sequence's relative error against the
// exact squared distance, 2^-58 covering FP32 underflow of tiny products.
// C absorbs the FP32 roundings of the two threshold sums, the FP64 norm and
// the centring error (|(a'-b') - (x_a - x_b)| <= 2^-53 (|a'|+|b'|)(1+2^-53)).
// Rows whose |a'|^2 exceeds 2^100 (or is not finite) get NaN terms, so every
// pair with them falls in the band; the filter is used for
// 2^-50 <= eps^2 <= 2^60 only (msj_api.hip runs the plain form otherwise).
//
// Layout.  Nodes are padded to 32-row boundaries (padded row space P).  The
// BF16 planes are stored in the operand order of V_MFMA_F32_32X32X16_BF16:
// for row block rb (32 rows) and k-step ks (16 dims), 64 lanes x 16 bytes,
// lane = row%32 + 32*(k%16 / 8), 8 consecutive dims per lane.  A tile slab is
// thus a linear copy into LDS and a linear, conflict-free read from it, and
// the same array serves as operand A (query rows) and operand B (candidates).
//
// Work.  A tile is 256 A rows of node X against 256 B rows of a node Y >= X
// of X's upper neighbour list (inside X only tiles on or above the diagonal,
// pairs p <= c).  Above d = 128 a workgroup of 8 waves computes one tile at a
// time, wave (wr, wc) a 128 x 64 block as 4 x 2 MFMA 32x32 tiles; up to
// d = 128 two workgroups of 4 waves share a CU, each computing a half tile
// (256 x 128), so one's epilogue overlaps the other's MFMAs; split-BF16 up
// to d = 128 runs quarter tiles (128 x 128, 64 x 64 per wave, 168 VGPRs),
// three workgroups per CU.  Tiles are
// grouped into super-blocks of up to 8 x 8 tiles (the work items);
// consecutive tile slots run on one XCD.  Slabs of 32 dimensions are
// copied global -> LDS directly into a double buffer: the next slab's copy
// is in flight while the current one runs on the matrix cores.  Split half
// tiles use no slabs: each wave loads its own fragments (the layout is the
// operand layout) from global memory straight into registers, the next
// k-step's loads in flight during the current one's MFMAs, with no barrier
// between waves (measured on 515,345 x 90: 87 -> 80 ms for the BF16 pass,
// 80 -> 72 ms for the cross products).
//
// Split-BF16 tiles are scored in BF16 first (high planes only, one third of
// the products): a 32 x 32 block whose pairs are all certain rejects under
// the BF16 bound (C of a BF16 accumulation of d products) is done.  Any
// other block (a pair that is not, a sampled item, NaN terms) adds the two
// cross products and is classified under the split-BF16 bound, as without
// this step: half tiles refine just those blocks, each wave loading both
// planes of the fragments they need; full tiles refine the whole 128 x 64
// block of a wave that has one.  The split bound covers the reordered
// accumulation (any order of the FP32 additions).  On this data most blocks
// pair rows that are far apart and stop after the BF16 pass.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>
#include <vector>

#include "msj_internal.h"

#pragma clang fp contract(off)

namespace msj {

namespace {

constexpr uint32_t TM = 256;     // rows per tile of the work-item grid, A and B
constexpr uint32_t KS_HALF = 8;  // k-steps (16 dims) up to which half tiles are used
constexpr uint32_t KSL = 2;      // 16-dim k-steps per LDS slab
constexpr uint32_t SBS = 8;      // tiles per side of a super-block
constexpr uint32_t MHITS = 128;  // per-wave LDS buffer of result keys
constexpr uint32_t KALIGN = 16 * KSL;

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx16 __attribute__((ext_vector_type(16)));

__device__ inline bf16x8 as_bf16x8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }

// FP32 -> BF16, round to nearest even (NaN kept quiet).
__device__ inline uint32_t bf16_bits(float f)
{
    uint32_t b = __float_as_uint(f);
    if ((b & 0x7FFFFFFFu) > 0x7F800000u)
        return (b >> 16) | 0x40u;
    return (b + 0x7FFFu + ((b >> 16) & 1u)) >> 16;
}
__device__ inline float bf16_value(uint32_t h) { return __uint_as_float(h << 16); }

// ------------------------------------------------------------ preparation
__global__ __launch_bounds__(256) void k_prow(const uint32_t *__restrict__ nstart, uint32_t M,
                                              uint32_t *__restrict__ prow)
{
    uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x < M)
        prow[x] = (nstart[x + 1] - nstart[x] + 31u) & ~31u;
}

__global__ __launch_bounds__(256) void k_u64_to_u32(const uint64_t *__restrict__ a, size_t n,
                                                    uint32_t *__restrict__ b)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        b[i] = (uint32_t)a[i];
}

__global__ __launch_bounds__(256) void k_fill_u32(uint32_t *__restrict__ p, size_t n, uint32_t v)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = v;
}

// Join position i (node order, Morton order inside a node) -> padded row.
__global__ __launch_bounds__(256) void k_pad_rows(const uint32_t *__restrict__ nstart, uint32_t M,
                                                  const uint32_t *__restrict__ pstart,
                                                  const uint32_t *__restrict__ jperm, uint32_t n,
                                                  uint32_t *__restrict__ pjoin,
                                                  uint32_t *__restrict__ porig)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    uint32_t lo = 0, hi = M;
    while (hi - lo > 1) {
        uint32_t mid = lo + (hi - lo) / 2;
        if (nstart[mid] <= i)
            lo = mid;
        else
            hi = mid;
    }
    const uint32_t r = pstart[lo] + (i - nstart[lo]);
    pjoin[r] = i;
    porig[r] = jperm[i];
}

// Column sums of xr over row blocks, then the means; every sum in a fixed
// order, so the centre is the same from run to run.  Workgroup b sums rows
// [b * rows_per, ...): thread t takes column t % ds and every Q-th row from
// row t / ds (Q = threads per column), and the Q partial sums are added in
// phase order.
__global__ __launch_bounds__(256) void k_col_partial(const double *__restrict__ xr, uint32_t n,
                                                     uint32_t ds, uint32_t rows_per,
                                                     double *__restrict__ part)
{
    __shared__ double ph[256];
    const uint32_t b = blockIdx.x, tid = threadIdx.x;
    const uint32_t r0 = b * rows_per, r1 = min(n, r0 + rows_per);
    const uint32_t Q = ds < 256 ? 256 / ds : 1u;
    for (uint32_t j0 = 0; j0 < ds; j0 += 256) {
        const uint32_t q = tid / (Q > 1 ? ds : 256u), j = j0 + tid % (Q > 1 ? ds : 256u);
        double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
        if (q < Q && j < ds) {
            uint32_t r = r0 + q;
            for (; r + 3 * Q < r1; r += 4 * Q) {
                a0 += xr[(size_t)r * ds + j];
                a1 += xr[(size_t)(r + Q) * ds + j];
                a2 += xr[(size_t)(r + 2 * Q) * ds + j];
                a3 += xr[(size_t)(r + 3 * Q) * ds + j];
            }
            for (; r < r1; r += Q)
                a0 += xr[(size_t)r * ds + j];
        }
        ph[tid] = (a0 + a1) + (a2 + a3);
        __syncthreads();
        if (tid < min(256u, ds - j0) && (Q > 1 || j0 + tid < ds)) {
            double acc = ph[tid];
            for (uint32_t k = 1; k < Q; ++k)
                acc += ph[k * ds + tid];
            part[(size_t)b * ds + j0 + tid] = acc;
        }
        __syncthreads();
    }
}

// mu[j] = (sum of the nb partial sums of column j = blockIdx.x) / n: each
// thread adds every 256th partial, then a fixed tree adds the threads' sums.
__global__ __launch_bounds__(256) void k_col_final(const double *__restrict__ part, uint32_t nb,
                                                   uint32_t ds, uint32_t n,
                                                   double *__restrict__ mu)
{
    __shared__ double ph[256];
    const uint32_t j = blockIdx.x, tid = threadIdx.x;
    double acc = 0.0;
    for (uint32_t b = tid; b < nb; b += 256)
        acc += part[(size_t)b * ds + j];
    ph[tid] = acc;
    __syncthreads();
    for (uint32_t o = 128; o > 0; o >>= 1) {
        if (tid < o)
            ph[tid] += ph[tid + o];
        __syncthreads();
    }
    if (tid == 0)
        mu[j] = ph[0] / (double)n;
}

// RPW padded rows per wave (64 / RPW lanes each, one 8-dimension group per
// lane): centred BF16 (and low BF16) fragments, and the row's threshold
// terms (alpha, alpha', |a'|^2 as FP32, 0).  The squared norm's rounding
// (any summation order) stays far inside the 2^-40 slack of the terms.
template <bool SPLIT, uint32_t RPW>
__global__ __launch_bounds__(256) void k_frag_prep(const double *__restrict__ xr, uint32_t ds,
                                                   uint32_t KS, const double *__restrict__ mu,
                                                   const uint32_t *__restrict__ pjoin,
                                                   uint32_t Ptot, double C,
                                                   uint4 *__restrict__ fh, uint4 *__restrict__ fl,
                                                   float4 *__restrict__ thr, double Cb,
                                                   float2 *__restrict__ thrb)
{
    constexpr uint32_t LPR = WAVE / RPW;  // lanes per row
    const uint32_t lane = threadIdx.x & (WAVE - 1);
    const uint32_t r = (blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE) * RPW + lane / LPR;
    const uint32_t sl = lane % LPR;
    const bool live = r < Ptot;  // whole waves stay for the shuffles
    const uint32_t j = live ? pjoin[r] : ~0u;
    const bool valid = j != ~0u;
    const uint32_t groups = KS * 2;  // 8-dim groups
    double na = 0.0;
    for (uint32_t g = sl; live && g < groups; g += LPR) {
        uint32_t hw[4] = {0, 0, 0, 0}, lw[4] = {0, 0, 0, 0};
        if (valid) {
#pragma unroll
            for (int e = 0; e < 8; ++e) {
                const uint32_t k = g * 8 + e;
                double a = 0.0;
                if (k < ds)
                    a = xr[(size_t)j * ds + k] - mu[k];
                na = fma(a, a, na);
                const uint32_t hb = bf16_bits((float)a);
                const double rem = a - (double)bf16_value(hb);  // exact (Sterbenz)
                const uint32_t lb = bf16_bits((float)rem);
                hw[e / 2] |= hb << (16 * (e & 1));
                lw[e / 2] |= lb << (16 * (e & 1));
            }
        }
        const size_t idx = ((size_t)(r / 32) * KS + g / 2) * WAVE + (r % 32) + 32 * (g & 1);
        fh[idx] = make_uint4(hw[0], hw[1], hw[2], hw[3]);
        if (SPLIT)
            fl[idx] = make_uint4(lw[0], lw[1], lw[2], lw[3]);
    }
#pragma unroll
    for (uint32_t o = LPR / 2; o > 0; o >>= 1)
        na += __shfl_xor(na, o, WAVE);
    if (live && sl == 0) {
        float4 t = make_float4(0.f, 0.f, 0.f, 0.f);
        float2 tb = make_float2(0.f, 0.f);  // the same terms under the BF16 bound
        if (valid) {
            if (na <= 0x1p100) {
                t.x = __double2float_ru((1.0 + C) * na * (1.0 + 0x1p-40));
                t.y = __double2float_rd((1.0 - C) * na * (1.0 - 0x1p-40));
                t.z = (float)na;
                tb.x = __double2float_ru((1.0 + Cb) * na * (1.0 + 0x1p-40));
                tb.y = __double2float_rd((1.0 - Cb) * na * (1.0 - 0x1p-40));
            } else {
                t.x = t.y = t.z = __builtin_nanf("");
                tb.x = tb.y = __builtin_nanf("");
            }
        }
        thr[r] = t;
        if (thrb)
            thrb[r] = tb;
    }
}

// ------------------------------------------------------------- work items
__global__ __launch_bounds__(256) void k_edge_src(const uint32_t *__restrict__ uoff, uint32_t M,
                                                  uint32_t *__restrict__ esrc)
{
    uint32_t x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= M)
        return;
    for (uint32_t e = uoff[x]; e < uoff[x + 1]; ++e)
        esrc[e] = x;
}

// Work is cut into 128 x 128 tiles of (A rows of X) x (B rows of Y), and the
// tiles into super-blocks of s x s tiles, s = min(8, tiles of X, tiles of
// Y); inside X only super-blocks and tiles on or above the diagonal.  The
// super-blocks are the work items; items are grouped by s (class 8 - s,
// largest s first) so that a tile slot t maps to its item arithmetically.
// Consecutive slots, which run concurrently on one XCD, lie in one
// super-block: 8 A tiles and 8 B tiles shared through that XCD's L2.
__global__ __launch_bounds__(256) void k_sblk_count(const uint32_t *__restrict__ nstart,
                                                    const uint32_t *__restrict__ esrc,
                                                    const uint32_t *__restrict__ uadj,
                                                    uint32_t E, uint32_t *__restrict__ cnt8,
                                                    uint64_t *__restrict__ tiles)
{
    uint32_t e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E)
        return;
    const uint32_t X = esrc[e], Y = uadj[e];
    const uint64_t TA = (nstart[X + 1] - nstart[X] + TM - 1) / TM;
    const uint64_t TB = (nstart[Y + 1] - nstart[Y] + TM - 1) / TM;
    const uint64_t sz = std::min<uint64_t>(SBS, std::min(TA, TB));
    const uint64_t SA = (TA + sz - 1) / sz, SB = (TB + sz - 1) / sz;
    const uint64_t nsb = X == Y ? SA * (SA + 1) / 2 : SA * SB;
    const uint32_t cls = SBS - (uint32_t)sz;
    for (uint32_t c = 0; c < SBS; ++c)
        cnt8[(size_t)c * E + e] = c == cls ? (uint32_t)nsb : 0u;
    tiles[e] = X == Y ? TA * (TA + 1) / 2 : TA * TB;
}

__global__ __launch_bounds__(256) void k_sblk_fill(const uint32_t *__restrict__ nstart,
                                                   const uint32_t *__restrict__ pstart,
                                                   const uint32_t *__restrict__ esrc,
                                                   const uint32_t *__restrict__ uadj,
                                                   const uint64_t *__restrict__ off8, uint32_t E,
                                                   uint64_t nitems, MItem *__restrict__ items)
{
    const uint64_t k = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nitems)
        return;
    uint64_t lo = 0, hi = (uint64_t)SBS * E;  // entry with off8[i] <= k < off8[i+1]
    while (hi - lo > 1) {
        uint64_t mid = lo + (hi - lo) / 2;
        if (off8[mid] <= k)
            lo = mid;
        else
            hi = mid;
    }
    const uint32_t cls = (uint32_t)(lo / E), e = (uint32_t)(lo % E);
    const uint64_t l = k - off8[lo];
    const uint32_t X = esrc[e], Y = uadj[e];
    const uint32_t sx = nstart[X + 1] - nstart[X], sy = nstart[Y + 1] - nstart[Y];
    const uint64_t sz = SBS - cls;
    const uint64_t TA = (sx + TM - 1) / TM, TB = (sy + TM - 1) / TM;
    const uint64_t SA = (TA + sz - 1) / sz, SB = (TB + sz - 1) / sz;
    uint64_t I, J;
    if (X == Y) {
        // row I holds super-blocks J = I .. SA-1; F(I) = I SA - I (I - 1) / 2
        uint64_t ilo = 0, ihi = SA;
        while (ihi - ilo > 1) {
            uint64_t mid = ilo + (ihi - ilo) / 2;
            if (mid * SA - mid * (mid - 1) / 2 <= l)
                ilo = mid;
            else
                ihi = mid;
        }
        I = ilo;
        J = I + (l - (I * SA - I * (I - 1) / 2));
    } else {
        I = l / SB;
        J = l % SB;
    }
    MItem w;
    w.qbeg = pstart[X] + (uint32_t)(I * sz * TM);
    w.qend = pstart[X] + sx;
    w.cbeg = pstart[Y] + (uint32_t)(J * sz * TM);
    w.cend = pstart[Y] + sy;
    w.self = (uint32_t)sz | (X == Y ? 0x100u : 0u) | (X == Y && I == J ? 0x200u : 0u);
    items[k] = w;
}

// --------------------------------------------------------------- the filter
// ctr: [0] result keys, [1] band entries, [2] max error (double bits),
// [3] verified pairs, [4] violations, [5] error sum (double), [6] pairs
// decided in FP64, [7] class-0 band entries, [8] wave blocks that took the
// cross products, [9] verified class-3 pairs.  band_budget != 0: workgroups
// stop taking items once [7] exceeds it (the BF16 bound was too loose; the
// host reruns with split-BF16).
//
// Band entries: (join position p << 32 | class << 30 | join position c) and
// the filter score.  Class 0: band pair, decided in FP64.  With DIAG
// (count_work): class 1, a certain accept (every one is verified); class 2,
// a certain reject of a sampled work item (verified); class 3 (REF), a pair
// of a sampled work item that the first, BF16 pass scores as a certain
// reject under the BF16 bound -- the test that lets a block skip the cross
// products (verified).
//
// Tile.  A workgroup (NW waves) computes a TMA x TN tile (256 x 256, or a
// 256 x 128 half tile with HALF, or a 128 x 128 quarter tile with DIRECT;
// REF: split-BF16, BF16 pass first); wave (wr, wc) the WRS x 64 block of
// rows WRS wr.. and columns 64 wc.., as NBI x 2 MFMA 32x32 tiles (128 or,
// DIRECT, 64 FP32 accumulators per lane).  Per slab of KL
// k-steps the workgroup copies the row blocks of fragments (8 of A, TN / 32
// of B; the high plane) from global memory straight into LDS, and each wave
// reads its 4 + 2 fragments per k-step from there; the next slab's copy
// runs into the other LDS buffer while the matrix cores work on the current
// one.  Split half tiles (DIRECT) load the 4 + 2 fragments per k-step into
// registers instead.
//
// Epilogue.  The A and B rows' terms are staged per tile; rows outside the
// item (padding, the next node) get +inf terms, so the fast pass -- one add,
// one fma, one compare per pair -- rejects them without a validity test.
// Only a 32-row block that holds a pair that is not a certain reject (or
// NaN-scored) takes the full pass: it tests validity explicitly, collects
// certain accepts into a lane mask and stages band entries in LDS.  Half
// tiles classify the blocks one at a time (unrolled), then emit band entries
// (a band block's accumulators go through a per-lane LDS column) and
// certain accepts in a loop that is not unrolled; full tiles run the fast
// pass over all four blocks first.
template <bool SPLIT, bool DIAG, bool HALF, bool REF>
__global__ __launch_bounds__(HALF ? 256 : 512, HALF ? (SPLIT && REF ? 3 : 2) : 1) void k_mfma_join(
    const uint4 *__restrict__ fh, const uint4 *__restrict__ fl, uint32_t KS,
    const float4 *__restrict__ thr, const float2 *__restrict__ thrb,
    const uint32_t *__restrict__ porig,
    const uint32_t *__restrict__ pjoin, const MItem *__restrict__ items,
    const SlotMap *__restrict__ smp,
    uint64_t per_xcd, float lo, float hi, int qbits, unsigned long long *__restrict__ out,
    unsigned long long cap, unsigned long long *__restrict__ ctr,
    unsigned long long *__restrict__ band, float *__restrict__ bscore, unsigned long long bcap,
    uint32_t sample_every, unsigned long long band_budget)
{
    constexpr uint32_t NP = SPLIT ? 2 : 1;
    constexpr uint32_t TN = HALF ? TM / 2 : TM;   // B rows of a workgroup tile
    constexpr uint32_t KL = SPLIT && HALF ? 1 : KSL;  // k-steps per LDS slab
    constexpr uint32_t RBA = TM / 32;             // A row blocks per workgroup tile
    constexpr uint32_t RBB = TN / 32;             // B row blocks per workgroup tile
    constexpr uint32_t SLA = RBA * KL * WAVE;     // uint4 per plane, A
    constexpr uint32_t SLB = RBB * KL * WAVE;     // uint4 per plane, B
    constexpr uint32_t UNITS = NP * (RBA + RBB) * KL;  // 1 KB copies per slab
    constexpr uint32_t NW = HALF ? 4 : 8;         // waves per workgroup
    constexpr uint32_t BHITS = 64;                // per-wave band staging
    static_assert(UNITS % NW == 0, "every wave issues the same number of copies");
    static_assert(!REF || SPLIT, "BF16-first refinement: split-BF16 only");
    // split half tiles: both passes read each wave's fragments straight from
    // global memory into registers (no LDS slab, no barrier per k-step)
    constexpr bool DIRECT = SPLIT && HALF && REF;
    // 32-row blocks per wave (DIRECT: 64 x 64 wave blocks, 128 x 128 quarter
    // tiles, fewer registers: three workgroups per CU) and the A rows of a
    // workgroup tile
    constexpr uint32_t NBI = DIRECT ? 2 : 4;
    constexpr uint32_t WRS = 32 * NBI;
    constexpr uint32_t TMA = 2 * WRS;
    // two slab buffers: the next slab is copied (global -> LDS, no VGPRs)
    // while the matrix cores work on the current one
    __shared__ uint4 sA[2][NP][DIRECT ? 1 : SLA];
    __shared__ uint4 sB[2][NP][DIRECT ? 1 : SLB];
    __shared__ float4 tA[TMA], tB[TN];
    __shared__ __align__(16) float tAy[TMA];  // tA[.].y, four rows per 16-byte read
    // REF: the rows' reject terms under the BF16 bound (first pass)
    __shared__ __align__(16) float tAyb[REF ? TMA : 4], tByb[REF ? TN : 4];
    __shared__ uint32_t oA[TMA], oB[TN], jA[TMA], jB[TN];
    __shared__ unsigned long long hsh[NW][MHITS];
    __shared__ unsigned long long bsh[NW][BHITS];
    __shared__ float ssh[NW][BHITS];
    __shared__ uint32_t s_stop;
    if (threadIdx.x == 0)
        s_stop = 0;

    const uint32_t tid = threadIdx.x, lane = tid & (WAVE - 1);
    // half tiles: the wave index in a scalar register
    const uint32_t wv = HALF ? __builtin_amdgcn_readfirstlane(tid / WAVE) : tid / WAVE;
    const uint32_t wr = HALF ? wv >> 1 : wv >> 2, wc = HALF ? wv & 1 : wv & 3, h = lane / 32;
    const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (WAVE - lane));
    const float INF = __builtin_inff();
    unsigned long long *hb = hsh[wv];
    unsigned long long *bb = bsh[wv];
    float *sb = ssh[wv];
    uint32_t fill = 0, bfill = 0, b0fill = 0;  // wave-uniform
    auto wave_sync = [&]() {
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    };
    auto flush = [&]() {  // result keys: one atomic per flush
        wave_sync();
        unsigned long long base = 0;
        if (lane == 0)
            base = atomicAdd(&ctr[0], (unsigned long long)fill);
        base = __shfl(base, 0, WAVE);
        for (uint32_t e = lane; e < fill; e += WAVE)
            if (base + e < cap)
                out[base + e] = hb[e];
        wave_sync();
        fill = 0;
    };
    auto bflush = [&]() {  // band entries
        wave_sync();
        unsigned long long base = 0;
        if (lane == 0) {
            base = atomicAdd(&ctr[1], (unsigned long long)bfill);
            if (b0fill)
                atomicAdd(&ctr[7], (unsigned long long)b0fill);
        }
        base = __shfl(base, 0, WAVE);
        for (uint32_t e = lane; e < bfill; e += WAVE)
            if (base + e < bcap) {
                band[base + e] = bb[e];
                bscore[base + e] = sb[e];
            }
        wave_sync();
        bfill = 0;
        b0fill = 0;
    };

    // Half-tile slots in a grid-stride loop over virtual block ids; id v
    // runs half slot (v % 8) * per_xcd + v / 8 (tile slot = half slot / 2,
    // B half = half slot % 2), so the slots of one XCD (the hardware deals
    // workgroups out round robin) are consecutive: one super-block.  Slots
    // outside their item (ragged node edges, below the diagonal) are
    // skipped.
    //
    // Tiles are software-pipelined: during the last slab of a tile the
    // workgroup already copies the next tile's first slab into the free LDS
    // buffer and loads its row terms into registers, so neither waits on
    // global memory when the next tile starts.
    const uint64_t vtotal = 8ull * per_xcd;
    struct Slot {
        uint64_t it;
        uint32_t ta, tbk, half, aq;
        bool in;
    };
    // (DIRECT: quarter slots, B half = u % 2, A half aq = u / 2 % 2)
    auto slot_of = [&](uint64_t v) {
        Slot S;
        const uint64_t u = (v % 8u) * per_xcd + v / 8u;
        const uint64_t t = DIRECT ? u >> 2 : HALF ? u >> 1 : u;
        S.half = HALF ? (uint32_t)(u & 1u) : 0u;
        S.aq = DIRECT ? (uint32_t)((u >> 1) & 1u) : 0u;
        // the slot map stays in memory (scalar loads once per tile): held
        // in SGPRs for the whole kernel it pushed other values into spills
        const SlotMap &sm = *smp;
        S.in = t < sm.slot_base[SBS];
        uint32_t cls = 0;
        while (S.in && cls + 1 < SBS && t >= sm.slot_base[cls + 1])
            ++cls;
        const uint32_t sz = SBS - cls;
        const uint64_t local = S.in ? t - sm.slot_base[cls] : 0;
        S.it = sm.item_base[cls] + local / (sz * sz);
        const uint32_t r = (uint32_t)(local % (sz * sz));
        S.ta = r / sz;
        S.tbk = r % sz;
        return S;
    };
    // first A row and first B row of a half (quarter) tile
    auto a0_of = [&](const Slot &S, const MItem &w) { return w.qbeg + S.ta * TM + S.aq * TMA; };
    auto b0_of = [&](const Slot &S, const MItem &w) { return w.cbeg + S.tbk * TM + S.half * TN; };
    // (a diagonal tile's quarter of A rows 128.. against B rows ..127 lies
    // below the diagonal)
    auto valid = [&](const Slot &S, const MItem &w) {
        return S.in && a0_of(S, w) < w.qend && b0_of(S, w) < w.cend &&
               !((w.self & 0x200u) && S.tbk < S.ta) &&
               !(DIRECT && (w.self & 0x200u) && S.tbk == S.ta && S.aq > S.half);
    };
    // slab copy into buffer `buf`: unit u = ((plane * (RBA + RBB)) + block)
    // * KL + kstep, blocks 0..RBA-1 of A then RBB of B, is one wave-wide
    // 16-byte-per-lane load straight into LDS (the lane layout of a fragment
    // block is its memory layout); wave wv issues units wv, wv + NW, ...
    // (np = 1: the high plane only, the first pass of REF)
    auto copy_slab = [&](uint32_t rbA, uint32_t rbB, uint32_t k0, uint32_t buf, uint32_t np) {
#pragma unroll
        for (uint32_t m = 0; m < UNITS / NW; ++m) {
            if (m * NP >= np * (UNITS / NW))
                break;  // units of the low plane (plane is the slowest index)
            const uint32_t u = wv + NW * m;
            const uint32_t kk = u % KL, blk = (u / KL) % (RBA + RBB), pl = u / (KL * (RBA + RBB));
            const bool op = blk >= RBA;
            const uint32_t b = op ? blk - RBA : blk;
            const uint4 *src =
                (pl ? fl : fh) + ((size_t)((op ? rbB : rbA) + b) * KS + k0 + kk) * WAVE + lane;
            uint4 *dst = (op ? &sB[buf][pl][0] : &sA[buf][pl][0]) + (b * KL + kk) * WAVE;
            __builtin_amdgcn_global_load_lds((const void *)src,
                                             (void __attribute__((address_space(3))) *)dst, 16,
                                             0, 0);
        }
    };
    // row terms: thread k < TM loads A row k; B row k is loaded by thread k
    // (half tiles: k < TN) or TM + k (full tiles); rows outside the item
    // (padding, the next node) get +inf terms
    float4 t4 = make_float4(0.f, 0.f, 0.f, 0.f), u4 = t4;
    float tyb = 0.f, uyb = 0.f;  // REF: reject terms under the BF16 bound
    uint32_t o4 = 0, j4 = 0, ou = 0, ju = 0;
    auto fetch_terms = [&](const Slot &S, const MItem &w) {
        if (tid < TMA) {
            const uint32_t ra = a0_of(S, w) + tid;
            t4 = thr[ra];
            if (REF)
                tyb = thrb[ra].y;
            if (ra >= w.qend) {
                t4 = make_float4(INF, INF, INF, INF);
                tyb = INF;
            }
            o4 = porig[ra];
            j4 = pjoin[ra];
        }
        if (HALF ? tid < TN : tid >= TM) {
            const uint32_t rb = b0_of(S, w) + (HALF ? tid : tid - TM);
            u4 = thr[rb];
            if (REF)
                uyb = thrb[rb].y;
            if (rb >= w.cend) {
                u4 = make_float4(INF, INF, INF, INF);
                uyb = INF;
            }
            ou = porig[rb];
            ju = pjoin[rb];
        }
    };
    // planes copied for a tile's first pass (REF: the high plane only)
    constexpr uint32_t NP0 = REF ? 1 : NP;

    // the first tile
    uint64_t v = blockIdx.x;
    Slot cs = {0, 0, 0, 0, 0, false};
    MItem w = {0, 0, 0, 0, 0};
    bool have = false;
    for (; v < vtotal; v += gridDim.x) {
        cs = slot_of(v);
        if (!cs.in)
            continue;
        w = items[cs.it];
        if (valid(cs, w)) {
            have = true;
            break;
        }
    }
    uint32_t buf = 0;  // LDS buffer of the current slab
    if (have) {
        fetch_terms(cs, w);
        if (!DIRECT)
            copy_slab(a0_of(cs, w) / 32, b0_of(cs, w) / 32, 0, 0, NP0);
    }
    while (have) {
        __syncthreads();  // the previous tile is done with tA .. jB
        if (tid < TMA) {
            tA[tid] = t4;
            oA[tid] = o4;
            jA[tid] = j4;
            tAy[tid] = t4.y;
            if (REF)
                tAyb[tid] = tyb;
        }
        if (HALF ? tid < TN : tid >= TM) {
            const uint32_t k = HALF ? tid : tid - TM;
            tB[k] = u4;
            oB[k] = ou;
            jB[k] = ju;
            if (REF)
                tByb[k] = uyb;
        }
        // the band counter is read every 32nd slot of a workgroup: one global
        // address polled by every workgroup on every tile is a hot spot
        if (tid == 0 && band_budget != 0 && ((v / gridDim.x) & 31u) == 0)
            s_stop = __hip_atomic_load(&ctr[7], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) >
                     band_budget;
        // a row of the tile with NaN terms (non-finite or huge values) sends
        // every 32-row block to the full pass
        const bool tile_nan = __syncthreads_or((tid < TMA && (t4.x != t4.x || t4.y != t4.y)) ||
                                               ((HALF ? tid < TN : tid >= TM) &&
                                                (u4.x != u4.x || u4.y != u4.y))) != 0;
        if (s_stop)
            break;  // uniform over the workgroup
        const uint64_t it = cs.it;
        const uint32_t qbeg = a0_of(cs, w), cb = b0_of(cs, w);
        const bool sampled = DIAG && (it % sample_every == 0);
        const bool self = (w.self & 0x100u) != 0;
        const uint32_t rbA = qbeg / 32, rbB = cb / 32;
        // the next slot's item, loaded now, tested before the last slab
        uint64_t vn = v + gridDim.x;
        Slot ns = {0, 0, 0, 0, 0, false};
        MItem nw = {0, 0, 0, 0, 0};
        if (vn < vtotal) {
            ns = slot_of(vn);
            if (ns.in)
                nw = items[ns.it];
        }
        bool nhave = false;
        // the next valid slot (a skipped slot costs one more item load)
        auto find_next = [&]() {
            while (vn < vtotal) {
                if (valid(ns, nw)) {
                    nhave = true;
                    break;
                }
                vn += gridDim.x;
                if (vn >= vtotal)
                    break;
                ns = slot_of(vn);
                if (ns.in)
                    nw = items[ns.it];
            }
        };

        floatx16 acc[NBI][2];
#pragma unroll
        for (int bi = 0; bi < (int)NBI; ++bi)
#pragma unroll
            for (int bj = 0; bj < 2; ++bj)
#pragma unroll
                for (int e = 0; e < 16; ++e)
                    acc[bi][bj][e] = 0.f;
        // LDS byte addresses of this lane's first A and B fragments (k-step 0,
        // buffer 0); fragment (block, k-step) lies (block * KL + kstep) KiB
        // further, the low plane SLA * 16 (A) or SLB * 16 (B) bytes, buffer 1
        // NP times that.
        const uint32_t ldsA = (uint32_t)(uintptr_t)(
            (__attribute__((address_space(3))) uint4 *)&sA[0][0][(4 * wr * KL) * WAVE + lane]);
        const uint32_t ldsB = (uint32_t)(uintptr_t)(
            (__attribute__((address_space(3))) uint4 *)&sB[0][0][(2 * wc * KL) * WAVE + lane]);
        static_assert(SPLIT && HALF ? (KL == 1 && SLA * 16 == 8192 && SLB * 16 == 4096)
                                    : (KL == 2 && SLA * 16 == 16384 && SLB * 16 == TN * 64),
                      "fragment read offsets below");
        bool wneed = true;  // REF: this wave's 128 x 64 block needs the cross products
        // One pass over the k-slabs.  PH 0: the only pass; REF: PH 1 the high
        // planes' products (BF16), PH 2 the cross products of a tile that
        // needs them.
        auto kpass = [&](auto phase) __attribute__((always_inline)) {
            constexpr int PH = decltype(phase)::value;
            for (uint32_t k0 = 0; k0 < KS; k0 += KL, buf ^= 1u) {
                // this wave's copies of the slab have landed; after the barrier,
                // everyone's have, and nobody still reads the other buffer
                asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
                __syncthreads();
                if (k0 + KL < KS) {
                    // in flight during this slab's MFMAs
                    copy_slab(rbA, rbB, k0 + KL, buf ^ 1u, PH == 1 ? 1u : NP);
                } else if (PH == 0) {
                    // last slab: find the next tile (a skipped slot costs one more
                    // item load) and start its first slab and row terms
                    while (vn < vtotal) {
                        if (valid(ns, nw)) {
                            nhave = true;
                            break;
                        }
                        vn += gridDim.x;
                        if (vn >= vtotal)
                            break;
                        ns = slot_of(vn);
                        if (ns.in)
                            nw = items[ns.it];
                    }
                    if (nhave) {
                        copy_slab(a0_of(ns, nw) / 32, b0_of(ns, nw) / 32, 0, buf ^ 1u, NP0);
                        fetch_terms(ns, nw);
                    }
                }
                // the second REF pass: only waves whose block needs the cross
                // products read fragments and issue MFMAs (the others free the
                // matrix cores of their SIMD); all copy and synchronise
#pragma unroll
                for (uint32_t kk = 0; kk < (PH == 2 && !wneed ? 0u : KL); ++kk) {
                    // the k-step's fragments in one block, waited for inside it
                    // (outputs early-clobber: no read can overwrite an address);
                    // the copy of the next slab stays in flight
                    const uint32_t va = ldsA + buf * (NP * SLA * 16) + kk * 1024;
                    const uint32_t vb = ldsB + buf * (NP * SLB * 16) + kk * 1024;
                    uint4 a0, a1, a2, a3, b0, b1, l0, l1, l2, l3, m0, m1;
                    if (SPLIT && HALF && PH == 1) {  // high plane, blocks 1 KiB apart
                        asm volatile(
                            "ds_read_b128 %0, %6\n"
                            "ds_read_b128 %1, %6 offset:1024\n"
                            "ds_read_b128 %2, %6 offset:2048\n"
                            "ds_read_b128 %3, %6 offset:3072\n"
                            "ds_read_b128 %4, %7\n"
                            "ds_read_b128 %5, %7 offset:1024\n"
                            "s_waitcnt lgkmcnt(0)"
                            : "=&v"(a0), "=&v"(a1), "=&v"(a2), "=&v"(a3), "=&v"(b0), "=&v"(b1)
                            : "v"(va), "v"(vb)
                            : "memory");
                        l0 = l1 = l2 = l3 = m0 = m1 = make_uint4(0, 0, 0, 0);
                    } else if (SPLIT && HALF) {  // one k-step per slab: blocks 1 KiB apart
                        asm volatile(
                            "ds_read_b128 %0, %12\n"
                            "ds_read_b128 %1, %12 offset:1024\n"
                            "ds_read_b128 %2, %12 offset:2048\n"
                            "ds_read_b128 %3, %12 offset:3072\n"
                            "ds_read_b128 %4, %13\n"
                            "ds_read_b128 %5, %13 offset:1024\n"
                            "ds_read_b128 %6, %12 offset:8192\n"
                            "ds_read_b128 %7, %12 offset:9216\n"
                            "ds_read_b128 %8, %12 offset:10240\n"
                            "ds_read_b128 %9, %12 offset:11264\n"
                            "ds_read_b128 %10, %13 offset:4096\n"
                            "ds_read_b128 %11, %13 offset:5120\n"
                            "s_waitcnt lgkmcnt(0)"
                            : "=&v"(a0), "=&v"(a1), "=&v"(a2), "=&v"(a3), "=&v"(b0), "=&v"(b1),
                              "=&v"(l0), "=&v"(l1), "=&v"(l2), "=&v"(l3), "=&v"(m0), "=&v"(m1)
                            : "v"(va), "v"(vb)
                            : "memory");
                    } else if (SPLIT && PH != 1) {  // two k-steps per slab: blocks 2 KiB apart
                        asm volatile(
                            "ds_read_b128 %0, %12\n"
                            "ds_read_b128 %1, %12 offset:2048\n"
                            "ds_read_b128 %2, %12 offset:4096\n"
                            "ds_read_b128 %3, %12 offset:6144\n"
                            "ds_read_b128 %4, %13\n"
                            "ds_read_b128 %5, %13 offset:2048\n"
                            "ds_read_b128 %6, %12 offset:16384\n"
                            "ds_read_b128 %7, %12 offset:18432\n"
                            "ds_read_b128 %8, %12 offset:20480\n"
                            "ds_read_b128 %9, %12 offset:22528\n"
                            "ds_read_b128 %10, %13 offset:16384\n"
                            "ds_read_b128 %11, %13 offset:18432\n"
                            "s_waitcnt lgkmcnt(0)"
                            : "=&v"(a0), "=&v"(a1), "=&v"(a2), "=&v"(a3), "=&v"(b0), "=&v"(b1),
                              "=&v"(l0), "=&v"(l1), "=&v"(l2), "=&v"(l3), "=&v"(m0), "=&v"(m1)
                            : "v"(va), "v"(vb)
                            : "memory");
                    } else {
                        asm volatile(
                            "ds_read_b128 %0, %6\n"
                            "ds_read_b128 %1, %6 offset:2048\n"
                            "ds_read_b128 %2, %6 offset:4096\n"
                            "ds_read_b128 %3, %6 offset:6144\n"
                            "ds_read_b128 %4, %7\n"
                            "ds_read_b128 %5, %7 offset:2048\n"
                            "s_waitcnt lgkmcnt(0)"
                            : "=&v"(a0), "=&v"(a1), "=&v"(a2), "=&v"(a3), "=&v"(b0), "=&v"(b1)
                            : "v"(va), "v"(vb)
                            : "memory");
                        l0 = l1 = l2 = l3 = m0 = m1 = make_uint4(0, 0, 0, 0);
                    }
                    const bf16x8 ah[4] = {as_bf16x8(a0), as_bf16x8(a1), as_bf16x8(a2),
                                          as_bf16x8(a3)};
                    const bf16x8 bh[2] = {as_bf16x8(b0), as_bf16x8(b1)};
                    if (PH != 2) {  // the second REF pass adds the cross products only
#pragma unroll
                        for (int bi = 0; bi < 4; ++bi)
#pragma unroll
                            for (int bj = 0; bj < 2; ++bj)
                                acc[bi][bj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                                    ah[bi], bh[bj], acc[bi][bj], 0, 0, 0);
                    }
                    if (SPLIT && PH != 1) {
                        const bf16x8 al[4] = {as_bf16x8(l0), as_bf16x8(l1), as_bf16x8(l2),
                                              as_bf16x8(l3)};
                        const bf16x8 bl[2] = {as_bf16x8(m0), as_bf16x8(m1)};
#pragma unroll
                        for (int bi = 0; bi < 4; ++bi)
#pragma unroll
                            for (int bj = 0; bj < 2; ++bj)
                                acc[bi][bj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                                    ah[bi], bl[bj], acc[bi][bj], 0, 0, 0);
#pragma unroll
                        for (int bi = 0; bi < 4; ++bi)
#pragma unroll
                            for (int bj = 0; bj < 2; ++bj)
                                acc[bi][bj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                                    al[bi], bh[bj], acc[bi][bj], 0, 0, 0);
                    }
                }
            }

        };
        uint32_t wmask = 0xFFu;  // split half tiles: the blocks that were refined
        bool refined = !REF;
        if constexpr (!REF) {
            kpass(std::integral_constant<int, 0>());
        } else {
            if constexpr (DIRECT) {
                // k-step k's 4 + 2 high-plane fragments of this wave, one
                // 16-byte load per lane each; two k-steps in flight (the
                // loads of k + 1 are issued before the MFMAs of k)
                const uint4 *pa = fh + ((size_t)(rbA + NBI * wr) * KS) * WAVE + lane;
                const uint4 *pb = fh + ((size_t)(rbB + 2 * wc) * KS) * WAVE + lane;
                auto ld = [&](uint32_t k, uint4 (&a)[NBI], uint4 (&b)[2]) __attribute__((always_inline)) {
#pragma unroll
                    for (int bi = 0; bi < (int)NBI; ++bi)
                        a[bi] = pa[((size_t)bi * KS + k) * WAVE];
#pragma unroll
                    for (int bj = 0; bj < 2; ++bj)
                        b[bj] = pb[((size_t)bj * KS + k) * WAVE];
                };
                auto mm = [&](const uint4 (&a)[NBI], const uint4 (&b)[2]) __attribute__((always_inline)) {
#pragma unroll
                    for (int bi = 0; bi < (int)NBI; ++bi)
#pragma unroll
                        for (int bj = 0; bj < 2; ++bj)
                            acc[bi][bj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                                as_bf16x8(a[bi]), as_bf16x8(b[bj]), acc[bi][bj], 0, 0, 0);
                };
                uint4 a0[NBI], b0[2], a1[NBI], b1[2];
                // waves in the BF16 pass issue at the lower priority, so a
                // wave in its block test, cross products or epilogue (the
                // latency-bound tail of a tile) issues first (64.96 -> 64.61
                // ms on 515,345 x 90)
                __builtin_amdgcn_s_setprio(0);
                ld(0, a0, b0);
                for (uint32_t k0 = 0; k0 < KS; k0 += 2) {  // KS is even
                    ld(k0 + 1, a1, b1);
                    mm(a0, b0);
                    if (k0 + 2 < KS)
                        ld(k0 + 2, a0, b0);
                    mm(a1, b1);
                }
                __builtin_amdgcn_s_setprio(1);
                // the next tile (its item was loaded at this tile's start):
                // its row terms arrive during this tile's test and epilogue
                find_next();
                if (nhave)
                    fetch_terms(ns, nw);
            } else {
                kpass(std::integral_constant<int, 1>());
            }
            // BF16 scores of the tile: if every pair is a certain reject
            // under the BF16 bound, the tile is done; otherwise (or when
            // sampled or NaN-scored) the cross products follow and the tile
            // is classified under the split-BF16 bound
            float tbb[2];
#pragma unroll
            for (int bj = 0; bj < 2; ++bj)
                tbb[bj] = tByb[64 * wc + 32 * bj + (lane & 31)];
            // Per 32 x 32 block (bit 2 bi + bj) the smallest reject score of
            // the BF16 pass: the lane's column c takes the smallest
            // RN(alpha'_b,p - 2 G) of its 16 rows, then adds alpha'_b,c (RN
            // is monotone: no pair of the column scores less; the bound of
            // these terms allows the rounding of alpha'_b,p - 2 G).  A block
            // whose smallest score exceeds hi holds certain rejects only
            // (half tiles refine the other blocks one by one; full tiles the
            // wave's whole 128 x 64 block).  Scores of padding rows and of
            // rows outside the item are +inf.
            float m = INF;
            uint32_t nmask = 0;
            auto min3 = [](float a, float b, float c) { return fminf(fminf(a, b), c); };
#pragma unroll
            for (int bi = 0; bi < (int)NBI; ++bi) {
                float s[2][16];
#pragma unroll
                for (int eg = 0; eg < 4; ++eg) {
                    const float4 ay4 = *(const float4 *)&tAyb[WRS * wr + 32 * bi + 8 * eg + 4 * h];
                    const float ay[4] = {ay4.x, ay4.y, ay4.z, ay4.w};
#pragma unroll
                    for (int ei = 0; ei < 4; ++ei)
#pragma unroll
                        for (int bj = 0; bj < 2; ++bj)
                            s[bj][4 * eg + ei] = fmaf(-2.f, acc[bi][bj][4 * eg + ei], ay[ei]);
                }
#pragma unroll
                for (int bj = 0; bj < 2; ++bj) {
                    const float *q = s[bj];
                    const float x0 = min3(q[0], q[1], q[2]), x1 = min3(q[3], q[4], q[5]);
                    const float x2 = min3(q[6], q[7], q[8]), x3 = min3(q[9], q[10], q[11]);
                    const float x4 = min3(q[12], q[13], q[14]);
                    const float sc = fminf(min3(x0, x1, x2), min3(x3, x4, q[15])) + tbb[bj];
                    if (HALF)
                        nmask |= (__builtin_amdgcn_ballot_w64(!(sc > hi)) != 0 ? 1u : 0u)
                                 << (2 * bi + bj);
                    m = fminf(m, sc);
                }
            }
            if constexpr (DIAG) {
                // sampled: every pair the BF16 score calls a certain reject
                // goes to the FP64 check as class 3 before the cross products
                // overwrite the accumulators
                if (sampled) {
                    const bool dg = self && cb < qbeg + TMA && qbeg < cb + TN;
#pragma unroll
                    for (int bi = 0; bi < (int)NBI; ++bi)  // constant register indices
#pragma unroll
                        for (int e = 0; e < 16; ++e) {
                            const uint32_t il = WRS * wr + 32 * bi + 8 * (e / 4) + 4 * h + (e % 4);
#pragma unroll
                            for (int bj = 0; bj < 2; ++bj) {
                                const uint32_t jl = 64 * wc + 32 * bj + (lane & 31);
                                const uint32_t p = qbeg + il, c = cb + jl;
                                const bool valid =
                                    p < w.qend && c < w.cend && (!dg || p <= c);
                                const float sc =
                                    fmaf(-2.f, acc[bi][bj][e], tAyb[il] + tByb[jl]);
                                const bool push = valid && sc > hi;
                                const uint64_t mb = __builtin_amdgcn_ballot_w64(push);
                                if (!mb)
                                    continue;
                                const uint32_t k = (uint32_t)__popcll(mb);
                                if (bfill + k > BHITS)
                                    bflush();
                                if (push) {
                                    const uint32_t idx = bfill + (uint32_t)__popcll(mb & lt);
                                    bb[idx] = ((unsigned long long)jA[il] << 32) |
                                              (3ull << 30) | jB[jl];
                                    sb[idx] = sc;
                                }
                                bfill += k;
                            }
                        }
                }
            }
            const bool need = sampled || tile_nan || !(m > hi);
            wneed = __builtin_amdgcn_ballot_w64(need) != 0;
            if constexpr (DIRECT) {
                // the 32 x 32 blocks to refine (all of a sampled or NaN-scored
                // tile): each wave loads both planes of the fragments its
                // blocks need straight into registers, k-step by k-step, and
                // adds the two cross products (no LDS, no barrier)
                const uint32_t wm = (sampled || tile_nan) ? (1u << (2 * NBI)) - 1u : nmask;
                wmask = wm;
                wneed = wm != 0;
                if (wm != 0) {
                    if (lane == 0)  // 32 x 32 blocks that take the cross products
                        atomicAdd(&ctr[8], (unsigned long long)__popc(wm));
                    uint32_t rows = 0, cols = 0;  // the wave's 32-row blocks and columns in wm
#pragma unroll
                    for (uint32_t bi = 0; bi < NBI; ++bi)
                        if ((wm >> (2 * bi)) & 3u)
                            rows |= 1u << bi;
#pragma unroll
                    for (uint32_t bj = 0; bj < 2; ++bj)
                        if (wm & (0x55u << bj))
                            cols |= 1u << bj;
                    const size_t ra = (size_t)(rbA + NBI * wr) * KS, rb = (size_t)(rbB + 2 * wc) * KS;
                    for (uint32_t k = 0; k < KS; ++k) {
                        uint4 ah[NBI], al[NBI], bh[2], bl[2];
#pragma unroll
                        for (int bi = 0; bi < (int)NBI; ++bi) {
                            ah[bi] = al[bi] = make_uint4(0, 0, 0, 0);
                            if ((rows >> bi) & 1u) {
                                const size_t o = (ra + (size_t)bi * KS + k) * WAVE + lane;
                                ah[bi] = fh[o];
                                al[bi] = fl[o];
                            }
                        }
#pragma unroll
                        for (int bj = 0; bj < 2; ++bj) {
                            bh[bj] = bl[bj] = make_uint4(0, 0, 0, 0);
                            if ((cols >> bj) & 1u) {
                                const size_t o = (rb + (size_t)bj * KS + k) * WAVE + lane;
                                bh[bj] = fh[o];
                                bl[bj] = fl[o];
                            }
                        }
#pragma unroll
                        for (int bi = 0; bi < (int)NBI; ++bi)
#pragma unroll
                            for (int bj = 0; bj < 2; ++bj)
                                if ((wm >> (2 * bi + bj)) & 1u)
                                    acc[bi][bj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                                        as_bf16x8(ah[bi]), as_bf16x8(bl[bj]), acc[bi][bj], 0, 0, 0);
#pragma unroll
                        for (int bi = 0; bi < (int)NBI; ++bi)
#pragma unroll
                            for (int bj = 0; bj < 2; ++bj)
                                if ((wm >> (2 * bi + bj)) & 1u)
                                    acc[bi][bj] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(
                                        as_bf16x8(al[bi]), as_bf16x8(bh[bj]), acc[bi][bj], 0, 0, 0);
                    }
                    refined = true;
                }
            } else if (__syncthreads_or(need) != 0) {
                if (wneed && lane == 0)  // 32 x 32 blocks that take the cross products
                    atomicAdd(&ctr[8], 8ull);
                // both planes again from the first slab (buffer buf is free:
                // the last slab used the other one)
                copy_slab(rbA, rbB, 0, buf, NP);
                kpass(std::integral_constant<int, 2>());
                refined = true;
            }
        }

        // a wave whose block the BF16 pass settled holds certain rejects only
        if (refined && wneed) {
            // ---- epilogue.  C[i][j] of MFMA tile (bi, bj) is in lane
            // j + 32h, element e, with i = 8 (e / 4) + 4 h + e % 4.
            // a half tile crossing the diagonal of a node with itself: pairs
            // p <= c only (there cb >= qbeg: doff = cb - qbeg is 0 or TN)
            const bool diag_tile = self && cb < qbeg + TMA && qbeg < cb + TN;
            const uint32_t doff = cb - qbeg;
            float4 tb[2];
#pragma unroll
            for (int bj = 0; bj < 2; ++bj)
                tb[bj] = tB[64 * wc + 32 * bj + (lane & 31)];
            if constexpr (HALF) {
                // half tiles: one 32-row block at a time through all passes
                constexpr uint32_t NB = 2;    // 32-column MFMA blocks per wave
                using Mask = uint32_t;        // 16 NB pair slots
                const uint32_t c0 = 64 * wc;  // first tile column of this wave
                __shared__ float gst[NW][8][WAVE];  // per-lane scores of band slots
                bool cvalid[NB];
#pragma unroll
                for (int bj = 0; bj < (int)NB; ++bj)
                    cvalid[bj] = cb + c0 + 32 * bj + (lane & 31) < w.cend;
                // band entries (rare): staged slot by slot, with their scores.
                // Per 32 x 32 MFMA block and half (8 slots per lane) the lane's
                // accumulators go to its own column of the wave's LDS stage, read
                // back by slot index in a loop that is not unrolled.
                auto band_run = [&](uint32_t bi, uint32_t bj, uint32_t hf, Mask bm, Mask amk) {
                    const uint32_t jl = c0 + 32 * bj + (lane & 31);
                    const float4 tbq = tB[jl];
#pragma unroll 1
                    for (uint32_t j = 0; j < 8; ++j) {
                        const uint32_t e = 8 * hf + j;
                        const uint32_t il = WRS * wr + 32 * bi + 8 * (e / 4) + 4 * h + (e % 4);
                        const bool push = (bm >> (NB * e + bj)) & 1u;
                        const uint64_t mb = __builtin_amdgcn_ballot_w64(push);
                        if (!mb)
                            continue;
                        const uint32_t k = (uint32_t)__popcll(mb);
                        if (bfill + k > BHITS)
                            bflush();
                        const bool a_ok = (amk >> (NB * e + bj)) & 1u;
                        bool inband = push;
                        if (push) {
                            const float4 ta = tA[il];
                            const float G = gst[wv][j][lane];
                            const uint32_t p = qbeg + il, c = cb + jl;
                            const bool valid = p < w.qend && c < w.cend && (!diag_tile || p <= c);
                            const bool r_ok = fmaf(-2.f, G, ta.y + tbq.y) > hi;
                            inband = valid && !a_ok && !r_ok;
                            const uint32_t cls = DIAG ? (a_ok ? 1u : (inband ? 0u : 2u)) : 0u;
                            const uint32_t idx = bfill + (uint32_t)__popcll(mb & lt);
                            bb[idx] = ((unsigned long long)jA[il] << 32) |
                                      ((unsigned long long)cls << 30) | jB[jl];
                            sb[idx] = fmaf(-2.f, G, ta.z + tbq.z);
                        }
                        bfill += k;
                        b0fill +=
                            DIAG ? (uint32_t)__popcll(__builtin_amdgcn_ballot_w64(inband)) : k;
                    }
                };
                // slots 8 hf .. 8 hf + 7 of column block bj: bits NB e + bj
                constexpr Mask HALF_COL = 0x5555u;

                // One 32-row block (bi) at a time, so that only its accumulators are
                // read out of the matrix registers together: the classification
                // into lane masks here (unrolled, constant register indices), the
                // rare band entries and the certain accepts in the loop after it
                // (not unrolled: one copy of that code, fewer scalar registers)
                Mask am4[4] = {0u, 0u, 0u, 0u}, bm4[4] = {0u, 0u, 0u, 0u};
#pragma unroll
                for (int bi = 0; bi < (int)NBI; ++bi) {
                    // split (REF): only the refined blocks; the others hold
                    // certain rejects of the BF16 pass
                    const uint32_t bsel = REF ? (wmask >> (2 * bi)) & 3u : 3u;
                    if (bsel == 0)
                        continue;
                    // fast pass: the block's smallest reject score (in a diagonal tile,
                    // of pairs p <= c); a block whose smallest score exceeds hi holds
                    // certain rejects only.  Scores of padding rows and of rows
                    // outside the item are +inf; NaN scores (NaN terms) are caught by
                    // tile_nan.
                    auto fast = [&](auto dg) __attribute__((always_inline)) {
                        constexpr bool DT = decltype(dg)::value;
                        float m = INF;
#pragma unroll
                        for (int eg = 0; eg < 4; ++eg) {
                            const uint32_t i0 = WRS * wr + 32 * bi + 8 * eg + 4 * h;
                            const float4 ay4 = *(const float4 *)&tAy[i0];
                            const float ay[4] = {ay4.x, ay4.y, ay4.z, ay4.w};
#pragma unroll
                            for (int ei = 0; ei < 4; ++ei)
#pragma unroll
                                for (int bj = 0; bj < (int)NB; ++bj) {
                                    const float g = acc[bi][bj][4 * eg + ei];
                                    float sc = fmaf(-2.f, g, ay[ei] + tb[bj].y);
                                    if ((DT && i0 + ei > c0 + 32 * bj + (lane & 31) + doff) ||
                                        ((bsel >> bj) & 1u) == 0)
                                        sc = INF;
                                    m = fminf(m, sc);
                                }
                        }
                        return m;
                    };
                    const float mn = diag_tile ? fast(std::true_type()) : fast(std::false_type());
                    const bool need = sampled || tile_nan || !(mn > hi);
                    if (!__builtin_amdgcn_ballot_w64(need))
                        continue;
                    // full pass: the block's 16 NB pairs per lane (slot NB e + bj)
                    // classified into lane bit masks of certain accepts and band
                    // entries (with DIAG also the verified classes); bitwise logic,
                    // no branches
                    Mask amask = 0, bmask = 0;
#pragma unroll
                    for (int e = 0; e < 16; ++e) {
                        const uint32_t il = WRS * wr + 32 * bi + 8 * (e / 4) + 4 * h + (e % 4);
                        const float4 ta = tA[il];
                        const uint32_t p = qbeg + il;
                        const bool pv = p < w.qend;
#pragma unroll
                        for (int bj = 0; bj < (int)NB; ++bj) {
                            const uint32_t c = cb + c0 + 32 * bj + (lane & 31);
                            const bool valid = pv & cvalid[bj] & (!diag_tile | (p <= c)) &
                                               (((bsel >> bj) & 1u) != 0);
                            const float G = acc[bi][bj][e];
                            const bool a_ok = valid & (fmaf(-2.f, G, ta.x + tb[bj].x) <= lo);
                            const bool r_ok = fmaf(-2.f, G, ta.y + tb[bj].y) > hi;
                            const bool inband = valid & !a_ok & !r_ok;
                            const bool push = DIAG ? (inband | a_ok | (sampled & valid)) : inband;
                            amask |= (Mask)a_ok << (NB * e + bj);
                            bmask |= (Mask)push << (NB * e + bj);
                        }
                    }
                    am4[bi] = amask;
                    bm4[bi] = bmask;
                }
#pragma unroll 1
                for (int bi = 0; bi < (int)NBI; ++bi) {
                    // register arrays need constant indices
                    const Mask amask = bi == 0 ? am4[0] : bi == 1 ? am4[1] : bi == 2 ? am4[2] : am4[3];
                    const Mask bmask = bi == 0 ? bm4[0] : bi == 1 ? bm4[1] : bi == 2 ? bm4[2] : bm4[3];
                    // band entries
                    if (__builtin_amdgcn_ballot_w64(bmask != 0)) {
                        floatx16 cur[NB];
                        switch (bi) {
                        case 0:
                            cur[0] = acc[0][0], cur[1] = acc[0][1];
                            break;
                        case 1:
                            cur[0] = acc[1][0], cur[1] = acc[1][1];
                            break;
                        case 2:
                            cur[0] = acc[NBI > 2 ? 2 : 0][0], cur[1] = acc[NBI > 2 ? 2 : 0][1];
                            break;
                        default:
                            cur[0] = acc[NBI > 3 ? 3 : 0][0], cur[1] = acc[NBI > 3 ? 3 : 0][1];
                            break;
                        }
#pragma unroll
                        for (int bj = 0; bj < (int)NB; ++bj)
#pragma unroll
                            for (int hf = 0; hf < 2; ++hf) {
                                const Mask col = HALF_COL << (NB * 8 * hf + bj);
                                if (__builtin_amdgcn_ballot_w64((bmask & col) != 0)) {
#pragma unroll
                                    for (int j = 0; j < 8; ++j)
                                        gst[wv][j][lane] = cur[bj][8 * hf + j];
                                    band_run(bi, bj, hf, bmask, amask);
                                }
                            }
                    }
                    // certain accepts: both orders (the self pair once)
                    Mask am = amask;
                    while (true) {
                        const bool has = am != 0;
                        const uint64_t m1 = __builtin_amdgcn_ballot_w64(has);
                        if (!m1)
                            break;
                        uint32_t il = 0, jl = 0;
                        if (has) {
                            const uint32_t k = (uint32_t)__builtin_ctzll((uint64_t)am);
                            am &= am - 1;
                            const uint32_t e = k / NB, bj = k % NB;
                            il = WRS * wr + 32 * bi + 8 * (e / 4) + 4 * h + (e % 4);
                            jl = c0 + 32 * bj + (lane & 31);
                        }
                        const bool two = has && (qbeg + il) != (cb + jl);
                        const uint64_t m2 = __builtin_amdgcn_ballot_w64(two);
                        const uint32_t tot = (uint32_t)(__popcll(m1) + __popcll(m2));
                        if (fill + tot > MHITS)
                            flush();
                        if (has) {
                            const unsigned long long po = oA[il], co = oB[jl];
                            const uint32_t idx =
                                fill + (uint32_t)(__popcll(m1 & lt) + __popcll(m2 & lt));
                            hb[idx] = (po << qbits) | co;
                            if (two)
                                hb[idx + 1] = (co << qbits) | po;
                        }
                        fill += tot;
                    }
                }
            } else {
                // fast pass: per 32-row block the smallest reject score of its pairs
                // (in a diagonal tile of pairs p <= c); a block whose smallest score
                // exceeds hi holds certain rejects only.  Scores of padding rows and
                // of rows outside the item are +inf; NaN scores (NaN terms) are
                // caught by tile_nan.
                float mn[4];
                auto fast = [&](auto dg) {
                    constexpr bool DT = decltype(dg)::value;
#pragma unroll
                    for (int bi = 0; bi < (int)NBI; ++bi) {
                        float m = INF;
#pragma unroll
                        for (int eg = 0; eg < 4; ++eg) {
                            const uint32_t i0 = WRS * wr + 32 * bi + 8 * eg + 4 * h;
                            const float4 ay4 = *(const float4 *)&tAy[i0];
                            const float ay[4] = {ay4.x, ay4.y, ay4.z, ay4.w};
#pragma unroll
                            for (int ei = 0; ei < 4; ++ei)
#pragma unroll
                                for (int bj = 0; bj < 2; ++bj) {
                                    float sc =
                                        fmaf(-2.f, acc[bi][bj][4 * eg + ei], ay[ei] + tb[bj].y);
                                    if (DT && i0 + ei > 64 * wc + 32 * bj + (lane & 31) + doff)
                                        sc = INF;
                                    m = fminf(m, sc);
                                }
                        }
                        mn[bi] = m;
                    }
                };
                if (diag_tile)
                    fast(std::true_type());
                else
                    fast(std::false_type());

                // full pass, for the 32-row blocks (bi) that need it: the block's
                // 32 pairs per lane (slot 2 e + bj) classified into lane bit masks of
                // certain accepts and band entries (with DIAG also the verified
                // classes)
                bool cvalid[2];
#pragma unroll
                for (int bj = 0; bj < 2; ++bj)
                    cvalid[bj] = cb + 64 * wc + 32 * bj + (lane & 31) < w.cend;
                uint32_t amask[4] = {0u, 0u, 0u, 0u}, bmask[4] = {0u, 0u, 0u, 0u};
#pragma unroll
                for (int bi = 0; bi < (int)NBI; ++bi) {
                    const bool need = sampled || tile_nan || !(mn[bi] > hi);
                    if (!__builtin_amdgcn_ballot_w64(need))
                        continue;
#pragma unroll
                    for (int e = 0; e < 16; ++e) {
                        const uint32_t il = WRS * wr + 32 * bi + 8 * (e / 4) + 4 * h + (e % 4);
                        const float4 ta = tA[il];
                        const uint32_t p = qbeg + il;
                        const bool pv = p < w.qend;
#pragma unroll
                        for (int bj = 0; bj < 2; ++bj) {
                            const uint32_t c = cb + 64 * wc + 32 * bj + (lane & 31);
                            const bool valid = pv && cvalid[bj] && (!diag_tile || p <= c);
                            const float G = acc[bi][bj][e];
                            const bool a_ok = valid && fmaf(-2.f, G, ta.x + tb[bj].x) <= lo;
                            const bool r_ok = fmaf(-2.f, G, ta.y + tb[bj].y) > hi;
                            const bool inband = valid && !a_ok && !r_ok;
                            const bool push =
                                DIAG ? (inband || a_ok || (sampled && valid)) : inband;
                            amask[bi] |= (uint32_t)a_ok << (2 * e + bj);
                            bmask[bi] |= (uint32_t)push << (2 * e + bj);
                        }
                    }
                }
                // band entries (rare): staged slot by slot, with their scores; one
                // 32-row block at a time (not unrolled: a copy of its accumulators)
#pragma unroll 1
                for (int bi = 0; bi < (int)NBI; ++bi) {
                    // register arrays need constant indices
                    const uint32_t bm =
                        bi == 0 ? bmask[0] : bi == 1 ? bmask[1] : bi == 2 ? bmask[2] : bmask[3];
                    if (!__builtin_amdgcn_ballot_w64(bm != 0))
                        continue;
                    const uint32_t amk =
                        bi == 0 ? amask[0] : bi == 1 ? amask[1] : bi == 2 ? amask[2] : amask[3];
                    floatx16 cur[2];
                    switch (bi) {
                    case 0:
                        cur[0] = acc[0][0], cur[1] = acc[0][1];
                        break;
                    case 1:
                        cur[0] = acc[1][0], cur[1] = acc[1][1];
                        break;
                    case 2:
                        cur[0] = acc[NBI > 2 ? 2 : 0][0], cur[1] = acc[NBI > 2 ? 2 : 0][1];
                        break;
                    default:
                        cur[0] = acc[NBI > 3 ? 3 : 0][0], cur[1] = acc[NBI > 3 ? 3 : 0][1];
                        break;
                    }
#pragma unroll
                    for (int e = 0; e < 16; ++e) {
                        const uint32_t il = WRS * wr + 32 * bi + 8 * (e / 4) + 4 * h + (e % 4);
#pragma unroll
                        for (int bj = 0; bj < 2; ++bj) {
                            const bool push = (bm >> (2 * e + bj)) & 1u;
                            const uint64_t mb = __builtin_amdgcn_ballot_w64(push);
                            if (!mb)
                                continue;
                            const uint32_t jl = 64 * wc + 32 * bj + (lane & 31);
                            const uint32_t k = (uint32_t)__popcll(mb);
                            if (bfill + k > BHITS)
                                bflush();
                            const bool a_ok = (amk >> (2 * e + bj)) & 1u;
                            bool inband = push;
                            if (push) {
                                const float4 ta = tA[il];
                                const float G = cur[bj][e];
                                const uint32_t p = qbeg + il, c = cb + jl;
                                const bool valid =
                                    p < w.qend && c < w.cend && (!diag_tile || p <= c);
                                const bool r_ok = fmaf(-2.f, G, ta.y + tb[bj].y) > hi;
                                inband = valid && !a_ok && !r_ok;
                                const uint32_t cls = DIAG ? (a_ok ? 1u : (inband ? 0u : 2u)) : 0u;
                                const uint32_t idx = bfill + (uint32_t)__popcll(mb & lt);
                                bb[idx] = ((unsigned long long)jA[il] << 32) |
                                          ((unsigned long long)cls << 30) | jB[jl];
                                sb[idx] = fmaf(-2.f, G, ta.z + tb[bj].z);
                            }
                            bfill += k;
                            b0fill += DIAG ? (uint32_t)__popcll(__builtin_amdgcn_ballot_w64(inband))
                                           : k;
                        }
                    }
                }
                // certain accepts: both orders (the self pair once)
#pragma unroll
                for (int bi = 0; bi < (int)NBI; ++bi) {
                    uint32_t am = amask[bi];
                    while (true) {
                        const bool has = am != 0;
                        const uint64_t m1 = __builtin_amdgcn_ballot_w64(has);
                        if (!m1)
                            break;
                        uint32_t il = 0, jl = 0;
                        if (has) {
                            const uint32_t k = (uint32_t)__builtin_ctz(am);
                            am &= am - 1;
                            const uint32_t e = k >> 1, bj = k & 1u;
                            il = WRS * wr + 32 * bi + 8 * (e / 4) + 4 * h + (e % 4);
                            jl = 64 * wc + 32 * bj + (lane & 31);
                        }
                        const bool two = has && (qbeg + il) != (cb + jl);
                        const uint64_t m2 = __builtin_amdgcn_ballot_w64(two);
                        const uint32_t tot = (uint32_t)(__popcll(m1) + __popcll(m2));
                        if (fill + tot > MHITS)
                            flush();
                        if (has) {
                            const unsigned long long po = oA[il], co = oB[jl];
                            const uint32_t idx =
                                fill + (uint32_t)(__popcll(m1 & lt) + __popcll(m2 & lt));
                            hb[idx] = (po << qbits) | co;
                            if (two)
                                hb[idx + 1] = (co << qbits) | po;
                        }
                        fill += tot;
                    }
                }
            }
        }
        if constexpr (REF && !DIRECT) {
            // the next tile: find it (a skipped slot costs one more item
            // load), start its first slab (high plane) and its row terms
            find_next();
            if (nhave) {
                copy_slab(a0_of(ns, nw) / 32, b0_of(ns, nw) / 32, 0, buf, NP0);
                fetch_terms(ns, nw);
            }
        }
        v = vn;
        cs = ns;
        w = nw;
        have = nhave;
    }
    // no copy into this workgroup's LDS may still be in flight at its end
    asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
    if (fill)
        flush();
    if (bfill)
        bflush();
}

// FP64 decision of the band: the plain form's sequence over the whole padded
// row (same decision as with early exit; the full sum is the score that the
// filter's estimate is compared with).
template <bool DIAG>
__global__ __launch_bounds__(256) void k_band_check(
    const double *__restrict__ xr, uint32_t ds, double eps2, const uint32_t *__restrict__ jperm,
    const unsigned long long *__restrict__ band, const float *__restrict__ bscore,
    unsigned long long bcap, int qbits, unsigned long long *__restrict__ out,
    unsigned long long cap, unsigned long long *__restrict__ ctr)
{
    const unsigned long long nb = min(ctr[1], bcap);
    const uint32_t lane = threadIdx.x & (WAVE - 1);
    const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (WAVE - lane));
    double emax = 0.0, esum = 0.0;
    unsigned long long checked = 0, viol = 0, decided = 0, checked3 = 0;
    const size_t stride = (size_t)gridDim.x * blockDim.x;
    // whole waves iterate together so the ballots below see every lane
    for (size_t base = (size_t)blockIdx.x * blockDim.x + (threadIdx.x & ~(WAVE - 1)); base < nb;
         base += stride) {
        const size_t i = base + lane;
        const bool have = i < nb;
        bool emit = false, two = false;
        unsigned long long ko = 0, kr = 0;
        if (have) {
            const unsigned long long e = band[i];
            const uint32_t ja = (uint32_t)(e >> 32), jb = (uint32_t)e & 0x3FFFFFFFu;
            const uint32_t cls = ((uint32_t)e >> 30) & 3u;
            const double *qa = xr + (size_t)ja * ds;
            const double *qb = xr + (size_t)jb * ds;
            double s = 0.0;
            for (uint32_t l = 0; l < ds; l += 4) {
                double t0 = qa[l] - qb[l];
                s = fma(t0, t0, s);
                double t1 = qa[l + 1] - qb[l + 1];
                double q1 = t1 * t1;
                s = s + q1;
                double t2 = qa[l + 2] - qb[l + 2];
                double q2 = t2 * t2;
                s = s + q2;
                double t3 = qa[l + 3] - qb[l + 3];
                double q3 = t3 * t3;
                s = s + q3;
            }
            const bool ok = !(s > eps2);
            const double err = fabs((double)bscore[i] - s);
            if (err == err && cls != 3) {  // class 3 holds BF16 scores
                emax = err > emax ? err : emax;
                esum += err;
            }
            if (cls == 0) {
                ++decided;
                emit = ok;
                two = ok && ja != jb;
                ko = (unsigned long long)jperm[ja];
                kr = (unsigned long long)jperm[jb];
            }
            if (DIAG && cls != 0) {
                ++checked;
                checked3 += cls == 3;
                if ((cls == 1 && !ok) || ((cls == 2 || cls == 3) && ok))
                    ++viol;
            }
        }
        const uint64_t m1 = __builtin_amdgcn_ballot_w64(emit);
        if (m1) {
            const uint64_t m2 = __builtin_amdgcn_ballot_w64(two);
            unsigned long long b0 = 0;
            if (lane == 0)
                b0 = atomicAdd(&ctr[0], (unsigned long long)(__popcll(m1) + __popcll(m2)));
            b0 = __shfl(b0, 0, WAVE);
            if (emit) {
                const unsigned long long o = b0 + __popcll(m1 & lt) + __popcll(m2 & lt);
                if (o < cap)
                    out[o] = (ko << qbits) | kr;
                if (two && o + 1 < cap)
                    out[o + 1] = (kr << qbits) | ko;
            }
        }
    }
    // wave reductions, one atomic each
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1) {
        double t = __shfl_xor(emax, o, WAVE);
        emax = t > emax ? t : emax;
        esum += __shfl_xor(esum, o, WAVE);
        checked += __shfl_xor(checked, o, WAVE);
        viol += __shfl_xor(viol, o, WAVE);
        decided += __shfl_xor(decided, o, WAVE);
        checked3 += __shfl_xor(checked3, o, WAVE);
    }
    if (lane == 0) {
        if (emax > 0.0)
            atomicMax(&ctr[2], (unsigned long long)__double_as_longlong(emax));
        if (esum > 0.0)
            atomicAdd((double *)&ctr[5], esum);
        if (checked)
            atomicAdd(&ctr[3], checked);
        if (viol)
            atomicAdd(&ctr[4], viol);
        if (decided)
            atomicAdd(&ctr[6], decided);
        if (checked3)
            atomicAdd(&ctr[9], checked3);
    }
}

__global__ void k_zero_ctr(unsigned long long *c, uint32_t n)
{
    if (threadIdx.x < n)
        c[threadIdx.x] = 0ull;
}

#define LAUNCH(kern, grid, block, ...)                                                   \
    do {                                                                                 \
        ::msj::check_grid((grid), (block));                                              \
        hipLaunchKernelGGL(kern, dim3(grid), dim3(block), 0, s, __VA_ARGS__);            \
        MSJ_HIP(hipGetLastError());                                                      \
    } while (0)

// Bound constant C of the threshold terms (see the file comment).  rnd:
// the rounding of the sum before the final one, in units of 2^-24
// (|a'|^2 + |b'|^2): 1 for RN(alpha_p + alpha_c); the BF16 pass of the
// split form rounds alpha'_b,p - 2 G instead, |alpha'_b,p - 2 G| <= 2.01
// (|a'|^2 + |b'|^2), and takes 3.
double bound_c(uint32_t d, bool split, double rnd = 1.0)
{
    const double u = 0x1p-8 * (1.0 + 0x1p-15);
    const double ep = split ? (3 * u * u + 2 * u * u * u + 2 * u * u * u * u) : (2 * u + u * u);
    const double m = split ? 3.0 * d : (double)d;
    const double t = 2.0 * m * 0x1p-24;
    const double gam = t / (1.0 - t);
    const double eg = ep + gam * (1 + u) * (1 + u) * (1 + 2 * u);
    return 1.01 * (eg + rnd * 0x1p-24) + 0x1p-40;
}

float round_down_f(double v)
{
    float f = (float)v;
    if ((double)f > v)
        f = std::nextafter(f, -INFINITY);
    return f;
}
float round_up_f(double v)
{
    float f = (float)v;
    if ((double)f < v)
        f = std::nextafter(f, INFINITY);
    return f;
}

}  // namespace

bool mfma_filter_applies(double eps, uint32_t n, uint32_t d)
{
    const double e2 = eps * eps;
    return e2 >= 0x1p-50 && e2 <= 0x1p60 && n < (1u << 30) && d <= 4096;
}

void MfmaRunner::prepare(const Index &ix, const JoinRunner &J, const double *xr, uint32_t n,
                         uint32_t d, double eps, bool split_bf16, hipStream_t s)
{
    (void)eps;
    split = split_bf16;
    const uint32_t M = ix.M;
    const uint32_t ds = padded_dim(d);
    const uint32_t Kp = (d + KALIGN - 1) / KALIGN * KALIGN;
    KS = Kp / 16;

    // padded row space
    prow.ensure(M);
    p64.ensure((size_t)M + 1);
    LAUNCH(k_prow, grid_for(M, 256), 256, ix.nstart.p, M, prow.p);
    scan.excl_u32_to_u64(prow.p, p64.p, M, s);
    uint64_t Pt = 0;
    MSJ_HIP(hipMemcpyAsync(&Pt, p64.p + M, sizeof(Pt), hipMemcpyDeviceToHost, s));
    MSJ_HIP(hipStreamSynchronize(s));
    if (Pt + TM >= (1ull << 31))
        throw Error(MSJ_E_OVERFLOW, "padded rows exceed 2^31");
    P = (uint32_t)Pt;
    const uint32_t Ptot = P + TM;
    pstart.ensure((size_t)M + 1);
    LAUNCH(k_u64_to_u32, grid_for((size_t)M + 1, 256), 256, p64.p, (size_t)M + 1, pstart.p);
    pjoin.ensure(Ptot);
    porig.ensure(Ptot);
    LAUNCH(k_fill_u32, grid_for(Ptot, 256), 256, pjoin.p, (size_t)Ptot, ~0u);
    LAUNCH(k_fill_u32, grid_for(Ptot, 256), 256, porig.p, (size_t)Ptot, 0u);
    LAUNCH(k_pad_rows, grid_for(n, 256), 256, ix.nstart.p, M, pstart.p, J.jperm.p, n, pjoin.p,
           porig.p);

    // centre
    const uint32_t nb = std::min<uint32_t>(1024, std::max<uint32_t>(1, (n + 255) / 256));
    const uint32_t rows_per = (n + nb - 1) / nb;
    colpart.ensure((size_t)nb * ds);
    mu.ensure(ds);
    LAUNCH(k_col_partial, nb, 256, xr, n, ds, rows_per, colpart.p);
    LAUNCH(k_col_final, ds, 256, colpart.p, nb, ds, n, mu.p);

    // fragments and threshold terms
    const double C = bound_c(d, split);
    const size_t nfrag = (size_t)(Ptot / 32) * KS * WAVE;
    fh.ensure(nfrag);
    if (split)
        fl.ensure(nfrag);
    thr.ensure(Ptot);
    // split-BF16 scores every tile in BF16 first: it also needs the rows'
    // terms under the BF16 bound
    const bool refine = split;
    if (refine)
        thrb.ensure(Ptot);
    const double Cb = bound_c(d, false, 3.0);
    // rows per wave: as many as leave a lane per 8-dimension group
    const uint32_t rpw = 2 * KS <= 16 ? 4u : (2 * KS <= 32 ? 2u : 1u);
    const unsigned fg = grid_for(Ptot, 4 * rpw);
#define MSJ_FRAG_PREP(SPL, R)                                                                    \
    LAUNCH((k_frag_prep<SPL, R>), fg, 256, xr, ds, KS, mu.p, pjoin.p, Ptot, C, fh.p,           \
           SPL ? fl.p : fh.p, thr.p, Cb, (SPL && refine) ? thrb.p : nullptr)
    if (split) {
        if (rpw == 4)
            MSJ_FRAG_PREP(true, 4);
        else if (rpw == 2)
            MSJ_FRAG_PREP(true, 2);
        else
            MSJ_FRAG_PREP(true, 1);
    } else {
        if (rpw == 4)
            MSJ_FRAG_PREP(false, 4);
        else if (rpw == 2)
            MSJ_FRAG_PREP(false, 2);
        else
            MSJ_FRAG_PREP(false, 1);
    }
#undef MSJ_FRAG_PREP

    // work items: super-blocks over the upper neighbour lists
    uint32_t E = 0;
    MSJ_HIP(hipMemcpyAsync(&E, J.uoff.p + M, sizeof(E), hipMemcpyDeviceToHost, s));
    MSJ_HIP(hipStreamSynchronize(s));
    nitems = 0;
    tiles = 0;
    smap = SlotMap{};
    if (E == 0)
        return;
    const size_t E8 = (size_t)SBS * E;
    esrc.ensure(E);
    cnt8.ensure(E8);
    etiles.ensure((size_t)E + 1);
    off8.ensure(E8 + 1);
    LAUNCH(k_edge_src, grid_for(M, 256), 256, J.uoff.p, M, esrc.p);
    LAUNCH(k_sblk_count, grid_for(E, 256), 256, ix.nstart.p, esrc.p, J.uadj.p, E, cnt8.p,
           etiles.p);
    scan.excl_u32_to_u64(cnt8.p, off8.p, E8, s);
    p64.ensure((size_t)E + 1);  // pstart is converted already; contents not needed
    scan.excl_u64_to_u64(etiles.p, p64.p, E, s);
    std::vector<uint64_t> cb(SBS + 1);
    for (uint32_t c = 0; c <= SBS; ++c)
        MSJ_HIP(hipMemcpyAsync(&cb[c], off8.p + (size_t)c * E, sizeof(uint64_t),
                               hipMemcpyDeviceToHost, s));
    MSJ_HIP(hipMemcpyAsync(&tiles, p64.p + E, sizeof(tiles), hipMemcpyDeviceToHost, s));
    MSJ_HIP(hipStreamSynchronize(s));
    nitems = cb[SBS];
    uint64_t slot = 0;
    for (uint32_t c = 0; c < SBS; ++c) {
        const uint64_t sz = SBS - c;
        smap.item_base[c] = cb[c];
        smap.slot_base[c] = slot;
        slot += (cb[c + 1] - cb[c]) * sz * sz;
    }
    smap.slot_base[SBS] = slot;
    if (slot >= (1ull << 40))
        throw Error(MSJ_E_OVERFLOW, "filter tile slots exceed 2^40");
    smapd.ensure(1);
    MSJ_HIP(hipMemcpyAsync(smapd.p, &smap, sizeof(SlotMap), hipMemcpyHostToDevice, s));
    items.ensure(std::max<uint64_t>(nitems, 1));
    if (nitems)
        LAUNCH(k_sblk_fill, grid_for(nitems, 256), 256, ix.nstart.p, pstart.p, esrc.p, J.uadj.p,
               off8.p, E, nitems, items.p);
}

uint64_t MfmaRunner::run(const Index &ix, JoinRunner &J, const double *xr, uint32_t n, uint32_t d,
                         double eps, uint64_t cap_hint, bool diag, hipStream_t s, hipEvent_t k0,
                         hipEvent_t kmid, hipEvent_t k1, uint64_t *launches, float *kernel_ms,
                         float *main_ms, msj_stats *st)
{
    const double eps2 = eps * eps;  // launcher.cu:3200
    const uint32_t ds = padded_dim(d);
    const double g = (ds + 4.0) * 0x1p-52;
    const float lo = round_down_f(eps2 / ((1.0 + 0x1p-23) * (1.0 + g)) * (1.0 - 0x1p-30) - 0x1p-58);
    const float hi = round_up_f(eps2 / ((1.0 - 0x1p-24) * (1.0 - g)) * (1.0 + 0x1p-30) + 0x1p-58);
    const int qbits = std::max(1, bits_for(n ? n - 1 : 0));
    // BF16 band budget: beyond it the BF16 bound is too loose for this data
    // and the join is rerun with split-BF16 (the kernel stops early)
    const uint64_t bf16_budget = std::max<uint64_t>(16ull * n, 1ull << 22);
    ctr.ensure(10);
    if (J.pairs.cap < cap_hint)
        J.pairs.ensure(cap_hint);
    *kernel_ms = 0.f;
    *main_ms = 0.f;
    for (int attempt = 0; attempt < 4; ++attempt) {
        // count_work: verify every pair of one item in sample_every (at least
        // 61, about 5e7 sampled pairs at most)
        const uint64_t pairs_bound = tiles * (uint64_t)TM * TM;
        const uint32_t sample_every = (uint32_t)std::min<uint64_t>(
            std::max<uint64_t>(61, pairs_bound / 50000000ull + 1), 0x7FFFFFFFull);
        const uint64_t budget = split ? 0 : bf16_budget;
        uint64_t bwant = std::max<uint64_t>(
            std::max<uint64_t>(last_band + last_band / 4, 16ull * n), 1ull << 20);
        if (!split)
            bwant = std::max<uint64_t>(bwant, bf16_budget + bf16_budget / 8);
        if (band.cap < bwant) {
            band.ensure(bwant);
            bscore.ensure(bwant);
        }
        // d <= 16 KS_HALF: half tiles (256 x 128, two workgroups per CU, the
        // epilogue of one overlaps the other's MFMAs); above: 256 x 256 tiles
        // (twice the operand reuse per slab copy)
        const bool half = KS <= KS_HALF;
        // slots: half tiles (split-BF16: quarter tiles, 128 x 128) or tiles
        const uint64_t slots = (half ? (split ? 4 : 2) : 1) * smap.slot_base[SBS];
        const uint64_t per_xcd = (slots + 7) / 8;
        const uint32_t grid =
            (uint32_t)std::max<uint64_t>(8, std::min<uint64_t>(per_xcd * 8, MAX_GRID));
        LAUNCH(k_zero_ctr, 1, 64, ctr.p, 10u);
        MSJ_HIP(hipEventRecord(k0, s));
        if (nitems) {
#define MSJ_MJ(SP, DG)                                                                          \
    do {                                                                                        \
        if (half)                                                                               \
            MSJ_MJH(SP, DG, true, SP, 256);                                                  \
        else                                                                                    \
            MSJ_MJH(SP, DG, false, SP, 512);                                                    \
    } while (0)
#define MSJ_MJH(SP, DG, HF, RF, NT)                                                             \
    LAUNCH((k_mfma_join<SP, DG, HF, RF>), grid, NT, fh.p, SP ? fl.p : fh.p, KS, thr.p,          \
           RF ? thrb.p : nullptr, porig.p, pjoin.p,                                             \
           items.p, smapd.p, per_xcd, lo, hi, qbits, (unsigned long long *)J.pairs.p,           \
           (unsigned long long)J.pairs.cap, ctr.p, (unsigned long long *)band.p, bscore.p,      \
           (unsigned long long)band.cap, sample_every, (unsigned long long)budget)
            if (split) {
                if (diag)
                    MSJ_MJ(true, true);
                else
                    MSJ_MJ(true, false);
            } else {
                if (diag)
                    MSJ_MJ(false, true);
                else
                    MSJ_MJ(false, false);
            }
#undef MSJ_MJ
#undef MSJ_MJH
            MSJ_HIP(hipEventRecord(kmid, s));
            if (diag)
                LAUNCH(k_band_check<true>, 2048, 256, xr, ds, eps2, J.jperm.p,
                       (const unsigned long long *)band.p, bscore.p,
                       (unsigned long long)band.cap, qbits, (unsigned long long *)J.pairs.p,
                       (unsigned long long)J.pairs.cap, ctr.p);
            else
                LAUNCH(k_band_check<false>, 2048, 256, xr, ds, eps2, J.jperm.p,
                       (const unsigned long long *)band.p, bscore.p,
                       (unsigned long long)band.cap, qbits, (unsigned long long *)J.pairs.p,
                       (unsigned long long)J.pairs.cap, ctr.p);
            *launches += 2;
        } else {
            MSJ_HIP(hipEventRecord(kmid, s));
        }
        MSJ_HIP(hipEventRecord(k1, s));
        unsigned long long h[10];
        MSJ_HIP(hipMemcpyAsync(h, ctr.p, sizeof(h), hipMemcpyDeviceToHost, s));
        MSJ_HIP(hipStreamSynchronize(s));
        float ms = 0.f, mms = 0.f;
        MSJ_HIP(hipEventElapsedTime(&ms, k0, k1));
        MSJ_HIP(hipEventElapsedTime(&mms, k0, kmid));
        *kernel_ms += ms;
        *main_ms += mms;
        const uint64_t R = h[0], nband = h[1];
        if (!split && h[7] > bf16_budget) {
            // escalate: the same items with split-BF16 operands
            prepare(ix, J, xr, n, d, eps, true, s);
            last_band = 0;
            continue;
        }
        if (R <= J.pairs.cap && nband <= band.cap) {
            last_band = nband;
            double emax, esum;
            std::memcpy(&emax, &h[2], sizeof(double));
            std::memcpy(&esum, &h[5], sizeof(double));
            st->band_pairs = h[6];
            st->max_score_err = emax;
            st->mean_score_err = nband ? esum / (double)nband : 0.0;
            st->filter_checked = h[3];
            st->filter_violations = h[4];
            st->filter_checked_bf16 = h[9];
            st->form_used = split ? MSJ_FORM_MFMA_SPLIT : MSJ_FORM_MFMA_BF16;
            // split: BF16 products on every tile, the two cross products on
            // the 32 x 32 blocks (64 per tile) that needed them (h[8])
            const double tile_flops = (double)TM * TM * 2.0 * (16.0 * KS);
            const bool refine = split;
            st->mfma_flops = !split ? tiles * tile_flops
                             : refine ? tiles * tile_flops + (double)h[8] * tile_flops / 32.0
                                      : tiles * tile_flops * 3.0;
            st->split_share_ppm =
                !split ? 0u
                : refine ? (uint32_t)std::min<uint64_t>(
                               1000000ull, tiles ? h[8] * 1000000ull / (64ull * tiles) : 0)
                         : 1000000u;
            return R;
        }
        if (nband > band.cap) {
            band.ensure(nband);
            bscore.ensure(nband);
        }
        if (R > J.pairs.cap)
            J.pairs.ensure(R);
    }
    throw Error(MSJ_E_OVERFLOW, "result count changed between identical join passes");
}

}  // namespace msj
