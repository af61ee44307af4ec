// msj_prims.hip -- device primitives: exclusive scan, stable LSD radix sort,
// reductions.  Integer work only; HBM-bound by design.
//
// Radix sort: 8-bit digits, 4096-key tiles.  Per pass: a tile histogram
// (LDS atomics), one exclusive scan over the digit-major histogram, and a
// stable scatter whose in-tile ranks come from wave ballots (8 ballots give
// each lane the mask of lanes holding the same digit), so equal keys keep
// their input order without any global atomics.
#include "msj_internal.h"

namespace msj {

namespace {

constexpr int SCAN_BLOCK = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_BLOCK * SCAN_ITEMS;

constexpr int RS_BLOCK = 256;
constexpr int RS_ROUNDS = 16;
constexpr int RS_TILE = RS_BLOCK * RS_ROUNDS;

template <class T>
__device__ inline T wave_incl_scan(T v, int lane)
{
#pragma unroll
    for (int o = 1; o < WAVE; o <<= 1) {
        T t = __shfl_up(v, o, WAVE);
        if (lane >= o)
            v += t;
    }
    return v;
}

template <class T>
__device__ inline T wave_sum(T v)
{
#pragma unroll
    for (int o = WAVE / 2; o > 0; o >>= 1)
        v += __shfl_down(v, o, WAVE);
    return v;
}

// Exclusive scan of one value per thread across the block; *total gets the
// block sum.  `lds` holds one T per wave.
template <class T>
__device__ inline T block_excl_scan(T v, T *lds, T *total)
{
    const int lane = threadIdx.x & (WAVE - 1);
    const int wave = threadIdx.x / WAVE;
    const int nw = blockDim.x / WAVE;
    T inc = wave_incl_scan(v, lane);
    if (lane == WAVE - 1)
        lds[wave] = inc;
    __syncthreads();
    T off = 0, tot = 0;
    for (int w = 0; w < nw; ++w) {
        T x = lds[w];
        if (w < wave)
            off += x;
        tot += x;
    }
    __syncthreads();
    *total = tot;
    return off + inc - v;
}

template <class Tin>
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_reduce(const Tin *__restrict__ in, size_t n,
                                                            uint64_t *__restrict__ part)
{
    __shared__ uint64_t lds[SCAN_BLOCK / WAVE];
    size_t base = (size_t)blockIdx.x * SCAN_TILE;
    uint64_t s = 0;
#pragma unroll
    for (int r = 0; r < SCAN_ITEMS; ++r) {
        size_t i = base + (size_t)r * SCAN_BLOCK + threadIdx.x;
        if (i < n)
            s += (uint64_t)in[i];
    }
    s = wave_sum(s);
    if ((threadIdx.x & (WAVE - 1)) == 0)
        lds[threadIdx.x / WAVE] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint64_t t = 0;
        for (int w = 0; w < SCAN_BLOCK / WAVE; ++w)
            t += lds[w];
        part[blockIdx.x] = t;
    }
}

// Block-local scan of a 2048-element tile plus the tile's offset.  Thread t
// owns elements t*8 .. t*8+7 of the tile.
template <class Tin, class Tout>
__global__ __launch_bounds__(SCAN_BLOCK) void k_scan_apply(const Tin *__restrict__ in,
                                                           Tout *__restrict__ out, size_t n,
                                                           const uint64_t *__restrict__ offs)
{
    __shared__ uint64_t lds[SCAN_BLOCK / WAVE];
    size_t base = (size_t)blockIdx.x * SCAN_TILE + (size_t)threadIdx.x * SCAN_ITEMS;
    uint64_t v[SCAN_ITEMS];
    uint64_t s = 0;
#pragma unroll
    for (int r = 0; r < SCAN_ITEMS; ++r) {
        size_t i = base + r;
        v[r] = (i < n) ? (uint64_t)in[i] : 0;
        s += v[r];
    }
    uint64_t tot;
    uint64_t off = block_excl_scan<uint64_t>(s, lds, &tot);
    off += offs ? offs[blockIdx.x] : 0;
#pragma unroll
    for (int r = 0; r < SCAN_ITEMS; ++r) {
        size_t i = base + r;
        if (i < n)
            out[i] = (Tout)off;
        off += v[r];
    }
    // the block holding element n-1 (block 0 when n == 0) writes the total
    size_t last_block = n ? (n - 1) / SCAN_TILE : 0;
    if ((size_t)blockIdx.x == last_block && threadIdx.x == SCAN_BLOCK - 1)
        out[n] = (Tout)off;
}

template <class Tin, class Tout>
void scan_level(Scanner &sc, int level, const Tin *in, Tout *out, size_t n, hipStream_t s)
{
    unsigned nb = grid_for(n, SCAN_TILE);
    if (nb == 1) {
        hipLaunchKernelGGL((k_scan_apply<Tin, Tout>), dim3(1), dim3(SCAN_BLOCK), 0, s, in, out, n,
                           (const uint64_t *)nullptr);
        MSJ_HIP(hipGetLastError());
        return;
    }
    if (level >= 4)
        throw Error(MSJ_E_OVERFLOW, "scan too deep");
    sc.part[level].ensure((size_t)nb + 1);
    uint64_t *part = sc.part[level].p;
    hipLaunchKernelGGL((k_scan_reduce<Tin>), dim3(nb), dim3(SCAN_BLOCK), 0, s, in, n, part);
    MSJ_HIP(hipGetLastError());
    // In place: every block of k_scan_apply reads its own tile into registers
    // before the barrier inside block_excl_scan and writes only that tile.
    scan_level<uint64_t, uint64_t>(sc, level + 1, part, part, nb, s);
    hipLaunchKernelGGL((k_scan_apply<Tin, Tout>), dim3(nb), dim3(SCAN_BLOCK), 0, s, in, out, n,
                       (const uint64_t *)part);
    MSJ_HIP(hipGetLastError());
}

// ------------------------------------------------------------- radix sort
// Lanes of a wave holding the same digit: 8 ballots, one per digit bit.
__device__ inline uint64_t digit_match(bool valid, uint32_t dg)
{
    uint64_t m = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
        bool bit = (dg >> b) & 1u;
        uint64_t bb = __ballot(bit);
        m &= bit ? bb : ~bb;
    }
    return m;
}

__global__ __launch_bounds__(RS_BLOCK) void k_rs_hist(const uint64_t *__restrict__ keys, size_t n,
                                                      int shift, uint32_t *__restrict__ hist,
                                                      uint32_t ntiles)
{
    __shared__ uint32_t cnt[256];
    const int lane = threadIdx.x & (WAVE - 1);
    const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (WAVE - lane));
    cnt[threadIdx.x] = 0;
    __syncthreads();
    const size_t base = (size_t)blockIdx.x * RS_TILE + threadIdx.x;
    // the tile's keys are loaded up front: one memory round trip per tile
    uint64_t kr[RS_ROUNDS];
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; ++r) {
        const size_t i = base + (size_t)r * RS_BLOCK;
        kr[r] = i < n ? keys[i] : 0ull;
    }
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; ++r) {
        bool valid = base + (size_t)r * RS_BLOCK < n;
        uint32_t dg = (uint32_t)(kr[r] >> shift) & 0xFFu;
        uint64_t m = digit_match(valid, dg);
        if (valid && (m & lt) == 0)  // one LDS atomic per digit per wave
            atomicAdd(&cnt[dg], (uint32_t)__popcll(m));
    }
    __syncthreads();
    hist[(size_t)threadIdx.x * ntiles + blockIdx.x] = cnt[threadIdx.x];
}

__global__ __launch_bounds__(RS_BLOCK) void k_rs_scatter(const uint64_t *__restrict__ kin,
                                                         const uint32_t *__restrict__ vin,
                                                         uint64_t *__restrict__ kout,
                                                         uint32_t *__restrict__ vout, size_t n,
                                                         int shift,
                                                         const uint64_t *__restrict__ hscan,
                                                         uint32_t ntiles)
{
    __shared__ uint64_t base[256];
    __shared__ uint32_t run[256];
    __shared__ uint32_t wcnt[RS_BLOCK / WAVE][256];
    const int t = threadIdx.x;
    const int lane = t & (WAVE - 1);
    const int wave = t / WAVE;
    base[t] = hscan[(size_t)t * ntiles + blockIdx.x];
    run[t] = 0;
#pragma unroll
    for (int w = 0; w < RS_BLOCK / WAVE; ++w)
        wcnt[w][t] = 0;
    __syncthreads();
    const uint64_t lt = (lane == 0) ? 0ull : ((~0ull) >> (WAVE - lane));
    const size_t tile0 = (size_t)blockIdx.x * RS_TILE + t;
    // the tile's keys and values are loaded up front (one memory round trip
    // per tile); the rounds below then rank and store them in input order
    uint64_t kr[RS_ROUNDS];
    uint32_t vr[RS_ROUNDS];
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; ++r) {
        const size_t i = tile0 + (size_t)r * RS_BLOCK;
        kr[r] = i < n ? kin[i] : 0ull;
        vr[r] = (i < n && vin) ? vin[i] : 0u;
    }
#pragma unroll
    for (int r = 0; r < RS_ROUNDS; ++r) {
        const size_t i = tile0 + (size_t)r * RS_BLOCK;
        bool valid = i < n;
        uint64_t key = kr[r];
        uint32_t val = vr[r];
        uint32_t dg = (uint32_t)(key >> shift) & 0xFFu;
        uint64_t m = digit_match(valid, dg);
        uint32_t rank = (uint32_t)__popcll(m & lt);
        if (valid && rank == 0)
            wcnt[wave][dg] = (uint32_t)__popcll(m);
        __syncthreads();
        if (valid) {
            uint64_t pos = base[dg] + run[dg] + rank;
            for (int w = 0; w < wave; ++w)
                pos += wcnt[w][dg];
            kout[pos] = key;
            if (vin)
                vout[pos] = val;
        }
        __syncthreads();
        uint32_t add = 0;
#pragma unroll
        for (int w = 0; w < RS_BLOCK / WAVE; ++w) {
            add += wcnt[w][t];
            wcnt[w][t] = 0;
        }
        run[t] += add;
        __syncthreads();
    }
}

template <class T>
__global__ void k_fill(T *p, size_t n, T v)
{
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n)
        p[i] = v;
}

__global__ __launch_bounds__(256) void k_reduce_sum_u64(const uint64_t *__restrict__ in, size_t n,
                                                        unsigned long long *dst)
{
    uint64_t s = 0;
    for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
         i += (size_t)gridDim.x * blockDim.x)
        s += in[i];
    s = wave_sum(s);
    if ((threadIdx.x & (WAVE - 1)) == 0 && s)
        atomicAdd(dst, (unsigned long long)s);
}

}  // namespace

void Scanner::excl_u32_to_u32(const uint32_t *in, uint32_t *out, size_t n, hipStream_t s)
{
    scan_level<uint32_t, uint32_t>(*this, 0, in, out, n, s);
}
void Scanner::excl_u32_to_u64(const uint32_t *in, uint64_t *out, size_t n, hipStream_t s)
{
    scan_level<uint32_t, uint64_t>(*this, 0, in, out, n, s);
}
void Scanner::excl_u64_to_u64(const uint64_t *in, uint64_t *out, size_t n, hipStream_t s)
{
    scan_level<uint64_t, uint64_t>(*this, 0, in, out, n, s);
}

void RadixSorter::sort(uint64_t *keys, uint32_t *vals, size_t n, int bits, hipStream_t s)
{
    if (n < 2 || bits <= 0)
        return;
    if (n >= (1ull << 32))
        throw Error(MSJ_E_OVERFLOW, "radix sort of 2^32 or more keys");
    uint32_t ntiles = (uint32_t)((n + RS_TILE - 1) / RS_TILE);
    k_alt.ensure(n);
    if (vals)
        v_alt.ensure(n);
    hist.ensure((size_t)256 * ntiles);
    hist_scan.ensure((size_t)256 * ntiles + 1);
    uint64_t *ka = keys, *kb = k_alt.p;
    uint32_t *va = vals, *vb = vals ? v_alt.p : nullptr;
    int passes = (bits + 7) / 8;
    for (int ps = 0; ps < passes; ++ps) {
        int shift = ps * 8;
        hipLaunchKernelGGL(k_rs_hist, dim3(ntiles), dim3(RS_BLOCK), 0, s, ka, n, shift, hist.p,
                           ntiles);
        MSJ_HIP(hipGetLastError());
        scan.excl_u32_to_u64(hist.p, hist_scan.p, (size_t)256 * ntiles, s);
        hipLaunchKernelGGL(k_rs_scatter, dim3(ntiles), dim3(RS_BLOCK), 0, s, ka, va, kb, vb, n,
                           shift, hist_scan.p, ntiles);
        MSJ_HIP(hipGetLastError());
        std::swap(ka, kb);
        std::swap(va, vb);
    }
    if (ka != keys) {
        MSJ_HIP(hipMemcpyAsync(keys, ka, n * sizeof(uint64_t), hipMemcpyDeviceToDevice, s));
        if (vals)
            MSJ_HIP(hipMemcpyAsync(vals, va, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    }
}

void reduce_sum_u64(const uint64_t *in, size_t n, unsigned long long *dst, hipStream_t s)
{
    hipLaunchKernelGGL(k_fill<unsigned long long>, dim3(1), dim3(64), 0, s, dst, (size_t)1, 0ull);
    unsigned g = grid_for(n, 256);
    if (g > 2048)
        g = 2048;
    hipLaunchKernelGGL(k_reduce_sum_u64, dim3(g), dim3(256), 0, s, in, n, dst);
    MSJ_HIP(hipGetLastError());
}

}  // namespace msj
