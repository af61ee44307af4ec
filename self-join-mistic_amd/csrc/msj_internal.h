// msj_internal.h -- shared declarations of the HIP implementation.
//
// Everything here is internal to libmsj.so; the public boundary is
// include/msj.h.  Device code in this library is compiled with
// -ffp-contract=off: every fused multiply-add in a parity-relevant
// expression is written as fma().
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "msj.h"

namespace msj {

// Internal failure carrying an msj_status; converted to a return code at the
// C ABI.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char *what, const char *file, int line);

// Runs fn(arg) and maps an Error / std::bad_alloc / std::exception to an
// msj_status, recording its text for msj_last_error() (the C ABI's rule:
// status codes, never an abort or an exception across the boundary).
int guarded_call(void (*fn)(void *), void *arg);
#define MSJ_HIP(x) ::msj::hip_check((x), #x, __FILE__, __LINE__)

// A fixed device region handed out front to back (256-byte aligned).  While
// one is installed on the calling thread (see ArenaScope), DevBuf takes its
// memory from it instead of hipMalloc; nothing is ever returned to it.
struct Arena {
    char *base = nullptr;
    size_t cap = 0, used = 0;
    void *take(size_t bytes)
    {
        const size_t at = (used + 255) & ~(size_t)255;
        if (at > cap || bytes > cap - at)
            throw Error(MSJ_E_NOMEM, "allocation-free scratch exhausted: " +
                                         std::to_string(at + bytes) + " bytes needed, " +
                                         std::to_string(cap) + " available");
        used = at + bytes;
        return base + at;
    }
};
extern thread_local Arena *g_arena;

// Installs an arena on this thread for the lifetime of the scope.
struct ArenaScope {
    Arena *prev;
    explicit ArenaScope(Arena *a) : prev(g_arena) { g_arena = a; }
    ~ArenaScope() { g_arena = prev; }
};

// Device allocation that only grows; released on destruction.  Under an
// installed Arena the memory is borrowed from it (and never freed here).
template <class T>
struct DevBuf {
    T *p = nullptr;
    size_t cap = 0;
    bool borrowed = false;
    DevBuf() = default;
    DevBuf(const DevBuf &) = delete;
    DevBuf &operator=(const DevBuf &) = delete;
    ~DevBuf() { release(); }
    void ensure(size_t n)
    {
        if (n <= cap)
            return;
        release();
        if (g_arena) {
            p = (T *)g_arena->take(n * sizeof(T));
            borrowed = true;
            cap = n;
            return;
        }
        hipError_t e = hipMalloc((void **)&p, n * sizeof(T));
        if (e != hipSuccess) {
            p = nullptr;
            (void)hipGetLastError();
            throw Error(MSJ_E_NOMEM, std::string("hipMalloc of ") + std::to_string(n * sizeof(T)) +
                                         " bytes failed: " + hipGetErrorString(e));
        }
        cap = n;
    }
    void release()
    {
        if (p && !borrowed)
            (void)hipFree(p);
        p = nullptr;
        cap = 0;
        borrowed = false;
    }
};

constexpr int WAVE = 64;

// ---------------------------------------------------------------- primitives
// Exclusive prefix sum: out[i] = sum(in[0..i)), out[n] = total.  out must hold
// n+1 elements.  Scratch is grown on demand.
struct Scanner {
    DevBuf<uint64_t> part[4];
    void excl_u32_to_u32(const uint32_t *in, uint32_t *out, size_t n, hipStream_t s);
    void excl_u32_to_u64(const uint32_t *in, uint64_t *out, size_t n, hipStream_t s);
    void excl_u64_to_u64(const uint64_t *in, uint64_t *out, size_t n, hipStream_t s);
};

// Stable LSD radix sort of 64-bit keys (optionally carrying 32-bit values)
// on the low `bits` bits, 8 bits per pass.
struct RadixSorter {
    DevBuf<uint64_t> k_alt;
    DevBuf<uint32_t> v_alt;
    DevBuf<uint32_t> hist;
    DevBuf<uint64_t> hist_scan;
    Scanner scan;
    void sort(uint64_t *keys, uint32_t *vals, size_t n, int bits, hipStream_t s);
};

// Sum of a u64 array into *dst (device), *dst cleared first.
void reduce_sum_u64(const uint64_t *in, size_t n, unsigned long long *dst, hipStream_t s);

inline unsigned grid_for(size_t n, unsigned block)
{
    size_t g = (n + block - 1) / block;
    return (unsigned)(g ? g : 1);
}

// A dispatch covers fewer than 2^32 work-items (the AQL packet's grid size
// is 32 bits): larger launches would be truncated silently, so they are
// refused here and the kernels that can be that large loop over their work.
constexpr unsigned MAX_GRID = 1u << 20;  // blocks of a grid-stride launch
inline void check_grid(size_t grid, unsigned block)
{
    if (grid == 0 || grid >= (1ull << 31) || (uint64_t)grid * block >= (1ull << 32))
        throw Error(MSJ_E_OVERFLOW, "kernel launch exceeds 2^32 work-items");
}
inline unsigned grid_stride(size_t n, unsigned block)
{
    size_t g = (n + block - 1) / block;
    return (unsigned)(g == 0 ? 1 : (g < MAX_GRID ? g : MAX_GRID));
}

inline int bits_for(uint64_t maxval)  // bits needed to represent 0..maxval
{
    int b = 0;
    while (b < 64 && (maxval >> b) != 0)
        ++b;
    return b;
}

// ------------------------------------------------------------- node net
// Final node net in HBM (the reference's std::vector<Node>, nodes.cuh:9-38,
// flattened to CSR).  Nodes are in the reference's node order; points within
// a node ascend by row id.
struct Index {
    uint32_t M = 0;              // nodes
    uint32_t edges = 0;          // adjacency entries (self included)
    uint64_t num_calcs = 0;      // sum of numCalcs
    DevBuf<uint32_t> perm;       // [n]   node-order position -> row (pointArray)
    DevBuf<uint32_t> nstart;     // [M+1] node boundaries
    DevBuf<uint32_t> aoff;       // [M+1] neighbour-node CSR offsets
    DevBuf<uint32_t> adj;        // [edges]
    DevBuf<uint64_t> sumsz;      // [M]   points in the neighbour nodes
    DevBuf<uint64_t> ncalcs;     // [M]   numCalcs = size * sumsz
};

struct LayerStats {
    uint32_t layers = 0, ncand = 0;
    uint32_t chosen[MSJ_MAX_LAYERS] = {};
    uint32_t nodes[MSJ_MAX_LAYERS] = {};
    uint64_t calcs[MSJ_MAX_LAYERS] = {};
    double sse[MSJ_MAX_LAYERS] = {};
};

// glibc random_r TYPE_3 stream (the generator behind the reference's rand(),
// utils.cu:77); re-seeded at the start of every join.
struct GlibcRand {
    int32_t r[31];
    int f = 3, b = 0;
    void seed(uint32_t s);
    int32_t next();
};

// kernel.cu:895-905 key of a reference-point partition from the FP64 squared
// distance D: floor(sqrt(D) / eps), IEEE sqrt and division, converted as
// cvt.rzi.u32.f64 does (truncate, saturate, NaN -> 0).
__device__ inline uint32_t rp_key(double D, double eps)
{
    const double q = floor(sqrt(D) / eps);
    if (!(q > 0.0))
        return 0u;
    if (q >= 4294967295.0)
        return 0xFFFFFFFFu;
    return (uint32_t)q;
}

// nodes.cu:130 key of an axis partition: ceil((x + 1e-7) / eps), converted
// as x86-64 host code does ((unsigned) of a double: cvttsd2si to 64 bits,
// keep the low half; out-of-range inputs give 0x8000000000000000).
__device__ inline uint32_t axis_key(double x, double eps)
{
    const double q = ceil((x + 0.0000001) / eps);
    if (!(q > -9223372036854775808.0 && q < 9223372036854775808.0))
        return 0u;
    return (uint32_t)(uint64_t)(int64_t)q;
}

// tree.cu:284 key of an axis partition of the tree/COSS index:
// ceil((x + 1e-7) / eps) + 10, converted as x86-64 host code does.
__device__ inline uint32_t coss_axis_key(double x, double eps)
{
    const double q = ceil((x + 0.0000001) / eps) + 10.0;
    if (!(q > -9223372036854775808.0 && q < 9223372036854775808.0))
        return 0u;
    return (uint32_t)(uint64_t)(int64_t)q;
}

// Test probe of an index build (host buffers, filled when set).
struct IndexProbe {
    uint32_t layer = ~0u;        // layer whose candidate keys are copied out
    uint32_t *keys = nullptr;    // [ncand * n] keys of every candidate, row order
    bool node_keys = false;      // keep each layer's (parent, key) of the children
    std::vector<uint32_t> cparent[MSJ_MAX_LAYERS], ckey[MSJ_MAX_LAYERS];
    std::vector<uint32_t> tuples;  // tree/COSS: the nodes' bin numbers, [M * layers]
};

// s after g steps of s = RN(s + c), for finite s >= 0 and c >= 0 normal,
// in a few steps per binade of s instead of g.  Inside one binade [2^(e-1),
// 2^e) with ulp u, c = q*u + r: with r != u/2 every step adds the same
// rounded multiple of u; with r == u/2 (a tie) the step rounds to an even
// significand and, once the significand is even, adds the same multiple of
// u again every step.  So after a step that lands in the binade, the next
// step that stays in it gives the steady increment, and the run of steps
// that keeps s below 2^e - u is added at once, exactly, in units of u.
__host__ __device__ inline double repeat_add(double s, double c, uint64_t g)
{
    if (!(c > 0.0))
        return s;
    while (g > 0) {
        s = s + c;
        if (--g == 0)
            break;
        int es, et;
        const double t = s + c;  // lands even if it was a tie
        --g;
        (void)frexp(s, &es);
        (void)frexp(t, &et);
        s = t;
        if (es != et || g == 0)
            continue;
        const double t2 = s + c;  // from an even significand: the steady step
        --g;
        (void)frexp(t2, &et);
        const double delta = t2 - s;  // exact when both lie in the binade
        s = t2;
        if (es != et)
            continue;
        if (delta == 0.0)  // c below half an ulp: s no longer moves
            break;
        const double u = ldexp(1.0, es - 53);
        const uint64_t room = (uint64_t)((ldexp(1.0, es) - u - s) / u);  // exact
        const uint64_t D = (uint64_t)(delta / u);
        uint64_t k = room / D;
        if (k > g)
            k = g;
        s = s + ldexp((double)(k * D), es - 53);
        g -= k;
    }
    return s;
}

struct IndexBuilder {
    IndexProbe *probe = nullptr;
    // per-layer scratch
    DevBuf<uint32_t> keys, keys2;  // [ncand*n] candidate keys, row order (this, next layer)
    DevBuf<double> test_rp, distmat, sigma, maxc, rpT;
    DevBuf<uint64_t> ck;         // composite keys
    DevBuf<uint32_t> flags, runidx, starts, runfirst, cnt_nodes;
    DevBuf<uint32_t> code, tab;  // dense candidate bins: point -> parent's first bin, counts
    DevBuf<double> sse, best_sse;
    DevBuf<uint32_t> best;       // per layer: the winning candidate
    DevBuf<uint32_t> vals, pnode, pnode2, perm2;
    DevBuf<uint8_t> nsplit, nsplit2;
    DevBuf<uint32_t> cparent, ckey, cstart, cfirst, deg, aoff2, adj2, aoff1, adj1;
    DevBuf<uint64_t> sumsz2, ncalcs2, tmp64;
    DevBuf<uint32_t> u32scalar;
    DevBuf<unsigned long long> u64scalar;
    // tree/COSS index
    DevBuf<uint32_t> cmm, cslots, prank, bnum, tup, cwalk;
    DevBuf<unsigned long long> cacc;
    DevBuf<double> cscore;
    Scanner scan;
    RadixSorter sorter;
    // Builds the candidate index of `xs` (SoA, row order, storage T) into
    // `out`: the node net, or the tree/COSS bins when p.index asks for them.
    void build(const void *xs, bool f32, uint32_t n, uint32_t d, double eps, const msj_params &p,
               Index &out, LayerStats &st, hipStream_t s);
};

// utils.cu:22-41 population sigma of `cols` columns, each summed sequentially
// in row order (bit-identical to the host loop); element (r, c) is
// A[r * rs + c * cs] (row-major: rs = cols, cs = 1; SoA: rs = 1, cs = rows).
void column_sigma(const double *A, uint32_t rows, uint32_t cols, size_t rs, size_t cs,
                  double *sigma, hipStream_t s);

// Candidate generator of the brute force: one node holding every row (in
// row order), adjacent to itself only, so every pair is a candidate.
void build_single_node(uint32_t n, Index &out, LayerStats &st, hipStream_t s);

// ------------------------------------------------------------------- join
// One wave: queries [qbeg, qend) of node x (offsets inside the node) against
// entries [cbeg, cend) of x's upper candidate list.
struct JoinItem {
    uint32_t x, qbeg, qend, cbeg, cend;
};

// Everything a wave needs to run the work items of one join (plain form).
struct JobView {
    const double *xr;            // FP64 rows in join order, ds doubles each
    uint32_t d, ds;
    double eps2;
    const uint32_t *nstart, *uoff, *uadj, *perm;  // perm: join position -> row id
    const uint32_t *ucum;        // per upper entry: first offset in the node's list
    const JoinItem *items;
    uint32_t nitems;
    int qbits;
    unsigned long long *out, *counter;
    unsigned long long cap;
};

struct JoinRunner {
    DevBuf<uint32_t> itcnt, itoff, uoff, uadj, ucum;
    DevBuf<uint64_t> ulen, tot64, trow;
    DevBuf<unsigned long long> tsum;
    DevBuf<JoinItem> items;
    DevBuf<unsigned long long> counter, work_ctr;
    DevBuf<uint64_t> pairs;      // (row << qbits | neighbour) keys
    DevBuf<uint32_t> jperm;      // join order: node order, Morton order inside a node
    DevBuf<uint64_t> jkeys;
    DevBuf<double> mm;
    DevBuf<double> pca;          // sample, covariance, directions, projections
    uint32_t nitems = 0;
    uint64_t tile_rows = 0;      // sum over upper entries (X, Y) of |X| + |Y|
    Scanner scan;
    RadixSorter sorter;
    // Join order of the points: nodes as in ix, points inside a node sorted
    // along a Morton curve of the first three columns (jperm).
    void order(const Index &ix, const void *xs, bool f32, uint32_t n, uint32_t d, hipStream_t s);
    // Upper neighbour lists and work items of the final node net.
    // plain_items = false: upper lists only (the MFMA form has its own items).
    void plan(const Index &ix, uint32_t n, uint32_t d, hipStream_t s, bool plain_items = true,
              uint64_t evals_per_item = 0);
    // Descriptor of this join for the kernels (after plan()).
    JobView view(const Index &ix, const double *xr, uint32_t n, uint32_t d, double eps) const;
    // Runs the plain FP64 distance kernel on the node-ordered row-major FP64
    // copy xr; returns the number of pairs R.  Grows `pairs` and reruns when
    // R exceeds the capacity.  With dims_evaluated, a counting build of the
    // kernel also sums the FP64 dimensions the early-exit test consumed.
    // progress != 0: the items run in 16 launches, each followed by a line
    // on stderr (long joins stay observable).
    uint64_t run_plain(const Index &ix, const double *xr, uint32_t n, uint32_t d, double eps,
                       uint64_t cap_hint, hipStream_t s, hipEvent_t k0, hipEvent_t k1,
                       uint64_t *launches, float *kernel_ms, uint64_t *dims_evaluated,
                       bool progress = false);
    // The plain kernel of the planned items, asynchronously on s (counter
    // zeroed first); the pair count is *counter once s gets there, and a
    // count above pairs.cap means a rerun with run_plain.
    void launch_plain(const Index &ix, const double *xr, uint32_t n, uint32_t d, double eps,
                      hipStream_t s);
    // Sorts the R pairs and writes counts[n], offsets[n+1], nbrs[R].
    void finish(uint32_t n, uint64_t R, uint32_t *counts, uint64_t *offsets, uint32_t *nbrs,
                hipStream_t s);
};

// ------------------------------------------------------------ MFMA filter
// A super-block of the filter: s x s tiles of 128 x 128 pairs, A rows from
// qbeg (node X ends at qend) against B rows from cbeg (node Y ends at cend),
// in the padded row space (nodes start on 32-row boundaries).  self: s in
// bits 0-7, bit 8 X == Y (pairs p <= c only), bit 9 a diagonal super-block
// (tiles on or above its diagonal only).
struct MItem {
    uint32_t qbeg, qend, cbeg, cend, self;
};

// Tile slots -> super-blocks: items of class c (s = 8 - c) start at
// item_base[c] and own slots [slot_base[c], slot_base[c+1]), s*s each.
struct SlotMap {
    unsigned long long slot_base[9];
    unsigned long long item_base[8];
};

// Precision form: pairs are scored on the matrix cores from BF16 (or
// split-BF16) copies of the centred rows, FP32 accumulation; a rigorous
// per-pair error bound turns eps^2 into a band, and only pairs inside the
// band are decided again with the FP64 sequence of the plain form.
struct MfmaRunner {
    DevBuf<uint32_t> pstart;     // [M+1] padded row offset of each node
    DevBuf<uint32_t> prow;       // [M]   roundup(size, 32)
    DevBuf<uint32_t> pjoin;      // [P]   padded row -> join position (~0u: padding)
    DevBuf<uint32_t> porig;      // [P]   padded row -> caller's row id
    DevBuf<uint4> fh, fl;        // fragments of the BF16 high / low parts
    DevBuf<float4> thr;          // [P]   per-row threshold terms
    DevBuf<float2> thrb;         // [P]   the same under the BF16 bound (split-BF16)
    DevBuf<double> mu, colpart;  // centre (column means) and partial sums
    DevBuf<uint64_t> p64;        // scan scratch
    DevBuf<uint32_t> cnt8;       // [8][E] super-blocks per (class, upper edge)
    DevBuf<uint64_t> etiles;     // per upper edge: 128 x 128 tiles
    DevBuf<uint64_t> off8;       // exclusive scan of cnt8
    DevBuf<uint32_t> esrc;       // upper edge -> source node
    DevBuf<MItem> items;
    DevBuf<uint64_t> band;       // band pairs (join positions, class bits)
    DevBuf<float> bscore;        // their filter scores
    DevBuf<unsigned long long> ctr;  // [0] pairs, [1] band, [2] err bits, [3] checked,
                                     // [4] violations, [5] err sum (double bits)
    uint32_t P = 0, KS = 0;
    uint64_t nitems = 0;         // super-blocks
    uint64_t tiles = 0;          // 128 x 128 tiles of the last prepare()
    SlotMap smap = {};
    DevBuf<SlotMap> smapd;       // the same in device memory (the kernel reads it there)
    bool split = false;
    Scanner scan;
    // Padded rows, centre, fragments and thresholds from the node-ordered
    // FP64 rows xr (join order of `J`), and the work items over J's upper
    // neighbour lists.
    void prepare(const Index &ix, const JoinRunner &J, const double *xr, uint32_t n,
                 uint32_t d, double eps, bool split_bf16, hipStream_t s);
    // Filter + band recheck; writes the result keys into J.pairs (grown and
    // rerun on overflow) and returns their number.
    // kernel_ms: filter + band kernels; main_ms: the filter kernel alone
    // (events k0 .. kmid).
    uint64_t run(const Index &ix, JoinRunner &J, const double *xr, uint32_t n, uint32_t d,
                 double eps, uint64_t cap_hint, bool diag, hipStream_t s, hipEvent_t k0,
                 hipEvent_t kmid, hipEvent_t k1, uint64_t *launches, float *kernel_ms,
                 float *main_ms, msj_stats *st);
    uint64_t last_band = 0;
};

// One join's device state and its two host steps (the body of
// nodeLauncher5, launcher.cu:3093-3532).
struct JoinCore {
    Index ix;
    IndexBuilder builder;
    JoinRunner join;
    MfmaRunner mfma;
    DevBuf<double> xr;           // FP64 rows in join order
    // Candidate generation and distance kernels on xs (SoA, row order);
    // rows: the same data row-major, or null.  Returns the pair count R
    // (keys in join.pairs).  cap: initial result-key capacity (0 under an
    // arena: half of what it has left).  Events ev[0..7] on stream s time
    // the stages; st receives the statistics.
    uint64_t run(const void *xs, const void *rows, bool f32, uint32_t n, uint32_t d, double eps,
                 const msj_params &prm, uint64_t cap, hipStream_t s, hipEvent_t *ev,
                 msj_stats &st);
    // Sorts the pairs into counts[n], offsets[n+1], nbrs[R] and synchronises.
    void finish(uint32_t n, uint64_t R, uint32_t *counts, uint64_t *offsets, uint32_t *nbrs,
                hipStream_t s, hipEvent_t *ev, msj_stats &st);
    // The reference's batch record (launcher.cu:3203-3233) of the last run
    // into st, once: finish computes it while the pair sort runs.
    void batch_record(msj_stats &st);
    std::vector<uint32_t> hstart;  // node starts and numCalcs of the last run
    std::vector<uint64_t> hcalcs;
    uint32_t rec_n = 0;
    double rec_eps = 0.0;
    uint64_t rec_calcs = 0;
    bool rec_pending = false;
};

// The filter's bounds hold for 2^-50 <= eps^2 <= 2^60, n < 2^30, d <= 4096.
bool mfma_filter_applies(double eps, uint32_t n, uint32_t d);

// Checks every field of msj_params (MSJ_E_INVAL with the reason).
void validate_params(const msj_params &p);

// Row-major device rows (n x d, 4- or 8-byte values) -> SoA device copy xs.
void rows_to_soa_device(const void *rows, uint32_t n, uint32_t d, bool f32, void *xs,
                        hipStream_t s);

// Row-major host data (n x d, 4- or 8-byte values) -> SoA device copy xs.
void load_soa(const void *data, uint32_t n, uint32_t d, bool f32, DevBuf<unsigned char> &xs,
              DevBuf<unsigned char> &staging, hipStream_t s);

// Batching form (msj_batch.hip); the C entry points are in msj_api.hip.
struct BatchEngine;
BatchEngine *batch_create(const msj_params &p);
void batch_destroy(BatchEngine *b);
void batch_load(BatchEngine *b, const msj_job *jobs, uint32_t njobs);
void batch_run(BatchEngine *b, msj_result *outs, msj_batch_stats *st);

// Row width of the node-ordered FP64 copy: d rounded up to a multiple of 4.
inline uint32_t padded_dim(uint32_t d) { return (d + 3u) & ~3u; }

// xr[i*ds + j] = FP64 value of column j of node-order row i (row perm[i]),
// ds = padded_dim(d), zero padding; from `rows` (row-major, storage type)
// when given, else from the columns `xs`.
void gather_node_rows(const void *xs, const void *rows, double *xr, bool f32,
                      const uint32_t *perm, uint32_t n, uint32_t d, hipStream_t s);

}  // namespace msj
