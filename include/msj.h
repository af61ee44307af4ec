/*
 * msj.h -- C ABI of the MI355X-native MiSTIC epsilon self-join.
 *
 * Drop-in boundary for the reference's GPU join path:
 *   struct neighborTable* nodeLauncher5(double* data, unsigned dim,
 *       unsigned numPoints, unsigned numRP, unsigned* pointArray, double epsilon)
 *   (reference include/launcher.cuh:77-82, src/launcher.cu:3093-3532; called
 *   once from main, src/main.cu:171-178).
 *
 * Plain C types only: no C++, HIP or torch types cross this boundary.  Every
 * entry point returns MSJ_OK or a negative msj_status and never aborts;
 * msj_last_error() gives the detail of the last failure on the calling thread.
 *
 * Data contract (as nodeLauncher5): host, row-major, n rows of d values,
 * dimensions already reordered by decreasing standard deviation (main.cu:106-122,
 * see msj_dim_order).  The library never writes caller memory except *out.
 *
 * Results: every ordered pair (p, q) with the reference's FP64 distance test
 * ||p - q||^2 <= eps^2 (kernel.cu:833-865), including (p, p); rows and
 * neighbour ids are the caller's row indices (0-based input order).
 */
#ifndef MSJ_H
#define MSJ_H

#include <stddef.h>
#include <stdint.h>

#include "msj_params.h"

#ifdef __cplusplus
extern "C" {
#endif

#define MSJ_ABI_VERSION 2u
#define MSJ_MAX_LAYERS 16

typedef enum msj_status {
    MSJ_OK = 0,
    MSJ_E_INVAL = -1,       /* bad argument (n, d, eps, params)               */
    MSJ_E_NOMEM = -2,       /* host or device allocation failed               */
    MSJ_E_HIP = -3,         /* a HIP runtime call failed                      */
    MSJ_E_OVERFLOW = -4,    /* a count exceeds what the ABI types can hold    */
    MSJ_E_UNSUPPORTED = -5, /* valid request this build does not implement    */
    MSJ_E_NODEV = -6        /* no usable HIP device                           */
} msj_status;

/* Storage precision of the input coordinates.  F32 inputs are widened
 * exactly to FP64 for every decision. */
typedef enum msj_storage { MSJ_STORAGE_F64 = 0, MSJ_STORAGE_F32 = 1 } msj_storage;

/* Distance-and-compare form.  All forms return the same pair set. */
typedef enum msj_form {
    MSJ_FORM_PLAIN = 0,      /* FP64 chunked early-exit test on every candidate */
    MSJ_FORM_MFMA_BF16 = 1,  /* BF16 MFMA Gram filter + FP64 band recheck       */
    MSJ_FORM_MFMA_SPLIT = 2  /* split-BF16 MFMA Gram filter + FP64 band recheck */
} msj_form;

/* Candidate generator: which pairs the distance test is applied to.  Every
 * generator yields the same pair set (the distance test decides); they differ
 * in how many candidates they evaluate. */
typedef enum msj_index {
    MSJ_INDEX_NODENET = 0,   /* the reference's node net (nodes.cu:5-356)        */
    MSJ_INDEX_BRUTE = 1,     /* every pair: one node holding all rows; the
                                verifier's ground truth (nodes.cu:749-885)     */
    MSJ_INDEX_COSS = 2       /* the reference's tree/COSS bins (tree.cu:3-996)   */
} msj_index;

/* Where msj_ctx_join leaves the result arrays. */
typedef enum msj_output {
    MSJ_OUT_HOST = 0,   /* library-allocated host arrays in msj_result          */
    MSJ_OUT_DEVICE = 1  /* arrays stay in HBM, owned by the context (see below) */
} msj_output;

typedef struct msj_params {
    uint32_t struct_size;        /* sizeof(msj_params), set by msj_default_params   */
    int32_t  device;             /* HIP device ordinal (params.cuh:40 CUDA_DEVICE)  */
    uint32_t maxrp;              /* MAXRP: node-net layers (params.cuh:50)          */
    uint32_t rp_per_layer;       /* RPPERLAYER, <= 32 (params.cuh:53)               */
    uint32_t numidim;            /* NUMIDIM (params.cuh:61); used for j < min(.,d)  */
    uint32_t seed;               /* glibc srand seed of the RP stream; 1 = unseeded */
    double   sample_per;         /* SAMPLE_PER (params.cuh:69)                      */
    uint64_t min_calcs_per_node; /* MIN_CALCS_PER_NODE (params.cuh:38)              */
    uint32_t form;               /* msj_form                                        */
    uint32_t storage;            /* msj_storage of the input                        */
    uint32_t output;             /* msj_output                                      */
    uint32_t verbose;            /* 1: per-layer progress on stdout                 */
    uint32_t count_work;         /* 1: plain form counts the FP64 dimensions
                                    evaluated; MFMA forms verify their filter
                                    classes in FP64 (both slower)                 */
    uint32_t index;              /* msj_index: candidate generator (default node net) */
    uint64_t page_pairs;         /* MSJ_OUT_HOST: neighbour ids per page of the
                                    pinned double buffer (0: 2^24)                 */
} msj_params;

typedef struct msj_stats {
    uint32_t num_nodes;          /* final node count                                */
    uint32_t num_layers;         /* layers built (numSplits)                        */
    uint32_t num_candidates;     /* candidate partitions per layer                  */
    uint32_t num_pages;          /* pages of the neighbour list copied to host      */
    uint64_t num_calcs;          /* candidate distance evaluations (nodes.cu:697)   */
    uint64_t total_pairs;        /* "Total results Set Size" (launcher.cu:3513)     */
    uint32_t chosen[MSJ_MAX_LAYERS];          /* selected candidate per layer       */
    uint32_t nodes_per_layer[MSJ_MAX_LAYERS];
    uint64_t calcs_per_layer[MSJ_MAX_LAYERS];
    double   sse[MSJ_MAX_LAYERS];             /* selection statistic of the winner  */
    uint32_t ref_num_batches;    /* batch count of the reference's plan (3203-3229) */
    uint32_t pad0;
    double   ref_var_calc;       /* varCalc of that plan (launcher.cu:3203)         */
    /* device time, HIP events on the compute stream, milliseconds */
    double   ms_index;           /* reference points, keys, node net, adjacency     */
    double   ms_join;            /* distance-and-compare kernels                    */
    double   ms_join_kernel;     /* the same, kernel launches only                  */
    double   ms_output;          /* pair sort, counts, offsets                      */
    double   ms_d2h;             /* result pages to host (MSJ_OUT_HOST), host clock */
    double   ms_total;           /* host wall clock of msj_ctx_join                 */
    uint64_t join_launches;      /* distance kernel launches (incl. retries)        */
    uint64_t dims_evaluated;     /* count_work: sum over evaluated pairs of the
                                    dimensions the early-exit test consumed       */
    uint64_t band_pairs;         /* MFMA forms: pairs re-decided in FP64            */
    double   max_score_err;      /* MFMA forms: max |filter score - FP64 score|     */
    double   mean_score_err;     /* MFMA forms: mean of the same over scored pairs  */
    uint64_t filter_checked;     /* MFMA forms with count_work: sampled candidate
                                    pairs whose filter class was verified in FP64  */
    uint64_t filter_violations;  /* of those, certain-accept/reject classes that
                                    FP64 contradicts (0 unless the bound is wrong) */
    uint32_t form_used;          /* msj_form the join actually ran                  */
    uint32_t split_share_ppm;    /* split-BF16 form: share of the tiles' pairs
                                    whose cross products ran, per million (the
                                    others were certain rejects in BF16)        */
    double   mfma_flops;         /* MFMA forms: matrix-core flops issued (tiles)    */
    double   ms_main_kernel;     /* the dominant distance kernel alone, HIP events:
                                    k_join_sym / k_join_hd (plain) or k_mfma_join   */
    uint64_t tile_rows;          /* sum over adjacent node pairs X <= Y of |X|+|Y|:
                                    rows read once per candidate tile; 8*d bytes
                                    each is the join's algorithmic HBM traffic     */
    uint64_t filter_checked_bf16;/* split-BF16 form with count_work: of the verified
                                    pairs, those the BF16 first pass scored as
                                    certain rejects (the test that lets a block
                                    skip the cross products)                      */
} msj_stats;

typedef struct msj_result {
    uint32_t  n;                 /* rows                                            */
    uint32_t  pad0;
    uint64_t  total_pairs;       /* R = sum(counts)                                 */
    uint32_t *counts;            /* [n]    neighbours of each row, self included    */
    uint64_t *offsets;           /* [n+1]  CSR offsets into nbrs                    */
    uint32_t *nbrs;              /* [R]    neighbour row ids, ascending per row     */
    uint32_t *node_order;        /* [n]    pointArray after the node-net build      */
    uint32_t  on_device;         /* 1: the four arrays above are device pointers    */
    uint32_t  pad1;
    msj_stats stats;
} msj_result;

typedef struct msj_ctx msj_ctx;

const char *msj_version(void);
const char *msj_strerror(int status);
const char *msj_last_error(void);
void        msj_default_params(msj_params *p);

/* One-shot join (replaces nodeLauncher5, launcher.cu:3093): upload, build the
 * node net, join, return host arrays.  Free with msj_result_free. */
int  msj_self_join(const void *data, uint32_t n, uint32_t d, double eps,
                   const msj_params *p, msj_result *out);
void msj_result_free(msj_result *r);

/* Context API: keeps the data resident in HBM across joins. */
int  msj_ctx_create(const msj_params *p, msj_ctx **ctx);
void msj_ctx_destroy(msj_ctx *ctx);
/* Replaces the context's parameters for later joins (form, candidate
 * generator, node-net knobs, output, ...); device and storage must stay as
 * they were at creation (MSJ_E_INVAL otherwise).  The loaded data stays. */
int  msj_ctx_set_params(msj_ctx *ctx, const msj_params *p);
/* Uploads n x d row-major values of p->storage precision. */
int  msj_ctx_load(msj_ctx *ctx, const void *data, uint32_t n, uint32_t d);
/* main.cu:42-50 + 106-122 in one call: n rows of d FP64 values in file
 * order (dimensions NOT yet reordered) are streamed to the device through a
 * bounded pinned buffer (rows may be a read-only mapping of the input file,
 * larger than host memory), the reference's per-dimension sigma is computed
 * on the device (bit-identical to stddev(), utils.cu:22-53) and the
 * dimensions are kept in decreasing-sigma order at p->storage precision.
 * order[d] / sigma[d] (may be NULL) receive that order and the sorted
 * sigmas, as msj_dim_order returns them. */
int  msj_ctx_load_unordered(msj_ctx *ctx, const double *rows, uint32_t n, uint32_t d,
                            uint32_t *order, double *sigma);
/* Runs the whole join on the resident data.  With MSJ_OUT_DEVICE the result
 * arrays are device pointers valid until the next join on, or the
 * destruction of, this context; msj_result_free does not free them. */
int  msj_ctx_join(msj_ctx *ctx, double eps, msj_result *out);

/* ---- Batching form: many independent joins of uneven sizes and dimensions.
 * No counterpart in the reference (its main runs one join per process); the
 * per-job semantics are msj_self_join's.  A run is a pipeline over four
 * streams: two builder threads build the jobs' node nets (largest first);
 * a job large enough to fill the GPU is joined as soon as its net is built,
 * overlapping the builds of the next jobs (the reference's NUMSTREAMS
 * overlap, params.cuh:63), in params.form (with an MFMA form, by the filter
 * and its FP64 band recheck where msj_self_join would use them; the
 * result is the same); the smaller jobs share one persistent plain kernel per
 * row width, replayed from a captured hipGraph, that drains a device work
 * queue of their work items; an output thread sorts each job's pairs as its
 * join completes.  Tile shapes of the persistent kernel (work per item,
 * workgroups per CU, and for rows wider than 32 the dimensions per slab)
 * are chosen per row width by a built-in autotuner on the first run and
 * kept by the batch. */
typedef struct msj_job {
    const void *data;            /* host, n x d row-major, dims sigma-ordered,
                                    precision of params.storage                   */
    uint32_t    n;
    uint32_t    d;
    double      eps;
} msj_job;

typedef struct msj_batch_stats {
    uint32_t njobs;
    uint32_t workers;            /* host threads / streams building node nets      */
    uint32_t grid_blocks;        /* persistent launch: workgroups                  */
    uint32_t classes;            /* distinct padded row widths                     */
    uint32_t class_ds[8];        /* per class: padded row width                    */
    uint32_t class_epi_log2[8];  /* per class: log2 of lane evaluations per item   */
    uint64_t work_items;         /* items in the device queue                      */
    uint64_t total_pairs;        /* sum over jobs                                   */
    uint64_t num_calcs;          /* sum over jobs                                   */
    double   ms_index;           /* host clock until every node net, join order and
                                    plan is built (joins run meanwhile)            */
    double   ms_join;            /* device time of the join stream: the solo joins
                                    and the persistent launch                      */
    double   ms_output;          /* host clock spent on pair sorts and CSR arrays
                                    (overlapping later joins)                      */
    double   ms_tune;            /* host clock: autotuning (first run of a class)   */
    double   ms_total;           /* host clock of msj_batch_run                     */
    uint32_t graph_replays;      /* launches through the captured graph             */
    uint32_t solo_jobs;          /* jobs joined alone (enough work to fill the GPU)  */
    uint32_t class_slab[8];      /* per class: dimensions per slab of the wide-row
                                    kernel (8 or 16; 0 for widths <= 32)         */
} msj_batch_stats;

typedef struct msj_batch msj_batch;

int  msj_batch_create(const msj_params *p, msj_batch **b);
void msj_batch_destroy(msj_batch *b);
/* Uploads the jobs' data (copied; the caller's buffers may be freed after). */
int  msj_batch_load(msj_batch *b, const msj_job *jobs, uint32_t njobs);
/* Joins every loaded job; outs has njobs entries (free each with
 * msj_result_free; with MSJ_OUT_DEVICE the arrays stay valid until the next
 * run or msj_batch_destroy).  st may be NULL. */
int  msj_batch_run(msj_batch *b, msj_result *outs, msj_batch_stats *st);
/* One-shot: create, load, run, destroy. */
int  msj_join_batch(const msj_job *jobs, uint32_t njobs, const msj_params *p, msj_result *outs);

/* ---- Allocation-free join.  No counterpart in the reference: for callers
 * that cannot allocate device memory (serving loops that must not call
 * hipMalloc, code running between its own allocations).  The same join as
 * msj_self_join, with the caller's device buffers in and out:
 *   rows        device, n x d row-major, dims sigma-ordered, p->storage
 *   counts      device [n]; offsets device [n+1]; node_order device [n]
 *   nbrs        device [nbrs_cap]; a larger result returns MSJ_E_OVERFLOW
 *               with stats->total_pairs = R (call again with room for R)
 *   stats       host, may be NULL
 * Every other device buffer comes from one file-scope device array of
 * MSJ_NOALLOC_SCRATCH_BYTES that each device holds from the library's first
 * kernel launch on it until the process ends; inputs whose buffers do not
 * fit return MSJ_E_OVERFLOW (the data copies alone need n*d*(4 or 8) +
 * n*roundup(d,4)*8 bytes) or MSJ_E_NOMEM (scratch exhausted, with the sizes
 * in msj_last_error).  Calls on one device share the array: they run one at
 * a time (a per-device lock), each clearing what the previous one used.
 * Synchronous on the device's null stream; p->output is ignored. */
#define MSJ_NOALLOC_SCRATCH_BYTES (256ull << 20)
int  msj_self_join_noalloc(const void *rows, uint32_t n, uint32_t d, double eps,
                           const msj_params *p, uint32_t *counts, uint64_t *offsets,
                           uint32_t *nbrs, uint64_t nbrs_cap, uint32_t *node_order,
                           msj_stats *stats);

/* utils.cu:22-53 stddev(): dimension order by decreasing population standard
 * deviation (ties: higher index first), as main.cu:111 uses it. */
int  msj_dim_order(const double *data, uint32_t n, uint32_t d, uint32_t *order, double *sigma);

#ifdef __cplusplus
}
#endif
#endif
