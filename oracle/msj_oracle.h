/*
 * msj_oracle.h -- CPU restatement of MiSTIC's epsilon self-join (node-net path).
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker and the CPU
 * baseline.  Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline
 * leg may load it.  The product (libmsj.so) never links, loads or calls it.
 *
 * Every function restates the reference path it cites (file:line in
 * bwd29/self-join-MiSTIC).  Plain C11 + OpenMP; the FP64 operation order is
 * fixed explicitly (build with -ffp-contract=off; fused steps call fma()).
 */
#ifndef MSJ_ORACLE_H
#define MSJ_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OR_MAXL 16

/* Distance operation sequences (reference kernel.cu:833-865).
 *   OR_SEQ_NVVM7  : per 4-wide chunk s=fma(d0,d0,s); s=RN(s+RN(dj^2)) j=1..3
 *                   (nvcc 11.x contraction of cachedDistanceCheck; default)
 *   OR_SEQ_ALLFMA : s=fma(dj,dj,s) for every j (contraction across statements)
 *   OR_SEQ_NOFMA  : s=RN(s+RN(dj^2)) for every j (no contraction)          */
enum { OR_SEQ_NVVM7 = 0, OR_SEQ_ALLFMA = 1, OR_SEQ_NOFMA = 2 };

typedef struct or_params {
    uint32_t maxrp;              /* MAXRP: node-net layers (params.cuh:50) = 6      */
    uint32_t rp_per_layer;       /* RPPERLAYER (params.cuh:53) = 32                 */
    uint32_t numidim;            /* NUMIDIM (params.cuh:61) = 6                     */
    uint32_t seed;               /* glibc srand seed; 1 == never seeded (default)   */
    double   sample_per;         /* SAMPLE_PER (params.cuh:69) = 0.01               */
    uint64_t min_calcs_per_node; /* MIN_CALCS_PER_NODE (params.cuh:38) = 100000     */
    uint32_t seq;                /* OR_SEQ_*                                        */
    uint32_t join;               /* 0: build the index only; 1: join               */
    uint32_t sample_every;       /* join only node-order positions % k == 0 (0,1: all) */
    int32_t  nthreads;           /* OpenMP threads, <= 0: runtime default           */
    uint32_t index;              /* candidate generator: 0 node net (nodes.cu),
                                    2 tree/COSS bins (tree.cu)                      */
} or_params;

typedef struct or_result {
    uint32_t n, d;
    uint32_t num_nodes;
    uint32_t num_layers;
    uint32_t num_candidates;     /* candidate partitions per layer (RPs + axes)     */
    uint32_t pad0;
    uint64_t num_calcs;          /* sum of node numCalcs (nodes.cu:697-706)         */
    uint64_t evals_done;         /* candidate distance evaluations performed        */
    uint64_t total_pairs;        /* ordered pairs incl. self ("totals")             */
    uint32_t *counts;            /* [n] neighbours per original row                 */
    uint64_t *offsets;           /* [n+1]                                           */
    uint32_t *nbrs;              /* [total_pairs] ascending original ids per row    */
    uint32_t *node_order;        /* [n] pointArray after buildNodeNet               */
    uint32_t *node_start;        /* [num_nodes+1] node boundaries in node order     */
    uint32_t *adj_off;           /* [num_nodes+1] neighbour-node CSR                */
    uint32_t *adj;               /* [adj_off[num_nodes]] neighbour node ids         */
    uint32_t *node_keys;         /* [num_nodes * num_layers] binNumbers per node;
                                    0xFFFFFFFF from the layer a node stopped splitting */
    uint32_t chosen[OR_MAXL];    /* selected candidate per layer                    */
    double   sse[OR_MAXL];       /* its sum of squared node-size deviations         */
    uint32_t nodes_per_layer[OR_MAXL];
    uint64_t calcs_per_layer[OR_MAXL];
    uint64_t flips_allfma;       /* result pairs decided differently by ALLFMA      */
    uint64_t flips_nofma;        /* result pairs decided differently by NOFMA       */
    double   t_construct, t_join, t_output;
} or_result;

void     or_default_params(or_params *p);

/* glibc random()/rand() TYPE_3 stream after srand(seed). */
void     or_glibc_rand(uint32_t seed, uint32_t count, int32_t *out);

/* utils.cu:22-53 stddev(): population sigma per dim, order by sigma descending
 * (stable ascending sort on the IEEE bit order, then reversed). */
void     or_dim_order(const double *A, uint32_t n, uint32_t d,
                      uint32_t *order, double *sigma);

/* utils.cu:66-137 createRPArray(), called once per layer (nodes.cu:65):
 * writes layers*rp_per_layer*d doubles. */
void     or_rp_layers(const double *data, uint32_t n, uint32_t d,
                      const or_params *p, uint32_t layers, double *rp_out);

/* kernel.cu:880-909 (RP keys) and nodes.cu:128-135 (axis keys, j < min(6,d)).
 * keys[t*n + i], t < nrp + naxis. */
void     or_keys(const double *data, uint32_t n, uint32_t d, double eps,
                 const double *rp, uint32_t nrp, uint32_t naxis, uint32_t *keys);

/* Final squared distance of one pair under a sequence (no early exit). */
double   or_seq_sum(const double *p, const double *q, uint32_t d, uint32_t seq);

/* Full pipeline: buildNodeNet (nodes.cu:5-356) + nodeByPoint5 join
 * (kernel.cu:992-1046) + neighbour table (launcher.cu:4-50,3418-3506). */
int      or_self_join(const double *data, uint32_t n, uint32_t d, double eps,
                      const or_params *p, or_result *out);

/* O(n^2) FP64 brute force with the same per-pair decision. */
int      or_brute_force(const double *data, uint32_t n, uint32_t d, double eps,
                        uint32_t seq, int32_t nthreads, or_result *out);

void     or_result_free(or_result *r);

#ifdef __cplusplus
}
#endif
#endif
