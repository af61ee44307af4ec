/*
 * msj_probe.h -- read-only probes of the index build, for parity tests.
 *
 * No counterpart in the reference's interface: there the intermediate state
 * of buildNodeNet lives in host vectors (allBinNumber, nodes.cu:75-135;
 * std::vector<Node>, nodes.cuh:9-38) that only its printf lines and the
 * commented-out verifier nodeForce (nodes.cu:749-885) look at.  These entry
 * points copy the same state out of the GPU build so that tests can compare
 * it bit for bit with a CPU restatement.  Same conventions as msj.h: plain C,
 * status codes, never aborts.
 */
#ifndef MSJ_PROBE_H
#define MSJ_PROBE_H

#include <stdint.h>

#include "msj.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct msj_index_probe {
    uint32_t  layer;       /* in:  layer whose candidate keys go to `keys`            */
    uint32_t  ncand;       /* out: candidate partitions per layer (RPs + axes)        */
    uint32_t  num_nodes;   /* out: final node count                                   */
    uint32_t  num_layers;  /* out: layers built                                       */
    uint32_t *keys;        /* in: NULL or room for (rp_per_layer + axes) * n entries,
                              axes = min(numidim, d) for the node net, min(maxrp, d)
                              for the tree/COSS index: key of every row under every
                              candidate at `layer` (kernel.cu:880-909; axis keys
                              nodes.cu:128-135 or tree.cu:284), candidate-major,
                              rows in the caller's order                           */
    uint32_t *node_keys;   /* in: NULL or room for n * maxrp entries: the final
                              nodes' key tuples (binNumbers, nodes.cuh:9-38, or the
                              tree's bin numbers key - skip, tree.cu:247), node
                              order, num_layers keys each; 0xFFFFFFFF from the layer
                              a node stopped splitting (nodes.cu:476-484)         */
    uint32_t *node_start;  /* in: NULL or room for n + 1 entries: node boundaries     */
    uint32_t *node_order;  /* in: NULL or room for n entries: pointArray              */
    uint32_t *adj_off;     /* in: NULL or room for n + 1 entries: neighbour-node CSR
                              offsets (self included)                              */
    uint32_t *adj;         /* in: NULL or room for adj_cap entries: neighbour node
                              ids of every node, in the order the join reads them
                              (ascending for the tree/COSS index)                 */
    uint64_t  adj_cap;     /* in: entries `adj` holds (more are not copied)           */
    uint64_t  num_edges;   /* out: adjacency entries of the index                    */
    uint64_t  num_calcs;   /* out: sum over nodes of |X| * sum of neighbour sizes    */
    uint32_t  chosen[MSJ_MAX_LAYERS];          /* out: selected candidate per layer  */
    uint32_t  nodes_per_layer[MSJ_MAX_LAYERS]; /* out: nodes (non-empty bins) per layer */
    double    sse[MSJ_MAX_LAYERS];             /* out: selection statistic of the winner */
} msj_index_probe;

/* Builds the candidate index of `data` (as msj_self_join takes it; the node
 * net, or the tree/COSS bins when p->index is MSJ_INDEX_COSS) on the GPU and
 * fills the requested probe buffers; runs no join. */
int msj_probe_index(const void *data, uint32_t n, uint32_t d, double eps,
                    const msj_params *p, msj_index_probe *probe);

/* The device arithmetic of a reference-point key on m squared distances:
 * q[i] = sqrt(D[i]) / eps in FP64 and key[i] = the key the index build
 * derives from D[i] (floor, then the saturating conversion of
 * kernel.cu:905).  Either output may be NULL. */
int msj_probe_key_math(const double *D, uint64_t m, double eps, double *q, uint32_t *key,
                       int device);

/* The tree/COSS statistic's run of empty bins: s after g steps of
 * s = RN(s + c) (FP64, s >= 0, c >= 0 normal), computed in a few steps per
 * binade as the index build's sequential sum does (tree.cu:272-275 adds
 * (0 - avg)^2 once per empty bin).  Host code, no device needed. */
double msj_probe_repeat_add(double s, double c, uint64_t g);

#ifdef __cplusplus
}
#endif
#endif
