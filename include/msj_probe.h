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
    uint32_t *keys;        /* in: NULL or room for (rp_per_layer + min(numidim, d)) * n
                              entries: key of every row under every candidate at
                              `layer` (kernel.cu:880-909, nodes.cu:128-135),
                              candidate-major, rows in the caller's order          */
    uint32_t *node_keys;   /* in: NULL or room for n * maxrp entries: the final
                              nodes' key tuples (binNumbers, nodes.cuh:9-38), node
                              order, num_layers keys each; 0xFFFFFFFF from the layer
                              a node stopped splitting (nodes.cu:476-484)         */
    uint32_t *node_start;  /* in: NULL or room for n + 1 entries: node boundaries     */
    uint32_t *node_order;  /* in: NULL or room for n entries: pointArray              */
} msj_index_probe;

/* Builds the node net of `data` (as msj_self_join takes it) on the GPU and
 * fills the requested probe buffers; runs no join. */
int msj_probe_index(const void *data, uint32_t n, uint32_t d, double eps,
                    const msj_params *p, msj_index_probe *probe);

/* The device arithmetic of a reference-point key on m squared distances:
 * q[i] = sqrt(D[i]) / eps in FP64 and key[i] = the key the index build
 * derives from D[i] (floor, then the saturating conversion of
 * kernel.cu:905).  Either output may be NULL. */
int msj_probe_key_math(const double *D, uint64_t m, double eps, double *q, uint32_t *key,
                       int device);

#ifdef __cplusplus
}
#endif
#endif
