/*
 * msj_nodelauncher.h -- the reference's own join entry point, served by
 * libmsj.so (built as libmsj_nodelauncher.so, C++ with OpenMP).
 *
 *   struct neighborTable *nodeLauncher5(double *data, unsigned int dim,
 *       unsigned int numPoints, unsigned int numRP, unsigned int *pointArray,
 *       double epsilon);
 *
 * replaces reference include/launcher.cuh:77-82 / src/launcher.cu:3093-3532
 * with the same signature, and fills the same result type, struct
 * neighborTable (reference include/utils.cuh:33-41), the way the reference's
 * batch loop does (launcher.cu:3382-3395, constructNeighborTable 4-50):
 *   - tables[p] is indexed by node-order position p (pointArray[p] is the
 *     caller's row id of position p after the call, nodes.cu:336-345);
 *   - entry 0 is the reference's placeholder {min = p, max = p, data =
 *     pointArray};
 *   - entry 1 holds p's neighbours as node-order positions (self included):
 *     vectdataPtr[1][vectindexmin[1] .. vectindexmax[1]], ascending.
 * The reference pages results in batches and appends one entry per batch in
 * which p had results; one batch here leaves one entry.  Its neighbour order
 * within an entry is that of an unstable device sort; here it is ascending.
 *
 * A translation unit that already has the reference's struct neighborTable
 * (its utils.cuh) defines MSJ_HAVE_NEIGHBOR_TABLE before including this
 * header; the layout below is the same.
 */
#ifndef MSJ_NODELAUNCHER_H
#define MSJ_NODELAUNCHER_H

#include <stdint.h>

#ifdef __cplusplus
#include <omp.h>

#include <vector>

#ifndef MSJ_HAVE_NEIGHBOR_TABLE
#define MSJ_HAVE_NEIGHBOR_TABLE
typedef struct neighborTable {
    unsigned int cntNDataArrays;
    std::vector<unsigned int> vectindexmin;
    std::vector<unsigned int> vectindexmax;
    std::vector<unsigned int *> vectdataPtr;
    omp_lock_t pointLock;
} neighborTable;
#endif

/* Aborts with a message on any library error, as the reference's assert()s
 * do (launcher.cu:3260 and others); use msj_self_join for status codes. */
struct neighborTable *nodeLauncher5(double *data, unsigned int dim, unsigned int numPoints,
                                    unsigned int numRP, unsigned int *pointArray,
                                    double epsilon);

extern "C" {
#else
struct neighborTable;
#endif

/* The same call for C callers and tests: returns NULL (and a status in *rc)
 * instead of aborting. */
struct neighborTable *msj_nodelauncher5(double *data, unsigned int dim, unsigned int numPoints,
                                        unsigned int *pointArray, double epsilon, int *rc);

/* Walks tables[0 .. numPoints): counts[p] = neighbours of position p over the
 * entries k >= 1; if ids is not NULL, those neighbours (node-order positions)
 * are appended to it in entry order.  Returns the total. */
uint64_t msj_neighbor_table_flatten(const struct neighborTable *tables, unsigned int numPoints,
                                    uint32_t *counts, uint32_t *ids);

/* Entry access for callers without the C++ type: cntNDataArrays of table p,
 * and entry k's inclusive range [*lo, *hi] into the returned array (NULL
 * when k is out of range). */
unsigned int msj_neighbor_table_entries(const struct neighborTable *tables, unsigned int p);
const unsigned int *msj_neighbor_table_entry(const struct neighborTable *tables, unsigned int p,
                                             unsigned int k, unsigned int *lo, unsigned int *hi);

/* Releases tables from nodeLauncher5 / msj_nodelauncher5 and the neighbour
 * array they point into (the reference never frees its tables). */
void msj_neighbor_table_free(struct neighborTable *tables, unsigned int numPoints);

#ifdef __cplusplus
}
#endif
#endif
