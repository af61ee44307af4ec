// nodelauncher5.cpp -- the reference's entry point nodeLauncher5 and its
// result type struct neighborTable (reference src/launcher.cu:3093-3532,
// include/launcher.cuh:77-82, include/utils.cuh:33-41) on top of the C ABI of
// libmsj.so.  Host C++ only; the join runs on the GPU inside libmsj.so.
//
// The reference fills the tables in its batch loop: entry 0 of every table
// is the placeholder {p, p, pointArray} (launcher.cu:3382-3395), then
// constructNeighborTable (launcher.cu:4-50) appends, per batch, the range of
// p's neighbours (node-order positions) in that batch's sorted pair array.
// Here there is one batch: entry 1 is p's whole neighbour list.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <new>
#include <vector>

#include "msj.h"
#include "msj_nodelauncher.h"

namespace {

// tables -> the neighbour array their entries point into
std::mutex g_mu;
std::map<const neighborTable *, unsigned int *> g_ids;

neighborTable *launch(double *data, unsigned int dim, unsigned int numPoints,
                      unsigned int *pointArray, double epsilon, int *rc)
{
    msj_params p;
    msj_default_params(&p);  // MAXRP, RPPERLAYER, NUMIDIM, SAMPLE_PER, ... (params.cuh)
    msj_result r;
    *rc = msj_self_join(data, numPoints, dim, epsilon, &p, &r);
    if (*rc != MSJ_OK)
        return nullptr;
    if (r.total_pairs > 0xFFFFFFFFull) {
        // the table's ranges are unsigned int (utils.cuh:36-37)
        msj_result_free(&r);
        *rc = MSJ_E_OVERFLOW;
        return nullptr;
    }
    // pointArray <- node order, as buildNodeNet leaves it (nodes.cu:336-345)
    std::memcpy(pointArray, r.node_order, sizeof(unsigned int) * numPoints);
    std::vector<unsigned int> pos(numPoints);  // row -> node-order position
    for (unsigned int i = 0; i < numPoints; ++i)
        pos[pointArray[i]] = i;
    // neighbour ids as node-order positions, grouped by query position
    // (launcher.cu:3468-3471: the pairs sorted by their first point)
    unsigned int *ids = (unsigned int *)std::malloc(sizeof(unsigned int) *
                                                    (r.total_pairs ? r.total_pairs : 1));
    neighborTable *tables =
        (neighborTable *)std::malloc(sizeof(neighborTable) * (numPoints ? numPoints : 1));
    if (!ids || !tables) {
        std::free(ids);
        std::free(tables);
        msj_result_free(&r);
        *rc = MSJ_E_NOMEM;
        return nullptr;
    }
    uint64_t w = 0;
    for (unsigned int i = 0; i < numPoints; ++i) {
        const unsigned int row = pointArray[i];
        const uint64_t b = r.offsets[row], e = r.offsets[row + 1];
        const uint64_t first = w;
        for (uint64_t k = b; k < e; ++k)
            ids[w++] = pos[r.nbrs[k]];
        std::sort(ids + first, ids + w);
        neighborTable *t = new (&tables[i]) neighborTable();
        t->cntNDataArrays = 1;
        t->vectindexmin.assign(1, i);
        t->vectindexmax.assign(1, i);
        t->vectdataPtr.assign(1, pointArray);
        if (e > b) {
            t->cntNDataArrays = 2;
            t->vectindexmin.push_back((unsigned int)first);
            t->vectindexmax.push_back((unsigned int)(w - 1));
            t->vectdataPtr.push_back(ids);
        }
        omp_init_lock(&t->pointLock);
    }
    std::printf("Total results Set Size: %llu \n", (unsigned long long)r.total_pairs);
    msj_result_free(&r);
    std::lock_guard<std::mutex> lk(g_mu);
    g_ids[tables] = ids;
    return tables;
}

}  // namespace

struct neighborTable *nodeLauncher5(double *data, unsigned int dim, unsigned int numPoints,
                                    unsigned int numRP, unsigned int *pointArray, double epsilon)
{
    (void)numRP;  // unused on this path in the reference too (main.cu:175)
    int rc = MSJ_OK;
    neighborTable *t = launch(data, dim, numPoints, pointArray, epsilon, &rc);
    if (!t) {
        std::fprintf(stderr, "nodeLauncher5: %s: %s\n", msj_strerror(rc),
                     rc == MSJ_E_OVERFLOW ? "more than 2^32 result pairs" : msj_last_error());
        std::abort();
    }
    return t;
}

extern "C" {

struct neighborTable *msj_nodelauncher5(double *data, unsigned int dim, unsigned int numPoints,
                                        unsigned int *pointArray, double epsilon, int *rc)
{
    int dummy = 0;
    return launch(data, dim, numPoints, pointArray, epsilon, rc ? rc : &dummy);
}

uint64_t msj_neighbor_table_flatten(const struct neighborTable *tables, unsigned int numPoints,
                                    uint32_t *counts, uint32_t *ids)
{
    uint64_t total = 0;
    for (unsigned int p = 0; p < numPoints; ++p) {
        const neighborTable &t = tables[p];
        uint64_t c = 0;
        for (unsigned int k = 1; k < t.cntNDataArrays; ++k) {
            const unsigned int lo = t.vectindexmin[k], hi = t.vectindexmax[k];
            for (unsigned int j = lo; j <= hi; ++j) {
                if (ids)
                    ids[total + c] = t.vectdataPtr[k][j];
                ++c;
            }
        }
        if (counts)
            counts[p] = (uint32_t)c;
        total += c;
    }
    return total;
}

unsigned int msj_neighbor_table_entries(const struct neighborTable *tables, unsigned int p)
{
    return tables[p].cntNDataArrays;
}

const unsigned int *msj_neighbor_table_entry(const struct neighborTable *tables, unsigned int p,
                                             unsigned int k, unsigned int *lo, unsigned int *hi)
{
    const neighborTable &t = tables[p];
    if (k >= t.cntNDataArrays)
        return nullptr;
    *lo = t.vectindexmin[k];
    *hi = t.vectindexmax[k];
    return t.vectdataPtr[k];
}

void msj_neighbor_table_free(struct neighborTable *tables, unsigned int numPoints)
{
    if (!tables)
        return;
    unsigned int *ids = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        auto it = g_ids.find(tables);
        if (it != g_ids.end()) {
            ids = it->second;
            g_ids.erase(it);
        }
    }
    for (unsigned int p = 0; p < numPoints; ++p) {
        omp_destroy_lock(&tables[p].pointLock);
        tables[p].~neighborTable();
    }
    std::free(tables);
    std::free(ids);
}

}  // extern "C"
