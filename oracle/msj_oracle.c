/*
 * msj_oracle.c -- CPU restatement of MiSTIC's node-net epsilon self-join.
 *
 * TEST INFRASTRUCTURE ONLY: the parity checker and the CPU baseline.  The
 * product library never links or calls this file.  Compile with
 * -ffp-contract=off: every fused multiply-add below is an explicit fma().
 *
 * Layout of this restatement (reference file:line in bwd29/self-join-MiSTIC):
 *   glibc rand()                 utils.cu:77 (unseeded => srand(1) stream)
 *   or_dim_order                 utils.cu:22-53  (stddev + thrust radix order)
 *   rp_create                    utils.cu:55-137 (createRPArray)
 *   or_keys                      kernel.cu:880-909, nodes.cu:128-135
 *   build_net                    nodes.cu:5-356  (buildNodeNet)
 *     init_nodes                 nodes.cu:358-449
 *     split_nodes                nodes.cu:452-633
 *     update_neighbors           nodes.cu:719-747
 *     update_calcs               nodes.cu:677-695
 *   build_coss                   tree.cu:3-593, 596-759, 886-996 (buildTree,
 *                                generateRanges / binarySearch)
 *   join                         kernel.cu:992-1046 + 833-865 (nodeByPoint5,
 *                                cachedDistanceCheck), launcher.cu:3093-3532
 *
 * Deliberate deviations, each result-neutral or a defect fix:
 *   - axis partitions exist only for j < min(NUMIDIM, d); the reference reads
 *     data[k*dim+j] for j < 6 regardless of d (nodes.cu:128-135), i.e. other
 *     rows' coordinates (or past the end of the array) when d < 6;
 *   - the 38 candidate partitions are compared on their node sizes only; the
 *     full split + neighbour linking (which the reference also performs for
 *     the losers) is done for the selected candidate;
 *   - equal SSE: the lowest candidate index wins (the reference takes the
 *     first thread to enter its omp critical section, nodes.cu:204-220).
 */
#include "msj_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define XMALLOC(T, n) ((T *)xmalloc(sizeof(T) * (size_t)(n)))

static void *xmalloc(size_t sz)
{
    void *p = malloc(sz ? sz : 1);
    if (!p) {
        fprintf(stderr, "msj_oracle: out of memory (%zu bytes)\n", sz);
        abort();
    }
    return p;
}

void or_default_params(or_params *p)
{
    memset(p, 0, sizeof(*p));
    p->maxrp = 6;
    p->rp_per_layer = 32;
    p->numidim = 6;
    p->seed = 1;
    p->sample_per = 0.01;
    p->min_calcs_per_node = 100000;
    p->seq = OR_SEQ_NVVM7;
    p->join = 1;
    p->sample_every = 1;
    p->nthreads = 0;
    p->index = 0;
}

/* ------------------------------------------------------------------------ */
/* glibc random_r TYPE_3 (degree 31, separation 3), the generator behind    */
/* rand() in utils.cu:77.                                                    */
/* ------------------------------------------------------------------------ */
typedef struct {
    int32_t r[31];
    int f, b;
} or_rng;

static int32_t rng_next(or_rng *g)
{
    uint32_t val = (uint32_t)g->r[g->f] + (uint32_t)g->r[g->b];
    g->r[g->f] = (int32_t)val;
    int32_t res = (int32_t)(val >> 1);
    if (++g->f >= 31) {
        g->f = 0;
        ++g->b;
    } else if (++g->b >= 31) {
        g->b = 0;
    }
    return res;
}

static void rng_seed(or_rng *g, uint32_t seed)
{
    if (seed == 0)
        seed = 1;
    int32_t word = (int32_t)seed;
    g->r[0] = word;
    for (int i = 1; i < 31; i++) {
        long hi = word / 127773;
        long lo = word % 127773;
        long w = 16807 * lo - 2836 * hi;
        if (w < 0)
            w += 2147483647;
        word = (int32_t)w;
        g->r[i] = word;
    }
    g->f = 3;
    g->b = 0;
    for (int i = 0; i < 310; i++)
        (void)rng_next(g);
}

void or_glibc_rand(uint32_t seed, uint32_t count, int32_t *out)
{
    or_rng g;
    rng_seed(&g, seed);
    for (uint32_t i = 0; i < count; i++)
        out[i] = rng_next(&g);
}

/* ------------------------------------------------------------------------ */
/* Conversions as the reference's compilers emit them.                      */
/* ------------------------------------------------------------------------ */

/* Device double -> unsigned (cvt.rzi.u32.f64): truncate, saturate, NaN -> 0. */
static inline uint32_t cvt_u32_sat(double q)
{
    if (!(q > 0.0))
        return 0u;
    if (q >= 4294967295.0)
        return 0xFFFFFFFFu;
    return (uint32_t)q;
}

/* x86-64 host double -> unsigned: cvttsd2si to 64 bits, keep the low half;
 * out-of-range inputs give the "integer indefinite" 0x8000000000000000. */
static inline uint32_t cvt_u32_x86(double q)
{
    if (!(q > -9223372036854775808.0 && q < 9223372036854775808.0))
        return 0u;
    return (uint32_t)(uint64_t)(int64_t)q;
}

/* x86-64 host double -> int: cvttsd2si to 32 bits; out of range -> INT_MIN. */
static inline int32_t cvt_i32_x86(double q)
{
    if (!(q > -2147483649.0 && q < 2147483648.0))
        return (int32_t)0x80000000u;
    return (int32_t)q;
}

/* Total order of thrust's radix sort on doubles (sign-flip of the IEEE bits). */
static inline uint64_t radix_key_f64(double x)
{
    uint64_t u;
    memcpy(&u, &x, 8);
    return (u >> 63) ? ~u : (u | 0x8000000000000000ull);
}

/* ------------------------------------------------------------------------ */
/* utils.cu:22-53 stddev()                                                   */
/* ------------------------------------------------------------------------ */
static void stddev_order(const double *A, uint32_t dim, uint32_t num_points,
                         uint32_t *order_desc, double *sigma_out)
{
    double *dev = XMALLOC(double, dim);
    for (uint32_t i = 0; i < dim; i++) {
        double mean = 0.0;
        for (uint32_t j = 0; j < num_points; j++)
            mean += A[(size_t)dim * j + i];
        mean /= num_points;
        double devmean = 0.0;
        for (uint32_t j = 0; j < num_points; j++) {
            double t = A[(size_t)dim * j + i] - mean;
            devmean += t * t; /* pow(x,2) is folded to x*x by gcc */
        }
        devmean /= num_points;
        dev[i] = sqrt(devmean);
    }
    /* thrust::sort_by_key (stable, radix order) ascending, then reversed. */
    uint32_t *asc = XMALLOC(uint32_t, dim);
    for (uint32_t i = 0; i < dim; i++)
        asc[i] = i;
    for (uint32_t i = 1; i < dim; i++) { /* stable insertion sort */
        uint32_t v = asc[i];
        uint64_t kv = radix_key_f64(dev[v]);
        uint32_t j = i;
        while (j > 0 && radix_key_f64(dev[asc[j - 1]]) > kv) {
            asc[j] = asc[j - 1];
            j--;
        }
        asc[j] = v;
    }
    for (uint32_t i = 0; i < dim; i++) {
        order_desc[i] = asc[dim - 1 - i];
        if (sigma_out)
            sigma_out[i] = dev[asc[dim - 1 - i]];
    }
    free(asc);
    free(dev);
}

void or_dim_order(const double *A, uint32_t n, uint32_t d, uint32_t *order, double *sigma)
{
    stddev_order(A, d, n, order, sigma);
}

/* ------------------------------------------------------------------------ */
/* utils.cu:66-137 createRPArray(data, numRP, dim, numPoints)               */
/* ------------------------------------------------------------------------ */
static void rp_create(const double *data, uint32_t numRP, uint32_t dim, uint64_t numPoints,
                      double sample_per, or_rng *g, double *RPArray)
{
    uint32_t sample_size = (uint32_t)cvt_u32_x86((double)numPoints * sample_per);
    uint32_t test_rp = numRP * 2;
    double *testRP = XMALLOC(double, (size_t)test_rp * dim);
    for (size_t i = 0; i < (size_t)test_rp * dim; i++)
        testRP[i] = (double)rng_next(g) / (double)2147483647;

    /* distances of the first sample_size rows to every candidate (utils.cu:55-64) */
    double *distmat = XMALLOC(double, (size_t)test_rp * sample_size);
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)sample_size; i++) {
        for (uint32_t j = 0; j < test_rp; j++) {
            const double *pt = &data[(size_t)i * dim];
            const double *rp = &testRP[(size_t)j * dim];
            double dist = 0;
            for (uint32_t k = 0; k < dim; k++) {
                double diff = rp[k] - pt[k];
                dist += diff * diff;
            }
            distmat[(size_t)i * test_rp + j] = sqrt(dist);
        }
    }
    uint32_t *order = XMALLOC(uint32_t, test_rp);
    stddev_order(distmat, test_rp, sample_size, order, NULL);

    for (uint32_t i = 0; i < numRP / 2; i++)
        for (uint32_t j = 0; j < dim; j++)
            RPArray[(size_t)i * dim + j] = testRP[(size_t)order[i] * dim + j];

    /* box reference points: the max corner over the sample, then corners */
    for (uint32_t i = 0; i < dim; i++)
        testRP[i] = 0;
    for (uint32_t i = 0; i < sample_size; i++)
        for (uint32_t j = 0; j < dim; j++)
            if (testRP[j] < data[(size_t)i * dim + j])
                testRP[j] = data[(size_t)i * dim + j];
    for (uint32_t i = 1; i < test_rp; i++) {
        /* unsigned step = pow(2,dim)/test_rp: 2^(d-6) for 6<=d<=37, else 0 */
        uint32_t step = cvt_u32_x86(pow(2, dim) / test_rp);
        for (uint32_t j = 0; j < dim; j++) {
            uint32_t div = (uint32_t)cvt_i32_x86(pow(2, j)); /* (int)pow(2,j) */
            uint32_t bit = ((i * step) / div) % 2u;
            testRP[(size_t)i * dim + j] = (double)bit * testRP[j];
        }
    }
    for (uint32_t i = 0; i < numRP / 2 + numRP % 2; i++)
        for (uint32_t j = 0; j < dim; j++)
            RPArray[(size_t)i * dim + j + (size_t)(numRP / 2) * dim] = testRP[(size_t)i * dim + j];

    free(order);
    free(distmat);
    free(testRP);
}

void or_rp_layers(const double *data, uint32_t n, uint32_t d, const or_params *p,
                  uint32_t layers, double *rp_out)
{
    or_rng g;
    rng_seed(&g, p->seed);
    for (uint32_t l = 0; l < layers; l++)
        rp_create(data, p->rp_per_layer, d, n, p->sample_per, &g,
                  rp_out + (size_t)l * p->rp_per_layer * d);
}

/* ------------------------------------------------------------------------ */
/* Keys: kernel.cu:895-905 (fma accumulation under nvcc contraction, IEEE   */
/* sqrt and division, saturating conversion) and nodes.cu:130 (host ceil).  */
/* ------------------------------------------------------------------------ */
void or_keys(const double *data, uint32_t n, uint32_t d, double eps, const double *rp,
             uint32_t nrp, uint32_t naxis, uint32_t *keys)
{
    #pragma omp parallel for schedule(static)
    for (int64_t ii = 0; ii < (int64_t)n; ii++) {
        size_t i = (size_t)ii;
        const double *x = &data[i * d];
        for (uint32_t t = 0; t < nrp; t++) {
            const double *r = &rp[(size_t)t * d];
            double dist = 0;
            for (uint32_t j = 0; j < d; j++) {
                double df = x[j] - r[j];
                dist = fma(df, df, dist);
            }
            keys[(size_t)t * n + i] = cvt_u32_sat(floor(sqrt(dist) / eps));
        }
        for (uint32_t a = 0; a < naxis; a++)
            keys[(size_t)(nrp + a) * n + i] = cvt_u32_x86(ceil((x[a] + 0.0000001) / eps));
    }
}

/* ------------------------------------------------------------------------ */
/* Distance decision, kernel.cu:833-865 with ILP = 4.                        */
/* Returns the sum at the point of decision; *accept = !(s > eps2).          */
/* ------------------------------------------------------------------------ */
static inline double seq_dist(const double *p, const double *q, uint32_t d, double eps2,
                              uint32_t seq, int early, int *accept)
{
    double s = 0.0;
    for (uint32_t l = 0; l < d; l += 4) {
        double t = p[l] - q[l];
        if (seq == OR_SEQ_NOFMA)
            s = s + t * t;
        else
            s = fma(t, t, s);
        for (uint32_t j = 1; j < 4 && l + j < d; j++) {
            double u = p[l + j] - q[l + j];
            if (seq == OR_SEQ_ALLFMA)
                s = fma(u, u, s);
            else
                s = s + u * u;
        }
        if (early && s > eps2) {
            *accept = 0;
            return s;
        }
    }
    *accept = !(s > eps2);
    return s;
}

double or_seq_sum(const double *p, const double *q, uint32_t d, uint32_t seq)
{
    int acc;
    return seq_dist(p, q, d, 0.0, seq, 0, &acc);
}

/* ------------------------------------------------------------------------ */
/* Stable LSD radix sort of u32 keys (optionally carrying u32 values),      */
/* the ordering of thrust::sort_by_key (nodes.cu:372,505).                  */
/* ------------------------------------------------------------------------ */
static void radix_sort_u32(uint32_t *k, uint32_t *v, size_t n, uint32_t *tk, uint32_t *tv)
{
    if (n < 2)
        return;
    uint32_t vary = 0; /* bits that differ between any key and the first */
    for (size_t i = 1; i < n; i++)
        vary |= k[i] ^ k[0];
    uint32_t *sk = k, *sv = v, *dk = tk, *dv = tv;
    for (int shift = 0; shift < 32; shift += 8) {
        if (((vary >> shift) & 0xFFu) == 0) /* digit constant: identity pass */
            continue;
        size_t cnt[257];
        memset(cnt, 0, sizeof(cnt));
        for (size_t i = 0; i < n; i++)
            cnt[((sk[i] >> shift) & 0xFFu) + 1]++;
        for (int b = 0; b < 256; b++)
            cnt[b + 1] += cnt[b];
        for (size_t i = 0; i < n; i++) {
            size_t pos = cnt[(sk[i] >> shift) & 0xFFu]++;
            dk[pos] = sk[i];
            if (v)
                dv[pos] = sv[i];
        }
        uint32_t *x = sk;
        sk = dk;
        dk = x;
        if (v) {
            x = sv;
            sv = dv;
            dv = x;
        }
    }
    if (sk != k) {
        memcpy(k, sk, n * sizeof(uint32_t));
        if (v)
            memcpy(v, sv, n * sizeof(uint32_t));
    }
}

/* ------------------------------------------------------------------------ */
/* Node net data structure (nodes.cuh:9-38).                                 */
/* ------------------------------------------------------------------------ */
typedef struct {
    uint32_t npts;
    uint32_t *pts; /* nodePoints: original row ids */
    uint32_t nkeys;
    uint32_t keys[OR_MAXL]; /* binNumbers */
    uint32_t *nbr;          /* neighborIndex, self first */
    uint32_t nnbr, capnbr;
    uint64_t numCalcs;
    int split;
    uint32_t index; /* nodeIndex */
} onode;

typedef struct {
    onode *v;
    uint32_t n, cap;
} ovec;

static void ovec_push(ovec *a, onode nd)
{
    if (a->n == a->cap) {
        a->cap = a->cap ? a->cap * 2 : 16;
        a->v = (onode *)realloc(a->v, sizeof(onode) * a->cap);
        if (!a->v)
            abort();
    }
    a->v[a->n++] = nd;
}

static void node_push_nbr(onode *nd, uint32_t x)
{
    if (nd->nnbr == nd->capnbr) {
        nd->capnbr = nd->capnbr ? nd->capnbr * 2 : 4;
        nd->nbr = (uint32_t *)realloc(nd->nbr, sizeof(uint32_t) * nd->capnbr);
        if (!nd->nbr)
            abort();
    }
    nd->nbr[nd->nnbr++] = x;
}

static void ovec_free(ovec *a)
{
    for (uint32_t i = 0; i < a->n; i++) {
        free(a->v[i].pts);
        free(a->v[i].nbr);
    }
    free(a->v);
    a->v = NULL;
    a->n = a->cap = 0;
}

/* newNode (nodes.cu:636-675): copies parent's keys, appends key, self link. */
static onode new_node(uint32_t npts, const uint32_t *pts, const onode *parent, uint32_t key,
                      uint32_t index)
{
    onode nd;
    memset(&nd, 0, sizeof(nd));
    nd.index = index;
    nd.npts = npts;
    if (parent) {
        nd.nkeys = parent->nkeys;
        memcpy(nd.keys, parent->keys, sizeof(uint32_t) * parent->nkeys);
    }
    nd.keys[nd.nkeys++] = key;
    node_push_nbr(&nd, index);
    nd.pts = XMALLOC(uint32_t, npts);
    memcpy(nd.pts, pts, sizeof(uint32_t) * npts);
    nd.split = 1;
    return nd;
}

/* Sibling links: consecutive distinct keys that differ by exactly one
 * (nodes.cu:414-435 and 557-578). */
static void link_siblings(onode *c, uint32_t m)
{
    if (m < 2)
        return;
    if (c[0].keys[c[0].nkeys - 1] == c[1].keys[c[1].nkeys - 1] - 1u)
        node_push_nbr(&c[0], c[1].index);
    for (uint32_t j = 1; j + 1 < m; j++) {
        uint32_t kj = c[j].keys[c[j].nkeys - 1];
        if (kj == c[j - 1].keys[c[j - 1].nkeys - 1] + 1u)
            node_push_nbr(&c[j], c[j - 1].index);
        if (kj == c[j + 1].keys[c[j + 1].nkeys - 1] - 1u)
            node_push_nbr(&c[j], c[j + 1].index);
    }
    if (c[m - 1].keys[c[m - 1].nkeys - 1] == c[m - 2].keys[c[m - 2].nkeys - 1] + 1u)
        node_push_nbr(&c[m - 1], c[m - 2].index);
}

/* updateNodeCalcs (nodes.cu:677-695) */
static void update_calcs(ovec *nodes)
{
    #pragma omp parallel for schedule(dynamic, 64)
    for (int64_t i = 0; i < (int64_t)nodes->n; i++) {
        onode *nd = &nodes->v[i];
        uint64_t s = 0;
        for (uint32_t j = 0; j < nd->nnbr; j++)
            s += (uint64_t)nodes->v[nd->nbr[j]].npts;
        nd->numCalcs = s * (uint64_t)nd->npts;
    }
}

static uint64_t total_calcs(const ovec *nodes)
{
    uint64_t t = 0;
    for (uint32_t i = 0; i < nodes->n; i++)
        t += nodes->v[i].numCalcs;
    return t;
}

/* initNodes (nodes.cu:358-449) */
static void init_nodes(uint32_t numPoints, const uint32_t *binNumberIn, ovec *out)
{
    uint32_t *key = XMALLOC(uint32_t, numPoints);
    uint32_t *pa = XMALLOC(uint32_t, numPoints);
    uint32_t *tk = XMALLOC(uint32_t, numPoints);
    uint32_t *tv = XMALLOC(uint32_t, numPoints);
    memcpy(key, binNumberIn, sizeof(uint32_t) * numPoints);
    for (uint32_t i = 0; i < numPoints; i++)
        pa[i] = i;
    radix_sort_u32(key, pa, numPoints, tk, tv);
    free(tk);
    free(tv);
    if (key[0] == key[numPoints - 1]) {
        ovec_push(out, new_node(numPoints, pa, NULL, key[0], 0));
    } else {
        uint32_t start = 0, m = 0;
        for (uint32_t i = 0; i < numPoints; i++) {
            if (i == numPoints - 1 || key[i] != key[i + 1]) {
                ovec_push(out, new_node(i - start + 1, pa + start, NULL, key[i], m));
                start = i + 1;
                m++;
            }
        }
        link_siblings(out->v, m);
    }
    update_calcs(out);
    free(key);
    free(pa);
}

/* splitNodes (nodes.cu:452-633) + updateNeighbors (nodes.cu:719-747) */
static void split_nodes(const uint32_t *allBin, const ovec *nodes, uint64_t min_calcs, ovec *out)
{
    uint32_t nn = nodes->n;
    ovec *tmp = (ovec *)calloc(nn, sizeof(ovec));
    #pragma omp parallel
    {
        uint32_t cap = 0;
        uint32_t *bn = NULL, *pts = NULL, *tk = NULL, *tv = NULL;
        #pragma omp for schedule(dynamic, 16)
        for (int64_t ii = 0; ii < (int64_t)nn; ii++) {
            const onode *P = &nodes->v[ii];
            if (P->numCalcs < min_calcs) {
                onode c = new_node(P->npts, P->pts, P, 0xFFFFFFFFu, 0);
                c.split = 0;
                ovec_push(&tmp[ii], c);
                continue;
            }
            if (P->npts > cap) {
                cap = P->npts;
                bn = (uint32_t *)realloc(bn, sizeof(uint32_t) * cap);
                pts = (uint32_t *)realloc(pts, sizeof(uint32_t) * cap);
                tk = (uint32_t *)realloc(tk, sizeof(uint32_t) * cap);
                tv = (uint32_t *)realloc(tv, sizeof(uint32_t) * cap);
            }
            for (uint32_t j = 0; j < P->npts; j++) {
                bn[j] = allBin[P->pts[j]];
                pts[j] = P->pts[j];
            }
            radix_sort_u32(bn, pts, P->npts, tk, tv);
            if (bn[0] == bn[P->npts - 1]) {
                ovec_push(&tmp[ii], new_node(P->npts, pts, P, bn[0], 0));
                continue;
            }
            uint32_t start = 0, m = 0;
            for (uint32_t j = 0; j < P->npts; j++) {
                if (j == P->npts - 1 || bn[j] != bn[j + 1]) {
                    ovec_push(&tmp[ii], new_node(j - start + 1, pts + start, P, bn[j], m));
                    start = j + 1;
                    m++;
                }
            }
            link_siblings(tmp[ii].v, m);
        }
        free(bn);
        free(pts);
        free(tk);
        free(tv);
    }
    /* linear node ids (nodes.cu:595-604) */
    uint32_t off = 0;
    for (uint32_t i = 0; i < nn; i++) {
        for (uint32_t j = 0; j < tmp[i].n; j++) {
            tmp[i].v[j].index += off;
            for (uint32_t k = 0; k < tmp[i].v[j].nnbr; k++)
                tmp[i].v[j].nbr[k] += off;
        }
        off += tmp[i].n;
    }
    /* updateNeighbors */
    #pragma omp parallel for schedule(dynamic, 16)
    for (int64_t ii = 0; ii < (int64_t)nn; ii++) {
        const onode *P = &nodes->v[ii];
        for (uint32_t j = 0; j < tmp[ii].n; j++) {
            onode *c = &tmp[ii].v[j];
            uint32_t nb = c->keys[c->nkeys - 1];
            for (uint32_t k = 1; k < P->nnbr; k++) {
                uint32_t q = P->nbr[k];
                for (uint32_t l = 0; l < tmp[q].n; l++) {
                    const onode *o = &tmp[q].v[l];
                    if (!o->split || !c->split) {
                        node_push_nbr(c, o->index);
                    } else {
                        uint32_t cb = o->keys[o->nkeys - 1];
                        if (cb + 1u == nb || cb - 1u == nb || cb == nb)
                            node_push_nbr(c, o->index);
                    }
                }
            }
        }
    }
    for (uint32_t i = 0; i < nn; i++) {
        for (uint32_t j = 0; j < tmp[i].n; j++)
            ovec_push(out, tmp[i].v[j]);
        free(tmp[i].v);
    }
    free(tmp);
    update_calcs(out);
}

/* Node sizes a candidate partition would produce, in node order, and the
 * selection statistic sum (size - N/#nodes)^2 (nodes.cu:194-199). */
static double candidate_sse(const uint32_t *allBin, const ovec *nodes, uint32_t numPoints,
                            int layer0, uint64_t min_calcs, uint32_t *nout)
{
    uint32_t cap = numPoints + 1;
    uint32_t *sizes = XMALLOC(uint32_t, cap);
    uint32_t ns = 0;
    uint32_t *bn = XMALLOC(uint32_t, numPoints);
    uint32_t *tk = XMALLOC(uint32_t, numPoints);
    if (layer0) {
        memcpy(bn, allBin, sizeof(uint32_t) * numPoints);
        radix_sort_u32(bn, NULL, numPoints, tk, NULL);
        uint32_t start = 0;
        for (uint32_t i = 0; i < numPoints; i++)
            if (i == numPoints - 1 || bn[i] != bn[i + 1]) {
                sizes[ns++] = i - start + 1;
                start = i + 1;
            }
    } else {
        for (uint32_t i = 0; i < nodes->n; i++) {
            const onode *P = &nodes->v[i];
            if (P->numCalcs < min_calcs) {
                sizes[ns++] = P->npts;
                continue;
            }
            for (uint32_t j = 0; j < P->npts; j++)
                bn[j] = allBin[P->pts[j]];
            radix_sort_u32(bn, NULL, P->npts, tk, NULL);
            uint32_t start = 0;
            for (uint32_t j = 0; j < P->npts; j++)
                if (j == P->npts - 1 || bn[j] != bn[j + 1]) {
                    sizes[ns++] = j - start + 1;
                    start = j + 1;
                }
        }
    }
    double avg = numPoints * 1.0 / ns;
    double sse = 0;
    for (uint32_t k = 0; k < ns; k++)
        sse += (sizes[k] - avg) * (sizes[k] - avg);
    free(sizes);
    free(bn);
    free(tk);
    *nout = ns;
    return sse;
}

/* buildNodeNet (nodes.cu:5-356) */
static void build_net(const double *data, uint32_t n, uint32_t d, double eps, const or_params *p,
                      or_result *res, ovec *final_nodes)
{
    or_rng g;
    rng_seed(&g, p->seed);
    uint32_t nrp = p->rp_per_layer;
    uint32_t naxis = p->numidim < d ? p->numidim : d;
    uint32_t ncand = nrp + naxis;
    res->num_candidates = ncand;
    double *rp = XMALLOC(double, (size_t)nrp * d);
    uint32_t *keys = XMALLOC(uint32_t, (size_t)ncand * n);
    ovec cur = {0};
    uint32_t layers = p->maxrp < OR_MAXL ? p->maxrp : OR_MAXL;
    for (uint32_t l = 0; l < layers; l++) {
        rp_create(data, nrp, d, n, p->sample_per, &g, rp);
        or_keys(data, n, d, eps, rp, nrp, naxis, keys);
        double *sse = XMALLOC(double, ncand);
        uint32_t *cnt = XMALLOC(uint32_t, ncand);
        #pragma omp parallel for schedule(dynamic, 1)
        for (int64_t t = 0; t < (int64_t)ncand; t++)
            sse[t] = candidate_sse(keys + (size_t)t * n, &cur, n, l == 0, p->min_calcs_per_node,
                                   &cnt[t]);
        uint32_t best = 0;
        for (uint32_t t = 1; t < ncand; t++)
            if (sse[t] < sse[best])
                best = t;
        ovec next = {0};
        if (l == 0)
            init_nodes(n, keys + (size_t)best * n, &next);
        else
            split_nodes(keys + (size_t)best * n, &cur, p->min_calcs_per_node, &next);
        ovec_free(&cur);
        cur = next;
        res->chosen[l] = best;
        res->sse[l] = sse[best];
        res->nodes_per_layer[l] = cur.n;
        res->calcs_per_layer[l] = total_calcs(&cur);
        free(sse);
        free(cnt);
    }
    res->num_layers = layers;
    free(rp);
    free(keys);
    *final_nodes = cur;
}

/* ------------------------------------------------------------------------ */
/* Tree / COSS bins: buildTree (tree.cu:3-593, with MINSQRS = true,          */
/* params.cuh:29) and the 3^L neighbour probes of generateRanges /           */
/* binarySearch (tree.cu:596-759, 886-996).                                  */
/*                                                                          */
/* Readings of the reference kept here (documented deviations aside):       */
/*  - axis candidates are maxRP of them (tree.cu:278), guarded j < d;        */
/*  - an axis candidate's sum of squares is not reset between layers unless  */
/*    its average bin count triggers the N^2 penalty (tree.cu:336-339); its  */
/*    first-layer start value (uninitialised in the reference) is 0;         */
/*  - a candidate whose layer would exceed 2^32 bins, or whose key span     */
/*    wraps the unsigned bin count maxBin + 2 - (minBin - 2) (both unsigned  */
/*    overflow and out-of-bounds writes in the reference) is not eligible;   */
/*    a layer with no eligible candidate fails the build (returns -3).       */
/* ------------------------------------------------------------------------ */
typedef struct {
    uint32_t skip, nb, nonempty;
    int eligible;
} coss_cand;

/* Bin address of every point under candidate t in layer cl (tree.cu:224-253). */
static inline uint64_t coss_addr(uint32_t key, uint32_t skip, uint32_t nb, uint32_t prank, int cl)
{
    uint32_t bn = key - skip;
    if (cl == 0)
        return bn;
    return (uint64_t)(uint32_t)((prank - 1u) * nb) + bn;
}

static int cmp_tuple_len;
static int cmp_tuples(const uint32_t *a, const uint32_t *b, int L)
{
    for (int l = 0; l < L; l++) {
        if (a[l] < b[l])
            return -1;
        if (a[l] > b[l])
            return 1;
    }
    return 0;
}

static int build_coss(const double *data, uint32_t n, uint32_t d, double eps, const or_params *p,
                      or_result *res, ovec *final_nodes)
{
    or_rng g;
    rng_seed(&g, p->seed);
    const uint32_t nrp = p->rp_per_layer;
    const uint32_t L = p->maxrp < OR_MAXL ? p->maxrp : OR_MAXL;
    const uint32_t naxis = L < d ? L : d;
    const uint32_t ncand = nrp + naxis;
    res->num_candidates = ncand;
    double *rp = XMALLOC(double, (size_t)nrp * d);
    uint32_t *keys = XMALLOC(uint32_t, (size_t)ncand * n);
    uint32_t *prank = XMALLOC(uint32_t, n); /* 1-based rank of the point's bin in the last layer */
    uint32_t *bnum = XMALLOC(uint32_t, (size_t)n * L); /* pointBinNumbers */
    double *sumsq = (double *)calloc(ncand, sizeof(double));
    double *score = XMALLOC(double, ncand);
    coss_cand *cc = (coss_cand *)calloc(ncand, sizeof(coss_cand));
    uint32_t prev_nonempty = 0;
    for (uint32_t i = 0; i < n; i++)
        prank[i] = 1;
    for (uint32_t cl = 0; cl < L; cl++) {
        rp_create(data, nrp, d, n, p->sample_per, &g, rp);
        /* RP keys: binningKernel (kernel.cu:895-905) */
        #pragma omp parallel for schedule(static)
        for (int64_t ii = 0; ii < (int64_t)n; ii++) {
            const double *x = &data[(size_t)ii * d];
            for (uint32_t t = 0; t < nrp; t++) {
                const double *r = &rp[(size_t)t * d];
                double dist = 0;
                for (uint32_t j = 0; j < d; j++) {
                    double df = x[j] - r[j];
                    dist = fma(df, df, dist);
                }
                keys[(size_t)t * n + ii] = cvt_u32_sat(floor(sqrt(dist) / eps));
            }
            /* axis keys: ceil((x + 0.0000001) / eps) + 10 (tree.cu:284) */
            for (uint32_t a = 0; a < naxis; a++)
                keys[(size_t)(nrp + a) * n + ii] = cvt_u32_x86(ceil((x[a] + 0.0000001) / eps) + 10);
        }
        for (uint32_t t = 0; t < nrp; t++)
            sumsq[t] = 0;
        #pragma omp parallel for schedule(dynamic, 1)
        for (int64_t tt = 0; tt < (int64_t)ncand; tt++) {
            const uint32_t t = (uint32_t)tt;
            const uint32_t *k = keys + (size_t)t * n;
            uint32_t maxB, minB;
            if (t < nrp) { /* tree.cu:181-196 */
                maxB = minB = k[0];
                for (uint32_t i = 0; i < n; i++) {
                    if (k[i] > maxB)
                        maxB = k[i];
                    if (k[i] < minB)
                        minB = k[i];
                }
            } else { /* tree.cu:281-287: int bounds seeded from row 0, unsigned compares */
                const uint32_t a = t - nrp;
                int32_t maxBi = cvt_i32_x86(ceil((data[a] + 0.000001) / eps) + 10);
                int32_t minBi = (int32_t)((uint32_t)maxBi - 1u);
                for (uint32_t i = 0; i < n; i++) {
                    if (k[i] > (uint32_t)maxBi)
                        maxBi = (int32_t)k[i];
                    if (k[i] < (uint32_t)minBi)
                        minBi = (int32_t)k[i];
                }
                maxB = (uint32_t)maxBi;
                minB = (uint32_t)minBi;
            }
            coss_cand c;
            c.skip = minB - 2u;
            c.nb = maxB + 2u - c.skip;
            const uint64_t count64 = cl == 0 ? c.nb : (uint64_t)prev_nonempty * c.nb;
            const int wraps = (uint64_t)(maxB - minB) + 4u > 0xFFFFFFFFull;
            c.eligible = !wraps && count64 > 0 && count64 <= 0xFFFFFFFFull;
            c.nonempty = 0;
            if (!c.eligible) {
                cc[t] = c;
                score[t] = INFINITY;
                continue;
            }
            uint32_t *bins = (uint32_t *)calloc(count64, sizeof(uint32_t));
            if (!bins)
                abort();
            for (uint32_t i = 0; i < n; i++) {
                uint64_t addr = coss_addr(k[i], c.skip, c.nb, prank[i], cl);
                if (bins[addr] == 0)
                    c.nonempty++;
                bins[addr]++;
            }
            const double avg = (double)((uint64_t)n / c.nonempty);
            if (t >= nrp && avg >= n / 3.0)
                sumsq[t] = (double)((uint64_t)n * n);
            double s = sumsq[t];
            for (uint64_t j = 0; j < count64; j++) {
                double e = bins[j] - avg;
                s += e * e;
            }
            sumsq[t] = s;
            score[t] = s;
            free(bins);
            cc[t] = c;
        }
        uint32_t best = 0; /* MINSQRS: the first smallest (tree.cu:395-402) */
        for (uint32_t t = 1; t < ncand; t++)
            if (score[best] > score[t])
                best = t;
        if (!cc[best].eligible) {
            free(rp); free(keys); free(prank); free(bnum); free(sumsq); free(score); free(cc);
            return -3;
        }
        /* apply: ranks of the non-empty bins (tree.cu:416-430), bin numbers */
        const coss_cand c = cc[best];
        const uint64_t count64 = cl == 0 ? c.nb : (uint64_t)prev_nonempty * c.nb;
        uint32_t *rankof = (uint32_t *)calloc(count64, sizeof(uint32_t));
        const uint32_t *k = keys + (size_t)best * n;
        uint64_t *addr = XMALLOC(uint64_t, n);
        for (uint32_t i = 0; i < n; i++) {
            addr[i] = coss_addr(k[i], c.skip, c.nb, prank[i], cl);
            rankof[addr[i]] = 1;
        }
        uint32_t cnt = 0;
        for (uint64_t j = 0; j < count64; j++)
            if (rankof[j])
                rankof[j] = ++cnt;
        for (uint32_t i = 0; i < n; i++) {
            prank[i] = rankof[addr[i]];
            bnum[(size_t)i * L + cl] = k[i] - c.skip;
        }
        free(addr);
        free(rankof);
        prev_nonempty = c.nonempty;
        res->chosen[cl] = best;
        res->sse[cl] = score[best];
        res->nodes_per_layer[cl] = c.nonempty;
    }
    res->num_layers = L;
    /* points by final bin (lexicographic in the bin numbers), stable by row
     * (tree.cu:472-503) */
    uint32_t *order = XMALLOC(uint32_t, n), *tk = XMALLOC(uint32_t, n), *tv = XMALLOC(uint32_t, n);
    uint32_t *kk = XMALLOC(uint32_t, n);
    for (uint32_t i = 0; i < n; i++) {
        order[i] = i;
        kk[i] = prank[i];
    }
    radix_sort_u32(kk, order, n, tk, tv);
    free(tk);
    free(tv);
    /* nodes = non-empty final bins */
    ovec out = {0};
    uint32_t start = 0;
    for (uint32_t i = 0; i < n; i++) {
        if (i == n - 1 || kk[i] != kk[i + 1]) {
            onode nd;
            memset(&nd, 0, sizeof(nd));
            nd.index = out.n;
            nd.npts = i - start + 1;
            nd.pts = XMALLOC(uint32_t, nd.npts);
            memcpy(nd.pts, order + start, sizeof(uint32_t) * nd.npts);
            nd.nkeys = L;
            for (uint32_t l = 0; l < L; l++)
                nd.keys[l] = bnum[(size_t)order[start] * L + l];
            nd.split = 1;
            ovec_push(&out, nd);
            start = i + 1;
        }
    }
    free(kk);
    free(order);
    /* neighbours: the 3^L probes of binarySearch (tree.cu:960-975), each a
     * binary search over the sorted bin-number tuples (bSearch, 886-917) */
    const uint32_t M = out.n;
    uint32_t *tup = XMALLOC(uint32_t, (size_t)M * L);
    for (uint32_t v = 0; v < M; v++)
        memcpy(tup + (size_t)v * L, out.v[v].keys, sizeof(uint32_t) * L);
    uint64_t probes = 1;
    for (uint32_t l = 0; l < L; l++)
        probes *= 3;
    #pragma omp parallel for schedule(dynamic, 16)
    for (int64_t vv = 0; vv < (int64_t)M; vv++) {
        onode *nd = &out.v[vv];
        uint32_t want[OR_MAXL];
        uint64_t sum = 0;
        for (uint64_t i = 0; i < probes; i++) {
            uint64_t q = i;
            for (uint32_t l = 0; l < L; l++) {
                want[l] = nd->keys[l] + (uint32_t)(q % 3) - 1u;
                q /= 3;
            }
            int64_t lo = 0, hi = (int64_t)M - 1, hit = -1;
            while (lo <= hi) {
                int64_t mid = (lo + hi) / 2;
                int c = cmp_tuples(tup + (size_t)mid * L, want, (int)L);
                if (c == 0) {
                    hit = mid;
                    break;
                }
                if (c < 0)
                    lo = mid + 1;
                else
                    hi = mid - 1;
            }
            if (hit >= 0) {
                node_push_nbr(nd, (uint32_t)hit);
                sum += out.v[hit].npts;
            }
        }
        nd->numCalcs = sum * (uint64_t)nd->npts; /* tree.cu:756 */
    }
    (void)cmp_tuple_len;
    free(tup);
    res->calcs_per_layer[L - 1] = total_calcs(&out);
    free(rp);
    free(keys);
    free(prank);
    free(bnum);
    free(sumsq);
    free(score);
    free(cc);
    *final_nodes = out;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Results                                                                   */
/* ------------------------------------------------------------------------ */
typedef struct {
    uint32_t *q; /* neighbour original ids, grouped per query */
    uint32_t *cnt; /* per query of the node */
    uint64_t len, cap;
} nodebuf;

static void nb_push(nodebuf *b, uint32_t v)
{
    if (b->len == b->cap) {
        b->cap = b->cap ? b->cap * 2 : 256;
        b->q = (uint32_t *)realloc(b->q, sizeof(uint32_t) * b->cap);
        if (!b->q)
            abort();
    }
    b->q[b->len++] = v;
}

static int cmp_u32(const void *a, const void *b)
{
    uint32_t x = *(const uint32_t *)a, y = *(const uint32_t *)b;
    return (x > y) - (x < y);
}

/* Scatter per-node buffers into the per-original-row CSR, rows sorted. */
static void assemble(or_result *res, uint32_t n, uint32_t nnodes, const uint32_t *node_start,
                     const uint32_t *order, nodebuf *bufs)
{
    res->counts = XMALLOC(uint32_t, n);
    res->offsets = XMALLOC(uint64_t, (size_t)n + 1);
    memset(res->counts, 0, sizeof(uint32_t) * n);
    for (uint32_t x = 0; x < nnodes; x++) {
        if (!bufs[x].cnt)
            continue;
        for (uint32_t i = node_start[x]; i < node_start[x + 1]; i++)
            res->counts[order[i]] = bufs[x].cnt[i - node_start[x]];
    }
    uint64_t acc = 0;
    for (uint32_t i = 0; i < n; i++) {
        res->offsets[i] = acc;
        acc += res->counts[i];
    }
    res->offsets[n] = acc;
    res->total_pairs = acc;
    res->nbrs = XMALLOC(uint32_t, acc);
    #pragma omp parallel for schedule(dynamic, 16)
    for (int64_t xx = 0; xx < (int64_t)nnodes; xx++) {
        nodebuf *b = &bufs[xx];
        if (!b->cnt)
            continue;
        uint64_t pos = 0;
        for (uint32_t i = node_start[xx]; i < node_start[xx + 1]; i++) {
            uint32_t row = order[i];
            uint32_t c = b->cnt[i - node_start[xx]];
            uint32_t *dst = res->nbrs + res->offsets[row];
            memcpy(dst, b->q + pos, sizeof(uint32_t) * c);
            qsort(dst, c, sizeof(uint32_t), cmp_u32);
            pos += c;
        }
        free(b->q);
        free(b->cnt);
        b->q = NULL;
        b->cnt = NULL;
    }
}

static inline int near_boundary(double s, double eps2)
{
    return fabs(s - eps2) <= 1e-9 * eps2;
}

int or_self_join(const double *data, uint32_t n, uint32_t d, double eps, const or_params *pin,
                 or_result *out)
{
    or_params p;
    if (pin)
        p = *pin;
    else
        or_default_params(&p);
    memset(out, 0, sizeof(*out));
    out->n = n;
    out->d = d;
    if (d == 0 || !(eps > 0.0) || p.maxrp == 0)
        return -1;
    if (p.nthreads > 0)
        omp_set_num_threads(p.nthreads);
    if (n == 0) {
        out->offsets = XMALLOC(uint64_t, 1);
        out->offsets[0] = 0;
        return 0;
    }

    double t0 = omp_get_wtime();
    ovec nodes = {0};
    if (p.index == 2) {
        if (build_coss(data, n, d, eps, &p, out, &nodes) != 0)
            return -3;
    } else
        build_net(data, n, d, eps, &p, out, &nodes);
    uint32_t nn = nodes.n;
    out->num_nodes = nn;
    out->num_calcs = total_calcs(&nodes);
    out->node_order = XMALLOC(uint32_t, n);
    out->node_start = XMALLOC(uint32_t, (size_t)nn + 1);
    out->adj_off = XMALLOC(uint32_t, (size_t)nn + 1);
    uint32_t pos = 0, apos = 0;
    for (uint32_t x = 0; x < nn; x++) {
        out->node_start[x] = pos;
        memcpy(out->node_order + pos, nodes.v[x].pts, sizeof(uint32_t) * nodes.v[x].npts);
        pos += nodes.v[x].npts;
        out->adj_off[x] = apos;
        apos += nodes.v[x].nnbr;
    }
    out->node_start[nn] = pos;
    out->adj_off[nn] = apos;
    out->adj = XMALLOC(uint32_t, apos);
    out->node_keys = XMALLOC(uint32_t, (size_t)nn * out->num_layers);
    for (uint32_t x = 0; x < nn; x++) {
        memcpy(out->adj + out->adj_off[x], nodes.v[x].nbr, sizeof(uint32_t) * nodes.v[x].nnbr);
        memcpy(out->node_keys + (size_t)x * out->num_layers, nodes.v[x].keys,
               sizeof(uint32_t) * out->num_layers);
    }
    ovec_free(&nodes);
    out->t_construct = omp_get_wtime() - t0;

    if (!p.join) {
        out->offsets = XMALLOC(uint64_t, (size_t)n + 1);
        memset(out->offsets, 0, sizeof(uint64_t) * ((size_t)n + 1));
        out->counts = XMALLOC(uint32_t, n);
        memset(out->counts, 0, sizeof(uint32_t) * n);
        return 0;
    }

    /* node-order copy of the data (launcher.cu:3136-3146) */
    double t1 = omp_get_wtime();
    double *dn = XMALLOC(double, (size_t)n * d);
    #pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < (int64_t)n; i++)
        memcpy(dn + (size_t)i * d, data + (size_t)out->node_order[i] * d, sizeof(double) * d);
    double eps2 = eps * eps; /* launcher.cu:3200 */
    uint32_t every = p.sample_every ? p.sample_every : 1;
    nodebuf *bufs = (nodebuf *)calloc(nn, sizeof(nodebuf));
    uint64_t evals = 0, fa = 0, fn = 0;
    #pragma omp parallel for schedule(dynamic, 1) reduction(+ : evals, fa, fn)
    for (int64_t xx = 0; xx < (int64_t)nn; xx++) {
        uint32_t x = (uint32_t)xx;
        uint32_t xs = out->node_start[x], xe = out->node_start[x + 1];
        nodebuf *b = &bufs[x];
        b->cnt = XMALLOC(uint32_t, xe - xs);
        for (uint32_t i = xs; i < xe; i++) {
            uint32_t c = 0;
            if (i % every == 0) {
                const double *pp = dn + (size_t)i * d;
                for (uint32_t a = out->adj_off[x]; a < out->adj_off[x + 1]; a++) {
                    uint32_t y = out->adj[a];
                    for (uint32_t j = out->node_start[y]; j < out->node_start[y + 1]; j++) {
                        const double *qq = dn + (size_t)j * d;
                        int acc;
                        double s = seq_dist(pp, qq, d, eps2, p.seq, 1, &acc);
                        if (near_boundary(s, eps2)) {
                            int a1, a2;
                            (void)seq_dist(pp, qq, d, eps2, OR_SEQ_ALLFMA, 0, &a1);
                            (void)seq_dist(pp, qq, d, eps2, OR_SEQ_NOFMA, 0, &a2);
                            fa += (a1 != acc);
                            fn += (a2 != acc);
                        }
                        if (acc) {
                            nb_push(b, out->node_order[j]);
                            c++;
                        }
                    }
                    evals += (uint64_t)(out->node_start[y + 1] - out->node_start[y]);
                }
            }
            b->cnt[i - xs] = c;
        }
    }
    free(dn);
    out->evals_done = evals;
    out->flips_allfma = fa;
    out->flips_nofma = fn;
    double t2 = omp_get_wtime();
    out->t_join = t2 - t1;
    assemble(out, n, nn, out->node_start, out->node_order, bufs);
    free(bufs);
    out->t_output = omp_get_wtime() - t2;
    return 0;
}

int or_brute_force(const double *data, uint32_t n, uint32_t d, double eps, uint32_t seq,
                   int32_t nthreads, or_result *out)
{
    memset(out, 0, sizeof(*out));
    out->n = n;
    out->d = d;
    if (d == 0 || !(eps > 0.0))
        return -1;
    if (nthreads > 0)
        omp_set_num_threads(nthreads);
    double t1 = omp_get_wtime();
    double eps2 = eps * eps;
    /* one "node" per row, so assemble() can be reused */
    uint32_t *ident = XMALLOC(uint32_t, (size_t)n + 1);
    for (uint32_t i = 0; i <= n; i++)
        ident[i] = i;
    nodebuf *bufs = (nodebuf *)calloc(n ? n : 1, sizeof(nodebuf));
    #pragma omp parallel for schedule(dynamic, 16)
    for (int64_t ii = 0; ii < (int64_t)n; ii++) {
        nodebuf *b = &bufs[ii];
        b->cnt = XMALLOC(uint32_t, 1);
        uint32_t c = 0;
        for (uint32_t j = 0; j < n; j++) {
            int acc;
            (void)seq_dist(data + (size_t)ii * d, data + (size_t)j * d, d, eps2, seq, 1, &acc);
            if (acc) {
                nb_push(b, j);
                c++;
            }
        }
        b->cnt[0] = c;
    }
    out->evals_done = (uint64_t)n * n;
    double t2 = omp_get_wtime();
    out->t_join = t2 - t1;
    assemble(out, n, n, ident, ident, bufs);
    free(bufs);
    free(ident);
    out->t_output = omp_get_wtime() - t2;
    return 0;
}

void or_result_free(or_result *r)
{
    if (!r)
        return;
    free(r->counts);
    free(r->offsets);
    free(r->nbrs);
    free(r->node_order);
    free(r->node_start);
    free(r->adj_off);
    free(r->adj);
    free(r->node_keys);
    memset(r, 0, sizeof(*r));
}
