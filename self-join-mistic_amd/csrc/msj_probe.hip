// msj_probe.hip -- read-only probes of the index build (include/msj_probe.h).
//
// They run the same device code as msj_self_join (IndexBuilder::build and
// its key helpers) and copy intermediate state to the caller, so tests can
// compare it with a CPU restatement: the candidate keys of one layer
// (reference kernel.cu:880-909 binningKernel, nodes.cu:128-135 axis keys),
// the final nodes' key tuples (nodes.cuh:9-38 binNumbers) and the device
// sqrt / division behind every key.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "msj_internal.h"
#include "msj_probe.h"

#pragma clang fp contract(off)

namespace {

using msj::Error;

__global__ __launch_bounds__(256) void k_key_math(const double *__restrict__ D, uint64_t m,
                                                  double eps, double *__restrict__ q,
                                                  uint32_t *__restrict__ key)
{
    const uint64_t stride = (uint64_t)gridDim.x * blockDim.x;
    for (uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m; i += stride) {
        const double v = D[i];
        q[i] = sqrt(v) / eps;
        key[i] = msj::rp_key(v, eps);
    }
}

struct DevSet {
    int prev = -1;
    explicit DevSet(int dev)
    {
        if (hipGetDevice(&prev) != hipSuccess)
            prev = -1;
        MSJ_HIP(hipSetDevice(dev));
    }
    ~DevSet()
    {
        if (prev >= 0)
            (void)hipSetDevice(prev);
    }
};

void check_device(int dev)
{
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
        (void)hipGetLastError();
        throw Error(MSJ_E_NODEV, "no HIP device visible");
    }
    if (dev < 0 || dev >= ndev)
        throw Error(MSJ_E_NODEV, "device ordinal out of range");
}

struct ProbeArgs {
    const void *data;
    uint32_t n, d;
    double eps;
    const msj_params *p;
    msj_index_probe *probe;
};

void probe_index(void *a_)
{
    ProbeArgs &a = *(ProbeArgs *)a_;
    if (!a.probe)
        throw Error(MSJ_E_INVAL, "probe is NULL");
    msj_params prm;
    if (a.p)
        prm = *a.p;
    else
        msj_default_params(&prm);
    if (prm.struct_size != sizeof(msj_params))
        throw Error(MSJ_E_INVAL, "msj_params.struct_size mismatch: call msj_default_params first");
    if (prm.maxrp == 0 || prm.maxrp > MSJ_MAX_LAYERS || prm.rp_per_layer == 0 ||
        prm.rp_per_layer > 32)
        throw Error(MSJ_E_INVAL, "maxrp must be in 1..16, rp_per_layer in 1..32");
    if (a.d == 0 || a.n == 0 || !a.data)
        throw Error(MSJ_E_INVAL, "probe needs n > 0, d > 0 and data");
    if (!(a.eps > 0.0) || !std::isfinite(a.eps))
        throw Error(MSJ_E_INVAL, "eps must be positive and finite");
    if ((uint64_t)a.n * a.d >= (1ull << 40))
        throw Error(MSJ_E_OVERFLOW, "n*d too large");
    check_device(prm.device);
    DevSet g(prm.device);
    hipStream_t s = nullptr;
    MSJ_HIP(hipStreamCreate(&s));
    struct StreamGuard {
        hipStream_t s;
        ~StreamGuard() { (void)hipStreamDestroy(s); }
    } sg{s};
    const bool f32 = prm.storage == MSJ_STORAGE_F32;
    msj::DevBuf<unsigned char> xs, staging;
    msj::load_soa(a.data, a.n, a.d, f32, xs, staging, s);
    staging.release();
    msj::IndexProbe pr;
    pr.layer = a.probe->layer;
    pr.keys = a.probe->keys;
    pr.node_keys = a.probe->node_keys != nullptr;
    msj::IndexBuilder B;
    B.probe = &pr;
    msj::Index ix;
    msj::LayerStats ls;
    B.build(xs.p, f32, a.n, a.d, a.eps, prm, ix, ls, s);
    MSJ_HIP(hipStreamSynchronize(s));
    const uint32_t M = ix.M, L = ls.layers;
    a.probe->ncand = ls.ncand;
    a.probe->num_nodes = M;
    a.probe->num_layers = L;
    if (a.probe->node_start)
        MSJ_HIP(hipMemcpy(a.probe->node_start, ix.nstart.p, sizeof(uint32_t) * ((size_t)M + 1),
                          hipMemcpyDeviceToHost));
    if (a.probe->node_order)
        MSJ_HIP(hipMemcpy(a.probe->node_order, ix.perm.p, sizeof(uint32_t) * a.n,
                          hipMemcpyDeviceToHost));
    for (uint32_t l = 0; l < MSJ_MAX_LAYERS; ++l) {
        a.probe->chosen[l] = ls.chosen[l];
        a.probe->nodes_per_layer[l] = ls.nodes[l];
        a.probe->sse[l] = ls.sse[l];
    }
    a.probe->num_edges = ix.edges;
    a.probe->num_calcs = ix.num_calcs;
    if (a.probe->adj_off)
        MSJ_HIP(hipMemcpy(a.probe->adj_off, ix.aoff.p, sizeof(uint32_t) * ((size_t)M + 1),
                          hipMemcpyDeviceToHost));
    if (a.probe->adj && a.probe->adj_cap)
        MSJ_HIP(hipMemcpy(a.probe->adj, ix.adj.p,
                          sizeof(uint32_t) * std::min<uint64_t>(a.probe->adj_cap, ix.edges),
                          hipMemcpyDeviceToHost));
    if (a.probe->node_keys && prm.index == MSJ_INDEX_COSS) {
        if (pr.tuples.size() != (size_t)M * L)
            throw Error(MSJ_E_HIP, "probe: tree/COSS bin numbers missing");
        std::copy(pr.tuples.begin(), pr.tuples.end(), a.probe->node_keys);
    } else if (a.probe->node_keys) {
        // walk each final node up its parents: layer l's children record the
        // index of their parent among layer l-1's nodes
        for (uint32_t c = 0; c < M; ++c) {
            uint32_t v = c;
            for (uint32_t l = L; l-- > 0;) {
                if (v >= pr.ckey[l].size())
                    throw Error(MSJ_E_HIP, "probe: inconsistent parent links");
                a.probe->node_keys[(size_t)c * L + l] = pr.ckey[l][v];
                v = pr.cparent[l][v];
            }
        }
    }
}

struct KeyMathArgs {
    const double *D;
    uint64_t m;
    double eps;
    double *q;
    uint32_t *key;
    int device;
};

void probe_key_math(void *a_)
{
    KeyMathArgs &a = *(KeyMathArgs *)a_;
    if (a.m && !a.D)
        throw Error(MSJ_E_INVAL, "D is NULL");
    if (!(a.eps > 0.0))
        throw Error(MSJ_E_INVAL, "eps must be positive");
    if (a.m == 0)
        return;
    check_device(a.device);
    DevSet g(a.device);
    msj::DevBuf<double> dD, dq;
    msj::DevBuf<uint32_t> dk;
    dD.ensure(a.m);
    dq.ensure(a.m);
    dk.ensure(a.m);
    MSJ_HIP(hipMemcpy(dD.p, a.D, sizeof(double) * a.m, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_key_math, dim3(msj::grid_stride(a.m, 256)), dim3(256), 0, nullptr, dD.p,
                       a.m, a.eps, dq.p, dk.p);
    MSJ_HIP(hipGetLastError());
    MSJ_HIP(hipDeviceSynchronize());
    if (a.q)
        MSJ_HIP(hipMemcpy(a.q, dq.p, sizeof(double) * a.m, hipMemcpyDeviceToHost));
    if (a.key)
        MSJ_HIP(hipMemcpy(a.key, dk.p, sizeof(uint32_t) * a.m, hipMemcpyDeviceToHost));
}

}  // namespace

extern "C" {

int msj_probe_index(const void *data, uint32_t n, uint32_t d, double eps, const msj_params *p,
                    msj_index_probe *probe)
{
    ProbeArgs a{data, n, d, eps, p, probe};
    return msj::guarded_call(probe_index, &a);
}

int msj_probe_key_math(const double *D, uint64_t m, double eps, double *q, uint32_t *key,
                       int device)
{
    KeyMathArgs a{D, m, eps, q, key, device};
    return msj::guarded_call(probe_key_math, &a);
}

double msj_probe_repeat_add(double s, double c, uint64_t g)
{
    return msj::repeat_add(s, c, g);
}

}  // extern "C"
