// msj_main -- drop-in for the reference CLI `build/main <file.bin> <dim> <numRP> <eps>`
// (reference src/main.cu:20-486).
//
// Same input (raw little-endian FP64, row-major, N = filesize/8/dim, a trailing
// partial row dropped, main.cu:42-50), same dimension reorder by decreasing
// standard deviation (main.cu:106-122), same stdout lines and the same
// whitespace-separated stderr record (main.cu:84, nodes.cu:28,329,
// launcher.cu:3120,3233,3515, main.cu:182,444).  The join runs in libmsj.so on
// the GPU through the C ABI.
//
// Extra options (after the four positional arguments):
//   --form plain|bf16|split   distance form (default plain)
//   --f32                     store coordinates as FP32 (widened exactly)
//   --device N                HIP device ordinal
//   --seed S                  srand seed of the reference-point stream
//   --out PREFIX              write PREFIX.counts.u32, PREFIX.offsets.u64,
//                             PREFIX.nbrs.u32 and PREFIX.node_order.u32
//   --verbose                 per-layer lines
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <string>
#include <vector>

#include "msj.h"

static double now_s()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

static int usage(const char *argv0)
{
    std::fprintf(stderr,
                 "usage: %s <file.bin> <dim> <numRP> <eps> [--form plain|bf16|split] [--f32]\n"
                 "          [--device N] [--seed S] [--out PREFIX] [--verbose]\n",
                 argv0);
    return 2;
}

template <class T>
static bool write_file(const std::string &path, const T *p, size_t n)
{
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f)
        return false;
    bool ok = std::fwrite(p, sizeof(T), n, f) == n;
    return std::fclose(f) == 0 && ok;
}

int main(int argc, char *argv[])
{
    if (argc < 5)
        return usage(argv[0]);
    const char *filename = argv[1];
    unsigned dim = (unsigned)std::atoi(argv[2]);
    unsigned numRP = (unsigned)std::atoi(argv[3]);
    double epsilon = 0;
    if (std::sscanf(argv[4], "%lf", &epsilon) != 1 || dim == 0) {
        std::fprintf(stderr, "msj_main: bad <dim> or <eps>\n");
        return 2;
    }
    msj_params prm;
    msj_default_params(&prm);
    std::string out_prefix;
    for (int i = 5; i < argc; ++i) {
        std::string a = argv[i];
        if (a == "--form" && i + 1 < argc) {
            std::string f = argv[++i];
            prm.form = f == "bf16" ? MSJ_FORM_MFMA_BF16 : f == "split" ? MSJ_FORM_MFMA_SPLIT
                                                                       : MSJ_FORM_PLAIN;
        } else if (a == "--f32") {
            prm.storage = MSJ_STORAGE_F32;
        } else if (a == "--device" && i + 1 < argc) {
            prm.device = std::atoi(argv[++i]);
        } else if (a == "--seed" && i + 1 < argc) {
            prm.seed = (uint32_t)std::strtoul(argv[++i], nullptr, 10);
        } else if (a == "--out" && i + 1 < argc) {
            out_prefix = argv[++i];
        } else if (a == "--verbose") {
            prm.verbose = 1;
        } else {
            return usage(argv[0]);
        }
    }

    double time0 = now_s();
    std::ifstream file(filename, std::ios::in | std::ios::binary);
    if (!file) {
        std::fprintf(stderr, "msj_main: cannot open %s\n", filename);
        return 1;
    }
    file.seekg(0, std::ios::end);
    size_t size = (size_t)file.tellg();
    file.seekg(0, std::ios::beg);
    uint64_t numPoints64 = size / sizeof(double) / dim;
    if (numPoints64 > 0xFFFFFFFEull) {
        std::fprintf(stderr, "msj_main: %llu points exceed the 32-bit point ids\n",
                     (unsigned long long)numPoints64);
        return 1;
    }
    unsigned numPoints = (unsigned)numPoints64;
    std::vector<double> A((size_t)numPoints * dim);
    file.read(reinterpret_cast<char *>(A.data()), (std::streamsize)(A.size() * sizeof(double)));
    file.close();
    double time00 = now_s();
    std::printf("\nTime to read in file: %f\n", time00 - time0);
    std::fprintf(stderr, "\n%f %u ", epsilon, numPoints);

    std::printf("\nNumber points: %d ", numPoints);
    std::printf("\nNumber Dimensions: %d ", dim);
    std::printf("\nNumber Reference Points: %d ", numRP);
    std::printf("\nNumber Concurent Streams: %d", 2);
    std::printf("\nBlock Size: %d, Kernel Blocks: %d", BLOCK_SIZE, KERNEL_BLOCKS);
    std::printf("\nUsing dynamic searching");
    std::printf("\nDistance Threshold: %f \n*********************************\n\n", epsilon);

    double time1 = now_s();
    // main.cu:106-122: order dimensions by decreasing sigma and reorder rows
    std::vector<uint32_t> order(dim);
    if (numPoints > 0) {
        int rc = msj_dim_order(A.data(), numPoints, dim, order.data(), nullptr);
        if (rc != MSJ_OK) {
            std::fprintf(stderr, "msj_main: %s: %s\n", msj_strerror(rc), msj_last_error());
            return 1;
        }
    } else {
        for (unsigned j = 0; j < dim; ++j)
            order[j] = j;
    }
    const bool f32 = prm.storage == MSJ_STORAGE_F32;
    std::vector<double> ordered64;
    std::vector<float> ordered32;
    if (f32)
        ordered32.resize((size_t)numPoints * dim);
    else
        ordered64.resize((size_t)numPoints * dim);
    for (size_t i = 0; i < numPoints; ++i)
        for (unsigned j = 0; j < dim; ++j) {
            double v = A[i * dim + order[j]];
            if (f32)
                ordered32[i * dim + j] = (float)v;
            else
                ordered64[i * dim + j] = v;
        }

    double time2 = now_s();
    std::fprintf(stderr, "%u ", (unsigned)RPPERLAYER);
    msj_result res;
    const void *data = f32 ? (const void *)ordered32.data() : (const void *)ordered64.data();
    int rc = msj_self_join(data, numPoints, dim, epsilon, &prm, &res);
    if (rc != MSJ_OK) {
        std::fprintf(stderr, "msj_main: %s: %s\n", msj_strerror(rc), msj_last_error());
        return 1;
    }
    const msj_stats &st = res.stats;
    for (unsigned l = 0; l < st.num_layers; ++l)
        std::printf("SubGraph 0 Layer %u Selecting RP %u with Nodes: %u and calcs: %llu and "
                    "stdev %f::\n",
                    l, st.chosen[l], st.nodes_per_layer[l],
                    (unsigned long long)st.calcs_per_layer[l], st.sse[l]);
    std::printf("Final graph has %u nodes, %llu calcs total\n", st.num_nodes,
                (unsigned long long)st.num_calcs);
    std::fprintf(stderr, "%u %u %llu ", st.num_layers, st.num_nodes,
                 (unsigned long long)st.num_calcs);
    const double constructT = st.ms_index / 1e3;
    std::printf("Node Construct time: %f\n", constructT);
    std::fprintf(stderr, "%f ", constructT);
    std::fprintf(stderr, "%u %f ", st.ref_num_batches, st.ref_var_calc);
    std::printf("Batchs: %u\n", st.ref_num_batches);
    const double kernelT = (st.ms_join + st.ms_output + st.ms_d2h) / 1e3;
    std::printf("Total results Set Size: %llu \n", (unsigned long long)res.total_pairs);
    std::printf("Kernel Time: %f\n", kernelT);
    std::fprintf(stderr, "%g %llu ", kernelT, (unsigned long long)res.total_pairs);
    double time3 = now_s();
    std::printf("Node Laucnher time: %f\n", time3 - time2);
    std::fprintf(stderr, "%f ", time3 - time2);

    if (!out_prefix.empty()) {
        bool ok = write_file(out_prefix + ".counts.u32", res.counts, numPoints) &&
                  write_file(out_prefix + ".offsets.u64", res.offsets, (size_t)numPoints + 1) &&
                  write_file(out_prefix + ".nbrs.u32", res.nbrs, (size_t)res.total_pairs) &&
                  write_file(out_prefix + ".node_order.u32", res.node_order, numPoints);
        if (!ok) {
            std::fprintf(stderr, "msj_main: writing %s.* failed\n", out_prefix.c_str());
            msj_result_free(&res);
            return 1;
        }
    }
    msj_result_free(&res);
    std::printf("Total Time: %f\n", now_s() - time1);
    std::fprintf(stderr, "%f ", now_s() - time1);
    return 0;
}
