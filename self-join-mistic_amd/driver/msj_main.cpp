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
// The input file is memory-mapped and streamed to the GPU once (the
// dimension order is computed there), so the host never holds a copy of the
// data: inputs larger than host memory work.
//
// Extra options (after the four positional arguments):
//   --form plain|bf16|split   distance form (default plain)
//   --index nodenet|brute|coss  candidate generator (default: the node net;
//                             coss: the tree's bins, tree.cu)
//   --f32                     store coordinates as FP32 (rounded to nearest)
//   --device N                HIP device ordinal
//   --seed S                  srand seed of the reference-point stream
//   --page-pairs K            neighbour ids per host page (0: library default)
//   --out PREFIX              write PREFIX.counts.u32, PREFIX.offsets.u64,
//                             PREFIX.nbrs.u32, PREFIX.node_order.u32 and
//                             PREFIX.manifest.json (what the files hold)
//   --strict                  a trailing partial row is an error (exit 1)
//   --dry-run                 check the input, print the header, no GPU work
//   --verbose                 per-layer lines
// Exit status: 0 success, 1 input or library error, 2 usage.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cerrno>
#include <chrono>
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <cstring>
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
                 "usage: %s <file.bin> <dim> <numRP> <eps> [--form plain|bf16|split]\n"
                 "          [--index nodenet|brute|coss] [--f32] [--device N] [--seed S]\n"
                 "          [--page-pairs K] [--out PREFIX] [--strict] [--dry-run] [--verbose]\n",
                 argv0);
    return 2;
}

template <class T>
static bool write_file(const std::string &path, const T *p, size_t n)
{
    FILE *f = std::fopen(path.c_str(), "wb");
    if (!f)
        return false;
    bool ok = n == 0 || std::fwrite(p, sizeof(T), n, f) == n;
    return std::fclose(f) == 0 && ok;
}

static std::string json_str(const std::string &s)
{
    std::string o = "\"";
    for (char c : s) {
        if (c == '"' || c == '\\') {
            o += '\\';
            o += c;
        } else if ((unsigned char)c < 0x20) {
            char b[8];
            std::snprintf(b, sizeof(b), "\\u%04x", (unsigned)(unsigned char)c);
            o += b;
        } else {
            o += c;
        }
    }
    return o + "\"";
}

static std::string base_name(const std::string &p)
{
    size_t k = p.find_last_of('/');
    return k == std::string::npos ? p : p.substr(k + 1);
}

// PREFIX.manifest.json: the shape, dtype and meaning of every array file.
static bool write_manifest(const std::string &prefix, const char *input, unsigned n, unsigned dim,
                           double eps, const msj_params &prm, const msj_result &res,
                           const std::vector<uint32_t> &order)
{
    uint64_t eps_bits;
    std::memcpy(&eps_bits, &eps, 8);
    const char *forms[] = {"plain", "bf16", "split"};
    const char *indexes[] = {"nodenet", "brute", "coss"};
    const msj_stats &st = res.stats;
    std::string j = "{\n";
    j += "  \"format\": \"msj-result-1\",\n";
    j += "  \"input\": " + json_str(input) + ",\n";
    j += "  \"n\": " + std::to_string(n) + ",\n";
    j += "  \"d\": " + std::to_string(dim) + ",\n";
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%.17g", eps);
    j += std::string("  \"eps\": ") + buf + ",\n";
    std::snprintf(buf, sizeof(buf), "%016" PRIx64, eps_bits);
    j += std::string("  \"eps_hex\": \"") + buf + "\",\n";
    j += std::string("  \"storage\": \"") + (prm.storage == MSJ_STORAGE_F32 ? "f32" : "f64") +
         "\",\n";
    j += std::string("  \"form\": \"") + forms[prm.form % 3] + "\",\n";
    j += std::string("  \"form_used\": \"") + forms[st.form_used % 3] + "\",\n";
    j += std::string("  \"index\": \"") + indexes[prm.index % 3] + "\",\n";
    j += "  \"dim_order\": [";
    for (size_t k = 0; k < order.size(); ++k)
        j += (k ? ", " : "") + std::to_string(order[k]);
    j += "],\n";
    j += "  \"total_pairs\": " + std::to_string(res.total_pairs) + ",\n";
    j += "  \"num_nodes\": " + std::to_string(st.num_nodes) + ",\n";
    j += "  \"num_calcs\": " + std::to_string(st.num_calcs) + ",\n";
    j += "  \"rows\": \"0-based row positions in the input file\",\n";
    j += "  \"neighbours\": \"per row ascending, self included: nbrs[offsets[r]:offsets[r+1]]\",\n";
    j += "  \"files\": {\n";
    struct F {
        const char *key, *suffix, *dtype;
        uint64_t count;
        const char *what;
    } files[] = {
        {"counts", ".counts.u32", "<u4", n, "neighbours of each row"},
        {"offsets", ".offsets.u64", "<u8", (uint64_t)n + 1, "CSR offsets into nbrs"},
        {"nbrs", ".nbrs.u32", "<u4", res.total_pairs, "neighbour row ids"},
        {"node_order", ".node_order.u32", "<u4", n,
         "pointArray: row of each node-order position"},
    };
    for (size_t k = 0; k < 4; ++k) {
        j += std::string("    \"") + files[k].key + "\": {\"path\": " +
             json_str(base_name(prefix) + files[k].suffix) + ", \"dtype\": \"" + files[k].dtype +
             "\", \"count\": " + std::to_string(files[k].count) + ", \"meaning\": \"" +
             files[k].what + "\"}" + (k < 3 ? ",\n" : "\n");
    }
    j += "  }\n}\n";
    FILE *f = std::fopen((prefix + ".manifest.json").c_str(), "w");
    if (!f)
        return false;
    bool ok = std::fwrite(j.data(), 1, j.size(), f) == j.size();
    return std::fclose(f) == 0 && ok;
}

struct Mapping {
    void *p = MAP_FAILED;
    size_t len = 0;
    int fd = -1;
    ~Mapping()
    {
        if (p != MAP_FAILED && len)
            munmap(p, len);
        if (fd >= 0)
            close(fd);
    }
};

static int fail(const char *what, int rc)
{
    std::fprintf(stderr, "msj_main: %s: %s: %s\n", what, msj_strerror(rc), msj_last_error());
    return 1;
}

int main(int argc, char *argv[])
{
    if (argc < 5)
        return usage(argv[0]);
    const char *filename = argv[1];
    char *endp = nullptr;
    long dim_l = std::strtol(argv[2], &endp, 10);
    if (*argv[2] == '\0' || *endp != '\0' || dim_l <= 0 || dim_l > 0x7FFFFFFFL) {
        std::fprintf(stderr, "msj_main: bad <dim> %s\n", argv[2]);
        return 2;
    }
    unsigned dim = (unsigned)dim_l;
    unsigned numRP = (unsigned)std::atoi(argv[3]);
    double epsilon = 0;
    if (std::sscanf(argv[4], "%lf", &epsilon) != 1 || !(epsilon > 0.0)) {
        std::fprintf(stderr, "msj_main: bad <eps> %s (must be positive)\n", argv[4]);
        return 2;
    }
    msj_params prm;
    msj_default_params(&prm);
    std::string out_prefix;
    bool strict = false, dry = false;
    for (int i = 5; i < argc; ++i) {
        std::string a = argv[i];
        if (a == "--form" && i + 1 < argc) {
            std::string f = argv[++i];
            if (f == "plain")
                prm.form = MSJ_FORM_PLAIN;
            else if (f == "bf16")
                prm.form = MSJ_FORM_MFMA_BF16;
            else if (f == "split")
                prm.form = MSJ_FORM_MFMA_SPLIT;
            else
                return usage(argv[0]);
        } else if (a == "--index" && i + 1 < argc) {
            std::string f = argv[++i];
            if (f == "nodenet")
                prm.index = MSJ_INDEX_NODENET;
            else if (f == "brute")
                prm.index = MSJ_INDEX_BRUTE;
            else if (f == "coss")
                prm.index = MSJ_INDEX_COSS;
            else
                return usage(argv[0]);
        } else if (a == "--f32") {
            prm.storage = MSJ_STORAGE_F32;
        } else if (a == "--device" && i + 1 < argc) {
            prm.device = std::atoi(argv[++i]);
        } else if (a == "--seed" && i + 1 < argc) {
            prm.seed = (uint32_t)std::strtoul(argv[++i], nullptr, 10);
        } else if (a == "--page-pairs" && i + 1 < argc) {
            prm.page_pairs = std::strtoull(argv[++i], nullptr, 10);
        } else if (a == "--out" && i + 1 < argc) {
            out_prefix = argv[++i];
        } else if (a == "--strict") {
            strict = true;
        } else if (a == "--dry-run") {
            dry = true;
        } else if (a == "--verbose") {
            prm.verbose = 1;
        } else {
            return usage(argv[0]);
        }
    }

    double time0 = now_s();
    Mapping map;
    map.fd = open(filename, O_RDONLY);
    if (map.fd < 0) {
        std::fprintf(stderr, "msj_main: cannot open %s: %s\n", filename, std::strerror(errno));
        return 1;
    }
    struct stat sb;
    if (fstat(map.fd, &sb) != 0 || !S_ISREG(sb.st_mode)) {
        std::fprintf(stderr, "msj_main: %s is not a regular file\n", filename);
        return 1;
    }
    const uint64_t size = (uint64_t)sb.st_size;
    const uint64_t row_bytes = (uint64_t)sizeof(double) * dim;
    const uint64_t numPoints64 = size / row_bytes;  // main.cu:50
    const uint64_t tail = size - numPoints64 * row_bytes;
    if (tail != 0) {
        if (strict) {
            std::fprintf(stderr,
                         "msj_main: %s: %llu bytes are not a multiple of a row (%u x 8 bytes): "
                         "%llu trailing bytes\n",
                         filename, (unsigned long long)size, dim, (unsigned long long)tail);
            return 1;
        }
        std::printf("Warning: %llu trailing bytes of %s ignored (not a whole row of %u doubles)\n",
                    (unsigned long long)tail, filename, dim);
    }
    if (numPoints64 > 0xFFFFFFFEull) {
        std::fprintf(stderr, "msj_main: %llu points exceed the 32-bit point ids\n",
                     (unsigned long long)numPoints64);
        return 1;
    }
    if (numPoints64 * dim >= (1ull << 40)) {
        std::fprintf(stderr, "msj_main: %llu points x %u dimensions exceed 2^40 values\n",
                     (unsigned long long)numPoints64, dim);
        return 1;
    }
    const unsigned numPoints = (unsigned)numPoints64;
    const size_t map_len = (size_t)(numPoints64 * row_bytes);
    if (map_len) {
        map.p = mmap(nullptr, map_len, PROT_READ, MAP_PRIVATE, map.fd, 0);
        if (map.p == MAP_FAILED) {
            std::fprintf(stderr, "msj_main: mmap of %s failed: %s\n", filename,
                         std::strerror(errno));
            return 1;
        }
        map.len = map_len;
        (void)madvise(map.p, map_len, MADV_SEQUENTIAL);
    }
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
    if (dry) {
        std::printf("Dry run: input checked, no join\n");
        std::fflush(stdout);
        return 0;
    }

    double time1 = now_s();
    msj_ctx *ctx = nullptr;
    int rc = msj_ctx_create(&prm, &ctx);
    if (rc != MSJ_OK)
        return fail("context", rc);
    struct CtxGuard {
        msj_ctx *c;
        ~CtxGuard() { msj_ctx_destroy(c); }
    } cg{ctx};
    // main.cu:106-122: dimensions by decreasing sigma, on the device
    std::vector<uint32_t> order(dim);
    rc = msj_ctx_load_unordered(ctx, (const double *)(map.len ? map.p : nullptr), numPoints, dim,
                                order.data(), nullptr);
    if (rc != MSJ_OK)
        return fail("load", rc);

    double time2 = now_s();
    std::fprintf(stderr, "%u ", (unsigned)RPPERLAYER);
    msj_result res;
    rc = msj_ctx_join(ctx, epsilon, &res);
    if (rc != MSJ_OK)
        return fail("join", rc);
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
                  write_file(out_prefix + ".node_order.u32", res.node_order, numPoints) &&
                  write_manifest(out_prefix, filename, numPoints, dim, epsilon, prm, res, order);
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
