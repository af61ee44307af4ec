// msj_verify -- the reference's verifier nodeForce (reference
// src/nodes.cu:749-885) as a standalone tool: the index join against a
// brute force over every pair, with the missing pairs listed.
//
//   msj_verify <file.bin> <dim> <eps> [--form plain|bf16|split] [--f32]
//              [--against PREFIX] [--list K] [--device N]
//
// The input is read as msj_main reads it (raw FP64 rows, dimensions reordered
// by decreasing sigma).  The index join is the node net with the chosen
// form, or, with --against PREFIX, the arrays a previous `msj_main --out
// PREFIX` wrote (PREFIX.counts.u32, .offsets.u64, .nbrs.u32).  The brute
// force evaluates every pair with the same FP64 distance test on the GPU
// (one node holding every row), so any difference is a pair the candidate
// generator (or the saved result) lost or invented.
//
// Output (stdout), as nodeForce prints it:
//   Missing M pairs:            then (row,neighbour) pairs, at most K listed
//   Extra E pairs:              pairs the index reports that brute force does not
//   Total NodeFORCE results Set Size: R_index , unique pairs: R_brute
// Exit status: 0 identical, 3 the sets differ, 1 error, 2 usage.
#include <fcntl.h>
#include <sys/mman.h>
#include <sys/stat.h>
#include <unistd.h>

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#include "msj.h"

static int usage(const char *a0)
{
    std::fprintf(stderr,
                 "usage: %s <file.bin> <dim> <eps> [--form plain|bf16|split] [--f32]\n"
                 "          [--against PREFIX] [--list K] [--device N]\n",
                 a0);
    return 2;
}

struct Csr {
    uint32_t n = 0;
    uint64_t R = 0;
    std::vector<uint32_t> counts;
    std::vector<uint64_t> offsets;
    std::vector<uint32_t> nbrs;
};

static Csr from_result(const msj_result &r)
{
    Csr c;
    c.n = r.n;
    c.R = r.total_pairs;
    c.counts.assign(r.counts, r.counts + r.n);
    c.offsets.assign(r.offsets, r.offsets + r.n + 1);
    c.nbrs.assign(r.nbrs, r.nbrs + r.total_pairs);
    return c;
}

template <class T>
static bool read_all(const std::string &path, std::vector<T> &v, size_t expect)
{
    FILE *f = std::fopen(path.c_str(), "rb");
    if (!f)
        return false;
    std::fseek(f, 0, SEEK_END);
    long sz = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (sz < 0 || (size_t)sz % sizeof(T) != 0 || (expect != (size_t)-1 && (size_t)sz != expect * sizeof(T))) {
        std::fclose(f);
        return false;
    }
    v.resize((size_t)sz / sizeof(T));
    bool ok = v.empty() || std::fread(v.data(), sizeof(T), v.size(), f) == v.size();
    std::fclose(f);
    return ok;
}

static bool load_saved(const std::string &prefix, uint32_t n, Csr &c)
{
    c.n = n;
    if (!read_all(prefix + ".counts.u32", c.counts, n) ||
        !read_all(prefix + ".offsets.u64", c.offsets, (size_t)n + 1) ||
        !read_all(prefix + ".nbrs.u32", c.nbrs, (size_t)-1))
        return false;
    c.R = c.nbrs.size();
    if (c.offsets[0] != 0 || c.offsets[n] != c.R)
        return false;
    for (uint32_t i = 0; i < n; ++i)
        if (c.offsets[i + 1] < c.offsets[i] || c.offsets[i + 1] - c.offsets[i] != c.counts[i])
            return false;
    return true;
}

static int fail(const char *what, int rc)
{
    std::fprintf(stderr, "msj_verify: %s: %s: %s\n", what, msj_strerror(rc), msj_last_error());
    return 1;
}

int main(int argc, char *argv[])
{
    if (argc < 4)
        return usage(argv[0]);
    const char *filename = argv[1];
    char *e = nullptr;
    long dim_l = std::strtol(argv[2], &e, 10);
    double eps = 0;
    if (*e || dim_l <= 0 || std::sscanf(argv[3], "%lf", &eps) != 1 || !(eps > 0.0))
        return usage(argv[0]);
    const unsigned dim = (unsigned)dim_l;
    msj_params prm;
    msj_default_params(&prm);
    std::string against;
    uint64_t list = 1000;
    for (int i = 4; i < argc; ++i) {
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
        } else if (a == "--f32") {
            prm.storage = MSJ_STORAGE_F32;
        } else if (a == "--against" && i + 1 < argc) {
            against = argv[++i];
        } else if (a == "--list" && i + 1 < argc) {
            list = std::strtoull(argv[++i], nullptr, 10);
        } else if (a == "--device" && i + 1 < argc) {
            prm.device = std::atoi(argv[++i]);
        } else {
            return usage(argv[0]);
        }
    }
    int fd = open(filename, O_RDONLY);
    if (fd < 0) {
        std::fprintf(stderr, "msj_verify: cannot open %s: %s\n", filename, std::strerror(errno));
        return 1;
    }
    struct stat sb;
    if (fstat(fd, &sb) != 0) {
        close(fd);
        return 1;
    }
    const uint64_t n64 = (uint64_t)sb.st_size / (8ull * dim);
    if (n64 > 0xFFFFFFFEull) {
        close(fd);
        std::fprintf(stderr, "msj_verify: too many points\n");
        return 1;
    }
    const uint32_t n = (uint32_t)n64;
    const size_t len = (size_t)n64 * 8 * dim;
    void *map = len ? mmap(nullptr, len, PROT_READ, MAP_PRIVATE, fd, 0) : nullptr;
    close(fd);
    if (map == MAP_FAILED) {
        std::fprintf(stderr, "msj_verify: mmap failed: %s\n", std::strerror(errno));
        return 1;
    }

    msj_ctx *ctx = nullptr;
    int rc = msj_ctx_create(&prm, &ctx);
    if (rc != MSJ_OK)
        return fail("context", rc);
    rc = msj_ctx_load_unordered(ctx, (const double *)map, n, dim, nullptr, nullptr);
    if (len)
        munmap(map, len);
    if (rc != MSJ_OK) {
        msj_ctx_destroy(ctx);
        return fail("load", rc);
    }
    Csr idx, bf;
    msj_result r;
    if (against.empty()) {
        rc = msj_ctx_join(ctx, eps, &r);
        if (rc != MSJ_OK) {
            msj_ctx_destroy(ctx);
            return fail("index join", rc);
        }
        idx = from_result(r);
        msj_result_free(&r);
    } else if (!load_saved(against, n, idx)) {
        msj_ctx_destroy(ctx);
        std::fprintf(stderr, "msj_verify: %s.{counts.u32,offsets.u64,nbrs.u32} do not hold a "
                     "result of %u rows\n", against.c_str(), n);
        return 1;
    }
    msj_params bp = prm;
    bp.index = MSJ_INDEX_BRUTE;
    bp.form = MSJ_FORM_PLAIN;
    rc = msj_ctx_set_params(ctx, &bp);
    if (rc == MSJ_OK)
        rc = msj_ctx_join(ctx, eps, &r);
    msj_ctx_destroy(ctx);
    if (rc != MSJ_OK)
        return fail("brute force", rc);
    bf = from_result(r);
    msj_result_free(&r);

    // both lists ascending per row: one merge per row
    std::vector<std::pair<uint32_t, uint32_t>> missing, extra;
    uint64_t n_missing = 0, n_extra = 0;
    for (uint32_t row = 0; row < n; ++row) {
        uint64_t a = idx.offsets[row], ae = idx.offsets[row + 1];
        uint64_t b = bf.offsets[row], be = bf.offsets[row + 1];
        while (a < ae || b < be) {
            if (b < be && (a >= ae || bf.nbrs[b] < idx.nbrs[a])) {
                if (missing.size() < list)
                    missing.emplace_back(row, bf.nbrs[b]);
                ++n_missing;
                ++b;
            } else if (a < ae && (b >= be || idx.nbrs[a] < bf.nbrs[b])) {
                if (extra.size() < list)
                    extra.emplace_back(row, idx.nbrs[a]);
                ++n_extra;
                ++a;
            } else {
                ++a;
                ++b;
            }
        }
    }
    std::printf("Missing %llu pairs:\n", (unsigned long long)n_missing);
    for (auto &p : missing)
        std::printf("(%u,%u),", p.first, p.second);
    std::printf("\nExtra %llu pairs:\n", (unsigned long long)n_extra);
    for (auto &p : extra)
        std::printf("(%u,%u),", p.first, p.second);
    std::printf("\nTotal NodeFORCE results Set Size: %llu , unique pairs: %llu\n",
                (unsigned long long)idx.R, (unsigned long long)bf.R);
    return (n_missing || n_extra) ? 3 : 0;
}
