// msj_csv2bin -- text table to the raw FP64 input of msj_main, as the
// reference's converter does (reference toCSV.cpp:51-85, despite its name):
//
//   msj_csv2bin <in.csv> <out.bin> [--dim D]
//
// Values are read as the reference's loop reads them (fscanf "%lf, " then
// "%lf "): a number, an optional comma, any whitespace (newlines included),
// repeated; reading stops at the first text that is not a number.  Every
// value is written as a little-endian FP64, in reading order.
//
// With --dim D the table is also checked: every line must hold exactly D
// values and reading must reach the end of the file (exit 3 otherwise, with
// the line number); without it the reference's behaviour is kept.
// Exit status: 0 success, 1 I/O error, 2 usage, 3 malformed table (--dim).
#include <cctype>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

static int usage(const char *a0)
{
    std::fprintf(stderr, "usage: %s <in.csv> <out.bin> [--dim D]\n", a0);
    return 2;
}

int main(int argc, char *argv[])
{
    if (argc != 3 && argc != 5)
        return usage(argv[0]);
    long dim = 0;
    if (argc == 5) {
        if (std::strcmp(argv[3], "--dim") != 0)
            return usage(argv[0]);
        char *e = nullptr;
        dim = std::strtol(argv[4], &e, 10);
        if (*e || dim <= 0)
            return usage(argv[0]);
    }
    FILE *in = std::fopen(argv[1], "rb");
    if (!in) {
        std::printf("No such File\n");  // toCSV.cpp:57
        std::fprintf(stderr, "msj_csv2bin: cannot open %s: %s\n", argv[1], std::strerror(errno));
        return 1;
    }
    std::string text;
    {
        std::vector<char> buf(1 << 20);
        size_t k;
        while ((k = std::fread(buf.data(), 1, buf.size(), in)) > 0)
            text.append(buf.data(), k);
        std::fclose(in);
    }
    std::vector<double> data;
    const char *p = text.c_str(), *end = p + text.size();
    size_t line = 1, on_line = 0;
    bool ragged = false;
    size_t bad_line = 0;
    auto end_line = [&]() {
        if (dim && on_line != 0 && on_line != (size_t)dim && !ragged) {
            ragged = true;
            bad_line = line;
        }
        ++line;
        on_line = 0;
    };
    while (true) {
        while (p < end && std::isspace((unsigned char)*p)) {  // scanf skips white space
            if (*p == '\n')
                end_line();
            ++p;
        }
        if (p >= end)
            break;
        char *q = nullptr;
        errno = 0;
        const double v = std::strtod(p, &q);  // the conversion of %lf
        if (q == p)
            break;  // not a number: the reference's loop ends here
        data.push_back(v);
        ++on_line;
        p = q;
        if (p < end && *p == ',')
            ++p;
    }
    const bool rest = p < end;
    if (on_line)
        end_line();
    std::printf("size of data: %zu\n", data.size());
    if (dim) {
        if (ragged) {
            std::fprintf(stderr, "msj_csv2bin: line %zu does not hold %ld values\n", bad_line, dim);
            return 3;
        }
        if (rest) {
            std::fprintf(stderr, "msj_csv2bin: text that is not a number at line %zu\n", line);
            return 3;
        }
        if (data.size() % (size_t)dim) {
            std::fprintf(stderr, "msj_csv2bin: %zu values are not rows of %ld\n", data.size(), dim);
            return 3;
        }
    }
    FILE *out = std::fopen(argv[2], "wb");
    if (!out) {
        std::fprintf(stderr, "msj_csv2bin: cannot create %s: %s\n", argv[2], std::strerror(errno));
        return 1;
    }
    bool ok = data.empty() || std::fwrite(data.data(), sizeof(double), data.size(), out) == data.size();
    if (std::fclose(out) != 0 || !ok) {
        std::fprintf(stderr, "msj_csv2bin: writing %s failed\n", argv[2]);
        return 1;
    }
    return 0;
}
