// rawdata_bench.cpp -- checks and times the driver's ASCII field writer.
//
//   rawdata_bench DIR N THREADS
//
// Writes small fields (plain and cylindrical-mirrored layouts, special values)
// with 1, 3, 8 and 64 threads and compares the bytes with the serial ostream
// text of save_ascii (src/output_styles.cpp:182-327), then formats one N x N
// field without writing it and writes it, each with THREADS threads and with
// one, printing MB/s.  Exit status 0 and a last line "OK" when every
// comparison matched.

#include "rawdata.h"
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <random>
#include <sstream>
using namespace brp_driver;

static std::string serial(const Field &f, int X, int Y, bool cyl, double dx, double dy) {
    std::ostringstream out;
    for (int j = 0; j <= Y; j++) {
        for (int i = cyl ? -X : 0; i <= X; i++)
            out << (double)i * dx << "\t" << (double)j * dy << "\t" << f.at(std::abs(i), j) << "\n";
        out << "\n";
    }
    return out.str();
}

static std::string slurp(const std::string &p) {
    std::ifstream in(p, std::ios::binary);
    std::ostringstream s;
    s << in.rdbuf();
    return s.str();
}

int main(int argc, char **argv) {
    const std::string dir = argv[1];
    const int big = std::atoi(argv[2]), threads = std::atoi(argv[3]);
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    // small fields with special values: equality with the serial ostream text
    for (int cyl = 0; cyl <= 1; cyl++) {
        const int X = 37, Y = 23;
        Field f;
        f.nx = X + 2; f.ny = Y + 2;
        f.v.resize((size_t)f.nx * f.ny);
        for (auto &v : f.v) v = u(rng) * std::pow(10.0, 12 * u(rng));
        f.v[3] = 0.0; f.v[4] = -0.0; f.v[5] = 1e-300; f.v[6] = -123456789.0; f.v[7] = 0.5;
        const double dx = 1e-4, dy = 1.25e-4;
        const std::string want = serial(f, X, Y, cyl, dx, dy);
        for (int t : {1, 3, 8, 64}) {
            double mn = 1e10, mx = -1e10;
            const std::string p = dir + "/f.txt";
            if (!write_ascii_field(p, f, X, Y, cyl, dx, dy, t, &mn, &mx)) return 2;
            if (slurp(p) != want) { std::printf("MISMATCH cyl=%d threads=%d\n", cyl, t); return 1; }
        }
    }
    // one 4096^2 field: throughput with `threads` threads and with one
    Field f;
    f.nx = big + 2; f.ny = big + 2;
    f.v.resize((size_t)f.nx * f.ny);
    for (auto &v : f.v) v = 1.0e6 * u(rng);
    for (int t : {threads, 1}) {  // formatting alone (no file system)
        const auto t0 = std::chrono::steady_clock::now();
        const long long n = format_ascii_field(f, big, big, false, 1e-4, 1e-4, t);
        const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::printf("format threads %d: %.1f MB in %.2f s = %.0f MB/s\n", t, n / 1e6, s, n / 1e6 / s);
    }
    for (int t : {threads, 1}) {
        const std::string p = dir + "/big.txt";
        const auto t0 = std::chrono::steady_clock::now();
        if (!write_ascii_field(p, f, big, big, false, 1e-4, 1e-4, t, nullptr, nullptr)) return 3;
        const double s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
        std::ifstream in(p, std::ios::binary | std::ios::ate);
        const double mb = (double)in.tellg() / 1e6;
        std::printf("threads %d: %.1f MB in %.2f s = %.0f MB/s\n", t, mb, s, mb / s);
        std::remove(p.c_str());
    }
    std::printf("OK\n");
    return 0;
}
