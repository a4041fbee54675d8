// Prints what cuda-rp_amd/csrc/rp_vfit.h builds, for tests/test_rp_tables.py:
//   rp_vfit_dump table K3      the V and Y tables for k3 = K3, one
//                              coefficient per line (%a, exact)
//   rp_vfit_dump cf S...       U(S) and W(S) from the continued fraction
//                              (re im re im, %La, exact)
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "rp_vfit.h"

int main(int argc, char **argv) {
    if (argc >= 3 && !std::strcmp(argv[1], "table")) {
        std::vector<double> t(2 * brp::kVIntervals * (brp::kVDegree + 1) * 2);
        brp::rp_vtable_build(std::strtod(argv[2], nullptr), t.data());
        for (double x : t) std::printf("%a\n", x);
        return 0;
    }
    if (argc >= 2 && !std::strcmp(argv[1], "cf")) {
        for (int i = 2; i < argc; i++) {
            const long double s = std::strtold(argv[i], nullptr);
            const brp::RpCxl u = brp::rp_impedance_cf(s, 3), w = brp::rp_impedance_cf(s, 5);
            std::printf("%La %La %La %La\n", u.re, u.im, w.re, w.im);
        }
        return 0;
    }
    std::fprintf(stderr, "usage: rp_vfit_dump table K3 | cf S...\n");
    return 2;
}
