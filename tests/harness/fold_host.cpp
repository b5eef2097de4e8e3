// Host build of the one-pass fold of the bucket kernels' field products
// (fold_est in gpu_msm_amd/csrc/field.hpp): reads lines
// "KMIN KMAX t0 .. t14 top" (15 normalised 28-bit limbs and the signed limb 15
// of a value v) and prints the 16 limbs of v - floor(v / 2p) 2p.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <iostream>
#include <string>

#include "field.hpp"

namespace {

template <int A, int B>
void fold_line(const char* p) {
    uint32_t t[gm::NL];
    int n = 0;
    for (int i = 0; i < gm::NL - 1; ++i) {
        unsigned x = 0;
        if (sscanf(p, " %u%n", &x, &n) != 1) return;
        p += n;
        t[i] = x;
    }
    long long top = 0;
    if (sscanf(p, " %lld", &top) != 1) return;
    gm::Fp r;
    gm::fold_est<A, B>(r, t, (int32_t)top);
    for (int i = 0; i < gm::NL; ++i) printf("%u%c", r.v[i], i + 1 < gm::NL ? ' ' : '\n');
}

}  // namespace

int main() {
    std::string line;
    while (std::getline(std::cin, line)) {
        int a = 0, b = 0, n = 0;
        if (sscanf(line.c_str(), "%d %d%n", &a, &b, &n) != 2) continue;
        const char* p = line.c_str() + n;
        if (a == -1 && b == 1)
            fold_line<-1, 1>(p);
        else if (a == -2 && b == 1)
            fold_line<-2, 1>(p);
        else if (a == 0 && b == 3)
            fold_line<0, 3>(p);
        else if (a == 0 && b == 5)
            fold_line<0, 5>(p);
        else
            printf("unsupported\n");
    }
    return 0;
}
