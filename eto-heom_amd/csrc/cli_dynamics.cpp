// etoheom_dynamics <file.key> -- drop-in for the reference CPU_dynamics executable
// (etoheom-cpu/src/CPU_dynamics.cpp:14-49): same header lines and per-site
// population blocks for sites 1 and 2, values in the default ostream format.
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "etoheom.h"

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <file.key>\n", argv[0]);
        return 2;
    }
    std::printf("Now is CPU version\n");
    std::printf("Running %s\n", argv[1]);
    std::printf("The bath type is \n");   // param_dynamics never sets bath_type
    const char* dev = std::getenv("ETOHEOM_DEVICE");
    eh_ctx* ctx = nullptr;
    int rc = eh_ctx_create(dev ? std::atoi(dev) : 0, &ctx);
    if (rc) {
        std::fprintf(stderr, "etoheom_dynamics: %s\n", eh_last_error());
        return 1;
    }
    const int sites[2] = {1, 2};
    int64_t n = 0;
    rc = eh_dynamics_key_run(ctx, argv[1], sites, 2, nullptr, 0, &n);
    std::vector<double> pop((size_t)(2 * (n > 0 ? n : 1)), 0.0);
    if (rc == EH_OK) rc = eh_dynamics_key_run(ctx, argv[1], sites, 2, pop.data(), n, &n);
    if (rc != EH_OK && rc != EH_ENAN) {
        std::fprintf(stderr, "etoheom_dynamics: %s\n", eh_last_error());
        eh_ctx_destroy(ctx);
        return 1;
    }
    for (int s = 0; s < 2; s++) {
        std::printf("site %d population: \n\n", s + 1);
        for (int64_t t = 0; t < n; t++) std::printf("%g\n", pop[(size_t)(s * n + t)]);
        std::printf("\n");
    }
    eh_ctx_destroy(ctx);
    return 0;
}
