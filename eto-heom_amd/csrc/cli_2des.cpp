// etoheom_2des <file.key> -- drop-in for the reference CPU_2DES executable
// (etoheom-cpu/src/CPU_2DES.cpp:14-96): same argument, same stdout, one line
// "%.6e %.6e" (scientific, precision 6) per readout point.  The sample loop runs
// all 3*n_sample propagations as one batch on GPU 0 (override with ETOHEOM_DEVICE).
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "etoheom.h"

int main(int argc, char** argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s <file.key>\n", argv[0]);
        return 2;
    }
    const char* dev = std::getenv("ETOHEOM_DEVICE");
    eh_ctx* ctx = nullptr;
    int rc = eh_ctx_create(dev ? std::atoi(dev) : 0, &ctx);
    if (rc) {
        std::fprintf(stderr, "etoheom_2des: %s\n", eh_last_error());
        return 1;
    }
    int64_t n = 0;
    rc = eh_2des_key_run(ctx, argv[1], nullptr, 0, &n);
    std::vector<double> P((size_t)(2 * (n > 0 ? n : 1)));
    if (rc == EH_OK) rc = eh_2des_key_run(ctx, argv[1], P.data(), n, &n);
    if (rc != EH_OK && rc != EH_ENAN) {
        std::fprintf(stderr, "etoheom_2des: %s\n", eh_last_error());
        eh_ctx_destroy(ctx);
        return 1;
    }
    for (int64_t i = 0; i < n; i++) std::printf("%.6e %.6e\n", P[(size_t)2 * i], P[(size_t)2 * i + 1]);
    eh_ctx_destroy(ctx);
    return 0;
}
