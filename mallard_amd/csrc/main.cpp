// `mallard_amd -i input.toml`: the reference's CLI driver (main.cpp:24-78)
// on the MI355X engine. Returns the reference's integer status.
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/mallard_amd.h"

int main(int argc, char** argv) {
    std::string input;
    for (int i = 1; i < argc; ++i) {
        if ((!std::strcmp(argv[i], "-i") || !std::strcmp(argv[i], "--input")) && i + 1 < argc) input = argv[++i];
    }
    if (input.empty()) {
        std::fprintf(stderr, "usage: %s -i input.toml\n", argv[0]);
        return 1;
    }
    mlr_ctx* ctx = nullptr;
    if (mlr_ctx_create_toml(input.c_str(), &ctx)) {
        std::fprintf(stderr, "init failed: %s\n", mlr_last_error(nullptr));
        return 1;
    }
    int64_t sz[15];
    mlr_get_sizes(ctx, sz);
    std::printf("cells %lld, faces %lld, device layout %.3f GB\n", (long long)sz[0], (long long)sz[2], sz[8] / 1e9);
    const auto t0 = std::chrono::steady_clock::now();
    const int rc = mlr_run(ctx);
    const double el = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    if (rc) {
        std::fprintf(stderr, "run failed: %s\n", mlr_last_error(ctx));
        mlr_ctx_destroy(ctx);
        return 1;
    }
    double t = 0;
    int64_t steps = 0;
    mlr_get_time(ctx, &t, &steps);
    std::printf("Solver finished: step %lld t = %.17g wall %.3f s (%.3e s / step / cell)\n", (long long)steps, t, el,
                steps ? el / steps / sz[0] : 0.0);
    mlr_ctx_destroy(ctx);
    return 0;
}
