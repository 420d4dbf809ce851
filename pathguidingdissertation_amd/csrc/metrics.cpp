// metrics.cpp — image error metrics on the host (Metrics.h:8-69 plus relMSE).
//
// MSE, RMSE and RRMSE keep the reference's FP32 arithmetic and sequential summation
// order, so the values match Metrics::calculate* bit for bit; relMSE, the
// convergence measure of the guiding study, is accumulated in FP64.
#include <cmath>
#include <cstdint>

#include "../../include/pg_tracer.h"

extern "C" pg_status pg_metrics(const float* test, const float* truth, uint64_t n_test, uint64_t n_truth,
                                int32_t width, int32_t height, pg_image_metrics* out) {
    if (!test || !truth || !out) return PG_ERR_INVALID;
    if (n_test != n_truth) {  // Metrics.h:10-13 returns -1
        out->mse = out->rmse = out->rrmse = -1.f;
        out->relmse = -1.0;
        return PG_ERR_INVALID;
    }
    if (width <= 0 || height <= 0 || (uint64_t)width * (uint64_t)height * 3u > n_test) return PG_ERR_INVALID;
    const int64_t n = (int64_t)width * height * 3;
    float sse = 0.f, sq = 0.f;
    double rel = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        float y = test[i], yh = truth[i];
        float d = y - yh;
        sq += yh * yh;
        sse += d * d;
        double dd = (double)y - (double)yh;
        rel += dd * dd / ((double)yh * (double)yh + 1e-2);
    }
    out->mse = sse / (float)n;
    out->rmse = std::sqrt(sse / (float)n);
    out->rrmse = out->rmse / (std::sqrt(sq) + 1e-2f);
    out->relmse = rel / (double)n;
    return PG_OK;
}
