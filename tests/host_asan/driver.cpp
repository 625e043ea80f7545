// Drives the library's host paths (plans, single-bandwidth complex64 and FP64 transforms,
// batched correlation and peak search, grouped transforms with their tuner, graph capture,
// schedule timing and graph cache) against the stand-in runtime of mock_runtime.cpp, under
// AddressSanitizer.  Exit status 0 and a launch count on the last line mean no report.
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "so3ft.h"

extern "C" long mock_launches();

static void chk(int rc, const char *what)
{
    if (rc) {
        fprintf(stderr, "%s: rc %d: %s\n", what, rc, so3ft_last_error());
        exit(1);
    }
}

static void run_plan(int64_t bw)
{
    so3ft_handle h = 0;
    chk(so3ft_init(bw, bw - 1, 0, 1, 0, 0, &h), "so3ft_init");
    const int64_t N = 2 * bw, N3 = N * N * N, nc = (4 * bw * bw * bw - bw) / 3, nb = bw * bw;
    const int n = 3;
    std::vector<float> c32(2 * nc * n), f32(2 * N3 * n);
    std::vector<double> c64(2 * nc * n), f64(2 * N3 * n), fl(2 * nb * n), gl(2 * nb * n), val(n);
    std::vector<int64_t> idx(n);
    chk(so3ft_isoft_many_c64(h, c32.data(), f32.data(), n, nullptr), "isoft_many_c64");
    chk(so3ft_soft_many_c64(h, f32.data(), c32.data(), n, nullptr), "soft_many_c64");
    chk(so3ft_isoft(h, c64.data(), f64.data(), nullptr), "isoft");
    chk(so3ft_soft(h, f64.data(), c64.data(), nullptr), "soft");
    chk(so3ft_cross_correlation_ylm_cmplx_many(h, fl.data(), gl.data(), f64.data(), n, nullptr), "correlation_many");
    chk(so3ft_cross_correlation_argmax_many(h, fl.data(), gl.data(), n, idx.data(), val.data(), nullptr),
        "correlation_argmax_many");
    chk(so3ft_destroy(h), "so3ft_destroy");
    printf("plan bw=%lld ok\n", (long long)bw);
}

// calls isoft + soft `calls` times; fresh: new output buffers every call (graph cache
// eviction and update)
static void run_group(std::vector<int64_t> bws, int calls, bool fresh)
{
    const int64_t n = (int64_t)bws.size();
    int64_t ct = 0, gt = 0;
    std::vector<int64_t> co(n), go(n);
    chk(so3ft_group_layout(bws.data(), n, co.data(), go.data(), &ct, &gt), "group_layout");
    int64_t h = 0;
    chk(so3ft_group_create(bws.data(), n, &h), "group_create");
    std::vector<float> c(2 * ct), f(2 * gt), c2(2 * ct);
    std::vector<std::vector<float>> keep;
    for (int r = 0; r < calls; ++r) {
        float *ff = f.data(), *cc = c2.data();
        if (fresh) {
            keep.emplace_back(2 * gt);
            ff = keep.back().data();
            keep.emplace_back(2 * ct);
            cc = keep.back().data();
        }
        chk(so3ft_group_isoft_c64(h, c.data(), ff, nullptr), "group_isoft_c64");
        chk(so3ft_group_soft_c64(h, ff, cc, nullptr), "group_soft_c64");
    }
    chk(so3ft_group_destroy(h), "group_destroy");
    printf("group of %lld ok\n", (long long)n);
}

int main()
{
    for (int64_t bw : {1, 2, 5, 16, 32, 40, 55, 64}) run_plan(bw);
    run_group({8, 16, 6, 12, 8, 4, 16, 24, 2, 8, 3}, 3, false);
    run_group({8, 16, 6, 12, 8, 4, 16, 24, 2, 8, 3}, 12, true);
    run_group({32, 40, 5, 55, 40}, 3, false);
    run_group({32, 40, 5, 55, 40}, 12, true);
    printf("launches %ld\n", mock_launches());
    return 0;
}
