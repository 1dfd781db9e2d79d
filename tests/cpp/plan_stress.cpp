// plan_stress.cpp -- drives the C ABI's plan machinery (streams, events, graph
// capture and replay, recompute lists, tuning table, fixed-scratch
// reconstruction) through many create / execute / profile / destroy cycles,
// with several plans of every path alive at once; every call's status is
// checked and the paths' totals are compared.
//
//   plan_stress [rounds]      exit status 0 when every call succeeded and
//                             the paths' totals agree
//
// The workload is synthetic and seeded: a balanced tree of 24 leaves and 600
// families whose counts span size buckets 20..96 (several size classes, so the
// main passes run concurrently on the side streams and the FP32/FP16 passes
// recompute some families in FP64).
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <random>
#include <vector>

#include "cafe_gpu.h"

static int g_fail = 0;

static void ok(int rc, const char* what)
{
    if (rc != CAFE_GPU_OK) {
        fprintf(stderr, "%s failed (%d): %s\n", what, rc, cafe_gpu_last_error());
        g_fail = 1;
        exit(1);
    }
}

int main(int argc, char** argv)
{
    const int rounds = argc > 1 ? atoi(argv[1]) : 3;
    // balanced tree: leaves 0..23, interior nodes pair them up level by level
    const int n_leaves = 24;
    std::vector<int> parent, leaf_col;
    std::vector<int> level;
    for (int i = 0; i < n_leaves; ++i) {
        parent.push_back(-1);
        leaf_col.push_back(i);
        level.push_back(i);
    }
    while (level.size() > 1) {
        std::vector<int> next;
        for (size_t i = 0; i + 1 < level.size(); i += 2) {
            const int u = (int)parent.size();
            parent.push_back(-1);
            leaf_col.push_back(-1);
            parent[level[i]] = u;
            parent[level[i + 1]] = u;
            next.push_back(u);
        }
        if (level.size() % 2) next.push_back(level.back());
        level = next;
    }
    const int n_nodes = (int)parent.size();
    std::vector<double> t(n_nodes, 0.0);
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> ul(5.0, 40.0);
    for (int i = 0; i < n_nodes; ++i)
        if (parent[i] >= 0) t[i] = floor(ul(rng) * 1000.0) / 1000.0;

    // families: counts around a per-family level, bounds from the bucket list
    const int F = 600;
    const int bucket_B[4] = {20, 32, 48, 96};
    std::vector<int> counts((size_t)F * n_leaves), bB(F), bR(F);
    for (int f = 0; f < F; ++f) {
        const int b = bucket_B[f % 4];
        bB[f] = bR[f] = b;
        std::uniform_int_distribution<int> base(0, b * 3 / 4 - 2);
        const int c0 = base(rng);
        for (int l = 0; l < n_leaves; ++l) {
            std::uniform_int_distribution<int> d(-2, 2);
            int c = c0 + d(rng);
            counts[(size_t)f * n_leaves + l] = c < 0 ? 0 : (c > b ? b : c);
        }
    }

    cafe_gpu_ctx* ctx = nullptr;
    ok(cafe_gpu_create(0, &ctx), "create");
    ok(cafe_gpu_set_tree(ctx, n_nodes, parent.data(), leaf_col.data()), "set_tree");
    ok(cafe_gpu_set_families(ctx, F, n_leaves, counts.data(), bB.data(), bR.data(), 0, 0), "set_families");
    int nb = 0;
    ok(cafe_gpu_bucket_count(ctx, &nb), "bucket_count");

    // T lambda points, one key per (point, distinct branch length)
    const int T = 6;
    std::vector<double> lam_q, t_q;
    std::vector<int> slot((size_t)T * n_nodes, 0);
    for (int p = 0; p < T; ++p) {
        const double lam = floor((0.0005 + 0.002 * p) * 1e9) / 1e9;
        for (int i = 0; i < n_nodes; ++i) {
            if (parent[i] < 0) continue;
            slot[(size_t)p * n_nodes + i] = (int)lam_q.size();
            lam_q.push_back(lam);
            t_q.push_back(t[i]);
        }
    }
    const int n_keys = (int)lam_q.size();
    ok(cafe_gpu_matrix_reset(ctx, 97, n_keys), "matrix_reset");
    const int stride = 96;
    std::vector<double> prior((size_t)nb * stride, 0.0);
    for (int b = 0; b < nb; ++b) {
        int B = 0, R = 0, cnt = 0;
        ok(cafe_gpu_bucket_info(ctx, b, &B, &R, &cnt), "bucket_info");
        for (int j = 0; j < R; ++j) prior[(size_t)b * stride + j] = (double)(float)(1.0 / R);
    }
    cafe_gpu_eval_desc d{};
    d.model = CAFE_GPU_MODEL_BASE;
    d.n_trials = T;
    d.n_cat = 1;
    d.node_slot = slot.data();
    d.prior = prior.data();
    d.prior_stride = stride;

    std::vector<double> totals[4];
    for (int r = 0; r < rounds; ++r) {
        // one plan per path, all alive together, executed interleaved
        cafe_gpu_plan* plans[4] = {};
        for (int path = 0; path < 4; ++path) {
            d.path = path;
            ok(cafe_gpu_plan_create(ctx, &d, n_keys, lam_q.data(), t_q.data(), CAFE_GPU_BUILD_RECURRENCE, 1,
                                    &plans[path]),
               "plan_create");
        }
        for (int rep = 0; rep < 3; ++rep)
            for (int path = 3; path >= 0; --path) ok(cafe_gpu_plan_execute(plans[path]), "plan_execute");
        for (int path = 0; path < 4; ++path) {
            std::vector<double> tot(T), lnl((size_t)T * F);
            long long fb = 0;
            ok(cafe_gpu_plan_results(plans[path], lnl.data(), nullptr, nullptr, tot.data(), &fb), "plan_results");
            double ms[4];
            int graph = 0;
            ok(cafe_gpu_plan_profile(plans[path], ms, &graph), "plan_profile");
            int ncls = 0;
            ok(cafe_gpu_plan_class_times(plans[path], 0, &ncls, nullptr, nullptr, nullptr, nullptr), "class_times");
            std::vector<int> nt(ncls), launch(6 * ncls);
            std::vector<long long> tiles(ncls);
            std::vector<double> cms(ncls);
            ok(cafe_gpu_plan_class_times(plans[path], ncls, &ncls, nt.data(), tiles.data(), cms.data(), launch.data()),
               "class_times");
            long long r64 = 0, rp = 0;
            ok(cafe_gpu_plan_recompute_counts(plans[path], &r64, &rp), "recompute_counts");
            if (r == 0) totals[path] = tot;
            printf("round %d path %d: graph=%d classes=%d total[0]=%.9f recomputed=%lld/%lld\n", r, path, graph,
                   ncls, tot[0], r64, rp);
        }
        // new keys for one plan, then destroy in a different order each round
        std::vector<double> lq2(lam_q);
        for (double& x : lq2) x = floor(x * 1.1 * 1e9) / 1e9;
        ok(cafe_gpu_plan_set_keys(plans[r % 4], lq2.data(), t_q.data()), "plan_set_keys");
        ok(cafe_gpu_plan_execute(plans[r % 4]), "plan_execute");
        ok(cafe_gpu_plan_synchronize(plans[r % 4]), "plan_synchronize");
        for (int k = 0; k < 4; ++k) ok(cafe_gpu_plan_destroy(plans[(k + r) % 4]), "plan_destroy");

        // a tuning entry per round, then eager evaluations on every path
        ok(cafe_gpu_tuning_set(ctx, 2, 2, n_leaves, 0, 2, 1, 1, 4 + r), "tuning_set");
        for (int path = 0; path < 4; ++path) {
            d.path = path;
            std::vector<double> tot(T);
            long long fb = 0;
            ok(cafe_gpu_matrix_build(ctx, n_keys, lam_q.data(), t_q.data(), 0, CAFE_GPU_BUILD_RECURRENCE),
               "matrix_build");
            ok(cafe_gpu_eval(ctx, &d, nullptr, nullptr, nullptr, tot.data(), &fb), "eval");
        }
        ok(cafe_gpu_tuning_clear(ctx), "tuning_clear");

        // reconstruction on both scratch kinds (all families at the first point)
        std::vector<double> rp((size_t)97, 1.0 / 96);
        std::vector<int> s1((size_t)F * n_nodes), s2((size_t)F * n_nodes);
        ok(cafe_gpu_set_families(ctx, F, n_leaves, counts.data(), nullptr, nullptr, 96, 96), "set_families");
        ok(cafe_gpu_reconstruct(ctx, slot.data(), rp.data(), 96, 96, 0, F, s1.data()), "reconstruct");
        ok(cafe_gpu_reconstruct_noalloc(ctx, slot.data(), rp.data(), 96, 96, 0, F, s2.data()), "reconstruct_noalloc");
        if (memcmp(s1.data(), s2.data(), sizeof(int) * s1.size()) != 0) {
            fprintf(stderr, "reconstructions differ\n");
            g_fail = 1;
        }
        ok(cafe_gpu_set_families(ctx, F, n_leaves, counts.data(), bB.data(), bR.data(), 0, 0), "set_families");
        nb = 0;
        ok(cafe_gpu_bucket_count(ctx, &nb), "bucket_count");
    }
    for (int path = 1; path < 4; ++path)
        for (int p = 0; p < T; ++p) {
            const double a = totals[path][p], b = totals[0][p];
            if (!(fabs(a - b) <= 1e-6 * fabs(b))) {
                fprintf(stderr, "path %d point %d: total %.12g vs plain %.12g\n", path, p, a, b);
                g_fail = 1;
            }
        }
    ok(cafe_gpu_destroy(ctx), "destroy");
    printf(g_fail ? "FAILED\n" : "plan_stress: ok\n");
    return g_fail;
}
