// api.hip -- host side of libcafe_gpu.so: context, planning and the C ABI of
// include/cafe_gpu.h.
#include <math.h>
#include <stdarg.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <numeric>
#include <string>
#include <vector>

#include "cafe_gpu.h"
#include "engine.h"

namespace cafe {
hipError_t launch_build(int algo, const double* d_lam, const double* d_t, const double* d_lg, int n,
                        int first_slot, int N, int Np, long long slot_elems, double* mt64, float* mt32,
                        hipStream_t stream);
hipError_t launch_fill_zero(void* p, long long bytes, hipStream_t stream);
hipError_t launch_prune(int path, int nt, int grid, const PruneParams& p, hipStream_t st);
hipError_t launch_compact(const TileDesc* tiles, const int* n_tiles_dev, int max_tiles,
                          const BucketDev* buckets, const int* perm, const int* flags, int F, int mask,
                          TileDesc* out_tiles, int* out_count, int* fam_count, hipStream_t st);
hipError_t launch_finalize(int model, int T, int K, int F, const double* val, const int* flag,
                           const double* cat_probs, double* lnl, double* cat, double* post, int* fail,
                           double* total, hipStream_t st);
}  // namespace cafe

using namespace cafe;

namespace {

thread_local std::string g_last_error;

int fail(int code, const char* fmt, ...)
{
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    g_last_error = buf;
    return code;
}

#define HIP_OR_FAIL(expr)                                                                          \
    do {                                                                                           \
        hipError_t _e = (expr);                                                                    \
        if (_e != hipSuccess)                                                                      \
            return fail(_e == hipErrorOutOfMemory ? CAFE_GPU_ENOMEM : CAFE_GPU_EDEVICE, "%s: %s", \
                        #expr, hipGetErrorString(_e));                                             \
    } while (0)

int round_up(int x, int m) { return (x + m - 1) / m * m; }

// Grow-only device buffer.
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    hipError_t ensure(size_t need)
    {
        if (need <= bytes) return hipSuccess;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        hipError_t e = hipMalloc(&p, need);
        if (e == hipSuccess) bytes = need;
        return e;
    }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <typename T> T* as() const { return (T*)p; }
};

struct BucketHost {
    int B, R, count, fam_begin, max_count;
};

}  // namespace

struct cafe_gpu_ctx {
    int device = 0;
    int n_cu = 256;
    hipStream_t stream = nullptr;
    hipEvent_t ev[6] = {};
    double timings[4] = {0, 0, 0, 0};
    bool build_timed = false;
    DevBuf lg;

    // matrix pool
    int N = 0, Np = 0, capacity = 0;
    long long slot_elems = 0;
    DevBuf mt64, mt32, key_lam, key_t;
    std::vector<char> built;

    // tree program
    int n_nodes = 0, tree_leaves = 0, n_slots = 0;
    std::vector<OpDev> ops;
    std::vector<ChildDev> kids;
    DevBuf d_ops, d_kids;

    // families
    int F = 0, n_leaves = 0;
    std::vector<int> perm;
    std::vector<BucketHost> buckets;
    DevBuf d_counts, d_perm;

    // error model
    int n_dev = 0, err_rows = 0;
    DevBuf d_err64, d_err32;

    // eval buffers
    DevBuf d_node_slot, d_prior64, d_prior32, d_catp, d_buckets, d_tiles, d_ntiles, d_counters;
    DevBuf d_fb_tiles, d_fb_count, d_val, d_flag, d_lnl, d_cat, d_post, d_fail, d_total, d_scratch,
        d_root;
};

namespace {

int check_ctx(cafe_gpu_ctx* ctx)
{
    if (!ctx) return fail(CAFE_GPU_EINVAL, "null context");
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(CAFE_GPU_EDEVICE, "hipSetDevice failed");
    return CAFE_GPU_OK;
}

// Sethi-Ullman order: evaluate the interior child that needs the most live
// partial vectors first, so at most ~log2(n) + 2 vectors are live at once.
struct Planner {
    const std::vector<std::vector<int>>& kids;
    const std::vector<int>& leaf_col;
    std::vector<int> need, slot_of;
    std::vector<char> used;
    std::vector<OpDev> ops;
    std::vector<ChildDev> out_kids;
    int max_slot = -1;

    Planner(const std::vector<std::vector<int>>& k, const std::vector<int>& lc)
        : kids(k), leaf_col(lc), need(k.size(), -1), slot_of(k.size(), -1) {}

    std::vector<int> interior_order(int u)
    {
        std::vector<int> in;
        for (int c : kids[u])
            if (leaf_col[c] < 0) in.push_back(c);
        std::stable_sort(in.begin(), in.end(), [&](int a, int b) { return need[a] > need[b]; });
        return in;
    }

    int compute_need(int u)
    {
        if (leaf_col[u] >= 0) return need[u] = 0;
        for (int c : kids[u]) compute_need(c);
        auto in = interior_order(u);
        int n = (int)in.size() + 1;   // children held + a fresh output slot
        for (size_t i = 0; i < in.size(); ++i) n = std::max(n, (int)i + need[in[i]]);
        return need[u] = n;
    }

    int alloc()
    {
        for (size_t i = 0; i < used.size(); ++i)
            if (!used[i]) {
                used[i] = 1;
                max_slot = std::max(max_slot, (int)i);
                return (int)i;
            }
        used.push_back(1);
        max_slot = std::max(max_slot, (int)used.size() - 1);
        return (int)used.size() - 1;
    }

    void emit(int u)
    {
        auto in = interior_order(u);
        for (int c : in) emit(c);
        const int out = alloc();
        OpDev op;
        op.node = u;
        op.out_slot = out;
        op.n_child = (int)kids[u].size();
        op.child_begin = (int)out_kids.size();
        for (int c : kids[u]) {   // product in descendant order
            ChildDev ch;
            ch.node = c;
            ch.leaf_col = leaf_col[c];
            ch.in_slot = leaf_col[c] >= 0 ? -1 : slot_of[c];
            ch.pad = 0;
            out_kids.push_back(ch);
        }
        ops.push_back(op);
        for (int c : in) used[slot_of[c]] = 0;
        slot_of[u] = out;
    }
};

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* cafe_gpu_last_error(void) { return g_last_error.c_str(); }

int cafe_gpu_create(int device, cafe_gpu_ctx** out)
{
    if (!out) return fail(CAFE_GPU_EINVAL, "null output pointer");
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= 0)
        return fail(CAFE_GPU_EDEVICE, "no HIP device available");
    if (device < 0 || device >= n) return fail(CAFE_GPU_EINVAL, "device %d out of range", device);
    HIP_OR_FAIL(hipSetDevice(device));
    auto* ctx = new cafe_gpu_ctx();
    ctx->device = device;
    int cu = 0;
    if (hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, device) == hipSuccess && cu > 0)
        ctx->n_cu = cu;
    if (hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess) {
        delete ctx;
        return fail(CAFE_GPU_EDEVICE, "hipStreamCreate failed");
    }
    for (auto& e : ctx->ev) (void)hipEventCreate(&e);
    // log-gamma of integers from the host libm: the reference's
    // lgamma cache (probability.cpp:69-75)
    std::vector<double> lg(kLgammaN);
    for (int i = 0; i < kLgammaN; ++i) lg[i] = lgamma((double)i);
    if (ctx->lg.ensure(sizeof(double) * kLgammaN) != hipSuccess ||
        hipMemcpy(ctx->lg.p, lg.data(), sizeof(double) * kLgammaN, hipMemcpyHostToDevice) != hipSuccess) {
        cafe_gpu_destroy(ctx);
        return fail(CAFE_GPU_ENOMEM, "log-gamma table upload failed");
    }
    *out = ctx;
    return CAFE_GPU_OK;
}

int cafe_gpu_destroy(cafe_gpu_ctx* ctx)
{
    if (!ctx) return CAFE_GPU_OK;
    (void)hipSetDevice(ctx->device);
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    DevBuf* bufs[] = {&ctx->lg, &ctx->mt64, &ctx->mt32, &ctx->key_lam, &ctx->key_t, &ctx->d_ops,
                      &ctx->d_kids, &ctx->d_counts, &ctx->d_perm, &ctx->d_err64, &ctx->d_err32,
                      &ctx->d_node_slot, &ctx->d_prior64, &ctx->d_prior32, &ctx->d_catp, &ctx->d_buckets,
                      &ctx->d_tiles, &ctx->d_ntiles, &ctx->d_counters, &ctx->d_fb_tiles, &ctx->d_fb_count,
                      &ctx->d_val, &ctx->d_flag, &ctx->d_lnl, &ctx->d_cat, &ctx->d_post, &ctx->d_fail,
                      &ctx->d_total, &ctx->d_scratch, &ctx->d_root};
    for (DevBuf* b : bufs) b->release();
    for (auto& e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return CAFE_GPU_OK;
}

int cafe_gpu_matrix_reset(cafe_gpu_ctx* ctx, int N, int capacity)
{
    if (int rc = check_ctx(ctx)) return rc;
    if (N < 2 || N * 2 + 2 > kLgammaN) return fail(CAFE_GPU_EINVAL, "matrix size %d out of range", N);
    if (capacity < 1) return fail(CAFE_GPU_EINVAL, "capacity must be positive");
    ctx->N = N;
    ctx->Np = round_up(N, kTile) + kTile;
    ctx->slot_elems = (long long)ctx->Np * ctx->Np;
    ctx->capacity = capacity;
    const size_t n64 = sizeof(double) * ctx->slot_elems * capacity;
    const size_t n32 = sizeof(float) * ctx->slot_elems * capacity;
    HIP_OR_FAIL(ctx->mt64.ensure(n64));
    HIP_OR_FAIL(ctx->mt32.ensure(n32));
    // padding rows/columns must read as zero
    HIP_OR_FAIL(launch_fill_zero(ctx->mt64.p, (long long)n64, ctx->stream));
    HIP_OR_FAIL(launch_fill_zero(ctx->mt32.p, (long long)n32, ctx->stream));
    HIP_OR_FAIL(hipStreamSynchronize(ctx->stream));
    ctx->built.assign(capacity, 0);
    return CAFE_GPU_OK;
}

int cafe_gpu_matrix_build(cafe_gpu_ctx* ctx, int n, const double* lambda_q, const double* t_q,
                          int first_slot, int algo)
{
    if (int rc = check_ctx(ctx)) return rc;
    if (ctx->N == 0) return fail(CAFE_GPU_ESTATE, "matrix_reset has not been called");
    if (n < 0 || first_slot < 0 || first_slot + n > ctx->capacity)
        return fail(CAFE_GPU_EINVAL, "slots %d..%d exceed capacity %d", first_slot, first_slot + n,
                    ctx->capacity);
    if (algo != CAFE_GPU_BUILD_CLOSED_FORM && algo != CAFE_GPU_BUILD_RECURRENCE)
        return fail(CAFE_GPU_EINVAL, "unknown build algorithm %d", algo);
    if (n == 0) return CAFE_GPU_OK;
    if (!lambda_q || !t_q) return fail(CAFE_GPU_EINVAL, "null key arrays");
    for (int i = 0; i < n; ++i)
        if (!(lambda_q[i] >= 0) || !(t_q[i] >= 0) || !std::isfinite(lambda_q[i]) || !std::isfinite(t_q[i]))
            return fail(CAFE_GPU_EINVAL, "key %d has an invalid lambda or branch length", i);
    HIP_OR_FAIL(ctx->key_lam.ensure(sizeof(double) * n));
    HIP_OR_FAIL(ctx->key_t.ensure(sizeof(double) * n));
    HIP_OR_FAIL(hipMemcpyAsync(ctx->key_lam.p, lambda_q, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
    HIP_OR_FAIL(hipMemcpyAsync(ctx->key_t.p, t_q, sizeof(double) * n, hipMemcpyHostToDevice, ctx->stream));
    HIP_OR_FAIL(hipEventRecord(ctx->ev[4], ctx->stream));
    HIP_OR_FAIL(launch_build(algo, ctx->key_lam.as<double>(), ctx->key_t.as<double>(), ctx->lg.as<double>(), n,
                             first_slot, ctx->N, ctx->Np, ctx->slot_elems, ctx->mt64.as<double>(),
                             ctx->mt32.as<float>(), ctx->stream));
    HIP_OR_FAIL(hipEventRecord(ctx->ev[5], ctx->stream));
    HIP_OR_FAIL(hipStreamSynchronize(ctx->stream));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ctx->ev[4], ctx->ev[5]);
    ctx->timings[0] = ms;
    for (int i = 0; i < n; ++i) ctx->built[first_slot + i] = 1;
    return CAFE_GPU_OK;
}

int cafe_gpu_matrix_get(cafe_gpu_ctx* ctx, int slot, double* out)
{
    if (int rc = check_ctx(ctx)) return rc;
    if (!out) return fail(CAFE_GPU_EINVAL, "null output");
    if (slot < 0 || slot >= ctx->capacity || !ctx->built[slot])
        return fail(CAFE_GPU_ENOTFOUND, "matrix slot %d has not been built", slot);
    const int N = ctx->N, Np = ctx->Np;
    std::vector<double> cm((size_t)N * Np);
    HIP_OR_FAIL(hipMemcpy(cm.data(), ctx->mt64.as<double>() + (long long)slot * ctx->slot_elems,
                          sizeof(double) * cm.size(), hipMemcpyDeviceToHost));
    for (int s = 0; s < N; ++s)
        for (int c = 0; c < N; ++c) out[(size_t)s * N + c] = cm[(size_t)c * Np + s];
    return CAFE_GPU_OK;
}

int cafe_gpu_set_tree(cafe_gpu_ctx* ctx, int n_nodes, const int* parent, const int* leaf_col)
{
    if (int rc = check_ctx(ctx)) return rc;
    if (n_nodes < 2 || !parent || !leaf_col) return fail(CAFE_GPU_EINVAL, "tree needs >= 2 nodes");
    std::vector<std::vector<int>> kids(n_nodes);
    std::vector<int> lc(leaf_col, leaf_col + n_nodes);
    int root = -1, n_leaves = 0;
    for (int i = 0; i < n_nodes; ++i) {
        if (parent[i] < 0) {
            if (root >= 0) return fail(CAFE_GPU_EINVAL, "tree has more than one root");
            root = i;
        } else if (parent[i] >= n_nodes || parent[i] == i) {
            return fail(CAFE_GPU_EINVAL, "node %d has an invalid parent", i);
        } else {
            kids[parent[i]].push_back(i);
        }
        if (lc[i] >= 0) ++n_leaves;
    }
    if (root < 0) return fail(CAFE_GPU_EINVAL, "tree has no root");
    for (int i = 0; i < n_nodes; ++i) {
        if (lc[i] >= 0 && !kids[i].empty()) return fail(CAFE_GPU_EINVAL, "leaf %d has children", i);
        if (lc[i] < 0 && kids[i].empty()) return fail(CAFE_GPU_EINVAL, "interior node %d has no children", i);
        if (lc[i] >= n_leaves) return fail(CAFE_GPU_EINVAL, "leaf column %d out of range", lc[i]);
        int d = 0;
        for (int p = i; p != root; p = parent[p])
            if (++d > n_nodes) return fail(CAFE_GPU_EINVAL, "tree has a cycle");
    }
    if (lc[root] >= 0) return fail(CAFE_GPU_EINVAL, "root cannot be a leaf");
    std::vector<char> seen(n_leaves, 0);
    for (int i = 0; i < n_nodes; ++i)
        if (lc[i] >= 0) {
            if (seen[lc[i]]) return fail(CAFE_GPU_EINVAL, "leaf column %d used twice", lc[i]);
            seen[lc[i]] = 1;
        }
    Planner pl(kids, lc);
    pl.compute_need(root);
    pl.used.clear();
    pl.emit(root);
    if (pl.max_slot + 1 > kMaxSlots)
        return fail(CAFE_GPU_EINVAL, "tree needs %d live partial vectors (max %d)", pl.max_slot + 1, kMaxSlots);
    ctx->n_nodes = n_nodes;
    ctx->tree_leaves = n_leaves;
    ctx->n_slots = pl.max_slot + 1;
    ctx->ops = pl.ops;
    ctx->kids = pl.out_kids;
    HIP_OR_FAIL(ctx->d_ops.ensure(sizeof(OpDev) * ctx->ops.size()));
    HIP_OR_FAIL(ctx->d_kids.ensure(sizeof(ChildDev) * ctx->kids.size()));
    HIP_OR_FAIL(hipMemcpy(ctx->d_ops.p, ctx->ops.data(), sizeof(OpDev) * ctx->ops.size(), hipMemcpyHostToDevice));
    HIP_OR_FAIL(hipMemcpy(ctx->d_kids.p, ctx->kids.data(), sizeof(ChildDev) * ctx->kids.size(),
                          hipMemcpyHostToDevice));
    return CAFE_GPU_OK;
}

int cafe_gpu_set_families(cafe_gpu_ctx* ctx, int F, int n_leaves, const int* counts, const int* bound_B,
                          const int* bound_R, int B, int R)
{
    if (int rc = check_ctx(ctx)) return rc;
    if (F < 1 || n_leaves < 1 || !counts) return fail(CAFE_GPU_EINVAL, "need at least one family and leaf");
    std::vector<int> fb(F), fr(F);
    for (int f = 0; f < F; ++f) {
        fb[f] = bound_B ? bound_B[f] : B;
        fr[f] = bound_R ? bound_R[f] : R;
        if (fb[f] < 1 || fr[f] < 1) return fail(CAFE_GPU_EINVAL, "family %d has a non-positive bound", f);
        for (int i = 0; i < n_leaves; ++i) {
            const int c = counts[(size_t)f * n_leaves + i];
            if (c < 0 || c > fb[f])
                return fail(CAFE_GPU_EINVAL, "family %d leaf %d count %d outside 0..%d", f, i, c, fb[f]);
        }
    }
    std::vector<int> order(F);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
        return fb[a] != fb[b] ? fb[a] < fb[b] : fr[a] < fr[b];
    });
    ctx->buckets.clear();
    for (int k = 0; k < F; ++k) {
        const int f = order[k];
        if (ctx->buckets.empty() || ctx->buckets.back().B != fb[f] || ctx->buckets.back().R != fr[f])
            ctx->buckets.push_back(BucketHost{fb[f], fr[f], 0, k, 0});
        auto& b = ctx->buckets.back();
        b.count++;
        for (int i = 0; i < n_leaves; ++i) b.max_count = std::max(b.max_count, counts[(size_t)f * n_leaves + i]);
    }
    std::vector<int> cT((size_t)n_leaves * F);
    for (int k = 0; k < F; ++k)
        for (int i = 0; i < n_leaves; ++i) cT[(size_t)i * F + k] = counts[(size_t)order[k] * n_leaves + i];
    HIP_OR_FAIL(ctx->d_counts.ensure(sizeof(int) * cT.size()));
    HIP_OR_FAIL(ctx->d_perm.ensure(sizeof(int) * F));
    HIP_OR_FAIL(hipMemcpy(ctx->d_counts.p, cT.data(), sizeof(int) * cT.size(), hipMemcpyHostToDevice));
    HIP_OR_FAIL(hipMemcpy(ctx->d_perm.p, order.data(), sizeof(int) * F, hipMemcpyHostToDevice));
    ctx->perm = order;
    ctx->F = F;
    ctx->n_leaves = n_leaves;
    return CAFE_GPU_OK;
}

int cafe_gpu_bucket_count(cafe_gpu_ctx* ctx, int* n)
{
    if (!ctx || !n) return fail(CAFE_GPU_EINVAL, "null argument");
    *n = (int)ctx->buckets.size();
    return CAFE_GPU_OK;
}

int cafe_gpu_bucket_info(cafe_gpu_ctx* ctx, int i, int* B, int* R, int* count)
{
    if (!ctx) return fail(CAFE_GPU_EINVAL, "null context");
    if (i < 0 || i >= (int)ctx->buckets.size()) return fail(CAFE_GPU_EINVAL, "bucket %d out of range", i);
    if (B) *B = ctx->buckets[i].B;
    if (R) *R = ctx->buckets[i].R;
    if (count) *count = ctx->buckets[i].count;
    return CAFE_GPU_OK;
}

int cafe_gpu_set_error_model(cafe_gpu_ctx* ctx, int n_dev, int n_rows, const double* rows)
{
    if (int rc = check_ctx(ctx)) return rc;
    if (!rows) {
        ctx->n_dev = 0;
        ctx->err_rows = 0;
        return CAFE_GPU_OK;
    }
    if (n_dev < 1 || n_dev % 2 == 0 || n_rows < 1)
        return fail(CAFE_GPU_EINVAL, "error model needs an odd number of deviations and >= 1 row");
    const size_t n = (size_t)n_dev * n_rows;
    std::vector<float> r32(rows, rows + n);
    HIP_OR_FAIL(ctx->d_err64.ensure(sizeof(double) * n));
    HIP_OR_FAIL(ctx->d_err32.ensure(sizeof(float) * n));
    HIP_OR_FAIL(hipMemcpy(ctx->d_err64.p, rows, sizeof(double) * n, hipMemcpyHostToDevice));
    HIP_OR_FAIL(hipMemcpy(ctx->d_err32.p, r32.data(), sizeof(float) * n, hipMemcpyHostToDevice));
    ctx->n_dev = n_dev;
    ctx->err_rows = n_rows;
    return CAFE_GPU_OK;
}

}  // extern "C"

namespace {

struct KernelClass {
    int nt;       // template size class
    int begin, end;
};

// Builds the work list for `points` parameter points and runs the main
// pruning path plus the FP64 / plain recomputation passes.  Results land in
// d_val / d_flag ([point][F]).
int run_prune(cafe_gpu_ctx* ctx, int path, int model, int n_points, const int* point_ids,
              double* root_out, int root_stride, long long* n_fallback)
{
    const int F = ctx->F;
    // buckets in descending work order
    std::vector<int> border(ctx->buckets.size());
    std::iota(border.begin(), border.end(), 0);
    std::stable_sort(border.begin(), border.end(), [&](int a, int b) {
        return std::max(ctx->buckets[a].B + 1, ctx->buckets[a].R) > std::max(ctx->buckets[b].B + 1, ctx->buckets[b].R);
    });
    int maxLR = 0;
    for (auto& b : ctx->buckets) maxLR = std::max(maxLR, std::max(round_up(b.B + 1, kTile), round_up(b.R, kTile)));

    auto class_of = [&](const BucketHost& b) {
        const int lp = round_up(b.B + 1, kTile), rp = round_up(b.R, kTile);
        if (path == CAFE_GPU_PATH_PLAIN64) return (std::max(b.B + 1, b.R) + 63) / 64;
        const int nt = std::max(lp, rp) / kTile;
        static const int scaled_sizes[] = {2, 4, 8, 13, 16, 24, 32};
        static const int f64_sizes[] = {2, 4, 8};
        if (path == CAFE_GPU_PATH_BATCHED32) {
            for (int s : scaled_sizes)
                if (nt <= s) return s;
            return -1;
        }
        for (int s : f64_sizes)
            if (nt <= s) return s;
        return 8;
    };
    // plain kernel row classes must be powers of two up to 16
    auto plain_class = [](int nq) { int c = 1; while (c < nq) c <<= 1; return c; };

    std::vector<TileDesc> tiles;
    std::vector<KernelClass> classes;
    std::vector<std::pair<int, int>> bucket_class;   // (class, bucket)
    for (int bi : border) {
        int c = class_of(ctx->buckets[bi]);
        if (path == CAFE_GPU_PATH_PLAIN64) c = plain_class(c);
        if (c < 0 || (path == CAFE_GPU_PATH_PLAIN64 && c > 16))
            return fail(CAFE_GPU_EINVAL, "bucket B=%d R=%d exceeds the supported size", ctx->buckets[bi].B,
                        ctx->buckets[bi].R);
        bucket_class.push_back({c, bi});
    }
    std::stable_sort(bucket_class.begin(), bucket_class.end(),
                     [](const std::pair<int, int>& a, const std::pair<int, int>& b) { return a.first > b.first; });
    for (auto& bc : bucket_class) {
        const BucketHost& b = ctx->buckets[bc.second];
        if (classes.empty() || classes.back().nt != bc.first)
            classes.push_back(KernelClass{bc.first, (int)tiles.size(), (int)tiles.size()});
        for (int pi = 0; pi < n_points; ++pi)
            for (int f0 = 0; f0 < b.count; f0 += kTile)
                tiles.push_back(TileDesc{bc.second, point_ids[pi], f0, std::min(kTile, b.count - f0)});
        classes.back().end = (int)tiles.size();
    }
    const int n_tiles = (int)tiles.size();
    const int n_cls = (int)classes.size();
    std::vector<int> counts(n_cls + 4, 0);
    for (int i = 0; i < n_cls; ++i) counts[i] = classes[i].end - classes[i].begin;

    HIP_OR_FAIL(ctx->d_tiles.ensure(sizeof(TileDesc) * std::max(1, n_tiles)));
    HIP_OR_FAIL(ctx->d_fb_tiles.ensure(sizeof(TileDesc) * 2 * std::max(1, n_tiles)));
    HIP_OR_FAIL(ctx->d_ntiles.ensure(sizeof(int) * (n_cls + 4)));
    HIP_OR_FAIL(ctx->d_counters.ensure(sizeof(int) * (n_cls + 8)));
    HIP_OR_FAIL(ctx->d_fb_count.ensure(sizeof(int) * 4));
    HIP_OR_FAIL(hipMemcpyAsync(ctx->d_tiles.p, tiles.data(), sizeof(TileDesc) * n_tiles, hipMemcpyHostToDevice,
                               ctx->stream));
    HIP_OR_FAIL(hipMemcpyAsync(ctx->d_ntiles.p, counts.data(), sizeof(int) * counts.size(), hipMemcpyHostToDevice,
                               ctx->stream));
    HIP_OR_FAIL(launch_fill_zero(ctx->d_counters.p, sizeof(int) * (n_cls + 8), ctx->stream));
    HIP_OR_FAIL(launch_fill_zero(ctx->d_fb_count.p, sizeof(int) * 4, ctx->stream));

    // per-workgroup scratch for partial vectors
    const int ldv = round_up(maxLR, 4) + 4;
    const int max_grid = std::max(64, ctx->n_cu * 16);
    const long long per_wg_mfma = (long long)ctx->n_slots * 16 * ldv;
    const long long per_wg_plain = (long long)ctx->n_slots * ldv;
    const size_t scratch_bytes =
        (size_t)max_grid * std::max(per_wg_mfma * (path == CAFE_GPU_PATH_BATCHED32 ? 8 : 8), per_wg_plain * 8);
    HIP_OR_FAIL(ctx->d_scratch.ensure(scratch_bytes));

    PruneParams p{};
    p.buckets = ctx->d_buckets.as<BucketDev>();
    p.ops = ctx->d_ops.as<OpDev>();
    p.n_ops = (int)ctx->ops.size();
    p.kids = ctx->d_kids.as<ChildDev>();
    p.n_nodes = ctx->n_nodes;
    p.node_slot = ctx->d_node_slot.as<int>();
    p.counts_T = ctx->d_counts.as<int>();
    p.F = F;
    p.perm = ctx->d_perm.as<int>();
    p.mt64 = ctx->mt64.as<double>();
    p.mt32 = ctx->mt32.as<float>();
    p.slot_elems = ctx->slot_elems;
    p.Np = ctx->Np;
    p.err64 = ctx->n_dev ? ctx->d_err64.as<double>() : nullptr;
    p.err32 = ctx->n_dev ? ctx->d_err32.as<float>() : nullptr;
    p.n_dev = ctx->n_dev;
    p.prior64 = ctx->d_prior64.as<double>();
    p.prior32 = ctx->d_prior32.as<float>();
    p.model = model;
    p.out_val = ctx->d_val.as<double>();
    p.out_flag = ctx->d_flag.as<int>();
    p.root_out = root_out;
    p.root_stride = root_stride;
    p.scratch = ctx->d_scratch.p;
    p.ldv = ldv;
    p.n_slots = ctx->n_slots;

    int* d_cnt = ctx->d_ntiles.as<int>();
    int* d_ctr = ctx->d_counters.as<int>();
    HIP_OR_FAIL(hipEventRecord(ctx->ev[0], ctx->stream));
    for (int ci = 0; ci < n_cls; ++ci) {
        PruneParams q = p;
        q.tiles = ctx->d_tiles.as<TileDesc>() + classes[ci].begin;
        q.n_tiles_dev = d_cnt + ci;
        q.tile_counter = d_ctr + ci;
        q.scratch_per_wg = path == CAFE_GPU_PATH_PLAIN64 ? per_wg_plain : per_wg_mfma;
        const int grid = std::max(1, std::min(counts[ci], max_grid));
        HIP_OR_FAIL(launch_prune(path, classes[ci].nt, grid, q, ctx->stream));
    }
    HIP_OR_FAIL(hipEventRecord(ctx->ev[1], ctx->stream));

    // recomputation passes (device-side lists, no host round trip)
    TileDesc* fb64 = ctx->d_fb_tiles.as<TileDesc>();
    TileDesc* fbplain = fb64 + std::max(1, n_tiles);
    int* fbc = ctx->d_fb_count.as<int>();   // [0] fp64 tiles, [1] plain tiles
    const TileDesc* plain_src = ctx->d_tiles.as<TileDesc>();
    const int* plain_src_n = nullptr;
    if (path == CAFE_GPU_PATH_BATCHED32) {
        for (int ci = 0; ci < n_cls; ++ci)
            HIP_OR_FAIL(launch_compact(ctx->d_tiles.as<TileDesc>() + classes[ci].begin, d_cnt + ci, counts[ci],
                                       ctx->d_buckets.as<BucketDev>(), ctx->d_perm.as<int>(), ctx->d_flag.as<int>(),
                                       F, kFlagRecompute64, fb64, fbc + 0, fbc + 2, ctx->stream));
        PruneParams q = p;
        q.tiles = fb64;
        q.n_tiles_dev = fbc + 0;
        q.tile_counter = d_ctr + n_cls;
        q.scratch_per_wg = per_wg_mfma;
        HIP_OR_FAIL(launch_prune(CAFE_GPU_PATH_BATCHED64, 8, std::min(max_grid, std::max(1, n_tiles)), q,
                                 ctx->stream));
        HIP_OR_FAIL(launch_compact(fb64, fbc + 0, n_tiles, ctx->d_buckets.as<BucketDev>(), ctx->d_perm.as<int>(),
                                   ctx->d_flag.as<int>(), F, kFlagRecomputePlain, fbplain, fbc + 1, fbc + 3, ctx->stream));
        plain_src_n = fbc + 1;
    } else if (path == CAFE_GPU_PATH_BATCHED64) {
        for (int ci = 0; ci < n_cls; ++ci)
            HIP_OR_FAIL(launch_compact(ctx->d_tiles.as<TileDesc>() + classes[ci].begin, d_cnt + ci, counts[ci],
                                       ctx->d_buckets.as<BucketDev>(), ctx->d_perm.as<int>(), ctx->d_flag.as<int>(),
                                       F, kFlagRecomputePlain, fbplain, fbc + 1, fbc + 3, ctx->stream));
        plain_src_n = fbc + 1;
    }
    (void)plain_src;
    if (plain_src_n) {
        PruneParams q = p;
        q.tiles = fbplain;
        q.n_tiles_dev = plain_src_n;
        q.tile_counter = d_ctr + n_cls + 1;
        q.scratch_per_wg = per_wg_plain;
        int nq = 1;
        while (nq * 64 < maxLR) nq <<= 1;
        HIP_OR_FAIL(launch_prune(CAFE_GPU_PATH_PLAIN64, nq, std::min(max_grid, std::max(1, n_tiles)), q,
                                 ctx->stream));
    }
    HIP_OR_FAIL(hipEventRecord(ctx->ev[2], ctx->stream));
    if (n_fallback) {
        int h[4] = {0, 0, 0, 0};
        HIP_OR_FAIL(hipMemcpyAsync(h, fbc, sizeof(int) * 4, hipMemcpyDeviceToHost, ctx->stream));
        HIP_OR_FAIL(hipStreamSynchronize(ctx->stream));
        // families recomputed: FP64 pass + plain pass
        *n_fallback = (long long)h[2] + (long long)h[3];
    }
    return CAFE_GPU_OK;
}

int prepare_eval(cafe_gpu_ctx* ctx, const cafe_gpu_eval_desc* d, int& P)
{
    if (!d) return fail(CAFE_GPU_EINVAL, "null descriptor");
    if (ctx->n_nodes == 0) return fail(CAFE_GPU_ESTATE, "set_tree has not been called");
    if (ctx->F == 0) return fail(CAFE_GPU_ESTATE, "set_families has not been called");
    if (ctx->N == 0) return fail(CAFE_GPU_ESTATE, "no matrices have been built");
    if (ctx->n_leaves != ctx->tree_leaves)
        return fail(CAFE_GPU_EINVAL, "families have %d leaves, tree has %d", ctx->n_leaves, ctx->tree_leaves);
    if (d->model != CAFE_GPU_MODEL_BASE && d->model != CAFE_GPU_MODEL_GAMMA)
        return fail(CAFE_GPU_EINVAL, "unknown model %d", d->model);
    if (d->path < CAFE_GPU_PATH_PLAIN64 || d->path > CAFE_GPU_PATH_BATCHED32)
        return fail(CAFE_GPU_EINVAL, "unknown path %d", d->path);
    const int K = d->n_cat;
    if (d->n_trials < 1 || K < 1 || (d->model == CAFE_GPU_MODEL_BASE && K != 1))
        return fail(CAFE_GPU_EINVAL, "bad trial/category counts");
    if (!d->node_slot || !d->prior) return fail(CAFE_GPU_EINVAL, "null node_slot or prior");
    if (d->model == CAFE_GPU_MODEL_GAMMA && !d->cat_probs) return fail(CAFE_GPU_EINVAL, "gamma needs cat_probs");
    const int h = ctx->n_dev ? (ctx->n_dev - 1) / 2 : 0;
    int maxR = 0;
    for (auto& b : ctx->buckets) {
        if (std::max(b.B, b.R) + 1 > ctx->N)
            return fail(CAFE_GPU_EINVAL, "bucket B=%d R=%d needs matrices of size %d (built %d)", b.B, b.R,
                        std::max(b.B, b.R) + 1, ctx->N);
        if (ctx->n_dev && (b.B >= ctx->err_rows || b.max_count + h > b.B))
            return fail(CAFE_GPU_EINVAL, "error model does not cover bucket B=%d (max count %d)", b.B, b.max_count);
        maxR = std::max(maxR, b.R);
    }
    if (d->prior_stride < maxR) return fail(CAFE_GPU_EINVAL, "prior_stride %d < max R %d", d->prior_stride, maxR);
    P = d->n_trials * K;
    const long long ns = (long long)P * ctx->n_nodes;
    std::vector<int> slot(d->node_slot, d->node_slot + ns);
    int root = -1;
    for (auto& op : ctx->ops) root = op.node;   // root is the last op
    for (long long i = 0; i < ns; ++i) {
        const int node = (int)(i % ctx->n_nodes);
        if (node == root) {
            slot[i] = 0;
            continue;
        }
        if (slot[i] < 0 || slot[i] >= ctx->capacity || !ctx->built[slot[i]])
            return fail(CAFE_GPU_ENOTFOUND, "node %d of point %lld uses matrix slot %d, which is not built", node,
                        i / ctx->n_nodes, slot[i]);
    }
    const int nb = (int)ctx->buckets.size();
    std::vector<double> pr64((size_t)nb * d->prior_stride, 0.0);
    std::vector<float> pr32(pr64.size());
    std::vector<BucketDev> bd(nb);
    for (int b = 0; b < nb; ++b) {
        const auto& bh = ctx->buckets[b];
        for (int j = 0; j < bh.R; ++j) pr64[(size_t)b * d->prior_stride + j] = d->prior[(size_t)b * d->prior_stride + j];
        bd[b].B = bh.B;
        bd[b].R = bh.R;
        bd[b].L = bh.B + 1;
        bd[b].Lp = round_up(bh.B + 1, kTile);
        bd[b].Rp = round_up(bh.R, kTile);
        bd[b].fam_begin = bh.fam_begin;
        bd[b].count = bh.count;
        bd[b].prior_off = b * d->prior_stride;
    }
    for (size_t i = 0; i < pr64.size(); ++i) pr32[i] = (float)pr64[i];
    HIP_OR_FAIL(ctx->d_node_slot.ensure(sizeof(int) * ns));
    HIP_OR_FAIL(ctx->d_prior64.ensure(sizeof(double) * pr64.size()));
    HIP_OR_FAIL(ctx->d_prior32.ensure(sizeof(float) * pr32.size()));
    HIP_OR_FAIL(ctx->d_buckets.ensure(sizeof(BucketDev) * nb));
    HIP_OR_FAIL(ctx->d_catp.ensure(sizeof(double) * K));
    HIP_OR_FAIL(hipMemcpyAsync(ctx->d_node_slot.p, slot.data(), sizeof(int) * ns, hipMemcpyHostToDevice, ctx->stream));
    HIP_OR_FAIL(hipMemcpyAsync(ctx->d_prior64.p, pr64.data(), sizeof(double) * pr64.size(), hipMemcpyHostToDevice,
                               ctx->stream));
    HIP_OR_FAIL(hipMemcpyAsync(ctx->d_prior32.p, pr32.data(), sizeof(float) * pr32.size(), hipMemcpyHostToDevice,
                               ctx->stream));
    HIP_OR_FAIL(hipMemcpyAsync(ctx->d_buckets.p, bd.data(), sizeof(BucketDev) * nb, hipMemcpyHostToDevice,
                               ctx->stream));
    std::vector<double> cp(K, 1.0);
    if (d->cat_probs) cp.assign(d->cat_probs, d->cat_probs + K);
    HIP_OR_FAIL(hipMemcpyAsync(ctx->d_catp.p, cp.data(), sizeof(double) * K, hipMemcpyHostToDevice, ctx->stream));
    const size_t PF = (size_t)P * ctx->F;
    HIP_OR_FAIL(ctx->d_val.ensure(sizeof(double) * PF));
    HIP_OR_FAIL(ctx->d_flag.ensure(sizeof(int) * PF));
    HIP_OR_FAIL(hipStreamSynchronize(ctx->stream));   // host vectors above go out of scope
    return CAFE_GPU_OK;
}

}  // namespace

extern "C" {

int cafe_gpu_eval(cafe_gpu_ctx* ctx, const cafe_gpu_eval_desc* d, double* lnl, double* cat, double* post,
                  double* total, long long* n_fallback)
{
    if (int rc = check_ctx(ctx)) return rc;
    int P = 0;
    if (int rc = prepare_eval(ctx, d, P)) return rc;
    std::vector<int> pts(P);
    std::iota(pts.begin(), pts.end(), 0);
    if (int rc = run_prune(ctx, d->path, d->model, P, pts.data(), nullptr, 0, n_fallback)) return rc;
    const int T = d->n_trials, K = d->n_cat, F = ctx->F;
    const size_t TF = (size_t)T * F;
    HIP_OR_FAIL(ctx->d_lnl.ensure(sizeof(double) * TF));
    HIP_OR_FAIL(ctx->d_cat.ensure(sizeof(double) * TF * K));
    HIP_OR_FAIL(ctx->d_post.ensure(sizeof(double) * TF * K));
    HIP_OR_FAIL(ctx->d_fail.ensure(sizeof(int) * TF));
    HIP_OR_FAIL(ctx->d_total.ensure(sizeof(double) * T));
    HIP_OR_FAIL(launch_finalize(d->model, T, K, F, ctx->d_val.as<double>(), ctx->d_flag.as<int>(),
                                ctx->d_catp.as<double>(), ctx->d_lnl.as<double>(), ctx->d_cat.as<double>(),
                                ctx->d_post.as<double>(), ctx->d_fail.as<int>(), ctx->d_total.as<double>(),
                                ctx->stream));
    HIP_OR_FAIL(hipEventRecord(ctx->ev[3], ctx->stream));
    if (lnl) HIP_OR_FAIL(hipMemcpyAsync(lnl, ctx->d_lnl.p, sizeof(double) * TF, hipMemcpyDeviceToHost, ctx->stream));
    if (cat && d->model == CAFE_GPU_MODEL_GAMMA)
        HIP_OR_FAIL(hipMemcpyAsync(cat, ctx->d_cat.p, sizeof(double) * TF * K, hipMemcpyDeviceToHost, ctx->stream));
    if (post && d->model == CAFE_GPU_MODEL_GAMMA)
        HIP_OR_FAIL(hipMemcpyAsync(post, ctx->d_post.p, sizeof(double) * TF * K, hipMemcpyDeviceToHost, ctx->stream));
    if (total) HIP_OR_FAIL(hipMemcpyAsync(total, ctx->d_total.p, sizeof(double) * T, hipMemcpyDeviceToHost, ctx->stream));
    HIP_OR_FAIL(hipStreamSynchronize(ctx->stream));
    float ms = 0;
    (void)hipEventElapsedTime(&ms, ctx->ev[0], ctx->ev[1]);
    ctx->timings[1] = ms;
    (void)hipEventElapsedTime(&ms, ctx->ev[1], ctx->ev[2]);
    ctx->timings[2] = ms;
    (void)hipEventElapsedTime(&ms, ctx->ev[2], ctx->ev[3]);
    ctx->timings[3] = ms;
    return CAFE_GPU_OK;
}

int cafe_gpu_root_vectors(cafe_gpu_ctx* ctx, const cafe_gpu_eval_desc* d, int trial, int cat_k, int f0, int nf,
                          double* out)
{
    if (int rc = check_ctx(ctx)) return rc;
    int P = 0;
    if (int rc = prepare_eval(ctx, d, P)) return rc;
    if (trial < 0 || trial >= d->n_trials || cat_k < 0 || cat_k >= d->n_cat)
        return fail(CAFE_GPU_EINVAL, "trial/category out of range");
    if (f0 < 0 || nf < 0 || f0 + nf > ctx->F || !out) return fail(CAFE_GPU_EINVAL, "family range out of range");
    int maxR = 0;
    for (auto& b : ctx->buckets) maxR = std::max(maxR, b.R);
    HIP_OR_FAIL(ctx->d_root.ensure(sizeof(double) * (size_t)ctx->F * maxR));
    const int point = trial * d->n_cat + cat_k;
    if (int rc = run_prune(ctx, d->path, d->model, 1, &point, ctx->d_root.as<double>(), maxR, nullptr)) return rc;
    std::vector<double> h((size_t)ctx->F * maxR);
    HIP_OR_FAIL(hipMemcpyAsync(h.data(), ctx->d_root.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, ctx->stream));
    HIP_OR_FAIL(hipStreamSynchronize(ctx->stream));
    // R of each family from its bucket
    std::vector<int> famR(ctx->F);
    for (auto& b : ctx->buckets)
        for (int k = 0; k < b.count; ++k) famR[ctx->perm[b.fam_begin + k]] = b.R;
    size_t o = 0;
    for (int f = f0; f < f0 + nf; ++f)
        for (int j = 0; j < famR[f]; ++j) out[o++] = h[(size_t)f * maxR + j];
    return CAFE_GPU_OK;
}

int cafe_gpu_last_timings(cafe_gpu_ctx* ctx, double* ms4)
{
    if (!ctx || !ms4) return fail(CAFE_GPU_EINVAL, "null argument");
    for (int i = 0; i < 4; ++i) ms4[i] = ctx->timings[i];
    return CAFE_GPU_OK;
}

}  // extern "C"
