// api.hip -- host side of libcafe_gpu.so: context, planning and the C ABI of
// include/cafe_gpu.h.
#include <math.h>
#include <stdarg.h>
#include <ctype.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <functional>
#include <map>
#include <mutex>
#include <numeric>
#include <string>
#include <vector>

#include "cafe_gpu.h"
#include "engine.h"

namespace cafe {
hipError_t launch_build(int algo, const double* d_lam, const double* d_t, const double* d_lg, int n,
                        int first_slot, int N, int Np, long long slot_elems, double* mt64, float* mt32,
                        float* lf32, hipStream_t stream, const unsigned char* copies, int top, int* band);
hipError_t launch_fill_zero(void* p, long long bytes, hipStream_t stream);
hipError_t launch_reconstruct(const ReconNode* nodes, int n_nodes, const int* kids, const int* parent,
                              const double* mt64, int N, int Np, long long slot_elems, const int* counts_T, int F,
                              const int* fam_sorted, int nf, const double* prior, int B, int Rr, double* L_all,
                              int* C_all, double* prod_all, int Lp, int* states, int* err, hipStream_t st);
hipError_t launch_fold(const double* mt64, const float* lf32, double* out64, float* out32, const double* err64,
                       const float* err32, int n_dev, int n_rows, int N, int Np, long long slot_elems,
                       int first_slot, int n, hipStream_t stream);
hipError_t launch_prune_plain(int nq, int grid, const PruneParams& p, hipStream_t st);
hipError_t launch_prune_mfma(bool fp64, const MfmaShape& s, int grid, const PruneParams& p, hipStream_t st);
hipError_t mfma_occupancy(bool fp64, const MfmaShape& s, int* per_cu);
hipError_t f16_dispatch(const MfmaShape& s, int grid, const PruneParams& p, hipStream_t st, int* per_cu);
hipError_t launch_split_f16(const float* mt32, _Float16* hi, _Float16* lo, int Np, long long slot_elems, int n_slots,
                            hipStream_t st);
int lane_rows_for(int rows);
int lane_max_slots(int lane_rows);
hipError_t lane_dispatch(int lr, int ns, int grid, size_t shm, const PruneParams& p, hipStream_t st, int* per_cu);
hipError_t launch_pack(const SegDesc* segs, int n_segs, const BucketDev* buckets, const int* perm,
                       const int* flags, int F, int mask, int* fam_list, TileDesc* out_tiles,
                       const int* class_off, int* class_count, int* fam_count, int tile, hipStream_t st);
hipError_t launch_finalize(int model, int T, int K, int F, const double* val, const int* flag,
                           const double* cat_probs, double* lnl, double* cat, double* post, int* fail,
                           double* total, hipStream_t st);
}  // namespace cafe

using namespace cafe;

// side streams per context: the size classes of one pruning pass (as many as
// the runtime's default hardware queues per process)
constexpr int kSideStreams = 4;

// Fixed device scratch of the entry points that allocate nothing
// (cafe_gpu_reconstruct_noalloc): a reconstruction's program, family list,
// state rows, error flags and per-family L, C and product rows.  Every device
// the library runs on holds its own copy for the life of the process.  In host
// code `scratch` names a host-side stand-in, never the device memory: its
// device address comes from hipGetSymbolAddress, once per device (see
// fixed_scratch_of), and only that address is handed to kernels and copies.
constexpr size_t kFixedScratchBytes = (size_t)32 << 20;
static __device__ __align__(256) unsigned char scratch[kFixedScratchBytes];

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
size_t align256(size_t x) { return (x + 255) / 256 * 256; }

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

// Device buffers of one evaluation (eager calls or a plan).
struct EvalBufs {
    DevBuf node_slot, prior64, prior32, catp, buckets, tiles, ntiles, counters, fb_tiles, fb_count, val,
        flag, lnl, cat, post, fail, total, scratch, root, segs, class_off, fam_list;
    void release_all()
    {
        DevBuf* all[] = {&node_slot, &prior64, &prior32, &catp, &buckets, &tiles, &ntiles, &counters, &fb_tiles,
                         &fb_count, &val, &flag, &lnl, &cat, &post, &fail, &total, &scratch, &root, &segs,
                         &class_off, &fam_list};
        for (DevBuf* b : all) b->release();
    }
};

}  // namespace

struct cafe_gpu_ctx {
    int device = 0;
    int n_cu = 256;
    hipStream_t stream = nullptr;
    // size classes of one pruning phase run concurrently on these (default
    // priority), forked from and joined back to `stream` with join_ev
    std::vector<hipStream_t> side;
    std::vector<hipEvent_t> join_ev;
    hipEvent_t ev[6] = {};
    double timings[4] = {0, 0, 0, 0};
    bool build_timed = false;
    DevBuf lg;

    // matrix pool
    int N = 0, Np = 0, capacity = 0;
    long long slot_elems = 0;
    DevBuf mt64, mt32, lf32, key_lam, key_t;
    DevBuf band32;         // per slot and 16-row tile of mt32: its nonzero k-tiles (band_entry)
    DevBuf mt16h, mt16l;   // split-FP16 halves of mt32 (BATCHED16; made by its evaluations)
    std::vector<char> built;   // per slot: 0 no, 1 every copy, 2 FP64 and the FP32 copies a plan read

    // tree program
    int n_nodes = 0, tree_leaves = 0, n_slots = 0, max_children = 2;
    std::vector<int> parent_h, leaf_col_h;   // the tree as set
    DevBuf d_recon;                          // reconstruction: program, families, prior, states, flags
    DevBuf d_recon_scratch;                  // reconstruction: per-family L, C and product rows
    char* fixed_scratch = nullptr;           // this device's copy of `scratch` (device address)
    std::vector<OpDev> ops;
    std::vector<ChildDev> kids;
    DevBuf d_ops, d_kids, d_steps;
    int n_steps = 0;
    DevBuf d_steps_lane;     // the same steps with output-reusing slots (split-FP16 kernel)
    int n_slots_lane = 0;
    DevBuf d_steps_hot;      // and with the vector the next op reads kept in registers (per-lane kernel)
    int n_slots_hot = 0;

    // families
    int F = 0, n_leaves = 0;
    std::vector<int> perm;
    std::vector<BucketHost> buckets;
    DevBuf d_counts, d_perm;

    // error model, and the pools with it folded into every matrix (leaf
    // branches read those)
    int n_dev = 0, err_rows = 0;
    DevBuf d_err64, d_err32, mt64e, mt32e;

    // buffers of eager evaluations (plans own their own)
    EvalBufs eager;
    long long pool_gen = 0;   // bumped by matrix_reset; plans check it
    // measured launch shapes: (path, tiles per vector, leaves) -> shape
    struct TuneEntry {
        int path, nt, leaves, kernel, ntw, w, lds, grid_mult;
        // the plan's size classes it was measured in (bit nt - 1 per class of
        // nt tiles); 0: any plan
        unsigned long long plan_classes;
    };
    std::vector<TuneEntry> tuning;
    // bumped by set_tree, set_families and by set_error_model when it adds or
    // removes the model or changes its shape (buffers a plan's graph points
    // into); plans check it.  Same-shape error-model rows are updated in place.
    long long input_gen = 0;
};

namespace {

int check_ctx(cafe_gpu_ctx* ctx)
{
    if (!ctx) return fail(CAFE_GPU_EINVAL, "null context");
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(CAFE_GPU_EDEVICE, "hipSetDevice failed");
    return CAFE_GPU_OK;
}

// Device address of this device's `scratch` array, resolved once per device
// with that device current (cafe_gpu_create makes it current first).
int fixed_scratch_of(int device, char** out)
{
    static std::mutex mu;
    static std::map<int, char*> addr;
    std::lock_guard<std::mutex> lock(mu);
    auto it = addr.find(device);
    if (it == addr.end()) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(CAFE_GPU_EDEVICE, "hipGetSymbolAddress(scratch) failed on device %d: %s", device,
                        hipGetErrorString(e));
        it = addr.emplace(device, (char*)p).first;
    }
    *out = it->second;
    return CAFE_GPU_OK;
}

// Pools with the error model folded in: sized with the matrix pool (outside
// any graph capture).
int ensure_fold_pools(cafe_gpu_ctx* ctx)
{
    if (!ctx->n_dev || !ctx->capacity) return CAFE_GPU_OK;
    HIP_OR_FAIL(ctx->mt64e.ensure(sizeof(double) * ctx->slot_elems * ctx->capacity));
    HIP_OR_FAIL(ctx->mt32e.ensure(sizeof(float) * ctx->slot_elems * ctx->capacity));
    return CAFE_GPU_OK;
}

// Enqueues the error-model fold of slots first .. first+n-1 (nothing without
// an error model) into the FP64 pool (fp64) and the FP32 one (fp32).
int enqueue_fold(cafe_gpu_ctx* ctx, int first, int n, hipStream_t st, bool fp64 = true, bool fp32 = true)
{
    if (!ctx->n_dev || n <= 0) return CAFE_GPU_OK;
    HIP_OR_FAIL(launch_fold(ctx->mt64.as<double>(), ctx->lf32.as<float>(), fp64 ? ctx->mt64e.as<double>() : nullptr,
                            fp32 ? ctx->mt32e.as<float>() : nullptr, ctx->d_err64.as<double>(),
                            ctx->d_err32.as<float>(), ctx->n_dev, ctx->err_rows, ctx->N, ctx->Np, ctx->slot_elems,
                            first, n, st));
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
    // reuse: a node's output may take one of its children's slots (a kernel
    // whose lanes read every child before writing the output, in order)
    bool reuse = false;
    // hot (with reuse): the last interior child a node's op reads is the op
    // just before it, so its vector is read from the kernel's hot registers
    // (in_slot kHotSlot) and takes no slot; so is every output that the next
    // op reads (out_slot kHotSlot), and the root's
    bool hot = false;

    Planner(const std::vector<std::vector<int>>& k, const std::vector<int>& lc, bool r = false, bool h = false)
        : kids(k), leaf_col(lc), need(k.size(), -1), slot_of(k.size(), -1), reuse(r || h), hot(h) {}

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

    void emit(int u, bool store = true)
    {
        auto in = interior_order(u);
        for (size_t i = 0; i < in.size(); ++i) emit(in[i], !hot || i + 1 < in.size());
        if (reuse)
            for (int c : in)
                if (slot_of[c] != kHotSlot) used[slot_of[c]] = 0;
        const int out = hot && !store ? kHotSlot : alloc();
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
        if (!reuse)
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
    ctx->side.assign(kSideStreams, nullptr);
    for (auto& st : ctx->side)
        if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) {
            cafe_gpu_destroy(ctx);
            return fail(CAFE_GPU_EDEVICE, "hipStreamCreate failed");
        }
    if (int rc = fixed_scratch_of(device, &ctx->fixed_scratch)) {
        const std::string msg = cafe_gpu_last_error();
        cafe_gpu_destroy(ctx);
        return fail(rc, "%s", msg.c_str());
    }
    ctx->join_ev.assign(3 * (kSideStreams + 1), nullptr);
    for (auto& e : ctx->join_ev) (void)hipEventCreateWithFlags(&e, hipEventDisableTiming);
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
    DevBuf* bufs[] = {&ctx->lg, &ctx->mt64, &ctx->mt32, &ctx->lf32, &ctx->key_lam, &ctx->key_t, &ctx->d_ops,
                      &ctx->d_kids, &ctx->d_steps, &ctx->d_steps_lane, &ctx->d_steps_hot, &ctx->d_counts, &ctx->d_perm, &ctx->d_err64,
                      &ctx->d_err32, &ctx->mt64e, &ctx->mt32e, &ctx->mt16h, &ctx->mt16l, &ctx->band32};
    for (DevBuf* b : bufs) b->release();
    ctx->eager.release_all();
    for (auto& e : ctx->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : ctx->join_ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& st : ctx->side)
        if (st) {
            (void)hipStreamSynchronize(st);
            (void)hipStreamDestroy(st);
        }
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
    delete ctx;
    return CAFE_GPU_OK;
}

int cafe_gpu_matrix_reset(cafe_gpu_ctx* ctx, int N, int capacity)
{
    if (int rc = check_ctx(ctx)) return rc;
    if (N < 2 || N * 2 + 2 > kLgammaN) return fail(CAFE_GPU_EINVAL, "matrix size %d out of range", N);
    if (capacity < 1) return fail(CAFE_GPU_EINVAL, "capacity must be positive");
    const int Np = round_up(N + kTile, 32);   // >= round_up(N, 16) + 16, whole 32-wide transpose tiles
    const long long slot_elems = (long long)Np * Np;
    const size_t n64 = sizeof(double) * slot_elems * capacity;
    const size_t n32 = sizeof(float) * slot_elems * capacity;
    // every plan made over the old pool is stale from here on, and until the
    // new pool is complete the context has none
    ctx->pool_gen++;
    ctx->N = ctx->Np = ctx->capacity = 0;
    ctx->slot_elems = 0;
    ctx->built.clear();
    HIP_OR_FAIL(ctx->mt64.ensure(n64));
    HIP_OR_FAIL(ctx->mt32.ensure(n32));
    HIP_OR_FAIL(ctx->lf32.ensure(n32));
    const size_t nband = sizeof(int) * (size_t)(Np / kTile) * capacity;
    HIP_OR_FAIL(ctx->band32.ensure(nband));
    if (ctx->n_dev) {
        HIP_OR_FAIL(ctx->mt64e.ensure(n64));
        HIP_OR_FAIL(ctx->mt32e.ensure(n32));
    }
    // padding rows/columns must read as zero
    HIP_OR_FAIL(launch_fill_zero(ctx->mt64.p, (long long)n64, ctx->stream));
    HIP_OR_FAIL(launch_fill_zero(ctx->mt32.p, (long long)n32, ctx->stream));
    HIP_OR_FAIL(launch_fill_zero(ctx->lf32.p, (long long)n32, ctx->stream));
    HIP_OR_FAIL(launch_fill_zero(ctx->band32.p, (long long)nband, ctx->stream));
    ctx->N = N;
    ctx->Np = Np;
    ctx->slot_elems = slot_elems;
    ctx->capacity = capacity;
    if (int rc = enqueue_fold(ctx, 0, capacity, ctx->stream)) return rc;
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
                             ctx->mt32.as<float>(), ctx->lf32.as<float>(), ctx->stream, nullptr, 0,
                             ctx->band32.as<int>()));
    if (int rc = enqueue_fold(ctx, first_slot, n, ctx->stream)) return rc;
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
        return fail(CAFE_GPU_ENOTFOUND, "matrix slot %d is not built", slot);
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
    // the per-lane kernel's program: the same order, the vector each op reads
    // from the op just before it in registers, the others in output-reusing
    // slots
    Planner ph(kids, lc, true, true);
    ph.compute_need(root);
    ph.used.clear();
    ph.emit(root, false);
    if (pl.max_slot + 1 > kMaxSlots)
        return fail(CAFE_GPU_EINVAL, "tree needs %d live partial vectors (max %d)", pl.max_slot + 1, kMaxSlots);
    ctx->input_gen++;
    ctx->n_nodes = n_nodes;
    ctx->tree_leaves = n_leaves;
    ctx->parent_h.assign(parent, parent + n_nodes);
    ctx->leaf_col_h.assign(leaf_col, leaf_col + n_nodes);
    ctx->max_children = 1;
    for (auto& k : kids) ctx->max_children = std::max(ctx->max_children, (int)k.size());
    ctx->n_slots = pl.max_slot + 1;
    ctx->ops = pl.ops;
    ctx->kids = pl.out_kids;
    HIP_OR_FAIL(ctx->d_ops.ensure(sizeof(OpDev) * ctx->ops.size()));
    HIP_OR_FAIL(ctx->d_kids.ensure(sizeof(ChildDev) * ctx->kids.size()));
    HIP_OR_FAIL(hipMemcpy(ctx->d_ops.p, ctx->ops.data(), sizeof(OpDev) * ctx->ops.size(), hipMemcpyHostToDevice));
    HIP_OR_FAIL(hipMemcpy(ctx->d_kids.p, ctx->kids.data(), sizeof(ChildDev) * ctx->kids.size(),
                          hipMemcpyHostToDevice));
    // the same program as a flat list of child steps
    auto flatten = [](const std::vector<OpDev>& ops, const std::vector<ChildDev>& kids) {
        std::vector<StepDev> steps;
        for (size_t o = 0; o < ops.size(); ++o) {
            const OpDev& op = ops[o];
            for (int k = 0; k < op.n_child; ++k) {
                const ChildDev& ch = kids[op.child_begin + k];
                int ctl = op.out_slot;
                if (k == 0) ctl |= kStepFirst;
                if (k == op.n_child - 1) ctl |= kStepLast;
                if (o + 1 == ops.size()) ctl |= kStepRoot;
                steps.push_back(StepDev{ch.node, ch.leaf_col, ch.in_slot, ctl});
            }
        }
        return steps;
    };
    const std::vector<StepDev> steps = flatten(ctx->ops, ctx->kids);
    ctx->n_steps = (int)steps.size();
    HIP_OR_FAIL(ctx->d_steps.ensure(sizeof(StepDev) * steps.size()));
    HIP_OR_FAIL(hipMemcpy(ctx->d_steps.p, steps.data(), sizeof(StepDev) * steps.size(), hipMemcpyHostToDevice));
    // the lane kernel's program: the same order, outputs reusing child slots
    Planner pr(kids, lc, true);
    pr.compute_need(root);
    pr.used.clear();
    pr.emit(root);
    const std::vector<StepDev> lsteps = flatten(pr.ops, pr.out_kids);
    ctx->n_slots_lane = pr.max_slot + 1;
    HIP_OR_FAIL(ctx->d_steps_lane.ensure(sizeof(StepDev) * lsteps.size()));
    HIP_OR_FAIL(hipMemcpy(ctx->d_steps_lane.p, lsteps.data(), sizeof(StepDev) * lsteps.size(),
                          hipMemcpyHostToDevice));
    const std::vector<StepDev> hsteps = flatten(ph.ops, ph.out_kids);
    ctx->n_slots_hot = std::max(1, ph.max_slot + 1);
    HIP_OR_FAIL(ctx->d_steps_hot.ensure(sizeof(StepDev) * hsteps.size()));
    HIP_OR_FAIL(hipMemcpy(ctx->d_steps_hot.p, hsteps.data(), sizeof(StepDev) * hsteps.size(),
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
    ctx->input_gen++;
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
        if (ctx->n_dev) ctx->input_gen++;
        ctx->n_dev = 0;
        ctx->err_rows = 0;
        return CAFE_GPU_OK;
    }
    if (n_dev < 1 || n_dev % 2 == 0 || n_rows < 1)
        return fail(CAFE_GPU_EINVAL, "error model needs an odd number of deviations and >= 1 row");
    if (n_dev != ctx->n_dev || n_rows != ctx->err_rows) ctx->input_gen++;   // same shape: in place
    const size_t n = (size_t)n_dev * n_rows;
    std::vector<float> r32(rows, rows + n);
    HIP_OR_FAIL(ctx->d_err64.ensure(sizeof(double) * n));
    HIP_OR_FAIL(ctx->d_err32.ensure(sizeof(float) * n));
    HIP_OR_FAIL(hipMemcpy(ctx->d_err64.p, rows, sizeof(double) * n, hipMemcpyHostToDevice));
    HIP_OR_FAIL(hipMemcpy(ctx->d_err32.p, r32.data(), sizeof(float) * n, hipMemcpyHostToDevice));
    ctx->n_dev = n_dev;
    ctx->err_rows = n_rows;
    // fold into every slot of an existing pool (later builds fold their own slots)
    if (int rc = ensure_fold_pools(ctx)) return rc;
    if (int rc = enqueue_fold(ctx, 0, ctx->capacity, ctx->stream)) return rc;
    HIP_OR_FAIL(hipStreamSynchronize(ctx->stream));
    return CAFE_GPU_OK;
}

}  // extern "C"


namespace {

// Kernel size class: buckets whose vectors have the same number of 16-row
// tiles (MFMA paths) or the same PLAIN64 rows per lane (PLAIN64 path).
struct KernelClass {
    int nt;              // 16-row tiles of the class's largest vector
    int nq;              // PLAIN64 rows per lane (row s = lane + 64 q)
    int slot_tiles;      // 16-row tiles of its largest stored (non-root) vector
    MfmaShape s32, s64;  // launch shapes of the FP32 and FP64 MFMA kernels
    MfmaShape s16;       // launch shape of the split-FP16 kernel (BATCHED16)
    MfmaShape r64;       // FP64 recompute of the FP32 pass (a few tiles: latency)
    int lane_rows = 0;   // > 0: the FP32 main pass runs k_prune_lane with this many rows (64-family tiles)
    bool plain = false;  // PLAIN64, or vectors over 64 tiles (1,024 rows): the main pass is k_prune_plain's (final FP64)
    int lane_slots = 0;  // its slot instance
    int lane_per_cu = 0; // its workgroups resident per CU
    int lane_grid_mult = 0;   // workgroups per resident slot (0: default)
    int fb32 = 1;        // FP32 matrix-core main pass: blocks of 16 families per tile (2: 32-family tiles)
    bool tuned = false;  // its shape came from the context's tuning table
    int r16 = 0;         // 16-family tiles its families fill (the recompute lists' capacity)
    int r_begin = 0;     // offset of its FP64 recompute list
    int begin, end;      // its main tiles
    int seg_begin = 0, seg_end = 0;   // its (bucket, point) segments
    // byte offsets of its global scratch in the main pass, the FP64 and the
    // PLAIN64 recompute: disjoint, since one class's recompute can run
    // beside another class's main pass
    size_t off_main = 0, off_r64 = 0, off_rp = 0;
    int stream = 0;      // side stream of its launches
};

// Host-side description of one pruning sequence over a tile list.
struct PrunePlan {
    int path = 0, model = 0;
    std::vector<KernelClass> classes;
    std::vector<int> class_counts;   // main tiles per class (+ padding)
    std::vector<int> plain_off;      // offsets of the one-family PLAIN64 recompute lists
    std::vector<int> main_order;     // main-pass launch order (longest first)
    int light_stream = 0;            // the side stream expected to finish its main-pass work first
    bool has_plain = false;          // a class runs its main pass on PLAIN64 (reads the FP64 pool)
    int plain_list_size = 0;
    int n_split = 0;                 // BATCHED16: pool slots converted to FP16 halves before pruning
    int n_tiles = 0, n_segs = 0, maxLR = 0;
    int r16_total = 0;               // capacity of the FP64 recompute lists (16-family tiles)
    int waves_per_cu = 16;           // persistent grid: resident waves per CU
    int n_cu = 256;
    long long per_wg_plain = 0;
    PruneParams base{};
    // fb_count layout: [0, n) FP64 recompute tiles per class, [n, 2n) PLAIN64
    // tiles per class, [2n] / [2n+1] families recomputed in each pass
    int fam_count_off() const { return 2 * (int)classes.size(); }
    // workgroups of a launch: resident waves per CU over the machine, no more
    // than the tiles, and a multiple of 8 (the MFMA kernels give each XCD a
    // contiguous share of the tiles) when there are at least 8
    int grid(int w, int max_tiles) const
    {
        const int g = std::max(1, std::min(n_cu * waves_per_cu / std::max(1, w), max_tiles));
        return g >= 8 ? g / 8 * 8 : g;
    }
    // a lane-kernel launch: as an MFMA launch with partials in LDS
    int grid_lane(int per_cu, int max_tiles, int mult = 0) const
    {
        const long long want = (long long)n_cu * std::max(1, per_cu) * (mult > 0 ? mult : 16);
        const int g = (int)std::max(1LL, std::min(want, (long long)max_tiles));
        return g >= 8 ? g / 8 * 8 : g;
    }
    // an MFMA launch: a multiple of the workgroups of its shape that are
    // resident at once.  Each workgroup takes an equal static share of the
    // tiles, so with many more workgroups than slots the dispatcher evens
    // out the per-CU load as slots free up (16x: the small-vector classes run
    // ~25% faster than with exactly one resident round).  Kernels with
    // partials in global scratch, which grows with the grid, take 2x.
    int grid(const MfmaShape& s, int max_tiles) const
    {
        const int per_cu = s.per_cu > 0 ? s.per_cu : std::max(1, waves_per_cu / std::max(1, s.w));
        const long long want = (long long)n_cu * per_cu * (s.grid_mult > 0 ? s.grid_mult : (s.lds ? 16 : 2));
        const int g = (int)std::max(1LL, std::min(want, (long long)max_tiles));
        return g >= 8 ? g / 8 * 8 : g;
    }
};

// MFMA launch shape for vectors of nt 16-row tiles: W waves per workgroup
// split the output tiles, NTW per wave, so that W NTW >= nt (one pass; the
// scaled FP32 kernel needs the whole vector in one pass).  Small vectors run
// one wave per family tile (throughput), large ones several (latency).
MfmaShape mfma_shape(int nt, bool fp64)
{
    MfmaShape s{1, 1, false, 0};
    if (nt <= 4) s = {std::max(1, nt), 1, false, 0};
    else if (nt <= 8) s = {(nt + 1) / 2, 2, false, 0};
    else if (nt <= 16) s = {(nt + 3) / 4, 4, false, 0};
    else if (nt <= 32) s = {(nt + 7) / 8, 8, false, 0};
    else s = {(nt + 15) / 16, 16, false, 0};   // up to 64 tiles
    return s;
}

// FP32 launch shape with 32-family tiles (two blocks of 16 per MFMA operand)
// for vectors of up to 32 tiles (kMaxTilesFb2): two output tiles per wave
// over 4, 8 or 16 waves (the products and factors of both blocks are held per
// output tile, so more tiles per wave would spill).
constexpr int kMaxTilesFb2 = 32;
MfmaShape mfma_shape_fb2(int nt)
{
    MfmaShape s{2, nt <= 8 ? 4 : nt <= 16 ? 8 : 16, false, 0};
    s.fb = 2;
    if (s.w == 16) s.grid_mult = 1;   // one resident round (global partials grow with the grid)
    return s;
}

// Built-in family-block choice of the FP32 kernel: 32-family tiles for the
// 17-32-tile vectors (one output tile of sixteen waves per 16 families would
// otherwise stream each matrix operand for one MFMA: 1.5x slower at B = 500)
// and for the only matrix-core class of a plan when it fills the device
// with at least one 32-family tile per CU (measured 1.1-1.4x faster: B = 100,
// 150; with 2,000 families, 9% slower); 16-family tiles for smaller vectors
// when several matrix-core classes run concurrently, where the shorter tile
// lists leave fewer waves resident beside the other classes and the step
// gains nothing.
int builtin_fb32(int path, int nt, bool only_mfma_class, long long fam_points, int n_cu)
{
    if (path != CAFE_GPU_PATH_BATCHED32 || nt > kMaxTilesFb2) return 1;
    return nt > 16 || (only_mfma_class && fam_points >= 32LL * n_cu) ? 2 : 1;
}

// FP64 recompute shape: its lists hold a few tiles, so the time is one tile's
// chain of steps; one or two output tiles per wave (four above 32 tiles) over
// up to 8 waves (16 above 16 tiles) keep each step's k-loop short, with room
// for four k-tiles of loads in flight.
MfmaShape mfma_shape_chain(int nt)
{
    if (nt <= 2) return {std::max(1, nt), 1, false, 0};
    if (nt <= 4) return {1, 4, false, 0};
    if (nt <= 8) return {1, 8, false, 0};
    if (nt <= 16) return {1, 16, false, 0};
    if (nt <= 32) return {2, 16, false, 0};
    return {4, 16, false, 0};
}

// Dynamic LDS of an MFMA launch without its partials: the bookkeeping and the
// tile's leaf counts.
size_t mfma_meta_only(const MfmaShape& s, bool fp64, int n_slots, int n_leaves)
{
    return (size_t)mfma_meta_bytes(n_slots, s.w, !fp64, s.fb) + (size_t)mfma_count_bytes(n_leaves, s.fb);
}

// Partial vectors go to LDS when the workgroup's share stays small enough for
// several workgroups per CU; the rest use the global scratch.  A latency
// shape (one workgroup per CU is enough: few tiles) may take up to 144 KB of
// the CU's 160 KB.
void place_partials(MfmaShape& s, bool fp64, int n_slots, int slot_tiles, int n_leaves, bool latency = false)
{
    const size_t meta = mfma_meta_only(s, fp64, n_slots, n_leaves);
    const size_t frag = (size_t)n_slots * slot_tiles * s.fb * 64 * 4 * (fp64 ? 8 : 4);
    if (latency)
        s.lds = (s.w <= 8 || (s.w == 16 && s.ntw == 1)) && meta + frag + 4096 <= 144 * 1024;
    else if (s.fb > 1)   // two workgroups per CU, or one of 8 waves
        s.lds = s.w <= 8 && meta + frag + 4096 <= (s.w >= 8 ? 144 : 76) * 1024;
    else
        s.lds = s.w <= 4 && meta + frag + 4096 <= 64 * 1024 && (meta + frag) / s.w <= 20 * 1024;
    s.shm = meta + (s.lds ? frag : 0);
}

// The split-FP16 kernel takes vectors of up to 32 tiles (eight waves of four
// tiles without spilling registers); BATCHED16 runs larger ones on the FP32
// matrix-core kernel.
constexpr int kF16MaxTiles = 32;
constexpr int kMfmaMaxTiles = 64;     // the matrix-core shapes' largest vectors (1,024 rows)
constexpr int kPlainMaxRows = 2048;   // k_prune_plain's largest instance: 32 rows per lane

// The split-FP16 kernel's partials: per 32-row block and lane three 16-byte
// B operands (hi, lo, bound).
size_t f16_frag_bytes(int n_slots, int slot_tiles) { return (size_t)3 * n_slots * ((slot_tiles + 1) / 2) * 64 * 16; }

void place_partials16(MfmaShape& s, int n_slots, int slot_tiles, int n_leaves)
{
    const size_t meta = (size_t)f16_meta_bytes(n_slots, s.w) + (size_t)mfma_count_bytes(n_leaves);
    const size_t frag = f16_frag_bytes(n_slots, slot_tiles);
    // two workgroups per CU (its registers allow two waves per SIMD)
    s.lds = s.w <= 4 && meta + frag + 4096 <= 76 * 1024;
    s.shm = meta + (s.lds ? frag : 0);
}

size_t f16_scratch_bytes(const MfmaShape& s, int n_slots, int slot_tiles, int grid)
{
    return s.lds ? 0 : (size_t)grid * f16_frag_bytes(n_slots, slot_tiles);
}

// Dynamic LDS of a lane-kernel launch: the tile's leaf counts.
size_t lane_shm(int n_leaves) { return (size_t)n_leaves * 64 * sizeof(int); }

// Global scratch of an MFMA launch (partials not in LDS), bytes.
size_t mfma_scratch_bytes(const MfmaShape& s, bool fp64, int n_slots, int slot_tiles, int grid)
{
    if (s.lds) return 0;
    return (size_t)grid * n_slots * slot_tiles * s.fb * 64 * 4 * (fp64 ? 8 : 4);
}

// The tuning entry for a class of a plan whose size classes are plan_classes
// (bit nt - 1 per class of nt tiles), if the context holds one: an entry
// measured in a plan of the same classes, else one for any plan.
const cafe_gpu_ctx::TuneEntry* find_tuning(const cafe_gpu_ctx* ctx, int path, int nt,
                                           unsigned long long plan_classes)
{
    const cafe_gpu_ctx::TuneEntry* any = nullptr;
    for (const auto& e : ctx->tuning) {
        if (e.path != path || e.nt != nt || e.leaves != ctx->n_leaves) continue;
        if (e.plan_classes == plan_classes) return &e;
        if (e.plan_classes == 0 && !any) any = &e;
    }
    return any;
}

// The matrix-core shape a tuning entry names for a class of nt tiles whose
// stored vectors have slot_tiles tiles, when that kernel instance exists,
// covers the class (tiles per wave x waves >= nt) and fits a CU (occupancy
// query; LDS partials only where they fit).
bool tuned_mfma_shape(cafe_gpu_ctx* ctx, int path, const cafe_gpu_ctx::TuneEntry& e, int nt, int slot_tiles,
                      MfmaShape& s)
{
    if ((e.kernel != 0 && e.kernel != 2) || path == CAFE_GPU_PATH_PLAIN64 || e.ntw < 1 || e.w < 1 || e.ntw * e.w < nt)
        return false;
    const bool f16 = path == CAFE_GPU_PATH_BATCHED16 && e.kernel == 0;
    if (f16 && nt > kF16MaxTiles) return false;
    const bool fp64 = path == CAFE_GPU_PATH_BATCHED64;
    // 32-family tiles: the FP32 kernel's main pass (BATCHED32, or BATCHED16
    // classes past the split-FP16 kernel's sizes)
    if (e.kernel == 2 && (fp64 || nt > kMaxTilesFb2 || (path == CAFE_GPU_PATH_BATCHED16 && nt <= kF16MaxTiles)))
        return false;
    s = MfmaShape{e.ntw, e.w, e.lds != 0, 0, 0};
    s.fb = e.kernel == 2 ? 2 : 1;
    if (f16) place_partials16(s, ctx->n_slots_lane, slot_tiles, ctx->n_leaves);
    else place_partials(s, fp64, ctx->n_slots, slot_tiles, ctx->n_leaves);
    if (e.lds && !s.lds) return false;   // LDS partials that do not fit
    if (!e.lds && s.lds) {               // global scratch asked for
        s.lds = false;
        s.shm = f16 ? (size_t)f16_meta_bytes(ctx->n_slots_lane, s.w) + (size_t)mfma_count_bytes(ctx->n_leaves)
                    : mfma_meta_only(s, fp64, ctx->n_slots, ctx->n_leaves);
    }
    const PruneParams none{};
    const hipError_t q = f16 ? f16_dispatch(s, 0, none, nullptr, &s.per_cu) : mfma_occupancy(fp64, s, &s.per_cu);
    if (q != hipSuccess || s.per_cu < 1) {
        (void)hipGetLastError();
        return false;   // no such instance, or it does not fit a CU
    }
    s.grid_mult = std::max(0, e.grid_mult);
    return true;
}

// Replaces a class's built-in launch shape by its tuning entry where the entry
// is usable (the kernel itself -- per lane or matrix cores -- was chosen when
// the class's tiles were made); otherwise the class keeps its shape.
int apply_tuning(cafe_gpu_ctx* ctx, int path, KernelClass& kc, unsigned long long plan_classes)
{
    const cafe_gpu_ctx::TuneEntry* e = find_tuning(ctx, path, kc.nt, plan_classes);
    if (!e || path == CAFE_GPU_PATH_PLAIN64) return CAFE_GPU_OK;
    if (kc.lane_rows > 0) {
        if (e->kernel == 1) {
            kc.lane_grid_mult = std::max(0, e->grid_mult);
            kc.tuned = true;
        }
        return CAFE_GPU_OK;
    }
    MfmaShape s;
    if (!tuned_mfma_shape(ctx, path, *e, kc.nt, kc.slot_tiles, s)) return CAFE_GPU_OK;
    if ((e->kernel == 2) != (kc.fb32 == 2)) return CAFE_GPU_OK;   // the tiles were cut for the other width
    if (path == CAFE_GPU_PATH_BATCHED16 && e->kernel == 0) kc.s16 = s;
    else if (path == CAFE_GPU_PATH_BATCHED64) kc.s64 = s;
    else kc.s32 = s;
    kc.tuned = true;
    return CAFE_GPU_OK;
}

// Validates a descriptor and uploads node slots, priors, bucket table and
// category probabilities into b; sizes the per-point outputs.
int upload_eval_inputs(cafe_gpu_ctx* ctx, EvalBufs& b, const cafe_gpu_eval_desc* d, int& P)
{
    if (!d) return fail(CAFE_GPU_EINVAL, "null descriptor");
    if (ctx->n_nodes == 0) return fail(CAFE_GPU_ESTATE, "set_tree has not been called");
    if (ctx->F == 0) return fail(CAFE_GPU_ESTATE, "set_families has not been called");
    if (ctx->N == 0) return fail(CAFE_GPU_ESTATE, "matrix_reset has not been called");
    if (ctx->n_leaves != ctx->tree_leaves)
        return fail(CAFE_GPU_EINVAL, "families have %d leaves, tree has %d", ctx->n_leaves, ctx->tree_leaves);
    if (d->model != CAFE_GPU_MODEL_BASE && d->model != CAFE_GPU_MODEL_GAMMA)
        return fail(CAFE_GPU_EINVAL, "unknown model %d", d->model);
    if (d->path < CAFE_GPU_PATH_PLAIN64 || d->path > CAFE_GPU_PATH_BATCHED16)
        return fail(CAFE_GPU_EINVAL, "unknown path %d", d->path);
    const int K = d->n_cat;
    if (d->n_trials < 1 || K < 1 || (d->model == CAFE_GPU_MODEL_BASE && K != 1))
        return fail(CAFE_GPU_EINVAL, "bad trial/category counts");
    if (!d->node_slot || !d->prior) return fail(CAFE_GPU_EINVAL, "null node_slot or prior");
    if (d->model == CAFE_GPU_MODEL_GAMMA && !d->cat_probs) return fail(CAFE_GPU_EINVAL, "gamma needs cat_probs");
    const int h = ctx->n_dev ? (ctx->n_dev - 1) / 2 : 0;
    int maxR = 0;
    for (auto& bk : ctx->buckets) {
        if (std::max(bk.B, bk.R) + 1 > ctx->N)
            return fail(CAFE_GPU_EINVAL, "bucket B=%d R=%d needs matrices of size %d (pool has %d)", bk.B, bk.R,
                        std::max(bk.B, bk.R) + 1, ctx->N);
        if (ctx->n_dev && (bk.B >= ctx->err_rows || bk.max_count + h > bk.B))
            return fail(CAFE_GPU_EINVAL, "error model does not cover bucket B=%d (max count %d)", bk.B,
                        bk.max_count);
        maxR = std::max(maxR, bk.R);
    }
    if (d->prior_stride < maxR) return fail(CAFE_GPU_EINVAL, "prior_stride %d < max R %d", d->prior_stride, maxR);
    P = d->n_trials * K;
    const long long ns = (long long)P * ctx->n_nodes;
    std::vector<int> slot(d->node_slot, d->node_slot + ns);
    const int root = ctx->ops.back().node;
    for (long long i = 0; i < ns; ++i) {
        const int node = (int)(i % ctx->n_nodes);
        if (node == root) {
            slot[i] = 0;
            continue;
        }
        if (slot[i] < 0 || slot[i] >= ctx->capacity || !ctx->built[slot[i]])
            return fail(CAFE_GPU_ENOTFOUND, "node %d of point %lld uses matrix slot %d, which is not built", node,
                        i / ctx->n_nodes, slot[i]);
        if (ctx->built[slot[i]] == 2 && (d->path == CAFE_GPU_PATH_BATCHED32 || d->path == CAFE_GPU_PATH_BATCHED16))
            return fail(CAFE_GPU_ENOTFOUND,
                        "matrix slot %d holds the FP64 matrix and only the FP32 copies a plan read; build it with "
                        "cafe_gpu_matrix_build for this path", slot[i]);
    }
    const int nb = (int)ctx->buckets.size();
    std::vector<double> pr64((size_t)nb * d->prior_stride, 0.0);
    std::vector<float> pr32(pr64.size());
    std::vector<BucketDev> bd(nb);
    for (int i = 0; i < nb; ++i) {
        const auto& bh = ctx->buckets[i];
        for (int j = 0; j < bh.R; ++j) pr64[(size_t)i * d->prior_stride + j] = d->prior[(size_t)i * d->prior_stride + j];
        bd[i].B = bh.B;
        bd[i].R = bh.R;
        bd[i].L = bh.B + 1;
        bd[i].Lp = round_up(bh.B + 1, kTile);
        bd[i].Rp = round_up(bh.R + 1, kTile);
        bd[i].fam_begin = bh.fam_begin;
        bd[i].count = bh.count;
        bd[i].prior_off = i * d->prior_stride;
    }
    for (size_t i = 0; i < pr64.size(); ++i) pr32[i] = (float)pr64[i];
    std::vector<double> cp(K, 1.0);
    if (d->cat_probs) cp.assign(d->cat_probs, d->cat_probs + K);
    HIP_OR_FAIL(b.node_slot.ensure(sizeof(int) * ns));
    HIP_OR_FAIL(b.prior64.ensure(sizeof(double) * pr64.size()));
    HIP_OR_FAIL(b.prior32.ensure(sizeof(float) * pr32.size()));
    HIP_OR_FAIL(b.buckets.ensure(sizeof(BucketDev) * nb));
    HIP_OR_FAIL(b.catp.ensure(sizeof(double) * K));
    HIP_OR_FAIL(hipMemcpy(b.node_slot.p, slot.data(), sizeof(int) * ns, hipMemcpyHostToDevice));
    HIP_OR_FAIL(hipMemcpy(b.prior64.p, pr64.data(), sizeof(double) * pr64.size(), hipMemcpyHostToDevice));
    HIP_OR_FAIL(hipMemcpy(b.prior32.p, pr32.data(), sizeof(float) * pr32.size(), hipMemcpyHostToDevice));
    HIP_OR_FAIL(hipMemcpy(b.buckets.p, bd.data(), sizeof(BucketDev) * nb, hipMemcpyHostToDevice));
    HIP_OR_FAIL(hipMemcpy(b.catp.p, cp.data(), sizeof(double) * K, hipMemcpyHostToDevice));
    const size_t PF = (size_t)P * ctx->F;
    HIP_OR_FAIL(b.val.ensure(sizeof(double) * PF));
    HIP_OR_FAIL(b.flag.ensure(sizeof(int) * PF));
    const size_t TF = (size_t)d->n_trials * ctx->F;
    HIP_OR_FAIL(b.lnl.ensure(sizeof(double) * TF));
    HIP_OR_FAIL(b.cat.ensure(sizeof(double) * TF * K));
    HIP_OR_FAIL(b.post.ensure(sizeof(double) * TF * K));
    HIP_OR_FAIL(b.fail.ensure(sizeof(int) * TF));
    HIP_OR_FAIL(b.total.ensure(sizeof(double) * d->n_trials));
    return CAFE_GPU_OK;
}

// Builds and uploads the tile list for `n_points` parameter points (heaviest
// classes and buckets first, grouped by kernel size class) and sizes the
// scratch.
int plan_prune(cafe_gpu_ctx* ctx, EvalBufs& b, int path, int model, int n_points, const int* point_ids,
               double* root_out, int root_stride, PrunePlan& pp)
{
    pp = PrunePlan();
    pp.path = path;
    pp.model = model;
    pp.n_cu = ctx->n_cu;
    int maxLR = 0;
    for (auto& bk : ctx->buckets)
        maxLR = std::max(maxLR, std::max(round_up(bk.B + 1, kTile), round_up(bk.R + 1, kTile)));
    pp.maxLR = maxLR;
    auto nq_of = [](const BucketHost& bk) {
        int c = 1;
        while (c * 64 < std::max(bk.B + 1, bk.R)) c <<= 1;
        return c;
    };
    auto nt_of = [](const BucketHost& bk) {   // root rows run 0..R
        return std::max(round_up(bk.B + 1, kTile), round_up(bk.R + 1, kTile)) / kTile;
    };
    std::vector<std::pair<int, int>> bucket_class;   // (class key, bucket)
    for (int bi = 0; bi < (int)ctx->buckets.size(); ++bi) {
        const BucketHost& bk = ctx->buckets[bi];
        // vectors (sizes 0..B) and root rows (sizes 0..R) up to 2,048 rows:
        // PLAIN64 covers 32 x 64; the MFMA shapes take up to 64 tiles of 16
        // and larger vectors run on PLAIN64 in every path
        if (std::max(bk.B + 1, bk.R + 1) > kPlainMaxRows)
            return fail(CAFE_GPU_EINVAL, "bucket B=%d R=%d exceeds the sizes this path supports", bk.B, bk.R);
        bucket_class.push_back({path == CAFE_GPU_PATH_PLAIN64 ? nq_of(bk) : nt_of(bk), bi});
    }
    // heavy classes (and, within a class, larger buckets) first
    std::stable_sort(bucket_class.begin(), bucket_class.end(), [&](const std::pair<int, int>& x,
                                                                   const std::pair<int, int>& y) {
        if (x.first != y.first) return x.first > y.first;
        return ctx->buckets[x.second].B > ctx->buckets[y.second].B;
    });
    // size classes the one-family-per-lane FP32 kernel takes: all vectors and
    // root rows within one of its row counts, and the tree's slots within its
    // instances
    std::map<int, int> key_rows;
    for (auto& bc : bucket_class) {
        const BucketHost& bk = ctx->buckets[bc.second];
        key_rows[bc.first] = std::max(key_rows[bc.first], std::max(bk.B + 1, bk.R + 1));
    }
    auto lane_rows_of = [&](int key) {
        if (path != CAFE_GPU_PATH_BATCHED32 && path != CAFE_GPU_PATH_BATCHED16) return 0;
        const int lr = lane_rows_for(key_rows[key]);
        return lr > 0 && ctx->n_slots_hot + 1 <= lane_max_slots(lr) ? lr : 0;
    };
    // size classes the matrix cores take (those the per-lane kernel does not),
    // and the family-likelihoods of each class
    int n_mfma_keys = 0;
    for (const auto& kr : key_rows)
        if (lane_rows_of(kr.first) == 0) ++n_mfma_keys;
    std::map<int, long long> key_fams;
    for (auto& bc : bucket_class) key_fams[bc.first] += (long long)ctx->buckets[bc.second].count * n_points;
    // the plan's size classes, for the tuning entries measured in a plan
    unsigned long long plan_classes = 0;
    for (const auto& kr : key_rows)
        if (kr.first >= 1 && kr.first <= 64) plan_classes |= 1ull << (kr.first - 1);
    std::vector<TileDesc> tiles;
    std::vector<SegDesc> segs;
    std::vector<int> cls_fams;
    int prev_key = -1;
    for (auto& bc : bucket_class) {
        const BucketHost& bk = ctx->buckets[bc.second];
        if (pp.classes.empty() || bc.first != prev_key) {
            KernelClass kc{};
            kc.begin = kc.end = (int)tiles.size();
            kc.seg_begin = kc.seg_end = (int)segs.size();
            // the kernel is chosen here: its tiles carry 64 families (per
            // lane) or 16 (matrix cores)
            kc.plain = path == CAFE_GPU_PATH_PLAIN64 || bc.first > kMfmaMaxTiles;
            kc.lane_rows = kc.plain ? 0 : lane_rows_of(bc.first);
            kc.fb32 = kc.lane_rows > 0 || kc.plain
                          ? 1
                          : builtin_fb32(path, bc.first, n_mfma_keys == 1, key_fams[bc.first], ctx->n_cu);
            if (const cafe_gpu_ctx::TuneEntry* te = kc.plain ? nullptr : find_tuning(ctx, path, bc.first, plan_classes)) {
                MfmaShape probe;
                if (tuned_mfma_shape(ctx, path, *te, bc.first, bc.first, probe)) {
                    kc.lane_rows = 0;
                    kc.fb32 = probe.fb;
                }
            }
            if (kc.lane_rows > 0) {
                kc.lane_slots = ctx->n_slots_hot + 1;   // the stored slots and the hot array
                const PruneParams none{};
                HIP_OR_FAIL(lane_dispatch(kc.lane_rows, kc.lane_slots, 0, lane_shm(ctx->n_leaves), none, nullptr,
                                          &kc.lane_per_cu));
                if (kc.lane_per_cu == 0) kc.lane_rows = 0;   // no instance fits: matrix-core kernel
            }
            pp.classes.push_back(kc);
            cls_fams.push_back(0);
            prev_key = bc.first;
        }
        KernelClass& kc = pp.classes.back();
        const int ci = (int)pp.classes.size() - 1;
        kc.nt = std::max(kc.nt, nt_of(bk));
        kc.nq = std::max(kc.nq, nq_of(bk));
        kc.slot_tiles = std::max(kc.slot_tiles, round_up(bk.B + 1, kTile) / kTile);
        const int tw = kc.lane_rows > 0 ? 64 : kTile * kc.fb32;   // families per tile
        for (int pi = 0; pi < n_points; ++pi) {
            for (int f0 = 0; f0 < bk.count; f0 += tw)
                tiles.push_back(TileDesc{bc.second, point_ids[pi], f0, std::min(tw, bk.count - f0)});
            for (int f0 = 0; f0 < bk.count; f0 += kSegChunk)
                segs.push_back(SegDesc{bc.second, point_ids[pi], ci, f0});
        }
        kc.r16 += n_points * ((bk.count + kTile - 1) / kTile);
        cls_fams[ci] += bk.count * n_points;
        kc.end = (int)tiles.size();
        kc.seg_end = (int)segs.size();
    }
    const int n_cls = (int)pp.classes.size();
    for (auto& kc : pp.classes) {
        if (kc.plain) {
            pp.has_plain = true;
            continue;
        }
        kc.s32 = kc.fb32 == 2 ? mfma_shape_fb2(kc.nt) : mfma_shape(kc.nt, false);
        kc.s64 = mfma_shape(kc.nt, true);
        kc.r64 = mfma_shape_chain(kc.nt);
        place_partials(kc.s32, false, ctx->n_slots, kc.slot_tiles, ctx->n_leaves);
        place_partials(kc.s64, true, ctx->n_slots, kc.slot_tiles, ctx->n_leaves);
        place_partials(kc.r64, true, ctx->n_slots, kc.slot_tiles, ctx->n_leaves, true);
        HIP_OR_FAIL(mfma_occupancy(false, kc.s32, &kc.s32.per_cu));
        HIP_OR_FAIL(mfma_occupancy(true, kc.s64, &kc.s64.per_cu));
        HIP_OR_FAIL(mfma_occupancy(true, kc.r64, &kc.r64.per_cu));
        if (path == CAFE_GPU_PATH_BATCHED16 && kc.nt <= kF16MaxTiles) {
            // the program whose outputs reuse a child's slot (the store's
            // maxima exchange orders all reads before the writes)
            kc.s16 = mfma_shape(kc.nt, false);
            place_partials16(kc.s16, ctx->n_slots_lane, kc.slot_tiles, ctx->n_leaves);
            const PruneParams none{};
            HIP_OR_FAIL(f16_dispatch(kc.s16, 0, none, nullptr, &kc.s16.per_cu));
            if (kc.s16.per_cu == 0 && kc.s16.lds) {   // LDS does not fit: global scratch
                kc.s16.lds = false;
                kc.s16.shm = (size_t)f16_meta_bytes(ctx->n_slots_lane, kc.s16.w) + (size_t)mfma_count_bytes(ctx->n_leaves);
                HIP_OR_FAIL(f16_dispatch(kc.s16, 0, none, nullptr, &kc.s16.per_cu));
            }
        }
        if (kc.r64.per_cu == 0 && kc.r64.lds) {   // the device does not fit that much LDS: global scratch
            kc.r64.lds = false;
            kc.r64.shm = mfma_meta_only(kc.r64, true, ctx->n_slots, ctx->n_leaves);
            HIP_OR_FAIL(mfma_occupancy(true, kc.r64, &kc.r64.per_cu));
        }
    }
    // Measured shapes from the context's tuning table replace the built-in
    // ones where they exist and fit the class.
    for (auto& kc : pp.classes)
        if (!kc.plain)
            if (int rc = apply_tuning(ctx, path, kc, plan_classes)) return rc;
    // Main-pass classes to side streams, longest first onto the least loaded
    // stream.  Estimated cost: contraction work (tiles x nt^2, plus a per-tile
    // overhead) and a floor for the serial chain of one large tile; a
    // per-lane class counts a quarter (it runs unpadded at the full vector
    // rate beside the matrix-core classes: measured 3.480 -> 3.459 ms per
    // step on the uneven-size workload against the full count, 3.52 at half).
    {
        std::vector<double> cost(n_cls);
        std::vector<int> order(n_cls);
        for (int ci = 0; ci < n_cls; ++ci) {
            const KernelClass& kc = pp.classes[ci];
            const double tiles = kc.r16;   // in 16-family tiles
            cost[ci] = (tiles * (kc.nt * kc.nt + 8.0) + 2500.0 * kc.nt * kc.nt) * (kc.lane_rows > 0 ? 0.25 : 1.0);
            order[ci] = ci;
        }
        std::stable_sort(order.begin(), order.end(), [&](int a, int b2) { return cost[a] > cost[b2]; });
        std::vector<double> load(kSideStreams, 0.0);
        for (int ci : order) {
            const int k = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            pp.classes[ci].stream = k;
            load[k] += cost[ci];
        }
        pp.main_order = order;
        // the stream that frees up first: that of the per-lane classes (at
        // full VALU rate they end well before their estimate), else the one
        // with the least estimated work
        pp.light_stream = (int)(std::min_element(load.begin(), load.begin() + std::max(1, std::min(kSideStreams, n_cls))) -
                                load.begin());
        for (int ci : order)
            if (pp.classes[ci].lane_rows > 0) {
                pp.light_stream = pp.classes[ci].stream;
                break;
            }
    }
    pp.n_tiles = (int)tiles.size();
    pp.class_counts.assign(n_cls + 4, 0);
    for (int i = 0; i < n_cls; ++i) pp.class_counts[i] = pp.classes[i].end - pp.classes[i].begin;
    // [0, n_cls] offsets of the 16-family recompute lists (room for all of a
    // class's families in 16-family tiles), [n_cls + 1, 2 n_cls + 1] of the
    // one-family PLAIN64 lists
    std::vector<int> class_off(2 * (n_cls + 1), 0);
    for (int i = 0; i < n_cls; ++i) {
        pp.classes[i].r_begin = class_off[i];
        class_off[i + 1] = class_off[i] + pp.classes[i].r16;
    }
    pp.r16_total = class_off[n_cls];
    for (int i = 0; i < n_cls; ++i) class_off[n_cls + 1 + i + 1] = class_off[n_cls + 1 + i] + cls_fams[i];
    pp.plain_list_size = class_off[2 * n_cls + 1];
    pp.plain_off.assign(class_off.begin() + n_cls + 1, class_off.end());
    pp.n_segs = (int)segs.size();

    HIP_OR_FAIL(b.tiles.ensure(sizeof(TileDesc) * std::max(1, pp.n_tiles)));
    HIP_OR_FAIL(b.fb_tiles.ensure(sizeof(TileDesc) * ((size_t)std::max(1, pp.r16_total) + std::max(1, pp.plain_list_size))));
    HIP_OR_FAIL(b.ntiles.ensure(sizeof(int) * (n_cls + 4)));
    HIP_OR_FAIL(b.counters.ensure(sizeof(int) * (3 * n_cls + 8)));
    HIP_OR_FAIL(b.fb_count.ensure(sizeof(int) * (2 * n_cls + 4)));
    HIP_OR_FAIL(b.segs.ensure(sizeof(SegDesc) * std::max(1, pp.n_segs)));
    HIP_OR_FAIL(b.class_off.ensure(sizeof(int) * 2 * (n_cls + 1)));
    {
        size_t maxp = 0;
        for (int i = 0; i < n_points; ++i) maxp = std::max(maxp, (size_t)point_ids[i] + 1);
        HIP_OR_FAIL(b.fam_list.ensure(sizeof(int) * maxp * (size_t)ctx->F));
    }
    HIP_OR_FAIL(hipMemcpy(b.tiles.p, tiles.data(), sizeof(TileDesc) * pp.n_tiles, hipMemcpyHostToDevice));
    HIP_OR_FAIL(hipMemcpy(b.ntiles.p, pp.class_counts.data(), sizeof(int) * pp.class_counts.size(),
                          hipMemcpyHostToDevice));
    HIP_OR_FAIL(hipMemcpy(b.segs.p, segs.data(), sizeof(SegDesc) * pp.n_segs, hipMemcpyHostToDevice));
    HIP_OR_FAIL(hipMemcpy(b.class_off.p, class_off.data(), sizeof(int) * 2 * (n_cls + 1), hipMemcpyHostToDevice));

    // Global scratch: per class and pass (classes of one pass run
    // concurrently, passes one after another).  PLAIN64 partials take one
    // region of n_slots x ldv doubles per workgroup unless they fit in LDS.
    const int ldv = round_up(maxLR, 4) + 4;
    pp.per_wg_plain = (long long)ctx->n_slots * ldv;
    const bool plain_in_lds = sizeof(double) * (size_t)ctx->n_slots * ldv <= 48 * 1024;
    auto plain_bytes = [&](int n) {
        return plain_in_lds || n <= 0 ? (size_t)0 : (size_t)pp.grid(1, n) * 8 * (size_t)pp.per_wg_plain;
    };
    size_t main_sz = 0, r64_sz = 0, rp_sz = 0;
    for (int ci = 0; ci < n_cls; ++ci) {
        KernelClass& kc = pp.classes[ci];
        const int n = pp.class_counts[ci];
        size_t mb = 0;
        if (kc.plain) mb = plain_bytes(n);
        else if (path == CAFE_GPU_PATH_BATCHED64)
            mb = mfma_scratch_bytes(kc.s64, true, ctx->n_slots, kc.slot_tiles, pp.grid(kc.s64, n));
        else if (kc.lane_rows > 0) mb = 0;
        else if (path == CAFE_GPU_PATH_BATCHED16 && kc.nt <= kF16MaxTiles) mb = f16_scratch_bytes(kc.s16, ctx->n_slots_lane, kc.slot_tiles, pp.grid(kc.s16, n));
        else mb = mfma_scratch_bytes(kc.s32, false, ctx->n_slots, kc.slot_tiles, pp.grid(kc.s32, n));
        kc.off_main = main_sz;
        main_sz += align256(mb);
        if ((path == CAFE_GPU_PATH_BATCHED32 || path == CAFE_GPU_PATH_BATCHED16) && !kc.plain) {
            kc.off_r64 = r64_sz;
            r64_sz += align256(mfma_scratch_bytes(kc.r64, true, ctx->n_slots, kc.slot_tiles, pp.grid(kc.r64, kc.r16)));
        }
        if (path != CAFE_GPU_PATH_PLAIN64 && !kc.plain) {
            kc.off_rp = rp_sz;
            rp_sz += align256(plain_bytes(pp.plain_off[ci + 1] - pp.plain_off[ci]));
        }
    }
    for (auto& kc : pp.classes) {
        kc.off_r64 += main_sz;
        kc.off_rp += main_sz + r64_sz;
    }
    HIP_OR_FAIL(b.scratch.ensure(std::max<size_t>(main_sz + r64_sz + rp_sz, 256)));

    PruneParams& p = pp.base;
    p.buckets = b.buckets.as<BucketDev>();
    p.ops = ctx->d_ops.as<OpDev>();
    p.n_ops = (int)ctx->ops.size();
    p.kids = ctx->d_kids.as<ChildDev>();
    p.steps = ctx->d_steps.as<StepDev>();
    p.n_steps = ctx->n_steps;
    p.n_leaves = ctx->n_leaves;
    p.n_nodes = ctx->n_nodes;
    p.node_slot = b.node_slot.as<int>();
    p.counts_T = ctx->d_counts.as<int>();
    p.F = ctx->F;
    p.perm = ctx->d_perm.as<int>();
    p.mt64 = ctx->mt64.as<double>();
    p.mt32 = ctx->mt32.as<float>();
    p.band = ctx->band32.as<int>();
    p.band_stride = ctx->Np / kTile;
    p.slot_elems = ctx->slot_elems;
    p.Np = ctx->Np;
    p.leaf64 = ctx->n_dev ? ctx->mt64e.as<double>() : ctx->mt64.as<double>();
    p.leaf32 = ctx->n_dev ? ctx->mt32e.as<float>() : ctx->lf32.as<float>();
    p.cm32 = ctx->lf32.as<float>();
    p.n_dev = ctx->n_dev;
    p.prior64 = b.prior64.as<double>();
    p.prior32 = b.prior32.as<float>();
    p.model = model;
    p.out_val = b.val.as<double>();
    p.out_flag = b.flag.as<int>();
    p.root_out = root_out;
    p.root_stride = root_stride;
    p.fam_list = b.fam_list.as<int>();
    p.vec_exp = 120 / std::max(1, ctx->max_children) - kMatExp;
    // split FP16: a factor is below 2^(kMatExp16 + a) and a node's product of
    // max_children factors stays inside FP32; vectors stay inside FP16
    p.vec_exp16 = std::min(15, 126 / std::max(1, ctx->max_children) - kMatExp16);
    if (path == CAFE_GPU_PATH_BATCHED16) {
        const size_t n16 = sizeof(_Float16) * (size_t)ctx->slot_elems * ctx->capacity;
        HIP_OR_FAIL(ctx->mt16h.ensure(n16));
        HIP_OR_FAIL(ctx->mt16l.ensure(n16));
        pp.n_split = ctx->capacity;
    }
    p.mt16h = ctx->mt16h.as<_Float16>();
    p.mt16l = ctx->mt16l.as<_Float16>();
    p.scratch = b.scratch.p;
    p.ldv = ldv;
    p.n_slots = ctx->n_slots;
    p.zero = 0;
    return CAFE_GPU_OK;
}

// Enqueues the pruning sequence: the main path per size class, then
// (device-side lists, no host round trip) the FP64 recompute of the FP32
// families the flush bound did not certify and the PLAIN64 recompute of
// families near the subnormal range.  late_build (when set) enqueues the
// FP64 pool, which only the recompute stages read: it runs on the side stream
// expected to finish its main-pass work first, after that work.
int enqueue_prune(cafe_gpu_ctx* ctx, EvalBufs& b, const PrunePlan& pp, hipStream_t st, hipEvent_t ev_main,
                  hipEvent_t ev_fb, bool concurrent, const std::vector<hipEvent_t>* cls_ev = nullptr,
                  const std::function<int(hipStream_t)>* late_build = nullptr)
{
    const int n_cls = (int)pp.classes.size();
    int* d_cnt = b.ntiles.as<int>();
    int* d_ctr = b.counters.as<int>();
    int* fbc = b.fb_count.as<int>();
    char* scratch = (char*)pp.base.scratch;
    HIP_OR_FAIL(launch_fill_zero(b.counters.p, sizeof(int) * (3 * n_cls + 8), st));
    HIP_OR_FAIL(launch_fill_zero(b.fb_count.p, sizeof(int) * (2 * n_cls + 4), st));
    // BATCHED16 reads the FP16 halves of the FP32 matrices
    if (pp.path == CAFE_GPU_PATH_BATCHED16)
        HIP_OR_FAIL(launch_split_f16(ctx->mt32.as<float>(), ctx->mt16h.as<_Float16>(), ctx->mt16l.as<_Float16>(),
                                     ctx->Np, ctx->slot_elems, pp.n_split, st));

    // Three stages per size class: the main pass, then (BATCHED32) the FP64
    // recompute of the families the FP32 bound did not certify, then
    // (BATCHED32/64) the PLAIN64 recompute of those near the subnormal range.
    // Each recompute stage packs its class's flagged families into a list
    // sized on the device and prunes it.
    TileDesc* fb64 = b.fb_tiles.as<TileDesc>();
    TileDesc* fbplain = fb64 + std::max(1, pp.r16_total);
    const int* coff = b.class_off.as<int>();
    const int fam_off = pp.fam_count_off();
    const SegDesc* segs = b.segs.as<SegDesc>();
    auto stage_main = [&](int ci, hipStream_t s) -> int {
        const KernelClass& kc = pp.classes[ci];
        const int n = pp.class_counts[ci];
        PruneParams q = pp.base;
        q.tiles = b.tiles.as<TileDesc>() + kc.begin;
        q.n_tiles_dev = d_cnt + ci;
        q.tile_counter = d_ctr + ci;
        q.slot_tiles = kc.slot_tiles;
        q.scratch = scratch + kc.off_main;
        if (kc.plain) {   // final FP64 values: nothing to recompute
            q.scratch_per_wg = pp.per_wg_plain;
            HIP_OR_FAIL(launch_prune_plain(kc.nq, pp.grid(1, n), q, s));
        } else if (pp.path == CAFE_GPU_PATH_BATCHED64) {
            HIP_OR_FAIL(launch_prune_mfma(true, kc.s64, pp.grid(kc.s64, n), q, s));
        } else if (kc.lane_rows > 0) {   // BATCHED32 / BATCHED16: small vectors, one family per lane
            q.steps = ctx->d_steps_hot.as<StepDev>();
            q.n_slots = ctx->n_slots_hot;
            HIP_OR_FAIL(lane_dispatch(kc.lane_rows, kc.lane_slots, pp.grid_lane(kc.lane_per_cu, n, kc.lane_grid_mult),
                                      lane_shm(ctx->n_leaves), q, s, nullptr));
        } else if (pp.path == CAFE_GPU_PATH_BATCHED16 && kc.nt <= kF16MaxTiles) {
            q.steps = ctx->d_steps_lane.as<StepDev>();
            q.n_slots = ctx->n_slots_lane;
            HIP_OR_FAIL(f16_dispatch(kc.s16, pp.grid(kc.s16, n), q, s, nullptr));
        } else {
            HIP_OR_FAIL(launch_prune_mfma(false, kc.s32, pp.grid(kc.s32, n), q, s));
        }
        return CAFE_GPU_OK;
    };
    auto stage_r64 = [&](int ci, hipStream_t s) -> int {
        if (pp.path != CAFE_GPU_PATH_BATCHED32 && pp.path != CAFE_GPU_PATH_BATCHED16) return CAFE_GPU_OK;
        const KernelClass& kc = pp.classes[ci];
        if (kc.plain) return CAFE_GPU_OK;
        HIP_OR_FAIL(launch_pack(segs + kc.seg_begin, kc.seg_end - kc.seg_begin, b.buckets.as<BucketDev>(),
                                pp.base.perm, b.flag.as<int>(), pp.base.F, kFlagRecompute64, b.fam_list.as<int>(),
                                fb64, coff, fbc, fbc + fam_off, kTile, s));
        PruneParams q = pp.base;
        q.tiles = fb64 + kc.r_begin;
        q.n_tiles_dev = fbc + ci;
        q.tile_counter = d_ctr + n_cls + ci;
        q.slot_tiles = kc.slot_tiles;
        q.scratch = scratch + kc.off_r64;
        HIP_OR_FAIL(launch_prune_mfma(true, kc.r64, pp.grid(kc.r64, kc.r16), q, s));
        return CAFE_GPU_OK;
    };
    auto stage_plain = [&](int ci, hipStream_t s) -> int {
        if (pp.path == CAFE_GPU_PATH_PLAIN64) return CAFE_GPU_OK;
        const KernelClass& kc = pp.classes[ci];
        if (kc.plain) return CAFE_GPU_OK;
        HIP_OR_FAIL(launch_pack(segs + kc.seg_begin, kc.seg_end - kc.seg_begin, b.buckets.as<BucketDev>(),
                                pp.base.perm, b.flag.as<int>(), pp.base.F, kFlagRecomputePlain,
                                b.fam_list.as<int>(), fbplain, coff + n_cls + 1, fbc + n_cls, fbc + fam_off + 1, 1,
                                s));
        PruneParams q = pp.base;
        q.tiles = fbplain + pp.plain_off[ci];
        q.n_tiles_dev = fbc + n_cls + ci;
        q.tile_counter = d_ctr + 2 * n_cls + ci;
        q.scratch = scratch + kc.off_rp;
        q.scratch_per_wg = pp.per_wg_plain;
        HIP_OR_FAIL(launch_prune_plain(kc.nq, pp.grid(1, pp.plain_off[ci + 1] - pp.plain_off[ci]), q, s));
        return CAFE_GPU_OK;
    };

    // Stage by stage over the classes: every main pass, then every FP64
    // recompute followed on the same stream by the class's PLAIN64 recompute.
    // When concurrent, the classes of one stage run on the side streams (the
    // main passes on their longest-first assignment, the recomputes round
    // robin in that order: measured 1.5% faster per step than placing the
    // recomputes by their vectors' length) and st waits for all of them
    // before the next stage: a recompute is a latency-bound chain on a few
    // tiles, and beside a main pass it would slow that pass down.
    const int S = concurrent ? std::min((int)ctx->side.size(), n_cls) : 0;
    if (late_build && S <= 1)
        if (int rc = (*late_build)(st)) return rc;
    // launch2 (may be null) follows every launch of the stage on the same
    // streams, before the join: the PLAIN64 pass of a class reads only the
    // flags its own main pass and FP64 pass wrote, on the same stream
    using Launch = std::function<int(int, hipStream_t)>;
    auto stage = [&](int idx, const Launch& launch, const Launch* launch2 = nullptr) -> int {
        if (S <= 1) {
            for (int ci = 0; ci < n_cls; ++ci) {
                if (cls_ev && idx == 0) HIP_OR_FAIL(hipEventRecord((*cls_ev)[ci], st));
                if (int rc = launch(ci, st)) return rc;
            }
            if (cls_ev && idx == 0) HIP_OR_FAIL(hipEventRecord((*cls_ev)[n_cls], st));
            if (launch2)
                for (int ci = 0; ci < n_cls; ++ci)
                    if (int rc = (*launch2)(ci, st)) return rc;
            return CAFE_GPU_OK;
        }
        hipEvent_t* ev = ctx->join_ev.data() + idx * (kSideStreams + 1);
        HIP_OR_FAIL(hipEventRecord(ev[0], st));
        for (int k = 0; k < S; ++k) HIP_OR_FAIL(hipStreamWaitEvent(ctx->side[k], ev[0], 0));
        for (int i = 0; i < n_cls; ++i) {
            const int ci = pp.main_order[i];
            const int k = idx == 0 ? pp.classes[ci].stream % S : i % S;
            if (int rc = launch(ci, ctx->side[k])) return rc;
        }
        if (launch2)
            for (int i = 0; i < n_cls; ++i)
                if (int rc = (*launch2)(pp.main_order[i], ctx->side[i % S])) return rc;
        if (idx == 0 && late_build)
            if (int rc = (*late_build)(ctx->side[pp.light_stream % S])) return rc;
        for (int k = 0; k < S; ++k) {
            HIP_OR_FAIL(hipEventRecord(ev[1 + k], ctx->side[k]));
            HIP_OR_FAIL(hipStreamWaitEvent(st, ev[1 + k], 0));
        }
        return CAFE_GPU_OK;
    };
    if (int rc = stage(0, stage_main)) return rc;
    if (ev_main) HIP_OR_FAIL(hipEventRecord(ev_main, st));
    const Launch plain = stage_plain;
    if (pp.path == CAFE_GPU_PATH_BATCHED32 || pp.path == CAFE_GPU_PATH_BATCHED16) {
        if (int rc = stage(1, stage_r64, &plain)) return rc;
    } else if (pp.path != CAFE_GPU_PATH_PLAIN64) {
        if (int rc = stage(2, stage_plain)) return rc;
    }
    if (ev_fb) HIP_OR_FAIL(hipEventRecord(ev_fb, st));
    return CAFE_GPU_OK;
}

int enqueue_finalize(cafe_gpu_ctx* ctx, EvalBufs& b, int model, int T, int K, hipStream_t st)
{
    HIP_OR_FAIL(launch_finalize(model, T, K, ctx->F, b.val.as<double>(), b.flag.as<int>(), b.catp.as<double>(),
                                b.lnl.as<double>(), b.cat.as<double>(), b.post.as<double>(), b.fail.as<int>(),
                                b.total.as<double>(), st));
    return CAFE_GPU_OK;
}

int copy_results(cafe_gpu_ctx* ctx, EvalBufs& b, const PrunePlan& pp, int model, int T, int K, double* lnl,
                 double* cat, double* post, double* total, long long* n_fallback)
{
    const size_t TF = (size_t)T * ctx->F;
    hipStream_t st = ctx->stream;
    if (lnl) HIP_OR_FAIL(hipMemcpyAsync(lnl, b.lnl.p, sizeof(double) * TF, hipMemcpyDeviceToHost, st));
    if (cat && model == CAFE_GPU_MODEL_GAMMA)
        HIP_OR_FAIL(hipMemcpyAsync(cat, b.cat.p, sizeof(double) * TF * K, hipMemcpyDeviceToHost, st));
    if (post && model == CAFE_GPU_MODEL_GAMMA)
        HIP_OR_FAIL(hipMemcpyAsync(post, b.post.p, sizeof(double) * TF * K, hipMemcpyDeviceToHost, st));
    if (total) HIP_OR_FAIL(hipMemcpyAsync(total, b.total.p, sizeof(double) * T, hipMemcpyDeviceToHost, st));
    int h[2] = {0, 0};
    if (n_fallback)
        HIP_OR_FAIL(hipMemcpyAsync(h, b.fb_count.as<int>() + pp.fam_count_off(), sizeof(int) * 2,
                                   hipMemcpyDeviceToHost, st));
    HIP_OR_FAIL(hipStreamSynchronize(st));
    if (n_fallback) *n_fallback = (long long)h[0] + (long long)h[1];
    return CAFE_GPU_OK;
}

double elapsed_ms(hipEvent_t a, hipEvent_t b)
{
    float ms = 0;
    if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.0;
    return ms;
}

// Pupko reconstruction of families f0 .. f0+nf-1 (cafe_gpu_reconstruct).  The
// device working memory -- program, child lists, parents, prior, a launch's
// family list, its state rows and the two error flags, then per family the
// L, C and product rows -- comes from the context's grow-only buffers when
// `fixed` is null, else from the fixed region [fixed, fixed + fixed_bytes),
// in which case nothing is allocated: launches are sized to fit the region
// and a family that does not fit on its own is rejected.
int reconstruct_impl(cafe_gpu_ctx* ctx, const int* node_slot, const double* prior, int B, int R, int f0, int nf,
                     int* states, char* fixed, size_t fixed_bytes)
{
    if (int rc = check_ctx(ctx)) return rc;
    if (!node_slot || !prior || !states) return fail(CAFE_GPU_EINVAL, "null argument");
    if (ctx->n_nodes == 0 || ctx->F == 0) return fail(CAFE_GPU_ESTATE, "tree and families must be set first");
    if (ctx->N == 0) return fail(CAFE_GPU_ESTATE, "matrix_reset has not been called");
    if (B < 0 || R < 0 || B + 1 > ctx->N) return fail(CAFE_GPU_EINVAL, "bounds B=%d R=%d do not fit matrices of %d", B, R, ctx->N);
    if (f0 < 0 || nf < 0 || f0 + nf > ctx->F) return fail(CAFE_GPU_EINVAL, "family range out of range");
    if (ctx->n_leaves != ctx->tree_leaves)
        return fail(CAFE_GPU_EINVAL, "families have %d leaves, tree has %d", ctx->n_leaves, ctx->tree_leaves);
    if (nf == 0) return CAFE_GPU_OK;
    const int n = ctx->n_nodes;
    const int Rr = std::min(B, R) + 1;   // gene_family_reconstructor.cpp:149-151
    const int Lp = B + 1;
    // device layout: program | kids | parent | prior | families | states | flags,
    // then the per-family L, product and C rows (the program has one record
    // per node, the child lists one entry per non-root node)
    auto up = [](size_t b) { return (b + 255) / 256 * 256; };
    const size_t fam_bytes = (size_t)Lp * ((size_t)n * (sizeof(double) + sizeof(int)) + sizeof(double));
    const size_t o_kids = up(sizeof(ReconNode) * n);
    const size_t o_par = o_kids + up(sizeof(int) * std::max(1, n - 1));
    const size_t o_pri = o_par + up(sizeof(int) * n);
    const size_t o_fam = o_pri + up(sizeof(double) * Rr);
    // families per launch: enough workgroups to fill the device several times
    // over, with the per-family scratch within 2 GB (or within the fixed region)
    size_t cap = std::min<size_t>(8192, std::max<size_t>(64, ((size_t)2 << 30) / fam_bytes));
    if (fixed) {
        // o_fam + up(4 c) + up(4 c n) + 256 + c fam_bytes <= fixed_bytes
        const size_t per = sizeof(int) + sizeof(int) * (size_t)n + fam_bytes;
        const size_t head = o_fam + 3 * 256 + 256;
        const size_t fit = fixed_bytes > head ? (fixed_bytes - head) / per : 0;
        if (fit < 1)
            return fail(CAFE_GPU_EINVAL,
                        "one family of %d nodes with vectors of %d sizes needs %zu bytes of scratch; the fixed "
                        "scratch holds %zu",
                        n, Lp, head + per, fixed_bytes);
        cap = std::min(cap, fit);
    }
    // children first (depth first, children in increasing index)
    std::vector<std::vector<int>> kids(n);
    int root = -1;
    for (int i = 0; i < n; ++i) {
        if (ctx->parent_h[i] < 0) root = i;
        else kids[ctx->parent_h[i]].push_back(i);
    }
    std::vector<ReconNode> prog;
    std::vector<int> kid_list;
    std::vector<std::pair<int, int>> stack{{root, 0}};
    while (!stack.empty()) {
        auto& top = stack.back();
        const int u = top.first;
        if (top.second < (int)kids[u].size()) {
            stack.push_back({kids[u][top.second++], 0});
            continue;
        }
        stack.pop_back();
        ReconNode nd{u, ctx->leaf_col_h[u], (int)kid_list.size(), (int)kids[u].size(), -1};
        for (int c : kids[u]) kid_list.push_back(c);
        if (u != root) {
            nd.slot = node_slot[u];
            if (nd.slot < 0 || nd.slot >= ctx->capacity)
                return fail(CAFE_GPU_EINVAL, "node %d: matrix slot %d outside the pool", u, nd.slot);
            if (!ctx->built[nd.slot])   // matrix_cache::get_matrix throws on a miss (matrix_cache.cpp:88-104)
                return fail(CAFE_GPU_ENOTFOUND, "node %d uses matrix slot %d, which is not built", u, nd.slot);
        }
        prog.push_back(nd);
    }
    // count-row (sorted) index of each caller family
    std::vector<int> sorted_of(ctx->F);
    for (int s = 0; s < ctx->F; ++s) sorted_of[ctx->perm[s]] = s;
    if (prog.size() != (size_t)n || kid_list.size() + 1 != (size_t)n)
        return fail(CAFE_GPU_EINVAL, "reconstruction program does not cover the tree (internal)");
    const int chunk = (int)std::min<size_t>((size_t)nf, cap);
    const size_t o_st = o_fam + up(sizeof(int) * chunk);
    const size_t o_err = o_st + up(sizeof(int) * (size_t)chunk * n);
    const size_t o_scr = o_err + 256;
    char* base;
    char* scr;
    if (fixed) {
        base = fixed;
        scr = fixed + o_scr;
        if (o_scr + (size_t)chunk * fam_bytes > fixed_bytes)
            return fail(CAFE_GPU_EINVAL, "fixed scratch layout overflows (internal)");
    } else {
        HIP_OR_FAIL(ctx->d_recon.ensure(o_scr));
        HIP_OR_FAIL(ctx->d_recon_scratch.ensure((size_t)chunk * fam_bytes));
        base = (char*)ctx->d_recon.p;
        scr = (char*)ctx->d_recon_scratch.p;
    }
    double* L_all = (double*)scr;   // doubles first: every row stays 8-byte aligned
    double* prod_all = L_all + (size_t)chunk * n * Lp;
    int* C_all = (int*)(prod_all + (size_t)chunk * Lp);
    hipStream_t st = ctx->stream;
    HIP_OR_FAIL(hipMemcpyAsync(base, prog.data(), sizeof(ReconNode) * prog.size(), hipMemcpyHostToDevice, st));
    if (!kid_list.empty())
        HIP_OR_FAIL(hipMemcpyAsync(base + o_kids, kid_list.data(), sizeof(int) * kid_list.size(), hipMemcpyHostToDevice, st));
    HIP_OR_FAIL(hipMemcpyAsync(base + o_par, ctx->parent_h.data(), sizeof(int) * n, hipMemcpyHostToDevice, st));
    HIP_OR_FAIL(hipMemcpyAsync(base + o_pri, prior, sizeof(double) * Rr, hipMemcpyHostToDevice, st));
    HIP_OR_FAIL(launch_fill_zero(base + o_err, 2 * sizeof(int), st));
    std::vector<int> fams(chunk);
    for (int c0 = 0; c0 < nf; c0 += chunk) {
        const int m = std::min(chunk, nf - c0);
        for (int k = 0; k < m; ++k) fams[k] = sorted_of[f0 + c0 + k];
        HIP_OR_FAIL(hipMemcpyAsync(base + o_fam, fams.data(), sizeof(int) * m, hipMemcpyHostToDevice, st));
        HIP_OR_FAIL(launch_reconstruct((const ReconNode*)base, (int)prog.size(), (const int*)(base + o_kids),
                                       (const int*)(base + o_par), ctx->mt64.as<double>(), ctx->N, ctx->Np,
                                       ctx->slot_elems, ctx->d_counts.as<int>(), ctx->F, (const int*)(base + o_fam), m,
                                       (const double*)(base + o_pri), B, Rr, L_all, C_all, prod_all, Lp,
                                       (int*)(base + o_st), (int*)(base + o_err), st));
        HIP_OR_FAIL(hipMemcpyAsync(states + (size_t)c0 * n, base + o_st, sizeof(int) * (size_t)m * n,
                                   hipMemcpyDeviceToHost, st));
        HIP_OR_FAIL(hipStreamSynchronize(st));   // fams and the state rows are reused by the next chunk
    }
    int flags[2] = {0, 0};
    HIP_OR_FAIL(hipMemcpyAsync(flags, base + o_err, sizeof(flags), hipMemcpyDeviceToHost, st));
    HIP_OR_FAIL(hipStreamSynchronize(st));
    if (flags[0]) return fail(CAFE_GPU_EINVAL, "a leaf count is outside the matrices (N = %d)", ctx->N);
    if (flags[1]) return fail(CAFE_GPU_EINVAL, "a reconstructed state fell outside 0..%d", B);
    return CAFE_GPU_OK;
}

}  // namespace

// One device-resident evaluation (see cafe_gpu_plan_* in cafe_gpu.h).
struct cafe_gpu_plan {
    cafe_gpu_ctx* ctx = nullptr;
    EvalBufs bufs;
    PrunePlan pp;
    int model = 0, path = 0, T = 0, K = 0, n_keys = 0, algo = 0;
    DevBuf d_lam, d_t;
    DevBuf d_copies;         // per key: the FP32 copies the plan's kernels read (kCopy* bits)
    int copy_top = 0;        // kCopyColTop: rows and columns of the per-lane classes
    long long pool_gen = 0, input_gen = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool graph_ok = false;
    hipEvent_t ev[5] = {};
    std::vector<hipEvent_t> cls_ev;   // around each main-pass class launch (profile runs)
    std::vector<double> cls_ms;       // their times in the last profile run
};

namespace {

// build -> prune -> recompute -> finalize; events bracket the phases when
// `timed` (ev[0] start, ev[1] after build, ev[2] after the main prune,
// ev[3] after the recompute passes, ev[4] after finalize)
int plan_enqueue(cafe_gpu_plan* pl, bool timed)
{
    cafe_gpu_ctx* ctx = pl->ctx;
    hipStream_t st = ctx->stream;
    // BATCHED32 reads the FP64 pool only in its recompute stages: concurrent
    // runs with the recurrence build make the FP32 pools first and the FP64
    // one beside the main pass
    const bool split = !timed && pl->algo != 0 && !pl->pp.has_plain &&
                       (pl->pp.path == CAFE_GPU_PATH_BATCHED32 || pl->pp.path == CAFE_GPU_PATH_BATCHED16);
    if (timed) HIP_OR_FAIL(hipEventRecord(pl->ev[0], st));
    HIP_OR_FAIL(launch_build(pl->algo, pl->d_lam.as<double>(), pl->d_t.as<double>(), ctx->lg.as<double>(), pl->n_keys,
                             0, ctx->N, ctx->Np, ctx->slot_elems, split ? nullptr : ctx->mt64.as<double>(),
                             ctx->mt32.as<float>(), ctx->lf32.as<float>(), st, pl->d_copies.as<unsigned char>(),
                             pl->copy_top, ctx->band32.as<int>()));
    // the folded FP32 leaf pool is read by the reduced paths only
    const bool reduced = pl->pp.path == CAFE_GPU_PATH_BATCHED32 || pl->pp.path == CAFE_GPU_PATH_BATCHED16;
    if (int rc = enqueue_fold(ctx, 0, pl->n_keys, st, !split, reduced)) return rc;
    if (timed) HIP_OR_FAIL(hipEventRecord(pl->ev[1], st));
    const std::function<int(hipStream_t)> build64 = [&](hipStream_t s) -> int {
        HIP_OR_FAIL(launch_build(pl->algo, pl->d_lam.as<double>(), pl->d_t.as<double>(), ctx->lg.as<double>(),
                                 pl->n_keys, 0, ctx->N, ctx->Np, ctx->slot_elems, ctx->mt64.as<double>(), nullptr,
                                 nullptr, s, nullptr, 0, nullptr));
        return enqueue_fold(ctx, 0, pl->n_keys, s, true, false);
    };
    // profile runs issue the size classes one after another (standalone kernel
    // times); executions and the captured graph run them concurrently
    if (int rc = enqueue_prune(ctx, pl->bufs, pl->pp, st, timed ? pl->ev[2] : nullptr, timed ? pl->ev[3] : nullptr,
                               !timed, timed ? &pl->cls_ev : nullptr, split ? &build64 : nullptr))
        return rc;
    if (int rc = enqueue_finalize(ctx, pl->bufs, pl->model, pl->T, pl->K, st)) return rc;
    if (timed) HIP_OR_FAIL(hipEventRecord(pl->ev[4], st));
    return CAFE_GPU_OK;
}

// Captures plan_enqueue into a graph; returns false (stream left clean) if the
// runtime refuses any step of the capture.
bool plan_capture(cafe_gpu_plan* pl, bool timed)
{
    hipStream_t st = pl->ctx->stream;
    if (hipStreamBeginCapture(st, hipStreamCaptureModeRelaxed) != hipSuccess) return false;
    const int rc = plan_enqueue(pl, timed);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(st, &g);
    if (rc != CAFE_GPU_OK || e != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        return false;
    }
    hipGraphExec_t ex = nullptr;
    if (hipGraphInstantiate(&ex, g, nullptr, nullptr, 0) != hipSuccess) {
        (void)hipGraphDestroy(g);
        (void)hipGetLastError();
        return false;
    }
    pl->graph = g;
    pl->exec = ex;
    return true;
}

// A plan's graph and launch parameters point into the context's pool and
// input buffers and were sized for its tree, families and error model.
int plan_current(cafe_gpu_plan* pl)
{
    if (pl->pool_gen != pl->ctx->pool_gen)
        return fail(CAFE_GPU_ESTATE, "matrix pool was reset after the plan was made");
    if (pl->input_gen != pl->ctx->input_gen)
        return fail(CAFE_GPU_ESTATE, "tree, families or error-model shape changed after the plan was made");
    return CAFE_GPU_OK;
}

// Slots 0 .. n_keys-1 hold the plan's matrices once an execution is enqueued:
// the FP64 pool and the FP32 copies the plan reads.
void plan_mark_built(cafe_gpu_plan* pl)
{
    for (int i = 0; i < pl->n_keys; ++i) pl->ctx->built[i] = 2;
}

// The FP32 copies each key's users read in this plan: a leaf branch's
// matrix gives leaf columns (c-major); an interior branch's matrix is an
// MFMA A operand (row-major; also what the split-FP16 halves come from) or,
// in the per-lane classes, columns of the top-left lane_rows x lane_rows
// block (c-major).  FP64 paths read no FP32 copy.
void plan_copies(const cafe_gpu_ctx* ctx, const cafe_gpu_eval_desc* d, const PrunePlan& pp, int P, int n_keys,
                 std::vector<unsigned char>& copies, int& top)
{
    copies.assign(n_keys, 0);
    top = 0;
    if (pp.path != CAFE_GPU_PATH_BATCHED32 && pp.path != CAFE_GPU_PATH_BATCHED16) return;
    bool mfma = false;
    for (const auto& kc : pp.classes) {
        if (kc.lane_rows > 0) top = std::max(top, kc.lane_rows);
        else mfma = true;
    }
    const unsigned char interior = (mfma ? kCopyRow : 0) | (top > 0 ? kCopyColTop : 0);
    const int root = ctx->ops.back().node;
    for (long long i = 0; i < (long long)P * ctx->n_nodes; ++i) {
        const int node = (int)(i % ctx->n_nodes);
        if (node == root) continue;
        const int s = d->node_slot[i];
        if (s >= 0 && s < n_keys) copies[s] |= ctx->leaf_col_h[node] >= 0 ? kCopyCol : interior;
    }
}

}  // namespace

extern "C" {

int cafe_gpu_eval(cafe_gpu_ctx* ctx, const cafe_gpu_eval_desc* d, double* lnl, double* cat, double* post,
                  double* total, long long* n_fallback)
{
    if (int rc = check_ctx(ctx)) return rc;
    EvalBufs& b = ctx->eager;
    int P = 0;
    if (int rc = upload_eval_inputs(ctx, b, d, P)) return rc;
    std::vector<int> pts(P);
    std::iota(pts.begin(), pts.end(), 0);
    PrunePlan pp;
    if (int rc = plan_prune(ctx, b, d->path, d->model, P, pts.data(), nullptr, 0, pp)) return rc;
    HIP_OR_FAIL(hipEventRecord(ctx->ev[0], ctx->stream));
    if (int rc = enqueue_prune(ctx, b, pp, ctx->stream, ctx->ev[1], ctx->ev[2], true)) return rc;
    if (int rc = enqueue_finalize(ctx, b, d->model, d->n_trials, d->n_cat, ctx->stream)) return rc;
    HIP_OR_FAIL(hipEventRecord(ctx->ev[3], ctx->stream));
    if (int rc = copy_results(ctx, b, pp, d->model, d->n_trials, d->n_cat, lnl, cat, post, total, n_fallback))
        return rc;
    ctx->timings[1] = elapsed_ms(ctx->ev[0], ctx->ev[1]);
    ctx->timings[2] = elapsed_ms(ctx->ev[1], ctx->ev[2]);
    ctx->timings[3] = elapsed_ms(ctx->ev[2], ctx->ev[3]);
    return CAFE_GPU_OK;
}

int cafe_gpu_root_vectors(cafe_gpu_ctx* ctx, const cafe_gpu_eval_desc* d, int trial, int cat_k, int f0, int nf,
                          double* out)
{
    if (int rc = check_ctx(ctx)) return rc;
    EvalBufs& b = ctx->eager;
    int P = 0;
    if (int rc = upload_eval_inputs(ctx, b, d, P)) return rc;
    if (trial < 0 || trial >= d->n_trials || cat_k < 0 || cat_k >= d->n_cat)
        return fail(CAFE_GPU_EINVAL, "trial/category out of range");
    if (f0 < 0 || nf < 0 || f0 + nf > ctx->F || !out) return fail(CAFE_GPU_EINVAL, "family range out of range");
    int maxR = 0;
    for (auto& bk : ctx->buckets) maxR = std::max(maxR, bk.R);
    HIP_OR_FAIL(b.root.ensure(sizeof(double) * (size_t)ctx->F * maxR));
    const int point = trial * d->n_cat + cat_k;
    PrunePlan pp;
    if (int rc = plan_prune(ctx, b, d->path, d->model, 1, &point, b.root.as<double>(), maxR, pp)) return rc;
    if (int rc = enqueue_prune(ctx, b, pp, ctx->stream, nullptr, nullptr, true)) return rc;
    std::vector<double> h((size_t)ctx->F * maxR);
    HIP_OR_FAIL(hipMemcpyAsync(h.data(), b.root.p, sizeof(double) * h.size(), hipMemcpyDeviceToHost, ctx->stream));
    HIP_OR_FAIL(hipStreamSynchronize(ctx->stream));
    std::vector<int> famR(ctx->F);
    for (auto& bk : ctx->buckets)
        for (int k = 0; k < bk.count; ++k) famR[ctx->perm[bk.fam_begin + k]] = bk.R;
    size_t o = 0;
    for (int f = f0; f < f0 + nf; ++f)
        for (int j = 0; j < famR[f]; ++j) out[o++] = h[(size_t)f * maxR + j];
    return CAFE_GPU_OK;
}

int cafe_gpu_reconstruct(cafe_gpu_ctx* ctx, const int* node_slot, const double* prior, int B, int R, int f0, int nf,
                         int* states)
{
    return reconstruct_impl(ctx, node_slot, prior, B, R, f0, nf, states, nullptr, 0);
}

int cafe_gpu_reconstruct_noalloc(cafe_gpu_ctx* ctx, const int* node_slot, const double* prior, int B, int R,
                                 int f0, int nf, int* states)
{
    if (int rc = check_ctx(ctx)) return rc;
    if (!ctx->fixed_scratch) return fail(CAFE_GPU_ESTATE, "fixed scratch not resolved for device %d", ctx->device);
    return reconstruct_impl(ctx, node_slot, prior, B, R, f0, nf, states, ctx->fixed_scratch, kFixedScratchBytes);
}

int cafe_gpu_tuning_set(cafe_gpu_ctx* ctx, int path, int tiles, int leaves, int kernel, int ntw, int w, int lds,
                        int grid_mult)
{
    if (!ctx) return fail(CAFE_GPU_EINVAL, "null context");
    if (path < CAFE_GPU_PATH_PLAIN64 || path > CAFE_GPU_PATH_BATCHED16 || tiles < 1 || leaves < 1 || kernel < 0 ||
        kernel > 2 || ntw < 0 || w < 0 || grid_mult < 0)
        return fail(CAFE_GPU_EINVAL, "invalid tuning entry");
    // the new entry applies to any plan: it replaces every entry of the class
    ctx->tuning.erase(std::remove_if(ctx->tuning.begin(), ctx->tuning.end(),
                                     [&](const cafe_gpu_ctx::TuneEntry& e) {
                                         return e.path == path && e.nt == tiles && e.leaves == leaves;
                                     }),
                      ctx->tuning.end());
    ctx->tuning.push_back(cafe_gpu_ctx::TuneEntry{path, tiles, leaves, kernel, ntw, w, lds ? 1 : 0, grid_mult, 0});
    return CAFE_GPU_OK;
}

int cafe_gpu_tuning_clear(cafe_gpu_ctx* ctx)
{
    if (!ctx) return fail(CAFE_GPU_EINVAL, "null context");
    ctx->tuning.clear();
    return CAFE_GPU_OK;
}

int cafe_gpu_tuning_load(cafe_gpu_ctx* ctx, const char* path, int* n_loaded)
{
    if (!ctx || !path) return fail(CAFE_GPU_EINVAL, "null argument");
    FILE* fh = fopen(path, "r");
    if (!fh) return fail(CAFE_GPU_EINVAL, "cannot open tuning file %s", path);
    std::vector<cafe_gpu_ctx::TuneEntry> entries;
    char line[512];
    int lineno = 0;
    while (fgets(line, sizeof(line), fh)) {
        ++lineno;
        char* hash = strchr(line, '#');
        if (hash) *hash = 0;
        cafe_gpu_ctx::TuneEntry e{};
        const int n = sscanf(line, "%d %d %d %d %d %d %d %d", &e.path, &e.nt, &e.leaves, &e.kernel, &e.ntw, &e.w,
                             &e.lds, &e.grid_mult);
        if (n <= 0) continue;   // blank or comment
        bool ok = n == 8 && e.path >= CAFE_GPU_PATH_PLAIN64 && e.path <= CAFE_GPU_PATH_BATCHED16 && e.nt >= 1 &&
                  e.leaves >= 1 && e.kernel >= 0 && e.kernel <= 2 && e.ntw >= 0 && e.w >= 0 && e.grid_mult >= 0;
        // optional "plan=n1,n2,...": the tiles per vector of every size class
        // of the plan the entry was measured in (the entry's own among them)
        if (ok) {
            if (const char* pc = strstr(line, "plan=")) {
                pc += 5;
                for (;;) {
                    char* end = nullptr;
                    const long v = strtol(pc, &end, 10);
                    if (end == pc || v < 1 || v > 64) {
                        ok = false;
                        break;
                    }
                    e.plan_classes |= 1ull << (v - 1);
                    pc = end;
                    if (*pc != ',') break;
                    ++pc;
                }
                ok = ok && (*pc == 0 || isspace((unsigned char)*pc)) && (e.plan_classes >> (e.nt - 1) & 1);
            }
        }
        if (!ok) {
            fclose(fh);
            return fail(CAFE_GPU_EINVAL, "%s:%d: malformed tuning entry", path, lineno);
        }
        e.lds = e.lds ? 1 : 0;
        entries.push_back(e);
    }
    fclose(fh);
    ctx->tuning = entries;
    if (n_loaded) *n_loaded = (int)entries.size();
    return CAFE_GPU_OK;
}

int cafe_gpu_last_timings(cafe_gpu_ctx* ctx, double* ms4)
{
    if (!ctx || !ms4) return fail(CAFE_GPU_EINVAL, "null argument");
    for (int i = 0; i < 4; ++i) ms4[i] = ctx->timings[i];
    return CAFE_GPU_OK;
}

int cafe_gpu_plan_create(cafe_gpu_ctx* ctx, const cafe_gpu_eval_desc* d, int n_keys, const double* lambda_q,
                         const double* t_q, int algo, int use_graph, cafe_gpu_plan** out)
{
    if (int rc = check_ctx(ctx)) return rc;
    if (!out) return fail(CAFE_GPU_EINVAL, "null output pointer");
    if (ctx->N == 0) return fail(CAFE_GPU_ESTATE, "matrix_reset has not been called");
    if (n_keys < 1 || n_keys > ctx->capacity)
        return fail(CAFE_GPU_EINVAL, "%d keys do not fit the pool capacity %d", n_keys, ctx->capacity);
    if (algo != CAFE_GPU_BUILD_CLOSED_FORM && algo != CAFE_GPU_BUILD_RECURRENCE)
        return fail(CAFE_GPU_EINVAL, "unknown build algorithm %d", algo);
    if (!lambda_q || !t_q) return fail(CAFE_GPU_EINVAL, "null key arrays");
    auto* pl = new cafe_gpu_plan();
    pl->ctx = ctx;
    // the plan builds its keys into slots 0 .. n_keys-1 at every execution:
    // they count as built while its inputs are validated, and are marked
    // built in the context only once an execution has been enqueued
    const std::vector<char> built_before = ctx->built;
    for (int i = 0; i < n_keys; ++i) ctx->built[i] = 1;
    int P = 0;
    int rc = upload_eval_inputs(ctx, pl->bufs, d, P);
    ctx->built = built_before;
    std::vector<int> pts(P);
    std::iota(pts.begin(), pts.end(), 0);
    if (rc == CAFE_GPU_OK) rc = plan_prune(ctx, pl->bufs, d->path, d->model, P, pts.data(), nullptr, 0, pl->pp);
    if (rc == CAFE_GPU_OK && (pl->d_lam.ensure(sizeof(double) * n_keys) != hipSuccess ||
                              pl->d_t.ensure(sizeof(double) * n_keys) != hipSuccess))
        rc = fail(CAFE_GPU_ENOMEM, "key upload buffers");
    if (rc != CAFE_GPU_OK) {
        pl->bufs.release_all();
        delete pl;
        return rc;
    }
    pl->model = d->model;
    pl->path = d->path;
    pl->T = d->n_trials;
    pl->K = d->n_cat;
    pl->n_keys = n_keys;
    pl->algo = algo;
    pl->pool_gen = ctx->pool_gen;
    pl->input_gen = ctx->input_gen;
    for (auto& e : pl->ev) (void)hipEventCreate(&e);
    pl->cls_ev.assign(pl->pp.classes.size() + 1, nullptr);
    for (auto& e : pl->cls_ev) (void)hipEventCreate(&e);
    {
        std::vector<unsigned char> copies;
        plan_copies(ctx, d, pl->pp, pl->T * pl->K, n_keys, copies, pl->copy_top);
        if (pl->d_copies.ensure(copies.size()) != hipSuccess ||
            hipMemcpy(pl->d_copies.p, copies.data(), copies.size(), hipMemcpyHostToDevice) != hipSuccess) {
            cafe_gpu_plan_destroy(pl);
            return fail(CAFE_GPU_ENOMEM, "plan copy table");
        }
    }
    *out = pl;
    if ((rc = cafe_gpu_plan_set_keys(pl, lambda_q, t_q)) != CAFE_GPU_OK) {
        cafe_gpu_plan_destroy(pl);
        *out = nullptr;
        return rc;
    }
    if (use_graph) pl->graph_ok = plan_capture(pl, false);   // eager launches if the runtime refuses
    return CAFE_GPU_OK;
}

int cafe_gpu_plan_set_keys(cafe_gpu_plan* pl, const double* lambda_q, const double* t_q)
{
    if (!pl || !lambda_q || !t_q) return fail(CAFE_GPU_EINVAL, "null argument");
    for (int i = 0; i < pl->n_keys; ++i)
        if (!(lambda_q[i] >= 0) || !(t_q[i] >= 0) || !std::isfinite(lambda_q[i]) || !std::isfinite(t_q[i]))
            return fail(CAFE_GPU_EINVAL, "key %d has an invalid lambda or branch length", i);
    HIP_OR_FAIL(hipSetDevice(pl->ctx->device));
    HIP_OR_FAIL(hipStreamSynchronize(pl->ctx->stream));
    HIP_OR_FAIL(hipMemcpy(pl->d_lam.p, lambda_q, sizeof(double) * pl->n_keys, hipMemcpyHostToDevice));
    HIP_OR_FAIL(hipMemcpy(pl->d_t.p, t_q, sizeof(double) * pl->n_keys, hipMemcpyHostToDevice));
    return CAFE_GPU_OK;
}

int cafe_gpu_plan_execute(cafe_gpu_plan* pl)
{
    if (!pl) return fail(CAFE_GPU_EINVAL, "null plan");
    if (int rc = plan_current(pl)) return rc;
    cafe_gpu_ctx* ctx = pl->ctx;
    if (pl->graph_ok) HIP_OR_FAIL(hipGraphLaunch(pl->exec, ctx->stream));
    else if (int rc = plan_enqueue(pl, false)) return rc;
    plan_mark_built(pl);
    return CAFE_GPU_OK;
}

int cafe_gpu_plan_synchronize(cafe_gpu_plan* pl)
{
    if (!pl) return fail(CAFE_GPU_EINVAL, "null plan");
    HIP_OR_FAIL(hipStreamSynchronize(pl->ctx->stream));
    return CAFE_GPU_OK;
}

int cafe_gpu_plan_results(cafe_gpu_plan* pl, double* lnl, double* cat, double* post, double* total,
                          long long* n_fallback)
{
    if (!pl) return fail(CAFE_GPU_EINVAL, "null plan");
    if (int rc = plan_current(pl)) return rc;
    return copy_results(pl->ctx, pl->bufs, pl->pp, pl->model, pl->T, pl->K, lnl, cat, post, total, n_fallback);
}

int cafe_gpu_plan_profile(cafe_gpu_plan* pl, double* ms4, int* graph_ok)
{
    if (!pl || !ms4) return fail(CAFE_GPU_EINVAL, "null argument");
    if (int rc = plan_current(pl)) return rc;
    cafe_gpu_ctx* ctx = pl->ctx;
    HIP_OR_FAIL(hipSetDevice(ctx->device));
    if (int rc = plan_enqueue(pl, true)) return rc;
    plan_mark_built(pl);
    HIP_OR_FAIL(hipStreamSynchronize(ctx->stream));
    if (graph_ok) *graph_ok = pl->graph_ok ? 1 : 0;
    ms4[0] = elapsed_ms(pl->ev[0], pl->ev[1]);
    ms4[1] = elapsed_ms(pl->ev[1], pl->ev[2]);
    ms4[2] = elapsed_ms(pl->ev[2], pl->ev[3]);
    ms4[3] = elapsed_ms(pl->ev[3], pl->ev[4]);
    pl->cls_ms.assign(pl->pp.classes.size(), 0.0);
    for (size_t i = 0; i < pl->pp.classes.size(); ++i) pl->cls_ms[i] = elapsed_ms(pl->cls_ev[i], pl->cls_ev[i + 1]);
    return CAFE_GPU_OK;
}

int cafe_gpu_plan_class_times(cafe_gpu_plan* pl, int cap, int* n_classes, int* class_nt, long long* class_tiles,
                              double* ms, int* launch)
{
    if (!pl || !n_classes) return fail(CAFE_GPU_EINVAL, "null argument");
    const int n = (int)pl->pp.classes.size();
    *n_classes = n;
    if (pl->cls_ms.size() != (size_t)n) return fail(CAFE_GPU_ESTATE, "cafe_gpu_plan_profile has not been run");
    for (int i = 0; i < n && i < cap; ++i) {
        if (class_nt) class_nt[i] = pl->pp.classes[i].nt;
        if (class_tiles) class_tiles[i] = pl->pp.class_counts[i];
        if (ms) ms[i] = pl->cls_ms[i];
        if (launch) {
            const KernelClass& kc = pl->pp.classes[i];
            const int n = pl->pp.class_counts[i];
            int* l = launch + 6 * i;
            if (kc.plain) {
                l[0] = 2;
                l[1] = kc.nq;
                l[2] = 1;
                l[3] = 0;
                l[4] = pl->pp.grid(1, n);
                l[5] = 1;
            } else if ((pl->path == CAFE_GPU_PATH_BATCHED32 || pl->path == CAFE_GPU_PATH_BATCHED16) &&
                       kc.lane_rows > 0) {
                l[0] = 1;
                l[1] = kc.lane_rows;
                l[2] = 1;
                l[3] = 0;
                l[4] = pl->pp.grid_lane(kc.lane_per_cu, n, kc.lane_grid_mult);
                l[5] = 64;
            } else {
                const bool f16 = pl->path == CAFE_GPU_PATH_BATCHED16 && kc.nt <= kF16MaxTiles;
                const MfmaShape& sh = pl->path == CAFE_GPU_PATH_BATCHED64 ? kc.s64 : f16 ? kc.s16 : kc.s32;
                l[0] = f16 ? 3 : 0;
                l[1] = sh.ntw;
                l[2] = sh.w;
                l[3] = sh.lds ? 1 : 0;
                l[4] = pl->pp.grid(sh, n);
                l[5] = kTile * sh.fb;
            }
        }
    }
    return CAFE_GPU_OK;
}

int cafe_gpu_plan_recompute_counts(cafe_gpu_plan* pl, long long* fp64, long long* plain)
{
    if (!pl) return fail(CAFE_GPU_EINVAL, "null plan");
    int h[2] = {0, 0};
    HIP_OR_FAIL(hipSetDevice(pl->ctx->device));
    HIP_OR_FAIL(hipMemcpyAsync(h, pl->bufs.fb_count.as<int>() + pl->pp.fam_count_off(), sizeof(int) * 2,
                               hipMemcpyDeviceToHost, pl->ctx->stream));
    HIP_OR_FAIL(hipStreamSynchronize(pl->ctx->stream));
    if (fp64) *fp64 = h[0];
    if (plain) *plain = h[1];
    return CAFE_GPU_OK;
}

int cafe_gpu_plan_destroy(cafe_gpu_plan* pl)
{
    if (!pl) return CAFE_GPU_OK;
    (void)hipSetDevice(pl->ctx->device);
    (void)hipStreamSynchronize(pl->ctx->stream);
    if (pl->exec) (void)hipGraphExecDestroy(pl->exec);
    if (pl->graph) (void)hipGraphDestroy(pl->graph);
    for (auto& e : pl->ev)
        if (e) (void)hipEventDestroy(e);
    for (auto& e : pl->cls_ev)
        if (e) (void)hipEventDestroy(e);
    pl->bufs.release_all();
    pl->d_lam.release();
    pl->d_t.release();
    pl->d_copies.release();
    delete pl;
    return CAFE_GPU_OK;
}

}  // extern "C"
