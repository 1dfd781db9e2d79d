/*
 * cafe_gpu.h -- C ABI of the MI355X likelihood engine for CAFE5's hot path.
 *
 * The engine replaces, for one GPU:
 *   - matrix_cache::precalculate_matrices / get_matrix
 *       (src/matrix_cache.h:76-77, src/matrix_cache.cpp:88-163)
 *   - inference_prune + compute_node_probability for every family
 *       (src/core.cpp:134-145, src/probability.cpp:175-330)
 *   - the per-family root reductions of base_model / gamma_model
 *       (src/base_model.cpp:53-100, src/gamma_core.cpp:97-237)
 *
 * Conventions
 *   - Every call returns an int status (CAFE_GPU_OK == 0).  On failure
 *     cafe_gpu_last_error() returns a message for the calling thread.
 *   - The caller owns all host buffers; the context owns device memory and
 *     one HIP stream.  Calls on one context are not thread safe.
 *   - Matrices follow matrix_cache semantics: keys are (lambda, t) already
 *     quantized and de-quantized as matrix_cache_key does
 *     (src/matrix_cache.h:44-63); entries are P(s -> c) of the linear
 *     birth-death process, row 0 = delta_0, saturated keys (1-2a < 0) give a
 *     zero matrix except (0,0).
 *   - A tree has n_nodes nodes; parent[i] == -1 marks the root; leaf_col[i]
 *     is a leaf's column in a family's count row (-1 for interior nodes);
 *     a node's children are taken in increasing node index, which the caller
 *     makes equal to the reference's descendant order.
 */
#ifndef CAFE_GPU_H
#define CAFE_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

enum {
    CAFE_GPU_OK = 0,
    CAFE_GPU_EINVAL = 1,     /* bad argument or shape */
    CAFE_GPU_ENOMEM = 2,     /* device allocation failed */
    CAFE_GPU_EDEVICE = 3,    /* HIP runtime / kernel launch error */
    CAFE_GPU_ESTATE = 4,     /* call order (e.g. eval before set_tree) */
    CAFE_GPU_ENOTFOUND = 5   /* matrix slot not built (get_matrix miss) */
};

/* Matrix build algorithms. */
enum {
    /* probability.cpp:104-148 term by term: log-binomials, exp, pow, summed
     * in ascending j, clamped to [0,1].  O(N^3) per matrix. */
    CAFE_GPU_BUILD_CLOSED_FORM = 0,
    /* The same matrix from the generating function of the process,
     * G_s(z) = G_1(z)^s with G_1 = (a + (1-2a) z)/(1 - a z):
     * P_s(c) = a P_{s-1}(c) + (1-2a) P_{s-1}(c-1) + a P_s(c-1).
     * All terms are positive, so the result agrees with the closed form to
     * ~1e-13 relative.  O(N^2) per matrix. */
    CAFE_GPU_BUILD_RECURRENCE = 1
};

/* Pruning paths. */
enum {
    /* per-family FP64 matrix-vector products in the reference's summation
     * order (the drop-in); no rescaling */
    CAFE_GPU_PATH_PLAIN64 = 0,
    /* families as columns, FP64 MFMA contractions; no rescaling */
    CAFE_GPU_PATH_BATCHED64 = 1,
    /* families as columns, FP32 MFMA contractions with per-family
     * power-of-two rescaling and an exponent carry; each family gets a bound
     * on what FP32 flushing could have lost, and families whose bound is not
     * below 1e-9 of their likelihood (or whose FP64 value would be
     * subnormal) are recomputed on the BATCHED64 path */
    CAFE_GPU_PATH_BATCHED32 = 2,
    /* families as columns, split-FP16 MFMA contractions (each operand as two
     * FP16 halves, three products with FP32 accumulation, about FP32
     * accuracy for in-range values) with the same per-family rescaling and
     * exponent carry; every partial vector carries a per-entry bound on what
     * FP16's narrow range can lose, propagated by a fourth MFMA, and families
     * whose bound is not below 1e-9 of their likelihood are recomputed on the
     * BATCHED64 path.  Vectors of at most 33 rows run the FP32 per-lane
     * kernel of BATCHED32 (no matrix-core work there). */
    CAFE_GPU_PATH_BATCHED16 = 3
};

enum {
    CAFE_GPU_MODEL_BASE = 0,   /* base_model: lnL_f = max_j log L_j + log prior_j */
    CAFE_GPU_MODEL_GAMMA = 1   /* gamma_model: K rate categories */
};

typedef struct cafe_gpu_ctx cafe_gpu_ctx;

/* --- context --------------------------------------------------------- */

/* Creates a context on HIP device `device` (no counterpart in the reference,
 * which has no device state). */
int cafe_gpu_create(int device, cafe_gpu_ctx **out);
int cafe_gpu_destroy(cafe_gpu_ctx *ctx);
const char *cafe_gpu_last_error(void);

/* --- transition matrices (matrix_cache) --------------------------------- */

/* Replaces matrix_cache(int N) (src/matrix_cache.h:91): sets the matrix
 * dimension N = max(R, B) + 1 and the pool capacity in matrices; drops all
 * built matrices.  2 <= N <= 2047 (the log-gamma table holds 0..4095).  On
 * every path, size classes whose vectors exceed 1,024 rows (64 tiles of 16)
 * run their main pass on the PLAIN64 kernel (final FP64 values, the
 * reference's summation order), the role of the reference's heap-allocated
 * path for large families (src/probability.cpp:239-310). */
int cafe_gpu_matrix_reset(cafe_gpu_ctx *ctx, int N, int capacity);

/* Replaces matrix_cache::precalculate_matrices (src/matrix_cache.cpp:113-163):
 * builds n matrices at the de-quantized keys (lambda_q[i], t_q[i]) into slots
 * first_slot .. first_slot + n - 1 with the given algorithm. */
int cafe_gpu_matrix_build(cafe_gpu_ctx *ctx, int n, const double *lambda_q,
                          const double *t_q, int first_slot, int algo);

/* Replaces matrix_cache::get_matrix (src/matrix_cache.cpp:88-105) for host
 * consumers: copies slot's N x N matrix, row-major out[s*N + c] = P(s -> c). */
int cafe_gpu_matrix_get(cafe_gpu_ctx *ctx, int slot, double *out);

/* --- model inputs -------------------------------------------------------- */

/* The species tree (clade, src/clade.h:25-121). */
int cafe_gpu_set_tree(cafe_gpu_ctx *ctx, int n_nodes, const int *parent,
                      const int *leaf_col);

/* Gene families (gene_family, src/gene_family.h): counts is F x n_leaves
 * row-major.  bound_B[f] / bound_R[f] are the family's max_family_size and
 * max_root_family_size (NULL: B / R for all).  Families with equal (B, R)
 * form one size bucket. */
int cafe_gpu_set_families(cafe_gpu_ctx *ctx, int F, int n_leaves, const int *counts,
                          const int *bound_B, const int *bound_R, int B, int R);

/* Number of size buckets and, for bucket i, its (B, R) and family count.
 * Buckets are numbered in ascending (B, R). */
int cafe_gpu_bucket_count(cafe_gpu_ctx *ctx, int *n);
int cafe_gpu_bucket_info(cafe_gpu_ctx *ctx, int i, int *B, int *R, int *count);

/* Leaf error model (error_model::get_probs, src/error_model.cpp:52-57):
 * rows[k * n_dev + i] = probability of deviation i - (n_dev-1)/2 for an
 * observed count k, k = 0 .. n_rows-1 (n_rows must exceed every family's B).
 * rows == NULL removes the error model. */
int cafe_gpu_set_error_model(cafe_gpu_ctx *ctx, int n_dev, int n_rows, const double *rows);

/* --- evaluation (infer_family_likelihoods) --------------------------------- */

typedef struct {
    int model;            /* CAFE_GPU_MODEL_* */
    int path;             /* CAFE_GPU_PATH_* */
    int n_trials;         /* T parameter points evaluated together */
    int n_cat;            /* K categories (1 for the base model) */
    /* node_slot[(t * K + k) * n_nodes + i]: matrix slot of the branch above
     * node i at trial t, category k (the root's entry is ignored) */
    const int *node_slot;
    /* cat_probs[k] (gamma only) */
    const double *cat_probs;
    /* prior[b * prior_stride + j] = double(root_equilibrium_distribution::
     * compute(j)) for bucket b, j < R_b (base_model.cpp:84, gamma_core.cpp:156) */
    const double *prior;
    int prior_stride;
} cafe_gpu_eval_desc;

/* One likelihood evaluation per trial point for every family
 * (base_model.cpp:53-100 / gamma_core.cpp:168-237).  Outputs, all optional:
 *   lnl[t * F + f]          per-family log-likelihood (base: the stash's
 *                           posterior_probability; gamma: log of the family
 *                           likelihood)
 *   cat[(t * F + f) * K + k] gamma category likelihood (max_j L_j prior_j) p_k
 *   post[(t * F + f) * K + k] gamma posterior (p_k applied twice, as
 *                           gamma_core.cpp:102)
 *   total[t]                -sum_f lnl (compensated sum), +inf when any
 *                           gamma family has a zero-sum root vector
 *   n_fallback              families recomputed on the FP64 path (BATCHED32)
 */
int cafe_gpu_eval(cafe_gpu_ctx *ctx, const cafe_gpu_eval_desc *desc, double *lnl,
                  double *cat, double *post, double *total, long long *n_fallback);

/* Root vectors of one trial/category for families f0 .. f0+nf-1 (FP64
 * paths only): out[(f - f0) * R_f + j] = L_root[j] as inference_prune
 * returns it (src/core.cpp:134-145).  Used by host consumers and tests. */
int cafe_gpu_root_vectors(cafe_gpu_ctx *ctx, const cafe_gpu_eval_desc *desc, int trial,
                          int cat, int f0, int nf, double *out);

/* --- ancestral reconstruction -------------------------------------------- */

/* Pupko's joint reconstruction (pupko_reconstructor::reconstruct_gene_family,
 * src/gene_family_reconstructor.cpp:30-189) of families f0 .. f0+nf-1 with
 * vectors of sizes 0..B and root sizes 1..min(B, R):
 *   node_slot[i]  matrix slot of the branch above node i (root ignored)
 *   prior[j]      double(root_equilibrium_distribution::compute(j)),
 *                 j = 0 .. min(B, R) (the root scores size j with compute(j))
 *   states[(f - f0) * n_nodes + i]  the reconstructed size of node i (a
 *                 leaf's own count)
 * Uses the FP64 matrices and the reference's order of every product, so the
 * states are the reference's.  The caller rejects all-zero matrices of
 * interior branches as the reference does (:84-85). */
int cafe_gpu_reconstruct(cafe_gpu_ctx *ctx, const int *node_slot, const double *prior, int B, int R,
                         int f0, int nf, int *states);

/* The same reconstruction with no device allocation (no counterpart in the
 * reference): for callers that cannot allocate device memory, such as code
 * under stream capture or a serving loop that forbids hipMalloc.  Arguments,
 * results and errors are those of cafe_gpu_reconstruct; the working memory is
 * a fixed 32 MiB device array compiled into the library (one per device, held
 * until the process ends) instead of the context's grow-only buffers, so
 * families run in launches sized to fit it.  A family whose rows alone need
 * more (nodes x (B + 1) x 12 bytes, about 32 MiB) is rejected with
 * CAFE_GPU_EINVAL before anything runs.  Every call on one device shares that
 * array: calls that overlap on one device, from any context or thread, race.
 * The inputs it reads (tree, families, matrices) are the context's, set up
 * beforehand by the calls that allocate. */
int cafe_gpu_reconstruct_noalloc(cafe_gpu_ctx *ctx, const int *node_slot, const double *prior, int B,
                                 int R, int f0, int nf, int *states);

/* --- device-resident plans (no reference counterpart) ----------------------
 * A plan keeps everything an evaluation needs on the device and replays it
 * as one HIP graph: matrix build for n_keys keys into pool slots
 * 0 .. n_keys-1, pruning on the descriptor's path (with its recompute passes),
 * the model reduction and the totals.  It serves the repeated calls of the
 * optimizer scorer (src/optimizer_scorer.cpp:22-36) and lambda-grid scans:
 * only the key values change between executions (plan_set_keys).  The node
 * slots of the descriptor must refer to slots < n_keys.  use_graph = 0 (or a
 * runtime that refuses the capture) runs the same launches eagerly.  A plan's
 * build writes each slot's FP64 matrix and only the FP32 copies the plan's
 * own kernels read (none on the FP64 paths); after an execution the FP64
 * paths, cafe_gpu_matrix_get and the reconstructions may use those slots as
 * they are, while an eager BATCHED32 or BATCHED16 evaluation of them returns
 * CAFE_GPU_ENOTFOUND until cafe_gpu_matrix_build rebuilds them. */
typedef struct cafe_gpu_plan cafe_gpu_plan;

int cafe_gpu_plan_create(cafe_gpu_ctx *ctx, const cafe_gpu_eval_desc *desc, int n_keys,
                         const double *lambda_q, const double *t_q, int algo, int use_graph,
                         cafe_gpu_plan **out);
/* New key values (same count) for the next execution. */
int cafe_gpu_plan_set_keys(cafe_gpu_plan *plan, const double *lambda_q, const double *t_q);
/* Enqueues one execution on the context stream and returns. */
int cafe_gpu_plan_execute(cafe_gpu_plan *plan);
int cafe_gpu_plan_synchronize(cafe_gpu_plan *plan);
/* Waits and copies the last execution's outputs (layouts as cafe_gpu_eval). */
int cafe_gpu_plan_results(cafe_gpu_plan *plan, double *lnl, double *cat, double *post,
                          double *total, long long *n_fallback);
/* Runs one execution with the plan's launches issued directly (not from the
 * graph) and HIP events between the phases; returns their times in ms:
 * [0] matrix build, [1] main prune, [2] recompute passes, [3] finalize.
 * *graph_ok = 1 when plan_execute replays a captured graph. */
int cafe_gpu_plan_profile(cafe_gpu_plan *plan, double *ms4, int *graph_ok);
/* Families of the last execution recomputed on BATCHED64 (FP32 results the
 * flush bound did not certify) and on PLAIN64 (FP64 values near the
 * subnormal range, redone in the reference's summation order). */
int cafe_gpu_plan_recompute_counts(cafe_gpu_plan *plan, long long *fp64, long long *plain);
/* Per size class of the main pruning pass, from the last
 * cafe_gpu_plan_profile run (no reference counterpart; measurement): the
 * class's 16-row tile count per vector, its number of tiles (of the
 * families per tile given in launch) and the milliseconds of its kernel
 * launch (HIP events on the context stream),
 * and, in launch[6 i .. 6 i + 5], that launch's shape: kernel (0 matrix-core
 * k_prune_mfma, 1 one-family-per-lane k_prune_lane, 2 PLAIN64 k_prune_plain),
 * 16-row tiles per wave (lane and PLAIN64 kernels: rows per lane), waves per
 * workgroup, partial vectors in LDS (0/1), workgroups, and families per tile
 * (matrix-core kernel: 16, or 32 with two 16-family blocks per operand).
 * Writes at most cap entries (any output may be null); *n_classes is the
 * number of classes. */
int cafe_gpu_plan_class_times(cafe_gpu_plan *plan, int cap, int *n_classes, int *class_nt,
                              long long *class_tiles, double *ms, int *launch);
int cafe_gpu_plan_destroy(cafe_gpu_plan *plan);

/* --- tile shapes (no reference counterpart) --------------------------------
 * Plans and evaluations pick a launch shape for every size class (the
 * buckets whose vectors have the same number of 16-row tiles): the
 * one-family-per-lane kernel or the matrix-core kernel, and for the latter
 * its families per tile (16, or 32 for the FP32 kernel on vectors of up to 32
 * tiles: each matrix operand then feeds two 16-family blocks), output tiles
 * per wave, waves per workgroup, partials in LDS or global scratch, and the
 * grid (workgroups per resident slot).  The context holds a
 * table of measured choices keyed by (path, tiles per vector, leaves of the
 * tree) that overrides the built-in rules; cafe5_amd/tune.py writes it as a
 * cache file on the GPU, one line per entry:
 *     path tiles leaves kernel ntw w lds grid_mult [measured ms] [plan=n1,n2,...]
 * (kernel 0 matrix core, 1 per lane, 2 matrix core with 32-family tiles;
 * '#' starts a comment).  plan= lists the tiles per vector of every size
 * class of the plan the entry was measured in (the classes run concurrently,
 * so a class's best shape depends on the others): such an entry applies only
 * to plans with exactly those classes, and is preferred there to an entry
 * without the list, which applies to any plan.  An entry whose
 * shape does not exist or does not fit the class is ignored, and the class
 * keeps its built-in shape. */
/* Replaces the table with the entries of a cache file; *n_loaded (may be
 * NULL) receives their number. */
int cafe_gpu_tuning_load(cafe_gpu_ctx *ctx, const char *path, int *n_loaded);
/* Adds an entry for any plan, replacing the entries of the same path, tiles
 * and leaves. */
int cafe_gpu_tuning_set(cafe_gpu_ctx *ctx, int path, int tiles, int leaves, int kernel, int ntw, int w,
                        int lds, int grid_mult);
/* Empties the table (built-in shapes only). */
int cafe_gpu_tuning_clear(cafe_gpu_ctx *ctx);

/* --- timing (measurement support; no reference counterpart) --------------- */

/* Milliseconds of the last eval's phases, measured with HIP events on the
 * context stream: [0] matrix build (last cafe_gpu_matrix_build), [1] main
 * prune kernel, [2] FP64 recompute, [3] finalize + totals. */
int cafe_gpu_last_timings(cafe_gpu_ctx *ctx, double *ms4);

#ifdef __cplusplus
}
#endif
#endif
