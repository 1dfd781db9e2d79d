/*
 * cafe_oracle.c -- FP64 CPU restatement of CAFE5's likelihood hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker); see cafe_oracle.h for the contract
 * and the list of reference files followed.  Built with -ffp-contract=off so
 * that every multiply and add rounds separately, as the reference's x86-64
 * build does.
 */
#include "cafe_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* probability.cpp:59 caches lgamma of integers below 1024 and calls lgamma()
 * directly above; both give glibc's lgamma(i), so one larger table holds the
 * same values. */
#define ORC_LG_N 4096
static double orc_lg[ORC_LG_N];
static int orc_ready = 0;

void orc_init(void)
{
    for (int i = 0; i < ORC_LG_N; ++i)
        orc_lg[i] = lgamma((double)i);
    orc_ready = 1;
}

/* probability.cpp:61-67 (lgamma2) */
static double orc_lgamma2(double n)
{
    if (n >= 0 && n < ORC_LG_N && (n - (int)n < 0.00000000001))
        return orc_lg[(int)n];
    return lgamma(n);
}

/* probability.cpp:82-91.  The 100x100 chooseln table of the reference holds
 * the same three-term expression, so it is not kept separately. */
double orc_chooseln(double n, double r)
{
    if (r == 0 || (n == 0 && r == 0))
        return 0;
    else if (n <= 0 || r <= 0)
        return log(0.0);
    return orc_lgamma2(n + 1) - orc_lgamma2(r + 1) - orc_lgamma2(n - r + 1);
}

/* probability.cpp:104-148: sum over j of exp(log-binomials + (s+c-2j) log a)
 * times coeff^j, summed in ascending j from 0.0, clamped to [0, 1]. */
double orc_birthdeath(int s, int c, double log_alpha, double coeff)
{
    int m = c < s ? c : s;
    int s_add_c = s + c;
    int s_add_c_sub_1 = s_add_c - 1;
    int s_sub_1 = s - 1;
    double result = 0.0;
    for (int j = 0; j <= m; j++) {
        double t = orc_chooseln(s, j) + orc_chooseln(s_add_c_sub_1 - j, s_sub_1)
                 + (s_add_c - 2 * j) * log_alpha;
        double e = exp(t);
        double p = e * pow(coeff, (double)j);
        result += p;
    }
    if (result > 1.0) result = 1.0;
    if (result < 0.0) result = 0.0;
    return result;
}

/* probability.cpp:150-167 */
double orc_transition(double lambda, double t, int parent_size, int child_size)
{
    double alpha = lambda * t / (1 + lambda * t);
    double coeff = 1 - 2 * alpha;
    double result = 0;
    if (coeff > 0 && coeff != 1)
        result = orc_birthdeath(parent_size, child_size, log(alpha), coeff);
    return result;
}

/* matrix_cache.cpp:107-111 */
int orc_is_saturated(double t, double lambda)
{
    double alpha = lambda * t / (1 + lambda * t);
    return (1 - 2 * alpha) < 0;
}

/* matrix_cache.h:49-62: truncating quantization, then de-quantization. */
double orc_key_lambda(double lambda)
{
    long q = (long)(lambda * 1000000000);
    return (double)q / 1000000000.0;
}

double orc_key_branch_length(double t)
{
    long q = (long)(t * 1000);
    return (double)q / 1000.0;
}

/* matrix_cache.cpp:78-85 */
static double orc_from_parent(double lambda, double t, int s, int c)
{
    if (s == 0)
        return c == 0 ? 1.0 : 0.0;
    return orc_transition(lambda, t, s, c);
}

/* One key of matrix_cache.cpp:113-163. */
static void orc_build_rows(double lambda_q, double t_q, int N, int s_lo, int s_hi,
                           double *out)
{
    int sat = orc_is_saturated(t_q, lambda_q);
    for (int s = s_lo; s < s_hi; ++s) {
        double *row = out + (size_t)s * N;
        for (int c = 0; c < N; ++c)
            row[c] = 0.0;
    }
    if (s_lo == 0) {
        out[0] = orc_from_parent(lambda_q, t_q, 0, 0);
        if (!sat)
            for (int c = 0; c < N; ++c)
                out[c] = orc_from_parent(lambda_q, t_q, 0, c);
    }
    if (sat)
        return;
    for (int s = s_lo > 1 ? s_lo : 1; s < s_hi; ++s)
        for (int c = 0; c < N; ++c)
            out[(size_t)s * N + c] = orc_from_parent(lambda_q, t_q, s, c);
}

void orc_build_matrix(double lambda_q, double t_q, int N, double *out)
{
    if (!orc_ready)
        orc_init();
    orc_build_rows(lambda_q, t_q, N, 0, N, out);
}

/* ------------------------------------------------------------------------ */
/* Pruning                                                                  */
/* ------------------------------------------------------------------------ */

/* matrix_cache.cpp:50-56: result[s-s_min] = sum_{c} M[s][c] v[c-c_min],
 * ascending c, accumulated from 0. */
void orc_matvec(const double *M, int N, const double *v, int s_min, int s_max,
                int c_min, int c_max, double *result)
{
    for (int s = s_min; s <= s_max; s++) {
        double acc = 0;
        for (int c = c_min; c <= c_max; c++)
            acc += M[(size_t)s * N + c] * v[c - c_min];
        result[s - s_min] = acc;
    }
}

int orc_leaf_vector(int size, const double *err, int n_dev, int B, double *v)
{
    if (size < 0 || size > B)
        return -1;
    if (err) {
        /* offset = count - (n_dev - 1) / 2; negative indices skipped, no
         * upper-bound check in the reference (count <= B - 1 required). */
        int offset = size - (n_dev - 1) / 2;
        const double *row = err + (size_t)size * n_dev;
        for (int i = 0; i < n_dev; ++i) {
            if (offset + i < 0)
                continue;
            if (offset + i > B)
                return -1;
            v[offset + i] = row[i];
        }
    } else {
        v[size] = 1.0;
    }
    return 0;
}

static int orc_find_root(const orc_tree *tree)
{
    int root = -1;
    for (int i = 0; i < tree->n_nodes; ++i)
        if (tree->parent[i] < 0) {
            if (root >= 0)
                return -1;
            root = i;
        }
    return root;
}

/* Children-before-parent order (the reference walks reverse level order,
 * clade.cpp:69-100; any order that finishes children first gives the same
 * values because each node only reads its children). */
static int orc_postorder(const orc_tree *tree, int *order)
{
    int n = tree->n_nodes;
    int *depth = (int *)malloc(sizeof(int) * n);
    int maxd = 0;
    for (int i = 0; i < n; ++i) {
        int d = 0, p = tree->parent[i];
        while (p >= 0) {
            ++d;
            p = tree->parent[p];
            if (d > n) {
                free(depth);
                return -1;
            }
        }
        depth[i] = d;
        if (d > maxd)
            maxd = d;
    }
    int k = 0;
    for (int d = maxd; d >= 0; --d)
        for (int i = 0; i < n; ++i)
            if (depth[i] == d)
                order[k++] = i;
    free(depth);
    return 0;
}

/* core.cpp:134-145 / compute_node_probability (probability.cpp:175-234) with
 * internal vectors of B + 1 entries and the root's child factors taken over
 * rows s_min..s_max (root_size_range); the root vector has root_len entries
 * and entry i is the product of the children's row s_min + i (0 past
 * s_max, as the reference's zero-filled result buffer gives). */
static int orc_prune_rows(const orc_tree *tree, const int *counts, const double *mat_pool,
                          const int *node_mat, int N, const double *err, int n_dev,
                          int B, int s_min, int s_max, int root_len, double *out_root)
{
    int n = tree->n_nodes;
    int root = orc_find_root(tree);
    if (root < 0 || B + 1 > N || s_max + 1 > N || s_min < 0 || s_max < s_min || root_len < 1)
        return -1;
    int *order = (int *)malloc(sizeof(int) * n);
    if (orc_postorder(tree, order) != 0) {
        free(order);
        return -1;
    }
    int L = B + 1;
    int rows = s_max - s_min + 1;
    /* core.cpp:137-139: root vector has R entries, others B+1, all zero. */
    int stride = L > root_len ? L : root_len;
    if (rows > stride)
        stride = rows;
    double *probs = (double *)calloc((size_t)n * stride, sizeof(double));
    double *result = (double *)malloc(sizeof(double) * stride);
    int rc = 0;
    for (int oi = 0; oi < n && rc == 0; ++oi) {
        int u = order[oi];
        double *v = probs + (size_t)u * stride;
        if (tree->leaf_col[u] >= 0) {
            rc = orc_leaf_vector(counts[tree->leaf_col[u]], err, n_dev, B, v);
            continue;
        }
        int is_root = (u == root);
        int len = is_root ? root_len : L;
        for (int i = 0; i < len; ++i)
            v[i] = 1;
        for (int ch = 0; ch < n; ++ch) {
            if (tree->parent[ch] != u)
                continue;
            const double *M = mat_pool + (size_t)node_mat[ch] * N * N;
            const double *vc = probs + (size_t)ch * stride;
            for (int i = 0; i < stride; ++i)
                result[i] = 0;
            if (is_root)   /* probability.cpp:206-219: rows s_min..s_max, cols 0..B */
                orc_matvec(M, N, vc, s_min, s_max, 0, B, result);
            else           /* probability.cpp:220-233: rows 0..B, cols 0..B */
                orc_matvec(M, N, vc, 0, B, 0, B, result);
            for (int i = 0; i < len; ++i)
                v[i] *= result[i];
        }
    }
    if (rc == 0)
        memcpy(out_root, probs + (size_t)root * stride, sizeof(double) * root_len);
    free(result);
    free(probs);
    free(order);
    return rc;
}

int orc_prune(const orc_tree *tree, const int *counts, const double *mat_pool,
              const int *node_mat, int N, const double *err, int n_dev,
              int B, int R, double *out_root)
{
    /* root sizes 1..R (core.cpp:140-142) */
    if (R + 1 > N)
        return -1;
    return orc_prune_rows(tree, counts, mat_pool, node_mat, N, err, n_dev, B, 1, R, R, out_root);
}

int orc_family_probability(const orc_tree *tree, const int *counts, const double *mat_pool,
                           const int *node_mat, int N, int B, int max_size, int r, double *out)
{
    /* probability.cpp:462-466 and :468-474: m = min(B, max + max(50, max / 5)) */
    int grow = max_size / 5 > 50 ? max_size / 5 : 50;
    int m = B < max_size + grow ? B : max_size + grow;
    if (r < 1 || m < 0)
        return -1;
    double root[2];
    /* the root vector's first entry is row r; the others are products with
     * the zero tail of the result buffer (probability.cpp:210-217) */
    if (orc_prune_rows(tree, counts, mat_pool, node_mat, N, NULL, 0, m, r, r, 2, root) != 0)
        return -1;
    *out = root[0] < root[1] ? root[1] : root[0];   /* max_element, :474 */
    return 0;
}

/* gene_family_reconstructor.cpp:30-189 */
int orc_reconstruct(const orc_tree *tree, const int *counts, const double *mat_pool, const int *node_mat,
                    int N, int B, int R, const double *prior, int *states, double *node_L)
{
    int n = tree->n_nodes;
    int root = orc_find_root(tree);
    if (root < 0 || B < 0 || R < 0 || B + 1 > N)
        return -1;
    int *order = (int *)malloc(sizeof(int) * n);
    if (orc_postorder(tree, order) != 0) {
        free(order);
        return -1;
    }
    int L1 = B + 1;
    int Rr = (B < R ? B : R) + 1;   /* :149-151 */
    double *L = (double *)calloc((size_t)n * L1, sizeof(double));
    int *C = (int *)calloc((size_t)n * L1, sizeof(int));
    int root_c = 0, rc = 0;
    for (int oi = 0; oi < n && rc == 0; ++oi) {
        int u = order[oi];
        double *Lu = L + (size_t)u * L1;
        int *Cu = C + (size_t)u * L1;
        if (tree->leaf_col[u] >= 0) {
            /* :30-46: the leaf's count for every parent size i */
            int obs = counts[tree->leaf_col[u]];
            if (obs < 0 || obs >= N) {
                rc = -1;
                break;
            }
            const double *M = mat_pool + (size_t)node_mat[u] * N * N;
            for (int i = 0; i < L1; ++i) {
                Lu[i] = M[(size_t)i * N + obs];
                Cu[i] = obs;
            }
            continue;
        }
        if (u == root) {
            /* :48-74: every i scans the same j = 1..Rr-1; C[0] keeps the
             * last scan's first maximum */
            for (int i = 1; i < Rr; ++i) {
                double max_val = -1;
                for (int j = 1; j < Rr; ++j) {
                    double value = 1.0;
                    for (int ch = 0; ch < n; ++ch)
                        if (tree->parent[ch] == u)
                            value *= L[(size_t)ch * L1 + j];
                    double val = value * prior[j];
                    if (val > max_val) {
                        max_val = val;
                        root_c = j;
                    }
                }
                Lu[i] = max_val;
            }
            continue;
        }
        const double *M = mat_pool + (size_t)node_mat[u] * N * N;
        int zero = 1;   /* matrix::is_zero */
        for (size_t e = 0; e < (size_t)N * N && zero; ++e)
            zero = M[e] == 0.0;
        if (zero) {
            rc = -2;
            break;
        }
        /* :76-111: i is the parent size, j this node's */
        for (int i = 0; i < L1; ++i) {
            int max_j = 0;
            double max_val = -1;
            for (int j = 0; j < L1; ++j) {
                double value = 1.0;
                for (int ch = 0; ch < n; ++ch)
                    if (tree->parent[ch] == u)
                        value *= L[(size_t)ch * L1 + j];
                double val = value * M[(size_t)i * N + j];
                if (val > max_val) {
                    max_j = j;
                    max_val = val;
                }
            }
            Lu[i] = max_val;
            Cu[i] = max_j;
        }
    }
    if (rc == 0) {
        /* :170-189: the root's pick, then each interior child's C at its
         * parent's state (parents before children: reverse post-order) */
        states[root] = root_c;
        for (int oi = n - 1; oi >= 0; --oi) {
            int u = order[oi];
            if (u == root)
                continue;
            if (tree->leaf_col[u] >= 0)
                states[u] = counts[tree->leaf_col[u]];
            else
                states[u] = C[(size_t)u * L1 + states[tree->parent[u]]];
        }
        if (node_L)
            memcpy(node_L, L, sizeof(double) * (size_t)n * L1);
    }
    free(C);
    free(L);
    free(order);
    return rc;
}

/* probability.cpp:517-527 */
double orc_pvalue(double v, const double *conddist, int n)
{
    int idx = n - 1;
    /* std::upper_bound: the first element greater than v */
    int lo = 0, hi = n;
    while (lo < hi) {
        int mid = lo + (hi - lo) / 2;
        if (v < conddist[mid])
            hi = mid;
        else
            lo = mid + 1;
    }
    if (lo != n)
        idx = lo;
    return idx / (double)n;
}

/* probability.cpp:529-541 */
double orc_find_best_pvalue(const double *root_probs, int n_root, int max_size,
                            const double *conddist, int n_dist, int n_sim)
{
    if (n_root < 1)
        return 0;
    double *pv = (double *)calloc((size_t)n_root, sizeof(double));
    int check = (int)rint(max_size * 1.25);
    for (int j = 0; j < check && j < n_root && j < n_dist; ++j)
        pv[j] = orc_pvalue(root_probs[j], conddist + (size_t)j * n_sim, n_sim);
    double best = pv[0];
    for (int j = 1; j < n_root; ++j)   /* std::max_element: the first largest */
        if (best < pv[j])
            best = pv[j];
    free(pv);
    return best;
}

/* ------------------------------------------------------------------------ */
/* Root reductions                                                          */
/* ------------------------------------------------------------------------ */

/* base_model.cpp:80-90 for one family */
static double orc_base_family(int R, const double *root, const double *prior)
{
    double best = 0;
    for (int j = 0; j < R; ++j) {
        double full = log(root[j]) + log(prior[j]);
        if (j == 0 || best < full)
            best = full;
    }
    return best;
}

double orc_base_reduce(int F, int R, const double *roots, const double *prior,
                       double *out_lnl)
{
    double sum = 0.0;
    for (int f = 0; f < F; ++f) {
        out_lnl[f] = orc_base_family(R, roots + (size_t)f * R, prior);
        sum += out_lnl[f];
    }
    return -sum;
}

/* gamma_core.cpp:143-165 (prune's reduction) + 97-109 (posterior) + 190-208
 * for one family.  Returns 0 on success, 1 on a zero-sum category. */
static int orc_gamma_family(int K, int R, const double *roots, const double *prior,
                            const double *cat_probs, double *cat, double *fam,
                            double *post, double *lnl)
{
    for (int k = 0; k < K; ++k) {
        const double *L = roots + (size_t)k * R;
        double s = 0.0;
        for (int j = 0; j < R; ++j)
            s += L[j];
        if (s == 0.0)
            return 1;
        double best = 0;
        for (int j = 0; j < R; ++j) {
            double full = L[j] * prior[j];
            if (j == 0 || best < full)
                best = full;
        }
        cat[k] = best * cat_probs[k];
    }
    double family = 0.0;
    for (int k = 0; k < K; ++k)
        family += cat[k];
    double denom = 0.0;
    for (int k = 0; k < K; ++k)
        denom += cat[k] * cat_probs[k];
    for (int k = 0; k < K; ++k)
        post[k] = (cat[k] * cat_probs[k]) / denom;
    *fam = family;
    *lnl = log(family);
    return 0;
}

double orc_gamma_reduce(int F, int K, int R, const double *roots,
                        const double *prior, const double *cat_probs,
                        double *out_cat, double *out_fam, double *out_post,
                        double *out_lnl, int *n_failed)
{
    int failed = 0;
    double sum = 0.0;
    for (int f = 0; f < F; ++f) {
        int bad = orc_gamma_family(K, R, roots + (size_t)f * K * R, prior, cat_probs,
                                   out_cat + (size_t)f * K, out_fam + f,
                                   out_post + (size_t)f * K, out_lnl + f);
        if (bad) {
            ++failed;
            out_lnl[f] = NAN;
        } else {
            sum += out_lnl[f];
        }
    }
    if (n_failed)
        *n_failed = failed;
    if (failed)
        return -log(0.0);
    return -sum;
}

/* ------------------------------------------------------------------------ */
/* Whole evaluation (what one infer_family_likelihoods call does)           */
/* ------------------------------------------------------------------------ */

typedef struct {
    double lambda_q, t_q;
} orc_key;

static int orc_key_find(const orc_key *keys, int n, double lq, double tq)
{
    for (int i = 0; i < n; ++i)
        if (keys[i].lambda_q == lq && keys[i].t_q == tq)
            return i;
    return -1;
}

double orc_evaluate(const orc_tree *tree, const double *node_lambda,
                    const double *node_t, int F, int n_leaves, const int *counts,
                    int B, int R, const double *prior, const double *err,
                    int n_dev, int n_cat, const double *multipliers,
                    const double *cat_probs, double *out_lnl, double *out_cat,
                    double *out_fam, double *out_post, int n_threads)
{
    if (!orc_ready)
        orc_init();
    int n = tree->n_nodes;
    int N = (R > B ? R : B) + 1;   /* base_model.cpp:65, gamma_core.cpp:185 */
    int K = n_cat > 0 ? n_cat : 1;
    int root = orc_find_root(tree);
    if (root < 0)
        return NAN;
#ifdef _OPENMP
    if (n_threads > 0)
        omp_set_num_threads(n_threads);
#else
    (void)n_threads;
#endif

    /* Key set as precalculate_matrices sees it: every lambda value (times
     * each multiplier) crossed with every distinct positive branch length
     * (lambda.cpp:69-83, clade.cpp:236-245, gamma_core.cpp:111-121). */
    double *lams = (double *)malloc(sizeof(double) * n);
    double *ts = (double *)malloc(sizeof(double) * n);
    int nl = 0, nt = 0;
    for (int i = 0; i < n; ++i) {
        if (i == root)
            continue;
        int seen = 0;
        for (int j = 0; j < nl; ++j)
            if (lams[j] == node_lambda[i])
                seen = 1;
        if (!seen)
            lams[nl++] = node_lambda[i];
        seen = 0;
        for (int j = 0; j < nt; ++j)
            if (ts[j] == node_t[i])
                seen = 1;
        if (!seen && node_t[i] > 0.0)
            ts[nt++] = node_t[i];
    }
    orc_key *keys = (orc_key *)malloc(sizeof(orc_key) * (size_t)nl * nt * K + 1);
    int nk = 0;
    for (int k = 0; k < K; ++k)
        for (int a = 0; a < nl; ++a)
            for (int b = 0; b < nt; ++b) {
                double lam = n_cat > 0 ? lams[a] * multipliers[k] : lams[a];
                double lq = orc_key_lambda(lam), tq = orc_key_branch_length(ts[b]);
                if (orc_key_find(keys, nk, lq, tq) < 0) {
                    keys[nk].lambda_q = lq;
                    keys[nk].t_q = tq;
                    ++nk;
                }
            }
    double *pool = (double *)malloc(sizeof(double) * (size_t)nk * N * N);
    /* matrix_cache.cpp:136: OpenMP over (key, row). */
    long total_rows = (long)nk * N;
#pragma omp parallel for schedule(static)
    for (long kr = 0; kr < total_rows; ++kr) {
        int key = (int)(kr / N), s = (int)(kr % N);
        orc_build_rows(keys[key].lambda_q, keys[key].t_q, N, s, s + 1,
                       pool + (size_t)key * N * N);
    }
    int *node_mat = (int *)malloc(sizeof(int) * (size_t)n * K);
    for (int k = 0; k < K; ++k)
        for (int i = 0; i < n; ++i) {
            if (i == root) {
                node_mat[k * n + i] = 0;
                continue;
            }
            double lam = n_cat > 0 ? node_lambda[i] * multipliers[k] : node_lambda[i];
            node_mat[k * n + i] = orc_key_find(keys, nk, orc_key_lambda(lam),
                                               orc_key_branch_length(node_t[i]));
        }

    double *roots = (double *)malloc(sizeof(double) * (size_t)F * K * R);
    int bad = 0;
#pragma omp parallel for schedule(static) reduction(+ : bad)
    for (int f = 0; f < F; ++f)
        for (int k = 0; k < K; ++k)
            if (orc_prune(tree, counts + (size_t)f * n_leaves, pool, node_mat + k * n, N,
                          err, n_dev, B, R, roots + ((size_t)f * K + k) * R) != 0)
                bad++;

    double total;
    if (bad) {
        total = NAN;
    } else if (n_cat == 0) {
        total = orc_base_reduce(F, R, roots, prior, out_lnl);
        if (out_fam)
            for (int f = 0; f < F; ++f)
                out_fam[f] = out_lnl[f];
    } else {
        int failed = 0;
        total = orc_gamma_reduce(F, K, R, roots, prior, cat_probs, out_cat, out_fam,
                                 out_post, out_lnl, &failed);
    }
    free(roots);
    free(node_mat);
    free(pool);
    free(keys);
    free(ts);
    free(lams);
    return total;
}

/* ------------------------------------------------------------------------ */
/* Discrete gamma (PAML; Yang 1994) -- published algorithms AS91 (Best &    */
/* Roberts 1975), AS32 (Bhattacharjee 1970) and AS70 (Odeh & Evans 1974).   */
/* The iteration steps and tolerances are those of the published routines,  */
/* which the reference uses (gamma.cpp:66-215); the multipliers depend on   */
/* them because the iterations stop at a tolerance of 0.5e-6.               */
/* ------------------------------------------------------------------------ */

/* AS70: upper-tail normal quantile by a rational approximation in
 * y = sqrt(-2 log p). */
static double orc_normal_quantile(double prob)
{
    static const double num[5] = {-.322232431088, -1, -.342242088547, -.0204231210245,
                                  -.453642210148e-4};
    static const double den[5] = {.0993484626060, .588581570495, .531103462366,
                                  .103537752850, .0038560700634};
    double tail = prob < 0.5 ? prob : 1 - prob;
    if (tail < 1e-20)
        return -9999;
    double y = sqrt(log(1 / (tail * tail)));
    double pn = num[4], pd = den[4];
    for (int i = 3; i >= 0; --i) {
        pn = pn * y + num[i];
        pd = pd * y + den[i];
    }
    double z = y + pn / pd;
    return prob < 0.5 ? -z : z;
}

/* AS32: regularized lower incomplete gamma ratio I(x, a), series for
 * x <= 1 or x < a, otherwise Legendre's continued fraction. */
static double orc_incomplete_gamma(double x, double a, double ln_gamma_a)
{
    const double tol = 1e-8, big = 1e30;
    if (x == 0)
        return 0;
    if (x < 0 || a <= 0)
        return -1;
    double lead = exp(a * log(x) - x - ln_gamma_a);
    if (!(x > 1 && x >= a)) {
        double sum = 1, term = 1, d = a;
        do {
            d++;
            term *= x / d;
            sum += term;
        } while (term > tol);
        return sum * (lead / a);
    }
    double u = 1 - a, w = u + x + 1, k = 0;
    double q[6] = {1, x, x + 1, x * w, 0, 0};
    double frac = q[2] / q[3];
    for (;;) {
        u++;
        w += 2;
        k++;
        double uk = u * k;
        q[4] = w * q[2] - uk * q[0];
        q[5] = w * q[3] - uk * q[1];
        if (q[5] != 0) {
            double next = q[4] / q[5];
            double gap = fabs(frac - next);
            if (gap <= tol && gap <= tol * next)
                return 1 - lead * frac;
            frac = next;
        }
        for (int i = 0; i < 4; ++i)
            q[i] = q[i + 2];
        if (!(fabs(q[4]) < big))
            for (int i = 0; i < 4; ++i)
                q[i] /= big;
    }
}

/* AS91: chi-square quantile with v degrees of freedom. */
static double orc_chi2_quantile(double p, double v)
{
    const double tol = .5e-6, ln2 = .6931471805;
    if (p < .000002 || p > .999998 || v <= 0)
        return -1;
    double g = lgamma(v / 2), half_v = v / 2, c = half_v - 1, ch;
    if (v < -1.24 * log(p)) {
        /* small-v start */
        ch = pow((p * half_v * exp(g + half_v * ln2)), 1 / half_v);
        if (ch - tol < 0)
            return ch;
    } else if (v <= .32) {
        /* v <= 0.32: Newton iterations on a rational start */
        double la = log(1 - p), prev;
        ch = 0.4;
        do {
            prev = ch;
            double p1 = 1 + ch * (4.67 + ch);
            double p2 = ch * (6.73 + ch * (6.66 + ch));
            double t = -0.5 + (4.67 + 2 * ch) / p1 - (6.73 + ch * (13.32 + 3 * ch)) / p2;
            ch -= (1 - exp(la + g + .5 * ch + c * ln2) * p2 / p1) / t;
        } while (!(fabs(prev / ch - 1) - .01 <= 0));
    } else {
        /* Wilson-Hilferty start */
        double x = orc_normal_quantile(p);
        double p1 = 0.222222 / v;
        ch = v * pow((x * sqrt(p1) + 1 - p1), 3.0);
        if (ch > 2.2 * v + 6)
            ch = -2 * (log(1 - p) - c * log(.5 * ch) + g);
    }
    /* seventh-order Taylor refinement until the relative step is below tol */
    double prev;
    do {
        prev = ch;
        double p1 = .5 * ch;
        double t = orc_incomplete_gamma(p1, half_v, g);
        if (t < 0)
            return -1;
        double p2 = p - t;
        t = p2 * exp(half_v * ln2 + g + p1 - c * log(ch));
        double b = t / ch;
        double a = 0.5 * t - b * c;
        double s1 = (210 + a * (140 + a * (105 + a * (84 + a * (70 + 60 * a))))) / 420;
        double s2 = (420 + a * (735 + a * (966 + a * (1141 + 1278 * a)))) / 2520;
        double s3 = (210 + a * (462 + a * (707 + 932 * a))) / 2520;
        double s4 = (252 + a * (672 + 1182 * a) + c * (294 + a * (889 + 1740 * a))) / 5040;
        double s5 = (84 + 264 * a + c * (175 + 606 * a)) / 2520;
        double s6 = (120 + c * (346 + 127 * c)) / 5040;
        ch += t * (1 + 0.5 * t * s1 - b * c * (s1 - b * (s2 - b * (s3 - b * (s4 - b * (s5 - b * s6))))));
    } while (fabs(prev / ch - 1) > tol);
    return ch;
}

/* gamma.cpp:15-52 (median branch) + gamma.cpp:225-241 */
void orc_get_gamma(int K, double alpha, double *freq, double *rate)
{
    double beta = alpha;
    double gap05 = 1.0 / (2.0 * K), factor = alpha / beta * K, t = 0;
    for (int i = 0; i < K; i++)
        rate[i] = orc_chi2_quantile((i * 2.0 + 1) * gap05, 2.0 * (alpha)) / (2.0 * (beta));
    for (int i = 0; i < K; i++)
        t += rate[i];
    for (int i = 0; i < K; i++)
        rate[i] *= factor / t;
    for (int i = 0; i < K; i++)
        freq[i] = 1.0 / K;
}

/* ------------------------------------------------------------------------ */
/* Root priors (root_equilibrium_distribution.cpp), float valued (.h:40)     */
/* ------------------------------------------------------------------------ */

/* poisson.cpp:21-24 */
static double orc_poisspdf(int x, double lambda)
{
    return exp(x * log(lambda) - lgamma(x + 1) - lambda);
}

/* .cpp:30-35 + build_percentages .cpp:70-79: uniform over 0..max_size-1 */
void orc_prior_uniform(int max_size, int R, double *prior)
{
    float pct = (float)1 / (float)max_size;
    for (int j = 0; j < R; ++j)
        prior[j] = j < max_size ? (double)pct : 0.0;
}

/* .cpp:56-68: shifted Poisson, truncated once ceil(pct*n) copies fill n,
 * plus 5 more percentages. */
void orc_prior_poisson(double poisson_lambda, int num_values, int R, double *prior)
{
    int cap = num_values + 64;
    double *pct = (double *)malloc(sizeof(double) * (size_t)(cap + 8));
    size_t filled = 0;
    int n = 0;
    for (int i = 0; filled < (size_t)num_values; ++i) {
        double p = orc_poisspdf(i, poisson_lambda);
        for (size_t j = 0; j < p * num_values; ++j)
            filled++;
        if (n >= cap) {
            cap *= 2;
            pct = (double *)realloc(pct, sizeof(double) * (size_t)(cap + 8));
        }
        pct[n++] = p;
    }
    for (int i = 0; i < 5; ++i) {
        pct[n] = orc_poisspdf(n, poisson_lambda);
        ++n;
    }
    for (int j = 0; j < R; ++j)
        prior[j] = j < n ? (double)(float)pct[j] : 0.0;
    free(pct);
}

/* .cpp:15-28 + build_percentages */
void orc_prior_map(int n, const int *sizes, const int *counts, int R, double *prior)
{
    long total = 0;
    int maxs = 0;
    for (int i = 0; i < n; ++i) {
        total += counts[i];
        if (counts[i] > 0 && sizes[i] > maxs)
            maxs = sizes[i];
    }
    for (int j = 0; j < R; ++j) {
        long c = 0;
        for (int i = 0; i < n; ++i)
            if (sizes[i] == j)
                c += counts[i];
        float pct = (float)c / (float)total;
        prior[j] = j <= maxs ? (double)pct : 0.0;
    }
}
