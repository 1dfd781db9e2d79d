// ref_driver.cpp -- C entry points over the compiled CAFE5 reference sources.
//
// TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together with
// the reference's own hot-path sources (read in place from the reference tree,
// never copied) into oracle/_ref/libcaferef.so.  Tests use it to pin the oracle
// (cafe_oracle.c) to the reference at ~1 ulp; nothing in the product loads it.
//
// Every function here only calls reference code: matrix_cache, inference_prune,
// base_model / gamma_model::infer_family_likelihoods, get_gamma, the
// root_equilibrium_distribution constructors, the p-value functions
// (create_family, compute_family_probabilities, pvalue, find_best_pvalue) and
// Pupko's reconstruction (initialize_at_node, reconstruct_gene_family).
#include <algorithm>
#include <chrono>
#include <fstream>
#include <cmath>
#include <cstring>
#include <random>
#include <set>
#include <sstream>
#include <string>
#include <vector>

#include "easylogging++.h"
INITIALIZE_EASYLOGGINGPP

#include "base_model.h"
#include "clade.h"
#include "core.h"
#include "error_model.h"
#include "gamma.h"
#include "gamma_core.h"
#include "gene_family.h"
#include "gene_family_reconstructor.h"
#include "io.h"
#include "lambda.h"
#include "matrix_cache.h"
#include "probability.h"
#include "root_equilibrium_distribution.h"

// The reference's executables define this engine (cafexp.cpp / test.cpp).
std::mt19937 randomizer_engine(10);
void init_lgamma_cache();
double chooseln(double n, double r);

#define REF_API extern "C" __attribute__((visibility("default")))

namespace {

// Access to the protected per-family records of the reference models.
struct base_probe : public base_model {
    using base_model::base_model;
    const std::vector<family_info_stash>& stash() const { return results; }
};
struct gamma_probe : public gamma_model {
    using gamma_model::gamma_model;
    const std::vector<family_info_stash>& stash() const { return results; }
};

std::vector<std::string> split_names(const char* names)
{
    std::vector<std::string> out;
    std::istringstream ist(names);
    std::string s;
    while (ist >> s) out.push_back(s);
    return out;
}

std::vector<gene_family> make_families(int F, const char* leaf_names, const int* counts)
{
    auto names = split_names(leaf_names);
    std::vector<gene_family> fams(F);
    for (int f = 0; f < F; ++f) {
        fams[f].set_id("F" + std::to_string(f));
        for (size_t i = 0; i < names.size(); ++i)
            fams[f].set_species_size(names[i], counts[(size_t)f * names.size() + i]);
    }
    return fams;
}

// error model rows 0..n_rows-1 with deviations -h..h (n_dev = 2h+1)
error_model* make_error_model(int n_dev, int n_rows, const double* rows, int max_size)
{
    if (!rows) return nullptr;
    auto* em = new error_model();
    std::vector<std::string> devs;
    for (int i = 0; i < n_dev; ++i) devs.push_back(std::to_string(i - (n_dev - 1) / 2));
    em->set_deviations(devs);
    em->set_max_family_size(max_size);
    for (int r = 0; r < n_rows; ++r)
        em->set_probabilities(r, std::vector<double>(rows + (size_t)r * n_dev, rows + (size_t)(r + 1) * n_dev));
    return em;
}

lambda* make_lambda(int n_lambdas, const double* lambdas, const char* lambda_tree)
{
    if (lambda_tree && lambda_tree[0]) {
        clade* lt = parse_newick(lambda_tree, true);
        auto m = lt->get_lambda_index_map();
        delete lt;
        return new multiple_lambda(m, std::vector<double>(lambdas, lambdas + n_lambdas));
    }
    return new single_lambda(lambdas[0]);
}

}  // namespace

REF_API void ref_init()
{
    init_lgamma_cache();
    // as the reference's test main does: no log output (test.cpp:3149-3168)
    el::Configurations conf;
    conf.setToDefault();
    conf.setGlobally(el::ConfigurationType::Enabled, "false");
    conf.setGlobally(el::ConfigurationType::ToFile, "false");
    el::Loggers::setDefaultConfigurations(conf, true);
}

REF_API double ref_chooseln(double n, double r) { return chooseln(n, r); }

REF_API double ref_birthdeath(int s, int c, double log_alpha, double coeff)
{
    return birthdeath_rate_with_log_alpha(s, c, log_alpha, coeff);
}

REF_API double ref_transition(double lambda, double t, int s, int c)
{
    return the_probability_of_going_from_parent_fam_size_to_c(lambda, t, s, c);
}

// matrix_cache(N).precalculate_matrices({lambda}, {t}); get_matrix(t, lambda)
REF_API int ref_build_matrix(double lambda, double t, int N, double* out)
{
    try {
        matrix_cache cache(N);
        cache.precalculate_matrices({lambda}, std::set<double>{t});
        const matrix* m = cache.get_matrix(t, lambda);
        for (int s = 0; s < N; ++s)
            for (int c = 0; c < N; ++c) out[(size_t)s * N + c] = m->get(s, c);
        return 0;
    } catch (...) {
        return -1;
    }
}

// inference_prune (core.cpp:134) on one family; out_root has R entries.
REF_API int ref_prune(const char* newick, const char* leaf_names, const int* counts,
                      double lambda, double multiplier, int B, int R,
                      int n_dev, int n_err_rows, const double* err_rows, double* out_root)
{
    try {
        clade* tree = parse_newick(newick);
        auto fams = make_families(1, leaf_names, counts);
        error_model* em = make_error_model(n_dev, n_err_rows, err_rows, B);
        single_lambda lam(lambda);
        matrix_cache cache(std::max(R, B) + 1);
        cache.precalculate_matrices({lambda * multiplier}, tree->get_branch_lengths());
        auto v = inference_prune(fams[0], cache, &lam, em, tree, multiplier, R, B);
        std::copy(v.begin(), v.end(), out_root);
        delete em;
        delete tree;
        return 0;
    } catch (...) {
        return -1;
    }
}

// root_equilibrium_distribution variants: kind 0 uniform(max_size),
// 1 Poisson(poisson_lambda, max_size), 2 map(sizes, counts)
static root_equilibrium_distribution* make_prior(int kind, int max_size, double poisson_lambda,
                                                 int n_map, const int* map_sizes, const int* map_counts)
{
    if (kind == 1) return new root_equilibrium_distribution(poisson_lambda, (size_t)max_size);
    if (kind == 2) {
        std::map<int, int> m;
        for (int i = 0; i < n_map; ++i) m[map_sizes[i]] = map_counts[i];
        return new root_equilibrium_distribution(m);
    }
    return new root_equilibrium_distribution((size_t)max_size);
}

REF_API void ref_prior(int kind, int max_size, double poisson_lambda, int n_map,
                       const int* map_sizes, const int* map_counts, int R, double* out)
{
    auto* p = make_prior(kind, max_size, poisson_lambda, n_map, map_sizes, map_counts);
    for (int j = 0; j < R; ++j) out[j] = p->compute(j);
    delete p;
}

// base_model::infer_family_likelihoods; returns -lnL (or +inf), out_lnl[F]
REF_API double ref_base_model(const char* newick, const char* leaf_names, int F, const int* counts,
                              int n_lambdas, const double* lambdas, const char* lambda_tree,
                              int B, int R, int prior_kind, int prior_max, double poisson_lambda,
                              int n_dev, int n_err_rows, const double* err_rows, double* out_lnl)
{
    try {
        clade* tree = parse_newick(newick);
        auto fams = make_families(F, leaf_names, counts);
        error_model* em = make_error_model(n_dev, n_err_rows, err_rows, B);
        lambda* lam = make_lambda(n_lambdas, lambdas, lambda_tree);
        auto* prior = make_prior(prior_kind, prior_max, poisson_lambda, 0, nullptr, nullptr);
        base_probe model(lam, tree, &fams, B, R, em);
        double v = model.infer_family_likelihoods(*prior, lam);
        const auto& st = model.stash();
        for (int f = 0; f < F && (size_t)f < st.size(); ++f) out_lnl[f] = st[f].posterior_probability;
        delete prior;
        delete lam;
        delete em;
        delete tree;
        return v;
    } catch (...) {
        return std::nan("");
    }
}

// gamma_model with explicit categories (gamma_core.h constructor 2);
// out arrays F*K (cat, post) and F (fam); returns -lnL or +inf
REF_API double ref_gamma_model(const char* newick, const char* leaf_names, int F, const int* counts,
                               int n_lambdas, const double* lambdas, const char* lambda_tree,
                               int B, int R, int K, const double* cat_probs, const double* multipliers,
                               int prior_kind, int prior_max, double poisson_lambda,
                               int n_dev, int n_err_rows, const double* err_rows,
                               double* out_cat, double* out_fam, double* out_post)
{
    try {
        clade* tree = parse_newick(newick);
        auto fams = make_families(F, leaf_names, counts);
        error_model* em = make_error_model(n_dev, n_err_rows, err_rows, B);
        lambda* lam = make_lambda(n_lambdas, lambdas, lambda_tree);
        auto* prior = make_prior(prior_kind, prior_max, poisson_lambda, 0, nullptr, nullptr);
        gamma_probe model(lam, tree, &fams, B, R,
                          std::vector<double>(cat_probs, cat_probs + K),
                          std::vector<double>(multipliers, multipliers + K), em);
        double v = model.infer_family_likelihoods(*prior, lam);
        const auto& st = model.stash();
        for (size_t i = 0; i < st.size() && i < (size_t)F * K; ++i) {
            out_cat[i] = st[i].category_likelihood;
            out_post[i] = st[i].posterior_probability;
            if (i % K == 0) out_fam[i / K] = st[i].family_likelihood;
        }
        delete prior;
        delete lam;
        delete em;
        delete tree;
        return v;
    } catch (...) {
        return std::nan("");
    }
}

REF_API void ref_get_gamma(int K, double alpha, double* freq, double* rate)
{
    std::vector<double> f(K), r(K);
    get_gamma(f, r, alpha);
    std::copy(f.begin(), f.end(), freq);
    std::copy(r.begin(), r.end(), rate);
}

// --- timing sessions: the base model is constructed once (its O(F^2)
// build_reference_list runs in the constructor, core.cpp:69-70) and then
// evaluated repeatedly, as the optimizer does.
namespace {
struct base_session {
    clade* tree;
    std::vector<gene_family> fams;
    single_lambda lam{0.0};
    root_equilibrium_distribution* prior;
    base_probe* model;
};
}  // namespace

REF_API void* ref_base_session_create(const char* newick, const char* leaf_names, int F, const int* counts,
                                      int B, int R, int prior_kind, int prior_max, double poisson_lambda)
{
    try {
        auto* s = new base_session();
        s->tree = parse_newick(newick);
        s->fams = make_families(F, leaf_names, counts);
        s->prior = make_prior(prior_kind, prior_max, poisson_lambda, 0, nullptr, nullptr);
        s->model = new base_probe(&s->lam, s->tree, &s->fams, B, R, nullptr);
        return s;
    } catch (...) {
        return nullptr;
    }
}

REF_API double ref_base_session_eval(void* h, double lambda, double* out_lnl)
{
    auto* s = (base_session*)h;
    s->lam.update(&lambda);
    double v = s->model->infer_family_likelihoods(*s->prior, &s->lam);
    if (out_lnl) {
        const auto& st = s->model->stash();
        for (size_t f = 0; f < st.size(); ++f) out_lnl[f] = st[f].posterior_probability;
    }
    return v;
}

REF_API void ref_base_session_free(void* h)
{
    auto* s = (base_session*)h;
    if (!s) return;
    delete s->model;
    delete s->prior;
    delete s->tree;
    delete s;
}

#ifdef _OPENMP
#include <omp.h>
#endif
REF_API int ref_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
    return omp_get_max_threads();
#else
    (void)n;
    return 1;
#endif
}

// compute_family_probabilities (probability.cpp:448-477) for F families whose
// clademaps hold counts[f] at the leaves (leaf_names order) and inner[f] at
// every interior node; out[F].
REF_API int ref_family_probabilities(const char* newick, const char* leaf_names, int F, const int* counts,
                                     const int* inner, double lambda, int B, int R, int root_size, double* out)
{
    try {
        clade* tree = parse_newick(newick);
        auto names = split_names(leaf_names);
        single_lambda lam(lambda);
        matrix_cache cache(std::max(R, B) + 1);
        cache.precalculate_matrices({lambda}, tree->get_branch_lengths());
        pvalue_parameters p = {tree, &lam, B, R, cache};
        std::vector<clademap<int>> sizes(F);
        for (int f = 0; f < F; ++f)
            tree->apply_prefix_order([&](const clade* c) {
                if (c->is_leaf()) {
                    const auto it = std::find(names.begin(), names.end(), c->get_taxon_name());
                    sizes[f][c] = counts[(size_t)f * names.size() + (it - names.begin())];
                } else {
                    sizes[f][c] = inner[f];
                }
            });
        auto v = compute_family_probabilities(p, sizes, root_size);
        std::copy(v.begin(), v.end(), out);
        delete tree;
        return 0;
    } catch (...) {
        return -1;
    }
}

// A pruning session for timing the reference's two per-evaluation costs
// apart: the matrix_cache of one lambda x multiplier over the tree's branch
// lengths is built (precalculate_matrices, timed), then each run prunes every
// family with inference_prune in an OpenMP loop over families (the loop of
// base_model / gamma_model).  Outputs: build seconds; run returns seconds.
struct prune_session {
    clade* tree = nullptr;
    std::vector<gene_family> fams;
    single_lambda* lam = nullptr;
    matrix_cache* cache = nullptr;
    double multiplier = 1.0;
    int B = 0, R = 0;
};

REF_API void* ref_prune_session_create(const char* newick, const char* leaf_names, int F, const int* counts,
                                       double lambda, double multiplier, int B, int R, double* out_build_s)
{
    try {
        auto* s = new prune_session();
        s->tree = parse_newick(newick);
        s->fams = make_families(F, leaf_names, counts);
        s->lam = new single_lambda(lambda);
        s->cache = new matrix_cache(std::max(R, B) + 1);
        s->multiplier = multiplier;
        s->B = B;
        s->R = R;
        const auto t0 = std::chrono::steady_clock::now();
        s->cache->precalculate_matrices({lambda * multiplier}, s->tree->get_branch_lengths());
        const auto t1 = std::chrono::steady_clock::now();
        if (out_build_s) *out_build_s = std::chrono::duration<double>(t1 - t0).count();
        return s;
    } catch (...) {
        return nullptr;
    }
}

REF_API double ref_prune_session_run(void* h)
{
    auto* s = (prune_session*)h;
    const int F = (int)s->fams.size();
    std::vector<double> first(F, 0.0);
    const auto t0 = std::chrono::steady_clock::now();
#pragma omp parallel for schedule(dynamic)
    for (int f = 0; f < F; ++f) {
        auto v = inference_prune(s->fams[f], *s->cache, s->lam, nullptr, s->tree, s->multiplier, s->R, s->B);
        first[f] = v.empty() ? 0.0 : v[0];
    }
    const auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

REF_API void ref_prune_session_free(void* h)
{
    auto* s = (prune_session*)h;
    if (!s) return;
    delete s->cache;
    delete s->lam;
    delete s->tree;
    delete s;
}

REF_API double ref_pvalue(double v, const double* conddist, int n)
{
    return pvalue(v, std::vector<double>(conddist, conddist + n));
}

// find_best_pvalue (probability.cpp:529-541) for a family whose largest count
// is max_size; conddist holds n_dist sorted rows of n_sim values.  The caller
// keeps rint(max_size * 1.25) within both n_root and n_dist.
REF_API double ref_find_best_pvalue(int max_size, const double* root_probs, int n_root, const double* conddist,
                                    int n_dist, int n_sim)
{
    gene_family fam;
    fam.set_species_size("A", max_size);
    std::vector<std::vector<double>> cd(n_dist);
    for (int j = 0; j < n_dist; ++j) cd[j].assign(conddist + (size_t)j * n_sim, conddist + (size_t)(j + 1) * n_sim);
    return find_best_pvalue(fam, std::vector<double>(root_probs, root_probs + n_root), cd);
}

// create_family (probability.cpp:355-375) after randomizer_engine.seed(seed):
// out[i] = the size drawn for leaf i (leaf_names order), n_draws families in a row.
REF_API int ref_create_families(const char* newick, const char* leaf_names, double lambda, int B, int R,
                                int N, unsigned seed, int root_size, int n_draws, int* out)
{
    try {
        clade* tree = parse_newick(newick);
        auto names = split_names(leaf_names);
        single_lambda lam(lambda);
        matrix_cache cache(N);
        cache.precalculate_matrices({lambda}, tree->get_branch_lengths());
        pvalue_parameters p = {tree, &lam, B, R, cache};
        randomizer_engine.seed(seed);
        for (int d = 0; d < n_draws; ++d) {
            auto sizes = create_family(p, root_size);
            for (size_t i = 0; i < names.size(); ++i)
                out[(size_t)d * names.size() + i] = sizes.at(tree->find_descendant(names[i]));
        }
        delete tree;
        return 0;
    } catch (...) {
        return -1;
    }
}

// pupko_reconstructor::reconstruct_gene_family (gene_family_reconstructor.cpp:
// 30-189) of one family with the prior of ref_prior's kinds; out[i] is the
// state of the node named node_names[i] (a leaf's own count).
REF_API int ref_reconstruct(const char* newick, const char* leaf_names, const char* node_names, const int* counts,
                            double lambda, int B, int R, int prior_kind, int prior_max, double poisson_lambda,
                            int n_map, const int* map_sizes, const int* map_counts, int* out)
{
    try {
        clade* tree = parse_newick(newick);
        auto fams = make_families(1, leaf_names, counts);
        single_lambda lam(lambda);
        matrix_cache cache(std::max(R, B) + 1);
        cache.precalculate_matrices({lambda}, tree->get_branch_lengths());
        auto* prior = make_prior(prior_kind, prior_max, poisson_lambda, n_map, map_sizes, map_counts);
        clademap<int> result;
        clademap<std::vector<int>> Cs;
        clademap<std::vector<double>> Ls;
        for (auto it = tree->reverse_level_begin(); it != tree->reverse_level_end(); ++it)
            pupko_reconstructor::initialize_at_node(*it, Cs, Ls, B, R);
        pupko_reconstructor::reconstruct_gene_family(&lam, tree, &fams[0], &cache, prior, result, Cs, Ls);
        auto names = split_names(node_names);
        for (size_t i = 0; i < names.size(); ++i) {
            const clade* c = tree->find_descendant(names[i]);
            out[i] = c->is_leaf() ? fams[0].get_species_size(names[i]) : result[c];
        }
        delete prior;
        delete tree;
        return 0;
    } catch (...) {
        return -1;
    }
}

// Pupko reconstruction of F families at one lambda (uniform prior over
// 1..R) with one matrix cache, families over OpenMP threads as the reference's
// base_model::reconstruct_ancestral_states does (base_model.cpp:145-160): the
// CPU side of the reconstruction timing.  out[f * n_nodes + i] as in
// ref_reconstruct.
REF_API int ref_reconstruct_many(const char* newick, const char* leaf_names, const char* node_names, int F,
                                 const int* counts, double lambda, int B, int R, int* out)
{
    try {
        clade* tree = parse_newick(newick);
        auto fams = make_families(F, leaf_names, counts);
        single_lambda lam(lambda);
        matrix_cache cache(std::max(R, B) + 1);
        cache.precalculate_matrices({lambda}, tree->get_branch_lengths());
        root_equilibrium_distribution prior((size_t)R);
        auto names = split_names(node_names);
        std::vector<const clade*> nodes;
        for (auto& n : names) nodes.push_back(tree->find_descendant(n));
        int bad = 0;
#pragma omp parallel for reduction(+ : bad)
        for (int f = 0; f < F; ++f) {
            try {
                clademap<int> result;
                clademap<std::vector<int>> Cs;
                clademap<std::vector<double>> Ls;
                for (auto it = tree->reverse_level_begin(); it != tree->reverse_level_end(); ++it)
                    pupko_reconstructor::initialize_at_node(*it, Cs, Ls, B, R);
                pupko_reconstructor::reconstruct_gene_family(&lam, tree, &fams[f], &cache, &prior, result, Cs, Ls);
                for (size_t i = 0; i < nodes.size(); ++i)
                    out[(size_t)f * nodes.size() + i] =
                        nodes[i]->is_leaf() ? fams[f].get_species_size(names[i]) : result[nodes[i]];
            } catch (...) {
                ++bad;
            }
        }
        delete tree;
        return bad ? -1 : 0;
    } catch (...) {
        return -1;
    }
}

// The files as the reference reads them (read_gene_families, io.cpp:134-217;
// read_error_model_file, :228-275) evaluated by its base model at one lambda
// with a uniform prior over 1..R: -lnL, out_lnl[F] in file order (F <= cap).
REF_API double ref_example(const char* newick, const char* families_path, const char* error_model_path,
                           double lambda, int B, int R, int cap, int* n_families, double* out_lnl)
{
    try {
        clade* tree = parse_newick(newick);
        std::ifstream ff(families_path);
        std::vector<gene_family> fams;
        read_gene_families(ff, tree, fams);
        error_model* em = nullptr;
        if (error_model_path && error_model_path[0]) {
            std::ifstream ef(error_model_path);
            em = new error_model();
            read_error_model_file(ef, em);
        }
        single_lambda lam(lambda);
        root_equilibrium_distribution prior((size_t)R);
        base_probe model(&lam, tree, &fams, B, R, em);
        const double v = model.infer_family_likelihoods(prior, &lam);
        const auto& st = model.stash();
        *n_families = (int)fams.size();
        for (size_t f = 0; f < st.size() && (int)f < cap; ++f) out_lnl[f] = st[f].posterior_probability;
        delete em;
        delete tree;
        return v;
    } catch (...) {
        *n_families = -1;
        return 0;
    }
}
