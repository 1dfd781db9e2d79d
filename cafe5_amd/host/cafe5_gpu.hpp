// cafe5_gpu.hpp -- C++ host interface of the MI355X likelihood engine.
//
// Mirrors the public interface of CAFE5's likelihood hot path -- the classes a
// caller of base_model / gamma_model::infer_family_likelihoods and of
// matrix_cache touches -- with the same names, argument meaning and error
// behaviour, and runs the work through the C ABI of include/cafe_gpu.h:
//
//   clade                       src/clade.h:25-121 (tree; Newick reader)
//   gene_family                 src/gene_family.h:28-86 (case-insensitive counts)
//   lambda / single_lambda /
//   multiple_lambda             src/lambda.h:21-100
//   error_model                 src/error_model.h:27-60
//   root_equilibrium_distribution  src/root_equilibrium_distribution.h:13-53
//   matrix, matrix_cache        src/matrix_cache.h:13-96
//   family_info_stash           src/core.h:19-31
//   base_model                  src/base_model.h:17-43 (infer_family_likelihoods)
//   gamma_model                 src/gamma_core.h:38-101 (infer_family_likelihoods,
//                               prune, set_alpha / get_gamma)
//
// Errors: std::runtime_error, as the reference; invalid parameters and
// saturated gamma families give +inf from infer_family_likelihoods.
#pragma once

#include <chrono>
#include <functional>
#include <map>
#include <memory>
#include <ostream>
#include <random>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

struct cafe_gpu_ctx;
struct cafe_gpu_plan;

namespace cafe5_gpu {

// ---------------------------------------------------------------------------
// device context (no reference counterpart)
// ---------------------------------------------------------------------------
class gpu_context {
public:
    explicit gpu_context(int device = 0);
    ~gpu_context();
    gpu_context(const gpu_context&) = delete;
    gpu_context& operator=(const gpu_context&) = delete;
    cafe_gpu_ctx* get() const { return _ctx; }
    // pruning path: 0 plain FP64, 1 batched FP64 MFMA, 2 batched FP32 MFMA, 3 batched split-FP16 MFMA
    int path = 2;
    // matrix build: 0 closed form (reference term order), 1 recurrence
    int build_algo = 1;
    // the object that last changed the context's tree, families or matrix
    // pool through this interface (a resident model rebuilds its plan when
    // it is not itself)
    const void* owner = nullptr;

private:
    cafe_gpu_ctx* _ctx = nullptr;
};

// throws std::runtime_error carrying cafe_gpu_last_error() when rc != 0
void check(int rc, const char* what);

// ---------------------------------------------------------------------------
// tree
// ---------------------------------------------------------------------------
class clade {
public:
    clade() = default;
    clade(std::string taxon_name, double length) : _taxon_name(std::move(taxon_name)), _branch_length(length) {}
    ~clade();
    clade(const clade&) = delete;
    clade& operator=(const clade&) = delete;

    const clade* get_parent() const { return _p_parent; }
    void add_descendant(clade* c);
    bool is_leaf() const { return _descendants.empty(); }
    bool is_root() const { return _p_parent == nullptr; }
    double get_branch_length() const;
    int get_lambda_index() const;
    std::string get_taxon_name() const { return _taxon_name; }
    std::vector<clade*>::const_iterator descendant_begin() const { return _descendants.begin(); }
    std::vector<clade*>::const_iterator descendant_end() const { return _descendants.end(); }
    // children before parents (clade.cpp:69-100)
    std::vector<const clade*> reverse_level_order() const;
    // distinct positive branch lengths (clade.cpp:236-245)
    std::set<double> get_branch_lengths() const;
    // node name -> lambda index - 1 (clade.cpp:194-204)
    std::map<std::string, int> get_lambda_index_map() const;
    const clade* find_descendant(const std::string& name) const;
    // node first, then its descendants left to right (clade.cpp:275-291)
    void apply_prefix_order(const std::function<void(const clade*)>& f) const;

    friend clade* parse_newick(const std::string& text, bool parse_to_lambdas);

private:
    void name_interior();
    clade* _p_parent = nullptr;
    std::string _taxon_name;
    double _branch_length = 0.0;
    int _lambda_index = 0;
    bool _is_lambda_clade = false;
    std::vector<clade*> _descendants;
};

// Newick text -> tree (clade.cpp:293-419 semantics: interior nodes named by
// their sorted leaf names, branch lengths (or lambda indices) after ':').
clade* parse_newick(const std::string& text, bool parse_to_lambdas = false);

// ---------------------------------------------------------------------------
// families, lambdas, error model, prior
// ---------------------------------------------------------------------------
class gene_family {
public:
    void set_id(std::string id) { _id = std::move(id); }
    std::string id() const { return _id; }
    void set_desc(std::string d) { _desc = std::move(d); }
    std::string desc() const { return _desc; }
    void set_species_size(const std::string& species, int count);
    // throws if the species is missing (gene_family.cpp:38-45)
    int get_species_size(const std::string& species) const;
    int get_max_size() const;

private:
    std::string _id, _desc;
    std::map<std::string, int> _sizes;   // lower-cased species -> count
};

class lambda {
public:
    virtual ~lambda() {}
    virtual lambda* multiply(double factor) const = 0;
    virtual double get_value_for_clade(const clade* c) const = 0;
    virtual bool is_valid() const = 0;
    virtual std::vector<double> values() const = 0;
    virtual lambda* clone() const = 0;
    // lambda.h: the optimizer's parameters in and out
    virtual void update(const double* values) = 0;
    virtual int count() const = 0;
    // which of the values a clade's branch takes
    virtual int index_of_clade(const clade* c) const = 0;
};

class single_lambda : public lambda {
public:
    explicit single_lambda(double lam) : _lambda(lam) {}
    double get_single_lambda() const { return _lambda; }
    lambda* multiply(double factor) const override { return new single_lambda(_lambda * factor); }
    double get_value_for_clade(const clade*) const override { return _lambda; }
    bool is_valid() const override { return _lambda > 0; }   // lambda.h:58-60
    std::vector<double> values() const override { return {_lambda}; }
    lambda* clone() const override { return new single_lambda(_lambda); }
    void update(double v) { _lambda = v; }
    void update(const double* values) override { _lambda = values[0]; }
    int count() const override { return 1; }
    int index_of_clade(const clade*) const override { return 0; }

private:
    double _lambda;
};

class multiple_lambda : public lambda {
public:
    multiple_lambda(std::map<std::string, int> name_to_index, std::vector<double> lambdas)
        : _index(std::move(name_to_index)), _lambdas(std::move(lambdas)) {}
    lambda* multiply(double factor) const override;
    double get_value_for_clade(const clade* c) const override;   // by clade name (lambda.cpp:32-40)
    bool is_valid() const override;                               // none negative (lambda.cpp:58-61)
    std::vector<double> values() const override { return _lambdas; }
    lambda* clone() const override { return new multiple_lambda(_index, _lambdas); }
    void update(const double* values) override { std::copy(values, values + _lambdas.size(), _lambdas.begin()); }
    int count() const override { return (int)_lambdas.size(); }
    int index_of_clade(const clade* c) const override;

private:
    std::map<std::string, int> _index;
    std::vector<double> _lambdas;
};

// lambda.cpp:69-83
std::vector<double> get_lambda_values(const lambda* p);

class error_model {
public:
    error_model() : _deviations{-1, 0, 1} {}
    void set_deviations(const std::vector<int>& d) { _deviations = d; }
    void set_max_family_size(size_t m) { _max_family_size = m; }
    // error_model.cpp:31-50 (validation and forward fill)
    void set_probabilities(size_t fam_size, const std::vector<double>& probs);
    // error_model.cpp:52-57
    std::vector<double> get_probs(size_t fam_size) const;
    size_t n_deviations() const { return _deviations.size(); }
    const std::vector<int>& deviations() const { return _deviations; }
    // rows set so far (error_model.h:55-57)
    size_t get_max_family_size() const { return _dists.size(); }
    // get_probs(k) is defined: a row was set for k, or k is within maxcnt
    bool covers(size_t k) const { return !_dists.empty() && (k < _dists.size() || k <= _max_family_size); }
    // the distinct last entries (the epsilons) of the rows, ascending
    // (error_model.cpp:59-67)
    std::vector<double> get_epsilons() const;
    // every row whose epsilon is within 1% of a key takes the key's value:
    // row 0 becomes {p0, 1 - e, e}, the others {e, 1 - 2e, e}, through
    // set_probabilities; keys are tried in ascending order against the row as
    // already replaced (error_model.cpp:79-109).  Rows must have three
    // deviations (std::runtime_error otherwise).
    void replace_epsilons(const std::map<double, double>* new_epsilons);
    // a model with one epsilon takes new_epsilon (error_model.cpp:69-77)
    void update_single_epsilon(double new_epsilon);

private:
    size_t _max_family_size = 0;
    std::vector<int> _deviations;
    std::vector<std::vector<double>> _dists;
};

class root_equilibrium_distribution {
public:
    explicit root_equilibrium_distribution(const std::map<int, int>& root_distribution);
    explicit root_equilibrium_distribution(size_t max_size);               // uniform
    root_equilibrium_distribution(double poisson_lambda, size_t num_values);
    // float valued (root_equilibrium_distribution.h:40)
    float compute(size_t n) const;

private:
    std::vector<double> _pct;
};

// ---------------------------------------------------------------------------
// transition matrices
// ---------------------------------------------------------------------------
class matrix {
public:
    explicit matrix(int size) : _size(size), _values((size_t)size * size) {}
    double get(int x, int y) const { return _values[(size_t)x * _size + y]; }
    void set(int x, int y, double v) { _values[(size_t)x * _size + y] = v; }
    int size() const { return _size; }
    // matrix_cache.cpp:32-58
    void multiply(const std::vector<double>& v, int s_min, int s_max, int c_min, int c_max, double* result) const;
    // a size y drawn from row x's entries 0..max-1 with randomizer_engine
    // (matrix_cache.cpp:60-66)
    int select_random_y(int x, int max) const;
    double* data() { return _values.data(); }
    bool is_zero() const;   // every entry 0

private:
    int _size;
    std::vector<double> _values;
};

class matrix_cache {
public:
    matrix_cache(int matrix_size, gpu_context& gpu);
    ~matrix_cache();
    // builds every missing (lambda, t) key on the GPU (matrix_cache.cpp:113-163)
    void precalculate_matrices(const std::vector<double>& lambdas, const std::set<double>& branch_lengths);
    // host copy of a built matrix; throws on a miss (matrix_cache.cpp:88-105)
    const matrix* get_matrix(double branch_length, double lambda) const;
    int get_cache_size() const { return (int)_slots.size(); }
    int get_matrix_size() const { return _matrix_size; }
    static bool is_saturated(double branch_length, double lambda);
    double get_from_parent_fam_size_to_c(double lambda, double branch_length, int parent_size, int child_size) const;
    void warn_on_saturation(std::ostream& ost) const;

    // engine slot of a key (quantized as matrix_cache_key does), -1 if absent
    int slot_of(double lambda, double branch_length) const;
    gpu_context& gpu() const { return _gpu; }

private:
    typedef std::pair<long, long> key_t;   // (branch_length * 1e3, lambda * 1e9), truncated
    static key_t make_key(double lambda, double branch_length);
    int _matrix_size;
    gpu_context& _gpu;
    int _capacity = 0;
    std::map<key_t, int> _slots;
    mutable std::map<int, std::unique_ptr<matrix>> _host;
};

// ---------------------------------------------------------------------------
// models
// ---------------------------------------------------------------------------
struct family_info_stash {
    family_info_stash() {}
    family_info_stash(std::string fam, double lam, double cat_lh, double fam_lh, double pp, bool signif)
        : family_id(std::move(fam)), lambda_multiplier(lam), category_likelihood(cat_lh),
          family_likelihood(fam_lh), posterior_probability(pp), significant(signif) {}
    std::string family_id;
    double lambda_multiplier = 0.0;
    double category_likelihood = 0.0;
    double family_likelihood = 0.0;
    double posterior_probability = 0.0;
    bool significant = false;
};
std::ostream& operator<<(std::ostream& ost, const family_info_stash& r);

class model {
public:
    model(lambda* p_lambda, const clade* p_tree, const std::vector<gene_family>* p_gene_families,
          int max_family_size, int max_root_family_size, error_model* p_error_model, gpu_context& gpu);
    virtual ~model() {}
    virtual double infer_family_likelihoods(const root_equilibrium_distribution& prior, const lambda* p_lambda) = 0;
    virtual void write_family_likelihoods(std::ostream& ost) = 0;
    const std::vector<family_info_stash>& get_results() const { return results; }
    lambda* get_lambda() const { return _p_lambda; }

    // engine plumbing: upload tree, families and error model; map every
    // branch of every (lambda x multiplier) point to its matrix slot
    void upload();
    // the error model's rows for counts 0..max_family_size as the engine takes
    // them (empty without an error model)
    std::vector<double> error_rows() const;
    // per-family (max_family_size, max_root_family_size) for the engine's size
    // buckets; empty: the model's bounds for every family
    void set_family_bounds(std::vector<int> bound_B, std::vector<int> bound_R);
    std::vector<int> node_slots(matrix_cache& cache, const lambda* lam, const std::vector<double>& multipliers);

protected:
    std::vector<double> prior_row(const root_equilibrium_distribution& prior) const;

    lambda* _p_lambda;
    const clade* _p_tree;
    const std::vector<gene_family>* _p_gene_families;
    int _max_family_size;
    int _max_root_family_size;
    error_model* _p_error_model;
    gpu_context& _gpu;
    std::vector<const clade*> _nodes;     // engine node order (children before parents)
    std::vector<int> _bound_B, _bound_R;  // per-family bounds (empty: the model's)
    std::vector<family_info_stash> results;
};

class base_model : public model {
public:
    base_model(lambda* p_lambda, const clade* p_tree, const std::vector<gene_family>* p_gene_families,
               int max_family_size, int max_root_family_size, error_model* p_error_model, gpu_context& gpu)
        : model(p_lambda, p_tree, p_gene_families, max_family_size, max_root_family_size, p_error_model, gpu) {}
    // base_model.cpp:53-100: -lnL, +inf for an invalid lambda; results hold
    // lnL_f in posterior_probability
    double infer_family_likelihoods(const root_equilibrium_distribution& prior, const lambda* p_lambda) override;
    void write_family_likelihoods(std::ostream& ost) override;   // base_model.cpp:102-109
};

class gamma_model : public model {
public:
    // explicit categories (gamma_core.h constructor 2)
    gamma_model(lambda* p_lambda, const clade* p_tree, const std::vector<gene_family>* p_gene_families,
                int max_family_size, int max_root_family_size, std::vector<double> gamma_categories,
                std::vector<double> multipliers, error_model* p_error_model, gpu_context& gpu);
    // n categories at shape alpha (gamma_core.cpp:24-34)
    gamma_model(lambda* p_lambda, const clade* p_tree, const std::vector<gene_family>* p_gene_families,
                int max_family_size, int max_root_family_size, int n_gamma_cats, double fixed_alpha,
                error_model* p_error_model, gpu_context& gpu);
    void set_alpha(double alpha);                 // gamma_core.cpp:61-67
    double get_alpha() const { return _alpha; }
    std::vector<double> get_lambda_multipliers() const { return _lambda_multipliers; }
    bool can_infer() const;                       // gamma_core.cpp:123-141
    // gamma_core.cpp:143-165 for one family: false on a zero-sum category
    bool prune(const gene_family& family, const root_equilibrium_distribution& prior, matrix_cache& calc,
               const lambda* p_lambda, std::vector<double>& category_likelihoods);
    double infer_family_likelihoods(const root_equilibrium_distribution& prior, const lambda* p_lambda) override;
    void write_family_likelihoods(std::ostream& ost) override;   // gamma_core.cpp:52-58

private:
    std::vector<double> _gamma_cat_probs;
    std::vector<double> _lambda_multipliers;
    double _alpha = 0.0;
};

// A base model whose evaluations stay on the device (no reference counterpart:
// the reference rebuilds its matrix cache and re-prunes every family on each
// call).  The first call uploads tree and families and captures one plan
// (matrix build for every lambda x branch-length key, pruning, reduction);
// later calls with new lambda values only set the plan's keys and replay it.
// The plan is rebuilt when the prior, the lambda count or the context's owner
// changed in between.
class resident_base_model : public base_model {
public:
    using base_model::base_model;
    ~resident_base_model() override;
    double infer_family_likelihoods(const root_equilibrium_distribution& prior, const lambda* p_lambda) override;

private:
    ::cafe_gpu_plan* _plan = nullptr;
    std::vector<double> _plan_prior;
    std::vector<double> _plan_err;   // error-model rows on the device
    std::vector<double> _t;   // the tree's branch lengths, key order
    int _n_lambda = 0;
};

// PAML discrete gamma rates (gamma.cpp:15-52, 225-241)
void get_gamma(std::vector<double>& v_freq, std::vector<double>& v_rate, double alpha);

// inference_prune (core.cpp:134-145) for one family on the engine: the root
// vector (max_root_family_size entries, root size j+1 at index j)
std::vector<double> inference_prune(const gene_family& gf, matrix_cache& calc, const lambda* p_lambda,
                                    const error_model* p_error_model, const clade* p_tree,
                                    double lambda_multiplier, int max_root_family_size, int max_family_size);

// ---------------------------------------------------------------------------
// p-values (probability.cpp:335-581)
// ---------------------------------------------------------------------------
template <typename T>
using clademap = std::map<const clade*, T>;

// The engine the reference's executables define (cafexp.cpp, test.cpp);
// create_family draws from it.
extern std::mt19937 randomizer_engine;

struct pvalue_parameters {
    const clade* p_tree;
    const lambda* p_lambda;
    const int max_family_size;
    const int max_root_family_size;
    const matrix_cache& cache;
};

// probability.cpp:481-507: a node's size drawn from its parent's matrix row
// (nothing for the root; leaves keep the draw, no error model)
void set_weighted_random_family_size(const clade* node, clademap<int>* sizemap, const lambda* p_lambda,
                                     int max_family_size, const matrix_cache& cache);
// probability.cpp:355-375: sizes drawn down the tree from root_family_size
clademap<int> create_family(pvalue_parameters p, int root_family_size);
// probability.cpp:379-428: per family, the likelihood of its leaf sizes given
// root size root_family_size, interior vectors stopping at
// min(max_family_size, max + max(50, max / 5)) for the largest size max in
// its map.  All families run in one engine evaluation (PLAIN64 path, the
// reference's summation order).
std::vector<double> compute_family_probabilities(pvalue_parameters p, const std::vector<clademap<int>>& sizes,
                                                 int root_family_size);
// probability.cpp:435-446: number_of_simulations drawn families, sorted
// family probabilities
std::vector<double> get_random_probabilities(pvalue_parameters p, int number_of_simulations, int root_family_size);
// probability.cpp:517-527
double pvalue(double v, const std::vector<double>& conddist);
// probability.cpp:529-541 (the loop also stops at the end of the root vector
// and of the distributions, where the reference would index past them)
double find_best_pvalue(const gene_family& fam, const std::vector<double>& root_probabilities,
                        const std::vector<std::vector<double>>& conditional_distribution);
// probability.cpp:544-581: the conditional distributions of every root size
// (their simulated families in one engine evaluation, drawn in the
// reference's order), then each family's best p-value
std::vector<double> compute_pvalues(pvalue_parameters p, const std::vector<gene_family>& families,
                                    int number_of_simulations);

// ---------------------------------------------------------------------------
// input and output files (io.cpp:134-297, user_data.cpp:105-118)
// ---------------------------------------------------------------------------
// Gene family table: the CAFE layout (a "Desc<TAB>Family ID<TAB>species..."
// header, then description, id and counts per line) or the CAFExp layout
// ("#name" lines naming the tree's nodes in column order, then counts and a
// final id column); blank lines are skipped; throws "No families found".
void read_gene_families(std::istream& input_file, const clade* p_tree, std::vector<gene_family>& gene_families);
// "maxcnt: n", "cntdiff: d...", then "size p..." rows
void read_error_model_file(std::istream& error_model_file, error_model* p_error_model);
void write_error_model_file(std::ostream& ost, const error_model& model);
// "size count" lines
std::map<int, int> read_rootdist(std::istream& rootdist_file);

// ---------------------------------------------------------------------------
// ancestral reconstruction (gene_family_reconstructor.cpp:30-189)
// ---------------------------------------------------------------------------
// Pupko's joint reconstruction of every family on the engine (one wave per
// family): per family, the reconstructed size of every interior node and the
// root (leaves keep their counts), the result of
// pupko_reconstructor::reconstruct_gene_family with vectors of sizes
// 0..max_family_size and root sizes 1..min(max_family_size,
// max_root_family_size).  The cache must hold every branch's matrix; an
// all-zero interior matrix throws "Zero matrix found" as the reference does.
std::vector<clademap<int>> reconstruct_gene_families(const lambda* p_lambda, const clade* p_tree,
                                                     const std::vector<gene_family>& families, matrix_cache& calc,
                                                     const root_equilibrium_distribution& prior, int max_family_size,
                                                     int max_root_family_size);

// ---------------------------------------------------------------------------
// optimizer (optimizer.h, optimizer.cpp, optimizer_scorer.h/.cpp)
// ---------------------------------------------------------------------------
class optimizer_scorer {
public:
    virtual ~optimizer_scorer() {}
    virtual std::vector<double> initial_guesses() = 0;
    virtual double calculate_score(const double* values) = 0;
    virtual void finalize(double* results) { (void)results; }
};

// optimizer.h:26-32 (the Standard strategy, the default build's)
struct optimizer_parameters {
    double neldermead_expansion = 2.0;
    double neldermead_reflection = 1.0;
    int neldermead_iterations = 300;
};

struct candidate {
    explicit candidate(int size) : values(size), score(0) {}
    std::vector<double> values;
    double score;
};

// FMinSearch and its steps (optimizer.cpp:61-322): the simplex of n + 1
// candidates sorted by score, reflection 1, expansion 2, contraction and
// shrink 0.5, tolerances 1e-6, initial steps of 5% (0.00025 from zero)
struct nelder_mead {
    nelder_mead(optimizer_scorer* scorer, int n);
    optimizer_scorer* scorer;
    int n;
    int maxiters = 250, iters = 0;
    double rho = 1, chi = 2, psi = 0.5, sigma = 0.5;
    double tolx = 1e-6, tolf = 1e-6;
    double delta = 0.05, zero_delta = 0.00025;
    std::vector<candidate> pool;
    std::vector<candidate*> candidates;   // sorted best first
    std::vector<double> x_mean, x_r, x_tmp;

    void sort();
    bool check_values() const;   // largest coordinate step between neighbours <= tolx
    bool check_scores() const;   // largest score gap to the best <= tolf
    bool threshold_achieved() const { return check_values() && check_scores(); }
    void init(const double* x0);
    void mean();
    double reflection();
    double expansion();
    double contract_outside();
    double contract_inside();
    void shrink();
    void set_last(const double* x, double f);
    // the search from x0; true when it stopped at the iteration cap
    bool minimize(const double* x0);
};

class optimizer {
public:
    struct result {
        std::vector<double> values;
        double score = 0;
        int num_iterations = 0;
        std::chrono::seconds duration{0};
    };
    explicit optimizer(optimizer_scorer* p_scorer) : _p_scorer(p_scorer) {}
    // optimizer.cpp:348-363: throws when the guesses score +inf
    std::vector<double> get_initial_guesses();
    // optimizer.cpp:546-573 with the Standard strategy (:365-388)
    result optimize(const optimizer_parameters& params);

private:
    optimizer_scorer* _p_scorer;
};
std::ostream& operator<<(std::ostream& ost, const optimizer::result& r);   // optimizer.cpp:575-601

// inference_optimizer_scorer + lambda_optimizer (optimizer_scorer.cpp:22-70):
// -lnL of the model at the given lambda values, NaN read as +inf
class lambda_optimizer : public optimizer_scorer {
public:
    lambda_optimizer(lambda* p_lambda, model* p_model, const root_equilibrium_distribution* p_distribution,
                     double longest_branch)
        : _p_lambda(p_lambda), _p_model(p_model), _p_distribution(p_distribution), _longest_branch(longest_branch) {}
    // normal draws around 0.002 scaled to the longest branch (:40-56)
    std::vector<double> initial_guesses() override;
    double calculate_score(const double* values) override;
    void finalize(double* results) override { _p_lambda->update(results); }

private:
    lambda* _p_lambda;
    model* _p_model;
    const root_equilibrium_distribution* _p_distribution;
    double _longest_branch;
};

// lambda_epsilon_optimizer (optimizer_scorer.h:82-107, optimizer_scorer.cpp:
// 72-108): the lambdas followed by the error model's epsilons.  A score puts
// the lambdas in place and replaces each epsilon of the previous call by the
// new one (error_model::replace_epsilons); finalize keeps the best lambdas and
// the single epsilon.  root_distribution_map is unused, as in the reference.
class lambda_epsilon_optimizer : public optimizer_scorer {
public:
    lambda_epsilon_optimizer(model* p_model, error_model* p_error_model,
                             const root_equilibrium_distribution* p_distribution,
                             const std::map<int, int>& root_distribution_map, lambda* p_lambda, double longest_branch)
        : _lambda_optimizer(p_lambda, p_model, p_distribution, longest_branch), _p_error_model(p_error_model),
          _p_lambda(p_lambda), _p_model(p_model), _p_distribution(p_distribution)
    {
        (void)root_distribution_map;
    }
    // the lambda guesses, then the model's current epsilons
    std::vector<double> initial_guesses() override;
    double calculate_score(const double* values) override;
    void finalize(double* results) override;

private:
    lambda_optimizer _lambda_optimizer;
    error_model* _p_error_model;
    lambda* _p_lambda;
    model* _p_model;
    const root_equilibrium_distribution* _p_distribution;
    std::vector<double> _current;   // the epsilons in the model now
};

// gamma_optimizer (optimizer_scorer.h:111-130, optimizer_scorer.cpp:111-145):
// the gamma shape alpha of a gamma model at its own lambda
class gamma_optimizer : public optimizer_scorer {
public:
    gamma_optimizer(gamma_model* p_model, const root_equilibrium_distribution* prior)
        : _p_gamma_model(p_model), _p_distribution(prior) {}
    // one draw of Gamma(shape 4, scale 0.25) (mean 1) from randomizer_engine
    std::vector<double> initial_guesses() override;
    // set_alpha, then -lnL of the model, NaN read as +inf
    double calculate_score(const double* values) override;
    void finalize(double* result) override { _p_gamma_model->set_alpha(*result); }
    double get_alpha() const { return _p_gamma_model->get_alpha(); }

private:
    gamma_model* _p_gamma_model;
    const root_equilibrium_distribution* _p_distribution;
};

// gamma_lambda_optimizer (optimizer_scorer.h:132-152, optimizer_scorer.cpp:
// 147-180): the lambdas followed by alpha; guesses in that order (the lambda
// draws first)
class gamma_lambda_optimizer : public optimizer_scorer {
public:
    gamma_lambda_optimizer(lambda* p_lambda, gamma_model* p_model, const root_equilibrium_distribution* p_distribution,
                           double longest_branch)
        : _lambda_optimizer(p_lambda, p_model, p_distribution, longest_branch), _gamma_optimizer(p_model, p_distribution),
          _p_lambda(p_lambda), _p_model(p_model), _p_distribution(p_distribution) {}
    std::vector<double> initial_guesses() override;
    double calculate_score(const double* values) override;
    void finalize(double* results) override;

private:
    lambda_optimizer _lambda_optimizer;
    gamma_optimizer _gamma_optimizer;
    lambda* _p_lambda;
    gamma_model* _p_model;
    const root_equilibrium_distribution* _p_distribution;
};

}  // namespace cafe5_gpu
