// cafe5_gpu.cpp -- C++ host interface of the MI355X likelihood engine; see
// cafe5_gpu.hpp for the reference interface each class mirrors.
#include "cafe5_gpu.hpp"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <deque>
#include <iomanip>
#include <limits>
#include <numeric>
#include <sstream>

#include "cafe_gpu.h"

namespace cafe5_gpu {

void check(int rc, const char* what)
{
    if (rc != CAFE_GPU_OK) throw std::runtime_error(std::string(what) + ": " + cafe_gpu_last_error());
}

gpu_context::gpu_context(int device) { check(cafe_gpu_create(device, &_ctx), "cafe_gpu_create"); }

gpu_context::~gpu_context() { cafe_gpu_destroy(_ctx); }

// ---------------------------------------------------------------------------
// clade
// ---------------------------------------------------------------------------
clade::~clade()
{
    for (clade* c : _descendants) delete c;
}

void clade::add_descendant(clade* c)
{
    c->_p_parent = this;
    _descendants.push_back(c);
}

double clade::get_branch_length() const
{
    if (_is_lambda_clade) throw std::runtime_error("Requested branch length from lambda tree");
    return _branch_length;
}

int clade::get_lambda_index() const
{
    if (!_is_lambda_clade) throw std::runtime_error("Requested lambda index from branch length tree");
    return _lambda_index;
}

std::vector<const clade*> clade::reverse_level_order() const
{
    // breadth first from this node, then reversed: children before parents
    std::vector<const clade*> bfs{this};
    for (size_t i = 0; i < bfs.size(); ++i)
        for (const clade* c : bfs[i]->_descendants) bfs.push_back(c);
    return std::vector<const clade*>(bfs.rbegin(), bfs.rend());
}

std::set<double> clade::get_branch_lengths() const
{
    std::set<double> out;
    for (const clade* c : reverse_level_order())
        if (c->get_branch_length() > 0.0) out.insert(c->get_branch_length());
    return out;
}

std::map<std::string, int> clade::get_lambda_index_map() const
{
    std::map<std::string, int> m;
    for (const clade* c : reverse_level_order()) m[c->get_taxon_name()] = c->get_lambda_index() - 1;
    return m;
}

const clade* clade::find_descendant(const std::string& name) const
{
    for (const clade* c : reverse_level_order())
        if (c->_taxon_name == name) return c;
    return nullptr;
}

void clade::apply_prefix_order(const std::function<void(const clade*)>& f) const
{
    // depth first with an explicit stack; children pushed right to left so
    // the leftmost is visited next
    std::vector<const clade*> stack{this};
    while (!stack.empty()) {
        const clade* c = stack.back();
        stack.pop_back();
        for (auto it = c->_descendants.rbegin(); it != c->_descendants.rend(); ++it) stack.push_back(*it);
        f(c);
    }
}

void clade::name_interior()
{
    for (clade* c : _descendants) c->name_interior();
    if (_descendants.empty()) return;
    std::vector<std::string> leaves;
    std::vector<const clade*> stack{this};
    while (!stack.empty()) {
        const clade* c = stack.back();
        stack.pop_back();
        if (c->_descendants.empty()) leaves.push_back(c->_taxon_name);
        for (const clade* d : c->_descendants) stack.push_back(d);
    }
    std::sort(leaves.begin(), leaves.end());
    _taxon_name.clear();
    for (auto& s : leaves) _taxon_name += s;
}

clade* parse_newick(const std::string& text, bool parse_to_lambdas)
{
    std::unique_ptr<clade> root(new clade());
    root->_is_lambda_clade = parse_to_lambdas;
    clade* cur = root.get();
    size_t i = 0;
    const size_t n = text.size();
    auto is_delim = [](char ch) { return ch == '(' || ch == ')' || ch == ',' || ch == ':' || ch == ';' || std::isspace((unsigned char)ch); };
    while (i < n) {
        const char ch = text[i];
        if (std::isspace((unsigned char)ch)) {
            ++i;
        } else if (ch == '(') {
            clade* c = new clade();
            c->_is_lambda_clade = parse_to_lambdas;
            cur->add_descendant(c);
            cur = c;
            ++i;
        } else if (ch == ',') {
            // text without the outer parentheses (clade.cpp:334-341): a comma
            // at the top level makes a new root whose first child is the
            // clade read so far
            if (cur == root.get()) {
                clade* old = root.release();
                root.reset(new clade());
                root->_is_lambda_clade = parse_to_lambdas;
                root->add_descendant(old);
            }
            clade* c = new clade();
            c->_is_lambda_clade = parse_to_lambdas;
            cur->_p_parent->add_descendant(c);
            cur = c;
            ++i;
        } else if (ch == ')') {
            if (!cur->_p_parent) throw std::runtime_error("Unbalanced parentheses in Newick text");
            cur = cur->_p_parent;
            ++i;
        } else if (ch == ';') {
            break;
        } else if (ch == ':') {
            ++i;
            size_t j = i;
            while (j < n && !is_delim(text[j])) ++j;
            const std::string num = text.substr(i, j - i);
            if (parse_to_lambdas) {
                cur->_lambda_index = (int)std::strtol(num.c_str(), nullptr, 0);
            } else {
                cur->_branch_length = std::atof(num.c_str());
            }
            i = j;
        } else {
            size_t j = i;
            while (j < n && !is_delim(text[j])) ++j;
            cur->_taxon_name = text.substr(i, j - i);
            i = j;
        }
    }
    root->name_interior();
    if (parse_to_lambdas && root->_lambda_index == 0) root->_lambda_index = 1;
    for (const clade* c : root->reverse_level_order()) {
        if (parse_to_lambdas) {
            if (c->_lambda_index < 1) throw std::runtime_error("Invalid lambda index set for " + c->_taxon_name);
        } else if (!c->is_root() && c->_branch_length <= 0) {
            throw std::runtime_error("Invalid branch length set for " + c->_taxon_name);
        }
    }
    return root.release();
}

// ---------------------------------------------------------------------------
// gene_family, lambda, error model, prior
// ---------------------------------------------------------------------------
static std::string lower(std::string s)
{
    for (auto& ch : s) ch = (char)std::tolower((unsigned char)ch);
    return s;
}

void gene_family::set_species_size(const std::string& species, int count) { _sizes[lower(species)] = count; }

int gene_family::get_species_size(const std::string& species) const
{
    auto it = _sizes.find(lower(species));
    if (it == _sizes.end()) throw std::runtime_error(species + " was not found in gene family " + _id);
    return it->second;
}

int gene_family::get_max_size() const
{
    int m = 0;
    for (auto& kv : _sizes) m = std::max(m, kv.second);
    return m;
}

lambda* multiple_lambda::multiply(double factor) const
{
    auto v = _lambdas;
    for (auto& x : v) x *= factor;
    return new multiple_lambda(_index, v);
}

double multiple_lambda::get_value_for_clade(const clade* c) const
{
    return _lambdas.at(_index.at(c->get_taxon_name()));
}

int multiple_lambda::index_of_clade(const clade* c) const
{
    return _index.at(c->get_taxon_name());
}

bool multiple_lambda::is_valid() const
{
    return std::none_of(_lambdas.begin(), _lambdas.end(), [](double d) { return d < 0; });
}

std::vector<double> get_lambda_values(const lambda* p) { return p->values(); }

static bool nearly_equal(double x, double y) { return std::abs(x - y) <= 0.01 * std::abs(x); }

void error_model::set_probabilities(size_t fam_size, const std::vector<double>& probs)
{
    if ((fam_size == 0 || _dists.empty()) && !nearly_equal(probs.at(0), 0.0))
        throw std::runtime_error("Cannot have a non-zero probability for family size 0 for negative deviation");
    if (!nearly_equal(std::accumulate(probs.begin(), probs.end(), 0.0), 1.0))
        throw std::runtime_error("Sum of probabilities must be equal to one");
    if (_dists.empty()) _dists.push_back(probs);
    if (_dists.size() <= fam_size) _dists.resize(fam_size + 1, _dists.back());
    _dists[fam_size] = probs;
}

std::vector<double> error_model::get_probs(size_t fam_size) const
{
    if (_dists.empty()) throw std::runtime_error("error model has no probabilities");
    if (fam_size >= _dists.size()) {
        if (fam_size <= _max_family_size) return _dists.back();
        throw std::runtime_error("family size " + std::to_string(fam_size) + " is beyond the error model");
    }
    return _dists[fam_size];
}

std::vector<double> error_model::get_epsilons() const
{
    std::set<double> eps;
    for (const auto& row : _dists) eps.insert(row.back());
    return std::vector<double>(eps.begin(), eps.end());
}

void error_model::replace_epsilons(const std::map<double, double>* new_epsilons)
{
    for (size_t k = 0; k < _dists.size(); ++k) {
        std::vector<double> row = _dists[k];
        if (row.size() != 3) throw std::runtime_error("replace_epsilons needs rows of three deviations");
        for (const auto& kv : *new_epsilons) {
            if (!nearly_equal(kv.first, row.back())) continue;
            row.back() = kv.second;
            if (k == 0) {
                row[1] = 1 - kv.second;
            } else {
                row.front() = kv.second;
                row[1] = 1 - kv.second * 2;
            }
            set_probabilities(k, row);
        }
    }
}

void error_model::update_single_epsilon(double new_epsilon)
{
    const std::vector<double> eps = get_epsilons();
    if (eps.size() != 1) throw std::runtime_error("update_single_epsilon needs a model with one epsilon");
    const std::map<double, double> m{{eps[0], new_epsilon}};
    replace_epsilons(&m);
}

root_equilibrium_distribution::root_equilibrium_distribution(const std::map<int, int>& rd)
{
    if (rd.empty()) throw std::runtime_error("No root distribution specified");
    long total = 0;
    int top = -1;
    for (auto& kv : rd) {
        total += kv.second;
        if (kv.second > 0) top = std::max(top, kv.first);
    }
    _pct.assign(top + 1, 0.0);
    for (int i = 0; i <= top; ++i) {
        auto it = rd.find(i);
        const int c = it == rd.end() ? 0 : it->second;
        _pct[i] = float(c) / float(total);
    }
}

root_equilibrium_distribution::root_equilibrium_distribution(size_t max_size)
{
    _pct.assign(max_size, float(1) / float(max_size));
}

static double poisspdf(int x, double lambda) { return std::exp(x * std::log(lambda) - std::lgamma(x + 1) - lambda); }

root_equilibrium_distribution::root_equilibrium_distribution(double poisson_lambda, size_t num_values)
{
    size_t filled = 0;
    for (int i = 0; filled < num_values; ++i) {
        const double pct = poisspdf(i, poisson_lambda);
        for (size_t j = 0; j < pct * num_values; ++j) ++filled;
        _pct.push_back(pct);
    }
    for (int i = 0; i < 5; ++i) _pct.push_back(poisspdf((int)_pct.size(), poisson_lambda));
}

float root_equilibrium_distribution::compute(size_t n) const
{
    if (n >= _pct.size()) return 0;
    return (float)_pct[n];
}

// ---------------------------------------------------------------------------
// matrices
// ---------------------------------------------------------------------------
void matrix::multiply(const std::vector<double>& v, int s_min, int s_max, int c_min, int c_max, double* result) const
{
    for (int s = s_min; s <= s_max; s++) {
        result[s - s_min] = 0;
        for (int c = c_min; c <= c_max; c++) result[s - s_min] += get(s, c) * v[c - c_min];
    }
}

bool matrix::is_zero() const
{
    return std::all_of(_values.begin(), _values.end(), [](double v) { return v == 0; });
}

int matrix::select_random_y(int x, int max) const
{
    const double* row = _values.data() + (size_t)x * _size;
    std::discrete_distribution<int> draw(row, row + max);
    return draw(randomizer_engine);
}

matrix_cache::key_t matrix_cache::make_key(double lambda, double branch_length)
{
    return key_t(long(branch_length * 1000), long(lambda * 1000000000));
}

matrix_cache::matrix_cache(int matrix_size, gpu_context& gpu) : _matrix_size(matrix_size), _gpu(gpu) {}

matrix_cache::~matrix_cache() {}

void matrix_cache::precalculate_matrices(const std::vector<double>& lambdas, const std::set<double>& branch_lengths)
{
    std::vector<key_t> fresh;
    for (double lam : lambdas)
        for (double t : branch_lengths) {
            const key_t k = make_key(lam, t);
            if (!_slots.count(k) && std::find(fresh.begin(), fresh.end(), k) == fresh.end()) fresh.push_back(k);
        }
    if (fresh.empty() && _capacity > 0) return;
    _gpu.owner = this;
    std::vector<key_t> build;
    const int need = (int)(_slots.size() + fresh.size());
    int first = (int)_slots.size();
    if (need > _capacity) {
        // grow the pool (this drops what is built) and rebuild every key
        _capacity = std::max(need, 2 * _capacity);
        check(cafe_gpu_matrix_reset(_gpu.get(), _matrix_size, _capacity), "cafe_gpu_matrix_reset");
        for (auto& kv : _slots) build.push_back(kv.first);
        std::sort(build.begin(), build.end(), [&](const key_t& a, const key_t& b) { return _slots[a] < _slots[b]; });
        first = 0;
        _host.clear();
    }
    for (auto& k : fresh) {
        _slots[k] = (int)_slots.size();
        build.push_back(k);
    }
    std::vector<double> lq, tq;
    for (auto& k : build) {
        tq.push_back(double(k.first) / 1000.0);
        lq.push_back(double(k.second) / 1000000000.0);
    }
    check(cafe_gpu_matrix_build(_gpu.get(), (int)build.size(), lq.data(), tq.data(), first, _gpu.build_algo),
          "cafe_gpu_matrix_build");
}

int matrix_cache::slot_of(double lambda, double branch_length) const
{
    auto it = _slots.find(make_key(lambda, branch_length));
    return it == _slots.end() ? -1 : it->second;
}

const matrix* matrix_cache::get_matrix(double branch_length, double lambda) const
{
    const int slot = slot_of(lambda, branch_length);
    if (slot < 0) {
        std::ostringstream ost;
        ost << "Failed to find matrix for " << _matrix_size << "," << branch_length << "," << lambda;
        throw std::runtime_error(ost.str());
    }
    auto it = _host.find(slot);
    if (it == _host.end()) {
        std::unique_ptr<matrix> m(new matrix(_matrix_size));
        check(cafe_gpu_matrix_get(_gpu.get(), slot, m->data()), "cafe_gpu_matrix_get");
        it = _host.emplace(slot, std::move(m)).first;
    }
    return it->second.get();
}

bool matrix_cache::is_saturated(double branch_length, double lambda)
{
    const double alpha = lambda * branch_length / (1 + lambda * branch_length);
    return (1 - 2 * alpha) < 0;
}

// P(s -> c) from the generating-function recurrence the engine uses:
// P_s(c) = a P_{s-1}(c) + (1-2a) P_{s-1}(c-1) + a P_s(c-1)
double matrix_cache::get_from_parent_fam_size_to_c(double lambda, double branch_length, int parent_size,
                                                   int child_size) const
{
    if (parent_size == 0) return child_size == 0 ? 1.0 : 0.0;
    const double lt = lambda * branch_length;
    const double alpha = lt / (1 + lt);
    const double coeff = 1 - 2 * alpha;
    if (!(coeff > 0 && coeff != 1)) return 0.0;
    std::vector<double> prev(child_size + 1, 0.0), cur(child_size + 1, 0.0);
    prev[0] = 1.0;
    for (int s = 1; s <= parent_size; ++s) {
        for (int c = 0; c <= child_size; ++c)
            cur[c] = alpha * prev[c] + (c ? coeff * prev[c - 1] + alpha * cur[c - 1] : 0.0);
        std::swap(prev, cur);
    }
    return std::min(std::max(prev[child_size], 0.0), 1.0);
}

void matrix_cache::warn_on_saturation(std::ostream& ost) const
{
    for (auto& kv : _slots) {
        const double t = double(kv.first.first) / 1000.0, l = double(kv.first.second) / 1000000000.0;
        if (is_saturated(t, l))
            ost << "WARNING: Saturated branch using lambda " << l << " on branch length " << t << std::endl;
    }
}

// ---------------------------------------------------------------------------
// models
// ---------------------------------------------------------------------------
std::ostream& operator<<(std::ostream& ost, const family_info_stash& r)
{
    ost << r.family_id << "\t" << r.lambda_multiplier << "\t" << r.category_likelihood << "\t"
        << r.family_likelihood;
    ost << "\t" << r.posterior_probability << "\t" << (r.significant ? "*" : "N/S");
    return ost;
}

model::model(lambda* p_lambda, const clade* p_tree, const std::vector<gene_family>* p_gene_families,
             int max_family_size, int max_root_family_size, error_model* p_error_model, gpu_context& gpu)
    : _p_lambda(p_lambda), _p_tree(p_tree), _p_gene_families(p_gene_families), _max_family_size(max_family_size),
      _max_root_family_size(max_root_family_size), _p_error_model(p_error_model), _gpu(gpu)
{
}

void model::set_family_bounds(std::vector<int> bound_B, std::vector<int> bound_R)
{
    _bound_B = std::move(bound_B);
    _bound_R = std::move(bound_R);
}

void model::upload()
{
    // engine node order: reverse of a breadth-first walk that enqueues
    // siblings last-to-first, so children precede parents and each parent's
    // children keep their descendant order in increasing index
    std::vector<const clade*> bfs{_p_tree};
    for (size_t i = 0; i < bfs.size(); ++i) {
        std::vector<const clade*> kids(bfs[i]->descendant_begin(), bfs[i]->descendant_end());
        for (auto it = kids.rbegin(); it != kids.rend(); ++it) bfs.push_back(*it);
    }
    _nodes.assign(bfs.rbegin(), bfs.rend());
    std::map<const clade*, int> index;
    for (size_t i = 0; i < _nodes.size(); ++i) index[_nodes[i]] = (int)i;
    std::vector<int> parent(_nodes.size()), leaf_col(_nodes.size(), -1);
    std::vector<std::string> leaves;
    for (size_t i = 0; i < _nodes.size(); ++i) {
        parent[i] = _nodes[i]->is_root() ? -1 : index.at(_nodes[i]->get_parent());
        if (_nodes[i]->is_leaf()) {
            leaf_col[i] = (int)leaves.size();
            leaves.push_back(_nodes[i]->get_taxon_name());
        }
    }
    _gpu.owner = this;
    check(cafe_gpu_set_tree(_gpu.get(), (int)_nodes.size(), parent.data(), leaf_col.data()), "cafe_gpu_set_tree");
    const int F = (int)_p_gene_families->size();
    std::vector<int> counts((size_t)F * leaves.size());
    for (int f = 0; f < F; ++f)
        for (size_t l = 0; l < leaves.size(); ++l)
            counts[(size_t)f * leaves.size() + l] = _p_gene_families->at(f).get_species_size(leaves[l]);
    if ((!_bound_B.empty() && (int)_bound_B.size() != F) || (!_bound_R.empty() && (int)_bound_R.size() != F))
        throw std::runtime_error("per-family bounds do not match the family count");
    check(cafe_gpu_set_families(_gpu.get(), F, (int)leaves.size(), counts.data(),
                                _bound_B.empty() ? nullptr : _bound_B.data(),
                                _bound_R.empty() ? nullptr : _bound_R.data(), _max_family_size,
                                _max_root_family_size),
          "cafe_gpu_set_families");
    const std::vector<double> rows = error_rows();
    if (!rows.empty())
        check(cafe_gpu_set_error_model(_gpu.get(), (int)_p_error_model->n_deviations(), _max_family_size + 1,
                                       rows.data()),
              "cafe_gpu_set_error_model");
    else
        check(cafe_gpu_set_error_model(_gpu.get(), 0, 0, nullptr), "cafe_gpu_set_error_model");
}

std::vector<double> model::error_rows() const
{
    if (!_p_error_model) return {};
    const int nd = (int)_p_error_model->n_deviations();
    std::vector<double> rows((size_t)(_max_family_size + 1) * nd);
    std::vector<double> pr;
    for (int k = 0; k <= _max_family_size; ++k) {
        // rows past the model's range repeat its last one (the reference
        // reads only observed counts' rows and leaves those undefined)
        if (_p_error_model->covers((size_t)k) || pr.empty()) pr = _p_error_model->get_probs(k);
        std::copy(pr.begin(), pr.end(), rows.begin() + (size_t)k * nd);
    }
    return rows;
}

std::vector<int> model::node_slots(matrix_cache& cache, const lambda* lam, const std::vector<double>& multipliers)
{
    std::vector<int> out;
    for (double m : multipliers) {
        std::unique_ptr<lambda> ml(lam->multiply(m));
        for (const clade* c : _nodes) {
            if (c->is_root()) {
                out.push_back(0);
                continue;
            }
            const int s = cache.slot_of(ml->get_value_for_clade(c), c->get_branch_length());
            if (s < 0) {
                std::ostringstream ost;
                ost << "Failed to find matrix for " << cache.get_matrix_size() << "," << c->get_branch_length()
                    << "," << ml->get_value_for_clade(c);
                throw std::runtime_error(ost.str());
            }
            out.push_back(s);
        }
    }
    return out;
}

std::vector<double> model::prior_row(const root_equilibrium_distribution& prior) const
{
    std::vector<double> row(_max_root_family_size);
    for (int j = 0; j < _max_root_family_size; ++j) row[j] = prior.compute(j);   // base_model.cpp:84
    return row;
}

double base_model::infer_family_likelihoods(const root_equilibrium_distribution& prior, const lambda*)
{
    // base_model.cpp:56-60: the member lambda is used, the argument ignored
    if (!_p_lambda->is_valid()) return -std::log(0.0);
    upload();
    matrix_cache calc(std::max(_max_root_family_size, _max_family_size) + 1, _gpu);
    calc.precalculate_matrices(get_lambda_values(_p_lambda), _p_tree->get_branch_lengths());
    std::vector<int> slots = node_slots(calc, _p_lambda, {1.0});
    std::vector<double> pr = prior_row(prior);
    cafe_gpu_eval_desc d{CAFE_GPU_MODEL_BASE, _gpu.path, 1, 1, slots.data(), nullptr, pr.data(),
                         _max_root_family_size};
    const size_t F = _p_gene_families->size();
    std::vector<double> lnl(F);
    double total = 0;
    check(cafe_gpu_eval(_gpu.get(), &d, lnl.data(), nullptr, nullptr, &total, nullptr), "cafe_gpu_eval");
    results.resize(F);
    for (size_t i = 0; i < F; ++i)
        results[i] = family_info_stash(_p_gene_families->at(i).id(), 0.0, 0.0, 0.0, lnl[i], false);
    return total;
}

void base_model::write_family_likelihoods(std::ostream& ost)
{
    ost << "#FamilyID\tLikelihood of Family" << std::endl;
    for (const auto& r : results) ost << r.family_id << "\t" << r.posterior_probability << std::endl;
}

gamma_model::gamma_model(lambda* p_lambda, const clade* p_tree, const std::vector<gene_family>* p_gene_families,
                         int max_family_size, int max_root_family_size, std::vector<double> gamma_categories,
                         std::vector<double> multipliers, error_model* p_error_model, gpu_context& gpu)
    : model(p_lambda, p_tree, p_gene_families, max_family_size, max_root_family_size, p_error_model, gpu),
      _gamma_cat_probs(std::move(gamma_categories)), _lambda_multipliers(std::move(multipliers))
{
}

gamma_model::gamma_model(lambda* p_lambda, const clade* p_tree, const std::vector<gene_family>* p_gene_families,
                         int max_family_size, int max_root_family_size, int n_gamma_cats, double fixed_alpha,
                         error_model* p_error_model, gpu_context& gpu)
    : model(p_lambda, p_tree, p_gene_families, max_family_size, max_root_family_size, p_error_model, gpu)
{
    _gamma_cat_probs.resize(n_gamma_cats);
    _lambda_multipliers.resize(n_gamma_cats);
    set_alpha(fixed_alpha);
}

void gamma_model::set_alpha(double alpha)
{
    _alpha = alpha;
    if (_gamma_cat_probs.size() > 1) get_gamma(_gamma_cat_probs, _lambda_multipliers, alpha);
}

bool gamma_model::can_infer() const
{
    if (!_p_lambda->is_valid()) return false;
    if (_alpha < 0) return false;
    auto v = get_lambda_values(_p_lambda);
    auto lengths = _p_tree->get_branch_lengths();
    const double longest = *std::max_element(lengths.begin(), lengths.end());
    const double largest = *std::max_element(v.begin(), v.end());
    return !matrix_cache::is_saturated(longest, largest);
}

bool gamma_model::prune(const gene_family& family, const root_equilibrium_distribution& prior, matrix_cache& calc,
                        const lambda* p_lambda, std::vector<double>& category_likelihoods)
{
    category_likelihoods.clear();
    const std::vector<gene_family> one{family};
    const std::vector<gene_family>* saved = _p_gene_families;
    _p_gene_families = &one;
    upload();
    _p_gene_families = saved;
    std::vector<int> slots = node_slots(calc, p_lambda, _lambda_multipliers);
    std::vector<double> pr = prior_row(prior);
    const int K = (int)_gamma_cat_probs.size();
    // per category: the root vector, then (gamma_core.cpp:149-162)
    // max_j L_j prior_j times p_k, false on a zero-sum vector
    cafe_gpu_eval_desc d{CAFE_GPU_MODEL_GAMMA, CAFE_GPU_PATH_PLAIN64, 1, K, slots.data(), _gamma_cat_probs.data(),
                         pr.data(), _max_root_family_size};
    std::vector<double> root(_max_root_family_size);
    for (int k = 0; k < K; ++k) {
        check(cafe_gpu_root_vectors(_gpu.get(), &d, 0, k, 0, 1, root.data()), "cafe_gpu_root_vectors");
        if (std::accumulate(root.begin(), root.end(), 0.0) == 0.0) return false;
        double best = 0;
        for (int j = 0; j < _max_root_family_size; ++j) {
            const double full = root[j] * (double)prior.compute(j);
            if (j == 0 || best < full) best = full;
        }
        category_likelihoods.push_back(best * _gamma_cat_probs[k]);
    }
    return true;
}

double gamma_model::infer_family_likelihoods(const root_equilibrium_distribution& prior, const lambda* p_lambda)
{
    results.clear();
    if (!can_infer()) return -std::log(0.0);
    upload();
    matrix_cache calc(std::max(_max_root_family_size, _max_family_size) + 1, _gpu);
    // prepare_matrices_for_simulation (gamma_core.cpp:111-121)
    std::vector<double> lams;
    for (double m : _lambda_multipliers) {
        std::unique_ptr<lambda> ml(_p_lambda->multiply(m));
        auto v = get_lambda_values(ml.get());
        lams.insert(lams.end(), v.begin(), v.end());
    }
    calc.precalculate_matrices(lams, _p_tree->get_branch_lengths());
    std::vector<int> slots = node_slots(calc, p_lambda, _lambda_multipliers);
    std::vector<double> pr = prior_row(prior);
    const int K = (int)_gamma_cat_probs.size();
    cafe_gpu_eval_desc d{CAFE_GPU_MODEL_GAMMA, _gpu.path, 1, K, slots.data(), _gamma_cat_probs.data(), pr.data(),
                         _max_root_family_size};
    const size_t F = _p_gene_families->size();
    std::vector<double> lnl(F), cat(F * K), post(F * K);
    double total = 0;
    check(cafe_gpu_eval(_gpu.get(), &d, lnl.data(), cat.data(), post.data(), &total, nullptr), "cafe_gpu_eval");
    if (std::isinf(total) && total > 0) return total;   // a zero-sum family (gamma_core.cpp:216-225)
    for (size_t i = 0; i < F; ++i) {
        double fam = 0.0;
        for (int k = 0; k < K; ++k) fam += cat[i * K + k];
        for (int k = 0; k < K; ++k)
            results.push_back(family_info_stash(_p_gene_families->at(i).id(), _lambda_multipliers[k], cat[i * K + k],
                                                fam, post[i * K + k], post[i * K + k] > 0.95));
    }
    return total;
}

void gamma_model::write_family_likelihoods(std::ostream& ost)
{
    ost << "#FamilyID\tGamma Cat Mean\tLikelihood of Category\tLikelihood of Family\tPosterior Probability\t"
           "Significant"
        << std::endl;
    for (const auto& r : results) ost << r << "\n";
}

std::vector<double> inference_prune(const gene_family& gf, matrix_cache& calc, const lambda* p_lambda,
                                    const error_model* p_error_model, const clade* p_tree, double lambda_multiplier,
                                    int max_root_family_size, int max_family_size)
{
    std::vector<gene_family> one{gf};
    // a throwaway model uploads the inputs; the caller's cache holds the matrices
    base_model m(const_cast<lambda*>(p_lambda), p_tree, &one, max_family_size, max_root_family_size,
                 const_cast<error_model*>(p_error_model), calc.gpu());
    m.upload();
    std::vector<int> slots = m.node_slots(calc, p_lambda, {lambda_multiplier});
    std::vector<double> pr(max_root_family_size, 1.0);
    cafe_gpu_eval_desc d{CAFE_GPU_MODEL_BASE, CAFE_GPU_PATH_PLAIN64, 1, 1, slots.data(), nullptr, pr.data(),
                         max_root_family_size};
    std::vector<double> root(max_root_family_size);
    check(cafe_gpu_root_vectors(calc.gpu().get(), &d, 0, 0, 0, 1, root.data()), "cafe_gpu_root_vectors");
    return root;
}

// ---------------------------------------------------------------------------
// discrete gamma (PAML; Yang 1994): published algorithms AS91 (Best &
// Roberts 1975) for chi-square quantiles, AS32 (Bhattacharjee 1970) for the
// incomplete gamma ratio and AS70 (Odeh & Evans 1974) for normal quantiles,
// with the iteration steps and tolerances the reference uses (gamma.cpp:66-215)
// ---------------------------------------------------------------------------
namespace {

double normal_quantile(double prob)
{
    static const double num[5] = {-.322232431088, -1, -.342242088547, -.0204231210245, -.453642210148e-4};
    static const double den[5] = {.0993484626060, .588581570495, .531103462366, .103537752850, .0038560700634};
    const double tail = prob < 0.5 ? prob : 1 - prob;
    if (tail < 1e-20) return -9999;
    const double y = std::sqrt(std::log(1 / (tail * tail)));
    double pn = num[4], pd = den[4];
    for (int i = 3; i >= 0; --i) {
        pn = pn * y + num[i];
        pd = pd * y + den[i];
    }
    const double z = y + pn / pd;
    return prob < 0.5 ? -z : z;
}

double incomplete_gamma_ratio(double x, double a, double ln_gamma_a)
{
    const double tol = 1e-8, big = 1e30;
    if (x == 0) return 0;
    if (x < 0 || a <= 0) return -1;
    const double lead = std::exp(a * std::log(x) - x - ln_gamma_a);
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
        const double uk = u * k;
        q[4] = w * q[2] - uk * q[0];
        q[5] = w * q[3] - uk * q[1];
        if (q[5] != 0) {
            const double next = q[4] / q[5];
            const double gap = std::fabs(frac - next);
            if (gap <= tol && gap <= tol * next) return 1 - lead * frac;
            frac = next;
        }
        for (int i = 0; i < 4; ++i) q[i] = q[i + 2];
        if (!(std::fabs(q[4]) < big))
            for (int i = 0; i < 4; ++i) q[i] /= big;
    }
}

double chi2_quantile(double p, double v)
{
    const double tol = .5e-6, ln2 = .6931471805;
    if (p < .000002 || p > .999998 || v <= 0) return -1;
    const double g = std::lgamma(v / 2), half_v = v / 2, c = half_v - 1;
    double ch;
    if (v < -1.24 * std::log(p)) {
        ch = std::pow((p * half_v * std::exp(g + half_v * ln2)), 1 / half_v);
        if (ch - tol < 0) return ch;
    } else if (v <= .32) {
        const double la = std::log(1 - p);
        double prev;
        ch = 0.4;
        do {
            prev = ch;
            const double p1 = 1 + ch * (4.67 + ch);
            const double p2 = ch * (6.73 + ch * (6.66 + ch));
            const double t = -0.5 + (4.67 + 2 * ch) / p1 - (6.73 + ch * (13.32 + 3 * ch)) / p2;
            ch -= (1 - std::exp(la + g + .5 * ch + c * ln2) * p2 / p1) / t;
        } while (!(std::fabs(prev / ch - 1) - .01 <= 0));
    } else {
        const double x = normal_quantile(p);
        const double p1 = 0.222222 / v;
        ch = v * std::pow((x * std::sqrt(p1) + 1 - p1), 3.0);
        if (ch > 2.2 * v + 6) ch = -2 * (std::log(1 - p) - c * std::log(.5 * ch) + g);
    }
    double prev;
    do {
        prev = ch;
        const double p1 = .5 * ch;
        double t = incomplete_gamma_ratio(p1, half_v, g);
        if (t < 0) return -1;
        const double p2 = p - t;
        t = p2 * std::exp(half_v * ln2 + g + p1 - c * std::log(ch));
        const double b = t / ch;
        const double a = 0.5 * t - b * c;
        const double s1 = (210 + a * (140 + a * (105 + a * (84 + a * (70 + 60 * a))))) / 420;
        const double s2 = (420 + a * (735 + a * (966 + a * (1141 + 1278 * a)))) / 2520;
        const double s3 = (210 + a * (462 + a * (707 + 932 * a))) / 2520;
        const double s4 = (252 + a * (672 + 1182 * a) + c * (294 + a * (889 + 1740 * a))) / 5040;
        const double s5 = (84 + 264 * a + c * (175 + 606 * a)) / 2520;
        const double s6 = (120 + c * (346 + 127 * c)) / 5040;
        ch += t * (1 + 0.5 * t * s1 - b * c * (s1 - b * (s2 - b * (s3 - b * (s4 - b * (s5 - b * s6))))));
    } while (std::fabs(prev / ch - 1) > tol);
    return ch;
}

}  // namespace

void get_gamma(std::vector<double>& v_freq, std::vector<double>& v_rate, double alpha)
{
    const int K = (int)v_freq.size();
    if (K == 0 || v_rate.size() != v_freq.size()) throw std::runtime_error("get_gamma: bad category vectors");
    const double beta = alpha;
    const double gap05 = 1.0 / (2.0 * K), factor = alpha / beta * K;
    double t = 0;
    for (int i = 0; i < K; i++) v_rate[i] = chi2_quantile((i * 2.0 + 1) * gap05, 2.0 * (alpha)) / (2.0 * (beta));
    for (int i = 0; i < K; i++) t += v_rate[i];
    for (int i = 0; i < K; i++) v_rate[i] *= factor / t;
    for (int i = 0; i < K; i++) v_freq[i] = 1.0 / K;
}

// ---------------------------------------------------------------------------
// p-values
// ---------------------------------------------------------------------------
std::mt19937 randomizer_engine;

void set_weighted_random_family_size(const clade* node, clademap<int>* sizemap, const lambda* p_lambda,
                                     int max_family_size, const matrix_cache& cache)
{
    if (node->is_root()) return;
    const int parent_size = (*sizemap)[node->get_parent()];
    int c = 0;   // a lost family stays lost
    if (parent_size > 0) {
        const double lam = p_lambda->get_value_for_clade(node);
        const double t = node->get_branch_length();
        const matrix* m = cache.get_matrix(t, lam);
        if (matrix_cache::is_saturated(t, lam)) {
            // the reference draws a uniform size here and then overwrites it
            // with the row draw below; the draw still advances the engine
            std::uniform_int_distribution<int> uniform(0, max_family_size - 1);
            (void)uniform(randomizer_engine);
        }
        c = m->select_random_y(parent_size, max_family_size);
    }
    (*sizemap)[node] = c;
}

clademap<int> create_family(pvalue_parameters p, int root_family_size)
{
    clademap<int> sizes;
    sizes[p.p_tree] = root_family_size;
    // no error model for simulated families (docs/architecture/0006)
    p.p_tree->apply_prefix_order([&](const clade* c) {
        set_weighted_random_family_size(c, &sizes, p.p_lambda, p.max_family_size, p.cache);
    });
    return sizes;
}

namespace {

// Root vectors of families whose leaf sizes come from clademaps, each with
// its own (B, R), on the PLAIN64 path (the reference's summation order):
// out[f] holds bound_R[f] entries, root size j+1 at index j.
std::vector<std::vector<double>> plain_root_vectors(const pvalue_parameters& p, const std::vector<gene_family>& fams,
                                                    std::vector<int> bound_B, std::vector<int> bound_R)
{
    std::vector<std::vector<double>> out(fams.size());
    if (fams.empty()) return out;
    matrix_cache& calc = const_cast<matrix_cache&>(p.cache);
    const int B = *std::max_element(bound_B.begin(), bound_B.end());
    const int R = *std::max_element(bound_R.begin(), bound_R.end());
    if (std::max(B, R) + 1 > calc.get_matrix_size())
        throw std::runtime_error("family bounds exceed the matrix size of the cache");
    std::vector<int> rows(bound_R);
    base_model m(const_cast<lambda*>(p.p_lambda), p.p_tree, &fams, B, R, nullptr, calc.gpu());
    m.set_family_bounds(std::move(bound_B), std::move(bound_R));
    m.upload();
    std::vector<int> slots = m.node_slots(calc, p.p_lambda, {1.0});
    int n_buckets = 0;
    check(cafe_gpu_bucket_count(calc.gpu().get(), &n_buckets), "cafe_gpu_bucket_count");
    std::vector<double> prior((size_t)std::max(1, n_buckets) * R, 1.0);   // root vectors only
    cafe_gpu_eval_desc d{CAFE_GPU_MODEL_BASE, CAFE_GPU_PATH_PLAIN64, 1, 1, slots.data(), nullptr, prior.data(), R};
    std::vector<double> flat(std::accumulate(rows.begin(), rows.end(), (size_t)0));
    check(cafe_gpu_root_vectors(calc.gpu().get(), &d, 0, 0, 0, (int)fams.size(), flat.data()),
          "cafe_gpu_root_vectors");
    size_t o = 0;
    for (size_t f = 0; f < fams.size(); ++f) {
        out[f].assign(flat.begin() + o, flat.begin() + o + rows[f]);
        o += rows[f];
    }
    return out;
}

// compute_family_probabilities with a root size per family
std::vector<double> family_probabilities(const pvalue_parameters& p, const std::vector<clademap<int>>& sizes,
                                         const std::vector<int>& root_sizes)
{
    const size_t F = sizes.size();
    std::vector<gene_family> fams(F);
    std::vector<int> bound_B(F), bound_R(root_sizes);
    for (size_t f = 0; f < F; ++f) {
        if (sizes[f].empty()) throw std::runtime_error("empty family size map");
        int mx = 0;
        for (const auto& kv : sizes[f]) mx = std::max(mx, kv.second);
        // probability.cpp:397-399: the interior vectors' bound
        bound_B[f] = std::min(p.max_family_size, mx + std::max(50, mx / 5));
        for (const auto& kv : sizes[f])   // gene_family::init_from_clademap
            if (kv.first->is_leaf()) fams[f].set_species_size(kv.first->get_taxon_name(), kv.second);
    }
    auto roots = plain_root_vectors(p, fams, bound_B, bound_R);
    std::vector<double> out(F);
    for (size_t f = 0; f < F; ++f) {
        // the root vector is (L(r), 0, ..., 0): its largest element
        const double v = roots[f][root_sizes[f] - 1];
        out[f] = v < 0.0 ? 0.0 : v;
    }
    return out;
}

}  // namespace

std::vector<double> compute_family_probabilities(pvalue_parameters p, const std::vector<clademap<int>>& sizes,
                                                 int root_family_size)
{
    if (root_family_size < 1) throw std::runtime_error("root family size must be positive");
    return family_probabilities(p, sizes, std::vector<int>(sizes.size(), root_family_size));
}

std::vector<double> get_random_probabilities(pvalue_parameters p, int number_of_simulations, int root_family_size)
{
    std::vector<clademap<int>> fams;
    for (int i = 0; i < number_of_simulations; ++i) fams.push_back(create_family(p, root_family_size));
    auto out = compute_family_probabilities(p, fams, root_family_size);
    std::sort(out.begin(), out.end());
    return out;
}

double pvalue(double v, const std::vector<double>& conddist)
{
    const auto above = std::upper_bound(conddist.begin(), conddist.end(), v);
    const long idx = above != conddist.end() ? long(above - conddist.begin()) : long(conddist.size()) - 1;
    return idx / (double)conddist.size();
}

double find_best_pvalue(const gene_family& fam, const std::vector<double>& root_probabilities,
                        const std::vector<std::vector<double>>& conditional_distribution)
{
    std::vector<double> pvalues(root_probabilities.size());
    const int check_to = (int)std::rint(fam.get_max_size() * 1.25);
    for (int j = 0; j < check_to && j < (int)pvalues.size() && j < (int)conditional_distribution.size(); ++j)
        pvalues[j] = pvalue(root_probabilities[j], conditional_distribution[j]);
    return pvalues.empty() ? 0.0 : *std::max_element(pvalues.begin(), pvalues.end());
}

std::vector<double> compute_pvalues(pvalue_parameters p, const std::vector<gene_family>& families,
                                    int number_of_simulations)
{
    const int R = p.max_root_family_size;
    // every root size's simulated families, drawn in the reference's order
    // (root size 1 first), then all of them in one evaluation
    std::vector<clademap<int>> sims;
    std::vector<int> roots;
    for (int i = 0; i < R; ++i)
        for (int k = 0; k < number_of_simulations; ++k) {
            sims.push_back(create_family(p, i + 1));
            roots.push_back(i + 1);
        }
    std::vector<std::vector<double>> conditional(R);
    if (!sims.empty()) {
        auto probs = family_probabilities(p, sims, roots);
        for (int i = 0; i < R; ++i) {
            conditional[i].assign(probs.begin() + (size_t)i * number_of_simulations,
                                  probs.begin() + (size_t)(i + 1) * number_of_simulations);
            std::sort(conditional[i].begin(), conditional[i].end());
        }
    }
    // the observed families' root vectors over sizes 1..R (probability.cpp:556-569)
    std::vector<double> result(families.size());
    if (families.empty()) return result;
    auto obs = plain_root_vectors(p, families, std::vector<int>(families.size(), p.max_family_size),
                                  std::vector<int>(families.size(), R));
    for (size_t i = 0; i < families.size(); ++i) result[i] = find_best_pvalue(families[i], obs[i], conditional);
    return result;
}

// ---------------------------------------------------------------------------
// input and output files
// ---------------------------------------------------------------------------
namespace {

std::vector<std::string> split_on(const std::string& text, char delim)
{
    std::vector<std::string> out;
    std::string cur;
    std::istringstream ist(text);
    while (std::getline(ist, cur, delim)) out.push_back(cur);
    return out;
}

}  // namespace

void read_gene_families(std::istream& input_file, const clade* p_tree, std::vector<gene_family>& gene_families)
{
    std::map<size_t, std::string> cafe_columns;   // CAFE layout: column -> species
    std::map<size_t, std::string> node_columns;   // CAFExp layout: column -> leaf name
    bool header = true;
    size_t hash_lines = 0;
    std::string line;
    while (std::getline(input_file, line)) {
        if (line.empty()) continue;
        const std::vector<std::string> cells = split_on(line, '\t');
        if (!node_columns.empty() && line[0] != '#') header = false;
        if (header) {
            if (line[0] == '#') {
                // one tree node per "#name" line, in column order; leaves only count
                if (!p_tree) throw std::runtime_error("No tree was provided.");
                std::string name = line.substr(1);
                if (!name.empty() && name.back() == '\r') name.pop_back();
                const clade* c = p_tree->find_descendant(name);
                if (!c) throw std::runtime_error(name + " not located in tree");
                if (c->is_leaf()) node_columns[hash_lines] = name;
                ++hash_lines;
            } else {
                header = false;
                if (node_columns.empty())
                    for (size_t i = 0; i < cells.size(); ++i) cafe_columns[i] = cells[i];
            }
            continue;
        }
        gene_family fam;
        for (size_t i = 0; i < cells.size(); ++i) {
            if (node_columns.empty()) {
                if (i == 0) fam.set_desc(cells[i]);
                else if (i == 1) fam.set_id(cells[i]);
                else fam.set_species_size(cafe_columns[i], std::atoi(cells[i].c_str()));
            } else {
                auto it = node_columns.find(i);
                if (it != node_columns.end()) fam.set_species_size(it->second, std::atoi(cells[i].c_str()));
                else if (i == cells.size() - 1) fam.set_id(cells[i]);
            }
        }
        gene_families.push_back(fam);
    }
    if (gene_families.empty()) throw std::runtime_error("No families found");
}

void read_error_model_file(std::istream& error_model_file, error_model* p_error_model)
{
    std::string line;
    while (std::getline(error_model_file, line)) {
        if (line.compare(0, 3, "max") == 0) {
            auto parts = split_on(line, ':');
            std::string v = parts.size() > 1 ? parts[1] : "";
            v.erase(std::remove_if(v.begin(), v.end(), [](unsigned char ch) { return std::isspace(ch); }), v.end());
            p_error_model->set_max_family_size((size_t)std::stoi(v));
        } else if (line.compare(0, 3, "cnt") == 0) {
            auto parts = split_on(line, ' ');
            if (parts.size() % 2 != 0)
                throw std::runtime_error("Number of different count differences in the error model (including 0) is "
                                         "not an odd number. Exiting...");
            std::vector<int> devs;
            for (size_t i = 1; i < parts.size(); ++i) devs.push_back(std::stoi(parts[i]));
            p_error_model->set_deviations(devs);
        } else {
            auto parts = split_on(line, ' ');
            if (parts.empty()) continue;
            const int size = std::stoi(parts[0]);
            std::vector<double> probs;
            for (size_t i = 1; i < parts.size(); ++i) probs.push_back(std::stod(parts[i]));
            p_error_model->set_probabilities((size_t)size, probs);
        }
    }
}

void write_error_model_file(std::ostream& ost, const error_model& model)
{
    ost << "maxcnt: " << model.get_max_family_size() - 1 << "\n";
    ost << "cntdiff:";
    for (int d : model.deviations()) ost << " " << d;
    ost << "\n";
    std::vector<double> last;
    for (size_t k = 0; k < model.get_max_family_size(); ++k) {
        const auto probs = model.get_probs(k);
        if (probs == last) continue;   // only the rows that change
        last = probs;
        ost << k;
        for (double v : probs) ost << " " << v;
        ost << std::endl;
    }
}

std::map<int, int> read_rootdist(std::istream& rootdist_file)
{
    std::map<int, int> out;
    std::string line;
    while (std::getline(rootdist_file, line)) {
        std::istringstream ist(line);
        int size = 0, count = 0;
        ist >> size >> count;
        out[size] = count;
    }
    return out;
}

// ---------------------------------------------------------------------------
// ancestral reconstruction
// ---------------------------------------------------------------------------
std::vector<clademap<int>> reconstruct_gene_families(const lambda* p_lambda, const clade* p_tree,
                                                     const std::vector<gene_family>& families, matrix_cache& calc,
                                                     const root_equilibrium_distribution& prior, int max_family_size,
                                                     int max_root_family_size)
{
    std::vector<clademap<int>> out(families.size());
    if (families.empty()) return out;
    base_model m(const_cast<lambda*>(p_lambda), p_tree, &families, max_family_size, max_root_family_size, nullptr,
                 calc.gpu());
    m.upload();
    std::vector<int> slots = m.node_slots(calc, p_lambda, {1.0});
    // gene_family_reconstructor.cpp:84-85, for the interior branches
    std::vector<const clade*> nodes = p_tree->reverse_level_order();   // every node once
    for (const clade* c : nodes)
        if (!c->is_leaf() && !c->is_root() &&
            calc.get_matrix(c->get_branch_length(), p_lambda->get_value_for_clade(c))->is_zero())
            throw std::runtime_error("Zero matrix found");
    const int Rr = std::min(max_family_size, max_root_family_size) + 1;
    std::vector<double> pr(Rr);
    for (int j = 0; j < Rr; ++j) pr[j] = prior.compute(j);   // the root scores size j with compute(j)
    // the engine's node order is the one upload() gave the tree
    std::vector<const clade*> order;
    {
        std::vector<const clade*> bfs{p_tree};
        for (size_t i = 0; i < bfs.size(); ++i) {
            std::vector<const clade*> kids(bfs[i]->descendant_begin(), bfs[i]->descendant_end());
            for (auto it = kids.rbegin(); it != kids.rend(); ++it) bfs.push_back(*it);
        }
        order.assign(bfs.rbegin(), bfs.rend());
    }
    const int n = (int)order.size();
    std::vector<int> states((size_t)families.size() * n);
    check(cafe_gpu_reconstruct(calc.gpu().get(), slots.data(), pr.data(), max_family_size, max_root_family_size, 0,
                               (int)families.size(), states.data()),
          "cafe_gpu_reconstruct");
    for (size_t f = 0; f < families.size(); ++f)
        for (int i = 0; i < n; ++i) out[f][order[i]] = states[f * n + i];
    return out;
}

// ---------------------------------------------------------------------------
// device-resident base model
// ---------------------------------------------------------------------------
resident_base_model::~resident_base_model()
{
    if (_plan) (void)cafe_gpu_plan_destroy(_plan);
}

double resident_base_model::infer_family_likelihoods(const root_equilibrium_distribution& prior, const lambda*)
{
    if (!_p_lambda->is_valid()) return -std::log(0.0);   // base_model.cpp:56-60
    const std::vector<double> lams = get_lambda_values(_p_lambda);
    const std::vector<double> pr = prior_row(prior);
    // keys: lambda i x branch length j at i * |t| + j, quantized as
    // matrix_cache_key does (matrix_cache.h:44-63)
    auto keys = [&](std::vector<double>& lq, std::vector<double>& tq) {
        lq.clear();
        tq.clear();
        for (double l : lams)
            for (double t : _t) {
                lq.push_back(double(long(l * 1000000000)) / 1000000000.0);
                tq.push_back(double(long(t * 1000)) / 1000.0);
            }
    };
    std::vector<double> lq, tq;
    // new error-model values of the same shape (an epsilon search) go to the
    // device in place: the plan folds them into its pools at every execution
    std::vector<double> err = error_rows();
    const bool err_in_place = err.size() == _plan_err.size() && !err.empty();
    if (!_plan || _gpu.owner != this || pr != _plan_prior || (int)lams.size() != _n_lambda ||
        (err != _plan_err && !err_in_place)) {
        if (_plan) {
            (void)cafe_gpu_plan_destroy(_plan);
            _plan = nullptr;
        }
        upload();
        const std::set<double> ts = _p_tree->get_branch_lengths();
        _t.assign(ts.begin(), ts.end());
        _n_lambda = (int)lams.size();
        const int n_keys = std::max(1, _n_lambda * (int)_t.size());
        check(cafe_gpu_matrix_reset(_gpu.get(), std::max(_max_root_family_size, _max_family_size) + 1, n_keys),
              "cafe_gpu_matrix_reset");
        std::vector<int> slots;
        for (const clade* c : _nodes) {
            if (c->is_root()) {
                slots.push_back(0);
                continue;
            }
            const int j = (int)(std::lower_bound(_t.begin(), _t.end(), c->get_branch_length()) - _t.begin());
            slots.push_back(_p_lambda->index_of_clade(c) * (int)_t.size() + j);
        }
        _plan_prior = pr;
        _plan_err = error_rows();
        cafe_gpu_eval_desc d{CAFE_GPU_MODEL_BASE, _gpu.path, 1, 1, slots.data(), nullptr, _plan_prior.data(),
                             _max_root_family_size};
        keys(lq, tq);
        if (lq.empty()) {   // a tree without positive branch lengths
            lq.push_back(0.0);
            tq.push_back(0.0);
        }
        check(cafe_gpu_plan_create(_gpu.get(), &d, n_keys, lq.data(), tq.data(), _gpu.build_algo, 1, &_plan),
              "cafe_gpu_plan_create");
    } else {
        keys(lq, tq);
        if (!lq.empty()) check(cafe_gpu_plan_set_keys(_plan, lq.data(), tq.data()), "cafe_gpu_plan_set_keys");
        if (err != _plan_err) {
            check(cafe_gpu_set_error_model(_gpu.get(), (int)_p_error_model->n_deviations(), _max_family_size + 1,
                                           err.data()),
                  "cafe_gpu_set_error_model");
            _plan_err.swap(err);
        }
    }
    check(cafe_gpu_plan_execute(_plan), "cafe_gpu_plan_execute");
    const size_t F = _p_gene_families->size();
    std::vector<double> lnl(F);
    double total = 0;
    check(cafe_gpu_plan_results(_plan, lnl.data(), nullptr, nullptr, &total, nullptr), "cafe_gpu_plan_results");
    results.resize(F);
    for (size_t i = 0; i < F; ++i)
        results[i] = family_info_stash(_p_gene_families->at(i).id(), 0.0, 0.0, 0.0, lnl[i], false);
    return total;
}

// ---------------------------------------------------------------------------
// optimizer
// ---------------------------------------------------------------------------
nelder_mead::nelder_mead(optimizer_scorer* s, int count)
    : scorer(s), n(count), pool(count + 1, candidate(count)), x_mean(count), x_r(count), x_tmp(count)
{
    for (auto& c : pool) candidates.push_back(&c);
}

void nelder_mead::sort()
{
    // n + 1 <= 16 candidates: libstdc++ sorts them by insertion, so equal
    // scores keep their order
    std::sort(candidates.begin(), candidates.end(), [](const candidate* a, const candidate* b) {
        return a->score < b->score;
    });
}

bool nelder_mead::check_values() const
{
    double big = -std::numeric_limits<double>::max();
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j)
            big = std::max(big, std::fabs(candidates[i + 1]->values[j] - candidates[i]->values[j]));
    return big <= tolx;
}

bool nelder_mead::check_scores() const
{
    double big = -std::numeric_limits<double>::max();
    for (int i = 1; i <= n; ++i) big = std::max(big, std::fabs(candidates[i]->score - candidates[0]->score));
    return big <= tolf;
}

void nelder_mead::init(const double* x0)
{
    for (int i = 0; i <= n; ++i) {
        // after a point that scored +inf the next step is 100x wider
        const double step = (i > 1 && std::isinf(candidates[i - 1]->score)) ? 1 + delta * 100 : 1 + delta;
        for (int j = 0; j < n; ++j)
            candidates[i]->values[j] = (i - 1 == j) ? (x0[j] ? step * x0[j] : zero_delta) : x0[j];
        candidates[i]->score = scorer->calculate_score(candidates[i]->values.data());
    }
    sort();
}

void nelder_mead::mean()
{
    for (int i = 0; i < n; ++i) {
        double s = 0;
        for (int j = 0; j < n; ++j) s += candidates[j]->values[i];
        x_mean[i] = s / n;
    }
}

double nelder_mead::reflection()
{
    for (int i = 0; i < n; ++i) x_r[i] = x_mean[i] + rho * (x_mean[i] - candidates[n]->values[i]);
    return scorer->calculate_score(x_r.data());
}

double nelder_mead::expansion()
{
    for (int i = 0; i < n; ++i) x_tmp[i] = x_mean[i] + chi * (x_r[i] - x_mean[i]);
    return scorer->calculate_score(x_tmp.data());
}

double nelder_mead::contract_outside()
{
    for (int i = 0; i < n; ++i) x_tmp[i] = x_mean[i] + psi * (x_r[i] - x_mean[i]);
    return scorer->calculate_score(x_tmp.data());
}

double nelder_mead::contract_inside()
{
    for (int i = 0; i < n; ++i) x_tmp[i] = x_mean[i] + psi * (x_mean[i] - candidates[n]->values[i]);
    return scorer->calculate_score(x_tmp.data());
}

void nelder_mead::shrink()
{
    for (int i = 1; i <= n; ++i) {
        for (int j = 0; j < n; ++j)
            candidates[i]->values[j] =
                candidates[0]->values[j] + sigma * (candidates[i]->values[j] - candidates[0]->values[j]);
        candidates[i]->score = scorer->calculate_score(candidates[i]->values.data());
    }
    sort();
}

void nelder_mead::set_last(const double* x, double f)
{
    std::copy(x, x + n, candidates.back()->values.begin());
    candidates.back()->score = f;
    sort();
}

bool nelder_mead::minimize(const double* x0)
{
    init(x0);
    int i = 0;
    for (; i < maxiters && !threshold_achieved(); ++i) {
        mean();
        const double fr = reflection();
        const double worst = candidates[n]->score;
        if (fr < candidates[0]->score) {
            const double fe = expansion();
            if (fe < fr) set_last(x_tmp.data(), fe);
            else set_last(x_r.data(), fr);
        } else if (fr >= worst) {
            if (fr > worst) {
                const double fc = contract_inside();
                if (fc < worst) set_last(x_tmp.data(), fc);
                else shrink();
            } else {
                const double fc = contract_outside();
                if (fc <= fr) set_last(x_tmp.data(), fc);
                else shrink();
            }
        } else {
            set_last(x_r.data(), fr);
        }
    }
    iters = i;
    return i == maxiters;
}

std::vector<double> optimizer::get_initial_guesses()
{
    // NUM_OPTIMIZER_INITIALIZATION_ATTEMPTS = 1 (CMakeLists.txt:13)
    auto initial = _p_scorer->initial_guesses();
    double first = _p_scorer->calculate_score(initial.data());
    for (int i = 0; std::isinf(first) && i < 1; ++i) {
        initial = _p_scorer->initial_guesses();
        first = _p_scorer->calculate_score(initial.data());
    }
    if (std::isinf(first)) throw std::runtime_error("Failed to initialize any reasonable values");
    return initial;
}

optimizer::result optimizer::optimize(const optimizer_parameters& params)
{
    const auto before = std::chrono::system_clock::now();
    result r;
    auto initial = get_initial_guesses();
    nelder_mead nm(_p_scorer, (int)initial.size());
    nm.chi = params.neldermead_expansion;
    nm.rho = params.neldermead_reflection;
    nm.maxiters = params.neldermead_iterations;
    nm.tolx = nm.tolf = 1e-6;
    nm.minimize(initial.data());
    r.score = nm.candidates[0]->score;
    r.values = nm.candidates[0]->values;
    r.num_iterations = nm.iters;
    r.duration = std::chrono::duration_cast<std::chrono::seconds>(std::chrono::system_clock::now() - before);
    return r;
}

std::ostream& operator<<(std::ostream& ost, const optimizer::result& r)
{
    if (r.score == -std::log(0.0)) {
        ost << "\nFailed to find reasonable starting values to initialize search." << std::endl;
        return ost;
    }
    const long s = (long)r.duration.count();
    ost << "\nCompleted " << r.num_iterations << " iterations" << std::endl;
    ost << "Time: " << s / 3600 << "H " << (s / 60) % 60 << "M " << s % 60 << "S" << std::endl;
    ost << "Best match" << (r.values.size() == 1 ? " is: " : "es are: ") << std::setw(15) << std::setprecision(14);
    for (size_t i = 0; i + 1 < r.values.size(); ++i) ost << r.values[i] << ',';
    if (!r.values.empty()) ost << r.values.back() << std::endl;
    ost << "Final -lnL: " << r.score << std::endl;
    return ost;
}

std::vector<double> lambda_optimizer::initial_guesses()
{
    // a normal draw centred so that lambda starts near 0.002, times
    // 1 / longest branch, redrawn while negative
    const double distmean = 0.002 / (1.0 / _longest_branch);
    std::normal_distribution<double> draw(distmean, 0.2);
    std::vector<double> out(_p_lambda->count());
    for (auto& v : out) {
        v = 1.0 / _longest_branch * draw(randomizer_engine);
        while (v < 0) v = 1.0 / _longest_branch * draw(randomizer_engine);
    }
    return out;
}

double lambda_optimizer::calculate_score(const double* values)
{
    _p_lambda->update(values);
    const double score = _p_model->infer_family_likelihoods(*_p_distribution, _p_lambda);
    return std::isnan(score) ? -std::log(0.0) : score;
}

std::vector<double> lambda_epsilon_optimizer::initial_guesses()
{
    std::vector<double> out = _lambda_optimizer.initial_guesses();
    _current = _p_error_model->get_epsilons();
    out.insert(out.end(), _current.begin(), _current.end());
    return out;
}

double lambda_epsilon_optimizer::calculate_score(const double* values)
{
    const double* eps = values + _p_lambda->count();
    _p_lambda->update(values);
    std::map<double, double> replacements;
    for (size_t i = 0; i < _current.size(); ++i) {
        replacements[_current[i]] = eps[i];
        _current[i] = eps[i];
    }
    _p_error_model->replace_epsilons(&replacements);
    const double score = _p_model->infer_family_likelihoods(*_p_distribution, _p_lambda);
    return std::isnan(score) ? -std::log(0.0) : score;
}

void lambda_epsilon_optimizer::finalize(double* results)
{
    _lambda_optimizer.finalize(results);
    _p_error_model->update_single_epsilon(results[_p_lambda->count()]);
}

std::vector<double> gamma_optimizer::initial_guesses()
{
    std::gamma_distribution<double> draw(4.0, 0.25);
    return {draw(randomizer_engine)};
}

double gamma_optimizer::calculate_score(const double* values)
{
    _p_gamma_model->set_alpha(*values);
    const double score = _p_gamma_model->infer_family_likelihoods(*_p_distribution, _p_gamma_model->get_lambda());
    return std::isnan(score) ? -std::log(0.0) : score;
}

std::vector<double> gamma_lambda_optimizer::initial_guesses()
{
    std::vector<double> out = _lambda_optimizer.initial_guesses();
    const std::vector<double> alpha = _gamma_optimizer.initial_guesses();
    out.insert(out.end(), alpha.begin(), alpha.end());
    return out;
}

double gamma_lambda_optimizer::calculate_score(const double* values)
{
    _p_lambda->update(values);
    _p_model->set_alpha(values[_p_lambda->count()]);
    const double score = _p_model->infer_family_likelihoods(*_p_distribution, _p_lambda);
    return std::isnan(score) ? -std::log(0.0) : score;
}

void gamma_lambda_optimizer::finalize(double* results)
{
    _lambda_optimizer.finalize(results);
    _gamma_optimizer.finalize(results + _p_lambda->count());
}

}  // namespace cafe5_gpu

// get_gamma for callers outside C++ (the Python mirror's gamma model): K
// category frequencies and rate multipliers at shape alpha; 0 on success.
extern "C" int cafe5_get_gamma(int K, double alpha, double* freq, double* rate)
{
    if (K < 1 || !freq || !rate) return 1;
    try {
        std::vector<double> f(K), r(K);
        cafe5_gpu::get_gamma(f, r, alpha);
        for (int i = 0; i < K; ++i) {
            freq[i] = f[i];
            rate[i] = r[i];
        }
        return 0;
    } catch (...) {
        return 1;
    }
}
