// adapter_test.cpp -- drives the C++ host interface (cafe5_amd/host) on a GPU.
//
//   adapter_test kats            the reference's known-answer tests, run
//                                through the mirrored classes; prints
//                                "PASS name" / "FAIL name ..." lines
//   adapter_test base FILE       evaluates the base or gamma model described
//                                in FILE and prints per-family values
//   adapter_test condprob FILE   compute_family_probabilities for the
//                                families of FILE at root size `root`:
//                                "PROB id value"
//   adapter_test io              the reference's file-reader tests (no GPU)
//   adapter_test example TREE FAMILIES LAMBDA B R [ERRORMODEL]
//                                reads the files as CAFE5 does and prints
//                                the base model's "TOTAL" and "LNL id v"
//   adapter_test optimize FILE   optimizes lambda for the base model of FILE
//                                (Standard Nelder-Mead, engine seeded `seed`)
//                                on a device-resident model: "START x...",
//                                one "EVAL x... score" per evaluation, then
//                                "RESULT iters score x..."
//   adapter_test pvalues FILE    compute_pvalues for the families of FILE
//                                with the engine seeded `seed`: the draws
//                                ("SIM r max c1 c2 ..."), each root size's
//                                sorted distribution ("COND r v ..."), then
//                                "PVALUE id value"
//
// FILE (written by tests/test_gpu_cpp_adapter.py):
//   newick <text>
//   lambda_tree <text>|-          (multi-lambda classes, "-" for one lambda)
//   lambdas v1 v2 ...
//   bounds B R
//   prior uniform N | poisson mu N
//   gamma K alpha | none
//   error eps | none
//   species name1 name2 ...
//   family id c1 c2 ...          (one line per family)
//   root r / seed s / sims n     (condprob, pvalues)
//   inner id v                   (condprob: interior sizes of family id)
#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <iterator>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include "cafe5_gpu.hpp"

using namespace cafe5_gpu;

static int failures = 0;

static void report(const std::string& name, bool ok, const std::string& detail = "")
{
    std::cout << (ok ? "PASS " : "FAIL ") << name << (detail.empty() ? "" : " " + detail) << std::endl;
    if (!ok) ++failures;
}

static bool close_to(double v, double want, double unit)
{
    return std::fabs(v - want) <= unit * (1 + 1e-9);
}

static int run_kats(gpu_context& gpu)
{
    // test.cpp:595-617
    {
        matrix_cache calc(5, gpu);
        single_lambda lam(0.05);
        calc.precalculate_matrices(get_lambda_values(&lam), std::set<double>{5});
        const matrix* m = calc.get_matrix(5, 0.05);
        const double want[5][5] = {{1, 0, 0, 0, 0},
                                   {0.2, 0.64, 0.128, 0.0256, 0.00512},
                                   {0.04, 0.256, 0.4608, 0.17408, 0.0512},
                                   {0.008, 0.0768, 0.26112, 0.36352, 0.187392},
                                   {0.0016, 0.02048, 0.1024, 0.249856, 0.305562}};
        bool ok = true;
        for (int i = 0; i < 5; ++i)
            for (int j = 0; j < 5; ++j) ok &= std::fabs(m->get(i, j) - want[i][j]) < 0.00001;
        report("probability_of_matrix", ok);
        report("matrix_cache_get_matrix_miss_throws", [&] {
            try {
                calc.get_matrix(6, 0.05);
                return false;
            } catch (std::runtime_error&) {
                return true;
            }
        }());
        report("get_from_parent_fam_size_to_c", close_to(calc.get_from_parent_fam_size_to_c(0.05, 5, 5, 9), 0.0152237, 1e-7),
               std::to_string(calc.get_from_parent_fam_size_to_c(0.05, 5, 5, 9)));
    }
    // test.cpp:856-863
    {
        matrix_cache calc(5, gpu);
        multiple_lambda lam(std::map<std::string, int>(), std::vector<double>{.1, .2, .3, .4});
        calc.precalculate_matrices(get_lambda_values(&lam), std::set<double>{1, 2, 3});
        report("precalculate_matrices_calculates_all_lambdas_all_branchlengths", calc.get_cache_size() == 12);
    }
    // test.cpp:1630-1644
    report("matrix_is_saturated", matrix_cache::is_saturated(25, 0.05) && !matrix_cache::is_saturated(25, 0.01));
    // test.cpp:1662-1683
    {
        std::unique_ptr<clade> tree(parse_newick("(A:1,B:3):7"));
        gene_family fam;
        fam.set_species_size("A", 3);
        fam.set_species_size("B", 6);
        single_lambda lam(0.03);
        matrix_cache cache(21, gpu);
        cache.precalculate_matrices({0.045}, std::set<double>{1.0, 3.0, 7.0});
        auto actual = inference_prune(fam, cache, &lam, nullptr, tree.get(), 1.5, 20, 20);
        const double want[] = {-17.2771, -10.0323, -5.0695, -4.91426, -5.86062, -7.75163, -10.7347, -14.2334,
                               -18.0458, -22.073, -26.2579, -30.5639, -34.9663, -39.4472, -43.9935, -48.595,
                               -53.2439, -57.9338, -62.6597, -67.4173};
        bool ok = actual.size() == 20;
        for (int i = 0; ok && i < 20; ++i) ok &= std::fabs(std::log(actual[i]) - want[i]) <= 1.2e-5 * (1 + std::fabs(want[i]));
        report("inference_prune", ok);
    }
    // test.cpp:461-488
    {
        std::unique_ptr<clade> tree(parse_newick("(A:1,B:1);"));
        std::vector<gene_family> families(4);
        const int counts[4][2] = {{1, 2}, {2, 1}, {3, 6}, {6, 3}};
        for (int f = 0; f < 4; ++f) {
            families[f].set_species_size("A", counts[f][0]);
            families[f].set_species_size("B", counts[f][1]);
        }
        single_lambda lam(0.01);
        root_equilibrium_distribution prior(100);
        for (int path = 0; path < 3; ++path) {
            gpu.path = path;
            base_model core(&lam, tree.get(), &families, 56, 8, nullptr, gpu);
            const double v = core.infer_family_likelihoods(prior, &lam);
            report("infer_processes_path" + std::to_string(path), std::fabs(v - 46.56632) < 1e-5, std::to_string(v));
        }
        gpu.path = 2;
        single_lambda bad(-0.01);
        base_model core(&bad, tree.get(), &families, 56, 8, nullptr, gpu);
        report("base_model_invalid_lambda_is_inf", std::isinf(core.infer_family_likelihoods(prior, &bad)));
    }
    // test.cpp:1269-1294
    {
        std::vector<gene_family> families(1);
        families[0].set_species_size("A", 3);
        families[0].set_species_size("B", 6);
        std::unique_ptr<clade> tree(parse_newick("(A:1,B:3):7"));
        single_lambda lam(0.005);
        std::map<int, int> rd{{1, 2}, {2, 2}, {3, 2}, {4, 2}, {5, 1}};
        root_equilibrium_distribution dist(rd);
        matrix_cache cache(11, gpu);
        cache.precalculate_matrices({0.0005, 0.0025}, std::set<double>{1, 3, 7});
        gamma_model model(&lam, tree.get(), &families, 10, 8, {0.01, 0.05}, {0.1, 0.5}, nullptr, gpu);
        std::vector<double> cat;
        const bool ok = model.prune(families[0], dist, cache, &lam, cat);
        report("gamma_model_prune", ok && cat.size() == 2 && std::fabs(std::log(cat[0]) + 23.04433) < 1e-5 &&
                                        std::fabs(std::log(cat[1]) + 16.68005) < 1e-5);
    }
    // test.cpp:1296-1311
    {
        std::vector<gene_family> families(1);
        families[0].set_species_size("A", 3);
        families[0].set_species_size("B", 6);
        std::unique_ptr<clade> tree(parse_newick("(A:1,B:3):7"));
        single_lambda lam(0.9);
        matrix_cache cache(11, gpu);
        cache.precalculate_matrices({0.09, 0.45}, std::set<double>{1, 3, 7});
        gamma_model model(&lam, tree.get(), &families, 10, 8, {1.0, 1.0}, {0.1, 0.5}, nullptr, gpu);
        std::vector<double> cat;
        root_equilibrium_distribution prior(100);
        report("gamma_model_prune_returns_false_if_saturated", !model.prune(families[0], prior, cache, &lam, cat));
        report("gamma_model_infer_saturated_is_inf", std::isinf(model.infer_family_likelihoods(prior, &lam)));
    }
    // test.cpp:490-505, 2690-2738
    {
        root_equilibrium_distribution u(10);
        std::map<int, int> m{{1, 3}, {2, 5}};
        root_equilibrium_distribution md(m);
        root_equilibrium_distribution p(0.75, 100);
        report("root_prior_values", std::fabs(u.compute(5) - 0.1) < 1e-7 && md.compute(1) == 0.375f &&
                                        std::fabs(p.compute(0) - 0.47237) < 1e-5 && p.compute(11) == 0.0f);
    }
    // test.cpp:1765-1804 (error model rows and forward fill)
    {
        error_model em;
        em.set_max_family_size(20);
        em.set_probabilities(0, {0.0, 0.8, 0.2});
        em.set_probabilities(1, {0.2, 0.6, 0.2});
        em.set_probabilities(20, {0.2, 0.6, 0.2});
        const auto r = em.get_probs(7);
        bool threw = false;
        try {
            error_model e2;
            e2.set_probabilities(0, {0.1, 0.8, 0.1});
        } catch (std::runtime_error&) {
            threw = true;
        }
        report("error_model_probabilities", r.size() == 3 && r[0] == 0.2 && r[1] == 0.6 && threw);
    }
    // gamma.cpp (no reference test pins the values; printed for the oracle)
    // test.cpp:1227-1243
    {
        std::vector<double> cd(10);
        double n = 0;
        for (auto& v : cd) v = (n += 0.01);
        report("pvalues", pvalue(0.05, cd) == 0.5 && pvalue(0.0001, cd) == 0.0 && pvalue(0.099, cd) == 0.9);
        const double v = pvalue(.35, {.1, .2, .3, .4, .5, .6, .7, .8, .9});
        report("pvalues_2", close_to(v, 3.0 / 9.0, 1e-12));
    }
    // test.cpp:2276-2300 (the Inference fixture's family: A = 1, B = 2)
    {
        gene_family fam;
        fam.set_species_size("A", 1);
        fam.set_species_size("B", 2);
        std::vector<std::vector<double>> cd(3, std::vector<double>{.1, .2, .3, .4});
        std::vector<double> root(3);
        root[0] = .15;
        const double a = find_best_pvalue(fam, root, cd);
        root[2] = .35;
        const double b = find_best_pvalue(fam, root, cd);
        report("find_best_pvalue", a == 0.25 && b == 0.25);
    }
    // test.cpp:636-654: the reference's own draw sequence (engine seeded 10)
    // over matrices built in the reference's term order
    {
        const int algo = gpu.build_algo;
        gpu.build_algo = 0;
        std::unique_ptr<clade> t(parse_newick("((A:1,B:1):1,(C:1,D:1):1);"));
        single_lambda lam(0.2);
        matrix_cache cache(15, gpu);
        cache.precalculate_matrices({0.2}, std::set<double>{1});
        pvalue_parameters p = {t.get(), &lam, 10, 8, cache};
        randomizer_engine.seed(10);
        auto sizes = create_family(p, 6);
        auto at = [&](const char* n) { return sizes.at(t->find_descendant(n)); };
        report("generate_family", at("A") == 5 && at("B") == 5 && at("C") == 7 && at("D") == 4,
               std::to_string(at("A")) + " " + std::to_string(at("B")) + " " + std::to_string(at("C")) + " " +
                   std::to_string(at("D")));
        gpu.build_algo = algo;
    }
    // test.cpp:1245-1266 (the Reconstruction fixture; interior sizes 5)
    {
        std::unique_ptr<clade> t(parse_newick("((A:1,B:3):7,(C:11,D:17):23);"));
        single_lambda lam(0.03);
        matrix_cache cache(100, gpu);
        cache.precalculate_matrices({0.03}, t->get_branch_lengths());
        pvalue_parameters p = {t.get(), &lam, 20, 15, cache};
        const std::map<std::string, int> leaf{{"A", 11}, {"B", 2}, {"C", 5}, {"D", 6}};
        clademap<int> cm;
        t->apply_prefix_order([&](const clade* c) { cm[c] = c->is_leaf() ? leaf.at(c->get_taxon_name()) : 5; });
        auto res = compute_family_probabilities(p, {cm}, 5);
        // doctest::Approx(1e-10) at its default epsilon
        const bool ok = res.size() == 1 && std::fabs(res[0] - 1e-10) < 1.1920928955078125e-05 * (1 + 1e-10);
        std::ostringstream ost;
        ost << std::setprecision(17) << (res.empty() ? -1.0 : res[0]);
        report("compute_family_probabilities", ok, ost.str());
    }
    // test.cpp:1090-1118 through the batched reconstruction
    {
        std::unique_ptr<clade> t(parse_newick("(A:1,B:3):7"));
        std::vector<gene_family> fams(1);
        fams[0].set_species_size("A", 3);
        fams[0].set_species_size("B", 6);
        single_lambda lam(0.005);
        matrix_cache cache(11, gpu);
        cache.precalculate_matrices(get_lambda_values(&lam), std::set<double>{1, 3, 7});
        std::map<int, int> rootdist;
        const int v[9] = {1, 2, 3, 4, 5, 4, 3, 2, 1};
        for (int i = 0; i < 9; ++i) rootdist[i] = v[i];
        root_equilibrium_distribution dist(rootdist);
        auto res = reconstruct_gene_families(&lam, t.get(), fams, cache, dist, 10, 8);
        const int ab = res[0].at(t.get());
        report("reconstruct_gene_family", ab == 4, std::to_string(ab));
    }
    // test.cpp:619-630
    {
        std::unique_ptr<clade> t(parse_newick("((A:1,B:1):1,(C:1,D:1):1);"));
        single_lambda lam(0.05);
        matrix_cache cache(100, gpu);
        cache.precalculate_matrices({0.05}, std::set<double>{1});
        pvalue_parameters p = {t.get(), &lam, 12, 8, cache};
        auto probs = get_random_probabilities(p, 10, 3);
        // Approx(0.001905924).scale(10000) at doctest's default epsilon
        const double tol = 1.1920928955078125e-05 * (10000 + 0.001905924);
        report("get_random_probabilities",
               probs.size() == 10 && std::is_sorted(probs.begin(), probs.end()) &&
                   std::fabs(probs[0] - 0.001905924) < tol);
    }
    // test.cpp:2265-2275 (the Inference fixture: engine seeded 10)
    {
        std::unique_ptr<clade> t(parse_newick("(A:1,B:1);"));
        single_lambda lam(0.05);
        std::vector<gene_family> fams(1);
        fams[0].set_id("TestFamily1");
        fams[0].set_species_size("A", 1);
        fams[0].set_species_size("B", 2);
        randomizer_engine.seed(10);
        matrix_cache cache(100, gpu);
        cache.precalculate_matrices(get_lambda_values(&lam), t->get_branch_lengths());
        pvalue_parameters p = {t.get(), &lam, 10, 8, cache};
        auto values = compute_pvalues(p, fams, 3);
        report("estimator_compute_pvalues", values.size() == 1 && values[0] == 0.0,
               values.empty() ? "" : std::to_string(values[0]));
    }
    // the reference's Nelder-Mead step tests (test.cpp:2802-3030) on a
    // scorer that multiplies the first two values (test.cpp:2844-2862)
    {
        struct product_scorer : optimizer_scorer {
            std::vector<double> initial_guesses() override { return {5, 3}; }
            double calculate_score(const double* v) override { return v[0] * v[1]; }
        } ms;
        auto fresh = [&]() { return nelder_mead(&ms, 2); };
        {
            auto nm = fresh();
            candidate* c[3] = {nm.candidates[0], nm.candidates[1], nm.candidates[2]};
            c[0]->score = 3, c[1]->score = 5, c[2]->score = 1;
            nm.sort();
            report("fminsearch_sort", nm.candidates[0] == c[2] && nm.candidates[1] == c[0] && nm.candidates[2] == c[1]);
        }
        {
            auto nm = fresh();
            const double v[3][2] = {{1, 3}, {2, 4}, {3, 5}};
            for (int i = 0; i < 3; ++i) nm.candidates[i]->values = {v[i][0], v[i][1]};
            nm.tolx = 3;
            const bool a = nm.check_values();
            nm.tolx = .5;
            report("fminsearch_checkV", a && !nm.check_values());
            nm.candidates[0]->score = 1, nm.candidates[1]->score = 3, nm.candidates[2]->score = 5;
            nm.tolf = 5;
            const bool b = nm.check_scores();
            nm.tolf = 1;
            report("fminsearch_checkF", b && !nm.check_scores());
        }
        {
            auto nm = fresh();
            auto init = ms.initial_guesses();
            nm.init(init.data());
            report("fminsearch_min_init", nm.candidates[0]->score == 15 && nm.candidates[1]->score == 15.75 &&
                                              close_to(nm.candidates[2]->score, 15.75, 1e-12));
        }
        {
            auto nm = fresh();
            nm.candidates[0]->values = {300, 12};
            nm.candidates[1]->values = {200, 44};
            nm.mean();
            report("fminsearch_x_mean", nm.x_mean[0] == 250 && nm.x_mean[1] == 28);
            const double fr = nm.reflection();   // the third candidate is (0, 0)
            report("fminsearch_x_reflection", nm.x_r[0] == 500 && nm.x_r[1] == 56 && fr == 28000);
            const double fe = nm.expansion();
            report("fminsearch_x_expansion", nm.x_tmp[0] == 750 && nm.x_tmp[1] == 84 && fe == 63000);
            const double fo = nm.contract_outside();
            report("fminsearch_x_contract_outside", nm.x_tmp[0] == 375 && nm.x_tmp[1] == 42 && fo == 15750);
            nm.candidates[2]->values = {26, 12};
            const double fi = nm.contract_inside();
            report("fminsearch_x_contract_inside", nm.x_tmp[0] == 362 && nm.x_tmp[1] == 36 && fi == 13032);
        }
        {
            auto nm = fresh();
            nm.candidates[0]->values = {300, 200};
            nm.candidates[1]->values = {42, 64};
            nm.candidates[2]->values = {26, 12};
            nm.shrink();
            report("fminsearch_x_shrink", nm.candidates[0]->values == std::vector<double>{300, 200} &&
                                              nm.candidates[1]->values == std::vector<double>{163, 106} &&
                                              nm.candidates[2]->values == std::vector<double>{171, 132});
        }
        {
            auto nm = fresh();
            nm.candidates[0]->values = {300, 200}, nm.candidates[0]->score = 2;
            nm.candidates[1]->values = {42, 64}, nm.candidates[1]->score = 4;
            nm.candidates[2]->values = {26, 12}, nm.candidates[2]->score = 6;
            const double x[2] = {99, 14};
            nm.set_last(x, 3);
            report("fminsearch_set_last_element", nm.candidates[0]->score == 2 && nm.candidates[1]->score == 3 &&
                                                      nm.candidates[2]->score == 4);
        }
        // test.cpp:2400-2416: guesses come from the scorer; +inf scores throw
        struct fixed_scorer : optimizer_scorer {
            bool fail = false;
            std::vector<double> initial_guesses() override { return {0.2}; }
            double calculate_score(const double*) override { return fail ? INFINITY : 1000; }
        } fs;
        optimizer opt(&fs);
        auto g = opt.get_initial_guesses();
        report("optimizer_gets_initial_guesses_from_scorer", g.size() == 1 && g[0] == 0.2);
        fs.fail = true;
        bool threw = false;
        try {
            opt.get_initial_guesses();
        } catch (std::runtime_error& e) {
            threw = std::string(e.what()) == "Failed to initialize any reasonable values";
        }
        report("optimizer_disallows_bad_initializations", threw);
    }
    {
        // test.cpp:2212-2234: a score replaces the previous epsilon by the
        // candidate's (the model's own score is not used)
        struct mock_model : model {
            using model::model;
            double infer_family_likelihoods(const root_equilibrium_distribution&, const lambda*) override { return 0; }
            void write_family_likelihoods(std::ostream&) override {}
        };
        const double initial_epsilon = 0.01;
        error_model err;
        err.set_probabilities(0, {.0, .99, initial_epsilon});
        err.set_probabilities(1, {initial_epsilon, .98, initial_epsilon});
        single_lambda lambda(0.05);
        mock_model mm(&lambda, nullptr, nullptr, 10, 10, &err, gpu);
        lambda_epsilon_optimizer opt(&mm, &err, nullptr, std::map<int, int>(), &lambda, 10);
        opt.initial_guesses();
        std::vector<double> values{0.05, 0.06};
        opt.calculate_score(values.data());
        const bool first = err.get_probs(0) == std::vector<double>({0, .94, .06});
        values[1] = 0.04;
        opt.calculate_score(values.data());
        report("lambda_epsilon_optimizer", first && err.get_probs(0) == std::vector<double>({0, .96, .04}));
    }
    {
        // test.cpp:777-792, :794-800: one guess per lambda, then alpha
        single_lambda sl(0.05);
        gamma_model m4(nullptr, nullptr, nullptr, 0, 5, 4, .25, nullptr, gpu);
        gamma_lambda_optimizer glo(&sl, &m4, nullptr, 5);
        const auto g = glo.initial_guesses();
        report("gamma_lambda_optimizer__provides_two_guesses",
               g.size() == 2 && g[0] > 0 && g[0] < 1 && g[1] > 0 && g[1] < 10);
        gamma_model m0(nullptr, nullptr, nullptr, 0, 0, 0, 0, nullptr, gpu);
        gamma_optimizer go(&m0, nullptr);
        report("gamma_optimizer__creates_single_initial_guess", go.initial_guesses().size() == 1);
    }
    {
        // test.cpp:2317-2332: a score puts lambda and alpha into the model
        std::unique_ptr<clade> tree(parse_newick("(A:1,B:1);"));
        single_lambda lam(0.05);
        std::vector<gene_family> fams(1);
        fams[0].set_id("TestFamily1");
        fams[0].set_species_size("A", 1);
        fams[0].set_species_size("B", 2);
        root_equilibrium_distribution prior((size_t)10);
        gamma_model m(&lam, tree.get(), &fams, 10, 10, std::vector<double>{0.3, 0.7}, std::vector<double>{0.5, 1.5},
                      nullptr, gpu);
        gamma_lambda_optimizer opt(&lam, &m, &prior, 7);
        std::vector<double> values{0.01, 0.25};
        const double score = opt.calculate_score(values.data());
        report("gamma_lambda_optimizer_updates_model_alpha_and_lambda",
               std::fabs(m.get_alpha() - 0.25) < 1e-5 && std::fabs(lam.get_single_lambda() - 0.01) < 1e-5 &&
                   std::isfinite(score));
    }
    {
        std::vector<double> f(4), r(4);
        get_gamma(f, r, 0.5);
        std::cout << std::setprecision(17) << "GAMMA 4 0.5";
        for (double x : r) std::cout << " " << x;
        std::cout << std::endl;
    }
    return failures ? 1 : 0;
}

static int run_model(gpu_context& gpu, const std::string& file)
{
    std::ifstream in(file);
    if (!in) throw std::runtime_error("cannot open " + file);
    std::string line, newick, lambda_tree = "-", prior_kind = "uniform";
    std::vector<double> lambdas;
    int B = 0, R = 0, K = 0, prior_n = 0;
    double alpha = 0, eps = -1, mu = 0;
    std::vector<std::string> species;
    std::vector<gene_family> families;
    while (std::getline(in, line)) {
        std::istringstream ist(line);
        std::string tag;
        ist >> tag;
        if (tag == "newick") ist >> newick;
        else if (tag == "lambda_tree") ist >> lambda_tree;
        else if (tag == "lambdas") { double v; while (ist >> v) lambdas.push_back(v); }
        else if (tag == "bounds") ist >> B >> R;
        else if (tag == "prior") {
            ist >> prior_kind;
            if (prior_kind == "poisson") ist >> mu;
            ist >> prior_n;
        } else if (tag == "gamma") {
            std::string k;
            ist >> k;
            if (k != "none") { K = std::stoi(k); ist >> alpha; }
        } else if (tag == "error") {
            std::string e;
            ist >> e;
            if (e != "none") eps = std::stod(e);
        } else if (tag == "species") { std::string s; while (ist >> s) species.push_back(s); }
        else if (tag == "path") ist >> gpu.path;
        else if (tag == "family") {
            gene_family f;
            std::string id;
            ist >> id;
            f.set_id(id);
            for (auto& s : species) { int c; ist >> c; f.set_species_size(s, c); }
            families.push_back(f);
        }
    }
    std::unique_ptr<clade> tree(parse_newick(newick));
    std::unique_ptr<lambda> lam;
    if (lambda_tree != "-") {
        std::unique_ptr<clade> lt(parse_newick(lambda_tree, true));
        lam.reset(new multiple_lambda(lt->get_lambda_index_map(), lambdas));
    } else {
        lam.reset(new single_lambda(lambdas.at(0)));
    }
    std::unique_ptr<error_model> em;
    if (eps >= 0) {
        em.reset(new error_model());
        em->set_max_family_size(B);
        em->set_probabilities(0, {0, 1 - eps, eps});
        em->set_probabilities(1, {eps, 1 - 2 * eps, eps});
    }
    std::unique_ptr<root_equilibrium_distribution> prior(
        prior_kind == "poisson" ? new root_equilibrium_distribution(mu, (size_t)prior_n)
                                : new root_equilibrium_distribution((size_t)prior_n));
    std::cout << std::setprecision(17);
    if (K > 0) {
        gamma_model m(lam.get(), tree.get(), &families, B, R, K, alpha, em.get(), gpu);
        const double total = m.infer_family_likelihoods(*prior, lam.get());
        std::cout << "TOTAL " << total << std::endl;
        for (auto& r : m.get_results())
            std::cout << "STASH " << r.family_id << " " << r.lambda_multiplier << " " << r.category_likelihood << " "
                      << r.family_likelihood << " " << r.posterior_probability << std::endl;
    } else {
        base_model m(lam.get(), tree.get(), &families, B, R, em.get(), gpu);
        const double total = m.infer_family_likelihoods(*prior, lam.get());
        std::cout << "TOTAL " << total << std::endl;
        for (auto& r : m.get_results()) std::cout << "LNL " << r.family_id << " " << r.posterior_probability << std::endl;
    }
    return 0;
}

// condprob / pvalues input: tree, one lambda, bounds, families and the
// mode's parameters
struct pvalue_input {
    std::string newick;
    double lambda = 0;
    int B = 0, R = 0, root = 1, sims = 0, algo = 1;
    unsigned seed = 0;
    std::vector<std::string> species;
    std::vector<gene_family> families;
    std::map<std::string, int> inner;
};

static pvalue_input read_pvalue_input(const std::string& file)
{
    std::ifstream in(file);
    if (!in) throw std::runtime_error("cannot open " + file);
    pvalue_input d;
    std::string line;
    while (std::getline(in, line)) {
        std::istringstream ist(line);
        std::string tag;
        ist >> tag;
        if (tag == "newick") ist >> d.newick;
        else if (tag == "lambdas") ist >> d.lambda;
        else if (tag == "bounds") ist >> d.B >> d.R;
        else if (tag == "root") ist >> d.root;
        else if (tag == "seed") ist >> d.seed;
        else if (tag == "sims") ist >> d.sims;
        else if (tag == "build") ist >> d.algo;
        else if (tag == "species") { std::string s; while (ist >> s) d.species.push_back(s); }
        else if (tag == "inner") { std::string id; int v; ist >> id >> v; d.inner[id] = v; }
        else if (tag == "family") {
            gene_family f;
            std::string id;
            ist >> id;
            f.set_id(id);
            for (auto& s : d.species) { int c; ist >> c; f.set_species_size(s, c); }
            d.families.push_back(f);
        }
    }
    return d;
}

static int run_condprob(gpu_context& gpu, const std::string& file)
{
    pvalue_input d = read_pvalue_input(file);
    gpu.build_algo = d.algo;
    std::unique_ptr<clade> tree(parse_newick(d.newick));
    single_lambda lam(d.lambda);
    matrix_cache cache(std::max(d.B, d.R) + 1, gpu);
    cache.precalculate_matrices({d.lambda}, tree->get_branch_lengths());
    pvalue_parameters p = {tree.get(), &lam, d.B, d.R, cache};
    std::vector<clademap<int>> sizes;
    for (auto& f : d.families) {
        clademap<int> cm;
        tree->apply_prefix_order([&](const clade* c) {
            cm[c] = c->is_leaf() ? f.get_species_size(c->get_taxon_name()) : d.inner.at(f.id());
        });
        sizes.push_back(cm);
    }
    auto probs = compute_family_probabilities(p, sizes, d.root);
    std::cout << std::setprecision(17);
    for (size_t i = 0; i < probs.size(); ++i) std::cout << "PROB " << d.families[i].id() << " " << probs[i] << "\n";
    return 0;
}

// Logs every point the optimizer scores.
struct logging_scorer : optimizer_scorer {
    explicit logging_scorer(optimizer_scorer& s) : inner(s) {}
    optimizer_scorer& inner;
    int n = 1;
    std::vector<double> initial_guesses() override
    {
        auto g = inner.initial_guesses();
        n = (int)g.size();
        std::cout << "START";
        for (double v : g) std::cout << " " << v;
        std::cout << "\n";
        return g;
    }
    double calculate_score(const double* x) override
    {
        const double v = inner.calculate_score(x);
        std::cout << "EVAL";
        for (int i = 0; i < n; ++i) std::cout << " " << x[i];
        std::cout << " " << v << "\n";
        return v;
    }
};

static int run_optimize(gpu_context& gpu, const std::string& file)
{
    pvalue_input d = read_pvalue_input(file);
    gpu.build_algo = d.algo;
    std::unique_ptr<clade> tree(parse_newick(d.newick));
    single_lambda lam(d.lambda);
    resident_base_model model(&lam, tree.get(), &d.families, d.B, d.R, nullptr, gpu);
    root_equilibrium_distribution prior((size_t)d.R);
    double longest = 0;
    for (double t : tree->get_branch_lengths()) longest = std::max(longest, t);
    lambda_optimizer scorer(&lam, &model, &prior, longest);
    logging_scorer log(scorer);
    std::cout << std::setprecision(17);
    randomizer_engine.seed(d.seed);
    optimizer opt(&log);
    optimizer_parameters params;
    auto r = opt.optimize(params);
    std::cout << "RESULT " << r.num_iterations << " " << r.score;
    for (double v : r.values) std::cout << " " << v;
    std::cout << "\n";
    return 0;
}

// lambda and epsilon together (lambda_epsilon_optimizer) on a device-resident
// base model with the symmetric +-1 error model {0, 1-e, e}, {e, 1-2e, e}
static int run_optimize_eps(gpu_context& gpu, const std::string& file, double epsilon)
{
    pvalue_input d = read_pvalue_input(file);
    gpu.build_algo = d.algo;
    std::unique_ptr<clade> tree(parse_newick(d.newick));
    single_lambda lam(d.lambda);
    error_model em;
    em.set_max_family_size(d.B);
    em.set_probabilities(0, {0, 1 - epsilon, epsilon});
    em.set_probabilities(1, {epsilon, 1 - 2 * epsilon, epsilon});
    resident_base_model model(&lam, tree.get(), &d.families, d.B, d.R, &em, gpu);
    root_equilibrium_distribution prior((size_t)d.R);
    double longest = 0;
    for (double t : tree->get_branch_lengths()) longest = std::max(longest, t);
    lambda_epsilon_optimizer scorer(&model, &em, &prior, std::map<int, int>(), &lam, longest);
    logging_scorer log(scorer);
    std::cout << std::setprecision(17);
    randomizer_engine.seed(d.seed);
    optimizer opt(&log);
    optimizer_parameters params;
    auto r = opt.optimize(params);
    std::cout << "RESULT " << r.num_iterations << " " << r.score;
    for (double v : r.values) std::cout << " " << v;
    std::cout << "\n";
    scorer.finalize(r.values.data());   // execute.cpp:101
    std::cout << "FINAL " << lam.get_value_for_clade(tree.get());
    for (double e : em.get_epsilons()) std::cout << " " << e;
    for (double v : em.get_probs(1)) std::cout << " " << v;
    std::cout << "\n";
    return 0;
}

// lambda and alpha together (gamma_lambda_optimizer) on a K-category gamma
// model
static int run_optimize_gamma(gpu_context& gpu, const std::string& file, int K)
{
    pvalue_input d = read_pvalue_input(file);
    gpu.build_algo = d.algo;
    std::unique_ptr<clade> tree(parse_newick(d.newick));
    single_lambda lam(d.lambda);
    gamma_model model(&lam, tree.get(), &d.families, d.B, d.R, K, 1.0, nullptr, gpu);
    root_equilibrium_distribution prior((size_t)d.R);
    double longest = 0;
    for (double t : tree->get_branch_lengths()) longest = std::max(longest, t);
    gamma_lambda_optimizer scorer(&lam, &model, &prior, longest);
    logging_scorer log(scorer);
    std::cout << std::setprecision(17);
    randomizer_engine.seed(d.seed);
    optimizer opt(&log);
    optimizer_parameters params;
    auto r = opt.optimize(params);
    std::cout << "RESULT " << r.num_iterations << " " << r.score;
    for (double v : r.values) std::cout << " " << v;
    std::cout << "\n";
    scorer.finalize(r.values.data());
    std::cout << "FINAL " << lam.get_single_lambda() << " " << model.get_alpha();
    for (double v : model.get_lambda_multipliers()) std::cout << " " << v;
    std::cout << "\n";
    return 0;
}

static int run_pvalues(gpu_context& gpu, const std::string& file)
{
    pvalue_input d = read_pvalue_input(file);
    gpu.build_algo = d.algo;
    std::unique_ptr<clade> tree(parse_newick(d.newick));
    single_lambda lam(d.lambda);
    matrix_cache cache(std::max(d.B, d.R) + 1, gpu);
    cache.precalculate_matrices({d.lambda}, tree->get_branch_lengths());
    pvalue_parameters p = {tree.get(), &lam, d.B, d.R, cache};
    std::cout << std::setprecision(17);
    // the draws compute_pvalues makes (root sizes ascending), with their
    // probabilities through compute_family_probabilities
    randomizer_engine.seed(d.seed);
    for (int r = 1; r <= d.R; ++r) {
        std::vector<clademap<int>> sims;
        for (int k = 0; k < d.sims; ++k) sims.push_back(create_family(p, r));
        for (auto& cm : sims) {
            int mx = 0;
            for (auto& kv : cm) mx = std::max(mx, kv.second);
            std::cout << "SIM " << r << " " << mx;
            for (auto& s : d.species) std::cout << " " << cm.at(tree->find_descendant(s));
            std::cout << "\n";
        }
        auto probs = compute_family_probabilities(p, sims, r);
        std::sort(probs.begin(), probs.end());
        std::cout << "COND " << r;
        for (double v : probs) std::cout << " " << v;
        std::cout << "\n";
    }
    randomizer_engine.seed(d.seed);
    auto pv = compute_pvalues(p, d.families, d.sims);
    for (size_t i = 0; i < pv.size(); ++i) std::cout << "PVALUE " << d.families[i].id() << " " << pv[i] << "\n";
    return 0;
}

static int run_io()
{
    // Newick text without the outer parentheses (clade.cpp:334-341): a
    // top-level comma makes a new root over the clades read
    auto kids = [](const clade* c) { return std::vector<const clade*>(c->descendant_begin(), c->descendant_end()); };
    {
        std::unique_ptr<clade> t(parse_newick("A:1,B:3;"));
        const auto d = kids(t.get());
        report("parse_newick_adds_a_root_over_top_level_siblings",
               t->is_root() && d.size() == 2 && d[0]->get_taxon_name() == "A" && d[1]->get_taxon_name() == "B" &&
                   d[0]->get_branch_length() == 1 && d[1]->get_branch_length() == 3 && t->get_taxon_name() == "AB");
    }
    {
        std::unique_ptr<clade> t(parse_newick("(A:1,B:1):2,(C:1,D:1):3;"));
        const auto d = kids(t.get());
        report("parse_newick_adds_a_root_over_top_level_clades",
               d.size() == 2 && kids(d[0]).size() == 2 && kids(d[1]).size() == 2 && d[0]->get_branch_length() == 2 &&
                   d[1]->get_branch_length() == 3 && t->find_descendant("C") != nullptr &&
                   t->find_descendant("A")->get_branch_length() == 1 && t->get_taxon_name() == "ABCD");
    }
    // test.cpp:367-426
    {
        std::istringstream ist("Desc\tFamily ID\tA\tB\tC\tD\n\t (null)1\t5\t10\t2\t6\n\t (null)2\t5\t10\t2\t6\n"
                               "\t (null)3\t5\t10\t2\t6\n\t (null)4\t5\t10\t2\t6");
        std::vector<gene_family> f;
        read_gene_families(ist, nullptr, f);
        report("read_gene_families_reads_cafe_files", f.size() == 4 && f[0].get_species_size("A") == 5 &&
                                                          f[0].get_species_size("B") == 10 &&
                                                          f[0].get_species_size("C") == 2 &&
                                                          f[0].get_species_size("D") == 6 && f[0].id() == " (null)1");
    }
    {
        std::istringstream ist("#A\n#B\n#AB\n#CD\n#C\n#ABCD\n#D\n35\t36\t35\t35\t36\t34\t34\t1\n98\t96\t97\t98\t98\t98\t98\t1\n");
        std::unique_ptr<clade> t(parse_newick("((A:1,B:1):1,(C:1,D:1):1);"));
        std::vector<gene_family> f;
        read_gene_families(ist, t.get(), f);
        report("read_gene_families_reads_simulation_files",
               f.size() == 2 && f[0].get_species_size("A") == 35 && f[0].get_species_size("B") == 36 &&
                   f[0].get_species_size("C") == 36 && f[0].get_species_size("D") == 34 && f[0].id() == "1");
    }
    {
        std::istringstream ist("");
        std::unique_ptr<clade> t(parse_newick("((A:1,B:1):1,(C:1,D:1):1);"));
        std::vector<gene_family> f;
        bool threw = false;
        try {
            read_gene_families(ist, t.get(), f);
        } catch (std::runtime_error& e) {
            threw = std::string(e.what()) == "No families found";
        }
        report("read_gene_families_throws_if_no_families_found", threw);
    }
    {
        std::istringstream ist("Desc\tFamily ID\tA\tB\tC\tD\n\n\n\n\t (null)1\t5\t10\t2\t6\n\n\n\n");
        std::vector<gene_family> f;
        read_gene_families(ist, nullptr, f);
        report("read_gene_families_skips_blank_lines", f.size() == 1);
    }
    // test.cpp:1552-1608
    {
        std::istringstream ist("maxcnt: 10\ncntdiff: -1 0 1\n0 0.0 0.8 0.2\n1 0.2 0.6 0.2\n2 0.2 0.6 0.2\n"
                               "3 0.2 0.6 0.2\n5 0.2 0.6 0.2\n");
        error_model m;
        read_error_model_file(ist, &m);
        auto is = [&](size_t k, std::vector<double> want) { return m.get_probs(k) == want; };
        report("read_error_model", is(0, {0.0, 0.8, 0.2}) && is(1, {0.2, 0.6, 0.2}) && is(4, {0.2, 0.6, 0.2}) &&
                                       is(7, {0.2, 0.6, 0.2}) && m.deviations() == std::vector<int>{-1, 0, 1});
    }
    {
        error_model m;
        m.set_probabilities(0, {.0, .7, .3});
        m.set_probabilities(1, {.2, .6, .2});
        m.set_probabilities(2, {.1, .8, .1});
        m.set_probabilities(3, {.2, .6, .2});
        std::ostringstream ost;
        write_error_model_file(ost, m);
        report("write_error_model",
               ost.str() == "maxcnt: 3\ncntdiff: -1 0 1\n0 0 0.7 0.3\n1 0.2 0.6 0.2\n2 0.1 0.8 0.1\n3 0.2 0.6 0.2\n");
    }
    {
        std::istringstream ist("1 1\n2 5\n3 10\n");
        auto rd = read_rootdist(ist);
        report("read_rootdist", rd.size() == 3 && rd[1] == 1 && rd[2] == 5 && rd[3] == 10);
    }
    {
        // test.cpp:1483-1495
        error_model m;
        m.set_probabilities(0, {.0, .7, .3});
        m.set_probabilities(1, {.2, .6, .2});
        m.set_probabilities(2, {.1, .8, .1});
        m.set_probabilities(3, {.2, .6, .2});
        report("error_model_get_epsilon", m.get_epsilons() == std::vector<double>({.1, .2, .3}));
    }
    {
        // test.cpp:1527-1549
        std::istringstream ist("maxcnt: 10\ncntdiff: -1 0 1\n0 0.0 0.8 0.2\n1 0.2 0.6 0.2\n2 0.2 0.6 0.2\n");
        error_model m;
        read_error_model_file(ist, &m);
        std::map<double, double> rep{{.2, .3}};
        m.replace_epsilons(&rep);
        const auto p0 = m.get_probs(0), p1 = m.get_probs(1);
        // a single epsilon: update_single_epsilon replaces it everywhere
        error_model u;
        u.set_probabilities(0, {0, .9, .1});
        u.set_probabilities(1, {.1, .8, .1});
        u.update_single_epsilon(.05);
        report("error_model_replace_epsilons", p0[1] == .7 && p0[2] == .3 && p1[0] == .3 && p1[1] == .4 &&
                                                   p1[2] == .3 && u.get_epsilons() == std::vector<double>({.05}) &&
                                                   u.get_probs(1)[1] == 1 - .05 * 2);
    }
    return failures ? 1 : 0;
}

static int run_example(gpu_context& gpu, char** a, int n)
{
    // TREE FAMILIES LAMBDA B R [ERRORMODEL]
    std::ifstream tf(a[0]);
    std::string newick((std::istreambuf_iterator<char>(tf)), std::istreambuf_iterator<char>());
    while (!newick.empty() && std::isspace((unsigned char)newick.back())) newick.pop_back();
    std::unique_ptr<clade> tree(parse_newick(newick));
    std::ifstream ff(a[1]);
    if (!ff) throw std::runtime_error(std::string("cannot open ") + a[1]);
    std::vector<gene_family> families;
    read_gene_families(ff, tree.get(), families);
    single_lambda lam(std::stod(a[2]));
    const int B = std::stoi(a[3]), R = std::stoi(a[4]);
    std::unique_ptr<error_model> em;
    if (n > 5) {
        std::ifstream ef(a[5]);
        em.reset(new error_model());
        read_error_model_file(ef, em.get());
    }
    gpu.path = 0;
    base_model m(&lam, tree.get(), &families, B, R, em.get(), gpu);
    root_equilibrium_distribution prior((size_t)R);
    const double total = m.infer_family_likelihoods(prior, &lam);
    std::cout << std::setprecision(17) << "TOTAL " << total << "\n";
    for (auto& r : m.get_results()) std::cout << "LNL " << r.family_id << " " << r.posterior_probability << "\n";
    return 0;
}

int main(int argc, char** argv)
{
    try {
        const std::string mode = argc > 1 ? argv[1] : "kats";
        if (mode == "io") return run_io();   // no device needed
        gpu_context gpu(0);
        if (mode == "example" && argc > 6) return run_example(gpu, argv + 2, argc - 2);
        if (mode == "kats") return run_kats(gpu);
        if (mode == "base" && argc > 2) return run_model(gpu, argv[2]);
        if (mode == "condprob" && argc > 2) return run_condprob(gpu, argv[2]);
        if (mode == "pvalues" && argc > 2) return run_pvalues(gpu, argv[2]);
        if (mode == "optimize" && argc > 2) {
            if (argc > 3) gpu.path = std::atoi(argv[3]);
            return run_optimize(gpu, argv[2]);
        }
        if (mode == "optimize_gamma" && argc > 3) {
            if (argc > 4) gpu.path = std::atoi(argv[4]);
            return run_optimize_gamma(gpu, argv[2], std::atoi(argv[3]));
        }
        if (mode == "optimize_eps" && argc > 3) {
            if (argc > 4) gpu.path = std::atoi(argv[4]);
            return run_optimize_eps(gpu, argv[2], std::atof(argv[3]));
        }
        std::cerr << "usage: adapter_test kats | base FILE | condprob FILE | pvalues FILE | optimize FILE [path] | "
                     "optimize_eps FILE EPSILON [path] | optimize_gamma FILE K [path]"
                  << std::endl;
        return 2;
    } catch (std::exception& e) {
        std::cerr << "error: " << e.what() << std::endl;
        return 3;
    }
}
