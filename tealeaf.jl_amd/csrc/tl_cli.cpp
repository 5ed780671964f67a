// tl_cli.cpp -- command-line driver over libtealeaf_hip, the counterpart of
// run.jl (Laura7089/TeaLeaf.jl run.jl:1-47):
//   tealeaf [-s cg|cheby|ppcg|jacobi] [-x N] [-y N] [-i tea.in] [--steps N]
//           [--device D] [-p tea.problems] [--graph]
// --graph runs CG iterations as hipGraph replays (TL_LAUNCH_GRAPH).
// -x/-y re-derive dx/dy from the domain (run.jl:35-40 leaves them stale).
// Each timestep prints one JSON line with the solver statistics and the
// field summary (vol, mass, ie, temp); at the end, if check_result is on, the
// QA line of fieldsummary (kernels.jl:119-133) against the problem file.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>

#include "../../include/tealeaf_hip.h"

static const char *solver_name(int s)
{
    static const char *n[] = {"cg", "cheby", "ppcg", "jacobi"};
    return (s >= 0 && s < 4) ? n[s] : "?";
}

struct Cb {
    int solver;
    std::chrono::steady_clock::time_point t;
    double temp = 0.0;
};

static void on_step(int step, const double s[4], const tl_solve_stats *st, void *user)
{
    Cb *cb = (Cb *)user;
    const auto now = std::chrono::steady_clock::now();
    const double ms = std::chrono::duration<double, std::milli>(now - cb->t).count();
    cb->t = now;
    cb->temp = s[3];
    std::printf("{\"step\": %d, \"solver\": \"%s\", \"iterations\": %d, \"cg_iters\": %d, \"cheby_iters\": %d, "
                "\"error\": %.17g, \"rel_resid\": %.6e, \"vol\": %.17g, \"mass\": %.17g, \"ie\": %.17g, "
                "\"temp\": %.17g, \"step_ms\": %.3f}\n",
                step, solver_name(cb->solver), st->iterations, st->cg_iters, st->cheby_iters, st->error,
                st->rel_resid, s[0], s[1], s[2], s[3], ms);
    std::fflush(stdout);
}

int main(int argc, char **argv)
{
    std::string infile = "tea.in", solver, extra, problems = "tea.problems";
    int device = 0, steps = 0, graph = 0;
    for (int i = 1; i < argc; ++i) {
        const std::string a = argv[i];
        auto next = [&]() -> std::string {
            if (i + 1 >= argc) {
                std::fprintf(stderr, "missing value for %s\n", a.c_str());
                std::exit(2);
            }
            return argv[++i];
        };
        if (a == "-s" || a == "--solver")
            solver = next();
        else if (a == "-x")
            extra += "\nx_cells=" + next();
        else if (a == "-y")
            extra += "\ny_cells=" + next();
        else if (a == "-i" || a == "--in-file")
            infile = next();
        else if (a == "--steps")
            steps = std::atoi(next().c_str());
        else if (a == "--device")
            device = std::atoi(next().c_str());
        else if (a == "-p" || a == "--problems")
            problems = next();
        else if (a == "--graph")
            graph = 1;
        else if (a == "-h" || a == "--help") {
            std::printf("usage: tealeaf [-s cg|cheby|ppcg|jacobi] [-x N] [-y N] [-i tea.in] [--steps N] "
                        "[--device D] [-p tea.problems] [--graph]\n");
            return 0;
        } else {
            std::fprintf(stderr, "unknown argument %s\n", a.c_str());
            return 2;
        }
    }
    std::ifstream f(infile);
    if (!f) {
        std::fprintf(stderr, "cannot open %s\n", infile.c_str());
        return 1;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    std::string text = ss.str() + extra + "\n";
    if (!solver.empty()) {
        if (solver != "cg" && solver != "cheby" && solver != "ppcg" && solver != "jacobi") {
            std::fprintf(stderr, "solver must be cg, cheby, ppcg or jacobi\n");
            return 2;
        }
        text += "use_" + solver + "\n";
    }
    tl_deck deck;
    if (tl_deck_parse(text.c_str(), &deck) != TL_OK) {
        std::fprintf(stderr, "could not parse %s\n", infile.c_str());
        return 1;
    }
    if (graph) deck.params.launch = TL_LAUNCH_GRAPH;
    tl_ctx *ctx = nullptr;
    int rc = tl_deck_create(&ctx, &deck, device);
    if (rc != TL_OK) {
        std::fprintf(stderr, "tl_deck_create: %s\n", tl_last_error(nullptr));
        return 1;
    }
    Cb cb{deck.params.solver, std::chrono::steady_clock::now()};
    rc = tl_run(ctx, &deck, nullptr, steps, on_step, &cb);
    if (rc != TL_OK) std::fprintf(stderr, "tl_run: %s\n", tl_last_error(ctx));
    if (rc == TL_OK && deck.check_result && (steps == 0 || steps == deck.end_step)) {
        double cv = 1.0, diff = 0.0;
        int found = 0;
        const int qrc = tl_checking_value(problems.c_str(), deck.x_cells, deck.y_cells, deck.end_step, &cv, &found);
        if (qrc == TL_OK) {
            if (!found) std::fprintf(stderr, "Problem was not found in the test problems file.\n");
            const int pass = tl_qa_check(cb.temp, cv, &diff);
            std::printf("{\"check\": \"%s\", \"checking_value\": %.17g, \"actual\": %.17g, \"qa_diff\": %.6e}\n",
                        pass ? "PASSED" : "FAILED", cv, cb.temp, diff);
        } else {
            std::fprintf(stderr, "cannot read problem file %s\n", problems.c_str());
        }
    }
    tl_destroy(ctx);
    return rc == TL_OK ? 0 : 1;
}
