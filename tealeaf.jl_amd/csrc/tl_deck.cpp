// tl_deck.cpp -- tea.in parser of libtealeaf_hip.
//
// Same deck format and keys as Settings(infile) in src/settings.jl:90-173 of
// Laura7089/TeaLeaf.jl: `state N k=v ...` lines, `use_<solver>` lines,
// `key=value` lines with underscores removed from the key and
// initial_timestep mapped to dtinit; lines starting with `*` or without `=`
// are skipped; unknown keys are counted as warnings and ignored.  Whitespace
// around tokens is tolerated.  State bounds are moved inwards by dx/100 and
// dy/100 (settings.jl:159-162) using the final cell size.
//
// Keys this build adds (the reference parser warns about and ignores them):
//   precision = fp64 | fp32_ir      stop_rule = reference | relative
//   rel_tol = <float>               warm_start = none | learned
//   ir_delta = <float>              fp32_kernels = auto | two | fused
//   warm_policy = backoff | every    ppcg_pairs = auto | off    cheby_pairs = auto | off
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cerrno>
#include <climits>
#include <cstdlib>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

#include "../../include/tealeaf_hip.h"

namespace {

std::string strip(const std::string &s)
{
    size_t a = 0, b = s.size();
    while (a < b && std::isspace((unsigned char)s[a])) ++a;
    while (b > a && std::isspace((unsigned char)s[b - 1])) --b;
    return s.substr(a, b - a);
}

bool parse_int(const std::string &v, long long *out)
{
    const std::string t = strip(v);
    if (t.empty()) return false;
    char *end = nullptr;
    errno = 0;
    const long long x = std::strtoll(t.c_str(), &end, 10);
    if (errno || *end) return false;
    *out = x;
    return true;
}

bool parse_float(const std::string &v, double *out)
{
    const std::string t = strip(v);
    if (t.empty()) return false;
    char *end = nullptr;
    errno = 0;
    const double x = std::strtod(t.c_str(), &end);
    if (errno == ERANGE || *end) return false;
    *out = x;
    return true;
}

bool parse_bool(const std::string &v, int *out)
{
    const std::string t = strip(v);
    if (t == "true" || t == "1") {
        *out = 1;
        return true;
    }
    if (t == "false" || t == "0") {
        *out = 0;
        return true;
    }
    return false;
}

bool to_int(long long x, int *out)
{
    if (x < INT_MIN || x > INT_MAX) return false;
    *out = (int)x;
    return true;
}

}  // namespace

extern "C" int tl_deck_parse(const char *text, tl_deck *d)
{
    if (!text || !d) return TL_E_INVALID;
    std::memset(d, 0, sizeof(*d));
    // Settings defaults, settings.jl:41-78
    tl_default_params(&d->params);
    d->end_step = LLONG_MAX;
    d->summary_frequency = 10;
    d->halo_depth = 2;
    d->check_result = 1;
    d->dt_init = 0.1;
    d->end_time = 10.0;
    d->x_cells = 10;
    d->y_cells = 10;
    d->xmin = 0.0;
    d->ymin = 0.0;
    d->xmax = 100.0;
    d->ymax = 100.0;

    std::vector<std::string> state_lines;
    std::istringstream in(text);
    std::string raw;
    while (std::getline(in, raw)) {
        const std::string line = strip(raw);
        if (line.compare(0, 5, "state") == 0) {
            state_lines.push_back(line);
            continue;
        }
        if (line.compare(0, 4, "use_") == 0) {
            const std::string s = strip(line.substr(4));
            if (s == "cg")
                d->params.solver = TL_CG;
            else if (s == "ppcg")
                d->params.solver = TL_PPCG;
            else if (s == "cheby")
                d->params.solver = TL_CHEBY;
            else if (s == "jacobi")
                d->params.solver = TL_JACOBI;
            else
                d->n_warnings++;
            continue;
        }
        if (line.empty() || line[0] == '*' || line.find('=') == std::string::npos) continue;
        const size_t eq = line.find('=');
        std::string key;
        for (char ch : line.substr(0, eq))
            if (ch != '_') key += ch;
        key = strip(key);
        const std::string val = line.substr(eq + 1);
        if (key == "initialtimestep") key = "dtinit";
        long long iv;
        double fv;
        int bv;
        bool ok = true;
        if (key == "endstep") {
            ok = parse_int(val, &iv);
            if (ok) d->end_step = iv;
        } else if (key == "presteps") {
            ok = parse_int(val, &iv) && to_int(iv, &d->params.presteps);
        } else if (key == "maxiters") {
            ok = parse_int(val, &iv) && to_int(iv, &d->params.max_iters);
        } else if (key == "coefficient") {
            ok = parse_int(val, &iv) && to_int(iv, &d->params.coefficient);
        } else if (key == "ppcginnersteps") {
            ok = parse_int(val, &iv) && to_int(iv, &d->params.ppcg_inner_steps);
        } else if (key == "summaryfrequency") {
            ok = parse_int(val, &iv) && to_int(iv, &d->summary_frequency);
        } else if (key == "halodepth") {
            ok = parse_int(val, &iv) && to_int(iv, &d->halo_depth);
        } else if (key == "xcells") {
            ok = parse_int(val, &iv) && to_int(iv, &d->x_cells);
        } else if (key == "ycells") {
            ok = parse_int(val, &iv) && to_int(iv, &d->y_cells);
        } else if (key == "errorswitch") {
            ok = parse_bool(val, &bv);
            if (ok) d->params.error_switch = bv;
        } else if (key == "checkresult") {
            ok = parse_bool(val, &bv);
            if (ok) d->check_result = d->params.check_result = bv;
        } else if (key == "eps") {
            ok = parse_float(val, &fv);
            if (ok) d->params.eps = fv;
        } else if (key == "dtinit") {
            ok = parse_float(val, &fv);
            if (ok) d->dt_init = fv;
        } else if (key == "endtime") {
            ok = parse_float(val, &fv);
            if (ok) d->end_time = fv;
        } else if (key == "epslim") {
            ok = parse_float(val, &fv);
            if (ok) d->params.epslim = fv;
        } else if (key == "xmin") {
            ok = parse_float(val, &d->xmin);
        } else if (key == "ymin") {
            ok = parse_float(val, &d->ymin);
        } else if (key == "xmax") {
            ok = parse_float(val, &d->xmax);
        } else if (key == "ymax") {
            ok = parse_float(val, &d->ymax);
        } else if (key == "precision") {
            const std::string v = strip(val);
            if (v == "fp64")
                d->params.precision = TL_FP64;
            else if (v == "fp32ir" || v == "fp32_ir")
                d->params.precision = TL_FP32_IR;
            else
                ok = false;
        } else if (key == "stoprule") {
            const std::string v = strip(val);
            if (v == "reference")
                d->params.stop_rule = TL_STOP_REFERENCE;
            else if (v == "relative")
                d->params.stop_rule = TL_STOP_RELATIVE;
            else
                ok = false;
        } else if (key == "reltol") {
            ok = parse_float(val, &fv);
            if (ok) d->params.rel_tol = fv;
        } else if (key == "irdelta") {
            ok = parse_float(val, &fv);
            if (ok) d->params.ir_delta = fv;
        } else if (key == "fp32kernels") {
            const std::string v = strip(val);
            if (v == "auto")
                d->params.fp32_kernels = TL_FP32_KERNELS_AUTO;
            else if (v == "two")
                d->params.fp32_kernels = TL_FP32_KERNELS_TWO;
            else if (v == "fused")
                d->params.fp32_kernels = TL_FP32_KERNELS_FUSED;
            else if (v == "pairs")
                d->params.fp32_kernels = TL_FP32_KERNELS_PAIRS;
            else
                ok = false;
        } else if (key == "warmpolicy") {
            const std::string v = strip(val);
            if (v == "backoff")
                d->params.warm_policy = TL_WARM_POLICY_BACKOFF;
            else if (v == "every")
                d->params.warm_policy = TL_WARM_POLICY_EVERY;
            else
                ok = false;
        } else if (key == "chebypairs") {
            const std::string v = strip(val);
            if (v == "auto")
                d->params.cheby_pairs = TL_CHEBY_PAIRS_AUTO;
            else if (v == "off")
                d->params.cheby_pairs = TL_CHEBY_PAIRS_OFF;
            else
                ok = false;
        } else if (key == "ppcgpairs") {
            const std::string v = strip(val);
            if (v == "auto")
                d->params.ppcg_pairs = TL_PPCG_PAIRS_AUTO;
            else if (v == "off")
                d->params.ppcg_pairs = TL_PPCG_PAIRS_OFF;
            else
                ok = false;
        } else if (key == "warmstart") {
            const std::string v = strip(val);
            if (v == "none")
                d->params.warm_start = TL_WARM_NONE;
            else if (v == "learned")
                d->params.warm_start = TL_WARM_LEARNED;
            else
                ok = false;
        } else {
            ok = false;
        }
        if (!ok) d->n_warnings++;  // settings.jl:126-128: warn and ignore
    }
    if (d->x_cells < 1 || d->y_cells < 1 || d->halo_depth < 1) return TL_E_PARSE;
    // settings.jl:132-133
    d->dx = (d->xmax - d->xmin) / d->x_cells;
    d->dy = (d->ymax - d->ymin) / d->y_cells;

    // readstate, settings.jl:142-173
    for (const std::string &line : state_lines) {
        std::istringstream ls(line);
        std::string tok;
        std::vector<std::string> parts;
        while (ls >> tok) parts.push_back(tok);
        long long num;
        if (parts.size() < 2 || !parse_int(parts[1], &num) || num < 1 || num > TL_MAX_STATES) return TL_E_PARSE;
        tl_state s;
        std::memset(&s, 0, sizeof(s));
        for (size_t q = 2; q < parts.size(); ++q) {
            const size_t eq = parts[q].find('=');
            if (eq == std::string::npos) return TL_E_PARSE;
            const std::string key = parts[q].substr(0, eq), val = parts[q].substr(eq + 1);
            double v = 0;
            const bool isnum = parse_float(val, &v);
            if (key == "density") {
                if (!isnum) return TL_E_PARSE;
                s.density = v;
            } else if (key == "energy") {
                if (!isnum) return TL_E_PARSE;
                s.energy = v;
            }
            if (num == 1) continue;  // state 1 is the background: geometry ignored
            if (key == "xmin") {
                if (!isnum) return TL_E_PARSE;
                s.xmin = v + d->dx / 100;
            } else if (key == "ymin") {
                if (!isnum) return TL_E_PARSE;
                s.ymin = v + d->dy / 100;
            } else if (key == "xmax") {
                if (!isnum) return TL_E_PARSE;
                s.xmax = v - d->dx / 100;
            } else if (key == "ymax") {
                if (!isnum) return TL_E_PARSE;
                s.ymax = v - d->dy / 100;
            } else if (key == "radius") {
                if (!isnum) return TL_E_PARSE;
                s.radius = v;
            } else if (key == "geometry") {
                if (val == "rectangle")
                    s.geometry = 0;
                else if (val == "circular")
                    s.geometry = 1;
                else if (val == "point")
                    s.geometry = 2;
                else
                    return TL_E_PARSE;
            }
        }
        s.defined = 1;
        if (d->nstates >= TL_MAX_STATES) return TL_E_PARSE;
        d->states[d->nstates++] = s;  // pushed in file order, settings.jl:97-100
    }
    if (d->nstates < 1) return TL_E_PARSE;
    return TL_OK;
}

// checkingvalue, settings.jl:180-196.  Lines are split on single spaces as the
// reference does (split(line, " ")); the first three fields are integers,
// the fourth the checking value.  An empty line ends nothing: it is skipped.
extern "C" int tl_checking_value(const char *problemfile, int xcells, int ycells, int endstep, double *value,
                                 int *found)
{
    if (!problemfile || !value || !found) return TL_E_INVALID;
    *value = 1.0;
    *found = 0;
    FILE *f = std::fopen(problemfile, "r");
    if (!f) return TL_E_IO;
    std::string line;
    int rc = TL_OK;
    for (int ch = 0; ch != EOF;) {
        line.clear();
        while ((ch = std::fgetc(f)) != EOF && ch != '\n') line.push_back((char)ch);
        if (!line.empty() && line.back() == '\r') line.pop_back();
        if (strip(line).empty()) continue;
        std::vector<std::string> tok;
        size_t a = 0;
        for (;;) {
            const size_t b = line.find(' ', a);
            tok.push_back(line.substr(a, b == std::string::npos ? std::string::npos : b - a));
            if (b == std::string::npos) break;
            a = b + 1;
        }
        long long v[3];
        double cv;
        if (tok.size() < 4 || !parse_int(tok[0], &v[0]) || !parse_int(tok[1], &v[1]) || !parse_int(tok[2], &v[2]) ||
            !parse_float(tok[3], &cv)) {
            rc = TL_E_PARSE;
            break;
        }
        if (v[0] == xcells && v[1] == ycells && v[2] == endstep) {
            *value = cv;
            *found = 1;
            break;
        }
    }
    std::fclose(f);
    return rc;
}

// fieldsummary's PASSED/FAILED rule, kernels.jl:127-132
extern "C" int tl_qa_check(double actual, double checking_value, double *qa_diff)
{
    const double d = std::fabs(100.0 * actual / checking_value - 100.0);
    if (qa_diff) *qa_diff = d;
    return d < 0.001 ? 1 : 0;
}
