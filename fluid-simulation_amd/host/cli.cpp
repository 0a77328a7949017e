// Command line of the headless driver.  The flags, their defaults and the
// error messages are the reference's interface (src/cli.h, src/cli.cu:7-101);
// the parser is a table: each flag's entry carries its metavariable, its
// help text and the action that stores its value, and the usage text is
// generated from the same table.
#include "cli.h"

#include <cstdlib>
#include <functional>
#include <iomanip>
#include <iostream>
#include <sstream>
#include <utility>
#include <vector>

#include "apps.h"

namespace wcsph_host {

namespace {

// Names of an enumeration's values, in one table for parsing and printing.
template <class E>
using NameTable = std::vector<std::pair<const char*, E>>;

const NameTable<RunOptions::Searcher>& searchers() {
    static const NameTable<RunOptions::Searcher> t = {
        {"mcmc", RunOptions::Searcher::Mcmc}, {"random", RunOptions::Searcher::Random},
        {"full", RunOptions::Searcher::Full}};
    return t;
}

const NameTable<RunOptions::StopKind>& stop_kinds() {
    static const NameTable<RunOptions::StopKind> t = {
        {"iters", RunOptions::StopKind::Iters}, {"sim-time", RunOptions::StopKind::SimTime},
        {"wall-time", RunOptions::StopKind::WallTime}};
    return t;
}

template <class E>
bool lookup(const NameTable<E>& t, const std::string& name, E& out) {
    for (const auto& [n, v] : t)
        if (name == n) {
            out = v;
            return true;
        }
    return false;
}

template <class E>
const char* name_of(const NameTable<E>& t, E v, const char* fallback) {
    for (const auto& [n, x] : t)
        if (x == v) return n;
    return fallback;
}

// "--stop KIND=VALUE": the kind from the table, the value a whole number
// string (strtod must consume all of it).
void set_stop(RunOptions& o, const std::string& spec) {
    const size_t eq = spec.find('=');
    if (eq == std::string::npos) throw std::invalid_argument("--stop must be iters=N, sim-time=T, or wall-time=T");
    const std::string kind = spec.substr(0, eq), value = spec.substr(eq + 1);
    char* end = nullptr;
    const double v = std::strtod(value.c_str(), &end);
    if (value.empty() || *end != '\0') throw std::invalid_argument("Invalid --stop value: " + value);
    if (!lookup(stop_kinds(), kind, o.stop_kind)) throw std::invalid_argument("Unknown --stop kind: " + kind);
    o.stop_value = v;
}

using Setter = std::function<void(RunOptions&, const std::string&)>;

struct Flag {
    std::vector<std::string> names;  // spellings, the first one shown in the usage
    std::string metavar;             // empty: the flag takes no value
    std::string help;
    Setter set;                      // called with the value ("" for a switch)
};

Setter store(std::optional<std::string> RunOptions::*field) {
    return [field](RunOptions& o, const std::string& v) { o.*field = v; };
}

const std::vector<Flag>& flag_table() {
    static const std::vector<Flag> table = {
        {{"--app"}, "NAME", "Scene (default dam_break_32k), one of the names listed below",
         [](RunOptions& o, const std::string& v) { o.app = v; }},
        {{"--mesh"}, "FILE", "OBJ mesh of the dragon obstacle (scenes that use it)", store(&RunOptions::mesh)},
        {{"--headless"}, "", "No window (the only mode of this driver)",
         [](RunOptions& o, const std::string&) { o.headless = true; }},
        {{"--snapshot-load"}, "FILE", "Load snapshot after the scene is set up", store(&RunOptions::snapshot_load)},
        {{"--snapshot-save"}, "FILE", "Save snapshot at stop", store(&RunOptions::snapshot_save)},
        {{"--frozen-config"}, "FILE", "KTT JSON config; disables the autotuner", store(&RunOptions::frozen_config)},
        {{"--searcher"}, "{mcmc|random|full}", "Default: mcmc (every searcher runs the project autotuner once)",
         [](RunOptions& o, const std::string& v) {
             if (!lookup(searchers(), v, o.searcher)) throw std::invalid_argument("Unknown searcher: " + v);
         }},
        {{"--tuning-budget"}, "FLOAT", "Default: 0.1 (0 disables tuning)",
         [](RunOptions& o, const std::string& v) { o.tuning_budget = std::strtof(v.c_str(), nullptr); }},
        {{"--stop"}, "iters=N | sim-time=T | wall-time=T", "Required in --headless; sim-time counts the dt used", set_stop},
        {{"--fixed-dt"}, "FLOAT", "Default: 0.01 (the dt hint of every step)",
         [](RunOptions& o, const std::string& v) { o.fixed_dt = std::strtof(v.c_str(), nullptr); }},
        {{"--warmup-iters"}, "N", "Default: 0",
         [](RunOptions& o, const std::string& v) { o.warmup_iters = std::atoi(v.c_str()); }},
        {{"--ktt-output"}, "PATH", "Prefix for the tuning results (PATH.json)", store(&RunOptions::ktt_output)},
        {{"--log-csv"}, "FILE", "Per-step CSV log (the last column is the dt used)", store(&RunOptions::log_csv)},
        {{"--log-metrics"}, "", "Add mean_speed,ke columns",
         [](RunOptions& o, const std::string&) { o.log_metrics = true; }},
        {{"--print-options"}, "", "Print the parsed options and exit",
         [](RunOptions& o, const std::string&) { o.print_options = true; }},
        {{"--print-scene"}, "", "Print the scene's particle counts and exit",
         [](RunOptions& o, const std::string&) { o.print_scene = true; }},
        {{"--help", "-h"}, "", "This text", nullptr},
    };
    return table;
}

const Flag* find_flag(const std::string& arg) {
    for (const Flag& f : flag_table())
        for (const std::string& n : f.names)
            if (arg == n) return &f;
    return nullptr;
}

void print_usage(const char* prog) {
    std::ostringstream u;
    u << "Usage: " << prog << " [OPTIONS]\n";
    for (const Flag& f : flag_table()) {
        const std::string head = f.names[0] + (f.metavar.empty() ? "" : " " + f.metavar);
        u << "  " << std::left << std::setw(34) << head << ' ' << f.help << '\n';
    }
    u << "Scenes:\n";
    for (const auto& name : app_names()) u << "  " << name << '\n';
    std::cerr << u.str();
}

}  // namespace

const char* searcher_name(RunOptions::Searcher s) { return name_of(searchers(), s, "full"); }

const char* stop_kind_name(RunOptions::StopKind k) { return name_of(stop_kinds(), k, "none"); }

RunOptions parse_cli(int argc, char** argv) {
    RunOptions opts;
    const std::vector<std::string> args(argv + 1, argv + argc);
    for (size_t k = 0; k < args.size(); ++k) {
        const Flag* f = find_flag(args[k]);
        if (!f) throw std::invalid_argument("Unknown argument: " + args[k]);
        if (!f->set) {  // --help
            print_usage(argv[0]);
            std::exit(EXIT_SUCCESS);
        }
        std::string value;
        if (!f->metavar.empty()) {
            if (k + 1 >= args.size()) throw std::invalid_argument(args[k] + " requires an argument");
            value = args[++k];
        }
        f->set(opts, value);
    }
    if (opts.headless && opts.stop_kind == RunOptions::StopKind::None && !opts.print_options && !opts.print_scene)
        throw std::invalid_argument("--stop is required in --headless mode");
    return opts;
}

}  // namespace wcsph_host
