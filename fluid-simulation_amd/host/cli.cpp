#include "cli.h"

#include <cstdlib>
#include <cstring>
#include <iostream>

#include "apps.h"

namespace wcsph_host {

namespace {

void print_usage(const char* prog) {
    std::cerr << "Usage: " << prog << " [OPTIONS]\n"
              << "  --app NAME                      Scene (default dam_break_32k); one of:\n";
    for (const auto& name : app_names()) std::cerr << "                                    " << name << "\n";
    std::cerr << "  --mesh FILE                     OBJ mesh of the dragon obstacle (scenes that use it)\n"
                 "  --headless                      No window (the only mode of this driver)\n"
                 "  --snapshot-load FILE            Load snapshot after the scene is set up\n"
                 "  --snapshot-save FILE            Save snapshot at stop\n"
                 "  --frozen-config FILE            KTT JSON config; disables the autotuner\n"
                 "  --searcher {mcmc|random|full}   Default: mcmc (all run the project autotuner once)\n"
                 "  --tuning-budget FLOAT           Default: 0.1 (0 disables tuning)\n"
                 "  --stop iters=N | sim-time=T | wall-time=T   Required in --headless\n"
                 "  --fixed-dt FLOAT                Default: 0.01\n"
                 "  --warmup-iters N                Default: 0\n"
                 "  --ktt-output PATH               Prefix for the tuning results (PATH.json)\n"
                 "  --log-csv FILE                  Per-step CSV log\n"
                 "  --log-metrics                   Add mean_speed,ke columns\n"
                 "  --print-options                 Print the parsed options and exit\n"
                 "  --print-scene                   Print the scene's particle counts and exit\n";
}

RunOptions::Searcher parse_searcher(const char* s) {
    if (std::strcmp(s, "mcmc") == 0) return RunOptions::Searcher::Mcmc;
    if (std::strcmp(s, "random") == 0) return RunOptions::Searcher::Random;
    if (std::strcmp(s, "full") == 0) return RunOptions::Searcher::Full;
    throw std::invalid_argument(std::string("Unknown searcher: ") + s);
}

void parse_stop(const char* s, RunOptions::StopKind& kind, double& value) {
    const char* eq = std::strchr(s, '=');
    if (!eq) throw std::invalid_argument("--stop must be iters=N, sim-time=T, or wall-time=T");
    const std::string key(s, eq - s);
    const char* val = eq + 1;
    char* end = nullptr;
    value = std::strtod(val, &end);
    if (end == val || *end != '\0') throw std::invalid_argument(std::string("Invalid --stop value: ") + val);
    if (key == "iters")
        kind = RunOptions::StopKind::Iters;
    else if (key == "sim-time")
        kind = RunOptions::StopKind::SimTime;
    else if (key == "wall-time")
        kind = RunOptions::StopKind::WallTime;
    else
        throw std::invalid_argument("Unknown --stop kind: " + key);
}

}  // namespace

const char* searcher_name(RunOptions::Searcher s) {
    switch (s) {
    case RunOptions::Searcher::Mcmc: return "mcmc";
    case RunOptions::Searcher::Random: return "random";
    default: return "full";
    }
}

const char* stop_kind_name(RunOptions::StopKind k) {
    switch (k) {
    case RunOptions::StopKind::Iters: return "iters";
    case RunOptions::StopKind::SimTime: return "sim-time";
    case RunOptions::StopKind::WallTime: return "wall-time";
    default: return "none";
    }
}

RunOptions parse_cli(int argc, char** argv) {
    RunOptions opts;
    auto need = [&](int& i, const char* flag) -> const char* {
        if (++i >= argc) throw std::invalid_argument(std::string(flag) + " requires an argument");
        return argv[i];
    };
    for (int i = 1; i < argc; ++i) {
        const char* arg = argv[i];
        if (std::strcmp(arg, "--headless") == 0) {
            opts.headless = true;
        } else if (std::strcmp(arg, "--app") == 0) {
            opts.app = need(i, arg);
        } else if (std::strcmp(arg, "--mesh") == 0) {
            opts.mesh = need(i, arg);
        } else if (std::strcmp(arg, "--snapshot-load") == 0) {
            opts.snapshot_load = need(i, arg);
        } else if (std::strcmp(arg, "--snapshot-save") == 0) {
            opts.snapshot_save = need(i, arg);
        } else if (std::strcmp(arg, "--frozen-config") == 0) {
            opts.frozen_config = need(i, arg);
        } else if (std::strcmp(arg, "--searcher") == 0) {
            opts.searcher = parse_searcher(need(i, arg));
        } else if (std::strcmp(arg, "--tuning-budget") == 0) {
            opts.tuning_budget = std::strtof(need(i, arg), nullptr);
        } else if (std::strcmp(arg, "--stop") == 0) {
            parse_stop(need(i, arg), opts.stop_kind, opts.stop_value);
        } else if (std::strcmp(arg, "--fixed-dt") == 0) {
            opts.fixed_dt = std::strtof(need(i, arg), nullptr);
        } else if (std::strcmp(arg, "--warmup-iters") == 0) {
            opts.warmup_iters = std::atoi(need(i, arg));
        } else if (std::strcmp(arg, "--ktt-output") == 0) {
            opts.ktt_output = need(i, arg);
        } else if (std::strcmp(arg, "--log-csv") == 0) {
            opts.log_csv = need(i, arg);
        } else if (std::strcmp(arg, "--log-metrics") == 0) {
            opts.log_metrics = true;
        } else if (std::strcmp(arg, "--print-options") == 0) {
            opts.print_options = true;
        } else if (std::strcmp(arg, "--print-scene") == 0) {
            opts.print_scene = true;
        } else if (std::strcmp(arg, "--help") == 0 || std::strcmp(arg, "-h") == 0) {
            print_usage(argv[0]);
            std::exit(EXIT_SUCCESS);
        } else {
            throw std::invalid_argument(std::string("Unknown argument: ") + arg);
        }
    }
    if (opts.headless && opts.stop_kind == RunOptions::StopKind::None && !opts.print_options && !opts.print_scene)
        throw std::invalid_argument("--stop is required in --headless mode");
    return opts;
}

}  // namespace wcsph_host
