// Command line of the headless driver: the reference's flag set
// (src/cli.h, src/cli.cu:7-101) plus the scene choice, which the reference
// makes by building one binary per app.
#pragma once

#include <optional>
#include <stdexcept>
#include <string>

namespace wcsph_host {

struct RunOptions {
    bool headless = false;
    std::optional<std::string> snapshot_load;
    std::optional<std::string> snapshot_save;
    std::optional<std::string> frozen_config;

    enum class Searcher { Mcmc, Random, Full } searcher = Searcher::Mcmc;
    float tuning_budget = 0.1f;

    enum class StopKind { None, Iters, SimTime, WallTime } stop_kind = StopKind::None;
    double stop_value = 0.0;

    float fixed_dt = 0.01f;
    int warmup_iters = 0;

    std::optional<std::string> ktt_output;
    std::optional<std::string> log_csv;
    bool log_metrics = false;

    // additions of this driver
    std::string app = "dam_break_32k";
    std::optional<std::string> mesh;  // OBJ file for the scenes' dragon obstacle
    bool print_options = false;       // print the parsed options and exit
    bool print_scene = false;         // build the scene on the host, print its sizes and exit
};

// Throws std::invalid_argument with the reference's messages; --help
// prints the usage and exits 0.
RunOptions parse_cli(int argc, char** argv);

const char* searcher_name(RunOptions::Searcher s);
const char* stop_kind_name(RunOptions::StopKind k);

}  // namespace wcsph_host
