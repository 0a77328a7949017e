// wcsph_headless: the reference's CLI-driven headless stepping
// (Application::init / run_headless / stop_reached, src/application.cu:24-57,
// 105-161) over the C++ simulator host.  One process, one scene, one GPU.
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <exception>
#include <fstream>
#include <iostream>
#include <memory>
#include <optional>
#include <string>
#include <utility>

#include "apps.h"
#include "cli.h"
#include "fluid_simulator.h"

using namespace wcsph_host;

namespace {

// Host-only stand-in used by --print-scene: builds the particles, never steps.
class SceneOnly final : public FluidSimulator {
public:
    using FluidSimulator::FluidSimulator;
    void update(float) override {}
};

void print_options(const RunOptions& o) {
    auto opt = [](const std::optional<std::string>& s) { return s ? "\"" + *s + "\"" : std::string("null"); };
    std::printf("{\"app\": \"%s\", \"headless\": %s, \"stop\": \"%s\", \"stop_value\": %.17g, \"fixed_dt\": %.17g, "
                "\"warmup_iters\": %d, \"searcher\": \"%s\", \"tuning_budget\": %.17g, \"snapshot_load\": %s, "
                "\"snapshot_save\": %s, \"frozen_config\": %s, \"ktt_output\": %s, \"log_csv\": %s, "
                "\"log_metrics\": %s, \"mesh\": %s}\n",
                o.app.c_str(), o.headless ? "true" : "false", stop_kind_name(o.stop_kind), o.stop_value,
                (double)o.fixed_dt, o.warmup_iters, searcher_name(o.searcher), (double)o.tuning_budget,
                opt(o.snapshot_load).c_str(), opt(o.snapshot_save).c_str(), opt(o.frozen_config).c_str(),
                opt(o.ktt_output).c_str(), opt(o.log_csv).c_str(), o.log_metrics ? "true" : "false",
                opt(o.mesh).c_str());
}

bool stop_reached(const RunOptions& o, long step, double sim_time, std::chrono::steady_clock::time_point t0) {
    switch (o.stop_kind) {
    case RunOptions::StopKind::Iters: return step >= (long)o.stop_value;
    case RunOptions::StopKind::SimTime: return sim_time >= o.stop_value;
    case RunOptions::StopKind::WallTime:
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count() >= o.stop_value;
    default: return false;
    }
}

void write_tuning(const std::string& prefix, const std::vector<std::pair<wcsph_variant, double>>& res) {
    std::ofstream f(prefix + ".json");
    if (!f) throw std::runtime_error("cannot write " + prefix + ".json");
    f << "[\n";
    for (size_t k = 0; k < res.size(); ++k) {
        const wcsph_variant& v = res[k].first;
        f << "  {\"block_size\": " << v.block_size << ", \"fluid_capacity\": " << v.fluid_capacity
          << ", \"boundary_capacity\": " << v.boundary_capacity << ", \"use_graph\": " << v.use_graph
          << ", \"neighbor_mode\": " << v.neighbor_mode << ", \"neighbor_groups\": " << v.neighbor_groups
          << ", \"cell_size_mult\": " << v.cell_size_mult << ", \"ms_per_step\": " << res[k].second << "}" << (k + 1 < res.size() ? "," : "") << "\n";
    }
    f << "]\n";
}

// Where a headless run stands: steps taken, physical time since the run
// began, and when it began (the wall-time stop rule's origin).
struct Progress {
    using clock = std::chrono::steady_clock;
    long steps = 0;
    double sim_time = 0.0;
    clock::time_point started = clock::now();
};

// One step's row of the headless CSV: the reference's columns
// (application.cu:110-114: step, sim_time, wall_dt_ms, scheduled_tune and,
// with --log-metrics, mean_speed and ke) and the dt the step used.
struct StepRow {
    long step;
    double sim_time, wall_ms;
    bool scheduled_tune;
    std::optional<std::pair<double, double>> metrics;  // mean speed, kinetic energy
    double dt;
};

class StepLog {
public:
    StepLog(const std::optional<std::string>& path, bool metrics) : metrics_(metrics) {
        if (!path) return;
        out_.open(*path);
        if (!out_) throw std::runtime_error("cannot write " + *path);
        out_.precision(9);
        out_ << "step,sim_time,wall_dt_ms,scheduled_tune" << (metrics_ ? ",mean_speed,ke" : "") << ",dt\n";
    }
    explicit operator bool() const { return out_.is_open(); }
    void write(const StepRow& r) {
        out_ << r.step << ',' << r.sim_time << ',' << r.wall_ms << ',' << (r.scheduled_tune ? 1 : 0);
        if (metrics_ && r.metrics) out_ << ',' << r.metrics->first << ',' << r.metrics->second;
        out_ << ',' << r.dt << '\n';
    }

private:
    std::ofstream out_;
    bool metrics_;
};

int run(const RunOptions& opts) {
    if (opts.print_options) {
        print_options(opts);
        if (opts.frozen_config) {
            const wcsph_variant base{256, 64, 32, 1, WCSPH_NEIGHBORS_PARTICLE, 16384, 1.f};
            const wcsph_variant fv = variant_from_ktt_config(*opts.frozen_config, base);
            std::printf("{\"frozen_block_size\": %u, \"frozen_cell_size_mult\": %.9g, \"frozen_neighbor_mode\": %u}\n",
                        fv.block_size, (double)fv.cell_size_mult, fv.neighbor_mode);
        }
        return EXIT_SUCCESS;
    }
    AppScene app = make_app(opts.app, opts.mesh);
    if (opts.print_scene) {
        SceneOnly s(app.opts, app.params.particle_radius);
        std::printf("{\"app\": \"%s\", \"fluid\": %u, \"boundary\": %u}\n", app.name.c_str(), s.get_fluid_particles(),
                    s.get_boundary_particles());
        return EXIT_SUCCESS;
    }
    if (!opts.headless) throw std::invalid_argument("this driver has no window: pass --headless");
    const unsigned n_fluid = app.opts.initial_fluid.empty()
                                 ? app.opts.grid_dims.x * app.opts.grid_dims.y * app.opts.grid_dims.z
                                 : (unsigned)(app.opts.initial_fluid.size() / 4);
    // a snapshot whose header does not fit the scene fails before any device work
    // (the checks and messages of Snapshot::load, snapshot.cu:51-80)
    if (opts.snapshot_load) {
        char name[64];
        uint32_t n = 0;
        if (wcsph_snapshot_info(opts.snapshot_load->c_str(), name, sizeof(name), &n) != WCSPH_OK)
            throw std::runtime_error(std::string("Snapshot load failed: ") + wcsph_snapshot_error());
        if (app.name != name)
            throw std::runtime_error("Snapshot load failed: snapshot: app_name mismatch (file='" + std::string(name) +
                                     "', expected='" + app.name + "')");
        if (n != n_fluid)
            throw std::runtime_error("Snapshot load failed: snapshot: fluid count mismatch (file=" +
                                     std::to_string(n) + ", expected=" + std::to_string(n_fluid) + ")");
    }

    // a frozen config is read (and its unmapped parameters named) before any
    // device work, so a bad file fails early
    if (opts.frozen_config) {
        std::vector<std::string> ignored;
        variant_from_ktt_config(*opts.frozen_config, wcsph_variant{}, &ignored);
        if (!ignored.empty()) {
            std::fprintf(stderr, "frozen config: %zu KTT parameters have no counterpart here:\n", ignored.size());
            for (const auto& s2 : ignored) std::fprintf(stderr, "  %s\n", s2.c_str());
        }
    }

    // Application::init (application.cu:24-57)
    SPHSimulator sim(app.opts, app.params);
    auto update_objects = [&](float delta) {
        if (app.update_objects) app.update_objects(sim, delta);
        sim.update(delta);
    };
    if (opts.snapshot_load) {
        const std::string err = sim.load_snapshot(*opts.snapshot_load, app.name);
        if (!err.empty()) throw std::runtime_error("Snapshot load failed: " + err);
    }
    for (int i = 0; i < opts.warmup_iters; ++i) update_objects(opts.fixed_dt);
    if (opts.frozen_config) {
        sim.set_frozen_config(variant_from_ktt_config(*opts.frozen_config, sim.variant()));
    } else if (opts.tuning_budget > 0.f) {
        // one-shot project autotuner in place of KTT's per-frame search;
        // every searcher name runs the same exhaustive pass over its space
        const auto res = sim.autotune(opts.fixed_dt, 5);
        if (opts.ktt_output) write_tuning(*opts.ktt_output, res);
    }

    // Application::run_headless (application.cu:105-143): step until the stop
    // rule holds, one CSV row per step.  The reference advances its sim_time
    // by --fixed-dt whatever dt the solver took (application.cu:124-125); here
    // the run's physical time is the sum of the dt actually used (CFL-limited
    // in adaptive mode), so --stop sim-time=T stops at physical time T, and
    // the row's last column is that step's dt.
    const double clock_at_start = sim.last_stats().sim_time;  // after warm-up (tuning restores its trials)
    StepLog csv(opts.log_csv, opts.log_metrics);
    Progress done;
    while (!stop_reached(opts, done.steps, done.sim_time, done.started)) {
        const auto before = Progress::clock::now();
        update_objects(opts.fixed_dt);
        const wcsph_stats& st = sim.last_stats();
        done.sim_time = st.sim_time - clock_at_start;
        if (csv) {
            StepRow row{done.steps, done.sim_time,
                        std::chrono::duration<double, std::milli>(Progress::clock::now() - before).count(),
                        sim.was_scheduled_step(), {}, st.last_dt};
            if (opts.log_metrics) row.metrics = sim.compute_state_metrics();
            csv.write(row);
        }
        ++done.steps;
    }
    if (opts.snapshot_save) {
        const std::string err = sim.save_snapshot(*opts.snapshot_save, app.name);
        if (!err.empty()) throw std::runtime_error("Snapshot save failed: " + err);
    }
    std::printf("{\"app\": \"%s\", \"steps\": %ld, \"sim_time\": %.9g, \"wall_s\": %.6f, \"fluid\": %u}\n",
                app.name.c_str(), done.steps, done.sim_time,
                std::chrono::duration<double>(Progress::clock::now() - done.started).count(),
                sim.get_fluid_particles());
    return EXIT_SUCCESS;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        return run(parse_cli(argc, argv));
    } catch (const std::exception& e) {
        std::cerr << "Exception thrown: " << e.what() << std::endl;
        return EXIT_FAILURE;
    }
}
