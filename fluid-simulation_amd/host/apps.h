// The scenes the headless driver can run: the reference's apps
// (app/*/application.cu: setup_scene + update_objects) under the
// reference physics, and the dam-break / obstacle-collapse benchmark
// workloads under the benchmark physics.
#pragma once

#include <functional>
#include <optional>
#include <string>
#include <vector>

#include "fluid_simulator.h"

namespace wcsph_host {

struct AppScene {
    std::string name;  // also the app_name of its snapshots
    wcsph_params params;
    FluidSimulator::opts_t opts;
    // Per-frame object update (TiltingBoxApp::update_objects moves the box);
    // empty for static scenes.
    std::function<void(SPHSimulator&, float delta)> update_objects;
};

std::vector<std::string> app_names();

// Builds the scene description on the host.  mesh: OBJ file of the dragon
// for the scenes that place one (the reference loads
// models/stanford-dragon.obj, which it does not ship); without it those
// scenes run without the dragon and say so on stderr.
AppScene make_app(const std::string& name, const std::optional<std::string>& mesh);

}  // namespace wcsph_host
