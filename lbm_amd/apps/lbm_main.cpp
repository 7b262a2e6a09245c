// Drop-in driver: the reference's time loop (src/main.cpp:30-73) over the
// GPU domain. Same config file / command line, same scenario XML and fluid
// mask input, same console lines, .vts output every timesteps-per-plot
// steps. The lattice is a run-time option here (`lattice`, default D3Q19)
// instead of the reference's compile-time #define.
//
//   lbm config.cfg [--timesteps N] [--storage f32] [--lattice D3Q27] ...
//   lbm --dry-run config.cfg   (parse and print the geometry, no device work)
#include <algorithm>
#include <chrono>
#include <cstdlib>
#include <cstring>
#include <iostream>
#include <string>
#include <vector>

#include "lbm_hip/io.hpp"
#include "lbm_hip/lbm.hpp"

namespace {

double now() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

template <class model>
int run(lbm::io::Config& cfg) {
    // Allocate collision operator
    auto collision = lbm::BGKCollision<model>(cfg.tau());
    lbm::DeviceOptions opt;
    opt.storage = cfg.storage_id();
    opt.layout = cfg.layout_id();
    auto domain = lbm::io::parse_scenario_file<model>(cfg.scenario_xml(), cfg, collision, opt);
    std::cout << cfg << std::endl;
    std::cout << "> Domain lengths (x,y,z): " << domain->xlength() << ", " << domain->ylength()
              << ", " << domain->zlength() << std::endl;
    std::cout << "Starting simulation..." << std::endl;

    domain->set_nonfluid_cells_nullcollide();
    double mlups_duration = 0.0;
    const double walltime = now();
    // The steps between two outputs (src/main.cpp:50-57: stream, swap, collide
    // per step, a VTK file every timesteps_per_plot steps) are enqueued as one
    // lbm_step call, which lets a bulk-only domain run them two per pass; the
    // fields written are the same bits as stepping one at a time.
    const uint64_t every = cfg.timesteps_per_plot();
    for (uint64_t t = 0; t < cfg.timesteps();) {
        uint64_t n = cfg.timesteps() - t;
        if (every) n = std::min<uint64_t>(n, every - t % every);
        const double start = now();
        domain->step(n);
        domain->sync();
        mlups_duration += now() - start;
        t += n;
        if (every && t % every == 0) lbm::io::write_vtk_file(*domain, cfg.output_dir(), cfg.output_filename(), t);
        std::cout << "\r" << static_cast<int>(static_cast<double>(t) / cfg.timesteps() * 100) << " %";
        std::cout.flush();
    }
    std::cout << "\nFinished!" << std::endl;
    std::cout << "Total runtime: " << now() - walltime << " seconds." << std::endl;
    // src/main.cpp:64-66 counts the ghost layer; interior cells are printed too
    const double ghosted = static_cast<double>(2 + domain->xlength()) * (2 + domain->ylength()) *
                           (2 + domain->zlength());
    const double interior = static_cast<double>(domain->xlength()) * domain->ylength() * domain->zlength();
    std::cout << "MLUPS: " << ghosted * cfg.timesteps() / (mlups_duration * 1e6) << std::endl;
    std::cout << "MLUPS (interior cells): " << interior * cfg.timesteps() / (mlups_duration * 1e6)
              << std::endl;
    return EXIT_SUCCESS;
}

template <class model>
int dry_run(lbm::io::Config& cfg) {
    // geometry only: parse the scenario into host-side boundary boxes
    std::ifstream in(cfg.scenario_xml());
    std::stringstream ss;
    ss << in.rdbuf();
    const auto doc = lbm::io::parse_xml(ss.str());
    const auto* sc = doc.child("scenario");
    if (!sc) throw std::logic_error("Scenario node missing!");
    const auto* d = sc->child("domain");
    if (!d) throw std::logic_error("Domain node is missing!");
    std::cout << "scenario " << sc->attr("name") << "\nlattice " << model::name << "\n";
    for (const auto& b : d->children)
        if (b.name == "boundary") std::cout << "boundary " << b.attr("extent") << " " << b.attr("condition") << "\n";
    return EXIT_SUCCESS;
}

}  // namespace

int main(int argc, char** argv) {
    try {
        bool dry = false;
        std::vector<char*> args;
        for (int i = 0; i < argc; ++i) {
            if (std::strcmp(argv[i], "--dry-run") == 0) dry = true;
            else args.push_back(argv[i]);
        }
        lbm::io::Config cfg(static_cast<int>(args.size()), args.data(), !dry);
        if (cfg.help) return 1;
        const std::string lat = cfg.lattice();
        if (dry) {
            if (lat == "D3Q15") return dry_run<lbm::model::d3q15>(cfg);
            if (lat == "D3Q27") return dry_run<lbm::model::d3q27>(cfg);
            return dry_run<lbm::model::d3q19>(cfg);
        }
        if (lat == "D3Q15") return run<lbm::model::d3q15>(cfg);
        if (lat == "D3Q19") return run<lbm::model::d3q19>(cfg);
        if (lat == "D3Q27") return run<lbm::model::d3q27>(cfg);
        throw std::logic_error("lattice must be D3Q15, D3Q19 or D3Q27");
    } catch (const std::exception& ex) {
        std::cerr << "An error occured: " << ex.what() << std::endl;
        return EXIT_FAILURE;
    }
}
