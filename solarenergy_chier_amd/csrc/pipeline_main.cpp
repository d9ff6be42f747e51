// qmc_pipeline — command-line drop-in for the reference's RayTracingPipeline executable (main.cpp;
// RayTracingPipeline/RayTracingPipeline.cpp:14-141, 145-175), on the host C++ facade of
// include/qmcrt_facade.hpp (RayTracingPipeline::run).
//
//   qmc_pipeline -c <configuration.json> -s <scene.scn> -h <heliostat_index.txt> -o <output_dir/>
//   (long forms --configuration_path, --scene_path, --heliostat_index_path, --output_path)
//
// Argument rules of Tool/ArgumentParser/ArgumentParser.cpp:22-94: each file must end with its suffix
// and open; the output directory must end with '/' (and, here, exist). An unknown option prints
// "Unknown arguments." and stops (exit 2); invalid paths print the reference's message (exit 1).
// The pipeline loads the configuration and the scene, processes it (seeded sample pools and group
// offsets, fresh offsets per trace: runs are reproducible), loads the task file, traces every listed
// heliostat in one launch, then for every receiver of the task writes
// <output><r>_receiver_before_smooth.txt, smooths the map (radius 5, ratio 0.05) and writes
// <output><r>_receiver_after_smooth.txt.
#include <cstdio>
#include <exception>

#include "qmcrt_facade.hpp"

int main(int argc, char **argv) {
    try {
        return RayTracingPipeline::run(argc, argv);
    } catch (const std::exception &e) {
        std::fflush(stdout);
        std::fprintf(stderr, "\n%s\n", e.what());
        return 1;
    }
}
