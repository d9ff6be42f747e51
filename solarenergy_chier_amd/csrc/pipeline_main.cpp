// qmc_pipeline — command-line drop-in for the reference's RayTracingPipeline executable (main.cpp;
// RayTracingPipeline/RayTracingPipeline.cpp:14-141, 145-175) on the C ABI of include/qmcrt.h.
//
//   qmc_pipeline -c <configuration.json> -s <scene.scn> -h <heliostat_index.txt> -o <output_dir/>
//   (long forms --configuration_path, --scene_path, --heliostat_index_path, --output_path)
//
// Argument rules of Tool/ArgumentParser/ArgumentParser.cpp:22-94: each file must end with its suffix
// and open; the output directory must end with '/' (and, here, exist). An unknown option prints
// "Unknown arguments." and stops (exit 2); invalid paths print the reference's message (exit 1).
// The pipeline loads the configuration and the scene, processes it, loads the task file, traces every
// listed heliostat in one launch (seeded sample pools and group offsets: runs are reproducible), then
// for every receiver of the task writes <output><r>_receiver_before_smooth.txt, smooths the map
// (radius 5, ratio 0.05) and writes <output><r>_receiver_after_smooth.txt.
#include <getopt.h>
#include <sys/stat.h>

#include <chrono>
#include <cstdio>
#include <fstream>
#include <set>
#include <stdexcept>
#include <string>
#include <vector>

#include "qmcrt.h"

namespace {

struct Arguments {
    std::string configuration_path, scene_path, heliostat_index_path, output_path;
};

void check_valid_file(const std::string &path, const std::string &suffix) {
    if (path.size() < suffix.size() || path.compare(path.size() - suffix.size(), suffix.size(), suffix) != 0)
        throw std::runtime_error("The path of file ' " + path + " ' does not consist of ' " + suffix + " ' suffix.");
    std::ifstream f(path.c_str());
    if (!f.good()) throw std::runtime_error("The file couldn't be opened. Please check your file path.");
}

void check_valid_directory(const std::string &path) {
    if (path.empty() || path.back() != '/')
        throw std::runtime_error("' " + path + " ' directory format is incorrect. Please add '/' at the end of the path.");
    struct stat info {};
    if (stat(path.c_str(), &info) != 0 || !S_ISDIR(info.st_mode))
        throw std::runtime_error("' " + path + " ' directory does not exist. Please check the path.");
}

// 0: arguments valid; 2: unknown option. Invalid paths throw.
int parse_arguments(int argc, char **argv, Arguments &a) {
    static const struct option long_options[] = {{"configuration_path", required_argument, nullptr, 'c'},
                                                 {"scene_path", required_argument, nullptr, 's'},
                                                 {"output_path", required_argument, nullptr, 'o'},
                                                 {"heliostat_index_path", required_argument, nullptr, 'h'},
                                                 {nullptr, 0, nullptr, 0}};
    optind = 1;
    int c, index = 0;
    while ((c = getopt_long(argc, argv, "c:s:o:h:", long_options, &index)) != -1) {
        switch (c) {
        case 'c': a.configuration_path = optarg; std::printf("\noption -c/configuration_path with value '%s'", optarg); break;
        case 's': a.scene_path = optarg; std::printf("\noption -s/scene_path with value '%s'", optarg); break;
        case 'o': a.output_path = optarg; std::printf("\noption -o/output_path with value '%s'", optarg); break;
        case 'h': a.heliostat_index_path = optarg; std::printf("\noption -h/heliostat_index_path with value '%s'", optarg); break;
        default: std::printf("\nUnknown arguments.\n"); return 2;
        }
    }
    check_valid_file(a.configuration_path, ".json");
    check_valid_file(a.scene_path, ".scn");
    check_valid_file(a.heliostat_index_path, ".txt");
    check_valid_directory(a.output_path);
    return 0;
}

void check(qmc_status st, const char *what) {
    if (st != QMC_OK) throw std::runtime_error(std::string(what) + ": " + qmc_status_string(st) + ": " + qmc_last_error());
}

float save_receiver(qmc_ctx *ctx, const qmc_scene *scene, int r, const std::string &path_and_name) {
    qmc_receiver_info info;
    check(qmc_scene_get_receiver_info(scene, r, &info), "receiver info");
    const int h = info.res_y, w = info.res_x;
    std::printf("\treceiver resolution: (%d, %d).\n", h, w);
    std::vector<float> img((size_t)h * w);
    qmc_image_stats st;
    check(qmc_flux_read(ctx, r, img.data(), (int64_t)img.size()), "flux read");
    check(qmc_image_save_text((path_and_name + "_receiver_before_smooth.txt").c_str(), h, w, img.data(), 2, &st),
          "save text");
    std::printf("\tMax value is %g. Sum value is %g.\n\tMax value is located at (%d, %d)\n", st.max_value, st.sum,
                st.max_row, st.max_col);
    const auto t0 = std::chrono::high_resolution_clock::now();
    check(qmc_smooth_receiver(ctx, r, 5, 0.05f), "smooth");
    check(qmc_sync(ctx), "sync");
    const auto t1 = std::chrono::high_resolution_clock::now();
    std::printf("\t[Smooth time]:  %lld microseconds.\n",
                (long long)std::chrono::duration_cast<std::chrono::microseconds>(t1 - t0).count());
    check(qmc_flux_read(ctx, r, img.data(), (int64_t)img.size()), "flux read");
    check(qmc_image_save_text((path_and_name + "_receiver_after_smooth.txt").c_str(), h, w, img.data(), 2, &st),
          "save text");
    std::printf("\tMax value is %g. Sum value is %g.\n\tMax value is located at (%d, %d)\n", st.max_value, st.sum,
                st.max_row, st.max_col);
    return st.max_value;
}

int run(int argc, char **argv) {
    std::printf("1. Start loading arguments...");
    Arguments a;
    const int parsed = parse_arguments(argc, argv, a);
    if (parsed != 0) return parsed;

    std::printf("\n\n2. Initialize solar scene...\n");
    qmc_scene *scene = nullptr;
    check(qmc_scene_create(&scene), "scene");
    std::printf("\t2.1 Load configuration from '%s'.\n", a.configuration_path.c_str());
    int32_t fields = 0;
    check(qmc_scene_load_config(scene, a.configuration_path.c_str(), &fields), "configuration");
    std::printf("\t2.2 Load scene file from '%s'.\n", a.scene_path.c_str());
    check(qmc_scene_load_file(scene, a.scene_path.c_str()), "scene file");
    std::printf("\t2.3 Process scene.\n");
    check(qmc_scene_process(scene), "process");
    std::printf("\t2.4 Load heliostats indexes from '%s'.\n", a.heliostat_index_path.c_str());
    qmc_scene_counts counts;
    check(qmc_scene_get_counts(scene, &counts), "counts");
    std::vector<int32_t> ids((size_t)counts.n_heliostats + 1);
    int32_t n = 0, clean = 0;
    check(qmc_task_load(scene, a.heliostat_index_path.c_str(), ids.data(), (int32_t)ids.size(), &n, &clean), "task file");
    ids.resize((size_t)n);
    if (!clean) std::printf("\tthe task file repeats or exceeds its declared heliostat indexes\n");
    std::set<int32_t> receivers;  // ascending receiver order
    for (int32_t h : ids) {
        check(qmc_scene_check_heliostat(scene, h), "heliostat index");
        qmc_heliostat_info hi;
        qmc_grid_desc gd;
        check(qmc_scene_get_heliostat_info(scene, h, &hi), "heliostat info");
        check(qmc_scene_get_grid_desc(scene, hi.grid, &gd), "grid");
        receivers.insert(gd.receiver);
    }

    std::printf("\n3. Start ray tracing...\n");
    qmc_ctx *ctx = nullptr;
    check(qmc_create(0, &ctx), "device");
    check(qmc_upload(ctx, scene), "upload");
    qmc_pool_params pools{2, 4};
    check(qmc_pool_generate(ctx, &pools), "sample pools");
    check(qmc_offsets_generate(ctx, 3), "group offsets");
    check(qmc_event_record(ctx, 0), "event");
    check(qmc_trace(ctx, ids.data(), n), "trace");
    check(qmc_event_record(ctx, 1), "event");
    check(qmc_sync(ctx), "sync");
    float ms = 0.0f;
    check(qmc_event_elapsed_ms(ctx, 0, 1, &ms), "event");
    std::printf("\t%d heliostats took %.3f milliseconds to ray tracing.\n", n, ms);

    std::printf("\n4. Save results in '%s' directory.\n", a.output_path.c_str());
    for (int32_t r : receivers) {
        std::printf("------Saving No.%d receiver------\n", r);
        save_receiver(ctx, scene, r, a.output_path + std::to_string(r));
        check(qmc_flux_clear(ctx, r), "clear");
    }
    check(qmc_destroy(ctx), "device");
    qmc_scene_destroy(scene);
    return 0;
}

}  // namespace

int main(int argc, char **argv) {
    try {
        return run(argc, argv);
    } catch (const std::exception &e) {
        std::fflush(stdout);
        std::fprintf(stderr, "\n%s\n", e.what());
        return 1;
    }
}
