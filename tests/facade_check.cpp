// The reference's host call pattern, compiled against include/qmcrt_facade.hpp and run by
// tests/test_facade.py.
//
//   facade_check host   <fixtures>   loader, configuration, task loader, host processing and index
//                                    checks (no GPU)
//   facade_check device <fixtures>   QuasiMonteCarloRayTracerTest.cpp:19-47,189-208 on the
//                                    qmcrt_test_scene fixture: load, process, rayTracing(solarScene, 0),
//                                    read the receiver's device image; prints the image sums
//
// Prints "check <name> ok" per check and "all ok" at the end; exits non-zero at the first failure.
#include <unistd.h>

#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

#include "qmcrt_facade.hpp"

namespace {

int failures = 0;

void expect(bool ok, const std::string &name) {
    std::printf("check %s %s\n", name.c_str(), ok ? "ok" : "FAILED");
    std::fflush(stdout);
    if (!ok) ++failures;
}

template <class F>
bool throws_with(F f, const std::string &needle) {
    try {
        f();
    } catch (const std::runtime_error &e) {
        const bool found = std::string(e.what()).find(needle) != std::string::npos;
        if (!found) std::printf("  message was: %s\n", e.what());
        return found;
    }
    return false;
}

void host_checks(const std::string &fix) {
    SceneConfiguration *cfg = SceneConfiguration::getInstance();
    // Tests/DataStructureTests/SceneConfigurationTest.cpp:32-68: 10 fields, then 8, then a missing file
    expect(cfg->loadConfiguration(fix + "/ds_test_configuration.json") == cfg->num_of_fields, "config_full");
    expect(cfg->getDni() == 500.0f && cfg->getNum_sunshape_groups() == 32 && cfg->getSun_dir().y == -1.0f,
           "config_values");
    expect(cfg->loadConfiguration(fix + "/ds_test_partial_configuration.json") == 8, "config_partial");
    expect(throws_with([&] { cfg->loadConfiguration(fix + "/nonexist_file.json"); }, "doesn't exist"),
           "config_missing_throws");

    SolarScene *scene = SolarScene::GetInstance();
    SceneLoader loader;
    // malformed scenes return false, the scene is left empty
    expect(!loader.SceneFileRead(scene, fix + "/ds_test_scene_bad_nonRE.scn"), "scene_bad_returns_false");
    expect(scene->getHeliostats().empty() && !scene->getLoaded_from_file(), "scene_bad_empty");
    expect(!loader.SceneFileRead(scene, fix + "/no_such_scene.scn"), "scene_missing_returns_false");

    cfg->loadConfiguration(fix + "/qmcrt_test_configuration.json");
    expect(loader.SceneFileRead(scene, fix + "/qmcrt_test_scene.scn"), "scene_good");
    expect(scene->getLoaded_from_file() && scene->getReceivers().size() == 1 && scene->getGrids().size() == 1 &&
               scene->getHeliostats().size() == 3 && scene->getGridNum() == 1 && scene->getGroundLength() == 1024.0f,
           "scene_counts");
    Heliostat *h1 = scene->getHeliostats()[1];
    expect(h1->getRowAndColumn().x == 3 && h1->getRowAndColumn().y == 2 && h1->getGap().x == 1.0f &&
               h1->getPosition().z == 12.0f && h1->getBelongingGridId() == 0,
           "heliostat_view");
    Grid *g = scene->getGrids()[0];
    expect(g->getStartHeliostatIndex() == 0 && g->getNumberOfHeliostats() == 3 && g->getBelongingReceiverIndex() == 0 &&
               g->getInterval().x == 10.0f,
           "grid_view");
    Receiver *r = scene->getReceivers()[0];
    expect(r->getType() == 0 && r->getSize().x == 16.0f && r->getDeviceImage() == nullptr, "receiver_view");

    // QuasiMonteCarloRayTracer.cpp:96-127
    QuasiMonteCarloRayTracer tracer;
    expect(throws_with([&] { tracer.checkValidHeliostatIndex(scene, 3); },
                       "The ray tracing heliostat index 3 is invalid.The heliostat index should between 0 and 3."),
           "index_above_throws");
    expect(throws_with([&] { tracer.checkValidHeliostatIndex(scene, -1); }, "index -1 is invalid"),
           "index_negative_throws");
    tracer.checkValidHeliostatIndex(scene, 2);
    expect(tracer.receiverGridCombination(1, 0) == 10, "receiver_grid_combination");
    expect(throws_with([&] { tracer.rayTracing(scene, 0); }, "process the scene first"), "trace_before_process_throws");

    SceneProcessor none;
    expect(throws_with([&] { none.processScene(scene); }, "No scene configuration"), "process_without_config_throws");
    SceneProcessor processor(cfg);
    expect(processor.set_grid_receiver_heliostat_content(scene), "host_processing");
    // 16 x 10 m at int(1/0.5) = 2 pixels per metre; heliostat 0's aim normal is (0, 0, -1)
    // (QuasiMonteCarloRayTracerTest.cpp:134-135)
    expect(scene->getReceivers()[0]->getResolution().x == 32 && scene->getReceivers()[0]->getResolution().y == 20,
           "receiver_resolution");
    expect(scene->getHeliostats()[0]->getNormal().z < -0.99f, "aim_normal");

    // RayTracingPipeline/TaskLoader.cpp:114-204: N then N distinct indices; a repeat returns false
    const std::string tasks_ok = "/tmp/facade_check_tasks_" + std::to_string((long)getpid()) + ".txt";
    const std::string tasks_rep = tasks_ok + ".rep.txt";
    std::FILE *f = std::fopen(tasks_ok.c_str(), "w");
    std::fprintf(f, "2\n0 2\n");
    std::fclose(f);
    f = std::fopen(tasks_rep.c_str(), "w");
    std::fprintf(f, "3\n2 2 1\n");
    std::fclose(f);
    TaskLoader tasks;
    expect(tasks.loadRayTracingHeliostatIndex(tasks_ok, *scene) && tasks.getHeliostatIndexesArray().size() == 2 &&
               tasks.getHeliostatIndexesArray().count(2) == 1 && tasks.getReceiverIndexesArray().size() == 1,
           "task_loader");
    expect(!tasks.loadRayTracingHeliostatIndex(tasks_rep, *scene), "task_loader_repeat_false");
    std::remove(tasks_ok.c_str());
    std::remove(tasks_rep.c_str());
    scene->clear();
    expect(scene->getHeliostats().empty() && scene->qmcContext() == nullptr, "clear");
}

double image_sum(Receiver *r) {
    const int2 res = r->getResolution();
    std::vector<float> h((size_t)res.x * res.y);
    if (hipMemcpy(h.data(), r->getDeviceImage(), sizeof(float) * h.size(), hipMemcpyDeviceToHost) != hipSuccess)
        throw std::runtime_error("hipMemcpy");
    double s = 0.0;
    for (float v : h) s += v;
    return s;
}

void device_checks(const std::string &fix) {
    // QMCRTracerTestFixture::SetUp (QuasiMonteCarloRayTracerTest.cpp:19-47), without the test-only
    // heliostat subclass
    SolarScene *solarScene = SolarScene::GetInstance();
    SceneConfiguration *sceneConfiguration = SceneConfiguration::getInstance();
    sceneConfiguration->loadConfiguration(fix + "/qmcrt_test_configuration.json");
    SceneLoader sceneLoader;
    expect(sceneLoader.SceneFileRead(solarScene, fix + "/qmcrt_test_scene.scn"), "load");
    SceneProcessor sceneProcessor(sceneConfiguration);
    expect(sceneProcessor.processScene(solarScene), "processScene");

    // TEST_F(QMCRTracerTestFixture, rayTracing) (QuasiMonteCarloRayTracerTest.cpp:189-208)
    QuasiMonteCarloRayTracer QMCRTracer;
    QMCRTracer.rayTracing(solarScene, 0);
    Receiver *receiver = solarScene->getReceivers()[0];
    const int2 resolution = receiver->getResolution();
    expect(resolution.x == 32 && resolution.y == 20 && receiver->getDeviceImage() != nullptr, "device_image");
    const double s1 = image_sum(receiver);
    std::printf("sum after rayTracing(0): %.9e\n", s1);
    expect(s1 > 0.0, "deposits");
    // rayTracing accumulates, with fresh offsets for the second call
    QMCRTracer.rayTracing(solarScene, 0);
    const double s2 = image_sum(receiver);
    std::printf("sum after a second rayTracing(0): %.9e\n", s2);
    expect(s2 > 1.9 * s1 && s2 < 2.1 * s1, "accumulates");
    expect(throws_with([&] { QMCRTracer.rayTracing(solarScene, 5); }, "index 5 is invalid"), "invalid_index_throws");
    QMCRTracer.rayTracing(solarScene, std::vector<int>{1, 2});
    const double s3 = image_sum(receiver);
    std::printf("sum after rayTracing({1, 2}): %.9e\n", s3);
    expect(s3 > s2, "task_list");
    // the pipeline's smoothing of the device image in place
    ImageSmoother::image_smooth(receiver->getDeviceImage(), 5, 0.05f, resolution.x, resolution.y);
    std::printf("sum after smoothing: %.9e\n", image_sum(receiver));
    receiver->Cclean_image_content();
    expect(image_sum(receiver) == 0.0, "clean_image");
    // TearDown
    solarScene->clear();
}

}  // namespace

int main(int argc, char **argv) {
    if (argc != 3) {
        std::fprintf(stderr, "usage: facade_check host|device <fixtures dir>\n");
        return 2;
    }
    const std::string mode = argv[1], fix = argv[2];
    try {
        if (mode == "host") host_checks(fix);
        else if (mode == "device") device_checks(fix);
        else return 2;
    } catch (const std::exception &e) {
        std::printf("exception: %s\n", e.what());
        return 1;
    }
    if (failures) return 1;
    std::printf("all ok\n");
    return 0;
}
