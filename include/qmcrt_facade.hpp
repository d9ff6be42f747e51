/*
 * qmcrt_facade.hpp — the reference's host C++ call surface, over the C ABI of qmcrt.h.
 *
 * Header-only; include it and link libqmcrt.so (and the HIP runtime, for the float3 / int2 vector
 * types of the reference's signatures and for device pointers). The classes keep the names, methods
 * and error behaviour of JachyLikeCoding/SolarEnergy_Chier's host code, so that its callers
 * (RayTracingPipeline.cpp:14-141, Tests/QuasiMonteCarloRayTracingTests/QuasiMonteCarloRayTracerTest.cpp)
 * compile and run against them:
 *
 *   SceneConfiguration   Scene/SceneConfiguration.h:12-57        singleton; loadConfiguration returns the
 *                                                                 number of fields set, throws on a
 *                                                                 missing file
 *   SolarScene           Scene/SolarScene.h:17-62                 singleton; receivers, grids, heliostats,
 *                                                                 sunray; owns the device context
 *   Receiver / Grid / Heliostat / Sunray   Scene/DataStructure/   read-only views of the scene's objects
 *   SceneLoader          Scene/SceneLoader.h:14-29                SceneFileRead returns false on a
 *                                                                 malformed scene
 *   SceneProcessor       Scene/SceneProcessor.h:11-33             processScene: geometry on the host, then
 *                                                                 the device copy, sample pools, offsets
 *   QuasiMonteCarloRayTracer  QuasiMonteCarloRayTracing/QuasiMonteCarloRayTracer.h:11-27
 *                                                                 rayTracing accumulates into the
 *                                                                 receiver's device image; invalid indices
 *                                                                 throw std::runtime_error with the
 *                                                                 reference's messages
 *   ImageSmoother        Tool/Smoother/ImageSmoother.cuh:18-22    in place on a device image
 *   ImageSaver           Tool/ImageSaver/ImageSaver.h:10-14       text format of the reference
 *   TaskLoader           RayTracingPipeline/TaskLoader.h:14-27
 *   RayTracingPipeline   RayTracingPipeline/RayTracingPipeline.h:11-19
 *
 * Where this build differs: the sample pools and group offsets are seeded (SceneProcessor::seeds)
 * instead of drawn from time(NULL), and each rayTracing call draws fresh offsets for its heliostat
 * from those seeds, as the reference does from cuRAND; scene objects are views of the loaded scene
 * (no user-defined Heliostat subclasses); the pipeline traces its task list in one launch.
 */
#ifndef QMCRT_FACADE_HPP
#define QMCRT_FACADE_HPP

#include <getopt.h>
#include <hip/hip_runtime.h>
#include <sys/stat.h>

#include <chrono>
#include <cstdio>
#include <fstream>
#include <iostream>
#include <memory>
#include <set>
#include <stdexcept>
#include <string>
#include <unordered_set>
#include <vector>

#include "qmcrt.h"

namespace qmcrt_facade {

inline void check(qmc_status st, const std::string &what) {
    if (st != QMC_OK) throw std::runtime_error(what + ": " + qmc_status_string(st) + ": " + qmc_last_error());
}

inline float3 f3(const float p[3]) { return make_float3(p[0], p[1], p[2]); }

// Seeds of the sample pools and group offsets (the reference seeds cuRAND with time(NULL)).
struct Seeds {
    uint64_t sunshape = 2, perturbation = 4, offsets = 3, resample = 1000;
};

// The process's default device context, for ImageSmoother calls on a device image that no scene owns.
inline qmc_ctx *default_ctx() {
    static std::unique_ptr<qmc_ctx, qmc_status (*)(qmc_ctx *)> ctx(nullptr, qmc_destroy);
    if (!ctx) {
        qmc_ctx *c = nullptr;
        check(qmc_create(0, &c), "device context");
        ctx.reset(c);
    }
    return ctx.get();
}

}  // namespace qmcrt_facade

class SolarScene;

// Scene/SceneConfiguration.h:12-57.
class SceneConfiguration {
public:
    static SceneConfiguration *getInstance() {
        static SceneConfiguration instance;
        return &instance;
    }

    // Scene/SceneConfiguration.cpp:39-69: fields absent from the file keep their values; returns the
    // number of the 10 fields the file sets. A missing file throws; so does malformed json.
    int loadConfiguration(std::string configuration_file_path) {
        std::ifstream in(configuration_file_path);
        if (in.fail()) {
            const std::string msg = "The file '" + configuration_file_path + "' doesn't exist";
            std::cerr << msg << std::endl;
            throw std::runtime_error(msg);
        }
        qmc_scene *s = nullptr;
        qmcrt_facade::check(qmc_scene_create(&s), "SceneConfiguration");
        qmc_status st = qmc_scene_set_config(s, &cfg_);
        int32_t n = 0;
        if (st == QMC_OK) st = qmc_scene_load_config(s, configuration_file_path.c_str(), &n);
        if (st == QMC_OK) st = qmc_scene_get_config(s, &cfg_);
        const std::string err = st == QMC_OK ? "" : qmc_last_error();
        qmc_scene_destroy(s);
        if (st != QMC_OK) throw std::runtime_error("SceneConfiguration::loadConfiguration: " + err);
        sun_dir_ = qmcrt_facade::f3(cfg_.sun_dir);
        if (n != num_of_fields) std::printf("Noted: not all fields are set. You may need to check your configuration file.\n");
        return n;
    }

    const float3 &getSun_dir() const { return sun_dir_; }
    float getDni() const { return cfg_.dni; }
    float getCsr() const { return cfg_.csr; }
    int getNum_sunshape_groups() const { return cfg_.num_sunshape_groups; }
    int getNum_sunshape_lights_per_group() const { return cfg_.num_sunshape_lights_per_group; }
    int getInverse_transform_sampling_groups() const { return cfg_.inverse_transform_sampling_groups; }
    float getReceiver_pixel_length() const { return cfg_.receiver_pixel_length; }
    float getDisturb_std() const { return cfg_.helio_disturb_std; }
    float getHelio_pixel_length() const { return cfg_.helio_pixel_length; }
    float getReflected_rate() const { return cfg_.helio_reflected_rate; }

    const int num_of_fields = 10;

    // The whole configuration as the C ABI's descriptor.
    const qmc_config &config() const { return cfg_; }
    void setConfig(const qmc_config &c) {
        cfg_ = c;
        sun_dir_ = qmcrt_facade::f3(cfg_.sun_dir);
    }

private:
    SceneConfiguration() {
        qmc_config_default(&cfg_);  // the reference's defaults (Scene/SceneConfiguration.h:37,41-51)
        sun_dir_ = qmcrt_facade::f3(cfg_.sun_dir);
    }
    qmc_config cfg_{};
    float3 sun_dir_{};
};

// View of receiver `index` of a SolarScene (Scene/DataStructure/Receiver/Receiver.cuh:11-85).
class Receiver {
public:
    int getType() const { return desc().type; }
    float3 getNormal() const { return qmcrt_facade::f3(desc().normal); }
    float3 getPosition() const { return qmcrt_facade::f3(desc().pos); }
    float3 getSize() const { return qmcrt_facade::f3(desc().size); }
    int getSurfaceIndex() const { return desc().face; }
    // After processing (RectangleReceiver.cu:98-101, CylinderReceiver.cu:12-15).
    float getPixelLength() const { return info().pixel_length; }
    int2 getResolution() const {
        const qmc_receiver_info i = info();
        return make_int2(i.res_x, i.res_y);
    }
    // The flux map on the device (res.x * res.y floats, row-major), after SceneProcessor::processScene.
    float *getDeviceImage() const;
    // Receiver::Cclean_image_content (Receiver.cu:33-45).
    void Cclean_image_content();

private:
    friend class SolarScene;
    Receiver(SolarScene *scene, int index) : scene_(scene), index_(index) {}
    qmc_receiver_desc desc() const;
    qmc_receiver_info info() const;
    SolarScene *scene_;
    int index_;
};

// View of grid `index` (Scene/DataStructure/Grid/Grid.cuh:12-60).
class Grid {
public:
    int getGridType() const { return desc().type; }
    float3 getPosition() const { return qmcrt_facade::f3(desc().pos); }
    float3 getSize() const { return qmcrt_facade::f3(desc().size); }
    float3 getInterval() const { return qmcrt_facade::f3(desc().interval); }
    int getHeliostatType() const { return desc().heliostat_type; }
    int getStartHeliostatIndex() const { return desc().start_heliostat; }
    int getNumberOfHeliostats() const { return desc().n_heliostats; }
    int getBelongingReceiverIndex() const { return desc().receiver; }

private:
    friend class SolarScene;
    Grid(SolarScene *scene, int index) : scene_(scene), index_(index) {}
    qmc_grid_desc desc() const;
    SolarScene *scene_;
    int index_;
};

// View of heliostat `index` (Scene/DataStructure/Heliostat/Heliostat.cuh:17-100).
class Heliostat {
public:
    float3 getPosition() const { return qmcrt_facade::f3(desc().pos); }
    float3 getSize() const { return qmcrt_facade::f3(desc().size); }
    // After processing: the aim normal (Heliostat.cu:15-20).
    float3 getNormal() const { return qmcrt_facade::f3(info().normal); }
    int2 getRowAndColumn() const {
        const qmc_heliostat_desc d = desc();
        return make_int2(d.rows, d.cols);
    }
    float2 getGap() const {
        const qmc_heliostat_desc d = desc();
        return make_float2(d.gap[0], d.gap[1]);
    }
    int getBelongingGridId() const { return desc().grid; }
    int getSubHelioSize() const { return info().sub_count; }
    std::vector<float> getSurfaceProperty() const {
        const qmc_heliostat_desc d = desc();
        return std::vector<float>(d.surface_property, d.surface_property + 6);
    }
    std::vector<float3> getCPULocalCenters() const { return local(desc().local_centers); }
    std::vector<float3> getCPULocalNormals() const { return local(desc().local_normals); }

private:
    friend class SolarScene;
    Heliostat(SolarScene *scene, int index) : scene_(scene), index_(index) {}
    qmc_heliostat_desc desc() const;
    qmc_heliostat_info info() const;
    std::vector<float3> local(const float *p) const {
        std::vector<float3> out;
        const qmc_heliostat_desc d = desc();
        if (p)
            for (int k = 0; k < d.rows * d.cols; ++k) out.push_back(make_float3(p[3 * k], p[3 * k + 1], p[3 * k + 2]));
        return out;
    }
    SolarScene *scene_;
    int index_;
};

// The sun of a scene (Scene/DataStructure/Sunray/Sunray.cuh): the configuration's values once the
// scene is processed.
class Sunray {
public:
    float3 getSunDirection() const { return qmcrt_facade::f3(cfg().sun_dir); }
    float getDNI() const { return cfg().dni; }
    float getCSR() const { return cfg().csr; }
    int getNumOfSunshapeGroups() const { return cfg().num_sunshape_groups; }
    int getNumOfSunshapeLightsPerGroup() const { return cfg().num_sunshape_lights_per_group; }
    float getReflectiveRate() const { return cfg().helio_reflected_rate; }

private:
    friend class SolarScene;
    explicit Sunray(SolarScene *scene) : scene_(scene) {}
    qmc_config cfg() const;
    SolarScene *scene_;
};

// Scene/SolarScene.h:17-62.
class SolarScene {
public:
    static SolarScene *GetInstance() {
        static SolarScene instance;
        return &instance;
    }
    ~SolarScene() { clear(); }

    // SolarScene::clear (Scene/SolarScene.cpp:31-46): drops every object and the device context.
    bool clear() {
        drop_views();
        if (ctx_) qmc_destroy(ctx_);
        ctx_ = nullptr;
        if (scene_) qmc_scene_destroy(scene_);
        scene_ = nullptr;
        ground_length_ = ground_width_ = 0.0f;
        grid_num_ = 0;
        loaded_from_file_ = false;
        return true;
    }

    float getGroundLength() const { return ground_length_; }
    void setGroundLength(float v) { ground_length_ = v; }
    float getGroundWidth() const { return ground_width_; }
    void setGroundWidth(float v) { ground_width_ = v; }
    int getGridNum() const { return grid_num_; }
    void setGridNum(int v) { grid_num_ = v; }
    bool getLoaded_from_file() const { return loaded_from_file_; }
    void SetLoaded_from_file(bool v) { loaded_from_file_ = v; }

    // Components as the C ABI's descriptors (the reference adds Receiver* / Grid* / Heliostat* objects).
    void addReceiver(const qmc_receiver_desc &d) {
        qmcrt_facade::check(qmc_scene_add_receiver(qmc(), &d), "SolarScene::addReceiver");
        refresh();
    }
    void addGrid(const qmc_grid_desc &d) {
        qmcrt_facade::check(qmc_scene_add_grid(qmc(), &d), "SolarScene::addGrid");
        refresh();
    }
    void addHeliostat(const qmc_heliostat_desc &d) {
        qmcrt_facade::check(qmc_scene_add_heliostat(qmc(), &d), "SolarScene::addHeliostat");
        refresh();
    }

    Sunray *getSunray() { return sunray_.get(); }
    std::vector<Grid *> &getGrids() { return grids_; }
    std::vector<Receiver *> &getReceivers() { return receivers_; }
    std::vector<Heliostat *> &getHeliostats() { return heliostats_; }

    // The host scene behind the objects (created on first use) and the device context that
    // SceneProcessor::processScene creates on `device()`.
    qmc_scene *qmc() {
        if (!scene_) qmcrt_facade::check(qmc_scene_create(&scene_), "SolarScene");
        return scene_;
    }
    const qmc_scene *qmc() const { return scene_; }
    qmc_ctx *qmcContext() const { return ctx_; }
    int device() const { return device_; }
    void setDevice(int device) { device_ = device; }

private:
    friend class SceneLoader;
    friend class SceneProcessor;
    SolarScene() : sunray_(new Sunray(this)) {}
    SolarScene(const SolarScene &) = delete;
    SolarScene &operator=(const SolarScene &) = delete;

    void drop_views() {
        receivers_.clear();
        grids_.clear();
        heliostats_.clear();
        own_receivers_.clear();
        own_grids_.clear();
        own_heliostats_.clear();
    }
    // Rebuilds the object views after the scene's content changed.
    void refresh() {
        drop_views();
        if (!scene_) return;
        qmc_scene_counts c{};
        qmcrt_facade::check(qmc_scene_get_counts(scene_, &c), "SolarScene");
        ground_length_ = c.ground_length;
        ground_width_ = c.ground_width;
        grid_num_ = c.grid_num_declared;
        for (int i = 0; i < c.n_receivers; ++i) {
            own_receivers_.emplace_back(new Receiver(this, i));
            receivers_.push_back(own_receivers_.back().get());
        }
        for (int i = 0; i < c.n_grids; ++i) {
            own_grids_.emplace_back(new Grid(this, i));
            grids_.push_back(own_grids_.back().get());
        }
        for (int i = 0; i < c.n_heliostats; ++i) {
            own_heliostats_.emplace_back(new Heliostat(this, i));
            heliostats_.push_back(own_heliostats_.back().get());
        }
    }
    qmc_ctx *ensure_ctx() {
        if (!ctx_) qmcrt_facade::check(qmc_create(device_, &ctx_), "device context");
        return ctx_;
    }

    qmc_scene *scene_ = nullptr;
    qmc_ctx *ctx_ = nullptr;
    int device_ = 0;
    float ground_length_ = 0.0f, ground_width_ = 0.0f;
    int grid_num_ = 0;
    bool loaded_from_file_ = false;
    std::unique_ptr<Sunray> sunray_;
    std::vector<std::unique_ptr<Receiver>> own_receivers_;
    std::vector<std::unique_ptr<Grid>> own_grids_;
    std::vector<std::unique_ptr<Heliostat>> own_heliostats_;
    std::vector<Receiver *> receivers_;
    std::vector<Grid *> grids_;
    std::vector<Heliostat *> heliostats_;
};

inline qmc_receiver_desc Receiver::desc() const {
    qmc_receiver_desc d{};
    qmcrt_facade::check(qmc_scene_get_receiver_desc(scene_->qmc(), index_, &d), "Receiver");
    return d;
}
inline qmc_receiver_info Receiver::info() const {
    qmc_receiver_info i{};
    if (qmc_scene_get_receiver_info(scene_->qmc(), index_, &i) != QMC_OK) i = qmc_receiver_info{};
    return i;
}
inline float *Receiver::getDeviceImage() const {
    qmc_ctx *c = scene_->qmcContext();
    if (!c) return nullptr;
    float *p = nullptr;
    qmcrt_facade::check(qmc_flux_device_ptr(c, index_, &p, nullptr, nullptr), "Receiver::getDeviceImage");
    return p;
}
inline void Receiver::Cclean_image_content() {
    qmc_ctx *c = scene_->qmcContext();
    if (!c) return;
    qmcrt_facade::check(qmc_flux_clear(c, index_), "Receiver::Cclean_image_content");
    qmcrt_facade::check(qmc_sync(c), "Receiver::Cclean_image_content");
}
inline qmc_grid_desc Grid::desc() const {
    qmc_grid_desc d{};
    qmcrt_facade::check(qmc_scene_get_grid_desc(scene_->qmc(), index_, &d), "Grid");
    return d;
}
inline qmc_heliostat_desc Heliostat::desc() const {
    qmc_heliostat_desc d{};
    qmcrt_facade::check(qmc_scene_get_heliostat_desc(scene_->qmc(), index_, &d), "Heliostat");
    return d;
}
inline qmc_heliostat_info Heliostat::info() const {
    qmc_heliostat_info i{};
    if (qmc_scene_get_heliostat_info(scene_->qmc(), index_, &i) != QMC_OK) i = qmc_heliostat_info{};
    return i;
}
inline qmc_config Sunray::cfg() const {
    qmc_config c{};
    if (!scene_->qmc() || qmc_scene_get_config(scene_->qmc(), &c) != QMC_OK) qmc_config_default(&c);
    return c;
}

// Scene/SceneLoader.h:14-29.
class SceneLoader {
public:
    // Scene/SceneLoader.cpp:243-353: replaces the scene's content with the file's; false (with the
    // reason on stderr) when the file cannot be read or breaks the format, the scene then empty.
    bool SceneFileRead(SolarScene *solarScene, std::string filepath) {
        if (!solarScene) return false;
        if (solarScene->ctx_) {
            qmc_destroy(solarScene->ctx_);  // the device copy belongs to the previous content
            solarScene->ctx_ = nullptr;
        }
        const qmc_status st = qmc_scene_load_file(solarScene->qmc(), filepath.c_str());
        solarScene->refresh();
        if (st != QMC_OK) {
            std::cerr << "Error caused by ' " << qmc_last_error() << " '." << std::endl;
            solarScene->SetLoaded_from_file(false);
            return false;
        }
        solarScene->SetLoaded_from_file(true);
        return true;
    }
};

// Scene/SceneProcessor.h:11-33.
class SceneProcessor {
public:
    SceneProcessor() : sceneConfiguration(nullptr) {}
    explicit SceneProcessor(SceneConfiguration *sceneConfigure) : sceneConfiguration(sceneConfigure) {}

    // Scene/SceneProcessor.cpp:9-11: the geometry (receivers, grid cell lists, aim normals, world
    // vertices, micro-facet layout) on the host, then the device copy, the sample pools and the group
    // offsets, with per-call resampling of the offsets (each rayTracing call draws fresh ones).
    bool processScene(SolarScene *solarScene) {
        return set_grid_receiver_heliostat_content(solarScene) && set_sunray_content(solarScene);
    }

    // SceneProcessor::set_grid_receiver_heliostat_content (SceneProcessor.cpp:13-39), on the host.
    bool set_grid_receiver_heliostat_content(SolarScene *solarScene) {
        require(solarScene);
        qmcrt_facade::check(qmc_scene_set_config(solarScene->qmc(), &sceneConfiguration->config()),
                            "SceneProcessor");
        qmcrt_facade::check(qmc_scene_process(solarScene->qmc()), "SceneProcessor::processScene");
        solarScene->refresh();
        return true;
    }

    // SceneProcessor::set_sunray_content (SceneProcessor.cpp:41-54, SceneProcessor.cu:147-271): the
    // device copy of the processed scene, the sun-shape and slope-error pools and the group offsets.
    bool set_sunray_content(SolarScene *solarScene) {
        require(solarScene);
        qmc_ctx *c = solarScene->ensure_ctx();
        qmcrt_facade::check(qmc_upload(c, solarScene->qmc()), "SceneProcessor: device copy of the scene");
        qmc_pool_params p{};
        p.seed_sunshape = seeds.sunshape;
        p.seed_perturbation = seeds.perturbation;
        p.sunshape_model = QMC_SUNSHAPE_REFERENCE;
        qmcrt_facade::check(qmc_pool_generate(c, &p), "SceneProcessor: sample pools");
        qmcrt_facade::check(qmc_offsets_generate(c, seeds.offsets), "SceneProcessor: group offsets");
        qmcrt_facade::check(qmc_offsets_resample(c, 1, seeds.resample), "SceneProcessor: group offsets");
        qmcrt_facade::check(qmc_sync(c), "SceneProcessor");
        return true;
    }

    SceneConfiguration *getSceneConfiguration() const { return sceneConfiguration; }
    void setSceneConfiguration(SceneConfiguration *c) { sceneConfiguration = c; }

    qmcrt_facade::Seeds seeds;

private:
    void require(SolarScene *solarScene) const {
        if (!sceneConfiguration)
            throw std::runtime_error("No scene configuration. Please load it first before process scene.");
        if (!solarScene) throw std::runtime_error("SceneProcessor: null scene");
    }
    SceneConfiguration *sceneConfiguration;
};

// QuasiMonteCarloRayTracing/QuasiMonteCarloRayTracer.h:11-27.
class QuasiMonteCarloRayTracer {
public:
    // QuasiMonteCarloRayTracer::rayTracing (QuasiMonteCarloRayTracer.cpp:15-91): traces one heliostat
    // into its receiver's device image, adding to what is there, with fresh group offsets.
    void rayTracing(SolarScene *solarScene, int heliostat_id) {
        checkValidHeliostatIndex(solarScene, heliostat_id);
        const int32_t id = heliostat_id;
        trace(solarScene, &id, 1);
    }
    // Every listed heliostat in one launch (each as one rayTracing call).
    void rayTracing(SolarScene *solarScene, const std::vector<int> &heliostat_ids) {
        for (int h : heliostat_ids) checkValidHeliostatIndex(solarScene, h);
        std::vector<int32_t> ids(heliostat_ids.begin(), heliostat_ids.end());
        trace(solarScene, ids.data(), (int32_t)ids.size());
    }

    // QuasiMonteCarloRayTracer.cpp:96-127: std::runtime_error with the reference's message when the
    // heliostat, its grid or the grid's receiver index is out of range.
    void checkValidHeliostatIndex(SolarScene *solarScene, int heliostat_id) {
        if (!solarScene) throw std::runtime_error("QuasiMonteCarloRayTracer: null scene");
        const qmc_status st = qmc_scene_check_heliostat(solarScene->qmc(), heliostat_id);
        if (st == QMC_E_RANGE) throw std::runtime_error(qmc_last_error());
        qmcrt_facade::check(st, "QuasiMonteCarloRayTracer::checkValidHeliostatIndex");
    }

    // QuasiMonteCarloRayTracer.cpp:130-132.
    int receiverGridCombination(int receiver_type, int grid_type) { return receiver_type * 10 + grid_type; }

private:
    static void trace(SolarScene *solarScene, const int32_t *ids, int32_t n) {
        qmc_ctx *c = solarScene->qmcContext();
        if (!c) throw std::runtime_error("QuasiMonteCarloRayTracer: process the scene first (SceneProcessor::processScene)");
        qmcrt_facade::check(qmc_trace(c, ids, n), "QuasiMonteCarloRayTracer::rayTracing");
        qmcrt_facade::check(qmc_sync(c), "QuasiMonteCarloRayTracer::rayTracing");
    }
};

// Tool/Smoother/ImageSmoother.cuh:18-22.
class ImageSmoother {
public:
    // ImageSmoother::image_smooth (ImageSmoother.cu:68-113): trimmed mean in place on a device image
    // of width x height floats, on the scene's device context when it has one.
    static void image_smooth(float *d_array, int kernel_radius, float trimmed_ratio, int width, int height) {
        qmc_ctx *c = SolarScene::GetInstance()->qmcContext();
        if (!c) c = qmcrt_facade::default_ctx();
        qmcrt_facade::check(qmc_smooth(c, d_array, width, height, kernel_radius, trimmed_ratio),
                            "ImageSmoother::image_smooth");
        qmcrt_facade::check(qmc_sync(c), "ImageSmoother::image_smooth");
    }
};

// Tool/ImageSaver/ImageSaver.h:10-14.
class ImageSaver {
public:
    // ImageSaver::saveText (ImageSaver.cpp:13-57): rows bottom-up, values below 1e-5 as 0, the others
    // with `precision` decimals; returns the largest value written.
    static float saveText(std::string filename, int height, int width, float *h_data, int precision = 2,
                          int rows_package = 10) {
        (void)rows_package;  // the reference's write batching; the file is the same
        qmc_image_stats st{};
        qmcrt_facade::check(qmc_image_save_text(filename.c_str(), height, width, h_data, precision, &st),
                            "ImageSaver::saveText");
        std::printf("\tMax value is %g. Sum value is %g.\n\tMax value is located at (%d, %d)\n", st.max_value, st.sum,
                    st.max_row, st.max_col);
        return st.max_value;
    }
};

// RayTracingPipeline/TaskLoader.h:14-27.
class TaskLoader {
public:
    // TaskLoader::loadRayTracingHeliostatIndex (TaskLoader.cpp:114-204): false when the file repeats or
    // exceeds its declared indices (the distinct ones are kept); malformed or short files throw.
    bool loadRayTracingHeliostatIndex(std::string task_path, SolarScene &solarScene) {
        qmc_scene_counts c{};
        qmcrt_facade::check(qmc_scene_get_counts(solarScene.qmc(), &c), "TaskLoader");
        std::vector<int32_t> ids((size_t)c.n_heliostats + 1);
        int32_t n = 0, clean = 0;
        qmcrt_facade::check(qmc_task_load(solarScene.qmc(), task_path.c_str(), ids.data(), (int32_t)ids.size(), &n, &clean),
                            "TaskLoader::loadRayTracingHeliostatIndex");
        heliostat_indexes.clear();
        receiver_indexes.clear();
        ordered_.assign(ids.begin(), ids.begin() + n);
        for (int32_t h : ordered_) {
            heliostat_indexes.insert(h);
            const int g = solarScene.getHeliostats()[(size_t)h]->getBelongingGridId();
            receiver_indexes.insert(solarScene.getGrids()[(size_t)g]->getBelongingReceiverIndex());
        }
        return clean != 0;
    }
    const std::unordered_set<int> &getHeliostatIndexesArray() const { return heliostat_indexes; }
    const std::unordered_set<int> &getReceiverIndexesArray() const { return receiver_indexes; }
    // The indices in file order.
    const std::vector<int> &getHeliostatIndexesInFileOrder() const { return ordered_; }

private:
    std::unordered_set<int> heliostat_indexes;
    std::unordered_set<int> receiver_indexes;
    std::vector<int> ordered_;
};

// RayTracingPipeline/RayTracingPipeline.h:11-19.
class RayTracingPipeline {
public:
    // RayTracingPipeline::rayTracing (RayTracingPipeline.cpp:14-141). Throws std::runtime_error on an
    // invalid path (the messages of Tool/ArgumentParser/ArgumentParser.cpp:21-47) and on unknown
    // options.
    static void rayTracing(int argc, char *argv[]) {
        if (run(argc, argv) != 0) throw std::runtime_error("Unknown arguments.");
    }

    // The same, returning an exit status: 0 done, 2 unknown option.
    static int run(int argc, char *argv[]) {
        std::printf("1. Start loading arguments...");
        std::string configuration_path, scene_path, heliostat_index_path, output_path;
        static const struct option long_options[] = {{"configuration_path", required_argument, nullptr, 'c'},
                                                     {"scene_path", required_argument, nullptr, 's'},
                                                     {"output_path", required_argument, nullptr, 'o'},
                                                     {"heliostat_index_path", required_argument, nullptr, 'h'},
                                                     {nullptr, 0, nullptr, 0}};
        optind = 1;
        int c, index = 0;
        while ((c = getopt_long(argc, argv, "c:s:o:h:", long_options, &index)) != -1) {
            switch (c) {
            case 'c': configuration_path = optarg; std::printf("\noption -c/configuration_path with value '%s'", optarg); break;
            case 's': scene_path = optarg; std::printf("\noption -s/scene_path with value '%s'", optarg); break;
            case 'o': output_path = optarg; std::printf("\noption -o/output_path with value '%s'", optarg); break;
            case 'h': heliostat_index_path = optarg; std::printf("\noption -h/heliostat_index_path with value '%s'", optarg); break;
            default: std::printf("\nUnknown arguments.\n"); return 2;
            }
        }
        check_valid_file(configuration_path, ".json");
        check_valid_file(scene_path, ".scn");
        check_valid_file(heliostat_index_path, ".txt");
        check_valid_directory(output_path);

        std::printf("\n\n2. Initialize solar scene...\n");
        std::printf("\t2.1 Load configuration from '%s'.\n", configuration_path.c_str());
        SceneConfiguration *sceneConfiguration = SceneConfiguration::getInstance();
        sceneConfiguration->loadConfiguration(configuration_path);
        std::printf("\t2.2 Load scene file from '%s'.\n", scene_path.c_str());
        SceneLoader sceneLoader;
        SolarScene *solarScene = SolarScene::GetInstance();
        if (!sceneLoader.SceneFileRead(solarScene, scene_path))
            throw std::runtime_error("The scene file '" + scene_path + "' could not be loaded.");
        std::printf("\t2.3 Process scene.\n");
        SceneProcessor sceneProcessor(sceneConfiguration);
        sceneProcessor.processScene(solarScene);
        std::printf("\t2.4 Load heliostats indexes from '%s'.\n", heliostat_index_path.c_str());
        TaskLoader taskLoader;
        if (!taskLoader.loadRayTracingHeliostatIndex(heliostat_index_path, *solarScene))
            std::printf("\tthe task file repeats or exceeds its declared heliostat indexes\n");

        std::printf("\n3. Start ray tracing...\n");
        QuasiMonteCarloRayTracer QMCRTracer;
        const auto t0 = std::chrono::high_resolution_clock::now();
        QMCRTracer.rayTracing(solarScene, taskLoader.getHeliostatIndexesInFileOrder());
        const auto t1 = std::chrono::high_resolution_clock::now();
        std::printf("\t%zu heliostats took %lld microseconds to ray tracing.\n",
                    taskLoader.getHeliostatIndexesInFileOrder().size(),
                    (long long)std::chrono::duration_cast<std::chrono::microseconds>(t1 - t0).count());

        std::printf("\n4. Save results in '%s' directory.\n", output_path.c_str());
        const std::set<int> receivers(taskLoader.getReceiverIndexesArray().begin(),
                                      taskLoader.getReceiverIndexesArray().end());
        for (int r : receivers) {
            std::printf("------Saving No.%d receiver------\n", r);
            saveReceiverResult(solarScene->getReceivers()[(size_t)r], output_path + std::to_string(r), r);
            solarScene->getReceivers()[(size_t)r]->Cclean_image_content();
        }
        solarScene->clear();
        return 0;
    }

private:
    // RayTracingPipeline::saveReceiverResult (RayTracingPipeline.cpp:145-175): the map before and after
    // trimmed-mean smoothing (radius 5, ratio 0.05), as text.
    static float saveReceiverResult(Receiver *receiver, std::string pathAndName, int receiverIndex) {
        (void)receiverIndex;
        const int2 resolution = receiver->getResolution();
        std::printf("\treceiver resolution: (%d, %d).\n", resolution.y, resolution.x);
        std::vector<float> h_array((size_t)resolution.x * resolution.y);
        float *d_array = receiver->getDeviceImage();
        qmcrt_facade::check(hipMemcpy(h_array.data(), d_array, sizeof(float) * h_array.size(), hipMemcpyDeviceToHost) ==
                                    hipSuccess ? QMC_OK : QMC_E_HIP,
                            "RayTracingPipeline: device image");
        float max_value = ImageSaver::saveText(pathAndName + "_receiver_before_smooth.txt", resolution.y, resolution.x,
                                               h_array.data());
        const auto t0 = std::chrono::high_resolution_clock::now();
        ImageSmoother::image_smooth(d_array, 5, 0.05f, resolution.x, resolution.y);
        const auto t1 = std::chrono::high_resolution_clock::now();
        std::printf("\t[Smooth time]:  %lld microseconds.\n",
                    (long long)std::chrono::duration_cast<std::chrono::microseconds>(t1 - t0).count());
        qmcrt_facade::check(hipMemcpy(h_array.data(), d_array, sizeof(float) * h_array.size(), hipMemcpyDeviceToHost) ==
                                    hipSuccess ? QMC_OK : QMC_E_HIP,
                            "RayTracingPipeline: device image");
        max_value = ImageSaver::saveText(pathAndName + "_receiver_after_smooth.txt", resolution.y, resolution.x,
                                         h_array.data());
        return max_value;
    }

    // Tool/ArgumentParser/ArgumentParser.cpp:21-47 (an output directory must also exist).
    static void check_valid_file(const std::string &path, const std::string &suffix) {
        if (path.size() < suffix.size() || path.compare(path.size() - suffix.size(), suffix.size(), suffix) != 0)
            throw std::runtime_error("The path of file ' " + path + " ' does not consist of ' " + suffix + " ' suffix.");
        std::ifstream f(path.c_str());
        if (!f.good()) throw std::runtime_error("The file couldn't be opened. Please check your file path.");
    }
    static void check_valid_directory(const std::string &path) {
        if (path.empty() || path.back() != '/')
            throw std::runtime_error("' " + path + " ' directory format is incorrect. Please add '/' at the end of the path.");
        struct stat info {};
        if (stat(path.c_str(), &info) != 0 || !S_ISDIR(info.st_mode))
            throw std::runtime_error("' " + path + " ' directory does not exist. Please check the path.");
    }
};

#endif  // QMCRT_FACADE_HPP
