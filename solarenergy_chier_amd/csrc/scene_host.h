// scene_host.h — host-side scene model behind the qmc_scene C ABI.
//
// Holds what the reference keeps in its SolarScene singleton (Scene/SolarScene.h:17-62) and
// what SceneProcessor derives from it (Scene/SceneProcessor.cpp:13-39): receiver vertices and
// resolution, grid cell lists, heliostat aim normals, world vertices, sub-heliostat quads and
// the micro-facet layout. The device upload (qmc_upload) packs these into SoA tables.
#pragma once

#include <string>
#include <vector>

#include "../../include/qmcrt.h"
#include "qmc_math.h"

namespace qmc {

struct HostReceiver {
    qmc_receiver_desc d{};
    // processed
    int res_x = 0, res_y = 0;
    float pixel_length = 0.0f;
    float area = 0.0f;          // pixel area used by the energy factor
    bool res_override = false;
    V3 local_normal{0, 0, 0};
    V3 normal{0, 0, 0};
    V3 vert[4]{};
    V3 focus_center{0, 0, 0};
};

struct HostGrid {
    qmc_grid_desc d{};
    int nx = 0, ny = 0, nz = 0;
    std::vector<int> cell_start;  // nx*ny*nz + 1
    std::vector<int> match;       // grid-relative sub-heliostat ids
    std::vector<V3> sub_verts;    // 3 per sub-heliostat (v0, v1, v2 order of the reference)
    // per listed sub-heliostat id: the cell range whose lists hold it (x0, x1, y0, y1, z0, z1; empty
    // when x0 > x1 or y0 > y1 or z0 > z1)
    std::vector<int> sub_box;
    int n_sub = 0;
};

struct HostHeliostat {
    qmc_heliostat_desc d{};
    std::vector<float> file_centers, file_normals;  // read method 1
    int type = 0;                 // 0 flat, 1 focus-flat (from the grid)
    // processed
    V3 normal{0, 0, 0};
    V3 aim{0, 0, 0};
    V3 vert[4]{};
    float focal = 0.0f;
    float pixel_length = 0.0f;
    int sub_rows = 0, sub_cols = 0;   // micro-facets per facet
    int n_micro = 0;                  // F
    int sub_count = 1;                // blocking quads
    int sub_begin = 0;                // grid-relative first sub id
    float len_row = 0.0f, len_col = 0.0f;  // facet extent along z (rows) and x (cols)
    std::vector<V3> loc_centers, loc_normals;  // focus-flat per facet (heliostat frame)
};

}  // namespace qmc

struct qmc_scene {
    qmc_config cfg{};
    bool cfg_loaded = false;
    float ground_length = 0.0f, ground_width = 0.0f;
    int grid_num = 0;
    std::vector<qmc::HostReceiver> receivers;
    std::vector<qmc::HostGrid> grids;
    std::vector<qmc::HostHeliostat> heliostats;
    bool processed = false;
};

namespace qmc {
void set_error(const std::string &msg);
qmc_status fail(qmc_status s, const std::string &msg);
// Energy factor of a heliostat (reference rule or resolution-override rule).
float heliostat_factor(const qmc_scene &s, int h);
// Factor rule and its scalar: rule 0 -> factor = DNI*x*x*rho/L, rule 1 -> DNI*x*rho/L.
void heliostat_factor_parts(const qmc_scene &s, int h, int &rule, float &x);
// Aim point of heliostat h of a processed scene for an aim offset (across, up) in the receiver.
V3 heliostat_aim(const qmc_scene &s, int h, const float off[2]);
// Aim points in double (3 per heliostat): for rectangle receivers focus + o0*across + o1*up with the
// float focus, unit edge directions and offsets promoted before the arithmetic (the differentiable
// model's parameterisation); for cylinder receivers the float aim point.
void aim_points_f64(const qmc_scene &s, std::vector<double> &out);
void sunshape_params(const qmc_config &c, std::vector<float> &cdf_xy, std::vector<float> &slopes, float abcig[5]);
// Table of the physical Buie sampler (bin edges in mrad, cumulative probability, end densities).
struct BuieTable {
    std::vector<double> theta, cdf, dens_lo, dens_hi;
};
void buie_physical_table(const qmc_config &c, BuieTable &t);
}  // namespace qmc
