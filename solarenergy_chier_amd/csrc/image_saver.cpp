// image_saver.cpp — text output of a flux map in the reference's ImageSaver format
// (Tool/ImageSaver/ImageSaver.cpp:13-57): rows written bottom-up (file row r = image row height-1-r),
// values below Epsilon (1e-5, compared in double) written as "0", others in fixed notation with
// `precision` decimals, comma-separated, one image row per line. The statistics follow the same loop:
// a float running maximum whose location is the last (row, col) where the value equals it, and a
// float sum of the values at or above Epsilon in file order.
#include <cstdio>
#include <string>
#include <vector>

#include "scene_host.h"

using qmc::fail;

extern "C" qmc_status qmc_image_save_text(const char *path, int32_t height, int32_t width, const float *data,
                                          int32_t precision, qmc_image_stats *stats) {
    if (!path || !data) return fail(QMC_E_INVALID_ARG, "qmc_image_save_text: null argument");
    if (height <= 0 || width <= 0) return fail(QMC_E_INVALID_ARG, "qmc_image_save_text: empty image");
    if (precision < 0 || precision > 9) return fail(QMC_E_INVALID_ARG, "qmc_image_save_text: precision must be 0..9");
    FILE *f = std::fopen(path, "wb");
    if (!f) return fail(QMC_E_IO, std::string("qmc_image_save_text: cannot open '") + path + "' for writing");
    const double eps = 1e-5;  // Util/global_constant.h:11
    float max_value = 0.0f, sum = 0.0f;
    int32_t max_row = -1, max_col = -1;
    std::string line;
    char buf[64];
    bool ok = true;
    for (int32_t r = 0; r < height && ok; ++r) {
        line.clear();
        const float *row = data + (size_t)(height - 1 - r) * width;
        for (int32_t c = 0; c < width; ++c) {
            const float v = row[c];
            if ((double)v < eps) {
                line += '0';
            } else {
                max_value = v > max_value ? v : max_value;
                if (max_value == v) {
                    max_row = r;
                    max_col = c;
                }
                sum += v;
                const int k = std::snprintf(buf, sizeof(buf), "%.*f", (int)precision, (double)v);
                line.append(buf, k > 0 ? (size_t)k : 0);
            }
            line += c != width - 1 ? ',' : '\n';
        }
        ok = std::fwrite(line.data(), 1, line.size(), f) == line.size();
    }
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) return fail(QMC_E_IO, std::string("qmc_image_save_text: write to '") + path + "' failed");
    if (stats) {
        stats->max_value = max_value;
        stats->sum = sum;
        stats->max_row = max_row;
        stats->max_col = max_col;
    }
    return QMC_OK;
}
