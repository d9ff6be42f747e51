// scene_host.cpp — host scene model: .json configuration, .scn scene DSL, task files and the
// geometry processing that precedes tracing. No GPU calls in this file.
//
// Reference behaviour restated here (citations are to JachyLikeCoding/SolarEnergy_Chier):
//   SceneConfiguration::loadConfiguration     Scene/SceneConfiguration.cpp:39-69
//   SceneLoader::SceneFileRead + checkScene    Scene/SceneLoader.cpp:200-353
//   scene grammar D R G H+ ((R G|G) H+)*       Tool/FileLoader/RegularExpressionTree.cpp:47-89
//   SceneProcessor geometry                    Scene/SceneProcessor.cpp:13-39
//   receivers                                  Scene/DataStructure/Receiver/{Rectangle,Cylinder}Receiver.cu
//   grid cell lists                            Scene/DataStructure/Grid/RectangleGrid.cu:8-117
//   heliostat normal / vertices                Scene/DataStructure/Heliostat/Heliostat.cu:15-40
//   focus-flat facets                          Scene/DataStructure/Heliostat/FocusFlatRectangleHelio.cu:123-227,333-339
//   task file                                  RayTracingPipeline/TaskLoader.cpp:114-204
#include "scene_host.h"

#include <algorithm>
#include <cctype>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <memory>
#include <set>
#include <sstream>
#include <stdexcept>

namespace qmc {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

qmc_status fail(qmc_status s, const std::string &msg) {
    g_last_error = msg;
    return s;
}

}  // namespace qmc

using namespace qmc;

extern "C" const char *qmc_last_error(void) { return g_last_error.c_str(); }

extern "C" const char *qmc_status_string(qmc_status s) {
    switch (s) {
        case QMC_OK: return "ok";
        case QMC_E_INVALID_ARG: return "invalid argument";
        case QMC_E_RANGE: return "index out of range";
        case QMC_E_HIP: return "HIP runtime error";
        case QMC_E_OOM: return "out of memory";
        case QMC_E_UNSUPPORTED: return "unsupported";
        case QMC_E_IO: return "I/O error";
        case QMC_E_PARSE: return "parse error";
        case QMC_E_STATE: return "invalid call order";
        default: return "unknown status";
    }
}

// ---------------------------------------------------------------------------------------------
// Configuration (.json)
// ---------------------------------------------------------------------------------------------
extern "C" qmc_status qmc_config_default(qmc_config *c) {
    if (!c) return fail(QMC_E_INVALID_ARG, "qmc_config_default: null config");
    // Scene/SceneConfiguration.h:37,41-51
    c->sun_dir[0] = 1.0f;
    c->sun_dir[1] = 0.0f;
    c->sun_dir[2] = 0.0f;
    c->dni = 1000.0f;
    c->csr = 0.1f;
    c->num_sunshape_groups = 128;
    c->num_sunshape_lights_per_group = 2048;
    c->inverse_transform_sampling_groups = 1024;
    c->receiver_pixel_length = 0.01f;
    c->helio_disturb_std = 0.001f;
    c->helio_pixel_length = 0.01f;
    c->helio_reflected_rate = 0.88f;
    return QMC_OK;
}

namespace {

// Minimal JSON reader: enough of RFC 8259 for configuration files (objects, arrays, numbers,
// strings, true/false/null). Any syntax error throws std::runtime_error.
struct JValue {
    enum Kind { Null, Bool, Num, Str, Arr, Obj } kind = Null;
    double num = 0.0;
    bool b = false;
    std::string str;
    std::vector<JValue> arr;
    std::vector<std::pair<std::string, JValue>> obj;
    const JValue *find(const std::string &k) const {
        for (auto &kv : obj)
            if (kv.first == k) return &kv.second;
        return nullptr;
    }
};

class JParser {
public:
    explicit JParser(const std::string &s) : s_(s) {}
    JValue parse_document() {
        JValue v = value();
        ws();
        if (i_ != s_.size()) err("trailing characters");
        return v;
    }

private:
    const std::string &s_;
    size_t i_ = 0;
    [[noreturn]] void err(const char *what) {
        throw std::runtime_error(std::string("json: ") + what + " at offset " + std::to_string(i_));
    }
    void ws() {
        while (i_ < s_.size() && (s_[i_] == ' ' || s_[i_] == '\t' || s_[i_] == '\n' || s_[i_] == '\r'))
            ++i_;
    }
    bool lit(const char *w) {
        size_t n = std::strlen(w);
        if (s_.compare(i_, n, w) == 0) {
            i_ += n;
            return true;
        }
        return false;
    }
    std::string string_() {
        if (s_[i_] != '"') err("expected string");
        ++i_;
        std::string out;
        while (i_ < s_.size() && s_[i_] != '"') {
            char c = s_[i_++];
            if (c == '\\') {
                if (i_ >= s_.size()) err("bad escape");
                char e = s_[i_++];
                switch (e) {
                    case 'n': out += '\n'; break;
                    case 't': out += '\t'; break;
                    case 'r': out += '\r'; break;
                    case 'b': out += '\b'; break;
                    case 'f': out += '\f'; break;
                    case 'u': i_ += 4; out += '?'; break;
                    default: out += e;
                }
            } else {
                out += c;
            }
        }
        if (i_ >= s_.size()) err("unterminated string");
        ++i_;
        return out;
    }
    JValue value() {
        ws();
        if (i_ >= s_.size()) err("unexpected end");
        JValue v;
        char c = s_[i_];
        if (c == '{') {
            v.kind = JValue::Obj;
            ++i_;
            ws();
            if (i_ < s_.size() && s_[i_] == '}') {
                ++i_;
                return v;
            }
            for (;;) {
                ws();
                if (i_ >= s_.size()) err("unexpected end in object");
                std::string k = string_();
                ws();
                if (i_ >= s_.size() || s_[i_] != ':') err("expected ':'");
                ++i_;
                JValue item = value();
                v.obj.emplace_back(k, std::move(item));
                ws();
                if (i_ < s_.size() && s_[i_] == ',') {
                    ++i_;
                    continue;
                }
                if (i_ < s_.size() && s_[i_] == '}') {
                    ++i_;
                    return v;
                }
                err("expected ',' or '}'");
            }
        }
        if (c == '[') {
            v.kind = JValue::Arr;
            ++i_;
            ws();
            if (i_ < s_.size() && s_[i_] == ']') {
                ++i_;
                return v;
            }
            for (;;) {
                v.arr.push_back(value());
                ws();
                if (i_ < s_.size() && s_[i_] == ',') {
                    ++i_;
                    continue;
                }
                if (i_ < s_.size() && s_[i_] == ']') {
                    ++i_;
                    return v;
                }
                err("expected ',' or ']'");
            }
        }
        if (c == '"') {
            v.kind = JValue::Str;
            v.str = string_();
            return v;
        }
        if (lit("true")) {
            v.kind = JValue::Bool;
            v.b = true;
            return v;
        }
        if (lit("false")) {
            v.kind = JValue::Bool;
            return v;
        }
        if (lit("null")) return v;
        if (c == '-' || (c >= '0' && c <= '9')) {
            size_t st = i_;
            if (s_[i_] == '-') ++i_;
            while (i_ < s_.size() && (std::isdigit((unsigned char)s_[i_]) || s_[i_] == '.' ||
                                      s_[i_] == 'e' || s_[i_] == 'E' || s_[i_] == '+' || s_[i_] == '-'))
                ++i_;
            std::string tok = s_.substr(st, i_ - st);
            char *end = nullptr;
            v.kind = JValue::Num;
            v.num = std::strtod(tok.c_str(), &end);
            if (!end || *end != '\0') err("bad number");
            return v;
        }
        err("unexpected character");
    }
};

bool jnum(const JValue *v, double &out) {
    if (!v) return false;
    if (v->kind == JValue::Num) {
        out = v->num;
        return true;
    }
    if (v->kind == JValue::Bool) {
        out = v->b ? 1.0 : 0.0;
        return true;
    }
    throw std::runtime_error("json: field is not a number");
}

}  // namespace

extern "C" qmc_status qmc_scene_load_config(qmc_scene *s, const char *path, int32_t *n_fields) {
    if (!s || !path) return fail(QMC_E_INVALID_ARG, "qmc_scene_load_config: null argument");
    std::ifstream f(path);
    if (f.fail()) return fail(QMC_E_IO, std::string("The file '") + path + "' doesn't exist");
    std::stringstream ss;
    ss << f.rdbuf();
    std::string text = ss.str();
    if (!s->cfg_loaded) {
        qmc_config_default(&s->cfg);
        s->cfg_loaded = true;
    }
    qmc_config c = s->cfg;
    int found = 0;
    try {
        JValue doc = JParser(text).parse_document();
        if (doc.kind != JValue::Obj) throw std::runtime_error("json: top level is not an object");
        double x;
        if (const JValue *sd = doc.find("sun_dir")) {
            if (sd->kind != JValue::Arr || sd->arr.size() < 3)
                throw std::runtime_error("json: sun_dir is not a 3-array");
            for (int k = 0; k < 3; ++k) {
                jnum(&sd->arr[k], x);
                c.sun_dir[k] = (float)x;
            }
            ++found;
        }
        if (jnum(doc.find("dni"), x)) { c.dni = (float)x; ++found; }
        if (jnum(doc.find("csr"), x)) { c.csr = (float)x; ++found; }
        if (jnum(doc.find("num_of_sunshape_groups"), x)) { c.num_sunshape_groups = (int)x; ++found; }
        if (jnum(doc.find("num_per_sunshape_group"), x)) { c.num_sunshape_lights_per_group = (int)x; ++found; }
        if (jnum(doc.find("inverse_transform_sampling_groups"), x)) { c.inverse_transform_sampling_groups = (int)x; ++found; }
        if (jnum(doc.find("receiver_pixel_length"), x)) { c.receiver_pixel_length = (float)x; ++found; }
        if (jnum(doc.find("helio_disturb_std"), x)) { c.helio_disturb_std = (float)x; ++found; }
        if (jnum(doc.find("helio_pixel_length"), x)) { c.helio_pixel_length = (float)x; ++found; }
        if (jnum(doc.find("helio_reflected_rate"), x)) { c.helio_reflected_rate = (float)x; ++found; }
    } catch (const std::exception &e) {
        return fail(QMC_E_PARSE, std::string(path) + ": " + e.what());
    }
    s->cfg = c;
    s->processed = false;
    if (n_fields) *n_fields = found;
    return QMC_OK;
}

extern "C" qmc_status qmc_scene_set_config(qmc_scene *s, const qmc_config *c) {
    if (!s || !c) return fail(QMC_E_INVALID_ARG, "qmc_scene_set_config: null argument");
    s->cfg = *c;
    s->cfg_loaded = true;
    s->processed = false;
    return QMC_OK;
}

extern "C" qmc_status qmc_scene_get_config(const qmc_scene *s, qmc_config *c) {
    if (!s || !c) return fail(QMC_E_INVALID_ARG, "qmc_scene_get_config: null argument");
    if (!s->cfg_loaded) return qmc_config_default(c);
    *c = s->cfg;
    return QMC_OK;
}

// ---------------------------------------------------------------------------------------------
// Scene object
// ---------------------------------------------------------------------------------------------
extern "C" qmc_status qmc_scene_create(qmc_scene **out) {
    if (!out) return fail(QMC_E_INVALID_ARG, "qmc_scene_create: null out");
    auto *s = new (std::nothrow) qmc_scene();
    if (!s) return fail(QMC_E_OOM, "qmc_scene_create: allocation failed");
    qmc_config_default(&s->cfg);
    *out = s;
    return QMC_OK;
}

extern "C" qmc_status qmc_scene_destroy(qmc_scene *s) {
    delete s;
    return QMC_OK;
}

static void scene_clear_content(qmc_scene *s) {
    s->receivers.clear();
    s->grids.clear();
    s->heliostats.clear();
    s->grid_num = 0;
    s->ground_length = s->ground_width = 0.0f;
    s->processed = false;
}

extern "C" qmc_status qmc_scene_add_receiver(qmc_scene *s, const qmc_receiver_desc *d) {
    if (!s || !d) return fail(QMC_E_INVALID_ARG, "qmc_scene_add_receiver: null argument");
    if (d->type != 0 && d->type != 1) return fail(QMC_E_UNSUPPORTED, "Receiver type has not defined.");
    HostReceiver r;
    r.d = *d;
    s->receivers.push_back(r);
    s->processed = false;
    return QMC_OK;
}

extern "C" qmc_status qmc_scene_add_grid(qmc_scene *s, const qmc_grid_desc *d) {
    if (!s || !d) return fail(QMC_E_INVALID_ARG, "qmc_scene_add_grid: null argument");
    if (d->type != 0) return fail(QMC_E_UNSUPPORTED, "Grid type has not defined.");
    if (d->heliostat_type != 0 && d->heliostat_type != 1)
        return fail(QMC_E_UNSUPPORTED, "Heliostat has not defined.");
    HostGrid g;
    g.d = *d;
    s->grids.push_back(g);
    s->grid_num = (int)s->grids.size();
    s->processed = false;
    return QMC_OK;
}

extern "C" qmc_status qmc_scene_add_heliostat(qmc_scene *s, const qmc_heliostat_desc *d) {
    if (!s || !d) return fail(QMC_E_INVALID_ARG, "qmc_scene_add_heliostat: null argument");
    HostHeliostat h;
    h.d = *d;
    int n = d->rows * d->cols;
    if (d->local_centers && n > 0) h.file_centers.assign(d->local_centers, d->local_centers + 3 * n);
    if (d->local_normals && n > 0) h.file_normals.assign(d->local_normals, d->local_normals + 3 * n);
    h.d.local_centers = nullptr;
    h.d.local_normals = nullptr;
    s->heliostats.push_back(std::move(h));
    s->processed = false;
    return QMC_OK;
}

// ---------------------------------------------------------------------------------------------
// .scn loader
// ---------------------------------------------------------------------------------------------
namespace {

struct SceneError : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct CorrespondingError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// DFA of the scene grammar: start -D-> ground -R-> receiver -G-> grid -H-> heliostat;
// heliostat -R/G/H-> receiver/grid/heliostat.
enum { S_START, S_GROUND, S_RECV, S_GRID, S_HELIO };
int dfa_step(int st, char c) {
    switch (st) {
        case S_START: if (c == 'D') return S_GROUND; break;
        case S_GROUND: if (c == 'R') return S_RECV; break;
        case S_RECV: if (c == 'G') return S_GRID; break;
        case S_GRID: if (c == 'H') return S_HELIO; break;
        case S_HELIO:
            if (c == 'R') return S_RECV;
            if (c == 'G') return S_GRID;
            if (c == 'H') return S_HELIO;
            break;
    }
    throw SceneError("Cannot step forward.");
}

int int_field(const char *name, std::istream &in) {
    std::string head;
    in >> head;
    if (head != name) throw SceneError(std::string("Miss ' ") + name + " ' property.\n");
    int v = 0;
    in >> v;
    return v;
}

void float3_field(const char *name, std::istream &in, float out[3]) {
    std::string head;
    in >> head;
    if (head != name) throw SceneError(std::string("Miss ' ") + name + " ' property.\n");
    in >> out[0] >> out[1] >> out[2];
}

void isolated_field(const char *name, std::istream &in) {
    std::string head;
    in >> head;
    if (head != name) throw SceneError(std::string("Miss ' ") + name + " ' property.\n");
}

void check_scene(const qmc_scene &s) {
    if (s.grid_num != (int)s.grids.size())
        throw CorrespondingError("The scene should contains " + std::to_string(s.grid_num) +
                                 " grids, but got " + std::to_string(s.grids.size()) + " grids.");
    for (size_t i = 0; i < s.grids.size(); ++i) {
        int expect = s.grids[i].d.n_heliostats;
        int exist = (i == s.grids.size() - 1) ? (int)s.heliostats.size()
                                               : s.grids[i + 1].d.start_heliostat;
        exist -= s.grids[i].d.start_heliostat;
        if (exist != expect)
            throw CorrespondingError("The " + std::to_string(i + 1) + "-thgrid should contain " +
                                     std::to_string(expect) + " heliostat(s), but got " +
                                     std::to_string(exist) + " heliostat(s).");
        int r = s.grids[i].d.receiver;
        if (i == 0 && r != 0)
            throw CorrespondingError("The 0-th grid is expected belonging to 0-th receiver, but got " +
                                     std::to_string(r) + ".");
        if (i > 0) {
            int prev = s.grids[i - 1].d.receiver;
            if (r != prev && r != prev + 1)
                throw CorrespondingError("The " + std::to_string(i) + "-th grid is expected belonging to " +
                                         std::to_string(prev) + "-th or " + std::to_string(prev + 1) +
                                         "-threceiver, but got " + std::to_string(r) + ".");
        }
    }
}

void load_scene_stream(qmc_scene *s, std::istream &in) {
    int mrows = 1, mcols = 1;
    float gap[2] = {0.0f, 0.0f};
    float surface[6] = {-1, -1, -1, -1, -1, -1};
    std::vector<float> centers, normals;
    int receiver_id = -1;
    int total_helio = 0;
    int helio_type = -1;
    int st = S_START;
    std::string head;
    while (in >> head) {
        if (head[0] == '#') {
            std::string rest;
            std::getline(in, rest);
            continue;
        }
        if (head == "gap") {
            in >> gap[0] >> gap[1];
        } else if (head == "matrix") {
            in >> mrows >> mcols;
        } else if (head == "surface_property") {
            for (int k = 0; k < 6; ++k) in >> surface[k];
        } else if (head == "ground") {
            st = dfa_step(st, 'D');
            in >> s->ground_length >> s->ground_width;
            s->grid_num = int_field("ngrid", in);
        } else if (head == "Recv") {
            st = dfa_step(st, 'R');
            qmc_receiver_desc d{};
            in >> d.type;
            if (d.type != 0 && d.type != 1) throw SceneError("Receiver type has not defined.\n");
            float3_field("pos", in, d.pos);
            float3_field("size", in, d.size);
            float3_field("norm", in, d.normal);
            d.face = int_field("face", in);
            isolated_field("end", in);
            HostReceiver r;
            r.d = d;
            s->receivers.push_back(r);
            ++receiver_id;
        } else if (head == "Grid") {
            st = dfa_step(st, 'G');
            qmc_grid_desc d{};
            in >> d.type;
            if (d.type != 0) throw SceneError("Grid type has not defined.\n");
            d.receiver = receiver_id;
            d.start_heliostat = total_helio;
            float3_field("pos", in, d.pos);
            float3_field("size", in, d.size);
            float3_field("inter", in, d.interval);
            d.n_heliostats = int_field("n", in);
            d.heliostat_type = int_field("type", in);
            isolated_field("end", in);
            HostGrid g;
            g.d = d;
            s->grids.push_back(g);
            helio_type = d.heliostat_type;
        } else if (head == "helio") {
            st = dfa_step(st, 'H');
            if (helio_type != 0 && helio_type != 1) throw SceneError("Heliostat has not defined.\n");
            HostHeliostat h;
            qmc_heliostat_desc &d = h.d;
            d.rows = mrows;
            d.cols = mcols;
            d.gap[0] = gap[0];
            d.gap[1] = gap[1];
            for (int k = 0; k < 6; ++k) d.surface_property[k] = surface[k];
            d.grid = (int)s->grids.size() - 1;
            in >> d.pos[0] >> d.pos[1] >> d.pos[2];
            in >> d.size[0] >> d.size[1] >> d.size[2];
            if ((int)surface[1] == 1) {
                h.file_centers = centers;
                h.file_normals = normals;
            }
            s->heliostats.push_back(std::move(h));
            centers.clear();
            normals.clear();
            ++total_helio;
        } else if (head == "subhelio") {
            std::string rest;
            std::getline(in, rest);
            int n = mrows * mcols;
            for (int k = 0; k < n; ++k) {
                float a, b, c;
                in >> a >> b >> c;
                centers.push_back(a);
                centers.push_back(b);
                centers.push_back(c);
                std::getline(in, rest);
            }
            for (int k = 0; k < n; ++k) {
                float a, b, c;
                in >> a >> b >> c;
                normals.push_back(a);
                normals.push_back(b);
                normals.push_back(c);
                std::getline(in, rest);
            }
        } else {
            st = dfa_step(st, '?');
        }
    }
    check_scene(*s);
}

}  // namespace

static qmc_status load_scene_text(qmc_scene *s, const std::string &text, const std::string &where) {
    scene_clear_content(s);
    std::stringstream in(text);
    try {
        load_scene_stream(s, in);
    } catch (const CorrespondingError &e) {
        scene_clear_content(s);
        return fail(QMC_E_PARSE, where + ": Error caused by ' " + e.what() + " '.");
    } catch (const std::exception &e) {
        scene_clear_content(s);
        return fail(QMC_E_PARSE, where + ": " + e.what());
    }
    return QMC_OK;
}

extern "C" qmc_status qmc_scene_load_file(qmc_scene *s, const char *path) {
    if (!s || !path) return fail(QMC_E_INVALID_ARG, "qmc_scene_load_file: null argument");
    std::ifstream f(path);
    if (f.fail()) {
        scene_clear_content(s);
        return fail(QMC_E_IO, std::string("Cannot open the file: ") + path);
    }
    std::stringstream ss;
    ss << f.rdbuf();
    return load_scene_text(s, ss.str(), path);
}

extern "C" qmc_status qmc_scene_load_text(qmc_scene *s, const char *text) {
    if (!s || !text) return fail(QMC_E_INVALID_ARG, "qmc_scene_load_text: null argument");
    return load_scene_text(s, text, "<text>");
}

// ---------------------------------------------------------------------------------------------
// Processing
// ---------------------------------------------------------------------------------------------
namespace {

V3 f3(const float p[3]) { return V3{p[0], p[1], p[2]}; }

void receiver_init(HostReceiver &r, int ppm) {
    r.pixel_length = 1.0f / float(ppm);
    const V3 size = f3(r.d.size);
    if (r.d.type == 0) {
        // RectangleReceiver::Cinit_vertex (RectangleReceiver.cu:40-96)
        switch (r.d.face) {
            case 0: r.local_normal = V3{0.0f, 0.0f, 1.0f}; break;
            case 1: r.local_normal = V3{1.0f, 0.0f, 0.0f}; break;
            case 2: r.local_normal = V3{0.0f, 0.0f, -1.0f}; break;
            case 3: r.local_normal = V3{-1.0f, 0.0f, 0.0f}; break;
            default: r.local_normal = V3{0.0f, 0.0f, 0.0f}; break;
        }
        V3 *v = r.vert;
        switch (r.d.face) {
            case 0:
                v[0] = V3{-size.x / 2, -size.y / 2, size.z / 2};
                v[1] = V3{-size.x / 2, size.y / 2, size.z / 2};
                v[2] = V3{size.x / 2, size.y / 2, size.z / 2};
                v[3] = V3{size.x / 2, -size.y / 2, size.z / 2};
                break;
            case 1:
                v[0] = V3{size.x / 2, -size.y / 2, size.z / 2};
                v[1] = V3{size.x / 2, size.y / 2, size.z / 2};
                v[2] = V3{size.x / 2, size.y / 2, -size.z / 2};
                v[3] = V3{size.x / 2, -size.y / 2, -size.z / 2};
                break;
            case 2:
                v[0] = V3{size.x / 2, -size.y / 2, -size.z / 2};
                v[1] = V3{size.x / 2, size.y / 2, -size.z / 2};
                v[2] = V3{-size.x / 2, size.y / 2, -size.z / 2};
                v[3] = V3{-size.x / 2, -size.y / 2, -size.z / 2};
                break;
            case 3:
                v[0] = V3{-size.x / 2, -size.y / 2, -size.z / 2};
                v[1] = V3{-size.x / 2, size.y / 2, -size.z / 2};
                v[2] = V3{-size.x / 2, size.y / 2, size.z / 2};
                v[3] = V3{-size.x / 2, -size.y / 2, size.z / 2};
                break;
            default:
                for (int k = 0; k < 4; ++k) v[k] = V3{0, 0, 0};
        }
        r.normal = normalize(f3(r.d.normal));
        if (r.d.tilt_deg != 0.0f) {
            // Extension: tilt the face about the local x axis (positive tilts it downward),
            // then apply the reference's yaw rotation.
            double a = (double)r.d.tilt_deg * 3.14159265358979323846 / 180.0;
            float ca = (float)cos(a), sa = (float)sin(a);
            for (int k = 0; k < 4; ++k) {
                V3 p = v[k];
                v[k] = V3{p.x, p.y * ca - p.z * sa, p.y * sa + p.z * ca};
            }
        }
        for (int k = 0; k < 4; ++k) v[k] = rotate_y(v[k], r.local_normal, r.normal);
        for (int k = 0; k < 4; ++k) v[k] = v[k] + f3(r.d.pos);
        r.focus_center = (v[0] + v[2]) / 2.0f;
        r.res_x = (int)(size.x * float(ppm));
        r.res_y = (int)(size.y * float(ppm));
        r.area = r.pixel_length * r.pixel_length;
        r.res_override = false;
        if (r.d.res_x > 0 && r.d.res_y > 0) {
            r.res_override = true;
            r.res_x = r.d.res_x;
            r.res_y = r.d.res_y;
            r.pixel_length = size.x / float(r.res_x);
            r.area = (size.x / float(r.res_x)) * (size.y / float(r.res_y));
        }
    } else {
        // CylinderReceiver::CInit / Cset_resolution (CylinderReceiver.cu:4-15)
        r.res_x = (int)std::ceil(2 * M_PI * size.x * float(ppm));
        r.res_y = (int)(size.y * float(ppm));
        r.area = r.pixel_length * r.pixel_length;
        r.res_override = false;
        r.normal = f3(r.d.normal);
        if (r.d.res_x > 0 && r.d.res_y > 0) {
            r.res_override = true;
            r.res_x = r.d.res_x;
            r.res_y = r.d.res_y;
            float circ = (float)(2 * M_PI * size.x);
            r.pixel_length = circ / float(r.res_x);
            r.area = (circ / float(r.res_x)) * (size.y / float(r.res_y));
        }
    }
}

bool inside_cylinder(const HostReceiver &r, V3 p) {
    V2 l{p.x - r.d.pos[0], p.z - r.d.pos[2]};
    return dot2(l, l) <= r.d.size[0] * r.d.size[0];
}

V3 receiver_focus(const HostReceiver &r, V3 hpos) {
    if (r.d.type == 0) return r.focus_center;
    // CylinderReceiver::getFocusCenter (CylinderReceiver.cu:18-31)
    V3 pos = f3(r.d.pos);
    if (inside_cylinder(r, hpos)) return pos;
    V3 dir = normalize(hpos - pos);
    float radius = r.d.size[0] / length2(V2{dir.x, dir.z});
    float x = pos.x + dir.x * radius;
    float z = pos.z + dir.z * radius;
    return V3{x, pos.y, z};
}

// Aim point with the aim-offset extension: focus centre + off[0]*across + off[1]*up, where for a
// rectangle across/up are the unit edge directions v3-v0 (width) and v1-v0 (height), and for a
// cylinder the horizontal tangent at the focus point and the axis (0,1,0). A zero offset returns
// the reference focus centre unchanged.
V3 aim_point(const HostReceiver &r, V3 hpos, const float off[2]) {
    V3 f = receiver_focus(r, hpos);
    if (off[0] == 0.0f && off[1] == 0.0f) return f;
    V3 across, up;
    if (r.d.type == 0) {
        across = normalize(r.vert[3] - r.vert[0]);
        up = normalize(r.vert[1] - r.vert[0]);
    } else {
        V3 rad = f - f3(r.d.pos);
        across = (rad.x == 0.0f && rad.z == 0.0f) ? V3{1.0f, 0.0f, 0.0f} : normalize(V3{-rad.z, 0.0f, rad.x});
        up = V3{0.0f, 1.0f, 0.0f};
    }
    return f + across * off[0] + up * off[1];
}

int sub_size(const HostHeliostat &h) {
    if (h.type == 1) return h.d.rows * h.d.cols;
    return h.d.sub_count_override > 0 ? h.d.sub_count_override : 1;
}

// RectangleGrid::Cinit + CGridHelioMatch (RectangleGrid.cu:8-12,27-117).
void grid_match(HostGrid &g, const std::vector<HostHeliostat> &hs) {
    V3 gpos = f3(g.d.pos), gsize = f3(g.d.size), gint = f3(g.d.interval);
    g.nx = (int)std::ceil(gsize.x / gint.x);
    g.ny = (int)std::ceil(gsize.y / gint.y);
    g.nz = (int)std::ceil(gsize.z / gint.z);
    const size_t ncell = (size_t)g.nx * g.ny * g.nz;
    // two passes (count, then fill) over the cell ranges of every heliostat's bounding sphere; each
    // cell lists sub-heliostat ids in heliostat order, as the reference's per-cell appends do
    struct Range {
        int x0, x1, y0, y1, z0, z1, first, n;
    };
    std::vector<Range> ranges;
    ranges.reserve(g.d.n_heliostats);
    int start_sub = 0;
    for (int i = g.d.start_heliostat; i < g.d.start_heliostat + g.d.n_heliostats; ++i) {
        V3 pos = f3(hs[i].d.pos);
        float radius = length(f3(hs[i].d.size)) / 2;
        V3 mn = pos - radius, mx = pos + radius;
        V3 a = (mn - gpos) / gint, b = (mx - gpos) / gint;
        const int ss = sub_size(hs[i]);
        ranges.push_back(Range{std::max(0, (int)a.x), std::min(g.nx - 1, (int)b.x), std::max(0, (int)a.y),
                               std::min(g.ny - 1, (int)b.y), std::max(0, (int)a.z), std::min(g.nz - 1, (int)b.z),
                               start_sub, ss});
        start_sub += ss;
    }
    g.sub_box.assign(6 * (size_t)start_sub, 0);
    for (const Range &r : ranges)
        for (int k = 0; k < r.n; ++k) {
            int *b = &g.sub_box[6 * (size_t)(r.first + k)];
            b[0] = r.x0, b[1] = r.x1, b[2] = r.y0, b[3] = r.y1, b[4] = r.z0, b[5] = r.z1;
        }
    g.cell_start.assign(ncell + 1, 0);
    for (const Range &r : ranges)
        for (int x = r.x0; x <= r.x1; ++x)
            for (int y = r.y0; y <= r.y1; ++y)
                for (int z = r.z0; z <= r.z1; ++z) g.cell_start[(size_t)x * g.ny * g.nz + (size_t)y * g.nz + z + 1] += r.n;
    for (size_t c = 0; c < ncell; ++c) g.cell_start[c + 1] += g.cell_start[c];
    g.match.assign((size_t)g.cell_start[ncell], 0);
    std::vector<int> fill(g.cell_start.begin(), g.cell_start.end() - 1);
    for (const Range &r : ranges)
        for (int x = r.x0; x <= r.x1; ++x)
            for (int y = r.y0; y <= r.y1; ++y)
                for (int z = r.z0; z <= r.z1; ++z) {
                    int &f = fill[(size_t)x * g.ny * g.nz + (size_t)y * g.nz + z];
                    for (int k = 0; k < r.n; ++k) g.match[f++] = r.first + k;
                }
}

void heliostat_process(HostHeliostat &h, V3 aim, V3 sun_dir, float pixel_length) {
    h.aim = aim;
    h.pixel_length = pixel_length;
    V3 pos = f3(h.d.pos), size = f3(h.d.size);
    h.focal = h.d.surface_property[0];
    if (h.type == 1 && h.focal < 0.0f) h.focal = length(aim - pos);  // FocusFlatRectangleHelio.cu:333-339
    // Heliostat::CSetNormal (Heliostat.cu:15-20)
    V3 reflect_dir = normalize(aim - pos);
    h.normal = normalize(reflect_dir - sun_dir);
    // Heliostat::CSetWorldVertex (Heliostat.cu:25-40)
    V3 v0 = V3{-size.x / 2, size.y / 2, -size.z / 2};
    V3 v[4];
    v[0] = v0;
    v[1] = V3{0, 0, size.z} + v0;
    v[2] = V3{size.x, 0, size.z} + v0;
    v[3] = V3{size.x, 0, 0} + v0;
    Frame fr = make_frame(h.normal, false);
    for (int k = 0; k < 4; ++k) h.vert[k] = frame_apply(fr, v[k]) + pos;
    // micro-facet layout (RectangleHelio.cu:84-116, FocusFlatRectangleHelio.cu:68-78)
    h.len_row = (size.z - h.d.gap[1] * (h.d.rows - 1)) / float(h.d.rows);
    h.len_col = (size.x - h.d.gap[0] * (h.d.cols - 1)) / float(h.d.cols);
    h.sub_rows = (int)(h.len_row / pixel_length);
    h.sub_cols = (int)(h.len_col / pixel_length);
    h.n_micro = h.d.rows * h.d.cols * h.sub_rows * h.sub_cols;
    h.sub_count = sub_size(h);
    h.loc_centers.clear();
    h.loc_normals.clear();
    if (h.type == 1) {
        // FocusFlatRectangleHelio::CGetSubHeliostatVertexes (FocusFlatRectangleHelio.cu:123-227)
        int n = h.d.rows * h.d.cols;
        h.loc_centers.resize(n);
        h.loc_normals.resize(n);
        int method = (int)h.d.surface_property[1];
        for (int r = 0; r < h.d.rows; ++r)
            for (int c = 0; c < h.d.cols; ++c) {
                int id = r * h.d.cols + c;
                if (method == 1 && (int)h.file_centers.size() >= 3 * n && (int)h.file_normals.size() >= 3 * n) {
                    h.loc_centers[id] = V3{h.file_centers[3 * id], h.file_centers[3 * id + 1], h.file_centers[3 * id + 2]};
                    V3 nn{h.file_normals[3 * id], h.file_normals[3 * id + 1], h.file_normals[3 * id + 2]};
                    h.loc_normals[id] = normalize(nn);
                } else {
                    V3 cc{c * (h.d.gap[0] + h.len_col) - size.x / 2 + h.len_col / 2, size.y / 2,
                          r * (h.d.gap[1] + h.len_row) - size.z / 2 + h.len_row / 2};
                    cc.y += (cc.x * cc.x + cc.z * cc.z) / (4 * h.focal);
                    h.loc_centers[id] = cc;
                    V3 nn{-1 / (2 * h.focal) * cc.x, 1.0f, -1 / (2 * h.focal) * cc.z};
                    h.loc_normals[id] = normalize(nn);
                }
            }
    }
}

void append_sub_vertices(const HostHeliostat &h, std::vector<V3> &out) {
    if (h.type == 0) {
        // RectangleHelio::CGetSubHeliostatVertexes (RectangleHelio.cu:14-18): one quad, gaps ignored
        out.push_back(h.vert[0]);
        out.push_back(h.vert[1]);
        out.push_back(h.vert[2]);
        return;
    }
    V3 pos = f3(h.d.pos);
    V3 local[3] = {V3{-h.len_col / 2, 0.0f, h.len_row / 2}, V3{-h.len_col / 2, 0.0f, -h.len_row / 2},
                   V3{h.len_col / 2, 0.0f, -h.len_row / 2}};
    Frame fh = make_frame(h.normal, true);
    for (int r = 0; r < h.d.rows; ++r)
        for (int c = 0; c < h.d.cols; ++c) {
            int id = r * h.d.cols + c;
            Frame ff = make_frame(h.loc_normals[id], true);
            for (int k = 0; k < 3; ++k) {
                V3 p = frame_apply(ff, local[k]) + h.loc_centers[id];
                p = frame_apply(fh, p) + pos;
                out.push_back(p);
            }
        }
}

}  // namespace

extern "C" qmc_status qmc_scene_process(qmc_scene *s) {
    if (!s) return fail(QMC_E_INVALID_ARG, "qmc_scene_process: null scene");
    const qmc_config &c = s->cfg;
    if (!(c.receiver_pixel_length > 0.0f) || !(c.helio_pixel_length > 0.0f))
        return fail(QMC_E_INVALID_ARG, "qmc_scene_process: pixel lengths must be positive");
    int ppm = (int)(1.0f / c.receiver_pixel_length);  // SceneProcessor.cpp:20
    if (ppm <= 0) return fail(QMC_E_INVALID_ARG, "qmc_scene_process: receiver_pixel_length > 1 gives 0 pixels per meter");
    V3 sun = f3(c.sun_dir);
    // validate structure
    for (size_t gi = 0; gi < s->grids.size(); ++gi) {
        const HostGrid &g = s->grids[gi];
        if (g.d.receiver < 0 || g.d.receiver >= (int)s->receivers.size())
            return fail(QMC_E_RANGE, "grid " + std::to_string(gi) + " belongs to receiver " +
                                         std::to_string(g.d.receiver) + " which does not exist");
        if (g.d.start_heliostat < 0 || g.d.n_heliostats < 0 ||
            g.d.start_heliostat + g.d.n_heliostats > (int)s->heliostats.size())
            return fail(QMC_E_RANGE, "grid " + std::to_string(gi) + " heliostat range is invalid");
        if (!(g.d.interval[0] > 0 && g.d.interval[1] > 0 && g.d.interval[2] > 0))
            return fail(QMC_E_INVALID_ARG, "grid " + std::to_string(gi) + " interval must be positive");
    }
    for (auto &h : s->heliostats) {
        if (h.d.rows <= 0 || h.d.cols <= 0)
            return fail(QMC_E_INVALID_ARG, "heliostat matrix must be positive");
        if (h.d.grid < 0 || h.d.grid >= (int)s->grids.size())
            return fail(QMC_E_RANGE, "heliostat belongs to a grid that does not exist");
        h.type = s->grids[h.d.grid].d.heliostat_type;
    }
    for (auto &g : s->grids) {
        grid_match(g, s->heliostats);
        HostReceiver &r = s->receivers[g.d.receiver];
        receiver_init(r, ppm);
        int sub = 0;
        g.sub_verts.clear();
        for (int i = 0; i < g.d.n_heliostats; ++i) {
            HostHeliostat &h = s->heliostats[g.d.start_heliostat + i];
            V3 aim = aim_point(r, f3(h.d.pos), h.d.aim_offset);
            heliostat_process(h, aim, sun, c.helio_pixel_length);
            h.sub_begin = sub;
            sub += h.sub_count;
            append_sub_vertices(h, g.sub_verts);
        }
        g.n_sub = sub;
    }
    s->processed = true;
    return QMC_OK;
}

namespace qmc {
float heliostat_factor(const qmc_scene &s, int h) {
    const HostHeliostat &hh = s.heliostats[h];
    const HostReceiver &r = s.receivers[s.grids[hh.d.grid].d.receiver];
    const qmc_config &c = s.cfg;
    if (!r.res_override) {
        // QuasiMonteCarloRayTracer.cpp:58-60
        float ratio = hh.pixel_length / r.pixel_length;
        return c.dni * ratio * ratio * c.helio_reflected_rate / float(c.num_sunshape_lights_per_group);
    }
    return c.dni * (hh.pixel_length * hh.pixel_length / r.area) * c.helio_reflected_rate /
           float(c.num_sunshape_lights_per_group);
}

void heliostat_factor_parts(const qmc_scene &s, int h, int &rule, float &x) {
    const HostHeliostat &hh = s.heliostats[h];
    const HostReceiver &r = s.receivers[s.grids[hh.d.grid].d.receiver];
    if (!r.res_override) {
        rule = 0;
        x = hh.pixel_length / r.pixel_length;
    } else {
        rule = 1;
        x = hh.pixel_length * hh.pixel_length / r.area;
    }
}

V3 heliostat_aim(const qmc_scene &s, int h, const float off[2]) {
    const HostHeliostat &hh = s.heliostats[h];
    const HostReceiver &r = s.receivers[s.grids[hh.d.grid].d.receiver];
    return aim_point(r, f3(hh.d.pos), off);
}

void aim_points_f64(const qmc_scene &s, std::vector<double> &out) {
    out.assign(3 * s.heliostats.size(), 0.0);
    for (size_t h = 0; h < s.heliostats.size(); ++h) {
        const HostHeliostat &hh = s.heliostats[h];
        const HostReceiver &r = s.receivers[s.grids[hh.d.grid].d.receiver];
        double *a = &out[3 * h];
        if (r.d.type != 0) {
            a[0] = hh.aim.x;
            a[1] = hh.aim.y;
            a[2] = hh.aim.z;
            continue;
        }
        const V3 f = receiver_focus(r, f3(hh.d.pos));
        const V3 across = normalize(r.vert[3] - r.vert[0]), up = normalize(r.vert[1] - r.vert[0]);
        const double o0 = hh.d.aim_offset[0], o1 = hh.d.aim_offset[1];
        a[0] = (double)f.x + (double)across.x * o0 + (double)up.x * o1;
        a[1] = (double)f.y + (double)across.y * o0 + (double)up.y * o1;
        a[2] = (double)f.z + (double)across.z * o0 + (double)up.z * o1;
    }
}
}  // namespace qmc

extern "C" qmc_status qmc_scene_set_aim_offsets(qmc_scene *s, const float *off, int32_t n) {
    if (!s || !off) return fail(QMC_E_INVALID_ARG, "qmc_scene_set_aim_offsets: null argument");
    if (n != 2 * (int)s->heliostats.size())
        return fail(QMC_E_INVALID_ARG, "qmc_scene_set_aim_offsets: need 2 floats per heliostat");
    for (size_t h = 0; h < s->heliostats.size(); ++h) {
        s->heliostats[h].d.aim_offset[0] = off[2 * h];
        s->heliostats[h].d.aim_offset[1] = off[2 * h + 1];
    }
    s->processed = false;
    return QMC_OK;
}

// ---------------------------------------------------------------------------------------------
// Getters
// ---------------------------------------------------------------------------------------------
extern "C" qmc_status qmc_scene_get_counts(const qmc_scene *s, qmc_scene_counts *o) {
    if (!s || !o) return fail(QMC_E_INVALID_ARG, "qmc_scene_get_counts: null argument");
    o->n_receivers = (int)s->receivers.size();
    o->n_grids = (int)s->grids.size();
    o->n_heliostats = (int)s->heliostats.size();
    o->grid_num_declared = s->grid_num;
    o->ground_length = s->ground_length;
    o->ground_width = s->ground_width;
    return QMC_OK;
}

#define CHECK_INDEX(v, n, what)                                                               \
    if ((v) < 0 || (v) >= (int)(n))                                                           \
        return fail(QMC_E_RANGE, std::string(what) + " index " + std::to_string(v) + " out of range [0, " + \
                                     std::to_string(n) + ")");

extern "C" qmc_status qmc_scene_get_receiver_desc(const qmc_scene *s, int32_t r, qmc_receiver_desc *o) {
    if (!s || !o) return fail(QMC_E_INVALID_ARG, "null argument");
    CHECK_INDEX(r, s->receivers.size(), "receiver");
    *o = s->receivers[r].d;
    return QMC_OK;
}

extern "C" qmc_status qmc_scene_get_grid_desc(const qmc_scene *s, int32_t g, qmc_grid_desc *o) {
    if (!s || !o) return fail(QMC_E_INVALID_ARG, "null argument");
    CHECK_INDEX(g, s->grids.size(), "grid");
    *o = s->grids[g].d;
    return QMC_OK;
}

extern "C" qmc_status qmc_scene_get_heliostat_desc(const qmc_scene *s, int32_t h, qmc_heliostat_desc *o) {
    if (!s || !o) return fail(QMC_E_INVALID_ARG, "null argument");
    CHECK_INDEX(h, s->heliostats.size(), "heliostat");
    const HostHeliostat &hh = s->heliostats[h];
    *o = hh.d;
    o->local_centers = hh.file_centers.empty() ? nullptr : hh.file_centers.data();
    o->local_normals = hh.file_normals.empty() ? nullptr : hh.file_normals.data();
    return QMC_OK;
}

static void put3(float o[3], V3 v) {
    o[0] = v.x;
    o[1] = v.y;
    o[2] = v.z;
}

extern "C" qmc_status qmc_scene_get_receiver_info(const qmc_scene *s, int32_t r, qmc_receiver_info *o) {
    if (!s || !o) return fail(QMC_E_INVALID_ARG, "null argument");
    CHECK_INDEX(r, s->receivers.size(), "receiver");
    if (!s->processed) return fail(QMC_E_STATE, "scene not processed");
    const HostReceiver &rr = s->receivers[r];
    std::memset(o, 0, sizeof(*o));
    o->type = rr.d.type;
    o->res_x = rr.res_x;
    o->res_y = rr.res_y;
    o->pixel_length = rr.pixel_length;
    o->factor_area = rr.area;
    if (rr.d.type == 0)
        for (int k = 0; k < 4; ++k) put3(o->vertices[k], rr.vert[k]);
    put3(o->normal, rr.normal);
    put3(o->focus_center, rr.focus_center);
    return QMC_OK;
}

extern "C" qmc_status qmc_scene_get_heliostat_info(const qmc_scene *s, int32_t h, qmc_heliostat_info *o) {
    if (!s || !o) return fail(QMC_E_INVALID_ARG, "null argument");
    CHECK_INDEX(h, s->heliostats.size(), "heliostat");
    if (!s->processed) return fail(QMC_E_STATE, "scene not processed");
    const HostHeliostat &hh = s->heliostats[h];
    std::memset(o, 0, sizeof(*o));
    put3(o->pos, f3(hh.d.pos));
    put3(o->size, f3(hh.d.size));
    put3(o->normal, hh.normal);
    put3(o->aim, hh.aim);
    for (int k = 0; k < 4; ++k) put3(o->vertices[k], hh.vert[k]);
    o->focal_length = hh.focal;
    o->type = hh.type;
    o->grid = hh.d.grid;
    o->n_micro = hh.n_micro;
    o->sub_rows = hh.sub_rows;
    o->sub_cols = hh.sub_cols;
    o->sub_begin = hh.sub_begin;
    o->sub_count = hh.sub_count;
    return QMC_OK;
}

extern "C" qmc_status qmc_scene_get_grid_info(const qmc_scene *s, int32_t g, qmc_grid_info *o) {
    if (!s || !o) return fail(QMC_E_INVALID_ARG, "null argument");
    CHECK_INDEX(g, s->grids.size(), "grid");
    if (!s->processed) return fail(QMC_E_STATE, "scene not processed");
    const HostGrid &gg = s->grids[g];
    o->nx = gg.nx;
    o->ny = gg.ny;
    o->nz = gg.nz;
    o->n_match = (int)gg.match.size();
    o->n_sub = gg.n_sub;
    return QMC_OK;
}

extern "C" qmc_status qmc_scene_get_grid_cells(const qmc_scene *s, int32_t g, int32_t *cell_start, int32_t *match) {
    if (!s) return fail(QMC_E_INVALID_ARG, "null argument");
    CHECK_INDEX(g, s->grids.size(), "grid");
    if (!s->processed) return fail(QMC_E_STATE, "scene not processed");
    const HostGrid &gg = s->grids[g];
    if (cell_start) std::memcpy(cell_start, gg.cell_start.data(), gg.cell_start.size() * sizeof(int));
    if (match && !gg.match.empty()) std::memcpy(match, gg.match.data(), gg.match.size() * sizeof(int));
    return QMC_OK;
}

extern "C" qmc_status qmc_scene_get_subheliostat_vertices(const qmc_scene *s, int32_t g, float *out,
                                                          int32_t cap, int32_t *n) {
    if (!s) return fail(QMC_E_INVALID_ARG, "null argument");
    CHECK_INDEX(g, s->grids.size(), "grid");
    if (!s->processed) return fail(QMC_E_STATE, "scene not processed");
    const HostGrid &gg = s->grids[g];
    int nv = (int)gg.sub_verts.size();
    if (n) *n = nv;
    if (out) {
        if (cap < nv) return fail(QMC_E_INVALID_ARG, "vertex buffer too small");
        for (int i = 0; i < nv; ++i) put3(out + 3 * i, gg.sub_verts[i]);
    }
    return QMC_OK;
}

extern "C" qmc_status qmc_scene_get_factor(const qmc_scene *s, int32_t h, float *factor) {
    if (!s || !factor) return fail(QMC_E_INVALID_ARG, "null argument");
    CHECK_INDEX(h, s->heliostats.size(), "heliostat");
    if (!s->processed) return fail(QMC_E_STATE, "scene not processed");
    *factor = heliostat_factor(*s, h);
    return QMC_OK;
}

extern "C" qmc_status qmc_scene_check_heliostat(const qmc_scene *s, int32_t h) {
    if (!s) return fail(QMC_E_INVALID_ARG, "null argument");
    // QuasiMonteCarloRayTracer::checkValidHeliostatIndex (QuasiMonteCarloRayTracer.cpp:96-127)
    std::string msg = "The ray tracing heliostat index " + std::to_string(h);
    const std::string suffix = " is invalid.";
    size_t nh = s->heliostats.size();
    if (h < 0 || h >= (int)nh)
        return fail(QMC_E_RANGE, msg + suffix + "The heliostat index should between 0 and " + std::to_string(nh) + ".");
    int g = s->heliostats[h].d.grid;
    msg += "with belonging grid index " + std::to_string(g);
    size_t ng = s->grids.size();
    if (g < 0 || g >= (int)ng)
        return fail(QMC_E_RANGE, msg + suffix + "The grid index should between 0 and " + std::to_string(ng) + ".");
    int r = s->grids[g].d.receiver;
    msg += "of belonging receiver index " + std::to_string(r);
    size_t nr = s->receivers.size();
    if (r < 0 || r >= (int)nr)
        return fail(QMC_E_RANGE, msg + suffix + "The receiver index should between 0 and " + std::to_string(nr) + ".");
    return QMC_OK;
}

// ---------------------------------------------------------------------------------------------
// Task file (RayTracingPipeline/TaskLoader.cpp:114-204)
// ---------------------------------------------------------------------------------------------
namespace {
// A character cursor that mimics `stream >> std::string` on a fully buffered stringstream:
// available() is the number of unread characters (rdbuf()->in_avail()), and a failed extraction
// leaves the target unchanged.
struct Cursor {
    const std::string &s;
    size_t i = 0;
    explicit Cursor(const std::string &str) : s(str) {}
    size_t available() const { return s.size() - i; }
    bool next(std::string &tok) {
        size_t j = i;
        while (j < s.size() && std::isspace((unsigned char)s[j])) ++j;
        if (j >= s.size()) {
            i = j;
            return false;
        }
        size_t k = j;
        while (k < s.size() && !std::isspace((unsigned char)s[k])) ++k;
        tok = s.substr(j, k - j);
        i = k;
        return true;
    }
};

bool parse_int_full(const std::string &tok, int &v) {
    // std::stoi semantics: leading whitespace skipped, optional sign, digits; reports how many
    // characters were used; throws when no digits.
    size_t pos = 0;
    try {
        v = std::stoi(tok, &pos);
    } catch (...) {
        return false;
    }
    return pos >= tok.size();
}
}  // namespace

extern "C" qmc_status qmc_task_load(const qmc_scene *s, const char *path, int32_t *ids, int32_t cap,
                                    int32_t *n, int32_t *clean) {
    if (!s || !path) return fail(QMC_E_INVALID_ARG, "qmc_task_load: null argument");
    std::ifstream f(path);
    if (f.fail()) return fail(QMC_E_IO, std::string("Cannot open the file '") + path + "'.");
    std::stringstream ss;
    ss << f.rdbuf();
    std::string text = ss.str();
    Cursor cur(text);
    std::string tok;
    cur.next(tok);
    int N = 0;
    if (!parse_int_full(tok, N) || N <= 0)
        return fail(QMC_E_PARSE, "The number of heliostats '" + tok + "' is invalid.\nPlease check the 1st element in the file '" +
                                     path + "'.");
    std::vector<int> order;
    std::set<int> seen;
    bool ok = true;
    const int nh = (int)s->heliostats.size();
    for (int i = 0; i < N; ++i) {
        if (cur.available() == 0)
            return fail(QMC_E_PARSE, std::to_string(N) + " heliostat(s) are expected. But only got " + std::to_string(i) +
                                         " heliostats.\nPlease check your file '" + path + "'.");
        cur.next(tok);
        int id = 0;
        if (!parse_int_full(tok, id) || id < 0 || id >= nh)
            return fail(QMC_E_PARSE, "The " + std::to_string(i + 1) + "-th heliostat index '" + tok +
                                         "' is invalid.\n Please check your file '" + path + "'.");
        if (seen.insert(id).second)
            order.push_back(id);
        else
            ok = false;
    }
    std::string left;
    if (cur.next(left)) ok = false;
    if (n) *n = (int)order.size();
    if (ids) {
        if (cap < (int)order.size()) return fail(QMC_E_INVALID_ARG, "qmc_task_load: id buffer too small");
        for (size_t k = 0; k < order.size(); ++k) ids[k] = order[k];
    }
    if (clean) *clean = ok ? 1 : 0;
    return QMC_OK;
}

// ---------------------------------------------------------------------------------------------
// Buie sun-shape table (Scene/SceneProcessor.cu:7-44,196-239)
// ---------------------------------------------------------------------------------------------
namespace qmc {
// cdf_xy: 2*(n+1) floats (theta in mrad, cumulative area); slopes: n floats;
// abcig: A, B, C, I (area below 4.65 mrad), gamma.
void sunshape_params(const qmc_config &c, std::vector<float> &cdf_xy, std::vector<float> &slopes, float abcig[5]) {
    const int n = c.inverse_transform_sampling_groups;
    const float csr = c.csr;
    const float upper = 4.65f;
    const float k = 0.9f * logf(13.5f * csr) * powf(csr, -0.3f);
    const float gamma = 2.2f * logf(0.52f * csr) * powf(csr, 0.43f) - 0.1f;
    auto tail = [&](float th) { return expf(k) / (gamma + 1) * powf(th, gamma + 1); };
    auto core = [&](float th) { return cosf(0.326f * th) / cosf(0.308f * th); };
    const float between = tail(9.3f) - tail(upper);
    const float len = upper / float(n);
    cdf_xy.assign(2 * (size_t)(n + 1), 0.0f);
    slopes.assign((size_t)n, 0.0f);
    float x = 0.0f;
    float prev = core(x);
    for (int i = 1; i <= n; ++i) {
        x += len;
        float cur = core(x);
        float sub = (cur + prev) / 2.0f * len;
        cdf_xy[2 * i] = x;
        cdf_xy[2 * i + 1] = sub + cdf_xy[2 * (i - 1) + 1];
        slopes[i - 1] = len / sub;
        prev = cur;
    }
    const float I = cdf_xy[2 * n + 1];
    abcig[0] = I + between;
    abcig[1] = (gamma + 1) / expf(k);
    abcig[2] = powf(upper, gamma + 1);
    abcig[3] = I;
    abcig[4] = gamma;
}

// Nodes and weights of n-point Gauss-Legendre quadrature on [-1, 1]: Newton's method on the Legendre
// polynomial P_n from the usual starting guesses cos(pi (i + 3/4) / (n + 1/2)).
static void gauss_legendre(int n, double *x, double *w) {
    for (int i = 0; i < n; ++i) {
        double z = std::cos(M_PI * (i + 0.75) / (n + 0.5)), dp = 1.0;
        for (int it = 0; it < 100; ++it) {
            double p0 = 1.0, p1 = 0.0;  // P_j(z), P_{j-1}(z)
            for (int j = 1; j <= n; ++j) {
                const double p2 = p1;
                p1 = p0;
                p0 = ((2.0 * j - 1.0) * z * p1 - (j - 1.0) * p2) / j;
            }
            dp = n * (z * p0 - p1) / (z * z - 1.0);
            const double dz = p0 / dp;
            z -= dz;
            if (std::fabs(dz) < 1e-15) break;
        }
        x[i] = z;
        w[i] = 2.0 / ((1.0 - z * z) * dp * dp);
    }
}

// Physical Buie sun shape (D. Buie, A. G. Monger, C. J. Dey, "Sunshape distributions for terrestrial
// solar simulations", Solar Energy 74 (2003) 113-122), theta in mrad: phi = cos(0.326 t)/cos(0.308 t)
// up to 4.65, e^k t^g from 4.65 to 43.6, k = 0.9 ln(13.5 x) x^-0.3, g = 2.2 ln(0.52 x) x^0.43 - 0.1
// (x = CSR, the reference's own formulas, Scene/SceneProcessor.cu:203-210, here in double). Angles are
// drawn over the sky with density phi(t) sin(t): 1024 equal bins over the core, 1024 geometric bins
// over the circumsolar region, each bin's mass by 8-point Gauss-Legendre (the profile is smooth inside
// each region), and the density at both bin ends for the in-bin inversion (table_inverse_cdf).
void buie_physical_table(const qmc_config &c, BuieTable &t) {
    constexpr int kCore = 1024, kTail = 1024, kGauss = 8;
    constexpr double kCoreEnd = 4.65, kTailEnd = 43.6;
    const double x = c.csr;
    const double kap = 0.9 * std::log(13.5 * x) * std::pow(x, -0.3);
    const double gam = 2.2 * std::log(0.52 * x) * std::pow(x, 0.43) - 0.1;
    auto core = [](double th) { return std::cos(0.326 * th) / std::cos(0.308 * th) * std::sin(th * 1e-3); };
    auto tail = [&](double th) { return std::exp(kap) * std::pow(th, gam) * std::sin(th * 1e-3); };
    double gx[kGauss], gw[kGauss];
    gauss_legendre(kGauss, gx, gw);
    const int n = kCore + kTail;
    t.theta.assign(n + 1, 0.0);
    t.cdf.assign(n + 1, 0.0);
    t.dens_lo.assign(n, 0.0);
    t.dens_hi.assign(n, 0.0);
    for (int i = 0; i <= kCore; ++i) t.theta[i] = kCoreEnd * i / kCore;
    for (int i = 1; i <= kTail; ++i) t.theta[kCore + i] = kCoreEnd * std::pow(kTailEnd / kCoreEnd, (double)i / kTail);
    t.theta[n] = kTailEnd;
    double total = 0.0;
    for (int i = 0; i < n; ++i) {
        const double a = t.theta[i], b = t.theta[i + 1], half = 0.5 * (b - a), mid = 0.5 * (a + b);
        const bool in_core = i < kCore;
        double mass = 0.0;
        for (int k = 0; k < kGauss; ++k) {
            const double th = mid + half * gx[k];
            mass += gw[k] * (in_core ? core(th) : tail(th));
        }
        total += mass * half;
        t.cdf[i + 1] = total;
        t.dens_lo[i] = in_core ? core(a) : tail(a);
        t.dens_hi[i] = in_core ? core(b) : tail(b);
    }
    for (int i = 1; i <= n; ++i) t.cdf[i] /= total;
    t.cdf[n] = 1.0;
}
}  // namespace qmc

extern "C" qmc_status qmc_sunshape_table_buie(const qmc_scene *s, double *theta, double *cdf, double *dens_lo,
                                              double *dens_hi, int32_t *n) {
    if (!s) return fail(QMC_E_INVALID_ARG, "qmc_sunshape_table_buie: null scene");
    if (!(s->cfg.csr > 0.0f)) return fail(QMC_E_INVALID_ARG, "qmc_sunshape_table_buie: need csr > 0");
    BuieTable t;
    buie_physical_table(s->cfg, t);
    const size_t nb = t.dens_lo.size();
    if (theta) std::memcpy(theta, t.theta.data(), (nb + 1) * sizeof(double));
    if (cdf) std::memcpy(cdf, t.cdf.data(), (nb + 1) * sizeof(double));
    if (dens_lo) std::memcpy(dens_lo, t.dens_lo.data(), nb * sizeof(double));
    if (dens_hi) std::memcpy(dens_hi, t.dens_hi.data(), nb * sizeof(double));
    if (n) *n = (int32_t)nb;
    return QMC_OK;
}

extern "C" qmc_status qmc_sunshape_table(const qmc_scene *s, float *cdf_xy, float *slopes, float *abc_i) {
    if (!s) return fail(QMC_E_INVALID_ARG, "qmc_sunshape_table: null scene");
    const qmc_config &c = s->cfg;
    if (c.inverse_transform_sampling_groups <= 0 || !(c.csr > 0.0f))
        return fail(QMC_E_INVALID_ARG, "qmc_sunshape_table: need inverse_transform_sampling_groups > 0 and csr > 0");
    std::vector<float> cdf, sl;
    float abcig[5];
    sunshape_params(c, cdf, sl, abcig);
    if (cdf_xy) std::memcpy(cdf_xy, cdf.data(), cdf.size() * sizeof(float));
    if (slopes) std::memcpy(slopes, sl.data(), sl.size() * sizeof(float));
    if (abc_i) std::memcpy(abc_i, abcig, sizeof(abcig));
    return QMC_OK;
}
