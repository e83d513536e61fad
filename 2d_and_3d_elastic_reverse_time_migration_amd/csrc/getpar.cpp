// getpar.cpp - parameter parsing compatible with the reference's CWP getpar
// usage in elastic_modeling_parameter_input()
// (3D_elastic_RTM/3D_elastic_modeling_parameter_sh:2-164).
//
// Accepted syntax (the subset of CWP getpar the reference relies on):
//   key=value            on the argument list
//   par=file             read further key=value pairs from `file`
//                        (whitespace separated, '#' starts a comment;
//                        values may be quoted with '...' or "...")
// The last occurrence of a key wins, as in getpar.  Keys the reference
// requires produce "<key> is not set!" (the reference's err() text) as an
// error return instead of exit.  The string keys s_velocity/s_velocity1/
// s_density are required only when the matching smooth_time_* is 0
// (parameter_sh:126-142).  Extension keys (fd_order, boundary, lt_forward,
// cpml_r, cpml_vmax, tile_x, tile_y, zchunk) are optional.
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <string>
#include <vector>

#include "common.hpp"

namespace erm {

namespace {

void tokenize(const std::string& text, std::vector<std::string>& out) {
    size_t i = 0, n = text.size();
    while (i < n) {
        while (i < n && (isspace((unsigned char)text[i]))) ++i;
        if (i >= n) break;
        if (text[i] == '#') {  // comment to end of line
            while (i < n && text[i] != '\n') ++i;
            continue;
        }
        std::string tok;
        while (i < n && !isspace((unsigned char)text[i])) {
            char ch = text[i];
            if (ch == '\'' || ch == '"') {
                char q = ch;
                ++i;
                while (i < n && text[i] != q) tok += text[i++];
                if (i < n) ++i;
            } else {
                tok += ch;
                ++i;
            }
        }
        out.push_back(tok);
    }
}

bool load_file(const std::string& path, std::vector<std::string>& toks, std::string* err) {
    std::ifstream f(path);
    if (!f) {
        if (err) *err = "can not open par file " + path;
        return false;
    }
    std::stringstream ss;
    ss << f.rdbuf();
    tokenize(ss.str(), toks);
    return true;
}

struct Table {
    std::map<std::string, std::string> kv;
    bool has(const char* k) const { return kv.count(k) != 0; }
    bool geti(const char* k, int* v) const {
        auto it = kv.find(k);
        if (it == kv.end()) return false;
        char* end = nullptr;
        long x = strtol(it->second.c_str(), &end, 10);
        if (end == it->second.c_str()) return false;
        *v = (int)x;
        return true;
    }
    bool getf(const char* k, float* v) const {
        auto it = kv.find(k);
        if (it == kv.end()) return false;
        char* end = nullptr;
        double x = strtod(it->second.c_str(), &end);
        if (end == it->second.c_str()) return false;
        *v = (float)x;
        return true;
    }
    bool gets(const char* k, char* dst) const {
        auto it = kv.find(k);
        if (it == kv.end()) return false;
        snprintf(dst, ERM_PATH_MAX, "%s", it->second.c_str());
        return true;
    }
};

}  // namespace

void params_default(erm_params* p) {
    memset(p, 0, sizeof(*p));
    p->fd_order = 12;
    p->boundary = ERM_BOUNDARY_DAMP;
    p->lt_forward = -1;  // -1: 2*lt/3 as the reference (resolved at create)
    p->cpml_r = 1e-3f;
    p->cpml_vmax = 0.f;
    p->GPU_N = 1;
    p->add_receiver_bool = 1;
    p->receiver_num_z = 1;
    p->shot_num_x = p->shot_num_y = p->shot_num_z = 1;
    p->xcd_map = 2;  // XCD y-band tile order: measured fastest at 296x296x168 with C-PML
}

int parse_par(int argc, const char* const* argv, erm_params* p, std::string* err) {
    std::vector<std::string> toks;
    for (int a = 0; a < argc; ++a) {
        if (!argv[a]) continue;
        std::string s = argv[a];
        if (s.rfind("par=", 0) == 0) {
            if (!load_file(s.substr(4), toks, err)) return ERM_E_IO;
        } else {
            toks.push_back(s);
        }
    }
    Table t;
    for (const std::string& tok : toks) {
        size_t e = tok.find('=');
        if (e == std::string::npos || e == 0) continue;  // getpar ignores non key=value words
        t.kv[tok.substr(0, e)] = tok.substr(e + 1);
    }
    params_default(p);
#define REQI(key, dst)                                   \
    if (!t.geti(key, &(dst))) {                          \
        if (err) *err = std::string(key) + " is not set!"; \
        return ERM_E_PARAM;                              \
    }
#define REQF(key, dst)                                   \
    if (!t.getf(key, &(dst))) {                          \
        if (err) *err = std::string(key) + " is not set!"; \
        return ERM_E_PARAM;                              \
    }
#define REQS(key, dst, msg)                              \
    if (!t.gets(key, dst)) {                             \
        if (err) *err = msg;                             \
        return ERM_E_PARAM;                              \
    }
    REQI("RTM", p->RTM);
    REQI("GPU_switch", p->GPU_switch);
    REQI("GPU_N", p->GPU_N);
    REQI("GPU_start", p->GPU_start);
    REQI("exchange_device_bool", p->exchange_device_bool);
    REQI("lt", p->lt);
    REQF("dt", p->dt);
    REQF("freq", p->freq);
    REQI("join_vs", p->join_vs);
    REQI("join_den", p->join_den);
    REQI("join_shot", p->join_shot);
    REQI("join_wavefield", p->join_wavefield);
    REQI("wavefield_interval", p->wavefield_interval);
    REQS("infile_shot_name", p->infile_shot_name, "infile_shot_name is not set!");
    REQI("nx", p->nx);
    REQI("ny", p->ny);
    REQI("nz", p->nz);
    REQF("dx", p->dx);
    REQF("dy", p->dy);
    REQF("dz", p->dz);
    REQI("shot_num_x", p->shot_num_x);
    REQI("shot_num_y", p->shot_num_y);
    REQI("shot_num_z", p->shot_num_z);
    REQI("shot_start_x", p->shot_start_x);
    REQI("shot_interval_x", p->shot_interval_x);
    REQI("shot_start_y", p->shot_start_y);
    REQI("shot_interval_y", p->shot_interval_y);
    REQI("shot_start_z", p->shot_start_z);
    REQI("shot_interval_z", p->shot_interval_z);
    REQI("receiver_num_x", p->receiver_num_x);
    REQI("receiver_num_y", p->receiver_num_y);
    REQI("receiver_num_z", p->receiver_num_z);
    REQI("receiver_start_x", p->receiver_start_x);
    REQI("receiver_interval_x", p->receiver_interval_x);
    REQI("receiver_start_y", p->receiver_start_y);
    REQI("receiver_interval_y", p->receiver_interval_y);
    REQI("receiver_start_z", p->receiver_start_z);
    REQI("receiver_interval_z", p->receiver_interval_z);
    REQI("receiver_offset", p->receiver_offset);
    REQI("receiver_mark", p->receiver_mark);
    REQI("boundary_up", p->bu);
    REQI("boundary_down", p->bd);
    REQI("boundary_left", p->bl);
    REQI("boundary_right", p->br);
    REQI("boundary_front", p->bf);
    REQI("boundary_back", p->bb);
    REQF("coe_attenuation", p->coe_attenuation);
    REQF("velocity_multiple", p->velocity_multiple);
    REQF("smooth_time_vp", p->smooth_time_vp);
    REQF("smooth_time_vs", p->smooth_time_vs);
    REQF("smooth_time_density", p->smooth_time_density);
    REQS("velocity", p->velocity, "can not read vp model!");
    REQS("velocity1", p->velocity1, "can not read vs model!");
    REQS("density", p->density, "can not read density model!");
    if (p->smooth_time_vp == 0) REQS("s_velocity", p->s_velocity, "can not read s_vp model!");
    if (p->smooth_time_vs == 0) REQS("s_velocity1", p->s_velocity1, "can not read s_vs model!");
    if (p->smooth_time_density == 0) REQS("s_density", p->s_density, "can not read s_density model!");
    REQI("add_receiver_bool", p->add_receiver_bool);
    REQI("cut_direct_wave", p->cut_direct_wave);
    REQI("laplace", p->laplace);
    REQI("laplace_compensate", p->laplace_compensate);
    REQI("vsp", p->vsp);
    REQI("vsp_precon", p->vsp_precon);
    REQI("precon", p->precon);
    REQI("precon_z1", p->precon_z1);
    REQI("precon_z2", p->precon_z2);
#undef REQI
#undef REQF
#undef REQS
    // extensions (optional)
    t.geti("fd_order", &p->fd_order);
    t.geti("boundary", &p->boundary);
    t.geti("lt_forward", &p->lt_forward);
    t.getf("cpml_r", &p->cpml_r);
    t.getf("cpml_vmax", &p->cpml_vmax);
    t.geti("tile_x", &p->tile_x);
    t.geti("tile_y", &p->tile_y);
    t.geti("zchunk", &p->zchunk);
    t.geti("xcd_map", &p->xcd_map);
    t.geti("poynting", &p->poynting);
    if (p->fd_order != 8 && p->fd_order != 12) {
        if (err) *err = "fd_order must be 8 or 12";
        return ERM_E_PARAM;
    }
    return ERM_OK;
}

}  // namespace erm
