// Scenario formats, host only (no device work): the wall-map loader, the
// splitting of walls at bin-cell borders, and the simulation parameter file.
//
// Walls.  A .map file holds one wall per record: start xyz, end xyz, normal
// xyz, whitespace separated (data/ChokePoint15.map).  Records are read until
// one is incomplete; a record whose last value ends the file without a line
// break is dropped as well (WallGroupData::LoadFromFile reads with an
// istream and keeps a record only while the stream is still good,
// include/OpenSteer/CUDA/WallGroupData.cu:73-103).  The KNN position of a
// wall is its midpoint 0.5f * (start + end).
//
// Splitting (WallGroupData::SplitWalls, WallGroupData.cu:105-206): walls are
// kept in a list; for the wall at the cursor, the cells of the lattice are
// tried in the lattice's cell order (y outer, z, x inner, the order of
// KNNBinDataV2::CreateCells, CUDA/KNNBinDataV2.cu:73-136); the first cell
// whose entry point splits the wall into two parts both longer than 0.001
// replaces the wall by those two parts, appended at the back of the list
// (their midpoints start + 0.5f * (entry - start) and entry + 0.5f * (end -
// entry)), and the cursor moves to the next wall; a wall no cell splits stays.
// The resulting order and midpoints decide the wall KNN's tie-breaking, so
// the arithmetic follows the reference operation for operation (fp32, no
// contraction).
//
// Parameters (SimulationParams::load, src/Simulation.cpp:7-250): a stream of
// whitespace-separated tokens, `key = value` pairs (the token after the key is
// skipped), a token starting with '#' skips the rest of its line, blocks
// <world> ... </world> and <group> ... </group> (with <path> ... </path>
// inside a group), up to </simulation>.  Unknown keys are an error (the
// reference asserts), as are a missing file, a malformed number and a
// missing terminator (on which the reference's loop does not end).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <list>
#include <sstream>
#include <string>
#include <vector>

#include "gs_internal.hpp"

struct gs_wall_set_s {
    std::vector<float> v[12];  // start xyz, end xyz, normal xyz, mid xyz
    size_t size() const { return v[0].size(); }
};

namespace gs {
namespace {

struct W {
    float s[3], e[3], n[3], m[3];
};

gs_wall_set_s* check_set(gs_wall_set w) {
    if (!reg_has(w)) fail(GS_ERR_INVALID_HANDLE, "wall set handle is invalid");
    return w;
}

std::vector<W> to_list(const gs_wall_set_s* w) {
    std::vector<W> out(w->size());
    for (size_t i = 0; i < out.size(); ++i)
        for (int a = 0; a < 3; ++a) {
            out[i].s[a] = w->v[a][i];
            out[i].e[a] = w->v[3 + a][i];
            out[i].n[a] = w->v[6 + a][i];
            out[i].m[a] = w->v[9 + a][i];
        }
    return out;
}

template <class It> void from_list(gs_wall_set_s* w, It b, It e) {
    for (auto& v : w->v) v.clear();
    for (It i = b; i != e; ++i)
        for (int a = 0; a < 3; ++a) {
            w->v[a].push_back(i->s[a]);
            w->v[3 + a].push_back(i->e[a]);
            w->v[6 + a].push_back(i->n[a]);
            w->v[9 + a].push_back(i->m[a]);
        }
}

// Segment start->end against the box [lo, hi]: slab clipping; on a hit the
// entry point start + (end - start) * t_entry.
bool segment_enters(const float* s, const float* e, const float* lo, const float* hi, float* ip) {
    float t0 = 0.f, t1 = 1.f;
    for (int a = 0; a < 3; ++a) {
        float ta, tb;
        const float d = e[a] - s[a];
        if (s[a] < e[a]) {
            if (s[a] > hi[a] || e[a] < lo[a]) return false;
            ta = (s[a] < lo[a]) ? (lo[a] - s[a]) / d : 0.f;
            tb = (e[a] > hi[a]) ? (hi[a] - s[a]) / d : 1.f;
        } else {
            if (e[a] > hi[a] || s[a] < lo[a]) return false;
            ta = (s[a] > hi[a]) ? (hi[a] - s[a]) / d : 0.f;
            tb = (e[a] < lo[a]) ? (lo[a] - s[a]) / d : 1.f;
        }
        if (ta > t0) t0 = ta;
        if (tb < t1) t1 = tb;
        if (t1 < t0) return false;
    }
    for (int a = 0; a < 3; ++a) ip[a] = s[a] + (e[a] - s[a]) * t0;
    return true;
}

float dist3(const float* a, const float* b) {
    const float x = a[0] - b[0], y = a[1] - b[1], z = a[2] - b[2];
    return std::sqrt(((x * x) + (y * y)) + (z * z));
}

struct Cell {
    float lo[3], hi[3];
};

// Cells of a lattice of `cells` over a box of `size` centred at the origin,
// in the order y (outer), z, x (inner).
std::vector<Cell> lattice_cells(const float* size, const uint32_t* cells) {
    float step[3];
    for (int a = 0; a < 3; ++a) step[a] = size[a] / (float)cells[a];
    std::vector<Cell> out;
    out.reserve((size_t)cells[0] * cells[1] * cells[2]);
    for (uint32_t iy = 0; iy < cells[1]; ++iy)
        for (uint32_t iz = 0; iz < cells[2]; ++iz)
            for (uint32_t ix = 0; ix < cells[0]; ++ix) {
                Cell c;
                const uint32_t idx[3] = {ix, iy, iz};
                for (int a = 0; a < 3; ++a) {
                    c.lo[a] = (float)idx[a] * step[a] - 0.5f * size[a];
                    c.hi[a] = c.lo[a] + step[a];
                }
                out.push_back(c);
            }
    return out;
}

void split_walls(gs_wall_set_s* w, const std::vector<Cell>& cl) {
    const float eps = 0.001f;
    std::vector<W> init = to_list(w);
    std::list<W> L(init.begin(), init.end());
    for (auto it = L.begin(); it != L.end();) {
        bool split = false;
        for (const Cell& c : cl) {
            float ip[3];
            if (!segment_enters(it->s, it->e, c.lo, c.hi, ip)) continue;
            if (!(dist3(it->s, ip) > eps && dist3(ip, it->e) > eps)) continue;
            W a, b;
            for (int k = 0; k < 3; ++k) {
                a.s[k] = it->s[k];
                a.e[k] = ip[k];
                a.m[k] = it->s[k] + (ip[k] - it->s[k]) * 0.5f;
                b.s[k] = ip[k];
                b.e[k] = it->e[k];
                b.m[k] = ip[k] + (it->e[k] - ip[k]) * 0.5f;
                a.n[k] = b.n[k] = it->n[k];
            }
            L.push_back(a);
            L.push_back(b);
            it = L.erase(it);
            split = true;
            break;
        }
        if (!split) ++it;
    }
    from_list(w, L.begin(), L.end());
}

// ---- parameter file -------------------------------------------------------

struct Tokens {
    std::string text;
    size_t at = 0;
    std::string path;
    bool next(std::string& tok) {
        for (;;) {
            while (at < text.size() && std::isspace((unsigned char)text[at])) ++at;
            if (at >= text.size()) return false;
            const size_t b = at;
            while (at < text.size() && !std::isspace((unsigned char)text[at])) ++at;
            tok = text.substr(b, at - b);
            if (tok[0] != '#') return true;
            // A comment: the rest of the line goes.
            while (at < text.size() && text[at] != '\n') ++at;
        }
    }
    std::string must(const char* where) {
        std::string t;
        if (!next(t)) fail(GS_ERR_INVALID_ARGUMENT, path + ": unexpected end of file in " + where);
        return t;
    }
    float f(const char* key) {
        const std::string t = must(key);
        char* end = nullptr;
        const float v = std::strtof(t.c_str(), &end);
        if (end == t.c_str() || *end) fail(GS_ERR_INVALID_ARGUMENT, path + ": bad number '" + t + "' for " + key);
        return v;
    }
    uint32_t u(const char* key) {
        const std::string t = must(key);
        char* end = nullptr;
        const unsigned long v = std::strtoul(t.c_str(), &end, 10);
        if (end == t.c_str() || *end || t[0] == '-')
            fail(GS_ERR_INVALID_ARGUMENT, path + ": bad unsigned integer '" + t + "' for " + key);
        return (uint32_t)v;
    }
    void f3(const char* key, float* v) {
        for (int a = 0; a < 3; ++a) v[a] = f(key);
    }
};

bool boolean(Tokens& T, const char* key) {
    const uint32_t v = T.u(key);
    if (v > 1) fail(GS_ERR_INVALID_ARGUMENT, T.path + ": " + key + " must be 0 or 1");
    return v != 0;
}

void load_group(Tokens& T, gs_group_params& g) {
    std::string key = T.must("<group>");
    while (key != "</group>") {
        if (key == "<path>") {
            // (The reference reads the block's first key as the delimiter of
            // "<path>" and then uses it as a key: the same tokens.)
            std::string pk = T.must("<path>");
            while (pk != "</path>") {
                T.must(pk.c_str());  // the delimiter
                if (pk == "point") {
                    if (g.path_points >= GS_PARAMS_MAX_PATH)
                        fail(GS_ERR_UNSUPPORTED, T.path + ": more than GS_PARAMS_MAX_PATH path points");
                    T.f3("point", g.path[g.path_points]);
                    ++g.path_points;
                } else if (pk == "cyclic") {
                    g.path_cyclic = boolean(T, "cyclic") ? 1u : 0u;
                } else if (pk == "radius") {
                    g.path_radius = T.f("radius");
                } else {
                    fail(GS_ERR_INVALID_ARGUMENT, T.path + ": unknown path key '" + pk + "'");
                }
                pk = T.must("<path>");
            }
        } else {
            T.must(key.c_str());  // the delimiter
            if (key == "numAgents") g.num_agents = T.u("numAgents");
            else if (key == "startPosition") T.f3("startPosition", g.start_position);
            else if (key == "minStartRadius") g.min_start_radius = T.f("minStartRadius");
            else if (key == "maxStartRadius") g.max_start_radius = T.f("maxStartRadius");
            else if (key == "maxSpeed") g.max_speed = T.f("maxSpeed");
            else if (key == "maxForce") g.max_force = T.f("maxForce");
            else if (key == "bodyColor") T.f3("bodyColor", g.body_color);
            else fail(GS_ERR_INVALID_ARGUMENT, T.path + ": unknown group key '" + key + "'");
        }
        key = T.must("<group>");
    }
}

void load_world(Tokens& T, gs_world_params& w) {
    std::string key = T.must("<world>");
    while (key != "</world>") {
        T.must(key.c_str());  // the delimiter
        if (key == "dimensions") T.f3("dimensions", w.dimensions);
        else if (key == "cells") {
            for (int a = 0; a < 3; ++a) w.cells[a] = T.u("cells");
        } else if (key == "obstacleCount") w.obstacle_count = T.u("obstacleCount");
        else if (key == "mapFilename") {
            const std::string f = T.must("mapFilename");
            if (f.size() >= sizeof(w.map_filename)) fail(GS_ERR_UNSUPPORTED, T.path + ": map file name too long");
            std::memcpy(w.map_filename, f.c_str(), f.size() + 1);
        } else {
            fail(GS_ERR_INVALID_ARGUMENT, T.path + ": unknown world key '" + key + "'");
        }
        key = T.must("<world>");
    }
}

void load_params(const char* path, gs_sim_params& P) {
    std::ifstream in(path, std::ios::binary);
    if (!in.is_open()) fail(GS_ERR_INVALID_ARGUMENT, std::string("cannot open parameter file ") + path);
    std::ostringstream ss;
    ss << in.rdbuf();
    Tokens T;
    T.text = ss.str();
    T.path = path;
    std::memset(&P, 0, sizeof P);
    struct FKey {
        const char* k;
        float* v;
    };
    struct UKey {
        const char* k;
        uint32_t* v;
    };
    const FKey fkeys[] = {{"maxPursuitPredictionTime", &P.max_pursuit_prediction_time},
                          {"minTimeToCollision", &P.min_time_to_collision},
                          {"minTimeToObstacle", &P.min_time_to_obstacle},
                          {"minTimeToWall", &P.min_time_to_wall},
                          {"pathPredictionTime", &P.path_prediction_time},
                          {"minSeparationDistance", &P.min_separation_distance},
                          {"maxSeparationDistance", &P.max_separation_distance},
                          {"minFlockingDistance", &P.min_flocking_distance},
                          {"maxFlockingDistance", &P.max_flocking_distance},
                          {"cosMaxFlockingAngle", &P.cos_max_flocking_angle},
                          {"weightAlignment", &P.weight_alignment},
                          {"weightCohesion", &P.weight_cohesion},
                          {"weightSeparation", &P.weight_separation},
                          {"weightPursuit", &P.weight_pursuit},
                          {"weightSeek", &P.weight_seek},
                          {"weightFollowPath", &P.weight_follow_path},
                          {"weightAvoidObstacles", &P.weight_avoid_obstacles},
                          {"weightAvoidWalls", &P.weight_avoid_walls},
                          {"weightAvoidNeighbors", &P.weight_avoid_neighbors}};
    const UKey ukeys[] = {{"seed", &P.seed},
                          {"knn", &P.knn},
                          {"kno", &P.kno},
                          {"knw", &P.knw},
                          {"searchRadius", &P.search_radius},
                          {"maskAlignment", &P.mask_alignment},
                          {"maskAntiPenetrationAgents", &P.mask_anti_penetration_agents},
                          {"maskCohesion", &P.mask_cohesion},
                          {"maskSeparation", &P.mask_separation},
                          {"maskSeek", &P.mask_seek},
                          {"maskFlee", &P.mask_flee},
                          {"maskPursuit", &P.mask_pursuit},
                          {"maskEvade", &P.mask_evade},
                          {"maskFollowPath", &P.mask_follow_path},
                          {"maskAvoidObstacles", &P.mask_avoid_obstacles},
                          {"maskAvoidWalls", &P.mask_avoid_walls},
                          {"maskAvoidNeighbors", &P.mask_avoid_neighbors}};
    std::string key = T.must("<simulation>");
    while (key != "</simulation>") {
        if (key == "<world>") {
            load_world(T, P.world);
        } else if (key == "<group>") {
            if (P.n_groups >= GS_PARAMS_MAX_GROUPS) fail(GS_ERR_UNSUPPORTED, T.path + ": more than GS_PARAMS_MAX_GROUPS groups");
            load_group(T, P.groups[P.n_groups]);
            ++P.n_groups;
        } else {
            T.must(key.c_str());  // the delimiter
            bool known = false;
            for (const FKey& f : fkeys)
                if (key == f.k) {
                    *f.v = T.f(f.k);
                    known = true;
                    break;
                }
            for (const UKey& u : ukeys)
                if (!known && key == u.k) {
                    *u.v = T.u(u.k);
                    known = true;
                    break;
                }
            if (!known && key == "avoidCloseNeighbors") {
                P.avoid_close_neighbors = boolean(T, "avoidCloseNeighbors") ? 1u : 0u;
                known = true;
            }
            if (!known) fail(GS_ERR_INVALID_ARGUMENT, T.path + ": unknown key '" + key + "'");
        }
        key = T.must("<simulation>");
    }
}

// ---- .map -----------------------------------------------------------------

void load_map(const char* path, gs_wall_set_s* w) {
    std::ifstream in(path, std::ios::binary);
    if (!in.is_open()) fail(GS_ERR_INVALID_ARGUMENT, std::string("cannot open map file ") + path);
    std::ostringstream ss;
    ss << in.rdbuf();
    const std::string text = ss.str();
    for (auto& v : w->v) v.clear();
    size_t at = 0;
    for (;;) {
        float r[9];
        bool complete = true, at_eof = false;
        for (int j = 0; j < 9 && complete; ++j) {
            while (at < text.size() && std::isspace((unsigned char)text[at])) ++at;
            if (at >= text.size()) {
                complete = false;
                break;
            }
            const char* b = text.c_str() + at;
            char* e = nullptr;
            r[j] = std::strtof(b, &e);
            if (e == b) {
                complete = false;
                break;
            }
            at += (size_t)(e - b);
            if (j == 8) at_eof = at >= text.size();
        }
        // A record is kept only if it is complete and the stream did not
        // reach the end of the file on its last value.
        if (!complete || at_eof) break;
        for (int a = 0; a < 3; ++a) {
            w->v[a].push_back(r[a]);
            w->v[3 + a].push_back(r[3 + a]);
            w->v[6 + a].push_back(r[6 + a]);
            w->v[9 + a].push_back(0.5f * (r[a] + r[3 + a]));
        }
    }
}

}  // namespace
}  // namespace gs

using namespace gs;

extern "C" {

int gs_wall_set_load_map(const char* path, gs_wall_set* out) {
    return guarded([&] {
        if (!path || !out) fail(GS_ERR_INVALID_ARGUMENT, "NULL argument");
        auto* w = new gs_wall_set_s;
        try {
            load_map(path, w);
        } catch (...) {
            delete w;
            throw;
        }
        reg_add(w);
        *out = w;
    });
}

int gs_wall_set_create(uint32_t n, const gs_walls* host, gs_wall_set* out) {
    return guarded([&] {
        if (!out || (n > 0 && !host)) fail(GS_ERR_INVALID_ARGUMENT, "NULL argument");
        const float* const* hw = host ? (const float* const*)host : nullptr;
        for (int j = 0; j < 9 && n; ++j)
            if (!hw[j]) fail(GS_ERR_INVALID_ARGUMENT, "wall array is NULL");
        auto* w = new gs_wall_set_s;
        for (int j = 0; j < 9; ++j) w->v[j].assign(n ? hw[j] : nullptr, n ? hw[j] + n : nullptr);
        for (int a = 0; a < 3; ++a) {
            w->v[9 + a].resize(n);
            for (uint32_t i = 0; i < n; ++i) w->v[9 + a][i] = 0.5f * (hw[a][i] + hw[3 + a][i]);
        }
        reg_add(w);
        *out = w;
    });
}

int gs_wall_set_split(gs_wall_set w, const float world_size[3], const uint32_t cells[3]) {
    return guarded([&] {
        check_set(w);
        if (!world_size || !cells) fail(GS_ERR_INVALID_ARGUMENT, "NULL argument");
        for (int a = 0; a < 3; ++a) {
            if (!(world_size[a] > 0.f) || !std::isfinite(world_size[a]) || cells[a] == 0)
                fail(GS_ERR_INVALID_ARGUMENT, "lattice size and cells must be positive");
        }
        if ((uint64_t)cells[0] * cells[1] * cells[2] > GS_MAX_SPLIT_CELLS)
            fail(GS_ERR_UNSUPPORTED, "lattice larger than GS_MAX_SPLIT_CELLS (2^24) cells for a wall split");
        split_walls(w, lattice_cells(world_size, cells));
    });
}

int gs_wall_set_split_cells(gs_wall_set w, uint32_t n_cells, const float* min_bounds, const float* max_bounds) {
    return guarded([&] {
        check_set(w);
        if (n_cells > 0 && (!min_bounds || !max_bounds)) fail(GS_ERR_INVALID_ARGUMENT, "NULL argument");
        std::vector<Cell> cl(n_cells);
        for (uint32_t i = 0; i < n_cells; ++i)
            for (int a = 0; a < 3; ++a) {
                cl[i].lo[a] = min_bounds[3 * (size_t)i + a];
                cl[i].hi[a] = max_bounds[3 * (size_t)i + a];
            }
        split_walls(w, cl);
    });
}

int gs_wall_set_set_mid(gs_wall_set w, const float* mid_x, const float* mid_y, const float* mid_z) {
    return guarded([&] {
        check_set(w);
        const size_t n = w->size();
        if (n && (!mid_x || !mid_y || !mid_z)) fail(GS_ERR_INVALID_ARGUMENT, "NULL argument");
        const float* m[3] = {mid_x, mid_y, mid_z};
        for (int a = 0; a < 3; ++a)
            if (n) std::memcpy(w->v[9 + a].data(), m[a], n * 4);
    });
}

int gs_wall_set_size(gs_wall_set w, uint32_t* n) {
    return guarded([&] {
        check_set(w);
        if (!n) fail(GS_ERR_INVALID_ARGUMENT, "out is NULL");
        *n = (uint32_t)w->size();
    });
}

int gs_wall_set_get(gs_wall_set w, gs_walls* out, float* mid_x, float* mid_y, float* mid_z) {
    return guarded([&] {
        check_set(w);
        const size_t n = w->size();
        if (out) {
            float* const* hw = (float* const*)out;
            for (int j = 0; j < 9; ++j)
                if (hw[j] && n) std::memcpy(hw[j], w->v[j].data(), n * 4);
        }
        float* m[3] = {mid_x, mid_y, mid_z};
        for (int a = 0; a < 3; ++a)
            if (m[a] && n) std::memcpy(m[a], w->v[9 + a].data(), n * 4);
    });
}

int gs_wall_set_destroy(gs_wall_set w) {
    return guarded([&] {
        if (!w) return;
        check_set(w);
        reg_del(w);
        delete w;
    });
}

int gs_sim_params_load(const char* path, gs_sim_params* out) {
    return guarded([&] {
        if (!path || !out) fail(GS_ERR_INVALID_ARGUMENT, "NULL argument");
        gs_sim_params p;
        load_params(path, p);
        *out = p;
    });
}

}  // extern "C"
