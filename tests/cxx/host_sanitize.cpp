// Host-code exercise for the AddressSanitizer / UndefinedBehaviorSanitizer
// build of the library's host side (device code not compiled, no GPU used):
// the scenario-file readers on every file of the data directory and on
// truncated and byte-mutated copies of them, wall splitting over the files'
// own lattices and over hand-made cell lists, wall sets from host arrays,
// and the argument and handle checks of every entry-point family (NULL,
// never-created and destroyed handles; the frame calls; a context without a
// device).  Any sanitizer report ends the process with a non-zero status.
//
//   host_sanitize DATADIR TMPDIR
#include <dirent.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "gpusteer.h"

namespace {

int g_checks = 0;
void expect(bool ok, const char* what) {
    ++g_checks;
    if (!ok) {
        fprintf(stderr, "FAILED: %s (last error: %s)\n", what, gs_last_error());
        exit(1);
    }
}

std::vector<std::string> files_with(const std::string& dir, const char* suffix) {
    std::vector<std::string> out;
    DIR* d = opendir(dir.c_str());
    if (!d) return out;
    while (dirent* e = readdir(d)) {
        const std::string n = e->d_name;
        const size_t m = strlen(suffix);
        if (n.size() > m && n.compare(n.size() - m, m, suffix) == 0) out.push_back(dir + "/" + n);
    }
    closedir(d);
    return out;
}

std::string slurp(const std::string& p) {
    FILE* f = fopen(p.c_str(), "rb");
    if (!f) return {};
    std::string s;
    char buf[4096];
    size_t k;
    while ((k = fread(buf, 1, sizeof buf, f)) > 0) s.append(buf, k);
    fclose(f);
    return s;
}

void spill(const std::string& p, const std::string& s) {
    FILE* f = fopen(p.c_str(), "wb");
    if (!f) {
        fprintf(stderr, "cannot write %s\n", p.c_str());
        exit(2);
    }
    fwrite(s.data(), 1, s.size(), f);
    fclose(f);
}

// Reads a wall set back through gs_wall_set_get into owned vectors.
void read_back(gs_wall_set w, uint32_t expect_n) {
    uint32_t n = 0;
    expect(gs_wall_set_size(w, &n) == GS_OK, "gs_wall_set_size");
    expect(expect_n == UINT32_MAX || n == expect_n, "wall count");
    std::vector<float> v(12 * (size_t)n + 1);
    gs_walls h;
    float** hp = (float**)&h;
    for (int j = 0; j < 9; ++j) hp[j] = v.data() + (size_t)j * n;
    expect(gs_wall_set_get(w, &h, v.data() + 9 * (size_t)n, v.data() + 10 * (size_t)n, v.data() + 11 * (size_t)n) ==
               GS_OK,
           "gs_wall_set_get");
    // Only some outputs requested.
    expect(gs_wall_set_get(w, nullptr, v.data(), nullptr, nullptr) == GS_OK, "gs_wall_set_get (mid x only)");
}

// Every map with every params file's lattice; a few fixed lattices too.
void maps_and_params(const std::string& dir) {
    const auto maps = files_with(dir, ".map");
    const auto params = files_with(dir, ".params");
    expect(!maps.empty() && !params.empty(), "data files present");
    // The .param file is another program's format: refused with a status.
    for (const auto& p : files_with(dir, ".param")) {
        gs_sim_params sp;
        expect(gs_sim_params_load(p.c_str(), &sp) == GS_ERR_INVALID_ARGUMENT, "foreign format refused");
    }
    std::vector<gs_sim_params> loaded;
    for (const auto& p : params) {
        gs_sim_params sp;
        const int st = gs_sim_params_load(p.c_str(), &sp);
        expect(st == GS_OK, ("gs_sim_params_load " + p).c_str());
        loaded.push_back(sp);
    }
    for (size_t mi = 0; mi < maps.size(); ++mi) {
        const std::string& m = maps[mi];
        gs_wall_set w = nullptr;
        expect(gs_wall_set_load_map(m.c_str(), &w) == GS_OK, ("gs_wall_set_load_map " + m).c_str());
        read_back(w, UINT32_MAX);
        const float ws[3] = {2048.f, 10.f, 2048.f};
        const uint32_t cells[3][3] = {{64, 1, 64}, {3, 2, 5}, {1, 1, 1}};
        for (const auto& c : cells) expect(gs_wall_set_split(w, ws, c) == GS_OK, "gs_wall_set_split");
        // Every params file's lattice over three of the maps.
        for (const auto& sp : loaded) {
            if (mi >= 3) break;
            const uint64_t C = (uint64_t)sp.world.cells[0] * sp.world.cells[1] * sp.world.cells[2];
            if (C == 0 || C > 300000) continue;
            gs_wall_set w2 = nullptr;
            expect(gs_wall_set_load_map(m.c_str(), &w2) == GS_OK, "reload");
            expect(gs_wall_set_split(w2, sp.world.dimensions, sp.world.cells) == GS_OK, "split at a file's lattice");
            read_back(w2, UINT32_MAX);
            expect(gs_wall_set_destroy(w2) == GS_OK, "destroy");
        }
        expect(gs_wall_set_destroy(w) == GS_OK, "gs_wall_set_destroy");
    }
}

// Truncations and byte mutations of every file: a status, never a crash.
void damaged_inputs(const std::string& dir, const std::string& tmp) {
    auto all = files_with(dir, ".map");
    for (const auto& p : files_with(dir, ".params")) all.push_back(p);
    uint64_t rng = 0x9e3779b97f4a7c15ull;
    auto next = [&]() {
        rng ^= rng << 13;
        rng ^= rng >> 7;
        rng ^= rng << 17;
        return rng;
    };
    int cases = 0, ok = 0;
    for (const auto& f : all) {
        const std::string s = slurp(f);
        const bool is_map = f.size() > 4 && f.compare(f.size() - 4, 4, ".map") == 0;
        const std::string out = tmp + (is_map ? "/damaged.map" : "/damaged.params");
        std::vector<std::string> variants;
        for (size_t cut : {(size_t)0, (size_t)1, s.size() / 7, s.size() / 2, s.size() - 1}) variants.push_back(s.substr(0, cut));
        for (int r = 0; r < 24; ++r) {
            std::string t = s;
            const int edits = 1 + (int)(next() % 4);
            for (int e = 0; e < edits && !t.empty(); ++e) {
                const size_t at = next() % t.size();
                const char pool[] = "0123456789.-+eE \n\t#<>/=abcxyz\x01\xff";
                t[at] = pool[next() % (sizeof pool - 1)];
            }
            variants.push_back(t);
        }
        // Huge numbers and deep nesting of the params grammar.
        if (!is_map) {
            variants.push_back("<simulation>\n<world>\n dimensions = 1e39 1e39 1e39\n</world>\n</simulation>\n");
            variants.push_back("<simulation>\n<group>\n<path>\n" + std::string(5000, '1') + "\n</path>\n</group>\n</simulation>\n");
            std::string many = "<simulation>\n";
            for (int g = 0; g < 40; ++g) many += "<group>\n numAgents = 1\n</group>\n";
            variants.push_back(many + "</simulation>\n");
            std::string path = "<simulation>\n<group>\n<path>\n";
            for (int q = 0; q < 80; ++q) path += "1 2 3\n";
            variants.push_back(path + "</path>\n</group>\n</simulation>\n");
        } else {
            variants.push_back("1 2 3 4 5 6 7 8 nan\n1e39 -1e39 0 0 0 0 0 0 1\n");
        }
        for (const auto& v : variants) {
            spill(out, v);
            ++cases;
            if (is_map) {
                gs_wall_set w = nullptr;
                const int st = gs_wall_set_load_map(out.c_str(), &w);
                if (st == GS_OK) {
                    ++ok;
                    const float ws[3] = {2048.f, 10.f, 2048.f};
                    const uint32_t c[3] = {16, 1, 16};
                    (void)gs_wall_set_split(w, ws, c);
                    read_back(w, UINT32_MAX);
                    expect(gs_wall_set_destroy(w) == GS_OK, "destroy damaged");
                } else {
                    expect(w == nullptr, "no set on failure");
                }
            } else {
                gs_sim_params sp;
                if (gs_sim_params_load(out.c_str(), &sp) == GS_OK) ++ok;
            }
        }
    }
    printf("damaged inputs: %d cases, %d accepted\n", cases, ok);
}

void wall_sets_from_arrays() {
    const uint32_t n = 5;
    std::vector<float> a(9 * n);
    for (uint32_t i = 0; i < n; ++i) {
        a[0 * n + i] = -10.f + 4.f * (float)i;  // start x
        a[2 * n + i] = -3.f;                    // start z
        a[3 * n + i] = a[0 * n + i] + 9.f;      // end x
        a[5 * n + i] = 7.f;                     // end z
        a[8 * n + i] = 1.f;                     // normal z
    }
    gs_walls h;
    float** hp = (float**)&h;
    for (int j = 0; j < 9; ++j) hp[j] = a.data() + (size_t)j * n;
    gs_wall_set w = nullptr;
    expect(gs_wall_set_create(n, &h, &w) == GS_OK, "gs_wall_set_create");
    std::vector<float> lo, hi;
    for (int cx = 0; cx < 4; ++cx)
        for (int cz = 0; cz < 3; ++cz) {
            lo.insert(lo.end(), {-12.f + 6.f * (float)cx, -1.f, -4.f + 4.f * (float)cz});
            hi.insert(hi.end(), {-6.f + 6.f * (float)cx, 1.f, 0.f + 4.f * (float)cz});
        }
    expect(gs_wall_set_split_cells(w, (uint32_t)(lo.size() / 3), lo.data(), hi.data()) == GS_OK, "split_cells");
    uint32_t m = 0;
    expect(gs_wall_set_size(w, &m) == GS_OK && m > n, "split added pieces");
    std::vector<float> mid(3 * (size_t)m, 1.f);
    expect(gs_wall_set_set_mid(w, mid.data(), mid.data() + m, mid.data() + 2 * m) == GS_OK, "set_mid");
    read_back(w, m);
    expect(gs_wall_set_split_cells(w, 0, nullptr, nullptr) == GS_OK, "split over no cells");
    expect(gs_wall_set_create(n, nullptr, &w) == GS_ERR_INVALID_ARGUMENT, "create without arrays");
    gs_wall_set empty = nullptr;
    expect(gs_wall_set_create(0, nullptr, &empty) == GS_OK, "empty set");
    read_back(empty, 0);
    expect(gs_wall_set_destroy(empty) == GS_OK, "destroy empty");
    const float bad_ws[3] = {0.f, 1.f, 1.f};
    const uint32_t c[3] = {1, 1, 1};
    expect(gs_wall_set_split(w, bad_ws, c) == GS_ERR_INVALID_ARGUMENT, "split with a zero size");
    const uint32_t huge[3] = {4096, 4096, 4096};
    const float ws[3] = {1.f, 1.f, 1.f};
    expect(gs_wall_set_split(w, ws, huge) == GS_ERR_UNSUPPORTED, "split over too many cells");
    expect(gs_wall_set_destroy(w) == GS_OK, "destroy");
    expect(gs_wall_set_size(w, &m) == GS_ERR_INVALID_HANDLE, "destroyed set rejected");
}

void handle_checks() {
    gs_context ctx = nullptr;
    expect(gs_context_create(0, nullptr, &ctx) == GS_ERR_DEVICE && ctx == nullptr, "no device, no context");
    void* bogus = (void*)0x1234;
    gs_group g = (gs_group)bogus;
    gs_knn k = (gs_knn)bogus;
    gs_bins b = (gs_bins)bogus;
    expect(gs_group_destroy(nullptr) == GS_ERR_INVALID_HANDLE, "group destroy NULL");
    expect(gs_group_destroy(g) == GS_ERR_INVALID_HANDLE, "group destroy bogus");
    expect(gs_update_knn_database(g, b) == GS_ERR_INVALID_HANDLE, "bins update bogus");
    expect(gs_find_knn(g, k, b, g, 1, 0) == GS_ERR_INVALID_HANDLE, "find bogus");
    expect(gs_steer_separation(g, k, g, 1.f, 2.f, 0.f, 1.f, 0, 0) == GS_ERR_INVALID_HANDLE, "separation bogus");
    expect(gs_update_group(g, 0.1f, 0) == GS_ERR_INVALID_HANDLE, "update bogus");
    expect(gs_knn_download(k, nullptr, nullptr) == GS_ERR_INVALID_HANDLE, "knn download bogus");
    expect(gs_step_boids(g, k, b, nullptr, 1) == GS_ERR_INVALID_HANDLE, "step bogus");
    gs_knn ks[8] = {k, k, k, k, k, k, k, k};
    expect(gs_step_choke(g, g, g, g, b, ks, nullptr, 1) == GS_ERR_INVALID_ARGUMENT, "choke without scenario");
    expect(gs_ensemble_step((gs_ensemble)bogus, 1, 0) == GS_ERR_INVALID_HANDLE, "ensemble bogus");
    expect(gs_calib_destroy((gs_calib)bogus) == GS_ERR_INVALID_HANDLE, "calib bogus");
    gs_frame f = nullptr;
    expect(gs_frame_begin(nullptr) == GS_ERR_INVALID_HANDLE, "frame begin NULL");
    expect(gs_frame_end(nullptr, &f) == GS_ERR_INVALID_HANDLE, "frame end NULL");
    expect(gs_frame_launch((gs_frame)bogus, 1) == GS_ERR_INVALID_HANDLE, "frame launch bogus");
    expect(gs_frame_destroy(nullptr) == GS_OK, "frame destroy NULL");
    expect(gs_wall_set_load_map(nullptr, nullptr) == GS_ERR_INVALID_ARGUMENT, "map NULL");
    expect(gs_wall_set_load_map("/nonexistent/file.map", nullptr) == GS_ERR_INVALID_ARGUMENT, "map out NULL");
    gs_wall_set w = nullptr;
    expect(gs_wall_set_load_map("/nonexistent/file.map", &w) != GS_OK && w == nullptr, "missing map");
    gs_sim_params sp;
    expect(gs_sim_params_load("/nonexistent/file.params", &sp) != GS_OK, "missing params");
    for (int s = 0; s <= GS_ERR_FRAME; ++s) expect(strlen(gs_status_string(s)) > 0, "status string");
    expect(gs_version() == GS_VERSION, "version");
}

}  // namespace

int main(int argc, char** argv) {
    if (argc != 3) {
        fprintf(stderr, "usage: %s DATADIR TMPDIR\n", argv[0]);
        return 2;
    }
    handle_checks();
    wall_sets_from_arrays();
    maps_and_params(argv[1]);
    damaged_inputs(argv[1], argv[2]);
    printf("host_sanitize: %d checks passed\n", g_checks);
    return 0;
}
