// A Boids plugin written against the reference's group API (the per-frame
// sequence of plugins/BoidsGroup.cpp:606-637: wrapWorld, updateKNNDatabase,
// findKNearestNeighbors, separation, cohesion, alignment, updateGroup), built
// on include/gpusteer_opensteer.hpp.  The GPU test drives it through files:
//
//   boids_group_plugin IN OUT STEPS MODE
//
// IN  (little-endian): uint32 n, k, cells[3], search_radius; float
//     world_size[3], dt, sep[4], coh[4], ali[4] (min_d, max_d, cos_max, weight);
//     then n AgentData records as float arrays px,py,pz, fx,fy,fz, speed,
//     sx,sy,sz, ux,uy,uz, radius, max_speed, max_force, mass.
// OUT: n x (px,py,pz, fx,fy,fz, speed) floats in agent-id order, then the
//     last KNN result: uint32 idx[k][n].
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "gpusteer_opensteer.hpp"

using namespace OpenSteer;

namespace {
template <class T> bool rd(FILE* f, T* p, size_t n) { return fread(p, sizeof(T), n, f) == n; }
}  // namespace

int main(int argc, char** argv) {
    if (argc != 5) {
        fprintf(stderr, "usage: %s IN OUT STEPS MODE\n", argv[0]);
        return 2;
    }
    FILE* in = fopen(argv[1], "rb");
    if (!in) {
        perror(argv[1]);
        return 2;
    }
    uint32_t n, k, cells[3], sr;
    float world[3], dt, sep[4], coh[4], ali[4];
    bool ok = rd(in, &n, 1) && rd(in, &k, 1) && rd(in, cells, 3) && rd(in, &sr, 1) && rd(in, world, 3) &&
              rd(in, &dt, 1) && rd(in, sep, 4) && rd(in, coh, 4) && rd(in, ali, 4);
    std::vector<float> f(17 * (size_t)n);
    ok = ok && rd(in, f.data(), f.size());
    fclose(in);
    if (!ok) {
        fprintf(stderr, "short input\n");
        return 2;
    }
    const int steps = atoi(argv[3]);
    try {
        gpusteer::Context::Default().SetMode(atoi(argv[4]));
        const uint3 wc = make_uint3(cells[0], cells[1], cells[2]);
        AgentGroup boids(wc, k);
        auto F = [&](int j, uint32_t i) { return f[(size_t)j * n + i]; };
        for (uint32_t i = 0; i < n; ++i) {
            AgentData ad;
            ad.id = i;
            ad.position = make_float4(F(0, i), F(1, i), F(2, i), 0.f);
            ad.direction = make_float4(F(3, i), F(4, i), F(5, i), 0.f);
            ad.speed = F(6, i);
            ad.side = make_float3(F(7, i), F(8, i), F(9, i));
            ad.up = make_float3(F(10, i), F(11, i), F(12, i));
            ad.radius = F(13, i);
            ad.maxSpeed = F(14, i);
            ad.maxForce = F(15, i);
            ad.mass = F(16, i);
            if (!boids.AddAgent(ad)) throw gpusteer::Error(GS_ERR_INVALID_ARGUMENT, "duplicate id");
        }
        KNNBinData bins(wc, make_float3(world[0], world[1], world[2]), sr);
        KNNData knnSelf(n, k);
        const float3 worldSize = make_float3(world[0], world[1], world[2]);
        for (int s = 0; s < steps; ++s) {
            wrapWorld(&boids, worldSize);
            updateKNNDatabase(&boids, &bins);
            findKNearestNeighbors(&boids, &knnSelf, &bins, &boids, sr);
            steerForSeparation(&boids, &knnSelf, &boids, sep[0], sep[1], sep[2], sep[3], 0);
            steerForCohesion(&boids, &knnSelf, &boids, coh[0], coh[1], coh[2], coh[3], 0);
            steerForAlignment(&boids, &knnSelf, &boids, ali[0], ali[1], ali[2], ali[3], 0);
            updateGroup(&boids, dt);
        }
        FILE* out = fopen(argv[2], "wb");
        if (!out) {
            perror(argv[2]);
            return 2;
        }
        for (uint32_t i = 0; i < n; ++i) {
            AgentData ad;
            if (!boids.GetDataForAgent(i, ad)) throw gpusteer::Error(GS_ERR_INVALID_ARGUMENT, "missing agent");
            const float rec[7] = {ad.position.x, ad.position.y, ad.position.z, ad.direction.x,
                                  ad.direction.y, ad.direction.z, ad.speed};
            fwrite(rec, sizeof(float), 7, out);
        }
        std::vector<uint> idx(k);
        std::vector<uint32_t> all((size_t)k * n);
        for (uint32_t i = 0; i < n; ++i) {
            knnSelf.getAgentData(i, idx.data(), nullptr);
            for (uint32_t j = 0; j < k; ++j) all[(size_t)j * n + i] = idx[j];
        }
        fwrite(all.data(), sizeof(uint32_t), all.size(), out);
        fclose(out);
    } catch (const gpusteer::Error& e) {
        fprintf(stderr, "gpusteer error %d: %s\n", e.status(), e.what());
        return 1;
    }
    return 0;
}
