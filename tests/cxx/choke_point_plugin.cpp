// The choke-point plugin of plugins/ChokePointNew.cpp written against the
// reference's API on include/gpusteer_opensteer.hpp: CtfWorld's set-up
// (:101-118: lattice, wall map split at the cell borders) and
// CtfGroup::update (:587-630) as the per-frame sequence.  The agents come
// from a file (their placement is the caller's), so the GPU test can compare
// the result with the oracle:
//
//   choke_point_plugin PARAMS MAPDIR AGENTS STEPS DT OUT
//
// AGENTS (little-endian): uint32 n; then n AgentData records as float arrays
//     px,py,pz, fx,fy,fz, speed, sx,sy,sz, ux,uy,uz, radius, max_speed,
//     max_force, mass.
// OUT: n x (px,py,pz, fx,fy,fz, speed, sx,sy,sz, ux,uy,uz) floats in
//     agent-id order.
// Every 30th frame the plugin's extractData (:413-435) writes
// Frames/cell_density_frame_<f>.dat and Frames/cell_avg_velocity_frame_<f>.dat
// and appends the selected cells' values (relative to the working directory).
#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

#include "gpusteer_opensteer.hpp"

using namespace OpenSteer;

namespace {
template <class T> bool rd(FILE* f, T* p, size_t n) { return fread(p, sizeof(T), n, f) == n; }
}  // namespace

int main(int argc, char** argv) {
    if (argc != 7) {
        fprintf(stderr, "usage: %s PARAMS MAPDIR AGENTS STEPS DT OUT\n", argv[0]);
        return 2;
    }
    FILE* in = fopen(argv[3], "rb");
    if (!in) {
        perror(argv[3]);
        return 2;
    }
    uint32_t n = 0;
    bool ok = rd(in, &n, 1);
    std::vector<float> f(17 * (size_t)n);
    ok = ok && rd(in, f.data(), f.size());
    fclose(in);
    if (!ok) {
        fprintf(stderr, "short input\n");
        return 2;
    }
    const int steps = atoi(argv[4]);
    const float elapsedTime = (float)atof(argv[5]);
    try {
        SimulationParams sp;
        sp.load(argv[1]);
        WorldParams& wp = sp.m_WorldParams;
        GroupParams& gp = sp.m_vecGroupParams.at(0);

        // CtfWorld
        KNNBinData bins(wp.m_u3Cells, wp.m_f3Dimensions, sp.m_nSearchRadius);
        WallGroup walls(wp.m_u3Cells, sp.m_nKNW);
        const std::string map = std::string(argv[2]) + "/" + wp.m_strMapFilename;
        if (!walls.LoadFromFile(map.c_str())) {
            fprintf(stderr, "cannot load %s\n", map.c_str());
            return 1;
        }
        walls.SplitWalls(bins.hvCells());
        walls.SyncDevice();
        updateKNNDatabase(&walls, &bins);

        // CtfGroup
        AgentGroup group(wp.m_u3Cells, sp.m_nKNN);
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
            if (!group.AddAgent(ad)) throw gpusteer::Error(GS_ERR_INVALID_ARGUMENT, "duplicate id");
        }
        KNNData knnSelf(n, sp.m_nKNN);
        KNNData knnWalls(n, sp.m_nKNW);
        PolylinePathwayCUDA path(gp.m_vecPathPoints, gp.m_fPathRadius, gp.m_bPathIsCyclic);

        // CtfSimulation's selected cells (plugins/ChokePointNew.cpp:342-350).
        const std::vector<uint> cells = {16512, 31872, 32896, 33920, 34944, 35968, 36992, 38016};
        uint frame = 1;
        for (int s = 0; s < steps; ++s) {
            wrapWorld(&group, wp.m_f3Dimensions);
            updateKNNDatabase(&group, &bins);
            // the reference's cudaMemset of pdAppliedKernels
            gpusteer::check(gs_clear_bits(group.device_handle()), "gs_clear_bits");
            findKNearestNeighbors(&group, &knnSelf, &bins, &group, sp.m_nSearchRadius);
            findKNearestNeighbors(&group, &knnWalls, &bins, &walls, sp.m_nSearchRadius);
            steerToAvoidWalls(&group, &knnWalls, &walls, sp.m_fMinTimeToWall, sp.m_fWeightAvoidWalls,
                              sp.m_nMaskAvoidWalls);
            steerForSeparation(&group, &knnSelf, &group, sp.m_fMinSeparationDistance, sp.m_fMaxSeparationDistance,
                               sp.m_fCosMaxFlockingAngle, sp.m_fWeightSeparation, GS_BIT_ANTI_PENETRATION_WALL);
            steerForAlignment(&group, &knnSelf, &group, sp.m_fMinFlockingDistance, sp.m_fMaxFlockingDistance,
                              sp.m_fCosMaxFlockingAngle, sp.m_fWeightAlignment, sp.m_nMaskAlignment);
            steerForCohesion(&group, &knnSelf, &group, sp.m_fMinFlockingDistance, sp.m_fMaxFlockingDistance,
                             sp.m_fCosMaxFlockingAngle, sp.m_fWeightCohesion, sp.m_nMaskCohesion);
            steerToFollowPath(&group, &path, sp.m_fPathPredictionTime, sp.m_fWeightFollowPath, sp.m_nMaskFollowPath);
            steerForSeek(&group, path.hvPoints()[1], sp.m_fWeightSeek, sp.m_nMaskSeek);
            updateGroup(&group, elapsedTime);
            antiPenetrationAgents(&group, &knnSelf, &group, sp.m_nMaskAntiPenetrationAgents);
            // extractData
            if (frame % 30 == 0) {
                char name[100];
                snprintf(name, sizeof name, "Frames/cell_density_frame_%u", frame);
                WriteCellDensity(name, &group, cells);
                snprintf(name, sizeof name, "Frames/cell_avg_velocity_frame_%u", frame);
                WriteAvgCellVelocity(name, &group, cells);
            }
            ++frame;
        }

        FILE* out = fopen(argv[6], "wb");
        if (!out) {
            perror(argv[6]);
            return 2;
        }
        for (uint32_t i = 0; i < n; ++i) {
            AgentData ad;
            if (!group.GetDataForAgent(i, ad)) throw gpusteer::Error(GS_ERR_INVALID_ARGUMENT, "missing agent");
            const float rec[13] = {ad.position.x, ad.position.y, ad.position.z, ad.direction.x, ad.direction.y,
                                   ad.direction.z, ad.speed, ad.side.x, ad.side.y, ad.side.z,
                                   ad.up.x, ad.up.y, ad.up.z};
            fwrite(rec, sizeof(float), 13, out);
        }
        fclose(out);
    } catch (const gpusteer::Error& e) {
        fprintf(stderr, "gpusteer error %d: %s\n", e.status(), e.what());
        return 1;
    }
    return 0;
}
