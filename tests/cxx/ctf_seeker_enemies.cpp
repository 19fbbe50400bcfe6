// The choke-point plugin's seeker reading the enemy group through the
// reference's host views, on include/gpusteer_opensteer.hpp: the enemy loop
// of CtfSeeker::clearPathToGoal (plugins/ChokePointGroup.cpp:886-894) and
// CtfSeeker::steerToEvadeAllDefenders (:1012-1018) --
//
//     AgentGroupData & agd = gEnemies->GetAgentGroupData();
//     for( size_t i = 0; i < gEnemies->Size(); i++ ) { agd.getAgentData( i, aData ); ... }
//
// -- after the enemies have been stepped on the device, plus the wall
// group's host vectors (WallGroupData::hvLine*, plugins/ChokePointNew.cpp:182-185),
// the enemy group's KNN database host view (BaseGroup::GetKNNDatabase) and a
// plugin's clear of the applied-behaviour bits through pdAppliedKernels()
// (the cudaMemset of plugins/ChokePointGroup.cpp:691, here hipMemset).
//
//   ctf_seeker_enemies PARAMS MAPDIR AGENTS STEPS DT SEEKER_X SEEKER_Z OUT
//
// AGENTS (little-endian): uint32 n; then the 17 float arrays px,py,pz,
//     fx,fy,fz, speed, sx,sy,sz, ux,uy,uz, radius, max_speed, max_force, mass.
// Each frame: KNN database of the enemies, k nearest enemies, separation,
// seek of the seeker's position, integrate (the reference CUDA kernels'
// formulas); then one more frame's steering without the integrate.
// OUT: n x (px,py,pz, speed, eDistance, eForwardDistance, appliedKernels
//     before the clear, after the clear) as floats / uint32 in the loop's
//     order; then the KNN database's hvCellIndices, hvAgentIndicesSorted,
//     hvCellIndicesSorted (n uint32 each) and hvPositionSorted (n x 4
//     floats) after a final updateKNNDatabase; then uint32 wall count and
//     per wall start, mid, end, normal (16 floats).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "gpusteer_opensteer.hpp"

using namespace OpenSteer;

namespace {
template <class T> bool rd(FILE* f, T* p, size_t n) { return fread(p, sizeof(T), n, f) == n; }
template <class T> void wr(FILE* f, const T* p, size_t n) { fwrite(p, sizeof(T), n, f); }
// The plugin's vector helpers (include/OpenSteer/VectorUtils.cuh), for its
// own host code.
float3 sub3(float3 a, float3 b) { return make_float3(a.x - b.x, a.y - b.y, a.z - b.z); }
float dot3(float3 a, float3 b) { return ((a.x * b.x) + (a.y * b.y)) + (a.z * b.z); }
float len3(float3 a) { return sqrtf(dot3(a, a)); }
float3 xyz(float4 a) { return make_float3(a.x, a.y, a.z); }
}  // namespace

int main(int argc, char** argv) {
    if (argc != 9) {
        fprintf(stderr, "usage: %s PARAMS MAPDIR AGENTS STEPS DT SEEKER_X SEEKER_Z OUT\n", argv[0]);
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
    const float3 seeker = make_float3((float)atof(argv[6]), 0.f, (float)atof(argv[7]));
    FILE* out = fopen(argv[8], "wb");
    if (!out) {
        perror(argv[8]);
        return 2;
    }
    try {
        SimulationParams sp;
        sp.load(argv[1]);
        WorldParams& wp = sp.m_WorldParams;
        KNNBinData bins(wp.m_u3Cells, wp.m_f3Dimensions, sp.m_nSearchRadius);
        WallGroup walls(wp.m_u3Cells, sp.m_nKNW);
        if (!walls.LoadFromFile((std::string(argv[2]) + "/" + wp.m_strMapFilename).c_str())) {
            fprintf(stderr, "cannot load the map\n");
            return 2;
        }
        walls.SplitWalls(bins.hvCells());
        walls.SyncDevice();
        updateKNNDatabase(&walls, &bins);

        AgentGroup enemies(wp.m_u3Cells, sp.m_nKNN);
        AgentGroup* gEnemies = &enemies;
        for (uint32_t i = 0; i < n; ++i) {
            auto at = [&](int j) { return f[(size_t)j * n + i]; };
            AgentData ad;
            ad.id = i;
            ad.position = make_float4(at(0), at(1), at(2), 0.f);
            ad.direction = make_float4(at(3), at(4), at(5), 0.f);
            ad.speed = at(6);
            ad.side = make_float3(at(7), at(8), at(9));
            ad.up = make_float3(at(10), at(11), at(12));
            ad.radius = at(13);
            ad.maxSpeed = at(14);
            ad.maxForce = at(15);
            ad.mass = at(16);
            enemies.AddAgent(ad);
        }
        KNNData knnSelf(n, sp.m_nKNN);
        for (int s = 0; s < steps; ++s) {
            updateKNNDatabase(gEnemies, &bins);
            findKNearestNeighbors(gEnemies, &knnSelf, &bins, gEnemies, sp.m_nSearchRadius);
            steerForSeparation(gEnemies, &knnSelf, gEnemies, sp.m_fMinSeparationDistance,
                               sp.m_fMaxSeparationDistance, sp.m_fCosMaxFlockingAngle, sp.m_fWeightSeparation,
                               sp.m_nMaskSeparation);
            steerForSeek(gEnemies, seeker, sp.m_fWeightSeek, sp.m_nMaskSeek);
            updateGroup(gEnemies, elapsedTime);
        }
        // One more frame's steering, not integrated: the agents' applied-
        // behaviour bits are set when the seeker reads them.
        updateKNNDatabase(gEnemies, &bins);
        findKNearestNeighbors(gEnemies, &knnSelf, &bins, gEnemies, sp.m_nSearchRadius);
        steerForSeparation(gEnemies, &knnSelf, gEnemies, sp.m_fMinSeparationDistance, sp.m_fMaxSeparationDistance,
                           sp.m_fCosMaxFlockingAngle, sp.m_fWeightSeparation, sp.m_nMaskSeparation);
        steerForSeek(gEnemies, seeker, sp.m_fWeightSeek, sp.m_nMaskSeek);
        // The seeker's enemy loop.
        const float3 seekerForward = make_float3(0.f, 0.f, 1.f);
        AgentData aData;
        AgentGroupData& agd = gEnemies->GetAgentGroupData();
        std::vector<uint32_t> bits_before(n);
        std::vector<float> rec;
        for (size_t i = 0; i < gEnemies->Size(); i++) {
            agd.getAgentData(i, aData);
            const float3 eOffset = sub3(xyz(aData.position), seeker);
            const float eDistance = len3(eOffset);
            const float eForwardDistance = dot3(seekerForward, eOffset);
            const float r[6] = {aData.position.x, aData.position.y, aData.position.z, aData.speed, eDistance,
                                eForwardDistance};
            rec.insert(rec.end(), r, r + 6);
            bits_before[i] = aData.appliedKernels;
        }
        // The plugin clears the applied-behaviour bits on the device.
        if (hipMemset(agd.pdAppliedKernels(), 0, gEnemies->Size() * sizeof(uint)) != hipSuccess ||
            hipDeviceSynchronize() != hipSuccess) {
            fprintf(stderr, "hipMemset failed\n");
            return 1;
        }
        gEnemies->SetSyncHost();
        for (size_t i = 0; i < gEnemies->Size(); i++) {
            agd.getAgentData(i, aData);
            wr(out, &rec[6 * i], 6);
            wr(out, &bits_before[i], 1);
            const uint32_t after = aData.appliedKernels;
            wr(out, &after, 1);
        }
        // The enemies' KNN database after a build of the final positions.
        updateKNNDatabase(gEnemies, &bins);
        KNNDatabase& db = gEnemies->GetKNNDatabase();
        wr(out, db.hvCellIndices().data(), n);
        wr(out, db.hvAgentIndicesSorted().data(), n);
        wr(out, db.hvCellIndicesSorted().data(), n);
        wr(out, (const float*)db.hvPositionSorted().data(), 4 * (size_t)n);
        // The walls' host vectors.
        WallGroupData const& wgd = walls.GetWallGroupData();
        const uint32_t nw = wgd.size();
        wr(out, &nw, 1);
        for (uint32_t i = 0; i < nw; ++i) {
            const float4 v[4] = {wgd.hvLineStart()[i], wgd.hvLineMid()[i], wgd.hvLineEnd()[i], wgd.hvLineNormal()[i]};
            wr(out, (const float*)v, 16);
        }
        // The device pointers exist for every field (storage order).
        if (!gEnemies->pdSpeed() || !gEnemies->pdRadius() || !gEnemies->pdPosition().x || !gEnemies->pdID() ||
            !agd.pdMass() || !walls.pdPosition().z || walls.pdSpeed() ||
            walls.GetWallGroupData().pdLineNormal().x == nullptr) {
            fprintf(stderr, "device pointer missing\n");
            return 1;
        }
    } catch (const gpusteer::Error& e) {
        fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    fclose(out);
    return 0;
}
