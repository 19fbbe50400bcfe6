// The multi-group plugin of plugins/MultiGroup.cpp written against the
// reference's API on include/gpusteer_opensteer.hpp: groups sharing one bin
// lattice, each with a KNN container per group and its path; a frame runs
// Group::preUpdate for every group (:546-556) and then Group::update for each
// group in turn (:558-611).  The agents come from a file (their placement is
// the caller's), so the GPU test can compare the result with the oracle:
//
//   multi_group_plugin PARAMS AGENTS STEPS DT OUT
//
// AGENTS (little-endian): uint32 groups; per group uint32 n, then n AgentData
//     records as float arrays px,py,pz, fx,fy,fz, speed, sx,sy,sz, ux,uy,uz,
//     radius, max_speed, max_force, mass.
// OUT: per group n x (px,py,pz, fx,fy,fz, speed, sx,sy,sz, ux,uy,uz) floats in
//     agent-id order.
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <vector>

#include "gpusteer_opensteer.hpp"

using namespace OpenSteer;

namespace {
template <class T> bool rd(FILE* f, T* p, size_t n) { return fread(p, sizeof(T), n, f) == n; }
}  // namespace

int main(int argc, char** argv) {
    if (argc != 6) {
        fprintf(stderr, "usage: %s PARAMS AGENTS STEPS DT OUT\n", argv[0]);
        return 2;
    }
    FILE* in = fopen(argv[2], "rb");
    if (!in) {
        perror(argv[2]);
        return 2;
    }
    uint32_t ng = 0;
    bool ok = rd(in, &ng, 1);
    std::vector<std::vector<float>> data(ng);
    std::vector<uint32_t> counts(ng);
    for (uint32_t gi = 0; ok && gi < ng; ++gi) {
        ok = rd(in, &counts[gi], 1);
        data[gi].resize(17 * (size_t)counts[gi]);
        ok = ok && rd(in, data[gi].data(), data[gi].size());
    }
    fclose(in);
    if (!ok) {
        fprintf(stderr, "short input\n");
        return 2;
    }
    const int steps = atoi(argv[3]);
    const float elapsedTime = (float)atof(argv[4]);
    try {
        SimulationParams sp;
        sp.load(argv[1]);
        WorldParams& wp = sp.m_WorldParams;
        if (sp.m_vecGroupParams.size() != ng) throw gpusteer::Error(GS_ERR_INVALID_ARGUMENT, "group count");
        KNNBinData bins(wp.m_u3Cells, wp.m_f3Dimensions, sp.m_nSearchRadius);

        // Group (:180-230): agents, a KNN container per group, the path.
        std::vector<std::unique_ptr<AgentGroup>> groups;
        std::vector<std::vector<std::unique_ptr<KNNData>>> knn(ng);
        std::vector<std::unique_ptr<PolylinePathwayCUDA>> paths(ng);
        for (uint32_t gi = 0; gi < ng; ++gi) {
            groups.emplace_back(new AgentGroup(wp.m_u3Cells, sp.m_nKNN));
            const std::vector<float>& f = data[gi];
            const uint32_t n = counts[gi];
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
                if (!groups[gi]->AddAgent(ad)) throw gpusteer::Error(GS_ERR_INVALID_ARGUMENT, "duplicate id");
            }
            for (uint32_t j = 0; j < ng; ++j) knn[gi].emplace_back(new KNNData(n, sp.m_nKNN));
            const GroupParams& gp = sp.m_vecGroupParams[gi];
            if (gp.m_vecPathPoints.size() >= 2)
                paths[gi].reset(new PolylinePathwayCUDA(gp.m_vecPathPoints, gp.m_fPathRadius, gp.m_bPathIsCyclic));
        }

        for (int s = 0; s < steps; ++s) {
            // Group::preUpdate for every group
            for (uint32_t gi = 0; gi < ng; ++gi) {
                wrapWorld(groups[gi].get(), wp.m_f3Dimensions);
                updateKNNDatabase(groups[gi].get(), &bins);
                // the reference's cudaMemset of pdAppliedKernels
                gpusteer::check(gs_clear_bits(groups[gi]->device_handle()), "gs_clear_bits");
            }
            // Group::update for each group in turn
            for (uint32_t gi = 0; gi < ng; ++gi) {
                AgentGroup* A = groups[gi].get();
                for (uint32_t i = 0; i < ng; ++i)
                    findKNearestNeighbors(A, knn[gi][i].get(), &bins, groups[i].get(), sp.m_nSearchRadius);
                for (uint32_t i = 0; i < ng; ++i) {
                    if (i == gi) continue;
                    steerToAvoidNeighbors(A, knn[gi][i].get(), groups[i].get(), sp.m_fMinTimeToCollision,
                                          sp.m_fMinSeparationDistance, sp.m_bAvoidCloseNeighbors,
                                          sp.m_fWeightAvoidNeighbors, sp.m_nMaskAvoidNeighbors);
                }
                for (uint32_t i = 0; i < ng; ++i)
                    steerForSeparation(A, knn[gi][i].get(), groups[i].get(), sp.m_fMinSeparationDistance,
                                       sp.m_fMaxSeparationDistance, sp.m_fCosMaxFlockingAngle, sp.m_fWeightSeparation,
                                       sp.m_nMaskSeparation);
                steerForAlignment(A, knn[gi][gi].get(), A, sp.m_fMinFlockingDistance, sp.m_fMaxFlockingDistance,
                                  sp.m_fCosMaxFlockingAngle, sp.m_fWeightAlignment, sp.m_nMaskAlignment);
                steerForCohesion(A, knn[gi][gi].get(), A, sp.m_fMinFlockingDistance, sp.m_fMaxFlockingDistance,
                                 sp.m_fCosMaxFlockingAngle, sp.m_fWeightCohesion, sp.m_nMaskCohesion);
                if (paths[gi]) {
                    steerToFollowPath(A, paths[gi].get(), sp.m_fPathPredictionTime, sp.m_fWeightFollowPath,
                                      sp.m_nMaskFollowPath);
                    steerForSeek(A, paths[gi]->hvPoints()[1], sp.m_fWeightSeek, sp.m_nMaskSeek);
                }
                updateGroup(A, elapsedTime);
                for (uint32_t i = 0; i < ng; ++i)
                    antiPenetrationAgents(A, knn[gi][i].get(), groups[i].get(), sp.m_nMaskAntiPenetrationAgents);
            }
        }

        FILE* out = fopen(argv[5], "wb");
        if (!out) {
            perror(argv[5]);
            return 2;
        }
        for (uint32_t gi = 0; gi < ng; ++gi)
            for (uint32_t i = 0; i < counts[gi]; ++i) {
                AgentData ad;
                if (!groups[gi]->GetDataForAgent(i, ad))
                    throw gpusteer::Error(GS_ERR_INVALID_ARGUMENT, "missing agent");
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
