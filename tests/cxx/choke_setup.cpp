// The scenario set-up of a choke-point plugin written against the
// reference's API (ChokePointNew's constructor, plugins/ChokePointNew.cpp:107-119:
// SimulationParams::load, KNNBinData, WallGroup::LoadFromFile,
// SplitWalls(hvCells()), SyncDevice, updateKNNDatabase), built on
// include/gpusteer_opensteer.hpp.  The GPU test drives it through files:
//
//   choke_setup PARAMS MAPDIR OUT
//
// OUT (little-endian): uint32 knw, cells[3]; float dimensions[3]; uint32 n
//     walls; then n x (start xyz, end xyz, normal xyz, mid xyz) floats, the
//     walls as the wall group holds them after SplitWalls and SyncDevice;
//     then the wall group's bin database read back from the device: uint32
//     key[n] (cell per wall), perm[n] (wall per sorted slot), float
//     sorted midpoints x[n], y[n], z[n].
#include <cstdio>
#include <string>
#include <vector>

#include "gpusteer_opensteer.hpp"

using namespace OpenSteer;

int main(int argc, char** argv) {
    if (argc != 4) {
        fprintf(stderr, "usage: %s PARAMS MAPDIR OUT\n", argv[0]);
        return 2;
    }
    try {
        SimulationParams sp;
        sp.load(argv[1]);
        WorldParams& wp = sp.m_WorldParams;
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
        FILE* out = fopen(argv[3], "wb");
        if (!out) {
            perror(argv[3]);
            return 2;
        }
        const uint32_t head[4] = {sp.m_nKNW, wp.m_u3Cells.x, wp.m_u3Cells.y, wp.m_u3Cells.z};
        const float dims[3] = {wp.m_f3Dimensions.x, wp.m_f3Dimensions.y, wp.m_f3Dimensions.z};
        const uint32_t n = walls.Size();
        fwrite(head, 4, 4, out);
        fwrite(dims, 4, 3, out);
        fwrite(&n, 4, 1, out);
        for (uint32_t i = 0; i < n; ++i) {
            float3 s, e, nrm, mid;
            walls.GetWall(i, s, e, nrm, mid);
            const float rec[12] = {s.x, s.y, s.z, e.x, e.y, e.z, nrm.x, nrm.y, nrm.z, mid.x, mid.y, mid.z};
            fwrite(rec, 4, 12, out);
        }
        std::vector<uint32_t> key(n), perm(n);
        std::vector<float> sorted[3];
        for (auto& a : sorted) a.resize(n);
        gpusteer::check(gs_group_download_bins(walls.device_handle(), key.data(), nullptr, perm.data(),
                                               sorted[0].data(), sorted[1].data(), sorted[2].data()),
                        "gs_group_download_bins");
        fwrite(key.data(), 4, n, out);
        fwrite(perm.data(), 4, n, out);
        for (auto& a : sorted) fwrite(a.data(), 4, n, out);
        fclose(out);
    } catch (const gpusteer::Error& e) {
        fprintf(stderr, "%s\n", e.what());
        return 1;
    }
    return 0;
}
