// Per-cell data extraction (the reference's only analysis output):
// ComputeCellDensity / ComputeAvgCellVelocity (CUDA/ProcessData.cu:52-237)
// and their text writers WriteCellDensity / WriteAvgCellVelocity
// (src/ExtractData.cpp:19-141).
//
// Both read the group's bin database of its last build: cell c holds the
// storage slots [start[c], start[c+1]) in ascending agent index, which is the
// reference's sorted order (a stable sort by cell).  Results are written in
// KNNBinDataV2's cell numbering, ix + iz X + iy X Z
// (CUDA/KNNBinDataV2.cu:103-135), the numbering the writers index.
#include <cinttypes>
#include <cstdio>
#include <string>
#include <vector>

#include "gs_ops.hpp"

namespace gs {

__device__ __forceinline__ uint32_t v2_index(const GridDev& g, uint32_t c) {
    const uint32_t iz = c % (uint32_t)g.dz;
    const uint32_t iy = (c / (uint32_t)g.dz) % (uint32_t)g.dy;
    const uint32_t ix = c / ((uint32_t)g.dz * (uint32_t)g.dy);
    return ix + iz * (uint32_t)g.dx + iy * (uint32_t)g.dx * (uint32_t)g.dz;
}

// One lane per cell: agents in the cell (0 when empty).
__global__ __launch_bounds__(kBlock) void k_cell_density(uint32_t C, GridDev g, const uint32_t* __restrict__ start,
                                                         uint32_t* __restrict__ out) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    out[v2_index(g, c)] = start[c + 1] - start[c];
}

// One lane per cell: the sum of forward * speed over the cell's agents in
// ascending agent index, divided by their count as a float; (0, 0, 0) for an
// empty cell.  The current forward and speed of each agent (its storage slot
// keeps them since the build).
__global__ __launch_bounds__(kBlock) void k_cell_velocity(uint32_t C, GridDev g, const uint32_t* __restrict__ start,
                                                          const float* __restrict__ fx, const float* __restrict__ fy,
                                                          const float* __restrict__ fz,
                                                          const float* __restrict__ speed, float* __restrict__ out) {
    const uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= C) return;
    const uint32_t t0 = start[c], t1 = start[c + 1];
    f3 sum = mk3(0.f, 0.f, 0.f);
    for (uint32_t t = t0; t < t1; ++t) sum = sum + ld3(fx, fy, fz, t) * speed[t];
    if (t1 > t0) sum = sum / (float)(t1 - t0);
    const uint32_t o = v2_index(g, c);
    out[3 * (size_t)o] = sum.x;
    out[3 * (size_t)o + 1] = sum.y;
    out[3 * (size_t)o + 2] = sum.z;
}

void launch_cell_density(gs_group_s* g, uint32_t* out) {
    const uint32_t C = g->bin_cells;
    if (C == 0) return;
    k_cell_density<<<blocks_for(C), kBlock, 0, g->ctx->stream>>>(C, g->bin_grid, g->start.p, out);
    GS_LAUNCHED("k_cell_density");
}

void launch_cell_velocity(gs_group_s* g, float* out) {
    const uint32_t C = g->bin_cells;
    if (C == 0) return;
    k_cell_velocity<<<blocks_for(C), kBlock, 0, g->ctx->stream>>>(C, g->bin_grid, g->start.p, g->F(3), g->F(4),
                                                                   g->F(5), g->F(6), out);
    GS_LAUNCHED("k_cell_velocity");
}

// A float as the reference's std::ostream << float prints it (default
// floatfield, precision 6: the %g conversion of the value as a double).
static std::string fmt_float(float v) {
    char b[48];
    snprintf(b, sizeof b, "%g", (double)v);
    return b;
}

// Appends one line to "<prefix>_<cell>.dat" (the reference opens its
// per-cell files in append mode, src/ExtractData.cpp:35-47).
static void append_line(const char* prefix, uint32_t cell, const std::string& line) {
    const std::string name = std::string(prefix) + "_" + std::to_string(cell) + ".dat";
    FILE* f = fopen(name.c_str(), "a");
    if (!f) fail(GS_ERR_INVALID_ARGUMENT, "cannot open " + name);
    fputs(line.c_str(), f);
    fclose(f);
}

// WriteCellDensity (src/ExtractData.cpp:19-88): each selected cell's count
// appended to "<selected_prefix>_<cell>.dat"; "<prefix>.dat" holds a 4-line
// header and rows i = 0 .. Z-2 (all Z rows with all_rows; the reference's
// loop stops one row short, :63) of X tab-terminated counts of cells
// i X + j.
void write_cell_density(const GridDev& G, const std::vector<uint32_t>& d, const char* prefix,
                        const uint32_t* selected, uint32_t n_selected, const char* selected_prefix, int all_rows) {
    for (uint32_t s = 0; s < n_selected; ++s) {
        if (selected[s] >= d.size()) fail(GS_ERR_INVALID_ARGUMENT, "selected cell out of range");
        append_line(selected_prefix, selected[s], std::to_string(d[selected[s]]) + "\n");
    }
    const std::string name = std::string(prefix) + ".dat";
    FILE* f = fopen(name.c_str(), "w");
    if (!f) fail(GS_ERR_INVALID_ARGUMENT, "cannot open " + name);
    fprintf(f, "%%\tCell densities\n%%\tworldCells:\tX: %d\tY: %d\tZ: %d\n%%\tData format: \n%%\tX\tY\tnumAgents\n",
            G.dx, G.dy, G.dz);
    const uint32_t rows = all_rows ? (uint32_t)G.dz : (uint32_t)G.dz - 1u;
    for (uint32_t i = 0; i < rows; ++i) {
        for (uint32_t j = 0; j < (uint32_t)G.dx; ++j) fprintf(f, "%" PRIu32 "\t", d[(size_t)i * G.dx + j]);
        fputc('\n', f);
    }
    fclose(f);
}

// WriteAvgCellVelocity (src/ExtractData.cpp:90-140): each selected cell's
// "v.x<TAB>v.z" appended to "<selected_prefix>_<cell>.dat"; "<prefix>.dat"
// holds a 4-line header and lines "i<TAB>j<TAB>v.x<TAB>v.z" for rows
// i = 0 .. Z-2 (all with all_rows) and j = 0 .. X-1 of cell i X + j.  The
// reference reads v.x at i Z + j (:134), the same cell when X = Z.
void write_cell_velocity(const GridDev& G, const std::vector<float>& v, const char* prefix,
                         const uint32_t* selected, uint32_t n_selected, const char* selected_prefix, int all_rows) {
    const size_t C = v.size() / 3;
    for (uint32_t s = 0; s < n_selected; ++s) {
        const uint32_t c = selected[s];
        if (c >= C) fail(GS_ERR_INVALID_ARGUMENT, "selected cell out of range");
        append_line(selected_prefix, c, fmt_float(v[3 * (size_t)c]) + "\t" + fmt_float(v[3 * (size_t)c + 2]) + "\n");
    }
    const std::string name = std::string(prefix) + ".dat";
    FILE* f = fopen(name.c_str(), "w");
    if (!f) fail(GS_ERR_INVALID_ARGUMENT, "cannot open " + name);
    fprintf(f, "%%\tCell average velocities\n%%\tworldCells:\tX: %d\tY: %d\tZ: %d\n%%\tData format: \n"
               "%%\tX\tZ\tV.x\tV.z\n",
            G.dx, G.dy, G.dz);
    const uint32_t rows = all_rows ? (uint32_t)G.dz : (uint32_t)G.dz - 1u;
    for (uint32_t i = 0; i < rows; ++i)
        for (uint32_t j = 0; j < (uint32_t)G.dx; ++j) {
            const size_t c = (size_t)i * G.dx + j;
            fprintf(f, "%" PRIu32 "\t%" PRIu32 "\t%s\t%s\n", i, j, fmt_float(v[3 * c]).c_str(),
                    fmt_float(v[3 * c + 2]).c_str());
        }
    fclose(f);
}

}  // namespace gs
