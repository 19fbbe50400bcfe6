// Bin database build: counting sort of a group's points by lattice cell, and
// the physical reorder of the group's state into that order.
//
// Replaces KNNBinningV2UpdateDBCUDA::run (include/OpenSteer/CUDA/KNNBinningV2.cu:87-162):
// BuildDB (cell key per agent, KNNBinningV2Kernel.cu:269-296), thrust::sort_by_key
// (:127-129), the cellStart memset (:132) and ReorderDB (KNNBinningV2Kernel.cu:314-379).
//
// Passes (all over storage slots, coalesced):
//   1. key      cell per point + atomic arrival rank within the cell (+ optional wrap)
//   2. scan     exclusive scan of the per-cell counts -> start[C+1]; clears counts
//   3. scatter  id_tmp[start[cell] + rank] = agent index (the arrival order)
//   4. reorder  per old slot: its final slot puts the members of each cell in
//               ascending agent index (a stable sort by cell, as thrust's);
//               every live field, the bits and the agent index are written
//               there in the alternate buffers, with the old <-> new maps.
// Because the previous build left the state in cell order and agents move
// little between builds, the reorder's writes land nearly sequentially.  The counts array is all-zero between builds (pass 2 clears it),
// so no memset is recorded in the stream.
#include "gs_internal.hpp"

namespace gs {

// Members of a cell are put in ascending agent index (the stable sort's
// order) by counting, for each member, the members with a smaller index: m
// reads per member of an m-member cell, issued by all members at once.
// Cells with more members than this keep the atomic arrival order instead
// (KNN results never depend on the order inside a cell; the per-cell
// velocity sum of such a cell does).
constexpr uint32_t kCanonCap = 65536;
constexpr int kScanItems = 8;
constexpr int kScanTile = kBlock * kScanItems;  // 2048 cells per scan block

__global__ __launch_bounds__(kBlock) void k_fill_u32(uint32_t* __restrict__ p, uint32_t n, uint32_t v) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ __launch_bounds__(kBlock) void k_iota(uint32_t* __restrict__ p, uint32_t n) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = i;
}

__global__ __launch_bounds__(kBlock) void k_invert(uint32_t n, const uint32_t* __restrict__ ids,
                                                   uint32_t* __restrict__ slot_of_id) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s < n) slot_of_id[ids[s]] = s;
}

void launch_invert_ids(gs_group_s* g) {
    if (g->n == 0) return;
    k_invert<<<blocks_for(g->n), kBlock, 0, g->ctx->stream>>>(g->n, g->ids(), g->slot_of_id.p);
    GS_LAUNCHED("k_invert");
}

void launch_fill_u32(uint32_t* p, uint32_t n, uint32_t v, hipStream_t st) {
    if (n == 0) return;
    k_fill_u32<<<blocks_for(n), kBlock, 0, st>>>(p, n, v);
    GS_LAUNCHED("k_fill_u32");
}

void launch_iota(uint32_t* p, uint32_t n, hipStream_t st) {
    if (n == 0) return;
    k_iota<<<blocks_for(n), kBlock, 0, st>>>(p, n);
    GS_LAUNCHED("k_iota");
}

__global__ __launch_bounds__(kBlock) void k_bin_key(uint32_t n, float* __restrict__ px, float* __restrict__ py,
                                                    float* __restrict__ pz, GridDev g, uint32_t* __restrict__ key,
                                                    uint32_t* __restrict__ rank, uint32_t* __restrict__ count,
                                                    int wrap_shape, float wcx, float wcy, float wcz, float wsx,
                                                    float wsy, float wsz, uint32_t* __restrict__ box_acc,
                                                    uint32_t* __restrict__ box_clear) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = i < n;
    const int lane = threadIdx.x & 63;
    // Inactive lanes get a key no cell has, so they form segments of their own.
    uint32_t c = 0xFFFFFFC0u | (uint32_t)lane;
    float x = 0.f, y = 0.f, z = 0.f;
    if (active) {
        x = px[i], y = py[i], z = pz[i];
        const float x0 = x, y0 = y, z0 = z;
        wrap_point(wrap_shape, wcx, wcy, wcz, wsx, wsy, wsz, x, y, z);
        // Only wrapped coordinates are written back.
        if (__float_as_uint(x) != __float_as_uint(x0)) px[i] = x;
        if (__float_as_uint(y) != __float_as_uint(y0)) py[i] = y;
        if (__float_as_uint(z) != __float_as_uint(z0)) pz[i] = z;
        c = cell_index(g, axis_cell(x, g.ox, g.sx, g.dx), axis_cell(y, g.oy, g.sy, g.dy),
                       axis_cell(z, g.oz, g.sz, g.dz));
        key[i] = c;
    }
    count_and_rank(active, c, i, count, rank);
    if (box_acc) {
        // The bounding box of the positions this build snapshots (after the
        // wrap); the other buffer is cleared for the next build.
        if (blockIdx.x == 0 && threadIdx.x < 6) box_clear[threadIdx.x] = threadIdx.x < 3 ? kBoxEmptyMin : kBoxEmptyMax;
        box_accumulate(active, x, y, z, box_acc);
    }
}

// Wavefront-then-block exclusive scan of one value per thread.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* lds_wave, uint32_t& total) {
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    uint32_t inc = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t t = __shfl_up(inc, d, 64);
        if (lane >= d) inc += t;
    }
    if (lane == 63) lds_wave[wave] = inc;
    __syncthreads();
    uint32_t wave_off = 0, sum = 0;
#pragma unroll
    for (int w = 0; w < kBlock / 64; ++w) {
        const uint32_t s = lds_wave[w];
        if (w < wave) wave_off += s;
        sum += s;
    }
    total = sum;
    return wave_off + inc - v;
}

// Pass 2a: each block scans kScanTile counts locally, writes local offsets
// into start and its tile sum into partial[block].
__global__ __launch_bounds__(kBlock) void k_scan_tiles(uint32_t C, const uint32_t* __restrict__ count,
                                                       uint32_t* __restrict__ start, uint32_t* __restrict__ partial) {
    __shared__ uint32_t lds[kBlock / 64];
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile + (uint64_t)threadIdx.x * kScanItems;
    uint32_t v[kScanItems];
    uint32_t sum = 0;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        v[j] = (base + j < C) ? count[base + j] : 0u;
        sum += v[j];
    }
    uint32_t total;
    uint32_t off = block_exclusive_scan(sum, lds, total);
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        if (base + j < C) start[base + j] = off;
        off += v[j];
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = total;
}

// Pass 2b (many tiles): one block scans the tile sums in place (exclusive)
// and writes the grand total to start[C].
__global__ __launch_bounds__(kBlock) void k_scan_partials(uint32_t nparts, uint32_t* __restrict__ partial,
                                                          uint32_t* __restrict__ start, uint32_t C) {
    __shared__ uint32_t lds[kBlock / 64];
    uint32_t carry = 0;
    for (uint32_t base = 0; base < nparts; base += kBlock) {
        const uint32_t i = base + threadIdx.x;
        const uint32_t v = i < nparts ? partial[i] : 0u;
        uint32_t total;
        const uint32_t ex = block_exclusive_scan(v, lds, total);
        if (i < nparts) partial[i] = carry + ex;
        carry += total;
        __syncthreads();
    }
    if (threadIdx.x == 0) start[C] = carry;
}

// Pass 2c: add the tile offsets and clear the counts for the next build.
// With few tiles (fused = 1) each block sums the tile totals before its own
// tile itself, which saves the launch of pass 2b; the last block then writes
// the grand total to start[C].
constexpr uint32_t kScanFusedTiles = 1024;
__global__ __launch_bounds__(kBlock) void k_scan_add(uint32_t C, const uint32_t* __restrict__ partial,
                                                     uint32_t* __restrict__ start, uint32_t* __restrict__ count,
                                                     int fused) {
    __shared__ uint32_t lds[kBlock / 64];
    uint32_t add;
    if (fused) {
        uint32_t part = 0;
        for (uint32_t i = threadIdx.x; i < blockIdx.x; i += kBlock) part += partial[i];
        (void)block_exclusive_scan(part, lds, add);
    } else {
        add = partial[blockIdx.x];
    }
    const uint64_t base = (uint64_t)blockIdx.x * kScanTile;
#pragma unroll
    for (int j = 0; j < kScanItems; ++j) {
        const uint64_t c = base + (uint64_t)j * kBlock + threadIdx.x;
        if (c < C) {
            start[c] += add;
            count[c] = 0u;
        }
    }
    if (fused && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) start[C] = add + partial[blockIdx.x];
}

// Pass 3: the agent index of every slot of the arrival order (slot t =
// start[cell] + arrival rank), which the reorder's canonical ranks read.
__global__ __launch_bounds__(kBlock) void k_scatter(uint32_t n, const uint32_t* __restrict__ key,
                                                    const uint32_t* __restrict__ rank,
                                                    const uint32_t* __restrict__ start,
                                                    const uint32_t* __restrict__ ids, uint32_t* __restrict__ id_tmp) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    // start[key]+rank < n whenever the counts were zero before the key pass;
    // the guard keeps a corrupted count array from writing out of bounds.
    const uint32_t t = start[key[i]] + rank[i];
    if (t < n) id_tmp[t] = ids[i];
}

// Pass 4, one lane per old slot s (every read coalesced, the writes land
// near s since agents rarely change cells): the agent's arrival slot is
// t = start[cell] + rank, its final slot the cell start plus the number of
// members of its cell with a smaller agent index.
__global__ __launch_bounds__(kBlock) void k_reorder(uint32_t n, const uint32_t* __restrict__ rank,
                                                    const uint32_t* __restrict__ id_tmp,
                                                    const uint32_t* __restrict__ key,
                                                    const uint32_t* __restrict__ start,
                                                    const uint32_t* __restrict__ ids, uint32_t* __restrict__ ids_new,
                                                    uint32_t* __restrict__ new_to_old,
                                                    uint32_t* __restrict__ old_to_new,
                                                    const uint32_t* __restrict__ bits, uint32_t* __restrict__ bits_new,
                                                    FieldSet fs, float4* __restrict__ spos,
                                                    float4* __restrict__ sfwd) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const uint32_t c = key[s];
    const uint32_t b = start[c], e = min(start[c + 1], n);
    const uint32_t t = b + rank[s];
    if (t >= n) return;  // only after a corrupted build (see k_scatter)
    const uint32_t id = ids[s];
    uint32_t d = t;
    if (e - b <= kCanonCap) {
        uint32_t r = 0;
        for (uint32_t j = b; j < e; ++j) r += (id_tmp[j] < id) ? 1u : 0u;
        d = b + r;
    }
    new_to_old[d] = s;
    old_to_new[s] = d;
    ids_new[d] = id;
    if (bits) bits_new[d] = bits[s];
#pragma unroll
    for (int f = 0; f < kFields; ++f)
        if (f < fs.nf) fs.dst[f][d] = fs.src[f][s];
    // fields 0..2 (positions) are always first in the set; a lazy build also
    // snapshots forward and speed (fields 3..6, passed after them)
    spos[d] = make_float4(fs.src[0][s], fs.src[1][s], fs.src[2][s], __uint_as_float(id));
    if (sfwd) sfwd[d] = make_float4(fs.src[3][s], fs.src[4][s], fs.src[5][s], fs.src[6][s]);
}

void launch_scan_counts(uint32_t* count, uint32_t* start, uint32_t C, hipStream_t st) {
    const uint32_t ntiles = (C + kScanTile - 1) / kScanTile;
    uint32_t* partial = start + C + 1;
    k_scan_tiles<<<ntiles, kBlock, 0, st>>>(C, count, start, partial);
    GS_LAUNCHED("k_scan_tiles");
    const int fused = ntiles <= kScanFusedTiles ? 1 : 0;
    if (!fused) {
        k_scan_partials<<<1, kBlock, 0, st>>>(ntiles, partial, start, C);
        GS_LAUNCHED("k_scan_partials");
    }
    k_scan_add<<<ntiles, kBlock, 0, st>>>(C, partial, start, count, fused);
    GS_LAUNCHED("k_scan_add");
}

void launch_scatter_reorder(uint32_t n, const uint32_t* key, const uint32_t* rank, const uint32_t* start,
                            const uint32_t* ids, uint32_t* ids_new, uint32_t* id_tmp,
                            uint32_t* new_to_old, uint32_t* old_to_new, const uint32_t* bits, uint32_t* bits_new,
                            const FieldSet& fs, float4* spos, hipStream_t st, float4* sfwd) {
    if (n == 0) return;
    k_scatter<<<blocks_for(n), kBlock, 0, st>>>(n, key, rank, start, ids, id_tmp);
    GS_LAUNCHED("k_scatter");
    k_reorder<<<blocks_for(n), kBlock, 0, st>>>(n, rank, id_tmp, key, start, ids, ids_new, new_to_old,
                                                old_to_new, bits, bits_new, fs, spos, sfwd);
    GS_LAUNCHED("k_reorder");
}

// lazy (fused Boids step only): move nothing but the snapshots (positions,
// and forward + speed into sfwd), the agent indices and the maps; every
// other field stays in the previous order until the flock kernel, which reads
// it through new_to_old and writes the step's results in the new order.
void launch_bins_build(gs_group_s* g, const gs_bins_s* b, int wrap_shape, const float* wc, const float* ws,
                       bool lazy, bool keys_ready) {
    const uint32_t n = g->n;
    const uint32_t C = b->cells;
    hipStream_t st = g->ctx->stream;
    const bool new_cells = !g->binned || g->bin_cells != C || g->count.cap < C;
    // Keys from the previous step's flock kernel (counts already taken):
    // only for the lattice that step was built on, never with a box to keep.
    if (keys_ready && (new_cells || g->bbox_wanted)) fail(GS_ERR_INVALID_ARGUMENT, "internal: stale fused keys");
    g->key.reserve(n);
    g->rank.reserve(n);
    g->id_tmp.reserve(n);
    g->new_to_old.reserve(n);
    g->old_to_new.reserve(n);
    g->spos.reserve(n);
    const uint32_t ntiles = (C + kScanTile - 1) / kScanTile;
    g->start.reserve((size_t)C + 1 + ntiles);  // partial tile sums live after start[C]
    if (new_cells) {
        g->count.reserve(C);
        // A fill kernel rather than hipMemsetAsync, so the build can be
        // captured into a graph.
        launch_fill_u32(g->count.p, C, 0u, st);
    }
    g->bin_cells = C;
    g->bin_grid = b->g;
    float wcv[3] = {0, 0, 0}, wsv[3] = {0, 0, 0};
    if (wc) memcpy(wcv, wc, sizeof wcv);
    if (ws) memcpy(wsv, ws, sizeof wsv);
    // Bounding box of this build's positions, once a capped search targets
    // the group.
    uint32_t *box_acc = nullptr, *box_clear = nullptr;
    if (g->bbox_wanted && g->bbox[0].p && g->bbox[1].p && n > 0) {
        g->bbox_cur ^= 1;
        box_acc = g->bbox[g->bbox_cur].p;
        box_clear = g->bbox[g->bbox_cur ^ 1].p;
    }
    g->bbox_fresh = box_acc != nullptr;
    if (n > 0 && !keys_ready) {
        k_bin_key<<<blocks_for(n), kBlock, 0, st>>>(n, g->F(0), g->F(1), g->F(2), b->g, g->key.p, g->rank.p,
                                                    g->count.p, wrap_shape, wcv[0], wcv[1], wcv[2], wsv[0], wsv[1],
                                                    wsv[2], box_acc, box_clear);
        GS_LAUNCHED("k_bin_key");
    }
    launch_scan_counts(g->count.p, g->start.p, C, st);
    if (n > 0) {
        FieldSet fs{};
        int nf = 0;
        // Fields holding one value in every slot are unchanged by any
        // permutation and stay where they are (positions always move: the
        // snapshot is built from them).
        const uint32_t moved = lazy ? 0u : g->live & ~(g->inv & ~kFPos);
        for (int f = 0; f < kFields; ++f)
            if (moved & (1u << f)) {
                fs.src[nf] = g->F(f);
                fs.dst[nf] = g->Falt(f);
                ++nf;
            }
        fs.nf = nf;
        if (lazy) {
            for (int f = 0; f < 7; ++f) fs.src[f] = g->F(f);
            g->sfwd.reserve(n);
        }
        const bool has_bits = !lazy && g->kind == GS_KIND_AGENTS && !g->bits_zero;
        launch_scatter_reorder(n, g->key.p, g->rank.p, g->start.p, g->ids(), g->ids_alt(), g->id_tmp.p,
                               g->new_to_old.p, g->old_to_new.p, has_bits ? g->bits() : nullptr,
                               has_bits ? g->bitsb[g->bcur ^ 1].p : nullptr, fs, g->spos.p, st,
                               lazy ? g->sfwd.p : nullptr);
        for (int f = 0; f < kFields; ++f)
            if (moved & (1u << f)) g->flip(f);
        if (has_bits) g->bcur ^= 1;
        g->icur ^= 1;
    }
    g->binned = true;
    g->snap_valid = true;
}

}  // namespace gs
