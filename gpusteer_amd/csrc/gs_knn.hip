// Exact k-nearest-neighbour search, binned and brute force.
//
// Semantics (both kernels): for each query a of group A, the k points b of
// group B with the smallest key (d2, b), d2 = ((ax-bx)^2 + (ay-by)^2) + (az-bz)^2
// in fp32, excluding a itself when A and B are the same group.  This is the
// order KNNBruteForceCUDAKernelV2 produces (strict '<' over an ascending scan:
// the lower index wins ties, include/OpenSteer/CUDA/KNNBruteForceCudaKernel.cu:245-270).
// Output is slot-major [k][nA] indexed by the query's own index
// (KNNBinningV2Kernel.cu:520-526), padded with UINT32_MAX / FLT_MAX (:451-457),
// distance = sqrtf(d2).
//
// Top-k selection: a register-resident sorted list of KMAX (d2, index) pairs
// with a branch-free compare-exchange ripple.  For k < KMAX the first KMAX-k
// entries hold a sentinel below every real key, so the list's last entry is
// always the k-th best and no register is ever indexed dynamically.
//
// Binned search replaces KNNBinningV2Kernel (KNNBinningV2Kernel.cu:417-528),
// which scans a fixed cell radius and is therefore approximate.  Here each
// query (one lane, queries in sorted/cell order so a wavefront's lanes share
// cells) scans the box of ring 1 around its cell (3 x 3 columns, or 5 x 5 for
// fine lattices, each column's z-run trimmed to the bound's sphere), then
// rings of growing radius while the k-th best d2 is not below the squared
// distance to every unscanned cell.  The bound starts at +inf or, when the
// container holds the exact result of the previous call for the same groups,
// at the largest current d2 of those k previous neighbours (a valid upper
// bound on the k-th distance: they are k distinct candidates).  Either way
// the result is the exact k nearest.
#include "gs_knn_core.hpp"

namespace gs {

// One lane per query slot s of A (storage = cell order, so a wavefront's lanes
// share cells and their candidate runs overlap in cache).  The flat path
// (RF == 1) takes its launch shape as template arguments: BS threads per
// block and APL queries per lane (a block covers APL consecutive runs of BS
// slots), both picked by the autotuner (gs_context_set_shape).
template <int KMAX, int RF, int BS = kBlock, int APL = 1>
__global__ __launch_bounds__(BS, RF == 1 ? (KMAX <= 8 ? 5 : (KMAX <= 16 ? 4 : 1)) : 1) void k_knn_binned(uint32_t nA, const float4* __restrict__ apos, uint32_t nB,
                                                       const float4* __restrict__ bpos,
                                                       const uint32_t* __restrict__ bstart, GridDev g, int self,
                                                       uint32_t k, int R0, float cap2, const uint32_t* __restrict__ seed,
                                                       const uint32_t* __restrict__ row_map,
                                                       const uint32_t* __restrict__ val_map,
                                                       uint32_t* __restrict__ out_idx, float* __restrict__ out_dist,
                                                       const uint32_t* __restrict__ box) {
    if constexpr (RF == 1) {
        // Slot a of the lane's APL slots (a block covers APL runs of BS slots).
        auto one = [&](uint32_t s) {
            TopK<KMAX> h;
            h.init(k);
            // Every lane stays to the end: the wavefront completes its lanes'
            // inexact queries together (lanes past nA only help).
            const bool active = s < nA;
            const float4 q = active ? apos[s] : make_float4(0.f, 0.f, 0.f, 0.f);
            FlatQuery Q{q.x, q.y, q.z, INFINITY, self ? s : UINT32_MAX, active};
            if (nB > 0) {
                Q.thr = fminf(cap2, (seed && active) ? seed_bound<KMAX>(seed, k, nA, row_map ? row_map[s] : s, nB,
                                                                        val_map, q.x, q.y, q.z, bpos)
                                                     : INFINITY);
                // A capped query that cannot reach B's bounding box has no
                // candidate (box: NULL for an uncapped search).
                if (box && active && box_beyond(box, q.x, q.y, q.z, cap2)) Q.thr = -1.f;
                GS_FLAT_LDS_NK(BS, KMAX)
                knn_search_flat<KMAX, BS>(h, Q, g, nB, bpos, bstart, flds);
            }
            if (active) h.store(k, nA, s, out_idx, out_dist);
        };
        if constexpr (APL == 1) {
            one(xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x);
        } else {
            for (int a = 0; a < APL; ++a) one((xcd_block(blockIdx.x, gridDim.x) * APL + a) * BS + threadIdx.x);
        }
    } else {
        static_assert(BS == kBlock && APL == 1, "the box walk has the default shape");
        const uint32_t s = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
        TopK<KMAX> h;
        h.init(k);
        if (s >= nA) return;
        const float4 q = apos[s];
        if (nB > 0) {
            const float thr = fminf(cap2, seed ? seed_bound<KMAX>(seed, k, nA, row_map ? row_map[s] : s, nB,
                                                                  val_map, q.x, q.y, q.z, bpos)
                                               : INFINITY);
            GS_DEFER_LDS
            knn_search<KMAX, RF>(h, thr, q.x, q.y, q.z, self ? s : UINT32_MAX, g, R0, bpos, bstart, lbuf);
        }
        h.store(k, nA, s, out_idx, out_dist);
    }
}

// Brute force: one lane per query slot, B staged through LDS in tiles of one
// block (positions and agent indices).
template <int KMAX>
__global__ __launch_bounds__(kBlock) void k_knn_brute(uint32_t nA, const float* __restrict__ ax,
                                                      const float* __restrict__ ay, const float* __restrict__ az,
                                                      uint32_t nB, const float* __restrict__ bx,
                                                      const float* __restrict__ by, const float* __restrict__ bz,
                                                      const uint32_t* __restrict__ bids, int self, uint32_t k,
                                                      float cap2, uint32_t* __restrict__ out_idx,
                                                      float* __restrict__ out_dist) {
    __shared__ float tx[kBlock], ty[kBlock], tz[kBlock];
    __shared__ uint32_t tid[kBlock];
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = q < nA;
    const float qx = active ? ax[q] : 0.f, qy = active ? ay[q] : 0.f, qz = active ? az[q] : 0.f;
    TopK<KMAX> h;
    h.init(k);
    for (uint32_t base = 0; base < nB; base += kBlock) {
        const uint32_t t = base + threadIdx.x;
        if (t < nB) {
            tx[threadIdx.x] = bx[t];
            ty[threadIdx.x] = by[t];
            tz[threadIdx.x] = bz[t];
            tid[threadIdx.x] = bids[t];
        }
        __syncthreads();
        const uint32_t m = min((uint32_t)kBlock, nB - base);
        if (active) {
            for (uint32_t j = 0; j < m; ++j) {
                const uint32_t b = base + j;
                if (self && b == q) continue;
                const float d2 = sqdist(qx, qy, qz, tx[j], ty[j], tz[j]);
                const uint64_t kk = make_key(d2, tid[j]);
                if (d2 <= cap2 && h.accepts(kk)) h.push(kk, b);
            }
        }
        __syncthreads();
    }
    if (active) h.store(k, nA, q, out_idx, out_dist);
}

// Binned-database semantics over a small or sparse B (walls, obstacles):
// the same queries and candidates as k_knn_binned (the position snapshots of
// the last builds, agent indices in .w) scanned exhaustively from LDS, so a
// small target set costs nB distance tests instead of a shell expansion over
// empty cells.  Identical results (the search is exact either way).
//
// The snapshot is in cell order, so each wave's 64 entries of a tile are
// spatially coherent: the wave takes their bounding box, and a wave skips a
// chunk when no lane's query can reach the box within its current bound
// (its k-th key's d2, or the cap).  The box distance is formed with the
// same per-axis subtractions and the same summation order as a candidate's
// d2, and every step is monotone, so a member's d2 is never below it: the
// skip drops only candidates the heap would reject.
template <int KMAX>
__global__ __launch_bounds__(kBlock) void k_knn_snap_brute(uint32_t nA, const float4* __restrict__ apos, uint32_t nB,
                                                           const float4* __restrict__ bpos, int self, uint32_t k,
                                                           float cap2, uint32_t* __restrict__ out_idx,
                                                           float* __restrict__ out_dist) {
    __shared__ float4 tile[kBlock];
    __shared__ float box[kBlock / 64][6];
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = s < nA;
    const float4 q = active ? apos[s] : make_float4(0.f, 0.f, 0.f, 0.f);
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    TopK<KMAX> h;
    h.init(k);
    for (uint32_t base = 0; base < nB; base += kBlock) {
        const uint32_t t = base + threadIdx.x;
        float lo[3] = {INFINITY, INFINITY, INFINITY}, hi[3] = {-INFINITY, -INFINITY, -INFINITY};
        if (t < nB) {
            const float4 p = bpos[t];
            tile[threadIdx.x] = p;
            lo[0] = hi[0] = p.x;
            lo[1] = hi[1] = p.y;
            lo[2] = hi[2] = p.z;
        }
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                lo[a] = fminf(lo[a], __shfl_xor(lo[a], off, 64));
                hi[a] = fmaxf(hi[a], __shfl_xor(hi[a], off, 64));
            }
        }
        if (lane == 0) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                box[wave][a] = lo[a];
                box[wave][3 + a] = hi[a];
            }
        }
        __syncthreads();
        const uint32_t m = min((uint32_t)kBlock, nB - base);
        for (uint32_t c0 = 0; c0 < m; c0 += 64) {
            const int c = c0 >> 6;
            const float gx = fmaxf(fmaxf(box[c][0] - q.x, q.x - box[c][3]), 0.f);
            const float gy = fmaxf(fmaxf(box[c][1] - q.y, q.y - box[c][4]), 0.f);
            const float gz = fmaxf(fmaxf(box[c][2] - q.z, q.z - box[c][5]), 0.f);
            const float d2box = ((gx * gx) + (gy * gy)) + (gz * gz);
            const bool need = active && d2box <= fminf(cap2, h.worst());
            if (__ballot(need) == 0) continue;
            const uint32_t c1 = min(c0 + 64, m);
            if (need) {
                for (uint32_t j = c0; j < c1; ++j) {
                    const uint32_t b = base + j;
                    if (self && b == s) continue;
                    const float4 p = tile[j];
                    const float d2 = sqdist4(q.x, q.y, q.z, p);
                    const uint64_t kk = make_key(d2, id_of(p));
                    if (d2 <= cap2 && h.accepts(kk)) h.push(kk, b);
                }
            }
        }
        __syncthreads();
    }
    if (active) h.store(k, nA, s, out_idx, out_dist);
}

void launch_knn_snap_brute(gs_group_s* A, gs_knn_s* out, gs_group_s* B, float cap2) {
    const uint32_t nA = A->n;
    if (nA == 0) return;
    hipStream_t st = A->ctx->stream;
    const int self = A == B ? 1 : 0;
#define GS_SB(KM)                                                                                             \
    k_knn_snap_brute<KM><<<blocks_for(nA), kBlock, 0, st>>>(nA, A->spos.p, B->n, B->spos.p, self, out->k, cap2, \
                                                            out->idx_alt(), out->dist_alt())
    if (out->k <= 8) GS_SB(8);
    else if (out->k <= 16) GS_SB(16);
    else GS_SB(32);
#undef GS_SB
    GS_LAUNCHED("k_knn_snap_brute");
}

// Moves a [k][nA] result from the previous storage order of A and/or B to the
// current one: row s_old -> row_map[s_old], value v_old -> val_map[v_old].
__global__ __launch_bounds__(kBlock) void k_knn_remap(uint32_t nA, uint32_t k, const uint32_t* __restrict__ in_idx,
                                                      const float* __restrict__ in_dist,
                                                      const uint32_t* __restrict__ row_map,
                                                      const uint32_t* __restrict__ val_map,
                                                      uint32_t* __restrict__ out_idx, float* __restrict__ out_dist) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nA) return;
    const uint32_t r = row_map ? row_map[s] : s;
    for (uint32_t j = 0; j < k; ++j) {
        uint32_t v = in_idx[(size_t)j * nA + s];
        if (v != UINT32_MAX && val_map) v = val_map[v];
        out_idx[(size_t)j * nA + r] = v;
        out_dist[(size_t)j * nA + r] = in_dist[(size_t)j * nA + s];
    }
}

template <int KMAX, int RF, int BS = kBlock, int APL = 1>
static void run_binned_shape(gs_group_s* A, gs_knn_s* out, gs_group_s* B, uint32_t R0, float cap2,
                             const uint32_t* seed, const uint32_t* row_map, const uint32_t* val_map, bool dist) {
    const uint32_t nA = A->n;
    k_knn_binned<KMAX, RF, BS, APL><<<blocks_for(nA, BS * APL), BS, 0, A->ctx->stream>>>(
        nA, A->spos.p, B->n, B->spos.p, B->start.p, B->bin_grid, A == B ? 1 : 0, out->k, (int)R0, cap2, seed,
        row_map, val_map, out->idx_alt(), dist ? out->dist_alt() : nullptr,
        (std::isfinite(cap2) && B->bbox_fresh) ? B->bbox[B->bbox_cur].p : nullptr);
    GS_LAUNCHED("k_knn_binned");
}

template <int KMAX, int RF>
static void run_binned(gs_group_s* A, gs_knn_s* out, gs_group_s* B, uint32_t R0, float cap2, const uint32_t* seed,
                       const uint32_t* row_map, const uint32_t* val_map, bool d) {
    // The k = 7 flat search runs in the context's launch shape.
    if constexpr (KMAX == 7 && RF == 1) {
        const LaunchShape sh = A->ctx->shape[GS_SHAPE_KNN];
        if (sh.block == 128 && sh.apl == 1) return run_binned_shape<7, 1, 128, 1>(A, out, B, R0, cap2, seed, row_map, val_map, d);
        if (sh.block == 128 && sh.apl == 2) return run_binned_shape<7, 1, 128, 2>(A, out, B, R0, cap2, seed, row_map, val_map, d);
        if (sh.block == 256 && sh.apl == 2) return run_binned_shape<7, 1, 256, 2>(A, out, B, R0, cap2, seed, row_map, val_map, d);
        if (sh.block == 512 && sh.apl == 1) return run_binned_shape<7, 1, 512, 1>(A, out, B, R0, cap2, seed, row_map, val_map, d);
        if (sh.block == 512 && sh.apl == 2) return run_binned_shape<7, 1, 512, 2>(A, out, B, R0, cap2, seed, row_map, val_map, d);
    }
    run_binned_shape<KMAX, RF>(A, out, B, R0, cap2, seed, row_map, val_map, d);
}

template <int KMAX>
static void run_binned_rf(gs_group_s* A, gs_knn_s* out, gs_group_s* B, uint32_t R0, float cap2,
                          const uint32_t* seed, const uint32_t* row_map, const uint32_t* val_map, bool d) {
    // search_radius selects the fast path's column radius: 3 x 3 columns for
    // lattices whose cell is about the k-th neighbour distance, 5 x 5 for
    // finer lattices.
    // The 3 x 3 path addresses B with 32-bit byte offsets (fewer than 2^27
    // points); larger groups take the 5 x 5 walk, which is exact as well.
    // Lists longer than 8 keep the 3 x 3 path on every lattice: the 5 x 5
    // walk holding 16 keys exceeds the register file and spills.
    if ((R0 >= 2 && KMAX <= 8) || B->n >= kFlatMaxPoints)
        run_binned<KMAX, 2>(A, out, B, R0, cap2, seed, row_map, val_map, d);
    else run_binned<KMAX, 1>(A, out, B, R0, cap2, seed, row_map, val_map, d);
}

void launch_knn_binned(gs_group_s* A, gs_knn_s* out, gs_group_s* B, uint32_t search_radius, float cap2,
                       const uint32_t* seed, const uint32_t* row_map, const uint32_t* val_map, bool dist) {
    if (A->n == 0) return;
    // k = 7 (the parameter files' default) gets lists of exactly 7 entries:
    // one compare-exchange fewer per insertion than the 8-entry lists.
    if (out->k == 7) run_binned_rf<7>(A, out, B, search_radius, cap2, seed, row_map, val_map, dist);
    else if (out->k <= 8) run_binned_rf<8>(A, out, B, search_radius, cap2, seed, row_map, val_map, dist);
    else if (out->k <= 16) run_binned_rf<16>(A, out, B, search_radius, cap2, seed, row_map, val_map, dist);
    else run_binned_rf<32>(A, out, B, search_radius, cap2, seed, row_map, val_map, dist);
}

void launch_knn_brute(gs_group_s* A, gs_knn_s* out, gs_group_s* B, float cap2) {
    const uint32_t nA = A->n;
    if (nA == 0) return;
    hipStream_t st = A->ctx->stream;
    const int self = A == B ? 1 : 0;
#define GS_BRUTE(KM)                                                                                              \
    k_knn_brute<KM><<<blocks_for(nA), kBlock, 0, st>>>(nA, A->F(0), A->F(1), A->F(2), B->n, B->F(0), B->F(1),    \
                                                       B->F(2), B->ids(), self, out->k, cap2, out->idx_alt(),    \
                                                       out->dist_alt())
    if (out->k <= 8) GS_BRUTE(8);
    else if (out->k <= 16) GS_BRUTE(16);
    else GS_BRUTE(32);
#undef GS_BRUTE
    GS_LAUNCHED("k_knn_brute");
}

void launch_knn_remap(gs_knn_s* out, uint32_t nA, const uint32_t* row_map, const uint32_t* val_map) {
    if (nA == 0) return;
    k_knn_remap<<<blocks_for(nA), kBlock, 0, out->ctx->stream>>>(nA, out->k, out->idx(), out->dist(), row_map,
                                                                 val_map, out->idx_alt(), out->dist_alt());
    GS_LAUNCHED("k_knn_remap");
}

}  // namespace gs
