// Exact KNN of a small query group by a split exhaustive scan.
//
// With few queries (a few thousand: the MultiGroup plugin's groups) a
// one-lane-per-query search runs as a few dozen wavefronts, and a sparse or
// distant target group makes each of them walk hundreds of cells serially.
// Here B's snapshot is cut into S contiguous ranges, and each (query tile,
// range) pair is a block of its own, so nA / 64 x S wavefronts share the chip:
//
//   k_knn_split_part: every lane scans one range for its query (the range's
//   candidates are the same for the whole block, staged through LDS) and keeps the range's k best keys (d2 bits, agent
//   index) in a register list; the S partial lists go to scratch, [S][nA][k].
//   A previous exact result held by the container bounds the scan as in the
//   binned search, so a range far from the query's neighbours inserts
//   nothing.
//   k_knn_split_merge: one wavefront per query; lane l holds range l's sorted
//   list and k rounds of a 64-bit wavefront minimum take the k best heads.
//
// The keys are the binned search's, every candidate of B is tested once and
// the ranges are disjoint, so the result is the same exact k nearest (same
// cap rule d2 <= cap2, the query's own slot excluded for a group with itself).
#include "gs_knn_core.hpp"

namespace gs {

template <int KMAX>
__global__ __launch_bounds__(kBlock) void k_knn_split_part(uint32_t nA, const float4* __restrict__ apos, uint32_t nB,
                                                           const float4* __restrict__ bpos, int self, uint32_t k,
                                                           float cap2, uint32_t chunk,
                                                           const uint32_t* __restrict__ seed,
                                                           const uint32_t* __restrict__ row_map,
                                                           const uint32_t* __restrict__ val_map,
                                                           uint64_t* __restrict__ part_key,
                                                           uint32_t* __restrict__ part_slot) {
    __shared__ float4 tile[kBlock];
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = s < nA;
    const float4 q = active ? apos[s] : make_float4(0.f, 0.f, 0.f, 0.f);
    const uint32_t skip = (self && active) ? s : UINT32_MAX;
    const uint32_t b0 = blockIdx.y * chunk, b1 = min(b0 + chunk, nB);
    TopK<KMAX> h;
    h.init(k);
    // A candidate can enter only with d2 <= lim: the cap and, with a seed
    // (the container's previous exact result), the largest current d2 to
    // those k neighbours, a bound on the k-th distance over all of B; then
    // the list's own k-th d2.  Inactive lanes take none.
    float lim = -INFINITY;
    if (active)
        lim = fminf(cap2, seed ? seed_bound<KMAX>(seed, k, nA, row_map ? row_map[s] : s, nB, val_map, q.x, q.y,
                                                  q.z, bpos)
                               : INFINITY);
    // The fused d2 filters against the widened limit; a passing candidate's
    // tile index goes to the lane's LDS column (the cursor advances only on a
    // pass), and the wavefront inserts the columns' entries with exact keys
    // when some lane's column is nearly full and at the end of each tile: a
    // lane's insertions no longer hold the other 63 lanes for every
    // candidate some lane takes.
    __shared__ uint32_t buf[kSplitBuf][kBlock];
    float limf = filter_limit(lim);
    int cnt = 0;
    auto flush = [&](uint32_t base) {
        for (int e = 0; __any(e < cnt); ++e) {
            if (e < cnt) {
                const uint32_t j = buf[e][threadIdx.x];
                const float4 p = tile[j];
                const float d2 = sqdist4(q.x, q.y, q.z, p);
                const uint64_t kk = make_key(d2, id_of(p));
                if (d2 <= lim && h.accepts(kk)) {
                    h.push(kk, base + j);
                    lim = fminf(cap2, h.worst());
                }
            }
        }
        cnt = 0;
        limf = filter_limit(lim);
    };
    // The range goes through LDS a block-sized tile at a time (one coalesced
    // load per thread); every lane then reads the same entry (a broadcast).
    // Tile entries past the range hold NaN coordinates, which fail every
    // filter, so groups of eight need no bounds test; the eight filters are
    // branch-free and independent.
    //
    // Each 64-entry chunk of a tile carries its bounding box; a wavefront
    // skips a chunk when no lane's query is within its limit of the box.
    // The box distance uses the same per-axis subtractions and summation
    // order as a candidate's exact d2 and every step is monotone, so it is
    // never above a member's d2: the skip drops only candidates with
    // d2 > lim, which the exact test would reject.
    __shared__ float box[kBlock / 64][6];
    const float nan = __builtin_nanf("");
    const uint32_t lane = threadIdx.x & 63u, wave = threadIdx.x >> 6;
    for (uint32_t base = b0; base < b1; base += kBlock) {
        const uint32_t m = min((uint32_t)kBlock, b1 - base);
        __syncthreads();
        const float4 mine = threadIdx.x < m ? bpos[base + threadIdx.x] : make_float4(nan, nan, nan, 0.f);
        tile[threadIdx.x] = mine;
        // fminf / fmaxf ignore the NaN of entries past the range; a chunk
        // with no entry keeps the empty box (+inf, -inf), which no query
        // reaches.
        float lo[3] = {fminf(mine.x, INFINITY), fminf(mine.y, INFINITY), fminf(mine.z, INFINITY)};
        float hi[3] = {fmaxf(mine.x, -INFINITY), fmaxf(mine.y, -INFINITY), fmaxf(mine.z, -INFINITY)};
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
        for (uint32_t j0 = 0; j0 < m; j0 += 8u) {
            if ((j0 & 63u) == 0u) {
                const uint32_t c = j0 >> 6;
                const float gx = fmaxf(fmaxf(box[c][0] - q.x, q.x - box[c][3]), 0.f);
                const float gy = fmaxf(fmaxf(box[c][1] - q.y, q.y - box[c][4]), 0.f);
                const float gz = fmaxf(fmaxf(box[c][2] - q.z, q.z - box[c][5]), 0.f);
                const float d2box = ((gx * gx) + (gy * gy)) + (gz * gz);
                if (!__any(d2box <= lim)) {
                    j0 += 56u;  // the rest of the chunk (the loop adds 8)
                    continue;
                }
            }
            float4 p[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) p[u] = tile[j0 + (uint32_t)u];
            bool pass[8];
#pragma unroll
            for (int u = 0; u < 8; ++u)
                pass[u] = (sqdist_filter(q.x, q.y, q.z, p[u]) <= limf) & (base + j0 + (uint32_t)u != skip);
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                buf[cnt][threadIdx.x] = j0 + (uint32_t)u;
                cnt += pass[u] ? 1 : 0;
            }
            if (__any(cnt > kSplitBuf - 8)) flush(base);
        }
        flush(base);
    }
    if (!active) return;
    // [range][nA][k]: a wavefront's lists are one contiguous block, and the
    // merge's lane reads its range's list as k consecutive entries.
    const size_t o0 = ((size_t)blockIdx.y * nA + s) * k;
#pragma unroll
    for (int j = 0; j < KMAX; ++j) {
        if (j >= KMAX - (int)k) {
            const size_t o = o0 + (uint32_t)(j - (KMAX - (int)k));
            part_key[o] = h.key[j];
            part_slot[o] = h.t[j];
        }
    }
}

template <int KMAX>
__global__ __launch_bounds__(kBlock) void k_knn_split_merge(uint32_t nA, uint32_t S, uint32_t k,
                                                            const uint64_t* __restrict__ part_key,
                                                            const uint32_t* __restrict__ part_slot,
                                                            uint32_t* __restrict__ out_idx,
                                                            float* __restrict__ out_dist) {
    const uint32_t s = blockIdx.x * (kBlock / 64) + (threadIdx.x >> 6);
    const uint32_t lane = threadIdx.x & 63u;
    if (s >= nA) return;  // the whole wavefront
    uint64_t key[KMAX];
    uint32_t t[KMAX];
#pragma unroll
    for (int j = 0; j < KMAX; ++j) {
        key[j] = kKeyEmpty;
        t[j] = UINT32_MAX;
        if (lane < S && (uint32_t)j < k) {
            const size_t o = ((size_t)lane * nA + s) * k + (uint32_t)j;
            key[j] = part_key[o];
            t[j] = part_slot[o];
        }
    }
    for (uint32_t r = 0; r < k; ++r) {
        // The smallest head of the wavefront's lists (keys are distinct
        // except for the empty key, which is the largest).
        uint64_t m = key[0];
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) {
            const uint64_t o = __shfl_xor(m, off, 64);
            m = o < m ? o : m;
        }
        const uint64_t own = __ballot(key[0] == m);
        const int owner = __ffsll((unsigned long long)own) - 1;
        const uint32_t slot = __shfl(t[0], owner, 64);
        if (lane == 0) {
            const bool pad = m == kKeyEmpty;
            out_idx[(size_t)r * nA + s] = pad ? UINT32_MAX : slot;
            out_dist[(size_t)r * nA + s] = pad ? FLT_MAX : sqrtf(__uint_as_float((uint32_t)(m >> 32)));
        }
        const bool adv = (int)lane == owner;
#pragma unroll
        for (int j = 0; j < KMAX - 1; ++j) {
            key[j] = adv ? key[j + 1] : key[j];
            t[j] = adv ? t[j + 1] : t[j];
        }
        key[KMAX - 1] = adv ? kKeyEmpty : key[KMAX - 1];
        t[KMAX - 1] = adv ? UINT32_MAX : t[KMAX - 1];
    }
}

uint32_t knn_split_ranges(uint32_t nA, uint32_t nB) {
    if (nA == 0) return 0;
    const uint32_t waves = (nA + 63u) / 64u;
    uint32_t S = (kSplitWaves + waves - 1u) / waves;
    S = std::min<uint32_t>(S, std::min<uint32_t>(64u, nB / kSplitMinRange));
    return S;
}

template <int KMAX>
static void run_split(gs_group_s* A, gs_knn_s* out, gs_group_s* B, float cap2, uint32_t S, const uint32_t* seed,
                      const uint32_t* row_map, const uint32_t* val_map) {
    const uint32_t nA = A->n, nB = B->n, k = out->k;
    const uint32_t chunk = (nB + S - 1u) / S;
    S = (nB + chunk - 1u) / chunk;  // no empty range
    const size_t parts = (size_t)S * k * nA;
    out->part_key.reserve(parts);
    out->part_slot.reserve(parts);
    hipStream_t st = A->ctx->stream;
    const dim3 grid(blocks_for(nA), S);
    k_knn_split_part<KMAX><<<grid, kBlock, 0, st>>>(nA, A->spos.p, nB, B->spos.p, A == B ? 1 : 0, k, cap2, chunk,
                                                    seed, row_map, val_map, out->part_key.p, out->part_slot.p);
    GS_LAUNCHED("k_knn_split_part");
    k_knn_split_merge<KMAX><<<(nA + kBlock / 64 - 1) / (kBlock / 64), kBlock, 0, st>>>(
        nA, S, k, out->part_key.p, out->part_slot.p, out->idx_alt(), out->dist_alt());
    GS_LAUNCHED("k_knn_split_merge");
}

void launch_knn_split(gs_group_s* A, gs_knn_s* out, gs_group_s* B, float cap2, const uint32_t* seed,
                      const uint32_t* row_map, const uint32_t* val_map) {
    if (A->n == 0) return;
    const uint32_t S = knn_split_ranges(A->n, B->n);
    if (S < 2u) {
        launch_knn_snap_brute(A, out, B, cap2);
        return;
    }
    if (out->k == 7) run_split<7>(A, out, B, cap2, S, seed, row_map, val_map);
    else if (out->k <= 8) run_split<8>(A, out, B, cap2, S, seed, row_map, val_map);
    else if (out->k <= 16) run_split<16>(A, out, B, cap2, S, seed, row_map, val_map);
    else run_split<32>(A, out, B, cap2, S, seed, row_map, val_map);
}

}  // namespace gs
