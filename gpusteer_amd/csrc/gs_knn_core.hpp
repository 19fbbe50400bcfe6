// Exact KNN building blocks shared by the stand-alone KNN kernels and the
// fused step kernel: register top-k list, lattice distance bounds, seeded
// search with deferred insertion.  Points are addressed by storage slot; ties
// are broken by agent index (the API's index), which the position snapshot
// carries in its fourth component.
#pragma once
#include "gs_internal.hpp"

namespace gs {

// Candidates a lane may defer (per-lane LDS buffer) before the wave merges
// them into the top-k lists.  Each entry is its 64-bit key (d2 bits, agent
// index) and its slot, in two planes: 12 bytes per entry.
constexpr int kKnnBuf = 12;
struct DeferBuf {
    uint2* key;     // [kKnnBuf][kBlock]
    uint32_t* slot; // [kKnnBuf][kBlock]
};
#define GS_DEFER_LDS                                          \
    __shared__ uint2 lds_defer_key[kKnnBuf * kBlock];         \
    __shared__ uint32_t lds_defer_slot[kKnnBuf * kBlock];     \
    const DeferBuf lbuf{lds_defer_key + threadIdx.x, lds_defer_slot + threadIdx.x};

// Sorted list of KMAX entries in registers: a 64-bit key (fp32 bits of d2 in
// the high word -- monotonic for d2 >= 0 -- and the agent index in the low
// word, so one unsigned compare orders by (d2, agent index)) plus the
// neighbour's storage slot.  For k < KMAX the first KMAX-k entries hold the
// sentinel key 0, below every real key, so the last entry is always the k-th
// best and no register is ever indexed dynamically.  Empty entries hold
// (+inf, UINT32_MAX).
constexpr uint64_t kKeyEmpty = (0x7F800000ull << 32) | 0xFFFFFFFFull;
__device__ __forceinline__ uint64_t make_key(float d2, uint32_t id) {
    return ((uint64_t)__float_as_uint(d2) << 32) | id;
}
template <int KMAX> struct TopK {
    uint64_t key[KMAX];
    uint32_t t[KMAX];
    __device__ __forceinline__ void init(uint32_t k) {
#pragma unroll
        for (int j = 0; j < KMAX; ++j) {
            key[j] = (j < KMAX - (int)k) ? 0ull : kKeyEmpty;
            t[j] = UINT32_MAX;
        }
    }
    __device__ __forceinline__ float worst() const { return __uint_as_float((uint32_t)(key[KMAX - 1] >> 32)); }
    __device__ __forceinline__ bool accepts(uint64_t kk) const { return kk < key[KMAX - 1]; }
    // Caller guarantees accepts(kk).  Branch-free compare-exchange ripple
    // from the back (few live registers).
    __device__ __forceinline__ void push(uint64_t kk, uint32_t tt) {
        key[KMAX - 1] = kk;
        t[KMAX - 1] = tt;
#pragma unroll
        for (int j = KMAX - 2; j >= 0; --j) {
            const bool sw = key[j + 1] < key[j];
            const uint64_t lo_k = sw ? key[j + 1] : key[j];
            const uint64_t hi_k = sw ? key[j] : key[j + 1];
            const uint32_t lo_t = sw ? t[j + 1] : t[j];
            const uint32_t hi_t = sw ? t[j] : t[j + 1];
            key[j] = lo_k;
            key[j + 1] = hi_k;
            t[j] = lo_t;
            t[j + 1] = hi_t;
        }
    }
    // The same insertion with every compare independent: kk goes to the
    // first position whose key exceeds it and every entry from there on
    // moves up one (8 64-bit compares instead of 13 for 7 entries).  Used by
    // the flat search's merge of 7-entry lists, where the extra live
    // compares fit the register budget (with 8 entries, or inlined at the
    // box walk's many sites, they spill).
    __device__ __forceinline__ void push_wide(uint64_t kk, uint32_t tt) {
        bool c[KMAX];
#pragma unroll
        for (int j = 0; j < KMAX; ++j) c[j] = kk < key[j];
#pragma unroll
        for (int j = KMAX - 1; j >= 1; --j) {
            key[j] = c[j - 1] ? key[j - 1] : (c[j] ? kk : key[j]);
            t[j] = c[j - 1] ? t[j - 1] : (c[j] ? tt : t[j]);
        }
        key[0] = c[0] ? kk : key[0];
        t[0] = c[0] ? tt : t[0];
    }
    // Writes column `row` of the slot-major [k][nA] output: neighbour slots
    // and distances (pads UINT32_MAX / FLT_MAX); no distances when out_dist
    // is NULL.
    __device__ __forceinline__ void store(uint32_t k, uint32_t nA, uint32_t row, uint32_t* __restrict__ out_idx,
                                          float* __restrict__ out_dist) const {
#pragma unroll
        for (int j = 0; j < KMAX; ++j) {
            if (j >= KMAX - (int)k) {
                const size_t o = (size_t)(j - (KMAX - (int)k)) * nA + row;
                const bool pad = t[j] == UINT32_MAX;
                out_idx[o] = t[j];
                if (out_dist) out_dist[o] = pad ? FLT_MAX : sqrtf(__uint_as_float((uint32_t)(key[j] >> 32)));
            }
        }
    }
};

__device__ __forceinline__ uint32_t id_of(const float4& p) { return __float_as_uint(p.w); }
__device__ __forceinline__ float sqdist4(float qx, float qy, float qz, const float4& p) {
    return sqdist(qx, qy, qz, p.x, p.y, p.z);
}
// d2 with fused multiply-adds: within a few ulps of sqdist4, for filters only
// (a key is always formed from sqdist4).
__device__ __forceinline__ float sqdist_filter(float qx, float qy, float qz, const float4& p) {
    const float dx = qx - p.x, dy = qy - p.y, dz = qz - p.z;
    return __builtin_fmaf(dx, dx, __builtin_fmaf(dy, dy, dz * dz));
}
// The filter's limit for an exact limit lim: widened by 2e-5 relative (far
// more than the few ulps between the two d2 forms) and 2^-100 absolute (their
// rounding in the subnormal range), so no candidate with d2 <= lim fails
// sqdist_filter <= filter_limit(lim).  -inf and +inf stay as they are.
__device__ __forceinline__ float filter_limit(float lim) { return __builtin_fmaf(lim, 1.0f + 2e-5f, 0x1p-100f); }

// Distance from coordinate q to the slab of cells [i0, i1] along one axis;
// boundary cells extend to infinity outward (points outside the lattice are
// clamped into them).  The rounding slack makes the gap an under-estimate.
__device__ __forceinline__ float axis_gap(float q, int i0, int i1, float o, float h, int d, float slack) {
    float gap = 0.f;
    if (i0 > 0) gap = fmaxf(gap, (o + (float)i0 * h) - q);
    if (i1 < d - 1) gap = fmaxf(gap, q - (o + (float)(i1 + 1) * h));
    return fmaxf(gap - slack, 0.f);
}

// Largest d2 from the query to the k neighbours of its previous exact result
// (a valid upper bound on the current k-th distance: they are k distinct
// candidates), or +inf when that result is incomplete.  The previous result
// may be held at the previous storage order of A (row = its old slot) and of
// B (val_map translates old B slots to current ones).  Unrolled over KMAX so
// all loads of a level are in flight together.
template <int KMAX>
__device__ __forceinline__ float seed_bound(const uint32_t* __restrict__ seed, uint32_t k, uint32_t nA, uint32_t row,
                                            uint32_t nB, const uint32_t* __restrict__ val_map, float qx, float qy,
                                            float qz, const float4* __restrict__ bpos) {
    // Entries j >= k repeat entry k-1, so every load is unconditional.
    uint32_t v[KMAX];
#pragma unroll
    for (int j = 0; j < KMAX; ++j) v[j] = seed[(size_t)min(j, (int)k - 1) * nA + row];
    bool ok = true;
#pragma unroll
    for (int j = 0; j < KMAX; ++j) ok = ok && v[j] < nB;
    if (!ok) return INFINITY;
    if (val_map) {
#pragma unroll
        for (int j = 0; j < KMAX; ++j) v[j] = val_map[v[j]];
    }
    float4 p[KMAX];
#pragma unroll
    for (int j = 0; j < KMAX; ++j) p[j] = bpos[v[j]];
    float m = 0.f;
#pragma unroll
    for (int j = 0; j < KMAX; ++j) m = fmaxf(m, sqdist4(qx, qy, qz, p[j]));
    return m;
}

// Distance from the query to the nearest cell outside the box of ring R
// around its cell (R cells each way in x and y, R * zr in z); `all` when that
// box covers the whole lattice.
__device__ __forceinline__ float outside_gap(const GridDev& g, int cx, int cy, int cz, int R, float qx, float qy,
                                             float qz, bool& all) {
    const int Z = R * g.zr;
    float dmin = INFINITY;
    all = true;
    if (cx - R > 0) { dmin = fminf(dmin, qx - (g.ox + (float)(cx - R) * g.hx)); all = false; }
    if (cx + R < g.dx - 1) { dmin = fminf(dmin, (g.ox + (float)(cx + R + 1) * g.hx) - qx); all = false; }
    if (cy - R > 0) { dmin = fminf(dmin, qy - (g.oy + (float)(cy - R) * g.hy)); all = false; }
    if (cy + R < g.dy - 1) { dmin = fminf(dmin, (g.oy + (float)(cy + R + 1) * g.hy) - qy); all = false; }
    if (cz - Z > 0) { dmin = fminf(dmin, qz - (g.oz + (float)(cz - Z) * g.hz)); all = false; }
    if (cz + Z < g.dz - 1) { dmin = fminf(dmin, (g.oz + (float)(cz + Z + 1) * g.hz) - qz); all = false; }
    return dmin - g.slack;
}
__device__ __forceinline__ bool exact_after(float lim, float dmin) {
    return dmin > 0.f && lim < dmin * dmin * (1.0f - 2e-5f);
}

// Merges every lane's deferred candidates into its list.  Called with the
// whole wavefront converged; iterates as long as any lane has entries left.
template <int KMAX>
__device__ __forceinline__ void flush_deferred(TopK<KMAX>& h, int& cnt, const DeferBuf& lbuf) {
    for (int e = 0; e < kKnnBuf; ++e) {
        if (!__any(e < cnt)) break;
        if (e < cnt) {
            const uint2 v = lbuf.key[e * kBlock];
            const uint64_t kk = ((uint64_t)v.x << 32) | v.y;
            if (h.accepts(kk)) h.push(kk, lbuf.slot[e * kBlock]);
        }
    }
    cnt = 0;
}

// Every point of the lattice, slots bstart[0] .. bstart[C] (C = the
// lattice's cells): the exact completion of a query whose next search box
// would hold more cells than the lattice holds points (sparse or tiny
// target sets on large lattices), at a cost bounded by the point count.
// Points already in the list (from the boxes scanned before) are skipped by
// slot; sentinel and empty entries hold no slot.
template <int KMAX>
__device__ void full_scan(TopK<KMAX>& h, float thr, float qx, float qy, float qz, uint32_t self_slot,
                          uint32_t t0, uint32_t t1, const float4* __restrict__ bpos) {
    float lim = fminf(thr, h.worst());
    for (uint32_t t = t0; t < t1; ++t) {
        const float4 p = bpos[t];
        const float d2 = sqdist4(qx, qy, qz, p);
        if (d2 <= lim && t != self_slot) {
            const uint64_t kk = make_key(d2, id_of(p));
            if (h.accepts(kk)) {
                bool dup = false;
#pragma unroll
                for (int j = 0; j < KMAX; ++j) dup = dup || h.t[j] == t;
                if (!dup) {
                    h.push(kk, t);
                    lim = fminf(thr, h.worst());
                }
            }
        }
    }
}
// Cells of the ring-R box (R cells each way in x and y, R * zr in z),
// clipped to the lattice, and whether that box holds more cells than twice
// the lattice's points (then a full scan is cheaper than the ring walk).
__device__ __forceinline__ bool box_beyond_points(const GridDev& g, int R, uint32_t points) {
    const uint64_t bx = (uint64_t)min(2 * R + 1, g.dx), by = (uint64_t)min(2 * R + 1, g.dy);
    const uint64_t bz = (uint64_t)min(2 * R * g.zr + 1, g.dz);
    return bx * by * bz > 2ull * points + 64ull;
}

// Generic shells R >= Rstart (rare path): scan the rings between the boxes of
// ring R - 1 and R (R cells each way in x and y, R * zr in z), each as
// contiguous z-runs of storage slots, skip runs farther than the bound, stop
// once the k-th best is below the distance to every unscanned cell.
template <int KMAX>
__device__ void shell_search(TopK<KMAX>& h, float thr, float qx, float qy, float qz, uint32_t self_slot,
                             const GridDev& g, int cx, int cy, int cz, int Rstart, int R0,
                             const float4* __restrict__ bpos, const uint32_t* __restrict__ bstart) {
    const float grow = 1.0f + 2e-5f;
    const uint32_t tlo = bstart[0], thi = bstart[(uint32_t)g.dx * (uint32_t)g.dy * (uint32_t)g.dz];
    for (int R = Rstart;; ++R) {
        if (R > 1 && box_beyond_points(g, R, thi - tlo)) {
            full_scan<KMAX>(h, thr, qx, qy, qz, self_slot, tlo, thi, bpos);
            return;
        }
        const int x0 = max(cx - R, 0), x1 = min(cx + R, g.dx - 1);
        const int y0 = max(cy - R, 0), y1 = min(cy + R, g.dy - 1);
        for (int ix = x0; ix <= x1; ++ix) {
            const bool xedge = (ix == cx - R) || (ix == cx + R);
            const float gx = axis_gap(qx, ix, ix, g.ox, g.hx, g.dx, g.slack);
            for (int iy = y0; iy <= y1; ++iy) {
                const bool edge = xedge || (iy == cy - R) || (iy == cy + R);
                const float gy = axis_gap(qy, iy, iy, g.oy, g.hy, g.dy, g.slack);
                const float gxy = gx * gx + gy * gy;
                const int Z = R * g.zr, Zin = (R - 1) * g.zr;
                for (int run = 0; run < (edge ? 1 : 2); ++run) {
                    int z0, z1;
                    if (edge) {
                        z0 = cz - Z;
                        z1 = cz + Z;
                    } else if (run == 0) {
                        z0 = cz - Z;
                        z1 = cz - Zin - 1;
                    } else {
                        z0 = cz + Zin + 1;
                        z1 = cz + Z;
                    }
                    z0 = max(z0, 0);
                    z1 = min(z1, g.dz - 1);
                    if (z0 > z1) continue;
                    const float gz = axis_gap(qz, z0, z1, g.oz, g.hz, g.dz, g.slack);
                    float lim = fminf(thr, h.worst());
                    if (gxy + gz * gz > lim * grow) continue;
                    const uint32_t c0 = cell_index(g, ix, iy, z0);
                    const uint32_t t0 = bstart[c0];
                    const uint32_t t1 = bstart[c0 + (uint32_t)(z1 - z0) + 1u];
                    for (uint32_t t = t0; t < t1; ++t) {
                        const float4 p = bpos[t];
                        const float d2 = sqdist4(qx, qy, qz, p);
                        if (d2 <= lim && t != self_slot) {
                            const uint64_t kk = make_key(d2, id_of(p));
                            if (h.accepts(kk)) {
                                h.push(kk, t);
                                lim = fminf(thr, h.worst());
                            }
                        }
                    }
                }
            }
        }
        if (R < R0) continue;
        bool all;
        const float dmin = outside_gap(g, cx, cy, cz, R, qx, qy, qz, all);
        if (all || exact_after(fminf(thr, h.worst()), dmin)) break;
    }
}

// Exact top-k of the query among the points of B.  B is in storage (cell)
// order: cell c holds slots [bstart[c], bstart[c+1]); bpos[t] =
// (x, y, z, agent index bits).  self_slot (A == B) is excluded.
//
// Fast path over the (2RF+1)^2 columns around the query's cell (RF = 1 or 2,
// a compile-time choice matched to the lattice).  For each column the z-run
// covers exactly the cells that intersect the bound's sphere (its half-chord
// sqrt(bound - gap_xy^2) around the query), capped to cz +- RF.  All runs are
// located first (their start loads in flight together); the column walk is
// unrolled and wave-uniform, and candidates are loaded four at a time.
// Candidates within the bound are appended to the lane's LDS buffer (lbuf,
// entries strided by kBlock) and merged into the list by the whole wavefront at column
// boundaries, so the insertion ripple runs about once per accepted candidate
// instead of once per candidate of any lane.  If the box of radius RF does
// not prove the list exact, the rare path continues with shells R > RF.
template <int KMAX, int RF>
__device__ __forceinline__ void knn_search(TopK<KMAX>& h, float thr, float qx, float qy, float qz,
                                           uint32_t self_slot, const GridDev& g, int R0,
                                           const float4* __restrict__ bpos, const uint32_t* __restrict__ bstart,
                                           const DeferBuf& lbuf) {
    constexpr int W = 2 * RF + 1;
    constexpr int NC = W * W;
    const int cx = axis_cell(qx, g.ox, g.sx, g.dx);
    const int cy = axis_cell(qy, g.oy, g.sy, g.dy);
    const int cz = axis_cell(qz, g.oz, g.sz, g.dz);
    const float grow = 1.0f + 2e-5f;
    float lim = fminf(thr, h.worst());
    const float bound = lim * grow;
    const int zlo = max(cz - RF * g.zr, 0), zhi = min(cz + RF * g.zr, g.dz - 1);
    // Locate the runs against the initial bound (tighter bounds found later
    // only make the scan do more than needed, never less).
    uint32_t r0[NC], r1[NC];
#pragma unroll
    for (int col = 0; col < NC; ++col) {
        const int ix = cx + col / W - RF, iy = cy + col % W - RF;
        r0[col] = r1[col] = 0u;
        if (ix >= 0 && ix < g.dx && iy >= 0 && iy < g.dy) {
            const float gx = axis_gap(qx, ix, ix, g.ox, g.hx, g.dx, g.slack);
            const float gy = axis_gap(qy, iy, iy, g.oy, g.hy, g.dy, g.slack);
            const float gxy = gx * gx + gy * gy;
            if (gxy <= bound) {
                int z0 = zlo, z1 = zhi;
                if (bound < INFINITY) {
                    // Cells of this column that can hold a point within the
                    // bound: |z - qz| <= sqrt(bound - gxy) (+ rounding slack).
                    const float half = sqrtf(bound - gxy) * grow + g.slack;
                    z0 = max(z0, axis_cell(qz - half, g.oz, g.sz, g.dz));
                    z1 = min(z1, axis_cell(qz + half, g.oz, g.sz, g.dz));
                }
                if (z0 <= z1) {
                    const uint32_t c0 = cell_index(g, ix, iy, z0);
                    r0[col] = bstart[c0];
                    r1[col] = bstart[c0 + (uint32_t)(z1 - z0) + 1u];
                }
            }
        }
    }
    int cnt = 0;
#pragma unroll
    for (int col = 0; col < NC; ++col) {
        const uint32_t t1 = r1[col];
        for (uint32_t t = r0[col]; t < t1; t += 4) {
            // Unconditional loads at clamped (in-range) slots keep all four in
            // flight together; only their use is predicated.
            float4 p[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) p[u] = bpos[min(t + u, t1 - 1u)];
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                if (t + u < t1) {
                    const float d2 = sqdist4(qx, qy, qz, p[u]);
                    if (d2 <= lim && t + u != self_slot) {
                        if (cnt < kKnnBuf) {
                            lbuf.key[cnt * kBlock] = make_uint2(__float_as_uint(d2), id_of(p[u]));
                            lbuf.slot[cnt * kBlock] = t + u;
                            ++cnt;
                        } else {
                            const uint64_t kk = make_key(d2, id_of(p[u]));
                            if (h.accepts(kk)) {
                                h.push(kk, t + u);
                                lim = fminf(thr, h.worst());
                            }
                        }
                    }
                }
            }
        }
        if (col < NC - 1 && __any(cnt >= kKnnBuf / 2)) {
            flush_deferred<KMAX>(h, cnt, lbuf);
            lim = fminf(thr, h.worst());
        }
    }
    flush_deferred<KMAX>(h, cnt, lbuf);
    bool all;
    const float dmin = outside_gap(g, cx, cy, cz, RF, qx, qy, qz, all);
    if (R0 <= RF && (all || exact_after(fminf(thr, h.worst()), dmin))) return;
    shell_search<KMAX>(h, thr, qx, qy, qz, self_slot, g, cx, cy, cz, RF + 1, R0, bpos, bstart);
}

// ---------------------------------------------------------------------------
// Flat-stream fast path over the 3 x 3 columns (lattices whose cell is about
// the k-th neighbour distance).
//
// The column walk above runs the columns in lockstep, so a wavefront pays,
// per column, for its lane with the longest z-run.  Here each lane instead
// walks its own run list -- the non-empty sphere-exact z-runs of its nine
// columns, compacted into LDS -- as one stream of candidate slots, eight per
// iteration (two chunks of four, each within one run).  Positions are read
// with buffer loads: one byte offset per chunk, the four candidates at
// immediate offsets, and the hardware range check makes a chunk that runs
// past the end of B read zeros instead of faulting (only candidates inside
// their run are used).  A candidate within the bound has its slot appended to
// the lane's LDS buffer -- the store is unconditional, only the count
// advances -- and its own slot passes like any other (d2 = 0); the merge
// recomputes each buffered candidate's key from its position, drops the
// query's own slot and inserts the rest.  The wavefront merges once any lane
// holds more than flat_buf - 8 entries, so an iteration can never overflow
// the buffer.
constexpr int kFlatBuf = 16;
// 16-entry lists defer 24: the first merge then mostly comes after the scan
// and takes the whole batch at once (their kernel holds 4 wavefronts per
// SIMD by its registers, which the larger pool still fits).
template <int KMAX> constexpr int flat_buf() { return KMAX == 16 ? 24 : kFlatBuf; }
constexpr int kFlatRuns = 10;  // nine columns, one of them split at the query
typedef __attribute__((address_space(3))) uint32_t lds_u32;
// One pool of flat_pool<KMAX>() words per lane, strided by the block size:
// the deferral buffer, the run list (first slots, then 16-bit lengths), and
// after the scan the cooperative shells' candidate triples (as many as the
// smallest pool, kFlatPool words, holds).
constexpr int kFlatPool = kFlatBuf + kFlatRuns + (kFlatRuns + 1) / 2;
template <int KMAX> constexpr int flat_pool() { return flat_buf<KMAX>() + kFlatRuns + (kFlatRuns + 1) / 2; }
constexpr int kCoopEntries = kFlatPool / 3;
struct FlatLds {
    uint32_t* pool;      // this lane's column
    uint32_t* slot;      // [flat_buf][block size]: deferred candidate slots
    uint32_t* run_t;     // [kFlatRuns][block size]: first slot of a non-empty run
    uint16_t* run_n;     // [kFlatRuns][block size]: its length
    uint32_t slot_addr;  // LDS byte address of the lane's first buffer entry
};
// For a block of BS threads searching KMAX-entry lists: lane columns strided
// by BS (the search functions' STRIDE template argument).
#define GS_FLAT_LDS_NK(BS, KMAX)                                                                        \
    __shared__ uint32_t lds_flat_pool[flat_pool<KMAX>() * (BS)];                                       \
    const FlatLds flds{lds_flat_pool + threadIdx.x, lds_flat_pool + threadIdx.x,                        \
                       lds_flat_pool + flat_buf<KMAX>() * (BS) + threadIdx.x,                           \
                       (uint16_t*)(lds_flat_pool + (flat_buf<KMAX>() + kFlatRuns) * (BS)) + threadIdx.x, \
                       (uint32_t)(size_t)(lds_u32*)(lds_flat_pool + threadIdx.x)};
#define GS_FLAT_LDS_N(BS) GS_FLAT_LDS_NK(BS, 8)
#define GS_FLAT_LDS GS_FLAT_LDS_N(kBlock)

// Buffer resource over the float4 snapshot of B (raw buffer, byte offsets,
// range-checked to n * 16 bytes).  Flat-path searches need n < 2^27.
constexpr uint32_t kFlatMaxPoints = 1u << 27;
__device__ __forceinline__ __amdgpu_buffer_rsrc_t pos_rsrc(const float4* p, uint32_t n) {
    return __builtin_amdgcn_make_buffer_rsrc((void*)p, (short)0, (int)(n * 16u), 0x00020000);
}
__device__ __forceinline__ float4 ld_pos(__amdgpu_buffer_rsrc_t r, uint32_t byte_off) {
    const auto v = __builtin_amdgcn_raw_buffer_load_b128(r, byte_off, 0, 0);
    return make_float4(__uint_as_float(v[0]), __uint_as_float(v[1]), __uint_as_float(v[2]),
                       __uint_as_float(v[3]));
}

// Cell of coordinate p along one axis for run bounds only: multiplication by
// the precomputed cells-per-length instead of lq's division, clamped.  Off by
// at most a few ulps of the cell coordinate from axis_cell, far inside the
// lattice's rounding slack that the callers add to every bound.
__device__ __forceinline__ int axis_cell_bound(float p, float o, float r, int d) {
    const float f = fminf(fmaxf((p - o) * r, 0.f), (float)(d - 1));
    return (int)f;
}

template <int KMAX, int STRIDE = kBlock>
__device__ __forceinline__ void flat_flush(TopK<KMAX>& h, int& cnt, const FlatLds& L, __amdgpu_buffer_rsrc_t rb,
                                           float qx, float qy, float qz, uint32_t self_slot, float thr,
                                           int first = 0) {
    // Entries from `first` on (the ones before were merged otherwise).
    for (int e0 = first; e0 < flat_buf<KMAX>(); e0 += 4) {
        if (!__any(e0 < cnt)) break;
        uint32_t sl[4];
        float4 p[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            sl[u] = L.slot[(e0 + u) * STRIDE];
            p[u] = ld_pos(rb, sl[u] * 16u);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            if (e0 + u < cnt && sl[u] != self_slot) {
                const float d2 = sqdist4(qx, qy, qz, p[u]);
                const uint64_t kk = make_key(d2, id_of(p[u]));
                if (d2 <= thr && h.accepts(kk)) {
                    if constexpr (KMAX == 7)
                        h.push_wide(kk, sl[u]);
                    else
                        h.push(kk, sl[u]);
                }
            }
        }
    }
    cnt = 0;
}

// The scan's only merge when the list is still empty and no lane of the
// wavefront holds more than eight entries -- the common case of a seeded
// bound with seven-entry lists: the seven previous neighbours, now and then
// one more.  The batch's exact keys are sorted at once by the 19-comparator
// network for eight inputs (Knuth, TAOCP vol. 3, sec. 5.3.4) and the seven
// smallest become the list: the same list as inserting them one by one
// (keys are distinct), for 19 compare-exchanges instead of up to 8 x 6.
// Missing entries, the query's own slot and keys beyond the bound sort last
// as empty entries.
__device__ __forceinline__ void sort_cx(uint64_t& ka, uint32_t& ta, uint64_t& kb, uint32_t& tb) {
    const bool sw = kb < ka;
    const uint64_t lo = sw ? kb : ka, hi = sw ? ka : kb;
    const uint32_t tl = sw ? tb : ta, th = sw ? ta : tb;
    ka = lo;
    kb = hi;
    ta = tl;
    tb = th;
}
template <int STRIDE = kBlock>
__device__ __forceinline__ void flat_sort_fresh(TopK<7>& h, int cnt, const FlatLds& L, __amdgpu_buffer_rsrc_t rb,
                                                float qx, float qy, float qz, uint32_t self_slot, float thr) {
    uint64_t key[8];
    uint32_t sl[8];
    float4 p[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        // Entries past cnt hold stale slots: their loads are range-checked
        // (zeros past the end of B) and their keys are discarded below.
        sl[u] = L.slot[u * STRIDE];
        p[u] = ld_pos(rb, sl[u] * 16u);
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        const float d2 = sqdist4(qx, qy, qz, p[u]);
        const bool use = u < cnt && sl[u] != self_slot && d2 <= thr;
        key[u] = use ? make_key(d2, id_of(p[u])) : kKeyEmpty;
        sl[u] = use ? sl[u] : UINT32_MAX;
    }
    constexpr int kNet[19][2] = {{0, 2}, {1, 3}, {4, 6}, {5, 7}, {0, 4}, {1, 5}, {2, 6}, {3, 7}, {0, 1}, {2, 3},
                                 {4, 5}, {6, 7}, {2, 4}, {3, 5}, {1, 4}, {3, 6}, {1, 2}, {3, 4}, {5, 6}};
#pragma unroll
    for (int c = 0; c < 19; ++c) sort_cx(key[kNet[c][0]], sl[kNet[c][0]], key[kNet[c][1]], sl[kNet[c][1]]);
#pragma unroll
    for (int j = 0; j < 7; ++j) {
        h.key[j] = key[j];
        h.t[j] = sl[j];
    }
}

// The first merge of a 16-entry list with k = 16 (no sentinel entries): the
// buffer's first sixteen entries become the list, sorted by Batcher's
// odd-even merge network for sixteen inputs (63 compare-exchanges, applied to
// the list's own registers) instead of up to 16 x 15 of insertion; the caller
// inserts the entries past the sixteenth.  Missing entries, the query's own
// slot and keys beyond the bound sort last as empty entries.
constexpr int kNet16[63][2] = {
    {0, 1},   {2, 3},   {0, 2},   {1, 3},   {1, 2},   {4, 5},   {6, 7},   {4, 6},   {5, 7},   {5, 6},
    {0, 4},   {2, 6},   {2, 4},   {1, 5},   {3, 7},   {3, 5},   {1, 2},   {3, 4},   {5, 6},   {8, 9},
    {10, 11}, {8, 10},  {9, 11},  {9, 10},  {12, 13}, {14, 15}, {12, 14}, {13, 15}, {13, 14}, {8, 12},
    {10, 14}, {10, 12}, {9, 13},  {11, 15}, {11, 13}, {9, 10},  {11, 12}, {13, 14}, {0, 8},   {4, 12},
    {4, 8},   {2, 10},  {6, 14},  {6, 10},  {2, 4},   {6, 8},   {10, 12}, {1, 9},   {5, 13},  {5, 9},
    {3, 11},  {7, 15},  {7, 11},  {3, 5},   {7, 9},   {11, 13}, {1, 2},   {3, 4},   {5, 6},   {7, 8},
    {9, 10},  {11, 12}, {13, 14}};
template <int STRIDE = kBlock>
__device__ __forceinline__ void flat_sort_fresh16(TopK<16>& h, int cnt, const FlatLds& L, __amdgpu_buffer_rsrc_t rb,
                                                  float qx, float qy, float qz, uint32_t self_slot, float thr) {
    static_assert(flat_buf<16>() >= 16, "the buffer fills the list");
#pragma unroll
    for (int h4 = 0; h4 < 16; h4 += 4) {
        uint32_t sl[4];
        float4 p[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            // Entries past cnt hold stale slots: their loads are range-checked
            // (zeros past the end of B) and their keys are discarded below.
            sl[u] = L.slot[(h4 + u) * STRIDE];
            p[u] = ld_pos(rb, sl[u] * 16u);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const float d2 = sqdist4(qx, qy, qz, p[u]);
            const bool use = h4 + u < cnt && sl[u] != self_slot && d2 <= thr;
            h.key[h4 + u] = use ? make_key(d2, id_of(p[u])) : kKeyEmpty;
            h.t[h4 + u] = use ? sl[u] : UINT32_MAX;
        }
    }
#pragma unroll
    for (int c = 0; c < 63; ++c) sort_cx(h.key[kNet16[c][0]], h.t[kNet16[c][0]], h.key[kNet16[c][1]], h.t[kNet16[c][1]]);
}

// ---------------------------------------------------------------------------
// Cooperative completion of one lane's query by its whole wavefront.
//
// A query whose ring-1 box (3 x 3 columns, zr cells each way in z) does not
// prove its list exact (its k-th neighbour lies beyond the box: sparse
// regions, a lattice finer than the neighbour distance) continues with the
// rings R = 2, 3, ... (boxes of R cells each way in x and y, R * zr in z).
// Done by its lane alone this is a serial walk of ~100 cells with a dependent
// pair of loads per cell, which holds its whole block for as long.  Instead
// the wavefront takes its failing queries one at a time: the cells of a shell
// are dealt round-robin to the 64 lanes, each lane writes the candidates of
// its cells within the query's current k-th distance to its column of the
// (then idle) flat-path LDS as (d2 bits, agent index, slot) triples, and the
// owner lane inserts them into its list.  Shells continue until the k-th best
// is below the distance to every unscanned cell.  A lane with more than
// kCoopEntries candidates in one shell sends the owner to the serial walk
// for the remaining shells (exact either way).  Call with the wavefront
// converged; `pool` is the calling lane's column of the flat-path LDS.
// Up to this many inexact queries per wavefront are completed cooperatively,
// one after another; more go to the lanes' own parallel walks.
constexpr int kCoopMaxQueries = 4;

template <int KMAX, int STRIDE = kBlock>
__device__ void coop_shells(TopK<KMAX>& h, int owner, float qx0, float qy0, float qz0, float thr0,
                            uint32_t self0, int cx0, int cy0, int cz0, uint32_t* pool, const GridDev& g,
                            const float4* __restrict__ bpos, const uint32_t* __restrict__ bstart) {
    const int lane = (int)(threadIdx.x & 63u);
    const float qx = __shfl(qx0, owner, 64), qy = __shfl(qy0, owner, 64), qz = __shfl(qz0, owner, 64);
    const float thr = __shfl(thr0, owner, 64);
    const uint32_t self_slot = (uint32_t)__shfl((int)self0, owner, 64);
    const int cx = __shfl(cx0, owner, 64), cy = __shfl(cy0, owner, 64), cz = __shfl(cz0, owner, 64);
    const float grow = 1.0f + 2e-5f;
    const uint32_t tlo = bstart[0], thi = bstart[(uint32_t)g.dx * (uint32_t)g.dy * (uint32_t)g.dz];
    for (int R = 2;; ++R) {
        // A box with more cells than the lattice has points: the owner scans
        // every point instead (the enumeration below visits the whole box).
        if ((uint64_t)(2 * R + 1) * (uint64_t)(2 * R + 1) * (uint64_t)(2 * R * g.zr + 1) > 2ull * (thi - tlo) + 64ull) {
            if (lane == owner) full_scan<KMAX>(h, thr, qx, qy, qz, self_slot, tlo, thi, bpos);
            return;
        }
        const float lim = fminf(thr, __shfl(h.worst(), owner, 64));
        int cnt = 0;
        bool overflow = false;
        // Ring R: the box of R cells each way in x and y and R * zr in z,
        // less the box of ring R - 1 (at most 2 * points + 64 cells here).
        const int W = 2 * R + 1, Z = R * g.zr, Zin = (R - 1) * g.zr, Wz = 2 * Z + 1;
        const int nbox = W * W * Wz;
        for (int i = lane; i < nbox; i += 64) {
            const int dx = i / (W * Wz) - R, dy = (i / Wz) % W - R, dz = i % Wz - Z;
            if (abs(dx) != R && abs(dy) != R && abs(dz) <= Zin) continue;
            const int ix = cx + dx, iy = cy + dy, iz = cz + dz;
            if (ix < 0 || ix >= g.dx || iy < 0 || iy >= g.dy || iz < 0 || iz >= g.dz) continue;
            const float gx = axis_gap(qx, ix, ix, g.ox, g.hx, g.dx, g.slack);
            const float gy = axis_gap(qy, iy, iy, g.oy, g.hy, g.dy, g.slack);
            const float gz = axis_gap(qz, iz, iz, g.oz, g.hz, g.dz, g.slack);
            if ((gx * gx + gy * gy) + gz * gz > lim * grow) continue;
            const uint32_t c = cell_index(g, ix, iy, iz);
            const uint32_t t1 = bstart[c + 1u];
            for (uint32_t t = bstart[c]; t < t1; ++t) {
                const float4 p = bpos[t];
                const float d2 = sqdist4(qx, qy, qz, p);
                if (d2 <= lim && t != self_slot) {
                    if (cnt < kCoopEntries) {
                        pool[(3 * cnt) * STRIDE] = __float_as_uint(d2);
                        pool[(3 * cnt + 1) * STRIDE] = id_of(p);
                        pool[(3 * cnt + 2) * STRIDE] = t;
                        ++cnt;
                    } else {
                        overflow = true;
                    }
                }
            }
        }
        if (__any(overflow)) {
            if (lane == owner) shell_search<KMAX>(h, thr, qx, qy, qz, self_slot, g, cx, cy, cz, R, 1, bpos, bstart);
            return;
        }
        for (uint64_t have = __ballot(cnt > 0); have != 0; have &= have - 1) {
            const int l = __ffsll((unsigned long long)have) - 1;
            const int c = __shfl(cnt, l, 64);
            if (lane == owner) {
                const uint32_t* col = pool + (l - lane);  // lane l's column
                for (int e = 0; e < c; ++e) {
                    const uint64_t kk = ((uint64_t)col[(3 * e) * STRIDE] << 32) | col[(3 * e + 1) * STRIDE];
                    if (h.accepts(kk)) h.push(kk, col[(3 * e + 2) * STRIDE]);
                }
            }
        }
        bool all;
        const float dmin = outside_gap(g, cx, cy, cz, R, qx, qy, qz, all);
        if (all || exact_after(fminf(thr, __shfl(h.worst(), owner, 64)), dmin)) break;
    }
}

// The slot run [x, y) of one column of the flat path's 3 x 3: the cells of
// the column (base = cell index of z = 0) whose z-range can hold a point
// within the bound, |z - qz| <= sqrt(bound - gxy) (widened by the growth
// factor, which covers the hardware square root's error, and the lattice
// slack), capped to [zlo, zhi]; (0, 0) for a column outside the lattice or
// beyond the bound.
__device__ __forceinline__ uint2 flat_column_run(bool valid, float gxy, float bound, float grow, float qz, int zlo,
                                                 int zhi, uint32_t base, const GridDev& g,
                                                 const uint32_t* __restrict__ bstart) {
    uint2 r = make_uint2(0u, 0u);
    if (valid && gxy <= bound) {
        int z0 = zlo, z1 = zhi;
        if (bound < INFINITY) {
            const float half = __builtin_amdgcn_sqrtf(bound - gxy) * grow + g.slack;
            z0 = max(z0, axis_cell_bound(qz - half, g.oz, g.rz, g.dz));
            z1 = min(z1, axis_cell_bound(qz + half, g.oz, g.rz, g.dz));
        }
        if (z0 <= z1) {
            const uint32_t c0 = base + (uint32_t)z0;
            r.x = bstart[c0];
            r.y = bstart[c0 + (uint32_t)(z1 - z0) + 1u];
        }
    }
    return r;
}

// One query of the flat path: position, bound (an upper bound on the k-th
// distance, or the cap), the query's own slot (UINT32_MAX: none), and whether
// the lane has a query (lanes without one scan nothing and only help complete
// other lanes' queries).
struct FlatQuery {
    float x, y, z, thr;
    uint32_t self;
    bool active;
};

// Requires 1 <= nB < kFlatMaxPoints and a minimum search radius of at most
// one cell.  Called by every lane of the wavefront.
template <int KMAX, int STRIDE = kBlock>
__device__ __forceinline__ void knn_search_flat(TopK<KMAX>& h, FlatQuery& Q, const GridDev& g, uint32_t nB,
                                                const float4* __restrict__ bpos, const uint32_t* __restrict__ bstart,
                                                const FlatLds& L) {
    const __amdgpu_buffer_rsrc_t rb = pos_rsrc(bpos, nB);
    const float grow = 1.0f + 2e-5f;
    int M = 0;
    bool long_run = false;
    {
        // The query's cell only centres the ring-1 box; every bound below is
        // geometric with the lattice slack, and the exactness test falls
        // back to the shell search, so the cheap mapping is exact in effect.
        const float qx = Q.x, qy = Q.y, qz = Q.z;
        const int cx = axis_cell_bound(qx, g.ox, g.rx, g.dx);
        const int cy = axis_cell_bound(qy, g.oy, g.ry, g.dy);
        const int cz = axis_cell_bound(qz, g.oz, g.rz, g.dz);
        const float bound = fminf(Q.thr, h.worst()) * grow;
        const int zlo = max(cz - g.zr, 0), zhi = min(cz + g.zr, g.dz - 1);
        // Per-axis squared gaps and cell-index bases of the columns at
        // offsets -1, 0, +1 (formed once per axis, not once per column).
        float gx2[3], gy2[3];
        bool vx[3], vy[3];
        uint32_t xbase[3], ybase[3];
#pragma unroll
        for (int o = 0; o < 3; ++o) {
            const int ix = cx + o - 1, iy = cy + o - 1;
            const float gx = axis_gap(qx, ix, ix, g.ox, g.hx, g.dx, g.slack);
            const float gy = axis_gap(qy, iy, iy, g.oy, g.hy, g.dy, g.slack);
            gx2[o] = gx * gx;
            gy2[o] = gy * gy;
            vx[o] = ix >= 0 && ix < g.dx;
            vy[o] = iy >= 0 && iy < g.dy;
            xbase[o] = (uint32_t)(ix * g.dy * g.dz);
            ybase[o] = (uint32_t)(iy * g.dz);
        }
        // Runs of the nine columns, plus a tenth: the query's own slot splits
        // the run that holds it (the centre column's, whenever the query lies
        // in the cell its coordinates map to), so the scan never buffers the
        // query itself; the merge still drops it wherever it appears.  Every
        // element is written unconditionally (selects, not branches), which
        // keeps the arrays in scalar registers.
        uint32_t r0[kFlatRuns], r1[kFlatRuns];
#pragma unroll
        for (int col = 0; col < 9; ++col) {
            const int ox = col / 3, oy = col % 3;
            const uint2 r = flat_column_run(vx[ox] && vy[oy], gx2[ox] + gy2[oy], bound, grow, qz, zlo, zhi,
                                            xbase[ox] + ybase[oy], g, bstart);
            r0[col] = r.x;
            r1[col] = r.y;
        }
        {
            const bool split = Q.self >= r0[4] && Q.self < r1[4];
            r0[9] = split ? Q.self + 1u : 0u;
            r1[9] = split ? r1[4] : 0u;
            r1[4] = split ? Q.self : r1[4];
        }
        // Compact the non-empty runs into the lane's LDS list.  A run longer
        // than the 16-bit length field sends the lane to the shell search
        // instead.
#pragma unroll
        for (int col = 0; col < kFlatRuns; ++col) {
            const uint32_t len = r1[col] - r0[col];
            long_run = long_run || len > 0xFFFFu;
            if (len != 0u) {
                L.run_t[M * STRIDE] = r0[col];
                L.run_n[M * STRIDE] = (uint16_t)len;
                ++M;
            }
        }
        // thr < 0: the query was found to have no candidate at all.
        if (long_run || !Q.active || Q.thr < 0.f) M = 0;
    }
    const uint32_t* rt = L.run_t;
    const uint16_t* rn = L.run_n;
    const float qx = Q.x, qy = Q.y, qz = Q.z, thr = Q.thr;
    const uint32_t self_slot = Q.self;
    const bool active = Q.active;
    const int cx = axis_cell_bound(qx, g.ox, g.rx, g.dx);
    const int cy = axis_cell_bound(qy, g.oy, g.ry, g.dy);
    const int cz = axis_cell_bound(qz, g.oz, g.rz, g.dz);
    float lim = fminf(thr, h.worst());
    // A chunk is (t, e): slots t .. min(t + 4, e) - 1 of the current run,
    // t < e; (0, 0) once the stream is exhausted.
    int m = 0;
    auto next_chunk = [&](uint32_t t, uint32_t e, uint32_t& tn, uint32_t& en) {
        tn = t + 4u;
        en = e;
        if (tn >= e) {
            if (m < M) {
                tn = rt[m * STRIDE];
                en = tn + rn[m * STRIDE];
                ++m;
            } else {
                tn = en = 0u;
            }
        }
    };
    uint32_t ta = 0u, ea = 0u, tb, eb;
    if (M > 0) {
        ta = rt[0];
        ea = ta + rn[0];
        m = 1;
    }
    next_chunk(ta, ea, tb, eb);
    // wa: LDS byte address of the lane's next free buffer entry (entries are
    // STRIDE words apart); the scan loop leaves the list's registers
    // untouched (merges happen between scan segments).
    constexpr uint32_t kEntry = STRIDE * 4u;
    uint32_t wa = L.slot_addr;
    const uint32_t wfull = L.slot_addr + (flat_buf<KMAX>() - 8) * kEntry;
    // The scan only filters: a candidate enters the buffer when its fused
    // d2 is within the bound widened by `grow` (more than the few ulps by
    // which the fused form can differ from the exact one), and the merge
    // recomputes the exact key and bound.
    float limf = filter_limit(lim);
    bool fresh = true;  // no merge yet: the list is empty
    for (;;) {
        while (__any(ta < ea) && !__any(wa > wfull)) {
            // Byte offsets of the two chunks; the per-candidate offsets are
            // immediates of the load instructions.
            uint32_t oa = ta * 16u, ob = tb * 16u;
            asm volatile("" : "+v"(oa), "+v"(ob));
            float4 pa[4], pb[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) pa[u] = ld_pos(rb, oa + 16u * u);
#pragma unroll
            for (int u = 0; u < 4; ++u) pb[u] = ld_pos(rb, ob + 16u * u);
            uint32_t tc, ec, td, ed;
            next_chunk(tb, eb, tc, ec);
            next_chunk(tc, ec, td, ed);
            const uint32_t na = ea - ta, nb = eb - tb;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const float d2 = sqdist_filter(qx, qy, qz, pa[u]);
                *(lds_u32*)(size_t)wa = ta + (uint32_t)u;
                wa += ((uint32_t)u < na && d2 <= limf) ? kEntry : 0u;
            }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const float d2 = sqdist_filter(qx, qy, qz, pb[u]);
                *(lds_u32*)(size_t)wa = tb + (uint32_t)u;
                wa += ((uint32_t)u < nb && d2 <= limf) ? kEntry : 0u;
            }
            ta = tc;
            ea = ec;
            tb = td;
            eb = ed;
        }
        int cnt = (int)((wa - L.slot_addr) / kEntry);
        const bool done = !__any(ta < ea);
        if constexpr (KMAX == 7) {
            if (fresh && done && !__any(cnt > 8)) {
                flat_sort_fresh<STRIDE>(h, cnt, L, rb, qx, qy, qz, self_slot, thr);
                break;
            }
        }
        // k = 16 (key[0] is a sentinel 0 for k < 16): the empty list takes
        // the first sixteen entries, sorted at once, and the rest by insertion.
        bool merged = false;
        if constexpr (KMAX == 16) {
            if (fresh && h.key[0] != 0ull) {
                flat_sort_fresh16<STRIDE>(h, cnt, L, rb, qx, qy, qz, self_slot, thr);
                flat_flush<KMAX, STRIDE>(h, cnt, L, rb, qx, qy, qz, self_slot, thr, 16);
                merged = true;
            }
        }
        if (!merged) flat_flush<KMAX, STRIDE>(h, cnt, L, rb, qx, qy, qz, self_slot, thr);
        fresh = false;
        wa = L.slot_addr;
        lim = fminf(thr, h.worst());
        limf = filter_limit(lim);
        if (done) break;
    }
    bool all;
    const float dmin = outside_gap(g, cx, cy, cz, 1, qx, qy, qz, all);
    const bool exact = !active || long_run || all || thr < 0.f || exact_after(fminf(thr, h.worst()), dmin);
    const uint64_t need = __ballot(!exact);
    if (__popcll(need) <= kCoopMaxQueries) {
        for (uint64_t left = need; left != 0; left &= left - 1) {
            const int owner = __ffsll((unsigned long long)left) - 1;
            coop_shells<KMAX, STRIDE>(h, owner, qx, qy, qz, thr, self_slot, cx, cy, cz, L.pool, g, bpos, bstart);
        }
    } else if (!exact) {
        // Many inexact queries in the wavefront (a lattice finer than the
        // neighbour distance or the cap): every lane walks its own rings, in
        // parallel.
        shell_search<KMAX>(h, thr, qx, qy, qz, self_slot, g, cx, cy, cz, 2, 1, bpos, bstart);
    }
    // A cell too full for the run list (more than 65535 points in one run):
    // the lane scans from its own cell outwards by itself, every lane of a
    // wavefront in parallel (a cooperative walk would give one lane a whole
    // cell per query).
    if (active && long_run) shell_search<KMAX>(h, thr, qx, qy, qz, self_slot, g, cx, cy, cz, 0, 1, bpos, bstart);
}

}  // namespace gs
