// Internal declarations shared by the HIP translation units of libgpusteer.
#pragma once

#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>
#include <cstring>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/gpusteer.h"

namespace gs {

// ---------------------------------------------------------------------------
// Errors: internal code throws gs::Error; every C entry point converts it to a
// status and records the message for gs_last_error().
struct Error : std::runtime_error {
    int status;
    Error(int s, const std::string& m) : std::runtime_error(m), status(s) {}
};
void set_last_error(const std::string& m);
// Runs f, turning exceptions into a status (and the message for
// gs_last_error()): the shape of every C entry point.
template <class F> int guarded(F&& f) {
    try {
        f();
        return GS_OK;
    } catch (const Error& e) {
        set_last_error(e.what());
        return e.status;
    } catch (const std::bad_alloc&) {
        set_last_error("host allocation failed");
        return GS_ERR_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        set_last_error(e.what());
        return GS_ERR_DEVICE;
    }
}
[[noreturn]] void fail(int status, const std::string& msg);
// Live-handle registry: every handle the library hands out is registered
// until destroyed, so a destroyed or foreign pointer is rejected without
// being dereferenced.
void reg_add(const void* h);
void reg_del(const void* h);
bool reg_has(const void* h);
void hip_check(hipError_t e, const char* what);
#define GS_HIP(call) ::gs::hip_check((call), #call)
// Checks the launch of the preceding kernel.
#define GS_LAUNCHED(name) ::gs::hip_check(hipGetLastError(), name)

constexpr int kBlock = 256;
inline unsigned blocks_for(uint64_t n, unsigned per = kBlock) { return (unsigned)((n + per - 1) / per); }

// ---------------------------------------------------------------------------
// Device memory: a typed buffer that grows on demand.  Owned by handles.
void* dev_alloc(size_t bytes);
void dev_free(void* p);
template <class T> struct DevBuf {
    T* p = nullptr;
    size_t cap = 0;  // elements
    void reserve(size_t n) {
        if (n <= cap) return;
        T* q = (T*)dev_alloc(n * sizeof(T));
        if (p) dev_free(p);
        p = q;
        cap = n;
    }
    void release() {
        if (p) dev_free(p);
        p = nullptr;
        cap = 0;
    }
};

// ---------------------------------------------------------------------------
// Bin-lattice geometry as kernels see it.  Cell mapping follows src/lq.c:185-208
// (ix = (int)(((x - ox) / sx) * dx)), index ix*dy*dz + iy*dz + iz (:175-176);
// points outside the lattice are clamped into the boundary cells.
struct GridDev {
    float ox, oy, oz;
    float sx, sy, sz;
    int dx, dy, dz;
    float hx, hy, hz;  // cell size (s/d)
    float rx, ry, rz;  // cells per unit length (d/s), for conservative run bounds
    float slack;       // absolute rounding allowance of the mapping
    // z cells per ring of the search boxes: the box of ring R spans R cells
    // each way in x and y and R * zr cells in z, so lattices with finer z
    // cells (shorter z-runs to trim) keep boxes of about equal extent.
    int zr;
};

// Workgroups are dealt to the chip's 8 XCDs round-robin (block b runs on XCD
// b % 8), and each XCD has its own L2.  Kernels whose neighbouring blocks
// read the same data (agents in cell order reading nearby cells) use this
// bijection of [0, nb) instead of blockIdx.x, so each XCD takes one
// contiguous range of the work and the shared neighbourhoods stay in its L2.
constexpr uint32_t kXcds = 8;
__device__ __forceinline__ uint32_t xcd_block(uint32_t b, uint32_t nb) {
    const uint32_t x = b % kXcds, i = b / kXcds;
    const uint32_t q = nb / kXcds, r = nb % kXcds;  // XCD x runs q + (x < r) blocks
    return x * q + min(x, r) + i;
}

// Order-preserving map of floats to unsigned integers (min / max of the
// integers are min / max of the floats) and back.
__device__ __forceinline__ uint32_t ord_of(float f) {
    const uint32_t u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float float_of_ord(uint32_t o) {
    return __uint_as_float((o & 0x80000000u) ? (o & 0x7FFFFFFFu) : ~o);
}
// An empty box: min above max.
constexpr uint32_t kBoxEmptyMin = 0xFFFFFFFFu, kBoxEmptyMax = 0u;
// Adds the active lanes' points to the box bb[6]: wavefront min / max, then
// across the block's wavefronts in LDS, then one atomic per bound from the
// block.  Called by every thread of a kBlock-thread block (it synchronises
// the block).
__device__ __forceinline__ void box_accumulate(bool active, float x, float y, float z, uint32_t* __restrict__ bb) {
    __shared__ uint32_t part[kBlock / 64][6];
    uint32_t lo[3] = {active ? ord_of(x) : kBoxEmptyMin, active ? ord_of(y) : kBoxEmptyMin,
                      active ? ord_of(z) : kBoxEmptyMin};
    uint32_t hi[3] = {active ? lo[0] : kBoxEmptyMax, active ? lo[1] : kBoxEmptyMax, active ? lo[2] : kBoxEmptyMax};
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            lo[a] = min(lo[a], (uint32_t)__shfl_xor((int)lo[a], d, 64));
            hi[a] = max(hi[a], (uint32_t)__shfl_xor((int)hi[a], d, 64));
        }
    }
    if ((threadIdx.x & 63) == 0) {
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            part[threadIdx.x >> 6][a] = lo[a];
            part[threadIdx.x >> 6][3 + a] = hi[a];
        }
    }
    __syncthreads();
    if (threadIdx.x < 6) {
        const uint32_t j = threadIdx.x;
        uint32_t v = part[0][j];
#pragma unroll
        for (int w = 1; w < kBlock / 64; ++w) v = j < 3 ? min(v, part[w][j]) : max(v, part[w][j]);
        // A block without points holds the empty bounds, which change nothing.
        if (j < 3 && v != kBoxEmptyMin) atomicMin(&bb[j], v);
        if (j >= 3 && v != kBoxEmptyMax) atomicMax(&bb[j], v);
    }
}
// Whether no point of the box bb can lie within d2 <= cap2 of the query:
// the squared gap from the query to the box exceeds cap2 by more than the
// rounding of either d2 form (2e-5 relative).  False for an empty or
// non-finite box (nothing is skipped on doubt).
__device__ __forceinline__ bool box_beyond(const uint32_t* __restrict__ bb, float qx, float qy, float qz,
                                           float cap2) {
    if (bb[0] > bb[3]) return false;
    float g2 = 0.f;
    const float q[3] = {qx, qy, qz};
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        const float gap = fmaxf(fmaxf(float_of_ord(bb[a]) - q[a], q[a] - float_of_ord(bb[3 + a])), 0.f);
        g2 += gap * gap;
    }
    return g2 > cap2 * (1.0f + 2e-5f);
}

__device__ __forceinline__ int axis_cell(float p, float o, float s, int d) {
    if (!(p >= o)) return 0;
    if (p >= o + s) return d - 1;
    int i = (int)(((p - o) / s) * (float)d);
    return i < d ? i : d - 1;
}
__device__ __forceinline__ uint32_t cell_index(const GridDev& g, int ix, int iy, int iz) {
    return (uint32_t)((ix * g.dy * g.dz) + (iy * g.dz) + iz);
}

// ---------------------------------------------------------------------------
// fp32 vector algebra in the reference's evaluation order
// (include/OpenSteer/VectorUtils.cuh).  The library is compiled with
// -ffp-contract=off and correctly rounded division and square root.
struct f3 {
    float x, y, z;
};
__device__ __forceinline__ f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
__device__ __forceinline__ f3 operator+(f3 a, f3 b) { return mk3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ f3 operator-(f3 a, f3 b) { return mk3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ f3 operator-(f3 a) { return mk3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ f3 operator*(f3 a, float s) { return mk3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ f3 operator/(f3 a, float s) { return mk3(a.x / s, a.y / s, a.z / s); }
__device__ __forceinline__ float dot3(f3 a, f3 b) { return ((a.x * b.x) + (a.y * b.y)) + (a.z * b.z); }
__device__ __forceinline__ float len3(f3 a) { return sqrtf(dot3(a, a)); }
__device__ __forceinline__ f3 normalize3(f3 a) {
    const float l = len3(a);
    return (l > 0.f) ? a / l : a;
}
__device__ __forceinline__ f3 cross3(f3 a, f3 b) {
    return mk3((a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x));
}
__device__ __forceinline__ bool zero3(f3 a) { return a.x == 0.f && a.y == 0.f && a.z == 0.f; }
__device__ __forceinline__ f3 perp3(f3 v, f3 unit) { return v - unit * dot3(v, unit); }
__device__ __forceinline__ f3 truncate3(f3 v, float m) {
    const float m2 = m * m;
    const float l2 = dot3(v, v);
    if (l2 <= m2) return v;
    return v * (m / sqrtf(l2));
}
__device__ __forceinline__ float sqdist(float ax, float ay, float az, float bx, float by, float bz) {
    const float dx = ax - bx, dy = ay - by, dz = az - bz;
    return ((dx * dx) + (dy * dy)) + (dz * dz);
}

// Agent arrays as kernels take them (device pointers).
typedef gs_agents Agents;
__device__ __forceinline__ f3 ld3(const float* x, const float* y, const float* z, uint32_t i) {
    return mk3(x[i], y[i], z[i]);
}
__device__ __forceinline__ void st3(float* x, float* y, float* z, uint32_t i, f3 v) {
    x[i] = v.x;
    y[i] = v.y;
    z[i] = v.z;
}

}  // namespace gs

// ---------------------------------------------------------------------------
// Handles.
namespace gs {
// Launch shape of a tuned kernel: threads per block and agent slots per lane.
struct LaunchShape {
    uint32_t block = 256;
    uint32_t apl = 1;
};
}  // namespace gs

struct gs_context_s {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    gs::LaunchShape shape[2];  // GS_SHAPE_KNN, GS_SHAPE_FLOCK
    uint32_t magic = 0x67736378u;
};

struct gs_bins_s {
    gs_context ctx = nullptr;
    gs::GridDev g{};
    uint32_t cells = 0;
    uint32_t magic = 0x67736269u;
};

// Groups keep their state in storage order: every bin build physically
// reorders all live fields into cell order (ascending agent index inside a
// cell), so neighbours are neighbours in memory and every kernel reads and
// writes coalesced.  Agent indices of the API stay stable through an
// id <-> slot map; downloads translate back to index order.
constexpr int kFields = 23;  // gs_agents float fields
// Live fields per kind: agents all 23; obstacles centre (0..2) + radius (19);
// walls midpoint (0..2) + start/end/normal (3..11).
constexpr uint32_t kFieldsAgents = (1u << kFields) - 1u;
// Field groups of gs_agents (field order of include/gpusteer.h).
constexpr uint32_t kFPos = 0x7u;                      // px py pz
constexpr uint32_t kFKinematic = 0xFFFFu;             // pos fwd speed acc side up (0..15)
constexpr uint32_t kFSteer = 0x7u << 16;              // gx gy gz
constexpr uint32_t kFConstants = 0xFu << 19;          // radius maxSpeed maxForce mass
constexpr uint32_t kFieldsObstacles = 0x7u | (1u << 19);
constexpr uint32_t kFieldsWalls = 0xFFFu;
constexpr int kWallField = 3;  // walls: start = F(3..5), end = F(6..8), normal = F(9..11)

struct gs_group_s {
    gs_context ctx = nullptr;
    int kind = GS_KIND_AGENTS;
    uint32_t n = 0;
    uint32_t live = 0;               // live field mask
    // Fields whose every slot holds the same bit pattern (inv_val): a
    // permutation leaves them unchanged, so bin builds do not move them.
    // Positions are never marked.  Kept by every launch that writes fields
    // (mark_written / mark_integrated) and cleared when device pointers are
    // handed out.
    uint32_t inv = 0;
    float inv_val[kFields] = {};
    bool bits_zero = false;          // every slot's bits are 0
    uint64_t membership_epoch = 0;   // changes whenever the members change
    uint64_t order_epoch = 0;        // changes whenever the storage order changes
    uint64_t prev_order_epoch = 0;   // order before the last reorder (0: none)
    uint32_t magic = 0x67736772u;

    // Double-buffered SoA fields; cur[j] selects the current buffer.
    gs::DevBuf<float> fb[kFields][2];
    uint8_t fcur[kFields] = {};
    gs::DevBuf<uint32_t> bitsb[2];
    uint8_t bcur = 0;
    gs::DevBuf<uint32_t> idb[2];     // agent index of each storage slot
    uint8_t icur = 0;
    gs::DevBuf<uint32_t> slot_of_id; // storage slot of each agent index
    // The last reorder: new slot -> old slot and old slot -> new slot.
    gs::DevBuf<uint32_t> new_to_old, old_to_new;

    // Bin database (KNNDatabase, include/OpenSteer/CUDA/KNNDatabase.cuh:30-38):
    // after a build, cell c holds storage slots [start[c], start[c+1]).
    bool binned = false;
    bool snap_valid = false;  // spos holds the positions of the last build
    gs::GridDev bin_grid{};
    uint32_t bin_cells = 0;
    gs::DevBuf<uint32_t> key, rank, id_tmp, start, count;
    // Positions at the last build in storage order, packed (x, y, z, agent
    // index bits): what KNN searches, like the reference's positionSorted.
    gs::DevBuf<float4> spos;
    // (forward, speed) at the last lazy build, in storage order (fused step
    // only; see launch_bins_build).
    gs::DevBuf<float4> sfwd;
    gs::DevBuf<uint32_t> flag;  // small device scratch (finite check)
    // Bounding box of the snapshot positions (a capped search skips a query
    // whose cap cannot reach it): order-preserving integer encodings of min
    // x, y, z then max x, y, z.  Kept only once a capped search has targeted
    // the group (bbox_wanted): each build's key pass then accumulates into
    // bbox[bbox_cur] (cleared by the build before) and clears the other
    // buffer for the next build.  bbox_fresh: bbox[bbox_cur] covers the
    // current snapshot.  An accumulation that was not cleared only widens
    // the box, so a stale box is still a valid bound.
    gs::DevBuf<uint32_t> bbox[2];
    uint8_t bbox_cur = 0;
    bool bbox_wanted = false;
    bool bbox_fresh = false;

    float* F(int j) const { return fb[j][fcur[j]].p; }
    float* Falt(int j) const { return fb[j][fcur[j] ^ 1].p; }
    void flip(int j) { fcur[j] ^= 1; }
    uint32_t* bits() const { return bitsb[bcur].p; }
    uint32_t* ids() const { return idb[icur].p; }
    uint32_t* ids_alt() const { return idb[icur ^ 1].p; }
    gs_agents dev_agents() const;
    // Bookkeeping of inv / bits_zero after a launch that wrote `fields`
    // (and the bits, when `bits` is set).
    void mark_written(uint32_t fields, bool bits) {
        inv &= ~fields;
        if (bits) bits_zero = false;
    }
    // After an integrate: kinematic fields rewritten, steering and bits zero.
    void mark_integrated() {
        inv &= ~kFKinematic;
        inv |= kFSteer;
        for (int j = 16; j < 19; ++j) inv_val[j] = 0.f;
        bits_zero = true;
    }
};

struct gs_knn_s {
    gs_context ctx = nullptr;
    uint32_t n = 0, k = 0;
    // Slot-major [k][n] in storage slots: row = slot of the query in A,
    // value = slot of the neighbour in B.  Double-buffered so a call can read
    // its previous result while writing the new one.
    gs::DevBuf<uint32_t> idxb[2];
    gs::DevBuf<float> distb[2];
    uint8_t cur = 0;
    // The contents are the exact result for (a, b) at these epochs.
    bool valid = false;
    const gs_group_s* a = nullptr;
    const gs_group_s* b = nullptr;
    uint64_t a_members = 0, b_members = 0, a_order = 0, b_order = 0;
    // Partial lists of the split exhaustive search, [ranges][n][k].
    gs::DevBuf<uint64_t> part_key;
    gs::DevBuf<uint32_t> part_slot;
    uint32_t magic = 0x67736b6eu;

    uint32_t* idx() const { return idxb[cur].p; }
    float* dist() const { return distb[cur].p; }
    uint32_t* idx_alt() const { return idxb[cur ^ 1].p; }
    float* dist_alt() const { return distb[cur ^ 1].p; }
};

// A polyline path (PolylinePathwayCUDA, CUDA/PolylinePathwayCUDA.cuh:11-68):
// host copies and one device buffer holding 7 arrays of m floats (point x, y,
// z, segment direction x, y, z, segment length; entry 0 of the last four is 0).
struct gs_path_s {
    gs_context ctx = nullptr;
    uint32_t m = 0;
    float radius = 0.f, total = 0.f;
    int cyclic = 0;
    std::vector<float> pts, nrm, len;  // [m][3], [m][3], [m]
    gs::DevBuf<float> dev;
    uint32_t magic = 0x67737061u;
};

namespace gs {
// Field pointers of one reorder (current -> alternate buffers).
struct FieldSet {
    const float* src[kFields];
    float* dst[kFields];
    int nf;
};

// Wave-aggregated histogram + rank: lanes holding the same cell in a run of
// consecutive lanes (storage is in cell order from the previous build) take
// one atomic per run; rank[i] = the lane's arrival position in its cell.
// Inactive lanes must hold a key no cell has, distinct per lane.
// The step's world wrap of one point: GS_WRAP_BOX is WrapWorldKernel's face
// teleport on all three axes (CUDA/WrapWorldCUDAKernel.cu:25-35),
// GS_WRAP_SPHERE sphericalWrapAround (plugins/Boids.cpp:200-206); any other
// shape leaves the point.
__device__ __forceinline__ void wrap_point(int shape, float wcx, float wcy, float wcz, float wsx, float wsy, float wsz,
                                           float& x, float& y, float& z) {
    if (shape == GS_WRAP_BOX) {
        const float hx = wsx * 0.5f, hy = wsy * 0.5f, hz = wsz * 0.5f;
        if (x < wcx - hx) x = wcx + hx;
        else if (x > wcx + hx) x = wcx - hx;
        if (y < wcy - hy) y = wcy + hy;
        else if (y > wcy + hy) y = wcy - hy;
        if (z < wcz - hz) z = wcz + hz;
        else if (z > wcz + hz) z = wcz - hz;
    } else if (shape == GS_WRAP_SPHERE) {
        const f3 off = mk3(x - wcx, y - wcy, z - wcz);
        const float r = len3(off);
        if (r > wsx) {
            const f3 p = mk3(x, y, z) + ((off / r) * wsx) * -2.f;
            x = p.x;
            y = p.y;
            z = p.z;
        }
    }
}

__device__ __forceinline__ void count_and_rank(bool active, uint32_t c, uint32_t i, uint32_t* __restrict__ count,
                                               uint32_t* __restrict__ rank) {
    const int lane = threadIdx.x & 63;
    const uint32_t prev = __shfl_up(c, 1, 64);
    const bool head = lane == 0 || prev != c;
    const uint64_t heads = __ballot(head);
    const uint64_t upto = lane == 63 ? ~0ull : ((1ull << (lane + 1)) - 1ull);
    const int head_lane = 63 - __clzll((unsigned long long)(heads & upto));
    const uint64_t after = heads & ~upto;
    const int run_end = after ? (__ffsll((unsigned long long)after) - 1) : 64;
    uint32_t base = 0;
    if (head && active) base = atomicAdd(&count[c], (uint32_t)(run_end - lane));
    base = __shfl(base, head_lane, 64);
    if (active) rank[i] = base + (uint32_t)(lane - head_lane);
}

// Exclusive scan of C counts into start[0..C] (start needs C + 1 + ceil(C /
// 2048) entries: the tile sums live after start[C]); clears the counts.
void launch_scan_counts(uint32_t* count, uint32_t* start, uint32_t C, hipStream_t st);
// Stable scatter of n keyed slots into cell order and the gather of every
// field in fs, bits and ids into the new order; writes the order maps and
// the packed position snapshot.
void launch_scatter_reorder(uint32_t n, const uint32_t* key, const uint32_t* rank, const uint32_t* start,
                            const uint32_t* ids, uint32_t* ids_new, uint32_t* id_tmp,
                            uint32_t* new_to_old, uint32_t* old_to_new, const uint32_t* bits, uint32_t* bits_new,
                            const FieldSet& fs, float4* spos, hipStream_t st, float4* sfwd = nullptr);

// Host-side launchers (one per operation) implemented in the .hip files.
void launch_fill_u32(uint32_t* p, uint32_t n, uint32_t v, hipStream_t st);
void launch_iota(uint32_t* p, uint32_t n, hipStream_t st);
// Bin build + reorder of g into cell order (optionally wrapping positions
// first); also writes the position snapshot spos.
// keys_ready: the previous step's flock kernel already did the key pass
// (launch_flock_update's next_bins) for this lattice and wrap.
void launch_bins_build(gs_group_s* g, const gs_bins_s* b, int wrap_shape, const float* wc, const float* ws,
                       bool lazy = false, bool keys_ready = false);
// KNN into out's alternate buffer (caller flips).  row_map / val_map translate
// a seed held at the previous storage order of A / B (NULL: same order).
// cap2: only points with d2 <= cap2 qualify (+inf: no cap).
// dist false: the alternate distance buffer is left unwritten (a caller that
// reads only the indices before the next search).
void launch_knn_binned(gs_group_s* A, gs_knn_s* out, gs_group_s* B, uint32_t search_radius, float cap2,
                       const uint32_t* seed, const uint32_t* row_map, const uint32_t* val_map, bool dist = true);
void launch_knn_brute(gs_group_s* A, gs_knn_s* out, gs_group_s* B, float cap2);
// The binned search's semantics (position snapshots of both builds) by an
// exhaustive scan, for small B.
void launch_knn_snap_brute(gs_group_s* A, gs_knn_s* out, gs_group_s* B, float cap2);
constexpr uint32_t kSmallTargetSet = 256;
// A sparse target set (fewer than one point per kSparseCells cells, at most
// kExhaustiveMax points) is also scanned exhaustively when a search that
// finds nothing would cover more than nB / 4 cells (the cap's box, or the
// whole lattice without a cap): nB distance tests beat the shell walk over
// empty cells.
constexpr uint32_t kExhaustiveMax = 4096;
constexpr uint32_t kSparseCells = 16;
// A small query group's search of another small group (at most kSplitMaxA
// queries, at most kSplitMaxPairs query-candidate pairs, B at most
// kSplitMaxB points) whose reach covers more than kSplitMinCells cells is
// done by the split exhaustive scan: B cut into ranges of at least
// kSplitMinRange points, about kSplitWaves wavefronts in all, at most 64
// ranges.
constexpr uint32_t kSplitMaxA = 16384;
constexpr uint32_t kSplitMaxB = 16384;
constexpr uint64_t kSplitMaxPairs = 1ull << 26;
constexpr uint32_t kSplitWaves = 4096;
constexpr uint32_t kSplitMinRange = 64;
constexpr uint64_t kSplitMinCells = 64;
constexpr int kSplitBuf = 16;  // deferred entries per lane
// Number of ranges for nA queries over nB points (< 2: no split).
uint32_t knn_split_ranges(uint32_t nA, uint32_t nB);
// The split exhaustive scan (falls back to the unsplit one below 2 ranges);
// seed / row_map / val_map as for launch_knn_binned.
void launch_knn_split(gs_group_s* A, gs_knn_s* out, gs_group_s* B, float cap2, const uint32_t* seed,
                      const uint32_t* row_map, const uint32_t* val_map);
// Rewrites out's contents from the previous storage order of A / B to the
// current one (into the alternate buffer; caller flips).
void launch_knn_remap(gs_knn_s* out, uint32_t nA, const uint32_t* row_old_to_new, const uint32_t* val_old_to_new);
void launch_flock(gs_group_s* A, gs_knn_s* knn, gs_group_s* B, const gs_flock_term* sep, const gs_flock_term* ali,
                  const gs_flock_term* coh, int mode);
void launch_seek(gs_group_s* A, const float* target, float w, uint32_t mask, int mode);
void launch_avoid_obstacles(gs_group_s* A, gs_group_s* O, gs_knn_s* knn, float min_time, float w, uint32_t mask);
void launch_avoid_walls(gs_group_s* A, gs_knn_s* knn, gs_group_s* W, float min_time, float w, uint32_t mask);
void launch_update(gs_group_s* A, float dt, int mode);
void launch_update_anti_penetration(gs_group_s* A, gs_knn_s* knn, gs_group_s* W, float dt, int closest);
void launch_anti_penetration_agents(gs_group_s* A, gs_knn_s* knn, gs_group_s* B, uint32_t mask);
void launch_follow_path(gs_group_s* A, const gs_path_s* p, float prediction_time, float w, uint32_t mask);
void launch_avoid_neighbors(gs_group_s* A, gs_knn_s* knn, gs_group_s* B, float min_time, float min_sep,
                            int avoid_close, float w, uint32_t mask, int mode);
// Per-cell extraction of a group's last bin build, in KNNBinDataV2's cell
// numbering: agents per cell (uint32 [C]) and mean velocity (float [C][3]).
void launch_cell_density(gs_group_s* g, uint32_t* out);
void launch_cell_velocity(gs_group_s* g, float* out);
void write_cell_density(const GridDev& G, const std::vector<uint32_t>& d, const char* prefix,
                        const uint32_t* selected, uint32_t n_selected, const char* selected_prefix, int all_rows);
void write_cell_velocity(const GridDev& G, const std::vector<float>& v, const char* prefix,
                         const uint32_t* selected, uint32_t n_selected, const char* selected_prefix, int all_rows);
void launch_wrap(gs_group_s* A, int shape, const float* c, const float* s);
void launch_clear_bits(gs_group_s* A);
void launch_check_finite(gs_group_s* A, uint32_t* first_bad_slot_host);
// Flocking + integrate of the Boids step from a current KNN result of A with
// itself; new positions / forwards go to the alternate field buffers (caller
// flips fields 0..5).
// next_bins: the same call steps again on this lattice, so this step also
// does the next step's key pass (wrap, cell keys, counts) for the positions
// it writes, and the next build is called with keys_ready.
void launch_flock_update(gs_group_s* A, gs_knn_s* knn, const gs_boids_params* p, bool lazy,
                         const gs_bins_s* next_bins = nullptr);
// Whether the fused step may build lazily (steering and bits zero, per-agent
// constants uniform).
bool flock_step_lazy_ok(const gs_group_s* A);
// Inverts the id map into slot_of_id.
void launch_invert_ids(gs_group_s* g);
// Allocation-free KNN (gs_noalloc.hip): locates the fixed scratch array on
// `device` (current), and runs the search of n device points into it.
void noalloc_init_device(int device);
void noalloc_find_knn(gs_context ctx, const gs_bins_s* bins, uint32_t n, const float* px, const float* py,
                      const float* pz, uint32_t k, uint32_t* idx, float* dist);
size_t noalloc_scratch_bytes();
}  // namespace gs
