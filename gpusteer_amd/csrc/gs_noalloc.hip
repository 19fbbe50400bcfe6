// Exact KNN of a point set in device memory without any allocation
// (gs_find_knn_noalloc).  The bin database (keys, ranks, counts, cell starts,
// order maps), the sorted position snapshot and the slot-space lists of one
// call live in `scratch`, one fixed device array of the library; the caller
// passes only the points and the output arrays.  The kernels are the ones of
// gs_update_knn_database + gs_find_knn (counting-sort build, exact binned
// search), run on a transient group and container whose buffers point into
// the array, plus one kernel here that writes the result in point indices.
#include <mutex>

#include "gs_internal.hpp"

namespace gs {

namespace {

constexpr size_t kN = GS_NOALLOC_MAX_POINTS;
constexpr size_t kC = GS_NOALLOC_MAX_CELLS;
constexpr size_t kK = GS_NOALLOC_MAX_K;
// Cell starts: C + 1 entries, then one partial sum per scan tile of 2048
// cells (launch_scan_counts).
constexpr size_t kStartWords = kC + 1 + (kC + 2047) / 2048;

constexpr size_t align256(size_t b) { return (b + 255) & ~size_t(255); }

// Byte offsets of the regions of the array.
constexpr size_t kOffSpos = 0;                                     // float4 [N]
constexpr size_t kOffU32 = kOffSpos + align256(16 * kN);           // 7 x uint32 [N]
constexpr size_t kU32Arrays = 7;  // key rank id_tmp new_to_old old_to_new ids ids_alt
constexpr size_t kOffPos = kOffU32 + kU32Arrays * align256(4 * kN); // 3 x float [N] (sorted positions)
constexpr size_t kOffCount = kOffPos + 3 * align256(4 * kN);        // uint32 [C]
constexpr size_t kOffStart = kOffCount + align256(4 * kC);          // uint32 [kStartWords]
constexpr size_t kOffIdx = kOffStart + align256(4 * kStartWords);   // uint32 [K][N] (slot space)
constexpr size_t kOffDist = kOffIdx + align256(4 * kK * kN);        // float [K][N]
constexpr size_t kScratchBytes = kOffDist + align256(4 * kK * kN);

}  // namespace

// The array itself.  Its name is a host-side stand-in, never a device
// address: host code uses only the per-device address that
// hipGetSymbolAddress returns (each device holds its own copy).
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static void* g_scratch_addr[kMaxDevices] = {};

// Called once per device by gs_context_create, with that device current.
void noalloc_init_device(int device) {
    if (device < 0 || device >= kMaxDevices) fail(GS_ERR_UNSUPPORTED, "device index beyond the library's table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_addr[device]) return;
    void* p = nullptr;
    GS_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    if (!p) fail(GS_ERR_DEVICE, "hipGetSymbolAddress returned no address for the scratch array");
    g_scratch_addr[device] = p;
}

static unsigned char* scratch_of(int device) {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    unsigned char* p = (device >= 0 && device < kMaxDevices) ? (unsigned char*)g_scratch_addr[device] : nullptr;
    if (!p) fail(GS_ERR_DEVICE, "scratch array of this device was not set up by gs_context_create");
    return p;
}

size_t noalloc_scratch_bytes() { return kScratchBytes; }

// Slot-space lists -> point indices: row = the query's point index, values
// = neighbour point indices (the .w of the snapshot holds each slot's point
// index).
__global__ __launch_bounds__(kBlock) void k_noalloc_to_points(uint32_t n, uint32_t k,
                                                              const float4* __restrict__ spos,
                                                              const uint32_t* __restrict__ sidx,
                                                              const float* __restrict__ sdist,
                                                              uint32_t* __restrict__ idx, float* __restrict__ dist) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const uint32_t q = __float_as_uint(spos[s].w);
    for (uint32_t j = 0; j < k; ++j) {
        const uint32_t v = sidx[(size_t)j * n + s];
        idx[(size_t)j * n + q] = v == UINT32_MAX ? UINT32_MAX : __float_as_uint(spos[v].w);
        dist[(size_t)j * n + q] = sdist[(size_t)j * n + s];
    }
}

template <class T> static void point_into(DevBuf<T>& b, unsigned char* base, size_t off, size_t cap) {
    b.p = (T*)(base + off);
    b.cap = cap;  // reserve() up to cap is a no-op: nothing is allocated
}

void noalloc_find_knn(gs_context ctx, const gs_bins_s* bins, uint32_t n, const float* px, const float* py,
                      const float* pz, uint32_t k, uint32_t* idx, float* dist) {
    if (n > kN) fail(GS_ERR_UNSUPPORTED, "gs_find_knn_noalloc: more than GS_NOALLOC_MAX_POINTS points");
    if ((size_t)bins->cells > kC) fail(GS_ERR_UNSUPPORTED, "gs_find_knn_noalloc: more than GS_NOALLOC_MAX_CELLS cells");
    if (k == 0 || k > kK) fail(GS_ERR_UNSUPPORTED, "gs_find_knn_noalloc: k must be 1..GS_NOALLOC_MAX_K");
    if (n == 0) return;
    if (!px || !py || !pz || !idx || !dist) fail(GS_ERR_INVALID_ARGUMENT, "gs_find_knn_noalloc: NULL array");
    unsigned char* base = scratch_of(ctx->device);
    hipStream_t st = ctx->stream;

    // A transient group over the caller's positions: only the positions are
    // live; the build writes its sorted copy, snapshot and maps into the
    // array.  Nothing here owns memory (DevBuf frees only on release()).
    gs_group_s g;
    g.ctx = ctx;
    g.kind = GS_KIND_AGENTS;
    g.n = n;
    g.live = kFPos;
    g.bits_zero = true;
    const float* in[3] = {px, py, pz};
    for (int a = 0; a < 3; ++a) {
        g.fb[a][0].p = const_cast<float*>(in[a]);  // read only: the build writes the alternate buffer
        g.fb[a][0].cap = n;
        point_into(g.fb[a][1], base, kOffPos + (size_t)a * align256(4 * kN), kN);
    }
    DevBuf<uint32_t>* u32[kU32Arrays] = {&g.key, &g.rank, &g.id_tmp, &g.new_to_old, &g.old_to_new,
                                         &g.idb[0], &g.idb[1]};
    for (size_t j = 0; j < kU32Arrays; ++j) point_into(*u32[j], base, kOffU32 + j * align256(4 * kN), kN);
    point_into(g.spos, base, kOffSpos, kN);
    point_into(g.count, base, kOffCount, kC);
    point_into(g.start, base, kOffStart, kStartWords);

    // The array keeps the previous call's contents: the identity order is
    // written here and the build clears the cell counts (a fill kernel, not
    // hipMemsetAsync, so the call can be captured into a graph).
    launch_iota(g.ids(), n, st);
    launch_bins_build(&g, bins, -1, nullptr, nullptr);

    gs_knn_s out;
    out.ctx = ctx;
    out.n = n;
    out.k = k;
    point_into(out.idxb[1], base, kOffIdx, kK * kN);
    point_into(out.distb[1], base, kOffDist, kK * kN);
    out.cur = 0;  // the search writes the alternate buffers idxb[1] / distb[1]
    launch_knn_binned(&g, &out, &g, 1, INFINITY, nullptr, nullptr, nullptr);
    k_noalloc_to_points<<<blocks_for(n), kBlock, 0, st>>>(n, k, g.spos.p, out.idxb[1].p, out.distb[1].p, idx, dist);
    GS_LAUNCHED("k_noalloc_to_points");
}

}  // namespace gs
