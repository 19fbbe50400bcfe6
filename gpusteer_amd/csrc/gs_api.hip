// C-ABI entry points of libgpusteer (declared in include/gpusteer.h).
//
// Each entry point validates its handles and sizes, converts internal errors
// to status codes (never aborting, unlike the reference's CUDA_SAFE_CALL /
// cutilCheckMsg, include/OpenSteer/CUDA/CUDAGlobals.cuh:11) and queues device
// work on the context's stream.
#include <atomic>
#include <mutex>
#include <string>
#include <memory>
#include <type_traits>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "gs_ensemble.hpp"
#include "gs_internal.hpp"

namespace gs {

static thread_local std::string g_last_error;
static void frame_note_failure(const std::string& msg);
// Every failing entry point records its message here (and, while a frame is
// recorded, marks the recording as failed).
void set_last_error(const std::string& m) {
    g_last_error = m;
    frame_note_failure(m);
}
void fail(int status, const std::string& msg) { throw Error(status, msg); }
void hip_check(hipError_t e, const char* what) {
    if (e == hipSuccess) return;
    if (e == hipErrorOutOfMemory) fail(GS_ERR_OUT_OF_MEMORY, std::string(what) + ": out of device memory");
    fail(GS_ERR_DEVICE, std::string(what) + ": " + hipGetErrorString(e));
}

// Frame recording on this thread (gs_frame_begin .. gs_frame_end), or NULL.
struct Recording;
static thread_local Recording* t_rec = nullptr;
// Fails a call that cannot be recorded into a frame while one is recorded.
static void no_frame(const char* what) {
    if (t_rec)
        fail(GS_ERR_FRAME, std::string(what) + " cannot be recorded into a frame (it copies, allocates, frees or "
                                               "synchronizes); issue it before gs_frame_begin or after gs_frame_end");
}

// Allocation-failure injection for tests: after `g_fail_after` more
// successful allocations, every allocation fails (negative = never).
static std::atomic<int64_t> g_fail_after{-1};
void* dev_alloc(size_t bytes) {
    if (t_rec)
        fail(GS_ERR_FRAME, "a device buffer would grow while recording a frame; issue the calls once unrecorded "
                           "first so every buffer has its size");
    int64_t left = g_fail_after.load();
    if (left >= 0) {
        if (left == 0) fail(GS_ERR_OUT_OF_MEMORY, "device allocation failed (injected)");
        g_fail_after.store(left - 1);
    }
    void* p = nullptr;
    if (bytes == 0) bytes = 4;
    hipError_t e = hipMalloc(&p, bytes);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        fail(GS_ERR_OUT_OF_MEMORY, "hipMalloc(" + std::to_string(bytes) + ") failed: " + hipGetErrorString(e));
    }
    return p;
}
void dev_free(void* p) {
    if (p && t_rec) fail(GS_ERR_FRAME, "device memory would be freed while recording a frame");
    if (p) (void)hipFree(p);
}

// Live-handle registry: a destroyed or foreign pointer is rejected instead of
// being dereferenced.
static std::mutex g_reg_mu;
static std::unordered_set<const void*>& registry() {
    static std::unordered_set<const void*> r;
    return r;
}
void reg_add(const void* h) {
    std::lock_guard<std::mutex> l(g_reg_mu);
    registry().insert(h);
}
void reg_del(const void* h) {
    std::lock_guard<std::mutex> l(g_reg_mu);
    registry().erase(h);
}
bool reg_has(const void* h) {
    if (!h) return false;
    std::lock_guard<std::mutex> l(g_reg_mu);
    return registry().count(h) != 0;
}
static std::atomic<uint64_t> g_epoch{1};
static uint64_t next_epoch() { return g_epoch.fetch_add(1); }

// While a frame is recorded, the first check of a group or KNN container
// snapshots its host-side state (see the frame section below).
static void frame_touch(gs_group_s* g);
static void frame_touch(gs_knn_s* k);
static void frame_touch(gs_bins_s* b);
static void frame_touch(gs_path_s* p);
static void frame_touch(const void*) {}

template <class H> static H* check(H* h, const char* what) {
    if (!reg_has(h)) fail(GS_ERR_INVALID_HANDLE, std::string("invalid ") + what + " handle");
    if (t_rec) frame_touch(h);
    return h;
}
static gs_group_s* check_group(gs_group g, const char* what, int kind = -1) {
    check(g, what);
    if (kind >= 0 && g->kind != kind) fail(GS_ERR_INVALID_HANDLE, std::string(what) + ": wrong group kind");
    return g;
}
static void use_device(gs_context ctx) { GS_HIP(hipSetDevice(ctx->device)); }


static GridDev make_grid(const float* o, const float* s, const uint32_t* c) {
    for (int a = 0; a < 3; ++a) {
        if (!(s[a] > 0.f) || !std::isfinite(s[a]) || !std::isfinite(o[a]))
            fail(GS_ERR_INVALID_ARGUMENT, "bin lattice size must be finite and positive");
        if (c[a] == 0) fail(GS_ERR_INVALID_ARGUMENT, "bin lattice needs at least one cell per axis");
    }
    const uint64_t C = (uint64_t)c[0] * c[1] * c[2];
    if (C > GS_MAX_CELLS) fail(GS_ERR_UNSUPPORTED, "bin lattice larger than GS_MAX_CELLS (2^28) cells");
    GridDev g;
    g.ox = o[0];
    g.oy = o[1];
    g.oz = o[2];
    g.sx = s[0];
    g.sy = s[1];
    g.sz = s[2];
    g.dx = (int)c[0];
    g.dy = (int)c[1];
    g.dz = (int)c[2];
    g.hx = g.sx / (float)g.dx;
    g.hy = g.sy / (float)g.dy;
    g.hz = g.sz / (float)g.dz;
    g.rx = (float)g.dx / g.sx;
    g.ry = (float)g.dy / g.sy;
    g.rz = (float)g.dz / g.sz;
    g.slack = 1e-5f * (std::fabs(g.ox) + std::fabs(g.oy) + std::fabs(g.oz) + g.sx + g.sy + g.sz);
    // z cells per x (or, for a single x cell, y) cell width, at least 1.
    const float hxy = g.dx > 1 ? g.hx : g.hy;
    g.zr = g.dz > 1 ? std::max(1, std::min(8, (int)std::ceil(hxy / g.hz * 0.999f))) : 1;
    return g;
}

static bool same_grid(const GridDev& a, const GridDev& b) { return memcmp(&a, &b, sizeof(GridDev)) == 0; }

static void require_knn_shape(const gs_knn_s* knn, const gs_group_s* A) {
    if (knn->ctx != A->ctx) fail(GS_ERR_INVALID_ARGUMENT, "KNN container belongs to another context");
    if (knn->n != A->n)
        fail(GS_ERR_SIZE_MISMATCH, "KNN container holds " + std::to_string(knn->n) + " rows, group has " +
                                       std::to_string(A->n));
}

static void h2d(void* d, const void* h, size_t bytes, hipStream_t st) {
    if (bytes) GS_HIP(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, st));
}
static void d2h(void* h, const void* d, size_t bytes, hipStream_t st) {
    if (bytes) GS_HIP(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, st));
}

// Records that out's current buffer holds the exact result for (A, B) now.
static void mark_result(gs_knn_s* out, const gs_group_s* A, const gs_group_s* B) {
    out->valid = true;
    out->a = A;
    out->b = B;
    out->a_members = A->membership_epoch;
    out->b_members = B->membership_epoch;
    out->a_order = A->order_epoch;
    out->b_order = B->order_epoch;
}

static bool holds_result_for(const gs_knn_s* out, const gs_group_s* A, const gs_group_s* B) {
    return out->valid && out->a == A && out->b == B && out->a_members == A->membership_epoch &&
           out->b_members == B->membership_epoch;
}

// Seed of a KNN call: the container's previous exact result for (A, B), if
// it is at most one reorder behind on either side.  row_map sends a current
// slot of A to its previous slot, val_map a previous slot of B to its
// current one.
struct Seed {
    const uint32_t* seed = nullptr;
    const uint32_t* row_map = nullptr;
    const uint32_t* val_map = nullptr;
};
static Seed seed_for(const gs_knn_s* out, const gs_group_s* A, const gs_group_s* B, uint32_t flags) {
    Seed s;
    if ((flags & GS_KNN_NO_SEED) || !holds_result_for(out, A, B)) return s;
    if (out->a_order != A->order_epoch) {
        if (A->prev_order_epoch == 0 || out->a_order != A->prev_order_epoch) return s;
        s.row_map = A->new_to_old.p;
    }
    if (out->b_order != B->order_epoch) {
        if (B->prev_order_epoch == 0 || out->b_order != B->prev_order_epoch) return s;
        s.val_map = B->old_to_new.p;
    }
    s.seed = out->idx();
    return s;
}

// Brings a container that holds (A, B)'s result to the current storage order
// of both groups (one reorder behind at most), or fails.
static void ensure_current(gs_knn_s* knn, const gs_group_s* A, const gs_group_s* B) {
    if (!holds_result_for(knn, A, B))
        fail(GS_ERR_INVALID_ARGUMENT, "KNN container does not hold a result for these groups (call gs_find_knn)");
    const uint32_t* row = nullptr;
    const uint32_t* val = nullptr;
    if (knn->a_order != A->order_epoch) {
        if (A->prev_order_epoch == 0 || knn->a_order != A->prev_order_epoch)
            fail(GS_ERR_NOT_BINNED, "KNN result is older than the last two bin builds of its group; recompute it");
        row = A->old_to_new.p;
    }
    if (knn->b_order != B->order_epoch) {
        if (B->prev_order_epoch == 0 || knn->b_order != B->prev_order_epoch)
            fail(GS_ERR_NOT_BINNED, "KNN result is older than the last two bin builds of its group; recompute it");
        val = B->old_to_new.p;
    }
    if (row || val) {
        launch_knn_remap(knn, A->n, row, val);
        knn->cur ^= 1;
        knn->a_order = A->order_epoch;
        knn->b_order = B->order_epoch;
    }
}

static void after_reorder(gs_group_s* g) {
    if (g->n == 0) return;
    g->prev_order_epoch = g->order_epoch;
    g->order_epoch = next_epoch();
}

template <class T> static std::vector<T> d2h_vec(const T* d, size_t n, hipStream_t st) {
    std::vector<T> v(n);
    if (n) GS_HIP(hipMemcpyAsync(v.data(), d, n * sizeof(T), hipMemcpyDeviceToHost, st));
    return v;
}

// ---- frames: the host-side state of recorded handles ------------------------
// A recorded call sequence depends on, and changes, the host-side bookkeeping
// of the groups and KNN containers it touches: sizes, current double buffers,
// bin database flags, which result a KNN container holds.  The recording
// snapshots each touched handle at its first check and again at the end;
// gs_frame_end puts the begin state back (nothing has run), and every replay
// requires the begin state and then installs the end state.  Epochs are
// unique names, so they are compared and installed up to renaming: a replay
// requires the current epochs to relate to each other as the recorded ones
// did (the same ones equal, different ones different, none where there was
// none), maps each recorded epoch to its current counterpart, and gives the
// epochs the recorded calls created fresh values.
struct Snap {
    std::vector<uint64_t> exact;   // compared bit for bit
    std::vector<uint64_t> epochs;  // compared up to renaming
};

// Visits the bookkeeping of a handle in a fixed order: reading appends the
// values to a Snap, writing installs them from one.  Device buffers (pointer
// and capacity) are compared, never written.
struct SnapIO {
    Snap* s;
    bool write;
    size_t ix = 0, ie = 0;
    template <class T> void val(T& v) {
        static_assert(sizeof(T) <= 8 && std::is_trivially_copyable<T>::value, "one word per value");
        if (write) {
            const uint64_t w = s->exact[ix++];
            memcpy(&v, &w, sizeof(T));
        } else {
            uint64_t w = 0;
            memcpy(&w, &v, sizeof(T));
            s->exact.push_back(w);
        }
    }
    void ep(uint64_t& v) {
        if (write)
            v = s->epochs[ie++];
        else
            s->epochs.push_back(v);
    }
    template <class T> void buf(gs::DevBuf<T>& b) {
        if (write) {
            ix += 2;
        } else {
            s->exact.push_back((uint64_t)(uintptr_t)b.p);
            s->exact.push_back((uint64_t)b.cap);
        }
    }
    void grid(GridDev& g) {
        static_assert(sizeof(GridDev) % 4 == 0, "GridDev is 4-byte fields");
        uint32_t w[sizeof(GridDev) / 4];
        memcpy(w, &g, sizeof w);
        for (auto& x : w) val(x);
        if (write) memcpy(&g, w, sizeof w);
    }
};

static void visit(gs_group_s* g, SnapIO& io) {
    io.val(g->kind);
    io.val(g->n);
    io.val(g->live);
    io.val(g->inv);
    for (auto& v : g->inv_val) io.val(v);
    io.val(g->bits_zero);
    io.ep(g->membership_epoch);
    io.ep(g->order_epoch);
    io.ep(g->prev_order_epoch);
    for (auto& c : g->fcur) io.val(c);
    io.val(g->bcur);
    io.val(g->icur);
    io.val(g->binned);
    io.val(g->snap_valid);
    io.grid(g->bin_grid);
    io.val(g->bin_cells);
    for (auto& f : g->fb)
        for (auto& b : f) io.buf(b);
    for (auto& b : g->bitsb) io.buf(b);
    for (auto& b : g->idb) io.buf(b);
    gs::DevBuf<uint32_t>* u[] = {&g->slot_of_id, &g->new_to_old, &g->old_to_new, &g->key,
                                 &g->rank,       &g->id_tmp,     &g->start,      &g->count, &g->flag};
    for (auto* b : u) io.buf(*b);
    io.buf(g->spos);
    io.buf(g->sfwd);
    io.val(g->bbox_cur);
    io.val(g->bbox_wanted);
    io.val(g->bbox_fresh);
    for (auto& b : g->bbox) io.buf(b);
}

static void visit(gs_knn_s* k, SnapIO& io) {
    io.val(k->n);
    io.val(k->k);
    io.val(k->cur);
    io.val(k->valid);
    io.val(k->a);
    io.val(k->b);
    io.ep(k->a_members);
    io.ep(k->b_members);
    io.ep(k->a_order);
    io.ep(k->b_order);
    for (auto& b : k->idxb) io.buf(b);
    for (auto& b : k->distb) io.buf(b);
    io.buf(k->part_key);
    io.buf(k->part_slot);
}

// A handle the recorded calls used: a group or KNN container with its
// bookkeeping, or a bin lattice or path, which only has to stay alive.
struct Touched {
    gs_group_s* g = nullptr;
    gs_knn_s* k = nullptr;
    const void* keep = nullptr;
    Snap begin, end;
    const void* handle() const { return g ? (const void*)g : k ? (const void*)k : keep; }
};

template <class F> static void visit_touched(Touched& t, F&& f) {
    if (t.g)
        f(t.g);
    else if (t.k)
        f(t.k);
}

static Snap snap_of(Touched& t) {
    Snap s;
    SnapIO io{&s, false};
    visit_touched(t, [&](auto* h) { visit(h, io); });
    return s;
}

static void install(Touched& t, Snap& s) {
    SnapIO io{&s, true};
    visit_touched(t, [&](auto* h) { visit(h, io); });
}

struct Recording {
    gs_context_s* ctx = nullptr;
    std::vector<Touched> touched;
    bool failed = false;
    std::string first_error;
};

static void frame_note_failure(const std::string& msg) {
    if (t_rec && !t_rec->failed) {
        t_rec->failed = true;
        t_rec->first_error = msg;
    }
}

template <class H> static void frame_touch_h(H* h, H* Touched::*member) {
    if (h->ctx != t_rec->ctx) fail(GS_ERR_FRAME, "a handle of another context was used while recording a frame");
    for (auto& t : t_rec->touched)
        if (t.*member == h) return;
    Touched t;
    t.*member = h;
    t.begin = snap_of(t);
    t_rec->touched.push_back(std::move(t));
}
static void frame_touch(gs_group_s* g) { frame_touch_h(g, &Touched::g); }
static void frame_touch(gs_knn_s* k) { frame_touch_h(k, &Touched::k); }
template <class H> static void frame_keep(H* h) {
    if (h->ctx != t_rec->ctx) fail(GS_ERR_FRAME, "a handle of another context was used while recording a frame");
    for (auto& t : t_rec->touched)
        if (t.keep == h) return;
    Touched t;
    t.keep = h;
    t_rec->touched.push_back(std::move(t));
}
static void frame_touch(gs_bins_s* b) { frame_keep(b); }
static void frame_touch(gs_path_s* p) { frame_keep(p); }

}  // namespace gs

struct gs_frame_s {
    gs_context ctx = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<gs::Touched> touched;
    uint32_t magic = 0x67736672u;
};

using namespace gs;

gs_agents gs_group_s::dev_agents() const {
    gs_agents a;
    float** fp = (float**)&a;
    for (int j = 0; j < kFields; ++j) fp[j] = F(j);
    a.bits = bits();
    return a;
}

extern "C" {

const char* gs_status_string(int status) {
    switch (status) {
        case GS_OK: return "ok";
        case GS_ERR_INVALID_ARGUMENT: return "invalid argument";
        case GS_ERR_INVALID_HANDLE: return "invalid handle";
        case GS_ERR_SIZE_MISMATCH: return "size mismatch";
        case GS_ERR_UNSUPPORTED: return "unsupported";
        case GS_ERR_NOT_BINNED: return "group has no current bin database";
        case GS_ERR_OUT_OF_MEMORY: return "out of device memory";
        case GS_ERR_DEVICE: return "device error";
        case GS_ERR_NONFINITE: return "non-finite state";
        case GS_ERR_FRAME: return "not recordable, or frame replayed in another state";
        default: return "unknown status";
    }
}

const char* gs_last_error(void) { return g_last_error.c_str(); }
int gs_version(void) { return GS_VERSION; }

/* Test hook (not part of the public header's stable surface): make device
 * allocations fail after n more successes; n < 0 disables. */
int gs_debug_fail_alloc_after(int64_t n) {
    g_fail_after.store(n);
    return GS_OK;
}

// ---- context -----------------------------------------------------------------
int gs_context_create(int device, void* hip_stream, gs_context* out) {
    return guarded([&] {
        no_frame("gs_context_create");
        if (!out) fail(GS_ERR_INVALID_ARGUMENT, "out is NULL");
        *out = nullptr;
        int count = 0;
        hipError_t e = hipGetDeviceCount(&count);
        if (e != hipSuccess || count == 0) {
            (void)hipGetLastError();
            fail(GS_ERR_DEVICE, "no HIP device available");
        }
        if (device < 0 || device >= count) fail(GS_ERR_INVALID_ARGUMENT, "device index out of range");
        GS_HIP(hipSetDevice(device));
        // Per-device set-up of the allocation-free entry point's array.
        noalloc_init_device(device);
        auto* c = new gs_context_s;
        c->device = device;
        if (hip_stream) {
            c->stream = (hipStream_t)hip_stream;
        } else {
            e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
            if (e != hipSuccess) {
                delete c;
                hip_check(e, "hipStreamCreateWithFlags");
            }
            c->own_stream = true;
        }
        reg_add(c);
        *out = c;
    });
}

int gs_context_destroy(gs_context ctx) {
    return guarded([&] {
        no_frame("gs_context_destroy");
        check(ctx, "context");
        use_device(ctx);
        (void)hipStreamSynchronize(ctx->stream);
        if (ctx->own_stream) (void)hipStreamDestroy(ctx->stream);
        reg_del(ctx);
        delete ctx;
    });
}

int gs_context_synchronize(gs_context ctx) {
    return guarded([&] {
        no_frame("gs_context_synchronize");
        check(ctx, "context");
        use_device(ctx);
        GS_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int gs_context_stream(gs_context ctx, void** s) {
    return guarded([&] {
        check(ctx, "context");
        if (!s) fail(GS_ERR_INVALID_ARGUMENT, "out is NULL");
        *s = (void*)ctx->stream;
    });
}

int gs_context_set_shape(gs_context ctx, int kernel, uint32_t block_threads, uint32_t agents_per_lane) {
    return guarded([&] {
        check(ctx, "context");
        if (kernel != GS_SHAPE_KNN && kernel != GS_SHAPE_FLOCK) fail(GS_ERR_INVALID_ARGUMENT, "unknown kernel");
        if ((block_threads != 128 && block_threads != 256 && block_threads != 512) ||
            (agents_per_lane != 1 && agents_per_lane != 2))
            fail(GS_ERR_UNSUPPORTED, "launch shape must be 128, 256 or 512 threads and 1 or 2 agents per lane");
        ctx->shape[kernel] = LaunchShape{block_threads, agents_per_lane};
    });
}

int gs_context_get_shape(gs_context ctx, int kernel, uint32_t* block_threads, uint32_t* agents_per_lane) {
    return guarded([&] {
        check(ctx, "context");
        if (kernel != GS_SHAPE_KNN && kernel != GS_SHAPE_FLOCK) fail(GS_ERR_INVALID_ARGUMENT, "unknown kernel");
        if (block_threads) *block_threads = ctx->shape[kernel].block;
        if (agents_per_lane) *agents_per_lane = ctx->shape[kernel].apl;
    });
}

// ---- bins --------------------------------------------------------------------
int gs_bins_create(gs_context ctx, const float origin[3], const float size[3], const uint32_t cells[3], gs_bins* out) {
    return guarded([&] {
        no_frame("gs_bins_create");
        check(ctx, "context");
        if (!origin || !size || !cells || !out) fail(GS_ERR_INVALID_ARGUMENT, "NULL argument");
        auto* b = new gs_bins_s;
        b->ctx = ctx;
        try {
            b->g = make_grid(origin, size, cells);
        } catch (...) {
            delete b;
            throw;
        }
        b->cells = cells[0] * cells[1] * cells[2];
        reg_add(b);
        *out = b;
    });
}

int gs_bins_create_centered(gs_context ctx, const uint32_t world_cells[3], const float world_size[3], gs_bins* out) {
    if (!world_size) return GS_ERR_INVALID_ARGUMENT;
    const float origin[3] = {-0.5f * world_size[0], -0.5f * world_size[1], -0.5f * world_size[2]};
    return gs_bins_create(ctx, origin, world_size, world_cells, out);
}

int gs_bins_destroy(gs_bins b) {
    return guarded([&] {
        no_frame("gs_bins_destroy");
        check(b, "bins");
        reg_del(b);
        delete b;
    });
}

int gs_bins_geometry(gs_bins b, float origin[3], float size[3], uint32_t cells[3]) {
    return guarded([&] {
        check(b, "bins");
        if (origin) {
            origin[0] = b->g.ox;
            origin[1] = b->g.oy;
            origin[2] = b->g.oz;
        }
        if (size) {
            size[0] = b->g.sx;
            size[1] = b->g.sy;
            size[2] = b->g.sz;
        }
        if (cells) {
            cells[0] = (uint32_t)b->g.dx;
            cells[1] = (uint32_t)b->g.dy;
            cells[2] = (uint32_t)b->g.dz;
        }
    });
}

// ---- groups ------------------------------------------------------------------
int gs_group_create(gs_context ctx, int kind, gs_group* out) {
    return guarded([&] {
        no_frame("gs_group_create");
        check(ctx, "context");
        if (!out) fail(GS_ERR_INVALID_ARGUMENT, "out is NULL");
        if (kind != GS_KIND_AGENTS && kind != GS_KIND_OBSTACLES && kind != GS_KIND_WALLS)
            fail(GS_ERR_INVALID_ARGUMENT, "unknown group kind");
        auto* g = new gs_group_s;
        g->ctx = ctx;
        g->kind = kind;
        g->live = kind == GS_KIND_AGENTS ? kFieldsAgents : kind == GS_KIND_OBSTACLES ? kFieldsObstacles : kFieldsWalls;
        g->membership_epoch = next_epoch();
        g->order_epoch = next_epoch();
        // The bounding-box pair of capped searches (12 words, allocated here
        // so that no later call allocates for it), both empty.
        use_device(ctx);
        for (auto& b : g->bbox) {
            b.reserve(6);
            launch_fill_u32(b.p, 3, kBoxEmptyMin, ctx->stream);
            launch_fill_u32(b.p + 3, 3, kBoxEmptyMax, ctx->stream);
        }
        reg_add(g);
        *out = g;
    });
}

static void release_group(gs_group_s* g) {
    for (auto& f : g->fb)
        for (auto& b : f) b.release();
    for (auto& b : g->bitsb) b.release();
    for (auto& b : g->idb) b.release();
    gs::DevBuf<uint32_t>* u[] = {&g->slot_of_id, &g->new_to_old, &g->old_to_new, &g->key, &g->rank,
                                 &g->id_tmp, &g->start,      &g->count,      &g->flag};
    for (auto* b : u) b->release();
    g->spos.release();
    g->sfwd.release();
    for (auto& b : g->bbox) b.release();
}

int gs_group_destroy(gs_group g) {
    return guarded([&] {
        no_frame("gs_group_destroy");
        check_group(g, "group");
        use_device(g->ctx);
        (void)hipStreamSynchronize(g->ctx->stream);
        release_group(g);
        reg_del(g);
        delete g;
    });
}

int gs_group_size(gs_group g, uint32_t* n) {
    return guarded([&] {
        check_group(g, "group");
        if (!n) fail(GS_ERR_INVALID_ARGUMENT, "out is NULL");
        *n = g->n;
    });
}

// New members in index order: allocate both buffers of every live field,
// identity id <-> slot maps, fresh epochs.
static void begin_members(gs_group_s* g, uint32_t n) {
    for (int j = 0; j < kFields; ++j)
        if (g->live & (1u << j)) {
            g->fb[j][0].reserve(n);
            g->fb[j][1].reserve(n);
        }
    if (g->kind == GS_KIND_AGENTS) {
        g->bitsb[0].reserve(n);
        g->bitsb[1].reserve(n);
    }
    g->idb[0].reserve(n);
    g->idb[1].reserve(n);
    g->slot_of_id.reserve(n);
    launch_iota(g->ids(), n, g->ctx->stream);
    g->n = n;
    g->inv = 0;
    g->bits_zero = false;
    g->binned = false;
    g->snap_valid = false;
    g->membership_epoch = next_epoch();
    g->order_epoch = next_epoch();
    g->prev_order_epoch = 0;
}

int gs_group_upload_agents(gs_group g, uint32_t n, const gs_agents* host) {
    return guarded([&] {
        no_frame("gs_group_upload_agents");
        check_group(g, "group", GS_KIND_AGENTS);
        if (n > 0 && !host) fail(GS_ERR_INVALID_ARGUMENT, "host arrays are NULL");
        const float* const* hf = host ? (const float* const*)host : nullptr;
        if (n > 0) {
            for (int j = 0; j < kFields; ++j)
                if (!hf[j]) fail(GS_ERR_INVALID_ARGUMENT, "agent array " + std::to_string(j) + " is NULL");
            if (!host->bits) fail(GS_ERR_INVALID_ARGUMENT, "bits array is NULL");
        }
        use_device(g->ctx);
        hipStream_t st = g->ctx->stream;
        begin_members(g, n);
        for (int j = 0; j < kFields && n; ++j) h2d(g->F(j), hf[j], (size_t)n * 4, st);
        if (n) h2d(g->bits(), host->bits, (size_t)n * 4, st);
        // Fields with one bit pattern in every slot (positions excluded).
        if (n) {
            for (int j = 3; j < kFields; ++j) {
                const uint32_t* w = (const uint32_t*)hf[j];
                bool same = true;
                for (uint32_t i = 1; i < n && same; ++i) same = w[i] == w[0];
                if (same) {
                    g->inv |= 1u << j;
                    g->inv_val[j] = hf[j][0];
                }
            }
            bool zero = true;
            for (uint32_t i = 0; i < n && zero; ++i) zero = host->bits[i] == 0u;
            g->bits_zero = zero;
        }
        GS_HIP(hipStreamSynchronize(st));
    });
}

int gs_group_download_agents(gs_group g, gs_agents* host) {
    return guarded([&] {
        no_frame("gs_group_download_agents");
        check_group(g, "group", GS_KIND_AGENTS);
        if (!host) fail(GS_ERR_INVALID_ARGUMENT, "host arrays are NULL");
        use_device(g->ctx);
        hipStream_t st = g->ctx->stream;
        const size_t n = g->n;
        if (!n) return;
        // Storage order -> agent index order on the host.
        auto ids = d2h_vec(g->ids(), n, st);
        std::vector<float> tmp(n);
        GS_HIP(hipStreamSynchronize(st));
        float* const* hf = (float* const*)host;
        for (int j = 0; j < kFields; ++j) {
            if (!hf[j]) continue;
            GS_HIP(hipMemcpyAsync(tmp.data(), g->F(j), n * 4, hipMemcpyDeviceToHost, st));
            GS_HIP(hipStreamSynchronize(st));
            for (size_t s = 0; s < n; ++s) hf[j][ids[s]] = tmp[s];
        }
        if (host->bits) {
            auto b = d2h_vec(g->bits(), n, st);
            GS_HIP(hipStreamSynchronize(st));
            for (size_t s = 0; s < n; ++s) host->bits[ids[s]] = b[s];
        }
    });
}

int gs_group_device_agents(gs_group g, gs_agents* dev) {
    return guarded([&] {
        check_group(g, "group");
        if (!dev) fail(GS_ERR_INVALID_ARGUMENT, "out is NULL");
        *dev = g->dev_agents();
        // The caller may write through these pointers.
        g->inv = 0;
        g->bits_zero = false;
    });
}

int gs_group_device_order(gs_group g, uint32_t** id_of_slot, uint32_t** slot_of_id) {
    return guarded([&] {
        check_group(g, "group");
        if (id_of_slot) *id_of_slot = g->ids();
        if (slot_of_id) {
            // Not maintained by the hot path: inverted on request.
            use_device(g->ctx);
            launch_invert_ids(g);
            *slot_of_id = g->slot_of_id.p;
        }
    });
}

int gs_group_upload_obstacles(gs_group g, uint32_t n, const float* cx, const float* cy, const float* cz,
                              const float* radius) {
    return guarded([&] {
        no_frame("gs_group_upload_obstacles");
        check_group(g, "group", GS_KIND_OBSTACLES);
        if (n > 0 && (!cx || !cy || !cz || !radius)) fail(GS_ERR_INVALID_ARGUMENT, "obstacle arrays are NULL");
        use_device(g->ctx);
        hipStream_t st = g->ctx->stream;
        begin_members(g, n);
        const float* src[4] = {cx, cy, cz, radius};
        const int dst[4] = {0, 1, 2, 19};
        for (int j = 0; j < 4; ++j) h2d(g->F(dst[j]), src[j], (size_t)n * 4, st);
        GS_HIP(hipStreamSynchronize(st));
    });
}

// Walls with their KNN positions (midpoints): given, or 0.5f*(start+end)
// (CUDA/WallGroupData.cu:87-89, WallGroup.h:31).
static void upload_walls(gs_group_s* g, uint32_t n, const float* const* hw, const float* const* given_mid) {
    std::vector<float> mid[3];
    const float* m[3];
    for (int a = 0; a < 3; ++a) {
        if (given_mid) {
            m[a] = given_mid[a];
        } else {
            mid[a].resize(n);
            for (uint32_t i = 0; i < n; ++i) mid[a][i] = 0.5f * (hw[a][i] + hw[3 + a][i]);
            m[a] = mid[a].data();
        }
    }
    use_device(g->ctx);
    hipStream_t st = g->ctx->stream;
    begin_members(g, n);
    for (int j = 0; j < 9 && n; ++j) h2d(g->F(kWallField + j), hw[j], (size_t)n * 4, st);
    for (int a = 0; a < 3 && n; ++a) h2d(g->F(a), m[a], (size_t)n * 4, st);
    GS_HIP(hipStreamSynchronize(st));
}

int gs_group_upload_walls(gs_group g, uint32_t n, const gs_walls* host) {
    return guarded([&] {
        no_frame("gs_group_upload_walls");
        check_group(g, "group", GS_KIND_WALLS);
        if (n > 0 && !host) fail(GS_ERR_INVALID_ARGUMENT, "wall arrays are NULL");
        const float* const* hw = host ? (const float* const*)host : nullptr;
        for (int j = 0; j < 9 && n; ++j)
            if (!hw[j]) fail(GS_ERR_INVALID_ARGUMENT, "wall array is NULL");
        upload_walls(g, n, hw, nullptr);
    });
}

int gs_group_upload_walls_mid(gs_group g, uint32_t n, const gs_walls* host, const float* mid_x, const float* mid_y,
                              const float* mid_z) {
    return guarded([&] {
        no_frame("gs_group_upload_walls_mid");
        check_group(g, "group", GS_KIND_WALLS);
        if (n > 0 && (!host || !mid_x || !mid_y || !mid_z)) fail(GS_ERR_INVALID_ARGUMENT, "wall arrays are NULL");
        const float* const* hw = host ? (const float* const*)host : nullptr;
        for (int j = 0; j < 9 && n; ++j)
            if (!hw[j]) fail(GS_ERR_INVALID_ARGUMENT, "wall array is NULL");
        const float* mid[3] = {mid_x, mid_y, mid_z};
        upload_walls(g, n, hw, mid);
    });
}

int gs_group_upload_wall_set(gs_group g, gs_wall_set w) {
    return guarded([&] {
        no_frame("gs_group_upload_wall_set");
        check_group(g, "group", GS_KIND_WALLS);
        uint32_t n = 0;
        if (gs_wall_set_size(w, &n) != GS_OK) fail(GS_ERR_INVALID_HANDLE, "wall set handle is invalid");
        std::vector<float> v[12];
        for (auto& a : v) a.resize(n);
        gs_walls out{v[0].data(), v[1].data(), v[2].data(), v[3].data(), v[4].data(),
                     v[5].data(), v[6].data(), v[7].data(), v[8].data()};
        if (gs_wall_set_get(w, &out, v[9].data(), v[10].data(), v[11].data()) != GS_OK)
            fail(GS_ERR_INVALID_HANDLE, "wall set handle is invalid");
        const float* hw[9];
        for (int j = 0; j < 9; ++j) hw[j] = v[j].data();
        const float* mid[3] = {v[9].data(), v[10].data(), v[11].data()};
        upload_walls(g, n, hw, mid);
    });
}

// ComputeCellDensity / ComputeAvgCellVelocity (CUDA/ProcessData.cu:52-129)
// into host memory through a temporary device buffer.
static void extract_to_host(gs_group_s* g, size_t count, void* host, bool velocity) {
    if (!g->binned || !g->snap_valid) fail(GS_ERR_NOT_BINNED, "group has no bin database (gs_update_knn_database)");
    use_device(g->ctx);
    hipStream_t st = g->ctx->stream;
    if (count == 0) return;
    DevBuf<uint32_t> dev;  // uint32 or float elements, 4 bytes each
    dev.reserve(count);
    try {
        if (velocity) launch_cell_velocity(g, (float*)dev.p);
        else launch_cell_density(g, dev.p);
        GS_HIP(hipMemcpyAsync(host, dev.p, count * 4, hipMemcpyDeviceToHost, st));
        GS_HIP(hipStreamSynchronize(st));
    } catch (...) {
        (void)hipStreamSynchronize(st);
        dev.release();
        throw;
    }
    dev.release();
}

int gs_group_cell_density(gs_group g, uint32_t* agents_per_cell) {
    return guarded([&] {
        no_frame("gs_group_cell_density");
        check_group(g, "group");
        if (!agents_per_cell) fail(GS_ERR_INVALID_ARGUMENT, "output is NULL");
        extract_to_host(g, g->bin_cells, agents_per_cell, false);
    });
}

int gs_group_cell_velocity(gs_group g, float* avg_velocity) {
    return guarded([&] {
        no_frame("gs_group_cell_velocity");
        check_group(g, "group", GS_KIND_AGENTS);
        if (!avg_velocity) fail(GS_ERR_INVALID_ARGUMENT, "output is NULL");
        extract_to_host(g, (size_t)g->bin_cells * 3, avg_velocity, true);
    });
}

int gs_write_cell_density(gs_group g, const char* prefix, const uint32_t* selected_cells, uint32_t n_selected,
                          const char* selected_prefix, int all_rows) {
    return guarded([&] {
        no_frame("gs_write_cell_density");
        check_group(g, "group");
        if (!prefix) fail(GS_ERR_INVALID_ARGUMENT, "prefix is NULL");
        if (n_selected && (!selected_cells || !selected_prefix))
            fail(GS_ERR_INVALID_ARGUMENT, "selected cells without a prefix");
        std::vector<uint32_t> d(g->bin_cells);
        extract_to_host(g, d.size(), d.data(), false);
        write_cell_density(g->bin_grid, d, prefix, selected_cells, n_selected, selected_prefix, all_rows);
    });
}

int gs_write_cell_velocity(gs_group g, const char* prefix, const uint32_t* selected_cells, uint32_t n_selected,
                           const char* selected_prefix, int all_rows) {
    return guarded([&] {
        no_frame("gs_write_cell_velocity");
        check_group(g, "group", GS_KIND_AGENTS);
        if (!prefix) fail(GS_ERR_INVALID_ARGUMENT, "prefix is NULL");
        if (n_selected && (!selected_cells || !selected_prefix))
            fail(GS_ERR_INVALID_ARGUMENT, "selected cells without a prefix");
        std::vector<float> v((size_t)g->bin_cells * 3);
        extract_to_host(g, v.size(), v.data(), true);
        write_cell_velocity(g->bin_grid, v, prefix, selected_cells, n_selected, selected_prefix, all_rows);
    });
}

int gs_group_download_bins(gs_group g, uint32_t* key, uint32_t* start, uint32_t* perm, float* spx, float* spy,
                           float* spz) {
    return guarded([&] {
        no_frame("gs_group_download_bins");
        check_group(g, "group");
        if (!g->binned || !g->snap_valid) fail(GS_ERR_NOT_BINNED, "group has no bin database");
        use_device(g->ctx);
        hipStream_t st = g->ctx->stream;
        const size_t n = g->n;
        const size_t C = g->bin_cells;
        auto st_h = d2h_vec(g->start.p, C + 1, st);
        auto ids = d2h_vec(g->ids(), n, st);
        auto sp = d2h_vec(g->spos.p, n, st);
        GS_HIP(hipStreamSynchronize(st));
        for (size_t t = 0; t < n; ++t) {
            if (spx) spx[t] = sp[t].x;
            if (spy) spy[t] = sp[t].y;
            if (spz) spz[t] = sp[t].z;
        }
        if (start) memcpy(start, st_h.data(), (C + 1) * 4);
        if (perm) memcpy(perm, ids.data(), n * 4);  // agent index of each sorted slot
        if (key)
            for (size_t c = 0; c < C; ++c)
                for (uint32_t s = st_h[c]; s < st_h[c + 1] && s < n; ++s) key[ids[s]] = (uint32_t)c;
    });
}

int gs_group_bin_geometry(gs_group g, float origin[3], float size[3], uint32_t cells[3]) {
    return guarded([&] {
        check_group(g, "group");
        if (!g->binned) fail(GS_ERR_NOT_BINNED, "group has no bin database");
        const GridDev& b = g->bin_grid;
        if (origin) origin[0] = b.ox, origin[1] = b.oy, origin[2] = b.oz;
        if (size) size[0] = b.sx, size[1] = b.sy, size[2] = b.sz;
        if (cells) cells[0] = (uint32_t)b.dx, cells[1] = (uint32_t)b.dy, cells[2] = (uint32_t)b.dz;
    });
}

int gs_group_check_finite(gs_group g, uint32_t* first_bad) {
    uint32_t bad = UINT32_MAX;
    int s = guarded([&] {
        no_frame("gs_group_check_finite");
        check_group(g, "group", GS_KIND_AGENTS);
        use_device(g->ctx);
        uint32_t slot = UINT32_MAX;
        launch_check_finite(g, &slot);
        if (slot != UINT32_MAX) {
            GS_HIP(hipMemcpyAsync(&bad, g->ids() + slot, 4, hipMemcpyDeviceToHost, g->ctx->stream));
            GS_HIP(hipStreamSynchronize(g->ctx->stream));
        }
    });
    if (first_bad) *first_bad = bad;
    if (s != GS_OK) return s;
    if (bad != UINT32_MAX) {
        set_last_error("non-finite state at agent " + std::to_string(bad));
        return GS_ERR_NONFINITE;
    }
    return GS_OK;
}

// ---- KNN container --------------------------------------------------------------
int gs_knn_create(gs_context ctx, uint32_t n, uint32_t k, gs_knn* out) {
    return guarded([&] {
        no_frame("gs_knn_create");
        check(ctx, "context");
        if (!out) fail(GS_ERR_INVALID_ARGUMENT, "out is NULL");
        if (k == 0) fail(GS_ERR_INVALID_ARGUMENT, "k must be at least 1");
        if (k > GS_MAX_K) fail(GS_ERR_UNSUPPORTED, "k larger than GS_MAX_K");
        use_device(ctx);
        auto* c = new gs_knn_s;
        c->ctx = ctx;
        c->n = n;
        c->k = k;
        try {
            for (int b = 0; b < 2; ++b) {
                c->idxb[b].reserve((size_t)n * k);
                c->distb[b].reserve((size_t)n * k);
            }
        } catch (...) {
            for (int b = 0; b < 2; ++b) {
                c->idxb[b].release();
                c->distb[b].release();
            }
            delete c;
            throw;
        }
        reg_add(c);
        *out = c;
    });
}

int gs_knn_destroy(gs_knn knn) {
    return guarded([&] {
        no_frame("gs_knn_destroy");
        check(knn, "knn");
        use_device(knn->ctx);
        (void)hipStreamSynchronize(knn->ctx->stream);
        for (int b = 0; b < 2; ++b) {
            knn->idxb[b].release();
            knn->distb[b].release();
        }
        knn->part_key.release();
        knn->part_slot.release();
        reg_del(knn);
        delete knn;
    });
}

int gs_knn_shape(gs_knn knn, uint32_t* n, uint32_t* k) {
    return guarded([&] {
        check(knn, "knn");
        if (n) *n = knn->n;
        if (k) *k = knn->k;
    });
}

int gs_knn_download(gs_knn knn, uint32_t* idx, float* dist) {
    return guarded([&] {
        no_frame("gs_knn_download");
        check(knn, "knn");
        if (!knn->valid) fail(GS_ERR_INVALID_ARGUMENT, "KNN container holds no result");
        if (!reg_has(knn->a) || !reg_has(knn->b)) fail(GS_ERR_INVALID_HANDLE, "the KNN result's groups are gone");
        gs_group_s* A = (gs_group_s*)knn->a;
        gs_group_s* B = (gs_group_s*)knn->b;
        use_device(knn->ctx);
        ensure_current(knn, A, B);
        hipStream_t st = knn->ctx->stream;
        const size_t n = knn->n, k = knn->k;
        auto si = d2h_vec(knn->idx(), n * k, st);
        auto sd = d2h_vec(knn->dist(), n * k, st);
        auto ida = d2h_vec(A->ids(), A->n, st);
        auto idb = d2h_vec(B->ids(), B->n, st);
        GS_HIP(hipStreamSynchronize(st));
        // Slot space -> agent index space: row = A's agent index, value = B's.
        for (size_t j = 0; j < k; ++j)
            for (size_t s = 0; s < n; ++s) {
                const uint32_t v = si[j * n + s];
                const size_t o = j * n + ida[s];
                if (idx) idx[o] = v == UINT32_MAX ? UINT32_MAX : idb[v];
                if (dist) dist[o] = sd[j * n + s];
            }
    });
}

int gs_knn_device(gs_knn knn, uint32_t** idx, float** dist) {
    return guarded([&] {
        check(knn, "knn");
        // The caller may now write the buffers: their contents can no longer
        // be trusted as an exact previous result.
        knn->valid = false;
        if (idx) *idx = knn->idx();
        if (dist) *dist = knn->dist();
    });
}

// ---- operator API ------------------------------------------------------------------
int gs_update_knn_database(gs_group g, gs_bins bins) {
    return guarded([&] {
        check_group(g, "group");
        check(bins, "bins");
        if (bins->ctx != g->ctx) fail(GS_ERR_INVALID_ARGUMENT, "bins belong to another context");
        use_device(g->ctx);
        launch_bins_build(g, bins, -1, nullptr, nullptr);
        after_reorder(g);
    });
}

// Strict radius cap r (lq.c:327: distanceSquared < radiusSquared, both fp32)
// as an inclusive bound on d2: the largest float below r*r.
static float cap2_of(float max_distance) {
    if (std::isnan(max_distance) || max_distance <= 0.f)
        fail(GS_ERR_INVALID_ARGUMENT, "max_distance must be positive (INFINITY for no cap)");
    if (std::isinf(max_distance)) return INFINITY;
    const float r2 = max_distance * max_distance;
    return std::nextafter(r2, 0.f);
}

// Cells of the box a capped search can reach (the whole lattice without a
// cap).
static uint64_t reach_cells(const GridDev& g, uint32_t cells, float max_distance) {
    if (std::isinf(max_distance)) return cells;
    const float h[3] = {g.hx, g.hy, g.hz};
    const int d[3] = {g.dx, g.dy, g.dz};
    uint64_t box = 1;
    for (int a = 0; a < 3; ++a) {
        const double r = std::ceil((double)max_distance / (double)h[a]);
        box *= (uint64_t)std::min<double>((double)d[a], 2.0 * r + 1.0);
    }
    return box;
}

// Whether the exhaustive scan of B's snapshot is cheaper than the binned
// search (the result is the same exact one either way).
static bool prefer_exhaustive(const gs_group_s* B, const GridDev& g, uint32_t cells, float max_distance) {
    if (B->n <= kSmallTargetSet) return true;
    if (B->n > kExhaustiveMax || (uint64_t)B->n * kSparseCells > cells) return false;
    return reach_cells(g, cells, max_distance) * 4 >= B->n;
}

// A small query group searching another small group over a wide reach: the
// split exhaustive scan fills the chip where a one-lane-per-query search
// would run as a few dozen wavefronts, each walking the cells around a query
// that the other group may be far from.  A group's search of itself (its
// points are where its queries are) and a narrowly capped search stay in the
// bins.
static bool prefer_split(const gs_group_s* A, const gs_group_s* B, const GridDev& g, uint32_t cells,
                         float max_distance) {
    return A != B && A->n <= kSplitMaxA && B->n <= kSplitMaxB && (uint64_t)A->n * B->n <= kSplitMaxPairs &&
           reach_cells(g, cells, max_distance) > kSplitMinCells && knn_split_ranges(A->n, B->n) >= 2u;
}

static void find_knn(gs_group A, gs_knn out, gs_bins bins, gs_group B, uint32_t search_radius, uint32_t flags,
                     float max_distance) {
    check_group(A, "group A");
    check_group(B, "group B");
    check(out, "knn");
    check(bins, "bins");
    if (A->ctx != B->ctx || bins->ctx != A->ctx) fail(GS_ERR_INVALID_ARGUMENT, "handles of different contexts");
    require_knn_shape(out, A);
    if (!A->binned || !B->binned || !A->snap_valid || !B->snap_valid)
        fail(GS_ERR_NOT_BINNED, "call gs_update_knn_database on both groups first");
    if (!same_grid(B->bin_grid, bins->g) || !same_grid(A->bin_grid, bins->g))
        fail(GS_ERR_NOT_BINNED, "groups were binned with another lattice");
    const float cap2 = cap2_of(max_distance);
    use_device(A->ctx);
    // B's builds from now on keep its bounding box, which lets the binned
    // search skip queries the cap cannot reach it from.
    if (std::isfinite(cap2)) B->bbox_wanted = true;
    if (!(flags & GS_KNN_BINNED) && prefer_split(A, B, bins->g, bins->cells, max_distance)) {
        const Seed sd = seed_for(out, A, B, flags);
        launch_knn_split(A, out, B, cap2, sd.seed, sd.row_map, sd.val_map);
    } else if (!(flags & GS_KNN_BINNED) && prefer_exhaustive(B, bins->g, bins->cells, max_distance)) {
        launch_knn_snap_brute(A, out, B, cap2);
    } else {
        const Seed sd = seed_for(out, A, B, flags);
        launch_knn_binned(A, out, B, search_radius, cap2, sd.seed, sd.row_map, sd.val_map);
    }
    out->cur ^= 1;
    mark_result(out, A, B);
}

int gs_find_knn_within(gs_group A, gs_knn out, gs_bins bins, gs_group B, uint32_t search_radius, uint32_t flags,
                       float max_distance) {
    return guarded([&] { find_knn(A, out, bins, B, search_radius, flags, max_distance); });
}

int gs_find_knn(gs_group A, gs_knn out, gs_bins bins, gs_group B, uint32_t search_radius, uint32_t flags) {
    return guarded([&] { find_knn(A, out, bins, B, search_radius, flags, INFINITY); });
}

int gs_find_knn_noalloc(gs_context ctx, gs_bins bins, uint32_t n, const float* px, const float* py, const float* pz,
                        uint32_t k, uint32_t* idx, float* dist) {
    return guarded([&] {
        check(ctx, "context");
        check(bins, "bins");
        if (bins->ctx != ctx) fail(GS_ERR_INVALID_ARGUMENT, "bins belong to another context");
        use_device(ctx);
        noalloc_find_knn(ctx, bins, n, px, py, pz, k, idx, dist);
    });
}

int gs_find_knn_brute(gs_group A, gs_knn out, gs_group B) {
    return guarded([&] {
        check_group(A, "group A");
        check_group(B, "group B");
        check(out, "knn");
        if (A->ctx != B->ctx) fail(GS_ERR_INVALID_ARGUMENT, "handles of different contexts");
        require_knn_shape(out, A);
        use_device(A->ctx);
        launch_knn_brute(A, out, B, INFINITY);
        out->cur ^= 1;
        mark_result(out, A, B);
    });
}

static int flock_call(gs_group A, gs_knn knn, gs_group B, const gs_flock_term* s, const gs_flock_term* a,
                      const gs_flock_term* c, int mode) {
    return guarded([&] {
        check_group(A, "group A", GS_KIND_AGENTS);
        check_group(B, "group B", GS_KIND_AGENTS);
        check(knn, "knn");
        require_knn_shape(knn, A);
        if (mode != GS_MODE_OPENSTEER && mode != GS_MODE_GPUREF) fail(GS_ERR_INVALID_ARGUMENT, "unknown mode");
        use_device(A->ctx);
        ensure_current(knn, A, B);
        launch_flock(A, knn, B, s, a, c, mode);
    });
}

int gs_steer_separation(gs_group A, gs_knn knn, gs_group B, float min_d, float max_d, float cos_max, float w,
                        uint32_t mask, int mode) {
    const gs_flock_term t{min_d, max_d, cos_max, w, mask};
    return flock_call(A, knn, B, &t, nullptr, nullptr, mode);
}
int gs_steer_alignment(gs_group A, gs_knn knn, gs_group B, float min_d, float max_d, float cos_max, float w,
                       uint32_t mask, int mode) {
    const gs_flock_term t{min_d, max_d, cos_max, w, mask};
    return flock_call(A, knn, B, nullptr, &t, nullptr, mode);
}
int gs_steer_cohesion(gs_group A, gs_knn knn, gs_group B, float min_d, float max_d, float cos_max, float w,
                      uint32_t mask, int mode) {
    const gs_flock_term t{min_d, max_d, cos_max, w, mask};
    return flock_call(A, knn, B, nullptr, nullptr, &t, mode);
}
int gs_steer_flocking(gs_group A, gs_knn knn, gs_group B, const gs_flock_term* s, const gs_flock_term* a,
                      const gs_flock_term* c, int mode) {
    return flock_call(A, knn, B, s, a, c, mode);
}

int gs_steer_seek(gs_group A, const float target[3], float w, uint32_t mask, int mode) {
    return guarded([&] {
        check_group(A, "group A", GS_KIND_AGENTS);
        if (!target) fail(GS_ERR_INVALID_ARGUMENT, "target is NULL");
        if (mode != GS_MODE_OPENSTEER && mode != GS_MODE_GPUREF) fail(GS_ERR_INVALID_ARGUMENT, "unknown mode");
        use_device(A->ctx);
        launch_seek(A, target, w, mask, mode);
    });
}

int gs_steer_avoid_obstacles(gs_group A, gs_group O, gs_knn knn, float min_time, float w, uint32_t mask) {
    return guarded([&] {
        check_group(A, "group A", GS_KIND_AGENTS);
        check_group(O, "obstacles", GS_KIND_OBSTACLES);
        check(knn, "knn");
        require_knn_shape(knn, A);
        use_device(A->ctx);
        ensure_current(knn, A, O);
        launch_avoid_obstacles(A, O, knn, min_time, w, mask);
    });
}

int gs_steer_avoid_walls(gs_group A, gs_knn knn, gs_group W, float min_time, float w, uint32_t mask) {
    return guarded([&] {
        check_group(A, "group A", GS_KIND_AGENTS);
        check_group(W, "walls", GS_KIND_WALLS);
        check(knn, "knn");
        require_knn_shape(knn, A);
        use_device(A->ctx);
        ensure_current(knn, A, W);
        launch_avoid_walls(A, knn, W, min_time, w, mask);
    });
}

int gs_update_group(gs_group A, float dt, int mode) {
    return guarded([&] {
        check_group(A, "group A", GS_KIND_AGENTS);
        if (mode != GS_MODE_OPENSTEER && mode != GS_MODE_GPUREF) fail(GS_ERR_INVALID_ARGUMENT, "unknown mode");
        use_device(A->ctx);
        launch_update(A, dt, mode);
    });
}

int gs_update_group_anti_penetration(gs_group A, gs_knn knn, gs_group W, float dt, int closest_point) {
    return guarded([&] {
        check_group(A, "group A", GS_KIND_AGENTS);
        check_group(W, "walls", GS_KIND_WALLS);
        check(knn, "knn");
        if (closest_point != GS_CLOSEST_CLAMPED && closest_point != GS_CLOSEST_AS_REFERENCE)
            fail(GS_ERR_INVALID_ARGUMENT, "unknown closest-point rule");
        require_knn_shape(knn, A);
        use_device(A->ctx);
        ensure_current(knn, A, W);
        launch_update_anti_penetration(A, knn, W, dt, closest_point);
    });
}

int gs_anti_penetration_agents(gs_group A, gs_knn knn, gs_group B, uint32_t mask) {
    return guarded([&] {
        check_group(A, "group A", GS_KIND_AGENTS);
        check_group(B, "group B", GS_KIND_AGENTS);
        check(knn, "knn");
        require_knn_shape(knn, A);
        if (B->ctx != A->ctx) fail(GS_ERR_INVALID_ARGUMENT, "groups belong to different contexts");
        use_device(A->ctx);
        ensure_current(knn, A, B);
        launch_anti_penetration_agents(A, knn, B, mask);
    });
}

int gs_steer_avoid_neighbors(gs_group A, gs_knn knn, gs_group B, float min_time_to_collision,
                             float min_separation_distance, int avoid_close_neighbors, float w, uint32_t mask,
                             int mode) {
    return guarded([&] {
        check_group(A, "group A", GS_KIND_AGENTS);
        check_group(B, "group B", GS_KIND_AGENTS);
        check(knn, "knn");
        if (mode != GS_MODE_OPENSTEER && mode != GS_MODE_GPUREF) fail(GS_ERR_INVALID_ARGUMENT, "unknown mode");
        require_knn_shape(knn, A);
        if (B->ctx != A->ctx) fail(GS_ERR_INVALID_ARGUMENT, "groups belong to different contexts");
        use_device(A->ctx);
        ensure_current(knn, A, B);
        launch_avoid_neighbors(A, knn, B, min_time_to_collision, min_separation_distance, avoid_close_neighbors ? 1 : 0,
                               w, mask, mode);
    });
}

// PolylinePathwayCUDA::Initialize (CUDA/PolylinePathwayCUDA.cu:17-60) on the
// host, in float, in its order: m = n (+1 closing on the first point when
// cyclic); per segment end i >= 1 the difference to the previous point, its
// length, its normalized direction (unchanged when the length is 0) and the
// running total.
int gs_path_create(gs_context ctx, uint32_t n_points, const float* points, float radius, int cyclic, gs_path* out) {
    return guarded([&] {
        no_frame("gs_path_create");
        check(ctx, "context");
        if (!out) fail(GS_ERR_INVALID_ARGUMENT, "out is NULL");
        if (n_points < 2 || !points) fail(GS_ERR_INVALID_ARGUMENT, "a path needs at least 2 points");
        if (n_points > (1u << 20)) fail(GS_ERR_UNSUPPORTED, "more than 2^20 path points");
        auto* p = new gs_path_s();
        p->ctx = ctx;
        p->radius = radius;
        p->cyclic = cyclic ? 1 : 0;
        const uint32_t m = n_points + (cyclic ? 1u : 0u);
        p->m = m;
        p->pts.assign((size_t)m * 3, 0.f);
        p->nrm.assign((size_t)m * 3, 0.f);
        p->len.assign(m, 0.f);
        float total = 0.f;
        for (uint32_t i = 0; i < m; ++i) {
            const uint32_t j = (cyclic && i == m - 1) ? 0 : i;
            for (int a = 0; a < 3; ++a) p->pts[3 * i + a] = points[3 * j + a];
            if (i == 0) continue;
            const float dx = p->pts[3 * i] - p->pts[3 * i - 3];
            const float dy = p->pts[3 * i + 1] - p->pts[3 * i - 2];
            const float dz = p->pts[3 * i + 2] - p->pts[3 * i - 1];
            const float L = std::sqrt(((dx * dx) + (dy * dy)) + (dz * dz));
            p->len[i] = L;
            p->nrm[3 * i] = L > 0.f ? dx / L : dx;
            p->nrm[3 * i + 1] = L > 0.f ? dy / L : dy;
            p->nrm[3 * i + 2] = L > 0.f ? dz / L : dz;
            total += L;
        }
        p->total = total;
        std::vector<float> packed((size_t)7 * m);
        for (uint32_t i = 0; i < m; ++i) {
            for (int a = 0; a < 3; ++a) {
                packed[(size_t)a * m + i] = p->pts[3 * i + a];
                packed[(size_t)(3 + a) * m + i] = p->nrm[3 * i + a];
            }
            packed[(size_t)6 * m + i] = p->len[i];
        }
        try {
            use_device(ctx);
            p->dev.reserve(packed.size());
            h2d(p->dev.p, packed.data(), packed.size() * sizeof(float), ctx->stream);
            GS_HIP(hipStreamSynchronize(ctx->stream));
        } catch (...) {
            p->dev.release();
            delete p;
            throw;
        }
        reg_add(p);
        *out = p;
    });
}

int gs_path_info(gs_path p, uint32_t* n_points, float* total_length) {
    return guarded([&] {
        check(p, "path");
        if (n_points) *n_points = p->m;
        if (total_length) *total_length = p->total;
    });
}

int gs_path_get(gs_path p, float* points, float* directions, float* lengths) {
    return guarded([&] {
        check(p, "path");
        if (points) memcpy(points, p->pts.data(), p->pts.size() * sizeof(float));
        if (directions) memcpy(directions, p->nrm.data(), p->nrm.size() * sizeof(float));
        if (lengths) memcpy(lengths, p->len.data(), p->len.size() * sizeof(float));
    });
}

int gs_path_destroy(gs_path p) {
    return guarded([&] {
        no_frame("gs_path_destroy");
        check(p, "path");
        use_device(p->ctx);
        (void)hipStreamSynchronize(p->ctx->stream);
        p->dev.release();
        reg_del(p);
        delete p;
    });
}

int gs_steer_follow_path(gs_group A, gs_path path, float prediction_time, float w, uint32_t mask) {
    return guarded([&] {
        check_group(A, "group A", GS_KIND_AGENTS);
        check(path, "path");
        if (path->ctx != A->ctx) fail(GS_ERR_INVALID_ARGUMENT, "path belongs to another context");
        use_device(A->ctx);
        launch_follow_path(A, path, prediction_time, w, mask);
    });
}

int gs_wrap_world(gs_group A, int shape, const float center[3], const float size_or_radius[3]) {
    return guarded([&] {
        check_group(A, "group A", GS_KIND_AGENTS);
        if (!center || !size_or_radius) fail(GS_ERR_INVALID_ARGUMENT, "NULL argument");
        if (shape != GS_WRAP_BOX && shape != GS_WRAP_SPHERE) fail(GS_ERR_INVALID_ARGUMENT, "unknown wrap shape");
        use_device(A->ctx);
        launch_wrap(A, shape, center, size_or_radius);
    });
}

int gs_clear_bits(gs_group A) {
    return guarded([&] {
        check_group(A, "group A", GS_KIND_AGENTS);
        use_device(A->ctx);
        launch_clear_bits(A);
    });
}

static void check_boids_args(gs_group A, gs_knn knn, gs_bins bins, const gs_boids_params* p) {
    check_group(A, "group A", GS_KIND_AGENTS);
    check(knn, "knn");
    check(bins, "bins");
    if (!p) fail(GS_ERR_INVALID_ARGUMENT, "params is NULL");
    if (bins->ctx != A->ctx) fail(GS_ERR_INVALID_ARGUMENT, "handles of different contexts");
    require_knn_shape(knn, A);
    if (p->mode != GS_MODE_OPENSTEER && p->mode != GS_MODE_GPUREF) fail(GS_ERR_INVALID_ARGUMENT, "unknown mode");
    if (p->wrap_shape != -1 && p->wrap_shape != GS_WRAP_BOX && p->wrap_shape != GS_WRAP_SPHERE)
        fail(GS_ERR_INVALID_ARGUMENT, "unknown wrap shape");
}

// One Boids step: build + reorder (wrap fused), seeded exact KNN of A with
// itself, then flocking + integrate (new positions / forwards into the
// alternate buffers, flipped afterwards).  dist false: the search writes
// neighbour slots only -- the flocking kernel and the next step's seed read
// no distances, so only the last step of a call stores them.
//
// keys_ready: the previous step of the same call did this step's key pass;
// next: another step of the call follows (this step then does its key pass
// in the flocking kernel, unless the group keeps a bounding box).  Returns
// whether it did.
static bool boids_step_once(gs_group_s* A, gs_knn_s* knn, gs_bins_s* bins, const gs_boids_params* p,
                            hipEvent_t* ev, bool dist = true, bool keys_ready = false, bool next = false) {
    hipStream_t st = A->ctx->stream;
    if (ev) GS_HIP(hipEventRecord(ev[0], st));
    const bool lazy = flock_step_lazy_ok(A);
    launch_bins_build(A, bins, p->wrap_shape, p->wrap_center, p->wrap_size, lazy, keys_ready);
    after_reorder(A);
    if (ev) GS_HIP(hipEventRecord(ev[1], st));
    const Seed sd = seed_for(knn, A, A, p->knn_flags);
    launch_knn_binned(A, knn, A, p->search_radius, INFINITY, sd.seed, sd.row_map, sd.val_map, dist);
    if (A->n) knn->cur ^= 1;
    mark_result(knn, A, A);
    if (ev) GS_HIP(hipEventRecord(ev[2], st));
    const bool fuse = next && A->n > 0 && !A->bbox_wanted;
    launch_flock_update(A, knn, p, lazy, fuse ? bins : nullptr);
    if (ev) GS_HIP(hipEventRecord(ev[3], st));
    return fuse;
}

int gs_step_choke(gs_group A0, gs_group A1, gs_group walls, gs_group obstacles, gs_bins bins, gs_knn const* knn,
                  const gs_scenario* s, uint32_t steps) {
    return guarded([&] {
        if (!knn || !s) fail(GS_ERR_INVALID_ARGUMENT, "knn or scenario is NULL");
        check_group(A0, "group 0", GS_KIND_AGENTS);
        check_group(A1, "group 1", GS_KIND_AGENTS);
        check_group(walls, "walls", GS_KIND_WALLS);
        check_group(obstacles, "obstacles", GS_KIND_OBSTACLES);
        check(bins, "bins");
        for (int j = 0; j < 8; ++j) check(knn[j], "knn");
        if (s->mode != GS_MODE_OPENSTEER && s->mode != GS_MODE_GPUREF) fail(GS_ERR_INVALID_ARGUMENT, "unknown mode");
        // Each call below is the C entry point itself, so the sequence is
        // the one a plugin issues; a failing call ends the step with its status.
        auto ok = [](int st) {
            if (st != GS_OK) throw Error(st, gs_last_error());
        };
        gs_group A[2] = {A0, A1};
        const gs_flock_term& t = s->choke_separation;
        const float cap = s->other_max_distance > 0.f ? s->other_max_distance : INFINITY;
        for (uint32_t step = 0; step < steps; ++step) {
            ok(gs_update_knn_database(A[0], bins));
            ok(gs_update_knn_database(A[1], bins));
            for (int g = 0; g < 2; ++g) {
                gs_knn const* k = knn + 4 * g;
                ok(gs_clear_bits(A[g]));
                ok(gs_find_knn(A[g], k[0], bins, A[g], 1, GS_KNN_DEFAULT));
                ok(gs_find_knn_within(A[g], k[1], bins, A[1 - g], 1, GS_KNN_DEFAULT, cap));
                ok(gs_find_knn(A[g], k[2], bins, walls, 1, GS_KNN_DEFAULT));
                ok(gs_find_knn(A[g], k[3], bins, obstacles, 1, GS_KNN_DEFAULT));
                ok(gs_steer_avoid_walls(A[g], k[2], walls, s->walls_min_time, s->walls_weight, 0));
                ok(gs_steer_avoid_obstacles(A[g], obstacles, k[3], s->obstacles_min_time, s->obstacles_weight, 0));
                ok(gs_steer_separation(A[g], k[0], A[g], t.min_d, t.max_d, t.cos_max, t.weight, t.mask, s->mode));
                ok(gs_steer_separation(A[g], k[1], A[1 - g], t.min_d, t.max_d, t.cos_max, t.weight, t.mask,
                                       s->mode));
                ok(gs_steer_seek(A[g], s->goal[g], s->seek_weight, s->seek_mask, s->mode));
            }
            ok(gs_update_group(A[0], s->dt, s->mode));
            ok(gs_update_group(A[1], s->dt, s->mode));
        }
    });
}

int gs_step_boids(gs_group A, gs_knn knn, gs_bins bins, const gs_boids_params* p, uint32_t steps) {
    return guarded([&] {
        check_boids_args(A, knn, bins, p);
        use_device(A->ctx);
        bool keys = false;
        for (uint32_t s = 0; s < steps; ++s) keys = boids_step_once(A, knn, bins, p, nullptr, s + 1 == steps, keys,
                                                                    s + 1 < steps);
    });
}

int gs_step_boids_timed(gs_group A, gs_knn knn, gs_bins bins, const gs_boids_params* p, uint32_t steps,
                        gs_boids_timing* timing) {
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    int status = guarded([&] {
        no_frame("gs_step_boids_timed");
        check_boids_args(A, knn, bins, p);
        if (!timing) fail(GS_ERR_INVALID_ARGUMENT, "timing is NULL");
        if (steps == 0) fail(GS_ERR_INVALID_ARGUMENT, "steps must be at least 1");
        use_device(A->ctx);
        // Timing-only events: no system-scope fence, so recording them does not
        // write back and invalidate L2 between the stages being timed.
        for (auto& e : ev) GS_HIP(hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
        double t[3] = {0, 0, 0};
        for (uint32_t s = 0; s < steps; ++s) {
            boids_step_once(A, knn, bins, p, ev);
            GS_HIP(hipEventSynchronize(ev[3]));
            for (int j = 0; j < 3; ++j) {
                float ms = 0;
                GS_HIP(hipEventElapsedTime(&ms, ev[j], ev[j + 1]));
                t[j] += ms;
            }
        }
        GS_HIP(hipStreamSynchronize(A->ctx->stream));
        timing->bin_build_ms = (float)(t[0] / steps);
        timing->knn_ms = (float)(t[1] / steps);
        timing->steer_integrate_ms = (float)(t[2] / steps);
        timing->step_ms = (float)((t[0] + t[1] + t[2]) / steps);
    });
    for (auto& e : ev)
        if (e) (void)hipEventDestroy(e);
    return status;
}


// ---- recorded frames -----------------------------------------------------------

int gs_frame_begin(gs_context ctx) {
    return guarded([&] {
        check(ctx, "context");
        if (t_rec) fail(GS_ERR_FRAME, "a frame is already being recorded on this thread");
        use_device(ctx);
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        GS_HIP(hipStreamIsCapturing(ctx->stream, &cs));
        if (cs != hipStreamCaptureStatusNone) fail(GS_ERR_FRAME, "the context's stream is already being captured");
        // Thread-local mode: HIP itself also refuses unsafe calls of this
        // thread (allocation, synchronization) while recording.
        GS_HIP(hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
        auto* r = new Recording;
        r->ctx = ctx;
        t_rec = r;
    });
}

int gs_frame_end(gs_context ctx, gs_frame* out) {
    return guarded([&] {
        check(ctx, "context");
        if (!t_rec || t_rec->ctx != ctx) fail(GS_ERR_FRAME, "no frame is being recorded on this context by this thread");
        if (out) *out = nullptr;
        std::unique_ptr<Recording> r(t_rec);
        t_rec = nullptr;
        use_device(ctx);
        hipGraph_t graph = nullptr;
        const hipError_t ce = hipStreamEndCapture(ctx->stream, &graph);
        // Nothing has run: every touched handle goes back to its begin state.
        for (auto& t : r->touched) {
            t.end = snap_of(t);
            install(t, t.begin);
        }
        if (ce != hipSuccess) {
            (void)hipGetLastError();
            if (graph) (void)hipGraphDestroy(graph);
            fail(GS_ERR_FRAME, std::string("the recording was invalidated: ") + hipGetErrorString(ce) +
                                   (r->failed ? " (first failing call: " + r->first_error + ")" : ""));
        }
        if (r->failed) {
            if (graph) (void)hipGraphDestroy(graph);
            fail(GS_ERR_FRAME, "a call failed while recording the frame: " + r->first_error);
        }
        if (!out) {
            (void)hipGraphDestroy(graph);
            fail(GS_ERR_INVALID_ARGUMENT, "out is NULL");
        }
        auto* f = new gs_frame_s;
        f->ctx = ctx;
        f->graph = graph;
        size_t nodes = 0;
        hipError_t e = hipGraphGetNodes(graph, nullptr, &nodes);
        if (e == hipSuccess && nodes > 0) e = hipGraphInstantiate(&f->exec, graph, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            (void)hipGraphDestroy(graph);
            delete f;
            hip_check(e, "hipGraphInstantiate");
        }
        f->touched = std::move(r->touched);
        reg_add(f);
        *out = f;
    });
}

// A replay's precondition and renaming: the current state of every touched
// handle against its begin snapshot.
static bool frame_matches(gs_frame_s* f, std::vector<Snap>& cur, std::unordered_map<uint64_t, uint64_t>& ren,
                          std::string& why) {
    std::unordered_map<uint64_t, uint64_t> back;
    cur.clear();
    for (auto& t : f->touched) {
        if (!reg_has(t.handle())) {
            why = "a handle the frame touches was destroyed";
            return false;
        }
        cur.push_back(snap_of(t));
        const Snap& b = t.begin;
        const Snap& c = cur.back();
        if (b.exact != c.exact) {
            why = std::string("a ") + (t.g ? "group" : t.k ? "KNN container" : "handle") +
                  " the frame touches has other sizes, buffers or current double buffers than when it was recorded";
            return false;
        }
        for (size_t i = 0; i < b.epochs.size(); ++i) {
            const uint64_t x = b.epochs[i], y = c.epochs[i];
            if ((x == 0) != (y == 0)) {
                why = "a bin database or KNN result the frame relies on is missing (or one it lacked is present)";
                return false;
            }
            if (x == 0) continue;
            auto fx = ren.emplace(x, y);
            auto by = back.emplace(y, x);
            if (fx.first->second != y || by.first->second != x) {
                why = "the KNN results the frame relies on are for other orders of their groups than when recorded";
                return false;
            }
        }
    }
    return true;
}

int gs_frame_launch(gs_frame f, uint32_t times) {
    return guarded([&] {
        check(f, "frame");
        no_frame("gs_frame_launch");
        check(f->ctx, "frame's context");
        use_device(f->ctx);
        std::vector<Snap> cur;
        for (uint32_t rep = 0; rep < times; ++rep) {
            std::unordered_map<uint64_t, uint64_t> ren;
            std::string why;
            if (!frame_matches(f, cur, ren, why))
                fail(GS_ERR_FRAME, "replay " + std::to_string(rep + 1) + " of " + std::to_string(times) + ": " + why);
            if (f->exec) GS_HIP(hipGraphLaunch(f->exec, f->ctx->stream));
            // Install the end state, recorded epochs renamed, new ones fresh.
            for (auto& t : f->touched) {
                Snap e = t.end;
                for (auto& x : e.epochs) {
                    if (x == 0) continue;
                    auto it = ren.find(x);
                    if (it == ren.end()) it = ren.emplace(x, next_epoch()).first;
                    x = it->second;
                }
                install(t, e);
            }
        }
    });
}

int gs_frame_destroy(gs_frame f) {
    return guarded([&] {
        if (!f) return;
        check(f, "frame");
        no_frame("gs_frame_destroy");
        if (reg_has(f->ctx)) use_device(f->ctx);
        if (f->exec) (void)hipGraphExecDestroy(f->exec);
        if (f->graph) (void)hipGraphDestroy(f->graph);
        reg_del(f);
        delete f;
    });
}

// ---- batched ensemble -----------------------------------------------------------

static gs_ensemble_s* check_ens(gs_ensemble e) { return check(e, "ensemble"); }

static void release_ensemble(gs_ensemble_s* e) {
    for (int p = 0; p < 2; ++p) {
        if (e->exec[p]) (void)hipGraphExecDestroy(e->exec[p]);
        if (e->graph[p]) (void)hipGraphDestroy(e->graph[p]);
        e->exec[p] = nullptr;
        e->graph[p] = nullptr;
    }
}

static void drop_graphs(gs_ensemble_s* e) { release_ensemble(e); }

int gs_ensemble_create(gs_context ctx, uint32_t ns, const gs_scenario* sc, uint32_t k, gs_ensemble* out) {
    return guarded([&] {
        no_frame("gs_ensemble_create");
        check(ctx, "context");
        if (!out) fail(GS_ERR_INVALID_ARGUMENT, "out is NULL");
        *out = nullptr;
        if (ns > 0 && !sc) fail(GS_ERR_INVALID_ARGUMENT, "scenarios is NULL");
        if (k < 1 || k > 8) fail(GS_ERR_UNSUPPORTED, "ensemble k must be 1..8");
        uint64_t slots = 0, cells = 0;
        for (uint32_t i = 0; i < ns; ++i) {
            const gs_scenario& d = sc[i];
            if (d.kind != GS_SCENARIO_BOIDS && d.kind != GS_SCENARIO_CHOKE)
                fail(GS_ERR_INVALID_ARGUMENT, "scenario " + std::to_string(i) + ": unknown kind");
            if (d.mode != GS_MODE_OPENSTEER && d.mode != GS_MODE_GPUREF)
                fail(GS_ERR_INVALID_ARGUMENT, "scenario " + std::to_string(i) + ": unknown mode");
            if (d.kind == GS_SCENARIO_BOIDS && d.wrap_shape != GS_WRAP_BOX && d.wrap_shape != GS_WRAP_SPHERE &&
                d.wrap_shape != -1)
                fail(GS_ERR_INVALID_ARGUMENT, "scenario " + std::to_string(i) + ": unknown wrap shape");
            if (d.kind == GS_SCENARIO_CHOKE &&
                (d.k_walls < 1 || d.k_walls > 8 || d.k_obstacles < 1 || d.k_obstacles > 8))
                fail(GS_ERR_UNSUPPORTED, "scenario " + std::to_string(i) + ": k_walls / k_obstacles must be 1..8");
            (void)make_grid(d.origin, d.size, d.cells);
            const int ng = d.kind == GS_SCENARIO_CHOKE ? 2 : 1;
            for (int g = 0; g < ng; ++g) {
                slots += (d.n[g] + kBlock - 1) / kBlock * (uint64_t)kBlock;
                cells += (uint64_t)d.cells[0] * d.cells[1] * d.cells[2] + 1;
            }
        }
        if (slots >= (1ull << 31) || cells >= (1ull << 31)) fail(GS_ERR_UNSUPPORTED, "ensemble too large");
        use_device(ctx);
        auto* e = new gs_ensemble_s();
        e->ctx = ctx;
        try {
            ens_init(e, ns, sc, k, make_grid);
            e->walls_h.assign(ns, {});
            e->obs_h.assign(ns, {});
            ens_alloc(e);
            ens_upload_geometry(e);
            GS_HIP(hipStreamSynchronize(ctx->stream));
        } catch (...) {
            ens_release(e);
            delete e;
            throw;
        }
        reg_add(e);
        *out = e;
    });
}

int gs_ensemble_destroy(gs_ensemble e) {
    return guarded([&] {
        no_frame("gs_ensemble_destroy");
        check_ens(e);
        reg_del(e);
        use_device(e->ctx);
        (void)hipStreamSynchronize(e->ctx->stream);
        release_ensemble(e);
        ens_release(e);
        delete e;
    });
}

int gs_ensemble_slots(gs_ensemble e, uint32_t* slots) {
    return guarded([&] {
        check_ens(e);
        if (!slots) fail(GS_ERR_INVALID_ARGUMENT, "slots is NULL");
        *slots = e->N;
    });
}

static const EnsPop& ens_pop(gs_ensemble_s* e, uint32_t scenario, uint32_t group) {
    if (scenario >= e->scen.size()) fail(GS_ERR_INVALID_ARGUMENT, "scenario index out of range");
    const int ng = e->scen[scenario].kind == GS_SCENARIO_CHOKE ? 2 : 1;
    if ((int)group >= ng) fail(GS_ERR_INVALID_ARGUMENT, "group index out of range");
    for (const auto& P : e->pops)
        if (P.scen == scenario && P.group == (int)group) return P;
    fail(GS_ERR_INVALID_ARGUMENT, "no such population");
}

int gs_ensemble_upload_agents(gs_ensemble e, uint32_t scenario, uint32_t group, const gs_agents* host) {
    return guarded([&] {
        no_frame("gs_ensemble_upload_agents");
        check_ens(e);
        const EnsPop& P = ens_pop(e, scenario, group);
        if (P.n && !host) fail(GS_ERR_INVALID_ARGUMENT, "host arrays are NULL");
        const float* const* hf = (const float* const*)host;
        for (int j = 0; j < kFields && P.n; ++j)
            if (!hf[j]) fail(GS_ERR_INVALID_ARGUMENT, "agent array is NULL");
        use_device(e->ctx);
        hipStream_t st = e->ctx->stream;
        GS_HIP(hipStreamSynchronize(st));
        std::vector<uint32_t> ids(P.n);
        for (uint32_t i = 0; i < P.n; ++i) ids[i] = i;
        for (int j = 0; j < kFields && P.n; ++j) h2d(e->F(j) + P.a0, hf[j], (size_t)P.n * 4, st);
        std::vector<uint32_t> bits(P.n, 0u);
        if (P.n && host->bits) memcpy(bits.data(), host->bits, (size_t)P.n * 4);
        if (P.n) {
            h2d(e->bits() + P.a0, bits.data(), (size_t)P.n * 4, st);
            h2d(e->ids() + P.a0, ids.data(), (size_t)P.n * 4, st);
        }
        GS_HIP(hipStreamSynchronize(st));
        e->seeded = false;
        // Which of the never-written fields hold one value in this population
        // (bit patterns), and so need not move in the counting sort.
        const uint32_t consts = (0x7u << 7) | (0xFu << 19);
        uint32_t uni = 0;
        for (int j = 0; j < kFields; ++j) {
            if (!(consts & (1u << j))) continue;
            const uint32_t* w = P.n ? (const uint32_t*)hf[j] : nullptr;
            bool same = true;
            for (uint32_t i = 1; i < P.n && same; ++i) same = w[i] == w[0];
            if (same) uni |= 1u << j;
        }
        if (e->pop_uniform.size() != e->pops.size()) e->pop_uniform.assign(e->pops.size(), 0u);
        e->pop_uniform[(size_t)(&P - e->pops.data())] = uni;
        uint32_t frozen = consts;
        for (uint32_t u : e->pop_uniform) frozen &= u;
        bool opensteer = false;
        for (const gs_scenario& d : e->scen) opensteer = opensteer || d.mode == GS_MODE_OPENSTEER;
        if (opensteer) frozen &= ~(0x7u << 7);  // the smoothed acceleration is written
        if (frozen != e->frozen) {
            drop_graphs(e);  // captured steps move the fields of the old set
            e->frozen = frozen;
        }
        // The uploaded steering and bits may be non-zero: the next step moves them.
        e->steer_zero = false;
    });
}

int gs_ensemble_download_agents(gs_ensemble e, uint32_t scenario, uint32_t group, gs_agents* host) {
    return guarded([&] {
        no_frame("gs_ensemble_download_agents");
        check_ens(e);
        const EnsPop& P = ens_pop(e, scenario, group);
        if (!host) fail(GS_ERR_INVALID_ARGUMENT, "host arrays are NULL");
        if (!P.n) return;
        use_device(e->ctx);
        hipStream_t st = e->ctx->stream;
        auto ids = d2h_vec(e->ids() + P.a0, P.n, st);
        GS_HIP(hipStreamSynchronize(st));
        for (uint32_t s = 0; s < P.n; ++s)
            if (ids[s] >= P.n) fail(GS_ERR_DEVICE, "ensemble id map corrupted");
        float* const* hf = (float* const*)host;
        for (int j = 0; j < kFields; ++j) {
            if (!hf[j]) continue;
            auto v = d2h_vec(e->F(j) + P.a0, P.n, st);
            GS_HIP(hipStreamSynchronize(st));
            for (uint32_t s = 0; s < P.n; ++s) hf[j][ids[s]] = v[s];
        }
        if (host->bits) {
            auto b = d2h_vec(e->bits() + P.a0, P.n, st);
            GS_HIP(hipStreamSynchronize(st));
            for (uint32_t s = 0; s < P.n; ++s) host->bits[ids[s]] = b[s];
        }
    });
}

static void ens_upload_walls(gs_ensemble e, uint32_t scenario, uint32_t n, const gs_walls* host,
                             const float* const mid[3]) {
    check_ens(e);
    if (scenario >= e->scen.size()) fail(GS_ERR_INVALID_ARGUMENT, "scenario index out of range");
    if (n && !host) fail(GS_ERR_INVALID_ARGUMENT, "wall arrays are NULL");
    const float* const* hw = (const float* const*)host;
    for (int j = 0; j < 9 && n; ++j)
        if (!hw[j]) fail(GS_ERR_INVALID_ARGUMENT, "wall array is NULL");
    if (mid && n && (!mid[0] || !mid[1] || !mid[2])) fail(GS_ERR_INVALID_ARGUMENT, "midpoint arrays are NULL");
    // 12 floats per wall: start, end, normal, KNN position.  Without given
    // positions a wall's is its midpoint 0.5f*(start+end)
    // (CUDA/WallGroupData.cu:87-89, WallGroup.h:31).
    auto& v = e->walls_h[scenario];
    v.assign((size_t)n * 12, 0.f);
    for (uint32_t i = 0; i < n; ++i) {
        float* r = &v[(size_t)i * 12];
        for (int j = 0; j < 9; ++j) r[j] = hw[j][i];
        for (int a = 0; a < 3; ++a) r[9 + a] = mid ? mid[a][i] : 0.5f * (r[a] + r[3 + a]);
    }
    use_device(e->ctx);
    drop_graphs(e);
    ens_upload_geometry(e);
    GS_HIP(hipStreamSynchronize(e->ctx->stream));
}

int gs_ensemble_upload_walls(gs_ensemble e, uint32_t scenario, uint32_t n, const gs_walls* host) {
    return guarded([&] {
        no_frame("gs_ensemble_upload_walls");
        ens_upload_walls(e, scenario, n, host, nullptr);
    });
}

int gs_ensemble_upload_walls_mid(gs_ensemble e, uint32_t scenario, uint32_t n, const gs_walls* host,
                                 const float* mid_x, const float* mid_y, const float* mid_z) {
    return guarded([&] {
        no_frame("gs_ensemble_upload_walls_mid");
        const float* mid[3] = {mid_x, mid_y, mid_z};
        ens_upload_walls(e, scenario, n, host, mid);
    });
}

int gs_ensemble_upload_obstacles(gs_ensemble e, uint32_t scenario, uint32_t n, const float* cx, const float* cy,
                                 const float* cz, const float* radius) {
    return guarded([&] {
        no_frame("gs_ensemble_upload_obstacles");
        check_ens(e);
        if (scenario >= e->scen.size()) fail(GS_ERR_INVALID_ARGUMENT, "scenario index out of range");
        if (n && (!cx || !cy || !cz || !radius)) fail(GS_ERR_INVALID_ARGUMENT, "obstacle arrays are NULL");
        auto& v = e->obs_h[scenario];
        v.assign((size_t)n * 4, 0.f);
        for (uint32_t i = 0; i < n; ++i) {
            v[(size_t)i * 4 + 0] = cx[i];
            v[(size_t)i * 4 + 1] = cy[i];
            v[(size_t)i * 4 + 2] = cz[i];
            v[(size_t)i * 4 + 3] = radius[i];
        }
        use_device(e->ctx);
        drop_graphs(e);
        ens_upload_geometry(e);
        GS_HIP(hipStreamSynchronize(e->ctx->stream));
    });
}

int gs_ensemble_download_knn(gs_ensemble e, uint32_t scenario, uint32_t group, uint32_t* idx) {
    return guarded([&] {
        no_frame("gs_ensemble_download_knn");
        check_ens(e);
        const EnsPop& P = ens_pop(e, scenario, group);
        if (!idx) fail(GS_ERR_INVALID_ARGUMENT, "idx is NULL");
        if (!e->seeded) fail(GS_ERR_NOT_BINNED, "the ensemble has no KNN result yet");
        if (!P.n) return;
        use_device(e->ctx);
        hipStream_t st = e->ctx->stream;
        auto ids = d2h_vec(e->ids(), e->N, st);
        std::vector<uint32_t> col((size_t)e->k * P.n);
        for (uint32_t j = 0; j < e->k; ++j)
            d2h(col.data() + (size_t)j * P.n, e->knn_self[e->kcur].p + (size_t)j * e->N + P.a0, (size_t)P.n * 4, st);
        GS_HIP(hipStreamSynchronize(st));
        for (uint32_t j = 0; j < e->k; ++j)
            for (uint32_t s = 0; s < P.n; ++s) {
                const uint32_t v = col[(size_t)j * P.n + s];
                idx[(size_t)j * P.n + ids[P.a0 + s]] = v == UINT32_MAX ? UINT32_MAX : ids[v];
            }
    });
}

int gs_ensemble_step(gs_ensemble e, uint32_t steps, uint32_t flags) {
    return guarded([&] {
        no_frame("gs_ensemble_step");
        check_ens(e);
        if (flags & ~(uint32_t)GS_ENSEMBLE_EAGER) fail(GS_ERR_INVALID_ARGUMENT, "unknown flags");
        use_device(e->ctx);
        hipStream_t st = e->ctx->stream;
        uint32_t left = steps;
        if (left && !e->seeded) {
            ens_step(e);  // the first step has no previous KNN to bound its search
            --left;
        }
        if (flags & GS_ENSEMBLE_EAGER) {
            for (; left; --left) ens_step(e);
        } else {
            while (left >= 2) {
                // Captured pairs of steps start at either parity of the ids
                // (every moved field flips with them; frozen ones stay put).
                const int par = e->icur;
                if (!e->exec[par]) {
                    GS_HIP(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
                    try {
                        ens_step(e);
                        ens_step(e);
                    } catch (...) {
                        hipGraph_t g = nullptr;
                        (void)hipStreamEndCapture(st, &g);
                        if (g) (void)hipGraphDestroy(g);
                        throw;
                    }
                    GS_HIP(hipStreamEndCapture(st, &e->graph[par]));
                    GS_HIP(hipGraphInstantiate(&e->exec[par], e->graph[par], nullptr, nullptr, 0));
                }
                GS_HIP(hipGraphLaunch(e->exec[par], st));
                left -= 2;
            }
            if (left) ens_step(e);
        }
        GS_HIP(hipGetLastError());
    });
}

// ---- differentiable calibration ------------------------------------------------

int gs_calib_create(gs_ensemble e, const gs_calib_config* cfg, gs_calib* out) {
    return guarded([&] {
        no_frame("gs_calib_create");
        check_ens(e);
        if (!cfg || !out) fail(GS_ERR_INVALID_ARGUMENT, "config / out is NULL");
        *out = nullptr;
        if (cfg->steps == 0 || cfg->checkpoint_every == 0)
            fail(GS_ERR_INVALID_ARGUMENT, "steps and checkpoint_every must be positive");
        if (!(cfg->max_force > 0.f) || !(cfg->mass > 0.f) || !(cfg->dt > 0.f))
            fail(GS_ERR_INVALID_ARGUMENT, "max_force, mass and dt must be positive");
        for (const auto& d : e->scen)
            if (d.kind != GS_SCENARIO_CHOKE) fail(GS_ERR_UNSUPPORTED, "calibration needs crowd scenarios only");
        use_device(e->ctx);
        auto* c = new gs_calib_s();
        try {
            calib_init(c, e, *cfg);
        } catch (...) {
            calib_release(c);
            delete c;
            throw;
        }
        reg_add(c);
        *out = c;
    });
}

int gs_calib_destroy(gs_calib c) {
    return guarded([&] {
        no_frame("gs_calib_destroy");
        check(c, "calibration");
        reg_del(c);
        use_device(c->ens->ctx);
        (void)hipStreamSynchronize(c->ens->ctx->stream);
        calib_release(c);
        delete c;
    });
}

int gs_calib_forward(gs_calib c, const gs_calib_params* params, double* loss) {
    return guarded([&] {
        no_frame("gs_calib_forward");
        check(c, "calibration");
        check_ens(c->ens);
        if (!params || !loss) fail(GS_ERR_INVALID_ARGUMENT, "params / loss is NULL");
        const float* p = (const float*)params;
        for (int j = 0; j < 14; ++j)
            if (!std::isfinite(p[j])) fail(GS_ERR_INVALID_ARGUMENT, "parameters must be finite");
        for (int g = 0; g < 2; ++g)
            if (!(params->radius[g] > 0.f) || !(params->max_speed[g] > 0.f))
                fail(GS_ERR_INVALID_ARGUMENT, "radius and max_speed must be positive");
        use_device(c->ens->ctx);
        *loss = calib_forward(c, *params);
    });
}

int gs_calib_backward(gs_calib c, gs_calib_params* grad) {
    return guarded([&] {
        no_frame("gs_calib_backward");
        check(c, "calibration");
        check_ens(c->ens);
        if (!grad) fail(GS_ERR_INVALID_ARGUMENT, "grad is NULL");
        if (!c->forward_done) fail(GS_ERR_INVALID_ARGUMENT, "call gs_calib_forward first");
        use_device(c->ens->ctx);
        calib_backward(c, *grad);
    });
}

int gs_calib_download_state(gs_calib c, uint32_t scenario, uint32_t group, gs_agents* host) {
    return guarded([&] {
        no_frame("gs_calib_download_state");
        check(c, "calibration");
        if (!host) fail(GS_ERR_INVALID_ARGUMENT, "host is NULL");
        if (!c->forward_done) fail(GS_ERR_INVALID_ARGUMENT, "call gs_calib_forward first");
        const EnsPop& P = ens_pop(c->ens, scenario, group);
        use_device(c->ens->ctx);
        calib_download(c, P, host);
    });
}

int gs_calib_timing(gs_calib c, float* forward_ms, float* backward_ms) {
    return guarded([&] {
        no_frame("gs_calib_timing");
        check(c, "calibration");
        if (forward_ms) *forward_ms = c->fwd_ms;
        if (backward_ms) *backward_ms = c->bwd_ms;
    });
}

}  // extern "C"
