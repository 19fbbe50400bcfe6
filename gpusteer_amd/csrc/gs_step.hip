// Boids step: the per-frame sequence of plugins/BoidsGroup.cpp:606-637
// (wrapWorld, updateKNNDatabase, findKNearestNeighbors, steerForSeparation /
// Alignment / Cohesion, updateGroup) as
//
//   bin build (gs_bins.hip) with the wrap fused into the key pass; it puts
//   the position snapshot in spos and, lazily, forward + speed in sfwd;
//   exact seeded KNN of the group with itself (gs_knn.hip);
//   k_flock_update: one lane per storage slot -> the three flocking terms
//   -> integrate -> new positions and forward vectors into the alternate
//   buffers, the agent's other fields in place (or, lazily, read at the
//   previous slot and written in the new order); when the same call steps
//   again, also the next step's wrap and key pass for the new positions, so
//   that step's build starts at the scan.
//
// The lazy form applies when steering and bits are zero and the per-agent
// constants are uniform (they need no moving): the build then moves no
// state field, and each field crosses memory once per step.
//
// Other agents are only ever read through the snapshot (positions) and the
// current forward buffers, which no lane writes, so the update is Jacobi (all
// forces from the step-t state) exactly like the separate launches, and the
// results are identical to gs_wrap_world + gs_update_knn_database +
// gs_find_knn + gs_steer_flocking + gs_update_group.
#include "gs_ops.hpp"

namespace gs {

// Group-wide facts the host tracks: steering and bits hold 0 in every slot
// (then they are neither read nor rewritten), and which per-agent constants
// hold one value in every slot (then that value is passed instead of read).
struct StepUniforms {
    uint32_t steer_zero;
    uint32_t has;  // bit 0 radius, 1 maxSpeed, 2 maxForce, 3 mass
    float v[4];
};

// Lazy step (the bin build moved only the snapshots): the agent's own
// fields are read at its previous slot new_to_old[s], forward and speed come
// from the (forward, speed) snapshot, and every result is written at s into
// the alternate buffers (acc only when it changes or is not uniform).
struct LazyIO {
    const uint32_t* new_to_old;  // NULL: not lazy
    const float4* sfwd;
    float *speed, *sx, *sy, *sz, *ux, *uy, *uz, *ax, *ay, *az;  // alternate buffers
    uint32_t write_acc;
};

// The next step's key pass, done by this step for the positions it writes
// (when the same call steps again on the same lattice): the world wrap of the
// next step's start, the cell key and the histogram count with arrival rank,
// exactly as k_bin_key would.  key == NULL: not done.
struct NextKeys {
    GridDev g;
    uint32_t *key, *rank, *count;
    int wrap_shape;
    float wc[3], ws[3];
};

// Lists longer than kFlockChunk entries are gathered and summed a chunk at a
// time, in the same slot order (so the same sums): register arrays for a
// whole 16-entry list would hold the kernel to three wavefronts per SIMD.
constexpr int kFlockChunk = 8;

// flock_core over the KNN row s, gathering kFlockChunk neighbours at a time
// (positions from spos; forward vectors from sfwd when given, else from
// fx / fy / fz, only when alignment is live).  Slots past the first pad
// reuse slot 0's neighbour so every load of a chunk is in flight together.
template <int KMAX, int MODE>
__device__ __forceinline__ void flock_chunked(const uint32_t* __restrict__ knn, uint32_t k, uint32_t n, uint32_t s,
                                              const float4* __restrict__ spos, const float4* __restrict__ sfwd,
                                              const float* __restrict__ fx, const float* __restrict__ fy,
                                              const float* __restrict__ fz, f3 pos, f3 fwd, float radius,
                                              const FlockTerm& sep, const FlockTerm& ali, const FlockTerm& coh,
                                              f3& steering, uint32_t& bits) {
    FlockAcc a = flock_begin(fwd, bits, sep, ali, coh);
    if (!flock_any(a)) return;
    uint32_t v[KMAX];
#pragma unroll
    for (int j = 0; j < KMAX; ++j) v[j] = knn[(size_t)min(j, (int)k - 1) * n + s];
    int m = (int)k;
#pragma unroll
    for (int j = KMAX - 1; j >= 0; --j)
        if (j < (int)k && v[j] >= n) m = j;
#pragma unroll
    for (int c = 0; c < KMAX; c += kFlockChunk) {
        if (c >= m) break;
        f3 bp[kFlockChunk], bf[kFlockChunk];
#pragma unroll
        for (int j = 0; j < kFlockChunk; ++j) {
            const uint32_t b = (c + j < m) ? v[c + j] : v[0];
            const float4 q = spos[b];
            bp[j] = mk3(q.x, q.y, q.z);
            bf[j] = mk3(0.f, 0.f, 0.f);
        }
        if (a.do_a) {
#pragma unroll
            for (int j = 0; j < kFlockChunk; ++j) {
                const uint32_t b = (c + j < m) ? v[c + j] : v[0];
                if (sfwd) {
                    const float4 f = sfwd[b];
                    bf[j] = mk3(f.x, f.y, f.z);
                } else {
                    bf[j] = ld3(fx, fy, fz, b);
                }
            }
        }
        flock_accumulate<MODE, kFlockChunk>(a, pos, fwd, radius, min(m - c, kFlockChunk), bp, bf, sep, ali, coh);
    }
    flock_finish(a, pos, fwd, sep, ali, coh, steering, bits);
}

// Returns the cell key of the new position when N.key is set.
template <int KMAX, int MODE>
__device__ __forceinline__ uint32_t flock_update_slot(uint32_t s, uint32_t n, Agents A, StepUniforms U, LazyIO Z,
                                                      const NextKeys& N,
                                                  float* __restrict__ npx, float* __restrict__ npy,
                                                  float* __restrict__ npz, float* __restrict__ nfx,
                                                  float* __restrict__ nfy, float* __restrict__ nfz,
                                                  const float4* __restrict__ spos, const uint32_t* __restrict__ knn,
                                                  uint32_t k, FlockTerm sep, FlockTerm ali, FlockTerm coh, float dt) {
    const bool lazy = Z.new_to_old != nullptr;
    // Own fields: at s (state in this order) or at the previous slot o.
    const uint32_t o = lazy ? Z.new_to_old[s] : s;
    // Step-t snapshot: positions in spos, forward vectors in the current
    // buffers (or the snapshot); results go to the alternate buffers.
    const float4 q = spos[s];
    const f3 pos = mk3(q.x, q.y, q.z);
    f3 fwd;
    float speed;
    if (lazy) {
        const float4 fs = Z.sfwd[s];
        fwd = mk3(fs.x, fs.y, fs.z);
        speed = fs.w;
    } else {
        fwd = ld3(A.fx, A.fy, A.fz, s);
        speed = A.speed[s];
    }
    uint32_t bits = U.steer_zero ? 0u : A.bits[s];
    f3 steering;
    const float radius = (U.has & 1u) ? U.v[0] : A.radius[s];
    const float max_speed = (U.has & 2u) ? U.v[1] : A.max_speed[s];
    const float max_force = (U.has & 4u) ? U.v[2] : A.max_force[s];
    const float mass = (U.has & 8u) ? U.v[3] : A.mass[s];
    if constexpr (KMAX > kFlockChunk) {
        steering = U.steer_zero ? mk3(0.f, 0.f, 0.f) : ld3(A.gx, A.gy, A.gz, s);
        flock_chunked<KMAX, MODE>(knn, k, n, s, spos, lazy ? Z.sfwd : nullptr, A.fx, A.fy, A.fz, pos, fwd, radius,
                                  sep, ali, coh, steering, bits);
    } else {
        f3 bp[KMAX], bf[KMAX];
        const int m = gather_neighbours<KMAX>(knn, k, n, s, n, spos, nullptr, nullptr, nullptr, A.fx, A.fy, A.fz,
                                              ali.on && !(bits & ali.mask), bp, bf, lazy ? Z.sfwd : nullptr);
        steering = U.steer_zero ? mk3(0.f, 0.f, 0.f) : ld3(A.gx, A.gy, A.gz, s);
        flock_core<MODE, KMAX>(pos, fwd, radius, m, bp, bf, sep, ali, coh, steering, bits);
    }
    // Integrate (clears steering and bits as updateGroup does).
    f3 p = pos, f = fwd;
    f3 side = ld3(A.sx, A.sy, A.sz, o), up = ld3(A.ux, A.uy, A.uz, o);
    f3 sacc = ld3(A.ax, A.ay, A.az, o);
    integrate<MODE>(steering, p, f, speed, sacc, side, up, max_force, max_speed, mass, dt);
    uint32_t c = 0u;
    if (N.key) {
        wrap_point(N.wrap_shape, N.wc[0], N.wc[1], N.wc[2], N.ws[0], N.ws[1], N.ws[2], p.x, p.y, p.z);
        const GridDev& g = N.g;
        c = cell_index(g, axis_cell(p.x, g.ox, g.sx, g.dx), axis_cell(p.y, g.oy, g.sy, g.dy),
                       axis_cell(p.z, g.oz, g.sz, g.dz));
        N.key[s] = c;
    }
    st3(npx, npy, npz, s, p);
    st3(nfx, nfy, nfz, s, f);
    if (lazy) {
        st3(Z.sx, Z.sy, Z.sz, s, side);
        st3(Z.ux, Z.uy, Z.uz, s, up);
        Z.speed[s] = speed;
        if (Z.write_acc) st3(Z.ax, Z.ay, Z.az, s, sacc);
    } else {
        st3(A.sx, A.sy, A.sz, s, side);
        st3(A.ux, A.uy, A.uz, s, up);
        if (MODE == GS_MODE_OPENSTEER) st3(A.ax, A.ay, A.az, s, sacc);
        A.speed[s] = speed;
    }
    if (!U.steer_zero) {
        st3(A.gx, A.gy, A.gz, s, mk3(0.f, 0.f, 0.f));
        A.bits[s] = 0u;
    }
    return c;
}

// One lane per agent slot; the launch shape (BS threads per block, APL slots
// per lane: a block covers APL consecutive runs of BS slots) is picked by the
// autotuner for k = 7 (gs_context_set_shape).
template <int KMAX, int MODE, int BS = kBlock, int APL = 1>
__global__ __launch_bounds__(BS) void k_flock_update(uint32_t n, Agents A, StepUniforms U, LazyIO Z, NextKeys N,
                                                     float* __restrict__ npx, float* __restrict__ npy,
                                                     float* __restrict__ npz, float* __restrict__ nfx,
                                                     float* __restrict__ nfy, float* __restrict__ nfz,
                                                     const float4* __restrict__ spos,
                                                     const uint32_t* __restrict__ knn, uint32_t k, FlockTerm sep,
                                                     FlockTerm ali, FlockTerm coh, float dt) {
    // Every lane of the wavefront takes part in the histogram count (lanes
    // past n with a key no cell has, as in k_bin_key).
    auto one = [&](uint32_t s) {
        const bool active = s < n;
        uint32_t c = 0xFFFFFFC0u | (threadIdx.x & 63u);
        if (active)
            c = flock_update_slot<KMAX, MODE>(s, n, A, U, Z, N, npx, npy, npz, nfx, nfy, nfz, spos, knn, k, sep,
                                              ali, coh, dt);
        if (N.key) count_and_rank(active, c, s, N.count, N.rank);
    };
    if constexpr (APL == 1) {
        one(xcd_block(blockIdx.x, gridDim.x) * BS + threadIdx.x);
    } else {
        for (int a = 0; a < APL; ++a) one((xcd_block(blockIdx.x, gridDim.x) * APL + a) * BS + threadIdx.x);
    }
}

static FlockTerm fterm(const gs_flock_term& t) { return FlockTerm{t.min_d, t.max_d, t.cos_max, t.weight, t.mask, 1}; }

static StepUniforms step_uniforms(const gs_group_s* A) {
    StepUniforms U{};
    U.steer_zero = (A->inv & kFSteer) == kFSteer && A->bits_zero &&
                   __builtin_bit_cast(uint32_t, A->inv_val[16]) == 0u &&
                   __builtin_bit_cast(uint32_t, A->inv_val[17]) == 0u &&
                   __builtin_bit_cast(uint32_t, A->inv_val[18]) == 0u;
    for (int c = 0; c < 4; ++c)
        if (A->inv & (1u << (19 + c))) {
            U.has |= 1u << c;
            U.v[c] = A->inv_val[19 + c];
        }
    return U;
}

bool flock_step_lazy_ok(const gs_group_s* A) {
    const StepUniforms U = step_uniforms(A);
    return A->kind == GS_KIND_AGENTS && U.steer_zero && U.has == 0xFu;
}

template <int KMAX>
static void run_flock_update(gs_group_s* A, gs_knn_s* knn, const gs_boids_params* p, bool lazy,
                             const gs_bins_s* next_bins) {
    const uint32_t n = A->n;
    hipStream_t st = A->ctx->stream;
    const Agents a = A->dev_agents();
    const StepUniforms U = step_uniforms(A);
    LazyIO Z{};
    const bool acc_written = p->mode == GS_MODE_OPENSTEER || !(A->inv & (0x7u << 7));
    if (lazy) {
        Z.new_to_old = A->new_to_old.p;
        Z.sfwd = A->sfwd.p;
        Z.speed = A->Falt(6);
        Z.sx = A->Falt(10);
        Z.sy = A->Falt(11);
        Z.sz = A->Falt(12);
        Z.ux = A->Falt(13);
        Z.uy = A->Falt(14);
        Z.uz = A->Falt(15);
        Z.ax = A->Falt(7);
        Z.ay = A->Falt(8);
        Z.az = A->Falt(9);
        Z.write_acc = acc_written ? 1u : 0u;
    }
    NextKeys N{};
    if (next_bins) {
        N.g = next_bins->g;
        N.key = A->key.p;
        N.rank = A->rank.p;
        N.count = A->count.p;
        N.wrap_shape = p->wrap_shape;
        for (int j = 0; j < 3; ++j) {
            N.wc[j] = p->wrap_center[j];
            N.ws[j] = p->wrap_size[j];
        }
    }
    const LaunchShape sh = A->ctx->shape[GS_SHAPE_FLOCK];
#define GS_FU(MODE, BS, APL)                                                                                         \
    k_flock_update<KMAX, MODE, BS, APL><<<blocks_for(n, BS * APL), BS, 0, st>>>(                                     \
        n, a, U, Z, N, A->Falt(0), A->Falt(1), A->Falt(2), A->Falt(3), A->Falt(4), A->Falt(5), A->spos.p, knn->idx(), \
        knn->k, fterm(p->separation), fterm(p->alignment), fterm(p->cohesion), p->dt)
#define GS_FU_SHAPES(MODE)                                                                                           \
    if constexpr (KMAX == 7) {                                                                                       \
        if (sh.block == 128 && sh.apl == 1) GS_FU(MODE, 128, 1);                                                     \
        else if (sh.block == 128 && sh.apl == 2) GS_FU(MODE, 128, 2);                                                \
        else if (sh.block == 256 && sh.apl == 2) GS_FU(MODE, 256, 2);                                                \
        else if (sh.block == 512 && sh.apl == 1) GS_FU(MODE, 512, 1);                                                \
        else if (sh.block == 512 && sh.apl == 2) GS_FU(MODE, 512, 2);                                                \
        else GS_FU(MODE, kBlock, 1);                                                                                 \
    } else {                                                                                                         \
        GS_FU(MODE, kBlock, 1);                                                                                      \
    }
    if (p->mode == GS_MODE_GPUREF) {
        GS_FU_SHAPES(GS_MODE_GPUREF)
    } else {
        GS_FU_SHAPES(GS_MODE_OPENSTEER)
    }
#undef GS_FU_SHAPES
#undef GS_FU
    GS_LAUNCHED("k_flock_update");
    // New positions and forward vectors (and, lazily, every other result)
    // are in the alternate buffers.
    for (int j = 0; j < 6; ++j) A->flip(j);
    if (lazy) {
        A->flip(6);
        for (int j = 10; j < 16; ++j) A->flip(j);
        if (acc_written)
            for (int j = 7; j < 10; ++j) A->flip(j);
    }
}

void launch_flock_update(gs_group_s* A, gs_knn_s* knn, const gs_boids_params* p, bool lazy,
                         const gs_bins_s* next_bins) {
    if (A->n == 0) return;
    if (knn->k == 7) run_flock_update<7>(A, knn, p, lazy, next_bins);
    else if (knn->k <= 8) run_flock_update<8>(A, knn, p, lazy, next_bins);
    else if (knn->k <= 16) run_flock_update<16>(A, knn, p, lazy, next_bins);
    else run_flock_update<32>(A, knn, p, lazy, next_bins);
    A->mark_integrated();
}

}  // namespace gs
