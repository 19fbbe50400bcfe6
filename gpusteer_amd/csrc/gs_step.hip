// Fused Boids step: the per-frame sequence of plugins/BoidsGroup.cpp:606-637
// (wrapWorld, updateKNNDatabase, findKNearestNeighbors, steerForSeparation /
// Alignment / Cohesion, updateGroup) in two launches after the bin build:
//
//   bin build (gs_bins.hip) with the wrap fused into the key pass and a
//   snapshot of positions and forward vectors gathered in cell order, plus the
//   inverse permutation;
//   k_boids: one lane per agent in cell order -> exact KNN against the
//   snapshot -> the three flocking terms from the snapshot -> integrate ->
//   write the agent's new state and its KNN column.
//
// Every read of another agent goes to the snapshot, every write goes to the
// agent's own slot of the live arrays, so the update is Jacobi (all forces
// from the step-t state) exactly like the separate launches, and the results
// are identical to gs_wrap_world + gs_update_knn_database + gs_find_knn +
// gs_steer_flocking + gs_update_group.
#include "gs_ops.hpp"

#include "gs_knn_core.hpp"

namespace gs {

template <int KMAX, int MODE>
__global__ __launch_bounds__(kBlock) void k_boids(uint32_t n, Agents A, const float* __restrict__ spx,
                                                  const float* __restrict__ spy, const float* __restrict__ spz,
                                                  const float* __restrict__ sfx, const float* __restrict__ sfy,
                                                  const float* __restrict__ sfz, const uint32_t* __restrict__ perm,
                                                  const uint32_t* __restrict__ inv,
                                                  const uint32_t* __restrict__ start, GridDev g, uint32_t k, int R0,
                                                  const uint32_t* __restrict__ seed, uint32_t* __restrict__ out_idx,
                                                  float* __restrict__ out_dist, FlockTerm sep, FlockTerm ali,
                                                  FlockTerm coh, float dt) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    const f3 pos = mk3(spx[s], spy[s], spz[s]);
    const f3 fwd = mk3(sfx[s], sfy[s], sfz[s]);
    const uint32_t q = perm[s];

    // Exact KNN of q within its own group (self excluded by sorted slot).
    TopK<KMAX> h;
    h.init(k);
    const float thr = seed ? seed_bound(seed, k, n, q, n, pos.x, pos.y, pos.z, inv, spx, spy, spz) : INFINITY;
    binned_search<KMAX>(h, thr, pos.x, pos.y, pos.z, s, g, R0, spx, spy, spz, perm, start);
    h.store(k, n, q, out_idx, out_dist);

    // Flocking from the snapshot, slot order.
    uint32_t bits = A.bits[q];
    f3 steering = ld3(A.gx, A.gy, A.gz, q);
    // The list's slots KMAX-k .. KMAX-1 are the k results in order; the
    // unrolled walk keeps every list access a register access.
    flock_core<MODE, KMAX>(pos, fwd, A.radius[q], k, sep, ali, coh, steering, bits,
                           [&](int j, bool want_fwd, f3& bp, f3& bf) {
                               if (j < KMAX - (int)k) return (int)kNbrSkip;
                               const uint32_t b = h.i[j];
                               if (b == UINT32_MAX) return (int)kNbrStop;
                               const uint32_t t = inv[b];
                               bp = mk3(spx[t], spy[t], spz[t]);
                               if (want_fwd) bf = mk3(sfx[t], sfy[t], sfz[t]);
                               return (int)kNbrValid;
                           });

    // Integrate (clears steering and bits as updateGroup does).
    f3 p = pos, f = fwd;
    f3 side = ld3(A.sx, A.sy, A.sz, q), up = ld3(A.ux, A.uy, A.uz, q);
    f3 sacc = ld3(A.ax, A.ay, A.az, q);
    float speed = A.speed[q];
    integrate<MODE>(steering, p, f, speed, sacc, side, up, A.max_force[q], A.max_speed[q], A.mass[q], dt);
    st3(A.px, A.py, A.pz, q, p);
    st3(A.fx, A.fy, A.fz, q, f);
    st3(A.sx, A.sy, A.sz, q, side);
    st3(A.ux, A.uy, A.uz, q, up);
    if (MODE == GS_MODE_OPENSTEER) st3(A.ax, A.ay, A.az, q, sacc);
    A.speed[q] = speed;
    st3(A.gx, A.gy, A.gz, q, mk3(0.f, 0.f, 0.f));
    A.bits[q] = 0u;
}

static FlockTerm fterm(const gs_flock_term& t) { return FlockTerm{t.min_d, t.max_d, t.cos_max, t.weight, t.mask, 1}; }

template <int KMAX>
static void run_boids(gs_group_s* A, gs_knn_s* knn, const gs_boids_params* p, const uint32_t* seed) {
    const uint32_t n = A->n;
    hipStream_t st = A->ctx->stream;
    const Agents a = A->dev_agents();
#define GS_BOIDS(MODE)                                                                                               \
    k_boids<KMAX, MODE><<<blocks_for(n), kBlock, 0, st>>>(                                                           \
        n, a, A->spx.p, A->spy.p, A->spz.p, A->sfx.p, A->sfy.p, A->sfz.p, A->perm.p, A->inv.p, A->start.p,          \
        A->bin_grid, knn->k, (int)p->search_radius, seed, knn->idx.p, knn->dist.p, fterm(p->separation),            \
        fterm(p->alignment), fterm(p->cohesion), p->dt)
    if (p->mode == GS_MODE_GPUREF) GS_BOIDS(GS_MODE_GPUREF);
    else GS_BOIDS(GS_MODE_OPENSTEER);
#undef GS_BOIDS
    GS_LAUNCHED("k_boids");
}

void launch_boids_fused(gs_group_s* A, gs_knn_s* knn, const gs_boids_params* p, bool use_seed) {
    if (A->n == 0) return;
    const uint32_t* seed = use_seed ? knn->idx.p : nullptr;
    if (knn->k <= 8) run_boids<8>(A, knn, p, seed);
    else if (knn->k <= 16) run_boids<16>(A, knn, p, seed);
    else run_boids<32>(A, knn, p, seed);
}

}  // namespace gs
