// Fused Boids step: the per-frame sequence of plugins/BoidsGroup.cpp:606-637
// (wrapWorld, updateKNNDatabase, findKNearestNeighbors, steerForSeparation /
// Alignment / Cohesion, updateGroup) in two launches after the bin build:
//
//   bin build (gs_bins.hip) with the wrap fused into the key pass; it leaves
//   the group's state in cell order;
//   k_boids: one lane per storage slot -> exact KNN against the state ->
//   the three flocking terms -> integrate -> new positions and forward
//   vectors into the alternate buffers, the agent's other fields in place.
//
// Other agents are only ever read through positions and forward vectors of
// the current buffers, which no lane writes, so the update is Jacobi (all
// forces from the step-t state) exactly like the separate launches, and the
// results are identical to gs_wrap_world + gs_update_knn_database +
// gs_find_knn + gs_steer_flocking + gs_update_group.
#include "gs_ops.hpp"

#include "gs_knn_core.hpp"

namespace gs {

template <int KMAX, int MODE>
__global__ __launch_bounds__(kBlock) void k_boids(uint32_t n, Agents A, float* __restrict__ npx,
                                                  float* __restrict__ npy, float* __restrict__ npz,
                                                  float* __restrict__ nfx, float* __restrict__ nfy,
                                                  float* __restrict__ nfz, const float4* __restrict__ spos,
                                                  const uint32_t* __restrict__ start, GridDev g, uint32_t k, int R0,
                                                  const uint32_t* __restrict__ seed,
                                                  const uint32_t* __restrict__ row_map,
                                                  const uint32_t* __restrict__ val_map,
                                                  uint32_t* __restrict__ out_idx, float* __restrict__ out_dist,
                                                  FlockTerm sep, FlockTerm ali, FlockTerm coh, float dt) {
    __shared__ uint2 lbuf[kKnnBuf * kBlock];
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    // A's current buffers are the step-t snapshot (just built in cell order);
    // new positions and forward vectors go to the alternate buffers.
    const float4 q = spos[s];
    const f3 pos = mk3(q.x, q.y, q.z);
    const f3 fwd = ld3(A.fx, A.fy, A.fz, s);

    // Exact KNN within the group (self excluded by slot).
    TopK<KMAX> h;
    h.init(k);
    const float thr =
        seed ? seed_bound(seed, k, n, row_map ? row_map[s] : s, n, val_map, pos.x, pos.y, pos.z, spos) : INFINITY;
    knn_search<KMAX>(h, thr, pos.x, pos.y, pos.z, s, g, R0, spos, start, lbuf + threadIdx.x);
    h.store(k, n, s, out_idx, out_dist);

    // Flocking from the snapshot in slot order.  The list's entries
    // KMAX-k .. KMAX-1 are the k results; the unrolled walk keeps every list
    // access a register access.
    uint32_t bits = A.bits[s];
    f3 steering = ld3(A.gx, A.gy, A.gz, s);
    flock_core<MODE, KMAX>(pos, fwd, A.radius[s], k, sep, ali, coh, steering, bits,
                           [&](int j, bool want_fwd, f3& bp, f3& bf) {
                               if (j < KMAX - (int)k) return (int)kNbrSkip;
                               const uint32_t t = h.t[j];
                               if (t == UINT32_MAX) return (int)kNbrStop;
                               const float4 b = spos[t];
                               bp = mk3(b.x, b.y, b.z);
                               if (want_fwd) bf = ld3(A.fx, A.fy, A.fz, t);
                               return (int)kNbrValid;
                           });

    // Integrate (clears steering and bits as updateGroup does).  Fields only
    // this lane reads are updated in place.
    f3 p = pos, f = fwd;
    f3 side = ld3(A.sx, A.sy, A.sz, s), up = ld3(A.ux, A.uy, A.uz, s);
    f3 sacc = ld3(A.ax, A.ay, A.az, s);
    float speed = A.speed[s];
    integrate<MODE>(steering, p, f, speed, sacc, side, up, A.max_force[s], A.max_speed[s], A.mass[s], dt);
    st3(npx, npy, npz, s, p);
    st3(nfx, nfy, nfz, s, f);
    st3(A.sx, A.sy, A.sz, s, side);
    st3(A.ux, A.uy, A.uz, s, up);
    if (MODE == GS_MODE_OPENSTEER) st3(A.ax, A.ay, A.az, s, sacc);
    A.speed[s] = speed;
    st3(A.gx, A.gy, A.gz, s, mk3(0.f, 0.f, 0.f));
    A.bits[s] = 0u;
}

static FlockTerm fterm(const gs_flock_term& t) { return FlockTerm{t.min_d, t.max_d, t.cos_max, t.weight, t.mask, 1}; }

template <int KMAX>
static void run_boids(gs_group_s* A, gs_knn_s* knn, const gs_boids_params* p, const uint32_t* seed,
                      const uint32_t* row_map, const uint32_t* val_map) {
    const uint32_t n = A->n;
    hipStream_t st = A->ctx->stream;
    const Agents a = A->dev_agents();
#define GS_BOIDS(MODE)                                                                                               \
    k_boids<KMAX, MODE><<<blocks_for(n), kBlock, 0, st>>>(                                                           \
        n, a, A->Falt(0), A->Falt(1), A->Falt(2), A->Falt(3), A->Falt(4), A->Falt(5), A->spos.p, A->start.p,        \
        A->bin_grid, knn->k, (int)p->search_radius, seed, row_map, val_map, knn->idx_alt(), knn->dist_alt(),        \
        fterm(p->separation), fterm(p->alignment), fterm(p->cohesion), p->dt)
    if (p->mode == GS_MODE_GPUREF) GS_BOIDS(GS_MODE_GPUREF);
    else GS_BOIDS(GS_MODE_OPENSTEER);
#undef GS_BOIDS
    GS_LAUNCHED("k_boids");
}

void launch_boids_fused(gs_group_s* A, gs_knn_s* knn, const gs_boids_params* p, const uint32_t* seed,
                        const uint32_t* row_map, const uint32_t* val_map) {
    if (A->n == 0) return;
    if (knn->k <= 8) run_boids<8>(A, knn, p, seed, row_map, val_map);
    else if (knn->k <= 16) run_boids<16>(A, knn, p, seed, row_map, val_map);
    else run_boids<32>(A, knn, p, seed, row_map, val_map);
}

}  // namespace gs
