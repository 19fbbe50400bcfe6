// Boids step: the per-frame sequence of plugins/BoidsGroup.cpp:606-637
// (wrapWorld, updateKNNDatabase, findKNearestNeighbors, steerForSeparation /
// Alignment / Cohesion, updateGroup) as
//
//   bin build (gs_bins.hip) with the wrap fused into the key pass; it leaves
//   the group's state in cell order and the position snapshot in spos;
//   exact seeded KNN of the group with itself (gs_knn.hip);
//   k_flock_update: one lane per storage slot -> the three flocking terms
//   -> integrate -> new positions and forward vectors into the alternate
//   buffers, the agent's other fields in place.
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

template <int KMAX, int MODE>
__global__ __launch_bounds__(kBlock) void k_flock_update(uint32_t n, Agents A, StepUniforms U, float* __restrict__ npx,
                                                         float* __restrict__ npy, float* __restrict__ npz,
                                                         float* __restrict__ nfx, float* __restrict__ nfy,
                                                         float* __restrict__ nfz, const float4* __restrict__ spos,
                                                         const uint32_t* __restrict__ knn, uint32_t k, FlockTerm sep,
                                                         FlockTerm ali, FlockTerm coh, float dt) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= n) return;
    // Step-t snapshot: positions in spos, forward vectors in the current
    // buffers; new positions / forward vectors go to the alternate buffers.
    const float4 q = spos[s];
    const f3 pos = mk3(q.x, q.y, q.z);
    const f3 fwd = ld3(A.fx, A.fy, A.fz, s);
    uint32_t bits = U.steer_zero ? 0u : A.bits[s];
    f3 bp[KMAX], bf[KMAX];
    const int m = gather_neighbours<KMAX>(knn, k, n, s, n, spos, nullptr, nullptr, nullptr, A.fx, A.fy, A.fz,
                                          ali.on && !(bits & ali.mask), bp, bf);
    f3 steering = U.steer_zero ? mk3(0.f, 0.f, 0.f) : ld3(A.gx, A.gy, A.gz, s);
    const float radius = (U.has & 1u) ? U.v[0] : A.radius[s];
    const float max_speed = (U.has & 2u) ? U.v[1] : A.max_speed[s];
    const float max_force = (U.has & 4u) ? U.v[2] : A.max_force[s];
    const float mass = (U.has & 8u) ? U.v[3] : A.mass[s];
    flock_core<MODE, KMAX>(pos, fwd, radius, m, bp, bf, sep, ali, coh, steering, bits);
    // Integrate (clears steering and bits as updateGroup does).
    f3 p = pos, f = fwd;
    f3 side = ld3(A.sx, A.sy, A.sz, s), up = ld3(A.ux, A.uy, A.uz, s);
    f3 sacc = ld3(A.ax, A.ay, A.az, s);
    float speed = A.speed[s];
    integrate<MODE>(steering, p, f, speed, sacc, side, up, max_force, max_speed, mass, dt);
    st3(npx, npy, npz, s, p);
    st3(nfx, nfy, nfz, s, f);
    st3(A.sx, A.sy, A.sz, s, side);
    st3(A.ux, A.uy, A.uz, s, up);
    if (MODE == GS_MODE_OPENSTEER) st3(A.ax, A.ay, A.az, s, sacc);
    A.speed[s] = speed;
    if (!U.steer_zero) {
        st3(A.gx, A.gy, A.gz, s, mk3(0.f, 0.f, 0.f));
        A.bits[s] = 0u;
    }
}

static FlockTerm fterm(const gs_flock_term& t) { return FlockTerm{t.min_d, t.max_d, t.cos_max, t.weight, t.mask, 1}; }

template <int KMAX>
static void run_flock_update(gs_group_s* A, gs_knn_s* knn, const gs_boids_params* p) {
    const uint32_t n = A->n;
    hipStream_t st = A->ctx->stream;
    const Agents a = A->dev_agents();
    StepUniforms U{};
    U.steer_zero = (A->inv & kFSteer) == kFSteer && A->bits_zero && __builtin_bit_cast(uint32_t, A->inv_val[16]) == 0u &&
                   __builtin_bit_cast(uint32_t, A->inv_val[17]) == 0u && __builtin_bit_cast(uint32_t, A->inv_val[18]) == 0u;
    for (int c = 0; c < 4; ++c)
        if (A->inv & (1u << (19 + c))) {
            U.has |= 1u << c;
            U.v[c] = A->inv_val[19 + c];
        }
#define GS_FU(MODE)                                                                                                  \
    k_flock_update<KMAX, MODE><<<blocks_for(n), kBlock, 0, st>>>(                                                    \
        n, a, U, A->Falt(0), A->Falt(1), A->Falt(2), A->Falt(3), A->Falt(4), A->Falt(5), A->spos.p, knn->idx(), knn->k, \
        fterm(p->separation), fterm(p->alignment), fterm(p->cohesion), p->dt)
    if (p->mode == GS_MODE_GPUREF) GS_FU(GS_MODE_GPUREF);
    else GS_FU(GS_MODE_OPENSTEER);
#undef GS_FU
    GS_LAUNCHED("k_flock_update");
}

void launch_flock_update(gs_group_s* A, gs_knn_s* knn, const gs_boids_params* p) {
    if (A->n == 0) return;
    if (knn->k <= 8) run_flock_update<8>(A, knn, p);
    else if (knn->k <= 16) run_flock_update<16>(A, knn, p);
    else run_flock_update<32>(A, knn, p);
    A->mark_integrated();
}

}  // namespace gs
