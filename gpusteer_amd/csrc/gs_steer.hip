// Steering behaviours, integration and world wrap: one lane per agent, SoA
// fp32 state read and written coalesced; neighbour data gathered through the
// KNN indices (spatially close agents, served from L2 / Infinity Cache).
//
// Every behaviour follows the reference kernels' accumulate convention
// (e.g. SteerForSeparationCUDAKernel.cu:74-75,124-134): skip the agent when
// bits & mask, add f*w into the steering accumulator as (f*w) + steering, and
// set the behaviour's bit when f*w is non-zero.
#include "gs_ops.hpp"

namespace gs {

template <int KMAX, int MODE>
__global__ __launch_bounds__(kBlock) void k_flock(uint32_t nA, Agents A, uint32_t nB, const float* __restrict__ bpx,
                                                  const float* __restrict__ bpy, const float* __restrict__ bpz,
                                                  const float* __restrict__ bfx, const float* __restrict__ bfy,
                                                  const float* __restrict__ bfz, const uint32_t* __restrict__ knn,
                                                  uint32_t k, FlockTerm sep, FlockTerm ali, FlockTerm coh) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nA) return;
    uint32_t bits = A.bits[i];
    const bool any = (sep.on && !(bits & sep.mask)) || (ali.on && !(bits & ali.mask)) ||
                     (coh.on && !(bits & coh.mask));
    if (!any) return;
    f3 bp[KMAX], bf[KMAX];
    const int m = gather_neighbours<KMAX>(knn, k, nA, i, nB, nullptr, bpx, bpy, bpz, bfx, bfy, bfz,
                                          ali.on && !(bits & ali.mask), bp, bf);
    f3 steering = ld3(A.gx, A.gy, A.gz, i);
    flock_core<MODE, KMAX>(ld3(A.px, A.py, A.pz, i), ld3(A.fx, A.fy, A.fz, i), A.radius[i], m, bp, bf, sep, ali,
                           coh, steering, bits);
    st3(A.gx, A.gy, A.gz, i, steering);
    A.bits[i] = bits;
}

// Seek: OPENSTEER = src/SteerLibrary.cpp:36-42 (target - position - velocity);
// GPUREF = SteerForSeekCUDAKernel.cu:52-61 (normalize(target - position - forward)).
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_seek(uint32_t n, Agents A, float tx, float ty, float tz, float w,
                                                 uint32_t mask) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t bits = A.bits[i];
    if (bits & mask) return;
    const f3 desired = mk3(tx, ty, tz) - ld3(A.px, A.py, A.pz, i);
    const f3 fwd = ld3(A.fx, A.fy, A.fz, i);
    const f3 f = (MODE == GS_MODE_GPUREF) ? normalize3(desired - fwd) : desired - fwd * A.speed[i];
    f3 steering = ld3(A.gx, A.gy, A.gz, i);
    apply(steering, bits, f, w, GS_BIT_SEEK);
    st3(A.gx, A.gy, A.gz, i, steering);
    A.bits[i] = bits;
}

// Obstacle avoidance over the k nearest obstacles: SteerToAvoidObstaclesKernel
// with findNextIntersectionWithSphere (CUDA/AvoidObstaclesCUDAKernel.cu:47-99,136-253).
__global__ __launch_bounds__(kBlock) void k_avoid_obstacles(uint32_t n, Agents A, uint32_t nO,
                                                            const float* __restrict__ ox, const float* __restrict__ oy,
                                                            const float* __restrict__ oz, const float* __restrict__ orad,
                                                            const uint32_t* __restrict__ knn, uint32_t k,
                                                            float min_time, float w, uint32_t mask) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t bits = A.bits[i];
    if (bits & mask) return;
    f3 steering = ld3(A.gx, A.gy, A.gz, i);
    avoid_obstacles_one(ld3(A.px, A.py, A.pz, i), ld3(A.fx, A.fy, A.fz, i), ld3(A.sx, A.sy, A.sz, i),
                        ld3(A.ux, A.uy, A.uz, i), A.radius[i], A.speed[i], knn + i, n, k, nO, ox, oy, oz, orad,
                        min_time, w, steering, bits);
    st3(A.gx, A.gy, A.gz, i, steering);
    A.bits[i] = bits;
}

// Wall avoidance over the k nearest wall segments: SteerToAvoidWallsCUDAKernel
// (CUDA/AvoidWallsCUDAKernel.cu:67-245) — three feelers, touching test, the
// anti-penetration move of touching agents (:215-223) and its bit rules.
__global__ __launch_bounds__(kBlock) void k_avoid_walls(uint32_t n, Agents A, uint32_t nW, gs_walls W,
                                                        const uint32_t* __restrict__ knn, uint32_t k, float min_time,
                                                        float w, uint32_t mask) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t bits = A.bits[i];
    if (bits & mask) return;
    f3 pos = ld3(A.px, A.py, A.pz, i), dir = ld3(A.fx, A.fy, A.fz, i);
    f3 steering = ld3(A.gx, A.gy, A.gz, i);
    avoid_walls_one(pos, dir, ld3(A.sx, A.sy, A.sz, i), A.radius[i], A.speed[i], knn + i, n, k, nW, W, min_time, w,
                    steering, bits);
    st3(A.gx, A.gy, A.gz, i, steering);
    st3(A.px, A.py, A.pz, i, pos);
    st3(A.fx, A.fy, A.fz, i, dir);
    A.bits[i] = bits;
}

template <int MODE>
__global__ __launch_bounds__(kBlock) void k_update(uint32_t n, Agents A, float dt) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    f3 pos = ld3(A.px, A.py, A.pz, i), fwd = ld3(A.fx, A.fy, A.fz, i);
    f3 side = ld3(A.sx, A.sy, A.sz, i), up = ld3(A.ux, A.uy, A.uz, i);
    f3 sacc = ld3(A.ax, A.ay, A.az, i);
    float speed = A.speed[i];
    integrate<MODE>(ld3(A.gx, A.gy, A.gz, i), pos, fwd, speed, sacc, side, up, A.max_force[i], A.max_speed[i],
                    A.mass[i], dt);
    st3(A.px, A.py, A.pz, i, pos);
    st3(A.fx, A.fy, A.fz, i, fwd);
    st3(A.sx, A.sy, A.sz, i, side);
    st3(A.ux, A.uy, A.uz, i, up);
    if (MODE == GS_MODE_OPENSTEER) st3(A.ax, A.ay, A.az, i, sacc);
    A.speed[i] = speed;
    st3(A.gx, A.gy, A.gz, i, mk3(0.f, 0.f, 0.f));
    A.bits[i] = 0u;
}

// World wrap: GS_WRAP_BOX = WrapWorldKernel face teleport
// (CUDA/WrapWorldCUDAKernel.cu:25-35, all three axes); GS_WRAP_SPHERE =
// sphericalWrapAround of plugins/Boids.cpp:200-206.
__global__ __launch_bounds__(kBlock) void k_wrap(uint32_t n, float* __restrict__ px, float* __restrict__ py,
                                                 float* __restrict__ pz, int shape, float cx, float cy, float cz,
                                                 float s0, float s1, float s2) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    f3 p = ld3(px, py, pz, i);
    if (shape == GS_WRAP_BOX) {
        const float hx = s0 * 0.5f, hy = s1 * 0.5f, hz = s2 * 0.5f;
        if (p.x < cx - hx) p.x = cx + hx;
        else if (p.x > cx + hx) p.x = cx - hx;
        if (p.y < cy - hy) p.y = cy + hy;
        else if (p.y > cy + hy) p.y = cy - hy;
        if (p.z < cz - hz) p.z = cz + hz;
        else if (p.z > cz + hz) p.z = cz - hz;
    } else {
        const f3 off = p - mk3(cx, cy, cz);
        const float r = len3(off);
        if (r > s0) p = p + ((off / r) * s0) * -2.f;
    }
    st3(px, py, pz, i, p);
}

__global__ __launch_bounds__(kBlock) void k_clear_bits(uint32_t n, uint32_t* __restrict__ bits) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) bits[i] = 0u;
}

__global__ __launch_bounds__(kBlock) void k_check_finite(uint32_t n, Agents A, uint32_t* __restrict__ first_bad) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const float* const* f = (const float* const*)&A;
    bool ok = true;
#pragma unroll
    for (int j = 0; j < 23; ++j) ok = ok && isfinite(f[j][i]);
    if (!ok) atomicMin(first_bad, i);
}

static FlockTerm term(const gs_flock_term* t) {
    FlockTerm r{0.f, 0.f, 0.f, 0.f, 0u, 0};
    if (t) r = FlockTerm{t->min_d, t->max_d, t->cos_max, t->weight, t->mask, 1};
    return r;
}

template <int KMAX>
static void run_flock(gs_group_s* A, gs_knn_s* knn, gs_group_s* B, const gs_flock_term* sep,
                      const gs_flock_term* ali, const gs_flock_term* coh, int mode) {
    hipStream_t st = A->ctx->stream;
    const Agents a = A->dev_agents();
#define GS_FLOCK(MODE)                                                                                              \
    k_flock<KMAX, MODE><<<blocks_for(A->n), kBlock, 0, st>>>(A->n, a, B->n, B->F(0), B->F(1), B->F(2), B->F(3),      \
                                                             B->F(4), B->F(5), knn->idx(), knn->k, term(sep),        \
                                                             term(ali), term(coh))
    if (mode == GS_MODE_GPUREF) GS_FLOCK(GS_MODE_GPUREF);
    else GS_FLOCK(GS_MODE_OPENSTEER);
#undef GS_FLOCK
    GS_LAUNCHED("k_flock");
}

void launch_flock(gs_group_s* A, gs_knn_s* knn, gs_group_s* B, const gs_flock_term* sep, const gs_flock_term* ali,
                  const gs_flock_term* coh, int mode) {
    if (A->n == 0) return;
    if (knn->k <= 8) run_flock<8>(A, knn, B, sep, ali, coh, mode);
    else if (knn->k <= 16) run_flock<16>(A, knn, B, sep, ali, coh, mode);
    else run_flock<32>(A, knn, B, sep, ali, coh, mode);
    A->mark_written(kFSteer, true);
}

void launch_seek(gs_group_s* A, const float* t, float w, uint32_t mask, int mode) {
    if (A->n == 0) return;
    hipStream_t st = A->ctx->stream;
    if (mode == GS_MODE_GPUREF)
        k_seek<GS_MODE_GPUREF><<<blocks_for(A->n), kBlock, 0, st>>>(A->n, A->dev_agents(), t[0], t[1], t[2], w, mask);
    else
        k_seek<GS_MODE_OPENSTEER><<<blocks_for(A->n), kBlock, 0, st>>>(A->n, A->dev_agents(), t[0], t[1], t[2], w,
                                                                       mask);
    GS_LAUNCHED("k_seek");
    A->mark_written(kFSteer, true);
}

void launch_avoid_obstacles(gs_group_s* A, gs_group_s* O, gs_knn_s* knn, float min_time, float w, uint32_t mask) {
    if (A->n == 0) return;
    k_avoid_obstacles<<<blocks_for(A->n), kBlock, 0, A->ctx->stream>>>(A->n, A->dev_agents(), O->n, O->F(0),
                                                                        O->F(1), O->F(2), O->F(19), knn->idx(),
                                                                        knn->k, min_time, w, mask);
    GS_LAUNCHED("k_avoid_obstacles");
    A->mark_written(kFSteer, true);
}

void launch_avoid_walls(gs_group_s* A, gs_knn_s* knn, gs_group_s* W, float min_time, float w, uint32_t mask) {
    if (A->n == 0) return;
    const int f0 = kWallField;
    gs_walls wd{W->F(f0), W->F(f0 + 1), W->F(f0 + 2), W->F(f0 + 3), W->F(f0 + 4),
                W->F(f0 + 5), W->F(f0 + 6), W->F(f0 + 7), W->F(f0 + 8)};
    k_avoid_walls<<<blocks_for(A->n), kBlock, 0, A->ctx->stream>>>(A->n, A->dev_agents(), W->n, wd, knn->idx(),
                                                                    knn->k, min_time, w, mask);
    GS_LAUNCHED("k_avoid_walls");
    // touching agents are moved and turned (positions, forward vectors)
    A->mark_written(kFSteer | 0x3Fu, true);
}

void launch_update(gs_group_s* A, float dt, int mode) {
    if (A->n == 0) return;
    hipStream_t st = A->ctx->stream;
    if (mode == GS_MODE_GPUREF)
        k_update<GS_MODE_GPUREF><<<blocks_for(A->n), kBlock, 0, st>>>(A->n, A->dev_agents(), dt);
    else
        k_update<GS_MODE_OPENSTEER><<<blocks_for(A->n), kBlock, 0, st>>>(A->n, A->dev_agents(), dt);
    GS_LAUNCHED("k_update");
    A->mark_integrated();
}

void launch_wrap(gs_group_s* A, int shape, const float* c, const float* s) {
    if (A->n == 0) return;
    k_wrap<<<blocks_for(A->n), kBlock, 0, A->ctx->stream>>>(A->n, A->F(0), A->F(1), A->F(2), shape, c[0], c[1], c[2],
                                                            s[0], s[1], s[2]);
    GS_LAUNCHED("k_wrap");
    A->mark_written(kFPos, false);
}

void launch_clear_bits(gs_group_s* A) {
    if (A->n == 0) return;
    k_clear_bits<<<blocks_for(A->n), kBlock, 0, A->ctx->stream>>>(A->n, A->bits());
    GS_LAUNCHED("k_clear_bits");
    A->bits_zero = true;
}

void launch_check_finite(gs_group_s* A, uint32_t* first_bad_host) {
    *first_bad_host = UINT32_MAX;
    if (A->n == 0) return;
    hipStream_t st = A->ctx->stream;
    A->flag.reserve(1);
    GS_HIP(hipMemcpyAsync(A->flag.p, first_bad_host, sizeof(uint32_t), hipMemcpyHostToDevice, st));
    k_check_finite<<<blocks_for(A->n), kBlock, 0, st>>>(A->n, A->dev_agents(), A->flag.p);
    GS_LAUNCHED("k_check_finite");
    GS_HIP(hipMemcpyAsync(first_bad_host, A->flag.p, sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    GS_HIP(hipStreamSynchronize(st));
}

}  // namespace gs
