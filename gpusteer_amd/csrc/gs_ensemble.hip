// Batched ensemble: many independent scenarios (Boids flocks and choke-point
// crowds of uneven agent counts) advanced together, every phase of a step one
// launch over all scenarios, the step sequence captured into a hipGraph.
//
// Layout.  Every group of every scenario is a population with its own slot
// range [a0, a0 + n) of the shared SoA arrays, padded to a multiple of kBlock
// so each block belongs to one population (its parameters are block-uniform
// scalar loads), and its own cell range [c0, c0 + cells] of one shared bin
// lattice table (the extra cell holds the padding slots, is never searched).
// One counting sort over all slots therefore keeps every population inside
// its own range and in cell order, exactly as a single-instance group build
// does, and the search of a population only ever visits its own cells.
//
// Step (same results, bit for bit, as the single-instance calls of each
// scenario, which the tests check):
//   e_bin_key       wrap (Boids) + cell key + histogram       all slots
//   scan, scatter, reorder (gs_bins.hip)                       all slots
//   e_knn           exact seeded KNN self (+ other group, walls and
//                   obstacles by brute force for crowds)       all agents
//   e_boids         flocking + integrate (Jacobi)              Boids blocks
//   e_choke_a(g)    clear bits, wall and obstacle avoidance    crowd group g
//   e_choke_b(g)    separation (own, other group), seek        crowd group g
//                   (a0, b0, a1, b1: the order of the per-group call
//                   sequence, plugins/ChokePointGroup.cpp:683-728)
//   e_choke_update  integrate                                  crowd blocks
// New positions and forward vectors always go to the alternate buffers and
// fields 0..5 flip after the step, so every buffer returns to its place every
// two steps: a captured pair of steps replays indefinitely.
#include <algorithm>
#include <cmath>
#include <vector>

#include "gs_ensemble.hpp"
#include "gs_knn_core.hpp"

namespace gs {

struct EnsDev {
    const EnsPop* pops;
    const EnsScen* scens;
    const uint32_t* block_pop;
};

__global__ __launch_bounds__(kBlock) void e_bin_key(EnsDev E, float* __restrict__ px, float* __restrict__ py,
                                                    float* __restrict__ pz, uint32_t* __restrict__ key,
                                                    uint32_t* __restrict__ rank, uint32_t* __restrict__ count,
                                                    uint32_t* __restrict__ box_acc, uint32_t* __restrict__ box_clear,
                                                    uint32_t nbox) {
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t s = blk * blockDim.x + threadIdx.x;
    const EnsPop& P = E.pops[E.block_pop[blk]];
    const bool agent = s < P.a0 + P.n;
    uint32_t c = P.c0 + P.cells;  // padding slots: the population's extra cell
    float x = 0.f, y = 0.f, z = 0.f;
    if (agent) {
        x = px[s], y = py[s], z = pz[s];
        const EnsScen& S = E.scens[P.scen];
        if (S.wrap_shape == GS_WRAP_BOX) {
            const float hx = S.ws[0] * 0.5f, hy = S.ws[1] * 0.5f, hz = S.ws[2] * 0.5f;
            const float x0 = x, y0 = y, z0 = z;
            if (x < S.wc[0] - hx) x = S.wc[0] + hx;
            else if (x > S.wc[0] + hx) x = S.wc[0] - hx;
            if (y < S.wc[1] - hy) y = S.wc[1] + hy;
            else if (y > S.wc[1] + hy) y = S.wc[1] - hy;
            if (z < S.wc[2] - hz) z = S.wc[2] + hz;
            else if (z > S.wc[2] + hz) z = S.wc[2] - hz;
            if (x != x0) px[s] = x;
            if (y != y0) py[s] = y;
            if (z != z0) pz[s] = z;
        } else if (S.wrap_shape == GS_WRAP_SPHERE) {
            const f3 off = mk3(x - S.wc[0], y - S.wc[1], z - S.wc[2]);
            const float r = len3(off);
            if (r > S.ws[0]) {
                const f3 p = mk3(x, y, z) + ((off / r) * S.ws[0]) * -2.f;
                x = p.x;
                y = p.y;
                z = p.z;
                px[s] = x;
                py[s] = y;
                pz[s] = z;
            }
        }
        const GridDev& g = P.g;
        c = P.c0 + cell_index(g, axis_cell(x, g.ox, g.sx, g.dx), axis_cell(y, g.oy, g.sy, g.dy),
                              axis_cell(z, g.oz, g.sz, g.dz));
    }
    key[s] = c;
    count_and_rank(true, c, s, count, rank);
    // Crowd populations' bounding boxes (block-uniform branches: a block
    // belongs to one population); the other buffer is cleared for the next
    // build.
    if (blockIdx.x == 0)
        for (uint32_t j = threadIdx.x; j < nbox; j += kBlock) box_clear[j] = j % 6 < 3 ? kBoxEmptyMin : kBoxEmptyMax;
    if (P.kind == GS_SCENARIO_CHOKE) box_accumulate(agent, x, y, z, box_acc + 6 * (size_t)E.block_pop[blk]);
}

// Brute-force top-k over a scenario's few walls (midpoints) or obstacles
// (centres), key (d2, local index): the same exact result the binned search
// of a wall / obstacle group returns.
template <int KMAX>
__device__ __forceinline__ void brute_small(TopK<KMAX>& h, float qx, float qy, float qz, uint32_t n,
                                            const float* __restrict__ x, const float* __restrict__ y,
                                            const float* __restrict__ z) {
    for (uint32_t j = 0; j < n; ++j) {
        const uint64_t kk = make_key(sqdist(qx, qy, qz, x[j], y[j], z[j]), j);
        if (h.accepts(kk)) h.push(kk, j);
    }
}

template <int KMAX>
__device__ __forceinline__ void store_idx(const TopK<KMAX>& h, uint32_t k, uint32_t N, uint32_t s,
                                          uint32_t* __restrict__ out, const uint32_t* __restrict__ map = nullptr) {
    const uint32_t row = map ? map[s] : s;
#pragma unroll
    for (int j = 0; j < KMAX; ++j)
        if (j >= KMAX - (int)k) {
            const uint32_t v = h.t[j];
            out[(size_t)(j - (KMAX - (int)k)) * N + row] = (map && v != UINT32_MAX) ? map[v] : v;
        }
}
// Wall / obstacle lists hold indices within the scenario: only the row moves.
template <int KMAX>
__device__ __forceinline__ void store_local(const TopK<KMAX>& h, uint32_t k, uint32_t N, uint32_t s,
                                            uint32_t* __restrict__ out, const uint32_t* __restrict__ map) {
    const uint32_t row = map ? map[s] : s;
#pragma unroll
    for (int j = 0; j < KMAX; ++j)
        if (j >= KMAX - (int)k) out[(size_t)(j - (KMAX - (int)k)) * N + row] = h.t[j];
}

struct EnsKnnOut {
    uint32_t* self;    // [k][N] global slots
    uint32_t* other;   // [k][N] global slots (crowds)
    uint32_t* walls;   // [kw][N] wall index within the scenario
    uint32_t* obs;     // [ko][N] obstacle index within the scenario
    const uint32_t* seed_self;   // previous results at the previous order, or NULL
    const uint32_t* seed_other;
    // NULL: rows and neighbour values are current (sorted) slots.  Otherwise
    // both are translated through it (sorted slot -> fixed index), for state
    // kept in a fixed order with a sorted snapshot for the search.
    const uint32_t* out_map;
    // The populations' bounding boxes of this build (6 words each).
    const uint32_t* box;
};

// KMAX: 7 for k = 7 (the lists of exactly 7 entries of the single-instance
// search), else 8.  Two passes: PART 0 searches every agent's own group (all
// blocks; the search alone, at the single-instance search's register budget
// of five wavefronts per SIMD), PART 1 the crowd agents' other group, walls
// and obstacles (crowd blocks only, from the `blocks` list).
template <int KMAX, int PART>
__global__ __launch_bounds__(kBlock, PART == 0 ? 5 : 1) void e_knn(
    EnsDev E, const uint32_t* __restrict__ blocks, uint32_t N, uint32_t k, const float4* __restrict__ spos,
    const uint32_t* __restrict__ start, const uint32_t* __restrict__ new_to_old,
    const uint32_t* __restrict__ old_to_new, EnsKnnOut out, const float* __restrict__ wmx,
    const float* __restrict__ wmy, const float* __restrict__ wmz, const float* __restrict__ ox,
    const float* __restrict__ oy, const float* __restrict__ oz) {
    GS_FLAT_LDS
    const uint32_t blk = PART == 0 ? xcd_block(blockIdx.x, gridDim.x) : blocks[xcd_block(blockIdx.x, gridDim.x)];
    const uint32_t s = blk * blockDim.x + threadIdx.x;
    const EnsPop& P = E.pops[E.block_pop[blk]];
    // Padding lanes stay through the searches (the wavefront completes its
    // lanes' inexact queries together) and return before the small-set part.
    const bool active = s < P.a0 + P.n;
    const float4 q = active ? spos[s] : make_float4(0.f, 0.f, 0.f, 0.f);
    if constexpr (PART == 0) {
        TopK<KMAX> h;
        h.init(k);
        const float thr = (out.seed_self && active) ? seed_bound<KMAX>(out.seed_self, k, N, new_to_old[s], N,
                                                                    old_to_new, q.x, q.y, q.z, spos)
                                                    : INFINITY;
        FlatQuery Q{q.x, q.y, q.z, thr, s, active};
        knn_search_flat<KMAX>(h, Q, P.g, N, spos, start + P.c0, flds);
        if (active) store_idx<KMAX>(h, k, N, s, out.self, out.out_map);
        return;
    }
    if (P.kind != GS_SCENARIO_CHOKE) return;
    const EnsPop& O = E.pops[P.other];
    if (O.n > 0) {
        TopK<KMAX> h;
        h.init(k);
        const float cap2 = E.scens[P.scen].other_cap2;
        float thr = fminf(cap2, (out.seed_other && active) ? seed_bound<KMAX>(out.seed_other, k, N, new_to_old[s], N,
                                                                           old_to_new, q.x, q.y, q.z, spos)
                                                        : INFINITY);
        // A query the cap cannot reach the other group's box from has no
        // candidate.
        if (active && box_beyond(out.box + 6 * (size_t)P.other, q.x, q.y, q.z, cap2)) thr = -1.f;
        FlatQuery Q{q.x, q.y, q.z, thr, UINT32_MAX, active};
        knn_search_flat<KMAX>(h, Q, O.g, N, spos, start + O.c0, flds);
        if (active) store_idx<KMAX>(h, k, N, s, out.other, out.out_map);
    } else if (active) {
        TopK<KMAX> h;
        h.init(k);
        store_idx<KMAX>(h, k, N, s, out.other, out.out_map);
    }
    if (!active) return;
    const EnsScen& S = E.scens[P.scen];
    // Lists of up to four entries (the crowds' 3 walls and 2 obstacles) take
    // 4-entry registers: each insertion ripples through 4 entries, not 8.
    auto small = [&](uint32_t kk, uint32_t n, const float* x, const float* y, const float* z, uint32_t* o) {
        if (kk <= 4) {
            TopK<4> h;
            h.init(kk);
            brute_small<4>(h, q.x, q.y, q.z, n, x, y, z);
            store_local<4>(h, kk, N, s, o, out.out_map);
        } else {
            TopK<8> h;
            h.init(kk);
            brute_small<8>(h, q.x, q.y, q.z, n, x, y, z);
            store_local<8>(h, kk, N, s, o, out.out_map);
        }
    };
    small(S.k_walls, S.nw, wmx + S.w0, wmy + S.w0, wmz + S.w0, out.walls);
    small(S.k_obs, S.no, ox + S.o0, oy + S.o0, oz + S.o0, out.obs);
}

// Side, up and (unless frozen) the smoothed acceleration are read only by the
// agent's own step, so the counting sort leaves them in the previous storage
// order: the step kernels read them at the agent's previous slot old[s] from
// the current buffers and the integrate writes them in the new order into
// the alternate buffers (the host flips them after the step).
struct OwnFields {
    const uint32_t* old;           // new slot -> previous slot
    float *sx, *sy, *sz, *ux, *uy, *uz;  // alternate buffers of side and up
    float *ax, *ay, *az;           // of the smoothed acceleration; NULL when frozen
};

// Boids: flocking + integrate from the step-t snapshot (as k_flock_update).
template <int MODE>
__device__ __forceinline__ void boids_one(const EnsScen& S, uint32_t s, uint32_t N, uint32_t k, Agents A,
                                          float* __restrict__ npx, float* __restrict__ npy, float* __restrict__ npz,
                                          float* __restrict__ nfx, float* __restrict__ nfy, float* __restrict__ nfz,
                                          const float4* __restrict__ spos, const uint32_t* __restrict__ knn,
                                          const OwnFields& O) {
    const float4 q = spos[s];
    const f3 pos = mk3(q.x, q.y, q.z);
    const f3 fwd = ld3(A.fx, A.fy, A.fz, s);
    uint32_t bits = A.bits[s];
    f3 bp[8], bf[8];
    const int m = gather_neighbours<8>(knn, k, N, s, N, spos, nullptr, nullptr, nullptr, A.fx, A.fy, A.fz,
                                       S.ali.on && !(bits & S.ali.mask), bp, bf);
    f3 steering = ld3(A.gx, A.gy, A.gz, s);
    flock_core<MODE, 8>(pos, fwd, A.radius[s], m, bp, bf, S.sep, S.ali, S.coh, steering, bits);
    f3 p = pos, f = fwd;
    const uint32_t o = O.old[s];
    f3 side = ld3(A.sx, A.sy, A.sz, o), up = ld3(A.ux, A.uy, A.uz, o);
    f3 sacc = ld3(A.ax, A.ay, A.az, O.ax ? o : s);
    float speed = A.speed[s];
    integrate<MODE>(steering, p, f, speed, sacc, side, up, A.max_force[s], A.max_speed[s], A.mass[s], S.dt);
    st3(npx, npy, npz, s, p);
    st3(nfx, nfy, nfz, s, f);
    st3(O.sx, O.sy, O.sz, s, side);
    st3(O.ux, O.uy, O.uz, s, up);
    if (O.ax) st3(O.ax, O.ay, O.az, s, sacc);  // every mode: the buffer flips
    A.speed[s] = speed;
    st3(A.gx, A.gy, A.gz, s, mk3(0.f, 0.f, 0.f));
    A.bits[s] = 0u;
}

__global__ __launch_bounds__(kBlock) void e_boids(EnsDev E, const uint32_t* __restrict__ blocks, uint32_t N,
                                                  uint32_t k, Agents A, float* __restrict__ npx,
                                                  float* __restrict__ npy, float* __restrict__ npz,
                                                  float* __restrict__ nfx, float* __restrict__ nfy,
                                                  float* __restrict__ nfz, const float4* __restrict__ spos,
                                                  const uint32_t* __restrict__ knn, OwnFields O) {
    const uint32_t b = blocks[xcd_block(blockIdx.x, gridDim.x)];
    const uint32_t s = b * kBlock + threadIdx.x;
    const EnsPop& P = E.pops[E.block_pop[b]];
    if (s >= P.a0 + P.n) return;
    const EnsScen& S = E.scens[P.scen];
    if (S.mode == GS_MODE_GPUREF)
        boids_one<GS_MODE_GPUREF>(S, s, N, k, A, npx, npy, npz, nfx, nfy, nfz, spos, knn, O);
    else
        boids_one<GS_MODE_OPENSTEER>(S, s, N, k, A, npx, npy, npz, nfx, nfy, nfz, spos, knn, O);
}

// Crowd phase a: clear the bits, avoid walls (moves touching agents), avoid
// obstacles — gs_clear_bits, gs_steer_avoid_walls, gs_steer_avoid_obstacles.
__global__ __launch_bounds__(kBlock) void e_choke_a(EnsDev E, const uint32_t* __restrict__ blocks, uint32_t N,
                                                    Agents A, const uint32_t* __restrict__ kwall,
                                                    const uint32_t* __restrict__ kobs, gs_walls W,
                                                    const float* __restrict__ ox, const float* __restrict__ oy,
                                                    const float* __restrict__ oz, const float* __restrict__ orad,
                                                    const uint32_t* __restrict__ old) {
    const uint32_t b = blocks[xcd_block(blockIdx.x, gridDim.x)];
    const uint32_t s = b * kBlock + threadIdx.x;
    const EnsPop& P = E.pops[E.block_pop[b]];
    if (s >= P.a0 + P.n) return;
    const EnsScen& S = E.scens[P.scen];
    uint32_t bits = 0u;
    f3 steering = ld3(A.gx, A.gy, A.gz, s);
    f3 pos = ld3(A.px, A.py, A.pz, s), dir = ld3(A.fx, A.fy, A.fz, s);
    const uint32_t o = old[s];  // side and up are in the previous order
    const f3 side = ld3(A.sx, A.sy, A.sz, o);
    const float radius = A.radius[s], speed = A.speed[s];
    const gs_walls Ws{W.sx + S.w0, W.sy + S.w0, W.sz + S.w0, W.ex + S.w0, W.ey + S.w0,
                      W.ez + S.w0, W.nx + S.w0, W.ny + S.w0, W.nz + S.w0};
    avoid_walls_one(pos, dir, side, radius, speed, kwall + s, N, S.k_walls, S.nw, Ws, S.mt_walls, S.w_walls,
                    steering, bits);
    avoid_obstacles_one(pos, dir, side, ld3(A.ux, A.uy, A.uz, o), radius, speed, kobs + s, N, S.k_obs, S.no,
                        ox + S.o0, oy + S.o0, oz + S.o0, orad + S.o0, S.mt_obs, S.w_obs, steering, bits);
    st3(A.px, A.py, A.pz, s, pos);
    st3(A.fx, A.fy, A.fz, s, dir);
    st3(A.gx, A.gy, A.gz, s, steering);
    A.bits[s] = bits;
}

template <int MODE>
__device__ __forceinline__ void choke_b_one(const EnsScen& S, int group, uint32_t s, uint32_t N, uint32_t k, Agents A,
                                            const uint32_t* __restrict__ kself, const uint32_t* __restrict__ kother) {
    uint32_t bits = A.bits[s];
    f3 steering = ld3(A.gx, A.gy, A.gz, s);
    const f3 pos = ld3(A.px, A.py, A.pz, s), fwd = ld3(A.fx, A.fy, A.fz, s);
    const float radius = A.radius[s];
    const FlockTerm off{0.f, 0.f, 0.f, 0.f, 0u, 0};
    const uint32_t* cols[2] = {kself, kother};
    for (int w = 0; w < 2; ++w) {
        if (bits & S.csep.mask) break;  // bits only accumulate
        f3 bp[8], bf[8];
        const int m = gather_neighbours<8>(cols[w], k, N, s, N, nullptr, A.px, A.py, A.pz, nullptr, nullptr,
                                           nullptr, false, bp, bf);
        flock_core<MODE, 8>(pos, fwd, radius, m, bp, bf, S.csep, off, off, steering, bits);
    }
    if (!(bits & S.mask_seek)) {
        const f3 desired = mk3(S.goal[group][0], S.goal[group][1], S.goal[group][2]) - pos;
        const f3 f = (MODE == GS_MODE_GPUREF) ? normalize3(desired - fwd) : desired - fwd * A.speed[s];
        apply(steering, bits, f, S.w_seek, GS_BIT_SEEK);
    }
    st3(A.gx, A.gy, A.gz, s, steering);
    A.bits[s] = bits;
}

// Crowd phase b: separation from the own group, from the other group, seek
// the group's goal — gs_steer_separation (x2), gs_steer_seek.
__global__ __launch_bounds__(kBlock) void e_choke_b(EnsDev E, const uint32_t* __restrict__ blocks, uint32_t N,
                                                    uint32_t k, Agents A, const uint32_t* __restrict__ kself,
                                                    const uint32_t* __restrict__ kother) {
    const uint32_t b = blocks[xcd_block(blockIdx.x, gridDim.x)];
    const uint32_t s = b * kBlock + threadIdx.x;
    const EnsPop& P = E.pops[E.block_pop[b]];
    if (s >= P.a0 + P.n) return;
    const EnsScen& S = E.scens[P.scen];
    if (S.mode == GS_MODE_GPUREF) choke_b_one<GS_MODE_GPUREF>(S, P.group, s, N, k, A, kself, kother);
    else choke_b_one<GS_MODE_OPENSTEER>(S, P.group, s, N, k, A, kself, kother);
}

template <int MODE>
__device__ __forceinline__ void update_one(float dt, uint32_t s, Agents A, float* __restrict__ npx,
                                           float* __restrict__ npy, float* __restrict__ npz, float* __restrict__ nfx,
                                           float* __restrict__ nfy, float* __restrict__ nfz, const OwnFields& O) {
    f3 pos = ld3(A.px, A.py, A.pz, s), fwd = ld3(A.fx, A.fy, A.fz, s);
    const uint32_t o = O.old[s];
    f3 side = ld3(A.sx, A.sy, A.sz, o), up = ld3(A.ux, A.uy, A.uz, o);
    f3 sacc = ld3(A.ax, A.ay, A.az, O.ax ? o : s);
    float speed = A.speed[s];
    integrate<MODE>(ld3(A.gx, A.gy, A.gz, s), pos, fwd, speed, sacc, side, up, A.max_force[s], A.max_speed[s],
                    A.mass[s], dt);
    st3(npx, npy, npz, s, pos);
    st3(nfx, nfy, nfz, s, fwd);
    st3(O.sx, O.sy, O.sz, s, side);
    st3(O.ux, O.uy, O.uz, s, up);
    if (O.ax) st3(O.ax, O.ay, O.az, s, sacc);  // every mode: the buffer flips
    A.speed[s] = speed;
    st3(A.gx, A.gy, A.gz, s, mk3(0.f, 0.f, 0.f));
    A.bits[s] = 0u;
}

__global__ __launch_bounds__(kBlock) void e_choke_update(EnsDev E, const uint32_t* __restrict__ blocks, Agents A,
                                                         float* __restrict__ npx, float* __restrict__ npy,
                                                         float* __restrict__ npz, float* __restrict__ nfx,
                                                         float* __restrict__ nfy, float* __restrict__ nfz,
                                                         OwnFields O) {
    const uint32_t b = blocks[xcd_block(blockIdx.x, gridDim.x)];
    const uint32_t s = b * kBlock + threadIdx.x;
    const EnsPop& P = E.pops[E.block_pop[b]];
    if (s >= P.a0 + P.n) return;
    const EnsScen& S = E.scens[P.scen];
    if (S.mode == GS_MODE_GPUREF) update_one<GS_MODE_GPUREF>(S.dt, s, A, npx, npy, npz, nfx, nfy, nfz, O);
    else update_one<GS_MODE_OPENSTEER>(S.dt, s, A, npx, npy, npz, nfx, nfy, nfz, O);
}

// ---------------------------------------------------------------------------
// Host side.

static FlockTerm eterm(const gs_flock_term& t) { return FlockTerm{t.min_d, t.max_d, t.cos_max, t.weight, t.mask, 1}; }

template <class T> static void h2d_vec(DevBuf<T>& d, const std::vector<T>& v, hipStream_t st) {
    d.reserve(v.size() ? v.size() : 1);
    if (!v.empty()) GS_HIP(hipMemcpyAsync(d.p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
}

void ens_init(gs_ensemble_s* e, uint32_t ns, const gs_scenario* sc, uint32_t k,
              GridDev (*make_grid)(const float*, const float*, const uint32_t*)) {
    e->k = k;
    e->scen.assign(sc, sc + ns);
    std::vector<EnsPop> pops;
    std::vector<EnsScen> scens(ns);
    uint32_t a = 0, c = 0;
    for (uint32_t i = 0; i < ns; ++i) {
        const gs_scenario& d = sc[i];
        EnsScen& S = scens[i];
        memset(&S, 0, sizeof S);
        S.kind = d.kind;
        S.mode = d.mode;
        S.dt = d.dt;
        S.wrap_shape = d.kind == GS_SCENARIO_BOIDS ? d.wrap_shape : -1;
        for (int j = 0; j < 3; ++j) {
            S.wc[j] = d.wrap_center[j];
            S.ws[j] = d.wrap_size[j];
            S.goal[0][j] = d.goal[0][j];
            S.goal[1][j] = d.goal[1][j];
        }
        S.sep = eterm(d.separation);
        S.ali = eterm(d.alignment);
        S.coh = eterm(d.cohesion);
        S.csep = eterm(d.choke_separation);
        S.w_seek = d.seek_weight;
        S.mask_seek = d.seek_mask;
        S.mt_walls = d.walls_min_time;
        S.w_walls = d.walls_weight;
        S.mt_obs = d.obstacles_min_time;
        S.w_obs = d.obstacles_weight;
        S.k_walls = d.k_walls;
        S.k_obs = d.k_obstacles;
        S.other_cap2 = d.other_max_distance > 0.f && std::isfinite(d.other_max_distance)
                           ? std::nextafter(d.other_max_distance * d.other_max_distance, 0.f)
                           : INFINITY;
        const GridDev g = make_grid(d.origin, d.size, d.cells);
        const uint32_t cells = d.cells[0] * d.cells[1] * d.cells[2];
        const int ng = d.kind == GS_SCENARIO_CHOKE ? 2 : 1;
        const int first = (int)pops.size();
        for (int gi = 0; gi < ng; ++gi) {
            EnsPop P;
            P.a0 = a;
            P.n = d.n[gi];
            P.c0 = c;
            P.cells = cells;
            P.g = g;
            P.kind = d.kind;
            P.group = gi;
            P.other = ng == 2 ? first + (1 - gi) : -1;
            P.scen = i;
            pops.push_back(P);
            const uint64_t padded = (uint64_t)((d.n[gi] + kBlock - 1) / kBlock) * kBlock;
            if ((uint64_t)a + padded >= kFlatMaxPoints)
                fail(GS_ERR_UNSUPPORTED, "ensemble larger than 2^27 agent slots");
            a += (uint32_t)padded;
            c += cells + 1;
        }
    }
    e->N = a;
    e->C = c;
    e->pops = pops;
    e->scens_host = scens;
    // Block lists per phase.
    std::vector<uint32_t> block_pop(a / kBlock);
    for (uint32_t p = 0; p < pops.size(); ++p) {
        const uint32_t b0 = pops[p].a0 / kBlock, nb = (pops[p].n + kBlock - 1) / kBlock;
        for (uint32_t b = 0; b < nb; ++b) {
            block_pop[b0 + b] = p;
            const EnsPop& P = pops[p];
            if (P.kind == GS_SCENARIO_BOIDS) e->blocks_boids.push_back(b0 + b);
            else {
                (P.group == 0 ? e->blocks_g0 : e->blocks_g1).push_back(b0 + b);
                e->blocks_choke.push_back(b0 + b);
            }
        }
    }
    e->block_pop_h = block_pop;
}

void ens_release(gs_ensemble_s* e) {
    e->d_pops.release();
    e->d_scens.release();
    for (auto* b : {&e->d_block_pop, &e->d_blocks_boids, &e->d_blocks_g0, &e->d_blocks_g1, &e->d_blocks_choke,
                    &e->knn_walls, &e->knn_obs, &e->key, &e->rank, &e->id_tmp, &e->new_to_old,
                    &e->old_to_new, &e->count, &e->start})
        b->release();
    for (auto& f : e->fb)
        for (auto& b : f) b.release();
    for (int b = 0; b < 2; ++b) {
        e->bitsb[b].release();
        e->idb[b].release();
        e->knn_self[b].release();
        e->knn_other[b].release();
        e->box[b].release();
    }
    e->spos.release();
    for (auto& b : e->walls) b.release();
    for (auto& b : e->obs) b.release();
}

void ens_alloc(gs_ensemble_s* e) {
    hipStream_t st = e->ctx->stream;
    const size_t N = e->N ? e->N : 1;
    for (int f = 0; f < kFields; ++f)
        for (int b = 0; b < 2; ++b) e->fb[f][b].reserve(N);
    for (int b = 0; b < 2; ++b) {
        e->bitsb[b].reserve(N);
        e->idb[b].reserve(N);
        e->knn_self[b].reserve(N * e->k);
        e->knn_other[b].reserve(N * e->k);
    }
    uint32_t kw = 1, ko = 1;
    for (auto& s : e->scens_host) {
        kw = std::max(kw, s.k_walls);
        ko = std::max(ko, s.k_obs);
    }
    e->kw_max = kw;
    e->ko_max = ko;
    e->knn_walls.reserve(N * kw);
    e->knn_obs.reserve(N * ko);
    e->key.reserve(N);
    e->rank.reserve(N);
    e->id_tmp.reserve(N);
    e->new_to_old.reserve(N);
    e->old_to_new.reserve(N);
    e->spos.reserve(N);
    const uint32_t ntiles = (e->C + 2047u) / 2048u;
    e->count.reserve(e->C ? e->C : 1);
    e->start.reserve((size_t)e->C + 1 + ntiles);
    launch_fill_u32(e->count.p, e->C, 0u, st);
    // Both box buffers start empty (min above max for every population).
    std::vector<uint32_t> box0(6 * std::max<size_t>(e->pops.size(), 1));
    for (size_t j = 0; j < box0.size(); ++j) box0[j] = j % 6 < 3 ? kBoxEmptyMin : kBoxEmptyMax;
    for (auto& b : e->box) h2d_vec(b, box0, st);
    GS_HIP(hipStreamSynchronize(st));  // box0 is read before it goes out of scope
    h2d_vec(e->d_pops, e->pops, st);
    h2d_vec(e->d_scens, e->scens_host, st);
    h2d_vec(e->d_block_pop, e->block_pop_h, st);
    h2d_vec(e->d_blocks_boids, e->blocks_boids, st);
    h2d_vec(e->d_blocks_g0, e->blocks_g0, st);
    h2d_vec(e->d_blocks_g1, e->blocks_g1, st);
    h2d_vec(e->d_blocks_choke, e->blocks_choke, st);
    // Padding slots: distinct ids per population (never searched).
    std::vector<uint32_t> ids(N, 0u);
    for (auto& P : e->pops) {
        const uint32_t end = (uint32_t)((P.n + kBlock - 1) / kBlock) * kBlock;
        for (uint32_t j = 0; j < end; ++j) ids[P.a0 + j] = j;
    }
    for (int b = 0; b < 2; ++b) {
        GS_HIP(hipMemcpyAsync(e->idb[b].p, ids.data(), N * 4, hipMemcpyHostToDevice, st));
        launch_fill_u32(e->bitsb[b].p, (uint32_t)N, 0u, st);
    }
    for (int f = 0; f < kFields; ++f)
        for (int b = 0; b < 2; ++b) launch_fill_u32((uint32_t*)e->fb[f][b].p, (uint32_t)N, 0u, st);
    e->walls_mid_ready = false;
}

void ens_upload_geometry(gs_ensemble_s* e) {
    hipStream_t st = e->ctx->stream;
    // Scenario ranges into the concatenated wall / obstacle arrays.
    std::vector<float> w[12], o[4];
    for (uint32_t i = 0; i < e->scens_host.size(); ++i) {
        EnsScen& S = e->scens_host[i];
        S.w0 = (uint32_t)w[0].size();
        S.nw = (uint32_t)e->walls_h[i].size() / 12;
        for (uint32_t j = 0; j < S.nw; ++j) {
            // start, end, normal, KNN position
            const float* r = &e->walls_h[i][(size_t)j * 12];
            for (int f = 0; f < 12; ++f) w[f].push_back(r[f]);
        }
        S.o0 = (uint32_t)o[0].size();
        S.no = (uint32_t)e->obs_h[i].size() / 4;
        for (uint32_t j = 0; j < S.no; ++j)
            for (int f = 0; f < 4; ++f) o[f].push_back(e->obs_h[i][(size_t)j * 4 + f]);
    }
    for (int f = 0; f < 12; ++f) h2d_vec(e->walls[f], w[f], st);
    for (int f = 0; f < 4; ++f) h2d_vec(e->obs[f], o[f], st);
    h2d_vec(e->d_scens, e->scens_host, st);
    e->walls_mid_ready = true;
}

static Agents ens_agents(gs_ensemble_s* e) {
    Agents a;
    float** p = (float**)&a;
    for (int f = 0; f < kFields; ++f) p[f] = e->F(f);
    a.bits = e->bits();
    return a;
}

static EnsDev ens_dev(gs_ensemble_s* e) {
    return EnsDev{e->d_pops.p, e->d_scens.p, e->d_block_pop.p};
}

// Exact KNN lists of every agent (own group, other group capped, walls,
// obstacles) for state held in a fixed order: the positions px/py/pz are
// binned into a sorted snapshot (the ensemble's bin buffers as scratch; its
// own state and KNN containers are not touched) and the lists are written in
// the fixed order (rows and neighbour values are fixed indices).  Seeds are
// The ensemble search kernels for the ensemble's k (7-entry lists for
// k = 7): own groups over all blocks, then the crowds' other lists.
#define GS_E_KNN(...)                                                                            \
    do {                                                                                         \
        if (e->k == 7)                                                                           \
            e_knn<7, 0><<<nblocks, kBlock, 0, st>>>(E, nullptr, __VA_ARGS__);                    \
        else                                                                                     \
            e_knn<8, 0><<<nblocks, kBlock, 0, st>>>(E, nullptr, __VA_ARGS__);                    \
        GS_LAUNCHED("e_knn");                                                                    \
        const unsigned ncb = (unsigned)e->blocks_choke.size();                                   \
        if (ncb && e->k == 7)                                                                    \
            e_knn<7, 1><<<ncb, kBlock, 0, st>>>(E, e->d_blocks_choke.p, __VA_ARGS__);            \
        else if (ncb)                                                                            \
            e_knn<8, 1><<<ncb, kBlock, 0, st>>>(E, e->d_blocks_choke.p, __VA_ARGS__);            \
    } while (0)

// previous lists in the same fixed order (or NULL).
void ens_search_fixed(gs_ensemble_s* e, float* px, float* py, float* pz, const uint32_t* ids_fixed,
                      uint32_t* ids_scratch, float* const scratch[3], const uint32_t* seed_self,
                      const uint32_t* seed_other, uint32_t* out_self, uint32_t* out_other, uint32_t* out_walls,
                      uint32_t* out_obs) {
    hipStream_t st = e->ctx->stream;
    const uint32_t N = e->N, nblocks = N / kBlock;
    if (N == 0) return;
    const EnsDev E = ens_dev(e);
    e->box_cur ^= 1;
    const uint32_t nbox = 6 * (uint32_t)e->pops.size();
    e_bin_key<<<nblocks, kBlock, 0, st>>>(E, px, py, pz, e->key.p, e->rank.p, e->count.p, e->box[e->box_cur].p,
                                          e->box[e->box_cur ^ 1].p, nbox);
    GS_LAUNCHED("e_bin_key");
    launch_scan_counts(e->count.p, e->start.p, e->C, st);
    FieldSet fs{};
    fs.src[0] = px;
    fs.src[1] = py;
    fs.src[2] = pz;
    for (int a = 0; a < 3; ++a) fs.dst[a] = scratch[a];
    fs.nf = 3;
    launch_scatter_reorder(N, e->key.p, e->rank.p, e->start.p, ids_fixed, ids_scratch, e->id_tmp.p,
                           e->new_to_old.p, e->old_to_new.p, nullptr, nullptr, fs, e->spos.p, st);
    EnsKnnOut ko;
    ko.self = out_self;
    ko.other = out_other;
    ko.walls = out_walls;
    ko.obs = out_obs;
    ko.seed_self = seed_self;
    ko.seed_other = seed_other;
    ko.out_map = e->new_to_old.p;
    ko.box = e->box[e->box_cur].p;
    GS_E_KNN(N, e->k, e->spos.p, e->start.p, e->new_to_old.p, e->old_to_new.p, ko,
                                      e->walls[9].p, e->walls[10].p, e->walls[11].p, e->obs[0].p, e->obs[1].p,
                                      e->obs[2].p);
    GS_LAUNCHED("e_knn");
}

// One step, queued on the ensemble's stream (capturable: kernels only).
void ens_step(gs_ensemble_s* e) {
    hipStream_t st = e->ctx->stream;
    const uint32_t N = e->N, nblocks = N / kBlock;
    if (N == 0) return;
    const EnsDev E = ens_dev(e);
    e->box_cur ^= 1;
    const uint32_t nbox = 6 * (uint32_t)e->pops.size();
    e_bin_key<<<nblocks, kBlock, 0, st>>>(E, e->F(0), e->F(1), e->F(2), e->key.p, e->rank.p, e->count.p,
                                          e->box[e->box_cur].p, e->box[e->box_cur ^ 1].p, nbox);
    GS_LAUNCHED("e_bin_key");
    launch_scan_counts(e->count.p, e->start.p, e->C, st);
    // Moved: every field but the frozen ones and the agent's own fields
    // (side, up, smoothed acceleration: see OwnFields), positions first, as
    // the reorder's snapshot reads them there.
    const uint32_t steer = 0x7u << 16;
    const uint32_t own = (0x3Fu << 10) | ((0x7u << 7) & ~e->frozen);
    const uint32_t moved = ((1u << kFields) - 1u) & ~e->frozen & ~(e->steer_zero ? steer : 0u) & ~own;
    FieldSet fs{};
    int nf = 0;
    for (int f = 0; f < kFields; ++f)
        if (moved & (1u << f)) {
            fs.src[nf] = e->F(f);
            fs.dst[nf] = e->Falt(f);
            ++nf;
        }
    fs.nf = nf;
    const bool move_bits = !e->steer_zero;
    launch_scatter_reorder(N, e->key.p, e->rank.p, e->start.p, e->ids(), e->idb[e->icur ^ 1].p, e->id_tmp.p,
                           e->new_to_old.p, e->old_to_new.p, move_bits ? e->bits() : nullptr,
                           move_bits ? e->bitsb[e->bcur ^ 1].p : nullptr, fs, e->spos.p, st);
    for (int f = 0; f < kFields; ++f)
        if (moved & (1u << f)) e->fcur[f] ^= 1;
    if (move_bits) e->bcur ^= 1;
    e->icur ^= 1;
    EnsKnnOut ko;
    ko.self = e->knn_self[e->kcur ^ 1].p;
    ko.other = e->knn_other[e->kcur ^ 1].p;
    ko.walls = e->knn_walls.p;
    ko.obs = e->knn_obs.p;
    ko.seed_self = e->seeded ? e->knn_self[e->kcur].p : nullptr;
    ko.seed_other = e->seeded ? e->knn_other[e->kcur].p : nullptr;
    ko.out_map = nullptr;
    ko.box = e->box[e->box_cur].p;
    GS_E_KNN(N, e->k, e->spos.p, e->start.p, e->new_to_old.p, e->old_to_new.p, ko,
                                      e->walls[9].p, e->walls[10].p, e->walls[11].p, e->obs[0].p, e->obs[1].p,
                                      e->obs[2].p);
    GS_LAUNCHED("e_knn");
    e->kcur ^= 1;
    e->seeded = true;
    const Agents A = ens_agents(e);
    const bool acc_own = (own >> 7) & 1u;
    const OwnFields O{e->new_to_old.p, e->Falt(10), e->Falt(11), e->Falt(12), e->Falt(13), e->Falt(14), e->Falt(15),
                      acc_own ? e->Falt(7) : nullptr, acc_own ? e->Falt(8) : nullptr, acc_own ? e->Falt(9) : nullptr};
    const uint32_t* kself = e->knn_self[e->kcur].p;
    const uint32_t* kother = e->knn_other[e->kcur].p;
    if (!e->blocks_boids.empty()) {
        e_boids<<<(unsigned)e->blocks_boids.size(), kBlock, 0, st>>>(E, e->d_blocks_boids.p, N, e->k, A, e->Falt(0),
                                                                      e->Falt(1), e->Falt(2), e->Falt(3), e->Falt(4),
                                                                      e->Falt(5), e->spos.p, kself, O);
        GS_LAUNCHED("e_boids");
    }
    if (!e->blocks_choke.empty()) {
        const gs_walls W{e->walls[0].p, e->walls[1].p, e->walls[2].p, e->walls[3].p, e->walls[4].p,
                         e->walls[5].p, e->walls[6].p, e->walls[7].p, e->walls[8].p};
        for (int g = 0; g < 2; ++g) {
            const auto& bl = g == 0 ? e->blocks_g0 : e->blocks_g1;
            const uint32_t* dbl = g == 0 ? e->d_blocks_g0.p : e->d_blocks_g1.p;
            if (bl.empty()) continue;
            e_choke_a<<<(unsigned)bl.size(), kBlock, 0, st>>>(E, dbl, N, A, e->knn_walls.p, e->knn_obs.p, W,
                                                              e->obs[0].p, e->obs[1].p, e->obs[2].p, e->obs[3].p,
                                                              e->new_to_old.p);
            GS_LAUNCHED("e_choke_a");
            e_choke_b<<<(unsigned)bl.size(), kBlock, 0, st>>>(E, dbl, N, e->k, A, kself, kother);
            GS_LAUNCHED("e_choke_b");
        }
        e_choke_update<<<(unsigned)e->blocks_choke.size(), kBlock, 0, st>>>(
            E, e->d_blocks_choke.p, A, e->Falt(0), e->Falt(1), e->Falt(2), e->Falt(3), e->Falt(4), e->Falt(5), O);
        GS_LAUNCHED("e_choke_update");
    }
    for (int f = 0; f < 6; ++f) e->fcur[f] ^= 1;
    for (int f = 7; f < 16; ++f)
        if (own & (1u << f)) e->fcur[f] ^= 1;
    // Every agent's steering and bits are now zero.
    e->steer_zero = true;
}

}  // namespace gs
