// Plugin-compat kernels: the per-step extras the choke-point plugins call
// around the integrate (plugins/ChokePointNew.cpp:620-629,
// plugins/ChokePointGroup.cpp:723-727):
//
//   k_update_anti_penetration  UpdateCUDAKernelNew (CUDA/UpdateCUDAKernelNew.cu:95-268)
//   k_anti_penetration_agents  AntiPenetrationAgentsCUDAKernel (CUDA/AntiPenetrationAgentsCUDAKernel.cu:57-109)
//   k_follow_path              FollowPathCUDAKernel (CUDA/SteerToFollowPathCUDAKernel.cu:189-287)
//                              over a PolylinePathwayCUDA (CUDA/PolylinePathwayCUDA.cu:17-60)
//   k_avoid_neighbors          SteerToAvoidCloseNeighborsCUDAKernel + SteerToAvoidNeighborsCUDAKernel
//                              (CUDA/SteerToAvoidNeighborsCUDAKernel.cu:93-402), the multi-group
//                              plugin's avoidance of the other groups (plugins/MultiGroup.cpp:571-579)
//
// One lane per agent, SoA fields read and written coalesced, like the other
// behaviours.  Walls are gathered through the agent's KNN column; the path
// (a few points, the same for every lane) is read with uniform loads.
#include "gs_ops.hpp"

namespace gs {

// ClosestPointOnLine (UpdateCUDAKernelNew.cu:73-81).  The reference's
// clamp(0.f, 1.f, t) has its arguments swapped and returns 1 for every t
// (:67-71,79); GS_CLOSEST_AS_REFERENCE keeps that (the point start + v * 1),
// GS_CLOSEST_CLAMPED clamps t into [0, 1].
__device__ __forceinline__ f3 closest_point_on_line(f3 p, f3 s, f3 e, int closest) {
    const f3 v = e - s;
    const f3 w = p - s;
    float t = dot3(w, v) / dot3(v, v);
    if (closest == GS_CLOSEST_AS_REFERENCE) t = 1.f;
    else t = t < 0.f ? 0.f : (t > 1.f ? 1.f : t);
    return s + v * t;
}

// The GPU-reference integrate (UpdateCUDAKernel's formulas) with wall
// anti-penetration between the speed limit and the Euler step: over the
// agent's k nearest walls in KNN order (stopping at a pad), the nearest
// crossing of position -> position + v dt and the LAST wall whose closest
// point lies within the radius (the reference keeps no minimum there,
// :215-221).  Within the radius: move by -normal * (d - r); crossing: keep
// the velocity's component perpendicular to that wall's normal; either sets
// GS_BIT_ANTI_PENETRATION_WALL.  Steering is zeroed, the bits are kept (the
// reference's reset is commented out, :153).  The KNN is read slot-major, as
// the search writes it (the reference indexes it agent-major, :150,187).  At
// speed 0 the side stays unchanged (the reference stores shared memory it
// never loaded, :155,261).
__global__ __launch_bounds__(kBlock) void k_update_anti_penetration(uint32_t n, Agents A, uint32_t nW, gs_walls W,
                                                                    const uint32_t* __restrict__ knn, uint32_t k,
                                                                    float dt, int closest) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    f3 pos = ld3(A.px, A.py, A.pz, i), fwd = ld3(A.fx, A.fy, A.fz, i);
    f3 side = ld3(A.sx, A.sy, A.sz, i), up = ld3(A.ux, A.uy, A.uz, i);
    const float radius = A.radius[i];
    const f3 force = truncate3(ld3(A.gx, A.gy, A.gz, i), A.max_force[i]);
    const f3 acc = force / A.mass[i];
    f3 new_vel = truncate3(fwd * A.speed[i] + acc * dt, A.max_speed[i]);
    const f3 future = pos + new_vel * dt;

    float line_d = FLT_MAX;
    bool line = false;
    f3 line_n = mk3(0.f, 0.f, 0.f);
    float col_d = FLT_MAX;
    bool col = false;
    f3 col_n = line_n;
    for (uint32_t j = 0; j < k; ++j) {
        const uint32_t l = knn[(size_t)j * n + i];
        if (l >= nW) break;
        const f3 ls = ld3(W.sx, W.sy, W.sz, l);
        const f3 le = ld3(W.ex, W.ey, W.ez, l);
        const f3 ln = ld3(W.nx, W.ny, W.nz, l);
        f3 ip;
        if (lines_intersect(pos, future, ls, le, ip)) {
            const float d = len3(pos - ip);
            if (d < col_d) {
                col_d = d;
                col = true;
                col_n = ln;
            }
        }
        const f3 cp = closest_point_on_line(pos, ls, le, closest) - pos;
        const float d2 = dot3(cp, cp);
        if (d2 < radius * radius) {
            line_d = sqrtf(d2);
            line = true;
            line_n = ln;
        }
    }
    uint32_t bits = A.bits[i];
    if (line) {
        pos = pos + (-line_n) * (line_d - radius);
        bits |= GS_BIT_ANTI_PENETRATION_WALL;
    }
    if (col) {
        new_vel = perp3(new_vel, col_n);
        bits |= GS_BIT_ANTI_PENETRATION_WALL;
    }
    const float speed = len3(new_vel);
    if (speed > 0.f) {
        fwd = new_vel / speed;
        side = normalize3(cross3(fwd, up));
        up = cross3(side, fwd);
    }
    pos = pos + new_vel * dt;
    st3(A.px, A.py, A.pz, i, pos);
    st3(A.fx, A.fy, A.fz, i, fwd);
    st3(A.sx, A.sy, A.sz, i, side);
    st3(A.ux, A.uy, A.uz, i, up);
    A.speed[i] = speed;
    st3(A.gx, A.gy, A.gz, i, mk3(0.f, 0.f, 0.f));
    A.bits[i] = bits;
}

// Push-apart from the nearest agent of B (the first KNN entry and its
// distance): when r_a + r_b - d >= 0, move along normalize(p_a - p_b) by the
// whole overlap if B's agent has a masked bit, else by half, and set
// GS_BIT_ANTI_PENETRATION_AGENT; masked agents stay.  Every lane reads the
// positions and bits before the call (A's current buffers, and B's) and
// writes A's into the alternate buffers, so A may be B.
__global__ __launch_bounds__(kBlock) void k_anti_penetration_agents(
    uint32_t nA, const float* __restrict__ px, const float* __restrict__ py, const float* __restrict__ pz,
    const float* __restrict__ radius, const uint32_t* __restrict__ bits_in, float* __restrict__ ox,
    float* __restrict__ oy, float* __restrict__ oz, uint32_t* __restrict__ bits_out,
    const uint32_t* __restrict__ knn0, const float* __restrict__ dist0, uint32_t nB, const float* __restrict__ bx,
    const float* __restrict__ by, const float* __restrict__ bz, const float* __restrict__ brad,
    const uint32_t* __restrict__ bbits, uint32_t mask) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nA) return;
    f3 pos = ld3(px, py, pz, i);
    uint32_t bits = bits_in[i];
    const uint32_t j = knn0[i];
    if (!(bits & mask) && j < nB) {
        const float overlap = (radius[i] + brad[j]) - dist0[i];
        if (!(overlap < 0.f)) {
            const f3 off = normalize3(pos - ld3(bx, by, bz, j));
            const float move = (bbits[j] & mask) ? overlap : 0.5f * overlap;
            pos = pos + off * move;
            bits |= GS_BIT_ANTI_PENETRATION_AGENT;
        }
    }
    st3(ox, oy, oz, i, pos);
    bits_out[i] = bits;
}

// A PolylinePathwayCUDA as kernels read it: m points, the unit direction
// and length of the segment ending at each point (entry 0 unused).
struct PathDev {
    const float *px, *py, *pz, *nx, *ny, *nz, *len;
    uint32_t m;
    float radius, total;
    int cyclic;
    __device__ __forceinline__ f3 pt(uint32_t i) const { return mk3(px[i], py[i], pz[i]); }
    __device__ __forceinline__ f3 nm(uint32_t i) const { return mk3(nx[i], ny[i], nz[i]); }
};

// pointToSegmentDistance (SteerToFollowPathCUDAKernel.cu:34-69).
__device__ __forceinline__ float point_to_segment(f3 p, f3 e0, f3 e1, f3 nrm, float seg_len, float& proj,
                                                  f3& chosen) {
    const f3 local = p - e0;
    proj = dot3(nrm, local);
    if (proj < 0.f) {
        chosen = e0;
        proj = 0.f;
        return len3(p - e0);
    }
    if (proj > seg_len) {
        chosen = e1;
        proj = seg_len;
        return len3(p - e1);
    }
    chosen = e0 + nrm * proj;
    return len3(p - chosen);
}

// mapPointToPathDistance (:71-102) for two points in one walk over the
// segments (each point's minimum is tracked separately, as two calls would).
__device__ __forceinline__ void path_distances(const PathDev& P, f3 a, f3 b, float& da, float& db) {
    float min_a = FLT_MAX, min_b = FLT_MAX, total = 0.f;
    da = 0.f;
    db = 0.f;
    for (uint32_t s = 1; s < P.m; ++s) {
        const f3 e0 = P.pt(s - 1), e1 = P.pt(s), nrm = P.nm(s);
        const float L = P.len[s];
        float proj;
        f3 chosen;
        float d = point_to_segment(a, e0, e1, nrm, L, proj, chosen);
        if (d < min_a) {
            min_a = d;
            da = total + proj;
        }
        d = point_to_segment(b, e0, e1, nrm, L, proj, chosen);
        if (d < min_b) {
            min_b = d;
            db = total + proj;
        }
        total += L;
    }
}

// mapPointToPath (:104-142): how far p lies outside the tube.
__device__ __forceinline__ float path_outside(const PathDev& P, f3 p) {
    float min_d = FLT_MAX;
    f3 on = mk3(0.f, 0.f, 0.f);
    for (uint32_t s = 1; s < P.m; ++s) {
        float proj;
        f3 chosen;
        const float d = point_to_segment(p, P.pt(s - 1), P.pt(s), P.nm(s), P.len[s], proj, chosen);
        if (d < min_d) {
            min_d = d;
            on = chosen;
        }
    }
    return len3(on - p) - P.radius;
}

// mapPathDistanceToPoint (:144-187).  A distance the segment walk does not
// place (rounding leaves it above the last segment) maps to the last point,
// where the reference returns an unset value.
__device__ __forceinline__ f3 path_point_at(const PathDev& P, float dist) {
    float remaining = dist;
    if (P.cyclic) {
        remaining = fmodf(dist, P.total);
    } else {
        if (dist < 0.f) return P.pt(0);
        if (dist >= P.total) return P.pt(P.m - 1);
    }
    for (uint32_t s = 1; s < P.m; ++s) {
        const float L = P.len[s];
        if (L < remaining) {
            remaining -= L;
        } else {
            const float ratio = remaining / L;
            const f3 a = P.pt(s - 1);
            return a + (P.pt(s) - a) * ratio;  // interpolate (VectorUtils.cuh:292-295)
        }
    }
    return P.pt(P.m - 1);
}

// Path following: predict position + forward * speed * t; when the
// prediction lies inside the tube and moves forward along the path, no
// force; otherwise seek the path point prediction-offset ahead of the
// agent's own path distance: normalize(target - position - forward) * w.
__global__ __launch_bounds__(kBlock) void k_follow_path(uint32_t n, Agents A, PathDev P, float prediction_time,
                                                        float w, uint32_t mask) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t bits = A.bits[i];
    if (bits & mask) return;
    const f3 pos = ld3(A.px, A.py, A.pz, i), fwd = ld3(A.fx, A.fy, A.fz, i);
    const float speed = A.speed[i];
    const float offset = prediction_time * speed;
    const f3 future = pos + (fwd * speed) * prediction_time;
    float now_d, future_d;
    path_distances(P, pos, future, now_d, future_d);
    const bool rightway = (offset > 0.f) ? (now_d < future_d) : (now_d > future_d);
    const float outside = path_outside(P, future);
    f3 steering = mk3(0.f, 0.f, 0.f);
    if (!((outside < 0.f) && rightway)) {
        const f3 target = path_point_at(P, now_d + offset);
        steering = (target - pos) - fwd;
    }
    f3 acc = ld3(A.gx, A.gy, A.gz, i);
    apply(acc, bits, normalize3(steering), w, GS_BIT_FOLLOW_PATH);
    st3(A.gx, A.gy, A.gz, i, acc);
    A.bits[i] = bits;
}

// Neighbour avoidance over A's KNN in B (B is only read, so it may be A).
// GS_MODE_GPUREF: with avoid_close, SteerToAvoidCloseNeighborsCUDAKernel
// (CUDA/SteerToAvoidNeighborsCUDAKernel.cu:93-187) first, then
// SteerToAvoidNeighborsCUDAKernel (:235-402) for agents without the
// AVOID_CLOSE_NEIGHBORS bit (SteerToAvoidNeighborsCUDA.cu:143), fused into
// one pass per agent.  The head-on rule reads the approach position the
// kernel's helper wrote last (the last neighbour passing the time test,
// chosen or not, :334-338); the parallel rule uses the threat's current
// position (the kernel reads an unset variable, :372).
// GS_MODE_OPENSTEER: SteerLibrary::steerToAvoidNeighbors
// (src/SteerLibrary.cpp:290-389) with steerToAvoidCloseNeighbors(v, 0, .)
// (:470-495): current distances, threshold 2 r, the chosen threat's
// approach position, the perpendicular rule when the threat is not faster,
// forces added unnormalized.
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_avoid_neighbors(uint32_t nA, Agents A, uint32_t nB,
                                                            const float* __restrict__ bpx, const float* __restrict__ bpy,
                                                            const float* __restrict__ bpz, const float* __restrict__ bfx,
                                                            const float* __restrict__ bfy, const float* __restrict__ bfz,
                                                            const float* __restrict__ bspeed,
                                                            const float* __restrict__ brad,
                                                            const uint32_t* __restrict__ knn,
                                                            const float* __restrict__ knn_dist, uint32_t k,
                                                            float min_time, float min_sep, int avoid_close, float w,
                                                            uint32_t mask) {
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nA) return;
    uint32_t bits = A.bits[i];
    if (bits & mask) return;
    const f3 pos = ld3(A.px, A.py, A.pz, i), dir = ld3(A.fx, A.fy, A.fz, i), side = ld3(A.sx, A.sy, A.sz, i);
    const float speed = A.speed[i], radius = A.radius[i];
    f3 acc = ld3(A.gx, A.gy, A.gz, i);
    bool done = false;
    if (MODE == GS_MODE_GPUREF) {
        if (avoid_close) {
            f3 steering = mk3(0.f, 0.f, 0.f);
            for (uint32_t j = 0; j < k; ++j) {
                const uint32_t b = knn[(size_t)j * nA + i];
                if (b >= nB) break;
                const float min_c = min_sep + (radius + brad[b]);
                if (knn_dist[(size_t)j * nA + i] < min_c) {
                    steering = perp3(-(ld3(bpx, bpy, bpz, b) - pos), dir);
                    break;
                }
            }
            apply(acc, bits, normalize3(steering), w, GS_BIT_AVOID_CLOSE_NEIGHBORS);
        }
        done = (bits & GS_BIT_AVOID_CLOSE_NEIGHBORS) != 0;
    } else {
        f3 sep = mk3(0.f, 0.f, 0.f);
        for (uint32_t j = 0; j < k; ++j) {
            const uint32_t b = knn[(size_t)j * nA + i];
            if (b >= nB) break;
            const float min_c = 0.f + (radius + brad[b]);
            const f3 off = ld3(bpx, bpy, bpz, b) - pos;
            if (len3(off) < min_c) {
                sep = perp3(-off, dir);
                break;
            }
        }
        if (!zero3(sep)) {
            apply(acc, bits, sep, w, GS_BIT_AVOID_CLOSE_NEIGHBORS);
            done = true;
        }
    }
    if (!done) {
        float min_t = min_time;
        bool found = false;
        f3 last_near = mk3(0.f, 0.f, 0.f), threat_near = last_near, t_dir = last_near, t_pos = last_near;
        float t_speed = 0.f;
        for (uint32_t j = 0; j < k; ++j) {
            const uint32_t b = knn[(size_t)j * nA + i];
            if (b >= nB) break;
            const f3 bp = ld3(bpx, bpy, bpz, b), bd = ld3(bfx, bfy, bfz, b);
            const float bs = bspeed[b];
            const float threshold = MODE == GS_MODE_GPUREF ? radius + brad[b] : radius * 2.f;
            // predictNearestApproachTime (:205-233)
            float time = 0.f;
            const f3 rel = bd * bs - dir * speed;
            const float rel_speed = len3(rel);
            if (rel_speed != 0.f) time = dot3(rel / rel_speed, pos - bp) / rel_speed;
            if (time >= 0.f && time < min_t) {
                const f3 my_near = pos + dir * (speed * time);
                const f3 his_near = bp + bd * (bs * time);
                last_near = his_near;
                if (len3(my_near - his_near) < threshold) {
                    min_t = time;
                    found = true;
                    t_dir = bd;
                    t_speed = bs;
                    t_pos = bp;
                    threat_near = his_near;
                }
            }
        }
        float steer = 0.f;
        if (found) {
            const float par = dot3(dir, t_dir);
            if (par < -0.707f) {
                const f3 off = (MODE == GS_MODE_GPUREF ? last_near : threat_near) - pos;
                steer = dot3(off, side) > 0.f ? -1.f : 1.f;
            } else if (par > 0.707f) {
                steer = dot3(t_pos - pos, side) > 0.f ? -1.f : 1.f;
            } else if (MODE == GS_MODE_GPUREF ? (speed <= t_speed) : (t_speed <= speed)) {
                steer = dot3(side, t_dir * t_speed) > 0.f ? -1.f : 1.f;
            }
        }
        if (MODE == GS_MODE_GPUREF) {
            if (found) apply(acc, bits, normalize3(side * steer), w, GS_BIT_AVOID_NEIGHBORS);
        } else {
            apply(acc, bits, side * steer, w, GS_BIT_AVOID_NEIGHBORS);
        }
    }
    st3(A.gx, A.gy, A.gz, i, acc);
    A.bits[i] = bits;
}

void launch_avoid_neighbors(gs_group_s* A, gs_knn_s* knn, gs_group_s* B, float min_time, float min_sep,
                            int avoid_close, float w, uint32_t mask, int mode) {
    if (A->n == 0) return;
    hipStream_t st = A->ctx->stream;
#define GS_AN(M)                                                                                                  \
    k_avoid_neighbors<M><<<blocks_for(A->n), kBlock, 0, st>>>(A->n, A->dev_agents(), B->n, B->F(0), B->F(1),     \
                                                              B->F(2), B->F(3), B->F(4), B->F(5), B->F(6),        \
                                                              B->F(19), knn->idx(), knn->dist(), knn->k, min_time, \
                                                              min_sep, avoid_close, w, mask)
    if (mode == GS_MODE_GPUREF) GS_AN(GS_MODE_GPUREF);
    else GS_AN(GS_MODE_OPENSTEER);
#undef GS_AN
    GS_LAUNCHED("k_avoid_neighbors");
    A->mark_written(kFSteer, true);
}

void launch_update_anti_penetration(gs_group_s* A, gs_knn_s* knn, gs_group_s* W, float dt, int closest) {
    if (A->n == 0) return;
    const int f0 = kWallField;
    gs_walls wd{W->F(f0), W->F(f0 + 1), W->F(f0 + 2), W->F(f0 + 3), W->F(f0 + 4),
                W->F(f0 + 5), W->F(f0 + 6), W->F(f0 + 7), W->F(f0 + 8)};
    k_update_anti_penetration<<<blocks_for(A->n), kBlock, 0, A->ctx->stream>>>(A->n, A->dev_agents(), W->n, wd,
                                                                                knn->idx(), knn->k, dt, closest);
    GS_LAUNCHED("k_update_anti_penetration");
    // kinematic fields rewritten, steering zeroed, bits kept (some set)
    A->mark_integrated();
    A->bits_zero = false;
}

void launch_anti_penetration_agents(gs_group_s* A, gs_knn_s* knn, gs_group_s* B, uint32_t mask) {
    if (A->n == 0) return;
    k_anti_penetration_agents<<<blocks_for(A->n), kBlock, 0, A->ctx->stream>>>(
        A->n, A->F(0), A->F(1), A->F(2), A->F(19), A->bits(), A->Falt(0), A->Falt(1), A->Falt(2),
        A->bitsb[A->bcur ^ 1].p, knn->idx(), knn->dist(), B->n, B->F(0), B->F(1), B->F(2), B->F(19), B->bits(),
        mask);
    GS_LAUNCHED("k_anti_penetration_agents");
    for (int j = 0; j < 3; ++j) A->flip(j);
    A->bcur ^= 1;
    A->mark_written(kFPos, true);
}

void launch_follow_path(gs_group_s* A, const gs_path_s* p, float prediction_time, float w, uint32_t mask) {
    if (A->n == 0) return;
    const float* d = p->dev.p;
    const uint32_t m = p->m;
    PathDev P{d, d + m, d + 2 * m, d + 3 * m, d + 4 * m, d + 5 * m, d + 6 * m, m, p->radius, p->total, p->cyclic};
    k_follow_path<<<blocks_for(A->n), kBlock, 0, A->ctx->stream>>>(A->n, A->dev_agents(), P, prediction_time, w,
                                                                    mask);
    GS_LAUNCHED("k_follow_path");
    A->mark_written(kFSteer, true);
}

}  // namespace gs
