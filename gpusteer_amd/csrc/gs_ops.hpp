// Per-agent device operations shared by the behaviour kernels and the fused
// step kernel.  Arithmetic follows the reference's evaluation order.
#pragma once
#include "gs_internal.hpp"

namespace gs {

struct FlockTerm {
    float min_d, max_d, cos_max, w;
    uint32_t mask;
    int on;
};

// Boid neighbourhood predicate: OPENSTEER = SteerLibrary::inBoidNeighborhood
// (src/SteerLibrary.cpp:506-544, squared distances, min = 3*radius when
// min_d < 0); GPUREF = FlockingCommon.cuh:4-30 (sqrt distances).  Both test
// the cone with dot(fwd, offset / sqrtf(d2)); the caller passes d = sqrtf(d2)
// and that dot product, which do not depend on the term, so a neighbour tested
// by several terms computes them once (same values, same results).
template <int MODE>
__device__ __forceinline__ bool in_hood_pre(float d2, float d, float cdot, float radius, const FlockTerm& t) {
    if (MODE == GS_MODE_GPUREF) {
        if (d < t.min_d) return true;
        if (d > t.max_d) return false;
        return cdot > t.cos_max;
    }
    const float md = (t.min_d < 0.f) ? radius * 3.f : t.min_d;
    if (d2 < (md * md)) return true;
    if (d2 > (t.max_d * t.max_d)) return false;
    return cdot > t.cos_max;
}

__device__ __forceinline__ void apply(f3& steering, uint32_t& bits, f3 f, float w, uint32_t bit) {
    const f3 fw = f * w;
    if (!zero3(fw)) bits |= bit;
    steering = fw + steering;
}

// Separation (src/SteerLibrary.cpp:552-585), alignment (:593-623) and cohesion
// (:632-662) over the KNN slots in slot order, stopping at the first pad
// (SteerForSeparationCUDAKernel.cu:98-121).  The three sums are independent
// of each other, so they are gathered in one pass over the neighbours and then
// applied in the order separation, alignment, cohesion with the masks tested
// against the bits as each previous term left them, exactly as three
// consecutive launches would.  The pass is split into flock_begin,
// flock_accumulate (any number of consecutive chunks of neighbours, in slot
// order) and flock_finish, so a long list can be gathered a chunk at a time.
struct FlockAcc {
    f3 ss, sa, sc;
    int ns, na, nc;
    bool do_s, do_a, do_c;
    bool fok;
    float cone_e;
};

__device__ __forceinline__ FlockAcc flock_begin(f3 fwd, uint32_t bits, const FlockTerm& sep, const FlockTerm& ali,
                                                const FlockTerm& coh) {
    FlockAcc a;
    // Bits only accumulate, so a term masked now stays masked.
    a.do_s = sep.on && !(bits & sep.mask);
    a.do_a = ali.on && !(bits & ali.mask);
    a.do_c = coh.on && !(bits & coh.mask);
    // The cone input dot3(fwd, off / d) only enters the comparisons
    // `> cos_max` of the three terms.  Its estimate fma-dot(fwd, off) times
    // the hardware reciprocal of d (GPU rule) or reciprocal square root of
    // d2 (OpenSteer rule, whose predicate compares squared distances, so d
    // itself is needed only for the exact value), each within 1 ulp, lies
    // within 12 u |fwd| of it (u = 2^-24: three roundings of the estimate's
    // dot, the reciprocal (with d's rounding) and its product, against four
    // on each term of the exact form, with |off| / d <= 1 + 3u).  An
    // estimate farther than E = 2^-18 (|fx| + |fy| + |fz|) + 2^-100 from
    // every threshold therefore decides all three comparisons as the exact
    // value does; otherwise, or for d2 outside [2^-80, 2^80] or a forward sum
    // above 2^40 (where underflow or overflow could void the bound; NaN fails
    // the tests too), the exact value is computed.
    const float fsum = (fabsf(fwd.x) + fabsf(fwd.y)) + fabsf(fwd.z);
    a.fok = fsum <= 0x1p40f;
    a.cone_e = __builtin_fmaf(fsum, 0x1p-18f, 0x1p-100f);
    a.ss = a.sa = a.sc = mk3(0.f, 0.f, 0.f);
    a.ns = a.na = a.nc = 0;
    return a;
}

__device__ __forceinline__ bool flock_any(const FlockAcc& a) { return a.do_s || a.do_a || a.do_c; }

// The next cnt neighbours (cnt <= CH) from the register arrays bp / bf.
template <int MODE, int CH>
__device__ __forceinline__ void flock_accumulate(FlockAcc& a, f3 pos, f3 fwd, float radius, int cnt, const f3* bp,
                                                 const f3* bf, const FlockTerm& sep, const FlockTerm& ali,
                                                 const FlockTerm& coh) {
#pragma unroll
    for (int j = 0; j < CH; ++j) {
        if (j < cnt) {
            const f3 off = bp[j] - pos;
            const float d2 = dot3(off, off);
            const bool gpuref = MODE == GS_MODE_GPUREF;
            float d = gpuref ? sqrtf(d2) : 0.f;
            float cdot = __builtin_fmaf(fwd.z, off.z, __builtin_fmaf(fwd.y, off.y, fwd.x * off.x)) *
                         (gpuref ? __builtin_amdgcn_rcpf(d) : __builtin_amdgcn_rsqf(d2));
            const bool sure = a.fok && d2 >= 0x1p-80f && d2 <= 0x1p80f && fabsf(cdot - sep.cos_max) > a.cone_e &&
                              fabsf(cdot - ali.cos_max) > a.cone_e && fabsf(cdot - coh.cos_max) > a.cone_e;
            if (!sure) {
                if (!gpuref) d = sqrtf(d2);
                cdot = dot3(fwd, off / d);
            }
            if (a.do_s && in_hood_pre<MODE>(d2, d, cdot, radius, sep)) {
                a.ss = a.ss + off / -d2;
                ++a.ns;
            }
            if (a.do_a && in_hood_pre<MODE>(d2, d, cdot, radius, ali)) {
                a.sa = a.sa + bf[j];
                ++a.na;
            }
            if (a.do_c && in_hood_pre<MODE>(d2, d, cdot, radius, coh)) {
                a.sc = a.sc + bp[j];
                ++a.nc;
            }
        }
    }
}

__device__ __forceinline__ void flock_finish(const FlockAcc& a, f3 pos, f3 fwd, const FlockTerm& sep,
                                             const FlockTerm& ali, const FlockTerm& coh, f3& steering,
                                             uint32_t& bits) {
    if (a.do_s) {
        f3 ss = a.ss;
        if (a.ns > 0) ss = normalize3(ss / (float)a.ns);
        apply(steering, bits, ss, sep.w, GS_BIT_SEPARATION);
    }
    if (a.do_a && !(bits & ali.mask)) {
        f3 sa = a.sa;
        if (a.na > 0) sa = normalize3((sa / (float)a.na) - fwd);
        apply(steering, bits, sa, ali.w, GS_BIT_ALIGNMENT);
    }
    if (a.do_c && !(bits & coh.mask)) {
        f3 sc = a.sc;
        if (a.nc > 0) sc = normalize3((sc / (float)a.nc) - pos);
        apply(steering, bits, sc, coh.w, GS_BIT_COHESION);
    }
}

// The whole pass over the first m neighbours: bp / bf hold their positions
// and forward vectors (register arrays of KMAX, walked fully unrolled).
template <int MODE, int KMAX>
__device__ __forceinline__ void flock_core(f3 pos, f3 fwd, float radius, int m, const f3* bp, const f3* bf,
                                           const FlockTerm& sep, const FlockTerm& ali, const FlockTerm& coh,
                                           f3& steering, uint32_t& bits) {
    FlockAcc a = flock_begin(fwd, bits, sep, ali, coh);
    if (!flock_any(a)) return;
    flock_accumulate<MODE, KMAX>(a, pos, fwd, radius, m, bp, bf, sep, ali, coh);
    flock_finish(a, pos, fwd, sep, ali, coh, steering, bits);
}

// Loads the KNN slots of row `row` ([k][nA] slot-major) and the neighbours'
// positions (and forward vectors when want_fwd) into register arrays, every
// level of loads in flight together.  Returns m, the number of neighbours
// before the first pad (>= nB).  Positions come from the packed snapshot when
// bpos is given, else from the SoA arrays bx/by/bz; forward vectors from the
// packed (forward, speed) snapshot when bfwd is given, else from bfx/bfy/bfz.
template <int KMAX>
__device__ __forceinline__ int gather_neighbours(const uint32_t* __restrict__ knn, uint32_t k, uint32_t nA,
                                                 uint32_t row, uint32_t nB, const float4* __restrict__ bpos,
                                                 const float* __restrict__ bx, const float* __restrict__ by,
                                                 const float* __restrict__ bz, const float* __restrict__ bfx,
                                                 const float* __restrict__ bfy, const float* __restrict__ bfz,
                                                 bool want_fwd, f3* bp, f3* bf,
                                                 const float4* __restrict__ bfwd = nullptr) {
    // Loads are unconditional at in-range indices (slots past the first pad
    // reuse slot 0's neighbour) so each level is in flight together; only
    // the first m results are used.
    uint32_t v[KMAX];
#pragma unroll
    for (int j = 0; j < KMAX; ++j) v[j] = knn[(size_t)min(j, (int)k - 1) * nA + row];
    int m = (int)k;
#pragma unroll
    for (int j = KMAX - 1; j >= 0; --j)
        if (j < (int)k && v[j] >= nB) m = j;
#pragma unroll
    for (int j = 0; j < KMAX; ++j) bp[j] = bf[j] = mk3(0.f, 0.f, 0.f);
    if (m == 0) return 0;
#pragma unroll
    for (int j = 0; j < KMAX; ++j) {
        const uint32_t b = (j < m) ? v[j] : v[0];
        if (bpos) {
            const float4 q = bpos[b];
            bp[j] = mk3(q.x, q.y, q.z);
        } else {
            bp[j] = ld3(bx, by, bz, b);
        }
    }
    if (want_fwd) {
#pragma unroll
        for (int j = 0; j < KMAX; ++j) {
            const uint32_t b = (j < m) ? v[j] : v[0];
            if (bfwd) {
                const float4 f = bfwd[b];
                bf[j] = mk3(f.x, f.y, f.z);
            } else {
                bf[j] = ld3(bfx, bfy, bfz, b);
            }
        }
    }
    return m;
}

// LinesIntersect in the XZ plane, include/OpenSteer/VectorUtils.cuh:390-428.
__device__ __forceinline__ bool lines_intersect(f3 sa, f3 ea, f3 sb, f3 eb, f3& ip) {
    const float eps = 0.001f;
    const float denom = ((eb.z - sb.z) * (ea.x - sa.x)) - ((eb.x - sb.x) * (ea.z - sa.z));
    const float numera = ((eb.x - sb.x) * (sa.z - sb.z)) - ((eb.z - sb.z) * (sa.x - sb.x));
    const float numerb = ((ea.x - sa.x) * (sa.z - sb.z)) - ((ea.z - sa.z) * (sa.x - sb.x));
    if (fabsf(denom) < eps) {
        if (fabsf(numera) < eps && fabsf(numerb) < eps) {
            ip = sa;
            return true;
        }
        return false;
    }
    const float ua = numera / denom;
    const float ub = numerb / denom;
    if (ua >= 0.f && ua <= 1.f && ub >= 0.f && ub <= 1.f) {
        ip = mk3(sa.x + ua * (ea.x - sa.x), 0.f, sa.z + ua * (ea.z - sa.z));
        return true;
    }
    return false;
}

// Obstacle avoidance of one agent against the first k entries of its KNN
// column (knn[j * stride], stopping at the first index >= nO):
// SteerToAvoidObstaclesKernel + findNextIntersectionWithSphere
// (CUDA/AvoidObstaclesCUDAKernel.cu:47-99,136-253).  Applies the result to
// steering / bits with the accumulate convention.
__device__ __forceinline__ void avoid_obstacles_one(f3 pos, f3 fwd, f3 side, f3 up, float radius, float speed,
                                                    const uint32_t* __restrict__ knn, uint32_t stride, uint32_t k,
                                                    uint32_t nO, const float* __restrict__ ox,
                                                    const float* __restrict__ oy, const float* __restrict__ oz,
                                                    const float* __restrict__ orad, float min_time, float w,
                                                    f3& steering, uint32_t& bits) {
    const float min_dist = min_time * speed;
    bool n_hit = false;
    float n_dist = 0.f;
    f3 n_pos = mk3(0.f, 0.f, 0.f);
    for (uint32_t j = 0; j < k; ++j) {
        const uint32_t o = knn[(size_t)j * stride];
        if (o >= nO) break;
        const f3 op = ld3(ox, oy, oz, o);
        const float orr = orad[o];
        bool x_hit = false;
        float x_dist = 0.f;
        const f3 go = op - pos;
        const f3 lc = mk3(dot3(go, side), dot3(go, up), dot3(go, fwd));
        const float b = -2.f * lc.z;
        const float c = ((lc.x * lc.x + lc.y * lc.y) + lc.z * lc.z) - (orr + radius) * (orr + radius);
        const float d = (b * b) - (4.f * c);
        if (!(d < 0.f)) {
            const float s = sqrtf(d);
            const float p = (-b + s) / 2.f;
            const float q = (-b - s) / 2.f;
            if (!((p < 0.f) && (q < 0.f))) {
                x_hit = true;
                x_dist = ((p > 0.f) && (q > 0.f)) ? ((p < q) ? p : q) : ((p > 0.f) ? p : q);
            }
        }
        if ((n_hit == false) || ((x_hit != false) && (x_dist < n_dist))) {
            n_hit = x_hit;
            n_dist = x_dist;
            n_pos = op;
        }
    }
    f3 f = mk3(0.f, 0.f, 0.f);
    if (n_hit && (n_dist < min_dist)) f = perp3(pos - n_pos, fwd);
    apply(steering, bits, f, w, GS_BIT_AVOID_OBSTACLES);
}

// Wall avoidance of one agent against the first k entries of its KNN column:
// SteerToAvoidWallsCUDAKernel (CUDA/AvoidWallsCUDAKernel.cu:67-245) — three
// feelers, the touching test, the anti-penetration move of a touching agent
// (:215-223, updates pos / dir) and its bit rules.  steering += w * force.
__device__ __forceinline__ void avoid_walls_one(f3& pos, f3& dir, f3 side, float radius, float speed,
                                                const uint32_t* __restrict__ knn, uint32_t stride, uint32_t k,
                                                uint32_t nW, const gs_walls& W, float min_time, float w,
                                                f3& steering_acc, uint32_t& bits) {
    const float fl = 4.f * radius;  // FEELER_LENGTH (:41) * radius
    const f3 vel = dir * speed;
    f3 feelers[3];
    feelers[0] = pos + vel * min_time;
    feelers[1] = pos + (side * fl + dir * fl);
    feelers[2] = pos + ((-side) * fl + dir * fl);
    int feeler = 0;
    float nearest_dist = FLT_MAX;
    uint32_t nearest_line = UINT32_MAX;
    f3 nearest_ip = mk3(0.f, 0.f, 0.f), nearest_normal = nearest_ip;
    bool touching = false;
    for (uint32_t j = 0; j < k; ++j) {
        const uint32_t l = knn[(size_t)j * stride];
        if (l >= nW) break;
        const f3 ln = ld3(W.nx, W.ny, W.nz, l);
        const f3 ls = ld3(W.sx, W.sy, W.sz, l);
        const f3 le = ld3(W.ex, W.ey, W.ez, l);
        f3 ip;
        if (lines_intersect(pos, pos + (-ln) * radius, ls, le, ip)) {
            touching = true;
            nearest_normal = ln;
            nearest_dist = len3(pos - ip);
        }
#pragma unroll
        for (int f = 0; f < 3; ++f) {
            if (lines_intersect(pos, feelers[f], ls, le, ip)) {
                const float dist = len3(pos - ip);
                if (dist < nearest_dist) {
                    nearest_dist = dist;
                    nearest_line = l;
                    nearest_ip = ip;
                    feeler = f;
                    nearest_normal = ln;
                }
            }
        }
    }
    f3 steering = mk3(0.f, 0.f, 0.f);
    if (touching) {
        const float pen = radius - nearest_dist;
        pos = pos + nearest_normal * pen;
        dir = normalize3(perp3(dir, nearest_normal));
        bits |= GS_BIT_ANTI_PENETRATION_WALL;
    } else if (nearest_line != UINT32_MAX) {
        const f3 fe = feeler == 0 ? feelers[0] : (feeler == 1 ? feelers[1] : feelers[2]);
        steering = nearest_normal * len3(fe - nearest_ip);
        bits |= GS_BIT_AVOID_WALLS;
    }
    steering = steering * w;
    steering_acc = steering + steering_acc;
}

// limitMaxDeviationAngle, include/OpenSteer/VectorUtils.cuh:305-364.
__device__ __forceinline__ f3 limit_max_deviation(f3 source, float cos_cone, f3 basis) {
    const float source_len = len3(source);
    if (source_len == 0.f) return source;
    const f3 direction = source / source_len;
    if (dot3(direction, basis) >= cos_cone) return source;
    const f3 unit_perp = normalize3(perp3(source, basis));
    const float perp_dist = sqrtf(1.f - (cos_cone * cos_cone));
    return ((basis * cos_cone) + (unit_perp * perp_dist)) * source_len;
}

// pow(range, 20) as double-precision repeated squaring narrowed to float.
__device__ __forceinline__ float pow20(float range) {
    const double r = (double)range;
    const double r2 = r * r;
    const double r4 = r2 * r2;
    const double r5 = r4 * r;
    const double r10 = r5 * r5;
    return (float)(r10 * r10);
}

__device__ __forceinline__ float clipf(float x, float lo, float hi) {
    if (x < lo) return lo;
    if (x > hi) return hi;
    return x;
}

// One integration step of agent i given its steering force.  Returns the new
// state through the references.
//   OPENSTEER: SimpleVehicle::applySteeringForce (src/SimpleVehicle.cpp:167-214),
//     adjustRawSteeringForce (:111-132), regenerateLocalSpace (:224-230),
//     regenerateOrthonormalBasisUF (include/OpenSteer/SimpleVehicle.h:190-217).
//   GPUREF: UpdateCUDAKernel (CUDA/UpdateCUDAKernel.cu:78-108).
template <int MODE>
__device__ __forceinline__ void integrate(f3 force, f3& pos, f3& fwd, float& speed, f3& sacc, f3& side, f3& up,
                                          float max_force, float max_speed, float mass, float dt) {
    f3 new_vel;
    if (MODE == GS_MODE_GPUREF) {
        force = truncate3(force, max_force);
        // x / 1 is x exactly (signed zeros included), so a unit mass skips
        // the three divisions.
        f3 acc = force;
        if (mass != 1.f) acc = force / mass;
        new_vel = truncate3(fwd * speed + acc * dt, max_speed);
        speed = len3(new_vel);
        if (speed > 0.f) {
            fwd = new_vel / speed;
            side = normalize3(cross3(fwd, up));
            up = cross3(side, fwd);
        }
        pos = pos + new_vel * dt;
    } else {
        f3 adjusted = force;
        const float max_adj = 0.2f * max_speed;
        if (!((speed > max_adj) || zero3(force))) {
            const float range = speed / max_adj;
            const float cosine = 1.f + (-1.f - 1.f) * pow20(range);
            adjusted = limit_max_deviation(force, cosine, fwd);
        }
        f3 new_acc = truncate3(adjusted, max_force);
        if (mass != 1.f) new_acc = new_acc / mass;
        if (dt > 0.f) {
            const float rate = clipf(clipf(9.f * dt, 0.15f, 0.4f), 0.f, 1.f);
            sacc = sacc + (new_acc - sacc) * rate;
        }
        new_vel = truncate3(fwd * speed + sacc * dt, max_speed);
        speed = len3(new_vel);
        pos = pos + new_vel * dt;
        if (speed > 0.f) {
            fwd = new_vel / speed;
            side = cross3(fwd, up);
            up = cross3(side, fwd);
        }
    }
}

}  // namespace gs
