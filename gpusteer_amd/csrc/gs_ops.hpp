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
// min_d < 0); GPUREF = FlockingCommon.cuh:4-30 (sqrt distances).
template <int MODE>
__device__ __forceinline__ bool in_hood(f3 offset, float d2, f3 fwd, float radius, const FlockTerm& t) {
    if (MODE == GS_MODE_GPUREF) {
        const float d = sqrtf(d2);
        if (d < t.min_d) return true;
        if (d > t.max_d) return false;
        return dot3(fwd, offset / d) > t.cos_max;
    }
    const float md = (t.min_d < 0.f) ? radius * 3.f : t.min_d;
    if (d2 < (md * md)) return true;
    if (d2 > (t.max_d * t.max_d)) return false;
    return dot3(fwd, offset / sqrtf(d2)) > t.cos_max;
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
// consecutive launches would.  nbr(j, want_fwd, bp, bf) fetches neighbour
// slot j and returns kNbrValid, kNbrStop at the first pad, or kNbrSkip for a
// slot to pass over.  NSLOT == 0 walks j = 0..k-1; NSLOT > 0 walks
// j = 0..NSLOT-1 fully unrolled (for register-resident neighbour lists).
enum { kNbrValid = 0, kNbrStop = 1, kNbrSkip = 2 };
template <int MODE, int NSLOT, class Nbr>
__device__ __forceinline__ void flock_core(f3 pos, f3 fwd, float radius, uint32_t k, const FlockTerm& sep,
                                           const FlockTerm& ali, const FlockTerm& coh, f3& steering,
                                           uint32_t& bits, Nbr nbr) {
    // Bits only accumulate, so a term masked now stays masked.
    const bool do_s = sep.on && !(bits & sep.mask);
    const bool do_a = ali.on && !(bits & ali.mask);
    const bool do_c = coh.on && !(bits & coh.mask);
    if (!do_s && !do_a && !do_c) return;
    f3 ss = mk3(0.f, 0.f, 0.f), sa = ss, sc = ss;
    int ns = 0, na = 0, nc = 0;
    auto visit = [&](f3 bp, f3 bf) {
        const f3 off = bp - pos;
        const float d2 = dot3(off, off);
        if (do_s && in_hood<MODE>(off, d2, fwd, radius, sep)) {
            ss = ss + off / -d2;
            ++ns;
        }
        if (do_a && in_hood<MODE>(off, d2, fwd, radius, ali)) {
            sa = sa + bf;
            ++na;
        }
        if (do_c && in_hood<MODE>(off, d2, fwd, radius, coh)) {
            sc = sc + bp;
            ++nc;
        }
    };
    if (NSLOT == 0) {
        for (uint32_t j = 0; j < k; ++j) {
            f3 bp, bf = mk3(0.f, 0.f, 0.f);
            if (nbr((int)j, do_a, bp, bf) != kNbrValid) break;
            visit(bp, bf);
        }
    } else {
        bool stopped = false;
#pragma unroll
        for (int j = 0; j < (NSLOT > 0 ? NSLOT : 1); ++j) {
            f3 bp = mk3(0.f, 0.f, 0.f), bf = bp;
            if (!stopped) {
                const int r = nbr(j, do_a, bp, bf);
                if (r == kNbrStop) stopped = true;
                else if (r == kNbrValid) visit(bp, bf);
            }
        }
    }
    if (do_s) {
        if (ns > 0) ss = normalize3(ss / (float)ns);
        apply(steering, bits, ss, sep.w, GS_BIT_SEPARATION);
    }
    if (do_a && !(bits & ali.mask)) {
        if (na > 0) sa = normalize3((sa / (float)na) - fwd);
        apply(steering, bits, sa, ali.w, GS_BIT_ALIGNMENT);
    }
    if (do_c && !(bits & coh.mask)) {
        if (nc > 0) sc = normalize3((sc / (float)nc) - pos);
        apply(steering, bits, sc, coh.w, GS_BIT_COHESION);
    }
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
        const f3 acc = force / mass;
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
        const f3 new_acc = truncate3(adjusted, max_force) / mass;
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
