// Exact KNN building blocks shared by the stand-alone KNN kernel and the
// fused step kernel: register top-k list, lattice distance bounds, seeded
// shell search.
#pragma once
#include "gs_internal.hpp"

namespace gs {

template <int KMAX> struct TopK {
    float d[KMAX];
    uint32_t i[KMAX];
    __device__ __forceinline__ void init(uint32_t k) {
#pragma unroll
        for (int j = 0; j < KMAX; ++j) {
            const bool sentinel = j < KMAX - (int)k;
            d[j] = sentinel ? -1.f : INFINITY;
            i[j] = sentinel ? 0u : UINT32_MAX;
        }
    }
    __device__ __forceinline__ float worst() const { return d[KMAX - 1]; }
    __device__ __forceinline__ bool accepts(float dd, uint32_t ii) const {
        return dd < d[KMAX - 1] || (dd == d[KMAX - 1] && ii < i[KMAX - 1]);
    }
    // Caller guarantees accepts(dd, ii).
    __device__ __forceinline__ void push(float dd, uint32_t ii) {
        d[KMAX - 1] = dd;
        i[KMAX - 1] = ii;
#pragma unroll
        for (int j = KMAX - 2; j >= 0; --j) {
            const bool sw = d[j + 1] < d[j] || (d[j + 1] == d[j] && i[j + 1] < i[j]);
            const float lo_d = sw ? d[j + 1] : d[j];
            const float hi_d = sw ? d[j] : d[j + 1];
            const uint32_t lo_i = sw ? i[j + 1] : i[j];
            const uint32_t hi_i = sw ? i[j] : i[j + 1];
            d[j] = lo_d;
            d[j + 1] = hi_d;
            i[j] = lo_i;
            i[j + 1] = hi_i;
        }
    }
    __device__ __forceinline__ void store(uint32_t k, uint32_t nA, uint32_t q, uint32_t* __restrict__ out_idx,
                                          float* __restrict__ out_dist) const {
#pragma unroll
        for (int j = 0; j < KMAX; ++j) {
            if (j >= KMAX - (int)k) {
                const size_t o = (size_t)(j - (KMAX - (int)k)) * nA + q;
                const bool pad = i[j] == UINT32_MAX;
                out_idx[o] = i[j];
                out_dist[o] = pad ? FLT_MAX : sqrtf(d[j]);
            }
        }
    }
};

// Distance from coordinate q to the slab of cells [i0, i1] along one axis;
// boundary cells extend to infinity outward (points outside the lattice are
// clamped into them).  The rounding slack makes the gap an under-estimate.
__device__ __forceinline__ float axis_gap(float q, int i0, int i1, float o, float h, int d, float slack) {
    float gap = 0.f;
    if (i0 > 0) gap = fmaxf(gap, (o + (float)i0 * h) - q);
    if (i1 < d - 1) gap = fmaxf(gap, q - (o + (float)(i1 + 1) * h));
    return fmaxf(gap - slack, 0.f);
}

// Largest d2 from the query to the k previous neighbours, measured on the
// same sorted snapshot the search scans (binv = sorted slot of each point),
// or +inf when the seed is incomplete.
__device__ __forceinline__ float seed_bound(const uint32_t* __restrict__ seed, uint32_t k, uint32_t nA, uint32_t q,
                                            uint32_t nB, float qx, float qy, float qz,
                                            const uint32_t* __restrict__ binv, const float* __restrict__ bspx,
                                            const float* __restrict__ bspy, const float* __restrict__ bspz) {
    float m = 0.f;
    for (uint32_t j = 0; j < k; ++j) {
        const uint32_t b = seed[(size_t)j * nA + q];
        if (b >= nB) return INFINITY;
        const uint32_t t = binv[b];
        m = fmaxf(m, sqdist(qx, qy, qz, bspx[t], bspy[t], bspz[t]));
    }
    return m;
}

// Scans the lattice around (qx,qy,qz) until the top-k is exact.  self_slot is
// the query's own sorted slot when A == B (excluded), else UINT32_MAX.
// visit(t) is called for every candidate sorted slot t with d2 <= bound.
template <int KMAX>
__device__ __forceinline__ void binned_search(TopK<KMAX>& h, float thr, float qx, float qy, float qz,
                                              uint32_t self_slot, const GridDev& g, int R0,
                                              const float* __restrict__ bspx, const float* __restrict__ bspy,
                                              const float* __restrict__ bspz, const uint32_t* __restrict__ bperm,
                                              const uint32_t* __restrict__ bstart) {
    const int cx = axis_cell(qx, g.ox, g.sx, g.dx);
    const int cy = axis_cell(qy, g.oy, g.sy, g.dy);
    const int cz = axis_cell(qz, g.oz, g.sz, g.dz);
    const float grow = 1.0f + 2e-5f;
    for (int R = 0;; ++R) {
        const int x0 = max(cx - R, 0), x1 = min(cx + R, g.dx - 1);
        const int y0 = max(cy - R, 0), y1 = min(cy + R, g.dy - 1);
        for (int ix = x0; ix <= x1; ++ix) {
            const bool xedge = (ix == cx - R) || (ix == cx + R);
            const float gx = axis_gap(qx, ix, ix, g.ox, g.hx, g.dx, g.slack);
            for (int iy = y0; iy <= y1; ++iy) {
                const bool edge = xedge || (iy == cy - R) || (iy == cy + R);
                const float gy = axis_gap(qy, iy, iy, g.oy, g.hy, g.dy, g.slack);
                const float gxy = gx * gx + gy * gy;
                // z-runs of this column inside shell R: the whole [cz-R, cz+R]
                // on an edge column, else its two end cells.
                for (int run = 0; run < (edge ? 1 : 2); ++run) {
                    int z0, z1;
                    if (edge) {
                        z0 = max(cz - R, 0);
                        z1 = min(cz + R, g.dz - 1);
                    } else {
                        z0 = z1 = (run == 0) ? cz - R : cz + R;
                        if (z0 < 0 || z0 >= g.dz) continue;
                    }
                    const float gz = axis_gap(qz, z0, z1, g.oz, g.hz, g.dz, g.slack);
                    float lim = fminf(thr, h.worst());
                    if (gxy + gz * gz > lim * grow) continue;
                    const uint32_t c0 = cell_index(g, ix, iy, z0);
                    const uint32_t t0 = bstart[c0];
                    const uint32_t t1 = bstart[c0 + (uint32_t)(z1 - z0) + 1u];
                    for (uint32_t t = t0; t < t1; ++t) {
                        if (t == self_slot) continue;
                        const float d2 = sqdist(qx, qy, qz, bspx[t], bspy[t], bspz[t]);
                        if (d2 <= lim) {
                            const uint32_t b = bperm[t];
                            if (h.accepts(d2, b)) {
                                h.push(d2, b);
                                lim = fminf(thr, h.worst());
                            }
                        }
                    }
                }
            }
        }
        if (R < R0) continue;
        // Stop when every unscanned cell is farther than the k-th best.
        float dmin = INFINITY;
        bool all = true;
        if (cx - R > 0) { dmin = fminf(dmin, qx - (g.ox + (float)(cx - R) * g.hx)); all = false; }
        if (cx + R < g.dx - 1) { dmin = fminf(dmin, (g.ox + (float)(cx + R + 1) * g.hx) - qx); all = false; }
        if (cy - R > 0) { dmin = fminf(dmin, qy - (g.oy + (float)(cy - R) * g.hy)); all = false; }
        if (cy + R < g.dy - 1) { dmin = fminf(dmin, (g.oy + (float)(cy + R + 1) * g.hy) - qy); all = false; }
        if (cz - R > 0) { dmin = fminf(dmin, qz - (g.oz + (float)(cz - R) * g.hz)); all = false; }
        if (cz + R < g.dz - 1) { dmin = fminf(dmin, (g.oz + (float)(cz + R + 1) * g.hz) - qz); all = false; }
        if (all) break;
        dmin -= g.slack;
        const float lim = fminf(thr, h.worst());
        if (dmin > 0.f && lim < dmin * dmin * (1.0f - 2e-5f)) break;
    }
}

}  // namespace gs
