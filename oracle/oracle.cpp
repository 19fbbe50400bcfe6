// CPU oracle for the GPUSteer per-step simulation path.
//
// TEST INFRASTRUCTURE ONLY.  This file is a plain C++ restatement of the
// reference's CPU path (OpenSteer lq.c binning, brute-force proximity,
// SteerLibrary forces, SimpleVehicle integration) and of the GPU-only kernel
// semantics that have no CPU counterpart (wall and obstacle avoidance, the
// GPU-reference update, the face wrap).  It exists so that tests/, the smoke
// check and bench.py's cpu_baseline leg can check and time the HIP path.  The
// product (libgpusteer.so) never links, loads or calls it.
//
// Every routine names the reference file:line it restates.  Arithmetic is
// written operation-for-operation in the reference's evaluation order and the
// file is compiled with -ffp-contract=off, so that every fp32 result is the
// IEEE result of the same sequence of correctly rounded operations the GPU
// kernels perform.
//
// Pins (see tests/test_oracle_*.py): closed forms for each force, brute force
// versus binned KNN on constructed ties, the real lq.c (compiled from the
// reference sources into oracle/_ref) for cell membership and radius queries,
// std::pow for the x^20 restatement, exact antipode / face mappings.
//
// Templated on the scalar type so that the same restatement runs in float
// (parity and CPU baseline) and double (drift bound and gradient checks).

#include <algorithm>
#include <cfloat>
#include <climits>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <type_traits>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

// ---------------------------------------------------------------------------
// Vector algebra: include/OpenSteer/VectorUtils.cuh:35-189.
template <class T> struct V3 { T x, y, z; };
template <class T> static inline V3<T> mk(T x, T y, T z) { return V3<T>{x, y, z}; }
template <class T> static inline V3<T> add(V3<T> a, V3<T> b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
template <class T> static inline V3<T> sub(V3<T> a, V3<T> b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
template <class T> static inline V3<T> neg(V3<T> a) { return mk(-a.x, -a.y, -a.z); }
template <class T> static inline V3<T> mul(V3<T> a, T s) { return mk(a.x * s, a.y * s, a.z * s); }
template <class T> static inline V3<T> dvd(V3<T> a, T s) { return mk(a.x / s, a.y / s, a.z / s); }
// VectorUtils.cuh:86-89: (x*x' + y*y') + z*z', left to right.
template <class T> static inline T dot(V3<T> a, V3<T> b) { return ((a.x * b.x) + (a.y * b.y)) + (a.z * b.z); }
// Square root and absolute value through overloadable names, so that the
// calibration model also runs on the tangent type below (float and double
// call std::sqrt / std::fabs unchanged).
template <class T> static inline T osqrt(T x) { return std::sqrt(x); }
template <class T> static inline T ofabs(T x) { return std::fabs(x); }
template <class T> static inline T len(V3<T> a) { return osqrt(dot(a, a)); }
// VectorUtils.cuh:115-120: normalize skips the divide when the length is 0.
template <class T> static inline V3<T> normalize(V3<T> a) {
    const T l = len(a);
    return (l > T(0)) ? dvd(a, l) : a;
}
// VectorUtils.cuh:131-136.
template <class T> static inline V3<T> cross(V3<T> a, V3<T> b) {
    return mk((a.y * b.z) - (a.z * b.y), (a.z * b.x) - (a.x * b.z), (a.x * b.y) - (a.y * b.x));
}
template <class T> static inline bool is_zero(V3<T> a) { return a.x == T(0) && a.y == T(0) && a.z == T(0); }
template <class T> static inline T distance(V3<T> a, V3<T> b) { return len(sub(a, b)); }
// VectorUtils.cuh:163-175.
template <class T> static inline V3<T> parallel_component(V3<T> v, V3<T> unit) { return mul(unit, dot(v, unit)); }
template <class T> static inline V3<T> perpendicular_component(V3<T> v, V3<T> unit) {
    return sub(v, parallel_component(v, unit));
}
// VectorUtils.cuh:181-189.
template <class T> static inline V3<T> truncate_length(V3<T> v, T max_len) {
    const T max2 = max_len * max_len;
    const T l2 = dot(v, v);
    if (l2 <= max2) return v;
    return mul(v, max_len / std::sqrt(l2));
}
// VectorUtils.cuh:305-344 (insideOrOutside == true: limitMaxDeviationAngle).
template <class T> static inline V3<T> limit_max_deviation_angle(V3<T> source, T cos_cone, V3<T> basis) {
    const T source_len = len(source);
    if (source_len == T(0)) return source;
    const V3<T> direction = dvd(source, source_len);
    const T cos_source = dot(direction, basis);
    if (cos_source >= cos_cone) return source;
    const V3<T> perp = perpendicular_component(source, basis);
    const V3<T> unit_perp = normalize(perp);
    const T perp_dist = std::sqrt(T(1) - (cos_cone * cos_cone));
    const V3<T> c0 = mul(basis, cos_cone);
    const V3<T> c1 = mul(unit_perp, perp_dist);
    return mul(add(c0, c1), source_len);
}
// Utilities.h:97 clip.
template <class T> static inline T clip(T x, T lo, T hi) {
    if (x < lo) return lo;
    if (x > hi) return hi;
    return x;
}
// Utilities.h:203-210 blendIntoAccumulator with VectorUtils.cuh:275-278 interpolate.
template <class T> static inline V3<T> blend(T rate, V3<T> new_value, V3<T> acc) {
    const T a = clip(rate, T(0), T(1));
    return add(acc, mul(sub(new_value, acc), a));
}
// pow(range, 20) of SimpleVehicle.cpp:129 (a float argument promoted to double
// by C++ overloading, the result narrowed back to float by interpolate's alpha
// parameter).  Restated as double-precision repeated squaring so that CPU and
// GPU produce the same bits: r^20 = ((r^2)^2 * r)^2)^2.  Pinned against
// std::pow in tests/test_oracle_math.py.
template <class T> static inline T pow20(T range) {
    const double r = double(range);
    const double r2 = r * r;
    const double r4 = r2 * r2;
    const double r5 = r4 * r;
    const double r10 = r5 * r5;
    return T(r10 * r10);
}

// LinesIntersect, include/OpenSteer/VectorUtils.cuh:390-428 (XZ plane).
template <class T> static inline bool lines_intersect(V3<T> sa, V3<T> ea, V3<T> sb, V3<T> eb, V3<T>& ip) {
    const T eps = T(0.001f);  // VectorUtils.cuh:8 EPSILON
    const T denom = ((eb.z - sb.z) * (ea.x - sa.x)) - ((eb.x - sb.x) * (ea.z - sa.z));
    const T numera = ((eb.x - sb.x) * (sa.z - sb.z)) - ((eb.z - sb.z) * (sa.x - sb.x));
    const T numerb = ((ea.x - sa.x) * (sa.z - sb.z)) - ((ea.z - sa.z) * (sa.x - sb.x));
    if (std::fabs(denom) < eps) {
        if (std::fabs(numera) < eps && std::fabs(numerb) < eps) {
            ip = sa;  // coincident
            return true;
        }
        return false;  // parallel
    }
    const T ua = numera / denom;
    const T ub = numerb / denom;
    if (ua >= T(0) && ua <= T(1) && ub >= T(0) && ub <= T(1)) {
        ip.x = sa.x + ua * (ea.x - sa.x);
        ip.y = T(0);
        ip.z = sa.z + ua * (ea.z - sa.z);
        return true;
    }
    return false;
}

// ---------------------------------------------------------------------------
// Bin lattice: src/lq.c:185-208 (lqBinForLocation), index order :175-176
// (ix*divy*divz + iy*divz + iz).  lq.c puts points outside the super-brick in
// one "other" bin; here they are clamped into the boundary cell of each axis,
// which keeps every KNN query exact because the search bound below treats a
// boundary cell as extending to infinity outward.
struct Grid {
    float ox, oy, oz;  // super-brick minimum corner
    float sx, sy, sz;  // super-brick size
    int dx, dy, dz;    // divisions
};
static inline int axis_cell(float p, float o, float s, int d) {
    if (!(p >= o)) return 0;          // below (or NaN): lower boundary cell
    if (p >= o + s) return d - 1;     // at/after the far face: upper boundary cell
    int i = (int)(((p - o) / s) * (float)d);  // lq.c:198-200
    return i < d ? i : d - 1;
}
static inline uint32_t cell_of(const Grid& g, float x, float y, float z) {
    const int ix = axis_cell(x, g.ox, g.sx, g.dx);
    const int iy = axis_cell(y, g.oy, g.sy, g.dy);
    const int iz = axis_cell(z, g.oz, g.sz, g.dz);
    return (uint32_t)((ix * g.dy * g.dz) + (iy * g.dz) + iz);
}

// Stable counting sort of agents by cell: the reference builds the same
// structure with BuildDB + thrust::sort_by_key + ReorderDB
// (include/OpenSteer/CUDA/KNNBinningV2.cu:87-162).  start has C+1 entries;
// cell c holds sorted slots [start[c], start[c+1]) in ascending agent index.
static void bins_build(const Grid& g, int n, const float* px, const float* py, const float* pz,
                       uint32_t* key, uint32_t* start, uint32_t* perm, float* spx, float* spy, float* spz) {
    const int C = g.dx * g.dy * g.dz;
    std::vector<uint32_t> count((size_t)C + 1, 0);
    for (int i = 0; i < n; ++i) {
        key[i] = cell_of(g, px[i], py[i], pz[i]);
        count[key[i]]++;
    }
    uint32_t run = 0;
    for (int c = 0; c < C; ++c) {
        start[c] = run;
        run += count[c];
    }
    start[C] = run;
    std::vector<uint32_t> cursor(start, start + C);
    for (int i = 0; i < n; ++i) {
        const uint32_t s = cursor[key[i]]++;
        perm[s] = (uint32_t)i;
        if (spx) {
            spx[s] = px[i];
            spy[s] = py[i];
            spz[s] = pz[i];
        }
    }
}

// ---------------------------------------------------------------------------
// KNN.  Key (d2, index) ascending; d2 = (dx*dx + dy*dy) + dz*dz in fp32.
// Brute force follows KNNBruteForceCUDAKernelV2
// (include/OpenSteer/CUDA/KNNBruteForceCudaKernel.cu:207-283): ascending scan
// of B with strict '<' against the k-th entry, so the lower index wins ties.
// Output is slot-major [k][nA] (KNNBinningV2Kernel.cu:520-526), padded with
// UINT32_MAX / FLT_MAX (:451-457); dist = sqrtf(d2).
struct Best {
    float d2;
    uint32_t idx;
};
static inline bool key_less(float d2a, uint32_t ia, float d2b, uint32_t ib) {
    return d2a < d2b || (d2a == d2b && ia < ib);
}
static inline void insert_best(Best* h, int k, float d2, uint32_t idx) {
    if (!key_less(d2, idx, h[k - 1].d2, h[k - 1].idx)) return;
    int j = k - 1;
    while (j > 0 && key_less(d2, idx, h[j - 1].d2, h[j - 1].idx)) {
        h[j] = h[j - 1];
        --j;
    }
    h[j].d2 = d2;
    h[j].idx = idx;
}
static inline float d2_of(float ax, float ay, float az, float bx, float by, float bz) {
    const float dx = ax - bx, dy = ay - by, dz = az - bz;
    return ((dx * dx) + (dy * dy)) + (dz * dz);
}
static inline void write_knn(const Best* h, int k, int nA, int q, uint32_t* idx, float* dist) {
    for (int j = 0; j < k; ++j) {
        if (h[j].idx == UINT32_MAX) {
            idx[(size_t)j * nA + q] = UINT32_MAX;
            dist[(size_t)j * nA + q] = FLT_MAX;
        } else {
            idx[(size_t)j * nA + q] = h[j].idx;
            dist[(size_t)j * nA + q] = std::sqrt(h[j].d2);
        }
    }
}

// Optional radius cap: only points with d2 < cap*cap (fp32, strict) qualify,
// the locality test of src/lq.c:327; cap = +inf for none.
static void knn_brute(int nA, const float* ax, const float* ay, const float* az, int nB, const float* bx,
                      const float* by, const float* bz, int self, int k, uint32_t* idx, float* dist,
                      const uint32_t* queries, int nq, float cap) {
    // queries (optional): compute only these query indices (results still
    // written to column q of the [k][nA] arrays).
    const int total = queries ? nq : nA;
    const float r2 = cap * cap;
#pragma omp parallel for schedule(dynamic, 64)
    for (int t = 0; t < total; ++t) {
        const int q = queries ? (int)queries[t] : t;
        std::vector<Best> h(k, Best{FLT_MAX, UINT32_MAX});
        for (int b = 0; b < nB; ++b) {
            if (self && b == q) continue;
            const float d2 = d2_of(ax[q], ay[q], az[q], bx[b], by[b], bz[b]);
            if (d2 < r2) insert_best(h.data(), k, d2, (uint32_t)b);
        }
        write_knn(h.data(), k, nA, q, idx, dist);
    }
}

// Exact KNN over the bin lattice: scan shells of cells of growing Chebyshev
// radius around the query's cell until the k-th best d2 is below the squared
// distance from the query to every unscanned cell.  Replaces the reference's
// fixed-radius scan (KNNBinningV2Kernel.cu:470-515), which is not exact; the
// result equals knn_brute by construction and is tested to.
static void knn_binned(const Grid& g, int nA, const float* ax, const float* ay, const float* az, int nB,
                       const float* sbx, const float* sby, const float* sbz, const uint32_t* bperm,
                       const uint32_t* bstart, int self, int k, uint32_t* idx, float* dist, float cap) {
    const float hx = g.sx / (float)g.dx, hy = g.sy / (float)g.dy, hz = g.sz / (float)g.dz;
    const float r2 = cap * cap;
#pragma omp parallel for schedule(dynamic, 256)
    for (int q = 0; q < nA; ++q) {
        std::vector<Best> h(k, Best{FLT_MAX, UINT32_MAX});
        const float qx = ax[q], qy = ay[q], qz = az[q];
        const int cx = axis_cell(qx, g.ox, g.sx, g.dx);
        const int cy = axis_cell(qy, g.oy, g.sy, g.dy);
        const int cz = axis_cell(qz, g.oz, g.sz, g.dz);
        for (int R = 0;; ++R) {
            for (int ix = cx - R; ix <= cx + R; ++ix) {
                if (ix < 0 || ix >= g.dx) continue;
                for (int iy = cy - R; iy <= cy + R; ++iy) {
                    if (iy < 0 || iy >= g.dy) continue;
                    const bool edge = (ix == cx - R || ix == cx + R || iy == cy - R || iy == cy + R);
                    for (int iz = cz - R; iz <= cz + R; iz += edge ? 1 : (R > 0 ? 2 * R : 1)) {
                        if (iz < 0 || iz >= g.dz) continue;
                        const uint32_t c = (uint32_t)((ix * g.dy * g.dz) + (iy * g.dz) + iz);
                        for (uint32_t s = bstart[c]; s < bstart[c + 1]; ++s) {
                            const uint32_t b = bperm[s];
                            if (self && (int)b == q) continue;
                            const float d2 = d2_of(qx, qy, qz, sbx[s], sby[s], sbz[s]);
                            if (d2 < r2) insert_best(h.data(), k, d2, b);
                        }
                    }
                }
            }
            // Distance from the query to the nearest unscanned cell (faces of
            // the scanned box that are not grid boundaries).
            float dmin = FLT_MAX;
            bool all = true;
            if (cx - R > 0) { dmin = std::min(dmin, qx - (g.ox + (float)(cx - R) * hx)); all = false; }
            if (cx + R < g.dx - 1) { dmin = std::min(dmin, (g.ox + (float)(cx + R + 1) * hx) - qx); all = false; }
            if (cy - R > 0) { dmin = std::min(dmin, qy - (g.oy + (float)(cy - R) * hy)); all = false; }
            if (cy + R < g.dy - 1) { dmin = std::min(dmin, (g.oy + (float)(cy + R + 1) * hy) - qy); all = false; }
            if (cz - R > 0) { dmin = std::min(dmin, qz - (g.oz + (float)(cz - R) * hz)); all = false; }
            if (cz + R < g.dz - 1) { dmin = std::min(dmin, (g.oz + (float)(cz + R + 1) * hz) - qz); all = false; }
            if (all) break;
            // Conservative slack absorbs the rounding of the cell mapping.
            dmin -= 1e-5f * (std::fabs(g.ox) + std::fabs(g.oy) + std::fabs(g.oz) + g.sx + g.sy + g.sz);
            if (dmin > 0.f && std::min(h[k - 1].d2, r2) < dmin * dmin * (1.f - 1e-5f)) break;
        }
        write_knn(h.data(), k, nA, q, idx, dist);
    }
}

// ---------------------------------------------------------------------------
// Agent state, struct-of-arrays.  Field meaning follows AgentData
// (include/OpenSteer/AgentData.h:13-42) plus SimpleVehicle's smoothed
// acceleration (include/OpenSteer/SimpleVehicle.h:335-345).
template <class T> struct Agents {
    T *px, *py, *pz;    // position
    T *fx, *fy, *fz;    // unit forward (direction)
    T *speed;
    T *ax, *ay, *az;    // smoothed acceleration
    T *sx, *sy, *sz;    // side
    T *ux, *uy, *uz;    // up
    T *gx, *gy, *gz;    // steering accumulator
    T *radius, *max_speed, *max_force, *mass;
    uint32_t* bits;     // applied-behaviour bits (CUDAKernelGlobals.cuh:11-31)
};
template <class T> static inline V3<T> P(const Agents<T>& a, int i) { return mk(a.px[i], a.py[i], a.pz[i]); }
template <class T> static inline V3<T> F(const Agents<T>& a, int i) { return mk(a.fx[i], a.fy[i], a.fz[i]); }
template <class T> static inline V3<T> S(const Agents<T>& a, int i) { return mk(a.sx[i], a.sy[i], a.sz[i]); }
template <class T> static inline V3<T> U(const Agents<T>& a, int i) { return mk(a.ux[i], a.uy[i], a.uz[i]); }
template <class T> static inline V3<T> G(const Agents<T>& a, int i) { return mk(a.gx[i], a.gy[i], a.gz[i]); }
template <class T> static inline void setG(Agents<T>& a, int i, V3<T> v) { a.gx[i] = v.x; a.gy[i] = v.y; a.gz[i] = v.z; }

enum { MODE_OPENSTEER = 0, MODE_GPUREF = 1 };
enum {
    BIT_SEEK = 1u << 2,
    BIT_SEPARATION = 1u << 4,
    BIT_ALIGNMENT = 1u << 5,
    BIT_COHESION = 1u << 6,
    BIT_AVOID_OBSTACLES = 1u << 7,
    BIT_AVOID_WALLS = 1u << 8,
    BIT_AVOID_NEIGHBORS = 1u << 9,
    BIT_AVOID_CLOSE_NEIGHBORS = 1u << 10,
    BIT_FOLLOW_PATH = 1u << 11,
    BIT_ANTI_PENETRATION_WALL = 1u << 12,
    BIT_ANTI_PENETRATION_AGENT = 1u << 13,
};

// Applies a weighted behaviour result with the reference kernels' mask / bit /
// accumulate convention (e.g. SteerForSeparationCUDAKernel.cu:74-75,124-134):
// steering = (f*w) + steering, bit set when f*w is non-zero.
template <class T> static inline void accumulate(Agents<T>& a, int i, V3<T> f, T w, uint32_t bit) {
    const V3<T> fw = mul(f, w);
    if (!is_zero(fw)) a.bits[i] |= bit;
    setG(a, i, add(fw, G(a, i)));
}

// Boid neighbourhood predicate.
//  OPENSTEER: SteerLibrary::inBoidNeighborhood (src/SteerLibrary.cpp:506-544),
//             squared-distance tests; min distance = 3*radius when min_d < 0
//             (the value SteerLibrary passes, :566,607,646).
//  GPUREF:    FlockingCommon.cuh:4-30, sqrt distance tests, explicit min_d.
template <class T>
static inline bool in_neighborhood(int mode, V3<T> pos, V3<T> fwd, T radius, V3<T> other, T min_d, T max_d, T cos_max) {
    const V3<T> offset = sub(other, pos);
    if (mode == MODE_GPUREF) {
        const T d = len(offset);
        if (d < min_d) return true;
        if (d > max_d) return false;
        return dot(fwd, dvd(offset, d)) > cos_max;
    }
    const T md = (min_d < T(0)) ? radius * T(3) : min_d;
    const T d2 = dot(offset, offset);
    if (d2 < (md * md)) return true;
    if (d2 > (max_d * max_d)) return false;
    const V3<T> unit = dvd(offset, std::sqrt(d2));
    return dot(fwd, unit) > cos_max;
}

enum { FLOCK_SEPARATION = 0, FLOCK_ALIGNMENT = 1, FLOCK_COHESION = 2 };
// Separation (src/SteerLibrary.cpp:552-585; SteerForSeparationCUDAKernel.cu:98-121),
// alignment (:593-623; SteerForAlignmentCUDAKernel.cu), cohesion (:632-662;
// SteerForCohesionCUDAKernel.cu).  Neighbours are the KNN slots in slot order,
// stopping at the first pad (SteerForSeparationCUDAKernel.cu:104).
// Reads the step-t snapshot of B (Jacobi), writes A's steering and bits.
template <class T>
static void steer_flock_one(int which, int mode, Agents<T>& A, int nA, const uint32_t* knn, int k, const Agents<T>& B,
                            int nB, T min_d, T max_d, T cos_max, T w, uint32_t mask) {
    const uint32_t bit = which == FLOCK_SEPARATION ? BIT_SEPARATION : which == FLOCK_ALIGNMENT ? BIT_ALIGNMENT : BIT_COHESION;
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nA; ++i) {
        if (A.bits[i] & mask) continue;
        const V3<T> pos = P(A, i), fwd = F(A, i);
        V3<T> steering = mk(T(0), T(0), T(0));
        int count = 0;
        for (int j = 0; j < k; ++j) {
            const uint32_t b = knn[(size_t)j * nA + i];
            if (b >= (uint32_t)nB) break;
            const V3<T> bp = P(B, (int)b);
            if (!in_neighborhood(mode, pos, fwd, A.radius[i], bp, min_d, max_d, cos_max)) continue;
            if (which == FLOCK_SEPARATION) {
                const V3<T> offset = sub(bp, pos);
                const T d2 = dot(offset, offset);
                steering = add(steering, dvd(offset, -d2));
            } else if (which == FLOCK_ALIGNMENT) {
                steering = add(steering, F(B, (int)b));
            } else {
                steering = add(steering, bp);
            }
            ++count;
        }
        if (count > 0) {
            const V3<T> mean = dvd(steering, (T)count);
            if (which == FLOCK_SEPARATION) steering = normalize(mean);
            else if (which == FLOCK_ALIGNMENT) steering = normalize(sub(mean, fwd));
            else steering = normalize(sub(mean, pos));
        }
        accumulate(A, i, steering, w, bit);
    }
}

// Seek.  OPENSTEER: src/SteerLibrary.cpp:36-42 (target - position - velocity).
// GPUREF: SteerForSeekCUDAKernel.cu:52-61 (normalize(target - position - forward)).
template <class T>
static void steer_seek(int mode, Agents<T>& A, int nA, const T* tx, const T* ty, const T* tz, int per_agent, T w,
                       uint32_t mask) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nA; ++i) {
        if (A.bits[i] & mask) continue;
        const int t = per_agent ? i : 0;
        const V3<T> desired = sub(mk(tx[t], ty[t], tz[t]), P(A, i));
        V3<T> f;
        if (mode == MODE_GPUREF) f = normalize(sub(desired, F(A, i)));
        else f = sub(desired, mul(F(A, i), A.speed[i]));
        accumulate(A, i, f, w, BIT_SEEK);
    }
}

// Obstacle avoidance: SteerToAvoidObstaclesKernel and its
// findNextIntersectionWithSphere (include/OpenSteer/CUDA/AvoidObstaclesCUDAKernel.cu:47-99,136-253).
// (The CPU group version, src/SteerLibrary.cpp:222-228, returns zero.)
template <class T>
static void avoid_obstacles(Agents<T>& A, int nA, const uint32_t* knn, int k, int nO, const T* ox, const T* oy,
                            const T* oz, const T* orad, T min_time, T w, uint32_t mask) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nA; ++i) {
        if (A.bits[i] & mask) continue;
        const V3<T> pos = P(A, i), side = S(A, i), up = U(A, i), fwd = F(A, i);
        const T radius = A.radius[i];
        const T min_dist = min_time * A.speed[i];
        bool n_hit = false, x_hit = false;
        T n_dist = T(0), x_dist = T(0);
        V3<T> n_pos = mk(T(0), T(0), T(0)), x_pos = n_pos;
        for (int j = 0; j < k; ++j) {
            const uint32_t o = knn[(size_t)j * nA + i];
            if (o >= (uint32_t)nO) break;
            const V3<T> op = mk(ox[o], oy[o], oz[o]);
            const T orr = orad[o];
            // findNextIntersectionWithSphere (:47-99)
            x_hit = false;
            x_pos = op;
            const V3<T> go = sub(op, pos);
            const V3<T> lc = mk(dot(go, side), dot(go, up), dot(go, fwd));
            const T b = T(-2) * lc.z;
            const T c = ((lc.x * lc.x + lc.y * lc.y) + lc.z * lc.z) - (orr + radius) * (orr + radius);
            const T d = (b * b) - (T(4) * c);
            if (!(d < T(0))) {
                const T s = std::sqrt(d);
                const T p = (-b + s) / T(2);
                const T q = (-b - s) / T(2);
                if (!((p < T(0)) && (q < T(0)))) {
                    x_hit = true;
                    x_dist = ((p > T(0)) && (q > T(0))) ? ((p < q) ? p : q) : ((p > T(0)) ? p : q);
                }
            }
            // :220-221 nearest selection
            if ((n_hit == false) || ((x_hit != false) && (x_dist < n_dist))) {
                n_hit = x_hit;
                n_dist = x_dist;
                n_pos = x_pos;
            }
        }
        V3<T> steering = mk(T(0), T(0), T(0));
        if (n_hit && (n_dist < min_dist)) steering = perpendicular_component(sub(pos, n_pos), fwd);
        accumulate(A, i, steering, w, BIT_AVOID_OBSTACLES);
    }
}

// Wall avoidance: SteerToAvoidWallsCUDAKernel
// (include/OpenSteer/CUDA/AvoidWallsCUDAKernel.cu:67-245), including its side
// effect of moving agents that touch a wall (:215-223) and its bit rules.
template <class T>
static void avoid_walls(Agents<T>& A, int nA, const uint32_t* knn, int k, int nW, const T* wsx, const T* wsy,
                        const T* wsz, const T* wex, const T* wey, const T* wez, const T* wnx, const T* wny,
                        const T* wnz, T min_time, T w, uint32_t mask) {
    const T feeler_len = T(4);  // :41 FEELER_LENGTH
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nA; ++i) {
        if (A.bits[i] & mask) continue;
        V3<T> pos = P(A, i), dir = F(A, i);
        const V3<T> side = S(A, i);
        const T radius = A.radius[i];
        const V3<T> vel = mul(dir, A.speed[i]);
        V3<T> feelers[3];
        feelers[0] = add(pos, mul(vel, min_time));
        feelers[1] = add(pos, add(mul(side, feeler_len * radius), mul(dir, feeler_len * radius)));
        feelers[2] = add(pos, add(mul(neg(side), feeler_len * radius), mul(dir, feeler_len * radius)));
        int feeler_index = 0;
        T nearest_dist = T(FLT_MAX);
        uint32_t nearest_line = UINT32_MAX;
        V3<T> nearest_ip = mk(T(0), T(0), T(0)), nearest_normal = nearest_ip;
        bool touching = false;
        for (int j = 0; j < k; ++j) {
            const uint32_t l = knn[(size_t)j * nA + i];
            if (l >= (uint32_t)nW) break;
            const V3<T> ln = mk(wnx[l], wny[l], wnz[l]);
            const V3<T> ls = mk(wsx[l], wsy[l], wsz[l]);
            const V3<T> le = mk(wex[l], wey[l], wez[l]);
            V3<T> ip;
            const V3<T> closest = add(pos, mul(neg(ln), radius));
            if (lines_intersect(pos, closest, ls, le, ip)) {
                touching = true;
                nearest_normal = ln;
                nearest_dist = distance(pos, ip);
            }
            for (int f = 0; f < 3; ++f) {
                if (lines_intersect(pos, feelers[f], ls, le, ip)) {
                    const T dist = distance(pos, ip);
                    if (dist < nearest_dist) {
                        nearest_dist = dist;
                        nearest_line = l;
                        nearest_ip = ip;
                        feeler_index = f;
                        nearest_normal = ln;
                    }
                }
            }
        }
        V3<T> steering = mk(T(0), T(0), T(0));
        if (touching) {
            const T pen = radius - nearest_dist;
            pos = add(pos, mul(nearest_normal, pen));
            dir = normalize(perpendicular_component(dir, nearest_normal));
            A.bits[i] |= BIT_ANTI_PENETRATION_WALL;
        } else if (nearest_line != UINT32_MAX) {
            const T pen = len(sub(feelers[feeler_index], nearest_ip));
            steering = mul(nearest_normal, pen);
            A.bits[i] |= BIT_AVOID_WALLS;
        }
        steering = mul(steering, w);
        setG(A, i, add(steering, G(A, i)));
        A.px[i] = pos.x; A.py[i] = pos.y; A.pz[i] = pos.z;
        A.fx[i] = dir.x; A.fy[i] = dir.y; A.fz[i] = dir.z;
    }
}

// Integrate one step.
//  OPENSTEER: SimpleVehicle::applySteeringForce (src/SimpleVehicle.cpp:167-214)
//    with adjustRawSteeringForce (:111-132), regenerateLocalSpace (:224-230)
//    and regenerateOrthonormalBasisUF (include/OpenSteer/SimpleVehicle.h:190-217;
//    right-handed, side = forward x up un-normalized, :193-194).  Curvature and
//    smoothed position are annotation-only and dropped.
//  GPUREF: UpdateCUDAKernel (include/OpenSteer/CUDA/UpdateCUDAKernel.cu:78-123):
//    no adjustment, no smoothing, normalized side.  When the new speed is 0
//    the side is left unchanged (the reference writes an unset value).
// Both clear the steering accumulator and the bits (UpdateCUDAKernel.cu:111,123).
template <class T> static void update(int mode, Agents<T>& A, int nA, T dt) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nA; ++i) {
        V3<T> force = G(A, i);
        V3<T> fwd = F(A, i);
        V3<T> pos = P(A, i);
        V3<T> up = U(A, i);
        V3<T> side = S(A, i);
        T speed = A.speed[i];
        const T max_force = A.max_force[i], max_speed = A.max_speed[i], mass = A.mass[i];
        V3<T> new_vel;
        if (mode == MODE_GPUREF) {
            force = truncate_length(force, max_force);
            const V3<T> acc = dvd(force, mass);
            new_vel = add(mul(fwd, speed), mul(acc, dt));
            new_vel = truncate_length(new_vel, max_speed);
            speed = len(new_vel);
            if (speed > T(0)) {
                fwd = dvd(new_vel, speed);
                side = normalize(cross(fwd, up));
                up = cross(side, fwd);
            }
            pos = add(pos, mul(new_vel, dt));
        } else {
            // adjustRawSteeringForce (:111-132)
            V3<T> adjusted = force;
            const T max_adj = T(0.2f) * max_speed;
            if (!((speed > max_adj) || is_zero(force))) {
                const T range = speed / max_adj;
                const T cosine = T(1) + (T(-1) - T(1)) * pow20(range);  // interpolate(pow(range,20), 1, -1)
                adjusted = limit_max_deviation_angle(force, cosine, fwd);
            }
            const V3<T> clipped = truncate_length(adjusted, max_force);
            const V3<T> new_acc = dvd(clipped, mass);
            V3<T> sacc = mk(A.ax[i], A.ay[i], A.az[i]);
            if (dt > T(0)) {
                const T rate = clip(T(9) * dt, T(0.15f), T(0.4f));
                sacc = blend(rate, new_acc, sacc);
            }
            new_vel = add(mul(fwd, speed), mul(sacc, dt));
            new_vel = truncate_length(new_vel, max_speed);
            speed = len(new_vel);
            pos = add(pos, mul(new_vel, dt));
            if (speed > T(0)) {
                fwd = dvd(new_vel, speed);
                side = cross(fwd, up);
                up = cross(side, fwd);
            }
            A.ax[i] = sacc.x; A.ay[i] = sacc.y; A.az[i] = sacc.z;
        }
        A.px[i] = pos.x; A.py[i] = pos.y; A.pz[i] = pos.z;
        A.fx[i] = fwd.x; A.fy[i] = fwd.y; A.fz[i] = fwd.z;
        A.sx[i] = side.x; A.sy[i] = side.y; A.sz[i] = side.z;
        A.ux[i] = up.x; A.uy[i] = up.y; A.uz[i] = up.z;
        A.speed[i] = speed;
        A.gx[i] = T(0); A.gy[i] = T(0); A.gz[i] = T(0);
        A.bits[i] = 0u;
    }
}

// World boundary.
//  WRAP_BOX:    WrapWorldKernel (include/OpenSteer/CUDA/WrapWorldCUDAKernel.cu:25-35)
//               face teleport, applied on all three axes of a box given by
//               centre and size (the choke-point-only z rule at :42-43 is not
//               restated).
//  WRAP_SPHERE: Boid::handleBoundary mode 1 (plugins/Boids.cpp:185-209):
//               OpenSteer's sphericalWrapAround, p + (offset/r)*R*-2 when r > R.
enum { WRAP_BOX = 0, WRAP_SPHERE = 1 };
template <class T> static void wrap_world(int shape, Agents<T>& A, int nA, const T* c, const T* size_or_radius) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nA; ++i) {
        V3<T> p = P(A, i);
        if (shape == WRAP_BOX) {
            const T hx = size_or_radius[0] * T(0.5f), hy = size_or_radius[1] * T(0.5f), hz = size_or_radius[2] * T(0.5f);
            if (p.x < c[0] - hx) p.x = c[0] + hx;
            else if (p.x > c[0] + hx) p.x = c[0] - hx;
            if (p.y < c[1] - hy) p.y = c[1] + hy;
            else if (p.y > c[1] + hy) p.y = c[1] - hy;
            if (p.z < c[2] - hz) p.z = c[2] + hz;
            else if (p.z > c[2] + hz) p.z = c[2] - hz;
        } else {
            const T R = size_or_radius[0];
            const V3<T> offset = sub(p, mk(c[0], c[1], c[2]));
            const T r = len(offset);
            if (r > R) p = add(p, mul(mul(dvd(offset, r), R), T(-2)));
        }
        A.px[i] = p.x; A.py[i] = p.y; A.pz[i] = p.z;
    }
}


// ---------------------------------------------------------------------------
// Plugin-compat kernels (the choke-point plugins' per-step extras).

// ClosestPointOnLine (include/OpenSteer/CUDA/UpdateCUDAKernelNew.cu:73-81).
// The reference calls clamp(0.f, 1.f, t) with its arguments in the wrong
// order (:79 against the clamp(x, a, b) of :67-71), which returns 1 for every
// t: its "closest point" is always start + v * 1.  CLOSEST_AS_REFERENCE keeps
// that; CLOSEST_CLAMPED clamps t into [0, 1] (the closest point of the
// segment, the evident intent).
enum { CLOSEST_CLAMPED = 0, CLOSEST_AS_REFERENCE = 1 };
template <class T> static inline V3<T> closest_point_on_line(V3<T> p, V3<T> s, V3<T> e, int closest) {
    const V3<T> v = sub(e, s);
    const V3<T> w = sub(p, s);
    T t = dot(w, v) / dot(v, v);
    if (closest == CLOSEST_AS_REFERENCE) t = T(1);
    else t = t < T(0) ? T(0) : (t > T(1) ? T(1) : t);
    return add(s, mul(v, t));
}

// UpdateCUDAKernelNew (include/OpenSteer/CUDA/UpdateCUDAKernelNew.cu:95-268):
// the GPU-reference integrate with wall anti-penetration between the speed
// limit and the Euler step.  Over the k nearest walls in KNN order (stopping
// at a pad): the nearest crossing of the segment position -> position +
// v dt (LinesIntersect, strict <) and the LAST wall whose closest point lies
// within the radius (no distance comparison, :215-221).  A wall within the
// radius moves the agent by -normal * (d - r); a crossing replaces the
// velocity by its component perpendicular to that wall's normal; either sets
// the ANTI_PENETRATION_WALL bit.  Steering is zeroed; the bits are NOT cleared
// (:153 is commented out).  The KNN is read slot-major [k][n] as written by
// the search (the reference reads it agent-major, :150, which mixes agents).
// At speed 0 the side is left unchanged (the reference writes shared memory
// it never loaded, :155,261).
template <class T>
static void update_anti_penetration(Agents<T>& A, int nA, const uint32_t* knn, int k, int nW, const T* wsx,
                                    const T* wsy, const T* wsz, const T* wex, const T* wey, const T* wez, const T* wnx,
                                    const T* wny, const T* wnz, T dt, int closest) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nA; ++i) {
        V3<T> pos = P(A, i), fwd = F(A, i), side = S(A, i), up = U(A, i);
        const T radius = A.radius[i];
        V3<T> force = truncate_length(G(A, i), A.max_force[i]);
        const V3<T> acc = dvd(force, A.mass[i]);
        V3<T> new_vel = add(mul(fwd, A.speed[i]), mul(acc, dt));
        new_vel = truncate_length(new_vel, A.max_speed[i]);
        const V3<T> future = add(pos, mul(new_vel, dt));
        T line_d = T(FLT_MAX);
        uint32_t line = UINT32_MAX;
        V3<T> line_n = mk(T(0), T(0), T(0));
        T col_d = T(FLT_MAX);
        uint32_t col = UINT32_MAX;
        V3<T> col_n = line_n;
        for (int j = 0; j < k; ++j) {
            const uint32_t l = knn[(size_t)j * nA + i];
            if (l >= (uint32_t)nW) break;
            const V3<T> ls = mk(wsx[l], wsy[l], wsz[l]);
            const V3<T> le = mk(wex[l], wey[l], wez[l]);
            const V3<T> ln = mk(wnx[l], wny[l], wnz[l]);
            V3<T> ip;
            if (lines_intersect(pos, future, ls, le, ip)) {
                const T d = distance(pos, ip);
                if (d < col_d) {
                    col_d = d;
                    col = l;
                    col_n = ln;
                }
            }
            const V3<T> cp = sub(closest_point_on_line(pos, ls, le, closest), pos);
            const T d2 = dot(cp, cp);
            if (d2 < radius * radius) {
                line_d = std::sqrt(d2);
                line = l;
                line_n = ln;
            }
        }
        if (line != UINT32_MAX) {
            pos = add(pos, mul(neg(line_n), line_d - radius));
            A.bits[i] |= BIT_ANTI_PENETRATION_WALL;
        }
        if (col != UINT32_MAX) {
            new_vel = perpendicular_component(new_vel, col_n);
            A.bits[i] |= BIT_ANTI_PENETRATION_WALL;
        }
        const T speed = len(new_vel);
        if (speed > T(0)) {
            fwd = dvd(new_vel, speed);
            side = normalize(cross(fwd, up));
            up = cross(side, fwd);
        }
        pos = add(pos, mul(new_vel, dt));
        A.px[i] = pos.x; A.py[i] = pos.y; A.pz[i] = pos.z;
        A.fx[i] = fwd.x; A.fy[i] = fwd.y; A.fz[i] = fwd.z;
        A.sx[i] = side.x; A.sy[i] = side.y; A.sz[i] = side.z;
        A.ux[i] = up.x; A.uy[i] = up.y; A.uz[i] = up.z;
        A.speed[i] = speed;
        A.gx[i] = T(0); A.gy[i] = T(0); A.gz[i] = T(0);
    }
}

// AntiPenetrationAgentsCUDAKernel (include/OpenSteer/CUDA/AntiPenetrationAgentsCUDAKernel.cu:57-109)
// with its launcher (AntiPenetrationAgentsCUDA.cu:47-96): every agent of A
// not masked, whose first KNN entry is an agent of B overlapping it
// (r_a + r_b - knn distance >= 0), moves away from that agent's position
// along the normalized offset by the whole overlap when the other agent's
// bits match the mask, else by half; ANTI_PENETRATION_AGENT bit set.  All
// reads (A's and B's positions and bits) are of the state before the call,
// also when A is B (the reference reads B through textures while writing a
// copy of A's positions).
template <class T>
static void anti_penetration_agents(Agents<T>& A, int nA, const uint32_t* knn, const float* knn_dist,
                                    const Agents<T>& B, int nB, uint32_t mask) {
    std::vector<T> ox(nA), oy(nA), oz(nA);
    std::vector<uint32_t> ob(nA);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nA; ++i) {
        V3<T> pos = P(A, i);
        uint32_t bits = A.bits[i];
        const uint32_t j = knn[i];
        if (!(bits & mask) && j < (uint32_t)nB) {
            const T overlap = (A.radius[i] + B.radius[j]) - T(knn_dist[i]);
            if (!(overlap < T(0))) {
                const V3<T> off = normalize(sub(pos, P(B, j)));
                const T move = (B.bits[j] & mask) ? overlap : T(0.5f) * overlap;
                pos = add(pos, mul(off, move));
                bits |= BIT_ANTI_PENETRATION_AGENT;
            }
        }
        ox[i] = pos.x; oy[i] = pos.y; oz[i] = pos.z;
        ob[i] = bits;
    }
    for (int i = 0; i < nA; ++i) {
        A.px[i] = ox[i]; A.py[i] = oy[i]; A.pz[i] = oz[i];
        A.bits[i] = ob[i];
    }
}

// PolylinePathwayCUDA::Initialize (include/OpenSteer/CUDA/PolylinePathwayCUDA.cu:17-60):
// m = n (+1 when cyclic, closing on the first point); per segment end i >= 1
// the unit direction normals[i], the length lengths[i] and the running total;
// entry 0 of both stays zero.  Returns the total length.
template <class T>
static T path_init(int n, const T* pts_in, int cyclic, T* pts, T* nrm, T* lens) {
    const int m = n + (cyclic ? 1 : 0);
    T total = T(0);
    for (int i = 0; i < m; ++i) {
        const int j = (cyclic && i == m - 1) ? 0 : i;
        for (int a = 0; a < 3; ++a) pts[3 * i + a] = pts_in[3 * j + a];
        if (i > 0) {
            const V3<T> d = sub(mk(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]),
                                mk(pts[3 * i - 3], pts[3 * i - 2], pts[3 * i - 1]));
            lens[i] = len(d);
            const V3<T> u = normalize(d);
            nrm[3 * i] = u.x; nrm[3 * i + 1] = u.y; nrm[3 * i + 2] = u.z;
            total += lens[i];
        } else {
            nrm[0] = nrm[1] = nrm[2] = T(0);
            lens[0] = T(0);
        }
    }
    return total;
}

template <class T> struct Path {
    int m;
    const T *pts, *nrm, *lens;
    T radius, total;
    int cyclic;
    V3<T> pt(int i) const { return mk(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]); }
    V3<T> nm(int i) const { return mk(nrm[3 * i], nrm[3 * i + 1], nrm[3 * i + 2]); }
};

// pointToSegmentDistance (CUDA/SteerToFollowPathCUDAKernel.cu:34-69).
template <class T>
static inline T point_to_segment(V3<T> p, V3<T> e0, V3<T> e1, V3<T> n, T seg_len, T& proj, V3<T>& chosen) {
    const V3<T> local = sub(p, e0);
    proj = dot(n, local);
    if (proj < T(0)) {
        chosen = e0;
        proj = T(0);
        return distance(p, e0);
    }
    if (proj > seg_len) {
        chosen = e1;
        proj = seg_len;
        return distance(p, e1);
    }
    chosen = add(e0, mul(n, proj));
    return distance(p, chosen);
}

// mapPointToPathDistance (:71-102).
template <class T> static T path_distance_of(const Path<T>& P_, V3<T> p) {
    T min_d = T(FLT_MAX), total = T(0), along = T(0);
    for (int i = 1; i < P_.m; ++i) {
        T proj;
        V3<T> chosen;
        const T d = point_to_segment(p, P_.pt(i - 1), P_.pt(i), P_.nm(i), P_.lens[i], proj, chosen);
        if (d < min_d) {
            min_d = d;
            along = total + proj;
        }
        total += P_.lens[i];
    }
    return along;
}

// mapPointToPath (:104-142): the nearest point on the path and how far p is
// outside the tube.
template <class T> static V3<T> map_point_to_path(const Path<T>& P_, V3<T> p, T& outside) {
    T min_d = T(FLT_MAX);
    V3<T> on = mk(T(0), T(0), T(0));
    for (int i = 1; i < P_.m; ++i) {
        T proj;
        V3<T> chosen;
        const T d = point_to_segment(p, P_.pt(i - 1), P_.pt(i), P_.nm(i), P_.lens[i], proj, chosen);
        if (d < min_d) {
            min_d = d;
            on = chosen;
        }
    }
    outside = distance(on, p) - P_.radius;
    return on;
}

// mapPathDistanceToPoint (:144-187).  A distance the segment walk does not
// place (rounding leaves `remaining` above the last segment's length) maps to
// the last point, where the reference returns an unset value.
template <class T> static V3<T> path_point_at(const Path<T>& P_, T dist) {
    T remaining = dist;
    if (P_.cyclic) {
        remaining = std::fmod(dist, P_.total);
    } else {
        if (dist < T(0)) return P_.pt(0);
        if (dist >= P_.total) return P_.pt(P_.m - 1);
    }
    for (int i = 1; i < P_.m; ++i) {
        const T seg = P_.lens[i];
        if (seg < remaining) {
            remaining -= seg;
        } else {
            const T ratio = remaining / seg;
            return add(P_.pt(i - 1), mul(sub(P_.pt(i), P_.pt(i - 1)), ratio));  // interpolate, VectorUtils.cuh:292-295
        }
    }
    return P_.pt(P_.m - 1);
}

// FollowPathCUDAKernel (CUDA/SteerToFollowPathCUDAKernel.cu:189-287).
template <class T> static void follow_path(Agents<T>& A, int nA, const Path<T>& P_, T prediction_time, T w, uint32_t mask) {
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nA; ++i) {
        if (A.bits[i] & mask) continue;
        const V3<T> pos = P(A, i), fwd = F(A, i);
        const T speed = A.speed[i];
        const T offset = prediction_time * speed;
        const V3<T> future = add(pos, mul(mul(fwd, speed), prediction_time));
        const T now_d = path_distance_of(P_, pos);
        const T future_d = path_distance_of(P_, future);
        const bool rightway = (offset > T(0)) ? (now_d < future_d) : (now_d > future_d);
        T outside;
        map_point_to_path(P_, future, outside);
        V3<T> steering = mk(T(0), T(0), T(0));
        if (!((outside < T(0)) && rightway)) {
            const V3<T> target = path_point_at(P_, now_d + offset);
            steering = sub(sub(target, pos), fwd);
        }
        accumulate(A, i, normalize(steering), w, BIT_FOLLOW_PATH);
    }
}

// ---------------------------------------------------------------------------
// Neighbour avoidance (unaligned collision avoidance).

// predictNearestApproachTime (CUDA/SteerToAvoidNeighborsCUDAKernel.cu:205-233,
// src/SteerLibrary.cpp:404-438).
template <class T>
static inline T nearest_approach_time(V3<T> pos, V3<T> dir, T speed, V3<T> opos, V3<T> odir, T ospeed) {
    const V3<T> my_vel = mul(dir, speed);
    const V3<T> other_vel = mul(odir, ospeed);
    const V3<T> rel = sub(other_vel, my_vel);
    const T rel_speed = len(rel);
    if (rel_speed == T(0)) return T(0);
    const V3<T> tangent = dvd(rel, rel_speed);
    const V3<T> rel_pos = sub(pos, opos);
    const T projection = dot(tangent, rel_pos);
    return projection / rel_speed;
}

// steerToAvoidNeighbors over the KNN of A in B (B's fields are only read, so
// B may be A).
//  GPUREF: SteerToAvoidNeighborsCUDA::run (CUDA/SteerToAvoidNeighborsCUDA.cu:85-150):
//    with avoid_close, SteerToAvoidCloseNeighborsCUDAKernel
//    (CUDA/SteerToAvoidNeighborsCUDAKernel.cu:93-187: the first neighbour in
//    KNN order whose KNN distance is below min_sep + r_a + r_b gives
//    perpendicularComponent(-offset, forward); normalized, weighted,
//    AVOID_CLOSE_NEIGHBORS bit), then SteerToAvoidNeighborsCUDAKernel (:235-402)
//    masked by mask | AVOID_CLOSE_NEIGHBORS: the neighbour with the earliest
//    nearest approach in [0, min_time) closer than r_a + r_b; head-on
//    (parallelness < -0.707) steers away from the threat position at nearest
//    approach -- as in the kernel, the value its helper wrote last, for the
//    last neighbour that passed the time test (:334-338), chosen or not;
//    parallel (> 0.707) away from the threat's position (the kernel reads an
//    unset variable there, :372; its current position is the CPU rule);
//    otherwise, when not faster than the threat, behind it.  side * steer is
//    normalized and weighted; nothing is added without a threat.
//  OPENSTEER: SteerLibrary::steerToAvoidNeighbors (src/SteerLibrary.cpp:290-389):
//    steerToAvoidCloseNeighbors(v, 0, others) first (:470-495: current
//    distances, min separation 0, the perpendicular component unnormalized,
//    AVOID_CLOSE_NEIGHBORS bit); otherwise the threat within min_time closer
//    than 2 r_a, the head-on rule with the chosen threat's approach position,
//    the perpendicular rule when the threat is not faster; side * steer added
//    unnormalized (zero without a threat); min_sep and avoid_close unused.
template <class T>
static void avoid_neighbors(int mode, Agents<T>& A, int nA, const uint32_t* knn, const float* knn_dist, int k,
                            const Agents<T>& B, int nB, T min_time, T min_sep, int avoid_close, T w, uint32_t mask) {
    const T angle = T(0.707f);
#pragma omp parallel for schedule(static)
    for (int i = 0; i < nA; ++i) {
        if (A.bits[i] & mask) continue;
        const V3<T> pos = P(A, i), dir = F(A, i), side = S(A, i);
        const T speed = A.speed[i], radius = A.radius[i];
        if (mode == MODE_GPUREF && avoid_close) {
            V3<T> steering = mk(T(0), T(0), T(0));
            for (int j = 0; j < k; ++j) {
                const uint32_t b = knn[(size_t)j * nA + i];
                if (b >= (uint32_t)nB) break;
                const T d = T(knn_dist[(size_t)j * nA + i]);
                const T min_c = min_sep + (radius + B.radius[b]);
                if (d < min_c) {
                    steering = perpendicular_component(neg(sub(P(B, b), pos)), dir);
                    break;
                }
            }
            accumulate(A, i, normalize(steering), w, BIT_AVOID_CLOSE_NEIGHBORS);
            if (A.bits[i] & BIT_AVOID_CLOSE_NEIGHBORS) continue;
        }
        if (mode == MODE_GPUREF && (A.bits[i] & BIT_AVOID_CLOSE_NEIGHBORS)) continue;
        if (mode == MODE_OPENSTEER) {
            V3<T> sep = mk(T(0), T(0), T(0));
            for (int j = 0; j < k; ++j) {
                const uint32_t b = knn[(size_t)j * nA + i];
                if (b >= (uint32_t)nB) break;
                const T min_c = T(0) + (radius + B.radius[b]);
                const V3<T> off = sub(P(B, b), pos);
                if (len(off) < min_c) {
                    sep = perpendicular_component(neg(off), dir);
                    break;
                }
            }
            if (!is_zero(sep)) {
                accumulate(A, i, sep, w, BIT_AVOID_CLOSE_NEIGHBORS);
                continue;
            }
        }
        T min_t = min_time;
        bool found = false;
        V3<T> last_near = mk(T(0), T(0), T(0)), threat_near = last_near;
        V3<T> t_dir = last_near, t_pos = last_near;
        T t_speed = T(0);
        for (int j = 0; j < k; ++j) {
            const uint32_t b = knn[(size_t)j * nA + i];
            if (b >= (uint32_t)nB) break;
            const V3<T> bp = P(B, b), bd = F(B, b);
            const T bs = B.speed[b];
            const T threshold = mode == MODE_GPUREF ? radius + B.radius[b] : radius * T(2);
            const T time = nearest_approach_time(pos, dir, speed, bp, bd, bs);
            if (time >= T(0) && time < min_t) {
                const V3<T> my_near = add(pos, mul(dir, speed * time));
                const V3<T> his_near = add(bp, mul(bd, bs * time));
                last_near = his_near;
                if (distance(my_near, his_near) < threshold) {
                    min_t = time;
                    found = true;
                    t_dir = bd;
                    t_speed = bs;
                    t_pos = bp;
                    threat_near = his_near;
                }
            }
        }
        T steer = T(0);
        if (found) {
            const T par = dot(dir, t_dir);
            if (par < -angle) {
                const V3<T> off = sub(mode == MODE_GPUREF ? last_near : threat_near, pos);
                steer = dot(off, side) > T(0) ? T(-1) : T(1);
            } else if (par > angle) {
                const V3<T> off = sub(t_pos, pos);
                steer = dot(off, side) > T(0) ? T(-1) : T(1);
            } else if (mode == MODE_GPUREF ? (speed <= t_speed) : (t_speed <= speed)) {
                const T sd = dot(side, mul(t_dir, t_speed));
                steer = sd > T(0) ? T(-1) : T(1);
            }
        }
        if (mode == MODE_GPUREF) {
            if (found) accumulate(A, i, normalize(mul(side, steer)), w, BIT_AVOID_NEIGHBORS);
        } else {
            accumulate(A, i, mul(side, steer), w, BIT_AVOID_NEIGHBORS);
        }
    }
}

// ---------------------------------------------------------------------------
// Calibration model (the differentiable crowd step) for gradient checks.
//
// One step per agent of group g, every neighbour read from the step's start
// (Jacobi), the reference CUDA kernels' formulas: wall avoidance with the
// anti-penetration move (CUDA/AvoidWallsCUDAKernel.cu:67-245), obstacle
// avoidance (CUDA/AvoidObstaclesCUDAKernel.cu:136-253), separation from the
// own and the other group, alignment and cohesion with the own group
// (CUDA/FlockingCommon.cuh:4-30 neighbourhoods, SteerForSeparation /
// Alignment / CohesionCUDAKernel.cu), seek (SteerForSeekCUDAKernel.cu:52-61),
// then UpdateCUDAKernel (CUDA/UpdateCUDAKernel.cu:78-108).  Loss: mean over
// agents of |p_T - goal|^2 + pen_w * sum over steps and over each agent's two
// KNN lists of max(0, r_i + r_j - d)^2.
//
// Every branch goes through a decision tape: recorded on a base run, replayed
// under perturbed parameters, so central differences measure the derivative
// of the branch the base run took (KNN lists are taped the same way).
// Forward-mode tangent of the calibration model (the FP64 analytic
// gradient).  A value carries its primal P -- float: the fp32 model's own
// arithmetic, so the primal trajectory is bit for bit the fp32 run's; double:
// the FP64 model -- and its derivative with respect to each of the 14
// parameters in double.  Every operation applies the textbook rule at its
// primal operands (sum, product, quotient, square root, absolute value), so
// the loss's tangents are the exact derivative of the branch the decision
// tape fixes, evaluated along the primal trajectory.  This is the forward
// (tangent-linear) form of the derivative the GPU's reverse-mode adjoint
// computes, derived independently of it.
constexpr int kC4Params = 14;
template <class P> struct Tan {
    P v;
    double d[kC4Params];
    Tan() : v(P(0)) { for (double& x : d) x = 0.0; }
    Tan(double x) : v(P(x)) { for (double& y : d) y = 0.0; }
    explicit operator double() const { return double(v); }
};
template <class P> static inline Tan<P> operator+(const Tan<P>& a, const Tan<P>& b) {
    Tan<P> r(0.0);
    r.v = a.v + b.v;
    for (int j = 0; j < kC4Params; ++j) r.d[j] = a.d[j] + b.d[j];
    return r;
}
template <class P> static inline Tan<P> operator-(const Tan<P>& a, const Tan<P>& b) {
    Tan<P> r(0.0);
    r.v = a.v - b.v;
    for (int j = 0; j < kC4Params; ++j) r.d[j] = a.d[j] - b.d[j];
    return r;
}
template <class P> static inline Tan<P> operator-(const Tan<P>& a) {
    Tan<P> r(0.0);
    r.v = -a.v;
    for (int j = 0; j < kC4Params; ++j) r.d[j] = -a.d[j];
    return r;
}
template <class P> static inline Tan<P> operator*(const Tan<P>& a, const Tan<P>& b) {
    Tan<P> r(0.0);
    r.v = a.v * b.v;
    const double av = double(a.v), bv = double(b.v);
    for (int j = 0; j < kC4Params; ++j) r.d[j] = a.d[j] * bv + av * b.d[j];
    return r;
}
template <class P> static inline Tan<P> operator/(const Tan<P>& a, const Tan<P>& b) {
    Tan<P> r(0.0);
    r.v = a.v / b.v;
    const double bv = double(b.v), q = double(a.v) / bv;
    for (int j = 0; j < kC4Params; ++j) r.d[j] = (a.d[j] - q * b.d[j]) / bv;
    return r;
}
template <class P> static inline Tan<P>& operator+=(Tan<P>& a, const Tan<P>& b) { return a = a + b; }
template <class P> static inline bool operator<(const Tan<P>& a, const Tan<P>& b) { return a.v < b.v; }
template <class P> static inline bool operator>(const Tan<P>& a, const Tan<P>& b) { return a.v > b.v; }
template <class P> static inline bool operator<=(const Tan<P>& a, const Tan<P>& b) { return a.v <= b.v; }
template <class P> static inline bool operator>=(const Tan<P>& a, const Tan<P>& b) { return a.v >= b.v; }
template <class P> static inline bool operator==(const Tan<P>& a, const Tan<P>& b) { return a.v == b.v; }
template <class P> static inline bool operator!=(const Tan<P>& a, const Tan<P>& b) { return a.v != b.v; }
// d sqrt(x) = dx / (2 sqrt(x)); at x = 0 (a zero length, which the model
// only reaches on branches that do not use it) the tangent is taken as 0.
template <class P> static inline Tan<P> osqrt(const Tan<P>& a) {
    Tan<P> r(0.0);
    r.v = std::sqrt(a.v);
    const double s = std::sqrt(double(a.v));
    if (s > 0.0)
        for (int j = 0; j < kC4Params; ++j) r.d[j] = a.d[j] / (2.0 * s);
    return r;
}
template <class P> static inline Tan<P> ofabs(const Tan<P>& a) { return a.v < P(0) ? -a : a; }

// A parameter's value as the model's scalar: T(v), and for the tangent type
// the unit derivative d(theta_j)/d(theta_j) = 1.
template <class T> struct C4Seed {
    static T of(double v, int) { return T(v); }
};
template <class P> struct C4Seed<Tan<P>> {
    static Tan<P> of(double v, int j) {
        Tan<P> t(v);
        t.d[j] = 1.0;
        return t;
    }
};

#pragma omp declare reduction(+ : Tan<float>, Tan<double> : omp_out = omp_out + omp_in)

struct C4Params {
    double w_sep[2], w_ali[2], w_coh[2], w_obs[2], w_wall[2], radius[2], vmax[2];
};
struct C4Config {
    double sep_min, sep_max, sep_cos, ali_min, ali_max, ali_cos, coh_min, coh_max, coh_cos;
    uint32_t sep_mask, ali_mask, coh_mask, seek_mask;
    double w_seek, mt_walls, mt_obs, max_force, mass, dt, pen_w;
    int k, kw, ko, steps;
    double other_cap;
};
struct C4Scenario {
    int n[2];
    std::vector<double> st[2][13];  // p f speed side up per group
    double goal[2][3];
    std::vector<double> w[9];       // walls: start end normal
    std::vector<double> o[4];       // obstacles: centre radius
    Grid grid;
    // tape: per step, per agent (both groups, group 0 first) decisions; KNN lists
    std::vector<std::vector<std::vector<uint8_t>>> tape;
    std::vector<std::vector<uint32_t>> knn;  // per step: own[k][n], other[k][n], walls[kw][n], obs[ko][n] per group
};

struct Dec {
    int mode;  // 1 record, 2 replay
    std::vector<uint8_t>* v;
    size_t at;
    bool operator()(bool c) {
        if (mode == 1) {
            v->push_back(c ? 1 : 0);
            return c;
        }
        return (*v)[at++] != 0;
    }
};

template <class T> static inline V3<T> normalize_d(V3<T> a, Dec& D) {
    const T l = len(a);
    return D(l > T(0)) ? dvd(a, l) : a;
}
template <class T> static inline V3<T> truncate_d(V3<T> v, T m, Dec& D) {
    const T l2 = dot(v, v);
    if (D(l2 <= m * m)) return v;
    return mul(v, m / osqrt(l2));
}
template <class T> static bool li_d(V3<T> sa, V3<T> ea, V3<T> sb, V3<T> eb, V3<T>& ip, Dec& D) {
    const T eps = T(0.001f);
    const T denom = ((eb.z - sb.z) * (ea.x - sa.x)) - ((eb.x - sb.x) * (ea.z - sa.z));
    const T numera = ((eb.x - sb.x) * (sa.z - sb.z)) - ((eb.z - sb.z) * (sa.x - sb.x));
    const T numerb = ((ea.x - sa.x) * (sa.z - sb.z)) - ((ea.z - sa.z) * (sa.x - sb.x));
    if (D(ofabs(denom) < eps)) {
        if (D(ofabs(numera) < eps && ofabs(numerb) < eps)) {
            ip = sa;
            return true;
        }
        return false;
    }
    const T ua = numera / denom;
    const T ub = numerb / denom;
    if (D(ua >= T(0) && ua <= T(1) && ub >= T(0) && ub <= T(1))) {
        ip = mk(sa.x + ua * (ea.x - sa.x), T(0), sa.z + ua * (ea.z - sa.z));
        return true;
    }
    return false;
}
template <class T> static bool hood_d(V3<T> off, T d2, V3<T> fwd, T mn, T mx, T cs, Dec& D) {
    const T d = osqrt(d2);
    if (D(d < mn)) return true;
    if (D(d > mx)) return false;
    return D(dot(fwd, dvd(off, d)) > cs);
}
template <class T> static void apply_d(V3<T>& G, uint32_t& bits, V3<T> f, T w, uint32_t bit, Dec& D) {
    const V3<T> fw = mul(f, w);
    if (D(!is_zero(fw))) bits |= bit;
    G = add(fw, G);
}

// One step of scenario sc from state X (both groups) into Y; returns the
// penetration loss of the step.  lists: this step's KNN lists (see C4Scenario).
template <class T>
static T c4_step(const C4Scenario& sc, const C4Params& th, const C4Config& cf, const std::vector<T>* X,
                 std::vector<T>* Y, const std::vector<uint32_t>* lists, std::vector<std::vector<uint8_t>>& tape,
                 int mode) {
    const int k = cf.k;
    T pen_total = T(0);
    for (int g = 0; g < 2; ++g) {
        const int n = sc.n[g], no = sc.n[1 - g];
        const std::vector<T>* Xg = X + 13 * g;
        const std::vector<T>* Xo = X + 13 * (1 - g);
        std::vector<T>* Yg = Y + 13 * g;
        const uint32_t* own = lists[4 * g + 0].data();
        const uint32_t* oth = lists[4 * g + 1].data();
        const uint32_t* kw = lists[4 * g + 2].data();
        const uint32_t* ko = lists[4 * g + 3].data();
        const T r = C4Seed<T>::of(th.radius[g], 10 + g), vmax = C4Seed<T>::of(th.vmax[g], 12 + g);
        T pen = T(0);
#pragma omp parallel for schedule(static) reduction(+ : pen)
        for (int i = 0; i < n; ++i) {
            Dec D{mode, &tape[(size_t)g * sc.n[0] + i], 0};
            if (mode == 1) tape[(size_t)g * sc.n[0] + i].clear();
            auto P3 = [&](const std::vector<T>* A, int base, int j) { return mk(A[base][j], A[base + 1][j], A[base + 2][j]); };
            V3<T> p = P3(Xg, 0, i), f = P3(Xg, 3, i), sd = P3(Xg, 7, i), up = P3(Xg, 10, i);
            const T s = Xg[6][i];
            uint32_t bits = 0;
            V3<T> G = mk(T(0), T(0), T(0));
            // walls
            {
                const T fl = T(4) * r;
                const V3<T> vel = mul(f, s);
                V3<T> fe[3] = {add(p, mul(vel, T(cf.mt_walls))), add(p, add(mul(sd, fl), mul(f, fl))),
                               add(p, add(mul(neg(sd), fl), mul(f, fl)))};
                int feeler = 0;
                T nearest_dist = T(FLT_MAX);
                bool have_line = false, touching = false;
                V3<T> nearest_ip = mk(T(0), T(0), T(0)), nearest_normal = nearest_ip;
                for (int j = 0; j < cf.kw; ++j) {
                    const uint32_t l = kw[(size_t)j * n + i];
                    if (D(l >= (uint32_t)sc.w[0].size())) break;
                    const V3<T> ln = mk(T(sc.w[6][l]), T(sc.w[7][l]), T(sc.w[8][l]));
                    const V3<T> ls = mk(T(sc.w[0][l]), T(sc.w[1][l]), T(sc.w[2][l]));
                    const V3<T> le = mk(T(sc.w[3][l]), T(sc.w[4][l]), T(sc.w[5][l]));
                    V3<T> ip;
                    if (li_d(p, add(p, mul(neg(ln), r)), ls, le, ip, D)) {
                        touching = true;
                        nearest_normal = ln;
                        nearest_dist = distance(p, ip);
                    }
                    for (int q = 0; q < 3; ++q) {
                        if (li_d(p, fe[q], ls, le, ip, D)) {
                            const T dist = distance(p, ip);
                            if (D(dist < nearest_dist)) {
                                nearest_dist = dist;
                                have_line = true;
                                nearest_ip = ip;
                                feeler = q;
                                nearest_normal = ln;
                            }
                        }
                    }
                }
                V3<T> steer = mk(T(0), T(0), T(0));
                if (D(touching)) {
                    const T pn_ = r - nearest_dist;
                    p = add(p, mul(nearest_normal, pn_));
                    f = normalize_d(perpendicular_component(f, nearest_normal), D);
                    bits |= BIT_ANTI_PENETRATION_WALL;
                } else if (D(have_line)) {
                    steer = mul(nearest_normal, len(sub(fe[feeler], nearest_ip)));
                    bits |= BIT_AVOID_WALLS;
                }
                G = add(mul(steer, C4Seed<T>::of(th.w_wall[g], 8 + g)), G);
            }
            // obstacles
            {
                const T min_dist = T(cf.mt_obs) * s;
                bool n_hit = false;
                T n_dist = T(0);
                V3<T> n_pos = mk(T(0), T(0), T(0));
                for (int j = 0; j < cf.ko; ++j) {
                    const uint32_t o = ko[(size_t)j * n + i];
                    if (D(o >= (uint32_t)sc.o[0].size())) break;
                    const V3<T> op = mk(T(sc.o[0][o]), T(sc.o[1][o]), T(sc.o[2][o]));
                    const T orr = T(sc.o[3][o]);
                    bool x_hit = false;
                    T x_dist = T(0);
                    const V3<T> go = sub(op, p);
                    const V3<T> lc = mk(dot(go, sd), dot(go, up), dot(go, f));
                    const T b = T(-2) * lc.z;
                    const T c = ((lc.x * lc.x + lc.y * lc.y) + lc.z * lc.z) - (orr + r) * (orr + r);
                    const T d = (b * b) - (T(4) * c);
                    if (D(!(d < T(0)))) {
                        const T sq = osqrt(d);
                        const T pp = (-b + sq) / T(2), qq = (-b - sq) / T(2);
                        if (D(!((pp < T(0)) && (qq < T(0))))) {
                            x_hit = true;
                            x_dist = D((pp > T(0)) && (qq > T(0))) ? (D(pp < qq) ? pp : qq) : (D(pp > T(0)) ? pp : qq);
                        }
                    }
                    if (D((n_hit == false) || ((x_hit != false) && (x_dist < n_dist)))) {
                        n_hit = x_hit;
                        n_dist = x_dist;
                        n_pos = op;
                    }
                }
                V3<T> fo = mk(T(0), T(0), T(0));
                if (D(n_hit && (n_dist < min_dist))) fo = perpendicular_component(sub(p, n_pos), f);
                apply_d(G, bits, fo, C4Seed<T>::of(th.w_obs[g], 6 + g), BIT_AVOID_OBSTACLES, D);
            }
            // separation: own group, other group
            for (int w = 0; w < 2; ++w) {
                if (D(!(bits & cf.sep_mask))) {
                    const uint32_t* lst = w == 0 ? own : oth;
                    const std::vector<T>* B = w == 0 ? Xg : Xo;
                    const int nb = w == 0 ? n : no;
                    V3<T> acc = mk(T(0), T(0), T(0));
                    int cnt = 0;
                    for (int j = 0; j < k; ++j) {
                        const uint32_t b = lst[(size_t)j * n + i];
                        if (D(b >= (uint32_t)nb)) break;
                        const V3<T> off = sub(P3(B, 0, b), p);
                        const T d2 = dot(off, off);
                        if (hood_d(off, d2, f, T(cf.sep_min), T(cf.sep_max), T(cf.sep_cos), D)) {
                            acc = add(acc, dvd(off, -d2));
                            ++cnt;
                        }
                    }
                    if (D(cnt > 0)) acc = normalize_d(dvd(acc, (T)cnt), D);
                    apply_d(G, bits, acc, C4Seed<T>::of(th.w_sep[g], 0 + g), BIT_SEPARATION, D);
                }
            }
            // alignment
            if (D(!(bits & cf.ali_mask))) {
                V3<T> acc = mk(T(0), T(0), T(0));
                int cnt = 0;
                for (int j = 0; j < k; ++j) {
                    const uint32_t b = own[(size_t)j * n + i];
                    if (D(b >= (uint32_t)n)) break;
                    const V3<T> off = sub(P3(Xg, 0, b), p);
                    if (hood_d(off, dot(off, off), f, T(cf.ali_min), T(cf.ali_max), T(cf.ali_cos), D)) {
                        acc = add(acc, P3(Xg, 3, b));
                        ++cnt;
                    }
                }
                if (D(cnt > 0)) acc = normalize_d(sub(dvd(acc, (T)cnt), f), D);
                apply_d(G, bits, acc, C4Seed<T>::of(th.w_ali[g], 2 + g), BIT_ALIGNMENT, D);
            }
            // cohesion
            if (D(!(bits & cf.coh_mask))) {
                V3<T> acc = mk(T(0), T(0), T(0));
                int cnt = 0;
                for (int j = 0; j < k; ++j) {
                    const uint32_t b = own[(size_t)j * n + i];
                    if (D(b >= (uint32_t)n)) break;
                    const V3<T> pb = P3(Xg, 0, b);
                    const V3<T> off = sub(pb, p);
                    if (hood_d(off, dot(off, off), f, T(cf.coh_min), T(cf.coh_max), T(cf.coh_cos), D)) {
                        acc = add(acc, pb);
                        ++cnt;
                    }
                }
                if (D(cnt > 0)) acc = normalize_d(sub(dvd(acc, (T)cnt), p), D);
                apply_d(G, bits, acc, C4Seed<T>::of(th.w_coh[g], 4 + g), BIT_COHESION, D);
            }
            // seek
            if (D(!(bits & cf.seek_mask))) {
                const V3<T> goal = mk(T(sc.goal[g][0]), T(sc.goal[g][1]), T(sc.goal[g][2]));
                apply_d(G, bits, normalize_d(sub(sub(goal, p), f), D), T(cf.w_seek), BIT_SEEK, D);
            }
            // integrate (UpdateCUDAKernel)
            const V3<T> Ft = truncate_d(G, T(cf.max_force), D);
            const V3<T> acc = dvd(Ft, T(cf.mass));
            const V3<T> v = truncate_d(add(mul(f, s), mul(acc, T(cf.dt))), vmax, D);
            const T sn = len(v);
            V3<T> fn = f, sdn = sd, upn = up;
            if (D(sn > T(0))) {
                fn = dvd(v, sn);
                sdn = normalize_d(cross(fn, up), D);
                upn = cross(sdn, fn);
            }
            const V3<T> pn = add(p, mul(v, T(cf.dt)));
            const V3<T> out[4] = {pn, fn, sdn, upn};
            for (int q = 0; q < 3; ++q) {
                Yg[q][i] = (&out[0].x)[q];
                Yg[3 + q][i] = (&out[1].x)[q];
                Yg[7 + q][i] = (&out[2].x)[q];
                Yg[10 + q][i] = (&out[3].x)[q];
            }
            Yg[6][i] = sn;
            // penetration of this step (step-start positions)
            const V3<T> p0 = P3(Xg, 0, i);
            for (int w = 0; w < 2; ++w) {
                const uint32_t* lst = w == 0 ? own : oth;
                const std::vector<T>* B = w == 0 ? Xg : Xo;
                const int nb = w == 0 ? n : no;
                const T R = r + C4Seed<T>::of(th.radius[w == 0 ? g : 1 - g], 10 + (w == 0 ? g : 1 - g));
                for (int j = 0; j < k; ++j) {
                    const uint32_t b = lst[(size_t)j * n + i];
                    if (D(b >= (uint32_t)nb)) break;
                    const T e = R - distance(P3(B, 0, b), p0);
                    if (D(e > T(0))) pen += e * e;
                }
            }
        }
        pen_total += pen;
    }
    return pen_total * T(cf.pen_w);
}

// KNN lists of one step from the state (positions rounded to fp32, as the
// GPU's), for both groups: own (exact), other (capped), walls / obstacles.
template <class T>
static void c4_lists(const C4Scenario& sc, const C4Config& cf, const std::vector<T>* X, std::vector<uint32_t>* lists) {
    std::vector<float> pos[2][3];
    for (int g = 0; g < 2; ++g)
        for (int a = 0; a < 3; ++a) {
            pos[g][a].resize(sc.n[g]);
            for (int i = 0; i < sc.n[g]; ++i) pos[g][a][i] = (float)(double)X[13 * g + a][i];
        }
    const int nw = (int)sc.w[0].size(), nob = (int)sc.o[0].size();
    std::vector<float> wm[3], oc[3];
    for (int a = 0; a < 3; ++a) {
        wm[a].resize(nw);
        for (int l = 0; l < nw; ++l) wm[a][l] = 0.5f * ((float)sc.w[a][l] + (float)sc.w[3 + a][l]);
        oc[a].resize(nob);
        for (int o = 0; o < nob; ++o) oc[a][o] = (float)sc.o[a][o];
    }
    for (int g = 0; g < 2; ++g) {
        const int n = sc.n[g], no = sc.n[1 - g];
        std::vector<float> dist;
        auto run = [&](int which, const float* bx, const float* by, const float* bz, int nb, int k, bool self,
                       float cap) {
            std::vector<uint32_t>& L = lists[4 * g + which];
            L.assign((size_t)k * n, UINT32_MAX);
            dist.assign((size_t)k * n, 0.f);
            knn_brute(n, pos[g][0].data(), pos[g][1].data(), pos[g][2].data(), nb, bx, by, bz, self ? 1 : 0, k,
                      L.data(), dist.data(), nullptr, 0, cap);
        };
        run(0, pos[g][0].data(), pos[g][1].data(), pos[g][2].data(), n, cf.k, true, INFINITY);
        run(1, pos[1 - g][0].data(), pos[1 - g][1].data(), pos[1 - g][2].data(), no, cf.k, false, (float)cf.other_cap);
        run(2, wm[0].data(), wm[1].data(), wm[2].data(), nw, cf.kw, false, INFINITY);
        run(3, oc[0].data(), oc[1].data(), oc[2].data(), nob, cf.ko, false, INFINITY);
    }
}

// Whole trajectory: returns the loss.  mode 1 records tape + lists, 2 replays.
template <class T>
static T c4_run_t(C4Scenario& sc, const C4Params& th, const C4Config& cf, int mode, double* final_state,
                  double* pos_tangent_rms) {
    std::vector<T> X[26], Y[26];
    for (int g = 0; g < 2; ++g)
        for (int c = 0; c < 13; ++c) {
            X[13 * g + c].assign(sc.st[g][c].begin(), sc.st[g][c].end());
            Y[13 * g + c].assign(sc.n[g], T(0));
        }
    if (mode == 1) {
        sc.tape.assign(cf.steps, std::vector<std::vector<uint8_t>>((size_t)sc.n[0] + sc.n[1]));
        sc.knn.assign((size_t)cf.steps * 8, {});
    }
    T loss = T(0);
    for (int t = 0; t < cf.steps; ++t) {
        std::vector<uint32_t>* lists = &sc.knn[(size_t)t * 8];
        if (mode == 1) c4_lists(sc, cf, X, lists);
        loss += c4_step(sc, th, cf, X, Y, lists, sc.tape[t], mode);
        for (int c = 0; c < 26; ++c) std::swap(X[c], Y[c]);
        if constexpr (!std::is_floating_point<T>::value) {
            // RMS over agents of the position's tangent |d p / d theta_j| after
            // step t: how a parameter change's effect on the state grows.
            if (pos_tangent_rms) {
                double acc[kC4Params] = {};
                for (int g = 0; g < 2; ++g)
                    for (int a = 0; a < 3; ++a)
                        for (int i = 0; i < sc.n[g]; ++i)
                            for (int j = 0; j < kC4Params; ++j) acc[j] += X[13 * g + a][i].d[j] * X[13 * g + a][i].d[j];
                for (int j = 0; j < kC4Params; ++j)
                    pos_tangent_rms[(size_t)t * kC4Params + j] = std::sqrt(acc[j] / double(sc.n[0] + sc.n[1]));
            }
        }
    }
    T term = T(0);
    for (int g = 0; g < 2; ++g)
        for (int i = 0; i < sc.n[g]; ++i) {
            const V3<T> d = sub(mk(X[13 * g][i], X[13 * g + 1][i], X[13 * g + 2][i]),
                                mk(T(sc.goal[g][0]), T(sc.goal[g][1]), T(sc.goal[g][2])));
            term += dot(d, d);
        }
    loss += term / T(sc.n[0] + sc.n[1]);
    if (final_state) {
        size_t at = 0;
        for (int g = 0; g < 2; ++g)
            for (int c = 0; c < 13; ++c)
                for (int i = 0; i < sc.n[g]; ++i) final_state[at++] = (double)X[13 * g + c][i];
    }
    return loss;
}
template <class T>
static double c4_run(C4Scenario& sc, const C4Params& th, const C4Config& cf, int mode, double* final_state) {
    return (double)c4_run_t<T>(sc, th, cf, mode, final_state, nullptr);
}

}  // namespace orc

// ---------------------------------------------------------------------------
// C exports for ctypes (tests, smoke check, bench cpu_baseline).
using namespace orc;

extern "C" {

typedef struct {
    float ox, oy, oz, sx, sy, sz;
    int dx, dy, dz;
} or_grid;

int or_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
void or_set_num_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

uint32_t or_cell_of(const or_grid* g, float x, float y, float z) {
    Grid G{g->ox, g->oy, g->oz, g->sx, g->sy, g->sz, g->dx, g->dy, g->dz};
    return cell_of(G, x, y, z);
}

void or_bins_build(const or_grid* g, int n, const float* px, const float* py, const float* pz, uint32_t* key,
                   uint32_t* start, uint32_t* perm, float* spx, float* spy, float* spz) {
    Grid G{g->ox, g->oy, g->oz, g->sx, g->sy, g->sz, g->dx, g->dy, g->dz};
    bins_build(G, n, px, py, pz, key, start, perm, spx, spy, spz);
}

void or_knn_brute(int nA, const float* ax, const float* ay, const float* az, int nB, const float* bx, const float* by,
                  const float* bz, int self, int k, uint32_t* idx, float* dist, const uint32_t* queries, int nq,
                  float cap) {
    knn_brute(nA, ax, ay, az, nB, bx, by, bz, self, k, idx, dist, queries, nq, cap);
}

void or_knn_binned(const or_grid* g, int nA, const float* ax, const float* ay, const float* az, int nB,
                   const float* bx, const float* by, const float* bz, int self, int k, uint32_t* idx, float* dist,
                   float cap) {
    Grid G{g->ox, g->oy, g->oz, g->sx, g->sy, g->sz, g->dx, g->dy, g->dz};
    const int C = g->dx * g->dy * g->dz;
    std::vector<uint32_t> key(nB), start((size_t)C + 1), perm(nB);
    std::vector<float> sx(nB), sy(nB), sz(nB);
    bins_build(G, nB, bx, by, bz, key.data(), start.data(), perm.data(), sx.data(), sy.data(), sz.data());
    knn_binned(G, nA, ax, ay, az, nB, sx.data(), sy.data(), sz.data(), perm.data(), start.data(), self, k, idx, dist,
               cap);
}

#define OR_AGENT_API(T, SUF)                                                                                        \
    typedef Agents<T> or_agents_##SUF;                                                                              \
    void or_steer_flock_##SUF(int which, int mode, or_agents_##SUF* A, int nA, const uint32_t* knn, int k,          \
                              const or_agents_##SUF* B, int nB, T min_d, T max_d, T cos_max, T w, uint32_t mask) { \
        steer_flock_one<T>(which, mode, *A, nA, knn, k, *B, nB, min_d, max_d, cos_max, w, mask);                    \
    }                                                                                                               \
    void or_steer_seek_##SUF(int mode, or_agents_##SUF* A, int nA, const T* tx, const T* ty, const T* tz,           \
                             int per_agent, T w, uint32_t mask) {                                                   \
        steer_seek<T>(mode, *A, nA, tx, ty, tz, per_agent, w, mask);                                                \
    }                                                                                                               \
    void or_avoid_obstacles_##SUF(or_agents_##SUF* A, int nA, const uint32_t* knn, int k, int nO, const T* ox,      \
                                  const T* oy, const T* oz, const T* orad, T min_time, T w, uint32_t mask) {        \
        avoid_obstacles<T>(*A, nA, knn, k, nO, ox, oy, oz, orad, min_time, w, mask);                                \
    }                                                                                                               \
    void or_avoid_walls_##SUF(or_agents_##SUF* A, int nA, const uint32_t* knn, int k, int nW, const T* sx,          \
                              const T* sy, const T* sz, const T* ex, const T* ey, const T* ez, const T* nx,         \
                              const T* ny, const T* nz, T min_time, T w, uint32_t mask) {                           \
        avoid_walls<T>(*A, nA, knn, k, nW, sx, sy, sz, ex, ey, ez, nx, ny, nz, min_time, w, mask);                  \
    }                                                                                                               \
    void or_update_##SUF(int mode, or_agents_##SUF* A, int nA, T dt) { update<T>(mode, *A, nA, dt); }               \
    void or_wrap_world_##SUF(int shape, or_agents_##SUF* A, int nA, const T* c, const T* s) {                       \
        wrap_world<T>(shape, *A, nA, c, s);                                                                         \
    }                                                                                                               \
    void or_update_anti_penetration_##SUF(or_agents_##SUF* A, int nA, const uint32_t* knn, int k, int nW,           \
                                          const T* sx, const T* sy, const T* sz, const T* ex, const T* ey,          \
                                          const T* ez, const T* nx, const T* ny, const T* nz, T dt, int closest) {  \
        update_anti_penetration<T>(*A, nA, knn, k, nW, sx, sy, sz, ex, ey, ez, nx, ny, nz, dt, closest);            \
    }                                                                                                               \
    void or_anti_penetration_agents_##SUF(or_agents_##SUF* A, int nA, const uint32_t* knn, const float* dist,       \
                                          const or_agents_##SUF* B, int nB, uint32_t mask) {                        \
        anti_penetration_agents<T>(*A, nA, knn, dist, *B, nB, mask);                                                \
    }                                                                                                               \
    T or_path_init_##SUF(int n, const T* pts_in, int cyclic, T* pts, T* nrm, T* lens) {                             \
        return path_init<T>(n, pts_in, cyclic, pts, nrm, lens);                                                     \
    }                                                                                                               \
    void or_follow_path_##SUF(or_agents_##SUF* A, int nA, int m, const T* pts, const T* nrm, const T* lens,         \
                              T radius, T total, int cyclic, T prediction_time, T w, uint32_t mask) {               \
        const Path<T> P_{m, pts, nrm, lens, radius, total, cyclic};                                                 \
        follow_path<T>(*A, nA, P_, prediction_time, w, mask);                                                       \
    }                                                                                                               \
    void or_avoid_neighbors_##SUF(int mode, or_agents_##SUF* A, int nA, const uint32_t* knn, const float* dist,    \
                                  int k, const or_agents_##SUF* B, int nB, T min_time, T min_sep, int avoid_close,  \
                                  T w, uint32_t mask) {                                                             \
        avoid_neighbors<T>(mode, *A, nA, knn, dist, k, *B, nB, min_time, min_sep, avoid_close, w, mask);            \
    }                                                                                                               \
    void or_path_point_at_##SUF(int m, const T* pts, const T* nrm, const T* lens, T total, int cyclic, T dist,      \
                                T* out) {                                                                           \
        const Path<T> P_{m, pts, nrm, lens, T(0), total, cyclic};                                                   \
        const V3<T> r = path_point_at<T>(P_, dist);                                                                 \
        out[0] = r.x; out[1] = r.y; out[2] = r.z;                                                                   \
    }                                                                                                               \
    T or_path_distance_of_##SUF(int m, const T* pts, const T* nrm, const T* lens, const T* p) {                     \
        const Path<T> P_{m, pts, nrm, lens, T(0), T(0), 0};                                                         \
        return path_distance_of<T>(P_, mk(p[0], p[1], p[2]));                                                       \
    }

OR_AGENT_API(float, f32)
OR_AGENT_API(double, f64)


static void c4_config_of(const double* cfgd, const int* cfgi, C4Config& cf) {
    cf.sep_min = cfgd[0]; cf.sep_max = cfgd[1]; cf.sep_cos = cfgd[2];
    cf.ali_min = cfgd[3]; cf.ali_max = cfgd[4]; cf.ali_cos = cfgd[5];
    cf.coh_min = cfgd[6]; cf.coh_max = cfgd[7]; cf.coh_cos = cfgd[8];
    cf.w_seek = cfgd[9]; cf.mt_walls = cfgd[10]; cf.mt_obs = cfgd[11]; cf.max_force = cfgd[12];
    cf.mass = cfgd[13]; cf.dt = cfgd[14]; cf.pen_w = cfgd[15]; cf.other_cap = cfgd[16];
    cf.sep_mask = (uint32_t)cfgi[0]; cf.ali_mask = (uint32_t)cfgi[1]; cf.coh_mask = (uint32_t)cfgi[2];
    cf.seek_mask = (uint32_t)cfgi[3]; cf.k = cfgi[4]; cf.kw = cfgi[5]; cf.ko = cfgi[6]; cf.steps = cfgi[7];
}

// Calibration model (gradient checks).  or_c4_create copies the scenario:
// state[2][13][n_g] (p f speed side up), goals[2][3], walls[9][nw], obstacles[4][no].
void* or_c4_create(int n0, int n1, const double* state0, const double* state1, const double* goals, int nw,
                   const double* walls, int no, const double* obstacles) {
    auto* sc = new C4Scenario();
    sc->n[0] = n0;
    sc->n[1] = n1;
    const double* st[2] = {state0, state1};
    for (int g = 0; g < 2; ++g)
        for (int c = 0; c < 13; ++c) sc->st[g][c].assign(st[g] + (size_t)c * sc->n[g], st[g] + (size_t)(c + 1) * sc->n[g]);
    for (int g = 0; g < 2; ++g)
        for (int a = 0; a < 3; ++a) sc->goal[g][a] = goals[g * 3 + a];
    for (int j = 0; j < 9; ++j) sc->w[j].assign(walls + (size_t)j * nw, walls + (size_t)(j + 1) * nw);
    for (int j = 0; j < 4; ++j) sc->o[j].assign(obstacles + (size_t)j * no, obstacles + (size_t)(j + 1) * no);
    return sc;
}
void or_c4_destroy(void* h) { delete (C4Scenario*)h; }
// params: 14 doubles (gs_calib_params order); config: see C4Config (doubles then ints).
// precision 32: the model in fp32 (the GPU's arithmetic, for recording the
// branch decisions the GPU takes); 64: in FP64.
double or_c4_run(void* h, const double* params, const double* cfgd, const int* cfgi, int mode, double* final_state,
                 int precision) {
    C4Params th;
    memcpy(&th, params, sizeof th);
    C4Config cf;
    c4_config_of(cfgd, cfgi, cf);
    if (precision == 32) return c4_run<float>(*(C4Scenario*)h, th, cf, mode, final_state);
    return c4_run<double>(*(C4Scenario*)h, th, cf, mode, final_state);
}
// The analytic gradient of the taped model (forward-mode tangents, replaying
// the tape and KNN lists of the last recording run): grad[14] = d loss /
// d params.  primal_precision 32: along the fp32 trajectory (the GPU's);
// 64: along the FP64 trajectory (the finite differences' function).
// pos_tangent_rms (optional, [steps][14]): RMS over agents of the position's
// tangent after each step.
double or_c4_tangent(void* h, const double* params, const double* cfgd, const int* cfgi, int primal_precision,
                     double* grad, double* pos_tangent_rms) {
    C4Params th;
    memcpy(&th, params, sizeof th);
    C4Config cf;
    c4_config_of(cfgd, cfgi, cf);
    if (primal_precision == 32) {
        const Tan<float> l = c4_run_t<Tan<float>>(*(C4Scenario*)h, th, cf, 2, nullptr, pos_tangent_rms);
        for (int j = 0; j < kC4Params; ++j) grad[j] = l.d[j];
        return double(l.v);
    }
    const Tan<double> l = c4_run_t<Tan<double>>(*(C4Scenario*)h, th, cf, 2, nullptr, pos_tangent_rms);
    for (int j = 0; j < kC4Params; ++j) grad[j] = l.d[j];
    return l.v;
}

float or_pow20(float r) { return pow20(r); }
void or_limit_max_deviation_angle(const float* src, float cosc, const float* basis, float* out) {
    V3<float> r = limit_max_deviation_angle(mk(src[0], src[1], src[2]), cosc, mk(basis[0], basis[1], basis[2]));
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
void or_truncate_length(const float* v, float m, float* out) {
    V3<float> r = truncate_length(mk(v[0], v[1], v[2]), m);
    out[0] = r.x; out[1] = r.y; out[2] = r.z;
}
int or_lines_intersect(const float* sa, const float* ea, const float* sb, const float* eb, float* ip) {
    V3<float> p{0, 0, 0};
    const int r = lines_intersect(mk(sa[0], sa[1], sa[2]), mk(ea[0], ea[1], ea[2]), mk(sb[0], sb[1], sb[2]),
                                  mk(eb[0], eb[1], eb[2]), p);
    ip[0] = p.x; ip[1] = p.y; ip[2] = p.z;
    return r;
}

}  // extern "C"
