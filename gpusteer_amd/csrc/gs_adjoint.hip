// Differentiable calibration of crowd scenarios: a fused forward step of
// steering + integration, its hand-written adjoint (reverse mode), and a
// checkpointed driver (gs_calib_* in include/gpusteer.h).
//
// Model, per agent i of group g at step t (Jacobi: every neighbour read is the
// step-t state), the reference CUDA kernels' formulas (GS_MODE_GPUREF):
//   walls      SteerToAvoidWallsCUDAKernel (CUDA/AvoidWallsCUDAKernel.cu:67-245),
//              including the anti-penetration move          weight w_wall[g]
//   obstacles  SteerToAvoidObstaclesKernel (AvoidObstaclesCUDAKernel.cu:136-253) w_obs[g]
//   separation from the own group and from the other group
//              (SteerForSeparationCUDAKernel.cu)             w_sep[g]
//   alignment, cohesion with the own group                   w_ali[g], w_coh[g]
//   seek the group's goal (SteerForSeekCUDAKernel.cu)        fixed weight
//   integrate  UpdateCUDAKernel (UpdateCUDAKernel.cu:78-108), max speed vmax[g]
// with agent radius radius[g].  KNN lists (own group k, other group capped,
// walls, obstacles) are computed from the step-t positions and held fixed
// under differentiation (they are piecewise constant), as are every other
// branch of the step (neighbourhood membership, wall / obstacle hits,
// truncation): the adjoint is the exact derivative of the branch the forward
// took.  The up vector is treated as constant (exact for motion in the plane
// normal to it, the crowd's case).
//
// Loss = mean over scenarios of [ mean_i |p_T,i - goal_i|^2
//        + pen_w * sum_t sum_i sum_{j in lists(i)} max(0, r_i + r_j - |p_j - p_i|)^2 ].
//
// Checkpoints every `checkpoint_every` steps; the backward pass recomputes
// each segment forward (KNN included) from its checkpoint, keeping the
// segment's states and lists, then runs the adjoint kernel backwards through
// it.  Every sum is taken in a fixed order, so the loss and the gradient are
// identical, bit for bit, from run to run: a neighbour's contribution to an
// agent's adjoint goes to its own slot of the step's reverse lists and the
// backward of the step before adds an agent's slots in (query, slot) order;
// loss and parameter gradient are summed per block (waves in order), and the
// per-block sums are added in block order at the end.
#include <algorithm>
#include <vector>

#include "gs_ensemble.hpp"

namespace gs {

constexpr int kState = 13;   // p(3) f(3) speed side(3) up(3)
constexpr int kAdj = 10;     // p(3) f(3) speed side(3)

struct CState {
    float* a[kState];
};
struct CAdj {
    float* a[kAdj];
};

struct CalibDev {
    const EnsPop* pops;
    const EnsScen* scens;
    const uint32_t* block_pop;
    gs_calib_params th;
    FlockTerm sep, ali, coh;   // distances and cone; weights from th
    float w_seek;
    uint32_t mask_seek, mask_sep;
    float mt_walls, mt_obs, max_force, mass, dt, pen_w;
    uint32_t k, N;
    float inv_S;               // 1 / number of scenarios
    // geometry (concatenated over scenarios)
    const float *wsx, *wsy, *wsz, *wex, *wey, *wez, *wnx, *wny, *wnz;
    const float *ox, *oy, *oz, *orad;
};

__device__ __forceinline__ f3 ldS(const CState& S, int base, uint32_t i) {
    return mk3(S.a[base][i], S.a[base + 1][i], S.a[base + 2][i]);
}

// ---- lines_intersect (XZ plane) with its adjoint --------------------------
struct LIRec {
    bool hit, coincident;
    float ua, denom;
};
__device__ __forceinline__ bool li_fwd(f3 sa, f3 ea, f3 sb, f3 eb, f3& ip, LIRec& r) {
    const float eps = 0.001f;
    const float denom = ((eb.z - sb.z) * (ea.x - sa.x)) - ((eb.x - sb.x) * (ea.z - sa.z));
    const float numera = ((eb.x - sb.x) * (sa.z - sb.z)) - ((eb.z - sb.z) * (sa.x - sb.x));
    const float numerb = ((ea.x - sa.x) * (sa.z - sb.z)) - ((ea.z - sa.z) * (sa.x - sb.x));
    r.coincident = false;
    r.denom = denom;
    r.ua = 0.f;
    if (fabsf(denom) < eps) {
        if (fabsf(numera) < eps && fabsf(numerb) < eps) {
            ip = sa;
            r.coincident = true;
            r.hit = true;
            return true;
        }
        r.hit = false;
        return false;
    }
    const float ua = numera / denom;
    const float ub = numerb / denom;
    r.ua = ua;
    if (ua >= 0.f && ua <= 1.f && ub >= 0.f && ub <= 1.f) {
        ip = mk3(sa.x + ua * (ea.x - sa.x), 0.f, sa.z + ua * (ea.z - sa.z));
        r.hit = true;
        return true;
    }
    r.hit = false;
    return false;
}
// Given d(loss)/d(ip), accumulate d/d(sa) and d/d(ea) (the wall is constant).
__device__ __forceinline__ void li_adj(f3 sa, f3 ea, f3 sb, f3 eb, const LIRec& r, f3 ipb, f3& sab, f3& eab) {
    if (r.coincident) {
        sab = sab + ipb;
        return;
    }
    const float dAx = ea.x - sa.x, dAz = ea.z - sa.z;
    const float bx = eb.x - sb.x, bz = eb.z - sb.z;
    const float ub = ipb.x * dAx + ipb.z * dAz;
    float dAxb = ipb.x * r.ua, dAzb = ipb.z * r.ua;
    const float numerab = ub / r.denom;
    const float denomb = -ub * r.ua / r.denom;
    const float czb = bx * numerab, cxb = -bz * numerab;
    dAxb += bz * denomb;
    dAzb += -bx * denomb;
    sab.x += ipb.x + cxb - dAxb;
    sab.z += ipb.z + czb - dAzb;
    eab.x += dAxb;
    eab.z += dAzb;
}

__device__ __forceinline__ f3 normalize_adj(f3 x, f3 yb) {
    const float l = len3(x);
    if (!(l > 0.f)) return mk3(0.f, 0.f, 0.f);
    const f3 y = x / l;
    return (yb - y * dot3(y, yb)) / l;
}
// y = truncate(v, m): returns d/dv, adds d/dm to mb.
__device__ __forceinline__ f3 truncate_adj(f3 v, float m, bool trunc, f3 yb, float& mb) {
    if (!trunc) return yb;
    const float l = len3(v);
    const f3 u = v / l;
    mb += dot3(u, yb);
    return (yb - u * dot3(u, yb)) * (m / l);
}
__device__ __forceinline__ f3 cross_adj_a(f3 b, f3 yb) { return cross3(b, yb); }  // y = a x b: d/da

// ---- per-agent forward with the branch record ------------------------------
struct C4Rec {
    // walls: last update event (0 none, 1 touching segment, 2 feeler) and
    // whether any touching event happened
    int w_event, w_touch;
    uint32_t w_wall;
    int w_feeler;
    LIRec w_li;
    // obstacles
    int o_apply;
    uint32_t o_idx;
    // flocking membership masks (bit j: list slot j counted) and counts
    uint32_t m_ss, m_so, m_al, m_co;
    int on_ss, on_so, on_al, on_co, on_seek;
    // integrate
    int tf, tv, moved;
};

struct C4In {
    f3 p, f, sd, up;
    float s, r, vmax;
};

template <bool REC>
__device__ __forceinline__ void walls_fwd(const CalibDev& D, const EnsScen& S, uint32_t N, uint32_t i,
                                          const uint32_t* __restrict__ kw, f3& p, f3& f, f3 sd, float s, float r,
                                          float w, f3& G, uint32_t& bits, C4Rec& rec) {
    const float fl = 4.f * r;
    const f3 vel = f * s;
    f3 fe[3];
    fe[0] = p + vel * D.mt_walls;
    fe[1] = p + (sd * fl + f * fl);
    fe[2] = p + ((-sd) * fl + f * fl);
    int feeler = 0;
    float nearest_dist = FLT_MAX;
    uint32_t nearest_line = UINT32_MAX;
    f3 nearest_ip = mk3(0.f, 0.f, 0.f), nearest_normal = nearest_ip;
    bool touching = false;
    rec.w_event = 0;
    rec.w_touch = 0;
    for (uint32_t j = 0; j < S.k_walls; ++j) {
        const uint32_t l = kw[(size_t)j * N + i];
        if (l >= S.nw) break;
        const uint32_t g = S.w0 + l;
        const f3 ln = mk3(D.wnx[g], D.wny[g], D.wnz[g]);
        const f3 ls = mk3(D.wsx[g], D.wsy[g], D.wsz[g]);
        const f3 le = mk3(D.wex[g], D.wey[g], D.wez[g]);
        f3 ip;
        LIRec lr;
        if (li_fwd(p, p + (-ln) * r, ls, le, ip, lr)) {
            touching = true;
            nearest_normal = ln;
            nearest_dist = len3(p - ip);
            rec.w_event = 1;
            rec.w_wall = l;
            rec.w_li = lr;
        }
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            if (li_fwd(p, fe[q], ls, le, ip, lr)) {
                const float dist = len3(p - ip);
                if (dist < nearest_dist) {
                    nearest_dist = dist;
                    nearest_line = l;
                    nearest_ip = ip;
                    feeler = q;
                    nearest_normal = ln;
                    rec.w_event = 2;
                    rec.w_wall = l;
                    rec.w_feeler = q;
                    rec.w_li = lr;
                }
            }
        }
    }
    rec.w_touch = touching ? 1 : 0;
    f3 steering = mk3(0.f, 0.f, 0.f);
    if (touching) {
        const float pen = r - nearest_dist;
        p = p + nearest_normal * pen;
        f = normalize3(perp3(f, nearest_normal));
        bits |= GS_BIT_ANTI_PENETRATION_WALL;
    } else if (nearest_line != UINT32_MAX) {
        const f3 fsel = feeler == 0 ? fe[0] : (feeler == 1 ? fe[1] : fe[2]);
        steering = nearest_normal * len3(fsel - nearest_ip);
        bits |= GS_BIT_AVOID_WALLS;
    }
    steering = steering * w;
    G = steering + G;
}

// GPUREF neighbourhood (CUDA/FlockingCommon.cuh:4-30).
__device__ __forceinline__ bool hood(f3 off, float d2, f3 fwd, const FlockTerm& t) {
    const float d = sqrtf(d2);
    if (d < t.min_d) return true;
    if (d > t.max_d) return false;
    return dot3(fwd, off / d) > t.cos_max;
}

// One agent's step.  Returns the new state; fills rec.
__device__ __forceinline__ void c4_fwd_agent(const CalibDev& D, const EnsPop& P, const EnsScen& S, uint32_t i,
                                             const CState& X, const uint32_t* __restrict__ kself,
                                             const uint32_t* __restrict__ kother, const uint32_t* __restrict__ kw,
                                             const uint32_t* __restrict__ ko, C4In& in, f3& p1, f3& f1, f3& G,
                                             f3& pn, f3& fn, float& sn, f3& sdn, f3& upn, C4Rec& rec) {
    const int g = P.group;
    in.p = ldS(X, 0, i);
    in.f = ldS(X, 3, i);
    in.s = X.a[6][i];
    in.sd = ldS(X, 7, i);
    in.up = ldS(X, 10, i);
    in.r = D.th.radius[g];
    in.vmax = D.th.max_speed[g];
    uint32_t bits = 0u;
    G = mk3(0.f, 0.f, 0.f);
    p1 = in.p;
    f1 = in.f;
    walls_fwd<true>(D, S, D.N, i, kw, p1, f1, in.sd, in.s, in.r, D.th.w_wall[g], G, bits, rec);
    // obstacles (CUDA/AvoidObstaclesCUDAKernel.cu:136-253)
    {
        const float min_dist = D.mt_obs * in.s;
        bool n_hit = false;
        float n_dist = 0.f;
        uint32_t n_o = UINT32_MAX;
        for (uint32_t j = 0; j < S.k_obs; ++j) {
            const uint32_t o = ko[(size_t)j * D.N + i];
            if (o >= S.no) break;
            const uint32_t go_ = S.o0 + o;
            const f3 op = mk3(D.ox[go_], D.oy[go_], D.oz[go_]);
            const float orr = D.orad[go_];
            bool x_hit = false;
            float x_dist = 0.f;
            const f3 go = op - p1;
            const f3 lc = mk3(dot3(go, in.sd), dot3(go, in.up), dot3(go, f1));
            const float b = -2.f * lc.z;
            const float c = ((lc.x * lc.x + lc.y * lc.y) + lc.z * lc.z) - (orr + in.r) * (orr + in.r);
            const float d = (b * b) - (4.f * c);
            if (!(d < 0.f)) {
                const float sq = sqrtf(d);
                const float pp = (-b + sq) / 2.f;
                const float qq = (-b - sq) / 2.f;
                if (!((pp < 0.f) && (qq < 0.f))) {
                    x_hit = true;
                    x_dist = ((pp > 0.f) && (qq > 0.f)) ? ((pp < qq) ? pp : qq) : ((pp > 0.f) ? pp : qq);
                }
            }
            if ((n_hit == false) || ((x_hit != false) && (x_dist < n_dist))) {
                n_hit = x_hit;
                n_dist = x_dist;
                n_o = go_;
            }
        }
        rec.o_apply = (n_hit && (n_dist < min_dist)) ? 1 : 0;
        rec.o_idx = n_o;
        f3 fo = mk3(0.f, 0.f, 0.f);
        if (rec.o_apply) fo = perp3(p1 - mk3(D.ox[n_o], D.oy[n_o], D.oz[n_o]), f1);
        apply(G, bits, fo, D.th.w_obs[g], GS_BIT_AVOID_OBSTACLES);
    }
    // separation: own group, then the other group
    const uint32_t* lists[2] = {kself, kother};
    for (int w = 0; w < 2; ++w) {
        uint32_t m = 0u;
        int on = 0;
        f3 acc = mk3(0.f, 0.f, 0.f);
        int cnt = 0;
        if (!(bits & D.mask_sep)) {
            on = 1;
            for (uint32_t j = 0; j < D.k; ++j) {
                const uint32_t b = lists[w][(size_t)j * D.N + i];
                if (b >= D.N) break;
                const f3 off = ldS(X, 0, b) - p1;
                const float d2 = dot3(off, off);
                if (hood(off, d2, f1, D.sep)) {
                    acc = acc + off / -d2;
                    ++cnt;
                    m |= 1u << j;
                }
            }
            if (cnt > 0) acc = normalize3(acc / (float)cnt);
            apply(G, bits, acc, D.th.w_sep[g], GS_BIT_SEPARATION);
        }
        if (w == 0) {
            rec.m_ss = m;
            rec.on_ss = on;
        } else {
            rec.m_so = m;
            rec.on_so = on;
        }
    }
    // alignment
    {
        uint32_t m = 0u;
        f3 acc = mk3(0.f, 0.f, 0.f);
        int cnt = 0;
        rec.on_al = !(bits & D.ali.mask) ? 1 : 0;
        if (rec.on_al) {
            for (uint32_t j = 0; j < D.k; ++j) {
                const uint32_t b = kself[(size_t)j * D.N + i];
                if (b >= D.N) break;
                const f3 off = ldS(X, 0, b) - p1;
                if (hood(off, dot3(off, off), f1, D.ali)) {
                    acc = acc + ldS(X, 3, b);
                    ++cnt;
                    m |= 1u << j;
                }
            }
            if (cnt > 0) acc = normalize3((acc / (float)cnt) - f1);
            apply(G, bits, acc, D.th.w_ali[g], GS_BIT_ALIGNMENT);
        }
        rec.m_al = m;
    }
    // cohesion
    {
        uint32_t m = 0u;
        f3 acc = mk3(0.f, 0.f, 0.f);
        int cnt = 0;
        rec.on_co = !(bits & D.coh.mask) ? 1 : 0;
        if (rec.on_co) {
            for (uint32_t j = 0; j < D.k; ++j) {
                const uint32_t b = kself[(size_t)j * D.N + i];
                if (b >= D.N) break;
                const f3 pb = ldS(X, 0, b);
                const f3 off = pb - p1;
                if (hood(off, dot3(off, off), f1, D.coh)) {
                    acc = acc + pb;
                    ++cnt;
                    m |= 1u << j;
                }
            }
            if (cnt > 0) acc = normalize3((acc / (float)cnt) - p1);
            apply(G, bits, acc, D.th.w_coh[g], GS_BIT_COHESION);
        }
        rec.m_co = m;
    }
    // seek (GPUREF: normalize(target - position - forward))
    rec.on_seek = !(bits & D.mask_seek) ? 1 : 0;
    if (rec.on_seek) {
        const f3 goal = mk3(S.goal[g][0], S.goal[g][1], S.goal[g][2]);
        apply(G, bits, normalize3((goal - p1) - f1), D.w_seek, GS_BIT_SEEK);
    }
    // integrate (UpdateCUDAKernel)
    const float l2 = dot3(G, G);
    rec.tf = l2 > D.max_force * D.max_force ? 1 : 0;
    const f3 Ft = truncate3(G, D.max_force);
    const f3 acc = Ft / D.mass;
    const f3 v1 = f1 * in.s + acc * D.dt;
    rec.tv = dot3(v1, v1) > in.vmax * in.vmax ? 1 : 0;
    const f3 v = truncate3(v1, in.vmax);
    sn = len3(v);
    fn = f1;
    sdn = in.sd;
    upn = in.up;
    rec.moved = sn > 0.f ? 1 : 0;
    if (rec.moved) {
        fn = v / sn;
        sdn = normalize3(cross3(fn, in.up));
        upn = cross3(sdn, fn);
    }
    pn = p1 + v * D.dt;
}

__device__ __forceinline__ uint32_t agent_block_pop(const CalibDev& D, uint32_t blk) { return D.block_pop[blk]; }

// Penetration term of step t for agent i: pairs with its own-group and
// other-group lists, positions of step t.
__device__ __forceinline__ float pen_loss(const CalibDev& D, const EnsPop& P, uint32_t i, const CState& X,
                                          const uint32_t* __restrict__ kself, const uint32_t* __restrict__ kother) {
    const f3 p = ldS(X, 0, i);
    const int g = P.group;
    float L = 0.f;
    for (int w = 0; w < 2; ++w) {
        const uint32_t* lst = w == 0 ? kself : kother;
        const float R = D.th.radius[g] + D.th.radius[w == 0 ? g : 1 - g];
        for (uint32_t j = 0; j < D.k; ++j) {
            const uint32_t b = lst[(size_t)j * D.N + i];
            if (b >= D.N) break;
            const float e = R - len3(ldS(X, 0, b) - p);
            if (e > 0.f) L += e * e;
        }
    }
    return L * D.pen_w;
}

// Adds the block's sum of v (waves in order) to *part, which the block owns.
__device__ inline void block_add(float v, double* __restrict__ part, double* lds) {
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = (double)v;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int w = 0; w < kBlock / 64; ++w) t += lds[w];
        *part += t;
    }
    __syncthreads();
}

// out[c] = sum over rows r (in order, then a fixed tree) of part[r * cols + c];
// one block per column.
__global__ __launch_bounds__(kBlock) void c4_reduce_rows(const double* __restrict__ part, uint32_t rows,
                                                         uint32_t cols, double* __restrict__ out) {
    __shared__ double lds[kBlock];
    const uint32_t c = blockIdx.x;
    double t = 0.0;
    for (uint32_t r = threadIdx.x; r < rows; r += kBlock) t += part[(size_t)r * cols + c];
    lds[threadIdx.x] = t;
    __syncthreads();
    for (int h = kBlock / 2; h > 0; h >>= 1) {
        if ((int)threadIdx.x < h) lds[threadIdx.x] += lds[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = lds[0];
}

// Reverse lists of one step: for every agent b, the entries (query i, list
// w, slot j) of the own-group (w = 0) and other-group (w = 1) lists that name
// b, as the segment [rstart[b], rstart[b+1]) ordered by entry code.  The
// backward writes entry e's contribution to slot slot_of[code(e)]; the
// backward of the step before adds agent b's segment in order.
__device__ inline uint32_t rev_code(uint32_t i, uint32_t w, uint32_t j, uint32_t k) {
    return (i * 2u + w) * k + j;
}

// Counts the entries naming each agent; an entry's arrival index among them
// goes to arrival[code] (the order of arrival varies, the count does not).
__global__ __launch_bounds__(kBlock) void c4_rev_count(CalibDev D, const uint32_t* __restrict__ kself,
                                                       const uint32_t* __restrict__ kother,
                                                       uint32_t* __restrict__ cnt, uint32_t* __restrict__ arrival) {
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t i = blk * blockDim.x + threadIdx.x;
    const EnsPop& P = D.pops[D.block_pop[blk]];
    if (i >= P.a0 + P.n) return;
    for (int w = 0; w < 2; ++w) {
        const uint32_t* lst = w == 0 ? kself : kother;
        for (uint32_t j = 0; j < D.k; ++j) {
            const uint32_t b = lst[(size_t)j * D.N + i];
            if (b >= D.N) break;
            arrival[rev_code(i, (uint32_t)w, j, D.k)] = atomicAdd(cnt + b, 1u);
        }
    }
}

// Places each entry's code in its agent's segment at its arrival index, and
// the agent beside it (rev_agent).
__global__ __launch_bounds__(kBlock) void c4_rev_fill(CalibDev D, const uint32_t* __restrict__ kself,
                                                      const uint32_t* __restrict__ kother,
                                                      const uint32_t* __restrict__ rstart,
                                                      const uint32_t* __restrict__ arrival, uint32_t* __restrict__ rev,
                                                      uint32_t* __restrict__ rev_agent) {
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t i = blk * blockDim.x + threadIdx.x;
    const EnsPop& P = D.pops[D.block_pop[blk]];
    if (i >= P.a0 + P.n) return;
    for (int w = 0; w < 2; ++w) {
        const uint32_t* lst = w == 0 ? kself : kother;
        for (uint32_t j = 0; j < D.k; ++j) {
            const uint32_t b = lst[(size_t)j * D.N + i];
            if (b >= D.N) break;
            const uint32_t c = rev_code(i, (uint32_t)w, j, D.k);
            const uint32_t pos = rstart[b] + arrival[c];
            rev[pos] = c;
            rev_agent[pos] = b;
        }
    }
}

// One thread per filled slot s: slot_of[code] = the code's rank in its
// agent's segment (the segments hold each code once), so the slot order is
// the code order, whatever the arrival.
__global__ __launch_bounds__(kBlock) void c4_rev_rank(uint32_t N, const uint32_t* __restrict__ rstart,
                                                      const uint32_t* __restrict__ rev,
                                                      const uint32_t* __restrict__ rev_agent,
                                                      uint32_t* __restrict__ slot_of) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= rstart[N]) return;
    const uint32_t b = rev_agent[s];
    const uint32_t s0 = rstart[b], s1 = rstart[b + 1];
    const uint32_t c = rev[s];
    uint32_t r = 0;
    for (uint32_t t = s0; t < s1; ++t) r += rev[t] < c ? 1u : 0u;
    slot_of[c] = s0 + r;
}

__global__ __launch_bounds__(kBlock) void c4_forward(CalibDev D, CState X, CState Y,
                                                     const uint32_t* __restrict__ kself,
                                                     const uint32_t* __restrict__ kother,
                                                     const uint32_t* __restrict__ kw, const uint32_t* __restrict__ ko,
                                                     double* __restrict__ loss_part) {
    __shared__ double lds[kBlock / 64];
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t i = blk * blockDim.x + threadIdx.x;
    const EnsPop& P = D.pops[D.block_pop[blk]];
    float lp = 0.f;
    if (i < P.a0 + P.n) {
        const EnsScen& S = D.scens[P.scen];
        C4In in;
        C4Rec rec;
        f3 p1, f1, G, pn, fn, sdn, upn;
        float sn;
        c4_fwd_agent(D, P, S, i, X, kself, kother, kw, ko, in, p1, f1, G, pn, fn, sn, sdn, upn, rec);
        Y.a[0][i] = pn.x; Y.a[1][i] = pn.y; Y.a[2][i] = pn.z;
        Y.a[3][i] = fn.x; Y.a[4][i] = fn.y; Y.a[5][i] = fn.z;
        Y.a[6][i] = sn;
        Y.a[7][i] = sdn.x; Y.a[8][i] = sdn.y; Y.a[9][i] = sdn.z;
        Y.a[10][i] = upn.x; Y.a[11][i] = upn.y; Y.a[12][i] = upn.z;
        lp = pen_loss(D, P, i, X, kself, kother) * D.inv_S;
    }
    if (loss_part) block_add(lp, loss_part + blockIdx.x, lds);
}

// Terminal term: loss += mean over agents of |p - goal|^2 (per scenario, / S);
// writes the adjoint seed dL/dp_T and zeroes the other adjoint components.
__global__ __launch_bounds__(kBlock) void c4_terminal(CalibDev D, CState X, CAdj A, double* __restrict__ loss_part) {
    __shared__ double lds[kBlock / 64];
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t i = blk * blockDim.x + threadIdx.x;
    const EnsPop& P = D.pops[D.block_pop[blk]];
    float l = 0.f;
#pragma unroll
    for (int c = 0; c < kAdj; ++c) A.a[c][i] = 0.f;
    if (i < P.a0 + P.n) {
        const EnsScen& S = D.scens[P.scen];
        const EnsPop& Q = D.pops[P.other];
        const float inv_n = 1.f / (float)(P.n + Q.n);
        const f3 goal = mk3(S.goal[P.group][0], S.goal[P.group][1], S.goal[P.group][2]);
        const f3 d = ldS(X, 0, i) - goal;
        l = dot3(d, d) * inv_n * D.inv_S;
        const f3 gp = d * (2.f * inv_n * D.inv_S);
        A.a[0][i] = gp.x;
        A.a[1][i] = gp.y;
        A.a[2][i] = gp.z;
    }
    if (loss_part) block_add(l, loss_part + blockIdx.x, lds);
}

// Parameter-gradient slots (gs_calib_params order).
enum { TW_SEP = 0, TW_ALI = 2, TW_COH = 4, TW_OBS = 6, TW_WALL = 8, TR = 10, TV = 12, kTheta = 14 };

// Adjoint of one step.  In: the step-t state X, the lists of step t, the
// adjoint of the step-(t+1) state as own_in plus the neighbour terms
// contrib_in summed over each agent's segment of rstart_in (the reverse lists
// of step t+1; none when rstart_in is NULL).  Out: the adjoint of the step-t
// state as own_out (each agent's own terms) and contrib_out (one slot per
// neighbour entry, slot_out[code]: p and f terms as two float4), and the
// block's parameter-gradient sums added to its row of grad_part.
__global__ __launch_bounds__(kBlock) void c4_backward(CalibDev D, CState X, const uint32_t* __restrict__ kself,
                                                      const uint32_t* __restrict__ kother,
                                                      const uint32_t* __restrict__ kw,
                                                      const uint32_t* __restrict__ ko, CAdj own_in,
                                                      const uint32_t* __restrict__ rstart_in,
                                                      const float4* __restrict__ contrib_in, CAdj own_out,
                                                      const uint32_t* __restrict__ slot_out,
                                                      float4* __restrict__ contrib_out,
                                                      double* __restrict__ grad_part) {
    __shared__ double lds[kBlock / 64];
    const uint32_t blk = xcd_block(blockIdx.x, gridDim.x);
    const uint32_t i = blk * blockDim.x + threadIdx.x;
    const EnsPop& P = D.pops[D.block_pop[blk]];
    float tg[kTheta];
#pragma unroll
    for (int c = 0; c < kTheta; ++c) tg[c] = 0.f;
    if (i < P.a0 + P.n) {
        const EnsScen& S = D.scens[P.scen];
        const int g = P.group;
        C4In in;
        C4Rec rec;
        f3 p1, f1, G, pn, fn, sdn, upn;
        float sn;
        c4_fwd_agent(D, P, S, i, X, kself, kother, kw, ko, in, p1, f1, G, pn, fn, sn, sdn, upn, rec);
        // neighbour terms of the step after, in segment order
        f3 accp = mk3(0.f, 0.f, 0.f), accf = accp;
        if (rstart_in) {
            const uint32_t s1 = rstart_in[i + 1];
            for (uint32_t s = rstart_in[i]; s < s1; ++s) {
                const float4 u = contrib_in[2 * (size_t)s], v = contrib_in[2 * (size_t)s + 1];
                accp = accp + mk3(u.x, u.y, u.z);
                accf = accf + mk3(u.w, v.x, v.y);
            }
        }
        auto own = [&](int c) { return own_in.a[c][i]; };
        f3 pnb = mk3(own(0), own(1), own(2)) + accp;
        f3 fnb = mk3(own(3), own(4), own(5)) + accf;
        const float snb = own(6);
        const f3 sdnb = mk3(own(7), own(8), own(9));
        f3 pb = mk3(0.f, 0.f, 0.f), fb = pb, sdb = pb;
        float sb = 0.f;
        f3 p1b = mk3(0.f, 0.f, 0.f), f1b = p1b;
        // ---- integrate
        const f3 Ft = truncate3(G, D.max_force);
        const f3 accv = Ft / D.mass;
        const f3 v1 = f1 * in.s + accv * D.dt;
        f3 vb = mk3(0.f, 0.f, 0.f);
        if (rec.moved) {
            const f3 cb = normalize_adj(cross3(fn, in.up), sdnb);  // sdn = normalize(fn x up), up constant
            fnb = fnb + cross_adj_a(in.up, cb);
            vb = vb + (fnb - fn * dot3(fn, fnb)) / sn;  // fn = v / |v|
            vb = vb + fn * snb;                         // sn = |v|
        } else {
            f1b = f1b + fnb;
            sdb = sdb + sdnb;
        }
        p1b = p1b + pnb;
        vb = vb + pnb * D.dt;
        float vmaxb = 0.f;
        const f3 v1b = truncate_adj(v1, in.vmax, rec.tv != 0, vb, vmaxb);
        tg[TV + g] += vmaxb;
        f1b = f1b + v1b * in.s;
        sb += dot3(f1, v1b);
        float fmb = 0.f;
        const f3 Gb = truncate_adj(G, D.max_force, rec.tf != 0, (v1b * D.dt) / D.mass, fmb);
        // ---- seek
        if (rec.on_seek) {
            const f3 goal = mk3(S.goal[g][0], S.goal[g][1], S.goal[g][2]);
            const f3 qb = normalize_adj((goal - p1) - f1, Gb * D.w_seek);
            p1b = p1b - qb;
            f1b = f1b - qb;
        }
        // ---- upstream adjoints of the neighbour terms (first pass over the lists)
        f3 nb_coh = mk3(0.f, 0.f, 0.f), nb_ali = nb_coh, ab_ss = nb_coh, ab_so = nb_coh;
        {
            f3 sc = mk3(0.f, 0.f, 0.f), sa = sc, s_ss = sc, s_so = sc;
            for (uint32_t j = 0; j < D.k; ++j) {
                const uint32_t b = kself[(size_t)j * D.N + i];
                if (b >= D.N) break;
                const uint32_t bit = 1u << j;
                if (!((rec.m_co | rec.m_al | rec.m_ss) & bit)) continue;
                const f3 pj = ldS(X, 0, b);
                if (rec.m_co & bit) sc = sc + pj;
                if (rec.m_al & bit) sa = sa + ldS(X, 3, b);
                if (rec.m_ss & bit) {
                    const f3 off = pj - p1;
                    s_ss = s_ss + off / -dot3(off, off);
                }
            }
            for (uint32_t j = 0; j < D.k; ++j) {
                const uint32_t b = kother[(size_t)j * D.N + i];
                if (b >= D.N) break;
                if (rec.m_so & (1u << j)) {
                    const f3 off = ldS(X, 0, b) - p1;
                    s_so = s_so + off / -dot3(off, off);
                }
            }
            if (rec.on_co && rec.m_co) {
                const float cnt = (float)__popc(rec.m_co);
                const f3 q = (sc / cnt) - p1;
                tg[TW_COH + g] += dot3(normalize3(q), Gb);
                const f3 qb = normalize_adj(q, Gb * D.th.w_coh[g]);
                p1b = p1b - qb;
                nb_coh = qb / cnt;
            }
            if (rec.on_al && rec.m_al) {
                const float cnt = (float)__popc(rec.m_al);
                const f3 q = (sa / cnt) - f1;
                tg[TW_ALI + g] += dot3(normalize3(q), Gb);
                const f3 qb = normalize_adj(q, Gb * D.th.w_ali[g]);
                f1b = f1b - qb;
                nb_ali = qb / cnt;
            }
            if (rec.on_ss && rec.m_ss) {
                const float cnt = (float)__popc(rec.m_ss);
                const f3 mean = s_ss / cnt;
                tg[TW_SEP + g] += dot3(normalize3(mean), Gb);
                ab_ss = normalize_adj(mean, Gb * D.th.w_sep[g]) / cnt;
            }
            if (rec.on_so && rec.m_so) {
                const float cnt = (float)__popc(rec.m_so);
                const f3 mean = s_so / cnt;
                tg[TW_SEP + g] += dot3(normalize3(mean), Gb);
                ab_so = normalize_adj(mean, Gb * D.th.w_sep[g]) / cnt;
            }
        }
        // ---- per-neighbour pass: separation, cohesion, alignment and the
        // step's penetration term, one accumulation per neighbour slot
        for (int w = 0; w < 2; ++w) {
            const uint32_t* lst = w == 0 ? kself : kother;
            const uint32_t msep = w == 0 ? rec.m_ss : rec.m_so;
            const f3 ab = w == 0 ? ab_ss : ab_so;
            const int go = w == 0 ? g : 1 - g;
            const float R = D.th.radius[g] + D.th.radius[go];
            for (uint32_t j = 0; j < D.k; ++j) {
                const uint32_t b = lst[(size_t)j * D.N + i];
                if (b >= D.N) break;
                const uint32_t bit = 1u << j;
                const f3 pj = ldS(X, 0, b);
                f3 cp = mk3(0.f, 0.f, 0.f), cf = cp;
                if (msep & bit) {
                    const f3 off = pj - p1;
                    const float d2 = dot3(off, off);
                    // term = -off / d2
                    const f3 offb = ab / -d2 + off * (2.f * dot3(off, ab) / (d2 * d2));
                    cp = cp + offb;
                    p1b = p1b - offb;
                }
                if (w == 0) {
                    if (rec.m_co & bit) cp = cp + nb_coh;
                    if (rec.m_al & bit) cf = cf + nb_ali;
                }
                const f3 dv = pj - in.p;
                const float d = len3(dv);
                const float e = R - d;
                if (e > 0.f && d > 0.f) {
                    const float eb = 2.f * D.pen_w * e * D.inv_S;
                    tg[TR + g] += eb;
                    tg[TR + go] += eb;
                    const f3 gpj = dv * (-eb / d);
                    cp = cp + gpj;
                    pb = pb - gpj;
                }
                const size_t so = slot_out[rev_code(i, (uint32_t)w, j, D.k)];
                contrib_out[2 * so] = make_float4(cp.x, cp.y, cp.z, cf.x);
                contrib_out[2 * so + 1] = make_float4(cf.y, cf.z, 0.f, 0.f);
            }
        }
        // ---- obstacles
        if (rec.o_apply) {
            const f3 c = mk3(D.ox[rec.o_idx], D.oy[rec.o_idx], D.oz[rec.o_idx]);
            const f3 vv = p1 - c;
            tg[TW_OBS + g] += dot3(perp3(vv, f1), Gb);
            const f3 Fb = Gb * D.th.w_obs[g];
            p1b = p1b + (Fb - f1 * dot3(f1, Fb));
            f1b = f1b - (Fb * dot3(vv, f1) + vv * dot3(f1, Fb));
        }
        // ---- walls
        {
            const float r = in.r;
            const float fl = 4.f * r;
            const f3 p = in.p, f = in.f, sd = in.sd;
            f3 fe[3];
            fe[0] = p + (f * in.s) * D.mt_walls;
            fe[1] = p + (sd * fl + f * fl);
            fe[2] = p + ((-sd) * fl + f * fl);
            f3 feb[3] = {mk3(0.f, 0.f, 0.f), mk3(0.f, 0.f, 0.f), mk3(0.f, 0.f, 0.f)};
            float rb = 0.f;
            pb = pb + p1b;  // p1 = p (+ n * pen when touching)
            if (rec.w_touch) {
                const uint32_t gw = S.w0 + rec.w_wall;
                const f3 n = mk3(D.wnx[gw], D.wny[gw], D.wnz[gw]);
                const f3 ls = mk3(D.wsx[gw], D.wsy[gw], D.wsz[gw]);
                const f3 le = mk3(D.wex[gw], D.wey[gw], D.wez[gw]);
                // p1 = p + n * (r - dist), f1 = normalize(perp(f, n))
                const float penb = dot3(n, p1b);
                rb += penb;
                const float distb = -penb;
                const f3 qb = normalize_adj(perp3(f, n), f1b);
                fb = fb + (qb - n * dot3(n, qb));
                const f3 ea = rec.w_event == 1 ? p + (-n) * r : fe[rec.w_feeler];
                f3 ip;
                LIRec lr;
                li_fwd(p, ea, ls, le, ip, lr);
                const f3 dv = p - ip;
                const float dl = len3(dv);
                if (dl > 0.f) {
                    const f3 u = dv * (distb / dl);
                    pb = pb + u;
                    f3 sab = mk3(0.f, 0.f, 0.f), eab = sab;
                    li_adj(p, ea, ls, le, lr, -u, sab, eab);
                    pb = pb + sab;
                    if (rec.w_event == 1) {
                        pb = pb + eab;
                        rb -= dot3(n, eab);
                    } else {
                        feb[rec.w_feeler] = feb[rec.w_feeler] + eab;
                    }
                }
            } else {
                fb = fb + f1b;
                if (rec.w_event == 2) {
                    const uint32_t gw = S.w0 + rec.w_wall;
                    const f3 n = mk3(D.wnx[gw], D.wny[gw], D.wnz[gw]);
                    const f3 ls = mk3(D.wsx[gw], D.wsy[gw], D.wsz[gw]);
                    const f3 le = mk3(D.wex[gw], D.wey[gw], D.wez[gw]);
                    const f3 ea = fe[rec.w_feeler];
                    f3 ip;
                    LIRec lr;
                    li_fwd(p, ea, ls, le, ip, lr);
                    const f3 dv = ea - ip;
                    const float dl = len3(dv);
                    tg[TW_WALL + g] += dot3(n * dl, Gb);
                    const float lb = dot3(n, Gb) * D.th.w_wall[g];
                    if (dl > 0.f) {
                        const f3 u = dv * (lb / dl);
                        feb[rec.w_feeler] = feb[rec.w_feeler] + u;
                        f3 sab = mk3(0.f, 0.f, 0.f), eab = sab;
                        li_adj(p, ea, ls, le, lr, -u, sab, eab);
                        pb = pb + sab;
                        feb[rec.w_feeler] = feb[rec.w_feeler] + eab;
                    }
                }
            }
            // feelers: fe0 = p + f*s*mt; fe1 = p + sd*fl + f*fl; fe2 = p - sd*fl + f*fl
            pb = pb + feb[0] + feb[1] + feb[2];
            fb = fb + feb[0] * (in.s * D.mt_walls) + (feb[1] + feb[2]) * fl;
            sb += D.mt_walls * dot3(f, feb[0]);
            sdb = sdb + (feb[1] - feb[2]) * fl;
            rb += 4.f * (dot3(feb[1], sd + f) + dot3(feb[2], f - sd));
            tg[TR + g] += rb;
        }
        own_out.a[0][i] = pb.x;
        own_out.a[1][i] = pb.y;
        own_out.a[2][i] = pb.z;
        own_out.a[3][i] = fb.x;
        own_out.a[4][i] = fb.y;
        own_out.a[5][i] = fb.z;
        own_out.a[6][i] = sb;
        own_out.a[7][i] = sdb.x;
        own_out.a[8][i] = sdb.y;
        own_out.a[9][i] = sdb.z;
    }
    // parameter gradient: the block's sums, added to its own row
#pragma unroll
    for (int c = 0; c < kTheta; ++c) block_add(tg[c], grad_part + (size_t)blockIdx.x * kTheta + c, lds);
}

}  // namespace gs

// ---------------------------------------------------------------------------
// Host side.

namespace gs {

// Ensemble field of each state component (gs_agents order).
static const int kStateField[kState] = {0, 1, 2, 3, 4, 5, 6, 10, 11, 12, 13, 14, 15};

// A state is one buffer of kState x N floats (component-major).
static CState cstate(gs_calib_s* c, std::vector<DevBuf<float>>& v, int idx) {
    CState s;
    for (int j = 0; j < kState; ++j) s.a[j] = v[(size_t)idx].p + (size_t)j * c->ens->N;
    return s;
}
static CAdj cadj(DevBuf<float>& buf, size_t N) {
    CAdj a;
    for (int j = 0; j < kAdj; ++j) a.a[j] = buf.p + (size_t)j * N;
    return a;
}
static FlockTerm cterm(const gs_flock_term& t) { return FlockTerm{t.min_d, t.max_d, t.cos_max, 0.f, t.mask, 1}; }

static CalibDev calib_dev(gs_calib_s* c) {
    gs_ensemble_s* e = c->ens;
    CalibDev D;
    D.pops = e->d_pops.p;
    D.scens = e->d_scens.p;
    D.block_pop = e->d_block_pop.p;
    D.th = c->theta;
    D.sep = cterm(c->cfg.separation);
    D.ali = cterm(c->cfg.alignment);
    D.coh = cterm(c->cfg.cohesion);
    D.w_seek = c->cfg.seek_weight;
    D.mask_seek = c->cfg.seek_mask;
    D.mask_sep = c->cfg.separation.mask;
    D.mt_walls = c->cfg.walls_min_time;
    D.mt_obs = c->cfg.obstacles_min_time;
    D.max_force = c->cfg.max_force;
    D.mass = c->cfg.mass;
    D.dt = c->cfg.dt;
    D.pen_w = c->cfg.penetration_weight;
    D.k = e->k;
    D.N = e->N;
    D.inv_S = 1.f / (float)std::max<size_t>(1, e->scen.size());
    D.wsx = e->walls[0].p;
    D.wsy = e->walls[1].p;
    D.wsz = e->walls[2].p;
    D.wex = e->walls[3].p;
    D.wey = e->walls[4].p;
    D.wez = e->walls[5].p;
    D.wnx = e->walls[6].p;
    D.wny = e->walls[7].p;
    D.wnz = e->walls[8].p;
    D.ox = e->obs[0].p;
    D.oy = e->obs[1].p;
    D.oz = e->obs[2].p;
    D.orad = e->obs[3].p;
    return D;
}

static void copy_state(gs_calib_s* c, std::vector<DevBuf<float>>& dst, int di, std::vector<DevBuf<float>>& src,
                       int si) {
    GS_HIP(hipMemcpyAsync(dst[(size_t)di].p, src[(size_t)si].p, (size_t)c->ens->N * kState * 4,
                          hipMemcpyDeviceToDevice, c->ens->ctx->stream));
}

// Lists of step slot j: own, other, walls, obstacles.
static uint32_t* list(gs_calib_s* c, int j, int w) { return c->lists[(size_t)j * 4 + w].p; }

static void calib_search(gs_calib_s* c, const CState& X, int slot, int seed_slot) {
    ens_search_fixed(c->ens, X.a[0], X.a[1], X.a[2], c->ids_fixed.p, c->ids_scratch.p,
                     (float* const[3]){c->pos_scratch[0].p, c->pos_scratch[1].p, c->pos_scratch[2].p},
                     seed_slot >= 0 ? list(c, seed_slot, 0) : nullptr, seed_slot >= 0 ? list(c, seed_slot, 1) : nullptr,
                     list(c, slot, 0), list(c, slot, 1), list(c, slot, 2), list(c, slot, 3));
}

void calib_init(gs_calib_s* c, gs_ensemble_s* e, const gs_calib_config& cfg) {
    c->ens = e;
    c->cfg = cfg;
    const uint32_t T = cfg.steps, K = cfg.checkpoint_every;
    c->nseg = (T + K - 1) / K;
    const size_t N = e->N ? e->N : 1;
    hipStream_t st = e->ctx->stream;
    auto alloc_states = [&](std::vector<DevBuf<float>>& v, size_t count) {
        v.resize(count);
        for (auto& b : v) b.reserve(N * kState);
    };
    alloc_states(c->x0, 2);  // initial state and the forward's final state
    alloc_states(c->ckpt, std::max<uint32_t>(1, c->nseg));
    alloc_states(c->seg, (size_t)std::max<uint32_t>(K, 1) + 1);
    // The forward keeps every step's KNN lists (integers, held fixed under
    // differentiation) when they fit in a modest share of HBM, so the
    // backward recomputes only the state steps; otherwise it recomputes the
    // lists of each segment too.
    const size_t list_bytes = N * 4 * (2 * (size_t)e->k + e->kw_max + e->ko_max);
    c->record_lists = (size_t)T * list_bytes <= (size_t)48 << 30;
    c->lists.resize((size_t)(c->record_lists ? std::max<uint32_t>(T, 2) : std::max<uint32_t>(K, 2)) * 4);
    for (size_t j = 0; j < c->lists.size(); ++j) {
        const int w = (int)(j % 4);
        c->lists[j].reserve(N * (w < 2 ? e->k : (w == 2 ? e->kw_max : e->ko_max)));
    }
    const size_t entries = N * 2 * (size_t)e->k;  // own-group and other-group list slots
    const size_t scan_words = N + 1 + (N + 2047) / 2048;  // launch_scan_counts' start + tile sums
    for (int b = 0; b < 2; ++b) {
        c->adj_own[b].reserve(N * kAdj);
        c->rstart[b].reserve(scan_words);
        c->contrib[b].reserve(2 * entries);
    }
    c->rev.reserve(entries);
    c->rev_agent.reserve(entries);
    c->slot_of.reserve(entries);
    c->rcnt.reserve(N);
    launch_fill_u32(c->rcnt.p, (uint32_t)N, 0u, st);
    c->loss_part.reserve(std::max<size_t>(N / kBlock, 1));
    c->grad_part.reserve(std::max<size_t>(N / kBlock, 1) * kTheta);
    c->ids_fixed.reserve(N);
    c->ids_scratch.reserve(N);
    for (int a = 0; a < 3; ++a) c->pos_scratch[a].reserve(N);
    c->loss.reserve(2);
    c->grad.reserve(kTheta);
    const size_t bytes = N * 4;
    for (int j = 0; j < kState; ++j)
        GS_HIP(hipMemcpyAsync(c->x0[0].p + (size_t)j * N, e->F(kStateField[j]), bytes, hipMemcpyDeviceToDevice, st));
    GS_HIP(hipMemcpyAsync(c->ids_fixed.p, e->ids(), bytes, hipMemcpyDeviceToDevice, st));
    for (auto& ev : c->ev) GS_HIP(hipEventCreateWithFlags(&ev, hipEventDisableSystemFence));
    // The fixed order is each population's agents sorted once by their
    // initial cell, so the neighbour gathers and adjoint accumulations of the
    // whole run touch nearby memory (agents stay near their start).
    std::vector<std::vector<float>> h(kState, std::vector<float>(N));
    std::vector<uint32_t> ids(N);
    for (int j = 0; j < kState; ++j)
        GS_HIP(hipMemcpyAsync(h[j].data(), c->x0[0].p + (size_t)j * N, bytes, hipMemcpyDeviceToHost, st));
    GS_HIP(hipMemcpyAsync(ids.data(), c->ids_fixed.p, bytes, hipMemcpyDeviceToHost, st));
    GS_HIP(hipStreamSynchronize(st));
    auto axis = [](float p, float o, float s, int d) {
        if (!(p >= o)) return 0;
        if (p >= o + s) return d - 1;
        const int q = (int)(((p - o) / s) * (float)d);
        return q < d ? q : d - 1;
    };
    std::vector<uint32_t> order, key;
    std::vector<float> tmp;
    std::vector<uint32_t> tmpi;
    for (const EnsPop& P : e->pops) {
        const GridDev& gd = P.g;
        order.resize(P.n);
        key.resize(P.n);
        for (uint32_t q = 0; q < P.n; ++q) {
            const uint32_t s = P.a0 + q;
            order[q] = q;
            key[q] = (uint32_t)(axis(h[0][s], gd.ox, gd.sx, gd.dx) * gd.dy * gd.dz +
                                axis(h[1][s], gd.oy, gd.sy, gd.dy) * gd.dz + axis(h[2][s], gd.oz, gd.sz, gd.dz));
        }
        std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) { return key[a] < key[b]; });
        for (int j = 0; j < kState; ++j) {
            tmp.assign(h[j].begin() + P.a0, h[j].begin() + P.a0 + P.n);
            for (uint32_t q = 0; q < P.n; ++q) h[j][P.a0 + q] = tmp[order[q]];
        }
        tmpi.assign(ids.begin() + P.a0, ids.begin() + P.a0 + P.n);
        for (uint32_t q = 0; q < P.n; ++q) ids[P.a0 + q] = tmpi[order[q]];
    }
    for (int j = 0; j < kState; ++j)
        GS_HIP(hipMemcpyAsync(c->x0[0].p + (size_t)j * N, h[j].data(), bytes, hipMemcpyHostToDevice, st));
    GS_HIP(hipMemcpyAsync(c->ids_fixed.p, ids.data(), bytes, hipMemcpyHostToDevice, st));
    GS_HIP(hipStreamSynchronize(st));
}

void calib_release(gs_calib_s* c) {
    for (auto& ev : c->ev)
        if (ev) (void)hipEventDestroy(ev);
    for (auto* v : {&c->x0, &c->ckpt, &c->seg})
        for (auto& b : *v) b.release();
    for (auto& b : c->lists) b.release();
    for (int b = 0; b < 2; ++b) {
        c->adj_own[b].release();
        c->rstart[b].release();
        c->contrib[b].release();
    }
    c->rev.release();
    c->rev_agent.release();
    c->slot_of.release();
    c->rcnt.release();
    c->loss_part.release();
    c->grad_part.release();
    c->ids_fixed.release();
    c->ids_scratch.release();
    for (int a = 0; a < 3; ++a) c->pos_scratch[a].release();
    c->loss.release();
    c->grad.release();
}

static void zero_words(void* p, size_t words, hipStream_t st) { launch_fill_u32((uint32_t*)p, (uint32_t)words, 0u, st); }

double calib_forward(gs_calib_s* c, const gs_calib_params& th) {
    gs_ensemble_s* e = c->ens;
    hipStream_t st = e->ctx->stream;
    c->theta = th;
    const CalibDev D = calib_dev(c);
    const uint32_t T = c->cfg.steps, K = c->cfg.checkpoint_every, nb = e->N / kBlock;
    zero_words(c->loss_part.p, 2 * std::max<uint32_t>(nb, 1), st);
    GS_HIP(hipEventRecord(c->ev[0], st));
    copy_state(c, c->seg, 0, c->x0, 0);
    int cur = 0;
    for (uint32_t t = 0; t < T; ++t) {
        if (t % K == 0) copy_state(c, c->ckpt, (int)(t / K), c->seg, cur);
        const int slot = c->record_lists ? (int)t : (int)(t & 1u);
        const int prev = c->record_lists ? (int)t - 1 : (slot ^ 1);
        calib_search(c, cstate(c, c->seg, cur), slot, t > 0 ? prev : -1);
        if (nb)
            c4_forward<<<nb, kBlock, 0, st>>>(D, cstate(c, c->seg, cur), cstate(c, c->seg, cur ^ 1), list(c, slot, 0),
                                              list(c, slot, 1), list(c, slot, 2), list(c, slot, 3),
                                              c->loss_part.p);
        GS_LAUNCHED("c4_forward");
        cur ^= 1;
    }
    // Terminal loss (its adjoint seed is rewritten by each backward).
    if (nb) c4_terminal<<<nb, kBlock, 0, st>>>(D, cstate(c, c->seg, cur), cadj(c->adj_own[0], e->N), c->loss_part.p);
    GS_LAUNCHED("c4_terminal");
    c4_reduce_rows<<<1, kBlock, 0, st>>>(c->loss_part.p, nb, 1, c->loss.p);
    GS_LAUNCHED("c4_reduce_rows");
    copy_state(c, c->x0, 1, c->seg, cur);
    GS_HIP(hipEventRecord(c->ev[1], st));
    double loss = 0.0;
    GS_HIP(hipMemcpyAsync(&loss, c->loss.p, sizeof(double), hipMemcpyDeviceToHost, st));
    GS_HIP(hipStreamSynchronize(st));
    GS_HIP(hipEventElapsedTime(&c->fwd_ms, c->ev[0], c->ev[1]));
    c->forward_done = true;
    return loss;
}

void calib_backward(gs_calib_s* c, gs_calib_params& grad) {
    gs_ensemble_s* e = c->ens;
    hipStream_t st = e->ctx->stream;
    const CalibDev D = calib_dev(c);
    const uint32_t T = c->cfg.steps, K = c->cfg.checkpoint_every, nb = e->N / kBlock;
    const size_t N = e->N;
    GS_HIP(hipEventRecord(c->ev[2], st));
    zero_words(c->grad_part.p, 2 * (size_t)std::max<uint32_t>(nb, 1) * kTheta, st);
    // Seed the adjoint of the final state (kept by the forward in x0[1]) here
    // rather than relying on the forward's seed: the sweep below reuses both
    // adjoint buffer pairs, so a second backward after one forward must start
    // again from the terminal term.  Its loss is not summed again.
    if (nb) c4_terminal<<<nb, kBlock, 0, st>>>(D, cstate(c, c->x0, 1), cadj(c->adj_own[0], N), nullptr);
    GS_LAUNCHED("c4_terminal");
    int a = 0;  // adj_own[a] (+ the reverse lists rstart[a] / contrib[a]) hold the adjoint of the state after the step
    bool have_in = false;  // the final state has no neighbour terms
    for (int sgi = (int)c->nseg - 1; sgi >= 0; --sgi) {
        const uint32_t t0 = (uint32_t)sgi * K;
        const uint32_t L = std::min(K, T - t0);
        copy_state(c, c->seg, 0, c->ckpt, sgi);
        // lists of step t0 + j: kept by the forward, or recomputed here
        auto slot_of = [&](int j) { return c->record_lists ? (int)t0 + j : j; };
        for (uint32_t j = 0; j < L; ++j) {
            if (!c->record_lists) calib_search(c, cstate(c, c->seg, (int)j), (int)j, j > 0 ? (int)j - 1 : -1);
            const int sl = slot_of((int)j);
            if (nb)
                c4_forward<<<nb, kBlock, 0, st>>>(D, cstate(c, c->seg, (int)j), cstate(c, c->seg, (int)j + 1),
                                                  list(c, sl, 0), list(c, sl, 1), list(c, sl, 2), list(c, sl, 3),
                                                  nullptr);
            GS_LAUNCHED("c4_forward");
        }
        for (int j = (int)L - 1; j >= 0; --j) {
            const int sl = slot_of(j);
            if (nb) {
                // reverse lists of step j into parity a ^ 1
                // (slot_of holds the arrival indices until the rank overwrites it)
                c4_rev_count<<<nb, kBlock, 0, st>>>(D, list(c, sl, 0), list(c, sl, 1), c->rcnt.p, c->slot_of.p);
                GS_LAUNCHED("c4_rev_count");
                launch_scan_counts(c->rcnt.p, c->rstart[a ^ 1].p, (uint32_t)N, st);  // leaves rcnt zero
                c4_rev_fill<<<nb, kBlock, 0, st>>>(D, list(c, sl, 0), list(c, sl, 1), c->rstart[a ^ 1].p,
                                                   c->slot_of.p, c->rev.p, c->rev_agent.p);
                GS_LAUNCHED("c4_rev_fill");
                c4_rev_rank<<<blocks_for(N * 2 * e->k), kBlock, 0, st>>>((uint32_t)N, c->rstart[a ^ 1].p, c->rev.p,
                                                                         c->rev_agent.p, c->slot_of.p);
                GS_LAUNCHED("c4_rev_rank");
                c4_backward<<<nb, kBlock, 0, st>>>(D, cstate(c, c->seg, j), list(c, sl, 0), list(c, sl, 1),
                                                   list(c, sl, 2), list(c, sl, 3), cadj(c->adj_own[a], N),
                                                   have_in ? c->rstart[a].p : nullptr, c->contrib[a].p,
                                                   cadj(c->adj_own[a ^ 1], N), c->slot_of.p, c->contrib[a ^ 1].p,
                                                   c->grad_part.p);
                GS_LAUNCHED("c4_backward");
            }
            have_in = true;
            a ^= 1;
        }
    }
    c4_reduce_rows<<<kTheta, kBlock, 0, st>>>(c->grad_part.p, nb, kTheta, c->grad.p);
    GS_LAUNCHED("c4_reduce_rows");
    GS_HIP(hipEventRecord(c->ev[3], st));
    double g[kTheta];
    GS_HIP(hipMemcpyAsync(g, c->grad.p, sizeof g, hipMemcpyDeviceToHost, st));
    GS_HIP(hipStreamSynchronize(st));
    GS_HIP(hipEventElapsedTime(&c->bwd_ms, c->ev[2], c->ev[3]));
    float* out = (float*)&grad;
    for (int q = 0; q < kTheta; ++q) out[q] = (float)g[q];
}

void calib_download(gs_calib_s* c, const EnsPop& P, gs_agents* host) {
    hipStream_t st = c->ens->ctx->stream;
    if (!P.n) return;
    std::vector<uint32_t> ids(P.n);
    GS_HIP(hipMemcpyAsync(ids.data(), c->ids_fixed.p + P.a0, (size_t)P.n * 4, hipMemcpyDeviceToHost, st));
    std::vector<float> v(P.n);
    float* const* hf = (float* const*)host;
    GS_HIP(hipStreamSynchronize(st));
    for (int j = 0; j < kState; ++j) {
        float* dst = hf[kStateField[j]];
        if (!dst) continue;
        GS_HIP(hipMemcpyAsync(v.data(), c->x0[1].p + (size_t)j * c->ens->N + P.a0, (size_t)P.n * 4,
                              hipMemcpyDeviceToHost, st));
        GS_HIP(hipStreamSynchronize(st));
        for (uint32_t s = 0; s < P.n; ++s) dst[ids[s]] = v[s];
    }
}

}  // namespace gs
