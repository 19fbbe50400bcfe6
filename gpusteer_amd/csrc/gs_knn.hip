// Exact k-nearest-neighbour search, binned and brute force.
//
// Semantics (both kernels): for each query a of group A, the k points b of
// group B with the smallest key (d2, b), d2 = ((ax-bx)^2 + (ay-by)^2) + (az-bz)^2
// in fp32, excluding a itself when A and B are the same group.  This is the
// order KNNBruteForceCUDAKernelV2 produces (strict '<' over an ascending scan:
// the lower index wins ties, include/OpenSteer/CUDA/KNNBruteForceCudaKernel.cu:245-270).
// Output is slot-major [k][nA] indexed by the query's own index
// (KNNBinningV2Kernel.cu:520-526), padded with UINT32_MAX / FLT_MAX (:451-457),
// distance = sqrtf(d2).
//
// Top-k selection: a register-resident sorted list of KMAX (d2, index) pairs
// with a branch-free compare-exchange ripple.  For k < KMAX the first KMAX-k
// entries hold a sentinel below every real key, so the list's last entry is
// always the k-th best and no register is ever indexed dynamically.
//
// Binned search replaces KNNBinningV2Kernel (KNNBinningV2Kernel.cu:417-528),
// which scans a fixed cell radius and is therefore approximate.  Here each
// query (one lane, queries in sorted/cell order so a wavefront's lanes share
// cells) scans shells of cells of growing Chebyshev radius, each shell as
// contiguous z-runs of the sorted arrays, skips runs whose box is farther
// than the current bound, and stops once the k-th best d2 is below the
// squared distance to every unscanned cell.  The bound starts at +inf or, when
// the container holds the exact result of the previous call for the same
// groups, at the largest current d2 of those k previous neighbours (a valid
// upper bound on the k-th distance: they are k distinct candidates).  Either
// way the result is the exact k nearest.
#include "gs_knn_core.hpp"

namespace gs {

template <int KMAX>
__global__ __launch_bounds__(kBlock) void k_knn_binned(
    uint32_t nA, const float* __restrict__ aspx, const float* __restrict__ aspy, const float* __restrict__ aspz,
    const uint32_t* __restrict__ aperm, uint32_t nB, const float* __restrict__ bspx, const float* __restrict__ bspy,
    const float* __restrict__ bspz, const uint32_t* __restrict__ bperm, const uint32_t* __restrict__ bstart,
    const uint32_t* __restrict__ binv, GridDev g, int self,
    uint32_t k, int R0, const uint32_t* __restrict__ seed, uint32_t* __restrict__ out_idx,
    float* __restrict__ out_dist) {
    const uint32_t s = blockIdx.x * blockDim.x + threadIdx.x;
    if (s >= nA) return;
    const float qx = aspx[s], qy = aspy[s], qz = aspz[s];
    const uint32_t q = aperm[s];
    TopK<KMAX> h;
    h.init(k);
    const float thr = seed ? seed_bound(seed, k, nA, q, nB, qx, qy, qz, binv, bspx, bspy, bspz) : INFINITY;
    binned_search<KMAX>(h, thr, qx, qy, qz, self ? s : UINT32_MAX, g, R0, bspx, bspy, bspz, bperm, bstart);
    h.store(k, nA, q, out_idx, out_dist);
}

// Brute force: one lane per query in index order, B staged through LDS in
// tiles of one block.
template <int KMAX>
__global__ __launch_bounds__(kBlock) void k_knn_brute(uint32_t nA, const float* __restrict__ ax,
                                                      const float* __restrict__ ay, const float* __restrict__ az,
                                                      uint32_t nB, const float* __restrict__ bx,
                                                      const float* __restrict__ by, const float* __restrict__ bz,
                                                      int self, uint32_t k, uint32_t* __restrict__ out_idx,
                                                      float* __restrict__ out_dist) {
    __shared__ float tx[kBlock], ty[kBlock], tz[kBlock];
    const uint32_t q = blockIdx.x * blockDim.x + threadIdx.x;
    const bool active = q < nA;
    const float qx = active ? ax[q] : 0.f, qy = active ? ay[q] : 0.f, qz = active ? az[q] : 0.f;
    TopK<KMAX> h;
    h.init(k);
    for (uint32_t base = 0; base < nB; base += kBlock) {
        const uint32_t t = base + threadIdx.x;
        if (t < nB) {
            tx[threadIdx.x] = bx[t];
            ty[threadIdx.x] = by[t];
            tz[threadIdx.x] = bz[t];
        }
        __syncthreads();
        const uint32_t m = min((uint32_t)kBlock, nB - base);
        if (active) {
            for (uint32_t j = 0; j < m; ++j) {
                const uint32_t b = base + j;
                if (self && b == q) continue;
                const float d2 = sqdist(qx, qy, qz, tx[j], ty[j], tz[j]);
                if (h.accepts(d2, b)) h.push(d2, b);
            }
        }
        __syncthreads();
    }
    if (active) h.store(k, nA, q, out_idx, out_dist);
}

template <int KMAX>
static void run_binned(gs_group_s* A, gs_knn_s* out, gs_group_s* B, uint32_t R0, const uint32_t* seed) {
    const uint32_t nA = A->n;
    k_knn_binned<KMAX><<<blocks_for(nA), kBlock, 0, A->ctx->stream>>>(
        nA, A->spx.p, A->spy.p, A->spz.p, A->perm.p, B->n, B->spx.p, B->spy.p, B->spz.p, B->perm.p, B->start.p,
        B->inv.p, B->bin_grid, A == B ? 1 : 0, out->k, (int)R0, seed, out->idx.p,
        out->dist.p);
    GS_LAUNCHED("k_knn_binned");
}

void launch_knn_binned(gs_group_s* A, gs_knn_s* out, gs_group_s* B, uint32_t search_radius, bool use_seed) {
    if (A->n == 0) return;
    const uint32_t* seed = use_seed ? out->idx.p : nullptr;
    // The seed is read and the result written in the same buffer: every lane
    // reads only its own column (query q) before writing it.
    if (out->k <= 8) run_binned<8>(A, out, B, search_radius, seed);
    else if (out->k <= 16) run_binned<16>(A, out, B, search_radius, seed);
    else run_binned<32>(A, out, B, search_radius, seed);
}

void launch_knn_brute(gs_group_s* A, gs_knn_s* out, gs_group_s* B) {
    const uint32_t nA = A->n;
    if (nA == 0) return;
    hipStream_t st = A->ctx->stream;
    const int self = A == B ? 1 : 0;
#define GS_BRUTE(KM)                                                                                              \
    k_knn_brute<KM><<<blocks_for(nA), kBlock, 0, st>>>(nA, A->f[0].p, A->f[1].p, A->f[2].p, B->n, B->f[0].p,     \
                                                       B->f[1].p, B->f[2].p, self, out->k, out->idx.p, out->dist.p)
    if (out->k <= 8) GS_BRUTE(8);
    else if (out->k <= 16) GS_BRUTE(16);
    else GS_BRUTE(32);
#undef GS_BRUTE
    GS_LAUNCHED("k_knn_brute");
}

}  // namespace gs
