// Batched-ensemble state shared by gs_ensemble.hip (kernels, step) and
// gs_api.hip (entry points).
#pragma once
#include <vector>

#include "gs_ops.hpp"

namespace gs {

struct EnsPop {
    uint32_t a0, n, c0, cells;  // slots [a0, a0 + n) are agents; cells c0 .. c0 + cells - 1
    GridDev g;
    int kind, group, other;     // other: the crowd's other population (-1 for Boids)
    uint32_t scen;
};

struct EnsScen {
    int kind, mode;
    float dt;
    int wrap_shape;
    float wc[3], ws[3];
    FlockTerm sep, ali, coh;    // Boids
    FlockTerm csep;             // crowd separation
    float goal[2][3];
    float w_seek;
    uint32_t mask_seek;
    float mt_walls, w_walls, mt_obs, w_obs;
    uint32_t k_walls, k_obs;
    float other_cap2;           // inclusive d2 bound of the other-group KNN
    uint32_t w0, nw, o0, no;    // ranges in the wall / obstacle arrays
};

}  // namespace gs

struct gs_ensemble_s {
    gs_context ctx = nullptr;
    uint32_t k = 0;
    uint32_t N = 0, C = 0;  // padded slots, cells (with one extra per population)
    std::vector<gs_scenario> scen;
    std::vector<gs::EnsPop> pops;
    std::vector<gs::EnsScen> scens_host;
    std::vector<uint32_t> block_pop_h, blocks_boids, blocks_g0, blocks_g1, blocks_choke;
    gs::DevBuf<gs::EnsPop> d_pops;
    gs::DevBuf<gs::EnsScen> d_scens;
    gs::DevBuf<uint32_t> d_block_pop, d_blocks_boids, d_blocks_g0, d_blocks_g1, d_blocks_choke;

    gs::DevBuf<float> fb[kFields][2];
    uint8_t fcur[kFields] = {};
    gs::DevBuf<uint32_t> bitsb[2], idb[2];
    uint8_t bcur = 0, icur = 0;
    gs::DevBuf<uint32_t> knn_self[2], knn_other[2];
    uint8_t kcur = 0;
    bool seeded = false;
    // Fields a step's counting sort need not move: the per-agent constants
    // (and the smoothed acceleration when no scenario integrates in the
    // OpenSteer mode) when every population holds one value for them, since
    // the sort keeps each population in its slot range; and, from the second
    // step after an upload on, the steering vector and the bits, which every
    // step leaves zero.  pop_uniform: per population, the fields uniform at
    // its last upload.
    std::vector<uint32_t> pop_uniform;
    uint32_t frozen = 0;   // frozen constants (fields 7..9, 19..22)
    bool steer_zero = false;
    uint32_t kw_max = 1, ko_max = 1;
    gs::DevBuf<uint32_t> knn_walls, knn_obs;
    // Bounding boxes of the crowd populations' snapshot positions, 6 words
    // per population (as a group's bbox pair: the key pass accumulates into
    // box[box_cur] and clears the other buffer): the other-group search skips
    // queries its cap cannot reach the other group's box from.
    gs::DevBuf<uint32_t> box[2];
    uint8_t box_cur = 0;
    gs::DevBuf<uint32_t> key, rank, id_tmp, new_to_old, old_to_new, count, start;
    gs::DevBuf<float4> spos;

    // Walls: start, end, normal, midpoint (12 arrays); obstacles: centre, radius.
    gs::DevBuf<float> walls[12], obs[4];
    std::vector<std::vector<float>> walls_h, obs_h;  // per scenario, 12 / 4 floats per entry
    bool walls_mid_ready = false;

    // Captured pairs of steps, one per buffer parity.
    hipGraph_t graph[2] = {nullptr, nullptr};
    hipGraphExec_t exec[2] = {nullptr, nullptr};
    uint32_t magic = 0x67736573u;

    float* F(int f) const { return fb[f][fcur[f]].p; }
    float* Falt(int f) const { return fb[f][fcur[f] ^ 1].p; }
    uint32_t* bits() const { return bitsb[bcur].p; }
    uint32_t* ids() const { return idb[icur].p; }
};

struct gs_calib_s {
    gs_ensemble_s* ens = nullptr;
    gs_calib_config cfg{};
    gs_calib_params theta{};
    bool forward_done = false;
    uint32_t nseg = 0;
    // fixed-order state: initial, checkpoints, the recomputed segment
    std::vector<gs::DevBuf<float>> x0, ckpt, seg;  // one buffer of 13 x N floats per state
    bool record_lists = false;  // the forward keeps every step's KNN lists
    gs::DevBuf<uint32_t> ids_fixed, ids_scratch;
    gs::DevBuf<float> pos_scratch[3];
    // KNN lists of the segment's steps: per step own k, other k, walls kw, obstacles ko
    std::vector<gs::DevBuf<uint32_t>> lists;
    // adjoint of a state: each agent's own terms (10 components x N),
    // double-buffered, plus the neighbour terms of the step's reverse lists
    // (rstart, one 8-float slot per list entry in contrib), double-buffered;
    // rev / rev_agent / slot_of / rcnt are the scratch of a reverse-list build
    gs::DevBuf<float> adj_own[2];
    gs::DevBuf<uint32_t> rstart[2], rev, rev_agent, slot_of, rcnt;
    gs::DevBuf<float4> contrib[2];
    // loss and gradient, and their per-block sums (added in block order)
    gs::DevBuf<double> loss, grad, loss_part, grad_part;
    int final_buf = 0;  // which of seg[0] / seg[1] holds the forward's final state
    hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
    float fwd_ms = 0.f, bwd_ms = 0.f;
    uint32_t magic = 0x67736361u;
};

namespace gs {
void calib_init(gs_calib_s* c, gs_ensemble_s* e, const gs_calib_config& cfg);
void calib_release(gs_calib_s* c);
double calib_forward(gs_calib_s* c, const gs_calib_params& th);
void calib_backward(gs_calib_s* c, gs_calib_params& grad);
void calib_download(gs_calib_s* c, const EnsPop& P, gs_agents* host);
void ens_init(gs_ensemble_s* e, uint32_t ns, const gs_scenario* sc, uint32_t k,
              GridDev (*make_grid)(const float*, const float*, const uint32_t*));
void ens_alloc(gs_ensemble_s* e);
// Frees every device buffer of the ensemble (its graphs are destroyed by the
// caller).
void ens_release(gs_ensemble_s* e);
void ens_upload_geometry(gs_ensemble_s* e);
void ens_step(gs_ensemble_s* e);
void ens_search_fixed(gs_ensemble_s* e, float* px, float* py, float* pz, const uint32_t* ids_fixed,
                      uint32_t* ids_scratch, float* const scratch[3], const uint32_t* seed_self,
                      const uint32_t* seed_other, uint32_t* out_self, uint32_t* out_other, uint32_t* out_walls,
                      uint32_t* out_obs);
}  // namespace gs
