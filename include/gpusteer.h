/*
 * gpusteer.h — C-ABI of the MI355X (gfx950) implementation of GPUSteer's
 * per-step simulation path: spatial-bin construction, exact binned and
 * brute-force k-nearest-neighbour search, the SteerLibrary forces and the
 * SimpleVehicle integrate step.
 *
 * Plain C types only.  Every entry point returns an int status (GS_OK == 0),
 * never aborts, and queues its device work on the stream of the context the
 * handle belongs to.  There is no global state: handles of different contexts
 * may be used from different threads.  A single handle must not be used by
 * two threads at once.
 *
 * Reference interface each entry point replaces (paths relative to the
 * reference repository; "GSL" = include/OpenSteer/CUDA/GroupSteerLibrary.cuh):
 *   gs_update_knn_database   GSL:129  updateKNNDatabase(BaseGroup*, KNNBinData*)
 *   gs_find_knn              GSL:138  findKNearestNeighbors(A, KNNData*, KNNBinData*, B, searchRadius)
 *   gs_find_knn_brute        KNNBruteForceCUDAV2::run (CUDA/KNNBruteForceCUDA.cu:163-187)
 *   gs_steer_separation      GSL:161  steerForSeparation(A, KNNData*, B, minD, maxD, cosA, w, mask)
 *   gs_steer_alignment       GSL:170  steerForAlignment(...)
 *   gs_steer_cohesion        GSL:179  steerForCohesion(...)
 *   gs_steer_avoid_obstacles GSL:102  steerToAvoidObstacles(A, ObstacleGroup*, KNNData*, minTime, w, mask)
 *   gs_steer_avoid_walls     GSL:120  steerToAvoidWalls(A, KNNData*, WallGroup*, minTime, w, mask)
 *   gs_steer_seek            GSL:52   steerForSeek(A, float3 target, w, mask)
 *   gs_update_group          GSL:189  updateGroup(A, elapsedTime)
 *   gs_wrap_world            GSL:43   wrapWorld(A, float3 worldSize)
 *   gs_clear_bits            plugins/ChokePointNew.cpp:595 (cudaMemset of pdAppliedKernels)
 *   gs_update_group_anti_penetration
 *                            GSL:198  updateGroupWithAntiPenetration(A, KNNData*, WallGroup*, elapsedTime)
 *   gs_anti_penetration_agents
 *                            GSL:216  antiPenetrationAgents(A, KNNData*, B, mask)
 *   gs_steer_follow_path     GSL:88   steerToFollowPath(A, PolylinePathwayCUDA*, predictionTime, w, mask)
 *   gs_steer_avoid_neighbors GSL:111  steerToAvoidNeighbors(A, KNNData*, B, minTimeToCollision,
 *                                     minSeparationDistance, bAvoidCloseNeighbors, w, mask)
 *   gs_path_*                include/OpenSteer/CUDA/PolylinePathwayCUDA.cuh:11-68
 *   gs_group_cell_density    ComputeCellDensity(AgentGroup*, uint*) (CUDA/ProcessData.cu:52-80)
 *   gs_group_cell_velocity   ComputeAvgCellVelocity(AgentGroup*, float3*) (CUDA/ProcessData.cu:83-129)
 *   gs_write_cell_density    WriteCellDensity(prefix, AgentGroup*, selected) (src/ExtractData.cpp:19-88)
 *   gs_write_cell_velocity   WriteAvgCellVelocity(prefix, AgentGroup*, selected) (src/ExtractData.cpp:90-140)
 *   gs_group_*               include/OpenSteer/AgentGroup.h:32-54, ObstacleGroup.h:23-37,
 *                            WallGroup.h:18-33 (AddAgent/SyncDevice/GetDataForAgent/pd*)
 *   gs_knn_*                 include/OpenSteer/CUDA/KNNData.cuh:31-54
 *   gs_bins_*                include/OpenSteer/CUDA/KNNBinDataV2.cuh:43-61
 *   gs_step_boids            the per-frame sequence of plugins/BoidsGroup.cpp:606-637
 *                            (wrap, DB, KNN, separation/alignment/cohesion, update)
 *                            fused into one call; no single reference counterpart.
 */
#ifndef GPUSTEER_H
#define GPUSTEER_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GS_VERSION 100

/* ---- status codes ------------------------------------------------------- */
enum {
    GS_OK = 0,
    GS_ERR_INVALID_ARGUMENT = 1, /* bad value (size 0 where not allowed, NULL pointer, bad enum) */
    GS_ERR_INVALID_HANDLE = 2,   /* NULL / destroyed / wrong-kind handle */
    GS_ERR_SIZE_MISMATCH = 3,    /* e.g. KNN container sized for another group */
    GS_ERR_UNSUPPORTED = 4,      /* e.g. k > GS_MAX_K, too many cells */
    GS_ERR_NOT_BINNED = 5,       /* KNN on a group without a current bin database */
    GS_ERR_OUT_OF_MEMORY = 6,    /* device allocation failed */
    GS_ERR_DEVICE = 7,           /* HIP runtime error (message in gs_last_error) */
    GS_ERR_NONFINITE = 8,        /* gs_group_check_finite found NaN/Inf */
    GS_ERR_FRAME = 9             /* call not recordable into a frame, or a frame replayed on handles in another state */
};

#define GS_MAX_K 32u
/* Largest bin lattice (cells), and the largest lattice gs_wall_set_split
 * splits over (it lists every cell on the host, 40 bytes each). */
#define GS_MAX_CELLS (1u << 28)
#define GS_MAX_SPLIT_CELLS (1u << 24)

/* Force / integrate semantics.
 * GS_MODE_OPENSTEER: the CPU OpenSteer formulas (src/SteerLibrary.cpp,
 *   src/SimpleVehicle.cpp): squared-distance neighbourhood tests with a
 *   minimum distance of 3 x radius when min_d < 0, unnormalized seek,
 *   applySteeringForce with steering adjustment and acceleration smoothing.
 * GS_MODE_GPUREF: the reference CUDA kernels (FlockingCommon.cuh,
 *   SteerForSeekCUDAKernel.cu, UpdateCUDAKernel.cu). */
enum { GS_MODE_OPENSTEER = 0, GS_MODE_GPUREF = 1 };

enum { GS_KIND_AGENTS = 0, GS_KIND_OBSTACLES = 1, GS_KIND_WALLS = 2 };

enum { GS_WRAP_BOX = 0, GS_WRAP_SPHERE = 1 };

/* gs_find_knn flags */
enum {
    GS_KNN_DEFAULT = 0,
    /* Do not use the container's previous (exact) result as a search bound. */
    GS_KNN_NO_SEED = 1u << 0,
    /* Search through the bins even where an exhaustive scan of B is chosen
     * (small query groups, small or sparse B); the result is the same. */
    GS_KNN_BINNED = 1u << 1
};

/* Behaviour bits, include/OpenSteer/CUDA/CUDAKernelGlobals.cuh:12-31. */
#define GS_BIT_PURSUE (1u << 0)
#define GS_BIT_EVADE (1u << 1)
#define GS_BIT_SEEK (1u << 2)
#define GS_BIT_FLEE (1u << 3)
#define GS_BIT_SEPARATION (1u << 4)
#define GS_BIT_ALIGNMENT (1u << 5)
#define GS_BIT_COHESION (1u << 6)
#define GS_BIT_AVOID_OBSTACLES (1u << 7)
#define GS_BIT_AVOID_WALLS (1u << 8)
#define GS_BIT_AVOID_NEIGHBORS (1u << 9)
#define GS_BIT_AVOID_CLOSE_NEIGHBORS (1u << 10)
#define GS_BIT_FOLLOW_PATH (1u << 11)
#define GS_BIT_ANTI_PENETRATION_WALL (1u << 12)
#define GS_BIT_ANTI_PENETRATION_AGENT (1u << 13)

/* ---- handles -------------------------------------------------------------- */
typedef struct gs_context_s* gs_context;
typedef struct gs_bins_s* gs_bins;
typedef struct gs_group_s* gs_group;
typedef struct gs_knn_s* gs_knn;

/* Agent state, struct of arrays (one float / uint32 per agent in each array).
 * Used for host upload/download and to expose the device arrays.  Field
 * meaning: include/OpenSteer/AgentData.h:13-42 plus SimpleVehicle's smoothed
 * acceleration (include/OpenSteer/SimpleVehicle.h:335-345). */
typedef struct {
    float *px, *py, *pz;    /* position */
    float *fx, *fy, *fz;    /* unit forward (direction) */
    float *speed;
    float *ax, *ay, *az;    /* smoothed acceleration */
    float *sx, *sy, *sz;    /* side */
    float *ux, *uy, *uz;    /* up */
    float *gx, *gy, *gz;    /* steering accumulator */
    float *radius, *max_speed, *max_force, *mass;
    uint32_t *bits;         /* applied-behaviour bits */
} gs_agents;

/* Wall segments (include/OpenSteer/CUDA/WallGroupData.cuh), SoA. */
typedef struct {
    float *sx, *sy, *sz; /* start */
    float *ex, *ey, *ez; /* end */
    float *nx, *ny, *nz; /* normal */
} gs_walls;

/* One flocking term, the arguments of steerFor{Separation,Alignment,Cohesion}. */
typedef struct {
    float min_d, max_d, cos_max, weight;
    uint32_t mask;
} gs_flock_term;

/* ---- errors ---------------------------------------------------------------- */
const char* gs_status_string(int status);
/* Detail of the last failure on the calling thread ("" if none). */
const char* gs_last_error(void);
int gs_version(void);

/* ---- context ---------------------------------------------------------------- */
/* hip_stream: a hipStream_t to queue work on, or NULL for a context-owned
 * non-blocking stream at default priority. */
int gs_context_create(int device, void* hip_stream, gs_context* out);
int gs_context_destroy(gs_context ctx);
int gs_context_synchronize(gs_context ctx);
int gs_context_stream(gs_context ctx, void** hip_stream);

/* ---- bin lattice (KNNBinData) ------------------------------------------------ */
/* origin = minimum corner, size = extent, cells = divisions per axis. */
int gs_bins_create(gs_context ctx, const float origin[3], const float size[3], const uint32_t cells[3], gs_bins* out);
/* The reference's centred form, KNNBinDataV2(worldCells, worldSize, searchRadius)
 * (CUDA/KNNBinDataV2.cu:30-42): origin = -size/2. */
int gs_bins_create_centered(gs_context ctx, const uint32_t world_cells[3], const float world_size[3], gs_bins* out);
int gs_bins_destroy(gs_bins bins);
int gs_bins_geometry(gs_bins bins, float origin[3], float size[3], uint32_t cells[3]);

/* ---- groups (AgentGroup / ObstacleGroup / WallGroup) ------------------------- */
int gs_group_create(gs_context ctx, int kind, gs_group* out);
int gs_group_destroy(gs_group g);
int gs_group_size(gs_group g, uint32_t* n);
/* Agents: every pointer of *host must be non-NULL (n > 0).  Synchronous. */
int gs_group_upload_agents(gs_group g, uint32_t n, const gs_agents* host);
int gs_group_download_agents(gs_group g, gs_agents* host);
/* Device pointers of the agent arrays, in storage order: each bin build
 * (gs_update_knn_database, gs_step_boids) reorders a group's state into
 * cell order, so storage slot s holds agent id_of_slot[s] (see
 * gs_group_device_order).  Valid until the next build, upload,
 * gs_step_boids or gs_anti_penetration_agents on the group (these write
 * the group's alternate buffers and switch to them).  For a wall group the
 * arrays px/py/pz hold the KNN positions (midpoints), fx/fy/fz the start,
 * speed/ax/ay the end and az/sx/sy the normal points (fields 3..11 of
 * gs_agents); for an obstacle group px/py/pz the centres and radius the
 * radii; the other pointers are NULL for those kinds. */
int gs_group_device_agents(gs_group g, gs_agents* dev);
/* The group's storage order: agent index of each slot and slot of each agent
 * index (device arrays of n entries; valid until the next build or upload). */
int gs_group_device_order(gs_group g, uint32_t** id_of_slot, uint32_t** slot_of_id);
/* Obstacles: centre and radius per obstacle (ObstacleGroup::AddObstacle). */
int gs_group_upload_obstacles(gs_group g, uint32_t n, const float* cx, const float* cy, const float* cz,
                              const float* radius);
/* Walls: segments; the KNN position of a wall is its midpoint (WallGroup.h:31). */
int gs_group_upload_walls(gs_group g, uint32_t n, const gs_walls* host);
/* Copy the bin database of a group (after gs_update_knn_database): any output
 * may be NULL.  key[n] cell per agent; start[C+1] first sorted slot per cell;
 * perm[n] agent index per sorted slot (ascending within a cell; a cell of
 * more than 65,536 agents keeps the order in which its agents were counted);
 * sorted positions spx/spy/spz[n]. */
int gs_group_download_bins(gs_group g, uint32_t* key, uint32_t* start, uint32_t* perm, float* spx, float* spy,
                           float* spz);
/* The lattice of the group's last bin build (KNNDatabase::worldCells,
 * include/OpenSteer/CUDA/KNNDatabase.cuh:80); GS_ERR_NOT_BINNED before the
 * first build.  Any output may be NULL. */
int gs_group_bin_geometry(gs_group g, float origin[3], float size[3], uint32_t cells[3]);
/* Scan the agent state for NaN/Inf: GS_NONFINITE if any, *first_bad = agent index. */
int gs_group_check_finite(gs_group g, uint32_t* first_bad);

/* ---- KNN container (KNNData) ----------------------------------------------------- */
int gs_knn_create(gs_context ctx, uint32_t n, uint32_t k, gs_knn* out);
int gs_knn_destroy(gs_knn knn);
int gs_knn_shape(gs_knn knn, uint32_t* n, uint32_t* k);
/* Slot-major [k][n] in agent indices: idx[j*n + i] is the agent index (in
 * B) of the j-th nearest neighbour of agent i (of A); pads are UINT32_MAX /
 * FLT_MAX.  Either output may be NULL. */
int gs_knn_download(gs_knn knn, uint32_t* idx, float* dist);
/* Device buffers of the current result in storage slots ([k][n]: row = slot
 * of the query in A, value = slot of the neighbour in B).  Marks the
 * container's contents as no longer trusted by later calls. */
int gs_knn_device(gs_knn knn, uint32_t** idx, float** dist);

/* ---- operator API (GroupSteerLibrary.cuh) ---------------------------------------- */
int gs_update_knn_database(gs_group g, gs_bins bins);
/* Exact k nearest of B for each agent of A, key (d^2, index).  A and B must
 * both have a current bin database built with `bins`.  search_radius is the
 * cell radius scanned before the first exactness test (the result does not
 * depend on it).  Unless GS_KNN_NO_SEED, a previous exact result held by `out`
 * for the same A, B and memberships bounds the search (result unchanged).
 * Unless GS_KNN_BINNED, a small A searching another small group B (at most
 * 16384 agents each, at most 2^26 pairs) over more than 64 cells of reach is
 * served by an exhaustive scan of B's snapshot split over many wavefronts,
 * and a small or sparse B by an exhaustive scan: the same exact result. */
int gs_find_knn(gs_group A, gs_knn out, gs_bins bins, gs_group B, uint32_t search_radius, uint32_t flags);
/* gs_find_knn restricted to points closer than max_distance (strict, the
 * locality test of src/lq.c:327: d^2 < r^2 in fp32); rows with fewer than k
 * such points are padded.  Gives the same forces as the uncapped search for
 * any behaviour whose neighbourhood lies inside the cap (e.g. separation
 * from another group, SteerLibrary.cpp:524-541), at a cost independent of
 * how far away the rest of that group is.  INFINITY = no cap. */
int gs_find_knn_within(gs_group A, gs_knn out, gs_bins bins, gs_group B, uint32_t search_radius, uint32_t flags,
                       float max_distance);
int gs_find_knn_brute(gs_group A, gs_knn out, gs_group B);
int gs_steer_separation(gs_group A, gs_knn knn, gs_group B, float min_d, float max_d, float cos_max, float weight,
                        uint32_t mask, int mode);
int gs_steer_alignment(gs_group A, gs_knn knn, gs_group B, float min_d, float max_d, float cos_max, float weight,
                       uint32_t mask, int mode);
int gs_steer_cohesion(gs_group A, gs_knn knn, gs_group B, float min_d, float max_d, float cos_max, float weight,
                      uint32_t mask, int mode);
/* The three flocking terms in one pass, applied in the order separation,
 * alignment, cohesion (bit-identical to the three calls in that order). */
int gs_steer_flocking(gs_group A, gs_knn knn, gs_group B, const gs_flock_term* separation,
                      const gs_flock_term* alignment, const gs_flock_term* cohesion, int mode);
int gs_steer_seek(gs_group A, const float target[3], float weight, uint32_t mask, int mode);
int gs_steer_avoid_obstacles(gs_group A, gs_group obstacles, gs_knn knn, float min_time_to_collision, float weight,
                             uint32_t mask);
int gs_steer_avoid_walls(gs_group A, gs_knn knn, gs_group walls, float min_time_to_collision, float weight,
                         uint32_t mask);
int gs_update_group(gs_group A, float elapsed_time, int mode);

/* ---- per-cell extraction (density and mean velocity) ---------------------- */
/* Of the group's last bin build (gs_update_knn_database), C = cells of its
 * lattice, in KNNBinDataV2's cell numbering ix + iz*X + iy*X*Z
 * (CUDA/KNNBinDataV2.cu:103-135).  Density: agents per cell.  Velocity:
 * the mean of forward * speed over each cell's agents (summed in agent
 * index order -- for a cell of more than 65,536 agents in the order its
 * agents were counted --, divided by the count), (0, 0, 0) for an empty
 * cell; the current forward and speed of the agents binned there. */
int gs_group_cell_density(gs_group g, uint32_t* agents_per_cell /* host [C] */);
int gs_group_cell_velocity(gs_group g, float* avg_velocity /* host [C][3] */);
/* The reference's text files: "<prefix>.dat" (overwritten) with its header
 * and rows i = 0 .. Z-2 of cells i*X + j (all Z rows when all_rows != 0; the
 * reference stops one row short); each selected cell's value appended to
 * "<selected_prefix>_<cell>.dat" (the reference uses "Frames/cell_density"
 * and "Frames/cell_avg_velocity").  Floats print as std::ostream's default
 * (%g). */
int gs_write_cell_density(gs_group g, const char* prefix, const uint32_t* selected_cells, uint32_t n_selected,
                          const char* selected_prefix, int all_rows);
int gs_write_cell_velocity(gs_group g, const char* prefix, const uint32_t* selected_cells, uint32_t n_selected,
                           const char* selected_prefix, int all_rows);

/* ---- plugin-compat kernels (the choke-point plugins' per-step extras) ----- */
/* Closest point of a wall segment for the within-radius test of
 * gs_update_group_anti_penetration.  The reference's ClosestPointOnLine
 * (CUDA/UpdateCUDAKernelNew.cu:73-81) swaps clamp's arguments and always
 * returns the point start + (end - start) * 1: GS_CLOSEST_AS_REFERENCE keeps
 * that; GS_CLOSEST_CLAMPED clamps the segment parameter into [0, 1]. */
enum { GS_CLOSEST_CLAMPED = 0, GS_CLOSEST_AS_REFERENCE = 1 };
/* UpdateCUDAKernelNew (CUDA/UpdateCUDAKernelNew.cu:95-268): gs_update_group's
 * GPU-reference integrate with wall anti-penetration over the k nearest walls
 * (knn: A to walls).  A wall within the radius moves the agent out along its
 * normal; crossing a wall during the step removes the velocity's component
 * along that wall's normal; either sets GS_BIT_ANTI_PENETRATION_WALL.
 * Steering is zeroed; unlike gs_update_group the bits are kept. */
int gs_update_group_anti_penetration(gs_group A, gs_knn knn, gs_group walls, float elapsed_time, int closest_point);
/* AntiPenetrationAgentsCUDAKernel (CUDA/AntiPenetrationAgentsCUDAKernel.cu:57-109):
 * each agent of A without a masked bit whose nearest agent of B (the first
 * KNN entry, at the KNN's distance) overlaps it moves apart by the overlap
 * (the whole of it when the other agent has a masked bit, else half) and
 * gets GS_BIT_ANTI_PENETRATION_AGENT.  Reads the state before the call, so B
 * may be A.  Allocates nothing per call. */
int gs_anti_penetration_agents(gs_group A, gs_knn knn, gs_group B, uint32_t mask);
/* Polyline path (PolylinePathwayCUDA): n_points xyz triples (>= 2), tube
 * radius; cyclic closes it on the first point. */
typedef struct gs_path_s* gs_path;
int gs_path_create(gs_context ctx, uint32_t n_points, const float* points, float radius, int cyclic, gs_path* out);
/* m = stored points (n_points + 1 when cyclic) and the total length. */
int gs_path_info(gs_path p, uint32_t* m, float* total_length);
/* Host copies (hvPoints / hvNormals / hvLengths): points[m][3], unit
 * segment directions[m][3] and lengths[m] of the segment ending at each point
 * (entry 0 is 0).  Any output may be NULL. */
int gs_path_get(gs_path p, float* points, float* directions, float* lengths);
int gs_path_destroy(gs_path p);
/* FollowPathCUDAKernel (CUDA/SteerToFollowPathCUDAKernel.cu:189-287): seek
 * the point prediction_time * speed ahead along the path unless the predicted
 * position is inside the tube and moving forward; sets GS_BIT_FOLLOW_PATH. */
int gs_steer_follow_path(gs_group A, gs_path path, float prediction_time, float weight, uint32_t mask);
/* Unaligned collision avoidance against B over A's KNN in B (B is only read;
 * it may be A).  GS_MODE_GPUREF = SteerToAvoidNeighborsCUDA
 * (CUDA/SteerToAvoidNeighborsCUDA.cu:85-150): with avoid_close_neighbors,
 * agents whose KNN distance to a neighbour is below min_separation_distance +
 * both radii steer away from it (GS_BIT_AVOID_CLOSE_NEIGHBORS); the others
 * steer sideways from the neighbour of earliest nearest approach within
 * min_time_to_collision that comes closer than both radii
 * (GS_BIT_AVOID_NEIGHBORS).  GS_MODE_OPENSTEER = SteerLibrary's
 * steerToAvoidNeighbors (src/SteerLibrary.cpp:290-389): close neighbours
 * always checked with separation 0, threshold 2 r, forces unnormalized. */
int gs_steer_avoid_neighbors(gs_group A, gs_knn knn, gs_group B, float min_time_to_collision,
                             float min_separation_distance, int avoid_close_neighbors, float weight, uint32_t mask,
                             int mode);
/* GS_WRAP_BOX: size_or_radius = box size (centred at center);
 * GS_WRAP_SPHERE: size_or_radius[0] = radius. */
int gs_wrap_world(gs_group A, int shape, const float center[3], const float size_or_radius[3]);
int gs_clear_bits(gs_group A);

/* ---- launch shapes ----------------------------------------------------------------
 * The k = 7 exact search (the 3 x 3 flat path of gs_find_knn* and the fused
 * step) and the k = 7 flocking + integrate kernel of gs_step_boids run in a
 * launch shape the caller may set per context: block_threads in {128, 256,
 * 512} threads per block and agents_per_lane in {1, 2} (a lane handles that
 * many agent slots, so a wavefront covers 64 x agents_per_lane agents).
 * Default 256 x 1.  Results do not depend on the shape.  gpusteer_amd's
 * autotuner picks the shapes by timing them on the workload.  Other values
 * return GS_ERR_UNSUPPORTED.  No counterpart in the reference (fixed
 * block size THREADSPERBLOCK = 256, include/OpenSteer/CUDA/CUDAKernelGlobals.cuh:9). */
enum { GS_SHAPE_KNN = 0, GS_SHAPE_FLOCK = 1 };
int gs_context_set_shape(gs_context ctx, int kernel, uint32_t block_threads, uint32_t agents_per_lane);
int gs_context_get_shape(gs_context ctx, int kernel, uint32_t* block_threads, uint32_t* agents_per_lane);

/* ---- fused Boids step ------------------------------------------------------------ */
typedef struct {
    int wrap_shape;               /* GS_WRAP_BOX / GS_WRAP_SPHERE, or -1 for none */
    float wrap_center[3];
    float wrap_size[3];
    uint32_t search_radius;
    uint32_t knn_flags;
    gs_flock_term separation, alignment, cohesion;
    float dt;
    int mode;
} gs_boids_params;
/* `steps` times: wrap -> bin database -> exact KNN (A with itself) ->
 * separation, alignment, cohesion -> integrate.  Same results as the separate
 * calls in that order. */
int gs_step_boids(gs_group A, gs_knn knn, gs_bins bins, const gs_boids_params* params, uint32_t steps);

/* gs_step_boids with HIP-event timing recorded on the context's stream:
 * average milliseconds per step of the bin build (key, scan, scatter,
 * reorder), of the exact KNN, and of the flocking + integrate kernel. */
typedef struct {
    float bin_build_ms;
    float knn_ms;
    float steer_integrate_ms;
    float step_ms;
} gs_boids_timing;
int gs_step_boids_timed(gs_group A, gs_knn knn, gs_bins bins, const gs_boids_params* params, uint32_t steps,
                        gs_boids_timing* timing);

/* ---- fused choke-point crowd step -------------------------------------------
 * `steps` times the per-frame sequence of plugins/ChokePointGroup.cpp:683-728
 * on two agent groups A[0], A[1] with a wall group and an obstacle group,
 * all binned over `bins` (walls and obstacles once, beforehand):
 * gs_update_knn_database on both groups; then per group g = 0, 1:
 * gs_clear_bits, gs_find_knn of A[g] with itself into knn[4g], of A[g] in
 * A[1-g] within s->other_max_distance into knn[4g+1] (no cap when <= 0),
 * of A[g] in the walls into knn[4g+2] and in the obstacles into knn[4g+3],
 * gs_steer_avoid_walls, gs_steer_avoid_obstacles, gs_steer_separation
 * (s->choke_separation) from A[g] and from A[1-g], gs_steer_seek of
 * s->goal[g]; then gs_update_group on both (s->dt, s->mode).  Same results
 * as those calls, issued from the library instead of the caller.  Reads
 * s's crowd fields only.  No single reference counterpart (a plugin's frame). */
typedef struct gs_scenario_s gs_scenario;
int gs_step_choke(gs_group A0, gs_group A1, gs_group walls, gs_group obstacles, gs_bins bins, gs_knn const* knn,
                  const gs_scenario* s, uint32_t steps);

/* ---- recorded frames --------------------------------------------------------------
 * A frame is a call sequence on one context recorded once into a HIP graph
 * and replayed: a plugin's per-frame calls (or a fused step) without the host
 * cost of issuing every launch again.  No counterpart in the reference, whose
 * plugins launch every kernel eagerly each frame (plugins/ChokePointNew.cpp,
 * plugins/ChokePointGroup.cpp:683-728).
 *
 * gs_frame_begin(ctx): the calls that follow on ctx, from this thread, are
 * recorded instead of run.  Recordable are the calls that only launch
 * kernels on existing handles: gs_update_knn_database, the gs_find_knn*
 * searches (gs_find_knn_noalloc included), the gs_steer_* calls,
 * gs_update_group*, gs_anti_penetration_agents, gs_wrap_world,
 * gs_clear_bits, gs_step_boids and gs_step_choke.  A call that would copy
 * between host and device, allocate or free device memory, or synchronize
 * (creating, destroying, uploading, downloading, extraction to the host,
 * gs_group_check_finite, gs_step_boids_timed, the ensemble and calibration
 * calls, a search whose buffers must first grow) fails with GS_ERR_FRAME
 * while recording; run the sequence once unrecorded first, so every buffer
 * has its size.  Work the caller queues on the context's stream meanwhile is
 * recorded too.
 * gs_frame_end(ctx, &f): ends the recording.  The recorded calls have not
 * run: every handle they touched is back in its state at gs_frame_begin.
 * Fails with GS_ERR_FRAME (and records nothing) if a call failed while
 * recording.
 * gs_frame_launch(f, times): runs the recorded calls `times` times on the
 * context's stream, with the same results as issuing them.  Each replay
 * requires every handle the calls touched to be in the state it had at
 * gs_frame_begin (the same sizes, device buffers and current double buffers,
 * and KNN containers holding results for the same groups and orders, up to
 * later reorders); otherwise it fails with GS_ERR_FRAME before launching.  A
 * Boids or crowd step swaps double buffers, so a frame of an even number of
 * steps replays any number of times.
 * gs_frame_destroy(f): releases the graph (handles are not touched). */
typedef struct gs_frame_s* gs_frame;
int gs_frame_begin(gs_context ctx);
int gs_frame_end(gs_context ctx, gs_frame* out);
int gs_frame_launch(gs_frame f, uint32_t times);
int gs_frame_destroy(gs_frame f);

/* ---- batched ensemble -------------------------------------------------------------
 * Many independent scenarios of uneven size stepped together: every phase of a
 * step is one launch over all scenarios and the steps are replayed from a
 * captured hipGraph.  Each scenario's results are identical, bit for bit, to
 * running it alone through the calls above:
 *   GS_SCENARIO_BOIDS  the gs_step_boids sequence (wrap, bins, KNN, flocking,
 *                      integrate) on one group;
 *   GS_SCENARIO_CHOKE  two groups; per group in order 0, 1: gs_clear_bits,
 *                      KNN to itself / to the other group (capped at
 *                      other_max_distance) / to the walls / to the
 *                      obstacles, gs_steer_avoid_walls,
 *                      gs_steer_avoid_obstacles, gs_steer_separation (own
 *                      group, other group), gs_steer_seek (the group's goal);
 *                      then gs_update_group on both
 *                      (plugins/ChokePointGroup.cpp:683-728).
 * No counterpart in the reference, which runs one scenario per process. */
typedef struct gs_ensemble_s* gs_ensemble;
enum { GS_SCENARIO_BOIDS = 0, GS_SCENARIO_CHOKE = 1 };
struct gs_scenario_s {
    int kind;                          /* GS_SCENARIO_BOIDS / GS_SCENARIO_CHOKE */
    uint32_t n[2];                     /* agents of group 0 (and group 1 for a crowd) */
    float origin[3], size[3];          /* this scenario's bin lattice */
    uint32_t cells[3];
    float dt;
    int mode;                          /* GS_MODE_* */
    /* Boids */
    int wrap_shape;                    /* GS_WRAP_BOX / GS_WRAP_SPHERE / -1 */
    float wrap_center[3], wrap_size[3];
    gs_flock_term separation, alignment, cohesion;
    /* choke-point crowd */
    gs_flock_term choke_separation;
    float goal[2][3];
    float seek_weight;
    uint32_t seek_mask;
    float walls_min_time, walls_weight, obstacles_min_time, obstacles_weight;
    uint32_t k_walls, k_obstacles;     /* 1..8 */
    float other_max_distance;          /* KNN to the other group: gs_find_knn_within cap (<= 0: none) */
};
/* k: neighbours per agent (1..8), the same for every scenario. */
int gs_ensemble_create(gs_context ctx, uint32_t n_scenarios, const gs_scenario* scenarios, uint32_t k,
                       gs_ensemble* out);
int gs_ensemble_destroy(gs_ensemble e);
/* Total slots (agents plus per-group padding to 256) of the ensemble. */
int gs_ensemble_slots(gs_ensemble e, uint32_t* slots);
int gs_ensemble_upload_agents(gs_ensemble e, uint32_t scenario, uint32_t group, const gs_agents* host);
int gs_ensemble_download_agents(gs_ensemble e, uint32_t scenario, uint32_t group, gs_agents* host);
int gs_ensemble_upload_walls(gs_ensemble e, uint32_t scenario, uint32_t n, const gs_walls* host);
/* The same with each wall's KNN position given (walls split at bin-cell
 * borders keep the midpoints SplitWalls gave them, WallGroupData.cu:185-190). */
int gs_ensemble_upload_walls_mid(gs_ensemble e, uint32_t scenario, uint32_t n, const gs_walls* host,
                                 const float* mid_x, const float* mid_y, const float* mid_z);
int gs_ensemble_upload_obstacles(gs_ensemble e, uint32_t scenario, uint32_t n, const float* cx, const float* cy,
                                 const float* cz, const float* radius);
/* The last KNN of a group with itself, [k][n] agent indices (pads UINT32_MAX). */
int gs_ensemble_download_knn(gs_ensemble e, uint32_t scenario, uint32_t group, uint32_t* idx);
enum { GS_ENSEMBLE_GRAPH = 0, GS_ENSEMBLE_EAGER = 1 };
/* Advance every scenario `steps` steps.  GS_ENSEMBLE_GRAPH replays captured
 * pairs of steps (the first step after an upload runs eagerly); the results
 * do not depend on the flag. */
int gs_ensemble_step(gs_ensemble e, uint32_t steps, uint32_t flags);

/* ---- differentiable calibration (adjoint) -----------------------------------------
 * Gradients of a trajectory loss over crowd (GS_SCENARIO_CHOKE) scenarios with
 * respect to per-group steering weights, agent radius and max speed.  One
 * step of the model, per agent of group g, with every neighbour read from the
 * step's start (Jacobi) and the reference CUDA kernels' formulas: wall
 * avoidance with the anti-penetration move (weight w_wall[g]), obstacle
 * avoidance (w_obs[g]), separation from the own and from the other group
 * (w_sep[g]), alignment and cohesion with the own group (w_ali[g], w_coh[g]),
 * seek of the group's goal, then the UpdateCUDAKernel integrate with max speed
 * max_speed[g]; agent radius radius[g].  KNN lists (own group k, other group
 * within the scenario's other_max_distance, walls, obstacles) come from the
 * step's start positions and are held fixed under differentiation, as are all
 * other branches.
 *   loss = mean over scenarios of [ mean_i |p_T,i - goal_i|^2
 *          + penetration_weight * sum_t sum_i sum_{j in the two lists of i}
 *            max(0, r_i + r_j - |p_j - p_i|)^2 ]
 * gs_calib_backward runs the hand-written adjoint backwards through the steps,
 * recomputing each checkpoint_every-step segment from its checkpoint.
 * No counterpart in the reference. */
typedef struct gs_calib_s* gs_calib;
typedef struct {
    float w_sep[2], w_ali[2], w_coh[2], w_obs[2], w_wall[2], radius[2], max_speed[2];
} gs_calib_params;
typedef struct {
    gs_flock_term separation, alignment, cohesion; /* distances, cone, mask (weights: gs_calib_params) */
    float seek_weight;
    uint32_t seek_mask;
    float walls_min_time, obstacles_min_time;
    float max_force, mass, dt;
    float penetration_weight;
    uint32_t steps, checkpoint_every;
} gs_calib_config;
/* The initial state is the ensemble's agent state at this call; every
 * scenario of e must be GS_SCENARIO_CHOKE.  e must outlive c. */
int gs_calib_create(gs_ensemble e, const gs_calib_config* config, gs_calib* out);
int gs_calib_destroy(gs_calib c);
int gs_calib_forward(gs_calib c, const gs_calib_params* params, double* loss);
/* After gs_calib_forward: d(loss)/d(params) at the forward's params. */
int gs_calib_backward(gs_calib c, gs_calib_params* grad);
/* After gs_calib_forward: the final state (position, forward, speed, side, up). */
int gs_calib_download_state(gs_calib c, uint32_t scenario, uint32_t group, gs_agents* host);
/* HIP-event milliseconds of the last forward and backward passes. */
int gs_calib_timing(gs_calib c, float* forward_ms, float* backward_ms);

/* ---- allocation-free KNN ------------------------------------------------------
 * No counterpart in the reference: for callers that cannot allocate device
 * memory (work under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  The result of gs_update_knn_database on a group holding
 * these points followed by gs_find_knn of that group with itself over `bins`:
 * for each point the k nearest other points by (d^2, index), idx[j*n + i] the
 * point index of the j-th nearest of point i, dist = sqrtf(d^2), pads
 * UINT32_MAX / FLT_MAX.  px, py, pz (n floats each) and idx, dist (k*n each)
 * are device arrays.  The bin database and the intermediate lists live in
 * one fixed device array of the library (about 13 MB on each device the
 * library runs on, held until the process ends, located once per device by
 * gs_context_create); the call allocates and frees nothing.  Limits:
 * n <= GS_NOALLOC_MAX_POINTS, lattice cells <= GS_NOALLOC_MAX_CELLS,
 * 1 <= k <= GS_NOALLOC_MAX_K (GS_ERR_UNSUPPORTED beyond).  Every call on a
 * device shares that device's array, which keeps the last call's contents:
 * calls that overlap on one device (other threads, other streams, other
 * contexts) race and must be serialized by the caller.  Queued on the
 * context's stream; returns before the work completes. */
#define GS_NOALLOC_MAX_POINTS 65536u
#define GS_NOALLOC_MAX_CELLS 131072u
#define GS_NOALLOC_MAX_K 16u
int gs_find_knn_noalloc(gs_context ctx, gs_bins bins, uint32_t n, const float* px, const float* py, const float* pz,
                        uint32_t k, uint32_t* idx, float* dist);

/* ---- Scenario formats (host only, no device work) ----------------------
 * Wall maps, wall splitting at bin-cell borders, simulation parameter files:
 * what the plugins load before the per-step path runs. */
typedef struct gs_wall_set_s* gs_wall_set;
/* WallGroupData::LoadFromFile (include/OpenSteer/CUDA/WallGroupData.cu:73-103):
 * records of start xyz, end xyz, normal xyz; midpoints 0.5f * (start + end). */
int gs_wall_set_load_map(const char* path, gs_wall_set* out);
/* A set from arrays (midpoints 0.5f * (start + end)). */
int gs_wall_set_create(uint32_t n, const gs_walls* host, gs_wall_set* out);
/* WallGroupData::SplitWalls (WallGroupData.cu:105-206) over the cells of
 * KNNBinDataV2::CreateCells (CUDA/KNNBinDataV2.cu:73-136): a lattice of
 * `cells` over a box of `world_size` centred at the origin, at most
 * GS_MAX_SPLIT_CELLS cells (GS_ERR_UNSUPPORTED beyond). */
int gs_wall_set_split(gs_wall_set w, const float world_size[3], const uint32_t cells[3]);
/* The same over any list of cells, tried in list order: cell i spans
 * [min_bounds[3i..3i+2], max_bounds[3i..3i+2]] (bin_cell::minBound / maxBound,
 * CUDA/KNNBinDataGlobals.cuh:7-13). */
int gs_wall_set_split_cells(gs_wall_set w, uint32_t n_cells, const float* min_bounds, const float* max_bounds);
/* Replaces the midpoints (KNN positions) of the set's n walls. */
int gs_wall_set_set_mid(gs_wall_set w, const float* mid_x, const float* mid_y, const float* mid_z);
int gs_wall_set_size(gs_wall_set w, uint32_t* n);
/* Copies the walls (any pointer may be NULL) into caller arrays of size n. */
int gs_wall_set_get(gs_wall_set w, gs_walls* out, float* mid_x, float* mid_y, float* mid_z);
int gs_wall_set_destroy(gs_wall_set w);
/* WallGroup::SyncDevice after LoadFromFile / SplitWalls
 * (include/OpenSteer/WallGroup.h:18-27): uploads the set with its own
 * midpoints as the walls' KNN positions. */
int gs_group_upload_wall_set(gs_group g, gs_wall_set w);
/* gs_group_upload_walls with explicit KNN positions (midpoints). */
int gs_group_upload_walls_mid(gs_group g, uint32_t n, const gs_walls* host, const float* mid_x, const float* mid_y,
                              const float* mid_z);

#define GS_PARAMS_MAX_GROUPS 32u
#define GS_PARAMS_MAX_PATH 64u
/* GroupParams (include/OpenSteer/Simulation.h:16-48). */
typedef struct {
    uint32_t num_agents;
    float max_speed, max_force;
    float start_position[3];
    float min_start_radius, max_start_radius;
    uint32_t path_points;
    float path[GS_PARAMS_MAX_PATH][3];
    float path_radius;
    uint32_t path_cyclic;
    float body_color[3];
} gs_group_params;
/* WorldParams (Simulation.h:50-64). */
typedef struct {
    float dimensions[3];
    uint32_t cells[3];
    uint32_t obstacle_count;
    char map_filename[256];
} gs_world_params;
/* SimulationParams (Simulation.h:66-120); fields a file does not set are 0. */
typedef struct {
    gs_world_params world;
    uint32_t n_groups;
    gs_group_params groups[GS_PARAMS_MAX_GROUPS];
    uint32_t seed, knn, kno, knw, search_radius;
    float max_pursuit_prediction_time, min_time_to_collision, min_time_to_obstacle, min_time_to_wall,
        path_prediction_time, min_separation_distance, max_separation_distance, min_flocking_distance,
        max_flocking_distance, cos_max_flocking_angle;
    uint32_t avoid_close_neighbors;
    float weight_alignment, weight_cohesion, weight_separation, weight_pursuit, weight_seek, weight_flee,
        weight_evade, weight_follow_path, weight_avoid_obstacles, weight_avoid_walls, weight_avoid_neighbors;
    uint32_t mask_alignment, mask_anti_penetration_agents, mask_cohesion, mask_separation, mask_pursuit,
        mask_seek, mask_flee, mask_evade, mask_follow_path, mask_avoid_obstacles, mask_avoid_walls,
        mask_avoid_neighbors;
} gs_sim_params;
/* SimulationParams::load (src/Simulation.cpp:7-250).  Unknown keys, bad
 * numbers, a missing file or a missing </simulation> return
 * GS_ERR_INVALID_ARGUMENT (the reference asserts or does not return). */
int gs_sim_params_load(const char* path, gs_sim_params* out);

#ifdef __cplusplus
}
#endif
#endif /* GPUSTEER_H */
