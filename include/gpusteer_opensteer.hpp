/*
 * gpusteer_opensteer.hpp — the reference's OpenSteer group API on top of the
 * gpusteer C-ABI, so a plugin written against the reference compiles with a
 * change of includes only.
 *
 * Re-created names (reference file:line):
 *   AgentData                     include/OpenSteer/AgentData.h:13-42
 *   ObstacleData                  include/OpenSteer/CUDA/ObstacleGroupData.cuh:10-14
 *   BaseGroup                     include/OpenSteer/BaseGroup.h:8-33 (GetKNNDatabase, Size, pd*)
 *   KNNDatabase                   include/OpenSteer/CUDA/KNNDatabase.cuh:30-123 (host view)
 *   AgentGroup                    include/OpenSteer/AgentGroup.h:12-56, src/AgentGroup.cpp:8-140
 *   AgentGroupData                include/OpenSteer/AgentGroupData.cuh:19-297 (getAgentData, size, pd*)
 *   ObstacleGroup                 include/OpenSteer/ObstacleGroup.h:10-43
 *   WallGroup                     include/OpenSteer/WallGroup.h:12-37 (LoadFromFile, SplitWalls, AddWall,
 *                                 GetWallGroupData)
 *   WallGroupData                 include/OpenSteer/CUDA/WallGroupData.cuh:12-58 (hvLine*, size, pdLine*)
 *   KNNData                       include/OpenSteer/CUDA/KNNData.cuh:10-55
 *   KNNBinData, bin_cell          include/OpenSteer/CUDA/KNNBinDataV2.cuh:43-61, KNNBinDataGlobals.cuh:7-13
 *   SimulationParams, WorldParams,
 *   GroupParams                   include/OpenSteer/Simulation.h:16-120, src/Simulation.cpp:7-250
 *   wrapWorld .. updateGroup      include/OpenSteer/CUDA/GroupSteerLibrary.cuh:43-189
 *
 * Differences a plugin can observe:
 *   - failures throw gpusteer::Error (the reference aborts through
 *     CUDA_SAFE_CALL); nothing is left half-applied on the device;
 *   - every group / container / lattice belongs to a gpusteer::Context; the
 *     constructors without one use a process-wide context on device 0;
 *   - the force and integrate formulas are those of the reference CUDA
 *     kernels (GS_MODE_GPUREF) unless gpusteer::Context::SetMode selects the
 *     CPU OpenSteer formulas (GS_MODE_OPENSTEER);
 *   - the binned KNN is exact (the reference's V2 kernel scans a fixed cell
 *     radius and can miss neighbours);
 *   - device arrays are struct-of-arrays in the library's storage order (a
 *     bin build puts each group's agents in cell order): pdSpeed(),
 *     pdRadius() and the other scalar fields are float* as in the reference,
 *     but slot s holds agent pdID()[s], not agent s; pdPosition(),
 *     pdDirection() and the other vector fields return gpusteer::DeviceVec3
 *     (three float arrays) where the reference returns float4* / float3*.
 *     Host views (getAgentData, hvLine*, KNNDatabase's hv*) are in agent
 *     index order as in the reference.
 */
#ifndef GPUSTEER_OPENSTEER_HPP
#define GPUSTEER_OPENSTEER_HPP

#include <hip/hip_vector_types.h>

#include <cstdint>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>

#include "gpusteer.h"

namespace gpusteer {

class Error : public std::runtime_error {
public:
    Error(int status, const std::string& what) : std::runtime_error(what), status_(status) {}
    int status() const { return status_; }

private:
    int status_;
};

inline void check(int status, const char* call) {
    if (status != GS_OK)
        throw Error(status, std::string(call) + ": " + gs_status_string(status) + ": " + gs_last_error());
}

class Context {
public:
    explicit Context(int device = 0, void* hip_stream = nullptr) {
        check(gs_context_create(device, hip_stream, &h_), "gs_context_create");
    }
    ~Context() { gs_context_destroy(h_); }
    Context(const Context&) = delete;
    Context& operator=(const Context&) = delete;
    gs_context handle() const { return h_; }
    void Synchronize() { check(gs_context_synchronize(h_), "gs_context_synchronize"); }
    int Mode() const { return mode_; }
    void SetMode(int mode) { mode_ = mode; }
    // Closest-point rule of updateGroupWithAntiPenetration: GS_CLOSEST_CLAMPED
    // (default) or GS_CLOSEST_AS_REFERENCE (the reference's wall-end point).
    int ClosestPoint() const { return closest_; }
    void SetClosestPoint(int rule) { closest_ = rule; }
    // The context used by constructors that are not given one.
    static Context& Default() {
        static Context ctx(0, nullptr);
        return ctx;
    }

private:
    gs_context h_ = nullptr;
    int mode_ = GS_MODE_GPUREF;
    int closest_ = GS_CLOSEST_CLAMPED;
};

// A device vector field as three float arrays (the library's struct of
// arrays), in storage order: what pdPosition() and the other vector-field
// accessors return where the reference returns float4* / float3*.
struct DeviceVec3 {
    float* x;
    float* y;
    float* z;
};

}  // namespace gpusteer

namespace OpenSteer {

typedef unsigned int uint;
typedef unsigned int id_type;

struct AgentData {
    id_type id = 0;
    float3 side = make_float3(0.f, 0.f, 0.f);
    float3 up = make_float3(0.f, 0.f, 0.f);
    float4 direction = make_float4(0.f, 0.f, 0.f, 0.f);
    float4 position = make_float4(0.f, 0.f, 0.f, 0.f);
    float4 steering = make_float4(0.f, 0.f, 0.f, 0.f);
    float speed = 0.f;
    float maxSpeed = 0.f;
    float maxForce = 0.f;
    float radius = 0.f;
    float mass = 0.f;
    uint appliedKernels = 0u;
};

struct ObstacleData {
    float4 position = make_float4(0.f, 0.f, 0.f, 0.f);
    float radius = 0.f;
};

// Common base of the three group kinds (include/OpenSteer/BaseGroup.h:8-33).
struct bin_cell {
    size_t index;      // index of the cell
    float3 position;   // centre
    float3 minBound;
    float3 maxBound;
};

class BaseGroup;

// The host view of a group's bin database after updateKNNDatabase
// (KNNDatabase, include/OpenSteer/CUDA/KNNDatabase.cuh:30-123): in the
// reference's cell numbering ix + iz X + iy X Z (CUDA/KNNBinDataV2.cu:103-135)
// and sort order (agents stably sorted by that cell index, the result of its
// thrust::sort_by_key), downloaded on first access after each build.  The
// device arrays of the database are the library's own (storage order, its
// cell numbering) and are not exposed.
class KNNDatabase {
public:
    KNNDatabase(BaseGroup& g, uint k, uint3 const& worldCells)
        : g_(&g), k_(k), cells_(worldCells.x * worldCells.y * worldCells.z), worldCells_(worldCells) {}
    uint const& k() const { return k_; }
    uint const& size() const { return n_; }
    uint const& cells() const { return cells_; }
    uint3 const& worldCells() const { return worldCells_; }
    // Cell of each agent (agent index order), the cells and agent indices in
    // sorted order, and the positions in sorted order (w = 0).
    std::vector<uint> const& hvCellIndices() { syncHost(); return cell_; }
    std::vector<uint> const& hvCellIndicesSorted() { syncHost(); return cell_sorted_; }
    std::vector<uint> const& hvAgentIndicesSorted() { syncHost(); return agent_sorted_; }
    std::vector<float4> const& hvPositionSorted() { syncHost(); return pos_sorted_; }
    void setSyncHost() { host_valid_ = false; }
    inline void syncHost();

private:
    BaseGroup* g_;
    uint k_, n_ = 0, cells_;
    uint3 worldCells_;
    std::vector<uint> cell_, cell_sorted_, agent_sorted_;
    std::vector<float4> pos_sorted_;
    bool host_valid_ = false;
};

class BaseGroup {
public:
    BaseGroup(int kind, uint3 const& worldCells, uint k, gpusteer::Context& ctx)
        : ctx_(&ctx), kind_(kind), worldCells_(worldCells), k_(k), db_(*this, k, worldCells) {
        gpusteer::check(gs_group_create(ctx.handle(), kind, &h_), "gs_group_create");
    }
    virtual ~BaseGroup() { gs_group_destroy(h_); }
    BaseGroup(const BaseGroup&) = delete;
    BaseGroup& operator=(const BaseGroup&) = delete;

    virtual uint const& Size() const { return n_; }
    uint3 const& WorldCells() const { return worldCells_; }
    uint K() const { return k_; }
    gpusteer::Context& context() const { return *ctx_; }
    // BaseGroup::GetKNNDatabase (include/OpenSteer/BaseGroup.h:24).
    KNNDatabase& GetKNNDatabase() { return db_; }
    // Uploads pending host edits; called by every operation before it runs.
    virtual void SyncDevice() {}
    gs_group device_handle() {
        SyncDevice();
        return h_;
    }
    // The handle without uploading pending host edits (host views).
    gs_group raw_handle() const { return h_; }
    // Called by every operation that may change the device state.
    virtual void SetSyncHost() { db_.setSyncHost(); }

    // BaseGroup::pdPosition / pdDirection / pdSpeed / pdRadius
    // (include/OpenSteer/BaseGroup.h:27-30): the group's device arrays in
    // storage order (slot s holds agent pdID()[s]), valid until the next
    // operation on the group.  A wall group's positions are its midpoints
    // and it has no direction, speed or radius (NULL, as WallGroup.h:29-31);
    // an obstacle group has positions and radii.
    gpusteer::DeviceVec3 pdPosition() { return vec3(0); }
    gpusteer::DeviceVec3 pdDirection() {
        return kind_ == GS_KIND_AGENTS ? vec3(3) : gpusteer::DeviceVec3{nullptr, nullptr, nullptr};
    }
    float* pdSpeed() { return kind_ == GS_KIND_AGENTS ? field(6) : nullptr; }
    float* pdRadius() { return kind_ == GS_KIND_WALLS ? nullptr : field(19); }
    // Agent index of each storage slot.
    uint* pdID() {
        uint32_t* ids = nullptr;
        gpusteer::check(gs_group_device_order(device_handle(), &ids, nullptr), "gs_group_device_order");
        return ids;
    }

protected:
    gs_agents device_arrays() {
        gs_agents a;
        gpusteer::check(gs_group_device_agents(device_handle(), &a), "gs_group_device_agents");
        return a;
    }
    // Field j of gs_agents (the order of include/gpusteer.h).
    float* field(int j) {
        const gs_agents a = device_arrays();
        return ((float* const*)&a)[j];
    }
    gpusteer::DeviceVec3 vec3(int j) {
        const gs_agents a = device_arrays();
        float* const* f = (float* const*)&a;
        return gpusteer::DeviceVec3{f[j], f[j + 1], f[j + 2]};
    }

    gpusteer::Context* ctx_;
    int kind_;
    gs_group h_ = nullptr;
    uint3 worldCells_;
    uint k_;
    uint n_ = 0;
    KNNDatabase db_;
};

inline void KNNDatabase::syncHost() {
    if (host_valid_) return;
    gs_group h = g_->raw_handle();
    uint32_t n = 0;
    gpusteer::check(gs_group_size(h, &n), "gs_group_size");
    uint32_t c[3];
    gpusteer::check(gs_group_bin_geometry(h, nullptr, nullptr, c), "gs_group_bin_geometry");
    std::vector<uint32_t> key(n), perm(n);
    std::vector<float> sx(n), sy(n), sz(n);
    gpusteer::check(gs_group_download_bins(h, key.data(), nullptr, perm.data(), sx.data(), sy.data(), sz.data()),
                    "gs_group_download_bins");
    // The library numbers cells ix*Y*Z + iy*Z + iz (src/lq.c's order); the
    // reference's KNNBinDataV2 numbers them ix + iz*X + iy*X*Z.
    const uint32_t X = c[0], Y = c[1], Z = c[2];
    n_ = n;
    worldCells_ = make_uint3(X, Y, Z);
    cells_ = X * Y * Z;
    cell_.resize(n);
    for (uint32_t a = 0; a < n; ++a) {
        const uint32_t ix = key[a] / (Y * Z), iy = (key[a] / Z) % Y, iz = key[a] % Z;
        cell_[a] = ix + iz * X + iy * X * Z;
    }
    std::vector<float4> pos(n);
    for (uint32_t t = 0; t < n; ++t) pos[perm[t]] = make_float4(sx[t], sy[t], sz[t], 0.f);
    // Stable sort of the agent indices by cell (ties in agent index order).
    std::vector<uint32_t> count((size_t)cells_ + 1, 0u);
    for (uint32_t a = 0; a < n; ++a) ++count[cell_[a] + 1];
    for (size_t i = 0; i < cells_; ++i) count[i + 1] += count[i];
    agent_sorted_.assign(n, 0u);
    for (uint32_t a = 0; a < n; ++a) agent_sorted_[count[cell_[a]]++] = a;
    cell_sorted_.resize(n);
    pos_sorted_.resize(n);
    for (uint32_t s = 0; s < n; ++s) {
        cell_sorted_[s] = cell_[agent_sorted_[s]];
        pos_sorted_[s] = pos[agent_sorted_[s]];
    }
    host_valid_ = true;
}

class AgentGroup;

// AgentGroupData (include/OpenSteer/AgentGroupData.cuh:19-297): a group's
// agents as plugins read them.  getAgentData(i, ad) is agent i of the group
// (insertion order, as the reference's arrays), synchronised from the device
// when an operation has run since the last read.  The pd* accessors return
// the device arrays in storage order (see BaseGroup::pdPosition).
class AgentGroupData {
public:
    explicit AgentGroupData(AgentGroup& g) : g_(&g) {}
    inline size_t size() const;
    inline void getAgentData(size_t const index, AgentData& ad);
    inline void syncHost();
    inline void syncDevice();
    inline void setSyncDevice();
    inline uint* pdAppliedKernels();
    inline uint* pdID();
    inline gpusteer::DeviceVec3 pdPosition();
    inline gpusteer::DeviceVec3 pdDirection();
    inline gpusteer::DeviceVec3 pdSide();
    inline gpusteer::DeviceVec3 pdUp();
    inline gpusteer::DeviceVec3 pdSteering();
    inline float* pdSpeed();
    inline float* pdMaxSpeed();
    inline float* pdMaxForce();
    inline float* pdRadius();
    inline float* pdMass();

private:
    AgentGroup* g_;
};

class AgentGroup : public BaseGroup {
    friend class AgentGroupData;

public:
    AgentGroup(uint3 const& worldCells, uint const knn, gpusteer::Context& ctx = gpusteer::Context::Default())
        : BaseGroup(GS_KIND_AGENTS, worldCells, knn, ctx), data_(*this) {}
    // AgentGroup::GetAgentGroupData (include/OpenSteer/AgentGroup.h:42).
    AgentGroupData& GetAgentGroupData() { return data_; }

    // Adds an agent unless its id is already present (src/AgentGroup.cpp:34-55).
    bool AddAgent(AgentData const& ad) {
        SyncHost();
        if (index_.count(ad.id)) return false;
        index_[ad.id] = agents_.size();
        agents_.push_back(ad);
        n_ = (uint)agents_.size();
        dirty_ = true;
        return true;
    }
    // Removes an agent, keeping the order of the others (src/AgentGroup.cpp:58-80).
    void RemoveAgent(id_type const id) {
        SyncHost();
        auto it = index_.find(id);
        if (it == index_.end()) return;
        agents_.erase(agents_.begin() + (long)it->second);
        index_.clear();
        for (size_t i = 0; i < agents_.size(); ++i) index_[agents_[i].id] = i;
        n_ = (uint)agents_.size();
        dirty_ = true;
    }
    void Clear() {
        agents_.clear();
        index_.clear();
        n_ = 0;
        dirty_ = true;
        host_valid_ = true;
    }
    // State of one agent as the device last left it (src/AgentGroup.cpp:96-112).
    bool GetDataForAgent(id_type const id, AgentData& ad) {
        SyncHost();
        auto it = index_.find(id);
        if (it == index_.end()) return false;
        ad = agents_[it->second];
        return true;
    }
    // Marks the host copies stale; the next host access downloads.
    void SetSyncHost() override {
        host_valid_ = false;
        BaseGroup::SetSyncHost();
    }

    void SyncDevice() override {
        if (!dirty_) return;
        const size_t n = agents_.size();
        std::vector<float> f(23 * (n ? n : 1));
        std::vector<uint32_t> bits(n ? n : 1);
        gs_agents a = soa(f.data(), bits.data(), n);
        for (size_t i = 0; i < n; ++i) {
            const AgentData& d = agents_[i];
            a.px[i] = d.position.x; a.py[i] = d.position.y; a.pz[i] = d.position.z;
            a.fx[i] = d.direction.x; a.fy[i] = d.direction.y; a.fz[i] = d.direction.z;
            a.speed[i] = d.speed;
            a.ax[i] = a.ay[i] = a.az[i] = 0.f;
            a.sx[i] = d.side.x; a.sy[i] = d.side.y; a.sz[i] = d.side.z;
            a.ux[i] = d.up.x; a.uy[i] = d.up.y; a.uz[i] = d.up.z;
            a.gx[i] = d.steering.x; a.gy[i] = d.steering.y; a.gz[i] = d.steering.z;
            a.radius[i] = d.radius; a.max_speed[i] = d.maxSpeed; a.max_force[i] = d.maxForce; a.mass[i] = d.mass;
            a.bits[i] = d.appliedKernels;
        }
        gpusteer::check(gs_group_upload_agents(h_, (uint32_t)n, &a), "gs_group_upload_agents");
        dirty_ = false;
        host_valid_ = true;
    }
    void SyncHost() {
        if (host_valid_ || dirty_) return;
        const size_t n = agents_.size();
        std::vector<float> f(23 * (n ? n : 1));
        std::vector<uint32_t> bits(n ? n : 1);
        gs_agents a = soa(f.data(), bits.data(), n);
        gpusteer::check(gs_group_download_agents(h_, &a), "gs_group_download_agents");
        for (size_t i = 0; i < n; ++i) {
            AgentData& d = agents_[i];
            d.position = make_float4(a.px[i], a.py[i], a.pz[i], 0.f);
            d.direction = make_float4(a.fx[i], a.fy[i], a.fz[i], 0.f);
            d.speed = a.speed[i];
            d.side = make_float3(a.sx[i], a.sy[i], a.sz[i]);
            d.up = make_float3(a.ux[i], a.uy[i], a.uz[i]);
            d.steering = make_float4(a.gx[i], a.gy[i], a.gz[i], 0.f);
            d.radius = a.radius[i]; d.maxSpeed = a.max_speed[i]; d.maxForce = a.max_force[i]; d.mass = a.mass[i];
            d.appliedKernels = a.bits[i];
        }
        host_valid_ = true;
    }

private:
    static gs_agents soa(float* f, uint32_t* bits, size_t n) {
        gs_agents a;
        float** p[23] = {&a.px, &a.py, &a.pz, &a.fx, &a.fy, &a.fz, &a.speed, &a.ax, &a.ay, &a.az, &a.sx, &a.sy,
                         &a.sz, &a.ux, &a.uy, &a.uz, &a.gx, &a.gy, &a.gz, &a.radius, &a.max_speed, &a.max_force,
                         &a.mass};
        const size_t stride = n ? n : 1;
        for (int j = 0; j < 23; ++j) *p[j] = f + j * stride;
        a.bits = bits;
        return a;
    }

    std::vector<AgentData> agents_;
    std::map<id_type, size_t> index_;
    bool dirty_ = false;
    bool host_valid_ = true;
    AgentGroupData data_;
};

inline size_t AgentGroupData::size() const { return g_->agents_.size(); }
inline void AgentGroupData::getAgentData(size_t const index, AgentData& ad) {
    g_->SyncHost();
    if (index < g_->agents_.size()) ad = g_->agents_[index];
}
inline void AgentGroupData::syncHost() { g_->SyncHost(); }
inline void AgentGroupData::syncDevice() { g_->SyncDevice(); }
// Host edits made through getAgentData's copies are not seen; AddAgent /
// RemoveAgent mark the group for upload themselves.  Kept for plugins that
// call it after editing their own copies.
inline void AgentGroupData::setSyncDevice() { g_->dirty_ = true; }
inline uint* AgentGroupData::pdAppliedKernels() { return g_->device_arrays().bits; }
inline uint* AgentGroupData::pdID() { return g_->pdID(); }
inline gpusteer::DeviceVec3 AgentGroupData::pdPosition() { return g_->vec3(0); }
inline gpusteer::DeviceVec3 AgentGroupData::pdDirection() { return g_->vec3(3); }
inline gpusteer::DeviceVec3 AgentGroupData::pdSide() { return g_->vec3(10); }
inline gpusteer::DeviceVec3 AgentGroupData::pdUp() { return g_->vec3(13); }
inline gpusteer::DeviceVec3 AgentGroupData::pdSteering() { return g_->vec3(16); }
inline float* AgentGroupData::pdSpeed() { return g_->field(6); }
inline float* AgentGroupData::pdMaxSpeed() { return g_->field(20); }
inline float* AgentGroupData::pdMaxForce() { return g_->field(21); }
inline float* AgentGroupData::pdRadius() { return g_->field(19); }
inline float* AgentGroupData::pdMass() { return g_->field(22); }

class ObstacleGroup : public BaseGroup {
public:
    ObstacleGroup(uint3 const& worldCells, uint const kno, gpusteer::Context& ctx = gpusteer::Context::Default())
        : BaseGroup(GS_KIND_OBSTACLES, worldCells, kno, ctx) {}
    void Clear() {
        obstacles_.clear();
        n_ = 0;
        dirty_ = true;
    }
    void AddObstacle(ObstacleData const& od) {
        obstacles_.push_back(od);
        n_ = (uint)obstacles_.size();
        dirty_ = true;
    }
    void RemoveObstacle(uint const index) {
        if (index >= obstacles_.size()) return;
        obstacles_.erase(obstacles_.begin() + index);
        n_ = (uint)obstacles_.size();
        dirty_ = true;
    }
    bool GetDataForObstacle(uint const index, ObstacleData& od) const {
        if (index >= obstacles_.size()) return false;
        od = obstacles_[index];
        return true;
    }
    void SyncDevice() override {
        if (!dirty_) return;
        const size_t n = obstacles_.size();
        std::vector<float> x(n + 1), y(n + 1), z(n + 1), r(n + 1);
        for (size_t i = 0; i < n; ++i) {
            x[i] = obstacles_[i].position.x;
            y[i] = obstacles_[i].position.y;
            z[i] = obstacles_[i].position.z;
            r[i] = obstacles_[i].radius;
        }
        gpusteer::check(gs_group_upload_obstacles(h_, (uint32_t)n, x.data(), y.data(), z.data(), r.data()),
                        "gs_group_upload_obstacles");
        dirty_ = false;
    }

private:
    std::vector<ObstacleData> obstacles_;
    bool dirty_ = false;
};

class WallGroup;

// WallGroupData (include/OpenSteer/CUDA/WallGroupData.cuh:12-58): the host
// vectors of the walls (start, midpoint, end, normal; w = 0) in wall order,
// and their device arrays (storage order, see BaseGroup::pdPosition).
class WallGroupData {
public:
    explicit WallGroupData(WallGroup& w) : w_(&w) {}
    std::vector<float4> const& hvLineStart() const { return start_; }
    std::vector<float4> const& hvLineMid() const { return mid_; }
    std::vector<float4> const& hvLineEnd() const { return end_; }
    std::vector<float4> const& hvLineNormal() const { return normal_; }
    uint const& size() const { return n_; }
    inline gpusteer::DeviceVec3 pdLineStart();
    inline gpusteer::DeviceVec3 pdLineMid();
    inline gpusteer::DeviceVec3 pdLineEnd();
    inline gpusteer::DeviceVec3 pdLineNormal();

private:
    friend class WallGroup;
    WallGroup* w_;
    std::vector<float4> start_, mid_, end_, normal_;
    uint n_ = 0;
};

// Wall segments (line start, end and normal; include/OpenSteer/CUDA/WallGroupData.cuh).
class WallGroup : public BaseGroup {
    friend class WallGroupData;

public:
    WallGroup(uint3 const& worldCells, uint const knw, gpusteer::Context& ctx = gpusteer::Context::Default())
        : BaseGroup(GS_KIND_WALLS, worldCells, knw, ctx), data_(*this) {
        gpusteer::check(gs_wall_set_create(0, nullptr, &set_), "gs_wall_set_create");
    }
    // WallGroup::GetWallGroupData (include/OpenSteer/WallGroup.h:25).
    WallGroupData& GetWallGroupData() { return data_; }
    ~WallGroup() { gs_wall_set_destroy(set_); }
    void Clear() { replace(nullptr, 0); }
    void AddWall(float3 const& start, float3 const& end, float3 const& normal) {
        // Appends to the current walls (their midpoints are kept).
        std::vector<float> v[12];
        get(v);
        const float w[9] = {start.x, start.y, start.z, end.x, end.y, end.z, normal.x, normal.y, normal.z};
        for (int j = 0; j < 9; ++j) v[j].push_back(w[j]);
        for (int a = 0; a < 3; ++a) v[9 + a].push_back(0.5f * (w[a] + w[3 + a]));
        rebuild(v);
    }
    // WallGroup::LoadFromFile (WallGroupData.cu:73-103): replaces the walls.
    bool LoadFromFile(char const* szFilename) {
        gs_wall_set w = nullptr;
        if (gs_wall_set_load_map(szFilename, &w) != GS_OK) return false;
        gs_wall_set_destroy(set_);
        set_ = w;
        refresh();
        return true;
    }
    // WallGroup::SplitWalls (WallGroupData.cu:105-206) over the given cells,
    // in their order (normally KNNBinData::hvCells()).
    void SplitWalls(std::vector<bin_cell> const& cells) {
        std::vector<float> lo(3 * cells.size()), hi(3 * cells.size());
        for (size_t i = 0; i < cells.size(); ++i) {
            lo[3 * i] = cells[i].minBound.x; lo[3 * i + 1] = cells[i].minBound.y; lo[3 * i + 2] = cells[i].minBound.z;
            hi[3 * i] = cells[i].maxBound.x; hi[3 * i + 1] = cells[i].maxBound.y; hi[3 * i + 2] = cells[i].maxBound.z;
        }
        gpusteer::check(gs_wall_set_split_cells(set_, (uint32_t)cells.size(), lo.data(), hi.data()),
                        "gs_wall_set_split_cells");
        refresh();
    }
    void SyncDevice() override {
        if (!dirty_) return;
        gpusteer::check(gs_group_upload_wall_set(h_, set_), "gs_group_upload_wall_set");
        dirty_ = false;
    }
    // Host copies (start, end, normal, mid per wall), as WallGroupData's host vectors.
    void GetWall(uint i, float3& start, float3& end, float3& normal, float3& mid) const {
        std::vector<float> v[12];
        get(v);
        start = make_float3(v[0][i], v[1][i], v[2][i]);
        end = make_float3(v[3][i], v[4][i], v[5][i]);
        normal = make_float3(v[6][i], v[7][i], v[8][i]);
        mid = make_float3(v[9][i], v[10][i], v[11][i]);
    }

private:
    gs_wall_set set_ = nullptr;
    bool dirty_ = false;

    void get(std::vector<float>* v) const {
        uint32_t n = 0;
        gpusteer::check(gs_wall_set_size(set_, &n), "gs_wall_set_size");
        for (int j = 0; j < 12; ++j) v[j].resize(n);
        gs_walls w{v[0].data(), v[1].data(), v[2].data(), v[3].data(), v[4].data(),
                   v[5].data(), v[6].data(), v[7].data(), v[8].data()};
        gpusteer::check(gs_wall_set_get(set_, &w, v[9].data(), v[10].data(), v[11].data()), "gs_wall_set_get");
    }
    void rebuild(std::vector<float>* v) {
        // A set from the arrays, then the midpoints restored (split walls keep theirs).
        const uint32_t n = (uint32_t)v[0].size();
        std::vector<float> pad[12];
        for (int j = 0; j < 12; ++j) pad[j] = v[j], pad[j].push_back(0.f);
        gs_walls w{pad[0].data(), pad[1].data(), pad[2].data(), pad[3].data(), pad[4].data(),
                   pad[5].data(), pad[6].data(), pad[7].data(), pad[8].data()};
        replace(&w, n, pad[9].data(), pad[10].data(), pad[11].data());
    }
    void replace(const gs_walls* w, uint32_t n, const float* mx = nullptr, const float* my = nullptr,
                 const float* mz = nullptr) {
        gs_wall_set s = nullptr;
        gpusteer::check(gs_wall_set_create(n, w, &s), "gs_wall_set_create");
        if (mx) gpusteer::check(gs_wall_set_set_mid(s, mx, my, mz), "gs_wall_set_set_mid");
        gs_wall_set_destroy(set_);
        set_ = s;
        refresh();
    }
    void refresh() {
        std::vector<float> v[12];
        get(v);
        const uint32_t n = (uint32_t)v[0].size();
        n_ = n;
        dirty_ = true;
        data_.n_ = n;
        std::vector<float4>* h[4] = {&data_.start_, &data_.end_, &data_.normal_, &data_.mid_};
        for (int q = 0; q < 4; ++q) {
            h[q]->resize(n);
            for (uint32_t i = 0; i < n; ++i) (*h[q])[i] = make_float4(v[3 * q][i], v[3 * q + 1][i], v[3 * q + 2][i], 0.f);
        }
    }
    WallGroupData data_;
};

// Wall fields in the group's device arrays: midpoint 0..2, start 3..5,
// end 6..8, normal 9..11 (gs_group_device_agents).
inline gpusteer::DeviceVec3 WallGroupData::pdLineStart() { return w_->vec3(3); }
inline gpusteer::DeviceVec3 WallGroupData::pdLineMid() { return w_->vec3(0); }
inline gpusteer::DeviceVec3 WallGroupData::pdLineEnd() { return w_->vec3(6); }
inline gpusteer::DeviceVec3 WallGroupData::pdLineNormal() { return w_->vec3(9); }

class KNNData {
public:
    KNNData(uint const nSize, uint const nK, gpusteer::Context& ctx = gpusteer::Context::Default()) : n_(nSize), k_(nK) {
        gpusteer::check(gs_knn_create(ctx.handle(), nSize, nK, &h_), "gs_knn_create");
    }
    ~KNNData() { gs_knn_destroy(h_); }
    KNNData(const KNNData&) = delete;
    KNNData& operator=(const KNNData&) = delete;
    uint k() const { return k_; }
    uint size() const { return n_; }
    void setSyncHost() { host_valid_ = false; }
    std::vector<uint> const& hvKNNIndices() { syncHost(); return idx_; }
    std::vector<float> const& hvKNNDistances() { syncHost(); return dist_; }
    // Neighbours of agent `index` (slots 0..k-1; CUDA/KNNData.cu:65-69).
    void getAgentData(uint const index, uint* pKNNIndices, float* pKNNDistances) {
        syncHost();
        for (uint j = 0; j < k_; ++j) {
            if (pKNNIndices) pKNNIndices[j] = idx_[(size_t)j * n_ + index];
            if (pKNNDistances) pKNNDistances[j] = dist_[(size_t)j * n_ + index];
        }
    }
    gs_knn handle() const { return h_; }

private:
    void syncHost() {
        if (host_valid_) return;
        idx_.resize((size_t)n_ * k_);
        dist_.resize((size_t)n_ * k_);
        gpusteer::check(gs_knn_download(h_, idx_.data(), dist_.data()), "gs_knn_download");
        host_valid_ = true;
    }
    gs_knn h_ = nullptr;
    uint n_, k_;
    std::vector<uint> idx_;
    std::vector<float> dist_;
    bool host_valid_ = false;
};

// Uniform lattice centred on the origin (CUDA/KNNBinDataV2.cu:30-42).
class KNNBinData {
public:
    KNNBinData(uint3 const& worldCells, float3 const& worldSize, uint const searchRadius,
               gpusteer::Context& ctx = gpusteer::Context::Default())
        : cells_(worldCells), size_(worldSize), radius_(searchRadius) {
        const uint32_t c[3] = {worldCells.x, worldCells.y, worldCells.z};
        const float s[3] = {worldSize.x, worldSize.y, worldSize.z};
        gpusteer::check(gs_bins_create_centered(ctx.handle(), c, s, &h_), "gs_bins_create_centered");
    }
    ~KNNBinData() { gs_bins_destroy(h_); }
    KNNBinData(const KNNBinData&) = delete;
    KNNBinData& operator=(const KNNBinData&) = delete;
    uint radius() const { return radius_; }
    uint3 const& WorldCells() const { return cells_; }
    float3 const& WorldSize() const { return size_; }
    uint getNumCells() const { return cells_.x * cells_.y * cells_.z; }
    // KNNBinDataV2::CreateCells (CUDA/KNNBinDataV2.cu:73-136): cells of the
    // lattice over a box centred at the origin, y outer, z, x inner.
    std::vector<bin_cell> hvCells() const {
#if defined(__clang__)
#pragma clang fp contract(off)  // each bound rounds as two operations, as the library's lattice
#endif
        const float3 step = make_float3(size_.x / cells_.x, size_.y / cells_.y, size_.z / cells_.z);
        std::vector<bin_cell> out;
        out.reserve((size_t)cells_.x * cells_.y * cells_.z);
        for (uint iy = 0; iy < cells_.y; ++iy)
            for (uint iz = 0; iz < cells_.z; ++iz)
                for (uint ix = 0; ix < cells_.x; ++ix) {
                    bin_cell c;
                    c.index = ix + iz * cells_.x + iy * cells_.z * cells_.x;
                    c.minBound = make_float3((float)ix * step.x - 0.5f * size_.x, (float)iy * step.y - 0.5f * size_.y,
                                             (float)iz * step.z - 0.5f * size_.z);
                    c.position = make_float3(c.minBound.x + 0.5f * step.x, c.minBound.y + 0.5f * step.y,
                                             c.minBound.z + 0.5f * step.z);
                    c.maxBound = make_float3(c.minBound.x + step.x, c.minBound.y + step.y, c.minBound.z + step.z);
                    out.push_back(c);
                }
        return out;
    }
    bool is3D() const { return cells_.y > 1; }
    gs_bins handle() const { return h_; }

private:
    gs_bins h_ = nullptr;
    uint3 cells_;
    float3 size_;
    uint radius_;
};

// ---- GroupSteerLibrary.cuh free functions -----------------------------------

inline void wrapWorld(AgentGroup* pAgentGroup, float3 const& worldSize) {
    const float c[3] = {0.f, 0.f, 0.f}, s[3] = {worldSize.x, worldSize.y, worldSize.z};
    gpusteer::check(gs_wrap_world(pAgentGroup->device_handle(), GS_WRAP_BOX, c, s), "gs_wrap_world");
    pAgentGroup->SetSyncHost();
}
inline void steerForSeek(AgentGroup* pAgentGroup, float3 const& target, float const fWeight, uint const doNotApplyWith) {
    const float t[3] = {target.x, target.y, target.z};
    gpusteer::check(gs_steer_seek(pAgentGroup->device_handle(), t, fWeight, doNotApplyWith,
                                  pAgentGroup->context().Mode()),
                    "gs_steer_seek");
    pAgentGroup->SetSyncHost();
}
inline void steerToAvoidObstacles(AgentGroup* pAgentGroup, ObstacleGroup* pObstacleGroup, KNNData* pKNNData,
                                  float const fMinTimeToCollision, float const fWeight, uint const doNotApplyWith) {
    gpusteer::check(gs_steer_avoid_obstacles(pAgentGroup->device_handle(), pObstacleGroup->device_handle(),
                                             pKNNData->handle(), fMinTimeToCollision, fWeight, doNotApplyWith),
                    "gs_steer_avoid_obstacles");
    pAgentGroup->SetSyncHost();
}
inline void steerToAvoidWalls(AgentGroup* pAgentGroup, KNNData* pKNNData, WallGroup* pWallGroup,
                              float const fMinTimeToCollision, float const fWeight, uint const doNotApplyWith) {
    gpusteer::check(gs_steer_avoid_walls(pAgentGroup->device_handle(), pKNNData->handle(),
                                         pWallGroup->device_handle(), fMinTimeToCollision, fWeight, doNotApplyWith),
                    "gs_steer_avoid_walls");
    pAgentGroup->SetSyncHost();
}
inline void updateKNNDatabase(BaseGroup* pGroup, KNNBinData* pKNNBinData) {
    gpusteer::check(gs_update_knn_database(pGroup->device_handle(), pKNNBinData->handle()), "gs_update_knn_database");
    pGroup->SetSyncHost();
}
inline void findKNearestNeighbors(AgentGroup* pAgentGroup, KNNData* pKNNData, KNNBinData* pKNNBinData,
                                  BaseGroup* pOtherGroup, uint const searchRadius) {
    gpusteer::check(gs_find_knn(pAgentGroup->device_handle(), pKNNData->handle(), pKNNBinData->handle(),
                                pOtherGroup->device_handle(), searchRadius, GS_KNN_DEFAULT),
                    "gs_find_knn");
    pKNNData->setSyncHost();
}
inline void steerForSeparation(AgentGroup* pAgentGroup, KNNData* pKNNData, AgentGroup* pOtherGroup,
                               float const minDistance, float const maxDistance, float const cosMaxAngle,
                               float const fWeight, uint const doNotApplyWith) {
    gpusteer::check(gs_steer_separation(pAgentGroup->device_handle(), pKNNData->handle(),
                                        pOtherGroup->device_handle(), minDistance, maxDistance, cosMaxAngle, fWeight,
                                        doNotApplyWith, pAgentGroup->context().Mode()),
                    "gs_steer_separation");
    pAgentGroup->SetSyncHost();
}
inline void steerForAlignment(AgentGroup* pAgentGroup, KNNData* pKNNData, AgentGroup* pOtherGroup,
                              float const minDistance, float const maxDistance, float const cosMaxAngle,
                              float const fWeight, uint const doNotApplyWith) {
    gpusteer::check(gs_steer_alignment(pAgentGroup->device_handle(), pKNNData->handle(),
                                       pOtherGroup->device_handle(), minDistance, maxDistance, cosMaxAngle, fWeight,
                                       doNotApplyWith, pAgentGroup->context().Mode()),
                    "gs_steer_alignment");
    pAgentGroup->SetSyncHost();
}
inline void steerForCohesion(AgentGroup* pAgentGroup, KNNData* pKNNData, AgentGroup* pOtherGroup,
                             float const minDistance, float const maxDistance, float const cosMaxAngle,
                             float const fWeight, uint const doNotApplyWith) {
    gpusteer::check(gs_steer_cohesion(pAgentGroup->device_handle(), pKNNData->handle(), pOtherGroup->device_handle(),
                                      minDistance, maxDistance, cosMaxAngle, fWeight, doNotApplyWith,
                                      pAgentGroup->context().Mode()),
                    "gs_steer_cohesion");
    pAgentGroup->SetSyncHost();
}
inline void updateGroup(AgentGroup* pAgentGroup, const float elapsedTime) {
    gpusteer::check(gs_update_group(pAgentGroup->device_handle(), elapsedTime, pAgentGroup->context().Mode()),
                    "gs_update_group");
    pAgentGroup->SetSyncHost();
}

// PolylinePathwayCUDA (CUDA/PolylinePathwayCUDA.cuh:11-68): the path's
// derived arrays live in the library (host copies and device buffer).
class PolylinePathwayCUDA {
public:
    PolylinePathwayCUDA(std::vector<float3> const& points, float const radius, bool const cyclic,
                        gpusteer::Context& ctx = gpusteer::Context::Default())
        : ctx_(&ctx) {
        Initialize(points, radius, cyclic);
    }
    ~PolylinePathwayCUDA() { gs_path_destroy(h_); }
    PolylinePathwayCUDA(const PolylinePathwayCUDA&) = delete;
    PolylinePathwayCUDA& operator=(const PolylinePathwayCUDA&) = delete;
    void Initialize(std::vector<float3> const& points, float const radius, bool const cyclic) {
        std::vector<float> xyz(3 * points.size());
        for (size_t i = 0; i < points.size(); ++i) {
            xyz[3 * i] = points[i].x;
            xyz[3 * i + 1] = points[i].y;
            xyz[3 * i + 2] = points[i].z;
        }
        gs_path p = nullptr;
        gpusteer::check(gs_path_create(ctx_->handle(), (uint32_t)points.size(), xyz.data(), radius, cyclic ? 1 : 0, &p),
                        "gs_path_create");
        if (h_) gs_path_destroy(h_);
        h_ = p;
        radius_ = radius;
        cyclic_ = cyclic;
        gpusteer::check(gs_path_info(h_, &count_, &total_), "gs_path_info");
        std::vector<float> pt(3 * count_), nm(3 * count_);
        lengths_.assign(count_, 0.f);
        gpusteer::check(gs_path_get(h_, pt.data(), nm.data(), lengths_.data()), "gs_path_get");
        points_.resize(count_);
        normals_.resize(count_);
        for (uint i = 0; i < count_; ++i) {
            points_[i] = make_float3(pt[3 * i], pt[3 * i + 1], pt[3 * i + 2]);
            normals_[i] = make_float3(nm[3 * i], nm[3 * i + 1], nm[3 * i + 2]);
        }
    }
    std::vector<float3> const& hvPoints() const { return points_; }
    std::vector<float3> const& hvNormals() const { return normals_; }
    std::vector<float> const& hvLengths() const { return lengths_; }
    uint const& numPoints() const { return count_; }
    float const& radius() const { return radius_; }
    bool const& cyclic() const { return cyclic_; }
    float const& totalPathLength() const { return total_; }
    gs_path handle() const { return h_; }

private:
    gpusteer::Context* ctx_;
    gs_path h_ = nullptr;
    uint count_ = 0;
    float total_ = 0.f, radius_ = 0.f;
    bool cyclic_ = false;
    std::vector<float3> points_, normals_;
    std::vector<float> lengths_;
};

inline void steerToFollowPath(AgentGroup* pAgentGroup, PolylinePathwayCUDA* pPath, float const predictionTime,
                              float const fWeight, uint const doNotApplyWith) {
    gpusteer::check(gs_steer_follow_path(pAgentGroup->device_handle(), pPath->handle(), predictionTime, fWeight,
                                         doNotApplyWith),
                    "gs_steer_follow_path");
    pAgentGroup->SetSyncHost();
}
inline void steerToAvoidNeighbors(AgentGroup* pAgentGroup, KNNData* pKNNData, AgentGroup* pOtherGroup,
                                  const float fMinTimeToCollision, float const fMinSeparationDistance,
                                  bool const bAvoidCloseNeighbors, float const fWeight, uint const doNotApplyWith) {
    gpusteer::check(gs_steer_avoid_neighbors(pAgentGroup->device_handle(), pKNNData->handle(),
                                             pOtherGroup->device_handle(), fMinTimeToCollision, fMinSeparationDistance,
                                             bAvoidCloseNeighbors ? 1 : 0, fWeight, doNotApplyWith,
                                             pAgentGroup->context().Mode()),
                    "gs_steer_avoid_neighbors");
    pAgentGroup->SetSyncHost();
}
inline void updateGroupWithAntiPenetration(AgentGroup* pAgentGroup, KNNData* pKNNData, WallGroup* pWallGroup,
                                           const float elapsedTime) {
    gpusteer::check(gs_update_group_anti_penetration(pAgentGroup->device_handle(), pKNNData->handle(),
                                                     pWallGroup->device_handle(), elapsedTime,
                                                     pAgentGroup->context().ClosestPoint()),
                    "gs_update_group_anti_penetration");
    pAgentGroup->SetSyncHost();
}
inline void antiPenetrationAgents(AgentGroup* pAgentGroup, KNNData* pKNNData, AgentGroup* pOtherGroup,
                                  uint const doNotApplyWith) {
    gpusteer::check(gs_anti_penetration_agents(pAgentGroup->device_handle(), pKNNData->handle(),
                                               pOtherGroup->device_handle(), doNotApplyWith),
                    "gs_anti_penetration_agents");
    pAgentGroup->SetSyncHost();
}

// ---- ProcessData.cu / ExtractData.h -----------------------------------------
// Of the group's last updateKNNDatabase, KNNBinDataV2's cell numbering.
inline void ComputeCellDensity(AgentGroup* pAgentGroup, uint* pAgentsPerCell) {
    gpusteer::check(gs_group_cell_density(pAgentGroup->device_handle(), pAgentsPerCell), "gs_group_cell_density");
}
inline void ComputeAvgCellVelocity(AgentGroup* pAgentGroup, float3* pAvgCellVelocity) {
    static_assert(sizeof(float3) == 3 * sizeof(float), "float3 must be packed");
    gpusteer::check(gs_group_cell_velocity(pAgentGroup->device_handle(), (float*)pAvgCellVelocity),
                    "gs_group_cell_velocity");
}
inline void WriteCellDensity(char const* szFilenamePrefix, AgentGroup* pAgentGroup,
                             std::vector<uint> vecSelectedCells = std::vector<uint>()) {
    gpusteer::check(gs_write_cell_density(pAgentGroup->device_handle(), szFilenamePrefix, vecSelectedCells.data(),
                                          (uint32_t)vecSelectedCells.size(), "Frames/cell_density", 0),
                    "gs_write_cell_density");
}
inline void WriteAvgCellVelocity(char const* szFilenamePrefix, AgentGroup* pAgentGroup,
                                 std::vector<uint> vecSelectedCells = std::vector<uint>()) {
    gpusteer::check(gs_write_cell_velocity(pAgentGroup->device_handle(), szFilenamePrefix, vecSelectedCells.data(),
                                           (uint32_t)vecSelectedCells.size(), "Frames/cell_avg_velocity", 0),
                    "gs_write_cell_velocity");
}

// ---- Simulation.h parameter classes -----------------------------------------

struct GroupParams {
    uint m_nNumAgents = 0;
    float m_fMaxSpeed = 0.f, m_fMaxForce = 0.f;
    float3 m_f3StartPosition = make_float3(0.f, 0.f, 0.f);
    float m_fMinStartRadius = 0.f, m_fMaxStartRadius = 0.f;
    std::vector<float3> m_vecPathPoints;
    float m_fPathRadius = 0.f;
    bool m_bPathIsCyclic = false;
    float3 m_f3BodyColor = make_float3(0.f, 0.f, 0.f);
};
struct WorldParams {
    float3 m_f3Dimensions = make_float3(0.f, 0.f, 0.f);
    uint3 m_u3Cells = make_uint3(0, 0, 0);
    uint m_nObstacleCount = 0;
    std::string m_strMapFilename;
};
class SimulationParams {
public:
    // SimulationParams::load (src/Simulation.cpp:7-250); throws
    // gpusteer::Error on an unknown key, a bad number, a missing file or a
    // missing </simulation> (the reference asserts or does not return).
    void load(char const* szFilename) {
        std::unique_ptr<gs_sim_params> p(new gs_sim_params());
        gpusteer::check(gs_sim_params_load(szFilename, p.get()), "gs_sim_params_load");
        const gs_sim_params& q = *p;
        m_WorldParams.m_f3Dimensions = make_float3(q.world.dimensions[0], q.world.dimensions[1], q.world.dimensions[2]);
        m_WorldParams.m_u3Cells = make_uint3(q.world.cells[0], q.world.cells[1], q.world.cells[2]);
        m_WorldParams.m_nObstacleCount = q.world.obstacle_count;
        m_WorldParams.m_strMapFilename = q.world.map_filename;
        m_vecGroupParams.clear();
        for (uint32_t i = 0; i < q.n_groups; ++i) {
            const gs_group_params& g = q.groups[i];
            GroupParams gp;
            gp.m_nNumAgents = g.num_agents;
            gp.m_fMaxSpeed = g.max_speed;
            gp.m_fMaxForce = g.max_force;
            gp.m_f3StartPosition = make_float3(g.start_position[0], g.start_position[1], g.start_position[2]);
            gp.m_fMinStartRadius = g.min_start_radius;
            gp.m_fMaxStartRadius = g.max_start_radius;
            for (uint32_t j = 0; j < g.path_points; ++j)
                gp.m_vecPathPoints.push_back(make_float3(g.path[j][0], g.path[j][1], g.path[j][2]));
            gp.m_fPathRadius = g.path_radius;
            gp.m_bPathIsCyclic = g.path_cyclic != 0;
            gp.m_f3BodyColor = make_float3(g.body_color[0], g.body_color[1], g.body_color[2]);
            m_vecGroupParams.push_back(gp);
        }
        m_nSeed = q.seed; m_nKNN = q.knn; m_nKNO = q.kno; m_nKNW = q.knw; m_nSearchRadius = q.search_radius;
        m_fMaxPursuitPredictionTime = q.max_pursuit_prediction_time;
        m_fMinTimeToCollision = q.min_time_to_collision;
        m_fMinTimeToObstacle = q.min_time_to_obstacle;
        m_fMinTimeToWall = q.min_time_to_wall;
        m_fPathPredictionTime = q.path_prediction_time;
        m_fMinSeparationDistance = q.min_separation_distance;
        m_fMaxSeparationDistance = q.max_separation_distance;
        m_fMinFlockingDistance = q.min_flocking_distance;
        m_fMaxFlockingDistance = q.max_flocking_distance;
        m_fCosMaxFlockingAngle = q.cos_max_flocking_angle;
        m_bAvoidCloseNeighbors = q.avoid_close_neighbors != 0;
        m_fWeightAlignment = q.weight_alignment; m_fWeightCohesion = q.weight_cohesion;
        m_fWeightSeparation = q.weight_separation; m_fWeightPursuit = q.weight_pursuit;
        m_fWeightSeek = q.weight_seek; m_fWeightFlee = q.weight_flee; m_fWeightEvade = q.weight_evade;
        m_fWeightFollowPath = q.weight_follow_path; m_fWeightAvoidObstacles = q.weight_avoid_obstacles;
        m_fWeightAvoidWalls = q.weight_avoid_walls; m_fWeightAvoidNeighbors = q.weight_avoid_neighbors;
        m_nMaskAlignment = q.mask_alignment; m_nMaskAntiPenetrationAgents = q.mask_anti_penetration_agents;
        m_nMaskCohesion = q.mask_cohesion; m_nMaskSeparation = q.mask_separation;
        m_nMaskPursuit = q.mask_pursuit; m_nMaskSeek = q.mask_seek; m_nMaskFlee = q.mask_flee;
        m_nMaskEvade = q.mask_evade; m_nMaskFollowPath = q.mask_follow_path;
        m_nMaskAvoidObstacles = q.mask_avoid_obstacles; m_nMaskAvoidWalls = q.mask_avoid_walls;
        m_nMaskAvoidNeighbors = q.mask_avoid_neighbors;
    }

    WorldParams m_WorldParams;
    std::vector<GroupParams> m_vecGroupParams;
    uint m_nSeed = 0, m_nKNN = 0, m_nKNO = 0, m_nKNW = 0, m_nSearchRadius = 0;
    float m_fMaxPursuitPredictionTime = 0.f, m_fMinTimeToCollision = 0.f, m_fMinTimeToObstacle = 0.f,
          m_fMinTimeToWall = 0.f, m_fPathPredictionTime = 0.f, m_fMinSeparationDistance = 0.f,
          m_fMaxSeparationDistance = 0.f, m_fMinFlockingDistance = 0.f, m_fMaxFlockingDistance = 0.f,
          m_fCosMaxFlockingAngle = 0.f;
    bool m_bAvoidCloseNeighbors = false;
    float m_fWeightAlignment = 0.f, m_fWeightCohesion = 0.f, m_fWeightSeparation = 0.f, m_fWeightPursuit = 0.f,
          m_fWeightSeek = 0.f, m_fWeightFlee = 0.f, m_fWeightEvade = 0.f, m_fWeightFollowPath = 0.f,
          m_fWeightAvoidObstacles = 0.f, m_fWeightAvoidWalls = 0.f, m_fWeightAvoidNeighbors = 0.f;
    uint m_nMaskAlignment = 0, m_nMaskAntiPenetrationAgents = 0, m_nMaskCohesion = 0, m_nMaskSeparation = 0,
         m_nMaskPursuit = 0, m_nMaskSeek = 0, m_nMaskFlee = 0, m_nMaskEvade = 0, m_nMaskFollowPath = 0,
         m_nMaskAvoidObstacles = 0, m_nMaskAvoidWalls = 0, m_nMaskAvoidNeighbors = 0;
};

}  // namespace OpenSteer

#endif  // GPUSTEER_OPENSTEER_HPP
