// flipsolver2d_hip.h -- C++ host API with the reference's class names, over
// the C-ABI of include/flipgpu.h.
//
// A caller of the reference's FlipSolver (FlipSolver2dLib/flipsolver2d.h:186-278)
// keeps its code: construct from FlipSolverParameters, add obstacles,
// sources, sinks and initial fluid, call stepFrame(), read
// markerParticles(), materialGrid(), fluidVelocityGrid(), fluidSdf(),
// solidSdf() and timeStats().  Every substep runs on the GPU; accessors
// download the device state on first use after a step.  The protected stage
// hooks of step() are virtual as in the reference; each default runs its
// stage on the device.  LinearSolver::solve (linearsolver.h:15-20) takes the
// reference's IndexedPressureParameters unit list and solves it on the GPU.
// Errors throw fghip::Error (the C-ABI code and message); the reference's
// own step path has no error returns.
#pragma once

#include <array>
#include <cstddef>
#include <cstdint>
#include <memory>
#include <stdexcept>
#include <string>
#include <sys/types.h>
#include <utility>
#include <vector>

#include "flipgpu.h"

namespace fghip {

class Error : public std::runtime_error {
public:
    Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

// geometry2d.h:8-63
class Vec3 {
public:
    Vec3(float x = 0.0f, float y = 0.0f, float z = 0.0f) : m_x(x), m_y(y), m_z(z) {}
    Vec3(std::pair<float, float>& p) : m_x(p.first), m_y(p.second), m_z(0.f) {}
    float& x() { return m_x; }
    float& y() { return m_y; }
    float& z() { return m_z; }
    const float& x() const { return m_x; }
    const float& y() const { return m_y; }
    const float& z() const { return m_z; }
    friend Vec3 operator-(Vec3 a, Vec3 b) { return Vec3(a.m_x - b.m_x, a.m_y - b.m_y, a.m_z - b.m_z); }
    friend Vec3 operator+(Vec3 a, Vec3 b) { return Vec3(a.m_x + b.m_x, a.m_y + b.m_y, a.m_z + b.m_z); }
    friend Vec3 operator*(Vec3 a, float s) { return Vec3(a.m_x * s, a.m_y * s, a.m_z * s); }
    friend Vec3 operator*(float s, Vec3 a) { return Vec3(s * a.m_x, s * a.m_y, s * a.m_z); }
    friend Vec3 operator/(Vec3 a, float s) { return Vec3(a.m_x / s, a.m_y / s, a.m_z / s); }
    float dot(Vec3 r) const { return m_x * r.m_x + m_y * r.m_y + m_z * r.m_z; }

private:
    float m_x, m_y, m_z;
};

// geometry2d.h:65-85: a closed polygon, vertices in metres (x along I)
class Geometry2d {
public:
    Geometry2d() = default;
    Geometry2d(std::vector<Vec3>& verts) : m_verts(verts) {}
    void addVertex(Vec3 v) { m_verts.push_back(v); }
    std::vector<Vec3> verts() const { return m_verts; }
    int vertextCount() const { return (int)m_verts.size(); }
    // geometry2d.cpp:54-73 on the host
    float signedDistance(float x, float y) const;
    float signedDistance(Vec3 p) const { return signedDistance(p.x(), p.y()); }

private:
    std::vector<Vec3> m_verts;
};

// obstacle.h
class Obstacle {
public:
    Obstacle(float friction, Geometry2d& geo) : m_friction(friction), m_geometry(geo) {}
    float friction() const { return m_friction; }
    void setFriction(float f) { m_friction = f; }
    Geometry2d& geometry() { return m_geometry; }
    const Geometry2d& geometry() const { return m_geometry; }
    void setGeometry(Geometry2d& g) { m_geometry = g; }

private:
    float m_friction;
    Geometry2d m_geometry;
};

// emitter.h: a source or an initial-fluid region.  The liquid solver uses
// its geometry, velocity and velocity transfer; temperature, concentration,
// fuel, viscosity and divergence are kept for callers (the gas solvers and
// viscosity are not part of this solver; divergence control is cleared every
// substep by centeredParamsToGrid, flipsolver2d.cpp:1382).
class Emitter {
public:
    Emitter(Geometry2d& geo) : m_geometry(geo) {}
    void setViscosity(float v) { m_viscosity = v; }
    float viscosity() const { return m_viscosity; }
    Geometry2d& geometry() { return m_geometry; }
    const Geometry2d& geometry() const { return m_geometry; }
    void setGeometry(Geometry2d& g) { m_geometry = g; }
    float temperature() const { return m_temperature; }
    void setTemperature(float t) { m_temperature = t; }
    float concentrartion() const { return m_concentration; }
    void setConcentrartion(float c) { m_concentration = c; }
    float divergence() const { return m_divergence; }
    void setDivergence(float d) { m_divergence = d; }
    float fuel() const { return m_fuel; }
    void setFuel(float f) { m_fuel = f; }
    Vec3 velocity() const { return m_velocity; }
    void setVelocity(Vec3 v) { m_velocity = v; }
    bool velocityTransfer() const { return m_transferVelocity; }
    void setVelocityTransfer(bool t) { m_transferVelocity = t; }

private:
    float m_viscosity = 0.f, m_temperature = 273.f, m_concentration = 1.f, m_divergence = 0.f, m_fuel = 1.f;
    bool m_transferVelocity = false;
    Vec3 m_velocity;
    Geometry2d m_geometry;
};

// sink.h
class Sink {
public:
    Sink(float divergence, Geometry2d& geo) : m_geo(geo), m_divergence(divergence) {}
    float divergence() const { return m_divergence; }
    void setDivergence(float d) { m_divergence = d; }
    Geometry2d& geo() { return m_geo; }
    const Geometry2d& geo() const { return m_geo; }
    void setGeo(const Geometry2d& g) { m_geo = g; }

private:
    Geometry2d m_geo;
    float m_divergence;
};

// flipsolver2d.h:29-56
enum SimulationMethod : char { SIMULATION_LIQUID, SIMULATION_SMOKE, SIMULATION_FIRE, SIMULATION_NBFLIP };
enum ParameterHandlingMethod : char { PARTICLE, HYBRID, GRID };

struct FlipSolverParameters {
    double fluidDensity = 1000.0;
    unsigned int seed = 42;
    double dx = 0.0;
    int particlesPerCell = 4;
    Vec3 globalAcceleration{9.81f, 0.f, 0.f};
    float resolution = 0.f;
    int fps = 30;
    int maxSubsteps = 10;
    float picRatio = 0.03f;
    float cflNumber = 5.f;
    float particleScale = 0.8f;
    int pcgIterLimit = 200;
    float domainSizeI = 1.f;
    float domainSizeJ = 1.f;
    int gridSizeI = 0;
    int gridSizeJ = 0;
    float sceneScale = 1.f;
    bool viscosityEnabled = false;
    bool useHeavyViscosity = false;
    SimulationMethod simulationMethod = SIMULATION_LIQUID;
    ParameterHandlingMethod parameterHandlingMethod = PARTICLE;
};

// flipsolver2d.h:58-72
enum SolverStage {
    ADVECTION = 0,
    DECOMPOSITION,
    DENSITY,
    PARTICLE_REBIN,
    PARTICLE_TO_GRID,
    GRID_UPDATE,
    AFTER_TRANSFER,
    PRESSURE,
    VISCOSITY,
    REPRESSURE,
    PARTICLE_UPDATE,
    PARTICLE_RESEED,
    SOLVER_STAGE_COUNT
};

// index2d.h
struct Index2d {
    ssize_t i = 0, j = 0;
    Index2d() = default;
    Index2d(ssize_t i_, ssize_t j_) : i(i_), j(j_) {}
};

// linearindexable2d.h:8-110: row-major i*sizeJ + j
class LinearIndexable2d {
public:
    LinearIndexable2d(size_t si = 0, size_t sj = 0) : m_sizeI((ssize_t)si), m_sizeJ((ssize_t)sj) {}
    ssize_t sizeI() const { return m_sizeI; }
    ssize_t sizeJ() const { return m_sizeJ; }
    size_t linearSize() const { return (size_t)(m_sizeI * m_sizeJ); }
    bool inBounds(ssize_t i, ssize_t j) const { return i >= 0 && i < m_sizeI && j >= 0 && j < m_sizeJ; }
    bool inBounds(Index2d x) const { return inBounds(x.i, x.j); }
    ssize_t linearIndex(ssize_t i, ssize_t j) const { return inBounds(i, j) ? i * m_sizeJ + j : -1; }
    ssize_t linearIndex(Index2d x) const { return linearIndex(x.i, x.j); }
    Index2d index2d(ssize_t lin) const { return Index2d(lin / m_sizeJ, lin % m_sizeJ); }
    // linearIdxOfOffset (linearindexable2d.h:103-106): plain linear offset
    ssize_t linearIdxOfOffset(ssize_t lin, ssize_t di, ssize_t dj) const { return lin + di * m_sizeJ + dj; }

protected:
    ssize_t m_sizeI, m_sizeJ;
};

// grid2d.h: row-major host grid; getAt clamps out-of-range indices (OOB_EXTEND)
template <class T>
class Grid2d : public LinearIndexable2d {
public:
    Grid2d(size_t si = 0, size_t sj = 0, T v = T()) : LinearIndexable2d(si, sj), m_data(si * sj, v) {}
    typename std::vector<T>::reference at(ssize_t i, ssize_t j) { return m_data[(size_t)(i * m_sizeJ + j)]; }
    typename std::vector<T>::const_reference at(ssize_t i, ssize_t j) const { return m_data[(size_t)(i * m_sizeJ + j)]; }
    typename std::vector<T>::reference at(Index2d x) { return at(x.i, x.j); }
    T getAt(ssize_t i, ssize_t j) const
    {
        i = i < 0 ? 0 : (i >= m_sizeI ? m_sizeI - 1 : i);
        j = j < 0 ? 0 : (j >= m_sizeJ ? m_sizeJ - 1 : j);
        return m_data[(size_t)(i * m_sizeJ + j)];
    }
    void fill(T v) { std::fill(m_data.begin(), m_data.end(), v); }
    std::vector<T>& data() { return m_data; }
    const std::vector<T>& data() const { return m_data; }

private:
    std::vector<T> m_data;
};

// materialgrid.h:6-20
enum FluidMaterial : int8_t { FLUID = 0x40, SOURCE = 0x41, SOLID = 0x20, SINK = 0x12, EMPTY = 0x10 };

class MaterialGrid : public Grid2d<int8_t> {
public:
    using Grid2d<int8_t>::Grid2d;
    static bool isFluid(int8_t m) { return (m & 0x40) != 0; }
    static bool isStrictFluid(int8_t m) { return m == FLUID; }
    static bool isEmpty(int8_t m) { return (m & 0x10) != 0; }
    static bool isSolid(int8_t m) { return (m & 0x20) != 0; }
    static bool isSource(int8_t m) { return m == SOURCE; }
    static bool isSink(int8_t m) { return m == SINK; }
    // the cell tests of materialgrid.cpp:42-127 (indices clamp at the edges)
    bool isFluid(ssize_t i, ssize_t j) const { return isFluid(getAt(i, j)); }
    bool isStrictFluid(ssize_t i, ssize_t j) const { return isStrictFluid(getAt(i, j)); }
    bool isEmpty(ssize_t i, ssize_t j) const { return isEmpty(getAt(i, j)); }
    bool isSolid(ssize_t i, ssize_t j) const { return isSolid(getAt(i, j)); }
    bool isSource(ssize_t i, ssize_t j) const { return isSource(getAt(i, j)); }
    bool isSink(ssize_t i, ssize_t j) const { return isSink(getAt(i, j)); }
    // materialgrid.cpp:137-140
    int nonsolidNeighborCount(ssize_t i, ssize_t j) const
    {
        return !isSolid(i - 1, j) + !isSolid(i + 1, j) + !isSolid(i, j - 1) + !isSolid(i, j + 1);
    }
};

class SdfGrid : public Grid2d<float> {
public:
    using Grid2d<float>::Grid2d;
};

// staggeredvelocitygrid.h:8-74: U is (I+1) x J, V is I x (J+1).
class StaggeredVelocityGrid : public LinearIndexable2d {
public:
    StaggeredVelocityGrid(size_t si = 0, size_t sj = 0)
        : LinearIndexable2d(si, sj), m_u(si + 1, sj), m_v(si, sj + 1), m_uValid(si + 1, sj, false),
          m_vValid(si, sj + 1, false)
    {
    }
    Grid2d<float>& velocityGridU() { return m_u; }
    Grid2d<float>& velocityGridV() { return m_v; }
    const Grid2d<float>& velocityGridU() const { return m_u; }
    const Grid2d<float>& velocityGridV() const { return m_v; }
    Grid2d<bool>& uSampleValidityGrid() { return m_uValid; }
    Grid2d<bool>& vSampleValidityGrid() { return m_vValid; }
    const Grid2d<bool>& uSampleValidityGrid() const { return m_uValid; }
    const Grid2d<bool>& vSampleValidityGrid() const { return m_vValid; }
    float& u(ssize_t i, ssize_t j) { return m_u.at(i, j); }
    float& v(ssize_t i, ssize_t j) { return m_v.at(i, j); }
    float getU(ssize_t i, ssize_t j) const { return m_u.at(i, j); }
    float getV(ssize_t i, ssize_t j) const { return m_v.at(i, j); }
    void setU(ssize_t i, ssize_t j, float x) { m_u.at(i, j) = x; }
    void setV(ssize_t i, ssize_t j, float x) { m_v.at(i, j) = x; }
    bool getUValidity(ssize_t i, ssize_t j) const { return m_uValid.at(i, j); }
    bool getVValidity(ssize_t i, ssize_t j) const { return m_vValid.at(i, j); }
    void setUValidity(ssize_t i, ssize_t j, bool x) { m_uValid.at(i, j) = x; }
    void setVValidity(ssize_t i, ssize_t j, bool x) { m_vValid.at(i, j) = x; }
    // Bilinear sample in cell units (staggeredvelocitygrid.cpp:158-167).
    Vec3 velocityAt(float i, float j) const;
    Vec3 velocityAt(Vec3 p) const { return velocityAt(p.x(), p.y()); }

private:
    Grid2d<float> m_u, m_v;
    Grid2d<bool> m_uValid, m_vValid;
};

// markerparticlesystem.h:18-148: the particles of one 3x3-cell bin
class ParticleBin {
public:
    size_t size() const { return m_positions.size(); }
    std::vector<Vec3>& positions() { return m_positions; }
    std::vector<Vec3>& velocities() { return m_velocities; }
    const std::vector<Vec3>& positions() const { return m_positions; }
    const std::vector<Vec3>& velocities() const { return m_velocities; }
    Vec3& particlePosition(size_t k) { return m_positions[k]; }
    Vec3& particleVelocity(size_t k) { return m_velocities[k]; }
    const std::vector<uint32_t>& ids() const { return m_ids; }
    size_t binIdx() const { return m_binIdx; }

private:
    friend class MarkerParticleSystem;
    std::vector<Vec3> m_positions, m_velocities;
    std::vector<uint32_t> m_ids;
    size_t m_binIdx = 0;
};

// markerparticlesystem.h:187-249: a host view of the device particles in
// the reference's 3x3-cell bins (each bin in (cell, id) order), plus the
// flat arrays in the device's canonical (cell, id) order.
class MarkerParticleSystem {
public:
    MarkerParticleSystem(size_t gridSizeI = 0, size_t gridSizeJ = 0, size_t binSize = 3);
    ParticleBin& binForGridIdx(size_t linIdx) { return m_bins.data()[(size_t)gridToBinIdx((ssize_t)linIdx)]; }
    ParticleBin& binForGridIdx(Index2d x) { return binForGridIdx(x.i, x.j); }
    ParticleBin& binForGridIdx(size_t i, size_t j) { return m_bins.data()[(size_t)gridToBinIdx((ssize_t)i, (ssize_t)j)]; }
    ParticleBin& binForGridPosition(Vec3 pos) { return binForGridIdx((size_t)pos.x(), (size_t)pos.y()); }
    ParticleBin& binForBinIdx(size_t linIdx) { return m_bins.data()[linIdx]; }
    ssize_t gridToBinIdx(Index2d x) const { return gridToBinIdx(x.i, x.j); }
    ssize_t gridToBinIdx(Vec3 pos) const { return gridToBinIdx((ssize_t)pos.x(), (ssize_t)pos.y()); }
    ssize_t gridToBinIdx(ssize_t linIdx) const
    {
        return gridToBinIdx(linIdx / m_grid.sizeJ(), linIdx % m_grid.sizeJ());
    }
    ssize_t gridToBinIdx(ssize_t i, ssize_t j) const { return m_bins.linearIndex(i / m_binSize, j / m_binSize); }
    Grid2d<ParticleBin>& bins() { return m_bins; }
    std::vector<Vec3>& positions(size_t binIdx) { return m_bins.data()[binIdx].positions(); }
    std::vector<Vec3>& velocities(size_t binIdx) { return m_bins.data()[binIdx].velocities(); }
    Index2d binIdxForIdx(Index2d x) const { return binIdxForIdx(x.i, x.j); }
    Index2d binIdxForIdx(ssize_t i, ssize_t j) const { return Index2d(i / m_binSize, j / m_binSize); }
    // markerparticlesystem.cpp:114-128: the 3x3 bins around a cell's bin, -1 outside
    std::array<ssize_t, 9> binsForGridCell(Index2d x) const { return binsForGridCell(x.i, x.j); }
    std::array<ssize_t, 9> binsForGridCell(ssize_t i, ssize_t j) const;
    size_t particleCount() const { return m_pos.size(); }
    // flat, canonical (cell, id) order
    const std::vector<Vec3>& allPositions() const { return m_pos; }
    const std::vector<Vec3>& allVelocities() const { return m_vel; }
    const std::vector<uint32_t>& ids() const { return m_ids; }

private:
    friend class FlipSolver;
    void assign(std::vector<Vec3> pos, std::vector<Vec3> vel, std::vector<uint32_t> ids);
    ssize_t m_binSize;
    LinearIndexable2d m_grid;
    Grid2d<ParticleBin> m_bins;
    std::vector<Vec3> m_pos, m_vel;
    std::vector<uint32_t> m_ids;
};

// flipsolver2d.h:88-181
class SolverStats {
public:
    using StageTimings = std::array<float, SOLVER_STAGE_COUNT>;
    int pressureIterations() const { return pressureIters_; }
    int densityIterations() const { return densityIters_; }
    int viscosityIterations() const { return 0; }
    int substepCount() const { return substeps_; }
    StageTimings timings() const { return times_; }
    float frameTime() const { return frameTime_; }
    double pressureResidual() const { return resid_; }

private:
    friend class FlipSolver;
    StageTimings times_{};
    int substeps_ = 0, pressureIters_ = 0, densityIters_ = 0;
    float frameTime_ = 0.f;
    double resid_ = 0.0;
};

// pressuredata.h:85-151: one fluid row of the pressure matrix
enum NeighborMask : uint8_t {
    I_NEG_NEIGHBOR_BIT = 1 << 0,
    I_POS_NEIGHBOR_BIT = 1 << 1,
    J_NEG_NEIGHBOR_BIT = 1 << 2,
    J_POS_NEIGHBOR_BIT = 1 << 3
};

struct IndexedPressureParameterUnit {
    void setNeighbor(NeighborMask bit, bool flag)
    {
        if (flag) fluidNeighborMask |= bit;
        else fluidNeighborMask &= (uint8_t)~bit;
    }
    // row times vec: s (n x_c - sum of the flagged neighbours at linear offsets)
    double multiply(const std::vector<double>& vec, const LinearIndexable2d& indexer, double scale) const;

    size_t unitIndex = (size_t)-1;
    uint8_t fluidNeighborMask = 0;
    uint8_t nonsolidNeighborCount = 0;
};

// pressuredata.h:153-240: the fluid rows in ascending unit index; every
// other row is the identity
class IndexedPressureParameters {
public:
    IndexedPressureParameters(size_t size, const LinearIndexable2d& indexer, double scale)
        : m_indexer(indexer), m_scale(scale)
    {
        m_data.reserve(size);
    }
    void add(IndexedPressureParameterUnit& unit) { m_data.push_back(unit); }
    void endThreadDataRange() {}   // the device solve has no thread ranges
    const std::vector<IndexedPressureParameterUnit>& data() const { return m_data; }
    const LinearIndexable2d& indexer() const { return m_indexer; }
    double scale() const { return m_scale; }
    // out = A in (FP64, on the host)
    void multiply(const std::vector<double>& in, std::vector<double>& out) const;
    // The assembled matrix of a material grid (getPressureProjectionMatrix,
    // flipsolver2d.cpp:797-887).
    static IndexedPressureParameters fromMaterials(const MaterialGrid& m, double scale);

private:
    LinearIndexable2d m_indexer;
    double m_scale;
    std::vector<IndexedPressureParameterUnit> m_data;
};

// The preconditioner is built on the device from the matrix rows (the
// multigrid V-cycle by default); this type keeps LinearSolver::solve's
// signature (PressureIPPCoeficients.h).
class IndexedIPPCoefficients {
public:
    IndexedIPPCoefficients() = default;
    explicit IndexedIPPCoefficients(const IndexedPressureParameters&) {}
};

class LinearSolver {
public:
    // How tol is read: the reference's max-abs residual bound
    // (linearsolver.cpp:53-59; met here by reaching ||b - A x||2 <= tol,
    // which bounds every component), or the FP64 relative residual
    // ||b - A x||2 / ||b||2.
    enum ToleranceMode { MAX_ABS, RELATIVE_L2 };

    explicit LinearSolver(fg_ctx* sharedCtx = nullptr, int device = 0);
    ~LinearSolver();
    LinearSolver(const LinearSolver&) = delete;
    LinearSolver& operator=(const LinearSolver&) = delete;
    void setToleranceMode(ToleranceMode m) { mode_ = m; }
    ToleranceMode toleranceMode() const { return mode_; }
    // Returns iterLimit when the solve did not meet tol (the reference's
    // failure test in FlipSolver::project, flipsolver2d.cpp:107), else the
    // inner PCG iterations, at most iterLimit - 1.  iterLimit caps the inner
    // iterations of each refinement sweep.
    int solve(const IndexedPressureParameters& matrixIn, const IndexedIPPCoefficients& precond,
              std::vector<double>& result, const std::vector<double>& vec, int iterLimit = 20, double tol = 1e-6);
    double lastResidual() const { return resid_; }
    int lastIterations() const { return iters_; }

private:
    fg_ctx* ctx_ = nullptr;
    bool ownCtx_ = false;
    ToleranceMode mode_ = MAX_ABS;
    double resid_ = 0.0;
    int iters_ = 0;
};

class FlipSolver {
public:
    // One device context per solver unless one is shared (several solvers
    // on one context must not step concurrently).
    explicit FlipSolver(const FlipSolverParameters* p, fg_ctx* sharedCtx = nullptr, int device = 0);
    virtual ~FlipSolver();
    FlipSolver(const FlipSolver&) = delete;
    FlipSolver& operator=(const FlipSolver&) = delete;

    size_t particleCount();
    size_t cellCount() const { return (size_t)params_.gridSizeI * params_.gridSizeJ; }
    int pcgIterationLimit() const { return params_.pcgIterLimit; }
    void stepFrame();
    size_t gridSizeI() const { return (size_t)params_.gridSizeI; }
    size_t gridSizeJ() const { return (size_t)params_.gridSizeJ; }
    void addGeometry(Obstacle& geometry);
    void addSource(Emitter& emitter);
    void addSink(Sink& geometry);
    void addInitialFluid(Emitter& emitter);
    int frameNumber() const { return frame_; }
    std::vector<Obstacle>& geometryObjects() { return obstacles_; }
    std::vector<Emitter>& sourceObjects() { return sources_; }
    std::vector<Sink>& sinkObjects() { return sinks_; }

    MarkerParticleSystem& markerParticles();
    const MaterialGrid& materialGrid() const;
    const StaggeredVelocityGrid& fluidVelocityGrid() const;
    const SdfGrid& fluidSdf() const;
    const SdfGrid& solidSdf() const;
    const SolverStats& timeStats() const { return stats_; }

    float stepDt() const { return stepDt_; }
    double dx() const { return params_.dx; }
    SimulationMethod simulationMethod() const { return params_.simulationMethod; }
    float domainSizeI() const { return params_.domainSizeI; }
    float domainSizeJ() const { return params_.domainSizeJ; }
    double fluidDensity() const { return params_.fluidDensity; }
    int particlesPerCell() const { return params_.particlesPerCell; }
    float picRatio() const { return params_.picRatio; }
    float cflNumber() const { return params_.cflNumber; }
    int maxSubsteps() const { return params_.maxSubsteps; }
    int fps() const { return params_.fps; }
    float frameDt() const { return 1.f / (float)params_.fps; }
    Vec3 globalAcceleration() const { return params_.globalAcceleration; }
    float sceneScale() const { return params_.sceneScale; }
    float lastFrameTime() const { return stats_.frameTime_; }
    float avgFrameTime() const { return frame_ ? (float)(totalFrameMs_ / frame_) : 0.f; }
    virtual void initAdditionalParameters() {}

    // The device scene behind this solver (for batching several solvers
    // with fg_batch_step_frame, or for fg_scene_get_state).
    fg_scene* deviceScene() { return scene_; }
    fg_ctx* deviceContext() { return ctx_; }

protected:
    // firstFrameInit (flipsolver2d.cpp:788-795): sinks, sources, solids,
    // initial fluid and seeding, on the device, on the first stepFrame().
    virtual void firstFrameInit();
    // One FlipSolver::step() with stepDt() (flipsolver2d.cpp:412-462): the
    // hooks below in the reference's order.  The first call of a frame
    // steps the whole frame on the device when no hook is overridden
    // (stepFrameOnDevice()), which keeps its launches back to back.
    virtual void step();
    // The stage hooks of step() (flipsolver2d.h:280-367); each default runs
    // its stage on the device (fg_scene_stage).  An override replaces or
    // wraps that stage; the scene is reachable through deviceScene().
    virtual void advect();
    virtual void rebinParticles();             // matrix snapshot, prune, rebin
    virtual void densityCorrectionStage();
    virtual void gridUpdate();                 // particleToGrid + updateSdf + updateMaterials
    virtual void particleToGrid();
    virtual void updateSdf();                  // updateSdf + updateMaterials + afterTransfer
    virtual void extrapolateLevelset();        // extrapolateLevelsetInside
    virtual void extrapolateVelocity();        // m_fluidVelocityGrid.extrapolate(10)
    virtual void applyBodyForces();            // saved grid + applyBodyForces
    virtual void project();                    // calcPressureRhs + solve + applyPressures (+ solids)
    virtual void particleUpdate();             // extrapolate(10) + G2P
    virtual void countParticles();             // countParticles + reseedParticles
    // true when a subclass overrides a stage hook (then frames step through
    // step() substep by substep)
    virtual bool stagesOverridden() const { return false; }
    void runStage(int stage);
    static void check(int rc);

private:
    void sync() const;
    void stepFrameOnDevice();

    FlipSolverParameters params_;
    fg_ctx* ctx_ = nullptr;
    bool ownCtx_ = false;
    fg_scene* scene_ = nullptr;
    std::vector<Obstacle> obstacles_;
    std::vector<Emitter> sources_;
    std::vector<Sink> sinks_;
    std::vector<Emitter> initialFluid_;
    int frame_ = 0;
    float stepDt_ = 0.f;
    double totalFrameMs_ = 0.0;
    SolverStats stats_;
    // host copies, refreshed lazily after a step
    mutable bool dirty_ = true;
    mutable MarkerParticleSystem particles_;
    mutable MaterialGrid material_;
    mutable StaggeredVelocityGrid velocity_;
    mutable SdfGrid fluidSdf_, solidSdf_;
};

// NBFlipParameters (nbflipsolver.h:8-11): no fields of its own.
struct NBFlipParameters : FlipSolverParameters {
};

// NBFlipSolver (nbflipsolver.h:13-80, nbflipsolver.cpp): the narrow-band
// variant.  Its step(), first frame and band rules run on the device
// (fg_scene_params.sim_method = FG_SIM_NBFLIP); the solve tolerance is the
// constructor's 1e-6 (nbflipsolver.cpp:22).  Its step() is whole: the stage
// hooks follow FlipSolver::step's order and refuse an NB-FLIP scene.
class NBFlipSolver : public FlipSolver {
public:
    explicit NBFlipSolver(const NBFlipParameters* p, fg_ctx* sharedCtx = nullptr, int device = 0);

protected:
    void step() override;
};

}  // namespace fghip
