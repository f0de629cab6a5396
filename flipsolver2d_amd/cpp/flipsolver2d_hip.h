// flipsolver2d_hip.h -- C++ host API with the reference's class names, over
// the C-ABI of include/flipgpu.h.
//
// A caller of the reference's FlipSolver (FlipSolver2dLib/flipsolver2d.h:186-278)
// keeps its code: construct from FlipSolverParameters, add obstacles and
// initial fluid, call stepFrame(), read markerParticles(), materialGrid(),
// fluidVelocityGrid(), fluidSdf(), solidSdf() and timeStats().  Every
// substep runs on the GPU; accessors download the device state on first use
// after a step.  LinearSolver::solve (linearsolver.h:15-20) runs the
// pressure system on the GPU.  Errors throw fghip::Error (the C-ABI code and
// message); the reference's own step path has no error returns.
#pragma once

#include <array>
#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "flipgpu.h"

namespace fghip {

struct Vec3 {
    float x = 0.f, y = 0.f, z = 0.f;
    Vec3() = default;
    Vec3(float x_, float y_, float z_ = 0.f) : x(x_), y(y_), z(z_) {}
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

class Error : public std::runtime_error {
public:
    Error(int code, const std::string& what) : std::runtime_error(what), code_(code) {}
    int code() const { return code_; }

private:
    int code_;
};

// Closed polygon, vertices in metres (x along I).
class Geometry2d {
public:
    Geometry2d() = default;
    explicit Geometry2d(std::vector<Vec3> verts) : verts_(std::move(verts)) {}
    const std::vector<Vec3>& verts() const { return verts_; }

private:
    std::vector<Vec3> verts_;
};

class Obstacle {
public:
    explicit Obstacle(Geometry2d g, float friction = 0.f) : geo_(std::move(g)), friction_(friction) {}
    const Geometry2d& geometry() const { return geo_; }
    float friction() const { return friction_; }

private:
    Geometry2d geo_;
    float friction_;
};

class Emitter {
public:
    explicit Emitter(Geometry2d g, float viscosity = 0.f) : geo_(std::move(g)), viscosity_(viscosity) {}
    const Geometry2d& geometry() const { return geo_; }
    float viscosity() const { return viscosity_; }

private:
    Geometry2d geo_;
    float viscosity_;
};

class Sink {
public:
    explicit Sink(Geometry2d g) : geo_(std::move(g)) {}
    const Geometry2d& geo() const { return geo_; }

private:
    Geometry2d geo_;
};

// Row-major (i*sizeJ + j) host grid.
template <class T>
class Grid2d {
public:
    Grid2d(size_t si = 0, size_t sj = 0, T v = T()) : si_(si), sj_(sj), data_(si * sj, v) {}
    size_t sizeI() const { return si_; }
    size_t sizeJ() const { return sj_; }
    T& at(size_t i, size_t j) { return data_[i * sj_ + j]; }
    const T& at(size_t i, size_t j) const { return data_[i * sj_ + j]; }
    std::vector<T>& data() { return data_; }
    const std::vector<T>& data() const { return data_; }

private:
    size_t si_, sj_;
    std::vector<T> data_;
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
};

class SdfGrid : public Grid2d<float> {
public:
    using Grid2d<float>::Grid2d;
};

// staggeredvelocitygrid.h:8-74: U is (I+1) x J, V is I x (J+1).
class StaggeredVelocityGrid {
public:
    StaggeredVelocityGrid(size_t si = 0, size_t sj = 0)
        : u_(si + 1, sj), v_(si, sj + 1), uValid_(si + 1, sj), vValid_(si, sj + 1), si_(si), sj_(sj) {}
    Grid2d<float>& velocityGridU() { return u_; }
    Grid2d<float>& velocityGridV() { return v_; }
    const Grid2d<float>& velocityGridU() const { return u_; }
    const Grid2d<float>& velocityGridV() const { return v_; }
    Grid2d<uint8_t>& uSampleValidityGrid() { return uValid_; }
    Grid2d<uint8_t>& vSampleValidityGrid() { return vValid_; }
    float getU(size_t i, size_t j) const { return u_.at(i, j); }
    float getV(size_t i, size_t j) const { return v_.at(i, j); }
    bool getUValidity(size_t i, size_t j) const { return uValid_.at(i, j) != 0; }
    bool getVValidity(size_t i, size_t j) const { return vValid_.at(i, j) != 0; }
    // Bilinear sample in cell units (staggeredvelocitygrid.cpp:158-167).
    Vec3 velocityAt(float i, float j) const;
    size_t sizeI() const { return si_; }
    size_t sizeJ() const { return sj_; }

private:
    Grid2d<float> u_, v_;
    Grid2d<uint8_t> uValid_, vValid_;
    size_t si_, sj_;
};

// The particles as flat arrays (cell units; canonical (cell, id) order).
class MarkerParticleSystem {
public:
    size_t particleCount() const { return pos_.size(); }
    std::vector<Vec3>& positions() { return pos_; }
    std::vector<Vec3>& velocities() { return vel_; }
    const std::vector<uint32_t>& ids() const { return ids_; }

private:
    friend class FlipSolver;
    std::vector<Vec3> pos_, vel_;
    std::vector<uint32_t> ids_;
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
    // FlipSolver::firstFrameInit (flipsolver2d.cpp:682-707, 788-795): solids,
    // initial fluid and seeding, on the first stepFrame().
    virtual void firstFrameInit();
    // One FlipSolver::step() with stepDt() (flipsolver2d.cpp:412-462).
    virtual void step();

private:
    void sync() const;
    static void check(int rc);

    FlipSolverParameters params_;
    fg_ctx* ctx_ = nullptr;
    bool ownCtx_ = false;
    fg_scene* scene_ = nullptr;
    std::vector<Obstacle> obstacles_;
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

// The pressure system as the solver assembles it (flipsolver2d.cpp:797-887):
// the 5-point operator s * (n_c x_c - sum of fluid neighbours) on fluid
// cells, identity elsewhere, carried here as the material grid and s.
struct IndexedPressureParameters {
    size_t sizeI = 0, sizeJ = 0;
    std::vector<int8_t> materials;   // row-major I x J
    double scale = 1.0;              // dt / (rho dx^2)
};

// The inverse-Poisson preconditioner is derived on the device from the
// same stencil; this tag keeps LinearSolver::solve's signature.
struct IndexedIPPCoefficients {};

class LinearSolver {
public:
    explicit LinearSolver(fg_ctx* sharedCtx = nullptr, int device = 0);
    ~LinearSolver();
    LinearSolver(const LinearSolver&) = delete;
    LinearSolver& operator=(const LinearSolver&) = delete;
    // Returns the inner PCG iterations.  tol is the FP64 relative residual
    // ||b - Ax||2 / ||b||2 reached (the reference's tol is a max-abs bound).
    int solve(const IndexedPressureParameters& matrixIn, const IndexedIPPCoefficients& precond,
              std::vector<double>& result, const std::vector<double>& vec, int iterLimit = 20, double tol = 1e-6);
    double lastResidual() const { return resid_; }

private:
    fg_ctx* ctx_ = nullptr;
    bool ownCtx_ = false;
    double resid_ = 0.0;
};

}  // namespace fghip
