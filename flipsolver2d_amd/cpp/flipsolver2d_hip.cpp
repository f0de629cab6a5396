// flipsolver2d_hip.cpp -- the reference-named C++ API over the C-ABI.
#include "flipsolver2d_hip.h"

#include <algorithm>
#include <cmath>

namespace fghip {

namespace {

// Grid2d::lerp (grid2d.h:186-216) on a host copy: clamp, half-cell
// neighbour choice, linear blends in the reference's order.
float lerp_host(const Grid2d<float>& g, float oi, float oj, float i, float j)
{
    const int si = (int)g.sizeI(), sj = (int)g.sizeJ();
    auto clampf = [](float v, float lo, float hi) { return v < lo ? lo : (hi < v ? hi : v); };
    auto at = [&](int a, int b) {
        a = std::min(std::max(a, 0), si - 1);
        b = std::min(std::max(b, 0), sj - 1);
        return g.at((size_t)a, (size_t)b);
    };
    auto mix = [](float a, float b, float f) { return (a * (1.0f - f)) + (b * f); };
    i = clampf(i + oi, 0.f, (float)(si - 1));
    j = clampf(j + oj, 0.f, (float)(sj - 1));
    const int ci = (int)std::floor(i), cj = (int)std::floor(j);
    const float fi = i - std::trunc(i), fj = j - std::trunc(j);
    const int nj = fj >= 0.5f ? cj + 1 : cj - 1;
    const int ni = fi >= 0.5f ? ci + 1 : ci - 1;
    const float iF = fi < 0.5f ? 0.5f - fi : fi - 0.5f;
    const float jF = fj < 0.5f ? 0.5f - fj : fj - 0.5f;
    const float v1 = mix(at(ci, cj), at(ni, cj), iF);
    const float v2 = mix(at(ci, nj), at(ni, nj), iF);
    return mix(v1, v2, jF);
}

void flatten(const std::vector<Geometry2d>& geos, std::vector<int32_t>& nv, std::vector<float>& v)
{
    nv.clear();
    v.clear();
    for (const Geometry2d& g : geos) {
        nv.push_back((int32_t)g.verts().size());
        for (const Vec3& p : g.verts()) {
            v.push_back(p.x);
            v.push_back(p.y);
        }
    }
}

}  // namespace

Vec3 StaggeredVelocityGrid::velocityAt(float i, float j) const
{
    return Vec3(lerp_host(u_, 0.5f, 0.f, i, j), lerp_host(v_, 0.f, 0.5f, i, j), 0.f);
}

void FlipSolver::check(int rc)
{
    if (rc != FG_OK) throw Error(rc, fg_last_error());
}

FlipSolver::FlipSolver(const FlipSolverParameters* p, fg_ctx* sharedCtx, int device)
{
    if (!p) throw Error(FG_E_INVALID, "FlipSolverParameters is NULL");
    params_ = *p;
    if (params_.gridSizeI <= 0 || params_.gridSizeJ <= 0) throw Error(FG_E_INVALID, "grid size must be > 0");
    if (!(params_.dx > 0.0)) params_.dx = params_.domainSizeI / params_.gridSizeI;
    if (params_.viscosityEnabled) throw Error(FG_E_INVALID, "viscosity is not part of this solver");
    if (params_.simulationMethod != SIMULATION_LIQUID) throw Error(FG_E_INVALID, "only the liquid solver is built");
    if (sharedCtx) {
        ctx_ = sharedCtx;
    } else {
        check(fg_ctx_create(device, 0, &ctx_));
        ownCtx_ = true;
    }
    fg_scene_params sp{};
    sp.size_i = params_.gridSizeI;
    sp.size_j = params_.gridSizeJ;
    sp.dx = params_.dx;
    sp.fluid_density = params_.fluidDensity;
    sp.gravity_i = params_.globalAcceleration.x;
    sp.gravity_j = params_.globalAcceleration.y;
    sp.pic_ratio = params_.picRatio;
    sp.cfl_number = params_.cflNumber;
    sp.particle_scale = params_.particleScale;
    sp.fps = params_.fps;
    sp.max_substeps = params_.maxSubsteps;
    sp.particles_per_cell = params_.particlesPerCell;
    sp.pcg_iter_limit = 0;   // the GPU solve runs to its residual target
    sp.density_mode = FG_DENSITY_OFF;
    sp.solver_tolerance = 1e-6;
    sp.max_particles = (int64_t)params_.gridSizeI * params_.gridSizeJ * params_.particlesPerCell;
    const int rc = fg_scene_create(ctx_, &sp, &scene_);
    if (rc != FG_OK) {
        const std::string msg = fg_last_error();
        if (ownCtx_) fg_ctx_destroy(ctx_);
        throw Error(rc, msg);
    }
    material_ = MaterialGrid(gridSizeI(), gridSizeJ(), EMPTY);
    velocity_ = StaggeredVelocityGrid(gridSizeI(), gridSizeJ());
    fluidSdf_ = SdfGrid(gridSizeI(), gridSizeJ());
    solidSdf_ = SdfGrid(gridSizeI(), gridSizeJ());
}

FlipSolver::~FlipSolver()
{
    if (scene_) fg_scene_destroy(scene_);
    if (ownCtx_ && ctx_) fg_ctx_destroy(ctx_);
}

void FlipSolver::addGeometry(Obstacle& geometry)
{
    if (frame_ > 0) throw Error(FG_E_STATE, "obstacles must be added before the first frame");
    obstacles_.push_back(geometry);
}

void FlipSolver::addSource(Emitter&)
{
    throw Error(FG_E_INVALID, "emitters (SOURCE cells and reseeding) are not part of this solver");
}

void FlipSolver::addSink(Sink&)
{
    throw Error(FG_E_INVALID, "sinks are not part of this solver");
}

void FlipSolver::addInitialFluid(Emitter& emitter)
{
    if (frame_ > 0) throw Error(FG_E_STATE, "initial fluid must be added before the first frame");
    initialFluid_.push_back(emitter);
}

void FlipSolver::firstFrameInit()
{
    std::vector<Geometry2d> solids, fluids;
    std::vector<float> friction;
    for (const Obstacle& o : obstacles_) {
        solids.push_back(o.geometry());
        friction.push_back(o.friction());
    }
    for (const Emitter& e : initialFluid_) fluids.push_back(e.geometry());
    std::vector<int32_t> snv, fnv;
    std::vector<float> sv, fv;
    flatten(solids, snv, sv);
    flatten(fluids, fnv, fv);
    if (snv.empty()) snv.push_back(0);
    if (fnv.empty()) fnv.push_back(0);
    if (sv.empty()) sv.assign(2, 0.f);
    if (fv.empty()) fv.assign(2, 0.f);
    check(fg_scene_init_geometry(scene_, (int32_t)solids.size(), snv.data(), sv.data(),
                                 friction.empty() ? nullptr : friction.data(), (int32_t)fluids.size(), fnv.data(),
                                 fv.data(), params_.seed));
    initAdditionalParameters();
    dirty_ = true;
}

void FlipSolver::step()
{
    fg_stats st{};
    check(fg_scene_substep(scene_, stepDt_, &st));
    dirty_ = true;
}

void FlipSolver::stepFrame()
{
    if (frame_ == 0) firstFrameInit();
    fg_stats st{};
    check(fg_scene_step_frame(scene_, &st));
    // fg_stage -> SolverStage (flipsolver2d.cpp:412-462 stage boundaries)
    static const int kMap[FG_STAGE_NUM] = {ADVECTION,     PARTICLE_REBIN,  DENSITY,        PARTICLE_TO_GRID,
                                           GRID_UPDATE,   AFTER_TRANSFER,  AFTER_TRANSFER, AFTER_TRANSFER,
                                           PRESSURE,      PARTICLE_UPDATE, PARTICLE_UPDATE, PARTICLE_RESEED};
    stats_ = SolverStats();
    for (int k = 0; k < FG_STAGE_NUM; k++) stats_.times_[kMap[k]] += (float)st.stage_ms[k];
    stats_.substeps_ = st.substeps;
    stats_.pressureIters_ = st.pressure_iters;
    stats_.densityIters_ = st.density_iters;
    stats_.frameTime_ = (float)st.frame_ms;
    stats_.resid_ = st.pressure_rel_resid;
    stepDt_ = st.last_dt;
    totalFrameMs_ += st.frame_ms;
    frame_++;
    dirty_ = true;
}

size_t FlipSolver::particleCount()
{
    int64_t n = 0;
    check(fg_scene_particle_count(scene_, &n));
    return (size_t)n;
}

void FlipSolver::sync() const
{
    if (!dirty_) return;
    const size_t I = gridSizeI(), J = gridSizeJ();
    const size_t cap = I * J * (size_t)params_.particlesPerCell;
    std::vector<float> px(cap), py(cap), vx(cap), vy(cap);
    std::vector<uint32_t> id(cap);
    std::vector<float> fsdf(I * J), ssdf(I * J);
    fg_scene_state st{};
    st.px = px.data();
    st.py = py.data();
    st.vx = vx.data();
    st.vy = vy.data();
    st.pid = id.data();
    st.u = velocity_.velocityGridU().data().data();
    st.v = velocity_.velocityGridV().data().data();
    st.u_valid = velocity_.uSampleValidityGrid().data().data();
    st.v_valid = velocity_.vSampleValidityGrid().data().data();
    st.material = material_.data().data();
    st.fluid_sdf = fluidSdf_.data().data();
    st.solid_sdf = solidSdf_.data().data();
    check(fg_scene_get_state(scene_, &st));
    const size_t n = (size_t)st.n_particles;
    particles_.pos_.resize(n);
    particles_.vel_.resize(n);
    particles_.ids_.assign(id.begin(), id.begin() + (ptrdiff_t)n);
    for (size_t k = 0; k < n; k++) {
        particles_.pos_[k] = Vec3(px[k], py[k], 0.f);
        particles_.vel_[k] = Vec3(vx[k], vy[k], 0.f);
    }
    dirty_ = false;
}

MarkerParticleSystem& FlipSolver::markerParticles()
{
    sync();
    return particles_;
}

const MaterialGrid& FlipSolver::materialGrid() const
{
    sync();
    return material_;
}

const StaggeredVelocityGrid& FlipSolver::fluidVelocityGrid() const
{
    sync();
    return velocity_;
}

const SdfGrid& FlipSolver::fluidSdf() const
{
    sync();
    return fluidSdf_;
}

const SdfGrid& FlipSolver::solidSdf() const
{
    sync();
    return solidSdf_;
}

LinearSolver::LinearSolver(fg_ctx* sharedCtx, int device)
{
    if (sharedCtx) {
        ctx_ = sharedCtx;
    } else {
        const int rc = fg_ctx_create(device, 0, &ctx_);
        if (rc != FG_OK) throw Error(rc, fg_last_error());
        ownCtx_ = true;
    }
}

LinearSolver::~LinearSolver()
{
    if (ownCtx_ && ctx_) fg_ctx_destroy(ctx_);
}

int LinearSolver::solve(const IndexedPressureParameters& m, const IndexedIPPCoefficients&, std::vector<double>& result,
                        const std::vector<double>& vec, int iterLimit, double tol)
{
    const size_t C = m.sizeI * m.sizeJ;
    if (m.materials.size() != C || vec.size() != C) throw Error(FG_E_INVALID, "system size mismatch");
    result.assign(C, 0.0);
    fg_solve_stats st{};
    const int rc = fg_poisson_solve(ctx_, (int32_t)m.sizeI, (int32_t)m.sizeJ, m.materials.data(), m.scale, vec.data(),
                                    result.data(), tol, iterLimit, &st);
    if (rc != FG_OK) throw Error(rc, fg_last_error());
    resid_ = st.rel_residual;
    return st.iterations;
}

}  // namespace fghip
