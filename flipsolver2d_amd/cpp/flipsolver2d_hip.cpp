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
    auto mix = [](float a, float b, float f) { return (a * (1.0f - f)) + (b * f); };
    i = clampf(i + oi, 0.f, (float)(si - 1));
    j = clampf(j + oj, 0.f, (float)(sj - 1));
    const int ci = (int)std::floor(i), cj = (int)std::floor(j);
    const float fi = i - std::trunc(i), fj = j - std::trunc(j);
    const int nj = fj >= 0.5f ? cj + 1 : cj - 1;
    const int ni = fi >= 0.5f ? ci + 1 : ci - 1;
    const float iF = fi < 0.5f ? 0.5f - fi : fi - 0.5f;
    const float jF = fj < 0.5f ? 0.5f - fj : fj - 0.5f;
    const float v1 = mix(g.getAt(ci, cj), g.getAt(ni, cj), iF);
    const float v2 = mix(g.getAt(ci, nj), g.getAt(ni, nj), iF);
    return mix(v1, v2, jF);
}

// Polygons as the C-ABI takes them (fg_polygons): counts and x, y pairs.
struct PolyPack {
    std::vector<int32_t> nv;
    std::vector<float> v;
    fg_polygons view() const { return fg_polygons{(int32_t)nv.size(), nv.data(), v.data()}; }
};

PolyPack pack(const std::vector<Geometry2d>& geos)
{
    PolyPack p;
    for (const Geometry2d& g : geos) {
        const std::vector<Vec3> vs = g.verts();
        p.nv.push_back((int32_t)vs.size());
        for (const Vec3& q : vs) {
            p.v.push_back(q.x());
            p.v.push_back(q.y());
        }
    }
    return p;
}

}  // namespace

// geometry2d.cpp:54-73 (the polygon distance of Inigo Quilez's formula)
float Geometry2d::signedDistance(float x, float y) const
{
    const size_t N = m_verts.size();
    if (N == 0) return 0.f;
    auto dot = [](float ax, float ay, float bx, float by) { return (ax * bx + ay * by) + 0.f * 0.f; };
    const float dx0 = x - m_verts[0].x(), dy0 = y - m_verts[0].y();
    float d = dot(dx0, dy0, dx0, dy0);
    float s = 1.0f;
    for (size_t i = 0, j = N - 1; i < N; j = i, i++) {
        const float ex = m_verts[j].x() - m_verts[i].x(), ey = m_verts[j].y() - m_verts[i].y();
        const float wx = x - m_verts[i].x(), wy = y - m_verts[i].y();
        const float h = std::clamp(dot(wx, wy, ex, ey) / dot(ex, ey, ex, ey), 0.0f, 1.0f);
        const float bx = wx - ex * h, by = wy - ey * h;
        d = std::min(d, dot(bx, by, bx, by));
        const bool c1 = y >= m_verts[i].y(), c2 = y < m_verts[j].y(), c3 = ex * wy > ey * wx;
        if ((c1 && c2 && c3) || (!c1 && !c2 && !c3)) s *= -1.0f;
    }
    return (float)(s * std::sqrt((double)d));
}

Vec3 StaggeredVelocityGrid::velocityAt(float i, float j) const
{
    return Vec3(lerp_host(m_u, 0.5f, 0.f, i, j), lerp_host(m_v, 0.f, 0.5f, i, j), 0.f);
}

MarkerParticleSystem::MarkerParticleSystem(size_t gridSizeI, size_t gridSizeJ, size_t binSize)
    : m_binSize((ssize_t)binSize), m_grid(gridSizeI, gridSizeJ),
      m_bins(gridSizeI / binSize + (gridSizeI % binSize != 0), gridSizeJ / binSize + (gridSizeJ % binSize != 0))
{
    for (size_t k = 0; k < m_bins.data().size(); k++) m_bins.data()[k].m_binIdx = k;
}

std::array<ssize_t, 9> MarkerParticleSystem::binsForGridCell(ssize_t i, ssize_t j) const
{
    std::array<ssize_t, 9> out;
    out.fill(-1);
    const Index2d c = binIdxForIdx(i, j);
    int k = 0;
    for (ssize_t di = -1; di <= 1; di++)
        for (ssize_t dj = -1; dj <= 1; dj++, k++)
            if (m_bins.inBounds(c.i + di, c.j + dj)) out[k] = m_bins.linearIndex(c.i + di, c.j + dj);
    return out;
}

void MarkerParticleSystem::assign(std::vector<Vec3> pos, std::vector<Vec3> vel, std::vector<uint32_t> ids)
{
    m_pos = std::move(pos);
    m_vel = std::move(vel);
    m_ids = std::move(ids);
    for (ParticleBin& b : m_bins.data()) {
        b.m_positions.clear();
        b.m_velocities.clear();
        b.m_ids.clear();
    }
    for (size_t k = 0; k < m_pos.size(); k++) {
        const ssize_t b = gridToBinIdx((ssize_t)std::floor(m_pos[k].x()), (ssize_t)std::floor(m_pos[k].y()));
        if (b < 0) continue;
        ParticleBin& bin = m_bins.data()[(size_t)b];
        bin.m_positions.push_back(m_pos[k]);
        bin.m_velocities.push_back(m_vel[k]);
        bin.m_ids.push_back(m_ids[k]);
    }
}

double IndexedPressureParameterUnit::multiply(const std::vector<double>& vec, const LinearIndexable2d& idx,
                                              double scale) const
{
    const ssize_t c = (ssize_t)unitIndex;
    auto at = [&](ssize_t k) { return (k >= 0 && k < (ssize_t)vec.size()) ? vec[(size_t)k] : 0.0; };
    return (scale * nonsolidNeighborCount) * at(c) +
           (-scale * (fluidNeighborMask & I_NEG_NEIGHBOR_BIT)) * at(idx.linearIdxOfOffset(c, -1, 0)) +
           (-scale * ((fluidNeighborMask & I_POS_NEIGHBOR_BIT) >> 1)) * at(idx.linearIdxOfOffset(c, 1, 0)) +
           (-scale * ((fluidNeighborMask & J_NEG_NEIGHBOR_BIT) >> 2)) * at(idx.linearIdxOfOffset(c, 0, -1)) +
           (-scale * ((fluidNeighborMask & J_POS_NEIGHBOR_BIT) >> 3)) * at(idx.linearIdxOfOffset(c, 0, 1));
}

void IndexedPressureParameters::multiply(const std::vector<double>& in, std::vector<double>& out) const
{
    out = in;
    for (const IndexedPressureParameterUnit& u : m_data) out[u.unitIndex] = u.multiply(in, m_indexer, m_scale);
}

// getPressureProjectionMatrix (flipsolver2d.cpp:797-887): per fluid cell,
// the neighbours (clamped at the edges) that are fluid or empty count as
// nonsolid; a fluid neighbour is flagged when its linear offset lies inside
// the grid (so a row-end cell couples to the next row's first cell).
IndexedPressureParameters IndexedPressureParameters::fromMaterials(const MaterialGrid& m, double scale)
{
    IndexedPressureParameters out(m.linearSize(), m, scale);
    const ssize_t I = m.sizeI(), J = m.sizeJ(), C = I * J;
    const ssize_t di[4] = {-1, 1, 0, 0}, dj[4] = {0, 0, -1, 1};
    const NeighborMask bit[4] = {I_NEG_NEIGHBOR_BIT, I_POS_NEIGHBOR_BIT, J_NEG_NEIGHBOR_BIT, J_POS_NEIGHBOR_BIT};
    for (ssize_t i = 0; i < I; i++)
        for (ssize_t j = 0; j < J; j++) {
            if (!m.isFluid(i, j)) continue;
            IndexedPressureParameterUnit u;
            const ssize_t c = m.linearIndex(i, j);
            u.unitIndex = (size_t)c;
            for (int d = 0; d < 4; d++) {
                if (m.isFluid(i + di[d], j + dj[d])) {
                    u.nonsolidNeighborCount++;
                    const ssize_t l = m.linearIdxOfOffset(c, di[d], dj[d]);
                    u.setNeighbor(bit[d], l >= 0 && l < C);
                } else if (m.isEmpty(i + di[d], j + dj[d])) {
                    u.nonsolidNeighborCount++;
                }
            }
            out.add(u);
        }
    return out;
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
    if (params_.viscosityEnabled && params_.useHeavyViscosity)
        throw Error(FG_E_INVALID, "the heavy viscosity model is not built (the light one is)");
    if (params_.viscosityEnabled && params_.parameterHandlingMethod != PARTICLE)
        throw Error(FG_E_INVALID, "viscosity is built for particle parameter handling only");
    if (params_.simulationMethod == SIMULATION_SMOKE || params_.simulationMethod == SIMULATION_FIRE)
        throw Error(FG_E_INVALID, "only the liquid solver is built");
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
    sp.gravity_i = params_.globalAcceleration.x();
    sp.gravity_j = params_.globalAcceleration.y();
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
    sp.sim_method = params_.simulationMethod == SIMULATION_NBFLIP ? FG_SIM_NBFLIP : FG_SIM_FLIP;
    sp.viscosity_enabled = params_.viscosityEnabled ? 1 : 0;
    if (sp.sim_method == FG_SIM_NBFLIP) sp.max_particles *= 2;   // band cells fill to ppc, crowded ones hold 2 ppc
    const int rc = fg_scene_create(ctx_, &sp, &scene_);
    if (rc != FG_OK) {
        const std::string msg = fg_last_error();
        if (ownCtx_) fg_ctx_destroy(ctx_);
        throw Error(rc, msg);
    }
    particles_ = MarkerParticleSystem(gridSizeI(), gridSizeJ(), 3);
    material_ = MaterialGrid(gridSizeI(), gridSizeJ(), EMPTY);
    velocity_ = StaggeredVelocityGrid(gridSizeI(), gridSizeJ());
    fluidSdf_ = SdfGrid(gridSizeI(), gridSizeJ());
    solidSdf_ = SdfGrid(gridSizeI(), gridSizeJ());
}

namespace {
// the parameters of an NBFlipSolver (a temporary that lives through the base
// constructor's initialiser)
struct NBParams {
    FlipSolverParameters p;
    const FlipSolverParameters* get() const { return &p; }
};
NBParams nbflip_params(const NBFlipParameters* p)
{
    if (!p) throw Error(FG_E_INVALID, "NBFlipParameters is NULL");
    NBParams r{*p};
    r.p.simulationMethod = SIMULATION_NBFLIP;
    return r;
}
}  // namespace

NBFlipSolver::NBFlipSolver(const NBFlipParameters* p, fg_ctx* sharedCtx, int device)
    : FlipSolver(nbflip_params(p).get(), sharedCtx, device)
{
}

// NBFlipSolver::step (nbflipsolver.cpp:26-64) as one device substep
void NBFlipSolver::step()
{
    fg_stats st{};
    check(fg_scene_substep(deviceScene(), stepDt(), &st));
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

void FlipSolver::addSource(Emitter& emitter)
{
    if (frame_ > 0) throw Error(FG_E_STATE, "sources must be added before the first frame");
    sources_.push_back(emitter);
}

void FlipSolver::addSink(Sink& sink)
{
    if (frame_ > 0) throw Error(FG_E_STATE, "sinks must be added before the first frame");
    sinks_.push_back(sink);
}

void FlipSolver::addInitialFluid(Emitter& emitter)
{
    if (frame_ > 0) throw Error(FG_E_STATE, "initial fluid must be added before the first frame");
    initialFluid_.push_back(emitter);
}

void FlipSolver::firstFrameInit()
{
    std::vector<Geometry2d> solids, sources, sinks, fluids;
    std::vector<float> friction, velocity, source_visc, fluid_visc;
    std::vector<uint8_t> transfer;
    for (const Obstacle& o : obstacles_) {
        solids.push_back(o.geometry());
        friction.push_back(o.friction());
    }
    for (const Emitter& e : sources_) {
        sources.push_back(e.geometry());
        velocity.push_back(e.velocity().x());
        velocity.push_back(e.velocity().y());
        transfer.push_back(e.velocityTransfer() ? 1 : 0);
        source_visc.push_back(e.viscosity());
    }
    for (const Sink& s : sinks_) sinks.push_back(s.geo());
    for (const Emitter& e : initialFluid_) {
        fluids.push_back(e.geometry());
        fluid_visc.push_back(e.viscosity());
    }
    const PolyPack ps = pack(solids), pr = pack(sources), pk = pack(sinks), pf = pack(fluids);
    fg_scene_objects o{};
    o.solids = ps.view();
    o.friction = friction.empty() ? nullptr : friction.data();
    o.sources = pr.view();
    o.source_velocity = velocity.empty() ? nullptr : velocity.data();
    o.source_transfer = transfer.empty() ? nullptr : transfer.data();
    o.sinks = pk.view();
    o.fluids = pf.view();
    o.seed = params_.seed;
    o.source_viscosity = source_visc.empty() ? nullptr : source_visc.data();
    o.fluid_viscosity = fluid_visc.empty() ? nullptr : fluid_visc.data();
    check(fg_scene_init_objects(scene_, &o));
    initAdditionalParameters();
    dirty_ = true;
}

void FlipSolver::runStage(int stage)
{
    check(fg_scene_stage(scene_, stage, stepDt_));
    dirty_ = true;
}

// The stage hooks: step() runs them in FlipSolver::step's order
// (flipsolver2d.cpp:412-462).
void FlipSolver::advect() { runStage(FG_STAGE_ADVECT); }
void FlipSolver::rebinParticles() { runStage(FG_STAGE_REBIN); }
void FlipSolver::densityCorrectionStage() { runStage(FG_STAGE_DENSITY); }
void FlipSolver::particleToGrid() { runStage(FG_STAGE_P2G); }
void FlipSolver::updateSdf() { runStage(FG_STAGE_SDF); }
void FlipSolver::gridUpdate()
{
    particleToGrid();
    updateSdf();
}
void FlipSolver::extrapolateLevelset() { runStage(FG_STAGE_LEVELSET); }
void FlipSolver::extrapolateVelocity() { runStage(FG_STAGE_EXTRAPOLATE); }
void FlipSolver::applyBodyForces() { runStage(FG_STAGE_FORCES); }
void FlipSolver::project()
{
    runStage(FG_STAGE_PROJECT);
    runStage(FG_STAGE_SOLIDS);
}
void FlipSolver::particleUpdate() { runStage(FG_STAGE_G2P); }
void FlipSolver::countParticles() { runStage(FG_STAGE_COUNT); }

void FlipSolver::step()
{
    advect();
    rebinParticles();
    densityCorrectionStage();
    gridUpdate();
    extrapolateLevelset();
    extrapolateVelocity();
    applyBodyForces();
    project();
    particleUpdate();
    countParticles();
}

void FlipSolver::stepFrameOnDevice()
{
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
}

// FlipSolver::stepFrame, flipsolver2d.cpp:464-500
void FlipSolver::stepFrame()
{
    if (frame_ == 0) firstFrameInit();
    if (!stagesOverridden()) {
        stepFrameOnDevice();
    } else {
        // the reference's loop: a CFL dt from the fastest particle, step()
        const float frameDt = 1.f / params_.fps;
        float t = 0.f;
        int count = 0;
        bool finished = false;
        stats_ = SolverStats();
        while (!finished) {
            float vel = 0.f;
            check(fg_scene_max_velocity(scene_, &vel));
            float dt = params_.cflNumber / (vel + 1e-15f);
            if (t + dt >= frameDt || count == params_.maxSubsteps - 1) {
                dt = frameDt - t;
                finished = true;
            } else if (t + 2.f * dt >= frameDt) {
                dt = 0.5f * (frameDt - t);
            }
            stepDt_ = dt;
            step();
            stats_.substeps_++;
            t += dt;
            count++;
            if (count > 50) break;
        }
    }
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
    std::vector<uint8_t> uval((I + 1) * J), vval(I * (J + 1));
    fg_scene_state st{};
    st.px = px.data();
    st.py = py.data();
    st.vx = vx.data();
    st.vy = vy.data();
    st.pid = id.data();
    // the grid as step() holds it after the last stage hook (between hooks
    // of an unfinished step, e.g. the projected grid after project())
    st.work_u = velocity_.velocityGridU().data().data();
    st.work_v = velocity_.velocityGridV().data().data();
    st.u_valid = uval.data();
    st.v_valid = vval.data();
    st.material = material_.data().data();
    st.fluid_sdf = fluidSdf_.data().data();
    st.solid_sdf = solidSdf_.data().data();
    check(fg_scene_get_state(scene_, &st));
    for (size_t k = 0; k < uval.size(); k++) velocity_.uSampleValidityGrid().data()[k] = uval[k] != 0;
    for (size_t k = 0; k < vval.size(); k++) velocity_.vSampleValidityGrid().data()[k] = vval[k] != 0;
    const size_t n = (size_t)st.n_particles;
    std::vector<Vec3> pos(n), vel(n);
    for (size_t k = 0; k < n; k++) {
        pos[k] = Vec3(px[k], py[k], 0.f);
        vel[k] = Vec3(vx[k], vy[k], 0.f);
    }
    particles_.assign(std::move(pos), std::move(vel), std::vector<uint32_t>(id.begin(), id.begin() + (ptrdiff_t)n));
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

// LinearSolver::solve (linearsolver.cpp:25-73) on the GPU: the unit list
// becomes one matrix byte per cell (fg_poisson_solve_matrix); x starts at
// zero as there.
int LinearSolver::solve(const IndexedPressureParameters& m, const IndexedIPPCoefficients&, std::vector<double>& result,
                        const std::vector<double>& vec, int iterLimit, double tol)
{
    const LinearIndexable2d& ix = m.indexer();
    const size_t C = ix.linearSize();
    if (vec.size() != C) throw Error(FG_E_INVALID, "system size mismatch");
    std::vector<uint8_t> code(C, 0);
    for (const IndexedPressureParameterUnit& u : m.data()) {
        if (u.unitIndex >= C) throw Error(FG_E_INVALID, "unit index out of range");
        code[u.unitIndex] = (uint8_t)(0x80 | ((u.nonsolidNeighborCount & 7) << 4) | (u.fluidNeighborMask & 15));
    }
    double rel = tol;
    if (mode_ == MAX_ABS) {
        // ||r||_2 <= tol bounds every |r_k| by tol
        double bb = 0.0;
        for (double b : vec) bb += b * b;
        rel = bb > 0 ? tol / std::sqrt(bb) : 1.0;
        if (rel >= 1.0) {   // x = 0 already meets it (linearsolver.cpp:35-40)
            result.assign(C, 0.0);
            resid_ = 0.0;
            iters_ = 0;
            return 0;
        }
    }
    result.assign(C, 0.0);
    fg_solve_stats st{};
    const int rc = fg_poisson_solve_matrix(ctx_, (int32_t)ix.sizeI(), (int32_t)ix.sizeJ(), code.data(), m.scale(),
                                           vec.data(), result.data(), rel, iterLimit, &st);
    if (rc != FG_OK && rc != FG_E_NUMERIC) throw Error(rc, fg_last_error());
    resid_ = st.rel_residual;
    iters_ = st.iterations;
    if (rc != FG_OK || !st.converged) return iterLimit;
    return std::min(st.iterations, std::max(iterLimit - 1, 0));
}

}  // namespace fghip
