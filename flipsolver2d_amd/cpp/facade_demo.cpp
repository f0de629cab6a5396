// facade_demo.cpp -- drives the reference-named C++ API the way the
// reference's callers do: AutoBench's runScene (AutoBench/benchmarkrunnerapplication.cpp:88-92)
// builds a scene and calls stepFrame(); FlipSolver::project
// (FlipSolver2dLib/flipsolver2d.cpp:93-126) assembles the pressure matrix and
// calls LinearSolver::solve; the GUI loads scene files with
// JsonSceneReader::loadJson (Utils/jsonscenereader.cpp:8-77).
//
//   facade_demo dam <res> <ppc> <frames> <out.bin>
//       writes: int64 n, f32 px[n], py[n], vx[n], vy[n], u32 id[n],
//               i8 material[I*J], f32 u[(I+1)*J], f32 v[I*(J+1)],
//               i32 substeps (last frame), f32 stepDt, f32 velocityAt(I/2, J/2).x
//   facade_demo hooks <res> <ppc> <frames> <out.bin>
//       the same through a subclass that overrides stage hooks (step() and
//       its hooks run substep by substep); same output
//   facade_demo solve <in.bin> <out.bin> <maxabs 0|1>
//       reads:  i32 I, J, f64 scale, f64 tol, i32 iterLimit, i8 material[I*J], f64 b[I*J]
//       writes: i32 return value, i32 iterations, f64 rel_residual, f64 x[I*J]
//   facade_demo json <scene.json> <frames> <out.bin>   (output as dam)
//   facade_demo jsoninfo <scene.json>    prints the parsed scene, one key per line
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "flipsolver2d_hip.h"
#include "json_scene.h"

using namespace fghip;

static Geometry2d box(double x0, double x1, double y0, double y1)
{
    const float a = (float)x0, b = (float)x1, c = (float)y0, e = (float)y1;
    std::vector<Vec3> v = {Vec3(a, c), Vec3(a, e), Vec3(b, e), Vec3(b, c)};
    return Geometry2d(v);
}

template <class T>
static void put(FILE* f, const T* p, size_t n)
{
    if (n && fwrite(p, sizeof(T), n, f) != n) {
        std::fprintf(stderr, "write failed\n");
        std::exit(3);
    }
}

template <class T>
static void get(FILE* f, T* p, size_t n)
{
    if (n && fread(p, sizeof(T), n, f) != n) {
        std::fprintf(stderr, "read failed\n");
        std::exit(3);
    }
}

// A caller's subclass that wraps a stage hook: frames then run step() and
// its hooks substep by substep, in the reference's structure.
class HookedSolver : public FlipSolver {
public:
    using FlipSolver::FlipSolver;
    int forceCalls = 0;

protected:
    void applyBodyForces() override
    {
        forceCalls++;
        FlipSolver::applyBodyForces();
    }
    bool stagesOverridden() const override { return true; }
};

static void dam_scene(FlipSolver& solver, int res)
{
    const double d = 1.0 / res;
    std::vector<Geometry2d> walls = {box(0.0, d, 0.0, 1.0), box(1.0 - d, 1.0, 0.0, 1.0), box(0.0, 1.0, 0.0, d),
                                     box(0.0, 1.0, 1.0 - d, 1.0)};
    for (Geometry2d& g : walls) {
        Obstacle o(0.f, g);
        solver.addGeometry(o);
    }
    Geometry2d fg = box(0.3 > d ? 0.3 : d, 1.0 - d, d, 0.4);
    Emitter fluid(fg);
    solver.addInitialFluid(fluid);
}

static int write_state(FlipSolver& solver, const char* out)
{
    MarkerParticleSystem& mp = solver.markerParticles();
    const int64_t n = (int64_t)mp.particleCount();
    std::vector<float> px(n), py(n), vx(n), vy(n);
    for (int64_t k = 0; k < n; k++) {
        px[k] = mp.allPositions()[k].x();
        py[k] = mp.allPositions()[k].y();
        vx[k] = mp.allVelocities()[k].x();
        vy[k] = mp.allVelocities()[k].y();
    }
    // the bin view holds every particle once, in its cell's 3x3 bin
    size_t binned = 0;
    for (size_t b = 0; b < mp.bins().linearSize(); b++) binned += mp.binForBinIdx(b).size();
    if (binned != (size_t)n) {
        std::fprintf(stderr, "bins hold %zu of %lld particles\n", binned, (long long)n);
        return 4;
    }
    FILE* f = std::fopen(out, "wb");
    if (!f) return 2;
    put(f, &n, 1);
    put(f, px.data(), n);
    put(f, py.data(), n);
    put(f, vx.data(), n);
    put(f, vy.data(), n);
    put(f, mp.ids().data(), n);
    const MaterialGrid& m = solver.materialGrid();
    put(f, m.data().data(), m.data().size());
    const StaggeredVelocityGrid& g = solver.fluidVelocityGrid();
    put(f, g.velocityGridU().data().data(), g.velocityGridU().data().size());
    put(f, g.velocityGridV().data().data(), g.velocityGridV().data().size());
    const int32_t subs = solver.timeStats().substepCount();
    const float dt = solver.stepDt();
    const float mid = g.velocityAt((float)solver.gridSizeI() * 0.5f, (float)solver.gridSizeJ() * 0.5f).x();
    put(f, &subs, 1);
    put(f, &dt, 1);
    put(f, &mid, 1);
    std::fclose(f);
    return 0;
}

static void run_frames(FlipSolver& solver, int frames)
{
    for (int k = 0; k < frames; k++) {
        solver.stepFrame();
        std::printf("frame %d: %zu particles, %d substeps, %d pressure iterations, %.2f ms\n", k,
                    solver.particleCount(), solver.timeStats().substepCount(), solver.timeStats().pressureIterations(),
                    solver.timeStats().frameTime());
        std::fflush(stdout);
    }
}

static FlipSolverParameters dam_params(int res, int ppc)
{
    FlipSolverParameters p;
    p.gridSizeI = res;
    p.gridSizeJ = res;
    p.dx = 1.0 / res;
    p.particlesPerCell = ppc;
    p.seed = 42;
    return p;
}

static int run_dam(int res, int ppc, int frames, const char* out, bool hooked)
{
    const FlipSolverParameters p = dam_params(res, ppc);
    if (hooked) {
        HookedSolver solver(&p);
        dam_scene(solver, res);
        run_frames(solver, frames);
        std::printf("applyBodyForces hook ran %d times\n", solver.forceCalls);
        if (solver.forceCalls == 0) return 5;
        return write_state(solver, out);
    }
    FlipSolver solver(&p);
    dam_scene(solver, res);
    run_frames(solver, frames);
    return write_state(solver, out);
}

// The project() pattern (flipsolver2d.cpp:93-126): the matrix from the
// materials, a solve with the iteration limit and tolerance, the failure
// test iters >= limit.
static int run_solve(const char* in, const char* out, bool maxabs)
{
    FILE* f = std::fopen(in, "rb");
    if (!f) return 2;
    int32_t I = 0, J = 0, limit = 0;
    double scale = 0, tol = 0;
    get(f, &I, 1);
    get(f, &J, 1);
    get(f, &scale, 1);
    get(f, &tol, 1);
    get(f, &limit, 1);
    MaterialGrid mat((size_t)I, (size_t)J, EMPTY);
    std::vector<double> b((size_t)I * J), x;
    get(f, mat.data().data(), mat.data().size());
    get(f, b.data(), b.size());
    std::fclose(f);
    const IndexedPressureParameters A = IndexedPressureParameters::fromMaterials(mat, scale);
    const IndexedIPPCoefficients M(A);
    LinearSolver solver;
    solver.setToleranceMode(maxabs ? LinearSolver::MAX_ABS : LinearSolver::RELATIVE_L2);
    const int32_t ret = solver.solve(A, M, x, b, limit, tol);
    if (ret >= limit) std::printf("Pressure solver solving failed!\n");
    const int32_t it = solver.lastIterations();
    const double r = solver.lastResidual();
    f = std::fopen(out, "wb");
    if (!f) return 2;
    put(f, &ret, 1);
    put(f, &it, 1);
    put(f, &r, 1);
    put(f, x.data(), x.size());
    std::fclose(f);
    std::printf("solve %dx%d: returned %d, %d iterations, rel residual %.3e, %zu fluid rows\n", I, J, ret, it, r,
                A.data().size());
    return 0;
}

static int run_json(const char* scene, int frames, const char* out)
{
    std::shared_ptr<FlipSolver> solver = JsonSceneReader::loadJson(scene);
    if (!solver) return 6;
    run_frames(*solver, frames);
    return write_state(*solver, out);
}

static void print_polys(const char* kind, const std::vector<Geometry2d>& gs)
{
    for (const Geometry2d& g : gs) {
        std::printf("%s", kind);
        for (const Vec3& v : g.verts()) std::printf(" %.9g %.9g", v.x(), v.y());
        std::printf("\n");
    }
}

static int run_jsoninfo(const char* scene)
{
    const SceneDescription d = JsonSceneReader::readScene(scene);
    const FlipSolverParameters& p = d.params;
    std::printf("simulationMethod %d\nfluidDensity %.17g\nseed %u\ndx %.17g\nparticlesPerCell %d\n",
                (int)p.simulationMethod, p.fluidDensity, p.seed, p.dx, p.particlesPerCell);
    std::printf("globalAcceleration %.9g %.9g\nresolution %.9g\nfps %d\nmaxSubsteps %d\npicRatio %.9g\n",
                p.globalAcceleration.x(), p.globalAcceleration.y(), p.resolution, p.fps, p.maxSubsteps, p.picRatio);
    std::printf("cflNumber %.9g\nparticleScale %.9g\npcgIterLimit %d\ndomainSizeI %.9g\ndomainSizeJ %.9g\n",
                p.cflNumber, p.particleScale, p.pcgIterLimit, p.domainSizeI, p.domainSizeJ);
    std::printf("gridSizeI %d\ngridSizeJ %d\nsceneScale %.9g\nviscosityEnabled %d\nparameterHandlingMethod %d\n",
                p.gridSizeI, p.gridSizeJ, p.sceneScale, (int)p.viscosityEnabled, (int)p.parameterHandlingMethod);
    std::vector<Geometry2d> g;
    for (const Obstacle& o : d.solids) {
        std::printf("friction %.9g\n", o.friction());
        g.push_back(o.geometry());
    }
    print_polys("solid", g);
    g.clear();
    for (const Emitter& e : d.sources) {
        std::printf("source_velocity %.9g %.9g %d\n", e.velocity().x(), e.velocity().y(), (int)e.velocityTransfer());
        std::printf("source_viscosity %.9g\n", e.viscosity());
        g.push_back(e.geometry());
    }
    print_polys("source", g);
    g.clear();
    for (const Sink& s : d.sinks) g.push_back(s.geo());
    print_polys("sink", g);
    g.clear();
    for (const Emitter& e : d.fluids) {
        std::printf("fluid_viscosity %.9g\n", e.viscosity());
        g.push_back(e.geometry());
    }
    print_polys("fluid", g);
    return 0;
}

int main(int argc, char** argv)
{
    try {
        if (argc == 6 && !std::strcmp(argv[1], "dam"))
            return run_dam(std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]), argv[5], false);
        if (argc == 6 && !std::strcmp(argv[1], "hooks"))
            return run_dam(std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]), argv[5], true);
        if (argc == 5 && !std::strcmp(argv[1], "solve")) return run_solve(argv[2], argv[3], std::atoi(argv[4]) != 0);
        if (argc == 5 && !std::strcmp(argv[1], "json")) return run_json(argv[2], std::atoi(argv[3]), argv[4]);
        if (argc == 3 && !std::strcmp(argv[1], "jsoninfo")) return run_jsoninfo(argv[2]);
    } catch (const Error& e) {
        std::fprintf(stderr, "error %d: %s\n", e.code(), e.what());
        return 1;
    }
    std::fprintf(stderr,
                 "usage: facade_demo dam|hooks <res> <ppc> <frames> <out> | solve <in> <out> <maxabs> | "
                 "json <scene> <frames> <out> | jsoninfo <scene>\n");
    return 2;
}
