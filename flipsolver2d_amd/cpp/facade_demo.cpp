// facade_demo.cpp -- drives the reference-named C++ API the way the
// reference's AutoBench drives FlipSolver (AutoBench/benchmarkrunnerapplication.cpp:88-92):
// build a dam-break scene, call stepFrame() K times, read the accessors.
//
//   facade_demo dam <res> <ppc> <frames> <out.bin>
//       writes: int64 n, f32 px[n], py[n], vx[n], vy[n], u32 id[n],
//               i8 material[I*J], f32 u[(I+1)*J], f32 v[I*(J+1)],
//               i32 substeps (last frame), f32 stepDt, f32 velocityAt(I/2, J/2).x
//   facade_demo solve <in.bin> <out.bin>
//       reads:  i32 I, J, f64 scale, f64 tol, i32 iterLimit, i8 material[I*J], f64 b[I*J]
//       writes: i32 iterations, f64 rel_residual, f64 x[I*J]
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "flipsolver2d_hip.h"

using namespace fghip;

static Geometry2d box(double x0, double x1, double y0, double y1)
{
    const float a = (float)x0, b = (float)x1, c = (float)y0, e = (float)y1;
    return Geometry2d({Vec3(a, c), Vec3(a, e), Vec3(b, e), Vec3(b, c)});
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

static int run_dam(int res, int ppc, int frames, const char* out)
{
    FlipSolverParameters p;
    p.gridSizeI = res;
    p.gridSizeJ = res;
    p.dx = 1.0 / res;
    p.particlesPerCell = ppc;
    p.seed = 42;
    FlipSolver solver(&p);
    // the dam break: one-cell walls, fluid x in [0.3, 1), y in [0, 0.4) clipped to the interior
    const double d = 1.0 / res;
    std::vector<Obstacle> walls = {Obstacle(box(0.0, d, 0.0, 1.0)), Obstacle(box(1.0 - d, 1.0, 0.0, 1.0)),
                                   Obstacle(box(0.0, 1.0, 0.0, d)), Obstacle(box(0.0, 1.0, 1.0 - d, 1.0))};
    for (Obstacle& w : walls) solver.addGeometry(w);
    Emitter fluid(box(0.3 > d ? 0.3 : d, 1.0 - d, d, 0.4));
    solver.addInitialFluid(fluid);
    for (int k = 0; k < frames; k++) {
        solver.stepFrame();
        std::printf("frame %d: %zu particles, %d substeps, %d pressure iterations, %.2f ms\n", k,
                    solver.particleCount(), solver.timeStats().substepCount(), solver.timeStats().pressureIterations(),
                    solver.timeStats().frameTime());
        std::fflush(stdout);
    }
    MarkerParticleSystem& mp = solver.markerParticles();
    const int64_t n = (int64_t)mp.particleCount();
    std::vector<float> px(n), py(n), vx(n), vy(n);
    for (int64_t k = 0; k < n; k++) {
        px[k] = mp.positions()[k].x;
        py[k] = mp.positions()[k].y;
        vx[k] = mp.velocities()[k].x;
        vy[k] = mp.velocities()[k].y;
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
    const float mid = g.velocityAt(res * 0.5f, res * 0.5f).x;
    put(f, &subs, 1);
    put(f, &dt, 1);
    put(f, &mid, 1);
    std::fclose(f);
    return 0;
}

static int run_solve(const char* in, const char* out)
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
    IndexedPressureParameters m;
    m.sizeI = (size_t)I;
    m.sizeJ = (size_t)J;
    m.scale = scale;
    m.materials.resize((size_t)I * J);
    std::vector<double> b((size_t)I * J), x;
    get(f, m.materials.data(), m.materials.size());
    get(f, b.data(), b.size());
    std::fclose(f);
    LinearSolver solver;
    const int32_t it = solver.solve(m, IndexedIPPCoefficients{}, x, b, limit, tol);
    const double r = solver.lastResidual();
    f = std::fopen(out, "wb");
    if (!f) return 2;
    put(f, &it, 1);
    put(f, &r, 1);
    put(f, x.data(), x.size());
    std::fclose(f);
    std::printf("solve %dx%d: %d iterations, rel residual %.3e\n", I, J, it, r);
    return 0;
}

int main(int argc, char** argv)
{
    try {
        if (argc == 6 && !std::strcmp(argv[1], "dam"))
            return run_dam(std::atoi(argv[2]), std::atoi(argv[3]), std::atoi(argv[4]), argv[5]);
        if (argc == 4 && !std::strcmp(argv[1], "solve")) return run_solve(argv[2], argv[3]);
    } catch (const Error& e) {
        std::fprintf(stderr, "error %d: %s\n", e.code(), e.what());
        return 1;
    }
    std::fprintf(stderr, "usage: facade_demo dam <res> <ppc> <frames> <out> | solve <in> <out>\n");
    return 2;
}
