/*
 * ref_harness.cpp -- C entry points over the reference's own building
 * blocks, compiled from the reference sources where they lie (see
 * oracle/Makefile).  TEST INFRASTRUCTURE ONLY: used to pin the oracle's
 * restatements.  Nothing here re-implements reference arithmetic except the
 * bookkeeping needed to feed LinearSolver::solve (unit lists + per-thread
 * data ranges, following flipsolver2d.cpp:797-945), whose inputs come from
 * the caller.
 *
 * FlipSolver itself (flipsolver2d.cpp) includes Eigen, which the reference
 * fetches at configure time and which is absent here, so it is not built.
 */
#include <cstdint>
#include <cstring>
#include <vector>

#include "grid2d.h"
#include "linearsolver.h"
#include "markerparticlesystem.h"
#include "materialgrid.h"
#include "mathfuncs.h"
#include "pressuredata.h"
#include "PressureIPPCoeficients.h"
#include "sdfgrid.h"
#include "staggeredvelocitygrid.h"
#include "threadpool.h"
#include "vmath.h"
#include "geometry2d.h"

extern "C" {

float ref_bspline(float v) { return simmath::bSpline(v); }
float ref_quadratic_bspline(float x, float y) { return simmath::quadraticBSpline(x, y); }
float ref_linear_hat(float v) { return simmath::linearHat(v); }
float ref_bilinear_hat(float x, float y) { return simmath::bilinearHat(x, y); }
float ref_frac(float v) { return simmath::frac(v); }
int ref_integr(float v) { return simmath::integr(v); }

int32_t ref_thread_count() { return (int32_t)ThreadPool::i()->threadCount(); }

int32_t ref_split_range(int64_t length, int64_t min_size, int64_t jobs, int64_t* starts, int64_t* ends)
{
    std::vector<Range> r = ThreadPool::i()->splitRange((size_t)length, (size_t)min_size, (size_t)jobs);
    for (size_t k = 0; k < r.size(); k++) {
        starts[k] = (int64_t)r[k].start;
        ends[k] = (int64_t)r[k].end;
    }
    return (int32_t)r.size();
}

void ref_lerp(const float* g, int32_t si, int32_t sj, float off_i, float off_j,
              const float* xi, const float* xj, int64_t n, float* out)
{
    Grid2d<float> grid(si, sj, 0.f, OOB_EXTEND, 0.f, Vec3(off_i, off_j));
    std::memcpy(grid.data().data(), g, sizeof(float) * (size_t)si * sj);
    for (int64_t k = 0; k < n; k++) out[k] = grid.interpolateAt(xi[k], xj[k]);
}

void ref_velocity_at(const float* u, const float* v, int32_t I, int32_t J,
                     const float* xi, const float* xj, int64_t n, float* ou, float* ov)
{
    StaggeredVelocityGrid g(I, J);
    std::memcpy(g.velocityGridU().data().data(), u, sizeof(float) * (size_t)(I + 1) * J);
    std::memcpy(g.velocityGridV().data().data(), v, sizeof(float) * (size_t)I * (J + 1));
    for (int64_t k = 0; k < n; k++) {
        Vec3 r = g.velocityAt(xi[k], xj[k]);
        ou[k] = r.x();
        ov[k] = r.y();
    }
}

void ref_closest_surface_point(const float* sdf, int32_t I, int32_t J,
                               const float* xi, const float* xj, int64_t n, float* oi, float* oj)
{
    SdfGrid g(I, J);
    std::memcpy(g.data().data(), sdf, sizeof(float) * (size_t)I * J);
    for (int64_t k = 0; k < n; k++) {
        Vec3 r = g.closestSurfacePoint(xi[k], xj[k]);
        oi[k] = r.x();
        oj[k] = r.y();
    }
}

void ref_extrapolate(float* g, uint8_t* valid, int32_t si, int32_t sj, int32_t radius)
{
    Grid2d<float> grid(si, sj, 0.f, OOB_EXTEND, 0.f);
    Grid2d<bool> flags(si, sj, false, OOB_CONST, true);
    std::memcpy(grid.data().data(), g, sizeof(float) * (size_t)si * sj);
    for (size_t k = 0; k < (size_t)si * sj; k++) flags.data()[k] = valid[k] != 0;
    simmath::breadthFirstExtrapolate(grid, flags, radius, 1, false);
    std::memcpy(g, grid.data().data(), sizeof(float) * (size_t)si * sj);
    for (size_t k = 0; k < (size_t)si * sj; k++) valid[k] = flags.data()[k] ? 1 : 0;
}

float ref_polygon_sdf(const float* verts, int32_t nverts, float x, float y)
{
    Geometry2d geo;
    for (int32_t k = 0; k < nverts; k++) geo.addVertex(Vec3(verts[2 * k], verts[2 * k + 1]));
    return geo.signedDistance(x, y);
}

/* Material predicates (materialgrid.h:15-20) as a bit set:
 * 1 fluid, 2 strict fluid, 4 empty, 8 solid, 16 source, 32 sink. */
int32_t ref_material_bits(int8_t code)
{
    const FluidMaterial m = static_cast<FluidMaterial>(code);
    return (fluidTest(m) ? 1 : 0) | (strictFluidTest(m) ? 2 : 0) | (emptyTest(m) ? 4 : 0) |
           (solidTest(m) ? 8 : 0) | (sourceTest(m) ? 16 : 0) | (sinkTest(m) ? 32 : 0);
}

static void fill_material(MaterialGrid& g, const int8_t* m, int32_t I, int32_t J)
{
    for (size_t k = 0; k < (size_t)I * J; k++) g.data()[k] = static_cast<FluidMaterial>(m[k]);
}

int32_t ref_nonsolid_count_lin(const int8_t* m, int32_t I, int32_t J, int64_t lin)
{
    MaterialGrid g(I, J, FluidMaterial::SINK);
    fill_material(g, m, I, J);
    return g.nonsolidNeighborCount((ssize_t)lin);
}

/* MaterialGrid::u/vSampleAffectedBySolid, materialgrid.cpp:152-166 */
void ref_sample_affected_by_solid(const int8_t* m, int32_t I, int32_t J, uint8_t* u_out, uint8_t* v_out)
{
    MaterialGrid g(I, J, FluidMaterial::SINK);
    fill_material(g, m, I, J);
    for (int32_t i = 0; i < I; i++)
        for (int32_t j = 0; j < J; j++) {
            u_out[(size_t)i * J + j] = g.uSampleAffectedBySolid(i, j);
            v_out[(size_t)i * J + j] = g.vSampleAffectedBySolid(i, j);
        }
}

void ref_bins_for_cell(int32_t I, int32_t J, int32_t i, int32_t j, int64_t* out9)
{
    MarkerParticleSystem mps(I, J, 3);
    std::array<ssize_t, 9> b = mps.binsForGridCell(i, j);
    for (int k = 0; k < 9; k++) out9[k] = b[k];
}

double ref_vops_maxabs(const double* v, int64_t n)
{
    std::vector<double> vec(v, v + n);
    return VOps::i().maxAbs(vec);
}

double ref_vops_dot(const double* a, const double* b, int64_t n)
{
    std::vector<double> x(a, a + n), y(b, b + n);
    return VOps::i().dot(x, y);
}

/* LinearSolver::solve (linearsolver.cpp:25-73) on a matrix given as per-cell
 * (fluid, nonsolid count, fluid-neighbour mask) and the reference IPP whose
 * coefficients are 1/(nonsolidNeighborCount(neighbour)*scale)
 * (flipsolver2d.cpp:924-927) from the reference MaterialGrid. */
int32_t ref_pcg(const int8_t* material, const uint8_t* fluid, const uint8_t* nsc, const uint8_t* mask,
                int32_t I, int32_t J, double scale, const double* b, double* x, int32_t iter_limit, double tol)
{
    LinearIndexable2d indexer(I, J);
    MaterialGrid mg(I, J, FluidMaterial::SINK);
    fill_material(mg, material, I, J);
    const size_t C = (size_t)I * J;
    IndexedPressureParameters A(C / 3, indexer, scale);
    IndexedIPPCoefficients M(C / 3, indexer);
    std::vector<Range> ranges = ThreadPool::i()->splitRange(C);
    size_t cur = 0;
    for (size_t l = 0; l < C; l++) {
        const bool cross = cur < ranges.size() && l >= ranges.at(cur).end;
        if (!fluid[l]) {
            if (cross) {
                A.endThreadDataRange();
                M.endThreadDataRange();
                cur++;
            }
            continue;
        }
        IndexedPressureParameterUnit pu;
        pu.unitIndex = l;
        pu.nonsolidNeighborCount = nsc[l];
        pu.fluidNeighborMask = mask[l];
        IndexedIPPCoefficientUnit mu;
        mu.unitIndex = l;
        mu.iNeg = 1.0 / (mg.nonsolidNeighborCount((ssize_t)l - J) * scale);
        mu.iPos = 1.0 / (mg.nonsolidNeighborCount((ssize_t)l + J) * scale);
        mu.jNeg = 1.0 / (mg.nonsolidNeighborCount((ssize_t)l - 1) * scale);
        mu.jPos = 1.0 / (mg.nonsolidNeighborCount((ssize_t)l + 1) * scale);
        if (cross) {
            A.endThreadDataRange();
            M.endThreadDataRange();
            cur++;
        }
        A.add(pu);
        M.add(mu);
    }
    if (cur < ranges.size()) {
        A.endThreadDataRange();
        M.endThreadDataRange();
    }
    LinearSolver solver;
    std::vector<double> res(C, 0.0), rhs(b, b + C);
    const int it = solver.solve(A, M, res, rhs, iter_limit, tol);
    std::memcpy(x, res.data(), sizeof(double) * C);
    return it;
}

} /* extern "C" */
