// fg_init.hip -- scene set-up and the source/reseed stages on the device.
//
// First-frame init (FlipSolver::firstFrameInit, flipsolver2d.cpp:788-795):
// one thread per cell runs the reference's four cell passes in its order --
// updateSinks (555-571), updateSources (532-553), updateSolids (502-530),
// updateInitialFluid (573-590) -- against the scene's polygons, each with the
// coordinate arithmetic of its pass, so the classification and the solid
// SDF equal the host restatement bit for bit.  Seeding (seedInitialFluid,
// 682-707): the strict-fluid cells' particle counts are scanned into slots
// (the counting sort's scan), and each cell writes its particles in (cell,
// id) order, quadrant-jittered by the counter-based RNG.
//
// Sources (FlipSolver::afterTransfer, 390-410, and reseedParticles, 627-680):
// after the grid update the faces U(i,j), V(i,j) of SOURCE cells whose
// emitter transfers velocity take the emitter's velocity; after
// countParticles a SOURCE cell holding fewer than ppc/2 particles gets the
// difference, appended after the live particles in cell order (a scan of
// the per-cell additions), at whole-cell uniform positions, with the final
// grid's velocity when the source transfers it.
#include "fg_kernels.h"

namespace fg {

// Geometry2d::signedDistance, geometry2d.cpp:54-73 (the polygon distance of
// Inigo Quilez's formula).  The zero z term of Vec3::dot is kept and the
// square root is taken in double, as there.
__device__ float polygon_sdf(const float2* v, int N, float px, float py)
{
    auto dot = [](float ax, float ay, float bx, float by) { return (ax * bx + ay * by) + 0.f * 0.f; };
    const float dx0 = px - v[0].x, dy0 = py - v[0].y;
    float sq = dot(dx0, dy0, dx0, dy0);
    float sgn = 1.0f;
    for (int i = 0, j = N - 1; i < N; j = i, i++) {
        const float ex = v[j].x - v[i].x, ey = v[j].y - v[i].y;
        const float wx = px - v[i].x, wy = py - v[i].y;
        const float h = fclamp(dot(wx, wy, ex, ey) / dot(ex, ey, ex, ey), 0.0f, 1.0f);
        const float bx = wx - ex * h, by = wy - ey * h;
        const float d = dot(bx, by, bx, by);
        sq = d < sq ? d : sq;   // std::min
        const bool xf = py >= v[i].y;
        const bool yf = py < v[j].y;
        const bool zf = ex * wy > ey * wx;
        if ((xf && yf && zf) || (!xf && !yf && !zf)) sgn *= -1.0f;
    }
    return (float)(sgn * sqrt((double)sq));
}

__global__ void __launch_bounds__(kBlock) k_init_cells(const SceneDev* scenes, InitGeo geo)
{
    const SceneDev& s = scenes[blockIdx.y];
    const int I = s.I, J = s.J, C = I * J;
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= C) return;
    const int i = c / J, j = c - (c / J) * J;
    const float dx = geo.dx;
    int8_t mat = M_EMPTY;
    int32_t eid = -1;
    auto poly = [&](int k) { return geo.verts + geo.voff[k]; };
    auto nv = [&](int k) { return geo.voff[k + 1] - geo.voff[k]; };
    // updateSinks: float centre
    {
        const float x = ((float)i + 0.5f) * dx, y = ((float)j + 0.5f) * dx;
        for (int k = geo.set[0].first; k < geo.set[0].first + geo.set[0].n; k++)
            if (polygon_sdf(poly(k), nv(k), x, y) <= 0.f) mat = M_SINK;
    }
    // updateSources: i*dx + dx/2 in float; the last source holding the cell wins
    {
        const float hdx = (float)((double)dx / 2.0);
        const float x = (float)i * dx + hdx, y = (float)j * dx + hdx;
        for (int k = geo.set[1].first, e = 0; k < geo.set[1].first + geo.set[1].n; k++, e++)
            if (polygon_sdf(poly(k), nv(k), x, y) <= 0.f) {
                mat = M_SOURCE;
                eid = e;
            }
    }
    // updateSolids: centre in double, distance in cells
    float dist = 3.40282347e+38f;
    int min_idx = 0;
    {
        const float x = (float)(((float)i + 0.5) * (double)dx), y = (float)(((float)j + 0.5) * (double)dx);
        for (int k = geo.set[2].first, o = 0; k < geo.set[2].first + geo.set[2].n; k++, o++) {
            const float sd = polygon_sdf(poly(k), nv(k), x, y) / dx;
            if (sd < dist) {
                min_idx = o;
                dist = sd;
            }
        }
    }
    int32_t sid = -1;
    if (dist < 0) {
        mat = M_SOLID;
        sid = min_idx;
    }
    float fsdf = 0.f, visc = 0.f;
    if (!s.nb) {
        // updateInitialFluid: float centre; the fluid's viscosity into the grid
        const float x = ((float)i + 0.5f) * dx, y = ((float)j + 0.5f) * dx;
        for (int k = geo.set[3].first; k < geo.set[3].first + geo.set[3].n; k++)
            if (polygon_sdf(poly(k), nv(k), x, y) <= 0.f) {
                mat = M_FLUID;
                visc = geo.visc[k];
            }
    } else {
        // NBFlipSolver::fluidSdfFromInitialFluid (nbflipsolver.cpp:291-318):
        // centre in double as updateSolids, the nearest fluid's distance in
        // cells; FLUID where negative, whatever the cell was
        const float x = (float)(((float)i + 0.5) * (double)dx), y = (float)(((float)j + 0.5) * (double)dx);
        fsdf = 3.40282347e+38f;
        int fid = -1;
        for (int k = geo.set[3].first; k < geo.set[3].first + geo.set[3].n; k++) {
            const float sd = polygon_sdf(poly(k), nv(k), x, y) / dx;
            if (sd < fsdf) {
                fsdf = sd;
                fid = k;
            }
        }
        if (fsdf < 0) {
            mat = M_FLUID;
            if (fid >= 0) visc = geo.visc[fid];   // the nearest fluid's viscosity
        }
        // updateGridFromSources' distance (320-349) at the cell corner
        // (i dx, j dx); the sources never move, so it is taken once here
        const float cx = (float)i * dx, cy = (float)j * dx;
        float sdist = 3.40282347e+38f;
        int sid = -1;
        for (int k = geo.set[1].first, e = 0; k < geo.set[1].first + geo.set[1].n; k++, e++) {
            const float sd = polygon_sdf(poly(k), nv(k), cx, cy) / dx;
            if (sd < sdist) {
                sdist = sd;
                sid = e;
            }
        }
        s.src_sdf[c] = sdist;
        if (s.visc_on) s.src_id[c] = sdist < 0 ? sid : -1;
        s.pcount[c] = 0;   // m_fluidParticleCounts.fill(0)
    }
    s.mat[c] = mat;
    s.sid[c] = sid;
    s.eid[c] = eid;
    s.ssdf[c] = dist;
    s.fsdf[c] = fsdf;
    if (s.visc_on) s.cvisc[c] = visc;
    s.x64[c] = 0.0;
    // seeding: this cell's particles (NB-FLIP's initialFluidSeed adds none,
    // nbflipsolver.cpp:254-289)
    s.cellCount[c] = (!s.nb && is_strict_fluid(mat)) ? s.ppc : 0;
}

// seedInitialFluid (flipsolver2d.cpp:682-707): particle k of a strict-fluid
// cell is jittered uniformly inside quadrant k % 4 of the cell; ids follow
// (cell, k) order, so the particles come out sorted.  Clears the cell
// counts the scan read (the sort's histogram starts from zero).
__global__ void __launch_bounds__(kBlock) k_seed(const SceneDev* scenes, uint32_t seed)
{
    const SceneDev& s = scenes[blockIdx.y];
    const int I = s.I, J = s.J, C = I * J;
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= C) return;
    const int cnt = s.cellCount[c];
    if (cnt == 0) return;
    s.cellCount[c] = 0;
    const int i = c / J, j = c - (c / J) * J;
    const int base = s.cellStart[c];
    for (int k = 0; k < cnt; k++) {
        const int64_t idx = (int64_t)base + k;
        if (idx >= s.cap) return;   // the host refuses a scene whose seeding overflows
        const int q = k & 3;
        const float qi = (float)(q >> 1), qj = (float)(q & 1);
        const float x = (float)i + 0.5f * (qi + jitter(seed, (uint32_t)c, (uint32_t)k, 0));
        const float y = (float)j + 0.5f * (qj + jitter(seed, (uint32_t)c, (uint32_t)k, 1));
        s.pos[idx] = make_float2(x, y);
        s.vel[idx] = make_float2(0.f, 0.f);
        s.pid[idx] = (uint32_t)idx;
        s.key[idx] = 0;
        // m_viscosityGrid.interpolateAt(pos), offset (1/2, 1/2) (flipsolver2d.cpp:37, 697-702)
        if (s.visc_on) s.pvisc[idx] = lerp_grid(s.cvisc, I, J, 0.5f, 0.5f, x, y);
    }
}

// Clear the cell counts of a set-up that stops after the scan (a scene
// whose seeding would overflow its capacity).
__global__ void __launch_bounds__(kBlock) k_clear_counts(const SceneDev* scenes)
{
    const SceneDev& s = scenes[blockIdx.y];
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c < s.I * s.J) s.cellCount[c] = 0;
}

// FlipSolver::afterTransfer, flipsolver2d.cpp:390-410: the working grid's
// faces U(i,j), V(i,j) of a SOURCE cell whose emitter transfers velocity.
__global__ void __launch_bounds__(kBlock) k_after_transfer(const SceneDev* scenes)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active || !s.has_sources) return;
    const int I = s.I, J = s.J, C = I * J;
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= C || s.mat[c] != M_SOURCE) return;
    const int e = s.eid[c];
    if (e < 0 || e >= s.n_emitters) return;
    const float4 em = s.emitters[e];
    if (em.z == 0.f) return;
    const int i = c / J, j = c - (c / J) * J;
    s.u[(size_t)i * J + j] = em.x;
    s.v[(size_t)i * (J + 1) + j] = em.y;
    s.uval[(size_t)i * J + j] = 1;
    s.vval[(size_t)i * (J + 1) + j] = 1;
}

// NBFlipSolver::reseedParticles (nbflipsolver.cpp:115-201): a SOURCE cell or
// a cell whose level set lies inside the band (narrow, resampling) offers
// ppc - count candidates (nb_candidates); candidate k, the counter RNG's k-th
// position, is skipped in a SOURCE cell above the resampling band or where
// the level set at the candidate is below the narrow band (nb_keep).
__device__ __forceinline__ int nb_candidates(const SceneDev& s, int c)
{
    const float sd = s.fsdf[c];
    const bool src = s.mat[c] == M_SOURCE;
    if (!(src || (sd < kNbResamplingBand && sd > kNbNarrowBand))) return 0;
    const int more = s.ppc - s.pcount[c];
    return more > 0 ? more : 0;
}
__device__ __forceinline__ bool nb_keep(const SceneDev& s, int c, float x, float y)
{
    const bool src = s.mat[c] == M_SOURCE;
    const float ns = lerp_grid(s.fsdf, s.I, s.J, 0.f, 0.f, x, y);
    return !((src && s.fsdf[c] > kNbResamplingBand) || ns < kNbNarrowBand);
}

// reseedParticles, part 1: the particles each cell gains -- FlipSolver's
// SOURCE cells, or NB-FLIP's band cells (the cell counts of countParticles
// are this substep's, k_sdf_materials).
__global__ void __launch_bounds__(kBlock) k_reseed_count(const SceneDev* scenes)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active || !(s.has_sources || s.nb)) return;
    const int C = s.I * s.J;
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= C) return;
    if (s.nb) {
        const int more = nb_candidates(s, c);
        int kept = 0;
        if (more > 0) {
            const uint32_t key = reseed_key(s.rseed, s.epoch);
            const int i = c / s.J, j = c - (c / s.J) * s.J;
            for (int k = 0; k < more; k++) {
                const float x = (float)i + jitter(key, (uint32_t)c, (uint32_t)k, 0);
                const float y = (float)j + jitter(key, (uint32_t)c, (uint32_t)k, 1);
                kept += nb_keep(s, c, x, y) ? 1 : 0;
            }
        }
        s.cellCount[c] = kept;
        return;
    }
    const int more = s.ppc / 2 - s.pcount[c];
    s.cellCount[c] = (s.mat[c] == M_SOURCE && more > 0) ? more : 0;
}

// part 2 (after the scan of those counts into cellStart): write them after
// the live particles, in cell order, with consecutive ids; the next
// substep's max |v|^2 takes their velocities in (atomicMax on the bits of
// non-negative floats orders them as the floats).
__global__ void __launch_bounds__(kBlock) k_reseed_write(const SceneDev* scenes)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active || !(s.has_sources || s.nb)) return;
    const int I = s.I, J = s.J, C = I * J;
    const int c = blockIdx.x * kBlock + threadIdx.x;
    if (c >= C) return;
    const int more = s.cellCount[c];
    if (more == 0) return;
    s.cellCount[c] = 0;
    const int i = c / J, j = c - (c / J) * J;
    const int64_t n = s.sc->n_particles;
    const int off = s.cellStart[c];
    const uint32_t key = reseed_key(s.rseed, s.epoch);
    const int e = s.eid[c];
    const bool transfer = e >= 0 && e < s.n_emitters && s.emitters[e].z != 0.f;
    float m = 0.f;
    // NB-FLIP: walk the candidates again and write the kept ones
    const int cand = s.nb ? nb_candidates(s, c) : more;
    for (int k = 0, a = 0; k < cand && a < more; k++) {
        const float x = (float)i + jitter(key, (uint32_t)c, (uint32_t)k, 0);
        const float y = (float)j + jitter(key, (uint32_t)c, (uint32_t)k, 1);
        if (s.nb && !nb_keep(s, c, x, y)) continue;
        const int64_t idx = n + off + a;
        if (idx >= s.cap) break;
        const float2 v = transfer ? velocity_at(s.uf, s.vf, I, J, x, y) : make_float2(0.f, 0.f);
        s.pos[idx] = make_float2(x, y);
        s.vel[idx] = v;
        s.pid[idx] = s.next_pid + (uint32_t)(off + a);
        s.key[idx] = 0;
        // the emitter's viscosity; NB-FLIP: else the grid's at the position
        if (s.visc_on)
            s.pvisc[idx] = (e >= 0 && e < s.n_emitters) ? s.emitters[e].w
                           : (s.nb && e == -1 ? lerp_grid(s.cvisc, I, J, 0.5f, 0.5f, x, y) : 0.f);
        const float v2 = v.x * v.x + v.y * v.y;
        m = v2 > m ? v2 : m;
        a++;
    }
    if (m > 0.f) atomicMax(&s.sc->max_v2_bits, __float_as_uint(m));
}

// part 3: commit the new count (reseeds beyond the capacity are dropped)
__global__ void k_reseed_commit(const SceneDev* scenes)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active || !(s.has_sources || s.nb) || threadIdx.x != 0) return;
    SceneScalars* sc = s.sc;
    const int64_t n = sc->n_particles, total = s.cellStart[s.I * s.J];
    const int64_t added = total < s.cap - n ? total : s.cap - n;
    sc->n_particles = (int32_t)(n + added);
    sc->reseeded = (int32_t)added;
    sc->reseed_dropped = (int32_t)(total - added);
}

void launch_init_cells(dim3 g, hipStream_t st, const SceneDev* sc, const InitGeo& geo)
{
    hipLaunchKernelGGL(k_init_cells, g, dim3(kBlock), 0, st, sc, geo);
}
void launch_seed(dim3 g, hipStream_t st, const SceneDev* sc, uint32_t seed)
{
    hipLaunchKernelGGL(k_seed, g, dim3(kBlock), 0, st, sc, seed);
}
void launch_clear_counts(dim3 g, hipStream_t st, const SceneDev* sc)
{
    hipLaunchKernelGGL(k_clear_counts, g, dim3(kBlock), 0, st, sc);
}
void launch_after_transfer(dim3 g, hipStream_t st, const SceneDev* sc)
{
    hipLaunchKernelGGL(k_after_transfer, g, dim3(kBlock), 0, st, sc);
}
void launch_reseed_count(dim3 g, hipStream_t st, const SceneDev* sc)
{
    hipLaunchKernelGGL(k_reseed_count, g, dim3(kBlock), 0, st, sc);
}
void launch_reseed_write(dim3 g, hipStream_t st, const SceneDev* sc)
{
    hipLaunchKernelGGL(k_reseed_write, g, dim3(kBlock), 0, st, sc);
}
void launch_reseed_commit(dim3 g, hipStream_t st, const SceneDev* sc)
{
    hipLaunchKernelGGL(k_reseed_commit, g, dim3(64), 0, st, sc);
}

}  // namespace fg
