// fg_grid.hip -- cell- and face-space kernels of the substep.
//
// Layout per scene: cells row-major (i*J + j); U faces (I+1) x J, V faces
// I x (J+1) (FlipSolver2dLib/staggeredvelocitygrid.cpp:6-14).  One thread
// per cell; blocks of a launch are remapped so each XCD gets a contiguous
// band of rows.
#include "fg_kernels.h"

namespace fg {

__device__ __forceinline__ int cell_index(int nbx)
{
    return xcd_tile(blockIdx.x, nbx) * kBlock + threadIdx.x;
}

// ---------------------------------------------------------------------------
// Matrix snapshot: FlipSolver::getPressureProjectionMatrix
// (flipsolver2d.cpp:797-887), stored matrix-free as one byte per cell:
// bits 0-3 fluid-neighbour mask (i-, i+, j-, j+; set only when the *linear*
// offset index is inside [0, I*J)), bits 4-6 nonsolid count n, bit 7 fluid.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_stencil(const SceneDev* scenes)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active) return;
    const int I = s.I, J = s.J, C = I * J;
    const int c = cell_index(gridDim.x);
    if (c >= C) return;
    const int i = c / J, j = c - (c / J) * J;
    const int8_t m = s.mat[c];
    uint8_t code = 0;
    if (is_fluid(m)) {
        const int di[4] = {-1, 1, 0, 0}, dj[4] = {0, 0, -1, 1};
        const int off[4] = {-J, J, -1, 1};
        int n = 0, mask = 0;
#pragma unroll
        for (int d = 0; d < 4; d++) {
            const int8_t mn = mat_ext(s.mat, I, J, i + di[d], j + dj[d]);
            if (is_fluid(mn)) {
                n++;
                const int ln = c + off[d];
                if (ln >= 0 && ln < C) mask |= 1 << d;
            } else if (is_empty(mn)) {
                n++;
            }
        }
        code = (uint8_t)(0x80 | (n << 4) | mask);
    }
    s.stencil[c] = code;
}

// ---------------------------------------------------------------------------
// P2G: FlipSolver::particleVelocityToGridThread (flipsolver2d.cpp:1329-1377).
// A particle reaches face U(i,j)/V(i,j) only if both quadratic B-spline
// weights exceed 1e-9, which confines it to cells (i-1..i+1, j-1..j+1): three
// contiguous (cell, id)-ordered particle ranges.  Deterministic gather, no
// atomics.  U row I and V column J are never written by the reference's
// cell loop and stay 0 / invalid.
//
// A block owns a 16 x 16 tile of cells and first stages the positions of
// the particles of the 18 x 18-cell window around it in LDS: in (cell, id)
// order each window row is one contiguous segment, so the loads are
// coalesced and every position is read from HBM about (18/16)^2 times
// instead of nine.  Each cell then sums over its 3 x 3 cells from LDS in the
// same order as before (bit-exact); velocities are read only for particles
// that pass the weight gate.  The capacity holds 2*ppc = 16 particles per
// window cell at 8 ppc; a fuller window reads positions from global memory.
// ---------------------------------------------------------------------------
constexpr int kP2gTile = 16;
// A tile's window (18 x 18 cells) is staged in LDS, positions and
// velocities, when it holds at most 8 particles a cell on average (41.5 KiB,
// three blocks per CU); a denser window reads them through the cache.
// Staging both at up to 16 a cell (83 KiB, one block per CU) was 5 % slower
// on the 256 mixed scenes, at 10 or 12 no better than at 8.
constexpr int kP2gPerCell = 8;
constexpr int kP2gCap = (kP2gTile + 2) * (kP2gTile + 2) * kP2gPerCell;

__global__ void __launch_bounds__(kBlock) k_p2g(const SceneDev* scenes)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active) return;
    const int I = s.I, J = s.J;
    const int tilesJ = (J + kP2gTile - 1) / kP2gTile;
    const int tiles = ((I + kP2gTile - 1) / kP2gTile) * tilesJ;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    if (tile >= tiles) return;
    const int ti = tile / tilesJ, tj = tile - (tile / tilesJ) * tilesJ;
    const int i0 = ti * kP2gTile, j0 = tj * kP2gTile;
    const int wr0 = i0 > 0 ? i0 - 1 : 0, wr1 = min(i0 + kP2gTile, I - 1);
    const int wc0 = j0 > 0 ? j0 - 1 : 0, wc1 = min(j0 + kP2gTile, J - 1);
    const int nrows = wr1 - wr0 + 1;   // <= 18
    __shared__ float2 spos[kP2gCap], svel[kP2gCap];
    __shared__ int seg_beg[kP2gTile + 2], seg_off[kP2gTile + 3];
    // this thread's face and its three rows' particle ranges, loaded now with
    // the window's bounds (clamped rows; a row outside the grid gets an
    // empty range)
    const int i = i0 + (threadIdx.x >> 4), j = j0 + (threadIdx.x & 15);
    const bool face = i < I && j < J;
    const int ic0 = min(i, I - 1), jc = min(j, J - 1);
    const int jl = jc > 0 ? jc - 1 : 0, jh = jc < J - 1 ? jc + 1 : J - 1;
    gptr<const int32_t> cs = gbl<const int32_t>(s.cellStart);
    int ra[3], re[3];
#pragma unroll
    for (int q = 0; q < 3; q++) {
        const int ii = ic0 + q - 1, icl = min(max(ii, 0), I - 1);
        const int a = cs[icl * J + jl], e = cs[icl * J + jh + 1];
        const bool in = ii >= 0 && ii < I;
        ra[q] = in ? a : 0;
        re[q] = in ? e : 0;
    }
    if (threadIdx.x < 64) {
        const int r = threadIdx.x;
        int beg = 0, cnt = 0;
        if (r < nrows) {
            const size_t row = (size_t)(wr0 + r) * J;
            beg = s.cellStart[row + wc0];
            cnt = s.cellStart[row + wc1 + 1] - beg;
        }
        int incl = cnt;
#pragma unroll
        for (int o = 1; o < 32; o <<= 1) {
            const int t = __shfl_up(incl, o, 64);
            if (r >= o) incl += t;
        }
        if (r < nrows) {
            seg_beg[r] = beg;
            seg_off[r + 1] = incl;
        }
        if (r == 0) seg_off[0] = 0;
    }
    __syncthreads();
    const int total = seg_off[nrows];
    const bool staged = total <= kP2gCap;
    if (staged) {
        constexpr int U = 4;   // loads in flight per thread
        for (int t0 = threadIdx.x; t0 < total; t0 += U * kBlock) {
            float2 v[U], w[U];
#pragma unroll
            for (int u = 0; u < U; u++) {
                const int t = t0 + u * kBlock;
                if (t >= total) continue;
                int lo = 0, hi = nrows - 1;   // the row whose offset range holds t
                while (lo < hi) {
                    const int mid = (lo + hi + 1) >> 1;
                    if (seg_off[mid] <= t) lo = mid;
                    else hi = mid - 1;
                }
                v[u] = s.pos[seg_beg[lo] + (t - seg_off[lo])];
                w[u] = s.vel[seg_beg[lo] + (t - seg_off[lo])];
            }
#pragma unroll
            for (int u = 0; u < U; u++)
                if (t0 + u * kBlock < total) {
                    spos[t0 + u * kBlock] = v[u];
                    svel[t0 + u * kBlock] = w[u];
                }
        }
    }
    __syncthreads();
    if (!face) return;
    const float fi = (float)i, fj = (float)j;
    const float fih = fi + 0.5f, fjh = fj + 0.5f;
    float uw = 1e-10f, vw = 1e-10f, uv = 0.0f, vv = 0.0f;
    for (int di = -1; di <= 1; di++) {
        const int ii = i + di;
        if (ii < 0 || ii >= I) continue;
        const int a = ra[di + 1], e = re[di + 1];
        const int shift = seg_off[ii - wr0] - seg_beg[ii - wr0];
        // chunks of 8 particles: all loads first (clamped to the row's last
        // particle, unconditional, pinned so that none is sunk into the
        // guarded sums), then the sums in order
        constexpr int U = 8;
        for (int k0 = a; k0 < e; k0 += U) {
            float2 pp[U], vel[U];
            if (staged) {   // uniform over the block
#pragma unroll
                for (int u = 0; u < U; u++) {
                    pp[u] = spos[min(k0 + u, e - 1) + shift];
                    vel[u] = svel[min(k0 + u, e - 1) + shift];
                }
            } else {
#pragma unroll
                for (int u = 0; u < U; u++) {
                    pp[u] = ldg2(s.pos, min(k0 + u, e - 1));
                    vel[u] = ldg2(s.vel, min(k0 + u, e - 1));
                }
            }
#pragma unroll
            for (int u = 0; u < U; u++) asm volatile("" : "+v"(pp[u].x), "+v"(pp[u].y), "+v"(vel[u].x), "+v"(vel[u].y));
            // branch-free: a particle that does not pass adds +0, which
            // leaves every sum as it was (the weight sums stay positive and
            // the velocity sums, starting at +0, are never -0)
#pragma unroll
            for (int u = 0; u < U; u++) {
                const float2 p = pp[u];
                const float wu = qbspline(p.x - fi, p.y - fjh);
                const float wv = qbspline(p.x - fih, p.y - fj);
                const bool g = (k0 + u < e) & (wu > 1e-9f) & (wv > 1e-9f);
                const float cu = wu * vel[u].x, cv = wv * vel[u].y;
                uw += g ? wu : 0.f;
                uv += g ? cu : 0.f;
                vw += g ? wv : 0.f;
                vv += g ? cv : 0.f;
            }
        }
    }
    // a face was reached iff its weight sum left 1e-10 (each weight that
    // passes exceeds 1e-9)
    s.u[(size_t)i * J + j] = uv / uw;
    s.uval[(size_t)i * J + j] = uw > 1e-10f;
    s.v[(size_t)i * (J + 1) + j] = vv / vw;
    s.vval[(size_t)i * (J + 1) + j] = vw > 1e-10f;
    if (i == I - 1) {
        s.u[(size_t)I * J + j] = 0.f;
        s.uval[(size_t)I * J + j] = 0;
    }
    if (j == J - 1) {
        s.v[(size_t)i * (J + 1) + J] = 0.f;
        s.vval[(size_t)i * (J + 1) + J] = 0;
    }
}

// ---------------------------------------------------------------------------
// Fluid SDF: FlipSolver::updateSdfThread (flipsolver2d.cpp:1276-1311) over
// the reference's window -- the 3x3 bins of 3x3 cells around the cell's bin
// (markerparticlesystem.cpp:9-20, 109-128) -- then updateMaterials
// (1053-1075) and the per-cell particle count of countParticles.  All nine
// cells of a bin share that window, so the bin's kSdfLanes lanes split the
// window's rows, stream their particles once with nine running minima each,
// and combine the minima by lane shuffles (min is exact and order-free, so
// the result equals the per-cell loop bit for bit).
// ---------------------------------------------------------------------------
// kExact: the fallback pass (k_sdf_exact) -- every bin scans its window and
// only the SDF is written.  Otherwise a bin whose cells all hold particles
// (with the shortcut on) skips its scan: each of its cells is inside.
// Most bins scan nothing (inside, or far from the liquid), so every load a
// lane needs besides the particles is issued at the top, in one round: the
// window rows' bounds, the bin's 3 x 4 cell boundaries (its cells' particle
// counts) and the materials of the cells the lane writes.
template <bool kExact>
__device__ __forceinline__ void sdf_bins(const SceneDev& s)
{
    static_assert(kSdfLanes >= 4 && kSdfLanes <= 64 && (kSdfLanes & (kSdfLanes - 1)) == 0,
                  "a bin's lanes are a power-of-two group inside one wave");
    const int I = s.I, J = s.J;
    const int nbI = (I + 2) / 3, nbJ = (J + 2) / 3;
    const int sub = threadIdx.x & (kSdfLanes - 1);
    const int bin = xcd_tile(blockIdx.x, gridDim.x) * (kBlock / kSdfLanes) + threadIdx.x / kSdfLanes;
    if (bin >= nbI * nbJ) return;   // all lanes of a bin leave together
    const int bi = bin / nbJ, bj = bin - (bin / nbJ) * nbJ;
    const int r0 = max(0, 3 * bi - 3), r1 = min(I, 3 * bi + 6);
    const int c0 = max(0, 3 * bj - 3), c1 = min(J, 3 * bj + 6);
    const int i0 = 3 * bi, j0 = 3 * bj;
    // this lane's window rows r0+sub, r0+sub+kSdfLanes, ... and cells
    // q = sub, sub+kSdfLanes, ... of the bin (q = 3a + b)
    constexpr int NR = (9 + kSdfLanes - 1) / kSdfLanes;
    gptr<const int32_t> cs = gbl<const int32_t>(s.cellStart);
    int rb[NR], re[NR];
#pragma unroll
    for (int q = 0; q < NR; q++) {   // clamped rows: every load unconditional
        const int r = r0 + sub + q * kSdfLanes, rc = min(r, r1 - 1);
        rb[q] = cs[rc * J + c0];
        re[q] = cs[rc * J + c1];
    }
    int bnd[3][4];   // cellStart at (i0+a, j0+k), clamped to the grid (index i*J + J is row i's end)
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int k = 0; k < 4; k++) bnd[a][k] = cs[min(i0 + a, I - 1) * J + min(j0 + k, J)];
    int8_t mq[NR];
    if (!kExact) {
        gptr<const int8_t> mat = gbl<const int8_t>(s.mat);
#pragma unroll
        for (int q = 0; q < NR; q++) {   // cell sub + q kSdfLanes (clamped: loads unconditional)
            const int cq = min(sub + q * kSdfLanes, 8), a = cq / 3, b = cq - a * 3;
            mq[q] = mat[min(i0 + a, I - 1) * J + min(j0 + b, J - 1)];
        }
    }
#pragma unroll
    for (int q = 0; q < NR; q++) {
        const int r = r0 + sub + q * kSdfLanes;
        if (r >= r1) re[q] = rb[q];
    }
    float cx[3], cy[3];
#pragma unroll
    for (int q = 0; q < 3; q++) {
        cx[q] = (float)(i0 + q) + 0.5f;
        cy[q] = (float)(j0 + q) + 0.5f;
    }
    float d2[3][3];
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) d2[a][b] = 3.402823466e+38f;
    // Every particle of a cell lies within sqrt(0.5) of its centre, so with
    // particleScale above that (sdf_short) a cell holding a particle has
    // sdf < 0 whatever its window holds; extrapolateLevelsetInside then
    // overwrites its value (when some cell is outside, which k_sdf_exact
    // checks).  A bin whose cells all hold particles skips its scan.
    // (NB-FLIP combines the level set with the advected one before the inside
    // extrapolation, so it needs every value: no shortcut there)
    bool full = false;
    if (!kExact && s.sdf_short && !s.nb) {
        int have = 0;
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int b = 0; b < 3; b++)
                have += bnd[a][b + 1] > bnd[a][b] || i0 + a >= I || j0 + b >= J;
        full = have == 9;
        if (full) {
#pragma unroll
            for (int q = 0; q < NR; q++) re[q] = rb[q];
        }
    }
#pragma unroll
    for (int q = 0; q < NR; q++) {
        const int e = re[q];
        constexpr int U = 24;   // loads in flight per lane (a row segment holds ~9 ppc)
        for (int k0 = rb[q]; k0 < e; k0 += U) {
            // past the row's end a lane re-reads its last particle: the
            // minimum of a value with itself changes nothing, so no branch
            float2 pp[U];
#pragma unroll
            for (int u = 0; u < U; u++) pp[u] = ldg2(s.pos, min(k0 + u, e - 1));
#pragma unroll
            for (int u = 0; u < U; u++) {
#pragma unroll
                for (int a = 0; a < 3; a++) {
                    const float dxp = pp[u].x - cx[a];
                    const float dx2 = dxp * dxp;
#pragma unroll
                    for (int b = 0; b < 3; b++) {
                        const float dyp = pp[u].y - cy[b];
                        const float nd = dx2 + dyp * dyp;
                        if (nd < d2[a][b]) d2[a][b] = nd;
                    }
                }
            }
        }
    }
#pragma unroll
    for (int a = 0; a < 3; a++)
#pragma unroll
        for (int b = 0; b < 3; b++) {
#pragma unroll
            for (int o = 1; o < kSdfLanes; o <<= 1) {
                const float t = __shfl_xor(d2[a][b], o, 64);
                if (t < d2[a][b]) d2[a][b] = t;
            }
        }
    // cell q of the bin is written by lane q mod kSdfLanes (q, a, b
    // compile-time, so the arrays stay in registers)
#pragma unroll
    for (int q = 0; q < 9; q++) {
        if ((q & (kSdfLanes - 1)) != sub) continue;
        const int a = q / 3, b = q - (q / 3) * 3;
        const int i = i0 + a, j = j0 + b;
        if (i >= I || j >= J) continue;
        const int c = i * J + j;
        const int n = bnd[a][b + 1] - bnd[a][b];
        // a skipped bin's cells are inside: any negative value holds their place
        const float sdf = full ? -s.pscale : sqrtf(d2[a][b]) - s.pscale;
        if (kExact) {
            if (n > 0) s.fsdf[c] = sdf;   // the cells the fast pass skipped
            continue;
        }
        s.fsdf[c] = sdf;
        s.pcount[c] = min(n, 2 * s.ppc);
        if (s.nb) continue;   // NB-FLIP updates the materials from its final level set
        int8_t m = mq[q / kSdfLanes];
        if (sdf < 0.f) {
            if (is_empty(m)) m = M_FLUID;
        } else {
            if (is_strict_fluid(m)) m = M_EMPTY;
        }
        s.mat[c] = m;
    }
}

__global__ void __launch_bounds__(kBlock) k_sdf_materials(const SceneDev* scenes)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active) return;
    sdf_bins<false>(s);
}

// The exact SDF of every cell that holds particles, for a scene in which no
// cell is outside (sdf > 0): extrapolateLevelsetInside then has no known
// cell and keeps every value, so the values the fast pass skipped must be
// the true ones.  Runs after k_levelset_init; otherwise exits at once.
__global__ void __launch_bounds__(kBlock) k_sdf_exact(const SceneDev* scenes)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active || !s.sdf_short || cst(s.sc)->sdf_known) return;
    sdf_bins<true>(s);
}

// ---------------------------------------------------------------------------
// Breadth-first extrapolation as a level-synchronous LDS computation.
//
// The reference's FIFO BFS (simmath::breadthFirstExtrapolate,
// mathfuncs.cpp:152-215; FlipSolver::extrapolateLevelsetInside,
// flipsolver2d.cpp:1433-1494) gives every reached cell the level L = its
// 8-connected distance to the known set, and sets its value from the
// neighbours of level < L only, all of which are final before any level-L
// cell is processed.  So levels can be computed band by band: a tile of
// kTile x kTile cells with an H-cell halo holds every cell a level <= L0+H
// tile value depends on.  Sums are FP64 in getNeighborhood order
// (linearindexable2d.h:55-71) exactly as the reference accumulates them.
// ---------------------------------------------------------------------------
constexpr int kTile = 32;
static_assert(kTile == kBfsTile, "the host sizes the BFS launches and tile flags by kBfsTile");
constexpr int16_t kNone = 0x7fff;     // outside the grid
constexpr int16_t kUnreached = 0x7ffe;

// Levels are expanded as row bitmasks (the region is at most 64 cells
// wide): one wave holds one region row per lane, dilates the frontier of
// level L with shifts and lane shuffles, and removes the reached set, so the
// expansion costs a few instructions per row and level and needs no
// barrier.  The value pass then visits only the cells of each level.
template <int H, bool LEVELSET, bool OUT = false>
__device__ void bfs_band(const float* gin, const uint8_t* vin, const int32_t* mkin, float* gout, uint8_t* vout,
                         int32_t* mkout, int si, int sj, int L0, int* more, uint8_t* tile_more)
{
    constexpr int R = kTile + 2 * H;
    static_assert(R <= 64, "a region row must fit one 64-bit mask");
    // LDS row strides chosen so that the cells of a region column (a level's
    // vertical side, visited by consecutive threads) fall in different banks:
    // an odd number of words per row, for the floats and for the int16
    // markers (two to a word)
    constexpr int RS = R + 1;
    constexpr int RSM = ((R + 2) / 2) % 2 == 1 ? R + 2 : R + 4;
    constexpr int T = 16;   // threads as a 16 x 16 layout for the load and store passes
    __shared__ int16_t mk[R * RSM];
    __shared__ float val[R * RS];
    __shared__ unsigned long long front[H + 1][R];   // front[s]: region cells at level L0 + s
    __shared__ unsigned long long open_s[R];
    __shared__ int16_t pre[H + 1][64];                // row offsets of level L0+s's cells
    __shared__ int16_t tot[H + 1];
    __shared__ int top_s;
    const int tx = threadIdx.x & (T - 1), ty = threadIdx.x >> 4;
    const int tilesJ = (sj + kTile - 1) / kTile;
    const int tile = blockIdx.x;
    const int ti = tile / tilesJ, tj = tile - (tile / tilesJ) * tilesJ;
    const int gi0 = ti * kTile - H, gj0 = tj * kTile - H;
    // the region: every load of the thread issued before any is used
    // (clamped indices, the loads unconditional; each value is pinned by an
    // empty asm after all of them, so none is sunk into a later branch)
    // Extrapolation (!LEVELSET) loads the halo's values only if the tile has
    // work: a tile with no unknown cell, or no known one within reach, is
    // copied, and needs the region's validity bytes and its own values alone.
    constexpr int NA = (R + T - 1) / T;
    float xr[NA][NA];
    int lr[NA][NA];
    bool okr[NA][NA];
#pragma unroll
    for (int a = 0; a < NA; a++)
#pragma unroll
        for (int b = 0; b < NA; b++) {
            const int ri = ty + a * T, rj = tx + b * T;
            const int gi = gi0 + ri, gj = gj0 + rj;
            okr[a][b] = ri < R && rj < R && gi >= 0 && gi < si && gj >= 0 && gj < sj;
            const size_t g = (size_t)min(max(gi, 0), si - 1) * sj + min(max(gj, 0), sj - 1);
            const bool inner = ri >= H && ri < H + kTile && rj >= H && rj < H + kTile;
            xr[a][b] = (LEVELSET || inner) ? gbl<const float>(gin)[g] : 0.f;
            if (LEVELSET) lr[a][b] = gbl<const int32_t>(mkin)[g];
            else lr[a][b] = (int)gbl<const uint8_t>(vin)[g];
        }
#pragma unroll
    for (int a = 0; a < NA; a++)
#pragma unroll
        for (int b = 0; b < NA; b++) {
            asm volatile("" : "+v"(xr[a][b]), "+v"(lr[a][b]));
            xr[a][b] = okr[a][b] ? xr[a][b] : 0.f;
            lr[a][b] = okr[a][b] ? lr[a][b] : -1;
        }
    bool has_front = false, has_open = false;
#pragma unroll
    for (int a = 0; a < NA; a++)
#pragma unroll
        for (int b = 0; b < NA; b++) {
            const int ri = ty + a * T, rj = tx + b * T;
            if (ri >= R || rj >= R) continue;
            int16_t m = kNone;
            if (lr[a][b] >= 0) {
                if (LEVELSET) m = lr[a][b] <= L0 ? (int16_t)lr[a][b] : kUnreached;
                else m = lr[a][b] ? 0 : kUnreached;
            }
            has_front |= m == L0;
            has_open |= m == kUnreached;
            mk[ri * RSM + rj] = m;
            val[ri * RS + rj] = xr[a][b];
        }
    // a tile cell reaches level L0+s only through a level-L0 cell within
    // distance s <= H, i.e. inside this region
    const bool work = __syncthreads_or(has_front) && __syncthreads_or(has_open);
    if (!LEVELSET && work) {   // the halo's values (the barriers below order these stores)
        float hv[NA][NA];
#pragma unroll
        for (int a = 0; a < NA; a++)
#pragma unroll
            for (int b = 0; b < NA; b++) {
                const int ri = ty + a * T, rj = tx + b * T;
                const bool inner = ri >= H && ri < H + kTile && rj >= H && rj < H + kTile;
                const int gi = gi0 + ri, gj = gj0 + rj;
                const size_t g = (size_t)min(max(gi, 0), si - 1) * sj + min(max(gj, 0), sj - 1);
                hv[a][b] = (okr[a][b] && !inner) ? gbl<const float>(gin)[g] : 0.f;
            }
#pragma unroll
        for (int a = 0; a < NA; a++)
#pragma unroll
            for (int b = 0; b < NA; b++) {
                const int ri = ty + a * T, rj = tx + b * T;
                const bool inner = ri >= H && ri < H + kTile && rj >= H && rj < H + kTile;
                if (ri < R && rj < R && !inner) val[ri * RS + rj] = hv[a][b];
            }
    }
    if (!work) {
        if (!LEVELSET) {
            for (int li = ty; li < kTile; li += T)
                for (int lj = tx; lj < kTile; lj += T) {
                    const int gi = ti * kTile + li, gj = tj * kTile + lj;
                    if (gi >= si || gj >= sj) continue;
                    const size_t g = (size_t)gi * sj + gj;
                    const int r = (li + H) * RS + (lj + H), rm = (li + H) * RSM + (lj + H);
                    gout[g] = val[r];
                    vout[g] = mk[rm] != kUnreached;
                }
        } else if (threadIdx.x == 0) {
            *tile_more = 0;
        }
        return;
    }
    // row masks of the open (unreached) cells and of level L0, one ballot
    // per region row (4 waves, consecutive LDS addresses)
    {
        const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
        for (int r = w; r < R; r += kBlock / 64) {
            const int16_t m = lane < R ? mk[r * RSM + lane] : kNone;
            const unsigned long long o = __ballot(m == kUnreached);
            const unsigned long long f = __ballot(m == (int16_t)L0);
            if (lane == 0) {
                open_s[r] = o;
                front[0][r] = f;
            }
        }
    }
    __syncthreads();
    // expand levels L0+1 .. L0+H: wave 0, lane = region row; per level also
    // the row offsets of its cells
    if (threadIdx.x < 64) {
        const int r = threadIdx.x;
        unsigned long long open = r < R ? open_s[r] : 0ull, f = r < R ? front[0][r] : 0ull;
        int top = 0;
        for (int st = 1; st <= H; st++) {
            const unsigned long long d = f | (f << 1) | (f >> 1);
            const unsigned long long up = __shfl_up(d, 1, 64), dn = __shfl_down(d, 1, 64);
            unsigned long long nf = d;
            if (r > 0) nf |= up;
            if (r < R - 1) nf |= dn;
            nf &= open;
            if (r >= R) nf = 0;
            open &= ~nf;
            if (r < R) front[st][r] = nf;
            // exclusive scan of the row counts by bit slices: one ballot per
            // bit of the count (<= 64), no lane shuffles
            const int cnt = __popcll(nf);
            const unsigned long long below = (1ull << r) - 1ull;
            int excl = 0, total = 0;
#pragma unroll
            for (int b = 0; b < 7; b++) {
                const unsigned long long m = __ballot((cnt >> b) & 1);
                excl += __popcll(m & below) << b;
                total += __popcll(m) << b;
            }
            pre[st][r] = (int16_t)excl;
            if (r == 0) tot[st] = (int16_t)total;
            f = nf;
            if (!__any(nf != 0ull)) break;
            top = st;
        }
        if (r == 0) top_s = top;
    }
    __syncthreads();
    const int top = top_s;
    // values, level by level: the cells of a level spread over all threads
    // (cell t of the level: the row whose offset range holds t, then the
    // (t - offset)-th set bit of that row's mask)
    for (int st = 1; st <= top; st++) {
        const int L = L0 + st;
        const int n = tot[st];
        for (int t = threadIdx.x; t < n; t += kBlock) {
            int lo = 0, hi = 63;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (pre[st][mid] <= t) lo = mid;
                else hi = mid - 1;
            }
            const int ri = lo;
            int k = t - pre[st][ri];
            unsigned long long m = front[st][ri];
            int rj = 0;
#pragma unroll
            for (int w = 32; w >= 1; w >>= 1) {
                const unsigned long long low = m & ((1ull << w) - 1ull);
                const int c = __popcll(low);
                if (k >= c) {
                    k -= c;
                    m >>= w;
                    rj += w;
                } else {
                    m = low;
                }
            }
            // the 8 neighbours in getNeighborhood order, loads first
            int16_t qm[8];
            float qv[8];
#pragma unroll
            for (int e = 0; e < 8; e++) {
                const int di = (e < 3) ? -1 : (e < 5 ? 0 : 1);
                const int dj = (e < 3) ? e - 1 : (e < 5 ? (e == 3 ? -1 : 1) : e - 6);
                const int qi = ri + di, qj = rj + dj;
                const bool ok = qi >= 0 && qi < R && qj >= 0 && qj < R;
                qm[e] = ok ? mk[qi * RSM + qj] : kNone;
                qv[e] = ok ? val[qi * RS + qj] : 0.f;
            }
            double avg = 0;
            int count = 0;
#pragma unroll
            for (int e = 0; e < 8; e++)
                if (qm[e] < L) {
                    avg += (double)qv[e];
                    count++;
                }
            if (LEVELSET && OUT) val[ri * RS + rj] = (float)(avg / count + 1.f);
            else if (LEVELSET) val[ri * RS + rj] = (float)(avg / count - 1.f);
            else val[ri * RS + rj] = (float)(avg / count);
            mk[ri * RSM + rj] = (int16_t)L;   // same-level neighbours read L or unreached: not below L either way
        }
        __syncthreads();
    }
    // write the tile
    bool reached_end = false;
    for (int li = ty; li < kTile; li += T)
        for (int lj = tx; lj < kTile; lj += T) {
            const int gi = ti * kTile + li, gj = tj * kTile + lj;
            if (gi >= si || gj >= sj) continue;
            const int r = (li + H) * RS + (lj + H), rm = (li + H) * RSM + (lj + H);
            const size_t g = (size_t)gi * sj + gj;
            const int16_t m = mk[rm];
            if (LEVELSET) {
                if (m > L0 && m <= L0 + H) {
                    gout[g] = val[r];
                    mkout[g] = m;
                    reached_end |= m == L0 + H;
                }
            } else {
                gout[g] = val[r];
                vout[g] = (m != kUnreached) ? 1 : 0;
            }
        }
    if (LEVELSET) {
        const bool any = __syncthreads_or(reached_end);
        if (threadIdx.x == 0) {
            *tile_more = any ? 1 : 0;
            if (any) *more = 1;
        }
    }
}

// Velocity extrapolation, radius 10 (StaggeredVelocityGrid::extrapolate,
// staggeredvelocitygrid.cpp:146-156): levels 1..11 in one band.
// blockIdx.z: 0 = U grid, 1 = V grid.  which: 0 = (u -> su), 1 = (u -> uf).
__global__ void __launch_bounds__(kBlock) k_extrapolate(const SceneDev* scenes, int which)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active) return;
    const bool isU = blockIdx.z == 0;
    const int si = isU ? s.I + 1 : s.I, sj = isU ? s.J : s.J + 1;
    const int tiles = ((si + kTile - 1) / kTile) * ((sj + kTile - 1) / kTile);
    if ((int)blockIdx.x >= tiles) return;
    const float* gin = isU ? s.u : s.v;
    const uint8_t* vin = isU ? (which == 0 ? s.uval : s.uvalS) : (which == 0 ? s.vval : s.vvalS);
    float* gout = isU ? (which == 0 ? s.su : s.uf) : (which == 0 ? s.sv : s.vf);
    uint8_t* vout = isU ? (which == 0 ? s.uvalS : s.uvalF) : (which == 0 ? s.vvalS : s.vvalF);
    bfs_band<11, false>(gin, vin, nullptr, gout, vout, nullptr, si, sj, 0, nullptr, nullptr);
}

// Markers (known cells: sdf > 0, level 0), one block per BFS tile, which
// records whether it holds a known cell (band 0's front); clears the band
// flags.  OUT: extrapolateLevelsetOutside (flipsolver2d.cpp:1496-1557),
// whose known cells hold less than I*J (as a float) and whose levels add 1.
template <bool OUT>
__global__ void __launch_bounds__(kBlock) k_levelset_init(const SceneDev* scenes)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active) return;
    const int I = s.I, J = s.J;
    const int tilesJ = (J + kTile - 1) / kTile;
    const int tiles = ((I + kTile - 1) / kTile) * tilesJ;
    const int tile = blockIdx.x;
    if (tile >= tiles) return;
    if (tile == 0)
        for (int k = threadIdx.x; k < kMaxBands; k += kBlock) s.lsFlags[k] = 0;
    const int ti = tile / tilesJ, tj = tile - ti * tilesJ;
    bool known = false;
    for (int q = threadIdx.x; q < kTile * kTile; q += kBlock) {
        const int i = ti * kTile + q / kTile, j = tj * kTile + q % kTile;
        if (i >= I || j >= J) continue;
        const int c = i * J + j;
        const float f = s.fsdf[c];
        const bool k = OUT ? f < (float)((uint64_t)I * (uint64_t)J) : f > 0.f;
        s.marker[c] = k ? 0 : kMaxLevel;
        known |= k;
    }
    if (__syncthreads_or(known) && threadIdx.x == 0) {
        s.lsTile[tile] = 1;
        if (!OUT) s.sc->sdf_known = 1;   // some cell is outside
    } else if (threadIdx.x == 0) {
        s.lsTile[tile] = 0;
    }
}

// One band of extrapolateLevelsetInside: levels L0+1 .. L0+16.  Runs only if
// the previous band reached its last level somewhere (lsFlags[band - 1]);
// records the bands that ran in sc->ls_bands.  A tile has work only if a
// level-L0 cell lies within its 16-cell halo, i.e. in it or one of its eight
// neighbours (the halo is narrower than a tile), and such a tile reached
// L0 in the band before: the others only record that they did not reach
// the band's end.
template <bool OUT>
__global__ void __launch_bounds__(kBlock) k_levelset_band(const SceneDev* scenes, int band)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active) return;
    if (band > 0 && cst(s.lsFlags)[band - 1] == 0) return;   // written by the previous band's launch
    const int tilesI = (s.I + kTile - 1) / kTile, tilesJ = (s.J + kTile - 1) / kTile;
    const int tiles = tilesI * tilesJ;
    const int tile = blockIdx.x;
    if (tile >= tiles) return;
    if (tile == 0 && threadIdx.x == 0) s.sc->ls_bands = band + 1;
    const uint8_t* fin = s.lsTile + (band & 1) * tiles;
    uint8_t* fout = s.lsTile + ((band + 1) & 1) * tiles;
    const int ti = tile / tilesJ, tj = tile - ti * tilesJ;
    bool near = false;
    if (threadIdx.x < 9) {
        const int ni = ti + threadIdx.x / 3 - 1, nj = tj + threadIdx.x % 3 - 1;
        near = ni >= 0 && ni < tilesI && nj >= 0 && nj < tilesJ && fin[ni * tilesJ + nj] != 0;
    }
    if (!__syncthreads_or(near)) {
        if (threadIdx.x == 0) fout[tile] = 0;
        return;
    }
    bfs_band<16, true, OUT>(s.fsdf, nullptr, s.marker, s.fsdf, nullptr, s.marker, s.I, s.J, band * 16,
                            &s.lsFlags[band], &fout[tile]);
}

// ---------------------------------------------------------------------------
// Density correction, part 1: FlipSolver::updateDensityGridThread
// (flipsolver2d.cpp:199-249) and calcDensityCorrectionRhs (993-1011).  The
// centred bilinear hat reaches a cell from particles of the 3x3 cells around
// it only, gathered in (cell, id) order; the right-hand side goes to b.
// Materials are the pre-update ones (the step runs this before gridUpdate).
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_density_rhs(const SceneDev* scenes)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active || s.density_mode == 0) return;
    const int I = s.I, J = s.J, C = I * J;
    const int c = cell_index(gridDim.x);
    if (c >= C) return;
    const int i = c / J, j = c - (c / J) * J;
    const float fi = (float)i, fj = (float)j;
    const float mass = s.pmass;
    float d = 0.f;
    const int j0 = j > 0 ? j - 1 : 0, j1 = j < J - 1 ? j + 1 : J - 1;
    for (int di = -1; di <= 1; di++) {
        const int ii = i + di;
        if (ii < 0 || ii >= I) continue;
        const int a = s.cellStart[ii * J + j0], e = s.cellStart[ii * J + j1 + 1];
        for (int k = a; k < e; k++) {
            const float2 p = s.pos[k];
            const float w = bilinear_hat(p.x - fi - 0.5f, p.y - fj - 0.5f);
            if (fabsf(w) > 1e-6f) d += w * mass;
        }
    }
    d /= s.cell_vol;
    const int8_t m = s.mat[c];
    if (is_fluid(m) && (is_empty(mat_ext(s.mat, I, J, i + 1, j)) || is_empty(mat_ext(s.mat, I, J, i - 1, j)) ||
                        is_empty(mat_ext(s.mat, I, J, i, j + 1)) || is_empty(mat_ext(s.mat, I, J, i, j - 1))))
        d = fclamp(d, (float)s.rho, 3.40282347e+38f);
    if (is_fluid(m)) {
        double r = (double)d / s.rho;
        r = r < 0.5 ? 0.5 : (1.5 < r ? 1.5 : r);
        s.b[c] = s.dens_rhs_scale * (1.0 - r);
    } else {
        s.b[c] = 0.0;
    }
}

// ---------------------------------------------------------------------------
// Density correction, part 2: adjustParticlesByDensityThread
// (flipsolver2d.cpp:261-303) when the scene's gate is open (mode on, or
// gated and the solve reached its tolerance), then re-key for the counting
// sort: particles keep (cell, id) order for every later gather.  The
// reference's index past the last row reads outside its vector; it reads 0
// here.  A particle pushed outside the grid is dropped.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_density_adjust(const SceneDev* scenes)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active || s.density_mode == 0 || (int)blockIdx.x >= s.pblocks) return;
    const SceneScalars* sc = s.sc;
    const double bn = sqrt(sc->bb), rn = sqrt(sc->rr64);
    const bool conv = !sc->numeric_bad && (bn == 0.0 || rn <= s.tol * bn);
    const bool apply = s.density_mode == 2 || conv;
    const int n = sc->n_particles;
    const int I = s.I, J = s.J;
    const int k = blockIdx.x * kBlock + threadIdx.x;
    uint32_t key = kDead;
    if (k < n) {
        float2 p = s.pos[k];
        if (apply) {
            const float x = p.x, y = p.y;
            const int ic = iclamp((int)(x - 0.5f), 0, I), jc = iclamp((int)(y - 0.5f), 0, J);
            const int i = iclamp((int)x, 0, I), j = iclamp((int)y, 0, J);
            auto P = [&](int a, int b) -> float {
                return (a < 0 || a >= I || b < 0 || b >= J) ? 0.f : (float)s.x64[(size_t)a * J + b];
            };
            const float pci = P(ic, j), pcj = P(i, jc);
            float pi = P(ic + 1, j), pj = P(i, jc + 1);
            if (is_solid(mat_ext(s.mat, I, J, ic + 1, j)) || is_solid(mat_ext(s.mat, I, J, ic, j))) pi = pci;
            if (is_solid(mat_ext(s.mat, I, J, i, jc + 1)) || is_solid(mat_ext(s.mat, I, J, i, jc))) pj = pcj;
            p.x = x + (pi - pci) * s.dens_scale;
            p.y = y + (pj - pcj) * s.dens_scale;
        }
        // a non-finite position (NaN pressure) is dropped before the
        // conversion to a cell index, like one outside the grid
        const bool fin = isfinite(p.x) && isfinite(p.y);
        const int pi_ = fin ? integr(p.x) : -1, pj_ = fin ? integr(p.y) : -1;
        if (pi_ >= 0 && pi_ < I && pj_ >= 0 && pj_ < J) key = (uint32_t)(pi_ * J + pj_);
        s.posA[k] = p;
        s.velA[k] = s.vel[k];
        s.key[k] = key;
    }
    const int lane = threadIdx.x & 63;
    const uint32_t prev = __shfl_up(key, 1, 64);
    const bool head = lane == 0 || prev != key;
    const unsigned long long heads = __ballot(head);
    const unsigned long long above = heads & ~((lane == 63) ? ~0ull : ((2ull << lane) - 1));
    const int end = above ? (__ffsll((long long)above) - 1) : 64;
    if (head && key != kDead) atomicAdd(&s.cellCount[key], end - lane);
    if (blockIdx.x == 0 && threadIdx.x == 0) s.sc->density_applied = apply ? 1 : 0;
}

// ---------------------------------------------------------------------------
// Saved grid + body forces: m_savedFluidVelocityGrid = m_fluidVelocityGrid;
// applyBodyForces (flipsolver2d.cpp:439-440, 1222-1241).  The extrapolated
// grid (su/sv) is the saved copy; the working grid gets the force.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_forces(const SceneDev* scenes)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active) return;
    const int nu = (s.I + 1) * s.J, nv = s.I * (s.J + 1);
    const int k = blockIdx.x * kBlock + threadIdx.x;
    const float du = s.force * s.g_i, dv = s.force * s.g_j;
    if (k < nu) s.u[k] = s.su[k] + du;
    if (k < nv) s.v[k] = s.sv[k] + dv;
}

// ---------------------------------------------------------------------------
// Pressure right-hand side: FlipSolver::calcPressureRhs + divergenceAt
// (flipsolver2d.cpp:759-764, 962-991).  Divergence is formed in float and
// widened, as there; the divergence-control term is zero because
// centeredParamsToGrid clears it before this read (flipsolver2d.cpp:1382).
// Also starts the solve: x = 0, d = 0, partials of ||b||^2.  Runs on the
// solver's block/chunk layout.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_rhs(const SceneDev* scenes)
{
    const SceneDev& s = scenes[blockIdx.y];
    const int blk = solver_block();
    if (!s.active) return;
    if (blk >= s.pcg_blocks) return;
    const int I = s.I, J = s.J;
    const Chunk ch = my_chunk(s, blk);
    double bb = 0.0;
    for (int c = ch.begin + threadIdx.x; c < ch.end; c += kBlock) {
        const int i = c / J, j = c - (c / J) * J;
        double val = 0.0;
        const int8_t m = s.mat[c];
        if (is_fluid(m)) {
            const double scale = s.sRhs;
            const float u0 = s.u[(size_t)i * J + j], u1 = s.u[(size_t)(i + 1) * J + j];
            const float v0 = s.v[(size_t)i * (J + 1) + j], v1 = s.v[(size_t)i * (J + 1) + j + 1];
            const float div = (((u1 - u0) + v1) - v0) + 0.f;
            val = -scale * (double)div;
            val -= (double)is_solid(mat_ext(s.mat, I, J, i - 1, j)) * scale * (double)(u0 - 0);
            val += (double)is_solid(mat_ext(s.mat, I, J, i + 1, j)) * scale * (double)(u1 - 0);
            val -= (double)is_solid(mat_ext(s.mat, I, J, i, j - 1)) * scale * (double)(v0 - 0);
            val += (double)is_solid(mat_ext(s.mat, I, J, i, j + 1)) * scale * (double)(v1 - 0);
        }
        s.b[c] = val;
        // with the V-cycle the solve starts from the previous substep's
        // pressure on fluid cells (zero elsewhere; k_cg_init falls back to
        // zero where that start is no better): the pressure changes little
        // between substeps, so the first residual is a fraction of ||b||.
        // With the density correction on, x holds this substep's density
        // solve instead, so the pressure solve starts from zero.
        if (!is_fluid(m) || s.precond != kPrecondMG || s.density_mode != 0) s.x64[c] = 0.0;
        bb += val * val;
    }
    double v[1] = {bb};
    put_partials<1>(s, kSlotBB, v, blk);
}

// ---------------------------------------------------------------------------
// Pressure application + solid friction: applyPressureThreadU/V
// (flipsolver2d.cpp:1133-1193, FP64 update of the float face values) then
// updateVelocityFromSolids (1077-1104, materialgrid.cpp:152-166).  Both
// touch faces U(i,j), V(i,j) of cell (i,j) only, so they fuse per cell.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) k_apply_solids(const SceneDev* scenes, int friction)
{
    const SceneDev& s = scenes[blockIdx.y];
    if (!s.active) return;
    const int I = s.I, J = s.J, C = I * J;
    const int c = cell_index(gridDim.x);
    if (c >= C) return;
    const int i = c / J, j = c - (c / J) * J;
    const double scale = s.sApply;
    const size_t fu = (size_t)i * J + j, fv = (size_t)i * (J + 1) + j;
    const double pc = s.x64[c];
    const int8_t mb = s.mat[c];
    float u = s.u[fu], v = s.v[fv];
    {
        const double pm = i - 1 < 0 ? 0.0 : s.x64[c - J];
        const int8_t ma = mat_ext(s.mat, I, J, i - 1, j);
        if (is_fluid(ma) || is_fluid(mb)) {
            if (is_solid(ma) || is_solid(mb)) u = 0.f;
            else u = (float)((double)u - scale * (pc - pm));
        } else {
            s.uvalS[fu] = 0;
        }
    }
    {
        const double pm = j - 1 < 0 ? 0.0 : s.x64[c - 1];
        const int8_t ma = mat_ext(s.mat, I, J, i, j - 1);
        if (is_fluid(ma) || is_fluid(mb)) {
            if (is_solid(ma) || is_solid(mb)) v = 0.f;
            else v = (float)((double)v - scale * (pc - pm));
        } else {
            s.vvalS[fv] = 0;
        }
    }
    // friction of the solids in the 3x3 neighbourhood; with no nonzero
    // friction the factor (1 - avg) is exactly 1 and the loads are skipped
    int n = 0;
    float avg = 0;
    if (friction && s.friction_any) {
        for (int di = -1; di < 2; di++)
            for (int dj = -1; dj < 2; dj++) {
                const int ii = i + di, jj = j + dj;
                if (ii >= 0 && ii < I && jj >= 0 && jj < J) {
                    const int id = s.sid[(size_t)ii * J + jj];
                    if (id != -1) {
                        avg += (id < s.n_friction) ? s.friction[id] : 0.f;
                        n++;
                    }
                }
            }
    }
    if (n > 0) {
        avg /= (float)n;
        auto sol = [&](int a, int b) { return is_solid(mat_ext(s.mat, I, J, a, b)); };
        if (sol(i, j + 1) || sol(i - 1, j + 1) || sol(i, j) || sol(i - 1, j) || sol(i, j - 1) || sol(i - 1, j - 1))
            u = u * (1 - avg);
        if (sol(i + 1, j) || sol(i + 1, j - 1) || sol(i, j) || sol(i, j - 1) || sol(i - 1, j) || sol(i - 1, j - 1))
            v = v * (1 - avg);
    }
    s.u[fu] = u;
    s.v[fv] = v;
}

}  // namespace fg

namespace fg {
#define FG_LAUNCH0(name) \
    void launch_##name(dim3 g, hipStream_t st, const SceneDev* sc) { fg_launch(k_##name, g, dim3(kBlock), st, sc); }
#define FG_LAUNCH1(name, T, a) \
    void launch_##name(dim3 g, hipStream_t st, const SceneDev* sc, T a) { fg_launch(k_##name, g, dim3(kBlock), st, sc, a); }
FG_LAUNCH0(stencil)
FG_LAUNCH0(p2g)
FG_LAUNCH0(sdf_materials)
FG_LAUNCH0(sdf_exact)
FG_LAUNCH1(extrapolate, int, which)
void launch_levelset_init(dim3 g, hipStream_t st, const SceneDev* sc, bool outside)
{
    if (outside) fg_launch(k_levelset_init<true>, g, dim3(kBlock), st, sc);
    else fg_launch(k_levelset_init<false>, g, dim3(kBlock), st, sc);
}
void launch_levelset_band(dim3 g, hipStream_t st, const SceneDev* sc, int band, bool outside)
{
    if (outside) fg_launch(k_levelset_band<true>, g, dim3(kBlock), st, sc, band);
    else fg_launch(k_levelset_band<false>, g, dim3(kBlock), st, sc, band);
}
FG_LAUNCH0(forces)
FG_LAUNCH0(rhs)
FG_LAUNCH1(apply_solids, int, friction)
FG_LAUNCH0(density_rhs)
FG_LAUNCH0(density_adjust)
}  // namespace fg
