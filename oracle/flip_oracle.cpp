/*
 * flip_oracle.cpp -- CPU restatement of FlipSolver::step() (reference
 * FlipSolver2dLib/flipsolver2d.cpp:412-462) and the functions it calls.
 *
 * TEST INFRASTRUCTURE ONLY (see flip_oracle.h).  Compiled IEEE-strict:
 * g++ -O2 -ffp-contract=off, no -ffast-math.
 *
 * Citations are "file:line" into the reference's FlipSolver2dLib/.
 */
#include "flip_oracle.h"

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <functional>
#include <limits>
#include <mutex>
#include <numeric>
#include <queue>
#include <thread>
#include <vector>

namespace {

/* ------------------------------------------------------------------ */
/* Small helpers restating mathfuncs.cpp / grid2d.h                     */
/* ------------------------------------------------------------------ */

/* mathfuncs.cpp:3-6  frac(v) = v - (long)v */
inline float frac_(float v) { return v - static_cast<float>(static_cast<long>(v)); }
/* mathfuncs.cpp:8-11 integr(v) = (int)floor(v) */
inline long integr(float v) { return static_cast<long>(static_cast<int>(std::floor(v))); }
/* mathfuncs.cpp:13-16 */
inline float lerp1(float a, float b, float f) { return (a * (1.0f - f)) + (b * f); }

inline bool is_fluid(int8_t m) { return (m & ORC_FLUID) != 0; }    /* materialgrid.h:15 */
inline bool is_strict_fluid(int8_t m) { return m == (int8_t)ORC_FLUID; }
inline bool is_empty(int8_t m) { return (m & ORC_EMPTY) != 0; }    /* materialgrid.h:17 */
inline bool is_solid(int8_t m) { return (m & ORC_SOLID) != 0; }
inline bool is_source(int8_t m) { return m == (int8_t)ORC_SOURCE; }
inline bool is_sink(int8_t m) { return m == (int8_t)ORC_SINK; }

/* Grid2d::getAt with OOB_EXTEND, grid2d.h:122-143 */
inline float get_ext(const float* g, long si, long sj, long i, long j)
{
    i = std::clamp<long>(i, 0, si - 1);
    j = std::clamp<long>(j, 0, sj - 1);
    return g[i * sj + j];
}
inline int8_t mat_ext(const int8_t* m, long I, long J, long i, long j)
{
    i = std::clamp<long>(i, 0, I - 1);
    j = std::clamp<long>(j, 0, J - 1);
    return m[i * J + j];
}

/* Grid2d::lerp, grid2d.h:186-216 (also simmath::lerpCenteredGrid,
 * mathfuncs.cpp:84-106, which is the same with a zero offset). */
float lerp_grid(const float* g, long si, long sj, float oi, float oj, float i, float j)
{
    i += oi;
    j += oj;
    i = std::clamp(i, 0.f, static_cast<float>(si - 1));
    j = std::clamp(j, 0.f, static_cast<float>(sj - 1));
    const long ci = integr(i), cj = integr(j);
    const float fi = frac_(i), fj = frac_(j);
    const long c2i = ci, c2j = fj >= 0.5f ? cj + 1 : cj - 1;
    const long c3i = fi >= 0.5f ? ci + 1 : ci - 1, c3j = fj >= 0.5f ? cj + 1 : cj - 1;
    const long c4i = fi >= 0.5f ? ci + 1 : ci - 1, c4j = cj;
    const float iF = fi < 0.5f ? 0.5f - fi : fi - 0.5f;
    const float jF = fj < 0.5f ? 0.5f - fj : fj - 0.5f;
    const float v1 = lerp1(get_ext(g, si, sj, ci, cj), get_ext(g, si, sj, c4i, c4j), iF);
    const float v2 = lerp1(get_ext(g, si, sj, c2i, c2j), get_ext(g, si, sj, c3i, c3j), iF);
    return lerp1(v1, v2, jF);
}

struct V2 { float x, y; };

/* StaggeredVelocityGrid::velocityAt, staggeredvelocitygrid.cpp:158-167; U is
 * (I+1)xJ with offset (0.5,0), V is Ix(J+1) with offset (0,0.5)
 * (staggeredvelocitygrid.cpp:6-14). */
inline V2 velocity_at(const float* u, const float* v, long I, long J, float x, float y)
{
    return V2{lerp_grid(u, I + 1, J, 0.5f, 0.f, x, y), lerp_grid(v, I, J + 1, 0.f, 0.5f, x, y)};
}

/* FlipSolver::rk4Integrate, flipsolver2d.cpp:1195-1203, with Vec3 operator
 * semantics of geometry2d.cpp:3-27 (scalar*vec multiplies each component). */
V2 rk4(const float* u, const float* v, long I, long J, V2 p, float dt)
{
    V2 a = velocity_at(u, v, I, J, p.x, p.y);
    const V2 k1{a.x * dt, a.y * dt};
    a = velocity_at(u, v, I, J, p.x + k1.x * 0.5f, p.y + k1.y * 0.5f);
    const V2 k2{a.x * dt, a.y * dt};
    a = velocity_at(u, v, I, J, p.x + k2.x * 0.5f, p.y + k2.y * 0.5f);
    const V2 k3{a.x * dt, a.y * dt};
    a = velocity_at(u, v, I, J, p.x + k3.x, p.y + k3.y);
    const V2 k4{a.x * dt, a.y * dt};
    const float sx = ((k1.x + k2.x * 2.f) + k3.x * 2.f) + k4.x;
    const float sy = ((k1.y + k2.y * 2.f) + k3.y * 2.f) + k4.y;
    const float sixth = 1.0f / 6.0f;
    return V2{p.x + sx * sixth, p.y + sy * sixth};
}

/* simmath::gradCenteredGrid called with float positions: the (ssize_t,
 * ssize_t) parameters truncate (mathfuncs.cpp:48-59, sdfgrid.cpp:20,33). */
inline V2 grad_trunc(const float* g, long I, long J, float x, float y)
{
    const long i = static_cast<long>(x), j = static_cast<long>(y);
    const float ip1 = get_ext(g, I, J, i + 1, j), im1 = get_ext(g, I, J, i - 1, j);
    const float jp1 = get_ext(g, I, J, i, j + 1), jm1 = get_ext(g, I, J, i, j - 1);
    return V2{(ip1 - im1) / 2.f, (jp1 - jm1) / 2.f};
}

/* SdfGrid::closestSurfacePoint, sdfgrid.cpp:15-46 */
V2 closest_surface_point(const float* sdf, long I, long J, V2 pos)
{
    V2 cp = pos;
    float value = lerp_grid(sdf, I, J, 0.f, 0.f, pos.x, pos.y);
    V2 grad = grad_trunc(sdf, I, J, pos.x, pos.y);
    for (int it = 0; it < 100; it++) {
        float alpha = 1;
        for (int k = 0; k < 10; k++) {
            const float t = alpha * value;
            const V2 q{cp.x - grad.x * t, cp.y - grad.y * t};
            const float vq = lerp_grid(sdf, I, J, 0.f, 0.f, q.x, q.y);
            if (std::abs(vq) < std::abs(value)) {
                cp = q;
                value = lerp_grid(sdf, I, J, 0.f, 0.f, q.x, q.y);
                grad = grad_trunc(sdf, I, J, q.x, q.y);
                if (std::abs(value) < 1e-5f) return cp;
            } else {
                alpha *= 0.7f;
            }
        }
    }
    return cp;
}

/* simmath::bSpline / quadraticBSpline, mathfuncs.cpp:32-43 */
inline float bspline(float value)
{
    value = std::abs(value);
    const float t1 = (0.75f - value * value) * static_cast<float>(value < 0.5f);
    const float t2 = ((0.5f * (1.5f - value)) * (1.5f - value)) *
                     static_cast<float>(value >= 0.5f && value < 1.5f);
    return t1 + t2;
}
inline float qbspline(float x, float y) { return (bspline(x) * bspline(y)) * bspline(0.f); }

/* simmath::linearHat / bilinearHat, mathfuncs.cpp:61-77 */
inline float linear_hat(float v)
{
    if (v >= 0 && v <= 1) return 1 - v;
    if (v < 0 && v >= -1) return 1 + v;
    return 0;
}
inline float bilinear_hat(float x, float y) { return (linear_hat(x) * linear_hat(y)) * linear_hat(0.f); }

/* ------------------------------------------------------------------ */
/* splitRange (threading/threadpool.cpp:41-76) and a tiny pool           */
/* ------------------------------------------------------------------ */
struct Range { int64_t start, end; };

std::vector<Range> split_range(int64_t length, int64_t min_size, int64_t jobs, int threads)
{
    std::vector<Range> parts;
    if (length == 0) return parts;
    int64_t num = (int64_t)threads * jobs;
    if (length < num) num = length;
    int64_t part = (int64_t)std::floor((double)(length / num));
    part = std::max(min_size, part);
    int64_t rem = 0;
    if (part * num < length) rem = length - part * num;
    int64_t cs = 0, ce = std::min(length, part);
    for (int64_t i = 0; i < num; i++) {
        if (i < rem) ce++;
        parts.push_back(Range{cs, ce});
        cs = ce;
        ce = std::min(cs + part, length);
        if (ce == cs) break;
    }
    return parts;
}

/* Fan-out/barrier pool used for the CPU-baseline timing (the reference's
 * ThreadPool::enqueue + wait, threadpool.h:42-59, threadpool.cpp:78-92).
 * Results never depend on it: every parallel site writes disjoint outputs
 * and reductions merge per-range partials in range order. */
class Pool {
public:
    static Pool& get(int threads)
    {
        static Pool* inst = nullptr;
        static std::mutex m;
        std::lock_guard<std::mutex> g(m);
        if (!inst || inst->n_ != threads) {
            delete inst;
            inst = new Pool(threads);
        }
        return *inst;
    }
    void run(const std::vector<Range>& rs, const std::function<void(size_t, Range)>& f)
    {
        if (n_ <= 1 || rs.size() <= 1) {
            for (size_t k = 0; k < rs.size(); k++) f(k, rs[k]);
            return;
        }
        uint64_t g;
        {
            std::unique_lock<std::mutex> l(mu_);
            job_ = &f;
            ranges_ = &rs;
            next_ = 0;
            pending_ = rs.size();
            g = ++gen_;
        }
        cv_.notify_all();
        work(g);
        std::unique_lock<std::mutex> l(mu_);
        done_.wait(l, [&] { return pending_ == 0; });
        job_ = nullptr;
        ranges_ = nullptr;
    }
    ~Pool()
    {
        {
            std::unique_lock<std::mutex> l(mu_);
            stop_ = true;
            gen_++;
        }
        cv_.notify_all();
        for (auto& t : th_) t.join();
    }

private:
    explicit Pool(int n) : n_(n)
    {
        for (int i = 1; i < n; i++) th_.emplace_back([this] { loop(); });
    }
    /* Hands out job indices of generation g only; a worker that wakes late
     * for a finished generation finds gen_ moved on and returns. */
    void work(uint64_t g)
    {
        for (;;) {
            size_t k;
            const std::function<void(size_t, Range)>* job;
            const std::vector<Range>* rs;
            {
                std::unique_lock<std::mutex> l(mu_);
                if (gen_ != g || !job_ || next_ >= ranges_->size()) return;
                k = next_++;
                job = job_;
                rs = ranges_;
            }
            (*job)(k, (*rs)[k]);
            std::unique_lock<std::mutex> l(mu_);
            if (--pending_ == 0) done_.notify_all();
        }
    }
    void loop()
    {
        uint64_t seen = 0;
        for (;;) {
            uint64_t g;
            {
                std::unique_lock<std::mutex> l(mu_);
                cv_.wait(l, [&] { return stop_ || gen_ != seen; });
                if (stop_) return;
                seen = g = gen_;
            }
            work(g);
        }
    }
    int n_;
    std::vector<std::thread> th_;
    std::mutex mu_;
    std::condition_variable cv_, done_;
    const std::function<void(size_t, Range)>* job_ = nullptr;
    const std::vector<Range>* ranges_ = nullptr;
    size_t next_ = 0;
    size_t pending_ = 0;
    uint64_t gen_ = 0;
    bool stop_ = false;
};

void parallel_for(int threads, int64_t n, int64_t jobs, const std::function<void(int64_t, int64_t)>& f)
{
    const int t = std::max(1, threads);
    auto rs = split_range(n, 1, jobs, t);
    Pool::get(t).run(rs, [&](size_t, Range r) { f(r.start, r.end); });
}

inline long cell_of(float x, float y, long J) { return integr(x) * J + integr(y); }

double now_ms()
{
    return std::chrono::duration<double, std::milli>(
               std::chrono::steady_clock::now().time_since_epoch()).count();
}

} // namespace

/* ================================================================== */
/* Exported point functions                                              */
/* ================================================================== */
extern "C" {

float orc_bspline(float v) { return bspline(v); }
float orc_quadratic_bspline(float x, float y) { return qbspline(x, y); }
float orc_linear_hat(float v) { return linear_hat(v); }

void orc_lerp(const float* g, int32_t si, int32_t sj, float off_i, float off_j,
              const float* xi, const float* xj, int64_t n, float* out)
{
    for (int64_t k = 0; k < n; k++) out[k] = lerp_grid(g, si, sj, off_i, off_j, xi[k], xj[k]);
}

void orc_velocity_at(const float* u, const float* v, int32_t I, int32_t J,
                     const float* xi, const float* xj, int64_t n, float* ou, float* ov)
{
    for (int64_t k = 0; k < n; k++) {
        const V2 a = velocity_at(u, v, I, J, xi[k], xj[k]);
        ou[k] = a.x;
        ov[k] = a.y;
    }
}

void orc_rk4(const float* u, const float* v, int32_t I, int32_t J, float dt,
             const float* xi, const float* xj, int64_t n, float* oi, float* oj)
{
    for (int64_t k = 0; k < n; k++) {
        const V2 r = rk4(u, v, I, J, V2{xi[k], xj[k]}, dt);
        oi[k] = r.x;
        oj[k] = r.y;
    }
}

void orc_closest_surface_point(const float* sdf, int32_t I, int32_t J,
                               const float* xi, const float* xj, int64_t n, float* oi, float* oj)
{
    for (int64_t k = 0; k < n; k++) {
        const V2 r = closest_surface_point(sdf, I, J, V2{xi[k], xj[k]});
        oi[k] = r.x;
        oj[k] = r.y;
    }
}

/* Geometry2d::signedDistance, geometry2d.cpp:54-73 (iq polygon SDF).  Vec3
 * dot products carry the zero z term; sqrt is taken in double as the
 * unqualified sqrt(float) there resolves to ::sqrt(double). */
float orc_polygon_sdf(const float* vt, int32_t N, float px, float py)
{
    auto dot = [](float ax, float ay, float bx, float by) { return (ax * bx + ay * by) + 0.f * 0.f; };
    float dx0 = px - vt[0], dy0 = py - vt[1];
    float sq = dot(dx0, dy0, dx0, dy0);
    float s = 1.0f;
    for (int32_t i = 0, j = N - 1; i < N; j = i, i++) {
        const float ex = vt[2 * j] - vt[2 * i], ey = vt[2 * j + 1] - vt[2 * i + 1];
        const float wx = px - vt[2 * i], wy = py - vt[2 * i + 1];
        const float h = std::clamp(dot(wx, wy, ex, ey) / dot(ex, ey, ex, ey), 0.0f, 1.0f);
        const float bx = wx - ex * h, by = wy - ey * h;
        sq = std::min(sq, dot(bx, by, bx, by));
        const bool xf = py >= vt[2 * i + 1];
        const bool yf = py < vt[2 * j + 1];
        const bool zf = ex * wy > ey * wx;
        if ((xf && yf && zf) || (!xf && !yf && !zf)) s *= -1.0f;
    }
    return static_cast<float>(s * std::sqrt(static_cast<double>(sq)));
}

int32_t orc_material_bits(int8_t m)
{
    return (is_fluid(m) ? 1 : 0) | (is_strict_fluid(m) ? 2 : 0) | (is_empty(m) ? 4 : 0) | (is_solid(m) ? 8 : 0) |
           (is_source(m) ? 16 : 0) | (is_sink(m) ? 32 : 0);
}

double orc_maxabs_replica(const double* v, int64_t n, int32_t threads)
{
    double mx = 0;
    for (const Range& r : split_range(n, 1, 1, threads)) {
        double ma = std::numeric_limits<double>::min();
        for (int64_t i = r.start; i < r.end; i++)
            if (std::abs(v[i]) > 0.0) ma = std::abs(v[i]);
        if (ma > mx) mx = std::abs(ma);
    }
    return mx;
}

double orc_dot_replica(const double* a, const double* b, int64_t n, int32_t threads)
{
    double res = 0;
    for (const Range& r : split_range(n, 1, 1, threads)) {
        double part = 0.0;
        for (int64_t i = r.start; i < r.end; i++) part += a[i] * b[i];
        res += part;
    }
    return res;
}

int32_t orc_split_range(int64_t length, int64_t min_size, int64_t jobs, int32_t threads,
                        int64_t* starts, int64_t* ends)
{
    auto rs = split_range(length, min_size, jobs, threads);
    for (size_t k = 0; k < rs.size(); k++) {
        starts[k] = rs[k].start;
        ends[k] = rs[k].end;
    }
    return (int32_t)rs.size();
}

/* ================================================================== */
/* Particles                                                             */
/* ================================================================== */

/* FlipSolver::advectThread, flipsolver2d.cpp:305-338.  Rebinning (the
 * RebinRecord bookkeeping) is replaced by the later sort on cell key. */
void orc_advect(const OrcParams* p, OrcState* s, float dt, uint8_t* alive)
{
    const long I = p->size_i, J = p->size_j;
    parallel_for(p->threads, s->n_particles, 1, [&](int64_t a, int64_t b) {
        for (int64_t k = a; k < b; k++) {
            V2 pos = rk4(s->u, s->v, I, J, V2{s->px[k], s->py[k]}, dt);
            if (lerp_grid(s->solid_sdf, I, J, 0.f, 0.f, pos.x, pos.y) < 0.f)
                pos = closest_surface_point(s->solid_sdf, I, J, pos);
            s->px[k] = pos.x;
            s->py[k] = pos.y;
            const long pi = integr(pos.x), pj = integr(pos.y);
            const bool inb = pi >= 0 && pi < I && pj >= 0 && pj < J;
            alive[k] = (!inb || is_sink(mat_ext(s->material, I, J, pi, pj))) ? 0 : 1;
        }
    });
}

/* Prune dead particles (flipsolver2d.cpp:353-359) and order the survivors by
 * (cell, stable id, input index).  cell_start gets I*J+1 offsets. */
void orc_sort_particles(const OrcParams* p, OrcState* s, const uint8_t* alive, int32_t* cell_start)
{
    const long J = p->size_j;
    const int64_t C = (int64_t)p->size_i * p->size_j;
    std::vector<int64_t> idx;
    idx.reserve(s->n_particles);
    for (int64_t k = 0; k < s->n_particles; k++) {
        if (alive && !alive[k]) continue;
        const long ci = integr(s->px[k]), cj = integr(s->py[k]);
        if (ci < 0 || ci >= p->size_i || cj < 0 || cj >= J) continue; /* moved out by density adjust */
        idx.push_back(k);
    }
    std::vector<long> key(s->n_particles);
    for (int64_t k : idx) key[k] = cell_of(s->px[k], s->py[k], J);
    std::sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) {
        if (key[a] != key[b]) return key[a] < key[b];
        if (s->pid[a] != s->pid[b]) return s->pid[a] < s->pid[b];
        return a < b; /* a caller may repeat an id: input order breaks the tie */
    });
    const int64_t n = (int64_t)idx.size();
    std::vector<float> px(n), py(n), vx(n), vy(n), pv(s->pvisc ? n : 0);
    std::vector<uint32_t> id(n);
    for (int64_t k = 0; k < n; k++) {
        const int64_t q = idx[k];
        px[k] = s->px[q]; py[k] = s->py[q]; vx[k] = s->vx[q]; vy[k] = s->vy[q]; id[k] = s->pid[q];
        if (s->pvisc) pv[k] = s->pvisc[q];
    }
    std::memcpy(s->px, px.data(), n * 4); std::memcpy(s->py, py.data(), n * 4);
    std::memcpy(s->vx, vx.data(), n * 4); std::memcpy(s->vy, vy.data(), n * 4);
    std::memcpy(s->pid, id.data(), n * 4);
    if (s->pvisc) std::memcpy(s->pvisc, pv.data(), n * 4);
    s->n_particles = n;
    if (cell_start) {
        std::fill(cell_start, cell_start + C + 1, 0);
        for (int64_t k = 0; k < n; k++) cell_start[cell_of(s->px[k], s->py[k], J) + 1]++;
        for (int64_t c = 0; c < C; c++) cell_start[c + 1] += cell_start[c];
    }
}

void orc_cell_start(const OrcParams* p, const OrcState* s, int32_t* cs)
{
    const long J = p->size_j;
    const int64_t C = (int64_t)p->size_i * p->size_j;
    std::fill(cs, cs + C + 1, 0);
    for (int64_t k = 0; k < s->n_particles; k++) cs[cell_of(s->px[k], s->py[k], J) + 1]++;
    for (int64_t c = 0; c < C; c++) cs[c + 1] += cs[c];
}

/* FlipSolver::maxParticleVelocity, flipsolver2d.cpp:1560-1574 */
float orc_max_particle_velocity(const OrcState* s)
{
    float m = std::numeric_limits<float>::min();
    for (int64_t k = 0; k < s->n_particles; k++) {
        const float q = s->vx[k] * s->vx[k] + s->vy[k] * s->vy[k];
        if (q > m) m = q;
    }
    return std::sqrt(m);
}

/* FlipSolver::countParticles, flipsolver2d.cpp:1021-1051: per cell, particles
 * beyond 2*ppc are erased; survivors are the smallest ids of the cell. */
int32_t orc_count_particles(const OrcParams* p, OrcState* s, const int32_t* cs)
{
    const int64_t C = (int64_t)p->size_i * p->size_j;
    const int32_t cap = 2 * p->particles_per_cell;
    std::vector<uint8_t> keep(s->n_particles, 1);
    int32_t erased = 0;
    for (int64_t c = 0; c < C; c++) {
        const int32_t n = cs[c + 1] - cs[c];
        s->particle_counts[c] = std::min(n, cap);
        for (int32_t r = cap; r < n; r++) {
            keep[cs[c] + r] = 0;
            erased++;
        }
    }
    if (erased) {
        int64_t w = 0;
        for (int64_t k = 0; k < s->n_particles; k++) {
            if (!keep[k]) continue;
            s->px[w] = s->px[k]; s->py[w] = s->py[k]; s->vx[w] = s->vx[k]; s->vy[w] = s->vy[k];
            s->pid[w] = s->pid[k];
            if (s->pvisc) s->pvisc[w] = s->pvisc[k];
            w++;
        }
        s->n_particles = w;
    }
    return erased;
}

/* FlipSolver::particleUpdate, flipsolver2d.cpp:361-388 */
void orc_g2p(const OrcParams* p, OrcState* s)
{
    const long I = p->size_i, J = p->size_j;
    const float pic = p->pic_ratio;
    parallel_for(1, s->n_particles, 1, [&](int64_t a, int64_t b) {
        for (int64_t k = a; k < b; k++) {
            const float x = s->px[k], y = s->py[k];
            const float ou = lerp_grid(s->saved_u, I + 1, J, 0.5f, 0.f, x, y);
            const float ov = lerp_grid(s->saved_v, I, J + 1, 0.f, 0.5f, x, y);
            const V2 nv = velocity_at(s->u, s->v, I, J, x, y);
            const float ax = nv.x * pic, ay = nv.y * pic;
            const float bx = ((s->vx[k] + nv.x) - ou) * (1.f - pic);
            const float by = ((s->vy[k] + nv.y) - ov) * (1.f - pic);
            s->vx[k] = ax + bx;
            s->vy[k] = ay + by;
        }
    });
}

/* ================================================================== */
/* Grid transfer and rebuild                                             */
/* ================================================================== */

/* FlipSolver::particleVelocityToGrid[Thread], flipsolver2d.cpp:1313-1377.
 * Only particles of the 3x3 cells around (i,j) can pass the dual 1e-9
 * weight gate, so the 9-bin window reduces to those cells; sums run over
 * rows i-1..i+1, each a contiguous (cell, id)-ordered range. */
void orc_p2g(const OrcParams* p, OrcState* s, const int32_t* cs)
{
    const long I = p->size_i, J = p->size_j;
    std::fill(s->u, s->u + (I + 1) * J, 0.f);
    std::fill(s->v, s->v + I * (J + 1), 0.f);
    std::fill(s->u_valid, s->u_valid + (I + 1) * J, 0);
    std::fill(s->v_valid, s->v_valid + I * (J + 1), 0);
    parallel_for(p->threads, I * J, 16, [&](int64_t a, int64_t b) {
        for (int64_t c = a; c < b; c++) {
            const long i = c / J, j = c % J;
            float uw = 1e-10f, vw = 1e-10f, uv = 0.0f, vv = 0.0f;
            bool uval = false, vval = false;
            for (long di = -1; di <= 1; di++) {
                const long ii = i + di;
                if (ii < 0 || ii >= I) continue;
                const long j0 = std::max<long>(j - 1, 0), j1 = std::min<long>(j + 1, J - 1);
                for (int32_t k = cs[ii * J + j0]; k < cs[ii * J + j1 + 1]; k++) {
                    const float x = s->px[k], y = s->py[k];
                    const float wu = qbspline(x - static_cast<float>(i), y - (static_cast<float>(j) + 0.5f));
                    const float wv = qbspline(x - (static_cast<float>(i) + 0.5f), y - static_cast<float>(j));
                    if (wu > 1e-9f && wv > 1e-9f) {
                        uw += wu;
                        uv += wu * s->vx[k];
                        uval = true;
                        vw += wv;
                        vv += wv * s->vy[k];
                        vval = true;
                    }
                }
            }
            s->u[i * J + j] = uv / uw;
            s->u_valid[i * J + j] = uval;
            s->v[i * (J + 1) + j] = vv / vw;
            s->v_valid[i * (J + 1) + j] = vval;
        }
    });
}

/* FlipSolver::updateSdfThread, flipsolver2d.cpp:1276-1311.  The window is the
 * reference's: the 3x3 bins (3x3 cells each, markerparticlesystem.cpp:9-20,
 * 109-128) around the cell's bin. */
void orc_fluid_sdf(const OrcParams* p, OrcState* s, const int32_t* cs)
{
    const long I = p->size_i, J = p->size_j;
    const float radius = p->particle_scale;
    parallel_for(p->threads, I * J, 1, [&](int64_t a, int64_t b) {
        for (int64_t c = a; c < b; c++) {
            const long i = c / J, j = c % J;
            const long bi = i / 3, bj = j / 3;
            const long r0 = std::max<long>(0, 3 * bi - 3), r1 = std::min<long>(I, 3 * bi + 6);
            const long c0 = std::max<long>(0, 3 * bj - 3), c1 = std::min<long>(J, 3 * bj + 6);
            float d2 = std::numeric_limits<float>::max();
            const float cx = static_cast<float>(i) + 0.5f, cy = static_cast<float>(j) + 0.5f;
            for (long r = r0; r < r1; r++) {
                for (int32_t k = cs[r * J + c0]; k < cs[r * J + c1]; k++) {
                    const float dxp = s->px[k] - cx, dyp = s->py[k] - cy;
                    const float nd = dxp * dxp + dyp * dyp;
                    if (nd < d2) d2 = nd;
                }
            }
            s->fluid_sdf[c] = std::sqrt(d2) - radius;
        }
    });
}

/* FlipSolver::updateMaterials, flipsolver2d.cpp:1053-1075 */
void orc_update_materials(const OrcParams* p, OrcState* s)
{
    const int64_t C = (int64_t)p->size_i * p->size_j;
    for (int64_t c = 0; c < C; c++) {
        if (s->fluid_sdf[c] < 0.f) {
            if (is_empty(s->material[c])) s->material[c] = (int8_t)ORC_FLUID;
        } else {
            if (is_strict_fluid(s->material[c])) s->material[c] = (int8_t)ORC_EMPTY;
        }
    }
}

/* LinearIndexable2d::getNeighborhood order, linearindexable2d.h:55-71 */
static inline void neighborhood(long si, long sj, long idx, long out[8])
{
    const long i = idx / sj, j = idx - (idx / sj) * sj;
    int o = 0;
    for (long di = -1; di <= 1; di++)
        for (long dj = -1; dj <= 1; dj++) {
            if (di == 0 && dj == 0) continue;
            const long ii = i + di, jj = j + dj;
            out[o++] = (ii < 0 || ii >= si || jj < 0 || jj >= sj) ? -1 : ii * sj + jj;
        }
}

/* FlipSolver::extrapolateLevelsetInside / extrapolateLevelsetOutside,
 * flipsolver2d.cpp:1433-1494 / 1496-1557 (FIFO BFS).  They differ only in
 * the known set (sdf > 0 inside; sdf < I*J as a float outside) and in the
 * step each level adds (-1 inside, +1 outside). */
static void levelset_bfs(float* g, int32_t I, int32_t J, bool outside)
{
    const long C = (long)I * J;
    const int MAXI = std::numeric_limits<int>::max();
    const float max_sdf = static_cast<float>(static_cast<size_t>(I) * static_cast<size_t>(J));
    std::vector<int> mk(C, MAXI);
    std::queue<long> wf;
    for (long c = 0; c < C; c++)
        if (outside ? g[c] < max_sdf : g[c] > 0.f) mk[c] = 0;
    long nb[8];
    for (long c = 0; c < C; c++) {
        if (mk[c] != 0) {
            neighborhood(I, J, c, nb);
            for (long q : nb) {
                if (q == -1) continue;
                if (mk[q] == 0) {
                    mk[c] = 1;
                    wf.push(c);
                    break;
                }
            }
        }
    }
    while (!wf.empty()) {
        const long c = wf.front();
        neighborhood(I, J, c, nb);
        double avg = 0;
        int count = 0;
        for (long q : nb) {
            if (q == -1) continue;
            if (mk[q] < mk[c]) {
                avg += g[q];
                count++;
            }
            if (mk[q] == MAXI && mk[c] <= 1e6) {
                mk[q] = mk[c] + 1;
                wf.push(q);
            }
        }
        g[c] = outside ? static_cast<float>(avg / count + 1.f) : static_cast<float>(avg / count - 1.f);
        wf.pop();
    }
}

void orc_extrapolate_levelset_inside(float* g, int32_t I, int32_t J) { levelset_bfs(g, I, J, false); }
void orc_extrapolate_levelset_outside(float* g, int32_t I, int32_t J) { levelset_bfs(g, I, J, true); }

/* simmath::breadthFirstExtrapolate, mathfuncs.cpp:152-215 (FIFO BFS) */
void orc_extrapolate(float* g, uint8_t* valid, int32_t si, int32_t sj, int32_t radius)
{
    const long C = (long)si * sj;
    const int MAXI = std::numeric_limits<int>::max();
    std::vector<int> mk(C, MAXI);
    std::queue<long> wf;
    for (long c = 0; c < C; c++)
        if (valid[c]) mk[c] = 0;
    long nb[8];
    for (long c = 0; c < C; c++) {
        if (mk[c] != 0) {
            neighborhood(si, sj, c, nb);
            for (long q : nb) {
                if (q == -1) continue;
                if (mk[q] == 0) {
                    mk[c] = 1;
                    wf.push(c);
                    break;
                }
            }
        }
    }
    while (!wf.empty()) {
        const long c = wf.front();
        neighborhood(si, sj, c, nb);
        double avg = 0;
        int count = 0;
        for (long q : nb) {
            if (q == -1) continue;
            if (mk[q] < mk[c]) {
                avg += g[q];
                count++;
            }
            if (mk[q] == MAXI && mk[c] <= radius) {
                mk[q] = mk[c] + 1;
                wf.push(q);
            }
        }
        g[c] = static_cast<float>(avg / count);
        valid[c] = 1;
        wf.pop();
    }
}

/* FlipSolver::applyBodyForces, flipsolver2d.cpp:1222-1241 */
void orc_body_forces(const OrcParams* p, OrcState* s, float dt)
{
    const long I = p->size_i, J = p->size_j;
    const float factor = static_cast<float>(dt / p->dx);
    const float du = factor * p->gravity_i, dv = factor * p->gravity_j;
    for (long k = 0; k < (I + 1) * J; k++) s->u[k] += du;
    for (long k = 0; k < I * (J + 1); k++) s->v[k] += dv;
}

/* FlipSolver::calcPressureRhs + divergenceAt, flipsolver2d.cpp:759-764,
 * 962-991.  The divergence-control term is zero: centeredParamsToGrid
 * clears it every substep before this read (flipsolver2d.cpp:1382). */
void orc_pressure_rhs(const OrcParams* p, const OrcState* s, double* rhs)
{
    const long I = p->size_i, J = p->size_j;
    const double scale = 1.f / p->dx;
    for (long i = 0; i < I; i++)
        for (long j = 0; j < J; j++) {
            const int8_t* m = s->material;
            if (is_fluid(m[i * J + j])) {
                const float u0 = s->u[i * J + j], u1 = s->u[(i + 1) * J + j];
                const float v0 = s->v[i * (J + 1) + j], v1 = s->v[i * (J + 1) + j + 1];
                const float div = (((u1 - u0) + v1) - v0) + 0.f;
                double val = -scale * static_cast<double>(div);
                val -= (double)is_solid(mat_ext(m, I, J, i - 1, j)) * scale * static_cast<double>(u0 - 0);
                val += (double)is_solid(mat_ext(m, I, J, i + 1, j)) * scale * static_cast<double>(u1 - 0);
                val -= (double)is_solid(mat_ext(m, I, J, i, j - 1)) * scale * static_cast<double>(v0 - 0);
                val += (double)is_solid(mat_ext(m, I, J, i, j + 1)) * scale * static_cast<double>(v1 - 0);
                rhs[i * J + j] = val;
            } else {
                rhs[i * J + j] = 0.0;
            }
        }
}

/* FlipSolver::applyPressureThreadU/V, flipsolver2d.cpp:1133-1193 */
void orc_apply_pressure(const OrcParams* p, OrcState* s, float dt, const double* pr)
{
    const long I = p->size_i, J = p->size_j;
    const double scale = dt / (p->fluid_density * p->dx);
    const int8_t* m = s->material;
    parallel_for(p->threads, I * J, 1, [&](int64_t a, int64_t b) {
        for (int64_t c = a; c < b; c++) {
            const long i = c / J, j = c % J;
            /* U face (i,j) between cells (i-1,j) and (i,j) */
            {
                const double pc = pr[c];
                const double pm = (i - 1 < 0) ? 0.0 : pr[c - J];
                const int8_t ma = mat_ext(m, I, J, i - 1, j), mb = m[c];
                if (is_fluid(ma) || is_fluid(mb)) {
                    if (is_solid(ma) || is_solid(mb)) s->u[i * J + j] = 0.f;
                    else s->u[i * J + j] = static_cast<float>(s->u[i * J + j] - scale * (pc - pm));
                } else {
                    s->u_valid[i * J + j] = 0;
                }
            }
            /* V face (i,j) between cells (i,j-1) and (i,j) */
            {
                const double pc = pr[c];
                const double pm = (j - 1 < 0) ? 0.0 : pr[c - 1];
                const int8_t ma = mat_ext(m, I, J, i, j - 1), mb = m[c];
                if (is_fluid(ma) || is_fluid(mb)) {
                    if (is_solid(ma) || is_solid(mb)) s->v[i * (J + 1) + j] = 0.f;
                    else s->v[i * (J + 1) + j] = static_cast<float>(s->v[i * (J + 1) + j] - scale * (pc - pm));
                } else {
                    s->v_valid[i * (J + 1) + j] = 0;
                }
            }
        }
    });
}

/* FlipSolver::updateVelocityFromSolids + validSolidNeighborIds +
 * MaterialGrid::u/vSampleAffectedBySolid, flipsolver2d.cpp:766-786,
 * 1077-1104, materialgrid.cpp:152-166 */
void orc_velocity_from_solids(const OrcParams* p, OrcState* s)
{
    const long I = p->size_i, J = p->size_j;
    const int8_t* m = s->material;
    auto sol = [&](long i, long j) { return is_solid(mat_ext(m, I, J, i, j)); };
    for (long i = 0; i < I; i++)
        for (long j = 0; j < J; j++) {
            int ids[9];
            int n = 0;
            for (long di = -1; di < 2; di++)
                for (long dj = -1; dj < 2; dj++) {
                    const long ii = i + di, jj = j + dj;
                    if (ii >= 0 && ii < I && jj >= 0 && jj < J) {
                        const int id = s->solid_id[ii * J + jj];
                        if (id != -1) ids[n++] = id;
                    }
                }
            if (n == 0) continue;
            float avg = 0;
            for (int k = 0; k < n; k++) avg += (ids[k] < p->n_obstacles && p->friction) ? p->friction[ids[k]] : 0.f;
            avg /= static_cast<float>(n);
            if (sol(i, j + 1) || sol(i - 1, j + 1) || sol(i, j) || sol(i - 1, j) || sol(i, j - 1) || sol(i - 1, j - 1))
                s->u[i * J + j] = s->u[i * J + j] * (1 - avg);
            if (sol(i + 1, j) || sol(i + 1, j - 1) || sol(i, j) || sol(i, j - 1) || sol(i - 1, j) || sol(i - 1, j - 1))
                s->v[i * (J + 1) + j] = s->v[i * (J + 1) + j] * (1 - avg);
        }
}

/* FlipSolver::updateDensityGrid[Thread], flipsolver2d.cpp:188-249 */
void orc_density_grid(const OrcParams* p, OrcState* s, const int32_t* cs)
{
    const long I = p->size_i, J = p->size_j;
    const float cell_volume = static_cast<float>(p->dx * p->dx * p->dx);
    const float mass = static_cast<float>((p->fluid_density * cell_volume) /
                                          static_cast<float>(p->particles_per_cell));
    const int8_t* m = s->material;
    parallel_for(p->threads, I * J, 8, [&](int64_t a, int64_t b) {
        for (int64_t c = a; c < b; c++) {
            const long i = c / J, j = c % J;
            float d = 0.f;
            for (long di = -1; di <= 1; di++) {
                const long ii = i + di;
                if (ii < 0 || ii >= I) continue;
                const long j0 = std::max<long>(j - 1, 0), j1 = std::min<long>(j + 1, J - 1);
                for (int32_t k = cs[ii * J + j0]; k < cs[ii * J + j1 + 1]; k++) {
                    const float w = bilinear_hat(s->px[k] - static_cast<float>(i) - 0.5f,
                                                 s->py[k] - static_cast<float>(j) - 0.5f);
                    if (std::abs(w) > 1e-6f) d += w * mass;
                }
            }
            d /= cell_volume;
            if (is_fluid(m[i * J + j]) &&
                (is_empty(mat_ext(m, I, J, i + 1, j)) || is_empty(mat_ext(m, I, J, i - 1, j)) ||
                 is_empty(mat_ext(m, I, J, i, j + 1)) || is_empty(mat_ext(m, I, J, i, j - 1))))
                d = std::clamp(d, static_cast<float>(p->fluid_density), std::numeric_limits<float>::max());
            s->density[i * J + j] = d;
        }
    });
}

/* FlipSolver::calcDensityCorrectionRhs, flipsolver2d.cpp:993-1011 */
void orc_density_rhs(const OrcParams* p, const OrcState* s, float dt, double* rhs)
{
    const int64_t C = (int64_t)p->size_i * p->size_j;
    const double scale = 1.0 / dt;
    for (int64_t c = 0; c < C; c++)
        rhs[c] = is_fluid(s->material[c])
                     ? scale * (1.0 - std::clamp(s->density[c] / p->fluid_density, 0.5, 1.5))
                     : 0.0;
}

/* FlipSolver::adjustParticlesByDensityThread, flipsolver2d.cpp:261-303.
 * linearIndex() of an index past the last row is -1 there and reads out of
 * the vector; it is read as 0 here. */
void orc_adjust_particles(const OrcParams* p, OrcState* s, float dt, const double* pr)
{
    const long I = p->size_i, J = p->size_j;
    const float scale = static_cast<float>((dt * dt) / (p->fluid_density * p->dx * p->dx));
    auto lin = [&](long i, long j) -> long { return (i < 0 || i >= I || j < 0 || j >= J) ? -1 : i * J + j; };
    auto P = [&](long l) -> float { return l < 0 ? 0.f : static_cast<float>(pr[l]); };
    const int8_t* m = s->material;
    for (int64_t k = 0; k < s->n_particles; k++) {
        const float x = s->px[k], y = s->py[k];
        const long ic = std::clamp<long>(static_cast<int>(x - 0.5f), 0, I);
        const long jc = std::clamp<long>(static_cast<int>(y - 0.5f), 0, J);
        const long i = std::clamp<long>(static_cast<int>(x), 0, I);
        const long j = std::clamp<long>(static_cast<int>(y), 0, J);
        const float pci = P(lin(ic, j)), pcj = P(lin(i, jc));
        float pi = P(lin(ic + 1, j)), pj = P(lin(i, jc + 1));
        if (is_solid(mat_ext(m, I, J, ic + 1, j)) || is_solid(mat_ext(m, I, J, ic, j))) pi = pci;
        if (is_solid(mat_ext(m, I, J, i, jc + 1)) || is_solid(mat_ext(m, I, J, i, jc))) pj = pcj;
        s->px[k] = x + (pi - pci) * scale;
        s->py[k] = y + (pj - pcj) * scale;
    }
}

/* ================================================================== */
/* Pressure matrix and solvers                                           */
/* ================================================================== */

/* FlipSolver::getPressureProjectionMatrix, flipsolver2d.cpp:797-887:
 * neighbour material tests clamp (OOB_EXTEND); the fluid-neighbour bit is
 * set only if the *linear* offset index is inside [0, I*J). */
void orc_build_matrix(const int8_t* m, int32_t I, int32_t J, uint8_t* n_out, uint8_t* mask_out, uint8_t* fluid_out)
{
    const long C = (long)I * J;
    for (long i = 0; i < I; i++)
        for (long j = 0; j < J; j++) {
            const long l = i * J + j;
            n_out[l] = 0;
            mask_out[l] = 0;
            fluid_out[l] = is_fluid(m[l]);
            if (!fluid_out[l]) continue;
            const long di[4] = {-1, 1, 0, 0}, dj[4] = {0, 0, -1, 1};
            const long off[4] = {-(long)J, (long)J, -1, 1};
            uint8_t n = 0, mask = 0;
            for (int d = 0; d < 4; d++) {
                const int8_t mn = mat_ext(m, I, J, i + di[d], j + dj[d]);
                if (is_fluid(mn)) {
                    n++;
                    const long ln = l + off[d];
                    if (ln >= 0 && ln < C) mask |= (uint8_t)(1u << d);
                } else if (is_empty(mn)) {
                    n++;
                }
            }
            n_out[l] = n;
            mask_out[l] = mask;
        }
}

/* IndexedPressureParameterUnit::multiply + IndexedPressureParameters::
 * multiplyThread, pressuredata.h:132-146, 215-238 */
static inline double ref_row(const uint8_t n, const uint8_t mask, long l, long J, long C, double scale, const double* x)
{
    auto X = [&](long q) { return (q >= 0 && q < C) ? x[q] : 0.0; };
    return (scale * n) * X(l) + (-scale * (mask & 1)) * X(l - J) + (-scale * ((mask & 2) >> 1)) * X(l + J) +
           (-scale * ((mask & 4) >> 2)) * X(l - 1) + (-scale * ((mask & 8) >> 3)) * X(l + 1);
}

static void spmv_mt(int threads, const uint8_t* n, const uint8_t* mask, const uint8_t* fluid, int32_t I,
                    int32_t J, double scale, const double* x, double* y)
{
    const long C = (long)I * J;
    parallel_for(threads, C, 1, [&](int64_t a, int64_t b) {
        for (long l = a; l < b; l++) y[l] = fluid[l] ? ref_row(n[l], mask[l], l, J, C, scale, x) : x[l];
    });
}

void orc_spmv(const uint8_t* n, const uint8_t* mask, const uint8_t* fluid, int32_t I, int32_t J,
              double scale, const double* x, double* y)
{
    spmv_mt(1, n, mask, fluid, I, J, scale, x, y);
}

/* MaterialGrid::nonsolidNeighborCount(ssize_t linIdx), materialgrid.cpp:129-140,
 * with LinearIndexable2d::index2d's truncating division (linearindexable2d.h:43-51). */
static inline int nonsolid_count_lin(const int8_t* m, long I, long J, long lin)
{
    const long i = lin / J, j = lin - i * J;
    return !is_solid(mat_ext(m, I, J, i - 1, j)) + !is_solid(mat_ext(m, I, J, i + 1, j)) +
           !is_solid(mat_ext(m, I, J, i, j - 1)) + !is_solid(mat_ext(m, I, J, i, j + 1));
}

int32_t orc_nonsolid_count_lin(const int8_t* m, int32_t I, int32_t J, int64_t lin)
{
    return nonsolid_count_lin(m, I, J, (long)lin);
}

/* Reference IPP: FlipSolver::getIPPCoefficients (flipsolver2d.cpp:889-945) +
 * IndexedIPPCoefficientUnit::multiply (PressureIPPCoeficients.h:23-40). */
static void ipp_ref_mt(int threads, const int8_t* m, const uint8_t* fluid, int32_t I, int32_t J, double scale,
                       const double* x, double* y)
{
    const long C = (long)I * J;
    auto X = [&](long q) { return (q >= 0 && q < C) ? x[q] : 0.0; };
    parallel_for(threads, C, 1, [&](int64_t a, int64_t b) {
    for (long l = a; l < b; l++) {
        if (!fluid[l]) {
            y[l] = x[l];
            continue;
        }
        const double iNeg = 1.0 / (nonsolid_count_lin(m, I, J, l - J) * scale);
        const double iPos = 1.0 / (nonsolid_count_lin(m, I, J, l + J) * scale);
        const double jNeg = 1.0 / (nonsolid_count_lin(m, I, J, l - 1) * scale);
        const double jPos = 1.0 / (nonsolid_count_lin(m, I, J, l + 1) * scale);
        const double diag = 1.0 + iNeg * iNeg + jNeg * jNeg;
        y[l] = diag * X(l) + jNeg * X(l - 1) + jPos * X(l + 1) + iNeg * X(l - J) + iPos * X(l + J);
    }
    });
}

void orc_ipp_ref_apply(const int8_t* m, const uint8_t* fluid, int32_t I, int32_t J, double scale,
                       const double* x, double* y)
{
    ipp_ref_mt(1, m, fluid, I, J, scale, x, y);
}

/* The inverse-Poisson preconditioner as inversepoissonpreconditioner.h:45-79
 * defines it: M = K K^T with K = I - L D^-1 (L the strictly lower part of A,
 * D its diagonal), the full sparse product.  K has 1 on the diagonal and
 * 1/n_k at a fluid row's lower fluid neighbours k (mask bits, lower =
 * smaller linear index).  For a fluid row c:
 *   z_c = (1 + sum_lower (1/n_k)^2) r_c + sum_lower r_k / n_k + sum_upper r_u / n_c
 *       + [i-(c) and j-(c-J+1)] r_{c-J+1} / n_{c-J}^2
 *       + [j-(c) and i-(c+J-1)] r_{c+J-1} / n_{c-1}^2
 * the last two being the fill-in at the anti-diagonal neighbours (rows c
 * and d share the lower neighbour e). */
static void ipp_rows(const uint8_t* n, const uint8_t* mask, const uint8_t* fluid, long J, long C, const double* x,
                     double* y, long from, long to)
{
    for (long l = from; l < to; l++) {
        if (!fluid[l]) {
            y[l] = x[l];
            continue;
        }
        const double wc = n[l] ? 1.0 / n[l] : 0.0;
        const long lo[2] = {l - (long)J, l - 1};
        const uint8_t lob[2] = {1, 4};
        const long up[2] = {l + (long)J, l + 1};
        const uint8_t upb[2] = {2, 8};
        double diag = 1.0, off = 0.0;
        for (int d = 0; d < 2; d++)
            if (mask[l] & lob[d]) {
                const double wk = n[lo[d]] ? 1.0 / n[lo[d]] : 0.0;
                diag += wk * wk;
                off += wk * x[lo[d]];
            }
        for (int d = 0; d < 2; d++)
            if (mask[l] & upb[d]) off += wc * x[up[d]];
        // fill-in: (c-J+1) shares c-J, (c+J-1) shares c-1
        if ((mask[l] & 1) && l - J + 1 < C && (mask[l - J + 1] & 4)) {
            const double we = n[l - J] ? 1.0 / n[l - J] : 0.0;
            off += we * we * x[l - J + 1];
        }
        if ((mask[l] & 4) && l + J - 1 < C && (mask[l + J - 1] & 1)) {
            const double we = n[l - 1] ? 1.0 / n[l - 1] : 0.0;
            off += we * we * x[l + J - 1];
        }
        y[l] = diag * x[l] + off;
    }
}

void orc_ipp_apply(const uint8_t* n, const uint8_t* mask, const uint8_t* fluid, int32_t I, int32_t J,
                   const double* x, double* y)
{
    ipp_rows(n, mask, fluid, J, (long)I * J, x, y, 0, (long)I * J);
}

/* LinearSolver::solve, linearsolver.cpp:25-73, with VOps (vmath.cpp:8-149):
 * dot = per-splitRange partials summed in range order; maxAbs keeps each
 * range's last nonzero |r_i| (the comparison is against the zero-initialised
 * output slot, vmath.cpp:100-111). */
int32_t orc_pcg_replica(const int8_t* m, int32_t I, int32_t J, double scale, const double* b, double* x,
                        int32_t iter_limit, double tol, int32_t threads, double* err_trace)
{
    const long C = (long)I * J;
    std::vector<uint8_t> n(C), mask(C), fluid(C);
    orc_build_matrix(m, I, J, n.data(), mask.data(), fluid.data());
    const auto rs = split_range(C, 1, 1, threads);
    Pool& pool = Pool::get(std::max(1, threads));
    std::vector<double> part(rs.size());
    auto dot = [&](const std::vector<double>& a, const std::vector<double>& c) {
        pool.run(rs, [&](size_t k, Range r) {
            double acc = 0.0;
            for (int64_t i = r.start; i < r.end; i++) acc += a[i] * c[i];
            part[k] = acc;
        });
        double res = 0;
        for (double v : part) res += v;
        return res;
    };
    auto maxabs = [&](const std::vector<double>& v) {
        pool.run(rs, [&](size_t k, Range r) {
            double ma = std::numeric_limits<double>::min();
            for (int64_t i = r.start; i < r.end; i++)
                if (std::abs(v[i]) > 0.0) ma = std::abs(v[i]);
            part[k] = ma;
        });
        double mx = 0;
        for (double ma : part)
            if (ma > mx) mx = std::abs(ma);
        return mx;
    };
    auto axpy = [&](std::vector<double>& out, const std::vector<double>& a, const std::vector<double>& c,
                    double val, bool sub) {
        pool.run(rs, [&](size_t, Range r) {
            if (sub)
                for (int64_t i = r.start; i < r.end; i++) out[i] = a[i] - c[i] * val;
            else
                for (int64_t i = r.start; i < r.end; i++) out[i] = a[i] + c[i] * val;
        });
    };
    std::fill(x, x + C, 0.0);
    bool zero = true;
    for (long i = 0; i < C; i++)
        if (std::abs(b[i]) > 1.0e-15) { zero = false; break; }
    if (zero) return 0;
    bool any_fluid = false;
    for (long i = 0; i < C; i++) any_fluid |= fluid[i] != 0;
    std::vector<double> res(b, b + C), aux = res, search = aux, X(C, 0.0);
    double sigma = dot(aux, res);
    for (int it = 0; it < iter_limit; it++) {
        if (any_fluid) spmv_mt(threads, n.data(), mask.data(), fluid.data(), I, J, scale, search.data(), aux.data());
        else aux = search;
        const double alpha = sigma / (dot(aux, search) + 1e-8);
        axpy(X, X, search, alpha, false);
        axpy(res, res, aux, alpha, true);
        const double err = maxabs(res);
        if (err_trace) err_trace[it] = err;
        if (err <= tol) {
            std::copy(X.begin(), X.end(), x);
            return it;
        }
        if (any_fluid) ipp_ref_mt(threads, m, fluid.data(), I, J, scale, res.data(), aux.data());
        else aux = res;
        const double ns = dot(aux, res);
        const double beta = ns / sigma;
        axpy(search, aux, search, beta, false);
        sigma = ns;
    }
    std::copy(X.begin(), X.end(), x);
    return iter_limit;
}

double orc_rel_residual(const int8_t* m, int32_t I, int32_t J, double scale, const double* b, const double* x)
{
    const long C = (long)I * J;
    std::vector<uint8_t> n(C), mask(C), fluid(C);
    orc_build_matrix(m, I, J, n.data(), mask.data(), fluid.data());
    std::vector<double> ax(C);
    orc_spmv(n.data(), mask.data(), fluid.data(), I, J, scale, x, ax.data());
    double rr = 0, bb = 0;
    for (long i = 0; i < C; i++) {
        rr += (b[i] - ax[i]) * (b[i] - ax[i]);
        bb += b[i] * b[i];
    }
    return bb > 0 ? std::sqrt(rr / bb) : std::sqrt(rr);
}

/* FP64 PCG with the symmetric IPP on the reference's own (A, b), run to a
 * tight relative residual: the "converged" pressure that the GPU path is
 * compared against. */
int32_t orc_pcg_converged(const int8_t* m, int32_t I, int32_t J, double scale, const double* b, double* x,
                          int32_t iter_limit, double rel_tol, double* rel_resid_out)
{
    const long C = (long)I * J;
    std::vector<uint8_t> n(C), mask(C), fluid(C);
    orc_build_matrix(m, I, J, n.data(), mask.data(), fluid.data());
    std::fill(x, x + C, 0.0);
    double bb = 0;
    for (long i = 0; i < C; i++) bb += b[i] * b[i];
    if (bb == 0) {
        if (rel_resid_out) *rel_resid_out = 0;
        return 0;
    }
    const double bn = std::sqrt(bb);
    std::vector<double> r(b, b + C), z(C), pp(C), q(C);
    orc_ipp_apply(n.data(), mask.data(), fluid.data(), I, J, r.data(), z.data());
    pp = z;
    double rz = 0;
    for (long i = 0; i < C; i++) rz += r[i] * z[i];
    int it = 0;
    for (; it < iter_limit; it++) {
        orc_spmv(n.data(), mask.data(), fluid.data(), I, J, scale, pp.data(), q.data());
        double pq = 0;
        for (long i = 0; i < C; i++) pq += pp[i] * q[i];
        const double alpha = rz / pq;
        double rr = 0;
        for (long i = 0; i < C; i++) {
            x[i] += alpha * pp[i];
            r[i] -= alpha * q[i];
            rr += r[i] * r[i];
        }
        if (std::sqrt(rr) <= rel_tol * bn) {
            it++;
            break;
        }
        orc_ipp_apply(n.data(), mask.data(), fluid.data(), I, J, r.data(), z.data());
        double rz2 = 0;
        for (long i = 0; i < C; i++) rz2 += r[i] * z[i];
        const double beta = rz2 / rz;
        rz = rz2;
        for (long i = 0; i < C; i++) pp[i] = z[i] + beta * pp[i];
    }
    if (rel_resid_out) *rel_resid_out = orc_rel_residual(m, I, J, scale, b, x);
    return it;
}

/* orc_pcg_converged on `threads` pool threads, for the CPU baseline at the
 * GPU path's accuracy: the products split by split_range, the dot products
 * summed as per-range partials in range order (so a thread count gives one
 * result; another count can differ in the last bits, and converges to the
 * same tolerance). */
static int32_t pcg_converged_mt(int threads, const int8_t* m, int32_t I, int32_t J, double scale, const double* b,
                                double* x, int32_t iter_limit, double rel_tol, double* rel_resid_out)
{
    const long C = (long)I * J;
    std::vector<uint8_t> n(C), mask(C), fluid(C);
    orc_build_matrix(m, I, J, n.data(), mask.data(), fluid.data());
    std::fill(x, x + C, 0.0);
    const auto rs = split_range(C, 1, 1, threads);
    std::vector<double> part(rs.size());
    Pool& pool = Pool::get(std::max(1, threads));
    auto reduce = [&](const std::function<double(long, long)>& f) {
        pool.run(rs, [&](size_t k, Range r) { part[k] = f(r.start, r.end); });
        double t = 0;
        for (double v : part) t += v;
        return t;
    };
    const double bb = reduce([&](long a, long e) {
        double t = 0;
        for (long i = a; i < e; i++) t += b[i] * b[i];
        return t;
    });
    if (bb == 0) {
        if (rel_resid_out) *rel_resid_out = 0;
        return 0;
    }
    const double bn = std::sqrt(bb);
    std::vector<double> r(b, b + C), z(C), pp(C), q(C);
    auto ipp = [&](const double* in, double* out) {
        parallel_for(threads, C, 1,
                     [&](int64_t a, int64_t e) { ipp_rows(n.data(), mask.data(), fluid.data(), J, C, in, out, a, e); });
    };
    ipp(r.data(), z.data());
    parallel_for(threads, C, 1, [&](int64_t a, int64_t e) { std::copy(z.begin() + a, z.begin() + e, pp.begin() + a); });
    double rz = reduce([&](long a, long e) {
        double t = 0;
        for (long i = a; i < e; i++) t += r[i] * z[i];
        return t;
    });
    int it = 0;
    for (; it < iter_limit; it++) {
        spmv_mt(threads, n.data(), mask.data(), fluid.data(), I, J, scale, pp.data(), q.data());
        const double pq = reduce([&](long a, long e) {
            double t = 0;
            for (long i = a; i < e; i++) t += pp[i] * q[i];
            return t;
        });
        const double alpha = rz / pq;
        const double rr = reduce([&](long a, long e) {
            double t = 0;
            for (long i = a; i < e; i++) {
                x[i] += alpha * pp[i];
                r[i] -= alpha * q[i];
                t += r[i] * r[i];
            }
            return t;
        });
        if (std::sqrt(rr) <= rel_tol * bn) {
            it++;
            break;
        }
        ipp(r.data(), z.data());
        const double rz2 = reduce([&](long a, long e) {
            double t = 0;
            for (long i = a; i < e; i++) t += r[i] * z[i];
            return t;
        });
        const double beta = rz2 / rz;
        rz = rz2;
        parallel_for(threads, C, 1, [&](int64_t a, int64_t e) {
            for (long i = a; i < e; i++) pp[i] = z[i] + beta * pp[i];
        });
    }
    if (rel_resid_out) *rel_resid_out = orc_rel_residual(m, I, J, scale, b, x);
    return it;
}

/* ================================================================== */
/* Composition: FlipSolver::step / stepFrame                              */
/* ================================================================== */

/* Substep split, flipsolver2d.cpp:476-490 */
float orc_frame_dt(float vel, float cfl, float frame_dt, float t, int32_t count, int32_t max_substeps,
                   int32_t* finished)
{
    float mss = cfl / (vel + 1e-15f);
    *finished = 0;
    if (t + mss >= frame_dt || count == (max_substeps - 1)) {
        mss = frame_dt - t;
        *finished = 1;
    } else if (t + 2.f * mss >= frame_dt) {
        mss = 0.5f * (frame_dt - t);
    }
    return mss;
}

static int32_t solve(const OrcParams* p, const int8_t* mat_a, double scale, const double* b, double* x,
                     double* rel_resid, bool* converged)
{
    const int32_t I = p->size_i, J = p->size_j;
    if (p->solver_mode == ORC_SOLVER_REF_REPLICA) {
        const int32_t it = orc_pcg_replica(mat_a, I, J, scale, b, x, p->pcg_iter_limit, p->project_tol,
                                           std::max(1, p->threads), nullptr);
        *rel_resid = orc_rel_residual(mat_a, I, J, scale, b, x);
        *converged = it < p->pcg_iter_limit;
        return it;
    }
    const int32_t cap = std::max<int32_t>(20000, 20 * std::max(I, J));
    const int32_t it = p->threads > 1
                           ? pcg_converged_mt(p->threads, mat_a, I, J, scale, b, x, cap, p->project_tol, rel_resid)
                           : orc_pcg_converged(mat_a, I, J, scale, b, x, cap, p->project_tol, rel_resid);
    *converged = *rel_resid <= p->project_tol * 10;
    return it;
}

/* FlipSolver::step, flipsolver2d.cpp:412-462 */
int32_t orc_substep(const OrcParams* p, OrcState* s, float dt, OrcStats* st)
{
    const long I = p->size_i, J = p->size_j;
    const int64_t C = (int64_t)I * J;
    double t0 = now_ms();
    auto stage = [&](int k) {
        const double t1 = now_ms();
        if (st) st->stage_ms[k] += t1 - t0;
        t0 = t1;
    };
    std::vector<uint8_t> alive(s->n_particles);
    /* 1. advect (ADVECTION) */
    orc_advect(p, s, dt, alive.data());
    int64_t killed = 0;
    for (uint8_t a : alive) killed += a == 0;
    stage(0);
    /* 2. the matrix is assembled from the materials as they are now
     * (flipsolver2d.cpp:417-418), before gridUpdate changes them */
    std::vector<int8_t> mat_a(s->material, s->material + C);
    const double sA = dt / (p->fluid_density * p->dx * p->dx);
    stage(1);
    /* 3. prune + rebin (PARTICLE_REBIN) */
    std::vector<int32_t> cs(C + 1);
    orc_sort_particles(p, s, alive.data(), cs.data());
    stage(3);
    /* 4. density correction (DENSITY) */
    if (p->density_mode != ORC_DENSITY_OFF && !p->viscosity_enabled) {   /* flipsolver2d.cpp:425 */
        std::vector<float> dens_local;
        float* dens = s->density;
        if (!dens) {
            dens_local.resize(C);
            dens = dens_local.data();
        }
        OrcState tmp = *s;
        tmp.density = dens;
        orc_density_grid(p, &tmp, cs.data());
        std::vector<double> rhs(C), x(C);
        orc_density_rhs(p, &tmp, dt, rhs.data());
        double rel = 0;
        bool conv = false;
        const int32_t it = solve(p, mat_a.data(), sA, rhs.data(), x.data(), &rel, &conv);
        if (st) {
            st->density_iters = std::max(st->density_iters, it);
            st->density_rel_resid = rel;
        }
        const bool apply = p->density_mode == ORC_DENSITY_ON || (p->density_mode == ORC_DENSITY_GATED && conv);
        if (apply) {
            orc_adjust_particles(p, s, dt, x.data());
            /* particles moved: re-key so every later gather is exact */
            orc_sort_particles(p, s, nullptr, cs.data());
            if (st) st->density_applied++;
        }
    }
    stage(2);
    /* 5. gridUpdate: P2G (+ centeredParamsToGrid's viscosity), SDF, materials
     * (PARTICLE_TO_GRID, GRID_UPDATE) */
    orc_p2g(p, s, cs.data());
    if (p->viscosity_enabled && s->pvisc && s->viscosity) orc_viscosity_grid(p, s, cs.data());
    stage(4);
    orc_fluid_sdf(p, s, cs.data());
    orc_update_materials(p, s);
    stage(5);
    /* 6. afterTransfer, level set, extrapolate */
    orc_after_transfer(p, s);
    orc_extrapolate_levelset_inside(s->fluid_sdf, I, J);
    orc_extrapolate(s->u, s->u_valid, I + 1, J, 10);
    orc_extrapolate(s->v, s->v_valid, I, J + 1, 10);
    /* 7. saved grid + body forces */
    std::memcpy(s->saved_u, s->u, sizeof(float) * (I + 1) * J);
    std::memcpy(s->saved_v, s->v, sizeof(float) * I * (J + 1));
    orc_body_forces(p, s, dt);
    stage(6);
    /* 8. project (PRESSURE) */
    {
        std::vector<double> rhs_local, pr_local;
        double* rhs = s->rhs;
        double* pr = s->pressure;
        if (!rhs) { rhs_local.resize(C); rhs = rhs_local.data(); }
        if (!pr) { pr_local.resize(C); pr = pr_local.data(); }
        orc_pressure_rhs(p, s, rhs);
        double rel = 0;
        bool conv = false;
        const int32_t it = solve(p, mat_a.data(), sA, rhs, pr, &rel, &conv);
        if (st) {
            st->pressure_iters = std::max(st->pressure_iters, it);
            st->pressure_rel_resid = rel;
        }
        orc_apply_pressure(p, s, dt, pr);
    }
    stage(7);
    /* 9. solids; with viscosity the light model and a second projection with
     * the same matrix (VISCOSITY, REPRESSURE: flipsolver2d.cpp:446-453);
     * extrapolate, G2P (PARTICLE_UPDATE) */
    orc_velocity_from_solids(p, s);
    if (p->viscosity_enabled && s->viscosity) {
        orc_viscosity_apply(p, s, dt, 1e-12);
        std::vector<double> rhs_local, pr_local;
        double* rhs = s->rhs;
        double* pr = s->pressure;
        if (!rhs) { rhs_local.resize(C); rhs = rhs_local.data(); }
        if (!pr) { pr_local.resize(C); pr = pr_local.data(); }
        orc_pressure_rhs(p, s, rhs);
        double rel = 0;
        bool conv = false;
        const int32_t it = solve(p, mat_a.data(), sA, rhs, pr, &rel, &conv);
        if (st) {
            st->pressure_iters = std::max(st->pressure_iters, it);
            st->pressure_rel_resid = rel;
        }
        orc_apply_pressure(p, s, dt, pr);
    }
    orc_extrapolate(s->u, s->u_valid, I + 1, J, 10);
    orc_extrapolate(s->v, s->v_valid, I, J + 1, 10);
    orc_g2p(p, s);
    stage(10);
    /* 10. count + reseed (PARTICLE_RESEED); reseeding only fills SOURCE cells */
    const int32_t erased = orc_count_particles(p, s, cs.data());
    orc_reseed(p, s);
    s->rng_epoch++;
    stage(11);
    if (st) {
        st->particles_killed += killed;
        st->particles_erased += erased;
        st->last_dt = dt;
        st->substeps++;
    }
    return 0;
}

/* FlipSolver::stepFrame, flipsolver2d.cpp:464-500 (firstFrameInit is the
 * scene constructor here) */
int32_t orc_step_frame(const OrcParams* p, OrcState* s, OrcStats* st)
{
    const float frame_dt = 1.f / p->fps;
    float t = 0.f;
    int32_t count = 0, finished = 0;
    while (!finished) {
        const float vel = orc_max_particle_velocity(s);
        const float dt = orc_frame_dt(vel, p->cfl_number, frame_dt, t, count, p->max_substeps, &finished);
        if (st) st->max_velocity = vel;
        if (p->sim_method == ORC_SIM_NBFLIP) orc_nb_substep(p, s, dt, st);
        else orc_substep(p, s, dt, st);
        t += dt;
        count++;
        if (count > 50) break;
    }
    return count;
}

/* ================================================================== */
/* First-frame init (flipsolver2d.cpp:502-590, 682-707)                   */
/* ================================================================== */
void orc_init_geometry(const OrcParams* p, OrcState* s, int32_t n_solids, const int32_t* snv, const float* sv,
                       int32_t n_fluids, const int32_t* fnv, const float* fv)
{
    const long I = p->size_i, J = p->size_j;
    const int64_t C = I * J;
    for (int64_t c = 0; c < C; c++) {
        s->material[c] = (int8_t)ORC_EMPTY;
        s->solid_id[c] = -1;
        s->solid_sdf[c] = 0.f;
    }
    /* updateSolids, flipsolver2d.cpp:502-530 */
    const float dxf = static_cast<float>(p->dx);
    for (long i = 0; i < I; i++)
        for (long j = 0; j < J; j++) {
            float dist = std::numeric_limits<float>::max();
            int min_idx = 0;
            const float* v = sv;
            for (int32_t k = 0; k < n_solids; k++) {
                const float x = static_cast<float>((static_cast<float>(i) + 0.5) * dxf);
                const float y = static_cast<float>((static_cast<float>(j) + 0.5) * dxf);
                const float sd = orc_polygon_sdf(v, snv[k], x, y) / dxf;
                if (sd < dist) {
                    min_idx = k;
                    dist = sd;
                }
                v += 2 * snv[k];
            }
            s->solid_sdf[i * J + j] = dist;
            if (dist < 0) {
                s->material[i * J + j] = (int8_t)ORC_SOLID;
                s->solid_id[i * J + j] = min_idx;
            }
        }
    /* updateInitialFluid, flipsolver2d.cpp:573-590 */
    for (long i = 0; i < I; i++)
        for (long j = 0; j < J; j++) {
            const float* v = fv;
            for (int32_t k = 0; k < n_fluids; k++) {
                const float x = (static_cast<float>(i) + 0.5f) * dxf;
                const float y = (static_cast<float>(j) + 0.5f) * dxf;
                if (orc_polygon_sdf(v, fnv[k], x, y) <= 0.f) s->material[i * J + j] = (int8_t)ORC_FLUID;
                v += 2 * fnv[k];
            }
        }
    std::fill(s->u, s->u + (I + 1) * J, 0.f);
    std::fill(s->v, s->v + I * (J + 1), 0.f);
    std::fill(s->u_valid, s->u_valid + (I + 1) * J, 0);
    std::fill(s->v_valid, s->v_valid + I * (J + 1), 0);
    std::fill(s->fluid_sdf, s->fluid_sdf + C, 0.f);
}

static inline uint32_t fmix32(uint32_t h)
{
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return h;
}

/* Counter-based uniform in [0,1): 24 random bits of a murmur3-finalised
 * hash of (seed, cell, k, axis). */
float orc_jitter(uint32_t seed, uint32_t cell, uint32_t k, uint32_t axis)
{
    const uint32_t h = fmix32(seed ^ fmix32(cell * 0x9e3779b1u ^ fmix32(k * 2u + axis + 0x632be5abu)));
    return static_cast<float>(h >> 8) * (1.0f / 16777216.0f);
}

/* firstFrameInit's cell passes (flipsolver2d.cpp:788-795) in the reference's
 * order: updateSinks (555-571), updateSources (532-553), updateSolids
 * (502-530), updateInitialFluid (573-590).  Each pass's coordinate
 * arithmetic is the reference's own (float, or double where it mixes in
 * 0.5 as a double). */
void orc_init_objects(const OrcParams* p, OrcState* s, int32_t n_solids, const int32_t* snv, const float* sv,
                      int32_t n_sources, const int32_t* srnv, const float* srv, int32_t n_sinks,
                      const int32_t* sknv, const float* skv, int32_t n_fluids, const int32_t* fnv, const float* fv)
{
    const long I = p->size_i, J = p->size_j;
    const int64_t C = I * J;
    for (int64_t c = 0; c < C; c++) {
        s->material[c] = (int8_t)ORC_EMPTY;
        s->solid_id[c] = -1;
        s->solid_sdf[c] = 0.f;
        if (s->emitter_id) s->emitter_id[c] = -1;
    }
    const float dx = static_cast<float>(p->dx);
    /* updateSinks */
    for (long i = 0; i < I; i++)
        for (long j = 0; j < J; j++) {
            const float* v = skv;
            for (int32_t k = 0; k < n_sinks; k++) {
                if (orc_polygon_sdf(v, sknv[k], (static_cast<float>(i) + 0.5f) * dx,
                                    (static_cast<float>(j) + 0.5f) * dx) <= 0.f)
                    s->material[i * J + j] = (int8_t)ORC_SINK;
                v += 2 * sknv[k];
            }
        }
    /* updateSources */
    const float hdx = static_cast<float>(dx / 2.0);
    for (long i = 0; i < I; i++)
        for (long j = 0; j < J; j++) {
            const float* v = srv;
            for (int32_t k = 0; k < n_sources; k++) {
                if (orc_polygon_sdf(v, srnv[k], static_cast<float>(i) * dx + hdx, static_cast<float>(j) * dx + hdx) <=
                    0.f) {
                    s->material[i * J + j] = (int8_t)ORC_SOURCE;
                    if (s->emitter_id) s->emitter_id[i * J + j] = k;
                }
                v += 2 * srnv[k];
            }
        }
    /* updateSolids */
    for (long i = 0; i < I; i++)
        for (long j = 0; j < J; j++) {
            float dist = std::numeric_limits<float>::max();
            int min_idx = 0;
            const float* v = sv;
            for (int32_t k = 0; k < n_solids; k++) {
                const float x = static_cast<float>((static_cast<float>(i) + 0.5) * dx);
                const float y = static_cast<float>((static_cast<float>(j) + 0.5) * dx);
                const float sd = orc_polygon_sdf(v, snv[k], x, y) / dx;
                if (sd < dist) {
                    min_idx = k;
                    dist = sd;
                }
                v += 2 * snv[k];
            }
            s->solid_sdf[i * J + j] = dist;
            if (dist < 0) {
                s->material[i * J + j] = (int8_t)ORC_SOLID;
                s->solid_id[i * J + j] = min_idx;
            }
        }
    /* updateInitialFluid: also the viscosity grid of the fluid cells */
    if (s->viscosity) std::fill(s->viscosity, s->viscosity + C, 0.f);
    for (long i = 0; i < I; i++)
        for (long j = 0; j < J; j++) {
            const float* v = fv;
            for (int32_t k = 0; k < n_fluids; k++) {
                if (orc_polygon_sdf(v, fnv[k], (static_cast<float>(i) + 0.5f) * dx, (static_cast<float>(j) + 0.5f) * dx) <=
                    0.f) {
                    s->material[i * J + j] = (int8_t)ORC_FLUID;
                    if (s->viscosity) s->viscosity[i * J + j] = p->fluid_viscosity ? p->fluid_viscosity[k] : 0.f;
                }
                v += 2 * fnv[k];
            }
        }
    std::fill(s->u, s->u + (I + 1) * J, 0.f);
    std::fill(s->v, s->v + I * (J + 1), 0.f);
    std::fill(s->u_valid, s->u_valid + (I + 1) * J, 0);
    std::fill(s->v_valid, s->v_valid + I * (J + 1), 0);
    std::fill(s->fluid_sdf, s->fluid_sdf + C, 0.f);
}

/* FlipSolver::afterTransfer, flipsolver2d.cpp:390-410 */
void orc_after_transfer(const OrcParams* p, OrcState* s)
{
    if (!s->emitter_id) return;
    const long I = p->size_i, J = p->size_j;
    for (long i = 0; i < I; i++)
        for (long j = 0; j < J; j++) {
            if (!is_source(s->material[i * J + j])) continue;
            const int e = s->emitter_id[i * J + j];
            if (e < 0 || e >= p->n_emitters || !p->emitter_transfer || !p->emitter_transfer[e]) continue;
            s->u[i * J + j] = static_cast<float>(p->emitter_velocity[2 * e] / p->dx);
            s->v[i * (J + 1) + j] = static_cast<float>(p->emitter_velocity[2 * e + 1] / p->dx);
            s->u_valid[i * J + j] = 1;
            s->v_valid[i * (J + 1) + j] = 1;
        }
}

float orc_reseed_jitter(uint32_t seed, uint32_t epoch, uint32_t cell, uint32_t k, uint32_t axis)
{
    return orc_jitter(fmix32(seed ^ fmix32(epoch * 0x9e3779b1u + 0x85ebca77u)), cell, k, axis);
}

/* FlipSolver::reseedParticles, flipsolver2d.cpp:627-680, with
 * jitteredPosInCell (1013-1019): whole-cell uniform positions from the
 * counter-based RNG of substep rng_epoch, velocity from the final grid when
 * the source transfers it (velocityAt, staggeredvelocitygrid.cpp:158-167),
 * appended after the survivors in cell order. */
int64_t orc_reseed(const OrcParams* p, OrcState* s)
{
    if (!s->emitter_id) return 0;
    const long I = p->size_i, J = p->size_j;
    const int32_t half = p->particles_per_cell / 2;
    int64_t added = 0;
    for (long i = 0; i < I; i++)
        for (long j = 0; j < J; j++) {
            const long c = i * J + j;
            const int32_t more = half - s->particle_counts[c];
            if (more <= 0 || !is_source(s->material[c])) continue;
            const int e = s->emitter_id[c];
            const bool transfer = e >= 0 && e < p->n_emitters && p->emitter_transfer && p->emitter_transfer[e];
            for (int32_t k = 0; k < more; k++) {
                if (s->n_particles >= s->capacity) return added;
                const float x = static_cast<float>(i) + orc_reseed_jitter(p->reseed_seed, s->rng_epoch, (uint32_t)c, k, 0);
                const float y = static_cast<float>(j) + orc_reseed_jitter(p->reseed_seed, s->rng_epoch, (uint32_t)c, k, 1);
                float vx = 0.f, vy = 0.f;
                if (transfer) {
                    vx = lerp_grid(s->u, I + 1, J, 0.5f, 0.f, x, y);
                    vy = lerp_grid(s->v, I, J + 1, 0.f, 0.5f, x, y);
                }
                const int64_t n = s->n_particles++;
                s->px[n] = x;
                s->py[n] = y;
                s->vx[n] = vx;
                s->vy[n] = vy;
                s->pid[n] = s->next_pid++;
                if (s->pvisc) s->pvisc[n] = p->emitter_viscosity ? p->emitter_viscosity[e] : 0.f;
                added++;
            }
        }
    return added;
}

/* seedInitialFluid (flipsolver2d.cpp:682-707) for strict-fluid cells, with
 * particle k of a cell in quadrant k%4 jittered uniformly inside it. */
int64_t orc_seed_particles(const OrcParams* p, OrcState* s, uint32_t seed)
{
    const long I = p->size_i, J = p->size_j;
    int64_t n = 0;
    for (long i = 0; i < I; i++)
        for (long j = 0; j < J; j++) {
            if (!is_strict_fluid(s->material[i * J + j])) continue;
            const uint32_t cell = (uint32_t)(i * J + j);
            for (int k = 0; k < p->particles_per_cell; k++) {
                if (n >= s->capacity) return -1;
                const int q = k & 3;
                const float qi = static_cast<float>(q >> 1), qj = static_cast<float>(q & 1);
                s->px[n] = static_cast<float>(i) + 0.5f * (qi + orc_jitter(seed, cell, (uint32_t)k, 0));
                s->py[n] = static_cast<float>(j) + 0.5f * (qj + orc_jitter(seed, cell, (uint32_t)k, 1));
                s->vx[n] = 0.f;
                s->vy[n] = 0.f;
                s->pid[n] = (uint32_t)n;
                /* m_viscosityGrid.interpolateAt(pos): offset (0.5, 0.5) (flipsolver2d.cpp:37, 697-702) */
                if (s->pvisc)
                    s->pvisc[n] = s->viscosity ? lerp_grid(s->viscosity, I, J, 0.5f, 0.5f, s->px[n], s->py[n]) : 0.f;
                n++;
            }
        }
    s->n_particles = n;
    return n;
}


/* ================================================================== */
/* Viscosity: the light model (viscositymodel.cpp:4-140)               */
/* ================================================================== */
void orc_viscosity_grid(const OrcParams* p, OrcState* s, const int32_t* cs)
{
    const long I = p->size_i, J = p->size_j;
    for (long i = 0; i < I; i++)
        for (long j = 0; j < J; j++) {
            /* the quadratic B-spline at (x - i, y - j) reaches cells i-2 .. i+1
             * (and j-2 .. j+1); their particles in (cell, id) order */
            float wt = 1e-10f, vv = 0.0f;
            const long j0 = std::max<long>(0, j - 2), j1 = std::min<long>(J - 1, j + 1);
            for (long a = std::max<long>(0, i - 2); a <= std::min<long>(I - 1, i + 1); a++)
                for (int32_t k = cs[a * J + j0]; k < cs[a * J + j1 + 1]; k++) {
                    const float w = qbspline(s->px[k] - static_cast<float>(i), s->py[k] - static_cast<float>(j));
                    if (w > 1e-9f) {
                        wt += w;
                        vv += w * s->pvisc[k];
                    }
                }
            float v = vv / wt;
            /* afterTransfer: a SOURCE cell takes its emitter's viscosity */
            const long c = i * J + j;
            if (is_source(s->material[c]) && s->emitter_id) {
                const int e = s->emitter_id[c];
                if (e >= 0 && e < p->n_emitters) v = p->emitter_viscosity ? p->emitter_viscosity[e] : 0.f;
            }
            s->viscosity[c] = v;
        }
}

void orc_viscosity_matrix(const int8_t* m, const float* visc, int32_t I, int32_t J, float dt, double* diag,
                          double* east, double* north)
{
    const double scale = dt;
    for (long i = 0; i < I; i++)
        for (long j = 0; j < J; j++) {
            const long c = i * J + j;
            const bool sc = is_solid(m[c]);
            diag[c] = sc ? 1.0 : 4.0 * (visc[c] * scale) + 1.0;
            /* a pair's entry is written by each non-solid end; the later one
             * in the (i, j) loop wins: the far end's value if it is not
             * solid, else the near end's */
            east[c] = 0.0;
            if (i + 1 < I) {
                const long b = c + J;
                const bool sb = is_solid(m[b]);
                if (!(sc && sb)) east[c] = (sc ? visc[b] : visc[c]) * scale;
            }
            north[c] = 0.0;
            if (j + 1 < J) {
                const long b = c + 1;
                const bool sb = is_solid(m[b]);
                if (!(sc && sb)) north[c] = (sc ? visc[b] : visc[c]) * scale;
            }
        }
}

static void visc_spmv(const double* d, const double* e, const double* nn, long I, long J, const double* x, double* y)
{
    for (long i = 0; i < I; i++)
        for (long j = 0; j < J; j++) {
            const long c = i * J + j;
            double v = d[c] * x[c];
            if (i + 1 < I) v += e[c] * x[c + J];
            if (i > 0) v += e[c - J] * x[c - J];
            if (j + 1 < J) v += nn[c] * x[c + 1];
            if (j > 0) v += nn[c - 1] * x[c - 1];
            y[c] = v;
        }
}

/* Eigen::ConjugateGradient with its DiagonalPreconditioner, from x = 0,
 * stopping when ||r||^2 < tol^2 ||b||^2 */
static int32_t visc_cg(const double* d, const double* e, const double* nn, long I, long J, const double* b, double* x,
                       double tol)
{
    const long C = I * J;
    std::vector<double> r(b, b + C), z(C), pp(C), q(C);
    std::fill(x, x + C, 0.0);
    double bb = 0;
    for (long c = 0; c < C; c++) bb += b[c] * b[c];
    if (bb == 0.0) return 0;
    const double thr = tol * tol * bb;
    double rr = bb;
    if (rr < thr) return 0;
    double rz = 0;
    for (long c = 0; c < C; c++) {
        pp[c] = r[c] / d[c];
        rz += r[c] * pp[c];
    }
    int32_t it = 0;
    while (it < 100000) {
        visc_spmv(d, e, nn, I, J, pp.data(), q.data());
        double pq = 0;
        for (long c = 0; c < C; c++) pq += pp[c] * q[c];
        const double alpha = rz / pq;
        rr = 0;
        for (long c = 0; c < C; c++) {
            x[c] += alpha * pp[c];
            r[c] -= alpha * q[c];
            rr += r[c] * r[c];
        }
        it++;
        if (rr < thr) break;
        double rzn = 0;
        for (long c = 0; c < C; c++) {
            z[c] = r[c] / d[c];
            rzn += r[c] * z[c];
        }
        const double beta = rzn / rz;
        rz = rzn;
        for (long c = 0; c < C; c++) pp[c] = z[c] + beta * pp[c];
    }
    return it;
}

int32_t orc_viscosity_apply(const OrcParams* p, OrcState* s, float dt, double rel_tol)
{
    const long I = p->size_i, J = p->size_j, C = I * J;
    std::vector<double> d(C), e(C), nn(C), b(C), x(C);
    orc_viscosity_matrix(s->material, s->viscosity, I, J, dt, d.data(), e.data(), nn.data());
    const float rho = static_cast<float>(p->fluid_density);
    int32_t most = 0;
    for (int which = 0; which < 2; which++) {
        float* g = which == 0 ? s->u : s->v;
        const long sj = which == 0 ? J : J + 1;
        for (long i = 0; i < I; i++)
            for (long j = 0; j < J; j++) b[i * J + j] = rho * g[i * sj + j];   /* fillRhs */
        most = std::max(most, visc_cg(d.data(), e.data(), nn.data(), I, J, b.data(), x.data(), rel_tol));
        for (long i = 0; i < I; i++)
            for (long j = 0; j < J; j++) g[i * sj + j] = static_cast<float>(x[i * J + j] / rho);   /* applyResult */
    }
    return most;
}

/* ================================================================== */
/* NB-FLIP: NBFlipSolver (nbflipsolver.cpp)                             */
/* ================================================================== */
/* m_narrowBand, m_combinationBand, m_resamplingBand (nbflipsolver.cpp:18-20) */
static constexpr float kNarrowBand = -3.f, kCombinationBand = -2.f, kResamplingBand = -1.f;

void orc_euler_advect(const float* u, const float* v, int32_t I, int32_t J, float dt, const float* in, int32_t si,
                      int32_t sj, float off_i, float off_j, float* out)
{
    for (long idx = 0; idx < (long)si * sj; idx++) {
        const long i = idx / sj, j = idx - (idx / sj) * sj;
        const V2 q = rk4(u, v, I, J, V2{static_cast<float>(i), static_cast<float>(j)}, -dt);
        out[idx] = lerp_grid(in, si, sj, off_i, off_j, q.x, q.y);
    }
}

void orc_nb_prune(const OrcParams* p, const OrcState* s, uint8_t* alive)
{
    const long I = p->size_i, J = p->size_j;
    const int32_t most = 2 * p->particles_per_cell;
    for (int64_t k = 0; k < s->n_particles; k++) {
        if (!alive[k]) continue;
        const float x = s->px[k], y = s->py[k];
        /* ssize_t of the position: a live particle is inside the grid, where
         * the conversion equals the floor */
        const long c = static_cast<long>(x) * J + static_cast<long>(y);
        const float sdf = lerp_grid(s->fluid_sdf, I, J, 0.f, 0.f, x, y);
        const bool src = is_source(s->material[c]);
        if (!src && sdf < kNarrowBand) {
            alive[k] = 0;
            continue;
        }
        if ((src || sdf < kResamplingBand) && s->particle_counts[c] > most) alive[k] = 0;
    }
}


void orc_nb_combine(const OrcParams* p, OrcState* s, const float* adv_u, const float* adv_v, const float* adv_sdf)
{
    const long I = p->size_i, J = p->size_j;
    const float flt_max = std::numeric_limits<float>::max();
    for (long c = 0; c < I * J; c++) {
        /* updateGridFromSources: std::min(dist, sdf) */
        const float dist = p->source_sdf ? p->source_sdf[c] : flt_max;
        float f = s->fluid_sdf[c];
        f = (f < dist) ? f : dist;
        /* combineLevelset: std::min(advected + h, sdf), h = 1 */
        const float a = adv_sdf[c] + 1.f;
        s->fluid_sdf[c] = (f < a) ? f : a;
    }
    /* combineVelocityGrid: nbCombine(advected, particle, sdf at the face) */
    for (long i = 0; i < I + 1; i++)
        for (long j = 0; j < J; j++) {
            const float sdf = lerp_grid(s->fluid_sdf, I, J, 0.f, 0.f, static_cast<float>(i), 0.5f + static_cast<float>(j));
            if (!(sdf > kCombinationBand)) s->u[i * J + j] = adv_u[i * J + j];
        }
    for (long i = 0; i < I; i++)
        for (long j = 0; j < J + 1; j++) {
            const float sdf = lerp_grid(s->fluid_sdf, I, J, 0.f, 0.f, 0.5f + static_cast<float>(i), static_cast<float>(j));
            if (!(sdf > kCombinationBand)) s->v[i * (J + 1) + j] = adv_v[i * (J + 1) + j];
        }
}

int64_t orc_nb_reseed(const OrcParams* p, OrcState* s)
{
    const long I = p->size_i, J = p->size_j;
    const int32_t ppc = p->particles_per_cell;
    int64_t added = 0;
    for (long i = 0; i < I; i++)
        for (long j = 0; j < J; j++) {
            const long c = i * J + j;
            const bool src = is_source(s->material[c]);
            const float sd = s->fluid_sdf[c];
            if (!(src || (sd < kResamplingBand && sd > kNarrowBand))) continue;
            const int32_t more = ppc - s->particle_counts[c];
            if (more <= 0) continue;
            const int e = s->emitter_id ? s->emitter_id[c] : -1;
            const bool transfer = e >= 0 && e < p->n_emitters && p->emitter_transfer && p->emitter_transfer[e];
            for (int32_t k = 0; k < more; k++) {
                const float x = static_cast<float>(i) + orc_reseed_jitter(p->reseed_seed, s->rng_epoch, (uint32_t)c, k, 0);
                const float y = static_cast<float>(j) + orc_reseed_jitter(p->reseed_seed, s->rng_epoch, (uint32_t)c, k, 1);
                const float ns = lerp_grid(s->fluid_sdf, I, J, 0.f, 0.f, x, y);
                if ((src && sd > kResamplingBand) || ns < kNarrowBand) continue;
                if (s->n_particles >= s->capacity) return added;
                float vx = 0.f, vy = 0.f;
                if (transfer) {
                    vx = lerp_grid(s->u, I + 1, J, 0.5f, 0.f, x, y);
                    vy = lerp_grid(s->v, I, J + 1, 0.f, 0.5f, x, y);
                }
                const int64_t n = s->n_particles++;
                s->px[n] = x;
                s->py[n] = y;
                s->vx[n] = vx;
                s->vy[n] = vy;
                s->pid[n] = s->next_pid++;
                /* the emitter's viscosity, else the grid's at the position */
                if (s->pvisc) {
                    if (e != -1) s->pvisc[n] = (e >= 0 && e < p->n_emitters && p->emitter_viscosity) ? p->emitter_viscosity[e] : 0.f;
                    else s->pvisc[n] = s->viscosity ? lerp_grid(s->viscosity, I, J, 0.5f, 0.5f, x, y) : 0.f;
                }
                added++;
            }
        }
    return added;
}

/* NBFlipSolver::centeredParamsToGridThread (nbflipsolver.cpp:441-504): the
 * cells i-2 .. i+1 x j-2 .. j+1 (where the spline reaches) in (cell, id) order */
void orc_nb_viscosity_grid(const OrcParams* p, OrcState* s, const int32_t* cs)
{
    const long I = p->size_i, J = p->size_j;
    for (long i = 0; i < I; i++)
        for (long j = 0; j < J; j++) {
            float cw = 1e-10f, vv = 0.f;
            bool known = false;
            const long j0 = std::max<long>(0, j - 2), j1 = std::min<long>(J - 1, j + 1);
            for (long a = std::max<long>(0, i - 2); a <= std::min<long>(I - 1, i + 1); a++)
                for (int32_t k = cs[a * J + j0]; k < cs[a * J + j1 + 1]; k++) {
                    const float w = qbspline(s->px[k] - static_cast<float>(i), s->py[k] - static_cast<float>(j));
                    if (std::abs(w) > 1e-6f) {
                        cw += w;
                        vv += w * s->pvisc[k];
                        known = true;
                    }
                }
            s->viscosity[i * J + j] = known ? vv / cw : vv;
        }
}

/* FlipSolver::afterTransfer's SOURCE cells (flipsolver2d.cpp:396-397),
 * updateGridFromSources' viscosity (nbflipsolver.cpp:338-346) and
 * combineCenteredGrids (400-419) */
void orc_nb_combine_viscosity(const OrcParams* p, OrcState* s, float* adv_visc)
{
    const long I = p->size_i, J = p->size_j;
    auto ev = [&](int e) { return (e >= 0 && e < p->n_emitters && p->emitter_viscosity) ? p->emitter_viscosity[e] : 0.f; };
    for (long c = 0; c < I * J; c++) {
        if (is_source(s->material[c]) && s->emitter_id && s->emitter_id[c] >= 0) s->viscosity[c] = ev(s->emitter_id[c]);
        if (p->source_id && p->source_id[c] >= 0) {
            s->viscosity[c] = ev(p->source_id[c]);
            adv_visc[c] = ev(p->source_id[c]);
        }
    }
    for (long i = 0; i < I; i++)
        for (long j = 0; j < J; j++) {
            const float sdf = lerp_grid(s->fluid_sdf, I, J, 0.f, 0.f, 0.5f + static_cast<float>(i),
                                        static_cast<float>(j + 0.5));
            if (!(sdf > kCombinationBand)) s->viscosity[i * J + j] = adv_visc[i * J + j];
        }
}

int32_t orc_nb_substep(const OrcParams* p, OrcState* s, float dt, OrcStats* st)
{
    const long I = p->size_i, J = p->size_j;
    const int64_t C = (int64_t)I * J;
    double t0 = now_ms();
    auto stage = [&](int k) {
        const double t1 = now_ms();
        if (st) st->stage_ms[k] += t1 - t0;
        t0 = t1;
    };
    /* 1. advect: FlipSolver::advect, pruneNarrowBand, then the Euler pass of
     * U, V and the level set over the grid as it is now (66-113) */
    std::vector<uint8_t> alive(s->n_particles);
    orc_advect(p, s, dt, alive.data());
    orc_nb_prune(p, s, alive.data());
    int64_t killed = 0;
    for (uint8_t a : alive) killed += a == 0;
    std::vector<float> adv_u((I + 1) * J), adv_v(I * (J + 1)), adv_sdf(C);
    orc_euler_advect(s->u, s->v, I, J, dt, s->u, I + 1, J, 0.5f, 0.f, adv_u.data());
    orc_euler_advect(s->u, s->v, I, J, dt, s->v, I, J + 1, 0.f, 0.5f, adv_v.data());
    orc_euler_advect(s->u, s->v, I, J, dt, s->fluid_sdf, I, J, 0.f, 0.f, adv_sdf.data());
    const bool visc = p->viscosity_enabled && s->viscosity && s->pvisc;
    std::vector<float> adv_visc(visc ? C : 0);
    if (visc) orc_euler_advect(s->u, s->v, I, J, dt, s->viscosity, I, J, 0.5f, 0.5f, adv_visc.data());
    stage(0);
    /* 2. prune + rebin */
    std::vector<int32_t> cs(C + 1);
    orc_sort_particles(p, s, alive.data(), cs.data());
    stage(3);
    /* 3. particleToGrid, extrapolate(10), the saved grid */
    orc_p2g(p, s, cs.data());
    if (visc) orc_nb_viscosity_grid(p, s, cs.data());
    orc_extrapolate(s->u, s->u_valid, I + 1, J, 10);
    orc_extrapolate(s->v, s->v_valid, I, J + 1, 10);
    std::memcpy(s->saved_u, s->u, sizeof(float) * (I + 1) * J);
    std::memcpy(s->saved_v, s->v, sizeof(float) * I * (J + 1));
    stage(4);
    /* 4. gridUpdate (213-226): updateSdf, extrapolateLevelsetOutside,
     * afterTransfer (FlipSolver's, updateGridFromSources, combine),
     * extrapolateLevelsetInside, updateMaterials, applyBodyForces */
    orc_fluid_sdf(p, s, cs.data());
    orc_extrapolate_levelset_outside(s->fluid_sdf, I, J);
    orc_after_transfer(p, s);
    orc_nb_combine(p, s, adv_u.data(), adv_v.data(), adv_sdf.data());
    if (visc) orc_nb_combine_viscosity(p, s, adv_visc.data());
    orc_extrapolate_levelset_inside(s->fluid_sdf, I, J);
    orc_update_materials(p, s);
    orc_body_forces(p, s, dt);
    stage(5);
    /* 5. the matrix from the materials as they are now, then project */
    std::vector<int8_t> mat_a(s->material, s->material + C);
    const double sA = dt / (p->fluid_density * p->dx * p->dx);
    stage(1);
    {
        std::vector<double> rhs_local, pr_local;
        double* rhs = s->rhs;
        double* pr = s->pressure;
        if (!rhs) { rhs_local.resize(C); rhs = rhs_local.data(); }
        if (!pr) { pr_local.resize(C); pr = pr_local.data(); }
        orc_pressure_rhs(p, s, rhs);
        double rel = 0;
        bool conv = false;
        const int32_t it = solve(p, mat_a.data(), sA, rhs, pr, &rel, &conv);
        if (st) {
            st->pressure_iters = std::max(st->pressure_iters, it);
            st->pressure_rel_resid = rel;
        }
        orc_apply_pressure(p, s, dt, pr);
    }
    stage(7);
    /* 6. solids; viscosity and the second projection (46-53); extrapolate, G2P */
    orc_velocity_from_solids(p, s);
    if (visc) {
        orc_viscosity_apply(p, s, dt, 1e-12);
        double* rhs = s->rhs;
        double* pr = s->pressure;
        std::vector<double> rhs_local, pr_local;
        if (!rhs) { rhs_local.resize(C); rhs = rhs_local.data(); }
        if (!pr) { pr_local.resize(C); pr = pr_local.data(); }
        orc_pressure_rhs(p, s, rhs);
        double rel = 0;
        bool conv = false;
        const int32_t it = solve(p, mat_a.data(), sA, rhs, pr, &rel, &conv);
        if (st) {
            st->pressure_iters = std::max(st->pressure_iters, it);
            st->pressure_rel_resid = rel;
        }
        orc_apply_pressure(p, s, dt, pr);
    }
    orc_extrapolate(s->u, s->u_valid, I + 1, J, 10);
    orc_extrapolate(s->v, s->v_valid, I, J + 1, 10);
    orc_g2p(p, s);
    stage(10);
    /* 7. countParticles + the band reseed */
    const int32_t erased = orc_count_particles(p, s, cs.data());
    orc_nb_reseed(p, s);
    s->rng_epoch++;
    stage(11);
    if (st) {
        st->particles_killed += killed;
        st->particles_erased += erased;
        st->last_dt = dt;
        st->substeps++;
    }
    return 0;
}

void orc_nb_init_objects(const OrcParams* p, OrcState* s, int32_t n_solids, const int32_t* snv, const float* sv,
                         int32_t n_sources, const int32_t* srnv, const float* srv, int32_t n_sinks,
                         const int32_t* sknv, const float* skv, int32_t n_fluids, const int32_t* fnv, const float* fv,
                         float* source_sdf, int32_t* source_id)
{
    const long I = p->size_i, J = p->size_j;
    /* updateSinks, updateSources, updateSolids (no updateInitialFluid) */
    orc_init_objects(p, s, n_solids, snv, sv, n_sources, srnv, srv, n_sinks, sknv, skv, 0, nullptr, nullptr);
    const float dx = static_cast<float>(p->dx);
    const float flt_max = std::numeric_limits<float>::max();
    for (long i = 0; i < I; i++)
        for (long j = 0; j < J; j++) {
            /* fluidSdfFromInitialFluid: centre in double, as updateSolids */
            float dist = flt_max;
            int fid = -1;
            const float* v = fv;
            for (int32_t k = 0; k < n_fluids; k++) {
                const float x = static_cast<float>((static_cast<float>(i) + 0.5) * dx);
                const float y = static_cast<float>((static_cast<float>(j) + 0.5) * dx);
                const float sd = orc_polygon_sdf(v, fnv[k], x, y) / dx;
                if (sd < dist) {
                    dist = sd;
                    fid = k;
                }
                v += 2 * fnv[k];
            }
            s->fluid_sdf[i * J + j] = dist;
            if (dist < 0) {
                s->material[i * J + j] = (int8_t)ORC_FLUID;
                if (s->viscosity && fid != -1) s->viscosity[i * J + j] = p->fluid_viscosity ? p->fluid_viscosity[fid] : 0.f;
            }
            /* updateGridFromSources' distance: the cell corner (i dx, j dx) */
            float sdist = flt_max;
            int sid = -1;
            v = srv;
            for (int32_t k = 0; k < n_sources; k++) {
                const float sd = orc_polygon_sdf(v, srnv[k], static_cast<float>(i) * dx, static_cast<float>(j) * dx) / dx;
                if (sd < sdist) {
                    sdist = sd;
                    sid = k;
                }
                v += 2 * srnv[k];
            }
            if (source_sdf) source_sdf[i * J + j] = sdist;
            if (source_id) source_id[i * J + j] = sdist < 0 ? sid : -1;
        }
    std::fill(s->particle_counts, s->particle_counts + I * J, 0);
    s->n_particles = 0;
    s->next_pid = 0;
    s->rng_epoch = 0;
}

} /* extern "C" */
