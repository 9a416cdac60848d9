// flip_oracle.cpp -- CPU restatement of Banana's 3-D FLIP substep (test infrastructure).
//
// THIS FILE IS THE ORACLE. It is test infrastructure only: the test suite, the smoke check and
// the CPU-baseline leg of bench.py load it as a checker. Nothing in the shipped library links,
// includes or calls it.
//
// It restates, in the same floating-point operation order, the reference solver
//   Simulation/ParticleSolvers/PICFluid/FLIP_Solver<3,float> + PIC_Solver<3,float>
// and the BananaCore helpers it calls. Every function cites the reference file:line it follows
// (paths relative to the reference checkout). The code is templated on the real type:
//   R = float  -> "parity mode": the reference's FP32 arithmetic and operation order;
//   R = double -> "FP64 oracle": the same algorithm in double precision.
//
// Deterministic by default (threads = 1): particle loops run in index order and every reduction
// is a serial left-to-right sum. With threads > 1 the particle scatter uses CAS float adds and
// the PCG reductions use fixed chunks, as the reference does with TBB (Scheduler.h:127-141,
// AtomicOperations.h:31-49); that mode exists to time the reference algorithm on host cores.
//
// Parity status of each function is recorded next to it ("pinned by" / "parity unpinned").
// Build: see oracle/Makefile (g++ -O2 -ffp-contract=off; x86-64 baseline ISA, so no FMA).

#include <algorithm>
#include <atomic>
#include <cmath>
#include <condition_variable>
#include <cstdint>
#include <cstring>
#include <functional>
#include <limits>
#include <mutex>
#include <thread>
#include <vector>

namespace orc {

//------------------------------------------------------------------------------------------------
// small thread pool: parallel_for over [0,n) split into `nthreads` contiguous chunks
//------------------------------------------------------------------------------------------------
class Pool {
public:
    explicit Pool(int n) : m_n(n < 1 ? 1 : n) {
        for(int t = 1; t < m_n; ++t) {
            m_workers.emplace_back([this, t]() { worker(t); });
        }
    }
    ~Pool() {
        {
            std::lock_guard<std::mutex> lk(m_mu);
            m_quit = true;
            ++m_gen;
        }
        m_cv.notify_all();
        for(auto& w : m_workers) w.join();
    }
    int size() const { return m_n; }
    // fn(begin, end, chunk_id)
    void run(size_t n, const std::function<void(size_t, size_t, int)>& fn) {
        if(m_n == 1 || n < 2) {
            fn(0, n, 0);
            return;
        }
        {
            std::lock_guard<std::mutex> lk(m_mu);
            m_fn   = &fn;
            m_size = n;
            m_done = 0;
            ++m_gen;
        }
        m_cv.notify_all();
        chunk(0);
        std::unique_lock<std::mutex> lk(m_mu);
        m_cvDone.wait(lk, [this]() { return m_done == m_n - 1; });
        m_fn = nullptr;
    }
    template<class F> void parallel_for(size_t n, F f) {
        run(n, [&](size_t b, size_t e, int) { for(size_t i = b; i < e; ++i) f(i); });
    }

private:
    void chunk(int t) {
        size_t b = m_size * (size_t)t / (size_t)m_n;
        size_t e = m_size * (size_t)(t + 1) / (size_t)m_n;
        (*m_fn)(b, e, t);
    }
    void worker(int t) {
        uint64_t seen = 0;
        for(;;) {
            std::unique_lock<std::mutex> lk(m_mu);
            m_cv.wait(lk, [&]() { return m_gen != seen; });
            seen = m_gen;
            if(m_quit) return;
            lk.unlock();
            chunk(t);
            lk.lock();
            ++m_done;
            if(m_done == m_n - 1) m_cvDone.notify_one();
        }
    }
    int                                            m_n;
    std::vector<std::thread>                       m_workers;
    std::mutex                                     m_mu;
    std::condition_variable                        m_cv, m_cvDone;
    uint64_t                                       m_gen  = 0;
    bool                                           m_quit = false;
    const std::function<void(size_t, size_t, int)>* m_fn  = nullptr;
    size_t                                         m_size = 0;
    int                                            m_done = 0;
};

//------------------------------------------------------------------------------------------------
template<class R> struct V3 {
    R x, y, z;
    R&       operator[](int d) { return d == 0 ? x : (d == 1 ? y : z); }
    const R& operator[](int d) const { return d == 0 ? x : (d == 1 ? y : z); }
};

template<class T> struct Arr3 {
    int            nx = 0, ny = 0, nz = 0;
    std::vector<T> d;
    void resize(int a, int b, int c, T v) {
        nx = a; ny = b; nz = c;
        d.assign((size_t)a * b * c, v);
    }
    void   assign(T v) { std::fill(d.begin(), d.end(), v); }
    size_t idx(int i, int j, int k) const { return ((size_t)k * ny + j) * nx + i; }  // Array.h:276
    T&       operator()(int i, int j, int k) { return d[idx(i, j, k)]; }
    const T& operator()(int i, int j, int k) const { return d[idx(i, j, k)]; }
    size_t   size() const { return d.size(); }
};

template<class R> constexpr R Tiny() { return std::numeric_limits<R>::min(); }  // Setup.h:335
template<class R> constexpr R Huge() { return std::numeric_limits<R>::max(); }  // Setup.h:336

// MathHelpers.h:515-529 -- get_barycentric with clamping (f forced to 0 / 1 at the clamps)
template<class R> inline void get_barycentric(R x, int& i, R& f, int i_low, int i_high) {
    R s = std::floor(x);
    i   = static_cast<int>(s);
    if(i < i_low) {
        i = i_low;
        f = 0;
    } else if(i > i_high - 2) {
        i = i_high - 2;
        f = 1;
    } else {
        f = (R)(x - s);
    }
}
// MathHelpers.h:548-573
template<class R> inline R lerp(R v0, R v1, R f) { return (R(1.0) - f) * v0 + f * v1; }
template<class R> inline R bilerp(R v00, R v10, R v01, R v11, R fx, R fy) {
    return lerp(lerp(v00, v10, fx), lerp(v01, v11, fx), fy);
}
template<class R> inline R trilerp(R v000, R v100, R v010, R v110, R v001, R v101, R v011, R v111, R fx, R fy, R fz) {
    return lerp(bilerp(v000, v100, v010, v110, fx, fy), bilerp(v001, v101, v011, v111, fx, fy), fz);
}
// MathHelpers.h:804-821 -- 2-point fraction_inside (value < 0 is inside)
template<class R> inline R fraction_inside(R l, R r) {
    if(l < 0 && r < 0) return R(1.0);
    if(l < 0 && r >= 0) return l / (l - r);
    if(l >= 0 && r < 0) return r / (r - l);
    return R(0);
}
// MathHelpers.h:790-797
template<class R> inline void cycle_array(R* arr, int size) {
    R t = arr[0];
    for(int i = 0; i < size - 1; ++i) arr[i] = arr[i + 1];
    arr[size - 1] = t;
}
// MathHelpers.h:825-895 -- 4-corner fraction_inside (square cut-cell area)
template<class R> R fraction_inside4(R phi_bl, R phi_br, R phi_tl, R phi_tr) {
    int inside_count = (phi_bl < 0 ? 1 : 0) + (phi_tl < 0 ? 1 : 0) + (phi_br < 0 ? 1 : 0) + (phi_tr < 0 ? 1 : 0);
    R   list[]       = {phi_bl, phi_br, phi_tr, phi_tl};
    if(inside_count == 4) {
        return R(1.0);
    } else if(inside_count == 3) {
        while(list[0] < 0) cycle_array(list, 4);
        R side0 = R(1.0) - fraction_inside(list[0], list[3]);
        R side1 = R(1.0) - fraction_inside(list[0], list[1]);
        return R(1.0) - R(0.5) * side0 * side1;
    } else if(inside_count == 2) {
        while(list[0] >= 0 || !(list[1] < 0 || list[2] < 0)) cycle_array(list, 4);
        if(list[1] < 0) {
            R side_left  = fraction_inside(list[0], list[3]);
            R side_right = fraction_inside(list[1], list[2]);
            return R(0.5) * (side_left + side_right);
        } else {
            R middle_point = R(0.25) * (list[0] + list[1] + list[2] + list[3]);
            if(middle_point < 0) {
                R area  = R(0);
                R side1 = R(1.0) - fraction_inside(list[0], list[3]);
                R side3 = R(1.0) - fraction_inside(list[2], list[3]);
                area += R(0.5) * side1 * side3;
                R side2 = R(1.0) - fraction_inside(list[2], list[1]);
                R side0 = R(1.0) - fraction_inside(list[0], list[1]);
                area += R(0.5) * side0 * side2;
                return R(1.0) - area;
            } else {
                R area  = R(0);
                R side0 = fraction_inside(list[0], list[1]);
                R side1 = fraction_inside(list[0], list[3]);
                area += R(0.5) * side0 * side1;
                R side2 = fraction_inside(list[2], list[1]);
                R side3 = fraction_inside(list[2], list[3]);
                area += R(0.5) * side2 * side3;
                return area;
            }
        }
    } else if(inside_count == 1) {
        while(list[0] >= 0) cycle_array(list, 4);
        R side0 = fraction_inside(list[0], list[3]);
        R side1 = fraction_inside(list[0], list[1]);
        return R(0.5) * side0 * side1;
    }
    return R(0);
}
template<class R> inline R clamp01(R a) { return a < R(0) ? R(0) : (a > R(1.0) ? R(1.0) : a); }  // MathHelpers.h:368-377

// glm::length of a 3-vector: sqrt((x*x + y*y) + z*z) -- glm compute_dot order (parity unpinned: glm absent)
template<class R> inline R len3(R x, R y, R z) { return std::sqrt(x * x + y * y + z * z); }

//------------------------------------------------------------------------------------------------
// Box boundary SDF, negative OUTSIDE the box (BoundaryObject.h:47 passes bNegativeInside=false)
// GeometryObjects.cpp:246-268 (BoxObject::signedDistance), NumberHelpers.h:338-345 (isInside)
//------------------------------------------------------------------------------------------------
template<class R> struct BoxSDF {
    R lo[3], hi[3];
    R signedDistance(const V3<R>& p, bool negInside) const {
        R    mind   = Huge<R>();
        bool inside = true;
        for(int d = 0; d < 3; ++d)
            if(p[d] < lo[d] || p[d] > hi[d]) inside = false;
        if(inside) {
            for(int d = 0; d < 3; ++d) mind = std::min(mind, std::min(p[d] - lo[d], hi[d] - p[d]));
            mind = -mind;
        } else {
            R cp[3];
            for(int d = 0; d < 3; ++d) cp[d] = std::max(std::min(p[d], hi[d]), lo[d]);
            mind = len3(p.x - cp[0], p.y - cp[1], p.z - cp[2]);
        }
        if(!negInside) mind = -mind;
        return mind;  // uniform scale 1
    }
    // GeometryObjects.cpp:37-85 (gradSignedDistance, 3-D branch, dx = 1e-4)
    V3<R> grad(const V3<R>& p, bool negInside) const {
        const R dx   = R(1e-4);
        auto    sd   = [&](R a, R b, R c) { return signedDistance(V3<R>{a, b, c}, negInside); };
        R       v000 = sd(p.x - dx, p.y - dx, p.z - dx);
        R       v001 = sd(p.x - dx, p.y - dx, p.z + dx);
        R       v010 = sd(p.x - dx, p.y + dx, p.z - dx);
        R       v011 = sd(p.x - dx, p.y + dx, p.z + dx);
        R       v100 = sd(p.x + dx, p.y - dx, p.z - dx);
        R       v101 = sd(p.x + dx, p.y - dx, p.z + dx);
        R       v110 = sd(p.x + dx, p.y + dx, p.z - dx);
        R       v111 = sd(p.x + dx, p.y + dx, p.z + dx);
        R ddx00 = v100 - v000, ddx10 = v110 - v010, ddx01 = v101 - v001, ddx11 = v111 - v011;
        R ddy00 = v010 - v000, ddy10 = v110 - v100, ddy01 = v011 - v001, ddy11 = v111 - v101;
        R ddz00 = v001 - v000, ddz10 = v101 - v100, ddz01 = v011 - v010, ddz11 = v111 - v110;
        V3<R> g;
        g.x = (((ddx00 + ddx01) + ddx10) + ddx11) * R(0.25);
        g.y = (((ddy00 + ddy01) + ddy10) + ddy11) * R(0.25);
        g.z = (((ddz00 + ddz01) + ddz10) + ddz11) * R(0.25);
        return g;
    }
};

//------------------------------------------------------------------------------------------------
// Additional boundary objects (ParticleSolver.cpp:242-298; GeometryObjects.cpp:93-245, 246-266,
// 379-383, 861-896; Animation.h; CubicSpline.h). A shape's distance: query point through the
// inverse of translate * rotate(axis, angle) * scale, then the box / unit-sphere distance with
// bNegativeInside = false, times the uniform scale. Static shapes and the domain box form one
// union group (min), each animated object a group of its own.
//------------------------------------------------------------------------------------------------
struct Shape {
    int    type  = 0;  // 0 box, 1 unit sphere
    bool   xform = false;
    float  inv[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
    float  scale   = 1;
    float  lo[3] = {-1, -1, -1}, hi[3] = {1, 1, 1};
    template<class R> R sd(const V3<R>& p0) const {
        V3<R> p = p0;
        if(xform) {
            const float* m = inv;
            p = V3<R>{((R(m[0]) * p0.x + R(m[1]) * p0.y) + R(m[2]) * p0.z) + R(m[3]),
                      ((R(m[4]) * p0.x + R(m[5]) * p0.y) + R(m[6]) * p0.z) + R(m[7]),
                      ((R(m[8]) * p0.x + R(m[9]) * p0.y) + R(m[10]) * p0.z) + R(m[11])};
        }
        R r;
        if(type == 0) {
            BoxSDF<R> b;
            for(int d = 0; d < 3; ++d) {
                b.lo[d] = R(lo[d]);
                b.hi[d] = R(hi[d]);
            }
            r = b.signedDistance(p, false);
        } else {
            r = -(len3(p.x, p.y, p.z) - R(1));
        }
        return R(scale) * r;
    }
};

// a clamped cubic (end slopes 0) or piecewise linear interpolant through (x_i, y_i), with the
// reference's extrapolation (the end polynomials continued); the cubic's second-derivative
// coefficients come from a dense Gaussian elimination of the end-conditioned system
struct Interp {
    std::vector<double> x, y, a, b, c;
    double              b0 = 0, c0 = 0;
    void build(const std::vector<double>& X, const std::vector<double>& Y, bool cubic) {
        x = X;
        y = Y;
        const size_t n = x.size();
        a.assign(n, 0);
        b.assign(n, 0);
        c.assign(n, 0);
        if(cubic) {
            std::vector<std::vector<double>> A(n, std::vector<double>(n + 1, 0.0));
            auto slope = [&](size_t i) { return (y[i + 1] - y[i]) / (x[i + 1] - x[i]); };
            A[0][0] = 2 * (x[1] - x[0]);
            A[0][1] = x[1] - x[0];
            A[0][n] = 3 * slope(0);
            for(size_t i = 1; i + 1 < n; ++i) {
                A[i][i - 1] = (x[i] - x[i - 1]) / 3;
                A[i][i]     = 2 * (x[i + 1] - x[i - 1]) / 3;
                A[i][i + 1] = (x[i + 1] - x[i]) / 3;
                A[i][n]     = slope(i) - slope(i - 1);
            }
            A[n - 1][n - 2] = x[n - 1] - x[n - 2];
            A[n - 1][n - 1] = 2 * (x[n - 1] - x[n - 2]);
            A[n - 1][n]     = -3 * slope(n - 2);
            for(size_t k = 0; k < n; ++k)  // no pivoting: the system is diagonally dominant
                for(size_t i = k + 1; i < n; ++i) {
                    const double f = A[i][k] / A[k][k];
                    for(size_t j = k; j <= n; ++j) A[i][j] -= f * A[k][j];
                }
            for(size_t i = n; i-- > 0;) {
                double v = A[i][n];
                for(size_t j = i + 1; j < n; ++j) v -= A[i][j] * b[j];
                b[i] = v / A[i][i];
            }
            for(size_t i = 0; i + 1 < n; ++i) {
                a[i] = (b[i + 1] - b[i]) / (3 * (x[i + 1] - x[i]));
                c[i] = slope(i) - (2 * b[i] + b[i + 1]) * (x[i + 1] - x[i]) / 3;
            }
        } else {
            for(size_t i = 0; i + 1 < n; ++i) c[i] = (y[i + 1] - y[i]) / (x[i + 1] - x[i]);
        }
        b0 = b[0];
        c0 = c[0];
        const double h = x[n - 1] - x[n - 2];
        c[n - 1]       = 3 * a[n - 2] * h * h + 2 * b[n - 2] * h + c[n - 2];
        a[n - 1]       = 0;
    }
    double operator()(double t) const {
        const size_t n = x.size();
        size_t       k = 0;
        while(k + 1 < n && x[k + 1] < t) ++k;  // the last knot below t (0 to the left of x_0)
        const double h = t - x[k];
        if(t < x[0]) return (b0 * h + c0) * h + y[0];
        if(t > x[n - 1]) {
            const double hr = t - x[n - 1];
            return (b[n - 1] * hr + c[n - 1]) * hr + y[n - 1];
        }
        return ((a[k] * h + b[k]) * h + c[k]) * h + y[k];
    }
};

}  // namespace orc

extern "C" {
struct orc_params {
    float    box_min[3], box_max[3];          // SimulationDomainBox.BoxMin / BoxMax
    float    cell_size;                        // CellSize
    float    ratio_cell_size_pradius;          // RatioCellSizePRadius
    uint32_t n_expand_cells;                   // NExpandCells
    float    min_timestep, max_timestep;       // MinTimestep, MaxTimestep
    float    cfl_factor;                       // CFLFactor
    float    frame_duration;                   // FrameDuration
    float    cg_relative_tolerance;            // CGRelativeTolerance
    uint32_t max_cg_iteration;                 // MaxCGIteration
    float    pic_flip_ratio;                   // PIC_FLIP_ratio
    uint32_t advection_steps;                  // AdvectionSteps
    int32_t  gravity_type;                     // 0 Earth, 1 Directional, 2 ToCenter, 3 FromCenter
    float    gravity_direction[3];             // GravityDirection
    float    gravity_center[3];                // GravityCenter
    int32_t  apply_gravity;                    // ApplyGravity
    int32_t  reflect_velocity_at_boundary;     // ReflectVelocityAtBoundary
    float    boundary_reflection_multiplier;   // BoundaryReflectionMultiplier
    int32_t  boundary_sdf_at_init;             // (extension) compute boundary SDF once at scene load
};
// a scene's AdditionalBoundaryObjects entry (same layout as the library's bnn_boundary_object)
struct orc_keyframe {
    uint32_t frame;
    float    translation[3];
    float    rotation_axis_angle[4];  // degrees
    float    scale;
    int32_t  has_rotation, has_scale;
};
struct orc_boundary_object {
    int32_t      geometry;  // 0 box, 1 sphere
    float        box_min[3], box_max[3];
    float        translation[3];
    float        rotation_axis_angle[4];  // degrees
    float        scale;
    int32_t      has_translation, has_rotation, has_scale;
    uint32_t     n_keyframes;
    orc_keyframe keyframes[16];
    int32_t      periodic;
    uint32_t     start_frame;
    int32_t      cubic_translation, cubic_rotation, cubic_scale;
};
}

namespace orc {

// one boundary object: its description, its animation channels and its current shape
struct BoundaryObj {
    orc_boundary_object d{};
    bool                dynamic = false, periodic = false;
    uint32_t            start = 0, maxFrame = 0;
    struct Key {
        uint32_t frame;
        double   t[3], r[4], s;
    };
    std::vector<Key> keys;
    Interp           tr[3], rot[4], sc;
    Shape            shape;

    void load(const orc_boundary_object& o) {
        d = o;
        shape.type = o.geometry == 1 ? 1 : 0;
        for(int k = 0; k < 3; ++k) {
            shape.lo[k] = o.box_min[k];
            shape.hi[k] = o.box_max[k];
        }
        const double deg = 3.14159265358979323846 / 180.0;
        Key          k0{0, {o.translation[0], o.translation[1], o.translation[2]}, {1, 1, 1, 0}, o.has_scale ? (double)o.scale : 1.0};
        if(o.rotation_axis_angle[3] != 0.f) {
            for(int k = 0; k < 3; ++k) k0.r[k] = o.rotation_axis_angle[k];
            k0.r[3] = o.rotation_axis_angle[3] * deg;
        }
        keys.assign(1, k0);
        for(uint32_t i = 0; i < o.n_keyframes; ++i) {
            const orc_keyframe& f = o.keyframes[i];
            Key                 k{f.frame, {f.translation[0], f.translation[1], f.translation[2]}, {1, 1, 1, 0}, f.has_scale ? (double)f.scale : 1.0};
            for(int q = 0; q < 4; ++q) k.r[q] = (f.has_rotation ? (double)f.rotation_axis_angle[q] : (q < 3 ? 1.0 : 0.0)) * deg;
            if(k.frame == 0) keys[0] = k;
            else keys.push_back(k);
        }
        dynamic  = o.n_keyframes > 0;
        periodic = o.periodic != 0;
        start    = o.start_frame;
        if(dynamic) {
            std::vector<double> x, t[3], r[4], s;
            for(const Key& k : keys) {
                if(periodic && k.frame < start) continue;
                maxFrame = std::max(maxFrame, k.frame);
                x.push_back(k.frame);
                for(int q = 0; q < 3; ++q) t[q].push_back(k.t[q]);
                for(int q = 0; q < 4; ++q) r[q].push_back(k.r[q]);
                s.push_back(k.s);
            }
            for(int q = 0; q < 3; ++q) tr[q].build(x, t[q], o.cubic_translation != 0);
            for(int q = 0; q < 4; ++q) rot[q].build(x, r[q], o.cubic_rotation != 0);
            sc.build(x, s, o.cubic_scale != 0);
        }
        shape.xform = false;
        if(o.has_translation || (o.has_rotation && o.rotation_axis_angle[3] != 0.f) || o.has_scale) place(keys[0], keys[0].s);
    }
    // Animation::getTransformation / getUniformScale at (frame + fraction) (Animation.h)
    void update(uint32_t frame, double fraction) {
        if(!dynamic) return;
        Key k = keys[0];
        if(!(keys.size() == 1 || (periodic && frame < start))) {
            uint32_t f = frame;
            if(periodic && f > maxFrame) f = (f - start) % (maxFrame - start) + start;
            const double x = f + fraction;
            for(int q = 0; q < 3; ++q) k.t[q] = tr[q](x);
            for(int q = 0; q < 4; ++q) k.r[q] = rot[q](x);
            k.s = sc(x);
        }
        double us = keys[0].s;
        if(keys.size() > 1) {
            uint32_t f = frame;
            if(periodic && f > maxFrame) f = f % maxFrame;
            us = sc(f + fraction);
        }
        place(k, us);
    }
    // the shape's inverse map: M = T R(axis, angle) S inverted by cofactors
    void place(const Key& k, double uniformScale) {
        const double n  = std::sqrt(k.r[0] * k.r[0] + k.r[1] * k.r[1] + k.r[2] * k.r[2]);
        const double ux = n > 0 ? k.r[0] / n : 0, uy = n > 0 ? k.r[1] / n : 0, uz = n > 0 ? k.r[2] / n : 0;
        const double c = std::cos(k.r[3]), sn = std::sin(k.r[3]);
        // R = c I + (1 - c) u u^T + sin [u]x
        double M[3][3] = {{c + (1 - c) * ux * ux, (1 - c) * ux * uy - sn * uz, (1 - c) * ux * uz + sn * uy},
                          {(1 - c) * uy * ux + sn * uz, c + (1 - c) * uy * uy, (1 - c) * uy * uz - sn * ux},
                          {(1 - c) * uz * ux - sn * uy, (1 - c) * uz * uy + sn * ux, c + (1 - c) * uz * uz}};
        for(auto& row : M)
            for(double& e : row) e *= k.s;
        const double det = M[0][0] * (M[1][1] * M[2][2] - M[1][2] * M[2][1]) - M[0][1] * (M[1][0] * M[2][2] - M[1][2] * M[2][0]) +
                           M[0][2] * (M[1][0] * M[2][1] - M[1][1] * M[2][0]);
        double I[3][3];
        for(int i = 0; i < 3; ++i)
            for(int j = 0; j < 3; ++j) {
                const int i1 = (j + 1) % 3, i2 = (j + 2) % 3, j1 = (i + 1) % 3, j2 = (i + 2) % 3;
                I[i][j] = (M[i1][j1] * M[i2][j2] - M[i1][j2] * M[i2][j1]) / det;
            }
        for(int i = 0; i < 3; ++i) {
            double off = 0;
            for(int j = 0; j < 3; ++j) {
                shape.inv[4 * i + j] = (float)I[i][j];
                off -= I[i][j] * k.t[j];
            }
            shape.inv[4 * i + 3] = (float)off;
        }
        shape.scale = (float)uniformScale;
        shape.xform = true;
    }
};

// uniform in [-1, 1) from (seed, p, q, component): 24 bits of a splitmix64 hash, so the value is
// exact in float and the same on the device
static inline uint64_t mix64o(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
static inline float rand11(uint64_t seed, uint32_t p, uint32_t q, int comp) {
    const uint64_t h = mix64o(seed ^ mix64o(((uint64_t)p << 32 | q) * 4u + (uint64_t)comp));
    return (float)(uint32_t)(h >> 40) * (1.0f / 8388608.0f) - 1.0f;
}

enum Field {
    F_U = 0, F_V, F_W,
    F_WEIGHT_U, F_WEIGHT_V, F_WEIGHT_W,
    F_VALID_U, F_VALID_V, F_VALID_W,
    F_EXTRAP_U, F_EXTRAP_V, F_EXTRAP_W,
    F_OLD_U, F_OLD_V, F_OLD_W,
    F_PHI, F_BOUNDARY_SDF, F_ACTIVE_IDX,
    F_TMP_U, F_TMP_V, F_TMP_W,
    F_COUNT
};

struct Base {
    virtual ~Base() {}
};

template<class R> class Solver : public Base {
public:
    using Vec = V3<R>;
    orc_params prm;
    Pool       pool;
    // derived parameters
    R     bmin[3], bmax[3], dx, invdx, halfdx, sdfRadius, particleRadius;
    int   n[3];
    R     gravity[3];
    BoxSDF<R> domain;
    // the scene's additional boundary objects (none: the domain box alone)
    std::vector<BoundaryObj> bobjs;
    // particles
    std::vector<Vec> pos, vel;
    // transfer scheme: 0 FLIP (FLIP_Solver), 1 APIC (APIC_Solver), 2 AFLIP (AFLIP_Solver); the
    // affine matrix of each particle, C[p][d] = the gradient of velocity component d (APIC_Data.h)
    int              scheme = 0;
    std::vector<std::array<Vec, 3>> C;
    // grid (PIC_Data.cpp:101-117, FLIP_Data.h:37-46)
    Arr3<R>    vels[3], weights[3], tmpVels[3], oldVels[3], dVels[3];
    Arr3<char> valids[3], extrapolates[3], tmpValids[3];
    Arr3<uint32_t> activeCellIdx;
    Arr3<R>    fluidSDF, boundarySDF;
    // system
    uint32_t                                    nActive = 0;
    std::vector<std::vector<std::pair<uint32_t, R>>> rows;  // sorted (col, val) per row
    std::vector<R>                              rhs, pressure;
    // pcg settings/outputs (PCGSolver.h:93-108)
    R        tolFactor;
    uint32_t maxIters;
    R        minDiagonalRatio = R(0.25);
    uint32_t outIters         = 0;
    R        outResidual      = 0;
    bool     outSuccess       = true;
    int      precond          = 2;  // 0 none, 1 Jacobi, 2 IC(0) "MICCL0_SYMMETRIC" (the reference's choice)
    // frame bookkeeping (ParticleSolverData.h:43-50)
    R        frameTime = 0;
    uint32_t finishedFrame = 0;
    uint32_t substepCount = 0;
    R        lastSubstep = 0;

    Solver(const orc_params& p, int threads) : prm(p), pool(threads) { makeReady(); }

    // ParticleSolverData.cpp:229-272 + PIC_Data.cpp:41-44, 128-144 + Grid.cpp:38-55
    void makeReady() {
        dx             = R(prm.cell_size);
        particleRadius = dx / R(prm.ratio_cell_size_pradius);
        // the grid (origin, extent, cell counts) is the reference's, computed in its float
        // arithmetic, in both precisions: the FP64 oracle solves the same discretisation
        for(int d = 0; d < 3; ++d) {
            const float exf = prm.cell_size * float(prm.n_expand_cells);
            const float lo = prm.box_min[d] - exf, hi = prm.box_max[d] + exf;
            bmin[d]      = R(lo);
            bmax[d]      = R(hi);
            n[d]         = static_cast<int>(std::ceil((hi - lo) / prm.cell_size));
            domain.lo[d] = R(prm.box_min[d]);
            domain.hi[d] = R(prm.box_max[d]);
        }
        sdfRadius = dx * R(1.01 * std::sqrt(3) / 2.0);
        invdx     = R(1.0) / dx;
        halfdx    = R(0.5) * dx;
        // gravity (ParticleSolverData.cpp:256-271)
        if(prm.gravity_type == 1) {
            R gx = R(prm.gravity_direction[0]), gy = R(prm.gravity_direction[1]), gz = R(prm.gravity_direction[2]);
            R l2 = gx * gx + gy * gy + gz * gz;
            if(l2 < std::numeric_limits<R>::epsilon()) {
                prm.gravity_type = 0;
            } else {
                R is = R(1) / std::sqrt(l2);  // glm::normalize = x * inversesqrt(dot(x, x))
                gravity[0] = R(9.81) * (gx * is);
                gravity[1] = R(9.81) * (gy * is);
                gravity[2] = R(9.81) * (gz * is);
            }
        }
        if(prm.gravity_type == 0) {
            gravity[0] = 0; gravity[1] = R(-9.81); gravity[2] = 0;
        }
        for(int d = 0; d < 3; ++d) {
            int e[3] = {0, 0, 0};
            e[d]     = 1;
            vels[d].resize(n[0] + e[0], n[1] + e[1], n[2] + e[2], R(0));
            weights[d].resize(n[0] + e[0], n[1] + e[1], n[2] + e[2], R(0));
            tmpVels[d].resize(n[0] + e[0], n[1] + e[1], n[2] + e[2], R(0));
            oldVels[d].resize(n[0] + e[0], n[1] + e[1], n[2] + e[2], R(0));
            dVels[d].resize(n[0] + e[0], n[1] + e[1], n[2] + e[2], R(0));
            valids[d].resize(n[0] + e[0], n[1] + e[1], n[2] + e[2], 0);
            extrapolates[d].resize(n[0] + e[0], n[1] + e[1], n[2] + e[2], 0);
            tmpValids[d].resize(n[0] + e[0], n[1] + e[1], n[2] + e[2], 0);
        }
        activeCellIdx.resize(n[0], n[1], n[2], 0u);
        fluidSDF.resize(n[0], n[1], n[2], R(0));
        boundarySDF.resize(n[0] + 1, n[1] + 1, n[2] + 1, R(0));  // zero-filled: PIC_Data.cpp:116
        tolFactor = std::max(R(prm.cg_relative_tolerance), R(1e-30));  // PCGSolver.cpp:59-64
        maxIters  = prm.max_cg_iteration;
        if(prm.boundary_sdf_at_init) computeBoundarySDF();
    }

    // Grid.h:197
    Vec gridCoord(const Vec& p) const { return Vec{(p.x - bmin[0]) / dx, (p.y - bmin[1]) / dx, (p.z - bmin[2]) / dx}; }

    // ArrayHelpers.cpp:91-102 (interpolateValueLinear, 3-D)
    R interp(const Arr3<R>& g, R px, R py, R pz) const {
        int i, j, k;
        R   fi, fj, fk;
        get_barycentric(px, i, fi, 0, g.nx);
        get_barycentric(py, j, fj, 0, g.ny);
        get_barycentric(pz, k, fk, 0, g.nz);
        return trilerp(g(i, j, k), g(i + 1, j, k), g(i, j + 1, k), g(i + 1, j + 1, k),
                       g(i, j, k + 1), g(i + 1, j, k + 1), g(i, j + 1, k + 1), g(i + 1, j + 1, k + 1), fi, fj, fk);
    }
    // PIC_Solver.cpp:1119-1133 (getVelocityFromGrid): gridPos - extra, extra = 0.5 except on the axis
    Vec velFromGrid(const Arr3<R>* g, const Vec& gp) const {
        Vec v;
        v.x = interp(g[0], gp.x, gp.y - R(0.5), gp.z - R(0.5));
        v.y = interp(g[1], gp.x - R(0.5), gp.y, gp.z - R(0.5));
        v.z = interp(g[2], gp.x - R(0.5), gp.y - R(0.5), gp.z);
        return v;
    }

    // PIC_Solver.cpp:1137-1144
    Vec trace_rk2(const Vec& p, R dt) const {
        Vec gp   = gridCoord(p);
        Vec pvel = velFromGrid(vels, gp);
        R   h    = R(0.5) * dt;
        Vec mid{p.x + h * pvel.x, p.y + h * pvel.y, p.z + h * pvel.z};
        gp   = gridCoord(mid);
        pvel = velFromGrid(vels, gp);
        return Vec{p.x + dt * pvel.x, p.y + dt * pvel.y, p.z + dt * pvel.z};
    }

    // the boundary objects after ParticleSolver::generateBoundaries (ParticleSolver.cpp:262-287):
    // group 0 = the domain box (and, with additional objects, every static one in a CSG union),
    // then one group per animated object
    int nGroups() const {
        int n = 1;
        for(const auto& o : bobjs) n += o.dynamic ? 1 : 0;
        return n;
    }
    R groupSD(int g, const Vec& p) const {
        if(g == 0) {
            R sd = domain.signedDistance(p, false);
            for(const auto& o : bobjs)
                if(!o.dynamic) sd = std::min(sd, o.shape.sd(p));
            return sd;
        }
        int k = 0;
        for(const auto& o : bobjs)
            if(o.dynamic && ++k == g) return o.shape.sd(p);
        return Huge<R>();
    }
    // GeometryObjects.cpp:37-85 on a group
    Vec groupGrad(int g, const Vec& p) const {
        if(g == 0 && bobjs.empty()) return domain.grad(p, false);
        const R dx   = R(1e-4);
        auto    sd   = [&](R a, R b, R c) { return groupSD(g, Vec{a, b, c}); };
        R       v000 = sd(p.x - dx, p.y - dx, p.z - dx), v001 = sd(p.x - dx, p.y - dx, p.z + dx);
        R       v010 = sd(p.x - dx, p.y + dx, p.z - dx), v011 = sd(p.x - dx, p.y + dx, p.z + dx);
        R       v100 = sd(p.x + dx, p.y - dx, p.z - dx), v101 = sd(p.x + dx, p.y - dx, p.z + dx);
        R       v110 = sd(p.x + dx, p.y + dx, p.z - dx), v111 = sd(p.x + dx, p.y + dx, p.z + dx);
        R ddx00 = v100 - v000, ddx10 = v110 - v010, ddx01 = v101 - v001, ddx11 = v111 - v011;
        R ddy00 = v010 - v000, ddy10 = v110 - v100, ddy01 = v011 - v001, ddy11 = v111 - v101;
        R ddz00 = v001 - v000, ddz10 = v101 - v100, ddz01 = v011 - v010, ddz11 = v111 - v110;
        return Vec{(((ddx00 + ddx01) + ddx10) + ddx11) * R(0.25), (((ddy00 + ddy01) + ddy10) + ddy11) * R(0.25),
                   (((ddz00 + ddz01) + ddz10) + ddz11) * R(0.25)};
    }

    // BoundaryObject.cpp:72-90 (constrainToBoundary with velocity), every boundary object in turn
    // (PIC_Solver.cpp:362-366)
    bool constrainToBoundary(Vec& p, Vec& v) const {
        bool changed = false;
        for(int grp = 0; grp < nGroups(); ++grp) {
            R phi = groupSD(grp, p);
            if(phi < 0) {
                Vec g  = groupGrad(grp, p);
                R   m2 = g.x * g.x + g.y * g.y + g.z * g.z;
                if(m2 > Tiny<R>()) {
                    R s = std::sqrt(m2);
                    g.x /= s; g.y /= s; g.z /= s;
                    p.x -= phi * g.x; p.y -= phi * g.y; p.z -= phi * g.z;
                    if(prm.reflect_velocity_at_boundary) {
                        // glm::reflect(I, N) = I - 2 dot(N, I) N
                        R dn = g.x * v.x + g.y * v.y + g.z * v.z;
                        R m  = R(prm.boundary_reflection_multiplier);
                        v    = Vec{(v.x - R(2) * dn * g.x) * m, (v.y - R(2) * dn * g.y) * m, (v.z - R(2) * dn * g.z) * m};
                        changed = true;
                    }
                }
            }
        }
        return changed;
    }
    void setBoundaryObjects(const orc_boundary_object* o, uint32_t n) {
        bobjs.assign(n, BoundaryObj{});
        for(uint32_t i = 0; i < n; ++i) bobjs[i].load(o[i]);
        if(prm.boundary_sdf_at_init) computeBoundarySDF();
    }
    // PIC_Solver::advanceScene (PIC_Solver.cpp:70-118)
    void advanceScene(R fraction) {
        bool dyn = false;
        for(auto& o : bobjs)
            if(o.dynamic) {
                o.update(finishedFrame, (double)fraction);
                dyn = true;
            }
        if(dyn) computeBoundarySDF();
    }

    // PIC_Solver.cpp:353-373
    void moveParticles(R dt) {
        const R sub = dt / R(prm.advection_steps);
        pool.parallel_for(pos.size(), [&](size_t p) {
            Vec  pp = pos[p], pv = vel[p];
            bool changed = false;
            for(uint32_t i = 0; i < prm.advection_steps; ++i) {
                pp = trace_rk2(pp, sub);
                if(constrainToBoundary(pp, pv)) changed = true;
            }
            pos[p] = pp;
            if(changed) vel[p] = pv;
        });
    }

    // BoundaryObject::constrainToBoundary(ppos) (BoundaryObject.cpp:55-67): position only
    void constrainPosition(Vec& p) const {
        Vec v{0, 0, 0};
        const int32_t keep = prm.reflect_velocity_at_boundary;
        const_cast<orc_params&>(prm).reflect_velocity_at_boundary = 0;
        constrainToBoundary(p, v);
        const_cast<orc_params&>(prm).reflect_velocity_at_boundary = keep;
    }

    // PIC_Solver::correctParticlePositions (PIC_Solver.cpp:377-485), 3-D branch: a repulsion from
    // every other particle of the 27 cells around the particle's cell, w = smooth_kernel(d^2, r^2)
    // (MathHelpers.h:661-664) with r = 2 particleRadius / sqrt(3), spring += w (p - q) / d; for a
    // pair closer than 0.01 r a random push (threshold / dt) * u, u in [-1, 1)^3; then advectionSteps
    // times pos += spring * (r * dt / advectionSteps * stiffness), each followed by the boundary
    // projection. Neighbours in ascending cell index (k, j, i) and, in a cell, ascending particle
    // index (the reference's cell lists are unordered).
    // Parity unpinned against the running reference: its random push comes from an unseeded
    // generator (NumberHelpers.h:146,211), so here u is a hash of (seed, p, q, component); and
    // smooth_kernel's pow(1 - x, 3) is written (t * t) * t.
    void correctParticlePositions(R dt, R stiffness, uint64_t seed) {
        const int nx = n[0], ny = n[1], nz = n[2];
        const size_t np_ = pos.size();
        std::vector<int64_t> cellOf(np_);
        std::vector<uint32_t> start((size_t)nx * ny * nz + 1, 0);
        auto cellIdx = [&](const Vec& p, int d) { return (int)((d == 0 ? p.x - bmin[0] : d == 1 ? p.y - bmin[1] : p.z - bmin[2]) / dx); };
        for(size_t p = 0; p < np_; ++p) {
            int c[3];
            for(int d = 0; d < 3; ++d) c[d] = std::min(std::max(cellIdx(pos[p], d), 0), n[d] - 1);
            cellOf[p] = ((int64_t)c[2] * ny + c[1]) * nx + c[0];
            ++start[cellOf[p] + 1];
        }
        for(size_t c = 0; c < (size_t)nx * ny * nz; ++c) start[c + 1] += start[c];
        std::vector<uint32_t> list(np_), fill(start.begin(), start.end() - 1);
        for(size_t p = 0; p < np_; ++p) list[fill[cellOf[p]]++] = (uint32_t)p;
        const R pr        = R(prm.cell_size) / R(prm.ratio_cell_size_pradius);
        const R radius    = R(2.0) * pr / R(std::sqrt(3));
        const R radius2   = radius * radius;
        const R threshold = R(0.01) * radius, threshold2 = threshold * threshold;
        const R substep   = dt / R(prm.advection_steps);
        const R K         = radius * substep * stiffness;
        std::vector<Vec> out(np_);
        pool.parallel_for(np_, [&](size_t p) {
            const Vec pp = pos[p];
            int       pc[3];
            for(int d = 0; d < 3; ++d) pc[d] = cellIdx(pp, d);  // getCellIdx<Int>: not clamped
            Vec spring{0, 0, 0};
            for(int k = -1; k <= 1; ++k)
                for(int j = -1; j <= 1; ++j)
                    for(int i = -1; i <= 1; ++i) {
                        const int ci = pc[0] + i, cj = pc[1] + j, ck = pc[2] + k;
                        if(ci < 0 || cj < 0 || ck < 0 || ci >= nx || cj >= ny || ck >= nz) continue;  // Grid::isValidCell
                        const size_t c = ((size_t)ck * ny + cj) * nx + ci;
                        for(uint32_t e = start[c]; e < start[c + 1]; ++e) {
                            const uint32_t q = list[e];
                            if(q == p) continue;
                            const Vec xpq{pp.x - pos[q].x, pp.y - pos[q].y, pp.z - pos[q].z};
                            const R   d2 = xpq.x * xpq.x + xpq.y * xpq.y + xpq.z * xpq.z;
                            const R   x2 = d2 / radius2;
                            const R   t  = R(1.0) - x2;
                            const R   w  = x2 > R(1.0) ? R(0) : (t * t) * t;
                            if(d2 > threshold2) {
                                const R s = std::sqrt(d2);
                                spring.x += w * xpq.x / s;
                                spring.y += w * xpq.y / s;
                                spring.z += w * xpq.z / s;
                            } else {
                                const R a = threshold / dt;
                                spring.x += a * R(rand11(seed, (uint32_t)p, q, 0));
                                spring.y += a * R(rand11(seed, (uint32_t)p, q, 1));
                                spring.z += a * R(rand11(seed, (uint32_t)p, q, 2));
                            }
                        }
                    }
            Vec np2 = pp;
            for(uint32_t it = 0; it < prm.advection_steps; ++it) {
                np2.x += spring.x * K;
                np2.y += spring.y * K;
                np2.z += spring.z * K;
                constrainPosition(np2);
            }
            out[p] = np2;
        });
        pos = out;
    }

    // float/double CAS add (AtomicOperations.h:31-49), used only when threads > 1. ORC_PLANT_RACE
    // (the ThreadSanitizer check's negative control, never in the library build) makes it a plain add.
    static void atomicAddR(R& dst, R v) {
#ifdef ORC_PLANT_RACE
        dst += v;
#else
        std::atomic_ref<R> a(dst);
        R cur = a.load(std::memory_order_relaxed);
        while(!a.compare_exchange_weak(cur, cur + v, std::memory_order_relaxed)) {}
#endif
    }

    // FLIP_Solver.cpp:57-95 (+ ArrayHelpers.cpp:50-76 getCoordinatesAndWeights)
    void mapParticles2Grid() {
        for(int d = 0; d < 3; ++d) {
            vels[d].assign(0);
            tmpVels[d].assign(0);
            valids[d].assign(0);
        }
        const bool par = pool.size() > 1;
        pool.parallel_for(pos.size(), [&](size_t p) {
            Vec gp = gridCoord(pos[p]);
            for(int d = 0; d < 3; ++d) {
                R q[3] = {gp.x - (d == 0 ? R(0) : R(0.5)), gp.y - (d == 1 ? R(0) : R(0.5)), gp.z - (d == 2 ? R(0) : R(0.5))};
                int ii[3];
                R   ff[3];
                get_barycentric(q[0], ii[0], ff[0], 0, vels[d].nx);
                get_barycentric(q[1], ii[1], ff[1], 0, vels[d].ny);
                get_barycentric(q[2], ii[2], ff[2], 0, vels[d].nz);
                R fi = ff[0], fj = ff[1], fk = ff[2];
                R w[8] = {(R(1.0) - fi) * (R(1.0) - fj) * (R(1.0) - fk), fi * (R(1.0) - fj) * (R(1.0) - fk),
                          (R(1.0) - fi) * fj * (R(1.0) - fk),            fi * fj * (R(1.0) - fk),
                          (R(1.0) - fi) * (R(1.0) - fj) * fk,            fi * (R(1.0) - fj) * fk,
                          (R(1.0) - fi) * fj * fk,                       fi * fj * fk};
                for(int c = 0; c < 8; ++c) {
                    int    i = ii[0] + (c & 1), j = ii[1] + ((c >> 1) & 1), k = ii[2] + ((c >> 2) & 1);
                    R      vd  = vel[p][d];
                    if(scheme != 0) {
                        // APIC_Solver.cpp:73-76: w (v_d + dot(C_d, gpos - ppos)), gpos =
                        // getWorldCoordinate(index + extra) = (index + extra) * dx + bMin
                        const R ex[3] = {d == 0 ? R(0) : R(0.5), d == 1 ? R(0) : R(0.5), d == 2 ? R(0) : R(0.5)};
                        const R gx = (R(i) + ex[0]) * dx + bmin[0], gy = (R(j) + ex[1]) * dx + bmin[1], gz = (R(k) + ex[2]) * dx + bmin[2];
                        const Vec& cd = C[p][d];
                        vd = vd + (cd.x * (gx - pos[p].x) + cd.y * (gy - pos[p].y) + cd.z * (gz - pos[p].z));
                    }
                    R      mom = w[c] * vd;
                    size_t id  = vels[d].idx(i, j, k);
                    if(par) {
                        atomicAddR(vels[d].d[id], mom);
                        atomicAddR(tmpVels[d].d[id], w[c]);
                    } else {
                        vels[d].d[id] += mom;
                        tmpVels[d].d[id] += w[c];
                    }
                }
            }
        });
        for(int d = 0; d < 3; ++d) {
            pool.parallel_for(vels[d].size(), [&](size_t i) {
                R wsum = tmpVels[d].d[i];
                if(wsum > Tiny<R>()) {
                    vels[d].d[i] /= wsum;
                    valids[d].d[i] = 1;
                }
            });
        }
    }

    // PIC_Solver.cpp:992-1090 (4 Jacobi layers over the interior index range)
    void extrapolate(Arr3<R>& grid, Arr3<R>& temp, Arr3<char>& valid, Arr3<char>& oldValid, Arr3<char>& extrap) {
        extrap.assign(0);
        temp.d = grid.d;
        for(int layers = 0; layers < 4; ++layers) {
            oldValid.d = valid.d;
            pool.parallel_for((size_t)(grid.nz - 2), [&](size_t kk) {
                int k = (int)kk + 1;
                for(int j = 1; j < grid.ny - 1; ++j)
                    for(int i = 1; i < grid.nx - 1; ++i) {
                        if(oldValid(i, j, k)) continue;
                        R        sum   = R(0);
                        uint32_t count = 0;
                        if(oldValid(i + 1, j, k)) { sum += grid(i + 1, j, k); ++count; }
                        if(oldValid(i - 1, j, k)) { sum += grid(i - 1, j, k); ++count; }
                        if(oldValid(i, j + 1, k)) { sum += grid(i, j + 1, k); ++count; }
                        if(oldValid(i, j - 1, k)) { sum += grid(i, j - 1, k); ++count; }
                        if(oldValid(i, j, k + 1)) { sum += grid(i, j, k + 1); ++count; }
                        if(oldValid(i, j, k - 1)) { sum += grid(i, j, k - 1); ++count; }
                        if(count > 0) {
                            temp(i, j, k)   = sum / static_cast<R>(count);
                            valid(i, j, k)  = 1;
                            extrap(i, j, k) = 1;
                        }
                    }
            });
            grid.d = temp.d;
        }
    }
    void extrapolateVelocity() {  // PIC_Solver.cpp:985-989
        for(int d = 0; d < 3; ++d) extrapolate(vels[d], tmpVels[d], valids[d], tmpValids[d], extrapolates[d]);
    }

    // ArrayHelpers.cpp:214-251 (interpolateGradient, 3-D)
    Vec interpolateGradient(R px, R py, R pz, const Arr3<R>& g) const {
        int i, j, k;
        R   fi, fj, fk;
        get_barycentric(px, i, fi, 0, g.nx);
        get_barycentric(py, j, fj, 0, g.ny);
        get_barycentric(pz, k, fk, 0, g.nz);
        R v000 = g(i, j, k), v001 = g(i, j, k + 1), v010 = g(i, j + 1, k), v011 = g(i, j + 1, k + 1);
        R v100 = g(i + 1, j, k), v101 = g(i + 1, j, k + 1), v110 = g(i + 1, j + 1, k), v111 = g(i + 1, j + 1, k + 1);
        R ddx00 = v100 - v000, ddx10 = v110 - v010, ddx01 = v101 - v001, ddx11 = v111 - v011;
        R ddy00 = v010 - v000, ddy10 = v110 - v100, ddy01 = v011 - v001, ddy11 = v111 - v101;
        R ddz00 = v001 - v000, ddz10 = v101 - v100, ddz01 = v011 - v010, ddz11 = v111 - v110;
        return Vec{bilerp(ddx00, ddx10, ddx01, ddx11, fj, fk), bilerp(ddy00, ddy10, ddy01, ddy11, fi, fk),
                   bilerp(ddz00, ddz10, ddz01, ddz11, fi, fj)};
    }

    // PIC_Solver.cpp:1094-1115 (component d reads the already-updated components < d)
    void constrainGridVelocity() {
        for(int d = 0; d < 3; ++d) {
            tmpVels[d].d = vels[d].d;
            Arr3<R>& a = vels[d];
            pool.parallel_for((size_t)a.nz, [&](size_t kk) {
                int k = (int)kk;
                for(int j = 0; j < a.ny; ++j)
                    for(int i = 0; i < a.nx; ++i) {
                        if(extrapolates[d](i, j, k) == 1 && weights[d](i, j, k) < Tiny<R>()) {
                            R gp[3] = {R(i) + R(0.5), R(j) + R(0.5), R(k) + R(0.5)};
                            gp[d]   = R(i * (d == 0) + j * (d == 1) + k * (d == 2));
                            Vec gpos{gp[0], gp[1], gp[2]};
                            Vec v   = velFromGrid(vels, gpos);
                            Vec nrm = interpolateGradient(gp[0], gp[1], gp[2], boundarySDF);
                            R   m2  = nrm.x * nrm.x + nrm.y * nrm.y + nrm.z * nrm.z;
                            if(m2 > Tiny<R>()) {
                                R s = std::sqrt(m2);
                                nrm.x /= s; nrm.y /= s; nrm.z /= s;
                                R dt = v.x * nrm.x + v.y * nrm.y + v.z * nrm.z;
                                v.x -= dt * nrm.x; v.y -= dt * nrm.y; v.z -= dt * nrm.z;
                                tmpVels[d](i, j, k) = v[d];
                            }
                        }
                    }
            });
            vels[d].d = tmpVels[d].d;
        }
    }

    // FLIP_Data.h:48-52
    void backupGridVelocity() {
        for(int d = 0; d < 3; ++d) oldVels[d].d = vels[d].d;
    }

    // PIC_Solver.cpp:505-529
    void addGravity(R dt) {
        if(!prm.apply_gravity) return;
        if(prm.gravity_type == 0) {
            R g = R(9.81) * dt;
            for(auto& x : vels[1].d) x -= g;
        } else if(prm.gravity_type == 1) {
            for(int d = 0; d < 3; ++d) {
                R g = gravity[d] * dt;
                for(auto& x : vels[d].d) x += g;
            }
        } else {
            // ToCenter / FromCenter: the reference loops over velocities[0].vsize() for every
            // component (PIC_Solver.cpp:521), i.e. it indexes each component by the u-face range.
            for(int d = 0; d < 3; ++d) {
                Arr3<R>& a = vels[0];
                for(int k = 0; k < a.nz; ++k)
                    for(int j = 0; j < a.ny; ++j)
                        for(int i = 0; i < a.nx; ++i) {
                            R gp[3] = {R(i) + R(0.5), R(j) + R(0.5), R(k) + R(0.5)};
                            gp[d]   = R(d == 0 ? i : (d == 1 ? j : k));
                            R w[3]  = {gp[0] * dx + bmin[0], gp[1] * dx + bmin[1], gp[2] * dx + bmin[2]};
                            R c[3]  = {R(prm.gravity_center[0]), R(prm.gravity_center[1]), R(prm.gravity_center[2])};
                            R dv[3];
                            for(int e = 0; e < 3; ++e) dv[e] = (prm.gravity_type == 2) ? c[e] - w[e] : w[e] - c[e];
                            R is = R(1) / std::sqrt(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
                            R gd = R(9.81) * (dv[d] * is);
                            if(vels[d].nx > i && vels[d].ny > j && vels[d].nz > k) vels[d](i, j, k) += gd * dt;
                        }
            }
        }
    }

    // PIC_Solver.cpp:543-597
    void computeCellWeights() {
        const auto& b = boundarySDF;
        for(int k = 0; k <= n[2]; ++k)
            for(int j = 0; j <= n[1]; ++j)
                for(int i = 0; i <= n[0]; ++i) {
                    if(i < weights[0].nx && j < weights[0].ny && k < weights[0].nz)
                        weights[0](i, j, k) = clamp01(R(1.0) - fraction_inside4(b(i, j, k), b(i, j + 1, k), b(i, j, k + 1), b(i, j + 1, k + 1)));
                    if(i < weights[1].nx && j < weights[1].ny && k < weights[1].nz)
                        weights[1](i, j, k) = clamp01(R(1.0) - fraction_inside4(b(i, j, k), b(i, j, k + 1), b(i + 1, j, k), b(i + 1, j, k + 1)));
                    if(i < weights[2].nx && j < weights[2].ny && k < weights[2].nz)
                        weights[2](i, j, k) = clamp01(R(1.0) - fraction_inside4(b(i, j, k), b(i, j + 1, k), b(i + 1, j, k), b(i + 1, j + 1, k)));
                }
    }

    static void atomicMinR(R& dst, R v) {
        std::atomic_ref<R> a(dst);
        R cur = a.load(std::memory_order_relaxed);
        while(v < cur && !a.compare_exchange_weak(cur, v, std::memory_order_relaxed)) {}
    }

    // PIC_Solver.cpp:601-689 (min is order independent, so the spin lock is replaced by a CAS min)
    void computeFluidSDF() {
        fluidSDF.assign(dx * R(3.0));
        const bool par = pool.size() > 1;
        pool.parallel_for(pos.size(), [&](size_t p) {
            const Vec& pp = pos[p];
            int c[3], lo[3], hi[3];
            for(int d = 0; d < 3; ++d) {
                c[d]  = static_cast<int>((pp[d] - bmin[d]) / dx);   // Grid.h:156-162
                lo[d] = std::max(0, c[d] - 1);
                hi[d] = std::min(c[d] + 1, n[d] - 1);
            }
            for(int k = lo[2]; k <= hi[2]; ++k)
                for(int j = lo[1]; j <= hi[1]; ++j)
                    for(int i = lo[0]; i <= hi[0]; ++i) {
                        // Grid.h:184-188: Vec3<R>(i+0.5, j+0.5, k+0.5) * dx + bMin
                        R sx = R(i + 0.5) * dx + bmin[0];
                        R sy = R(j + 0.5) * dx + bmin[1];
                        R sz = R(k + 0.5) * dx + bmin[2];
                        R phi = len3(sx - pp.x, sy - pp.y, sz - pp.z) - sdfRadius;
                        if(par) {
                            atomicMinR(fluidSDF(i, j, k), phi);
                        } else if(phi < fluidSDF(i, j, k)) {
                            fluidSDF(i, j, k) = phi;
                        }
                    }
        });
        // solid extension (PIC_Solver.cpp:671-687)
        const auto& b = boundarySDF;
        for(int k = 0; k < n[2]; ++k)
            for(int j = 0; j < n[1]; ++j)
                for(int i = 0; i < n[0]; ++i) {
                    if(fluidSDF(i, j, k) < halfdx) {
                        R s = R(0.125) * (b(i, j, k) + b(i + 1, j, k) + b(i, j + 1, k) + b(i + 1, j + 1, k) + b(i, j, k + 1) +
                                          b(i + 1, j, k + 1) + b(i, j + 1, k + 1) + b(i + 1, j + 1, k + 1));
                        if(s < 0) fluidSDF(i, j, k) = -halfdx;
                    }
                }
    }

    void addElement(uint32_t r, uint32_t c, R v) {  // SparseMatrix.cpp:87-96
        auto& row = rows[r];
        auto  it  = std::lower_bound(row.begin(), row.end(), c, [](const std::pair<uint32_t, R>& a, uint32_t b) { return a.first < b; });
        if(it != row.end() && it->first == c) it->second += v;
        else row.insert(it, {c, v});
    }

    // PIC_Solver.cpp:693-893. Cells with phi == 0 exactly are skipped (the reference would write
    // row 0 for them through activeCellIdx's zero fill; see the design notes).
    void computeSystem(R dt) {
        activeCellIdx.assign(0u);
        nActive = 0;
        for(size_t i = 0; i < fluidSDF.size(); ++i)
            if(fluidSDF.d[i] < 0) activeCellIdx.d[i] = nActive++;
        rows.assign(nActive, {});
        rhs.assign(nActive, R(0));
        pressure.assign(nActive, R(0));
        const auto &u = vels[0], &v = vels[1], &w = vels[2];
        const auto &wu = weights[0], &wv = weights[1], &ww = weights[2];
        for(int k = 1; k < n[2] - 1; ++k)
            for(int j = 1; j < n[1] - 1; ++j)
                for(int i = 1; i < n[0] - 1; ++i) {
                    const R pc = fluidSDF(i, j, k);
                    if(!(pc < 0)) continue;
                    const R pr = fluidSDF(i + 1, j, k), pl = fluidSDF(i - 1, j, k);
                    const R pt = fluidSDF(i, j + 1, k), pb = fluidSDF(i, j - 1, k);
                    const R pf = fluidSDF(i, j, k + 1), pn = fluidSDF(i, j, k - 1);
                    R rv = R(0);
                    const R tr = wu(i + 1, j, k) * dt; rv -= wu(i + 1, j, k) * u(i + 1, j, k);
                    const R tl = wu(i, j, k) * dt;     rv += wu(i, j, k) * u(i, j, k);
                    const R tt = wv(i, j + 1, k) * dt; rv -= wv(i, j + 1, k) * v(i, j + 1, k);
                    const R tb = wv(i, j, k) * dt;     rv += wv(i, j, k) * v(i, j, k);
                    const R tf = ww(i, j, k + 1) * dt; rv -= ww(i, j, k + 1) * w(i, j, k + 1);
                    const R tn = ww(i, j, k) * dt;     rv += ww(i, j, k) * w(i, j, k);
                    const uint32_t ic = activeCellIdx(i, j, k);
                    rhs[ic]           = rv;
                    R ct              = R(0);
                    auto side = [&](R pnb, R term, int a, int b2, int c2) {
                        if(pnb < 0) {
                            ct += term;
                            addElement(ic, activeCellIdx(a, b2, c2), -term);
                        } else {
                            R theta = std::max(R(0.01), fraction_inside(pc, pnb));
                            ct += term / theta;
                        }
                    };
                    side(pr, tr, i + 1, j, k);
                    side(pl, tl, i - 1, j, k);
                    side(pt, tt, i, j + 1, k);
                    side(pb, tb, i, j - 1, k);
                    side(pf, tf, i, j, k + 1);
                    side(pn, tn, i, j, k - 1);
                    addElement(ic, ic, ct);
                }
    }

    //--------------------------------------------------------------------------------------------
    // PCG (PCGSolver.cpp:143-199, 221-269, 398-466)
    //--------------------------------------------------------------------------------------------
    std::vector<R>        icInvDiag, icColValue, icADiag;
    std::vector<uint32_t> icColIndex, icColStart;
    std::vector<R>        jacobi;

    R dotv(const std::vector<R>& a, const std::vector<R>& b) {
        if(pool.size() == 1) {
            R s = 0;
            for(size_t i = 0; i < a.size(); ++i) s += a[i] * b[i];
            return s;
        }
        std::vector<R> part(pool.size(), R(0));
        pool.run(a.size(), [&](size_t b0, size_t e0, int t) {
            R s = 0;
            for(size_t i = b0; i < e0; ++i) s += a[i] * b[i];
            part[t] = s;
        });
        R s = 0;
        for(R x : part) s += x;
        return s;
    }
    R maxAbs(const std::vector<R>& a) {
        R m = 0;
        for(R x : a) m = std::max(m, std::abs(x));
        return m;
    }
    void spmv(const std::vector<R>& x, std::vector<R>& y) {  // SparseMatrix.cpp:220-235
        y.resize(nActive);
        pool.parallel_for(nActive, [&](size_t i) {
            R t = 0;
            for(const auto& e : rows[i]) t += e.second * x[e.first];
            y[i] = t;
        });
    }
    void formIC0() {  // PCGSolver.cpp:398-466
        const uint32_t N = nActive;
        icInvDiag.assign(N, 0); icADiag.assign(N, 0); icColStart.assign(N + 1, 0);
        icColValue.clear(); icColIndex.clear();
        for(uint32_t i = 0; i < N; ++i) {
            icColStart[i] = (uint32_t)icColIndex.size();
            for(const auto& e : rows[i]) {
                if(e.first > i) {
                    icColIndex.push_back(e.first);
                    icColValue.push_back(e.second);
                } else if(e.first == i) {
                    icInvDiag[i] = icADiag[i] = e.second;
                }
            }
        }
        icColStart[N] = (uint32_t)icColIndex.size();
        for(uint32_t k = 0; k < N; ++k) {
            if(icADiag[k] < Tiny<R>()) continue;
            if(icInvDiag[k] < minDiagonalRatio * icADiag[k]) icInvDiag[k] = R(1.0) / std::sqrt(icADiag[k]);
            else icInvDiag[k] = R(1.0) / std::sqrt(icInvDiag[k]);
            for(uint32_t p = icColStart[k]; p < icColStart[k + 1]; ++p) icColValue[p] *= icInvDiag[k];
            for(uint32_t p = icColStart[k]; p < icColStart[k + 1]; ++p) {
                uint32_t j    = icColIndex[p];
                R        mult = icColValue[p];
                icInvDiag[j] -= mult * icColValue[p];
                uint32_t a = p + 1, b = icColStart[j], as = icColStart[k + 1], bs = icColStart[j + 1];
                while(a < as && b < bs) {
                    if(icColIndex[b] == icColIndex[a]) {
                        icColValue[b] -= mult * icColValue[a];
                        ++a; ++b;
                    } else if(icColIndex[b] < icColIndex[a]) {
                        ++b;
                    } else {
                        ++a;
                    }
                }
            }
        }
    }
    // PCGSolver.cpp:301-313. Rows with a zero diagonal exist (solid-extension cells in the wall band
    // whose six faces all have weight 0); the reference's 1/a_ii would be inf there, the IC(0) it
    // actually uses skips them (PCGSolver.cpp:419-421). Their preconditioner entry is 0 here.
    void formJacobi() {
        jacobi.assign(nActive, R(0));
        for(uint32_t i = 0; i < nActive; ++i)
            for(const auto& e : rows[i])
                if(e.first == i) jacobi[i] = e.second > R(0) ? R(1.0) / e.second : R(0);  // zero rows: see below
    }
    void applyPrecond(const std::vector<R>& x, std::vector<R>& out) {  // PCGSolver.cpp:221-269
        if(precond == 0) { out = x; return; }
        if(precond == 1) {
            out.resize(x.size());
            for(size_t i = 0; i < x.size(); ++i) out[i] = jacobi[i] * x[i];
            return;
        }
        out = x;
        const uint32_t N = nActive;
        for(uint32_t i = 0; i < N; ++i) {
            out[i] *= icInvDiag[i];
            for(uint32_t j = icColStart[i]; j < icColStart[i + 1]; ++j) out[icColIndex[j]] -= icColValue[j] * out[i];
        }
        if(N == 0) return;
        uint32_t i = N;
        do {
            --i;
            for(uint32_t j = icColStart[i]; j < icColStart[i + 1]; ++j) out[i] -= icColValue[j] * out[icColIndex[j]];
            out[i] *= icInvDiag[i];
        } while(i > 0);
    }
    bool solve() {  // PCGSolver.cpp:143-199 (solve_precond, zero initial guess)
        std::vector<R> z, s, r;
        pressure.assign(nActive, R(0));
        r           = rhs;
        outResidual = maxAbs(r);
        if(outResidual < std::numeric_limits<R>::min()) {
            outIters   = 0;
            outSuccess = true;
            return true;
        }
        R tol = tolFactor * outResidual;
        if(precond == 2) formIC0();
        if(precond == 1) formJacobi();
        applyPrecond(r, z);
        R rho = dotv(z, r);
        if(rho < std::numeric_limits<R>::min() || std::isnan(rho)) {
            outIters   = 0;
            outSuccess = false;
            return false;
        }
        s = z;
        for(uint32_t it = 0; it < maxIters; ++it) {
            spmv(s, z);
            R tmp = dotv(s, z);
            if(tmp < std::numeric_limits<R>::min()) {
                outIters   = it + 1;
                outSuccess = true;
                return true;
            }
            R alpha = rho / tmp;
            for(size_t i = 0; i < s.size(); ++i) pressure[i] += alpha * s[i];
            for(size_t i = 0; i < z.size(); ++i) r[i] += -alpha * z[i];
            outResidual = maxAbs(r);
            if(outResidual < tol) {
                outIters   = it + 1;
                outSuccess = true;
                return true;
            }
            applyPrecond(r, z);
            R rho_new = dotv(z, r);
            R beta    = rho_new / rho;
            for(size_t i = 0; i < z.size(); ++i) z[i] += beta * s[i];
            s.swap(z);
            rho = rho_new;
        }
        outIters   = maxIters;
        outSuccess = false;
        return false;
    }

    // PIC_Solver.cpp:911-981
    void updateProjectedVelocity(R dt) {
        for(int d = 0; d < 3; ++d) valids[d].assign(0);
        auto P = [&](int i, int j, int k, R phi) { return phi < 0 ? pressure[activeCellIdx(i, j, k)] : R(0); };
        for(int k = 0; k < n[2]; ++k)
            for(int j = 0; j < n[1]; ++j)
                for(int i = 0; i < n[0]; ++i) {
                    const R pc = fluidSDF(i, j, k);
                    const R pl = i > 0 ? fluidSDF(i - 1, j, k) : R(0);
                    const R pb = j > 0 ? fluidSDF(i, j - 1, k) : R(0);
                    const R pn = k > 0 ? fluidSDF(i, j, k - 1) : R(0);
                    const R prc = P(i, j, k, pc);
                    const R prl = i > 0 ? P(i - 1, j, k, pl) : R(0);
                    const R prb = j > 0 ? P(i, j - 1, k, pb) : R(0);
                    const R prn = k > 0 ? P(i, j, k - 1, pn) : R(0);
                    if(i > 0 && (pc < 0 || pl < 0) && weights[0](i, j, k) > 0) {
                        R theta = std::max(R(0.01), fraction_inside(pl, pc));
                        vels[0](i, j, k) -= dt * (prc - prl) / theta;
                        valids[0](i, j, k) = 1;
                    }
                    if(j > 0 && (pc < 0 || pb < 0) && weights[1](i, j, k) > 0) {
                        R theta = std::max(R(0.01), fraction_inside(pb, pc));
                        vels[1](i, j, k) -= dt * (prc - prb) / theta;
                        valids[1](i, j, k) = 1;
                    }
                    if(k > 0 && weights[2](i, j, k) > 0 && (pc < 0 || pn < 0)) {
                        R theta = std::max(R(0.01), fraction_inside(pn, pc));
                        vels[2](i, j, k) -= dt * (prc - prn) / theta;
                        valids[2](i, j, k) = 1;
                    }
                }
        for(int d = 0; d < 3; ++d)
            for(size_t i = 0; i < vels[d].size(); ++i)
                if(valids[d].d[i] == 0) vels[d].d[i] = 0;
    }

    // PIC_Solver.cpp:533-539
    bool pressureProjection(R dt) {
        computeCellWeights();
        computeFluidSDF();
        computeSystem(dt);
        bool ok = solve();
        updateProjectedVelocity(dt);
        return ok;
    }

    // FLIP_Solver.cpp:115-121
    void computeChangesGridVelocity() {
        for(int d = 0; d < 3; ++d)
            for(size_t i = 0; i < vels[d].size(); ++i) dVels[d].d[i] = vels[d].d[i] - oldVels[d].d[i];
    }

    // ArrayHelpers.cpp:275-302 (interpolateGradientValue, 3-D): the gradient of the trilinear
    // interpolant of g over the cell of q, divided by the cell size; the eight corner terms are
    // added in the reference's order
    Vec interpGrad(const Arr3<R>& g, R qx, R qy, R qz) const {
        int i, j, k;
        R   fi, fj, fk;
        get_barycentric(qx, i, fi, 0, g.nx);
        get_barycentric(qy, j, fj, 0, g.ny);
        get_barycentric(qz, k, fk, 0, g.nz);
        const R v000 = g(i, j, k), v001 = g(i, j, k + 1), v010 = g(i, j + 1, k), v011 = g(i, j + 1, k + 1);
        const R v100 = g(i + 1, j, k), v101 = g(i + 1, j, k + 1), v110 = g(i + 1, j + 1, k), v111 = g(i + 1, j + 1, k + 1);
        const R a = R(1.0) - fi, b = R(1.0) - fj, c = R(1.0) - fk;
        Vec     gr;
        gr.x = v000 * (-b * c) + v001 * (-b * fk) + v010 * (-fj * c) + v011 * (-fj * fk) + v100 * (b * c) + v101 * (b * fk) +
               v110 * (fj * c) + v111 * (fj * fk);
        gr.y = v000 * (-a * c) + v001 * (-a * fk) + v010 * (a * c) + v011 * (a * fk) + v100 * (-fi * c) + v101 * (-fi * fk) +
               v110 * (fi * c) + v111 * (fi * fk);
        gr.z = v000 * (-a * b) + v001 * (a * b) + v010 * (-a * fj) + v011 * (a * fj) + v100 * (-fi * b) + v101 * (fi * b) +
               v110 * (-fi * fj) + v111 * (fi * fj);
        gr.x /= dx;
        gr.y /= dx;
        gr.z /= dx;
        return gr;
    }
    // APIC_Solver.cpp:106-115: row d of C is the gradient of velocity component d at gridPos - extra
    std::array<Vec, 3> affineFromGrid(const Arr3<R>* g, const Vec& gp) const {
        return {interpGrad(g[0], gp.x, gp.y - R(0.5), gp.z - R(0.5)), interpGrad(g[1], gp.x - R(0.5), gp.y, gp.z - R(0.5)),
                interpGrad(g[2], gp.x - R(0.5), gp.y - R(0.5), gp.z)};
    }

    // FLIP_Solver.cpp:99-111, 125-133; APIC_Solver.cpp:94-102; AFLIP_Solver.cpp:94-113
    void mapGrid2Particles() {
        const R r = R(prm.pic_flip_ratio);
        if(scheme == 1) {  // APIC: PIC velocity and the affine matrix from the grid
            pool.parallel_for(pos.size(), [&](size_t p) {
                Vec gp = gridCoord(pos[p]);
                vel[p] = velFromGrid(vels, gp);
                C[p]   = affineFromGrid(vels, gp);
            });
            return;
        }
        if(scheme == 2) {  // AFLIP: the FLIP blend for both the velocity and the affine matrix
            pool.parallel_for(pos.size(), [&](size_t p) {
                Vec                gp  = gridCoord(pos[p]);
                Vec                gv  = velFromGrid(vels, gp);
                Vec                dgv = velFromGrid(dVels, gp);
                std::array<Vec, 3> gc  = affineFromGrid(vels, gp);
                std::array<Vec, 3> dgc = affineFromGrid(dVels, gp);
                Vec                pv  = vel[p];
                std::array<Vec, 3> pc  = C[p];
                for(int d = 0; d < 3; ++d) vel[p][d] = (R(1.0) - r) * gv[d] + r * (pv[d] + dgv[d]);
                for(int d = 0; d < 3; ++d)
                    for(int e = 0; e < 3; ++e) C[p][d][e] = (R(1.0) - r) * gc[d][e] + r * (pc[d][e] + dgc[d][e]);
            });
            return;
        }
        pool.parallel_for(pos.size(), [&](size_t p) {
            Vec gp  = gridCoord(pos[p]);
            Vec gv  = velFromGrid(vels, gp);
            Vec dgv = velFromGrid(dVels, gp);
            Vec pv  = vel[p];
            for(int d = 0; d < 3; ++d) vel[p][d] = (R(1.0) - r) * gv[d] + r * (pv[d] + dgv[d]);
        });
    }

    // FLIP_Solver.cpp:42-53 (FLIP and AFLIP, AFLIP_Solver.cpp:43-45); APIC_Solver.cpp:41-48
    bool advanceVelocity(R dt) {
        if(scheme == 1) {
            mapParticles2Grid();
            addGravity(dt);
            bool ok = pressureProjection(dt);
            extrapolateVelocity();
            constrainGridVelocity();
            mapGrid2Particles();
            return ok;
        }
        mapParticles2Grid();
        extrapolateVelocity();
        constrainGridVelocity();
        backupGridVelocity();
        addGravity(dt);
        bool ok = pressureProjection(dt);
        extrapolateVelocity();
        constrainGridVelocity();
        computeChangesGridVelocity();
        mapGrid2Particles();
        return ok;
    }

    // one body of the advanceFrame substep loop (PIC_Solver.cpp:292-294), CorrectPosition = false
    bool step(R dt) {
        moveParticles(dt);
        return advanceVelocity(dt);
    }

    // PIC_Solver.cpp:342-349
    R timestepCFL() {
        R maxVel = Tiny<R>();
        for(int d = 0; d < 3; ++d) maxVel = std::max(maxVel, maxAbs(vels[d].d));
        R ts = maxVel > Tiny<R>() ? R(prm.cfl_factor) * (dx / maxVel) : Huge<R>();
        return std::min(std::max(ts, R(prm.min_timestep)), R(prm.max_timestep));
    }

    // PIC_Solver.cpp:267-311 (no dynamic objects, no frame IO)
    bool advanceFrame() {
        const R fd = R(prm.frame_duration);
        frameTime    = 0;
        substepCount = 0;
        bool ok      = true;
        while(frameTime < fd) {
            if(finishedFrame > 0) advanceScene(frameTime / fd);  // PIC_Solver.cpp:279-281
            R sub = timestepCFL();
            R rem = fd - frameTime;
            if(frameTime + sub >= fd) sub = rem;
            else if(frameTime + R(1.5) * sub >= fd) sub = rem * R(0.5);
            ok &= step(sub);
            lastSubstep = sub;
            frameTime += sub;
            ++substepCount;
        }
        ++finishedFrame;
        return ok;
    }

    // PIC_Solver.cpp:1158-1168: the smallest distance over the boundary objects
    void computeBoundarySDF() {
        const int ng = nGroups();
        for(int k = 0; k <= n[2]; ++k)
            for(int j = 0; j <= n[1]; ++j)
                for(int i = 0; i <= n[0]; ++i) {
                    Vec w{R(i) * dx + bmin[0], R(j) * dx + bmin[1], R(k) * dx + bmin[2]};
                    R   m = Huge<R>();
                    for(int g = 0; g < ng; ++g) m = std::min(m, groupSD(g, w));
                    boundarySDF(i, j, k) = m;
                }
    }

    //--------------------------------------------------------------------------------------------
    template<class T> void* fieldPtr(int f, size_t& count, int& kind) {
        // kind: 0 = R, 1 = char, 2 = uint32
        kind = 0;
        switch(f) {
            case F_U: case F_V: case F_W: count = vels[f].size(); return vels[f].d.data();
            case F_WEIGHT_U: case F_WEIGHT_V: case F_WEIGHT_W: count = weights[f - F_WEIGHT_U].size(); return weights[f - F_WEIGHT_U].d.data();
            case F_VALID_U: case F_VALID_V: case F_VALID_W: kind = 1; count = valids[f - F_VALID_U].size(); return valids[f - F_VALID_U].d.data();
            case F_EXTRAP_U: case F_EXTRAP_V: case F_EXTRAP_W: kind = 1; count = extrapolates[f - F_EXTRAP_U].size(); return extrapolates[f - F_EXTRAP_U].d.data();
            case F_OLD_U: case F_OLD_V: case F_OLD_W: count = oldVels[f - F_OLD_U].size(); return oldVels[f - F_OLD_U].d.data();
            case F_PHI: count = fluidSDF.size(); return fluidSDF.d.data();
            case F_BOUNDARY_SDF: count = boundarySDF.size(); return boundarySDF.d.data();
            case F_ACTIVE_IDX: kind = 2; count = activeCellIdx.size(); return activeCellIdx.d.data();
            case F_TMP_U: case F_TMP_V: case F_TMP_W: count = tmpVels[f - F_TMP_U].size(); return tmpVels[f - F_TMP_U].d.data();
        }
        count = 0;
        return nullptr;
    }
};

//------------------------------------------------------------------------------------------------
// Cell binning: Grid::collectIndexToCells + getParticleIdxSortedByCell (Grid.cpp:92-114, 283-309).
// The reference's order inside a cell is its spin-lock push_back order (nondeterministic); the
// oracle defines it as ascending particle index (a stable counting sort by linear cell index).
// Pinned by: the Grid known-answer test (Grid/_Grid.Test.hpp:35-44), brute-force counts and the
// permutation / sortedness invariants in tests/test_oracle_binning.py.
//------------------------------------------------------------------------------------------------
static inline int cellIndexAxis(float p, float bmin, float dx) { return static_cast<int>((p - bmin) / dx); }

//------------------------------------------------------------------------------------------------
// Morton z-order key of NeighborSearch::z_sort (NeighborSearch.cpp:44-86, DataStructures.h:32,
// Morton.h: morton3D_64_encode = m3D_e_sLUT<uint_fast64_t, uint_fast32_t>, Morton3D.h:54-66).
// On LP64 Linux uint_fast32_t is 64 bits wide, so the table loop runs over 8 bytes and the 64-bit
// result keeps the low 21.33 bits of each coordinate interleave (x at bit 3b, y 3b+1, z 3b+2).
// Restated bit by bit from the interleave definition (a for-loop, not the lookup tables).
//------------------------------------------------------------------------------------------------
static inline uint64_t spread3(uint64_t v, int nbits) {
    uint64_t r = 0;
    for(int b = 0; b < nbits; ++b) r |= ((v >> b) & 1ull) << (3 * b);
    return r;
}
static inline uint64_t mortonKeyOracle(const float* x, float invCellSize) {
    uint64_t c[3];
    for(int d = 0; d < 3; ++d) {
        float   tmp = x[d] - float(1.2345);
        int     k   = tmp >= 0 ? static_cast<int>(invCellSize * tmp) : static_cast<int>(invCellSize * tmp) - 1;
        int64_t v   = static_cast<int64_t>(k) - (static_cast<int64_t>(std::numeric_limits<int>::lowest()) + 1);
        c[d]        = static_cast<uint64_t>(v);
    }
    // m3D_e_sLUT with an 8-byte coordinate: answer = answer << 24 | chunk(byte i), i = 7..0.
    // Only the chunks of bytes 0, 1 and the low 16 bits of byte 2's chunk survive in 64 bits.
    uint64_t answer = 0;
    for(int i = 8; i > 0; --i) {
        int      shift = (i - 1) * 8;
        uint64_t chunk = spread3((c[2] >> shift) & 0xFF, 8) << 2 | spread3((c[1] >> shift) & 0xFF, 8) << 1 |
                         spread3((c[0] >> shift) & 0xFF, 8);
        answer = (answer << 24) | chunk;
    }
    return answer;
}

//------------------------------------------------------------------------------------------------
// Seeded scene generator (replaces the reference's unseeded ParticleGenerator,
// ParticleGenerator.hpp:87-136): one particle per sub-cell of the grid's sub-cell lattice
// (ppa^3 per cell) whose centre lies inside an object, jittered uniformly inside its sub-cell.
// Objects are visited in order; a sub-cell already claimed by an earlier object is skipped.
//------------------------------------------------------------------------------------------------
struct GridDesc {
    float bmin[3], dx;
    int   n[3];
    explicit GridDesc(const orc_params& p) {
        dx = p.cell_size;
        for(int d = 0; d < 3; ++d) {
            float ex = dx * float(p.n_expand_cells);
            bmin[d]  = p.box_min[d] - ex;
            float bx = p.box_max[d] + ex;
            n[d]     = static_cast<int>(std::ceil((bx - bmin[d]) / dx));
        }
    }
};
static inline uint64_t splitmix64(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
static inline double rand01(uint64_t seed, uint64_t id, uint32_t stream) {
    uint64_t h = splitmix64(seed ^ splitmix64(id * 0x2545F4914F6CDD1Dull + stream));
    return (double)(h >> 11) * (1.0 / 9007199254740992.0);
}

}  // namespace orc

extern "C" {

struct orc_scene_spec {
    int32_t  n_boxes;
    float    box_min[8][3], box_max[8][3], box_vel[8][3];
    int32_t  n_spheres;
    float    sphere_center[8][3], sphere_radius[8], sphere_vel[8][3];
    float    vel_sigma;        // > 0: add N(0, sigma^2) per component
    int32_t  ppa;              // particles per axis per cell (2 -> 8 per cell)
    float    jitter;           // 1.0 = uniform over the whole sub-cell
};

// returns the particle count; pos/vel may be null (count only)
uint64_t orc_generate_scene(const orc_params* prm, const orc_scene_spec* sp, uint64_t seed, float* pos, float* vel) {
    using namespace orc;
    GridDesc      tmp(*prm);  // the solver's grid (ParticleSolverData.cpp:247-253, Grid.cpp:38-55)
    const int     ppa = sp->ppa > 0 ? sp->ppa : 2;
    const double  h   = (double)tmp.dx / ppa;
    const int64_t NX = (int64_t)tmp.n[0] * ppa, NY = (int64_t)tmp.n[1] * ppa, NZ = (int64_t)tmp.n[2] * ppa;
    uint64_t      count = 0;
    auto center = [&](int64_t m, int d) { return (double)tmp.bmin[d] + ((double)m + 0.5) * h; };
    auto inBox  = [&](int b, const double c[3]) {
        for(int d = 0; d < 3; ++d)
            if(c[d] < (double)sp->box_min[b][d] || c[d] > (double)sp->box_max[b][d]) return false;
        return true;
    };
    auto inSph = [&](int s, const double c[3]) {
        double r2 = 0;
        for(int d = 0; d < 3; ++d) {
            double e = c[d] - (double)sp->sphere_center[s][d];
            r2 += e * e;
        }
        return r2 <= (double)sp->sphere_radius[s] * (double)sp->sphere_radius[s];
    };
    const int nobj = sp->n_boxes + sp->n_spheres;
    for(int o = 0; o < nobj; ++o) {
        double lo[3], hi[3];
        for(int d = 0; d < 3; ++d) {
            if(o < sp->n_boxes) {
                lo[d] = sp->box_min[o][d];
                hi[d] = sp->box_max[o][d];
            } else {
                int s = o - sp->n_boxes;
                lo[d] = (double)sp->sphere_center[s][d] - sp->sphere_radius[s];
                hi[d] = (double)sp->sphere_center[s][d] + sp->sphere_radius[s];
            }
        }
        int64_t mlo[3], mhi[3], MN[3] = {NX, NY, NZ};
        for(int d = 0; d < 3; ++d) {
            mlo[d] = std::max<int64_t>(0, (int64_t)std::floor((lo[d] - tmp.bmin[d]) / h - 0.5) - 1);
            mhi[d] = std::min<int64_t>(MN[d] - 1, (int64_t)std::ceil((hi[d] - tmp.bmin[d]) / h - 0.5) + 1);
        }
        for(int64_t mz = mlo[2]; mz <= mhi[2]; ++mz)
            for(int64_t my = mlo[1]; my <= mhi[1]; ++my)
                for(int64_t mx = mlo[0]; mx <= mhi[0]; ++mx) {
                    double c[3] = {center(mx, 0), center(my, 1), center(mz, 2)};
                    bool   in   = o < sp->n_boxes ? inBox(o, c) : inSph(o - sp->n_boxes, c);
                    if(!in) continue;
                    bool claimed = false;
                    for(int q = 0; q < o && !claimed; ++q) claimed = q < sp->n_boxes ? inBox(q, c) : inSph(q - sp->n_boxes, c);
                    if(claimed) continue;
                    if(pos) {
                        uint64_t gid = ((uint64_t)mz * (uint64_t)NY + (uint64_t)my) * (uint64_t)NX + (uint64_t)mx;
                        for(int d = 0; d < 3; ++d) {
                            double u = rand01(seed, gid, (uint32_t)d);
                            pos[count * 3 + d] = (float)(c[d] + (u - 0.5) * (double)sp->jitter * h);
                            double v0 = o < sp->n_boxes ? sp->box_vel[o][d] : sp->sphere_vel[o - sp->n_boxes][d];
                            if(sp->vel_sigma > 0) {
                                double u1 = rand01(seed, gid, 3u + 2u * d), u2 = rand01(seed, gid, 4u + 2u * d);
                                double g  = std::sqrt(-2.0 * std::log(1.0 - u1)) * std::cos(6.283185307179586 * u2);
                                v0 += (double)sp->vel_sigma * g;
                            }
                            vel[count * 3 + d] = (float)v0;
                        }
                    }
                    ++count;
                }
    }
    return count;
}

void* orc_create(const orc_params* p, int precision, int threads) {
    if(precision == 64) return static_cast<orc::Base*>(new orc::Solver<double>(*p, threads));
    return static_cast<orc::Base*>(new orc::Solver<float>(*p, threads));
}
void orc_destroy(void* h) { delete static_cast<orc::Base*>(h); }

}  // extern "C"

namespace {
template<class F> auto dispatch(void* h, F&& f) {
    auto* b = static_cast<orc::Base*>(h);
    if(auto* s = dynamic_cast<orc::Solver<double>*>(b)) return f(*s);
    return f(*dynamic_cast<orc::Solver<float>*>(b));
}
}  // namespace

extern "C" {

void orc_grid_dims(void* h, int32_t* n) {
    dispatch(h, [&](auto& s) { for(int d = 0; d < 3; ++d) n[d] = s.n[d]; return 0; });
}
void orc_set_precond(void* h, int32_t kind) {
    dispatch(h, [&](auto& s) { s.precond = kind; return 0; });
}
void orc_set_cg(void* h, double tol, uint32_t maxIters) {
    dispatch(h, [&](auto& s) {
        using R = std::decay_t<decltype(s.dx)>;
        s.tolFactor = std::max(R(tol), R(1e-30));
        s.maxIters  = maxIters;
        return 0;
    });
}
void orc_set_particles(void* h, const float* pos, const float* vel, uint64_t n) {
    dispatch(h, [&](auto& s) {
        s.pos.resize(n);
        s.vel.resize(n);
        s.C.assign(n, {});
        for(uint64_t i = 0; i < n; ++i)
            for(int d = 0; d < 3; ++d) {
                s.pos[i][d] = pos[3 * i + d];
                s.vel[i][d] = vel[3 * i + d];
            }
        return 0;
    });
}
// transfer scheme: 0 FLIP, 1 APIC, 2 AFLIP
void orc_set_scheme(void* h, int32_t scheme) {
    dispatch(h, [&](auto& s) {
        s.scheme = scheme;
        return 0;
    });
}
// affine matrices, 9 floats per particle, row d = C_d (the gradient of velocity component d)
void orc_set_affine(void* h, const float* c) {
    dispatch(h, [&](auto& s) {
        using R = std::decay_t<decltype(s.dx)>;
        for(size_t p = 0; p < s.C.size(); ++p)
            for(int d = 0; d < 3; ++d)
                for(int e = 0; e < 3; ++e) s.C[p][d][e] = (R)c[9 * p + 3 * d + e];
        return 0;
    });
}
void orc_get_affine(void* h, float* c) {
    dispatch(h, [&](auto& s) {
        for(size_t p = 0; p < s.C.size(); ++p)
            for(int d = 0; d < 3; ++d)
                for(int e = 0; e < 3; ++e) c[9 * p + 3 * d + e] = (float)s.C[p][d][e];
        return 0;
    });
}
uint64_t orc_num_particles(void* h) {
    return dispatch(h, [&](auto& s) { return (uint64_t)s.pos.size(); });
}
void orc_get_particles(void* h, float* pos, float* vel) {
    dispatch(h, [&](auto& s) {
        for(size_t i = 0; i < s.pos.size(); ++i)
            for(int d = 0; d < 3; ++d) {
                if(pos) pos[3 * i + d] = (float)s.pos[i][d];
                if(vel) vel[3 * i + d] = (float)s.vel[i][d];
            }
        return 0;
    });
}
// copies a grid field as double (R fields), int8 (flags) or uint32 (active index); returns count
uint64_t orc_get_field(void* h, int32_t f, void* dst) {
    return dispatch(h, [&](auto& s) -> uint64_t {
        using R = std::decay_t<decltype(s.dx)>;
        size_t count;
        int    kind;
        void*  p = s.template fieldPtr<R>(f, count, kind);
        if(!p || !dst) return count;
        if(kind == 0) for(size_t i = 0; i < count; ++i) ((double*)dst)[i] = (double)((R*)p)[i];
        if(kind == 1) std::memcpy(dst, p, count);
        if(kind == 2) std::memcpy(dst, p, count * 4);
        return count;
    });
}
uint64_t orc_set_field(void* h, int32_t f, const void* src) {
    return dispatch(h, [&](auto& s) -> uint64_t {
        using R = std::decay_t<decltype(s.dx)>;
        size_t count;
        int    kind;
        void*  p = s.template fieldPtr<R>(f, count, kind);
        if(!p) return 0;
        if(kind == 0) for(size_t i = 0; i < count; ++i) ((R*)p)[i] = (R)((const double*)src)[i];
        if(kind == 1) std::memcpy(p, src, count);
        if(kind == 2) std::memcpy(p, src, count * 4);
        return count;
    });
}
uint32_t orc_num_active(void* h) {
    return dispatch(h, [&](auto& s) { return s.nActive; });
}
void orc_get_system(void* h, double* rhs, double* pressure) {
    dispatch(h, [&](auto& s) {
        for(uint32_t i = 0; i < s.nActive; ++i) {
            if(rhs) rhs[i] = (double)s.rhs[i];
            if(pressure) pressure[i] = (double)s.pressure[i];
        }
        return 0;
    });
}
// row entries of the assembled matrix, 7 slots per row (col = -1 when unused)
void orc_get_matrix(void* h, int32_t* cols, double* vals) {
    dispatch(h, [&](auto& s) {
        for(uint32_t i = 0; i < s.nActive; ++i) {
            for(int c = 0; c < 7; ++c) { cols[7 * i + c] = -1; vals[7 * i + c] = 0; }
            int c = 0;
            for(const auto& e : s.rows[i]) {
                if(c < 7) { cols[7 * i + c] = (int32_t)e.first; vals[7 * i + c] = (double)e.second; }
                ++c;
            }
        }
        return 0;
    });
}
// IC(0) of the assembled system (PCGSolver.cpp:398-466): the inverse square roots of the modified
// diagonal, one per active cell
void orc_ic0_factor(void* h, double* inv_diag) {
    dispatch(h, [&](auto& s) {
        s.formIC0();
        for(uint32_t i = 0; i < s.nActive; ++i) inv_diag[i] = (double)s.icInvDiag[i];
        return 0;
    });
}
// z = M^-1 r with the selected preconditioner (PCGSolver.cpp:221-269), r and z in active-cell order
// (rounded to the oracle's precision on the way in)
void orc_apply_precond(void* h, const double* r, double* z) {
    dispatch(h, [&](auto& s) {
        using R = std::decay_t<decltype(s.dx)>;
        if(s.precond == 2) s.formIC0();
        if(s.precond == 1) s.formJacobi();
        std::vector<R> in(s.nActive), out;
        for(uint32_t i = 0; i < s.nActive; ++i) in[i] = (R)r[i];
        s.applyPrecond(in, out);
        for(uint32_t i = 0; i < s.nActive; ++i) z[i] = (double)out[i];
        return 0;
    });
}
void orc_pcg_stats(void* h, uint32_t* iters, double* residual, int32_t* success) {
    dispatch(h, [&](auto& s) {
        *iters    = s.outIters;
        *residual = (double)s.outResidual;
        *success  = s.outSuccess ? 1 : 0;
        return 0;
    });
}
int32_t orc_step(void* h, float dt) {
    return dispatch(h, [&](auto& s) { return s.step(dt) ? 0 : 1; });
}
int32_t orc_advance_velocity(void* h, float dt) {
    return dispatch(h, [&](auto& s) {
        using R = std::decay_t<decltype(s.dx)>;
        return s.advanceVelocity((R)dt) ? 0 : 1;
    });
}
int32_t orc_advance_frame(void* h, float* last_substep, uint32_t* substeps) {
    return dispatch(h, [&](auto& s) {
        int r = s.advanceFrame() ? 0 : 1;
        if(last_substep) *last_substep = (float)s.lastSubstep;
        if(substeps) *substeps = s.substepCount;
        return r;
    });
}
// individual stages, named after the reference methods
void orc_correct_particle_positions(void* h, float dt, float stiffness, uint64_t seed) {
    dispatch(h, [&](auto& s) {
        using R = std::decay_t<decltype(s.dx)>;
        s.correctParticlePositions((R)dt, (R)stiffness, seed);
        return 0;
    });
}
void orc_move_particles(void* h, float dt) { dispatch(h, [&](auto& s) { s.moveParticles(dt); return 0; }); }
void orc_map_particles_to_grid(void* h) { dispatch(h, [&](auto& s) { s.mapParticles2Grid(); return 0; }); }
void orc_extrapolate_velocity(void* h) { dispatch(h, [&](auto& s) { s.extrapolateVelocity(); return 0; }); }
void orc_constrain_grid_velocity(void* h) { dispatch(h, [&](auto& s) { s.constrainGridVelocity(); return 0; }); }
void orc_backup_grid_velocity(void* h) { dispatch(h, [&](auto& s) { s.backupGridVelocity(); return 0; }); }
void orc_add_gravity(void* h, float dt) { dispatch(h, [&](auto& s) { s.addGravity(dt); return 0; }); }
void orc_compute_cell_weights(void* h) { dispatch(h, [&](auto& s) { s.computeCellWeights(); return 0; }); }
void orc_compute_fluid_sdf(void* h) { dispatch(h, [&](auto& s) { s.computeFluidSDF(); return 0; }); }
void orc_compute_system(void* h, float dt) { dispatch(h, [&](auto& s) { s.computeSystem(dt); return 0; }); }
int32_t orc_solve_system(void* h) { return dispatch(h, [&](auto& s) { return s.solve() ? 0 : 1; }); }
void orc_update_projected_velocity(void* h, float dt) { dispatch(h, [&](auto& s) { s.updateProjectedVelocity(dt); return 0; }); }
void orc_compute_changes_grid_velocity(void* h) { dispatch(h, [&](auto& s) { s.computeChangesGridVelocity(); return 0; }); }
void orc_map_grid_to_particles(void* h) { dispatch(h, [&](auto& s) { s.mapGrid2Particles(); return 0; }); }
void orc_compute_boundary_sdf(void* h) { dispatch(h, [&](auto& s) { s.computeBoundarySDF(); return 0; }); }
void orc_set_boundary_objects(void* h, const orc_boundary_object* o, uint32_t n) {
    dispatch(h, [&](auto& s) {
        s.setBoundaryObjects(o, n);
        return 0;
    });
}
void orc_advance_scene(void* h, float fraction) {
    dispatch(h, [&](auto& s) {
        s.advanceScene(fraction);
        return 0;
    });
}
void orc_set_finished_frame(void* h, uint32_t f) {
    dispatch(h, [&](auto& s) {
        s.finishedFrame = f;
        return 0;
    });
}
// the signed distance of group g of the boundary objects at the given points (n x 3)
void orc_boundary_sd(void* h, int g, const float* p, uint64_t n, double* out) {
    dispatch(h, [&](auto& s) {
        using Rt = std::decay_t<decltype(s.dx)>;
        for(uint64_t i = 0; i < n; ++i) out[i] = (double)s.groupSD(g, orc::V3<Rt>{Rt(p[3 * i]), Rt(p[3 * i + 1]), Rt(p[3 * i + 2])});
        return 0;
    });
}
float orc_timestep_cfl(void* h) { return dispatch(h, [&](auto& s) { return (float)s.timestepCFL(); }); }

// stable counting sort by linear cell index; out_perm[i] = source index of sorted slot i,
// cell_start has ncells+1 entries (exclusive scan of the per-cell counts)
void orc_bin_particles(const float* pos, uint64_t np, const float* bmin, float dx, const int32_t* n, uint32_t* cell_key,
                       uint32_t* out_perm, uint32_t* cell_start) {
    const uint64_t C = (uint64_t)n[0] * n[1] * n[2];
    std::vector<uint32_t> count(C + 1, 0);
    for(uint64_t p = 0; p < np; ++p) {
        int c[3];
        for(int d = 0; d < 3; ++d) {
            c[d] = orc::cellIndexAxis(pos[3 * p + d], bmin[d], dx);
            c[d] = std::min(std::max(c[d], 0), n[d] - 1);  // Grid::getValidCellIdx clamp (Grid.h:164-176)
        }
        uint32_t key = (uint32_t)(((uint64_t)c[2] * n[1] + c[1]) * n[0] + c[0]);
        cell_key[p]  = key;
        ++count[key];
    }
    uint32_t acc = 0;
    for(uint64_t c = 0; c < C; ++c) {
        cell_start[c] = acc;
        acc += count[c];
    }
    cell_start[C] = acc;
    std::vector<uint32_t> cur(cell_start, cell_start + C);
    for(uint64_t p = 0; p < np; ++p) out_perm[cur[cell_key[p]]++] = (uint32_t)p;
}
void orc_cell_index(const float* pos, uint64_t np, const float* bmin, float dx, int32_t* out_ijk) {
    for(uint64_t p = 0; p < np; ++p)
        for(int d = 0; d < 3; ++d) out_ijk[3 * p + d] = orc::cellIndexAxis(pos[3 * p + d], bmin[d], dx);
}
void orc_morton_keys(const float* pos, uint64_t np, float inv_cell_size, uint64_t* keys) {
    for(uint64_t p = 0; p < np; ++p) keys[p] = orc::mortonKeyOracle(pos + 3 * p, inv_cell_size);
}
// NeighborSearch::z_sort permutation made stable (ties by ascending index)
void orc_morton_sort(const float* pos, uint64_t np, float inv_cell_size, uint32_t* perm) {
    std::vector<uint64_t> k(np);
    for(uint64_t p = 0; p < np; ++p) k[p] = orc::mortonKeyOracle(pos + 3 * p, inv_cell_size);
    for(uint64_t p = 0; p < np; ++p) perm[p] = (uint32_t)p;
    std::stable_sort(perm, perm + np, [&](uint32_t a, uint32_t b) { return k[a] < k[b]; });
}

}  // extern "C"
