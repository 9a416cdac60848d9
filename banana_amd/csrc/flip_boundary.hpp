// flip_boundary.hpp -- boundary objects on the host: the domain box plus the scene's additional
// boundary objects (box or sphere geometry, a static transform, optional key-frame animation), and
// the device set (BndSet, flip_device.hpp) they turn into for the current frame.
//
// Reference: ParticleSolver::generateBoundaries (ParticleSolver.cpp:242-298: static objects merged
// into one CSG union, dynamic ones kept apart), GeometryObject transforms (GeometryObjects.cpp:
// 93-245: translate * rotate(axis, angle) * uniform scale, the inverse applied to query points,
// distances scaled by the uniform scale), Animation (Animation.h: per-channel splines over the key
// frames, periodic wrap) and its CubicSpline (CubicSpline.h: cubic with zero end slopes, or linear).
#pragma once

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <stdexcept>
#include <vector>

#include "../../include/bnn_flip.h"
#include "flip_device.hpp"

namespace bnn {

// one channel of a key-frame animation: cubic interpolation with first derivative 0 at both ends
// (a tridiagonal system for the quadratic coefficients b), or piecewise linear; outside the knots a
// quadratic continues each end (b at the end knot, the slope there)
class Spline1 {
public:
    void set(const std::vector<double>& x, const std::vector<double>& y, bool cubic) {
        X = x;
        Y = y;
        const size_t n = X.size();
        a.assign(n, 0.0);
        b.assign(n, 0.0);
        c.assign(n, 0.0);
        if(n < 2) return;
        if(cubic) {
            // rows i = 1 .. n-2: (h_{i-1}/3) b_{i-1} + (2/3)(h_{i-1} + h_i) b_i + (h_i/3) b_{i+1} = s_i - s_{i-1}
            // ends (slope 0): 2 h_0 b_0 + h_0 b_1 = 3 s_0 ;  h_{n-2} b_{n-2} + 2 h_{n-2} b_{n-1} = -3 s_{n-2}
            std::vector<double> lo(n, 0.0), di(n, 0.0), up(n, 0.0), r(n, 0.0);
            auto                h = [&](size_t i) { return X[i + 1] - X[i]; };
            auto                sl = [&](size_t i) { return (Y[i + 1] - Y[i]) / h(i); };
            for(size_t i = 1; i + 1 < n; ++i) {
                lo[i] = h(i - 1) / 3.0;
                di[i] = 2.0 * (X[i + 1] - X[i - 1]) / 3.0;
                up[i] = h(i) / 3.0;
                r[i]  = sl(i) - sl(i - 1);
            }
            di[0]     = 2.0 * h(0);
            up[0]     = h(0);
            r[0]      = 3.0 * sl(0);
            lo[n - 1] = h(n - 2);
            di[n - 1] = 2.0 * h(n - 2);
            r[n - 1]  = -3.0 * sl(n - 2);
            // Thomas algorithm
            for(size_t i = 1; i < n; ++i) {
                const double m = lo[i] / di[i - 1];
                di[i] -= m * up[i - 1];
                r[i] -= m * r[i - 1];
            }
            b[n - 1] = r[n - 1] / di[n - 1];
            for(size_t i = n - 1; i-- > 0;) b[i] = (r[i] - up[i] * b[i + 1]) / di[i];
            for(size_t i = 0; i + 1 < n; ++i) {
                a[i] = (b[i + 1] - b[i]) / (3.0 * h(i));
                c[i] = sl(i) - (2.0 * b[i] + b[i + 1]) * h(i) / 3.0;
            }
        } else {
            for(size_t i = 0; i + 1 < n; ++i) c[i] = (Y[i + 1] - Y[i]) / (X[i + 1] - X[i]);
        }
        b0 = b[0];
        c0 = c[0];
        // right of the last knot: b[n-1] (the end condition's curvature, 0 when linear) and the slope
        // of the last piece there
        const double hl = X[n - 1] - X[n - 2];
        c[n - 1]        = 3.0 * a[n - 2] * hl * hl + 2.0 * b[n - 2] * hl + c[n - 2];
        a[n - 1]        = 0.0;
    }
    double operator()(double x) const {
        const size_t n = X.size();
        if(n == 1) return Y[0];
        size_t k = (size_t)(std::lower_bound(X.begin(), X.end(), x) - X.begin());
        k        = k > 0 ? k - 1 : 0;
        const double h = x - X[k];
        if(x < X[0]) return (b0 * h + c0) * h + Y[0];
        if(x > X[n - 1]) return (b[n - 1] * h + c[n - 1]) * h + Y[n - 1];
        return ((a[k] * h + b[k]) * h + c[k]) * h + Y[k];
    }

private:
    std::vector<double> X, Y, a, b, c;
    double              b0 = 0, c0 = 0;
};

struct KeyFrameD {
    uint32_t frame = 0;
    double   t[3]  = {0, 0, 0};
    double   r[4]  = {1, 1, 1, 0};  // axis (any length), angle in radians
    double   s     = 1;
};

class BoundaryObjectHost {
public:
    int32_t geometry = BNN_GEOMETRY_BOX;
    float   lo[3] = {-1, -1, -1}, hi[3] = {1, 1, 1};

    // from the ABI description (angles in degrees there, as in the scene files)
    void load(const bnn_boundary_object& o) {
        if(o.geometry != BNN_GEOMETRY_BOX && o.geometry != BNN_GEOMETRY_SPHERE)
            throw std::invalid_argument("boundary object geometry must be a box or a sphere");
        if(o.n_keyframes > BNN_MAX_KEYFRAMES) throw std::invalid_argument("too many key frames");
        geometry = o.geometry;
        for(int d = 0; d < 3; ++d) {
            lo[d] = o.box_min[d];
            hi[d] = o.box_max[d];
        }
        keys.assign(1, KeyFrameD{});
        KeyFrameD& k0 = keys[0];
        for(int d = 0; d < 3; ++d) k0.t[d] = o.translation[d];
        if(o.rotation_axis_angle[3] != 0.f)  // GeometryObject::setRotation ignores a zero angle
            for(int d = 0; d < 4; ++d) k0.r[d] = d < 3 ? o.rotation_axis_angle[d] : o.rotation_axis_angle[3] * (M_PI / 180.0);
        k0.s        = o.has_scale ? o.scale : 1.0;
        if(k0.s == 0.0) throw std::invalid_argument("a boundary object has scale 0");
        transformed = o.has_translation || (o.has_rotation && o.rotation_axis_angle[3] != 0.f) || o.has_scale;
        dynamic     = o.n_keyframes > 0;
        periodic    = o.periodic != 0;
        startFrame  = o.start_frame;
        for(uint32_t i = 0; i < o.n_keyframes; ++i) {
            const bnn_keyframe& f = o.keyframes[i];
            KeyFrameD           k;
            k.frame = f.frame;
            for(int d = 0; d < 3; ++d) k.t[d] = f.translation[d];
            // all four components in radians, the default (1, 1, 1, 0) included (Animation key frames,
            // GeometryObjects.cpp:187-193); only the direction of the axis matters
            const float dflt[4] = {1.f, 1.f, 1.f, 0.f};
            for(int d = 0; d < 4; ++d) k.r[d] = (f.has_rotation ? f.rotation_axis_angle[d] : dflt[d]) * (M_PI / 180.0);
            k.s = f.has_scale ? f.scale : 1.0;
            if(k.s == 0.0) throw std::invalid_argument("a key frame has scale 0");
            if(k.frame == 0) keys[0] = k;
            else keys.push_back(k);
        }
        if(dynamic) makeReady(o.cubic_translation != 0, o.cubic_rotation != 0, o.cubic_scale != 0);
        if(transformed) setTransform(matrixAt(keys[0]), keys[0].s);
    }

    bool isDynamic() const { return dynamic; }

    // GeometryObject::updateTransformation(frame, fraction) for an animated object
    void update(uint32_t frame, double fraction) {
        if(!dynamic) return;
        KeyFrameD k;
        uint32_t  f = frame;
        if(keys.size() == 1 || (periodic && frame < startFrame)) {
            k = keys[0];
        } else {
            if(periodic && f > maxFrame) f = ((f - startFrame) % (maxFrame - startFrame)) + startFrame;
            const double x = (double)f + fraction;
            for(int d = 0; d < 3; ++d) k.t[d] = tr[d](x);
            for(int d = 0; d < 4; ++d) k.r[d] = rot[d](x);
        }
        // Animation::getUniformScale wraps periodic frames by frame % maxFrame
        double s = keys[0].s;
        if(keys.size() > 1) {
            uint32_t fs = frame;
            if(periodic && fs > maxFrame) fs = fs % maxFrame;
            s = sc((double)fs + fraction);
        }
        k.s = keys.size() == 1 || (periodic && frame < startFrame) ? k.s : sc((double)f + fraction);
        setTransform(matrixAt(k), s);
    }

    BndShape shape() const {
        BndShape b{};
        b.type  = geometry == BNN_GEOMETRY_SPHERE ? 1 : 0;
        b.xform = transformed ? 1 : 0;
        for(int i = 0; i < 12; ++i) b.inv[i] = inv[i];
        b.scale = transformed ? scale : 1.0f;
        for(int d = 0; d < 3; ++d) {
            b.lo[d] = lo[d];
            b.hi[d] = hi[d];
        }
        return b;
    }

private:
    std::vector<KeyFrameD> keys;
    bool                   transformed = false, dynamic = false, periodic = false;
    uint32_t               startFrame = 0, maxFrame = 0;
    Spline1                tr[3], rot[4], sc;
    float                  inv[12] = {1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0};
    float                  scale   = 1.0f;

    void makeReady(bool cubicT, bool cubicR, bool cubicS) {
        std::vector<double> x, t[3], r[4], s;
        for(const KeyFrameD& k : keys) {
            if(periodic && k.frame < startFrame) continue;
            maxFrame = std::max(maxFrame, k.frame);
            x.push_back((double)k.frame);
            for(int d = 0; d < 3; ++d) t[d].push_back(k.t[d]);
            for(int d = 0; d < 4; ++d) r[d].push_back(k.r[d]);
            s.push_back(k.s);
        }
        if(x.size() < 2) throw std::invalid_argument("an animation needs at least two key frames in use");
        for(size_t i = 1; i < x.size(); ++i)
            if(!(x[i] > x[i - 1])) throw std::invalid_argument("key frames must have increasing frame numbers");
        if(periodic && maxFrame <= startFrame) throw std::invalid_argument("periodic animation: the last key frame must follow StartFrame");
        for(int d = 0; d < 3; ++d) tr[d].set(x, t[d], cubicT);
        for(int d = 0; d < 4; ++d) rot[d].set(x, r[d], cubicR);
        sc.set(x, s, cubicS);
    }

    // translate(t) * rotate(angle, normalize(axis)) * scale(s), as a 3 x 4 affine map (double)
    struct Affine3 {
        double m[3][3], t[3];
    };
    static Affine3 matrixAt(const KeyFrameD& k) {
        Affine3      A{};
        const double len = std::sqrt(k.r[0] * k.r[0] + k.r[1] * k.r[1] + k.r[2] * k.r[2]);
        const double ax = len > 0 ? k.r[0] / len : 0, ay = len > 0 ? k.r[1] / len : 0, az = len > 0 ? k.r[2] / len : 0;
        const double c = std::cos(k.r[3]), s = std::sin(k.r[3]), oc = 1.0 - c;
        // rotation about the unit axis (Rodrigues), row-major
        const double R[3][3] = {{c + oc * ax * ax, oc * ax * ay - s * az, oc * ax * az + s * ay},
                                {oc * ay * ax + s * az, c + oc * ay * ay, oc * ay * az - s * ax},
                                {oc * az * ax - s * ay, oc * az * ay + s * ax, c + oc * az * az}};
        for(int i = 0; i < 3; ++i) {
            for(int j = 0; j < 3; ++j) A.m[i][j] = R[i][j] * k.s;
            A.t[i] = k.t[i];
        }
        return A;
    }
    // the inverse of p -> M p + t with M = s R: p' = (R^T / s)(p - t)
    void setTransform(const Affine3& A, double s) {
        double ms = 0;  // the matrix's own scale (|first column|), for the inverse
        for(int i = 0; i < 3; ++i) ms += A.m[i][0] * A.m[i][0];
        ms = std::sqrt(ms);
        for(int i = 0; i < 3; ++i) {
            double acc = 0;
            for(int j = 0; j < 3; ++j) {
                const double v = A.m[j][i] / (ms * ms);  // (s R)^T / s^2 = R^T / s
                inv[4 * i + j] = (float)v;
                acc -= v * A.t[j];
            }
            inv[4 * i + 3] = (float)acc;
        }
        scale       = (float)s;
        transformed = true;
    }
};

// the device set for the domain box and the additional objects at their current transforms:
// with additional objects, the static ones (the domain box first) form one CSG union group and each
// dynamic object a group of its own (ParticleSolver.cpp:262-287); without, the domain box alone
static inline BndSet make_bnd_set(const float domLo[3], const float domHi[3], const std::vector<BoundaryObjectHost>& objs) {
    BndSet B{};
    BndShape dom{};
    dom.type  = 0;
    dom.xform = 0;
    dom.scale = 1.0f;
    for(int d = 0; d < 3; ++d) {
        dom.lo[d] = domLo[d];
        dom.hi[d] = domHi[d];
    }
    if(objs.size() + 1 > (size_t)kMaxBndShapes) throw std::invalid_argument("too many boundary objects");
    B.s[B.nShapes++] = dom;
    for(const auto& o : objs)
        if(!o.isDynamic()) B.s[B.nShapes++] = o.shape();
    B.end[B.nGroups++] = B.nShapes;
    for(const auto& o : objs)
        if(o.isDynamic()) {
            B.s[B.nShapes++]   = o.shape();
            B.end[B.nGroups++] = B.nShapes;
        }
    return B;
}

}  // namespace bnn
