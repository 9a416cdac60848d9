// flip_device.hpp -- device-side data views and exact-order helper math for the FLIP kernels.
//
// Everything that must match the reference bit for bit (cell indices, the fluid SDF and its
// labels, trilinear weights, extrapolation averages) is written with the reference's operation
// order and compiled without FP contraction (-ffp-contract=off), with IEEE division and sqrt.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace bnn {

constexpr float kFltMin = 1.17549435e-38f;  // std::numeric_limits<float>::min() (Tiny<float>, Setup.h:335)
constexpr int   kBrick  = 8;                // pressure bricks are 8x8x8 cells
constexpr int   kBrickCells = 512;

// Grid geometry of one scene (Grid.h:34-232): x-fastest arrays, bMin grown by NExpandCells.
struct GridGeom {
    int   nx, ny, nz;
    float bminx, bminy, bminz;
    float dx, halfdx, sdfRadius;
    int   nbx, nby, nbz;  // bricks per axis
};

// Particle arrays (structure of arrays). gx/gy/gz = (p - bMin) / dx (Grid.h:197), refreshed by the
// binning permute and valid until the particles move.
// Particle arrays, structure of arrays. gp / gv pack what the grid transfers read per particle, in
// the cell-sorted order the binning permute writes them: gp = (grid coordinates (p - bMin) / dx, vx),
// gv = (vy, vz); the velocity arrays stay the state (the copy in gp / gv is refreshed at each binning).
struct Particles {
    float *   px, *py, *pz, *vx, *vy, *vz;
    float4*   gp;
    float2*   gv;
    uint32_t* id;
};

// Domain box used by BoundaryObject::constrainToBoundary (BoundaryObject.cpp:72-90)
struct BoxDev {
    float lo[3], hi[3];
};

__device__ __forceinline__ size_t lin3(int i, int j, int k, int sx, int sy) {
    return ((size_t)k * (size_t)sy + (size_t)j) * (size_t)sx + (size_t)i;
}

// MathHelpers.h:515-529: get_barycentric with clamping
__device__ __forceinline__ void bary(float x, int i_high, int& i, float& f) {
    float s = floorf(x);
    i       = (int)s;
    if(i < 0) {
        i = 0;
        f = 0.f;
    } else if(i > i_high - 2) {
        i = i_high - 2;
        f = 1.f;
    } else {
        f = x - s;
    }
}

// MathHelpers.h:548-573
__device__ __forceinline__ float lerpf(float v0, float v1, float f) { return (1.0f - f) * v0 + f * v1; }
__device__ __forceinline__ float bilerpf(float v00, float v10, float v01, float v11, float fx, float fy) {
    return lerpf(lerpf(v00, v10, fx), lerpf(v01, v11, fx), fy);
}

// ArrayHelpers.cpp:91-102: interpolateValueLinear on an array with dims (sx, sy, sz)
__device__ __forceinline__ float interp3(const float* __restrict__ g, int sx, int sy, int sz, float x, float y, float z) {
    int   i, j, k;
    float fi, fj, fk;
    bary(x, sx, i, fi);
    bary(y, sy, j, fj);
    bary(z, sz, k, fk);
    const size_t b   = lin3(i, j, k, sx, sy);
    const size_t sxy = (size_t)sx * sy;
    float v000 = g[b], v100 = g[b + 1], v010 = g[b + sx], v110 = g[b + sx + 1];
    float v001 = g[b + sxy], v101 = g[b + sxy + 1], v011 = g[b + sxy + sx], v111 = g[b + sxy + sx + 1];
    return lerpf(bilerpf(v000, v100, v010, v110, fi, fj), bilerpf(v001, v101, v011, v111, fi, fj), fk);
}

// Same stencil on the difference (a - b): computeChangesGridVelocity followed by the FLIP
// interpolation (FLIP_Solver.cpp:115-133), each corner difference formed exactly as the
// reference's dVels array would hold it.
__device__ __forceinline__ float interp3_diff(const float* __restrict__ a, const float* __restrict__ o, int sx, int sy, int sz,
                                              float x, float y, float z) {
    int   i, j, k;
    float fi, fj, fk;
    bary(x, sx, i, fi);
    bary(y, sy, j, fj);
    bary(z, sz, k, fk);
    const size_t b   = lin3(i, j, k, sx, sy);
    const size_t sxy = (size_t)sx * sy;
    float v000 = a[b] - o[b], v100 = a[b + 1] - o[b + 1], v010 = a[b + sx] - o[b + sx], v110 = a[b + sx + 1] - o[b + sx + 1];
    float v001 = a[b + sxy] - o[b + sxy], v101 = a[b + sxy + 1] - o[b + sxy + 1];
    float v011 = a[b + sxy + sx] - o[b + sxy + sx], v111 = a[b + sxy + sx + 1] - o[b + sxy + sx + 1];
    return lerpf(bilerpf(v000, v100, v010, v110, fi, fj), bilerpf(v001, v101, v011, v111, fi, fj), fk);
}

// ArrayHelpers.cpp:275-302 (interpolateGradientValue): the gradient of the trilinear interpolant of
// the eight corner values over the cell, the corner terms added in the reference's order, divided
// by the cell size
__device__ __forceinline__ void grad_value8(float v000, float v001, float v010, float v011, float v100, float v101, float v110, float v111,
                                            float fi, float fj, float fk, float dx, float& gx, float& gy, float& gz) {
#pragma clang fp contract(off)
    const float a = 1.0f - fi, b = 1.0f - fj, c = 1.0f - fk;
    gx = v000 * (-b * c) + v001 * (-b * fk) + v010 * (-fj * c) + v011 * (-fj * fk) + v100 * (b * c) + v101 * (b * fk) + v110 * (fj * c) +
         v111 * (fj * fk);
    gy = v000 * (-a * c) + v001 * (-a * fk) + v010 * (a * c) + v011 * (a * fk) + v100 * (-fi * c) + v101 * (-fi * fk) + v110 * (fi * c) +
         v111 * (fi * fk);
    gz = v000 * (-a * b) + v001 * (a * b) + v010 * (-a * fj) + v011 * (a * fj) + v100 * (-fi * b) + v101 * (fi * b) + v110 * (-fi * fj) +
         v111 * (fi * fj);
    gx /= dx;
    gy /= dx;
    gz /= dx;
}

// interpolateGradientValue of grid g (dims sx, sy, sz) at grid coordinate (x, y, z); with o, of the
// difference g - o formed per corner (the dVels array of computeChangesGridVelocity)
__device__ __forceinline__ void interp3_grad(const float* __restrict__ g, const float* __restrict__ o, int sx, int sy, int sz, float x,
                                             float y, float z, float dx, float& gx, float& gy, float& gz) {
    int   i, j, k;
    float fi, fj, fk;
    bary(x, sx, i, fi);
    bary(y, sy, j, fj);
    bary(z, sz, k, fk);
    const size_t b = lin3(i, j, k, sx, sy), sxy = (size_t)sx * sy;
    const size_t i000 = b, i001 = b + sxy, i010 = b + sx, i011 = b + sxy + sx, i100 = b + 1, i101 = b + sxy + 1, i110 = b + sx + 1,
                 i111 = b + sxy + sx + 1;
    if(o) {
        grad_value8(g[i000] - o[i000], g[i001] - o[i001], g[i010] - o[i010], g[i011] - o[i011], g[i100] - o[i100], g[i101] - o[i101],
                    g[i110] - o[i110], g[i111] - o[i111], fi, fj, fk, dx, gx, gy, gz);
    } else {
        grad_value8(g[i000], g[i001], g[i010], g[i011], g[i100], g[i101], g[i110], g[i111], fi, fj, fk, dx, gx, gy, gz);
    }
}

// MathHelpers.h:804-821
template<class T> __device__ __forceinline__ T fraction_inside2(T l, T r) {
    if(l < 0 && r < 0) return T(1.0);
    if(l < 0 && r >= 0) return l / (l - r);
    if(l >= 0 && r < 0) return r / (r - l);
    return T(0);
}

__device__ __forceinline__ void cycle4(float* a) {
    float t = a[0];
    a[0]    = a[1];
    a[1]    = a[2];
    a[2]    = a[3];
    a[3]    = t;
}

// MathHelpers.h:825-895: square cut-cell fraction from four corner SDF values
__device__ inline float fraction_inside4(float phi_bl, float phi_br, float phi_tl, float phi_tr) {
    int   inside = (phi_bl < 0 ? 1 : 0) + (phi_tl < 0 ? 1 : 0) + (phi_br < 0 ? 1 : 0) + (phi_tr < 0 ? 1 : 0);
    float l[4]   = {phi_bl, phi_br, phi_tr, phi_tl};
    if(inside == 4) return 1.0f;
    if(inside == 3) {
        while(l[0] < 0) cycle4(l);
        float side0 = 1.0f - fraction_inside2(l[0], l[3]);
        float side1 = 1.0f - fraction_inside2(l[0], l[1]);
        return 1.0f - 0.5f * side0 * side1;
    }
    if(inside == 2) {
        while(l[0] >= 0 || !(l[1] < 0 || l[2] < 0)) cycle4(l);
        if(l[1] < 0) {
            float sl = fraction_inside2(l[0], l[3]);
            float sr = fraction_inside2(l[1], l[2]);
            return 0.5f * (sl + sr);
        }
        float mid = 0.25f * (l[0] + l[1] + l[2] + l[3]);
        if(mid < 0) {
            float area  = 0.f;
            float side1 = 1.0f - fraction_inside2(l[0], l[3]);
            float side3 = 1.0f - fraction_inside2(l[2], l[3]);
            area += 0.5f * side1 * side3;
            float side2 = 1.0f - fraction_inside2(l[2], l[1]);
            float side0 = 1.0f - fraction_inside2(l[0], l[1]);
            area += 0.5f * side0 * side2;
            return 1.0f - area;
        }
        float area  = 0.f;
        float side0 = fraction_inside2(l[0], l[1]);
        float side1 = fraction_inside2(l[0], l[3]);
        area += 0.5f * side0 * side1;
        float side2 = fraction_inside2(l[2], l[1]);
        float side3 = fraction_inside2(l[2], l[3]);
        area += 0.5f * side2 * side3;
        return area;
    }
    if(inside == 1) {
        while(l[0] >= 0) cycle4(l);
        float side0 = fraction_inside2(l[0], l[3]);
        float side1 = fraction_inside2(l[0], l[1]);
        return 0.5f * side0 * side1;
    }
    return 0.f;
}

__device__ __forceinline__ float clamp01f(float a) { return a < 0.f ? 0.f : (a > 1.0f ? 1.0f : a); }

// std::min / std::max semantics (first argument wins ties, including +0/-0)
__device__ __forceinline__ float smin(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float smax(float a, float b) { return (a < b) ? b : a; }

// glm::length of a 3-vector in glm's dot order: sqrt((x*x + y*y) + z*z)
__device__ __forceinline__ float len3f(float x, float y, float z) { return sqrtf(x * x + y * y + z * z); }

// BoxObject::signedDistance with bNegativeInside = false (GeometryObjects.cpp:246-268)
__device__ __forceinline__ float box_sd_out(const BoxDev& b, float x, float y, float z) {
    const float p[3]   = {x, y, z};
    bool        inside = true;
#pragma unroll
    for(int d = 0; d < 3; ++d)
        if(p[d] < b.lo[d] || p[d] > b.hi[d]) inside = false;
    float mind;
    if(inside) {
        mind = 3.40282347e+38f;
#pragma unroll
        for(int d = 0; d < 3; ++d) mind = smin(mind, smin(p[d] - b.lo[d], b.hi[d] - p[d]));
        mind = -mind;
    } else {
        float cp[3];
#pragma unroll
        for(int d = 0; d < 3; ++d) cp[d] = smax(smin(p[d], b.hi[d]), b.lo[d]);
        mind = len3f(x - cp[0], y - cp[1], z - cp[2]);
    }
    return -mind;
}

// GeometryObject::gradSignedDistance, 3-D branch, dx = 1e-4 (GeometryObjects.cpp:37-85)
__device__ inline void box_grad_out(const BoxDev& b, float x, float y, float z, float& gx, float& gy, float& gz) {
    const float h    = 1e-4f;
    float       v000 = box_sd_out(b, x - h, y - h, z - h);
    float       v001 = box_sd_out(b, x - h, y - h, z + h);
    float       v010 = box_sd_out(b, x - h, y + h, z - h);
    float       v011 = box_sd_out(b, x - h, y + h, z + h);
    float       v100 = box_sd_out(b, x + h, y - h, z - h);
    float       v101 = box_sd_out(b, x + h, y - h, z + h);
    float       v110 = box_sd_out(b, x + h, y + h, z - h);
    float       v111 = box_sd_out(b, x + h, y + h, z + h);
    float ddx00 = v100 - v000, ddx10 = v110 - v010, ddx01 = v101 - v001, ddx11 = v111 - v011;
    float ddy00 = v010 - v000, ddy10 = v110 - v100, ddy01 = v011 - v001, ddy11 = v111 - v101;
    float ddz00 = v001 - v000, ddz10 = v101 - v100, ddz01 = v011 - v010, ddz11 = v111 - v110;
    gx = (((ddx00 + ddx01) + ddx10) + ddx11) * 0.25f;
    gy = (((ddy00 + ddy01) + ddy10) + ddy11) * 0.25f;
    gz = (((ddz00 + ddz01) + ddz10) + ddz11) * 0.25f;
}

//--------------------------------------------------------------------------------------------------
// Boundary objects (BoundaryObject.h:43-48, BoundaryObject.cpp:55-90, GeometryObjects.cpp:37-85,
// 246-266, 379-383, 861-896). A boundary object's signed distance is its geometry's with
// bNegativeInside = false: positive inside the object, where the fluid may be. Static objects are
// merged into one CSG union (min of their distances, ParticleSolver.cpp:262-287); dynamic objects
// follow as groups of their own. A group = shapes [end[g-1], end[g]).
//--------------------------------------------------------------------------------------------------
constexpr int kMaxBndShapes = 8;
struct BndShape {
    int32_t type;     // 0 box (lo, hi in object space), 1 unit sphere
    int32_t xform;    // 1: object space = inv * world (3 x 4, row-major), distances times scale
    float   inv[12];
    float   scale;
    float   lo[3], hi[3];
};
struct BndSet {
    int32_t  nShapes, nGroups;
    int32_t  end[kMaxBndShapes];
    BndShape s[kMaxBndShapes];
};

// BoxObject / SphereObject::signedDistance(p, false) (GeometryObjects.cpp:246-266, 379-383)
__device__ __forceinline__ float bnd_shape_sd(const BndShape& s, float x, float y, float z) {
    if(s.xform) {
        const float* m = s.inv;
        const float  px = ((m[0] * x + m[1] * y) + m[2] * z) + m[3];
        const float  py = ((m[4] * x + m[5] * y) + m[6] * z) + m[7];
        const float  pz = ((m[8] * x + m[9] * y) + m[10] * z) + m[11];
        x = px;
        y = py;
        z = pz;
    }
    float r;
    if(s.type == 0) {
        BoxDev b;
        for(int d = 0; d < 3; ++d) {
            b.lo[d] = s.lo[d];
            b.hi[d] = s.hi[d];
        }
        r = box_sd_out(b, x, y, z);
    } else {
        r = -(len3f(x, y, z) - 1.0f);
    }
    return s.scale * r;
}
// the signed distance of group g: the CSG union of its shapes (CSGObject::signedDistance, Union)
__device__ __forceinline__ float bnd_group_sd(const BndSet& B, int g, float x, float y, float z) {
    const int b = g == 0 ? 0 : B.end[g - 1], e = B.end[g];
    float     sd = bnd_shape_sd(B.s[b], x, y, z);
    for(int i = b + 1; i < e; ++i) sd = smin(sd, bnd_shape_sd(B.s[i], x, y, z));
    return sd;
}
// GeometryObject::gradSignedDistance (GeometryObjects.cpp:37-85), dx = 1e-4, on group g
__device__ inline void bnd_group_grad(const BndSet& B, int g, float x, float y, float z, float& gx, float& gy, float& gz) {
    const float h    = 1e-4f;
    float       v000 = bnd_group_sd(B, g, x - h, y - h, z - h);
    float       v001 = bnd_group_sd(B, g, x - h, y - h, z + h);
    float       v010 = bnd_group_sd(B, g, x - h, y + h, z - h);
    float       v011 = bnd_group_sd(B, g, x - h, y + h, z + h);
    float       v100 = bnd_group_sd(B, g, x + h, y - h, z - h);
    float       v101 = bnd_group_sd(B, g, x + h, y - h, z + h);
    float       v110 = bnd_group_sd(B, g, x + h, y + h, z - h);
    float       v111 = bnd_group_sd(B, g, x + h, y + h, z + h);
    float ddx00 = v100 - v000, ddx10 = v110 - v010, ddx01 = v101 - v001, ddx11 = v111 - v011;
    float ddy00 = v010 - v000, ddy10 = v110 - v100, ddy01 = v011 - v001, ddy11 = v111 - v101;
    float ddz00 = v001 - v000, ddz10 = v101 - v100, ddz01 = v011 - v010, ddz11 = v111 - v110;
    gx = (((ddx00 + ddx01) + ddx10) + ddx11) * 0.25f;
    gy = (((ddy00 + ddy01) + ddy10) + ddy11) * 0.25f;
    gz = (((ddz00 + ddz01) + ddz10) + ddz11) * 0.25f;
}
// BoundaryObject::constrainToBoundary for every group in order (PIC_Solver.cpp:362-366); returns
// true when a reflection changed the velocity
__device__ inline bool bnd_constrain(const BndSet& B, float& x, float& y, float& z, float& vx, float& vy, float& vz, int reflect,
                                     float reflMult) {
    bool changed = false;
    for(int g = 0; g < B.nGroups; ++g) {
        const float phi = bnd_group_sd(B, g, x, y, z);
        if(phi < 0) {
            float gx, gy, gz;
            bnd_group_grad(B, g, x, y, z, gx, gy, gz);
            const float m2 = gx * gx + gy * gy + gz * gz;
            if(m2 > kFltMin) {
                const float s = sqrtf(m2);
                gx /= s;
                gy /= s;
                gz /= s;
                x -= phi * gx;
                y -= phi * gy;
                z -= phi * gz;
                if(reflect) {
                    const float dn = gx * vx + gy * vy + gz * vz;
                    vx             = (vx - 2.f * dn * gx) * reflMult;
                    vy             = (vy - 2.f * dn * gy) * reflMult;
                    vz             = (vz - 2.f * dn * gz) * reflMult;
                    changed        = true;
                }
            }
        }
    }
    return changed;
}

// wave64 reductions
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for(int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ float wave_max(float v) {
#pragma unroll
    for(int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ double wave_maxd(double v) {
#pragma unroll
    for(int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o, 64));
    return v;
}

}  // namespace bnn
