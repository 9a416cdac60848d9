// flip_kernels.hpp -- gfx950 kernels of the FLIP substep (included once, by flip_solver.hip).
//
// Layout in HBM (per scene, addressed through its SceneDesc):
//   particles   SoA float px..vz, grid coordinates gx..gz, uint32 id; cell-sorted after binning
//   MAC faces   u (nx+1,ny,nz), v (nx,ny+1,nz), w (nx,ny,nz+1), x-fastest (Array.h:276)
//   cells       phi (nx,ny,nz) float; cell_start (nx*ny*nz + 1) uint32
//   pressure    8^3-cell bricks holding >= 1 fluid cell, 512 cells contiguous per brick
//               (brick-major, x-fastest inside a brick): vectors, matrix diagonal and the three
//               "upper" couplings (+x,+y,+z); a brick's 6 neighbours come from a slot table.
// Every kernel runs over a work list of (scene, local block) items; one launch serves any number
// of scenes. Elementwise kernels use 256-thread blocks, brick kernels one 512-thread block per brick.
#pragma once

#include "flip_scene.hpp"

namespace bnn {

constexpr int kBlock = 256;

#define BNN_ITEM()                                \
    const WorkItem   it_ = W[blockIdx.x];         \
    const SceneDesc& S   = SD[it_.scene];         \
    const uint32_t   lb  = it_.local;             \
    (void)lb

//--------------------------------------------------------------------------------------------------
// face addressing: component d of the MAC grid, and the flat "all faces" item space
// (blocks of u, then of v, then of w)
//--------------------------------------------------------------------------------------------------
struct FaceView {
    float *  a, *t, *old, *wgt;
    uint8_t *val, *ext;
    int      sx, sy, sz;
};

__device__ __forceinline__ FaceView face_view(const SceneDesc& S, int d) {
    const GridGeom& g = S.g;
    if(d == 0) return FaceView{S.u, S.ut, S.uo, S.wu, S.vu, S.eu, g.nx + 1, g.ny, g.nz};
    if(d == 1) return FaceView{S.v, S.vt, S.vo, S.wv, S.vv, S.ev, g.nx, g.ny + 1, g.nz};
    return FaceView{S.w, S.wt, S.wo, S.ww, S.vw, S.ew, g.nx, g.ny, g.nz + 1};
}

__device__ __forceinline__ uint32_t face_blocks(const GridGeom& g, int d) {
    const size_t n = (size_t)(g.nx + (d == 0)) * (size_t)(g.ny + (d == 1)) * (size_t)(g.nz + (d == 2));
    return (uint32_t)((n + kBlock - 1) / kBlock);
}

// local block of the all-faces space -> component and flat face index
__device__ __forceinline__ void face_item(const GridGeom& g, uint32_t lb, int& d, size_t& t) {
    const uint32_t b0 = face_blocks(g, 0), b1 = face_blocks(g, 1);
    if(lb < b0) {
        d = 0;
    } else if(lb < b0 + b1) {
        d = 1;
        lb -= b0;
    } else {
        d = 2;
        lb -= b0 + b1;
    }
    t = (size_t)lb * kBlock + threadIdx.x;
}

//==================================================================================================
// moveParticles: RK2 through the grid + domain-box projection (PIC_Solver.cpp:353-373, 1137-1144;
// BoundaryObject.cpp:72-90)
//==================================================================================================
__device__ __forceinline__ void vel_from_grid(const SceneDesc& S, float gpx, float gpy, float gpz, float& ux, float& uy, float& uz) {
    const GridGeom& g = S.g;
    ux                = interp3(S.u, g.nx + 1, g.ny, g.nz, gpx, gpy - 0.5f, gpz - 0.5f);
    uy                = interp3(S.v, g.nx, g.ny + 1, g.nz, gpx - 0.5f, gpy, gpz - 0.5f);
    uz                = interp3(S.w, g.nx, g.ny, g.nz + 1, gpx - 0.5f, gpy - 0.5f, gpz);
}

__global__ void __launch_bounds__(kBlock) k_move(const SceneDesc* __restrict__ SD, const WorkItem* __restrict__ W) {
    BNN_ITEM();
    const uint32_t p = lb * kBlock + threadIdx.x;
    if(p >= S.np) return;
    const GridGeom&  g = S.g;
    const Particles& P = S.P[S.pcur];
    float            x = P.px[p], y = P.py[p], z = P.pz[p];
    float            vx = P.vx[p], vy = P.vy[p], vz = P.vz[p];
    bool             changed = false;
    const float      sub     = S.dt / (float)S.advSteps;
    for(uint32_t s = 0; s < S.advSteps; ++s) {
        // trace_rk2
        float ux, uy, uz;
        vel_from_grid(S, (x - g.bminx) / g.dx, (y - g.bminy) / g.dx, (z - g.bminz) / g.dx, ux, uy, uz);
        const float h  = 0.5f * sub;
        const float mx = x + h * ux, my = y + h * uy, mz = z + h * uz;
        vel_from_grid(S, (mx - g.bminx) / g.dx, (my - g.bminy) / g.dx, (mz - g.bminz) / g.dx, ux, uy, uz);
        x = x + sub * ux;
        y = y + sub * uy;
        z = z + sub * uz;
        // constrainToBoundary(ppos, pvel)
        const float phi = box_sd_out(S.dom, x, y, z);
        if(phi < 0) {
            float gx, gy, gz;
            box_grad_out(S.dom, x, y, z, gx, gy, gz);
            const float m2 = gx * gx + gy * gy + gz * gz;
            if(m2 > kFltMin) {
                const float sq = sqrtf(m2);
                gx /= sq;
                gy /= sq;
                gz /= sq;
                x -= phi * gx;
                y -= phi * gy;
                z -= phi * gz;
                if(S.reflect) {
                    const float dn = gx * vx + gy * vy + gz * vz;
                    vx             = (vx - 2.f * dn * gx) * S.reflMult;
                    vy             = (vy - 2.f * dn * gy) * S.reflMult;
                    vz             = (vz - 2.f * dn * gz) * S.reflMult;
                    changed        = true;
                }
            }
        }
    }
    P.px[p] = x;
    P.py[p] = y;
    P.pz[p] = z;
    if(changed) {
        P.vx[p] = vx;
        P.vy[p] = vy;
        P.vz[p] = vz;
    }
}

//==================================================================================================
// Binning: stable counting sort by linear cell index (Grid.cpp:92-114, 283-309)
//==================================================================================================
__device__ __forceinline__ int cell_axis(float p, float bmin, float dx, int n) {
    int c = (int)((p - bmin) / dx);               // Grid.h:156-162 (truncation toward zero)
    return c < 0 ? 0 : (c > n - 1 ? n - 1 : c);  // Grid::getValidCellIdx clamp
}

__global__ void __launch_bounds__(kBlock) k_cell_key(const SceneDesc* __restrict__ SD, const WorkItem* __restrict__ W) {
    BNN_ITEM();
    const uint32_t p = lb * kBlock + threadIdx.x;
    if(p >= S.np) return;
    const GridGeom&  g = S.g;
    const Particles& P = S.P[S.pcur];
    const int        i = cell_axis(P.px[p], g.bminx, g.dx, g.nx);
    const int        j = cell_axis(P.py[p], g.bminy, g.dx, g.ny);
    const int        k = cell_axis(P.pz[p], g.bminz, g.dx, g.nz);
    const uint32_t   c = (uint32_t)lin3(i, j, k, g.nx, g.ny);
    S.key[p]           = c;
    atomicAdd(&S.cellCount[c], 1u);
}

// counts return to zero here (each particle takes one slot), so the count array needs no clearing
__global__ void __launch_bounds__(kBlock) k_scatter(const SceneDesc* __restrict__ SD, const WorkItem* __restrict__ W) {
    BNN_ITEM();
    const uint32_t p = lb * kBlock + threadIdx.x;
    if(p >= S.np) return;
    const uint32_t c    = S.key[p];
    const uint32_t slot = S.cellStart[c] + atomicSub(&S.cellCount[c], 1u) - 1u;
    S.perm[slot]        = p;
}

// make each cell's segment ascending: the counting sort becomes stable (ties by particle index)
__global__ void __launch_bounds__(kBlock) k_segsort(const SceneDesc* __restrict__ SD, const WorkItem* __restrict__ W) {
    BNN_ITEM();
    const uint32_t c = lb * kBlock + threadIdx.x;
    if(c >= (uint32_t)(S.g.nx * S.g.ny * S.g.nz)) return;
    const uint32_t b = S.cellStart[c], e = S.cellStart[c + 1];
    if(e - b < 2) return;
    uint32_t* perm = S.perm;
    for(uint32_t i = b + 1; i < e; ++i) {
        uint32_t v = perm[i];
        uint32_t j = i;
        while(j > b && perm[j - 1] > v) {
            perm[j] = perm[j - 1];
            --j;
        }
        perm[j] = v;
    }
}

// gather into the other particle buffer in sorted order and refresh the grid coordinates
__global__ void __launch_bounds__(kBlock) k_permute(const SceneDesc* __restrict__ SD, const WorkItem* __restrict__ W) {
    BNN_ITEM();
    const uint32_t i = lb * kBlock + threadIdx.x;
    if(i >= S.np) return;
    const GridGeom&  g   = S.g;
    const Particles& in  = S.P[S.pcur];
    const Particles& out = S.P[S.pcur ^ 1u];
    const uint32_t   s   = S.perm[i];
    const float      x = in.px[s], y = in.py[s], z = in.pz[s];
    out.px[i]          = x;
    out.py[i]          = y;
    out.pz[i]          = z;
    out.vx[i]          = in.vx[s];
    out.vy[i]          = in.vy[s];
    out.vz[i]          = in.vz[s];
    out.id[i]          = in.id[s];
    out.gx[i]          = (x - g.bminx) / g.dx;
    out.gy[i]          = (y - g.bminy) / g.dx;
    out.gz[i]          = (z - g.bminz) / g.dx;
}

//--------------------------------------------------------------------------------------------------
// multi-segment exclusive scan of uint32 (3 phases); tile = 256 threads x 8 items. The work item's
// "scene" field indexes the segment table here.
//--------------------------------------------------------------------------------------------------
constexpr int kScanItems = 8;
constexpr int kScanTile  = kBlock * kScanItems;

__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t* sh, uint32_t& total) {
    const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
    uint32_t  x    = v;
#pragma unroll
    for(int o = 1; o < 64; o <<= 1) {
        uint32_t y = __shfl_up(x, o, 64);
        if(lane >= o) x += y;
    }
    if(lane == 63) sh[wid] = x;
    __syncthreads();
    uint32_t off = 0;
    total        = 0;
    for(int w = 0; w < (int)(blockDim.x >> 6); ++w) {
        if(w < wid) off += sh[w];
        total += sh[w];
    }
    __syncthreads();
    return off + x - v;
}

__global__ void __launch_bounds__(kBlock) k_scan_reduce(const ScanSeg* __restrict__ segs, const WorkItem* __restrict__ W,
                                                        uint32_t* __restrict__ part) {
    __shared__ uint32_t sh[kBlock / 64];
    const WorkItem      it   = W[blockIdx.x];
    const ScanSeg&      sg   = segs[it.scene];
    const size_t        base = (size_t)it.local * kScanTile + (size_t)threadIdx.x * kScanItems;
    uint32_t            s    = 0;
#pragma unroll
    for(int i = 0; i < kScanItems; ++i)
        if(base + i < sg.n) s += sg.in[base + i];
    uint32_t total;
    block_exclusive_scan(s, sh, total);
    if(threadIdx.x == 0) part[sg.partBase + it.local] = total;
}

// one block per segment: exclusive scan of its tile partials, in chunks of 256
__global__ void __launch_bounds__(kBlock) k_scan_partials(const ScanSeg* __restrict__ segs, uint32_t* __restrict__ part) {
    __shared__ uint32_t sh[kBlock / 64];
    const ScanSeg&      sg    = segs[blockIdx.x];
    const uint32_t      np    = (sg.n + kScanTile - 1) / kScanTile;
    uint32_t*           p     = part + sg.partBase;
    uint32_t            carry = 0;
    for(uint32_t b = 0; b < np; b += kBlock) {
        const uint32_t i = b + threadIdx.x;
        const uint32_t v = i < np ? p[i] : 0u;
        uint32_t       tot;
        const uint32_t ex = block_exclusive_scan(v, sh, tot);
        if(i < np) p[i] = carry + ex;
        carry += tot;
    }
    if(threadIdx.x == 0 && sg.total) *sg.total = carry;
}

__global__ void __launch_bounds__(kBlock) k_scan_down(const ScanSeg* __restrict__ segs, const WorkItem* __restrict__ W,
                                                      const uint32_t* __restrict__ part) {
    __shared__ uint32_t sh[kBlock / 64];
    const WorkItem      it   = W[blockIdx.x];
    const ScanSeg&      sg   = segs[it.scene];
    const size_t        base = (size_t)it.local * kScanTile + (size_t)threadIdx.x * kScanItems;
    uint32_t            v[kScanItems];
    uint32_t            s = 0;
#pragma unroll
    for(int i = 0; i < kScanItems; ++i) {
        v[i] = base + i < sg.n ? sg.in[base + i] : 0u;
        s += v[i];
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan(s, sh, total) + part[sg.partBase + it.local];
#pragma unroll
    for(int i = 0; i < kScanItems; ++i) {
        if(base + i < sg.n) sg.out[base + i] = run;
        run += v[i];
    }
}

//==================================================================================================
// mapParticles2Grid as a gather over cell-sorted particles (FLIP_Solver.cpp:57-95,
// ArrayHelpers.cpp:50-76). Thread (i,j,k) owns u(i,j,k), v(i,j,k), w(i,j,k); it visits the particles
// of cells [i-1,i+1]x[j-1,j+1]x[k-1,k+1] in sorted order and keeps each particle's exact trilinear
// weight for the face when the face is one of the particle's 8 stencil corners.
//==================================================================================================
__device__ __forceinline__ bool stencil_weight(float q, int i_high, int target, float& wgt) {
    int   i;
    float f;
    bary(q, i_high, i, f);
    if(target == i) {
        wgt = 1.0f - f;
        return true;
    }
    if(target == i + 1) {
        wgt = f;
        return true;
    }
    return false;
}

__global__ void __launch_bounds__(kBlock) k_p2g(const SceneDesc* __restrict__ SD, const WorkItem* __restrict__ W) {
    BNN_ITEM();
    const GridGeom& g  = S.g;
    const uint32_t  t  = lb * kBlock + threadIdx.x;
    const int       ex = g.nx + 1, ey = g.ny + 1, ez = g.nz + 1;
    if(t >= (uint32_t)(ex * ey * ez)) return;
    const Particles& P   = S.P[S.pcur];
    const int        i   = t % ex, j = (t / ex) % ey, k = t / (ex * ey);
    const bool       doU = j < g.ny && k < g.nz;
    const bool       doV = i < g.nx && k < g.nz;
    const bool       doW = i < g.nx && j < g.ny;
    float            mu = 0.f, wu = 0.f, mv = 0.f, wv = 0.f, mw = 0.f, ww = 0.f;
    const int        x0 = max(i - 1, 0), x1 = min(i + 1, g.nx - 1);
    for(int z = max(k - 1, 0); z <= min(k + 1, g.nz - 1); ++z)
        for(int y = max(j - 1, 0); y <= min(j + 1, g.ny - 1); ++y) {
            const size_t   r  = lin3(0, y, z, g.nx, g.ny);
            const uint32_t pb = S.cellStart[r + x0], pe = S.cellStart[r + x1 + 1];
            for(uint32_t p = pb; p < pe; ++p) {
                const float gx = P.gx[p], gy = P.gy[p], gz = P.gz[p];
                if(doU) {
                    float a, b, c;
                    if(stencil_weight(gx, g.nx + 1, i, a) && stencil_weight(gy - 0.5f, g.ny, j, b) &&
                       stencil_weight(gz - 0.5f, g.nz, k, c)) {
                        const float wt = a * b * c;
                        mu += wt * P.vx[p];
                        wu += wt;
                    }
                }
                if(doV) {
                    float a, b, c;
                    if(stencil_weight(gx - 0.5f, g.nx, i, a) && stencil_weight(gy, g.ny + 1, j, b) &&
                       stencil_weight(gz - 0.5f, g.nz, k, c)) {
                        const float wt = a * b * c;
                        mv += wt * P.vy[p];
                        wv += wt;
                    }
                }
                if(doW) {
                    float a, b, c;
                    if(stencil_weight(gx - 0.5f, g.nx, i, a) && stencil_weight(gy - 0.5f, g.ny, j, b) &&
                       stencil_weight(gz, g.nz + 1, k, c)) {
                        const float wt = a * b * c;
                        mw += wt * P.vz[p];
                        ww += wt;
                    }
                }
            }
        }
    if(doU) {
        const size_t id = lin3(i, j, k, g.nx + 1, g.ny);
        const bool   ok = wu > kFltMin;
        S.u[id]         = ok ? mu / wu : 0.f;
        S.vu[id]        = ok ? 1 : 0;
    }
    if(doV) {
        const size_t id = lin3(i, j, k, g.nx, g.ny + 1);
        const bool   ok = wv > kFltMin;
        S.v[id]         = ok ? mv / wv : 0.f;
        S.vv[id]        = ok ? 1 : 0;
    }
    if(doW) {
        const size_t id = lin3(i, j, k, g.nx, g.ny);
        const bool   ok = ww > kFltMin;
        S.w[id]         = ok ? mw / ww : 0.f;
        S.vw[id]        = ok ? 1 : 0;
    }
}

//==================================================================================================
// extrapolateVelocity: one Jacobi layer of all three components (PIC_Solver.cpp:992-1090). Even
// layers read the velocity array and write the temp array, odd layers the reverse; the valid flags
// alternate with the per-component temp flags the same way.
//==================================================================================================
__global__ void __launch_bounds__(kBlock) k_extrap_layer(const SceneDesc* __restrict__ SD, const WorkItem* __restrict__ W, int layer) {
    BNN_ITEM();
    int    d;
    size_t t;
    face_item(S.g, lb, d, t);
    const FaceView F = face_view(S, d);
    const int      sx = F.sx, sy = F.sy, sz = F.sz;
    if(t >= (size_t)sx * sy * sz) return;
    // temp flags of component d live in vtmp at the component's own offset
    const size_t   tmpOff = d == 0 ? 0 : (d == 1 ? (size_t)(S.g.nx + 1) * S.g.ny * S.g.nz
                                                  : (size_t)(S.g.nx + 1) * S.g.ny * S.g.nz + (size_t)S.g.nx * (S.g.ny + 1) * S.g.nz);
    uint8_t*       vt   = S.vtmp + tmpOff;
    const bool     even = (layer & 1) == 0;
    const float*   in   = even ? F.a : F.t;
    float*         out  = even ? F.t : F.a;
    const uint8_t* vin  = even ? F.val : vt;
    uint8_t*       vout = even ? vt : F.val;
    const int      i = (int)(t % sx), j = (int)((t / sx) % sy), k = (int)(t / ((size_t)sx * sy));
    const bool     interior = i >= 1 && i <= sx - 2 && j >= 1 && j <= sy - 2 && k >= 1 && k <= sz - 2;
    const uint8_t  v0       = vin[t];
    if(interior && !v0) {
        const size_t sxy   = (size_t)sx * sy;
        float        sum   = 0.f;
        uint32_t     count = 0;
        if(vin[t + 1]) { sum += in[t + 1]; ++count; }
        if(vin[t - 1]) { sum += in[t - 1]; ++count; }
        if(vin[t + sx]) { sum += in[t + sx]; ++count; }
        if(vin[t - sx]) { sum += in[t - sx]; ++count; }
        if(vin[t + sxy]) { sum += in[t + sxy]; ++count; }
        if(vin[t - sxy]) { sum += in[t - sxy]; ++count; }
        if(count > 0) {
            out[t]   = sum / (float)count;
            vout[t]  = 1;
            F.ext[t] = 1;
            return;
        }
    }
    out[t]  = in[t];
    vout[t] = v0;
    if(layer == 0) F.ext[t] = 0;
}

//==================================================================================================
// constrainGridVelocity, component d (PIC_Solver.cpp:1094-1115, ArrayHelpers.cpp:214-251); writes
// the temp array, which k_face_copy then copies back, so component d sees the updated components < d
//==================================================================================================
__device__ __forceinline__ void interp_gradient(const float* __restrict__ b, int sx, int sy, int sz, float x, float y, float z,
                                                float& gx, float& gy, float& gz) {
    int   i, j, k;
    float fi, fj, fk;
    bary(x, sx, i, fi);
    bary(y, sy, j, fj);
    bary(z, sz, k, fk);
    auto  at   = [&](int a, int c, int e) { return b[lin3(a, c, e, sx, sy)]; };
    float v000 = at(i, j, k), v001 = at(i, j, k + 1), v010 = at(i, j + 1, k), v011 = at(i, j + 1, k + 1);
    float v100 = at(i + 1, j, k), v101 = at(i + 1, j, k + 1), v110 = at(i + 1, j + 1, k), v111 = at(i + 1, j + 1, k + 1);
    gx = bilerpf(v100 - v000, v110 - v010, v101 - v001, v111 - v011, fj, fk);
    gy = bilerpf(v010 - v000, v110 - v100, v011 - v001, v111 - v101, fi, fk);
    gz = bilerpf(v001 - v000, v101 - v100, v011 - v010, v111 - v110, fi, fj);
}

__global__ void __launch_bounds__(kBlock) k_constrain(const SceneDesc* __restrict__ SD, const WorkItem* __restrict__ W, int d) {
    BNN_ITEM();
    if(!S.constrainOn) return;
    const FaceView  F = face_view(S, d);
    const GridGeom& g = S.g;
    const size_t    t = (size_t)lb * kBlock + threadIdx.x;
    if(t >= (size_t)F.sx * F.sy * F.sz) return;
    float val = F.a[t];
    if(F.ext[t] == 1 && F.wgt[t] < kFltMin) {
        const int i = (int)(t % F.sx), j = (int)((t / F.sx) % F.sy), k = (int)(t / ((size_t)F.sx * F.sy));
        float     gp[3] = {(float)i + 0.5f, (float)j + 0.5f, (float)k + 0.5f};
        gp[d]           = d == 0 ? (float)i : (d == 1 ? (float)j : (float)k);
        float vx, vy, vz, nx, ny, nz;
        vel_from_grid(S, gp[0], gp[1], gp[2], vx, vy, vz);
        interp_gradient(S.bsdf, g.nx + 1, g.ny + 1, g.nz + 1, gp[0], gp[1], gp[2], nx, ny, nz);
        const float m2 = nx * nx + ny * ny + nz * nz;
        if(m2 > kFltMin) {
            const float sq = sqrtf(m2);
            nx /= sq;
            ny /= sq;
            nz /= sq;
            const float dd = vx * nx + vy * ny + vz * nz;
            vx -= dd * nx;
            vy -= dd * ny;
            vz -= dd * nz;
            val = d == 0 ? vx : (d == 1 ? vy : vz);
        }
    }
    F.t[t] = val;
}

// copy of component d: temp -> velocity (after k_constrain)
__global__ void __launch_bounds__(kBlock) k_face_restore(const SceneDesc* __restrict__ SD, const WorkItem* __restrict__ W, int d) {
    BNN_ITEM();
    if(!S.constrainOn) return;
    const FaceView F = face_view(S, d);
    const size_t   t = (size_t)lb * kBlock + threadIdx.x;
    if(t >= (size_t)F.sx * F.sy * F.sz) return;
    F.a[t] = F.t[t];
}

// backupGridVelocity (FLIP_Data.h:48-52): old = velocity, all components
__global__ void __launch_bounds__(kBlock) k_backup(const SceneDesc* __restrict__ SD, const WorkItem* __restrict__ W) {
    BNN_ITEM();
    int    d;
    size_t t;
    face_item(S.g, lb, d, t);
    const FaceView F = face_view(S, d);
    if(t >= (size_t)F.sx * F.sy * F.sz) return;
    F.old[t] = F.a[t];
}

//==================================================================================================
// addGravity (PIC_Solver.cpp:505-529), all components. Earth / Directional add a constant (gravAdd
// = g*dt per component, formed on the host as the reference does). ToCenter / FromCenter: the
// reference walks velocities[0]'s index range for every component (PIC_Solver.cpp:521), so a face of
// component d is updated only when its (i,j,k) lies inside u's index range.
//==================================================================================================
__global__ void __launch_bounds__(kBlock) k_gravity(const SceneDesc* __restrict__ SD, const WorkItem* __restrict__ W) {
    BNN_ITEM();
    if(!S.applyGravity) return;
    const GridGeom& g = S.g;
    int             d;
    size_t          t;
    face_item(g, lb, d, t);
    const FaceView F = face_view(S, d);
    if(t >= (size_t)F.sx * F.sy * F.sz) return;
    if(S.gravityType == 0) {
        if(d == 1) F.a[t] -= S.gravAdd[1];
        return;
    }
    if(S.gravityType == 1) {
        F.a[t] += S.gravAdd[d];
        return;
    }
    const int i = (int)(t % F.sx), j = (int)((t / F.sx) % F.sy), k = (int)(t / ((size_t)F.sx * F.sy));
    if(i >= g.nx + 1 || j >= g.ny || k >= g.nz) return;
    float gp[3] = {(float)i + 0.5f, (float)j + 0.5f, (float)k + 0.5f};
    gp[d]       = d == 0 ? (float)i : (d == 1 ? (float)j : (float)k);
    const float wx = gp[0] * g.dx + g.bminx, wy = gp[1] * g.dx + g.bminy, wz = gp[2] * g.dx + g.bminz;
    const int   toC = S.gravityType == 2;
    float       dv[3];
    dv[0]          = toC ? S.gravCenter[0] - wx : wx - S.gravCenter[0];
    dv[1]          = toC ? S.gravCenter[1] - wy : wy - S.gravCenter[1];
    dv[2]          = toC ? S.gravCenter[2] - wz : wz - S.gravCenter[2];
    const float is = 1.0f / sqrtf(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
    F.a[t] += (9.81f * (dv[d] * is)) * S.dt;
}

//==================================================================================================
// computeCellWeights (PIC_Solver.cpp:543-597); counts faces with weight < FLT_MIN (solid faces)
//==================================================================================================
__global__ void __launch_bounds__(kBlock) k_weights(const SceneDesc* __restrict__ SD, const WorkItem* __restrict__ W) {
    BNN_ITEM();
    const GridGeom& g  = S.g;
    const int       sx = g.nx + 1, sy = g.ny + 1, sz = g.nz + 1;
    const size_t    t  = (size_t)lb * kBlock + threadIdx.x;
    if(t >= (size_t)sx * sy * sz) return;
    const int    i = (int)(t % sx), j = (int)((t / sx) % sy), k = (int)(t / ((size_t)sx * sy));
    const float* b = S.bsdf;
    auto         B = [&](int a, int c, int e) { return b[lin3(a, c, e, sx, sy)]; };
    uint32_t     solid = 0;
    if(j < g.ny && k < g.nz) {
        float w = clamp01f(1.0f - fraction_inside4(B(i, j, k), B(i, j + 1, k), B(i, j, k + 1), B(i, j + 1, k + 1)));
        S.wu[lin3(i, j, k, g.nx + 1, g.ny)] = w;
        solid += w < kFltMin;
    }
    if(i < g.nx && k < g.nz) {
        float w = clamp01f(1.0f - fraction_inside4(B(i, j, k), B(i, j, k + 1), B(i + 1, j, k), B(i + 1, j, k + 1)));
        S.wv[lin3(i, j, k, g.nx, g.ny + 1)] = w;
        solid += w < kFltMin;
    }
    if(i < g.nx && j < g.ny) {
        float w = clamp01f(1.0f - fraction_inside4(B(i, j, k), B(i, j + 1, k), B(i + 1, j, k), B(i + 1, j + 1, k)));
        S.ww[lin3(i, j, k, g.nx, g.ny)] = w;
        solid += w < kFltMin;
    }
    if(solid) atomicAdd(&S.ctrl->counters[kCtrSolid], solid);
}

// computeBoundarySDF with the domain box as the boundary (PIC_Solver.cpp:1158-1168)
__global__ void __launch_bounds__(kBlock) k_boundary_sdf(const SceneDesc* __restrict__ SD, const WorkItem* __restrict__ W) {
    BNN_ITEM();
    const GridGeom& g  = S.g;
    const int       sx = g.nx + 1, sy = g.ny + 1, sz = g.nz + 1;
    const size_t    t  = (size_t)lb * kBlock + threadIdx.x;
    if(t >= (size_t)sx * sy * sz) return;
    const int i = (int)(t % sx), j = (int)((t / sx) % sy), k = (int)(t / ((size_t)sx * sy));
    S.bsdf[t]   = box_sd_out(S.dom, (float)i * g.dx + g.bminx, (float)j * g.dx + g.bminy, (float)k * g.dx + g.bminz);
}

//==================================================================================================
// computeFluidSDF as a gather (PIC_Solver.cpp:601-689): phi(c) = min over particles whose clamped
// 3x3x3 window holds c of |centre(c) - p| - sdfRadius; min is exact, so the result is bit-identical
// to the reference's locked scatter. Then the solid extension.
//==================================================================================================
__global__ void __launch_bounds__(kBlock) k_fluid_sdf(const SceneDesc* __restrict__ SD, const WorkItem* __restrict__ W) {
    BNN_ITEM();
    const GridGeom& g = S.g;
    const uint32_t  t = lb * kBlock + threadIdx.x;
    if(t >= (uint32_t)(g.nx * g.ny * g.nz)) return;
    const Particles& P    = S.P[S.pcur];
    const int        i    = t % g.nx, j = (t / g.nx) % g.ny, k = t / (g.nx * g.ny);
    const float      sx   = ((float)i + 0.5f) * g.dx + g.bminx;
    const float      sy   = ((float)j + 0.5f) * g.dx + g.bminy;
    const float      sz   = ((float)k + 0.5f) * g.dx + g.bminz;
    float            best = g.dx * 3.0f;
    for(int z = max(k - 1, 0); z <= min(k + 1, g.nz - 1); ++z)
        for(int y = max(j - 1, 0); y <= min(j + 1, g.ny - 1); ++y)
            for(int x = max(i - 1, 0); x <= min(i + 1, g.nx - 1); ++x) {
                const size_t   c        = lin3(x, y, z, g.nx, g.ny);
                const uint32_t pb       = S.cellStart[c], pe = S.cellStart[c + 1];
                const bool     boundary = x == 0 || y == 0 || z == 0 || x == g.nx - 1 || y == g.ny - 1 || z == g.nz - 1;
                for(uint32_t p = pb; p < pe; ++p) {
                    const float ppx = P.px[p], ppy = P.py[p], ppz = P.pz[p];
                    if(boundary) {
                        // particles clamped into a boundary cell: use their true (unclamped) window
                        const int ci = (int)((ppx - g.bminx) / g.dx), cj = (int)((ppy - g.bminy) / g.dx),
                                  ck = (int)((ppz - g.bminz) / g.dx);
                        if(i < ci - 1 || i > ci + 1 || j < cj - 1 || j > cj + 1 || k < ck - 1 || k > ck + 1) continue;
                    }
                    const float v = len3f(sx - ppx, sy - ppy, sz - ppz) - g.sdfRadius;
                    if(v < best) best = v;
                }
            }
    if(best < g.halfdx) {
        const int   sxn = g.nx + 1, syn = g.ny + 1;
        auto        B   = [&](int a, int c, int e) { return S.bsdf[lin3(a, c, e, sxn, syn)]; };
        const float s   = 0.125f * (B(i, j, k) + B(i + 1, j, k) + B(i, j + 1, k) + B(i + 1, j + 1, k) + B(i, j, k + 1) +
                                  B(i + 1, j, k + 1) + B(i, j + 1, k + 1) + B(i + 1, j + 1, k + 1));
        if(s < 0) best = -g.halfdx;
    }
    S.phi[t] = best;
}

//==================================================================================================
// Bricks of a level. Level 0: bricks holding a fluid cell; level l > 0: parents of the active
// bricks of level l-1. Slot tables: brickSlot (brick grid -> slot or -1), slotBrick, and the six
// neighbour slots of every active brick.
//==================================================================================================
// one block (512 threads) per fine brick: flag bricks holding a fluid cell, count fluid cells
__global__ void __launch_bounds__(512) k_brick_flags(const SceneDesc* __restrict__ SD, const WorkItem* __restrict__ W) {
    BNN_ITEM();
    const GridGeom&  g  = S.g;
    const LevelDesc& L  = S.mg[0];
    const int        b  = (int)lb;
    const int        bx = b % g.nbx, by = (b / g.nbx) % g.nby, bz = b / (g.nbx * g.nby);
    const int        l  = threadIdx.x;
    const int        i = bx * 8 + (l & 7), j = by * 8 + ((l >> 3) & 7), k = bz * 8 + (l >> 6);
    int              fluid = 0;
    if(i < g.nx && j < g.ny && k < g.nz) fluid = S.phi[lin3(i, j, k, g.nx, g.ny)] < 0.f;
    const int cnt = __syncthreads_count(fluid);
    if(l == 0) {
        L.flag[b] = cnt > 0 ? 1u : 0u;
        if(cnt) atomicAdd(&S.ctrl->counters[kCtrFluid], (uint32_t)cnt);
    }
}

// parents of the active bricks of level lev-1 (thread per slot of level lev-1)
__global__ void __launch_bounds__(kBlock) k_mg_flag_parents(const SceneDesc* __restrict__ SD, const WorkItem* __restrict__ W, int lev) {
    BNN_ITEM();
    const LevelDesc& F = S.mg[lev - 1];
    const LevelDesc& C = S.mg[lev];
    const uint32_t   s = lb * kBlock + threadIdx.x;
    if(s >= F.nSlots) return;
    const int b  = F.slotBrick[s];
    const int bx = b % F.nbx, by = (b / F.nbx) % F.nby, bz = b / (F.nbx * F.nby);
    C.flag[((bz >> 1) * C.nby + (by >> 1)) * C.nbx + (bx >> 1)] = 1u;
}

// slot tables from the scanned flags. Flags of coarse levels are cleared here for their next build.
__global__ void __launch_bounds__(kBlock) k_level_slots(const SceneDesc* __restrict__ SD, const WorkItem* __restrict__ W, int lev) {
    BNN_ITEM();
    const LevelDesc& L = S.mg[lev];
    const uint32_t   b = lb * kBlock + threadIdx.x;
    if(b >= L.nb) return;
    if(L.flag[b]) {
        L.brickSlot[b]         = (int32_t)L.scan[b];
        L.slotBrick[L.scan[b]] = (int32_t)b;
        if(lev > 0) L.flag[b] = 0u;
    } else {
        L.brickSlot[b] = -1;
    }
}

// neighbour slots per active brick: -x, +x, -y, +y, -z, +z
__global__ void __launch_bounds__(kBlock) k_level_nbrs(const SceneDesc* __restrict__ SD, const WorkItem* __restrict__ W, int lev) {
    BNN_ITEM();
    const LevelDesc& L = S.mg[lev];
    const uint32_t   s = lb * kBlock + threadIdx.x;
    if(s >= L.nSlots) return;
    const int b  = L.slotBrick[s];
    const int bx = b % L.nbx, by = (b / L.nbx) % L.nby, bz = b / (L.nbx * L.nby);
    auto      at = [&](int x, int y, int z) -> int32_t {
        if(x < 0 || y < 0 || z < 0 || x >= L.nbx || y >= L.nby || z >= L.nbz) return -1;
        return L.brickSlot[(z * L.nby + y) * L.nbx + x];
    };
    L.nbr[6 * s + 0] = at(bx - 1, by, bz);
    L.nbr[6 * s + 1] = at(bx + 1, by, bz);
    L.nbr[6 * s + 2] = at(bx, by - 1, bz);
    L.nbr[6 * s + 3] = at(bx, by + 1, bz);
    L.nbr[6 * s + 4] = at(bx, by, bz - 1);
    L.nbr[6 * s + 5] = at(bx, by, bz + 1);
}

// vector index of the neighbour of local cell l in direction dir (-1 when absent)
__device__ __forceinline__ int32_t nbr_index(const int32_t* __restrict__ nbr, int32_t s, int l, int dir) {
    const int lx = l & 7, ly = (l >> 3) & 7, lz = l >> 6;
    switch(dir) {
        case 0: return lx > 0 ? s * 512 + l - 1 : (nbr[6 * s + 0] >= 0 ? nbr[6 * s + 0] * 512 + l + 7 : -1);
        case 1: return lx < 7 ? s * 512 + l + 1 : (nbr[6 * s + 1] >= 0 ? nbr[6 * s + 1] * 512 + l - 7 : -1);
        case 2: return ly > 0 ? s * 512 + l - 8 : (nbr[6 * s + 2] >= 0 ? nbr[6 * s + 2] * 512 + l + 56 : -1);
        case 3: return ly < 7 ? s * 512 + l + 8 : (nbr[6 * s + 3] >= 0 ? nbr[6 * s + 3] * 512 + l - 56 : -1);
        case 4: return lz > 0 ? s * 512 + l - 64 : (nbr[6 * s + 4] >= 0 ? nbr[6 * s + 4] * 512 + l + 448 : -1);
        default: return lz < 7 ? s * 512 + l + 64 : (nbr[6 * s + 5] >= 0 ? nbr[6 * s + 5] * 512 + l - 448 : -1);
    }
}

//==================================================================================================
// computeSystem (PIC_Solver.cpp:693-893), matrix-free form. Per fluid interior cell: rhs (FP64,
// negative divergence), diagonal and the +x/+y/+z couplings (-w*dt when both cells are fluid), in
// FP64 and rounded to FP32. Cells outside [1, n-2] keep an empty row, as in the reference.
//==================================================================================================
__global__ void __launch_bounds__(512) k_assemble(const SceneDesc* __restrict__ SD, const WorkItem* __restrict__ W) {
    BNN_ITEM();
    const GridGeom&  g  = S.g;
    const LevelDesc& L  = S.mg[0];
    const int        s  = (int)lb, l = threadIdx.x;
    const int        b  = L.slotBrick[s];
    const int        bx = b % g.nbx, by = (b / g.nbx) % g.nby, bz = b / (g.nbx * g.nby);
    const int        i = bx * 8 + (l & 7), j = by * 8 + ((l >> 3) & 7), k = bz * 8 + (l >> 6);
    const size_t     o = (size_t)s * 512 + l;
    double           rhs = 0, diag = 0, cx = 0, cy = 0, cz = 0;
    const bool       interior = i >= 1 && j >= 1 && k >= 1 && i <= g.nx - 2 && j <= g.ny - 2 && k <= g.nz - 2;
    if(interior) {
        auto        P  = [&](int a, int c, int e) { return S.phi[lin3(a, c, e, g.nx, g.ny)]; };
        const float pc = P(i, j, k);
        if(pc < 0) {
            const int    ux = g.nx + 1, vy = g.ny + 1;
            const double wr = S.wu[lin3(i + 1, j, k, ux, g.ny)], wl = S.wu[lin3(i, j, k, ux, g.ny)];
            const double wt = S.wv[lin3(i, j + 1, k, g.nx, vy)], wb = S.wv[lin3(i, j, k, g.nx, vy)];
            const double wf = S.ww[lin3(i, j, k + 1, g.nx, g.ny)], wn = S.ww[lin3(i, j, k, g.nx, g.ny)];
            rhs -= wr * (double)S.u[lin3(i + 1, j, k, ux, g.ny)];
            rhs += wl * (double)S.u[lin3(i, j, k, ux, g.ny)];
            rhs -= wt * (double)S.v[lin3(i, j + 1, k, g.nx, vy)];
            rhs += wb * (double)S.v[lin3(i, j, k, g.nx, vy)];
            rhs -= wf * (double)S.w[lin3(i, j, k + 1, g.nx, g.ny)];
            rhs += wn * (double)S.w[lin3(i, j, k, g.nx, g.ny)];
            const double ddt  = (double)S.dt;
            const double pcd  = (double)pc;
            auto         side = [&](float pn, double w, double& coupling, bool upper) {
                const double term = w * ddt;
                if(pn < 0) {
                    diag += term;
                    if(upper) coupling = -term;
                } else {
                    const double th = fraction_inside2(pcd, (double)pn);
                    diag += term / (th > 0.01 ? th : 0.01);
                }
            };
            double dummy = 0;
            side(P(i + 1, j, k), wr, cx, true);
            side(P(i - 1, j, k), wl, dummy, false);
            side(P(i, j + 1, k), wt, cy, true);
            side(P(i, j - 1, k), wb, dummy, false);
            side(P(i, j, k + 1), wf, cz, true);
            side(P(i, j, k - 1), wn, dummy, false);
            // couplings into boundary-layer cells (empty rows) are dropped
            if(i + 1 > g.nx - 2) cx = 0;
            if(j + 1 > g.ny - 2) cy = 0;
            if(k + 1 > g.nz - 2) cz = 0;
        }
    }
    S.rhs64[o] = rhs;
    if(S.refined) {
        S.diag64[o] = diag;
        S.cx64[o]   = cx;
        S.cy64[o]   = cy;
        S.cz64[o]   = cz;
    }
    S.rhs32[o]   = (float)rhs;
    L.diag[o]    = (float)diag;
    L.cx[o]      = (float)cx;
    L.cy[o]      = (float)cy;
    L.cz[o]      = (float)cz;
    L.invdiag[o] = diag > 0 ? (float)(1.0 / diag) : 0.f;
}

//==================================================================================================
// PCG (PCGSolver.cpp:143-199) with FP32 vectors and FP64 scalars. Per-scene scalars and the stop
// flag live in the scene's control block; every kernel returns at once for a scene whose flag is
// set, so the host queues iterations in chunks and reads the flags once per chunk.
// Reductions are two-level and deterministic: each block (one brick) writes one partial at the
// scene's fixed offset (redBase + brick slot); a one-block-per-scene fold combines that scene's
// partials in slot order. Kernel boundaries order the two, so no device-scope fences are needed.
//==================================================================================================
template<int NT> __device__ __forceinline__ double block_sum(double v, double* sh) {
    v = wave_sum(v);
    if((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0;
    if(threadIdx.x == 0)
        for(int i = 0; i < NT / 64; ++i) t += sh[i];
    __syncthreads();
    return t;  // valid in thread 0
}
template<int NT> __device__ __forceinline__ double block_max(double v, double* sh) {
    v = wave_maxd(v);
    if((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = v;
    __syncthreads();
    double t = 0;
    if(threadIdx.x == 0)
        for(int i = 0; i < NT / 64; ++i) t = fmax(t, sh[i]);
    __syncthreads();
    return t;
}

// fold one scene's partials p[redBase, redBase + nSlots) in fixed order with a 1024-thread block;
// the sum (or max) is valid in thread 0
__device__ __forceinline__ double fold(const SceneDesc& S, const double* __restrict__ p, bool isMax, double* sh) {
    const uint32_t b = S.redBase, e = S.redBase + S.mg[0].nSlots;
    double         a = 0;
    for(uint32_t i = b + threadIdx.x; i < e; i += 1024) a = isMax ? fmax(a, p[i]) : a + p[i];
    return isMax ? block_max<1024>(a, sh) : block_sum<1024>(a, sh);
}

// per-scene start of a pressure solve (one thread per listed scene)
__global__ void k_solve_begin(const SceneDesc* __restrict__ SD, const uint32_t* __restrict__ scenes, uint32_t n) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if(t >= n) return;
    const SceneDesc& S  = SD[scenes[t]];
    OuterState&      os = S.ctrl->outer;
    os.done             = 0;
    os.success          = 0;
    os.passes           = 0;
    os.totalIter        = 0;
    os.res              = 0;
    os.bmax             = 0;
    os.tol              = S.refineTol;
    S.ctrl->pcg.done    = 1;
}

// start of an inner PCG: scenes whose outer loop is finished stay stopped
__global__ void k_pcg_reset(const SceneDesc* __restrict__ SD, const uint32_t* __restrict__ scenes, uint32_t n) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if(t >= n) return;
    const SceneDesc& S  = SD[scenes[t]];
    PcgState&        st = S.ctrl->pcg;
    st.done             = S.ctrl->outer.done ? 1u : 0u;
    st.iter             = 0;
    st.success          = 0;
    st.maxIter          = S.maxIter;
    st.relTol           = S.pcgRelTol;
}

// init: x = 0, r = b, s = 0, z = M^-1 r (Jacobi); partials of z.r and max|b|
__global__ void __launch_bounds__(512) k_pcg_init(const SceneDesc* __restrict__ SD, const WorkItem* __restrict__ W,
                                                  double* __restrict__ pa, double* __restrict__ pb) {
    __shared__ double sh[8];
    BNN_ITEM();
    if(S.ctrl->pcg.done) return;
    const size_t o  = (size_t)lb * 512 + threadIdx.x;
    const float  bv = S.rhs32[o];
    S.x[o]          = 0.f;
    S.r[o]          = bv;
    S.s0[o]         = 0.f;
    double rz = 0, m = (double)fabsf(bv);
    if(S.jacobi) {
        const float zv = S.mg[0].invdiag[o] * bv;
        S.z[o]         = zv;
        rz             = (double)zv * (double)bv;
        rz             = block_sum<512>(rz, sh);
    }
    m = block_max<512>(m, sh);
    if(threadIdx.x == 0) {
        if(S.jacobi) pa[S.redBase + lb] = rz;
        pb[S.redBase + lb] = m;
    }
}

__global__ void __launch_bounds__(1024) k_pcg_init_fin(const SceneDesc* __restrict__ SD, const uint32_t* __restrict__ scenes,
                                                       const double* __restrict__ pa, const double* __restrict__ pb) {
    __shared__ double sh[16];
    const SceneDesc&  S  = SD[scenes[blockIdx.x]];
    PcgState*         st = &S.ctrl->pcg;
    if(st->done) return;
    const double A = fold(S, pa, false, sh);
    const double B = fold(S, pb, true, sh);
    if(threadIdx.x != 0) return;
    st->bmax    = (float)B;
    st->res     = (float)B;
    st->iter    = 0;
    st->beta    = 0;
    st->rho     = A;
    st->success = 0;
    st->tol     = st->relTol * (float)B;  // PCGSolver.cpp:160
    if((float)B < kFltMin) {              // PCGSolver.cpp:155-158
        st->done    = 1;
        st->success = 1;
    } else if(A < (double)kFltMin || isnan(A)) {  // PCGSolver.cpp:164-167
        st->done = 1;
    } else if(st->maxIter == 0) {
        st->done = 1;
    }
}

// s_out = z + beta s_in ; q = A s_out ; partial s_out . q  (parity p: s_in = p ? s1 : s0)
__global__ void __launch_bounds__(512) k_pcg_spmv(const SceneDesc* __restrict__ SD, const WorkItem* __restrict__ W, int parity,
                                                  double* __restrict__ pa) {
    __shared__ double sh[8];
    BNN_ITEM();
    if(S.ctrl->pcg.done) return;
    const LevelDesc& L    = S.mg[0];
    const float      beta = (float)S.ctrl->pcg.beta;
    const float*     sin  = parity ? S.s1 : S.s0;
    float*           sout = parity ? S.s0 : S.s1;
    const float*     z    = S.z;
    const int        s = (int)lb, l = threadIdx.x;
    const size_t     o   = (size_t)s * 512 + l;
    auto             Sv  = [&](int32_t idx) { return z[idx] + beta * sin[idx]; };
    const float      sc  = z[o] + beta * sin[o];
    const int32_t    ixm = nbr_index(L.nbr, s, l, 0), ixp = nbr_index(L.nbr, s, l, 1);
    const int32_t    iym = nbr_index(L.nbr, s, l, 2), iyp = nbr_index(L.nbr, s, l, 3);
    const int32_t    izm = nbr_index(L.nbr, s, l, 4), izp = nbr_index(L.nbr, s, l, 5);
    float            acc = 0.f;
    if(izm >= 0) acc += L.cz[izm] * Sv(izm);
    if(iym >= 0) acc += L.cy[iym] * Sv(iym);
    if(ixm >= 0) acc += L.cx[ixm] * Sv(ixm);
    acc += L.diag[o] * sc;
    if(ixp >= 0) acc += L.cx[o] * Sv(ixp);
    if(iyp >= 0) acc += L.cy[o] * Sv(iyp);
    if(izp >= 0) acc += L.cz[o] * Sv(izp);
    sout[o]        = sc;
    S.q[o]         = acc;
    const double d = block_sum<512>((double)sc * (double)acc, sh);
    if(threadIdx.x == 0) pa[S.redBase + lb] = d;
}

// tmp = s.q ; alpha = rho / tmp (PCGSolver.cpp:174-180)
__global__ void __launch_bounds__(1024) k_pcg_alpha(const SceneDesc* __restrict__ SD, const uint32_t* __restrict__ scenes,
                                                    const double* __restrict__ pa) {
    __shared__ double sh[16];
    const SceneDesc&  S  = SD[scenes[blockIdx.x]];
    PcgState*         st = &S.ctrl->pcg;
    if(st->done) return;
    const double A = fold(S, pa, false, sh);
    if(threadIdx.x != 0) return;
    st->tmp = A;
    if(A < (double)kFltMin) {
        st->done    = 1;
        st->success = 1;
        st->iter += 1;
    } else {
        st->alpha = st->rho / A;
    }
}

// x += alpha s ; r -= alpha q ; z = M^-1 r (Jacobi) ; partials z.r and max|r|
__global__ void __launch_bounds__(512) k_pcg_update(const SceneDesc* __restrict__ SD, const WorkItem* __restrict__ W, int parity,
                                                    double* __restrict__ pa, double* __restrict__ pb) {
    __shared__ double sh[8];
    BNN_ITEM();
    if(S.ctrl->pcg.done) return;
    const float  alpha = (float)S.ctrl->pcg.alpha;
    const float* s     = parity ? S.s0 : S.s1;  // the direction k_pcg_spmv just wrote
    const size_t o     = (size_t)lb * 512 + threadIdx.x;
    const float  xv    = S.x[o] + alpha * s[o];
    const float  rv    = S.r[o] + (-alpha) * S.q[o];
    S.x[o]             = xv;
    S.r[o]             = rv;
    double rz = 0, m = (double)fabsf(rv);
    if(S.jacobi) {  // z = M^-1 r fused here; the multigrid preconditioner runs after this kernel
        const float zv = S.mg[0].invdiag[o] * rv;
        S.z[o]         = zv;
        rz             = (double)zv * (double)rv;
        rz             = block_sum<512>(rz, sh);
    }
    m = block_max<512>(m, sh);
    if(threadIdx.x == 0) {
        if(S.jacobi) pa[S.redBase + lb] = rz;
        pb[S.redBase + lb] = m;
    }
}

// partial z.r after the V-cycle wrote z
__global__ void __launch_bounds__(512) k_pcg_dot_zr(const SceneDesc* __restrict__ SD, const WorkItem* __restrict__ W,
                                                    double* __restrict__ pa) {
    __shared__ double sh[8];
    BNN_ITEM();
    if(S.ctrl->pcg.done) return;
    const size_t o = (size_t)lb * 512 + threadIdx.x;
    const double d = block_sum<512>((double)S.z[o] * (double)S.r[o], sh);
    if(threadIdx.x == 0) pa[S.redBase + lb] = d;
}

// res = max|r| ; stop test (PCGSolver.cpp:182-187) ; beta = rz / rho (:190-194)
__global__ void __launch_bounds__(1024) k_pcg_beta(const SceneDesc* __restrict__ SD, const uint32_t* __restrict__ scenes,
                                                   const double* __restrict__ pa, const double* __restrict__ pb) {
    __shared__ double sh[16];
    const SceneDesc&  S  = SD[scenes[blockIdx.x]];
    PcgState*         st = &S.ctrl->pcg;
    if(st->done) return;
    const double A = fold(S, pa, false, sh);
    const double B = fold(S, pb, true, sh);
    if(threadIdx.x != 0) return;
    st->iter += 1;
    st->res = (float)B;
    if((float)B < st->tol) {
        st->done    = 1;
        st->success = 1;
    } else if(st->iter >= st->maxIter) {
        st->done = 1;
    } else {
        st->beta = A / st->rho;
        st->rho  = A;
    }
}

//--------------------------------------------------------------------------------------------------
// FP64 defect correction: r64 = b64 - A64 x64, max|r64|, max|b64|, r32 = (float) r64 (the inner
// right-hand side). The first pass (x64 = 0) also clears x64.
//--------------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(512) k_residual64(const SceneDesc* __restrict__ SD, const WorkItem* __restrict__ W, int first,
                                                    double* __restrict__ pa, double* __restrict__ pb) {
    __shared__ double sh[8];
    BNN_ITEM();
    if(S.ctrl->outer.done) return;
    const LevelDesc& L  = S.mg[0];
    const int        s  = (int)lb, l = threadIdx.x;
    const size_t     o  = (size_t)s * 512 + l;
    double           ax = 0.0;
    if(!first) {
        const double* x   = S.x64;
        const int32_t ixm = nbr_index(L.nbr, s, l, 0), ixp = nbr_index(L.nbr, s, l, 1);
        const int32_t iym = nbr_index(L.nbr, s, l, 2), iyp = nbr_index(L.nbr, s, l, 3);
        const int32_t izm = nbr_index(L.nbr, s, l, 4), izp = nbr_index(L.nbr, s, l, 5);
        if(izm >= 0) ax += S.cz64[izm] * x[izm];
        if(iym >= 0) ax += S.cy64[iym] * x[iym];
        if(ixm >= 0) ax += S.cx64[ixm] * x[ixm];
        ax += S.diag64[o] * x[o];
        if(ixp >= 0) ax += S.cx64[o] * x[ixp];
        if(iyp >= 0) ax += S.cy64[o] * x[iyp];
        if(izp >= 0) ax += S.cz64[o] * x[izp];
    } else {
        S.x64[o] = 0.0;
    }
    const double bv = S.rhs64[o];
    const double rv = bv - ax;
    S.rhs32[o]      = (float)rv;
    const double m  = block_max<512>(fabs(rv), sh);
    const double bm = block_max<512>(fabs(bv), sh);
    if(threadIdx.x == 0) {
        pa[S.redBase + lb] = m;
        pb[S.redBase + lb] = bm;
    }
}

// outer stop test per scene: max|r| <= tol * max|b| (the first pass records max|b|), or the pass cap
__global__ void __launch_bounds__(1024) k_residual64_fin(const SceneDesc* __restrict__ SD, const uint32_t* __restrict__ scenes,
                                                         const double* __restrict__ pa, const double* __restrict__ pb, int first) {
    __shared__ double sh[16];
    const SceneDesc&  S  = SD[scenes[blockIdx.x]];
    OuterState*       os = &S.ctrl->outer;
    if(os->done) return;
    const double M  = fold(S, pa, true, sh);
    const double BM = fold(S, pb, true, sh);
    if(threadIdx.x != 0) return;
    os->res = M;
    if(first) os->bmax = BM;
    if(os->bmax < 1.17549435e-38 || M <= os->tol * os->bmax) {
        os->done    = 1;
        os->success = 1;
    } else if(os->passes >= S.maxOuter) {
        os->done = 1;
    }
}

__global__ void __launch_bounds__(512) k_accumulate64(const SceneDesc* __restrict__ SD, const WorkItem* __restrict__ W) {
    BNN_ITEM();
    if(S.ctrl->outer.done) return;
    const size_t o = (size_t)lb * 512 + threadIdx.x;
    S.x64[o] += (double)S.x[o];
}

// after an inner solve: count the pass and its iterations; an inner solve that made no progress
// ends the outer loop
__global__ void k_outer_next(const SceneDesc* __restrict__ SD, const uint32_t* __restrict__ scenes, uint32_t n) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if(t >= n) return;
    const SceneDesc& S  = SD[scenes[t]];
    OuterState&      os = S.ctrl->outer;
    if(os.done) return;
    os.passes += 1;
    os.totalIter += S.ctrl->pcg.iter;
    if(S.ctrl->pcg.iter == 0) os.done = 1;
}

//==================================================================================================
// Multigrid preconditioner (extension; the reference's IC(0) is sequential). Level l+1 aggregates
// 2x2x2 cells of level l (piecewise-constant prolongation P); its operator is the Galerkin product
// P^T A P, which for a 7-point operator is again 7-point:
//   diag_C = sum_c diag_c + 2 * (couplings between children of C)
//   cxp_C  = sum of the 4 +x couplings leaving C's children with odd local x, etc.
// The V-cycle is symmetric (red-black Gauss-Seidel: red then black before the coarse correction,
// black then red after), so M^-1 is symmetric positive definite and usable inside CG.
//==================================================================================================
__device__ __forceinline__ int32_t level_cell_index(const LevelDesc& L, int i, int j, int k) {
    if(i < 0 || j < 0 || k < 0 || i >= L.nx || j >= L.ny || k >= L.nz) return -1;
    const int32_t s = L.brickSlot[((k >> 3) * L.nby + (j >> 3)) * L.nbx + (i >> 3)];
    return s < 0 ? -1 : s * 512 + ((k & 7) * 8 + (j & 7)) * 8 + (i & 7);
}

// Galerkin coefficients of level lev (one block per brick of level lev, thread = coarse cell)
__global__ void __launch_bounds__(512) k_mg_coarsen(const SceneDesc* __restrict__ SD, const WorkItem* __restrict__ W, int lev) {
    BNN_ITEM();
    const LevelDesc& F = S.mg[lev - 1];
    const LevelDesc& C = S.mg[lev];
    const int        s = (int)lb, l = threadIdx.x;
    const int        b  = C.slotBrick[s];
    const int        bx = b % C.nbx, by = (b / C.nbx) % C.nby, bz = b / (C.nbx * C.nby);
    const int        I = bx * 8 + (l & 7), J = by * 8 + ((l >> 3) & 7), K = bz * 8 + (l >> 6);
    float            d = 0.f, cx = 0.f, cy = 0.f, cz = 0.f, inner = 0.f;
    for(int c = 0; c < 8; ++c) {
        const int     a0 = c & 1, a1 = (c >> 1) & 1, a2 = (c >> 2) & 1;
        const int32_t f  = level_cell_index(F, 2 * I + a0, 2 * J + a1, 2 * K + a2);
        if(f < 0) continue;
        d += F.diag[f];
        if(a0) cx += F.cx[f];
        else inner += F.cx[f];
        if(a1) cy += F.cy[f];
        else inner += F.cy[f];
        if(a2) cz += F.cz[f];
        else inner += F.cz[f];
    }
    d += 2.f * inner;
    const size_t o = (size_t)s * 512 + l;
    C.diag[o]      = d;
    C.cx[o]        = cx;
    C.cy[o]        = cy;
    C.cz[o]        = cz;
    C.invdiag[o]   = d > 0.f ? 1.0f / d : 0.f;
}

__device__ __forceinline__ float level_offdiag_sum(const LevelDesc& L, int s, int l, const float* __restrict__ x) {
    const int32_t ixm = nbr_index(L.nbr, s, l, 0), ixp = nbr_index(L.nbr, s, l, 1);
    const int32_t iym = nbr_index(L.nbr, s, l, 2), iyp = nbr_index(L.nbr, s, l, 3);
    const int32_t izm = nbr_index(L.nbr, s, l, 4), izp = nbr_index(L.nbr, s, l, 5);
    const size_t  o   = (size_t)s * 512 + l;
    float         acc = 0.f;
    if(izm >= 0) acc += L.cz[izm] * x[izm];
    if(iym >= 0) acc += L.cy[iym] * x[iym];
    if(ixm >= 0) acc += L.cx[ixm] * x[ixm];
    if(ixp >= 0) acc += L.cx[o] * x[ixp];
    if(iyp >= 0) acc += L.cy[o] * x[iyp];
    if(izp >= 0) acc += L.cz[o] * x[izp];
    return acc;
}

__device__ __forceinline__ int cell_color(const LevelDesc& L, int s, int l) {
    const int b  = L.slotBrick[s];
    const int bx = b % L.nbx, by = (b / L.nbx) % L.nby, bz = b / (L.nbx * L.nby);
    return ((bx * 8 + (l & 7)) + (by * 8 + ((l >> 3) & 7)) + (bz * 8 + (l >> 6))) & 1;
}

// one red-black half sweep on level lev: cells of `color` get x = (b - sum a x_nbr) / diag. With
// zeroInit the other colour is set to 0 first (x starts at 0 in every V-cycle). Sweeps at or above
// the scene's own sweep count are skipped.
__global__ void __launch_bounds__(512) k_mg_smooth(const SceneDesc* __restrict__ SD, const WorkItem* __restrict__ W, int lev,
                                                   int color, int zeroInit, int sweep) {
    BNN_ITEM();
    if(S.ctrl->pcg.done || sweep >= S.mgSweeps) return;
    const LevelDesc& L = S.mg[lev];
    const int        s = (int)lb, l = threadIdx.x;
    const size_t     o = (size_t)s * 512 + l;
    if(cell_color(L, s, l) != color) {
        if(zeroInit) L.x[o] = 0.f;
        return;
    }
    const float off = zeroInit ? 0.f : level_offdiag_sum(L, s, l, L.x);
    L.x[o]          = (L.b[o] - off) * L.invdiag[o];
}

// r = b - A x on level lev
__global__ void __launch_bounds__(512) k_mg_residual(const SceneDesc* __restrict__ SD, const WorkItem* __restrict__ W, int lev) {
    BNN_ITEM();
    if(S.ctrl->pcg.done) return;
    const LevelDesc& L = S.mg[lev];
    const int        s = (int)lb, l = threadIdx.x;
    const size_t     o = (size_t)s * 512 + l;
    L.r[o]             = L.b[o] - (L.diag[o] * L.x[o] + level_offdiag_sum(L, s, l, L.x));
}

// b of level lev+1 = sum of the residuals of each coarse cell's 8 children (P^T r); one block per
// brick of level lev+1
__global__ void __launch_bounds__(512) k_mg_restrict(const SceneDesc* __restrict__ SD, const WorkItem* __restrict__ W, int lev) {
    BNN_ITEM();
    if(S.ctrl->pcg.done) return;
    const LevelDesc& F   = S.mg[lev];
    const LevelDesc& C   = S.mg[lev + 1];
    const int        s   = (int)lb, l = threadIdx.x;
    const int        b   = C.slotBrick[s];
    const int        bx  = b % C.nbx, by = (b / C.nbx) % C.nby, bz = b / (C.nbx * C.nby);
    const int        I   = bx * 8 + (l & 7), J = by * 8 + ((l >> 3) & 7), K = bz * 8 + (l >> 6);
    float            sum = 0.f;
    for(int c = 0; c < 8; ++c) {
        const int32_t f = level_cell_index(F, 2 * I + (c & 1), 2 * J + ((c >> 1) & 1), 2 * K + ((c >> 2) & 1));
        if(f >= 0) sum += F.r[f];
    }
    C.b[(size_t)s * 512 + l] = sum;
}

// x of level lev += omega * x of level lev+1 at the parent cell (P x_c, scaled)
__global__ void __launch_bounds__(512) k_mg_prolong(const SceneDesc* __restrict__ SD, const WorkItem* __restrict__ W, int lev) {
    BNN_ITEM();
    if(S.ctrl->pcg.done) return;
    const LevelDesc& F  = S.mg[lev];
    const LevelDesc& C  = S.mg[lev + 1];
    const int        s  = (int)lb, l = threadIdx.x;
    const int        b  = F.slotBrick[s];
    const int        bx = b % F.nbx, by = (b / F.nbx) % F.nby, bz = b / (F.nbx * F.nby);
    const int        i = bx * 8 + (l & 7), j = by * 8 + ((l >> 3) & 7), k = bz * 8 + (l >> 6);
    const size_t     o = (size_t)s * 512 + l;
    if(F.invdiag[o] == 0.f) return;  // not an unknown
    const int32_t c = level_cell_index(C, i >> 1, j >> 1, k >> 1);
    if(c >= 0) F.x[o] += S.mgOmega * C.x[c];
}

// coarsest level of each listed scene (its last level, always a single brick): x = 0, then
// `coarsestSweeps` red-black sweeps and as many black-red sweeps with x held in LDS, one thread
// per cell, one block per scene
__global__ void __launch_bounds__(512) k_mg_coarsest(const SceneDesc* __restrict__ SD, const uint32_t* __restrict__ scenes) {
    __shared__ float xs[512];
    const SceneDesc& S = SD[scenes[blockIdx.x]];
    if(S.ctrl->pcg.done) return;
    const LevelDesc& L      = S.mg[S.nLevels - 1];
    const int        sweeps = S.coarsestSweeps;
    const int        l      = threadIdx.x;
    const int        lx = l & 7, ly = (l >> 3) & 7, lz = l >> 6;
    const float      b = L.b[l], inv = L.invdiag[l];
    const float      cxp = L.cx[l], cyp = L.cy[l], czp = L.cz[l];
    const float      cxm = lx > 0 ? L.cx[l - 1] : 0.f, cym = ly > 0 ? L.cy[l - 8] : 0.f, czm = lz > 0 ? L.cz[l - 64] : 0.f;
    const int        color = cell_color(L, 0, l);
    xs[l]                  = 0.f;
    __syncthreads();
    for(int it = 0; it < 2 * sweeps; ++it) {
        for(int half = 0; half < 2; ++half) {
            const int c = (it < sweeps) ? half : 1 - half;
            if(color == c) {
                float off = 0.f;
                if(lz > 0) off += czm * xs[l - 64];
                if(ly > 0) off += cym * xs[l - 8];
                if(lx > 0) off += cxm * xs[l - 1];
                if(lx < 7) off += cxp * xs[l + 1];
                if(ly < 7) off += cyp * xs[l + 8];
                if(lz < 7) off += czp * xs[l + 64];
                xs[l] = (b - off) * inv;
            }
            __syncthreads();
        }
    }
    L.x[l] = xs[l];
}

//==================================================================================================
// updateProjectedVelocity (PIC_Solver.cpp:911-981), one thread per face, all components; faces the
// reference does not update are zeroed with valid = 0 (its serial pass, :973-980). Pressure from
// the FP64 solution (refined solves) or the FP32 one.
//==================================================================================================
template<class PT> __device__ __forceinline__ PT pressure_at(const SceneDesc& S, const PT* x, int i, int j, int k) {
    const GridGeom& g = S.g;
    const int32_t   s = S.mg[0].brickSlot[((k >> 3) * g.nby + (j >> 3)) * g.nbx + (i >> 3)];
    return x[(size_t)s * 512 + ((k & 7) * 8 + (j & 7)) * 8 + (i & 7)];
}

template<class PT>
__device__ __forceinline__ bool project_face(const SceneDesc& S, const FaceView& F, const PT* px, int d, int i, int j, int k, size_t t) {
    const GridGeom& g  = S.g;
    const int       ai = d == 0 ? i : (d == 1 ? j : k);  // index along the face normal
    const int       an = d == 0 ? g.nx : (d == 1 ? g.ny : g.nz);
    if(!(ai > 0 && ai < an && F.wgt[t] > 0.f)) return false;
    const int   li = i - (d == 0), lj = j - (d == 1), lk = k - (d == 2);
    const float pc = S.phi[lin3(i, j, k, g.nx, g.ny)];
    const float pl = S.phi[lin3(li, lj, lk, g.nx, g.ny)];
    if(!(pc < 0 || pl < 0)) return false;
    const PT prc = pc < 0 ? pressure_at<PT>(S, px, i, j, k) : PT(0);
    const PT prl = pl < 0 ? pressure_at<PT>(S, px, li, lj, lk) : PT(0);
    PT       th  = fraction_inside2((PT)pl, (PT)pc);
    th           = th > PT(0.01) ? th : PT(0.01);  // std::max(0.01, th)
    F.a[t]       = (float)((PT)F.a[t] - (PT)S.dt * (prc - prl) / th);
    return true;
}

__global__ void __launch_bounds__(kBlock) k_update_velocity(const SceneDesc* __restrict__ SD, const WorkItem* __restrict__ W) {
    BNN_ITEM();
    int    d;
    size_t t;
    face_item(S.g, lb, d, t);
    const FaceView F = face_view(S, d);
    if(t >= (size_t)F.sx * F.sy * F.sz) return;
    const int i = (int)(t % F.sx), j = (int)((t / F.sx) % F.sy), k = (int)(t / ((size_t)F.sx * F.sy));
    const bool upd = S.useX64 ? project_face<double>(S, F, S.x64, d, i, j, k, t) : project_face<float>(S, F, S.x, d, i, j, k, t);
    if(!upd) F.a[t] = 0.f;
    F.val[t] = upd ? 1 : 0;
}

//==================================================================================================
// mapGrid2Particles with the FLIP blend (FLIP_Solver.cpp:99-133, MathHelpers.h:549-551); the
// velocity change of computeChangesGridVelocity is formed per stencil corner
//==================================================================================================
__global__ void __launch_bounds__(kBlock) k_g2p(const SceneDesc* __restrict__ SD, const WorkItem* __restrict__ W) {
    BNN_ITEM();
    const uint32_t p = lb * kBlock + threadIdx.x;
    if(p >= S.np) return;
    const GridGeom&  g  = S.g;
    const Particles& P  = S.P[S.pcur];
    const float      gx = P.gx[p], gy = P.gy[p], gz = P.gz[p];
    float            ux, uy, uz;
    vel_from_grid(S, gx, gy, gz, ux, uy, uz);
    const float dx = interp3_diff(S.u, S.uo, g.nx + 1, g.ny, g.nz, gx, gy - 0.5f, gz - 0.5f);
    const float dy = interp3_diff(S.v, S.vo, g.nx, g.ny + 1, g.nz, gx - 0.5f, gy, gz - 0.5f);
    const float dz = interp3_diff(S.w, S.wo, g.nx, g.ny, g.nz + 1, gx - 0.5f, gy - 0.5f, gz);
    const float a  = 1.0f - S.ratio;
    P.vx[p]        = a * ux + S.ratio * (P.vx[p] + dx);
    P.vy[p]        = a * uy + S.ratio * (P.vy[p] + dy);
    P.vz[p]        = a * uz + S.ratio * (P.vz[p] + dz);
}

// max |u|, |v|, |w| per scene (timestepCFL, PIC_Solver.cpp:342-349) via atomicMax on the bits
__global__ void __launch_bounds__(kBlock) k_maxabs(const SceneDesc* __restrict__ SD, const WorkItem* __restrict__ W) {
    __shared__ float sh[kBlock / 64];
    BNN_ITEM();
    int    d;
    size_t t;
    face_item(S.g, lb, d, t);
    const FaceView F = face_view(S, d);
    float          m = t < (size_t)F.sx * F.sy * F.sz ? fabsf(F.a[t]) : 0.f;
    m                = wave_max(m);
    if((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    if(threadIdx.x == 0) {
        for(int i = 1; i < kBlock / 64; ++i) m = fmaxf(m, sh[i]);
        atomicMax(&S.ctrl->counters[kCtrMaxVel], __float_as_uint(m));
    }
}

//==================================================================================================
// per-scene control: clear counters, gather control blocks for one readback
//==================================================================================================
__global__ void k_ctrl_zero(const SceneDesc* __restrict__ SD, const uint32_t* __restrict__ scenes, uint32_t n, int first, int count) {
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if(t >= n * (uint32_t)count) return;
    SD[scenes[t / count]].ctrl->counters[first + (int)(t % count)] = 0u;
}

__global__ void k_ctrl_gather(const SceneDesc* __restrict__ SD, const uint32_t* __restrict__ scenes, uint32_t n,
                              uint32_t* __restrict__ out) {
    constexpr uint32_t kWords = sizeof(SceneCtrl) / 4;
    const uint32_t     t      = blockIdx.x * blockDim.x + threadIdx.x;
    if(t >= n * kWords) return;
    const uint32_t* src = reinterpret_cast<const uint32_t*>(SD[scenes[t / kWords]].ctrl);
    out[t]              = src[t % kWords];
}

}  // namespace bnn
