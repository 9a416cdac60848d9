// flip_kernels.hpp -- gfx950 kernels of the FLIP substep (included once, by flip_solver.hip).
// Each kernel is a functor launched through k_one / k_many (flip_launch.hpp).
//
// Layout in HBM (one scene):
//   particles   SoA float px..vz, grid coordinates gx..gz, uint32 id; kept cell-sorted after binning
//   MAC faces   u (nx+1,ny,nz), v (nx,ny+1,nz), w (nx,ny,nz+1), x-fastest (Array.h:276)
//   cells       phi (nx,ny,nz) float; cell_start (nx*ny*nz + 1) uint32
//   pressure    8^3-cell bricks holding >= 1 fluid cell, 512 cells contiguous per brick
//               (brick-major, x-fastest inside a brick): vectors, matrix diagonal and the three
//               "upper" couplings (+x,+y,+z); a brick's 6 neighbours come from a slot table.
#pragma once

#include "flip_device.hpp"
#include "flip_launch.hpp"

namespace bnn {

constexpr int kBlock = 256;

//==================================================================================================
// moveParticles: RK2 through the grid + domain-box projection (PIC_Solver.cpp:353-373, 1137-1144;
// BoundaryObject.cpp:72-90)
//==================================================================================================
struct VelGrid {
    const float *u, *v, *w;
};

// APIC / AFLIP affine matrices, nine arrays in the particles' storage order: c[3 d + e] = C_d[e]
struct Affine {
    float* c[9];
};

__device__ __forceinline__ void vel_from_grid(const VelGrid& G, const GridGeom& g, float gpx, float gpy, float gpz, float& ux,
                                              float& uy, float& uz) {
    ux = interp3(G.u, g.nx + 1, g.ny, g.nz, gpx, gpy - 0.5f, gpz - 0.5f);
    uy = interp3(G.v, g.nx, g.ny + 1, g.nz, gpx - 0.5f, gpy, gpz - 0.5f);
    uz = interp3(G.w, g.nx, g.ny, g.nz + 1, gpx - 0.5f, gpy - 0.5f, gpz);
}

struct k_move {
    static constexpr int kThreads = kBlock;
    static constexpr bool kByAddress = true;
    Particles P;
    uint32_t n;
    VelGrid G;
    GridGeom g;
    BndSet bnd;
    float dt;
    uint32_t advSteps;
    int reflect;
    float reflMult;
    __device__ void operator()(Blk blk) const {
        const uint32_t p = blk.id * blockDim.x + threadIdx.x;
        if(p >= n) return;
        float x = P.px[p], y = P.py[p], z = P.pz[p];
        float vx = P.vx[p], vy = P.vy[p], vz = P.vz[p];
        bool  changed = false;
        const float sub = dt / (float)advSteps;
        for(uint32_t s = 0; s < advSteps; ++s) {
            // trace_rk2
            float ux, uy, uz;
            vel_from_grid(G, g, (x - g.bminx) / g.dx, (y - g.bminy) / g.dx, (z - g.bminz) / g.dx, ux, uy, uz);
            const float h  = 0.5f * sub;
            const float mx = x + h * ux, my = y + h * uy, mz = z + h * uz;
            vel_from_grid(G, g, (mx - g.bminx) / g.dx, (my - g.bminy) / g.dx, (mz - g.bminz) / g.dx, ux, uy, uz);
            x = x + sub * ux;
            y = y + sub * uy;
            z = z + sub * uz;
            // constrainToBoundary(ppos, pvel) of every boundary object
            if(bnd_constrain(bnd, x, y, z, vx, vy, vz, reflect, reflMult)) changed = true;
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
};

// (i, j, k) of the linear index t of an sx x sy x sz array (x fastest), in 32-bit arithmetic: every
// grid array here has fewer than 2^32 entries, and 64-bit division is many instructions on the GPU
__device__ __forceinline__ void unlin3(uint32_t t, int sx, int sy, int& i, int& j, int& k) {
    const uint32_t r = t / (uint32_t)sx;
    i                = (int)(t - r * (uint32_t)sx);
    j                = (int)(r % (uint32_t)sy);
    k                = (int)(r / (uint32_t)sy);
}

//==================================================================================================
// Binning: stable counting sort by linear cell index (Grid.cpp:92-114, 283-309)
//==================================================================================================
__device__ __forceinline__ int cell_axis(float p, float bmin, float dx, int n) {
    int c = (int)((p - bmin) / dx);  // Grid.h:156-162 (truncation toward zero)
    return c < 0 ? 0 : (c > n - 1 ? n - 1 : c);  // Grid::getValidCellIdx clamp
}

// Runs of equal keys inside a wavefront (particles arrive almost cell-sorted from the previous
// substep, so a run usually covers a whole cell): the run's first lane does one atomic for the run.
// Returns the run length in the leader lane, 0 elsewhere; `leader` gets the run's first lane.
__device__ __forceinline__ uint32_t key_run(uint32_t c, bool active, int& leader) {
    const int      lane  = (int)(threadIdx.x & 63);
    const uint32_t prev  = __shfl_up(c, 1, 64);
    const bool     first = active && (lane == 0 || prev != c);
    const uint64_t heads = __ballot(first);
    const uint64_t live  = __ballot(active);
    const uint64_t upTo  = lane == 63 ? ~0ull : ((2ull << lane) - 1ull);
    leader               = 63 - __clzll(heads & upTo);
    if(!first) return 0u;
    const uint64_t after = heads & ~upTo;           // later run heads
    const int      end   = after ? __ffsll((unsigned long long)after) - 1 : 64 - __clzll(live);
    return (uint32_t)(end - lane);
}

struct k_cell_key {
    static constexpr int kThreads = kBlock;
    const float* __restrict__ px;
    const float* __restrict__ py;
    const float* __restrict__ pz;
    uint32_t n;
    GridGeom g;
    uint32_t* __restrict__ key;
    uint32_t* __restrict__ count;
    uint8_t* __restrict__  occ;  // per 8^3 cell brick: 1 when a particle lies in it (cleared before)
    __device__ void operator()(Blk blk) const {
        const uint32_t p      = blk.id * blockDim.x + threadIdx.x;
        const bool     active = p < n;
        uint32_t       c = 0xFFFFFFFFu, brick = 0;
        if(active) {
            const int i = cell_axis(px[p], g.bminx, g.dx, g.nx);
            const int j = cell_axis(py[p], g.bminy, g.dx, g.ny);
            const int k = cell_axis(pz[p], g.bminz, g.dx, g.nz);
            c           = (uint32_t)lin3(i, j, k, g.nx, g.ny);
            brick       = (uint32_t)(((k >> 3) * g.nby + (j >> 3)) * g.nbx + (i >> 3));
            key[p]      = c;
        }
        int            leader;
        const uint32_t len = key_run(c, active, leader);
        if(len) {
            atomicAdd(&count[c], len);
            occ[brick] = 1;
        }
    }
};

// counts return to zero here (each particle takes one slot), so the count array needs no clearing;
// a run's lanes take consecutive slots in lane order
struct k_scatter {
    static constexpr int kThreads = kBlock;
    const uint32_t* __restrict__ key;
    uint32_t n;
    const uint32_t* __restrict__ cell_start;
    uint32_t* __restrict__ count;
    uint32_t* __restrict__ perm;
    __device__ void operator()(Blk blk) const {
        const uint32_t p      = blk.id * blockDim.x + threadIdx.x;
        const bool     active = p < n;
        const uint32_t c      = active ? key[p] : 0xFFFFFFFFu;
        int            leader;
        const uint32_t len  = key_run(c, active, leader);
        uint32_t       left = 0;  // count before the run's subtraction
        if(len) left = atomicSub(&count[c], len);
        left = __shfl(left, leader, 64);
        const uint32_t runLen = __shfl(len, leader, 64);
        if(active) perm[cell_start[c] + left - runLen + (uint32_t)((int)(threadIdx.x & 63) - leader)] = p;
    }
};

// make each cell's segment ascending: the counting sort becomes stable (ties by particle index)
struct k_segsort {
    static constexpr int kThreads = kBlock;
    const uint32_t* __restrict__ cell_start;
    uint32_t ncells;
    uint32_t* __restrict__ perm;
    __device__ void operator()(Blk blk) const {
        const uint32_t c = blk.id * blockDim.x + threadIdx.x;
        if(c >= ncells) return;
        const uint32_t b = cell_start[c], e = cell_start[c + 1];
        if(e - b < 2) return;
        if(e - b <= 16) {
            // up to 16 entries in registers (padding sorts to the end): most cells arrive in order and
            // are left as they are; the others are sorted by odd-even transposition (the entries are
            // distinct particle indices, so the order is the one the insertion sort below gives)
            const uint32_t n = e - b;
            uint32_t       r[16];
#pragma unroll
            for(uint32_t k = 0; k < 16; ++k) r[k] = k < n ? perm[b + k] : 0xFFFFFFFFu;
            bool sorted = true;
#pragma unroll
            for(uint32_t k = 1; k < 16; ++k) sorted &= r[k - 1] <= r[k];
            if(sorted) return;
#pragma unroll
            for(int ph = 0; ph < 16; ++ph) {
#pragma unroll
                for(int k = ph & 1; k + 1 < 16; k += 2) {
                    const uint32_t lo = min(r[k], r[k + 1]), hi = max(r[k], r[k + 1]);
                    r[k]              = lo;
                    r[k + 1]          = hi;
                }
            }
#pragma unroll
            for(uint32_t k = 0; k < 16; ++k)
                if(k < n) perm[b + k] = r[k];
            return;
        }
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
};

struct k_permute {
    static constexpr int kThreads = kBlock;
    Particles in;
    Particles out;
    const uint32_t* __restrict__ perm;
    uint32_t n;
    GridGeom g;
    __device__ void operator()(Blk blk) const {
        const uint32_t i = blk.id * blockDim.x + threadIdx.x;
        if(i >= n) return;
        const uint32_t s = perm[i];
        const float    x = in.px[s], y = in.py[s], z = in.pz[s];
        const float    vx = in.vx[s], vy = in.vy[s], vz = in.vz[s];
        out.px[i]        = x;
        out.py[i]        = y;
        out.pz[i]        = z;
        out.vx[i]        = vx;
        out.vy[i]        = vy;
        out.vz[i]        = vz;
        out.id[i]        = in.id[s];
        out.gp[i]        = make_float4((x - g.bminx) / g.dx, (y - g.bminy) / g.dx, (z - g.bminz) / g.dx, vx);
        out.gv[i]        = make_float2(vy, vz);
    }
};

// the velocities' copy in gp.w / gv refreshed from the velocity arrays (particles already binned)
struct k_pack_velocity {
    static constexpr int kThreads = kBlock;
    Particles P;
    uint32_t n;
    __device__ void operator()(Blk blk) const {
        const uint32_t i = blk.id * blockDim.x + threadIdx.x;
        if(i >= n) return;
        P.gp[i].w = P.vx[i];
        P.gv[i]   = make_float2(P.vy[i], P.vz[i]);
    }
};

//==================================================================================================
// Z-order sort of the particles (PIC_Solver.cpp:315-328 -> NeighborSearch::z_sort,
// NeighborSearch.cpp:45-86). Keys are built in the caller's particle order; a stable LSD radix sort
// on 8-bit digits gives the permutation; the particles' external ids are renamed by it (storage on
// the device stays cell-sorted, so no particle data moves).
//==================================================================================================
// NeighborSearch::cell_index (NeighborSearch.cpp:47-55) and the shift of z_value (:60-70): cell of
// x - 1.2345 through the reciprocal cell size (floor for negative offsets), then minus (INT_MIN + 1).
__device__ __forceinline__ uint64_t zsort_coord(float x, float inv) {
    const float tmp = x - 1.2345f;
    const int   k   = tmp >= 0.f ? (int)(inv * tmp) : (int)(inv * tmp) - 1;
    return (uint64_t)((int64_t)k + 2147483647ll);
}
// bit b of v (b < 21) to bit 3b
__device__ __forceinline__ uint64_t spread_by_3(uint64_t v) {
    v &= 0x1FFFFFull;
    v = (v | v << 32) & 0x001F00000000FFFFull;
    v = (v | v << 16) & 0x001F0000FF0000FFull;
    v = (v | v << 8) & 0x100F00F00F00F00Full;
    v = (v | v << 4) & 0x10C30C30C30C30C3ull;
    v = (v | v << 2) & 0x1249249249249249ull;
    return v;
}
// morton3D_64_encode (Morton3D.h:54-66) keeps, in 64 bits, bit b of coordinate d at bit 3b + d:
// 22 bits of x, 21 of y and z.
__device__ __forceinline__ uint64_t zsort_key(float x, float y, float z, float inv) {
    const uint64_t cx = zsort_coord(x, inv), cy = zsort_coord(y, inv), cz = zsort_coord(z, inv);
    return spread_by_3(cx) | spread_by_3(cy) << 1 | spread_by_3(cz) << 2 | ((cx >> 21) & 1ull) << 63;
}

// keys in external order, split in 32-bit halves; values = external index
struct k_zsort_keys {
    static constexpr int kThreads = kBlock;
    Particles P;
    uint32_t n;
    float inv;
    uint32_t* __restrict__ klo;
    uint32_t* __restrict__ khi;
    uint32_t* __restrict__ val;
    __device__ void operator()(Blk blk) const {
        const uint32_t p = blk.id * blockDim.x + threadIdx.x;
        if(p >= n) return;
        const uint64_t k = zsort_key(P.px[p], P.py[p], P.pz[p], inv);
        const uint32_t e = P.id[p];
        klo[e]           = (uint32_t)k;
        khi[e]           = (uint32_t)(k >> 32);
        val[e]           = e;
    }
};

// AND and OR of all keys (digits on which every key agrees need no pass); acc = {and, or}, preset
// to {~0, 0}
struct k_zsort_andor {
    static constexpr int kThreads = kBlock;
    const uint32_t* __restrict__ klo;
    const uint32_t* __restrict__ khi;
    uint32_t n;
    unsigned long long* __restrict__ acc;
    __device__ void operator()(Blk blk) const {
        __shared__ unsigned long long sa[kBlock / 64], so[kBlock / 64];
        unsigned long long a = ~0ull, o = 0ull;
        for(uint32_t i = blk.id * blockDim.x + threadIdx.x; i < n; i += blk.n * blockDim.x) {
            const unsigned long long k = (unsigned long long)khi[i] << 32 | klo[i];
            a &= k;
            o |= k;
        }
        for(int off = 32; off > 0; off >>= 1) {
            a &= __shfl_xor(a, off, 64);
            o |= __shfl_xor(o, off, 64);
        }
        if((threadIdx.x & 63) == 0) {
            sa[threadIdx.x >> 6] = a;
            so[threadIdx.x >> 6] = o;
        }
        __syncthreads();
        if(threadIdx.x == 0) {
            for(int w = 1; w < kBlock / 64; ++w) {
                a &= sa[w];
                o |= so[w];
            }
            atomicAnd(&acc[0], a);
            atomicOr(&acc[1], o);
        }
    }
};

constexpr int kRsItems = 16;                  // keys per thread per tile
constexpr int kRsTile  = kBlock * kRsItems;   // keys per tile (one workgroup)
static_assert(kBlock == 256, "the radix passes use one thread per digit value");

// digit histogram of one tile, stored digit-major: hist[d * ntiles + tile]
struct k_rs_hist {
    static constexpr int kThreads = kBlock;
    const uint32_t* __restrict__ kw;
    uint32_t n;
    int shift;
    uint32_t ntiles;
    uint32_t* __restrict__ hist;
    __device__ void operator()(Blk blk) const {
        __shared__ uint32_t h[256];
        h[threadIdx.x] = 0;
        __syncthreads();
        const size_t base = (size_t)blk.id * kRsTile;
        for(int c = 0; c < kRsItems; ++c) {
            const size_t i = base + (size_t)c * kBlock + threadIdx.x;
            if(i < n) atomicAdd(&h[(kw[i] >> shift) & 255u], 1u);
        }
        __syncthreads();
        hist[(size_t)threadIdx.x * ntiles + blk.id] = h[threadIdx.x];
    }
};

// stable scatter of one tile: offs = exclusive scan of hist; inside the tile, keys keep their order
// (wave ballots give the rank among equal digits, earlier waves and chunks come first)
struct k_rs_scatter {
    static constexpr int kThreads = kBlock;
    const uint32_t* __restrict__ klo;
    const uint32_t* __restrict__ khi;
    const uint32_t* __restrict__ val;
    uint32_t n;
    int word;
    int shift;
    uint32_t ntiles;
    const uint32_t* __restrict__ offs;
    uint32_t* __restrict__ olo;
    uint32_t* __restrict__ ohi;
    uint32_t* __restrict__ oval;
    __device__ void operator()(Blk blk) const {
        constexpr int       kWaves = kBlock / 64;
        __shared__ uint32_t next[256];
        __shared__ uint32_t wcnt[kWaves][256];
        const int           t = threadIdx.x, lane = t & 63, wv = t >> 6;
        next[t] = offs[(size_t)t * ntiles + blk.id];
        for(int w = 0; w < kWaves; ++w) wcnt[w][t] = 0;
        __syncthreads();
        const uint64_t below = (1ull << lane) - 1ull;
        for(int c = 0; c < kRsItems; ++c) {
            const size_t i   = (size_t)blk.id * kRsTile + (size_t)c * kBlock + t;
            const bool   act = i < n;
            uint32_t     lo = 0, hi = 0, v = 0, d = 0;
            if(act) {
                lo = klo[i];
                hi = khi[i];
                v  = val[i];
                d  = ((word ? hi : lo) >> shift) & 255u;
            }
            uint64_t peers = __ballot(act);
            for(int b = 0; b < 8; ++b) {
                const uint64_t m = __ballot((d >> b) & 1u);
                peers &= ((d >> b) & 1u) ? m : ~m;
            }
            const uint32_t rank = (uint32_t)__popcll(peers & below);
            if(act && rank == 0) wcnt[wv][d] = (uint32_t)__popcll(peers);
            __syncthreads();
            if(act) {
                uint32_t pos = next[d] + rank;
                for(int w = 0; w < wv; ++w) pos += wcnt[w][d];
                olo[pos]  = lo;
                ohi[pos]  = hi;
                oval[pos] = v;
            }
            __syncthreads();
            uint32_t tot = 0;
            for(int w = 0; w < kWaves; ++w) {
                tot += wcnt[w][t];
                wcnt[w][t] = 0;
            }
            next[t] += tot;
            __syncthreads();
        }
    }
};

//==================================================================================================
// PIC_Solver::correctParticlePositions (PIC_Solver.cpp:377-485), 3-D: each particle is pushed away
// from the other particles of the 27 cells around its cell (Grid::getCellIdx, not clamped; cells
// outside the grid skipped) with w = smooth_kernel(d^2, r^2) (MathHelpers.h:661-664, (t*t)*t here),
// spring += w (p - q) / d; a pair closer than the threshold gets a random push (threshold / dt) u
// with u in [-1, 1)^3 from a hash of (seed, p, q, component), p and q being the particles' user
// indices. Then advectionSteps times pos += spring * K followed by the domain-box projection
// (BoundaryObject::constrainToBoundary(ppos), BoundaryObject.cpp:55-67). Neighbours are visited cell
// by cell (k, j, i ascending) and, in a cell, in storage order. Output to separate arrays (the
// neighbours' positions are read while others are written).
//==================================================================================================
__device__ __forceinline__ uint64_t mix64d(uint64_t x) {
    x += 0x9E3779B97F4A7C15ull;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
    return x ^ (x >> 31);
}
__device__ __forceinline__ float rand11(uint64_t seed, uint32_t p, uint32_t q, int comp) {
    const uint64_t h = mix64d(seed ^ mix64d(((uint64_t)p << 32 | q) * 4u + (uint64_t)comp));
    return (float)(uint32_t)(h >> 40) * (1.0f / 8388608.0f) - 1.0f;
}

struct CorrectParams {
    float    radius2, threshold, threshold2, push, K;  // push = threshold / dt
    uint32_t advSteps;
    uint64_t seed;
};

struct k_correct_positions {
    static constexpr int kThreads = kBlock;
    Particles P;
    const uint32_t* __restrict__ cell_start;
    uint32_t n;
    GridGeom g;
    BndSet bnd;
    CorrectParams c;
    float* __restrict__ ox;
    float* __restrict__ oy;
    float* __restrict__ oz;
    __device__ void operator()(Blk blk) const {
    #pragma clang fp contract(off)
        const uint32_t p = blk.id * blockDim.x + threadIdx.x;
        if(p >= n) return;
        const float    x = P.px[p], y = P.py[p], z = P.pz[p];
        const uint32_t pid = P.id[p];
        const int      ci = (int)((x - g.bminx) / g.dx), cj = (int)((y - g.bminy) / g.dx), ck = (int)((z - g.bminz) / g.dx);
        float          sx = 0.f, sy = 0.f, sz = 0.f;
        for(int k = ck - 1; k <= ck + 1; ++k) {
            if(k < 0 || k >= g.nz) continue;
            for(int j = cj - 1; j <= cj + 1; ++j) {
                if(j < 0 || j >= g.ny) continue;
                for(int i = ci - 1; i <= ci + 1; ++i) {
                    if(i < 0 || i >= g.nx) continue;
                    const size_t   cell = lin3(i, j, k, g.nx, g.ny);
                    const uint32_t qb = cell_start[cell], qe = cell_start[cell + 1];
                    for(uint32_t q = qb; q < qe; ++q) {
                        if(q == p) continue;
                        const float dx = x - P.px[q], dy = y - P.py[q], dz = z - P.pz[q];
                        const float d2 = dx * dx + dy * dy + dz * dz;
                        const float x2 = d2 / c.radius2;
                        const float t  = 1.0f - x2;
                        const float w  = x2 > 1.0f ? 0.f : (t * t) * t;
                        if(d2 > c.threshold2) {
                            const float s = sqrtf(d2);
                            sx += w * dx / s;
                            sy += w * dy / s;
                            sz += w * dz / s;
                        } else {
                            const uint32_t qid = P.id[q];
                            sx += c.push * rand11(c.seed, pid, qid, 0);
                            sy += c.push * rand11(c.seed, pid, qid, 1);
                            sz += c.push * rand11(c.seed, pid, qid, 2);
                        }
                    }
                }
            }
        }
        float nx_ = x, ny_ = y, nz_ = z;
        for(uint32_t it = 0; it < c.advSteps; ++it) {
            nx_ += sx * c.K;
            ny_ += sy * c.K;
            nz_ += sz * c.K;
            float vx = 0.f, vy = 0.f, vz = 0.f;  // constrainToBoundary(ppos): no velocity
            bnd_constrain(bnd, nx_, ny_, nz_, vx, vy, vz, 0, 1.f);
        }
        ox[p] = nx_;
        oy[p] = ny_;
        oz[p] = nz_;
    }
};

// identity elements of the AND / OR key accumulators
struct k_zsort_acc_init {
    static constexpr int kThreads = 64;
    unsigned long long* __restrict__ acc;
    __device__ void operator()(Blk blk) const {
        if(threadIdx.x == 0) {
            acc[0] = ~0ull;
            acc[1] = 0ull;
        }
    }
};

// rank[perm[q]] = q, then every particle's external id becomes its rank
struct k_zsort_rank {
    static constexpr int kThreads = kBlock;
    const uint32_t* __restrict__ perm;
    uint32_t n;
    uint32_t* __restrict__ rank;
    __device__ void operator()(Blk blk) const {
        const uint32_t q = blk.id * blockDim.x + threadIdx.x;
        if(q < n) rank[perm[q]] = q;
    }
};
struct k_zsort_rename {
    static constexpr int kThreads = kBlock;
    uint32_t* __restrict__ id;
    uint32_t n;
    const uint32_t* __restrict__ rank;
    __device__ void operator()(Blk blk) const {
        const uint32_t p = blk.id * blockDim.x + threadIdx.x;
        if(p < n) id[p] = rank[id[p]];
    }
};

//--------------------------------------------------------------------------------------------------
// device-wide exclusive scan of uint32 (3 phases); tile = 256 threads x 8 items
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

struct k_scan_reduce {
    static constexpr int kThreads = kBlock;
    const uint32_t* __restrict__ in;
    size_t n;
    uint32_t* __restrict__ partial;
    __device__ void operator()(Blk blk) const {
        __shared__ uint32_t sh[kBlock / 64];
        const size_t        base = (size_t)blk.id * kScanTile + (size_t)threadIdx.x * kScanItems;
        uint32_t            s    = 0;
    #pragma unroll
        for(int i = 0; i < kScanItems; ++i)
            if(base + i < n) s += in[base + i];
        uint32_t total;
        block_exclusive_scan(s, sh, total);
        if(threadIdx.x == 0) partial[blk.id] = total;
    }
};

// single block: exclusive scan of the per-tile partials, in chunks of 256
struct k_scan_partials {
    static constexpr int kThreads = kBlock;
    uint32_t* partial;
    uint32_t np;
    uint32_t* total_out;
    __device__ void operator()(Blk blk) const {
        __shared__ uint32_t sh[kBlock / 64];
        uint32_t            carry = 0;
        for(uint32_t b = 0; b < np; b += kBlock) {
            const uint32_t i = b + threadIdx.x;
            const uint32_t v = i < np ? partial[i] : 0u;
            uint32_t       tot;
            const uint32_t ex = block_exclusive_scan(v, sh, tot);
            if(i < np) partial[i] = carry + ex;
            carry += tot;
        }
        if(threadIdx.x == 0 && total_out) *total_out = carry;
    }
};

struct k_scan_down {
    static constexpr int kThreads = kBlock;
    const uint32_t* __restrict__ in;
    size_t n;
    const uint32_t* __restrict__ partial;
    uint32_t* __restrict__ out;
    __device__ void operator()(Blk blk) const {
        __shared__ uint32_t sh[kBlock / 64];
        const size_t        base = (size_t)blk.id * kScanTile + (size_t)threadIdx.x * kScanItems;
        uint32_t            v[kScanItems];
        uint32_t            s = 0;
    #pragma unroll
        for(int i = 0; i < kScanItems; ++i) {
            v[i] = base + i < n ? in[base + i] : 0u;
            s += v[i];
        }
        uint32_t total;
        uint32_t run = block_exclusive_scan(s, sh, total) + partial[blk.id];
    #pragma unroll
        for(int i = 0; i < kScanItems; ++i) {
            if(base + i < n) out[base + i] = run;
            run += v[i];
        }
    }
};

// Tile (tx, ty, tz) of an ntx x nty x ntz grid of tiles for workgroup blk: blocks are dispatched
// round-robin over the 8 XCDs, so each XCD is handed a contiguous run of tiles and shares the reads
// of neighbouring tiles through its own L2.
__device__ __forceinline__ void tile_of_block(Blk blk, int ntx, int nty, int ntz, int& tx, int& ty, int& tz) {
    const uint32_t nt = (uint32_t)(ntx * nty * ntz), b = blk.id, xcd = b % 8u, q = nt / 8u, r = nt % 8u;
    const uint32_t t = xcd * q + min(xcd, r) + b / 8u;
    tx = (int)(t % (uint32_t)ntx);
    ty = (int)((t / (uint32_t)ntx) % (uint32_t)nty);
    tz = (int)(t / (uint32_t)(ntx * nty));
}

//==================================================================================================
// mapParticles2Grid as a gather over cell-sorted particles (FLIP_Solver.cpp:57-95,
// ArrayHelpers.cpp:50-76). Thread (i,j,k) owns u(i,j,k), v(i,j,k), w(i,j,k); it visits the particles
// of cells [i-1,i+1]x[j-1,j+1]x[k-1,k+1] in sorted order and keeps each particle's exact trilinear
// weight for the face when the face is one of the particle's 8 stencil corners.
//==================================================================================================
struct FaceOut {
    float *u, *v, *w;
    uint8_t *vu, *vv, *vw;
};

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

//--------------------------------------------------------------------------------------------------
// P2G as a gather that streams along z. A thread owns the nodes (i, j, k0 .. k0+kSeg-1) of all
// three face grids and reads each particle row z = k0-1 .. k0+kSeg once: a particle binned to
// row z reaches only nodes z-1 .. z+1 along z (its z base is z-1 or z for the half-shifted stencil,
// z for the node-aligned one, clamps included), which a three-node register window holds. Node k
// still receives rows k-1, k, k+1 in that order, each row's y rows in order and each y row's
// particles in cell-sorted order, with the weight (wx*wy)*wz and the momentum w*v rounded before
// each add (no contraction): the per-node sums are those of a per-node gather over its 27 cells.
//--------------------------------------------------------------------------------------------------
// The segment length kSeg (the z extent of a thread's tile: 4, 8 or 16) changes which thread sums a
// node, never the order of its sum, so every length gives the same bits; the batch autotuner picks it.

struct P2gAcc {
    float mu, wu, mv, wv, mw, ww;
};

// weights (1-f) at base and f at base+1; d = base - (z-1) picks the window slots (0,1) or (1,2).
// vlo / vhi: the velocity carried to the base node and to base+1 (the particle's velocity; with the
// affine term of APIC / AFLIP, v + dot(C_d, node - particle) per node). A zero ab (no
// contribution) adds exact zeros.
__device__ __forceinline__ void p2g_window_add(int d, float ab, float f, float vlo, float vhi, float& m0, float& w0, float& m1,
                                               float& w1, float& m2, float& w2) {
#pragma clang fp contract(off)
    const float wl = ab * (1.0f - f), wh = ab * f;
    const bool  lo = d == 0;
    const float c0 = lo ? wl : 0.f, c1 = lo ? wh : wl, c2 = lo ? 0.f : wh;
    const float v1 = lo ? vhi : vlo;
    m0 += c0 * vlo;
    w0 += c0;
    m1 += c1 * v1;
    w1 += c1;
    m2 += c2 * vhi;
    w2 += c2;
}

// v_d + dot(C_d, gpos - ppos) (APIC_Solver.cpp:73-76) for the node at grid index (xi + ex, yi + ey,
// zi + ez): gpos = getWorldCoordinate = (index + extra) * dx + bMin, the dot product in glm's order
__device__ __forceinline__ float affine_vel(float v, const float* c, float xi, float yi, float zi, float ex, float ey, float ez,
                                            const GridGeom& g, float px, float py, float pz) {
#pragma clang fp contract(off)
    const float gxw = (xi + ex) * g.dx + g.bminx, gyw = (yi + ey) * g.dx + g.bminy, gzw = (zi + ez) * g.dx + g.bminz;
    return v + ((c[0] * (gxw - px) + c[1] * (gyw - py)) + c[2] * (gzw - pz));
}

__device__ __forceinline__ void p2g_emit(const P2gAcc& a, int i, int j, int k, bool uCol, bool vCol, bool wCol, const GridGeom& g,
                                         const FaceOut& out) {
    if(uCol && k < g.nz) {
        const size_t id = lin3(i, j, k, g.nx + 1, g.ny);
        const bool   ok = a.wu > kFltMin;
        out.u[id]       = ok ? a.mu / a.wu : 0.f;
        out.vu[id]      = ok ? 1 : 0;
    }
    if(vCol && k < g.nz) {
        const size_t id = lin3(i, j, k, g.nx, g.ny + 1);
        const bool   ok = a.wv > kFltMin;
        out.v[id]       = ok ? a.mv / a.wv : 0.f;
        out.vv[id]      = ok ? 1 : 0;
    }
    if(wCol) {
        const size_t id = lin3(i, j, k, g.nx, g.ny);
        const bool   ok = a.ww > kFltMin;
        out.w[id]       = ok ? a.mw / a.ww : 0.f;
        out.vw[id]      = ok ? 1 : 0;
    }
}

// stencil weight of node `target` for coordinate q when no clamp can apply (q's floor is base, the
// weights (1 - f, f) sit at base and base + 1): the operations of bary + stencil_weight without the
// clamps, so the value is the same; 0 outside the stencil
__device__ __forceinline__ float stencil_weight_interior(float q, float target) {
    const float s = floorf(q), f = q - s;
    return s == target ? 1.0f - f : (s == target - 1.0f ? f : 0.f);
}

// A particle binned to row z, seen from column (i, j): abU / abV / abW the x-y weight products of the
// three face grids (0 when the column is not in the particle's stencil or the grid's row is off: onU
// etc.), fh / dh and fn / dn the z weight of the half-shifted (u, v) and node-aligned (w) stencils and
// the window slot of their base. kInterior: no barycentric clamp can apply.
template<bool kInterior>
__device__ __forceinline__ void p2g_col_weights(float gx, float gy, float gz, bool onU, bool onV, bool onW, const GridGeom& g, int i,
                                                int j, int z, float fi, float fj, float& abU, float& abV, float& abW, float& fh,
                                                float& fn, int& dh, int& dn) {
#pragma clang fp contract(off)
    if(kInterior) {
        const float axn = stencil_weight_interior(gx, fi), axh = stencil_weight_interior(gx - 0.5f, fi);
        const float byn = stencil_weight_interior(gy, fj), byh = stencil_weight_interior(gy - 0.5f, fj);
        const float qh = gz - 0.5f, sh = floorf(qh);
        fh  = qh - sh;
        dh  = sh == (float)(z - 1) ? 0 : 1;
        fn  = gz - floorf(gz);
        dn  = 1;  // the node-aligned base is the row itself
        abU = onU ? axn * byh : 0.f;
        abV = onV ? axh * byn : 0.f;
        abW = onW ? axh * byh : 0.f;
    } else {
        float      axn = 0.f, axh = 0.f, byn = 0.f, byh = 0.f;
        const bool oxn = stencil_weight(gx, g.nx + 1, i, axn);
        const bool oxh = stencil_weight(gx - 0.5f, g.nx, i, axh);
        const bool oyn = stencil_weight(gy, g.ny + 1, j, byn);
        const bool oyh = stencil_weight(gy - 0.5f, g.ny, j, byh);
        int        kh, kn;
        bary(gz - 0.5f, g.nz, kh, fh);
        bary(gz, g.nz + 1, kn, fn);
        dh  = kh - (z - 1);
        dn  = kn - (z - 1);
        abU = (onU && oxn && oyh) ? axn * byh : 0.f;
        abV = (onV && oxh && oyn) ? axh * byn : 0.f;
        abW = (onW && oxh && oyh) ? axh * byh : 0.f;
    }
}

// one column segment of P2G: nodes (i, j, k0 .. k1-1). kInterior: every cell the segment reads lies
// in [1, n-2] on each axis, so no barycentric clamp can apply to its particles.
// FLIP / PIC: the coordinates and velocities of kBatch particles are loaded together (one memory
// round trip per batch instead of two or more per particle) and then added in order; a batch slot
// past the row's end repeats its last particle with every weight 0, which adds exact zeros (a sum
// here starts at +0 and never becomes -0), so the sums are those of the particle-by-particle loop.
template<bool kInterior, bool kAffine>
__device__ __forceinline__ void p2g_column(const Particles& P, const Affine& A, const uint32_t* __restrict__ cell_start, const GridGeom& g,
                                           const FaceOut& out, int i, int j, int k0, int k1) {
#pragma clang fp contract(off)
    constexpr int kBatch = 4;
    const bool uCol = j < g.ny, vCol = i < g.nx, wCol = i < g.nx && j < g.ny;
    const int  x0 = max(i - 1, 0), x1 = min(i + 1, g.nx - 1), xu = min(i, g.nx - 1);
    const int  y0 = max(j - 1, 0), y1 = min(j + 1, g.ny - 1);
    const int  zs = max(k0 - 1, 0), ze = min(k1, g.nz - 1);
    const float fi = (float)i, fj = (float)j;
    P2gAcc     a0 = {}, a1 = {}, a2 = {};  // nodes z-1, z, z+1 while row z is read
    for(int z = zs; z <= ze; ++z) {
        for(int y = y0; y <= y1; ++y) {
            const size_t   r    = lin3(0, y, z, g.nx, g.ny);
            const uint32_t pb   = cell_start[r + x0], pe = cell_start[r + x1 + 1], pu = cell_start[r + xu + 1];
            const bool     rowU = uCol, rowV = vCol && y <= j, rowW = wCol;
            if constexpr(!kAffine) {
                for(uint32_t p0 = pb; p0 < pe; p0 += kBatch) {
                    float qx[kBatch], qy[kBatch], qz[kBatch], ux[kBatch], uy[kBatch], uz[kBatch];
#pragma unroll
                    for(int q = 0; q < kBatch; ++q) {
                        const uint32_t p = min(p0 + (uint32_t)q, pe - 1u);
                        const float4   a = P.gp[p];
                        const float2   b = P.gv[p];
                        qx[q]            = a.x;
                        qy[q]            = a.y;
                        qz[q]            = a.z;
                        ux[q]            = a.w;
                        uy[q]            = b.x;
                        uz[q]            = b.y;
                    }
#pragma unroll
                    for(int q = 0; q < kBatch; ++q) {
                        const uint32_t p    = p0 + (uint32_t)q;
                        const bool     live = p < pe;
                        float          abU, abV, abW, fh, fn;
                        int            dh, dn;
                        p2g_col_weights<kInterior>(qx[q], qy[q], qz[q], live && rowU && p < pu, live && rowV, live && rowW, g, i, j, z, fi,
                                                   fj, abU, abV, abW, fh, fn, dh, dn);
                        // a zero weight adds exact zeros with v taken as 0
                        const float vx = abU != 0.f ? ux[q] : 0.f, vy = abV != 0.f ? uy[q] : 0.f, vz = abW != 0.f ? uz[q] : 0.f;
                        p2g_window_add(dh, abU, fh, vx, vx, a0.mu, a0.wu, a1.mu, a1.wu, a2.mu, a2.wu);
                        p2g_window_add(dh, abV, fh, vy, vy, a0.mv, a0.wv, a1.mv, a1.wv, a2.mv, a2.wv);
                        p2g_window_add(dn, abW, fn, vz, vz, a0.mw, a0.ww, a1.mw, a1.ww, a2.mw, a2.ww);
                    }
                }
            } else {
                for(uint32_t p = pb; p < pe; ++p) {
                    float abU, abV, abW, fh, fn;
                    int   dh, dn;
                    const float4 gq = P.gp[p];
                    p2g_col_weights<kInterior>(gq.x, gq.y, gq.z, rowU && p < pu, rowV, rowW, g, i, j, z, fi, fj, abU, abV, abW, fh, fn, dh,
                                               dn);
                    // a zero weight adds exact zeros whatever v is, so v is read only when it counts
                    const float vx = abU != 0.f ? P.vx[p] : 0.f, vy = abV != 0.f ? P.vy[p] : 0.f, vz = abW != 0.f ? P.vz[p] : 0.f;
                    // the base node along z of the half-shifted (u, v) and node-aligned (w) stencils
                    const float zh = (float)(z - 1 + dh), zn = (float)(z - 1 + dn);
                    const float px = P.px[p], py = P.py[p], pz = P.pz[p];
                    float       cu[3] = {0.f, 0.f, 0.f}, cv[3] = {0.f, 0.f, 0.f}, cw[3] = {0.f, 0.f, 0.f};
                    if(abU != 0.f)
                        for(int e = 0; e < 3; ++e) cu[e] = A.c[e][p];
                    if(abV != 0.f)
                        for(int e = 0; e < 3; ++e) cv[e] = A.c[3 + e][p];
                    if(abW != 0.f)
                        for(int e = 0; e < 3; ++e) cw[e] = A.c[6 + e][p];
                    p2g_window_add(dh, abU, fh, affine_vel(vx, cu, fi, fj, zh, 0.f, 0.5f, 0.5f, g, px, py, pz),
                                   affine_vel(vx, cu, fi, fj, zh + 1.0f, 0.f, 0.5f, 0.5f, g, px, py, pz), a0.mu, a0.wu, a1.mu, a1.wu,
                                   a2.mu, a2.wu);
                    p2g_window_add(dh, abV, fh, affine_vel(vy, cv, fi, fj, zh, 0.5f, 0.f, 0.5f, g, px, py, pz),
                                   affine_vel(vy, cv, fi, fj, zh + 1.0f, 0.5f, 0.f, 0.5f, g, px, py, pz), a0.mv, a0.wv, a1.mv, a1.wv,
                                   a2.mv, a2.wv);
                    p2g_window_add(dn, abW, fn, affine_vel(vz, cw, fi, fj, zn, 0.5f, 0.5f, 0.f, g, px, py, pz),
                                   affine_vel(vz, cw, fi, fj, zn + 1.0f, 0.5f, 0.5f, 0.f, g, px, py, pz), a0.mw, a0.ww, a1.mw, a1.ww,
                                   a2.mw, a2.ww);
                }
            }
        }
        if(z - 1 >= k0) p2g_emit(a0, i, j, z - 1, uCol, vCol, wCol, g, out);  // node z-1 is complete
        a0 = a1;
        a1 = a2;
        a2 = P2gAcc{};
    }
    // nodes ze and ze+1 receive no further rows
    if(ze >= k0 && ze < k1) p2g_emit(a0, i, j, ze, uCol, vCol, wCol, g, out);
    if(ze + 1 >= k0 && ze + 1 < k1) p2g_emit(a1, i, j, ze + 1, uCol, vCol, wCol, g, out);
}

// the 8 stencil nodes of one particle on face grid c (extent sx x sy x sz): momentum w * v and weight
// w = (wx * wy) * wz added to the tile's sums, acc[c * 513 + node] = (momentum, weight), for the nodes
// the tile owns; the others go to the unused slot c * 513 + 512. The 8 nodes are distinct, so all 8
// sums are read before any is written back (one LDS round trip per component, no branches).
__device__ __forceinline__ void p2g_tile_add(float2* acc, int c, int ix, float fx, int iy, float fy, int iz, float fz, float v, int x0, int y0,
                                             int z0, int3 lim) {
#pragma clang fp contract(off)
    const float wx[2] = {1.0f - fx, fx}, wy[2] = {1.0f - fy, fy}, wz[2] = {1.0f - fz, fz};
    // byte offset of a node within the component's 512 owned nodes, per axis; an axis the tile does
    // not own (outside the tile or past the grid: lim = min(8, size - tile origin)) adds 4096, so the
    // sum of the three, clamped to 4096, is the node's offset or the unused slot's
    uint32_t ox[2], oy[2], oz[2];
#pragma unroll
    for(int d = 0; d < 2; ++d) {
        const uint32_t i = (uint32_t)(ix + d - x0), j = (uint32_t)(iy + d - y0), k = (uint32_t)(iz + d - z0);
        ox[d]            = i < (uint32_t)lim.x ? i * 8u : 4096u;
        oy[d]            = j < (uint32_t)lim.y ? j * 64u : 4096u;
        oz[d]            = k < (uint32_t)lim.z ? k * 512u : 4096u;
    }
    char* const base = reinterpret_cast<char*>(acc + c * 513);
    float2*     n[8];
    float       w[8];
#pragma unroll
    for(int e = 0; e < 8; ++e) {
        const int dx = e & 1, dy = (e >> 1) & 1, dz = e >> 2;
        n[e]         = reinterpret_cast<float2*>(base + min(ox[dx] + oy[dy] + oz[dz], 4096u));
        w[e]         = (wx[dx] * wy[dy]) * wz[dz];
    }
    float2 a[8];
#pragma unroll
    for(int e = 0; e < 8; ++e) a[e] = *n[e];
#pragma unroll
    for(int e = 0; e < 8; ++e) {
        a[e].x += w[e] * v;
        a[e].y += w[e];
    }
#pragma unroll
    for(int e = 0; e < 8; ++e) *n[e] = a[e];
}

//--------------------------------------------------------------------------------------------------
// P2G (FLIP / PIC) as a deterministic scatter: one wavefront per 8^3 tile of nodes, every particle's
// stencil weights computed once. A particle binned to cell c reaches nodes c-1 .. c+1 on each axis
// (clamps included), so the tile's nodes [x0, x0+8) x [y0, y0+8) x [z0, z0+8) of the three face grids
// receive only particles of the cells [x0-1, x0+8] x .. (10^3, clipped to the grid), which it
// visits in 27 rounds, one colour class (cx mod 3, cy mod 3, cz mod 3) per round. Two cells of one
// colour lie 3 or more cells apart on some axis and reach no common node, so in a round each lane
// owns a cell (at most 4 x 4 x 4 per colour) and adds its particles, in their sorted order, to the
// tile's momentum and weight sums in LDS with no other lane touching those nodes. A node's sum is
// therefore taken in a fixed order: by colour, then by particle; each term is the reference's
// weight (wx * wy) * wz and momentum w * v, rounded before the add (no contraction). Contributions to
// nodes outside the tile are left to the tile that owns them.
//--------------------------------------------------------------------------------------------------
struct k_p2g_tile {
    static constexpr int kThreads = 64;
    Particles P;
    const uint32_t* __restrict__ cell_start;
    GridGeom g;
    FaceOut out;
    const uint8_t* __restrict__ occ;  // particle occupancy of the 8^3 cell bricks (binning)
    __device__ void operator()(Blk blk) const {
#pragma clang fp contract(off)
        __shared__ float2 acc[3 * 513];  // [component][node (x fastest), + 1 unused]: (momentum, weight)
        int tx, ty, tz;
        tile_of_block(blk, (g.nx + 8) / 8, (g.ny + 8) / 8, (g.nz + 8) / 8, tx, ty, tz);
        const int x0 = tx * 8, y0 = ty * 8, z0 = tz * 8, lane = threadIdx.x;
        const int cx0 = max(x0 - 1, 0), cx1 = min(x0 + 8, g.nx - 1);  // cells read
        const int cy0 = max(y0 - 1, 0), cy1 = min(y0 + 8, g.ny - 1);
        const int cz0 = max(z0 - 1, 0), cz1 = min(z0 + 8, g.nz - 1);
        // nodes owned per axis and component (face grids (nx+1, ny, nz), (nx, ny+1, nz), (nx, ny, nz+1));
        // the tile origin never lies past a face grid's size, so each is in 0 .. 8
        const int  mx = min(8, g.nx - x0), my = min(8, g.ny - y0), mz = min(8, g.nz - z0);
        const int3 limU = make_int3(min(8, g.nx + 1 - x0), my, mz), limV = make_int3(mx, min(8, g.ny + 1 - y0), mz),
                   limW = make_int3(mx, my, min(8, g.nz + 1 - z0));
        for(int e = lane; e < 3 * 513; e += 64) acc[e] = make_float2(0.f, 0.f);
        bool any = false;
        for(int bz = cz0 >> 3; bz <= (cz1 >> 3); ++bz)
            for(int by = cy0 >> 3; by <= (cy1 >> 3); ++by)
                for(int bx = cx0 >> 3; bx <= (cx1 >> 3); ++bx) any |= occ[(bz * g.nby + by) * g.nbx + bx] != 0;
        __syncthreads();
        if(any) {
            // the particle range of the lane's cell in a colour round (empty when the lane has no cell)
            auto range = [&](int colour, uint32_t& pb, uint32_t& pe) {
                const int kx = colour % 3, ky = (colour / 3) % 3, kz = colour / 9;
                // the (lane & 3)-th, ((lane >> 2) & 3)-th, (lane >> 4)-th cell of this colour on each axis
                const int cx = cx0 + (kx - cx0 % 3 + 3) % 3 + 3 * (lane & 3);
                const int cy = cy0 + (ky - cy0 % 3 + 3) % 3 + 3 * ((lane >> 2) & 3);
                const int cz = cz0 + (kz - cz0 % 3 + 3) % 3 + 3 * (lane >> 4);
                pb = pe = 0;
                if(cx <= cx1 && cy <= cy1 && cz <= cz1) {
                    const size_t cell = lin3(cx, cy, cz, g.nx, g.ny);
                    pb                = cell_start[cell];
                    pe                = cell_start[cell + 1];
                }
            };
            // up to 8 particles [p0, pe) of a cell: coordinates and velocities (a slot past pe repeats the last)
            struct Batch {
                float qx[8], qy[8], qz[8], ux[8], uy[8], uz[8];
            };
            auto load = [&](uint32_t p0, uint32_t pe, Batch& B) {
#pragma unroll
                for(int q = 0; q < 8; ++q) {
                    const uint32_t p = min(p0 + (uint32_t)q, pe - 1u);
                    const float4   a = P.gp[p];
                    const float2   b = P.gv[p];
                    B.qx[q]          = a.x;
                    B.qy[q]          = a.y;
                    B.qz[q]          = a.z;
                    B.ux[q]          = a.w;
                    B.uy[q]          = b.x;
                    B.uz[q]          = b.y;
                }
            };
            auto add = [&](const Batch& B, uint32_t p0, uint32_t pe) {
#pragma unroll
                for(int q = 0; q < 8; ++q) {
                    if(p0 + (uint32_t)q >= pe) break;
                    // node-aligned (n) and half-shifted (h) stencils on each axis: base and fraction
                    int   ixn, ixh, iyn, iyh, izn, izh;
                    float fxn, fxh, fyn, fyh, fzn, fzh;
                    bary(B.qx[q], g.nx + 1, ixn, fxn);
                    bary(B.qx[q] - 0.5f, g.nx, ixh, fxh);
                    bary(B.qy[q], g.ny + 1, iyn, fyn);
                    bary(B.qy[q] - 0.5f, g.ny, iyh, fyh);
                    bary(B.qz[q], g.nz + 1, izn, fzn);
                    bary(B.qz[q] - 0.5f, g.nz, izh, fzh);
                    // u: x node-aligned, y and z half-shifted; v: y node-aligned; w: z node-aligned
                    p2g_tile_add(acc, 0, ixn, fxn, iyh, fyh, izh, fzh, B.ux[q], x0, y0, z0, limU);
                    p2g_tile_add(acc, 1, ixh, fxh, iyn, fyn, izh, fzh, B.uy[q], x0, y0, z0, limV);
                    p2g_tile_add(acc, 2, ixh, fxh, iyh, fyh, izn, fzn, B.uz[q], x0, y0, z0, limW);
                }
            };
            // software pipeline: a round's first batch is loaded during the round before, its cell
            // range two rounds before
            Batch    cur, nxt;
            uint32_t cb, ce, nb, ne, nnb = 0, nne = 0;
            range(0, cb, ce);
            range(1, nb, ne);
            if(cb < ce) load(cb, ce, cur);
            for(int colour = 0; colour < 27; ++colour) {
                if(colour + 1 < 27 && nb < ne) load(nb, ne, nxt);
                if(colour + 2 < 27) range(colour + 2, nnb, nne);
                for(uint32_t p0 = cb; p0 < ce; p0 += 8) {
                    if(p0 != cb) load(p0, ce, cur);  // cells of more than 8 particles
                    add(cur, p0, ce);
                }
                __syncthreads();
                cur = nxt;
                cb  = nb;
                ce  = ne;
                nb  = nnb;
                ne  = nne;
            }
        }
        // the tile's nodes: u = momentum / weight where the weight is positive, else 0 and invalid
        for(int e = lane; e < 512; e += 64) {
            const int i = x0 + (e & 7), j = y0 + ((e >> 3) & 7), k = z0 + (e >> 6);
            if(i <= g.nx && j < g.ny && k < g.nz) {
                const size_t id = lin3(i, j, k, g.nx + 1, g.ny);
                const float  m = acc[e].x, w = acc[e].y;
                const bool   ok = w > kFltMin;
                out.u[id]       = ok ? m / w : 0.f;
                out.vu[id]      = ok ? 1 : 0;
            }
            if(i < g.nx && j <= g.ny && k < g.nz) {
                const size_t id = lin3(i, j, k, g.nx, g.ny + 1);
                const float  m = acc[513 + e].x, w = acc[513 + e].y;
                const bool   ok = w > kFltMin;
                out.v[id]       = ok ? m / w : 0.f;
                out.vv[id]      = ok ? 1 : 0;
            }
            if(i < g.nx && j < g.ny && k <= g.nz) {
                const size_t id = lin3(i, j, k, g.nx, g.ny);
                const float  m = acc[1026 + e].x, w = acc[1026 + e].y;
                const bool   ok = w > kFltMin;
                out.w[id]       = ok ? m / w : 0.f;
                out.vw[id]      = ok ? 1 : 0;
            }
        }
    }
};

template<bool kAffine, int kSeg>
struct k_p2g_col {
    static constexpr int kThreads = kBlock;
    Particles P;
    Affine A;
    const uint32_t* __restrict__ cell_start;
    GridGeom g;
    FaceOut out;
    const uint8_t* __restrict__ occ;  // particle occupancy of the 8^3 cell bricks (binning)
    __device__ void operator()(Blk blk) const {
        const int      ex = g.nx + 1, ey = g.ny + 1, ez = g.nz + 1;
        const uint32_t cols = (uint32_t)ex * (uint32_t)ey;
        const uint32_t nseg = (uint32_t)(ez + kSeg - 1) / kSeg;
        // consecutive blocks (neighbouring j rows, which read the same particle rows) stay on one XCD
        const uint32_t nblk = blk.n, b = blk.id, xcd = b % 8u, q = nblk / 8u, rr = nblk % 8u;
        const uint32_t t    = (xcd * q + min(xcd, rr) + b / 8u) * blockDim.x + threadIdx.x;
        if(t >= cols * nseg) return;
        const int  i = (int)(t % (uint32_t)ex), j = (int)((t / (uint32_t)ex) % (uint32_t)ey), seg = (int)(t / cols);
        const int  k0 = seg * kSeg, k1 = min(k0 + kSeg, ez);  // nodes [k0, k1)
        // cells read: x in [i-1, i+1], y in [j-1, j+1], z in [k0-1, k1] (clamped to the grid). With no
        // particle in any brick covering them every node of the segment has weight 0: u = 0, invalid.
        {
            const int x0 = max(i - 1, 0), x1 = min(i + 1, g.nx - 1), y0 = max(j - 1, 0), y1 = min(j + 1, g.ny - 1);
            const int zs = max(k0 - 1, 0), ze = min(k1, g.nz - 1);
            bool      any = false;
            for(int bz = zs >> 3; bz <= (ze >> 3); ++bz)
                for(int by = y0 >> 3; by <= (y1 >> 3); ++by)
                    for(int bx = x0 >> 3; bx <= (x1 >> 3); ++bx) any |= occ[(bz * g.nby + by) * g.nbx + bx] != 0;
            if(!any) {
                const P2gAcc zero{};
                for(int k = k0; k < k1; ++k) p2g_emit(zero, i, j, k, j < g.ny, i < g.nx, i < g.nx && j < g.ny, g, out);
                return;
            }
        }
        const bool interior = i >= 2 && i <= g.nx - 3 && j >= 2 && j <= g.ny - 3 && k0 >= 2 && k1 <= g.nz - 2;
        if(interior) p2g_column<true, kAffine>(P, A, cell_start, g, out, i, j, k0, k1);
        else p2g_column<false, kAffine>(P, A, cell_start, g, out, i, j, k0, k1);
    }
};

//==================================================================================================
// extrapolateVelocity of one component (PIC_Solver.cpp:985-1090): 4 Jacobi layers. Layer l fills an
// invalid face that is not on the array's outer layer with the mean of its neighbours valid before
// the layer (order +x, -x, +y, -y, +z, -z) and makes it valid and extrapolated.
//
// One workgroup per 8^3 brick of faces. A face changes only through faces within 4 face steps, so
// the brick and a halo of 4 faces on every side (a 16^3 tile) are read into LDS once and the four
// layers run there, in place with layer stamps: 0 invalid, 1 valid on entry, l + 2 filled in layer l;
// layer l reads the stamps 1 .. l + 1, so a face filled in the same layer is never read by it and the
// values equal the two-buffer Jacobi's bit for bit. A halo face at distance d from the brick can go
// wrong (its own neighbours beyond the tile are missing) only from layer 4 - d on, too late to reach
// the brick. Faces on the tile's outer shell are never computed.
//
// Other workgroups read this brick's faces as their halo while it runs, so what they read as valid
// on entry is never changed: a face filled here gets its value, ext = 1 and val = gen (2 or 3, the
// host alternates it per call), which this launch's loaders read as invalid; the other generation,
// left by the previous call, reads as valid and is rewritten as 1. Every other face of the brick gets
// ext = 0. Any nonzero val therefore means valid.
//==================================================================================================
struct k_extrap_tile {
    static constexpr int kThreads = 256;  // one (x, y) column of the tile per thread, its 16 faces in z
    static constexpr int kManyBlocks = 4;  // batched: at most 128 registers (a tighter bound spills)
    float* __restrict__ a;
    uint8_t* __restrict__ val;
    uint8_t* __restrict__ ext;
    int sx;
    int sy;
    int sz;
    uint8_t gen;
    __device__ void operator()(Blk blk) const {
        __shared__ float   as[4096];  // 16^3, x fastest
        __shared__ uint8_t ss[4096];
        int bx, by, bz;
        tile_of_block(blk, (sx + 7) / 8, (sy + 7) / 8, (sz + 7) / 8, bx, by, bz);
        const int    t = threadIdx.x, x = t & 15, y = t >> 4;  // faces (x, y, 0 .. 15)
        const int    gi = bx * 8 - 4 + x, gj = by * 8 - 4 + y, gk0 = bz * 8 - 4;
        const bool   inXY = gi >= 0 && gi < sx && gj >= 0 && gj < sy;
        const bool   brickXY = x >= 4 && x < 12 && y >= 4 && y < 12;
        const size_t sxy = (size_t)sx * (size_t)sy, col = (size_t)gj * (size_t)sx + (size_t)gi;
        auto         off = [&](int z) { return (size_t)(gk0 + z) * sxy + col; };
        uint8_t      v0[16];  // val on entry
        // the thread's own column: mark (0 not valid, 1 valid on entry, layer + 2 filled by that layer)
        // and value; its z neighbours are read from here, its x and y neighbours from the LDS copy
        uint8_t      cs[16];
        float        ca[16];
        int          anyValid = 0, anyOpen = 0;
#pragma unroll
        for(int z = 0; z < 16; ++z) {
            const int gk = gk0 + z;
            v0[z]        = (inXY && gk >= 0 && gk < sz) ? val[off(z)] : (uint8_t)0;
        }
#pragma unroll
        for(int z = 0; z < 16; ++z) {
            const uint8_t st = (v0[z] != 0 && v0[z] != gen) ? 1 : 0;
            ss[(z << 8) | (y << 4) | x] = st;
            cs[z]                       = st;
            ca[z]                       = 0.f;
            anyValid |= st;
            if(!st && brickXY && z >= 4 && z < 12 && gk0 + z < sz && inXY) anyOpen = 1;
        }
        const bool work = __syncthreads_or(anyValid) != 0 && __syncthreads_or(anyOpen) != 0;
        if(work) {
            // the values of the faces valid on entry (the others are written before they are read)
#pragma unroll
            for(int z = 0; z < 16; ++z)
                if(cs[z]) {
                    ca[z]                       = a[off(z)];
                    as[(z << 8) | (y << 4) | x] = ca[z];
                }
            __syncthreads();
            // faces that can be filled: not on the tile's outer shell, not on the array's outer layer
            uint32_t open = 0;
#pragma unroll
            for(int z = 1; z <= 14; ++z) {
                const int gk = gk0 + z;
                if(x >= 1 && x <= 14 && y >= 1 && y <= 14 && gi >= 1 && gi <= sx - 2 && gj >= 1 && gj <= sy - 2 && gk >= 1 && gk <= sz - 2 && !cs[z])
                    open |= 1u << z;
            }
            // Layer l (0 .. 3) updates only the faces within 3 - l cells (per axis) of the brick: a face
            // farther out reaches the brick through fewer than the layers left, so leaving it unfilled
            // changes no brick face, and every face a layer reads lies within the previous layer's region.
#pragma unroll
            for(int layer = 0; layer < 4; ++layer) {
                const uint8_t  lim = (uint8_t)(layer + 1);
                const int      lo = layer + 1, hi = 14 - layer;
                const uint32_t act = (x >= lo && x <= hi && y >= lo && y <= hi) ? open & (((2u << hi) - 1u) & ~((1u << lo) - 1u)) : 0u;
                float          res[16];
                uint32_t       fill = 0;
                // all reads of the layer first (a face this layer fills is never read by it), then the writes
#pragma unroll
                for(int z = 1; z <= 14; ++z) {
                    res[z] = 0.f;
                    if(!(act & (1u << z))) continue;
                    const int q   = (z << 8) | (y << 4) | x;
                    float     sum = 0.f;  // starts at +0 and never becomes -0, so adding +0 for a neighbour not taken is exact
                    uint32_t  count = 0;
                    // neighbours in the order +x, -x, +y, -y, +z, -z
                    const uint8_t m[6]  = {ss[q + 1], ss[q - 1], ss[q + 16], ss[q - 16], cs[z + 1], cs[z - 1]};
                    const float   av[6] = {as[q + 1], as[q - 1], as[q + 16], as[q - 16], ca[z + 1], ca[z - 1]};
#pragma unroll
                    for(int e = 0; e < 6; ++e) {
                        const bool ok = m[e] >= 1 && m[e] <= lim;
                        sum += ok ? av[e] : 0.f;
                        count += ok ? 1u : 0u;
                    }
                    if(count > 0) {
                        res[z] = sum / (float)count;
                        fill |= 1u << z;
                    }
                }
#pragma unroll
                for(int z = 1; z <= 14; ++z) {
                    if(!(fill & (1u << z))) continue;
                    const int q = (z << 8) | (y << 4) | x;
                    as[q]       = res[z];
                    ss[q]       = (uint8_t)(layer + 2);
                    ca[z]       = res[z];
                    cs[z]       = (uint8_t)(layer + 2);
                }
                open &= ~fill;
                if(layer < 3) __syncthreads();
            }
        }
        if(!brickXY || !inXY) return;
#pragma unroll
        for(int z = 4; z < 12; ++z) {
            if(gk0 + z >= sz) continue;
            const size_t  o  = off(z);
            const uint8_t st = cs[z];
            ext[o]           = st >= 2 ? 1 : 0;
            if(st >= 2) {
                a[o]   = ca[z];
                val[o] = gen;
            } else if(st == 1 && v0[z] != 1) {
                val[o] = 1;
            }
        }
    }
};

//==================================================================================================
// constrainGridVelocity, component d (PIC_Solver.cpp:1094-1115, ArrayHelpers.cpp:214-251)
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

struct k_constrain {
    static constexpr int kThreads = kBlock;
    VelGrid G;
    GridGeom g;
    int d;
    const uint8_t* __restrict__ ext;
    const float* __restrict__ weight;
    const float* __restrict__ bsdf;
    float* __restrict__ out;
    // the new velocity of face t (extrapolated, zero weight) into out[t]
    __device__ void face(size_t t, int sx, int sy) const {
        const float* cur = d == 0 ? G.u : (d == 1 ? G.v : G.w);
        float        val = cur[t];
        int i, j, k;
        unlin3((uint32_t)t, sx, sy, i, j, k);
        float gp[3] = {(float)i + 0.5f, (float)j + 0.5f, (float)k + 0.5f};
        gp[d]       = d == 0 ? (float)i : (d == 1 ? (float)j : (float)k);
        float vx, vy, vz, nx, ny, nz;
        vel_from_grid(G, g, gp[0], gp[1], gp[2], vx, vy, vz);
        interp_gradient(bsdf, g.nx + 1, g.ny + 1, g.nz + 1, gp[0], gp[1], gp[2], nx, ny, nz);
        const float m2 = nx * nx + ny * ny + nz * nz;
        if(m2 > kFltMin) {
            const float s = sqrtf(m2);
            nx /= s;
            ny /= s;
            nz /= s;
            const float dt = vx * nx + vy * ny + vz * nz;
            vx -= dt * nx;
            vy -= dt * ny;
            vz -= dt * nz;
            val = d == 0 ? vx : (d == 1 ? vy : vz);
        }
        out[t] = val;
    }
    // four faces per thread: the flags and weights read as one 4-byte and one 16-byte load
    __device__ void operator()(Blk blk) const {
        const int    sx = g.nx + (d == 0), sy = g.ny + (d == 1), sz = g.nz + (d == 2);
        const size_t N  = (size_t)sx * sy * sz;
        const size_t t0 = 4 * ((size_t)blk.id * blockDim.x + threadIdx.x);
        if(t0 >= N) return;
        if(t0 + 4 <= N) {
            const uchar4 e = *reinterpret_cast<const uchar4*>(ext + t0);
            const float4 w = *reinterpret_cast<const float4*>(weight + t0);
            if(e.x == 1 && w.x < kFltMin) face(t0, sx, sy);
            if(e.y == 1 && w.y < kFltMin) face(t0 + 1, sx, sy);
            if(e.z == 1 && w.z < kFltMin) face(t0 + 2, sx, sy);
            if(e.w == 1 && w.w < kFltMin) face(t0 + 3, sx, sy);
        } else {
            for(size_t t = t0; t < N; ++t)
                if(ext[t] == 1 && weight[t] < kFltMin) face(t, sx, sy);
        }
    }
};

// the second half of a component's constrain: the faces k_constrain computed (extrapolated, zero
// weight) take their new values; every other face keeps its velocity
struct k_constrain_apply {
    static constexpr int kThreads = kBlock;
    float* __restrict__ vel;
    const float* __restrict__ tmp;
    const uint8_t* __restrict__ ext;
    const float* __restrict__ weight;
    size_t n;
    // four faces per thread, as in k_constrain
    __device__ void operator()(Blk blk) const {
        const size_t t0 = 4 * ((size_t)blk.id * blockDim.x + threadIdx.x);
        if(t0 >= n) return;
        if(t0 + 4 <= n) {
            const uchar4 e = *reinterpret_cast<const uchar4*>(ext + t0);
            const float4 w = *reinterpret_cast<const float4*>(weight + t0);
            if(e.x == 1 && w.x < kFltMin) vel[t0] = tmp[t0];
            if(e.y == 1 && w.y < kFltMin) vel[t0 + 1] = tmp[t0 + 1];
            if(e.z == 1 && w.z < kFltMin) vel[t0 + 2] = tmp[t0 + 2];
            if(e.w == 1 && w.w < kFltMin) vel[t0 + 3] = tmp[t0 + 3];
        } else {
            for(size_t t = t0; t < n; ++t)
                if(ext[t] == 1 && weight[t] < kFltMin) vel[t] = tmp[t];
        }
    }
};

//==================================================================================================
// addGravity (PIC_Solver.cpp:505-529): Earth / Directional add a constant per component
//==================================================================================================
// c = a * b elementwise
struct k_mul_arrays {
    static constexpr int kThreads = kBlock;
    const float* __restrict__ a;
    const float* __restrict__ b;
    float* __restrict__ c;
    size_t n;
    __device__ void operator()(Blk blk) const {
        const size_t i = (size_t)blk.id * blockDim.x + threadIdx.x;
        if(i < n) c[i] = a[i] * b[i];
    }
};

struct k_add_const {
    static constexpr int kThreads = kBlock;
    float* __restrict__ a;
    size_t n;
    float add;
    int subtract;
    __device__ void operator()(Blk blk) const {
        const size_t t = (size_t)blk.id * blockDim.x + threadIdx.x;
        if(t >= n) return;
        if(subtract) a[t] -= add;
        else a[t] += add;
    }
};

// ToCenter / FromCenter: the reference walks velocities[0]'s index range for every component
// (PIC_Solver.cpp:521); indices outside component d's array are skipped here.
struct k_gravity_center {
    static constexpr int kThreads = kBlock;
    float* __restrict__ a;
    GridGeom g;
    int d;
    float cx;
    float cy;
    float cz;
    int toCenter;
    float dt;
    __device__ void operator()(Blk blk) const {
        const int    sx = g.nx + 1, sy = g.ny, sz = g.nz;
        const size_t t  = (size_t)blk.id * blockDim.x + threadIdx.x;
        if(t >= (size_t)sx * sy * sz) return;
        int i, j, k;
        unlin3((uint32_t)t, sx, sy, i, j, k);
        const int ax = g.nx + (d == 0), ay = g.ny + (d == 1), az = g.nz + (d == 2);
        if(i >= ax || j >= ay || k >= az) return;
        float gp[3] = {(float)i + 0.5f, (float)j + 0.5f, (float)k + 0.5f};
        gp[d]       = d == 0 ? (float)i : (d == 1 ? (float)j : (float)k);
        const float wx = gp[0] * g.dx + g.bminx, wy = gp[1] * g.dx + g.bminy, wz = gp[2] * g.dx + g.bminz;
        float       dv[3];
        dv[0]          = toCenter ? cx - wx : wx - cx;
        dv[1]          = toCenter ? cy - wy : wy - cy;
        dv[2]          = toCenter ? cz - wz : wz - cz;
        const float is = 1.0f / sqrtf(dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2]);
        a[lin3(i, j, k, ax, ay)] += (9.81f * (dv[d] * is)) * dt;
    }
};

//==================================================================================================
// computeCellWeights (PIC_Solver.cpp:543-597); counts faces with weight < FLT_MIN (solid faces)
//==================================================================================================
struct k_weights {
    static constexpr int kThreads = kBlock;
    const float* __restrict__ b;
    GridGeom g;
    float* __restrict__ wu;
    float* __restrict__ wv;
    float* __restrict__ ww;
    uint32_t* solidFaces;
    __device__ void operator()(Blk blk) const {
        const int    sx = g.nx + 1, sy = g.ny + 1, sz = g.nz + 1;
        const size_t t  = (size_t)blk.id * blockDim.x + threadIdx.x;
        if(t >= (size_t)sx * sy * sz) return;
        int i, j, k;
        unlin3((uint32_t)t, sx, sy, i, j, k);
        auto      B = [&](int a, int c, int e) { return b[lin3(a, c, e, sx, sy)]; };
        uint32_t  solid = 0;
        if(j < g.ny && k < g.nz) {
            float w = clamp01f(1.0f - fraction_inside4(B(i, j, k), B(i, j + 1, k), B(i, j, k + 1), B(i, j + 1, k + 1)));
            wu[lin3(i, j, k, g.nx + 1, g.ny)] = w;
            solid += w < kFltMin;
        }
        if(i < g.nx && k < g.nz) {
            float w = clamp01f(1.0f - fraction_inside4(B(i, j, k), B(i, j, k + 1), B(i + 1, j, k), B(i + 1, j, k + 1)));
            wv[lin3(i, j, k, g.nx, g.ny + 1)] = w;
            solid += w < kFltMin;
        }
        if(i < g.nx && j < g.ny) {
            float w = clamp01f(1.0f - fraction_inside4(B(i, j, k), B(i, j + 1, k), B(i + 1, j, k), B(i + 1, j + 1, k)));
            ww[lin3(i, j, k, g.nx, g.ny)] = w;
            solid += w < kFltMin;
        }
        if(solid) atomicAdd(solidFaces, solid);
    }
};

// computeBoundarySDF: the smallest signed distance over the boundary objects (PIC_Solver.cpp:1158-1168)
struct k_boundary_sdf {
    static constexpr int kThreads = kBlock;
    float* __restrict__ b;
    GridGeom g;
    BndSet bnd;
    __device__ void operator()(Blk blk) const {
        const int    sx = g.nx + 1, sy = g.ny + 1, sz = g.nz + 1;
        const size_t t  = (size_t)blk.id * blockDim.x + threadIdx.x;
        if(t >= (size_t)sx * sy * sz) return;
        int i, j, k;
        unlin3((uint32_t)t, sx, sy, i, j, k);
        const float x = (float)i * g.dx + g.bminx, y = (float)j * g.dx + g.bminy, z = (float)k * g.dx + g.bminz;
        float       m = 3.40282347e+38f;  // Huge<float>
        for(int q = 0; q < bnd.nGroups; ++q) m = smin(m, bnd_group_sd(bnd, q, x, y, z));
        b[t] = m;
    }
};

//==================================================================================================
// computeFluidSDF as a gather (PIC_Solver.cpp:601-689): phi(c) = min over particles whose clamped
// 3x3x3 window holds c of |centre(c) - p| - sdfRadius; min is exact, so the result is bit-identical
// to the reference's locked scatter. Then the solid extension. A work-group owns an 8x8x4 tile of
// cells; tiles are handed out so that each XCD (blocks are dispatched round-robin over 8 XCDs) works
// on a contiguous run of tiles and shares neighbour-cell reads through its own L2. A cell visits its
// own cell first and stops as soon as no farther class of neighbour can lower the minimum, so most
// cells read only their own particles (staging the 10x10x6 halo in LDS cost more than it saved).
//==================================================================================================
constexpr int kTileX = 8, kTileY = 8, kTileZ = 4;
constexpr int kTileThreads = kTileX * kTileY * kTileZ;  // 256

struct k_fluid_sdf {
    static constexpr int kThreads = kTileThreads;
    Particles P;
    const uint32_t* __restrict__ cell_start;
    GridGeom g;
    const float* __restrict__ bsdf;
    float* __restrict__ phi;
    const uint8_t* __restrict__ occ;  // particle occupancy of the 8^3 cell bricks (the same binning)
    __device__ void operator()(Blk blk) const {
        const int ntx = (g.nx + kTileX - 1) / kTileX, nty = (g.ny + kTileY - 1) / kTileY, ntz = (g.nz + kTileZ - 1) / kTileZ;
        int       tx, ty, tz;
        tile_of_block(blk, ntx, nty, ntz, tx, ty, tz);
        const int  x0 = tx * kTileX, y0 = ty * kTileY, z0 = tz * kTileZ;
        const int  i = x0 + (int)threadIdx.x % kTileX, j = y0 + ((int)threadIdx.x / kTileX) % kTileY, k = z0 + (int)threadIdx.x / (kTileX * kTileY);
        if(i >= g.nx || j >= g.ny || k >= g.nz) return;
        const float sx      = ((float)i + 0.5f) * g.dx + g.bminx;
        const float sy      = ((float)j + 0.5f) * g.dx + g.bminy;
        const float sz      = ((float)k + 0.5f) * g.dx + g.bminz;
        float       best    = g.dx * 3.0f;
        float       bestLen = 3.40282347e+38f;  // smallest |centre - p| so far
        // The own cell first, then its 6 face, 12 edge and 8 corner neighbours: no particle of a class-c
        // neighbour is nearer to the centre than 0, 0.5, 0.7071 or 0.8660 dx (less 1e-3 dx for the
        // rounding of particle-to-cell assignment), so once the nearest particle found beats a class's
        // bound, that class and all farther ones cannot lower the minimum. min() is exact, so visiting
        // fewer cells gives the same bits.
        const float bound[4] = {0.f, 0.499f * g.dx, 0.7061068f * g.dx, 0.8650254f * g.dx};
        // no particle in the bricks that hold the 3x3x3 cells around this one: nothing to visit, the
        // distance stays at its initial value (the scan below would find every cell empty)
        bool near = false;
        for(int bz = max(k - 1, 0) >> 3; bz <= (min(k + 1, g.nz - 1) >> 3); ++bz)
            for(int by = max(j - 1, 0) >> 3; by <= (min(j + 1, g.ny - 1) >> 3); ++by)
                for(int bx = max(i - 1, 0) >> 3; bx <= (min(i + 1, g.nx - 1) >> 3); ++bx) near |= occ[(bz * g.nby + by) * g.nbx + bx] != 0;
        for(int cls = 0; cls < 4 && near; ++cls) {
            if(bestLen < bound[cls]) break;
            for(int dz = -1; dz <= 1; ++dz)
                for(int dy = -1; dy <= 1; ++dy)
                    for(int dxo = -1; dxo <= 1; ++dxo) {
                        if(abs(dz) + abs(dy) + abs(dxo) != cls) continue;
                        const int x = i + dxo, y = j + dy, z = k + dz;
                        if(x < 0 || y < 0 || z < 0 || x >= g.nx || y >= g.ny || z >= g.nz) continue;
                        const bool boundary = x == 0 || y == 0 || z == 0 || x == g.nx - 1 || y == g.ny - 1 || z == g.nz - 1;
                        auto       visit    = [&](float ppx, float ppy, float ppz) {
                            if(boundary) {
                                // particles clamped into a boundary cell: use their true (unclamped) window
                                const int ci = (int)((ppx - g.bminx) / g.dx), cj = (int)((ppy - g.bminy) / g.dx),
                                          ck = (int)((ppz - g.bminz) / g.dx);
                                if(i < ci - 1 || i > ci + 1 || j < cj - 1 || j > cj + 1 || k < ck - 1 || k > ck + 1) return;
                            }
                            const float len = len3f(sx - ppx, sy - ppy, sz - ppz);
                            const float v   = len - g.sdfRadius;
                            if(v < best) best = v;
                            if(len < bestLen) bestLen = len;
                        };
                        const size_t   c  = lin3(x, y, z, g.nx, g.ny);
                        const uint32_t pb = cell_start[c], pe = cell_start[c + 1];
                        for(uint32_t p = pb; p < pe; ++p) visit(P.px[p], P.py[p], P.pz[p]);
                    }
        }
        if(best < g.halfdx) {
            const int   sxn = g.nx + 1, syn = g.ny + 1;
            auto        B   = [&](int a, int c, int e) { return bsdf[lin3(a, c, e, sxn, syn)]; };
            const float s   = 0.125f * (B(i, j, k) + B(i + 1, j, k) + B(i, j + 1, k) + B(i + 1, j + 1, k) + B(i, j, k + 1) +
                                      B(i + 1, j, k + 1) + B(i, j + 1, k + 1) + B(i + 1, j + 1, k + 1));
            if(s < 0) best = -g.halfdx;
        }
        phi[lin3(i, j, k, g.nx, g.ny)] = best;
    }
};

//==================================================================================================
// Pressure bricks
//==================================================================================================
// one block (512 threads) per brick: flag bricks holding a fluid cell, count fluid cells
struct k_brick_flags {
    static constexpr int kThreads = 512;
    const float* __restrict__ phi;
    GridGeom g;
    uint32_t* __restrict__ flag;
    uint32_t* __restrict__ nFluid;
    __device__ void operator()(Blk blk) const {
        const int b  = blk.id;
        const int bx = b % g.nbx, by = (b / g.nbx) % g.nby, bz = b / (g.nbx * g.nby);
        const int l  = threadIdx.x;
        const int i = bx * 8 + (l & 7), j = by * 8 + ((l >> 3) & 7), k = bz * 8 + (l >> 6);
        int fluid = 0;
        if(i < g.nx && j < g.ny && k < g.nz) fluid = phi[lin3(i, j, k, g.nx, g.ny)] < 0.f;
        const int cnt = __syncthreads_count(fluid);
        if(l == 0) {
            flag[b] = cnt > 0 ? 1u : 0u;
            if(cnt) atomicAdd(nFluid, (uint32_t)cnt);
        }
    }
};

struct k_brick_slots {
    static constexpr int kThreads = kBlock;
    const uint32_t* __restrict__ flag;
    const uint32_t* __restrict__ scan;
    uint32_t nb;
    int32_t* __restrict__ brickSlot;
    int32_t* __restrict__ slotBrick;
    __device__ void operator()(Blk blk) const {
        const uint32_t b = blk.id * blockDim.x + threadIdx.x;
        if(b >= nb) return;
        if(flag[b]) {
            brickSlot[b]        = (int32_t)scan[b];
            slotBrick[scan[b]] = (int32_t)b;
        } else {
            brickSlot[b] = -1;
        }
    }
};

// neighbour slots per active brick: -x, +x, -y, +y, -z, +z
struct k_brick_nbrs {
    static constexpr int kThreads = kBlock;
    const int32_t* __restrict__ slotBrick;
    uint32_t nSlots;
    const int32_t* __restrict__ brickSlot;
    GridGeom g;
    int32_t* __restrict__ nbr;
    __device__ void operator()(Blk blk) const {
        const uint32_t s = blk.id * blockDim.x + threadIdx.x;
        if(s >= nSlots) return;
        const int b  = slotBrick[s];
        const int bx = b % g.nbx, by = (b / g.nbx) % g.nby, bz = b / (g.nbx * g.nby);
        auto      at = [&](int x, int y, int z) -> int32_t {
            if(x < 0 || y < 0 || z < 0 || x >= g.nbx || y >= g.nby || z >= g.nbz) return -1;
            return brickSlot[(z * g.nby + y) * g.nbx + x];
        };
        nbr[6 * s + 0] = at(bx - 1, by, bz);
        nbr[6 * s + 1] = at(bx + 1, by, bz);
        nbr[6 * s + 2] = at(bx, by - 1, bz);
        nbr[6 * s + 3] = at(bx, by + 1, bz);
        nbr[6 * s + 4] = at(bx, by, bz - 1);
        nbr[6 * s + 5] = at(bx, by, bz + 1);
    }
};

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
struct SystemOut {
    double *rhs64, *diag64, *cx64, *cy64, *cz64;  // cx64..: may be null (FP32 mode)
    float * rhs32, *diag32, *cx32, *cy32, *cz32, *invdiag32;
};

struct k_assemble {
    static constexpr int kThreads = 512;
    const int32_t* __restrict__ slotBrick;
    GridGeom g;
    const float* __restrict__ phi;
    VelGrid G;
    const float* __restrict__ wu;
    const float* __restrict__ wv;
    const float* __restrict__ ww;
    float dt;
    SystemOut S;
    __device__ void operator()(Blk blk) const {
        const int s = blk.id, l = threadIdx.x;
        const int b = slotBrick[s];
        const int bx = b % g.nbx, by = (b / g.nbx) % g.nby, bz = b / (g.nbx * g.nby);
        const int i = bx * 8 + (l & 7), j = by * 8 + ((l >> 3) & 7), k = bz * 8 + (l >> 6);
        const size_t o = (size_t)s * 512 + l;
        double rhs = 0, diag = 0, cx = 0, cy = 0, cz = 0;
        const bool interior = i >= 1 && j >= 1 && k >= 1 && i <= g.nx - 2 && j <= g.ny - 2 && k <= g.nz - 2;
        if(interior) {
            auto P = [&](int a, int c, int e) { return phi[lin3(a, c, e, g.nx, g.ny)]; };
            const float pc = P(i, j, k);
            if(pc < 0) {
                const int    ux = g.nx + 1, vy = g.ny + 1;
                const double wr = wu[lin3(i + 1, j, k, ux, g.ny)], wl = wu[lin3(i, j, k, ux, g.ny)];
                const double wt = wv[lin3(i, j + 1, k, g.nx, vy)], wb = wv[lin3(i, j, k, g.nx, vy)];
                const double wf = ww[lin3(i, j, k + 1, g.nx, g.ny)], wn = ww[lin3(i, j, k, g.nx, g.ny)];
                rhs -= wr * (double)G.u[lin3(i + 1, j, k, ux, g.ny)];
                rhs += wl * (double)G.u[lin3(i, j, k, ux, g.ny)];
                rhs -= wt * (double)G.v[lin3(i, j + 1, k, g.nx, vy)];
                rhs += wb * (double)G.v[lin3(i, j, k, g.nx, vy)];
                rhs -= wf * (double)G.w[lin3(i, j, k + 1, g.nx, g.ny)];
                rhs += wn * (double)G.w[lin3(i, j, k, g.nx, g.ny)];
                const double ddt = (double)dt;
                const double pcd = (double)pc;
                auto side = [&](float pn, double w, double& coupling, bool upper) {
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
        if(S.rhs64) S.rhs64[o] = rhs;
        if(S.diag64) {
            S.diag64[o] = diag;
            S.cx64[o]   = cx;
            S.cy64[o]   = cy;
            S.cz64[o]   = cz;
        }
        S.rhs32[o]     = (float)rhs;
        S.diag32[o]    = (float)diag;
        S.cx32[o]      = (float)cx;
        S.cy32[o]      = (float)cy;
        S.cz32[o]      = (float)cz;
        S.invdiag32[o] = diag > 0 ? (float)(1.0 / diag) : 0.f;
    }
};

//==================================================================================================
// PCG (PCGSolver.cpp:143-199) with Jacobi preconditioning, FP32 vectors and FP64 scalars.
// Scalars and the stop flag live in device memory; every kernel returns at once after the stop
// flag is set, so the host can queue iterations in chunks without reading back per iteration.
// Reductions: per-block partials, then a one-block kernel folds them in fixed order.
//==================================================================================================
struct PcgState {
    double   rho, alpha, beta, tmp;
    float    tol, res, bmax, pad0;
    uint32_t iter, maxIter, done, success;
};

// Reductions are two-level and deterministic: each block of a main kernel reduces a fixed set of
// bricks (b, b + the block count, ... or a contiguous chunk) and writes one partial with a plain store; a
// one-block fold kernel then combines the partials in index order. Kernel boundaries order the two,
// so no device-scope fences are needed inside the main kernels.
constexpr int kMaxRedBlocks = 2048;

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

// fold n partials in fixed order (1024-thread block); sum (or max) valid in thread 0
__device__ __forceinline__ double fold(const double* __restrict__ p, int n, bool isMax, double* sh) {
    double a = 0;
    for(int i = threadIdx.x; i < n; i += 1024) a = isMax ? fmax(a, p[i]) : a + p[i];
    return isMax ? block_max<1024>(a, sh) : block_sum<1024>(a, sh);
}

// init: x = 0, r = b, s = 0, z = M^-1 r; partials of z.r and max|b|
struct k_pcg_init {
    static constexpr int kThreads = 512;
    const float* __restrict__ b;
    const float* __restrict__ invdiag;
    float* __restrict__ x;
    float* __restrict__ r;
    float* __restrict__ z;
    float* __restrict__ s;
    uint32_t nSlots;
    double* __restrict__ pa;
    double* __restrict__ pb;
    int jacobi;
    __device__ void operator()(Blk blk) const {
        __shared__ double sh[8];
        double            rz = 0, m = 0;
        for(uint32_t slot = blk.id; slot < nSlots; slot += blk.n) {
            const size_t o  = (size_t)slot * 512 + threadIdx.x;
            const float  bv = b[o];
            x[o]            = 0.f;
            r[o]            = bv;
            s[o]            = 0.f;
            if(jacobi) {
                const float zv = invdiag[o] * bv;
                z[o]           = zv;
                rz += (double)zv * (double)bv;
            }
            m = fmax(m, (double)fabsf(bv));
        }
        if(jacobi) rz = block_sum<512>(rz, sh);
        m = block_max<512>(m, sh);
        if(threadIdx.x == 0) {
            if(jacobi) pa[blk.id] = rz;
            pb[blk.id] = m;
        }
    }
};

// pa: na partials of z.r, pb: nb partials of max|b|
struct k_pcg_init_fin {
    static constexpr int kThreads = 1024;
    PcgState* st;
    const double* __restrict__ pa;
    int na;
    const double* __restrict__ pb;
    int nb;
    float relTol;
    uint32_t maxIter;
    __device__ void operator()(Blk blk) const {
        __shared__ double sh[16];
        const double      A = fold(pa, na, false, sh);
        const double      B = fold(pb, nb, true, sh);
        if(threadIdx.x != 0) return;
        st->bmax    = (float)B;
        st->res     = (float)B;
        st->iter    = 0;
        st->maxIter = maxIter;
        st->beta    = 0;
        st->rho     = A;
        st->done    = 0;
        st->success = 0;
        st->tol     = relTol * (float)B;  // PCGSolver.cpp:160
        if((float)B < kFltMin) {          // PCGSolver.cpp:155-158
            st->done    = 1;
            st->success = 1;
        } else if(A < (double)kFltMin || isnan(A)) {  // PCGSolver.cpp:164-167
            st->done = 1;
        } else if(maxIter == 0) {
            st->done = 1;
        }
    }
};

// s_out = z + beta s_in ; q = A s_out ; partial s_out . q
// the slots of this block when the grid shares n slots: one contiguous chunk per block, and the
// chunks of one XCD's blocks (equal block index mod 8) adjacent, so that brick faces read from
// neighbouring bricks are mostly in that XCD's L2
__device__ __forceinline__ void block_chunk(Blk blk, uint32_t n, uint32_t& s0, uint32_t& s1) {
    const uint32_t G = blk.n, b = blk.id, xcd = b % 8u, q = G / 8u, r = G % 8u;
    const uint32_t v = xcd * q + min(xcd, r) + b / 8u;
    s0               = (uint32_t)(((uint64_t)n * v) / G);
    s1               = (uint32_t)(((uint64_t)n * (v + 1)) / G);
}

struct k_pcg_spmv {
    static constexpr int kThreads = 512;
    const int32_t* __restrict__ nbr;
    const float* __restrict__ z;
    const float* __restrict__ sin;
    float* __restrict__ sout;
    float* __restrict__ q;
    const float* __restrict__ diag;
    const float* __restrict__ cx;
    const float* __restrict__ cy;
    const float* __restrict__ cz;
    uint32_t nSlots;
    const PcgState* __restrict__ st;
    double* __restrict__ pa;
    __device__ void operator()(Blk blk) const {
        __shared__ double sh[8];
        if(st && st->done) return;
        const float beta = (float)st->beta;
        const int   l    = threadIdx.x;
        double      d    = 0;
        uint32_t    s0, s1;
        block_chunk(blk, nSlots, s0, s1);
        for(uint32_t slot = s0; slot < s1; ++slot) {
            const int     s   = (int)slot;
            const size_t  o   = (size_t)s * 512 + l;
            auto          S   = [&](int32_t idx) { return z[idx] + beta * sin[idx]; };
            const float   sc  = z[o] + beta * sin[o];
            const int32_t ixm = nbr_index(nbr, s, l, 0), ixp = nbr_index(nbr, s, l, 1);
            const int32_t iym = nbr_index(nbr, s, l, 2), iyp = nbr_index(nbr, s, l, 3);
            const int32_t izm = nbr_index(nbr, s, l, 4), izp = nbr_index(nbr, s, l, 5);
            // a missing neighbour reads the cell itself with a zero coupling (adds exactly nothing), so
            // all loads are unconditional
            const int32_t oi = (int32_t)o;
            const int32_t jxm = ixm >= 0 ? ixm : oi, jxp = ixp >= 0 ? ixp : oi, jym = iym >= 0 ? iym : oi;
            const int32_t jyp = iyp >= 0 ? iyp : oi, jzm = izm >= 0 ? izm : oi, jzp = izp >= 0 ? izp : oi;
            const float   czm = cz[jzm], cym = cy[jym], cxm = cx[jxm], cxp = cx[o], cyp = cy[o], czp = cz[o];
            const float   szm = S(jzm), sym = S(jym), sxm = S(jxm), sxp = S(jxp), syp = S(jyp), szp = S(jzp);
            float         acc = 0.f;
            acc += (izm >= 0 ? czm : 0.f) * szm;
            acc += (iym >= 0 ? cym : 0.f) * sym;
            acc += (ixm >= 0 ? cxm : 0.f) * sxm;
            acc += diag[o] * sc;
            acc += (ixp >= 0 ? cxp : 0.f) * sxp;
            acc += (iyp >= 0 ? cyp : 0.f) * syp;
            acc += (izp >= 0 ? czp : 0.f) * szp;
            sout[o] = sc;
            q[o]    = acc;
            d += (double)sc * (double)acc;
        }
        d = block_sum<512>(d, sh);
        if(threadIdx.x == 0) pa[blk.id] = d;
    }
};

// tmp = s.q ; alpha = rho / tmp (PCGSolver.cpp:174-180)
struct k_pcg_alpha {
    static constexpr int kThreads = 1024;
    PcgState* st;
    const double* __restrict__ pa;
    int n;
    __device__ void operator()(Blk blk) const {
        __shared__ double sh[16];
        if(st && st->done) return;
        const double A = fold(pa, n, false, sh);
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
};

// x += alpha s ; r -= alpha q ; z = M^-1 r ; partials z.r and max|r|
struct k_pcg_update {
    static constexpr int kThreads = 512;
    const float* __restrict__ s;
    const float* __restrict__ q;
    const float* __restrict__ invdiag;
    float* __restrict__ x;
    float* __restrict__ r;
    float* __restrict__ z;
    uint32_t nSlots;
    const PcgState* __restrict__ st;
    double* __restrict__ pa;
    double* __restrict__ pb;
    int jacobi;
    __device__ void operator()(Blk blk) const {
        __shared__ double sh[8];
        if(st && st->done) return;
        const float alpha = (float)st->alpha;
        double      rz = 0, m = 0;
        for(uint32_t slot = blk.id; slot < nSlots; slot += blk.n) {
            const size_t o  = (size_t)slot * 512 + threadIdx.x;
            const float  xv = x[o] + alpha * s[o];
            const float  rv = r[o] + (-alpha) * q[o];
            x[o]            = xv;
            r[o]            = rv;
            if(jacobi) {  // z = M^-1 r fused here; the multigrid preconditioner runs after this kernel
                const float zv = invdiag[o] * rv;
                z[o]           = zv;
                rz += (double)zv * (double)rv;
            }
            m = fmax(m, (double)fabsf(rv));
        }
        if(jacobi) rz = block_sum<512>(rz, sh);
        m = block_max<512>(m, sh);
        if(threadIdx.x == 0) {
            if(jacobi) pa[blk.id] = rz;
            pb[blk.id] = m;
        }
    }
};

// res = max|r| ; stop test (PCGSolver.cpp:182-187) ; beta = rz / rho (:190-194)
// pa: na partials of z.r, pb: nb partials of max|r|
struct k_pcg_beta {
    static constexpr int kThreads = 1024;
    PcgState* st;
    const double* __restrict__ pa;
    int na;
    const double* __restrict__ pb;
    int nb;
    __device__ void operator()(Blk blk) const {
        __shared__ double sh[16];
        if(st && st->done) return;
        const double A = fold(pa, na, false, sh);
        const double B = fold(pb, nb, true, sh);
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
};

//--------------------------------------------------------------------------------------------------
// FP64 defect correction: r64 = b64 - A64 x64, max|r64|, max|b64|, r32 = (float) r64 (inner rhs)
//--------------------------------------------------------------------------------------------------
struct OuterState {
    double res, bmax;
};

struct k_residual64 {
    static constexpr int kThreads = 512;
    const int32_t* __restrict__ nbr;
    const double* __restrict__ x;
    const double* __restrict__ b;
    const double* __restrict__ diag;
    const double* __restrict__ cx;
    const double* __restrict__ cy;
    const double* __restrict__ cz;
    float* __restrict__ r32;
    uint32_t nSlots;
    double* __restrict__ pa;
    double* __restrict__ pb;
    int first;
    __device__ void operator()(Blk blk) const {
        __shared__ double sh[8];
        const int         l = threadIdx.x;
        double            m = 0, bm = 0;
        for(uint32_t slot = blk.id; slot < nSlots; slot += blk.n) {
            const int    s  = (int)slot;
            const size_t o  = (size_t)s * 512 + l;
            double       ax = 0.0;
            if(!first) {
                const int32_t ixm = nbr_index(nbr, s, l, 0), ixp = nbr_index(nbr, s, l, 1);
                const int32_t iym = nbr_index(nbr, s, l, 2), iyp = nbr_index(nbr, s, l, 3);
                const int32_t izm = nbr_index(nbr, s, l, 4), izp = nbr_index(nbr, s, l, 5);
                if(izm >= 0) ax += cz[izm] * x[izm];
                if(iym >= 0) ax += cy[iym] * x[iym];
                if(ixm >= 0) ax += cx[ixm] * x[ixm];
                ax += diag[o] * x[o];
                if(ixp >= 0) ax += cx[o] * x[ixp];
                if(iyp >= 0) ax += cy[o] * x[iyp];
                if(izp >= 0) ax += cz[o] * x[izp];
            }
            const double bv = b[o];
            const double rv = bv - ax;
            r32[o]          = (float)rv;
            m               = fmax(m, fabs(rv));
            bm              = fmax(bm, fabs(bv));
        }
        m  = block_max<512>(m, sh);
        bm = block_max<512>(bm, sh);
        if(threadIdx.x == 0) {
            pa[blk.id] = m;
            pb[blk.id] = bm;
        }
    }
};

struct k_residual64_fin {
    static constexpr int kThreads = 1024;
    OuterState* ost;
    const double* __restrict__ pa;
    const double* __restrict__ pb;
    int n;
    __device__ void operator()(Blk blk) const {
        __shared__ double sh[16];
        const double      M  = fold(pa, n, true, sh);
        const double      BM = fold(pb, n, true, sh);
        if(threadIdx.x == 0) {
            ost->res  = M;
            ost->bmax = BM;
        }
    }
};

// the previous solve's pressure as the initial guess of the next refined solve: cell l of new slot s
// takes the old value of the same cell when it is fluid now (phi < 0) and its brick had a slot then
// (prevBrickSlot: the old brick -> slot table), else 0
struct k_warm_start {
    static constexpr int kThreads = 512;
    const int32_t* __restrict__ slotBrick;
    const int32_t* __restrict__ prevBrickSlot;
    const double* __restrict__  xprev;
    const float* __restrict__   phi;
    GridGeom                    g;
    uint32_t                    nSlots;
    double* __restrict__        x;
    __device__ void operator()(Blk blk) const {
        const int l = threadIdx.x;
        for(uint32_t slot = blk.id; slot < nSlots; slot += blk.n) {
            const int32_t b  = slotBrick[slot];
            const int32_t os = prevBrickSlot[b];
            const int     bx = b % g.nbx, by = (b / g.nbx) % g.nby, bz = b / (g.nbx * g.nby);
            const int     i = bx * 8 + (l & 7), j = by * 8 + ((l >> 3) & 7), k = bz * 8 + (l >> 6);
            const bool    fluid = i < g.nx && j < g.ny && k < g.nz && phi[lin3(i, j, k, g.nx, g.ny)] < 0.f;
            x[(size_t)slot * 512 + l] = (fluid && os >= 0) ? xprev[(size_t)os * 512 + l] : 0.0;
        }
    }
};

struct k_accumulate64 {
    static constexpr int kThreads = 512;
    double* __restrict__ x64;
    const float* __restrict__ d;
    uint32_t nSlots;
    int first;
    __device__ void operator()(Blk blk) const {
        for(uint32_t slot = blk.id; slot < nSlots; slot += blk.n) {
            const size_t o = (size_t)slot * 512 + threadIdx.x;
            x64[o]         = first ? (double)d[o] : x64[o] + (double)d[o];
        }
    }
};

//==================================================================================================
// updateProjectedVelocity (PIC_Solver.cpp:911-981), one thread per face; faces the reference does
// not update are zeroed with valid = 0 (its serial pass, :973-980)
//==================================================================================================
template<class PT> struct PressView {
    const PT*      x;
    const int32_t* brickSlot;
};

template<class PT> __device__ __forceinline__ PT pressure_at(const PressView<PT>& pv, const GridGeom& g, int i, int j, int k) {
    const int32_t s = pv.brickSlot[((k >> 3) * g.nby + (j >> 3)) * g.nbx + (i >> 3)];
    return pv.x[(size_t)s * 512 + ((k & 7) * 8 + (j & 7)) * 8 + (i & 7)];
}

template<class PT>
struct k_update_velocity {
    static constexpr int kThreads = kBlock;
    float* __restrict__ vel;
    uint8_t* __restrict__ valid;
    const float* __restrict__ weight;
    const float* __restrict__ phi;
    PressView<PT> pv;
    GridGeom g;
    int d;
    float dt;
    __device__ void operator()(Blk blk) const {
        const int    sx = g.nx + (d == 0), sy = g.ny + (d == 1), sz = g.nz + (d == 2);
        const size_t t  = (size_t)blk.id * blockDim.x + threadIdx.x;
        if(t >= (size_t)sx * sy * sz) return;
        int i, j, k;
        unlin3((uint32_t)t, sx, sy, i, j, k);
        const int ai = d == 0 ? i : (d == 1 ? j : k);       // index along the face normal
        const int an = d == 0 ? g.nx : (d == 1 ? g.ny : g.nz);
        bool      upd = false;
        if(ai > 0 && ai < an && weight[t] > 0.f) {
            const int   li = i - (d == 0), lj = j - (d == 1), lk = k - (d == 2);
            const float pc = phi[lin3(i, j, k, g.nx, g.ny)];
            const float pl = phi[lin3(li, lj, lk, g.nx, g.ny)];
            if(pc < 0 || pl < 0) {
                const PT prc = pc < 0 ? pressure_at(pv, g, i, j, k) : PT(0);
                const PT prl = pl < 0 ? pressure_at(pv, g, li, lj, lk) : PT(0);
                PT       th  = fraction_inside2((PT)pl, (PT)pc);
                th           = th > PT(0.01) ? th : PT(0.01);  // std::max(0.01, th)
                vel[t]       = (float)((PT)vel[t] - (PT)dt * (prc - prl) / th);
                upd          = true;
            }
        }
        if(!upd) vel[t] = 0.f;
        valid[t] = upd ? 1 : 0;
    }
};

//==================================================================================================
// mapGrid2Particles with the FLIP blend (FLIP_Solver.cpp:99-133, MathHelpers.h:549-551)
//==================================================================================================
struct OldGrid {
    const float *u, *v, *w;
};

struct k_g2p {
    static constexpr int kThreads = kBlock;
    Particles P;
    uint32_t n;
    VelGrid G;
    OldGrid O;
    GridGeom g;
    float ratio;
    __device__ void operator()(Blk blk) const {
        const uint32_t p = blk.id * blockDim.x + threadIdx.x;
        if(p >= n) return;
        const float4 gq = P.gp[p];
        const float  gx = gq.x, gy = gq.y, gz = gq.z;
        float       ux, uy, uz;
        vel_from_grid(G, g, gx, gy, gz, ux, uy, uz);
        const float dx = interp3_diff(G.u, O.u, g.nx + 1, g.ny, g.nz, gx, gy - 0.5f, gz - 0.5f);
        const float dy = interp3_diff(G.v, O.v, g.nx, g.ny + 1, g.nz, gx - 0.5f, gy, gz - 0.5f);
        const float dz = interp3_diff(G.w, O.w, g.nx, g.ny, g.nz + 1, gx - 0.5f, gy - 0.5f, gz);
        const float a  = 1.0f - ratio;
        P.vx[p]        = a * ux + ratio * (P.vx[p] + dx);
        P.vy[p]        = a * uy + ratio * (P.vy[p] + dy);
        P.vz[p]        = a * uz + ratio * (P.vz[p] + dz);
    }
};

//==================================================================================================
// APIC / AFLIP (APIC_Solver.cpp:42-125, AFLIP_Solver.cpp:43-140): each particle carries an affine
// matrix C (row d = the gradient of velocity component d), stored as nine arrays in the particles'
// storage order and moved with them by the binning permute.
//==================================================================================================
struct k_permute_affine {
    static constexpr int kThreads = kBlock;
    Affine in;
    Affine out;
    const uint32_t* __restrict__ perm;
    uint32_t n;
    __device__ void operator()(Blk blk) const {
        const uint32_t i = blk.id * blockDim.x + threadIdx.x;
        if(i >= n) return;
        const uint32_t s = perm[i];
    #pragma unroll
        for(int k = 0; k < 9; ++k) out.c[k][i] = in.c[k][s];
    }
};

// APIC mapGrid2Particles (APIC_Solver.cpp:94-115): v = the interpolated grid velocity, C_d = the
// gradient of component d at gridPos - extra_d. AFLIP (AFLIP_Solver.cpp:94-140): both blended like
// FLIP, lerp(grid, particle + change, PIC_FLIP_ratio), the changes taken from vel - backup.
template<int kScheme>
struct k_g2p_affine {
    static constexpr int kThreads = kBlock;
    Particles P;
    Affine A;
    uint32_t n;
    VelGrid G;
    OldGrid O;
    GridGeom g;
    float ratio;
    __device__ void operator()(Blk blk) const {
    #pragma clang fp contract(off)
        const uint32_t p = blk.id * blockDim.x + threadIdx.x;
        if(p >= n) return;
        const float4 gq = P.gp[p];
        const float  gx = gq.x, gy = gq.y, gz = gq.z;
        float       v[3], gc[9];
        vel_from_grid(G, g, gx, gy, gz, v[0], v[1], v[2]);
        interp3_grad(G.u, nullptr, g.nx + 1, g.ny, g.nz, gx, gy - 0.5f, gz - 0.5f, g.dx, gc[0], gc[1], gc[2]);
        interp3_grad(G.v, nullptr, g.nx, g.ny + 1, g.nz, gx - 0.5f, gy, gz - 0.5f, g.dx, gc[3], gc[4], gc[5]);
        interp3_grad(G.w, nullptr, g.nx, g.ny, g.nz + 1, gx - 0.5f, gy - 0.5f, gz, g.dx, gc[6], gc[7], gc[8]);
        if(kScheme == 1) {
            P.vx[p] = v[0];
            P.vy[p] = v[1];
            P.vz[p] = v[2];
    #pragma unroll
            for(int k = 0; k < 9; ++k) A.c[k][p] = gc[k];
            return;
        }
        float dc[9];
        interp3_grad(G.u, O.u, g.nx + 1, g.ny, g.nz, gx, gy - 0.5f, gz - 0.5f, g.dx, dc[0], dc[1], dc[2]);
        interp3_grad(G.v, O.v, g.nx, g.ny + 1, g.nz, gx - 0.5f, gy, gz - 0.5f, g.dx, dc[3], dc[4], dc[5]);
        interp3_grad(G.w, O.w, g.nx, g.ny, g.nz + 1, gx - 0.5f, gy - 0.5f, gz, g.dx, dc[6], dc[7], dc[8]);
        const float dvx = interp3_diff(G.u, O.u, g.nx + 1, g.ny, g.nz, gx, gy - 0.5f, gz - 0.5f);
        const float dvy = interp3_diff(G.v, O.v, g.nx, g.ny + 1, g.nz, gx - 0.5f, gy, gz - 0.5f);
        const float dvz = interp3_diff(G.w, O.w, g.nx, g.ny, g.nz + 1, gx - 0.5f, gy - 0.5f, gz);
        const float a   = 1.0f - ratio;
        P.vx[p]         = a * v[0] + ratio * (P.vx[p] + dvx);
        P.vy[p]         = a * v[1] + ratio * (P.vy[p] + dvy);
        P.vz[p]         = a * v[2] + ratio * (P.vz[p] + dvz);
    #pragma unroll
        for(int k = 0; k < 9; ++k) A.c[k][p] = a * gc[k] + ratio * (A.c[k][p] + dc[k]);
    }
};

// max |a| over an array (timestepCFL, PIC_Solver.cpp:342-349); result via atomicMax on the bits
struct k_maxabs {
    static constexpr int kThreads = kBlock;
    const float* __restrict__ a;
    size_t n;
    uint32_t* out;
    __device__ void operator()(Blk blk) const {
        __shared__ float sh[kBlock / 64];
        float            m = 0.f;
        for(size_t t = (size_t)blk.id * blockDim.x + threadIdx.x; t < n; t += (size_t)blk.n * blockDim.x) m = fmaxf(m, fabsf(a[t]));
        m = wave_max(m);
        if((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
        __syncthreads();
        if(threadIdx.x == 0) {
            for(int i = 1; i < kBlock / 64; ++i) m = fmaxf(m, sh[i]);
            atomicMax(out, __float_as_uint(m));
        }
    }
};

}  // namespace bnn

namespace bnn {

//==================================================================================================
// Multigrid preconditioner (extension; the reference's IC(0) is sequential). Level l+1 aggregates
// 2x2x2 cells of level l (piecewise-constant prolongation P); its operator is the Galerkin product
// P^T A P, which for a 7-point operator is again 7-point:
//   diag_C = sum_c diag_c + 2 * (couplings between children of C)
//   cxp_C  = sum of the 4 +x couplings leaving C's children with odd local x, etc.
// Every level stores 8^3-cell bricks (slot tables as on the fine level). The V-cycle is symmetric
// (red-black Gauss-Seidel: red then black before the coarse correction, black then red after), so
// M^-1 is symmetric positive definite and usable inside CG.
//==================================================================================================
struct LevelDev {
    int            nx, ny, nz, nbx, nby, nbz;
    const int32_t* brickSlot;
    const int32_t* slotBrick;
    const int32_t* nbr;
    const float *  diag, *cx, *cy, *cz, *invdiag;
    float *        x, *b, *r;
    uint32_t       nSlots;
    const int32_t* ptab;  // per slot: parent coarse brick and its six neighbours (8 ints), or null
};

__device__ __forceinline__ int32_t level_cell_index(const LevelDev& L, int i, int j, int k) {
    if(i < 0 || j < 0 || k < 0 || i >= L.nx || j >= L.ny || k >= L.nz) return -1;
    const int32_t s = L.brickSlot[((k >> 3) * L.nby + (j >> 3)) * L.nbx + (i >> 3)];
    return s < 0 ? -1 : s * 512 + ((k & 7) * 8 + (j & 7)) * 8 + (i & 7);
}

// Every coarse level's active-brick flags in one pass: a brick of level l is active when a finest
// active brick lies below it (the parent of an active brick is active, level by level), and the
// level-l brick of finest brick (bx, by, bz) is (bx >> l, by >> l, bz >> l). Each writes the value 1,
// so concurrent writes to one flag agree.
constexpr int kMaxMgLevels = 16;
struct k_mg_flag_levels {
    static constexpr int kThreads = kBlock;
    const int32_t* __restrict__ fineSlotBrick;
    uint32_t nFine;
    int fnbx;
    int fnby;
    int nLevels;  // coarse levels 1 .. nLevels
    int cnbx[kMaxMgLevels];
    int cnby[kMaxMgLevels];
    uint32_t* flag[kMaxMgLevels];
    __device__ void operator()(Blk blk) const {
        const uint32_t s = blk.id * blockDim.x + threadIdx.x;
        if(s >= nFine) return;
        const int b = fineSlotBrick[s];
        const int bx = b % fnbx, by = (b / fnbx) % fnby, bz = b / (fnbx * fnby);
        for(int l = 1; l <= nLevels; ++l) flag[l][((bz >> l) * cnby[l] + (by >> l)) * cnbx[l] + (bx >> l)] = 1u;
    }
};

// Galerkin coefficients of the coarse level (one block per coarse brick, thread = coarse cell)
struct k_mg_coarsen {
    static constexpr int kThreads = 512;
    LevelDev F;
    LevelDev C;
    float* __restrict__ cdiag;
    float* __restrict__ ccx;
    float* __restrict__ ccy;
    float* __restrict__ ccz;
    float* __restrict__ cinv;
    __device__ void operator()(Blk blk) const {
        const int s = blk.id, l = threadIdx.x;
        const int b = C.slotBrick[s];
        const int bx = b % C.nbx, by = (b / C.nbx) % C.nby, bz = b / (C.nbx * C.nby);
        const int I = bx * 8 + (l & 7), J = by * 8 + ((l >> 3) & 7), K = bz * 8 + (l >> 6);
        float     d = 0.f, cx = 0.f, cy = 0.f, cz = 0.f, inner = 0.f;
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
        cdiag[o]       = d;
        ccx[o]         = cx;
        ccy[o]         = cy;
        ccz[o]         = cz;
        cinv[o]        = d > 0.f ? 1.0f / d : 0.f;
    }
};

// sum of the six couplings times x. A missing neighbour reads the cell's own x (finite) with a zero
// coupling, which adds exactly nothing, so every load is issued before the first use.
__device__ __forceinline__ float level_offdiag_sum(const LevelDev& L, int s, int l, const float* __restrict__ x) {
    const int32_t ixm = nbr_index(L.nbr, s, l, 0), ixp = nbr_index(L.nbr, s, l, 1);
    const int32_t iym = nbr_index(L.nbr, s, l, 2), iyp = nbr_index(L.nbr, s, l, 3);
    const int32_t izm = nbr_index(L.nbr, s, l, 4), izp = nbr_index(L.nbr, s, l, 5);
    const int32_t o   = s * 512 + l;
    const int32_t jxm = ixm >= 0 ? ixm : o, jxp = ixp >= 0 ? ixp : o, jym = iym >= 0 ? iym : o;
    const int32_t jyp = iyp >= 0 ? iyp : o, jzm = izm >= 0 ? izm : o, jzp = izp >= 0 ? izp : o;
    const float   czm = L.cz[jzm], cym = L.cy[jym], cxm = L.cx[jxm], cxp = L.cx[o], cyp = L.cy[o], czp = L.cz[o];
    const float   xzm = x[jzm], xym = x[jym], xxm = x[jxm], xxp = x[jxp], xyp = x[jyp], xzp = x[jzp];
    float         acc = 0.f;
    acc += (izm >= 0 ? czm : 0.f) * xzm;
    acc += (iym >= 0 ? cym : 0.f) * xym;
    acc += (ixm >= 0 ? cxm : 0.f) * xxm;
    acc += (ixp >= 0 ? cxp : 0.f) * xxp;
    acc += (iyp >= 0 ? cyp : 0.f) * xyp;
    acc += (izp >= 0 ? czp : 0.f) * xzp;
    return acc;
}

__device__ __forceinline__ int cell_color(const LevelDev& L, int s, int l) {
    const int b = L.slotBrick[s];
    const int bx = b % L.nbx, by = (b / L.nbx) % L.nby, bz = b / (L.nbx * L.nby);
    return ((bx * 8 + (l & 7)) + (by * 8 + ((l >> 3) & 7)) + (bz * 8 + (l >> 6))) & 1;
}

// Launch order -> brick slot, a bijection on [0, n) that gives each XCD (the blocks with equal
// b % 8) one contiguous range of slots: slots are ordered by brick index, so a brick's neighbours,
// whose faces the stencils read, mostly sit in the same XCD's L2.
__device__ __forceinline__ uint32_t xcd_slot(uint32_t b, uint32_t n) {
    const uint32_t xcd = b % 8u, q = n / 8u, r = n % 8u;
    return xcd * q + min(xcd, r) + b / 8u;
}

// The coarse bricks a fine brick's cells and halo can have as parents: its own parent (index 6) and
// that parent's six neighbours (-x, +x, -y, +y, -z, +z), or -1. Stored per fine slot (8 ints, the
// last unused) when the hierarchy is built.
struct k_mg_parent_table {
    static constexpr int kThreads = kBlock;
    LevelDev F;
    LevelDev C;
    int32_t* __restrict__ tab;
    __device__ void operator()(Blk blk) const {
        const uint32_t s = blk.id * blockDim.x + threadIdx.x;
        if(s >= F.nSlots) return;
        const int     b  = F.slotBrick[s];
        const int     bx = b % F.nbx, by = (b / F.nbx) % F.nby, bz = b / (F.nbx * F.nby);
        const int32_t cs = C.brickSlot[((bz >> 1) * C.nby + (by >> 1)) * C.nbx + (bx >> 1)];
        for(int d = 0; d < 6; ++d) tab[8 * s + d] = cs >= 0 ? C.nbr[6 * cs + d] : -1;
        tab[8 * s + 6] = cs;
        tab[8 * s + 7] = -1;
    }
};

// xpre + omega * x_coarse(parent) for an unknown, xpre otherwise (x + omega P x_c). cs is the coarse
// brick holding the parent of fine cell (i, j, k) (-1: none). An unknown's parent always lies inside
// the coarse grid; for other cells the coarse value is loaded but not used. The loads are
// unconditional (coarse cell 0 stands in when there is no parent brick).
__device__ __forceinline__ float mg_prolonged(const LevelDev& F, const float* __restrict__ xpre, const LevelDev& C, int32_t cs,
                                              float omega, int32_t o, int i, int j, int k) {
    const int32_t c  = cs < 0 ? 0 : cs * 512 + (((k >> 1) & 7) * 8 + ((j >> 1) & 7)) * 8 + ((i >> 1) & 7);
    const float   xc = C.x[c];
    const float   xv = xpre[o];
    const bool    un = F.invdiag[o] != 0.f && cs >= 0;
    return un ? xv + omega * xc : xv;
}

// one cell of the brick smoother: its LDS index and its row of A (a coupling to a missing neighbour
// is 0 and that neighbour's halo entry holds 0, so the six terms are added unconditionally)
struct SmoothCell {
    int   c;
    float czm, cym, cxm, cxp, cyp, czp, bv, inv;
};

template<int kProlong>
__device__ __forceinline__ SmoothCell smooth_cell_setup(const LevelDev& L, const LevelDev& C, int32_t parent, float omega, int s, int l,
                                                        int bx, int by, int bz, float* xs) {
    const int     lx = l & 7, ly = (l >> 3) & 7, lz = l >> 6;
    const int     c  = ((lz + 1) * 10 + (ly + 1)) * 10 + (lx + 1);
    const size_t  o  = (size_t)s * 512 + l;
    const int32_t ixm = nbr_index(L.nbr, s, l, 0), ixp = nbr_index(L.nbr, s, l, 1);
    const int32_t iym = nbr_index(L.nbr, s, l, 2), iyp = nbr_index(L.nbr, s, l, 3);
    const int32_t izm = nbr_index(L.nbr, s, l, 4), izp = nbr_index(L.nbr, s, l, 5);
    SmoothCell    R;
    R.c   = c;
    const float   czm = L.cz[izm >= 0 ? izm : (int32_t)o], cym = L.cy[iym >= 0 ? iym : (int32_t)o];
    const float   cxm = L.cx[ixm >= 0 ? ixm : (int32_t)o], cxp = L.cx[o], cyp = L.cy[o], czp = L.cz[o];
    R.czm = izm >= 0 ? czm : 0.f;
    R.cym = iym >= 0 ? cym : 0.f;
    R.cxm = ixm >= 0 ? cxm : 0.f;
    R.cxp = ixp >= 0 ? cxp : 0.f;
    R.cyp = iyp >= 0 ? cyp : 0.f;
    R.czp = izp >= 0 ? czp : 0.f;
    R.bv  = L.b[o];
    R.inv = L.invdiag[o];
    xs[c] = kProlong == 1 ? mg_prolonged(L, L.r, C, parent, omega, (int32_t)o, bx * 8 + lx, by * 8 + ly, bz * 8 + lz)
                          : (kProlong == 2 ? L.r[o] : 0.f);
    return R;
}

// the brick's one-cell halo (6 faces x 64 cells) in LDS: the neighbouring bricks' starting values
// (0 for the pre-smoother and for missing neighbours). A wave's 64 entries share one face.
template<int kProlong>
__device__ __forceinline__ void smooth_halo(const LevelDev& L, const LevelDev& C, int32_t parent, float omega, int s, int bx, int by,
                                            int bz, float* xs) {
    for(int h = threadIdx.x; h < 384; h += blockDim.x) {
        const int d = h >> 6, p = h & 63, a = p & 7, e = p >> 3;
        int       lx, ly, lz;  // the brick cell next to the halo entry
        if(d < 2) {
            lx = d == 0 ? 0 : 7;
            ly = a;
            lz = e;
        } else if(d < 4) {
            lx = a;
            ly = d == 2 ? 0 : 7;
            lz = e;
        } else {
            lx = a;
            ly = e;
            lz = d == 4 ? 0 : 7;
        }
        const int sx = d == 0 ? -1 : d == 1 ? 1 : 0, sy = d == 2 ? -1 : d == 3 ? 1 : 0, sz = d == 4 ? -1 : d == 5 ? 1 : 0;
        const int c  = ((lz + 1 + sz) * 10 + (ly + 1 + sy)) * 10 + (lx + 1 + sx);
        float     v  = 0.f;
        if(kProlong == 2) {
            const int32_t n = nbr_index(L.nbr, s, (lz * 8 + ly) * 8 + lx, d);
            if(n >= 0) v = L.r[n];
        } else if(kProlong == 1) {
            const int32_t n = nbr_index(L.nbr, s, (lz * 8 + ly) * 8 + lx, d);
            if(n >= 0) {
                // the neighbour brick's parent: this brick's parent, or its neighbour across the face
                // when this brick is on the parent's face (both uniform over the wave)
                const int  i = bx * 8 + lx + sx, j = by * 8 + ly + sy, k = bz * 8 + lz + sz;
                const bool cross = ((i >> 4) != (bx >> 1)) || ((j >> 4) != (by >> 1)) || ((k >> 4) != (bz >> 1));
                const int32_t cs = cross ? L.ptab[8 * s + d] : parent;
                v = mg_prolonged(L, L.r, C, cs, omega, n, i, j, k);
            }
        }
        xs[c] = v;
    }
}

// four cells of one colour: their neighbours all have the other colour, so every neighbour value is
// read before any of the four is written (the reads need not wait for the writes)
__device__ __forceinline__ void smooth_cells_update4(const SmoothCell* R, float* xs) {
    float v[4][6];
#pragma unroll
    for(int q = 0; q < 4; ++q) {
        v[q][0] = xs[R[q].c - 100];
        v[q][1] = xs[R[q].c - 10];
        v[q][2] = xs[R[q].c - 1];
        v[q][3] = xs[R[q].c + 1];
        v[q][4] = xs[R[q].c + 10];
        v[q][5] = xs[R[q].c + 100];
    }
#pragma unroll
    for(int q = 0; q < 4; ++q) {
        float acc = 0.f;
        acc += R[q].czm * v[q][0];
        acc += R[q].cym * v[q][1];
        acc += R[q].cxm * v[q][2];
        acc += R[q].cxp * v[q][3];
        acc += R[q].cyp * v[q][4];
        acc += R[q].czp * v[q][5];
        xs[R[q].c] = (R[q].bv - acc) * R[q].inv;
    }
}

__device__ __forceinline__ void smooth_cell_update(const SmoothCell& R, float* xs) {
    float acc = 0.f;
    acc += R.czm * xs[R.c - 100];
    acc += R.cym * xs[R.c - 10];
    acc += R.cxm * xs[R.c - 1];
    acc += R.cxp * xs[R.c + 1];
    acc += R.cyp * xs[R.c + 10];
    acc += R.czp * xs[R.c + 100];
    xs[R.c] = (R.bv - acc) * R.inv;
}

// coarse-grid correction in place: L.r (the pre-smoothed x) += omega * P x_coarse on every unknown.
// One block per brick, two cells per thread; each cell's parent lies in the brick's parent brick.
struct k_mg_prolong {
    static constexpr int kThreads = 256;
    LevelDev F;
    LevelDev C;
    float omega;
    const PcgState* __restrict__ st;
    __device__ void operator()(Blk blk) const {
        if(st && st->done) return;
        for(uint32_t v = blk.id; v < F.nSlots; v += blk.n) {
            const int     s      = (int)xcd_slot(v, F.nSlots);
            const int     b      = F.slotBrick[s];
            const int     bx     = b % F.nbx, by = (b / F.nbx) % F.nby, bz = b / (F.nbx * F.nby);
            const int32_t parent = F.ptab[8 * s + 6];
            for(int h = 0; h < 2; ++h) {
                const int    l = (int)threadIdx.x + 256 * h;
                const size_t o = (size_t)s * 512 + l;
                F.r[o]         = mg_prolonged(F, F.r, C, parent, omega, (int32_t)o, bx * 8 + (l & 7), by * 8 + ((l >> 3) & 7), bz * 8 + (l >> 6));
            }
        }
    }
};

// Brick-hybrid Gauss-Seidel: one block per brick runs `sweeps` red-black sweeps (pre-smoother) or
// black-red sweeps (post-smoother) on its 512 cells in LDS, with the x of neighbouring bricks held at
// their values on entry (Jacobi between bricks). For the brick's own system A_BB the red-black and
// black-red sequences are transposes of each other, so the pre- and post-smoothers are adjoint and
// the V-cycle stays symmetric. Pre-smoother (kProlong == 0): starts from x = 0 (every V-cycle does)
// and writes L.r. Post-smoother: starts from the coarse-grid-corrected L.r + omega * P x_coarse for
// the brick and its halo, formed on the fly (kProlong == 1: no extra launch, for the small coarse
// levels) or already added by k_mg_prolong (kProlong == 2: the finest level, where a streaming pass
// is cheaper than the dependent loads), and writes L.x. Input and output are different arrays, so
// no block reads what another writes. Each of the 256 threads owns
// an x-adjacent pair of cells, one of each colour, so every thread updates one cell per half-sweep.
// kDot (finest level, post-smoother): also writes the brick's partial of z.r (z = the result,
// r = L.b) to pz[slot], the PCG's next rho.
template<int kProlong, int kDot>
__device__ __forceinline__ void mg_smooth_one(const LevelDev& L, const LevelDev& C, float omega, int sweeps, double* __restrict__ pz, int s,
                                              float* xs, double* sh) {
    const int t = threadIdx.x, ly = (t >> 2) & 7, lz = t >> 5;
    const int l0 = (lz * 8 + ly) * 8 + 2 * (t & 3);        // even x; its colour is (ly + lz) & 1
    const int lr = ((ly + lz) & 1) ? l0 + 1 : l0, lb = ((ly + lz) & 1) ? l0 : l0 + 1;  // red (colour 0), black
    const int     b      = L.slotBrick[s];
    const int     bx     = b % L.nbx, by = (b / L.nbx) % L.nby, bz = b / (L.nbx * L.nby);
    const int32_t parent = kProlong == 1 ? L.ptab[8 * s + 6] : -1;
    const SmoothCell R = smooth_cell_setup<kProlong>(L, C, parent, omega, s, lr, bx, by, bz, xs);
    const SmoothCell K = smooth_cell_setup<kProlong>(L, C, parent, omega, s, lb, bx, by, bz, xs);
    smooth_halo<kProlong>(L, C, parent, omega, s, bx, by, bz, xs);
    __syncthreads();
    for(int it = 0; it < sweeps; ++it) {
        smooth_cell_update(kProlong ? K : R, xs);
        __syncthreads();
        smooth_cell_update(kProlong ? R : K, xs);
        __syncthreads();
    }
    float*      out             = kProlong ? L.x : L.r;
    const float zr = xs[R.c], zb = xs[K.c];
    out[(size_t)s * 512 + lr]   = zr;
    out[(size_t)s * 512 + lb]   = zb;
    if(kDot) {
        const double d = block_sum<256>((double)zr * (double)R.bv + (double)zb * (double)K.bv, sh);
        if(t == 0) pz[s] = d;
    }
    __syncthreads();  // xs is reused by the workgroup's next brick
}

template<int kProlong, int kDot>
struct k_mg_smooth_brick {
    static constexpr int kThreads = 256;
    LevelDev L;
    LevelDev C;
    float omega;
    int sweeps;
    const PcgState* __restrict__ st;
    double* __restrict__ pz;
    __device__ void operator()(Blk blk) const {
        __shared__ float xs[10 * 10 * 10];  // the brick and a one-cell halo, x-fastest
        __shared__ double sh[4];
        if(st && st->done) return;
        for(uint32_t v = blk.id; v < L.nSlots; v += blk.n)
            mg_smooth_one<kProlong, kDot>(L, C, omega, sweeps, pz, (int)xcd_slot(v, L.nSlots), xs, sh);
    }
};

// The PCG update fused with the V-cycle's first step on the finest level (multigrid): per brick,
// x += alpha s and r -= alpha q (PCGSolver.cpp:178-181), the brick's max|r| to pm[slot], then the
// pre-smoother on b = r, which reads only the brick's own cells (it starts from x = 0). The finest
// level's r array is the PCG's q buffer: this brick's q is read before the smoother overwrites it.
// Per cell the same arithmetic as k_pcg_update followed by k_mg_smooth_brick<0, 0>.
struct k_pcg_update_pre {
    static constexpr int kThreads = 256;
    const float* __restrict__ s;
    float* __restrict__ x;
    LevelDev L;  // the finest level: L.b = r, L.r = q
    float omega;
    int sweeps;
    const PcgState* __restrict__ st;
    double* __restrict__ pm;
    __device__ void operator()(Blk blk) const {
        __shared__ float  xs[10 * 10 * 10];
        __shared__ double sh[4];
        if(st && st->done) return;
        const float alpha = (float)st->alpha;
        const int   t     = threadIdx.x;
        for(uint32_t v = blk.id; v < L.nSlots; v += blk.n) {
            const int slot = (int)xcd_slot(v, L.nSlots);
            double    m    = 0;
            for(int h = 0; h < 2; ++h) {
                const size_t o  = (size_t)slot * 512 + t + 256 * h;
                const float  xv = x[o] + alpha * s[o];
                const float  rv = L.b[o] + (-alpha) * L.r[o];
                x[o]            = xv;
                L.b[o]          = rv;
                m               = fmax(m, (double)fabsf(rv));
            }
            m = block_max<256>(m, sh);
            if(t == 0) pm[slot] = m;
            __syncthreads();  // the brick's r, written by other threads of the workgroup, is the smoother's b
            mg_smooth_one<0, 0>(L, L, omega, sweeps, nullptr, slot, xs, sh);
        }
    }
};

// r = b - A x on level F with the pre-smoothed x (held in F.r), summed over each 2x2x2 group into
// the coarse right-hand side (P^T r). A group's eight cells lie in one fine brick, so the block of a
// fine brick writes the 64 coarse cells it covers; r itself is never stored. Coarse cells without an
// active child brick keep the zeros they get when the hierarchy is built. 256 threads, two cells each.
struct k_mg_residual_restrict {
    static constexpr int kThreads = 256;
    LevelDev F;
    LevelDev C;
    const PcgState* __restrict__ st;
    __device__ void operator()(Blk blk) const {
        __shared__ float rs[512];
        if(st && st->done) return;
        const int t = threadIdx.x;
        for(uint32_t v = blk.id; v < F.nSlots; v += blk.n) {
            const int s = (int)xcd_slot(v, F.nSlots);
            const int b = F.slotBrick[s];
            const int bx = b % F.nbx, by = (b / F.nbx) % F.nby, bz = b / (F.nbx * F.nby);
            for(int h = 0; h < 2; ++h) {
                const int    l = t + 256 * h;
                const size_t o = (size_t)s * 512 + l;
                rs[l]          = F.b[o] - (F.diag[o] * F.r[o] + level_offdiag_sum(F, s, l, F.r));
            }
            __syncthreads();
            if(t < 64) {
                const int32_t cs = F.ptab[8 * s + 6];  // the parent brick
                const int qx = t & 3, qy = (t >> 2) & 3, qz = t >> 4;  // coarse cell inside the brick's 4x4x4
                const int I = bx * 4 + qx, J = by * 4 + qy, K = bz * 4 + qz;
                float     sum = 0.f;
                for(int ch = 0; ch < 8; ++ch) {
                    const int fi = 2 * I + (ch & 1), fj = 2 * J + ((ch >> 1) & 1), fk = 2 * K + ((ch >> 2) & 1);
                    if(fi >= F.nx || fj >= F.ny || fk >= F.nz) continue;
                    sum += rs[((fk & 7) * 8 + (fj & 7)) * 8 + (fi & 7)];
                }
                if(cs >= 0 && I < C.nx && J < C.ny && K < C.nz) C.b[(size_t)cs * 512 + ((K & 7) * 8 + (J & 7)) * 8 + (I & 7)] = sum;
            }
            __syncthreads();  // rs is reused by the block's next brick
        }
    }
};

// coarsest level: x = 0, then `sweeps` red-black sweeps followed by `sweeps` black-red sweeps
// (symmetric), all inside one block
struct k_mg_coarsest {
    static constexpr int kThreads = 1024;
    LevelDev L;
    int sweeps;
    const PcgState* __restrict__ st;
    __device__ void operator()(Blk blk) const {
        if(st && st->done) return;
        const uint32_t n = L.nSlots * 512;
        for(uint32_t t = threadIdx.x; t < n; t += blockDim.x) L.x[t] = 0.f;
        __syncthreads();
        for(int it = 0; it < 2 * sweeps; ++it) {
            for(int half = 0; half < 2; ++half) {
                const int color = (it < sweeps) ? half : 1 - half;
                for(uint32_t t = threadIdx.x; t < n; t += blockDim.x) {
                    const int s = (int)(t >> 9), l = (int)(t & 511);
                    if(cell_color(L, s, l) != color) continue;
                    L.x[t] = (L.b[t] - level_offdiag_sum(L, s, l, L.x)) * L.invdiag[t];
                }
                __syncthreads();
            }
        }
    }
};

// partial z.r for the PCG (after the V-cycle wrote z)
//==================================================================================================
// The reference's preconditioner: natural-order incomplete Cholesky IC(0) with the 0.25 pivot
// safeguard (PCGSolver.cpp:398-466, minDiagonalRatio PCGSolver.h:88-89) and its two triangular
// solves (PCGSolver.cpp:243-269), level-scheduled over the hyperplanes i + j + k = h. On the 7-point
// stencil every lower neighbour of a cell (-x, -y, -z) lies on plane h - 1 and every upper neighbour
// on h + 1, so one launch per plane, in plane order, reproduces the sequential recurrences; each
// cell applies its neighbours' terms in the reference's order (by ascending natural index: -z, -y,
// -x going forward; +x, +y, +z coming back), with the same FP32 roundings. 3n - 2 dependent launches
// per sweep: latency-bound, used to compare iterates with the reference's preconditioner, not for
// speed. Cells outside the system (solid, air, the outer layer) have a zero diagonal and a zero
// factor, so they contribute nothing and get z = 0.
//==================================================================================================
struct Ic0Dev {
    const int32_t* brickSlot;
    const float *  diag, *cx, *cy, *cz;  // A: diagonal and the couplings to +x, +y, +z
    float *        inv, *lx, *ly, *lz;   // 1/sqrt(pivot); L entries below the diagonal (column-scaled)
};

__device__ __forceinline__ int32_t sys_cell(const int32_t* __restrict__ brickSlot, const GridGeom& g, int i, int j, int k) {
    if(i < 0 || j < 0 || k < 0) return -1;
    const int32_t s = brickSlot[((k >> 3) * g.nby + (j >> 3)) * g.nbx + (i >> 3)];
    return s < 0 ? -1 : s * 512 + ((k & 7) * 8 + (j & 7)) * 8 + (i & 7);
}

// the cell (i, j, k = h - i - j) of plane h owned by this thread, or -1
__device__ __forceinline__ int32_t ic0_plane_cell(Blk blk, const Ic0Dev& D, const GridGeom& g, int h, int& i, int& j, int& k) {
    const uint32_t t = blk.id * blockDim.x + threadIdx.x;
    if(t >= (uint32_t)g.nx * (uint32_t)g.ny) return -1;
    i = (int)(t % (uint32_t)g.nx);
    j = (int)(t / (uint32_t)g.nx);
    k = h - i - j;
    if(k < 0 || k >= g.nz) return -1;
    return sys_cell(D.brickSlot, g, i, j, k);
}

// pivot d = a_oo - L(o,-z)^2 - L(o,-y)^2 - L(o,-x)^2 (the column updates of the lower neighbours,
// in their processing order), inv = 1/sqrt(d), or 1/sqrt(a_oo) when d < 0.25 a_oo; then the column
// below the diagonal is scaled: L(+x,o) = a(o,+x) * inv, and likewise +y, +z
struct k_ic0_factor_plane {
    static constexpr int kThreads = kBlock;
    Ic0Dev D;
    GridGeom g;
    int h;
    __device__ void operator()(Blk blk) const {
    #pragma clang fp contract(off)
        int           i, j, k;
        const int32_t o = ic0_plane_cell(blk, D, g, h, i, j, k);
        if(o < 0) return;
        const float a = D.diag[o];
        float       inv;
        if(a < kFltMin) {
            inv = a;  // row skipped by the factorisation (PCGSolver.cpp:419-421): its couplings are zero
        } else {
            const int32_t mz = sys_cell(D.brickSlot, g, i, j, k - 1), my = sys_cell(D.brickSlot, g, i, j - 1, k),
                          mx = sys_cell(D.brickSlot, g, i - 1, j, k);
            float d = a;
            if(mz >= 0) {
                const float l = D.lz[mz];
                d -= l * l;
            }
            if(my >= 0) {
                const float l = D.ly[my];
                d -= l * l;
            }
            if(mx >= 0) {
                const float l = D.lx[mx];
                d -= l * l;
            }
            inv = d < 0.25f * a ? 1.0f / sqrtf(a) : 1.0f / sqrtf(d);
        }
        D.inv[o] = inv;
        const bool skipped = a < kFltMin;  // an unscaled zero column
        D.lx[o]  = skipped ? D.cx[o] : D.cx[o] * inv;
        D.ly[o]  = skipped ? D.cy[o] : D.cy[o] * inv;
        D.lz[o]  = skipped ? D.cz[o] : D.cz[o] * inv;
    }
};

// forward solve, plane h: y_o = (((r_o - L(o,-z) y_-z) - L(o,-y) y_-y) - L(o,-x) y_-x) * inv_o
struct k_ic0_lower_plane {
    static constexpr int kThreads = kBlock;
    Ic0Dev D;
    GridGeom g;
    int h;
    const float* __restrict__ r;
    float* __restrict__ y;
    const PcgState* __restrict__ st;
    __device__ void operator()(Blk blk) const {
    #pragma clang fp contract(off)
        if(st && st->done) return;
        int           i, j, k;
        const int32_t o = ic0_plane_cell(blk, D, g, h, i, j, k);
        if(o < 0) return;
        const int32_t mz = sys_cell(D.brickSlot, g, i, j, k - 1), my = sys_cell(D.brickSlot, g, i, j - 1, k),
                      mx = sys_cell(D.brickSlot, g, i - 1, j, k);
        float v = r[o];
        if(mz >= 0) v -= D.lz[mz] * y[mz];
        if(my >= 0) v -= D.ly[my] * y[my];
        if(mx >= 0) v -= D.lx[mx] * y[mx];
        y[o] = v * D.inv[o];
    }
};

// backward solve in place, plane h (descending): z_o = (((y_o - L(+x,o) z_+x) - L(+y,o) z_+y) -
// L(+z,o) z_+z) * inv_o
struct k_ic0_upper_plane {
    static constexpr int kThreads = kBlock;
    Ic0Dev D;
    GridGeom g;
    int h;
    float* __restrict__ z;
    const PcgState* __restrict__ st;
    __device__ void operator()(Blk blk) const {
    #pragma clang fp contract(off)
        if(st && st->done) return;
        int           i, j, k;
        const int32_t o = ic0_plane_cell(blk, D, g, h, i, j, k);
        if(o < 0) return;
        const int32_t px = i + 1 < g.nx ? sys_cell(D.brickSlot, g, i + 1, j, k) : -1;
        const int32_t py = j + 1 < g.ny ? sys_cell(D.brickSlot, g, i, j + 1, k) : -1;
        const int32_t pz = k + 1 < g.nz ? sys_cell(D.brickSlot, g, i, j, k + 1) : -1;
        float         v  = z[o];
        if(px >= 0) v -= D.lx[o] * z[px];
        if(py >= 0) v -= D.ly[o] * z[py];
        if(pz >= 0) v -= D.lz[o] * z[pz];
        z[o] = v * D.inv[o];
    }
};

struct k_pcg_dot_zr {
    static constexpr int kThreads = 512;
    const float* __restrict__ z;
    const float* __restrict__ r;
    uint32_t nSlots;
    const PcgState* __restrict__ st;
    double* __restrict__ pa;
    __device__ void operator()(Blk blk) const {
        __shared__ double sh[8];
        if(st && st->done) return;
        double d = 0;
        for(uint32_t slot = blk.id; slot < nSlots; slot += blk.n) {
            const size_t o = (size_t)slot * 512 + threadIdx.x;
            d += (double)z[o] * (double)r[o];
        }
        d = block_sum<512>(d, sh);
        if(threadIdx.x == 0) pa[blk.id] = d;
    }
};

}  // namespace bnn

namespace bnn {

// coarsest level when it is a single brick (no neighbour bricks): x = 0, then `sweeps` red-black
// sweeps and `sweeps` black-red sweeps with x held in LDS. One wavefront: each lane owns four red and
// four black cells (x-adjacent pairs on z = z0, z0+2, z0+4, z0+6), so a half-sweep ends with a
// single-wave barrier. The LDS array has a zero halo, so couplings to cells outside the brick (zero)
// are added unconditionally and change nothing.
struct k_mg_coarsest1 {
    static constexpr int kThreads = 64;
    LevelDev L;
    int sweeps;
    const PcgState* __restrict__ st;
    __device__ void operator()(Blk blk) const {
        __shared__ float xs[10 * 10 * 10];
        if(st && st->done) return;
        const int t = threadIdx.x, px = 2 * (t & 3), ly = (t >> 2) & 7, lz0 = t >> 5;
        const int shift = (ly + lz0) & 1;  // colour of the even-x cell of every pair of this lane
        for(int c = t; c < 1000; c += 64) xs[c] = 0.f;
        SmoothCell R[4], K[4];  // red (colour 0) and black cells
    #pragma unroll
        for(int q = 0; q < 4; ++q) {
            const int lz = lz0 + 2 * q;
    #pragma unroll
            for(int e = 0; e < 2; ++e) {
                const int  lx = px + e, l = (lz * 8 + ly) * 8 + lx;
                SmoothCell C;
                C.c   = ((lz + 1) * 10 + (ly + 1)) * 10 + (lx + 1);
                C.czm = lz > 0 ? L.cz[l - 64] : 0.f;
                C.cym = ly > 0 ? L.cy[l - 8] : 0.f;
                C.cxm = lx > 0 ? L.cx[l - 1] : 0.f;
                C.cxp = lx < 7 ? L.cx[l] : 0.f;
                C.cyp = ly < 7 ? L.cy[l] : 0.f;
                C.czp = lz < 7 ? L.cz[l] : 0.f;
                C.bv  = L.b[l];
                C.inv = L.invdiag[l];
                if((e ^ shift) == 0) R[q] = C;
                else K[q] = C;
            }
        }
        __syncthreads();
        for(int it = 0; it < 2 * sweeps; ++it) {
            if(it < sweeps) {  // red-black (uniform branch)
                smooth_cells_update4(R, xs);
                __syncthreads();
                smooth_cells_update4(K, xs);
            } else {  // black-red
                smooth_cells_update4(K, xs);
                __syncthreads();
                smooth_cells_update4(R, xs);
            }
            __syncthreads();
        }
    #pragma unroll
        for(int q = 0; q < 4; ++q) {
            const int lz = lz0 + 2 * q;
    #pragma unroll
            for(int e = 0; e < 2; ++e) {
                const int lx = px + e;
                L.x[(lz * 8 + ly) * 8 + lx] = xs[((lz + 1) * 10 + (ly + 1)) * 10 + (lx + 1)];
            }
        }
    }
};

}  // namespace bnn
