// flip_launch.hpp -- how the substep's kernels are launched, for one scene or for many at once.
//
// Every kernel is a functor: its fields are the launch arguments and operator()(Blk) is the work of
// one workgroup, which sees only its index among its scene's workgroups (Blk::id) and their count
// (Blk::n), never blockIdx or gridDim. The same functor then runs
//   - alone:   k_one<K>, one workgroup per scene block (the functor is the kernel argument), or
//   - batched: k_many<K>, one launch over a flat work list of many scenes. Scene s owns the
//              items [first[s], first[s] + count[s]) and its functor is ks[s]; a workgroup runs
//              several items of one XCD residue in turn. first[s] is a multiple of 8, so an item's
//              index mod 8 (the XCD it runs on) is the same in both forms.
// A scene's blocks therefore do exactly the same arithmetic in either form, and its results are the
// same bits alone and in any batch.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

namespace bnn {

struct Blk {
    uint32_t id, n;  // this workgroup's index among its scene's workgroups, and their count
};

// Single scene. The functor is the kernel's only argument, so it lies at offset 0 of the
// kernel-argument segment, and it is called in place there (constant memory, read with scalar loads
// even where it indexes one of its arrays at run time, as k_move does with the boundary shapes). Used
// as the by-value parameter instead, such a functor is first copied to per-lane scratch. The call is
// inlined by force here and in k_many: the larger functors exceed the inliner's threshold, and a real
// call doubles their registers.
template<class K> __global__ void __launch_bounds__(K::kThreads) k_one(K k) {
    (void)k;
    const K* kp = (const K*)__builtin_amdgcn_kernarg_segment_ptr();  // an address-space conversion
    [[clang::always_inline]] (*kp)(Blk{blockIdx.x, gridDim.x});
}

// Scenes per batched launch at most (the scene search below is two rounds of 64 lanes)
constexpr uint32_t kMaxLaunchScenes = 4096;

// Workgroups per CU that the batched form of K is compiled to leave room for: K::kManyBlocks where the
// functor names it (a register bound for functors whose batched form, a loop over items, otherwise
// takes many more registers than the single-scene form), else 1 (no bound).
template<class K, class = void> struct many_blocks {
    static constexpr int value = 1;
};
template<class K> struct many_blocks<K, std::void_t<decltype(K::kManyBlocks)>> {
    static constexpr int value = K::kManyBlocks;
};

// Batched launch. Items are the scenes' blocks laid end to end (scene s: items first[s] ..
// first[s] + count[s] - 1, first[s] a multiple of 8); workgroup w runs the items 8 r + (w mod 8) for
// rows r in [(w / 8) per, (w / 8 + 1) per) one after the other, so an item keeps its XCD residue and
// one scene search serves `per` items.
template<class K>
__global__ void __launch_bounds__(K::kThreads, many_blocks<K>::value) k_many(const K* __restrict__ ks, const uint32_t* __restrict__ first,
                                                      const uint32_t* __restrict__ count, uint32_t nScenes, uint32_t rows, uint32_t per) {
    const uint32_t x = blockIdx.x & 7u, r0 = (blockIdx.x >> 3) * per;
    const uint32_t r1 = min(r0 + per, rows);
    if(r0 >= r1) return;
    // the scene of the first item: the last s with first[s] <= i0, found by the first wavefront in
    // two rounds (64 buckets of ceil(nScenes / 64) scenes, then the scenes of one bucket)
    __shared__ uint32_t s_scene;
    const uint32_t      i0 = 8u * r0 + x;
    if(threadIdx.x < 64) {
        const uint32_t lane = threadIdx.x, stride = (nScenes + 63u) / 64u;
        const uint32_t b0   = lane * stride;
        const uint64_t m1   = __ballot(b0 < nScenes && first[b0] <= i0);
        const uint32_t base = ((uint32_t)__popcll(m1) - 1u) * stride;
        if(stride == 1u) {  // at most 64 scenes: the first round found it
            if(lane == 0) s_scene = base;
        } else {
            const uint32_t sc = base + lane;
            const uint64_t m2 = __ballot(lane < stride && sc < nScenes && first[sc] <= i0);
            if(lane == 0) s_scene = base + (uint32_t)__popcll(m2) - 1u;
        }
    }
    __syncthreads();
    uint32_t s = __builtin_amdgcn_readfirstlane(s_scene);
    for(uint32_t r = r0; r < r1; ++r) {
        const uint32_t i = 8u * r + x;
        while(s + 1u < nScenes && first[s + 1u] <= i) ++s;
        s                    = __builtin_amdgcn_readfirstlane(s);
        const uint32_t local = i - first[s], n = count[s];
        // the functor array is not written while the kernel runs: read through the constant address
        // space, its fields are scalar loads into scalar registers, as in k_one
        const K* kp = (const K*)((const __attribute__((address_space(4))) K*)ks + s);
        if(local < n) [[clang::always_inline]] (*kp)(Blk{local, n});  // else: padding up to the next scene's multiple of 8
        if(r + 1u < r1) __syncthreads();      // the next item may reuse this item's LDS
    }
}

// dst[i] = v for i < n words
struct k_fill32 {
    static constexpr int kThreads = 256;
    uint32_t* __restrict__ dst;
    size_t    n;
    uint32_t  v;
    __device__ void operator()(Blk B) const {
        for(size_t i = (size_t)B.id * kThreads + threadIdx.x; i < n; i += (size_t)B.n * kThreads) dst[i] = v;
    }
};

// dst[i] = src[i] for i < n words (non-overlapping); 16-byte accesses when both ends allow
struct k_copy32 {
    static constexpr int kThreads = 256;
    uint32_t* __restrict__       dst;
    const uint32_t* __restrict__ src;
    size_t                       n;
    __device__ void operator()(Blk B) const {
        const size_t stride = (size_t)B.n * kThreads, t0 = (size_t)B.id * kThreads + threadIdx.x;
        if((((uintptr_t)dst | (uintptr_t)src) & 15u) == 0) {
            const size_t n4 = n / 4;
            uint4*       d4 = reinterpret_cast<uint4*>(dst);
            const uint4* s4 = reinterpret_cast<const uint4*>(src);
            for(size_t i = t0; i < n4; i += stride) d4[i] = s4[i];
            for(size_t i = n4 * 4 + t0; i < n; i += stride) dst[i] = src[i];
        } else {
            for(size_t i = t0; i < n; i += stride) dst[i] = src[i];
        }
    }
};

// workgroups for a fill or copy of n words: enough to cover n in one pass up to a cap, after which
// the workgroups stride
static inline uint32_t fill_blocks(size_t n) {
    const size_t b = (n + 1023) / 1024;
    return (uint32_t)(b < 1 ? 1 : (b > 4096 ? 4096 : b));
}

// readbacks of a batched flush: entry e copies words[e] 32-bit words from src[e] to the staging
// buffer at out + offset[e] (one workgroup of 64 per entry, at most 64 words each)
__global__ void __launch_bounds__(64) k_gather_words(const uint32_t* const* __restrict__ src, const uint32_t* __restrict__ words,
                                                     const uint32_t* __restrict__ offset, uint32_t* __restrict__ out) {
    const uint32_t e = blockIdx.x;
    if(threadIdx.x < words[e]) out[offset[e] + threadIdx.x] = src[e][threadIdx.x];
}

}  // namespace bnn
