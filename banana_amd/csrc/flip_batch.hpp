// flip_batch.hpp -- host machinery that runs many scenes' substeps as one stream of fused launches.
//
// A scene's host code is a coroutine (Co<T>) that suspends wherever it must read a device value
// (brick counts, multigrid level sizes, PCG stop flags, residual norms). Alone, a sync point waits
// on the scene's stream and the coroutine never suspends. In a batch, every scene records its kernel
// launches and readbacks into a Recorder and suspends at its sync point; when all scenes are
// suspended the Recorder flushes: it merges the scenes' launch sequences kernel by kernel into
// launches over a flat work list (k_many in flip_launch.hpp), copies every recorded readback with
// one gather kernel and one device-to-host copy, waits once, and resumes every scene.
//
// Flushes whose launches repeat exactly (the same kernels over the same scenes with the same
// arguments, as in the PCG iterations of one solve) are captured as a hipGraph the second time they
// are seen and replayed from then on.
#pragma once

#include <hip/hip_runtime.h>

#include <chrono>
#include <coroutine>
#include <cstring>
#include <exception>
#include <optional>
#include <string>
#include <unordered_map>
#include <utility>
#include <vector>

#include "flip_launch.hpp"

namespace bnn {

//--------------------------------------------------------------------------------------------------
// Co<T>: a lazily started coroutine task; co_await runs it and yields its value, exceptions included
//--------------------------------------------------------------------------------------------------
template<class T> struct Co;

struct CoPromiseBase {
    std::coroutine_handle<> cont;
    std::exception_ptr      ex;
    std::suspend_always     initial_suspend() noexcept { return {}; }
    struct Final {
        bool await_ready() noexcept { return false; }
        template<class P> std::coroutine_handle<> await_suspend(std::coroutine_handle<P> h) noexcept {
            auto c = h.promise().cont;
            return c ? c : std::noop_coroutine();
        }
        void await_resume() noexcept {}
    };
    Final final_suspend() noexcept { return {}; }
    void  unhandled_exception() { ex = std::current_exception(); }
};

template<class T> struct CoPromise : CoPromiseBase {
    std::optional<T> value;
    Co<T>            get_return_object();
    void             return_value(T v) { value.emplace(std::move(v)); }
};
template<> struct CoPromise<void> : CoPromiseBase {
    Co<void> get_return_object();
    void     return_void() {}
};

template<class T = void> struct [[nodiscard]] Co {
    using promise_type = CoPromise<T>;
    std::coroutine_handle<promise_type> h;
    explicit Co(std::coroutine_handle<promise_type> hh) : h(hh) {}
    Co(Co&& o) noexcept : h(std::exchange(o.h, {})) {}
    Co(const Co&)            = delete;
    Co& operator=(const Co&) = delete;
    ~Co() {
        if(h) h.destroy();
    }
    bool                    await_ready() const noexcept { return false; }
    std::coroutine_handle<> await_suspend(std::coroutine_handle<> c) noexcept {
        h.promise().cont = c;
        return h;
    }
    T await_resume() { return result(); }
    bool done() const { return h.done(); }
    T    result() {
        if(h.promise().ex) std::rethrow_exception(h.promise().ex);
        if constexpr(!std::is_void_v<T>) return std::move(*h.promise().value);
    }
};
template<class T> inline Co<T> CoPromise<T>::get_return_object() { return Co<T>(std::coroutine_handle<CoPromise<T>>::from_promise(*this)); }
inline Co<void>                CoPromise<void>::get_return_object() { return Co<void>(std::coroutine_handle<CoPromise<void>>::from_promise(*this)); }

// a sync point: ready at once (the caller has already waited) or suspends, leaving its handle in
// *slot for the batch driver to resume after the flush
struct SyncAwait {
    std::coroutine_handle<>* slot;
    bool                     await_ready() const noexcept { return slot == nullptr; }
    void                     await_suspend(std::coroutine_handle<> h) noexcept { *slot = h; }
    void                     await_resume() const noexcept {}
};

// runs a task that cannot suspend (a single scene) to completion
template<class T> T run_now(Co<T>&& c) {
    c.h.resume();
    if(!c.h.done()) throw std::logic_error("a single-scene task suspended");
    return c.result();
}

//--------------------------------------------------------------------------------------------------
// Recorder: launches and readbacks of many scenes ("lanes"), merged and issued at flush time
//--------------------------------------------------------------------------------------------------
class Recorder {
public:
    using LaunchFn = void (*)(const void* ks, const uint32_t* first, const uint32_t* count, uint32_t nScenes, uint32_t items,
                              uint32_t per, hipStream_t s);
    // one scene's launch of the same kernel: k_one with the functor (host bytes) as its argument
    using LaunchOneFn = void (*)(const void* k, uint32_t blocks, hipStream_t s);

    explicit Recorder(hipStream_t s) : stream(s) {}
    ~Recorder() {
        clearGraphs();
        if(ev0) (void)hipEventDestroy(ev0);
        if(ev1) (void)hipEventDestroy(ev1);
        if(dblob) (void)hipFree(dblob);
        if(hblob) (void)hipHostFree(hblob);
    }
    Recorder(const Recorder&)            = delete;
    Recorder& operator=(const Recorder&) = delete;

    // lanes are slots for scenes; a lane is taken for the length of one task and then given back
    uint32_t takeLane() {
        if(!freeLanes.empty()) {
            const uint32_t l = freeLanes.back();
            freeLanes.pop_back();
            return l;
        }
        lanes.emplace_back();
        return (uint32_t)lanes.size() - 1;
    }
    void giveLane(uint32_t l) {
        Lane& L = lanes[l];
        if(!L.ops.empty() || !L.rbs.empty()) throw std::logic_error("a lane was given back with work recorded");
        freeLanes.push_back(l);
    }

    template<class K> void add(uint32_t lane, const K& k, uint32_t nblk) {
        static_assert(std::is_trivially_copyable_v<K>, "kernel functors are copied byte for byte");
        if(nblk == 0) return;
        Lane& L = lanes[lane];
        if(!L.rbs.empty()) throw std::logic_error("a launch was recorded after a readback in one flush");
        const size_t off = (L.bytes.size() + alignof(K) - 1) / alignof(K) * alignof(K);
        L.bytes.resize(off + sizeof(K));
        std::memcpy(L.bytes.data() + off, &k, sizeof(K));
        L.ops.push_back(Op{&launchMany<K>, &launchOne<K>, (uint32_t)sizeof(K), (uint32_t)alignof(K), (uint32_t)K::kThreads, nblk, off});
    }
    void readback(uint32_t lane, void* host, const void* dev, size_t bytes) {
        if(bytes % 4 || bytes > 256) throw std::logic_error("readbacks are whole words, at most 64");
        lanes[lane].rbs.push_back(Rb{host, dev, (uint32_t)(bytes / 4)});
    }
    // the identity of kernel K's batched launch (to name it in timedFns)
    template<class K> static LaunchFn fnOf() { return &launchMany<K>; }

    // drops everything recorded (after a failed flush)
    void discard() {
        for(Lane& L : lanes) {
            L.ops.clear();
            L.bytes.clear();
            L.rbs.clear();
        }
    }

    bool pending() const {
        for(const Lane& L : lanes)
            if(!L.ops.empty() || !L.rbs.empty()) return true;
        return false;
    }

    // issues everything recorded, waits for it and delivers the readbacks
    void flushAndSync() {
        if(!pending()) return;
        ++flushes;
        buildBlob();
        bool timed = false;
        for(const Launch& L : launches)
            for(LaunchFn f : timedFns) timed |= L.fn == f;
        bool viaGraph = false;
        if(useGraphs && !timed && launches.size() >= kMinGraphLaunches) viaGraph = graphLaunch();
        if(!viaGraph) {
            upload(dblob);
            issue(dblob, timed);
        }
        if(nRb) {
            hipLaunchKernelGGL(k_gather_words, dim3(nRb), dim3(64), 0, stream, reinterpret_cast<const uint32_t* const*>(dblob + rbSrcOff),
                               reinterpret_cast<const uint32_t*>(dblob + rbWordsOff), reinterpret_cast<const uint32_t*>(dblob + rbOffOff),
                               reinterpret_cast<uint32_t*>(dblob + rbOutOff));
            check(hipGetLastError());
            check(hipMemcpyAsync(hblob + rbOutOff, dblob + rbOutOff, (size_t)rbWords * 4, hipMemcpyDeviceToHost, stream));
        }
        check(hipStreamSynchronize(stream));
        if(timed) {
            float ms = 0;
            check(hipEventElapsedTime(&ms, ev0, ev1));
            timedMs += ms;
        }
        for(Lane& L : lanes) {
            for(const Rb& r : L.rbs) {
                std::memcpy(r.host, hblob + rbOutOff + (size_t)r.outWord * 4, (size_t)r.words * 4);
            }
            L.ops.clear();
            L.bytes.clear();
            L.rbs.clear();
        }
    }

    void clearGraphs() {
        for(auto& kv : graphs) {
            if(kv.second.exec) (void)hipGraphExecDestroy(kv.second.exec);
            if(kv.second.dev) (void)hipFree(kv.second.dev);
        }
        graphs.clear();
    }

    bool useGraphs = true;
    // items per workgroup: the launch keeps at least minGroups workgroups, each running at most
    // maxPer items (1: one workgroup per item)
    uint32_t maxPer = 16, minGroups = 16384;
    uint32_t perFor(uint32_t items) const {
        const uint32_t rows = (items + 7u) / 8u, q = std::max<uint32_t>(minGroups / 8u, 1u);
        return std::max<uint32_t>(1u, std::min<uint32_t>(maxPer, rows / q));
    }
    // kernels whose launches are timed with events (the first to the last such launch of a flush,
    // summed into timedMs); flushes with a timed launch are issued directly, never from a graph
    std::vector<LaunchFn> timedFns;
    double                timedMs = 0;
    // counters since creation: flushes, launches issued, graph replays and captures, scene launches
    // merged into them, and the number of lanes with work summed over flushes
    uint64_t flushes = 0, kernelLaunches = 0, graphReplays = 0, graphCaptures = 0, mergedOps = 0, laneFlushes = 0;

private:
    static constexpr size_t kMinGraphLaunches = 8;
    static constexpr size_t kMaxGraphs        = 64;
    static constexpr uint64_t kMaxItems       = 1ull << 31;  // work-items per launch

    struct Op {
        LaunchFn    fn;
        LaunchOneFn one;
        uint32_t    size, align, threads, nblk;
        size_t      off;
    };
    struct Rb {
        void*       host;
        const void* dev;
        uint32_t    words, outWord = 0;
    };
    struct Lane {
        std::vector<Op>      ops;
        std::vector<uint8_t> bytes;
        std::vector<Rb>      rbs;
    };
    struct Launch {
        LaunchFn    fn;
        LaunchOneFn one;  // used when the launch holds one scene (its functor at bodiesOff)
        size_t      bodiesOff, firstOff, countOff;
        uint32_t    nScenes, blocks;
    };
    struct Graph {
        std::vector<uint8_t> blob;
        std::vector<Launch>  launches;
        uint8_t*             dev  = nullptr;
        hipGraphExec_t       exec = nullptr;
        uint64_t             lastUse = 0;
    };

    // items = the last scene's end; per = items per workgroup and XCD residue (rows of 8 items)
    template<class K> static void launchOne(const void* k, uint32_t blocks, hipStream_t s) {
        hipLaunchKernelGGL(k_one<K>, dim3(blocks), dim3(K::kThreads), 0, s, *static_cast<const K*>(k));
    }
    template<class K>
    static void launchMany(const void* ks, const uint32_t* first, const uint32_t* count, uint32_t nScenes, uint32_t items, uint32_t per,
                           hipStream_t s) {
        const uint32_t rows = (items + 7u) / 8u, groups = 8u * ((rows + per - 1u) / per);
        hipLaunchKernelGGL(k_many<K>, dim3(groups), dim3(K::kThreads), 0, s, static_cast<const K*>(ks), first, count, nScenes, rows, per);
    }

    static void check(hipError_t e) {
        if(e != hipSuccess) throw std::runtime_error(std::string("batched launch: ") + hipGetErrorString(e));
    }

    size_t put(size_t bytes, size_t align) {
        const size_t off = (used + align - 1) / align * align;
        used             = off + bytes;
        if(used > staging.size()) staging.resize(std::max(used, staging.size() * 2));
        return off;
    }

    // merges the lanes' launch sequences: each round takes the kernel most common among the lanes'
    // next launches and issues it for every lane whose next launch it is (a lane's own order is kept)
    void buildBlob() {
        used = 0;
        launches.clear();
        const uint32_t        n = (uint32_t)lanes.size();
        std::vector<uint32_t> head(n, 0);
        std::vector<uint32_t> members;
        std::unordered_map<LaunchFn, uint32_t> freq;
        for(;;) {
            freq.clear();
            LaunchFn best = nullptr;
            uint32_t bestN = 0;
            for(uint32_t l = 0; l < n; ++l) {
                if(head[l] >= lanes[l].ops.size()) continue;
                const uint32_t c = ++freq[lanes[l].ops[head[l]].fn];
                if(c > bestN) {
                    bestN = c;
                    best  = lanes[l].ops[head[l]].fn;
                }
            }
            if(!best) break;
            members.clear();
            for(uint32_t l = 0; l < n; ++l)
                if(head[l] < lanes[l].ops.size() && lanes[l].ops[head[l]].fn == best) members.push_back(l);
            mergedOps += members.size();
            // split so that no launch exceeds kMaxItems work-items
            size_t s0 = 0;
            while(s0 < members.size()) {
                const Op& o0 = lanes[members[s0]].ops[head[members[s0]]];
                size_t    s1 = s0;
                uint64_t  blk = 0;
                while(s1 < members.size()) {
                    const Op&      o   = lanes[members[s1]].ops[head[members[s1]]];
                    const uint64_t pad = (o.nblk + 7u) & ~7u;
                    if(s1 > s0 && ((blk + pad) * o.threads > kMaxItems || s1 - s0 >= kMaxLaunchScenes)) break;
                    blk += pad;
                    ++s1;
                }
                const uint32_t m = (uint32_t)(s1 - s0);
                Launch         L{best, o0.one, put((size_t)o0.size * m, std::max<size_t>(o0.align, 16)), 0, 0, m, 0};
                for(uint32_t t = 0; t < m; ++t) {
                    const Lane& ln = lanes[members[s0 + t]];
                    const Op&   o  = ln.ops[head[members[s0 + t]]];
                    std::memcpy(staging.data() + L.bodiesOff + (size_t)t * o.size, ln.bytes.data() + o.off, o.size);
                }
                L.firstOff = put((size_t)m * 4, 16);
                L.countOff = put((size_t)m * 4, 16);
                uint32_t acc = 0;
                for(uint32_t t = 0; t < m; ++t) {
                    const Op& o = lanes[members[s0 + t]].ops[head[members[s0 + t]]];
                    std::memcpy(staging.data() + L.firstOff + (size_t)t * 4, &acc, 4);
                    std::memcpy(staging.data() + L.countOff + (size_t)t * 4, &o.nblk, 4);
                    acc += t + 1 < m ? ((o.nblk + 7u) & ~7u) : o.nblk;
                }
                L.blocks = acc;
                launches.push_back(L);
                s0 = s1;
            }
            for(uint32_t l : members) ++head[l];
        }
        // readback table: source pointers, word counts, output offsets, then the output words
        nRb     = 0;
        rbWords = 0;
        for(Lane& L : lanes)
            for(Rb& r : L.rbs) {
                r.outWord = rbWords;
                rbWords += r.words;
                ++nRb;
            }
        rbSrcOff = rbWordsOff = rbOffOff = rbOutOff = 0;
        if(nRb) {
            rbSrcOff   = put((size_t)nRb * 8, 16);
            rbWordsOff = put((size_t)nRb * 4, 16);
            rbOffOff   = put((size_t)nRb * 4, 16);
            uint32_t e = 0;
            for(Lane& L : lanes)
                for(Rb& r : L.rbs) {
                    const uint64_t src = (uint64_t)(uintptr_t)r.dev;
                    std::memcpy(staging.data() + rbSrcOff + (size_t)e * 8, &src, 8);
                    std::memcpy(staging.data() + rbWordsOff + (size_t)e * 4, &r.words, 4);
                    std::memcpy(staging.data() + rbOffOff + (size_t)e * 4, &r.outWord, 4);
                    ++e;
                }
            rbOutOff = put((size_t)rbWords * 4, 16);
        }
        kernelLaunches += launches.size();
        for(const Lane& L : lanes) laneFlushes += (!L.ops.empty() || !L.rbs.empty()) ? 1 : 0;
        // pinned and device copies large enough for the blob (both idle: every flush ends in a wait)
        if(used > cap) {
            if(dblob) check(hipFree(dblob));
            if(hblob) check(hipHostFree(hblob));
            dblob = nullptr;
            hblob = nullptr;
            cap   = used + used / 2 + 4096;
            check(hipMalloc((void**)&dblob, cap));
            check(hipHostMalloc((void**)&hblob, cap, hipHostMallocDefault));
        }
        std::memcpy(hblob, staging.data(), used);
    }

    void upload(uint8_t* dst) {
        // the launch part of the blob (the readback output words are written on the device)
        const size_t n = nRb ? rbOutOff : used;
        if(n) check(hipMemcpyAsync(dst, hblob, n, hipMemcpyHostToDevice, stream));
    }
    void issue(const uint8_t* base, bool timed = false) {
        if(timed && !ev0) {
            check(hipEventCreate(&ev0));
            check(hipEventCreate(&ev1));
        }
        bool started = false;
        for(const Launch& L : launches) {
            bool t = false;
            if(timed)
                for(LaunchFn f : timedFns) t |= L.fn == f;
            if(t && !started) {
                check(hipEventRecord(ev0, stream));
                started = true;
            }
            if(L.nScenes == 1)  // the functor travels as the kernel argument (host staging copy)
                L.one(staging.data() + L.bodiesOff, L.blocks, stream);
            else
                L.fn(base + L.bodiesOff, reinterpret_cast<const uint32_t*>(base + L.firstOff),
                     reinterpret_cast<const uint32_t*>(base + L.countOff), L.nScenes, L.blocks, perFor(L.blocks), stream);
            if(t) check(hipEventRecord(ev1, stream));
        }
        check(hipGetLastError());
    }

    static uint64_t hashBytes(const uint8_t* p, size_t n) {
        uint64_t h = 1469598103934665603ull;
        size_t   i = 0;
        for(; i + 8 <= n; i += 8) {
            uint64_t w;
            std::memcpy(&w, p + i, 8);
            h = (h ^ w) * 1099511628211ull;
            h ^= h >> 29;
        }
        for(; i < n; ++i) h = (h ^ p[i]) * 1099511628211ull;
        return h;
    }

    // launches the flush's kernels from a captured graph when this exact flush has been seen before;
    // returns false when the caller must issue them itself
    bool graphLaunch() {
        const size_t n = nRb ? rbOutOff : used;
        uint64_t     h = hashBytes(staging.data(), n);
        for(const Launch& L : launches) h = (h ^ (uint64_t)(uintptr_t)L.fn) * 1099511628211ull ^ L.blocks;
        ++tick;
        auto it = graphs.find(h);
        if(it != graphs.end()) {
            Graph& G = it->second;
            const bool same = G.blob.size() == n && std::memcmp(G.blob.data(), staging.data(), n) == 0 && sameLaunches(G.launches);
            if(same) {
                G.lastUse = tick;
                if(!G.exec) capture(G);
                check(hipGraphLaunch(G.exec, stream));
                ++graphReplays;
                if(nRb) upload(dblob);  // the readback table lives in the shared blob
                return true;
            }
            return false;  // a hash collision: issue directly
        }
        if(graphs.size() >= kMaxGraphs) evict();
        Graph& G   = graphs[h];
        G.blob.assign(staging.data(), staging.data() + n);
        G.launches = launches;
        G.lastUse  = tick;
        return false;
    }
    bool sameLaunches(const std::vector<Launch>& a) const {
        if(a.size() != launches.size()) return false;
        for(size_t i = 0; i < a.size(); ++i)
            if(a[i].fn != launches[i].fn || a[i].blocks != launches[i].blocks || a[i].bodiesOff != launches[i].bodiesOff) return false;
        return true;
    }
    void capture(Graph& G) {
        check(hipMalloc((void**)&G.dev, std::max<size_t>(G.blob.size(), 16)));
        check(hipMemcpyAsync(G.dev, hblob, G.blob.size(), hipMemcpyHostToDevice, stream));
        check(hipStreamSynchronize(stream));
        hipGraph_t graph = nullptr;
        check(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed));
        issue(G.dev);
        check(hipStreamEndCapture(stream, &graph));
        const hipError_t e = hipGraphInstantiate(&G.exec, graph, nullptr, nullptr, 0);
        (void)hipGraphDestroy(graph);
        check(e);
        ++graphCaptures;
    }
    void evict() {
        auto victim = graphs.begin();
        for(auto it = graphs.begin(); it != graphs.end(); ++it)
            if(it->second.lastUse < victim->second.lastUse) victim = it;
        if(victim->second.exec) (void)hipGraphExecDestroy(victim->second.exec);
        if(victim->second.dev) (void)hipFree(victim->second.dev);
        graphs.erase(victim);
    }

    hipStream_t                         stream;
    std::vector<Lane>                   lanes;
    std::vector<uint32_t>               freeLanes;
    std::vector<uint8_t>                staging;
    std::vector<Launch>                 launches;
    size_t                              used = 0, cap = 0;
    uint8_t *                           dblob = nullptr, *hblob = nullptr;
    uint32_t                            nRb = 0, rbWords = 0;
    size_t                              rbSrcOff = 0, rbWordsOff = 0, rbOffOff = 0, rbOutOff = 0;
    std::unordered_map<uint64_t, Graph> graphs;
    uint64_t                            tick = 0;
    hipEvent_t                          ev0 = nullptr, ev1 = nullptr;
};

}  // namespace bnn
