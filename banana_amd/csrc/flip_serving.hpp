// flip_serving.hpp -- many independent scenes on one GPU (included by flip_solver.hip after the
// FlipSolver class): a worker pool that runs scenes concurrently, each scene on its own HIP stream,
// and a request server with HBM-footprint admission and latency accounting.
//
// Every scene keeps its own device state and stream, so a scene's results are the same bits
// whether it runs alone or next to others; concurrency comes from several host workers keeping
// several streams (hardware queues) busy while each of them waits on its own readbacks.
#pragma once

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <functional>
#include <map>
#include <optional>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

namespace bnn {

// device bytes a FlipSolver with these parameters and n particles may hold: the grid, particles
// and the pressure system sized for every brick active (its largest size)
static size_t scene_footprint(const bnn_flip_params& prm, uint64_t n) {
    const GridDims G = grid_dims(prm);
    const size_t   nx = G.n[0], ny = G.n[1], nz = G.n[2];
    const size_t   C = nx * ny * nz, N = (nx + 1) * (ny + 1) * (nz + 1);
    const size_t   nU = (nx + 1) * ny * nz, nV = nx * (ny + 1) * nz, nW = nx * ny * (nz + 1), F = nU + nV + nW;
    size_t         nbx = (nx + 7) / 8, nby = (ny + 7) / 8, nbz = (nz + 7) / 8, nb = nbx * nby * nbz;
    size_t         b = 16 * F + 2 * F + 12 * C + 4 * N + 4 + 4 * (blocks(std::max(C, nb), kScanTile) + 1) + nb +
               256 + 12 * nb + 2 * sizeof(PcgState);
    const size_t slotsCap = nb + nb / 4 + 16;
    b += 12 * 4 * 512 * slotsCap + 28 * slotsCap + 2 * 8 * kMaxRedBlocks;
    b += (prm.pressure_mode == BNN_PRESSURE_REFINED ? 6 : 1) * 8 * 512 * slotsCap;
    if(prm.preconditioner == BNN_PRECOND_MG) {
        size_t lx = nx, ly = ny, lz = nz;
        while(lx > 8 || ly > 8 || lz > 8) {
            lx = (lx + 1) / 2, ly = (ly + 1) / 2, lz = (lz + 1) / 2;
            const size_t lnb = ((lx + 7) / 8) * ((ly + 7) / 8) * ((lz + 7) / 8), cap = lnb + lnb / 4 + 4;
            b += 12 * lnb + 28 * cap + 8 * 4 * 512 * cap;
        }
    }
    // per particle: two sets of 6 float arrays, the 16 + 8-byte packed transfer records and the id
    // (the binning writes the other set), the cell key and the permutation
    if(n) b += 112 * (n + n / 8 + 1024);
    return b;
}

// fixed set of host threads; run(n, f) calls f(i) for i < n across the threads and returns when all
// calls have finished
class WorkerPool {
public:
    explicit WorkerPool(unsigned threads) {
        for(unsigned t = 0; t < std::max(threads, 1u); ++t) th.emplace_back([this] { loop(); });
    }
    ~WorkerPool() {
        {
            std::unique_lock<std::mutex> lk(m);
            stop = true;
        }
        cv.notify_all();
        for(auto& t : th) t.join();
    }
    unsigned size() const { return (unsigned)th.size(); }
    void     run(uint32_t n, const std::function<void(uint32_t)>& f) {
        std::unique_lock<std::mutex> lk(m);
        job  = &f;
        next = 0;
        todo = n;
        left = n;
        ++epoch;
        cv.notify_all();
        done.wait(lk, [&] { return left == 0; });
        job = nullptr;
    }

private:
    std::vector<std::thread>            th;
    std::mutex                          m;
    std::condition_variable             cv, done;
    const std::function<void(uint32_t)>* job  = nullptr;
    uint32_t                            next = 0, todo = 0, left = 0;
    uint64_t                            epoch = 0;
    bool                                stop  = false;
    void                                loop() {
        std::unique_lock<std::mutex> lk(m);
        for(;;) {
            cv.wait(lk, [&] { return stop || (job && next < todo); });
            if(stop) return;
            const uint32_t i = next++;
            const auto*    f = job;
            lk.unlock();
            (*f)(i);
            lk.lock();
            if(--left == 0) done.notify_all();
        }
    }
};

//==================================================================================================
// fused batches: many scenes' tasks as one stream of merged launches (flip_batch.hpp)
//==================================================================================================
class BatchRunner {
public:
    // a scene task in flight: the scene records into the runner until the task ends
    struct Task {
        FlipSolver*             s = nullptr;
        std::optional<Co<bool>> co;
        std::coroutine_handle<> next;
        uint32_t                lane        = 0;
        uint32_t                barrierWait = 0;  // rounds spent waiting at the current barrier
        int32_t                 savedTiming = 0;
        bool                    done        = false;
        bool                    ok          = false;
        bnn_status              status      = BNN_OK;
        std::string             err;
    };

    // launch shapes the autotuner chooses among; none of them changes a result bit
    struct Shape {
        int      p2gSeg     = 8;
        uint32_t chunkFirst = 8, chunkMax = 8;
        uint32_t maxPer     = 16;     // work items per workgroup of a batched launch, at most ...
        uint32_t minGroups  = 16384;  // ... keeping at least this many workgroups
        bool     graphs     = true;
    };

    BatchRunner() {
        HIP_OK(hipGetDevice(&device));
        HIP_OK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        rec = std::make_unique<Recorder>(stream);
    }
    ~BatchRunner() {
        rec.reset();
        if(stream) (void)hipStreamDestroy(stream);
    }
    BatchRunner(const BatchRunner&)            = delete;
    BatchRunner& operator=(const BatchRunner&) = delete;

    uint32_t     barrierRounds = 0xFFFFFFFFu;  // rounds a scene may wait at a barrier
    Recorder&    recorder() { return *rec; }
    const Shape& shape() const { return cur; }
    void         setShape(const Shape& sh) { cur = sh; }

    // starts `co` (a task of scene s); the scene must be idle (every single-scene call leaves its
    // stream idle) and on this runner's device
    void start(Task& t, FlipSolver* s, Co<bool>&& co) {
        if(s->device != device) throw Error(BNN_ERR_INVALID_ARG, "a scene of the batch lives on another device");
        if(s->rec) throw Error(BNN_ERR_STATE, "a scene is already running in a batch");
        t.s           = s;
        t.lane        = rec->takeLane();
        t.savedTiming = s->prm.timing;
        s->prm.timing = 0;  // per-scene events do not apply to merged launches
        s->rec        = rec.get();
        s->lane       = t.lane;
        s->pending    = {};
        s->p2gSeg        = cur.p2gSeg;
        s->pcgChunkFirst = cur.chunkFirst;
        s->pcgChunkMax   = cur.chunkMax;
        t.co.emplace(std::move(co));
        t.next        = t.co->h;
        t.done        = false;
        t.barrierWait = 0;
    }

    // resumes every task waiting at a sync point (tasks started since the last round run to their
    // first one), then flushes; tasks that ended are then marked done and detached from the runner
    // (their last launches were in the flush)
    void round(const std::vector<Task*>& tasks) {
        Recorder* const prev = tl_batch;
        tl_batch             = rec.get();
        // scenes at a barrier wait while any other scene can still advance (at most barrierRounds
        // rounds: requests that keep arriving must not hold them back for ever)
        bool others = false;
        for(Task* t : tasks) others |= !t->done && t->next && !t->s->atBarrier;
        for(Task* t : tasks) {
            if(t->done || !t->next) continue;
            if(others && t->s->atBarrier && ++t->barrierWait <= barrierRounds) continue;
            const std::coroutine_handle<> h = t->next;
            t->next                         = {};
            t->s->pending                   = {};
            t->s->atBarrier                 = false;
            t->barrierWait                  = 0;
            h.resume();
            if(!t->co->done()) {
                t->next = t->s->pending;
                if(!t->next) {
                    tl_batch = prev;
                    throw std::logic_error("a batched task stopped outside a sync point");
                }
            }
        }
        try {
            rec->useGraphs = cur.graphs;
            rec->maxPer    = cur.maxPer;
            rec->minGroups = cur.minGroups;
            rec->flushAndSync();
        } catch(const std::exception& e) {
            tl_batch = prev;
            rec->discard();
            (void)hipStreamSynchronize(stream);
            for(Task* t : tasks)
                if(!t->done) fail(*t, BNN_ERR_DEVICE, e.what());
            return;
        }
        tl_batch = prev;
        for(Task* t : tasks)
            if(!t->done && t->co->done()) collect(*t);
    }

    // runs every task to its end
    void runAll(const std::vector<Task*>& tasks) {
        for(;;) {
            bool live = false;
            for(Task* t : tasks) live |= !t->done;
            if(!live) return;
            round(tasks);
        }
    }

private:
    void detach(Task& t) {
        t.s->rec        = nullptr;
        t.s->pending    = {};
        t.s->atBarrier  = false;
        t.s->prm.timing = t.savedTiming;
        rec->giveLane(t.lane);
        t.co.reset();
        t.next = {};
        t.done = true;
    }
    void collect(Task& t) {
        try {
            t.ok     = t.co->result();
            t.status = BNN_OK;
        } catch(const Error& e) {
            t.status = e.code;
            t.err    = e.what();
        } catch(const std::exception& e) {
            t.status = BNN_ERR_DEVICE;
            t.err    = e.what();
        }
        detach(t);
    }
    void fail(Task& t, bnn_status code, const std::string& msg) {
        t.status = code;
        t.err    = msg;
        detach(t);
    }

    int                       device = 0;
    hipStream_t               stream = nullptr;
    std::unique_ptr<Recorder> rec;
    Shape                     cur;
};

// The autotuner of a batch: the first batched calls try launch shapes, one per call: P2G segment
// lengths 8, 4 and 16 (scored by the P2G kernel time, measured with events) together with PCG
// stop-flag chunks of 8, 4 and 16 iterations and at most 16, 4 and 1 work items per workgroup
// (scored by the call's wall time less the P2G time), then graphs off at the best shape. The best
// shape is kept for every later call with the same signature (scene count and total cells, in
// powers of two).
class BatchTuner {
public:
    struct Trial {
        BatchRunner::Shape shape;
        double             p2gMs = 0, restMs = 0;
    };
    // the shape for the next call of this signature
    BatchRunner::Shape next(uint64_t sig) {
        Entry& e = table[sig];
        if(e.fixed) return e.best;
        const size_t k = e.trials.size();
        if(k < 3) {
            BatchRunner::Shape sh;
            sh.p2gSeg     = k == 0 ? 8 : (k == 1 ? 4 : 16);
            sh.chunkFirst = sh.chunkMax = k == 0 ? 8u : (k == 1 ? 4u : 16u);
            sh.maxPer     = k == 0 ? 16u : (k == 1 ? 4u : 1u);
            sh.graphs     = true;
            return sh;
        }
        BatchRunner::Shape sh = bestOfThree(e);
        sh.graphs             = false;
        return sh;
    }
    // records the outcome of a call made with `sh`
    void report(uint64_t sig, const BatchRunner::Shape& sh, double p2gMs, double wallMs) {
        Entry& e = table[sig];
        if(e.fixed) return;
        e.trials.push_back(Trial{sh, p2gMs, std::max(0.0, wallMs - p2gMs)});
        if(e.trials.size() == 4) {
            BatchRunner::Shape b = bestOfThree(e);
            const Trial&       g = e.trials[3];
            double             graphRest = 1e300;
            for(size_t i = 0; i < 3; ++i)
                if(e.trials[i].shape.chunkFirst == b.chunkFirst && e.trials[i].shape.maxPer == b.maxPer) graphRest = e.trials[i].restMs;
            b.graphs = !(g.restMs < graphRest);
            e.best   = b;
            e.fixed  = true;
        }
    }
    bool fixed(uint64_t sig) const {
        auto it = table.find(sig);
        return it != table.end() && it->second.fixed;
    }
    std::vector<Trial> trials(uint64_t sig) const {
        auto it = table.find(sig);
        return it == table.end() ? std::vector<Trial>{} : it->second.trials;
    }

private:
    struct Entry {
        std::vector<Trial> trials;
        BatchRunner::Shape best;
        bool               fixed = false;
    };
    static BatchRunner::Shape bestOfThree(const Entry& e) {
        size_t bp = 0, bc = 0;
        for(size_t i = 1; i < 3; ++i) {
            if(e.trials[i].p2gMs < e.trials[bp].p2gMs) bp = i;
            if(e.trials[i].restMs < e.trials[bc].restMs) bc = i;
        }
        BatchRunner::Shape sh;
        sh.p2gSeg     = e.trials[bp].shape.p2gSeg;
        sh.chunkFirst = e.trials[bc].shape.chunkFirst;
        sh.chunkMax   = e.trials[bc].shape.chunkMax;
        sh.maxPer     = e.trials[bc].shape.maxPer;
        sh.graphs     = true;
        return sh;
    }
    std::unordered_map<uint64_t, Entry> table;
};

// the batched launch of P2G: the affine transfer's column gather with segment seg (4, 8 or 16), or
// the FLIP / PIC tile scatter (which has no segment)
static inline Recorder::LaunchFn p2g_launch_fn(int seg, bool affine) {
    if(affine)
        return seg == 4 ? Recorder::fnOf<k_p2g_col<true, 4>>() : (seg == 16 ? Recorder::fnOf<k_p2g_col<true, 16>>() : Recorder::fnOf<k_p2g_col<true, 8>>());
    return Recorder::fnOf<k_p2g_tile>();
}

static inline uint64_t batch_signature(const std::vector<FlipSolver*>& sc) {
    uint64_t cells = 0;
    for(const FlipSolver* s : sc) cells += s->nC;
    auto lg = [](uint64_t v) {
        uint64_t l = 0;
        while(v > 1) {
            v >>= 1;
            ++l;
        }
        return l;
    };
    return lg(sc.size()) << 32 | lg(cells);
}

//==================================================================================================
// request server
//==================================================================================================
// With batching (the default) one engine thread runs every admitted request as a task of a fused
// batch (BatchRunner): each round it admits the ready requests (up to max_batch in flight), resumes
// the tasks to their next sync point and issues all their launches as merged launches, so requests
// that arrive while others run join at the next sync point. Worker threads only create scenes.
// Without batching, one worker runs one request at a time through the scene's own stream.
class Server {
public:
    explicit Server(const bnn_server_config& c) : cfg(c) {
        if(!cfg.hbm_budget_bytes) cfg.hbm_budget_bytes = 200000000000ull;
        if(!cfg.workers) cfg.workers = 4;
        if(!cfg.batching) cfg.workers = 1;
        if(!cfg.max_batch) cfg.max_batch = 1024;
        if(cfg.workers > 16) throw Error(BNN_ERR_INVALID_ARG, "at most 16 server workers");
        t0 = std::chrono::steady_clock::now();
        if(cfg.batching) {
            HIP_OK(hipGetDevice(&device));
            engine = std::thread([this] { engineLoop(); });
        }
        for(uint32_t w = 0; w < cfg.workers; ++w) threads.emplace_back([this] { workerLoop(); });
    }

    ~Server() {
        {
            std::unique_lock<std::mutex> lk(m);
            stopping = true;
        }
        cv.notify_all();
        for(auto& t : threads) t.join();
        if(engine.joinable()) engine.join();
        scenes.clear();
    }

    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }

    uint64_t addScene(const bnn_flip_params& p, const bnn_scene_spec& spec, uint64_t seed) {
        const uint64_t n     = generate_scene(p, spec, seed, nullptr, nullptr, 0);
        const size_t   bytes = scene_footprint(p, n);
        if(bytes > cfg.hbm_budget_bytes) throw Error(BNN_ERR_REJECTED, "scene footprint exceeds the HBM budget");
        auto s   = std::make_unique<Served>();
        s->prm   = p;
        s->spec  = spec;
        s->seed  = seed;
        s->np    = n;
        s->bytes = bytes;
        std::unique_lock<std::mutex> lk(m);
        const uint64_t               id = ++lastSceneId;
        s->id                           = id;
        waiting.push_back(id);
        scenes[id] = std::move(s);
        lk.unlock();
        cv.notify_all();
        return id;
    }

    void releaseScene(uint64_t id) {
        std::unique_ptr<Served> dead;
        {
            std::unique_lock<std::mutex> lk(m);
            auto                         it = scenes.find(id);
            if(it == scenes.end()) throw Error(BNN_ERR_INVALID_ARG, "unknown scene");
            Served& s = *it->second;
            s.release = true;
            if(s.state == Served::Waiting) {
                // never admitted: its queued requests cannot run any more
                failQueuedLocked(id, BNN_ERR_STATE);
                waiting.erase(std::remove(waiting.begin(), waiting.end(), id), waiting.end());
                dead = std::move(it->second);
                scenes.erase(it);
            } else if(s.state == Served::Ready && !s.busy && s.queued == 0) {
                dead = takeLocked(it);
            }
        }
        cv.notify_all();
        doneCv.notify_all();
    }

    uint64_t submit(uint64_t sceneId, int kind, float dt) {
        if(kind != BNN_REQ_STEP && kind != BNN_REQ_PRESSURE) throw Error(BNN_ERR_INVALID_ARG, "unknown request kind");
        if(!(dt > 0)) throw Error(BNN_ERR_INVALID_ARG, "dt must be > 0");
        std::unique_lock<std::mutex> lk(m);
        auto                         it = scenes.find(sceneId);
        if(it == scenes.end() || it->second->release) throw Error(BNN_ERR_INVALID_ARG, "unknown scene");
        const uint64_t id = ++lastReqId;
        Req&           r  = reqs[id];
        r.res.request_id  = id;
        r.res.scene_id    = sceneId;
        r.res.kind        = kind;
        r.res.t_submit    = now();
        r.dt              = dt;
        it->second->queued += 1;
        queue.push_back(id);
        ++st.submitted;
        if(firstSubmit < 0) firstSubmit = r.res.t_submit;
        lk.unlock();
        cv.notify_all();
        return id;
    }

    bool wait(uint64_t id, double timeout, bnn_request_result* out) {
        std::unique_lock<std::mutex> lk(m);
        if(reqs.find(id) == reqs.end()) throw Error(BNN_ERR_INVALID_ARG, "unknown request");
        auto isDone = [&] { return reqs[id].done; };
        if(timeout < 0) doneCv.wait(lk, isDone);
        else if(!doneCv.wait_for(lk, std::chrono::duration<double>(timeout), isDone)) return false;
        if(out) *out = reqs[id].res;
        return true;
    }

    bool drain(double timeout) {
        std::unique_lock<std::mutex> lk(m);
        auto idle = [&] { return queue.empty() && running == 0 && creating == 0; };
        if(timeout < 0) doneCv.wait(lk, idle);
        else if(!doneCv.wait_for(lk, std::chrono::duration<double>(timeout), idle)) return false;
        return true;
    }

    void getParticles(uint64_t id, float* pos, float* vel, uint64_t* n) {
        FlipSolver* sc = nullptr;
        {
            std::unique_lock<std::mutex> lk(m);
            auto                         it = scenes.find(id);
            if(it == scenes.end()) throw Error(BNN_ERR_INVALID_ARG, "unknown scene");
            Served& s = *it->second;
            if(s.state != Served::Ready || s.busy || !s.solver) throw Error(BNN_ERR_STATE, "scene is not resident and idle");
            s.busy = true;
            sc     = s.solver.get();
        }
        std::string msg;
        bnn_status  code = BNN_OK;
        try {
            if(n) *n = sc->np;
            if(pos || vel) sc->getParticles(pos, vel);
        } catch(const Error& e) {
            code = e.code;
            msg  = e.what();
        }
        std::unique_ptr<Served> dead;
        {
            std::unique_lock<std::mutex> lk(m);
            auto                         it = scenes.find(id);
            if(it != scenes.end()) {
                it->second->busy = false;
                if(it->second->release && it->second->queued == 0) dead = takeLocked(it);
            }
        }
        cv.notify_all();
        if(code != BNN_OK) throw Error(code, msg);
    }

    bnn_server_stats stats() {
        std::unique_lock<std::mutex> lk(m);
        bnn_server_stats             o = st;
        o.scenes_resident = o.scenes_waiting = 0;
        for(auto& kv : scenes) (kv.second->state == Served::Waiting ? o.scenes_waiting : o.scenes_resident) += 1;
        o.resident_bytes      = residentBytes;
        o.peak_resident_bytes = peakBytes;
        std::vector<double> lat(latencies);
        double              sum = 0, mx = 0;
        for(double l : lat) {
            sum += l;
            mx = std::max(mx, l);
        }
        std::sort(lat.begin(), lat.end());
        auto pct = [&](double q) {
            if(lat.empty()) return 0.0;
            size_t k = (size_t)std::ceil(q * (double)lat.size());
            k        = k ? k - 1 : 0;
            return lat[std::min(k, lat.size() - 1)];
        };
        o.latency_p50_ms  = pct(0.50) * 1e3;
        o.latency_p99_ms  = pct(0.99) * 1e3;
        o.latency_mean_ms = lat.empty() ? 0 : sum / lat.size() * 1e3;
        o.latency_max_ms  = mx * 1e3;
        o.throughput_rps  = (firstSubmit >= 0 && lastDone > firstSubmit) ? (double)lat.size() / (lastDone - firstSubmit) : 0;
        o.mean_batch      = cfg.batching ? (roundsSince ? (double)lanesSince / (double)roundsSince : 0.0) : 1.0;
        o.busy_s          = busy;
        return o;
    }

    void resetStats() {
        std::unique_lock<std::mutex> lk(m);
        st = bnn_server_stats{};
        latencies.clear();
        firstSubmit = lastDone = -1;
        busy                   = 0;
        lanesSince = roundsSince = 0;
    }

    std::string lastError() const {
        std::unique_lock<std::mutex> lk(errMutex);
        return err;
    }

private:
    struct Served {
        enum State { Waiting, Creating, Ready } state = Waiting;
        uint64_t                    id = 0;
        bnn_flip_params             prm{};
        bnn_scene_spec              spec{};
        uint64_t                    seed = 0, np = 0;
        size_t                      bytes = 0;
        std::unique_ptr<FlipSolver> solver;
        bool                        busy    = false;
        bool                        release = false;
        uint32_t                    queued  = 0;  // requests submitted and not finished
    };
    struct Req {
        bnn_request_result res{};
        float              dt   = 0;
        bool               done = false;
    };

    bnn_server_config                           cfg;
    int                                         device = 0;
    std::thread                                 engine;
    uint64_t                                    lanesSince = 0, roundsSince = 0;  // requests in flight summed over engine rounds
    std::chrono::steady_clock::time_point       t0;
    mutable std::mutex                          m, errMutex;
    std::condition_variable                     cv, doneCv;
    std::vector<std::thread>                    threads;
    std::map<uint64_t, std::unique_ptr<Served>> scenes;
    std::deque<uint64_t>                        waiting;  // scenes in arrival order
    std::deque<uint64_t>                        queue;    // requests in arrival order
    std::unordered_map<uint64_t, Req>           reqs;
    uint64_t                                    lastSceneId = 0, lastReqId = 0;
    size_t                                      residentBytes = 0, peakBytes = 0;
    uint32_t                                    running = 0, creating = 0;
    bool                                        stopping = false;
    bnn_server_stats                            st{};
    std::vector<double>                         latencies;
    double                                      firstSubmit = -1, lastDone = -1, busy = 0;
    std::string                                 err;

    void setError(const std::string& e) {
        std::unique_lock<std::mutex> lk(errMutex);
        err = e;
    }

    std::unique_ptr<Served> takeLocked(std::map<uint64_t, std::unique_ptr<Served>>::iterator it) {
        std::unique_ptr<Served> s = std::move(it->second);
        scenes.erase(it);
        residentBytes -= s->bytes;
        return s;
    }

    Served* findLocked(uint64_t id) {
        auto it = scenes.find(id);
        return it == scenes.end() ? nullptr : it->second.get();
    }

    // the first waiting scene (arrival order) whose footprint fits; later, smaller ones may pass
    Served* admitLocked() {
        for(auto it = waiting.begin(); it != waiting.end(); ++it) {
            Served* sp = findLocked(*it);
            if(!sp) continue;
            Served& s = *sp;
            if(residentBytes + s.bytes <= cfg.hbm_budget_bytes) {
                residentBytes += s.bytes;
                peakBytes = std::max(peakBytes, residentBytes);
                s.state   = Served::Creating;
                waiting.erase(it);
                return &s;
            }
        }
        return nullptr;
    }

    // the first queued request whose scene is resident and idle; requests of a scene run in
    // submission order (a later request never passes an earlier one of the same scene)
    bool pickLocked(uint64_t& rid) {
        std::unordered_map<uint64_t, bool> seen;
        for(auto it = queue.begin(); it != queue.end(); ++it) {
            const uint64_t sid = reqs[*it].res.scene_id;
            Served*        sp  = findLocked(sid);
            if(sp && !seen[sid] && sp->state == Served::Ready && !sp->busy) {
                Served& s = *sp;
                s.busy = true;
                rid    = *it;
                queue.erase(it);
                return true;
            }
            seen[sid] = true;
        }
        return false;
    }

    bool admissibleLocked() {
        for(uint64_t id : waiting) {
            const Served* s = findLocked(id);
            if(s && residentBytes + s->bytes <= cfg.hbm_budget_bytes) return true;
        }
        return false;
    }
    bool runnableLocked() {
        for(uint64_t id : queue) {
            const Served* s = findLocked(reqs[id].res.scene_id);
            if(s && s->state == Served::Ready && !s->busy) return true;
        }
        return false;
    }

    void failQueuedLocked(uint64_t sid, bnn_status code) {
        for(auto it = queue.begin(); it != queue.end();) {
            Req& r = reqs[*it];
            if(r.res.scene_id == sid) {
                r.res.status = code;
                r.res.t_done = now();
                r.done       = true;
                ++st.failed;
                it = queue.erase(it);
            } else {
                ++it;
            }
        }
    }

    void workerLoop() {
        for(;;) {
            Served*  toCreate = nullptr;
            uint64_t rid      = 0;
            bool     haveReq  = false;
            {
                std::unique_lock<std::mutex> lk(m);
                for(;;) {
                    if(admissibleLocked() || (!cfg.batching && runnableLocked())) break;
                    if(stopping && cfg.batching) return;  // the engine fails what cannot run
                    if(stopping) {
                        if(creating || running) {  // another worker may still unblock queued requests
                            cv.wait(lk);
                            continue;
                        }
                        for(uint64_t id : queue) {  // nothing can run them any more
                            Req& r       = reqs[id];
                            r.res.status = BNN_ERR_STATE;
                            r.res.t_done = now();
                            r.done       = true;
                            ++st.failed;
                        }
                        queue.clear();
                        doneCv.notify_all();
                        cv.notify_all();
                        return;
                    }
                    cv.wait(lk);
                }
                toCreate = admitLocked();
                if(toCreate) ++creating;
                else if(!cfg.batching && (haveReq = pickLocked(rid))) ++running;
            }
            if(toCreate) create(*toCreate);
            if(haveReq) run(rid);
        }
    }

    // one request in flight on the engine
    struct Flight {
        uint64_t                    rid = 0, sid = 0;
        FlipSolver*                 sv  = nullptr;
        int                         kind = 0;
        float                       dt   = 0;
        double                      t1   = 0;
        BatchRunner::Task           task;
    };

    void engineLoop() {
        std::string                          msg;
        std::vector<std::unique_ptr<Flight>> fl;
        try {
            HIP_OK(hipSetDevice(device));
            BatchRunner R;
            R.barrierRounds = (uint32_t)std::max(cfg.barrier_rounds, 0);
            BatchRunner::Shape sh;  // fixed shape: requests come and go, so no per-size tuning
            sh.chunkFirst = 4;
            sh.chunkMax   = 16;
            sh.maxPer     = 4;
            R.setShape(sh);
            std::vector<BatchRunner::Task*> ptrs;
            for(;;) {
                std::vector<Flight*> fresh;
                {
                    std::unique_lock<std::mutex> lk(m);
                    for(;;) {
                        uint64_t rid = 0;
                        while(fl.size() < cfg.max_batch && (fl.empty() || !cfg.max_batch_mcells || cellsInFlight(fl) < 1e6 * cfg.max_batch_mcells) &&
                              pickLocked(rid)) {
                            auto f  = std::make_unique<Flight>();
                            Req& r  = reqs[rid];
                            f->rid  = rid;
                            f->sid  = r.res.scene_id;
                            f->sv   = findLocked(f->sid)->solver.get();  // busy scenes are never erased
                            f->kind = r.res.kind;
                            f->dt   = r.dt;
                            ++running;
                            fresh.push_back(f.get());
                            fl.push_back(std::move(f));
                        }
                        if(!fl.empty()) break;
                        if(stopping && creating == 0 && !admissibleLocked()) {
                            for(uint64_t id : queue) {  // nothing can run them any more
                                Req& r       = reqs[id];
                                r.res.status = BNN_ERR_STATE;
                                r.res.t_done = now();
                                r.done       = true;
                                ++st.failed;
                            }
                            queue.clear();
                            doneCv.notify_all();
                            return;
                        }
                        cv.wait(lk);
                    }
                    const double tn = now();
                    for(Flight* f : fresh) {
                        Req& r           = reqs[f->rid];
                        r.res.t_start    = tn;
                        r.res.batch_size = (uint32_t)fl.size();
                        f->t1            = tn;
                    }
                }
                if(fl.size() == 1 && fresh.size() == 1) {
                    // a lone request (nothing else in flight or ready): it runs directly on its scene's
                    // stream, without the recorder's per-sync-point flush; requests that arrive meanwhile
                    // join the next round. The last batched round ended with a wait, and a direct run
                    // ends with one, so the scene's work stays in order across the two modes.
                    Flight* f = fresh[0];
                    try {
                        f->task.ok = run_now(f->kind == BNN_REQ_STEP ? f->sv->step(f->dt) : f->sv->projectionRequest(f->dt));
                    } catch(const Error& e) {
                        f->task.status = e.code;
                        f->task.err    = e.what();
                        (void)hipStreamSynchronize(f->sv->stream);
                    } catch(const std::exception& e) {
                        f->task.status = BNN_ERR_DEVICE;
                        f->task.err    = e.what();
                        (void)hipStreamSynchronize(f->sv->stream);
                    }
                    f->task.done = true;
                    {
                        std::unique_lock<std::mutex> lk(m);
                        lanesSince += 1;
                        roundsSince += 1;
                    }
                    std::unique_ptr<Flight> one = std::move(fl[0]);
                    fl.clear();
                    finish(*one);
                    continue;
                }
                for(Flight* f : fresh) {
                    try {
                        R.start(f->task, f->sv, f->kind == BNN_REQ_STEP ? f->sv->step(f->dt) : f->sv->projectionRequest(f->dt));
                    } catch(const Error& e) {
                        f->task.done   = true;
                        f->task.status = e.code;
                        f->task.err    = e.what();
                    }
                }
                ptrs.clear();
                for(auto& f : fl) ptrs.push_back(&f->task);
                R.round(ptrs);
                std::vector<std::unique_ptr<Flight>> keep, ended;
                for(auto& f : fl) (f->task.done ? ended : keep).push_back(std::move(f));
                fl.swap(keep);
                {
                    std::unique_lock<std::mutex> lk(m);
                    lanesSince += ptrs.size();
                    roundsSince += 1;
                }
                for(auto& f : ended) finish(*f);
            }
        } catch(const std::exception& e) {
            msg = e.what();
        }
        setError("serving engine stopped: " + msg);
        for(auto& f : fl) {  // requests in flight when the engine failed
            f->sv->rec       = nullptr;
            f->sv->pending   = {};
            f->sv->atBarrier = false;
            f->task.done   = true;
            f->task.status = BNN_ERR_DEVICE;
            f->task.err    = msg;
            finish(*f);
        }
        std::unique_lock<std::mutex> lk(m);
        for(uint64_t id : queue) {
            Req& r       = reqs[id];
            r.res.status = BNN_ERR_DEVICE;
            r.res.t_done = now();
            r.done       = true;
            ++st.failed;
        }
        queue.clear();
        doneCv.notify_all();
    }

    static double cellsInFlight(const std::vector<std::unique_ptr<Flight>>& fl) {
        double c = 0;
        for(const auto& f : fl) c += (double)f->sv->nC;
        return c;
    }

    // records a finished engine request and frees its scene
    void finish(Flight& f) {
        const BatchRunner::Task& t    = f.task;
        bnn_status               code = t.status;
        if(code != BNN_OK) setError(t.err);
        std::unique_ptr<Served> dead;
        {
            std::unique_lock<std::mutex> lk(m);
            const double t2 = now();
            busy += t2 - f.t1;
            Req& r       = reqs[f.rid];
            r.res.t_done = t2;
            r.res.status = code;
            if(code == BNN_OK) {
                r.res.pcg_iterations = f.sv->st.pcg_iterations;
                r.res.pcg_success    = f.sv->st.pcg_success;
                r.res.pcg_residual   = f.sv->st.pcg_residual;
                if(!t.ok && f.sv->prm.exit_if_pressure_projection_failed) r.res.status = BNN_ERR_NOT_CONVERGED;
            }
            r.done = true;
            if(r.res.status == BNN_OK) {
                ++st.completed;
                latencies.push_back(r.res.t_done - r.res.t_submit);
                lastDone = std::max(lastDone, r.res.t_done);
            } else {
                ++st.failed;
            }
            ++st.batches;
            auto it = scenes.find(f.sid);
            if(it != scenes.end()) {
                Served& sc = *it->second;
                sc.busy    = false;
                sc.queued -= 1;
                if(sc.release && sc.queued == 0) dead = takeLocked(it);
            }
            --running;
        }
        cv.notify_all();
        doneCv.notify_all();
    }

    void create(Served& s) {
        std::string msg;
        try {
            auto               sv = std::make_unique<FlipSolver>(s.prm);
            std::vector<float> pos(s.np * 3), vel(s.np * 3);
            generate_scene(s.prm, s.spec, s.seed, pos.data(), vel.data(), s.np);
            sv->setParticles(pos.data(), vel.data(), s.np);
            sv->sortParticles();  // as bnn_flip_generate_scene
            s.solver = std::move(sv);
        } catch(const std::exception& e) {
            msg = e.what();
        }
        std::unique_ptr<Served> dead;
        {
            std::unique_lock<std::mutex> lk(m);
            --creating;
            s.state = Served::Ready;
            if(!msg.empty()) {  // could not be created: fail its requests, free its budget
                setError(msg);
                failQueuedLocked(s.id, BNN_ERR_OUT_OF_MEMORY);
                s.queued = 0;
                dead     = takeLocked(scenes.find(s.id));
            } else if(s.release && s.queued == 0) {
                dead = takeLocked(scenes.find(s.id));
            }
        }
        cv.notify_all();
        doneCv.notify_all();
    }

    void run(uint64_t rid) {
        FlipSolver* sv = nullptr;
        int         kind;
        float       dt;
        uint64_t    sid;
        {
            std::unique_lock<std::mutex> lk(m);
            Req& r           = reqs[rid];
            sid              = r.res.scene_id;
            Served* sp       = findLocked(sid);  // busy scenes are never erased
            if(!sp || !sp->solver) {              // cannot happen while pickLocked marks it busy
                r.res.status = BNN_ERR_STATE;
                r.res.t_done = now();
                r.done       = true;
                ++st.failed;
                --running;
                lk.unlock();
                cv.notify_all();
                doneCv.notify_all();
                return;
            }
            sv               = sp->solver.get();
            kind             = r.res.kind;
            dt               = r.dt;
            r.res.t_start    = now();
            r.res.batch_size = running;
        }
        const double t1   = now();
        bnn_status   code = BNN_OK;
        bool         ok   = true;
        try {
            ok = run_now(kind == BNN_REQ_STEP ? sv->step(dt) : sv->projectionRequest(dt));
        } catch(const Error& e) {
            code = e.code;
            setError(e.what());
        } catch(const std::exception& e) {
            code = BNN_ERR_DEVICE;
            setError(e.what());
        }
        if(code != BNN_OK) (void)hipStreamSynchronize(sv->stream);  // nothing of this request stays queued
        std::unique_ptr<Served> dead;
        {
            std::unique_lock<std::mutex> lk(m);
            const double t2 = now();
            busy += t2 - t1;
            Req& r       = reqs[rid];
            r.res.t_done = t2;
            r.res.status = code;
            if(code == BNN_OK) {
                r.res.pcg_iterations = sv->st.pcg_iterations;
                r.res.pcg_success    = sv->st.pcg_success;
                r.res.pcg_residual   = sv->st.pcg_residual;
                if(!ok && sv->prm.exit_if_pressure_projection_failed) r.res.status = BNN_ERR_NOT_CONVERGED;
            }
            r.done = true;
            if(r.res.status == BNN_OK) {
                ++st.completed;
                latencies.push_back(r.res.t_done - r.res.t_submit);
                lastDone = std::max(lastDone, r.res.t_done);
            } else {
                ++st.failed;
            }
            ++st.batches;
            auto it = scenes.find(sid);
            if(it != scenes.end()) {
                Served& s = *it->second;
                s.busy    = false;
                s.queued -= 1;
                if(s.release && s.queued == 0) dead = takeLocked(it);
            }
            --running;
        }
        cv.notify_all();
        doneCv.notify_all();
    }
};

}  // namespace bnn
