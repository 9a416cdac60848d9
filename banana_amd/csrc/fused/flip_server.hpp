// flip_server.hpp -- request loop over resident scenes: HBM-footprint admission, continuous
// batching of step / pressure-solve requests across worker streams, latency accounting.
#pragma once

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "flip_engine.hpp"

namespace bnn {

class Server {
public:
    explicit Server(const bnn_server_config& c) : cfg(c) {
        if(!cfg.hbm_budget_bytes) cfg.hbm_budget_bytes = 200000000000ull;
        if(!cfg.max_batch) cfg.max_batch = 4096;
        if(!cfg.workers) cfg.workers = 2;
        if(cfg.workers > 8) throw Error(BNN_ERR_INVALID_ARG, "at most 8 server workers");
        t0 = std::chrono::steady_clock::now();
        HIP_OK(hipStreamCreateWithFlags(&ioStream, hipStreamNonBlocking));
        for(uint32_t w = 0; w < cfg.workers; ++w) runners.emplace_back(new Runner());
        for(uint32_t w = 0; w < cfg.workers; ++w) threads.emplace_back([this, w] { workerLoop(w); });
    }

    ~Server() {
        {
            std::unique_lock<std::mutex> lk(m);
            stopping = true;
        }
        cv.notify_all();
        for(auto& t : threads) t.join();
        scenes.clear();
        runners.clear();
        if(ioStream) (void)hipStreamDestroy(ioStream);
    }

    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }

    uint64_t addScene(const bnn_flip_params& p, const bnn_scene_spec& spec, uint64_t seed) {
        const uint64_t n     = generate_scene(p, spec, seed, nullptr, nullptr, 0);
        const size_t   bytes = Scene::footprint(p, n);
        if(bytes > cfg.hbm_budget_bytes) throw Error(BNN_ERR_REJECTED, "scene footprint exceeds the HBM budget");
        auto s   = std::make_unique<Served>();
        s->prm   = p;
        s->spec  = spec;
        s->seed  = seed;
        s->np    = n;
        s->bytes = bytes;
        std::unique_lock<std::mutex> lk(m);
        const uint64_t id = ++lastSceneId;
        s->id             = id;
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
                waiting.erase(std::remove(waiting.begin(), waiting.end(), id), waiting.end());
                dead = std::move(it->second);
                scenes.erase(it);
            } else if(s.state == Served::Ready && !s.busy && s.queued == 0) {
                dead = takeLocked(it);
            }
        }
        cv.notify_all();
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
        auto                         it = reqs.find(id);
        if(it == reqs.end()) throw Error(BNN_ERR_INVALID_ARG, "unknown request");
        auto done = [&] { return reqs[id].done; };
        if(timeout < 0) doneCv.wait(lk, done);
        else if(!doneCv.wait_for(lk, std::chrono::duration<double>(timeout), done)) return false;
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
        Scene* sc = nullptr;
        {
            std::unique_lock<std::mutex> lk(m);
            auto                         it = scenes.find(id);
            if(it == scenes.end()) throw Error(BNN_ERR_INVALID_ARG, "unknown scene");
            Served& s = *it->second;
            if(s.state != Served::Ready || s.busy) throw Error(BNN_ERR_STATE, "scene is not resident and idle");
            s.busy = true;
            sc     = s.scene.get();
        }
        try {
            if(n) *n = sc->np;
            std::unique_lock<std::mutex> io(ioMutex);
            if(pos || vel) sc->getParticles(pos, vel, ioStream);
        } catch(...) {
            finishIo(id);
            throw;
        }
        finishIo(id);
    }

    bnn_server_stats stats() {
        std::unique_lock<std::mutex> lk(m);
        bnn_server_stats             o = st;
        o.scenes_resident = o.scenes_waiting = 0;
        for(auto& kv : scenes) (kv.second->state == Served::Waiting ? o.scenes_waiting : o.scenes_resident) += 1;
        o.resident_bytes      = residentBytes;
        o.peak_resident_bytes = peakBytes;
        std::vector<double> lat;
        lat.reserve(latencies.size());
        double sum = 0, mx = 0;
        for(double l : latencies) {
            lat.push_back(l);
            sum += l;
            mx = std::max(mx, l);
        }
        std::sort(lat.begin(), lat.end());
        auto pct = [&](double q) {
            if(lat.empty()) return 0.0;
            size_t k = (size_t)std::ceil(q * (double)lat.size()) - 1;
            return lat[std::min(k, lat.size() - 1)];
        };
        o.latency_p50_ms  = pct(0.50) * 1e3;
        o.latency_p99_ms  = pct(0.99) * 1e3;
        o.latency_mean_ms = lat.empty() ? 0 : sum / lat.size() * 1e3;
        o.latency_max_ms  = mx * 1e3;
        o.throughput_rps  = (lastDone > firstSubmit && firstSubmit >= 0) ? (double)latencies.size() / (lastDone - firstSubmit) : 0;
        o.mean_batch      = batchCount ? (double)batchReqs / (double)batchCount : 0;
        o.busy_s          = busy;
        return o;
    }

    void resetStats() {
        std::unique_lock<std::mutex> lk(m);
        st          = bnn_server_stats{};
        latencies.clear();
        firstSubmit = -1;
        lastDone    = -1;
        batchCount = batchReqs = 0;
        busy                   = 0;
    }

    std::string lastError() const {
        std::unique_lock<std::mutex> lk(errMutex);
        return err;
    }
    void setError(const std::string& e) {
        std::unique_lock<std::mutex> lk(errMutex);
        err = e;
    }

private:
    struct Served {
        enum State { Waiting, Creating, Ready } state = Waiting;
        uint64_t               id = 0;
        bnn_flip_params        prm{};
        bnn_scene_spec         spec{};
        uint64_t               seed = 0, np = 0;
        size_t                 bytes  = 0;
        std::unique_ptr<Scene> scene;
        bool                   busy    = false;
        bool                   release = false;
        uint32_t               queued  = 0;  // requests submitted and not finished
    };
    struct Req {
        bnn_request_result res{};
        float              dt   = 0;
        bool               done = false;
    };

    bnn_server_config                          cfg;
    std::chrono::steady_clock::time_point      t0;
    mutable std::mutex                         m, errMutex;
    std::mutex                                 ioMutex;
    std::condition_variable                    cv, doneCv;
    std::vector<std::unique_ptr<Runner>>       runners;
    std::vector<std::thread>                   threads;
    hipStream_t                                ioStream = nullptr;
    std::map<uint64_t, std::unique_ptr<Served>> scenes;
    std::deque<uint64_t>                       waiting;  // scenes in arrival order
    std::deque<uint64_t>                       queue;    // requests in arrival order
    std::unordered_map<uint64_t, Req>          reqs;
    uint64_t                                   lastSceneId = 0, lastReqId = 0;
    size_t                                     residentBytes = 0, peakBytes = 0;
    uint32_t                                   running = 0, creating = 0;
    bool                                       stopping = false;
    bnn_server_stats                           st{};
    std::vector<double>                        latencies;
    double                                     firstSubmit = -1, lastDone = -1, busy = 0;
    uint64_t                                   batchCount = 0, batchReqs = 0;
    std::string                                err;

    std::unique_ptr<Served> takeLocked(std::map<uint64_t, std::unique_ptr<Served>>::iterator it) {
        std::unique_ptr<Served> s = std::move(it->second);
        scenes.erase(it);
        residentBytes -= s->bytes;
        return s;
    }

    void finishIo(uint64_t id) {
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
    }

    // scenes whose footprint fits: in arrival order, later (smaller) scenes may pass a blocked one
    void admitLocked(std::vector<Served*>& out) {
        for(auto it = waiting.begin(); it != waiting.end();) {
            Served& s = *scenes[*it];
            if(residentBytes + s.bytes <= cfg.hbm_budget_bytes) {
                residentBytes += s.bytes;
                peakBytes = std::max(peakBytes, residentBytes);
                s.state   = Served::Creating;
                out.push_back(&s);
                it = waiting.erase(it);
            } else {
                ++it;
            }
        }
    }

    // ready requests in arrival order, one per scene, skipping scenes that are busy, not resident or
    // already in this batch (later requests of such a scene wait, keeping per-scene order)
    void pickLocked(std::vector<uint64_t>& batch) {
        const uint32_t               cap = cfg.batching ? cfg.max_batch : 1u;
        std::vector<uint64_t>        keep;
        std::unordered_map<uint64_t, bool> blocked;
        while(!queue.empty()) {
            const uint64_t id = queue.front();
            queue.pop_front();
            const uint64_t sid = reqs[id].res.scene_id;
            Served&        s   = *scenes[sid];
            if(batch.size() < cap && s.state == Served::Ready && !s.busy && !blocked[sid]) {
                s.busy = true;
                batch.push_back(id);
            } else {
                keep.push_back(id);
            }
            blocked[sid] = true;
        }
        for(auto it = keep.rbegin(); it != keep.rend(); ++it) queue.push_front(*it);
    }

    bool hasWorkLocked() {
        for(uint64_t id : waiting)
            if(residentBytes + scenes[id]->bytes <= cfg.hbm_budget_bytes) return true;
        for(uint64_t id : queue) {
            const Served& s = *scenes[reqs[id].res.scene_id];
            if(s.state == Served::Ready && !s.busy) return true;
        }
        return false;
    }

    void createScene(Runner& R, Served& s) {
        auto sc = std::make_unique<Scene>(s.prm, R.stream);
        std::vector<float> pos(s.np * 3), vel(s.np * 3);
        generate_scene(s.prm, s.spec, s.seed, pos.data(), vel.data(), s.np);
        sc->setParticles(pos.data(), vel.data(), s.np, R.stream);
        if(s.prm.boundary_sdf_at_init) {
            R.bind({sc.get()});
            R.boundarySDF(Runner::kAll);
            R.sync();
        }
        s.scene = std::move(sc);
    }

    void workerLoop(uint32_t w) {
        Runner& R = *runners[w];
        for(;;) {
            std::vector<Served*>  toCreate;
            std::vector<uint64_t> batch;
            {
                std::unique_lock<std::mutex> lk(m);
                cv.wait(lk, [&] { return stopping || hasWorkLocked(); });
                if(!hasWorkLocked()) {  // stopping
                    if(creating || running) {  // other workers may still unblock queued requests
                        cv.wait(lk);
                        continue;
                    }
                    for(uint64_t id : queue) {         // nothing can run them any more
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
                admitLocked(toCreate);
                creating += (uint32_t)toCreate.size();
                pickLocked(batch);
                running += batch.empty() ? 0 : 1;
            }
            for(Served* s : toCreate) {
                std::string msg;
                try {
                    createScene(R, *s);
                } catch(const std::exception& e) {
                    msg = e.what();
                }
                {
                    std::unique_lock<std::mutex> lk(m);
                    --creating;
                    if(msg.empty()) {
                        s->state = Served::Ready;
                    } else {
                        // the scene could not be created: fail its queued requests
                        setError(msg);
                        residentBytes -= s->bytes;
                        s->bytes = 0;
                        s->state = Served::Ready;
                        s->release = true;
                        failQueuedLocked(s->id);
                    }
                }
                cv.notify_all();
                doneCv.notify_all();
            }
            if(!batch.empty()) runBatch(R, batch);
        }
    }

    void failQueuedLocked(uint64_t sid) {
        for(auto it = queue.begin(); it != queue.end();) {
            Req& r = reqs[*it];
            if(r.res.scene_id == sid) {
                r.res.status = BNN_ERR_DEVICE;
                r.res.t_done = now();
                r.done       = true;
                ++st.failed;
                it = queue.erase(it);
            } else {
                ++it;
            }
        }
        auto s = scenes.find(sid);
        if(s != scenes.end()) {
            s->second->queued = 0;
            takeLocked(s);
        }
    }

    void runBatch(Runner& R, const std::vector<uint64_t>& batch) {
        std::vector<Scene*>   sc;
        std::vector<int>      kind;
        std::vector<uint64_t> sid;
        {
            std::unique_lock<std::mutex> lk(m);
            for(uint64_t id : batch) {
                Req& r = reqs[id];
                sid.push_back(r.res.scene_id);
                Scene* x = scenes[r.res.scene_id]->scene.get();
                x->curDt = r.dt;
                sc.push_back(x);
                kind.push_back(r.res.kind);
                r.res.t_start    = now();
                r.res.batch_size = (uint32_t)batch.size();
            }
        }
        const double t1   = now();
        bnn_status   code = BNN_OK;
        std::string  msg;
        try {
            R.bind(sc);
            std::vector<uint32_t> stepIds, presIds;
            for(uint32_t i = 0; i < sc.size(); ++i) (kind[i] == BNN_REQ_STEP ? stepIds : presIds).push_back(i);
            if(!stepIds.empty()) R.step(R.sel(stepIds));
            if(!presIds.empty()) {
                const int ps = R.sel(presIds);
                R.markDesc();
                R.beginStats();
                R.pressureProjection(ps);
                R.endStats(ps);
            }
            R.sync();
        } catch(const Error& e) {
            code = e.code;
            msg  = e.what();
        } catch(const std::exception& e) {
            code = BNN_ERR_DEVICE;
            msg  = e.what();
        }
        if(!msg.empty()) setError(msg);
        std::vector<std::unique_ptr<Served>> dead;
        {
            std::unique_lock<std::mutex> lk(m);
            const double t2 = now();
            busy += t2 - t1;
            ++batchCount;
            batchReqs += batch.size();
            ++st.batches;
            for(size_t i = 0; i < batch.size(); ++i) {
                Req& r = reqs[batch[i]];
                r.res.t_done = t2;
                r.res.status = code;
                if(code == BNN_OK) {
                    r.res.pcg_iterations = sc[i]->st.pcg_iterations;
                    r.res.pcg_success    = sc[i]->st.pcg_success;
                    r.res.pcg_residual   = sc[i]->st.pcg_residual;
                    if(!sc[i]->lastOk && sc[i]->prm.exit_if_pressure_projection_failed) r.res.status = BNN_ERR_NOT_CONVERGED;
                }
                r.done = true;
                if(r.res.status == BNN_OK) {
                    ++st.completed;
                    latencies.push_back(r.res.t_done - r.res.t_submit);
                    lastDone = std::max(lastDone, r.res.t_done);
                } else {
                    ++st.failed;
                }
                auto it = scenes.find(sid[i]);
                if(it != scenes.end()) {
                    Served& s = *it->second;
                    s.busy    = false;
                    s.queued -= 1;
                    if(s.release && s.queued == 0) dead.push_back(takeLocked(it));
                }
            }
            --running;
        }
        dead.clear();
        cv.notify_all();
        doneCv.notify_all();
    }
};

}  // namespace bnn
