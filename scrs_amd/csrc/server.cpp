// server.cpp - serving loop: admission by payload footprint, continuous
// batching through the batch engine, or one codec call per job.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <mutex>
#include <thread>
#include <tuple>
#include <vector>

#include "device.hpp"
#include "engine.hpp"
#include "status.hpp"

using namespace scrs;

extern "C" uint64_t scrs_now_ns(void) {
    return static_cast<uint64_t>(
        std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now().time_since_epoch())
            .count());
}

namespace {

struct Request {
    scrs_job job{};
    std::vector<uint16_t> idx;
    std::vector<const uint8_t *> pay;
    uint64_t tag = 0, t_submit = 0, t_start = 0, bytes = 0;
};

struct InFlight {
    hipEvent_t ev = nullptr;
    hipEvent_t start = nullptr;  // timing: enqueued before the batch's work
    std::vector<Request *> reqs;
    uint64_t bytes = 0;
    uint64_t ring_end = 0;  // plain mode: pointer-ring position released on completion
    scrs_status status{};
};

using Key = std::tuple<int32_t, uint64_t, uint64_t, uint64_t>;

uint64_t footprint(const scrs_job &j) {
    const uint64_t k = j.profile.k, m = j.profile.m, L = j.profile.symbol_len;
    return j.op == SCRS_JOB_ENCODE ? (k + m) * L : 2 * k * L;
}

}  // namespace

struct scrs_server {
    scrs_server_config cfg{};
    int device = 0;
    hipStream_t stream = nullptr;   // plain mode; batched mode's first stream
    // Batched mode: consecutive batches go to the streams in turn, so one
    // batch's copy and solve overlap the previous batch's multiply-XOR.
    static constexpr int kBatchStreams = 2;
    static constexpr size_t kBatchesInFlight = 3;  // below the engine's staging slots
    hipStream_t streams[kBatchStreams] = {};
    int next_stream = 0;
    scrs_batch_engine *eng = nullptr;
    // Plain mode: codecs per geometry and a ring of device pointer arrays.
    std::map<Key, scrs_batch_encoder *> encs;
    std::map<Key, scrs_incr_encoder *> incrs;
    std::map<Key, scrs_batch_decoder *> decs;
    DevBuf ring;
    HostBuf ring_host;
    uint64_t ring_head = 0, ring_tail = 0, ring_cap = 0;
    std::vector<hipEvent_t> events;

    std::mutex mu;
    std::condition_variable cv_submit, cv_done;
    std::deque<Request *> pending;
    std::deque<scrs_completion> done;
    std::deque<InFlight> inflight;
    uint64_t inflight_bytes = 0, outstanding = 0;
    scrs_server_stats counters{};
    std::atomic<bool> stop{false};
    std::thread th;

    hipEvent_t take_event() {
        if (!events.empty()) {
            hipEvent_t e = events.back();
            events.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        (void)hipEventCreate(&e);
        return e;
    }

    void finish(InFlight &f, uint64_t t) {
        float ms = 0;
        const bool timed = f.start && hipEventElapsedTime(&ms, f.start, f.ev) == hipSuccess;
        if (f.start) events.push_back(f.start);
        std::lock_guard<std::mutex> lock(mu);
        if (timed) {
            counters.gpu_ns += static_cast<uint64_t>(ms * 1e6);
            counters.max_gpu_ns = std::max<uint64_t>(counters.max_gpu_ns, static_cast<uint64_t>(ms * 1e6));
        }
        for (Request *r : f.reqs) {
            scrs_completion c;
            c.tag = r->tag;
            c.status = f.status;
            c.t_submit_ns = r->t_submit;
            c.t_start_ns = r->t_start;
            c.t_done_ns = t;
            done.push_back(c);
            delete r;
        }
        inflight_bytes -= f.bytes;
        cv_done.notify_all();
    }

    scrs_status plain_one(Request *r, uint64_t *ring_end);
    void loop();
};

scrs_status scrs_server::plain_one(Request *r, uint64_t *ring_end) {
    const scrs_job &j = r->job;
    const Key key{j.profile.engine, j.profile.k, j.profile.m, j.profile.symbol_len};
    scrs_status s;
    if (j.op == SCRS_JOB_ENCODE) {
        if (j.profile.engine == SCRS_ENGINE_TOWER) {
            scrs_incr_encoder *&e = incrs[key];
            if (!e && (s = scrs_incr_encoder_create(&j.profile, &e)).code != SCRS_OK) return s;
            return scrs_incr_encode_batch_dev(e, j.src, j.dst, 1, stream);
        }
        scrs_batch_encoder *&e = encs[key];
        if (!e && (s = scrs_batch_encoder_create(&j.profile, &e)).code != SCRS_OK) return s;
        return scrs_encode_batch_dev(e, j.src, j.dst, 1, stream);
    }
    scrs_batch_decoder *&d = decs[key];
    if (!d && (s = scrs_batch_decoder_create(&j.profile, &d)).code != SCRS_OK) return s;
    const uint64_t bytes = j.profile.k * 8;
    if (bytes > ring_cap) return st(SCRS_ERR_CAPACITY, ring_cap);
    // Wrap, then wait until completed jobs have released enough of the ring.
    uint64_t start = ring_head % ring_cap;
    if (start + bytes > ring_cap) ring_head += ring_cap - start, start = 0;
    while (ring_head + bytes - ring_tail > ring_cap && !inflight.empty()) {
        InFlight &f = inflight.front();
        (void)hipEventSynchronize(f.ev);
        if (f.ring_end) ring_tail = f.ring_end;  // encodes hold no ring space
        events.push_back(f.ev);
        finish(f, scrs_now_ns());
        inflight.pop_front();
    }
    std::memcpy(ring_host.as<uint8_t>() + start, r->pay.data(), bytes);
    hipError_t err = hipMemcpyAsync(ring.as<uint8_t>() + start, ring_host.as<uint8_t>() + start, bytes,
                                    hipMemcpyHostToDevice, stream);
    if (err != hipSuccess) return st(SCRS_ERR_DEVICE, err);
    ring_head += bytes;
    *ring_end = ring_head;
    return scrs_decode_batch_ptrs_dev(d, r->idx.data(),
                                      reinterpret_cast<const uint8_t *const *>(ring.as<uint8_t>() + start), j.dst, 1,
                                      stream);
}

void scrs_server::loop() {
    DeviceGuard g(device);
    const uint64_t max_jobs = cfg.max_batch_jobs ? cfg.max_batch_jobs : 4096;
    const uint64_t budget = cfg.hbm_budget ? cfg.hbm_budget : (8ull << 30);
    std::vector<scrs_job> jobs;
    while (true) {
        bool progressed = false;
        // Completions.  Plain mode runs on one stream, so its work finishes
        // in launch order (the pointer ring relies on it); batches on
        // several streams may finish in any order.
        if (cfg.batched) {
            for (auto it = inflight.begin(); it != inflight.end();) {
                if (hipEventQuery(it->ev) != hipSuccess) {
                    ++it;
                    continue;
                }
                events.push_back(it->ev);
                finish(*it, scrs_now_ns());
                it = inflight.erase(it);
                progressed = true;
            }
        }
        while (!cfg.batched && !inflight.empty() && hipEventQuery(inflight.front().ev) == hipSuccess) {
            InFlight &f = inflight.front();
            ring_tail = f.ring_end ? f.ring_end : ring_tail;
            events.push_back(f.ev);
            finish(f, scrs_now_ns());
            inflight.pop_front();
            progressed = true;
        }
        // Admission: arrival order, while the footprint fits (one job is
        // always admitted when nothing is in flight).
        std::vector<Request *> batch;
        uint64_t bytes = 0;
        const size_t max_inflight = cfg.batched ? kBatchesInFlight : 1u << 20;
        if (inflight.size() < max_inflight) {
            std::lock_guard<std::mutex> lock(mu);
            while (!pending.empty() && batch.size() < (cfg.batched ? max_jobs : 1)) {
                Request *r = pending.front();
                if (inflight_bytes + bytes + r->bytes > budget && (inflight_bytes + bytes) > 0) break;
                pending.pop_front();
                batch.push_back(r);
                bytes += r->bytes;
            }
            inflight_bytes += bytes;
        }
        if (!batch.empty()) {
            progressed = true;
            const uint64_t t = scrs_now_ns();
            for (Request *r : batch) r->t_start = t;
            InFlight f;
            f.reqs = batch;
            f.bytes = bytes;
            f.start = take_event();
            hipStream_t s = stream;
            if (cfg.batched) {
                s = streams[next_stream];
                next_stream = (next_stream + 1) % kBatchStreams;
            }
            (void)hipEventRecord(f.start, s);
            if (cfg.batched) {
                jobs.clear();
                for (Request *r : batch) jobs.push_back(r->job);
                uint64_t failed = 0;
                f.status = scrs_batch_engine_submit(eng, jobs.data(), jobs.size(), s, &failed);
            } else {
                f.status = plain_one(batch.front(), &f.ring_end);
            }
            f.ev = take_event();
            (void)hipEventRecord(f.ev, s);
            const uint64_t t1 = scrs_now_ns();
            std::lock_guard<std::mutex> lock(mu);
            counters.batches++;
            counters.jobs += batch.size();
            counters.host_ns += t1 - t;
            counters.max_batch_jobs = std::max<uint64_t>(counters.max_batch_jobs, batch.size());
            inflight.push_back(std::move(f));
        }
        if (progressed) continue;
        if (inflight.empty()) {
            std::unique_lock<std::mutex> lock(mu);
            if (stop && pending.empty()) break;
            cv_submit.wait_for(lock, std::chrono::milliseconds(2), [&] { return stop.load() || !pending.empty(); });
            if (stop && pending.empty()) break;
        } else {
            std::this_thread::yield();
        }
    }
}

static scrs_status validate_job(const scrs_job &j) {
    scrs_profile p;
    scrs_status s = scrs_profile_resolve(j.profile.engine, j.profile.k, j.profile.m, j.profile.symbol_len, &p);
    if (s.code != SCRS_OK) return s;
    if (p.engine == SCRS_ENGINE_AFFT) return st(SCRS_CONFIG_UNSUPPORTED_MODE, p.engine);
    if (j.op == SCRS_JOB_ENCODE) return (j.src && j.dst) ? ok() : st(SCRS_ERR_INVALID_ARGUMENT, 2);
    if (j.op != SCRS_JOB_DECODE) return st(SCRS_ERR_INVALID_ARGUMENT, 2);
    if (!j.indices || !j.payloads || !j.dst) return st(SCRS_ERR_INVALID_ARGUMENT, 2);
    const uint64_t k = p.k, n = p.k + p.m;
    std::vector<uint64_t> seen((n + 63) / 64, 0);
    for (uint64_t t = 0; t < k; t++) {
        const uint64_t i = j.indices[t];
        if (i >= n) return st(SCRS_DECODE_INDEX_OUT_OF_RANGE, i, n);
        if (seen[i >> 6] & (1ull << (i & 63))) return st(SCRS_DECODE_DUPLICATE_INDEX, i);
        seen[i >> 6] |= 1ull << (i & 63);
        if (!j.payloads[t]) return st(SCRS_ERR_INVALID_ARGUMENT, 2);
    }
    return ok();
}

extern "C" scrs_status scrs_server_create(const scrs_server_config *cfg, scrs_server **out) {
    if (!out) return st(SCRS_ERR_INVALID_ARGUMENT, 2);
    *out = nullptr;
    auto *s = new scrs_server();
    if (cfg) s->cfg = *cfg;
    (void)hipGetDevice(&s->device);
    hipError_t err = hipStreamCreateWithFlags(&s->stream, hipStreamNonBlocking);
    s->streams[0] = s->stream;
    for (int i = 1; i < scrs_server::kBatchStreams && err == hipSuccess && s->cfg.batched; i++)
        err = hipStreamCreateWithFlags(&s->streams[i], hipStreamNonBlocking);
    scrs_status r = ok();
    if (err == hipSuccess && s->cfg.batched) r = scrs_batch_engine_create(&s->cfg.batch, &s->eng);
    if (err == hipSuccess && !s->cfg.batched) {
        s->ring_cap = 32ull << 20;
        if ((err = s->ring.ensure(s->ring_cap)) == hipSuccess) err = s->ring_host.ensure(s->ring_cap);
    }
    if (err != hipSuccess || r.code != SCRS_OK) {
        if (s->eng) scrs_batch_engine_destroy(s->eng);
        for (int i = 1; i < scrs_server::kBatchStreams; i++)
            if (s->streams[i]) (void)hipStreamDestroy(s->streams[i]);
        if (s->stream) (void)hipStreamDestroy(s->stream);
        delete s;
        return err != hipSuccess ? st(SCRS_ERR_DEVICE, err) : r;
    }
    s->th = std::thread([s] { s->loop(); });
    *out = s;
    return ok();
}

extern "C" void scrs_server_destroy(scrs_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lock(s->mu);
        s->stop = true;
    }
    s->cv_submit.notify_all();
    if (s->th.joinable()) s->th.join();
    DeviceGuard g(s->device);
    for (InFlight &f : s->inflight) {
        (void)hipEventSynchronize(f.ev);
        s->events.push_back(f.ev);
        for (Request *r : f.reqs) delete r;
    }
    for (Request *r : s->pending) delete r;
    for (int i = 0; i < scrs_server::kBatchStreams; i++)
        if (s->streams[i]) (void)hipStreamSynchronize(s->streams[i]);
    for (hipEvent_t e : s->events) (void)hipEventDestroy(e);
    for (auto &kv : s->encs) scrs_batch_encoder_destroy(kv.second);
    for (auto &kv : s->incrs) scrs_incr_encoder_destroy(kv.second);
    for (auto &kv : s->decs) scrs_batch_decoder_destroy(kv.second);
    if (s->eng) scrs_batch_engine_destroy(s->eng);
    for (int i = 1; i < scrs_server::kBatchStreams; i++)
        if (s->streams[i]) (void)hipStreamDestroy(s->streams[i]);
    (void)hipStreamDestroy(s->stream);
    delete s;
}

extern "C" scrs_status scrs_server_submit(scrs_server *s, const scrs_job *job, uint64_t tag) {
    if (!s) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    if (!job) return st(SCRS_ERR_INVALID_ARGUMENT, 2);
    scrs_status v = validate_job(*job);
    if (v.code != SCRS_OK) return v;
    auto *r = new Request();
    r->job = *job;
    if (job->op == SCRS_JOB_DECODE) {
        const uint64_t k = job->profile.k;
        r->idx.assign(job->indices, job->indices + k);
        r->pay.assign(job->payloads, job->payloads + k);
        r->job.indices = r->idx.data();
        r->job.payloads = r->pay.data();
    }
    r->tag = tag;
    r->bytes = footprint(*job);
    r->t_submit = scrs_now_ns();
    {
        std::lock_guard<std::mutex> lock(s->mu);
        s->pending.push_back(r);
        s->outstanding++;
    }
    s->cv_submit.notify_one();
    return ok();
}

extern "C" scrs_status scrs_server_poll(scrs_server *s, scrs_completion *out, uint64_t cap, uint64_t *n,
                                        int32_t wait) {
    if (!s) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    if (!n || (!out && cap)) return st(SCRS_ERR_INVALID_ARGUMENT, 2);
    std::unique_lock<std::mutex> lock(s->mu);
    if (wait) s->cv_done.wait(lock, [&] { return !s->done.empty() || s->outstanding == 0; });
    uint64_t c = 0;
    while (c < cap && !s->done.empty()) {
        out[c++] = s->done.front();
        s->done.pop_front();
    }
    s->outstanding -= c;
    *n = c;
    return ok();
}

extern "C" scrs_status scrs_server_replay(scrs_server *s, const scrs_job *jobs, const uint64_t *arrival_ns,
                                          uint64_t count, scrs_completion *completions) {
    if (!s) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    if (count && (!jobs || !arrival_ns || !completions)) return st(SCRS_ERR_INVALID_ARGUMENT, 2);
    for (uint64_t i = 0; i < count; i++) {
        scrs_status v = validate_job(jobs[i]);
        if (v.code != SCRS_OK) return v;
    }
    std::vector<uint8_t> seen(count, 0);
    uint64_t got = 0;
    std::vector<scrs_completion> buf(1024);
    auto drain = [&](int32_t wait) {
        uint64_t n = 0;
        (void)scrs_server_poll(s, buf.data(), buf.size(), &n, wait);
        for (uint64_t i = 0; i < n; i++)
            if (buf[i].tag < count && !seen[buf[i].tag]) {
                seen[buf[i].tag] = 1;
                completions[buf[i].tag] = buf[i];
                got++;
            }
    };
    const uint64_t t0 = scrs_now_ns();
    for (uint64_t i = 0; i < count; i++) {
        while (scrs_now_ns() - t0 < arrival_ns[i]) {
        }
        scrs_status r = scrs_server_submit(s, &jobs[i], i);
        if (r.code != SCRS_OK) return r;
        if ((i & 63) == 63) drain(0);
    }
    while (got < count) drain(1);
    return ok();
}

extern "C" scrs_status scrs_server_get_stats(scrs_server *s, scrs_server_stats *out) {
    if (!s) return st(SCRS_ERR_INVALID_ARGUMENT, 1);
    if (!out) return st(SCRS_ERR_INVALID_ARGUMENT, 2);
    std::lock_guard<std::mutex> lock(s->mu);
    *out = s->counters;
    return ok();
}
