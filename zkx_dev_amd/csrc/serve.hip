// Continuous-batching dwell server over the batch engine.
//
// Requests name a dwell resident in a device input pool (offset, P, R,
// template).  A scheduler thread admits pending requests in arrival order
// while the HBM budget allows (13 bytes per sample: the 8-byte
// pulse-compressed workspace plus the 4-byte threshold and 1-byte detection
// outputs) and launches them as one engine run on the first free lane; each
// lane has its own stream, so batches formed while others are still on the
// GPU run concurrently.  Requests that arrive while every lane is busy wait
// and join the next batch.  Completion is detected by polling the lane's
// event; the latency of a request is completion time minus submit time.
//
// Outputs live in a device arena of budget/13 cells (threshold float +
// detections uint8 at the same cell offset).  With keep_outputs the caller
// reads them through zkx_server_result and frees them with
// zkx_server_release; otherwise they are freed as soon as the batch ends.
// No reference counterpart: the reference's Task1 chain is one dwell per
// call (Task1/pipeline/src/task1_runner.cu:224-301).
#include "runtime.hpp"

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <new>
#include <thread>
#include <unordered_map>
#include <vector>

namespace {

using Clock = std::chrono::steady_clock;

struct Request {
    long long id;
    long long input_offset;
    int P, R, tmpl;
    long long cells;
    long long out_off = -1;
    Clock::time_point submitted;
    Clock::time_point done;
    int state = 0;  // 0 queued, 1 running, 2 done, 3 failed
    int error = 0;
};

struct InFlight {
    int lane;
    hipEvent_t ev;
    std::vector<long long> ids;
    bool busy = false;
};

// First-fit range allocator over [0, cap) cells with coalescing.
class Ranges {
public:
    explicit Ranges(long long cap) { free_[0] = cap; }
    long long take(long long n)
    {
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= n) {
                const long long off = it->first, len = it->second;
                free_.erase(it);
                if (len > n) free_[off + n] = len - n;
                return off;
            }
        }
        return -1;
    }
    void give(long long off, long long n)
    {
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }

private:
    std::map<long long, long long> free_;
};

}  // namespace

struct zkx_server {
    zkx_engine* engine = nullptr;
    int device = 0;
    const float2* pool = nullptr;
    long long pool_elems = 0;
    long long budget_bytes = 0;
    long long max_cells = 0;
    int max_batch = 64;
    bool keep = false;
    float* thr = nullptr;
    unsigned char* det = nullptr;
    std::vector<hipStream_t> streams;
    std::vector<InFlight> lanes;
    Ranges* ranges = nullptr;
    long long reserved_bytes = 0, peak_reserved = 0;
    long long next_id = 1;
    long long batches = 0, dwells_done = 0;
    std::deque<long long> queue;
    std::unordered_map<long long, Request> reqs;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::atomic<bool> stop{false};
    bool hold = false;
    std::thread worker;
    int fatal = 0;
};

namespace {

void finish_lane(zkx_server& s, InFlight& f, int status)
{
    const Clock::time_point now = Clock::now();
    for (long long id : f.ids) {
        auto it = s.reqs.find(id);
        if (it == s.reqs.end()) continue;
        Request& r = it->second;
        r.done = now;
        r.state = status ? 3 : 2;
        r.error = status;
        ++s.dwells_done;
        if (!s.keep || status) {
            s.ranges->give(r.out_off, r.cells);
            r.out_off = -1;
            s.reserved_bytes -= 13 * r.cells;
        } else {
            // outputs stay until release; the 8-byte workspace share is free now
            s.reserved_bytes -= 8 * r.cells;
        }
    }
    f.ids.clear();
    f.busy = false;
}

void worker_loop(zkx_server* sp)
{
    zkx_server& s = *sp;
    (void)hipSetDevice(s.device);
    std::vector<zkx_dwell_desc> descs;
    std::vector<long long> ids;
    std::unique_lock<std::mutex> lk(s.mu);
    while (!s.stop.load()) {
        bool progressed = false;
        // 1. retire finished batches
        for (InFlight& f : s.lanes) {
            if (!f.busy) continue;
            const hipError_t q = hipEventQuery(f.ev);
            if (q == hipErrorNotReady) continue;
            finish_lane(s, f, q == hipSuccess ? 0 : zkx::hip_status_to_code(q));
            progressed = true;
        }
        if (progressed) s.cv_done.notify_all();
        // 2. launch a batch on a free lane
        InFlight* free_lane = nullptr;
        for (InFlight& f : s.lanes)
            if (!f.busy) {
                free_lane = &f;
                break;
            }
        if (free_lane && !s.queue.empty() && !s.hold) {
            descs.clear();
            ids.clear();
            while (!s.queue.empty() && static_cast<int>(ids.size()) < s.max_batch) {
                Request& r = s.reqs[s.queue.front()];
                const long long need = 13 * r.cells;
                if (s.reserved_bytes + need > s.budget_bytes) break;  // strict arrival order
                const long long off = s.ranges->take(r.cells);
                if (off < 0) break;
                r.out_off = off;
                s.reserved_bytes += need;
                s.peak_reserved = std::max(s.peak_reserved, s.reserved_bytes);
                r.state = 1;
                zkx_dwell_desc d;
                d.num_pulses = r.P;
                d.range_bins = r.R;
                d.template_id = r.tmpl;
                d.reserved = 0;
                d.input_offset = r.input_offset;
                d.output_offset = off;
                descs.push_back(d);
                ids.push_back(r.id);
                s.queue.pop_front();
            }
            if (!ids.empty()) {
                InFlight& f = *free_lane;
                f.ids = ids;
                f.busy = true;
                ++s.batches;
                lk.unlock();
                int rc = zkx_engine_run(s.engine, f.lane, reinterpret_cast<const zkx_complex64*>(s.pool),
                                        descs.data(), static_cast<int>(descs.size()), s.thr, s.det, nullptr,
                                        s.streams[f.lane]);
                if (!rc) {
                    const hipError_t e = hipEventRecord(f.ev, s.streams[f.lane]);
                    if (e != hipSuccess) rc = zkx::hip_status_to_code(e);
                }
                lk.lock();
                if (rc) {
                    finish_lane(s, f, rc);
                    s.cv_done.notify_all();
                }
                continue;
            }
        }
        // 3. wait: briefly while batches run (poll their events), else for work
        bool any_busy = false;
        for (InFlight& f : s.lanes) any_busy |= f.busy;
        if (any_busy) {
            lk.unlock();
            std::this_thread::yield();
            lk.lock();
        } else {
            s.cv_work.wait_for(lk, std::chrono::milliseconds(5));
        }
    }
    // drain
    for (InFlight& f : s.lanes)
        if (f.busy) {
            (void)hipEventSynchronize(f.ev);
            finish_lane(s, f, 0);
        }
    s.cv_done.notify_all();
}

void destroy_server(zkx_server* s)
{
    if (!s) return;
    if (s->worker.joinable()) {
        s->stop.store(true);
        s->cv_work.notify_all();
        s->worker.join();
    }
    (void)hipSetDevice(s->device);
    for (InFlight& f : s->lanes)
        if (f.ev) (void)hipEventDestroy(f.ev);
    for (hipStream_t st : s->streams)
        if (st) (void)hipStreamDestroy(st);
    if (s->thr) (void)hipFree(s->thr);
    if (s->det) (void)hipFree(s->det);
    delete s->ranges;
    delete s;
}

}  // namespace

using zkx::set_last_error;

extern "C" {

int zkx_server_create(zkx_server** out, zkx_engine* engine, const zkx_complex64* input_pool, long long pool_elems,
                      long long hbm_budget_bytes, int lanes, int max_batch_dwells, int keep_outputs)
{
    zkx::clear_last_error();
    if (!out || !engine || !input_pool) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "server needs an engine and an input pool");
    *out = nullptr;
    if (pool_elems < 1 || hbm_budget_bytes < 13 * 16 || lanes < 1 || max_batch_dwells < 1)
        ZKX_FAIL(ZKX_INVALID_ARGUMENT, "server pool, budget, lanes and batch size must be positive");
    zkx_server* s = new (std::nothrow) zkx_server();
    if (!s) ZKX_FAIL(ZKX_OUT_OF_MEMORY, "server allocation failed");
    s->engine = engine;
    s->pool = reinterpret_cast<const float2*>(input_pool);
    s->pool_elems = pool_elems;
    s->budget_bytes = hbm_budget_bytes;
    s->max_cells = hbm_budget_bytes / 13;
    s->max_batch = max_batch_dwells;
    s->keep = keep_outputs != 0;
    int rc = ZKX_SUCCESS;
    hipError_t err = hipGetDevice(&s->device);
    if (err == hipSuccess) err = hipMalloc(reinterpret_cast<void**>(&s->thr), sizeof(float) * s->max_cells);
    if (err == hipSuccess) err = hipMalloc(reinterpret_cast<void**>(&s->det), s->max_cells);
    s->streams.assign(lanes, nullptr);
    s->lanes.resize(lanes);
    for (int i = 0; i < lanes && err == hipSuccess; ++i) {
        err = hipStreamCreateWithFlags(&s->streams[i], hipStreamNonBlocking);
        if (err == hipSuccess) err = hipEventCreateWithFlags(&s->lanes[i].ev, hipEventDisableTiming);
        s->lanes[i].lane = i;
    }
    if (err != hipSuccess) {
        rc = zkx::hip_status_to_code(err);
        set_last_error(rc, hipGetErrorString(err));
        destroy_server(s);
        return rc;
    }
    s->ranges = new (std::nothrow) Ranges(s->max_cells);
    if (!s->ranges) {
        destroy_server(s);
        ZKX_FAIL(ZKX_OUT_OF_MEMORY, "server allocation failed");
    }
    try {
        s->worker = std::thread(worker_loop, s);
    } catch (...) {
        destroy_server(s);
        ZKX_FAIL(ZKX_RUNTIME_ERROR, "server thread could not start");
    }
    *out = s;
    return ZKX_SUCCESS;
}

int zkx_server_destroy(zkx_server* s)
{
    zkx::clear_last_error();
    if (!s) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "server handle is null");
    destroy_server(s);
    return ZKX_SUCCESS;
}

int zkx_server_submit(zkx_server* s, long long input_offset, int num_pulses, int range_bins, int template_id,
                      long long* ticket)
{
    zkx::clear_last_error();
    if (!s || !ticket) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "server and ticket must be non-null");
    if (num_pulses < 16 || num_pulses > 256 || (num_pulses & (num_pulses - 1)) || range_bins < 1)
        ZKX_FAIL(ZKX_INVALID_SHAPE, "request shape: P a power of two in [16, 256], R positive");
    const long long cells = static_cast<long long>(num_pulses) * range_bins;
    if (input_offset < 0 || input_offset + cells > s->pool_elems)
        ZKX_FAIL(ZKX_INVALID_ARGUMENT, "request lies outside the input pool");
    if (13 * cells > s->budget_bytes) ZKX_FAIL(ZKX_OUT_OF_MEMORY, "request larger than the HBM budget");
    {
        // shape and template checks up front, so one bad request never fails a batch
        zkx_dwell_desc d{num_pulses, range_bins, template_id, 0, input_offset, 0};
        int nfft = 0, segs = 0;
        const int rc = zkx_engine_plan(s->engine, &d, 1, &nfft, &segs);
        if (rc) return rc;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    if (s->fatal) ZKX_FAIL(s->fatal, "server stopped after a device error");
    Request r;
    r.id = s->next_id++;
    r.input_offset = input_offset;
    r.P = num_pulses;
    r.R = range_bins;
    r.tmpl = template_id;
    r.cells = cells;
    r.submitted = Clock::now();
    s->reqs.emplace(r.id, r);
    s->queue.push_back(r.id);
    *ticket = r.id;
    s->cv_work.notify_one();
    return ZKX_SUCCESS;
}

// Wait until the request is done (timeout_ms < 0: forever).  Returns
// ZKX_SUCCESS with *done set, or the request's error.
int zkx_server_wait(zkx_server* s, long long ticket, int timeout_ms, int* done)
{
    zkx::clear_last_error();
    if (!s || !done) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "unknown ticket");
    auto pred = [&] { return it->second.state >= 2; };
    if (timeout_ms < 0) s->cv_done.wait(lk, pred);
    else s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), pred);
    *done = it->second.state >= 2;
    if (it->second.state == 3) ZKX_FAIL(it->second.error, "request failed on the device");
    return ZKX_SUCCESS;
}

// Latency (us, submit to completion) and, with keep_outputs, device pointers
// to the request's [P][R] threshold and detections.
int zkx_server_result(zkx_server* s, long long ticket, double* latency_us, const float** threshold,
                      const uint8_t** detections)
{
    zkx::clear_last_error();
    if (!s) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "server handle is null");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "unknown ticket");
    const Request& r = it->second;
    if (r.state < 2) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "request not finished");
    if (r.state == 3) ZKX_FAIL(r.error, "request failed on the device");
    if (latency_us) latency_us[0] = std::chrono::duration<double, std::micro>(r.done - r.submitted).count();
    if (threshold) *threshold = r.out_off >= 0 ? s->thr + r.out_off : nullptr;
    if (detections) *detections = r.out_off >= 0 ? s->det + r.out_off : nullptr;
    return ZKX_SUCCESS;
}

// Copy a kept request's [P][R] threshold and detections into caller buffers
// (device-to-device on `stream`).
int zkx_server_copy_result(zkx_server* s, long long ticket, float* threshold, uint8_t* detections, void* stream)
{
    zkx::clear_last_error();
    if (!s || !threshold || !detections) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "null argument");
    long long off = -1, cells = 0;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->reqs.find(ticket);
        if (it == s->reqs.end()) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "unknown ticket");
        if (it->second.state != 2 || it->second.out_off < 0)
            ZKX_FAIL(ZKX_INVALID_ARGUMENT, "request has no kept outputs");
        off = it->second.out_off;
        cells = it->second.cells;
    }
    hipStream_t st = static_cast<hipStream_t>(stream);
    ZKX_HIP_CHECK(hipMemcpyAsync(threshold, s->thr + off, sizeof(float) * cells, hipMemcpyDeviceToDevice, st));
    ZKX_HIP_CHECK(hipMemcpyAsync(detections, s->det + off, cells, hipMemcpyDeviceToDevice, st));
    return ZKX_SUCCESS;
}

int zkx_server_release(zkx_server* s, long long ticket)
{
    zkx::clear_last_error();
    if (!s) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "server handle is null");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(ticket);
    if (it == s->reqs.end()) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "unknown ticket");
    if (it->second.state < 2) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "request not finished");
    if (it->second.out_off >= 0) {
        s->ranges->give(it->second.out_off, it->second.cells);
        s->reserved_bytes -= 5 * it->second.cells;
    }
    s->reqs.erase(it);
    s->cv_work.notify_one();
    return ZKX_SUCCESS;
}

// Hold (1) or resume (0) batch formation; queued requests wait while held.
int zkx_server_hold(zkx_server* s, int hold)
{
    zkx::clear_last_error();
    if (!s) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "server handle is null");
    std::lock_guard<std::mutex> lk(s->mu);
    s->hold = hold != 0;
    s->cv_work.notify_one();
    return ZKX_SUCCESS;
}

// stats: [0] batches launched, [1] dwells finished, [2] queued now,
// [3] reserved bytes now, [4] peak reserved bytes
int zkx_server_stats(zkx_server* s, long long* stats)
{
    zkx::clear_last_error();
    if (!s || !stats) ZKX_FAIL(ZKX_INVALID_ARGUMENT, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    stats[0] = s->batches;
    stats[1] = s->dwells_done;
    stats[2] = static_cast<long long>(s->queue.size());
    stats[3] = s->reserved_bytes;
    stats[4] = s->peak_reserved;
    return ZKX_SUCCESS;
}

// Open-loop load: submit request i at arrival_us[i] (microseconds after the
// call starts) from this thread, then wait for all; latency_us[i] receives
// its submit-to-completion time.  Requests are released as they finish.
int zkx_server_replay(zkx_server* s, const double* arrival_us, const long long* input_offset,
                      const int* num_pulses, const int* range_bins, const int* template_id, int n,
                      double* latency_us)
{
    zkx::clear_last_error();
    if (!s || n < 0 || (n > 0 && (!arrival_us || !input_offset || !num_pulses || !range_bins || !template_id ||
                                  !latency_us)))
        ZKX_FAIL(ZKX_INVALID_ARGUMENT, "null argument");
    std::vector<long long> tickets(n, 0);
    const Clock::time_point t0 = Clock::now();
    for (int i = 0; i < n; ++i) {
        const Clock::time_point due = t0 + std::chrono::nanoseconds(static_cast<long long>(arrival_us[i] * 1e3));
        while (Clock::now() < due) {
        }
        const int rc = zkx_server_submit(s, input_offset[i], num_pulses[i], range_bins[i], template_id[i], &tickets[i]);
        if (rc) return rc;
    }
    for (int i = 0; i < n; ++i) {
        int done = 0;
        int rc = zkx_server_wait(s, tickets[i], -1, &done);
        if (!rc) rc = zkx_server_result(s, tickets[i], &latency_us[i], nullptr, nullptr);
        if (rc) return rc;
        if (s->keep && (rc = zkx_server_release(s, tickets[i]))) return rc;
        if (!s->keep) {
            std::lock_guard<std::mutex> lk(s->mu);
            s->reqs.erase(tickets[i]);
        }
    }
    return ZKX_SUCCESS;
}

}  // extern "C"
