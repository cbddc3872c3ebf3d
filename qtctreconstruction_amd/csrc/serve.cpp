// serve.cpp — serving form: a request loop that reconstructs volumes from
// device-resident sinograms with continuous batching.
//
//   * Admission by HBM budget: the server owns one device pool (a first-fit
//     sub-allocator over a single allocation); a job is admitted when its
//     filtered-sinogram workspace and its output fit. The first job in
//     admission order that does not fit stops admission, so a large job is
//     not passed over once it reaches the head.
//   * Admission order (ct_hip_server_set_policy): FIFO, or by estimated
//     device time minus a credit for time waited (the default). A batch
//     stops growing at a cap on its estimated device time (or a quarter of
//     the queued work, if more): the jobs of a batch complete together, so a
//     small job batched with a large one would wait for it. Ordering by size lets the many small jobs pass the few
//     large ones; the waiting credit bounds how long a large job can be
//     passed over. The device rate behind the estimates is measured: updates
//     completed per millisecond of time with a batch in flight.
//   * Continuous batching: up to `slots` batches are in flight at once, each
//     on its own stream. Jobs that arrive while batches run are admitted into
//     the next free slot as soon as memory frees up, so their persistent
//     launches overlap with the tail of the running ones on the device.
//   * Device-resident state: sinograms stay where the client put them, filter
//     banks / trig tables are cached in the server's context, outputs stay in
//     the pool until the client releases them.
//   * Each job is reconstructed with the same kernels and the same per-voxel
//     summation order as alone, so its result does not depend on the mix.
//
// Latency = submit -> completion observed by the worker (host clock).
#include "ctx.h"

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <thread>

using namespace cthip;
using namespace cthost;

namespace {

double now_ms() {
    using namespace std::chrono;
    return duration<double, std::milli>(steady_clock::now().time_since_epoch()).count();
}

// First-fit sub-allocator with coalescing over [0, size).
class Pool {
public:
    void init(size_t size) {
        free_.clear();
        free_[0] = size;
        size_ = size;
        used_ = 0;
    }
    bool alloc(size_t bytes, size_t* off) {
        bytes = (bytes + 255) & ~size_t(255);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second >= bytes) {
                *off = it->first;
                const size_t rest = it->second - bytes;
                const size_t start = it->first + bytes;
                free_.erase(it);
                if (rest) free_[start] = rest;
                used_ += bytes;
                sizes_[*off] = bytes;
                return true;
            }
        }
        return false;
    }
    void release(size_t off) {
        auto sz = sizes_.find(off);
        if (sz == sizes_.end()) return;
        size_t start = off, len = sz->second;
        used_ -= len;
        sizes_.erase(sz);
        auto next = free_.lower_bound(start);
        if (next != free_.end() && start + len == next->first) {
            len += next->second;
            next = free_.erase(next);
        }
        if (next != free_.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == start) {
                start = prev->first;
                len += prev->second;
                free_.erase(prev);
            }
        }
        free_[start] = len;
    }
    static size_t rounded(size_t b) { return (b + 255) & ~size_t(255); }
    size_t size() const { return size_; }
    size_t used() const { return used_; }

private:
    std::map<size_t, size_t> free_;
    std::map<size_t, size_t> sizes_;
    size_t size_ = 0, used_ = 0;
};

struct Job {
    uint64_t ticket = 0;
    ct_fbp_job d{};
    std::vector<float> angles;
    size_t out_off = 0, filt_off = 0, out_bytes = 0, filt_bytes = 0;
    double updates = 0;  // nz * bins^2 * na voxel-updates
    double t_submit = 0, t_admit = 0, t_done = 0;
    int state = 0;  // 0 queued, 1 running, 2 done, 3 failed
    int slot = -1;
    int readers = 0;  // downloads in progress (the output may not be released meanwhile)
};

struct Slot {
    BatchSlot bs;
    std::vector<Job*> jobs;
    bool busy = false;
};

}  // namespace

struct ct_hip_server {
    ct_hip_ctx* ctx = nullptr;
    char* base = nullptr;  // pool memory
    Pool pool;
    int max_batch = 16;
    int flags = CT_BATCH_GRAPH | CT_BATCH_AUTOTUNE;
    std::vector<Slot> slots;
    // admission policy (ct_hip_server_set_policy)
    int policy = CT_SERVE_SIZE_AGED;
    double batch_cap_ms = 5.0, age_weight = 0.25;
    // device rate estimate: updates completed per ms with a batch in flight
    double busy_ms = 0, busy_start = -1, done_updates = 0;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Job*> queue;
    std::deque<Job*> done;
    std::unordered_map<uint64_t, Job*> jobs;
    uint64_t next_ticket = 1;
    bool stop = false;
    std::thread worker;
    std::string err;
    // downloads of finished outputs: pinned staging on a stream of their own,
    // so a client's download neither takes the context's lock nor waits
    // behind the worker's launches
    std::unique_ptr<cthost::Stager> dl;
    // statistics
    uint64_t batches = 0, completed = 0, max_inflight_jobs = 0, overlapped_launches = 0;
    double busy_since = -1, first_submit = -1, last_done = 0;
};

namespace {

void fail_job(ct_hip_server* s, Job* j, const std::string& msg) {
    j->state = 3;
    j->t_done = now_ms();
    s->err = msg;
    s->done.push_back(j);
}

// Updates per ms: measured once 20 ms of busy time have been seen, else a
// figure typical of a full device (5e12 voxel-updates/s).
double rate_per_ms(const ct_hip_server* s) {
    return s->busy_ms > 20.0 ? s->done_updates / s->busy_ms : 5e9;
}

// Puts the queue in admission order (stable, so FIFO among equal keys).
void order_queue(ct_hip_server* s) {
    if (s->policy != CT_SERVE_SIZE_AGED || s->queue.size() < 2) return;
    const double t = now_ms(), r = rate_per_ms(s), a = s->age_weight;
    std::vector<std::pair<double, Job*>> key;
    key.reserve(s->queue.size());
    for (Job* j : s->queue) key.emplace_back(j->updates / r - a * (t - j->t_submit), j);
    std::stable_sort(key.begin(), key.end(),
                     [](const std::pair<double, Job*>& x, const std::pair<double, Job*>& y) {
                         return x.first < y.first;
                     });
    for (size_t i = 0; i < key.size(); ++i) s->queue[i] = key[i].second;
}

// Called with s->mu held. Admits queued jobs, in admission order, into slot k
// and launches them as one batch.
bool launch_into(ct_hip_server* s, int k, std::unique_lock<std::mutex>& lk) {
    Slot& sl = s->slots[k];
    std::vector<Job*> batch;
    order_queue(s);
    // the cap grows with the backlog (a quarter of the queued work), so a long
    // queue drains in large batches, which keep the device fuller
    double queued = 0;
    for (const Job* j : s->queue) queued += j->updates;
    const double cap =
        s->batch_cap_ms > 0 ? std::max(s->batch_cap_ms * rate_per_ms(s), 0.25 * queued) : 0.0;
    double work = 0;
    while (!s->queue.empty() && (int)batch.size() < s->max_batch) {
        Job* j = s->queue.front();
        if (!batch.empty() && cap > 0 && work + j->updates > cap) break;
        size_t oo = 0, fo = 0;
        const size_t need_bytes = Pool::rounded(j->out_bytes) + Pool::rounded(j->filt_bytes);
        if (need_bytes > s->pool.size()) {  // can never fit
            s->queue.pop_front();
            fail_job(s, j, "job needs " + std::to_string(need_bytes) + " B, pool holds " +
                               std::to_string(s->pool.size()) + " B");
            continue;
        }
        if (!s->pool.alloc(j->out_bytes, &oo)) break;
        if (!s->pool.alloc(j->filt_bytes, &fo)) {
            s->pool.release(oo);
            break;
        }
        s->queue.pop_front();
        j->out_off = oo;
        j->filt_off = fo;
        j->d.d_out = reinterpret_cast<float*>(s->base + oo);
        j->d.angles_deg = j->angles.data();
        j->state = 1;
        j->slot = k;
        j->t_admit = now_ms();
        batch.push_back(j);
        work += j->updates;
    }
    if (batch.empty()) return false;
    std::vector<ct_fbp_job> desc(batch.size());
    std::vector<float4*> filts(batch.size());
    for (size_t i = 0; i < batch.size(); ++i) {
        desc[i] = batch[i]->d;
        filts[i] = reinterpret_cast<float4*>(s->base + batch[i]->filt_off);
    }
    int inflight_slots = 0;
    for (auto& o : s->slots) inflight_slots += o.busy ? 1 : 0;
    lk.unlock();
    std::string msg;
    {
        std::lock_guard<std::recursive_mutex> g(s->ctx->mu);
        try {
            bind_device(s->ctx);
            slot_init(s->ctx, &sl.bs, true);
            enqueue_batch(s->ctx, &sl.bs, desc.data(), (int)desc.size(), filts.data(), s->flags);
        } catch (const std::exception& e) {
            msg = e.what();
        }
    }
    lk.lock();
    if (!msg.empty()) {
        for (Job* j : batch) {
            s->pool.release(j->out_off);
            s->pool.release(j->filt_off);
            fail_job(s, j, msg);
        }
        s->cv_done.notify_all();
        return true;
    }
    sl.jobs = std::move(batch);
    if (inflight_slots == 0) s->busy_start = now_ms();
    sl.busy = true;
    s->batches++;
    if (inflight_slots > 0) s->overlapped_launches++;
    uint64_t inflight = 0;
    for (auto& o : s->slots)
        if (o.busy) inflight += o.jobs.size();
    if (inflight > s->max_inflight_jobs) s->max_inflight_jobs = inflight;
    return true;
}

void worker_loop(ct_hip_server* s) {
    (void)hipSetDevice(s->ctx->device);
    std::unique_lock<std::mutex> lk(s->mu);
    while (!s->stop) {
        bool progress = false;
        for (auto& sl : s->slots) {
            if (!sl.busy) continue;
            const hipError_t q = hipEventQuery(sl.bs.ev_done);
            if (q == hipErrorNotReady) continue;
            const double t = now_ms();
            for (Job* j : sl.jobs) {
                s->pool.release(j->filt_off);
                if (q == hipSuccess) {
                    j->state = 2;
                    j->t_done = t;
                    s->done.push_back(j);
                    s->completed++;
                    s->done_updates += j->updates;
                } else {
                    s->pool.release(j->out_off);
                    fail_job(s, j, std::string("batch failed: ") + hipGetErrorString(q));
                }
            }
            s->last_done = t;
            sl.jobs.clear();
            sl.busy = false;
            progress = true;
            bool any = false;
            for (auto& o : s->slots) any |= o.busy;
            if (!any && s->busy_start >= 0) {
                s->busy_ms += t - s->busy_start;
                s->busy_start = -1;
            }
        }
        if (progress) s->cv_done.notify_all();
        for (int k = 0; k < (int)s->slots.size(); ++k)
            if (!s->slots[k].busy && !s->queue.empty()) progress |= launch_into(s, k, lk);
        if (!progress) {
            bool any_busy = false;
            for (auto& sl : s->slots) any_busy |= sl.busy;
            // poll completions every 50 us while work is in flight; otherwise sleep until a submit
            if (any_busy)
                s->cv_work.wait_for(lk, std::chrono::microseconds(50));
            else
                s->cv_work.wait_for(lk, std::chrono::milliseconds(20));
        }
    }
}

}  // namespace

extern "C" {

int ct_hip_server_create(int device, size_t pool_bytes, int max_batch, int slots,
                         ct_hip_server** out) {
    if (!out || pool_bytes == 0 || max_batch <= 0 || max_batch > kMaxBatch || slots <= 0 ||
        slots > 8)
        return CT_ERR_INVALID;
    *out = nullptr;
    ct_hip_ctx* c = nullptr;
    int rc = ct_hip_create(device, 0, &c);
    if (rc != CT_OK) return rc;
    auto* s = new (std::nothrow) ct_hip_server();
    if (!s) {
        ct_hip_destroy(c);
        return CT_ERR_NOMEM;
    }
    s->ctx = c;
    void* p = nullptr;
    rc = ct_hip_alloc(c, pool_bytes, &p);
    if (rc != CT_OK) {
        delete s;
        ct_hip_destroy(c);
        return rc;
    }
    s->base = static_cast<char*>(p);
    s->pool.init(pool_bytes);
    s->max_batch = max_batch;
    s->slots.resize(slots);
    s->dl.reset(new cthost::Stager());
    rc = guarded(c, [&] { s->dl->init(copier(c)); });
    if (rc != CT_OK) {
        s->dl.reset();
        ct_hip_free(c, p);
        delete s;
        ct_hip_destroy(c);
        return rc;
    }
    s->worker = std::thread(worker_loop, s);
    *out = s;
    return CT_OK;
}

void ct_hip_server_destroy(ct_hip_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    {
        std::lock_guard<std::recursive_mutex> g(s->ctx->mu);
        for (auto& sl : s->slots) {
            try {
                slot_release(s->ctx, &sl.bs);
            } catch (...) {
            }
        }
    }
    for (auto& kv : s->jobs) delete kv.second;
    s->dl.reset();  // before the context, whose copy threads it uses
    ct_hip_destroy(s->ctx);
    delete s;
}

int ct_hip_server_set_flags(ct_hip_server* s, int flags) {
    if (!s) return CT_ERR_INVALID;
    std::lock_guard<std::mutex> g(s->mu);
    s->flags = flags;
    return CT_OK;
}

int ct_hip_server_set_filter_backend(ct_hip_server* s, int backend) {
    if (!s) return CT_ERR_INVALID;
    // the worker forms batches under the context's lock, so a change applies
    // to the batches formed after it returns
    return ct_hip_set_filter_backend(s->ctx, backend);
}

int ct_hip_server_set_policy(ct_hip_server* s, int policy, double batch_cap_ms,
                             double age_weight) {
    if (!s || (policy != CT_SERVE_FIFO && policy != CT_SERVE_SIZE_AGED)) return CT_ERR_INVALID;
    if (!(batch_cap_ms >= 0) || !(age_weight >= 0)) return CT_ERR_INVALID;
    std::lock_guard<std::mutex> g(s->mu);
    s->policy = policy;
    s->batch_cap_ms = batch_cap_ms;
    s->age_weight = age_weight;
    return CT_OK;
}

int ct_hip_server_submit(ct_hip_server* s, const ct_fbp_job* job, uint64_t* ticket) {
    if (!s || !job || !ticket || !job->d_sinos || !job->angles_deg) return CT_ERR_INVALID;
    if (job->nz <= 0 || job->na <= 0 || job->bins <= 0) return CT_ERR_INVALID;
    const Geom g = geometry(job->bins);
    if (g.nfft > kMaxFFT) return CT_ERR_UNSUPPORTED;
    Job* j = new (std::nothrow) Job();
    if (!j) return CT_ERR_NOMEM;
    j->d = *job;
    j->angles.assign(job->angles_deg, job->angles_deg + job->na);
    j->out_bytes = sizeof(float) * (size_t)job->nz * job->bins * job->bins;
    j->filt_bytes = job_filt_bytes(*job);
    j->updates = (double)job->nz * job->bins * (double)job->bins * job->na;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        j->ticket = s->next_ticket++;
        j->t_submit = now_ms();
        if (s->first_submit < 0) s->first_submit = j->t_submit;
        s->jobs[j->ticket] = j;
        s->queue.push_back(j);
        *ticket = j->ticket;
    }
    s->cv_work.notify_one();
    return CT_OK;
}

// Waits up to timeout_ms for any finished job. Returns CT_OK with its ticket,
// device output and latency (ms), 1 on timeout, or the job's failure code.
int ct_hip_server_wait(ct_hip_server* s, double timeout_ms, uint64_t* ticket, const float** d_out,
                       double* latency_ms) {
    if (!s || !ticket) return CT_ERR_INVALID;
    std::unique_lock<std::mutex> lk(s->mu);
    // a negative or NaN timeout polls; past about a day it is taken as a day
    const double t_ms = timeout_ms > 0 ? std::min(timeout_ms, 8.64e7) : 0.0;
    const auto deadline =
        std::chrono::steady_clock::now() + std::chrono::microseconds((long long)(t_ms * 1000));
    while (s->done.empty()) {
        if (s->cv_done.wait_until(lk, deadline) == std::cv_status::timeout && s->done.empty()) return 1;
    }
    Job* j = s->done.front();
    s->done.pop_front();
    *ticket = j->ticket;
    if (d_out) *d_out = j->state == 2 ? j->d.d_out : nullptr;
    if (latency_ms) *latency_ms = j->t_done - j->t_submit;
    return j->state == 2 ? CT_OK : CT_ERR_INVALID;
}

// Returns a finished job's output memory to the pool.
int ct_hip_server_release(ct_hip_server* s, uint64_t ticket) {
    if (!s) return CT_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(ticket);
    if (it == s->jobs.end()) return CT_ERR_INVALID;
    Job* j = it->second;
    if (j->state == 1 || j->state == 0) return CT_ERR_INVALID;
    if (j->readers > 0) {  // a download still reads the output
        s->err = "release: ticket " + std::to_string(ticket) + " is being downloaded";
        return CT_ERR_INVALID;
    }
    if (j->state == 2) s->pool.release(j->out_off);
    delete j;
    s->jobs.erase(it);
    s->cv_work.notify_one();
    return CT_OK;
}

int ct_hip_server_download(ct_hip_server* s, uint64_t ticket, float* host_out, size_t bytes) {
    if (!s || !host_out) return CT_ERR_INVALID;
    const float* src = nullptr;
    size_t nb = 0;
    Job* j = nullptr;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->jobs.find(ticket);
        if (it == s->jobs.end() || it->second->state != 2) return CT_ERR_INVALID;
        j = it->second;
        src = j->d.d_out;
        nb = j->out_bytes;
        if (bytes < nb) return CT_ERR_INVALID;
        j->readers++;  // release waits for the copy (it refuses the ticket meanwhile)
    }
    // the job's batch has completed (state 2), so its output is final
    int rc = CT_OK;
    try {
        (void)hipSetDevice(s->ctx->device);
        s->dl->download(host_out, src, nb, nullptr);
    } catch (const StatusError& e) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->err = e.what();
        rc = e.code;
    } catch (const std::exception& e) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->err = e.what();
        rc = CT_ERR_INVALID;
    }
    std::lock_guard<std::mutex> lk(s->mu);
    j->readers--;
    return rc;
}

// stats[0] batches, [1] completed jobs, [2] max jobs in flight,
// [3] launches made while another batch was in flight, [4] pool bytes in use,
// [5] pool size, [6] ms from first submit to last completion
int ct_hip_server_stats(ct_hip_server* s, double* stats, int n) {
    if (!s || !stats) return CT_ERR_INVALID;
    std::lock_guard<std::mutex> lk(s->mu);
    const double v[7] = {(double)s->batches, (double)s->completed, (double)s->max_inflight_jobs,
                         (double)s->overlapped_launches, (double)s->pool.used(),
                         (double)s->pool.size(),
                         s->first_submit < 0 ? 0.0 : s->last_done - s->first_submit};
    for (int i = 0; i < n && i < 7; ++i) stats[i] = v[i];
    return CT_OK;
}

const char* ct_hip_server_last_error(const ct_hip_server* s) { return s ? s->err.c_str() : ""; }

}  // extern "C"
