// Commit server: continuous admission of commit jobs under an HBM budget.
//
// Jobs queue in arrival order. A scheduler thread admits the head of the queue
// as soon as an in-flight slot is free and the arena can hold the job's
// buffers, enqueues its work on the least-loaded of n_streams device streams
// (trace fill or copy -> LDE -> leaves -> tree levels -> group proofs -> root
// and proofs to pinned host memory -> event), and completes jobs as their
// events fire. Later arrivals are admitted while earlier jobs still run, so
// the device overlaps the work of many jobs (continuous batching).
//
// The caller-facing semantics of a job are those of commit_witness followed by
// MerkleTreeGL::getGroupProof for its queries (starks_api.cpp:683-723,
// merkleTreeGL.cpp:145-175); results do not depend on what else is running.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "../../include/pil2_commit_hip.h"
#include "commit_internal.h"
#include "ntt_kernels.h"
#include "poseidon2_kernels.h"

namespace {

const uint64_t kAlign = 4096;
const int64_t kErrBudget = -30, kErrCancelled = -31, kErrNotQueued = -32, kErrUnknownJob = -33,
              kErrNotDone = -34, kErrArena = -35;

uint64_t align_up(uint64_t x) { return (x + kAlign - 1) & ~(kAlign - 1); }

// First-fit allocator over one device allocation; free blocks coalesce.
class Arena {
  public:
    uint64_t size = 0, used = 0, peak = 0;
    char *base = nullptr;
    void init(char *b, uint64_t s)
    {
        base = b;
        size = s;
        free_.clear();
        if (s) free_[0] = s;
    }
    bool alloc(uint64_t bytes, uint64_t *off)
    {
        bytes = align_up(bytes ? bytes : 1);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            *off = it->first;
            uint64_t rest = it->second - bytes, at = it->first + bytes;
            free_.erase(it);
            if (rest) free_[at] = rest;
            used += bytes;
            if (used > peak) peak = used;
            return true;
        }
        return false;
    }
    void release(uint64_t off, uint64_t bytes)
    {
        bytes = align_up(bytes ? bytes : 1);
        used -= bytes;
        auto it = free_.emplace(off, bytes).first;
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
    std::map<uint64_t, uint64_t> free_;
};

enum class St { Queued, Running, Done, Released, Failed, Cancelled };

struct Layout {  // byte sizes of a job's buffers
    uint64_t lde, nodes, trace, scratch, idx, proofs, pw;
    uint64_t resident() const { return align_up(lde) + align_up(nodes); }
    uint64_t transient() const { return align_up(trace) + align_up(scratch) + align_up(idx) + align_up(proofs); }
};

Layout job_layout(uint64_t nb, uint64_t nbe, uint64_t nc, uint64_t arity, uint64_t nq, uint64_t llv, bool dev_trace)
{
    Layout L;
    uint64_t n = 1ull << nb, ne = 1ull << nbe;
    L.lde = 8 * ne * nc;
    L.nodes = 8 * pil2hip_tree_num_elements(ne, arity);
    L.trace = dev_trace ? 0 : 8 * n * nc;
    L.scratch = 8 * n * nc;
    L.pw = pil2hip_proof_words(ne, nc, arity, llv);
    L.idx = 8 * nq;
    L.proofs = 8 * nq * L.pw;
    return L;
}

struct Job {
    uint64_t id = 0;
    Pil2ServeRequest req{};
    std::vector<uint64_t> queries;
    Layout L{};
    St st = St::Queued;
    int64_t status = 0;
    double t_submit = 0, t_admit = 0, t_done = 0;
    uint64_t off_res = 0, off_tmp = 0;
    int slot = -1;
    uint32_t stream = 0;
    double work = 0;
    int opens = 0;
    uint64_t *d_lde = nullptr, *d_nodes = nullptr;
};

struct Slot {
    hipEvent_t ev = nullptr;
    uint64_t *pinned = nullptr;  // [queries | root (4) | proofs]
    uint64_t words = 0;
    bool busy = false;
};

struct Server {
    Pil2ServeConfig cfg{};
    int device = 0;
    Arena arena;
    std::vector<hipStream_t> streams;
    std::vector<double> stream_load;
    std::vector<Slot> slots;
    std::vector<uint64_t *> old_pinned;  // outgrown staging buffers, freed at destroy
    hipStream_t open_stream = nullptr;

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::thread thread;
    bool stop = false;
    std::unordered_map<uint64_t, Job> jobs;
    std::deque<uint64_t> queue;
    std::vector<uint64_t> running;
    std::deque<Pil2ServeCompletion> completions;
    uint64_t next_id = 0;
    Pil2ServeStats stats{};
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();

    double now_us() const
    {
        return std::chrono::duration<double, std::micro>(std::chrono::steady_clock::now() - t0).count();
    }

    void complete(Job &j, int64_t status)
    {
        j.status = status;
        j.t_done = now_us();
        Pil2ServeCompletion c{};
        c.job_id = j.id;
        c.tag = j.req.tag;
        c.status = status;
        c.t_submit_us = j.t_submit;
        c.t_admit_us = j.t_admit;
        c.t_done_us = j.t_done;
        c.stream = j.stream;
        completions.push_back(c);
        if (status == 0) stats.completed++;
        else if (status == kErrCancelled) stats.cancelled++;
        else stats.failed++;
        cv_done.notify_all();
    }

    void free_transient(Job &j)
    {
        if (j.L.transient()) arena.release(j.off_tmp, j.L.transient());
    }
    void free_resident(Job &j)
    {
        arena.release(j.off_res, j.L.resident());
        stats.resident_jobs--;
    }

    // Admits the job if a slot and arena space are available; returns false otherwise.
    bool try_admit(Job &j)
    {
        int s = -1;
        for (size_t i = 0; i < slots.size(); i++)
            if (!slots[i].busy) {
                s = (int)i;
                break;
            }
        if (s < 0) return false;
        uint64_t off_res, off_tmp = 0;
        if (!arena.alloc(j.L.resident(), &off_res)) return false;
        if (j.L.transient() && !arena.alloc(j.L.transient(), &off_tmp)) {
            arena.release(off_res, j.L.resident());
            return false;
        }
        uint64_t need = j.queries.size() + 4 + j.queries.size() * j.L.pw;
        Slot &sl = slots[s];
        if (sl.words < need) {  // grow the slot's staging buffer (rare)
            uint64_t *p = nullptr;
            if (hipHostMalloc((void **)&p, need * 8, hipHostMallocDefault) != hipSuccess) {
                arena.release(off_res, j.L.resident());
                if (j.L.transient()) arena.release(off_tmp, j.L.transient());
                return false;
            }
            if (sl.pinned) old_pinned.push_back(sl.pinned);
            sl.pinned = p;
            sl.words = need;
        }
        j.off_res = off_res;
        j.off_tmp = off_tmp;
        j.slot = s;
        sl.busy = true;
        j.st = St::Running;
        j.t_admit = now_us();
        stats.admitted++;
        stats.resident_jobs++;
        uint32_t best = 0;
        for (uint32_t k = 1; k < streams.size(); k++)
            if (stream_load[k] < stream_load[best]) best = k;
        j.stream = best;
        stream_load[best] += j.work;
        int64_t rc = enqueue(j, sl, streams[best]);
        if (rc) {  // launch-time failure: the job completes with the error
            stream_load[best] -= j.work;
            (void)hipStreamSynchronize(streams[best]);
            sl.busy = false;
            free_transient(j);
            free_resident(j);
            j.st = St::Failed;
            complete(j, rc);
            return true;
        }
        running.push_back(j.id);
        if (running.size() > stats.peak_running) stats.peak_running = (uint32_t)running.size();
        return true;
    }

    int64_t enqueue(Job &j, Slot &sl, hipStream_t st)
    {
        const Pil2ServeRequest &r = j.req;
        char *res = arena.base + j.off_res, *tmp = arena.base + j.off_tmp;
        j.d_lde = (uint64_t *)res;
        j.d_nodes = (uint64_t *)(res + align_up(j.L.lde));
        const uint64_t *d_trace = r.d_trace;
        char *p = tmp;
        uint64_t n = 1ull << r.n_bits, ne = 1ull << r.n_bits_ext, nq = j.queries.size();
        if (!d_trace) {
            uint64_t *t = (uint64_t *)p;
            p += align_up(j.L.trace);
            if (r.h_trace) {
                if (hipMemcpyAsync(t, r.h_trace, 8 * n * r.n_cols, hipMemcpyHostToDevice, st) != hipSuccess)
                    return api_fail(kErrArena, "trace copy failed");
            } else if (n * r.n_cols) {
                int e = ntt_launch_fill_uniform(t, n * r.n_cols, r.seed, 0, st);
                if (e) return api_fail(-100 - e, "fill launch failed");
            }
            d_trace = t;
        }
        uint64_t *d_scratch = (uint64_t *)p;
        p += align_up(j.L.scratch);
        uint64_t *d_idx = (uint64_t *)p;
        p += align_up(j.L.idx);
        uint64_t *d_proofs = (uint64_t *)p;
        int64_t rc = commit_enqueue(j.d_lde, j.d_nodes, d_trace, d_scratch, cfg.arity, r.n_bits, r.n_bits_ext,
                                    r.n_cols, st);
        if (rc) return rc;
        uint64_t nt = pil2hip_tree_num_elements(ne, cfg.arity);
        uint64_t *h_q = sl.pinned, *h_root = sl.pinned + nq, *h_pr = h_root + 4;
        if (nq) {
            memcpy(h_q, j.queries.data(), 8 * nq);
            if (hipMemcpyAsync(d_idx, h_q, 8 * nq, hipMemcpyHostToDevice, st) != hipSuccess)
                return api_fail(kErrArena, "query copy failed");
            rc = pil2hip_group_proofs(d_proofs, j.d_lde, 1, ne, r.n_cols, j.d_nodes, ne, cfg.arity, cfg.llv, d_idx,
                                      nq, st);
            if (rc) return rc;
            if (hipMemcpyAsync(h_pr, d_proofs, 8 * nq * j.L.pw, hipMemcpyDeviceToHost, st) != hipSuccess)
                return api_fail(kErrArena, "proof copy failed");
        }
        if (hipMemcpyAsync(h_root, j.d_nodes + nt - 4, 32, hipMemcpyDeviceToHost, st) != hipSuccess)
            return api_fail(kErrArena, "root copy failed");
        if (hipEventRecord(sl.ev, st) != hipSuccess) return api_fail(kErrArena, "event record failed");
        return 0;
    }

    void finish(Job &j, hipError_t q)
    {
        Slot &sl = slots[j.slot];
        uint64_t nq = j.queries.size();
        int64_t status = 0;
        if (q != hipSuccess) {
            status = -100 - (int64_t)q;
        } else {
            if (j.req.h_root) memcpy(j.req.h_root, sl.pinned + nq, 32);
            if (j.req.h_proofs && nq) memcpy(j.req.h_proofs, sl.pinned + nq + 4, 8 * nq * j.L.pw);
        }
        sl.busy = false;
        stream_load[j.stream] -= j.work;
        free_transient(j);
        if (status || !j.req.keep_resident) {
            free_resident(j);
            j.st = status ? St::Failed : St::Released;
        } else {
            j.st = St::Done;
        }
        complete(j, status);
    }

    void loop()
    {
        (void)hipSetDevice(device);
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            bool progressed = false;
            for (size_t i = 0; i < running.size();) {
                Job &j = jobs[running[i]];
                hipError_t q = hipEventQuery(slots[j.slot].ev);
                if (q == hipErrorNotReady) {
                    i++;
                    continue;
                }
                finish(j, q);
                running.erase(running.begin() + i);
                progressed = true;
            }
            while (!queue.empty() && running.size() < cfg.max_inflight) {
                Job &j = jobs[queue.front()];
                if (!try_admit(j)) break;
                queue.pop_front();
                progressed = true;
            }
            if (stop && running.empty()) {
                while (!queue.empty()) {  // cancel what was never admitted
                    Job &j = jobs[queue.front()];
                    queue.pop_front();
                    j.st = St::Cancelled;
                    complete(j, kErrCancelled);
                }
                break;
            }
            if (running.empty()) {
                cv_work.wait_for(lk, std::chrono::milliseconds(5));
            } else if (!progressed) {
                lk.unlock();
                std::this_thread::sleep_for(std::chrono::microseconds(20));
                lk.lock();
            }
        }
    }
};

int64_t check_request(const Server *s, const Pil2ServeRequest *r)
{
    if (!r) return api_fail(-13, "null request");
    if (r->n_bits_ext < r->n_bits) return api_fail(-10, "nBitsExt < nBits");
    if (r->n_bits_ext > NTT_MAX_LOG) return api_fail(-11, "nBitsExt > 28 is not supported");
    if (r->n_cols > 65535) return api_fail(-12, "nCols > 65535 is not supported");
    if (r->n_queries && !r->h_queries) return api_fail(-13, "null queries");
    uint64_t ne = 1ull << r->n_bits_ext;
    for (uint64_t q = 0; q < r->n_queries; q++)
        if (r->h_queries[q] >= ne) return api_fail(-28, "query index out of range");
    (void)s;
    return 0;
}

}  // namespace

extern "C" {

uint64_t pil2hip_serve_job_bytes(uint64_t n_bits, uint64_t n_bits_ext, uint64_t n_cols, uint64_t arity,
                                 uint64_t n_queries, uint64_t llv, bool device_trace)
{
    Layout L = job_layout(n_bits, n_bits_ext, n_cols, arity, n_queries, llv, device_trace);
    return L.resident() + L.transient();
}

void *pil2hip_serve_create(const Pil2ServeConfig *cfg)
{
    if (!cfg) {
        api_fail(-13, "null config");
        return nullptr;
    }
    if (cfg->arity < 2 || cfg->arity > 4) {
        api_fail(-7, "unsupported arity (Poseidon2 supports 2, 3, 4)");
        return nullptr;
    }
    if (cfg->n_streams < 1 || cfg->n_streams > 16 || cfg->max_inflight < 1 || cfg->max_inflight > 256 ||
        cfg->llv > 8) {
        api_fail(-29, "bad server config (1..16 streams, 1..256 in flight, llv <= 8)");
        return nullptr;
    }
    Server *s = new Server();
    s->cfg = *cfg;
    s->stats.budget_bytes = cfg->hbm_budget_bytes;
    if (hipGetDevice(&s->device) != hipSuccess) {
        delete s;
        api_fail(-8, "no HIP device");
        return nullptr;
    }
    char *base = nullptr;
    if (cfg->hbm_budget_bytes && hipMalloc((void **)&base, cfg->hbm_budget_bytes) != hipSuccess) {
        delete s;
        api_fail(kErrArena, "cannot allocate the HBM arena");
        return nullptr;
    }
    s->arena.init(base, cfg->hbm_budget_bytes);
    s->streams.resize(cfg->n_streams);
    s->stream_load.assign(cfg->n_streams, 0.0);
    hipError_t ce = hipSuccess;
    for (auto &st : s->streams)
        if (ce == hipSuccess) ce = hipStreamCreateWithFlags(&st, hipStreamNonBlocking);
    if (ce == hipSuccess) ce = hipStreamCreateWithFlags(&s->open_stream, hipStreamNonBlocking);
    s->slots.resize(cfg->max_inflight);
    for (auto &sl : s->slots)
        if (ce == hipSuccess) ce = hipEventCreateWithFlags(&sl.ev, hipEventDisableTiming);
    if (ce != hipSuccess) {  // no thread yet: release what exists and report
        for (auto st : s->streams)
            if (st) (void)hipStreamDestroy(st);
        if (s->open_stream) (void)hipStreamDestroy(s->open_stream);
        for (auto &sl : s->slots)
            if (sl.ev) (void)hipEventDestroy(sl.ev);
        if (base) (void)hipFree(base);
        delete s;
        api_fail(-100 - (int64_t)ce, std::string("server streams/events: ") + hipGetErrorString(ce));
        return nullptr;
    }
    s->thread = std::thread([s] { s->loop(); });
    return s;
}

int64_t pil2hip_serve_submit(void *srv, const Pil2ServeRequest *req)
{
    Server *s = (Server *)srv;
    if (!s) return api_fail(-13, "null server");
    int64_t rc = check_request(s, req);
    if (rc) return rc;
    Layout L = job_layout(req->n_bits, req->n_bits_ext, req->n_cols, s->cfg.arity, req->n_queries, s->cfg.llv,
                          req->d_trace != nullptr);
    if (L.resident() + L.transient() > s->cfg.hbm_budget_bytes)
        return api_fail(kErrBudget, "job needs more HBM than the server budget");
    std::lock_guard<std::mutex> lk(s->mu);
    uint64_t id = s->next_id++;
    Job &j = s->jobs[id];
    j.id = id;
    j.req = *req;
    j.queries.assign(req->h_queries, req->h_queries + req->n_queries);
    j.req.h_queries = nullptr;
    j.L = L;
    j.t_submit = s->now_us();
    j.work = (double)(1ull << req->n_bits_ext) * (double)(req->n_cols + 1);
    s->queue.push_back(id);
    s->stats.submitted++;
    s->cv_work.notify_all();
    return (int64_t)id;
}

int64_t pil2hip_serve_poll(void *srv, Pil2ServeCompletion *out, uint64_t max, int32_t timeout_ms)
{
    Server *s = (Server *)srv;
    if (!s || (!out && max)) return api_fail(-13, "null argument");
    std::unique_lock<std::mutex> lk(s->mu);
    if (s->completions.empty() && timeout_ms > 0)
        s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), [s] { return !s->completions.empty(); });
    uint64_t k = 0;
    while (k < max && !s->completions.empty()) {
        out[k++] = s->completions.front();
        s->completions.pop_front();
    }
    return (int64_t)k;
}

int64_t pil2hip_serve_cancel(void *srv, uint64_t job_id)
{
    Server *s = (Server *)srv;
    if (!s) return api_fail(-13, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(job_id);
    if (it == s->jobs.end()) return api_fail(kErrUnknownJob, "unknown job id");
    Job &j = it->second;
    if (j.st != St::Queued) return api_fail(kErrNotQueued, "job already admitted");
    for (auto q = s->queue.begin(); q != s->queue.end(); ++q)
        if (*q == job_id) {
            s->queue.erase(q);
            break;
        }
    j.st = St::Cancelled;
    s->complete(j, kErrCancelled);
    s->cv_work.notify_all();
    return 0;
}

int64_t pil2hip_serve_open(void *srv, uint64_t job_id, const uint64_t *h_queries, uint64_t n_queries,
                           uint64_t *h_proofs)
{
    Server *s = (Server *)srv;
    if (!s) return api_fail(-13, "null server");
    if (n_queries == 0) return 0;
    if (!h_queries || !h_proofs) return api_fail(-13, "null argument");
    uint64_t off = 0, bytes = 0, pw = 0, ne = 0, nc = 0;
    uint64_t *d_lde = nullptr, *d_nodes = nullptr;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        auto it = s->jobs.find(job_id);
        if (it == s->jobs.end()) return api_fail(kErrUnknownJob, "unknown job id");
        Job &j = it->second;
        if (j.st != St::Done) return api_fail(kErrNotDone, "job is not completed and resident");
        ne = 1ull << j.req.n_bits_ext;
        nc = j.req.n_cols;
        for (uint64_t q = 0; q < n_queries; q++)
            if (h_queries[q] >= ne) return api_fail(-28, "query index out of range");
        pw = j.L.pw;
        bytes = align_up(8 * n_queries) + 8 * n_queries * pw;
        if (!s->arena.alloc(bytes, &off)) return api_fail(kErrBudget, "no arena space for the openings");
        j.opens++;
        d_lde = j.d_lde;
        d_nodes = j.d_nodes;
    }
    uint64_t *d_idx = (uint64_t *)(s->arena.base + off);
    uint64_t *d_out = (uint64_t *)(s->arena.base + off + align_up(8 * n_queries));
    hipStream_t st = s->open_stream;
    int64_t rc = 0;
    if (hipMemcpyAsync(d_idx, h_queries, 8 * n_queries, hipMemcpyHostToDevice, st) != hipSuccess)
        rc = api_fail(kErrArena, "query copy failed");
    if (!rc)
        rc = pil2hip_group_proofs(d_out, d_lde, 1, ne, nc, d_nodes, ne, s->cfg.arity, s->cfg.llv, d_idx, n_queries, st);
    if (!rc && hipMemcpyAsync(h_proofs, d_out, 8 * n_queries * pw, hipMemcpyDeviceToHost, st) != hipSuccess)
        rc = api_fail(kErrArena, "proof copy failed");
    hipError_t e = hipStreamSynchronize(st);
    if (!rc && e != hipSuccess) rc = api_fail(-100 - (int64_t)e, "openings failed");
    std::lock_guard<std::mutex> lk(s->mu);
    s->arena.release(off, bytes);
    s->jobs[job_id].opens--;
    s->cv_work.notify_all();
    return rc;
}

int64_t pil2hip_serve_release(void *srv, uint64_t job_id)
{
    Server *s = (Server *)srv;
    if (!s) return api_fail(-13, "null server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(job_id);
    if (it == s->jobs.end()) return api_fail(kErrUnknownJob, "unknown job id");
    Job &j = it->second;
    if (j.st == St::Done) {
        if (j.opens) return api_fail(kErrNotDone, "openings in progress");
        s->free_resident(j);
        j.st = St::Released;
    } else if (j.st == St::Queued || j.st == St::Running) {
        return api_fail(kErrNotDone, "job not completed");
    }
    s->jobs.erase(it);  // released, failed and cancelled jobs are forgotten
    s->cv_work.notify_all();
    return 0;
}

int64_t pil2hip_serve_stats(void *srv, Pil2ServeStats *out)
{
    Server *s = (Server *)srv;
    if (!s || !out) return api_fail(-13, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    s->stats.queued = s->queue.size();
    s->stats.running = s->running.size();
    s->stats.bytes_in_use = s->arena.used;
    s->stats.peak_bytes = s->arena.peak;
    *out = s->stats;
    return 0;
}

double pil2hip_serve_now_us(void *srv)
{
    Server *s = (Server *)srv;
    return s ? s->now_us() : 0.0;
}

void pil2hip_serve_destroy(void *srv)
{
    Server *s = (Server *)srv;
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv_work.notify_all();
    }
    if (s->thread.joinable()) s->thread.join();
    for (auto st : s->streams) {
        (void)hipStreamSynchronize(st);
        (void)hipStreamDestroy(st);
    }
    (void)hipStreamDestroy(s->open_stream);
    for (auto &sl : s->slots) {
        (void)hipEventDestroy(sl.ev);
        if (sl.pinned) (void)hipHostFree(sl.pinned);
    }
    for (auto p : s->old_pinned) (void)hipHostFree(p);
    if (s->arena.base) (void)hipFree(s->arena.base);
    delete s;
}

}  // extern "C"
