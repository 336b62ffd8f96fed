// server.cpp -- continuous batching of LD-score requests over resident panels.
//
// No counterpart in the reference (its nearest analogue is one
// compute_l2_from_bed call per chromosome shard from a browser worker,
// ldsc-web/src/worker.rs:664-872).
//
// Requests enter a bounded queue.  `max_inflight` worker threads each own a
// device workspace and HIP stream; a worker takes the oldest queued request
// that fits the HBM budget (first fit in arrival order; once the oldest
// request has waited longer than kStarveSecs, nothing may overtake it) and
// runs it on its own stream.  Requests that arrive while others run are
// admitted as soon as a slot and memory free up, and their kernels overlap on
// the device with those already running.
//
// HBM accounting.  A workspace keeps its buffers from request to request
// (grow-only), so the budget is charged with what each worker holds: while
// idle, its workspace's device bytes; while running, those plus the
// request's estimate (l2_workspace_bytes over the request's own band, an
// upper bound of what the request can add).  After a request the charge
// falls back to the bytes the workspace then holds.  A worker frees its
// buffers when that would exceed the budget, and an idle worker frees them
// when a queued request fits nowhere, so retained buffers never keep a
// request out for good.
//
// Cancel and deadline.  A queued request is cancelled at once; a running one
// stops at the next phase boundary (after the tiling and planning, after the
// band), once its device work has drained.  With max_queue_ms a request that
// waits longer than that in the queue fails with LDSC_ERR_TIMEOUT.
//
// Band batching (batch_window_us > 0): a worker prepares its request
// (tiling, planning, uploads: l2_begin) on its own stream and then hands the
// band contraction to the batcher.  The first worker to arrive leads: while
// the handed-in tiles would not fill the chip it waits, up to the window,
// for the other workers still preparing; then it puts the tiles of every
// request handed in so far into one launch per band form
// (band_batch_launch, after each request's preparation, on one of a few
// streams in turn); requests handed in meanwhile go into the next launch.  Each
// worker then reduces and copies back its own result (l2_end).
//
// Each request's result is independent of what else is in flight or shares
// its launch: a tile reads and writes only its own request's buffers, and
// every reduction runs in a fixed order.
#include <algorithm>
#include <chrono>
#include <atomic>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>

#include "device.h"

namespace {

constexpr double kStarveSecs = 2.0;

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

struct Request {
    uint64_t id = 0;
    const ldsc_panel* panel = nullptr;
    std::vector<ldsc_snp> snps;
    ldsc_l2_config cfg{};
    std::vector<double> annot;
    uint64_t K = 1;
    double* l2_out = nullptr;
    double* maf_out = nullptr;
    size_t bytes = 0;  // device bytes the request may add to a workspace
    double t_arrival = 0, t_done = 0;
    int state = 0;  // 0 queued, 1 running, 2 done
    int rc = LDSC_OK;
    std::string err;
    std::atomic<bool> cancel{false};
};

// A worker thread's device state, under ldsc_server::mu.
struct Worker {
    ldsc::Workspace* ws = nullptr;
    int device = -1;
    size_t charge = 0;  // bytes of the budget charged to this worker
    bool trim = false;  // free the workspace's buffers before the next take
};

}  // namespace

// Leader / follower batching of band launches (see the file comment).
struct BandBatcher {
    struct Entry {
        ldsc::L2Call* call;
        bool launched = false;
        int rc = LDSC_OK;
        std::string err;
    };
    // Launches rotate over kSlots argument tables, each with its own stream,
    // so that consecutive launches can overlap on the device as the
    // per-request launches of separate workers do.
    struct Device {
        static constexpr int kSlots = 4;
        hipStream_t streams[kSlots] = {};
        ldsc::BandBatchSlot slots[kSlots];
        int next = 0;
    };
    // A launch with this many tiles already fills the chip (one wave per
    // SIMD on 256 CUs): the leader then stops waiting for more requests.
    static constexpr uint64_t kFillTiles = 1024;
    std::mutex mu;
    std::condition_variable cv;
    std::vector<Entry*> pending;
    bool leading = false;
    uint32_t preparing = 0;  // workers between taking a request and handing in its band
    double window = 0.0;     // seconds
    std::unordered_map<int, Device> devices;
    uint64_t launches = 0, batched = 0;
    uint32_t max_batch = 0;

    // what one round of launches did, added to the counters under mu
    struct Tally {
        uint64_t launches = 0, batched = 0;
        uint32_t max_batch = 0;
    };

    int run(ldsc::L2Call* call, char* err, uint64_t err_len);
    Tally launch_all(std::vector<Entry*>& take);
    void destroy();
};

// Runs on the leader alone (mu released): it owns `devices` meanwhile.
BandBatcher::Tally BandBatcher::launch_all(std::vector<Entry*>& take) {
    // one launch per (device, form) group, at most BandBatchSlot::kMaxCalls each
    Tally tally;
    std::vector<bool> used(take.size(), false);
    for (size_t i = 0; i < take.size(); ++i) {
        if (used[i]) continue;
        std::vector<Entry*> grp;
        std::vector<ldsc::L2Call*> calls;
        const int dev = take[i]->call->panel->device, form = take[i]->call->form;
        for (size_t j = i; j < take.size() && calls.size() < ldsc::BandBatchSlot::kMaxCalls; ++j)
            if (!used[j] && take[j]->call->panel->device == dev && take[j]->call->form == form) {
                used[j] = true;
                grp.push_back(take[j]);
                calls.push_back(take[j]->call);
            }
        char err[512] = {0};
        int rc = LDSC_OK;
        Device& d = devices[dev];
        const int k = d.next;
        d.next = (d.next + 1) % Device::kSlots;
        hipError_t e = hipSetDevice(dev);
        if (e == hipSuccess && !d.streams[k])
            e = hipStreamCreateWithFlags(&d.streams[k], hipStreamNonBlocking);
        if (e != hipSuccess)
            rc = ldsc::fail(err, sizeof(err), LDSC_ERR_DEVICE, "batch stream: %s", hipGetErrorString(e));
        else
            rc = ldsc::band_batch_launch(calls.data(), (uint32_t)calls.size(), &d.slots[k],
                                         d.streams[k], err, sizeof(err));
        for (Entry* en : grp) {
            en->rc = rc;
            if (rc) en->err = err;
        }
        if (rc == LDSC_OK) {
            ++tally.launches;
            if (grp.size() > 1) tally.batched += grp.size();
            tally.max_batch = std::max<uint32_t>(tally.max_batch, (uint32_t)grp.size());
        }
    }
    return tally;
}

int BandBatcher::run(ldsc::L2Call* call, char* err, uint64_t err_len) {
    Entry me{call};
    std::unique_lock<std::mutex> lk(mu);
    --preparing;
    pending.push_back(&me);
    cv.notify_all();
    while (!me.launched) {
        if (leading) {
            cv.wait(lk);
            continue;
        }
        leading = true;
        const auto until = std::chrono::steady_clock::now() +
                           std::chrono::duration_cast<std::chrono::steady_clock::duration>(
                               std::chrono::duration<double>(window));
        auto pending_tiles = [&] {
            uint64_t t = 0;
            for (const Entry* en : pending) t += en->call->ba.ntiles;
            return t;
        };
        // wait for the requests still being prepared while the launch would
        // not yet fill the chip
        while (preparing > 0 && pending_tiles() < kFillTiles &&
               cv.wait_until(lk, until) != std::cv_status::timeout) {
        }
        std::vector<Entry*> take;
        take.swap(pending);
        lk.unlock();
        const Tally t = launch_all(take);
        lk.lock();
        launches += t.launches;
        batched += t.batched;
        max_batch = std::max(max_batch, t.max_batch);
        for (Entry* en : take) en->launched = true;
        leading = false;
        cv.notify_all();
    }
    if (me.rc != LDSC_OK) return ldsc::fail(err, err_len, me.rc, "%s", me.err.c_str());
    return LDSC_OK;
}

void BandBatcher::destroy() {
    for (auto& kv : devices) {
        (void)hipSetDevice(kv.first);
        for (hipStream_t& st : kv.second.streams)
            if (st) {
                (void)hipStreamSynchronize(st);
                (void)hipStreamDestroy(st);
                st = nullptr;
            }
        for (auto& sl : kv.second.slots) sl.release();
    }
    devices.clear();
}

struct ldsc_server {
    ldsc_server_config cfg{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> queue;
    std::unordered_map<uint64_t, std::unique_ptr<Request>> reqs;
    uint64_t next_id = 1;
    size_t in_use = 0, in_use_max = 0;  // sum of the workers' charges
    bool stop = false;
    std::vector<std::thread> workers;
    BandBatcher batcher;
    uint64_t done_count = 0, cancelled = 0, timed_out = 0, trims = 0;

    void worker_main();
    Request* take_locked(Worker& w);
    void expire_locked(double now);
    void finish_locked(Request* r, int rc, const char* msg);
    int run_request(Request* r, ldsc::Workspace* ws, char* err, uint64_t err_len);
};

// One request on workspace ws: prepare, band (batched or alone), finish.
int ldsc_server::run_request(Request* r, ldsc::Workspace* ws, char* err, uint64_t err_len) {
    const bool batch = cfg.batch_window_us > 0;
    if (batch) {
        std::lock_guard<std::mutex> lk(batcher.mu);
        ++batcher.preparing;
    }
    ldsc::L2Call call;
    int rc = ldsc::l2_begin(r->panel, r->snps.data(), r->snps.size(), &r->cfg,
                            r->annot.empty() ? nullptr : r->annot.data(), r->K, r->l2_out,
                            r->maf_out, nullptr, nullptr, nullptr, ws, &call, err, err_len);
    // a cancel seen at a phase boundary: the work already queued on the
    // workspace's streams drains before the workspace serves another request
    auto cancelled = [&]() {
        (void)hipStreamSynchronize(ws->stream);
        (void)hipStreamSynchronize(ws->copy_stream);
        (void)hipStreamSynchronize(ws->up_stream);
        return ldsc::fail(err, err_len, LDSC_ERR_CANCELLED, "request %llu cancelled",
                          (unsigned long long)r->id);
    };
    const bool stop_now = rc == LDSC_OK && !call.done && r->cancel.load();
    const bool batched =
        batch && rc == LDSC_OK && !call.done && !stop_now && !ldsc::is_sketch_form(call.form);
    if (batch && !batched) {
        std::lock_guard<std::mutex> lk(batcher.mu);
        --batcher.preparing;
        batcher.cv.notify_all();
    }
    if (rc != LDSC_OK || call.done) return rc;
    if (stop_now) return cancelled();
    if (batched) {
        rc = batcher.run(&call, err, err_len);
    } else {
        rc = ldsc::l2_band(&call, err, err_len);
        if (rc == LDSC_OK) {
            std::lock_guard<std::mutex> lk(batcher.mu);
            ++batcher.launches;
            batcher.max_batch = std::max<uint32_t>(batcher.max_batch, 1);
        }
    }
    if (rc != LDSC_OK) return rc;
    if (r->cancel.load()) {
        // the band may run on a batch stream: it recorded ev[4] after it
        (void)hipEventSynchronize(ws->ev[4]);
        return cancelled();
    }
    return ldsc::l2_end(&call, err, err_len);
}

void ldsc_server::finish_locked(Request* r, int rc, const char* msg) {
    r->state = 2;
    r->rc = rc;
    r->err = msg;
    r->t_done = now_s();
}

// Queued requests past the queue deadline fail with LDSC_ERR_TIMEOUT.
void ldsc_server::expire_locked(double now) {
    if (cfg.max_queue_ms == 0) return;
    const double limit = cfg.max_queue_ms * 1e-3;
    bool any = false;
    for (auto it = queue.begin(); it != queue.end();) {
        if (now - (*it)->t_arrival > limit) {
            finish_locked(*it, LDSC_ERR_TIMEOUT, "request waited longer than max_queue_ms");
            ++timed_out;
            it = queue.erase(it);
            any = true;
        } else {
            ++it;
        }
    }
    if (any) cv_done.notify_all();
}

// The oldest queued request that fits the budget next to what the other
// workers hold, with this worker's own buffers counted as they stand.
Request* ldsc_server::take_locked(Worker& w) {
    expire_locked(now_s());
    if (queue.empty()) return nullptr;
    const size_t others = in_use - w.charge;
    const size_t room = cfg.hbm_budget_bytes > others ? cfg.hbm_budget_bytes - others : 0;
    const bool head_starving = now_s() - queue.front()->t_arrival > kStarveSecs;
    for (auto it = queue.begin(); it != queue.end(); ++it) {
        const size_t held = (*it)->panel->device == w.device ? w.charge : 0;
        if (held + (*it)->bytes <= room) {
            Request* r = *it;
            queue.erase(it);
            return r;
        }
        if (head_starving) break;
    }
    // nothing fits: buffers this idle worker keeps are given back
    if (w.charge > 0) w.trim = true;
    return nullptr;
}

void ldsc_server::worker_main() {
    Worker w;
    std::unique_lock<std::mutex> lk(mu);
    for (;;) {
        Request* r = nullptr;
        cv_work.wait(lk, [&] {
            if (stop) return true;
            r = take_locked(w);
            return r != nullptr || w.trim;
        });
        if (!r && w.trim) {
            w.trim = false;
            lk.unlock();
            if (w.ws) w.ws->trim();
            lk.lock();
            in_use -= w.charge;
            w.charge = w.ws ? w.ws->bytes() : 0;
            in_use += w.charge;
            ++trims;
            cv_work.notify_all();
            continue;
        }
        if (!r) break;  // stop
        // charge: the buffers held (on this device) plus what the request adds
        const bool same = w.ws && w.device == r->panel->device;
        in_use -= w.charge;
        w.charge = (same ? w.charge : 0) + r->bytes;
        in_use += w.charge;
        in_use_max = std::max(in_use_max, in_use);
        r->state = 1;
        lk.unlock();
        if (w.ws && !same) {
            w.ws->destroy();
            delete w.ws;
            w.ws = nullptr;
        }
        if (!w.ws) {
            w.ws = new ldsc::Workspace();
            if (w.ws->init(r->panel->device) != hipSuccess) {
                w.ws->destroy();
                delete w.ws;
                w.ws = nullptr;
            }
            w.device = r->panel->device;
        }
        char err[512] = {0};
        int rc;
        if (!w.ws) {
            rc = ldsc::fail(err, sizeof(err), LDSC_ERR_DEVICE, "could not create a device workspace");
        } else {
            rc = run_request(r, w.ws, err, sizeof(err));
        }
        // the charge falls to what the workspace now holds; over the budget
        // (next to the other workers) its buffers are freed
        size_t held = w.ws ? w.ws->bytes() : 0;
        lk.lock();
        if (in_use - w.charge + held > cfg.hbm_budget_bytes) {
            lk.unlock();
            w.ws->trim();
            held = w.ws->bytes();
            lk.lock();
            ++trims;
        }
        in_use -= w.charge;
        w.charge = held;
        in_use += w.charge;
        ++done_count;
        if (rc == LDSC_ERR_CANCELLED) ++cancelled;
        finish_locked(r, rc, err);
        cv_done.notify_all();
        cv_work.notify_all();
    }
    in_use -= w.charge;
    lk.unlock();
    if (w.ws) {
        w.ws->destroy();
        delete w.ws;
    }
}

extern "C" int ldsc_server_create(const ldsc_server_config* cfg, ldsc_server** out, char* err,
                                  uint64_t err_len) {
    if (!cfg || !out) return ldsc::fail(err, err_len, LDSC_ERR_INVALID_ARG, "NULL argument");
    if (cfg->max_inflight < 1 || cfg->max_inflight > 64)
        return ldsc::fail(err, err_len, LDSC_ERR_INVALID_ARG, "max_inflight must be in [1, 64]");
    if (cfg->queue_capacity < 1)
        return ldsc::fail(err, err_len, LDSC_ERR_INVALID_ARG, "queue_capacity must be >= 1");
    if (cfg->batch_window_us > 1000000)
        return ldsc::fail(err, err_len, LDSC_ERR_INVALID_ARG, "batch_window_us must be <= 1000000");
    if (cfg->hbm_budget_bytes == 0)
        return ldsc::fail(err, err_len, LDSC_ERR_INVALID_ARG, "hbm_budget_bytes must be > 0");
    auto* s = new ldsc_server();
    s->cfg = *cfg;
    s->batcher.window = cfg->batch_window_us * 1e-6;
    for (uint32_t i = 0; i < cfg->max_inflight; ++i)
        s->workers.emplace_back([s] { s->worker_main(); });
    *out = s;
    return LDSC_OK;
}

extern "C" int ldsc_server_destroy(ldsc_server* s) {
    if (!s) return LDSC_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        // queued (not yet running) requests are abandoned
        for (Request* r : s->queue)
            s->finish_locked(r, LDSC_ERR_CANCELLED, "server destroyed before the request ran");
        s->queue.clear();
    }
    s->cv_work.notify_all();
    for (auto& t : s->workers) t.join();
    s->batcher.destroy();
    delete s;
    return LDSC_OK;
}

extern "C" int ldsc_server_stats_get(ldsc_server* s, ldsc_server_stats* out) {
    if (!s || !out) return LDSC_ERR_INVALID_ARG;
    std::memset(out, 0, sizeof(*out));
    {
        std::lock_guard<std::mutex> lk(s->mu);
        out->requests_done = s->done_count;
        out->requests_cancelled = s->cancelled;
        out->requests_timed_out = s->timed_out;
        out->hbm_charged = s->in_use;
        out->hbm_charged_max = s->in_use_max;
        out->workspace_trims = s->trims;
    }
    std::lock_guard<std::mutex> lk(s->batcher.mu);
    out->band_launches = s->batcher.launches;
    out->batched_requests = s->batcher.batched;
    out->max_batch = s->batcher.max_batch;
    return LDSC_OK;
}

extern "C" int ldsc_l2_submit(ldsc_server* s, const ldsc_panel* panel, const ldsc_snp* snps,
                              uint64_t m, const ldsc_l2_config* cfg, const double* annot,
                              uint64_t K, double* l2_out, double* maf_out, uint64_t* request_id,
                              char* err, uint64_t err_len) {
    if (!s || !panel || !cfg || !request_id)
        return ldsc::fail(err, err_len, LDSC_ERR_INVALID_ARG, "NULL argument");
    int rc = ldsc::check_config(cfg, err, err_len);
    if (rc) return rc;
    if (m > 0 && (!snps || !l2_out))
        return ldsc::fail(err, err_len, LDSC_ERR_INVALID_ARG, "NULL SNP or output array");
    if (K < 1 || (!annot && K != 1))
        return ldsc::fail(err, err_len, LDSC_ERR_ANNOT, "K must be 1 without annotations, >= 1 with");
    auto r = std::make_unique<Request>();
    r->panel = panel;
    r->snps.assign(snps, snps + m);
    r->cfg = *cfg;
    if (annot) r->annot.assign(annot, annot + m * K);
    r->K = K;
    r->l2_out = l2_out;
    r->maf_out = maf_out;
    // the estimate uses the request's own band: its pair bounds and tiles
    {
        std::vector<uint64_t> lo;
        const uint64_t tiles = ldsc::band_tile_count(snps, m, cfg, lo);
        r->bytes = ldsc::l2_workspace_bytes(panel, m, cfg, K, tiles, annot != nullptr);
    }
    r->t_arrival = now_s();
    if (r->bytes > s->cfg.hbm_budget_bytes)
        return ldsc::fail(err, err_len, LDSC_ERR_OOM,
                          "request needs ~%llu device bytes, over the %llu-byte budget",
                          (unsigned long long)r->bytes,
                          (unsigned long long)s->cfg.hbm_budget_bytes);
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) return ldsc::fail(err, err_len, LDSC_ERR_INVALID_ARG, "server is stopping");
        if (s->queue.size() >= s->cfg.queue_capacity)
            return ldsc::fail(err, err_len, LDSC_ERR_BUSY, "request queue is full (%u)",
                              s->cfg.queue_capacity);
        r->id = s->next_id++;
        *request_id = r->id;
        s->queue.push_back(r.get());
        s->reqs.emplace(r->id, std::move(r));
    }
    s->cv_work.notify_one();
    return LDSC_OK;
}

static int collect_locked(ldsc_server* s, uint64_t id, double* latency, char* err,
                          uint64_t err_len) {
    auto it = s->reqs.find(id);
    Request* r = it->second.get();
    if (latency) *latency = r->t_done - r->t_arrival;
    int rc = r->rc;
    if (rc != LDSC_OK) ldsc::fail(err, err_len, rc, "%s", r->err.c_str());
    s->reqs.erase(it);
    return rc;
}

extern "C" int ldsc_l2_poll(ldsc_server* s, uint64_t id, double* latency, char* err,
                            uint64_t err_len) {
    if (!s) return ldsc::fail(err, err_len, LDSC_ERR_INVALID_ARG, "NULL server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end())
        return ldsc::fail(err, err_len, LDSC_ERR_NOT_FOUND, "unknown request %llu",
                          (unsigned long long)id);
    s->expire_locked(now_s());
    if (it->second->state != 2) return LDSC_ERR_PENDING;
    return collect_locked(s, id, latency, err, err_len);
}

extern "C" int ldsc_l2_wait(ldsc_server* s, uint64_t id, double* latency, char* err,
                            uint64_t err_len) {
    if (!s) return ldsc::fail(err, err_len, LDSC_ERR_INVALID_ARG, "NULL server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end())
        return ldsc::fail(err, err_len, LDSC_ERR_NOT_FOUND, "unknown request %llu",
                          (unsigned long long)id);
    Request* r = it->second.get();
    // a queued request's deadline is enforced here too, whether or not a
    // worker looks at the queue meanwhile
    while (r->state != 2) {
        if (s->cfg.max_queue_ms && r->state == 0) {
            const double due = r->t_arrival + s->cfg.max_queue_ms * 1e-3;
            s->cv_done.wait_for(lk, std::chrono::duration<double>(std::max(0.0, due - now_s()) + 1e-4));
            s->expire_locked(now_s());
        } else {
            s->cv_done.wait(lk);
        }
    }
    return collect_locked(s, id, latency, err, err_len);
}

extern "C" int ldsc_l2_cancel(ldsc_server* s, uint64_t id, char* err, uint64_t err_len) {
    if (!s) return ldsc::fail(err, err_len, LDSC_ERR_INVALID_ARG, "NULL server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->reqs.find(id);
    if (it == s->reqs.end())
        return ldsc::fail(err, err_len, LDSC_ERR_NOT_FOUND, "unknown request %llu",
                          (unsigned long long)id);
    Request* r = it->second.get();
    if (r->state == 0) {
        for (auto q = s->queue.begin(); q != s->queue.end(); ++q)
            if (*q == r) {
                s->queue.erase(q);
                break;
            }
        s->finish_locked(r, LDSC_ERR_CANCELLED, "request cancelled while queued");
        ++s->cancelled;
        s->cv_done.notify_all();
        s->cv_work.notify_all();
    } else if (r->state == 1) {
        r->cancel.store(true);
    }
    return LDSC_OK;
}
