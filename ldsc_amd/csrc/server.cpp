// server.cpp -- continuous batching of LD-score requests over resident panels.
//
// No counterpart in the reference (its nearest analogue is one
// compute_l2_from_bed call per chromosome shard from a browser worker,
// ldsc-web/src/worker.rs:664-872).
//
// Requests enter a bounded queue.  `max_inflight` worker threads each own a
// device workspace and HIP stream; a worker takes the oldest queued request
// whose device-memory estimate fits the remaining HBM budget (first fit in
// arrival order; once the oldest request has waited longer than
// kStarveSecs, nothing may overtake it), runs it on its own stream and
// returns its budget.  Requests that arrive while others run are admitted as
// soon as a slot and memory free up, and their kernels overlap on the device
// with those already running.  Each request's result is independent of what
// else is in flight: the computation of one request touches only its own
// workspace, and every reduction runs in a fixed order.
#include <chrono>
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
    size_t bytes = 0;
    double t_arrival = 0, t_done = 0;
    int state = 0;  // 0 queued, 1 running, 2 done
    int rc = LDSC_OK;
    std::string err;
};

}  // namespace

struct ldsc_server {
    ldsc_server_config cfg{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Request*> queue;
    std::unordered_map<uint64_t, std::unique_ptr<Request>> reqs;
    uint64_t next_id = 1;
    size_t in_use = 0;
    bool stop = false;
    std::vector<std::thread> workers;

    void worker_main();
    Request* take_locked();
};

Request* ldsc_server::take_locked() {
    if (queue.empty()) return nullptr;
    const size_t room = cfg.hbm_budget_bytes > in_use ? cfg.hbm_budget_bytes - in_use : 0;
    const bool head_starving = now_s() - queue.front()->t_arrival > kStarveSecs;
    for (auto it = queue.begin(); it != queue.end(); ++it) {
        if ((*it)->bytes <= room) {
            Request* r = *it;
            queue.erase(it);
            return r;
        }
        if (head_starving) break;
    }
    return nullptr;
}

void ldsc_server::worker_main() {
    ldsc::Workspace* ws = nullptr;
    int ws_device = -1;
    for (;;) {
        Request* r = nullptr;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv_work.wait(lk, [&] { return stop || (r = take_locked()) != nullptr; });
            if (!r) break;  // stop
            in_use += r->bytes;
            r->state = 1;
        }
        if (!ws || ws_device != r->panel->device) {
            if (ws) ldsc::release_workspace(ws);
            ws = ldsc::acquire_workspace(r->panel->device);
            ws_device = r->panel->device;
        }
        char err[512] = {0};
        int rc;
        if (!ws) {
            rc = ldsc::fail(err, sizeof(err), LDSC_ERR_DEVICE, "could not create a device workspace");
        } else {
            rc = ldsc::run_l2(r->panel, r->snps.data(), r->snps.size(), &r->cfg,
                              r->annot.empty() ? nullptr : r->annot.data(), r->K, r->l2_out,
                              r->maf_out, nullptr, nullptr, nullptr, ws, err, sizeof(err));
        }
        {
            std::lock_guard<std::mutex> lk(mu);
            in_use -= r->bytes;
            r->rc = rc;
            r->err = err;
            r->t_done = now_s();
            r->state = 2;
        }
        cv_done.notify_all();
        cv_work.notify_all();
    }
    if (ws) ldsc::release_workspace(ws);
}

extern "C" int ldsc_server_create(const ldsc_server_config* cfg, ldsc_server** out, char* err,
                                  uint64_t err_len) {
    if (!cfg || !out) return ldsc::fail(err, err_len, LDSC_ERR_INVALID_ARG, "NULL argument");
    if (cfg->max_inflight < 1 || cfg->max_inflight > 64)
        return ldsc::fail(err, err_len, LDSC_ERR_INVALID_ARG, "max_inflight must be in [1, 64]");
    if (cfg->queue_capacity < 1)
        return ldsc::fail(err, err_len, LDSC_ERR_INVALID_ARG, "queue_capacity must be >= 1");
    auto* s = new ldsc_server();
    s->cfg = *cfg;
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
        for (Request* r : s->queue) {
            r->state = 2;
            r->rc = LDSC_ERR_INVALID_ARG;
            r->err = "server destroyed before the request ran";
        }
        s->queue.clear();
    }
    s->cv_work.notify_all();
    for (auto& t : s->workers) t.join();
    delete s;
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
    r->bytes = ldsc::l2_workspace_bytes(panel, m, cfg, K);
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

static int finish_locked(ldsc_server* s, uint64_t id, double* latency, char* err,
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
    if (it->second->state != 2) return LDSC_ERR_PENDING;
    return finish_locked(s, id, latency, err, err_len);
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
    s->cv_done.wait(lk, [&] { return r->state == 2; });
    return finish_locked(s, id, latency, err, err_len);
}
