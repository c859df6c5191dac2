// Serving: a request loop with continuous batching over the batched passes.
//
// No counterpart in the reference, which runs one script on one AIG per
// process and copies the AIG to the host around every rw (aig_manager.cu:
// 647-672).  Here a worker thread owns one device context.  Jobs (an AIG + a
// script such as "b; rw; rwz; b") queue on submission; the worker admits them
// in arrival order while their estimated HBM footprint fits the budget, keeps
// each admitted AIG on the device for its whole script, and on every round
// runs each command kind once as a batched pass over all active jobs whose
// next command it is.  Jobs that arrive while others run join at the next
// round.  A job's result equals running its script alone (the batched passes
// are exact per member).
#include "gls_internal.h"
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

namespace {

enum Cmd : int { CMD_ST = 0, CMD_RW = 1, CMD_RWZ = 2, CMD_B = 3, CMD_KINDS = 4 };

// device bytes per object while a pass runs (cut sets, tables, scratch), with headroom
constexpr int64_t kBytesPerObj = 640;

bool parse_script(const char *script, std::vector<int> &cmds, std::string &err) {
    cmds.clear();
    std::string s = script ? script : "";
    size_t i = 0;
    while (i <= s.size()) {
        size_t j = s.find(';', i);
        if (j == std::string::npos) j = s.size();
        std::string c = s.substr(i, j - i);
        // normalise blanks
        std::string t;
        bool sp = false;
        for (char ch : c) {
            if (ch == ' ' || ch == '\t' || ch == '\n') { sp = !t.empty(); continue; }
            if (sp) t += ' ';
            sp = false;
            t += ch;
        }
        if (!t.empty()) {
            if (t == "st" || t == "strash") cmds.push_back(CMD_ST);
            else if (t == "rw" || t == "rewrite") cmds.push_back(CMD_RW);
            else if (t == "rwz" || t == "rw -z" || t == "rewrite -z") cmds.push_back(CMD_RWZ);
            else if (t == "b" || t == "balance") cmds.push_back(CMD_B);
            else { err = "unknown command '" + t + "'"; return false; }
        }
        i = j + 1;
    }
    return true;
}

} // namespace

struct gls_job {
    int64_t id = 0;
    std::vector<int> cmds;
    size_t pc = 0;
    int32_t nPIs = 0;
    std::vector<uint32_t> f0, f1, pos;    // input until admitted, result when done (AIGER literals)
    gls_aig *aig = nullptr;
    gls_job_info info = {};   // read by clients: written only under gls_server::mu
    // worker-only state, published into info under the lock when the job ends
    int32_t wstatus = 1;
    gls_stats wres = {};
};

struct gls_server {
    gls_ctx *ctx = nullptr;
    int64_t budget = 0, used = 0;
    int32_t maxActive = 0;
    std::mutex mu;
    std::condition_variable cv;      // worker: new work / stop
    std::condition_variable doneCv;  // waiters: a job finished
    bool running = true;
    int64_t nextId = 1;
    std::deque<std::shared_ptr<gls_job>> pending;
    std::vector<std::shared_ptr<gls_job>> active;
    std::map<int64_t, std::shared_ptr<gls_job>> jobs;
    std::string err;
    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    std::thread worker;
    double now_ms() const {
        return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    void run();
};

void gls_server::run() {
    for (;;) {
        std::vector<std::shared_ptr<gls_job>> admitted;
        {
            std::unique_lock<std::mutex> lk(mu);
            cv.wait(lk, [&] { return !running || !pending.empty() || !active.empty(); });
            if (!running && pending.empty() && active.empty()) return;
            // admission in arrival order, by estimated device footprint
            while (!pending.empty() && (int32_t)active.size() < maxActive) {
                auto &j = pending.front();
                if (!active.empty() && used + j->info.footprint_bytes > budget) break;
                used += j->info.footprint_bytes;
                j->info.status = 1;
                j->info.admit_ms = now_ms();
                active.push_back(j);
                admitted.push_back(j);
                pending.pop_front();
            }
        }
        // upload the admitted jobs (arrays checked and converted at submit,
        // taken over by the handle; one stream synchronisation for all of
        // them): they stay on the device until done
        for (auto &j : admitted) {
            j->aig = aig_adopt(ctx, j->nPIs, j->f0, j->f1, j->pos);
            const int rc = aig_upload_async(j->aig);
            if (rc != GLS_OK) { j->wstatus = rc; j->pc = j->cmds.size(); }
        }
        if (!admitted.empty() && hipStreamSynchronize(ctx->stream) != hipSuccess)
            for (auto &j : admitted) { j->wstatus = GLS_ERR_HIP; j->pc = j->cmds.size(); }
        for (auto &j : admitted)
            if (j->aig) {
                // the host copy is stale once a pass runs; free it now
                std::vector<uint32_t>().swap(j->aig->h_f0);
                std::vector<uint32_t>().swap(j->aig->h_f1);
                j->aig->host_valid = false;
            }
        // one batched pass per command kind over the jobs whose next command it is
        std::vector<gls_aig *> group[CMD_KINDS];
        std::vector<gls_job *> owners[CMD_KINDS];
        for (auto &j : active) {
            if (j->pc >= j->cmds.size() || j->wstatus < 0) continue;
            group[j->cmds[j->pc]].push_back(j->aig);
            owners[j->cmds[j->pc]].push_back(j.get());
        }
        for (int k = 0; k < CMD_KINDS; k++) {
            if (group[k].empty()) continue;
            const int32_t cnt = (int32_t)group[k].size();
            int rc = k == CMD_ST    ? gls_batch_strash(ctx, group[k].data(), cnt, 0)
                     : k == CMD_RW  ? gls_batch_rewrite(ctx, group[k].data(), cnt, 0, 0)
                     : k == CMD_RWZ ? gls_batch_rewrite(ctx, group[k].data(), cnt, 1, 0)
                                    : gls_batch_balance(ctx, group[k].data(), cnt, 1, 1);
            for (gls_job *j : owners[k]) {
                if (rc != GLS_OK) { j->wstatus = rc; j->pc = j->cmds.size(); }
                else j->pc++;
            }
            if (rc != GLS_OK) {
                std::lock_guard<std::mutex> lk(mu);
                err = gls_last_error(ctx);
            }
        }
        // finished jobs: results to the host, device memory back to the budget
        std::vector<std::shared_ptr<gls_job>> finished;
        for (auto &j : active)
            if (j->pc >= j->cmds.size()) finished.push_back(j);
        // the result arrays are filled before the job is published as done;
        // clients read them only after seeing status 2 under the lock
        for (auto &j : finished) {
            if (j->wstatus >= 0 && j->aig) {
                gls_stats &s = j->wres;
                gls_aig_stats(j->aig, &s);
                j->f0.resize(s.nObjs);
                j->f1.resize(s.nObjs);
                j->pos.resize(s.nPOs);
                int rc = gls_aig_to_arrays(j->aig, j->f0.data(), j->f1.data(), j->pos.data(), GLS_LIT_AIGER);
                j->wstatus = rc == GLS_OK ? 2 : rc;
            }
            if (j->aig) gls_aig_free(j->aig);
            j->aig = nullptr;
        }
        {
            std::lock_guard<std::mutex> lk(mu);
            for (auto &j : finished) {
                if (j->wstatus == 2) {
                    j->info.nPIs = j->wres.nPIs;
                    j->info.nPOs = j->wres.nPOs;
                    j->info.nAnds = j->wres.nAnds;
                    j->info.levels = j->wres.levels;
                }
                j->info.status = j->wstatus;
                j->info.done_ms = now_ms();
                used -= j->info.footprint_bytes;
                active.erase(std::find(active.begin(), active.end(), j));
            }
        }
        if (!finished.empty()) doneCv.notify_all();
    }
}

extern "C" int gls_server_create(int device, int64_t hbm_budget_bytes, int32_t max_active, gls_server **out) {
    if (!out || hbm_budget_bytes <= 0 || max_active <= 0) return GLS_ERR_ARG;
    gls_server *s = new gls_server();
    int rc = gls_ctx_create(device, &s->ctx);
    if (rc != GLS_OK) { delete s; return rc; }
    if (!gls_ctx_has_device(s->ctx)) { gls_ctx_destroy(s->ctx); delete s; return GLS_ERR_NODEV; }
    s->budget = hbm_budget_bytes;
    s->maxActive = max_active;
    s->worker = std::thread([s] { s->run(); });
    *out = s;
    return GLS_OK;
}

extern "C" int gls_server_submit(gls_server *s, int32_t nPIs, int32_t nPOs, int32_t nAnds, const uint32_t *fanin0,
                                 const uint32_t *fanin1, const uint32_t *pos, int32_t lit_convention,
                                 const char *script, int64_t *job_out) {
    if (!s || !job_out || nPIs < 0 || nPOs < 0 || nAnds < 0 || (nAnds > 0 && (!fanin0 || !fanin1)) ||
        (nPOs > 0 && !pos))
        return GLS_ERR_ARG;
    auto j = std::make_shared<gls_job>();
    std::string perr;
    if (!parse_script(script, j->cmds, perr)) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->err = perr;
        return GLS_ERR_ARG;
    }
    const int32_t n = nPIs + nAnds + 1;
    j->nPIs = nPIs;
    j->f0.assign(n, 0);
    j->f1.assign(n, 0);
    j->pos.assign(nPOs, 0);
    auto conv = [&](uint32_t l) { return (lit_convention == GLS_LIT_CULS && l < 2) ? 1u - l : l; };
    for (int32_t v = nPIs + 1; v < n; v++) {
        uint32_t a = conv(fanin0[v]), b = conv(fanin1[v]);
        if ((int64_t)(a >> 1) >= v || (int64_t)(b >> 1) >= v) {
            std::lock_guard<std::mutex> lk(s->mu);
            s->err = "submit: AND " + std::to_string(v) + " has a fanin that is not smaller than itself";
            return GLS_ERR_ARG;
        }
        j->f0[v] = std::min(a, b);
        j->f1[v] = std::max(a, b);
    }
    for (int32_t i = 0; i < nPOs; i++) {
        j->pos[i] = conv(pos[i]);
        if ((int64_t)(j->pos[i] >> 1) >= n) {
            std::lock_guard<std::mutex> lk(s->mu);
            s->err = "submit: PO out of range";
            return GLS_ERR_ARG;
        }
    }
    j->info.nAnds_in = nAnds;
    j->info.footprint_bytes = (int64_t)n * kBytesPerObj;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (j->info.footprint_bytes > s->budget) {
            s->err = "submit: the job needs more device memory than the server's budget";
            return GLS_ERR_NOMEM;
        }
        j->id = s->nextId++;
        j->info.submit_ms = s->now_ms();
        s->jobs[j->id] = j;
        s->pending.push_back(j);
        *job_out = j->id;
    }
    s->cv.notify_one();
    return GLS_OK;
}

extern "C" int gls_server_wait(gls_server *s, int64_t job, double timeout_ms, gls_job_info *info) {
    if (!s) return GLS_ERR_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->jobs.find(job);
    if (it == s->jobs.end()) return GLS_ERR_ARG;
    auto j = it->second;
    auto finished = [&] { return j->info.status == 2 || j->info.status < 0; };
    if (timeout_ms < 0) s->doneCv.wait(lk, finished);
    else if (!s->doneCv.wait_for(lk, std::chrono::duration<double, std::milli>(timeout_ms), finished)) {
        if (info) *info = j->info;
        return GLS_ERR_TIMEOUT;
    }
    if (info) *info = j->info;
    return j->info.status < 0 ? j->info.status : GLS_OK;
}

extern "C" int gls_server_result(gls_server *s, int64_t job, uint32_t *fanin0, uint32_t *fanin1, uint32_t *pos,
                                 int32_t lit_convention) {
    if (!s) return GLS_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(job);
    if (it == s->jobs.end() || it->second->info.status != 2) return GLS_ERR_ARG;
    const gls_job &j = *it->second;
    auto conv = [&](uint32_t l) { return (lit_convention == GLS_LIT_CULS && l < 2) ? 1u - l : l; };
    for (size_t v = 0; v < j.f0.size(); v++) {
        const bool and_ = (int64_t)v > j.nPIs;
        if (fanin0) fanin0[v] = and_ ? conv(j.f0[v]) : 0u;
        if (fanin1) fanin1[v] = and_ ? conv(j.f1[v]) : 0u;
    }
    if (pos)
        for (size_t i = 0; i < j.pos.size(); i++) pos[i] = conv(j.pos[i]);
    return GLS_OK;
}

extern "C" int gls_server_release(gls_server *s, int64_t job) {
    if (!s) return GLS_ERR_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->jobs.find(job);
    if (it == s->jobs.end()) return GLS_ERR_ARG;
    if (it->second->info.status != 2 && it->second->info.status >= 0) return GLS_ERR_ARG;   // still queued or running
    s->jobs.erase(it);
    return GLS_OK;
}

// The text is copied per calling thread, so the worker replacing the
// server's message cannot invalidate what a client holds.
extern "C" const char *gls_server_last_error(gls_server *s) {
    static thread_local std::string copy;
    if (!s) return "";
    std::lock_guard<std::mutex> lk(s->mu);
    copy = s->err;
    return copy.c_str();
}

extern "C" void gls_server_destroy(gls_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->running = false;
    }
    s->cv.notify_all();
    if (s->worker.joinable()) s->worker.join();
    gls_ctx_destroy(s->ctx);
    delete s;
}
