// wtv_server.cpp -- serving loop around the solver (no reference counterpart).
//
// Requests are reduced graphs already resident in HBM.  A server owns S worker
// slots; each slot is a host thread with its own HIP stream (default priority).
// Admission is FIFO: when a slot is free it takes the oldest waiting requests
// that fit max_batch, max_batch_nodes and the HBM budget (footprint of the
// graphs plus the solver workspace of the batch), and solves them together:
// one request goes straight to the plain solver, several are gathered into
// contiguous buffers, solved as one union (wtv_batch_solve semantics, every
// graph keeps its own source height) and scattered back.  Requests that arrive
// while other batches are on the device are admitted as slots and memory free
// up, and batches in different slots overlap on the device.  Results do not
// depend on the mix: each graph's flow and labels are those of a lone solve.
#include <hip/hip_runtime.h>

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

#include "wtv_internal.h"

namespace wtv {
namespace {

inline void cks(hipError_t e, const char* where) {
    if (e == hipSuccess) return;
    fail(WTV_EDEVICE, std::string(kPrefix) + "HIP at " + where + ": " + hipGetErrorString(e));
}

struct Job {
    uint64_t ticket;
    wtv_request req;
    uint64_t footprint;
    double submit_s;
    hipEvent_t ready = nullptr;  // recorded on req.stream at submit
    bool done = false;
    wtv_result res{};
    ~Job() { if (ready) (void)hipEventDestroy(ready); }
};

struct DevMem {
    void* p = nullptr;
    ~DevMem() { if (p) (void)hipFree(p); }
};

// Waits for the stream on scope exit, also when an exception unwinds: the
// gather buffers declared before it must outlive every copy queued on it.
struct SyncOnExit {
    hipStream_t s;
    ~SyncOnExit() { (void)hipStreamSynchronize(s); }
};

}  // namespace

uint64_t request_footprint(uint32_t n) {
    // graph 36 B + plain-solver workspace 26 B per node + batched extras
    // (graph id 4, union ids 16, gathered nbr/res/term/partition 37) + slack
    return static_cast<uint64_t>(n) * (36ull + 26ull + 20ull + 37ull) + (1ull << 20);
}

class Server {
  public:
    explicit Server(const wtv_server_config& c) : cfg_(c) {
        // four slots: concurrent latency-bound solves overlap on the device; more
        // than the four hardware queues of a process did not lower latency further
        if (cfg_.slots <= 0) cfg_.slots = 4;
        if (cfg_.max_batch <= 0) cfg_.max_batch = 64;
        if (cfg_.max_batch_nodes == 0) cfg_.max_batch_nodes = 1ull << 30;
        if (cfg_.max_batch_nodes > 0x7FFFFFFFull) cfg_.max_batch_nodes = 0x7FFFFFFFull;
        if (cfg_.hbm_budget_bytes == 0) cfg_.hbm_budget_bytes = 200ull * 1000ull * 1000ull * 1000ull;
        cks(hipGetDevice(&device_), "hipGetDevice");
        t0_ = std::chrono::steady_clock::now();
        for (int i = 0; i < cfg_.slots; ++i) workers_.emplace_back([this, i]() { worker(i); });
    }
    ~Server() {
        {
            std::lock_guard<std::mutex> lk(mu_);
            stop_ = true;
        }
        cv_work_.notify_all();
        for (auto& t : workers_) t.join();
    }

    uint64_t submit(const wtv_request& r) {
        if (r.n && (!r.d_nbr || !r.d_res || !r.d_term || !r.d_partition))
            fail(WTV_EINVAL, std::string(kPrefix) + "null request pointer");
        if (r.n > 0x7FFFFFFFu) fail(WTV_EINVAL, std::string(kPrefix) + ">INT32_MAX reduced nodes unsupported");
        const uint64_t fp = request_footprint(r.n);
        if (fp > cfg_.hbm_budget_bytes) fail(WTV_ENOMEM, std::string(kPrefix) + "request exceeds the HBM budget");
        std::lock_guard<std::mutex> lk(mu_);
        if (stop_) fail(WTV_EINVAL, std::string(kPrefix) + "server is stopping");
        auto job = std::make_shared<Job>();
        if (r.stream) {
            // the solve waits for what the client queued on its stream so far
            cks(hipEventCreateWithFlags(&job->ready, hipEventDisableTiming), "ready event");
            cks(hipEventRecord(job->ready, static_cast<hipStream_t>(r.stream)), "ready record");
        }
        job->ticket = next_ticket_++;
        job->req = r;
        job->footprint = fp;
        job->submit_s = now();
        jobs_[job->ticket] = job;
        queue_.push_back(job);
        cv_work_.notify_one();
        return job->ticket;
    }

    wtv_result wait(uint64_t ticket) {
        std::unique_lock<std::mutex> lk(mu_);
        auto it = jobs_.find(ticket);
        if (it == jobs_.end()) fail(WTV_EINVAL, std::string(kPrefix) + "unknown ticket");
        auto job = it->second;
        cv_done_.wait(lk, [&]() { return job->done; });
        jobs_.erase(ticket);
        return job->res;
    }

  private:
    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
    }

    // Oldest-first admission of everything that fits; called with mu_ held.
    std::vector<std::shared_ptr<Job>> take_batch() {
        std::vector<std::shared_ptr<Job>> b;
        uint64_t nodes = 0, bytes = 0;
        while (!queue_.empty() && static_cast<int>(b.size()) < cfg_.max_batch) {
            auto& j = queue_.front();
            if (!b.empty() && nodes + j->req.n > cfg_.max_batch_nodes) break;
            if (used_bytes_ + bytes + j->footprint > cfg_.hbm_budget_bytes) break;
            nodes += j->req.n;
            bytes += j->footprint;
            b.push_back(j);
            queue_.pop_front();
        }
        used_bytes_ += bytes;
        return b;
    }

    void worker(int slot) {
        (void)hipSetDevice(device_);
        hipStream_t s = nullptr;
        (void)hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
        for (;;) {
            std::vector<std::shared_ptr<Job>> batch;
            {
                std::unique_lock<std::mutex> lk(mu_);
                cv_work_.wait(lk, [&]() {
                    if (stop_ && queue_.empty()) return true;
                    if (queue_.empty()) return false;
                    return used_bytes_ + queue_.front()->footprint <= cfg_.hbm_budget_bytes;
                });
                if (queue_.empty()) break;
                batch = take_batch();
            }
            const double start = now();
            std::vector<int64_t> flows(batch.size(), 0);
            std::vector<int> status(batch.size(), WTV_OK);
            const int st = attempt([&]() { run(batch, flows, s); });
            if (st != WTV_OK) {
                if (batch.size() == 1) {
                    status[0] = st;
                } else {
                    // a failed batch leaves every request untouched (results are
                    // scattered back only on success): solve each alone, so only
                    // the request at fault reports the error
                    for (size_t i = 0; i < batch.size(); ++i) {
                        std::vector<std::shared_ptr<Job>> one{batch[i]};
                        std::vector<int64_t> f1(1, 0);
                        status[i] = attempt([&]() { run(one, f1, s); });
                        flows[i] = f1[0];
                    }
                }
            }
            const double end = now();
            {
                std::lock_guard<std::mutex> lk(mu_);
                for (size_t i = 0; i < batch.size(); ++i) {
                    auto& j = batch[i];
                    used_bytes_ -= j->footprint;
                    j->res.status = status[i];
                    j->res.batch_size = static_cast<int32_t>(batch.size());
                    j->res.slot = slot;
                    j->res.raw_flow = flows[i];
                    j->res.footprint_bytes = j->footprint;
                    j->res.submit_s = j->submit_s;
                    j->res.start_s = start;
                    j->res.end_s = end;
                    j->done = true;
                }
            }
            cv_done_.notify_all();
            cv_work_.notify_all();
        }
        if (s) (void)hipStreamDestroy(s);
    }

    template <class F>
    static int attempt(F&& f) {
        try {
            f();
            return WTV_OK;
        } catch (const Error& e) {
            return e.code;
        } catch (...) {
            return WTV_EINVAL;
        }
    }

    void run(const std::vector<std::shared_ptr<Job>>& batch, std::vector<int64_t>& flows, hipStream_t s) {
        wtv_stats st{};
        for (const auto& j : batch)
            if (j->ready) cks(hipStreamWaitEvent(s, j->ready, 0), "ready wait");
        if (batch.size() == 1) {
            const wtv_request& r = batch[0]->req;
            pr_solve(r.n, r.d_nbr, r.d_res, r.d_term, r.d_partition, true, cfg_.params, st, s);
            flows[0] = st.raw_flow;
            return;
        }
        const uint32_t G = static_cast<uint32_t>(batch.size());
        std::vector<uint64_t> off(G + 1, 0);
        for (uint32_t g = 0; g < G; ++g) off[g + 1] = off[g] + batch[g]->req.n;
        const size_t N = off[G];
        DevMem nbr, res, term, part;
        SyncOnExit sync{s};
        cks(hipMalloc(&nbr.p, std::max<size_t>(N, 1) * 16u), "gather nbr");
        cks(hipMalloc(&res.p, std::max<size_t>(N, 1) * 16u), "gather res");
        cks(hipMalloc(&term.p, std::max<size_t>(N, 1) * 4u), "gather term");
        cks(hipMalloc(&part.p, std::max<size_t>(N, 1)), "gather partition");
        char* pn = static_cast<char*>(nbr.p);
        char* pr = static_cast<char*>(res.p);
        char* pt = static_cast<char*>(term.p);
        char* pp = static_cast<char*>(part.p);
        for (uint32_t g = 0; g < G; ++g) {
            const wtv_request& r = batch[g]->req;
            if (!r.n) continue;
            cks(hipMemcpyAsync(pn + off[g] * 16u, r.d_nbr, r.n * 16ull, hipMemcpyDeviceToDevice, s), "gather");
            cks(hipMemcpyAsync(pr + off[g] * 16u, r.d_res, r.n * 16ull, hipMemcpyDeviceToDevice, s), "gather");
            cks(hipMemcpyAsync(pt + off[g] * 4u, r.d_term, r.n * 4ull, hipMemcpyDeviceToDevice, s), "gather");
        }
        batch_solve(G, off.data(), reinterpret_cast<int32_t*>(nbr.p), reinterpret_cast<uint32_t*>(res.p),
                    reinterpret_cast<int32_t*>(term.p), reinterpret_cast<uint8_t*>(part.p), true, flows.data(),
                    cfg_.params, st, s);
        for (uint32_t g = 0; g < G; ++g) {
            const wtv_request& r = batch[g]->req;
            if (!r.n) continue;
            cks(hipMemcpyAsync(r.d_res, pr + off[g] * 16u, r.n * 16ull, hipMemcpyDeviceToDevice, s), "scatter");
            cks(hipMemcpyAsync(r.d_term, pt + off[g] * 4u, r.n * 4ull, hipMemcpyDeviceToDevice, s), "scatter");
            cks(hipMemcpyAsync(r.d_partition, pp + off[g], r.n, hipMemcpyDeviceToDevice, s), "scatter");
        }
        cks(hipStreamSynchronize(s), "scatter sync");
    }

    wtv_server_config cfg_;
    int device_ = 0;
    std::chrono::steady_clock::time_point t0_;
    std::mutex mu_;
    std::condition_variable cv_work_, cv_done_;
    std::deque<std::shared_ptr<Job>> queue_;
    std::map<uint64_t, std::shared_ptr<Job>> jobs_;
    uint64_t next_ticket_ = 1;
    uint64_t used_bytes_ = 0;
    bool stop_ = false;
    std::vector<std::thread> workers_;
};

Server* server_new(const wtv_server_config& c) { return new Server(c); }
void server_delete(Server* s) { delete s; }
uint64_t server_submit(Server* s, const wtv_request& r) { return s->submit(r); }
wtv_result server_wait(Server* s, uint64_t ticket) { return s->wait(ticket); }

}  // namespace wtv
