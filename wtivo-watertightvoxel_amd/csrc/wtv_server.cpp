// wtv_server.cpp -- serving loop around the solver (no reference counterpart).
//
// Requests are reduced graphs already resident in HBM.  A server runs S arena
// sessions (session_serve, wtv_solver.hip), each a host thread with its own
// HIP stream (default priority) and its own node arena in HBM.  A session
// admits waiting requests at its global relabels: continuous batching, so a
// request that arrives while others are being solved joins them on the
// device at that session's next relabel, as arena nodes and graph slots
// allow, and an idle session takes it at once.  Each graph leaves at the
// first relabel after which it has no active vertex: its partition,
// residuals and terminals are written to the request's buffers.  Admission
// is FIFO by arena footprint; the arenas together stay within the HBM
// budget.  Results do not depend on the mix: each graph's flow and labels are
// those of a lone solve.  A request whose graph is invalid fails alone at
// admission; if a shared solve fails, its requests are solved again one by
// one, so only the request at fault reports the error.
#include <hip/hip_runtime.h>

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

#include "wtv_internal.h"

namespace wtv {
namespace {

inline void cks(hipError_t e, const char* where) {
    if (e == hipSuccess) return;
    fail(WTV_EDEVICE, std::string(kPrefix) + "HIP at " + where + ": " + hipGetErrorString(e));
}

inline uint64_t aligned_nodes(uint64_t n) { return (n + 1023u) / 1024u * 1024u; }

struct Ticket {
    hipEvent_t ready = nullptr;  // recorded on the request's stream at submit
    bool done = false;
    wtv_result res{};
    ~Ticket() { if (ready) (void)hipEventDestroy(ready); }
};

}  // namespace

uint64_t request_footprint(uint32_t n, bool pull) {
    // the arena nodes a request occupies (rounded up to the seed sweep's
    // 1024-node tile) times the arena bytes per node
    return aligned_nodes(n) * arena_bytes_per_node(pull);
}

class Server final : public SessionQueue {
  public:
    explicit Server(const wtv_server_config& c) : cfg_(c) {
        if (cfg_.slots <= 0) cfg_.slots = 4;
        if (cfg_.max_batch <= 0) cfg_.max_batch = 64;
        if (cfg_.max_batch_nodes == 0) cfg_.max_batch_nodes = 1ull << 27;
        if (cfg_.max_batch_nodes > 0x7FFFFFFFull) cfg_.max_batch_nodes = 0x7FFFFFFFull;
        if (cfg_.hbm_budget_bytes == 0) cfg_.hbm_budget_bytes = 200ull * 1000ull * 1000ull * 1000ull;
        // each session's arena: at most max_batch_nodes, and all of them
        // together within the HBM budget
        const uint64_t per =
            cfg_.hbm_budget_bytes / static_cast<uint64_t>(cfg_.slots) / arena_bytes_per_node(pull_enabled(cfg_.params));
        capacity_ = std::min<uint64_t>(cfg_.max_batch_nodes, per) / 1024u * 1024u;
        if (capacity_ < 1024u) fail(WTV_ENOMEM, std::string(kPrefix) + "HBM budget below one arena tile");
        cks(hipGetDevice(&device_), "hipGetDevice");
        t0_ = std::chrono::steady_clock::now();
        live_ = cfg_.slots;
        for (int i = 0; i < cfg_.slots; ++i) workers_.emplace_back([this, i]() { worker(i); });
    }
    ~Server() override {
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
        if (aligned_nodes(r.n) > capacity_)
            fail(WTV_ENOMEM, std::string(kPrefix) + "request exceeds the HBM budget of one session's arena");
        auto t = std::make_shared<Ticket>();
        if (r.stream) {
            // the solve waits for what the client queued on its stream so far
            cks(hipEventCreateWithFlags(&t->ready, hipEventDisableTiming), "ready event");
            cks(hipEventRecord(t->ready, static_cast<hipStream_t>(r.stream)), "ready record");
        }
        std::lock_guard<std::mutex> lk(mu_);
        if (stop_) fail(WTV_EINVAL, std::string(kPrefix) + "server is stopping");
        if (live_ == 0) fail(WTV_ENOMEM, std::string(kPrefix) + "no serving session could allocate its arena");
        SessionJob j;
        j.ticket = next_ticket_++;
        j.req = r;
        j.ready = t->ready;
        t->res.footprint_bytes = request_footprint(r.n, pull_enabled(cfg_.params));
        t->res.submit_s = now();
        tickets_[j.ticket] = t;
        queue_.push_back(j);
        cv_work_.notify_all();
        return j.ticket;
    }

    wtv_result wait(uint64_t ticket) {
        std::unique_lock<std::mutex> lk(mu_);
        auto it = tickets_.find(ticket);
        if (it == tickets_.end()) fail(WTV_EINVAL, std::string(kPrefix) + "unknown ticket");
        auto t = it->second;
        cv_done_.wait(lk, [&]() { return t->done; });
        tickets_.erase(ticket);
        return t->res;
    }

    // ---- SessionQueue
    std::vector<SessionJob> take(uint64_t free_nodes, uint32_t free_slots, bool resident_any) override {
        std::lock_guard<std::mutex> lk(mu_);
        std::vector<SessionJob> out;
        scan(free_nodes, free_slots, resident_any, &out);
        return out;
    }
    bool waiting(uint64_t free_nodes, uint32_t free_slots, bool resident_any) override {
        std::lock_guard<std::mutex> lk(mu_);
        return scan(free_nodes, free_slots, resident_any, nullptr);
    }
    bool wait_nonempty() override {
        std::unique_lock<std::mutex> lk(mu_);
        cv_work_.wait(lk, [&]() { return stop_ || !queue_.empty(); });
        return !queue_.empty();
    }
    void finish(uint64_t ticket, int status, int64_t flow, int32_t batch_size, int32_t slot, double start_s,
                double end_s) override {
        {
            std::lock_guard<std::mutex> lk(mu_);
            auto it = tickets_.find(ticket);
            if (it == tickets_.end()) return;
            wtv_result& r = it->second->res;
            r.status = status;
            r.batch_size = batch_size;
            r.slot = slot;
            r.raw_flow = flow;
            r.start_s = std::max(start_s, r.submit_s);
            r.end_s = end_s;
            it->second->done = true;
        }
        cv_done_.notify_all();
    }
    void requeue_alone(const std::vector<SessionJob>& jobs) override {
        {
            std::lock_guard<std::mutex> lk(mu_);
            for (size_t i = jobs.size(); i-- > 0;) {
                SessionJob j = jobs[i];
                j.alone = true;
                queue_.push_front(j);
            }
        }
        cv_work_.notify_all();
    }
    double now() const override {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
    }

  private:
    // Oldest-first requests that fit one after another into free_nodes and
    // free_slots (into *out when given); whether there is at least one.
    bool scan(uint64_t free_nodes, uint32_t free_slots, bool resident_any, std::vector<SessionJob>* out) {
        uint64_t room = free_nodes;
        uint32_t slots = std::min<uint32_t>(free_slots, static_cast<uint32_t>(cfg_.max_batch));
        bool any = false;
        while (!queue_.empty() && slots > 0) {
            const SessionJob& j = queue_.front();
            if (j.alone) {
                // solved with nothing else resident
                if (resident_any || any) break;
                if (aligned_nodes(j.req.n) > room) break;
                any = true;
                if (!out) break;
                out->push_back(j);
                queue_.pop_front();
                break;
            }
            const uint64_t need = aligned_nodes(j.req.n);
            if (need > room) break;
            room -= need;
            --slots;
            any = true;
            if (!out) break;
            out->push_back(j);
            queue_.pop_front();
        }
        return any;
    }

    void worker(int slot) {
        (void)hipSetDevice(device_);
        hipStream_t s = nullptr;
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess) s = nullptr;
        try {
            session_serve(*this, cfg_.params, static_cast<uint32_t>(capacity_),
                          static_cast<uint32_t>(cfg_.max_batch), s, slot);
        } catch (const Error& e) {
            // the session could not start (its arena did not fit): the
            // others serve the queue; with none left, the waiting requests
            // fail instead of waiting forever
            std::vector<uint64_t> dropped;
            {
                std::lock_guard<std::mutex> lk(mu_);
                if (--live_ == 0) {
                    for (const SessionJob& j : queue_) dropped.push_back(j.ticket);
                    queue_.clear();
                }
            }
            for (uint64_t t : dropped) finish(t, e.code, 0, 0, slot, now(), now());
        }
        if (s) (void)hipStreamDestroy(s);
    }

    wtv_server_config cfg_;
    uint64_t capacity_ = 0;  // arena nodes per session
    int device_ = 0;
    std::chrono::steady_clock::time_point t0_;
    std::mutex mu_;
    std::condition_variable cv_work_, cv_done_;
    std::deque<SessionJob> queue_;
    std::map<uint64_t, std::shared_ptr<Ticket>> tickets_;
    uint64_t next_ticket_ = 1;
    int live_ = 0;  // sessions serving the queue
    bool stop_ = false;
    std::vector<std::thread> workers_;
};

Server* server_new(const wtv_server_config& c) { return new Server(c); }
void server_delete(Server* s) { delete s; }
uint64_t server_submit(Server* s, const wtv_request& r) { return s->submit(r); }
wtv_result server_wait(Server* s, uint64_t ticket) { return s->wait(ticket); }

}  // namespace wtv
