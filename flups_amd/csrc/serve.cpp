// Serving loop and batches on one GPU. Neither has a counterpart in the
// reference, whose solver serves one problem per process.
//
// Server: solve requests of mixed shapes, boundary conditions, kernels and
// precisions are admitted as worker threads and device memory free up.
// Solvers (plan, Green's function on the device, stage arrays) are cached by
// the full problem description and evicted least-recently-used when the HBM
// budget needs room. Every cached solver has its own stream; a worker owns a
// solver for the duration of one solve, so requests that arrive while others
// run overlap on the device. A request whose solver would not fit waits; later
// requests may still run on solvers that are already cached (no new memory),
// so a large request is not starved by small ones.
//
// Batch: a fixed list of problems, set up once. Each problem's solve is recorded as a list of
// kernel launches (group.hpp); launches of the same kernel at the same step of different
// problems become one grouped launch whose blocks are mapped to the problems through an offset
// table, so a batch of 64 problems of 5-9 kernels each runs as a few hundred launches of
// blocks from many problems instead of one launch per problem and step. The grouped launches
// form one hipGraph (each depends on the launches of the previous step of its problems) that
// is replayed with a single launch. The previous form, one graph branch per problem, is kept
// for comparison (run_branches).
#define FLUPS_BUILDING_LIBRARY 1
#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "../../include/flups.h"
#include "group.hpp"
#include "solver.hpp"

namespace flups {

LaunchSink*& launch_sink() {
    thread_local LaunchSink* sink = nullptr;
    return sink;
}

namespace {

using Clock = std::chrono::steady_clock;

std::string key_of(const FLUPS_ServeRequest& r) {
    char b[640];
    std::snprintf(b, sizeof(b), "n%d,%d,%d c%d bc%d%d%d%d%d%d l%d h%a,%a,%a L%a,%a,%a g%d al%a p%d d%d", r.nglob[0],
                  r.nglob[1], r.nglob[2], r.center, r.bc[0][0], r.bc[0][1], r.bc[1][0], r.bc[1][1], r.bc[2][0], r.bc[2][1],
                  r.lda, r.h[0], r.h[1], r.h[2], r.L[0], r.L[1], r.L[2], r.green, r.alpha, r.fp32, r.order_diff);
    return b;
}

void validate(const FLUPS_ServeRequest& r) {
    FLUPS_REQUIRE(r.lda >= 1 && r.lda <= 3, "request lda must be 1..3");
    FLUPS_REQUIRE(r.type == 0 || r.type == 1, "request type must be STD (0) or ROT (1)");
    for (int d = 0; d < 3; ++d) FLUPS_REQUIRE(r.nglob[d] >= 1, "request nglob must be >= 1");
}

std::unique_ptr<SolverImpl> make_solver(const FLUPS_ServeRequest& r) {
    Topo t;
    t.axis = 0;
    t.lda = r.lda;
    t.nf = 1;
    t.alignment = sizeof(double);   // dense caller arrays: (lda, n2, n1, n0), no row padding
    for (int d = 0; d < 3; ++d) t.nglob[d] = r.nglob[d];
    t.cmpt_sizes();
    std::vector<int> bc(6 * r.lda);
    for (int d = 0; d < 3; ++d)
        for (int s = 0; s < 2; ++s)
            for (int c = 0; c < r.lda; ++c) bc[(d * 2 + s) * r.lda + c] = r.bc[d][s];
    const int center[3] = {r.center, r.center, r.center};
    auto s = std::make_unique<SolverImpl>(&t, bc.data(), r.h, r.L, r.order_diff, center);
    s->set_green(r.green);
    if (r.alpha > 0) s->set_alpha(r.alpha);
    s->set_fp32(r.fp32 != 0);
    return s;
}

struct Cached {
    std::string key;
    std::unique_ptr<SolverImpl> solver;
    hipStream_t stream = nullptr;
    size_t bytes = 0;   // reserved before setup, measured after
    bool busy = true;
    bool ready = false;
    uint64_t last_use = 0;
    ~Cached() {
        solver.reset();
        if (stream) (void)hipStreamDestroy(stream);
    }
};

struct Job {
    long id = 0;
    FLUPS_ServeRequest req{};
    std::string key;
    size_t need = 0;
    const double* rhs = nullptr;
    double* field = nullptr;
    Clock::time_point t_submit, t_done;
    int state = 0;   // 0 queued, 1 running, 2 done, 3 failed
    std::string error;
};

double pct(std::vector<double> v, double p) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double x = p / 100.0 * (double)(v.size() - 1);
    const size_t i = (size_t)x;
    const double f = x - (double)i;
    return i + 1 < v.size() ? v[i] * (1 - f) + v[i + 1] * f : v[i];
}

}  // namespace

class Server {
   public:
    Server(size_t budget, int nworkers, int per_key) : budget_(budget), per_key_(std::max(1, per_key)) {
        HIP_CHECK(hipGetDevice(&device_));
        for (int i = 0; i < std::max(1, nworkers); ++i) workers_.emplace_back([this] { work(); });
    }
    ~Server() {
        {
            std::lock_guard<std::mutex> lk(m_);
            stop_ = true;
            for (Job* j : queue_) {
                j->state = 3;
                j->error = "server shut down";
            }
            queue_.clear();
        }
        cv_work_.notify_all();
        cv_done_.notify_all();
        for (auto& t : workers_) t.join();
        cache_.clear();
    }

    long submit(const FLUPS_ServeRequest& r, const double* rhs, double* field) {
        validate(r);
        FLUPS_REQUIRE(rhs && field, "request buffers must not be NULL");
        auto j = std::make_unique<Job>();
        j->req = r;
        j->key = key_of(r);
        {
            std::lock_guard<std::mutex> lk(m_);
            auto it = need_.find(j->key);
            if (it != need_.end()) j->need = it->second;
        }
        if (!j->need) {   // host-side plan only: sizes for the admission decision
            auto s = make_solver(r);
            j->need = s->predicted_device_bytes(r.type);
        }
        FLUPS_REQUIRE(j->need <= budget_, "request needs %zu bytes of device memory, above the server budget %zu", j->need,
                      budget_);
        j->rhs = rhs;
        j->field = field;
        std::lock_guard<std::mutex> lk(m_);
        need_[j->key] = j->need;
        j->id = next_id_++;
        j->t_submit = Clock::now();
        if (!any_) {
            first_submit_ = j->t_submit;
            any_ = true;
        }
        queue_.push_back(j.get());
        const long id = j->id;
        jobs_[id] = std::move(j);
        cv_work_.notify_one();
        return id;
    }

    // 1 = done, 0 = failed (message in *err); removes the request
    int wait(long id, double* latency_ms, std::string* err) {
        std::unique_lock<std::mutex> lk(m_);
        auto it = jobs_.find(id);
        FLUPS_REQUIRE(it != jobs_.end(), "unknown request %ld", id);
        Job* j = it->second.get();
        cv_done_.wait(lk, [&] { return j->state >= 2; });
        const int ok = j->state == 2;
        if (latency_ms) *latency_ms = std::chrono::duration<double, std::milli>(j->t_done - j->t_submit).count();
        if (err) *err = j->error;
        jobs_.erase(it);
        return ok;
    }

    void drain() {
        std::unique_lock<std::mutex> lk(m_);
        cv_done_.wait(lk, [&] {
            for (auto& kv : jobs_)
                if (kv.second->state < 2) return false;
            return true;
        });
    }

    void stats(double out[12]) {
        std::lock_guard<std::mutex> lk(m_);
        const double win = any_ && done_ ? std::chrono::duration<double>(last_done_ - first_submit_).count() : 0.0;
        out[0] = (double)lat_.size();
        out[1] = (double)failed_;
        out[2] = pct(lat_, 50);
        out[3] = pct(lat_, 99);
        double sum = 0;
        for (double v : lat_) sum += v;
        out[4] = lat_.empty() ? 0.0 : sum / (double)lat_.size();
        out[5] = win > 0 ? (double)lat_.size() / win : 0.0;
        out[6] = (double)hits_;
        out[7] = (double)misses_;
        out[8] = (double)evictions_;
        out[9] = (double)peak_;
        out[10] = (double)used_;
        out[11] = (double)cache_.size();
    }

    void reset_stats() {
        std::lock_guard<std::mutex> lk(m_);
        lat_.clear();
        failed_ = hits_ = misses_ = evictions_ = 0;
        peak_ = used_;
        any_ = done_ = false;
    }

   private:
    // under the lock: the first queued request that can start now
    bool pick(Job*& job, Cached*& slot, bool& create, std::vector<std::unique_ptr<Cached>>& evict) {
        bool blocked = false;   // an earlier request waits for memory: later ones may not take new memory
        for (auto it = queue_.begin(); it != queue_.end(); ++it) {
            Job* j = *it;
            Cached* idle = nullptr;
            int count = 0;
            for (auto& e : cache_)
                if (e->key == j->key) {
                    ++count;
                    if (!e->busy && e->ready && !idle) idle = e.get();
                }
            if (idle) {
                idle->busy = true;
                job = j;
                slot = idle;
                create = false;
                ++hits_;
                queue_.erase(it);
                return true;
            }
            if (blocked || count >= per_key_) continue;
            // room for a new solver: evict idle solvers, least recently used first, if that is enough
            size_t idle_bytes = 0;
            for (auto& e : cache_)
                if (!e->busy) idle_bytes += e->bytes;
            if (used_ - idle_bytes + j->need > budget_) {
                blocked = true;
                continue;
            }
            while (used_ + j->need > budget_) {
                auto lru = cache_.end();
                for (auto e = cache_.begin(); e != cache_.end(); ++e)
                    if (!(*e)->busy && (lru == cache_.end() || (*e)->last_use < (*lru)->last_use)) lru = e;
                if (lru == cache_.end()) break;
                used_ -= (*lru)->bytes;
                evict.push_back(std::move(*lru));
                cache_.erase(lru);
                ++evictions_;
            }
            if (used_ + j->need > budget_) {
                blocked = true;
                continue;
            }
            auto c = std::make_unique<Cached>();
            c->key = j->key;
            c->bytes = j->need;
            used_ += j->need;
            peak_ = std::max(peak_, used_);
            slot = c.get();
            cache_.push_back(std::move(c));
            job = j;
            create = true;
            ++misses_;
            queue_.erase(it);
            return true;
        }
        return false;
    }

    void work() {
        (void)hipSetDevice(device_);
        std::unique_lock<std::mutex> lk(m_);
        while (true) {
            Job* j = nullptr;
            Cached* c = nullptr;
            bool create = false;
            std::vector<std::unique_ptr<Cached>> evict;
            cv_work_.wait(lk, [&] { return stop_ || pick(j, c, create, evict); });
            if (!j) return;   // stop
            j->state = 1;
            lk.unlock();
            evict.clear();   // frees device memory outside the lock
            std::string err;
            try {
                if (create) {
                    c->solver = make_solver(j->req);
                    HIP_CHECK(hipStreamCreate(&c->stream));
                    c->solver->set_stream(c->stream);
                    c->solver->setup();
                    c->solver->prepare(j->req.type);
                }
                c->solver->solve(j->field, j->rhs, j->req.type);
            } catch (const Error& e) {
                err = e.msg;
            } catch (const std::exception& e) {
                err = e.what();
            }
            lk.lock();
            j->t_done = Clock::now();
            if (create) {
                const size_t real = err.empty() ? c->solver->device_bytes() : 0;
                used_ = used_ - c->bytes + real;
                peak_ = std::max(peak_, used_);
                c->bytes = real;
                c->ready = err.empty();
            }
            c->busy = false;
            c->last_use = ++tick_;
            std::unique_ptr<Cached> dead;
            if (!c->ready) {   // failed set-up: drop the slot
                for (auto e = cache_.begin(); e != cache_.end(); ++e)
                    if (e->get() == c) {
                        used_ -= (*e)->bytes;
                        dead = std::move(*e);
                        cache_.erase(e);
                        break;
                    }
            }
            if (err.empty()) {
                j->state = 2;
                lat_.push_back(std::chrono::duration<double, std::milli>(j->t_done - j->t_submit).count());
                last_done_ = j->t_done;
                done_ = true;
            } else {
                j->state = 3;
                j->error = err;
                ++failed_;
            }
            cv_done_.notify_all();
            cv_work_.notify_all();
            if (dead) {
                lk.unlock();
                dead.reset();
                lk.lock();
            }
        }
    }

    size_t budget_;
    int per_key_;
    int device_ = 0;
    std::mutex m_;
    std::condition_variable cv_work_, cv_done_;
    std::deque<Job*> queue_;
    std::map<long, std::unique_ptr<Job>> jobs_;
    std::map<std::string, size_t> need_;
    std::vector<std::unique_ptr<Cached>> cache_;
    std::vector<std::thread> workers_;
    size_t used_ = 0, peak_ = 0;
    long next_id_ = 1;
    uint64_t tick_ = 0;
    bool stop_ = false, any_ = false, done_ = false;
    std::vector<double> lat_;
    long hits_ = 0, misses_ = 0, evictions_ = 0, failed_ = 0;
    Clock::time_point first_submit_, last_done_;
};

class Batch {
   public:
    Batch(int n, const FLUPS_ServeRequest* reqs, const double* const* rhs, double* const* field) {
        try {
            init(n, reqs, rhs, field);
        } catch (...) {   // a failed set-up releases what it made
            release();
            throw;
        }
    }
    ~Batch() { release(); }

   private:
    void init(int n, const FLUPS_ServeRequest* reqs, const double* const* rhs, double* const* field) {
        FLUPS_REQUIRE(n >= 1, "a batch needs at least one problem");
        HIP_CHECK(hipStreamCreate(&main_));
        for (int i = 0; i < n; ++i) {
            validate(reqs[i]);
            FLUPS_REQUIRE(rhs[i] && field[i], "batch buffers must not be NULL");
            FLUPS_REQUIRE(is_device(rhs[i]) && is_device(field[i]), "batch problem %d: buffers must be device memory", i);
            items_.emplace_back();   // owned by the batch from here on (released on failure)
            Item& it = items_.back();
            it.solver = make_solver(reqs[i]);
            HIP_CHECK(hipStreamCreate(&it.stream));
            it.solver->set_stream(it.stream);
            it.solver->setup();
            it.solver->prepare(reqs[i].type);
            HIP_CHECK(hipEventCreateWithFlags(&it.done, hipEventDisableTiming));
            it.rhs = rhs[i];
            it.field = field[i];
            it.type = reqs[i].type;
        }
        HIP_CHECK(hipEventCreateWithFlags(&fork_, hipEventDisableTiming));
        HIP_CHECK(hipEventCreate(&t0_));
        HIP_CHECK(hipEventCreate(&t1_));
        // one graph: fork from the main stream, one branch per problem, join. Branches are
        // recorded largest problem first (algorithmic bytes; FLUPS_BATCH_ORDER=0 keeps the
        // given order), so that the small ones fill in behind the large ones.
        std::vector<int> order(items_.size());
        for (size_t i = 0; i < order.size(); ++i) order[i] = (int)i;
        const char* bo = std::getenv("FLUPS_BATCH_ORDER");
        if (!bo || std::atoi(bo) != 0)
            std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
                return items_[a].solver->algorithmic_bytes(items_[a].type) > items_[b].solver->algorithmic_bytes(items_[b].type);
            });
        HIP_CHECK(hipStreamBeginCapture(main_, hipStreamCaptureModeThreadLocal));
        HIP_CHECK(hipEventRecord(fork_, main_));
        for (int oi : order) {
            auto& it = items_[oi];
            HIP_CHECK(hipStreamWaitEvent(it.stream, fork_, 0));
            it.solver->enqueue(it.field, it.rhs, it.type);
            HIP_CHECK(hipEventRecord(it.done, it.stream));
            HIP_CHECK(hipStreamWaitEvent(main_, it.done, 0));
        }
        HIP_CHECK(hipStreamEndCapture(main_, &graph_));
        HIP_CHECK(hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0));
        build_grouped();
    }
    void release() {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        if (main_ && hipStreamIsCapturing(main_, &cs) == hipSuccess && cs == hipStreamCaptureStatusActive) {
            hipGraph_t g = nullptr;   // a set-up that failed during the capture
            (void)hipStreamEndCapture(main_, &g);
            if (g) (void)hipGraphDestroy(g);
        }
        (void)hipGetLastError();
        if (gexec_) (void)hipGraphExecDestroy(gexec_);
        if (ggraph_) (void)hipGraphDestroy(ggraph_);
        if (dtab_) (void)hipFree(dtab_);
        if (exec_) (void)hipGraphExecDestroy(exec_);
        if (graph_) (void)hipGraphDestroy(graph_);
        for (auto& it : items_) {
            it.solver.reset();
            if (it.done) (void)hipEventDestroy(it.done);
            if (it.stream) (void)hipStreamDestroy(it.stream);
        }
        if (fork_) (void)hipEventDestroy(fork_);
        if (t0_) (void)hipEventDestroy(t0_);
        if (t1_) (void)hipEventDestroy(t1_);
        if (main_) (void)hipStreamDestroy(main_);
        gexec_ = exec_ = nullptr;
        ggraph_ = graph_ = nullptr;
        dtab_ = nullptr;
        fork_ = t0_ = t1_ = nullptr;
        main_ = nullptr;
        items_.clear();
    }

   public:
    // grouped launches (the default) or one graph branch per problem
    void run(bool grouped = true) {
        HIP_CHECK(hipEventRecord(t0_, main_));
        HIP_CHECK(hipGraphLaunch(grouped ? gexec_ : exec_, main_));
        HIP_CHECK(hipEventRecord(t1_, main_));
        HIP_CHECK(hipStreamSynchronize(main_));
        float ms = 0;
        HIP_CHECK(hipEventElapsedTime(&ms, t0_, t1_));
        last_ms_ = ms;
    }
    int launches(bool grouped) const { return grouped ? nlaunch_grouped_ : nlaunch_solo_; }
    // the same problems one after the other on the main stream, without the graph
    void run_sequential() {
        for (auto& it : items_) {
            it.solver->enqueue(it.field, it.rhs, it.type);
            HIP_CHECK(hipStreamSynchronize(it.stream));
        }
    }
    double last_ms() const { return last_ms_; }
    size_t device_bytes() const {
        size_t b = 0;
        for (auto& it : items_) b += it.solver->device_bytes();
        return b;
    }

   private:
    static bool is_device(const void* p) {
        hipPointerAttribute_t a;
        if (hipPointerGetAttributes(&a, p) != hipSuccess) {
            (void)hipGetLastError();
            return false;
        }
        return a.type == hipMemoryTypeDevice;
    }

    // record every problem's solve, group the launches by (step, kernel, block size, dynamic
    // LDS), and build the graph of grouped launches
    void build_grouped() {
        const int np = (int)items_.size();
        std::vector<std::vector<LaunchRec>> recs(np);
        for (int p = 0; p < np; ++p) {
            LaunchSink sink;
            launch_sink() = &sink;
            try {
                items_[p].solver->enqueue(items_[p].field, items_[p].rhs, items_[p].type);
            } catch (...) {
                launch_sink() = nullptr;
                throw;
            }
            launch_sink() = nullptr;
            FLUPS_REQUIRE(sink.same == 0, "unbalanced launch stages while recording batch problem %d", p);
            recs[p] = std::move(sink.recs);
            nlaunch_solo_ += (int)recs[p].size();
        }
        // groups: key -> (problem, record) pairs, larger grids first within a group
        // clusters: the problems sorted by algorithmic bytes and cut into `ncl` runs of about equal
        // size; launches are grouped within a cluster only, so that a grouped launch holds problems
        // of about the same work and one slow problem does not hold back the steps of small ones
        std::vector<int> order(np), cl(np, 0);
        for (int p = 0; p < np; ++p) order[p] = p;
        std::stable_sort(order.begin(), order.end(), [&](int a, int b) {
            return items_[a].solver->algorithmic_bytes(items_[a].type) > items_[b].solver->algorithmic_bytes(items_[b].type);
        });
        const char* re = std::getenv("FLUPS_BATCH_RATIO");
        const double ratio = std::max(1.0, re ? std::atof(re) : kRatio);
        const char* ce = std::getenv("FLUPS_BATCH_CLUSTERS");
        const int ncl = std::max(1, std::min(np, ce ? std::atoi(ce) : kClusters));
        std::vector<int> rank(np);
        for (int i = 0; i < np; ++i) {
            cl[order[i]] = (int)((long long)i * ncl / np);
            rank[order[i]] = i;
        }
        struct Key {
            int cluster;
            int stage;
            const void* fn;
            int threads;
            size_t smem, argsz;
            bool operator<(const Key& o) const {
                if (stage != o.stage) return stage < o.stage;
                if (cluster != o.cluster) return cluster < o.cluster;
                if (fn != o.fn) return fn < o.fn;
                if (threads != o.threads) return threads < o.threads;
                if (smem != o.smem) return smem < o.smem;
                return argsz < o.argsz;
            }
        };
        std::map<Key, std::vector<std::pair<int, int>>> groups;
        for (int p = 0; p < np; ++p)
            for (int r = 0; r < (int)recs[p].size(); ++r) {
                const LaunchRec& L = recs[p][r];
                FLUPS_REQUIRE(L.gx >= 1 && L.gy >= 1 && L.threads >= 1, "batch problem %d: empty launch recorded", p);
                groups[Key{cl[p], L.stage, L.grouped, L.threads, L.smem, L.args.size()}].push_back({p, r});
            }
        struct Chunk {
            std::vector<std::pair<int, int>> ent;
            size_t tab = 0;   // byte offset of its Args table
            int stride = 0;
        };
        std::vector<Chunk> chunks;
        size_t tab_bytes = 0;
        for (auto& kv : groups) {
            auto& v = kv.second;
            std::stable_sort(v.begin(), v.end(), [&](const std::pair<int, int>& a, const std::pair<int, int>& b) {
                const LaunchRec& x = recs[a.first][a.second];
                const LaunchRec& y = recs[b.first][b.second];
                return (long long)x.gx * x.gy > (long long)y.gx * y.gy;
            });
            // a chunk holds launches whose grids are within a factor `ratio` of its largest one
            auto blocks = [&](const std::pair<int, int>& e) {
                const LaunchRec& x = recs[e.first][e.second];
                return (long long)x.gx * x.gy;
            };
            for (size_t i = 0; i < v.size();) {
                size_t j = i + 1;
                while (j < v.size() && j - i < (size_t)kGroupMax && (double)blocks(v[j]) * ratio >= (double)blocks(v[i])) ++j;
                Chunk c;
                c.ent.assign(v.begin() + i, v.begin() + j);
                i = j;
                if (c.ent.size() > 1) {
                    c.stride = (int)((kv.first.argsz + 15) / 16 * 16);
                    c.tab = tab_bytes;
                    tab_bytes += (size_t)c.stride * c.ent.size();
                }
                chunks.push_back(std::move(c));
            }
        }
        // within a step, the launches of the larger problems first
        std::stable_sort(chunks.begin(), chunks.end(), [&](const Chunk& a, const Chunk& b) {
            const int sa = recs[a.ent[0].first][a.ent[0].second].stage, sb = recs[b.ent[0].first][b.ent[0].second].stage;
            if (sa != sb) return sa < sb;
            return rank[a.ent[0].first] < rank[b.ent[0].first];
        });
        std::vector<char> host(std::max<size_t>(tab_bytes, 16), 0);
        for (auto& c : chunks)
            for (size_t e = 0; c.ent.size() > 1 && e < c.ent.size(); ++e) {
                const LaunchRec& L = recs[c.ent[e].first][c.ent[e].second];
                std::memcpy(host.data() + c.tab + e * c.stride, L.args.data(), L.args.size());
            }
        HIP_CHECK(hipMalloc(&dtab_, host.size()));
        HIP_CHECK(hipMemcpy(dtab_, host.data(), host.size(), hipMemcpyHostToDevice));

        // graph: chunks in step order; each depends on the chunks that hold the previous
        // step of any of its problems
        HIP_CHECK(hipGraphCreate(&ggraph_, 0));
        std::map<std::pair<int, int>, std::vector<hipGraphNode_t>> at;   // (problem, step) -> nodes
        std::vector<std::vector<int>> steps(np);                          // steps of each problem, ascending
        for (int p = 0; p < np; ++p) {
            for (const LaunchRec& L : recs[p])
                if (steps[p].empty() || steps[p].back() != L.stage) steps[p].push_back(L.stage);
        }
        auto prev_step = [&](int p, int st) {
            int best = -1;
            for (int x : steps[p])
                if (x < st) best = x;
            return best;
        };
        for (auto& c : chunks) {
            const LaunchRec& L0 = recs[c.ent[0].first][c.ent[0].second];
            std::vector<hipGraphNode_t> deps;
            for (auto& e : c.ent) {
                const int ps = prev_step(e.first, L0.stage);
                if (ps < 0) continue;
                for (hipGraphNode_t n : at[{e.first, ps}])
                    if (std::find(deps.begin(), deps.end(), n) == deps.end()) deps.push_back(n);
            }
            hipKernelNodeParams kp{};
            GroupArgs g{};
            void* args[1];
            kp.blockDim = dim3(L0.threads);
            kp.sharedMemBytes = (unsigned)L0.smem;
            if (c.ent.size() == 1) {
                kp.func = const_cast<void*>(L0.solo);
                kp.gridDim = dim3(L0.gx, L0.gy);
                args[0] = const_cast<char*>(L0.args.data());
            } else {
                g.tab = static_cast<const char*>(dtab_) + c.tab;
                g.stride = c.stride;
                g.n = (int)c.ent.size();
                long long tot = 0;
                for (int e = 0; e < g.n; ++e) {
                    const LaunchRec& L = recs[c.ent[e].first][c.ent[e].second];
                    g.start[e] = (int)tot;
                    g.gx[e] = L.gx;
                    tot += (long long)L.gx * L.gy;
                }
                FLUPS_REQUIRE(tot < (1LL << 31), "grouped launch of %lld blocks", tot);
                g.start[g.n] = (int)tot;
                kp.func = const_cast<void*>(L0.grouped);
                kp.gridDim = dim3((unsigned)tot);
                args[0] = &g;
            }
            kp.kernelParams = args;
            kp.extra = nullptr;
            hipGraphNode_t node;
            HIP_CHECK(hipGraphAddKernelNode(&node, ggraph_, deps.data(), deps.size(), &kp));
            ++nlaunch_grouped_;
            for (auto& e : c.ent) at[{e.first, L0.stage}].push_back(node);
        }
        HIP_CHECK(hipGraphInstantiate(&gexec_, ggraph_, nullptr, nullptr, 0));
    }
   private:
    struct Item {
        std::unique_ptr<SolverImpl> solver;
        hipStream_t stream = nullptr;
        hipEvent_t done = nullptr;
        const double* rhs = nullptr;
        double* field = nullptr;
        int type = 0;
    };
    std::vector<Item> items_;
    hipStream_t main_ = nullptr;
    hipEvent_t fork_ = nullptr, t0_ = nullptr, t1_ = nullptr;
    hipGraph_t graph_ = nullptr, ggraph_ = nullptr;
    hipGraphExec_t exec_ = nullptr, gexec_ = nullptr;
    // grouping policy, measured on the 64-problem mixes of the batch benchmark: launches are grouped
    // across all problems (one cluster) but only with launches of the same grid (ratio 1); wider
    // groups hold back their small problems behind the large ones (3.0 ms for the random mix with
    // equal grids only, 3.6 ms when any grids are grouped)
    static constexpr int kClusters = 1;
    static constexpr double kRatio = 1.0;
    void* dtab_ = nullptr;   // Args tables of the grouped launches
    int nlaunch_solo_ = 0, nlaunch_grouped_ = 0;
    double last_ms_ = 0;
};

}  // namespace flups

// ---- C entry points ----
struct FLUPS_ServerS {
    flups::Server s;
    FLUPS_ServerS(size_t b, int w, int k) : s(b, w, k) {}
};
struct FLUPS_BatchS {
    flups::Batch b;
    FLUPS_BatchS(int n, const FLUPS_ServeRequest* r, const double* const* x, double* const* y) : b(n, r, x, y) {}
};

namespace flups {
void report_error(const char* where, const std::string& msg);   // abi.cpp
}

template <typename F>
static auto sguard(const char* where, F&& f) -> decltype(f()) {
    try {
        return f();
    } catch (const flups::Error& e) {
        flups::report_error(where, e.msg);
    } catch (const std::exception& e) {
        flups::report_error(where, e.what());
    }
    return decltype(f())();
}

extern "C" {

FLUPS_Server* flups_server_new(size_t hbm_budget_bytes, int workers, int solvers_per_key) {
    return sguard("flups_server_new", [&]() -> FLUPS_Server* {
        return reinterpret_cast<FLUPS_Server*>(new FLUPS_ServerS(hbm_budget_bytes, workers, solvers_per_key));
    });
}
void flups_server_free(FLUPS_Server* s) { delete reinterpret_cast<FLUPS_ServerS*>(s); }
long flups_server_submit(FLUPS_Server* s, const FLUPS_ServeRequest* req, const double* rhs, double* field) {
    return sguard("flups_server_submit", [&]() -> long {
        FLUPS_REQUIRE(s && req, "NULL server or request");
        return reinterpret_cast<FLUPS_ServerS*>(s)->s.submit(*req, rhs, field);
    });
}
int flups_server_wait(FLUPS_Server* s, long ticket, double* latency_ms) {
    return sguard("flups_server_wait", [&]() -> int {
        FLUPS_REQUIRE(s, "NULL server");
        std::string err;
        const int ok = reinterpret_cast<FLUPS_ServerS*>(s)->s.wait(ticket, latency_ms, &err);
        FLUPS_REQUIRE(ok, "request %ld failed: %s", ticket, err.c_str());
        return 1;
    });
}
void flups_server_drain(FLUPS_Server* s) {
    sguard("flups_server_drain", [&]() {
        FLUPS_REQUIRE(s, "NULL server");
        reinterpret_cast<FLUPS_ServerS*>(s)->s.drain();
        return 0;
    });
}
void flups_server_stats(FLUPS_Server* s, double out[12]) {
    sguard("flups_server_stats", [&]() {
        FLUPS_REQUIRE(s && out, "NULL server or output");
        reinterpret_cast<FLUPS_ServerS*>(s)->s.stats(out);
        return 0;
    });
}
size_t flups_server_request_bytes(const FLUPS_ServeRequest* req) {
    return sguard("flups_server_request_bytes", [&]() -> size_t {
        FLUPS_REQUIRE(req, "NULL request");
        flups::validate(*req);
        return flups::make_solver(*req)->predicted_device_bytes(req->type);
    });
}
void flups_server_reset_stats(FLUPS_Server* s) {
    if (s) reinterpret_cast<FLUPS_ServerS*>(s)->s.reset_stats();
}

FLUPS_Batch* flups_batch_new(int n, const FLUPS_ServeRequest* reqs, const double* const* rhs, double* const* field) {
    return sguard("flups_batch_new", [&]() -> FLUPS_Batch* {
        FLUPS_REQUIRE(reqs && rhs && field, "NULL batch description");
        return reinterpret_cast<FLUPS_Batch*>(new FLUPS_BatchS(n, reqs, rhs, field));
    });
}
void flups_batch_free(FLUPS_Batch* b) { delete reinterpret_cast<FLUPS_BatchS*>(b); }
int flups_batch_run(FLUPS_Batch* b) {
    return sguard("flups_batch_run", [&]() -> int {
        FLUPS_REQUIRE(b, "NULL batch");
        reinterpret_cast<FLUPS_BatchS*>(b)->b.run();
        return 1;
    });
}
int flups_batch_run_branches(FLUPS_Batch* b) {
    return sguard("flups_batch_run_branches", [&]() -> int {
        FLUPS_REQUIRE(b, "NULL batch");
        reinterpret_cast<FLUPS_BatchS*>(b)->b.run(false);
        return 1;
    });
}
int flups_batch_launches(FLUPS_Batch* b, int grouped) {
    return b ? reinterpret_cast<FLUPS_BatchS*>(b)->b.launches(grouped != 0) : 0;
}
int flups_batch_run_sequential(FLUPS_Batch* b) {
    return sguard("flups_batch_run_sequential", [&]() -> int {
        FLUPS_REQUIRE(b, "NULL batch");
        reinterpret_cast<FLUPS_BatchS*>(b)->b.run_sequential();
        return 1;
    });
}
double flups_batch_last_ms(FLUPS_Batch* b) { return b ? reinterpret_cast<FLUPS_BatchS*>(b)->b.last_ms() : 0.0; }
size_t flups_batch_device_bytes(FLUPS_Batch* b) { return b ? reinterpret_cast<FLUPS_BatchS*>(b)->b.device_bytes() : 0; }

}  // extern "C"
