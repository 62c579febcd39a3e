// server.hip — continuous-batching request loop for beamlet solves.
//
// The reference has no batching or serving mode (one beam per process).  This
// runtime serves many independent beam solves on one GPU:
//  * requests are queued by bfpn_server_submit from any thread;
//  * a preparer thread admits the queued request with the least total work
//    (steps x cells) as soon as a slot is free and its device memory fits the
//    HBM budget (state + source scratch of
//    the beam, plus its physics tables unless a cached copy matches), builds
//    the tables and the beam's context -- host work of up to seconds that
//    overlaps the GPU work of the running beams;
//  * physics tables (material tables, eq15 coefficients, split-BF16 source
//    operands, CN factors) are shared between beams with equal grid, step,
//    material and flags, and kept in an LRU cache inside the budget;
//  * every admitted beam keeps its whole phase space resident and runs on its
//    own HIP stream; the scheduler thread enqueues slices of depth steps,
//    shortest remaining work (steps x cells) first, to as many beams as it
//    takes to keep `fill_cells` cells in flight, and polls completion events,
//    so small requests finish early, work of different beams overlaps on the
//    device and newly arrived requests join while others are running;
//  * a finished beam's IDD history is copied out, its memory freed, and the
//    slot handed to the next request.
// Results do not depend on what else is running: beams share only read-only
// tables, so a solve in the mix equals the same solve alone, bit for bit.

#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <list>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "solver.hpp"

namespace bfpn {
namespace {

using Clock = std::chrono::steady_clock;

double seconds_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double>(b - a).count();
}

struct Request {
    uint64_t id = 0;
    bfpn_grid grid{};
    bfpn_phys phys{};
    std::string data;
    int steps = 0;
    size_t ctx_bytes = 0;
    Clock::time_point t_submit, t_admit, t_done;
    std::vector<LayerSpec> layers;  // empty: one material (data); with a map: its materials
    std::vector<double> slab_end;   // material map (empty: none), as in bfpn_create_mapped
    std::vector<unsigned char> map;
    std::unique_ptr<bfpn_ctx> ctx;
    std::vector<std::shared_ptr<PhysicsBase>> physics;  // one table set per layer
    int issued = 0;
    hipEvent_t slice_done = nullptr;
    bool in_flight = false;
    int status = 0;  // 0 queued, 1 running, 2 done, 3 failed
    std::string error;
    std::vector<double> idd;
};

struct CacheEntry {
    std::shared_ptr<PhysicsBase> physics;
    size_t bytes = 0;
};

}  // namespace
}  // namespace bfpn

using namespace bfpn;

struct bfpn_server {
    bfpn_server_opts opts{};
    std::mutex mu;
    std::condition_variable cv_done;  // a request finished
    std::condition_variable cv_work;  // work for the scheduler
    std::condition_variable cv_prep;  // work for the preparer
    std::deque<std::shared_ptr<Request>> pending;  // submitted
    std::deque<std::shared_ptr<Request>> ready;    // prepared, not yet scheduled
    std::vector<std::shared_ptr<Request>> active;  // scheduled (only the scheduler mutates it)
    int preparing = 0;
    std::map<uint64_t, std::shared_ptr<Request>> finished;
    std::map<std::string, CacheEntry> cache;
    std::list<std::string> lru;  // front = most recent
    size_t used = 0;             // bytes of admitted contexts + cached physics
    uint64_t next_id = 1;
    bool stopping = false;
    std::thread scheduler, preparer;
    bfpn_server_stats stats{};

    int admitted() const { return preparing + static_cast<int>(ready.size() + active.size()); }

    void touch(const std::string& key) {
        lru.remove(key);
        lru.push_front(key);
    }

    // Frees cached physics tables that no admitted beam uses, oldest first,
    // until `need` more bytes fit; the entries in `keep` (the tables of the
    // request being admitted) are never freed.  Called with the lock held.
    void evict_for(size_t need, const std::vector<std::string>& keep) {
        for (auto it = lru.rbegin(); it != lru.rend() && used + need > opts.hbm_budget_bytes;) {
            auto c = cache.find(*it);
            const bool kept = std::find(keep.begin(), keep.end(), *it) != keep.end();
            if (c != cache.end() && !kept && c->second.physics.use_count() == 1) {
                used -= c->second.bytes;
                cache.erase(c);
                it = std::list<std::string>::reverse_iterator(lru.erase(std::next(it).base()));
            } else {
                ++it;
            }
        }
    }

    void fail_locked(const std::shared_ptr<Request>& r, const std::string& msg) {
        r->status = 3;
        r->error = msg;
        r->t_done = Clock::now();
        finished[r->id] = r;
        ++stats.failed;
        cv_done.notify_all();
    }

    static double total_cells(const Request& r) {
        const double nyz = static_cast<double>(r.grid.Ny + 1) * (r.grid.Nz + 1);
        return static_cast<double>(r.steps) * r.grid.Ng * nyz * r.grid.Nmu * r.grid.Nom;
    }

    // Preparer thread: admission, physics tables and contexts; the queued
    // request with the least total work first (like the scheduler's order).
    void prepare_loop() {
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            if (pending.empty() || admitted() >= opts.max_active) {
                if (stopping && pending.empty()) break;
                cv_prep.wait(lk);
                continue;
            }
            auto pick = std::min_element(pending.begin(), pending.end(),
                                         [](const std::shared_ptr<Request>& a,
                                            const std::shared_ptr<Request>& b) {
                                             return total_cells(*a) < total_cells(*b);
                                         });
            auto req = *pick;
            // the request's table sets: one per layer (one for a single material)
            struct Tables {
                std::string key, data;
                bfpn_phys phys;
                std::shared_ptr<PhysicsBase> physics;  // cached, or built below
                size_t estimate = 0;                   // bytes charged when not cached
            };
            std::vector<Tables> tables;
            if (req->layers.empty()) {
                tables.push_back({physics_key(req->grid, req->phys, req->data, opts.precision),
                                  req->data, req->phys, nullptr, 0});
            } else {
                for (const LayerSpec& l : req->layers) {
                    const bfpn_phys q = layer_phys(req->phys, l);
                    tables.push_back({physics_key(req->grid, q, l.data_path, opts.precision),
                                      l.data_path, q, nullptr, 0});
                }
            }
            std::vector<std::string> keys;
            size_t phys_need = 0;
            for (Tables& t : tables) {
                keys.push_back(t.key);
                const bool repeat = std::count(keys.begin(), keys.end(), t.key) > 1;
                const auto hit = cache.find(t.key);
                if (hit == cache.end() && !repeat) {
                    t.estimate = physics_bytes(req->grid, t.phys, opts.precision);
                    phys_need += t.estimate;
                }
            }
            const size_t need = req->ctx_bytes + phys_need;
            if (used + need > opts.hbm_budget_bytes) evict_for(need, keys);
            if (used + need > opts.hbm_budget_bytes) {
                if (admitted() > 0) {  // wait for a running beam to finish
                    cv_prep.wait(lk);
                    continue;
                }
                pending.erase(pick);
                fail_locked(req, "request does not fit the HBM budget (" +
                                     std::to_string(need >> 20) + " MiB > " +
                                     std::to_string((opts.hbm_budget_bytes - used) >> 20) +
                                     " MiB free)");
                continue;
            }
            pending.erase(pick);
            used += need;
            ++preparing;
            for (Tables& t : tables) {  // evict_for kept these entries
                const auto hit = cache.find(t.key);
                if (hit != cache.end()) t.physics = hit->second.physics;
            }
            lk.unlock();
            std::string err;
            try {
                for (size_t k = 0; k < tables.size(); ++k) {
                    Tables& t = tables[k];
                    for (size_t j = 0; j < k && !t.physics; ++j)  // a material repeated in depth
                        if (tables[j].key == t.key) t.physics = tables[j].physics;
                    if (!t.physics) t.physics = make_physics(req->grid, t.phys, t.data, opts.precision);
                    req->physics.push_back(t.physics);
                }
                if (req->layers.empty())
                    req->ctx.reset(make_context(req->physics.front(), req->grid, req->phys));
                else if (!req->map.empty())
                    req->ctx.reset(make_mapped_context(req->grid, req->phys, req->layers,
                                                       req->slab_end, req->map, req->physics));
                else
                    req->ctx.reset(make_layered_context(req->grid, req->phys, req->layers,
                                                        req->physics));
                BFPN_HIP(hipEventCreateWithFlags(&req->slice_done, hipEventDisableTiming));
            } catch (const std::exception& e) {
                err = e.what();
            }
            lk.lock();
            --preparing;
            if (!err.empty()) {
                used -= need;
                req->ctx.reset();
                req->physics.clear();
                fail_locked(req, err);
                cv_work.notify_one();
                continue;
            }
            for (Tables& t : tables) {
                if (t.estimate) {  // newly built: keep the larger of estimate and real size
                    const size_t charged = std::max(t.estimate, t.physics->device_bytes());
                    used = used - t.estimate + charged;
                    cache[t.key] = CacheEntry{t.physics, charged};
                }
                touch(t.key);
            }
            req->status = 1;
            req->t_admit = Clock::now();
            ready.push_back(req);
            stats.peak_active = std::max(stats.peak_active, static_cast<int>(ready.size() + active.size()));
            cv_work.notify_one();
        }
    }

    // Enqueues the next slice of a running request: slice_steps depth steps
    // of a beam of up to 5e7 cells, proportionally fewer (at least one) for
    // larger beams, so a slice costs about the same whatever the beam and a
    // newly arrived small request never waits long behind a large one.
    void issue(Request& r) {
        const double scale = 5e7 / static_cast<double>(std::max<size_t>(r.ctx->cells, 1));
        const int per = std::max(1, std::min(opts.slice_steps,
                                             static_cast<int>(opts.slice_steps * scale)));
        const int n = std::min(per, r.steps - r.issued);
        r.ctx->advance(n);
        BFPN_HIP(hipEventRecord(r.slice_done, r.ctx->done_stream));
        r.issued += n;
        r.in_flight = true;
    }

    void finish(const std::shared_ptr<Request>& r) {
        std::vector<double> idd(r->steps);
        int n = 0;
        r->ctx->history(idd.data(), nullptr, r->steps, &n);
        idd.resize(n);
        const size_t bytes = r->ctx_bytes;
        (void)hipEventDestroy(r->slice_done);
        r->slice_done = nullptr;
        r->ctx.reset();
        r->physics.clear();
        std::lock_guard<std::mutex> g(mu);
        r->idd.swap(idd);
        r->status = 2;
        r->t_done = Clock::now();
        used -= bytes;
        ++stats.completed;
        finished[r->id] = r;
        cv_done.notify_all();
        cv_prep.notify_one();
    }

    void fail(const std::shared_ptr<Request>& r, const std::string& msg) {
        if (r->slice_done) (void)hipEventDestroy(r->slice_done);
        r->slice_done = nullptr;
        r->ctx.reset();
        r->physics.clear();
        std::lock_guard<std::mutex> g(mu);
        used -= r->ctx_bytes;
        fail_locked(r, msg);
        cv_prep.notify_one();
    }

    static double remaining_cells(const Request& r) {
        return static_cast<double>(r.steps - r.issued) * static_cast<double>(r.ctx->cells);
    }

    // Scheduler thread: slices, shortest remaining work first.
    void schedule_loop() {
        std::unique_lock<std::mutex> lk(mu);
        for (;;) {
            while (!ready.empty()) {
                active.push_back(ready.front());
                ready.pop_front();
            }
            if (active.empty()) {
                if (stopping && pending.empty() && preparing == 0) break;
                cv_work.wait(lk);
                continue;
            }
            lk.unlock();
            bool progressed = false;
            std::vector<std::shared_ptr<Request>> done;
            double in_flight = 0.0;
            std::vector<Request*> idle;
            for (auto& r : active) {
                try {
                    if (r->in_flight) {
                        const hipError_t q = hipEventQuery(r->slice_done);
                        if (q == hipErrorNotReady) {
                            in_flight += static_cast<double>(r->ctx->cells);
                            continue;
                        }
                        BFPN_HIP(q);
                        r->in_flight = false;
                        progressed = true;
                    }
                    if (r->issued >= r->steps) {
                        done.push_back(r);
                        continue;
                    }
                    idle.push_back(r.get());
                } catch (const std::exception& e) {
                    (void)hipGetLastError();
                    fail(r, e.what());
                    progressed = true;
                }
            }
            std::sort(idle.begin(), idle.end(), [](const Request* a, const Request* b) {
                return remaining_cells(*a) < remaining_cells(*b);
            });
            for (Request* r : idle) {
                if (in_flight > 0.0 && in_flight >= static_cast<double>(opts.fill_cells)) break;
                try {
                    issue(*r);
                    in_flight += static_cast<double>(r->ctx->cells);
                    progressed = true;
                } catch (const std::exception& e) {
                    (void)hipGetLastError();
                    for (auto& sp : active)
                        if (sp.get() == r) fail(sp, e.what());
                    progressed = true;
                }
            }
            for (auto& r : done) {
                try {
                    finish(r);
                } catch (const std::exception& e) {
                    fail(r, e.what());
                }
                progressed = true;
            }
            lk.lock();
            active.erase(std::remove_if(active.begin(), active.end(),
                                        [](const std::shared_ptr<Request>& r) { return r->status >= 2; }),
                         active.end());
            if (!progressed) {
                lk.unlock();
                std::this_thread::sleep_for(std::chrono::microseconds(50));
                lk.lock();
            }
        }
    }
};

extern "C" {

BFPN_API bfpn_status bfpn_server_create(const bfpn_server_opts* opts, bfpn_server** out) {
    return guarded([&] {
        BFPN_REQUIRE(opts && out, "null argument");
        BFPN_REQUIRE(opts->max_active >= 1 && opts->slice_steps >= 1, "max_active and slice_steps must be >= 1");
        BFPN_REQUIRE(opts->precision == BFPN_FP32 || opts->precision == BFPN_FP64, "bad precision");
        auto* s = new bfpn_server();
        s->opts = *opts;
        if (s->opts.hbm_budget_bytes == 0) {
            size_t free_b = 0, total_b = 0;
            BFPN_HIP(hipMemGetInfo(&free_b, &total_b));
            s->opts.hbm_budget_bytes = free_b - free_b / 10;
        }
        if (s->opts.fill_cells == 0) s->opts.fill_cells = 40000000;
        s->scheduler = std::thread([s] { s->schedule_loop(); });
        s->preparer = std::thread([s] { s->prepare_loop(); });
        *out = s;
    });
}

namespace {
// Queues a request (layers empty: one material, tables from data_path; with
// a map, layers holds its materials).
void submit_request(bfpn_server* s, const bfpn_grid* grid, const bfpn_phys* phys,
                    const char* data_path, std::vector<LayerSpec> layers, double t_final,
                    uint64_t* id, std::vector<double> slab_end = {},
                    std::vector<unsigned char> map = {}) {
    BFPN_REQUIRE(s && grid && phys && id, "null argument");
    validate_params(*grid, *phys);
    const long long nyz = static_cast<long long>(grid->Ny + 1) * (grid->Nz + 1);
    if (!map.empty()) {
        validate_map(layers, slab_end, map, nyz);
        BFPN_REQUIRE(!phys->legacy_energy && !phys->lite_memory,
                     "material maps support the eq15 energy model with the full state only");
    } else if (!layers.empty()) {
        validate_layers(layers);
    }
    auto r = std::make_shared<Request>();
    r->grid = *grid;
    r->phys = *phys;
    r->data = data_path ? data_path : "./data";
    r->layers = std::move(layers);
    r->slab_end = std::move(slab_end);
    r->map = std::move(map);
    const double dt = grid->Lx / grid->Nx;
    r->steps = static_cast<int>(std::ceil(t_final / dt - 1e-12));
    BFPN_REQUIRE(r->steps >= 1, "t_final must cover at least one depth step");
    r->ctx_bytes = context_bytes(*grid, *phys, s->opts.precision);
    if (!r->map.empty())  // packed copies of the smaller materials' nodes
        r->ctx_bytes += map_bytes(*grid, *phys, r->map, static_cast<int>(r->layers.size()),
                                  s->opts.precision);
    r->t_submit = Clock::now();
    std::lock_guard<std::mutex> g(s->mu);
    BFPN_REQUIRE(!s->stopping, "server is shutting down");
    r->id = s->next_id++;
    s->pending.push_back(r);
    ++s->stats.submitted;
    *id = r->id;
    s->cv_prep.notify_one();
}
}  // namespace

BFPN_API bfpn_status bfpn_server_submit_layered(bfpn_server* s, const bfpn_grid* grid,
                                                const bfpn_phys* phys, const bfpn_layer* layers,
                                                int nlayers, double t_final, uint64_t* id) {
    return guarded([&] {
        BFPN_REQUIRE(layers && nlayers > 0, "null argument");
        std::vector<LayerSpec> spec;
        for (int k = 0; k < nlayers; ++k) {
            BFPN_REQUIRE(layers[k].material && layers[k].data_path, "null layer field");
            spec.push_back({layers[k].x_end, layers[k].material, layers[k].density,
                            layers[k].data_path});
        }
        submit_request(s, grid, phys, nullptr, std::move(spec), t_final, id);
    });
}

BFPN_API bfpn_status bfpn_server_submit_mapped(bfpn_server* s, const bfpn_grid* grid,
                                               const bfpn_phys* phys, const bfpn_layer* materials,
                                               int nmat, const double* slab_end, int nslabs,
                                               const unsigned char* map, double t_final,
                                               uint64_t* id) {
    return guarded([&] {
        BFPN_REQUIRE(grid && materials && slab_end && map && nmat > 0 && nslabs > 0, "null argument");
        std::vector<LayerSpec> mats;
        for (int k = 0; k < nmat; ++k) {
            BFPN_REQUIRE(materials[k].material && materials[k].data_path, "null material field");
            mats.push_back({0.0, materials[k].material, materials[k].density, materials[k].data_path});
        }
        const long long nyz = static_cast<long long>(grid->Ny + 1) * (grid->Nz + 1);
        submit_request(s, grid, phys, nullptr, std::move(mats), t_final, id,
                       std::vector<double>(slab_end, slab_end + nslabs),
                       std::vector<unsigned char>(map, map + nslabs * nyz));
    });
}

BFPN_API bfpn_status bfpn_server_submit(bfpn_server* s, const bfpn_grid* grid, const bfpn_phys* phys,
                                        const char* data_path, double t_final, uint64_t* id) {
    return guarded([&] { submit_request(s, grid, phys, data_path, {}, t_final, id); });
}

BFPN_API int bfpn_server_poll(bfpn_server* s, uint64_t id) {
    if (!s) return -1;
    std::lock_guard<std::mutex> g(s->mu);
    auto it = s->finished.find(id);
    if (it != s->finished.end()) return it->second->status;
    for (auto& r : s->active)
        if (r->id == id) return 1;
    for (auto& r : s->ready)
        if (r->id == id) return 1;
    for (auto& r : s->pending)
        if (r->id == id) return 0;
    return 0;  // being prepared
}

BFPN_API bfpn_status bfpn_server_wait(bfpn_server* s, uint64_t id, double timeout_s, double* idd,
                                      int capacity, int* nsteps, double* times3) {
    return guarded([&] {
        BFPN_REQUIRE(s, "null server");
        std::unique_lock<std::mutex> lk(s->mu);
        const auto deadline = Clock::now() + std::chrono::duration_cast<Clock::duration>(
                                                 std::chrono::duration<double>(timeout_s > 0 ? timeout_s : 1e9));
        while (!s->finished.count(id)) {
            if (s->cv_done.wait_until(lk, deadline) == std::cv_status::timeout && !s->finished.count(id))
                throw Error(BFPN_E_INVALID, "timed out waiting for request " + std::to_string(id));
        }
        auto r = s->finished[id];
        s->finished.erase(id);
        if (times3) {
            times3[0] = r->status == 2 ? seconds_between(r->t_submit, r->t_admit) : 0.0;
            times3[1] = r->status == 2 ? seconds_between(r->t_admit, r->t_done) : 0.0;
            times3[2] = seconds_between(r->t_submit, r->t_done);
        }
        if (r->status != 2) throw Error(BFPN_E_INVALID, "request " + std::to_string(id) + " failed: " + r->error);
        const int n = std::min<int>(capacity, static_cast<int>(r->idd.size()));
        if (idd && n > 0) std::copy(r->idd.begin(), r->idd.begin() + n, idd);
        if (nsteps) *nsteps = static_cast<int>(r->idd.size());
    });
}

BFPN_API bfpn_status bfpn_server_get_stats(bfpn_server* s, bfpn_server_stats* out) {
    return guarded([&] {
        BFPN_REQUIRE(s && out, "null argument");
        std::lock_guard<std::mutex> g(s->mu);
        *out = s->stats;
        out->active = static_cast<int>(s->active.size() + s->ready.size()) + s->preparing;
        out->queued = static_cast<int>(s->pending.size());
        out->bytes_in_use = s->used;
        out->cached_tables = static_cast<int>(s->cache.size());
    });
}

BFPN_API bfpn_status bfpn_server_estimate(const bfpn_grid* grid, const bfpn_phys* phys,
                                          int precision, size_t* ctx_bytes, size_t* phys_bytes) {
    return guarded([&] {
        BFPN_REQUIRE(grid && phys && ctx_bytes && phys_bytes, "null argument");
        BFPN_REQUIRE(precision == BFPN_FP32 || precision == BFPN_FP64, "unknown precision");
        validate_params(*grid, *phys);
        *ctx_bytes = context_bytes(*grid, *phys, precision);
        *phys_bytes = physics_bytes(*grid, *phys, precision);
    });
}

BFPN_API void bfpn_server_destroy(bfpn_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stopping = true;
        s->cv_work.notify_one();
        s->cv_prep.notify_one();
    }
    if (s->preparer.joinable()) s->preparer.join();
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->cv_work.notify_one();
    }
    if (s->scheduler.joinable()) s->scheduler.join();
    delete s;
}

}  // extern "C"
