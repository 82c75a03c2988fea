// server.hip -- serving loop for fit requests (continuous batching on one GPU).
//
// Requests are queued with an arrival time on the server clock.  A scheduling round
// (vcl_server_progress) retires completed step batches, finishes requests that ran all their
// iterations, admits arrived requests while a slot is free and the request's arena fits the
// HBM budget, and queues the next steps of active requests, each on its own stream, so the
// steps of different requests overlap on the device.  In grouped mode the requests due to step
// in a round are split by step signature and each part steps as one vcl_group (shared launches
// over a device-side work table) on a stream of its own; a group is kept while its requests
// keep stepping together and rebuilt when they change.  Policy FIFO admits in arrival order and
// steps every active request; SRPT admits the arrived request with the least work first and
// steps only the few active requests with the least remaining work.  Every
// request's state (mesh, mip texture, lobes, Adam moments, intermediates) stays resident in one
// arena until the request is released; arenas are recycled through a pool so that admission
// and release do not call hipMalloc/hipFree (hipFree synchronises the whole device) in the
// steady state.
#include "internal.hpp"
#include <algorithm>
#include <cmath>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <thread>
#include <vector>

namespace {

using clk = std::chrono::steady_clock;

// Device arenas, best-fit reuse of released blocks (no splitting); blocks are only freed to
// make room when a new block would pass the budget.
struct ArenaPool {
    size_t budget = 0, in_use = 0, cached = 0, peak = 0;
    std::multimap<size_t, void*> free_blocks;   // size -> ptr
    std::map<void*, size_t> size_of;

    // 0: ok; 1: does not fit the budget now; <0: HIP error
    int get(size_t bytes, void** out) {
        auto it = free_blocks.lower_bound(bytes);
        if (it != free_blocks.end() && it->first <= bytes + bytes / 4) {
            *out = it->second;
            cached -= it->first;
            in_use += it->first;
            free_blocks.erase(it);
            peak = std::max(peak, in_use);
            return 0;
        }
        if (in_use + bytes > budget) return 1;
        while (in_use + cached + bytes > budget && !free_blocks.empty()) {
            auto big = std::prev(free_blocks.end());
            if (hipFree(big->second) != hipSuccess) return -1;
            cached -= big->first;
            size_of.erase(big->second);
            free_blocks.erase(big);
        }
        void* p = nullptr;
        if (hipMalloc(&p, bytes) != hipSuccess) return 1;
        size_of[p] = bytes;
        in_use += bytes;
        peak = std::max(peak, in_use);
        *out = p;
        return 0;
    }
    void put(void* p) {
        size_t b = size_of.at(p);
        in_use -= b;
        cached += b;
        free_blocks.insert({b, p});
    }
    void release_all() {
        for (auto& kv : size_of) (void)hipFree(kv.first);
        size_of.clear();
        free_blocks.clear();
        in_use = cached = 0;
    }
};

struct Req {
    uint64_t ticket;
    vcl_fit_desc d;
    int iterations;
    double arrival, admit = -1, finish = -1;
    size_t bytes = 0;
    void* arena = nullptr;
    vcl_fit* fit = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev = nullptr;
    int launched = 0, done = 0;
    uint64_t sig = 0;    // step signature (grouped mode)
    double cost = 0;     // work per iteration (pixels + 8 per triangle, over all views)
    float loss = 0.f;
    int status = VCL_OK;
    bool reported = false;
};

// A built group: its requests (by ticket, ascending), work tables from the pool, stream and the
// event after its last queued steps.
struct GroupSlot {
    std::vector<uint64_t> key;
    vcl_group* g = nullptr;
    void* table = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t ev = nullptr;
    int uses = 0;
    bool mark = false;
};

}  // namespace

struct vcl_server {
    vcl_server_config cfg;
    clk::time_point t0;
    ArenaPool pool;
    uint64_t next_ticket = 1;
    std::deque<Req*> queue;                 // submitted, not admitted (FIFO by submission)
    std::vector<Req*> active;               // admitted, stepping
    std::vector<Req*> finished;             // all iterations done (arena held until release)
    std::vector<hipStream_t> streams;       // idle streams for reuse
    std::map<uint64_t, vcl_tuning> tunings; // autotuned schedule per size class
    std::vector<GroupSlot*> groups;         // grouped mode: built groups
    double occ_integral = 0, occ_last_t = 0;

    double now() const { return std::chrono::duration<double>(clk::now() - t0).count(); }
};

static void occupancy_sample(vcl_server* S) {
    double t = S->now();
    S->occ_integral += (double)S->pool.in_use * (t - S->occ_last_t);
    S->occ_last_t = t;
}

static void release_req(vcl_server* S, Req* r) {
    if (r->fit) { vcl_fit_destroy(r->fit); r->fit = nullptr; }
    if (r->ev) { (void)hipEventDestroy(r->ev); r->ev = nullptr; }
    if (r->stream) { S->streams.push_back(r->stream); r->stream = nullptr; }
    if (r->arena) { occupancy_sample(S); S->pool.put(r->arena); r->arena = nullptr; }
}

// Size class of a request for the tuning table: pixels over all views in buckets of 4x
// (256^2, 512^2, 1024^2, 2048^2 ...) and triangles in buckets of 8x; the tuned kernels
// (raster tiling, texel splats) depend on these, not on the texture size or lobe count.
static uint64_t size_class(const vcl_fit_desc* d) {
    auto bucket = [](double x, double base) { return (uint64_t)std::lround(std::log(x < 1.0 ? 1.0 : x) / std::log(base)); };
    return bucket((double)d->width * d->height * d->views, 4.0) | bucket((double)d->T, 8.0) << 8;
}

extern "C" int vcl_server_create(const vcl_server_config* c, vcl_server** out) {
    VCL_REQUIRE(c && out && c->hbm_budget > 0 && c->max_active >= 1 && c->steps_per_round >= 1);
    vcl_server* S = new vcl_server();
    S->cfg = *c;
    S->pool.budget = c->hbm_budget;
    S->t0 = clk::now();
    *out = S;
    return VCL_OK;
}

static void drop_group(vcl_server* S, GroupSlot* G) {
    if (G->stream) (void)hipStreamSynchronize(G->stream);
    if (G->g) vcl_group_destroy(G->g);
    if (G->ev) (void)hipEventDestroy(G->ev);
    if (G->stream) S->streams.push_back(G->stream);
    if (G->table) S->pool.put(G->table);
    delete G;
}

extern "C" int vcl_server_destroy(vcl_server* S) {
    if (!S) return VCL_OK;
    for (GroupSlot* G : S->groups) drop_group(S, G);
    S->groups.clear();
    for (Req* r : S->active) if (r->stream) (void)hipStreamSynchronize(r->stream);
    for (auto* v : {&S->active, &S->finished}) {
        for (Req* r : *v) { release_req(S, r); delete r; }
        v->clear();
    }
    for (Req* r : S->queue) delete r;
    S->queue.clear();
    for (hipStream_t s : S->streams) (void)hipStreamDestroy(s);
    S->pool.release_all();
    delete S;
    return VCL_OK;
}

extern "C" int vcl_server_reset_clock(vcl_server* S) {
    VCL_REQUIRE(S);
    S->t0 = clk::now();
    S->occ_integral = 0;
    S->occ_last_t = 0;
    return VCL_OK;
}

extern "C" double vcl_server_clock(const vcl_server* S) { return S ? S->now() : 0.0; }

extern "C" int vcl_server_submit(vcl_server* S, const vcl_fit_desc* d, int iterations, double arrival_s, uint64_t* ticket) {
    VCL_REQUIRE(S && d && iterations >= 1);
    size_t bytes = vcl_fit_arena_bytes(d);
    VCL_REQUIRE(bytes > 0);
    if (bytes > S->cfg.hbm_budget) return VCL_ERR_OOM;
    Req* r = new Req();
    r->ticket = S->next_ticket++;
    r->d = *d;
    r->iterations = iterations;
    r->arrival = arrival_s;
    r->bytes = bytes;
    r->cost = (double)d->width * d->height * d->views + 8.0 * d->T * d->views;
    S->queue.push_back(r);
    if (ticket) *ticket = r->ticket;
    return VCL_OK;
}

static int launch_steps(vcl_server* S, Req* r) {
    int k = std::min(S->cfg.steps_per_round, r->iterations - r->launched);
    for (int i = 0; i < k; i++) {
        int st = S->cfg.use_graph ? vcl_fit_step_graph(r->fit, (vcl_stream_t)r->stream) : vcl_fit_step(r->fit, (vcl_stream_t)r->stream);
        if (st) return st;
    }
    r->launched += k;
    VCL_HIP_CHECK(hipEventRecord(r->ev, r->stream));
    return VCL_OK;
}

// Grouped mode: steps the due requests (idle, with iterations left) as one group per step
// signature.  A group whose requests are exactly one part is reused; groups not reused this
// round are dropped once their queued work has completed.
static int launch_grouped(vcl_server* S, const std::vector<Req*>& due) {
    // one part per (step signature, pixel bucket of 4x): requests of like size step in lock
    // step well, while a group of a large and a small request would hold the small one to the
    // large one's pace at every launch
    std::map<std::pair<uint64_t, uint64_t>, std::vector<Req*>> parts;
    const double max_px = S->cfg.group_max_pixels > 0 ? (double)S->cfg.group_max_pixels : 1e300;
    for (Req* r : due) {
        if ((double)r->d.width * r->d.height * r->d.views > max_px) {
            int st = launch_steps(S, r);
            if (st) return st;
            continue;
        }
        parts[{r->sig, size_class(&r->d) & 255}].push_back(r);
    }
    for (GroupSlot* G : S->groups) G->mark = false;
    std::vector<std::pair<GroupSlot*, std::vector<Req*>*>> plan;
    for (auto& kv : parts) {
        std::vector<Req*>& m = kv.second;
        std::sort(m.begin(), m.end(), [](const Req* a, const Req* b) { return a->ticket < b->ticket; });
        std::vector<uint64_t> key;
        for (Req* r : m) key.push_back(r->ticket);
        GroupSlot* found = nullptr;
        for (GroupSlot* G : S->groups)
            if (!G->mark && G->key == key) { found = G; break; }
        if (found) found->mark = true;
        plan.push_back({found, &m});
    }
    // drop the groups that were not reused and have nothing in flight
    for (size_t i = 0; i < S->groups.size();) {
        GroupSlot* G = S->groups[i];
        if (!G->mark && (!G->ev || hipEventQuery(G->ev) == hipSuccess)) {
            drop_group(S, G);
            S->groups.erase(S->groups.begin() + i);
        } else {
            i++;
        }
    }
    for (auto& pm : plan) {
        std::vector<Req*>& m = *pm.second;
        GroupSlot* G = pm.first;
        if (!G) {
            std::vector<vcl_fit*> fits;
            for (Req* r : m) fits.push_back(r->fit);
            size_t bytes = vcl_group_table_bytes(fits.data(), (int)fits.size());
            if (!bytes) return VCL_ERR_INVALID;
            G = new GroupSlot();
            for (Req* r : m) G->key.push_back(r->ticket);
            occupancy_sample(S);
            int g = S->pool.get(bytes, &G->table);
            if (g != 0) {
                delete G;
                if (g < 0) return VCL_ERR_HIP;
                // the tables do not fit the budget now: these requests step on their own
                for (Req* r : m) { int st = launch_steps(S, r); if (st) return st; }
                continue;
            }
            if (!S->streams.empty()) { G->stream = S->streams.back(); S->streams.pop_back(); }
            else if (hipStreamCreateWithFlags(&G->stream, hipStreamNonBlocking) != hipSuccess) { drop_group(S, G); return VCL_ERR_HIP; }
            if (hipEventCreateWithFlags(&G->ev, hipEventDisableTiming) != hipSuccess) { drop_group(S, G); return VCL_ERR_HIP; }
            int st = vcl_group_create_in(fits.data(), (int)fits.size(), G->table, bytes, (vcl_stream_t)G->stream, &G->g);
            if (st) { drop_group(S, G); return st; }
            G->mark = true;
            S->groups.push_back(G);
        }
        int k = S->cfg.steps_per_round;
        for (Req* r : m) k = std::min(k, r->iterations - r->launched);
        for (int i = 0; i < k; i++) {
            // eager on a group's first use (a one-off group is not worth a capture), then replayed
            int st = S->cfg.use_graph && G->uses > 0 ? vcl_group_step_graph(G->g, (vcl_stream_t)G->stream)
                                                     : vcl_group_step(G->g, (vcl_stream_t)G->stream);
            if (st) return st;
        }
        G->uses++;
        VCL_HIP_CHECK(hipEventRecord(G->ev, G->stream));
        for (Req* r : m) {
            r->launched += k;
            VCL_HIP_CHECK(hipEventRecord(r->ev, G->stream));
        }
    }
    return VCL_OK;
}

extern "C" int vcl_server_progress(vcl_server* S, int* pending) {
    VCL_REQUIRE(S);
    // 1. retire completed step batches; finish requests
    for (size_t i = 0; i < S->active.size();) {
        Req* r = S->active[i];
        hipError_t q = hipEventQuery(r->ev);
        if (q == hipErrorNotReady) { i++; continue; }
        if (q != hipSuccess) return VCL_ERR_HIP;
        r->done = r->launched;
        if (r->done < r->iterations) { i++; continue; }
        r->finish = S->now();
        float* lp = (float*)vcl_fit_buffer(r->fit, "loss");
        VCL_HIP_CHECK(hipMemcpyAsync(&r->loss, lp, sizeof(float), hipMemcpyDeviceToHost, r->stream));
        VCL_HIP_CHECK(hipStreamSynchronize(r->stream));
        S->active.erase(S->active.begin() + i);
        S->finished.push_back(r);
        if (S->cfg.auto_release) release_req(S, r);
    }
    // 2. admit arrived requests while a slot is free and the arena fits the budget: FIFO, or
    //    under VCL_SCHED_SRPT the arrived request with the least total work first
    double t = S->now();
    const bool srpt = S->cfg.policy == VCL_SCHED_SRPT;
    while ((int)S->active.size() < S->cfg.max_active) {
        auto pick = S->queue.end();
        for (auto it = S->queue.begin(); it != S->queue.end(); ++it) {
            if ((*it)->arrival > t) { if (!srpt) break; continue; }
            if (!srpt) { pick = it; break; }
            if (pick == S->queue.end() || (*it)->cost * (*it)->iterations < (*pick)->cost * (*pick)->iterations) pick = it;
        }
        if (pick == S->queue.end()) break;
        Req* r = *pick;
        occupancy_sample(S);
        void* arena = nullptr;
        int g = S->pool.get(r->bytes, &arena);
        if (g < 0) return VCL_ERR_HIP;
        if (g > 0) break;   // wait for memory to be released
        S->queue.erase(pick);
        r->arena = arena;
        if (!S->streams.empty()) { r->stream = S->streams.back(); S->streams.pop_back(); }
        else VCL_HIP_CHECK(hipStreamCreateWithFlags(&r->stream, hipStreamNonBlocking));
        VCL_HIP_CHECK(hipEventCreateWithFlags(&r->ev, hipEventDisableTiming));
        int st = vcl_fit_create_in(&r->d, arena, r->bytes, (vcl_stream_t)r->stream, &r->fit);
        if (st) {
            r->status = st;
            r->finish = S->now();
            S->finished.push_back(r);
            release_req(S, r);
            continue;
        }
        if (S->cfg.autotune_steps > 0) {
            // the schedule of this size class: tuned on its first request, reused afterwards
            uint64_t key = size_class(&r->d);
            auto it = S->tunings.find(key);
            if (it != S->tunings.end()) st = vcl_fit_set_tuning(r->fit, &it->second);
            else {
                // the snapshot of the tuning trials comes from the pool, counted against the
                // budget, and the trials run on the request's own stream
                vcl_tuning t;
                size_t sb = vcl_fit_autotune_scratch_bytes(r->fit);
                void* snap = nullptr;
                int g2 = S->pool.get(sb, &snap);
                if (g2 < 0) return VCL_ERR_HIP;
                if (g2 > 0) st = VCL_ERR_OOM;
                else {
                    st = vcl_fit_autotune_in(r->fit, S->cfg.autotune_steps, snap, sb, (vcl_stream_t)r->stream, &t, nullptr);
                    S->pool.put(snap);
                }
                if (!st) S->tunings[key] = t;
            }
            if (st) return st;
        }
        if (S->cfg.grouped && (st = vcl_fit_step_signature(r->fit, &r->sig))) return st;
        r->admit = S->now();
        S->active.push_back(r);
    }
    // 3. queue the next steps: under FIFO the idle ones of the max_stepping earliest-admitted
    //    requests (all when 0), under VCL_SCHED_SRPT those of the max_stepping requests with the
    //    least remaining work (the others stay resident, paused)
    if (S->cfg.grouped) {
        std::vector<Req*> due;
        if (srpt) {
            std::vector<Req*> order(S->active);
            std::stable_sort(order.begin(), order.end(), [](const Req* a, const Req* b) {
                return a->cost * (a->iterations - a->done) < b->cost * (b->iterations - b->done);
            });
            int lim = S->cfg.max_stepping > 0 ? S->cfg.max_stepping : (int)order.size();
            for (int i = 0; i < (int)order.size() && i < lim; i++)
                if (order[i]->launched == order[i]->done && order[i]->launched < order[i]->iterations) due.push_back(order[i]);
        } else {
            const int lim = S->cfg.max_stepping > 0 ? S->cfg.max_stepping : (int)S->active.size();
            for (int i = 0; i < (int)S->active.size() && i < lim; i++) {
                Req* r = S->active[i];
                if (r->launched == r->done && r->launched < r->iterations) due.push_back(r);
            }
        }
        int st = launch_grouped(S, due);
        if (st) return st;
    } else if (srpt) {
        std::vector<Req*> order(S->active);
        std::stable_sort(order.begin(), order.end(), [](const Req* a, const Req* b) {
            return a->cost * (a->iterations - a->done) < b->cost * (b->iterations - b->done);
        });
        int lim = S->cfg.max_stepping > 0 ? S->cfg.max_stepping : (int)order.size();
        for (int i = 0; i < (int)order.size() && i < lim; i++) {
            Req* r = order[i];
            if (r->launched == r->done && r->launched < r->iterations) {
                int st = launch_steps(S, r);
                if (st) return st;
            }
        }
    } else {
        // FIFO: the max_stepping earliest-admitted requests (all when 0)
        const int lim = S->cfg.max_stepping > 0 ? S->cfg.max_stepping : (int)S->active.size();
        for (int i = 0; i < (int)S->active.size() && i < lim; i++) {
            Req* r = S->active[i];
            if (r->launched == r->done && r->launched < r->iterations) {
                int st = launch_steps(S, r);
                if (st) return st;
            }
        }
    }
    if (pending) *pending = (int)(S->queue.size() + S->active.size());
    return VCL_OK;
}

extern "C" int vcl_server_drain(vcl_server* S) {
    VCL_REQUIRE(S);
    int pending = 1;
    while (pending) {
        int st = vcl_server_progress(S, &pending);
        if (st) return st;
        if (pending) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
    return VCL_OK;
}

extern "C" int vcl_server_finished(vcl_server* S, vcl_request_stats* out, int max, int* n) {
    VCL_REQUIRE(S && n && (max == 0 || out));
    int k = 0;
    for (Req* r : S->finished) {
        if (r->reported || k >= max) continue;
        vcl_request_stats& o = out[k++];
        o.ticket = r->ticket;
        o.arrival_s = r->arrival;
        o.admit_s = r->admit;
        o.finish_s = r->finish;
        o.iterations = r->done;
        o.final_loss = r->loss;
        o.device_bytes = r->bytes;
        o.status = r->status;
        r->reported = true;
    }
    *n = k;
    return VCL_OK;
}

static Req* find_finished(vcl_server* S, uint64_t ticket) {
    for (Req* r : S->finished) if (r->ticket == ticket) return r;
    return nullptr;
}

extern "C" int vcl_server_read(vcl_server* S, uint64_t ticket, const char* name, void* host, size_t bytes) {
    VCL_REQUIRE(S && name && host);
    Req* r = find_finished(S, ticket);
    VCL_REQUIRE(r && r->fit);
    return vcl_fit_read(r->fit, name, host, bytes, (vcl_stream_t)r->stream);
}

extern "C" int vcl_server_release(vcl_server* S, uint64_t ticket) {
    VCL_REQUIRE(S);
    for (size_t i = 0; i < S->finished.size(); i++) {
        Req* r = S->finished[i];
        if (r->ticket != ticket) continue;
        release_req(S, r);
        delete r;
        S->finished.erase(S->finished.begin() + i);
        return VCL_OK;
    }
    return VCL_ERR_INVALID;
}

extern "C" int vcl_server_memory(vcl_server* S, size_t* in_use, size_t* peak, size_t* cached, double* mean_in_use) {
    VCL_REQUIRE(S);
    occupancy_sample(S);
    if (in_use) *in_use = S->pool.in_use;
    if (peak) *peak = S->pool.peak;
    if (cached) *cached = S->pool.cached;
    if (mean_in_use) *mean_in_use = S->occ_last_t > 0 ? S->occ_integral / S->occ_last_t : 0.0;
    return VCL_OK;
}

extern "C" int vcl_server_tunings(vcl_server* S, uint64_t* keys, vcl_tuning* out, int max, int* n) {
    VCL_REQUIRE(S && n && (max == 0 || (keys && out)));
    int k = 0;
    for (auto& kv : S->tunings) {
        if (k >= max) break;
        keys[k] = kv.first;
        out[k] = kv.second;
        k++;
    }
    *n = k;
    return VCL_OK;
}

extern "C" int vcl_server_load_tunings(vcl_server* S, const uint64_t* keys, const vcl_tuning* t, int n) {
    VCL_REQUIRE(S && n >= 0 && (n == 0 || (keys && t)));
    for (int i = 0; i < n; i++) {
        vcl_tuning saved = g_vcl_tuning;
        int st = vcl_set_tuning(&t[i]);   // validates
        g_vcl_tuning = saved;
        if (st) return st;
        S->tunings[keys[i]] = t[i];
    }
    return VCL_OK;
}

