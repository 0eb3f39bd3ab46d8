// server.cpp - continuous-batching request engine over resident worlds.
//
// One worker thread owns the device work of a server:
//   1. it retires launches whose completion event has fired (latency is
//      stamped there) and frees their worlds;
//   2. while fewer than `streams` launches are in flight it builds the next
//      launch from the pending queue in arrival order: a request is taken if
//      its world is idle, not yet in this launch, no earlier pending request
//      of the same world was passed over, and the world is (or can be made)
//      resident within the HBM budget by evicting idle worlds (those the
//      pending queue names last, then the cheapest, then the least recently
//      used).  With a budget and a backlog, up to three quarters of the
//      launch first come from requests whose worlds are resident already,
//      within the next kLookahead launches' worth of the queue (each world's
//      requests still in order);
//   3. the taken radiance requests go into one batched radiance launch and
//      the glossy requests into one batched glossy launch, on the next stream,
//      followed by a completion event.
// Requests are therefore overlapped across streams and packed into launches
// as they arrive, while every request computes what it would alone.
//
// HBM budget.  With a budget every world lives either in the server's arena
// (one device allocation of the budget's size, made at creation and handed
// out first-fit with coalescing) or as a pinned host snapshot; worlds added
// while the arena has room move into it.  Evictions (device -> snapshot)
// and restores (snapshot -> arena) are copies queued on the server's copy
// stream, in stream order with each other; a launch that needs restored
// worlds waits for the latest of those restores with one event and sets up
// the rest of their blocks (one batched kernel) on its own stream, so the
// copy stream carries copies only, the worker never blocks on a swap and no
// allocation happens while serving.  While every launch slot is busy the
// worker restores the worlds of the next requests ahead of time (prefetch).
//
// Direct mode (the plain engine): each request is its own lumx_radiance_update
// or lumx_glossy call, with kernel arguments by value, queued on one stream in
// arrival order as soon as its world is resident, up to kDirectDepth calls
// deep, with no completion wait between them: the reference's one dispatch
// per request (src/renderer/src/internal_render.cpp:1204-1206).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstdint>
#include <deque>
#include <iterator>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "host_internal.hpp"

using namespace lumx;

namespace {

using Clock = std::chrono::steady_clock;

double seconds_since(Clock::time_point t0, Clock::time_point t) {
    return std::chrono::duration<double>(t - t0).count();
}

struct Req {
    lumx_request spec;
    Clock::time_point t_submit, t_done;
    bool done = false;
    int status = LUMX_OK;
};

struct ServedWorld {
    lumx_world* w = nullptr;
    int busy = 0;
    unsigned long long last_use = 0;
    size_t off = 0;  // arena offset while resident (with a budget)
    bool static_kept = false;  // its snapshot holds its static state, taken in this server
    hipEvent_t ready = nullptr;  // recorded on the copy stream after its last restore
    bool ready_pending = false;  // no launch has waited for that restore yet
    unsigned long long ready_seq = 0;  // copy-stream order of that restore
    bool init_pending = false;   // restored, the rest of its block not yet set up
};

struct Flight {
    std::vector<Req*> reqs;
    std::vector<int> worlds;
    hipEvent_t done = nullptr;
    int stream = 0;
};

// Calls queued at once by the direct engine.
constexpr int kDirectDepth = 256;
// Launches' worth of queued requests a batched launch looks through for
// worlds that are resident already.
constexpr int kLookahead = 16;

// First-fit ranges of the arena, coalesced on release.
struct Arena {
    char* base = nullptr;
    size_t bytes = 0;
    std::map<size_t, size_t> free;  // offset -> size
    bool take(size_t need, size_t* off) {
        for (auto it = free.begin(); it != free.end(); ++it) {
            if (it->second < need) continue;
            *off = it->first;
            const size_t rest = it->second - need, at = it->first + need;
            free.erase(it);
            if (rest) free[at] = rest;
            return true;
        }
        return false;
    }
    // Takes [off, off + need) if it lies inside one free range.
    bool take_at(size_t off, size_t need) {
        auto it = free.upper_bound(off);
        if (it == free.begin()) return false;
        --it;
        const size_t f0 = it->first, f1 = it->first + it->second;
        if (off < f0 || off + need > f1) return false;
        free.erase(it);
        if (off > f0) free[f0] = off - f0;
        if (f1 > off + need) free[off + need] = f1 - (off + need);
        return true;
    }
    void give(size_t off, size_t size) {
        auto next = free.lower_bound(off);
        if (next != free.end() && off + size == next->first) {
            size += next->second;
            next = free.erase(next);
        }
        if (next != free.begin()) {
            auto prev = std::prev(next);
            if (prev->first + prev->second == off) {
                prev->second += size;
                return;
            }
        }
        free[off] = size;
    }
};

// A run of the arena in address order: a free gap (k = -1) or resident world k.
struct ArenaSeg {
    size_t off, size;
    int k;
};

}  // namespace

struct lumx_server {
    lumx_context* ctx = nullptr;
    lumx_server_desc desc{};
    Launcher* L = nullptr;
    std::vector<hipStream_t> streams;
    std::vector<hipEvent_t> free_events;
    int next_stream = 0;
    hipStream_t copy_stream = nullptr;  // evictions and restores
    unsigned long long copy_seq = 0;    // restores recorded on it so far
    Arena arena;
    std::vector<ArenaSeg> segs;  // the arena as runs, while segs_valid
    bool segs_valid = false;     // cleared by every change of the arena

    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::vector<ServedWorld> worlds;
    std::deque<Req*> pending;
    std::vector<std::unique_ptr<Req>> tickets;
    std::vector<Flight> flights;
    bool stop = false;
    // Something changed since prefetch last ran (a launch retired or was
    // queued, a request arrived): what it could not restore then may fit now.
    bool prefetch_ready = true;
    unsigned long long use_clock = 0;
    size_t resident_bytes = 0;
    lumx_server_stats stats{};
    std::thread worker;
};

static bool resident(const lumx_world* w) { return w->blocks != nullptr; }

// Threads a request adds to its batched launch, bounded above: a radiance
// request takes n_probes * rays / K threads padded to the workgroup (K >= 1,
// workgroups of at most 1024), a glossy one 256 per 16 x 16 pixel tile.  A
// launch's total must stay below 2^31.
static const long long kLaunchThreads = (1ll << 31) - 1;
static long long request_threads(const lumx_request& r, const lumx_world* w) {
    if (r.kind == LUMX_REQ_RADIANCE) {
        const long long n = r.requests_dev ? r.n_requests : w->n_blocks;
        const long long probes = n / std::max(r.params.size, 1);
        return (probes * r.params.rays + 1023) / 1024 * 1024;
    }
    return (long long)((r.width + 15) / 16) * ((r.height + 15) / 16) * 256;
}

// Queue position of each world's next pending request (SIZE_MAX: none): the
// queue is the future the server can see, and evicting a world it names soon
// only brings that world back.  Taken once per launch being built: requests
// leaving the queue meanwhile belong to worlds that launch pins anyway.
static std::vector<size_t> next_use(const lumx_server* s) {
    std::vector<size_t> next(s->worlds.size(), SIZE_MAX);
    size_t pos = 0;
    for (const Req* r : s->pending) {
        size_t& nx = next[r->spec.world];
        if (nx == SIZE_MAX) nx = pos;
        ++pos;
    }
    return next;
}

// Make world i resident within the budget (caller holds the lock; `taken`
// marks worlds already in the launch being built, `next` is next_use()).
// The world needs one contiguous arena range: among the windows of that size
// made only of free space and idle resident worlds, the one whose worlds the
// pending queue names last (then: whose eviction copies the fewest bytes,
// then: whose worlds were used least recently) is cleared and taken.
// Nothing is evicted when no such window exists (busy worlds split the
// arena); the request then waits.  The restore is queued on the copy stream.
static bool ensure_resident(lumx_server* s, int i, const std::vector<char>& taken, const std::vector<size_t>& next) {
    ServedWorld& sw = s->worlds[i];
    if (resident(sw.w)) return true;
    const size_t need = sw.w->device_bytes;
    if (!s->desc.hbm_budget) {
        // No budget: a world evicted before it was added comes back into a
        // block of its own.
        if (lumx_world_restore(sw.w) != LUMX_OK) return false;
        s->resident_bytes += need;
        s->stats.restores++;
        return true;
    }
    size_t off = 0;
    if (s->arena.take(need, &off)) {
        s->segs_valid = false;
    } else {
        // The arena as segments in address order: free gaps (k = -1) and
        // resident worlds; pinned ones (busy, taken) end a window.  Kept
        // between calls while the arena is unchanged (a launch being built
        // asks for many worlds that do not fit).
        if (!s->segs_valid) {
            s->segs.clear();
            for (const auto& f : s->arena.free) s->segs.push_back({f.first, f.second, -1});
            for (int k = 0; k < (int)s->worlds.size(); ++k) {
                const ServedWorld& o = s->worlds[k];
                if (resident(o.w) && !o.w->dev_own) s->segs.push_back({o.off, o.w->device_bytes, k});
            }
            std::sort(s->segs.begin(), s->segs.end(),
                      [](const ArenaSeg& x, const ArenaSeg& y) { return x.off < y.off; });
            s->segs_valid = true;
        }
        const std::vector<ArenaSeg>& seg = s->segs;
        auto pinned = [&](const ArenaSeg& g) {
            if (g.k < 0) return false;
            const ServedWorld& o = s->worlds[g.k];
            return g.k == i || o.busy > 0 || taken[g.k];
        };
        // Best window: the one whose soonest-needed world is needed last
        // (none at all is best), then the fewest bytes copied out, then the
        // least recently used.
        int best = -1;
        size_t best_cost = 0, best_soonest = 0;
        unsigned long long best_age = 0;
        for (int a = 0; a < (int)seg.size(); ++a) {
            if (pinned(seg[a])) continue;
            size_t got = 0, cost = 0, soonest = SIZE_MAX;
            unsigned long long age = 0;
            int b = a;
            for (; b < (int)seg.size() && got < need; ++b) {
                // The segments tile the arena; a pinned world ends the window.
                if (pinned(seg[b]) || (b > a && seg[b].off != seg[b - 1].off + seg[b - 1].size)) break;
                got = seg[b].off + seg[b].size - seg[a].off;
                if (seg[b].k >= 0) {
                    cost += seg[b].size;
                    age = std::max(age, s->worlds[seg[b].k].last_use);
                    soonest = std::min(soonest, next[seg[b].k]);
                }
            }
            if (got < need) continue;
            const bool better = best < 0 || soonest > best_soonest ||
                                (soonest == best_soonest && (cost < best_cost || (cost == best_cost && age < best_age)));
            if (better) {
                best = a;
                best_cost = cost;
                best_soonest = soonest;
                best_age = age;
            }
        }
        if (best < 0) return false;
        const size_t start = seg[best].off;
        const std::vector<ArenaSeg> window(seg.begin() + best, seg.end());
        s->segs_valid = false;
        for (int b = 0; b < (int)window.size() && window[b].off < start + need; ++b) {
            if (window[b].k < 0) continue;
            ServedWorld& v = s->worlds[window[b].k];
            // Serving changes only fields: after the world's first snapshot in
            // this server only its field goes out.
            // (A world restored ahead of a launch that never ran sends back
            // the field it received: the set-up it skipped is not needed.)
            if (world_snapshot_async(v.w, s->copy_stream, v.static_kept) != LUMX_OK) return false;
            v.static_kept = true;
            v.init_pending = false;
            s->arena.give(v.off, v.w->device_bytes);
            s->resident_bytes -= v.w->device_bytes;
            s->stats.evictions++;
        }
        if (!s->arena.take_at(start, need)) return false;
        off = start;
    }
    // Batched engine: only the copies go on the copy stream; the kernel that
    // sets up the rest of the block runs on the stream of the world's next
    // launch, so the copy stream carries copies alone and no launch waits
    // behind another world's set-up.  The direct engine's one launch stream
    // runs its calls back to back, so there the set-up stays on the copy
    // stream, beside them.
    const bool defer = !s->desc.direct;
    const int rc = defer ? world_attach_copy_async(sw.w, s->arena.base + off, false, s->copy_stream)
                         : world_attach_async(sw.w, s->arena.base + off, false, s->copy_stream);
    if (rc != LUMX_OK) {
        s->arena.give(off, need);
        s->segs_valid = false;
        return false;
    }
    sw.off = off;
    s->resident_bytes += need;
    s->stats.restores++;
    sw.init_pending = defer;
    // Launches of this world wait for its restore (and, the copy stream being
    // in order, for the evictions queued before it).
    if (hipEventRecord(sw.ready, s->copy_stream) != hipSuccess) return false;
    sw.ready_pending = true;
    sw.ready_seq = ++s->copy_seq;
    return true;
}

// Launches on `st` of flight f start after its worlds' restores, and a
// restored world's block is set up on `st` first.  A world is in one flight
// at a time (the direct engine's calls share one stream), so whatever uses
// the world next is ordered after that set-up.
static int order_after_restores(lumx_server* s, const Flight& f, hipStream_t st) {
    // The copy stream runs in order: waiting for the latest of the flight's
    // restores waits for all of them.
    std::vector<lumx_world*> init;
    ServedWorld* last = nullptr;
    for (int wi : f.worlds) {
        ServedWorld& sw = s->worlds[wi];
        if (sw.ready_pending && (!last || sw.ready_seq > last->ready_seq)) last = &sw;
        if (sw.init_pending) init.push_back(sw.w);
    }
    if (last) HIP_TRY(hipStreamWaitEvent(st, last->ready, 0));
    for (int wi : f.worlds) s->worlds[wi].ready_pending = false;
    if (init.empty()) return LUMX_OK;
    // One set-up launch for all of them.
    int rc = world_attach_init_many(init.data(), (int)init.size(), st);
    if (rc) return rc;
    for (int wi : f.worlds) s->worlds[wi].init_pending = false;
    return LUMX_OK;
}

// While every launch slot is taken: restore the worlds of the next pending
// requests (at most one launch's worth), evicting only idle worlds that none
// of those requests names, so their launch finds them resident.
static void prefetch(lumx_server* s) {
    const int nw = (int)s->worlds.size();
    std::vector<char> keep(nw, 0);
    int n = 0;
    for (auto it = s->pending.begin(); it != s->pending.end() && n < s->desc.max_batch; ++it, ++n)
        keep[(*it)->spec.world] = 1;
    const std::vector<size_t> next = next_use(s);
    n = 0;
    for (auto it = s->pending.begin(); it != s->pending.end() && n < s->desc.max_batch; ++it, ++n) {
        const int wi = (*it)->spec.world;
        if (resident(s->worlds[wi].w)) continue;
        if (!ensure_resident(s, wi, keep, next)) break;
    }
}

// One request as its own single-world call (direct mode).
static int launch_direct(lumx_server* s, Flight& f) {
    hipStream_t st = s->streams[f.stream];
    int rc = order_after_restores(s, f, st);
    if (rc) return rc;
    const lumx_request& r = f.reqs[0]->spec;
    lumx_world* w = s->worlds[f.worlds[0]].w;
    if (r.kind == LUMX_REQ_RADIANCE)
        rc = lumx_radiance_update(w, r.requests_dev, r.requests_dev ? r.n_requests : 0, &r.params, st);
    else
        rc = lumx_glossy(w, r.matnorm_dev, r.depth_dev, r.width, r.height, &r.cam, r.max_dist, r.max_steps,
                         r.out_dev, nullptr, st);
    if (rc) return rc;
    HIP_TRY(hipEventRecord(f.done, st));
    return LUMX_OK;
}

static int launch_flight(lumx_server* s, Flight& f) {
    hipStream_t st = s->streams[f.stream];
    int rc0 = order_after_restores(s, f, st);
    if (rc0) return rc0;
    std::vector<lumx_world*> rw, gw;
    std::vector<const int32_t*> rq;
    std::vector<int> rn;
    std::vector<lumx_radiance_params> rp;
    std::vector<const uint8_t*> gm;
    std::vector<const float*> gd;
    std::vector<int> gwid, ght, gms;
    std::vector<float> gmd;
    std::vector<lumx_camera> gc;
    std::vector<float*> go;
    for (size_t k = 0; k < f.reqs.size(); ++k) {
        const lumx_request& r = f.reqs[k]->spec;
        lumx_world* w = s->worlds[f.worlds[k]].w;
        if (r.kind == LUMX_REQ_RADIANCE) {
            rw.push_back(w);
            rq.push_back(r.requests_dev);
            rn.push_back(r.requests_dev ? r.n_requests : 0);
            rp.push_back(r.params);
        } else {
            gw.push_back(w);
            gm.push_back(r.matnorm_dev);
            gd.push_back(r.depth_dev);
            gwid.push_back(r.width);
            ght.push_back(r.height);
            gc.push_back(r.cam);
            gmd.push_back(r.max_dist);
            gms.push_back(r.max_steps);
            go.push_back(r.out_dev);
        }
    }
    bool any_list = false;
    for (const int32_t* p : rq) any_list |= (p != nullptr);
    int rc = launcher_radiance(s->L, rw.data(), (int)rw.size(), any_list ? rq.data() : nullptr, rn.data(), rp.data(),
                               false, st, true);
    if (rc) return rc;
    rc = launcher_glossy(s->L, gw.data(), (int)gw.size(), gm.data(), gd.data(), gwid.data(), ght.data(), gc.data(),
                         gmd.data(), gms.data(), go.data(), st);
    if (rc) return rc;
    HIP_TRY(hipEventRecord(f.done, st));
    return LUMX_OK;
}

static void worker_loop(lumx_server* s) {
    hipSetDevice(s->ctx->device);
    std::unique_lock<std::mutex> lk(s->mu);
    while (true) {
        // 1. retire finished launches (the direct engine's are in stream
        // order: stop at the first one still running)
        bool retired = false;
        for (size_t k = 0; k < s->flights.size();) {
            Flight& f = s->flights[k];
            if (hipEventQuery(f.done) == hipSuccess) {
                const Clock::time_point now = Clock::now();
                for (size_t j = 0; j < f.reqs.size(); ++j) {
                    Req* r = f.reqs[j];
                    r->t_done = now;
                    r->done = true;
                    s->worlds[f.worlds[j]].busy--;
                    s->stats.completed++;
                    if (r->spec.kind == LUMX_REQ_RADIANCE) {
                        const lumx_world* w = s->worlds[f.worlds[j]].w;
                        long long n = r->spec.requests_dev ? r->spec.n_requests : w->n_blocks;
                        s->stats.probe_rays += (n / r->spec.params.size) * (long long)r->spec.params.rays;
                    } else {
                        s->stats.pixels += (long long)r->spec.width * r->spec.height;
                    }
                }
                s->free_events.push_back(f.done);
                s->flights.erase(s->flights.begin() + k);
                retired = true;
            } else if (s->desc.direct) {
                break;
            } else {
                ++k;
            }
        }
        if (retired) {
            s->prefetch_ready = true;
            s->cv_done.notify_all();
        }
        if (s->stop && s->pending.empty() && s->flights.empty()) break;

        // 2. build the next launch
        bool launched = false;
        const int depth = s->desc.direct ? kDirectDepth : s->desc.streams;
        if ((int)s->flights.size() < depth && !s->pending.empty()) {
            const int nw = (int)s->worlds.size();
            std::vector<char> taken(nw, 0), blocked(nw, 0);
            // Windows only shrink while a launch is built (its worlds are
            // pinned as they join it): once a world of `no_room` bytes found
            // none, no larger world will.
            std::vector<size_t> next;
            size_t no_room = SIZE_MAX;
            Flight f;
            long long threads[2] = {0, 0};  // radiance, glossy threads of the launch being built
            // Request *it of world wi joins the launch; the next one in the queue.
            auto admit = [&](std::deque<Req*>::iterator it, int wi, int kind, long long t) {
                taken[wi] = 1;
                threads[kind] += t;
                s->worlds[wi].busy++;
                s->worlds[wi].last_use = ++s->use_clock;
                f.reqs.push_back(*it);
                f.worlds.push_back(wi);
                return s->pending.erase(it);
            };
            // With a backlog (more than two launches' worth queued): first,
            // among the next kLookahead launches' worth of requests, those
            // whose worlds are resident now (no swap before they can run), for
            // up to three quarters of the launch; then the rest in queue
            // order, below, so the head of the queue moves every launch.  A
            // world skipped here is skipped for the rest of this pass, so each
            // world's requests keep their order.  Without a backlog requests
            // go strictly in queue order, which keeps the latency tail short.
            if (!s->desc.direct && s->desc.hbm_budget && s->desc.max_batch >= 4 &&
                s->pending.size() > 2 * (size_t)s->desc.max_batch) {
                std::vector<char> skip(nw, 0);
                int seen = 0;
                const int horizon = kLookahead * s->desc.max_batch;
                const int first_pass = s->desc.max_batch - s->desc.max_batch / 4;
                for (auto it = s->pending.begin(); it != s->pending.end() && seen < horizon &&
                                                   (int)f.reqs.size() < first_pass;
                     ++seen) {
                    Req* r = *it;
                    const int wi = r->spec.world;
                    bool ok = !skip[wi] && !taken[wi] && s->worlds[wi].busy == 0 && resident(s->worlds[wi].w);
                    if (ok && !f.reqs.empty())
                        ok = s->worlds[wi].w->desc.radiance_format ==
                             s->worlds[f.worlds[0]].w->desc.radiance_format;
                    const int kind = r->spec.kind == LUMX_REQ_RADIANCE ? 0 : 1;
                    const long long t = request_threads(r->spec, s->worlds[wi].w);
                    if (ok && threads[kind] + t > kLaunchThreads) ok = false;
                    if (!ok) {
                        skip[wi] = 1;
                        ++it;
                        continue;
                    }
                    it = admit(it, wi, kind, t);
                }
            }
            for (auto it = s->pending.begin(); it != s->pending.end() && (int)f.reqs.size() < s->desc.max_batch;) {
                Req* r = *it;
                const int wi = r->spec.world;
                if (s->desc.direct) {
                    // Arrival order, one request per call; a world's calls
                    // are ordered by the stream, so it need not be idle.
                    if (!f.reqs.empty()) break;
                    if (!resident(s->worlds[wi].w) && !ensure_resident(s, wi, taken, next_use(s))) break;
                    admit(it, wi, r->spec.kind == LUMX_REQ_RADIANCE ? 0 : 1, 0);
                    break;
                }
                bool ok = !blocked[wi] && !taken[wi] && s->worlds[wi].busy == 0;
                if (ok && !f.reqs.empty()) {
                    // one radiance format per launch
                    ok = s->worlds[wi].w->desc.radiance_format == s->worlds[f.worlds[0]].w->desc.radiance_format;
                }
                const int kind = r->spec.kind == LUMX_REQ_RADIANCE ? 0 : 1;
                const long long t = request_threads(r->spec, s->worlds[wi].w);
                if (ok && threads[kind] + t > kLaunchThreads) {
                    // The launch is full: this request and every later one of
                    // its world wait for the next launch.
                    ok = false;
                }
                if (ok && !resident(s->worlds[wi].w)) {
                    const size_t need = s->worlds[wi].w->device_bytes;
                    if (need >= no_room) {
                        ok = false;
                    } else {
                        if (next.empty()) next = next_use(s);
                        ok = ensure_resident(s, wi, taken, next);
                        if (!ok) no_room = need;
                    }
                }
                if (!ok) {
                    blocked[wi] = 1;
                    ++it;
                    continue;
                }
                it = admit(it, wi, kind, t);
            }
            if (!f.reqs.empty()) {
                int rc = LUMX_OK;
                if (s->free_events.empty()) {
                    hipEvent_t e = nullptr;
                    if (hipEventCreateWithFlags(&e, hipEventDisableTiming) == hipSuccess)
                        s->free_events.push_back(e);
                    else
                        rc = fail(LUMX_EDEVICE, "completion event creation failed");
                }
                if (rc == LUMX_OK) {
                    f.done = s->free_events.back();
                    s->free_events.pop_back();
                    f.stream = s->next_stream;
                    s->next_stream = (s->next_stream + 1) % (int)s->streams.size();
                    rc = s->desc.direct ? launch_direct(s, f) : launch_flight(s, f);
                    s->stats.launches++;
                }
                if (rc != LUMX_OK) {
                    // Record the failure on the requests; their worlds are
                    // released once whatever part of the launch was queued
                    // has finished.
                    if (f.done) (void)hipStreamSynchronize(s->streams[f.stream]);
                    const Clock::time_point now = Clock::now();
                    for (size_t j = 0; j < f.reqs.size(); ++j) {
                        f.reqs[j]->status = rc;
                        f.reqs[j]->t_done = now;
                        f.reqs[j]->done = true;
                        s->worlds[f.worlds[j]].busy--;
                    }
                    if (f.done) s->free_events.push_back(f.done);
                    s->cv_done.notify_all();
                } else {
                    s->flights.push_back(std::move(f));
                }
                launched = true;
            }
        }
        if (launched) {
            s->prefetch_ready = true;
            continue;
        }
        if (!s->desc.direct && s->desc.hbm_budget && (int)s->flights.size() >= depth && !s->pending.empty() &&
            s->prefetch_ready) {
            s->prefetch_ready = false;
            prefetch(s);
        }
        // 3. wait: poll in-flight events briefly, or sleep until a submission
        if (!s->flights.empty()) {
            lk.unlock();
            std::this_thread::sleep_for(std::chrono::microseconds(5));
            lk.lock();
        } else {
            s->cv_work.wait_for(lk, std::chrono::milliseconds(2));
        }
    }
}

extern "C" {

int lumx_server_create(lumx_context_t ctx, const lumx_server_desc* d, lumx_server_t* out) {
    if (!ctx || !d || !out) return fail(LUMX_EINVAL, "null argument");
    if (d->max_batch < 1 || d->max_batch > 4096 || d->streams < 1 || d->streams > 16)
        return fail(LUMX_EINVAL, "max_batch must be in [1, 4096] and streams in [1, 16]");
    HIP_TRY(hipSetDevice(ctx->device));
    lumx_server* s = new lumx_server();
    s->ctx = ctx;
    s->desc = *d;
    int rc = launcher_create(ctx, d->max_batch, 4 * d->streams + 4, &s->L);
    if (rc) {
        delete s;
        return rc;
    }
    rc = launcher_set_shape(s->L, d->block, d->rays_per_lane);
    if (rc) {
        launcher_destroy(s->L);
        delete s;
        return rc;
    }
    const int n_streams = d->direct ? 1 : d->streams;
    bool ok = hipStreamCreateWithFlags(&s->copy_stream, hipStreamNonBlocking) == hipSuccess;
    if (ok && d->hbm_budget) {
        // The arena: every byte of world state the server may hold.
        void* p = nullptr;
        if (hipMalloc(&p, d->hbm_budget) != hipSuccess) {
            if (s->copy_stream) hipStreamDestroy(s->copy_stream);
            launcher_destroy(s->L);
            delete s;
            return fail(LUMX_ENOMEM, "HBM budget of %zu bytes could not be allocated", d->hbm_budget);
        }
        s->arena.base = (char*)p;
        s->arena.bytes = d->hbm_budget;
        s->arena.free[0] = d->hbm_budget;
    }
    for (int k = 0; ok && k < n_streams; ++k) {
        hipStream_t st;
        ok = hipStreamCreateWithFlags(&st, hipStreamNonBlocking) == hipSuccess;
        if (ok) s->streams.push_back(st);
    }
    if (!ok) {
        for (hipStream_t t : s->streams) hipStreamDestroy(t);
        if (s->copy_stream) hipStreamDestroy(s->copy_stream);
        if (s->arena.base) hipFree(s->arena.base);
        launcher_destroy(s->L);
        delete s;
        return fail(LUMX_EDEVICE, "stream creation failed");
    }
    s->worker = std::thread(worker_loop, s);
    *out = s;
    return LUMX_OK;
}

int lumx_server_destroy(lumx_server_t s) {
    if (!s) return fail(LUMX_EINVAL, "null server");
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    s->worker.join();
    hipSetDevice(s->ctx->device);
    for (hipStream_t t : s->streams) {
        hipStreamSynchronize(t);
        hipStreamDestroy(t);
    }
    int status = LUMX_OK;
    // Worlds restored for launches that never ran: their blocks set up now.
    for (ServedWorld& sw : s->worlds) {
        if (!sw.init_pending || !resident(sw.w)) continue;
        if (world_attach_init(sw.w, s->copy_stream) != LUMX_OK) status = LUMX_EDEVICE;
        sw.init_pending = false;
    }
    hipStreamSynchronize(s->copy_stream);
    if (s->arena.base) {
        // Worlds in the arena move into blocks of their own (the same layout,
        // so one device copy each); a world that cannot stays on the host.
        for (ServedWorld& sw : s->worlds) {
            if (!resident(sw.w) || sw.w->dev_own) continue;
            void* mem = nullptr;
            if (hipMalloc(&mem, sw.w->device_bytes) == hipSuccess &&
                hipMemcpy(mem, sw.w->dev_mem, sw.w->device_bytes, hipMemcpyDeviceToDevice) == hipSuccess) {
                world_rebase(sw.w, mem, true);
            } else {
                if (mem) hipFree(mem);
                if (world_snapshot_async(sw.w, s->copy_stream, sw.static_kept) != LUMX_OK) status = LUMX_EDEVICE;
                hipStreamSynchronize(s->copy_stream);
            }
        }
        hipFree(s->arena.base);
    }
    hipStreamDestroy(s->copy_stream);
    for (ServedWorld& sw : s->worlds)
        if (sw.ready) hipEventDestroy(sw.ready);
    for (hipEvent_t e : s->free_events) hipEventDestroy(e);
    launcher_destroy(s->L);
    delete s;
    if (status) return fail(status, "a world could not leave the server's arena");
    return LUMX_OK;
}

int lumx_server_add_world(lumx_server_t s, lumx_world_t w, int* id) {
    if (!s || !w || !id) return fail(LUMX_EINVAL, "null argument");
    if (w->ctx != s->ctx) return fail(LUMX_EINVAL, "world belongs to another context");
    std::lock_guard<std::mutex> lk(s->mu);
    s->segs_valid = false;  // the arena may change below
    for (const ServedWorld& o : s->worlds)
        if (o.w == w) return fail(LUMX_EINVAL, "world already added");
    if (s->desc.hbm_budget) {
        if (w->device_bytes > s->desc.hbm_budget)
            return fail(LUMX_EINVAL, "world needs %zu bytes, more than the budget of %zu", w->device_bytes,
                        s->desc.hbm_budget);
    }
    HIP_TRY(hipSetDevice(s->ctx->device));
    if (s->desc.hbm_budget && resident(w)) {
        // With a budget every world is served from the arena: a resident
        // world moves there while the arena has room (one device copy of its
        // block), else to the host, into its pinned snapshot.
        HIP_TRY(hipDeviceSynchronize());  // work on the world may be queued on any stream
        // Its pinned snapshot now, so no eviction allocates while serving.
        int rc = world_snapshot_reserve(w, std::max(w->counter, w->prev_counter), s->copy_stream);
        if (rc) return rc;
        size_t off = 0;
        if (s->arena.take(w->device_bytes, &off)) {
            // Its static state into the snapshot now (before the move, so a
            // failure leaves the world as it was): evictions while serving
            // then copy only the field (a first copy into a pinned snapshot
            // has been seen to hold the host for milliseconds).
            rc = world_snapshot_static(w);
            if (rc) {
                s->arena.give(off, w->device_bytes);
                return rc;
            }
            hipError_t e = hipMemcpy(s->arena.base + off, w->dev_mem, w->device_bytes, hipMemcpyDeviceToDevice);
            if (e != hipSuccess) {
                s->arena.give(off, w->device_bytes);
                return fail(LUMX_EDEVICE, "moving a world into the arena: %s", hipGetErrorString(e));
            }
            if (w->dev_own) hipFree(w->dev_mem);
            world_rebase(w, s->arena.base + off, false);
        } else {
            rc = lumx_world_evict(w);
            if (rc) return rc;
        }
    }
    ServedWorld sw;
    sw.w = w;
    if (resident(w) && !w->dev_own) sw.off = (size_t)((char*)w->dev_mem - s->arena.base);
    // A world on the host holds a full snapshot; one in the arena, its static state.
    sw.static_kept = s->desc.hbm_budget != 0 && w->snap_static_valid;
    HIP_TRY(hipEventCreateWithFlags(&sw.ready, hipEventDisableTiming));
    s->worlds.push_back(sw);
    if (resident(w)) s->resident_bytes += w->device_bytes;
    *id = (int)s->worlds.size() - 1;
    return LUMX_OK;
}

int lumx_server_submit(lumx_server_t s, const lumx_request* r, long long* ticket) {
    if (!s || !r || !ticket) return fail(LUMX_EINVAL, "null argument");
    if (r->kind != LUMX_REQ_RADIANCE && r->kind != LUMX_REQ_GLOSSY) return fail(LUMX_EINVAL, "request kind invalid");
    std::unique_ptr<Req> q(new Req());
    q->spec = *r;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (r->world < 0 || r->world >= (int)s->worlds.size()) return fail(LUMX_EINVAL, "world id %d invalid", r->world);
        // Everything a launch would reject is rejected here, so one bad
        // request cannot fail the others packed into its launch.
        if (r->kind == LUMX_REQ_RADIANCE) {
            const int rays = r->params.rays;
            if (rays <= 0 || rays > 1024 || (rays & (rays - 1)))
                return fail(LUMX_EINVAL, "served radiance rays must be a power of two <= 1024");
            int rc = radiance_params_check(&r->params);
            if (rc) return rc;
            if (r->requests_dev && r->n_requests < 0) return fail(LUMX_EINVAL, "n_requests invalid");
            const lumx_world* w = s->worlds[r->world].w;
            if (r->requests_dev && r->n_requests / r->params.size > w->n_blocks)
                return fail(LUMX_EINVAL, "%d requests at stride %d dispatch more probes than the world's %d",
                            r->n_requests, r->params.size, w->n_blocks);
        } else if (!r->matnorm_dev || !r->depth_dev || !r->out_dev || r->width <= 0 || r->height <= 0 ||
                   r->width > 16384 || r->height > 16384) {
            return fail(LUMX_EINVAL, "glossy request buffers or size invalid");
        } else if (!(r->max_dist > 0.0f && r->max_dist <= 65536.0f) || r->max_steps < 0) {
            return fail(LUMX_EINVAL, "glossy max_dist must be in (0, 65536] and max_steps >= 0");
        }
        if (request_threads(*r, s->worlds[r->world].w) > kLaunchThreads)
            return fail(LUMX_EINVAL, "request too large for one launch (2^31 threads)");
        q->t_submit = Clock::now();
        *ticket = (long long)s->tickets.size();
        s->pending.push_back(q.get());
        s->prefetch_ready = true;
        s->tickets.push_back(std::move(q));
        s->stats.submitted++;
    }
    s->cv_work.notify_one();
    return LUMX_OK;
}

int lumx_server_wait(lumx_server_t s, long long ticket, double* latency_s) {
    if (!s) return fail(LUMX_EINVAL, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    if (ticket < 0 || ticket >= (long long)s->tickets.size()) return fail(LUMX_EINVAL, "ticket invalid");
    Req* r = s->tickets[ticket].get();
    s->cv_done.wait(lk, [&] { return r->done; });
    if (latency_s) *latency_s = seconds_since(r->t_submit, r->t_done);
    if (r->status != LUMX_OK) return fail(r->status, "request %lld failed", ticket);
    return LUMX_OK;
}

int lumx_server_drain(lumx_server_t s) {
    if (!s) return fail(LUMX_EINVAL, "null server");
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return s->pending.empty() && s->flights.empty(); });
    return LUMX_OK;
}

int lumx_server_get_stats(lumx_server_t s, lumx_server_stats* out) {
    if (!s || !out) return fail(LUMX_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *out = s->stats;
    out->resident_bytes = s->resident_bytes;
    return LUMX_OK;
}

int lumx_server_run_trace(lumx_server_t s, int n, const lumx_request* reqs, const double* arrival_s,
                          double* latency_s, double* wall_s) {
    if (!s || n < 0 || (n > 0 && (!reqs || !arrival_s))) return fail(LUMX_EINVAL, "bad argument");
    for (int i = 1; i < n; ++i)
        if (arrival_s[i] < arrival_s[i - 1]) return fail(LUMX_EINVAL, "arrival times must be non-decreasing");
    std::vector<long long> tickets(n, -1);
    std::atomic<int> submit_rc(LUMX_OK);
    const Clock::time_point t0 = Clock::now();
    std::thread submitter([&] {
        for (int i = 0; i < n; ++i) {
            std::this_thread::sleep_until(t0 + std::chrono::duration_cast<Clock::duration>(
                                                    std::chrono::duration<double>(arrival_s[i])));
            int rc = lumx_server_submit(s, &reqs[i], &tickets[i]);
            if (rc != LUMX_OK) {
                submit_rc = rc;
                return;
            }
        }
    });
    submitter.join();
    if (submit_rc != LUMX_OK) {
        lumx_server_drain(s);
        return submit_rc;
    }
    int status = LUMX_OK;
    Clock::time_point last = t0;
    for (int i = 0; i < n; ++i) {
        double lat = 0.0;
        int rc = lumx_server_wait(s, tickets[i], &lat);
        if (rc != LUMX_OK) status = rc;
        if (latency_s) latency_s[i] = lat;
        std::lock_guard<std::mutex> lk(s->mu);
        last = std::max(last, s->tickets[tickets[i]]->t_done);
    }
    if (wall_s) *wall_s = seconds_since(t0 + std::chrono::duration_cast<Clock::duration>(
                                                  std::chrono::duration<double>(n ? arrival_s[0] : 0.0)),
                                        last);
    return status;
}

}  // extern "C"
