// Serving loop: resident scenes, HBM-budget admission and continuous
// batching of render requests over one shared pool of path slots.
//
// A request of P passes over W x H pixels is a sequence of one-sample work
// units (pass k, pixel range) in the order k = 0 .. P-1, pixels ascending;
// a ReSTIR DI request's units are whole passes, since its spatial reuse reads
// neighbouring pixels of the same pass and its temporal reuse the previous
// pass.  Every step builds one batch:
//   1. admit queued requests in arrival order while their target
//      (16 B/pixel) and ReSTIR state (224 B/pixel) fit the HBM budget next
//      to the path pools;
//   2. pick the integrator configuration (scene, bounce_count,
//      light_sampling, generator) of the oldest unfinished request and gather
//      the unfinished requests that share it;
//   3. give each of them an equal share of the pool's slots in arrival
//      order, then hand what is left to the oldest first; each share is taken
//      from the request's cursor as chunks (pass k, pixels [a, b));
//   4. enqueue the wavefront kernels once over the pool (generate per chunk,
//      extend 0, the ReSTIR passes per chunk, then shade / shadow / extend per
//      bounce for all chunks together, accumulate per chunk in the order
//      taken) on the next render lane;
//   5. once `depth` batches are in flight (or no batch could be formed),
//      wait for the oldest, then download the requests it finished and keep
//      their device memory for later requests.
// Batches in flight overlap like renders on the context's lanes: a batch's
// last, latency-bound bounces run beside the next batch's first; each
// accumulates after the batch before it, so every pixel still takes its
// passes in order, and a ReSTIR request's passes in one batch run after its
// passes in the batch before (an event per request).
// A path's radiance depends only on (scene, camera, pixel, pass, seed,
// configuration) and each pixel accumulates its passes in order with
// (previous * k + new) / (k + 1), so a request's image is the same bits
// alone or in any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <new>
#include <vector>

#include "abi_guard.h"
#include "runtime_internal.h"

namespace {

// ReSTIR state per pixel: two G-buffer records of 5 float4 and four
// reservoir arrays (initial, temporal, final of this and the last pass).
constexpr uint64_t RESTIR_BYTES_PER_PIXEL = 10 * sizeof(float4) + 4 * sizeof(uint4);
// Per path slot of a batch with ReSTIR requests: final and initial
// reservoirs and the (pixel, pass) key.
constexpr uint64_t RESTIR_SLOT_BYTES = 2 * sizeof(uint4) + sizeof(uint2);

// A device allocation held by a request, or kept for the next one.
struct Block {
    void* ptr = nullptr;
    uint64_t bytes = 0;
};

struct Request {
    uint64_t id = 0;
    tt_server_request desc{};
    uint32_t width = 0, height = 0;
    uint32_t state = TT_REQUEST_QUEUED;
    // Scheduling cursor: next unit is (next_pass, next_pixel).
    uint32_t next_pass = 0, next_pixel = 0;
    uint32_t passes_done = 0;  // passes fully accumulated
    Block target;              // RGBA32F
    Block restir;              // ReSTIR G-buffers and reservoirs
    hipEvent_t restir_done = nullptr;  // after the last enqueued use of `restir`
    bool restir_recorded = false;
    bool cleared = false;      // the target has been zeroed by a batch
    std::vector<float> image;  // host copy once done
    uint64_t submitted_ns = 0, admitted_ns = 0, done_ns = 0;

    uint64_t pixels() const { return uint64_t(width) * height; }
    bool is_restir() const { return desc.constants.light_sampling == TT_LIGHT_SAMPLING_RESTIR; }
    uint64_t target_bytes() const { return pixels() * sizeof(float4); }
    uint64_t restir_bytes() const { return is_restir() ? pixels() * RESTIR_BYTES_PER_PIXEL : 0; }
    uint64_t resident_bytes() const { return target_bytes() + restir_bytes(); }
    bool scheduled_all() const { return next_pass >= desc.passes; }
};

struct Chunk {
    Request* request;
    uint32_t pass, pixel0, count, base;
};

// A batch enqueued and not yet waited for.
struct Batch {
    uint32_t lane;
    std::vector<Chunk> chunks;
    uint32_t paths;
    bool cached;              // ran with a radiance cache: its counts are read when it ends
    uint32_t train_capacity;  // its training record slots
    bool judged;              // with judge records
};

// The radiance cache serving a scene slot.
struct CacheSlot {
    tt_mlp* cache = nullptr;
    tt_cache_settings settings{};
};

// The path buffers a lane's shadow pass reads: the batch's, with the side
// stream's own traversal spill area.
tt::PathBuffers side_buffers_of(const tt::PathBuffers& b, const tt::Lane& L) {
    tt::PathBuffers s = b;
    s.spill = L.side_spill;
    return s;
}

bool same_config(const Request& a, const Request& b) {
    return a.desc.scene == b.desc.scene && a.desc.rng == b.desc.rng &&
           a.desc.constants.bounce_count == b.desc.constants.bounce_count &&
           a.desc.constants.light_sampling == b.desc.constants.light_sampling;
}

}  // namespace

struct tt_server {
    tt_context* context = nullptr;
    uint64_t budget = 0;
    uint32_t pool_paths = 0;
    std::vector<const tt_scene*> scenes;
    std::deque<Request*> queued;
    std::vector<Request*> active;  // arrival order
    std::map<uint64_t, std::unique_ptr<Request>> requests;
    uint64_t next_id = 1;
    uint64_t used = 0;          // blocks held by admitted requests
    std::vector<Block> spare;   // blocks of finished requests
    uint64_t spare_bytes = 0;
    bool restir_slots = false;  // a ReSTIR request was submitted: pools carry slot reservoirs
    uint32_t depth = 1;      // batches in flight, each with its own path pool (lanes 0 .. depth - 1)
    uint32_t next_slot = 0;  // lane of the next batch
    std::deque<Batch> inflight;
    std::map<uint32_t, CacheSlot> caches;  // scene slot -> its radiance cache
    std::chrono::steady_clock::time_point busy_since;  // first batch of the current busy period
    std::chrono::steady_clock::time_point origin;
    tt_server_stats stats{};
    double judge_traced_sum = 0.0, judge_cached_sum = 0.0;  // over judge records, per channel

    uint64_t now_ns() const {
        return uint64_t(std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now() - origin)
                            .count());
    }
    uint64_t pool_bytes(bool restir = false) const {
        return uint64_t(pool_paths) * (tt::PATH_BYTES + (restir || restir_slots ? RESTIR_SLOT_BYTES : 0)) * depth;
    }
};

namespace {

using tt::fail;

// A block of `bytes` for an admitted request: a kept block of the same size
// class (up to 1.5x) when one fits the budget, else a new allocation, after
// releasing kept blocks (hipFree waits for the device) only while the budget
// needs their room.
tt_status acquire(tt_server* sv, uint64_t bytes, Block& out) {
    tt_context* c = sv->context;
    int best = -1;
    for (size_t i = 0; i < sv->spare.size(); i++) {
        const Block& b = sv->spare[i];
        if (b.bytes >= bytes && 2 * b.bytes <= 3 * bytes && (best < 0 || b.bytes < sv->spare[best].bytes))
            best = int(i);
    }
    if (best >= 0 && sv->used + sv->spare[best].bytes + sv->pool_bytes() <= sv->budget) {
        out = sv->spare[best];
        sv->spare.erase(sv->spare.begin() + best);
        sv->spare_bytes -= out.bytes;
        sv->used += out.bytes;
        return TT_OK;
    }
    while (!sv->spare.empty() && sv->used + sv->spare_bytes + bytes + sv->pool_bytes() > sv->budget) {
        auto largest = std::max_element(sv->spare.begin(), sv->spare.end(),
                                        [](const Block& a, const Block& b) { return a.bytes < b.bytes; });
        TT_HIP(c, hipFree(largest->ptr));
        sv->spare_bytes -= largest->bytes;
        sv->spare.erase(largest);
    }
    out.bytes = bytes;
    TT_HIP(c, hipMalloc(&out.ptr, bytes));
    sv->used += bytes;
    return TT_OK;
}

void release(tt_server* sv, Block& b) {
    if (!b.ptr) return;
    sv->used -= b.bytes;
    sv->spare.push_back(b);
    sv->spare_bytes += b.bytes;
    b = Block{};
}

// Takes up to `want` path slots of work from r's cursor.  ReSTIR requests
// take whole passes, at least one when it fits the rest of the pool.
void take(tt_server* sv, Request* r, uint32_t want, uint32_t& base, std::vector<Chunk>& out) {
    const uint32_t px = uint32_t(r->pixels());
    if (r->is_restir()) {
        uint32_t taken = 0;
        while (!r->scheduled_all() && base + px <= sv->pool_paths && (taken == 0 || taken + px <= want)) {
            out.push_back(Chunk{r, r->next_pass, 0u, px, base});
            base += px;
            taken += px;
            r->next_pass++;
        }
        return;
    }
    while (want > 0 && !r->scheduled_all()) {
        uint32_t n = std::min<uint32_t>(want, px - r->next_pixel);
        out.push_back(Chunk{r, r->next_pass, r->next_pixel, n, base});
        base += n;
        want -= n;
        r->next_pixel += n;
        if (r->next_pixel == px) {
            r->next_pixel = 0;
            r->next_pass++;
        }
    }
}

uint64_t remaining_paths(const Request* r) {
    if (r->scheduled_all()) return 0;
    return uint64_t(r->desc.passes - r->next_pass) * r->pixels() - r->next_pixel;
}

tt_status admit(tt_server* sv) {
    tt_context* c = sv->context;
    while (!sv->queued.empty()) {
        Request* r = sv->queued.front();
        if (sv->used + r->resident_bytes() + sv->pool_bytes() > sv->budget) break;
        // The target is zeroed by the first batch that uses it, on that
        // batch's lane: nothing here waits for the batches in flight.
        tt_status st = acquire(sv, r->target_bytes(), r->target);
        if (st == TT_OK && r->is_restir()) {
            st = acquire(sv, r->restir_bytes(), r->restir);
            if (st == TT_OK) TT_HIP(c, hipEventCreateWithFlags(&r->restir_done, hipEventDisableTiming));
        }
        if (st != TT_OK) {
            release(sv, r->target);
            release(sv, r->restir);
            return st;
        }
        r->state = TT_REQUEST_ADMITTED;
        r->admitted_ns = sv->now_ns();
        sv->queued.pop_front();
        sv->active.push_back(r);
    }
    return TT_OK;
}

// ReSTIR buffers of request r for pass `pass`, its pixels at path slots
// [base, base + pixels): the same state a render of the same passes on a
// fresh target uses (current record and reservoirs alternate per pass,
// history from the second pass on).
tt::RestirBuffers restir_buffers(const Request& r, uint32_t pass, uint32_t base) {
    const size_t px = size_t(r.pixels());
    float4* g = static_cast<float4*>(r.restir.ptr);
    uint4* res = reinterpret_cast<uint4*>(g + 10 * px);
    tt::RestirBuffers rb{};
    rb.gbuffer[0] = g;
    rb.gbuffer[1] = g + 5 * px;
    rb.initial = res;
    rb.temporal = res + px;
    rb.final_[0] = res + 2 * px;
    rb.final_[1] = res + 3 * px;
    rb.cur = pass & 1u;
    rb.history_valid = pass > 0 ? 1u : 0u;
    rb.candidates = r.desc.restir_candidates ? r.desc.restir_candidates : 32u;
    rb.temporal_reuse = r.desc.restir_temporal ? 1u : 0u;
    rb.spatial_reuse = r.desc.restir_spatial ? 1u : 0u;
    rb.base = base;
    return rb;
}

void request_free(Request& r) {
    if (r.target.ptr) (void)hipFree(r.target.ptr);
    if (r.restir.ptr) (void)hipFree(r.restir.ptr);
    if (r.restir_done) (void)hipEventDestroy(r.restir_done);
    r.target = Block{};
    r.restir = Block{};
    r.restir_done = nullptr;
}

}  // namespace

extern "C" {

tt_status tt_server_create(tt_context* c, const tt_server_desc* desc, tt_server** out) try {
    if (!c || !out) return fail(c, TT_INVALID_ARGUMENT, "null argument");
    *out = nullptr;
    tt_server* sv = new (std::nothrow) tt_server();
    if (!sv) return fail(c, TT_OUT_OF_MEMORY, "server allocation");
    sv->context = c;
    tt::context_retain(c);
    sv->budget = desc && desc->hbm_budget_bytes ? desc->hbm_budget_bytes : (64ull << 30);
    sv->pool_paths = desc && desc->pool_paths ? desc->pool_paths : (4u << 20);
    // Batches in flight: as asked, else as many as the context has lanes
    // when their pools take at most half the budget, else one.
    const uint64_t pool = uint64_t(sv->pool_paths) * tt::PATH_BYTES;
    if (desc && desc->batches_in_flight) {
        sv->depth = std::min(desc->batches_in_flight, c->lane_count);
    } else {
        sv->depth = c->lane_count;
        while (sv->depth > 1 && 2 * pool * sv->depth > sv->budget) sv->depth--;
    }
    sv->origin = std::chrono::steady_clock::now();
    if (sv->pool_bytes() > sv->budget) {
        tt_server_destroy(sv);
        return fail(c, TT_OVER_BUDGET, "the path pool alone exceeds the HBM budget");
    }
    *out = sv;
    return TT_OK;
} catch (...) {
    return tt::exception_status(c ? &c->error : nullptr);
}

void tt_server_destroy(tt_server* sv) try {
    if (!sv) return;
    (void)hipSetDevice(sv->context->device);
    (void)hipStreamSynchronize(tt::api_stream(sv->context));  // also waits for batches in flight
    for (auto& kv : sv->requests) request_free(*kv.second);
    for (Block& b : sv->spare) (void)hipFree(b.ptr);
    tt_context* c = sv->context;
    delete sv;
    tt::context_release(c);
} catch (...) {
}

tt_status tt_server_add_scene(tt_server* sv, const tt_scene* s, uint32_t* slot) try {
    if (!sv || !s || !slot) return fail(sv ? sv->context : nullptr, TT_INVALID_ARGUMENT, "null argument");
    if (s->context != sv->context) return fail(sv->context, TT_STALE_HANDLE, "scene from another context");
    *slot = uint32_t(sv->scenes.size());
    sv->scenes.push_back(s);
    return TT_OK;
} catch (...) {
    return tt::exception_status(sv && sv->context ? &sv->context->error : nullptr);
}

tt_status tt_server_set_cache(tt_server* sv, uint32_t slot, tt_mlp* cache, const tt_cache_settings* settings) try {
    if (!sv) return TT_INVALID_ARGUMENT;
    tt_context* c = sv->context;
    if (slot >= sv->scenes.size()) return fail(c, TT_INVALID_ARGUMENT, "unknown scene slot");
    if (!cache) {
        sv->caches.erase(slot);
        return TT_OK;
    }
    if (tt::mlp_context(cache) != c) return fail(c, TT_STALE_HANDLE, "cache from another context");
    CacheSlot cs;
    cs.cache = cache;
    if (settings) cs.settings = *settings;
    tt::CachePlan plan;
    const char* why = nullptr;
    if (tt::cache_plan(&cs.settings, sv->pool_paths, plan, &why) != TT_OK) return fail(c, TT_INVALID_ARGUMENT, why);
    sv->caches[slot] = cs;
    return TT_OK;
} catch (...) {
    return tt::exception_status(sv && sv->context ? &sv->context->error : nullptr);
}

tt_status tt_server_submit(tt_server* sv, const tt_server_request* q, uint64_t* id) try {
    if (!sv || !q || !id) return fail(sv ? sv->context : nullptr, TT_INVALID_ARGUMENT, "null argument");
    tt_context* c = sv->context;
    const tt_constants& k = q->constants;
    if (q->scene >= sv->scenes.size()) return fail(c, TT_INVALID_ARGUMENT, "unknown scene slot");
    if (q->passes == 0 || k.screen_size[0] == 0 || k.screen_size[1] == 0)
        return fail(c, TT_INVALID_ARGUMENT, "empty request");
    if (k.bounce_count == 0 || k.bounce_count > 31) return fail(c, TT_INVALID_ARGUMENT, "bounce_count in [1, 31]");
    if (k.light_sampling != TT_LIGHT_SAMPLING_ON_HIT && k.light_sampling != TT_LIGHT_SAMPLING_NEXT_EVENT_ESTIMATION &&
        k.light_sampling != TT_LIGHT_SAMPLING_RESTIR)
        return fail(c, TT_UNSUPPORTED, "light sampling mode not supported");
    if (q->rng > 1) return fail(c, TT_INVALID_ARGUMENT, "unknown generator");
    if (q->restir_candidates > 1024) return fail(c, TT_INVALID_ARGUMENT, "restir_candidates > 1024");
    if (sv->scenes[q->scene]->view.texture_count > 0)
        return fail(c, TT_UNSUPPORTED, "textured scenes are not served (their level of detail needs each path's camera)");
    auto r = std::make_unique<Request>();
    r->desc = *q;
    r->width = k.screen_size[0];
    r->height = k.screen_size[1];
    if (r->pixels() >= (1ull << 30)) return fail(c, TT_INVALID_ARGUMENT, "request too large");
    if (r->is_restir() && r->pixels() > sv->pool_paths)
        return fail(c, TT_OVER_BUDGET, "a ReSTIR request renders whole passes, which must fit the path pool");
    if (r->resident_bytes() + sv->pool_bytes(r->is_restir()) > sv->budget)
        return fail(c, TT_OVER_BUDGET, "the request's target does not fit the HBM budget next to the path pool");
    if (r->is_restir()) sv->restir_slots = true;
    r->id = sv->next_id++;
    r->submitted_ns = sv->now_ns();
    sv->queued.push_back(r.get());
    *id = r->id;
    sv->requests[r->id] = std::move(r);
    return TT_OK;
} catch (...) {
    return tt::exception_status(sv && sv->context ? &sv->context->error : nullptr);
}

}  // extern "C"

namespace {

// Enqueues one batch over the pool on the server's next lane; false in
// `formed` when no request has work left to schedule.
tt_status enqueue_batch(tt_server* sv, bool& formed) {
    formed = false;
    tt_context* c = sv->context;
    // The oldest unfinished request sets the batch's configuration.
    Request* lead = nullptr;
    for (Request* r : sv->active)
        if (!r->scheduled_all()) {
            lead = r;
            break;
        }
    if (!lead) return TT_OK;
    auto t0 = std::chrono::steady_clock::now();
    std::vector<Request*> group;
    for (Request* r : sv->active)
        if (!r->scheduled_all() && same_config(*r, *lead)) group.push_back(r);

    const uint32_t lane = sv->next_slot;
    tt::PathBuffers b;
    tt_status st = tt::ensure_paths(c, lane, sv->pool_paths, b);
    if (st != TT_OK) return st;
    tt::Lane& L = c->lanes[lane];
    // The radiance cache of the group's scene: queues sized for a full pool
    // before any request's cursor moves.
    const tt_scene* scene = sv->scenes[lead->desc.scene];
    auto cache_it = sv->caches.find(lead->desc.scene);
    tt_mlp* cache = cache_it != sv->caches.end() ? cache_it->second.cache : nullptr;
    const tt_cache_settings* cache_settings = cache ? &cache_it->second.settings : nullptr;
    if (cache) {
        tt::CachePlan full;
        const char* why = nullptr;
        if (tt::cache_plan(cache_settings, sv->pool_paths, full, &why) != TT_OK) return fail(c, TT_INVALID_ARGUMENT, why);
        st = tt::ensure_cache(c, lane, sv->pool_paths, full.train_capacity, full.judge_capacity, b.cache);
        if (st != TT_OK) return st;
        st = tt::mlp_set_scene_bounds(cache, scene->bounds_lo, scene->bounds_hi);
        if (st != TT_OK) return st;
    }
    const bool restir = lead->is_restir();
    if (restir && L.restir_slot_capacity < sv->pool_paths) {
        if (L.restir_slot_memory) {
            TT_HIP(c, hipStreamSynchronize(tt::api_stream(c)));
            TT_HIP(c, hipFree(L.restir_slot_memory));
            L.restir_slot_memory = nullptr;
            L.restir_slot_capacity = 0;
        }
        TT_HIP(c, hipMalloc(&L.restir_slot_memory, size_t(sv->pool_paths) * RESTIR_SLOT_BYTES));
        L.restir_slot_capacity = sv->pool_paths;
    }

    std::vector<Chunk> chunks;
    uint32_t base = 0;
    uint32_t share = std::max<uint32_t>(sv->pool_paths / uint32_t(group.size()), 1u);
    for (Request* r : group) take(sv, r, uint32_t(std::min<uint64_t>(share, remaining_paths(r))), base, chunks);
    for (Request* r : group) take(sv, r, sv->pool_paths - base, base, chunks);
    const uint32_t paths = base;
    // This batch's shares: capacities no larger than the full pool's.
    tt::CachePlan plan{};
    if (cache) {
        const char* why = nullptr;
        (void)tt::cache_plan(cache_settings, paths, plan, &why);
    }

    const tt_constants& lk = lead->desc.constants;
    b.max_bounces = lk.bounce_count;
    auto chunk_args = [&](const Chunk& ch) {
        tt_constants k = ch.request->desc.constants;
        k.sample_count = 1;
        k.frame_index = ch.pass;
        k.accumulated_frame_count = ch.pass + 1;
        return tt::frame_args(k, ch.request->width, ch.request->height, ch.request->desc.rng, ch.request->desc.seed);
    };

    // As a render: after the API stream's work so far, on the lane's stream,
    // with shadow rays on its side stream when the context has side streams;
    // accumulation after the previous batch's or render's.
    const tt::Lane* other = c->prev_lane >= 0 && uint32_t(c->prev_lane) != lane ? &c->lanes[c->prev_lane] : nullptr;
    hipStream_t stream = L.stream;
    TT_HIP(c, hipEventRecord(c->api_mark, c->stream));
    c->api_dirty = false;
    TT_HIP(c, hipStreamWaitEvent(stream, c->api_mark, 0));
    hipStream_t side = c->side_streams ? L.side : stream;
    const bool shadows = lk.light_sampling != TT_LIGHT_SAMPLING_ON_HIT && scene->view.emissive_count > 0;
    // New targets start from zeros (pass 0 computes previous * 0 + new).
    for (const Chunk& ch : chunks) {
        Request* r = ch.request;
        if (r->cleared) continue;
        TT_HIP(c, tt::launch_zero(stream, static_cast<uint32_t*>(r->target.ptr), uint32_t(r->pixels() * 4)));
        r->cleared = true;
    }
    TT_HIP(c, tt::launch_zero(stream, L.counters, tt::C_WORDS));
    TT_HIP(c, tt::launch_zero_work(stream, L.work, tt::W_SETS));
    for (const Chunk& ch : chunks)
        TT_HIP(c, tt::launch_generate_chunk(stream, chunk_args(ch), b, ch.base, ch.pixel0, ch.count));
    tt::FrameArgs f = chunk_args(chunks.front());  // configuration fields are the group's
    if (cache) tt::apply_cache_plan(plan, f);
    TT_HIP(c, tt::launch_extend(stream, scene->view, b, 0, c->grid));
    if (restir && scene->view.emissive_count > 0) {
        // Each chunk is a whole pass: its ReSTIR passes over its request's
        // state, then its reservoirs and slot keys at its slots for shade 0.
        uint4* slot_final = static_cast<uint4*>(L.restir_slot_memory);
        uint4* slot_initial = slot_final + L.restir_slot_capacity;
        uint2* slot_key = reinterpret_cast<uint2*>(slot_initial + L.restir_slot_capacity);
        for (const Chunk& ch : chunks) {
            Request* r = ch.request;
            // The request's passes in earlier batches (other lanes) are done
            // with its state.
            if (r->restir_recorded) TT_HIP(c, hipStreamWaitEvent(stream, r->restir_done, 0));
            TT_HIP(c, tt::launch_zero(stream, L.counters + tt::C_RESTIR_VIS, 1));
            TT_HIP(c, tt::launch_zero_work(stream, L.work + tt::W_RESTIR * tt::WORK_SET, 1));
            tt::RestirBuffers rb = restir_buffers(*r, ch.pass, ch.base);
            TT_HIP(c, tt::launch_restir(stream, scene->view, chunk_args(ch), b, rb, c->grid));
            const uint4* initial = (rb.temporal_reuse || rb.spatial_reuse) ? rb.initial : rb.final_[rb.cur];
            TT_HIP(c, hipMemcpyAsync(slot_final + ch.base, rb.final_[rb.cur], size_t(ch.count) * sizeof(uint4),
                                     hipMemcpyDeviceToDevice, stream));
            TT_HIP(c, hipMemcpyAsync(slot_initial + ch.base, initial, size_t(ch.count) * sizeof(uint4),
                                     hipMemcpyDeviceToDevice, stream));
            TT_HIP(c, tt::launch_slot_keys(stream, slot_key + ch.base, 0u, ch.pass, ch.count));
            TT_HIP(c, hipEventRecord(r->restir_done, stream));
            r->restir_recorded = true;
        }
        b.restir_final = slot_final;
        b.restir_initial = slot_initial;
        b.restir_key = slot_key;
    }
    for (uint32_t bounce = 0; bounce < lk.bounce_count; bounce++) {
        if (shadows && bounce > 0 && side != stream) TT_HIP(c, hipStreamWaitEvent(stream, L.join, 0));
        TT_HIP(c, tt::launch_shade(stream, scene->view, f, b, bounce, paths));
        if (shadows) {
            if (side != stream) {
                TT_HIP(c, hipEventRecord(L.fork, stream));
                TT_HIP(c, hipStreamWaitEvent(side, L.fork, 0));
            }
            TT_HIP(c, tt::launch_shadow(side, scene->view, side_buffers_of(b, L), bounce, c->grid));
            if (side != stream) TT_HIP(c, hipEventRecord(L.join, side));
        }
        if (bounce + 1 < lk.bounce_count) TT_HIP(c, tt::launch_extend(stream, scene->view, b, bounce + 1, c->grid));
    }
    if (shadows && side != stream) TT_HIP(c, hipStreamWaitEvent(stream, L.join, 0));
    if (cache) {
        // As a cached render: the queries' radiance, and the cache against
        // the judge records, with the weights of the last training enqueued.
        st = tt::mlp_infer_device(cache, stream, b.cache.query_rec, b.cache.query_scale, L.counters + tt::C_CACHE_QUERY,
                                  paths, nullptr, b.radiance);
        if (st != TT_OK) return st;
        if (plan.judge_capacity) {
            st = tt::mlp_judge_device(cache, stream, b.cache.train_rec + 4 * size_t(b.cache.judge_base),
                                      b.cache.train_target + b.cache.judge_base, L.counters + tt::C_CACHE_JUDGE,
                                      plan.judge_capacity, L.judge_pred, L.judge_stats);
            if (st != TT_OK) return st;
        }
    }
    if (other && other->used) TT_HIP(c, hipStreamWaitEvent(stream, other->accumulated, 0));
    for (const Chunk& ch : chunks)
        TT_HIP(c, tt::launch_accumulate_chunk(stream, chunk_args(ch), b, ch.base, ch.pixel0, ch.count,
                                              static_cast<float4*>(ch.request->target.ptr)));
    TT_HIP(c, hipEventRecord(L.accumulated, stream));
    if (cache) {
        // Train on the batch's records; the lane's queues are free again once
        // training has read them, so the lane's end waits for it.
        st = tt::mlp_train_device(cache, stream, b.cache.train_rec, b.cache.train_target, L.counters + tt::C_CACHE_TRAIN,
                                  plan.train_capacity, plan.train_steps, true, nullptr, false,
                                  plan.ordered ? paths * plan.records : 0u);
        if (st != TT_OK) return st;
        st = tt::mlp_wait_trained(cache, stream);
        if (st != TT_OK) return st;
    }
    TT_HIP(c, hipEventRecord(L.done, stream));
    L.used = true;
    L.pending = true;
    c->prev_lane = int(lane);
    c->next_lane = (lane + 1u) % c->lane_count;
    sv->next_slot = (lane + 1u) % sv->depth;
    if (sv->inflight.empty()) sv->busy_since = t0;
    sv->inflight.push_back(Batch{lane, std::move(chunks), paths, cache != nullptr, plan.train_capacity,
                                 cache != nullptr && plan.judge_capacity > 0});
    formed = true;
    return TT_OK;
}

// Waits for the oldest batch in flight, then downloads the requests it
// finished and keeps their blocks for later requests.
tt_status finish_batch(tt_server* sv, uint32_t& done) {
    tt_context* c = sv->context;
    Batch batch = std::move(sv->inflight.front());
    sv->inflight.pop_front();
    TT_HIP(c, hipEventSynchronize(c->lanes[batch.lane].done));
    sv->stats.batches++;
    sv->stats.paths += batch.paths;
    if (batch.cached) {
        // The lane is not reused before its batch is finished here.
        const tt::Lane& L = c->lanes[batch.lane];
        uint32_t counts[3] = {0u, 0u, 0u};
        TT_HIP(c, hipMemcpy(counts, L.counters + tt::C_CACHE_QUERY, sizeof(counts), hipMemcpyDeviceToHost));
        sv->stats.cache_queries += counts[0];
        sv->stats.cache_samples += std::min(counts[1], batch.train_capacity);
        if (batch.judged) {
            double js[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
            TT_HIP(c, hipMemcpy(js, L.judge_stats, sizeof(js), hipMemcpyDeviceToHost));
            sv->stats.judge_samples += uint64_t(js[0]);
            sv->judge_cached_sum += js[1] / 3.0;
            sv->judge_traced_sum += js[2] / 3.0;
        }
    }
    if (sv->inflight.empty())  // the GPU idles from here until the next batch
        sv->stats.busy_ms +=
            std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - sv->busy_since).count();

    // Passes completed by this batch, then finished requests.
    for (const Chunk& ch : batch.chunks)
        if (ch.pixel0 + ch.count == ch.request->pixels()) ch.request->passes_done = ch.pass + 1;
    std::vector<Request*> still;
    for (Request* r : sv->active) {
        if (r->passes_done < r->desc.passes) {
            still.push_back(r);
            continue;
        }
        // Its last chunk was in this batch: no batch in flight touches it.
        r->image.resize(size_t(r->pixels()) * 4);
        TT_HIP(c, hipMemcpy(r->image.data(), r->target.ptr, r->target_bytes(), hipMemcpyDeviceToHost));
        release(sv, r->target);
        release(sv, r->restir);
        if (r->restir_done) {
            (void)hipEventDestroy(r->restir_done);
            r->restir_done = nullptr;
        }
        r->state = TT_REQUEST_DONE;
        r->done_ns = sv->now_ns();
        done++;
    }
    sv->active.swap(still);
    return TT_OK;
}

}  // namespace

extern "C" {

tt_status tt_server_step(tt_server* sv, uint32_t* completed) try {
    if (!sv) return TT_INVALID_ARGUMENT;
    tt_context* c = sv->context;
    if (completed) *completed = 0;
    TT_HIP(c, hipSetDevice(c->device));
    tt_status st = admit(sv);
    if (st != TT_OK) return st;
    bool formed = false;
    st = enqueue_batch(sv, formed);
    if (st != TT_OK) return st;
    // Keep up to `depth` batches in flight; with nothing new to enqueue,
    // finish the oldest.
    uint32_t done = 0;
    while (!sv->inflight.empty() && (sv->inflight.size() >= sv->depth || !formed)) {
        st = finish_batch(sv, done);
        if (st != TT_OK) return st;
        if (!formed) break;
    }
    if (completed) *completed = done;
    return TT_OK;
} catch (...) {
    return tt::exception_status(sv && sv->context ? &sv->context->error : nullptr);
}

tt_status tt_server_status(tt_server* sv, uint64_t id, tt_request_status* out) try {
    if (!sv || !out) return fail(sv ? sv->context : nullptr, TT_INVALID_ARGUMENT, "null argument");
    auto it = sv->requests.find(id);
    if (it == sv->requests.end()) return fail(sv->context, TT_STALE_HANDLE, "unknown request");
    const Request& r = *it->second;
    out->state = r.state;
    out->passes_done = r.passes_done;
    out->submitted_ns = r.submitted_ns;
    out->admitted_ns = r.admitted_ns;
    out->done_ns = r.done_ns;
    return TT_OK;
} catch (...) {
    return tt::exception_status(sv && sv->context ? &sv->context->error : nullptr);
}

tt_status tt_server_take_result(tt_server* sv, uint64_t id, float* rgba, size_t bytes) try {
    if (!sv || !rgba) return fail(sv ? sv->context : nullptr, TT_INVALID_ARGUMENT, "null argument");
    auto it = sv->requests.find(id);
    if (it == sv->requests.end()) return fail(sv->context, TT_STALE_HANDLE, "unknown request");
    Request& r = *it->second;
    if (r.state != TT_REQUEST_DONE) return fail(sv->context, TT_INVALID_ARGUMENT, "request not done");
    if (bytes < r.target_bytes()) return fail(sv->context, TT_INVALID_ARGUMENT, "result buffer too small");
    std::memcpy(rgba, r.image.data(), r.target_bytes());
    sv->requests.erase(it);
    return TT_OK;
} catch (...) {
    return tt::exception_status(sv && sv->context ? &sv->context->error : nullptr);
}

tt_status tt_server_stats_get(tt_server* sv, tt_server_stats* out) try {
    if (!sv || !out) return fail(sv ? sv->context : nullptr, TT_INVALID_ARGUMENT, "null argument");
    *out = sv->stats;
    out->queued = sv->queued.size();
    out->active = sv->active.size();
    uint64_t done = 0;
    for (auto& kv : sv->requests)
        if (kv.second->state == TT_REQUEST_DONE) done++;
    out->done = done;
    out->budget_used_bytes = sv->used + sv->spare_bytes + sv->pool_bytes();
    if (sv->stats.judge_samples > 0) {
        out->judge_traced = sv->judge_traced_sum / double(sv->stats.judge_samples);
        out->judge_cached = sv->judge_cached_sum / double(sv->stats.judge_samples);
    }
    return TT_OK;
} catch (...) {
    return tt::exception_status(sv && sv->context ? &sv->context->error : nullptr);
}

}  // extern "C"
