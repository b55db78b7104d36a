// Batched frames captured in one HIP graph, and the launch-shape autotuner.
//
// A batch owns `lanes` lanes: each lane is a context of its own (scratch arena)
// with its own stream.  Frames are spread over the lanes largest first onto the
// least-loaded lane, so up to `lanes` frames are in flight at once, small frames
// fill the GPU together, and the lanes finish close together.  The first run
// of a frame list goes eagerly (it grows the lane arenas to the frames' sizes);
// it is then captured once into a graph — fork from the batch's origin stream
// to the lanes, the frames in order on their lanes, join back — and later runs
// of the same list are one hipGraphLaunch.  Every stage of a frame sizes its
// launches from device counters (no host synchronisation inside a frame), which
// is what makes the capture possible.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "common.hpp"

struct fsr_batch {
    fsr_context* parent = nullptr;
    std::vector<fsr_context*> lanes;
    std::vector<hipStream_t> streams;
    std::vector<hipEvent_t> join;
    hipEvent_t fork = nullptr, in_ready = nullptr, out_ready = nullptr;
    hipStream_t origin = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    // key of the captured list
    std::vector<fsr_frame_desc> key;
    int key_method = -1;
    fsr_aniso_param key_aniso{};
    bool key_has_aniso = false;
    uint64_t* key_ts = nullptr;
};

using namespace fsr;

namespace {

constexpr int kLaneFieldWorkgroupsPerCu = 2;

void release_graph(fsr_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    b->key.clear();
}

// Lane of every frame: largest first, each onto the lane with the least work so
// far (longest-processing-time rule), with N + V/16 as the cost of a frame
// (particle gathers plus the dense-grid passes).  Returns frames in enqueue
// order with their lanes.
std::vector<std::pair<uint32_t, uint32_t>> schedule(const fsr_frame_desc* frames, uint32_t n,
                                                    uint32_t L) {
    std::vector<uint32_t> order(n);
    std::vector<double> cost(n);
    for (uint32_t f = 0; f < n; ++f) {
        order[f] = f;
        const fsr_grid_info& g = frames[f].sc_grid;
        cost[f] = (double)frames[f].num_particles +
                  (double)g.resolution.x * g.resolution.y * g.resolution.z / 16.0;
    }
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t c) { return cost[a] > cost[c]; });
    std::vector<double> load(L, 0.0);
    std::vector<std::pair<uint32_t, uint32_t>> out;
    out.reserve(n);
    for (uint32_t f : order) {
        const uint32_t l = (uint32_t)(std::min_element(load.begin(), load.end()) - load.begin());
        load[l] += cost[f];
        out.emplace_back(f, l);
    }
    return out;
}

fsr_status enqueue(fsr_batch* b, fsr_method method, const fsr_aniso_param* aniso,
                   const fsr_frame_desc* frames, uint32_t n, uint64_t* ts) {
    const uint32_t L = (uint32_t)b->lanes.size();
    FSR_HIP(hipEventRecord(b->fork, b->origin));
    for (uint32_t l = 0; l < L; ++l) FSR_HIP(hipStreamWaitEvent(b->streams[l], b->fork, 0));
    for (const auto& fl : schedule(frames, n, L)) {
        const uint32_t f = fl.first, l = fl.second;
        const fsr_frame_desc& d = frames[f];
        hipStream_t s = b->streams[l];
        if (ts) FSR_TRY(device_timestamp(ts + 2ull * f, s));
        if (d.src_pos && d.num_particles) {
            FSR_HIP(hipMemcpyAsync(d.pos, d.src_pos, 12ull * d.num_particles,
                                   hipMemcpyDeviceToDevice, s));
            FSR_HIP(hipMemcpyAsync(d.dens, d.src_dens, 4ull * d.num_particles,
                                   hipMemcpyDeviceToDevice, s));
        }
        FSR_TRY(fsr_reconstruct_frame(b->lanes[l], method, &d.sim, aniso, &d.sp_grid, &d.sc_grid,
                                      d.pos, d.dens, d.num_particles, d.phi, d.tri_pos, d.tri_nrm,
                                      d.tri_capacity, d.stats, s));
        if (ts) FSR_TRY(device_timestamp(ts + 2ull * f + 1, s));
    }
    for (uint32_t l = 0; l < L; ++l) {
        FSR_HIP(hipEventRecord(b->join[l], b->streams[l]));
        FSR_HIP(hipStreamWaitEvent(b->origin, b->join[l], 0));
    }
    return FSR_OK;
}

bool same_list(const fsr_batch* b, fsr_method method, const fsr_aniso_param* aniso,
               const fsr_frame_desc* frames, uint32_t n, uint64_t* ts) {
    if (!b->exec || b->key.size() != n || b->key_method != (int)method || b->key_ts != ts)
        return false;
    if ((aniso != nullptr) != b->key_has_aniso) return false;
    if (aniso && std::memcmp(aniso, &b->key_aniso, sizeof(fsr_aniso_param)) != 0) return false;
    return std::memcmp(frames, b->key.data(), sizeof(fsr_frame_desc) * n) == 0;
}

}  // namespace

extern "C" {

fsr_status fsr_batch_create(fsr_context* ctx, uint32_t lanes, fsr_batch** out) {
    if (!ctx || !out || lanes == 0 || lanes > 64) {
        set_error("fsr_batch_create: need a context, 1..64 lanes and an output pointer");
        return FSR_ERR_INVALID_ARG;
    }
    *out = nullptr;
    FSR_HIP(hipSetDevice(ctx->device));
    fsr_batch* b = new fsr_batch();
    b->parent = ctx;
    auto fail = [&](fsr_status st) {
        fsr_batch_destroy(b);
        return st;
    };
    for (uint32_t l = 0; l < lanes; ++l) {
        fsr_context* c = nullptr;
        fsr_status st = fsr_create(&c, ctx->device);
        if (st != FSR_OK) return fail(st);
        std::copy(ctx->tune, ctx->tune + TUNE_NUM, c->tune);
        // Lanes share the GPU: unless a launch shape was chosen, a lane's
        // persistent field kernel takes 2 workgroups per CU (of the 6 a lone
        // frame uses), so the field kernels of concurrent lanes run side by
        // side instead of queueing behind each other's workgroups
        // (fsr_batch_autotune picks another value when it is faster).
        if (c->tune[TUNE_FIELD] == 0) c->tune[TUNE_FIELD] = kLaneFieldWorkgroupsPerCu;
        b->lanes.push_back(c);
        hipStream_t s;
        hipError_t e = hipStreamCreateWithFlags(&s, hipStreamNonBlocking);
        if (e != hipSuccess) return fail(hip_fail(e, "hipStreamCreate", __FILE__, __LINE__));
        b->streams.push_back(s);
        hipEvent_t ev;
        e = hipEventCreateWithFlags(&ev, hipEventDisableTiming);
        if (e != hipSuccess) return fail(hip_fail(e, "hipEventCreate", __FILE__, __LINE__));
        b->join.push_back(ev);
    }
    hipError_t e = hipStreamCreateWithFlags(&b->origin, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->fork, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->in_ready, hipEventDisableTiming);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&b->out_ready, hipEventDisableTiming);
    if (e != hipSuccess) return fail(hip_fail(e, "batch stream/event", __FILE__, __LINE__));
    *out = b;
    return FSR_OK;
}

fsr_status fsr_batch_destroy(fsr_batch* b) {
    if (!b) return FSR_OK;
    if (b->parent) (void)hipSetDevice(b->parent->device);
    (void)hipDeviceSynchronize();
    release_graph(b);
    for (hipStream_t s : b->streams) (void)hipStreamDestroy(s);
    for (hipEvent_t ev : b->join) (void)hipEventDestroy(ev);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->in_ready) (void)hipEventDestroy(b->in_ready);
    if (b->out_ready) (void)hipEventDestroy(b->out_ready);
    if (b->origin) (void)hipStreamDestroy(b->origin);
    for (fsr_context* c : b->lanes) fsr_destroy(c);
    delete b;
    return FSR_OK;
}

fsr_status fsr_batch_run(fsr_batch* b, fsr_method method, const fsr_aniso_param* aniso,
                         const fsr_frame_desc* frames, uint32_t n, uint64_t* ts, void* stream) {
    if (!b || (n && !frames)) {
        set_error("fsr_batch_run: null argument");
        return FSR_ERR_INVALID_ARG;
    }
    FSR_HIP(hipSetDevice(b->parent->device));
    hipStream_t caller = (hipStream_t)stream;
    // order the batch after the caller's prior work, and the caller after the batch
    FSR_HIP(hipEventRecord(b->in_ready, caller));
    FSR_HIP(hipStreamWaitEvent(b->origin, b->in_ready, 0));
    if (same_list(b, method, aniso, frames, n, ts)) {
        FSR_HIP(hipGraphLaunch(b->exec, b->origin));
    } else {
        release_graph(b);
        // eager run: grows the lane arenas and produces this call's results
        FSR_TRY(enqueue(b, method, aniso, frames, n, ts));
        FSR_HIP(hipStreamSynchronize(b->origin));
        for (fsr_context* c : b->lanes) c->capturing = true;
        hipError_t e = hipStreamBeginCapture(b->origin, hipStreamCaptureModeRelaxed);
        fsr_status st = e == hipSuccess ? FSR_OK : hip_fail(e, "hipStreamBeginCapture", __FILE__, __LINE__);
        if (st == FSR_OK) st = enqueue(b, method, aniso, frames, n, ts);
        hipGraph_t g = nullptr;
        hipError_t e2 = hipStreamEndCapture(b->origin, &g);
        for (fsr_context* c : b->lanes) c->capturing = false;
        if (st != FSR_OK) {
            if (g) (void)hipGraphDestroy(g);
            return st;
        }
        if (e2 != hipSuccess) return hip_fail(e2, "hipStreamEndCapture", __FILE__, __LINE__);
        b->graph = g;
        FSR_HIP(hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
        b->key.assign(frames, frames + n);
        b->key_method = (int)method;
        b->key_has_aniso = aniso != nullptr;
        if (aniso) b->key_aniso = *aniso;
        b->key_ts = ts;
    }
    FSR_HIP(hipEventRecord(b->out_ready, b->origin));
    FSR_HIP(hipStreamWaitEvent(caller, b->out_ready, 0));
    return FSR_OK;
}

fsr_status fsr_batch_autotune(fsr_batch* b, fsr_method method, const fsr_aniso_param* aniso,
                              const fsr_frame_desc* frames, uint32_t n, uint32_t repeats,
                              uint32_t* choice, void* stream) {
    if (!b || !frames || n == 0 || repeats == 0 || repeats > 64) {
        set_error("fsr_batch_autotune: need a batch, frames and 1..64 repeats");
        return FSR_ERR_INVALID_ARG;
    }
    FSR_HIP(hipSetDevice(b->parent->device));
    hipStream_t caller = (hipStream_t)stream;
    release_graph(b);
    FSR_HIP(hipEventRecord(b->in_ready, caller));
    FSR_HIP(hipStreamWaitEvent(b->origin, b->in_ready, 0));
    // one eager run sizes the lane arenas
    FSR_TRY(enqueue(b, method, aniso, frames, n, nullptr));
    FSR_HIP(hipStreamSynchronize(b->origin));
    hipEvent_t e0, e1;
    FSR_HIP(hipEventCreate(&e0));
    hipError_t he = hipEventCreate(&e1);
    if (he != hipSuccess) {
        (void)hipEventDestroy(e0);
        return hip_fail(he, "hipEventCreate", __FILE__, __LINE__);
    }
    // workgroups per CU of a lane's field kernel (0: the kernel's resident
    // count); the whole batch, eager, median of `repeats` passes
    static const int kCand[] = {2, 1, 3, 4, 0};
    int best = kLaneFieldWorkgroupsPerCu;
    float best_t = 0.f;
    fsr_status st = FSR_OK;
    for (int cand : kCand) {
        for (fsr_context* c : b->lanes) c->tune[TUNE_FIELD] = cand;
        std::vector<float> t;
        for (uint32_t r = 0; r < repeats && st == FSR_OK; ++r) {
            he = hipEventRecord(e0, b->origin);
            if (he == hipSuccess) st = enqueue(b, method, aniso, frames, n, nullptr);
            if (he == hipSuccess && st == FSR_OK) he = hipEventRecord(e1, b->origin);
            if (he == hipSuccess && st == FSR_OK) he = hipEventSynchronize(e1);
            float ms = 0.f;
            if (he == hipSuccess && st == FSR_OK) he = hipEventElapsedTime(&ms, e0, e1);
            if (he != hipSuccess && st == FSR_OK) st = hip_fail(he, "batch autotune timing", __FILE__, __LINE__);
            t.push_back(ms);
        }
        if (st != FSR_OK) break;
        std::sort(t.begin(), t.end());
        const float med = t[t.size() / 2];
        if (cand == kCand[0] || med < 0.99f * best_t) {
            best_t = med;
            best = cand;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    for (fsr_context* c : b->lanes) c->tune[TUNE_FIELD] = best;
    if (choice) *choice = (uint32_t)best;
    FSR_HIP(hipEventRecord(b->out_ready, b->origin));
    FSR_HIP(hipStreamWaitEvent(caller, b->out_ready, 0));
    return st;
}

fsr_status fsr_set_launch_shape(fsr_context* ctx, const uint32_t* wg) {
    if (!ctx || !wg) {
        set_error("fsr_set_launch_shape: null argument");
        return FSR_ERR_INVALID_ARG;
    }
    for (int c = 0; c < TUNE_NUM; ++c) ctx->tune[c] = (int)std::min<uint32_t>(wg[c], 64u);
    return FSR_OK;
}

fsr_status fsr_autotune(fsr_context* ctx, fsr_method method, const fsr_aniso_param* aniso,
                        const fsr_frame_desc* d, uint32_t repeats, uint32_t* choice,
                        void* stream) {
    if (!ctx || !d || repeats == 0 || repeats > 64) {
        set_error("fsr_autotune: need a context, a frame and 1..64 repeats");
        return FSR_ERR_INVALID_ARG;
    }
    FSR_HIP(hipSetDevice(ctx->device));
    hipStream_t s = (hipStream_t)stream;
    // 0 = each kernel's resident count (the default); the others limit it
    static const int kCand[] = {0, 2, 3, 4};
    // stage whose time decides each class
    const int stage_of[TUNE_NUM] = {FSR_STAGE_FIELD, FSR_STAGE_MC, FSR_STAGE_PREP, FSR_STAGE_BAND};
    const bool used[TUNE_NUM] = {true, true, method == FSR_METHOD_YU13, true};
    std::vector<float> ms((size_t)repeats * FSR_NUM_STAGES);
    auto measure = [&](int stage, float* out) -> fsr_status {
        FSR_TRY(fsr_set_profiling(ctx, repeats));
        for (uint32_t r = 0; r < repeats; ++r) {
            if (d->src_pos && d->num_particles) {
                FSR_HIP(hipMemcpyAsync(d->pos, d->src_pos, 12ull * d->num_particles,
                                       hipMemcpyDeviceToDevice, s));
                FSR_HIP(hipMemcpyAsync(d->dens, d->src_dens, 4ull * d->num_particles,
                                       hipMemcpyDeviceToDevice, s));
            }
            FSR_TRY(fsr_reconstruct_frame(ctx, method, &d->sim, aniso, &d->sp_grid, &d->sc_grid,
                                          d->pos, d->dens, d->num_particles, d->phi, d->tri_pos,
                                          d->tri_nrm, d->tri_capacity, d->stats, s));
        }
        uint32_t got = 0;
        FSR_TRY(fsr_stage_times(ctx, ms.data(), repeats, &got));
        FSR_TRY(fsr_set_profiling(ctx, 0));
        std::vector<float> v;
        for (uint32_t r = 0; r < got; ++r) v.push_back(ms[(size_t)r * FSR_NUM_STAGES + stage]);
        std::sort(v.begin(), v.end());
        *out = v.empty() ? 0.f : v[v.size() / 2];
        return FSR_OK;
    };
    // one warm run sizes the arena
    float dummy;
    FSR_TRY(measure(FSR_STAGE_FIELD, &dummy));
    for (int c = 0; c < TUNE_NUM; ++c) {
        if (!used[c]) continue;
        // the default is timed first and kept unless a candidate beats it by
        // more than 1% (stage medians of `repeats` frames)
        int best = 0;
        float best_t = 0.f;
        for (int cand : kCand) {
            ctx->tune[c] = cand;
            float t;
            FSR_TRY(measure(stage_of[c], &t));
            if (cand == 0 || t < 0.99f * best_t) {
                best_t = t;
                best = cand;
            }
        }
        ctx->tune[c] = best;
    }
    if (choice)
        for (int c = 0; c < TUNE_NUM; ++c) choice[c] = (uint32_t)ctx->tune[c];
    return FSR_OK;
}

}  // extern "C"
