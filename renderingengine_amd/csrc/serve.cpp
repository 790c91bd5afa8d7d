// Request loop: jobs arrive at any time, are admitted in arrival order while
// their HBM footprint fits the budget next to the resident scenes, and run
// overlapped on the device.
//
//   frames     (BSDF-only / NEE / MIS) go to the batch engine, whose steps
//              run on the server's engine stream: their paths share one
//              launch sequence per step with every other admitted frame
//   sequences  (ReSTIR DI, frames rendered back to back with reservoir and
//              Film history kept on the device) take a renderer slot; each
//              slot records on a stream of its own, so sequences overlap the
//              engine's steps and each other
//
// Footprints (bytes charged to the budget from admission to release):
//   frame      Film (16 B per pixel) + per-sample values (16 B per path)
//   sequence   the renderer's per-pixel buffers with ReSTIR reservoirs and
//              surfaces, times the job's pixels, + its copied-out Film
// Resident scenes (hierarchies, triangles, the engine arena's copies) are
// charged once at add_scene.  The engine's path pool and the sequence
// renderers are allocated at creation, outside the budget.
//
// Nothing in the loop allocates device memory: job outputs come from arenas
// carved at creation, the renderers' ReSTIR buffers and traversal workspaces
// are sized before the first job.  A finished job is noticed by the first
// pump that queries its event; that host time is its finish time.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "arena.hpp"
#include "rex_internal.hpp"

namespace rex {
namespace {

constexpr uint32_t kMaxSequenceFrames = 64;
constexpr uint32_t kCountCopyWords = 48;  // rays [0, 16), shadows [32, 48)

int64_t now_ns() {
    return std::chrono::duration_cast<std::chrono::nanoseconds>(std::chrono::steady_clock::now().time_since_epoch())
        .count();
}

struct ServerJob {
    uint32_t kind = REX_JOB_FRAME;
    uint32_t scene = 0;
    uint32_t frames = 1;
    rex_frame_desc frame{};
    std::vector<rex_camera> cameras;
    rex_restir_desc restir{};
    bool ownRestir = false;
    uint32_t state = 0;  // 0 queued, 1 admitted, 2 finished
    uint64_t footprint = 0;
    int64_t submitNs = 0, admitNs = 0, finishNs = 0;
    uint64_t batchJob = 0;
    int slot = -1;
    void* film = nullptr;
    uint64_t rays = 0;
};

struct Slot {
    rex_renderer renderer = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    uint32_t* counts = nullptr;  // pinned: kCountCopyWords per frame
    uint64_t job = 0;
};

}  // namespace

struct Server {
    rex_server_desc desc{};
    uint64_t budget = 0;
    rex_batch batch = nullptr;
    hipStream_t engineStream = nullptr;
    std::vector<rex_scene> scenes;
    uint32_t maxDepth = 0;
    uint64_t residentBytes = 0;
    std::vector<Slot> slots;
    bool ready = false;
    DeviceArena sequenceOut;
    std::map<uint64_t, ServerJob> jobs;
    std::deque<uint64_t> frameQueue, sequenceQueue;
    std::vector<uint64_t> runningFrames;
    uint64_t nextId = 1, submitted = 0, admitted = 0, finished = 0, released = 0;
    uint64_t admittedBytes = 0, peakAdmittedBytes = 0, sequenceRays = 0, sequenceFrames = 0;
    size_t engineArenaBytes = 0;

    rex_status prepare();
    rex_status start_sequence(Slot& sl, ServerJob& j);
    rex_status admit();
    rex_status pump(uint32_t* finishedOut);
    void release_all();
    uint64_t footprint(const ServerJob& j) const;
    bool busy() const {
        if (!runningFrames.empty() || batch_busy(batch)) return true;
        for (const Slot& sl : slots) {
            if (sl.job != 0) return true;
        }
        return false;
    }
};

uint64_t Server::footprint(const ServerJob& j) const {
    const uint64_t pixels = static_cast<uint64_t>(j.frame.width) * j.frame.height;
    if (j.kind == REX_JOB_FRAME) return RangeAllocator::round(pixels * 16u) + pixels * j.frame.spp * 16u;
    return pixels * renderer_bytes_per_pixel(true) + RangeAllocator::round(pixels * 16u);
}

// The renderer slots, created at the first submit once every scene is known.
rex_status Server::prepare() {
    if (ready) return REX_SUCCESS;
    if (scenes.empty()) return fail(REX_MISSING_SCENE, "no resident scene");
    slots.resize(desc.sequenceSlots);
    for (Slot& sl : slots) {
        if (rex_status st = rex_renderer_create(scenes[0], desc.sequenceMaxPixels, 1u, &sl.renderer)) return st;
        if (rex_status st = renderer_prepare(sl.renderer, maxDepth, true)) return st;
        if (rex_status st = check_hip(hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking), "slot stream"))
            return st;
        if (rex_status st = check_hip(hipEventCreateWithFlags(&sl.done, hipEventDisableTiming), "slot event"))
            return st;
        if (rex_status st = check_hip(hipHostMalloc(reinterpret_cast<void**>(&sl.counts),
                                                    kMaxSequenceFrames * kCountCopyWords * sizeof(uint32_t)),
                                      "slot counts")) {
            return st;
        }
    }
    if (rex_status st = check_hip(hipDeviceSynchronize(), "server prepare")) return st;
    ready = true;
    return REX_SUCCESS;
}

// Records the whole sequence on the slot's stream: frames back to back (each
// followed by a copy of its queue sizes), the last Film copied out, the
// completion event.
rex_status Server::start_sequence(Slot& sl, ServerJob& j) {
    hipStream_t st = sl.stream;
    if (rex_status s = rex_renderer_bind_scene(sl.renderer, scenes[j.scene], st)) return s;
    if (j.ownRestir) {
        if (rex_status s = rex_renderer_set_restir(sl.renderer, &j.restir)) return s;
    } else {
        rex_restir_desc d{};
        d.candidates = 32;
        d.neighbors = 5;
        d.maxM = 20;
        d.maxAge = 20;
        d.temporal = 1;
        d.neighborMode = REX_RESTIR_NEIGHBORS_RING;
        d.radius = 30;
        if (rex_status s = rex_renderer_set_restir(sl.renderer, &d)) return s;
    }
    const uint32_t* counts = renderer_counts(sl.renderer);
    for (uint32_t k = 0; k < j.frames; ++k) {
        rex_frame_desc fd = j.frame;
        fd.camera = j.cameras[k];
        fd.sampleIndex = j.frame.sampleIndex + k;
        if (rex_status s = rex_render_frame(sl.renderer, &fd, 0, st)) return s;
        if (rex_status s = check_hip(hipMemcpyAsync(sl.counts + k * kCountCopyWords, counts,
                                                    kCountCopyWords * sizeof(uint32_t), hipMemcpyDeviceToHost, st),
                                     "sequence counts")) {
            return s;
        }
    }
    const size_t filmBytes = static_cast<size_t>(j.frame.width) * j.frame.height * sizeof(float4);
    if (rex_status s = check_hip(hipMemcpyAsync(j.film, renderer_film(sl.renderer), filmBytes,
                                                hipMemcpyDeviceToDevice, st),
                                 "sequence film")) {
        return s;
    }
    if (rex_status s = check_hip(hipEventRecord(sl.done, st), "sequence event")) return s;
    sequenceFrames += j.frames;
    return REX_SUCCESS;
}

// Arrival order across both queues; a sequence waits for a free slot without
// holding up frames behind it, and everything waits while the earliest
// admissible job does not fit the budget.
rex_status Server::admit() {
    for (;;) {
        int freeSlot = -1;
        for (size_t i = 0; i < slots.size(); ++i) {
            if (slots[i].job == 0) {
                freeSlot = static_cast<int>(i);
                break;
            }
        }
        const uint64_t fid = frameQueue.empty() ? 0 : frameQueue.front();
        const uint64_t sid = (sequenceQueue.empty() || freeSlot < 0) ? 0 : sequenceQueue.front();
        uint64_t id = fid;
        if (sid != 0 && (id == 0 || jobs.at(sid).submitNs < jobs.at(id).submitNs)) id = sid;
        if (id == 0) return REX_SUCCESS;
        ServerJob& j = jobs.at(id);
        if (residentBytes + admittedBytes + j.footprint > budget) return REX_SUCCESS;
        if (j.kind == REX_JOB_FRAME) {
            if (rex_status st = rex_batch_submit(batch, j.scene, &j.frame, &j.batchJob)) return st;
            frameQueue.pop_front();
            runningFrames.push_back(id);
        } else {
            j.film = sequenceOut.alloc(static_cast<size_t>(j.frame.width) * j.frame.height * sizeof(float4));
            if (j.film == nullptr) return REX_SUCCESS;  // earlier sequences' Films not yet released
            sequenceQueue.pop_front();
            Slot& sl = slots[freeSlot];
            sl.job = id;
            j.slot = freeSlot;
            if (rex_status st = start_sequence(sl, j)) return st;
        }
        j.state = 1;
        j.admitNs = now_ns();
        admittedBytes += j.footprint;
        peakAdmittedBytes = std::max(peakAdmittedBytes, admittedBytes);
        ++admitted;
    }
}

rex_status Server::pump(uint32_t* finishedOut) {
    uint32_t nf = 0;
    if (finishedOut) *finishedOut = 0;
    if (!ready) return REX_SUCCESS;
    // Sequences.
    for (Slot& sl : slots) {
        if (sl.job == 0) continue;
        const hipError_t q = hipEventQuery(sl.done);
        if (q == hipErrorNotReady) continue;
        if (rex_status st = check_hip(q, "sequence event")) return st;
        ServerJob& j = jobs.at(sl.job);
        const uint64_t pixels = static_cast<uint64_t>(j.frame.width) * j.frame.height;
        uint64_t rays = 0;
        for (uint32_t k = 0; k < j.frames; ++k) {
            const uint32_t* c = sl.counts + k * kCountCopyWords;
            for (uint32_t b = 0; b < 16u; ++b) rays += static_cast<uint64_t>(c[b]) + c[32u + b];
            rays += pixels;  // the frame's winner-visibility any-hit batch
        }
        j.rays = rays;
        sequenceRays += rays;
        j.state = 2;
        j.finishNs = now_ns();
        j.slot = -1;
        sl.job = 0;
        ++nf;
        ++finished;
    }
    // Frames.
    for (size_t i = 0; i < runningFrames.size();) {
        ServerJob& j = jobs.at(runningFrames[i]);
        rex_batch_job_info info;
        if (rex_status st = rex_batch_job(batch, j.batchJob, &info)) return st;
        if (info.state == 3u) {
            j.state = 2;
            j.finishNs = now_ns();
            j.film = info.film;
            runningFrames[i] = runningFrames.back();
            runningFrames.pop_back();
            ++nf;
            ++finished;
        } else {
            ++i;
        }
    }
    if (rex_status st = admit()) return st;
    if (batch_busy(batch)) {
        if (rex_status st = rex_batch_step(batch, engineStream)) return st;
    }
    if (finishedOut) *finishedOut = nf;
    return REX_SUCCESS;
}

void Server::release_all() {
    (void)hipDeviceSynchronize();
    for (Slot& sl : slots) {
        if (sl.renderer) (void)rex_renderer_destroy(sl.renderer);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.counts) (void)hipHostFree(sl.counts);
        sl = Slot{};
    }
    slots.clear();
    if (batch) (void)rex_batch_destroy(batch);
    batch = nullptr;
    if (engineStream) (void)hipStreamDestroy(engineStream);
    engineStream = nullptr;
    sequenceOut.destroy();
    jobs.clear();
}

}  // namespace rex

struct rex_server_t {
    rex::Server s;
};

using rex::check_hip;
using rex::fail;

extern "C" {

rex_status rex_server_create(const rex_server_desc* desc, rex_server* out) {
    if (out == nullptr) return fail(REX_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    rex_server_desc d{};
    if (desc != nullptr) d = *desc;
    if (d.sequenceSlots == 0u) d.sequenceSlots = 4u;
    if (d.sequenceSlots > 16u) return fail(REX_INVALID_ARGUMENT, "at most 16 sequence slots");
    if (d.sequenceMaxPixels == 0u) d.sequenceMaxPixels = 1920u * 1080u;
    if (d.sequenceOutputMiB == 0u) d.sequenceOutputMiB = 2048u;
    rex_server rs = new (std::nothrow) rex_server_t();
    if (rs == nullptr) return fail(REX_INVALID_ARGUMENT, "host allocation failed");
    rex::Server& s = rs->s;
    s.desc = d;
    s.budget = d.budgetBytes != 0u ? d.budgetBytes : (200ull << 30);
    rex_batch_desc bd{};
    bd.pathCapacity = d.pathCapacity;
    bd.maxJobs = d.maxJobs;
    bd.useGraphs = d.useGraphs;
    bd.shadeBlocksPerCu = d.shadeBlocksPerCu;
    bd.outputMiB = d.outputMiB;
    s.engineArenaBytes = static_cast<size_t>(d.outputMiB != 0u ? d.outputMiB : 16384u) << 20;
    rex_status st = rex_batch_create(&bd, &s.batch);
    if (st == REX_SUCCESS) {
        st = check_hip(hipStreamCreateWithFlags(&s.engineStream, hipStreamNonBlocking), "engine stream");
    }
    if (st == REX_SUCCESS) {
        st = check_hip(s.sequenceOut.create(static_cast<size_t>(d.sequenceOutputMiB) << 20), "sequence arena");
    }
    if (st != REX_SUCCESS) {
        s.release_all();
        delete rs;
        return st;
    }
    *out = rs;
    return REX_SUCCESS;
}

rex_status rex_server_destroy(rex_server server) {
    if (server == nullptr) return REX_SUCCESS;
    server->s.release_all();
    delete server;
    return REX_SUCCESS;
}

rex_status rex_server_add_scene(rex_server server, rex_scene scene, uint32_t* sceneIndex) {
    if (server == nullptr || sceneIndex == nullptr) return fail(REX_INVALID_ARGUMENT, "NULL argument");
    if (scene == nullptr) return fail(REX_MISSING_SCENE, "scene is NULL");
    rex::Server& s = server->s;
    if (s.ready) return fail(REX_INVALID_ARGUMENT, "scenes are added before the first submit");
    uint32_t idx = 0;
    if (rex_status st = rex_batch_add_scene(s.batch, scene, &idx)) return st;
    const rex::Scene& sc = scene->s;
    const uint64_t arenaCopy = static_cast<uint64_t>(sc.wide4qCount) * sizeof(rex::Wide4Q) +
                               static_cast<uint64_t>(sc.triCount) *
                                   (sizeof(rex::TriPos) + sizeof(rex::TriAttr) + sizeof(uint32_t));
    s.residentBytes += sc.deviceBytes + arenaCopy;
    s.maxDepth = std::max(s.maxDepth, sc.view.wide4Depth);
    s.scenes.push_back(scene);
    *sceneIndex = idx;
    return REX_SUCCESS;
}

rex_status rex_server_submit(rex_server server, const rex_server_job_desc* job, uint64_t* jobId) {
    if (server == nullptr || job == nullptr || jobId == nullptr) return fail(REX_INVALID_ARGUMENT, "NULL argument");
    rex::Server& s = server->s;
    if (job->sceneIndex >= s.scenes.size()) return fail(REX_MISSING_SCENE, "scene index not resident");
    const rex_frame_desc& f = job->frame;
    const uint64_t pixels = static_cast<uint64_t>(f.width) * f.height;
    if (pixels == 0 || f.spp == 0) return fail(REX_INVALID_ARGUMENT, "empty frame");
    if (f.width > 16384u || f.height > 16384u) return fail(REX_INVALID_ARGUMENT, "image dimension above 16384");
    if (f.maxBounce == 0 || f.maxBounce > 16u) return fail(REX_INVALID_ARGUMENT, "maxBounce must be in 1..16");
    if (f.queueMode != REX_QUEUE_ATOMIC_APPEND) return fail(REX_UNSUPPORTED, "served jobs use atomic appends");
    rex::ServerJob j;
    j.kind = job->kind;
    j.scene = job->sceneIndex;
    j.frame = f;
    if (job->kind == REX_JOB_FRAME) {
        if (f.estimator > REX_ESTIMATOR_MIS) return fail(REX_UNSUPPORTED, "frames use BSDF-only, NEE or MIS");
        if (pixels * f.spp >= 0xffffffffull) return fail(REX_UNSUPPORTED, "more than 2^32 paths in one frame");
        const uint64_t out = rex::RangeAllocator::round(pixels * 16u) + pixels * f.spp * 16u;
        if (out > s.engineArenaBytes) return fail(REX_UNSUPPORTED, "frame outputs exceed the engine's output arena");
        j.frames = 1;
    } else if (job->kind == REX_JOB_RESTIR_SEQUENCE) {
        if (job->frames == 0u || job->frames > rex::kMaxSequenceFrames) {
            return fail(REX_INVALID_ARGUMENT, "sequence length must be in 1..64");
        }
        if (f.estimator != REX_ESTIMATOR_RESTIR || f.spp != 1u) {
            return fail(REX_INVALID_ARGUMENT, "sequences render ReSTIR frames at one sample per pixel");
        }
        if (pixels > s.desc.sequenceMaxPixels) return fail(REX_UNSUPPORTED, "sequence larger than a renderer slot");
        if (job->restir != nullptr && job->restir->estimatorMode != REX_RESTIR_ESTIMATOR_BIASED) {
            // Its (K + 1)-per-pixel visibility queries would be allocated
            // inside the loop; rex_render_frame renders such frames.
            return fail(REX_UNSUPPORTED, "request-loop sequences use the biased ReSTIR estimator");
        }
        if (pixels * 16u > s.sequenceOut.ranges.total()) {
            return fail(REX_UNSUPPORTED, "sequence Film exceeds the sequence output arena");
        }
        j.frames = job->frames;
        j.cameras.assign(job->frames, f.camera);
        if (job->cameras != nullptr) std::memcpy(j.cameras.data(), job->cameras, job->frames * sizeof(rex_camera));
        if (job->restir != nullptr) {
            j.restir = *job->restir;
            j.ownRestir = true;
        }
    } else {
        return fail(REX_INVALID_ARGUMENT, "unknown job kind");
    }
    j.footprint = s.footprint(j);
    if (s.residentBytes + j.footprint > s.budget) return fail(REX_UNSUPPORTED, "job footprint exceeds the budget");
    if (rex_status st = s.prepare()) return st;
    j.submitNs = rex::now_ns();
    const uint64_t id = s.nextId++;
    s.jobs.emplace(id, std::move(j));
    (job->kind == REX_JOB_FRAME ? s.frameQueue : s.sequenceQueue).push_back(id);
    ++s.submitted;
    *jobId = id;
    return REX_SUCCESS;
}

rex_status rex_server_pump(rex_server server, uint32_t* finished) {
    if (server == nullptr) return fail(REX_INVALID_ARGUMENT, "server is NULL");
    try {
        return server->s.pump(finished);
    } catch (const std::exception& e) {
        return fail(REX_RECORDING_FAILED, std::string("internal: ") + e.what());
    }
}

rex_status rex_server_drain(rex_server server) {
    if (server == nullptr) return fail(REX_INVALID_ARGUMENT, "server is NULL");
    rex::Server& s = server->s;
    for (;;) {
        if (rex_status st = rex_server_pump(server, nullptr)) return st;
        const bool queued = !s.frameQueue.empty() || !s.sequenceQueue.empty();
        if (!queued && !s.busy()) return REX_SUCCESS;
        if (queued && !s.busy()) {
            return fail(REX_INVALID_ARGUMENT, "budget or output arenas full: release finished jobs before draining");
        }
    }
}

rex_status rex_server_job(rex_server server, uint64_t jobId, rex_server_job_info* info) {
    if (server == nullptr || info == nullptr) return fail(REX_INVALID_ARGUMENT, "NULL argument");
    auto it = server->s.jobs.find(jobId);
    if (it == server->s.jobs.end()) return fail(REX_INVALID_ARGUMENT, "unknown job");
    const rex::ServerJob& j = it->second;
    std::memset(info, 0, sizeof(*info));
    info->state = j.state;
    info->kind = j.kind;
    info->width = j.frame.width;
    info->height = j.frame.height;
    info->footprintBytes = j.footprint;
    info->submitNs = j.submitNs;
    info->admitNs = j.admitNs;
    info->finishNs = j.finishNs;
    info->rays = j.rays;
    info->film = j.state == 2 ? j.film : nullptr;
    return REX_SUCCESS;
}

rex_status rex_server_release(rex_server server, uint64_t jobId) {
    if (server == nullptr) return fail(REX_INVALID_ARGUMENT, "server is NULL");
    rex::Server& s = server->s;
    auto it = s.jobs.find(jobId);
    if (it == s.jobs.end()) return fail(REX_INVALID_ARGUMENT, "unknown job");
    rex::ServerJob& j = it->second;
    if (j.state != 2) return fail(REX_INVALID_ARGUMENT, "job has not finished");
    if (j.kind == REX_JOB_FRAME) {
        if (rex_status st = rex_batch_release(s.batch, j.batchJob)) return st;
    } else {
        s.sequenceOut.release(j.film);
    }
    s.admittedBytes -= j.footprint;
    ++s.released;
    s.jobs.erase(it);
    return REX_SUCCESS;
}

rex_status rex_server_get_stats(rex_server server, rex_server_stats* stats) {
    if (server == nullptr || stats == nullptr) return fail(REX_INVALID_ARGUMENT, "NULL argument");
    const rex::Server& s = server->s;
    std::memset(stats, 0, sizeof(*stats));
    rex_batch_stats bs;
    if (rex_status st = rex_batch_get_stats(s.batch, &bs)) return st;
    stats->submitted = s.submitted;
    stats->admitted = s.admitted;
    stats->finished = s.finished;
    stats->released = s.released;
    stats->queuedJobs = s.frameQueue.size() + s.sequenceQueue.size();
    stats->residentBytes = s.residentBytes;
    stats->admittedBytes = s.admittedBytes;
    stats->peakAdmittedBytes = s.peakAdmittedBytes;
    stats->budgetBytes = s.budget;
    stats->engineSteps = bs.steps;
    stats->engineRays = bs.raysTraced;
    stats->sequenceRays = s.sequenceRays;
    stats->sequenceFrames = s.sequenceFrames;
    return REX_SUCCESS;
}

}  // extern "C"
