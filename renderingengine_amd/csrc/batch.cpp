// Batch engine: the paths of many render jobs in one pool, traced and
// shaded by one shared launch sequence per step (continuous batching at path
// granularity).
//
// Device state:
//   arena      the resident scenes' compressed 4-wide nodes (references relocated),
//              triangles, triangle attributes and emitter map, one array
//              each; SceneRoot per scene (root reference + box)
//   pool       PathState[capacity] and a free list of slots; every path in
//              the pool has one ray in a queue or is about to retire
//   queues     rays[2][capacity] + rayPath + raySceneOf, hits, shadow items
//              + shadowSceneOf, retire lists[2][capacity]
//   jobs       BatchJob[maxJobs]: frame constants, shading view, Film and
//              per-sample values of each admitted job; retired-path counts
//   generation the admitted jobs' sample streams in admission order (global
//              path index -> job, pixel, sample) and the start cursor
//
// Step (parity c), all on one stream, graph-capturable:
//   traversal  the previous Shade's shadow queries + rays[c] (one launch)
//   finalize   paths the previous Shade retired: sample value into the
//              job's sample array, slot onto the free list, job count
//   resolve    the jobs whose every path has retired: Film running mean over
//              their samples in order (bit-identical to a plain frame)
//   reset      free list top raised, the counters Shade fills cleared
//   Shade      rays[c] -> rays[c ^ 1], shadow queries, retire list
//   regen      new paths of the generation stream into free slots, their
//              primary rays appended to rays[c ^ 1]
// Only Shade appends with atomics (one per queue per block); finalize and
// regen write at positions derived from the counts, since atomics on one
// address serialize.
// The pool stays full while any admitted job has paths left to start, so a
// step's launches are never short of work until the last job drains.  The
// host learns the retired counts and the cursor a few steps late through
// asynchronous copies; nothing waits for the device between steps.
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

hipError_t launch_relocate_wide4q(Wide4Q* dst, const Wide4Q* src, uint32_t n, uint32_t nodeBase, uint32_t triBase,
                                 hipStream_t stream);
hipError_t launch_iota_u32(uint32_t* dst, uint32_t n, hipStream_t stream);

constexpr uint32_t kStaging = 2;  // steps whose uploads / readbacks may be in flight

struct ArenaScene {
    Scene* src = nullptr;
    uint32_t w4Base = 0, w4Count = 0, triBase = 0, triCount = 0;
    SceneRoot root{};
    ShadingView sh{};  // the scene's tables; emitter map indexed by arena triangle
};

struct Job {
    uint32_t scene = 0;
    FrameConst f{};
    float4* film = nullptr;     // output arena, from admission to release
    float4* samples = nullptr;
    uint64_t outBytes = 0;      // Film + samples in the arena
    uint32_t pathCount = 0;
    uint32_t slot = kInvalid;
    uint32_t state = 0;  // 0 queued, 1 running, 2 resolve issued, 3 finished
    uint64_t genStart = 0;
    uint64_t submitStep = 0, admitStep = 0, finishStep = 0;
    uint32_t retired = 0;       // retired paths as of the latest readback
    hipEvent_t done = nullptr;  // recorded after the step that resolves it
};

struct Staging {
    // Host pinned.  Uploads of one step: new job records, the generation
    // list, control words, the resolve list.  Readback of one step: retired
    // counts per job slot and the control words.
    BatchJob* jobs = nullptr;       // maxJobs
    uint32_t* jobSlots = nullptr;   // maxJobs
    uint32_t* genJob = nullptr;     // maxJobs
    uint64_t* genStart = nullptr;   // maxJobs + 1
    uint32_t* resolve = nullptr;    // resolveJob[maxJobs], resolveStart[maxJobs + 1]
    uint32_t* ctl = nullptr;        // kCtlWords (host-written words)
    uint32_t* doneBack = nullptr;   // kDoneStripes x maxJobs (readback)
    uint32_t* ctlBack = nullptr;    // kCtlWords (readback)
    hipEvent_t done = nullptr;
    bool pending = false;
    uint64_t stepIssued = 0;        // the step whose results the readback holds
};

struct Batch {
    uint32_t capacity = 0, maxJobs = 0;
    bool useGraphs = false;
    DeviceArena outputs;
    uint32_t shadeBpc = 0;
    bool extended = false;  // some scene has sphere lights or a lat-long map
    uint32_t reserve = 0;   // free slots kept out of regeneration (path budget = capacity - reserve)
    // Online autotuner: windows of kTuneWarm + kTuneSteps busy steps, one per
    // candidate (Shade grid x graphs at the full budget, then smaller path
    // budgets with the best of those); score = traversal queries per second
    // of the window's timed steps.
    struct TuneCandidate {
        uint32_t bpc;
        bool graphs;
        uint32_t reserve;
        double raysPerSec;
    };
    static constexpr uint32_t kTuneWarm = 2, kTuneSteps = 4;
    bool tuning = false, tuned = false, windowOpen = false;
    std::vector<TuneCandidate> candidates;
    uint32_t tuneIndex = 0, windowSteps = 0;
    uint64_t windowRays = 0;
    std::chrono::steady_clock::time_point windowStart;
    int cus = 1;
    // arena
    Wide4Q* w4 = nullptr;  // compressed 4-wide nodes of every resident scene
    TriPos* tris = nullptr;
    TriAttr* attrs = nullptr;
    LeafBox* leafBox = nullptr;  // exact leaf boxes (per triangle) of every resident scene
    uint32_t* emitter = nullptr;
    uint32_t w4Cap = 0, triCap = 0, w4Count = 0, triCount = 0, depth = 0;
    SceneRoot* dRoots = nullptr;
    uint32_t rootsCap = 0;
    std::vector<ArenaScene> scenes;
    SceneView view{};
    // pool and queues
    PathState* paths = nullptr;
    F4* rays[2] = {nullptr, nullptr};
    uint32_t* rayPath[2] = {nullptr, nullptr};
    uint32_t* raySceneOf[2] = {nullptr, nullptr};
    QueueHit* hits = nullptr;
    ShadowItem* shadows = nullptr;
    uint32_t* shadowSceneOf = nullptr;
    uint32_t* retire[2] = {nullptr, nullptr};
    uint32_t* freeList = nullptr;
    uint32_t* counts = nullptr;  // kCnt* queue counters
    uint32_t* ctl = nullptr;     // kCtl* words
    BatchJob* dJobs = nullptr;
    uint32_t* doneCount = nullptr;  // kDoneStripes x maxJobs
    BatchJob* admitJobs = nullptr;  // staged records of this step's admissions
    uint32_t* admitSlots = nullptr;
    uint32_t* genJob = nullptr;
    uint64_t* genStart = nullptr;
    uint32_t* resolveJob = nullptr;
    uint32_t* resolveStart = nullptr;
    TraceWorkspace ws;
    Staging staging[kStaging];
    uint32_t stagingNext = 0;
    hipGraphExec_t graphs[2] = {nullptr, nullptr};
    hipStream_t graphStream = nullptr;
    // host bookkeeping
    std::map<uint64_t, Job> jobs;
    std::deque<uint64_t> queue;        // submitted, not admitted
    std::deque<uint64_t> generating;   // admitted, in admission order, not known fully started
    std::vector<uint64_t> slotJob;     // job id per device slot (0: free)
    std::vector<uint32_t> freeSlots;
    uint64_t genTotal = 0, cursorSeen = 0;
    uint64_t nextJob = 1, step = 0, launchedSteps = 0, finished = 0, running = 0;
    uint32_t cur = 0;

    BatchDevice device() const {
        BatchDevice d;
        d.jobs = dJobs;
        d.paths = paths;
        d.freeList = freeList;
        d.ctl = ctl;
        d.genJob = genJob;
        d.genStart = genStart;
        d.doneCount = doneCount;
        d.resolveJob = resolveJob;
        d.resolveStart = resolveStart;
        d.jobStride = maxJobs;
        return d;
    }
    void release();
    rex_status grow_arena(uint32_t needW4, uint32_t needTri);
    rex_status add_scene(Scene& s, uint32_t& index);
    rex_status record_step(hipStream_t st, uint32_t parity);
    rex_status step_once(hipStream_t st);
    rex_status tune_before(hipStream_t st);
    rex_status tune_after(hipStream_t st);
    void apply(const TuneCandidate& c);
    void keep_best(uint32_t upTo);
    rex_status device_rays(hipStream_t st, uint64_t& rays);
};

void Batch::apply(const TuneCandidate& c) {
    if (c.bpc != shadeBpc || c.graphs != useGraphs) {
        // Captured steps hold the old Shade grid.
        for (hipGraphExec_t& g : graphs) {
            if (g) (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    }
    shadeBpc = c.bpc;
    useGraphs = c.graphs;
    reserve = c.reserve;
}

rex_status Batch::device_rays(hipStream_t st, uint64_t& rays) {
    uint32_t w[2] = {0, 0};
    if (rex_status s = check_hip(hipStreamSynchronize(st), "tune sync")) return s;
    if (rex_status s = check_hip(hipMemcpy(w, ctl + kCtlRaysLo, sizeof(w), hipMemcpyDeviceToHost), "tune read")) return s;
    rays = (uint64_t)w[0] | ((uint64_t)w[1] << 32);
    return REX_SUCCESS;
}

// Opens the next candidate's window once the engine is busy (jobs running
// and the two uncaptured steps done).
rex_status Batch::tune_before(hipStream_t st) {
    if (!tuning) return REX_SUCCESS;
    if (windowSteps == 0) {
        if (running == 0 || launchedSteps < 2) return REX_SUCCESS;
        apply(candidates[tuneIndex]);
        windowOpen = true;
    }
    if (windowSteps == kTuneWarm) {
        if (rex_status s = device_rays(st, windowRays)) return s;
        windowStart = std::chrono::steady_clock::now();
    }
    return REX_SUCCESS;
}

void Batch::keep_best(uint32_t upTo) {
    TuneCandidate best = candidates[0];
    for (uint32_t k = 1; k < upTo; ++k) {
        if (candidates[k].raysPerSec > best.raysPerSec) best = candidates[k];
    }
    apply(best);
}

// Closes a window after its timed steps; after the last candidate the best
// setting stays.  If the work runs out inside a window, the best finished
// candidate stays.
rex_status Batch::tune_after(hipStream_t st) {
    if (!tuning || !windowOpen) return REX_SUCCESS;
    ++windowSteps;
    if (windowSteps < kTuneWarm + kTuneSteps) {
        if (running == 0) {
            tuning = false;
            tuned = true;
            windowOpen = false;
            windowSteps = 0;
            keep_best(tuneIndex > 0u ? tuneIndex : 1u);
        }
        return REX_SUCCESS;
    }
    uint64_t rays = 0;
    if (rex_status s = device_rays(st, rays)) return s;
    const double secs = std::chrono::duration<double>(std::chrono::steady_clock::now() - windowStart).count();
    candidates[tuneIndex].raysPerSec = secs > 0.0 ? (double)(rays - windowRays) / secs : 0.0;
    windowSteps = 0;
    windowOpen = false;
    ++tuneIndex;
    if (tuneIndex == 4u) {
        // Path budgets for the best Shade grid / graphs setting.
        TuneCandidate best = candidates[0];
        for (uint32_t k = 1; k < 4u; ++k) {
            if (candidates[k].raysPerSec > best.raysPerSec) best = candidates[k];
        }
        candidates.push_back({best.bpc, best.graphs, capacity / 2u, 0.0});
        candidates.push_back({best.bpc, best.graphs, capacity / 4u, 0.0});
    }
    if (tuneIndex == candidates.size()) {
        keep_best(tuneIndex);
        tuning = false;
        tuned = true;
    } else {
        apply(candidates[0]);  // until the next window opens
    }
    return REX_SUCCESS;
}

void Batch::release() {
    for (hipGraphExec_t& g : graphs) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
    if (graphStream) (void)hipStreamDestroy(graphStream);
    graphStream = nullptr;
    for (Staging& s : staging) {
        if (s.done) (void)hipEventDestroy(s.done);
        void* host[] = {s.jobs, s.jobSlots, s.genJob, s.genStart, s.resolve, s.ctl, s.doneBack, s.ctlBack};
        for (void* p : host) {
            if (p) (void)hipHostFree(p);
        }
        s = Staging{};
    }
    for (auto& kv : jobs) {
        if (kv.second.done) (void)hipEventDestroy(kv.second.done);
    }
    jobs.clear();
    outputs.destroy();
    void* ptrs[] = {w4, tris, attrs, leafBox, emitter, dRoots, paths, rays[0], rays[1], rayPath[0], rayPath[1],
                    raySceneOf[0], raySceneOf[1], hits, shadows, shadowSceneOf, retire[0], retire[1], freeList,
                    counts, ctl, dJobs, doneCount, admitJobs, admitSlots, genJob, genStart, resolveJob,
                    resolveStart};
    for (void* p : ptrs) {
        if (p) (void)hipFree(p);
    }
    ws.release();
}

rex_status Batch::grow_arena(uint32_t needW4, uint32_t needTri) {
    if (needW4 <= w4Cap && needTri <= triCap) return REX_SUCCESS;
    if (rex_status st = check_hip(hipDeviceSynchronize(), "arena grow sync")) return st;
    const uint32_t newW4 = std::max(needW4, w4Cap * 2u);
    const uint32_t newTri = std::max(needTri, triCap * 2u);
    Wide4Q* nw = nullptr;
    TriPos* nt = nullptr;
    TriAttr* na = nullptr;
    uint32_t* ne = nullptr;
    LeafBox* nb = nullptr;
    if (rex_status st = check_hip(hipMalloc(&nb, (size_t)newTri * sizeof(LeafBox)), "hipMalloc arena")) return st;
    if (rex_status st = check_hip(hipMalloc(&nw, (size_t)newW4 * sizeof(Wide4Q)), "hipMalloc arena")) {
        (void)hipFree(nb);
        return st;
    }
    if (rex_status st = check_hip(hipMalloc(&nt, (size_t)newTri * sizeof(TriPos)), "hipMalloc arena")) {
        (void)hipFree(nb);
        (void)hipFree(nw);
        return st;
    }
    if (rex_status st = check_hip(hipMalloc(&na, (size_t)newTri * sizeof(TriAttr)), "hipMalloc arena")) {
        (void)hipFree(nb);
        (void)hipFree(nw);
        (void)hipFree(nt);
        return st;
    }
    if (rex_status st = check_hip(hipMalloc(&ne, (size_t)newTri * sizeof(uint32_t)), "hipMalloc arena")) {
        (void)hipFree(nb);
        (void)hipFree(nw);
        (void)hipFree(nt);
        (void)hipFree(na);
        return st;
    }
    if (w4Count) (void)hipMemcpy(nw, w4, (size_t)w4Count * sizeof(Wide4Q), hipMemcpyDeviceToDevice);
    if (triCount) {
        (void)hipMemcpy(nt, tris, (size_t)triCount * sizeof(TriPos), hipMemcpyDeviceToDevice);
        (void)hipMemcpy(na, attrs, (size_t)triCount * sizeof(TriAttr), hipMemcpyDeviceToDevice);
        (void)hipMemcpy(ne, emitter, (size_t)triCount * sizeof(uint32_t), hipMemcpyDeviceToDevice);
        (void)hipMemcpy(nb, leafBox, (size_t)triCount * sizeof(LeafBox), hipMemcpyDeviceToDevice);
    }
    (void)hipFree(w4);
    (void)hipFree(tris);
    (void)hipFree(attrs);
    (void)hipFree(emitter);
    (void)hipFree(leafBox);
    leafBox = nb;
    w4 = nw;
    tris = nt;
    attrs = na;
    emitter = ne;
    w4Cap = newW4;
    triCap = newTri;
    return check_hip(hipDeviceSynchronize(), "arena copy");
}

rex_status Batch::add_scene(Scene& s, uint32_t& index) {
    if (!s.shade.ready) return fail(REX_MISSING_SCENE, "scene has no shading tables (rex_scene_set_shading)");
    if (s.triCount == 0 || s.wide4qCount == 0) return fail(REX_INVALID_ARGUMENT, "empty scene");
    if (running != 0 || !queue.empty()) return fail(REX_INVALID_ARGUMENT, "scenes are added while no job is in flight");
    if ((uint64_t)triCount + s.triCount > kLeafFirstMask) return fail(REX_UNSUPPORTED, "arena holds 2^28 triangles");
    if (rex_status st = grow_arena(w4Count + s.wide4qCount, triCount + s.triCount)) return st;
    ArenaScene a;
    a.src = &s;
    a.w4Base = w4Count;
    a.w4Count = s.wide4qCount;
    a.triBase = triCount;
    a.triCount = s.triCount;
    const uint32_t r = s.view.rootRefQ;
    a.root.rootRef = is_leaf_ref(r) ? ((r & ~kLeafFirstMask) | ((r & kLeafFirstMask) + a.triBase)) : r + a.w4Base;
    a.sh = s.shade.view();
    extended = extended || s.shade.extended;
    for (int k = 0; k < 3; ++k) {
        a.root.rootMin[k] = s.view.rootMin[k];
        a.root.rootMax[k] = s.view.rootMax[k];
    }
    if (rex_status st = check_hip(launch_relocate_wide4q(w4 + a.w4Base, s.dWide4Q, s.wide4qCount, a.w4Base, a.triBase,
                                                        nullptr),
                                  "arena relocate")) {
        return st;
    }
    if (rex_status st = check_hip(hipMemcpy(tris + a.triBase, s.dTris, (size_t)s.triCount * sizeof(TriPos),
                                            hipMemcpyDeviceToDevice),
                                  "arena copy")) {
        return st;
    }
    if (rex_status st = check_hip(hipMemcpy(attrs + a.triBase, s.dAttrs, (size_t)s.triCount * sizeof(TriAttr),
                                            hipMemcpyDeviceToDevice),
                                  "arena copy")) {
        return st;
    }
    if (rex_status st = check_hip(hipMemcpy(leafBox + a.triBase, s.dLeafBox, (size_t)s.triCount * sizeof(LeafBox),
                                            hipMemcpyDeviceToDevice),
                                  "arena copy")) {
        return st;
    }
    if (rex_status st = check_hip(hipMemcpy(emitter + a.triBase, s.shade.dEmitterOfTri,
                                            (size_t)s.triCount * sizeof(uint32_t), hipMemcpyDeviceToDevice),
                                  "arena copy")) {
        return st;
    }
    w4Count += s.wide4qCount;
    triCount += s.triCount;
    depth = std::max(depth, s.view.wide4Depth);
    scenes.push_back(a);
    for (ArenaScene& e : scenes) e.sh.emitterOfTri = emitter;
    if (scenes.size() > rootsCap) {
        const uint32_t cap = std::max<uint32_t>(64u, rootsCap * 2u);
        SceneRoot* nr = nullptr;
        if (rex_status st = check_hip(hipMalloc(&nr, cap * sizeof(SceneRoot)), "hipMalloc roots")) return st;
        if (dRoots) (void)hipFree(dRoots);
        dRoots = nr;
        rootsCap = cap;
    }
    std::vector<SceneRoot> roots(scenes.size());
    for (size_t i = 0; i < scenes.size(); ++i) roots[i] = scenes[i].root;
    if (rex_status st = check_hip(hipMemcpy(dRoots, roots.data(), roots.size() * sizeof(SceneRoot),
                                            hipMemcpyHostToDevice),
                                  "roots upload")) {
        return st;
    }
    view = SceneView{};
    view.tris = tris;
    view.attrs = attrs;
    view.wide4q = w4;
    view.leafBox = leafBox;
    view.nodeCount = w4Count;
    view.triCount = triCount;
    view.wide4Depth = depth;
    // Graphs captured before this scene saw the old arena pointers.
    for (hipGraphExec_t& g : graphs) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
    index = static_cast<uint32_t>(scenes.size() - 1);
    return check_hip(hipDeviceSynchronize(), "arena add");
}

// The launch sequence of one step for ray-queue parity `parity`.
rex_status Batch::record_step(hipStream_t st, uint32_t parity) {
    const uint32_t c = parity, n = parity ^ 1u;
    const BatchDevice d = device();
    const uint32_t grid = static_cast<uint32_t>(cus) * 8u;
    uint32_t* const rayCount[2] = {counts + kCntRays0, counts + kCntRays1};
    uint32_t* const retireCount[2] = {counts + kCntRetire0, counts + kCntRetire1};
    uint32_t* const shadowCount = counts + kCntShadow;
    if (rex_status s = check_hip(launch_queue_mixed_arena(view, ws, shadows, shadowCount, shadowSceneOf, rays[c],
                                                          rayCount[c], raySceneOf[c], dRoots, 2u * capacity, hits,
                                                          paths, st),
                                 "batch traversal")) {
        return s;
    }
    if (rex_status s = check_hip(launch_batch_finalize(d, retire[n], retireCount[n], grid, st), "finalize")) return s;
    if (rex_status s = check_hip(launch_batch_resolve(d, grid, st), "batch resolve")) return s;
    if (rex_status s = check_hip(launch_batch_reset(ctl, retireCount[n], rayCount[n], shadowCount, rayCount[c], st),
                                 "reset")) {
        return s;
    }
    ShadeLaunch L{};
    L.capacity = capacity;
    L.s = view;
    L.rays = rays[c];
    L.rayPath = rayPath[c];
    L.count = rayCount[c];
    L.hits = hits;
    L.paths = paths;
    L.nextRays = rays[n];
    L.nextPath = rayPath[n];
    L.nextCount = rayCount[n];
    L.shadows = shadows;
    L.shadowCount = shadowCount;
    L.jobs = dJobs;
    L.nextSceneOf = raySceneOf[n];
    L.shadowSceneOf = shadowSceneOf;
    L.blocksPerCu = shadeBpc;
    L.extended = extended;
    L.retire = retire[c];
    L.retireCount = retireCount[c];
    if (rex_status s = check_hip(launch_shade(L, st), "batch shade")) return s;
    return check_hip(launch_batch_regen(d, rays[n], rayPath[n], raySceneOf[n], rayCount[n], grid, st), "regen");
}

rex_status Batch::step_once(hipStream_t st) {
    Staging& sg = staging[stagingNext];
    if (sg.pending) {
        // The readback of the step that last used this staging slot.
        if (rex_status s = check_hip(hipEventSynchronize(sg.done), "staging wait")) return s;
        sg.pending = false;
        const uint64_t cursor = (uint64_t)sg.ctlBack[kCtlCursorLo] | ((uint64_t)sg.ctlBack[kCtlCursorHi] << 32);
        if (cursor > cursorSeen) cursorSeen = cursor;
        for (uint32_t slot = 0; slot < maxJobs; ++slot) {
            const uint64_t id = slotJob[slot];
            if (id == 0) continue;
            Job& j = jobs.at(id);
            // Counts read back before the job's admission belong to the
            // slot's previous job.
            if (j.state == 1 && j.admitStep <= sg.stepIssued) {
                uint32_t r = 0;
                for (uint32_t k = 0; k < kDoneStripes; ++k) r += sg.doneBack[k * maxJobs + slot];
                j.retired = r;
            }
        }
    }
    // Jobs whose every path has retired (as of the latest readback).
    uint32_t* resolveList = sg.resolve;
    uint32_t* resolveStartH = sg.resolve + maxJobs;
    uint32_t nResolve = 0;
    resolveStartH[0] = 0;
    std::vector<uint64_t> resolving;
    for (uint32_t slot = 0; slot < maxJobs && running != 0; ++slot) {
        const uint64_t id = slotJob[slot];
        if (id == 0) continue;
        const Job& j = jobs.at(id);
        if (j.state != 1 || j.retired != j.pathCount) continue;
        resolveList[nResolve] = slot;
        resolveStartH[nResolve + 1] = resolveStartH[nResolve] + j.f.width * j.f.height;
        ++nResolve;
        resolving.push_back(id);
    }
    // Admissions: jobs take free slots in submission order while their
    // outputs fit the arena.
    uint32_t nAdmit = 0;
    while (!queue.empty() && !freeSlots.empty()) {
        const uint64_t id = queue.front();
        Job& j = jobs.at(id);
        char* out = static_cast<char*>(outputs.alloc(j.outBytes));
        if (out == nullptr) break;
        queue.pop_front();
        const size_t filmBytes = RangeAllocator::round((size_t)j.f.width * j.f.height * sizeof(float4));
        j.film = reinterpret_cast<float4*>(out);
        j.samples = reinterpret_cast<float4*>(out + filmBytes);
        j.slot = freeSlots.back();
        freeSlots.pop_back();
        slotJob[j.slot] = id;
        j.genStart = genTotal;
        genTotal += j.pathCount;
        j.state = 1;
        j.admitStep = step;
        j.retired = 0;
        ++running;
        generating.push_back(id);
        BatchJob& bj = sg.jobs[nAdmit];
        bj = BatchJob{};
        bj.f = j.f;
        bj.sh = scenes[j.scene].sh;
        bj.scene = j.scene;
        bj.pixelCount = j.f.width * j.f.height;
        bj.pathCount = j.pathCount;
        bj.film = reinterpret_cast<F4*>(j.film);
        bj.samples = reinterpret_cast<F4*>(j.samples);
        sg.jobSlots[nAdmit] = j.slot;
        ++nAdmit;
        if (rex_status s = check_hip(hipMemsetAsync(j.film, 0, (size_t)bj.pixelCount * sizeof(float4), st),
                                     "film clear")) {
            return s;
        }
    }
    // Generation entries the device may still draw from.
    while (!generating.empty()) {
        const Job& j = jobs.at(generating.front());
        if (j.genStart + j.pathCount > cursorSeen) break;
        generating.pop_front();
    }
    if (running == 0 && nAdmit == 0) return REX_SUCCESS;  // idle
    uint32_t nGen = 0;
    for (uint64_t id : generating) {
        const Job& j = jobs.at(id);
        sg.genJob[nGen] = j.slot;
        sg.genStart[nGen] = j.genStart;
        ++nGen;
    }
    sg.genStart[nGen] = genTotal;
    sg.ctl[kCtlGenCount] = nGen;
    sg.ctl[kCtlGenTotalLo] = static_cast<uint32_t>(genTotal);
    sg.ctl[kCtlGenTotalHi] = static_cast<uint32_t>(genTotal >> 32);
    sg.ctl[kCtlResolveCount] = nResolve;
    if (nAdmit) {
        if (rex_status s = check_hip(hipMemcpyAsync(admitJobs, sg.jobs, nAdmit * sizeof(BatchJob),
                                                    hipMemcpyHostToDevice, st),
                                     "job upload")) {
            return s;
        }
        if (rex_status s = check_hip(hipMemcpyAsync(admitSlots, sg.jobSlots, nAdmit * sizeof(uint32_t),
                                                    hipMemcpyHostToDevice, st),
                                     "job upload")) {
            return s;
        }
        if (rex_status s = check_hip(launch_batch_admit(device(), dJobs, admitJobs, admitSlots, nAdmit, st),
                                     "admission")) {
            return s;
        }
    }
    if (nGen) {
        if (rex_status s = check_hip(hipMemcpyAsync(genJob, sg.genJob, nGen * sizeof(uint32_t), hipMemcpyHostToDevice,
                                                    st),
                                     "generation upload")) {
            return s;
        }
        if (rex_status s = check_hip(hipMemcpyAsync(genStart, sg.genStart, (nGen + 1) * sizeof(uint64_t),
                                                    hipMemcpyHostToDevice, st),
                                     "generation upload")) {
            return s;
        }
    }
    if (nResolve) {
        if (rex_status s = check_hip(hipMemcpyAsync(resolveJob, resolveList, nResolve * sizeof(uint32_t),
                                                    hipMemcpyHostToDevice, st),
                                     "resolve upload")) {
            return s;
        }
        if (rex_status s = check_hip(hipMemcpyAsync(resolveStart, resolveStartH, (nResolve + 1) * sizeof(uint32_t),
                                                    hipMemcpyHostToDevice, st),
                                     "resolve upload")) {
            return s;
        }
    }
    if (rex_status s = check_hip(hipMemcpyAsync(ctl + kCtlGenCount, sg.ctl + kCtlGenCount, sizeof(uint32_t),
                                                hipMemcpyHostToDevice, st),
                                 "control upload")) {
        return s;
    }
    sg.ctl[kCtlReserve] = reserve;
    if (rex_status s = check_hip(hipMemcpyAsync(ctl + kCtlGenTotalLo, sg.ctl + kCtlGenTotalLo, 4 * sizeof(uint32_t),
                                                hipMemcpyHostToDevice, st),
                                 "control upload")) {
        return s;
    }
    const uint32_t parity = cur;
    // The first steps run uncaptured: they size the traversal workspace.
    if (useGraphs && launchedSteps >= 2) {
        if (graphs[parity] == nullptr) {
            if (graphStream == nullptr) {
                if (rex_status s = check_hip(hipStreamCreateWithFlags(&graphStream, hipStreamNonBlocking),
                                             "graph stream")) {
                    return s;
                }
            }
            hipGraph_t g = nullptr;
            if (rex_status s = check_hip(hipStreamBeginCapture(graphStream, hipStreamCaptureModeThreadLocal),
                                         "begin capture")) {
                return s;
            }
            const rex_status rs = record_step(graphStream, parity);
            hipError_t e = hipStreamEndCapture(graphStream, &g);
            if (rs != REX_SUCCESS) {
                if (g) (void)hipGraphDestroy(g);
                return rs;
            }
            if (rex_status s = check_hip(e, "end capture")) return s;
            e = hipGraphInstantiate(&graphs[parity], g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (rex_status s = check_hip(e, "graph instantiate")) return s;
        }
        if (rex_status s = check_hip(hipGraphLaunch(graphs[parity], st), "graph launch")) return s;
    } else {
        if (rex_status s = record_step(st, parity)) return s;
    }
    // Readback for a later step.
    if (rex_status s = check_hip(hipMemcpyAsync(sg.doneBack, doneCount, kDoneStripes * maxJobs * sizeof(uint32_t),
                                                hipMemcpyDeviceToHost, st),
                                 "count readback")) {
        return s;
    }
    if (rex_status s = check_hip(hipMemcpyAsync(sg.ctlBack, ctl, kCtlWords * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                                st),
                                 "control readback")) {
        return s;
    }
    if (rex_status s = check_hip(hipEventRecord(sg.done, st), "step event")) return s;
    sg.pending = true;
    sg.stepIssued = step;
    for (uint64_t id : resolving) {
        Job& j = jobs.at(id);
        j.state = 2;
        j.finishStep = step;
        if (j.done == nullptr) {
            if (rex_status s = check_hip(hipEventCreateWithFlags(&j.done, hipEventDisableTiming), "job event"))
                return s;
        }
        if (rex_status s = check_hip(hipEventRecord(j.done, st), "job event")) return s;
        slotJob[j.slot] = 0;
        freeSlots.push_back(j.slot);
        --running;
        ++finished;
    }
    stagingNext = (stagingNext + 1) % kStaging;
    cur ^= 1u;
    ++step;
    ++launchedSteps;
    return REX_SUCCESS;
}

}  // namespace rex

struct rex_batch_t {
    rex::Batch b;
};

using rex::check_hip;
using rex::fail;

namespace rex {
bool batch_busy(rex_batch batch) { return !batch->b.queue.empty() || batch->b.running != 0; }
}  // namespace rex

extern "C" {

rex_status rex_batch_create(const rex_batch_desc* desc, rex_batch* out) {
    if (out == nullptr) return fail(REX_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    rex_batch_desc d{};
    if (desc != nullptr) d = *desc;
    const uint32_t capacity = d.pathCapacity != 0u ? d.pathCapacity : (16u << 20);
    const uint32_t maxJobs = d.maxJobs != 0u ? d.maxJobs : 256u;
    const size_t outputBytes = (size_t)(d.outputMiB != 0u ? d.outputMiB : 16384u) << 20;
    if (capacity >= 0x40000000u) return fail(REX_UNSUPPORTED, "path capacity above 2^30");
    if (maxJobs > 65536u) return fail(REX_INVALID_ARGUMENT, "maxJobs above 65536");
    if (d.shadeBlocksPerCu > 64u) return fail(REX_INVALID_ARGUMENT, "Shade grid above 64 blocks per CU");
    rex_batch rb = new (std::nothrow) rex_batch_t();
    if (rb == nullptr) return fail(REX_INVALID_ARGUMENT, "host allocation failed");
    rex::Batch& b = rb->b;
    b.capacity = capacity;
    b.maxJobs = maxJobs;
    b.useGraphs = d.useGraphs != 0u;
    b.shadeBpc = d.shadeBlocksPerCu;
    if (d.autotune != 0u) {
        b.tuning = true;
        b.candidates = {{16u, false, 0u, 0.0}, {8u, false, 0u, 0.0}, {16u, true, 0u, 0.0}, {8u, true, 0u, 0.0}};
    }
    int dev = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&b.cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (b.cus <= 0) b.cus = 1;
    const size_t cap = capacity;
    struct Alloc {
        void** p;
        size_t bytes;
    } allocs[] = {
        {reinterpret_cast<void**>(&b.paths), cap * sizeof(rex::PathState)},
        {reinterpret_cast<void**>(&b.rays[0]), cap * 2 * sizeof(rex::F4)},
        {reinterpret_cast<void**>(&b.rays[1]), cap * 2 * sizeof(rex::F4)},
        {reinterpret_cast<void**>(&b.rayPath[0]), cap * sizeof(uint32_t)},
        {reinterpret_cast<void**>(&b.rayPath[1]), cap * sizeof(uint32_t)},
        {reinterpret_cast<void**>(&b.raySceneOf[0]), cap * sizeof(uint32_t)},
        {reinterpret_cast<void**>(&b.raySceneOf[1]), cap * sizeof(uint32_t)},
        {reinterpret_cast<void**>(&b.hits), cap * sizeof(rex::QueueHit)},
        {reinterpret_cast<void**>(&b.shadows), cap * sizeof(rex::ShadowItem)},
        {reinterpret_cast<void**>(&b.shadowSceneOf), cap * sizeof(uint32_t)},
        {reinterpret_cast<void**>(&b.retire[0]), cap * sizeof(uint32_t)},
        {reinterpret_cast<void**>(&b.retire[1]), cap * sizeof(uint32_t)},
        {reinterpret_cast<void**>(&b.freeList), cap * sizeof(uint32_t)},
        {reinterpret_cast<void**>(&b.counts), rex::kCntWords * sizeof(uint32_t)},
        {reinterpret_cast<void**>(&b.ctl), rex::kCtlWords * sizeof(uint32_t)},
        {reinterpret_cast<void**>(&b.dJobs), (size_t)maxJobs * sizeof(rex::BatchJob)},
        {reinterpret_cast<void**>(&b.doneCount), (size_t)rex::kDoneStripes * maxJobs * sizeof(uint32_t)},
        {reinterpret_cast<void**>(&b.admitJobs), (size_t)maxJobs * sizeof(rex::BatchJob)},
        {reinterpret_cast<void**>(&b.admitSlots), (size_t)maxJobs * sizeof(uint32_t)},
        {reinterpret_cast<void**>(&b.genJob), (size_t)maxJobs * sizeof(uint32_t)},
        {reinterpret_cast<void**>(&b.genStart), ((size_t)maxJobs + 1) * sizeof(uint64_t)},
        {reinterpret_cast<void**>(&b.resolveJob), (size_t)maxJobs * sizeof(uint32_t)},
        {reinterpret_cast<void**>(&b.resolveStart), ((size_t)maxJobs + 1) * sizeof(uint32_t)},
    };
    for (const Alloc& a : allocs) {
        if (rex_status st = check_hip(hipMalloc(a.p, a.bytes), "hipMalloc batch buffer")) {
            b.release();
            delete rb;
            return st;
        }
    }
    if (rex_status st = check_hip(b.outputs.create(outputBytes), "hipMalloc output arena")) {
        b.release();
        delete rb;
        return st;
    }
    for (rex::Staging& s : b.staging) {
        struct HostAlloc {
            void** p;
            size_t bytes;
        } hallocs[] = {
            {reinterpret_cast<void**>(&s.jobs), maxJobs * sizeof(rex::BatchJob)},
            {reinterpret_cast<void**>(&s.jobSlots), maxJobs * sizeof(uint32_t)},
            {reinterpret_cast<void**>(&s.genJob), maxJobs * sizeof(uint32_t)},
            {reinterpret_cast<void**>(&s.genStart), (maxJobs + 1) * sizeof(uint64_t)},
            {reinterpret_cast<void**>(&s.resolve), (2 * maxJobs + 1) * sizeof(uint32_t)},
            {reinterpret_cast<void**>(&s.ctl), rex::kCtlWords * sizeof(uint32_t)},
            {reinterpret_cast<void**>(&s.doneBack), rex::kDoneStripes * maxJobs * sizeof(uint32_t)},
            {reinterpret_cast<void**>(&s.ctlBack), rex::kCtlWords * sizeof(uint32_t)},
        };
        rex_status st = REX_SUCCESS;
        for (const HostAlloc& a : hallocs) {
            if (st == REX_SUCCESS) st = check_hip(hipHostMalloc(a.p, a.bytes), "hipHostMalloc staging");
        }
        if (st == REX_SUCCESS) st = check_hip(hipEventCreateWithFlags(&s.done, hipEventDisableTiming), "staging event");
        if (st != REX_SUCCESS) {
            b.release();
            delete rb;
            return st;
        }
        std::memset(s.ctl, 0, rex::kCtlWords * sizeof(uint32_t));
    }
    rex_status st = check_hip(hipMemset(b.counts, 0, rex::kCntWords * sizeof(uint32_t)), "counter clear");
    if (st == REX_SUCCESS) st = check_hip(hipMemset(b.ctl, 0, rex::kCtlWords * sizeof(uint32_t)), "control clear");
    if (st == REX_SUCCESS) {
        st = check_hip(hipMemset(b.doneCount, 0, rex::kDoneStripes * maxJobs * sizeof(uint32_t)), "count clear");
    }
    if (st == REX_SUCCESS) st = check_hip(hipMemcpy(b.ctl + rex::kCtlFree, &capacity, sizeof(uint32_t),
                                                    hipMemcpyHostToDevice),
                                          "control init");
    if (st == REX_SUCCESS) st = check_hip(rex::launch_iota_u32(b.freeList, capacity, nullptr), "free list init");
    if (st == REX_SUCCESS) st = check_hip(hipDeviceSynchronize(), "batch init");
    if (st != REX_SUCCESS) {
        b.release();
        delete rb;
        return st;
    }
    b.slotJob.assign(maxJobs, 0);
    for (uint32_t i = 0; i < maxJobs; ++i) b.freeSlots.push_back(maxJobs - 1u - i);
    *out = rb;
    return REX_SUCCESS;
}

rex_status rex_batch_destroy(rex_batch batch) {
    if (batch == nullptr) return REX_SUCCESS;
    (void)hipDeviceSynchronize();
    batch->b.release();
    delete batch;
    return REX_SUCCESS;
}

rex_status rex_batch_add_scene(rex_batch batch, rex_scene scene, uint32_t* sceneIndex) {
    if (batch == nullptr || sceneIndex == nullptr) return fail(REX_INVALID_ARGUMENT, "NULL argument");
    if (scene == nullptr) return fail(REX_MISSING_SCENE, "scene is NULL");
    try {
        return batch->b.add_scene(scene->s, *sceneIndex);
    } catch (const std::exception& e) {
        return fail(REX_RECORDING_FAILED, std::string("internal: ") + e.what());
    }
}

rex_status rex_batch_submit(rex_batch batch, uint32_t sceneIndex, const rex_frame_desc* fd, uint64_t* jobId) {
    if (batch == nullptr || fd == nullptr || jobId == nullptr) return fail(REX_INVALID_ARGUMENT, "NULL argument");
    rex::Batch& b = batch->b;
    if (sceneIndex >= b.scenes.size()) return fail(REX_MISSING_SCENE, "scene index not resident");
    const uint64_t pixels = static_cast<uint64_t>(fd->width) * fd->height;
    if (pixels == 0 || fd->spp == 0) return fail(REX_INVALID_ARGUMENT, "empty frame");
    if (fd->width > 16384u || fd->height > 16384u) return fail(REX_INVALID_ARGUMENT, "image dimension above 16384");
    if (fd->maxBounce == 0 || fd->maxBounce > 16u) return fail(REX_INVALID_ARGUMENT, "maxBounce must be in 1..16");
    if (fd->estimator > REX_ESTIMATOR_MIS) return fail(REX_UNSUPPORTED, "batched jobs use BSDF-only, NEE or MIS");
    if (fd->queueMode != REX_QUEUE_ATOMIC_APPEND) return fail(REX_UNSUPPORTED, "batched jobs use atomic appends");
    const uint64_t paths = pixels * fd->spp;
    if (paths >= 0xffffffffull) return fail(REX_UNSUPPORTED, "more than 2^32 paths in one job");
    rex::Job j;
    j.scene = sceneIndex;
    rex::FrameConst& f = j.f;
    for (int k = 0; k < 3; ++k) {
        f.camPos[k] = fd->camera.position[k];
        f.camFwd[k] = fd->camera.forward[k];
        f.camRight[k] = fd->camera.right[k];
        f.camUp[k] = fd->camera.up[k];
    }
    f.tanHalfFov = fd->camera.tanHalfFov;
    f.aspect = fd->camera.aspect;
    f.width = fd->width;
    f.height = fd->height;
    f.sampleIndex = fd->sampleIndex;
    f.maxBounce = fd->maxBounce;
    f.spp = fd->spp;
    f.seedLow = static_cast<uint32_t>(fd->seed);
    f.seedHigh = static_cast<uint32_t>(fd->seed >> 32);
    f.streamTag = fd->streamTag != 0u ? fd->streamTag : 0x57a20001u;
    f.estimator = fd->estimator;
    const bool defaults = fd->rrStart == 0.0f && fd->rrMinimum == 0.0f && fd->rrMaximum == 0.0f;
    f.rrStart = defaults ? 3.0f : fd->rrStart;
    f.rrMinimum = defaults ? 0.05f : fd->rrMinimum;
    f.rrMaximum = defaults ? 0.95f : fd->rrMaximum;
    f.rayEpsilon = fd->rayEpsilon != 0.0f ? fd->rayEpsilon : 1.0e-4f;
    if (!(f.rayEpsilon > 0.0f)) return fail(REX_INVALID_ARGUMENT, "ray epsilon must be positive");
    f.pathCount = static_cast<uint32_t>(paths);
    j.pathCount = static_cast<uint32_t>(paths);
    j.outBytes = rex::RangeAllocator::round(pixels * sizeof(float4)) + paths * sizeof(float4);
    if (j.outBytes > b.outputs.ranges.total()) return fail(REX_UNSUPPORTED, "job outputs exceed the output arena");
    j.submitStep = b.step;
    const uint64_t id = b.nextJob++;
    b.jobs.emplace(id, j);
    b.queue.push_back(id);
    *jobId = id;
    return REX_SUCCESS;
}

rex_status rex_batch_step(rex_batch batch, void* stream) {
    if (batch == nullptr) return fail(REX_INVALID_ARGUMENT, "batch is NULL");
    if (batch->b.scenes.empty()) return fail(REX_MISSING_SCENE, "no resident scene");
    try {
        rex::Batch& b = batch->b;
        const hipStream_t st = static_cast<hipStream_t>(stream);
        if (rex_status s = b.tune_before(st)) return s;
        if (rex_status s = b.step_once(st)) return s;
        return b.tune_after(st);
    } catch (const std::exception& e) {
        return fail(REX_RECORDING_FAILED, std::string("internal: ") + e.what());
    }
}

rex_status rex_batch_run(rex_batch batch, void* stream) {
    if (batch == nullptr) return fail(REX_INVALID_ARGUMENT, "batch is NULL");
    rex::Batch& b = batch->b;
    while (!b.queue.empty() || b.running != 0) {
        if (rex_status st = rex_batch_step(batch, stream)) return st;
        if (b.running == 0 && !b.queue.empty()) {
            (void)hipStreamSynchronize(static_cast<hipStream_t>(stream));
            return fail(REX_INVALID_ARGUMENT, "output arena full: release finished jobs before running the rest");
        }
    }
    return check_hip(hipStreamSynchronize(static_cast<hipStream_t>(stream)), "batch sync");
}

rex_status rex_batch_job(rex_batch batch, uint64_t jobId, rex_batch_job_info* info) {
    if (batch == nullptr || info == nullptr) return fail(REX_INVALID_ARGUMENT, "NULL argument");
    auto it = batch->b.jobs.find(jobId);
    if (it == batch->b.jobs.end()) return fail(REX_INVALID_ARGUMENT, "unknown job");
    rex::Job& j = it->second;
    if (j.state == 2 && j.done != nullptr && hipEventQuery(j.done) == hipSuccess) j.state = 3;
    std::memset(info, 0, sizeof(*info));
    info->state = j.state;
    info->width = j.f.width;
    info->height = j.f.height;
    info->submitStep = j.submitStep;
    info->finishStep = j.finishStep;
    info->film = j.film;
    return REX_SUCCESS;
}

rex_status rex_batch_release(rex_batch batch, uint64_t jobId) {
    if (batch == nullptr) return fail(REX_INVALID_ARGUMENT, "batch is NULL");
    auto it = batch->b.jobs.find(jobId);
    if (it == batch->b.jobs.end()) return fail(REX_INVALID_ARGUMENT, "unknown job");
    rex::Job& j = it->second;
    if (j.state < 2) return fail(REX_INVALID_ARGUMENT, "job has not finished");
    if (j.done) {
        if (rex_status st = check_hip(hipEventSynchronize(j.done), "job sync")) return st;
        (void)hipEventDestroy(j.done);
    }
    batch->b.outputs.release(j.film);
    batch->b.jobs.erase(it);
    return REX_SUCCESS;
}

rex_status rex_batch_get_stats(rex_batch batch, rex_batch_stats* stats) {
    if (batch == nullptr || stats == nullptr) return fail(REX_INVALID_ARGUMENT, "NULL argument");
    const rex::Batch& b = batch->b;
    std::memset(stats, 0, sizeof(*stats));
    stats->steps = b.launchedSteps;
    stats->jobsFinished = b.finished;
    uint64_t inFlight = 0;
    for (const auto& kv : b.jobs) {
        if (kv.second.state == 1) inFlight += kv.second.pathCount - kv.second.retired;
    }
    stats->pathsInFlight = inFlight;
    stats->queuedJobs = b.queue.size();
    stats->shadeBlocksPerCu = b.shadeBpc;
    stats->traceBlocksPerCu = 0;
    stats->graphs = b.useGraphs ? 1u : 0u;
    stats->autotune = b.tuning ? 1u : (b.tuned ? 2u : 0u);
    stats->pathBudget = b.capacity - b.reserve;
    stats->outputBytesInUse = b.outputs.ranges.in_use();
    stats->outputBytesTotal = b.outputs.ranges.total();
    uint32_t words[rex::kCtlWords];
    if (rex_status st = check_hip(hipDeviceSynchronize(), "stats sync")) return st;
    if (rex_status st = check_hip(hipMemcpy(words, b.ctl, sizeof(words), hipMemcpyDeviceToHost), "stats read")) {
        return st;
    }
    stats->raysTraced = (uint64_t)words[rex::kCtlRaysLo] | ((uint64_t)words[rex::kCtlRaysHi] << 32);
    return REX_SUCCESS;
}

}  // extern "C"
