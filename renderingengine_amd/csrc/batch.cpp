// Batch engine: many render jobs through one shared set of wavefront
// launches per step (continuous batching).
//
// Device state:
//   arena      the resident scenes' 4-wide nodes (references relocated),
//              triangles, triangle attributes and emitter map, one array
//              each; SceneRoot per scene (root reference + box)
//   paths      PathState[capacity]: a job holds a contiguous slot range
//   queues     rays[2][capacity] + rayPath + raySceneOf, hits, shadow items
//              + shadowSceneOf (one ray and at most one shadow query per path)
//   jobs       BatchJob[maxJobs]: frame constants, shading view, slot range,
//              Film of each admitted job (unit)
//   lists      admitted / resolved job slots with their prefix sums
//
// Step s (cur = s & 1), all on one stream, no host readback:
//   RayGen of the units admitted at s -> appended to rays[cur]
//   traversal: shadow queries of step s-1's Shade + rays[cur] (one launch)
//   Resolve + Film of the units admitted at s - maxBounce (their last
//   Shade ran at step s-1, its shadow queries just now)
//   zero the counters of rays[cur ^ 1] and of the shadow queue
//   Shade of rays[cur] -> rays[cur ^ 1] and shadow queries
// Every path advances one bounce per step, so a unit admitted at step a
// finishes at step a + maxBounce without the device reporting it.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <deque>
#include <map>
#include <new>
#include <string>
#include <vector>

#include "rex_internal.hpp"

namespace rex {

hipError_t launch_relocate_wide4(Wide4* dst, const Wide4* src, uint32_t n, uint32_t nodeBase, uint32_t triBase,
                                 hipStream_t stream);
hipError_t launch_fill_u32(uint32_t* dst, uint32_t n, uint32_t value, hipStream_t stream);

namespace {

constexpr uint32_t kStaging = 4;  // host staging buffers in flight

struct ArenaScene {
    Scene* src = nullptr;
    uint32_t w4Base = 0, w4Count = 0, triBase = 0, triCount = 0;
    SceneRoot root{};
    ShadingView sh{};  // the scene's tables; emitter map indexed by arena triangle
};

// One admitted (or queued) sample range of a job.
struct Unit {
    uint64_t job = 0;
    uint32_t s0 = 0, s1 = 0;  // sample range [s0, s1) of the job's spp
    uint32_t slot = kInvalid;
    uint32_t pathBase = 0, paths = 0;
    uint64_t admitStep = 0;
};

struct Job {
    uint32_t scene = 0;
    rex_frame_desc frame{};
    FrameConst f{};
    float4* film = nullptr;
    uint32_t state = 0;  // 0 queued, 1 running, 2 resolve issued, 3 finished
    uint32_t unitsLeft = 0;
    uint32_t nextUnit = 0;       // next sample range to admit
    uint32_t unitSpp = 0;        // samples per unit
    bool running = false;        // a unit of this job is admitted and unresolved
    uint64_t submitStep = 0, finishStep = 0;
    hipEvent_t done = nullptr;  // recorded after the step that resolves the last range
};

// First-fit allocator of path-slot ranges.
struct RangeAlloc {
    std::map<uint32_t, uint32_t> free;  // start -> length
    void reset(uint32_t n) {
        free.clear();
        if (n) free[0] = n;
    }
    bool take(uint32_t n, uint32_t& start) {
        for (auto it = free.begin(); it != free.end(); ++it) {
            if (it->second >= n) {
                start = it->first;
                const uint32_t rest = it->second - n;
                free.erase(it);
                if (rest) free[start + n] = rest;
                return true;
            }
        }
        return false;
    }
    void give(uint32_t start, uint32_t n) {
        auto next = free.lower_bound(start);
        if (next != free.end() && start + n == next->first) {
            n += next->second;
            free.erase(next);
        }
        auto it = free.lower_bound(start);
        if (it != free.begin()) {
            auto prev = std::prev(it);
            if (prev->first + prev->second == start) {
                prev->second += n;
                return;
            }
        }
        free[start] = n;
    }
};

struct Staging {
    // Host pinned: job records of this step's admissions, the admit and
    // resolve lists, copied to the device at the start of the step.
    BatchJob* jobs = nullptr;      // maxJobs
    uint32_t* jobSlots = nullptr;  // maxJobs: device slot of jobs[i]
    uint32_t* lists = nullptr;     // admitCount, admitJob[maxJobs], admitStart[maxJobs + 1], resolve... (see offsets)
    hipEvent_t done = nullptr;
    bool pending = false;
};

}  // namespace

struct Batch {
    uint32_t capacity = 0, maxJobs = 0;
    bool useGraphs = false;
    uint32_t shadeBpc = 0, traceBpc = 0;
    int cus = 1;
    // arena
    Wide4* w4 = nullptr;
    TriPos* tris = nullptr;
    TriAttr* attrs = nullptr;
    uint32_t* emitter = nullptr;
    uint32_t w4Cap = 0, triCap = 0, w4Count = 0, triCount = 0, depth = 0;
    SceneRoot* dRoots = nullptr;
    uint32_t rootsCap = 0;
    std::vector<ArenaScene> scenes;
    SceneView view{};
    // paths and queues
    PathState* paths = nullptr;
    F4* rays[2] = {nullptr, nullptr};
    uint32_t* rayPath[2] = {nullptr, nullptr};
    uint32_t* raySceneOf[2] = {nullptr, nullptr};
    QueueHit* hits = nullptr;
    ShadowItem* shadows = nullptr;
    uint32_t* shadowSceneOf = nullptr;
    // counters: [0], [1] ray queues, [2] shadow queue
    uint32_t* counts = nullptr;
    BatchJob* dJobs = nullptr;
    uint32_t* dLists = nullptr;  // admitCount, admitJob, admitStart, resolveCount, resolveJob, resolveStart
    TraceWorkspace ws;
    Staging staging[kStaging];
    uint32_t stagingNext = 0;
    hipGraphExec_t graphs[2] = {nullptr, nullptr};
    hipStream_t graphStream = nullptr;
    // host bookkeeping
    std::map<uint64_t, Job> jobs;
    std::deque<uint64_t> queue;       // jobs with units left to admit
    std::vector<Unit> active;         // admitted units
    std::vector<uint32_t> freeSlots;  // device job slots
    RangeAlloc alloc;
    uint64_t nextJob = 1, step = 0, launchedSteps = 0, finished = 0, pathsHeld = 0;
    uint32_t cur = 0;

    // list offsets inside dLists / staging lists (words)
    uint32_t offAdmitCount() const { return 0; }
    uint32_t offAdmitJob() const { return 1; }
    uint32_t offAdmitStart() const { return 1 + maxJobs; }
    uint32_t offResolveCount() const { return 2 + 2 * maxJobs; }
    uint32_t offResolveJob() const { return 3 + 2 * maxJobs; }
    uint32_t offResolveStart() const { return 3 + 3 * maxJobs; }
    uint32_t listWords() const { return 4 + 4 * maxJobs; }

    void release();
    rex_status grow_arena(uint32_t needW4, uint32_t needTri);
    rex_status add_scene(Scene& s, uint32_t& index);
    rex_status record_step(hipStream_t st, uint32_t parity);
    rex_status step_once(hipStream_t st);
};

void Batch::release() {
    for (hipGraphExec_t& g : graphs) {
        if (g) (void)hipGraphExecDestroy(g);
        g = nullptr;
    }
    if (graphStream) (void)hipStreamDestroy(graphStream);
    graphStream = nullptr;
    for (Staging& s : staging) {
        if (s.done) (void)hipEventDestroy(s.done);
        if (s.jobs) (void)hipHostFree(s.jobs);
        if (s.jobSlots) (void)hipHostFree(s.jobSlots);
        if (s.lists) (void)hipHostFree(s.lists);
        s = Staging{};
    }
    for (auto& kv : jobs) {
        if (kv.second.film) (void)hipFree(kv.second.film);
        if (kv.second.done) (void)hipEventDestroy(kv.second.done);
    }
    jobs.clear();
    void* ptrs[] = {w4, tris, attrs, emitter, dRoots, paths, rays[0], rays[1], rayPath[0], rayPath[1],
                    raySceneOf[0], raySceneOf[1], hits, shadows, shadowSceneOf, counts, dJobs, dLists};
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
    Wide4* nw = nullptr;
    TriPos* nt = nullptr;
    TriAttr* na = nullptr;
    uint32_t* ne = nullptr;
    if (rex_status st = check_hip(hipMalloc(&nw, (size_t)newW4 * sizeof(Wide4)), "hipMalloc arena")) return st;
    if (rex_status st = check_hip(hipMalloc(&nt, (size_t)newTri * sizeof(TriPos)), "hipMalloc arena")) {
        (void)hipFree(nw);
        return st;
    }
    if (rex_status st = check_hip(hipMalloc(&na, (size_t)newTri * sizeof(TriAttr)), "hipMalloc arena")) {
        (void)hipFree(nw);
        (void)hipFree(nt);
        return st;
    }
    if (rex_status st = check_hip(hipMalloc(&ne, (size_t)newTri * sizeof(uint32_t)), "hipMalloc arena")) {
        (void)hipFree(nw);
        (void)hipFree(nt);
        (void)hipFree(na);
        return st;
    }
    if (w4Count) (void)hipMemcpy(nw, w4, (size_t)w4Count * sizeof(Wide4), hipMemcpyDeviceToDevice);
    if (triCount) {
        (void)hipMemcpy(nt, tris, (size_t)triCount * sizeof(TriPos), hipMemcpyDeviceToDevice);
        (void)hipMemcpy(na, attrs, (size_t)triCount * sizeof(TriAttr), hipMemcpyDeviceToDevice);
        (void)hipMemcpy(ne, emitter, (size_t)triCount * sizeof(uint32_t), hipMemcpyDeviceToDevice);
    }
    (void)hipFree(w4);
    (void)hipFree(tris);
    (void)hipFree(attrs);
    (void)hipFree(emitter);
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
    if (s.triCount == 0 || s.wide4Count == 0) return fail(REX_INVALID_ARGUMENT, "empty scene");
    if (!active.empty()) return fail(REX_INVALID_ARGUMENT, "scenes are added while no job is in flight");
    if ((uint64_t)triCount + s.triCount > kLeafFirstMask) return fail(REX_UNSUPPORTED, "arena holds 2^28 triangles");
    if (rex_status st = grow_arena(w4Count + s.wide4Count, triCount + s.triCount)) return st;
    ArenaScene a;
    a.src = &s;
    a.w4Base = w4Count;
    a.w4Count = s.wide4Count;
    a.triBase = triCount;
    a.triCount = s.triCount;
    const uint32_t r = s.view.rootRef;
    a.root.rootRef = is_leaf_ref(r) ? ((r & ~kLeafFirstMask) | ((r & kLeafFirstMask) + a.triBase)) : r + a.w4Base;
    a.sh = s.shade.view();
    for (int k = 0; k < 3; ++k) {
        a.root.rootMin[k] = s.view.rootMin[k];
        a.root.rootMax[k] = s.view.rootMax[k];
    }
    if (rex_status st = check_hip(launch_relocate_wide4(w4 + a.w4Base, s.dWide4, s.wide4Count, a.w4Base, a.triBase,
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
    if (rex_status st = check_hip(hipMemcpy(emitter + a.triBase, s.shade.dEmitterOfTri,
                                            (size_t)s.triCount * sizeof(uint32_t), hipMemcpyDeviceToDevice),
                                  "arena copy")) {
        return st;
    }
    w4Count += s.wide4Count;
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
    view.wide4 = w4;
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
    uint32_t* lists = dLists;
    BatchRaygenArgs rg;
    rg.jobs = dJobs;
    rg.admitJob = lists + offAdmitJob();
    rg.admitStart = lists + offAdmitStart();
    rg.admitCount = lists + offAdmitCount();
    rg.paths = paths;
    rg.rays = rays[c];
    rg.rayPath = rayPath[c];
    rg.raySceneOf = raySceneOf[c];
    rg.rayCount = counts + c;
    const uint32_t gridBlocks = static_cast<uint32_t>(cus) * 8u;
    if (rex_status s = check_hip(launch_batch_raygen(rg, gridBlocks, st), "batch raygen")) return s;
    if (rex_status s = check_hip(launch_queue_mixed_arena(view, ws, shadows, counts + 2, shadowSceneOf, rays[c],
                                                          counts + c, raySceneOf[c], dRoots, 2u * capacity, hits,
                                                          paths, st),
                                 "batch traversal")) {
        return s;
    }
    BatchResolveArgs rv;
    rv.jobs = dJobs;
    rv.resolveJob = lists + offResolveJob();
    rv.resolveStart = lists + offResolveStart();
    rv.resolveCount = lists + offResolveCount();
    rv.paths = paths;
    if (rex_status s = check_hip(launch_batch_resolve(rv, gridBlocks, st), "batch resolve")) return s;
    if (rex_status s = check_hip(launch_fill_u32(counts + n, 1u, 0u, st), "counter reset")) return s;
    if (rex_status s = check_hip(launch_fill_u32(counts + 2, 1u, 0u, st), "counter reset")) return s;
    ShadeLaunch L{};
    L.bounce = 0;
    L.capacity = capacity;
    L.s = view;
    L.rays = rays[c];
    L.rayPath = rayPath[c];
    L.count = counts + c;
    L.hits = hits;
    L.paths = paths;
    L.nextRays = rays[n];
    L.nextPath = rayPath[n];
    L.nextCount = counts + n;
    L.shadows = shadows;
    L.shadowCount = counts + 2;
    L.surfaces = nullptr;
    L.jobs = dJobs;
    L.nextSceneOf = raySceneOf[n];
    L.shadowSceneOf = shadowSceneOf;
    L.blocksPerCu = shadeBpc;
    L.traceBlocksPerCu = traceBpc;
    return check_hip(launch_shade(L, st), "batch shade");
}

rex_status Batch::step_once(hipStream_t st) {
    Staging& sg = staging[stagingNext];
    if (sg.pending) {
        if (rex_status s = check_hip(hipEventSynchronize(sg.done), "staging wait")) return s;
        sg.pending = false;
    }
    // Units whose last Shade ran at the previous step are resolved now.
    std::vector<Unit> finishing, keep;
    for (const Unit& u : active) {
        const Job& j = jobs.at(u.job);
        if (step >= u.admitStep + j.f.maxBounce) finishing.push_back(u);
        else keep.push_back(u);
    }
    uint32_t* L = sg.lists;
    L[offResolveCount()] = static_cast<uint32_t>(finishing.size());
    L[offResolveStart()] = 0;
    for (size_t i = 0; i < finishing.size(); ++i) {
        const Job& j = jobs.at(finishing[i].job);
        L[offResolveJob() + i] = finishing[i].slot;
        L[offResolveStart() + i + 1] = L[offResolveStart() + i] + j.f.width * j.f.height;
    }
    // Admissions: one unit per job per step, while slots and paths last.
    uint32_t nAdmit = 0;
    L[offAdmitStart()] = 0;
    std::vector<uint32_t> freedAfter;  // path ranges released after this step's resolve
    std::deque<uint64_t> requeue;
    while (!queue.empty() && nAdmit < maxJobs && !freeSlots.empty()) {
        const uint64_t id = queue.front();
        Job& j = jobs.at(id);
        if (j.running) {  // its previous range has not been resolved yet
            queue.pop_front();
            requeue.push_back(id);
            continue;
        }
        const uint32_t pixels = j.f.width * j.f.height;
        const uint32_t s0 = j.nextUnit * j.unitSpp;
        const uint32_t s1 = std::min(j.frame.spp, s0 + j.unitSpp);
        const uint32_t n = pixels * (s1 - s0);
        uint32_t base = 0;
        if (!alloc.take(n, base)) break;  // wait for capacity (keeps submission order)
        queue.pop_front();
        Unit u;
        u.job = id;
        u.s0 = s0;
        u.s1 = s1;
        u.slot = freeSlots.back();
        freeSlots.pop_back();
        u.pathBase = base;
        u.paths = n;
        u.admitStep = step;
        BatchJob& bj = sg.jobs[nAdmit];
        bj = BatchJob{};
        bj.f = j.f;
        bj.f.spp = s1 - s0;
        bj.f.sampleIndex = j.f.sampleIndex + s0;
        bj.f.pathCount = n;
        bj.sh = scenes[j.scene].sh;
        bj.pathBase = base;
        bj.scene = j.scene;
        bj.pixelCount = pixels;
        bj.film = reinterpret_cast<F4*>(j.film);
        sg.jobSlots[nAdmit] = u.slot;
        L[offAdmitJob() + nAdmit] = u.slot;
        L[offAdmitStart() + nAdmit + 1] = L[offAdmitStart() + nAdmit] + n;
        ++nAdmit;
        if (j.nextUnit == 0) {
            if (rex_status s = check_hip(hipMemsetAsync(j.film, 0, (size_t)pixels * sizeof(float4), st), "film clear"))
                return s;
        }
        j.nextUnit++;
        j.state = 1;
        j.running = true;
        pathsHeld += n;
        keep.push_back(u);
        if (j.nextUnit * j.unitSpp < j.frame.spp) requeue.push_back(id);
    }
    for (uint64_t id : requeue) queue.push_back(id);
    L[offAdmitCount()] = nAdmit;
    if (nAdmit == 0 && finishing.empty() && keep.empty()) return REX_SUCCESS;  // idle
    // Uploads: new job records, then the lists.
    for (uint32_t i = 0; i < nAdmit; ++i) {
        if (rex_status s = check_hip(hipMemcpyAsync(dJobs + sg.jobSlots[i], sg.jobs + i, sizeof(BatchJob),
                                                    hipMemcpyHostToDevice, st),
                                     "job upload")) {
            return s;
        }
    }
    if (rex_status s = check_hip(hipMemcpyAsync(dLists, L, listWords() * sizeof(uint32_t), hipMemcpyHostToDevice, st),
                                 "list upload")) {
        return s;
    }
    const uint32_t parity = cur;
    // The first steps run uncaptured: they size the traversal workspace.
    if (useGraphs && launchedSteps >= 2) {
        if (graphs[parity] == nullptr) {
            // Capture on a side stream ordered after the uploads.
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
    if (rex_status s = check_hip(hipEventRecord(sg.done, st), "step event")) return s;
    sg.pending = true;
    for (const Unit& u : finishing) {
        Job& j = jobs.at(u.job);
        alloc.give(u.pathBase, u.paths);
        pathsHeld -= u.paths;
        freeSlots.push_back(u.slot);
        j.running = false;
        if (--j.unitsLeft == 0) {
            j.state = 2;
            j.finishStep = step;
            if (j.done == nullptr) {
                if (rex_status s = check_hip(hipEventCreateWithFlags(&j.done, hipEventDisableTiming), "job event"))
                    return s;
            }
            if (rex_status s = check_hip(hipEventRecord(j.done, st), "job event")) return s;
            ++finished;
        }
    }
    active.swap(keep);
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

extern "C" {

rex_status rex_batch_create(const rex_batch_desc* desc, rex_batch* out) {
    if (out == nullptr) return fail(REX_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    rex_batch_desc d{};
    if (desc != nullptr) d = *desc;
    const uint32_t capacity = d.pathCapacity != 0u ? d.pathCapacity : (16u << 20);
    const uint32_t maxJobs = d.maxJobs != 0u ? d.maxJobs : 256u;
    if (capacity >= 0x40000000u) return fail(REX_UNSUPPORTED, "path capacity above 2^30");
    if (maxJobs > 65536u) return fail(REX_INVALID_ARGUMENT, "maxJobs above 65536");
    if (d.shadeBlocksPerCu > 64u || d.traceBlocksPerCu > 64u) return fail(REX_INVALID_ARGUMENT, "grid size above 64 blocks per CU");
    rex_batch rb = new (std::nothrow) rex_batch_t();
    if (rb == nullptr) return fail(REX_INVALID_ARGUMENT, "host allocation failed");
    rex::Batch& b = rb->b;
    b.capacity = capacity;
    b.maxJobs = maxJobs;
    b.useGraphs = d.useGraphs != 0u;
    b.shadeBpc = d.shadeBlocksPerCu;
    b.traceBpc = d.traceBlocksPerCu;
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
        {reinterpret_cast<void**>(&b.counts), 64 * sizeof(uint32_t)},
        {reinterpret_cast<void**>(&b.dJobs), (size_t)maxJobs * sizeof(rex::BatchJob)},
        {reinterpret_cast<void**>(&b.dLists), (size_t)b.listWords() * sizeof(uint32_t)},
    };
    for (const Alloc& a : allocs) {
        if (rex_status st = check_hip(hipMalloc(a.p, a.bytes), "hipMalloc batch buffer")) {
            b.release();
            delete rb;
            return st;
        }
    }
    for (rex::Staging& s : b.staging) {
        rex_status st = check_hip(hipHostMalloc(reinterpret_cast<void**>(&s.jobs), maxJobs * sizeof(rex::BatchJob)),
                                  "hipHostMalloc staging");
        if (st == REX_SUCCESS)
            st = check_hip(hipHostMalloc(reinterpret_cast<void**>(&s.jobSlots), maxJobs * sizeof(uint32_t)),
                           "hipHostMalloc staging");
        if (st == REX_SUCCESS)
            st = check_hip(hipHostMalloc(reinterpret_cast<void**>(&s.lists), b.listWords() * sizeof(uint32_t)),
                           "hipHostMalloc staging");
        if (st == REX_SUCCESS) st = check_hip(hipEventCreateWithFlags(&s.done, hipEventDisableTiming), "staging event");
        if (st != REX_SUCCESS) {
            b.release();
            delete rb;
            return st;
        }
    }
    if (rex_status st = check_hip(hipMemset(b.counts, 0, 64 * sizeof(uint32_t)), "counter clear")) {
        b.release();
        delete rb;
        return st;
    }
    b.alloc.reset(capacity);
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
    if (pixels > b.capacity) return fail(REX_INVALID_ARGUMENT, "one sample of the frame exceeds the path capacity");
    if (fd->width > 16384u || fd->height > 16384u) return fail(REX_INVALID_ARGUMENT, "image dimension above 16384");
    if (fd->maxBounce == 0 || fd->maxBounce > 16u) return fail(REX_INVALID_ARGUMENT, "maxBounce must be in 1..16");
    if (fd->estimator > REX_ESTIMATOR_MIS) return fail(REX_UNSUPPORTED, "batched jobs use BSDF-only, NEE or MIS");
    rex::Job j;
    j.scene = sceneIndex;
    j.frame = *fd;
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
    j.unitSpp = static_cast<uint32_t>(std::min<uint64_t>(fd->spp, b.capacity / pixels));
    j.unitsLeft = (fd->spp + j.unitSpp - 1u) / j.unitSpp;
    if (rex_status st = check_hip(hipMalloc(reinterpret_cast<void**>(&j.film), pixels * sizeof(float4)), "hipMalloc film"))
        return st;
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
        return batch->b.step_once(static_cast<hipStream_t>(stream));
    } catch (const std::exception& e) {
        return fail(REX_RECORDING_FAILED, std::string("internal: ") + e.what());
    }
}

rex_status rex_batch_run(rex_batch batch, void* stream) {
    if (batch == nullptr) return fail(REX_INVALID_ARGUMENT, "batch is NULL");
    rex::Batch& b = batch->b;
    while (!b.queue.empty() || !b.active.empty()) {
        if (rex_status st = rex_batch_step(batch, stream)) return st;
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
    if (j.film) (void)hipFree(j.film);
    batch->b.jobs.erase(it);
    return REX_SUCCESS;
}

rex_status rex_batch_get_stats(rex_batch batch, rex_batch_stats* stats) {
    if (batch == nullptr || stats == nullptr) return fail(REX_INVALID_ARGUMENT, "NULL argument");
    const rex::Batch& b = batch->b;
    std::memset(stats, 0, sizeof(*stats));
    stats->steps = b.launchedSteps;
    stats->jobsFinished = b.finished;
    stats->pathsInFlight = b.pathsHeld;
    stats->queuedJobs = b.queue.size();
    stats->shadeBlocksPerCu = b.shadeBpc;
    stats->traceBlocksPerCu = b.traceBpc;
    stats->graphs = b.useGraphs ? 1u : 0u;
    return REX_SUCCESS;
}

}  // extern "C"
