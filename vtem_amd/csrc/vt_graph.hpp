// Loops of identical iterations replayed as a HIP graph.
//
// A solver loop that enqueues k kernels per iteration costs the host one launch call per
// kernel.  vt_graph_loop captures `chunk` iterations once, on a private stream of the
// current device, instantiates the graph and replays it on the caller's stream: one
// host call per replay.  Iterations must be identical (same kernels, same arguments) and
// enqueue kernels only: no memsets (memset nodes have given wrong results on replays),
// copies, allocations or synchronisation.
//
// An executable graph stays alive until an event recorded after its last replay has
// completed; the next graph loop destroys the finished ones.  Any number of host threads
// may call it.
#pragma once
#include "vt_common.hpp"
#include <mutex>
#include <vector>

struct VtGraphRetire {
    hipGraphExec_t exec;
    hipEvent_t done;
};

inline std::mutex &vt_graph_mutex()
{
    static std::mutex m;
    return m;
}

inline std::vector<VtGraphRetire> &vt_graph_pending()
{
    static std::vector<VtGraphRetire> v;
    return v;
}

// destroy the executable graphs whose last replay has completed
inline void vt_graph_sweep()
{
    std::lock_guard<std::mutex> lk(vt_graph_mutex());
    std::vector<VtGraphRetire> &v = vt_graph_pending();
    size_t k = 0;
    for (size_t i = 0; i < v.size(); ++i) {
        if (hipEventQuery(v[i].done) == hipSuccess) {
            (void)hipGraphExecDestroy(v[i].exec);
            (void)hipEventDestroy(v[i].done);
        } else {
            v[k++] = v[i];
        }
    }
    v.resize(k);
}

// this thread's capture stream on the current device (created once, default priority)
inline hipStream_t vt_capture_stream(hipError_t *err)
{
    constexpr int MAXDEV = 64;
    thread_local hipStream_t cap[MAXDEV] = {};
    int dev = 0;
    *err = hipGetDevice(&dev);
    if (*err != hipSuccess) return nullptr;
    if (dev < 0 || dev >= MAXDEV) {
        *err = hipErrorInvalidDevice;
        return nullptr;
    }
    if (!cap[dev]) {
        *err = hipStreamCreateWithFlags(&cap[dev], hipStreamNonBlocking);
        if (*err != hipSuccess) {
            cap[dev] = nullptr;
            return nullptr;
        }
    }
    return cap[dev];
}

// body(stream) enqueues one iteration on `stream` and returns a status (VT_OK = 0).
// Runs `total` iterations on s: with chunk >= 2, total/chunk replays of a captured chunk
// of at most total/2 iterations and the remainder launched directly; chunk 0 or a chunk
// that would hold fewer than 2 iterations: all directly.
template <class Body>
int vt_graph_loop(hipStream_t s, int total, int chunk, Body &&body)
{
    if (chunk >= 2 && total < 2 * chunk) chunk = total / 2;
    if (chunk < 2) {
        for (int i = 0; i < total; ++i) {
            const int st = body(s);
            if (st != VT_OK) return st;
        }
        return VT_OK;
    }
    vt_graph_sweep();
    hipError_t e = hipSuccess;
    hipStream_t cap = vt_capture_stream(&e);
    if (!cap) return (int)e;
    e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return (int)e;
    int st = VT_OK;
    for (int i = 0; i < chunk && st == VT_OK; ++i) st = body(cap);
    hipGraph_t graph = nullptr;
    e = hipStreamEndCapture(cap, &graph);
    if (st != VT_OK || e != hipSuccess) {
        if (graph) (void)hipGraphDestroy(graph);
        return st != VT_OK ? st : (int)e;
    }
    hipGraphExec_t exec = nullptr;
    e = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
    (void)hipGraphDestroy(graph);
    if (e != hipSuccess) return (int)e;
    const int reps = total / chunk;
    for (int i = 0; i < reps && e == hipSuccess; ++i) e = hipGraphLaunch(exec, s);
    // retire the executable graph once its last replay is done
    hipEvent_t done = nullptr;
    hipError_t er = hipEventCreateWithFlags(&done, hipEventDisableTiming);
    if (er == hipSuccess) er = hipEventRecord(done, s);
    if (er == hipSuccess) {
        std::lock_guard<std::mutex> lk(vt_graph_mutex());
        vt_graph_pending().push_back({exec, done});
    } else {
        (void)hipStreamSynchronize(s);
        (void)hipGraphExecDestroy(exec);
        if (done) (void)hipEventDestroy(done);
    }
    if (e != hipSuccess) return (int)e;
    if (er != hipSuccess) return (int)er;
    for (int i = reps * chunk; i < total; ++i) {
        st = body(s);
        if (st != VT_OK) return st;
    }
    return VT_OK;
}
