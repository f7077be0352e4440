// Minimal host-only stand-in for the HIP runtime API, used ONLY to compile
// libsipg_hip's host orchestration (abi.cpp, solvers.cpp, batch.cpp,
// serve.cpp, basis.cpp) for a CPU test build: "device" memory is host
// memory.  Test infrastructure; never part of the product build.
//
// Two execution models.  Immediate (the default): every call takes effect
// when it is made.  Deferred (mock_deferred() = true, the batch and server
// drivers): asynchronous work -- launches made through mock_run,
// hipMemcpyAsync, hipMemsetAsync, hipFreeAsync and graph replays -- is queued
// in issue order and runs only when the host synchronises:
// hipStreamSynchronize, hipDeviceSynchronize, hipStreamDestroy, a synchronous
// hipMemcpy / hipMemset, hipFree and hipHostFree drain the queue, and
// hipEventSynchronize drains it up to that event's record.  A host buffer
// handed to an async copy is read when the copy runs (pinned-memory
// semantics), and a replay reads its graph when it runs, so a host buffer
// released or rewritten before its copy, a graph destroyed with replays
// pending, or host code reading a device-to-host copy before waiting for it
// shows up under AddressSanitizer or in the results.
#pragma once
#include <climits>
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <deque>
#include <functional>
#include <vector>

typedef int hipError_t;
enum { hipSuccess = 0, hipErrorOutOfMemory = 2, hipErrorInvalidValue = 1 };
typedef struct mock_stream *hipStream_t;
struct int2 { int x, y; };
struct mock_event { long seq; };
typedef struct mock_event *hipEvent_t;
enum hipMemcpyKind { hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice, hipMemcpyDefault };
#define hipEventDisableTiming 2

// ---- the queue of deferred work and stream capture ----------------------
struct mock_item {
    long seq;
    std::function<void()> fn;
};
inline bool &mock_deferred() {
    static bool on = false;
    return on;
}
inline std::deque<mock_item> &mock_queue() {
    static std::deque<mock_item> q;
    return q;
}
inline long &mock_seq() {
    static long s = 0;
    return s;
}
inline std::vector<std::function<void()>> *&mock_capture() {
    static std::vector<std::function<void()>> *cap = nullptr;
    return cap;
}
// runs the queued work issued up to and including sequence number `upto`
inline void mock_drain(long upto = LONG_MAX) {
    auto &q = mock_queue();
    while (!q.empty() && q.front().seq <= upto) {
        std::function<void()> f = std::move(q.front().fn);
        q.pop_front();
        f();
    }
}
inline size_t mock_pending() { return mock_queue().size(); }
// asynchronous work: recorded while a stream is captured, else queued
// (deferred model) or run now
inline void mock_run(std::function<void()> f) {
    if (mock_capture()) mock_capture()->push_back(std::move(f));
    else if (mock_deferred()) mock_queue().push_back(mock_item{++mock_seq(), std::move(f)});
    else f();
}

inline const char *hipGetErrorString(hipError_t) { return "mock hip error"; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipGetDeviceCount(int *c) { *c = 1; return hipSuccess; }
inline hipError_t hipSetDevice(int) { return hipSuccess; }
inline hipError_t hipGetDevice(int *d) { *d = 0; return hipSuccess; }
struct hipDeviceProp_t { char gcnArchName[256]; };
inline hipError_t hipGetDeviceProperties(hipDeviceProp_t *p, int) {
    std::strcpy(p->gcnArchName, "mock");
    return hipSuccess;
}
inline hipError_t hipDeviceSynchronize() { mock_drain(); return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t) { mock_drain(); return hipSuccess; }
template <typename T> inline hipError_t hipMalloc(T **p, size_t n) {
    *p = (T *)std::calloc(n ? n : 1, 1);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
inline hipError_t hipMalloc(void **p, size_t n) { *p = std::calloc(n ? n : 1, 1); return *p ? hipSuccess : hipErrorOutOfMemory; }
template <typename T> inline hipError_t hipHostMalloc(T **p, size_t n) { return hipMalloc(p, n); }
template <typename T> inline hipError_t hipMallocAsync(T **p, size_t n, hipStream_t) { return hipMalloc(p, n); }
inline hipError_t hipFree(void *p) { mock_drain(); std::free(p); return hipSuccess; }
inline hipError_t hipFreeAsync(void *p, hipStream_t) { mock_run([p]() { std::free(p); }); return hipSuccess; }
inline hipError_t hipHostFree(void *p) { mock_drain(); std::free(p); return hipSuccess; }
inline hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind) {
    mock_drain();
    std::memmove(d, s, n);
    return hipSuccess;
}
inline hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind, hipStream_t) {
    mock_run([d, s, n]() { std::memmove(d, s, n); });
    return hipSuccess;
}
inline hipError_t hipMemset(void *d, int v, size_t n) { mock_drain(); std::memset(d, v, n); return hipSuccess; }
inline hipError_t hipMemsetAsync(void *d, int v, size_t n, hipStream_t) {
    mock_run([d, v, n]() { std::memset(d, v, n); });
    return hipSuccess;
}
inline hipError_t hipEventCreate(hipEvent_t *e) { *e = new mock_event{0}; return hipSuccess; }
inline hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { return hipEventCreate(e); }
inline hipError_t hipEventDestroy(hipEvent_t e) { delete e; return hipSuccess; }
// the event marks the work issued so far
inline hipError_t hipEventRecord(hipEvent_t e, hipStream_t) {
    if (mock_capture()) return hipErrorInvalidValue;     // not used inside a capture
    e->seq = mock_seq();
    return hipSuccess;
}
inline hipError_t hipEventSynchronize(hipEvent_t e) { mock_drain(e->seq); return hipSuccess; }
inline hipError_t hipEventElapsedTime(float *ms, hipEvent_t, hipEvent_t) { *ms = 0.f; return hipSuccess; }

// ---- streams, pinned memory and graphs ----------------------------------
// A stream is a tag (one queue serves all streams: issue order is kept
// across streams); "capture" records the work, and a graph replays it.
#define hipStreamNonBlocking 1
#define hipHostMallocDefault 0
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal, hipStreamCaptureModeThreadLocal, hipStreamCaptureModeRelaxed };
typedef struct mock_graph *hipGraph_t;
typedef struct mock_graph *hipGraphExec_t;
typedef struct mock_node *hipGraphNode_t;
struct mock_graph { std::vector<std::function<void()>> work; };
inline hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned) { *s = (hipStream_t)std::malloc(8); return hipSuccess; }
inline hipError_t hipStreamDestroy(hipStream_t s) { mock_drain(); std::free(s); return hipSuccess; }
template <typename T> inline hipError_t hipHostMalloc(T **p, size_t n, unsigned) { return hipMalloc(p, n); }
inline hipError_t hipStreamBeginCapture(hipStream_t, hipStreamCaptureMode) {
    if (mock_capture()) return hipErrorInvalidValue;
    mock_capture() = new std::vector<std::function<void()>>();
    return hipSuccess;
}
inline hipError_t hipStreamEndCapture(hipStream_t, hipGraph_t *g) {
    if (!mock_capture()) return hipErrorInvalidValue;
    *g = new mock_graph{std::move(*mock_capture())};
    delete mock_capture();
    mock_capture() = nullptr;
    return hipSuccess;
}
inline hipError_t hipGraphInstantiate(hipGraphExec_t *e, hipGraph_t g, hipGraphNode_t *, char *, size_t) {
    *e = new mock_graph{g->work};
    return hipSuccess;
}
// each node of the replay runs from the executable graph when it is reached
inline hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t) {
    for (size_t k = 0; k < e->work.size(); ++k) mock_run([e, k]() { e->work[k](); });
    return hipSuccess;
}
inline hipError_t hipGraphDestroy(hipGraph_t g) { delete g; return hipSuccess; }
inline hipError_t hipGraphExecDestroy(hipGraphExec_t e) { delete e; return hipSuccess; }
