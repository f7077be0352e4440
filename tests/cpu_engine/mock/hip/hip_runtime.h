// Minimal host-only stand-in for the HIP runtime API, used ONLY to compile
// libsipg_hip's host orchestration (abi.cpp, solvers.cpp, basis.cpp) for a
// CPU test build: "device" memory is host memory, streams are synchronous.
// Test infrastructure; never part of the product build.
#pragma once
#include <cstddef>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <vector>

typedef int hipError_t;
enum { hipSuccess = 0, hipErrorOutOfMemory = 2, hipErrorInvalidValue = 1 };
typedef struct mock_stream *hipStream_t;
struct int2 { int x, y; };
typedef struct mock_event *hipEvent_t;
enum hipMemcpyKind { hipMemcpyHostToDevice, hipMemcpyDeviceToHost, hipMemcpyDeviceToDevice, hipMemcpyDefault };
#define hipEventDisableTiming 2

inline const char *hipGetErrorString(hipError_t) { return "mock hip error"; }
inline hipError_t hipGetLastError() { return hipSuccess; }
inline hipError_t hipGetDeviceCount(int *c) { *c = 1; return hipSuccess; }
inline hipError_t hipSetDevice(int) { return hipSuccess; }
inline hipError_t hipDeviceSynchronize() { return hipSuccess; }
inline hipError_t hipStreamSynchronize(hipStream_t) { return hipSuccess; }
template <typename T> inline hipError_t hipMalloc(T **p, size_t n) {
    *p = (T *)std::calloc(n ? n : 1, 1);
    return *p ? hipSuccess : hipErrorOutOfMemory;
}
inline hipError_t hipMalloc(void **p, size_t n) { *p = std::calloc(n ? n : 1, 1); return *p ? hipSuccess : hipErrorOutOfMemory; }
template <typename T> inline hipError_t hipHostMalloc(T **p, size_t n) { return hipMalloc(p, n); }
template <typename T> inline hipError_t hipMallocAsync(T **p, size_t n, hipStream_t) { return hipMalloc(p, n); }
inline hipError_t hipFree(void *p) { std::free(p); return hipSuccess; }
inline hipError_t hipFreeAsync(void *p, hipStream_t) { std::free(p); return hipSuccess; }
inline hipError_t hipHostFree(void *p) { std::free(p); return hipSuccess; }
inline hipError_t hipMemcpy(void *d, const void *s, size_t n, hipMemcpyKind) { std::memmove(d, s, n); return hipSuccess; }
inline hipError_t hipMemcpyAsync(void *d, const void *s, size_t n, hipMemcpyKind, hipStream_t) { std::memmove(d, s, n); return hipSuccess; }
inline hipError_t hipMemset(void *d, int v, size_t n) { std::memset(d, v, n); return hipSuccess; }
inline hipError_t hipMemsetAsync(void *d, int v, size_t n, hipStream_t) { std::memset(d, v, n); return hipSuccess; }
inline hipError_t hipEventCreate(hipEvent_t *e) { *e = (hipEvent_t)std::malloc(8); return hipSuccess; }
inline hipError_t hipEventCreateWithFlags(hipEvent_t *e, unsigned) { return hipEventCreate(e); }
inline hipError_t hipEventDestroy(hipEvent_t e) { std::free(e); return hipSuccess; }
inline hipError_t hipEventRecord(hipEvent_t, hipStream_t) { return hipSuccess; }
inline hipError_t hipEventSynchronize(hipEvent_t) { return hipSuccess; }
inline hipError_t hipEventElapsedTime(float *ms, hipEvent_t, hipEvent_t) { *ms = 0.f; return hipSuccess; }

// streams, pinned memory and graphs: a stream is a tag; "capture" records
// the work the mock launchers would run, and a graph replays it
#define hipStreamNonBlocking 1
#define hipHostMallocDefault 0
enum hipStreamCaptureMode { hipStreamCaptureModeGlobal, hipStreamCaptureModeThreadLocal, hipStreamCaptureModeRelaxed };
typedef struct mock_graph *hipGraph_t;
typedef struct mock_graph *hipGraphExec_t;
typedef struct mock_node *hipGraphNode_t;
struct mock_graph { std::vector<std::function<void()>> work; };
inline std::vector<std::function<void()>> *&mock_capture() {
    static std::vector<std::function<void()>> *cap = nullptr;
    return cap;
}
// runs `f` now, or records it while a stream is being captured
inline void mock_run(std::function<void()> f) {
    if (mock_capture()) mock_capture()->push_back(std::move(f));
    else f();
}
inline hipError_t hipStreamCreateWithFlags(hipStream_t *s, unsigned) { *s = (hipStream_t)std::malloc(8); return hipSuccess; }
inline hipError_t hipStreamDestroy(hipStream_t s) { std::free(s); return hipSuccess; }
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
inline hipError_t hipGraphLaunch(hipGraphExec_t e, hipStream_t) {
    for (auto &f : e->work) f();
    return hipSuccess;
}
inline hipError_t hipGraphDestroy(hipGraph_t g) { delete g; return hipSuccess; }
inline hipError_t hipGraphExecDestroy(hipGraphExec_t e) { delete e; return hipSuccess; }
