// serve.cpp -- continuous batching of independent SIPG solves (FP64).
//
// Requests (mesh, degree, penalty, Fourier forcing, tolerance) are queued
// by sipg_server_submit and admitted, in arrival order, as soon as device
// memory within the server's HBM budget and a slot are free.  Every
// admitted request's CG state stays on the device; one call of
// sipg_server_progress
//   1. reads the per-request states left by the previous replay,
//   2. confirms each finished recursion on a true residual (or resumes it),
//   3. retires finished requests (their solution stays until taken),
//   4. admits queued requests that now fit (RHS and initial residual are
//      built on the device),
//   5. replays kServeIters CG iterations of every running request from a
//      captured hipGraph (the batched kernels of kernels_batch.hip over
//      device-resident work lists), and waits for it.
// Under SIPG_METHOD_PCG the iterations are block-Jacobi PCG ones (the
// batched preconditioner of kernels_batch_pcg.hip after each x/r update).
// A request's arithmetic does not depend on the other requests in flight:
// its reductions run over its own blocks and chunks in a fixed order, so a
// result is the same alone and in any mix.
//
// Memory: six pools (x, b, r, p, Ap, and z, the increments since the last
// true residual), and under PCG a seventh (M^-1 r), share one offset space;
// a request of n unknowns reserves the same range of n elements in each,
// 48 n (56 n) bytes in all.  The budget bounds the pools together.
#include "sipg_internal.hpp"

#include <algorithm>
#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <deque>
#include <map>
#include <vector>

using namespace sipg_internal;

namespace {

constexpr int kServeIters = 8;       // iterations per replay (even)
constexpr int64_t kChunk = 8192;
constexpr int kMaxSlots = 4096;      // requests on the device at once
constexpr int kPools = 7;            // x, b, r, p, Ap, z (increments), zq (M^-1 r, PCG)
enum { SET_ACTIVE = 0, SET_INIT = 1, SET_REPL = 2, NSETS = 3 };
enum { QUEUED, RUNNING, DONE };
constexpr double kServeReplace = 3e-3;  // true residual at each 300x drop

struct Request {
    int64_t id = 0;
    int dim = 0, p = 0;
    double pen = 0;
    int nmodes = 0;
    double a[64];
    int m[64], n[64];
    double tol = 0;
    int tol_mode = SIPG_TOL_REL, maxit = 0;
    int phase = QUEUED;
    int slot = -1;
    int64_t off = -1, len = 0;
    int nrep_req = 0;                // replacements requested so far
    bool check_req = false;          // the last one requested is a final check
    double check_ref = -1.0;         // smallest true r.r of the failed final checks
    int check_stale = 0;             // failed checks without a 10% gain on it
    int rep_it = 0;                  // iteration count when the last true residual was requested
    int best_it = 0;                 // iteration count at the best failed final check
    double rec_min = -1.0;           // lowest polled recursive r.r since the last true residual
    bool cancel = false;            // cancelled while running: retired at the next progress
    sipg_request_times times{-1.0, -1.0, -1.0};
    sipg_cg_info info{};
};

struct SlotRec {                     // per-slot data uploaded at admission
    batch_inst inst;
    batch_ctl ctl;
};

}  // namespace

struct sipg_server {
    hipStream_t stream = nullptr;
    bool own_stream = false;
    double budget = 0;                       // bytes, all pools
    int method = SIPG_METHOD_CG;
    int npools = 6;                          // 7 under PCG
    int64_t cap = 0;                         // elements per pool
    double *pool[kPools] = {};
    std::map<int64_t, int64_t> freel;        // offset -> length of free ranges
    int64_t used = 0;

    std::vector<int> free_slots;
    std::vector<int64_t> slot_req;
    batch_inst *d_inst = nullptr;
    cg_state *d_state = nullptr;
    batch_ctl *d_ctl = nullptr;
    double *d_modes = nullptr;               // 192 doubles per slot
    cg_state *h_state = nullptr;             // pinned

    int shape[kMaxDegree + 1];
    int item_cap[NSETS][kMaxDegree + 1] = {};
    batch_item *d_items[NSETS][kMaxDegree + 1] = {};
    // per-item partial sums (active set: p.Ap, and r.r and r.z under PCG;
    // admission set: r.r and r.z under PCG), offsets per degree
    int item_base[NSETS][kMaxDegree + 1] = {};
    int item_total[NSETS] = {};
    int chunk_cap = 0;
    batch_chunk *d_chunks[NSETS] = {};
    int *d_list[NSETS] = {};
    int2 *d_item_range[NSETS] = {}, *d_chunk_range[NSETS] = {};
    int *d_counts = nullptr;                 // NSETS x (2 + kMaxDegree + 1)
    double *d_item_part = nullptr, *d_chunk_part = nullptr, *d_part2 = nullptr, *d_part3 = nullptr;
    double *d_item_rr[NSETS] = {}, *d_item_rz[NSETS] = {};
    char *h_stage[NSETS] = {};
    size_t stage_bytes[NSETS] = {};
    hipEvent_t stage_ev[NSETS] = {};
    bool stage_used[NSETS] = {};
    std::map<int, double *> d_tab;

    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<int> graph_degrees;
    bool graph_stale = true;
    bool replayed = false;                   // h_state holds a replay's outcome

    std::map<int64_t, Request> reqs;
    std::deque<int64_t> queue, done_q;
    std::vector<int64_t> active;             // request ids RUNNING
    int64_t next_id = 1;

    std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
    double timeout = 0.0;                    // seconds from submission; 0: none
    double now() const { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); }

    static constexpr int NC = 2 + kMaxDegree + 1;
    int *count(int set, int k) { return d_counts + set * NC + k; }
};

namespace {

template <typename Tp>
int dmalloc(Tp **p, size_t n) {
    if (n == 0) n = 1;
    const hipError_t e = hipMalloc((void **)p, n * sizeof(Tp));
    if (e != hipSuccess) {
        set_last_hip_error(e);
        return SIPG_ERR_NOMEM;
    }
    return SIPG_OK;
}

// ---- offset-space allocator (first fit, coalescing) ------------------------
int64_t range_alloc(sipg_server &S, int64_t n) {
    for (auto it = S.freel.begin(); it != S.freel.end(); ++it) {
        if (it->second < n) continue;
        const int64_t off = it->first, len = it->second;
        S.freel.erase(it);
        if (len > n) S.freel[off + n] = len - n;
        S.used += n;
        return off;
    }
    return -1;
}

void range_free(sipg_server &S, int64_t off, int64_t n) {
    S.used -= n;
    auto nx = S.freel.lower_bound(off);
    if (nx != S.freel.end() && off + n == nx->first) {
        n += nx->second;
        nx = S.freel.erase(nx);
    }
    if (nx != S.freel.begin()) {
        auto pv = std::prev(nx);
        if (pv->first + pv->second == off) {
            pv->second += n;
            return;
        }
    }
    S.freel[off] = n;
}

// grows a set's item buffer of degree p to hold n items
int ensure_items(sipg_server &S, int set, int p, int n) {
    if (n <= S.item_cap[set][p]) return SIPG_OK;
    int c = std::max(1024, S.item_cap[set][p]);
    while (c < n) c *= 2;
    SIPG_HIP_TRY(hipStreamSynchronize(S.stream));
    if (S.d_items[set][p]) (void)hipFree(S.d_items[set][p]);
    S.d_items[set][p] = nullptr;
    S.item_cap[set][p] = 0;
    int rc = dmalloc(&S.d_items[set][p], c);
    if (rc) return rc;
    S.item_cap[set][p] = c;
    if (set == SET_REPL) {
        S.graph_stale = true;                       // the graph holds the buffer
        return SIPG_OK;
    }
    // partial-sum offsets follow the set's capacities
    S.item_total[set] = 0;
    for (int q = 1; q <= kMaxDegree; ++q) {
        S.item_base[set][q] = S.item_total[set];
        S.item_total[set] += S.item_cap[set][q];
    }
    double **bufs[] = {&S.d_item_rr[set], &S.d_item_rz[set], &S.d_item_part};
    for (int k = 0; k < (set == SET_ACTIVE ? 3 : 2); ++k) {
        if (*bufs[k]) (void)hipFree(*bufs[k]);
        *bufs[k] = nullptr;
        if ((rc = dmalloc(bufs[k], S.item_total[set]))) return rc;
    }
    if (set == SET_ACTIVE) S.graph_stale = true;
    return SIPG_OK;
}

int ensure_stage(sipg_server &S, int set, size_t bytes) {
    if (bytes <= S.stage_bytes[set]) return SIPG_OK;
    if (S.stage_used[set]) SIPG_HIP_TRY(hipEventSynchronize(S.stage_ev[set]));
    if (S.h_stage[set]) (void)hipHostFree(S.h_stage[set]);
    S.h_stage[set] = nullptr;
    S.stage_bytes[set] = 0;
    const size_t b = std::max(bytes, (size_t)1 << 20);
    const hipError_t e = hipHostMalloc((void **)&S.h_stage[set], b, hipHostMallocDefault);
    if (e != hipSuccess) {
        set_last_hip_error(e);
        return SIPG_ERR_NOMEM;
    }
    S.stage_bytes[set] = b;
    return SIPG_OK;
}

// Builds and uploads the work lists of a set of requests (asynchronously,
// through the set's pinned staging area).  Extra per-slot records (for
// admissions) are appended to the same staging area.
int upload_set(sipg_server &S, int set, const std::vector<int64_t> &ids,
               const std::vector<std::pair<int, SlotRec>> *recs = nullptr,
               const std::vector<int> *mode_slots = nullptr) {
    std::vector<std::vector<batch_item>> items(kMaxDegree + 1);
    std::vector<batch_chunk> chunks;
    std::vector<int> list;
    std::vector<int2> item_range(kMaxSlots, int2{0, 0}), chunk_range(kMaxSlots, int2{0, 0});
    for (int64_t id : ids) {
        const Request &R = S.reqs.at(id);
        auto &v = items[R.p];
        const int s0 = (int)v.size();
        const int BX = kBatchShapeBX[S.shape[R.p]], BY = kBatchShapeBY[S.shape[R.p]];
        for (int by = 0; by < (R.dim + BY - 1) / BY; ++by)
            for (int bx = 0; bx < (R.dim + BX - 1) / BX; ++bx) v.push_back({R.slot, bx * BX, by * BY, 0});
        item_range[R.slot] = int2{s0, (int)v.size()};   // based below
        list.push_back(R.slot);
        const int c0 = (int)chunks.size();
        for (int64_t a = R.off; a < R.off + R.len; a += kChunk)
            chunks.push_back({R.slot, 0, a, std::min(a + kChunk, R.off + R.len)});
        chunk_range[R.slot] = int2{c0, (int)chunks.size()};
    }
    int rc;
    for (int p = 1; p <= kMaxDegree; ++p)
        if ((rc = ensure_items(S, set, p, (int)items[p].size()))) return rc;
    if (set != SET_REPL)
        for (int64_t id : ids) {
            const Request &R = S.reqs.at(id);
            item_range[R.slot].x += S.item_base[set][R.p];
            item_range[R.slot].y += S.item_base[set][R.p];
        }
    if ((int)chunks.size() > S.chunk_cap) return SIPG_ERR_BUDGET;
    size_t bytes = 64 * (kMaxDegree + 8);
    for (auto &v : items) bytes += v.size() * sizeof(batch_item);
    bytes += chunks.size() * sizeof(batch_chunk) + list.size() * sizeof(int) + 2 * kMaxSlots * sizeof(int2) +
             S.NC * sizeof(int);
    if (recs) bytes += recs->size() * (sizeof(SlotRec) + 16);
    if (mode_slots) bytes += mode_slots->size() * (192 * sizeof(double) + 16);
    if ((rc = ensure_stage(S, set, bytes))) return rc;
    if (S.stage_used[set]) SIPG_HIP_TRY(hipEventSynchronize(S.stage_ev[set]));
    char *h = S.h_stage[set];
    size_t o = 0;
    auto put = [&](const void *src, size_t n, void *dst) -> int {
        if (n == 0) return SIPG_OK;
        std::memcpy(h + o, src, n);
        SIPG_HIP_TRY(hipMemcpyAsync(dst, h + o, n, hipMemcpyHostToDevice, S.stream));
        o += (n + 15) & ~size_t(15);
        return SIPG_OK;
    };
    std::vector<int> counts(S.NC, 0);
    counts[0] = (int)list.size();
    counts[1] = (int)chunks.size();
    for (int p = 1; p <= kMaxDegree; ++p) {
        counts[2 + p] = (int)items[p].size();
        if ((rc = put(items[p].data(), items[p].size() * sizeof(batch_item), S.d_items[set][p]))) return rc;
    }
    if ((rc = put(chunks.data(), chunks.size() * sizeof(batch_chunk), S.d_chunks[set])) ||
        (rc = put(list.data(), list.size() * sizeof(int), S.d_list[set])) ||
        (rc = put(item_range.data(), item_range.size() * sizeof(int2), S.d_item_range[set])) ||
        (rc = put(chunk_range.data(), chunk_range.size() * sizeof(int2), S.d_chunk_range[set])) ||
        (rc = put(counts.data(), counts.size() * sizeof(int), S.count(set, 0))))
        return rc;
    if (recs)
        for (auto &r : *recs) {
            if ((rc = put(&r.second.inst, sizeof(batch_inst), S.d_inst + r.first)) ||
                (rc = put(&r.second.ctl, sizeof(batch_ctl), S.d_ctl + r.first)))
                return rc;
        }
    if (mode_slots)
        for (int slot : *mode_slots) {
            const Request &R = S.reqs.at(S.slot_req[slot]);
            double buf[192];
            std::memset(buf, 0, sizeof(buf));
            std::memcpy(buf, R.a, R.nmodes * sizeof(double));
            std::memcpy(buf + 64, R.m, R.nmodes * sizeof(int));
            std::memcpy((int *)(buf + 64) + 64, R.n, R.nmodes * sizeof(int));
            if ((rc = put(buf, sizeof(buf), S.d_modes + (size_t)slot * 192))) return rc;
        }
    SIPG_HIP_TRY(hipEventRecord(S.stage_ev[set], S.stream));
    S.stage_used[set] = true;
    return SIPG_OK;
}

int apply_set(sipg_server &S, int set, const std::vector<int> &degrees, const double *x, double *y) {
    for (int p : degrees) {
        int rc = launch_batch_apply(p, S.shape[p], x, y, S.d_inst, S.d_items[set][p], S.count(set, 2 + p),
                                    S.item_cap[set][p], nullptr, nullptr, 0, S.stream);
        if (rc) return rc;
    }
    return SIPG_OK;
}

// z = M^-1 r over the items of a set (PCG; pv non-null: also p = z), with
// per-item r.r and r.z; states non-null: finished requests are skipped
int prec_set(sipg_server &S, int set, const std::vector<int> &degrees, double *pv, const cg_state *states,
             int slot) {
    for (int p : degrees) {
        int rc = launch_batch_prec(p, S.shape[p], S.pool[2], S.pool[6], pv, S.d_inst, S.d_items[set][p],
                                   S.count(set, 2 + p), S.item_cap[set][p], S.d_item_rr[set] + S.item_base[set][p],
                                   S.d_item_rz[set] + S.item_base[set][p], states, slot, S.stream);
        if (rc) return rc;
    }
    return SIPG_OK;
}

std::vector<int> degrees_of(const sipg_server &S, const std::vector<int64_t> &ids) {
    std::vector<char> has(kMaxDegree + 1, 0);
    for (int64_t id : ids) has[S.reqs.at(id).p] = 1;
    std::vector<int> d;
    for (int p = 1; p <= kMaxDegree; ++p)
        if (has[p]) d.push_back(p);
    return d;
}

int capture_graph(sipg_server &S, const std::vector<int> &degrees) {
    if (S.exec) (void)hipGraphExecDestroy(S.exec);
    if (S.graph) (void)hipGraphDestroy(S.graph);
    S.exec = nullptr;
    S.graph = nullptr;
    double *X = S.pool[0], *Rv = S.pool[2], *Pv = S.pool[3], *AP = S.pool[4];
    const int A = SET_ACTIVE;
    int rc = SIPG_OK;
    SIPG_HIP_TRY(hipStreamBeginCapture(S.stream, hipStreamCaptureModeRelaxed));
    for (int it = 0; it < kServeIters && !rc; ++it) {
        const int slot = it & 1;
        for (int p : degrees) {
            if ((rc = launch_batch_apply(p, S.shape[p], Pv, AP, S.d_inst, S.d_items[A][p], S.count(A, 2 + p),
                                         S.item_cap[A][p], S.d_item_part + S.item_base[A][p], S.d_state, slot,
                                         S.stream)))
                break;
        }
        if (rc) break;
        if ((rc = launch_batch_alpha(S.d_list[A], S.count(A, 0), kMaxSlots, S.d_item_range[A], S.d_item_part,
                                     S.d_state, slot, S.stream)) ||
            (rc = launch_batch_update_xr(S.pool[5], Pv, Rv, AP, S.d_chunks[A], S.count(A, 1), S.chunk_cap,
                                         S.d_chunk_part, S.d_state, slot, S.stream)))
            break;
        if (it == 0) {
            // reliable update of the replacement set: x += z, r = b - A x
            const int Q = SET_REPL;
            if ((rc = launch_batch_fold(X, S.pool[5], S.d_chunks[Q], S.count(Q, 1), S.chunk_cap, S.stream)) ||
                (rc = apply_set(S, Q, degrees, X, AP)) ||
                (rc = launch_batch_residual(S.pool[1], AP, Rv, nullptr, S.d_chunks[Q], S.count(Q, 1), S.chunk_cap,
                                            S.d_part2, S.d_part3, S.stream)) ||
                (rc = launch_batch_mark(S.d_list[Q], S.count(Q, 0), kMaxSlots, S.d_state, slot, S.stream)))
                break;
        }
        if (S.method == SIPG_METHOD_PCG) {
            // z = M^-1 r (after any replacement), beta from r.z, p = z + beta p
            if ((rc = prec_set(S, A, degrees, nullptr, S.d_state, slot)) ||
                (rc = launch_batch_beta_pcg(S.d_list[A], S.count(A, 0), kMaxSlots, S.d_item_range[A],
                                            S.d_item_rr[A], S.d_item_rz[A], S.d_state, slot, S.stream)) ||
                (rc = launch_batch_update_p(S.pool[6], Pv, S.d_chunks[A], S.count(A, 1), S.chunk_cap, S.d_state,
                                            slot, S.stream)))
                break;
            continue;
        }
        if ((rc = launch_batch_beta(S.d_list[A], S.count(A, 0), kMaxSlots, S.d_chunk_range[A], S.d_chunk_part,
                                    S.d_chunk_range[SET_REPL], S.d_part2, S.d_state, slot, S.stream)) ||
            (rc = launch_batch_update_p(Rv, Pv, S.d_chunks[A], S.count(A, 1), S.chunk_cap, S.d_state, slot,
                                        S.stream)))
            break;
    }
    const hipError_t ec = hipStreamEndCapture(S.stream, &S.graph);
    if (rc) return rc;
    if (ec != hipSuccess) {
        set_last_hip_error(ec);
        return SIPG_ERR_HIP;
    }
    const hipError_t ei = hipGraphInstantiate(&S.exec, S.graph, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
        set_last_hip_error(ei);
        return SIPG_ERR_HIP;
    }
    S.graph_degrees = degrees;
    S.graph_stale = false;
    return SIPG_OK;
}

void finish(sipg_server &S, Request &R, const cg_state &st, bool cancelled = false) {
    const cg_slot &c = st.s[0];
    int status = SIPG_OK;
    if (cancelled) status = SIPG_ERR_CANCELLED;
    else if (st.err == 2) status = SIPG_ERR_BREAKDOWN;
    else if (st.err == 3) status = SIPG_ERR_NONFINITE;
    else if (!(std::sqrt(c.rr) <= st.stop) || !c.checked) status = SIPG_NOT_CONVERGED;
    R.times.completed = S.now();
    R.info.status = status;
    R.info.iters = c.it;
    R.info.replacements = st.nrep;
    R.info.pad = 0;
    R.info.res_norm = std::sqrt(c.rr);
    R.info.true_res_norm = (c.checked || st.nrep > 0) ? std::sqrt(st.rr_ref) : -1.0;
    R.info.b_norm = std::sqrt(st.bb);
    R.phase = DONE;
    S.slot_req[R.slot] = -1;
    S.free_slots.push_back(R.slot);
    R.slot = -1;
    S.done_q.push_back(R.id);
}

// A queued request that is cancelled or timed out completes without having
// run (it holds no slot and no memory).
void finish_queued(sipg_server &S, Request &R) {
    R.info = sipg_cg_info{};
    R.info.status = SIPG_ERR_CANCELLED;
    R.info.res_norm = -1.0;
    R.info.true_res_norm = -1.0;
    R.info.b_norm = -1.0;
    R.times.completed = S.now();
    R.phase = DONE;
    S.done_q.push_back(R.id);
}

bool timed_out(const sipg_server &S, const Request &R, double now) {
    return S.timeout > 0.0 && now - R.times.submitted > S.timeout;
}

}  // namespace

extern "C" {

int sipg_server_destroy(sipg_server *S) {
    if (!S) return SIPG_OK;
    if (S->stream) (void)hipStreamSynchronize(S->stream);
    if (S->exec) (void)hipGraphExecDestroy(S->exec);
    if (S->graph) (void)hipGraphDestroy(S->graph);
    for (double *p : S->pool)
        if (p) (void)hipFree(p);
    void *dev[] = {S->d_inst, S->d_state, S->d_ctl, S->d_modes, S->d_counts, S->d_item_part,
                   S->d_chunk_part, S->d_part2, S->d_part3};
    for (void *q : dev)
        if (q) (void)hipFree(q);
    for (int s = 0; s < NSETS; ++s) {
        if (S->d_item_rr[s]) (void)hipFree(S->d_item_rr[s]);
        if (S->d_item_rz[s]) (void)hipFree(S->d_item_rz[s]);
        for (int p = 0; p <= kMaxDegree; ++p)
            if (S->d_items[s][p]) (void)hipFree(S->d_items[s][p]);
        void *q[] = {S->d_chunks[s], S->d_list[s], S->d_item_range[s], S->d_chunk_range[s]};
        for (void *v : q)
            if (v) (void)hipFree(v);
        if (S->h_stage[s]) (void)hipHostFree(S->h_stage[s]);
        if (S->stage_ev[s]) (void)hipEventDestroy(S->stage_ev[s]);
    }
    for (auto &kv : S->d_tab) (void)hipFree(kv.second);
    if (S->h_state) (void)hipHostFree(S->h_state);
    if (S->own_stream && S->stream) (void)hipStreamDestroy(S->stream);
    delete S;
    return SIPG_OK;
}

int sipg_server_create(sipg_server **out, double hbm_budget_bytes, void *stream) {
    if (!out) return SIPG_ERR_INVALID;
    *out = nullptr;
    if (!(hbm_budget_bytes >= 1024.0) || !std::isfinite(hbm_budget_bytes)) return SIPG_ERR_INVALID;
    sipg_server *S = new sipg_server();
    auto fail = [&](int code) { sipg_server_destroy(S); return code; };
    hipError_t e;
    if (stream) {
        S->stream = (hipStream_t)stream;
    } else {
        if ((e = hipStreamCreateWithFlags(&S->stream, hipStreamNonBlocking)) != hipSuccess) {
            set_last_hip_error(e);
            return fail(SIPG_ERR_HIP);
        }
        S->own_stream = true;
    }
    S->budget = hbm_budget_bytes;
    S->cap = (int64_t)(hbm_budget_bytes / (S->npools * sizeof(double)));
    S->freel[0] = S->cap;
    for (int p = 0; p <= kMaxDegree; ++p) S->shape[p] = 1;   // 8x4 until set otherwise
    int rc;
    for (int k = 0; k < S->npools; ++k)
        if ((rc = dmalloc(&S->pool[k], S->cap))) return fail(rc);
    S->chunk_cap = (int)std::min<int64_t>(S->cap / kChunk + kMaxSlots, INT_MAX / 2);
    if ((rc = dmalloc(&S->d_inst, kMaxSlots)) || (rc = dmalloc(&S->d_state, kMaxSlots)) ||
        (rc = dmalloc(&S->d_ctl, kMaxSlots)) ||
        (rc = dmalloc(&S->d_modes, (size_t)kMaxSlots * 192)) || (rc = dmalloc(&S->d_counts, NSETS * S->NC)) ||
        (rc = dmalloc(&S->d_chunk_part, S->chunk_cap)) || (rc = dmalloc(&S->d_part2, S->chunk_cap)) ||
        (rc = dmalloc(&S->d_part3, S->chunk_cap)))
        return fail(rc);
    for (int s = 0; s < NSETS; ++s) {
        if ((rc = dmalloc(&S->d_chunks[s], S->chunk_cap)) || (rc = dmalloc(&S->d_list[s], kMaxSlots)) ||
            (rc = dmalloc(&S->d_item_range[s], kMaxSlots)) || (rc = dmalloc(&S->d_chunk_range[s], kMaxSlots)))
            return fail(rc);
        if ((e = hipEventCreateWithFlags(&S->stage_ev[s], hipEventDisableTiming)) != hipSuccess) {
            set_last_hip_error(e);
            return fail(SIPG_ERR_HIP);
        }
    }
    if ((e = hipMemset(S->d_counts, 0, NSETS * S->NC * sizeof(int))) != hipSuccess ||
        (e = hipMemset(S->d_state, 0, kMaxSlots * sizeof(cg_state))) != hipSuccess) {
        set_last_hip_error(e);
        return fail(SIPG_ERR_HIP);
    }
    if ((e = hipHostMalloc((void **)&S->h_state, kMaxSlots * sizeof(cg_state), hipHostMallocDefault)) !=
        hipSuccess) {
        set_last_hip_error(e);
        return fail(SIPG_ERR_NOMEM);
    }
    for (int p = 1; p <= kMaxDegree; ++p)
        if ((rc = ensure_items(*S, SET_ACTIVE, p, 1024))) return fail(rc);
    S->slot_req.assign(kMaxSlots, -1);
    for (int s = kMaxSlots - 1; s >= 0; --s) S->free_slots.push_back(s);
    *out = S;
    return SIPG_OK;
}

int sipg_server_set_shape(sipg_server *S, int p, int bx, int by) {
    if (!S || p < 1 || p > kMaxDegree) return SIPG_ERR_INVALID;
    for (int k = 0; k < kBatchShapes; ++k) {
        if (kBatchShapeBX[k] != bx || kBatchShapeBY[k] != by) continue;
        if (!batch_shape_supported(p, k)) return SIPG_ERR_UNSUPPORTED;
        if (S->method == SIPG_METHOD_PCG && !batch_prec_supported(p, 5.0 * (p + 1) * (p + 1), k))
            return SIPG_ERR_UNSUPPORTED;
        for (auto &kv : S->reqs)
            if (kv.second.p == p && kv.second.phase == RUNNING)
                return SIPG_ERR_INVALID;       // only while no request of degree p runs
        S->shape[p] = k;
        S->graph_stale = true;
        return SIPG_OK;
    }
    return SIPG_ERR_UNSUPPORTED;
}

int sipg_server_set_method(sipg_server *S, int method) {
    if (!S || (method != SIPG_METHOD_CG && method != SIPG_METHOD_PCG) || !S->reqs.empty()) return SIPG_ERR_INVALID;
    const int np = method == SIPG_METHOD_PCG ? 7 : 6;
    if (method == SIPG_METHOD_PCG)
        for (int p = 1; p <= kMaxDegree; ++p)
            if (!batch_prec_supported(p, 5.0 * (p + 1) * (p + 1), S->shape[p])) return SIPG_ERR_UNSUPPORTED;
    if (np != S->npools) {
        // the same budget over a different number of pools
        SIPG_HIP_TRY(hipStreamSynchronize(S->stream));
        for (double *&pl : S->pool) {
            if (pl) (void)hipFree(pl);
            pl = nullptr;
        }
        S->npools = np;
        S->cap = (int64_t)(S->budget / (np * sizeof(double)));
        S->freel.clear();
        S->freel[0] = S->cap;
        S->used = 0;
        for (int k = 0; k < np; ++k) {
            const int rc = dmalloc(&S->pool[k], S->cap);
            if (rc) {
                // no pools: every later submit reports SIPG_ERR_BUDGET
                for (double *&pl : S->pool) {
                    if (pl) (void)hipFree(pl);
                    pl = nullptr;
                }
                S->cap = 0;
                S->freel.clear();
                return rc;
            }
        }
    }
    S->method = method;
    S->graph_stale = true;
    return SIPG_OK;
}

int sipg_server_submit(sipg_server *S, int dim, int p, double pen, int nmodes, const double *a, const int *m,
                       const int *n, double tol, int tol_mode, int maxit, int64_t *id) {
    if (!S || dim < 1 || p < 1 || !std::isfinite(pen) || nmodes < 0 || nmodes > 64 ||
        (nmodes > 0 && (!a || !m || !n)) || !(tol >= 0.0) || (tol_mode != SIPG_TOL_ABS && tol_mode != SIPG_TOL_REL))
        return SIPG_ERR_INVALID;
    if (p > kMaxDegree) return SIPG_ERR_UNSUPPORTED;
    if ((int64_t)6 * dim * (p + 1) * (p + 1) > (int64_t)INT_MAX) return SIPG_ERR_UNSUPPORTED;
    if (S->method == SIPG_METHOD_PCG && !batch_prec_supported(p, pen, S->shape[p])) return SIPG_ERR_UNSUPPORTED;
    const int64_t len = (int64_t)dim * dim * (p + 1) * (p + 1);
    if (len > S->cap) return SIPG_ERR_BUDGET;             // would never fit
    Request R;
    R.id = S->next_id++;
    R.dim = dim;
    R.p = p;
    R.pen = pen;
    R.nmodes = nmodes;
    for (int k = 0; k < nmodes; ++k) {
        R.a[k] = a[k];
        R.m[k] = m[k];
        R.n[k] = n[k];
    }
    R.tol = tol;
    R.tol_mode = tol_mode;
    R.maxit = (int)std::min<int64_t>(maxit > 0 ? (int64_t)maxit : len, INT_MAX);
    R.len = len;
    R.times.submitted = S->now();
    S->reqs[R.id] = R;
    S->queue.push_back(R.id);
    if (id) *id = R.id;
    return SIPG_OK;
}

int sipg_server_progress(sipg_server *S, int *n_done) {
    if (!S) return SIPG_ERR_INVALID;
    int rc;
    bool lists_changed = false;
    // 1-3: outcome of the previous replay; the replacement set of the next
    // one (reliable updates: every 300x drop of the recursive residual, and
    // once it meets the tolerance -- the request then finishes only on
    // that true residual)
    std::vector<int64_t> repl;
    const double now = S->now();
    // 0: queued requests past the time limit complete as cancelled
    if (S->timeout > 0.0) {
        std::deque<int64_t> keep;
        for (int64_t id : S->queue) {
            Request &R = S->reqs.at(id);
            if (timed_out(*S, R, now)) finish_queued(*S, R);
            else keep.push_back(id);
        }
        S->queue.swap(keep);
    }
    if (S->replayed) {
        std::vector<int64_t> still;
        for (int64_t id : S->active) {
            Request &R = S->reqs.at(id);
            const cg_state &st = S->h_state[R.slot];
            const cg_slot &c = st.s[0];
            if (R.cancel || timed_out(*S, R, now)) {      // retired with its iterate so far
                finish(*S, R, st, true);
                lists_changed = true;
                continue;
            }
            if (st.err || c.done) {
                finish(*S, R, st);
                lists_changed = true;
                continue;
            }
            if (st.nrep != R.nrep_req) {                  // the last one is not in yet
                still.push_back(id);
                continue;
            }
            if (R.check_req) {
                // kStallChecks (PCG: kStallChecksPcg) failed final checks
                // that gain less than 10% on the best one, with the best one
                // stall_window iterations old, mean the FP64 floor: the
                // request ends there
                R.check_req = false;
                const bool gain = R.check_ref < 0.0 || st.rr_ref <= 0.81 * R.check_ref;
                if (!gain) ++R.check_stale;
                else R.best_it = c.it;
                if (R.check_stale >= (S->method == SIPG_METHOD_PCG ? kStallChecksPcg : kStallChecks) &&
                    c.it - R.best_it >= stall_window(c.it) && floor_in_reach(st.rr_ref, st.bb)) {
                    finish(*S, R, st);
                    lists_changed = true;
                    continue;
                }
                if (gain) R.check_ref = st.rr_ref;
            }
            // PCG: a recursion that does not fall 4x (in r.r) below the last
            // true residual within stall_window iterations has stalled at
            // the FP64 floor without meeting the tolerance: the request
            // ends there
            if (S->method == SIPG_METHOD_PCG && !(c.flags & kSlotWantCheck)) {
                R.rec_min = R.rec_min < 0.0 ? c.rr : std::min(R.rec_min, c.rr);
                if (floor_in_reach(st.rr_ref, st.bb) && R.rec_min > 0.25 * st.rr_ref &&
                    c.it - R.rep_it >= stall_window(c.it)) {
                    finish(*S, R, st);
                    lists_changed = true;
                    continue;
                }
            }
            still.push_back(id);
            const bool check = (c.flags & kSlotWantCheck) &&
                               (c.rr <= 0.0625 * st.rr_ref || c.it - R.rep_it >= force_check_interval(c.it));
            const bool periodic = c.rr <= kServeReplace * kServeReplace * st.rr_ref;
            if (check || periodic) {
                repl.push_back(id);
                ++R.nrep_req;
                R.check_req = check;
                R.rep_it = c.it;
                R.rec_min = -1.0;
            }
        }
        S->active.swap(still);
        S->replayed = false;
    }
    // 4: admissions, in arrival order while memory and slots last
    std::vector<int64_t> admitted;
    std::vector<std::pair<int, SlotRec>> recs;
    std::vector<int> slots;
    while (!S->queue.empty() && !S->free_slots.empty()) {
        Request &R = S->reqs.at(S->queue.front());
        const int64_t off = range_alloc(*S, R.len);
        if (off < 0) break;
        R.off = off;
        R.slot = S->free_slots.back();
        S->free_slots.pop_back();
        S->slot_req[R.slot] = R.id;
        R.phase = RUNNING;
        R.times.admitted = S->now();
        S->queue.pop_front();
        admitted.push_back(R.id);
        slots.push_back(R.slot);
        recs.push_back({R.slot, SlotRec{batch_inst{R.off, R.dim, R.dim, R.pen}, batch_ctl{R.tol, R.tol_mode, R.maxit}}});
        if (!S->d_tab.count(R.p)) {
            std::vector<double> tab;
            build_tables(R.p, tab);
            double *d = nullptr;
            if ((rc = dmalloc(&d, tab.size()))) return rc;
            SIPG_HIP_TRY(hipMemcpy(d, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice));
            S->d_tab[R.p] = d;
        }
    }
    if (!admitted.empty()) {
        if ((rc = upload_set(*S, SET_INIT, admitted, &recs, &slots))) return rc;
        for (int64_t id : admitted) {
            const Request &R = S->reqs.at(id);
            SIPG_HIP_TRY(hipMemsetAsync(S->pool[0] + R.off, 0, R.len * sizeof(double), S->stream));
            SIPG_HIP_TRY(hipMemsetAsync(S->pool[4] + R.off, 0, R.len * sizeof(double), S->stream));
            SIPG_HIP_TRY(hipMemsetAsync(S->pool[5] + R.off, 0, R.len * sizeof(double), S->stream));
            double *md = S->d_modes + (size_t)R.slot * 192;
            if ((rc = launch_rhs_fourier(SIPG_F64, R.p, R.dim, S->d_tab[R.p], R.nmodes, md, (const int *)(md + 64),
                                         (const int *)(md + 64) + 64, S->pool[1] + R.off, S->stream)))
                return rc;
        }
        // x = 0: r = p = b (PCG: z = M^-1 r, p = z); initial states
        const int I = SET_INIT;
        const bool pcg = S->method == SIPG_METHOD_PCG;
        if ((rc = launch_batch_residual(S->pool[1], S->pool[4], S->pool[2], S->pool[3], S->d_chunks[I],
                                        S->count(I, 1), S->chunk_cap, S->d_part2, S->d_part3, S->stream)) ||
            (pcg && (rc = prec_set(*S, I, degrees_of(*S, admitted), S->pool[3], nullptr, 0))) ||
            (rc = launch_batch_init(S->d_list[I], S->count(I, 0), kMaxSlots, S->d_chunk_range[I], S->d_part2,
                                    S->d_part3, S->d_state, S->d_ctl, pcg ? S->d_item_range[I] : nullptr,
                                    pcg ? S->d_item_rz[I] : nullptr, S->stream)))
            return rc;
        for (int64_t id : admitted) S->active.push_back(id);
        lists_changed = true;
    }
    // 5: lists, graph, one replay
    if ((rc = upload_set(*S, SET_REPL, repl))) return rc;
    if (lists_changed) {
        if ((rc = upload_set(*S, SET_ACTIVE, S->active))) return rc;
        const std::vector<int> deg = degrees_of(*S, S->active);
        if (deg != S->graph_degrees) S->graph_stale = true;
    }
    if (!S->active.empty()) {
        if (S->graph_stale && (rc = capture_graph(*S, degrees_of(*S, S->active)))) return rc;
        SIPG_HIP_TRY(hipGraphLaunch(S->exec, S->stream));
        SIPG_HIP_TRY(hipMemcpyAsync(S->h_state, S->d_state, kMaxSlots * sizeof(cg_state), hipMemcpyDeviceToHost,
                                    S->stream));
        S->replayed = true;
    }
    SIPG_HIP_TRY(hipStreamSynchronize(S->stream));
    if (n_done) *n_done = (int)S->done_q.size();
    return SIPG_OK;
}

int sipg_server_take(sipg_server *S, int64_t *id, sipg_cg_info *info, double *x_host, int64_t n_cap) {
    return sipg_server_take_ex(S, id, info, x_host, n_cap, nullptr);
}

int sipg_server_peek(const sipg_server *S, int64_t *id, int64_t *n) {
    if (!S || !id) return SIPG_ERR_INVALID;
    if (S->done_q.empty()) return SIPG_ERR_NOT_FOUND;
    *id = S->done_q.front();
    if (n) *n = S->reqs.at(*id).len;
    return SIPG_OK;
}

int sipg_server_cancel(sipg_server *S, int64_t id) {
    if (!S) return SIPG_ERR_INVALID;
    auto it = S->reqs.find(id);
    if (it == S->reqs.end() || it->second.phase == DONE) return SIPG_ERR_NOT_FOUND;
    Request &R = it->second;
    if (R.phase == QUEUED) {
        S->queue.erase(std::find(S->queue.begin(), S->queue.end(), id));
        finish_queued(*S, R);
    } else {
        R.cancel = true;
    }
    return SIPG_OK;
}

int sipg_server_set_timeout(sipg_server *S, double seconds) {
    if (!S || !(seconds >= 0.0) || !std::isfinite(seconds)) return SIPG_ERR_INVALID;
    S->timeout = seconds;
    return SIPG_OK;
}

int sipg_server_take_ex(sipg_server *S, int64_t *id, sipg_cg_info *info, double *x_host, int64_t n_cap,
                        sipg_request_times *times) {
    if (!S || !id) return SIPG_ERR_INVALID;
    if (S->done_q.empty()) return SIPG_ERR_NOT_FOUND;
    const int64_t rid = S->done_q.front();
    Request &R = S->reqs.at(rid);
    if (x_host && R.off < 0) {                          // never ran: the initial guess x = 0
        if (n_cap < R.len) return SIPG_ERR_INVALID;
        std::fill(x_host, x_host + R.len, 0.0);
    } else if (x_host) {
        if (n_cap < R.len) return SIPG_ERR_INVALID;
        const int64_t noe = (int64_t)R.dim * R.dim, nq = (int64_t)(R.p + 1) * (R.p + 1);
        // x + z: the increments after the last true residual (none unless the
        // request stopped on its iteration cap)
        std::vector<double> tmp(R.len), z(R.len);
        SIPG_HIP_TRY(hipStreamSynchronize(S->stream));
        SIPG_HIP_TRY(hipMemcpy(tmp.data(), S->pool[0] + R.off, R.len * sizeof(double), hipMemcpyDeviceToHost));
        SIPG_HIP_TRY(hipMemcpy(z.data(), S->pool[5] + R.off, R.len * sizeof(double), hipMemcpyDeviceToHost));
        for (int64_t e = 0; e < noe; ++e)
            for (int64_t q = 0; q < nq; ++q) x_host[q * noe + e] = tmp[e * nq + q] + z[e * nq + q];
    }
    *id = rid;
    if (info) *info = R.info;
    if (times) *times = R.times;
    if (R.off >= 0) range_free(*S, R.off, R.len);
    S->reqs.erase(rid);
    S->done_q.pop_front();
    return SIPG_OK;
}

int sipg_server_stats(const sipg_server *S, int *queued, int *running, int *done_held, double *bytes_used) {
    if (!S) return SIPG_ERR_INVALID;
    if (queued) *queued = (int)S->queue.size();
    if (running) *running = (int)S->active.size();
    if (done_held) *done_held = (int)S->done_q.size();
    if (bytes_used) *bytes_used = (double)S->used * S->npools * sizeof(double);
    return SIPG_OK;
}

}  // extern "C"
