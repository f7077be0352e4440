// batch.cpp -- many independent SIPG Poisson solves at once (FP64).
//
// The reference solves one mesh per process (its degree tables are process
// globals, device_const.hpp:10-18); a batch here holds any number of
// instances, each with its own mesh size, degree and penalty.  All vectors
// are concatenated (element-major within an instance).  The solver runs
// every instance's CG in the same launches:
//
//   per iteration  one fused apply + p.Ap launch per degree present (over a
//                  device list of (instance, block) items), then alpha per
//                  instance, x/r update (+ r.r per chunk), beta per instance
//                  and the p update, over device lists of chunks;
//   hipGraph       kGraphIters iterations are captured once per solve and
//                  replayed; the lists and their lengths live in device
//                  memory, so shrinking them as instances finish keeps the
//                  graph valid;
//   host loop      one replay in flight ahead; per-instance states are
//                  polled one replay behind.  The host picks the instances
//                  whose r is to be replaced by b - A x in the first
//                  iteration of a coming replay (reliable updates, as in
//                  sipg_cg_ex: periodically, and once the recursion meets
//                  the tolerance, which then finishes only on that true
//                  residual), and finished instances leave the lists.
//
// The block shape of the operator kernel is chosen per degree by timing the
// candidates on the batch's own instances (sipg_batch_tune).
#include "sipg_internal.hpp"

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

using namespace sipg_internal;

namespace {

constexpr int64_t kChunk = 8192;   // vector elements per chunk
constexpr int kGraphIters = 8;     // even: every replay starts on slot 0
enum { SET_ACTIVE = 0, SET_ALL = 1, SET_REPL = 2, NSETS = 3 };

struct Lists {
    std::vector<std::vector<batch_item>> items;   // per group
    std::vector<batch_chunk> chunks;
    std::vector<int> list;
    std::vector<int2> item_range, chunk_range;    // per instance id
};

}  // namespace

struct sipg_batch {
    int ninst = 0;
    std::vector<int> dimx, dimy, p;
    std::vector<double> pen;
    std::vector<int64_t> off;            // ninst + 1
    hipStream_t stream = nullptr;
    bool own_stream = false;
    bool graphs = true;                  // replay captured iterations (sipg_batch_set_graphs)

    struct Group {
        int p = 0, shape = -1;
        std::vector<int> inst;           // instance ids of this degree
        int cap = 0;                     // items of all instances, chosen shape
        int base = 0;                    // first item_part index of the group
        double us[kBatchShapes] = {0, 0, 0};
        batch_item *d_items[NSETS] = {nullptr, nullptr, nullptr};
    };
    std::vector<Group> groups;
    std::vector<int> group_of;

    batch_inst *d_inst = nullptr;
    cg_state *d_state = nullptr;
    cg_state *h_state = nullptr;         // pinned, 2 x ninst
    batch_ctl *d_ctl = nullptr;
    double *ws = nullptr;                // R, P, AP, Z (increments since the last true residual)
    double *d_item_part = nullptr, *d_chunk_part = nullptr, *d_part2 = nullptr, *d_part3 = nullptr;
    double *d_item_rr = nullptr, *d_item_rz = nullptr;   // PCG: per-item r.r and r.z
    double *d_zq = nullptr;              // PCG: z = M^-1 r (allocated on first use)
    int item_total = 0, chunk_cap = 0;
    int *d_counts = nullptr;             // NSETS x (2 + ngroups)
    int *d_list[NSETS] = {};
    batch_chunk *d_chunks[NSETS] = {};
    int2 *d_item_range[NSETS] = {}, *d_chunk_range[NSETS] = {};
    // pinned upload staging per set, and the event of its last upload
    char *h_stage[NSETS] = {};
    size_t stage_bytes = 0;
    hipEvent_t stage_ev[NSETS] = {};
    bool stage_used[NSETS] = {false, false, false};
    std::map<int, double *> d_tab;       // degree -> FP64 tables (RHS)
    double *d_modes = nullptr;           // RHS forcing scratch
    bool tuned = false;

    int64_t total() const { return off[ninst]; }
    int ncount() const { return 2 + (int)groups.size(); }
    int *count(int set, int k) { return d_counts + set * ncount() + k; }
    double *R() { return ws; }
    double *Pv() { return ws + total(); }
    double *AP() { return ws + 2 * total(); }
    double *Z() { return ws + 3 * total(); }
};

namespace {

int items_for(const sipg_batch &b, int i, int shape, std::vector<batch_item> &out) {
    const int BX = kBatchShapeBX[shape], BY = kBatchShapeBY[shape];
    const int nbx = (b.dimx[i] + BX - 1) / BX, nby = (b.dimy[i] + BY - 1) / BY;
    for (int by = 0; by < nby; ++by)
        for (int bx = 0; bx < nbx; ++bx) out.push_back({i, bx * BX, by * BY, 0});
    return nbx * nby;
}

void build_lists(const sipg_batch &b, const std::vector<char> &in_set, Lists &L) {
    const int ng = (int)b.groups.size();
    L.items.assign(ng, {});
    L.chunks.clear();
    L.list.clear();
    L.item_range.assign(b.ninst, int2{0, 0});
    L.chunk_range.assign(b.ninst, int2{0, 0});
    for (int g = 0; g < ng; ++g) {
        const auto &G = b.groups[g];
        for (int i : G.inst) {
            if (!in_set[i]) continue;
            const int s0 = (int)L.items[g].size();
            items_for(b, i, G.shape, L.items[g]);
            L.item_range[i] = int2{G.base + s0, G.base + (int)L.items[g].size()};
        }
    }
    for (int i = 0; i < b.ninst; ++i) {
        if (!in_set[i]) continue;
        L.list.push_back(i);
        const int c0 = (int)L.chunks.size();
        for (int64_t a = b.off[i]; a < b.off[i + 1]; a += kChunk)
            L.chunks.push_back({i, 0, a, std::min(a + kChunk, b.off[i + 1])});
        L.chunk_range[i] = int2{c0, (int)L.chunks.size()};
    }
}

// Copies the lists of `set` to the device through its pinned staging area,
// asynchronously on the batch stream.  The staging area is reused only
// after its previous copy has completed.
int upload_lists(sipg_batch &b, int set, const Lists &L) {
    if (b.stage_used[set]) SIPG_HIP_TRY(hipEventSynchronize(b.stage_ev[set]));
    char *h = b.h_stage[set];
    size_t o = 0;
    auto put = [&](const void *src, size_t bytes, void *dst) -> int {
        if (bytes == 0) return SIPG_OK;
        std::memcpy(h + o, src, bytes);
        SIPG_HIP_TRY(hipMemcpyAsync(dst, h + o, bytes, hipMemcpyHostToDevice, b.stream));
        o += (bytes + 15) & ~size_t(15);
        return SIPG_OK;
    };
    int rc;
    std::vector<int> counts(b.ncount());
    counts[0] = (int)L.list.size();
    counts[1] = (int)L.chunks.size();
    for (size_t g = 0; g < b.groups.size(); ++g) {
        counts[2 + g] = (int)L.items[g].size();
        if ((rc = put(L.items[g].data(), L.items[g].size() * sizeof(batch_item), b.groups[g].d_items[set])))
            return rc;
    }
    if ((rc = put(L.chunks.data(), L.chunks.size() * sizeof(batch_chunk), b.d_chunks[set]))) return rc;
    if ((rc = put(L.list.data(), L.list.size() * sizeof(int), b.d_list[set]))) return rc;
    if ((rc = put(L.item_range.data(), L.item_range.size() * sizeof(int2), b.d_item_range[set]))) return rc;
    if ((rc = put(L.chunk_range.data(), L.chunk_range.size() * sizeof(int2), b.d_chunk_range[set]))) return rc;
    if ((rc = put(counts.data(), counts.size() * sizeof(int), b.count(set, 0)))) return rc;
    SIPG_HIP_TRY(hipEventRecord(b.stage_ev[set], b.stream));
    b.stage_used[set] = true;
    return SIPG_OK;
}

// y = A x over the instances of a list set (no p.Ap)
int apply_set(sipg_batch &b, int set, const double *x, double *y) {
    for (size_t g = 0; g < b.groups.size(); ++g) {
        auto &G = b.groups[g];
        int rc = launch_batch_apply(G.p, G.shape, x, y, b.d_inst, G.d_items[set], b.count(set, 2 + (int)g),
                                    G.cap, nullptr, nullptr, 0, b.stream);
        if (rc) return rc;
    }
    return SIPG_OK;
}

// z = M^-1 r over the items of a list set (PCG; pv non-null: also p = z)
int prec_set(sipg_batch &b, int set, double *pv, const cg_state *states, int slot) {
    for (size_t g = 0; g < b.groups.size(); ++g) {
        auto &G = b.groups[g];
        int rc = launch_batch_prec(G.p, G.shape, b.R(), b.d_zq, pv, b.d_inst, G.d_items[set],
                                   b.count(set, 2 + (int)g), G.cap, b.d_item_rr + G.base, b.d_item_rz + G.base,
                                   states, slot, b.stream);
        if (rc) return rc;
    }
    return SIPG_OK;
}

// One CG iteration of the active instances.  With `repl`, the instances of
// the replacement set get r = b - A x between the x/r update and the p
// update (reliable update; the set is empty in most replays).  With `pcg`,
// z = M^-1 r follows (after any replacement), beta comes from r.z, and
// p = z + beta p.
int enqueue_iteration(sipg_batch &b, double *X, const double *Bv, int slot, bool repl, bool pcg) {
    const int A = SET_ACTIVE, Q = SET_REPL;
    int rc;
    for (size_t g = 0; g < b.groups.size(); ++g) {
        auto &G = b.groups[g];
        if ((rc = launch_batch_apply(G.p, G.shape, b.Pv(), b.AP(), b.d_inst, G.d_items[A],
                                     b.count(A, 2 + (int)g), G.cap, b.d_item_part + G.base, b.d_state, slot,
                                     b.stream)))
            return rc;
    }
    if ((rc = launch_batch_alpha(b.d_list[A], b.count(A, 0), b.ninst, b.d_item_range[A], b.d_item_part,
                                 b.d_state, slot, b.stream)))
        return rc;
    // increments go to z; x changes only at a true residual (group update)
    if ((rc = launch_batch_update_xr(b.Z(), b.Pv(), b.R(), b.AP(), b.d_chunks[A], b.count(A, 1), b.chunk_cap,
                                     b.d_chunk_part, b.d_state, slot, b.stream)))
        return rc;
    if (repl) {
        if ((rc = launch_batch_fold(X, b.Z(), b.d_chunks[Q], b.count(Q, 1), b.chunk_cap, b.stream)) ||
            (rc = apply_set(b, Q, X, b.AP())) ||
            (rc = launch_batch_residual(Bv, b.AP(), b.R(), nullptr, b.d_chunks[Q], b.count(Q, 1), b.chunk_cap,
                                        b.d_part2, b.d_part3, b.stream)) ||
            (rc = launch_batch_mark(b.d_list[Q], b.count(Q, 0), b.ninst, b.d_state, slot, b.stream)))
            return rc;
    }
    if (pcg) {
        if ((rc = prec_set(b, A, nullptr, b.d_state, slot)) ||
            (rc = launch_batch_beta_pcg(b.d_list[A], b.count(A, 0), b.ninst, b.d_item_range[A], b.d_item_rr,
                                        b.d_item_rz, b.d_state, slot, b.stream)))
            return rc;
        return launch_batch_update_p(b.d_zq, b.Pv(), b.d_chunks[A], b.count(A, 1), b.chunk_cap, b.d_state, slot,
                                     b.stream);
    }
    if ((rc = launch_batch_beta(b.d_list[A], b.count(A, 0), b.ninst, b.d_chunk_range[A], b.d_chunk_part,
                                b.d_chunk_range[Q], b.d_part2, b.d_state, slot, b.stream)))
        return rc;
    return launch_batch_update_p(b.R(), b.Pv(), b.d_chunks[A], b.count(A, 1), b.chunk_cap, b.d_state, slot,
                                 b.stream);
}

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

// allocations that depend on the chosen shapes
int alloc_lists(sipg_batch &b) {
    int rc;
    b.item_total = 0;
    for (auto &G : b.groups) {
        std::vector<batch_item> tmp;
        G.cap = 0;
        for (int i : G.inst) G.cap += items_for(b, i, G.shape, tmp);
        G.base = b.item_total;
        b.item_total += G.cap;
        for (int s = 0; s < NSETS; ++s) {
            if (G.d_items[s]) (void)hipFree(G.d_items[s]);
            G.d_items[s] = nullptr;
            if ((rc = dmalloc(&G.d_items[s], G.cap))) return rc;
        }
    }
    for (double **q : {&b.d_item_part, &b.d_item_rr, &b.d_item_rz}) {
        if (*q) (void)hipFree(*q);
        *q = nullptr;
        if ((rc = dmalloc(q, b.item_total))) return rc;
    }
    // pinned staging: everything one set can hold
    size_t bytes = 0;
    for (auto &G : b.groups) bytes += ((size_t)G.cap * sizeof(batch_item) + 15) & ~size_t(15);
    bytes += ((size_t)b.chunk_cap * sizeof(batch_chunk) + 15) & ~size_t(15);
    bytes += ((size_t)b.ninst * (sizeof(int) + 2 * sizeof(int2)) + 64) & ~size_t(15);
    bytes += (size_t)b.ncount() * sizeof(int) + 64;
    if (bytes > b.stage_bytes) {
        for (int s = 0; s < NSETS; ++s) {
            if (b.stage_used[s]) SIPG_HIP_TRY(hipEventSynchronize(b.stage_ev[s]));
            if (b.h_stage[s]) (void)hipHostFree(b.h_stage[s]);
            b.h_stage[s] = nullptr;
            const hipError_t e = hipHostMalloc((void **)&b.h_stage[s], bytes, hipHostMallocDefault);
            if (e != hipSuccess) {
                set_last_hip_error(e);
                return SIPG_ERR_NOMEM;
            }
        }
        b.stage_bytes = bytes;
    }
    std::vector<char> all(b.ninst, 1);
    Lists L;
    build_lists(b, all, L);
    if ((rc = upload_lists(b, SET_ALL, L))) return rc;
    SIPG_HIP_TRY(hipStreamSynchronize(b.stream));
    return SIPG_OK;
}

// Times each candidate block shape on all instances of each degree and
// keeps the fastest (2 warm-up + 5 timed applies per candidate).
// Tuning results kept for the life of the process, keyed by the device's
// architecture, the degree and the size class of that degree's work (the
// binary order of its element count; batches are FP64): a later batch of
// the same kind takes the cached shape without timing.  An explicit
// sipg_batch_tune always times, and refreshes the entry.
struct TuneKey {
    std::string arch;
    int p, size_class;
    bool operator<(const TuneKey &o) const {
        return std::tie(arch, p, size_class) < std::tie(o.arch, o.p, o.size_class);
    }
};
struct TuneVal {
    int shape;
    double us[kBatchShapes];
};
std::mutex g_tune_mu;
std::map<TuneKey, TuneVal> g_tune;

std::string device_arch() {
    int dev = 0;
    hipDeviceProp_t prop{};
    if (hipGetDevice(&dev) != hipSuccess || hipGetDeviceProperties(&prop, dev) != hipSuccess) return "?";
    return std::string(prop.gcnArchName);
}

TuneKey tune_key(const sipg_batch &b, const sipg_batch::Group &G, const std::string &arch) {
    int64_t noe = 0;
    for (int i : G.inst) noe += (int64_t)b.dimx[i] * b.dimy[i];
    int cls = 0;
    while ((noe >> cls) > 1) ++cls;
    return TuneKey{arch, G.p, cls};
}

int tune(sipg_batch &b, bool force) {
    int rc;
    const std::string arch = device_arch();
    if (!force) {
        // every degree in the cache: no timing at all
        std::lock_guard<std::mutex> lk(g_tune_mu);
        bool all = true;
        for (auto &G : b.groups) all = all && g_tune.count(tune_key(b, G, arch));
        if (all) {
            for (auto &G : b.groups) {
                const TuneVal &v = g_tune.at(tune_key(b, G, arch));
                G.shape = v.shape;
                std::copy(v.us, v.us + kBatchShapes, G.us);
            }
            b.tuned = true;
            return alloc_lists(b);
        }
    }
    hipEvent_t e0, e1;
    SIPG_HIP_TRY(hipEventCreate(&e0));
    SIPG_HIP_TRY(hipEventCreate(&e1));
    int *d_cnt = nullptr;
    batch_item *d_tmp = nullptr;
    size_t tmp_cap = 0;
    if ((rc = dmalloc(&d_cnt, 1))) return rc;
    for (auto &G : b.groups) {
        double best = 1e300;
        int best_shape = -1;
        for (int sh = 0; sh < kBatchShapes; ++sh) {
            G.us[sh] = 0.0;
            if (!batch_shape_supported(G.p, sh)) continue;
            std::vector<batch_item> items;
            for (int i : G.inst) items_for(b, i, sh, items);
            if (items.size() > tmp_cap) {
                if (d_tmp) (void)hipFree(d_tmp);
                d_tmp = nullptr;
                if ((rc = dmalloc(&d_tmp, items.size()))) break;
                tmp_cap = items.size();
            }
            const int n = (int)items.size();
            SIPG_HIP_TRY(hipMemcpy(d_tmp, items.data(), items.size() * sizeof(batch_item), hipMemcpyHostToDevice));
            SIPG_HIP_TRY(hipMemcpy(d_cnt, &n, sizeof(int), hipMemcpyHostToDevice));
            for (int r = 0; r < 2 && !rc; ++r)
                rc = launch_batch_apply(G.p, sh, b.R(), b.AP(), b.d_inst, d_tmp, d_cnt, n, nullptr, nullptr, 0,
                                        b.stream);
            SIPG_HIP_TRY(hipEventRecord(e0, b.stream));
            for (int r = 0; r < 5 && !rc; ++r)
                rc = launch_batch_apply(G.p, sh, b.R(), b.AP(), b.d_inst, d_tmp, d_cnt, n, nullptr, nullptr, 0,
                                        b.stream);
            SIPG_HIP_TRY(hipEventRecord(e1, b.stream));
            SIPG_HIP_TRY(hipEventSynchronize(e1));
            if (rc) break;
            float ms = 0.f;
            SIPG_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
            G.us[sh] = 1e3 * ms / 5.0;
            if (G.us[sh] < best) {
                best = G.us[sh];
                best_shape = sh;
            }
        }
        if (rc) break;
        G.shape = best_shape;
        TuneVal v{best_shape, {}};
        std::copy(G.us, G.us + kBatchShapes, v.us);
        std::lock_guard<std::mutex> lk(g_tune_mu);
        g_tune[tune_key(b, G, arch)] = v;
    }
    if (d_tmp) (void)hipFree(d_tmp);
    (void)hipFree(d_cnt);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    b.tuned = true;
    return alloc_lists(b);
}

int ensure_tuned(sipg_batch &b) { return b.tuned ? SIPG_OK : tune(b, false); }

int batch_solve(sipg_batch *b, void *x, const void *rhs, const sipg_cg_opts *opts, sipg_cg_info *info, bool pcg) {
    if (!b || !x || !rhs || !opts) return SIPG_ERR_INVALID;
    if (!(opts->tol >= 0.0) || (opts->tol_mode != SIPG_TOL_ABS && opts->tol_mode != SIPG_TOL_REL))
        return SIPG_ERR_INVALID;
    if (!(opts->replace_factor >= 0.0 && opts->replace_factor < 1.0)) return SIPG_ERR_INVALID;
    if (overlaps(x, rhs, (size_t)b->total() * 8)) return SIPG_ERR_ALIAS;
    int rc = ensure_tuned(*b);
    if (rc) return rc;
    sipg_batch &B = *b;
    if (pcg) {
        for (const auto &G : B.groups)
            for (int i : G.inst)
                if (!batch_prec_supported(G.p, B.pen[i], G.shape)) return SIPG_ERR_UNSUPPORTED;
        if (!B.d_zq && (rc = dmalloc(&B.d_zq, (size_t)B.total()))) return rc;
    }
    const int ni = B.ninst;
    double *X = (double *)x;
    const double *Bv = (const double *)rhs;
    const bool true_stop = opts->true_residual_stop != 0;
    std::vector<int> maxit(ni);
    std::vector<batch_ctl> ctl(ni);
    for (int i = 0; i < ni; ++i) {
        const int64_t n = B.off[i + 1] - B.off[i];
        maxit[i] = (int)std::min<int64_t>(opts->maxit > 0 ? (int64_t)opts->maxit : n, INT_MAX);
        ctl[i] = batch_ctl{opts->tol, opts->tol_mode, maxit[i]};
    }
    SIPG_HIP_TRY(hipMemcpy(B.d_ctl, ctl.data(), ni * sizeof(batch_ctl), hipMemcpyHostToDevice));

    // r = b - A x, p = r, z = 0, initial states
    SIPG_HIP_TRY(hipMemsetAsync(B.Z(), 0, (size_t)B.total() * sizeof(double), B.stream));
    if ((rc = apply_set(B, SET_ALL, X, B.AP()))) return rc;
    if ((rc = launch_batch_residual(Bv, B.AP(), B.R(), B.Pv(), B.d_chunks[SET_ALL], B.count(SET_ALL, 1),
                                    B.chunk_cap, B.d_part2, B.d_part3, B.stream)))
        return rc;
    if (pcg && (rc = prec_set(B, SET_ALL, B.Pv(), nullptr, 0))) return rc;      // z = M^-1 r, p = z
    if ((rc = launch_batch_init(B.d_list[SET_ALL], B.count(SET_ALL, 0), ni, B.d_chunk_range[SET_ALL], B.d_part2,
                                B.d_part3, B.d_state, B.d_ctl, pcg ? B.d_item_range[SET_ALL] : nullptr,
                                pcg ? B.d_item_rz : nullptr, B.stream)))
        return rc;
    cg_state *hs0 = B.h_state, *hs1 = B.h_state + ni;
    SIPG_HIP_TRY(hipMemcpyAsync(hs0, B.d_state, ni * sizeof(cg_state), hipMemcpyDeviceToHost, B.stream));
    SIPG_HIP_TRY(hipStreamSynchronize(B.stream));

    enum { ACTIVE, FINISHED };
    std::vector<char> phase(ni, ACTIVE);
    int n_open = 0;
    for (int i = 0; i < ni; ++i) {
        const cg_slot &c = hs0[i].s[0];
        if (hs0[i].err || c.done) phase[i] = FINISHED;      // nothing to do: r = b - A x was exact
        else ++n_open;
    }
    std::vector<char> in_set(ni);
    Lists L;
    auto upload_active = [&]() {
        for (int i = 0; i < ni; ++i) in_set[i] = phase[i] != FINISHED;
        build_lists(B, in_set, L);
        return upload_lists(B, SET_ACTIVE, L);
    };
    std::vector<char> none(ni, 0);
    Lists Q;
    build_lists(B, none, Q);
    if ((rc = upload_active()) || (rc = upload_lists(B, SET_REPL, Q))) return rc;

    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t ev[2] = {nullptr, nullptr};
    auto cleanup = [&]() {
        (void)hipStreamSynchronize(B.stream);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (auto &e : ev)
            if (e) (void)hipEventDestroy(e);
    };
    if (n_open > 0) {
        SIPG_HIP_TRY(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
        SIPG_HIP_TRY(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
    }
    if (n_open > 0 && B.graphs) {
        SIPG_HIP_TRY(hipStreamBeginCapture(B.stream, hipStreamCaptureModeRelaxed));
        for (int it = 0; it < kGraphIters && !rc; ++it) rc = enqueue_iteration(B, X, Bv, it & 1, it == 0, pcg);
        const hipError_t ec = hipStreamEndCapture(B.stream, &graph);
        if (rc || ec != hipSuccess) {
            if (ec != hipSuccess) set_last_hip_error(ec);
            cleanup();
            return rc ? rc : SIPG_ERR_HIP;
        }
        const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
        if (ei != hipSuccess) {
            set_last_hip_error(ei);
            cleanup();
            return SIPG_ERR_HIP;
        }
    }
    // One replay in flight ahead; state copy j is processed while replay
    // j+1 runs, and whatever it decides (replacement set, shorter lists) is
    // enqueued before replay j+2 and copy j+1.
    int64_t replays = 0;
    const int64_t max_replays =
        ((int64_t)*std::max_element(maxit.begin(), maxit.end()) + kGraphIters) / kGraphIters + 64;
    int finished_since = 0;
    int64_t open_work = 0;
    for (int i = 0; i < ni; ++i)
        if (phase[i] != FINISHED) open_work += B.off[i + 1] - B.off[i];
    int64_t done_work = 0;
    std::vector<int> nrep_req(ni, 0);
    // final checks: the smallest true r.r of the failed ones, how many failed
    // checks gained less than 10% on it, and whether one is in flight;
    // kStallChecks (PCG: kStallChecksPcg) such checks, with the best one
    // stall_window iterations old, mean the FP64 floor -- the instance ends
    // there, not converged
    std::vector<double> check_ref(ni, -1.0);
    std::vector<int> check_stale(ni, 0);
    std::vector<int> best_it(ni, 0);      // iteration count at the best failed check
    std::vector<char> check_req(ni, 0);
    std::vector<int> rep_it(ni, 0);       // iteration count when the last true residual was requested
    std::vector<double> rec_min(ni, -1.0);  // lowest polled recursive r.r since then
    bool stopped = false;
    const double rf2 = opts->replace_factor * opts->replace_factor;
    // one replay of the captured iterations, or the same launches made
    // directly when graphs are off
    auto launch_replay = [&]() -> int {
        if (exec) {
            SIPG_HIP_TRY(hipGraphLaunch(exec, B.stream));
        } else {
            for (int it = 0; it < kGraphIters; ++it) {
                const int r = enqueue_iteration(B, X, Bv, it & 1, it == 0, pcg);
                if (r) return r;
            }
        }
        ++replays;
        return SIPG_OK;
    };
    auto copy_states = [&](int k) -> int {
        SIPG_HIP_TRY(hipMemcpyAsync(k ? hs1 : hs0, B.d_state, ni * sizeof(cg_state), hipMemcpyDeviceToHost,
                                    B.stream));
        SIPG_HIP_TRY(hipEventRecord(ev[k], B.stream));
        return SIPG_OK;
    };
    if (n_open > 0) {
        int cur = 0;
        if ((rc = launch_replay()) || (rc = copy_states(0)) || (rc = launch_replay())) {
            cleanup();
            return rc;
        }
        for (;;) {
            const hipError_t es = hipEventSynchronize(ev[cur]);
            if (es != hipSuccess) {          // cleanup() still drains the stream and frees the graph
                set_last_hip_error(es);
                rc = SIPG_ERR_HIP;
                break;
            }
            const cg_state *hs = cur ? hs1 : hs0;
            std::vector<char> repl(ni, 0);
            bool any_repl = false;
            for (int i = 0; i < ni; ++i) {
                if (phase[i] == FINISHED) continue;
                const cg_slot &c = hs[i].s[0];
                if (hs[i].err || c.done) {
                    phase[i] = FINISHED;
                    --n_open;
                    ++finished_since;
                    done_work += B.off[i + 1] - B.off[i];
                    continue;
                }
                // a replacement is requested once the previous one shows in
                // the polled state; the final one once the recursion has
                // fallen 4x below the last true residual (as in sipg_cg_ex)
                if (hs[i].nrep != nrep_req[i]) continue;
                if (check_req[i]) {
                    check_req[i] = 0;
                    const bool gain = check_ref[i] < 0.0 || hs[i].rr_ref <= 0.81 * check_ref[i];
                    if (!gain) ++check_stale[i];
                    else best_it[i] = c.it;
                    if (check_stale[i] >= (pcg ? kStallChecksPcg : kStallChecks) &&
                        c.it - best_it[i] >= stall_window(c.it) && floor_in_reach(hs[i].rr_ref, hs[i].bb)) {
                        phase[i] = FINISHED;
                        stopped = true;
                        --n_open;
                        ++finished_since;
                        done_work += B.off[i + 1] - B.off[i];
                        continue;
                    }
                    if (gain) check_ref[i] = hs[i].rr_ref;
                }
                // PCG: a recursion that does not fall 4x (in r.r) below the
                // last true residual within stall_window iterations has
                // stalled at the FP64 floor without meeting the tolerance
                if (pcg && true_stop && !(c.flags & kSlotWantCheck)) {
                    rec_min[i] = rec_min[i] < 0.0 ? c.rr : std::min(rec_min[i], c.rr);
                    if (floor_in_reach(hs[i].rr_ref, hs[i].bb) && rec_min[i] > 0.25 * hs[i].rr_ref &&
                        c.it - rep_it[i] >= stall_window(c.it)) {
                        phase[i] = FINISHED;
                        stopped = true;
                        --n_open;
                        ++finished_since;
                        done_work += B.off[i + 1] - B.off[i];
                        continue;
                    }
                }
                const bool check = true_stop && (c.flags & kSlotWantCheck) &&
                                   (c.rr <= 0.0625 * hs[i].rr_ref || c.it - rep_it[i] >= force_check_interval(c.it));
                const bool periodic = rf2 > 0.0 && c.rr <= rf2 * hs[i].rr_ref;
                if (check || periodic) {
                    repl[i] = 1;
                    any_repl = true;
                    ++nrep_req[i];
                    check_req[i] = check ? 1 : 0;
                    rep_it[i] = c.it;
                    rec_min[i] = -1.0;
                }
            }
            // the replacement set of replay j+2 (empty most of the time)
            if (any_repl || Q.list.size()) {
                build_lists(B, repl, Q);
                if ((rc = upload_lists(B, SET_REPL, Q))) break;
            }
            // shorter lists once a tenth of the open work has finished (at
            // once for stopped instances, which are not finished on the device)
            if (finished_since > 0 && n_open > 0 && (stopped || 10 * done_work >= open_work)) {
                stopped = false;
                if ((rc = upload_active())) break;
                finished_since = 0;
                open_work -= done_work;
                done_work = 0;
            }
            if ((rc = copy_states(cur ^ 1))) break;
            if (n_open == 0) break;
            if (replays >= max_replays) {
                rc = SIPG_NOT_CONVERGED;
                break;
            }
            if ((rc = launch_replay())) break;
            cur ^= 1;
        }
    }
    cleanup();
    if (rc && rc != SIPG_NOT_CONVERGED) return rc;
    // increments after each instance's last true residual (none unless it
    // stopped on its iteration cap)
    int frc = launch_batch_fold(X, B.Z(), B.d_chunks[SET_ALL], B.count(SET_ALL, 1), B.chunk_cap, B.stream);
    if (frc) return frc;
    SIPG_HIP_TRY(hipStreamSynchronize(B.stream));
    SIPG_HIP_TRY(hipMemcpy(hs0, B.d_state, ni * sizeof(cg_state), hipMemcpyDeviceToHost));
    int status = rc;
    for (int i = 0; i < ni; ++i) {
        const cg_state &f = hs0[i];
        const cg_slot &c = f.s[0];
        int st = SIPG_OK;
        if (f.err == 2) st = SIPG_ERR_BREAKDOWN;
        else if (f.err == 3) st = SIPG_ERR_NONFINITE;
        else if (!(std::sqrt(c.rr) <= f.stop) || (true_stop && !c.checked)) st = SIPG_NOT_CONVERGED;
        if (info) {
            info[i].status = st;
            info[i].iters = c.it;
            info[i].replacements = f.nrep;
            info[i].pad = 0;
            info[i].res_norm = std::sqrt(c.rr);
            info[i].true_res_norm = (c.checked || f.nrep > 0) ? std::sqrt(f.rr_ref) : -1.0;
            info[i].b_norm = std::sqrt(f.bb);
        }
        if (st != SIPG_OK && status == SIPG_OK) status = st;
    }
    return status;
}

}  // namespace

extern "C" {

int sipg_batch_destroy(sipg_batch *b) {
    if (!b) return SIPG_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (auto &G : b->groups)
        for (int s = 0; s < NSETS; ++s)
            if (G.d_items[s]) (void)hipFree(G.d_items[s]);
    void *dev[] = {b->d_inst, b->d_state, b->d_ctl, b->ws, b->d_item_part, b->d_chunk_part,
                   b->d_part2, b->d_part3, b->d_counts, b->d_modes, b->d_item_rr, b->d_item_rz, b->d_zq};
    for (void *q : dev)
        if (q) (void)hipFree(q);
    for (int s = 0; s < NSETS; ++s) {
        void *q[] = {b->d_list[s], b->d_chunks[s], b->d_item_range[s], b->d_chunk_range[s]};
        for (void *v : q)
            if (v) (void)hipFree(v);
        if (b->h_stage[s]) (void)hipHostFree(b->h_stage[s]);
        if (b->stage_ev[s]) (void)hipEventDestroy(b->stage_ev[s]);
    }
    for (auto &kv : b->d_tab) (void)hipFree(kv.second);
    if (b->h_state) (void)hipHostFree(b->h_state);
    if (b->own_stream && b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return SIPG_OK;
}

int sipg_batch_create(sipg_batch **out, int ninst, const int *dimx, const int *dimy, const int *p,
                      const double *pen, void *stream) {
    if (!out) return SIPG_ERR_INVALID;
    *out = nullptr;
    if (ninst < 1 || !dimx || !dimy || !p || !pen) return SIPG_ERR_INVALID;
    for (int i = 0; i < ninst; ++i) {
        if (dimx[i] < 1 || dimy[i] < 1 || p[i] < 1 || !std::isfinite(pen[i])) return SIPG_ERR_INVALID;
        if (p[i] > kMaxDegree) return SIPG_ERR_UNSUPPORTED;
        if ((int64_t)6 * dimx[i] * (p[i] + 1) * (p[i] + 1) > (int64_t)INT_MAX) return SIPG_ERR_UNSUPPORTED;
    }
    sipg_batch *b = new sipg_batch();
    auto fail = [&](int code) { sipg_batch_destroy(b); return code; };
    b->ninst = ninst;
    b->dimx.assign(dimx, dimx + ninst);
    b->dimy.assign(dimy, dimy + ninst);
    b->p.assign(p, p + ninst);
    b->pen.assign(pen, pen + ninst);
    b->off.assign(ninst + 1, 0);
    for (int i = 0; i < ninst; ++i)
        b->off[i + 1] = b->off[i] + (int64_t)dimx[i] * dimy[i] * (p[i] + 1) * (p[i] + 1);
    hipError_t e;
    if (stream) {
        b->stream = (hipStream_t)stream;
    } else {
        // graph capture needs a stream other than the null stream
        if ((e = hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking)) != hipSuccess) {
            set_last_hip_error(e);
            return fail(SIPG_ERR_HIP);
        }
        b->own_stream = true;
    }
    // degree groups in increasing degree
    std::map<int, std::vector<int>> by_p;
    for (int i = 0; i < ninst; ++i) by_p[p[i]].push_back(i);
    b->group_of.assign(ninst, 0);
    for (auto &kv : by_p) {
        sipg_batch::Group G;
        G.p = kv.first;
        G.inst = kv.second;
        G.shape = batch_shape_supported(G.p, 0) ? 0 : 1;
        for (int i : G.inst) b->group_of[i] = (int)b->groups.size();
        b->groups.push_back(G);
    }
    int rc;
    std::vector<batch_inst> hi(ninst);
    for (int i = 0; i < ninst; ++i) hi[i] = batch_inst{b->off[i], dimx[i], dimy[i], pen[i]};
    b->chunk_cap = 0;
    for (int i = 0; i < ninst; ++i) b->chunk_cap += (int)((b->off[i + 1] - b->off[i] + kChunk - 1) / kChunk);
    if ((rc = dmalloc(&b->d_inst, ninst)) || (rc = dmalloc(&b->d_state, ninst)) ||
        (rc = dmalloc(&b->d_ctl, ninst)) ||
        (rc = dmalloc(&b->ws, 4 * (size_t)b->total())) || (rc = dmalloc(&b->d_chunk_part, b->chunk_cap)) ||
        (rc = dmalloc(&b->d_part2, b->chunk_cap)) || (rc = dmalloc(&b->d_part3, b->chunk_cap)) ||
        (rc = dmalloc(&b->d_counts, NSETS * b->ncount())) || (rc = dmalloc(&b->d_modes, 192)))
        return fail(rc);
    for (int s = 0; s < NSETS; ++s) {
        if ((rc = dmalloc(&b->d_list[s], ninst)) || (rc = dmalloc(&b->d_chunks[s], b->chunk_cap)) ||
            (rc = dmalloc(&b->d_item_range[s], ninst)) || (rc = dmalloc(&b->d_chunk_range[s], ninst)))
            return fail(rc);
        if ((e = hipEventCreateWithFlags(&b->stage_ev[s], hipEventDisableTiming)) != hipSuccess) {
            set_last_hip_error(e);
            return fail(SIPG_ERR_HIP);
        }
    }
    if ((e = hipHostMalloc((void **)&b->h_state, 2 * (size_t)ninst * sizeof(cg_state), hipHostMallocDefault)) !=
        hipSuccess) {
        set_last_hip_error(e);
        return fail(SIPG_ERR_NOMEM);
    }
    if ((e = hipMemcpy(b->d_inst, hi.data(), ninst * sizeof(batch_inst), hipMemcpyHostToDevice)) != hipSuccess ||
        (e = hipMemset(b->ws, 0, 4 * (size_t)b->total() * sizeof(double))) != hipSuccess) {
        set_last_hip_error(e);
        return fail(SIPG_ERR_HIP);
    }
    for (auto &G : b->groups) {
        std::vector<double> tab;
        build_tables(G.p, tab);
        double *d = nullptr;
        if ((rc = dmalloc(&d, tab.size()))) return fail(rc);
        b->d_tab[G.p] = d;
        if ((e = hipMemcpy(d, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice)) != hipSuccess) {
            set_last_hip_error(e);
            return fail(SIPG_ERR_HIP);
        }
    }
    if ((rc = alloc_lists(*b))) return fail(rc);
    *out = b;
    return SIPG_OK;
}

int sipg_batch_size(const sipg_batch *b, int64_t *total_n) {
    if (!b || !total_n) return SIPG_ERR_INVALID;
    *total_n = b->total();
    return SIPG_OK;
}

int sipg_batch_offset(const sipg_batch *b, int inst, int64_t *offset, int64_t *n) {
    if (!b || inst < 0 || inst >= b->ninst) return SIPG_ERR_INVALID;
    if (offset) *offset = b->off[inst];
    if (n) *n = b->off[inst + 1] - b->off[inst];
    return SIPG_OK;
}

int sipg_batch_tune(sipg_batch *b) {
    if (!b) return SIPG_ERR_INVALID;
    return tune(*b, true);
}

int sipg_batch_tuning(const sipg_batch *b, int p, int *bx, int *by, double *us_per_apply) {
    if (!b) return SIPG_ERR_INVALID;
    for (auto &G : b->groups) {
        if (G.p != p) continue;
        if (bx) *bx = kBatchShapeBX[G.shape];
        if (by) *by = kBatchShapeBY[G.shape];
        if (us_per_apply) *us_per_apply = b->tuned ? G.us[G.shape] : -1.0;
        return SIPG_OK;
    }
    return SIPG_ERR_NOT_FOUND;
}

int sipg_batch_set_shape(sipg_batch *b, int p, int bx, int by) {
    if (!b) return SIPG_ERR_INVALID;
    int shape = -1;
    for (int k = 0; k < kBatchShapes; ++k)
        if (kBatchShapeBX[k] == bx && kBatchShapeBY[k] == by) shape = k;
    if (shape < 0 || !batch_shape_supported(p, shape)) return SIPG_ERR_UNSUPPORTED;
    for (auto &G : b->groups) {
        if (G.p != p) continue;
        SIPG_HIP_TRY(hipStreamSynchronize(b->stream));
        G.shape = shape;
        b->tuned = true;          // an explicit choice replaces tuning
        return alloc_lists(*b);
    }
    return SIPG_ERR_NOT_FOUND;
}

int sipg_batch_set_graphs(sipg_batch *b, int use_graphs) {
    if (!b || (use_graphs != 0 && use_graphs != 1)) return SIPG_ERR_INVALID;
    b->graphs = use_graphs != 0;
    return SIPG_OK;
}

int sipg_batch_apply(sipg_batch *b, const void *x, void *y) {
    if (!b || !x || !y) return SIPG_ERR_INVALID;
    if (overlaps(x, y, (size_t)b->total() * 8)) return SIPG_ERR_ALIAS;
    int rc = ensure_tuned(*b);
    if (rc) return rc;
    if ((rc = apply_set(*b, SET_ALL, (const double *)x, (double *)y))) return rc;
    SIPG_HIP_TRY(hipStreamSynchronize(b->stream));
    return SIPG_OK;
}

int sipg_batch_rhs_fourier(sipg_batch *b, int inst, int nmodes, const double *a, const int *m, const int *n,
                           void *rhs) {
    if (!b || !rhs || inst < 0 || inst >= b->ninst || nmodes < 0 || nmodes > 64 ||
        (nmodes > 0 && (!a || !m || !n)))
        return SIPG_ERR_INVALID;
    if (b->dimx[inst] != b->dimy[inst]) return SIPG_ERR_INVALID;
    double *da = b->d_modes;
    int *dm = (int *)(b->d_modes + 64);
    int *dn = dm + 64;
    if (nmodes > 0) {
        SIPG_HIP_TRY(hipMemcpyAsync(da, a, nmodes * sizeof(double), hipMemcpyHostToDevice, b->stream));
        SIPG_HIP_TRY(hipMemcpyAsync(dm, m, nmodes * sizeof(int), hipMemcpyHostToDevice, b->stream));
        SIPG_HIP_TRY(hipMemcpyAsync(dn, n, nmodes * sizeof(int), hipMemcpyHostToDevice, b->stream));
    }
    int rc = launch_rhs_fourier(SIPG_F64, b->p[inst], b->dimx[inst], b->d_tab[b->p[inst]], nmodes, da, dm, dn,
                                (double *)rhs + b->off[inst], b->stream);
    if (rc) return rc;
    SIPG_HIP_TRY(hipStreamSynchronize(b->stream));
    return SIPG_OK;
}

int sipg_batch_upload(sipg_batch *b, int inst, const double *host, void *dev) {
    if (!b || !host || !dev || inst < 0 || inst >= b->ninst) return SIPG_ERR_INVALID;
    const int64_t noe = (int64_t)b->dimx[inst] * b->dimy[inst], nq = (int64_t)(b->p[inst] + 1) * (b->p[inst] + 1);
    std::vector<double> tmp(noe * nq);
    for (int64_t e = 0; e < noe; ++e)
        for (int64_t q = 0; q < nq; ++q) tmp[e * nq + q] = host[q * noe + e];
    SIPG_HIP_TRY(hipStreamSynchronize(b->stream));
    SIPG_HIP_TRY(hipMemcpy((double *)dev + b->off[inst], tmp.data(), tmp.size() * 8, hipMemcpyHostToDevice));
    return SIPG_OK;
}

int sipg_batch_download(sipg_batch *b, int inst, const void *dev, double *host) {
    if (!b || !host || !dev || inst < 0 || inst >= b->ninst) return SIPG_ERR_INVALID;
    const int64_t noe = (int64_t)b->dimx[inst] * b->dimy[inst], nq = (int64_t)(b->p[inst] + 1) * (b->p[inst] + 1);
    std::vector<double> tmp(noe * nq);
    SIPG_HIP_TRY(hipStreamSynchronize(b->stream));
    SIPG_HIP_TRY(hipMemcpy(tmp.data(), (const double *)dev + b->off[inst], tmp.size() * 8, hipMemcpyDeviceToHost));
    for (int64_t e = 0; e < noe; ++e)
        for (int64_t q = 0; q < nq; ++q) host[q * noe + e] = tmp[e * nq + q];
    return SIPG_OK;
}

int sipg_batch_cg(sipg_batch *b, void *x, const void *rhs, const sipg_cg_opts *opts, sipg_cg_info *info) {
    return batch_solve(b, x, rhs, opts, info, false);
}

int sipg_batch_pcg(sipg_batch *b, void *x, const void *rhs, const sipg_cg_opts *opts, sipg_cg_info *info) {
    return batch_solve(b, x, rhs, opts, info, true);
}

}  // extern "C"
