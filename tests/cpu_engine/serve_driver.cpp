// CPU test driver for the serving loop (serve.cpp against the mock runtime
// and the CPU stand-ins of the batched launchers).  Prints JSON lines for
// tests/test_cpu_engine.py.
#include "../../include/sipg_hip.h"

#include <hip/hip_runtime.h>   // the mock runtime (its execution model)

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <map>
#include <vector>


struct Req {
    int dim, p;
    double tol;
    int nm;
    double a[8];
    int m[8], n[8];
};

static std::vector<Req> make_requests(int count) {
    std::vector<Req> out;
    unsigned long long s = 99;
    auto next = [&]() { s = s * 6364136223846793005ULL + 1442695040888963407ULL; return s >> 33; };
    for (int i = 0; i < count; ++i) {
        Req r{};
        r.dim = 3 + (int)(next() % 10);
        r.p = 2 + (int)(next() % 5);
        r.tol = (next() % 2) ? 1e-10 : 1e-8;
        sipg_forcing_draw(1000 + i, 1, 4, 4, &r.nm, r.a, r.m, r.n);
        out.push_back(r);
    }
    return out;
}

// solves `reqs` on one server; returns solutions by request index
static std::map<int, std::vector<double>> serve(const std::vector<Req> &reqs, double budget, int *max_queued,
                                                int *rc_out, std::vector<sipg_cg_info> *infos,
                                                int method = SIPG_METHOD_CG) {
    std::map<int, std::vector<double>> sol;
    sipg_server *S = nullptr;
    int rc = sipg_server_create(&S, budget, nullptr);
    if (rc) { *rc_out = rc; return sol; }
    if ((rc = sipg_server_set_method(S, method))) { *rc_out = rc; sipg_server_destroy(S); return sol; }
    std::map<int64_t, int> index;
    for (size_t i = 0; i < reqs.size(); ++i) {
        const Req &r = reqs[i];
        int64_t id = 0;
        rc = sipg_server_submit(S, r.dim, r.p, 5.0 * (r.p + 1) * (r.p + 1), r.nm, r.a, r.m, r.n, r.tol,
                                SIPG_TOL_REL, 0, &id);
        if (rc) { *rc_out = rc; sipg_server_destroy(S); return sol; }
        index[id] = (int)i;
    }
    *max_queued = 0;
    if (infos) infos->assign(reqs.size(), sipg_cg_info{});
    for (int round = 0; round < 100000 && sol.size() < reqs.size(); ++round) {
        int nd = 0;
        if ((rc = sipg_server_progress(S, &nd))) break;
        int q = 0;
        sipg_server_stats(S, &q, nullptr, nullptr, nullptr);
        if (q > *max_queued) *max_queued = q;
        for (;;) {
            int64_t id = 0;
            sipg_cg_info info{};
            std::vector<double> x(64 * 64 * 49);
            const int t = sipg_server_take(S, &id, &info, x.data(), (int64_t)x.size());
            if (t == SIPG_ERR_NOT_FOUND) break;
            if (t) { rc = t; break; }
            const int k = index[id];
            x.resize((size_t)reqs[k].dim * reqs[k].dim * (reqs[k].p + 1) * (reqs[k].p + 1));
            sol[k] = x;
            if (infos) (*infos)[k] = info;
        }
        if (rc) break;
    }
    *rc_out = rc;
    sipg_server_destroy(S);
    return sol;
}

int main() {
    // asynchronous work runs only when the server synchronises (the
    // runtime's ordering, with host buffers and graphs read late)
    mock_deferred() = true;
    const std::vector<Req> reqs = make_requests(24);
    int64_t demand = 0;
    for (auto &r : reqs) demand += (int64_t)r.dim * r.dim * (r.p + 1) * (r.p + 1);
    // budget for about a third of the total demand (48 bytes per unknown)
    const double budget = 48.0 * demand / 3.0;
    int maxq = 0, rc = 0;
    std::vector<sipg_cg_info> infos;
    auto mix = serve(reqs, budget, &maxq, &rc, &infos);
    printf("{\"case\":\"serve_mix\",\"rc\":%d,\"done\":%zu,\"total\":%zu,\"max_queued\":%d}\n", rc, mix.size(),
           reqs.size(), maxq);
    for (size_t k = 0; k < reqs.size(); ++k) {
        const Req &r = reqs[k];
        std::vector<Req> one{r};
        int mq = 0, rc1 = 0;
        auto alone = serve(one, budget, &mq, &rc1, nullptr);
        double same = -1.0;
        if (mix.count(k) && alone.count(0)) {
            same = 0.0;
            for (size_t i = 0; i < mix[k].size(); ++i) same = std::fmax(same, std::fabs(mix[k][i] - alone[0][i]));
        }
        printf("{\"case\":\"serve_one\",\"k\":%zu,\"rc\":%d,\"status\":%d,\"iters\":%d,\"tol\":%.0e,"
               "\"true_rel\":%.3e,\"alone_maxdiff\":%.3e}\n",
               k, rc1, infos.size() > k ? infos[k].status : -1, infos.size() > k ? infos[k].iters : -1, r.tol,
               infos.size() > k && infos[k].b_norm > 0 ? infos[k].true_res_norm / infos[k].b_norm : -1.0, same);
    }
    // the same stream under PCG (56 bytes per unknown: the same share of
    // the demand fits)
    std::vector<sipg_cg_info> pinfos;
    int pq = 0, prc = 0;
    auto pmix = serve(reqs, 56.0 * demand / 3.0, &pq, &prc, &pinfos, SIPG_METHOD_PCG);
    printf("{\"case\":\"serve_pcg_mix\",\"rc\":%d,\"done\":%zu,\"total\":%zu,\"max_queued\":%d}\n", prc,
           pmix.size(), reqs.size(), pq);
    for (size_t k = 0; k < reqs.size(); ++k) {
        double same = -1.0, dcg = -1.0, ncg = 0.0;
        if (k % 4 == 0) {           // alone vs in the mix, for a sample
            std::vector<Req> one{reqs[k]};
            int mq = 0, rc1 = 0;
            auto alone = serve(one, 56.0 * demand / 3.0, &mq, &rc1, nullptr, SIPG_METHOD_PCG);
            if (pmix.count(k) && alone.count(0)) {
                same = 0.0;
                for (size_t i = 0; i < pmix[k].size(); ++i)
                    same = std::fmax(same, std::fabs(pmix[k][i] - alone[0][i]));
            }
        }
        if (pmix.count(k) && mix.count(k)) {
            dcg = 0.0;
            for (size_t i = 0; i < pmix[k].size(); ++i) {
                dcg += (pmix[k][i] - mix[k][i]) * (pmix[k][i] - mix[k][i]);
                ncg += mix[k][i] * mix[k][i];
            }
            dcg = std::sqrt(dcg / ncg);
        }
        printf("{\"case\":\"serve_pcg_one\",\"k\":%zu,\"status\":%d,\"iters\":%d,\"cg_iters\":%d,"
               "\"tol\":%.0e,\"true_rel\":%.3e,\"rel_vs_cg\":%.3e,\"alone_maxdiff\":%.3e}\n",
               k, pinfos.size() > k ? pinfos[k].status : -1, pinfos.size() > k ? pinfos[k].iters : -1,
               infos.size() > k ? infos[k].iters : -1, reqs[k].tol,
               pinfos.size() > k && pinfos[k].b_norm > 0 ? pinfos[k].true_res_norm / pinfos[k].b_norm : -1.0, dcg,
               same);
    }
    // refusals: a non-standard penalty under PCG; a method change while a
    // request is held
    {
        sipg_server *S = nullptr;
        int rc_pen = -1, rc_busy = -1, rc_create = sipg_server_create(&S, 1e7, nullptr);
        if (!rc_create) {
            sipg_server_set_method(S, SIPG_METHOD_PCG);
            int64_t id = 0;
            const Req &r = reqs[0];
            rc_pen = sipg_server_submit(S, r.dim, r.p, 81.0, r.nm, r.a, r.m, r.n, 1e-8, SIPG_TOL_REL, 0, &id);
            sipg_server_submit(S, r.dim, r.p, 5.0 * (r.p + 1) * (r.p + 1), r.nm, r.a, r.m, r.n, 1e-8,
                               SIPG_TOL_REL, 0, &id);
            rc_busy = sipg_server_set_method(S, SIPG_METHOD_CG);
            sipg_server_destroy(S);
        }
        printf("{\"case\":\"serve_pcg_refusals\",\"create\":%d,\"pen\":%d,\"busy\":%d}\n", rc_create, rc_pen,
               rc_busy);
    }
    // cancellation: one request cancelled in the queue, one while running,
    // the rest served; then a time limit that retires running requests
    for (int variant = 0; variant < 2; ++variant) {
        sipg_server *S = nullptr;
        int64_t need = 0;
        for (int k = 0; k < 6; ++k) need = std::max<int64_t>(need, (int64_t)reqs[k].dim * reqs[k].dim * (reqs[k].p + 1) * (reqs[k].p + 1));
        // room for the largest and a little more: the others queue
        if (sipg_server_create(&S, 48.0 * 1.2 * need, nullptr)) return 1;
        std::vector<int64_t> ids;
        for (int k = 0; k < 6; ++k) {
            const Req &r = reqs[k];
            int64_t id = 0;
            sipg_server_submit(S, r.dim, r.p, 5.0 * (r.p + 1) * (r.p + 1), r.nm, r.a, r.m, r.n, 1e-10, SIPG_TOL_REL,
                               0, &id);
            ids.push_back(id);
        }
        int c_queued = -1, c_running = -1, c_again = -1, c_unknown = -1;
        if (variant == 0) c_queued = sipg_server_cancel(S, ids[5]);       // still queued
        int nd = 0;
        sipg_server_progress(S, &nd);                                       // admits and runs a replay
        if (variant == 0) {
            c_running = sipg_server_cancel(S, ids[0]);                      // running
            c_again = sipg_server_cancel(S, ids[5]);                        // already completed
            c_unknown = sipg_server_cancel(S, 987654);
        } else {
            sipg_server_set_timeout(S, 1e-12);                              // everything is overdue now
        }
        std::map<int64_t, std::pair<sipg_cg_info, sipg_request_times>> got;
        for (int round = 0; round < 100000 && got.size() < ids.size(); ++round) {
            if (sipg_server_progress(S, &nd)) break;
            for (;;) {
                int64_t id = 0;
                sipg_cg_info info{};
                sipg_request_times t{};
                std::vector<double> x(64 * 64 * 49);
                if (sipg_server_take_ex(S, &id, &info, x.data(), (int64_t)x.size(), &t)) break;
                got[id] = {info, t};
            }
        }
        for (size_t k = 0; k < ids.size(); ++k) {
            const auto &g = got[ids[k]];
            printf("{\"case\":\"serve_%s\",\"k\":%zu,\"status\":%d,\"iters\":%d,\"submitted\":%.6f,\"admitted\":%.6f,"
                   "\"completed\":%.6f,\"c_queued\":%d,\"c_running\":%d,\"c_again\":%d,\"c_unknown\":%d}\n",
                   variant == 0 ? "cancel" : "timeout", k, g.first.status, g.first.iters, g.second.submitted,
                   g.second.admitted, g.second.completed, c_queued, c_running, c_again, c_unknown);
        }
        sipg_server_destroy(S);
    }
    // nothing may be left queued once every call has returned
    printf("{\"case\":\"deferred_queue\",\"pending\":%zu}\n", mock_pending());
    return 0;
}
