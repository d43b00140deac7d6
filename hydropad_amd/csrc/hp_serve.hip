// hp_serve.hip -- a request loop over a batch (continuous batching): box
// evolutions arrive at given times, are admitted in arrival order as box
// slots and HBM budget allow, advance one batched step at a time together
// with everything else resident, and retire as soon as their own step count
// is reached, freeing memory for the next arrivals.  Each request's initial
// state is the seeded planar Sod tube, generated on the device.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <thread>
#include <vector>

#include "../../include/hydropad_gpu.h"
#include "hp_batch.h"

namespace {

double quantile(std::vector<double> v, double q)
{
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    size_t k = (size_t)(q * (double)v.size());      // nearest rank, 0-based ceil(q n) - 1
    if (k >= v.size()) k = v.size() - 1;
    if (q * (double)v.size() == (double)k && k > 0) k -= 1;
    return v[k];
}

}  // namespace

extern "C" int hp_serve_run(const hp_request* reqs, int nreq, size_t hbm_budget, int max_boxes, int use_graph,
                            double* const* result_fields, hp_request_result* results, hp_serve_stats* stats)
{
    if (!reqs || nreq < 0 || !results || max_boxes < 1) return HP_E_ARG;
    for (int i = 0; i < nreq; ++i)
        if (reqs[i].n < 3 || reqs[i].nsteps < 1 || (reqs[i].precision != HP_FP64 && reqs[i].precision != HP_MIXED) ||
            !(reqs[i].cour > 0.0) || reqs[i].newton_limit < 0)
            return HP_E_ARG;
    hp_batch* b = nullptr;
    int rc = hp_batch_create(max_boxes, hbm_budget, &b);
    if (rc) return rc;
    if (use_graph) hp_batch_use_graph(b, 1);

    std::vector<int> order(nreq);
    for (int i = 0; i < nreq; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(),
                     [&](int a, int c) { return reqs[a].arrival_s < reqs[c].arrival_s; });
    for (int i = 0; i < nreq; ++i) {
        memset(&results[i], 0, sizeof(results[i]));
        results[i].id = reqs[i].id;
        results[i].arrival_s = reqs[i].arrival_s;
    }
    std::vector<int> owner(max_boxes, -1);
    std::vector<hp_box_status> st(max_boxes);
    // The step loop keeps one step queued ahead: step n+1 is enqueued, with an
    // asynchronous copy of the statuses behind it, before the host waits for
    // step n's copy, so the GPU never idles on a per-step synchronisation.
    // owner_at[k] is the slot -> request map when status copy k was queued: a
    // slot freed and re-admitted since then is not read from that copy.
    std::vector<int> owner_at[2] = {std::vector<int>(max_boxes, -1), std::vector<int>(max_boxes, -1)};
    int cslot = 0;
    bool inflight = false;
    const auto t0 = std::chrono::steady_clock::now();
    auto now = [&]() { return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count(); };
    int next = 0, completed = 0, nactive = 0;
    int64_t steps = 0, max_conc = 0;
    size_t peak = 0;
    double cells = 0.0;
    while (completed < nreq && rc == HP_OK) {
        // admit arrivals in order while slots and budget allow
        while (next < nreq && reqs[order[next]].arrival_s <= now()) {
            const int i = order[next];
            const hp_request& r = reqs[i];
            const int na = r.n + 6;
            if (hp_batch_box_bytes(na, na, na, 0) > hbm_budget) {     // can never fit
                results[i].rc = HP_E_NOMEM;
                results[i].admit_s = results[i].done_s = now();
                ++completed;
                ++next;
                continue;
            }
            int id = -1;
            const int a = hp_batch_add_sod(b, r.n, r.precision, r.seed, r.cour, r.nsteps, &id);
            if (a == HP_E_NOMEM && nactive == 0) {                     // nothing to wait for: the
                results[i].rc = HP_E_NOMEM;                             // device itself is out of memory
                results[i].admit_s = results[i].done_s = now();
                ++completed;
                ++next;
                continue;
            }
            if (a == HP_E_NOMEM) break;                                // wait for memory / a slot
            if (a != HP_OK) { rc = a; break; }
            if (r.newton_limit > 0 && (rc = hp_batch_set_newton_limit(b, id, r.newton_limit)) != HP_OK) break;
            results[i].admit_s = now();
            owner[id] = i;
            ++nactive;
            ++next;
        }
        if (rc) break;
        if (nactive == 0) {
            if (next < nreq) {
                const double wait = reqs[order[next]].arrival_s - now();
                if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 1e-3)));
            }
            continue;
        }
        max_conc = std::max<int64_t>(max_conc, nactive);
        peak = std::max(peak, hp_batch_bytes_used(b));
        rc = hp_batch_step(b, 1);
        if (rc) break;
        ++steps;
        rc = hp::batch_status_async(b, cslot);
        if (rc) break;
        owner_at[cslot] = owner;
        const bool have_prev = inflight;
        const int pslot = cslot ^ 1;
        inflight = true;
        cslot = pslot;
        if (!have_prev) continue;                       // nothing to read yet: queue the next step first
        rc = hp::batch_status_wait(b, pslot, st.data(), nullptr);
        if (rc) break;
        const double tdone = now();
        for (int id = 0; id < max_boxes; ++id) {
            if (owner[id] < 0 || owner_at[pslot][id] != owner[id] || !st[id].done) continue;
            const int i = owner[id];
            const hp_request& r = reqs[i];
            results[i].done_s = tdone;
            results[i].steps = st[id].steps_done;
            results[i].t = st[id].t;
            results[i].newton_nonconv = st[id].newton_nonconv;
            if (st[id].newton_nonconv) results[i].rc = HP_E_NEWTON;   // the reference stops (riemann.F90:60-65)
            if (result_fields && result_fields[5 * i]) {
                rc = hp_batch_download(b, id, result_fields[5 * i], result_fields[5 * i + 1], result_fields[5 * i + 2],
                                       result_fields[5 * i + 3], result_fields[5 * i + 4]);
                if (rc) break;
            }
            rc = hp_batch_remove(b, id);
            if (rc) break;
            owner[id] = -1;
            --nactive;
            ++completed;
            cells += (double)r.n * r.n * r.n * (double)r.nsteps;
        }
    }
    const double wall = now();
    hp_batch_destroy(b);
    if (stats) {
        std::vector<double> lat;
        double sum = 0.0;
        for (int i = 0; i < nreq; ++i)
            if (results[i].rc == HP_OK && results[i].done_s > 0.0) {
                lat.push_back(results[i].done_s - results[i].arrival_s);
                sum += lat.back();
            }
        memset(stats, 0, sizeof(*stats));
        stats->wall_s = wall;
        stats->cell_updates = cells;
        stats->p50_s = quantile(lat, 0.50);
        stats->p99_s = quantile(lat, 0.99);
        stats->mean_s = lat.empty() ? 0.0 : sum / (double)lat.size();
        stats->batched_steps = steps;
        stats->max_concurrent = max_conc;
        stats->peak_bytes = (double)peak;
        stats->completed = (int64_t)lat.size();
    }
    return rc;
}
