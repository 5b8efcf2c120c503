// Autotuner of the tree evaluators' occupancy choice (lanes per instance).
//
// For a batch of different trees the level-parallel evaluator gives each
// instance G lanes of a wavefront; how many pays depends on the trees' widths
// and depths and on the batch size.  tdyn_batch_autotune times short probe
// rollouts of the caller's own initial state under each candidate setting
// (fixed lanes, ranges of the automatic per-instance choice, the one-lane
// streaming evaluator, the tree sweeps), keeps the fastest, and remembers it
// in a cache file keyed by the batch's shape, so a later batch of the same
// shape is set up without probing.
#include "tdyn_host.hpp"

#include <cstdio>
#include <cstring>
#include <string>
#include <vector>
#include <algorithm>

using namespace tdyn;
using namespace tdyn::host;

namespace {

struct Candidate { int mode, lanes, lmin, lmax; };

int ilog2(double x) { int k = 0; while (x >= 2.0) { x *= 0.5; ++k; } return k; }

// Shape of the batch: instances, mean bodies, deepest body and mean widest
// level (log2 buckets), whether trees differ, whether they carry the extended
// bodies (reversed, Euler-angle, BendStretch, Bushing) and whether known udot
// is set (which rules out the streaming evaluator).
std::string shape_key(const tdyn_batch* b) {
    double w = 0;
    for (int t = 0; t < b->n_trees; ++t) w += b->h_tree_maxw.empty() ? 1 : b->h_tree_maxw[t];
    w /= std::max(1, b->n_trees);
    char buf[192];
    snprintf(buf, sizeof buf, "tree-lanes n%d nb%d depth%d width%d trees%d ext%d known%d", ilog2(b->n_inst),
             ilog2((double)b->total_b / b->n_inst - 1.0), ilog2(b->dmax + 1.0), ilog2(w),
             b->n_trees > 1 ? 1 : 0, (b->any_rev || b->any_euler || b->any_bendstretch || b->any_bushing) ? 1 : 0,
             b->known_p ? 1 : 0);
    return buf;
}

bool cache_lookup(const char* path, const std::string& key, Candidate& c) {
    if (!path) return false;
    FILE* f = fopen(path, "r");
    if (!f) return false;
    char line[512];
    bool found = false;
    while (fgets(line, sizeof line, f)) {
        char* tab = strchr(line, '\t');
        if (!tab) continue;
        *tab = 0;
        Candidate x;
        if (key == line && sscanf(tab + 1, "%d %d %d %d", &x.mode, &x.lanes, &x.lmin, &x.lmax) == 4) {
            c = x; found = true;          // the last entry for a key wins
        }
    }
    fclose(f);
    return found;
}

void cache_store(const char* path, const std::string& key, const Candidate& c) {
    if (!path) return;
    FILE* f = fopen(path, "a");
    if (!f) return;
    fprintf(f, "%s\t%d %d %d %d\n", key.c_str(), c.mode, c.lanes, c.lmin, c.lmax);
    fclose(f);
}

int apply(tdyn_batch* b, const Candidate& c) {
    int rc = tdyn_batch_set_evaluator(b, c.mode);
    if (!rc) rc = tdyn_batch_set_tree_lanes(b, c.lanes);
    if (!rc) rc = tdyn_batch_set_tree_lane_range(b, c.lmin, c.lmax);
    return rc;
}

// One probe rollout to t_probe; its duration in ms (events on the stream).
int probe(tdyn_batch* b, const double* q, const double* u, const double* fapp, const tdyn_integrator_cfg* cfg,
          double t_probe, hipStream_t s, float& ms) {
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return fail(TDYN_E_HIP, "hipEventCreate");
    if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); return fail(TDYN_E_HIP, "hipEventCreate"); }
    int rc = TDYN_OK;
    std::vector<int> st(b->n_inst);
    if (hipEventRecord(e0, s) != hipSuccess) rc = fail(TDYN_E_HIP, "hipEventRecord");
    if (!rc) rc = tdyn_rollout_begin(b, q, u, fapp, 0.0, nullptr, t_probe, cfg, s);
    for (int round = 0; !rc && round < 1000; ++round) {
        rc = tdyn_rollout_advance(b, 1 << 20, s);
        if (rc) break;
        if (hipMemcpyAsync(st.data(), b->status.p, st.size() * sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess) { rc = fail(TDYN_E_HIP, "probe status"); break; }
        bool running = false;
        for (int x : st) running |= x == 0;
        if (!running) break;
    }
    if (!rc && (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess))
        rc = fail(TDYN_E_HIP, "probe timing");
    (void)hipEventDestroy(e0); (void)hipEventDestroy(e1);
    return rc;
}

} // namespace

int tdyn_batch_autotune(tdyn_batch* b, const double* q, const double* u, const double* fapp,
                        const tdyn_integrator_cfg* cfg, double t_probe, const char* cache_path,
                        int32_t* chosen, hipStream_t s) {
    if (!b || !q || !u || !cfg) return fail(TDYN_E_ARG, "tdyn_batch_autotune: batch, q, u and cfg are required");
    if (!(t_probe > 0)) return fail(TDYN_E_ARG, "tdyn_batch_autotune: t_probe must be positive");
    std::vector<Candidate> cands;
    if (b->motion || b->any_cons || b->any_wit) {
        // prescribed motion, constraints and event witnesses run on the tree
        // sweeps only
        cands.push_back({1, 0, 0, 32});
    } else if (!b->stream_ok) {
        // one model for the batch (chain evaluator or tree sweeps): nothing
        // to choose
        cands.push_back({0, 0, 0, 32});
    } else {
        cands = {{3, 0, 16, 32}, {3, 0, 8, 32}, {3, 0, 16, 64}, {3, 0, 4, 16},
                 {3, 8, 0, 32}, {3, 16, 0, 32}, {3, 32, 0, 32}, {3, 1, 0, 32}, {1, 0, 0, 32}};
        // known udot: not on the one-lane streaming evaluator
        if (b->known_p) cands.erase(std::remove_if(cands.begin(), cands.end(),
                                                   [](const Candidate& c) { return c.lanes == 1; }), cands.end());
    }
    const std::string key = shape_key(b);
    Candidate best = cands[0];
    int from_cache = 0;
    Candidate hit{};
    auto same = [&](const Candidate& c) {
        return c.mode == hit.mode && c.lanes == hit.lanes && c.lmin == hit.lmin && c.lmax == hit.lmax;
    };
    // a cached choice is used only if it is one of this batch's candidates
    if (cands.size() > 1 && cache_lookup(cache_path, key, hit) && std::any_of(cands.begin(), cands.end(), same)) {
        best = hit;
        from_cache = 1;
    } else if (cands.size() > 1) {
        float best_ms = 0;
        for (size_t i = 0; i < cands.size(); ++i) {
            int rc = apply(b, cands[i]);
            if (rc) return rc;
            float ms = 0, m2 = 0;
            rc = probe(b, q, u, fapp, cfg, t_probe, s, ms);
            if (!rc) rc = probe(b, q, u, fapp, cfg, t_probe, s, m2);
            if (rc) return rc;
            ms = std::min(ms, m2);
            if (i == 0 || ms < best_ms) { best_ms = ms; best = cands[i]; }
        }
        cache_store(cache_path, key, best);
    }
    const int rc = apply(b, best);
    if (rc) return rc;
    b->rollout_ready = false;       // the probes' state is not the caller's
    if (chosen) { chosen[0] = best.mode; chosen[1] = best.lanes; chosen[2] = best.lmin; chosen[3] = best.lmax;
                  chosen[4] = from_cache; }
    return TDYN_OK;
}
