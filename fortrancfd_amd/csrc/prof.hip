// prof.hip — built-in kernel timing with HIP events on the context stream.
// Enabled by fcfd_prof_enable; fcfd_prof_read sums the elapsed time and the
// launch count per kernel region.  Used by bench.py for the live roofline.
#include <algorithm>

#include "fcfd_internal.h"

using namespace fcfd;

namespace fcfd {

ProfScope::ProfScope(fcfd_ctx *c_, int reg_) : c(c_), reg(reg_), ev(-1) {
    if (!c->prof.on) return;
    Prof &P = c->prof;
    if (P.used + 2 > P.pool.size()) {
        size_t grow = std::max<size_t>(256, P.pool.size());
        for (size_t i = 0; i < grow; ++i) {
            hipEvent_t e;
            if (hipEventCreate(&e) != hipSuccess) return;
            P.pool.push_back(e);
        }
    }
    ev = (int)P.used;
    P.used += 2;
    hipEventRecord(P.pool[ev], c->stream);
}

ProfScope::~ProfScope() {
    if (ev < 0) return;
    hipEventRecord(c->prof.pool[ev + 1], c->stream);
    if (c->capture_kind >= 0)   // inside a V-cycle graph: read back after each replay
        c->vgraph_prof[c->capture_kind][c->capture_inner].push_back({reg, ev});
    else
        c->prof.recs.push_back({reg, ev});
}

}  // namespace fcfd

static const char *kRegionNames[NREG] = {
    "ls_adv_stage", "redis_pass_a", "redis_pass_b", "mass_sums", "mpls_update", "ccd_lines", "curv",
    "diff_source", "predictor", "ppe_src", "projection", "node_vel", "bc_ghost", "mg_smooth_l0",
    "mg_smooth_coarse", "mg_restrict_l0", "mg_restrict_coarse", "mg_prolong_l0", "mg_prolong_coarse",
    "mg_resid_l0", "mg_coarse_cg", "fcg_vectors", "misc"};

extern "C" int fcfd_prof_nregions(void) { return NREG; }

extern "C" const char *fcfd_prof_region_name(int r) { return (r >= 0 && r < NREG) ? kRegionNames[r] : ""; }

namespace fcfd {
int mg_drop_graphs(fcfd_ctx *c);
// fold the event pairs of a just-replayed V-cycle graph into the totals
int prof_graph_accum(fcfd_ctx *c, int kind, int inner) {
    if (!c->prof.on) return FCFD_OK;
    for (auto &pr : c->vgraph_prof[kind][inner]) {
        float t = 0.f;
        FCFD_CHECK(hipEventElapsedTime(&t, c->prof.pool[pr.second], c->prof.pool[pr.second + 1]));
        c->prof_gms[pr.first] += t;
        c->prof_gcnt[pr.first] += 1;
    }
    return FCFD_OK;
}
}  // namespace fcfd

extern "C" int fcfd_prof_enable(fcfd_ctx *c, int on) {
    if (!c) return FCFD_ERR_ARG;
    FCFD_CHECK(hipStreamSynchronize(c->stream));
    // captured V-cycles hold events of the pool (or none): capture again
    FCFD_TRY(mg_drop_graphs(c));
    c->prof.on = on;
    c->prof.recs.clear();
    c->prof.used = 0;
    for (int r = 0; r < NREG; ++r) { c->prof_gms[r] = 0.0; c->prof_gcnt[r] = 0; }
    return FCFD_OK;
}

extern "C" int fcfd_prof_read(fcfd_ctx *c, double *ms, int64_t *count, int n) {
    if (!c || !ms || !count || n < NREG) return FCFD_ERR_ARG;
    FCFD_CHECK(hipStreamSynchronize(c->stream));
    for (int r = 0; r < NREG; ++r) { ms[r] = c->prof_gms[r]; count[r] = c->prof_gcnt[r]; }
    for (const ProfRec &R : c->prof.recs) {
        float t = 0.f;
        FCFD_CHECK(hipEventElapsedTime(&t, c->prof.pool[R.ev], c->prof.pool[R.ev + 1]));
        ms[R.reg] += t;
        count[R.reg] += 1;
    }
    return FCFD_OK;
}
