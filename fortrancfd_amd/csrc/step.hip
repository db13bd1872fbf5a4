// step.hip — C-ABI entry points named after the reference subroutines and
// the orchestration of one time step (`Two phase flows/newOpenMP/3D+IBM/src/
// main.f90:15-27`, `interface_solver/interface_solver.f90:25-34`,
// `ns_solver/ns_solvers.f90:1-31`, `ns_solver/admas_bashforth_solver.f90:254-281`).
#include <algorithm>
#include <cmath>

#include "fcfd_internal.h"

using namespace fcfd;

namespace fcfd {
int ns_setup_only(fcfd_ctx *c);
int mass_sums(fcfd_ctx *c, const double *phi, int with_lam);
int fetch_result(fcfd_ctx *c, int n);
}  // namespace fcfd

#define CTX_OK(c) \
    if (!(c)) return FCFD_ERR_ARG

extern "C" int fcfd_ls_funs(fcfd_ctx *c) {
    CTX_OK(c);
    return ls_props(c, c->f[FCFD_PHI], c->f[FCFD_HEAVY], c->f[FCFD_DELTA], c->f[FCFD_SIGN], nullptr, nullptr);
}

extern "C" int fcfd_rho_mu(fcfd_ctx *c) {
    CTX_OK(c);
    return ls_props(c, c->f[FCFD_PHI], c->f[FCFD_HEAVY], c->f[FCFD_DELTA], c->f[FCFD_SIGN], c->f[FCFD_RHO],
                    c->f[FCFD_MU]);
}

extern "C" int fcfd_ls_mv(fcfd_ctx *c, double *mass, double *vol) {
    CTX_OK(c);
    return ls_mv(c, mass, vol);
}

extern "C" int fcfd_level_set_rk3_solver(fcfd_ctx *c) {
    CTX_OK(c);
    return ls_advect(c);
}

extern "C" int fcfd_level_set_rk3_redis(fcfd_ctx *c, int btn, int *iters) {
    CTX_OK(c);
    if (btn != 0 && btn != 1) return FCFD_ERR_ARG;
    return redis(c, btn, iters);
}

extern "C" int fcfd_mass_preserving_level_set(fcfd_ctx *c) {
    CTX_OK(c);
    return mpls(c);
}

extern "C" int fcfd_surface_norms(fcfd_ctx *c) {
    CTX_OK(c);
    return surface_norms(c, c->f[FCFD_PHI], 1);
}

extern "C" int fcfd_curv(fcfd_ctx *c) {
    CTX_OK(c);
    FCFD_TRY(surface_norms(c, c->f[FCFD_PHI], 1));
    return curv_kernel(c);
}

extern "C" int fcfd_ppe_sor_init(fcfd_ctx *c) {
    CTX_OK(c);
    return mg_setup(c);
}

extern "C" int fcfd_ppe_solve_rhs(fcfd_ctx *c, double rel_tol, int precision, int *cycles, double *rel) {
    CTX_OK(c);
    if (precision != FCFD_PPE_FP64 && precision != FCFD_PPE_MIXED) return FCFD_ERR_ARG;
    if (!(rel_tol > 0.0)) return FCFD_ERR_ARG;
    return ppe_solve(c, rel_tol, precision, cycles, rel);
}

extern "C" int fcfd_ns_momentum_correction(fcfd_ctx *c) {
    CTX_OK(c);
    return projection(c);
}

extern "C" int fcfd_ppe_sor_solver(fcfd_ctx *c, double rel_tol, int precision, int *cycles, double *rel) {
    CTX_OK(c);
    if (precision != FCFD_PPE_FP64 && precision != FCFD_PPE_MIXED) return FCFD_ERR_ARG;
    if (!(rel_tol > 0.0)) return FCFD_ERR_ARG;
    FCFD_TRY(ppe_src(c));
    int rc = ppe_solve(c, rel_tol, precision, cycles, rel);
    if (rc != FCFD_OK && rc != FCFD_ERR_NOT_CONVERGED) return rc;
    FCFD_TRY(projection(c));
    return rc;
}

extern "C" int fcfd_ns_ab_predictor(fcfd_ctx *c) {
    CTX_OK(c);
    return ns_predictor(c, c->f[FCFD_U], c->f[FCFD_V], c->f[FCFD_W]);
}

extern "C" int fcfd_uccd(fcfd_ctx *c) {
    CTX_OK(c);
    double **F = c->f;
    const int fx[3] = {FCFD_NX, FCFD_NY, FCFD_NZ}, fxx[3] = {FCFD_NXX, FCFD_NYY, FCFD_NZZ};
    const int vel[3] = {FCFD_NU, FCFD_NV, FCFD_NW};
    for (int d = 0; d < c->G.dim; ++d) FCFD_TRY(uccd_dir(c, d, F[FCFD_PHI], F[vel[d]], F[fx[d]], F[fxx[d]]));
    return FCFD_OK;
}

extern "C" int fcfd_node_vel(fcfd_ctx *c) {
    CTX_OK(c);
    return node_vel(c);
}

extern "C" int fcfd_ns_check_convergence_div(fcfd_ctx *c) {
    CTX_OK(c);
    return div_check(c, 1);
}

// manager_switch for the fields whose old level the step reads, plus the
// viscous / body-force source: sec_part1..3 read only old-level fields, which
// equal the current fields at this point of the step.
extern "C" int fcfd_switch(fcfd_ctx *c) {
    CTX_OK(c);
    double **F = c->f;
    FCFD_TRY(copy_field(c, F[FCFD_PHI_OLD], F[FCFD_PHI]));
    FCFD_TRY(copy_field(c, F[FCFD_U_OLD], F[FCFD_U]));
    FCFD_TRY(copy_field(c, F[FCFD_V_OLD], F[FCFD_V]));
    std::swap(F[FCFD_SRC_U_OLD], F[FCFD_SRC_U]);
    std::swap(F[FCFD_SRC_V_OLD], F[FCFD_SRC_V]);
    if (c->G.dim == 3) {
        FCFD_TRY(copy_field(c, F[FCFD_W_OLD], F[FCFD_W]));
        std::swap(F[FCFD_SRC_W_OLD], F[FCFD_SRC_W]);
    }
    return ns_diff_source(c);
}

extern "C" int fcfd_interface_solver(fcfd_ctx *c) {
    CTX_OK(c);
    FCFD_TRY(ls_advect(c));
    int iv = c->pr.redis_interval > 0 ? c->pr.redis_interval : 5;
    c->diag.redis_iters = 0;
    if (c->iter % iv == 0) FCFD_TRY(redis(c, 1, nullptr));
    return mpls(c);
}

// ns_init: rho_mu and, with surface tension on, curv of the current phi
static int ns_init(fcfd_ctx *c) {
    FCFD_TRY(fcfd_rho_mu(c));
    if (c->pr.btn_sf > 0.0) FCFD_TRY(fcfd_curv(c));
    return FCFD_OK;
}

extern "C" int fcfd_ns_solver(fcfd_ctx *c) {
    CTX_OK(c);
    FCFD_TRY(ns_init(c));
    FCFD_TRY(mg_setup(c));
    const int np = c->pr.picard_iters > 0 ? c->pr.picard_iters : 4;
    const double tol = c->pr.ppe_tol > 0.0 ? c->pr.ppe_tol : 1e-10;
    int total = 0, rc = FCFD_OK;
    for (int it = 0; it < np; ++it) {
        // S is stored by the last Picard iteration only (the earlier ones'
        // S would be overwritten unread)
        FCFD_TRY(ns_predictor(c, c->f[FCFD_U], c->f[FCFD_V], c->f[FCFD_W], it == np - 1));
        FCFD_TRY(ppe_src(c));
        int cyc = 0;
        // mean(p) = 0 after the last Picard solve (the earlier ones' p enters
        // only gradients and the next solve's start)
        int r = ppe_solve(c, tol, c->pr.ppe_precision, &cyc, nullptr, it == np - 1);
        if (r != FCFD_OK && r != FCFD_ERR_NOT_CONVERGED) return r;
        if (r != FCFD_OK) rc = r;
        total += cyc;
        FCFD_TRY(projection(c));
    }
    c->diag.ppe_cycles_step = total;
    FCFD_TRY(div_check(c, 0));
    FCFD_TRY(node_vel(c));
    return rc;
}

extern "C" int fcfd_step(fcfd_ctx *c) {
    CTX_OK(c);
    c->time = c->time + c->pr.dt;
    c->iter = c->iter + 1;
    FCFD_TRY(fcfd_switch(c));
    FCFD_TRY(fcfd_interface_solver(c));
    int rc = fcfd_ns_solver(c);
    if (rc != FCFD_OK && rc != FCFD_ERR_NOT_CONVERGED) return rc;
    // ls_mv of the main loop: mass and volume of the new phi
    FCFD_TRY(mass_sums(c, c->f[FCFD_PHI], 0));
    FCFD_TRY(gather_to_host(c, c->diag_host, c->red.result, RES_STRIDE, DIAG_W, c->nb));
    c->diag_pending = 1;
    return rc;
}

extern "C" int fcfd_problem_init(fcfd_ctx *c, int redis0) {
    CTX_OK(c);
    double **F = c->f;
    FCFD_TRY(bc_scalar(c, F[FCFD_PHI]));
    FCFD_TRY(velbc(c, F[FCFD_U], F[FCFD_V], F[FCFD_W], 0));
    if (redis0) FCFD_TRY(redis(c, 0, nullptr));
    FCFD_TRY(node_vel(c));
    FCFD_TRY(ns_init(c));
    // p%loc%init: old = now
    FCFD_TRY(copy_field(c, F[FCFD_PHI_OLD], F[FCFD_PHI]));
    FCFD_TRY(copy_field(c, F[FCFD_U_OLD], F[FCFD_U]));
    FCFD_TRY(copy_field(c, F[FCFD_V_OLD], F[FCFD_V]));
    FCFD_TRY(copy_field(c, F[FCFD_SRC_U_OLD], F[FCFD_SRC_U]));
    FCFD_TRY(copy_field(c, F[FCFD_SRC_V_OLD], F[FCFD_SRC_V]));
    if (c->G.dim == 3) {
        FCFD_TRY(copy_field(c, F[FCFD_W_OLD], F[FCFD_W]));
        FCFD_TRY(copy_field(c, F[FCFD_SRC_W_OLD], F[FCFD_SRC_W]));
    }
    // ns_ab_setup
    FCFD_TRY(ns_diff_source(c));
    FCFD_TRY(ns_setup_only(c));
    // ls_mv, imass = mass
    double mass = 0.0, vol = 0.0;
    FCFD_TRY(ls_mv(c, &mass, &vol));
    c->pr.imass = mass;
    c->diag.imass = mass;
    for (int q = 0; q < c->nb; ++q) {
        c->diag_host[(long)q * DIAG_W] = c->red.host[q * RES_STRIDE];
        c->diag_host[(long)q * DIAG_W + 1] = c->red.host[q * RES_STRIDE + 1];
    }
    if (c->nb > 1) {   // every case of a batched stack keeps its own initial mass
        for (int q = 0; q < c->nb; ++q) c->cases[q].imass = c->red.host[q * RES_STRIDE];
        FCFD_TRY(upload_cases(c));
    }
    c->time = 0.0;
    c->iter = 0;
    return FCFD_OK;
}
