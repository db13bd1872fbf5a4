/*
 * fcfd.h — C ABI of libfcfd, the MI355X (gfx950) implementation of the
 * two-phase incompressible Navier-Stokes step of the reference
 * (`Two phase flows/newOpenMP/{2D,3D+IBM}`): WENO5 level-set advection with
 * SSP-RK3, mass-preserving level-set reinitialisation and volume correction,
 * the variable-coefficient (1/rho) pressure Poisson solve and the velocity
 * projection, plus the step glue that closes the time step.
 *
 * The reference has no FFI layer: its hot path is a set of argument-less
 * internal subroutines that mutate the global `type(manager) :: p`
 * (`.../src/main.f90:1-6`).  Here a context object plays the role of `p`:
 * it owns device-resident fields in the reference layout (fcfd_types.h) and
 * every entry point named fcfd_<subroutine> replaces the reference subroutine
 * of that name.  Host arrays are only borrowed by fcfd_push / fcfd_pull.
 * All calls are stream ordered on the context's HIP stream and must come from
 * one host thread per context (the reference calls them from serial code).
 * Every function returns an fcfd_status code; 0 is success.
 */
#ifndef FCFD_H
#define FCFD_H

#include "fcfd_types.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct fcfd_ctx fcfd_ctx;

/* Create a context for one grid on the current HIP device (`manager_init` +
 * `job_init`, `.../3D+IBM/lib/data/tree.f90:108-227`, `branches.f90:70-140`).
 * Fields start at zero. */
int fcfd_create(const fcfd_grid *grid, const fcfd_params *params, fcfd_ctx **out);
int fcfd_destroy(fcfd_ctx *ctx);
/* Batched 2D cases: ncases independent 2D problems on the same grid (grid->dim
 * = 2, nz = 1) in one context.  No counterpart in the reference, which runs
 * one case per process; this is the multi-case context of the batched
 * rising-bubble workload.  Case q is slice k = q of every field: a field
 * holds ncases consecutive (nx+6)(ny+6) planes, each with its own ghosts, and
 * fcfd_push / fcfd_pull move all of them.  Every entry point below advances
 * all cases in the same launches; reductions, the MPLS and reinitialisation
 * stop rules and the pressure solve's stopping test are per case, so each
 * case ends as it would alone.  params holds what the cases share; cases[q]
 * their density / viscosity ratios, Re, We, Fr and MPLS mass (problem_init
 * resets imass). */
int fcfd_create_batch(const fcfd_grid *grid, const fcfd_params *params, int ncases, const fcfd_case *cases,
                      fcfd_ctx **out);
/* Number of cases of a context (1 for fcfd_create). */
int fcfd_batch_size(const fcfd_ctx *ctx);
/* Replace / read the per-case parameters of a batched context (FCFD_ERR_ARG otherwise). */
int fcfd_set_cases(fcfd_ctx *ctx, const fcfd_case *cases);
int fcfd_get_cases(const fcfd_ctx *ctx, fcfd_case *cases);
/* Per-case mass and volume after the last fcfd_step (or fcfd_problem_init)
 * and the MPLS reference mass, into arrays of fcfd_batch_size entries (any
 * may be NULL). */
int fcfd_get_case_diag(fcfd_ctx *ctx, double *mass, double *vol, double *imass);
/* 1 when densities in [rho_min, rho_max] and the averaged face coefficients
 * 2 / (rho_a + rho_b) of the second multigrid level fit FP16's normal range
 * (with a small margin), else 0.  Host-only; no counterpart in the
 * reference.  params.ppe_rho_half is honoured only when the context's
 * density range [min(1, rho_12), max(1, rho_12)] over its cases passes this
 * test; otherwise the FP32 copies are read.  A density pushed by the caller
 * (the standalone Poisson problem) must itself lie in such a range when
 * ppe_rho_half is set. */
int fcfd_ppe_fp16_density_ok(double rho_min, double rho_max);
/* Replace the parameters (dt, tolerances ...) of an existing context. */
int fcfd_set_params(fcfd_ctx *ctx, const fcfd_params *params);
/* Thread-block tile of one stencil kernel (enum fcfd_tile_kernel): tx
 * threads along i (64, 128 or 256) times ty along j, tx * ty <= 256, and the
 * order in which the XCDs take the tiles: slab = 0 plane by plane, slab = S
 * > 0 in slabs of S tile rows, each XCD walking the planes of its slab (the
 * z-stencil's neighbour planes then stay in that XCD's L2).  Default 64 x 4,
 * slab 0.  No counterpart in the reference (its loops are OpenMP over k);
 * the tile autotuner sets these.  FCFD_ERR_ARG for another shape or a slab
 * outside 0..4096; slab may be NULL in fcfd_get_tile. */
int fcfd_set_tile(fcfd_ctx *ctx, int kernel, int tx, int ty, int slab);
int fcfd_get_tile(const fcfd_ctx *ctx, int kernel, int *tx, int *ty, int *slab);
/* Number of float64 values of one ghosted field. */
int64_t fcfd_field_size(const fcfd_ctx *ctx);

/* Host <-> device copies of one full ghosted field (borrowed host arrays). */
int fcfd_push(fcfd_ctx *ctx, int field, const double *host);
int fcfd_pull(fcfd_ctx *ctx, int field, double *host);
/* Device-to-device copies (e.g. from a torch tensor's storage). */
int fcfd_push_device(fcfd_ctx *ctx, int field, const double *dev);
int fcfd_pull_device(fcfd_ctx *ctx, int field, double *dev);
/* Wait for all work queued on the context's stream. */
int fcfd_sync(fcfd_ctx *ctx);

int fcfd_get_diag(fcfd_ctx *ctx, fcfd_diag *out);
int fcfd_set_time(fcfd_ctx *ctx, double time, int iter);
int fcfd_set_imass(fcfd_ctx *ctx, double imass);

/* ---- boundary conditions (`.../3D+IBM/src/others/bcs.f90`) ---- */
/* bc(dat): zero-gradient ghost copy of one field (bcs.f90:113-154). */
int fcfd_bc(fcfd_ctx *ctx, int field);
/* velbc(u,v,w): no-slip / slip ghost rules of vel%now (bcs.f90:156-369). */
int fcfd_velbc(fcfd_ctx *ctx);

/* ---- level-set properties (`.../interface_solver/others.f90`) ---- */
/* ls_funs: H, delta, sign of phi over the ghosted box (others.f90:56-94). */
int fcfd_ls_funs(fcfd_ctx *ctx);
/* rho_mu: ls_funs + rho, mu (others.f90:96-120). */
int fcfd_rho_mu(fcfd_ctx *ctx);
/* ls_mv: rho_mu + mass and volume sums (others.f90:122-176). */
int fcfd_ls_mv(fcfd_ctx *ctx, double *mass, double *vol);

/* ls_mv without allocation: ls_funs + rho_mu over the ghosted box of phi
 * (H, delta, sign, rho, mu; any output may be NULL) and the mass and volume
 * sums into sums[2q], sums[2q + 1] for each case q, with the kernels of
 * fcfd_ls_mv and bit for bit its results.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  It takes
 * no context: every pointer is the caller's device memory in the layout of
 * fcfd_types.h, and the call allocates nothing.  Its reduction partials live
 * in a fixed device array of the library (1.6 MB per device, held until the
 * process ends), so ncases is at most FCFD_LS_MV_NOALLOC_MAX_CASES
 * (FCFD_ERR_ARG beyond); calls that overlap on one device race on that array.
 * grid->dim = 3: one case (ncases = 1, cases = NULL) with params->rho_12 /
 * mu_12.  grid->dim = 2: ncases consecutive planes as a batched context holds
 * them, with the ratios of cases[q] when ncases > 1 (host array).  params
 * supplies ls_wid (and the ratios of a single case).  stream is a
 * hipStream_t (NULL: the null stream); the call only queues work on it. */
#define FCFD_LS_MV_NOALLOC_MAX_CASES 64
int fcfd_ls_mv_noalloc(const fcfd_grid *grid, const fcfd_params *params, int ncases, const fcfd_case *cases,
                       const double *phi, double *heavy, double *delta, double *sign, double *rho, double *mu,
                       double *sums, void *stream);

/* ---- level-set advection (`.../interface_solver/level_set_rk3_solver.f90:1-219`) ---- */
/* level_set_rk3_solver: SSP-RK3 of phi with the cell velocity nvel (fields
 * NU/NV/NW hold nvel%old of the reference); the source is the WENO5-JS flux
 * split, or with params.ls_scheme = FCFD_LS_UCCD the non-conservative
 * upwind-CCD form -(u phi_x + v phi_y + w phi_z) (the commented alternative
 * of level_set_rk3_solver.f90:109,123,137,155-157). */
int fcfd_level_set_rk3_solver(fcfd_ctx *ctx);
/* Factor s on NU/NV/NW in the WENO5 advection (fcfd_level_set_rk3_solver,
 * its FP32 form and its adjoint): the velocity of a step is s (NU, NV, NW).
 * Default 1.  No counterpart in the reference: it lets a prescribed
 * velocity A c(t) u(x) be set once and modulated per step without copies.
 * The UCCD scheme rejects s != 1 (FCFD_ERR_ARG at the advection call). */
int fcfd_set_velocity_scale(fcfd_ctx *ctx, double s);
/* One SSP-RK3 WENO5 step (the scheme of fcfd_level_set_rk3_solver) in FP32
 * on the caller's device arrays of the context's ghosted layout: phi (in/out,
 * ghosts refreshed by the zero-gradient copy after each stage) and the cell
 * velocity nu, nv[, nw] (times the velocity scale).  dt from the params.  No
 * counterpart in the reference (its advection is FP64 only); the first call
 * allocates four FP32 scratch fields in the context.  FCFD_ERR_ARG for the
 * UCCD scheme. */
int fcfd_level_set_rk3_solver_f32(fcfd_ctx *ctx, float *phi, const float *nu, const float *nv, const float *nw);
/* Discrete adjoint of one fcfd_level_set_rk3_solver step (WENO5 scheme, the
 * velocity s (NU, NV, NW) of the context frozen over the step).  No
 * counterpart in the reference: the differentiation entry point of the
 * advection.
 * Device pointers, ghosted field layout:
 *   phi     the level set at the start of the step (interior read);
 *   lambda  in: dJ/dphi after the step; out: dJ/dphi before it (interior
 *           read and written, ghosts untouched);
 *   u/v/wbar  when not NULL, dJ/d(NU, NV, NW) of this step (the stored
 *           fields, i.e. s times the derivative with respect to the
 *           velocity) is ADDED to them (interior).
 * Returns FCFD_ERR_ARG for the UCCD scheme. */
int fcfd_ls_adv_adjoint_step(fcfd_ctx *ctx, const double *phi, double *lambda, double *ubar, double *vbar,
                             double *wbar);
/* The same adjoint in FP32 for the FP32 step (fcfd_level_set_rk3_solver_f32):
 * phi, lambda, u/v/wbar as above in float, and the caller's FP32 cell
 * velocity nu, nv[, nw] of that step (times the context's velocity scale);
 * its reconstructions use the FP32 forward's weights.  No counterpart in the
 * reference.  The first call allocates seven FP32 scratch fields in the
 * context.  FCFD_ERR_ARG for the UCCD scheme. */
int fcfd_ls_adv_adjoint_step_f32(fcfd_ctx *ctx, const float *phi, float *lambda, float *ubar, float *vbar,
                                 float *wbar, const float *nu, const float *nv, const float *nw);

/* root_solve("uccd") (`.../lib/solver/ccd_solvers.f90:161-210,308-341,400-412`)
 * of phi along every interior line, upwind by NU/NV/NW (nvel%old): phi_x ->
 * NX/NY/NZ, the averaged second derivative -> NXX/NYY/NZZ.  The derivative
 * kernel of the UCCD advection source selected by params.ls_scheme. */
int fcfd_uccd(fcfd_ctx *ctx);

/* ---- reinitialisation (`.../interface_solver/level_set_rk3_redis.f90:1-410`) ---- */
/* level_set_rk3_redis(btn): btn = 1 maintains the distance property with the
 * mass-constrained lambda; btn = 0 is the initial redistancing.  *iters gets
 * the pseudo-step count (may be NULL). */
int fcfd_level_set_rk3_redis(fcfd_ctx *ctx, int btn, int *iters);

/* ---- mass correction (`.../interface_solver/mass_preserving_level_set.f90:1-54`) ---- */
int fcfd_mass_preserving_level_set(fcfd_ctx *ctx);

/* ---- CCD derivatives and curvature (`.../interface_solver/others_normals.f90`) ---- */
/* surface_norms: CCD first/second and cross derivatives of phi (others_normals.f90:1-83). */
int fcfd_surface_norms(fcfd_ctx *ctx);
/* curv: surface_norms + curvature (others_normals.f90:133-167). */
int fcfd_curv(fcfd_ctx *ctx);

/* ---- pressure Poisson + projection (`.../ns_solver/ns_ppe_solver.f90`) ---- */
/* ppe_sor_init: build the 1/rho operator hierarchy from rho (ns_ppe_solver.f90:1-65). */
int fcfd_ppe_sor_init(fcfd_ctx *ctx);
/* ppe_sor_solver: src = div(u*)/dt, multigrid solve of the ppe_sor_init
 * operator to ||r||/||f|| <= rel_tol (mean(p) = 0), then
 * ns_momentum_correction, as the reference's ppe_sor_solver ends with it
 * (ns_ppe_solver.f90:67-172).  precision: FCFD_PPE_FP64 or FCFD_PPE_MIXED.
 * Returns FCFD_ERR_NOT_CONVERGED after params.ppe_max_cycles. */
int fcfd_ppe_sor_solver(fcfd_ctx *ctx, double rel_tol, int precision, int *cycles, double *rel_resid);
/* The solve alone: A p = PPE_SRC (minus its mean) with the operator of the
 * last fcfd_ppe_sor_init, starting from P, to ||r||/||f|| <= rel_tol, then
 * mean(p) = 0.  No source evaluation and no projection: the standalone
 * variable-coefficient Poisson problem.  (The reference has no standalone
 * solver; this is the loop body of ppe_sor_solver, ns_ppe_solver.f90:289-361.) */
int fcfd_ppe_solve_rhs(fcfd_ctx *ctx, double rel_tol, int precision, int *cycles, double *rel_resid);
/* ns_momentum_correction: u -= dt grad p / rho_face, then velbc (ns_ppe_solver.f90:174-205). */
int fcfd_ns_momentum_correction(fcfd_ctx *ctx);

/* ---- momentum predictor glue (`.../ns_solver/`) ---- */
/* ns_linearize + ns_ab_setup + ns_ab_predictor (ns_solvers.f90:33-57,
 * admas_bashforth_solver.f90:230-311, ns_ab_adv.f90:1-73, ns_ab_diff.f90:1-198):
 * the viscous / body-force / surface-tension part (sec_part1..3) depends only
 * on start-of-step fields, so it is evaluated once per step by fcfd_switch
 * (and fcfd_problem_init) from PHI_OLD, U_OLD.. and the curvature/normals of
 * the previous ns_init; this call adds the upwind advection of the current
 * iterate: SRC_* gets adv + diff, U/V/W the AB2 predictor. */
int fcfd_ns_ab_predictor(fcfd_ctx *ctx);
/* node_vel: MAC -> cell velocity + nvelbc (ns_solvers.f90:128-147). */
int fcfd_node_vel(fcfd_ctx *ctx);
/* ns_check_convergence_div: max / mean |div u| into the diagnostics. */
int fcfd_ns_check_convergence_div(fcfd_ctx *ctx);
/* ns_solver: ns_init, picard_iters x (predictor, PPE, projection), node_vel
 * (ns_solvers.f90:1-31, admas_bashforth_solver.f90:254-281). */
int fcfd_ns_solver(fcfd_ctx *ctx);

/* ---- whole step ---- */
/* manager_switch (tree.f90:229-252): old <- now for the time-levelled fields,
 * then the step's viscous / body-force / surface-tension source of the
 * momentum predictor (ns_ab_diff.f90:1-198, see fcfd_ns_ab_predictor). */
int fcfd_switch(fcfd_ctx *ctx);
/* mpls(): advection, reinitialisation every redis_interval steps, mass
 * correction (interface_solver.f90:25-34). */
int fcfd_interface_solver(fcfd_ctx *ctx);
/* one pass of the main loop body (main.f90:15-27): time += dt, iter += 1,
 * switch, interface_solver, ns_solver, ls_mv. */
int fcfd_step(fcfd_ctx *ctx);
/* problem_init after the initial condition is pushed (problem_init.f90:109-133
 * without IBM and plotting): bc, velbc, [redis(0)], node_vel, ns_init,
 * p%loc%init, ns_ab_setup, ls_mv, imass = mass. */
int fcfd_problem_init(fcfd_ctx *ctx, int redis0);

/* ---- MPLS_RisingBubble (`MPLS_RisingBubble/20190612/`, 3D, two bubbles) ----
 * The reference's second code base: each bubble keeps its own level set
 * (FCFD_PHI1, FCFD_PHI2), advected with the upwind combined compact scheme
 * (M-DRPCCD coefficients) on interior lines, maintained and corrected to its
 * own mass, then merged by max; phi > 0 is the bubble (rho = rho_12 H + 1 - H,
 * AMURHO.F90:18-22).  Parameters come from fcfd_params: rho_12 / mu_12 =
 * RATIO_RHO / RATIO_AMU, re / we / fr, dt, rdt, ls_wid = INTERFACE_WIDTH,
 * btn_sf / btn_g = ST_FORCE / G_FORCE, no-slip walls (ubc = vbc = wbc = 1),
 * ppe_tol / ppe_precision for the pressure solve, which replaces the
 * reference's SOR (PRESS_POISSON) by the multigrid of fcfd_ppe_solve_rhs on
 * the same operator.  3D contexts only (FCFD_ERR_ARG otherwise). */
int fcfd_m3_get_state(fcfd_ctx *ctx, fcfd_m3_state *s);
int fcfd_m3_set_state(fcfd_ctx *ctx, const fcfd_m3_state *s);
/* AMURHO (AMURHO.F90:1-46): HEAVY, DELTA, RHO, MU of PHI. */
int fcfd_m3_amurho(fcfd_ctx *ctx);
/* LS_CCD (FIND_CURVATURE.F90:99-199): CCD normals NX..NZ, second and cross
 * derivatives, CURV of PHI. */
int fcfd_m3_ls_ccd(fcfd_ctx *ctx);
/* LS_RK3_SOLVER + CELL_FACE_VELOCITY (LS_SOLVER.F90:39-228): SSP-RK3 of PHI
 * with the UCCD source by the cell velocity of U, V, W (into NU, NV, NW). */
int fcfd_m3_ls_rk3_solver(fcfd_ctx *ctx);
/* LS_REDISTANCE(PHI, 1) (LS_REDISTANCE.F90:1-211): pseudo-time steps with the
 * RIS_MASS constraint until the pseudo time exceeds 2 ls_wid. */
int fcfd_m3_ls_redistance(fcfd_ctx *ctx);
/* MPLS(im) (LS_SOLVER.F90:257-321): five corrections of PHI to mass im. */
int fcfd_m3_mpls(fcfd_ctx *ctx, double im);
/* SOLVE_LS (LS_SOLVER.F90:1-37): both bubbles advected, maintained every
 * fifth step (LS_MAINTAIN.F90) and corrected, merged, reinitialised. */
int fcfd_m3_solve_ls(fcfd_ctx *ctx);
/* SOLVE_NS (SOLVE_NS.F90:1-49): AMURHO, curvature, up to four Picard
 * iterations of second-order upwind source + Adams-Bashforth
 * (SECOND_ORDER_UPWIND.F90), pressure solve, projection, no-slip velocity
 * ghosts, with the reference's early exit. */
int fcfd_m3_solve_ns(fcfd_ctx *ctx);
/* FIND_MASS (FIND_MASS.F90:1-162), level-set sums; the first call sets the
 * bubble targets from FCFD_PHI1, FCFD_PHI2. */
int fcfd_m3_find_mass(fcfd_ctx *ctx);
/* STORE_DATA (STORE_DATA.F90:1-33). */
int fcfd_m3_store_data(fcfd_ctx *ctx);
/* the part of SETUP after the initial level sets are pushed (SETUP.F90:
 * 210-222): ghost copies, reinitialisation, AMURHO, STORE_DATA, the first
 * momentum source, STORE_DATA, FIND_MASS. */
int fcfd_m3_setup(fcfd_ctx *ctx);
/* one pass of the main loop (MAIN.F90:10-26 without output): time, iter,
 * SOLVE_LS, SOLVE_NS, FIND_MASS, STORE_DATA. */
int fcfd_m3_step(fcfd_ctx *ctx);

/* ---- multi-block layout (`DomainDecomposition/3D/lib/pointer/ptr_roots.f90`,
 * block ranges of `.../lib/data/branches.f90:84-121`) ----
 * A grid cut into gx x gy x gz blocks (the reference's grid_x, grid_y,
 * grid_z), block (bx, by, bz) owning the 1-based ranges
 * is = bx N / gx + 1 .. ie = (bx + 1) N / gx per axis, each stored as a
 * ghosted array (ghost width ghc; none in z in 2D) in the column-major
 * reference layout.  The blocks of one field lie back to back in a caller's
 * device array of fcfd_blocks_stack_size doubles (a "stack"); a block's
 * offset is given by fcfd_blocks_range.  Every call that takes a stream only
 * queues work on it (NULL: the null stream). */
typedef struct fcfd_blocks fcfd_blocks;
#define FCFD_BLOCKS_MAX_PER_AXIS 16
/* ptrpart_init (`ptr_roots.f90:58-99`) with job_init's ranges: the layout of
 * grid cut into gx x gy x gz blocks (1..FCFD_BLOCKS_MAX_PER_AXIS each; gz = 1
 * in 2D), periodic along the axes whose flag is set (xper, yper, zper of the
 * reference); params->dt is the step of fcfd_blocks_level_set_rk3_solver
 * (params may be NULL otherwise).  FCFD_ERR_ARG when a block is thinner than
 * the ghost width (the reference stops there too, branches.f90:115-118). */
int fcfd_blocks_create(const fcfd_grid *grid, const fcfd_params *params, int gx, int gy, int gz, int xper,
                       int yper, int zper, fcfd_blocks **out);
int fcfd_blocks_destroy(fcfd_blocks *blocks);
/* doubles in one stack (-1 for NULL) */
int64_t fcfd_blocks_stack_size(const fcfd_blocks *blocks);
/* the 1-based interior ranges is, ie, js, je, ks, ke of block (bx, by, bz)
 * and the offset of its array in a stack */
int fcfd_blocks_range(const fcfd_blocks *blocks, int bx, int by, int bz, int32_t range[6], int64_t *offset);
/* every cell of every block (ghosts included) from the ghosted global field
 * at the same index / the blocks' interiors back into the global field. */
int fcfd_blocks_scatter(fcfd_blocks *blocks, const double *field, double *stack, void *stream);
int fcfd_blocks_gather(fcfd_blocks *blocks, const double *stack, double *field, void *stream);
/* job_bc(0) (`.../lib/data/branches.f90:240-290`) on every block: the
 * zero-gradient ghost copy on all six faces of each block. */
int fcfd_blocks_bc(fcfd_blocks *blocks, double *stack, void *stream);
/* ptrpart_sync (`ptr_roots.f90:157-320`): the halo exchange.  Every ghost
 * cell takes the value of the cell of its global index in the block that
 * owns it (wrapped on periodic axes; outside the domain on another axis, the
 * boundary block's own ghost).  Face ghosts equal the reference's after one
 * of its passes; edges and corners equal its settled state (the reference
 * copies them from neighbours that its loop may not have refreshed yet). */
int fcfd_blocks_sync(fcfd_blocks *blocks, double *stack, void *stream);
/* ptrpart_check_periodic (`ptr_roots.f90:322-404`): into *err the sum over
 * the periodic axes of the squared differences between the periodic ghost
 * layers and the cells they mirror (0 after fcfd_blocks_sync); synchronises
 * the stream. */
int fcfd_blocks_check_per(fcfd_blocks *blocks, const double *stack, double *err, void *stream);
/* The WENO5 SSP-RK3 step of fcfd_level_set_rk3_solver run block by block as
 * the multi-block code runs it (`DomainDecomposition/3D/src/interface_solver/
 * level_set_rk3_solver.f90:1-80`: per-block stage, job_bc(0), sync after
 * each stage), on stacks: phi in/out, the cell velocity nu, nv[, nw] (ghosts
 * filled, e.g. by fcfd_blocks_scatter + fcfd_blocks_sync).  Blocks read their
 * neighbours' data through the ghosts of shared faces and the clamped
 * zero-gradient values on physical faces, so on a non-periodic layout the
 * result equals fcfd_level_set_rk3_solver on the whole field bit for bit.
 * The first call allocates four scratch stacks. */
int fcfd_blocks_level_set_rk3_solver(fcfd_blocks *blocks, double *phi, const double *nu, const double *nv,
                                     const double *nw, void *stream);

/* ---- built-in kernel timing (no reference counterpart; replaces the
 * reference's system_clock buckets ls_adv / ls_red / ppe / ns,
 * `.../lib/data/branches.f90:36`, printed by print_CPU_info) ---- */
int fcfd_prof_nregions(void);
const char *fcfd_prof_region_name(int region);
/* on != 0: record a HIP event pair around every kernel region from now on
 * (clears earlier records); on == 0: stop recording. */
int fcfd_prof_enable(fcfd_ctx *ctx, int on);
/* summed milliseconds and launch counts per region since fcfd_prof_enable;
 * n must be >= fcfd_prof_nregions(). */
int fcfd_prof_read(fcfd_ctx *ctx, double *ms, int64_t *count, int n);

#ifdef __cplusplus
}
#endif
#endif /* FCFD_H */
