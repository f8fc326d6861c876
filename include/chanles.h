/*
 * chanles.h -- C ABI of the MI355X channel-flow LES time step.
 *
 * One context holds one channel resident on one GPU.  Host arrays use the
 * reference's Fortran shapes and column-major order (x fastest), so a
 * Fortran caller passes its arrays unchanged (see the iso_c_binding module
 * channel_rl_amd/fortran/chanles_mod.f90):
 *   U (nx,nyg,nzg)  V (nxg,ny,nzg)  W (nxg,nyg,nz)  nu_t, P (nxg,nyg,nzg)
 *   alpha_x (nx,2,nzg)      with nxg = nx+1, nyg = ny+1, nzg = nz+1
 *   rhs_u (2:nx-1,2:nyg-1,2:nzg-1)  rhs_v (2:nxg-1,2:ny-1,2:nzg-1)
 *   rhs_w (2:nxg-1,2:nyg-1,2:nz-1)                 (interior, packed)
 * (code/initialization.f90:133-136,156-159,467,476).
 *
 * Every call returns CHAN_OK (0) or a negative status; chan_last_error()
 * gives the message.  The Fortran wrappers turn a non-zero status into
 * `Stop`, the reference's error convention (code/initialization.f90:66-74).
 * Calls are enqueued on the context's stream; calls that return a value to
 * the host synchronise.  A context is not re-entrant.
 */
#ifndef CHANLES_H
#define CHANLES_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct chan_ctx chan_ctx;
typedef struct chan_mlp chan_mlp;
typedef struct chan_ensemble chan_ensemble;

enum { CHAN_OK = 0, CHAN_EINVAL = -1, CHAN_EHIP = -2, CHAN_ESTATE = -3, CHAN_ENOMEM = -4 };

/* arithmetic of the bulk (stencil/SGS/RK) arrays; the Poisson path is FP64 */
enum { CHAN_FP64 = 0, CHAN_FP32_BULK = 1 };

/* field selectors for chan_set_field / chan_get_field */
enum {
    CHAN_U = 0, CHAN_V = 1, CHAN_W = 2, CHAN_NU_T = 3, CHAN_P = 4, CHAN_ALPHA_X = 5,
    CHAN_RHS_U = 6, CHAN_RHS_V = 7, CHAN_RHS_W = 8
};

/* wall-stress source for the wall model (the reference's alpha_x hook) */
enum { CHAN_WALL_GIVEN = 0, CHAN_WALL_EQUILIBRIUM = 1, CHAN_WALL_MLP = 2 };

/* The reference's input-file parameters (code/input_output.f90:20-91) plus
 * the grid (faces; code/input_output.f90:108-124 / the .init mesh). */
typedef struct {
    int nx, ny, nz;            /* face counts, the input line `nx ny nz`   */
    const double* x;           /* nx face coordinates                      */
    const double* y;           /* ny                                       */
    const double* z;           /* nz                                       */
    double nu;                 /* viscosity                                */
    double CFL;                /* CFL, or fixed dt = -CFL when negative    */
    double dPdx;               /* pressure gradient                        */
    int x_mass_cte;            /* 1: constant mass flow in x by the
                                  reference's functional (the mean of U on
                                  rows 3 and nyg-2); 2: constant bulk flow
                                  (no reference counterpart); 0: none      */
    int LES_model;             /* 0 none, 2 dynamic Smagorinsky, 3 AMD     */
    int Dirichlet_nu_t;        /* 1 Dirichlet, 0 Neumann ghost for nu_t,
                                  2 Dirichlet with wall offset (AMD)       */
    int precision;             /* CHAN_FP64 or CHAN_FP32_BULK              */
    int device;                /* HIP device ordinal                       */
} chan_config;

/* Allocates the device state and sets up the constants of
 * code/initialization.f90:187-427 (grids, kxx/kzz, Dyy, weights, RK3).
 * nx-2 and nz-2 (the periodic points) must factor into 2, 3 and 5 and be at
 * most 2048 (CHAN_EINVAL otherwise).
 * On failure *out still holds a context carrying the message for
 * chan_last_error; release it with chan_destroy. */
int  chan_create(const chan_config* cfg, chan_ctx** out);
void chan_destroy(chan_ctx* ctx);
const char* chan_last_error(const chan_ctx* ctx);

/* host <-> device copies in the shapes listed above */
int chan_set_field(chan_ctx* ctx, int which, const double* src);
int chan_get_field(chan_ctx* ctx, int which, double* dst);
/* scalars: dt, t, dPdx, Qflow_x_0 */
int chan_set_scalars(chan_ctx* ctx, double dt, double t, double dPdx, double Qflow_x_0);
int chan_get_scalars(chan_ctx* ctx, double out4[4]);

/* code/boundary_conditions.f90:20 apply_boundary_conditions (on U,V,W) */
int chan_apply_boundary_conditions(chan_ctx* ctx);
/* code/subgrid.f90:23 compute_eddy_viscosity(U,V,W,nu_t): LES_model 2 dynamic
 * Smagorinsky (code/subgrid.f90:379-537), 3 AMD (code/subgrid.f90:632-874) */
int chan_compute_eddy_viscosity(chan_ctx* ctx);
/* code/wallmodel.f90:23 compute_wall_model(U,V,W,nu_t) from alpha_x */
int chan_compute_wall_model(chan_ctx* ctx);
/* code/equations.f90:24,174,327 compute_rhs_u/_v/_w into CHAN_RHS_* */
int chan_compute_rhs(chan_ctx* ctx);
/* code/projection.f90:20 compute_projection_step (U,V,W updated, P kept) */
int chan_compute_projection_step(chan_ctx* ctx);
/* code/projection.f90:74 solve_poisson_equation on P in place */
int chan_solve_poisson_equation(chan_ctx* ctx);
/* code/pressure.f90:21 compute_pressure: the true pressure of the current
 * U, V, W and nu_t into CHAN_P (Laplacian(p) = div of the momentum
 * right-hand sides, dP/dy = d/dy((nu + nu_t) dV/dy) at the walls), with
 * every periodic image and wall ghost row set.  The next projection
 * overwrites CHAN_P with the pseudo-pressure. */
int chan_compute_pressure(chan_ctx* ctx);
/* chan_compute_pressure without any allocation, for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No reference counterpart.  The wall data
 * that chan_compute_pressure allocates on its first call live instead in one
 * library-wide device array of 2^21 doubles (16 MiB) per GPU, held until the
 * process ends: a channel needs 2*nxg*nzg of them (CHAN_EINVAL beyond).  All
 * calls on one GPU share that array, so calls that overlap on one GPU (from
 * any contexts or threads) race; the results of a call do not depend on what
 * an earlier call left there. */
int chan_compute_pressure_noalloc(chan_ctx* ctx);
/* code/mass_flow.f90:24 compute_mean_mass_flow_U(U,Q) */
int chan_compute_mean_mass_flow_U(chan_ctx* ctx, double* Q);
/* code/initialization.f90:442: Qflow_x_0 = compute_mean_mass_flow_U(U) */
int chan_initialize_mass_flow(chan_ctx* ctx);
/* code/mass_flow.f90:113 compute_dPx_for_constant_mass_flow(U,dPdx) */
int chan_compute_dPx_for_constant_mass_flow(chan_ctx* ctx, double* dPdx);
/* code/time_integration.f90:102 compute_dt; dt may be NULL (no host sync) */
int chan_compute_dt(chan_ctx* ctx, double* dt);
/* code/time_integration.f90:24 compute_time_step_RK3 (graph-captured) */
int chan_compute_time_step_RK3(chan_ctx* ctx);
/* code/projection.f90:223 check_divergence: signed max and max |div| */
int chan_check_divergence(chan_ctx* ctx, double* max_div, double* max_abs_div);

/* Wall-stress source: CHAN_WALL_GIVEN keeps alpha_x as set by the host (the
 * reference's RL hook, code/smarties_stat.f90:231-258); CHAN_WALL_EQUILIBRIUM
 * evaluates the equilibrium log law U+ = ln(y+)/kappa + B at height h_wm
 * (h_wm <= 0: first off-wall cell) before each substep's wall model.  No
 * reference counterpart: the log-law closure is new. */
int chan_set_wall_model(chan_ctx* ctx, int mode, double h_wm, double kappa, double B);
/* Limit on the wall model's viscosity (no reference counterpart; default 0:
 * the reference's formula unchanged).  With limit > 0 the total viscosity
 * alpha_x / (dU/dy) that code/wallmodel.f90:30-35 sets at each wall face is
 * held in [0, limit (yg(2)-yg(1))^2 / dt]: where the resolved wall gradient
 * vanishes or reverses the formula asks for an unbounded or negative
 * viscosity, and the explicit step then diverges. */
int chan_set_wall_viscosity_limit(chan_ctx* ctx, double limit);
/* evaluate the equilibrium wall stress into alpha_x once */
int chan_compute_equilibrium_wall_stress(chan_ctx* ctx, double h_wm, double kappa, double B);

/* code/statistics.f90:52-64: xz-plane means over unique points, rows
 * j=1..ny of U, V, W, U^2, V^2, W^2, UV (out: 7*ny doubles). */
int chan_plane_statistics(chan_ctx* ctx, double* out);

/* code/statistics.f90:95-130: mean wall shear (nu + nu_t) dU/dy over each
 * wall (nu_t of the context, as left by the last step), u_tau =
 * sqrt((tau_b + tau_t) / Ly) and Re_tau = u_tau Ly / 2 / nu.  The reference
 * sums the bottom wall twice; the top wall uses its own rows here.
 * out: tau_bottom, tau_top, u_tau, Re_tau. */
int chan_wall_shear(chan_ctx* ctx, double out4[4]);

/* code/smarties_stat.f90:24-178 on the device: the states (6 per agent and
 * wall) and rewards of nagents agents per wall, agent a at x face
 * a*(nx-2)/nagents + 2, z centre 2, from U and nu_t of the current state.
 * draws (2*nagents*2): per (wall, agent), wall-major, the two uniform numbers
 * the reference draws with RANDOM_NUMBER: the first < 1/2 picks sampling row
 * 3, else 4; the second weighs that row against the next.  states
 * [wall][agent][6] = (dU l / u_tau, U/u_tau - dU l / u_tau ln(l u_tau / nu),
 * U, l, u_tau, dPdx_ref); rewards [wall][agent].  nagents must divide
 * nx-2.  Replaces the per-step host copy of the wall planes. */
int chan_rl_observe(chan_ctx* ctx, int nagents, double dPdx_ref, const double* draws, double* states,
                    double* rewards);
/* code/smarties_stat.f90:190,213,231-258: actions = alpha_x at each agent
 * times factors [wall][agent], interpolated linearly in x between agents
 * (periodic), x images, then the z interpolation of the reference, into
 * alpha_x on the device. */
int chan_rl_act(chan_ctx* ctx, int nagents, const double* factors);

/* code/app_main.f90:108-118: nsteps x (compute_dt; compute_time_step_RK3)
 * without host synchronisation. */
int chan_run_steps(chan_ctx* ctx, int nsteps);
int chan_synchronize(chan_ctx* ctx);

/* Measurement (the reference times nmonitor steps with MPI_WTIME,
 * code/monitor.f90:57,95,102; here per kernel): runs nsteps of
 * compute_dt + compute_time_step_RK3 without graph capture, bracketing every
 * kernel launch with HIP events on the context's stream, and keeps the
 * per-kernel totals, read back with chan_profile_count / chan_profile_entry. */
int chan_profile_steps(chan_ctx* ctx, int nsteps);
int chan_profile_count(const chan_ctx* ctx);
int chan_profile_entry(const chan_ctx* ctx, int idx, char* name, int len, double* total_ms, long* launches);
/* Measurement: a plain streaming copy of nbytes on the context's device,
 * timed over reps launches with HIP events (achievable HBM bandwidth, and
 * the reference point that scales the hardware traffic counters). */
int chan_stream_copy(chan_ctx* ctx, long nbytes, int reps, double* ms_per_copy);

/* ---------------------------------------------------------------------
 * The learned wall-stress closure.  No reference counterpart: it produces
 * the alpha_x the reference receives from its RL agent
 * (code/smarties_stat.f90:231-258).  A 4 -> 64 -> 64 -> 64 -> 2 network with
 * SiLU activations, BF16 weights and FP32 accumulation, evaluated and trained
 * by the library's own kernels (matrix-core hidden layers, Adam).
 *
 * Per wall cell (x face i, z centre k, wall; the entries of alpha_x) the
 * features are, before standardisation,
 *   f0 = asinh(h|u_par|/nu/10), f1 = asinh(h^2 d|u_par|/dn/nu/10),
 *   f2 = asinh(y+_prev/10) with y+_prev = h sqrt(|alpha_x|)/nu (the previous
 *   wall stress), f3 = asinh(h^3 dpdx/nu^2/10),
 * and the outputs o0 = log(1 + h^2 tau_par/nu^2), o1 = h^2 tau_perp/nu^2 /
 * exp(o0): the stress along and across u_par, rotated back to
 * (tau_wx, tau_wz).  The equilibrium log law (tau_par = u_tau^2, tau_perp = 0)
 * is the training target, the yardstick of the network's error and the fallback.
 *
 * Parameters: 8770 FP32 values, W1[4][64] b1[64] W2[64][64] b2[64]
 * W3[64][64] b3[64] W4[64][2] b4[2] (W[k][n], input k, output n); and 12
 * standardisation constants: feature means and scales (4 + 4), output means
 * and scales (2 + 2).  Features/outputs passed to the calls below are raw
 * (the network standardises them).  X is n x 4, Y is n x 2, row-major.
 *
 * Channels that use a network as their wall model (chan_set_wall_network)
 * are its users: chan_mlp_set_params, chan_mlp_fit_normalization and
 * chan_mlp_train first wait for the work already queued on the users'
 * streams, and the users' next steps wait for the change to finish, so a
 * change never overlaps a step that reads the network.  chan_mlp_destroy
 * refuses (CHAN_ESTATE, nothing released) while the network has users. */
int chan_mlp_create(int device, unsigned long long seed, chan_mlp** out);
int chan_mlp_destroy(chan_mlp* net);
const char* chan_mlp_last_error(const chan_mlp* net);
int chan_mlp_get_params(chan_mlp* net, float* params /* 8770 or NULL */, float* norm /* 12 or NULL */);
int chan_mlp_set_params(chan_mlp* net, const float* params, const float* norm);
/* standardisation constants from data: means and standard deviations */
int chan_mlp_fit_normalization(chan_mlp* net, const float* X, const float* Y, int n);
int chan_mlp_forward(chan_mlp* net, const float* X, int n, float* Y);
/* gradient of L = 1/(2n) sum (y - t)^2 over standardised outputs, no update */
int chan_mlp_gradients(chan_mlp* net, const float* X, const float* Y, int n, float* grad, double* loss);
/* Adam (lr, beta 0.9/0.999, eps 1e-8), shuffled minibatches (batch a
 * multiple of 32); epoch_loss (epochs values or NULL) gets each epoch's
 * mean loss during training */
int chan_mlp_train(chan_mlp* net, const float* X, const float* Y, int n, int epochs, int batch, float lr,
                   unsigned long long seed, double* epoch_loss);

/* Samples of every wall cell of the channel's current state at height h_wm
 * (<= 0: first off-wall cell) with viscosity nu: features X (2*nx*nzg x 4),
 * equilibrium targets Y (2*nx*nzg x 2) and, if tau is not NULL, the
 * equilibrium stress u_tau^2 (2*nx*nzg), in alpha_x order.  The lagged-stress
 * feature uses the context's alpha_x. */
int chan_wall_samples(chan_ctx* ctx, double h_wm, double nu, double kappa, double B, float* X, float* Y,
                      double* tau);
/* the network used by CHAN_WALL_MLP (not owned; same device as ctx; NULL
 * detaches the current one unless CHAN_WALL_MLP is selected; the channel is
 * a user of the network until it is detached or destroyed) and the largest
 * relative deviation
 * |(tau_par, tau_perp) - (u_tau^2, 0)| / u_tau^2 from the equilibrium stress
 * it may have before a cell falls back to the equilibrium value */
int chan_set_wall_network(chan_ctx* ctx, chan_mlp* net, double max_rel_err);
/* evaluate the selected wall-stress source (equilibrium or network) once on
 * the current state into alpha_x (CHAN_WALL_GIVEN: nothing to do) */
int chan_evaluate_wall_stress(chan_ctx* ctx);
/* of the last CHAN_WALL_MLP evaluation: mean relative deviation from the
 * equilibrium stress over the wall cells, fraction of cells that fell back */
int chan_wall_model_stats(chan_ctx* ctx, double out2[2]);

/* ---------------------------------------------------------------------
 * Ensembles: many independent channels (any sizes, precisions and wall
 * models; one device) advanced together.  No reference counterpart (the
 * reference runs one channel per MPI job).  Members of one shape and the
 * same step settings (precision, dynamic Smagorinsky, wall-model kind,
 * mass-flow mode, tile shapes) form a group whose step is one launch per
 * stage for all of its members, each member's arguments in a device table;
 * other members run their own kernels.  One step of every group and lone
 * member is captured into one graph of independent branches;
 * chan_ensemble_run_steps launches that graph once per step (compute_dt +
 * compute_time_step_RK3 of every member).  Results are bit for bit those of
 * the members stepped alone.  Members stay usable on their own: the ensemble
 * orders its steps after the members' pending work and the members' later
 * work after its steps.  Changing a member's settings (wall model, tile
 * shape) or stepping it alone is detected and the graphs are captured again.
 * The ensemble does not own the members; destroy it before them. */
int  chan_ensemble_create(chan_ctx* const* members, int n, chan_ensemble** out);
void chan_ensemble_destroy(chan_ensemble* ens);
const char* chan_ensemble_last_error(const chan_ensemble* ens);
int chan_ensemble_size(const chan_ensemble* ens);
int chan_ensemble_run_steps(chan_ensemble* ens, int nsteps);
/* grouping on (1, the default) or off (0: every member runs its own
 * kernels, one branch per member); the number of groups and lone members
 * the current settings give */
int chan_ensemble_set_grouping(chan_ensemble* ens, int on);
int chan_ensemble_group_count(chan_ensemble* ens);
int chan_ensemble_synchronize(chan_ensemble* ens);

/* Tile shapes and the autotuner.  No reference counterpart (the reference's
 * loops have no tiles).  CHAN_KERNEL_RHS selects the fused RHS + RK kernel,
 * whose shapes differ in (x, y) tile and z-chunk length; CHAN_KERNEL_DSM the
 * dynamic Smagorinsky pass A (x, y tile; its z chunk follows the grid).
 * chan_autotune times every shape of both kernels on this channel (reps
 * launches each: the RHS kernel's substep-2 form writing into the rhs
 * buffers, the DSM pass A writing |S| and its partial sums; the state is not
 * changed), keeps the fastest of each for the step and, when times_ms is not
 * NULL, writes the average ms per launch of the RHS shapes followed by the
 * DSM shapes, at most max_out values.  A channel whose LES model is not the
 * dynamic Smagorinsky (LES_model != 2) has no DSM pass: its DSM shapes are
 * not timed and their times are 0.  DSM shapes change the order of the
 * plane sums (results agree to rounding, not bit for bit). */
enum { CHAN_KERNEL_RHS = 0, CHAN_KERNEL_DSM = 1 };
int chan_tile_shape_count(int kernel);
int chan_set_tile_shape(chan_ctx* ctx, int kernel, int shape);
int chan_get_tile_shape(const chan_ctx* ctx, int kernel);  /* shape index, or < 0 */
int chan_autotune(chan_ctx* ctx, int reps, double* times_ms, int max_out);

#ifdef __cplusplus
}
#endif
#endif
