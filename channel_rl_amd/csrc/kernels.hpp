// kernels.hpp -- host-callable launchers of the channel kernels.
#pragma once
#include "common.hpp"
#include "wallmodel.hpp"

struct chan_mlp;

namespace chan {

// field kinds for the ghost fill
enum FieldKind { KIND_U = 0, KIND_V = 1, KIND_W = 2, KIND_C_DIRICHLET = 3, KIND_C_NEUMANN = 4 };

template <typename T>
struct EqAlphaArgs;

// Work buffers of the dynamic Smagorinsky model (box layout, type T).
template <typename T>
struct DsmBufs {
    T* smag;       // |S| (box layout)
    double* part;  // per-block partial sums: [nyg][2][nblk]
    double* ratio; // <LM>/<MM> per row (index j-1); also LM, MM at +nyg, +2nyg
    int nblk;      // partial-sum blocks per row
    int kchunk;    // z planes per block of the fused pass A
    int shape;     // tile shape of the fused pass A (dsm_shape_count)
    // reciprocal spacings, 1-based: 1/(x(i)-x(i-1)), 1/(xg(i)-xg(i-1)), ...
    const T *rdx, *rdy, *rdz, *rdxg, *rdyg, *rdzg;
    // compute_dt's maxima over the pass's output cells, when dtpart is set:
    // per block of the pass, [3][block] (code/time_integration.f90:113-121),
    // with the centre coordinates they divide (FP64, 1-based)
    double* dtpart;
    const double *xg, *yg, *zg;
    const double* mins3;  // compute_dt's smallest spacings (its viscous limit)
    int nblk_dt;          // blocks of the pass (dtpart's row length)
    // with ntab set (and the wall rows from the wall model, fuse_wm), pass B
    // is left to the consumer of nu_t: the ratio launch writes per row j the
    // coefficients (A, B) at ntab[c*nyg + j-1] with nu_t = max(A v, B) for
    // the value v that smag then holds at a cell, and a small kernel puts the
    // wall model's nu_t into smag's wall rows and completes the periodic
    // images of the others; nu_t itself is not written
    T* ntab;
};

template <typename T>
void launch_bc(T* F, const Dims& d, int kind, hipStream_t s);

// 1-based metric tables of the RHS kernel: reciprocal spacings
// r[i] = 1/(a(i)-a(i-1)) (0 where undefined) and the y interpolation weights
template <typename T>
struct RhsTabs {
    const T *rdx, *rdy, *rdz, *rdxg, *rdyg, *rdzg;
    const T *w0, *w1;
};

// mode 0: rhs only into R*; mode 1..3: RK substep (Wray, code/time_integration.f90:34-86)
// ntab: nut holds the dynamic model's deferred pass-B values (DsmBufs::ntab)
template <typename T>
void launch_rhs_rk(const T* U, const T* V, const T* W, const T* nut, const RhsTabs<T>& m, const Dims& d,
                   T nu, const Scalars* sc, int mode, const T* Uo, const T* Vo, const T* Wo, T* F1u,
                   T* F1v, T* F1w, T* Ou, T* Ov, T* Ow, double c1, double c2, int shape, hipStream_t s,
                   const T* ntab = nullptr);
// number of RHS tile shapes (shape index 0 .. count-1)
int rhs_shape_count();

template <typename T>
void launch_dsm(const T* U, const T* V, const T* W, T* nut, const Metric<T>& m, const Dims& d,
                DsmBufs<T>& b, int dirichlet, int fuse_wm, const double* alpha_x, T nu, const Scalars* sc,
                hipStream_t s, const EqAlphaArgs<T>* eq = nullptr);

// AMD eddy viscosity (amd.hip), dirichlet = Dirichlet_nu_t (0, 1, 2).  With
// fuse_wm, eq set: the wall rows' launch also computes the equilibrium wall
// stress into alpha first; dtpart set (fuse_wm, no wall viscosity limit):
// the two launches also take compute_dt ([3][amd_blocks] partials; md's
// centre coordinates, mins3 the smallest spacings) and set sc->dt
template <typename T>
void launch_amd(const T* U, const T* V, const T* W, T* nut, const Metric<T>& m, const Dims& d, int dirichlet,
                const Scalars* sc, int fuse_wm, const double* alpha, T nu, hipStream_t s,
                const EqAlphaArgs<T>* eq = nullptr, double* dtpart = nullptr, const Metric<double>* md = nullptr,
                const double* mins3 = nullptr);
int amd_blocks(const Dims& d);

// fused pass A of the dynamic Smagorinsky model (dsm_fused.hip): z-chunk
// length for this grid and device, and the partial-sum blocks per row it gives
template <typename T>
int dsm_fused_kchunk(const Dims& d, int device, int shape);
template <typename T>
int dsm_fused_nblk(const Dims& d, int kchunk, int shape);
template <typename T>
int dsm_fused_blocks(const Dims& d, int kchunk, int shape);
int dsm_shape_count();
template <typename T>
void launch_dsm_fused(const T* U, const T* V, const T* W, const Metric<T>& m, const Dims& d, DsmBufs<T>& b,
                      hipStream_t s);

template <typename T>
void launch_wall_model(const T* U, T* nut, const double* alpha_x, const Metric<T>& m, const Dims& d, T nu,
                       const Scalars* sc, hipStream_t s);


template <typename T>
void launch_equilibrium_alpha(const T* U, double* alpha_x, const Dims& d, double nu, const WallRows& ws,
                              double kappa, double B, hipStream_t s);

// ------------------------------------------------------ grouped launches
// One launch per stage for a group of same-shape channels: each run_*
// launcher takes one channel's arguments by value (tab = nullptr, nmem = 1)
// or a device table of nmem members' arguments; the other parameters are
// the group's common shape and launch settings.
struct DsmRatioArgs {
    const double* part;
    double* ratio;
    // with dtpart set, one more block reduces the DSM pass's compute_dt
    // maxima (nblk_dt blocks) and sets sc->dt as k_dt_final does
    const double* dtpart;
    int nblk_dt;
    Scalars* sc;
    const double* mins3;
    // with eq_alpha set, further blocks compute the equilibrium wall stress
    // of k_equilibrium_alpha from eq_U (float when fp32) into eq_alpha
    const void* eq_U;
    int fp32;
    double* eq_alpha;
    double eq_nu, eq_kappa, eq_B, eq_yc;
    int eq_jb;
    double eq_wb, eq_hb, eq_ht;
    // with ntab set, the row blocks also write pass B's coefficients
    // (DsmBufs::ntab; float when fp32), the wall rows passed through
    void* ntab;
};
template <typename T>
struct DsmNutArgs {
    T* nut;
    const T* smag;
    const double* ratio;
    const T* U;
    const double* alpha;
    Metric<T> m;
    T nu;
    int dirichlet;
    const Scalars* sc;
};
template <typename T>
struct EqAlphaArgs {
    const T* U;
    double* alpha;
    double nu;
    WallRows ws;
    double kappa, B, yc;
};
template <typename T>
struct MassArgs {
    T* U;
    double* part;
    Scalars* sc;
    Metric<double> m;
    // mode 1 with p set: U is the velocity before the projection by p, and
    // the functional is taken of the projected values (k_project's arithmetic)
    const double* p;
};
template <typename T>
struct DtArgs {
    const T *U, *V, *W;
    Metric<double> m;
    double* part;
    Scalars* sc;
    const double* mins3;
};
template <typename T>
struct DsmArgs {
    const T *U, *V, *W;
    DsmBufs<T> b;
};
template <typename T>
struct RhsArgs {
    const T *U, *V, *W, *nt;
    const T* ntab;  // nt holds deferred pass-B values with these row coefficients (or null)
    RhsTabs<T> tb;
    T nu;
    const Scalars* sc;
    const T *Uo, *Vo, *Wo;
    T *F1u, *F1v, *F1w, *Ou, *Ov, *Ow;
};
template <typename T>
struct WmArgs {
    const void* net;  // the network's inference image (mlp_device_net)
    const T *U, *W;
    double* alpha;
    const Scalars* sc;
    Metric<double> m;
    double nu, h_wm, kappa, B, yc, max_err;
    double* stats;
    int x_mass_cte;
};

void run_dsm_ratio(const DsmRatioArgs& one, const DsmRatioArgs* tab, int nmem, const Dims& d, int nblk,
                   hipStream_t s);
template <typename T>
void run_dsm_nut(const DsmNutArgs<T>& one, const DsmNutArgs<T>* tab, int nmem, const Dims& d, int fuse_wm,
                 hipStream_t s);
// the rest of a deferred pass B (DsmBufs::ntab) in the |S| box passed as nut
template <typename T>
void run_dsm_edges(const DsmNutArgs<T>& one, const DsmNutArgs<T>* tab, int nmem, const Dims& d, hipStream_t s);
template <typename T>
void run_equilibrium_alpha(const EqAlphaArgs<T>& one, const EqAlphaArgs<T>* tab, int nmem, const Dims& d,
                           hipStream_t s);
// mode 1: the reference's functional (code/mass_flow.f90:113-166); mode 2:
// the bulk mean of code/mass_flow.f90:24-62 (constant bulk flow)
template <typename T>
void run_mass_flow_correct(const MassArgs<T>& one, const MassArgs<T>* tab, int nmem, const Dims& d, int nblk,
                           int mode, int apply, int advance, hipStream_t s);
template <typename T>
void run_dt(const DtArgs<T>& one, const DtArgs<T>* tab, int nmem, const Dims& d, int nblk, hipStream_t s);
template <typename T>
void run_dsm_fused(const DsmArgs<T>& one, const DsmArgs<T>* tab, int nmem, const Dims& d, int kchunk, int shape,
                   hipStream_t s);
template <typename T>
void run_rhs_rk(const RhsArgs<T>& one, const RhsArgs<T>* tab, int nmem, const Dims& d, int mode, double c1,
                double c2, int shape, hipStream_t s);
template <typename T>
void run_wm_mlp(const WmArgs<T>& one, const WmArgs<T>* tab, int nmem, const Dims& d, hipStream_t s);

template <typename T>
void launch_mass_flow_init(const T* U, const Metric<double>& md, const Dims& d, double* part, int nblk,
                           Scalars* sc, hipStream_t s);

template <typename T>
void launch_mass_flow_correct(T* U, const Metric<double>& md, const Dims& d, double* part, int nblk, Scalars* sc,
                              int mode, int apply, hipStream_t s, int advance);
// the functional of mode 1 over the projection of U by p, before it is
// applied (the shift then goes into the projection: launch_project's shift)
template <typename T>
void launch_mass_flow_of_projection(const T* U, const double* p, const Metric<double>& md, const Dims& d,
                                    double* part, int nblk, Scalars* sc, hipStream_t s);

template <typename T>
void launch_dt(const T* U, const T* V, const T* W, const Metric<double>& md, const Dims& d, double* part,
               int nblk, Scalars* sc, const double* mins3, hipStream_t s);

template <typename T>
void launch_divergence_check(const T* U, const T* V, const T* W, const Metric<double>& md, const Dims& d,
                             double* part, int nblk, Scalars* sc, hipStream_t s);

template <typename T>
void launch_plane_stats(const T* U, const T* V, const T* W, const Metric<double>& md, const Dims& d,
                        double* out, hipStream_t s);

template <typename T>
void launch_wall_shear(const T* U, const T* nut, const Metric<double>& md, const Dims& d, double nu, double* part,
                       int nblk, double* out, hipStream_t s);

// the true pressure (pressure.hip)
template <typename T>
void launch_pressure_rhs(const T* V, const T* nut, const T* Ru, const T* Rv, const T* Rw, const Metric<double>& m,
                         const Dims& d, double nu, double* bc1, double* bc2, double* P, hipStream_t s);
void launch_pressure_ghosts(double* P, const double* bc1, const double* bc2, const Metric<double>& m, const Dims& d,
                            hipStream_t s);

// the RL agents' observations and actions (rl.hip)
template <typename T>
void launch_rl_observe(const T* U, const T* nut, const Metric<double>& m, const Dims& d, double nu, int na, int nxa,
                       double dPdx_ref, const double* draws, double* states, double* rewards, hipStream_t s);
void launch_rl_actions(double* alpha, const double* factors, const Dims& d, int na, int nxa, double* acts,
                       hipStream_t s);

template <typename T>
void launch_scale_shift(T* F, const Dims& d, hipStream_t s);

// ---------------------------------------------------------------- Poisson
struct FftPlan {
    int n;          // transform length
    int nr;         // number of radix passes
    int r[12];      // radices, first pass first
    const double2* tw;  // exp(-2 pi i t / n), t = 0..n-1 (device)
};

struct PoissonBufs {
    double2* ph;        // spectral pressure [ny-1][nkx][nzp], kz fastest
    double* p;          // pseudo-pressure (box layout, double)
    const double* kxx;  // nxp
    const double* kzz;  // nzp
    const double* dd;   // Dyy diagonal, rows j=2..nyg-1 at [0..ny-2]
    const double* dl;   // Dyy(j+1,j), j=2..nyg-2
    const double* du;   // Dyy(j,j+1)
    FftPlan fx, fz;
    int bx, bz;         // transforms per block in the x and z passes
    double* tri;        // pivots 1/den_j of the y elimination, [row][mode] (tall systems only)
};

template <typename T>
struct FwdArgs {
    const T *U, *V, *W;
    Metric<double> m;
    PoissonBufs pb;
};
template <typename T>
struct ProjArgs {
    T *U, *V, *W;
    const double* p;
    Metric<double> m;
    // with shift set, U also takes the mass-flow shift sc->last_dPdx
    // (code/time_integration.f90:91) in the same store
    const Scalars* shift;
};
// grouped forms of the projection's stages.  pb0 (forward) and `one`
// (solve, inverse) are any member's buffers in a grouped launch: the group's
// transform plans and batch sizes come from there, the members' pointers
// from tab (tab_pb for the z pass of the forward transform).
template <typename T>
void run_poisson_forward(const FwdArgs<T>& one, const FwdArgs<T>* tab, const PoissonBufs* tab_pb, int nmem,
                         const Dims& d, const PoissonBufs& pb0, hipStream_t s);
void run_poisson_solve(const PoissonBufs& one, const PoissonBufs* tab, int nmem, const Dims& d, hipStream_t s);
void run_poisson_inverse(const PoissonBufs& one, const PoissonBufs* tab, int nmem, const Dims& d, hipStream_t s);
template <typename T>
void run_project(const ProjArgs<T>& one, const ProjArgs<T>* tab, int nmem, const Dims& d, hipStream_t s);

template <typename T>
void launch_poisson_forward(const T* U, const T* V, const T* W, const Metric<double>& md, const Dims& d,
                            PoissonBufs& pb, hipStream_t s);
void launch_poisson_solve(const Dims& d, PoissonBufs& pb, hipStream_t s);
// the elimination's pivots of every mode (once per channel)
// the y solve keeps systems of up to 96 rows in registers with pivots computed
// in the kernel; taller systems need the pivot table PoissonBufs::tri
bool tridiag_needs_table(const Dims& d);
void launch_tridiag_setup(const Dims& d, PoissonBufs& pb, hipStream_t s);
void launch_poisson_inverse(const Dims& d, PoissonBufs& pb, hipStream_t s);
// forward x pass of a given centre field (for the standalone Poisson solve)
void launch_poisson_forward_field(const double* rhs, const Dims& d, PoissonBufs& pb, hipStream_t s);

// the learned wall model (mlp.hip).  mlp_device_net gives a network's
// inference image and device.  A channel that reads the network attaches
// its stream (mlp_attach, mlp_detach on release) and, before each launch
// that reads it, waits on mlp_ready: parameter changes then never overlap
// its steps.
bool mlp_device_net(const ::chan_mlp* net, const void** img, int* device);
hipEvent_t mlp_ready(const ::chan_mlp* net);
hipError_t mlp_attach(::chan_mlp* net, hipStream_t s);
void mlp_detach(::chan_mlp* net, hipStream_t s);
void launch_mlp_infer(const void* img, const float* X, float* Y, int n, hipStream_t s);
template <typename T>
void launch_wm_samples(const T* U, const T* W, const double* alpha, const Scalars* sc, const Metric<double>& m,
                       const Dims& d, double nu, double h_wm, double kappa, double B, int x_mass_cte, float* X, float* Y,
                       double* Tq, hipStream_t s);
long wm_mlp_groups(const Dims& d);
template <typename T>
void launch_wm_mlp(const void* net, const T* U, const T* W, double* alpha, const Scalars* sc,
                   const Metric<double>& m, const Dims& d, double nu, double h_wm, double kappa, double B,
                   int x_mass_cte, double max_err, double* stats, hipStream_t s);

template <typename T>
void launch_project(T* U, T* V, T* W, const double* p, const Metric<double>& md, const Dims& d, hipStream_t s,
                    const Scalars* shift = nullptr);

}  // namespace chan
