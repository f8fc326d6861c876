// chanles.hip -- the C ABI (include/chanles.h): context set-up, host<->device
// copies in the reference's Fortran shapes, and the per-subroutine entry
// points, including the graph-captured RK3 step.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/chanles.h"
#include "kernels.hpp"
#include "prof.hpp"
#include "wallmodel.hpp"

using namespace chan;

namespace chan {
thread_local Prof* g_prof = nullptr;
}

namespace {

// ------------------------------------------------------------ pack/unpack
// Fortran-packed host layout (n1,n2,n3) placed at box offset (o,o,o).
template <typename T>
__global__ void k_unpack(const double* src, T* dst, Dims d, int n1, int n2, int n3, int o)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long tot = (long)n1 * n2 * n3;
    if (q >= tot) return;
    const int i = (int)(q % n1), j = (int)((q / n1) % n2), k = (int)(q / ((long)n1 * n2));
    dst[at(d, i + o, j + o, k + o)] = T(src[q]);
}

template <typename T>
__global__ void k_pack(const T* src, double* dst, Dims d, int n1, int n2, int n3, int o)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long tot = (long)n1 * n2 * n3;
    if (q >= tot) return;
    const int i = (int)(q % n1), j = (int)((q / n1) % n2), k = (int)(q / ((long)n1 * n2));
    dst[q] = (double)src[at(d, i + o, j + o, k + o)];
}

template <typename T>
__global__ void k_fill(T* a, long n, T v)
{
    const long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (q < n) a[q] = v;
}

__global__ void k_advance_time(Scalars* sc) { sc->t = sc->t + sc->dt; }

// plain streaming copy (16-byte vectors), the bandwidth reference
// Each block copies 4 x 256 consecutive 16-byte elements: four coalesced
// loads in flight per lane, then the four stores.
__global__ void __launch_bounds__(256) k_stream_copy(const double2* __restrict__ a, double2* __restrict__ b, long n)
{
    const long base = (long)blockIdx.x * 1024 + threadIdx.x;
    double2 v[4];
#pragma unroll
    for (int r = 0; r < 4; r++)
        if (base + r * 256 < n) v[r] = a[base + r * 256];
#pragma unroll
    for (int r = 0; r < 4; r++)
        if (base + r * 256 < n) b[base + r * 256] = v[r];
}

struct Shape { int n1, n2, n3, o; };

}  // namespace

// Device memory of chan_compute_pressure_noalloc: the wall data of the true
// pressure's Neumann rows (two nxg x nzg planes) for callers that cannot
// allocate.  One copy per GPU, held until the process ends; its address on a
// device comes from hipGetSymbolAddress with that device current (the name
// itself is a host-side stand-in).
static constexpr long NOALLOC_SCRATCH = 1L << 21;  // doubles, 16 MiB
static __device__ __align__(256) double scratch[NOALLOC_SCRATCH];

struct chan_ctx {
    std::string err;
    int prec = CHAN_FP64;
    int device = 0;
    Dims d{};
    // host copies of the set-up constants
    std::vector<double> x, y, z, xg, yg, zg, w0, w1, kxx, kzz, dd, dl, du;
    double dxmin = 0, dymin = 0, dzmin = 0;
    double rk_t[3]{}, rk_coef[3][3]{};
    double nu = 0, CFL = 0, dPdx0 = 0;
    int x_mass_cte = 0, LES_model = 0, Dirichlet_nu_t = 1;
    int wall_mode = CHAN_WALL_GIVEN;
    double h_wm = 0.0, kappa = 0.41, Bc = 5.2;
    hipStream_t stream = nullptr;
    // device
    Metric<double> md{};
    Metric<float> mf{};
    std::vector<void*> allocs;
    void* state[3][3] = {};  // [buffer][U,V,W]
    int cur = 0;
    void* F1[3] = {};
    void* rhs[3] = {};
    void* nut = nullptr;
    void* dsm[13] = {};
    const void* rcp[2][6] = {};  // reciprocal spacings [double, float][dx, dy, dz, dxg, dyg, dzg]
    double* part = nullptr;
    int nblk_dsm = 0, nblk_red = 0, dsm_kchunk = 16, dsm_shape = 0;
    double* ratio = nullptr;
    double* alpha = nullptr;
    double* stage = nullptr;
    Scalars* sc = nullptr;
    double* mins3 = nullptr;
    PoissonBufs pb{};
    hipGraphExec_t graph[3] = {nullptr, nullptr, nullptr};
    // the same steps with compute_dt folded into substep 1 (dt_fused)
    hipGraphExec_t graph_dt[3] = {nullptr, nullptr, nullptr};
    double* dtpart = nullptr;  // the DSM pass's compute_dt maxima, [3][blocks]
    int nblk_dt = 0;
    // substeps 1 and 2 of a step leave the dynamic model's pass B to the RHS
    // kernel (DsmBufs::ntab): its row coefficients, [2][nyg] of the bulk type
    void* ntab = nullptr;
    bool defer_nut = true;
    double wm_limit = 0.0;  // the wall viscosity limit (a copy of sc->wm_nu_limit)
    bool use_graph = true;
    int rhs_shape = 0;  // index into the RHS kernel's tile shapes
    // learned wall model
    chan_mlp* net = nullptr;
    double wm_max_err = 0.05;
    double* wm_stats = nullptr;   // per 16-cell group: error sums, fallbacks
    float* wm_samples = nullptr;  // 2*nx*nzg x (4 + 2) floats, on demand
    long generation = 0;          // bumped whenever captured steps go stale
    double* pbc = nullptr;        // wall data of the true pressure, 2 planes nxg x nzg
    double* scratch = nullptr;    // this device's copy of `scratch` (no-alloc entry point)
    Prof prof;
    std::vector<std::pair<std::string, std::pair<double, long>>> prof_out;
};

namespace {

int fail(chan_ctx* c, int code, const std::string& msg)
{
    if (c) c->err = msg;
    return code;
}

#define HIPCHK(ctx, call)                                                                          \
    do {                                                                                           \
        hipError_t e_ = (call);                                                                    \
        if (e_ != hipSuccess)                                                                      \
            return fail(ctx, CHAN_EHIP, std::string(#call) + ": " + hipGetErrorString(e_));       \
    } while (0)

int dalloc(chan_ctx* c, void** p, size_t bytes)
{
    hipError_t e = hipMalloc(p, bytes ? bytes : 8);
    if (e != hipSuccess) return fail(c, CHAN_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    c->allocs.push_back(*p);
    e = hipMemset(*p, 0, bytes ? bytes : 8);
    if (e != hipSuccess) return fail(c, CHAN_EHIP, std::string("hipMemset: ") + hipGetErrorString(e));
    return CHAN_OK;
}

template <typename V>
int upload(chan_ctx* c, const V& v, const double** out)
{
    void* p;
    int rc = dalloc(c, &p, v.size() * sizeof(double));
    if (rc) return rc;
    HIPCHK(c, hipMemcpy(p, v.data(), v.size() * sizeof(double), hipMemcpyHostToDevice));
    *out = (const double*)p;
    return CHAN_OK;
}

int upload_f(chan_ctx* c, const std::vector<double>& v, const float** out)
{
    std::vector<float> f(v.begin(), v.end());
    void* p;
    int rc = dalloc(c, &p, f.size() * sizeof(float));
    if (rc) return rc;
    HIPCHK(c, hipMemcpy(p, f.data(), f.size() * sizeof(float), hipMemcpyHostToDevice));
    *out = (const float*)p;
    return CHAN_OK;
}

bool factor_plan(int n, FftPlan& pl)
{
    pl.n = n;
    pl.nr = 0;
    int m = n;
    while (m > 1) {
        int r = (m % 4 == 0) ? 4 : (m % 2 == 0) ? 2 : (m % 3 == 0) ? 3 : (m % 5 == 0) ? 5 : 0;
        if (r == 0 || pl.nr >= 12) return false;
        pl.r[pl.nr++] = r;
        m /= r;
    }
    return true;
}

// the equilibrium wall model's sampling rows: the first centre row jb >= 2
// whose successor lies at or above y(1) + h_wm (at most nyg-2), and the
// linear weight of row jb; h_wm = 0 samples row 2 at its own height
WallRows wall_rows(const chan_ctx* c, double h_wm)
{
    const std::vector<double>&y = c->y, &yg = c->yg;
    const int ny = c->d.ny, nyg = c->d.nyg;
    WallRows ws{2, 1.0, yg[2] - y[1], y[ny] - yg[nyg - 1]};
    if (h_wm > 0.0) {
        const double yt = y[1] + h_wm;
        while (ws.jb < nyg - 2 && yg[ws.jb + 1] < yt) ws.jb++;
        ws.wb = (yg[ws.jb + 1] - yt) / (yg[ws.jb + 1] - yg[ws.jb]);
        ws.hb = h_wm;
        ws.ht = h_wm;
    }
    return ws;
}

int make_twiddles(chan_ctx* c, FftPlan& pl)
{
    std::vector<double2> tw(pl.n);
    const long double pi = 3.141592653589793238462643383279502884L;
    for (int t = 0; t < pl.n; t++) {
        long double a = -2.0L * pi * (long double)t / (long double)pl.n;
        tw[t] = make_double2((double)cosl(a), (double)sinl(a));
    }
    void* p;
    int rc = dalloc(c, &p, tw.size() * sizeof(double2));
    if (rc) return rc;
    HIPCHK(c, hipMemcpy(p, tw.data(), tw.size() * sizeof(double2), hipMemcpyHostToDevice));
    pl.tw = (const double2*)p;
    return CHAN_OK;
}

size_t esize(const chan_ctx* c) { return c->prec == CHAN_FP64 ? sizeof(double) : sizeof(float); }

Shape field_shape(const chan_ctx* c, int which)
{
    const Dims& d = c->d;
    switch (which) {
    case CHAN_U: return {d.nx, d.nyg, d.nzg, 1};
    case CHAN_V: return {d.nxg, d.ny, d.nzg, 1};
    case CHAN_W: return {d.nxg, d.nyg, d.nz, 1};
    case CHAN_NU_T: case CHAN_P: return {d.nxg, d.nyg, d.nzg, 1};
    case CHAN_ALPHA_X: return {d.nx, 2, d.nzg, 1};
    case CHAN_RHS_U: return {d.nx - 2, d.nyg - 2, d.nzg - 2, 2};
    case CHAN_RHS_V: return {d.nxg - 2, d.ny - 2, d.nzg - 2, 2};
    case CHAN_RHS_W: return {d.nxg - 2, d.nyg - 2, d.nz - 2, 2};
    }
    return {0, 0, 0, 0};
}

void* field_ptr(chan_ctx* c, int which)
{
    switch (which) {
    case CHAN_U: case CHAN_V: case CHAN_W: return c->state[c->cur][which];
    case CHAN_NU_T: return c->nut;
    case CHAN_P: return c->pb.p;
    case CHAN_RHS_U: case CHAN_RHS_V: case CHAN_RHS_W: return c->rhs[which - CHAN_RHS_U];
    }
    return nullptr;
}

template <typename T>
DsmBufs<T> dsm_bufs(chan_ctx* c)
{
    DsmBufs<T> b;
    b.smag = (T*)c->dsm[9];
    b.part = c->part;
    b.ratio = c->ratio;
    b.nblk = c->nblk_dsm;
    b.kchunk = c->dsm_kchunk;
    b.shape = c->dsm_shape;
    const int pi = sizeof(T) == sizeof(double) ? 0 : 1;
    b.rdx = (const T*)c->rcp[pi][0]; b.rdy = (const T*)c->rcp[pi][1]; b.rdz = (const T*)c->rcp[pi][2];
    b.rdxg = (const T*)c->rcp[pi][3]; b.rdyg = (const T*)c->rcp[pi][4]; b.rdzg = (const T*)c->rcp[pi][5];
    b.dtpart = nullptr;  // set by the step that folds compute_dt into the pass
    b.ntab = nullptr;    // set by the substeps that defer pass B
    b.xg = c->md.xg; b.yg = c->md.yg; b.zg = c->md.zg;
    b.mins3 = c->mins3;
    b.nblk_dt = c->nblk_dt;
    return b;
}

template <typename T>
const Metric<T>& metric(const chan_ctx* c);
template <>
const Metric<double>& metric<double>(const chan_ctx* c) { return c->md; }
template <>
const Metric<float>& metric<float>(const chan_ctx* c) { return c->mf; }

template <typename T>
RhsTabs<T> rhs_tabs(const chan_ctx* c)
{
    const int pi = sizeof(T) == sizeof(double) ? 0 : 1;
    const Metric<T>& m = metric<T>(c);
    RhsTabs<T> t;
    t.rdx = (const T*)c->rcp[pi][0]; t.rdy = (const T*)c->rcp[pi][1]; t.rdz = (const T*)c->rcp[pi][2];
    t.rdxg = (const T*)c->rcp[pi][3]; t.rdyg = (const T*)c->rcp[pi][4]; t.rdzg = (const T*)c->rcp[pi][5];
    t.w0 = m.w0; t.w1 = m.w1;
    return t;
}

// ---------------------------------------------------------- stage helpers

template <typename T>
void enqueue_bc(chan_ctx* c, int buf)
{
    T** s = (T**)c->state[buf];
    launch_bc<T>(s[0], c->d, KIND_U, c->stream);
    launch_bc<T>(s[1], c->d, KIND_V, c->stream);
    launch_bc<T>(s[2], c->d, KIND_W, c->stream);
}

// with_eq: the DSM path also computes the equilibrium wall stress of this
// state into c->alpha (in its ratio launch, before the pass that reads it).
// defer: the DSM path leaves its pass B to the RHS kernel, c->nut is not
// written (the RHS then reads the |S| box with the coefficients c->ntab)
template <typename T>
void enqueue_eddy_viscosity(chan_ctx* c, int buf, int fuse_wm, bool with_dt = false, bool with_eq = false,
                            bool defer = false)
{
    T** s = (T**)c->state[buf];
    const T nu = T(c->nu);
    if (c->LES_model == 2) {
        DsmBufs<T> b = dsm_bufs<T>(c);
        if (with_dt) b.dtpart = c->dtpart;
        if (defer) b.ntab = (T*)c->ntab;
        const EqAlphaArgs<T> eq{s[0], c->alpha, c->nu, wall_rows(c, c->h_wm), c->kappa, c->Bc,
                                loglaw_crossover(c->kappa, c->Bc)};
        launch_dsm<T>(s[0], s[1], s[2], (T*)c->nut, metric<T>(c), c->d, b, c->Dirichlet_nu_t == 1, fuse_wm,
                      c->alpha, nu, c->sc, c->stream, with_eq ? &eq : nullptr);
    } else if (c->LES_model == 3) {
        const EqAlphaArgs<T> eq{s[0], c->alpha, c->nu, wall_rows(c, c->h_wm), c->kappa, c->Bc,
                                loglaw_crossover(c->kappa, c->Bc)};
        launch_amd<T>(s[0], s[1], s[2], (T*)c->nut, metric<T>(c), c->d, c->Dirichlet_nu_t, c->sc, fuse_wm, c->alpha,
                      nu, c->stream, with_eq ? &eq : nullptr, with_dt ? c->dtpart : nullptr, &c->md, c->mins3);
    } else {
        k_fill<T><<<cdiv(c->d.nbox, 256), 256, 0, c->stream>>>((T*)c->nut, c->d.nbox, T(0));
        if (fuse_wm) launch_wall_model<T>(s[0], (T*)c->nut, c->alpha, metric<T>(c), c->d, nu, c->sc, c->stream);
    }
}

// with_mass_flow: the step's last projection, also applying the correction
// of mass-flow mode 1 (code/time_integration.f90:88-92): its functional is
// taken of the projected values before they are stored, and the shift goes
// into the projection's own store (the same values as projecting, then
// shifting: one pass over U fewer)
template <typename T>
void enqueue_projection(chan_ctx* c, int buf, bool with_mass_flow = false)
{
    T** s = (T**)c->state[buf];
    launch_poisson_forward<T>(s[0], s[1], s[2], c->md, c->d, c->pb, c->stream);
    launch_poisson_solve(c->d, c->pb, c->stream);
    launch_poisson_inverse(c->d, c->pb, c->stream);
    if (with_mass_flow)
        launch_mass_flow_of_projection<T>(s[0], c->pb.p, c->md, c->d, c->part, c->nblk_red, c->sc, c->stream);
    launch_project<T>(s[0], s[1], s[2], c->pb.p, c->md, c->d, c->stream, with_mass_flow ? c->sc : nullptr);
}

// code/time_integration.f90:24-96 with three rotating state buffers:
// X (start of step, also Uo), Y, Z; substeps X->Y, Y->Z, Z->Y.
// dt_fused: compute_dt of the step's start state is taken by the DSM pass
// of substep 1 and its ratio kernel (dt_fusable), which need it first
template <typename T>
void enqueue_rk3(chan_ctx* c, int X, bool dt_fused = false)
{
    const int Y = (X + 1) % 3, Z = (X + 2) % 3;
    const int in[3] = {X, Y, Z}, out[3] = {Y, Z, Y};
    T** Xs = (T**)c->state[X];
    for (int s = 0; s < 3; s++) {
        T** I = (T**)c->state[in[s]];
        T** O = (T**)c->state[out[s]];
        // the equilibrium wall stress rides in a launch of the eddy viscosity
        // (the DSM ratio launch, or the AMD wall rows' launch)
        const bool eq_in_dsm = c->wall_mode == CHAN_WALL_EQUILIBRIUM && (c->LES_model == 2 || c->LES_model == 3);
        if (c->wall_mode == CHAN_WALL_EQUILIBRIUM && !eq_in_dsm) {
            launch_equilibrium_alpha<T>(I[0], c->alpha, c->d, c->nu, wall_rows(c, c->h_wm), c->kappa, c->Bc, c->stream);
        } else if (c->wall_mode == CHAN_WALL_MLP) {
            const void* net = nullptr;
            int ndev = 0;
            if (mlp_device_net(c->net, &net, &ndev))
                launch_wm_mlp<T>(net, I[0], I[2], c->alpha, c->sc, c->md, c->d, c->nu, c->h_wm, c->kappa, c->Bc,
                                 c->x_mass_cte, c->wm_max_err, c->wm_stats, c->stream);
        }
        // nu_t is read by the RHS only, except after the last substep
        const bool defer = c->defer_nut && c->LES_model == 2 && s < 2;
        enqueue_eddy_viscosity<T>(c, in[s], 1, dt_fused && s == 0, eq_in_dsm, defer);
        double c1, c2 = 0.0;
        if (s == 0) c1 = c->rk_coef[0][0];
        else if (s == 1) { c1 = c->rk_coef[1][0]; c2 = c->rk_coef[1][1]; }
        else { c1 = c->rk_coef[2][0]; c2 = c->rk_coef[2][2]; }
        launch_rhs_rk<T>(I[0], I[1], I[2], defer ? (const T*)c->dsm[9] : (const T*)c->nut, rhs_tabs<T>(c), c->d,
                         T(c->nu), c->sc, s + 1, Xs[0], Xs[1], Xs[2], (T*)c->F1[0], (T*)c->F1[1], (T*)c->F1[2], O[0],
                         O[1], O[2], c1, c2, c->rhs_shape, c->stream, defer ? (const T*)c->ntab : nullptr);
        enqueue_projection<T>(c, out[s], s == 2 && c->x_mass_cte == 1);
    }
    // the mass-flow correction's final reduction also advances the time (mode
    // 1: in the last projection)
    if (c->x_mass_cte == 2) {
        T** O = (T**)c->state[Y];
        launch_mass_flow_correct<T>(O[0], c->md, c->d, c->part, c->nblk_red, c->sc, c->x_mass_cte, 1, c->stream, 1);
    } else if (c->x_mass_cte != 1) {
        k_advance_time<<<1, 1, 0, c->stream>>>(c->sc);
    }
}

// the captured steps bake in launch shapes and options: drop them on change
void drop_graphs(chan_ctx* c)
{
    for (int q = 0; q < 3; q++) {
        if (c->graph[q]) { (void)hipGraphExecDestroy(c->graph[q]); c->graph[q] = nullptr; }
        if (c->graph_dt[q]) { (void)hipGraphExecDestroy(c->graph_dt[q]); c->graph_dt[q] = nullptr; }
    }
    c->generation++;
}

int check_launch(chan_ctx* c)
{
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail(c, CHAN_EHIP, std::string("kernel launch: ") + hipGetErrorString(e));
    return CHAN_OK;
}

template <typename T>
int rk3_step(chan_ctx* c, bool dt_fused)
{
    const int X = c->cur;
    if (!c->use_graph) {
        enqueue_rk3<T>(c, X, dt_fused);
        int rc = check_launch(c);
        if (rc) return rc;
    } else {
        hipGraphExec_t& gx = dt_fused ? c->graph_dt[X] : c->graph[X];
        if (!gx) {
            hipGraph_t g;
            HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            enqueue_rk3<T>(c, X, dt_fused);
            HIPCHK(c, hipStreamEndCapture(c->stream, &g));
            hipError_t e = hipGraphInstantiate(&gx, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) return fail(c, CHAN_EHIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        }
        HIPCHK(c, hipGraphLaunch(gx, c->stream));
    }
    c->cur = (X + 1) % 3;
    return CHAN_OK;
}

template <typename T>
void enqueue_dt(chan_ctx* c)
{
    T** s = (T**)c->state[c->cur];
    launch_dt<T>(s[0], s[1], s[2], c->md, c->d, c->part, c->nblk_red, c->sc, c->mins3, c->stream);
}

int set_device(chan_ctx* c)
{
    HIPCHK(c, hipSetDevice(c->device));
    return CHAN_OK;
}

// the channel's next launches after the last change of its wall network
int net_order(chan_ctx* c)
{
    if (c->wall_mode == CHAN_WALL_MLP && c->net) HIPCHK(c, hipStreamWaitEvent(c->stream, mlp_ready(c->net), 0));
    return CHAN_OK;
}

// the DSM pass A tile shape q with its z chunk and partial-sum block count
void set_dsm_shape(chan_ctx* c, int q)
{
    c->dsm_shape = q;
    if (c->prec == CHAN_FP64) {
        c->dsm_kchunk = dsm_fused_kchunk<double>(c->d, c->device, q);
        c->nblk_dsm = dsm_fused_nblk<double>(c->d, c->dsm_kchunk, q);
        c->nblk_dt = dsm_fused_blocks<double>(c->d, c->dsm_kchunk, q);
    } else {
        c->dsm_kchunk = dsm_fused_kchunk<float>(c->d, c->device, q);
        c->nblk_dsm = dsm_fused_nblk<float>(c->d, c->dsm_kchunk, q);
        c->nblk_dt = dsm_fused_blocks<float>(c->d, c->dsm_kchunk, q);
    }
}

// compute_dt folded into the step's first DSM pass: the dynamic Smagorinsky
// model (its pass reads every cell compute_dt reads) and a wall stress that
// does not read dt before that pass (the learned model's pressure-gradient
// feature does)
// compute_dt folded into the step's first eddy-viscosity launches: the DSM
// pass A with its ratio launch, or the AMD kernel with its wall rows' launch
// (whose wall viscosity limit, if any, reads dt in that launch); not with the
// learned wall model, whose kernel runs before them and reads dt
bool dt_fusable(const chan_ctx* c)
{
    if (c->wall_mode == CHAN_WALL_MLP) return false;
    return c->LES_model == 2 || (c->LES_model == 3 && c->wm_limit == 0.0);
}

// the address of this device's copy of `scratch`, looked up once per device
// with that device current (the caller has made it current)
constexpr int SCRATCH_DEVICES = 64;
std::mutex g_scratch_mu;
double* g_scratch[SCRATCH_DEVICES] = {};

int scratch_address(chan_ctx* c)
{
    if (c->device < 0 || c->device >= SCRATCH_DEVICES)
        return fail(c, CHAN_EINVAL, "device ordinal beyond the no-alloc scratch table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch[c->device]) {
        void* p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(c, CHAN_EHIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch[c->device] = static_cast<double*>(p);
    }
    c->scratch = g_scratch[c->device];
    return CHAN_OK;
}

}  // namespace

// ================================================================ C ABI

extern "C" {

const char* chan_last_error(const chan_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int chan_create(const chan_config* cfg, chan_ctx** out)
{
    if (!cfg || !out) return CHAN_EINVAL;
    *out = nullptr;
    chan_ctx* c = new chan_ctx();
    *out = c;
    const int nx = cfg->nx, ny = cfg->ny, nz = cfg->nz;
    // code/initialization.f90:66-74 restrictions, one domain (nprocs = 1)
    if (nx % 2 != 0) return fail(c, CHAN_EINVAL, "Error: nx must be even");
    if (nz % 2 != 0) return fail(c, CHAN_EINVAL, "Error: nz must be even");
    if (nx < 6 || nz < 6 || ny < 4) return fail(c, CHAN_EINVAL, "Error: grid too small (nx,nz >= 6, ny >= 4)");
    if (nz - 2 < 2) return fail(c, CHAN_EINVAL, "Error: nslices_z must be at least 2");
    if (!cfg->x || !cfg->y || !cfg->z) return fail(c, CHAN_EINVAL, "Error: missing grid coordinates");
    if (cfg->precision != CHAN_FP64 && cfg->precision != CHAN_FP32_BULK)
        return fail(c, CHAN_EINVAL, "Error: invalid precision");
    if (cfg->x_mass_cte < 0 || cfg->x_mass_cte > 2)
        return fail(c, CHAN_EINVAL, "Error: x_mass_cte must be 0, 1 (the reference's functional) or 2 (bulk flow)");
    if (cfg->LES_model != 0 && cfg->LES_model != 2 && cfg->LES_model != 3)
        return fail(c, CHAN_EINVAL, "Error: LES_model must be 0 (none), 2 (dynamic Smagorinsky) or 3 (AMD)");
    if (cfg->Dirichlet_nu_t < 0 || cfg->Dirichlet_nu_t > 2 || (cfg->Dirichlet_nu_t == 2 && cfg->LES_model != 3))
        return fail(c, CHAN_EINVAL, "Error: Dirichlet_nu_t must be 1 or 0 (2: AMD only)");
    c->prec = cfg->precision;
    c->device = cfg->device;
    c->nu = cfg->nu; c->CFL = cfg->CFL; c->dPdx0 = cfg->dPdx;
    c->x_mass_cte = cfg->x_mass_cte; c->LES_model = cfg->LES_model; c->Dirichlet_nu_t = cfg->Dirichlet_nu_t;
    if (const char* e = getenv("CHANLES_NO_GRAPH")) c->use_graph = (e[0] == '0');
    if (const char* e = getenv("CHANLES_NO_DEFERRED_NUT")) c->defer_nut = (e[0] == '0');

    Dims& d = c->d;
    d.nx = nx; d.ny = ny; d.nz = nz;
    d.nxg = nx + 1; d.nyg = ny + 1; d.nzg = nz + 1;
    d.nxp = nx - 2; d.nzp = nz - 2; d.nkx = d.nxp / 2 + 1;
    d.sx = (d.nxg + 15) / 16 * 16;
    d.sxy = (long)d.sx * d.nyg;
    d.nbox = d.sxy * d.nzg;

    // grids, code/initialization.f90:187-240 (1-based vectors, element 0 unused)
    int nxg = d.nxg, nyg = d.nyg, nzg = d.nzg;
    c->x.assign(nx + 2, 0.0); c->y.assign(ny + 2, 0.0); c->z.assign(nz + 2, 0.0);
    for (int i = 1; i <= nx; i++) c->x[i] = cfg->x[i - 1];
    for (int j = 1; j <= ny; j++) c->y[j] = cfg->y[j - 1];
    for (int k = 1; k <= nz; k++) c->z[k] = cfg->z[k - 1];
    auto centres = [](const std::vector<double>& f, int n, std::vector<double>& g) {
        const int nm = n - 1;
        std::vector<double> m(nm + 1);
        for (int i = 1; i <= nm; i++) m[i] = 0.5 * (f[i] + f[i + 1]);
        g.assign(nm + 3, 0.0);
        for (int i = 1; i <= nm; i++) g[i + 1] = m[i];
        g[1] = m[1] - 2.0 * (m[1] - f[1]);
        g[nm + 2] = m[nm] + 2.0 * (f[n] - m[nm]);
    };
    centres(c->x, nx, c->xg); centres(c->y, ny, c->yg); centres(c->z, nz, c->zg);
    auto minspacing = [](const std::vector<double>& g, int ng) {
        double m = g[2] - g[1];
        for (int i = 2; i < ng; i++) m = std::fmin(m, g[i + 1] - g[i]);
        return m;
    };
    c->dxmin = minspacing(c->xg, nxg); c->dymin = minspacing(c->yg, nyg); c->dzmin = minspacing(c->zg, nzg);
    for (int i = 1; i <= nx - 1; i++)
        if (!(c->x[i + 1] > c->x[i])) return fail(c, CHAN_EINVAL, "Error: x faces must increase");
    for (int j = 1; j <= ny - 1; j++)
        if (!(c->y[j + 1] > c->y[j])) return fail(c, CHAN_EINVAL, "Error: y faces must increase");
    for (int k = 1; k <= nz - 1; k++)
        if (!(c->z[k + 1] > c->z[k])) return fail(c, CHAN_EINVAL, "Error: z faces must increase");

    // modified wavenumbers, dx = dxmin (code/initialization.f90:250-308)
    const double pi = 4.0 * std::atan(1.0);
    auto modified = [&](int np, double dx, std::vector<double>& kk) {
        kk.assign(np, 0.0);
        int ch = (int)std::ceil((double)np / 2.0);
        for (int i = 0; i <= ch && i < np; i++) kk[i] = 2.0 * (std::cos(2.0 * pi * (double)i / (double)np) - 1.0) / (dx * dx);
        for (int i = ch + 1; i <= np - 1; i++)
            kk[i] = 2.0 * (std::cos(2.0 * pi * (double)(-np + i) / (double)np) - 1.0) / (dx * dx);
    };
    modified(d.nxp, c->dxmin, c->kxx);
    modified(d.nzp, c->dzmin, c->kzz);

    // Dyy rows j=2..nyg-1 packed at 0..ny-2 (code/initialization.f90:355-385)
    const std::vector<double>&Y = c->y, &YG = c->yg;
    auto abc = [&](int j, double& a, double& b, double& cc) {
        a = 1.0 / (Y[j] - Y[j - 1]) / (YG[j + 1] - YG[j]);
        b = 1.0 / (Y[j] - Y[j - 1]) * (-1.0 / (YG[j + 1] - YG[j]) - 1.0 / (YG[j] - YG[j - 1]));
        cc = 1.0 / (Y[j] - Y[j - 1]) / (YG[j] - YG[j - 1]);
    };
    const int nr = ny - 1;
    c->dd.assign(nr, 0.0); c->dl.assign(nr, 0.0); c->du.assign(nr, 0.0);
    for (int j = 3; j <= nyg - 2; j++) {
        double a, b, cc;
        abc(j, a, b, cc);
        c->dd[j - 2] = b;
        c->du[j - 2] = a;       // Dyy(j,j+1)
        c->dl[j - 3] = cc;      // Dyy(j,j-1) = DL(j-1)
    }
    {
        double a, b, cc;
        abc(2, a, b, cc);
        c->dd[0] = b + cc;
        c->du[0] = a;
        abc(nyg - 1, a, b, cc);
        c->dd[nr - 1] = a + b;
        c->dl[nr - 2] = cc;
    }
    c->w0.assign(ny + 2, 0.0); c->w1.assign(ny + 2, 0.0);
    for (int j = 1; j <= ny; j++) {
        c->w0[j] = (YG[j + 1] - Y[j]) / (YG[j + 1] - YG[j]);
        c->w1[j] = 1.0 - c->w0[j];
    }
    // Wray RK3 (code/initialization.f90:415-427)
    c->rk_t[0] = 8.0 / 15.0; c->rk_t[1] = 2.0 / 3.0; c->rk_t[2] = 1.0;
    std::memset(c->rk_coef, 0, sizeof(c->rk_coef));
    c->rk_coef[0][0] = 8.0 / 15.0;
    c->rk_coef[1][0] = 1.0 / 4.0; c->rk_coef[1][1] = 5.0 / 12.0;
    c->rk_coef[2][0] = 1.0 / 4.0; c->rk_coef[2][1] = 0.0; c->rk_coef[2][2] = 3.0 / 4.0;

    FftPlan fx{}, fz{};
    if (!factor_plan(d.nxp, fx) || !factor_plan(d.nzp, fz))
        return fail(c, CHAN_EINVAL, "Error: nx-2 and nz-2 must factor into 2, 3 and 5");
    // one block holds a whole line (and its batch) in LDS
    if (d.nxp > 2048 || d.nzp > 2048) return fail(c, CHAN_EINVAL, "Error: nx-2 and nz-2 must be at most 2048");
    // the reductions flatten a box's cells into 32-bit indices
    if (d.nbox > 0x7fffffffL) return fail(c, CHAN_EINVAL, "Error: the grid has more than 2^31 box cells");

    int rc = set_device(c);
    if (rc) return rc;
    HIPCHK(c, hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    if ((rc = scratch_address(c))) return rc;

    // metrics
    if ((rc = upload(c, c->x, &c->md.x)) || (rc = upload(c, c->y, &c->md.y)) || (rc = upload(c, c->z, &c->md.z)) ||
        (rc = upload(c, c->xg, &c->md.xg)) || (rc = upload(c, c->yg, &c->md.yg)) ||
        (rc = upload(c, c->zg, &c->md.zg)) || (rc = upload(c, c->w0, &c->md.w0)) ||
        (rc = upload(c, c->w1, &c->md.w1)))
        return rc;
    if ((rc = upload_f(c, c->x, &c->mf.x)) || (rc = upload_f(c, c->y, &c->mf.y)) ||
        (rc = upload_f(c, c->z, &c->mf.z)) || (rc = upload_f(c, c->xg, &c->mf.xg)) ||
        (rc = upload_f(c, c->yg, &c->mf.yg)) || (rc = upload_f(c, c->zg, &c->mf.zg)) ||
        (rc = upload_f(c, c->w0, &c->mf.w0)) || (rc = upload_f(c, c->w1, &c->mf.w1)))
        return rc;

    const size_t fb = (size_t)d.nbox * esize(c);
    for (int b = 0; b < 3; b++)
        for (int q = 0; q < 3; q++)
            if ((rc = dalloc(c, &c->state[b][q], fb))) return rc;
    for (int q = 0; q < 3; q++) {
        if ((rc = dalloc(c, &c->F1[q], fb))) return rc;
        if ((rc = dalloc(c, &c->rhs[q], fb))) return rc;
    }
    if ((rc = dalloc(c, &c->nut, fb))) return rc;
    if (c->LES_model == 2 && (rc = dalloc(c, &c->dsm[9], fb))) return rc;  // |S|
    if (c->LES_model == 2 && (rc = dalloc(c, &c->ntab, 2 * (size_t)d.nyg * sizeof(double)))) return rc;
    {
        // 1-based reciprocal spacings for the RHS and the dynamic model
        const std::vector<double>* f[6] = {&c->x, &c->y, &c->z, &c->xg, &c->yg, &c->zg};
        const int nf[6] = {nx, ny, nz, nxg, nyg, nzg};
        for (int q = 0; q < 6; q++) {
            std::vector<double> r(nf[q] + 2, 0.0);
            for (int i = 2; i <= nf[q]; i++) r[i] = 1.0 / ((*f[q])[i] - (*f[q])[i - 1]);
            const double* pd;
            const float* pf;
            if ((rc = upload(c, r, &pd)) || (rc = upload_f(c, r, &pf))) return rc;
            c->rcp[0][q] = pd;
            c->rcp[1][q] = pf;
        }
        c->md.rdx = (const double*)c->rcp[0][0]; c->md.rdy = (const double*)c->rcp[0][1];
        c->md.rdz = (const double*)c->rcp[0][2];
        c->md.rdxg = (const double*)c->rcp[0][3]; c->md.rdyg = (const double*)c->rcp[0][4];
        c->md.rdzg = (const double*)c->rcp[0][5];
        c->mf.rdx = (const float*)c->rcp[1][0]; c->mf.rdy = (const float*)c->rcp[1][1];
        c->mf.rdz = (const float*)c->rcp[1][2];
        c->mf.rdxg = (const float*)c->rcp[1][3]; c->mf.rdyg = (const float*)c->rcp[1][4];
        c->mf.rdzg = (const float*)c->rcp[1][5];
    }
    // partial sums: DSM rows use one slot per (x,z) block of the fused
    // kernel; room for the largest count any tile shape gives
    size_t nblk_max = 0, nblk_dt_max = 0;
    for (int q = dsm_shape_count() - 1; q >= 0; q--) {
        set_dsm_shape(c, q);
        nblk_max = std::max(nblk_max, (size_t)c->nblk_dsm);
        nblk_dt_max = std::max(nblk_dt_max, (size_t)c->nblk_dt);
    }
    if (c->LES_model == 3) nblk_dt_max = amd_blocks(d);
    if ((c->LES_model == 2 || c->LES_model == 3) &&
        (rc = dalloc(c, (void**)&c->dtpart, 3 * nblk_dt_max * sizeof(double))))
        return rc;
    // reduction blocks: about eight (y, z) rows each, at most 1024
    c->nblk_red = (int)std::max(64L, std::min(1024L, (long)(d.nyg - 2) * (d.nzg - 2) / 8));
    size_t npart = std::max((size_t)(d.nyg + 1) * 2 * nblk_max, (size_t)3 * c->nblk_red);
    if ((rc = dalloc(c, (void**)&c->part, npart * sizeof(double)))) return rc;
    if ((rc = dalloc(c, (void**)&c->ratio, (size_t)3 * d.nyg * sizeof(double)))) return rc;
    if ((rc = dalloc(c, (void**)&c->alpha, (size_t)d.nx * 2 * d.nzg * sizeof(double)))) return rc;
    if ((rc = dalloc(c, (void**)&c->wm_stats, (size_t)2 * wm_mlp_groups(d) * sizeof(double)))) return rc;
    if ((rc = dalloc(c, (void**)&c->stage, (size_t)d.nbox * sizeof(double)))) return rc;
    if ((rc = dalloc(c, (void**)&c->sc, sizeof(Scalars)))) return rc;
    if ((rc = dalloc(c, (void**)&c->pb.p, (size_t)d.nbox * sizeof(double)))) return rc;
    if ((rc = dalloc(c, (void**)&c->pb.ph, (size_t)(d.ny - 1) * d.nzp * d.nkx * sizeof(double2)))) return rc;
    if ((rc = upload(c, c->kxx, &c->pb.kxx)) || (rc = upload(c, c->kzz, &c->pb.kzz)) ||
        (rc = upload(c, c->dd, &c->pb.dd)) || (rc = upload(c, c->dl, &c->pb.dl)) || (rc = upload(c, c->du, &c->pb.du)))
        return rc;
    std::vector<double> mins = {c->dxmin, c->dymin, c->dzmin};
    const double* mp;
    if ((rc = upload(c, mins, &mp))) return rc;
    c->mins3 = (double*)mp;
    c->pb.fx = fx;
    c->pb.fz = fz;
    if ((rc = make_twiddles(c, c->pb.fx)) || (rc = make_twiddles(c, c->pb.fz))) return rc;
    c->pb.bx = std::max(1, std::min(16, 2048 / d.nxp));
    c->pb.bz = std::max(1, std::min(16, 2048 / d.nzp));
    if (tridiag_needs_table(d)) {
        if ((rc = dalloc(c, (void**)&c->pb.tri, (size_t)(d.ny - 1) * d.nzp * d.nkx * sizeof(double)))) return rc;
        launch_tridiag_setup(d, c->pb, c->stream);
        if ((rc = check_launch(c))) return rc;
    }
    Scalars s0{};
    s0.dt = 0.0; s0.t = 0.0; s0.dPdx = cfg->dPdx; s0.Qflow_x_0 = 0.0;
    s0.cfl = cfg->CFL; s0.nu = cfg->nu; s0.dt_max = 1.0e-3;
    HIPCHK(c, hipMemcpy(c->sc, &s0, sizeof(s0), hipMemcpyHostToDevice));
    HIPCHK(c, hipDeviceSynchronize());
    return CHAN_OK;
}

void chan_destroy(chan_ctx* c)
{
    if (!c) return;
    // a context rejected by chan_create's argument checks holds no device
    // resources: release it without touching the runtime
    if (!c->stream && c->allocs.empty()) {
        delete c;
        return;
    }
    (void)hipSetDevice(c->device);
    for (int q = 0; q < 3; q++) {
        if (c->graph[q]) (void)hipGraphExecDestroy(c->graph[q]);
        if (c->graph_dt[q]) (void)hipGraphExecDestroy(c->graph_dt[q]);
    }
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    if (c->net) mlp_detach(c->net, c->stream);
    for (void* p : c->allocs) (void)hipFree(p);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

int chan_set_field(chan_ctx* c, int which, const double* src)
{
    if (!c || !src) return CHAN_EINVAL;
    int rc = set_device(c);
    if (rc) return rc;
    Shape s = field_shape(c, which);
    if (s.n1 <= 0) return fail(c, CHAN_EINVAL, "chan_set_field: unknown field");
    const size_t n = (size_t)s.n1 * s.n2 * s.n3;
    if (which == CHAN_ALPHA_X) {
        HIPCHK(c, hipMemcpyAsync(c->alpha, src, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return CHAN_OK;
    }
    HIPCHK(c, hipMemcpyAsync(c->stage, src, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
    void* dst = field_ptr(c, which);
    if (which == CHAN_P || c->prec == CHAN_FP64)
        k_unpack<double><<<cdiv(n, 256), 256, 0, c->stream>>>(c->stage, (double*)dst, c->d, s.n1, s.n2, s.n3, s.o);
    else
        k_unpack<float><<<cdiv(n, 256), 256, 0, c->stream>>>(c->stage, (float*)dst, c->d, s.n1, s.n2, s.n3, s.o);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CHAN_OK;
}

int chan_get_field(chan_ctx* c, int which, double* dst)
{
    if (!c || !dst) return CHAN_EINVAL;
    int rc = set_device(c);
    if (rc) return rc;
    Shape s = field_shape(c, which);
    if (s.n1 <= 0) return fail(c, CHAN_EINVAL, "chan_get_field: unknown field");
    const size_t n = (size_t)s.n1 * s.n2 * s.n3;
    if (which == CHAN_ALPHA_X) {
        HIPCHK(c, hipMemcpyAsync(dst, c->alpha, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return CHAN_OK;
    }
    void* src = field_ptr(c, which);
    if (which == CHAN_P || c->prec == CHAN_FP64)
        k_pack<double><<<cdiv(n, 256), 256, 0, c->stream>>>((const double*)src, c->stage, c->d, s.n1, s.n2, s.n3, s.o);
    else
        k_pack<float><<<cdiv(n, 256), 256, 0, c->stream>>>((const float*)src, c->stage, c->d, s.n1, s.n2, s.n3, s.o);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(dst, c->stage, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CHAN_OK;
}

int chan_set_scalars(chan_ctx* c, double dt, double t, double dPdx, double Q)
{
    if (!c) return CHAN_EINVAL;
    Scalars s0{};
    int rc = set_device(c);
    if (rc) return rc;
    HIPCHK(c, hipMemcpy(&s0, c->sc, sizeof(s0), hipMemcpyDeviceToHost));
    s0.dt = dt; s0.t = t; s0.dPdx = dPdx; s0.Qflow_x_0 = Q;
    HIPCHK(c, hipMemcpy(c->sc, &s0, sizeof(s0), hipMemcpyHostToDevice));
    return CHAN_OK;
}

int chan_get_scalars(chan_ctx* c, double out[4])
{
    if (!c || !out) return CHAN_EINVAL;
    Scalars s0{};
    int rc = set_device(c);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(&s0, c->sc, sizeof(s0), hipMemcpyDeviceToHost));
    out[0] = s0.dt; out[1] = s0.t; out[2] = s0.dPdx; out[3] = s0.Qflow_x_0;
    return CHAN_OK;
}

#define DISPATCH(c, fn, ...) ((c)->prec == CHAN_FP64 ? fn<double>(__VA_ARGS__) : fn<float>(__VA_ARGS__))

int chan_apply_boundary_conditions(chan_ctx* c)
{
    if (!c) return CHAN_EINVAL;
    int rc = set_device(c);
    if (rc) return rc;
    DISPATCH(c, enqueue_bc, c, c->cur);
    return check_launch(c);
}

int chan_compute_eddy_viscosity(chan_ctx* c)
{
    if (!c) return CHAN_EINVAL;
    int rc = set_device(c);
    if (rc) return rc;
    DISPATCH(c, enqueue_eddy_viscosity, c, c->cur, 0);
    return check_launch(c);
}

int chan_compute_wall_model(chan_ctx* c)
{
    if (!c) return CHAN_EINVAL;
    int rc = set_device(c);
    if (rc) return rc;
    if (c->prec == CHAN_FP64)
        launch_wall_model<double>((double*)c->state[c->cur][0], (double*)c->nut, c->alpha, c->md, c->d, c->nu,
                                  c->sc, c->stream);
    else
        launch_wall_model<float>((float*)c->state[c->cur][0], (float*)c->nut, c->alpha, c->mf, c->d, (float)c->nu,
                                 c->sc, c->stream);
    return check_launch(c);
}

int chan_compute_rhs(chan_ctx* c)
{
    if (!c) return CHAN_EINVAL;
    int rc = set_device(c);
    if (rc) return rc;
    void** s = c->state[c->cur];
    if (c->prec == CHAN_FP64) {
        double** S = (double**)s;
        launch_rhs_rk<double>(S[0], S[1], S[2], (double*)c->nut, rhs_tabs<double>(c), c->d, c->nu, c->sc, 0, nullptr, nullptr,
                              nullptr, nullptr, nullptr, nullptr, (double*)c->rhs[0], (double*)c->rhs[1],
                              (double*)c->rhs[2], 0.0, 0.0, c->rhs_shape, c->stream);
    } else {
        float** S = (float**)s;
        launch_rhs_rk<float>(S[0], S[1], S[2], (float*)c->nut, rhs_tabs<float>(c), c->d, (float)c->nu, c->sc, 0, nullptr, nullptr,
                             nullptr, nullptr, nullptr, nullptr, (float*)c->rhs[0], (float*)c->rhs[1],
                             (float*)c->rhs[2], 0.0, 0.0, c->rhs_shape, c->stream);
    }
    return check_launch(c);
}

int chan_compute_projection_step(chan_ctx* c)
{
    if (!c) return CHAN_EINVAL;
    int rc = set_device(c);
    if (rc) return rc;
    DISPATCH(c, enqueue_projection, c, c->cur);
    return check_launch(c);
}

int chan_solve_poisson_equation(chan_ctx* c)
{
    if (!c) return CHAN_EINVAL;
    int rc = set_device(c);
    if (rc) return rc;
    launch_poisson_forward_field(c->pb.p, c->d, c->pb, c->stream);
    launch_poisson_solve(c->d, c->pb, c->stream);
    launch_poisson_inverse(c->d, c->pb, c->stream);
    return check_launch(c);
}

extern "C++" template <typename T>
void enqueue_pressure(chan_ctx* c, double* bc1)
{
    T** S = (T**)c->state[c->cur];
    T** R = (T**)c->rhs;
    launch_rhs_rk<T>(S[0], S[1], S[2], (const T*)c->nut, rhs_tabs<T>(c), c->d, T(c->nu), c->sc, 0, nullptr, nullptr,
                     nullptr, nullptr, nullptr, nullptr, R[0], R[1], R[2], 0.0, 0.0, c->rhs_shape, c->stream);
    double* bc2 = bc1 + (long)c->d.nxg * c->d.nzg;
    launch_pressure_rhs<T>(S[1], (const T*)c->nut, R[0], R[1], R[2], c->md, c->d, c->nu, bc1, bc2, c->pb.p,
                           c->stream);
    launch_poisson_forward_field(c->pb.p, c->d, c->pb, c->stream);
    launch_poisson_solve(c->d, c->pb, c->stream);
    launch_poisson_inverse(c->d, c->pb, c->stream);
    launch_pressure_ghosts(c->pb.p, bc1, bc2, c->md, c->d, c->stream);
}

int chan_compute_pressure(chan_ctx* c)
{
    if (!c) return CHAN_EINVAL;
    int rc = set_device(c);
    if (rc) return rc;
    if (!c->pbc && (rc = dalloc(c, (void**)&c->pbc, (size_t)2 * c->d.nxg * c->d.nzg * sizeof(double)))) return rc;
    DISPATCH(c, enqueue_pressure, c, c->pbc);
    return check_launch(c);
}

int chan_compute_pressure_noalloc(chan_ctx* c)
{
    if (!c) return CHAN_EINVAL;
    int rc = set_device(c);
    if (rc) return rc;
    if (!c->scratch) return fail(c, CHAN_ESTATE, "chan_compute_pressure_noalloc: no scratch address on this device");
    const long need = 2L * c->d.nxg * c->d.nzg;
    if (need > NOALLOC_SCRATCH)
        return fail(c, CHAN_EINVAL, "chan_compute_pressure_noalloc: the wall data need 2*nxg*nzg = " +
                                        std::to_string(need) + " doubles, more than the " +
                                        std::to_string(NOALLOC_SCRATCH) + " of the library's scratch");
    // every wall-data entry the kernels read is written earlier in the same
    // call: nothing needs clearing
    DISPATCH(c, enqueue_pressure, c, c->scratch);
    return check_launch(c);
}

int chan_initialize_mass_flow(chan_ctx* c)
{
    if (!c) return CHAN_EINVAL;
    int rc = set_device(c);
    if (rc) return rc;
    if (c->prec == CHAN_FP64)
        launch_mass_flow_init<double>((double*)c->state[c->cur][0], c->md, c->d, c->part, c->nblk_red, c->sc,
                                      c->stream);
    else
        launch_mass_flow_init<float>((float*)c->state[c->cur][0], c->md, c->d, c->part, c->nblk_red, c->sc,
                                     c->stream);
    return check_launch(c);
}

int chan_compute_mean_mass_flow_U(chan_ctx* c, double* Q)
{
    if (!c || !Q) return CHAN_EINVAL;
    double keep[4];
    int rc = chan_get_scalars(c, keep);
    if (rc) return rc;
    if ((rc = chan_initialize_mass_flow(c))) return rc;
    double now[4];
    if ((rc = chan_get_scalars(c, now))) return rc;
    *Q = now[3];
    return chan_set_scalars(c, keep[0], keep[1], keep[2], keep[3]);
}

int chan_compute_dPx_for_constant_mass_flow(chan_ctx* c, double* dPdx)
{
    if (!c) return CHAN_EINVAL;
    int rc = set_device(c);
    if (rc) return rc;
    double keep[4];
    if ((rc = chan_get_scalars(c, keep))) return rc;
    const int mode = c->x_mass_cte == 2 ? 2 : 1;
    if (c->prec == CHAN_FP64)
        launch_mass_flow_correct<double>((double*)c->state[c->cur][0], c->md, c->d, c->part, c->nblk_red, c->sc, mode,
                                         0, c->stream, 0);
    else
        launch_mass_flow_correct<float>((float*)c->state[c->cur][0], c->md, c->d, c->part, c->nblk_red, c->sc, mode, 0,
                                        c->stream, 0);
    if ((rc = check_launch(c))) return rc;
    Scalars s0{};
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(&s0, c->sc, sizeof(s0), hipMemcpyDeviceToHost));
    if (dPdx) *dPdx = s0.last_dPdx;
    return chan_set_scalars(c, keep[0], keep[1], keep[2], keep[3]);
}

int chan_compute_dt(chan_ctx* c, double* dt)
{
    if (!c) return CHAN_EINVAL;
    int rc = set_device(c);
    if (rc) return rc;
    DISPATCH(c, enqueue_dt, c);
    if ((rc = check_launch(c))) return rc;
    if (dt) {
        double s[4];
        if ((rc = chan_get_scalars(c, s))) return rc;
        *dt = s[0];
    }
    return CHAN_OK;
}

int chan_compute_time_step_RK3(chan_ctx* c)
{
    if (!c) return CHAN_EINVAL;
    int rc = set_device(c);
    if (rc) return rc;
    if ((rc = net_order(c))) return rc;
    return DISPATCH(c, rk3_step, c, false);
}

int chan_check_divergence(chan_ctx* c, double* max_div, double* max_abs)
{
    if (!c) return CHAN_EINVAL;
    int rc = set_device(c);
    if (rc) return rc;
    void** s = c->state[c->cur];
    if (c->prec == CHAN_FP64)
        launch_divergence_check<double>((double*)s[0], (double*)s[1], (double*)s[2], c->md, c->d, c->part,
                                        c->nblk_red, c->sc, c->stream);
    else
        launch_divergence_check<float>((float*)s[0], (float*)s[1], (float*)s[2], c->md, c->d, c->part, c->nblk_red,
                                       c->sc, c->stream);
    if ((rc = check_launch(c))) return rc;
    Scalars s0{};
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(&s0, c->sc, sizeof(s0), hipMemcpyDeviceToHost));
    if (max_div) *max_div = s0.maxdiv;
    if (max_abs) *max_abs = s0.maxabsdiv;
    return CHAN_OK;
}

int chan_set_wall_model(chan_ctx* c, int mode, double h_wm, double kappa, double B)
{
    if (!c) return CHAN_EINVAL;
    if (mode != CHAN_WALL_GIVEN && mode != CHAN_WALL_EQUILIBRIUM && mode != CHAN_WALL_MLP)
        return fail(c, CHAN_EINVAL, "invalid wall mode");
    if (mode == CHAN_WALL_MLP && !c->net) return fail(c, CHAN_ESTATE, "no wall network set (chan_set_wall_network)");
    if (!(kappa > 0)) return fail(c, CHAN_EINVAL, "kappa must be positive");
    c->wall_mode = mode; c->h_wm = h_wm; c->kappa = kappa; c->Bc = B;
    drop_graphs(c);
    return CHAN_OK;
}

int chan_set_wall_viscosity_limit(chan_ctx* c, double limit)
{
    if (!c) return CHAN_EINVAL;
    if (!(limit >= 0.0)) return fail(c, CHAN_EINVAL, "the wall viscosity limit must be >= 0");
    int rc = set_device(c);
    if (rc) return rc;
    Scalars s0{};
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(&s0, c->sc, sizeof(s0), hipMemcpyDeviceToHost));
    s0.wm_nu_limit = limit;
    HIPCHK(c, hipMemcpy(c->sc, &s0, sizeof(s0), hipMemcpyHostToDevice));
    c->wm_limit = limit;
    return CHAN_OK;
}

int chan_compute_equilibrium_wall_stress(chan_ctx* c, double h_wm, double kappa, double B)
{
    if (!c) return CHAN_EINVAL;
    int rc = set_device(c);
    if (rc) return rc;
    if (c->prec == CHAN_FP64)
        launch_equilibrium_alpha<double>((double*)c->state[c->cur][0], c->alpha, c->d, c->nu, wall_rows(c, h_wm), kappa,
                                         B, c->stream);
    else
        launch_equilibrium_alpha<float>((float*)c->state[c->cur][0], c->alpha, c->d, c->nu, wall_rows(c, h_wm), kappa,
                                         B, c->stream);
    return check_launch(c);
}

int chan_plane_statistics(chan_ctx* c, double* out)
{
    if (!c || !out) return CHAN_EINVAL;
    int rc = set_device(c);
    if (rc) return rc;
    void** s = c->state[c->cur];
    if (c->prec == CHAN_FP64)
        launch_plane_stats<double>((double*)s[0], (double*)s[1], (double*)s[2], c->md, c->d, c->stage, c->stream);
    else
        launch_plane_stats<float>((float*)s[0], (float*)s[1], (float*)s[2], c->md, c->d, c->stage, c->stream);
    if ((rc = check_launch(c))) return rc;
    HIPCHK(c, hipMemcpyAsync(out, c->stage, (size_t)7 * c->d.ny * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CHAN_OK;
}

int chan_wall_shear(chan_ctx* c, double out4[4])
{
    if (!c || !out4) return CHAN_EINVAL;
    int rc = set_device(c);
    if (rc) return rc;
    void** s = c->state[c->cur];
    double* dout = c->stage;  // two doubles of the staging area
    if (c->prec == CHAN_FP64)
        launch_wall_shear<double>((double*)s[0], (double*)c->nut, c->md, c->d, c->nu, c->part, c->nblk_red, dout,
                                  c->stream);
    else
        launch_wall_shear<float>((float*)s[0], (float*)c->nut, c->md, c->d, c->nu, c->part, c->nblk_red, dout,
                                 c->stream);
    if ((rc = check_launch(c))) return rc;
    double t[2];
    HIPCHK(c, hipMemcpyAsync(t, dout, sizeof(t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    // code/statistics.f90:126-130 with Ly the channel height
    const double Ly = c->y[c->d.ny] - c->y[1];
    const double ut = std::sqrt((t[0] + t[1]) / Ly);
    out4[0] = t[0];
    out4[1] = t[1];
    out4[2] = ut;
    out4[3] = ut * Ly / 2.0 / c->nu;
    return CHAN_OK;
}

static int rl_check(chan_ctx* c, int na)
{
    if (na < 1 || (c->d.nx - 2) % na != 0) return fail(c, CHAN_EINVAL, "nagents must divide nx-2");
    if (c->d.nyg < 6) return fail(c, CHAN_EINVAL, "the agents sample rows up to 5: ny >= 5 needed");
    if ((long)22 * na > c->d.nbox) return fail(c, CHAN_EINVAL, "too many agents for the staging area");
    return CHAN_OK;
}

int chan_rl_observe(chan_ctx* c, int nagents, double dPdx_ref, const double* draws, double* states, double* rewards)
{
    if (!c || !draws || !states || !rewards) return CHAN_EINVAL;
    int rc = set_device(c);
    if (rc || (rc = rl_check(c, nagents))) return rc;
    const int na = nagents, nxa = (c->d.nx - 2) / na;
    double* dd = c->stage;
    double* ds = dd + 4L * na;
    double* dr = ds + 12L * na;
    HIPCHK(c, hipMemcpyAsync(dd, draws, 4L * na * sizeof(double), hipMemcpyHostToDevice, c->stream));
    void** s = c->state[c->cur];
    if (c->prec == CHAN_FP64)
        launch_rl_observe<double>((double*)s[0], (double*)c->nut, c->md, c->d, c->nu, na, nxa, dPdx_ref, dd, ds, dr,
                                  c->stream);
    else
        launch_rl_observe<float>((float*)s[0], (float*)c->nut, c->md, c->d, c->nu, na, nxa, dPdx_ref, dd, ds, dr,
                                 c->stream);
    if ((rc = check_launch(c))) return rc;
    HIPCHK(c, hipMemcpyAsync(states, ds, 12L * na * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(rewards, dr, 2L * na * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CHAN_OK;
}

int chan_rl_act(chan_ctx* c, int nagents, const double* factors)
{
    if (!c || !factors) return CHAN_EINVAL;
    int rc = set_device(c);
    if (rc || (rc = rl_check(c, nagents))) return rc;
    const int na = nagents, nxa = (c->d.nx - 2) / na;
    double* df = c->stage + 18L * na;
    double* da = df + 2L * na;
    HIPCHK(c, hipMemcpyAsync(df, factors, 2L * na * sizeof(double), hipMemcpyHostToDevice, c->stream));
    launch_rl_actions(c->alpha, df, c->d, na, nxa, da, c->stream);
    if ((rc = check_launch(c))) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CHAN_OK;
}

int chan_run_steps(chan_ctx* c, int nsteps)
{
    if (!c || nsteps < 0) return CHAN_EINVAL;
    int rc = set_device(c);
    if (rc) return rc;
    const bool fuse = dt_fusable(c);
    for (int n = 0; n < nsteps; n++) {
        if (fuse) {
            // compute_dt + compute_time_step_RK3 with the former inside the
            // step (the same dt: its maxima are exact in any order)
            if ((rc = net_order(c))) return rc;
            if ((rc = DISPATCH(c, rk3_step, c, true))) return rc;
            continue;
        }
        DISPATCH(c, enqueue_dt, c);
        if ((rc = check_launch(c))) return rc;
        if ((rc = chan_compute_time_step_RK3(c))) return rc;
    }
    return CHAN_OK;
}

int chan_profile_steps(chan_ctx* c, int nsteps)
{
    if (!c || nsteps < 1) return CHAN_EINVAL;
    int rc = set_device(c);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->prof.acc.clear();
    const bool g = c->use_graph;
    c->use_graph = false;
    g_prof = &c->prof;
    const bool fuse = dt_fusable(c);
    for (int n = 0; n < nsteps && rc == CHAN_OK; n++) {
        if (fuse) {
            rc = net_order(c);
            if (!rc) rc = DISPATCH(c, rk3_step, c, true);
        } else {
            DISPATCH(c, enqueue_dt, c);
            rc = check_launch(c);
            if (!rc) rc = chan_compute_time_step_RK3(c);
        }
        hipError_t e = hipStreamSynchronize(c->stream);
        if (!rc && e != hipSuccess) rc = fail(c, CHAN_EHIP, std::string("profile sync: ") + hipGetErrorString(e));
        c->prof.collect();
    }
    g_prof = nullptr;
    c->use_graph = g;
    c->prof_out.assign(c->prof.acc.begin(), c->prof.acc.end());
    return rc;
}

int chan_profile_count(const chan_ctx* c) { return c ? (int)c->prof_out.size() : 0; }

int chan_wall_samples(chan_ctx* c, double h_wm, double nu, double kappa, double B, float* X, float* Y, double* tau)
{
    if (!c || !X || !Y) return CHAN_EINVAL;
    if (!(nu > 0.0) || !(kappa > 0.0)) return fail(c, CHAN_EINVAL, "wall samples: nu and kappa must be positive");
    int rc = set_device(c);
    if (rc) return rc;
    const long ncell = 2L * c->d.nx * c->d.nzg;
    // features (4), targets (2) as floats, then the stress as doubles
    if (!c->wm_samples && (rc = dalloc(c, (void**)&c->wm_samples, (size_t)ncell * 8 * sizeof(float)))) return rc;
    float* dX = c->wm_samples;
    float* dY = c->wm_samples + ncell * 4;
    double* dT = (double*)(c->wm_samples + ncell * 6);
    void** s = c->state[c->cur];
    if (c->prec == CHAN_FP64)
        launch_wm_samples<double>((double*)s[0], (double*)s[2], c->alpha, c->sc, c->md, c->d, nu, h_wm, kappa, B,
                                  c->x_mass_cte, dX, dY, dT, c->stream);
    else
        launch_wm_samples<float>((float*)s[0], (float*)s[2], c->alpha, c->sc, c->md, c->d, nu, h_wm, kappa, B,
                                 c->x_mass_cte, dX, dY, dT, c->stream);
    if ((rc = check_launch(c))) return rc;
    HIPCHK(c, hipMemcpyAsync(X, dX, (size_t)ncell * 4 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(Y, dY, (size_t)ncell * 2 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    if (tau) HIPCHK(c, hipMemcpyAsync(tau, dT, (size_t)ncell * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CHAN_OK;
}

int chan_evaluate_wall_stress(chan_ctx* c)
{
    if (!c) return CHAN_EINVAL;
    int rc = set_device(c);
    if (rc) return rc;
    void** s = c->state[c->cur];
    if (c->wall_mode == CHAN_WALL_EQUILIBRIUM) {
        if (c->prec == CHAN_FP64)
            launch_equilibrium_alpha<double>((double*)s[0], c->alpha, c->d, c->nu, wall_rows(c, c->h_wm), c->kappa,
                                         c->Bc, c->stream);
        else
            launch_equilibrium_alpha<float>((float*)s[0], c->alpha, c->d, c->nu, wall_rows(c, c->h_wm), c->kappa,
                                         c->Bc, c->stream);
    } else if (c->wall_mode == CHAN_WALL_MLP) {
        const void* net = nullptr;
        int ndev = 0;
        if (!mlp_device_net(c->net, &net, &ndev)) return fail(c, CHAN_ESTATE, "no wall network set");
        if ((rc = net_order(c))) return rc;
        if (c->prec == CHAN_FP64)
            launch_wm_mlp<double>(net, (double*)s[0], (double*)s[2], c->alpha, c->sc, c->md, c->d, c->nu,
                                  c->h_wm, c->kappa, c->Bc, c->x_mass_cte, c->wm_max_err, c->wm_stats, c->stream);
        else
            launch_wm_mlp<float>(net, (float*)s[0], (float*)s[2], c->alpha, c->sc, c->md, c->d, c->nu, c->h_wm,
                                 c->kappa, c->Bc, c->x_mass_cte, c->wm_max_err, c->wm_stats, c->stream);
    }
    return check_launch(c);
}

int chan_set_wall_network(chan_ctx* c, chan_mlp* net, double max_rel_err)
{
    if (!c) return CHAN_EINVAL;
    if (!(max_rel_err >= 0.0)) return fail(c, CHAN_EINVAL, "max_rel_err must be >= 0");
    if (net) {
        const void* img = nullptr;
        int dev = -1;
        if (!mlp_device_net(net, &img, &dev)) return fail(c, CHAN_EINVAL, "invalid network");
        if (dev != c->device) return fail(c, CHAN_EINVAL, "the network lives on another device");
    }
    if (!net && c->wall_mode == CHAN_WALL_MLP) return fail(c, CHAN_ESTATE, "the MLP wall model is in use");
    if (net != c->net) {
        int rc = set_device(c);
        if (rc) return rc;
        if (net) HIPCHK(c, mlp_attach(net, c->stream));
        if (c->net) mlp_detach(c->net, c->stream);
    }
    c->net = net;
    c->wm_max_err = max_rel_err;
    drop_graphs(c);
    return CHAN_OK;
}

int chan_wall_model_stats(chan_ctx* c, double out2[2])
{
    if (!c || !out2) return CHAN_EINVAL;
    int rc = set_device(c);
    if (rc) return rc;
    const long g = wm_mlp_groups(c->d);
    std::vector<double> h((size_t)2 * g);
    HIPCHK(c, hipMemcpyAsync(h.data(), c->wm_stats, h.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double e = 0.0, f = 0.0;
    for (long q = 0; q < g; q++) { e += h[q]; f += h[g + q]; }
    const double ncell = 2.0 * c->d.nx * c->d.nzg;
    out2[0] = e / ncell;
    out2[1] = f / ncell;
    return CHAN_OK;
}

int chan_tile_shape_count(int kernel)
{
    return kernel == CHAN_KERNEL_RHS ? rhs_shape_count() : (kernel == CHAN_KERNEL_DSM ? dsm_shape_count() : 0);
}

int chan_set_tile_shape(chan_ctx* c, int kernel, int shape)
{
    if (!c) return CHAN_EINVAL;
    if (kernel != CHAN_KERNEL_RHS && kernel != CHAN_KERNEL_DSM) return fail(c, CHAN_EINVAL, "unknown kernel for tile shapes");
    if (shape < 0 || shape >= chan_tile_shape_count(kernel)) return fail(c, CHAN_EINVAL, "tile shape out of range");
    if (kernel == CHAN_KERNEL_RHS) c->rhs_shape = shape;
    else set_dsm_shape(c, shape);
    drop_graphs(c);
    return CHAN_OK;
}

int chan_get_tile_shape(const chan_ctx* c, int kernel)
{
    if (!c) return CHAN_EINVAL;
    if (kernel == CHAN_KERNEL_RHS) return c->rhs_shape;
    if (kernel == CHAN_KERNEL_DSM) return c->dsm_shape;
    return CHAN_EINVAL;
}

extern "C++" {
namespace {
template <typename T>
void tune_rhs(chan_ctx* c)
{
    T** I = (T**)c->state[c->cur];
    T** R = (T**)c->rhs;
    launch_rhs_rk<T>(I[0], I[1], I[2], (const T*)c->nut, rhs_tabs<T>(c), c->d, T(c->nu), c->sc, 2, I[0], I[1], I[2],
                     (T*)c->F1[0], (T*)c->F1[1], (T*)c->F1[2], R[0], R[1], R[2], c->rk_coef[1][0], c->rk_coef[1][1],
                     c->rhs_shape, c->stream);
}
template <typename T>
void tune_dsm(chan_ctx* c)
{
    T** I = (T**)c->state[c->cur];
    DsmBufs<T> b = dsm_bufs<T>(c);
    launch_dsm_fused<T>(I[0], I[1], I[2], metric<T>(c), c->d, b, c->stream);
}
int tune_launch(chan_ctx* c, int kernel)
{
    if (kernel == CHAN_KERNEL_DSM && c->LES_model != 2)
        return fail(c, CHAN_ESTATE, "autotune: the DSM kernel runs only with LES_model 2");
    if (kernel == CHAN_KERNEL_RHS) {
        if (c->prec == CHAN_FP64) tune_rhs<double>(c);
        else tune_rhs<float>(c);
    } else {
        if (c->prec == CHAN_FP64) tune_dsm<double>(c);
        else tune_dsm<float>(c);
    }
    return check_launch(c);
}
}  // namespace
}

// every shape of each tuned kernel: one untimed launch, then reps timed ones
// (the RHS kernel's substep-2 form writes into the rhs buffers and the DSM
// pass A only |S| and its partial sums, so the state is unchanged); the
// fastest shape of each kernel is kept
int chan_autotune(chan_ctx* c, int reps, double* times_ms, int max_out)
{
    if (!c || reps < 1) return CHAN_EINVAL;
    int rc = set_device(c);
    if (rc) return rc;
    hipEvent_t e0, e1;
    HIPCHK(c, hipEventCreate(&e0));
    HIPCHK(c, hipEventCreate(&e1));
    int out = 0;
    for (int kernel : {CHAN_KERNEL_RHS, CHAN_KERNEL_DSM}) {
        const int n = chan_tile_shape_count(kernel);
        if (kernel == CHAN_KERNEL_DSM && c->LES_model != 2) {
            // no dynamic Smagorinsky, no |S| buffer: nothing to time
            for (int q = 0; q < n; q++, out++)
                if (times_ms && out < max_out) times_ms[out] = 0.0;
            continue;
        }
        const int keep = chan_get_tile_shape(c, kernel);
        int best = keep;
        double best_ms = 1e300;
        for (int q = 0; q < n && rc == CHAN_OK; q++, out++) {
            if (kernel == CHAN_KERNEL_RHS) c->rhs_shape = q;
            else set_dsm_shape(c, q);
            float ms = 0.f;
            for (int r = 0; r <= reps && rc == CHAN_OK; r++) {
                if (r == 1) (void)hipEventRecord(e0, c->stream);
                rc = tune_launch(c, kernel);
            }
            (void)hipEventRecord(e1, c->stream);
            hipError_t e = hipEventSynchronize(e1);
            if (!rc && e != hipSuccess) rc = fail(c, CHAN_EHIP, std::string("autotune: ") + hipGetErrorString(e));
            (void)hipEventElapsedTime(&ms, e0, e1);
            const double avg = (double)ms / reps;
            if (times_ms && out < max_out) times_ms[out] = avg;
            if (avg < best_ms) { best_ms = avg; best = q; }
        }
        const int pick = rc == CHAN_OK ? best : keep;
        if (kernel == CHAN_KERNEL_RHS) c->rhs_shape = pick;
        else set_dsm_shape(c, pick);
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    drop_graphs(c);
    return rc;
}

int chan_stream_copy(chan_ctx* c, long nbytes, int reps, double* ms_per_copy)
{
    if (!c || nbytes < 16 || reps < 1) return CHAN_EINVAL;
    int rc = set_device(c);
    if (rc) return rc;
    const long n = nbytes / 16;
    const long grid = (n + 1023) / 1024;
    if (grid > 0x7fffffffL) return fail(c, CHAN_EINVAL, "stream copy: size too large");
    void *a = nullptr, *b = nullptr;
    HIPCHK(c, hipMalloc(&a, n * 16));
    hipError_t e = hipMalloc(&b, n * 16);
    if (e != hipSuccess) {
        (void)hipFree(a);
        return fail(c, CHAN_ENOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    hipEvent_t e0, e1;
    (void)hipEventCreate(&e0);
    (void)hipEventCreate(&e1);
    (void)hipMemsetAsync(a, 0, n * 16, c->stream);
    k_stream_copy<<<(unsigned)grid, 256, 0, c->stream>>>((const double2*)a, (double2*)b, n);
    (void)hipEventRecord(e0, c->stream);
    for (int r = 0; r < reps; r++) k_stream_copy<<<(unsigned)grid, 256, 0, c->stream>>>((const double2*)a, (double2*)b, n);
    (void)hipEventRecord(e1, c->stream);
    e = hipEventSynchronize(e1);
    float ms = 0.f;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    if (e != hipSuccess) return fail(c, CHAN_EHIP, std::string("stream copy: ") + hipGetErrorString(e));
    if (ms_per_copy) *ms_per_copy = ms / reps;
    return CHAN_OK;
}

int chan_profile_entry(const chan_ctx* c, int idx, char* name, int len, double* total_ms, long* launches)
{
    if (!c || idx < 0 || idx >= (int)c->prof_out.size()) return CHAN_EINVAL;
    const auto& e = c->prof_out[idx];
    if (name && len > 0) {
        std::strncpy(name, e.first.c_str(), (size_t)len - 1);
        name[len - 1] = 0;
    }
    if (total_ms) *total_ms = e.second.first;
    if (launches) *launches = e.second.second;
    return CHAN_OK;
}

int chan_synchronize(chan_ctx* c)
{
    if (!c) return CHAN_EINVAL;
    int rc = set_device(c);
    if (rc) return rc;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return CHAN_OK;
}


// ------------------------------------------------------------------ ensemble
// Many independent channels (any sizes and precisions, one device) advanced
// together.  Members of one shape and the same step settings form a group:
// its step is one launch per stage for all of its members (the kernels'
// grouped form, each member's arguments in a device table), so a group of
// 30 small channels costs the launches of one channel.  Members that fit no
// group (another SGS model, a shape of their own) run their own kernels.
// One ensemble step (every group and lone member a branch between a fork
// and a join) is captured into a graph; three graphs cover the rotation of
// the members' state buffers, each with the argument tables of its phase.
struct chan_ensemble {
    std::string err;
    std::vector<chan_ctx*> ch;
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t fork = nullptr, done = nullptr;
    std::vector<hipEvent_t> join;
    hipGraphExec_t graph[3] = {nullptr, nullptr, nullptr};
    void* tabmem[3] = {nullptr, nullptr, nullptr};  // argument tables of each graph
    std::vector<int> cur0;    // members' rotation when the graphs were captured
    std::vector<long> gen0;   // members' generation then
    long step = 0;            // steps since the capture base
    int grouping = 1;         // 0: every member runs its own kernels
    std::vector<std::vector<int>> groups;  // members by group (a lone member: a group of one)
};

static int efail(chan_ensemble* e, int code, const std::string& msg)
{
    if (e) e->err = msg;
    return code;
}

static void ens_drop(chan_ensemble* e)
{
    for (int q = 0; q < 3; q++) {
        if (e->graph[q]) { (void)hipGraphExecDestroy(e->graph[q]); e->graph[q] = nullptr; }
        if (e->tabmem[q]) { (void)hipFree(e->tabmem[q]); e->tabmem[q] = nullptr; }
    }
    e->cur0.clear();
    e->gen0.clear();
    e->step = 0;
}

extern "C++" {
namespace {

// Argument tables of one graph, built in two passes over the same enqueue
// code: the first collects every stage's table in host memory (nothing is
// launched), the second, after the tables are in device memory, hands out
// their device addresses in the same order and launches.
struct TabSink {
    bool launch = false;
    std::vector<unsigned char> host;
    std::vector<size_t> offs;
    size_t next = 0;
    unsigned char* dev = nullptr;
    template <typename A>
    const A* put(const std::vector<A>& v)
    {
        if (!launch) {
            const size_t o = (host.size() + 255) & ~(size_t)255;
            host.resize(o + v.size() * sizeof(A));
            std::memcpy(host.data() + o, v.data(), v.size() * sizeof(A));
            offs.push_back(o);
            return nullptr;
        }
        return reinterpret_cast<const A*>(dev + offs[next++]);
    }
};

// the step settings that decide a channel's kernels and launch shapes
bool same_group(const chan_ctx* a, const chan_ctx* b)
{
    return a->prec == b->prec && a->d.nx == b->d.nx && a->d.ny == b->d.ny && a->d.nz == b->d.nz &&
           a->LES_model == 2 && b->LES_model == 2 && a->wall_mode == b->wall_mode && a->x_mass_cte == b->x_mass_cte &&
           a->rhs_shape == b->rhs_shape && a->dsm_shape == b->dsm_shape && a->dsm_kchunk == b->dsm_kchunk &&
           a->nblk_dsm == b->nblk_dsm &&
           a->nblk_red == b->nblk_red && a->pb.bx == b->pb.bx && a->pb.bz == b->pb.bz;
}

template <typename T>
__global__ void k_advance_time_g(const MassArgs<T>* tab) { tab[blockIdx.x].sc->t = tab[blockIdx.x].sc->t + tab[blockIdx.x].sc->dt; }

// compute_dt + compute_time_step_RK3 of every member of a group, one launch
// per stage (the arithmetic of enqueue_dt + enqueue_rk3 per member)
template <typename T>
int enqueue_group_step(chan_ensemble* e, const std::vector<int>& idx, TabSink& sink, hipStream_t s)
{
    const int n = (int)idx.size();
    chan_ctx* c0 = e->ch[idx[0]];
    const Dims& d = c0->d;
    std::vector<chan_ctx*> cs(n);
    std::vector<int> X(n);
    for (int q = 0; q < n; q++) {
        cs[q] = e->ch[idx[q]];
        X[q] = cs[q]->cur;
    }
    auto st = [&](int q, int buf) { return (T**)cs[q]->state[buf]; };
    {
        std::vector<DtArgs<T>> a(n);
        for (int q = 0; q < n; q++) {
            T** u = st(q, X[q]);
            a[q] = DtArgs<T>{u[0], u[1], u[2], cs[q]->md, cs[q]->part, cs[q]->sc, cs[q]->mins3};
        }
        const DtArgs<T>* t = sink.put(a);
        if (sink.launch) run_dt<T>(a[0], t, n, d, c0->nblk_red, s);
    }
    for (int sub = 0; sub < 3; sub++) {
        auto inb = [&](int q) { return sub == 0 ? X[q] : (sub == 1 ? (X[q] + 1) % 3 : (X[q] + 2) % 3); };
        auto outb = [&](int q) { return sub == 1 ? (X[q] + 2) % 3 : (X[q] + 1) % 3; };
        // wall stress into alpha_x
        if (c0->wall_mode == CHAN_WALL_EQUILIBRIUM) {
            std::vector<EqAlphaArgs<T>> a(n);
            for (int q = 0; q < n; q++) {
                chan_ctx* c = cs[q];
                a[q] = EqAlphaArgs<T>{st(q, inb(q))[0], c->alpha, c->nu, wall_rows(c, c->h_wm), c->kappa, c->Bc,
                                      loglaw_crossover(c->kappa, c->Bc)};
            }
            const EqAlphaArgs<T>* t = sink.put(a);
            if (sink.launch) run_equilibrium_alpha<T>(a[0], t, n, d, s);
        } else if (c0->wall_mode == CHAN_WALL_MLP) {
            std::vector<WmArgs<T>> a(n);
            for (int q = 0; q < n; q++) {
                chan_ctx* c = cs[q];
                const void* net = nullptr;
                int ndev = 0;
                if (!mlp_device_net(c->net, &net, &ndev)) return efail(e, CHAN_ESTATE, "a member's wall network is gone");
                T** u = st(q, inb(q));
                a[q] = WmArgs<T>{net, u[0], u[2], c->alpha, c->sc, c->md, c->nu, c->h_wm, c->kappa, c->Bc,
                                 loglaw_crossover(c->kappa, c->Bc), c->wm_max_err, c->wm_stats, c->x_mass_cte};
            }
            const WmArgs<T>* t = sink.put(a);
            if (sink.launch) run_wm_mlp<T>(a[0], t, n, d, s);
        }
        // dynamic Smagorinsky (pass A, plane ratios, pass B with the wall rows;
        // pass B left to the RHS in substeps 1 and 2, as enqueue_rk3 does)
        const bool defer = c0->defer_nut && sub < 2;
        {
            std::vector<DsmArgs<T>> a(n);
            std::vector<DsmRatioArgs> r(n);
            std::vector<DsmNutArgs<T>> v(n);
            for (int q = 0; q < n; q++) {
                chan_ctx* c = cs[q];
                T** u = st(q, inb(q));
                const DsmBufs<T> b = dsm_bufs<T>(c);
                a[q] = DsmArgs<T>{u[0], u[1], u[2], b};
                r[q] = DsmRatioArgs{b.part, b.ratio};
                r[q].fp32 = sizeof(T) == 4;
                if (defer) r[q].ntab = c->ntab;
                v[q] = DsmNutArgs<T>{defer ? b.smag : (T*)c->nut, b.smag, b.ratio, u[0], c->alpha, metric<T>(c),
                                     T(c->nu), c->Dirichlet_nu_t == 1, c->sc};
            }
            const DsmArgs<T>* ta = sink.put(a);
            const DsmRatioArgs* tr = sink.put(r);
            const DsmNutArgs<T>* tv = sink.put(v);
            if (sink.launch) {
                run_dsm_fused<T>(a[0], ta, n, d, c0->dsm_kchunk, c0->dsm_shape, s);
                run_dsm_ratio(r[0], tr, n, d, c0->nblk_dsm, s);
                if (defer) run_dsm_edges<T>(v[0], tv, n, d, s);
                else run_dsm_nut<T>(v[0], tv, n, d, 1, s);
            }
        }
        // RHS + RK combine
        {
            double c1, c2 = 0.0;
            if (sub == 0) c1 = c0->rk_coef[0][0];
            else if (sub == 1) { c1 = c0->rk_coef[1][0]; c2 = c0->rk_coef[1][1]; }
            else { c1 = c0->rk_coef[2][0]; c2 = c0->rk_coef[2][2]; }
            std::vector<RhsArgs<T>> a(n);
            for (int q = 0; q < n; q++) {
                chan_ctx* c = cs[q];
                T** I = st(q, inb(q));
                T** O = st(q, outb(q));
                T** Xs = st(q, X[q]);
                a[q] = RhsArgs<T>{I[0], I[1], I[2], defer ? (const T*)c->dsm[9] : (const T*)c->nut,
                                  defer ? (const T*)c->ntab : nullptr, rhs_tabs<T>(c), T(c->nu), c->sc, Xs[0], Xs[1],
                                  Xs[2], (T*)c->F1[0], (T*)c->F1[1], (T*)c->F1[2], O[0], O[1], O[2]};
            }
            const RhsArgs<T>* t = sink.put(a);
            if (sink.launch) run_rhs_rk<T>(a[0], t, n, d, sub + 1, c1, c2, c0->rhs_shape, s);
        }
        // projection of the new state
        {
            std::vector<FwdArgs<T>> f(n);
            std::vector<PoissonBufs> pbv(n);
            std::vector<ProjArgs<T>> pr(n);
            for (int q = 0; q < n; q++) {
                chan_ctx* c = cs[q];
                T** O = st(q, outb(q));
                f[q] = FwdArgs<T>{O[0], O[1], O[2], c->md, c->pb};
                pbv[q] = c->pb;
                pr[q] = ProjArgs<T>{O[0], O[1], O[2], c->pb.p, c->md};
            }
            const FwdArgs<T>* tf = sink.put(f);
            const PoissonBufs* tp = sink.put(pbv);
            const ProjArgs<T>* tr = sink.put(pr);
            if (sink.launch) {
                run_poisson_forward<T>(f[0], tf, tp, n, d, c0->pb, s);
                run_poisson_solve(c0->pb, tp, n, d, s);
                run_poisson_inverse(c0->pb, tp, n, d, s);
                run_project<T>(pr[0], tr, n, d, s);
            }
        }
    }
    // constant mass flow (its final reduction also advances the time), or the time advance alone
    {
        std::vector<MassArgs<T>> a(n);
        for (int q = 0; q < n; q++) a[q] = MassArgs<T>{st(q, (X[q] + 1) % 3)[0], cs[q]->part, cs[q]->sc, cs[q]->md};
        const MassArgs<T>* t = sink.put(a);
        if (sink.launch) {
            if (c0->x_mass_cte == 1 || c0->x_mass_cte == 2)
                run_mass_flow_correct<T>(a[0], t, n, d, c0->nblk_red, c0->x_mass_cte, 1, 1, s);
            else k_advance_time_g<T><<<n, 1, 0, s>>>(t);
        }
    }
    return CHAN_OK;
}

// one ensemble step of every group and lone member (sink: see TabSink)
int enqueue_ensemble_step(chan_ensemble* e, TabSink& sink, hipError_t& h)
{
    for (const std::vector<int>& g : e->groups) {
        chan_ctx* c0 = e->ch[g[0]];
        if (sink.launch) {
            h = hipStreamWaitEvent(c0->stream, e->fork, 0);
            if (h != hipSuccess) return CHAN_EHIP;
        }
        if (g.size() > 1) {
            const int rc = c0->prec == CHAN_FP64 ? enqueue_group_step<double>(e, g, sink, c0->stream)
                                                 : enqueue_group_step<float>(e, g, sink, c0->stream);
            if (rc) return rc;
        } else if (sink.launch) {
            DISPATCH(c0, enqueue_dt, c0);
            DISPATCH(c0, enqueue_rk3, c0, c0->cur);
        }
        if (sink.launch) {
            h = hipEventRecord(e->join[g[0]], c0->stream);
            if (h == hipSuccess) h = hipStreamWaitEvent(e->stream, e->join[g[0]], 0);
            if (h != hipSuccess) return CHAN_EHIP;
        }
    }
    return CHAN_OK;
}

void ens_form_groups(chan_ensemble* e)
{
    e->groups.clear();
    const int n = (int)e->ch.size();
    std::vector<int> taken(n, 0);
    for (int q = 0; q < n; q++) {
        if (taken[q]) continue;
        std::vector<int> g{q};
        taken[q] = 1;
        if (e->grouping)
            for (int r = q + 1; r < n; r++)
                if (!taken[r] && same_group(e->ch[q], e->ch[r])) {
                    g.push_back(r);
                    taken[r] = 1;
                }
        e->groups.push_back(g);
    }
}

}  // namespace
}  // extern "C++"

int chan_ensemble_create(chan_ctx* const* members, int n, chan_ensemble** out)
{
    if (!out) return CHAN_EINVAL;
    *out = nullptr;
    if (!members || n < 1) return CHAN_EINVAL;
    chan_ensemble* e = new chan_ensemble();
    *out = e;
    for (int q = 0; q < n; q++) {
        if (!members[q]) return efail(e, CHAN_EINVAL, "null member");
        for (int r = 0; r < q; r++)
            if (members[r] == members[q]) return efail(e, CHAN_EINVAL, "a channel appears twice");
        if (members[q]->device != members[0]->device) return efail(e, CHAN_EINVAL, "members on different devices");
        e->ch.push_back(members[q]);
    }
    e->device = members[0]->device;
    hipError_t h = hipSetDevice(e->device);
    if (h == hipSuccess) h = hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking);
    if (h == hipSuccess) h = hipEventCreateWithFlags(&e->fork, hipEventDisableTiming);
    if (h == hipSuccess) h = hipEventCreateWithFlags(&e->done, hipEventDisableTiming);
    e->join.assign(n, nullptr);
    for (int q = 0; q < n && h == hipSuccess; q++) h = hipEventCreateWithFlags(&e->join[q], hipEventDisableTiming);
    if (h != hipSuccess) return efail(e, CHAN_EHIP, std::string("ensemble set-up: ") + hipGetErrorString(h));
    return CHAN_OK;
}

void chan_ensemble_destroy(chan_ensemble* e)
{
    if (!e) return;
    (void)hipSetDevice(e->device);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    // the members' streams wait on this ensemble's events (run_steps orders
    // their later work after its steps): let them pass those waits before
    // the events go
    for (chan_ctx* c : e->ch)
        if (c && c->stream) (void)hipStreamSynchronize(c->stream);
    ens_drop(e);
    for (hipEvent_t ev : e->join)
        if (ev) (void)hipEventDestroy(ev);
    if (e->fork) (void)hipEventDestroy(e->fork);
    if (e->done) (void)hipEventDestroy(e->done);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
}

const char* chan_ensemble_last_error(const chan_ensemble* e) { return e ? e->err.c_str() : "null ensemble"; }

int chan_ensemble_size(const chan_ensemble* e) { return e ? (int)e->ch.size() : 0; }

int chan_ensemble_set_grouping(chan_ensemble* e, int on)
{
    if (!e) return CHAN_EINVAL;
    hipError_t h = hipSetDevice(e->device);
    if (h == hipSuccess) h = hipStreamSynchronize(e->stream);
    if (h != hipSuccess) return efail(e, CHAN_EHIP, std::string("ensemble grouping: ") + hipGetErrorString(h));
    e->grouping = on ? 1 : 0;
    ens_drop(e);
    return CHAN_OK;
}

int chan_ensemble_group_count(chan_ensemble* e)
{
    if (!e) return CHAN_EINVAL;
    ens_form_groups(e);
    return (int)e->groups.size();
}

int chan_ensemble_run_steps(chan_ensemble* e, int nsteps)
{
    if (!e || nsteps < 0) return CHAN_EINVAL;
    hipError_t h = hipSetDevice(e->device);
    if (h != hipSuccess) return efail(e, CHAN_EHIP, hipGetErrorString(h));
    const int n = (int)e->ch.size();
    // the graphs are valid while every member's rotation and settings are
    // where the capture left them
    bool stale = e->cur0.empty();
    for (int q = 0; q < n && !stale; q++)
        stale = e->ch[q]->generation != e->gen0[q] || e->ch[q]->cur != (e->cur0[q] + e->step) % 3;
    if (stale) {
        // a table may still be read by a step in flight
        h = hipStreamSynchronize(e->stream);
        if (h != hipSuccess) return efail(e, CHAN_EHIP, std::string("ensemble sync: ") + hipGetErrorString(h));
        ens_drop(e);
        for (chan_ctx* c : e->ch) { e->cur0.push_back(c->cur); e->gen0.push_back(c->generation); }
        ens_form_groups(e);
    }
    // order after whatever the members' own streams hold and after the last
    // change of their wall networks
    for (int q = 0; q < n; q++) {
        h = hipEventRecord(e->join[q], e->ch[q]->stream);
        if (h == hipSuccess) h = hipStreamWaitEvent(e->stream, e->join[q], 0);
        if (h == hipSuccess && e->ch[q]->wall_mode == CHAN_WALL_MLP && e->ch[q]->net)
            h = hipStreamWaitEvent(e->stream, mlp_ready(e->ch[q]->net), 0);
        if (h != hipSuccess) return efail(e, CHAN_EHIP, std::string("ensemble order: ") + hipGetErrorString(h));
    }
    for (int it = 0; it < nsteps; it++) {
        const int o = (int)(e->step % 3);
        if (!e->graph[o]) {
            // pass 1: the argument tables of this phase, into device memory
            TabSink sink;
            int rc = enqueue_ensemble_step(e, sink, h);
            if (rc) { ens_drop(e); return rc == CHAN_EHIP ? efail(e, rc, hipGetErrorString(h)) : rc; }
            if (!sink.host.empty()) {
                h = hipMalloc(&e->tabmem[o], sink.host.size());
                if (h == hipSuccess)
                    h = hipMemcpy(e->tabmem[o], sink.host.data(), sink.host.size(), hipMemcpyHostToDevice);
                if (h != hipSuccess) {
                    ens_drop(e);
                    return efail(e, CHAN_EHIP, std::string("ensemble tables: ") + hipGetErrorString(h));
                }
            }
            // pass 2: the launches, captured
            sink.launch = true;
            sink.dev = (unsigned char*)e->tabmem[o];
            hipGraph_t g;
            h = hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal);
            if (h == hipSuccess) h = hipEventRecord(e->fork, e->stream);
            if (h == hipSuccess) {
                rc = enqueue_ensemble_step(e, sink, h);
                if (rc && h == hipSuccess) h = hipErrorUnknown;
            }
            hipError_t h2 = hipStreamEndCapture(e->stream, &g);
            if (h == hipSuccess) h = h2;
            if (h == hipSuccess) {
                h = hipGraphInstantiate(&e->graph[o], g, nullptr, nullptr, 0);
                (void)hipGraphDestroy(g);
            }
            if (h != hipSuccess) {
                ens_drop(e);
                return efail(e, CHAN_EHIP, std::string("ensemble capture: ") + hipGetErrorString(h) +
                                               (e->err.empty() ? "" : " (" + e->err + ")"));
            }
        }
        h = hipGraphLaunch(e->graph[o], e->stream);
        if (h != hipSuccess) return efail(e, CHAN_EHIP, std::string("ensemble launch: ") + hipGetErrorString(h));
        for (chan_ctx* c : e->ch) c->cur = (c->cur + 1) % 3;
        e->step++;
    }
    // members' later work is ordered after the ensemble's steps
    h = hipEventRecord(e->done, e->stream);
    for (int q = 0; q < n && h == hipSuccess; q++) h = hipStreamWaitEvent(e->ch[q]->stream, e->done, 0);
    if (h != hipSuccess) return efail(e, CHAN_EHIP, std::string("ensemble order: ") + hipGetErrorString(h));
    return CHAN_OK;
}

int chan_ensemble_synchronize(chan_ensemble* e)
{
    if (!e) return CHAN_EINVAL;
    hipError_t h = hipSetDevice(e->device);
    if (h == hipSuccess) h = hipStreamSynchronize(e->stream);
    if (h != hipSuccess) return efail(e, CHAN_EHIP, std::string("ensemble sync: ") + hipGetErrorString(h));
    return CHAN_OK;
}

}  // extern "C"
