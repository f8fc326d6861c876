// stencil.hip -- ghost fill, RHS + Wray RK3 combine, dynamic Smagorinsky,
// wall model, mass-flow forcing, dt and divergence reductions.
//
// FP64 arithmetic keeps the reference's operation order and is compiled
// without FMA contraction, so the stencil stages reproduce the reference's
// rounding.  Reductions are two-level (fixed block partials, then one block
// in fixed order) and therefore deterministic from run to run.
#include "kernels.hpp"
#include "prof.hpp"
#include "wallmodel.hpp"
#include <algorithm>
#include <type_traits>

namespace chan {

static constexpr int RED_THREADS = 256;

// ------------------------------------------------------------ ghost fill

// code/boundary_conditions.f90:20-48 as one pass: the reference applies
// periodic x, periodic z, then the y wall condition; the composition maps
// every ghost cell to sign * F(interior image).  Sources are never ghosts,
// so the pass is free of read/write races.
template <typename T>
__global__ void k_bc(T* F, Dims d, int kind)
{
    const bool xf = kind == KIND_U, zf = kind == KIND_W, yf = kind == KIND_V;
    const int n1 = xf ? d.nx : d.nxg, n2 = yf ? d.ny : d.nyg, n3 = zf ? d.nz : d.nzg;
    const int gx = xf ? 2 : 3, gz = zf ? 2 : 3;
    const long ny_cnt = 2L * n1 * n3;
    const long nz_cnt = (long)gz * n1 * (n2 - 2);
    const long nx_cnt = (long)gx * (n2 - 2) * (n3 - gz);
    long q = (long)blockIdx.x * blockDim.x + threadIdx.x;
    int i, j, k;
    auto ghost_x = [&](int g) { return g == 0 ? 1 : (xf ? n1 : (g == 1 ? n1 - 1 : n1)); };
    auto ghost_z = [&](int g) { return g == 0 ? 1 : (zf ? n3 : (g == 1 ? n3 - 1 : n3)); };
    if (q < ny_cnt) {
        i = (int)(q % n1) + 1;
        long r = q / n1;
        k = (int)(r % n3) + 1;
        j = (r / n3) == 0 ? 1 : n2;
    } else if ((q -= ny_cnt) < nz_cnt) {
        i = (int)(q % n1) + 1;
        long r = q / n1;
        j = (int)(r % (n2 - 2)) + 2;
        k = ghost_z((int)(r / (n2 - 2)));
    } else if ((q -= nz_cnt) < nx_cnt) {
        j = (int)(q % (n2 - 2)) + 2;
        long r = q / (n2 - 2);
        k = (int)(r % (n3 - gz)) + 2;
        i = ghost_x((int)(r / (n3 - gz)));
    } else {
        return;
    }
    const int si = xf ? img_face(i, n1) : img_centre(i, n1);
    const int sk = zf ? img_face(k, n3) : img_centre(k, n3);
    int sj = j;
    T sign = T(1);
    if (j == 1 || j == n2) {
        if (yf) { F[at(d, i, j, k)] = T(0); return; }
        sj = (j == 1) ? 2 : n2 - 1;
        sign = (kind == KIND_C_NEUMANN) ? T(1) : T(-1);
    }
    F[at(d, i, j, k)] = sign * F[at(d, si, sj, sk)];
}

template <typename T>
void launch_bc(T* F, const Dims& d, int kind, hipStream_t s)
{
    const bool xf = kind == KIND_U, zf = kind == KIND_W, yf = kind == KIND_V;
    const int n1 = xf ? d.nx : d.nxg, n2 = yf ? d.ny : d.nyg, n3 = zf ? d.nz : d.nzg;
    const int gx = xf ? 2 : 3, gz = zf ? 2 : 3;
    long tot = 2L * n1 * n3 + (long)gz * n1 * (n2 - 2) + (long)gx * (n2 - 2) * (n3 - gz);
    PROF_BEGIN("k_bc", s);
    k_bc<T><<<cdiv(tot, 256), 256, 0, s>>>(F, d, kind);
    PROF_END(s);
}

// The RHS + RK combine kernel is rhs.hip.

// --------------------------------------------------------------- SGS: DSM
// Pass A (filters, L_ij, M_ij, plane partial sums, |S|) is dsm_fused.hip.

// <L:M>/<M:M> per row: one wave per row, the blocks' partials summed in a
// fixed order (strided per lane, then a butterfly across the wave), so the
// result does not depend on scheduling; no LDS and no barrier.  The row
// waves also write pass B's coefficients when it is deferred (ar.ntab).
// With the DSM pass's compute_dt maxima (ar.dtpart), one more block reduces
// them and sets the step's dt (the maximum does not depend on the order);
// with ar.eq_alpha, the blocks after it compute the equilibrium wall stress.
__global__ void __launch_bounds__(64) k_dsm_ratio(DsmRatioArgs one, const DsmRatioArgs* tab, int nblk, Dims d)
{
    const GBlk g = gblk(1);
    const DsmRatioArgs ar = member_args(one, tab, g.m);
    const int t = threadIdx.x;
    const int nyg = d.nyg;
    const int neq0 = nyg - 2 + (ar.dtpart ? 1 : 0);  // the equilibrium stress's first block
    if (g.x >= neq0) {
        // the equilibrium wall stress, 64 cells per block (k_equilibrium_alpha)
        const int q = (g.x - neq0) * 64 + t;
        if (q >= 2 * d.nx * d.nzg) return;
        const WallRows ws{ar.eq_jb, ar.eq_wb, ar.eq_hb, ar.eq_ht};
        if (ar.fp32)
            eq_alpha_cell<float>(q, (const float*)ar.eq_U, ar.eq_alpha, d, ar.eq_nu, ar.eq_kappa, ar.eq_B, ar.eq_yc, ws);
        else
            eq_alpha_cell<double>(q, (const double*)ar.eq_U, ar.eq_alpha, d, ar.eq_nu, ar.eq_kappa, ar.eq_B, ar.eq_yc,
                                  ws);
        return;
    }
    if (g.x == nyg - 2) {
        const double* dp = ar.dtpart;
        const int nb = ar.nblk_dt;
        double m0 = 0.0, m1 = 0.0, m2 = 0.0;
        for (int q = t; q < nb; q += 64) {
            m0 = fmax(m0, dp[q]);
            m1 = fmax(m1, dp[nb + q]);
            m2 = fmax(m2, dp[2L * nb + q]);
        }
        for (int m = 32; m > 0; m >>= 1) {
            m0 = fmax(m0, __shfl_xor(m0, m, 64));
            m1 = fmax(m1, __shfl_xor(m1, m, 64));
            m2 = fmax(m2, __shfl_xor(m2, m, 64));
        }
        if (t == 0) ar.sc->dt = dt_from_maxima(m0, m1, m2, ar.sc, ar.mins3);
        return;
    }
    const double* part = ar.part;
    double* ratio = ar.ratio;
    const int j = g.x + 2;
    double a = 0.0, c = 0.0;
    for (int q = t; q < nblk; q += 64) {
        a += part[((long)(j - 1) * 2 + 0) * nblk + q];
        c += part[((long)(j - 1) * 2 + 1) * nblk + q];
    }
    for (int m = 32; m > 0; m >>= 1) {
        a += __shfl_xor(a, m, 64);
        c += __shfl_xor(c, m, 64);
    }
    if (t == 0) {
        const double rt = a / c;
        ratio[j - 1] = rt;
        ratio[nyg + j - 1] = a;
        ratio[2 * nyg + j - 1] = c;
        if (ar.ntab) {
            // deferred pass B: row j (and the ghost row beyond rows 2 and
            // nyg-1) as k_dsm_nut computes it from the value there; the
            // wall-model rows 1, 2, nyg-1, nyg hold nu_t and pass through
            auto put = [&](auto* tab, int row) {
                using V = std::remove_pointer_t<decltype(tab)>;
                const bool wm = row <= 2 || row >= nyg - 1;
                tab[row - 1] = wm ? V(1) : V(-0.5) * V(rt);
                tab[nyg + row - 1] = wm ? V(-INFINITY) : V(0);
            };
            auto rows = [&](auto* tab) {
                put(tab, j);
                if (j == 2) put(tab, 1);
                if (j == nyg - 1) put(tab, nyg);
            };
            if (ar.fp32) rows((float*)ar.ntab);
            else rows((double*)ar.ntab);
        }
    }
}

// code/subgrid.f90:512-535: nu_t = max(-1/2 <LM>/<MM> |S|, 0), y ghost by
// Dirichlet_nu_t; with fuse_wm the wall rows come from the wall model
// (code/wallmodel.f90:39-42), which the reference applies right after.
// Each lane takes one (x, y) position of NUT_PLANES consecutive z planes and
// issues the planes' loads before any store (restrict pointers: the stores
// cannot alias them), so a wave keeps several loads in flight.
static constexpr int NUT_PLANES = 4;

template <typename T>
__global__ void __launch_bounds__(256) k_dsm_nut(DsmNutArgs<T> one, const DsmNutArgs<T>* tab, Dims d, int fuse_wm)
{
    const int gz = (d.nzg + NUT_PLANES - 1) / NUT_PLANES;
    const GBlk g = gblk(gz);
    const DsmNutArgs<T> a = member_args(one, tab, g.m);
    T* __restrict__ nut = a.nut;
    const T* __restrict__ smag = a.smag;
    // one flattened (x, y) plane per block row, so that every lane has a cell
    const int q = g.x * blockDim.x + threadIdx.x;
    if (q >= d.nxg * d.nyg) return;
    const int jm = q / d.nxg;
    const int i = q - jm * d.nxg + 1, j = jm + 1;
    const int k0 = g.z * NUT_PLANES + 1;
    const int np = min(NUT_PLANES, d.nzg - k0 + 1);
    const long p0 = at(d, i, j, k0);
    if (fuse_wm && (j <= 2 || j >= d.nyg - 1)) {
        for (int t = 0; t < np; t++)
            nut[p0 + t * d.sxy] =
                wm_nut_at<T>(a.U, a.alpha, a.m, d, a.nu, img_centre(i, d.nxg), k0 + t, j <= 2 ? 0 : 1, a.sc);
        return;
    }
    const int jj = (j == 1) ? 2 : (j == d.nyg ? d.nyg - 1 : j);
    const long s0 = at(d, img_centre(i, d.nxg), jj, 1);
    T v[NUT_PLANES];
#pragma unroll
    for (int t = 0; t < NUT_PLANES; t++) {
        // planes past the box read the last one and are not stored
        const int k = k0 + min(t, np - 1);
        v[t] = smag[s0 + (long)(img_centre(k, d.nzg) - 1) * d.sxy];
    }
    const T r = T(a.ratio[jj - 1]);
    const bool flip = jj != j && a.dirichlet;
#pragma unroll
    for (int t = 0; t < NUT_PLANES; t++) {
        const T x = fmax(T(-0.5) * r * v[t], T(0));
        if (t < np) nut[p0 + t * d.sxy] = flip ? -x : x;
    }
}

// The rest of a deferred pass B (DsmBufs::ntab): in the |S| box itself,
// which pass A filled on the unique cells of rows 2..nyg-1,
//   the wall model's nu_t on rows 1, 2, nyg-1, nyg (the values k_dsm_nut
//   writes there with fuse_wm), every cell with its periodic images; then
//   the x image columns 1, nxg-1, nxg and z image planes 1, nzg-1, nzg of
//   rows 3..nyg-2, from their unique cells.
// No thread reads a cell that another one writes.
template <typename T>
__global__ void __launch_bounds__(256) k_dsm_edges(DsmNutArgs<T> one, const DsmNutArgs<T>* tab, Dims d)
{
    const GBlk g = gblk(1);
    const DsmNutArgs<T> a = member_args(one, tab, g.m);
    T* S = a.nut;  // the |S| box (the launcher passes it as nut)
    const long q = (long)g.x * blockDim.x + threadIdx.x;
    const int nxz = d.nxg * d.nzg;
    if (q < 2L * nxz) {
        const int side = (int)(q / nxz), r = (int)(q - (long)side * nxz);
        const int k = r / d.nxg + 1, i = r - (k - 1) * d.nxg + 1;
        const T v = wm_nut_at<T>(a.U, a.alpha, a.m, d, a.nu, img_centre(i, d.nxg), k, side, a.sc);
        const int j0 = side == 0 ? 1 : d.nyg - 1;
        S[at(d, i, j0, k)] = v;
        S[at(d, i, j0 + 1, k)] = v;
        return;
    }
    const int npos = 3 * d.nzg + 3 * (d.nxg - 3);
    const long t = q - 2L * nxz;
    if (t >= (long)npos * (d.nyg - 4)) return;
    const int row = (int)(t / npos), pos = (int)(t - (long)row * npos);
    const int j = row + 3;
    int i, k;
    if (pos < 3 * d.nzg) {
        const int c = pos / d.nzg;
        i = c == 0 ? 1 : d.nxg - 2 + c;
        k = pos - c * d.nzg + 1;
    } else {
        const int p2 = pos - 3 * d.nzg, c = p2 / (d.nxg - 3);
        k = c == 0 ? 1 : d.nzg - 2 + c;
        i = p2 - c * (d.nxg - 3) + 2;
    }
    S[at(d, i, j, k)] = S[at(d, img_centre(i, d.nxg), j, img_centre(k, d.nzg))];
}

template <typename T>
void run_dsm_edges(const DsmNutArgs<T>& one, const DsmNutArgs<T>* tab, int nmem, const Dims& d, hipStream_t s)
{
    const int npos = 3 * d.nzg + 3 * (d.nxg - 3);
    const long n = 2L * d.nxg * d.nzg + (long)npos * (d.nyg - 4);
    PROF_BEGIN("k_dsm_edges", s);
    k_dsm_edges<T><<<dim3(cdiv(n, 256), 1, nmem), 256, 0, s>>>(one, tab, d);
    PROF_END(s);
}

void run_dsm_ratio(const DsmRatioArgs& one, const DsmRatioArgs* tab, int nmem, const Dims& d, int nblk,
                   hipStream_t s)
{
    PROF_BEGIN("k_dsm_ratio", s);
    const int neq = one.eq_alpha ? cdiv(2L * d.nx * d.nzg, 64) : 0;
    k_dsm_ratio<<<dim3(d.nyg - 2 + (one.dtpart ? 1 : 0) + neq, 1, nmem), 64, 0, s>>>(one, tab, nblk, d);
    PROF_END(s);
}

template <typename T>
void run_dsm_nut(const DsmNutArgs<T>& one, const DsmNutArgs<T>* tab, int nmem, const Dims& d, int fuse_wm,
                 hipStream_t s)
{
    dim3 grd(cdiv((long)d.nxg * d.nyg, 256), 1, (unsigned)cdiv(d.nzg, NUT_PLANES) * nmem);
    PROF_BEGIN("k_dsm_nut", s);
    k_dsm_nut<T><<<grd, 256, 0, s>>>(one, tab, d, fuse_wm);
    PROF_END(s);
}

template <typename T>
void launch_dsm(const T* U, const T* V, const T* W, T* nut, const Metric<T>& m, const Dims& d, DsmBufs<T>& b,
                int dirichlet, int fuse_wm, const double* alpha_x, T nu, const Scalars* sc, hipStream_t s,
                const EqAlphaArgs<T>* eq)
{
    launch_dsm_fused<T>(U, V, W, m, d, b, s);
    DsmRatioArgs ra{b.part, b.ratio, b.dtpart, b.nblk_dt, const_cast<Scalars*>(sc), b.dtpart ? b.mins3 : nullptr};
    ra.fp32 = sizeof(T) == 4;
    if (eq) {
        // the equilibrium wall stress in the same launch (it is read first by
        // the pass B that follows)
        ra.eq_U = eq->U;
        ra.eq_alpha = eq->alpha;
        ra.eq_nu = eq->nu; ra.eq_kappa = eq->kappa; ra.eq_B = eq->B; ra.eq_yc = eq->yc;
        ra.eq_jb = eq->ws.jb; ra.eq_wb = eq->ws.wb; ra.eq_hb = eq->ws.hb; ra.eq_ht = eq->ws.ht;
    }
    // pass B deferred to the consumer of nu_t (the wall rows from the wall
    // model only; the step has them so)
    const bool defer = b.ntab && fuse_wm;
    if (defer) ra.ntab = b.ntab;
    run_dsm_ratio(ra, nullptr, 1, d, b.nblk, s);
    const DsmNutArgs<T> na{nut, b.smag, b.ratio, U, alpha_x, m, nu, dirichlet, sc};
    if (defer) {
        DsmNutArgs<T> ea = na;
        ea.nut = b.smag;  // completed in place
        run_dsm_edges<T>(ea, nullptr, 1, d, s);
    } else {
        run_dsm_nut<T>(na, nullptr, 1, d, fuse_wm, s);
    }
}

// code/wallmodel.f90:23-46 on its own: wall rows from alpha_x, then the
// x-periodic fill of every other row of nu_t.
template <typename T>
__global__ void k_wall_model(const T* U, T* nut, const double* alpha, Metric<T> m, Dims d, T nu, const Scalars* sc)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > d.nxg || j > d.nyg) return;
    const bool wallrow = (j <= 2 || j >= d.nyg - 1);
    if (wallrow) {
        nut[at(d, i, j, k)] = wm_nut_at<T>(U, alpha, m, d, nu, img_centre(i, d.nxg), k, j <= 2 ? 0 : 1, sc);
    } else if (i == 1 || i >= d.nxg - 1) {
        nut[at(d, i, j, k)] = nut[at(d, img_centre(i, d.nxg), j, k)];
    }
}

template <typename T>
void launch_wall_model(const T* U, T* nut, const double* alpha_x, const Metric<T>& m, const Dims& d, T nu,
                       const Scalars* sc, hipStream_t s)
{
    dim3 blk(64, 4, 1), grd(cdiv(d.nxg, 64), cdiv(d.nyg, 4), d.nzg);
    PROF_BEGIN("k_wall_model", s);
    k_wall_model<T><<<grd, blk, 0, s>>>(U, nut, alpha_x, m, d, nu, sc);
    PROF_END(s);
}

// ------------------------------------------------ equilibrium wall stress


template <typename T>
__global__ void __launch_bounds__(256) k_equilibrium_alpha(EqAlphaArgs<T> one, const EqAlphaArgs<T>* tab, Dims d)
{
    const GBlk g = gblk(1);
    const EqAlphaArgs<T> a = member_args(one, tab, g.m);
    const int q = g.x * blockDim.x + threadIdx.x;
    if (q >= 2 * d.nx * d.nzg) return;
    eq_alpha_cell<T>(q, a.U, a.alpha, d, a.nu, a.kappa, a.B, a.yc, a.ws);
}

template <typename T>
void launch_equilibrium_alpha(const T* U, double* alpha_x, const Dims& d, double nu, const WallRows& ws,
                              double kappa, double B, hipStream_t s)
{
    run_equilibrium_alpha<T>(EqAlphaArgs<T>{U, alpha_x, nu, ws, kappa, B, loglaw_crossover(kappa, B)}, nullptr, 1, d,
                             s);
}

template <typename T>
void run_equilibrium_alpha(const EqAlphaArgs<T>& one, const EqAlphaArgs<T>* tab, int nmem, const Dims& d,
                           hipStream_t s)
{
    PROF_BEGIN("k_equilibrium_alpha", s);
    k_equilibrium_alpha<T><<<dim3(cdiv(2L * d.nx * d.nzg, 256), 1, nmem), 256, 0, s>>>(one, tab, d);
    PROF_END(s);
}

// ------------------------------------------------------------- reductions

// block sum of two doubles into part[blk], part[nblk + blk]
__device__ __forceinline__ void block_sum2(double a, double b, double* part, int nblk, int blk)
{
    __shared__ double red[2][RED_THREADS];
    const int t = threadIdx.x;
    red[0][t] = a; red[1][t] = b;
    __syncthreads();
    for (int st = RED_THREADS / 2; st > 0; st >>= 1) {
        if (t < st) { red[0][t] += red[0][t + st]; red[1][t] += red[1][t + st]; }
        __syncthreads();
    }
    if (t == 0) { part[blk] = red[0][0]; part[nblk + blk] = red[1][0]; }
}

__device__ __forceinline__ void final_sum2(const double* part, int nblk, double& a, double& b)
{
    __shared__ double red[2][RED_THREADS];
    const int t = threadIdx.x;
    double x = 0.0, y = 0.0;
    for (int q = t; q < nblk; q += RED_THREADS) { x += part[q]; y += part[nblk + q]; }
    red[0][t] = x; red[1][t] = y;
    __syncthreads();
    for (int st = RED_THREADS / 2; st > 0; st >>= 1) {
        if (t < st) { red[0][t] += red[0][t + st]; red[1][t] += red[1][t + st]; }
        __syncthreads();
    }
    a = red[0][0]; b = red[1][0];
}

// code/mass_flow.f90:24-62 (trapezoid bulk flow between yg(2) and yg(nyg-1))
template <typename T>
__global__ void k_massflow_u(const T* U, Metric<double> m, Dims d, double* part, int nblk)
{
    const int n1 = d.nx - 2, n2 = d.nyg - 3, n3 = d.nzp;
    const int tot = n1 * n2 * n3;
    double q = 0.0, nrm = 0.0;
    for (int t = blockIdx.x * RED_THREADS + threadIdx.x; t < tot; t += nblk * RED_THREADS) {
        const int r = t / n1;
        const int i = t - r * n1 + 2;
        const int j = r % n2 + 3;
        const int k = r / n2 + 2;
        const double dy = m.yg[j] - m.yg[j - 1];
        q += ((double)U[at(d, i, j, k)] + (double)U[at(d, i, j - 1, k)]) * 0.5 * dy;
        nrm += (1.0 + 1.0) * 0.5 * dy;
    }
    block_sum2(q, nrm, part, nblk, blockIdx.x);
}

__global__ void k_massflow_u_final(const double* part, int nblk, Scalars* sc)
{
    double a, b;
    final_sum2(part, nblk, a, b);
    if (threadIdx.x == 0) sc->Qflow_x_0 = a / b;
}

template <typename T>
void launch_mass_flow_init(const T* U, const Metric<double>& md, const Dims& d, double* part, int nblk, Scalars* sc,
                           hipStream_t s)
{
    PROF_BEGIN("k_massflow_u", s);
    k_massflow_u<T><<<nblk, RED_THREADS, 0, s>>>(U, md, d, part, nblk);
    PROF_END(s);
    PROF_BEGIN("k_massflow_u_final", s);
    k_massflow_u_final<<<1, RED_THREADS, 0, s>>>(part, nblk, sc);
    PROF_END(s);
}

// code/mass_flow.f90:113-166: mean of 0.5(U(:,3,:)+U(:,nyg-2,:)) over the
// unique x faces and z centres; dPdx = Qflow_x_0 - mean.
template <typename T>
__global__ void k_dpx_partial(MassArgs<T> one, const MassArgs<T>* tab, Dims d, int nblk)
{
    const GBlk g = gblk(1);
    const MassArgs<T> ar = member_args(one, tab, g.m);
    const T* U = ar.U;
    const double* p = ar.p;
    const int n1 = d.nx - 2, n3 = d.nzp, tot = n1 * n3;
    // U at face i of row j, or its projection by p as k_project computes it
    auto u = [&](int i, int j, int k) {
        const long q = at(d, i, j, k);
        return p ? T((double)U[q] - (p[at(d, i + 1, j, k)] - p[q]) * ar.m.rdxg[i + 1]) : U[q];
    };
    double s = 0.0, c = 0.0;
    for (int t = g.x * RED_THREADS + threadIdx.x; t < tot; t += nblk * RED_THREADS) {
        const int kq = t / n1;
        const int i = t - kq * n1 + 2, k = kq + 2;
        s += 0.5 * ((double)u(i, 3, k) + (double)u(i, d.nyg - 2, k));
        c += 1.0;
    }
    block_sum2(s, c, ar.part, nblk, g.x);
}

template <typename T>
__global__ void k_dpx_final(MassArgs<T> one, const MassArgs<T>* tab, int nblk, int advance)
{
    const GBlk g = gblk(1);
    const MassArgs<T> ar = member_args(one, tab, g.m);
    Scalars* sc = ar.sc;
    double a, b;
    final_sum2(ar.part, nblk, a, b);
    if (threadIdx.x == 0) {
        sc->last_dPdx = sc->Qflow_x_0 - a / b;
        sc->dPdx = 0.0;  // code/time_integration.f90:92
        if (advance) sc->t = sc->t + sc->dt;  // the step's time advance, one launch fewer
    }
}

// U(2:nx-1,2:nyg-1,2:nzg-1) += dPdx (code/time_integration.f90:91).  Each
// lane takes SHIFT_PLANES consecutive z planes of one (x, y) position, all
// loads issued before the first store.
static constexpr int SHIFT_PLANES = 4;

template <typename T>
__global__ void k_shift_u(MassArgs<T> one, const MassArgs<T>* tab, Dims d)
{
    const GBlk g = gblk((d.nzg - 3 + SHIFT_PLANES - 1) / SHIFT_PLANES);
    const MassArgs<T> ar = member_args(one, tab, g.m);
    T* U = ar.U;
    const Scalars* sc = ar.sc;
    const int i = g.x * blockDim.x + threadIdx.x + 2;
    const int j = g.y * blockDim.y + threadIdx.y + 2;
    const int k0 = g.z * SHIFT_PLANES + 2;
    if (i > d.nx - 1 || j > d.nyg - 1) return;
    const int np = min(SHIFT_PLANES, d.nzg - 2 - k0 + 1);
    const long p0 = at(d, i, j, k0);
    T v[SHIFT_PLANES];
#pragma unroll
    for (int t = 0; t < SHIFT_PLANES; t++) v[t] = U[p0 + min(t, np - 1) * d.sxy];
    const T s = T(sc->last_dPdx);
#pragma unroll
    for (int t = 0; t < SHIFT_PLANES; t++)
        if (t < np) store_img(U, d, 0, i, j, k0 + t, v[t] + s);
}

// Constant bulk flow (mass-flow mode 2; no reference counterpart: the
// reference's functional above holds the mean of two near-wall rows, under
// which the bulk of a developing flow drifts): the shift that restores the
// trapezoid bulk mean of code/mass_flow.f90:24-62, taken over every interior
// row, to Qflow_x_0.  The shift moves that mean by exactly itself.
template <typename T>
__global__ void k_bulk_partial(MassArgs<T> one, const MassArgs<T>* tab, Dims d, int nblk)
{
    const GBlk g = gblk(1);
    const MassArgs<T> ar = member_args(one, tab, g.m);
    const T* U = ar.U;
    const double* yg = ar.m.yg;
    const int n1 = d.nx - 2, n2 = d.nyg - 3, n3 = d.nzp;
    const int tot = n1 * n2 * n3;
    double q = 0.0, nrm = 0.0;
    // term t is the cell (t % n1 + 2, (t / n1) % n2 + 3, t / (n1 n2) + 2); a
    // thread's terms are S apart, so its indices follow by carries
    const StrideCarry sc = stride_carry(nblk * RED_THREADS, n1, n2);
    int t = g.x * RED_THREADS + threadIdx.x;
    CellIdx c = cell_of(t, n1, n2);
    for (; t < tot; t += nblk * RED_THREADS) {
        const int i = c.i + 2, j = c.j + 3, k = c.k + 2;
        const double dy = yg[j] - yg[j - 1];
        q += ((double)U[at(d, i, j, k)] + (double)U[at(d, i, j - 1, k)]) * 0.5 * dy;
        nrm += (1.0 + 1.0) * 0.5 * dy;
        c = advance(c, sc, n1, n2);
    }
    block_sum2(q, nrm, ar.part, nblk, g.x);
}

template <typename T>
void run_mass_flow_correct(const MassArgs<T>& one, const MassArgs<T>* tab, int nmem, const Dims& d, int nblk,
                           int mode, int apply, int advance, hipStream_t s)
{
    PROF_BEGIN("k_dpx_partial", s);
    if (mode == 2) k_bulk_partial<T><<<dim3(nblk, 1, nmem), RED_THREADS, 0, s>>>(one, tab, d, nblk);
    else k_dpx_partial<T><<<dim3(nblk, 1, nmem), RED_THREADS, 0, s>>>(one, tab, d, nblk);
    PROF_END(s);
    PROF_BEGIN("k_dpx_final", s);
    k_dpx_final<T><<<dim3(1, 1, nmem), RED_THREADS, 0, s>>>(one, tab, nblk, advance);
    PROF_END(s);
    if (apply) {
        // canonical cells only (k up to nzg-2), images written by store_img
        const dim3 blk = row_block(d.nx - 2);
        const dim3 grd(cdiv(d.nx - 2, blk.x), cdiv(d.nyg - 2, blk.y), (unsigned)cdiv(d.nzg - 3, SHIFT_PLANES) * nmem);
        PROF_BEGIN("k_shift_u", s);
        k_shift_u<T><<<grd, blk, 0, s>>>(one, tab, d);
        PROF_END(s);
    }
}

template <typename T>
void launch_mass_flow_correct(T* U, const Metric<double>& md, const Dims& d, double* part, int nblk, Scalars* sc,
                              int mode, int apply, hipStream_t s, int advance)
{
    run_mass_flow_correct<T>(MassArgs<T>{U, part, sc, md}, nullptr, 1, d, nblk, mode, apply, advance, s);
}

template <typename T>
void launch_mass_flow_of_projection(const T* U, const double* p, const Metric<double>& md, const Dims& d,
                                    double* part, int nblk, Scalars* sc, hipStream_t s)
{
    // mode 1, no shift applied here, the step's time advanced; one lane per
    // term of the functional ((nx-2) nzp of them), at most nblk blocks
    const int nb = std::min(nblk, cdiv((long)(d.nx - 2) * d.nzp, RED_THREADS));
    run_mass_flow_correct<T>(MassArgs<T>{const_cast<T*>(U), part, sc, md, p}, nullptr, 1, d, nb, 1, 0, 1, s);
}

// code/time_integration.f90:102-150 (max of dx/|u| as the reference; max is
// order independent, so the result is exact).
__device__ __forceinline__ void block_max3(double a, double b, double c, double* part, int nblk, int blk)
{
    __shared__ double red[3][RED_THREADS];
    const int t = threadIdx.x;
    red[0][t] = a; red[1][t] = b; red[2][t] = c;
    __syncthreads();
    for (int st = RED_THREADS / 2; st > 0; st >>= 1) {
        if (t < st) {
            red[0][t] = fmax(red[0][t], red[0][t + st]);
            red[1][t] = fmax(red[1][t], red[1][t + st]);
            red[2][t] = fmax(red[2][t], red[2][t + st]);
        }
        __syncthreads();
    }
    if (t == 0) {
        part[blk] = red[0][0];
        part[nblk + blk] = red[1][0];
        part[2 * nblk + blk] = red[2][0];
    }
}

// The unique cells of the reference's loop (i, j, k in 2..nxg-1, 2..nyg-1,
// 2..nzg-1), flattened with x fastest, are split into nblk equal runs, one per
// block; every lane takes a cell (rows of a few dozen cells would leave most
// of a 256-lane block idle), and each lane's next four cells are loaded
// before any is used.  The maximum does not depend on the order, so the
// result is the same as any other traversal.
template <typename T>
__global__ void k_dt_partial(DtArgs<T> one, const DtArgs<T>* tab, Dims d, int nblk)
{
    const GBlk g = gblk(1);
    const DtArgs<T> a = member_args(one, tab, g.m);
    const T *U = a.U, *V = a.V, *W = a.W;
    const Metric<double>& m = a.m;
    const int n1 = d.nxg - 2, n2 = d.nyg - 2;
    const int n = n1 * n2 * (d.nzg - 2);
    const int per = (n + nblk - 1) / nblk;
    const int q0 = g.x * per, q1 = min(q0 + per, n);
    double lu = 0.0, lv = 0.0, lw = 0.0;
    constexpr int E = 4;
    for (int qb = q0 + (int)threadIdx.x; qb < q1; qb += E * RED_THREADS) {
        T u[E], v[E], w[E];
        int ii[E], jj[E], kk[E];
#pragma unroll
        for (int e = 0; e < E; e++) {
            const int q = min(qb + e * RED_THREADS, q1 - 1);
            const int r = q / n1;
            ii[e] = q - r * n1 + 2;
            jj[e] = r % n2 + 2;
            kk[e] = r / n2 + 2;
            const long o = at(d, ii[e], jj[e], kk[e]);
            u[e] = U[o];
            v[e] = V[o];
            w[e] = W[o];
        }
#pragma unroll
        for (int e = 0; e < E; e++) {
            // cells past the run repeat its last cell, which changes no maximum
            const int i = ii[e], j = jj[e], k = kk[e];
            lu = fmax(lu, (m.xg[i + 1] - m.xg[i]) / fabs((double)u[e]));
            lv = fmax(lv, (m.yg[j + 1] - m.yg[j]) / fabs((double)v[e]));
            lw = fmax(lw, (m.zg[k + 1] - m.zg[k]) / fabs((double)w[e]));
        }
    }
    block_max3(lu, lv, lw, a.part, nblk, g.x);
}


template <typename T>
__global__ void k_dt_final(DtArgs<T> one, const DtArgs<T>* tab, int nblk)
{
    const GBlk g = gblk(1);
    const DtArgs<T> ar = member_args(one, tab, g.m);
    const double* part = ar.part;
    Scalars* sc = ar.sc;
    const double* mins3 = ar.mins3;
    __shared__ double red[3][RED_THREADS];
    const int t = threadIdx.x;
    double a = 0.0, b = 0.0, c = 0.0;
    for (int q = t; q < nblk; q += RED_THREADS) {
        a = fmax(a, part[q]); b = fmax(b, part[nblk + q]); c = fmax(c, part[2 * nblk + q]);
    }
    red[0][t] = a; red[1][t] = b; red[2][t] = c;
    __syncthreads();
    for (int st = RED_THREADS / 2; st > 0; st >>= 1) {
        if (t < st) {
            red[0][t] = fmax(red[0][t], red[0][t + st]);
            red[1][t] = fmax(red[1][t], red[1][t + st]);
            red[2][t] = fmax(red[2][t], red[2][t + st]);
        }
        __syncthreads();
    }
    if (t == 0) sc->dt = dt_from_maxima(red[0][0], red[1][0], red[2][0], sc, mins3);
}

template <typename T>
void run_dt(const DtArgs<T>& one, const DtArgs<T>* tab, int nmem, const Dims& d, int nblk, hipStream_t s)
{
    PROF_BEGIN("k_dt_partial", s);
    k_dt_partial<T><<<dim3(nblk, 1, nmem), RED_THREADS, 0, s>>>(one, tab, d, nblk);
    PROF_END(s);
    PROF_BEGIN("k_dt_final", s);
    k_dt_final<T><<<dim3(1, 1, nmem), RED_THREADS, 0, s>>>(one, tab, nblk);
    PROF_END(s);
}

template <typename T>
void launch_dt(const T* U, const T* V, const T* W, const Metric<double>& md, const Dims& d, double* part, int nblk,
               Scalars* sc, const double* mins3, hipStream_t s)
{
    run_dt<T>(DtArgs<T>{U, V, W, md, part, sc, mins3}, nullptr, 1, d, nblk, s);
}

// code/projection.f90:223-245 (signed max) plus max |div|
template <typename T>
__global__ void k_div_partial(const T* U, const T* V, const T* W, Metric<double> m, Dims d, double* part, int nblk)
{
    const int n1 = d.nxg - 2, n2 = d.nyg - 2, n3 = d.nzg - 2, tot = n1 * n2 * n3;
    double mx = 0.0, ma = 0.0;
    for (int t = blockIdx.x * RED_THREADS + threadIdx.x; t < tot; t += nblk * RED_THREADS) {
        const int r = t / n1;
        const int i = t - r * n1 + 2;
        const int j = r % n2 + 2;
        const int k = r / n2 + 2;
        double v = ((double)U[at(d, i, j, k)] - (double)U[at(d, i - 1, j, k)]) / (m.x[i] - m.x[i - 1]) +
                   ((double)V[at(d, i, j, k)] - (double)V[at(d, i, j - 1, k)]) / (m.y[j] - m.y[j - 1]) +
                   ((double)W[at(d, i, j, k)] - (double)W[at(d, i, j, k - 1)]) / (m.z[k] - m.z[k - 1]);
        mx = fmax(mx, v);
        ma = fmax(ma, fabs(v));
    }
    block_max3(mx, ma, 0.0, part, nblk, blockIdx.x);
}

__global__ void k_div_final(const double* part, int nblk, Scalars* sc)
{
    __shared__ double red[2][RED_THREADS];
    const int t = threadIdx.x;
    double a = 0.0, b = 0.0;
    for (int q = t; q < nblk; q += RED_THREADS) { a = fmax(a, part[q]); b = fmax(b, part[nblk + q]); }
    red[0][t] = a; red[1][t] = b;
    __syncthreads();
    for (int st = RED_THREADS / 2; st > 0; st >>= 1) {
        if (t < st) { red[0][t] = fmax(red[0][t], red[0][t + st]); red[1][t] = fmax(red[1][t], red[1][t + st]); }
        __syncthreads();
    }
    if (t == 0) { sc->maxdiv = red[0][0]; sc->maxabsdiv = red[1][0]; }
}

template <typename T>
void launch_divergence_check(const T* U, const T* V, const T* W, const Metric<double>& md, const Dims& d,
                             double* part, int nblk, Scalars* sc, hipStream_t s)
{
    PROF_BEGIN("k_div_partial", s);
    k_div_partial<T><<<nblk, RED_THREADS, 0, s>>>(U, V, W, md, d, part, nblk);
    PROF_END(s);
    PROF_BEGIN("k_div_final", s);
    k_div_final<<<1, RED_THREADS, 0, s>>>(part, nblk, sc);
    PROF_END(s);
}

// code/statistics.f90:52-64 as plane means over the unique periodic points:
// out[q*ny + j-1] for q = U, V, W, U2, V2, W2, UV.  One block per row.
template <typename T>
__global__ void k_plane_stats(const T* U, const T* V, const T* W, Metric<double> m, Dims d, double* out)
{
    __shared__ double red[7][RED_THREADS];
    const int j = blockIdx.x + 1;
    const int t = threadIdx.x;
    double s[7] = {0, 0, 0, 0, 0, 0, 0};
    const long n1 = d.nxp, n3 = d.nzp, tot = n1 * n3;
    for (long q = t; q < tot; q += RED_THREADS) {
        const int i = (int)(q % n1) + 2, k = (int)(q / n1) + 2;
        const double u = U[at(d, i, j, k)], v = V[at(d, i, j, k)], w = W[at(d, i, j, k)];
        s[0] += u; s[3] += u * u;
        s[1] += v; s[4] += v * v;
        s[2] += w; s[5] += w * w;
        const double uy0 = m.w0[j] * (double)U[at(d, i - 1, j, k)] + m.w1[j] * (double)U[at(d, i - 1, j + 1, k)];
        const double uy1 = m.w0[j] * (double)U[at(d, i, j, k)] + m.w1[j] * (double)U[at(d, i, j + 1, k)];
        s[6] += v * (0.5 * (uy0 + uy1));
    }
    for (int c = 0; c < 7; c++) red[c][t] = s[c];
    __syncthreads();
    for (int st = RED_THREADS / 2; st > 0; st >>= 1) {
        if (t < st)
            for (int c = 0; c < 7; c++) red[c][t] += red[c][t + st];
        __syncthreads();
    }
    if (t == 0)
        for (int c = 0; c < 7; c++) out[(long)c * d.ny + j - 1] = red[c][0] / ((double)d.nxp * (double)d.nzp);
}

template <typename T>
void launch_plane_stats(const T* U, const T* V, const T* W, const Metric<double>& md, const Dims& d, double* out,
                        hipStream_t s)
{
    PROF_BEGIN("k_plane_stats", s);
    k_plane_stats<T><<<d.ny, RED_THREADS, 0, s>>>(U, V, W, md, d, out);
    PROF_END(s);
}

// code/statistics.f90:95-124: mean wall shear stress (nu + nu_t) dU/dy at
// each wall over the unique x faces and z centres, nu_t taken to the wall
// face by the x average and the weighted y interpolation of
// code/interpolation.f90.  The reference evaluates the bottom wall for both
// sums; here the top wall uses its own rows (U(nyg-1) - U(nyg)) / (yg(nyg) -
// yg(nyg-1)) and its own wall face (y face ny).  out: tau_bottom, tau_top.
template <typename T>
__global__ void k_wall_shear_partial(const T* U, const T* nut, Metric<double> m, Dims d, double nu, double* part,
                                     int nblk)
{
    const long n1 = d.nx - 2, n3 = d.nzp, tot = n1 * n3;
    const int jt = d.nyg - 1, ny = d.ny;
    double sb = 0.0, st = 0.0;
    for (long t = (long)blockIdx.x * RED_THREADS + threadIdx.x; t < tot; t += (long)nblk * RED_THREADS) {
        const int i = (int)(t % n1) + 2, k = (int)(t / n1) + 2;
        auto tx = [&](int j) { return 0.5 * ((double)nut[at(d, i, j, k)] + (double)nut[at(d, i + 1, j, k)]); };
        const double nb = m.w0[1] * tx(1) + m.w1[1] * tx(2);
        const double nt = m.w0[ny] * tx(jt) + m.w1[ny] * tx(jt + 1);
        sb += ((double)U[at(d, i, 2, k)] - (double)U[at(d, i, 1, k)]) / (m.yg[2] - m.yg[1]) * (nu + nb);
        st += ((double)U[at(d, i, jt, k)] - (double)U[at(d, i, jt + 1, k)]) / (m.yg[jt + 1] - m.yg[jt]) * (nu + nt);
    }
    block_sum2(sb, st, part, nblk, blockIdx.x);
}

__global__ void k_wall_shear_final(const double* part, int nblk, double ncell, double* out)
{
    double a, b;
    final_sum2(part, nblk, a, b);
    if (threadIdx.x == 0) { out[0] = a / ncell; out[1] = b / ncell; }
}

template <typename T>
void launch_wall_shear(const T* U, const T* nut, const Metric<double>& md, const Dims& d, double nu, double* part,
                       int nblk, double* out, hipStream_t s)
{
    PROF_BEGIN("k_wall_shear", s);
    k_wall_shear_partial<T><<<nblk, RED_THREADS, 0, s>>>(U, nut, md, d, nu, part, nblk);
    k_wall_shear_final<<<1, RED_THREADS, 0, s>>>(part, nblk, (double)(d.nx - 2) * (double)d.nzp, out);
    PROF_END(s);
}

// ------------------------------------------------------- instantiations
#define CHAN_INST(T)                                                                                            \
    template void launch_bc<T>(T*, const Dims&, int, hipStream_t);                                              \
    template void launch_dsm<T>(const T*, const T*, const T*, T*, const Metric<T>&, const Dims&, DsmBufs<T>&,   \
                                int, int, const double*, T, const Scalars*, hipStream_t, const EqAlphaArgs<T>*); \
    template void launch_wall_model<T>(const T*, T*, const double*, const Metric<T>&, const Dims&, T,           \
                                       const Scalars*, hipStream_t);                                            \
    template void launch_equilibrium_alpha<T>(const T*, double*, const Dims&, double, const WallRows&,        \
                                              double, double, hipStream_t);                                     \
    template void launch_mass_flow_init<T>(const T*, const Metric<double>&, const Dims&, double*, int,          \
                                           Scalars*, hipStream_t);                                              \
    template void launch_mass_flow_correct<T>(T*, const Metric<double>&, const Dims&, double*, int, Scalars*,   \
                                              int, int, hipStream_t, int);                                      \
    template void launch_mass_flow_of_projection<T>(const T*, const double*, const Metric<double>&, const Dims&, \
                                                    double*, int, Scalars*, hipStream_t);                        \
    template void launch_dt<T>(const T*, const T*, const T*, const Metric<double>&, const Dims&, double*, int,  \
                               Scalars*, const double*, hipStream_t);                                           \
    template void launch_divergence_check<T>(const T*, const T*, const T*, const Metric<double>&, const Dims&, \
                                             double*, int, Scalars*, hipStream_t);                              \
    template void launch_plane_stats<T>(const T*, const T*, const T*, const Metric<double>&, const Dims&,       \
                                        double*, hipStream_t);                                                  \
    template void launch_wall_shear<T>(const T*, const T*, const Metric<double>&, const Dims&, double, double*, \
                                       int, double*, hipStream_t);                                              \
    template void run_dsm_nut<T>(const DsmNutArgs<T>&, const DsmNutArgs<T>*, int, const Dims&, int,             \
                                 hipStream_t);                                                                  \
    template void run_dsm_edges<T>(const DsmNutArgs<T>&, const DsmNutArgs<T>*, int, const Dims&, hipStream_t);  \
    template void run_equilibrium_alpha<T>(const EqAlphaArgs<T>&, const EqAlphaArgs<T>*, int, const Dims&,      \
                                           hipStream_t);                                                        \
    template void run_mass_flow_correct<T>(const MassArgs<T>&, const MassArgs<T>*, int, const Dims&, int, int,  \
                                           int, int, hipStream_t);                                              \
    template void run_dt<T>(const DtArgs<T>&, const DtArgs<T>*, int, const Dims&, int, hipStream_t);

CHAN_INST(double)
CHAN_INST(float)

}  // namespace chan
