// amd.hip -- the anisotropic minimum dissipation (AMD) eddy viscosity
// (code/subgrid.f90:632-874), LES_model 3:
//   nu_t = max(-C N / D, 0), C = 0.3,
//   N = sum_ijk (dh_k u_i)(dh_k u_j) S_ij   (dh: undivided difference)
//   D = sum_lm (d_l u_m)^2
// A block owns a TX x TY tile of (x,y) centres and marches in z; U, V, W of
// planes k-1, k, k+1 over the tile plus a one-cell halo sit in a four-slot
// LDS ring (the next plane is fetched into registers while one is
// computed), so each value is read from HBM once per tile and the ~80
// stencil operands of a centre come from LDS.  Each unique centre writes
// its periodic images and, on the wall rows, the ghost rows per
// Dirichlet_nu_t, so no boundary pass follows.  The staggered differences
// reach the centre through the reference's face-to-centre averages in the
// order of its interpolate calls (code/interpolation.f90:24-95), and the
// terms are summed in the order of its statements; divisions by spacings
// are multiplications by their reciprocals (an ulp per factor).
#include <algorithm>
#include "kernels.hpp"
#include "prof.hpp"
#include "wallmodel.hpp"

namespace chan {

// reciprocal spacings at the centre (i, j) and per plane k:
// rdx = 1/(x(i)-x(i-1)), rdxg0/1 = 1/(xg(i)-xg(i-1)), 1/(xg(i+1)-xg(i)), ...
template <typename T>
struct AmdXY { T rdx, rdxg0, rdxg1, rdy, rdyg0, rdyg1; };
template <typename T>
struct AmdZ { T rdz, rdzg0, rdzg1; };

// the model at centre (i, j, k) from F(f, a, b, c) (0 U, 1 V, 2 W)
template <typename T, typename A>
__device__ __forceinline__ T amd_gen(const A& F, const AmdXY<T>& g, const AmdZ<T>& gz, int i, int j, int k)
{
    const T h = T(0.5), qq = T(0.25);
    auto u = [&](int a, int b, int c) { return F(0, a, b, c); };
    auto v = [&](int a, int b, int c) { return F(1, a, b, c); };
    auto w = [&](int a, int b, int c) { return F(2, a, b, c); };
    auto ax = [&](T a, T b) { return h * (a + b); };
    // reciprocal centre spacings by offset: xg(i)-xg(i-1) at a = i, xg(i+1)-xg(i) at a = i+1
    auto rxg = [&](int a) { return a == i ? g.rdxg0 : g.rdxg1; };
    auto ryg = [&](int b) { return b == j ? g.rdyg0 : g.rdyg1; };
    auto rzg = [&](int c) { return c == k ? gz.rdzg0 : gz.rdzg1; };
    // rate of strain at the centre (code/subgrid.f90:1027-1059)
    const T S11 = (u(i, j, k) - u(i - 1, j, k)) * g.rdx;
    const T S22 = (v(i, j, k) - v(i, j - 1, k)) * g.rdy;
    const T S33 = (w(i, j, k) - w(i, j, k - 1)) * gz.rdz;
    auto e4 = [&](int a, int b) {
        return h * ((u(a - 1, b, k) - u(a - 1, b - 1, k)) * ryg(b) + (v(a, b - 1, k) - v(a - 1, b - 1, k)) * rxg(a));
    };
    auto e5 = [&](int a, int c) {
        return h * ((u(a - 1, j, c) - u(a - 1, j, c - 1)) * rzg(c) + (w(a, j, c - 1) - w(a - 1, j, c - 1)) * rxg(a));
    };
    auto e6 = [&](int b, int c) {
        return h * ((v(i, b - 1, c) - v(i, b - 1, c - 1)) * rzg(c) + (w(i, b, c - 1) - w(i, b - 1, c - 1)) * ryg(b));
    };
    const T S12 = qq * (e4(i, j) + e4(i + 1, j) + e4(i, j + 1) + e4(i + 1, j + 1));
    const T S13 = qq * (e5(i, k) + e5(i + 1, k) + e5(i, k + 1) + e5(i + 1, k + 1));
    const T S23 = qq * (e6(j, k) + e6(j + 1, k) + e6(j, k + 1) + e6(j + 1, k + 1));
    // undivided differences at their natural positions
    auto dUx = [&](int a, int b, int c) { return u(a, b, c) - u(a - 1, b, c); };
    auto dUy = [&](int a, int b, int c) { return u(a, b, c) - u(a, b - 1, c); };
    auto dUz = [&](int a, int b, int c) { return u(a, b, c) - u(a, b, c - 1); };
    auto dVx = [&](int a, int b, int c) { return v(a, b, c) - v(a - 1, b, c); };
    auto dVy = [&](int a, int b, int c) { return v(a, b, c) - v(a, b - 1, c); };
    auto dVz = [&](int a, int b, int c) { return v(a, b, c) - v(a, b, c - 1); };
    auto dWx = [&](int a, int b, int c) { return w(a, b, c) - w(a - 1, b, c); };
    auto dWy = [&](int a, int b, int c) { return w(a, b, c) - w(a, b - 1, c); };
    auto dWz = [&](int a, int b, int c) { return w(a, b, c) - w(a, b, c - 1); };
    const T Uxh = dUx(i, j, k);
    const T Uyh = ax(ax(dUy(i - 1, j, k), dUy(i, j, k)), ax(dUy(i - 1, j + 1, k), dUy(i, j + 1, k)));
    const T Uzh = ax(ax(dUz(i - 1, j, k), dUz(i, j, k)), ax(dUz(i - 1, j, k + 1), dUz(i, j, k + 1)));
    const T Vyh = dVy(i, j, k);
    const T Vxh = ax(ax(dVx(i, j - 1, k), dVx(i, j, k)), ax(dVx(i + 1, j - 1, k), dVx(i + 1, j, k)));
    const T Vzh = ax(ax(dVz(i, j - 1, k), dVz(i, j, k)), ax(dVz(i, j - 1, k + 1), dVz(i, j, k + 1)));
    const T Wzh = dWz(i, j, k);
    const T Wxh = ax(ax(dWx(i, j, k - 1), dWx(i, j, k)), ax(dWx(i + 1, j, k - 1), dWx(i + 1, j, k)));
    const T Wyh = ax(ax(dWy(i, j, k - 1), dWy(i, j, k)), ax(dWy(i, j + 1, k - 1), dWy(i, j + 1, k)));
    const T ry0 = g.rdyg0, ry1 = g.rdyg1, rz0 = gz.rdzg0, rz1 = gz.rdzg1, rx0 = g.rdxg0, rx1 = g.rdxg1;
    const T Uyd = ax(ax(dUy(i - 1, j, k) * ry0, dUy(i, j, k) * ry0), ax(dUy(i - 1, j + 1, k) * ry1, dUy(i, j + 1, k) * ry1));
    const T Uzd = ax(ax(dUz(i - 1, j, k) * rz0, dUz(i, j, k) * rz0), ax(dUz(i - 1, j, k + 1) * rz1, dUz(i, j, k + 1) * rz1));
    const T Vxd = ax(ax(dVx(i, j - 1, k) * rx0, dVx(i, j, k) * rx0), ax(dVx(i + 1, j - 1, k) * rx1, dVx(i + 1, j, k) * rx1));
    const T Vzd = ax(ax(dVz(i, j - 1, k) * rz0, dVz(i, j, k) * rz0), ax(dVz(i, j - 1, k + 1) * rz1, dVz(i, j, k + 1) * rz1));
    const T Wxd = ax(ax(dWx(i, j, k - 1) * rx0, dWx(i, j, k) * rx0), ax(dWx(i + 1, j, k - 1) * rx1, dWx(i + 1, j, k) * rx1));
    const T Wyd = ax(ax(dWy(i, j, k - 1) * ry0, dWy(i, j, k) * ry0), ax(dWy(i, j + 1, k - 1) * ry1, dWy(i, j + 1, k) * ry1));
    // cross terms formed before the last average (code/subgrid.f90:770-779,803-812,836-845)
    const T z12 = ax(ax(dUz(i - 1, j, k), dUz(i, j, k)) * ax(dVz(i, j - 1, k), dVz(i, j, k)),
                     ax(dUz(i - 1, j, k + 1), dUz(i, j, k + 1)) * ax(dVz(i, j - 1, k + 1), dVz(i, j, k + 1)));
    const T y13 = ax(ax(dUy(i - 1, j, k), dUy(i, j, k)) * ax(dWy(i, j, k - 1), dWy(i, j, k)),
                     ax(dUy(i - 1, j + 1, k), dUy(i, j + 1, k)) * ax(dWy(i, j + 1, k - 1), dWy(i, j + 1, k)));
    const T x23 = ax(ax(dVx(i, j - 1, k), dVx(i, j, k)) * ax(dWx(i, j, k - 1), dWx(i, j, k)),
                     ax(dVx(i + 1, j - 1, k), dVx(i + 1, j, k)) * ax(dWx(i + 1, j, k - 1), dWx(i + 1, j, k)));
    const T two = T(2);
    T N = Uxh * Uxh * S11;
    N = N + Uyh * Uyh * S11;
    N = N + Uzh * Uzh * S11;
    N = N + Vyh * Vyh * S22;
    N = N + Vxh * Vxh * S22;
    N = N + Vzh * Vzh * S22;
    N = N + (Wzh * Wzh) * S33;
    N = N + Wxh * Wxh * S33;
    N = N + Wyh * Wyh * S33;
    N = N + two * (Vxh * Uxh) * S12;
    N = N + two * (Uyh * Vyh) * S12;
    N = N + two * z12 * S12;
    N = N + two * (Wxh * Uxh) * S13;
    N = N + two * (Uzh * Wzh) * S13;
    N = N + two * y13 * S13;
    N = N + two * (Wyh * Vyh) * S23;
    N = N + two * (Vzh * Wzh) * S23;
    N = N + two * x23 * S23;
    T D = Uxh * Uxh * (g.rdx * g.rdx);
    D = D + Uyd * Uyd;
    D = D + Uzd * Uzd;
    D = D + Vyh * Vyh * (g.rdy * g.rdy);
    D = D + Vxd * Vxd;
    D = D + Vzd * Vzd;
    D = D + (Wzh * Wzh) * (gz.rdz * gz.rdz);
    D = D + Wxd * Wxd;
    D = D + Wyd * Wyd;
    // clipping; 0/0 (uniform neighbourhood) is NaN and clips to 0
    return fmax(-T(0.3) * N / D, T(0));
}

// The unique centres i = 2..nxg-2, j = 2..nyg-1, k = 2..nzg-2 from the LDS
// ring; each writes its periodic images (x: 1 <- nxg-2, nxg-1 <- 2,
// nxg <- 3; z alike) and rows 2 and nyg-1 the wall ghost rows 1 and nyg per
// Dirichlet_nu_t (1 mirror, 0 copy, 2 mirror plus the offset
// 0.38 sqrt(dPdx) (y(2)-y(1)), code/subgrid.f90:864-871).  With fuse_wm the
// four wall rows are left to k_amd_wall_rows.
// With DT (dt.part set), the launch also takes compute_dt's maxima over its
// cells (code/time_integration.f90:113-121): per column the smallest |U|
// and |V| (a correctly rounded quotient falls as its divisor grows, so the
// largest dx/|U| is dx over the smallest |U|, exactly) and per plane the
// largest dz/|W|, with the terms of the periodic image column nxg-1 and
// plane nzg-1 from the canonical values they copy; one slot per block.
static constexpr int AMD_KMAX = 32;
struct AmdDt {
    double* part;  // [3][blocks of the launch]
    const double *xg, *yg, *zg;
};
template <typename T, int TX, int TY, bool DT>
__global__ void __launch_bounds__(TX* TY) k_amd(const T* __restrict__ U, const T* __restrict__ V,
                                                const T* __restrict__ W, T* __restrict__ nut, Metric<T> m, Dims d,
                                                int dirichlet, const Scalars* sc, int fuse_wm, int kchunk, AmdDt dt)
{
    constexpr int PX = TX + 2, PY = TY + 2, NP = PX * PY, NT = TX * TY, NL = (NP + NT - 1) / NT;
    __shared__ T sh[3][4][NP];  // [field][slot][position]
    __shared__ T zt[3][AMD_KMAX];  // rdz(k), rdzg(k), rdzg(k+1) at k = kb + q
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * TX + tx;
    const GBlk gb = gblk_xcd((d.nzg - 3 + kchunk - 1) / kchunk);
    const int i0 = gb.x * TX + 2, j0 = gb.y * TY + 2;
    const int kb = gb.z * kchunk + 2;
    const int ke = min(kb + kchunk - 1, d.nzg - 2);
    const int i = i0 + tx, j = j0 + ty;
    const T* src[3] = {U, V, W};
    // this thread's share of a halo plane: LDS index and box offset (-1: outside the box)
    int lq[NL];
    long lo[NL];
#pragma unroll
    for (int r = 0; r < NL; r++) {
        const int q = tid + r * NT;
        const int ii = i0 - 1 + q % PX, jj = j0 - 1 + q / PX;
        lq[r] = q < NP ? q : -1;
        lo[r] = (q < NP && ii <= d.nxg && jj <= d.nyg) ? at(d, ii, jj, 1) : -1;
    }
    T pre[NL][3];
    auto gload = [&](int p) {
        const long po = (long)(p - 1) * d.sxy;
#pragma unroll
        for (int r = 0; r < NL; r++)
#pragma unroll
            for (int f = 0; f < 3; f++) pre[r][f] = lo[r] >= 0 ? src[f][lo[r] + po] : T(0);
    };
    auto slot = [&](int p) { return (p - kb + 4) & 3; };
    auto sstore = [&](int p) {
        const int s = slot(p);
#pragma unroll
        for (int r = 0; r < NL; r++)
            if (lq[r] >= 0)
#pragma unroll
                for (int f = 0; f < 3; f++) sh[f][s][lq[r]] = pre[r][f];
    };
    auto F = [&](int f, int a, int b, int c) -> T { return sh[f][slot(c)][(a - i0 + 1) + PX * (b - j0 + 1)]; };
    AmdXY<T> g;
    {
        const int ic = min(i, d.nxg - 2), jc = min(j, d.nyg - 1);
        g.rdx = m.rdx[ic]; g.rdxg0 = m.rdxg[ic]; g.rdxg1 = m.rdxg[ic + 1];
        g.rdy = m.rdy[jc]; g.rdyg0 = m.rdyg[jc]; g.rdyg1 = m.rdyg[jc + 1];
    }
    for (int q = tid; q < AMD_KMAX; q += NT) {
        const int kk = min(kb + q, d.nzg - 1);
        zt[0][q] = m.rdz[kk];
        zt[1][q] = m.rdzg[kk];
        zt[2][q] = m.rdzg[kk + 1];
    }
    const bool own = i <= d.nxg - 2 && j <= d.nyg - 1;  // a unique centre
    const bool act = own && !(fuse_wm && (j == 2 || j == d.nyg - 1));
    T dt_mu = T(INFINITY), dt_mv = T(INFINITY);
    double dt_lw = 0.0;
    const bool ghost = !fuse_wm && (j == 2 || j == d.nyg - 1);
    const T off = dirichlet == 2 ? T(2.0 * 0.38 * sqrt(sc->dPdx) * (double)(m.y[2] - m.y[1]) / 2.0) : T(0);
    int xs[3], nxs = 1;
    xs[0] = i;
    if (i == d.nxg - 2) xs[nxs++] = 1;
    if (i == 2) xs[nxs++] = d.nxg - 1;
    if (i == 3) xs[nxs++] = d.nxg;
    gload(kb - 1);
    sstore(kb - 1);
    gload(kb);
    sstore(kb);
    gload(kb + 1);
    __syncthreads();
    for (int k = kb; k <= ke; k++) {
        // plane k+1 into the slot of plane k-3 (free), then the next fetch
        sstore(k + 1);
        if (k + 2 <= ke + 1) gload(k + 2);
        __syncthreads();
        if (DT && own) {
            dt_mu = fmin(dt_mu, fabs(F(0, i, j, k)));
            dt_mv = fmin(dt_mv, fabs(F(1, i, j, k)));
            const double aw = fabs((double)F(2, i, j, k));
            dt_lw = fmax(dt_lw, (dt.zg[k + 1] - dt.zg[k]) / aw);
            if (k == 2) dt_lw = fmax(dt_lw, (dt.zg[d.nzg] - dt.zg[d.nzg - 1]) / aw);
        }
        if (act) {
            const AmdZ<T> gz{zt[0][k - kb], zt[1][k - kb], zt[2][k - kb]};
            const T val = amd_gen<T>(F, g, gz, i, j, k);
            int zs[3], nzs = 1;
            zs[0] = k;
            if (k == d.nzg - 2) zs[nzs++] = 1;
            if (k == 2) zs[nzs++] = d.nzg - 1;
            if (k == 3) zs[nzs++] = d.nzg;
            const int jg = j == 2 ? 1 : d.nyg;
            const T gv = dirichlet == 1 ? -val : (dirichlet == 2 ? -val + off : val);
            for (int a = 0; a < nxs; a++)
                for (int c = 0; c < nzs; c++) {
                    nut[at(d, xs[a], j, zs[c])] = val;
                    if (ghost) nut[at(d, xs[a], jg, zs[c])] = gv;
                }
        }
    }
    if constexpr (DT) {
        // the block's maxima: per lane, per wave (butterfly), over the waves
        double m3[3] = {0.0, 0.0, dt_lw};
        if (own) {
            m3[0] = (dt.xg[i + 1] - dt.xg[i]) / (double)dt_mu;
            if (i == 2) m3[0] = fmax(m3[0], (dt.xg[d.nxg] - dt.xg[d.nxg - 1]) / (double)dt_mu);
            m3[1] = (dt.yg[j + 1] - dt.yg[j]) / (double)dt_mv;
        }
        for (int s = 32; s > 0; s >>= 1)
            for (int c = 0; c < 3; c++) m3[c] = fmax(m3[c], __shfl_xor(m3[c], s, 64));
        __shared__ double wmax[NT / 64][3];
        if ((tid & 63) == 0)
            for (int c = 0; c < 3; c++) wmax[tid >> 6][c] = m3[c];
        __syncthreads();
        if (tid < 3) {
            double v = 0.0;
            for (int w = 0; w < NT / 64; w++) v = fmax(v, wmax[w][tid]);
            const int nb = gb.gx * gb.gy * gb.gz;
            dt.part[(long)tid * nb + (gb.z * gb.gy + gb.y) * gb.gx + gb.x] = v;
        }
    }
}

// The four wall rows from the wall model, every box cell (i, k) of them
// (code/wallmodel.f90:23-46 at the x faces around the canonical centre):
// one block per (z centre k, wall), whose threads first compute the
// equilibrium wall stress of the row's faces when eq.alpha is set (the
// values k_equilibrium_alpha computes: each face by one thread, then a
// barrier), then nu_t of rows 1, 2 or nyg-1, nyg.  With dtpart set, one
// more block reduces k_amd's compute_dt maxima and sets sc->dt as
// k_dt_final does (the launcher sets it only when the wall rows do not read
// dt: no wall viscosity limit).
struct AmdWallDt {
    const double* part;
    int nblk;
    const double* mins3;
};
template <typename T>
__global__ void __launch_bounds__(256) k_amd_wall(const T* U, T* nut, double* alpha, Metric<T> m, Dims d, T nu,
                                                  Scalars* sc, EqAlphaArgs<T> eq, AmdWallDt dt)
{
    const int t = threadIdx.x;
    if ((int)blockIdx.x == 2 * d.nzg) {
        double m3[3] = {0.0, 0.0, 0.0};
        for (int q = t; q < dt.nblk; q += blockDim.x)
            for (int c = 0; c < 3; c++) m3[c] = fmax(m3[c], dt.part[(long)c * dt.nblk + q]);
        for (int s = 32; s > 0; s >>= 1)
            for (int c = 0; c < 3; c++) m3[c] = fmax(m3[c], __shfl_xor(m3[c], s, 64));
        __shared__ double wmax[4][3];
        if ((t & 63) == 0)
            for (int c = 0; c < 3; c++) wmax[t >> 6][c] = m3[c];
        __syncthreads();
        if (t == 0) {
            for (int w = 1; w < (int)blockDim.x / 64; w++)
                for (int c = 0; c < 3; c++) m3[c] = fmax(m3[c], wmax[w][c]);
            sc->dt = dt_from_maxima(m3[0], m3[1], m3[2], sc, dt.mins3);
        }
        return;
    }
    const int k = (int)blockIdx.x % d.nzg + 1, wall = (int)blockIdx.x / d.nzg;
    if (eq.alpha) {
        for (int f = t; f < d.nx; f += blockDim.x)
            eq_alpha_cell<T>(f + d.nx * ((k - 1) + d.nzg * wall), U, eq.alpha, d, eq.nu, eq.kappa, eq.B, eq.yc, eq.ws);
        __syncthreads();
    }
    const int j0 = wall == 0 ? 1 : d.nyg - 1;
    for (int i = t + 1; i <= d.nxg; i += blockDim.x) {
        const T v = wm_nut_at<T>(U, alpha, m, d, nu, img_centre(i, d.nxg), k, wall, sc);
        nut[at(d, i, j0, k)] = v;
        nut[at(d, i, j0 + 1, k)] = v;
    }
}

// the AMD kernel's tile (32x8 centres) and z chunk: at most AMD_KMAX
// planes, enough blocks for two rounds on 256 CUs
static constexpr int AMD_TX = 32, AMD_TY = 8;
static dim3 amd_grid(const Dims& d, int& kc)
{
    const int nk = d.nzg - 3;
    const long tiles = (long)cdiv(d.nxg - 3, AMD_TX) * cdiv(d.nyg - 2, AMD_TY);
    kc = (int)std::min<long>(AMD_KMAX, std::max<long>(4, (tiles * nk + 2047) / 2048));
    kc = std::min(kc, nk);
    return dim3(cdiv(d.nxg - 3, AMD_TX), cdiv(d.nyg - 2, AMD_TY), (unsigned)cdiv(nk, kc));
}

template <typename T>
void launch_amd(const T* U, const T* V, const T* W, T* nut, const Metric<T>& m, const Dims& d, int dirichlet,
                const Scalars* sc, int fuse_wm, const double* alpha, T nu, hipStream_t s, const EqAlphaArgs<T>* eq,
                double* dtpart, const Metric<double>* md, const double* mins3)
{
    constexpr int TX = AMD_TX, TY = AMD_TY;
    int kc;
    const dim3 blk(TX, TY, 1), grd = amd_grid(d, kc);
    // compute_dt rides along only with the wall rows' launch after it
    const bool with_dt = dtpart && fuse_wm;
    const AmdDt adt{with_dt ? dtpart : nullptr, md ? md->xg : nullptr, md ? md->yg : nullptr, md ? md->zg : nullptr};
    PROF_BEGIN("k_amd", s);
    if (with_dt) k_amd<T, TX, TY, true><<<grd, blk, 0, s>>>(U, V, W, nut, m, d, dirichlet, sc, fuse_wm, kc, adt);
    else k_amd<T, TX, TY, false><<<grd, blk, 0, s>>>(U, V, W, nut, m, d, dirichlet, sc, fuse_wm, kc, adt);
    PROF_END(s);
    if (fuse_wm) {
        EqAlphaArgs<T> ea{};
        if (eq) ea = *eq;
        const AmdWallDt wdt{dtpart, (int)(grd.x * grd.y * grd.z), mins3};
        PROF_BEGIN("k_amd_wall", s);
        k_amd_wall<T><<<2 * d.nzg + (with_dt ? 1 : 0), 256, 0, s>>>(U, nut, const_cast<double*>(alpha), m, d, nu,
                                                                    const_cast<Scalars*>(sc), ea, wdt);
        PROF_END(s);
    }
}

// blocks of k_amd (the compute_dt partials' row length)
int amd_blocks(const Dims& d)
{
    int kc;
    const dim3 g = amd_grid(d, kc);
    return (int)(g.x * g.y * g.z);
}

template void launch_amd<double>(const double*, const double*, const double*, double*, const Metric<double>&,
                                 const Dims&, int, const Scalars*, int, const double*, double, hipStream_t,
                                 const EqAlphaArgs<double>*, double*, const Metric<double>*, const double*);
template void launch_amd<float>(const float*, const float*, const float*, float*, const Metric<float>&, const Dims&,
                                int, const Scalars*, int, const double*, float, hipStream_t,
                                const EqAlphaArgs<float>*, double*, const Metric<double>*, const double*);

}  // namespace chan
