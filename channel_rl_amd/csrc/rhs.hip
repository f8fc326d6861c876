// rhs.hip -- right-hand sides of u, v, w (code/equations.f90:24-468) and the
// Wray RK3 combine (code/time_integration.f90:42-44,59-61,76-81).
//
// A block owns a TX x TY tile of (x,y) cells and marches in z.  U, V, W and
// nu_t for planes k-1, k, k+1 of the tile plus a one-cell halo sit in LDS
// (a four-slot ring), so every value is read from HBM once per tile and
// the ~120 stencil operands of a cell come from LDS.  The arithmetic is the
// reference's operation order (no FMA contraction in FP64).
#include <algorithm>
#include "kernels.hpp"
#include "prof.hpp"

namespace chan {

// Metric factors of one cell: reciprocal spacings from the 1-based tables
// r[i] = 1/(a(i)-a(i-1)) of the faces and centres.  The reference divides by
// the spacings (code/equations.f90:41-160); multiplying by the correctly
// rounded reciprocal instead differs by at most an ulp per factor and turns
// ~36 FP64 divisions per cell into multiplications.  x and y factors depend
// on the thread's (i, j) only and are kept in registers over the z march;
// z factors are uniform per plane.
template <typename T>
struct RhsXY {
    T rx_f, rx_fp;         // 1/(x(i)-x(i-1)), 1/(x(i+1)-x(i))
    T rx_c, rx_cp;         // 1/(xg(i)-xg(i-1)), 1/(xg(i+1)-xg(i))
    T ry_f, ry_fp;         // 1/(y(j)-y(j-1)), 1/(y(j+1)-y(j))
    T ry_c, ry_cp;         // 1/(yg(j)-yg(j-1)), 1/(yg(j+1)-yg(j))
    T w0j, w1j, w0m, w1m;  // weight_y_0/1 at j and j-1
};
template <typename T>
struct RhsZ {
    T rz_f, rz_fp;         // 1/(z(k)-z(k-1)), 1/(z(k+1)-z(k))
    T rz_c, rz_cp;         // 1/(zg(k)-zg(k-1)), 1/(zg(k+1)-zg(k))
};

// A(f, i, j, k) returns field f (0 U, 1 V, 2 W, 3 nu_t) at Fortran (i,j,k).
// code/equations.f90:24-167: convection in divergence form, variable-viscosity
// diffusion with dx_1 = x(i)-x(i-1), dx_2 = x(i+1)-x(i), dx_3 = xg(i+1)-xg(i),
// dy_3 = y(j)-y(j-1), dz_1 = zg(k)-zg(k-1), dz_2 = zg(k+1)-zg(k), dz_3 = z(k)-z(k-1)
template <typename T, typename A>
__device__ __forceinline__ T rhs_u_gen(const A& F, const RhsXY<T>& g, const RhsZ<T>& gz, T nu, T dPdx, int i, int j,
                                       int k)
{
    auto u = [&](int a, int b, int c) { return F(0, a, b, c); };
    auto v = [&](int a, int b, int c) { return F(1, a, b, c); };
    auto w = [&](int a, int b, int c) { return F(2, a, b, c); };
    auto n = [&](int a, int b, int c) { return F(3, a, b, c); };
    const T h = T(0.5), q = T(0.25);
    T uc = h * (u(i, j, k) + u(i + 1, j, k));
    T ucm = h * (u(i - 1, j, k) + u(i, j, k));
    T t1 = -(uc * uc - ucm * ucm) * g.rx_cp;
    T uvp = (g.w0j * u(i, j, k) + g.w1j * u(i, j + 1, k)) * (h * (v(i, j, k) + v(i + 1, j, k)));
    T uvm = (g.w0m * u(i, j - 1, k) + g.w1m * u(i, j, k)) * (h * (v(i, j - 1, k) + v(i + 1, j - 1, k)));
    T t2 = -(uvp - uvm) * g.ry_f;
    T uwp = (h * (u(i, j, k) + u(i, j, k + 1))) * (h * (w(i, j, k) + w(i + 1, j, k)));
    T uwm = (h * (u(i, j, k - 1) + u(i, j, k))) * (h * (w(i, j, k - 1) + w(i + 1, j, k - 1)));
    T t3 = -(uwp - uwm) * gz.rz_f;
    T nu_x1 = nu + n(i, j, k);
    T nu_x2 = nu + n(i + 1, j, k);
    T nu_y1 = nu + h * (g.w0m * n(i, j - 1, k) + g.w1m * n(i, j, k) + g.w0m * n(i + 1, j - 1, k) +
                        g.w1m * n(i + 1, j, k));
    T nu_y2 = nu + h * (g.w0j * n(i, j, k) + g.w1j * n(i, j + 1, k) + g.w0j * n(i + 1, j, k) +
                        g.w1j * n(i + 1, j + 1, k));
    T nu_z1 = nu + q * (n(i, j, k) + n(i, j, k - 1) + n(i + 1, j, k) + n(i + 1, j, k - 1));
    T nu_z2 = nu + q * (n(i, j, k) + n(i, j, k + 1) + n(i + 1, j, k) + n(i + 1, j, k + 1));
    T dudy_p = (u(i, j + 1, k) - u(i, j, k)) * g.ry_cp;
    T dudy_m = (u(i, j, k) - u(i, j - 1, k)) * g.ry_c;
    T t4 = g.rx_cp * (nu_x2 * g.rx_fp * (u(i + 1, j, k) - u(i, j, k)) -
                      nu_x1 * g.rx_f * (u(i, j, k) - u(i - 1, j, k))) +
           g.ry_f * (nu_y2 * dudy_p - nu_y1 * dudy_m) +
           gz.rz_f * (nu_z2 * gz.rz_cp * (u(i, j, k + 1) - u(i, j, k)) -
                      nu_z1 * gz.rz_c * (u(i, j, k) - u(i, j, k - 1)));
    return (((t1 + t2) + t3) + t4) + dPdx;
}

// code/equations.f90:174-320, with dx_1 = xg(i)-xg(i-1), dx_2 = xg(i+1)-xg(i),
// dx_3 = x(i)-x(i-1), dy_1 = y(j)-y(j-1), dy_2 = y(j+1)-y(j), dy_3 = yg(j+1)-yg(j),
// dz as for u
template <typename T, typename A>
__device__ __forceinline__ T rhs_v_gen(const A& F, const RhsXY<T>& g, const RhsZ<T>& gz, T nu, int i, int j, int k)
{
    auto u = [&](int a, int b, int c) { return F(0, a, b, c); };
    auto v = [&](int a, int b, int c) { return F(1, a, b, c); };
    auto w = [&](int a, int b, int c) { return F(2, a, b, c); };
    auto n = [&](int a, int b, int c) { return F(3, a, b, c); };
    const T h = T(0.5);
    T vc = h * (v(i, j, k) + v(i, j + 1, k));
    T vcm = h * (v(i, j - 1, k) + v(i, j, k));
    T t1 = -(vc * vc - vcm * vcm) * g.ry_cp;
    T uvp = (g.w0j * u(i, j, k) + g.w1j * u(i, j + 1, k)) * (h * (v(i, j, k) + v(i + 1, j, k)));
    T uvm = (g.w0j * u(i - 1, j, k) + g.w1j * u(i - 1, j + 1, k)) * (h * (v(i - 1, j, k) + v(i, j, k)));
    T t2 = -(uvp - uvm) * g.rx_f;
    T vwp = (h * (v(i, j, k) + v(i, j, k + 1))) * (g.w0j * w(i, j, k) + g.w1j * w(i, j + 1, k));
    T vwm = (h * (v(i, j, k - 1) + v(i, j, k))) * (g.w0j * w(i, j, k - 1) + g.w1j * w(i, j + 1, k - 1));
    T t3 = -(vwp - vwm) * gz.rz_f;
    T nu_x1 = nu + h * (g.w0j * n(i - 1, j, k) + g.w1j * n(i - 1, j + 1, k) + g.w0j * n(i, j, k) +
                        g.w1j * n(i, j + 1, k));
    T nu_x2 = nu + h * (g.w0j * n(i, j, k) + g.w1j * n(i, j + 1, k) + g.w0j * n(i + 1, j, k) +
                        g.w1j * n(i + 1, j + 1, k));
    T nu_y1 = nu + n(i, j, k);
    T nu_y2 = nu + n(i, j + 1, k);
    T nu_z1 = nu + h * (g.w0j * n(i, j, k - 1) + g.w1j * n(i, j + 1, k - 1) + g.w0j * n(i, j, k) +
                        g.w1j * n(i, j + 1, k));
    T nu_z2 = nu + h * (g.w0j * n(i, j, k) + g.w1j * n(i, j + 1, k) + g.w0j * n(i, j, k + 1) +
                        g.w1j * n(i, j + 1, k + 1));
    T t4 = g.rx_f * (nu_x2 * g.rx_cp * (v(i + 1, j, k) - v(i, j, k)) -
                     nu_x1 * g.rx_c * (v(i, j, k) - v(i - 1, j, k))) +
           g.ry_cp * (nu_y2 * g.ry_fp * (v(i, j + 1, k) - v(i, j, k)) -
                      nu_y1 * g.ry_f * (v(i, j, k) - v(i, j - 1, k))) +
           gz.rz_f * (nu_z2 * gz.rz_cp * (v(i, j, k + 1) - v(i, j, k)) -
                      nu_z1 * gz.rz_c * (v(i, j, k) - v(i, j, k - 1)));
    return ((t1 + t2) + t3) + t4;
}

// code/equations.f90:327-468, with dx as for v, dy_3 = y(j)-y(j-1),
// dz_1 = z(k)-z(k-1), dz_2 = z(k+1)-z(k), dz_3 = zg(k+1)-zg(k)
template <typename T, typename A>
__device__ __forceinline__ T rhs_w_gen(const A& F, const RhsXY<T>& g, const RhsZ<T>& gz, T nu, int i, int j, int k)
{
    auto u = [&](int a, int b, int c) { return F(0, a, b, c); };
    auto v = [&](int a, int b, int c) { return F(1, a, b, c); };
    auto w = [&](int a, int b, int c) { return F(2, a, b, c); };
    auto n = [&](int a, int b, int c) { return F(3, a, b, c); };
    const T h = T(0.5), q = T(0.25);
    T wc = h * (w(i, j, k) + w(i, j, k + 1));
    T wcm = h * (w(i, j, k - 1) + w(i, j, k));
    T t1 = -(wc * wc - wcm * wcm) * gz.rz_cp;
    T uwp = (h * (u(i, j, k) + u(i, j, k + 1))) * (h * (w(i, j, k) + w(i + 1, j, k)));
    T uwm = (h * (u(i - 1, j, k) + u(i - 1, j, k + 1))) * (h * (w(i - 1, j, k) + w(i, j, k)));
    T t2 = -(uwp - uwm) * g.rx_f;
    T vwp = (h * (v(i, j, k) + v(i, j, k + 1))) * (g.w0j * w(i, j, k) + g.w1j * w(i, j + 1, k));
    T vwm = (h * (v(i, j - 1, k) + v(i, j - 1, k + 1))) * (g.w0m * w(i, j - 1, k) + g.w1m * w(i, j, k));
    T t3 = -(vwp - vwm) * g.ry_f;
    T nu_x1 = nu + q * (n(i, j, k) + n(i, j, k + 1) + n(i - 1, j, k) + n(i - 1, j, k + 1));
    T nu_x2 = nu + q * (n(i, j, k) + n(i, j, k + 1) + n(i + 1, j, k) + n(i + 1, j, k + 1));
    T nu_y1 = nu + h * (g.w0m * n(i, j - 1, k) + g.w1m * n(i, j, k) + g.w0m * n(i, j - 1, k + 1) +
                        g.w1m * n(i, j, k + 1));
    T nu_y2 = nu + h * (g.w0j * n(i, j, k) + g.w1j * n(i, j + 1, k) + g.w0j * n(i, j, k + 1) +
                        g.w1j * n(i, j + 1, k + 1));
    T nu_z1 = nu + n(i, j, k);
    T nu_z2 = nu + n(i, j, k + 1);
    T dwdy_p = (w(i, j + 1, k) - w(i, j, k)) * g.ry_cp;
    T dwdy_m = (w(i, j, k) - w(i, j - 1, k)) * g.ry_c;
    T t4 = g.rx_f * (nu_x2 * g.rx_cp * (w(i + 1, j, k) - w(i, j, k)) -
                     nu_x1 * g.rx_c * (w(i, j, k) - w(i - 1, j, k))) +
           g.ry_f * (nu_y2 * dwdy_p - nu_y1 * dwdy_m) +
           gz.rz_cp * (nu_z2 * gz.rz_fp * (w(i, j, k + 1) - w(i, j, k)) -
                       nu_z1 * gz.rz_f * (w(i, j, k) - w(i, j, k - 1)));
    return ((t1 + t2) + t3) + t4;
}

// MODE 0: rhs_u/v/w into Ou/Ov/Ow on the reference's interior ranges;
// MODE 1: Wray substep 1; MODE 2: substeps 2 and 3.  In RK modes the x and
// z ranges stop at the last unique cell (nxg-2, nzg-2): the image columns
// the reference computes there are overwritten by the boundary conditions
// that follow.
//
// Pipelining: while plane k is computed from the LDS ring (four slots,
// planes k-1, k, k+1 in use), the loads of planes k+2 and k+3 are in flight
// into two register buffers; plane k+2 goes to the free slot (plane k-2's)
// right after, so one barrier per plane separates the ring's writes from
// its reads, and every load has two planes of work to arrive.
// In substeps 2 and 3 the thread's Uo and F1 values are fetched before the
// stencil as well.
// Every global load is issued unconditionally (positions outside the box
// read a valid address and select 0; the last planes of a chunk re-read
// plane ke+1): the memory counter drains in issue order, and with a fixed
// number of loads per plane the compiler's wait for the oldest buffer
// leaves the newer loads in flight.  A load skipped on some path would
// make every such wait a wait for all outstanding loads.
static constexpr int RHS_KMAX = 32;  // longest z chunk a block marches

// Register budget: three waves per SIMD (168 VGPRs) in FP64 (two for the
// 512-thread tile, which one block per CU holds by its LDS), four in FP32.
template <typename T, int TX, int TY, int MODE, bool G>
__global__ void __launch_bounds__(TX* TY)
    __attribute__((amdgpu_waves_per_eu(sizeof(T) == 8 ? (TX * TY > 256 ? 2 : 3) : 4)))
    k_rhs_rk(RhsArgs<T> one, const RhsArgs<T>* tab, Dims d, T c1, T c2, int kchunk)
{
    constexpr int PX = TX + 2, PY = TY + 2, NP = PX * PY, NT = TX * TY, NL = (NP + NT - 1) / NT;
    __shared__ T sh[4][4][NP];  // [field][slot][position]
    // z reciprocal spacings of the block's planes kb..ke+1, read from LDS per
    // plane: a global load there would wait behind the prefetched planes
    constexpr int KZ = RHS_KMAX + 2;
    __shared__ T zt[2][KZ];  // 1/(z(k)-z(k-1)), 1/(zg(k)-zg(k-1)) at k = kb + q
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * TX + tx;
    const int imax = MODE == 0 ? d.nxg - 1 : d.nxg - 2;
    const int kmax = MODE == 0 ? d.nzg - 1 : d.nzg - 2;
    const GBlk gb = gblk_xcd((kmax - 1 + kchunk - 1) / kchunk);
    const RhsArgs<T> arg = member_args<G>(one, tab, gb.m);
    const T* __restrict__ U = arg.U;
    const T* __restrict__ V = arg.V;
    const T* __restrict__ W = arg.W;
    const T* __restrict__ nt = arg.nt;
    const RhsTabs<T>& tb = arg.tb;
    const T nu = arg.nu;
    const Scalars* sc = arg.sc;
    const T *Uo = arg.Uo, *Vo = arg.Vo, *Wo = arg.Wo;
    T *F1u = arg.F1u, *F1v = arg.F1v, *F1w = arg.F1w, *Ou = arg.Ou, *Ov = arg.Ov, *Ow = arg.Ow;
    const TileId tile{gb.x, gb.y, gb.z};
    const int i0 = tile.x * TX + 2, j0 = tile.y * TY + 2;
    const int kb = tile.z * kchunk + 2;
    const int ke = min(kb + kchunk - 1, kmax);
    const int i = i0 + tx, j = j0 + ty;
    const T* src[4] = {U, V, W, nt};
    // this thread's share of a halo plane: LDS index (-1: none), box offset
    // within a plane (a valid one, 0, where the position is outside the
    // box) and whether the position is inside
    int lq[NL];
    long lo[NL];
    bool lin[NL];
#pragma unroll
    for (int r = 0; r < NL; r++) {
        const int q = tid + r * NT;
        const int lx = q % PX, ly = q / PX;
        const int ii = i0 - 1 + lx, jj = j0 - 1 + ly;
        lq[r] = q < NP ? q : -1;
        lin[r] = q < NP && ii <= d.nxg && jj <= d.nyg;
        lo[r] = lin[r] ? at(d, ii, jj, 1) : 0;
    }
    // two register buffers: the plane stored at the end of an iteration was
    // fetched two iterations earlier
    T preA[NL][4], preB[NL][4];
    auto gload = [&](int p, T (&pre)[NL][4]) {
        const long po = (long)(p - 1) * d.sxy;
#pragma unroll
        for (int r = 0; r < NL; r++)
#pragma unroll
            for (int f = 0; f < 4; f++) {
                const T x = src[f][lo[r] + po];
                pre[r][f] = lin[r] ? x : T(0);
            }
    };
    auto slot = [&](int p) { return (p - kb + 4) & 3; };
    // deferred pass B of the dynamic model (arg.ntab): the nu_t plane holds
    // the value v of code/subgrid.f90:512-535 before the row's ratio, and
    // nu_t = max(A v, B) with the coefficients of the halo row, the same
    // operations as pass B itself (A = -1/2 ratio, B = 0; the wall-model
    // rows hold nu_t and pass through with A = 1, B = -inf)
    const T* ntab = arg.ntab;
    __shared__ T ncf[2][PY];
    if (ntab) {
        for (int q = tid; q < 2 * PY; q += NT) {
            const int c = q / PY, ly = q - c * PY;
            const int jj = j0 - 1 + ly;
            ncf[c][ly] = jj <= d.nyg ? ntab[c * d.nyg + jj - 1] : T(0);
        }
        __syncthreads();
    }
    auto sstore = [&](int p, const T (&pre)[NL][4]) {
        const int s = slot(p);
#pragma unroll
        for (int r = 0; r < NL; r++)
            if (lq[r] >= 0) {
#pragma unroll
                for (int f = 0; f < 3; f++) sh[f][s][lq[r]] = pre[r][f];
                T v = pre[r][3];
                if (ntab) {
                    const int ly = lq[r] / PX;
                    v = fmax(ncf[0][ly] * v, ncf[1][ly]);
                }
                sh[3][s][lq[r]] = v;
            }
    };
    auto F = [&](int f, int a, int b, int c) -> T { return sh[f][slot(c)][(a - i0 + 1) + PX * (b - j0 + 1)]; };
    const T dPdx = T(sc->dPdx);
    const T dt = T(sc->dt);
    RhsXY<T> g;
    const int ic = min(i, imax), jc = min(j, d.nyg - 1);
    {
        g.rx_f = tb.rdx[ic]; g.rx_fp = tb.rdx[ic + 1]; g.rx_c = tb.rdxg[ic]; g.rx_cp = tb.rdxg[ic + 1];
        g.ry_f = tb.rdy[jc]; g.ry_fp = tb.rdy[jc + 1]; g.ry_c = tb.rdyg[jc]; g.ry_cp = tb.rdyg[jc + 1];
        g.w0j = tb.w0[jc]; g.w1j = tb.w1[jc]; g.w0m = tb.w0[jc - 1]; g.w1m = tb.w1[jc - 1];
    }
    for (int q = tid; q < KZ; q += TX * TY) {
        const int kk = min(kb + q, d.nzg);
        zt[0][q] = tb.rdz[kk];
        zt[1][q] = tb.rdzg[kk];
    }
    gload(kb - 1, preA);
    sstore(kb - 1, preA);
    gload(kb, preA);
    sstore(kb, preA);
    gload(kb + 1, preA);
    sstore(kb + 1, preA);
    gload(min(kb + 2, ke + 1), preA);
    __syncthreads();
    const bool act = i <= imax && j <= d.nyg - 1;
    const bool du0 = act && i <= d.nx - 1, dv0 = act && j <= d.ny - 1;
    // the thread's own cell in plane 1 (clamped into the box for inactive
    // threads, which load but do not store)
    const long pc0 = at(d, ic, jc, 1);
    // the y and x images of a stored value: the wall mirror rows (U, W), the
    // first and last unique x columns; z images by plane, below
    const bool yimg = j == 2 || j == d.nyg - 1;
    const bool ximg_f = i == d.nx - 1 || i == 2;
    const bool ximg_c = i == d.nxg - 2 || i == 2 || i == 3;
    // one plane of output from the ring (planes k-1, k, k+1).  The thread's
    // own Uo, F1 loads are issued before the prefetch (fetch): memory
    // counters drain in issue order, so waiting for them does not wait for
    // the prefetched planes.
    auto plane = [&](int k, auto&& fetch) {
        const long p = pc0 + (long)(k - 1) * d.sxy;
        const bool du = du0 && (k <= d.nzg - 1);
        const bool dv = dv0 && (k <= d.nzg - 1);
        const bool dw = act && (k <= d.nz - 1);
        T uo = T(0), vo = T(0), wo = T(0), fu = T(0), fv = T(0), fw = T(0);
        // own-cell values read once and the substep-1 F written once: marked
        // non-temporal so that they do not displace the tiles' halo planes in L2
        if constexpr (MODE == 2) {
            uo = __builtin_nontemporal_load(&Uo[p]); fu = __builtin_nontemporal_load(&F1u[p]);
            vo = __builtin_nontemporal_load(&Vo[p]); fv = __builtin_nontemporal_load(&F1v[p]);
            wo = __builtin_nontemporal_load(&Wo[p]); fw = __builtin_nontemporal_load(&F1w[p]);
        }
        asm volatile("" ::: "memory");  // the own-cell loads before the prefetch
        fetch();
        // an empty asm with a memory clobber keeps the compiler from sinking
        // the own-cell loads into the branches that use them (after the
        // prefetch, where waiting for them would wait for the prefetch too)
        asm volatile("" ::: "memory");
        RhsZ<T> gz;
        gz.rz_f = zt[0][k - kb]; gz.rz_fp = zt[0][k - kb + 1]; gz.rz_c = zt[1][k - kb]; gz.rz_cp = zt[1][k - kb + 1];
        // each component is computed and stored before the next one, which
        // keeps the live registers of one stencil at a time
        const bool zimg_c = k == d.nzg - 2 || k == 2 || k == 3;
        const bool zimg_f = k == d.nz - 1 || k == 2;
        auto put = [&](int kind, bool on, bool img, T r, T old, T f1, T* O, T* F1) {
            if constexpr (MODE == 0) {
                if (on) O[p] = r;
            } else {
                T nv;
                if constexpr (MODE == 1) {
                    // U = Uo + dt*rk_coef(1,1)*Fu1 with Uo = U at substep 1
                    nv = old + dt * c1 * r;
                    if (on) __builtin_nontemporal_store(r, &F1[p]);
                } else {
                    // U = Uo + dt*(c1*Fu1 + c2*Fu_s); rk_coef(3,2) = 0 so the second stage's F is not kept
                    nv = old + dt * (c1 * f1 + c2 * r);
                }
                // the new field with its boundary-condition images: the
                // canonical store, then the images of the rare cells that have any
                if (on) O[p] = nv;
                if (on && img) store_img(O, d, kind, i, j, k, nv);
            }
        };
        // (the branches hold no loads: the number of loads per plane is fixed)
        if (du) {
            const T r = rhs_u_gen<T>(F, g, gz, nu, dPdx, i, j, k);
            put(0, true, ximg_f || yimg || zimg_c, r, MODE == 1 ? F(0, i, j, k) : uo, fu, Ou, F1u);
        }
        if (dv) {
            const T r = rhs_v_gen<T>(F, g, gz, nu, i, j, k);
            put(1, true, ximg_c || zimg_c, r, MODE == 1 ? F(1, i, j, k) : vo, fv, Ov, F1v);
        }
        if (dw) {
            const T r = rhs_w_gen<T>(F, g, gz, nu, i, j, k);
            put(2, true, ximg_c || yimg || zimg_f, r, MODE == 1 ? F(2, i, j, k) : wo, fw, Ow, F1w);
        }
    };
    // unrolled by two so the register buffers alternate without copies;
    // plane k+2 goes to the slot of plane k-2, last read before the
    // previous barrier
    for (int k = kb; k <= ke; k += 2) {
        plane(k, [&] { gload(min(k + 3, ke + 1), preB); });
        if (k + 2 <= ke + 1) sstore(k + 2, preA);
        __syncthreads();
        if (k + 1 > ke) break;
        plane(k + 1, [&] { gload(min(k + 4, ke + 1), preA); });
        if (k + 3 <= ke + 1) sstore(k + 3, preB);
        __syncthreads();
    }
}

// Tile shapes (x cells, y cells) and z-chunk lengths the launcher offers;
// chan_autotune times them on the channel at hand.  kchunk 0: the chunk that
// fills the last round of blocks on this grid and device
// (round_filling_chunk with the kernel's occupancy), counting the tiles of
// every member of a grouped launch (each cell is computed once whatever the
// chunk, so the results do not depend on it).
struct RhsShape { int tx, ty, kchunk; };
static const RhsShape kRhsShapes[] = {
    {64, 4, 16}, {64, 4, 8}, {64, 8, 8}, {64, 8, 16}, {32, 8, 8}, {32, 8, 16}, {64, 4, 32}, {64, 8, 32},
    {64, 4, 0},  {32, 8, 0}, {64, 8, 0},
};

template <typename T, int TX, int TY>
int rhs_round_chunk(int ni, int nj, int nk, int nmem)
{
    int per_cu = 0, ncu = 0, dev = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_rhs_rk<T, TX, TY, 2, false>, TX * TY, 0) != hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1)
        ncu = 1;
    return round_filling_chunk((long)cdiv(ni, TX) * cdiv(nj, TY) * nmem, nk, (long)per_cu * ncu, 6, 32);
}
int rhs_shape_count() { return (int)(sizeof(kRhsShapes) / sizeof(kRhsShapes[0])); }

template <typename T>
void run_rhs_rk(const RhsArgs<T>& one, const RhsArgs<T>* tab, int nmem, const Dims& d, int mode, double c1,
                double c2, int shape, hipStream_t s)
{
    const RhsShape sh = kRhsShapes[(shape >= 0 && shape < rhs_shape_count()) ? shape : 0];
    const int ni = (mode == 0 ? d.nxg - 1 : d.nxg - 2) - 1;
    const int nk = (mode == 0 ? d.nzg - 1 : d.nzg - 2) - 1;
    dim3 blk(sh.tx, sh.ty, 1);
    PROF_BEGIN("k_rhs_rk", s);
#define RHS_MODE(TX, TY, M)                                                                                         \
    do {                                                                                                            \
        if (tab) k_rhs_rk<T, TX, TY, M, true><<<grd, blk, 0, s>>>(one, tab, d, T(c1), T(c2), kc);                   \
        else k_rhs_rk<T, TX, TY, M, false><<<grd, blk, 0, s>>>(one, tab, d, T(c1), T(c2), kc);                      \
    } while (0)
#define RHS_GO(TX, TY)                                                                                              \
    do {                                                                                                            \
        const int kc = std::min(RHS_KMAX, sh.kchunk > 0 ? sh.kchunk : rhs_round_chunk<T, TX, TY>(ni, d.nyg - 2, nk, nmem)); \
        dim3 grd(cdiv(ni, TX), cdiv(d.nyg - 2, TY), (unsigned)cdiv(nk, kc) * nmem);                                 \
        if (mode == 0) RHS_MODE(TX, TY, 0);                                                                         \
        else if (mode == 1) RHS_MODE(TX, TY, 1);                                                                    \
        else RHS_MODE(TX, TY, 2);                                                                                   \
    } while (0)
    if (sh.tx == 64 && sh.ty == 4) RHS_GO(64, 4);
    else if (sh.tx == 64 && sh.ty == 8) RHS_GO(64, 8);
    else RHS_GO(32, 8);
#undef RHS_GO
#undef RHS_MODE
    PROF_END(s);
}

template <typename T>
void launch_rhs_rk(const T* U, const T* V, const T* W, const T* nut, const RhsTabs<T>& m, const Dims& d, T nu,
                   const Scalars* sc, int mode, const T* Uo, const T* Vo, const T* Wo, T* F1u, T* F1v, T* F1w,
                   T* Ou, T* Ov, T* Ow, double c1, double c2, int shape, hipStream_t s, const T* ntab)
{
    run_rhs_rk<T>(RhsArgs<T>{U, V, W, nut, ntab, m, nu, sc, Uo, Vo, Wo, F1u, F1v, F1w, Ou, Ov, Ow}, nullptr, 1, d,
                  mode, c1, c2, shape, s);
}

template void launch_rhs_rk<double>(const double*, const double*, const double*, const double*,
                                    const RhsTabs<double>&, const Dims&, double, const Scalars*, int, const double*,
                                    const double*, const double*, double*, double*, double*, double*, double*,
                                    double*, double, double, int, hipStream_t, const double*);
template void run_rhs_rk<double>(const RhsArgs<double>&, const RhsArgs<double>*, int, const Dims&, int, double,
                                 double, int, hipStream_t);
template void run_rhs_rk<float>(const RhsArgs<float>&, const RhsArgs<float>*, int, const Dims&, int, double, double,
                                int, hipStream_t);
template void launch_rhs_rk<float>(const float*, const float*, const float*, const float*, const RhsTabs<float>&,
                                   const Dims&, float, const Scalars*, int, const float*, const float*, const float*,
                                   float*, float*, float*, float*, float*, float*, double, double, int, hipStream_t,
                                   const float*);

}  // namespace chan
