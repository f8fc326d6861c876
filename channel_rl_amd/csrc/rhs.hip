// rhs.hip -- right-hand sides of u, v, w (code/equations.f90:24-468) and the
// Wray RK3 combine (code/time_integration.f90:42-44,59-61,76-81).
//
// A block owns a TX x TY tile of (x,y) cells and marches in z.  U, V, W and
// nu_t for planes k-1, k, k+1 of the tile plus a one-cell halo sit in LDS
// (a three-plane ring), so every value is read from HBM once per tile and
// the ~120 stencil operands of a cell come from LDS.  The arithmetic is the
// reference's operation order (no FMA contraction in FP64).
#include "kernels.hpp"
#include "prof.hpp"

namespace chan {

// A(f, i, j, k) returns field f (0 U, 1 V, 2 W, 3 nu_t) at Fortran (i,j,k).
template <typename T, typename A>
__device__ __forceinline__ T rhs_u_gen(const A& F, const Metric<T>& m, T nu, T dPdx, int i, int j, int k)
{
    auto u = [&](int a, int b, int c) { return F(0, a, b, c); };
    auto v = [&](int a, int b, int c) { return F(1, a, b, c); };
    auto w = [&](int a, int b, int c) { return F(2, a, b, c); };
    auto n = [&](int a, int b, int c) { return F(3, a, b, c); };
    const T h = T(0.5), q = T(0.25), one = T(1);
    const T *x = m.x, *y = m.y, *z = m.z, *xg = m.xg, *yg = m.yg, *zg = m.zg, *w0 = m.w0, *w1 = m.w1;
    T uc = h * (u(i, j, k) + u(i + 1, j, k));
    T ucm = h * (u(i - 1, j, k) + u(i, j, k));
    T t1 = -(uc * uc - ucm * ucm) / (xg[i + 1] - xg[i]);
    T uvp = (w0[j] * u(i, j, k) + w1[j] * u(i, j + 1, k)) * (h * (v(i, j, k) + v(i + 1, j, k)));
    T uvm = (w0[j - 1] * u(i, j - 1, k) + w1[j - 1] * u(i, j, k)) * (h * (v(i, j - 1, k) + v(i + 1, j - 1, k)));
    T t2 = -(uvp - uvm) / (y[j] - y[j - 1]);
    T uwp = (h * (u(i, j, k) + u(i, j, k + 1))) * (h * (w(i, j, k) + w(i + 1, j, k)));
    T uwm = (h * (u(i, j, k - 1) + u(i, j, k))) * (h * (w(i, j, k - 1) + w(i + 1, j, k - 1)));
    T t3 = -(uwp - uwm) / (z[k] - z[k - 1]);
    T dz_1 = zg[k] - zg[k - 1], dz_2 = zg[k + 1] - zg[k], dz_3 = z[k] - z[k - 1];
    T dy_3 = y[j] - y[j - 1];
    T dx_1 = x[i] - x[i - 1], dx_2 = x[i + 1] - x[i], dx_3 = xg[i + 1] - xg[i];
    T nu_x1 = nu + n(i, j, k);
    T nu_x2 = nu + n(i + 1, j, k);
    T nu_y1 = nu + h * (w0[j - 1] * n(i, j - 1, k) + w1[j - 1] * n(i, j, k) + w0[j - 1] * n(i + 1, j - 1, k) +
                        w1[j - 1] * n(i + 1, j, k));
    T nu_y2 = nu + h * (w0[j] * n(i, j, k) + w1[j] * n(i, j + 1, k) + w0[j] * n(i + 1, j, k) +
                        w1[j] * n(i + 1, j + 1, k));
    T nu_z1 = nu + q * (n(i, j, k) + n(i, j, k - 1) + n(i + 1, j, k) + n(i + 1, j, k - 1));
    T nu_z2 = nu + q * (n(i, j, k) + n(i, j, k + 1) + n(i + 1, j, k) + n(i + 1, j, k + 1));
    T dudy_p = (u(i, j + 1, k) - u(i, j, k)) / (yg[j + 1] - yg[j]);
    T dudy_m = (u(i, j, k) - u(i, j - 1, k)) / (yg[j] - yg[j - 1]);
    T t4 = one / dx_3 * (nu_x2 * one / dx_2 * (u(i + 1, j, k) - u(i, j, k)) -
                         nu_x1 * one / dx_1 * (u(i, j, k) - u(i - 1, j, k))) +
           one / dy_3 * (nu_y2 * dudy_p - nu_y1 * dudy_m) +
           one / dz_3 * (nu_z2 * one / dz_2 * (u(i, j, k + 1) - u(i, j, k)) -
                         nu_z1 * one / dz_1 * (u(i, j, k) - u(i, j, k - 1)));
    return (((t1 + t2) + t3) + t4) + dPdx;
}

template <typename T, typename A>
__device__ __forceinline__ T rhs_v_gen(const A& F, const Metric<T>& m, T nu, int i, int j, int k)
{
    auto u = [&](int a, int b, int c) { return F(0, a, b, c); };
    auto v = [&](int a, int b, int c) { return F(1, a, b, c); };
    auto w = [&](int a, int b, int c) { return F(2, a, b, c); };
    auto n = [&](int a, int b, int c) { return F(3, a, b, c); };
    const T h = T(0.5), one = T(1);
    const T *x = m.x, *y = m.y, *z = m.z, *xg = m.xg, *yg = m.yg, *zg = m.zg, *w0 = m.w0, *w1 = m.w1;
    T vc = h * (v(i, j, k) + v(i, j + 1, k));
    T vcm = h * (v(i, j - 1, k) + v(i, j, k));
    T t1 = -(vc * vc - vcm * vcm) / (yg[j + 1] - yg[j]);
    T uvp = (w0[j] * u(i, j, k) + w1[j] * u(i, j + 1, k)) * (h * (v(i, j, k) + v(i + 1, j, k)));
    T uvm = (w0[j] * u(i - 1, j, k) + w1[j] * u(i - 1, j + 1, k)) * (h * (v(i - 1, j, k) + v(i, j, k)));
    T t2 = -(uvp - uvm) / (x[i] - x[i - 1]);
    T vwp = (h * (v(i, j, k) + v(i, j, k + 1))) * (w0[j] * w(i, j, k) + w1[j] * w(i, j + 1, k));
    T vwm = (h * (v(i, j, k - 1) + v(i, j, k))) * (w0[j] * w(i, j, k - 1) + w1[j] * w(i, j + 1, k - 1));
    T t3 = -(vwp - vwm) / (z[k] - z[k - 1]);
    T dz_1 = zg[k] - zg[k - 1], dz_2 = zg[k + 1] - zg[k], dz_3 = z[k] - z[k - 1];
    T dy_1 = y[j] - y[j - 1], dy_2 = y[j + 1] - y[j], dy_3 = yg[j + 1] - yg[j];
    T dx_1 = xg[i] - xg[i - 1], dx_2 = xg[i + 1] - xg[i], dx_3 = x[i] - x[i - 1];
    T nu_x1 = nu + h * (w0[j] * n(i - 1, j, k) + w1[j] * n(i - 1, j + 1, k) + w0[j] * n(i, j, k) +
                        w1[j] * n(i, j + 1, k));
    T nu_x2 = nu + h * (w0[j] * n(i, j, k) + w1[j] * n(i, j + 1, k) + w0[j] * n(i + 1, j, k) +
                        w1[j] * n(i + 1, j + 1, k));
    T nu_y1 = nu + n(i, j, k);
    T nu_y2 = nu + n(i, j + 1, k);
    T nu_z1 = nu + h * (w0[j] * n(i, j, k - 1) + w1[j] * n(i, j + 1, k - 1) + w0[j] * n(i, j, k) +
                        w1[j] * n(i, j + 1, k));
    T nu_z2 = nu + h * (w0[j] * n(i, j, k) + w1[j] * n(i, j + 1, k) + w0[j] * n(i, j, k + 1) +
                        w1[j] * n(i, j + 1, k + 1));
    T t4 = one / dx_3 * (nu_x2 * one / dx_2 * (v(i + 1, j, k) - v(i, j, k)) -
                         nu_x1 * one / dx_1 * (v(i, j, k) - v(i - 1, j, k))) +
           one / dy_3 * (nu_y2 * one / dy_2 * (v(i, j + 1, k) - v(i, j, k)) -
                         nu_y1 * one / dy_1 * (v(i, j, k) - v(i, j - 1, k))) +
           one / dz_3 * (nu_z2 * one / dz_2 * (v(i, j, k + 1) - v(i, j, k)) -
                         nu_z1 * one / dz_1 * (v(i, j, k) - v(i, j, k - 1)));
    return ((t1 + t2) + t3) + t4;
}

template <typename T, typename A>
__device__ __forceinline__ T rhs_w_gen(const A& F, const Metric<T>& m, T nu, int i, int j, int k)
{
    auto u = [&](int a, int b, int c) { return F(0, a, b, c); };
    auto v = [&](int a, int b, int c) { return F(1, a, b, c); };
    auto w = [&](int a, int b, int c) { return F(2, a, b, c); };
    auto n = [&](int a, int b, int c) { return F(3, a, b, c); };
    const T h = T(0.5), q = T(0.25), one = T(1);
    const T *x = m.x, *y = m.y, *z = m.z, *xg = m.xg, *yg = m.yg, *zg = m.zg, *w0 = m.w0, *w1 = m.w1;
    T wc = h * (w(i, j, k) + w(i, j, k + 1));
    T wcm = h * (w(i, j, k - 1) + w(i, j, k));
    T t1 = -(wc * wc - wcm * wcm) / (zg[k + 1] - zg[k]);
    T uwp = (h * (u(i, j, k) + u(i, j, k + 1))) * (h * (w(i, j, k) + w(i + 1, j, k)));
    T uwm = (h * (u(i - 1, j, k) + u(i - 1, j, k + 1))) * (h * (w(i - 1, j, k) + w(i, j, k)));
    T t2 = -(uwp - uwm) / (x[i] - x[i - 1]);
    T vwp = (h * (v(i, j, k) + v(i, j, k + 1))) * (w0[j] * w(i, j, k) + w1[j] * w(i, j + 1, k));
    T vwm = (h * (v(i, j - 1, k) + v(i, j - 1, k + 1))) * (w0[j - 1] * w(i, j - 1, k) + w1[j - 1] * w(i, j, k));
    T t3 = -(vwp - vwm) / (y[j] - y[j - 1]);
    T dz_1 = z[k] - z[k - 1], dz_2 = z[k + 1] - z[k], dz_3 = zg[k + 1] - zg[k];
    T dy_3 = y[j] - y[j - 1];
    T dx_1 = xg[i] - xg[i - 1], dx_2 = xg[i + 1] - xg[i], dx_3 = x[i] - x[i - 1];
    T nu_x1 = nu + q * (n(i, j, k) + n(i, j, k + 1) + n(i - 1, j, k) + n(i - 1, j, k + 1));
    T nu_x2 = nu + q * (n(i, j, k) + n(i, j, k + 1) + n(i + 1, j, k) + n(i + 1, j, k + 1));
    T nu_y1 = nu + h * (w0[j - 1] * n(i, j - 1, k) + w1[j - 1] * n(i, j, k) + w0[j - 1] * n(i, j - 1, k + 1) +
                        w1[j - 1] * n(i, j, k + 1));
    T nu_y2 = nu + h * (w0[j] * n(i, j, k) + w1[j] * n(i, j + 1, k) + w0[j] * n(i, j, k + 1) +
                        w1[j] * n(i, j + 1, k + 1));
    T nu_z1 = nu + n(i, j, k);
    T nu_z2 = nu + n(i, j, k + 1);
    T dwdy_p = (w(i, j + 1, k) - w(i, j, k)) / (yg[j + 1] - yg[j]);
    T dwdy_m = (w(i, j, k) - w(i, j - 1, k)) / (yg[j] - yg[j - 1]);
    T t4 = one / dx_3 * (nu_x2 * one / dx_2 * (w(i + 1, j, k) - w(i, j, k)) -
                         nu_x1 * one / dx_1 * (w(i, j, k) - w(i - 1, j, k))) +
           one / dy_3 * (nu_y2 * dwdy_p - nu_y1 * dwdy_m) +
           one / dz_3 * (nu_z2 * one / dz_2 * (w(i, j, k + 1) - w(i, j, k)) -
                         nu_z1 * one / dz_1 * (w(i, j, k) - w(i, j, k - 1)));
    return ((t1 + t2) + t3) + t4;
}

template <int TX, int TY>
struct RhsTile {
    static constexpr int PX = TX + 2, PY = TY + 2, NP = PX * PY;
};

// mode 0: rhs_u/v/w into Ou/Ov/Ow on the reference's interior ranges;
// mode s = 1,2,3: Wray substep s.  In RK modes the x and z ranges stop at the
// last unique cell (nxg-2, nzg-2): the image columns the reference computes
// there are overwritten by the boundary conditions that follow.
template <typename T, int TX, int TY>
__global__ void __launch_bounds__(TX* TY)
    k_rhs_rk(const T* __restrict__ U, const T* __restrict__ V, const T* __restrict__ W, const T* __restrict__ nt,
             Metric<T> m, Dims d, T nu, const Scalars* sc, int mode, const T* Uo, const T* Vo, const T* Wo, T* F1u,
             T* F1v, T* F1w, T* Ou, T* Ov, T* Ow, T c1, T c2, int kchunk)
{
    using Tl = RhsTile<TX, TY>;
    constexpr int PX = Tl::PX, NP = Tl::NP;
    __shared__ T sh[4][3][NP];
    const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * TX + tx;
    const int i0 = blockIdx.x * TX + 2, j0 = blockIdx.y * TY + 2;
    const int imax = mode == 0 ? d.nxg - 1 : d.nxg - 2;
    const int kmax = mode == 0 ? d.nzg - 1 : d.nzg - 2;
    const int kb = blockIdx.z * kchunk + 2;
    const int ke = min(kb + kchunk - 1, kmax);
    const int i = i0 + tx, j = j0 + ty;
    const T* src[4] = {U, V, W, nt};
    auto slot = [&](int p) { return (p - kb + 3) % 3; };
    auto load = [&](int p) {
        const int s = slot(p);
        for (int q = tid; q < NP; q += TX * TY) {
            const int lx = q % PX, ly = q / PX;
            const int ii = i0 - 1 + lx, jj = j0 - 1 + ly;
            const bool ok = ii <= d.nxg && jj <= d.nyg;
            const long o = ok ? at(d, ii, jj, p) : 0;
            for (int f = 0; f < 4; f++) sh[f][s][q] = ok ? src[f][o] : T(0);
        }
    };
    auto F = [&](int f, int a, int b, int c) -> T { return sh[f][slot(c)][(a - i0 + 1) + PX * (b - j0 + 1)]; };
    const T dPdx = T(sc->dPdx);
    const T dt = T(sc->dt);
    load(kb - 1);
    load(kb);
    for (int k = kb; k <= ke; k++) {
        load(k + 1);
        __syncthreads();
        if (i <= imax && j <= d.nyg - 1) {
            const long p = at(d, i, j, k);
            const bool du = (i <= d.nx - 1) && (k <= d.nzg - 1);
            const bool dv = (j <= d.ny - 1) && (k <= d.nzg - 1);
            const bool dw = (k <= d.nz - 1);
            T ru = T(0), rv = T(0), rw = T(0);
            if (du) ru = rhs_u_gen<T>(F, m, nu, dPdx, i, j, k);
            if (dv) rv = rhs_v_gen<T>(F, m, nu, i, j, k);
            if (dw) rw = rhs_w_gen<T>(F, m, nu, i, j, k);
            if (mode == 0) {
                if (du) Ou[p] = ru;
                if (dv) Ov[p] = rv;
                if (dw) Ow[p] = rw;
            } else if (mode == 1) {
                // U = Uo + dt*rk_coef(1,1)*Fu1 with Uo = U at substep 1; the
                // new field is stored with its boundary-condition images
                if (du) { F1u[p] = ru; store_img(Ou, d, 0, i, j, k, F(0, i, j, k) + dt * c1 * ru); }
                if (dv) { F1v[p] = rv; store_img(Ov, d, 1, i, j, k, F(1, i, j, k) + dt * c1 * rv); }
                if (dw) { F1w[p] = rw; store_img(Ow, d, 2, i, j, k, F(2, i, j, k) + dt * c1 * rw); }
            } else {
                // U = Uo + dt*(c1*Fu1 + c2*Fu_s); rk_coef(3,2) = 0 so the second stage's F is not kept
                if (du) store_img(Ou, d, 0, i, j, k, Uo[p] + dt * (c1 * F1u[p] + c2 * ru));
                if (dv) store_img(Ov, d, 1, i, j, k, Vo[p] + dt * (c1 * F1v[p] + c2 * rv));
                if (dw) store_img(Ow, d, 2, i, j, k, Wo[p] + dt * (c1 * F1w[p] + c2 * rw));
            }
        }
        __syncthreads();
    }
}

static constexpr int RHS_TX = 64, RHS_TY = 4, RHS_KCHUNK = 16;

template <typename T>
void launch_rhs_rk(const T* U, const T* V, const T* W, const T* nut, const Metric<T>& m, const Dims& d, T nu,
                   const Scalars* sc, int mode, const T* Uo, const T* Vo, const T* Wo, T* F1u, T* F1v, T* F1w,
                   T* Ou, T* Ov, T* Ow, double c1, double c2, hipStream_t s)
{
    const int ni = (mode == 0 ? d.nxg - 1 : d.nxg - 2) - 1;
    const int nk = (mode == 0 ? d.nzg - 1 : d.nzg - 2) - 1;
    dim3 blk(RHS_TX, RHS_TY, 1);
    dim3 grd(cdiv(ni, RHS_TX), cdiv(d.nyg - 2, RHS_TY), cdiv(nk, RHS_KCHUNK));
    PROF_BEGIN("k_rhs_rk", s);
    k_rhs_rk<T, RHS_TX, RHS_TY><<<grd, blk, 0, s>>>(U, V, W, nut, m, d, nu, sc, mode, Uo, Vo, Wo, F1u, F1v, F1w, Ou,
                                                    Ov, Ow, T(c1), T(c2), RHS_KCHUNK);
    PROF_END(s);
}

template void launch_rhs_rk<double>(const double*, const double*, const double*, const double*,
                                    const Metric<double>&, const Dims&, double, const Scalars*, int, const double*,
                                    const double*, const double*, double*, double*, double*, double*, double*,
                                    double*, double, double, hipStream_t);
template void launch_rhs_rk<float>(const float*, const float*, const float*, const float*, const Metric<float>&,
                                   const Dims&, float, const Scalars*, int, const float*, const float*, const float*,
                                   float*, float*, float*, float*, float*, float*, double, double, hipStream_t);

}  // namespace chan
