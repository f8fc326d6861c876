// amd.hip -- the anisotropic minimum dissipation (AMD) eddy viscosity
// (code/subgrid.f90:632-874), LES_model 3:
//   nu_t = max(-C N / D, 0), C = 0.3,
//   N = sum_ijk (dh_k u_i)(dh_k u_j) S_ij   (dh: undivided difference)
//   D = sum_lm (d_l u_m)^2
// One thread per box cell: the unique interior cells compute the model from
// U, V, W in L1/L2 (the reads of neighbouring threads overlap), image cells
// copy their canonical cell's value and the wall ghost rows follow
// Dirichlet_nu_t, so no boundary pass follows.  The staggered differences
// reach the centre through the reference's face-to-centre averages in the
// order of its interpolate calls (code/interpolation.f90:24-95), and the
// terms are summed in the order of its statements.
#include "kernels.hpp"
#include "prof.hpp"
#include "wallmodel.hpp"

namespace chan {

template <typename T>
__device__ __forceinline__ T amd_at(const T* U, const T* V, const T* W, const Metric<T>& m, const Dims& d, int i,
                                    int j, int k)
{
    const T *x = m.x, *y = m.y, *z = m.z, *xg = m.xg, *yg = m.yg, *zg = m.zg;
    const T h = T(0.5), qq = T(0.25);
    auto u = [&](int a, int b, int c) { return U[at(d, a, b, c)]; };
    auto v = [&](int a, int b, int c) { return V[at(d, a, b, c)]; };
    auto w = [&](int a, int b, int c) { return W[at(d, a, b, c)]; };
    auto ax = [&](T a, T b) { return h * (a + b); };
    // rate of strain at the centre (code/subgrid.f90:1027-1059)
    const T S11 = (u(i, j, k) - u(i - 1, j, k)) / (x[i] - x[i - 1]);
    const T S22 = (v(i, j, k) - v(i, j - 1, k)) / (y[j] - y[j - 1]);
    const T S33 = (w(i, j, k) - w(i, j, k - 1)) / (z[k] - z[k - 1]);
    auto e4 = [&](int a, int b) {
        return h * ((u(a - 1, b, k) - u(a - 1, b - 1, k)) / (yg[b] - yg[b - 1]) +
                    (v(a, b - 1, k) - v(a - 1, b - 1, k)) / (xg[a] - xg[a - 1]));
    };
    auto e5 = [&](int a, int c) {
        return h * ((u(a - 1, j, c) - u(a - 1, j, c - 1)) / (zg[c] - zg[c - 1]) +
                    (w(a, j, c - 1) - w(a - 1, j, c - 1)) / (xg[a] - xg[a - 1]));
    };
    auto e6 = [&](int b, int c) {
        return h * ((v(i, b - 1, c) - v(i, b - 1, c - 1)) / (zg[c] - zg[c - 1]) +
                    (w(i, b, c - 1) - w(i, b - 1, c - 1)) / (yg[b] - yg[b - 1]));
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
    const T dyg0 = yg[j] - yg[j - 1], dyg1 = yg[j + 1] - yg[j];
    const T dzg0 = zg[k] - zg[k - 1], dzg1 = zg[k + 1] - zg[k];
    const T dxg0 = xg[i] - xg[i - 1], dxg1 = xg[i + 1] - xg[i];
    const T Uyd = ax(ax(dUy(i - 1, j, k) / dyg0, dUy(i, j, k) / dyg0), ax(dUy(i - 1, j + 1, k) / dyg1, dUy(i, j + 1, k) / dyg1));
    const T Uzd = ax(ax(dUz(i - 1, j, k) / dzg0, dUz(i, j, k) / dzg0), ax(dUz(i - 1, j, k + 1) / dzg1, dUz(i, j, k + 1) / dzg1));
    const T Vxd = ax(ax(dVx(i, j - 1, k) / dxg0, dVx(i, j, k) / dxg0), ax(dVx(i + 1, j - 1, k) / dxg1, dVx(i + 1, j, k) / dxg1));
    const T Vzd = ax(ax(dVz(i, j - 1, k) / dzg0, dVz(i, j, k) / dzg0), ax(dVz(i, j - 1, k + 1) / dzg1, dVz(i, j, k + 1) / dzg1));
    const T Wxd = ax(ax(dWx(i, j, k - 1) / dxg0, dWx(i, j, k) / dxg0), ax(dWx(i + 1, j, k - 1) / dxg1, dWx(i + 1, j, k) / dxg1));
    const T Wyd = ax(ax(dWy(i, j, k - 1) / dyg0, dWy(i, j, k) / dyg0), ax(dWy(i, j + 1, k - 1) / dyg1, dWy(i, j + 1, k) / dyg1));
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
    const T dx = x[i] - x[i - 1], dy = y[j] - y[j - 1], dz = z[k] - z[k - 1];
    T D = Uxh * Uxh / (dx * dx);
    D = D + Uyd * Uyd;
    D = D + Uzd * Uzd;
    D = D + Vyh * Vyh / (dy * dy);
    D = D + Vxd * Vxd;
    D = D + Vzd * Vzd;
    D = D + (Wzh * Wzh) / (dz * dz);
    D = D + Wxd * Wxd;
    D = D + Wyd * Wyd;
    // clipping; 0/0 (uniform neighbourhood) is NaN and clips to 0
    return fmax(-T(0.3) * N / D, T(0));
}

// every box cell: interior from the model at its canonical image, wall ghost
// rows per Dirichlet_nu_t (1 mirror, 0 copy, 2 mirror plus the offset
// 0.38 sqrt(dPdx) (y(2)-y(1)), code/subgrid.f90:864-871); with fuse_wm the
// wall rows come from the wall model instead
template <typename T>
__global__ void __launch_bounds__(256) k_amd(const T* U, const T* V, const T* W, T* nut, Metric<T> m, Dims d,
                                             int dirichlet, const Scalars* sc, int fuse_wm, const double* alpha, T nu)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > d.nxg || j > d.nyg) return;
    const long p = at(d, i, j, k);
    if (fuse_wm && (j <= 2 || j >= d.nyg - 1)) {
        nut[p] = wm_nut_at<T>(U, alpha, m, d, nu, img_centre(i, d.nxg), k, j <= 2 ? 0 : 1, sc);
        return;
    }
    const int jj = (j == 1) ? 2 : (j == d.nyg ? d.nyg - 1 : j);
    const T v = amd_at<T>(U, V, W, m, d, img_centre(i, d.nxg), jj, img_centre(k, d.nzg));
    T out = v;
    if (jj != j) {
        if (dirichlet == 1) out = -v;
        else if (dirichlet == 2) out = -v + T(2.0 * 0.38 * sqrt(sc->dPdx) * (double)(m.y[2] - m.y[1]) / 2.0);
    }
    nut[p] = out;
}

template <typename T>
void launch_amd(const T* U, const T* V, const T* W, T* nut, const Metric<T>& m, const Dims& d, int dirichlet,
                const Scalars* sc, int fuse_wm, const double* alpha, T nu, hipStream_t s)
{
    dim3 blk(64, 4, 1), grd(cdiv(d.nxg, 64), cdiv(d.nyg, 4), d.nzg);
    PROF_BEGIN("k_amd", s);
    k_amd<T><<<grd, blk, 0, s>>>(U, V, W, nut, m, d, dirichlet, sc, fuse_wm, alpha, nu);
    PROF_END(s);
}

template void launch_amd<double>(const double*, const double*, const double*, double*, const Metric<double>&,
                                 const Dims&, int, const Scalars*, int, const double*, double, hipStream_t);
template void launch_amd<float>(const float*, const float*, const float*, float*, const Metric<float>&, const Dims&,
                                int, const Scalars*, int, const double*, float, hipStream_t);

}  // namespace chan
