// pressure.hip -- the true pressure for output (code/pressure.f90:21-340):
// Laplacian(p) = div(rhs_u, rhs_v, rhs_w) with the wall conditions
// dP/dy = d/dy((nu + nu_t) dV/dy).  The solve reuses the projection's
// FFT/tridiagonal path: the reference adds the transformed wall data to the
// first and last row of every mode (code/pressure.f90:137-139); by linearity
// the same terms are added to those rows of the right-hand side before the
// transform.
#include "kernels.hpp"
#include "prof.hpp"

namespace chan {

// code/pressure.f90:272-336: one thread per interior (i, k) of both walls;
// bc planes (nxg x nzg), entries (i, k) at (i-1) + nxg (k-1)
template <typename T>
__global__ void k_pressure_bc(const T* V, const T* nut, Metric<double> m, Dims d, double nu, double* bc1, double* bc2)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 2;
    const int k = blockIdx.y + 2;
    if (i > d.nxg - 1 || k > d.nzg - 1) return;
    const double *y = m.y, *yg = m.yg;
    const int ny = d.ny, nyg = d.nyg;
    auto v = [&](int j) { return (double)V[at(d, i, j, k)]; };
    auto n = [&](int j) { return (double)nut[at(d, i, j, k)]; };
    const double dzeta = 1.0;
    {
        const double g = (-3.0 * y[1] + 4.0 * y[2] - y[3]) / (2.0 * dzeta);
        const double g2 = (2.0 * y[1] - 5.0 * y[2] + 4.0 * y[3] - 1.0 * y[4]) / (dzeta * dzeta);
        const double nui = nu + 0.5 * (n(1) + n(2));
        const double dnu = (n(2) - n(1)) / (yg[2] - yg[1]);
        const double dV = (-3.0 * v(1) + 4.0 * v(2) - v(3)) / (2.0 * dzeta) * 1.0 / g;
        double ddV = (2.0 * v(1) - 5.0 * v(2) + 4.0 * v(3) - 1.0 * v(4)) / (dzeta * dzeta);
        ddV = (ddV - dV * g2) / (g * g);
        bc1[(i - 1) + (long)d.nxg * (k - 1)] = dnu * dV + nui * ddV;
    }
    {
        const double g = (3.0 * y[ny] - 4.0 * y[ny - 1] + y[ny - 2]) / (2.0 * dzeta);
        const double g2 = (2.0 * y[ny] - 5.0 * y[ny - 1] + 4.0 * y[ny - 2] - 1.0 * y[ny - 3]) / (dzeta * dzeta);
        const double nui = nu + 0.5 * (n(nyg) + n(nyg - 1));
        const double dnu = (n(nyg) - n(nyg - 1)) / (yg[nyg] - yg[nyg - 1]);
        const double dV = (3.0 * v(ny) - 4.0 * v(ny - 1) + v(ny - 2)) / (2.0 * dzeta) * 1.0 / g;
        double ddV = (2.0 * v(ny) - 5.0 * v(ny - 1) + 4.0 * v(ny - 2) - 1.0 * v(ny - 3)) / (dzeta * dzeta);
        ddV = (ddV - dV * g2) / (g * g);
        bc2[(i - 1) + (long)d.nxg * (k - 1)] = dnu * dV + nui * ddV;
    }
}

// code/pressure.f90:42-67: divergence of the right-hand sides at the unique
// centres, rhs_u x-periodic (face 1 = face nx-1), rhs_v's wall faces from
// bc, rhs_w z-periodic (face 1 = face nz-1); plus the wall-data terms of
// rows 2 and nyg-1 (coef_bc_1 (yg(2)-yg(1)) bc_1 and -coef_bc_2 (yg(nyg) -
// yg(nyg-1)) bc_2, code/initialization.f90:370-385)
template <typename T>
__global__ void k_pressure_rhs(const T* Ru, const T* Rv, const T* Rw, const double* bc1, const double* bc2,
                               Metric<double> m, Dims d, double* P)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 2;
    const int j = blockIdx.y + 2;
    const int k = blockIdx.z + 2;
    if (i > d.nxg - 2 || j > d.nyg - 1 || k > d.nzg - 2) return;
    const double *x = m.x, *y = m.y, *z = m.z, *yg = m.yg;
    const long b = (i - 1) + (long)d.nxg * (k - 1);
    auto ru = [&](int a) { return (double)Ru[at(d, a == 1 ? d.nx - 1 : a, j, k)]; };
    auto rv = [&](int a) { return a == 1 ? bc1[b] : (a == d.ny ? bc2[b] : (double)Rv[at(d, i, a, k)]); };
    auto rw = [&](int a) { return (double)Rw[at(d, i, j, a == 1 ? d.nz - 1 : a)]; };
    double r = (ru(i) - ru(i - 1)) / (x[i] - x[i - 1]) + (rv(j) - rv(j - 1)) / (y[j] - y[j - 1]) +
               (rw(k) - rw(k - 1)) / (z[k] - z[k - 1]);
    const int nyg = d.nyg;
    if (j == 2) r = r + 1.0 / (y[2] - y[1]) / (yg[2] - yg[1]) * (yg[2] - yg[1]) * bc1[b];
    if (j == nyg - 1)
        r = r - 1.0 / (y[nyg - 1] - y[nyg - 2]) / (yg[nyg] - yg[nyg - 1]) * (yg[nyg] - yg[nyg - 1]) * bc2[b];
    P[at(d, i, j, k)] = r;
}

// code/pressure.f90:153-167: wall ghost rows P(1) = P(2) - dy bc_1,
// P(nyg) = P(nyg-1) + dy bc_2, and every periodic image (x and z: 1 <-
// ng-2, ng-1 <- 2, ng <- 3), each cell from its canonical cell
__global__ void k_pressure_ghosts(double* P, const double* bc1, const double* bc2, Metric<double> m, Dims d)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 1;
    const int k = blockIdx.z + 1;
    if (i > d.nxg || j > d.nyg) return;
    const int ic = img_centre(i, d.nxg), kc = img_centre(k, d.nzg);
    const bool canon = (ic == i) && (kc == k);
    const double* yg = m.yg;
    const long b = (ic - 1) + (long)d.nxg * (kc - 1);
    double v;
    if (j == 1) v = P[at(d, ic, 2, kc)] - (yg[2] - yg[1]) * bc1[b];
    else if (j == d.nyg) v = P[at(d, ic, d.nyg - 1, kc)] + (yg[d.nyg] - yg[d.nyg - 1]) * bc2[b];
    else if (!canon) v = P[at(d, ic, j, kc)];
    else return;
    P[at(d, i, j, k)] = v;
}

template <typename T>
void launch_pressure_rhs(const T* V, const T* nut, const T* Ru, const T* Rv, const T* Rw, const Metric<double>& m,
                         const Dims& d, double nu, double* bc1, double* bc2, double* P, hipStream_t s)
{
    PROF_BEGIN("k_pressure_bc", s);
    k_pressure_bc<T><<<dim3(cdiv(d.nxg - 2, 64), d.nzg - 2), 64, 0, s>>>(V, nut, m, d, nu, bc1, bc2);
    PROF_END(s);
    PROF_BEGIN("k_pressure_rhs", s);
    k_pressure_rhs<T><<<dim3(cdiv(d.nxg - 3, 64), d.nyg - 2, d.nzg - 3), 64, 0, s>>>(Ru, Rv, Rw, bc1, bc2, m, d, P);
    PROF_END(s);
}

void launch_pressure_ghosts(double* P, const double* bc1, const double* bc2, const Metric<double>& m, const Dims& d,
                            hipStream_t s)
{
    dim3 blk(64, 4, 1), grd(cdiv(d.nxg, 64), cdiv(d.nyg, 4), d.nzg);
    PROF_BEGIN("k_pressure_ghosts", s);
    k_pressure_ghosts<<<grd, blk, 0, s>>>(P, bc1, bc2, m, d);
    PROF_END(s);
}

template void launch_pressure_rhs<double>(const double*, const double*, const double*, const double*, const double*,
                                          const Metric<double>&, const Dims&, double, double*, double*, double*,
                                          hipStream_t);
template void launch_pressure_rhs<float>(const float*, const float*, const float*, const float*, const float*,
                                         const Metric<double>&, const Dims&, double, double*, double*, double*,
                                         hipStream_t);

}  // namespace chan
