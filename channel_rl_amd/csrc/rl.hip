// rl.hip -- the RL agents' observations and actions on the GPU
// (code/smarties_stat.f90:24-258).  The reference copies the wall planes of
// U and nu_t to the host every step to build its agents' states; here one
// small kernel builds states and rewards on the device (the host receives
// 14 values per agent pair) and a second one turns the agents' actions into
// alpha_x, so the wall-stress hook stays on the GPU.  The reference draws
// its sampling rows and heights with an unseeded RANDOM_NUMBER; the caller
// passes those draws in.
#include "kernels.hpp"
#include "prof.hpp"

namespace chan {

// One thread per (agent, wall).  Agent a (0-based) sits at x face
// a*nxa + 2, z centre 2 (code/smarties_stat.f90:61-108 bottom, :120-159 top).
// draws[(wall*na + a)*2 + 0] < 1/2 selects sampling row 3 else 4, draws[.. + 1]
// is the interpolation weight between that row and the next.
template <typename T>
__global__ void k_rl_observe(const T* U, const T* nut, Metric<double> m, Dims d, double nu, int na, int nxa,
                             double dPdx_ref, const double* draws, double* states, double* rewards)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= 2 * na) return;
    const int wall = t / na, a = t % na;
    const double* yg = m.yg;
    const int nyg = d.nyg;
    const int ib = a * nxa + 2;  // (i-1)*nx_nagents + 2 with i = a + 1
    auto u = [&](int i, int j, int k) { return (double)U[at(d, i, j, k)]; };
    auto n = [&](int i, int j, int k) { return (double)nut[at(d, i, j, k)]; };
    const int yind = draws[t * 2 + 0] < 0.5 ? 3 : 4;
    const double yr = draws[t * 2 + 1];
    double ut, l, Uloc, dUloc;
    if (wall == 0) {
        ut = sqrt((nu + 0.25 * (n(ib + 1, 2, 2) + n(ib, 2, 2) + n(ib + 1, 1, 2) + n(ib, 1, 2))) *
                  (u(ib, 2, 2) - u(ib, 1, 2)) / (yg[2] - yg[1]));
        l = yr * yg[yind] + (1.0 - yr) * yg[yind + 1];
        auto lin = [&](int i, int k) { return yr * u(i, yind, k) + ((1.0 - yr) * u(i, yind + 1, k)); };
        Uloc = 0.2 * lin(ib, 2) + 0.2 * lin(ib - 1, 2) + 0.2 * lin(ib + 1, 2) + 0.2 * lin(ib, 1) + 0.2 * lin(ib, 3);
        dUloc = 0.2 *
                (u(ib, yind + 1, 2) - u(ib, yind, 2) + u(ib - 1, yind + 1, 2) - u(ib - 1, yind, 2) +
                 u(ib + 1, yind + 1, 2) - u(ib + 1, yind, 2) + u(ib, yind + 1, 1) - u(ib, yind, 1) +
                 u(ib, yind + 1, 3) - u(ib, yind, 3)) /
                (yg[yind + 1] - yg[yind]);
    } else {
        // the reference divides by yg(2)-yg(1) here as at the bottom wall
        ut = sqrt((nu + 0.25 * (n(ib + 1, nyg, 2) + n(ib, nyg, 2) + n(ib + 1, nyg - 1, 2) + n(ib, nyg - 1, 2))) *
                  (u(ib, nyg - 1, 2) - u(ib, nyg, 2)) / (yg[2] - yg[1]));
        const int jh = nyg - yind + 1, jl = nyg - yind;
        l = 2.0 - (yr * yg[jh] + (1.0 - yr) * yg[jl]);
        auto lin = [&](int i, int k) { return yr * u(i, jh, k) + ((1.0 - yr) * u(i, jl, k)); };
        Uloc = 0.2 * lin(ib, 2) + 0.2 * lin(ib - 1, 2) + 0.2 * lin(ib + 1, 2) + 0.2 * lin(ib, 1) + 0.2 * lin(ib, 3);
        dUloc = 0.2 *
                (u(ib, jl, 2) - u(ib, jh, 2) + u(ib - 1, jl, 2) - u(ib - 1, jh, 2) + u(ib + 1, jl, 2) -
                 u(ib + 1, jh, 2) + u(ib, jl, 1) - u(ib, jh, 1) + u(ib, jl, 3) - u(ib, jh, 3)) /
                (yg[jh] - yg[jl]);
    }
    double* s = states + (long)t * 6;
    s[0] = (dUloc * l) / ut;
    s[1] = Uloc / ut - (dUloc * l) / ut * log(l * ut / nu);
    s[2] = Uloc;
    s[3] = l;
    s[4] = ut;
    s[5] = dPdx_ref;
    // code/smarties_stat.f90:161-178 (0.25 is a default-real constant there:
    // exact in both precisions)
    const double sq = sqrt(dPdx_ref);
    double tw;
    if (wall == 0)
        tw = (nu + 0.25 * (n(ib, 1, 2) + n(ib + 1, 1, 2) + n(ib, 2, 2) + n(ib + 1, 2, 2))) *
             (u(ib, 2, 2) - u(ib, 1, 2)) / (yg[2] - yg[1]);
    else
        tw = (nu + 0.25 * (n(ib, nyg - 1, 2) + n(ib + 1, nyg - 1, 2) + n(ib, nyg, 2) + n(ib + 1, nyg, 2))) *
             (u(ib, nyg - 1, 2) - u(ib, nyg, 2)) / (yg[nyg] - yg[nyg - 1]);
    rewards[t] = 100.0 - 5000.0 * fabs(sqrt(tw) - sq) / sq;
}

// code/smarties_stat.f90:231-258: each agent's action (a factor on its own
// alpha_x, :190/:213) interpolated linearly in x between neighbouring agents
// (periodic), x-periodic images, then the z interpolation between plane 2
// and plane nzg, evaluated as the reference's expressions are:
// (a * real(m)) / real(n) in double.
__global__ void k_rl_actions(const double* alpha, const double* factors, Dims d, int na, int nxa, double* acts)
{
    for (int q = threadIdx.x; q < 2 * na; q += blockDim.x) {
        const int wall = q / na, a = q % na;
        acts[q] = alpha[(a * nxa + 2 - 1) + (long)d.nx * (wall + 2L * (2 - 1))] * factors[q];
    }
}

__global__ void k_rl_fill(double* alpha, const double* acts, Dims d, int na, int nxa)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 1;  // x face
    const int k = blockIdx.y + 1, wall = blockIdx.z;
    if (i > d.nx) return;
    // canonical face for the x images: 1 <- nx-1, nx <- 2
    const int ic = i == 1 ? d.nx - 1 : (i == d.nx ? 2 : i);
    const int r = ic - 2;  // a*nxa + (j-1)
    const int a = r / nxa, j = r % nxa + 1;
    const int an = a + 1 < na ? a + 1 : 0;
    const double v = (acts[wall * na + a] * (double)(nxa - j + 1)) / (double)nxa +
                     (acts[wall * na + an] * (double)(j - 1)) / (double)nxa;
    // z: planes 2..nzg-2 (i_z = k-1 = 1..nzp) mix plane 2 and plane nzg,
    // both v after the x fill; planes 1, nzg-1, nzg keep v
    double out = v;
    if (k >= 2 && k <= d.nzg - 2) {
        const int iz = k - 1, nzp = d.nzp;
        out = (v * (double)(nzp - iz + 1)) / (double)nzp + (v * (double)(iz - 1)) / (double)nzp;
    }
    alpha[(i - 1) + (long)d.nx * (wall + 2L * (k - 1))] = out;
}

template <typename T>
void launch_rl_observe(const T* U, const T* nut, const Metric<double>& m, const Dims& d, double nu, int na, int nxa,
                       double dPdx_ref, const double* draws, double* states, double* rewards, hipStream_t s)
{
    k_rl_observe<T><<<cdiv(2 * na, 64), 64, 0, s>>>(U, nut, m, d, nu, na, nxa, dPdx_ref, draws, states, rewards);
}

void launch_rl_actions(double* alpha, const double* factors, const Dims& d, int na, int nxa, double* acts,
                       hipStream_t s)
{
    k_rl_actions<<<1, 256, 0, s>>>(alpha, factors, d, na, nxa, acts);
    dim3 blk(64), grd(cdiv(d.nx, 64), d.nzg, 2);
    k_rl_fill<<<grd, blk, 0, s>>>(alpha, acts, d, na, nxa);
}

template void launch_rl_observe<double>(const double*, const double*, const Metric<double>&, const Dims&, double,
                                        int, int, double, const double*, double*, double*, hipStream_t);
template void launch_rl_observe<float>(const float*, const float*, const Metric<double>&, const Dims&, double, int,
                                       int, double, const double*, double*, double*, hipStream_t);

}  // namespace chan
