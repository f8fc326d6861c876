// poisson.hip -- pseudo-pressure Poisson solve and projection
// (code/projection.f90:20-214).
//
//   x pass : divergence of u* computed on load, two z-rows packed into one
//            complex FFT of length nxp (real-to-complex by the two-for-one
//            split), modes kx = 0..nxp/2 kept.
//   z pass : complex FFT of length nzp per (j, kx), in place.
//   y solve: one thread per (kx,kz) mode, Thomas elimination of the real
//            tridiagonal (Dyy + kxx + kzz) on the complex right-hand side;
//            the 00 mode's first diagonal is scaled by 3/2 as in the
//            reference.
//   inverse: z pass, then complex-to-real x pass scaled by 1/(nxp nzp),
//            written with the periodic images P(nxg-1)=P(2), P(nzg-1)=P(2).
// Spectral layout ph[j-2][kz][kx] keeps kx fastest so that the z pass reads
// contiguous kx runs and the y solve reads one mode per lane, coalesced.
// FFTs are Stockham autosort passes of radix 2/3/4/5 in LDS.
#include "kernels.hpp"
#include "prof.hpp"

namespace chan {

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ double2 cmul(double2 a, double2 b)
{
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
// multiply by sign*i
__device__ __forceinline__ double2 cmuli(double2 a, int sign)
{
    return sign < 0 ? make_double2(a.y, -a.x) : make_double2(-a.y, a.x);
}

__device__ __forceinline__ double2 twid(const FftPlan& pl, int t, int sign)
{
    double2 w = pl.tw[t];
    return sign < 0 ? w : cconj(w);
}

// In-LDS batched Stockham FFT of B transforms of length pl.n.
// NB layout: element (n,b) at n*B + b; otherwise at b*N + n.
// Result is left in `a` (tmp is scratch of the same size).
template <bool NB>
__device__ void block_fft(double2* a, double2* tmp, const FftPlan& pl, int B, int sign)
{
    const int N = pl.n;
    double2* src = a;
    double2* dst = tmp;
    int Ns = 1;
    for (int p = 0; p < pl.nr; p++) {
        const int R = pl.r[p];
        const int NR = N / R;
        const int nbf = NR * B;
        const int tws = N / (Ns * R);
        for (int q = threadIdx.x; q < nbf; q += blockDim.x) {
            int b, jj;
            if (NB) { b = q % B; jj = q / B; } else { jj = q % NR; b = q / NR; }
            const int k = jj % Ns;
            double2 v[5];
            for (int r = 0; r < R; r++) v[r] = NB ? src[(jj + r * NR) * B + b] : src[b * N + jj + r * NR];
            if (Ns > 1)
                for (int r = 1; r < R; r++) v[r] = cmul(v[r], twid(pl, r * k * tws, sign));
            double2 o[5];
            if (R == 2) {
                o[0] = cadd(v[0], v[1]);
                o[1] = csub(v[0], v[1]);
            } else if (R == 4) {
                double2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
                double2 s13 = cadd(v[1], v[3]), d13 = cmuli(csub(v[1], v[3]), sign);
                o[0] = cadd(s02, s13);
                o[1] = cadd(d02, d13);
                o[2] = csub(s02, s13);
                o[3] = csub(d02, d13);
            } else if (R == 3) {
                const double c = -0.5, sn = 0.86602540378443864676;
                double2 s12 = cadd(v[1], v[2]), d12 = csub(v[1], v[2]);
                double2 t1 = make_double2(v[0].x + c * s12.x, v[0].y + c * s12.y);
                double2 t2 = cmuli(make_double2(sn * d12.x, sn * d12.y), sign);
                o[0] = cadd(v[0], s12);
                o[1] = cadd(t1, t2);
                o[2] = csub(t1, t2);
            } else {  // generic radix (5) from the table
                const int st = N / R;
                for (int qq = 0; qq < R; qq++) {
                    double2 acc = v[0];
                    for (int r = 1; r < R; r++) acc = cadd(acc, cmul(v[r], twid(pl, ((r * qq) % R) * st, sign)));
                    o[qq] = acc;
                }
            }
            const int idxD = (jj / Ns) * Ns * R + k;
            for (int r = 0; r < R; r++) {
                const int n = idxD + r * Ns;
                if (NB) dst[n * B + b] = o[r]; else dst[b * N + n] = o[r];
            }
        }
        __syncthreads();
        double2* t = src; src = dst; dst = t;
        Ns *= R;
    }
    if (src != a) {
        for (int q = threadIdx.x; q < N * B; q += blockDim.x) a[q] = src[q];
        __syncthreads();
    }
}

// divergence of the staggered velocity at centre (i,j,k), code/projection.f90:45-53
template <typename T>
__device__ __forceinline__ double div_at(const T* U, const T* V, const T* W, const Metric<double>& m, const Dims& d,
                                         int i, int j, int k)
{
    return ((double)U[at(d, i, j, k)] - (double)U[at(d, i - 1, j, k)]) / (m.x[i] - m.x[i - 1]) +
           ((double)V[at(d, i, j, k)] - (double)V[at(d, i, j - 1, k)]) / (m.y[j] - m.y[j - 1]) +
           ((double)W[at(d, i, j, k)] - (double)W[at(d, i, j, k - 1)]) / (m.z[k] - m.z[k - 1]);
}

// split the packed transform Z = F[a + i b] into F[a], F[b] for kx = 0..nkx-1
__device__ __forceinline__ void store_split(const double2* z, int N, int nkx, double2* outA, double2* outB, int b_,
                                            bool valid)
{
    for (int q = threadIdx.x; q < nkx; q += blockDim.x) {
        if (!valid) continue;
        double2 Z = z[b_ * N + q];
        double2 Zc = cconj(z[b_ * N + ((N - q) % N)]);
        double2 s = cadd(Z, Zc), df = csub(Z, Zc);
        outA[q] = make_double2(0.5 * s.x, 0.5 * s.y);
        outB[q] = make_double2(0.5 * df.y, -0.5 * df.x);
    }
}

template <typename T, bool FROM_FIELD>
__global__ void k_fftx_fwd(const T* U, const T* V, const T* W, const double* rhs, Metric<double> m, Dims d,
                           PoissonBufs pb)
{
    extern __shared__ double2 lds[];
    const int N = d.nxp, B = pb.bx;
    double2* a = lds;
    double2* tmp = lds + (long)N * B;
    const int npairs = (d.ny - 1) * (d.nzp / 2);
    const int pair0 = blockIdx.x * B;
    for (int q = threadIdx.x; q < N * B; q += blockDim.x) {
        const int b = q / N, n = q % N;
        const int pr = pair0 + b;
        double2 val = make_double2(0.0, 0.0);
        if (pr < npairs) {
            const int j = pr / (d.nzp / 2) + 2;
            const int kA = 2 * (pr % (d.nzp / 2)) + 2;
            const int i = n + 2;
            if (FROM_FIELD) {
                val = make_double2(rhs[at(d, i, j, kA)], rhs[at(d, i, j, kA + 1)]);
            } else {
                val = make_double2(div_at<T>(U, V, W, m, d, i, j, kA), div_at<T>(U, V, W, m, d, i, j, kA + 1));
            }
        }
        a[q] = val;
    }
    __syncthreads();
    block_fft<false>(a, tmp, pb.fx, B, -1);
    for (int b = 0; b < B; b++) {
        const int pr = pair0 + b;
        if (pr >= npairs) break;
        const int jr = pr / (d.nzp / 2);
        const int kA = 2 * (pr % (d.nzp / 2));
        double2* outA = pb.ph + ((long)jr * d.nzp + kA) * d.nkx;
        double2* outB = outA + d.nkx;
        store_split(a, N, d.nkx, outA, outB, b, true);
    }
}

template <int SIGN>
__global__ void k_fftz(Dims d, PoissonBufs pb)
{
    extern __shared__ double2 lds[];
    const int N = d.nzp, B = pb.bz;
    double2* a = lds;
    double2* tmp = lds + (long)N * B;
    const int kx0 = blockIdx.x * B;
    const int jr = blockIdx.y;
    double2* base = pb.ph + (long)jr * d.nzp * d.nkx;
    for (int q = threadIdx.x; q < N * B; q += blockDim.x) {
        const int b = q % B, n = q / B;
        const int kx = kx0 + b;
        a[q] = kx < d.nkx ? base[(long)n * d.nkx + kx] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    block_fft<true>(a, tmp, pb.fz, B, SIGN);
    for (int q = threadIdx.x; q < N * B; q += blockDim.x) {
        const int b = q % B, n = q / B;
        const int kx = kx0 + b;
        if (kx < d.nkx) base[(long)n * d.nkx + kx] = a[q];
    }
}

// code/projection.f90:90-113: (Dyy + kxx + kzz) p = r per mode, Thomas
// elimination (the matrix is real and weakly diagonally dominant).
static constexpr int TRI_THREADS = 64;

__global__ void __launch_bounds__(TRI_THREADS) k_tridiag(Dims d, PoissonBufs pb)
{
    extern __shared__ double cp[];  // [row][lane]
    const int lane = threadIdx.x;
    const int mode = blockIdx.x * TRI_THREADS + lane;
    const int nmodes = d.nkx * d.nzp;
    if (mode >= nmodes) return;
    const int kx = mode % d.nkx, kz = mode / d.nkx;
    const int n = d.ny - 1;
    const long plane = (long)nmodes;
    double2* ph = pb.ph + mode;
    const double lx = pb.kxx[kx], lz = pb.kzz[kz];
    double den = (pb.dd[0] + lx) + lz;
    if (kx == 0 && kz == 0) den = 3.0 / 2.0 * den;
    double inv = 1.0 / den;
    double c = pb.du[0] * inv;
    cp[lane] = c;
    double2 r = ph[0];
    double2 dp = make_double2(r.x * inv, r.y * inv);
    ph[0] = dp;
    for (int row = 1; row < n; row++) {
        const double l = pb.dl[row - 1];
        den = ((pb.dd[row] + lx) + lz) - l * c;
        inv = 1.0 / den;
        r = ph[row * plane];
        dp = make_double2((r.x - l * dp.x) * inv, (r.y - l * dp.y) * inv);
        ph[row * plane] = dp;
        if (row < n - 1) {
            c = pb.du[row] * inv;
            cp[row * TRI_THREADS + lane] = c;
        }
    }
    double2 xn = dp;
    for (int row = n - 2; row >= 0; row--) {
        const double cr = cp[row * TRI_THREADS + lane];
        double2 dr = ph[row * plane];
        xn = make_double2(dr.x - cr * xn.x, dr.y - cr * xn.y);
        ph[row * plane] = xn;
    }
}

__global__ void k_fftx_inv(Dims d, PoissonBufs pb)
{
    extern __shared__ double2 lds[];
    const int N = d.nxp, B = pb.bx;
    double2* a = lds;
    double2* tmp = lds + (long)N * B;
    const int npairs = (d.ny - 1) * (d.nzp / 2);
    const int pair0 = blockIdx.x * B;
    const int half = N / 2;
    for (int q = threadIdx.x; q < N * B; q += blockDim.x) {
        const int b = q / N, n = q % N;
        const int pr = pair0 + b;
        double2 z = make_double2(0.0, 0.0);
        if (pr < npairs) {
            const int jr = pr / (d.nzp / 2);
            const int kA = 2 * (pr % (d.nzp / 2));
            const double2* inA = pb.ph + ((long)jr * d.nzp + kA) * d.nkx;
            const double2* inB = inA + d.nkx;
            double2 xa, xb;
            if (n <= half) {
                xa = inA[n]; xb = inB[n];
                if (n == 0 || n == half) { xa.y = 0.0; xb.y = 0.0; }
            } else {
                xa = cconj(inA[N - n]); xb = cconj(inB[N - n]);
            }
            z = make_double2(xa.x - xb.y, xa.y + xb.x);
        }
        a[q] = z;
    }
    __syncthreads();
    block_fft<false>(a, tmp, pb.fx, B, +1);
    const double scale = (double)((long)d.nxp * d.nzp);
    for (int q = threadIdx.x; q < N * B; q += blockDim.x) {
        const int b = q / N, n = q % N;
        const int pr = pair0 + b;
        if (pr >= npairs) continue;
        const int j = pr / (d.nzp / 2) + 2;
        const int kA = 2 * (pr % (d.nzp / 2)) + 2;
        const double pa = a[q].x / scale, pbv = a[q].y / scale;
        const int i = n + 2;
        pb.p[at(d, i, j, kA)] = pa;
        pb.p[at(d, i, j, kA + 1)] = pbv;
        if (n == 0) {
            pb.p[at(d, d.nxg - 1, j, kA)] = pa;
            pb.p[at(d, d.nxg - 1, j, kA + 1)] = pbv;
        }
        if (kA == 2) {
            pb.p[at(d, i, j, d.nzg - 1)] = pa;
            if (n == 0) pb.p[at(d, d.nxg - 1, j, d.nzg - 1)] = pa;
        }
    }
}

static size_t fft_lds(int N, int B) { return (size_t)2 * N * B * sizeof(double2); }

template <typename T>
void launch_poisson_forward(const T* U, const T* V, const T* W, const Metric<double>& md, const Dims& d,
                            PoissonBufs& pb, hipStream_t s)
{
    const int npairs = (d.ny - 1) * (d.nzp / 2);
    PROF_BEGIN("k_fftx_fwd", s);
    k_fftx_fwd<T, false><<<cdiv(npairs, pb.bx), 256, fft_lds(d.nxp, pb.bx), s>>>(U, V, W, nullptr, md, d, pb);
    PROF_END(s);
    PROF_BEGIN("k_fftz", s);
    k_fftz<-1><<<dim3(cdiv(d.nkx, pb.bz), d.ny - 1), 256, fft_lds(d.nzp, pb.bz), s>>>(d, pb);
    PROF_END(s);
}

void launch_poisson_forward_field(const double* rhs, const Dims& d, PoissonBufs& pb, hipStream_t s)
{
    const int npairs = (d.ny - 1) * (d.nzp / 2);
    Metric<double> none{};
    PROF_BEGIN("k_fftx_fwd", s);
    k_fftx_fwd<double, true><<<cdiv(npairs, pb.bx), 256, fft_lds(d.nxp, pb.bx), s>>>(nullptr, nullptr, nullptr, rhs,
                                                                                       none, d, pb);
    PROF_END(s);
    PROF_BEGIN("k_fftz", s);
    k_fftz<-1><<<dim3(cdiv(d.nkx, pb.bz), d.ny - 1), 256, fft_lds(d.nzp, pb.bz), s>>>(d, pb);
    PROF_END(s);
}

void launch_poisson_solve(const Dims& d, PoissonBufs& pb, hipStream_t s)
{
    const int nmodes = d.nkx * d.nzp;
    PROF_BEGIN("k_tridiag", s);
    k_tridiag<<<cdiv(nmodes, TRI_THREADS), TRI_THREADS, (size_t)(d.ny - 1) * TRI_THREADS * sizeof(double), s>>>(d,
                                                                                                              pb);
    PROF_END(s);
}

void launch_poisson_inverse(const Dims& d, PoissonBufs& pb, hipStream_t s)
{
    PROF_BEGIN("k_fftz", s);
    k_fftz<+1><<<dim3(cdiv(d.nkx, pb.bz), d.ny - 1), 256, fft_lds(d.nzp, pb.bz), s>>>(d, pb);
    PROF_END(s);
    const int npairs = (d.ny - 1) * (d.nzp / 2);
    PROF_BEGIN("k_fftx_inv", s);
    k_fftx_inv<<<cdiv(npairs, pb.bx), 256, fft_lds(d.nxp, pb.bx), s>>>(d, pb);
    PROF_END(s);
}

// code/projection.f90:192-214: u -= grad p on the interior.
template <typename T>
__global__ void __launch_bounds__(256) k_project(T* U, T* V, T* W, const double* __restrict__ p, Metric<double> m,
                                                  Dims d)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x + 2;
    const int j = blockIdx.y * blockDim.y + threadIdx.y + 2;
    const int k = blockIdx.z + 2;
    if (i > d.nxg - 1 || j > d.nyg - 1) return;
    const long q = at(d, i, j, k);
    const double p0 = p[q];
    if (i <= d.nx - 1 && k <= d.nzg - 1)
        U[q] = T((double)U[q] - (p[at(d, i + 1, j, k)] - p0) / (m.xg[i + 1] - m.xg[i]));
    if (j <= d.ny - 1 && k <= d.nzg - 1)
        V[q] = T((double)V[q] - (p[at(d, i, j + 1, k)] - p0) / (m.yg[j + 1] - m.yg[j]));
    if (k <= d.nz - 1) W[q] = T((double)W[q] - (p[at(d, i, j, k + 1)] - p0) / (m.zg[k + 1] - m.zg[k]));
}

template <typename T>
void launch_project(T* U, T* V, T* W, const double* p, const Metric<double>& md, const Dims& d, hipStream_t s)
{
    dim3 blk(64, 4, 1), grd(cdiv(d.nxg - 2, 64), cdiv(d.nyg - 2, 4), d.nzg - 2);
    PROF_BEGIN("k_project", s);
    k_project<T><<<grd, blk, 0, s>>>(U, V, W, p, md, d);
    PROF_END(s);
    launch_bc<T>(U, d, KIND_U, s);
    launch_bc<T>(V, d, KIND_V, s);
    launch_bc<T>(W, d, KIND_W, s);
}

template void launch_poisson_forward<double>(const double*, const double*, const double*, const Metric<double>&,
                                             const Dims&, PoissonBufs&, hipStream_t);
template void launch_poisson_forward<float>(const float*, const float*, const float*, const Metric<double>&,
                                            const Dims&, PoissonBufs&, hipStream_t);
template void launch_project<double>(double*, double*, double*, const double*, const Metric<double>&, const Dims&,
                                     hipStream_t);
template void launch_project<float>(float*, float*, float*, const double*, const Metric<double>&, const Dims&,
                                    hipStream_t);

}  // namespace chan
