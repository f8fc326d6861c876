// poisson.hip -- pseudo-pressure Poisson solve and projection
// (code/projection.f90:20-214).
//
//   x pass : divergence of u* computed on load, two z-rows packed into one
//            complex FFT of length nxp (real-to-complex by the two-for-one
//            split), modes kx = 0..nxp/2 kept.
//   z pass : complex FFT of length nzp per (j, kx) line, in place.
//   y solve: Thomas elimination of the real tridiagonal (Dyy + kxx + kzz)
//            per mode, real and imaginary parts on separate lanes; the 00
//            mode's first diagonal is scaled by 3/2 as in the reference.
//   inverse: z pass, then complex-to-real x pass scaled by 1/(nxp nzp),
//            written with the periodic images P(nxg-1)=P(2), P(nzg-1)=P(2).
// Spectral layout ph[j-2][kx][kz]: the z pass reads whole contiguous lines,
// the y solve reads one mode per lane (kz fastest), the x passes touch
// 2B-long kz runs.  FFTs are in-place Stockham passes of radix 2/3/4/5 in
// LDS, staged through registers, with the twiddle table copied to LDS.
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

static constexpr int FFT_THREADS = 256;
static constexpr int FFT_MAXNB = 2048;  // B*N per block

// R-point DFT in registers (exp(sign 2 pi i / R) convention)
template <int R>
__device__ __forceinline__ void dft(double2 (&v)[R], const double2* tw, int N, int sign)
{
    if constexpr (R == 2) {
        double2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 4) {
        double2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
        double2 s13 = cadd(v[1], v[3]), d13 = cmuli(csub(v[1], v[3]), sign);
        v[0] = cadd(s02, s13);
        v[1] = cadd(d02, d13);
        v[2] = csub(s02, s13);
        v[3] = csub(d02, d13);
    } else if constexpr (R == 3) {
        const double c = -0.5, sn = 0.86602540378443864676;
        double2 s12 = cadd(v[1], v[2]), d12 = csub(v[1], v[2]);
        double2 t1 = make_double2(v[0].x + c * s12.x, v[0].y + c * s12.y);
        double2 t2 = cmuli(make_double2(sn * d12.x, sn * d12.y), sign);
        v[0] = cadd(v[0], s12);
        v[1] = cadd(t1, t2);
        v[2] = csub(t1, t2);
    } else {  // generic (5) from the table
        double2 o[R];
        const int st = N / R;
#pragma unroll
        for (int qq = 0; qq < R; qq++) {
            double2 acc = v[0];
#pragma unroll
            for (int r = 1; r < R; r++) {
                double2 w = tw[((r * qq) % R) * st];
                acc = cadd(acc, cmul(v[r], sign < 0 ? w : cconj(w)));
            }
            o[qq] = acc;
        }
#pragma unroll
        for (int r = 0; r < R; r++) v[r] = o[r];
    }
}

// Quotient by a block-uniform divisor 1 <= dv <= 2^16 for 0 <= q < 2^16: one
// high multiply by m = floor((2^32-1)/dv) + 1 (exact in that range; checked
// exhaustively for dv <= 2048 on the host).  Replaces the ~40-instruction
// integer division in the butterfly index arithmetic.
struct Quo16 {
    unsigned m;
    int dv;
};
__device__ __forceinline__ Quo16 quo16(int dv) { return Quo16{dv > 1 ? 0xFFFFFFFFu / (unsigned)dv + 1u : 0u, dv}; }
__device__ __forceinline__ int quo(int q, const Quo16& v) { return v.dv > 1 ? (int)__umulhi((unsigned)q, v.m) : q; }

// One in-place Stockham pass of radix R over B transforms (element (n,b) at
// b*ld + n).  Every thread loads all of its butterflies' inputs, the block
// synchronises, then the outputs are stored: no second buffer is needed.
template <int R>
__device__ __forceinline__ void fft_pass(double2* a, const double2* tw, int N, int B, int ld, int Ns, int sign)
{
    constexpr int MB = (FFT_MAXNB / R + FFT_THREADS - 1) / FFT_THREADS;
    const int NR = N / R, nbf = NR * B, tws = N / (Ns * R);
    const Quo16 qNR = quo16(NR), qNs = quo16(Ns);
    double2 v[MB][R];
    int jjv[MB], bv[MB];
#pragma unroll
    for (int t = 0; t < MB; t++) {
        const int q = threadIdx.x + t * FFT_THREADS;
        bv[t] = quo(q, qNR);
        jjv[t] = q - bv[t] * NR;
        if (q < nbf) {
#pragma unroll
            for (int r = 0; r < R; r++) v[t][r] = a[bv[t] * ld + jjv[t] + r * NR];
        }
    }
    __syncthreads();
#pragma unroll
    for (int t = 0; t < MB; t++) {
        const int q = threadIdx.x + t * FFT_THREADS;
        if (q < nbf) {
            const int jj = jjv[t], b = bv[t];
            const int js = quo(jj, qNs);
            const int k = jj - js * Ns;
            if (Ns > 1) {
#pragma unroll
                for (int r = 1; r < R; r++) {
                    double2 w = tw[r * k * tws];
                    v[t][r] = cmul(v[t][r], sign < 0 ? w : cconj(w));
                }
            }
            dft<R>(v[t], tw, N, sign);
            const int idxD = js * Ns * R + k;
#pragma unroll
            for (int r = 0; r < R; r++) a[b * ld + idxD + r * Ns] = v[t][r];
        }
    }
    __syncthreads();
}

// In-place batched Stockham FFT of B transforms of length pl.n held in `a`
// (element (n,b) at b*ld + n), twiddles tw[t] = exp(-2 pi i t/N) in LDS:
// passes p0..p1-1 of the plan (all by default).
__device__ void block_fft(double2* a, const double2* tw, const FftPlan& pl, int B, int ld, int sign, int p0 = 0,
                          int p1 = -1)
{
    if (p1 < 0) p1 = pl.nr;
    int Ns = 1;
    for (int p = 0; p < p0; p++) Ns *= pl.r[p];
    for (int p = p0; p < p1; p++) {
        const int R = pl.r[p];
        if (R == 4) fft_pass<4>(a, tw, pl.n, B, ld, Ns, sign);
        else if (R == 2) fft_pass<2>(a, tw, pl.n, B, ld, Ns, sign);
        else if (R == 3) fft_pass<3>(a, tw, pl.n, B, ld, Ns, sign);
        else fft_pass<5>(a, tw, pl.n, B, ld, Ns, sign);
        Ns *= R;
    }
}

__device__ __forceinline__ void load_twiddles(double2* tw, const double2* src, int n)
{
    for (int q = threadIdx.x; q < n; q += blockDim.x) tw[q] = src[q];
}

// pair pr -> (row jr = j-2, first z index kA-2) of the two packed z-rows
__device__ __forceinline__ void pair_rows(const Dims& d, int pr, int& jr, int& kz)
{
    jr = pr / (d.nzp / 2);
    kz = 2 * (pr % (d.nzp / 2));
}

// The block's pairs of packed z rows, per pair b < B <= 16: the box index
// of x centre 1 on the first row (-1 past the last pair), the spectrum
// offset of its (jr, kx = 0, kz) and the y and z reciprocal spacings, taken
// once per block instead of per element (a pair index takes an integer
// division, a box index 64-bit products).  Filled by the first B threads;
// the caller synchronises.
struct PairTab {
    long box[16], spec[16];
    double rdy[16], rdz0[16], rdz1[16];
    bool zfirst[16];  // the pair starts at the first unique z centre
};
__device__ __forceinline__ void fill_pairs(PairTab& t, const Dims& d, const Metric<double>* m, int pair0, int B,
                                           int npairs)
{
    const int b = threadIdx.x;
    if (b >= B) return;
    const int pr = min(pair0 + b, npairs - 1);
    int jr, kz;
    pair_rows(d, pr, jr, kz);
    const int j = jr + 2, kA = kz + 2;
    t.box[b] = pair0 + b < npairs ? at(d, 1, j, kA) : -1L;
    t.spec[b] = (long)jr * d.nkx * d.nzp + kz;
    t.zfirst[b] = kA == 2;
    if (m) {
        t.rdy[b] = m->rdy[j];
        t.rdz0[b] = m->rdz[kA];
        t.rdz1[b] = m->rdz[kA + 1];
    }
}

template <typename T, bool FROM_FIELD, bool G>
__global__ void __launch_bounds__(FFT_THREADS) k_fftx_fwd(FwdArgs<T> one, const FwdArgs<T>* tab, const double* rhs,
                                                          Dims d)
{
    // consecutive blocks share V rows (row j-1 of one is row j of a block
    // nzp/2B before it): XCD-contiguous order keeps both in one L2
    const GBlk gb = gblk_xcd(1);
    // the group's plans and batch sizes from the launch's arguments, this
    // member's buffers from its table entry
    const FwdArgs<T> arg = member_args<G>(one, tab, gb.m);
    const T *U = arg.U, *V = arg.V, *W = arg.W;
    const Metric<double>& m = arg.m;
    const PoissonBufs& pb = one.pb;
    double2* const ph = arg.pb.ph;
    extern __shared__ double2 lds[];
    const int N = d.nxp, B = pb.bx, ld = N + 1;
    const Quo16 qN = quo16(N), qB = quo16(B);
    double2* tw = lds;
    double2* a = lds + N;
    const int npairs = (d.ny - 1) * (d.nzp / 2);
    const int pair0 = gb.x * B;
    __shared__ PairTab pt;
    fill_pairs(pt, d, FROM_FIELD ? nullptr : &m, pair0, B, npairs);
    load_twiddles(tw, arg.pb.fx.tw, N);
    __syncthreads();
    const long sx = d.sx, sxy = d.sxy;
    for (int q = threadIdx.x; q < N * B; q += blockDim.x) {
        const int b = quo(q, qN), n = q - b * N;
        const long base = pt.box[b];
        double2 val = make_double2(0.0, 0.0);
        if (base >= 0) {
            // the divergence of the staggered velocity at centres (n+2, j, kA)
            // and (n+2, j, kA+1), code/projection.f90:45-53, with the
            // spacings' reciprocals (within an ulp of the reference's divisions)
            const long p = base + n + 1, p1 = p + sxy;
            if (FROM_FIELD) {
                val = make_double2(rhs[p], rhs[p1]);
            } else {
                const double rx = m.rdx[n + 2], ry = pt.rdy[b];
                val.x = (((double)U[p] - (double)U[p - 1]) * rx + ((double)V[p] - (double)V[p - sx]) * ry) +
                        ((double)W[p] - (double)W[p - sxy]) * pt.rdz0[b];
                val.y = (((double)U[p1] - (double)U[p1 - 1]) * rx + ((double)V[p1] - (double)V[p1 - sx]) * ry) +
                        ((double)W[p1] - (double)W[p]) * pt.rdz1[b];
            }
        }
        a[b * ld + n] = val;
    }
    __syncthreads();
    block_fft(a, tw, pb.fx, B, ld, -1);
    // split Z = F[a + i b] into F[a], F[b]; consecutive lanes write consecutive kz
    for (int q = threadIdx.x; q < d.nkx * B; q += blockDim.x) {
        const int kx = quo(q, qB), b = q - kx * B;
        if (pt.box[b] < 0) continue;
        const double2 Z = a[b * ld + kx];
        const double2 Zc = cconj(a[b * ld + ((N - kx) % N)]);
        const double2 s = cadd(Z, Zc), df = csub(Z, Zc);
        double2* out = ph + pt.spec[b] + (long)kx * d.nzp;
        out[0] = make_double2(0.5 * s.x, 0.5 * s.y);
        out[1] = make_double2(0.5 * df.y, -0.5 * df.x);
    }
}

// First Stockham pass of radix R (span 1) on values produced in registers:
// butterfly jj of transform b takes elements n = jj + r N/R, each from
// elem(valid, b, n) (a global load, or a value computed from loads), and
// leaves its outputs in LDS where the pass over LDS would put them.  Every
// load is issued before the first DFT.  BFAST: consecutive lanes take
// consecutive transforms b (else consecutive jj), whichever makes the
// elements' addresses consecutive.
template <int R, int SIGN, bool BFAST, typename F>
__device__ __forceinline__ void fft_first_pass(double2* a, const double2* tw, int N, int B, int ld, const F& elem)
{
    constexpr int MB = (FFT_MAXNB / R + FFT_THREADS - 1) / FFT_THREADS;
    const int NR = N / R, nbf = NR * B;
    const Quo16 qd = quo16(BFAST ? B : NR);
    double2 v[MB][R];
    // lanes past the last butterfly (q >= nbf) load the elements of
    // butterfly 0 and drop them: with BFAST their jj would pass N/R, their
    // n pass N
    auto split = [&](int q, int& b, int& jj) {
        if (q >= nbf) q = 0;
        if (BFAST) {
            jj = quo(q, qd);
            b = q - jj * B;
        } else {
            b = quo(q, qd);
            jj = q - b * NR;
        }
    };
#pragma unroll
    for (int t = 0; t < MB; t++) {
        const int q = threadIdx.x + t * FFT_THREADS;
        int b, jj;
        split(q, b, jj);
#pragma unroll
        for (int r = 0; r < R; r++) v[t][r] = elem(q < nbf, b, jj + r * NR);
    }
    if constexpr (R == 5) __syncthreads();  // the radix-5 DFT reads the twiddle table
#pragma unroll
    for (int t = 0; t < MB; t++) {
        const int q = threadIdx.x + t * FFT_THREADS;
        if (q < nbf) {
            int b, jj;
            split(q, b, jj);
            dft<R>(v[t], tw, N, SIGN);
#pragma unroll
            for (int r = 0; r < R; r++) a[b * ld + jj * R + r] = v[t][r];
        }
    }
}

// Last Stockham pass of radix R (span N/R) from LDS to registers: output n
// = k + r N/R of transform b goes to store(b, n, value); consecutive lanes
// take consecutive k.
template <int R, int SIGN, typename S>
__device__ __forceinline__ void fft_last_pass(const double2* a, const double2* tw, int N, int B, int ld, const S& store)
{
    constexpr int MB = (FFT_MAXNB / R + FFT_THREADS - 1) / FFT_THREADS;
    const int NR = N / R, nbf = NR * B;
    const Quo16 qNR = quo16(NR);
    double2 v[MB][R];
#pragma unroll
    for (int t = 0; t < MB; t++) {
        const int q = threadIdx.x + t * FFT_THREADS;
        const int b = quo(q, qNR), k = q - b * NR;
        if (q < nbf) {
#pragma unroll
            for (int r = 0; r < R; r++) v[t][r] = a[b * ld + k + r * NR];
        }
    }
#pragma unroll
    for (int t = 0; t < MB; t++) {
        const int q = threadIdx.x + t * FFT_THREADS;
        const int b = quo(q, qNR), k = q - b * NR;
        if (q < nbf) {
#pragma unroll
            for (int r = 1; r < R; r++) {
                const double2 w = tw[r * k];
                v[t][r] = cmul(v[t][r], SIGN < 0 ? w : cconj(w));
            }
            dft<R>(v[t], tw, N, SIGN);
#pragma unroll
            for (int r = 0; r < R; r++) store(b, k + r * NR, v[t][r]);
        }
    }
}

// the passes above for the plan's radix R (block-uniform)
template <int SIGN, bool BFAST, typename F>
__device__ __forceinline__ void fft_first_pass_r(int R, double2* a, const double2* tw, int N, int B, int ld,
                                                 const F& elem)
{
    if (R == 4) fft_first_pass<4, SIGN, BFAST>(a, tw, N, B, ld, elem);
    else if (R == 2) fft_first_pass<2, SIGN, BFAST>(a, tw, N, B, ld, elem);
    else if (R == 3) fft_first_pass<3, SIGN, BFAST>(a, tw, N, B, ld, elem);
    else fft_first_pass<5, SIGN, BFAST>(a, tw, N, B, ld, elem);
}
template <int SIGN, typename S>
__device__ __forceinline__ void fft_last_pass_r(int R, const double2* a, const double2* tw, int N, int B, int ld,
                                                const S& store)
{
    if (R == 4) fft_last_pass<4, SIGN>(a, tw, N, B, ld, store);
    else if (R == 2) fft_last_pass<2, SIGN>(a, tw, N, B, ld, store);
    else if (R == 3) fft_last_pass<3, SIGN>(a, tw, N, B, ld, store);
    else fft_last_pass<5, SIGN>(a, tw, N, B, ld, store);
}

template <int SIGN, bool G>
__global__ void __launch_bounds__(FFT_THREADS) k_fftz(PoissonBufs one, const PoissonBufs* tab, Dims d)
{
    const GBlk gb = gblk(1);
    const PoissonBufs& pb = one;  // the group's plans and batch sizes
    const PoissonBufs mb = member_args<G>(one, tab, gb.m);  // this member's buffers
    extern __shared__ double2 lds[];
    const FftPlan& pl = pb.fz;
    const int N = d.nzp, B = pb.bz, ld = N + 1;
    double2* tw = lds;
    double2* a = lds + N;
    const long nlines = (long)(d.ny - 1) * d.nkx;
    const long line0 = (long)gb.x * B;
    const long nrows = nlines - line0;  // lines of this block that exist (>= 1)
    double2* ln = mb.ph + line0 * N;
    load_twiddles(tw, mb.fz.tw, N);
    if (pl.nr >= 2) {
        // the first and the last pass run between registers and global
        // memory, the others in LDS: two LDS sweeps and four barriers fewer
        // than loading into LDS first and storing from it last
        fft_first_pass_r<SIGN, false>(pl.r[0], a, tw, N, B, ld, [&](bool ok, int b, int n) {
            return ok && b < nrows ? ln[(long)b * N + n] : make_double2(0.0, 0.0);
        });
        __syncthreads();
        block_fft(a, tw, pl, B, ld, SIGN, 1, pl.nr - 1);
        fft_last_pass_r<SIGN>(pl.r[pl.nr - 1], a, tw, N, B, ld, [&](int b, int n, double2 val) {
            if (b < nrows) ln[(long)b * N + n] = val;
        });
        return;
    }
    // a single pass (N <= 5): through LDS
    const Quo16 qN = quo16(N);
    for (int q = threadIdx.x; q < N * B; q += blockDim.x) {
        const int b = quo(q, qN), n = q - b * N;
        a[b * ld + n] = b < nrows ? ln[(long)b * N + n] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    block_fft(a, tw, pl, B, ld, SIGN);
    for (int q = threadIdx.x; q < N * B; q += blockDim.x) {
        const int b = quo(q, qN), n = q - b * N;
        if (b < nrows) ln[(long)b * N + n] = a[b * ld + n];
    }
}

// code/projection.f90:90-113: (Dyy + kxx + kzz) p = r per (kx, kz) mode by
// the Thomas algorithm, real and imaginary parts on neighbouring lanes.  The
// matrix (with the 00 mode's first row scaled by 3/2, code/projection.f90:107)
// is real, so the elimination's pivots 1/den_j depend on the mode only, and
// the modified upper diagonal is c'_j = Dyy_j,j+1 / den_j.
//
// Systems of at most TRI_KEEP rows (k_tridiag_tw) live in registers, half a
// system per lane: the right-hand sides are loaded first, all in flight at
// once; the pivot chain (one division per row, no global memory) runs while
// they arrive and keeps every 1/den_j in LDS; the sweeps then touch registers
// and LDS only and each row is written once.  A wave can hold few loads in
// flight, so a pivot table in memory would cost more than the divisions.
// Taller systems (k_tridiag<true, 32>) keep their last 32 rows in registers
// and stream the rows before them through memory, with pivots from a
// [row][mode] table that k_tridiag_setup computes once per run by the same
// arithmetic.  (k_tridiag<false, ..>, the one-ended register solve, is the
// form the two-ended kernel replaced; it is not instantiated.)
static constexpr int TRI_THREADS = 64;
static constexpr int TRI_GROUP = 8;
static constexpr int TRI_KEEP = 192;

__global__ void __launch_bounds__(TRI_THREADS) k_tridiag_setup(Dims d, PoissonBufs pb)
{
    const int nmodes = d.nkx * d.nzp;
    const int mode = blockIdx.x * blockDim.x + threadIdx.x;
    if (mode >= nmodes) return;
    const int kx = mode / d.nzp, kz = mode % d.nzp;
    const int n = d.ny - 1;
    const double lx = pb.kxx[kx], lz = pb.kzz[kz];
    double den = (pb.dd[0] + lx) + lz;
    if (kx == 0 && kz == 0) den = 3.0 / 2.0 * den;
    double inv = 1.0 / den;
    double c = pb.du[0] * inv;
    pb.tri[mode] = inv;
    for (int row = 1; row < n; row++) {
        den = ((pb.dd[row] + lx) + lz) - pb.dl[row - 1] * c;
        inv = 1.0 / den;
        c = row < n - 1 ? pb.du[row] * inv : 0.0;
        pb.tri[(long)row * nmodes + mode] = inv;
    }
}

// one lane per (mode, real/imaginary part); modes are kx-major, kz fastest
// Systems of at most 2H rows, two-ended: the top half of each system is
// eliminated downwards and the bottom half upwards at the same time, on
// neighbouring lanes running one instruction stream, which halves the
// dependent chain and doubles the waves (four lanes per mode: real and
// imaginary part, top and bottom).  Step u of a lane is row u (top) or row
// n-1-u (bottom); with the couplings A (towards the rows already
// eliminated) and C (towards the rest) both halves run
//   den = (D + kxx + kzz) - A c',  c' = C / den,  y = (r - A y') / den,
// the top half over rows 0..hn-1 (hn = ceil(n/2)), the bottom over rows
// n-1..hn.  The two meet in the 2x2 system of rows hn-1 and hn:
//   x(hn-1) + c'_top x(hn) = y_top,  x(hn) + c'_bot x(hn-1) = y_bot,
// which both lanes of a pair solve alike; each half then substitutes back
// towards its end.  Pivots stay in LDS, rows in registers.  The unrolled
// steps run without branches: a lane's rows are reached by a pointer that
// moves by one row per step, the matrix entries sit in LDS in each half's
// step order, and the steps past a half's end (odd n) keep the chain's
// values by selection and re-store the half's last row with its own value.
// (An empty asm on each new value keeps the compiler from moving its
// computation into a branch of the selection.)
// Register budget: two waves per SIMD (the 96-row half of a 192-row system
// holds 192 VGPRs of rows; left free, the compiler takes 450 and one wave).
template <int H, bool G>
__global__ void __launch_bounds__(TRI_THREADS) __attribute__((amdgpu_waves_per_eu(2)))
    k_tridiag_tw(PoissonBufs one, const PoissonBufs* tab, Dims d)
{
    const GBlk gb = gblk(1);
    const PoissonBufs pb = member_args<G>(one, tab, gb.m);
    const long g = (long)gb.x * TRI_THREADS + threadIdx.x;
    const int nmodes = d.nkx * d.nzp;
    const int n = d.ny - 1;
    const int hn = (n + 1) / 2;
    const bool bot = (g & 1) != 0;
    const int hl = bot ? n - hn : hn;  // this lane's rows
    const int mode = (int)(min(g, 4L * nmodes - 1) >> 2);
    const int comp = (int)((g >> 1) & 1);
    const int kx = mode / d.nzp, kz = mode - kx * d.nzp;
    const double lx = pb.kxx[kx], lz = pb.kzz[kz];
    // step u of the top half: row u, couplings (A, C) = (l_u, u_u); of the
    // bottom half (slot H+u): row n-1-u, (A, C) = (u_r, l_r); an identity
    // row past a half's end
    __shared__ double sD[2 * H], sA[2 * H], sC[2 * H];
    // 1/den per lane; for the 64- and 96-row halves per (mode, half) instead:
    // the real and imaginary lanes of a half (lanes t and t^2) compute the
    // same chain and share a slot (each writes the same value, each reads
    // back what the pair wrote), which halves the LDS that limits those
    // forms' occupancy (for the shorter ones the shared slots measured
    // slower: 19.0 -> 21.3 us at 97 rows)
    constexpr bool SHARE = H >= 64;
    __shared__ double piv[H][SHARE ? TRI_THREADS / 2 : TRI_THREADS];
    const int ps = SHARE ? (((threadIdx.x >> 2) << 1) | (threadIdx.x & 1)) : threadIdx.x;
    for (int q = threadIdx.x; q < 2 * H; q += TRI_THREADS) {
        const bool b = q >= H;
        const int u = b ? q - H : q;
        const int r = b ? n - 1 - u : u;
        const bool in = u < (b ? n - hn : hn);
        const double l = in && r > 0 ? pb.dl[r - 1] : 0.0;
        const double up = in && r < n - 1 ? pb.du[r] : 0.0;
        sD[q] = in ? pb.dd[r] : 1.0;
        sA[q] = b ? up : l;
        sC[q] = b ? l : up;
    }
    __syncthreads();
    if (g >= 4L * nmodes) return;
    const double* D = sD + (bot ? H : 0);
    const double* A = sA + (bot ? H : 0);
    const double* C = sC + (bot ? H : 0);
    const long plane = 2L * nmodes;
    const long step = bot ? -plane : plane;  // one row further into the half
    double* p0 = reinterpret_cast<double*>(pb.ph) + 2L * mode + comp + (bot ? (long)(n - 1) * plane : 0L);
    double* plast = p0 + (long)(hl - 1) * step;
    // the rows of this half, every load issued first (steps past the half
    // re-read its last row and are not used)
    double keep[H];
    {
        const double* pr = p0;
        int hq = hl;
        // for the long halves each loop below compares the steps against its
        // own opaque copy of hl and the stores take a fresh address chain:
        // left free, the compiler keeps the H comparisons of the first loop
        // as scalar masks for all four and the loads' addresses for the
        // stores (2H registers each), which spills the 96-row form
        if constexpr (SHARE) asm("" : "+v"(pr), "+v"(hq));
#pragma unroll
        for (int u = 0; u < H; u++) {
            keep[u] = *pr;
            pr = u + 1 < hq ? pr + step : pr;
        }
    }
    // keeps the compiler from sinking the loads below the chain
    asm volatile("" ::: "memory");
    // pivot chain; row 0's diagonal of the 00 mode scaled by 3/2
    double c = 0.0;
    const bool scale0 = !bot && kx == 0 && kz == 0;
    int hc = hl;
    if constexpr (SHARE) asm("" : "+v"(hc));
#pragma unroll
    for (int u = 0; u < H; u++) {
        double den = ((D[u] + lx) + lz) - A[u] * c;
        if (u == 0) den = scale0 ? 3.0 / 2.0 * den : den;
        const double inv = 1.0 / den;
        piv[u][ps] = inv;
        double cn = C[u] * inv;
        asm("" : "+v"(cn));
        c = u < hc ? cn : c;
    }
    asm volatile("" ::: "memory");
    double y = 0.0;
    int hf = hl;
    if constexpr (SHARE) asm("" : "+v"(hf));
#pragma unroll
    for (int u = 0; u < H; u++) {
        double yn = (keep[u] - A[u] * y) * piv[u][ps];
        asm("" : "+v"(yn));
        y = u < hf ? yn : y;
        keep[u] = y;
    }
    // the two middle rows from the pair's last values (lanes g and g^1)
    const double co = __shfl_xor(c, 1, 64), yo = __shfl_xor(y, 1, 64);
    const double ct = bot ? co : c, yt = bot ? yo : y, cb = bot ? c : co, yb = bot ? y : yo;
    const double xt = (yt - ct * yb) / (1.0 - ct * cb);  // x(hn-1)
    const double xb = yb - cb * xt;                       // x(hn)
    double xn = bot ? xb : xt;
    *plast = xn;
    // back towards the end of the half: x = y - c' x(previous); pw moves
    // with the valid steps only, from row hl-2 of the half
    double* pw = p0 + (long)max(hl - 2, 0) * step;
    int hb = hl - 1;
    if constexpr (SHARE) asm("" : "+v"(pw), "+v"(plast), "+v"(hb));
#pragma unroll
    for (int u = H - 2; u >= 0; u--) {
        double xu = keep[u] - C[u] * piv[u][ps] * xn;
        asm("" : "+v"(xu));
        const bool in = u < hb;
        xn = in ? xu : xn;
        *(in ? pw : plast) = xn;
        if (u > 0) pw = in ? pw - step : pw;
    }
}

// KEEP register rows; a system of more than TRI_KEEP rows keeps its last 32
// in registers (four waves per SIMD) and STREAMS the rows before them
template <bool STREAMED, int KEEP, bool GRP>
__global__ void __launch_bounds__(TRI_THREADS) k_tridiag(PoissonBufs one, const PoissonBufs* tab, Dims d)
{
    const GBlk gb = gblk(1);
    const PoissonBufs pb = member_args<GRP>(one, tab, gb.m);
    const long g = (long)gb.x * TRI_THREADS + threadIdx.x;
    const int nmodes = d.nkx * d.nzp;
    const int n = d.ny - 1;
    const int nk = n < KEEP ? n : KEEP;  // rows n0..n-1 stay in registers
    const int n0 = STREAMED ? n - nk : 0;
    const int mode = (int)(min(g, 2L * nmodes - 1) >> 1);
    const int kx = mode / d.nzp, kz = mode - kx * d.nzp;
    // the mode's wavenumbers before the right-hand sides: loads complete in
    // issue order, so the pivot chain can then run while the rows arrive
    const double lx = STREAMED ? 0.0 : pb.kxx[kx], lz = STREAMED ? 0.0 : pb.kzz[kz];
    // the register rows' matrix entries, read by every lane at the same
    // address; slots past the last row hold an identity row so that the
    // unrolled chains run without branches (their values there are not used)
    __shared__ double sd[KEEP], sl[KEEP], su[KEEP];
    // 1/den of the register rows for the block's modes (both lanes of a mode
    // write the same value; each lane reads back what it wrote)
    __shared__ double piv[KEEP][TRI_THREADS / 2];
    for (int u = threadIdx.x; u < KEEP; u += TRI_THREADS) {
        const int rw = n0 + u;
        sd[u] = u < nk ? pb.dd[rw] : 1.0;
        sl[u] = u < nk && rw > 0 ? pb.dl[rw - 1] : 0.0;
        su[u] = u < nk && rw < n - 1 ? pb.du[rw] : 0.0;
    }
    __syncthreads();
    if (g >= 2L * nmodes) return;
    const long plane = 2L * nmodes;
    double* ph = reinterpret_cast<double*>(pb.ph) + g;
    const double* ti = STREAMED ? pb.tri + mode : nullptr;  // the table exists for tall systems only
    double* pv = &piv[0][threadIdx.x >> 1];
    constexpr int PS = TRI_THREADS / 2;  // row stride of pv
    double keep[KEEP];
    double dp = 0.0;
    int row = 0;
    // streamed rows: groups of TRI_GROUP rows with their table pivots and
    // off-diagonal entries, the next group's loads in flight while one is
    // computed; every load is unconditional (rows clamped into 0..n0-1, the
    // extra slots unused) so that no branch makes the compiler drain the
    // memory counter, which completes in issue order
    constexpr int G = TRI_GROUP;
    double ra[G], va[G], ca[G], rb[G], vb[G], cb[G];
    // dir +1: rows r0.. with l_j (0 on row 0); dir -1: rows r0, r0-1, .. with u_j
    auto load_rows = [&](int r0, int dir, double (&r)[G], double (&v)[G], double (&c)[G]) {
#pragma unroll
        for (int u = 0; u < G; u++) {
            const int rw = min(max(r0 + dir * u, 0), n0 - 1);
            r[u] = ph[(long)rw * plane];
            v[u] = ti[(long)rw * nmodes];
            c[u] = dir > 0 ? (rw > 0 ? pb.dl[rw - 1] : 0.0) : pb.du[rw];
        }
    };
    auto forward = [&](int r0, const double (&r)[G], const double (&v)[G], const double (&l)[G]) {
#pragma unroll
        for (int u = 0; u < G; u++) {
            if (r0 + u < n0) {
                dp = (r[u] - l[u] * dp) * v[u];
                ph[(r0 + u) * plane] = dp;
            }
        }
    };
    // back substitution, rows r0, r0-1, ...: x(j) = dp(j) - c'(j) x(j+1)
    auto backward = [&](int r0, const double (&r)[G], const double (&v)[G], const double (&up)[G], double& x) {
#pragma unroll
        for (int u = 0; u < G; u++) {
            if (r0 - u >= 0) {
                x = r[u] - up[u] * v[u] * x;
                ph[(r0 - u) * plane] = x;
            }
        }
    };
    if (STREAMED) {
        load_rows(0, 1, ra, va, ca);
        for (; row < n0; row += 2 * G) {
            load_rows(row + G, 1, rb, vb, cb);
            forward(row, ra, va, ca);
            load_rows(row + 2 * G, 1, ra, va, ca);
            forward(row + G, rb, vb, cb);
        }
        // the register rows with their table pivots (slots past the last row
        // re-read row n-1 and are not used)
#pragma unroll
        for (int u = 0; u < KEEP; u++) {
            const long rw = min(n0 + u, n - 1);
            keep[u] = ph[rw * plane];
            pv[u * PS] = ti[rw * nmodes];
        }
    } else {
        // right-hand sides first, then the pivot chain while they arrive: the
        // arithmetic of k_tridiag_setup (on the first row the lower term is 0 * 0)
#pragma unroll
        for (int u = 0; u < KEEP; u++) keep[u] = ph[(long)min(u, n - 1) * plane];
        // keeps the compiler from sinking the loads below the chain
        asm volatile("" ::: "memory");
        double c = 0.0;
#pragma unroll
        for (int u = 0; u < KEEP; u++) {
            double den = ((sd[u] + lx) + lz) - sl[u] * c;
            if (u == 0) den = kx == 0 && kz == 0 ? 3.0 / 2.0 * den : den;
            const double inv = 1.0 / den;
            pv[u * PS] = inv;
            c = su[u] * inv;
        }
    }
    // the sweeps read the pivots back from LDS (no forwarding in registers)
    asm volatile("" ::: "memory");
#pragma unroll
    for (int u = 0; u < KEEP; u++) {
        dp = (keep[u] - sl[u] * dp) * pv[u * PS];
        keep[u] = dp;
    }
    // back substitution: x(n-1) = dp(n-1), x(j) = dp(j) - c'(j) x(j+1)
    double xn = 0.0;
#pragma unroll
    for (int u = KEEP - 1; u >= 0; u--) {
        if (u < nk) {
            xn = u == nk - 1 ? keep[u] : keep[u] - su[u] * pv[u * PS] * xn;
            ph[(n0 + u) * plane] = xn;
        }
    }
    if (STREAMED) {
        load_rows(n0 - 1, -1, ra, va, ca);
        for (row = n0 - 1; row >= 0; row -= 2 * G) {
            load_rows(row - G, -1, rb, vb, cb);
            backward(row, ra, va, ca, xn);
            load_rows(row - 2 * G, -1, ra, va, ca);
            backward(row - G, rb, vb, cb, xn);
        }
    }
}

template <bool G>
__global__ void __launch_bounds__(FFT_THREADS) k_fftx_inv(PoissonBufs one, const PoissonBufs* tab, Dims d)
{
    const GBlk gb = gblk(1);
    const PoissonBufs& pb = one;  // the group's plans and batch sizes
    const PoissonBufs mb = member_args<G>(one, tab, gb.m);  // this member's buffers
    double* const P = mb.p;
    extern __shared__ double2 lds[];
    const int N = d.nxp, B = pb.bx, ld = N + 1;
    const Quo16 qN = quo16(N);
    double2* tw = lds;
    double2* a = lds + N;
    const int npairs = (d.ny - 1) * (d.nzp / 2);
    const int pair0 = gb.x * B;
    const int half = N / 2;
    __shared__ PairTab pt;
    fill_pairs(pt, d, nullptr, pair0, B, npairs);
    load_twiddles(tw, mb.fx.tw, N);
    __syncthreads();
    // the unnormalised inverse transforms' factor, as a reciprocal (within an
    // ulp of the division)
    const double rscale = 1.0 / (double)((long)d.nxp * d.nzp);
    // element n of transform b (two packed z rows): mode n of both rows, or
    // the conjugate of mode N-n above the Nyquist mode; imaginary parts of
    // modes 0 and N/2 dropped.  Consecutive lanes read consecutive kz of one
    // kx line (BFAST); the loads unconditional, from a clamped pair.
    auto elem = [&](bool ok, int b, int n) {
        const int mode = n <= half ? n : N - n;
        const double2* in = mb.ph + pt.spec[b] + (long)mode * d.nzp;
        double2 xa = in[0], xb = in[1];
        if (n == 0 || n == half) {
            xa.y = 0.0;
            xb.y = 0.0;
        }
        if (n > half) {
            xa = cconj(xa);
            xb = cconj(xb);
        }
        const double2 val = make_double2(xa.x - xb.y, xa.y + xb.x);
        return ok && pt.box[b] >= 0 ? val : make_double2(0.0, 0.0);
    };
    // value n of transform b to P(n+2) of both rows, with the periodic images
    // P(nxg-1) = P(2) and, for the first z row, P(.., nzg-1) = P(.., 2)
    const long sxy = d.sxy, zimg = (long)(d.nzg - 3) * d.sxy;
    auto store = [&](int b, int n, double2 val) {
        const long base = pt.box[b];
        if (base < 0) return;
        const long p = base + n + 1, pi = base + d.nxg - 2;  // x centres n+2 and nxg-1
        const double pa = val.x * rscale, pbv = val.y * rscale;
        P[p] = pa;
        P[p + sxy] = pbv;
        if (n == 0) {
            P[pi] = pa;
            P[pi + sxy] = pbv;
        }
        if (pt.zfirst[b]) {
            P[p + zimg] = pa;
            if (n == 0) P[pi + zimg] = pa;
        }
    };
    if (pb.fx.nr >= 2) {
        // first and last pass between registers and global memory
        fft_first_pass_r<+1, true>(pb.fx.r[0], a, tw, N, B, ld, elem);
        __syncthreads();
        block_fft(a, tw, pb.fx, B, ld, +1, 1, pb.fx.nr - 1);
        fft_last_pass_r<+1>(pb.fx.r[pb.fx.nr - 1], a, tw, N, B, ld, store);
        return;
    }
    for (int q = threadIdx.x; q < N * B; q += blockDim.x) {
        const int b = quo(q, qN), n = q - b * N;
        a[b * ld + n] = elem(true, b, n);
    }
    __syncthreads();
    block_fft(a, tw, pb.fx, B, ld, +1);
    for (int q = threadIdx.x; q < N * B; q += blockDim.x) {
        const int b = quo(q, qN), n = q - b * N;
        store(b, n, a[b * ld + n]);
    }
}

static size_t fft_lds(int N, int B) { return (size_t)(N + (size_t)(N + 1) * B) * sizeof(double2); }

template <typename T>
void run_poisson_forward(const FwdArgs<T>& one, const FwdArgs<T>* tab, const PoissonBufs* tab_pb, int nmem,
                         const Dims& d, const PoissonBufs& pb0, hipStream_t s)
{
    const int npairs = (d.ny - 1) * (d.nzp / 2);
    PROF_BEGIN("k_fftx_fwd", s);
    if (tab) {
        FwdArgs<T> rep{};
        rep.pb = pb0;
        k_fftx_fwd<T, false, true><<<dim3(cdiv(npairs, pb0.bx), 1, nmem), FFT_THREADS, fft_lds(d.nxp, pb0.bx), s>>>(
            rep, tab, nullptr, d);
    }
    else
        k_fftx_fwd<T, false, false><<<dim3(cdiv(npairs, pb0.bx), 1, 1), FFT_THREADS, fft_lds(d.nxp, pb0.bx), s>>>(
            one, nullptr, nullptr, d);
    PROF_END(s);
    const dim3 gz(cdiv((long)(d.ny - 1) * d.nkx, pb0.bz), 1, nmem);
    PROF_BEGIN("k_fftz", s);
    if (tab) k_fftz<-1, true><<<gz, FFT_THREADS, fft_lds(d.nzp, pb0.bz), s>>>(pb0, tab_pb, d);
    else k_fftz<-1, false><<<gz, FFT_THREADS, fft_lds(d.nzp, pb0.bz), s>>>(one.pb, nullptr, d);
    PROF_END(s);
}

template <typename T>
void launch_poisson_forward(const T* U, const T* V, const T* W, const Metric<double>& md, const Dims& d,
                            PoissonBufs& pb, hipStream_t s)
{
    run_poisson_forward<T>(FwdArgs<T>{U, V, W, md, pb}, nullptr, nullptr, 1, d, pb, s);
}

void launch_poisson_forward_field(const double* rhs, const Dims& d, PoissonBufs& pb, hipStream_t s)
{
    const int npairs = (d.ny - 1) * (d.nzp / 2);
    FwdArgs<double> one{nullptr, nullptr, nullptr, Metric<double>{}, pb};
    PROF_BEGIN("k_fftx_fwd", s);
    k_fftx_fwd<double, true, false><<<cdiv(npairs, pb.bx), FFT_THREADS, fft_lds(d.nxp, pb.bx), s>>>(one, nullptr, rhs,
                                                                                                  d);
    PROF_END(s);
    PROF_BEGIN("k_fftz", s);
    k_fftz<-1, false><<<cdiv((long)(d.ny - 1) * d.nkx, pb.bz), FFT_THREADS, fft_lds(d.nzp, pb.bz), s>>>(pb, nullptr,
                                                                                                      d);
    PROF_END(s);
}

bool tridiag_needs_table(const Dims& d) { return d.ny - 1 > TRI_KEEP; }

void launch_tridiag_setup(const Dims& d, PoissonBufs& pb, hipStream_t s)
{
    k_tridiag_setup<<<cdiv((long)d.nkx * d.nzp, TRI_THREADS), TRI_THREADS, 0, s>>>(d, pb);
}

void run_poisson_solve(const PoissonBufs& one, const PoissonBufs* tab, int nmem, const Dims& d, hipStream_t s)
{
    const long lanes = 2L * d.nkx * d.nzp;
    PROF_BEGIN("k_tridiag", s);
    const int n = d.ny - 1;
    const dim3 g1(cdiv(lanes, TRI_THREADS), 1, nmem), g2(cdiv(2 * lanes, TRI_THREADS), 1, nmem);
    if (n > TRI_KEEP) {
        if (tab) k_tridiag<true, 32, true><<<g1, TRI_THREADS, 0, s>>>(one, tab, d);
        else k_tridiag<true, 32, false><<<g1, TRI_THREADS, 0, s>>>(one, nullptr, d);
    } else if (n > 128) {
        if (tab) k_tridiag_tw<96, true><<<g2, TRI_THREADS, 0, s>>>(one, tab, d);
        else k_tridiag_tw<96, false><<<g2, TRI_THREADS, 0, s>>>(one, nullptr, d);
    } else if (n > 96) {
        if (tab) k_tridiag_tw<64, true><<<g2, TRI_THREADS, 0, s>>>(one, tab, d);
        else k_tridiag_tw<64, false><<<g2, TRI_THREADS, 0, s>>>(one, nullptr, d);
    } else if (n > 64) {
        if (tab) k_tridiag_tw<48, true><<<g2, TRI_THREADS, 0, s>>>(one, tab, d);
        else k_tridiag_tw<48, false><<<g2, TRI_THREADS, 0, s>>>(one, nullptr, d);
    } else if (n > 32) {
        if (tab) k_tridiag_tw<32, true><<<g2, TRI_THREADS, 0, s>>>(one, tab, d);
        else k_tridiag_tw<32, false><<<g2, TRI_THREADS, 0, s>>>(one, nullptr, d);
    } else {
        if (tab) k_tridiag_tw<16, true><<<g2, TRI_THREADS, 0, s>>>(one, tab, d);
        else k_tridiag_tw<16, false><<<g2, TRI_THREADS, 0, s>>>(one, nullptr, d);
    }
    PROF_END(s);
}

void launch_poisson_solve(const Dims& d, PoissonBufs& pb, hipStream_t s) { run_poisson_solve(pb, nullptr, 1, d, s); }

void run_poisson_inverse(const PoissonBufs& one, const PoissonBufs* tab, int nmem, const Dims& d, hipStream_t s)
{
    PROF_BEGIN("k_fftz", s);
    const dim3 gz(cdiv((long)(d.ny - 1) * d.nkx, one.bz), 1, nmem);
    if (tab) k_fftz<+1, true><<<gz, FFT_THREADS, fft_lds(d.nzp, one.bz), s>>>(one, tab, d);
    else k_fftz<+1, false><<<gz, FFT_THREADS, fft_lds(d.nzp, one.bz), s>>>(one, nullptr, d);
    PROF_END(s);
    const int npairs = (d.ny - 1) * (d.nzp / 2);
    const dim3 gx(cdiv(npairs, one.bx), 1, nmem);
    PROF_BEGIN("k_fftx_inv", s);
    if (tab) k_fftx_inv<true><<<gx, FFT_THREADS, fft_lds(d.nxp, one.bx), s>>>(one, tab, d);
    else k_fftx_inv<false><<<gx, FFT_THREADS, fft_lds(d.nxp, one.bx), s>>>(one, nullptr, d);
    PROF_END(s);
}

void launch_poisson_inverse(const Dims& d, PoissonBufs& pb, hipStream_t s)
{
    run_poisson_inverse(pb, nullptr, 1, d, s);
}

// code/projection.f90:192-214: u -= grad p on the canonical cells, each
// stored with its boundary-condition images (the reference applies
// apply_boundary_conditions right after, code/time_integration.f90:49).
template <typename T, bool G>
__global__ void __launch_bounds__(256) k_project(ProjArgs<T> one, const ProjArgs<T>* tab, Dims d)
{
    // XCD-contiguous tile order: the plane k+1 that a tile reads is the
    // plane k of a tile that runs on the same XCD, so p is fetched once
    const GBlk gb = gblk_xcd(d.nzg - 3);
    const ProjArgs<T> arg = member_args<G>(one, tab, gb.m);
    T *U = arg.U, *V = arg.V, *W = arg.W;
    const double* __restrict__ p = arg.p;
    const Metric<double>& m = arg.m;
    const TileId t{gb.x, gb.y, gb.z};
    const int i = t.x * blockDim.x + threadIdx.x + 2;
    const int j = t.y * blockDim.y + threadIdx.y + 2;
    const int k = t.z + 2;
    if (i > d.nxg - 2 || j > d.nyg - 1) return;
    const long q = at(d, i, j, k);
    const double p0 = p[q];
    // the spacings' reciprocals: within an ulp of the reference's divisions
    const T un = T((double)U[q] - (p[at(d, i + 1, j, k)] - p0) * m.rdxg[i + 1]);
    store_img(U, d, 0, i, j, k, arg.shift ? un + T(arg.shift->last_dPdx) : un);
    if (j <= d.ny - 1)
        store_img(V, d, 1, i, j, k, T((double)V[q] - (p[at(d, i, j + 1, k)] - p0) * m.rdyg[j + 1]));
    store_img(W, d, 2, i, j, k, T((double)W[q] - (p[at(d, i, j, k + 1)] - p0) * m.rdzg[k + 1]));
}

template <typename T>
void run_project(const ProjArgs<T>& one, const ProjArgs<T>* tab, int nmem, const Dims& d, hipStream_t s)
{
    const dim3 blk = row_block(d.nxg - 3);
    const dim3 grd(cdiv(d.nxg - 3, blk.x), cdiv(d.nyg - 2, blk.y), (unsigned)(d.nzg - 3) * nmem);
    PROF_BEGIN("k_project", s);
    if (tab) k_project<T, true><<<grd, blk, 0, s>>>(one, tab, d);
    else k_project<T, false><<<grd, blk, 0, s>>>(one, nullptr, d);
    PROF_END(s);
}

template <typename T>
void launch_project(T* U, T* V, T* W, const double* p, const Metric<double>& md, const Dims& d, hipStream_t s,
                    const Scalars* shift)
{
    run_project<T>(ProjArgs<T>{U, V, W, p, md, shift}, nullptr, 1, d, s);
}

template void launch_poisson_forward<double>(const double*, const double*, const double*, const Metric<double>&,
                                             const Dims&, PoissonBufs&, hipStream_t);
template void launch_poisson_forward<float>(const float*, const float*, const float*, const Metric<double>&,
                                            const Dims&, PoissonBufs&, hipStream_t);
template void run_poisson_forward<double>(const FwdArgs<double>&, const FwdArgs<double>*, const PoissonBufs*, int,
                                          const Dims&, const PoissonBufs&, hipStream_t);
template void run_poisson_forward<float>(const FwdArgs<float>&, const FwdArgs<float>*, const PoissonBufs*, int,
                                         const Dims&, const PoissonBufs&, hipStream_t);
template void run_project<double>(const ProjArgs<double>&, const ProjArgs<double>*, int, const Dims&, hipStream_t);
template void run_project<float>(const ProjArgs<float>&, const ProjArgs<float>*, int, const Dims&, hipStream_t);
template void launch_project<double>(double*, double*, double*, const double*, const Metric<double>&, const Dims&,
                                     hipStream_t, const Scalars*);
template void launch_project<float>(float*, float*, float*, const double*, const Metric<double>&, const Dims&,
                                    hipStream_t, const Scalars*);

}  // namespace chan
