// dsm_fused.hip -- dynamic Smagorinsky pass A in one kernel
// (code/subgrid.f90:379-510): centre velocities, rate of strain |S| S_ij,
// filtered staggered velocities, the Leonard tensor L_ij, the model tensor
// M_ij and the per-row partial sums of L:M and M:M, plus |S| for pass B.
//
// Design: a block owns an x-y tile of TXO x TYO output cells and marches in
// z over a chunk of planes.  Threads map one-to-one onto the tile plus a halo
// of one ("columns"), x fastest.  Level 0 = U,V,W of the tile plus a halo
// of two, three planes, in LDS, loaded with periodic index wrap (no ghost
// cells are read).  Per plane, two phases separated by two barriers:
//   1  per column: the y-z taps of the staggered-velocity filter at the
//      column's level-0 position (the row ends also at the halo position
//      beyond them), their x taps by cross-lane moves (filtered velocities
//      to LDS, three planes); level 1: centre velocities a, |S| S_ij, and
//      the z taps of the filter inputs (running sums in registers, the
//      finished plane to LDS);
//   2  per output: the y taps of the filter inputs from LDS, the x taps by
//      cross-lane moves, hat(S) from the filtered velocities, L:M and M:M;
//      the next level-0 plane is stored and the one after it fetched.
// The test filter is separable (y weights one-sided on rows 2 and nyg-1 as
// the reference).  Spacing divisions are multiplications by precomputed
// reciprocals.  The sums run in a different order than the reference, so the
// model agrees with it to rounding, not bit for bit.
// Global traffic: U,V,W read once per tile (plus halo), |S| written once.
#include "kernels.hpp"
#include "prof.hpp"

namespace chan {

// longest z chunk a block marches (round_filling_chunk's upper bound below)
constexpr int DSM_KMAX = 64;

template <int TXO, int TYO>
struct DsmTile {
    static constexpr int CX = TXO + 2, CY = TYO + 2, NC = CX * CY;  // columns
    static constexpr int LX = TXO + 4, LY = TYO + 4, NL = LX * LY;  // level-0 plane
};

// Cross-lane moves along x.  Columns are x-fastest with CX columns per row
// and CX divides the wave, so lane l's x neighbours are lanes l-1 and l+1; a
// whole-wave shift by one lane (DPP wave_shr:1 / wave_shl:1) delivers them.
// Only the ends of a row (halo columns cx = 0 and CX-1, which produce no
// output) receive a lane of another row, or 0 at the ends of the wave.
// The whole wave must be active.
constexpr int DPP_WAVE_SHL1 = 0x130, DPP_WAVE_SHR1 = 0x138;
template <int CTRL>
__device__ inline double lane_shift(double v)
{
    const unsigned long long b = __builtin_bit_cast(unsigned long long, v);
    const int lo = __builtin_amdgcn_update_dpp(0, (int)(unsigned)b, CTRL, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_update_dpp(0, (int)(unsigned)(b >> 32), CTRL, 0xf, 0xf, true);
    return __builtin_bit_cast(double, (unsigned long long)(unsigned)lo | ((unsigned long long)(unsigned)hi << 32));
}
template <int CTRL>
__device__ inline float lane_shift(float v)
{
    return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(0, __builtin_bit_cast(int, v), CTRL, 0xf, 0xf, true));
}

// Off-diagonal rate of strain of code/subgrid.f90:1027-1070: the average
// over the four surrounding edges of (dA/db + dB/da)/2,
//   1/4 sum_{da,db} 1/2 ((A(da-1,db) - A(da-1,db-1)) rb[db] + (B(da,db-1) - B(da-1,db-1)) ra[da]),
// with the sums regrouped by metric factor (four products instead of eight,
// one scaling instead of five).  A(a, b): a in {-1,0}, b in {-1,0,1};
// B(a, b): a in {-1,0,1}, b in {-1,0}; ra/rb the reciprocal spacings along
// B's and A's difference direction.
template <typename T, typename FA, typename FB>
__device__ __forceinline__ T strain_offdiag(const FA& A, const FB& B, T ra0, T ra1, T rb0, T rb1)
{
    const T sA0 = (A(-1, 0) - A(-1, -1)) + (A(0, 0) - A(0, -1));
    const T sA1 = (A(-1, 1) - A(-1, 0)) + (A(0, 1) - A(0, 0));
    const T sB0 = (B(0, -1) - B(-1, -1)) + (B(0, 0) - B(-1, 0));
    const T sB1 = (B(1, -1) - B(0, -1)) + (B(1, 0) - B(0, 0));
    return T(0.125) * ((rb0 * sA0 + rb1 * sA1) + (ra0 * sB0 + ra1 * sB1));
}

// Square root of a sum of squares x >= 0: the hardware reciprocal square
// root refined by the Newton-Raphson (Goldschmidt) steps of the library
// sequence, without its rescaling of tiny and huge arguments (a strain
// magnitude squared stays far inside the normal range; below 2^-960, where
// the hardware estimate would overflow, the result is 0, as for x = 0).
// Within an ulp of the correctly rounded root, in 13 instructions instead
// of the library's 17.
__device__ __forceinline__ double sqrt_sum_sq(double x)
{
    const double r = __builtin_amdgcn_rsq(x);
    double g = x * r, h = 0.5 * r;
    const double e = __builtin_fma(-h, g, 0.5);
    g = __builtin_fma(g, e, g);
    h = __builtin_fma(h, e, h);
    double d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    d = __builtin_fma(-g, g, x);
    g = __builtin_fma(d, h, g);
    return x > 0x1p-960 ? g : 0.0;
}
__device__ __forceinline__ float sqrt_sum_sq(float x) { return sqrtf(x); }

// The register budget is held to two waves per SIMD (the FP64 kernels need
// it; the FP32 ones fit in fewer registers and run three or four).
template <typename T, int TXO, int TYO, bool G>
__global__ void __launch_bounds__((TXO + 2) * (TYO + 2)) __attribute__((amdgpu_waves_per_eu(sizeof(T) == 8 ? 2 : 4)))
    k_dsm_fused(DsmArgs<T> one, const DsmArgs<T>* tab, Dims d, int kchunk)
{
    // The model's sums are reordered against the reference anyway (separable
    // filter, per-row partials); contracting a*b+c into FMAs here saves a
    // third of the FP64 instructions and changes results only by rounding.
#pragma clang fp contract(fast)
    using Tl = DsmTile<TXO, TYO>;
    constexpr int CX = Tl::CX, NC = Tl::NC, LX = Tl::LX, NL = Tl::NL;
    static_assert(64 % CX == 0 && NC % 64 == 0, "rows of columns must tile the waves");
    __shared__ T l0[3][3][NL];  // [slot][U,V,W][position]
    __shared__ T lf[3][3][NC];  // [slot][Uf,Vf,Wf][column]
    __shared__ T zq[15][NC];    // z-combined filter inputs: a(3), a_a a_b (6), |S|S_ij (6)

    const int tid = threadIdx.x;
    const int cx = tid % CX, cy = tid / CX;
    const GBlk gb = gblk_xcd((d.nzg - 2 + kchunk - 1) / kchunk);
    const DsmArgs<T> arg = member_args<G>(one, tab, gb.m);
    const T* __restrict__ U = arg.U;
    const T* __restrict__ V = arg.V;
    const T* __restrict__ W = arg.W;
    const DsmBufs<T> b = arg.b;
    const TileId tile{gb.x, gb.y, gb.z};
    const int i0 = tile.x * TXO + 2, j0 = tile.y * TYO + 2;
    const int kb = tile.z * kchunk + 2;
    const int ke = min(kb + kchunk - 1, d.nzg - 1);
    const int ci = i0 - 1 + cx, cj = j0 - 1 + cy;  // column position (x unwrapped)
    const int nxp = d.nxp, nzp = d.nzp;
    // periodic wrap onto the unique points 2..nxp+1; every index used here is
    // within one period of that range
    auto wx = [&](int q) { return q < 2 ? q + nxp : (q > nxp + 1 ? q - nxp : q); };
    auto wz = [&](int q) { return q < 2 ? q + nzp : (q > nzp + 1 ? q - nzp : q); };
    const T h = T(0.5);
    // x, y and z reciprocal spacings of the block's columns, rows and
    // planes, in LDS: read per plane from here rather than from global
    // memory, whose loads would wait behind the velocity prefetch (memory
    // counters drain in issue order, so waiting for a metric load issued
    // after the prefetch waits for the prefetch too).  Level 1 uses the
    // canonical x image, the outputs the true index (code/subgrid.f90:456);
    // the z factors of plane p = kb-1+q sit at q (p's periodic image, which
    // is p itself on the block's output planes).
    __shared__ T mx[6][CX];  // rdx, rdxg(i), rdxg(i+1) at wx(ci); the same at ci
    __shared__ T my[3][TYO + 2];  // rdy, rdyg(j), rdyg(j+1)
    __shared__ T mz[3][DSM_KMAX + 3];  // rdz(p), rdzg(p), rdzg(p+1)
    __shared__ double mzd[DSM_KMAX + 3];  // zg(p+1) - zg(p), for compute_dt
    for (int q = tid; q < ke - kb + 3; q += NC) {
        const int pc = wz(kb - 1 + q);
        mz[0][q] = b.rdz[pc];
        mz[1][q] = b.rdzg[pc];
        mz[2][q] = b.rdzg[pc + 1];
        if (b.dtpart) mzd[q] = b.zg[pc + 1] - b.zg[pc];
    }
    if (tid < CX) {
        const int ic = wx(i0 - 1 + tid), it = min(max(i0 - 1 + tid, 2), d.nxg - 1);
        mx[0][tid] = b.rdx[ic];
        mx[1][tid] = b.rdxg[ic];
        mx[2][tid] = b.rdxg[ic + 1];
        mx[3][tid] = b.rdx[it];
        mx[4][tid] = b.rdxg[it];
        mx[5][tid] = b.rdxg[it + 1];
    } else if (tid < CX + TYO + 2) {
        const int r = tid - CX, jt = min(max(j0 - 1 + r, 2), d.nyg - 1);
        my[0][r] = b.rdy[jt];
        my[1][r] = b.rdyg[jt];
        my[2][r] = b.rdyg[jt + 1];
    }

    auto slot = [&](int p) { return (p - kb + 3) % 3; };  // ring slot of plane p >= kb-2

    // this thread's share of a level-0 plane: box offset within a plane
    // (-1 outside the walls) and whether V exists there; plane p's values are
    // fetched into registers (gload) one iteration before they are stored
    // into the ring (sstore), so the loads overlap the previous plane's work
    constexpr int NLD = (NL + NC - 1) / NC;
    long lo[NLD];
    bool lv[NLD];
#pragma unroll
    for (int r = 0; r < NLD; r++) {
        const int q = tid + r * NC;
        const int lx = q % LX, ly = q / LX;
        const int ii = wx(i0 - 2 + lx), jj = j0 - 2 + ly;
        const bool in = q < NL && jj >= 1 && jj <= d.nyg;
        lo[r] = in ? at(d, ii, jj, 1) : -1;
        lv[r] = in && jj <= d.ny;
    }
    T pre[NLD][3];
    auto gload = [&](int p) {
        const long po = (long)(wz(p) - 1) * d.sxy;
#pragma unroll
        for (int r = 0; r < NLD; r++) {
            T u = T(0), v = T(0), w = T(0);
            if (lo[r] >= 0) {
                u = U[lo[r] + po];
                w = W[lo[r] + po];
                if (lv[r]) v = V[lo[r] + po];
            }
            pre[r][0] = u;
            pre[r][1] = v;
            pre[r][2] = w;
        }
    };
    auto sstore = [&](int p) {
        const int s = slot(p);
#pragma unroll
        for (int r = 0; r < NLD; r++) {
            const int q = tid + r * NC;
            if (q < NL) {
                l0[s][0][q] = pre[r][0];
                l0[s][1][q] = pre[r][1];
                l0[s][2][q] = pre[r][2];
            }
        }
    };
    auto load_plane = [&](int p) {
        gload(p);
        sstore(p);
    };
    auto L0 = [&](int f, int p, int dx, int dy) -> T { return l0[slot(p)][f][(cx + 1 + dx) + LX * (cy + 1 + dy)]; };

    // y and z taps of the filter of the staggered velocities (code/subgrid.f90:
    // 140-190 over the whole y range, one-sided at the first/last row of each
    // field).  Each column takes them at its own level-0 position (cx+1,
    // cy+1) and, for the x taps beyond the row ends, the first and last
    // column also at the halo position next to them (lx = 0, LX-1; the other
    // lanes repeat their own).  The y taps of a plane are taken once and
    // enter the z taps of three planes through two running sums (ga: the next
    // plane's, two terms in; gb: the one after, one term in).  Every
    // direction carries the weights (1, 4, 1), the factor 1/6 of each taken
    // once at the end (1/216).  The one-sided rows (the first and last row
    // of each field, weights (0, 4, 2) and (2, 4, 0)) read their inner
    // neighbour for both taps: the y weights become a choice of rows, fixed
    // per column.  Rows outside a field (the walls' halo rows, V's row nyg)
    // produce values that no output reads.
    int ym[2], yp[2];  // [U and W, V]: LDS offsets of the rows below and above
    for (int t = 0; t < 2; t++) {
        const int n2 = t == 0 ? d.nyg : d.ny;
        ym[t] = cj == 1 ? LX : -LX;
        yp[t] = cj == n2 ? -LX : LX;
    }
    const int qown = (cx + 1) + LX * (cy + 1);
    const int qext = (cx == 0 ? 0 : (cx == CX - 1 ? LX - 1 : cx + 1)) + LX * (cy + 1);
    T gA[2][3], gB[2][3];  // running sums [own, halo position][U,V,W]
    auto ycomb = [&](int p, T out[2][3]) {
        const int s = slot(p);
        for (int f = 0; f < 3; f++) {
            const T* P = l0[s][f];
            const int m = ym[f == 1 ? 1 : 0], q = yp[f == 1 ? 1 : 0];
            out[0][f] = (P[qown + m] + P[qown + q]) + T(4) * P[qown];
            out[1][f] = (P[qext + m] + P[qext + q]) + T(4) * P[qext];
        }
    };
    // filtered staggered velocities of the column at plane p into LDS: the
    // running sums hold planes p-1 and p (ga) and p (gb), the y taps of plane
    // p+1 complete the z taps; x taps from the neighbouring lanes (the halo
    // position at the row ends).  Every lane of the wave must take part.
    // The sums are updated in place with their roles swapped (gb becomes the
    // next call's ga, ga its gb): callers alternate the two arrays, which
    // keeps the march free of register copies.
    auto filt_vel = [&](int p, T (&ga)[2][3], T (&gb)[2][3]) {
        T yc[2][3];
        ycomb(p + 1, yc);
        const int sf = slot(p);
        for (int f = 0; f < 3; f++) {
            const T own = ga[0][f] + yc[0][f];
            const T ext = ga[1][f] + yc[1][f];
            for (int e = 0; e < 2; e++) {
                gb[e][f] = gb[e][f] + T(4) * yc[e][f];
                ga[e][f] = yc[e][f];
            }
            // the moves run on every lane before the selection: inside a
            // conditional they would read disabled lanes as 0
            const T sl = lane_shift<DPP_WAVE_SHR1>(own), sr = lane_shift<DPP_WAVE_SHL1>(own);
            const T left = cx == 0 ? ext : sl;
            const T right = cx == CX - 1 ? ext : sr;
            lf[sf][f][tid] = ((left + right) + T(4) * own) * T(1.0 / 216.0);
        }
    };

    const bool out_col = cx >= 1 && cx <= TXO && cy >= 1 && cy <= TYO && ci <= d.nxg - 1 && cj <= d.nyg - 1;
    // compute_dt over this column's output cells (with b.dtpart): the
    // smallest |U| and |V| (x and y spacings are the column's own, and a
    // correctly rounded quotient falls as its divisor grows, so the largest
    // dx/|U| is dx over the smallest |U|, exactly), and the largest dz/|W|
    // plane by plane (the z spacings differ in their last bits)
    T dt_mu = T(INFINITY), dt_mv = T(INFINITY);
    double dt_lw = 0.0;
    const int pa[6] = {0, 1, 2, 0, 0, 1}, pb[6] = {0, 1, 2, 1, 2, 2};

    // level 1 of this column at plane p: the filter inputs Q = a (3),
    // a_a a_b (6), |S| S_ij (6); |S| of the output planes to global memory
    // Rows outside 2..nyg-1 (the wall ghost rows and beyond) compute on the
    // zeros loaded there and finite metric factors: their inputs enter the
    // y taps of rows 2 and nyg-1 only with weight 0 and the other taps not
    // at all, so no branch is needed around them.
    auto level1 = [&](int p, T Q[15]) {
        T a[3], ss[6], sm;
        {
            // code/subgrid.f90:394-396 centre velocities
            a[0] = h * (L0(0, p, -1, 0) + L0(0, p, 0, 0));
            a[1] = h * (L0(1, p, 0, -1) + L0(1, p, 0, 0));
            a[2] = h * (L0(2, p - 1, 0, 0) + L0(2, p, 0, 0));
            // code/subgrid.f90:1027-1070 at the canonical image (wx(ci), cj, pc)
            T s[6];
            s[0] = (L0(0, p, 0, 0) - L0(0, p, -1, 0)) * mx[0][cx];
            s[1] = (L0(1, p, 0, 0) - L0(1, p, 0, -1)) * my[0][cy];
            const int zq_ = p - kb + 1;
            s[2] = (L0(2, p, 0, 0) - L0(2, p - 1, 0, 0)) * mz[0][zq_];
            const T rz0 = mz[1][zq_], rz1 = mz[2][zq_];
            s[3] = strain_offdiag([&](int a, int c) { return L0(0, p, a, c); },
                                  [&](int a, int c) { return L0(1, p, a, c); }, mx[1][cx], mx[2][cx], my[1][cy],
                                  my[2][cy]);
            s[4] = strain_offdiag([&](int a, int c) { return L0(0, p + c, a, 0); },
                                  [&](int a, int c) { return L0(2, p + c, a, 0); }, mx[1][cx], mx[2][cx], rz0, rz1);
            s[5] = strain_offdiag([&](int a, int c) { return L0(1, p + c, 0, a); },
                                  [&](int a, int c) { return L0(2, p + c, 0, a); }, my[1][cy], my[2][cy], rz0, rz1);
            T acc = T(2) * s[0] * s[0];
            acc = acc + T(2) * s[1] * s[1];
            acc = acc + T(2) * s[2] * s[2];
            acc = acc + T(4) * s[3] * s[3];
            acc = acc + T(4) * s[4] * s[4];
            acc = acc + T(4) * s[5] * s[5];
            sm = sqrt_sum_sq(acc);
            for (int c = 0; c < 6; c++) ss[c] = sm * s[c];
        }
        for (int c = 0; c < 3; c++) Q[c] = a[c];
        for (int c = 0; c < 6; c++) {
            Q[3 + c] = a[pa[c]] * a[pb[c]];
            Q[9 + c] = ss[c];
        }
        if (out_col && p >= kb && p <= ke) {
            b.smag[at(d, ci, cj, p)] = sm;
            if (b.dtpart) {
                dt_mu = fmin(dt_mu, fabs(L0(0, p, 0, 0)));
                dt_mv = fmin(dt_mv, fabs(L0(1, p, 0, 0)));
                dt_lw = fmax(dt_lw, mzd[p - kb + 1] / fabs((double)L0(2, p, 0, 0)));
            }
        }
    };

    // z taps of the filter inputs as running sums: for plane k (planes k-1
    // and k in) and for plane k+1 (plane k in), in zA and zB with roles
    // alternating as in filt_vel
    T zA[15], zB[15], Q[15];
    // pre-roll: level 0 planes kb-2..kb+2, level 1 at kb-1 and kb
    load_plane(kb - 2);
    load_plane(kb - 1);
    load_plane(kb);
    __syncthreads();
    {
        T yc[2][3];
        ycomb(kb - 2, yc);
        for (int e = 0; e < 2; e++)
            for (int f = 0; f < 3; f++) gA[e][f] = yc[e][f];
        ycomb(kb - 1, yc);
        for (int e = 0; e < 2; e++)
            for (int f = 0; f < 3; f++) {
                gA[e][f] = gA[e][f] + T(4) * yc[e][f];
                gB[e][f] = yc[e][f];
            }
    }
    filt_vel(kb - 1, gA, gB);
    level1(kb - 1, Q);
    // the 15 filter inputs carry the test filter's weights (1, 4, 1) per
    // direction, the factor 1/6 of each left for the end (1/216 in all):
    // each tap set is then one add and one fused multiply-add
    for (int c = 0; c < 15; c++) zA[c] = Q[c];
    __syncthreads();
    load_plane(kb + 1);
    __syncthreads();
    filt_vel(kb, gB, gA);
    level1(kb, Q);
    for (int c = 0; c < 15; c++) {
        zA[c] = zA[c] + T(4) * Q[c];
        zB[c] = Q[c];
    }
    __syncthreads();
    load_plane(kb + 2);
    if (kb + 3 <= ke + 2) gload(kb + 3);
    __syncthreads();

    double acc_lm = 0.0, acc_mm = 0.0;
    // y taps of the filter inputs for this column's row (one-sided at rows 2
    // and nyg-1, whose weights (0, 4, 2) and (2, 4, 0) read the inner
    // neighbour for both taps), taken on the output rows 1..TYO only
    const int om0 = cx + CX * max(cy - 1, 0), op0 = cx + CX * min(cy + 1, Tl::CY - 1);
    const int om = cj == 2 ? op0 : om0, op = cj == d.nyg - 1 ? om0 : op0;

    // one output plane k; za/zb and ga/gb as described above (roles swap)
    auto plane = [&](int k, T (&za)[15], T (&zb)[15], T (&ga)[2][3], T (&gb)[2][3]) {
        // phase 1: l0 holds planes k..k+2
        filt_vel(k + 1, ga, gb);
        level1(k + 1, Q);
        T zown[15];  // z taps of the filter inputs at plane k
        for (int c = 0; c < 15; c++) {
            zown[c] = za[c] + Q[c];
            zb[c] = zb[c] + T(4) * Q[c];
            za[c] = Q[c];
        }
        for (int c = 0; c < 15; c++) zq[c][tid] = zown[c];
        __syncthreads();
        // phase 2: x-y taps of the 15 filter inputs, every lane (cross-lane moves)
        T tp[15];
        // on the halo rows (cy 0 and CY-1, clamped indices) the taps are
        // taken as well and never used: no branch
        for (int c = 0; c < 15; c++) tp[c] = (zq[c][om] + zq[c][op]) + T(4) * zown[c];
        for (int c = 0; c < 15; c++) {
            const T y = tp[c];
            tp[c] = (lane_shift<DPP_WAVE_SHR1>(y) + lane_shift<DPP_WAVE_SHL1>(y)) + T(4) * y;
        }
        if (out_col) {
            // hat(S) from the filtered velocities at the true index (code/subgrid.f90:456)
            auto uf = [&](int dx, int dy, int dz) { return lf[slot(k + dz)][0][(cx + dx) + CX * (cy + dy)]; };
            auto vf = [&](int dx, int dy, int dz) { return lf[slot(k + dz)][1][(cx + dx) + CX * (cy + dy)]; };
            auto wf = [&](int dx, int dy, int dz) { return lf[slot(k + dz)][2][(cx + dx) + CX * (cy + dy)]; };
            T sh[6];
            sh[0] = (uf(0, 0, 0) - uf(-1, 0, 0)) * mx[3][cx];
            sh[1] = (vf(0, 0, 0) - vf(0, -1, 0)) * my[0][cy];
            sh[2] = (wf(0, 0, 0) - wf(0, 0, -1)) * mz[0][k - kb + 1];
            const T rz0 = mz[1][k - kb + 1], rz1 = mz[2][k - kb + 1];
            sh[3] = strain_offdiag([&](int a, int c) { return uf(a, c, 0); }, [&](int a, int c) { return vf(a, c, 0); },
                                   mx[4][cx], mx[5][cx], my[1][cy], my[2][cy]);
            sh[4] = strain_offdiag([&](int a, int c) { return uf(a, 0, c); }, [&](int a, int c) { return wf(a, 0, c); },
                                   mx[4][cx], mx[5][cx], rz0, rz1);
            sh[5] = strain_offdiag([&](int a, int c) { return vf(0, a, c); }, [&](int a, int c) { return wf(0, a, c); },
                                   my[1][cy], my[2][cy], rz0, rz1);
            T acc = T(2) * sh[0] * sh[0];
            acc = acc + T(2) * sh[1] * sh[1];
            acc = acc + T(2) * sh[2] * sh[2];
            acc = acc + T(4) * sh[3] * sh[3];
            acc = acc + T(4) * sh[4] * sh[4];
            acc = acc + T(4) * sh[5] * sh[5];
            const T shm4 = T(4) * sqrt_sum_sq(acc);
            // tp holds 216 times the filtered inputs: L = s (T_ab - s T_a T_b)
            // and M = 4|hat S| hat S_ab - s T(|S| S_ab), s = 1/216, the outer
            // s of L taken once on L:M
            const T sf = T(1.0 / 216.0);
            const T sa[3] = {sf * tp[0], sf * tp[1], sf * tp[2]};
            T lm = T(0), mm = T(0);
            for (int q = 0; q < 6; q++) {
                const T L = tp[3 + q] - sa[pa[q]] * tp[pb[q]];
                const T M = shm4 * sh[q] - sf * tp[9 + q];
                const T wgt = q < 3 ? T(1) : T(2);
                lm = lm + wgt * (L * M);
                mm = mm + wgt * (M * M);
            }
            acc_lm += (double)(sf * lm);
            acc_mm += (double)mm;
        }
        // next level-0 plane into the slot of plane k (last read in phase 1)
        if (k + 3 <= ke + 2) sstore(k + 3);
        if (k + 4 <= ke + 2) gload(k + 4);
        __syncthreads();
    };
    // unrolled by two so the running sums alternate arrays without copies
    for (int k = kb; k <= ke; k += 2) {
        plane(k, zA, zB, gA, gB);
        if (k + 1 > ke) break;
        plane(k + 1, zB, zA, gB, gA);
    }
    // per-row partial sums over the block's columns, in a fixed order
    double* red = reinterpret_cast<double*>(&l0[0][0][0]);
    red[tid] = out_col ? acc_lm : 0.0;
    red[NC + tid] = out_col ? acc_mm : 0.0;
    __syncthreads();
    if (tid < TYO) {
        const int cyy = tid + 1, j = j0 + tid;
        if (j <= d.nyg - 1) {
            double sl = 0.0, sm2 = 0.0;
            for (int c = 1; c <= TXO; c++) {
                sl += red[c + CX * cyy];
                sm2 += red[NC + c + CX * cyy];
            }
            const int blk = tile.z * gb.gx + tile.x;
            b.part[((long)(j - 1) * 2 + 0) * b.nblk + blk] = sl;
            b.part[((long)(j - 1) * 2 + 1) * b.nblk + blk] = sm2;
        }
    }
    if (b.dtpart) {
        // the block's compute_dt maxima: per lane, then per wave (butterfly),
        // then over the waves; one slot per block of the member's grid
        double m3[3] = {0.0, 0.0, dt_lw};
        if (out_col) {
            m3[0] = (b.xg[ci + 1] - b.xg[ci]) / (double)dt_mu;
            m3[1] = (b.yg[cj + 1] - b.yg[cj]) / (double)dt_mv;
        }
        for (int m = 32; m > 0; m >>= 1)
            for (int c = 0; c < 3; c++) m3[c] = fmax(m3[c], __shfl_xor(m3[c], m, 64));
        __shared__ double wmax[NC / 64][3];
        if ((tid & 63) == 0)
            for (int c = 0; c < 3; c++) wmax[tid >> 6][c] = m3[c];
        __syncthreads();
        if (tid < 3) {
            double v = 0.0;
            for (int w = 0; w < NC / 64; w++) v = fmax(v, wmax[w][tid]);
            const int nb = gb.gx * gb.gy * gb.gz;
            b.dtpart[(long)tid * nb + (tile.z * gb.gy + tile.y) * gb.gx + tile.x] = v;
        }
    }
}

// Tile shapes, (x outputs, y outputs): a row of TXO + 2 columns divides the
// wave.  Shape 0 has 30 outputs per row (32 columns, half a wave): the least
// halo per output on long rows; shape 1 has 14 (16 columns) and twice the
// rows: the same halo per output, and less of the last x tile left empty on
// short rows (33, 65 or 97 cells: 60, 90, 120 columns of 30 against 42, 70,
// 98 of 14).  FP64: shapes 0 and 1 are 512-thread blocks (144 KB of LDS, one
// block of 8 waves per CU, whose two barriers per plane stall the whole CU);
// shape 2, 14 x 14 outputs in 256 threads (72 KB), puts two blocks of 4
// waves on a CU, whose barriers overlap, for 7% more halo columns.  FP32:
// shapes 0 and 1 are 256-thread blocks (40 KB), shape 2 is 30 x 14 outputs
// in 512 threads (72 KB), the least halo per output.  Shape 3, both
// precisions: 14 x 6 outputs in 128 threads (FP64 39 KB: four blocks of two
// waves per CU), 1.52 columns per output, for small grids, where it gives
// one channel more blocks.
struct DsmShape { int txo, tyo; };
template <typename T>
constexpr DsmShape dsm_shape(int q)
{
    return q == 3 ? DsmShape{14, 6}
                  : (sizeof(T) == 8 ? (q == 0 ? DsmShape{30, 14} : (q == 1 ? DsmShape{14, 30} : DsmShape{14, 14}))
                                    : (q == 0 ? DsmShape{30, 6} : (q == 1 ? DsmShape{14, 14} : DsmShape{30, 14})));
}
int dsm_shape_count() { return 4; }

template <typename T, int TXO, int TYO>
int dsm_per_cu()
{
    int per_cu = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k_dsm_fused<T, TXO, TYO, false>, (TXO + 2) * (TYO + 2),
                                                     0) != hipSuccess ||
        per_cu < 1)
        per_cu = 1;
    return per_cu;
}

// z-chunk length filling the last round of blocks (e.g. 192x97x192 FP64,
// 49 tiles per chunk on 256 slots: 39 planes, 5 chunks, one round, against
// 20 planes, 10 chunks, two rounds).
template <typename T>
int dsm_fused_kchunk(const Dims& d, int device, int shape)
{
    int ncu = 0;
    const DsmShape sh = dsm_shape<T>(shape);
    const int per_cu = shape == 0   ? dsm_per_cu<T, dsm_shape<T>(0).txo, dsm_shape<T>(0).tyo>()
                       : shape == 1 ? dsm_per_cu<T, dsm_shape<T>(1).txo, dsm_shape<T>(1).tyo>()
                       : shape == 2 ? dsm_per_cu<T, dsm_shape<T>(2).txo, dsm_shape<T>(2).tyo>()
                                    : dsm_per_cu<T, dsm_shape<T>(3).txo, dsm_shape<T>(3).tyo>();
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, device) != hipSuccess || ncu < 1) ncu = 1;
    const long tiles = (long)cdiv(d.nxg - 2, sh.txo) * cdiv(d.nyg - 2, sh.tyo);
    return round_filling_chunk(tiles, d.nzg - 2, (long)per_cu * ncu, 8, DSM_KMAX);
}

template <typename T>
int dsm_fused_nblk(const Dims& d, int kchunk, int shape)
{
    return cdiv(d.nxg - 2, dsm_shape<T>(shape).txo) * cdiv(d.nzg - 2, kchunk);
}

template <typename T>
void run_dsm_fused(const DsmArgs<T>& one, const DsmArgs<T>* tab, int nmem, const Dims& d, int kchunk, int shape,
                   hipStream_t s)
{
    if (kchunk < 1 || kchunk > DSM_KMAX) {
        fprintf(stderr, "chanles: DSM z chunk %d outside 1..%d\n", kchunk, DSM_KMAX);
        abort();
    }
    PROF_BEGIN("k_dsm_fused", s);
#define DSM_GO(Q)                                                                                                   \
    do {                                                                                                            \
        constexpr DsmShape sh = dsm_shape<T>(Q);                                                                    \
        dim3 blk((sh.txo + 2) * (sh.tyo + 2));                                                                      \
        dim3 grd(cdiv(d.nxg - 2, sh.txo), cdiv(d.nyg - 2, sh.tyo), (unsigned)cdiv(d.nzg - 2, kchunk) * nmem);       \
        if (tab) k_dsm_fused<T, sh.txo, sh.tyo, true><<<grd, blk, 0, s>>>(one, tab, d, kchunk);                     \
        else k_dsm_fused<T, sh.txo, sh.tyo, false><<<grd, blk, 0, s>>>(one, tab, d, kchunk);                        \
    } while (0)
    if (shape == 3) DSM_GO(3);
    else if (shape == 2) DSM_GO(2);
    else if (shape == 1) DSM_GO(1);
    else DSM_GO(0);
#undef DSM_GO
    PROF_END(s);
}

template <typename T>
void launch_dsm_fused(const T* U, const T* V, const T* W, const Metric<T>& m, const Dims& d, DsmBufs<T>& b,
                      hipStream_t s)
{
    (void)m;
    run_dsm_fused<T>(DsmArgs<T>{U, V, W, b}, nullptr, 1, d, b.kchunk, b.shape, s);
}

template int dsm_fused_kchunk<double>(const Dims&, int, int);
template int dsm_fused_kchunk<float>(const Dims&, int, int);
template <typename T>
int dsm_fused_blocks(const Dims& d, int kchunk, int shape)
{
    return cdiv(d.nxg - 2, dsm_shape<T>(shape).txo) * cdiv(d.nyg - 2, dsm_shape<T>(shape).tyo) *
           cdiv(d.nzg - 2, kchunk);
}

template int dsm_fused_blocks<double>(const Dims&, int, int);
template int dsm_fused_blocks<float>(const Dims&, int, int);
template int dsm_fused_nblk<double>(const Dims&, int, int);
template int dsm_fused_nblk<float>(const Dims&, int, int);
template void launch_dsm_fused<double>(const double*, const double*, const double*, const Metric<double>&,
                                       const Dims&, DsmBufs<double>&, hipStream_t);
template void launch_dsm_fused<float>(const float*, const float*, const float*, const Metric<float>&, const Dims&,
                                      DsmBufs<float>&, hipStream_t);
template void run_dsm_fused<double>(const DsmArgs<double>&, const DsmArgs<double>*, int, const Dims&, int, int,
                                    hipStream_t);
template void run_dsm_fused<float>(const DsmArgs<float>&, const DsmArgs<float>*, int, const Dims&, int, int,
                                   hipStream_t);

}  // namespace chan
