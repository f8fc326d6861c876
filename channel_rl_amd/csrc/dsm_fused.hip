// dsm_fused.hip -- dynamic Smagorinsky pass A in one kernel
// (code/subgrid.f90:379-510): centre velocities, rate of strain |S| S_ij,
// filtered staggered velocities, the Leonard tensor L_ij, the model tensor
// M_ij and the per-row partial sums of L:M and M:M, plus |S| for pass B.
//
// Design: a block owns an x-y tile of TXO x TYO output cells and marches in
// z over a chunk of planes.  Threads map one-to-one onto the tile plus a halo
// of one ("columns").  Per plane:
//   level 0  U,V,W of the tile plus a halo of two, three planes, in LDS,
//            loaded with periodic index wrap (no ghost cells are read);
//   level 1  per column: centre velocities a, |S| S_ij (in registers, three
//            planes), and the filtered staggered velocities (in LDS, three
//            planes);
//   filter   the test filter is separable: the z taps are combined in
//            registers, then the x and y taps read the combined plane from
//            LDS (y weights one-sided on rows 2 and nyg-1 as the reference).
// Global traffic: U,V,W read once per tile (plus halo), |S| written once.
#include "kernels.hpp"
#include "prof.hpp"

namespace chan {

template <typename T, int TXO, int TYO>
struct DsmTile {
    static constexpr int CX = TXO + 2, CY = TYO + 2, NC = CX * CY;  // columns
    static constexpr int LX = TXO + 4, LY = TYO + 4, NL = LX * LY;  // level-0 plane
};

template <typename T, int TXO, int TYO>
__global__ void __launch_bounds__((TXO + 2) * (TYO + 2))
    k_dsm_fused(const T* __restrict__ U, const T* __restrict__ V, const T* __restrict__ W, Metric<T> m, Dims d,
                DsmBufs<T> b, int kchunk)
{
    using Tl = DsmTile<T, TXO, TYO>;
    constexpr int CX = Tl::CX, NC = Tl::NC, LX = Tl::LX, NL = Tl::NL;
    __shared__ T l0[3][3][NL];  // [slot][U,V,W][position]
    __shared__ T lf[3][3][NC];  // [slot][Uf,Vf,Wf][column]
    __shared__ T zq[15][NC];    // z-combined filter inputs: a(3), a_a a_b (6), |S|S_ij (6)

    const int tid = threadIdx.x;
    const int cx = tid % CX, cy = tid / CX;
    const int i0 = blockIdx.x * TXO + 2, j0 = blockIdx.y * TYO + 2;
    const int kb = blockIdx.z * kchunk + 2;
    const int ke = min(kb + kchunk - 1, d.nzg - 1);
    const int ci = i0 - 1 + cx, cj = j0 - 1 + cy;  // column position (x unwrapped)
    const int nxp = d.nxp, nzp = d.nzp;
    auto wx = [&](int q) { int r = (q - 2) % nxp; return 2 + (r < 0 ? r + nxp : r); };
    auto wz = [&](int q) { int r = (q - 2) % nzp; return 2 + (r < 0 ? r + nzp : r); };
    const T h = T(0.5), qq = T(0.25);
    const T s6 = T(1.0 / 6.0), s23 = T(2.0 / 3.0), s13 = T(1.0 / 3.0);
    const T fil[3][3] = {{T(1.0 / 36.0), T(1.0 / 9.0), T(1.0 / 36.0)},
                         {T(1.0 / 9.0), T(4.0 / 9.0), T(1.0 / 9.0)},
                         {T(1.0 / 36.0), T(1.0 / 9.0), T(1.0 / 36.0)}};

    auto slot0 = [&](int p) { return (p - kb + 3) % 3; };  // level-0 ring slot of plane p >= kb-2
    auto slotf = [&](int p) { return (p - kb + 3) % 3; };  // filtered-velocity ring slot

    auto load_plane = [&](int p) {
        const int s = slot0(p), pz = wz(p);
        for (int q = tid; q < NL; q += NC) {
            const int lx = q % LX, ly = q / LX;
            const int ii = wx(i0 - 2 + lx), jj = j0 - 2 + ly;
            T u = T(0), v = T(0), w = T(0);
            if (jj >= 1 && jj <= d.nyg) {
                const long o = at(d, ii, jj, pz);
                u = U[o];
                w = W[o];
                if (jj <= d.ny) v = V[o];
            }
            l0[s][0][q] = u;
            l0[s][1][q] = v;
            l0[s][2][q] = w;
        }
    };
    // level-0 accessor: field f at column offset (dx,dy) on plane p
    auto L0 = [&](int f, int p, int dx, int dy) -> T { return l0[slot0(p)][f][(cx + 1 + dx) + LX * (cy + 1 + dy)]; };

    // level 1 of this column at plane p: a (3), |S| S_ij (6), |S|; filtered velocities to LDS
    auto level1 = [&](int p, T a[3], T ss[6], T& sm) {
        const int pc = wz(p), ic = wx(ci);
        for (int c = 0; c < 3; c++) a[c] = T(0);
        for (int c = 0; c < 6; c++) ss[c] = T(0);
        sm = T(0);
        const bool valid_row = cj >= 2 && cj <= d.nyg - 1;
        if (valid_row) {
            // code/subgrid.f90:394-396 centre velocities
            a[0] = h * (L0(0, p, -1, 0) + L0(0, p, 0, 0));
            a[1] = h * (L0(1, p, 0, -1) + L0(1, p, 0, 0));
            a[2] = h * (L0(2, p - 1, 0, 0) + L0(2, p, 0, 0));
            // code/subgrid.f90:1027-1070 at the canonical image (ic, cj, pc)
            const T *x = m.x, *y = m.y, *z = m.z, *xg = m.xg, *yg = m.yg, *zg = m.zg;
            const int jj = cj;
            T s[6];
            s[0] = (L0(0, p, 0, 0) - L0(0, p, -1, 0)) / (x[ic] - x[ic - 1]);
            s[1] = (L0(1, p, 0, 0) - L0(1, p, 0, -1)) / (y[jj] - y[jj - 1]);
            s[2] = (L0(2, p, 0, 0) - L0(2, p - 1, 0, 0)) / (z[pc] - z[pc - 1]);
            // e4(a,b) at edge (ic+a, jj+b), a,b in {0,1}
            auto e4 = [&](int da, int db) {
                return h * ((L0(0, p, da - 1, db) - L0(0, p, da - 1, db - 1)) / (yg[jj + db] - yg[jj + db - 1]) +
                            (L0(1, p, da, db - 1) - L0(1, p, da - 1, db - 1)) / (xg[ic + da] - xg[ic + da - 1]));
            };
            auto e5 = [&](int da, int dc) {
                return h * ((L0(0, p + dc, da - 1, 0) - L0(0, p + dc - 1, da - 1, 0)) / (zg[pc + dc] - zg[pc + dc - 1]) +
                            (L0(2, p + dc - 1, da, 0) - L0(2, p + dc - 1, da - 1, 0)) / (xg[ic + da] - xg[ic + da - 1]));
            };
            auto e6 = [&](int db, int dc) {
                return h * ((L0(1, p + dc, 0, db - 1) - L0(1, p + dc - 1, 0, db - 1)) / (zg[pc + dc] - zg[pc + dc - 1]) +
                            (L0(2, p + dc - 1, 0, db) - L0(2, p + dc - 1, 0, db - 1)) / (yg[jj + db] - yg[jj + db - 1]));
            };
            s[3] = qq * (e4(0, 0) + e4(1, 0) + e4(0, 1) + e4(1, 1));
            s[4] = qq * (e5(0, 0) + e5(1, 0) + e5(0, 1) + e5(1, 1));
            s[5] = qq * (e6(0, 0) + e6(1, 0) + e6(0, 1) + e6(1, 1));
            T acc = T(0);
            acc = acc + T(2) * s[0] * s[0];
            acc = acc + T(2) * s[1] * s[1];
            acc = acc + T(2) * s[2] * s[2];
            acc = acc + T(4) * s[3] * s[3];
            acc = acc + T(4) * s[4] * s[4];
            acc = acc + T(4) * s[5] * s[5];
            sm = sqrt(acc);
            for (int c = 0; c < 6; c++) ss[c] = sm * s[c];
        }
        // code/subgrid.f90:438-440 filtered staggered velocities (full y range)
        const int sf = slotf(p);
        for (int f = 0; f < 3; f++) {
            const int n2 = (f == 1) ? d.ny : d.nyg;
            T acc = T(0);
            if (cj >= 1 && cj <= n2) {
                for (int da = -1; da <= 1; da++)
                    for (int dc = -1; dc <= 1; dc++) {
                        T yv;
                        if (cj == 1)
                            yv = s23 * L0(f, p + dc, da, 0) + s13 * L0(f, p + dc, da, 1);
                        else if (cj == n2)
                            yv = s13 * L0(f, p + dc, da, -1) + s23 * L0(f, p + dc, da, 0);
                        else
                            yv = s6 * L0(f, p + dc, da, -1) + s23 * L0(f, p + dc, da, 0) + s6 * L0(f, p + dc, da, 1);
                        acc = acc + yv * fil[da + 1][dc + 1];
                    }
            }
            lf[sf][f][tid] = acc;
        }
    };

    // pre-roll: level 0 planes kb-2..kb, level 1 at kb-1 and kb
    T am[3], sm_m[6], a0[3], s0[6], ap[3], sp[6];
    T smag_m, smag_0, smag_p;
    load_plane(kb - 2);
    load_plane(kb - 1);
    load_plane(kb);
    __syncthreads();
    level1(kb - 1, am, sm_m, smag_m);
    __syncthreads();
    load_plane(kb + 1);
    __syncthreads();
    level1(kb, a0, s0, smag_0);

    const bool out_col = cx >= 1 && cx <= TXO && cy >= 1 && cy <= TYO && ci <= d.nxg - 1 && cj <= d.nyg - 1;
    double acc_lm = 0.0, acc_mm = 0.0;
    const int pa[6] = {0, 1, 2, 0, 0, 1}, pb[6] = {0, 1, 2, 1, 2, 2};

    for (int k = kb; k <= ke; k++) {
        __syncthreads();
        load_plane(k + 2);
        __syncthreads();
        level1(k + 1, ap, sp, smag_p);
        // z taps of the filter inputs
        for (int c = 0; c < 3; c++) zq[c][tid] = s6 * am[c] + s23 * a0[c] + s6 * ap[c];
        for (int c = 0; c < 6; c++) {
            zq[3 + c][tid] = s6 * (am[pa[c]] * am[pb[c]]) + s23 * (a0[pa[c]] * a0[pb[c]]) + s6 * (ap[pa[c]] * ap[pb[c]]);
            zq[9 + c][tid] = s6 * sm_m[c] + s23 * s0[c] + s6 * sp[c];
        }
        __syncthreads();
        if (out_col) {
            const int j = cj;
            const T wym = (j == 2) ? T(0) : (j == d.nyg - 1 ? s13 : s6);
            const T wy0 = (j == 2 || j == d.nyg - 1) ? s23 : s23;
            const T wyp = (j == d.nyg - 1) ? T(0) : (j == 2 ? s13 : s6);
            const T wyv[3] = {wym, wy0, wyp};
            const T wxv[3] = {s6, s23, s6};
            T F[15];
            for (int q = 0; q < 15; q++) F[q] = T(0);
            for (int dy = -1; dy <= 1; dy++) {
                if (wyv[dy + 1] == T(0)) continue;
                for (int dx = -1; dx <= 1; dx++) {
                    const T wgt = wxv[dx + 1] * wyv[dy + 1];
                    const int col = (cx + dx) + CX * (cy + dy);
                    for (int q = 0; q < 15; q++) F[q] = F[q] + wgt * zq[q][col];
                }
            }
            // hat(S) from the filtered velocities at the true index (code/subgrid.f90:456)
            const T *x = m.x, *y = m.y, *z = m.z, *xg = m.xg, *yg = m.yg, *zg = m.zg;
            const int i = ci;
            auto uf = [&](int dx, int dy, int dz) { return lf[slotf(k + dz)][0][(cx + dx) + CX * (cy + dy)]; };
            auto vf = [&](int dx, int dy, int dz) { return lf[slotf(k + dz)][1][(cx + dx) + CX * (cy + dy)]; };
            auto wf = [&](int dx, int dy, int dz) { return lf[slotf(k + dz)][2][(cx + dx) + CX * (cy + dy)]; };
            T sh[6];
            sh[0] = (uf(0, 0, 0) - uf(-1, 0, 0)) / (x[i] - x[i - 1]);
            sh[1] = (vf(0, 0, 0) - vf(0, -1, 0)) / (y[j] - y[j - 1]);
            sh[2] = (wf(0, 0, 0) - wf(0, 0, -1)) / (z[k] - z[k - 1]);
            auto e4 = [&](int da, int db) {
                return h * ((uf(da - 1, db, 0) - uf(da - 1, db - 1, 0)) / (yg[j + db] - yg[j + db - 1]) +
                            (vf(da, db - 1, 0) - vf(da - 1, db - 1, 0)) / (xg[i + da] - xg[i + da - 1]));
            };
            auto e5 = [&](int da, int dc) {
                return h * ((uf(da - 1, 0, dc) - uf(da - 1, 0, dc - 1)) / (zg[k + dc] - zg[k + dc - 1]) +
                            (wf(da, 0, dc - 1) - wf(da - 1, 0, dc - 1)) / (xg[i + da] - xg[i + da - 1]));
            };
            auto e6 = [&](int db, int dc) {
                return h * ((vf(0, db - 1, dc) - vf(0, db - 1, dc - 1)) / (zg[k + dc] - zg[k + dc - 1]) +
                            (wf(0, db, dc - 1) - wf(0, db - 1, dc - 1)) / (yg[j + db] - yg[j + db - 1]));
            };
            sh[3] = qq * (e4(0, 0) + e4(1, 0) + e4(0, 1) + e4(1, 1));
            sh[4] = qq * (e5(0, 0) + e5(1, 0) + e5(0, 1) + e5(1, 1));
            sh[5] = qq * (e6(0, 0) + e6(1, 0) + e6(0, 1) + e6(1, 1));
            T acc = T(0);
            acc = acc + T(2) * sh[0] * sh[0];
            acc = acc + T(2) * sh[1] * sh[1];
            acc = acc + T(2) * sh[2] * sh[2];
            acc = acc + T(4) * sh[3] * sh[3];
            acc = acc + T(4) * sh[4] * sh[4];
            acc = acc + T(4) * sh[5] * sh[5];
            const T shm = sqrt(acc);
            T L[6], M[6];
            for (int q = 0; q < 6; q++) {
                L[q] = F[3 + q] - F[pa[q]] * F[pb[q]];
                M[q] = T(4) * shm * sh[q] - F[9 + q];
            }
            T lm = L[0] * M[0] + L[1] * M[1] + L[2] * M[2] + T(2) * (L[3] * M[3] + L[4] * M[4] + L[5] * M[5]);
            T mm = M[0] * M[0] + M[1] * M[1] + M[2] * M[2] + T(2) * (M[3] * M[3] + M[4] * M[4] + M[5] * M[5]);
            acc_lm += (double)lm;
            acc_mm += (double)mm;
            b.smag[at(d, i, j, k)] = smag_0;
        }
        // rotate the level-1 ring
        for (int c = 0; c < 3; c++) { am[c] = a0[c]; a0[c] = ap[c]; }
        for (int c = 0; c < 6; c++) { sm_m[c] = s0[c]; s0[c] = sp[c]; }
        smag_m = smag_0;
        smag_0 = smag_p;
    }
    // per-row partial sums over the block's columns, in a fixed order
    __syncthreads();
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
            const int blk = blockIdx.z * gridDim.x + blockIdx.x;
            b.part[((long)(j - 1) * 2 + 0) * b.nblk + blk] = sl;
            b.part[((long)(j - 1) * 2 + 1) * b.nblk + blk] = sm2;
        }
    }
    (void)smag_m;
}

static constexpr int DSM_TXO = 30, DSM_TYO = 6, DSM_KCHUNK = 16;

int dsm_fused_nblk(const Dims& d)
{
    return cdiv(d.nxg - 2, DSM_TXO) * cdiv(d.nzg - 2, DSM_KCHUNK);
}

template <typename T>
void launch_dsm_fused(const T* U, const T* V, const T* W, const Metric<T>& m, const Dims& d, DsmBufs<T>& b,
                      hipStream_t s)
{
    dim3 blk((DSM_TXO + 2) * (DSM_TYO + 2));
    dim3 grd(cdiv(d.nxg - 2, DSM_TXO), cdiv(d.nyg - 2, DSM_TYO), cdiv(d.nzg - 2, DSM_KCHUNK));
    PROF_BEGIN("k_dsm_fused", s);
    k_dsm_fused<T, DSM_TXO, DSM_TYO><<<grd, blk, 0, s>>>(U, V, W, m, d, b, DSM_KCHUNK);
    PROF_END(s);
}

template void launch_dsm_fused<double>(const double*, const double*, const double*, const Metric<double>&,
                                       const Dims&, DsmBufs<double>&, hipStream_t);
template void launch_dsm_fused<float>(const float*, const float*, const float*, const Metric<float>&, const Dims&,
                                      DsmBufs<float>&, hipStream_t);

}  // namespace chan
