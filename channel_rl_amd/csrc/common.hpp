// common.hpp -- device-side layout and helpers shared by the channel kernels.
//
// Every 3-D field lives in one "box" layout: x fastest with pitch sx >= nxg,
// then y (nyg rows), then z (nzg planes), so U (nx,nyg,nzg), V (nxg,ny,nzg),
// W (nxg,nyg,nz) and the centre fields (nxg,nyg,nzg) share one index map and
// every stencil uses the same offsets.  Formulas use the reference's 1-based
// Fortran indices (code/initialization.f90:91-136).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace chan {

struct Dims {
    int nx, ny, nz;       // face counts (reference input line)
    int nxg, nyg, nzg;    // centres + 2 ghosts
    int nxp, nzp, nkx;    // unique periodic points, kept x modes (nxp/2+1)
    int sx;               // x pitch of the box (elements)
    long sxy;             // plane pitch sx*nyg
    long nbox;            // sxy*nzg
};

__host__ __device__ inline long at(const Dims& d, int i, int j, int k)
{
    return (long)(i - 1) + (long)d.sx * ((long)(j - 1) + (long)d.nyg * (long)(k - 1));
}

// Grid metrics, 1-based (element 0 unused), in the arithmetic type T.
template <typename T>
struct Metric {
    const T *x, *y, *z;      // faces
    const T *xg, *yg, *zg;   // centres + ghosts
    const T *w0, *w1;        // weight_y_0/1 (code/initialization.f90:398-400)
    const T *rdx, *rdy, *rdz;     // reciprocal face spacings 1/(x(i)-x(i-1)), ...
    const T *rdxg, *rdyg, *rdzg;  // reciprocal centre spacings 1/(xg(i)-xg(i-1)), ...
};

// Periodic image maps (code/boundary_conditions.f90:60-79; z as its global
// pattern).  Faces: 1 <- n-1, n <- 2.  Centres: 1 <- ng-2, ng-1 <- 2, ng <- 3.
__host__ __device__ inline int img_face(int i, int n)
{
    return i == 1 ? n - 1 : (i == n ? 2 : i);
}
__host__ __device__ inline int img_centre(int i, int ng)
{
    return i == 1 ? ng - 2 : (i == ng - 1 ? 2 : (i == ng ? 3 : i));
}

// Store v at the canonical cell (i,j,k) of a velocity field and at every
// ghost cell whose boundary-condition source it is, so that a producer
// kernel leaves the field exactly as code/boundary_conditions.f90:20-48
// would: periodic images in x and z, and in y the wall mirror -v for U and
// W (V's wall rows stay 0).  kind: 0 U (x faces), 1 V, 2 W (z faces).
template <typename T>
__device__ __forceinline__ void store_img(T* F, const Dims& d, int kind, int i, int j, int k, T v)
{
    int xs[2], zs[2], ys[2];
    T sg[2];
    int nx_ = 1, nz_ = 1, ny_ = 1;
    xs[0] = i;
    if (kind == 0) {  // faces: 1 <- nx-1, nx <- 2
        if (i == d.nx - 1) xs[nx_++] = 1;
        if (i == 2) xs[nx_++] = d.nx;
    } else {          // centres: 1 <- nxg-2, nxg-1 <- 2, nxg <- 3
        if (i == d.nxg - 2) xs[nx_++] = 1;
        if (i == 2) xs[nx_++] = d.nxg - 1;
        if (i == 3) xs[nx_++] = d.nxg;
    }
    zs[0] = k;
    if (kind == 2) {
        if (k == d.nz - 1) zs[nz_++] = 1;
        if (k == 2) zs[nz_++] = d.nz;
    } else {
        if (k == d.nzg - 2) zs[nz_++] = 1;
        if (k == 2) zs[nz_++] = d.nzg - 1;
        if (k == 3) zs[nz_++] = d.nzg;
    }
    ys[0] = j;
    sg[0] = v;
    if (kind != 1) {
        if (j == 2) { ys[ny_] = 1; sg[ny_++] = -v; }
        if (j == d.nyg - 1) { ys[ny_] = d.nyg; sg[ny_++] = -v; }
    }
    for (int a = 0; a < nx_; a++)
        for (int c = 0; c < nz_; c++)
            for (int b = 0; b < ny_; b++) F[at(d, xs[a], ys[b], zs[c])] = sg[b];
}

// Device scalars shared by the kernels of one channel (dt and dPdx are read
// from here, so a captured step needs no host round trip).
struct Scalars {
    double dt, t, dPdx, Qflow_x_0;
    double cfl, nu, dt_max;
    double wm_nu_limit;  // > 0: the wall model's viscosity limit (see wm_nut_at); 0: none
    double maxdiv, maxabsdiv, last_dPdx, pad2;
};

// code/time_integration.f90:123-148 from the three maxima of dx/|u|: the
// convective and viscous limits, the cap, a negative CFL as a fixed step
__device__ inline double dt_from_maxima(double mu, double mv, double mw, const Scalars* sc, const double* mins3)
{
    const double cfl = sc->cfl, nu = sc->nu;
    double dconv = fmin(fmin(cfl * mu, cfl * mv), cfl * mw);
    double dvis = fmin(fmin(cfl * (mins3[0] * mins3[0]) / nu, cfl * (mins3[1] * mins3[1]) / nu),
                       cfl * (mins3[2] * mins3[2]) / nu);
    double dt = fmin(dconv, dvis);
    dt = fmin(dt, sc->dt_max);
    if (cfl < 0) dt = -cfl;
    return dt;
}

inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

// A linear cell index t = i + n1 (j + n2 k) (0-based, x fastest) walked in
// steps of a fixed stride S: with S = sr n1 + s1 and sr = sk n2 + sj the
// next cell is (i + s1, j + sj, k + sk) plus at most one carry out of i and
// one out of j, so a loop needs no division per cell.
struct CellIdx { int i, j, k; };
struct StrideCarry { int s1, sj, sk; };
__host__ __device__ inline CellIdx cell_of(int t, int n1, int n2)
{
    const int r = t / n1;
    return CellIdx{t - r * n1, r % n2, r / n2};
}
__host__ __device__ inline StrideCarry stride_carry(int S, int n1, int n2)
{
    const int sr = S / n1;
    return StrideCarry{S % n1, sr % n2, sr / n2};
}
__host__ __device__ inline CellIdx advance(CellIdx c, const StrideCarry& s, int n1, int n2)
{
    c.i += s.s1;
    const int ci = c.i >= n1;
    if (ci) c.i -= n1;
    c.j += s.sj + ci;
    const int cj = c.j >= n2;
    if (cj) c.j -= n2;
    c.k += s.sk + cj;
    return c;
}

// A 256-thread block over x rows of n points: 64 x 4, or 32 x 8 where a
// 64-wide tile would leave more lanes idle (96 points: a half-empty second
// tile, a quarter of the launch's lanes).
inline dim3 row_block(int n)
{
    return (cdiv(n, 32) * 32 < cdiv(n, 64) * 64) ? dim3(32, 8, 1) : dim3(64, 4, 1);
}

// z-chunk length in [lo, hi] for a kernel whose blocks each march `kc` of
// nk planes over one of `tiles` (x,y) tiles, with two planes of pre-roll:
// blocks run in rounds of `slots` (resident blocks per CU x CUs), so the
// chunk that minimises rounds x (planes + 2) keeps the last round as full as
// it can be.
inline int round_filling_chunk(long tiles, int nk, long slots, int lo, int hi)
{
    int best = lo;
    long best_cost = -1;
    for (int kc = lo; kc <= hi; kc++) {
        const long rounds = (tiles * cdiv(nk, kc) + slots - 1) / slots;
        const long cost = rounds * ((kc < nk ? kc : nk) + 2);
        if (best_cost < 0 || cost < best_cost) {
            best_cost = cost;
            best = kc;
        }
    }
    return best;
}

// Tile order across the 8 XCDs of MI355X.  The dispatcher deals workgroups
// to the XCDs in turn (workgroup n to XCD n mod 8) and each XCD has its own
// L2, so neighbouring tiles in launch order land in different L2s and their
// shared halo rows are fetched from HBM once per XCD.  Remapping the linear
// workgroup id so that each XCD receives a contiguous run of tiles (x
// fastest, then y, then z) keeps neighbours, and their halos, in one L2.
// A bijection on 0..n-1 for any n.
__device__ inline int xcd_linear()
{
    const int gx = gridDim.x, gy = gridDim.y;
    const int n = gx * gy * gridDim.z;
    const int h = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    const int q = n / 8, r = n % 8, x = h % 8;
    return x * q + (x < r ? x : r) + h / 8;
}
struct TileId { int x, y, z; };
__device__ inline TileId xcd_tile()
{
    const int gx = gridDim.x, gy = gridDim.y;
    const int l = xcd_linear();
    return TileId{l % gx, (l / gx) % gy, l / (gx * gy)};
}

// Grouped launches.  A group is several channels of one shape, each with its
// own buffers and scalars, advanced by one launch per stage: the launch's
// grid stacks the members' grids along z (gz planes of blocks each) and
// tab[m] holds member m's arguments in device memory.  A launch for one
// channel passes its arguments by value, tab = nullptr and gz = gridDim.z;
// its member is 0.  GBlk is a block's position within its member's grid.
struct GBlk {
    int x, y, z;     // block within the member's grid
    int gx, gy, gz;  // the member's grid
    int m;           // member
};

// hardware order: member = blockIdx.z / gz
__device__ inline GBlk gblk(int gz)
{
    const int z = blockIdx.z;
    return GBlk{(int)blockIdx.x, (int)blockIdx.y, z % gz, (int)gridDim.x, (int)gridDim.y, gz, z / gz};
}

// XCD-contiguous order (xcd_linear) over the whole launch, then split into
// member and tile (x fastest, then y, then z): every member's tiles stay
// contiguous, so a small member runs on one or two XCDs and its halos stay
// in their L2s.  For one channel this is xcd_tile().
__device__ inline GBlk gblk_xcd(int gz)
{
    const int gx = gridDim.x, gy = gridDim.y;
    const int l = xcd_linear();
    const int per = gx * gy * gz;
    const int m = l / per, t = l - m * per;
    return GBlk{t % gx, (t / gx) % gy, t / (gx * gy), gx, gy, gz, m};
}

// a member's arguments: its table entry in a grouped launch, else the launch's
template <typename A>
__device__ __forceinline__ A member_args(const A& one, const A* __restrict__ tab, int m)
{
    return tab ? tab[m] : one;
}
// the same chosen at compile time, for kernels instantiated per form: the
// one-channel form keeps its arguments in the kernel-argument segment, from
// which the compiler reloads them where they are used instead of holding
// them in registers across the kernel
template <bool G, typename A>
__device__ __forceinline__ A member_args(const A& one, const A* __restrict__ tab, int m)
{
    if constexpr (G) return tab[m];
    else return one;
}
// by reference, for arguments a kernel indexes at run time (an array inside
// the struct): a copy would live in scratch memory
template <bool G, typename A>
__device__ __forceinline__ const A& member_ref(const A& one, const A* __restrict__ tab, int m)
{
    if constexpr (G) return tab[m];
    else return one;
}

}  // namespace chan
