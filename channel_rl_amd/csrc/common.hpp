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

// Device scalars shared by the kernels of one channel (dt and dPdx are read
// from here, so a captured step needs no host round trip).
struct Scalars {
    double dt, t, dPdx, Qflow_x_0;
    double cfl, nu, dt_max, pad;
    double maxdiv, maxabsdiv, last_dPdx, pad2;
};

inline int cdiv(long a, long b) { return (int)((a + b - 1) / b); }

}  // namespace chan
