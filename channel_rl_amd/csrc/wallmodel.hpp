// wallmodel.hpp -- wall-stress closures shared by the equilibrium and the
// learned (MLP) wall models: the log law and the sampling of the flow at the
// matching height.
#pragma once
#include <cmath>
#include "common.hpp"

namespace chan {

// Log law U+ = ln(y+)/kappa + B above the crossover with U+ = y+, linear
// below; Newton on f(ut) = ut (ln(h ut/nu)/kappa + B) - |u|.
// yc is the crossover y+ where ln(y+)/kappa + B = y+ (computed once on the host).
// guess > 0 (finite) starts Newton there instead of at the viscous estimate:
// f is increasing and convex on the log branch, so either start converges to
// the same root (to the 1e-15 stopping test); a close guess saves iterations.
__device__ inline double loglaw_utau(double u, double h, double nu, double kappa, double B, double yc,
                                     double guess = 0.0)
{
    double au = fabs(u);
    if (au == 0.0) return 0.0;
    double ut = sqrt(nu * au / h);
    if (h * ut / nu <= yc) return ut;
    double g = (guess > 0.0 && guess < 1e300) ? guess : ut;
    for (int it = 0; it < 60; it++) {
        double lg = log(h * g / nu);
        double f = g * (lg / kappa + B) - au;
        double fp = lg / kappa + B + 1.0 / kappa;
        double gn = g - f / fp;
        gn = gn > 0.5 * g ? gn : 0.5 * g;
        if (fabs(gn - g) <= 1e-15 * gn) { g = gn; break; }
        g = gn;
    }
    return g;
}

// code/wallmodel.f90:32-37: wall-model eddy viscosity at a centre from the
// two adjacent x faces; wall = 0 bottom (rows 1,2), 1 top (rows nyg-1,nyg).
// With sc->wm_nu_limit = L > 0 (no reference counterpart) each face's total
// viscosity alpha/(dU/dy) is held in [0, L dy^2/dt], dy = yg(2)-yg(1): the
// formula asks for an unbounded viscosity where the resolved wall gradient
// vanishes and a negative one where it reverses, either of which breaks the
// explicit step (its diffusion number passes the stability bound); L = 0
// keeps the reference's formula as it is.
template <typename T>
__device__ inline T wm_nut_at(const T* U, const double* alpha, const Metric<T>& m, const Dims& d, T nu,
                                       int ic, int k, int wall, const Scalars* sc)
{
    const T dyw = m.yg[2] - m.yg[1];
    const double lim = sc ? sc->wm_nu_limit : 0.0;
    const T numax = lim > 0.0 ? T(lim * (double)dyw * (double)dyw / sc->dt) : T(0);
    auto tau = [&](int f) {
        const T a = T(alpha[(f - 1) + (long)d.nx * (wall + 2L * (k - 1))]);
        const T du = wall == 0 ? (U[at(d, f, 2, k)] - U[at(d, f, 1, k)])
                               : (U[at(d, f, d.nyg - 1, k)] - U[at(d, f, d.nyg, k)]);
        T tot = a / (du / dyw);
        if (lim > 0.0) tot = (tot >= T(0)) ? (tot <= numax ? tot : numax) : T(0);
        return tot - nu;
    };
    return T(0.5) * (tau(ic - 1) + tau(ic));
}

inline double loglaw_crossover(double kappa, double B)
{
    double yc = 11.0;
    for (int it = 0; it < 50; it++) yc = yc - (yc - std::log(yc) / kappa - B) / (1.0 - 1.0 / (kappa * yc));
    return yc;
}

// The flow seen by a wall cell at the matching height.  Wall cells are the
// entries of alpha_x (nx, 2, nzg): x face i, z centre k, wall 0 (bottom) or
// 1 (top) (code/wallmodel.f90:23-46).  U lives there; W (x centres, z faces)
// is averaged to that point.  Below/above the height the two bracketing
// centre rows are interpolated linearly; h_wm <= 0 takes the first off-wall
// centre (the reference's wall-model point).
struct WallSample {
    double u, w;      // wall-parallel velocity components at the height
    double dudy;      // d|u_par|/dn between the bracketing rows, n away from the wall
    double h;         // the height
};

template <typename T>
__device__ inline WallSample wall_sample(const T* U, const T* W, const Metric<double>& m, const Dims& d, int i, int k,
                                         int wall, double h_wm)
{
    const double *y = m.y, *yg = m.yg;
    int jb = 2;
    double wb = 1.0, hb = yg[2] - y[1];
    if (h_wm > 0.0) {
        const double yt = y[1] + h_wm;
        while (jb < d.nyg - 2 && yg[jb + 1] < yt) jb++;
        wb = (yg[jb + 1] - yt) / (yg[jb + 1] - yg[jb]);
        hb = h_wm;
    }
    // z centre k as its canonical image for the W average over faces k-1, k
    const int kc = k == 1 ? d.nzg - 2 : (k == d.nzg ? 3 : k);
    auto uw_at = [&](int j, double& u, double& w) {
        u = (double)U[at(d, i, j, k)];
        w = 0.25 * ((double)W[at(d, i, j, kc - 1)] + (double)W[at(d, i, j, kc)] + (double)W[at(d, i + 1, j, kc - 1)] +
                    (double)W[at(d, i + 1, j, kc)]);
    };
    const int ja = wall == 0 ? jb : d.nyg + 1 - jb;       // nearer row
    const int jn = wall == 0 ? jb + 1 : d.nyg - jb;       // farther row
    double ua, wa, un, wn;
    uw_at(ja, ua, wa);
    uw_at(jn, un, wn);
    WallSample s;
    s.u = wb * ua + (1.0 - wb) * un;
    s.w = wb * wa + (1.0 - wb) * wn;
    const double dn = wall == 0 ? yg[jb + 1] - yg[jb] : yg[ja] - yg[jn];
    s.dudy = (sqrt(un * un + wn * wn) - sqrt(ua * ua + wa * wa)) / dn;
    s.h = (h_wm > 0.0 || wall == 0) ? hb : y[d.ny] - yg[d.nyg - 1];
    return s;
}

// Where the equilibrium wall model samples U: h_wm above each wall,
// interpolated between centre rows jb and jb+1 with weight wb on row jb
// (h_wm = 0: row 2 itself, wb = 1); hb is the bottom sampling height, ht the
// top one.  Depends on the grid only, so the host computes it.
struct WallRows {
    int jb;
    double wb, hb, ht;
};

// The equilibrium wall stress of one wall cell, q = (x face) + nx ((z centre)
// + nzg wall), from U at the matching height, in alpha_x (code/wallmodel.f90:23-46)
template <typename T>
__device__ inline void eq_alpha_cell(int q, const T* U, double* alpha, const Dims& d, double nu, double kappa,
                                     double B, double yc, const WallRows& ws)
{
    const int i = q % d.nx + 1, r = q / d.nx;
    const int k = r % d.nzg + 1, wall = r / d.nzg;
    const int jb = ws.jb;
    const double wb = ws.wb;
    double uw, h;
    if (wall == 0) {
        uw = wb * (double)U[at(d, i, jb, k)] + (1.0 - wb) * (double)U[at(d, i, jb + 1, k)];
        h = ws.hb;
    } else {
        const int jt = d.nyg + 1 - jb;
        uw = wb * (double)U[at(d, i, jt, k)] + (1.0 - wb) * (double)U[at(d, i, jt - 1, k)];
        h = ws.ht;
    }
    // Newton starts from the cell's previous u_tau (|alpha_x| = u_tau^2): the
    // flow moves little in a step, so a few iterations reach the same root
    const long ia = (i - 1) + (long)d.nx * (wall + 2L * (k - 1));
    const double t = loglaw_utau(uw, h, nu, kappa, B, yc, sqrt(fabs(alpha[ia])));
    alpha[ia] = (uw >= 0 ? 1.0 : -1.0) * t * t;
}

}  // namespace chan
