/*
 * oracle.c -- FP64 CPU restatement of the channel_RL LES time step.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP library.  It is built by
 * oracle/build.sh into oracle/liborc.so and loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg, nowhere else.
 *
 * Every function restates one reference subroutine and cites it.  Loop
 * ranges, temporaries and the order of floating-point operations follow the
 * reference (compiled here without FMA contraction), so that the oracle is
 * the reference's arithmetic, not a re-derivation.  Where the reference runs
 * under MPI the whole-domain (decomposed) semantics are restated: periodic
 * images instead of halo messages, one sum instead of Allreduce.
 */
#include "oracle.h"

#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

/* Loops whose iterations write disjoint outputs, each output by the same
 * operations in the same order as the serial loop, run over OpenMP threads
 * when built with -fopenmp: the results are bit-identical for any thread
 * count.  Sums and maxima over the domain stay serial. */
#ifdef _OPENMP
#define PAR_FOR _Pragma("omp parallel for schedule(static)")
#define PAR_FOR2 _Pragma("omp parallel for collapse(2) schedule(static)")
#else
#define PAR_FOR
#define PAR_FOR2
#endif

/* Wall time per stage of the RK3 step and of compute_dt, accumulated over
 * calls (for the CPU baseline's per-stage report; not part of the checks):
 * 0 equilibrium wall stress, 1 eddy viscosity, 2 wall model, 3 right-hand
 * sides, 4 RK combine, 5 boundary conditions, 6 projection, 7 mass-flow
 * correction, 8 compute_dt. */
static double g_stage[ORC_NSTAGE];
static double wall_now(void)
{
#ifdef _OPENMP
    return omp_get_wtime();
#else
    struct timespec ts;
    clock_gettime(CLOCK_MONOTONIC, &ts);
    return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
#endif
}
void orc_stage_times(double* out, int reset)
{
    for (int q = 0; q < ORC_NSTAGE; q++) {
        if (out) out[q] = g_stage[q];
        if (reset) g_stage[q] = 0.0;
    }
}
#define STAGE(q, stmt)                         \
    do {                                       \
        const double t0_ = wall_now();         \
        stmt;                                  \
        g_stage[q] += wall_now() - t0_;        \
    } while (0)

int orc_num_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

struct orc {
    int nx, ny, nz, nxg, nyg, nzg, nxp, nzp;
    double *x, *y, *z, *xm, *ym, *zm, *xg, *yg, *zg;
    double dxmin, dymin, dzmin;
    double *w0, *w1;              /* weight_y_0/1, index 1..ny          */
    double *dyy_d, *dyy_l, *dyy_u; /* Dyy(j,j), Dyy(j+1,j), Dyy(j,j+1)  */
    double *kxx, *kzz;            /* 0..nxp-1, 0..nzp-1                  */
    double nu, CFL, dPdx, Qflow_x_0, dt, t;
    int x_mass_cte, LES_model, Dirichlet_nu_t;
    int wall_mode;                /* 1: equilibrium alpha_x before each substep */
    double h_wm, kappa, Bc;
    double rk_t[3], rk_coef[3][3];
    /* state */
    double *U, *V, *W, *nu_t, *P, *alpha_x;
    double *Uo, *Vo, *Wo;
    double *Fu[3], *Fv[3], *Fw[3];
    /* scratch (box = nxg*nyg*nzg) */
    double *t0, *t1, *t2, *t3;
    double *Uf, *Vf, *Wf, *Sbuf, *Sijb, *Lij, *Mij, *Sh;
};

/* 1-based Fortran index into an array of leading dims n1,n2 */
#define IX(i, j, k, n1, n2) \
    ((size_t)((i) - 1) + (size_t)(n1) * ((size_t)((j) - 1) + (size_t)(n2) * (size_t)((k) - 1)))
#define AU(a, i, j, k) (a)[IX(i, j, k, o->nx, o->nyg)]
#define AV(a, i, j, k) (a)[IX(i, j, k, o->nxg, o->ny)]
#define AW(a, i, j, k) (a)[IX(i, j, k, o->nxg, o->nyg)]
#define AC(a, i, j, k) (a)[IX(i, j, k, o->nxg, o->nyg)]   /* centre box */

static double* dalloc(size_t n) { return (double*)calloc(n ? n : 1, sizeof(double)); }

static size_t nU(const orc* o) { return (size_t)o->nx * o->nyg * o->nzg; }
static size_t nV(const orc* o) { return (size_t)o->nxg * o->ny * o->nzg; }
static size_t nW(const orc* o) { return (size_t)o->nxg * o->nyg * o->nz; }
static size_t nC(const orc* o) { return (size_t)o->nxg * o->nyg * o->nzg; }

/* ------------------------------------------------------------------ setup */

/* code/initialization.f90:187-240 (grids), 250-259 (dx=dxmin, nxp),
 * 293-308 (kxx,kzz), 355-385 (Dyy), 398-400 (weights), 415-427 (RK3). */
orc* orc_create(int nx, int ny, int nz, const double* x, const double* y, const double* z,
                double nu, double CFL, double dPdx, int x_mass_cte, int LES_model,
                int Dirichlet_nu_t)
{
    orc* o = (orc*)calloc(1, sizeof(orc));
    o->nx = nx; o->ny = ny; o->nz = nz;
    o->nxg = nx + 1; o->nyg = ny + 1; o->nzg = nz + 1;
    o->nxp = nx - 2; o->nzp = nz - 2;
    o->nu = nu; o->CFL = CFL; o->dPdx = dPdx;
    o->x_mass_cte = x_mass_cte; o->LES_model = LES_model; o->Dirichlet_nu_t = Dirichlet_nu_t;

    o->x = dalloc(nx + 1); o->y = dalloc(ny + 1); o->z = dalloc(nz + 1);
    o->xm = dalloc(nx + 1); o->ym = dalloc(ny + 1); o->zm = dalloc(nz + 1);
    o->xg = dalloc(nx + 2); o->yg = dalloc(ny + 2); o->zg = dalloc(nz + 2);
    for (int i = 1; i <= nx; i++) o->x[i] = x[i - 1];
    for (int j = 1; j <= ny; j++) o->y[j] = y[j - 1];
    for (int k = 1; k <= nz; k++) o->z[k] = z[k - 1];
    int nxm = nx - 1, nym = ny - 1, nzm = nz - 1;
    for (int i = 1; i <= nxm; i++) o->xm[i] = 0.5 * (o->x[i] + o->x[i + 1]);
    for (int j = 1; j <= nym; j++) o->ym[j] = 0.5 * (o->y[j] + o->y[j + 1]);
    for (int k = 1; k <= nzm; k++) o->zm[k] = 0.5 * (o->z[k] + o->z[k + 1]);
    for (int i = 1; i <= nxm; i++) o->xg[i + 1] = o->xm[i];
    o->xg[1] = o->xm[1] - 2.0 * (o->xm[1] - o->x[1]);
    o->xg[nxm + 2] = o->xm[nxm] + 2.0 * (o->x[nx] - o->xm[nxm]);
    for (int j = 1; j <= nym; j++) o->yg[j + 1] = o->ym[j];
    o->yg[1] = o->ym[1] - 2.0 * (o->ym[1] - o->y[1]);
    o->yg[nym + 2] = o->ym[nym] + 2.0 * (o->y[ny] - o->ym[nym]);
    for (int k = 1; k <= nzm; k++) o->zg[k + 1] = o->zm[k];
    o->zg[1] = o->zm[1] - 2.0 * (o->zm[1] - o->z[1]);
    o->zg[nzm + 2] = o->zm[nzm] + 2.0 * (o->z[nz] - o->zm[nzm]);

    o->dxmin = o->xg[2] - o->xg[1];
    for (int i = 2; i < o->nxg; i++) o->dxmin = fmin(o->dxmin, o->xg[i + 1] - o->xg[i]);
    o->dymin = o->yg[2] - o->yg[1];
    for (int j = 2; j < o->nyg; j++) o->dymin = fmin(o->dymin, o->yg[j + 1] - o->yg[j]);
    o->dzmin = o->zg[2] - o->zg[1];
    for (int k = 2; k < o->nzg; k++) o->dzmin = fmin(o->dzmin, o->zg[k + 1] - o->zg[k]);

    /* modified wavenumbers, dx = dxmin (uniform x/z assumed) */
    const double pi = 4.0 * atan(1.0);
    double dx = o->dxmin, dz = o->dzmin;
    o->kxx = dalloc(o->nxp); o->kzz = dalloc(o->nzp);
    int cx = (int)ceil((double)o->nxp / 2.0), cz = (int)ceil((double)o->nzp / 2.0);
    for (int i = 0; i <= cx && i < o->nxp; i++)
        o->kxx[i] = 2.0 * (cos(2.0 * pi * (double)i / (double)o->nxp) - 1.0) / (dx * dx);
    for (int i = cx + 1; i <= o->nxp - 1; i++)
        o->kxx[i] = 2.0 * (cos(2.0 * pi * (double)(-o->nxp + i) / (double)o->nxp) - 1.0) / (dx * dx);
    for (int k = 0; k <= cz && k < o->nzp; k++)
        o->kzz[k] = 2.0 * (cos(2.0 * pi * (double)k / (double)o->nzp) - 1.0) / (dz * dz);
    for (int k = cz + 1; k <= o->nzp - 1; k++)
        o->kzz[k] = 2.0 * (cos(2.0 * pi * (double)(-o->nzp + k) / (double)o->nzp) - 1.0) / (dz * dz);

    /* Dyy: interior rows, then homogeneous-Neumann wall rows */
    int nyg = o->nyg;
    o->dyy_d = dalloc(nyg + 2); o->dyy_l = dalloc(nyg + 2); o->dyy_u = dalloc(nyg + 2);
    double* y_ = o->y; double* yg_ = o->yg;
    for (int j = 3; j <= nyg - 2; j++) {
        double a = 1.0 / (y_[j] - y_[j - 1]) / (yg_[j + 1] - yg_[j]);
        double b = 1.0 / (y_[j] - y_[j - 1]) * (-1.0 / (yg_[j + 1] - yg_[j]) - 1.0 / (yg_[j] - yg_[j - 1]));
        double c = 1.0 / (y_[j] - y_[j - 1]) / (yg_[j] - yg_[j - 1]);
        o->dyy_u[j] = a;        /* Dyy(j,j+1) */
        o->dyy_l[j - 1] = c;    /* Dyy(j,j-1) stored at the column index j-1 */
        o->dyy_d[j] = b;
    }
    {
        int j = 2;
        double a = 1.0 / (y_[j] - y_[j - 1]) / (yg_[j + 1] - yg_[j]);
        double b = 1.0 / (y_[j] - y_[j - 1]) * (-1.0 / (yg_[j + 1] - yg_[j]) - 1.0 / (yg_[j] - yg_[j - 1]));
        double c = 1.0 / (y_[j] - y_[j - 1]) / (yg_[j] - yg_[j - 1]);
        o->dyy_d[2] = b + c;
        o->dyy_u[2] = a;
    }
    {
        int j = nyg - 1;
        double a = 1.0 / (y_[j] - y_[j - 1]) / (yg_[j + 1] - yg_[j]);
        double b = 1.0 / (y_[j] - y_[j - 1]) * (-1.0 / (yg_[j + 1] - yg_[j]) - 1.0 / (yg_[j] - yg_[j - 1]));
        double c = 1.0 / (y_[j] - y_[j - 1]) / (yg_[j] - yg_[j - 1]);
        o->dyy_d[nyg - 1] = a + b;
        o->dyy_l[nyg - 2] = c;
    }

    o->w0 = dalloc(ny + 1); o->w1 = dalloc(ny + 1);
    for (int j = 1; j <= ny; j++) {
        o->w0[j] = (yg_[j + 1] - y_[j]) / (yg_[j + 1] - yg_[j]);
        o->w1[j] = 1.0 - o->w0[j];
    }

    o->rk_t[0] = 8.0 / 15.0; o->rk_t[1] = 2.0 / 3.0; o->rk_t[2] = 1.0;
    memset(o->rk_coef, 0, sizeof(o->rk_coef));
    o->rk_coef[0][0] = 8.0 / 15.0;
    o->rk_coef[1][0] = 1.0 / 4.0; o->rk_coef[1][1] = 5.0 / 12.0;
    o->rk_coef[2][0] = 1.0 / 4.0; o->rk_coef[2][1] = 0.0; o->rk_coef[2][2] = 3.0 / 4.0;

    o->U = dalloc(nU(o)); o->V = dalloc(nV(o)); o->W = dalloc(nW(o));
    o->nu_t = dalloc(nC(o)); o->P = dalloc(nC(o));
    o->alpha_x = dalloc((size_t)nx * 2 * o->nzg);
    o->Uo = dalloc(nU(o)); o->Vo = dalloc(nV(o)); o->Wo = dalloc(nW(o));
    for (int s = 0; s < 3; s++) {
        o->Fu[s] = dalloc(nU(o)); o->Fv[s] = dalloc(nV(o)); o->Fw[s] = dalloc(nW(o));
    }
    size_t nc = nC(o);
    o->t0 = dalloc(nc); o->t1 = dalloc(nc); o->t2 = dalloc(nc); o->t3 = dalloc(nc);
    o->Uf = dalloc(nc); o->Vf = dalloc(nc); o->Wf = dalloc(nc); o->Sbuf = dalloc(nc);
    o->Sh = dalloc(nc);
    o->Sijb = dalloc(6 * nc); o->Lij = dalloc(6 * nc); o->Mij = dalloc(6 * nc);
    return o;
}

void orc_destroy(orc* o)
{
    if (!o) return;
    double* p[] = {o->x, o->y, o->z, o->xm, o->ym, o->zm, o->xg, o->yg, o->zg, o->w0, o->w1,
                   o->dyy_d, o->dyy_l, o->dyy_u, o->kxx, o->kzz, o->U, o->V, o->W, o->nu_t,
                   o->P, o->alpha_x, o->Uo, o->Vo, o->Wo, o->t0, o->t1, o->t2, o->t3, o->Uf,
                   o->Vf, o->Wf, o->Sbuf, o->Sh, o->Sijb, o->Lij, o->Mij};
    for (size_t i = 0; i < sizeof(p) / sizeof(p[0]); i++) free(p[i]);
    for (int s = 0; s < 3; s++) { free(o->Fu[s]); free(o->Fv[s]); free(o->Fw[s]); }
    free(o);
}

void orc_get_grid(const orc* o, double* xg, double* yg, double* zg, double* w0, double* w1,
                  double* mins3)
{
    for (int i = 1; i <= o->nxg; i++) xg[i - 1] = o->xg[i];
    for (int j = 1; j <= o->nyg; j++) yg[j - 1] = o->yg[j];
    for (int k = 1; k <= o->nzg; k++) zg[k - 1] = o->zg[k];
    for (int j = 1; j <= o->ny; j++) { w0[j - 1] = o->w0[j]; w1[j - 1] = o->w1[j]; }
    mins3[0] = o->dxmin; mins3[1] = o->dymin; mins3[2] = o->dzmin;
}

void orc_get_poisson(const orc* o, double* d, double* l, double* u, double* kxx, double* kzz)
{
    /* rows j = 2..nyg-1 packed at 0..ny-2; l,u hold Dyy(j+1,j), Dyy(j,j+1) for j=2..nyg-2 */
    for (int j = 2; j <= o->nyg - 1; j++) d[j - 2] = o->dyy_d[j];
    for (int j = 2; j <= o->nyg - 2; j++) { l[j - 2] = o->dyy_l[j]; u[j - 2] = o->dyy_u[j]; }
    memcpy(kxx, o->kxx, sizeof(double) * o->nxp);
    memcpy(kzz, o->kzz, sizeof(double) * o->nzp);
}

void orc_get_rk(const orc* o, double* rk_t3, double* rk_coef9)
{
    for (int s = 0; s < 3; s++) rk_t3[s] = o->rk_t[s];
    for (int s = 0; s < 3; s++)
        for (int q = 0; q < 3; q++) rk_coef9[3 * s + q] = o->rk_coef[s][q];
}

/* ------------------------------------------------------------------ state */

static double* field_ptr(const orc* o, int which)
{
    switch (which) {
    case 0: return o->U; case 1: return o->V; case 2: return o->W;
    case 3: return o->nu_t; case 4: return o->P; case 5: return o->alpha_x;
    }
    return 0;
}

long orc_field_size(const orc* o, int which)
{
    switch (which) {
    case 0: return (long)nU(o); case 1: return (long)nV(o); case 2: return (long)nW(o);
    case 3: case 4: return (long)nC(o);
    case 5: return (long)o->nx * 2 * o->nzg;
    }
    return 0;
}

void orc_set_field(orc* o, int which, const double* src)
{
    memcpy(field_ptr(o, which), src, sizeof(double) * (size_t)orc_field_size(o, which));
}

void orc_get_field(const orc* o, int which, double* dst)
{
    memcpy(dst, field_ptr(o, which), sizeof(double) * (size_t)orc_field_size(o, which));
}

void orc_set_scalars(orc* o, double dt, double t, double dPdx, double Q)
{
    o->dt = dt; o->t = t; o->dPdx = dPdx; o->Qflow_x_0 = Q;
}

void orc_get_scalars(const orc* o, double* out)
{
    out[0] = o->dt; out[1] = o->t; out[2] = o->dPdx; out[3] = o->Qflow_x_0;
}

/* -------------------------------------------------------- periodic helpers */

/* code/boundary_conditions.f90:60-79 (x) and :84-226 collapsed to the
 * global index pattern (z): faces F(1)=F(n-1), F(n)=F(2); centres
 * F(1)=F(ng-2), F(ng-1)=F(2), F(ng)=F(3).  Arrays of leading dims n1,n2,n3. */
static void periodic_x(double* F, int n1, int n2, int n3, int faces)
{
    for (int k = 1; k <= n3; k++)
        for (int j = 1; j <= n2; j++) {
            if (faces) {
                F[IX(1, j, k, n1, n2)] = F[IX(n1 - 1, j, k, n1, n2)];
                F[IX(n1, j, k, n1, n2)] = F[IX(2, j, k, n1, n2)];
            } else {
                F[IX(1, j, k, n1, n2)] = F[IX(n1 - 2, j, k, n1, n2)];
                F[IX(n1 - 1, j, k, n1, n2)] = F[IX(2, j, k, n1, n2)];
                F[IX(n1, j, k, n1, n2)] = F[IX(3, j, k, n1, n2)];
            }
        }
}

static void periodic_z(double* F, int n1, int n2, int n3, int faces)
{
    for (int j = 1; j <= n2; j++)
        for (int i = 1; i <= n1; i++) {
            if (faces) {
                F[IX(i, j, 1, n1, n2)] = F[IX(i, j, n3 - 1, n1, n2)];
                F[IX(i, j, n3, n1, n2)] = F[IX(i, j, 2, n1, n2)];
            } else {
                F[IX(i, j, 1, n1, n2)] = F[IX(i, j, n3 - 2, n1, n2)];
                F[IX(i, j, n3 - 1, n1, n2)] = F[IX(i, j, 2, n1, n2)];
                F[IX(i, j, n3, n1, n2)] = F[IX(i, j, 3, n1, n2)];
            }
        }
}

/* code/boundary_conditions.f90:239-254 (Dirichlet) and :266-281 (Neumann) */
static void wall_y(double* F, int n1, int n2, int n3, int faces, int mode /*1 Dirichlet 0 Neumann*/)
{
    for (int k = 1; k <= n3; k++)
        for (int i = 1; i <= n1; i++) {
            if (faces) {
                if (mode) { F[IX(i, 1, k, n1, n2)] = 0.0; F[IX(i, n2, k, n1, n2)] = 0.0; }
                else {
                    F[IX(i, 1, k, n1, n2)] = F[IX(i, 2, k, n1, n2)];
                    F[IX(i, n2, k, n1, n2)] = F[IX(i, n2 - 1, k, n1, n2)];
                }
            } else {
                if (mode) {
                    F[IX(i, 1, k, n1, n2)] = -F[IX(i, 2, k, n1, n2)];
                    F[IX(i, n2, k, n1, n2)] = -F[IX(i, n2 - 1, k, n1, n2)];
                } else {
                    F[IX(i, 1, k, n1, n2)] = F[IX(i, 2, k, n1, n2)];
                    F[IX(i, n2, k, n1, n2)] = F[IX(i, n2 - 1, k, n1, n2)];
                }
            }
        }
}

/* code/boundary_conditions.f90:20-48 */
static void apply_bc(const orc* o, double* U, double* V, double* W)
{
    periodic_x(U, o->nx, o->nyg, o->nzg, 1);
    periodic_x(V, o->nxg, o->ny, o->nzg, 0);
    periodic_x(W, o->nxg, o->nyg, o->nz, 0);
    periodic_z(U, o->nx, o->nyg, o->nzg, 0);
    periodic_z(V, o->nxg, o->ny, o->nzg, 0);
    periodic_z(W, o->nxg, o->nyg, o->nz, 1);
    wall_y(U, o->nx, o->nyg, o->nzg, 0, 1);
    wall_y(V, o->nxg, o->ny, o->nzg, 1, 1);
    wall_y(W, o->nxg, o->nyg, o->nz, 0, 1);
}

void orc_apply_boundary_conditions(orc* o) { apply_bc(o, o->U, o->V, o->W); }

/* ---------------------------------------------------------------- RHS u,v,w */

/* code/equations.f90:24-167 (interpolations: code/interpolation.f90:24-105) */
void orc_compute_rhs_u(const orc* o, const double* U, const double* V, const double* W, double* rhs)
{
    const double *x = o->x, *y = o->y, *z = o->z, *xg = o->xg, *yg = o->yg, *zg = o->zg;
    const double *w0 = o->w0, *w1 = o->w1, *nt = o->nu_t;
    const double nu = o->nu;
#define NT(i, j, k) AC(nt, i, j, k)
    PAR_FOR
    for (int k = 2; k <= o->nzg - 1; k++) {
        double dz_1 = zg[k] - zg[k - 1], dz_2 = zg[k + 1] - zg[k], dz_3 = z[k] - z[k - 1];
        for (int j = 2; j <= o->nyg - 1; j++) {
            double dy_3 = y[j] - y[j - 1];
            for (int i = 2; i <= o->nx - 1; i++) {
                /* 1) -du^2/dx, u interpolated to centres and squared */
                double uc = 0.5 * (AU(U, i, j, k) + AU(U, i + 1, j, k));
                double ucm = 0.5 * (AU(U, i - 1, j, k) + AU(U, i, j, k));
                double t1 = -(uc * uc - ucm * ucm) / (xg[i + 1] - xg[i]);
                /* 2) -duv/dy at y faces: u weighted in y, v averaged in x */
                double uvp = (w0[j] * AU(U, i, j, k) + w1[j] * AU(U, i, j + 1, k)) *
                             (0.5 * (AV(V, i, j, k) + AV(V, i + 1, j, k)));
                double uvm = (w0[j - 1] * AU(U, i, j - 1, k) + w1[j - 1] * AU(U, i, j, k)) *
                             (0.5 * (AV(V, i, j - 1, k) + AV(V, i + 1, j - 1, k)));
                double t2 = -(uvp - uvm) / (y[j] - y[j - 1]);
                /* 3) -duw/dz at z faces */
                double uwp = (0.5 * (AU(U, i, j, k) + AU(U, i, j, k + 1))) *
                             (0.5 * (AW(W, i, j, k) + AW(W, i + 1, j, k)));
                double uwm = (0.5 * (AU(U, i, j, k - 1) + AU(U, i, j, k))) *
                             (0.5 * (AW(W, i, j, k - 1) + AW(W, i + 1, j, k - 1)));
                double t3 = -(uwp - uwm) / (z[k] - z[k - 1]);
                /* 4) div((nu+nu_t) grad u) */
                double dx_1 = x[i] - x[i - 1], dx_2 = x[i + 1] - x[i], dx_3 = xg[i + 1] - xg[i];
                double nu_x1 = nu + NT(i, j, k);
                double nu_x2 = nu + NT(i + 1, j, k);
                double nu_y1 = nu + 0.5 * (w0[j - 1] * NT(i, j - 1, k) + w1[j - 1] * NT(i, j, k) +
                                           w0[j - 1] * NT(i + 1, j - 1, k) + w1[j - 1] * NT(i + 1, j, k));
                double nu_y2 = nu + 0.5 * (w0[j] * NT(i, j, k) + w1[j] * NT(i, j + 1, k) +
                                           w0[j] * NT(i + 1, j, k) + w1[j] * NT(i + 1, j + 1, k));
                double nu_z1 = nu + 0.25 * (NT(i, j, k) + NT(i, j, k - 1) + NT(i + 1, j, k) + NT(i + 1, j, k - 1));
                double nu_z2 = nu + 0.25 * (NT(i, j, k) + NT(i, j, k + 1) + NT(i + 1, j, k) + NT(i + 1, j, k + 1));
                double dudy_p = (AU(U, i, j + 1, k) - AU(U, i, j, k)) / (yg[j + 1] - yg[j]);
                double dudy_m = (AU(U, i, j, k) - AU(U, i, j - 1, k)) / (yg[j] - yg[j - 1]);
                double t4 =
                    1.0 / dx_3 * (nu_x2 * 1.0 / dx_2 * (AU(U, i + 1, j, k) - AU(U, i, j, k)) -
                                  nu_x1 * 1.0 / dx_1 * (AU(U, i, j, k) - AU(U, i - 1, j, k))) +
                    1.0 / dy_3 * (nu_y2 * dudy_p - nu_y1 * dudy_m) +
                    1.0 / dz_3 * (nu_z2 * 1.0 / dz_2 * (AU(U, i, j, k + 1) - AU(U, i, j, k)) -
                                  nu_z1 * 1.0 / dz_1 * (AU(U, i, j, k) - AU(U, i, j, k - 1)));
                AU(rhs, i, j, k) = (((t1 + t2) + t3) + t4) + o->dPdx;
            }
        }
    }
}

/* code/equations.f90:174-320 */
void orc_compute_rhs_v(const orc* o, const double* U, const double* V, const double* W, double* rhs)
{
    const double *x = o->x, *y = o->y, *z = o->z, *xg = o->xg, *yg = o->yg, *zg = o->zg;
    const double *w0 = o->w0, *w1 = o->w1, *nt = o->nu_t;
    const double nu = o->nu;
    PAR_FOR
    for (int k = 2; k <= o->nzg - 1; k++) {
        double dz_1 = zg[k] - zg[k - 1], dz_2 = zg[k + 1] - zg[k], dz_3 = z[k] - z[k - 1];
        for (int j = 2; j <= o->ny - 1; j++) {
            double dy_1 = y[j] - y[j - 1], dy_2 = y[j + 1] - y[j], dy_3 = yg[j + 1] - yg[j];
            for (int i = 2; i <= o->nxg - 1; i++) {
                /* 1) -dv^2/dy, v averaged to centres */
                double vc = 0.5 * (AV(V, i, j, k) + AV(V, i, j + 1, k));
                double vcm = 0.5 * (AV(V, i, j - 1, k) + AV(V, i, j, k));
                double t1 = -(vc * vc - vcm * vcm) / (yg[j + 1] - yg[j]);
                /* 2) -duv/dx at x faces */
                double uvp = (w0[j] * AU(U, i, j, k) + w1[j] * AU(U, i, j + 1, k)) *
                             (0.5 * (AV(V, i, j, k) + AV(V, i + 1, j, k)));
                double uvm = (w0[j] * AU(U, i - 1, j, k) + w1[j] * AU(U, i - 1, j + 1, k)) *
                             (0.5 * (AV(V, i - 1, j, k) + AV(V, i, j, k)));
                double t2 = -(uvp - uvm) / (x[i] - x[i - 1]);
                /* 3) -dvw/dz at z faces */
                double vwp = (0.5 * (AV(V, i, j, k) + AV(V, i, j, k + 1))) *
                             (w0[j] * AW(W, i, j, k) + w1[j] * AW(W, i, j + 1, k));
                double vwm = (0.5 * (AV(V, i, j, k - 1) + AV(V, i, j, k))) *
                             (w0[j] * AW(W, i, j, k - 1) + w1[j] * AW(W, i, j + 1, k - 1));
                double t3 = -(vwp - vwm) / (z[k] - z[k - 1]);
                /* 4) viscous */
                double dx_1 = xg[i] - xg[i - 1], dx_2 = xg[i + 1] - xg[i], dx_3 = x[i] - x[i - 1];
                double nu_x1 = nu + 0.5 * (w0[j] * NT(i - 1, j, k) + w1[j] * NT(i - 1, j + 1, k) +
                                           w0[j] * NT(i, j, k) + w1[j] * NT(i, j + 1, k));
                double nu_x2 = nu + 0.5 * (w0[j] * NT(i, j, k) + w1[j] * NT(i, j + 1, k) +
                                           w0[j] * NT(i + 1, j, k) + w1[j] * NT(i + 1, j + 1, k));
                double nu_y1 = nu + NT(i, j, k);
                double nu_y2 = nu + NT(i, j + 1, k);
                double nu_z1 = nu + 0.5 * (w0[j] * NT(i, j, k - 1) + w1[j] * NT(i, j + 1, k - 1) +
                                           w0[j] * NT(i, j, k) + w1[j] * NT(i, j + 1, k));
                double nu_z2 = nu + 0.5 * (w0[j] * NT(i, j, k) + w1[j] * NT(i, j + 1, k) +
                                           w0[j] * NT(i, j, k + 1) + w1[j] * NT(i, j + 1, k + 1));
                double t4 =
                    1.0 / dx_3 * (nu_x2 * 1.0 / dx_2 * (AV(V, i + 1, j, k) - AV(V, i, j, k)) -
                                  nu_x1 * 1.0 / dx_1 * (AV(V, i, j, k) - AV(V, i - 1, j, k))) +
                    1.0 / dy_3 * (nu_y2 * 1.0 / dy_2 * (AV(V, i, j + 1, k) - AV(V, i, j, k)) -
                                  nu_y1 * 1.0 / dy_1 * (AV(V, i, j, k) - AV(V, i, j - 1, k))) +
                    1.0 / dz_3 * (nu_z2 * 1.0 / dz_2 * (AV(V, i, j, k + 1) - AV(V, i, j, k)) -
                                  nu_z1 * 1.0 / dz_1 * (AV(V, i, j, k) - AV(V, i, j, k - 1)));
                AV(rhs, i, j, k) = ((t1 + t2) + t3) + t4;
            }
        }
    }
}

/* code/equations.f90:327-468 */
void orc_compute_rhs_w(const orc* o, const double* U, const double* V, const double* W, double* rhs)
{
    const double *x = o->x, *y = o->y, *z = o->z, *xg = o->xg, *yg = o->yg, *zg = o->zg;
    const double *w0 = o->w0, *w1 = o->w1, *nt = o->nu_t;
    const double nu = o->nu;
    PAR_FOR
    for (int k = 2; k <= o->nz - 1; k++) {
        double dz_1 = z[k] - z[k - 1], dz_2 = z[k + 1] - z[k], dz_3 = zg[k + 1] - zg[k];
        for (int j = 2; j <= o->nyg - 1; j++) {
            double dy_3 = y[j] - y[j - 1];
            for (int i = 2; i <= o->nxg - 1; i++) {
                /* 1) -dw^2/dz */
                double wc = 0.5 * (AW(W, i, j, k) + AW(W, i, j, k + 1));
                double wcm = 0.5 * (AW(W, i, j, k - 1) + AW(W, i, j, k));
                double t1 = -(wc * wc - wcm * wcm) / (zg[k + 1] - zg[k]);
                /* 2) -duw/dx at x faces */
                double uwp = (0.5 * (AU(U, i, j, k) + AU(U, i, j, k + 1))) *
                             (0.5 * (AW(W, i, j, k) + AW(W, i + 1, j, k)));
                double uwm = (0.5 * (AU(U, i - 1, j, k) + AU(U, i - 1, j, k + 1))) *
                             (0.5 * (AW(W, i - 1, j, k) + AW(W, i, j, k)));
                double t2 = -(uwp - uwm) / (x[i] - x[i - 1]);
                /* 3) -dvw/dy at y faces */
                double vwp = (0.5 * (AV(V, i, j, k) + AV(V, i, j, k + 1))) *
                             (w0[j] * AW(W, i, j, k) + w1[j] * AW(W, i, j + 1, k));
                double vwm = (0.5 * (AV(V, i, j - 1, k) + AV(V, i, j - 1, k + 1))) *
                             (w0[j - 1] * AW(W, i, j - 1, k) + w1[j - 1] * AW(W, i, j, k));
                double t3 = -(vwp - vwm) / (y[j] - y[j - 1]);
                /* 4) viscous */
                double dx_1 = xg[i] - xg[i - 1], dx_2 = xg[i + 1] - xg[i], dx_3 = x[i] - x[i - 1];
                double nu_x1 = nu + 0.25 * (NT(i, j, k) + NT(i, j, k + 1) + NT(i - 1, j, k) + NT(i - 1, j, k + 1));
                double nu_x2 = nu + 0.25 * (NT(i, j, k) + NT(i, j, k + 1) + NT(i + 1, j, k) + NT(i + 1, j, k + 1));
                double nu_y1 = nu + 0.5 * (w0[j - 1] * NT(i, j - 1, k) + w1[j - 1] * NT(i, j, k) +
                                           w0[j - 1] * NT(i, j - 1, k + 1) + w1[j - 1] * NT(i, j, k + 1));
                double nu_y2 = nu + 0.5 * (w0[j] * NT(i, j, k) + w1[j] * NT(i, j + 1, k) +
                                           w0[j] * NT(i, j, k + 1) + w1[j] * NT(i, j + 1, k + 1));
                double nu_z1 = nu + NT(i, j, k);
                double nu_z2 = nu + NT(i, j, k + 1);
                double dwdy_p = (AW(W, i, j + 1, k) - AW(W, i, j, k)) / (yg[j + 1] - yg[j]);
                double dwdy_m = (AW(W, i, j, k) - AW(W, i, j - 1, k)) / (yg[j] - yg[j - 1]);
                double t4 =
                    1.0 / dx_3 * (nu_x2 * 1.0 / dx_2 * (AW(W, i + 1, j, k) - AW(W, i, j, k)) -
                                  nu_x1 * 1.0 / dx_1 * (AW(W, i, j, k) - AW(W, i - 1, j, k))) +
                    1.0 / dy_3 * (nu_y2 * dwdy_p - nu_y1 * dwdy_m) +
                    1.0 / dz_3 * (nu_z2 * 1.0 / dz_2 * (AW(W, i, j, k + 1) - AW(W, i, j, k)) -
                                  nu_z1 * 1.0 / dz_1 * (AW(W, i, j, k) - AW(W, i, j, k - 1)));
                AW(rhs, i, j, k) = ((t1 + t2) + t3) + t4;
            }
        }
    }
#undef NT
}

/* ------------------------------------------------------------------- SGS */

/* code/subgrid.f90:140-190 -- filter_xzy: 3x3 Simpson in x-z times
 * (1/6,2/3,1/6) in y, one-sided (2/3,1/3) on the first/last row.  Output
 * position (ii-1, j, kk-1) of an (n1-2, n2, n3-2) array, as the reference
 * stores it through its (2:n1-1, :, 2:n3-1) section argument. */
static const double FIL[3][3] = {{1.0 / 36.0, 1.0 / 9.0, 1.0 / 36.0},
                                 {1.0 / 9.0, 4.0 / 9.0, 1.0 / 9.0},
                                 {1.0 / 36.0, 1.0 / 9.0, 1.0 / 36.0}};

void orc_filter_xzy(int n1, int n2, int n3, const double* in, double* out)
{
    int m1 = n1 - 2, m2 = n2;
#define IN(i, j, k) in[IX(i, j, k, n1, n2)]
#define OUT(i, j, k) out[IX(i, j, k, m1, m2)]
    const long nout = (long)m1 * m2 * (n3 - 2);
    PAR_FOR
    for (long q = 0; q < nout; q++) out[q] = 0.0;
    /* each ii accumulates into its own output column OUT(ii-1, :, :) */
    PAR_FOR
    for (int ii = 2; ii <= n1 - 1; ii++)
        for (int i = -1; i <= 1; i++)
            for (int kk = 2; kk <= n3 - 1; kk++)
                for (int k = -1; k <= 1; k++) {
                    double f = FIL[i + 1][k + 1];
                    for (int j = 2; j <= n2 - 1; j++)
                        OUT(ii - 1, j, kk - 1) = OUT(ii - 1, j, kk - 1) +
                            (1.0 / 6.0 * IN(ii + i, j - 1, kk + k) + 2.0 / 3.0 * IN(ii + i, j, kk + k) +
                             1.0 / 6.0 * IN(ii + i, j + 1, kk + k)) * f;
                    OUT(ii - 1, 1, kk - 1) = OUT(ii - 1, 1, kk - 1) +
                        (2.0 / 3.0 * IN(ii + i, 1, kk + k) + 1.0 / 3.0 * IN(ii + i, 2, kk + k)) * f;
                    OUT(ii - 1, n2, kk - 1) = OUT(ii - 1, n2, kk - 1) +
                        (1.0 / 3.0 * IN(ii + i, n2 - 1, kk + k) + 2.0 / 3.0 * IN(ii + i, n2, kk + k)) * f;
                }
#undef IN
#undef OUT
}

/* filter a centre-box field restricted to rows 2..nyg-1 (the reference's
 * filter_xzy(T(1:nxg,2:nyg-1,1:nzg), Out(2:nxg-1,2:nyg-1,2:nzg-1))) into
 * the box positions (2..nxg-1, 2..nyg-1, 2..nzg-1) of dst. */
static void filter_box_rows(const orc* o, const double* src, double* dst, double* tmp_in, double* tmp_out)
{
    int n1 = o->nxg, n2 = o->nyg - 2, n3 = o->nzg;
    PAR_FOR
    for (int k = 1; k <= n3; k++)
        for (int j = 1; j <= n2; j++)
            for (int i = 1; i <= n1; i++) tmp_in[IX(i, j, k, n1, n2)] = AC(src, i, j + 1, k);
    orc_filter_xzy(n1, n2, n3, tmp_in, tmp_out);
    PAR_FOR
    for (int k = 2; k <= n3 - 1; k++)
        for (int j = 1; j <= n2; j++)
            for (int i = 2; i <= n1 - 1; i++) AC(dst, i, j + 1, k) = tmp_out[IX(i - 1, j, k - 1, n1 - 2, n2)];
}

/* code/subgrid.f90:1014-1072 -- rate of strain at centres; Sij holds six
 * centre-box components, S the magnitude sqrt(2 SijSij) (computed as the
 * reference's S**0.5d0 with sqrt). */
static void compute_Sij_full(const orc* o, const double* U, const double* V, const double* W,
                             int nUx, int nVy, int nWz, double* Sij, double* S)
{
    /* nUx / nVy / nWz: leading extents of the staggered inputs, which for the
     * filtered velocities are the same as for U, V, W. */
    (void)nUx; (void)nVy; (void)nWz;
    const double *x = o->x, *y = o->y, *z = o->z, *xg = o->xg, *yg = o->yg, *zg = o->zg;
    size_t nc = nC(o);
    double *S1 = Sij, *S2 = Sij + nc, *S3 = Sij + 2 * nc, *S4 = Sij + 3 * nc, *S5 = Sij + 4 * nc,
           *S6 = Sij + 5 * nc;
    PAR_FOR
    for (int k = 2; k <= o->nzg - 1; k++)
        for (int j = 2; j <= o->nyg - 1; j++)
            for (int i = 2; i <= o->nxg - 1; i++) {
                AC(S1, i, j, k) = (AU(U, i, j, k) - AU(U, i - 1, j, k)) / (x[i] - x[i - 1]);
                AC(S2, i, j, k) = (AV(V, i, j, k) - AV(V, i, j - 1, k)) / (y[j] - y[j - 1]);
                AC(S3, i, j, k) = (AW(W, i, j, k) - AW(W, i, j, k - 1)) / (z[k] - z[k - 1]);
            }
    PAR_FOR
    for (int k = 2; k <= o->nzg; k++)
        for (int j = 2; j <= o->nyg; j++)
            for (int i = 2; i <= o->nxg; i++) {
                AC(S4, i, j, k) = 0.5 * ((AU(U, i - 1, j, k) - AU(U, i - 1, j - 1, k)) / (yg[j] - yg[j - 1]) +
                                         (AV(V, i, j - 1, k) - AV(V, i - 1, j - 1, k)) / (xg[i] - xg[i - 1]));
                AC(S5, i, j, k) = 0.5 * ((AU(U, i - 1, j, k) - AU(U, i - 1, j, k - 1)) / (zg[k] - zg[k - 1]) +
                                         (AW(W, i, j, k - 1) - AW(W, i - 1, j, k - 1)) / (xg[i] - xg[i - 1]));
                AC(S6, i, j, k) = 0.5 * ((AV(V, i, j - 1, k) - AV(V, i, j - 1, k - 1)) / (zg[k] - zg[k - 1]) +
                                         (AW(W, i, j, k - 1) - AW(W, i, j - 1, k - 1)) / (yg[j] - yg[j - 1]));
            }
    /* edge -> centre, in place in increasing index order (reads only later
     * points, so every read sees an edge value).  One loop per component,
     * each split across threads along the axis its stencil does not touch:
     * S4 (x,y stencil) by k, S5 (x,z) by j, S6 (y,z) by i; within a thread
     * the order stays increasing. */
    PAR_FOR
    for (int k = 2; k <= o->nzg - 1; k++)
        for (int j = 2; j <= o->nyg - 1; j++)
            for (int i = 2; i <= o->nxg - 1; i++)
                AC(S4, i, j, k) = 0.25 * (AC(S4, i, j, k) + AC(S4, i + 1, j, k) + AC(S4, i, j + 1, k) + AC(S4, i + 1, j + 1, k));
    PAR_FOR
    for (int j = 2; j <= o->nyg - 1; j++)
        for (int k = 2; k <= o->nzg - 1; k++)
            for (int i = 2; i <= o->nxg - 1; i++)
                AC(S5, i, j, k) = 0.25 * (AC(S5, i, j, k) + AC(S5, i + 1, j, k) + AC(S5, i, j, k + 1) + AC(S5, i + 1, j, k + 1));
    PAR_FOR
    for (int i = 2; i <= o->nxg - 1; i++)
        for (int k = 2; k <= o->nzg - 1; k++)
            for (int j = 2; j <= o->nyg - 1; j++)
                AC(S6, i, j, k) = 0.25 * (AC(S6, i, j, k) + AC(S6, i, j + 1, k) + AC(S6, i, j, k + 1) + AC(S6, i, j + 1, k + 1));
    PAR_FOR
    for (int k = 2; k <= o->nzg - 1; k++)
        for (int j = 2; j <= o->nyg - 1; j++)
            for (int i = 2; i <= o->nxg - 1; i++) {
                double s = 0.0;
                s = s + 2.0 * AC(S1, i, j, k) * AC(S1, i, j, k);
                s = s + 2.0 * AC(S2, i, j, k) * AC(S2, i, j, k);
                s = s + 2.0 * AC(S3, i, j, k) * AC(S3, i, j, k);
                s = s + 4.0 * AC(S4, i, j, k) * AC(S4, i, j, k);
                s = s + 4.0 * AC(S5, i, j, k) * AC(S5, i, j, k);
                s = s + 4.0 * AC(S6, i, j, k) * AC(S6, i, j, k);
                AC(S, i, j, k) = sqrt(s);
            }
}

void orc_compute_Sij(const orc* o, const double* U, const double* V, const double* W, double* Sij, double* S)
{
    compute_Sij_full(o, U, V, W, o->nx, o->ny, o->nz, Sij, S);
}

/* code/subgrid.f90:438-440 -- filter a staggered field over its whole y
 * range (ghost rows included) into the same-shaped output, interior x/z. */
static void filter_full(const double* F, int n1, int n2, int n3, double* Ff, double* tmp)
{
    orc_filter_xzy(n1, n2, n3, F, tmp);
    PAR_FOR
    for (int k = 2; k <= n3 - 1; k++)
        for (int j = 1; j <= n2; j++)
            for (int i = 2; i <= n1 - 1; i++) Ff[IX(i, j, k, n1, n2)] = tmp[IX(i - 1, j, k - 1, n1 - 2, n2)];
}

/* code/subgrid.f90:379-537 -- dynamic Smagorinsky (Lilly).  L and M are
 * returned as 6 centre-box components valid on (2:nxg-1,2:nyg-1,2:nzg-1),
 * LM/MM as the xz-plane sums (index j-1 for j=2..nyg-1). */
static void dsm_core(orc* o, const double* U, const double* V, const double* W, double* nu_t,
                     double* L, double* M, double* LMp, double* MMp)
{
    int nxg = o->nxg, nyg = o->nyg, nzg = o->nzg;
    size_t nc = nC(o);
    double *uc = o->t0, *vc = o->t1, *wc = o->t2, *tmp = o->t3;
    double *Sij = o->Sijb, *S = o->Sbuf, *Sh = o->Sh;
    double* tin = dalloc(nc);
    double* tout = dalloc(nc);

    /* Part 1: centre velocities (faces -> centres), periodic fill */
    memset(uc, 0, nc * sizeof(double)); memset(vc, 0, nc * sizeof(double)); memset(wc, 0, nc * sizeof(double));
    PAR_FOR
    for (int k = 1; k <= nzg; k++)
        for (int j = 1; j <= nyg; j++)
            for (int i = 2; i <= nxg - 1; i++) AC(uc, i, j, k) = 0.5 * (AU(U, i - 1, j, k) + AU(U, i, j, k));
    PAR_FOR
    for (int k = 1; k <= nzg; k++)
        for (int j = 2; j <= nyg - 1; j++)
            for (int i = 1; i <= nxg; i++) AC(vc, i, j, k) = 0.5 * (AV(V, i, j - 1, k) + AV(V, i, j, k));
    PAR_FOR
    for (int k = 2; k <= nzg - 1; k++)
        for (int j = 1; j <= nyg; j++)
            for (int i = 1; i <= nxg; i++) AC(wc, i, j, k) = 0.5 * (AW(W, i, j, k - 1) + AW(W, i, j, k));
    periodic_x(uc, nxg, nyg, nzg, 0);
    periodic_z(wc, nxg, nyg, nzg, 0);

    /* products u_i u_j (periodic by construction), their filter hat(u_i u_j) */
    const double* comp[3] = {uc, vc, wc};
    const int pa[6] = {0, 1, 2, 0, 0, 1}, pb[6] = {0, 1, 2, 1, 2, 2};
    for (int q = 0; q < 6; q++) {
        double* Lq = L + q * nc;
        memset(tmp, 0, nc * sizeof(double));
        PAR_FOR
        for (int k = 1; k <= nzg; k++)
            for (int j = 2; j <= nyg - 1; j++)
                for (int i = 1; i <= nxg; i++)
                    AC(tmp, i, j, k) = AC(comp[pa[q]], i, j, k) * AC(comp[pb[q]], i, j, k);
        filter_box_rows(o, tmp, Lq, tin, tout);
    }
    /* hat(u_i) into Sij components 0..2 */
    for (int c = 0; c < 3; c++) filter_box_rows(o, comp[c], Sij + c * nc, tin, tout);
    PAR_FOR2
    for (int q = 0; q < 6; q++)
        for (int k = 2; k <= nzg - 1; k++)
            for (int j = 2; j <= nyg - 1; j++)
                for (int i = 2; i <= nxg - 1; i++)
                    AC(L + q * nc, i, j, k) = AC(L + q * nc, i, j, k) -
                                              AC(Sij + pa[q] * nc, i, j, k) * AC(Sij + pb[q] * nc, i, j, k);

    /* Part 2: filtered staggered velocities, periodic fill, hat(S) */
    double *Uf = o->Uf, *Vf = o->Vf, *Wf = o->Wf;
    filter_full(U, o->nx, nyg, nzg, Uf, tout);
    periodic_x(Uf, o->nx, nyg, nzg, 1); periodic_z(Uf, o->nx, nyg, nzg, 0);
    filter_full(V, nxg, o->ny, nzg, Vf, tout);
    periodic_x(Vf, nxg, o->ny, nzg, 0); periodic_z(Vf, nxg, o->ny, nzg, 0);
    filter_full(W, nxg, nyg, o->nz, Wf, tout);
    periodic_x(Wf, nxg, nyg, o->nz, 0); periodic_z(Wf, nxg, nyg, o->nz, 1);
    compute_Sij_full(o, Uf, Vf, Wf, o->nx, o->ny, o->nz, Sij, Sh);
    const double fs2 = 2.0 * 2.0; /* fil_size**2, fil_size = 2 (code/subgrid.f90:156) */
    PAR_FOR2
    for (int q = 0; q < 6; q++)
        for (int k = 2; k <= nzg - 1; k++)
            for (int j = 2; j <= nyg - 1; j++)
                for (int i = 2; i <= nxg - 1; i++)
                    AC(M + q * nc, i, j, k) = fs2 * AC(Sh, i, j, k) * AC(Sij + q * nc, i, j, k);

    /* unfiltered S_ij, |S|; M -= hat(|S| S_ij) */
    compute_Sij_full(o, U, V, W, o->nx, o->ny, o->nz, Sij, S);
    for (int q = 0; q < 6; q++) {
        memset(tmp, 0, nc * sizeof(double));
        PAR_FOR
        for (int k = 2; k <= nzg - 1; k++)
            for (int j = 2; j <= nyg - 1; j++)
                for (int i = 2; i <= nxg - 1; i++) AC(tmp, i, j, k) = AC(S, i, j, k) * AC(Sij + q * nc, i, j, k);
        periodic_x(tmp, nxg, nyg, nzg, 0);
        periodic_z(tmp, nxg, nyg, nzg, 0);
        filter_box_rows(o, tmp, Sij + q * nc, tin, tout);
    }
    PAR_FOR2
    for (int q = 0; q < 6; q++)
        for (int k = 2; k <= nzg - 1; k++)
            for (int j = 2; j <= nyg - 1; j++)
                for (int i = 2; i <= nxg - 1; i++)
                    AC(M + q * nc, i, j, k) = AC(M + q * nc, i, j, k) - AC(Sij + q * nc, i, j, k);

    /* Part 3: plane averages of L:M and M:M (plane = i 2..nxg-1, k 2..nzg-1) */
    if (nu_t) memset(nu_t, 0, nc * sizeof(double));
#define LL(q) AC(L + (q) * nc, i, j, k)
#define MMq(q) AC(M + (q) * nc, i, j, k)
    PAR_FOR
    for (int j = 2; j <= nyg - 1; j++) {
        double slm = 0.0, smm = 0.0;
        for (int k = 2; k <= nzg - 1; k++)
            for (int i = 2; i <= nxg - 1; i++) {
                double lm = LL(0) * MMq(0) + LL(1) * MMq(1) + LL(2) * MMq(2) +
                            2.0 * (LL(3) * MMq(3) + LL(4) * MMq(4) + LL(5) * MMq(5));
                slm = slm + lm;
            }
        for (int k = 2; k <= nzg - 1; k++)
            for (int i = 2; i <= nxg - 1; i++) {
                double mm = MMq(0) * MMq(0) + MMq(1) * MMq(1) + MMq(2) * MMq(2) +
                            2.0 * (MMq(3) * MMq(3) + MMq(4) * MMq(4) + MMq(5) * MMq(5));
                smm = smm + mm;
            }
        if (LMp) LMp[j - 1] = slm;
        if (MMp) MMp[j - 1] = smm;
        if (nu_t) {
            double r = slm / smm;
            for (int k = 2; k <= nzg - 1; k++)
                for (int i = 2; i <= nxg - 1; i++) AC(nu_t, i, j, k) = -0.5 * r * AC(S, i, j, k);
        }
    }
#undef LL
#undef MMq
    if (nu_t) {
        periodic_x(nu_t, nxg, nyg, nzg, 0);
        periodic_z(nu_t, nxg, nyg, nzg, 0);
        /* clipping; 0/0 planes (M:M = 0) give NaN, clipped to 0 (fmax) */
        PAR_FOR
        for (size_t q = 0; q < nc; q++) nu_t[q] = fmax(nu_t[q], 0.0);
        wall_y(nu_t, nxg, nyg, nzg, 0, o->Dirichlet_nu_t == 1);
    }
    free(tin); free(tout);
}

void orc_dsm_parts(orc* o, const double* U, const double* V, const double* W, double* L, double* M,
                   double* LMp, double* MMp)
{
    dsm_core(o, U, V, W, 0, L, M, LMp, MMp);
}

/* code/subgrid.f90:632-874 -- anisotropic minimum dissipation (AMD) at one
 * interior centre (i,j,k):
 *   nu_t = -C N / D, C = 0.3,
 *   N = sum (dh_k u_i)(dh_k u_j) S_ij (dh: the undivided difference, i.e. the
 *       spacing times the derivative), D = sum (d_l u_m)^2,
 * with every staggered difference carried to the centre by the reference's
 * face-to-centre averages (code/interpolation.f90:24-95, plain averages) in
 * the order its calls use them, and the terms accumulated in the order of
 * its statements.  ax(a, b) = 0.5*(a + b). */
#define AX(a, b) (0.5 * ((a) + (b)))
static double amd_point(const orc* o, const double* U, const double* V, const double* W, const double* Sij, int i,
                        int j, int k)
{
    const double *x = o->x, *y = o->y, *z = o->z, *xg = o->xg, *yg = o->yg, *zg = o->zg;
    const size_t nc = nC(o);
    const double S11 = AC(Sij, i, j, k), S22 = AC(Sij + nc, i, j, k), S33 = AC(Sij + 2 * nc, i, j, k);
    const double S12 = AC(Sij + 3 * nc, i, j, k), S13 = AC(Sij + 4 * nc, i, j, k), S23 = AC(Sij + 5 * nc, i, j, k);
    /* undivided differences at their natural positions */
#define dUx(a, b, c) (AU(U, a, b, c) - AU(U, (a) - 1, b, c))
#define dUy(a, b, c) (AU(U, a, b, c) - AU(U, a, (b) - 1, c))
#define dUz(a, b, c) (AU(U, a, b, c) - AU(U, a, b, (c) - 1))
#define dVx(a, b, c) (AV(V, a, b, c) - AV(V, (a) - 1, b, c))
#define dVy(a, b, c) (AV(V, a, b, c) - AV(V, a, (b) - 1, c))
#define dVz(a, b, c) (AV(V, a, b, c) - AV(V, a, b, (c) - 1))
#define dWx(a, b, c) (AW(W, a, b, c) - AW(W, (a) - 1, b, c))
#define dWy(a, b, c) (AW(W, a, b, c) - AW(W, a, (b) - 1, c))
#define dWz(a, b, c) (AW(W, a, b, c) - AW(W, a, b, (c) - 1))
    /* U: x then y / x then z; V: y then x / y then z; W: z then x / z then y */
    const double Uxh = dUx(i, j, k);
    const double Uyh = AX(AX(dUy(i - 1, j, k), dUy(i, j, k)), AX(dUy(i - 1, j + 1, k), dUy(i, j + 1, k)));
    const double Uzh = AX(AX(dUz(i - 1, j, k), dUz(i, j, k)), AX(dUz(i - 1, j, k + 1), dUz(i, j, k + 1)));
    const double Vyh = dVy(i, j, k);
    const double Vxh = AX(AX(dVx(i, j - 1, k), dVx(i, j, k)), AX(dVx(i + 1, j - 1, k), dVx(i + 1, j, k)));
    const double Vzh = AX(AX(dVz(i, j - 1, k), dVz(i, j, k)), AX(dVz(i, j - 1, k + 1), dVz(i, j, k + 1)));
    const double Wzh = dWz(i, j, k);
    const double Wxh = AX(AX(dWx(i, j, k - 1), dWx(i, j, k)), AX(dWx(i + 1, j, k - 1), dWx(i + 1, j, k)));
    const double Wyh = AX(AX(dWy(i, j, k - 1), dWy(i, j, k)), AX(dWy(i, j + 1, k - 1), dWy(i, j + 1, k)));
    /* derivatives: the differences divided by their spacings before averaging */
#define q(a, b) ((a) / (b))
    const double Uyd = AX(AX(q(dUy(i - 1, j, k), yg[j] - yg[j - 1]), q(dUy(i, j, k), yg[j] - yg[j - 1])),
                          AX(q(dUy(i - 1, j + 1, k), yg[j + 1] - yg[j]), q(dUy(i, j + 1, k), yg[j + 1] - yg[j])));
    const double Uzd = AX(AX(q(dUz(i - 1, j, k), zg[k] - zg[k - 1]), q(dUz(i, j, k), zg[k] - zg[k - 1])),
                          AX(q(dUz(i - 1, j, k + 1), zg[k + 1] - zg[k]), q(dUz(i, j, k + 1), zg[k + 1] - zg[k])));
    const double Vxd = AX(AX(q(dVx(i, j - 1, k), xg[i] - xg[i - 1]), q(dVx(i, j, k), xg[i] - xg[i - 1])),
                          AX(q(dVx(i + 1, j - 1, k), xg[i + 1] - xg[i]), q(dVx(i + 1, j, k), xg[i + 1] - xg[i])));
    const double Vzd = AX(AX(q(dVz(i, j - 1, k), zg[k] - zg[k - 1]), q(dVz(i, j, k), zg[k] - zg[k - 1])),
                          AX(q(dVz(i, j - 1, k + 1), zg[k + 1] - zg[k]), q(dVz(i, j, k + 1), zg[k + 1] - zg[k])));
    const double Wxd = AX(AX(q(dWx(i, j, k - 1), xg[i] - xg[i - 1]), q(dWx(i, j, k), xg[i] - xg[i - 1])),
                          AX(q(dWx(i + 1, j, k - 1), xg[i + 1] - xg[i]), q(dWx(i + 1, j, k), xg[i + 1] - xg[i])));
    const double Wyd = AX(AX(q(dWy(i, j, k - 1), yg[j] - yg[j - 1]), q(dWy(i, j, k), yg[j] - yg[j - 1])),
                          AX(q(dWy(i, j + 1, k - 1), yg[j + 1] - yg[j]), q(dWy(i, j + 1, k), yg[j + 1] - yg[j])));
#undef q
    /* cross terms formed before the last average */
    const double z12 = AX(AX(dUz(i - 1, j, k), dUz(i, j, k)) * AX(dVz(i, j - 1, k), dVz(i, j, k)),
                          AX(dUz(i - 1, j, k + 1), dUz(i, j, k + 1)) * AX(dVz(i, j - 1, k + 1), dVz(i, j, k + 1)));
    const double y13 = AX(AX(dUy(i - 1, j, k), dUy(i, j, k)) * AX(dWy(i, j, k - 1), dWy(i, j, k)),
                          AX(dUy(i - 1, j + 1, k), dUy(i, j + 1, k)) * AX(dWy(i, j + 1, k - 1), dWy(i, j + 1, k)));
    const double x23 = AX(AX(dVx(i, j - 1, k), dVx(i, j, k)) * AX(dWx(i, j, k - 1), dWx(i, j, k)),
                          AX(dVx(i + 1, j - 1, k), dVx(i + 1, j, k)) * AX(dWx(i + 1, j, k - 1), dWx(i + 1, j, k)));
    double N = Uxh * Uxh * S11;
    N = N + Uyh * Uyh * S11;
    N = N + Uzh * Uzh * S11;
    N = N + Vyh * Vyh * S22;
    N = N + Vxh * Vxh * S22;
    N = N + Vzh * Vzh * S22;
    N = N + (Wzh * Wzh) * S33;
    N = N + Wxh * Wxh * S33;
    N = N + Wyh * Wyh * S33;
    N = N + 2.0 * (Vxh * Uxh) * S12;
    N = N + 2.0 * (Uyh * Vyh) * S12;
    N = N + 2.0 * z12 * S12;
    N = N + 2.0 * (Wxh * Uxh) * S13;
    N = N + 2.0 * (Uzh * Wzh) * S13;
    N = N + 2.0 * y13 * S13;
    N = N + 2.0 * (Wyh * Vyh) * S23;
    N = N + 2.0 * (Vzh * Wzh) * S23;
    N = N + 2.0 * x23 * S23;
    const double dx = x[i] - x[i - 1], dy = y[j] - y[j - 1], dz = z[k] - z[k - 1];
    double D = Uxh * Uxh / (dx * dx);
    D = D + Uyd * Uyd;
    D = D + Uzd * Uzd;
    D = D + Vyh * Vyh / (dy * dy);
    D = D + Vxd * Vxd;
    D = D + Vzd * Vzd;
    D = D + (Wzh * Wzh) / (dz * dz);
    D = D + Wxd * Wxd;
    D = D + Wyd * Wyd;
#undef dUx
#undef dUy
#undef dUz
#undef dVx
#undef dVy
#undef dVz
#undef dWx
#undef dWy
#undef dWz
    return -0.3 * N / D;
}
#undef AX

static void amd_core(orc* o, const double* U, const double* V, const double* W, double* nu_t)
{
    const int nxg = o->nxg, nyg = o->nyg, nzg = o->nzg;
    compute_Sij_full(o, U, V, W, o->nx, o->ny, o->nz, o->Sijb, o->Sbuf);
    memset(nu_t, 0, nC(o) * sizeof(double));
    PAR_FOR
    for (int k = 2; k <= nzg - 1; k++)
        for (int j = 2; j <= nyg - 1; j++)
            for (int i = 2; i <= nxg - 1; i++) AC(nu_t, i, j, k) = amd_point(o, U, V, W, o->Sijb, i, j, k);
    /* clipping; 0/0 (a uniform neighbourhood) gives NaN, clipped to 0 (fmax) */
    PAR_FOR
    for (int k = 2; k <= nzg - 1; k++)
        for (int j = 2; j <= nyg - 1; j++)
            for (int i = 2; i <= nxg - 1; i++) AC(nu_t, i, j, k) = fmax(AC(nu_t, i, j, k), 0.0);
    periodic_x(nu_t, nxg, nyg, nzg, 0);
    periodic_z(nu_t, nxg, nyg, nzg, 0);
    if (o->Dirichlet_nu_t == 2) {
        /* code/subgrid.f90:868-869: Dirichlet with a wall offset */
        const double off = 2.0 * 0.38 * sqrt(o->dPdx) * (o->y[2] - o->y[1]) / 2.0;
        for (int k = 1; k <= nzg; k++)
            for (int i = 1; i <= nxg; i++) {
                AC(nu_t, i, 1, k) = -AC(nu_t, i, 2, k) + off;
                AC(nu_t, i, nyg, k) = -AC(nu_t, i, nyg - 1, k) + off;
            }
    } else {
        wall_y(nu_t, nxg, nyg, nzg, 0, o->Dirichlet_nu_t == 1);
    }
}

/* code/subgrid.f90:23-47 */
void orc_compute_eddy_viscosity(orc* o, const double* U, const double* V, const double* W, double* nu_t)
{
    if (o->LES_model == 2) {
        dsm_core(o, U, V, W, nu_t, o->Lij, o->Mij, 0, 0);
    } else if (o->LES_model == 3) {
        amd_core(o, U, V, W, nu_t);
    } else {
        memset(nu_t, 0, nC(o) * sizeof(double));
    }
}

/* code/wallmodel.f90:23-46 -- nu_t in wall rows 1,2 / nyg-1,nyg such that
 * (nu+nu_t) dU/dy = alpha_x, averaged from x faces to centres, then x-periodic. */
void orc_compute_wall_model(orc* o, const double* U, double* nu_t)
{
    int nx = o->nx, nxg = o->nxg, nyg = o->nyg, nzg = o->nzg;
    const double* yg = o->yg;
    double* tb = dalloc((size_t)nx * nzg);
    double* tt = dalloc((size_t)nx * nzg);
    for (int k = 1; k <= nzg; k++)
        for (int i = 1; i <= nx; i++) {
            tb[(i - 1) + (size_t)nx * (k - 1)] =
                o->alpha_x[IX(i, 1, k, nx, 2)] / ((AU(U, i, 2, k) - AU(U, i, 1, k)) / (yg[2] - yg[1])) - o->nu;
            tt[(i - 1) + (size_t)nx * (k - 1)] =
                o->alpha_x[IX(i, 2, k, nx, 2)] / ((AU(U, i, nyg - 1, k) - AU(U, i, nyg, k)) / (yg[2] - yg[1])) - o->nu;
        }
    for (int k = 1; k <= nzg; k++)
        for (int i = 2; i <= nx; i++) {
            double vb = 0.5 * (tb[(i - 2) + (size_t)nx * (k - 1)] + tb[(i - 1) + (size_t)nx * (k - 1)]);
            double vt = 0.5 * (tt[(i - 2) + (size_t)nx * (k - 1)] + tt[(i - 1) + (size_t)nx * (k - 1)]);
            AC(nu_t, i, 1, k) = vb; AC(nu_t, i, 2, k) = vb;
            AC(nu_t, i, nyg - 1, k) = vt; AC(nu_t, i, nyg, k) = vt;
        }
    periodic_x(nu_t, nxg, nyg, nzg, 0);
    free(tb); free(tt);
}

/* ------------------------------------------------------------- projection */

/* mixed-radix complex FFT (decimation in time), sign -1 forward, +1 backward,
 * unnormalised, as the DFT FFTW computes for the reference's plans
 * (code/initialization.f90:287-290). */
static int smallest_factor(int n)
{
    if (n % 4 == 0) return 4;
    if (n % 2 == 0) return 2;
    for (int p = 3; p * p <= n; p += 2)
        if (n % p == 0) return p;
    return n;
}

static void fft_rec(int n, const cplx* in, int stride, cplx* out, const cplx* tw, int tws, int N, cplx* work)
{
    /* tw: exp(sign 2 pi i t / N), t = 0..N-1; this level has size n = N / tws */
    if (n == 1) { out[0] = in[0]; return; }
    int p = smallest_factor(n), m = n / p;
    for (int r = 0; r < p; r++)
        fft_rec(m, in + (size_t)r * stride, stride * p, out + (size_t)r * m, tw, tws * p, N, work);
    cplx t[64];
    cplx* tt = p <= 64 ? t : (cplx*)malloc(sizeof(cplx) * p);
    for (int q = 0; q < n; q++) work[q] = out[q];
    for (int k = 0; k < m; k++) {
        for (int r = 0; r < p; r++) tt[r] = work[(size_t)r * m + k] * tw[((long)r * k * tws) % N];
        for (int q = 0; q < p; q++) {
            cplx s = 0.0;
            for (int r = 0; r < p; r++) s += tt[r] * tw[(((long)r * q) % p) * (long)m * tws % N];
            out[k + (size_t)q * m] = s;
        }
    }
    if (tt != t) free(tt);
}

static cplx* twiddles(int n, int sign)
{
    const double pi = 4.0 * atan(1.0);
    cplx* tw = (cplx*)malloc(sizeof(cplx) * n);
    for (int t = 0; t < n; t++) {
        double ang = sign * 2.0 * pi * (double)t / (double)n;
        tw[t] = cos(ang) + I * sin(ang);
    }
    return tw;
}

void orc_fft(int n, double* data, int sign)
{
    cplx* in = (cplx*)malloc(sizeof(cplx) * n);
    cplx* out = (cplx*)malloc(sizeof(cplx) * n);
    cplx* work = (cplx*)malloc(sizeof(cplx) * n);
    cplx* tw = twiddles(n, sign);
    for (int i = 0; i < n; i++) in[i] = data[2 * i] + I * data[2 * i + 1];
    fft_rec(n, in, 1, out, tw, 1, n, work);
    for (int i = 0; i < n; i++) { data[2 * i] = creal(out[i]); data[2 * i + 1] = cimag(out[i]); }
    free(in); free(out); free(work); free(tw);
}

/* LAPACK ZGTSV (Gaussian elimination with partial pivoting), restated from
 * its published algorithm; the reference calls it at code/projection.f90:110.
 * n unknowns, one right-hand side; interleaved complex arrays. */
static double cabs1(cplx z) { return fabs(creal(z)) + fabs(cimag(z)); }

static int zgtsv_c(int n, cplx* DL, cplx* D, cplx* DU, cplx* B)
{
    for (int k = 0; k < n - 1; k++) {
        if (DL[k] == 0.0) {
            if (D[k] == 0.0) return k + 1;
        } else if (cabs1(D[k]) >= cabs1(DL[k])) {
            cplx mult = DL[k] / D[k];
            D[k + 1] = D[k + 1] - mult * DU[k];
            B[k + 1] = B[k + 1] - mult * B[k];
            if (k < n - 2) DL[k] = 0.0;
        } else {
            cplx mult = D[k] / DL[k];
            D[k] = DL[k];
            cplx temp = D[k + 1];
            D[k + 1] = DU[k] - mult * temp;
            if (k < n - 2) {
                DL[k] = DU[k + 1];
                DU[k + 1] = -mult * DL[k];
            }
            DU[k] = temp;
            temp = B[k];
            B[k] = B[k + 1];
            B[k + 1] = temp - mult * B[k + 1];
        }
    }
    if (D[n - 1] == 0.0) return n;
    B[n - 1] = B[n - 1] / D[n - 1];
    if (n > 1) B[n - 2] = (B[n - 2] - DU[n - 2] * B[n - 1]) / D[n - 2];
    for (int k = n - 3; k >= 0; k--) B[k] = (B[k] - DU[k] * B[k + 1] - DL[k] * B[k + 2]) / D[k];
    return 0;
}

int orc_zgtsv(int n, double* dl, double* d, double* du, double* b)
{
    cplx *DL = (cplx*)malloc(sizeof(cplx) * (n > 1 ? n - 1 : 1)), *D = (cplx*)malloc(sizeof(cplx) * n),
         *DU = (cplx*)malloc(sizeof(cplx) * (n > 1 ? n - 1 : 1)), *B = (cplx*)malloc(sizeof(cplx) * n);
    for (int i = 0; i < n; i++) { D[i] = d[2 * i] + I * d[2 * i + 1]; B[i] = b[2 * i] + I * b[2 * i + 1]; }
    for (int i = 0; i < n - 1; i++) { DL[i] = dl[2 * i] + I * dl[2 * i + 1]; DU[i] = du[2 * i] + I * du[2 * i + 1]; }
    int info = zgtsv_c(n, DL, D, DU, B);
    for (int i = 0; i < n; i++) { b[2 * i] = creal(B[i]); b[2 * i + 1] = cimag(B[i]); }
    free(DL); free(D); free(DU); free(B);
    return info;
}

/* code/projection.f90:74-128 with code/projection.f90:133-151 collapsed to
 * the periodic images P(nxg-1)=P(2), P(:,:,nzg-1)=P(:,:,2).  rhs_p is a
 * centre box; in: divergence on (2:nxg-1,2:nyg-1,2:nzg-1); out: pressure. */
/* 2-D forward transform of a plane over the unique points (i,k) = (2..nxp+1,
 * 2..nzp+1) of an (nxg x nzg) array, into hat[kz*nxp + kx] */
static void plane_fft(const orc* o, const double* F, cplx* hat, cplx* line, cplx* lout, cplx* work, const cplx* twx,
                      const cplx* twz)
{
    int nxp = o->nxp, nzp = o->nzp, nxg = o->nxg;
    for (int k = 0; k < nzp; k++) {
        for (int i = 0; i < nxp; i++) line[i] = F[(size_t)(i + 1) + (size_t)nxg * (k + 1)];
        fft_rec(nxp, line, 1, lout, twx, 1, nxp, work);
        for (int i = 0; i < nxp; i++) hat[(size_t)k * nxp + i] = lout[i];
    }
    for (int i = 0; i < nxp; i++) {
        for (int k = 0; k < nzp; k++) line[k] = hat[(size_t)k * nxp + i];
        fft_rec(nzp, line, 1, lout, twz, 1, nzp, work);
        for (int k = 0; k < nzp; k++) hat[(size_t)k * nxp + i] = lout[k];
    }
}

static void poisson_core(orc* o, double* rhs_p, const double* bc1, const double* bc2);

void orc_solve_poisson_equation(orc* o, double* rhs_p) { poisson_core(o, rhs_p, 0, 0); }

/* code/projection.f90:74-128; with bc1/bc2 (nxg x nzg planes, 1-based) also
 * the inhomogeneous Neumann rows of code/pressure.f90:134-139 */
static void poisson_core(orc* o, double* rhs_p, const double* bc1, const double* bc2)
{
    int nxp = o->nxp, nzp = o->nzp, nyg = o->nyg, nxg = o->nxg, nzg = o->nzg;
    int nr = nyg - 2;
    size_t plane = (size_t)nxp * nzp;
    cplx* hat = (cplx*)malloc(sizeof(cplx) * plane * nr);   /* [j][kz][kx] */
    cplx *twxf = twiddles(nxp, -1), *twzf = twiddles(nzp, -1), *twxb = twiddles(nxp, +1), *twzb = twiddles(nzp, +1);
    const int lmax = nxp > nzp ? nxp : nzp;
    /* planes, modes and lines are independent: each iteration has its own
     * line buffers */
    PAR_FOR
    for (int j = 2; j <= nyg - 1; j++) {
        cplx* h = hat + plane * (j - 2);
        cplx* line = (cplx*)malloc(sizeof(cplx) * 3 * (size_t)lmax);
        cplx *lout = line + lmax, *work = lout + lmax;
        for (int k = 0; k < nzp; k++) {
            for (int i = 0; i < nxp; i++) line[i] = AC(rhs_p, i + 2, j, k + 2);
            fft_rec(nxp, line, 1, lout, twxf, 1, nxp, work);
            for (int i = 0; i < nxp; i++) h[(size_t)k * nxp + i] = lout[i];
        }
        for (int i = 0; i < nxp; i++) {
            for (int k = 0; k < nzp; k++) line[k] = h[(size_t)k * nxp + i];
            fft_rec(nzp, line, 1, lout, twzf, 1, nzp, work);
            for (int k = 0; k < nzp; k++) h[(size_t)k * nxp + i] = lout[k];
        }
        free(line);
    }
    cplx *b1h = 0, *b2h = 0;
    double cb1 = 0.0, cb2 = 0.0;
    if (bc1 && bc2) {
        b1h = (cplx*)malloc(sizeof(cplx) * plane);
        b2h = (cplx*)malloc(sizeof(cplx) * plane);
        cplx* line = (cplx*)malloc(sizeof(cplx) * 3 * (size_t)lmax);
        plane_fft(o, bc1, b1h, line, line + lmax, line + 2 * lmax, twxf, twzf);
        plane_fft(o, bc2, b2h, line, line + lmax, line + 2 * lmax, twxf, twzf);
        free(line);
        /* coef_bc_1 = c at j = 2, coef_bc_2 = a at j = nyg-1
         * (code/initialization.f90:370-385) */
        const double *y = o->y, *yg = o->yg;
        cb1 = 1.0 / (y[2] - y[1]) / (yg[2] - yg[1]);
        cb2 = 1.0 / (y[nyg - 1] - y[nyg - 2]) / (yg[nyg] - yg[nyg - 1]);
    }
    PAR_FOR
    for (int kz = 0; kz < nzp; kz++) {
        cplx* D = (cplx*)malloc(sizeof(cplx) * 4 * (size_t)nr);
        cplx *DL = D + nr, *DU = DL + nr, *B = DU + nr;
        for (int kx = 0; kx < nxp; kx++) {
            for (int j = 2; j <= nyg - 1; j++) D[j - 2] = (o->dyy_d[j] + o->kxx[kx]) + o->kzz[kz];
            for (int j = 2; j <= nyg - 2; j++) { DL[j - 2] = o->dyy_l[j]; DU[j - 2] = o->dyy_u[j]; }
            if (kx == 0 && kz == 0) D[0] = 3.0 / 2.0 * D[0];
            for (int j = 0; j < nr; j++) B[j] = hat[plane * j + (size_t)kz * nxp + kx];
            if (b1h) {
                B[0] = B[0] + cb1 * (o->yg[2] - o->yg[1]) * b1h[(size_t)kz * nxp + kx];
                B[nr - 1] = B[nr - 1] - cb2 * (o->yg[nyg] - o->yg[nyg - 1]) * b2h[(size_t)kz * nxp + kx];
            }
            zgtsv_c(nr, DL, D, DU, B);
            for (int j = 0; j < nr; j++) hat[plane * j + (size_t)kz * nxp + kx] = B[j];
        }
        free(D);
    }
    const double scale = (double)((long)nxp * nzp);
    PAR_FOR
    for (int j = 2; j <= nyg - 1; j++) {
        cplx* h = hat + plane * (j - 2);
        cplx* line = (cplx*)malloc(sizeof(cplx) * 3 * (size_t)lmax);
        cplx *lout = line + lmax, *work = lout + lmax;
        for (int i = 0; i < nxp; i++) {
            for (int k = 0; k < nzp; k++) line[k] = h[(size_t)k * nxp + i];
            fft_rec(nzp, line, 1, lout, twzb, 1, nzp, work);
            for (int k = 0; k < nzp; k++) h[(size_t)k * nxp + i] = lout[k];
        }
        for (int k = 0; k < nzp; k++) {
            for (int i = 0; i < nxp; i++) line[i] = h[(size_t)k * nxp + i];
            fft_rec(nxp, line, 1, lout, twxb, 1, nxp, work);
            for (int i = 0; i < nxp; i++) AC(rhs_p, i + 2, j, k + 2) = creal(lout[i] / scale);
        }
        free(line);
    }
    /* periodic images of the pseudo-pressure */
    for (int k = 2; k <= nzg; k++)
        for (int j = 2; j <= nyg - 1; j++) AC(rhs_p, nxg - 1, j, k) = AC(rhs_p, 2, j, k);
    for (int j = 2; j <= nyg - 1; j++)
        for (int i = 2; i <= nxg - 1; i++) AC(rhs_p, i, j, nzg - 1) = AC(rhs_p, i, j, 2);
    free(twxf); free(twzf); free(twxb); free(twzb);
    free(hat); free(b1h); free(b2h);
}

/* code/pressure.f90:236-340 compute_pressure_BC: bc = d/dy((nu + nu_t) dV/dy)
 * at each wall, one-sided second-order differences with metric factors;
 * planes (nxg x nzg), set on i = 2..nxg-1, k = 2..nzg-1. */
static void pressure_bc(const orc* o, const double* V, const double* nut, double* bc1, double* bc2)
{
    int nxg = o->nxg, nyg = o->nyg, ny = o->ny, nzg = o->nzg;
    const double *y = o->y, *yg = o->yg, nu = o->nu, dzeta = 1.0;
    double g = (-3.0 * y[1] + 4.0 * y[2] - y[3]) / (2.0 * dzeta);
    double g2 = (2.0 * y[1] - 5.0 * y[2] + 4.0 * y[3] - 1.0 * y[4]) / (dzeta * dzeta);
    for (int k = 2; k <= nzg - 1; k++)
        for (int i = 2; i <= nxg - 1; i++) {
            double nui = nu + 0.5 * (AC(nut, i, 1, k) + AC(nut, i, 2, k));
            double dnu = (AC(nut, i, 2, k) - AC(nut, i, 1, k)) / (yg[2] - yg[1]);
            double dV = (-3.0 * AV(V, i, 1, k) + 4.0 * AV(V, i, 2, k) - AV(V, i, 3, k)) / (2.0 * dzeta) * 1.0 / g;
            double ddV = (2.0 * AV(V, i, 1, k) - 5.0 * AV(V, i, 2, k) + 4.0 * AV(V, i, 3, k) - 1.0 * AV(V, i, 4, k)) /
                         (dzeta * dzeta);
            ddV = (ddV - dV * g2) / (g * g);
            bc1[(size_t)(i - 1) + (size_t)nxg * (k - 1)] = dnu * dV + nui * ddV;
        }
    g = (3.0 * y[ny] - 4.0 * y[ny - 1] + y[ny - 2]) / (2.0 * dzeta);
    g2 = (2.0 * y[ny] - 5.0 * y[ny - 1] + 4.0 * y[ny - 2] - 1.0 * y[ny - 3]) / (dzeta * dzeta);
    for (int k = 2; k <= nzg - 1; k++)
        for (int i = 2; i <= nxg - 1; i++) {
            double nui = nu + 0.5 * (AC(nut, i, nyg, k) + AC(nut, i, nyg - 1, k));
            double dnu = (AC(nut, i, nyg, k) - AC(nut, i, nyg - 1, k)) / (yg[nyg] - yg[nyg - 1]);
            double dV = (3.0 * AV(V, i, ny, k) - 4.0 * AV(V, i, ny - 1, k) + AV(V, i, ny - 2, k)) / (2.0 * dzeta) *
                        1.0 / g;
            double ddV = (2.0 * AV(V, i, ny, k) - 5.0 * AV(V, i, ny - 1, k) + 4.0 * AV(V, i, ny - 2, k) -
                          1.0 * AV(V, i, ny - 3, k)) / (dzeta * dzeta);
            ddV = (ddV - dV * g2) / (g * g);
            bc2[(size_t)(i - 1) + (size_t)nxg * (k - 1)] = dnu * dV + nui * ddV;
        }
}

/* code/pressure.f90:21-176 compute_pressure: the pressure of the current
 * U, V, W and nu_t from Laplacian(p) = div(rhs_u, rhs_v, rhs_w) with the
 * wall conditions dP/dy = d/dy(nu dV/dy).  The ghost planes follow the
 * whole-domain periodic images (x: 1 <- nxg-2, nxg-1 <- 2, nxg <- 3; z the
 * same), and the wall ghost rows P(1) = P(2) - dy bc_1, P(nyg) = P(nyg-1) +
 * dy bc_2 on the interior (i, k) as the reference intends. */
void orc_compute_pressure(orc* o, double* P)
{
    int nx = o->nx, ny = o->ny, nz = o->nz, nxg = o->nxg, nyg = o->nyg, nzg = o->nzg;
    const double *x = o->x, *y = o->y, *z = o->z, *yg = o->yg;
    size_t nc = nC(o);
    double* ru = dalloc(nc);
    double* rv = dalloc(nc);
    double* rw = dalloc(nc);
    double* bc1 = dalloc((size_t)nxg * nzg);
    double* bc2 = dalloc((size_t)nxg * nzg);
    pressure_bc(o, o->V, o->nu_t, bc1, bc2);
    /* rhs of u, v, w on their interior ranges (code/equations.f90), in the
     * fields' own shapes */
    orc_compute_rhs_u(o, o->U, o->V, o->W, ru);
    orc_compute_rhs_v(o, o->U, o->V, o->W, rv);
    orc_compute_rhs_w(o, o->U, o->V, o->W, rw);
#define RU(i, j, k) ru[IX(i, j, k, nx, nyg)]
#define RV(i, j, k) rv[IX(i, j, k, nxg, ny)]
#define RW(i, j, k) rw[IX(i, j, k, nxg, nyg)]
    /* boundaries: x-periodic rhs_u (faces), wall rows of rhs_v from bc,
     * z-periodic rhs_w (faces) */
    for (int k = 1; k <= nzg; k++)
        for (int j = 1; j <= nyg; j++) { RU(1, j, k) = RU(nx - 1, j, k); RU(nx, j, k) = RU(2, j, k); }
    for (int k = 2; k <= nzg - 1; k++)
        for (int i = 2; i <= nxg - 1; i++) {
            RV(i, 1, k) = bc1[(size_t)(i - 1) + (size_t)nxg * (k - 1)];
            RV(i, ny, k) = bc2[(size_t)(i - 1) + (size_t)nxg * (k - 1)];
        }
    for (int j = 1; j <= nyg; j++)
        for (int i = 1; i <= nxg; i++) { RW(i, j, 1) = RW(i, j, nz - 1); RW(i, j, nz) = RW(i, j, 2); }
    memset(P, 0, nc * sizeof(double));
    PAR_FOR
    for (int k = 2; k <= nzg - 1; k++)
        for (int j = 2; j <= nyg - 1; j++)
            for (int i = 2; i <= nxg - 1; i++)
                AC(P, i, j, k) = (RU(i, j, k) - RU(i - 1, j, k)) / (x[i] - x[i - 1]) +
                                 (RV(i, j, k) - RV(i, j - 1, k)) / (y[j] - y[j - 1]) +
                                 (RW(i, j, k) - RW(i, j, k - 1)) / (z[k] - z[k - 1]);
#undef RU
#undef RV
#undef RW
    poisson_core(o, P, bc1, bc2);
    for (int k = 2; k <= nzg - 1; k++)
        for (int i = 2; i <= nxg - 1; i++) {
            AC(P, i, 1, k) = AC(P, i, 2, k) - (yg[2] - yg[1]) * bc1[(size_t)(i - 1) + (size_t)nxg * (k - 1)];
            AC(P, i, nyg, k) = AC(P, i, nyg - 1, k) + (yg[nyg] - yg[nyg - 1]) * bc2[(size_t)(i - 1) + (size_t)nxg * (k - 1)];
        }
    periodic_x(P, nxg, nyg, nzg, 0);
    periodic_z(P, nxg, nyg, nzg, 0);
    free(ru); free(rv); free(rw); free(bc1); free(bc2);
}

/* code/projection.f90:20-55 and :192-214 */
void orc_compute_projection_step(orc* o)
{
    int nxg = o->nxg, nyg = o->nyg, nzg = o->nzg;
    const double *x = o->x, *y = o->y, *z = o->z, *xg = o->xg, *yg = o->yg, *zg = o->zg;
    double *U = o->U, *V = o->V, *W = o->W, *P = o->P;
    PAR_FOR
    for (int k = 2; k <= nzg - 1; k++)
        for (int j = 2; j <= nyg - 1; j++)
            for (int i = 2; i <= nxg - 1; i++)
                AC(P, i, j, k) = (AU(U, i, j, k) - AU(U, i - 1, j, k)) / (x[i] - x[i - 1]) +
                                 (AV(V, i, j, k) - AV(V, i, j - 1, k)) / (y[j] - y[j - 1]) +
                                 (AW(W, i, j, k) - AW(W, i, j, k - 1)) / (z[k] - z[k - 1]);
    orc_solve_poisson_equation(o, P);
    PAR_FOR
    for (int k = 2; k <= nzg - 1; k++)
        for (int j = 2; j <= nyg - 1; j++)
            for (int i = 2; i <= o->nx - 1; i++)
                AU(U, i, j, k) = AU(U, i, j, k) - (AC(P, i + 1, j, k) - AC(P, i, j, k)) / (xg[i + 1] - xg[i]);
    PAR_FOR
    for (int k = 2; k <= nzg - 1; k++)
        for (int j = 2; j <= o->ny - 1; j++)
            for (int i = 2; i <= nxg - 1; i++)
                AV(V, i, j, k) = AV(V, i, j, k) - (AC(P, i, j + 1, k) - AC(P, i, j, k)) / (yg[j + 1] - yg[j]);
    PAR_FOR
    for (int k = 2; k <= o->nz - 1; k++)
        for (int j = 2; j <= nyg - 1; j++)
            for (int i = 2; i <= nxg - 1; i++)
                AW(W, i, j, k) = AW(W, i, j, k) - (AC(P, i, j, k + 1) - AC(P, i, j, k)) / (zg[k + 1] - zg[k]);
}

/* code/projection.f90:223-245 (signed max as the reference, plus max |div|) */
void orc_check_divergence(const orc* o, double* max_div, double* max_abs)
{
    const double *x = o->x, *y = o->y, *z = o->z;
    const double *U = o->U, *V = o->V, *W = o->W;
    double d = 0.0, a = 0.0;
    for (int k = 2; k <= o->nzg - 1; k++)
        for (int j = 2; j <= o->nyg - 1; j++)
            for (int i = 2; i <= o->nxg - 1; i++) {
                double v = (AU(U, i, j, k) - AU(U, i - 1, j, k)) / (x[i] - x[i - 1]) +
                           (AV(V, i, j, k) - AV(V, i, j - 1, k)) / (y[j] - y[j - 1]) +
                           (AW(W, i, j, k) - AW(W, i, j, k - 1)) / (z[k] - z[k - 1]);
                d = fmax(d, v);
                a = fmax(a, fabs(v));
            }
    *max_div = d; *max_abs = a;
}

/* ------------------------------------------------------------- mass flow */

/* code/mass_flow.f90:24-62, whole domain: k over unique planes 2..nzg-2 */
double orc_compute_mean_mass_flow_U(const orc* o, const double* U)
{
    const double* yg = o->yg;
    double q = 0.0, norm = 0.0;
    for (int k = 2; k <= o->nzg - 2; k++)
        for (int j = 3; j <= o->nyg - 1; j++)
            for (int i = 2; i <= o->nx - 1; i++) {
                q = q + (AU(U, i, j, k) + AU(U, i, j - 1, k)) * 0.5 * (yg[j] - yg[j - 1]);
                norm = norm + (1.0 + 1.0) * 0.5 * (yg[j] - yg[j - 1]);
            }
    return q / norm;
}

/* code/mass_flow.f90:113-166: dPdx = Qflow_x_0 - mean(0.5(U(:,3,:)+U(:,nyg-2,:))) */
double orc_compute_dPx_for_constant_mass_flow(const orc* o, const double* U)
{
    double s = 0.0, norm = 0.0;
    for (int k = 2; k <= o->nzg - 2; k++)
        for (int i = 2; i <= o->nx - 1; i++) {
            s = s + 0.5 * (AU(U, i, 3, k) + AU(U, i, o->nyg - 2, k));
            norm = norm + 1.0;
        }
    return o->Qflow_x_0 - s / norm;
}

/* ----------------------------------------------------------------- dt */

/* code/time_integration.f90:102-150 (Max of dx/|u| as the reference) */
static double compute_dt_core(orc* o);
double orc_compute_dt(orc* o)
{
    const double t0 = wall_now();
    const double dt = compute_dt_core(o);
    g_stage[8] += wall_now() - t0;
    return dt;
}

static double compute_dt_core(orc* o)
{
    const double *xg = o->xg, *yg = o->yg, *zg = o->zg;
    double lU = 0.0, lV = 0.0, lW = 0.0;
    for (int i = 2; i <= o->nxg - 1; i++)
        for (int j = 2; j <= o->nyg - 1; j++)
            for (int k = 2; k <= o->nzg - 1; k++) {
                lU = fmax(lU, (xg[i + 1] - xg[i]) / fabs(AU(o->U, i, j, k)));
                lV = fmax(lV, (yg[j + 1] - yg[j]) / fabs(AV(o->V, i, j, k)));
                lW = fmax(lW, (zg[k + 1] - zg[k]) / fabs(AW(o->W, i, j, k)));
            }
    double dcu = o->CFL * lU, dcv = o->CFL * lV, dcw = o->CFL * lW;
    double dconv = fmin(fmin(dcu, dcv), dcw);
    double dvu = o->CFL * (o->dxmin * o->dxmin) / o->nu;
    double dvv = o->CFL * (o->dymin * o->dymin) / o->nu;
    double dvw = o->CFL * (o->dzmin * o->dzmin) / o->nu;
    double dvis = fmin(fmin(dvu, dvv), dvw);
    double dt = fmin(dconv, dvis);
    dt = fmin(dt, 1.0e-3);
    if (o->CFL < 0) dt = -o->CFL;
    o->dt = dt;
    return dt;
}

/* --------------------------------------------------------------- RK3 step */

void orc_set_wall_model(orc* o, int mode, double h_wm, double kappa, double B)
{
    o->wall_mode = mode; o->h_wm = h_wm; o->kappa = kappa; o->Bc = B;
}

/* code/time_integration.f90:24-96 */
void orc_compute_time_step_RK3(orc* o)
{
    double to = o->t;
    memcpy(o->Uo, o->U, nU(o) * sizeof(double));
    memcpy(o->Vo, o->V, nV(o) * sizeof(double));
    memcpy(o->Wo, o->W, nW(o) * sizeof(double));
    const double dt = o->dt;
    for (int s = 0; s < 3; s++) {
        if (o->wall_mode == 1) STAGE(0, orc_equilibrium_alpha(o, o->U, o->h_wm, o->kappa, o->Bc));
        STAGE(1, orc_compute_eddy_viscosity(o, o->U, o->V, o->W, o->nu_t));
        STAGE(2, orc_compute_wall_model(o, o->U, o->nu_t));
        STAGE(3, {
            orc_compute_rhs_u(o, o->U, o->V, o->W, o->Fu[s]);
            orc_compute_rhs_v(o, o->U, o->V, o->W, o->Fv[s]);
            orc_compute_rhs_w(o, o->U, o->V, o->W, o->Fw[s]);
        });
        const double* c = o->rk_coef[s];
        const double t_rk = wall_now();
        PAR_FOR
        for (int k = 2; k <= o->nzg - 1; k++)
            for (int j = 2; j <= o->nyg - 1; j++)
                for (int i = 2; i <= o->nx - 1; i++) {
                    size_t q = IX(i, j, k, o->nx, o->nyg);
                    if (s == 0) o->U[q] = o->Uo[q] + dt * c[0] * o->Fu[0][q];
                    else if (s == 1) o->U[q] = o->Uo[q] + dt * (c[0] * o->Fu[0][q] + c[1] * o->Fu[1][q]);
                    else o->U[q] = o->Uo[q] + dt * (c[0] * o->Fu[0][q] + c[1] * o->Fu[1][q] + c[2] * o->Fu[2][q]);
                }
        PAR_FOR
        for (int k = 2; k <= o->nzg - 1; k++)
            for (int j = 2; j <= o->ny - 1; j++)
                for (int i = 2; i <= o->nxg - 1; i++) {
                    size_t q = IX(i, j, k, o->nxg, o->ny);
                    if (s == 0) o->V[q] = o->Vo[q] + dt * c[0] * o->Fv[0][q];
                    else if (s == 1) o->V[q] = o->Vo[q] + dt * (c[0] * o->Fv[0][q] + c[1] * o->Fv[1][q]);
                    else o->V[q] = o->Vo[q] + dt * (c[0] * o->Fv[0][q] + c[1] * o->Fv[1][q] + c[2] * o->Fv[2][q]);
                }
        PAR_FOR
        for (int k = 2; k <= o->nz - 1; k++)
            for (int j = 2; j <= o->nyg - 1; j++)
                for (int i = 2; i <= o->nxg - 1; i++) {
                    size_t q = IX(i, j, k, o->nxg, o->nyg);
                    if (s == 0) o->W[q] = o->Wo[q] + dt * c[0] * o->Fw[0][q];
                    else if (s == 1) o->W[q] = o->Wo[q] + dt * (c[0] * o->Fw[0][q] + c[1] * o->Fw[1][q]);
                    else o->W[q] = o->Wo[q] + dt * (c[0] * o->Fw[0][q] + c[1] * o->Fw[1][q] + c[2] * o->Fw[2][q]);
                }
        g_stage[4] += wall_now() - t_rk;
        o->t = to + o->rk_t[s] * dt;
        STAGE(5, orc_apply_boundary_conditions(o));
        STAGE(6, orc_compute_projection_step(o));
        STAGE(5, orc_apply_boundary_conditions(o));
    }
    const double t_mf = wall_now();
    if (o->x_mass_cte == 1 || o->x_mass_cte == 2) {
        /* mode 2, constant bulk flow (no reference counterpart): the shift
         * that restores the bulk mean of code/mass_flow.f90:24-62 */
        o->dPdx = o->x_mass_cte == 1 ? orc_compute_dPx_for_constant_mass_flow(o, o->U)
                                     : o->Qflow_x_0 - orc_compute_mean_mass_flow_U(o, o->U);
        PAR_FOR
        for (int k = 2; k <= o->nzg - 1; k++)
            for (int j = 2; j <= o->nyg - 1; j++)
                for (int i = 2; i <= o->nx - 1; i++) AU(o->U, i, j, k) = AU(o->U, i, j, k) + o->dPdx;
        o->dPdx = 0.0;
        orc_apply_boundary_conditions(o);
    }
    g_stage[7] += wall_now() - t_mf;
}

/* ---------------------------------------------- equilibrium wall stress */

/* Log law U+ = ln(y+)/kappa + B above the crossover y+_c where it meets
 * U+ = y+, linear below.  Newton on f(ut) = ut (ln(h ut/nu)/kappa + B) - u. */
double orc_loglaw_utau(double u, double h, double nu, double kappa, double B)
{
    double au = fabs(u);
    if (au == 0.0) return 0.0;
    double ut = sqrt(nu * au / h);           /* viscous-sublayer value */
    double yc = 11.0;                        /* crossover y+_c: solve yc = ln(yc)/kappa + B */
    for (int it = 0; it < 50; it++) yc = yc - (yc - log(yc) / kappa - B) / (1.0 - 1.0 / (kappa * yc));
    if (h * ut / nu <= yc) return ut;
    double g = ut;
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

/* alpha_x(i,1,k) (bottom) and alpha_x(i,2,k) (top): signed tau_w from U at
 * the matching height; h_wm <= 0 selects the first off-wall centre. */
void orc_equilibrium_alpha(orc* o, const double* U, double h_wm, double kappa, double B)
{
    int nx = o->nx, nyg = o->nyg, ny = o->ny;
    const double *y = o->y, *yg = o->yg;
    int jb = 2;
    double wb = 1.0, hb = yg[2] - y[1];
    if (h_wm > 0.0) {
        double yt = y[1] + h_wm;
        jb = 2;
        while (jb < nyg - 2 && yg[jb + 1] < yt) jb++;
        wb = (yg[jb + 1] - yt) / (yg[jb + 1] - yg[jb]);
        hb = h_wm;
    }
    /* every (k, i) writes its own two entries */
    PAR_FOR
    for (int k = 1; k <= o->nzg; k++)
        for (int i = 1; i <= nx; i++) {
            double ub = wb * AU(U, i, jb, k) + (1.0 - wb) * AU(U, i, jb + 1, k);
            int jt = nyg + 1 - jb;  /* mirror row */
            double ut_ = wb * AU(U, i, jt, k) + (1.0 - wb) * AU(U, i, jt - 1, k);
            double tb = orc_loglaw_utau(ub, hb, o->nu, kappa, B);
            double tt = orc_loglaw_utau(ut_, (h_wm > 0.0 ? hb : y[ny] - yg[nyg - 1]), o->nu, kappa, B);
            o->alpha_x[IX(i, 1, k, nx, 2)] = (ub >= 0 ? 1.0 : -1.0) * tb * tb;
            o->alpha_x[IX(i, 2, k, nx, 2)] = (ut_ >= 0 ? 1.0 : -1.0) * tt * tt;
        }
}

/* --------------------------------------------------------- statistics */

/* xz-plane means over the unique periodic points, the quantities of
 * code/statistics.f90:52-64: per row j = 1..ny the means of U, V, W, their
 * squares, and <u v> with u interpolated to V's position (y-weights, then the
 * x average).  out = [U, V, W, U2, V2, W2, UV] x ny. */
/* code/statistics.f90:95-130 with the top wall on its own rows (the
 * reference evaluates the bottom wall twice): out = tau_b, tau_t, u_tau,
 * Re_tau, from U and the oracle's nu_t. */
void orc_wall_shear(const orc* o, double* out)
{
    int nx = o->nx, ny = o->ny, nyg = o->nyg, nzg = o->nzg;
    const double *U = o->U, *nu_t = o->nu_t, *yg = o->yg, *w0 = o->w0, *w1 = o->w1;
    double sb = 0.0, st = 0.0;
    for (int k = 2; k <= nzg - 2; k++)
        for (int i = 2; i <= nx - 1; i++) {
#define TX(j) (0.5 * (AC(nu_t, i, j, k) + AC(nu_t, i + 1, j, k)))
            double nb = w0[1] * TX(1) + w1[1] * TX(2);
            double nt = w0[ny] * TX(nyg - 1) + w1[ny] * TX(nyg);
#undef TX
            sb += (AU(U, i, 2, k) - AU(U, i, 1, k)) / (yg[2] - yg[1]) * (o->nu + nb);
            st += (AU(U, i, nyg - 1, k) - AU(U, i, nyg, k)) / (yg[nyg] - yg[nyg - 1]) * (o->nu + nt);
        }
    double nq = (double)(nx - 2) * (double)o->nzp;
    double Ly = o->y[ny] - o->y[1];
    out[0] = sb / nq;
    out[1] = st / nq;
    out[2] = sqrt((out[0] + out[1]) / Ly);
    out[3] = out[2] * Ly / 2.0 / o->nu;
}

void orc_plane_stats(const orc* o, double* out)
{
    int nx = o->nx, ny = o->ny, nxg = o->nxg, nzg = o->nzg, nz = o->nz;
    const double *U = o->U, *V = o->V, *W = o->W, *w0 = o->w0, *w1 = o->w1;
    double nq = (double)o->nxp * (double)o->nzp;
    for (int j = 1; j <= ny; j++) {
        double su = 0, sv = 0, sw = 0, su2 = 0, sv2 = 0, sw2 = 0, suv = 0;
        for (int k = 2; k <= nzg - 2; k++)
            for (int i = 2; i <= nx - 1; i++) { double u = AU(U, i, j, k); su += u; su2 += u * u; }
        for (int k = 2; k <= nz - 1; k++)
            for (int i = 2; i <= nxg - 2; i++) { double w = AW(W, i, j, k); sw += w; sw2 += w * w; }
        for (int k = 2; k <= nzg - 2; k++)
            for (int i = 2; i <= nxg - 2; i++) {
                double v = AV(V, i, j, k);
                sv += v; sv2 += v * v;
                double uy0 = w0[j] * AU(U, i - 1, j, k) + w1[j] * AU(U, i - 1, j + 1, k);
                double uy1 = w0[j] * AU(U, i, j, k) + w1[j] * AU(U, i, j + 1, k);
                suv += v * (0.5 * (uy0 + uy1));
            }
        out[0 * ny + j - 1] = su / nq; out[1 * ny + j - 1] = sv / nq; out[2 * ny + j - 1] = sw / nq;
        out[3 * ny + j - 1] = su2 / nq; out[4 * ny + j - 1] = sv2 / nq; out[5 * ny + j - 1] = sw2 / nq;
        out[6 * ny + j - 1] = suv / nq;
    }
}
