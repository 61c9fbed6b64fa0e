/*
 * lee_oracle.c -- FP64 CPU oracle of the reference FB_p_v time step.
 *
 * TEST INFRASTRUCTURE ONLY (see lee_oracle.h).  The product path never links this file.
 *
 * Every routine restates one reference routine and cites it.  Expressions keep the
 * reference's association order (the reference's C++ parses e.g. `-a*b*c` as
 * ((-a)*b)*c), so that an FP64 device path that follows the same order and does not
 * contract to FMA reproduces these results bit for bit.  Build with -ffp-contract=off.
 *
 * Layout: fields are SoA with x fastest, idx = i + nx*(j + ny*k) (the reference stores
 * f[i][j][k] with k fastest; the arithmetic per point does not depend on the layout).
 */
#include "lee_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static char g_err[512];
const char *orc_error(void) { return g_err; }
#define FAIL(...) do { snprintf(g_err, sizeof g_err, __VA_ARGS__); return NULL; } while (0)

#define REF_PI 3.1415926 /* mygeneral.h:17 */

/* ------------------------------------------------------------------------------------ */
/* grid object: calculation_2D (calculation_2D.h:25-130)                                 */
/* ------------------------------------------------------------------------------------ */
typedef struct {
    int nx, ny, nz;
    double *un, *vn, *wn, *pn, *unn, *vnn, *wnn, *pnn;
    double *X, *Y, *Z;       /* current (frame) coordinates */
    double *ax, *ay, *az;    /* axisbound_X/Y/Z: coordinates as read from file */
    double *Xx, *Yy, *Zz;    /* X_over_x, Y_over_y, Z_over_z (central differences) */
    double *Jac;             /* 3-D Jacobian (interior) */
    double dxi, dyi, dzi;    /* diff_x = 1/dx ... (calculation_2D.cpp:20-22) */
    double dx1, dy1, dz1;    /* diff_x1 = dx ... (calculation_2D.cpp:23-25) */
    double dt;
    /* cross derivatives X_over_y ... are scalars hard-set to 0 (calculation_2D.cpp:102-104) */
    double X_y, X_z, Y_x, Y_z, Z_x, Z_y;
    /* Terrain-following extension (no reference counterpart).  Physical height
     *   Z(i,j,k) = Z[k] + h(i,j) * tk[k],  tk[k] = (Z[nz-1] - Z[k]) / (Z[nz-1] - Z[0])
     * in the air (the top stays flat), tk = 1 in the porous layer (a rigid shift).  With the
     * reference's central differences the metric terms factor exactly as
     *   Z_over_z = Zz[k] * sh(i,j),  Z_over_x = hx(i,j) * tk[k],  Z_over_y = hy(i,j) * tk[k],
     * sh = 1 - h/D (air) or 1 (porous), hx, hy = central differences of h. */
    int terrain, pore_layer, nh;
    double hills[32];
    double *h, *sh, *hx, *hy, *tk;
} ogrid;

#define IDX(g, i, j, k) ((size_t)(i) + (size_t)(g)->nx * ((size_t)(j) + (size_t)(g)->ny * (size_t)(k)))

/* PML auxiliary slab (air.h:56-58): q[0..3] = Qu, Qv, Qw, Qp over `n` planes of one axis. */
typedef struct { int n; double *q[4]; } oslab;

struct orc_sim {
    lee_deck d;
    lee_options opt;          /* extensions: integrator, terrain, log wind */
    ogrid air, pore;
    int south_porous, east_tdbc;
    double gamma, Pav, Aav, c;
    /* wind (air.cpp:40-256) */
    int vmethod; double vcoef;
    double *U, *V, *W;
    /* PML (air.h:56-69) */
    int judge[3];
    int I1, I2, J1, J2, J3, K1, K2;    /* IMAX1.. after cal_PML_var; J1 is the gated JMAX1 */
    long pml_counter; int pml_time;
    double Rx, Ry, Rz, ax_, ay_, az_;  /* resistivities and alphas */
    double px1, px2, W1, W2, py1, py2, L1, L2, pz1, pz2, H1, H2;
    oslab sx1, sx2, sy1, sy2, sz1, sz2;
    /* TDBC (calculation_2D.h:71-75, airpore.cpp:252-254) */
    double Zinf, tau;
    double *Psi;                       /* [i + nx*k][6] */
    /* porous coefficients (porous.cpp:40-51) */
    double p_eff, p_beta, p_gama, p_invbeta;
    double p_E, p_E2;                  /* RK4 integrating factors e^-beta, e^-beta/2 */
    /* moving frame (mygeneral.h:62-68) */
    int mf_judge, mf_jmax, lead, trail, ijkmax; double ratio;
    int mf_count;
    /* receivers */
    int nrec; int *ri, *rj, *rk;
    /* RK4 registers (extension): increments and accumulators of the fields [domain][u v w p]
     * and of the auxiliaries, the auxiliaries of the stage state, and those of level n */
    double *rk_k[2][4], *rk_a[2][4];
    double *qk[6][4], *qa[6][4], *qs[6][4], *qsave[6][4];
    /* embedded porous obstacle (air.h:58-67): curve kind, index box istar..kstop, material
     * id per box node (0 air, 1 canopy, 2 trunk) and the per-material coefficients
     * crpo_* (index 0 hill, 1 canopy, 2 trunk) */
    int cr_judge, ob[6];
    int *cr_uc;
    double cr_gama[3], cr_invb[3], cr_eff[3];
    /* time-series source (calculation_2D.h ps, calculation_2D.cpp:220-235) and the step
     * index of airpore.cpp:572 (index_pressure, steps done) */
    double *src_ps; long src_n; int src_period;
    long index_pressure;
    /* ground impedance (extension): South face, the East TDBC recursion with K poles, closing
     * a ground interface plane k = 0 that carries the porous equations of the ground's top */
    int ground, gK;
    double g_gama, g_eff, g_ib;
    double gZinf, g_tb, g_inv, g_decay[8], g_gain[8];
    double *Gpsi;                      /* [i + nx*j][gK] */
};

static const double TD_A[6] = {8.226223820e-01, 3.732304064e-01, 9.106232372e-02,
                               4.819445335e-02, 4.225250937e-01, 1.837970986e-01};
static const double TD_R[6] = {8.681310138e-01, 1.953722814e-01, 1.108067950e-02,
                               2.566498621e-03, 1.457297589e-04, 4.648583789e-02};

static double *dalloc(size_t n, double fill)
{
    double *p = (double *)malloc((n ? n : 1) * sizeof(double));
    if (!p) return NULL;
    for (size_t i = 0; i < n; ++i) p[i] = fill;
    return p;
}

/* calculation_2D::calculation_2D (calculation_2D.cpp:12-236), single domain. */
static int grid_init(ogrid *g, int nx, int ny, int nz, const double *x, const double *y,
                     const double *z, double dt, double dx, double dy, double dz, int nanfill)
{
    memset(g, 0, sizeof *g);
    g->nx = nx; g->ny = ny; g->nz = nz;
    g->dt = dt;
    g->dxi = 1.0 / dx; g->dyi = 1.0 / dy; g->dzi = 1.0 / dz;
    g->dx1 = dx; g->dy1 = dy; g->dz1 = dz;
    size_t n = (size_t)nx * ny * nz;
    double nnfill = nanfill ? NAN : 0.0;
    g->un = dalloc(n, 0); g->vn = dalloc(n, 0); g->wn = dalloc(n, 0); g->pn = dalloc(n, 0);
    g->unn = dalloc(n, nnfill); g->vnn = dalloc(n, nnfill);
    g->wnn = dalloc(n, nnfill); g->pnn = dalloc(n, nnfill);
    g->Jac = dalloc(n, 0);
    g->X = dalloc(nx, 0); g->Y = dalloc(ny, 0); g->Z = dalloc(nz, 0);
    g->ax = dalloc(nx, 0); g->ay = dalloc(ny, 0); g->az = dalloc(nz, 0);
    g->Xx = dalloc(nx, 0); g->Yy = dalloc(ny, 0); g->Zz = dalloc(nz, 0);
    if (!g->un || !g->pnn || !g->Jac || !g->Zz) return -1;
    for (int i = 0; i < nx; ++i) g->X[i] = g->ax[i] = x[i];
    for (int j = 0; j < ny; ++j) g->Y[j] = g->ay[j] = y[j];
    for (int k = 0; k < nz; ++k) g->Z[k] = g->az[k] = z[k];
    return 0;
}

static void grid_free(ogrid *g)
{
    double **ps[] = {&g->un, &g->vn, &g->wn, &g->pn, &g->unn, &g->vnn, &g->wnn, &g->pnn,
                     &g->Jac, &g->X, &g->Y, &g->Z, &g->ax, &g->ay, &g->az, &g->Xx, &g->Yy, &g->Zz,
                     &g->h, &g->sh, &g->hx, &g->hy, &g->tk};
    for (size_t i = 0; i < sizeof ps / sizeof ps[0]; ++i) { free(*ps[i]); *ps[i] = NULL; }
}

/* Geo_over_var (df_over_var.cpp:5-9). */
static double geo(double a, double b, double c, double d, double jac) { return (a * b - c * d) * jac; }

typedef struct { double xX, xY, xZ, yX, yY, yZ, zX, zY, zZ; } metric9;

/* Z_over_z, Z_over_x, Z_over_y at (i,j,k): the reference's scalars and 1-D array, or the
 * terrain factors. */
static void z_derivs(const ogrid *g, int i, int j, int k, double *Zz, double *Zx, double *Zy)
{
    if (!g->terrain) { *Zz = g->Zz[k]; *Zx = g->Z_x; *Zy = g->Z_y; return; }
    size_t c = (size_t)i + (size_t)g->nx * j;
    *Zz = g->Zz[k] * g->sh[c];
    *Zx = g->hx[c] * g->tk[k];
    *Zy = g->hy[c] * g->tk[k];
}

/* The nine inverse-metric terms exactly as air::cal_fp computes them (air.cpp:1975-1992). */
static metric9 point_metric(const ogrid *g, int i, int j, int k)
{
    metric9 m;
    double J = g->Jac[IDX(g, i, j, k)];
    double Zz, Z_x, Z_y;
    z_derivs(g, i, j, k, &Zz, &Z_x, &Z_y);
    m.xX = geo(g->Yy[j], Zz, g->Y_z, Z_y, J);
    m.xY = geo(g->X_z, Z_y, g->X_y, Zz, J);
    m.xZ = geo(g->X_y, g->Y_z, g->X_z, g->Yy[j], J);
    m.yX = geo(g->Y_z, Z_x, g->Y_x, Zz, J);
    m.yY = geo(g->Xx[i], Zz, g->X_z, Z_x, J);
    m.yZ = geo(g->X_z, g->Y_x, g->Xx[i], g->Y_z, J);
    m.zX = geo(g->Y_x, Z_y, g->Yy[j], Z_x, J);
    m.zY = geo(g->X_y, Z_x, g->Xx[i], Z_y, J);
    m.zZ = geo(g->Xx[i], g->Yy[j], g->X_y, g->Y_x, J);
    return m;
}

/* Sum of the Gaussian hills at world (x, y). */
static double terrain_h(const ogrid *g, double x, double y)
{
    double h = 0.0;
    for (int n = 0; n < g->nh; ++n) {
        const double *H = &g->hills[4 * n];
        double dx = x - H[0], dy = y - H[1];
        h += H[2] * exp(-(dx * dx + dy * dy) / (2.0 * H[3] * H[3]));
    }
    return h;
}

/* Terrain tables of the current frame (X, Y, Z as set for the frame). */
static void terrain_tables(ogrid *g)
{
    int nx = g->nx, ny = g->ny, nz = g->nz;
    double top = g->Z[nz - 1], D = g->Z[nz - 1] - g->Z[0];
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            size_t c = (size_t)i + (size_t)nx * j;
            g->h[c] = terrain_h(g, g->X[i], g->Y[j]);
            g->sh[c] = g->pore_layer ? 1.0 : 1.0 - g->h[c] / D;
        }
    for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
            size_t c = (size_t)i + (size_t)nx * j;
            g->hx[c] = (i > 0 && i < nx - 1) ? (g->h[c + 1] - g->h[c - 1]) * g->dxi * 0.5 : 0.0;
            g->hy[c] = (j > 0 && j < ny - 1) ? (g->h[c + nx] - g->h[c - nx]) * g->dyi * 0.5 : 0.0;
        }
    for (int k = 0; k < nz; ++k) g->tk[k] = g->pore_layer ? 1.0 : (top - g->Z[k]) / D;
}

/* Physical height of node (i,j,k). */
static double z_phys(const ogrid *g, int i, int j, int k)
{
    if (!g->terrain) return g->Z[k];
    return g->Z[k] + g->h[(size_t)i + (size_t)g->nx * j] * g->tk[k];
}

/* Metric part of calculation_2D::SetMovingFrame (calculation_2D.cpp:410-433). */
static void grid_metrics(ogrid *g, int N, int trail)
{
    if (N != 0)
        for (int j = 0; j < g->ny; ++j) g->Y[j] = g->ay[j] + N * (trail - 1) * g->dy1;
    for (int i = 1; i < g->nx - 1; ++i) g->Xx[i] = (g->X[i + 1] - g->X[i - 1]) * g->dxi * 0.5;
    for (int j = 1; j < g->ny - 1; ++j) g->Yy[j] = (g->Y[j + 1] - g->Y[j - 1]) * g->dyi * 0.5;
    for (int k = 1; k < g->nz - 1; ++k) g->Zz[k] = (g->Z[k + 1] - g->Z[k - 1]) * g->dzi * 0.5;
    if (g->terrain) terrain_tables(g);
    for (int k = 1; k < g->nz - 1; ++k)
        for (int j = 1; j < g->ny - 1; ++j)
            for (int i = 1; i < g->nx - 1; ++i) {
                double Zz, Z_x, Z_y;
                z_derivs(g, i, j, k, &Zz, &Z_x, &Z_y);
                double J = g->Xx[i] * g->Yy[j] * Zz + Z_x * g->X_y * g->Y_z +
                           g->Y_x * Z_y * g->X_z;
                J += -g->Xx[i] * Z_y * g->Y_z - g->Y_x * g->X_y * Zz -
                     Z_x * g->Yy[j] * g->X_z;
                g->Jac[IDX(g, i, j, k)] = 1.0 / J;
            }
}

/* Field shift of calculation_2D::SetMovingFrame (calculation_2D.cpp:458-481), one rank:
 * the right-neighbour lead planes (calculation_2D.cpp:482-495) do not exist. */
static void shift_field(double *f, int nx, int ny, int nz, int temp_J, int lead, int trail)
{
    for (int k = 0; k < nz; ++k)
        for (int i = 0; i < nx; ++i)
            for (int j = 0; j < ny; ++j) {
                size_t id = (size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k);
                if (j < temp_J + 1 && j + lead - 1 < ny - 1)
                    f[id] = f[id + (size_t)nx * (size_t)(trail - 1)];
                else if (j >= temp_J + 1)
                    f[id] = 0.0;
            }
}

static void grid_frame(ogrid *g, int N, int mf_jmax, int lead, int trail)
{
    grid_metrics(g, N, trail);
    int temp_J = (N == 0) ? 0 : mf_jmax - lead;
    if (temp_J != 0) {
        shift_field(g->un, g->nx, g->ny, g->nz, temp_J, lead, trail);
        shift_field(g->vn, g->nx, g->ny, g->nz, temp_J, lead, trail);
        shift_field(g->wn, g->nx, g->ny, g->nz, temp_J, lead, trail);
        shift_field(g->pn, g->nx, g->ny, g->nz, temp_J, lead, trail);
        for (int k = 0; k < g->nz; ++k)
            for (int i = 0; i < g->nx; ++i) g->pn[IDX(g, i, 0, k)] = g->pn[IDX(g, i, 1, k)];
    }
}

/* Initial condition (calculation_2D.cpp:292-314).  Over terrain the source height is taken
 * above the ground under the source and the distance is physical. */
static void grid_ic(ogrid *g, const double *src, double gw)
{
    double gc = 4.0 * log(2.0) / pow(gw, 2);
    double zs = src[2] + (g->terrain ? terrain_h(g, src[0], src[1]) : 0.0);
    for (int k = 0; k < g->nz; ++k)
        for (int j = 0; j < g->ny; ++j)
            for (int i = 0; i < g->nx; ++i) {
                double x = g->X[i] - src[0], y = g->Y[j] - src[1], z = z_phys(g, i, j, k) - zs;
                double r = sqrt(pow(x, 2) + pow(y, 2) + pow(z, 2));
                size_t id = IDX(g, i, j, k);
                g->un[id] = g->vn[id] = g->wn[id] = 0.0;
                g->pn[id] = 1.0 * exp(-gc * r * r);
            }
}

/* ------------------------------------------------------------------------------------ */
/* wind: SpeedProfile (air.cpp:40-256)                                                   */
/* ------------------------------------------------------------------------------------ */
static double wind_alpha2(double y, double z, double *Rr)
{
    double alpha;
    *Rr = sqrt(pow(y - 5.0, 2) + pow(z - 2.0, 2));
    if (fabs(y - 5.0) <= 1.0e-6 && fabs(z - 2.0) <= 1.0e-6) alpha = 0.0;
    else if (y - 5.0 >= 0.0) alpha = asin((y - 5.0) / *Rr);
    else alpha = asin((y - 5.0) / *Rr) + 3.1415926;
    return alpha;
}
static double wind_U(int m, double b, double x, double y, double z)
{
    (void)x; (void)z;
    switch (m) { case 5: return b * y; default: return 0.0; }
}
static double wind_V(int m, double b, double x, double y, double z)
{
    (void)x;
    switch (m) {
    case 1: if (z >= 2.8) return b * 2.8; else if (z >= 0.0) return b * z; else return 0.0;
    case 2: { double Rr, a = wind_alpha2(y, z, &Rr); return -1.0 * 2.0 / (Rr + 2.0) * sin(a); }
    case 5: return b * log(fabs(z) / 0.1 + 1);
    default: return 0.0;
    }
}
static double wind_W(int m, double b, double x, double y, double z)
{
    (void)x;
    switch (m) {
    case 2: { double Rr, a = wind_alpha2(y, z, &Rr); return 1.0 * 2.0 / (Rr + 2.0) * cos(a); }
    case 5: return b * y;
    default: return 0.0;
    }
}

/* air::SetWindProfile (air.cpp:1286-1304). */
static void set_wind(orc_sim *s, int N)
{
    ogrid *g = &s->air;
    if (N != 0) {
        /* PML_y1 follows the frame (air.cpp:1290).  The reference indexes axisbound_Y with
         * the gated JMAX1, which is 0 while the y1 layer is off (an out-of-bounds read);
         * the value is only used once the layer is on, so the header width is used. */
        if (s->J3 > 0) s->py1 = g->ay[s->J3 - 1] + N * (s->trail - 1) * g->dy1;
        if (s->J2 > 0) s->py2 = g->ay[g->ny - s->J2] + N * (s->trail - 1) * g->dy1;
    }
    for (int k = 0; k < g->nz; ++k)
        for (int j = 0; j < g->ny; ++j)
            for (int i = 0; i < g->nx; ++i) {
                double Z0 = 0.0;
                size_t id = IDX(g, i, j, k);
                if (s->vmethod == 7) {
                    /* logarithmic profile along +y over the local ground (extension) */
                    double zg = z_phys(g, i, j, k) - z_phys(g, i, j, 0);
                    s->U[id] = 0.0;
                    s->V[id] = zg > s->opt.wind_z0 ? s->opt.wind_ustar / 0.41 * log(zg / s->opt.wind_z0) : 0.0;
                    s->W[id] = 0.0;
                    continue;
                }
                s->U[id] = wind_U(s->vmethod, s->vcoef, g->X[i], g->Y[j], g->Z[k] - Z0);
                s->V[id] = wind_V(s->vmethod, s->vcoef, g->X[i], g->Y[j], g->Z[k] - Z0);
                s->W[id] = wind_W(s->vmethod, s->vcoef, g->X[i], g->Y[j], g->Z[k] - Z0);
            }
}

/* ------------------------------------------------------------------------------------ */
/* PML slabs                                                                             */
/* ------------------------------------------------------------------------------------ */
#define QX(sl, f, ii, j, k) ((sl)->q[f][(size_t)(ii) + (size_t)(sl)->n * ((size_t)(j) + (size_t)s->air.ny * (size_t)(k))])
#define QY(sl, f, i, jj, k) ((sl)->q[f][(size_t)(i) + (size_t)s->air.nx * ((size_t)(jj) + (size_t)(sl)->n * (size_t)(k))])
#define QZ(sl, f, i, j, kk) ((sl)->q[f][(size_t)(i) + (size_t)s->air.nx * ((size_t)(j) + (size_t)s->air.ny * (size_t)(kk))])

static int slab_alloc(oslab *sl, int n, size_t plane)
{
    sl->n = n;
    for (int f = 0; f < 4; ++f) {
        sl->q[f] = n ? dalloc((size_t)n * plane, 0.0) : NULL;
        if (n && !sl->q[f]) return -1;
    }
    return 0;
}
static void slab_free(oslab *sl) { for (int f = 0; f < 4; ++f) { free(sl->q[f]); sl->q[f] = NULL; } }

/* PML judge of one axis from the absorbing flags (airpore.cpp:74-105, 182-247). */
static int pml_judge(int hi_absorbing, int lo_absorbing)
{
    int num = (hi_absorbing ? 1 : 0) + (lo_absorbing ? 10 : 0);
    switch (num) { case 1: return 2; case 10: return 1; case 11: return 3; default: return 0; }
}

/* air::cal_PML_var (air.cpp:509-937) for one rank covering the whole domain. */
static const char *pml_setup(orc_sim *s)
{
    const lee_deck *d = &s->d;
    ogrid *g = &s->air;
    int nx = g->nx, ny = g->ny, nz = g->nz;
    int *h = (int *)d->air_hdr;
    s->I1 = h[1]; s->I2 = h[2]; s->J1 = h[4]; s->J2 = h[5]; s->J3 = h[4]; s->K1 = h[7]; s->K2 = h[8];
    s->judge[0] = pml_judge(d->bc[LEE_FACE_F] == LEE_BC_ABSORBING, d->bc[LEE_FACE_B] == LEE_BC_ABSORBING);
    s->judge[1] = pml_judge(d->bc[LEE_FACE_E] == LEE_BC_ABSORBING, d->bc[LEE_FACE_W] == LEE_BC_ABSORBING);
    s->judge[2] = pml_judge(d->bc[LEE_FACE_N] == LEE_BC_ABSORBING, d->bc[LEE_FACE_S] == LEE_BC_ABSORBING);
    int nI1 = 0, nI2 = 0, nJ1 = 0, nJ2 = 0, nK1 = 0, nK2 = 0;
    /* x */
    if (s->judge[0]) {
        s->Rx = d->pml_R[0]; s->ax_ = d->pml_alpha[0];
        int jx = s->judge[0];
        if (jx == 1 || jx == 3) {
            if (s->I1 < 2 || s->I1 > nx) return "x1 PML width out of range";
            s->W1 = 1.0 / fabs(g->ax[s->I1 - 1] - g->ax[0]); s->px1 = g->ax[s->I1 - 1];
            nI1 = (nx - 1 >= s->I1) ? s->I1 + 1 : nx;
        }
        if (jx == 2 || jx == 3) {
            if (s->I2 < 2 || s->I2 > nx) return "x2 PML width out of range";
            s->W2 = 1.0 / fabs(g->ax[nx - 1] - g->ax[nx - s->I2]); s->px2 = g->ax[nx - s->I2];
            nI2 = (0 <= nx - s->I2 - 1) ? s->I2 + 1 : nx;
        }
        if (jx == 1) s->I2 = 0;
        if (jx == 2) s->I1 = 0;
    } else { s->I1 = 0; s->I2 = 0; }
    /* y */
    if (s->judge[1]) {
        s->Ry = d->pml_R[1]; s->ay_ = d->pml_alpha[1];
        int jy = s->judge[1];
        if (jy == 1 || jy == 3) {
            if (s->J1 < 2 || s->J1 > ny) return "y1 PML width out of range";
            s->L1 = 1.0 / fabs(g->ay[s->J1 - 1] - g->ay[0]); s->py1 = g->ay[s->J1 - 1];
            s->pml_time = d->pml_y_time; s->pml_counter = 0;
            nJ1 = (ny - 1 >= s->J1) ? s->J1 + 1 : ny;
        }
        if (jy == 2 || jy == 3) {
            if (s->J2 < 2 || s->J2 > ny) return "y2 PML width out of range";
            s->L2 = 1.0 / fabs(g->ay[ny - 1] - g->ay[ny - s->J2]); s->py2 = g->ay[ny - s->J2];
            nJ2 = (0 <= ny - s->J2 - 1) ? s->J2 + 1 : ny;
        }
        if (jy == 1) s->J2 = 0;
        if (jy == 2) s->J1 = 0;
    } else { s->J1 = 0; s->J2 = 0; }
    if (!(s->judge[1] == 1 || s->judge[1] == 3)) s->J3 = 0;
    /* z */
    if (s->judge[2]) {
        s->Rz = d->pml_R[2]; s->az_ = d->pml_alpha[2];
        int jz = s->judge[2];
        if (jz == 1 || jz == 3) {
            if (s->K1 < 2 || s->K1 > nz) return "z1 PML width out of range";
            s->H1 = 1.0 / fabs(g->az[s->K1 - 1] - g->az[0]); s->pz1 = g->az[s->K1 - 1];
            nK1 = (nz - 1 >= s->K1) ? s->K1 + 1 : nz;
        }
        if (jz == 2 || jz == 3) {
            if (s->K2 < 2 || s->K2 > nz) return "z2 PML width out of range";
            s->H2 = 1.0 / fabs(g->az[nz - 1] - g->az[nz - s->K2]); s->pz2 = g->az[nz - s->K2];
            nK2 = (0 <= nz - s->K2 - 1) ? s->K2 + 1 : nz;
        }
        if (jz == 1) s->K2 = 0;
        if (jz == 2) s->K1 = 0;
    } else { s->K1 = 0; s->K2 = 0; }
    size_t px = (size_t)ny * nz, py = (size_t)nx * nz, pz = (size_t)nx * ny;
    if (slab_alloc(&s->sx1, nI1, px) || slab_alloc(&s->sx2, nI2, px) ||
        slab_alloc(&s->sy1, nJ1, py) || slab_alloc(&s->sy2, nJ2, py) ||
        slab_alloc(&s->sz1, nK1, pz) || slab_alloc(&s->sz2, nK2, pz))
        return "out of memory";
    return NULL;
}

/* y1 time gate at the top of air::cal_fp (air.cpp:1965-1969). */
void orc_tick_gate(orc_sim *s)
{
    if (s->judge[1] == 1 || s->judge[1] == 3) {
        s->pml_counter = s->pml_counter + 1;
        if (s->pml_counter >= s->pml_time) s->J1 = s->J3; else s->J1 = 0;
    }
}

/* ------------------------------------------------------------------------------------ */
/* air::cal_fp (air.cpp:1953-2322) -- pressure increment at interior points              */
/* ------------------------------------------------------------------------------------ */
typedef struct { double c1, c2, c3, UX, UY, UZ, VX, VY, VZ, WX, WY, WZ; } windterms;

/* Contravariant wind coefficients and wind gradients (air.cpp:1993-2025 / 1515-1547). */
static windterms wind_terms(const orc_sim *s, const metric9 *m, int i, int j, int k)
{
    const ogrid *g = &s->air;
    windterms t;
    size_t id = IDX(g, i, j, k);
    size_t xp = IDX(g, i + 1, j, k), xm = IDX(g, i - 1, j, k);
    size_t yp = IDX(g, i, j + 1, k), ym = IDX(g, i, j - 1, k);
    size_t zp = IDX(g, i, j, k + 1), zm = IDX(g, i, j, k - 1);
    const double *U = s->U, *V = s->V, *W = s->W;
    double dxi = g->dxi, dyi = g->dyi, dzi = g->dzi, dt = g->dt;
    t.c1 = (U[id] * m->xX + V[id] * m->xY + W[id] * m->xZ) * dxi * dt;
    t.c2 = (U[id] * m->yX + V[id] * m->yY + W[id] * m->yZ) * dyi * dt;
    t.c3 = (U[id] * m->zX + V[id] * m->zY + W[id] * m->zZ) * dzi * dt;
#define GRAD(F, a, b, c) (m->a * (F[xp] - F[xm]) * dxi * 0.5 + m->b * (F[yp] - F[ym]) * dyi * 0.5 + m->c * (F[zp] - F[zm]) * dzi * 0.5)
    t.UX = GRAD(U, xX, yX, zX); t.UY = GRAD(U, xY, yY, zY); t.UZ = GRAD(U, xZ, yZ, zZ);
    t.VX = GRAD(V, xX, yX, zX); t.VY = GRAD(V, xY, yY, zY); t.VZ = GRAD(V, xZ, yZ, zZ);
    t.WX = GRAD(W, xX, yX, zX); t.WY = GRAD(W, xY, yY, zY); t.WZ = GRAD(W, xZ, yZ, zZ);
#undef GRAD
    return t;
}

/* Forward difference of a full-grid field at (i,j,k) along x, y, z. */
#define FDX(F) ((F)[IDX(g, i + 1, j, k)] - (F)[IDX(g, i, j, k)])
#define FDY(F) ((F)[IDX(g, i, j + 1, k)] - (F)[IDX(g, i, j, k)])
#define FDZ(F) ((F)[IDX(g, i, j, k + 1)] - (F)[IDX(g, i, j, k)])
#define BDX(F) ((F)[IDX(g, i, j, k)] - (F)[IDX(g, i - 1, j, k)])
#define BDY(F) ((F)[IDX(g, i, j, k)] - (F)[IDX(g, i, j - 1, k)])
#define BDZ(F) ((F)[IDX(g, i, j, k)] - (F)[IDX(g, i, j, k - 1)])
/* central differences of the convective terms on the RK4 path (extension): the reference's
 * one-sided pair advects the velocities across the flow downwind, which grows at up to
 * 2|V|/dy; the central form is neutral */
#define CDX(F) (((F)[IDX(g, i + 1, j, k)] - (F)[IDX(g, i - 1, j, k)]) * 0.5)
#define CDY(F) (((F)[IDX(g, i, j + 1, k)] - (F)[IDX(g, i, j - 1, k)]) * 0.5)
#define CDZ(F) (((F)[IDX(g, i, j, k + 1)] - (F)[IDX(g, i, j, k - 1)]) * 0.5)

static double sigma_of(double R, double alpha, double t) { return R * pow(t, alpha); }
static double beta_of(double alpha, double t) { return 1.0 + 25.0 * pow(t, alpha); }

/* material id of the obstacle at frame-local air node (i, j, k): the reference tests the
 * rank-global index against the box fixed at construction (air.cpp:1907-1909, 2294-2296),
 * so the obstacle keeps its index position when the frame moves. */
static int cr_at(const orc_sim *s, int in, int jn, int kn)
{
    if (s->cr_judge == 0) return 0;
    if (in < s->ob[0] || in > s->ob[1] || jn < s->ob[2] || jn > s->ob[3] || kn < s->ob[4] || kn > s->ob[5]) return 0;
    int ni = s->ob[1] - s->ob[0] + 1, nj = s->ob[3] - s->ob[2] + 1;
    return s->cr_uc[(size_t)(in - s->ob[0]) + (size_t)ni * ((size_t)(jn - s->ob[2]) + (size_t)nj * (kn - s->ob[4]))];
}

static double air_dp_point(orc_sim *s, const double *u, const double *v, const double *w,
                           const double *p, int i, int j, int k)
{
    const ogrid *g = &s->air;
    int nx = g->nx, ny = g->ny, nz = g->nz;
    double dxi = g->dxi, dyi = g->dyi, dzi = g->dzi, dt = g->dt;
    double gam = s->gamma, Pav = s->Pav;
    int in = i, jn = j, kn = k; /* one rank: local == global */
    metric9 m = point_metric(g, i, j, k);
    windterms t = wind_terms(s, &m, i, j, k);
    size_t id = IDX(g, i, j, k);
    double pc = p[id];
    double r;
    const int cen = s->opt.integrator == LEE_INT_RK4;
    if (cen) r = -t.c1 * CDX(p) - t.c2 * CDY(p) - t.c3 * CDZ(p);
    else r = -t.c1 * BDX(p) - t.c2 * BDY(p) - t.c3 * BDZ(p);
    r += -gam * pc * dt * (t.UX + t.VY + t.WZ);
    r += -gam * Pav * dt * (m.xX * FDX(u) * dxi + m.yX * FDY(u) * dyi + m.zX * FDZ(u) * dzi
                          + m.xY * FDX(v) * dxi + m.yY * FDY(v) * dyi + m.zY * FDZ(v) * dzi
                          + m.xZ * FDX(w) * dxi + m.yZ * FDY(w) * dyi + m.zZ * FDZ(w) * dzi);

    /* slab-field forward differences at slab-local indices */
#define XS_F(sl, f, ii, a, b, c) a * (QX(sl, f, (ii) + 1, j, k) - QX(sl, f, ii, j, k)) * dxi + \
                                  b * (QX(sl, f, ii, j + 1, k) - QX(sl, f, ii, j, k)) * dyi + \
                                  c * (QX(sl, f, ii, j, k + 1) - QX(sl, f, ii, j, k)) * dzi
#define YS_F(sl, f, jj, a, b, c) a * (QY(sl, f, i + 1, jj, k) - QY(sl, f, i, jj, k)) * dxi + \
                                  b * (QY(sl, f, i, (jj) + 1, k) - QY(sl, f, i, jj, k)) * dyi + \
                                  c * (QY(sl, f, i, jj, k + 1) - QY(sl, f, i, jj, k)) * dzi
#define ZS_F(sl, f, kk, a, b, c) a * (QZ(sl, f, i + 1, j, kk) - QZ(sl, f, i, j, kk)) * dxi + \
                                  b * (QZ(sl, f, i, j + 1, kk) - QZ(sl, f, i, j, kk)) * dyi + \
                                  c * (QZ(sl, f, i, j, (kk) + 1) - QZ(sl, f, i, j, kk)) * dzi
    double sx1 = 0, bx1 = 1, sx2 = 0, bx2 = 1, sy1 = 0, by1 = 1, sy2 = 0, by2 = 1;
    double sz1, bz1, sz2, bz2;
    int ii = 0, jj = 0, kk = 0;
    oslab *X1 = &s->sx1, *X2 = &s->sx2, *Y1 = &s->sy1, *Y2 = &s->sy2, *Z1 = &s->sz1, *Z2 = &s->sz2;
    /* PML X1 (air.cpp:2044-2057) */
    if (in < s->I1) {
        double t1 = (s->px1 - g->X[i]) * s->W1;
        sx1 = sigma_of(s->Rx, s->ax_, t1); bx1 = beta_of(s->ax_, t1);
        r += (1.0 - 1.0 / bx1) * (gam * Pav * (m.xX * FDX(u) * dxi + m.yX * FDY(u) * dyi + m.zX * FDZ(u) * dzi)) * dt
           - (sx1 * (pc + gam * Pav * (XS_F(X1, 1, i, m.xY, m.yY, m.zY) + XS_F(X1, 2, i, m.xZ, m.yZ, m.zZ)))) * dt;
    }
    /* PML X2 (air.cpp:2059-2074) */
    if (in >= nx - s->I2) {
        ii = i - (nx - X2->n);
        double t2 = (g->X[i] - s->px2) * s->W2;
        sx2 = sigma_of(s->Rx, s->ax_, t2); bx2 = beta_of(s->ax_, t2);
        r += (1.0 - 1.0 / bx2) * (gam * Pav * (m.xX * FDX(u) * dxi + m.yX * FDY(u) * dyi + m.zX * FDZ(u) * dzi)) * dt
           - (sx2 * (pc + gam * Pav * (XS_F(X2, 1, ii, m.xY, m.yY, m.zY) + XS_F(X2, 2, ii, m.xZ, m.yZ, m.zZ)))) * dt;
    }
    /* PML Y1 and Y1 x X (air.cpp:2076-2116) */
    if (jn < s->J1) {
        double t1 = (s->py1 - g->Y[j]) * s->L1;
        sy1 = sigma_of(s->Ry, s->ay_, t1); by1 = beta_of(s->ay_, t1);
        r += (1.0 - 1.0 / by1) * (gam * Pav * (m.xY * FDX(v) * dxi + m.yY * FDY(v) * dyi + m.zY * FDZ(v) * dzi)) * dt
           - sy1 * (pc + gam * Pav * (YS_F(Y1, 0, j, m.xX, m.yX, m.zX) + YS_F(Y1, 2, j, m.xZ, m.yZ, m.zZ))) * dt;
        if (in < s->I1) {
            double tx = (s->px1 - g->X[i]) * s->W1;
            sx1 = sigma_of(s->Rx, s->ax_, tx); bx1 = beta_of(s->ax_, tx);
            r += -(QX(X1, 3, i, j, k) * sx1 * sy1 + gam * Pav * (
                     (1 / bx1 - 1) * sy1 * (XS_F(X1, 0, i, m.xX, m.yX, m.zX))
                   + (1 / by1 - 1) * sx1 * (XS_F(X1, 1, i, m.xY, m.yY, m.zY)))) * dt;
        } else if (in >= nx - s->I2) {
            ii = i - (nx - X2->n);
            double tx = (g->X[i] - s->px2) * s->W2;
            sx2 = sigma_of(s->Rx, s->ax_, tx); bx2 = beta_of(s->ax_, tx);
            r += -(QX(X2, 3, ii, j, k) * sx2 * sy1 + gam * Pav * (
                     (1 / bx2 - 1) * sy1 * (XS_F(X2, 0, ii, m.xX, m.yX, m.zX))
                   + (1 / by1 - 1) * sx2 * (XS_F(X2, 1, ii, m.xY, m.yY, m.zY)))) * dt;
        }
    }
    /* PML Y2 and Y2 x X (air.cpp:2118-2158) */
    if (jn >= ny - s->J2) {
        jj = j - (ny - Y2->n);
        double t2 = (g->Y[j] - s->py2) * s->L2;
        sy2 = sigma_of(s->Ry, s->ay_, t2); by2 = beta_of(s->ay_, t2);
        r += (1 - 1 / by2) * (gam * Pav * (m.xY * FDX(v) * dxi + m.yY * FDY(v) * dyi + m.zY * FDZ(v) * dzi)) * dt
           - sy2 * (pc + gam * Pav * (YS_F(Y2, 0, jj, m.xX, m.yX, m.zX) + YS_F(Y2, 2, jj, m.xZ, m.yZ, m.zZ))) * dt;
        if (in < s->I1) {
            double tx = (s->px1 - g->X[i]) * s->W1;
            sx1 = sigma_of(s->Rx, s->ax_, tx); bx1 = beta_of(s->ax_, tx);
            r += -(QX(X1, 3, i, j, k) * sx1 * sy2 + gam * Pav * (
                     (1 / bx1 - 1) * sy2 * (XS_F(X1, 0, i, m.xX, m.yX, m.zX))
                   + (1 / by2 - 1) * sx1 * (XS_F(X1, 1, i, m.xY, m.yY, m.zY)))) * dt;
        } else if (in >= nx - s->I2) {
            ii = i - (nx - X2->n);
            double tx = (g->X[i] - s->px2) * s->W2;
            sx2 = sigma_of(s->Rx, s->ax_, tx); bx2 = beta_of(s->ax_, tx);
            r += -(QX(X2, 3, ii, j, k) * sx2 * sy2 + gam * Pav * (
                     (1 / bx2 - 1) * sy2 * (XS_F(X2, 0, ii, m.xX, m.yX, m.zX))
                   + (1 / by2 - 1) * sx2 * (XS_F(X2, 1, ii, m.xY, m.yY, m.zY)))) * dt;
        }
    }
    /* PML Z1 and cross terms (air.cpp:2160-2225) */
    if (kn < s->K1) {
        double t1 = (s->pz1 - g->Z[k]) * s->H1;
        sz1 = sigma_of(s->Rz, s->az_, t1); bz1 = beta_of(s->az_, t1);
        r += (1.0 - 1.0 / bz1) * gam * Pav * (m.xZ * FDX(w) * dxi + m.yZ * FDY(w) * dyi + m.zZ * FDZ(w) * dzi) * dt
           - sz1 * (pc + gam * Pav * (ZS_F(Z1, 0, k, m.xX, m.yX, m.zX) + ZS_F(Z1, 1, k, m.xY, m.yY, m.zY))) * dt;
        if (in < s->I1) {
            double tx = (s->px1 - g->X[i]) * s->W1;
            sx1 = sigma_of(s->Rx, s->ax_, tx); bx1 = beta_of(s->ax_, tx);
            r += -(QX(X1, 3, i, j, k) * sx1 * sz1 + gam * Pav * (
                     (1 / bx1 - 1) * sz1 * (XS_F(X1, 0, i, m.xX, m.yX, m.zX))
                   + (1 / bz1 - 1) * sx1 * (XS_F(X1, 2, i, m.xZ, m.yZ, m.zZ)))) * dt;
        } else if (in >= nx - s->I2) {
            ii = i - (nx - X2->n);
            double tx = (g->X[i] - s->px2) * s->W2;
            sx2 = sigma_of(s->Rx, s->ax_, tx); bx2 = beta_of(s->ax_, tx);
            r += -(QX(X2, 3, ii, j, k) * sx2 * sz1 + gam * Pav * (
                     (1 / bx2 - 1) * sz1 * (XS_F(X2, 0, ii, m.xX, m.yX, m.zX))
                   + (1 / bz1 - 1) * sx2 * (XS_F(X2, 2, ii, m.xZ, m.yZ, m.zZ)))) * dt;
        }
        if (jn < s->J1) {
            double ty = (s->py1 - g->Y[j]) * s->L1;
            sy1 = sigma_of(s->Ry, s->ay_, ty); by1 = beta_of(s->ay_, ty);
            r += -(QY(Y1, 3, i, j, k) * sy1 * sz1 + gam * Pav * (
                     (1 / by1 - 1) * sz1 * (YS_F(Y1, 1, j, m.xY, m.yY, m.zY))
                   + (1 / bz1 - 1) * sy1 * (YS_F(Y1, 2, j, m.xZ, m.yZ, m.zZ)))) * dt;
        } else if (jn >= ny - s->J2) {
            jj = j - (ny - Y2->n);
            double ty = (g->Y[j] - s->py2) * s->L2;
            sy2 = sigma_of(s->Ry, s->ay_, ty); by2 = beta_of(s->ay_, ty);
            r += -(QY(Y2, 3, i, jj, k) * sy2 * sz1 + gam * Pav * (
                     (1 / by2 - 1) * sz1 * (YS_F(Y2, 1, jj, m.xY, m.yY, m.zY))
                   + (1 / bz1 - 1) * sy2 * (YS_F(Y2, 2, jj, m.xZ, m.yZ, m.zZ)))) * dt;
        }
    }
    /* PML Z2 and cross terms (air.cpp:2227-2292) */
    if (kn >= nz - s->K2) {
        kk = k - (nz - Z2->n);
        double t2 = (g->Z[k] - s->pz2) * s->H2;
        sz2 = sigma_of(s->Rz, s->az_, t2); bz2 = beta_of(s->az_, t2);
        r += (1.0 - 1.0 / bz2) * gam * Pav * (m.xZ * FDX(w) * dxi + m.yZ * FDY(w) * dyi + m.zZ * FDZ(w) * dzi) * dt
           - sz2 * (pc + gam * Pav * (ZS_F(Z2, 0, kk, m.xX, m.yX, m.zX) + ZS_F(Z2, 1, kk, m.xY, m.yY, m.zY))) * dt;
        if (in < s->I1) {
            double tx = (s->px1 - g->X[i]) * s->W1;
            sx1 = sigma_of(s->Rx, s->ax_, tx); bx1 = beta_of(s->ax_, tx);
            r += -(QX(X1, 3, i, j, k) * sx1 * sz2 + gam * Pav * (
                     (1 / bx1 - 1) * sz2 * (XS_F(X1, 0, i, m.xX, m.yX, m.zX))
                   + (1 / bz2 - 1) * sx1 * (XS_F(X1, 2, i, m.xZ, m.yZ, m.zZ)))) * dt;
        } else if (in >= nx - s->I2) {
            ii = i - (nx - X2->n);
            double tx = (g->X[i] - s->px2) * s->W2;
            sx2 = sigma_of(s->Rx, s->ax_, tx); bx2 = beta_of(s->ax_, tx);
            r += -(QX(X2, 3, ii, j, k) * sx2 * sz2 + gam * Pav * (
                     (1 / bx2 - 1) * sz2 * (XS_F(X2, 0, ii, m.xX, m.yX, m.zX))
                   + (1 / bz2 - 1) * sx2 * (XS_F(X2, 2, ii, m.xZ, m.yZ, m.zZ)))) * dt;
        }
        if (jn < s->J1) {
            double ty = (s->py1 - g->Y[j]) * s->L1;
            sy1 = sigma_of(s->Ry, s->ay_, ty); by1 = beta_of(s->ay_, ty);
            r += -(QY(Y1, 3, i, j, k) * sy1 * sz2 + gam * Pav * (
                     (1 / by1 - 1) * sz2 * (YS_F(Y1, 1, j, m.xY, m.yY, m.zY))
                   + (1 / bz2 - 1) * sy1 * (YS_F(Y1, 2, j, m.xZ, m.yZ, m.zZ)))) * dt;
        } else if (jn >= ny - s->J2) {
            jj = j - (ny - Y2->n);
            double ty = (g->Y[j] - s->py2) * s->L2;
            sy2 = sigma_of(s->Ry, s->ay_, ty); by2 = beta_of(s->ay_, ty);
            r += -(QY(Y2, 3, i, jj, k) * sy2 * sz2 + gam * Pav * (
                     (1 / by2 - 1) * sz2 * (YS_F(Y2, 1, jj, m.xY, m.yY, m.zY))
                   + (1 / bz2 - 1) * sy2 * (YS_F(Y2, 2, jj, m.xZ, m.yZ, m.zZ)))) * dt;
        }
    }
#undef XS_F
#undef YS_F
#undef ZS_F
    /* inside the obstacle the increment is the rigid-frame porous continuity equation with
     * the full nine-term divergence (air.cpp:2294-2318); mask 1 takes the canopy, mask 2 the
     * trunk coefficients */
    int mo = cr_at(s, in, jn, kn);
    if (mo == 1 || mo == 2)
        r = -s->cr_gama[mo] * (m.xX * FDX(u) * dxi + m.yX * FDY(u) * dyi + m.zX * FDZ(u) * dzi
                             + m.xY * FDX(v) * dxi + m.yY * FDY(v) * dyi + m.zY * FDZ(v) * dzi
                             + m.xZ * FDX(w) * dxi + m.yZ * FDY(w) * dyi + m.zZ * FDZ(w) * dzi);
    return r;
}

/* ------------------------------------------------------------------------------------ */
/* air::cal_fuvw (air.cpp:1476-1949) -- velocity increments at interior points            */
/* ------------------------------------------------------------------------------------ */
static void air_duvw_point(orc_sim *s, const double *u, const double *v, const double *w,
                           const double *p, int i, int j, int k, double *ou, double *ov, double *ow)
{
    const ogrid *g = &s->air;
    int nx = g->nx, ny = g->ny, nz = g->nz;
    double dxi = g->dxi, dyi = g->dyi, dzi = g->dzi, dt = g->dt, Aav = s->Aav;
    int in = i, jn = j, kn = k;
    metric9 m = point_metric(g, i, j, k);
    windterms t = wind_terms(s, &m, i, j, k);
    size_t id = IDX(g, i, j, k);
    double uc = u[id], vc = v[id], wc = w[id];
    double du, dv, dw;
    const int cen = s->opt.integrator == LEE_INT_RK4;
    if (cen) du = -t.c1 * CDX(u) - t.c2 * CDY(u) - t.c3 * CDZ(u);
    else du = -t.c1 * FDX(u) - t.c2 * FDY(u) - t.c3 * FDZ(u);
    du += -uc * t.UX * dt - vc * t.UY * dt - wc * t.UZ * dt;
    du += -Aav * dt * (m.xX * BDX(p) * dxi + m.yX * BDY(p) * dyi + m.zX * BDZ(p) * dzi);
    if (cen) dv = -t.c1 * CDX(v) - t.c2 * CDY(v) - t.c3 * CDZ(v);
    else dv = -t.c1 * FDX(v) - t.c2 * FDY(v) - t.c3 * FDZ(v);
    dv += -uc * t.VX * dt - vc * t.VY * dt - wc * t.VZ * dt;
    dv += -Aav * dt * (m.xY * BDX(p) * dxi + m.yY * BDY(p) * dyi + m.zY * BDZ(p) * dzi);
    if (cen) dw = -t.c1 * CDX(w) - t.c2 * CDY(w) - t.c3 * CDZ(w);
    else dw = -t.c1 * FDX(w) - t.c2 * FDY(w) - t.c3 * FDZ(w);
    dw += -uc * t.WX * dt - vc * t.WY * dt - wc * t.WZ * dt;
    dw += -Aav * dt * (m.xZ * BDX(p) * dxi + m.yZ * BDY(p) * dyi + m.zZ * BDZ(p) * dzi);

#define XS_B(sl, f, ii, a, b, c) (a * (QX(sl, f, ii, j, k) - QX(sl, f, (ii) - 1, j, k)) * dxi + \
                                  b * (QX(sl, f, ii, j, k) - QX(sl, f, ii, j - 1, k)) * dyi + \
                                  c * (QX(sl, f, ii, j, k) - QX(sl, f, ii, j, k - 1)) * dzi)
#define YS_B(sl, f, jj, a, b, c) (a * (QY(sl, f, i, jj, k) - QY(sl, f, i - 1, jj, k)) * dxi + \
                                  b * (QY(sl, f, i, jj, k) - QY(sl, f, i, (jj) - 1, k)) * dyi + \
                                  c * (QY(sl, f, i, jj, k) - QY(sl, f, i, jj, k - 1)) * dzi)
#define ZS_B(sl, f, kk, a, b, c) (a * (QZ(sl, f, i, j, kk) - QZ(sl, f, i - 1, j, kk)) * dxi + \
                                  b * (QZ(sl, f, i, j, kk) - QZ(sl, f, i, j - 1, kk)) * dyi + \
                                  c * (QZ(sl, f, i, j, kk) - QZ(sl, f, i, j, (kk) - 1)) * dzi)
#define GP(a, b, c) (a * BDX(p) * dxi + b * BDY(p) * dyi + c * BDZ(p) * dzi)
    double sx1, bx1, sx2, bx2, sy1, by1, sy2, by2, sz1, bz1, sz2, bz2;
    int ii, jj, kk;
    oslab *X1 = &s->sx1, *X2 = &s->sx2, *Y1 = &s->sy1, *Y2 = &s->sy2, *Z1 = &s->sz1, *Z2 = &s->sz2;
    /* PML X1 (air.cpp:1579-1598) */
    if (in < s->I1) {
        double tx = (s->px1 - g->X[i]) * s->W1;
        sx1 = sigma_of(s->Rx, s->ax_, tx); bx1 = beta_of(s->ax_, tx);
        du += -sx1 * (uc) * dt + (1.0 - 1.0 / bx1) * (Aav * GP(m.xX, m.yX, m.zX)) * dt;
        dv += -sx1 * (vc + Aav * XS_B(X1, 3, i, m.xY, m.yY, m.zY)) * dt;
        dw += -sx1 * (wc + Aav * XS_B(X1, 3, i, m.xZ, m.yZ, m.zZ)) * dt;
    }
    /* PML X2 (air.cpp:1600-1619) */
    if (in >= nx - s->I2) {
        ii = i - (nx - X2->n);
        double tx = (g->X[i] - s->px2) * s->W2;
        sx2 = sigma_of(s->Rx, s->ax_, tx); bx2 = beta_of(s->ax_, tx);
        du += -sx2 * (uc) * dt + (1.0 - 1.0 / bx2) * (Aav * GP(m.xX, m.yX, m.zX)) * dt;
        dv += -sx2 * (vc + Aav * XS_B(X2, 3, ii, m.xY, m.yY, m.zY)) * dt;
        dw += -sx2 * (wc + Aav * XS_B(X2, 3, ii, m.xZ, m.yZ, m.zZ)) * dt;
    }
    /* PML Y1 and Y1 x X (air.cpp:1621-1673) */
    if (jn < s->J1) {
        double ty = (s->py1 - g->Y[j]) * s->L1;
        sy1 = sigma_of(s->Ry, s->ay_, ty); by1 = beta_of(s->ay_, ty);
        du += -sy1 * (uc + Aav * YS_B(Y1, 3, j, m.xX, m.yX, m.zX)) * dt;
        dv += (1.0 - 1.0 / by1) * Aav * GP(m.xY, m.yY, m.zY) * dt - sy1 * vc * dt;
        dw += -sy1 * (wc + Aav * YS_B(Y1, 3, j, m.xZ, m.yZ, m.zZ)) * dt;
        if (in < s->I1) {
            double tx = (s->px1 - g->X[i]) * s->W1;
            sx1 = sigma_of(s->Rx, s->ax_, tx); bx1 = beta_of(s->ax_, tx);
            du += -(QX(X1, 0, i, j, k) * sx1 * sy1 + (1 / bx1 - 1) * sy1 * Aav * XS_B(X1, 3, i, m.xX, m.yX, m.zX)) * dt;
            dv += -(QX(X1, 1, i, j, k) * sx1 * sy1 + (1 / by1 - 1) * sx1 * Aav * XS_B(X1, 3, i, m.xY, m.yY, m.zY)) * dt;
            dw += -QX(X1, 2, i, j, k) * sx1 * sy1 * dt;
        } else if (in >= nx - s->I2) {
            ii = i - (nx - X2->n);
            double tx = (g->X[i] - s->px2) * s->W2;
            sx2 = sigma_of(s->Rx, s->ax_, tx); bx2 = beta_of(s->ax_, tx);
            du += -(QX(X2, 0, ii, j, k) * sx2 * sy1 + (1 / bx2 - 1) * sy1 * Aav * XS_B(X2, 3, ii, m.xX, m.yX, m.zX)) * dt;
            dv += -(QX(X2, 1, ii, j, k) * sx2 * sy1 + (1 / by1 - 1) * sx2 * Aav * XS_B(X2, 3, ii, m.xY, m.yY, m.zY)) * dt;
            dw += -QX(X2, 2, ii, j, k) * sx2 * sy1 * dt;
        }
    }
    /* PML Y2 and Y2 x X (air.cpp:1675-1728) */
    if (jn >= ny - s->J2) {
        jj = j - (ny - Y2->n);
        double ty = (g->Y[j] - s->py2) * s->L2;
        sy2 = sigma_of(s->Ry, s->ay_, ty); by2 = beta_of(s->ay_, ty);
        du += -sy2 * (uc + Aav * YS_B(Y2, 3, jj, m.xX, m.yX, m.zX)) * dt;
        dv += (1 - 1 / by2) * Aav * GP(m.xY, m.yY, m.zY) * dt - sy2 * vc * dt;
        dw += -sy2 * (wc + Aav * YS_B(Y2, 3, jj, m.xZ, m.yZ, m.zZ)) * dt;
        if (in < s->I1) {
            double tx = (s->px1 - g->X[i]) * s->W1;
            sx1 = sigma_of(s->Rx, s->ax_, tx); bx1 = beta_of(s->ax_, tx);
            du += -(QX(X1, 0, i, j, k) * sx1 * sy2 + (1.0 / bx1 - 1.0) * sy2 * Aav * XS_B(X1, 3, i, m.xX, m.yX, m.zX)) * dt;
            dv += -(QX(X1, 1, i, j, k) * sx1 * sy2 + (1.0 / by2 - 1.0) * sx1 * Aav * XS_B(X1, 3, i, m.xY, m.yY, m.zY)) * dt;
            dw += -QX(X1, 2, i, j, k) * sx1 * sy2 * dt;
        } else if (in >= nx - s->I2) {
            ii = i - (nx - X2->n);
            double tx = (g->X[i] - s->px2) * s->W2;
            sx2 = sigma_of(s->Rx, s->ax_, tx); bx2 = beta_of(s->ax_, tx);
            du += -(QX(X2, 0, ii, j, k) * sx2 * sy2 + (1.0 / bx2 - 1.0) * sy2 * Aav * XS_B(X2, 3, ii, m.xX, m.yX, m.zX)) * dt;
            dv += -(QX(X2, 1, ii, j, k) * sx2 * sy2 + (1.0 / by2 - 1.0) * sx2 * Aav * XS_B(X2, 3, ii, m.xY, m.yY, m.zY)) * dt;
            dw += -QX(X2, 2, ii, j, k) * sx2 * sy2 * dt;
        }
    }
    /* PML Z1 and cross terms (air.cpp:1730-1816) */
    if (kn < s->K1) {
        double tz = (s->pz1 - g->Z[k]) * s->H1;
        sz1 = sigma_of(s->Rz, s->az_, tz); bz1 = beta_of(s->az_, tz);
        du += -sz1 * (uc + Aav * ZS_B(Z1, 3, k, m.xX, m.yX, m.zX)) * dt;
        dv += -sz1 * (vc + Aav * ZS_B(Z1, 3, k, m.xY, m.yY, m.zY)) * dt;
        dw += (1.0 - 1.0 / bz1) * Aav * GP(m.xZ, m.yZ, m.zZ) * dt - sz1 * wc * dt;
        if (in < s->I1) {
            double tx = (s->px1 - g->X[i]) * s->W1;
            sx1 = sigma_of(s->Rx, s->ax_, tx); bx1 = beta_of(s->ax_, tx);
            du += -(QX(X1, 0, i, j, k) * sx1 * sz1 + (1.0 / bx1 - 1.0) * sz1 * Aav * XS_B(X1, 3, i, m.xX, m.yX, m.zX)) * dt;
            dv += -QX(X1, 1, i, j, k) * sx1 * sz1 * dt;
            dw += -(QX(X1, 2, i, j, k) * sx1 * sz1 + (1.0 / bz1 - 1.0) * sx1 * Aav * XS_B(X1, 3, i, m.xZ, m.yZ, m.zZ)) * dt;
        } else if (in >= nx - s->I2) {
            ii = i - (nx - X2->n);
            double tx = (g->X[i] - s->px2) * s->W2;
            sx2 = sigma_of(s->Rx, s->ax_, tx); bx2 = beta_of(s->ax_, tx);
            du += -(QX(X2, 0, ii, j, k) * sx2 * sz1 + (1.0 / bx2 - 1.0) * sz1 * Aav * XS_B(X2, 3, ii, m.xX, m.yX, m.zX)) * dt;
            dv += -QX(X2, 1, ii, j, k) * sx2 * sz1 * dt;
            dw += -(QX(X2, 2, ii, j, k) * sx2 * sz1 + (1.0 / bz1 - 1.0) * sx2 * Aav * XS_B(X2, 3, ii, m.xZ, m.yZ, m.zZ)) * dt;
        }
        if (jn < s->J1) {
            double ty = (s->py1 - g->Y[j]) * s->L1;
            sy1 = sigma_of(s->Ry, s->ay_, ty); by1 = beta_of(s->ay_, ty);
            du += -QY(Y1, 0, i, j, k) * sy1 * sz1 * dt;
            dv += -(QY(Y1, 1, i, j, k) * sy1 * sz1 + (1.0 / by1 - 1.0) * sz1 * Aav * YS_B(Y1, 3, j, m.xY, m.yY, m.zY)) * dt;
            dw += -(QY(Y1, 2, i, j, k) * sy1 * sz1 + (1.0 / bz1 - 1.0) * sy1 * Aav * YS_B(Y1, 3, j, m.xZ, m.yZ, m.zZ)) * dt;
        } else if (jn >= ny - s->J2) {
            jj = j - (ny - Y2->n);
            double ty = (g->Y[j] - s->py2) * s->L2;
            sy2 = sigma_of(s->Ry, s->ay_, ty); by2 = beta_of(s->ay_, ty);
            du += -QY(Y2, 0, i, jj, k) * sy2 * sz1 * dt;
            dv += -(QY(Y2, 1, i, jj, k) * sy2 * sz1 + (1.0 / by2 - 1.0) * sz1 * Aav * YS_B(Y2, 3, jj, m.xY, m.yY, m.zY)) * dt;
            dw += -(QY(Y2, 2, i, jj, k) * sy2 * sz1 + (1.0 / bz1 - 1.0) * sy2 * Aav * YS_B(Y2, 3, jj, m.xZ, m.yZ, m.zZ)) * dt;
        }
    }
    /* PML Z2 and cross terms (air.cpp:1818-1905) */
    if (kn >= nz - s->K2) {
        kk = k - (nz - Z2->n);
        double tz = (g->Z[k] - s->pz2) * s->H2;
        sz2 = sigma_of(s->Rz, s->az_, tz); bz2 = beta_of(s->az_, tz);
        du += -sz2 * (uc + Aav * ZS_B(Z2, 3, kk, m.xX, m.yX, m.zX)) * dt;
        dv += -sz2 * (vc + Aav * ZS_B(Z2, 3, kk, m.xY, m.yY, m.zY)) * dt;
        dw += (1.0 - 1.0 / bz2) * Aav * GP(m.xZ, m.yZ, m.zZ) * dt - sz2 * wc * dt;
        if (in < s->I1) {
            double tx = (s->px1 - g->X[i]) * s->W1;
            sx1 = sigma_of(s->Rx, s->ax_, tx); bx1 = beta_of(s->ax_, tx);
            du += -(QX(X1, 0, i, j, k) * sx1 * sz2 + (1.0 / bx1 - 1.0) * sz2 * Aav * XS_B(X1, 3, i, m.xX, m.yX, m.zX)) * dt;
            dv += -QX(X1, 1, i, j, k) * sx1 * sz2 * dt;
            dw += -(QX(X1, 2, i, j, k) * sx1 * sz2 + (1.0 / bz2 - 1.0) * sx1 * Aav * XS_B(X1, 3, i, m.xZ, m.yZ, m.zZ)) * dt;
        } else if (in >= nx - s->I2) {
            ii = i - (nx - X2->n);
            double tx = (g->X[i] - s->px2) * s->W2;
            sx2 = sigma_of(s->Rx, s->ax_, tx); bx2 = beta_of(s->ax_, tx);
            du += -(QX(X2, 0, ii, j, k) * sx2 * sz2 + (1.0 / bx2 - 1.0) * sz2 * Aav * XS_B(X2, 3, ii, m.xX, m.yX, m.zX)) * dt;
            dv += -QX(X2, 1, ii, j, k) * sx2 * sz2 * dt;
            dw += -(QX(X2, 2, ii, j, k) * sx2 * sz2 + (1.0 / bz2 - 1.0) * sx2 * Aav * XS_B(X2, 3, ii, m.xZ, m.yZ, m.zZ)) * dt;
        }
        if (jn < s->J1) {
            double ty = (s->py1 - g->Y[j]) * s->L1;
            sy1 = sigma_of(s->Ry, s->ay_, ty); by1 = beta_of(s->ay_, ty);
            du += -QY(Y1, 0, i, j, k) * sy1 * sz2 * dt;
            dv += -(QY(Y1, 1, i, j, k) * sy1 * sz2 + (1.0 / by1 - 1.0) * sz2 * Aav * YS_B(Y1, 3, j, m.xY, m.yY, m.zY)) * dt;
            dw += -(QY(Y1, 2, i, j, k) * sy1 * sz2 + (1.0 / bz2 - 1.0) * sy1 * Aav * YS_B(Y1, 3, j, m.xZ, m.yZ, m.zZ)) * dt;
        } else if (jn >= ny - s->J2) {
            jj = j - (ny - Y2->n);
            double ty = (g->Y[j] - s->py2) * s->L2;
            sy2 = sigma_of(s->Ry, s->ay_, ty); by2 = beta_of(s->ay_, ty);
            du += -QY(Y2, 0, i, jj, k) * sy2 * sz2 * dt;
            dv += -(QY(Y2, 1, i, jj, k) * sy2 * sz2 + (1.0 / by2 - 1.0) * sz2 * Aav * YS_B(Y2, 3, jj, m.xY, m.yY, m.zY)) * dt;
            dw += -(QY(Y2, 2, i, jj, k) * sy2 * sz2 + (1.0 / bz2 - 1.0) * sy2 * Aav * YS_B(Y2, 3, jj, m.xZ, m.yZ, m.zZ)) * dt;
        }
    }
#undef XS_B
#undef YS_B
#undef ZS_B
    /* inside the obstacle: semi-implicit Darcy drag of the material (air.cpp:1907-1945) */
    int mo = cr_at(s, in, jn, kn);
    if (mo == 1 || mo == 2) {
        double ib = s->cr_invb[mo], f = -1.0 * ib * dt * s->cr_eff[mo];
        du = f * GP(m.xX, m.yX, m.zX) + uc * ib - uc;
        dv = f * GP(m.xY, m.yY, m.zY) + vc * ib - vc;
        dw = f * GP(m.xZ, m.yZ, m.zZ) + wc * ib - wc;
    }
#undef GP
    *ou = du; *ov = dv; *ow = dw;
}

void orc_air_dp(orc_sim *s, const double *u, const double *v, const double *w,
                const double *p, double *dp)
{
    const ogrid *g = &s->air;
#pragma omp parallel for schedule(static)
    for (int k = 1; k < g->nz - 1; ++k)
        for (int j = 1; j < g->ny - 1; ++j)
            for (int i = 1; i < g->nx - 1; ++i)
                dp[IDX(g, i, j, k)] = air_dp_point(s, u, v, w, p, i, j, k);
}

void orc_air_duvw(orc_sim *s, const double *u, const double *v, const double *w,
                  const double *p, double *du, double *dv, double *dw)
{
    const ogrid *g = &s->air;
#pragma omp parallel for schedule(static)
    for (int k = 1; k < g->nz - 1; ++k)
        for (int j = 1; j < g->ny - 1; ++j)
            for (int i = 1; i < g->nx - 1; ++i) {
                size_t id = IDX(g, i, j, k);
                air_duvw_point(s, u, v, w, p, i, j, k, &du[id], &dv[id], &dw[id]);
            }
}

/* ------------------------------------------------------------------------------------ */
/* porous stencils (porous.cpp:77-143)                                                   */
/* ------------------------------------------------------------------------------------ */
static void pore_fp_on(orc_sim *s, const double *u, const double *v, const double *w, double *out)
{
    ogrid *g = &s->pore;
    double dxi = g->dxi, dyi = g->dyi, dzi = g->dzi;
#pragma omp parallel for schedule(static)
    for (int k = 1; k < g->nz - 1; ++k)
        for (int j = 1; j < g->ny - 1; ++j)
            for (int i = 1; i < g->nx - 1; ++i) {
                metric9 m = point_metric(g, i, j, k);
                out[IDX(g, i, j, k)] = -s->p_gama * (m.xX * dxi * FDX(u) + m.yX * dyi * FDY(u) + m.zX * dzi * FDZ(u) +
                                                       m.xY * dxi * FDX(v) + m.yY * dyi * FDY(v) + m.zY * dzi * FDZ(v) +
                                                       m.xZ * dxi * FDX(w) + m.yZ * dyi * FDY(w) + m.zZ * dzi * FDZ(w));
            }
}

static void pore_fuvw_on(orc_sim *s, const double *u, const double *v, const double *w,
                         const double *p, double *ou, double *ov, double *ow)
{
    ogrid *g = &s->pore;
    double dxi = g->dxi, dyi = g->dyi, dzi = g->dzi, dt = g->dt;
    double ib = s->p_invbeta, eff = s->p_eff;
#pragma omp parallel for schedule(static)
    for (int k = 1; k < g->nz - 1; ++k)
        for (int j = 1; j < g->ny - 1; ++j)
            for (int i = 1; i < g->nx - 1; ++i) {
                metric9 m = point_metric(g, i, j, k);
                size_t id = IDX(g, i, j, k);
                ou[id] = -1.0 * ib * dt * eff * (m.xX * BDX(p) * dxi + m.yX * BDY(p) * dyi + m.zX * BDZ(p) * dzi) + u[id] * ib - u[id];
                ov[id] = -1.0 * ib * dt * eff * (m.xY * BDX(p) * dxi + m.yY * BDY(p) * dyi + m.zY * BDZ(p) * dzi) + v[id] * ib - v[id];
                ow[id] = -1.0 * ib * dt * eff * (m.xZ * BDX(p) * dxi + m.yZ * BDY(p) * dyi + m.zZ * BDZ(p) * dzi) + w[id] * ib - w[id];
            }
}

static void pore_fp(orc_sim *s) { ogrid *g = &s->pore; pore_fp_on(s, g->un, g->vn, g->wn, g->pnn); }
static void pore_fuvw(orc_sim *s)
{
    ogrid *g = &s->pore;
    pore_fuvw_on(s, g->un, g->vn, g->wn, g->pnn, g->unn, g->vnn, g->wnn);
}
void orc_pore_dp(orc_sim *s, const double *u, const double *v, const double *w, double *dp)
{
    if (s->south_porous) pore_fp_on(s, u, v, w, dp);
}
void orc_pore_duvw(orc_sim *s, const double *u, const double *v, const double *w, const double *p,
                   double *du, double *dv, double *dw)
{
    if (s->south_porous) pore_fuvw_on(s, u, v, w, p, du, dv, dw);
}

/* FB add loops of calculation_2D::Cal_pressure / Cal_velocity (calculation_2D.cpp:598-608, 564-574). */
static void add_interior(ogrid *g, double *fnn, const double *fn)
{
#pragma omp parallel for schedule(static)
    for (int k = 1; k < g->nz - 1; ++k)
        for (int j = 1; j < g->ny - 1; ++j)
            for (int i = 1; i < g->nx - 1; ++i) {
                size_t id = IDX(g, i, j, k);
                fnn[id] = fn[id] + fnn[id];
            }
}

/* ------------------------------------------------------------------------------------ */
/* boundary conditions (calculation_2D.cpp:646-864)                                     */
/* ------------------------------------------------------------------------------------ */
enum { FN, FS, FE, FW, FF, FB };

static void bc_p(ogrid *g, int face)
{
    int nx = g->nx, ny = g->ny, nz = g->nz;
    double *p = g->pnn;
    switch (face) {
    case FW: for (int k = 0; k < nz; ++k) for (int i = 0; i < nx; ++i) p[IDX(g, i, 0, k)] = p[IDX(g, i, 1, k)]; break;
    case FE:
        for (int k = 0; k < nz; ++k)
            for (int i = 0; i < nx; ++i) {
                int jn = ny - 1;
                p[IDX(g, i, jn, k)] = 15.0 / 8.0 * p[IDX(g, i, jn - 1, k)] - 5.0 / 4.0 * p[IDX(g, i, jn - 2, k)] + 3.0 / 8.0 * p[IDX(g, i, jn - 3, k)];
            }
        break;
    case FS: for (int j = 0; j < ny; ++j) for (int i = 0; i < nx; ++i) p[IDX(g, i, j, 0)] = p[IDX(g, i, j, 1)]; break;
    case FN: for (int j = 0; j < ny; ++j) for (int i = 0; i < nx; ++i) p[IDX(g, i, j, nz - 1)] = p[IDX(g, i, j, nz - 2)]; break;
    case FF: for (int k = 0; k < nz; ++k) for (int j = 0; j < ny; ++j) p[IDX(g, 0, j, k)] = p[IDX(g, 1, j, k)]; break;
    case FB: for (int k = 0; k < nz; ++k) for (int j = 0; j < ny; ++j) p[IDX(g, nx - 1, j, k)] = p[IDX(g, nx - 2, j, k)]; break;
    }
}

static void bc_v(ogrid *g, int face)
{
    int nx = g->nx, ny = g->ny, nz = g->nz;
    double *u = g->unn, *v = g->vnn, *w = g->wnn;
    switch (face) {
    case FW:
        for (int k = 0; k < nz; ++k) for (int i = 0; i < nx; ++i) {
            u[IDX(g, i, 0, k)] = u[IDX(g, i, 1, k)]; v[IDX(g, i, 0, k)] = 0.0; w[IDX(g, i, 0, k)] = w[IDX(g, i, 1, k)];
        }
        break;
    case FE:
        for (int k = 0; k < nz; ++k) for (int i = 0; i < nx; ++i) {
            u[IDX(g, i, ny - 1, k)] = u[IDX(g, i, ny - 2, k)]; v[IDX(g, i, ny - 1, k)] = 0.0; w[IDX(g, i, ny - 1, k)] = w[IDX(g, i, ny - 2, k)];
        }
        break;
    case FS:
        for (int j = 0; j < ny; ++j) for (int i = 0; i < nx; ++i) {
            u[IDX(g, i, j, 0)] = u[IDX(g, i, j, 1)]; v[IDX(g, i, j, 0)] = v[IDX(g, i, j, 1)]; w[IDX(g, i, j, 0)] = 0.0;
        }
        break;
    case FN:
        for (int j = 0; j < ny; ++j) for (int i = 0; i < nx; ++i) {
            u[IDX(g, i, j, nz - 1)] = u[IDX(g, i, j, nz - 2)]; v[IDX(g, i, j, nz - 1)] = v[IDX(g, i, j, nz - 2)]; w[IDX(g, i, j, nz - 1)] = 0.0;
        }
        break;
    case FF:
        for (int k = 0; k < nz; ++k) for (int j = 0; j < ny; ++j) {
            u[IDX(g, 0, j, k)] = 0.0; v[IDX(g, 0, j, k)] = v[IDX(g, 1, j, k)]; w[IDX(g, 0, j, k)] = w[IDX(g, 1, j, k)];
        }
        break;
    case FB:
        for (int k = 0; k < nz; ++k) for (int j = 0; j < ny; ++j) {
            u[IDX(g, nx - 1, j, k)] = 0.0; v[IDX(g, nx - 1, j, k)] = v[IDX(g, nx - 2, j, k)]; w[IDX(g, nx - 1, j, k)] = w[IDX(g, nx - 2, j, k)];
        }
        break;
    }
}

/* East TDBC: calculation_2D::UpdateBC_velocity(EastBC,Z,tau), Cal_TDBC, Update_Psi
 * (calculation_2D.cpp:986-1028). */
/* Ground impedance (extension).  The plane k = 0 is the ground's interface plane, as the
 * porous ground's top plane is when the air reads it through the interface copy
 * (calculation_2D.cpp:878-886): its pressure and tangential velocities follow the porous
 * equations of the ground's top material (porous.cpp:77-143 with porous.cpp:44-50), and its
 * normal velocity w (half a cell below the plane) follows the porous momentum equation with
 * the pressure half a cell further down given by the impedance, Z(s) = Zinf (1 + sum_k
 * a_k / (r_k + s tau)), through the TDBC recursion of calculation_2D.cpp:986-1028 with K
 * poles.  The interior nodes of the plane are updated like stencil output, before the faces
 * (N E W F B, there is no South face), which then treat the plane's edge nodes as on every
 * other plane.  The plane's metric is the porous top plane's, Z_over_z = Zz of the plane k = 1:
 * ex = 1/X_over_x, ey = 1/Y_over_y, ez = 1/Z_over_z and, on the terrain-following grid, the
 * cross terms zX = -(hx ex) ez, zY = -(hy ey) ez of the inverse Jacobian of a layer shifted
 * rigidly with the ground (Z_over_x = hx, Z_over_y = hy; porous.cpp's grid under
 * airpore.cpp's hills, sh = tk = 1 there), not the air column's compressed one: the closure
 * stands for that plane.  The z derivatives of the plane are one-sided towards k = 1 (the
 * velocities' pressure gradient d/dX = ex d/dx + zX d/dz takes (p1 - p0) dz^-1).  With no
 * hill (zX = zY = 0) every term reduces to the flat closure's, bit for bit. */
static void ground_metric(const ogrid *g, int i, int j, double *ex, double *ey, double *ez, double *zX, double *zY)
{
    *ex = 1.0 / g->Xx[i];
    *ey = 1.0 / g->Yy[j];
    *ez = 1.0 / g->Zz[1];
    if (!g->terrain) { *zX = 0.0; *zY = 0.0; return; }
    size_t c = (size_t)i + (size_t)g->nx * j;
    *zX = -(g->hx[c] * *ex) * *ez;
    *zY = -(g->hy[c] * *ey) * *ez;
}

static void bc_p_ground(orc_sim *s)
{
    ogrid *g = &s->air;
    for (int j = 1; j < g->ny - 1; ++j)
        for (int i = 1; i < g->nx - 1; ++i) {
            size_t e = IDX(g, i, j, 0);
            double ex, ey, ez, zX, zY;
            ground_metric(g, i, j, &ex, &ey, &ez, &zX, &zY);
            double div;
            if (!g->terrain) {
                div = ex * g->dxi * (g->un[IDX(g, i + 1, j, 0)] - g->un[e]) +
                      ey * g->dyi * (g->vn[IDX(g, i, j + 1, 0)] - g->vn[e]) +
                      ez * g->dzi * (g->wn[IDX(g, i, j, 1)] - g->wn[e]);
            } else {
                double dux = g->un[IDX(g, i + 1, j, 0)] - g->un[e], duz = g->un[IDX(g, i, j, 1)] - g->un[e];
                double dvy = g->vn[IDX(g, i, j + 1, 0)] - g->vn[e], dvz = g->vn[IDX(g, i, j, 1)] - g->vn[e];
                div = ((ex * g->dxi * dux + zX * g->dzi * duz) + (ey * g->dyi * dvy + zY * g->dzi * dvz)) +
                      ez * g->dzi * (g->wn[IDX(g, i, j, 1)] - g->wn[e]);
            }
            g->pnn[e] = g->pn[e] + (-s->g_gama * div);
        }
}

static void bc_v_ground(orc_sim *s)
{
    ogrid *g = &s->air;
    int nx = g->nx;
    double ib = s->g_ib, dt = g->dt, eff = s->g_eff;
    double zi = s->gZinf / s->g_inv;             /* Zinf (1 + tb sum_k a_k) */
    for (int j = 1; j < g->ny - 1; ++j)
        for (int i = 1; i < nx - 1; ++i) {
            size_t e = IDX(g, i, j, 0);
            double ex, ey, ez, zX, zY;
            ground_metric(g, i, j, &ex, &ey, &ez, &zX, &zY);
            double cpl = 2.0 * dt * eff * ez * g->dzi;   /* dt/rho_e over half a cell */
            if (!g->terrain) {
                g->unn[e] = g->un[e] + (-1.0 * ib * dt * eff * (ex * (g->pnn[e] - g->pnn[IDX(g, i - 1, j, 0)]) * g->dxi) + g->un[e] * ib - g->un[e]);
                g->vnn[e] = g->vn[e] + (-1.0 * ib * dt * eff * (ey * (g->pnn[e] - g->pnn[IDX(g, i, j - 1, 0)]) * g->dyi) + g->vn[e] * ib - g->vn[e]);
            } else {
                double pz = (g->pnn[IDX(g, i, j, 1)] - g->pnn[e]) * g->dzi;
                double gx = ex * (g->pnn[e] - g->pnn[IDX(g, i - 1, j, 0)]) * g->dxi + zX * pz;
                double gy = ey * (g->pnn[e] - g->pnn[IDX(g, i, j - 1, 0)]) * g->dyi + zY * pz;
                g->unn[e] = g->un[e] + (-1.0 * ib * dt * eff * gx + g->un[e] * ib - g->un[e]);
                g->vnn[e] = g->vn[e] + (-1.0 * ib * dt * eff * gy + g->vn[e] * ib - g->vn[e]);
            }
            double *Psi = &s->Gpsi[((size_t)i + (size_t)nx * j) * s->gK];
            double temp_psi = 0;
            for (int q = 0; q < s->gK; ++q) temp_psi += s->g_decay[q] * Psi[q];
            double hist = s->gZinf * temp_psi;
            double w = (g->wn[e] - cpl * (g->pnn[e] - hist)) * ib / (1.0 + ib * cpl * zi);
            g->wnn[e] = w;
            double vg = -w;
            for (int q = 0; q < s->gK; ++q) Psi[q] = s->g_decay[q] * Psi[q] + (s->g_gain[q] * vg) * s->g_tb;
        }
}

static void bc_v_tdbc(orc_sim *s)
{
    ogrid *g = &s->air;
    int nx = g->nx, ny = g->ny, nz = g->nz;
    double tb = g->dt / s->tau;
    for (int i = 0; i < nx; ++i)
        for (int k = 0; k < nz; ++k) {
            size_t e = IDX(g, i, ny - 1, k), e1 = IDX(g, i, ny - 2, k);
            g->unn[e] = g->unn[e1];
            g->vnn[e] = 0.0;
            g->wnn[e] = g->wnn[e1];
            double *Psi = &s->Psi[((size_t)i + (size_t)nx * k) * 6];
            double temp_at = 0, temp_psi = 0;
            for (int q = 0; q < 6; ++q) {
                temp_at += TD_A[q] * tb;
                temp_psi += exp(-1 * TD_R[q] * tb) * Psi[q];
            }
            double vv = 1.0 / (1 + temp_at) * (g->pnn[e] / s->Zinf - temp_psi);
            g->vnn[e] = vv;
            for (int q = 0; q < 6; ++q) Psi[q] = exp(-1 * TD_R[q] * tb) * Psi[q] + TD_A[q] * vv * tb;
        }
}

/* Air-porous ghost-plane coupling (calculation_2D.cpp:865-985), copy branch only. */
static int iface_ratio_ok(orc_sim *s) { return (int)(s->air.dz1 / s->pore.dz1) < 5.0; }

static void iface_p(orc_sim *s)
{
    ogrid *a = &s->air, *q = &s->pore;
    if (!iface_ratio_ok(s)) return;
    for (int i = 0; i < a->nx; ++i)
        for (int j = 0; j < a->ny; ++j) {
            a->pnn[IDX(a, i, j, 0)] = q->pnn[IDX(q, i, j, q->nz - 2)];
            q->pnn[IDX(q, i, j, q->nz - 1)] = a->pnn[IDX(a, i, j, 1)];
        }
}

static void iface_v(orc_sim *s)
{
    ogrid *a = &s->air, *q = &s->pore;
    if (!iface_ratio_ok(s)) return;
    for (int i = 0; i < a->nx; ++i)
        for (int j = 0; j < a->ny; ++j) {
            a->unn[IDX(a, i, j, 0)] = q->unn[IDX(q, i, j, q->nz - 2)];
            q->unn[IDX(q, i, j, q->nz - 1)] = a->unn[IDX(a, i, j, 1)];
            a->vnn[IDX(a, i, j, 0)] = q->vnn[IDX(q, i, j, q->nz - 2)];
            q->vnn[IDX(q, i, j, q->nz - 1)] = a->vnn[IDX(a, i, j, 1)];
            a->wnn[IDX(a, i, j, 0)] = q->wnn[IDX(q, i, j, q->nz - 2)];
            q->wnn[IDX(q, i, j, q->nz - 1)] = a->wnn[IDX(a, i, j, 1)];
        }
}

/* ------------------------------------------------------------------------------------ */
/* PML auxiliaries: air::Update_PML_Qvw / Update_PML_Qp (air.cpp:2323-2439)              */
/* ------------------------------------------------------------------------------------ */
static void q_add_x(orc_sim *s, oslab *sl, int f, int i0, const double *src)
{
    ogrid *g = &s->air;
    for (int ii = 0; ii < sl->n; ++ii)
        for (int j = 0; j < g->ny; ++j)
            for (int k = 0; k < g->nz; ++k)
                QX(sl, f, ii, j, k) += g->dt * src[IDX(g, ii + i0, j, k)];
}
static void q_add_y(orc_sim *s, oslab *sl, int f, int j0, const double *src)
{
    ogrid *g = &s->air;
    for (int jj = 0; jj < sl->n; ++jj)
        for (int k = 0; k < g->nz; ++k)
            for (int i = 0; i < g->nx; ++i)
                QY(sl, f, i, jj, k) += g->dt * src[IDX(g, i, jj + j0, k)];
}
static void q_add_z(orc_sim *s, oslab *sl, int f, int k0, const double *src)
{
    ogrid *g = &s->air;
    for (int kk = 0; kk < sl->n; ++kk)
        for (int i = 0; i < g->nx; ++i)
            for (int j = 0; j < g->ny; ++j)
                QZ(sl, f, i, j, kk) += g->dt * src[IDX(g, i, j, kk + k0)];
}

void orc_update_q(orc_sim *s, int which, const double *u, const double *v, const double *w, const double *p)
{
    ogrid *g = &s->air;
    int y1on = (s->sy1.n > 0) && (s->pml_counter >= s->pml_time);
    const double *src[4] = {u, v, w, p};
    int f0 = which == 0 ? 3 : 0, f1 = which == 0 ? 4 : 3;
    for (int f = f0; f < f1; ++f) {
        q_add_x(s, &s->sx1, f, 0, src[f]);
        q_add_x(s, &s->sx2, f, g->nx - s->sx2.n, src[f]);
        if (y1on) q_add_y(s, &s->sy1, f, 0, src[f]);
        q_add_y(s, &s->sy2, f, g->ny - s->sy2.n, src[f]);
        q_add_z(s, &s->sz1, f, 0, src[f]);
        q_add_z(s, &s->sz2, f, g->nz - s->sz2.n, src[f]);
    }
}

/* PML part of a frame change: air::SetQMove (air.cpp:2447-2730), one rank. */
static void shift_slab_x(orc_sim *s, oslab *sl, int temp_J)
{
    ogrid *g = &s->air;
    for (int f = 0; f < 4; ++f)
        for (int ii = 0; ii < sl->n; ++ii)
            for (int j = 0; j < g->ny; ++j)
                for (int k = 0; k < g->nz; ++k) {
                    if (j < temp_J + 1 && j + s->lead - 1 < g->ny - 1) QX(sl, f, ii, j, k) = QX(sl, f, ii, j + s->trail - 1, k);
                    else if (j >= temp_J + 1) QX(sl, f, ii, j, k) = 0.0;
                }
}
static void shift_slab_z(orc_sim *s, oslab *sl, int temp_J)
{
    ogrid *g = &s->air;
    for (int f = 0; f < 4; ++f)
        for (int kk = 0; kk < sl->n; ++kk)
            for (int i = 0; i < g->nx; ++i)
                for (int j = 0; j < g->ny; ++j) {
                    if (j < temp_J + 1 && j + s->lead - 1 < g->ny - 1) QZ(sl, f, i, j, kk) = QZ(sl, f, i, j + s->trail - 1, kk);
                    else if (j >= temp_J + 1) QZ(sl, f, i, j, kk) = 0.0;
                }
}

/* RK4 path (extension): the y layers' auxiliaries move with the frame too.  SetQMove shifts
 * only the x and z layers (air.cpp:2522-2729), so at the corners where a y layer meets an x
 * or z layer the cross terms combine shifted and unshifted auxiliaries, and a run with PML on
 * a y face and on an x or z face grows from frame to frame.  Here a y-layer node takes the
 * value of the node it receives the field from (the field rule of the shift), or zero when
 * that node lies outside the layer (it enters the layer now). */
static void shift_slab_y(orc_sim *s, oslab *sl, int lo, int temp_J)
{
    ogrid *g = &s->air;
    int sh = s->trail - 1;
    for (int f = 0; f < 4; ++f)
        for (int jj = 0; jj < sl->n; ++jj) {
            int j = lo + jj;
            for (int k = 0; k < g->nz; ++k)
                for (int i = 0; i < g->nx; ++i) {
                    if (j < temp_J + 1 && j + s->lead - 1 < g->ny - 1)
                        QY(sl, f, i, jj, k) = (jj + sh < sl->n) ? QY(sl, f, i, jj + sh, k) : 0.0;
                    else if (j >= temp_J + 1)
                        QY(sl, f, i, jj, k) = 0.0;
                }
        }
}

static void set_qmove(orc_sim *s, int N)
{
    if (N == 0) {
        oslab *all[6] = {&s->sx1, &s->sx2, &s->sy1, &s->sy2, &s->sz1, &s->sz2};
        size_t plane[6] = {(size_t)s->air.ny * s->air.nz, (size_t)s->air.ny * s->air.nz,
                           (size_t)s->air.nx * s->air.nz, (size_t)s->air.nx * s->air.nz,
                           (size_t)s->air.nx * s->air.ny, (size_t)s->air.nx * s->air.ny};
        for (int q = 0; q < 6; ++q)
            for (int f = 0; f < 4; ++f)
                if (all[q]->n) memset(all[q]->q[f], 0, sizeof(double) * plane[q] * all[q]->n);
        return;
    }
    int temp_J = s->mf_jmax - s->lead;
    if (temp_J == 0) return;
    if (s->sx1.n) shift_slab_x(s, &s->sx1, temp_J);
    if (s->sx2.n) shift_slab_x(s, &s->sx2, temp_J);
    if (s->sz1.n) shift_slab_z(s, &s->sz1, temp_J);
    if (s->sz2.n) shift_slab_z(s, &s->sz2, temp_J);
    if (s->opt.integrator == LEE_INT_RK4) {
        if (s->sy1.n) shift_slab_y(s, &s->sy1, 0, temp_J);
        if (s->sy2.n) shift_slab_y(s, &s->sy2, s->air.ny - s->sy2.n, temp_J);
    }
}

/* ------------------------------------------------------------------------------------ */
/* receivers (calculation_2D.cpp:317-352, 387-403; airpore.cpp:413-458)                  */
/* ------------------------------------------------------------------------------------ */
static void locate(const orc_sim *s, double rx, double ry, double rz, int *oi, int *oj, int *ok)
{
    const ogrid *g = &s->air;
    int IMAX = g->nx, JMAX = g->ny, KMAX = g->nz, IJK = s->ijkmax;
    int ii = -1, jj = -1, kk = -1;
    double Yf = g->ay[JMAX - 1] - g->ay[0];
    int N = (int)(ry / Yf + 1.0e-6);
    for (int i = 0; i < IMAX; ++i) if (rx - g->ax[i] <= g->dx1 / 8.0) { ii = i; break; }
    int NN = N * (JMAX - 1);
    if (NN >= IJK) { ii = jj = kk = 0; }
    else {
        for (int j = NN; j < IJK && j - NN < JMAX; ++j)
            if (ry - N * Yf - g->ay[j - NN] <= g->dy1 / 8.0) { jj = j; break; }
        for (int k = 0; k < KMAX; ++k) if (rz - g->az[k] <= g->dz1 / 8.0) { kk = k; break; }
        if (kk < 0) { ii = jj = kk = 0; }
    }
    /* airpore.cpp:455-456 (an index the reference leaves undefined counts as out of domain) */
    if ((ii == 0 && jj == 0 && kk == 0) || ii < 0 || jj < 0 || kk < 0) { ii = 1; jj = 1; kk = 1; }
    if (ii > IMAX || jj > IJK || kk > KMAX) { ii = 1; jj = 1; kk = 1; }
    *oi = ii; *oj = jj; *ok = kk;
}

static const double REC_YZ[15][2] = {{-0.23, 1.75}, {-0.46, 0.46}, {-0.1, 0.1}, {1.75, 0.23}, {1.75, 0.46},
                                     {0.6, 0.01}, {0.7, 0.01}, {0.8, 0.01}, {0.9, 0.01}, {1.0, 0.01},
                                     {1.2, 0.01}, {1.4, 0.01}, {1.6, 0.01}, {1.8, 0.01}, {2.0, 0.01}};

static double sample(const orc_sim *s, int r)
{
    const ogrid *g = &s->air;
    int N = s->mf_count, sh = N * (s->trail - 1);
    int ii = s->ri[r], jj = s->rj[r], kk = s->rk[r];
    double t = 0.0;
    if (jj >= sh && jj < g->ny + sh) {
        int jl = jj - sh;
        if (jl > 0 && jl < g->ny - 1 && ii > 0 && ii < g->nx - 1 && kk > 0 && kk < g->nz - 1)
            t = g->pn[IDX(g, ii, jl, kk)];
    }
    if (fabs(t) < 1.0e-100) t = 0.0;
    return t;
}

/* ------------------------------------------------------------------------------------ */
/* construction                                                                          */
/* ------------------------------------------------------------------------------------ */
static int terrain_alloc(ogrid *g, const lee_options *o, int pore_layer)
{
    g->terrain = 1;
    g->pore_layer = pore_layer;
    g->nh = o->n_hills;
    memcpy(g->hills, o->hills, sizeof g->hills);
    size_t n2 = (size_t)g->nx * g->ny;
    g->h = dalloc(n2, 0); g->sh = dalloc(n2, 1); g->hx = dalloc(n2, 0); g->hy = dalloc(n2, 0);
    g->tk = dalloc(g->nz, 1);
    return (g->h && g->sh && g->hx && g->hy && g->tk) ? 0 : -1;
}

orc_sim *orc_create(const lee_deck *d, int nrec, const double *rec, int fill_nn_nan)
{
    lee_options o;
    memset(&o, 0, sizeof o);
    o.precision = 64;
    o.n_receivers = nrec;
    o.receivers = rec;
    return orc_create_opt(d, &o, fill_nn_nan);
}

/* air::cal_hill_var (air.cpp:938-1105) with the material derivation of airpore.cpp:282-293:
 * coefficients, index box and material mask.  Coordinates are those of the air input
 * files (axisbound_*, calculation_2D.cpp:154-200) for the frame-0 index range. */
static const char *obstacle_setup(orc_sim *s)
{
    const lee_deck *d = &s->d;
    s->cr_judge = d->hill_judge;
    if (s->cr_judge == 0) return NULL;
    for (int m = 0; m < 3; ++m) {
        double q = d->hill_mat[m][0], por = d->hill_mat[m][1], res = d->hill_mat[m][2];
        double eff_density = q * q / por / s->Aav;      /* airpore.cpp:286 */
        double Kp = 1.0 / s->gamma / s->Pav;             /* airpore.cpp:287 */
        s->cr_eff[m] = 1.0 / eff_density;                /* air.cpp:958 */
        double Beta = res * d->dt * s->cr_eff[m];        /* air.cpp:962 */
        s->cr_gama[m] = d->dt / Kp / por;                /* air.cpp:963 */
        s->cr_invb[m] = 1.0 / (1.0 + Beta);              /* air.cpp:964 */
    }
    const double *ax = d->air_x, *ay = d->air_y, *az = d->air_z;
    int n[3] = {s->air.nx, s->air.ny, s->air.nz};
    const double *axes[3] = {ax, ay, az};
    double h[3] = {d->dx, d->dy, d->dz};
    for (int a = 0; a < 3; ++a) {                        /* air.cpp:987-1005 */
        int lo = 0, hi = 0;
        for (int q = 0; q < n[a]; ++q) {
            if (fabs(axes[a][q] - d->hill_box[2 * a]) <= h[a] / 8.0) lo = q;
            if (fabs(axes[a][q] - d->hill_box[2 * a + 1]) <= h[a] / 8.0) { hi = q; break; }
        }
        s->ob[2 * a] = lo; s->ob[2 * a + 1] = hi;
    }
    int ni = s->ob[1] - s->ob[0] + 1, nj = s->ob[3] - s->ob[2] + 1, nk = s->ob[5] - s->ob[4] + 1;
    if (ni <= 0 || nj <= 0 || nk <= 0) return "the obstacle bounds select an empty index box";
    s->cr_uc = (int *)calloc((size_t)ni * nj * nk, sizeof(int));
    if (!s->cr_uc) return "oom";
    const double *c = d->hill_coef;
    double A = c[0], B = c[1], C = c[2], D = c[3], E = c[4], F = c[5], G = c[6], H = c[7], I = c[8], J = c[9], K = c[10];
    int ii = 0, jj = 0;
    for (int i = s->ob[0]; i <= s->ob[1]; ++i)
        for (int j = s->ob[2]; j <= s->ob[3]; ++j)
            for (int k = s->ob[4]; k <= s->ob[5]; ++k) {
                int *m = &s->cr_uc[(size_t)(i - s->ob[0]) + (size_t)ni * ((size_t)(j - s->ob[2]) + (size_t)nj * (k - s->ob[4]))];
                double X = ax[i], Y = ay[j], Z = az[k], len, hgt;
                *m = 0;
                switch (s->cr_judge) {
                case 1:   /* sphere of radius b about (c, d, e), air.cpp:1020-1031 */
                    len = pow(X - C, 2) + pow(Y - D, 2) + pow(Z - E, 2);
                    if (sqrt(len) <= B) *m = 1;
                    break;
                case 2:   /* box [a,b] x [c,d] x [e,f], air.cpp:1034-1047 */
                    if (X >= A && X <= B && Y >= C && Y <= D && Z >= E && Z <= F) *m = 1;
                    break;
                case 3:   /* three rows of cylinders of radius a, pitch b, air.cpp:1050-1069 */
                    for (jj = 0; jj < d->hill_points; jj++) {
                        len = pow(X - C, 2) + pow(Y - (D + jj * B), 2);
                        if (sqrt(len) <= A && Z <= E) *m = 1;
                        len = pow(X - F, 2) + pow(Y - (G + jj * B), 2);
                        if (sqrt(len) <= A && Z <= H) *m = 1;
                        len = pow(X - I, 2) + pow(Y - (J + jj * B), 2);
                        if (sqrt(len) <= A && Z <= K) *m = 1;
                    }
                    break;
                case 4:   /* g x f array of cylinders of radius a, air.cpp:1071-1087 */
                    for (ii = 0; ii < G; ii++)
                        for (jj = 0; jj < F; jj++) {
                            len = pow(X - (B + ii * E), 2) + pow(Y - (C + jj * D), 2);
                            if (sqrt(len) <= A) *m = 1;
                        }
                    break;
                case 5:   /* tree at (0, 3): cone canopy 0.5..3 m, trunk r 0.1 m, air.cpp:1089-1103 */
                    for (ii = 0; ii < 1; ii++)
                        for (jj = 0; jj < 1; jj++) {
                            len = pow(X - (0.0 + ii * 2.0), 2) + pow(Y - (3.0 + jj * 2.0), 2);
                            hgt = -25 / (9 - 1e-50) * sqrt(len) + 3.0;
                            if (sqrt(len) <= 0.9 && Z <= hgt && Z >= 0.5) *m = 1;
                            if (sqrt(len) <= 0.1 && Z <= 1.5) *m = 2;
                        }
                    break;
                }
            }
    if (s->cr_judge == 5)   /* second trunk pass with the counters the loops left (air.cpp:1106-1115) */
        for (int i = s->ob[0]; i <= s->ob[1]; ++i)
            for (int j = s->ob[2]; j <= s->ob[3]; ++j)
                for (int k = s->ob[4]; k <= s->ob[5]; ++k) {
                    double len = pow(ax[i] - (0.0 + ii * 2.0), 2) + pow(ay[j] - (3.0 + jj * 2.0), 2);
                    if (sqrt(len) <= 0.1 && az[k] <= 1.5)
                        s->cr_uc[(size_t)(i - s->ob[0]) + (size_t)ni * ((size_t)(j - s->ob[2]) + (size_t)nj * (k - s->ob[4]))] = 2;
                }
    return NULL;
}

orc_sim *orc_create_opt(const lee_deck *d, const lee_options *opt, int fill_nn_nan)
{
    g_err[0] = 0;
    int nrec = opt->n_receivers;
    const double *rec = opt->receivers;
    if (opt->integrator != LEE_INT_DECK && opt->integrator != LEE_INT_RK4) FAIL("unknown integrator");
    if (opt->n_hills < 0 || opt->n_hills > LEE_MAX_HILLS) FAIL("n_hills must be 0..%d", LEE_MAX_HILLS);
    for (int n = 0; n < opt->n_hills; ++n)
        if (!(opt->hills[4 * n + 3] > 0)) FAIL("hill sigma must be positive");
    if (d->velocity_method == 7 && !(opt->wind_z0 > 0)) FAIL("velocity_method 7 needs wind_z0 > 0");
    if (opt->integrator == LEE_INT_RK4 && d->bc[LEE_FACE_E] == LEE_BC_TDBC) FAIL("the TDBC is not defined for RK4");
    if (opt->integrator == LEE_INT_DECK && d->scheme != 2) FAIL("only scheme 2 (FB_p_v) is implemented by the reference");
    if (d->bc[LEE_FACE_N] != LEE_BC_RIGID && d->bc[LEE_FACE_N] != LEE_BC_ABSORBING) FAIL("North must be rigid or absorbing");
    if (d->hill_judge < 0 || d->hill_judge > 5) FAIL("curve_judge %d has no mask", d->hill_judge);
    if (d->hill_judge != 0 && (opt->integrator == LEE_INT_RK4 || opt->n_hills > 0))
        FAIL("embedded porous obstacles run with the FB_p_v step on a separable grid");
    if (d->restart != 0) FAIL("restart is not supported");
    if (d->velocity_method == 3 || d->velocity_method == 6) FAIL("wind methods 3 and 6 read uninitialised members");
    orc_sim *s = (orc_sim *)calloc(1, sizeof *s);
    if (!s) FAIL("oom");
    s->d = *d;
    s->opt = *opt;
    const int *h = d->air_hdr;
    s->mf_judge = d->mf_judge; s->lead = d->mf_lead_dj; s->trail = d->mf_trail_dj; s->ratio = d->mf_ratio;
    int JM = h[3], IJK = h[9];
    /* airpore.cpp:542 divides by lead_DJ-1 whatever the frame judge */
    if (s->lead < 2) { free(s); FAIL("lead_DJ must be >= 2"); }
    if (s->mf_judge == 1) {
        if ((IJK - d->mf_jmax) % (s->lead - 1) != 0) {
            int Nl = (IJK - d->mf_jmax) / (s->lead - 1);
            IJK = d->mf_jmax + (Nl + 1) * (s->lead - 1);
        }
        JM = d->mf_jmax;
    } else {
        IJK = JM;
    }
    s->mf_jmax = JM; s->ijkmax = IJK;
    s->gamma = d->gamma; s->Pav = d->Pav; s->c = d->c;
    s->Aav = pow(d->c, 2) / d->gamma / d->Pav;
    s->vmethod = d->velocity_method; s->vcoef = d->velocity_coef;
    if (grid_init(&s->air, h[0], JM, h[6], d->air_x, d->air_y, d->air_z, d->dt, d->dx, d->dy, d->dz, fill_nn_nan)) {
        free(s); FAIL("oom");
    }
    if (opt->n_hills > 0 && terrain_alloc(&s->air, opt, 0)) { orc_destroy(s); FAIL("oom"); }
    size_t n = (size_t)h[0] * JM * h[6];
    s->U = dalloc(n, 0); s->V = dalloc(n, 0); s->W = dalloc(n, 0);
    const char *e = pml_setup(s);
    if (e) { orc_destroy(s); FAIL("%s", e); }
    e = obstacle_setup(s);
    if (e) { orc_destroy(s); FAIL("%s", e); }
    /* porous ground */
    s->south_porous = d->bc[LEE_FACE_S] == LEE_BC_POROUS;
    if (s->south_porous) {
        if (!d->has_pore) { orc_destroy(s); FAIL("South porous needs coordi_pore files"); }
        const int *ph = d->pore_hdr;
        int pjm = s->mf_judge == 1 ? d->mf_jmax : ph[3];
        if (ph[0] != h[0] || pjm != JM) { orc_destroy(s); FAIL("porous grid must share the air (i,j) index space"); }
        if (grid_init(&s->pore, ph[0], pjm, ph[6], d->pore_x, d->pore_y, d->pore_z, d->dt, d->pore_dx, d->pore_dy, d->pore_dz, fill_nn_nan)) {
            orc_destroy(s); FAIL("oom");
        }
        if (opt->n_hills > 0 && terrain_alloc(&s->pore, opt, 1)) { orc_destroy(s); FAIL("oom"); }
        double Aav = s->Aav;
        double eff_density = d->pore_q * d->pore_q / d->pore_porosity / Aav; /* airpore.cpp:250 */
        double Kp = 1.0 / d->gamma / d->Pav;                                  /* airpore.cpp:251 */
        s->p_eff = 1.0 / eff_density;                                          /* porous.cpp:44 */
        s->p_beta = d->pore_resistivity * d->dt * s->p_eff;                   /* porous.cpp:48 */
        s->p_gama = d->dt / Kp / d->pore_porosity;                            /* porous.cpp:49 */
        s->p_invbeta = 1.0 / (s->p_beta + 1.0);                               /* porous.cpp:50 */
        /* RK4 takes the Darcy drag exactly (integrating factor): e^-beta per step */
        s->p_E = exp(-s->p_beta);
        s->p_E2 = exp(-0.5 * s->p_beta);
    } else if (d->bc[LEE_FACE_S] != LEE_BC_RIGID && d->bc[LEE_FACE_S] != LEE_BC_ABSORBING) {
        orc_destroy(s); FAIL("South must be rigid, absorbing or porous");
    }
    s->east_tdbc = d->bc[LEE_FACE_E] == LEE_BC_TDBC;
    {
        double rho = d->gamma * d->Pav / 340.29 / 340.29;                       /* airpore.cpp:252-254 */
        s->Zinf = rho * 340.29 * d->pore_q / d->pore_porosity;
        s->tau = d->gamma * rho * d->pore_q * d->pore_q / d->pore_resistivity / d->pore_porosity;
    }
    s->Psi = dalloc((size_t)h[0] * h[6] * 6, 0.0); /* zeroed: the reference leaves Psi uninitialised */
    if (opt->ground_poles > 0) {
        if (opt->integrator == LEE_INT_RK4 || d->bc[LEE_FACE_S] != LEE_BC_RIGID || d->mf_judge == 1 || d->hill_judge) {
            orc_destroy(s); FAIL("the ground impedance replaces a rigid South face of an FB_p_v run in a fixed frame");
        }
        if (opt->ground_poles > LEE_MAX_POLES || !(opt->ground_zinf > 0) || !(opt->ground_tau > 0)) {
            orc_destroy(s); FAIL("bad ground impedance");
        }
        if (!(d->pore_q > 0) || !(d->pore_porosity > 0) || !(d->pore_resistivity >= 0)) {
            orc_destroy(s); FAIL("the ground impedance needs the top material (pore_q, pore_porosity, pore_resistivity)");
        }
        {
            double eff_density = d->pore_q * d->pore_q / d->pore_porosity / s->Aav;
            double Kp = 1.0 / d->gamma / d->Pav;
            s->g_eff = 1.0 / eff_density;
            s->g_gama = d->dt / Kp / d->pore_porosity;
            s->g_ib = 1.0 / (d->pore_resistivity * d->dt * s->g_eff + 1.0);
        }
        s->ground = 1; s->gK = opt->ground_poles; s->gZinf = opt->ground_zinf;
        s->g_tb = d->dt / opt->ground_tau;
        double at = 0;
        for (int q = 0; q < s->gK; ++q) {
            if (!(opt->ground_a[q] >= 0) || !(opt->ground_r[q] >= 0)) { orc_destroy(s); FAIL("bad ground poles"); }
            at += opt->ground_a[q] * s->g_tb;
            s->g_decay[q] = exp(-1 * opt->ground_r[q] * s->g_tb);
            s->g_gain[q] = opt->ground_a[q];
        }
        s->g_inv = 1.0 / (1 + at);
        s->Gpsi = dalloc((size_t)h[0] * JM * s->gK, 0.0);
    }
    /* receivers */
    if (nrec > 0) {
        s->nrec = nrec;
    } else {
        s->nrec = (d->rcv[1] == 0) ? 1 : 5;
    }
    s->ri = (int *)calloc(s->nrec, sizeof(int)); s->rj = (int *)calloc(s->nrec, sizeof(int)); s->rk = (int *)calloc(s->nrec, sizeof(int));
    for (int r = 0; r < s->nrec; ++r) {
        double rx, ry, rz;
        if (nrec > 0) { rx = rec[3 * r]; ry = rec[3 * r + 1]; rz = rec[3 * r + 2]; }
        else if (r < 15) { rx = d->rcv[0]; ry = REC_YZ[r][0]; rz = REC_YZ[r][1]; }
        else { rx = d->rcv[0]; ry = d->rcv[1]; rz = d->rcv[2]; }
        locate(s, rx, ry, rz, &s->ri[r], &s->rj[r], &s->rk[r]);
    }
    /* initial condition (airpore.cpp:947, 956-958); terrain heights of frame 0 first */
    if (s->air.terrain) terrain_tables(&s->air);
    if (s->south_porous && s->pore.terrain) terrain_tables(&s->pore);
    grid_ic(&s->air, d->src, d->gauss_width);
    if (s->south_porous) grid_ic(&s->pore, d->src, d->gauss_width);
    s->mf_count = -1;
    return s;
}

void orc_destroy(orc_sim *s)
{
    if (!s) return;
    grid_free(&s->air); grid_free(&s->pore);
    free(s->U); free(s->V); free(s->W); free(s->Psi);
    slab_free(&s->sx1); slab_free(&s->sx2); slab_free(&s->sy1); slab_free(&s->sy2);
    slab_free(&s->sz1); slab_free(&s->sz2);
    for (int d = 0; d < 2; ++d)
        for (int f = 0; f < 4; ++f) { free(s->rk_k[d][f]); free(s->rk_a[d][f]); }
    for (int q = 0; q < 6; ++q)
        for (int f = 0; f < 4; ++f) { free(s->qk[q][f]); free(s->qa[q][f]); free(s->qs[q][f]); }
    free(s->ri); free(s->rj); free(s->rk);
    free(s->cr_uc);
    free(s->src_ps);
    free(s->Gpsi);
    free(s);
}

void orc_get_ground_psi(const orc_sim *s, double *psi)
{
    if (s->ground) memcpy(psi, s->Gpsi, (size_t)s->air.nx * s->air.ny * s->gK * sizeof(double));
}

int orc_obstacle(const orc_sim *s, int *box, int *mask)
{
    if (s->cr_judge == 0) return 0;
    for (int q = 0; q < 6; ++q) box[q] = s->ob[q];
    if (mask) {
        size_t n = (size_t)(s->ob[1] - s->ob[0] + 1) * (s->ob[3] - s->ob[2] + 1) * (s->ob[5] - s->ob[4] + 1);
        memcpy(mask, s->cr_uc, n * sizeof(int));
    }
    return s->cr_judge;
}

void orc_dims(const orc_sim *s, int *dims)
{
    dims[0] = s->air.nx; dims[1] = s->air.ny; dims[2] = s->air.nz;
    dims[3] = s->south_porous ? s->pore.nx : 0; dims[4] = s->south_porous ? s->pore.ny : 0;
    dims[5] = s->south_porous ? s->pore.nz : 0; dims[6] = s->ijkmax; dims[7] = s->nrec;
}

int orc_receiver_indices(const orc_sim *s, int *ijk)
{
    for (int r = 0; r < s->nrec; ++r) { ijk[3 * r] = s->ri[r]; ijk[3 * r + 1] = s->rj[r]; ijk[3 * r + 2] = s->rk[r]; }
    return s->nrec;
}

/* airpore::SetMovingFrame (airpore.cpp:881-897). */
int orc_frame_setup(orc_sim *s, int N)
{
    grid_frame(&s->air, N, s->mf_jmax, s->lead, s->trail);
    set_wind(s, N);
    set_qmove(s, N);
    if (s->south_porous) grid_frame(&s->pore, N, s->mf_jmax, s->lead, s->trail);
    s->mf_count = N;
    return 0;
}

/* Last frame index of the do-while in airpore::get_output (airpore.cpp:542, 624):
 * (int)((JMAX-move_frame.JMAX)/(lead_DJ-1)+1.e-6) with JMAX = IJKMAX and move_frame.JMAX the
 * input value (integer division first, then truncation toward zero). */
static int last_frame(const orc_sim *s)
{
    double x = (double)((s->ijkmax - s->d.mf_jmax) / (s->lead - 1)) + 1.e-6;
    return (int)x;
}

/* MF_limit (airpore.cpp:542-560); the first frame is clamped to time_domain. */
int orc_frame_steps(const orc_sim *s, int mf, int temp_time)
{
    const lee_deck *d = &s->d;
    int lim;
    if (mf == last_frame(s) || s->ijkmax == d->mf_jmax) {
        lim = d->time_domain - temp_time;
    } else if (mf == 0) {
        lim = (int)((d->mf_jmax - 1) * d->dy / d->c / d->dt * s->ratio);
        if (lim > d->time_domain) lim = d->time_domain;
    } else {
        lim = (int)((s->trail - 1) * d->dy / d->c / d->dt);
        double l1 = 0.0, l2 = 0.0;
        if (mf % 10 == 0) { l1 = (s->trail - 1) * d->dy / d->c / d->dt - lim; l1 = l1 * 10.0; }
        if (mf % 100 == 0) { l2 = (s->trail - 1) * d->dy / d->c / d->dt - lim; l2 = l2 * 100 - (int)(l2 * 10) * 10; }
        lim = lim + (int)l1 + (int)l2;
        if (temp_time + lim > d->time_domain) lim = d->time_domain - temp_time;
    }
    return lim;
}

int orc_frame_count(const orc_sim *s)
{
    if (s->mf_judge != 1) return 1;
    int n = last_frame(s) + 1;
    return n < 1 ? 1 : n;
}

/* One time step: airpore::Cal_pressure, Cal_velocity, UpdateInitialCond (airpore.cpp:589-614). */
static void update_source(orc_sim *s, double *p);

static void one_step(orc_sim *s)
{
    ogrid *a = &s->air;
    /* --- pressure (airpore.cpp:366-398) --- */
    orc_tick_gate(s);
    orc_air_dp(s, a->un, a->vn, a->wn, a->pn, a->pnn);
    add_interior(a, a->pnn, a->pn);
    if (s->ground) bc_p_ground(s);
    bc_p(a, FN); bc_p(a, FE); bc_p(a, FW); bc_p(a, FF); bc_p(a, FB);
    if (s->ground) {
        /* no South face: the ground plane's interior nodes are stencil output */
    } else if (!s->south_porous) {
        bc_p(a, FS);
    } else {
        ogrid *q = &s->pore;
        pore_fp(s);
        add_interior(q, q->pnn, q->pn);
        bc_p(q, FS); bc_p(q, FE); bc_p(q, FW); bc_p(q, FF); bc_p(q, FB);
        iface_p(s);
    }
    orc_update_q(s, 0, NULL, NULL, NULL, a->pnn);
    /* --- velocity (airpore.cpp:324-364) --- */
    orc_air_duvw(s, a->un, a->vn, a->wn, a->pnn, a->unn, a->vnn, a->wnn);
    add_interior(a, a->unn, a->un); add_interior(a, a->vnn, a->vn); add_interior(a, a->wnn, a->wn);
    if (s->ground) bc_v_ground(s);
    bc_v(a, FN);
    if (s->east_tdbc) bc_v_tdbc(s); else bc_v(a, FE);
    bc_v(a, FW); bc_v(a, FF); bc_v(a, FB);
    if (s->ground) {
        /* no South face */
    } else if (!s->south_porous) {
        bc_v(a, FS);
    } else {
        ogrid *q = &s->pore;
        pore_fuvw(s);
        add_interior(q, q->unn, q->un); add_interior(q, q->vnn, q->vn); add_interior(q, q->wnn, q->wn);
        bc_v(q, FS); bc_v(q, FE); bc_v(q, FW); bc_v(q, FF); bc_v(q, FB);
        iface_v(s);
    }
    orc_update_q(s, 1, a->unn, a->vnn, a->wnn, NULL);
    update_source(s, a->pnn);
    /* --- nn -> n (calculation_2D.cpp:354-369) --- */
    size_t n = (size_t)a->nx * a->ny * a->nz;
    memcpy(a->un, a->unn, n * 8); memcpy(a->vn, a->vnn, n * 8);
    memcpy(a->wn, a->wnn, n * 8); memcpy(a->pn, a->pnn, n * 8);
    if (s->south_porous) {
        ogrid *q = &s->pore;
        size_t m = (size_t)q->nx * q->ny * q->nz;
        memcpy(q->un, q->unn, m * 8); memcpy(q->vn, q->vnn, m * 8);
        memcpy(q->wn, q->wnn, m * 8); memcpy(q->pn, q->pnn, m * 8);
    }
}

/* ------------------------------------------------------------------------------------ */
/* RK4 (extension; the reference parses scheme 5 but never integrates with it,           */
/* calculation_2D.cpp:551-612)                                                           */
/* ------------------------------------------------------------------------------------ */
/* Porous velocity right-hand side times dt without the drag: porous.cpp:112-135 as the ODE
 * du/dt = -(1/rho_e) grad p - (sigma/rho_e) u, whose stiff linear drag the RK4 step takes
 * exactly through the integrating factor e^{-beta t/dt} (Lawson's RK4), so dt*f here is
 * -dt*eff*grad p only. */
static void pore_rk_duvw_on(orc_sim *s, const double *u, const double *v, const double *w,
                            const double *p, double *ou, double *ov, double *ow)
{
    ogrid *g = &s->pore;
    double dxi = g->dxi, dyi = g->dyi, dzi = g->dzi, dt = g->dt;
    double f = (-dt) * s->p_eff;
#pragma omp parallel for schedule(static)
    for (int k = 1; k < g->nz - 1; ++k)
        for (int j = 1; j < g->ny - 1; ++j)
            for (int i = 1; i < g->nx - 1; ++i) {
                metric9 m = point_metric(g, i, j, k);
                size_t id = IDX(g, i, j, k);
                ou[id] = f * (m.xX * BDX(p) * dxi + m.yX * BDY(p) * dyi + m.zX * BDZ(p) * dzi);
                ov[id] = f * (m.xY * BDX(p) * dxi + m.yY * BDY(p) * dyi + m.zY * BDZ(p) * dzi);
                ow[id] = f * (m.xZ * BDX(p) * dxi + m.yZ * BDY(p) * dyi + m.zZ * BDZ(p) * dzi);
            }
}

static const double RK_A[3] = {0.5, 0.5, 1.0};
static const double RK_B[4] = {1.0 / 6.0, 1.0 / 3.0, 1.0 / 3.0, 1.0 / 6.0};

/* x_next / acc update of one value at stage st (0..3): acc accumulates x_n + sum b k; the
 * next stage state is x_n + a k, and after the last stage x_{n+1} = acc + b4 k. */
static double rk_combine(int st, double xn, double k, double *acc)
{
    if (st == 0) *acc = xn + RK_B[0] * k;
    else if (st < 3) *acc = *acc + RK_B[st] * k;
    return st < 3 ? xn + RK_A[st] * k : *acc + RK_B[3] * k;
}

/* Lawson (integrating-factor) RK4 for x' = -(beta/dt) x + N(x), k = dt N(stage):
 *   stage 0: acc = E x_n + (E/6) k1,      next = E2 x_n + (E2/2) k1
 *   stage 1: acc += (E2/3) k2,            next = E2 x_n + (1/2) k2
 *   stage 2: acc += (E2/3) k3,            next = E x_n + E2 k3
 *   stage 3: x_{n+1} = acc + (1/6) k4,    E = e^-beta, E2 = e^-beta/2
 * (classical RK4 in the variable e^{beta t/dt} x).  E = E2 = 1 gives rk_combine's values. */
static double rk_combine_if(int st, double xn, double k, double *acc, double E, double E2)
{
    const double c[3] = {E2, E2, E}, a[3] = {E2 * 0.5, 0.5, E2}, w[4] = {E / 6.0, E2 / 3.0, E2 / 3.0, 1.0 / 6.0};
    if (st == 0) *acc = E * xn + w[0] * k;
    else if (st < 3) *acc = *acc + w[st] * k;
    return st < 3 ? c[st] * xn + a[st] * k : *acc + w[3] * k;
}

/* Face conditions of the state held in the nn arrays, in the reference's face order. */
static void rk_faces(orc_sim *s)
{
    ogrid *a = &s->air;
    bc_p(a, FN); bc_p(a, FE); bc_p(a, FW); bc_p(a, FF); bc_p(a, FB);
    if (!s->south_porous) {
        bc_p(a, FS);
    } else {
        ogrid *q = &s->pore;
        bc_p(q, FS); bc_p(q, FE); bc_p(q, FW); bc_p(q, FF); bc_p(q, FB);
        iface_p(s);
    }
    bc_v(a, FN); bc_v(a, FE); bc_v(a, FW); bc_v(a, FF); bc_v(a, FB);
    if (!s->south_porous) {
        bc_v(a, FS);
    } else {
        ogrid *q = &s->pore;
        bc_v(q, FS); bc_v(q, FE); bc_v(q, FW); bc_v(q, FF); bc_v(q, FB);
        iface_v(s);
    }
}

static int rk_alloc(orc_sim *s)
{
    if (s->rk_k[0][0]) return 0;
    ogrid *g[2] = {&s->air, &s->pore};
    for (int d = 0; d < 2; ++d) {
        if (d == 1 && !s->south_porous) break;
        size_t n = (size_t)g[d]->nx * g[d]->ny * g[d]->nz;
        for (int f = 0; f < 4; ++f) {
            s->rk_k[d][f] = dalloc(n, 0); s->rk_a[d][f] = dalloc(n, 0);
            if (!s->rk_k[d][f] || !s->rk_a[d][f]) return -1;
        }
    }
    for (int q = 0; q < 6; ++q) {
        long n = orc_q_size(s, q);
        for (int f = 0; f < 4; ++f) {
            s->qk[q][f] = dalloc(n, 0); s->qa[q][f] = dalloc(n, 0); s->qs[q][f] = dalloc(n, 0);
            if (!s->qk[q][f] || !s->qa[q][f] || !s->qs[q][f]) return -1;
        }
    }
    return 0;
}

/* One classical RK4 step.  The stage state lives in the nn arrays (fields, with its face
 * values) and in qs (auxiliaries); level n stays in the n arrays and the slabs. */
static void rk4_step(orc_sim *s)
{
    ogrid *g[2] = {&s->air, &s->pore};
    oslab *sl[6] = {&s->sx1, &s->sx2, &s->sy1, &s->sy2, &s->sz1, &s->sz2};
    const int nd = s->south_porous ? 2 : 1;
    orc_tick_gate(s);
    const int y1on = (s->sy1.n > 0) && (s->pml_counter >= s->pml_time);
    for (int d = 0; d < nd; ++d) {
        size_t n = (size_t)g[d]->nx * g[d]->ny * g[d]->nz;
        memcpy(g[d]->unn, g[d]->un, n * 8); memcpy(g[d]->vnn, g[d]->vn, n * 8);
        memcpy(g[d]->wnn, g[d]->wn, n * 8); memcpy(g[d]->pnn, g[d]->pn, n * 8);
    }
    for (int q = 0; q < 6; ++q)
        for (int f = 0; f < 4; ++f) {
            long n = orc_q_size(s, q);
            if (n) memcpy(s->qs[q][f], sl[q]->q[f], n * 8);
            s->qsave[q][f] = sl[q]->q[f];
        }
    for (int st = 0; st < 4; ++st) {
        ogrid *a = &s->air;
        /* increments of the stage state (the PML terms read the stage auxiliaries) */
        for (int q = 0; q < 6; ++q) for (int f = 0; f < 4; ++f) sl[q]->q[f] = s->qs[q][f];
        orc_air_dp(s, a->unn, a->vnn, a->wnn, a->pnn, s->rk_k[0][3]);
        orc_air_duvw(s, a->unn, a->vnn, a->wnn, a->pnn, s->rk_k[0][0], s->rk_k[0][1], s->rk_k[0][2]);
        if (s->south_porous) {
            ogrid *p = &s->pore;
            pore_fp_on(s, p->unn, p->vnn, p->wnn, s->rk_k[1][3]);
            pore_rk_duvw_on(s, p->unn, p->vnn, p->wnn, p->pnn, s->rk_k[1][0], s->rk_k[1][1], s->rk_k[1][2]);
        }
        for (int q = 0; q < 6; ++q) for (int f = 0; f < 4; ++f) sl[q]->q[f] = s->qsave[q][f];
        /* auxiliaries: dQ/dt = f at every slab node, f of the stage state (face values included) */
        for (int q = 0; q < 6; ++q) {
            oslab *b = sl[q];
            if (!b->n || (q == 2 && !y1on)) continue;
            int ax = q >> 1;
            int lo = (q & 1) ? (ax == 0 ? a->nx : ax == 1 ? a->ny : a->nz) - b->n : 0;
            const double *src[4] = {a->unn, a->vnn, a->wnn, a->pnn};
            long cnt = orc_q_size(s, q);
#pragma omp parallel for schedule(static)
            for (long e = 0; e < cnt; ++e) {
                long A = ax == 0 ? b->n : a->nx, B = ax == 1 ? b->n : a->ny;
                int i = (int)(e % A), j = (int)((e / A) % B), k = (int)(e / (A * B));
                if (ax == 0) i += lo; else if (ax == 1) j += lo; else k += lo;
                size_t id = IDX(a, i, j, k);
                for (int f = 0; f < 4; ++f) {
                    double kq = a->dt * src[f][id];
                    s->qs[q][f][e] = rk_combine(st, b->q[f][e], kq, &s->qa[q][f][e]);
                }
            }
        }
        /* fields at interior points */
        for (int d = 0; d < nd; ++d) {
            ogrid *G = g[d];
            double *xn[4] = {G->un, G->vn, G->wn, G->pn}, *xs[4] = {G->unn, G->vnn, G->wnn, G->pnn};
#pragma omp parallel for schedule(static)
            for (int k = 1; k < G->nz - 1; ++k)
                for (int j = 1; j < G->ny - 1; ++j)
                    for (int i = 1; i < G->nx - 1; ++i) {
                        size_t id = IDX(G, i, j, k);
                        for (int f = 0; f < 4; ++f)
                            xs[f][id] = (d == 1 && f < 3)
                                            ? rk_combine_if(st, xn[f][id], s->rk_k[d][f][id], &s->rk_a[d][f][id], s->p_E, s->p_E2)
                                            : rk_combine(st, xn[f][id], s->rk_k[d][f][id], &s->rk_a[d][f][id]);
                    }
        }
        rk_faces(s);
    }
    /* the last stage state is level n+1 */
    for (int d = 0; d < nd; ++d) {
        size_t n = (size_t)g[d]->nx * g[d]->ny * g[d]->nz;
        memcpy(g[d]->un, g[d]->unn, n * 8); memcpy(g[d]->vn, g[d]->vnn, n * 8);
        memcpy(g[d]->wn, g[d]->wnn, n * 8); memcpy(g[d]->pn, g[d]->pnn, n * 8);
    }
    for (int q = 0; q < 6; ++q)
        for (int f = 0; f < 4; ++f) {
            long n = orc_q_size(s, q);
            if (n) memcpy(sl[q]->q[f], s->qs[q][f], n * 8);
        }
}

/* calculation_2D::UpdateSource (calculation_2D.cpp:371-385) with N = index_pressure: every
 * node of the air pressure at the new level gains ps[N/23] * cal_InitialPressure(...) when
 * N % 23 == 0 (the period is a parameter here).  Called after the velocity phase, before
 * nn -> n (the call site airpore.cpp:608-611); on the RK4 path after the step. */
static void update_source(orc_sim *s, double *p)
{
    long N = s->index_pressure;
    if (s->src_n == 0 || N % s->src_period != 0 || N / s->src_period >= s->src_n) return;
    ogrid *g = &s->air;
    double gc = 4.0 * log(2.0) / pow(s->d.gauss_width, 2);      /* calculation_2D.cpp:26 */
    const double *src = s->d.src;
    for (int i = 0; i < g->nx; i++)
        for (int j = 0; j < g->ny; j++)
            for (int k = 0; k < g->nz; k++) {
                double x = g->X[i] - src[0], y = g->Y[j] - src[1], z = g->Z[k] - src[2];
                double r = sqrt(pow(x, 2) + pow(y, 2) + pow(z, 2));   /* calculation_2D.cpp:292-298 */
                size_t id = IDX(g, i, j, k);
                p[id] = p[id] + s->src_ps[N / s->src_period] * (1.0 * exp(-gc * r * r));
            }
}

int orc_set_source_series(orc_sim *s, const double *ps, long n, int period)
{
    if (n < 0 || (n > 0 && (!ps || period < 1))) return -1;
    if (n > 0 && s->air.terrain) return -1;
    free(s->src_ps);
    s->src_ps = NULL; s->src_n = 0; s->src_period = 0;
    if (n == 0) return 0;
    s->src_ps = (double *)malloc((size_t)n * sizeof(double));
    if (!s->src_ps) return -1;
    memcpy(s->src_ps, ps, (size_t)n * sizeof(double));
    s->src_n = n; s->src_period = period;
    return 0;
}

void orc_step(orc_sim *s, int nsteps, double *rec)
{
    if (s->opt.integrator == LEE_INT_RK4 && rk_alloc(s)) return;
    for (int t = 0; t < nsteps; ++t) {
        if (rec) for (int r = 0; r < s->nrec; ++r) rec[(size_t)t * s->nrec + r] = sample(s, r);
        s->index_pressure += 1;
        if (s->opt.integrator == LEE_INT_RK4) {
            rk4_step(s);
            update_source(s, s->air.pn);
        } else {
            one_step(s);
        }
    }
}

int orc_run(orc_sim *s, double *rec)
{
    int mf = 0, temp_time = 0, nmax = last_frame(s);
    do {
        int lim = orc_frame_steps(s, mf, temp_time);
        orc_frame_setup(s, mf);
        if (lim > 0) orc_step(s, lim, rec ? rec + (size_t)temp_time * s->nrec : NULL);
        mf = mf + 1;
        temp_time = temp_time + (lim > 0 ? lim : 0);
    } while (mf <= nmax && temp_time < s->d.time_domain && s->mf_judge == 1);
    return temp_time;
}

void orc_get_state(const orc_sim *s, int domain, double *u, double *v, double *w, double *p)
{
    const ogrid *g = domain ? &s->pore : &s->air;
    size_t n = (size_t)g->nx * g->ny * g->nz;
    if (u) memcpy(u, g->un, n * 8);
    if (v) memcpy(v, g->vn, n * 8);
    if (w) memcpy(w, g->wn, n * 8);
    if (p) memcpy(p, g->pn, n * 8);
}

void orc_set_state(orc_sim *s, int domain, const double *u, const double *v, const double *w, const double *p)
{
    ogrid *g = domain ? &s->pore : &s->air;
    size_t n = (size_t)g->nx * g->ny * g->nz;
    if (u) memcpy(g->un, u, n * 8);
    if (v) memcpy(g->vn, v, n * 8);
    if (w) memcpy(g->wn, w, n * 8);
    if (p) memcpy(g->pn, p, n * 8);
}

void orc_get_psi(const orc_sim *s, double *psi)
{
    memcpy(psi, s->Psi, sizeof(double) * (size_t)s->air.nx * s->air.nz * 6);
}

long orc_q_size(const orc_sim *s, int slab)
{
    const oslab *all[6] = {&s->sx1, &s->sx2, &s->sy1, &s->sy2, &s->sz1, &s->sz2};
    size_t plane[6] = {(size_t)s->air.ny * s->air.nz, (size_t)s->air.ny * s->air.nz,
                       (size_t)s->air.nx * s->air.nz, (size_t)s->air.nx * s->air.nz,
                       (size_t)s->air.nx * s->air.ny, (size_t)s->air.nx * s->air.ny};
    return (long)(all[slab]->n * plane[slab]);
}
void orc_q_get(const orc_sim *s, int slab, int field, double *buf)
{
    const oslab *all[6] = {&s->sx1, &s->sx2, &s->sy1, &s->sy2, &s->sz1, &s->sz2};
    long n = orc_q_size(s, slab);
    if (n) memcpy(buf, all[slab]->q[field], n * 8);
}
void orc_q_set(orc_sim *s, int slab, int field, const double *buf)
{
    oslab *all[6] = {&s->sx1, &s->sx2, &s->sy1, &s->sy2, &s->sz1, &s->sz2};
    long n = orc_q_size(s, slab);
    if (n) memcpy(all[slab]->q[field], buf, n * 8);
}
/* The nine inverse-metric terms and the Jacobian at an interior node (xX xY xZ yX yY yZ
 * zX zY zZ J) of the air (domain 0) or porous (1) grid. */
void orc_metric(const orc_sim *s, int domain, int i, int j, int k, double *out)
{
    const ogrid *g = domain ? &s->pore : &s->air;
    metric9 m = point_metric(g, i, j, k);
    double v[10] = {m.xX, m.xY, m.xZ, m.yX, m.yY, m.yZ, m.zX, m.zY, m.zZ, g->Jac[IDX(g, i, j, k)]};
    memcpy(out, v, sizeof v);
}

void orc_wind(const orc_sim *s, double *U, double *V, double *W)
{
    size_t n = (size_t)s->air.nx * s->air.ny * s->air.nz;
    memcpy(U, s->U, n * 8); memcpy(V, s->V, n * 8); memcpy(W, s->W, n * 8);
}

/* FFT (FFT.cpp:15-82): bit reversal then iterative radix-2 butterflies with the
 * recursive twiddle update c1,c2 <- sqrt((1 +- c1)/2). */
/* Host threads of the stencil loops (OpenMP; every loop writes disjoint points, so results
 * do not depend on the count). */
void orc_set_threads(int n) { if (n > 0) omp_set_num_threads(n); }
int orc_get_threads(void) { return omp_get_max_threads(); }

void orc_fft(int dir, int m, double *x, double *y)
{
    long n = 1;
    for (int i = 0; i < m; ++i) n *= 2;
    long i2 = n >> 1, j = 0;
    for (long i = 0; i < n - 1; ++i) {
        if (i < j) { double tx = x[i], ty = y[i]; x[i] = x[j]; y[i] = y[j]; x[j] = tx; y[j] = ty; }
        long k = i2;
        while (k <= j) { j -= k; k >>= 1; }
        j += k;
    }
    double c1 = -1.0, c2 = 0.0;
    long l2 = 1;
    for (int l = 0; l < m; ++l) {
        long l1 = l2;
        l2 <<= 1;
        double u1 = 1.0, u2 = 0.0;
        for (long jj = 0; jj < l1; ++jj) {
            for (long i = jj; i < n; i += l2) {
                long i1 = i + l1;
                double t1 = u1 * x[i1] - u2 * y[i1];
                double t2 = u1 * y[i1] + u2 * x[i1];
                x[i1] = x[i] - t1; y[i1] = y[i] - t2;
                x[i] += t1; y[i] += t2;
            }
            double z = u1 * c1 - u2 * c2;
            u2 = u1 * c2 + u2 * c1;
            u1 = z;
        }
        c2 = sqrt((1.0 - c1) / 2.0);
        if (dir == 1) c2 = -c2;
        c1 = sqrt((1.0 + c1) / 2.0);
    }
    if (dir == 1) for (long i = 0; i < n; ++i) { x[i] /= n; y[i] /= n; }
}
