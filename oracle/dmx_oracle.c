/*
 * dmx_oracle.c -- CPU FP64 restatement of the reference MHD step.
 *
 * TEST INFRASTRUCTURE ONLY (see dmx_oracle.h): the checker for the HIP path.
 *
 * Each function names the reference routine (file:line, paths relative to
 * marcjoos-cea/dumses-hybrid/src) whose arithmetic it restates.  Expressions
 * keep the reference's evaluation order, because the goal is bitwise equality
 * with the compiled reference: Fortran evaluates a+b+c as (a+b)+c and a*b/c
 * as (a*b)/c; flang lowers MAX(a,b) to (a > b ? a : b), MIN(a,b) to
 * (a < b ? a : b), SIGN(1,x) to copysign(1,x) and real MOD to fmod (all
 * measured on amdflang 22).  Build with -ffp-contract=off.
 *
 * Storage: every 3D array is allocated over the widest reference range
 * [-2, n+4] per direction (bfc/dq/dbfc use iu2+1 = n+4) and zero-initialised,
 * as init_param zeroes the reference arrays (init.f90:74-81).
 */
#include "dmx_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NG 3

static const double zero = 0.0, one = 1.0, two = 2.0, half = 0.5,
                    forth = 0.25;
static const double smallr = 1e-10, smallc = 1e-10;

static inline double fmx(double a, double b) { return a > b ? a : b; }
static inline double fmn(double a, double b) { return a < b ? a : b; }
static inline double sgn1(double x) { return copysign(1.0, x); }

/* variable indices, 0-based (commons.f90:72-80) */
enum { IR = 0, IU, IV, IW, IP, IA, IB, IC };

struct dmxo {
    dmx_params p;
    int nx, ny, nz;
    long sx, sy, sz, n3;          /* widest extents n+7 */
    double dx, dy, dz;
    double *x, *y, *z;            /* x[i+2], i in [-2, nx+3] */
    double *uin;                  /* 11 vars */
    double *qin;                  /* 8 vars */
    double *bfc;                  /* 3 */
    double *dq;                   /* [dir][var] 3x8 */
    double *dbfc;                 /* [dir][comp] 3x3 */
    double *ex, *ey, *ez;
    double *qm, *qp;              /* [dir][var] */
    double *qrt, *qrb, *qlt, *qlb;/* [dir][var] */
    double *fgod;                 /* [dir][var] */
    double *fpre;                 /* [dir] */
    double *gstar;                /* 7: r u v w b c p */
    double *flux;                 /* [dir][var] */
    double *emfx, *emfy, *emfz;
    double *uold;                 /* 11 vars (fargo) */
    double *duin;                 /* 8 vars (fargo) */
};

#define IX(s, i, j, k) \
    ((long)((i) + 2) + (s)->sx * ((long)((j) + 2) + (s)->sy * (long)((k) + 2)))
#define F3(s, arr, v, i, j, k) ((arr)[(long)(v) * (s)->n3 + IX(s, i, j, k)])
#define U(i, j, k, v) F3(s, s->uin, v, i, j, k)
#define Q(i, j, k, v) F3(s, s->qin, v, i, j, k)
#define BFC(i, j, k, c) F3(s, s->bfc, c, i, j, k)
#define DQ(i, j, k, v, d) F3(s, s->dq, (d) * 8 + (v), i, j, k)
#define DBFC(i, j, k, c, d) F3(s, s->dbfc, (d) * 3 + (c), i, j, k)
#define EX(i, j, k) F3(s, s->ex, 0, i, j, k)
#define EY(i, j, k) F3(s, s->ey, 0, i, j, k)
#define EZ(i, j, k) F3(s, s->ez, 0, i, j, k)
#define QM(i, j, k, v, d) F3(s, s->qm, (d) * 8 + (v), i, j, k)
#define QP(i, j, k, v, d) F3(s, s->qp, (d) * 8 + (v), i, j, k)
#define QRT(i, j, k, v, d) F3(s, s->qrt, (d) * 8 + (v), i, j, k)
#define QRB(i, j, k, v, d) F3(s, s->qrb, (d) * 8 + (v), i, j, k)
#define QLT(i, j, k, v, d) F3(s, s->qlt, (d) * 8 + (v), i, j, k)
#define QLB(i, j, k, v, d) F3(s, s->qlb, (d) * 8 + (v), i, j, k)
#define FG(i, j, k, v, d) F3(s, s->fgod, (d) * 8 + (v), i, j, k)
#define FPRE(i, j, k, d) F3(s, s->fpre, d, i, j, k)
#define GST(i, j, k, v) F3(s, s->gstar, v, i, j, k)
#define FLX(i, j, k, v, d) F3(s, s->flux, (d) * 8 + (v), i, j, k)
#define EMFX(i, j, k) F3(s, s->emfx, 0, i, j, k)
#define EMFY(i, j, k) F3(s, s->emfy, 0, i, j, k)
#define EMFZ(i, j, k) F3(s, s->emfz, 0, i, j, k)
#define UOLD(i, j, k, v) F3(s, s->uold, v, i, j, k)
#define DUIN(i, j, k, v) F3(s, s->duin, v, i, j, k)
#define XC(i) (s->x[(i) + 2])

static double* zalloc(long n) { return (double*)calloc((size_t)n, sizeof(double)); }

dmxo* dmxo_create(const dmx_params* p)
{
    dmxo* s = (dmxo*)calloc(1, sizeof(dmxo));
    s->p = *p;
    s->nx = p->nx; s->ny = p->ny; s->nz = p->nz;
    s->sx = p->nx + 7; s->sy = p->ny + 7; s->sz = p->nz + 7;
    s->n3 = s->sx * s->sy * s->sz;
    /* init_grid, init.f90:438-440 and :464-505 (one rank: nxglob = nx) */
    s->dx = (p->xmax - p->xmin) / p->nx;
    s->dy = (p->ymax - p->ymin) / p->ny;
    s->dz = (p->zmax - p->zmin) / p->nz;
    s->x = zalloc(s->sx); s->y = zalloc(s->sy); s->z = zalloc(s->sz);
    for (int i = -2; i <= p->nx + 3; i++) s->x[i + 2] = p->xmin + s->dx / 2 + (double)(i - 1) * s->dx;
    for (int j = -2; j <= p->ny + 3; j++) s->y[j + 2] = p->ymin + s->dy / 2 + (double)(j - 1) * s->dy;
    for (int k = -2; k <= p->nz + 3; k++) s->z[k + 2] = p->zmin + s->dz / 2 + (double)(k - 1) * s->dz;
    long n = s->n3;
    s->uin = zalloc(11 * n); s->qin = zalloc(8 * n); s->bfc = zalloc(3 * n);
    s->dq = zalloc(24 * n); s->dbfc = zalloc(9 * n);
    s->ex = zalloc(n); s->ey = zalloc(n); s->ez = zalloc(n);
    s->qm = zalloc(24 * n); s->qp = zalloc(24 * n);
    s->qrt = zalloc(24 * n); s->qrb = zalloc(24 * n);
    s->qlt = zalloc(24 * n); s->qlb = zalloc(24 * n);
    s->fgod = zalloc(24 * n); s->fpre = zalloc(3 * n); s->gstar = zalloc(7 * n);
    s->flux = zalloc(24 * n);
    s->emfx = zalloc(n); s->emfy = zalloc(n); s->emfz = zalloc(n);
    s->uold = zalloc(11 * n); s->duin = zalloc(8 * n);
    return s;
}

void dmxo_destroy(dmxo* s)
{
    if (!s) return;
    double* arrs[] = {s->x, s->y, s->z, s->uin, s->qin, s->bfc, s->dq, s->dbfc,
                      s->ex, s->ey, s->ez, s->qm, s->qp, s->qrt, s->qrb, s->qlt,
                      s->qlb, s->fgod, s->fpre, s->gstar, s->flux, s->emfx,
                      s->emfy, s->emfz, s->uold, s->duin};
    for (size_t a = 0; a < sizeof(arrs) / sizeof(arrs[0]); a++) free(arrs[a]);
    free(s);
}

/* reference layout: uin(-2:nx+3, -2:ny+3, -2:nz+3, 11), i fastest */
void dmxo_load(dmxo* s, const double* uin)
{
    long gx = s->nx + 6, gy = s->ny + 6, gz = s->nz + 6;
    for (int v = 0; v < 11; v++)
        for (int k = -2; k <= s->nz + 3; k++)
            for (int j = -2; j <= s->ny + 3; j++)
                for (int i = -2; i <= s->nx + 3; i++)
                    U(i, j, k, v) = uin[(i + 2) + gx * ((j + 2) + gy * ((k + 2) + gz * v))];
}

void dmxo_store(const dmxo* s, double* uin)
{
    long gx = s->nx + 6, gy = s->ny + 6, gz = s->nz + 6;
    for (int v = 0; v < 11; v++)
        for (int k = -2; k <= s->nz + 3; k++)
            for (int j = -2; j <= s->ny + 3; j++)
                for (int i = -2; i <= s->nx + 3; i++)
                    uin[(i + 2) + gx * ((j + 2) + gy * ((k + 2) + gz * v))] = U(i, j, k, v);
}

/* ------------------------------------------------------------------------ */
/* primitive(dt), scheme.f90:20-91 (Cartesian)                               */
/* ------------------------------------------------------------------------ */
static void primitive(dmxo* s, double dt)
{
    const dmx_params* p = &s->p;
    double gamma = p->gamma, ciso = p->ciso, Omega0 = p->omega0;
    double smallp;
    if (p->iso) smallp = smallr * (ciso * ciso);
    else        smallp = smallr * (smallc * smallc) / gamma;
#pragma omp parallel for schedule(static)
    for (int k = -2; k <= s->nz + 3; k++)
        for (int j = -2; j <= s->ny + 3; j++)
            for (int i = -2; i <= s->nx + 3; i++) {
                double r = U(i, j, k, IR);
                Q(i, j, k, IR) = fmx(r, smallr);
                Q(i, j, k, IU) = U(i, j, k, IU) / r;
                Q(i, j, k, IV) = U(i, j, k, IV) / r;
                Q(i, j, k, IW) = U(i, j, k, IW) / r;
                Q(i, j, k, IA) = half * (U(i, j, k, IA) + U(i, j, k, IA + 3));
                Q(i, j, k, IB) = half * (U(i, j, k, IB) + U(i, j, k, IB + 3));
                Q(i, j, k, IC) = half * (U(i, j, k, IC) + U(i, j, k, IC + 3));
                if (p->iso) {
                    Q(i, j, k, IP) = r * (ciso * ciso);
                } else {
                    double u = Q(i, j, k, IU), v = Q(i, j, k, IV), w = Q(i, j, k, IW);
                    double A = Q(i, j, k, IA), B = Q(i, j, k, IB), C = Q(i, j, k, IC);
                    double eken = half * ((u * u + v * v) + w * w);
                    double emag = half * ((A * A + B * B) + C * C);
                    double eint = (U(i, j, k, IP) - emag) / r - eken;
                    Q(i, j, k, IP) = fmx((gamma - one) * Q(i, j, k, IR) * eint, smallp);
                }
                if (Omega0 > zero) {
                    double dvx = two * Omega0 * Q(i, j, k, IV);
                    double dvy = -half * Omega0 * Q(i, j, k, IU);
                    Q(i, j, k, IU) = Q(i, j, k, IU) + dvx * dt * half;
                    Q(i, j, k, IV) = Q(i, j, k, IV) + dvy * dt * half;
                }
            }
}

/* comp_speed, physics.f90:148-164 */
static inline double comp_speed(double gamma, double rho, double pr, double v,
                                double bn, double bt1, double bt2)
{
    double c2 = gamma * pr / rho;
    double b2 = (bn * bn + bt1 * bt1) + bt2 * bt2;
    double d2 = half * (c2 + b2 / rho);
    double cf = sqrt(d2 + sqrt(d2 * d2 - c2 * bn * bn / rho));
    return cf + fabs(v);
}

/* compute_dt, physics.f90:18-132 (OACC=0 branch, NDIM=3, Cartesian) */
double dmxo_compute_dt(dmxo* s)
{
    const dmx_params* p = &s->p;
    primitive(s, zero);
    double dt = p->courant * s->dx / smallc;
    double g = p->gamma;
    int shear = (p->omega0 > zero) && !p->fargo;
    double vshear = 1.5 * p->omega0 * (p->xmax - p->xmin) * half;
#pragma omp parallel for schedule(static) reduction(min : dt)
    for (int k = 1; k <= s->nz; k++)
        for (int j = 1; j <= s->ny; j++)
            for (int i = 1; i <= s->nx; i++) {
                double rho = Q(i, j, k, IR), vx = Q(i, j, k, IU), vy = Q(i, j, k, IV),
                       vz = Q(i, j, k, IW), pr = Q(i, j, k, IP), bx = Q(i, j, k, IA),
                       by = Q(i, j, k, IB), bz = Q(i, j, k, IC);
                double velx = comp_speed(g, rho, pr, vx, bx, by, bz);
                if (s->nx == 1) velx = zero;
                double vely = comp_speed(g, rho, pr, vy, by, bz, bx);
                if (shear) vely = vely + vshear;
                if (s->ny == 1) vely = zero;
                double velz = comp_speed(g, rho, pr, vz, bz, bx, by);
                if (s->nz == 1) velz = zero;
                double c = one / ((velx / s->dx + vely / s->dy) + velz / s->dz);
                dt = fmn(dt, c);
            }
    return p->courant * dt;
}

/* fc_magnetic, scheme.f90:193-249 */
static void fc_magnetic(dmxo* s)
{
    int iu1 = -2, iu2 = s->nx + 3, ju1 = -2, ju2 = s->ny + 3, ku1 = -2, ku2 = s->nz + 3;
#pragma omp parallel for schedule(static)
    for (int k = ku1; k <= ku2 + 1; k++)
        for (int j = ju1; j <= ju2 + 1; j++)
            for (int i = iu1; i <= iu2 + 1; i++) {
                if (k <= ku2 && j <= ju2)
                    BFC(i, j, k, 0) = (i <= iu2) ? U(i, j, k, IA) : U(i - 1, j, k, IA + 3);
                if (k <= ku2 && i <= iu2)
                    BFC(i, j, k, 1) = (j <= ju2) ? U(i, j, k, IB) : U(i, j - 1, k, IB + 3);
                if (j <= ju2 && i <= iu2)
                    BFC(i, j, k, 2) = (k <= ku2) ? U(i, j, k, IC) : U(i, j, k - 1, IC + 3);
            }
}

/* TVD slope of slope(), scheme.f90:452-477 (st = 1 minmod, 2 MC) */
static inline double tvd_slope(double qm1, double q0, double qp1, int st)
{
    double dlft = (double)st * (q0 - qm1);
    double drgt = (double)st * (qp1 - q0);
    double dcen = half * (dlft + drgt) / (double)st;
    double dsgn = sgn1(dcen);
    double slop = fmn(fabs(dlft), fabs(drgt));
    double dlim = slop;
    if ((dlft * drgt) <= zero) dlim = zero;
    return dsgn * fmn(dlim, fabs(dcen));
}

/* slope, scheme.f90:253-656 (3D branch :440-653) */
static void slope(dmxo* s)
{
    int st = s->p.slope_type;
    int nx = s->nx, ny = s->ny, nz = s->nz;
    int ilo = -1, ihi = nx + 2, jlo = -1, jhi = ny + 2, klo = -1, khi = nz + 2;
    if (st == 0) {
        memset(s->dq, 0, sizeof(double) * 24 * s->n3);
        memset(s->dbfc, 0, sizeof(double) * 9 * s->n3);
        return;
    }
    int stl = st < 2 ? st : 2;
#pragma omp parallel for schedule(static)
    for (int k = klo; k <= khi; k++)
        for (int j = jlo; j <= jhi; j++)
            for (int i = ilo; i <= ihi; i++) {
                for (int v = 0; v < 8; v++) {
                    if (st < 3) {
                        DQ(i, j, k, v, 0) = tvd_slope(Q(i - 1, j, k, v), Q(i, j, k, v), Q(i + 1, j, k, v), st);
                        DQ(i, j, k, v, 1) = tvd_slope(Q(i, j - 1, k, v), Q(i, j, k, v), Q(i, j + 1, k, v), st);
                        DQ(i, j, k, v, 2) = tvd_slope(Q(i, j, k - 1, v), Q(i, j, k, v), Q(i, j, k + 1, v), st);
                    } else {
                        /* multidimensional limiter, scheme.f90:483-555: min/max of
                         * the 27 differences, in the order of the reference
                         * (i fastest within each k-plane, planes k-1, k, k+1) */
                        double q0 = Q(i, j, k, v);
                        double vmin = 0, vmax = 0;
                        int first = 1;
                        for (int dk = -1; dk <= 1; dk++)
                            for (int di = -1; di <= 1; di++)
                                for (int dj = -1; dj <= 1; dj++) {
                                    double d = Q(i + di, j + dj, k + dk, v) - q0;
                                    if (first) { vmin = d; vmax = d; first = 0; }
                                    else { vmin = fmn(vmin, d); vmax = fmx(vmax, d); }
                                }
                        double dfx = half * (Q(i + 1, j, k, v) - Q(i - 1, j, k, v));
                        double dfy = half * (Q(i, j + 1, k, v) - Q(i, j - 1, k, v));
                        double dfz = half * (Q(i, j, k + 1, v) - Q(i, j, k - 1, v));
                        double dff = half * ((fabs(dfx) + fabs(dfy)) + fabs(dfz));
                        double slop = (dff > zero) ? fmn(one, fmn(fabs(vmin), fabs(vmax)) / dff) : one;
                        DQ(i, j, k, v, 0) = slop * dfx;
                        DQ(i, j, k, v, 1) = slop * dfy;
                        DQ(i, j, k, v, 2) = slop * dfz;
                    }
                }
            }
    /* face-B transverse slopes, scheme.f90:561-652 */
#pragma omp parallel for schedule(static)
    for (int k = klo; k <= khi + 1; k++)
        for (int j = jlo; j <= jhi + 1; j++)
            for (int i = ilo; i <= ihi + 1; i++) {
                if (k <= khi && j <= jhi) { /* Bx along y and z */
                    DBFC(i, j, k, 0, 1) = tvd_slope(BFC(i, j - 1, k, 0), BFC(i, j, k, 0), BFC(i, j + 1, k, 0), stl);
                    DBFC(i, j, k, 0, 2) = tvd_slope(BFC(i, j, k - 1, 0), BFC(i, j, k, 0), BFC(i, j, k + 1, 0), stl);
                }
                if (k <= khi && i <= ihi) { /* By along x and z */
                    DBFC(i, j, k, 1, 0) = tvd_slope(BFC(i - 1, j, k, 1), BFC(i, j, k, 1), BFC(i + 1, j, k, 1), stl);
                    DBFC(i, j, k, 1, 2) = tvd_slope(BFC(i, j, k - 1, 1), BFC(i, j, k, 1), BFC(i, j, k + 1, 1), stl);
                }
                if (j <= jhi && i <= ihi) { /* Bz along x and y */
                    DBFC(i, j, k, 2, 0) = tvd_slope(BFC(i - 1, j, k, 2), BFC(i, j, k, 2), BFC(i + 1, j, k, 2), stl);
                    DBFC(i, j, k, 2, 1) = tvd_slope(BFC(i, j - 1, k, 2), BFC(i, j, k, 2), BFC(i, j + 1, k, 2), stl);
                }
            }
}

/* trace3d, scheme.f90:1149-1711 (Cartesian) */
static void trace3d(dmxo* s, double dt)
{
    const dmx_params* p = &s->p;
    int nx = s->nx, ny = s->ny, nz = s->nz;
    int ilo = -1, ihi = nx + 2, jlo = -1, jhi = ny + 2, klo = -1, khi = nz + 2;
    int iu2 = nx + 3, ju2 = ny + 3, ku2 = nz + 3;
    double gamma = p->gamma, Omega0 = p->omega0;
    double smallp = smallr * (smallc * smallc) / gamma;
    double dtdx = dt / s->dx, dtdy = dt / s->dy, dtdz = dt / s->dz, dx = s->dx;
    int shearon = (Omega0 > zero) && !p->fargo;

    /* edge electric fields, scheme.f90:1194-1239 */
#pragma omp parallel for schedule(static)
    for (int k = klo; k <= ku2; k++)
        for (int j = jlo; j <= ju2; j++)
            for (int i = ilo; i <= iu2; i++) {
                double v = forth * (((Q(i, j - 1, k - 1, IV) + Q(i, j - 1, k, IV)) + Q(i, j, k - 1, IV)) + Q(i, j, k, IV));
                double w = forth * (((Q(i, j - 1, k - 1, IW) + Q(i, j - 1, k, IW)) + Q(i, j, k - 1, IW)) + Q(i, j, k, IW));
                double B = half * (BFC(i, j, k - 1, 1) + BFC(i, j, k, 1));
                double C = half * (BFC(i, j - 1, k, 2) + BFC(i, j, k, 2));
                double e = v * C - w * B;
                if (shearon) e = e + (-1.5 * Omega0 * XC(i)) * C;
                EX(i, j, k) = e;
                double u = forth * (((Q(i - 1, j, k - 1, IU) + Q(i - 1, j, k, IU)) + Q(i, j, k - 1, IU)) + Q(i, j, k, IU));
                w = forth * (((Q(i - 1, j, k - 1, IW) + Q(i - 1, j, k, IW)) + Q(i, j, k - 1, IW)) + Q(i, j, k, IW));
                double A = half * (BFC(i, j, k - 1, 0) + BFC(i, j, k, 0));
                C = half * (BFC(i - 1, j, k, 2) + BFC(i, j, k, 2));
                EY(i, j, k) = w * A - u * C;
                u = forth * (((Q(i - 1, j - 1, k, IU) + Q(i - 1, j, k, IU)) + Q(i, j - 1, k, IU)) + Q(i, j, k, IU));
                v = forth * (((Q(i - 1, j - 1, k, IV) + Q(i - 1, j, k, IV)) + Q(i, j - 1, k, IV)) + Q(i, j, k, IV));
                A = half * (BFC(i, j - 1, k, 0) + BFC(i, j, k, 0));
                B = half * (BFC(i - 1, j, k, 1) + BFC(i, j, k, 1));
                e = u * B - v * A;
                if (shearon) e = e - (-1.5 * Omega0 * half * (XC(i) + XC(i - 1))) * A;
                EZ(i, j, k) = e;
            }

    /* predictor and face/edge states, scheme.f90:1241-1707 */
#pragma omp parallel for schedule(static)
    for (int k = klo; k <= khi; k++)
        for (int j = jlo; j <= jhi; j++)
            for (int i = ilo; i <= ihi; i++) {
                double r = Q(i, j, k, IR), u = Q(i, j, k, IU), v = Q(i, j, k, IV),
                       w = Q(i, j, k, IW), pr = Q(i, j, k, IP), A = Q(i, j, k, IA),
                       B = Q(i, j, k, IB), C = Q(i, j, k, IC);
                double drx = half * DQ(i, j, k, IR, 0), dux = half * DQ(i, j, k, IU, 0),
                       dvx = half * DQ(i, j, k, IV, 0), dwx = half * DQ(i, j, k, IW, 0),
                       dpx = half * DQ(i, j, k, IP, 0), dBx = half * DQ(i, j, k, IB, 0),
                       dCx = half * DQ(i, j, k, IC, 0);
                double dry = half * DQ(i, j, k, IR, 1), duy = half * DQ(i, j, k, IU, 1),
                       dvy = half * DQ(i, j, k, IV, 1), dwy = half * DQ(i, j, k, IW, 1),
                       dpy = half * DQ(i, j, k, IP, 1), dAy = half * DQ(i, j, k, IA, 1),
                       dCy = half * DQ(i, j, k, IC, 1);
                double drz = half * DQ(i, j, k, IR, 2), duz = half * DQ(i, j, k, IU, 2),
                       dvz = half * DQ(i, j, k, IV, 2), dwz = half * DQ(i, j, k, IW, 2),
                       dpz = half * DQ(i, j, k, IP, 2), dAz = half * DQ(i, j, k, IA, 2),
                       dBz = half * DQ(i, j, k, IB, 2);
                double AL = BFC(i, j, k, 0), AR = BFC(i + 1, j, k, 0);
                double BL = BFC(i, j, k, 1), BR = BFC(i, j + 1, k, 1);
                double CL = BFC(i, j, k, 2), CR = BFC(i, j, k + 1, 2);
                double dAx = half * (AR - AL), dBy = half * (BR - BL), dCz = half * (CR - CL);
                double dALy = half * DBFC(i, j, k, 0, 1), dARy = half * DBFC(i + 1, j, k, 0, 1);
                double dALz = half * DBFC(i, j, k, 0, 2), dARz = half * DBFC(i + 1, j, k, 0, 2);
                double dBLx = half * DBFC(i, j, k, 1, 0), dBRx = half * DBFC(i, j + 1, k, 1, 0);
                double dBLz = half * DBFC(i, j, k, 1, 2), dBRz = half * DBFC(i, j + 1, k, 1, 2);
                double dCLx = half * DBFC(i, j, k, 2, 0), dCRx = half * DBFC(i, j, k + 1, 2, 0);
                double dCLy = half * DBFC(i, j, k, 2, 1), dCRy = half * DBFC(i, j, k + 1, 2, 1);
                double ELL = EX(i, j, k), ELR = EX(i, j, k + 1), ERL = EX(i, j + 1, k), ERR = EX(i, j + 1, k + 1);
                double FLL = EY(i, j, k), FLR = EY(i, j, k + 1), FRL = EY(i + 1, j, k), FRR = EY(i + 1, j, k + 1);
                double GLL = EZ(i, j, k), GLR = EZ(i, j + 1, k), GRL = EZ(i + 1, j, k), GRR = EZ(i + 1, j + 1, k);

                double sr0 = ((-u * drx - dux * r) * dtdx + (-v * dry - dvy * r) * dtdy) + (-w * drz - dwz * r) * dtdz;
                double su0 = ((-u * dux - ((dpx + B * dBx) + C * dCx) / r) * dtdx + (-v * duy + B * dAy / r) * dtdy)
                           + (-w * duz + C * dAz / r) * dtdz;
                double sv0 = ((-u * dvx + A * dBx / r) * dtdx + (-v * dvy - ((dpy + A * dAy) + C * dCy) / r) * dtdy)
                           + (-w * dvz + C * dBz / r) * dtdz;
                double sw0 = ((-u * dwx + A * dCx / r) * dtdx + (-v * dwy + B * dCy / r) * dtdy)
                           + (-w * dwz - ((dpz + A * dAz) + B * dBz) / r) * dtdz;
                double sp0 = ((-u * dpx - dux * gamma * pr) * dtdx + (-v * dpy - dvy * gamma * pr) * dtdy)
                           + (-w * dpz - dwz * gamma * pr) * dtdz;
                double sA0 = (((u * dBy + B * duy) - v * dAy) - A * dvy) * dtdy + (((u * dCz + C * duz) - w * dAz) - A * dwz) * dtdz;
                double sB0 = (((v * dAx + A * dvx) - u * dBx) - B * dux) * dtdx + (((v * dCz + C * dvz) - w * dBz) - B * dwz) * dtdz;
                double sC0 = (((w * dAx + A * dwx) - u * dCx) - C * dux) * dtdx + (((w * dBy + B * dwy) - v * dCy) - C * dvy) * dtdy;
                if (shearon) {
                    double shear = -1.5 * Omega0 * XC(i);
                    sr0 = sr0 - shear * dry * dtdy;
                    su0 = su0 - shear * duy * dtdy;
                    sv0 = sv0 - shear * dvy * dtdy;
                    sw0 = sw0 - shear * dwy * dtdy;
                    sp0 = sp0 - shear * dpy * dtdy;
                    sA0 = sA0 - shear * dAy * dtdy;
                    sB0 = (sB0 + (shear * dAx - 1.5 * Omega0 * A * dx) * dtdx) + shear * dBz * dtdz;
                    sC0 = sC0 - shear * dCy * dtdy;
                }
                double sAL0 = (GLR - GLL) * dtdy * half - (FLR - FLL) * dtdz * half;
                double sAR0 = (GRR - GRL) * dtdy * half - (FRR - FRL) * dtdz * half;
                double sBL0 = -(GRL - GLL) * dtdx * half + (ELR - ELL) * dtdz * half;
                double sBR0 = -(GRR - GLR) * dtdx * half + (ERR - ERL) * dtdz * half;
                double sCL0 = (FRL - FLL) * dtdx * half - (ERL - ELL) * dtdy * half;
                double sCR0 = (FRR - FLR) * dtdx * half - (ERR - ELR) * dtdy * half;

                r = r + sr0; u = u + su0; v = v + sv0; w = w + sw0; pr = pr + sp0;
                A = A + sA0; B = B + sB0; C = C + sC0;
                AL = AL + sAL0; AR = AR + sAR0; BL = BL + sBL0; BR = BR + sBR0;
                CL = CL + sCL0; CR = CR + sCR0;

#define SETQ(ARR, d, vr, vu, vv, vw, vp, vA, vB, vC)              \
    do {                                                          \
        ARR(i, j, k, IR, d) = fmx(smallr, (vr));                  \
        ARR(i, j, k, IU, d) = (vu);                               \
        ARR(i, j, k, IV, d) = (vv);                               \
        ARR(i, j, k, IW, d) = (vw);                               \
        ARR(i, j, k, IP, d) = fmx(smallp, (vp));                  \
        ARR(i, j, k, IA, d) = (vA);                               \
        ARR(i, j, k, IB, d) = (vB);                               \
        ARR(i, j, k, IC, d) = (vC);                               \
    } while (0)
                /* face states, scheme.f90:1489-1559 */
                SETQ(QP, 0, r - drx, u - dux, v - dvx, w - dwx, pr - dpx, AL, B - dBx, C - dCx);
                SETQ(QM, 0, r + drx, u + dux, v + dvx, w + dwx, pr + dpx, AR, B + dBx, C + dCx);
                SETQ(QP, 1, r - dry, u - duy, v - dvy, w - dwy, pr - dpy, A - dAy, BL, C - dCy);
                SETQ(QM, 1, r + dry, u + duy, v + dvy, w + dwy, pr + dpy, A + dAy, BR, C + dCy);
                SETQ(QP, 2, r - drz, u - duz, v - dvz, w - dwz, pr - dpz, A - dAz, B - dBz, CL);
                SETQ(QM, 2, r + drz, u + duz, v + dvz, w + dwz, pr + dpz, A + dAz, B + dBz, CR);
                /* x-edge corner states, scheme.f90:1561-1607 */
                SETQ(QRT, 0, r + (dry + drz), u + (duy + duz), v + (dvy + dvz), w + (dwy + dwz), pr + (dpy + dpz), A + (dAy + dAz), BR + dBRz, CR + dCRy);
                SETQ(QRB, 0, r + (dry - drz), u + (duy - duz), v + (dvy - dvz), w + (dwy - dwz), pr + (dpy - dpz), A + (dAy - dAz), BR - dBRz, CL + dCLy);
                SETQ(QLT, 0, r + (-dry + drz), u + (-duy + duz), v + (-dvy + dvz), w + (-dwy + dwz), pr + (-dpy + dpz), A + (-dAy + dAz), BL + dBLz, CR - dCRy);
                SETQ(QLB, 0, r + (-dry - drz), u + (-duy - duz), v + (-dvy - dvz), w + (-dwy - dwz), pr + (-dpy - dpz), A + (-dAy - dAz), BL - dBLz, CL - dCLy);
                /* y-edge corner states, scheme.f90:1609-1655 */
                SETQ(QRT, 1, r + (drx + drz), u + (dux + duz), v + (dvx + dvz), w + (dwx + dwz), pr + (dpx + dpz), AR + dARz, B + (dBx + dBz), CR + dCRx);
                SETQ(QRB, 1, r + (drx - drz), u + (dux - duz), v + (dvx - dvz), w + (dwx - dwz), pr + (dpx - dpz), AR - dARz, B + (dBx - dBz), CL + dCLx);
                SETQ(QLT, 1, r + (-drx + drz), u + (-dux + duz), v + (-dvx + dvz), w + (-dwx + dwz), pr + (-dpx + dpz), AL + dALz, B + (-dBx + dBz), CR - dCRx);
                SETQ(QLB, 1, r + (-drx - drz), u + (-dux - duz), v + (-dvx - dvz), w + (-dwx - dwz), pr + (-dpx - dpz), AL - dALz, B + (-dBx - dBz), CL - dCLx);
                /* z-edge corner states, scheme.f90:1657-1703 */
                SETQ(QRT, 2, r + (drx + dry), u + (dux + duy), v + (dvx + dvy), w + (dwx + dwy), pr + (dpx + dpy), AR + dARy, BR + dBRx, C + (dCx + dCy));
                SETQ(QRB, 2, r + (drx - dry), u + (dux - duy), v + (dvx - dvy), w + (dwx - dwy), pr + (dpx - dpy), AR - dARy, BL + dBLx, C + (dCx - dCy));
                SETQ(QLT, 2, r + (-drx + dry), u + (-dux + duy), v + (-dvx + dvy), w + (-dwx + dwy), pr + (-dpx + dpy), AL + dALy, BR - dBRx, C + (-dCx + dCy));
                SETQ(QLB, 2, r + (-drx - dry), u + (-dux - duy), v + (-dvx - dvy), w + (-dwx - dwy), pr + (-dpx - dpy), AL - dALy, BL - dBLx, C + (-dCx - dCy));
#undef SETQ
            }
}

/* ------------------------------------------------------------------------ */
/* 1D Riemann solvers, solver.f90:393-935.                                   */
/* q = [r, p, u, a, v, b, w, c] (solver_utils, solver.f90:31-57)             */
/* ------------------------------------------------------------------------ */
static void phys_flux(double r, double ploc, double u, double v, double w,
                      double a, double b, double c, double entho,
                      double* f, double* cons)
{
    /* left/right flux and conservative state shared by llf, hll, upwind
     * (solver.f90:430-452 and :541-563) */
    double ecin = half * ((u * u + v * v) + w * w) * r;
    double emag = half * ((a * a + b * b) + c * c);
    double etot = (ploc * entho + ecin) + emag;
    double ptot = ploc + emag;
    f[0] = r * u;
    f[1] = (etot + ptot) * u - a * ((u * a + v * b) + w * c);
    f[2] = (r * u * u + ptot) - a * a;
    f[3] = zero;
    f[4] = r * u * v - a * b;
    f[5] = b * u - a * v;
    f[6] = r * u * w - a * c;
    f[7] = c * u - a * w;
    cons[0] = r; cons[1] = etot; cons[2] = r * u; cons[3] = a;
    cons[4] = r * v; cons[5] = b; cons[6] = r * w; cons[7] = c;
}

static inline double cfast(double gamma, double ploc, double r, double a,
                           double b, double c)
{
    double c2 = gamma * ploc / r;
    double b2 = (a * a + b * b) + c * c;
    double d2 = half * (c2 + b2 / r);
    return sqrt(d2 + sqrt(d2 * d2 - c2 * a * a / r));
}

static void hlld1d(const double* ql, const double* qr, double gamma,
                   double ploc, double proc, double* fg, double* st)
{
    /* hlld, solver.f90:618-836 */
    double rl = ql[0], ul = ql[2], al = ql[3], vl = ql[4], bl = ql[5], wl = ql[6], cl = ql[7];
    double rr = qr[0], ur = qr[2], ar = qr[3], vr = qr[4], br = qr[5], wr = qr[6], cr = qr[7];
    double entho = one / (gamma - one);
    double a = half * (al + ar);
    double sgnm = sgn1(a);
    double ecinl = half * ((ul * ul + vl * vl) + wl * wl) * rl;
    double emagl = half * ((a * a + bl * bl) + cl * cl);
    double etotl = (ploc * entho + ecinl) + emagl;
    double ptotl = ploc + emagl;
    double vdotbl = (ul * a + vl * bl) + wl * cl;
    double ecinr = half * ((ur * ur + vr * vr) + wr * wr) * rr;
    double emagr = half * ((a * a + br * br) + cr * cr);
    double etotr = (proc * entho + ecinr) + emagr;
    double ptotr = proc + emagr;
    double vdotbr = (ur * a + vr * br) + wr * cr;
    double cfastl = cfast(gamma, ploc, rl, a, bl, cl);
    double cfastr = cfast(gamma, proc, rr, a, br, cr);
    double sl = fmn(ul, ur) - fmx(cfastl, cfastr);
    double sr = fmx(ul, ur) + fmx(cfastl, cfastr);
    double rcl = rl * (ul - sl);
    double rcr = rr * (sr - ur);
    double ustar = ((rcr * ur + rcl * ul) + (ptotl - ptotr)) / (rcr + rcl);
    double ptotstar = ((rcr * ptotl + rcl * ptotr) + rcl * rcr * (ul - ur)) / (rcr + rcl);
    double vstarl, bstarl, wstarl, cstarl, vstarr, bstarr, wstarr, cstarr;
    double rstarl = rl * (sl - ul) / (sl - ustar);
    double estar = rl * (sl - ul) * (sl - ustar) - a * a;
    double el = rl * (sl - ul) * (sl - ul) - a * a;
    if (estar == zero) {
        vstarl = vl; bstarl = bl; wstarl = wl; cstarl = cl;
    } else {
        vstarl = vl - a * bl * (ustar - ul) / estar;
        bstarl = bl * el / estar;
        wstarl = wl - a * cl * (ustar - ul) / estar;
        cstarl = cl * el / estar;
    }
    double vdotbstarl = (ustar * a + vstarl * bstarl) + wstarl * cstarl;
    double etotstarl = ((((sl - ul) * etotl - ptotl * ul) + ptotstar * ustar) + a * (vdotbl - vdotbstarl)) / (sl - ustar);
    double sqrrstarl = sqrt(rstarl);
    double calfvenl = fabs(a) / sqrrstarl;
    double sal = ustar - calfvenl;
    double rstarr = rr * (sr - ur) / (sr - ustar);
    estar = rr * (sr - ur) * (sr - ustar) - a * a;
    double er = rr * (sr - ur) * (sr - ur) - a * a;
    if (estar == zero) {
        vstarr = vr; bstarr = br; wstarr = wr; cstarr = cr;
    } else {
        vstarr = vr - a * br * (ustar - ur) / estar;
        bstarr = br * er / estar;
        wstarr = wr - a * cr * (ustar - ur) / estar;
        cstarr = cr * er / estar;
    }
    double vdotbstarr = (ustar * a + vstarr * bstarr) + wstarr * cstarr;
    double etotstarr = ((((sr - ur) * etotr - ptotr * ur) + ptotstar * ustar) + a * (vdotbr - vdotbstarr)) / (sr - ustar);
    double sqrrstarr = sqrt(rstarr);
    double calfvenr = fabs(a) / sqrrstarr;
    double sar = ustar + calfvenr;
    double den = sqrrstarl + sqrrstarr;
    double vss = ((sqrrstarl * vstarl + sqrrstarr * vstarr) + sgnm * (bstarr - bstarl)) / den;
    double wss = ((sqrrstarl * wstarl + sqrrstarr * wstarr) + sgnm * (cstarr - cstarl)) / den;
    double bss = ((sqrrstarl * bstarr + sqrrstarr * bstarl) + sgnm * sqrrstarl * sqrrstarr * (vstarr - vstarl)) / den;
    double css = ((sqrrstarl * cstarr + sqrrstarr * cstarl) + sgnm * sqrrstarl * sqrrstarr * (wstarr - wstarl)) / den;
    double vdotbss = (ustar * a + vss * bss) + wss * css;
    double etotssl = etotstarl - sgnm * sqrrstarl * (vdotbstarl - vdotbss);
    double etotssr = etotstarr + sgnm * sqrrstarr * (vdotbstarr - vdotbss);
    double ro, uo, vo, wo, bo, co, ptoto, etoto, vdotbo;
    if (sl > zero) {
        ro = rl; uo = ul; vo = vl; wo = wl; bo = bl; co = cl; ptoto = ptotl; etoto = etotl; vdotbo = vdotbl;
    } else if (sal > zero) {
        ro = rstarl; uo = ustar; vo = vstarl; wo = wstarl; bo = bstarl; co = cstarl; ptoto = ptotstar; etoto = etotstarl; vdotbo = vdotbstarl;
    } else if (ustar > zero) {
        ro = rstarl; uo = ustar; vo = vss; wo = wss; bo = bss; co = css; ptoto = ptotstar; etoto = etotssl; vdotbo = vdotbss;
    } else if (sar > zero) {
        ro = rstarr; uo = ustar; vo = vss; wo = wss; bo = bss; co = css; ptoto = ptotstar; etoto = etotssr; vdotbo = vdotbss;
    } else if (sr > zero) {
        ro = rstarr; uo = ustar; vo = vstarr; wo = wstarr; bo = bstarr; co = cstarr; ptoto = ptotstar; etoto = etotstarr; vdotbo = vdotbstarr;
    } else {
        ro = rr; uo = ur; vo = vr; wo = wr; bo = br; co = cr; ptoto = ptotr; etoto = etotr; vdotbo = vdotbr;
    }
    fg[0] = ro * uo;
    fg[1] = (etoto + ptoto) * uo - a * vdotbo;
    fg[2] = ro * uo * uo - a * a;
    fg[3] = zero;
    fg[4] = ro * uo * vo - a * bo;
    fg[5] = bo * uo - a * vo;
    fg[6] = ro * uo * wo - a * co;
    fg[7] = co * uo - a * wo;
    st[0] = ro; st[1] = uo; st[2] = vo; st[3] = wo; st[4] = bo; st[5] = co; st[6] = ptoto;
}

static void simple1d(int solver, const double* ql, const double* qr,
                     double gamma, double ploc, double proc, double* fg)
{
    /* llf (solver.f90:393-498), hll (:504-612), upwind (:842-935) */
    double rl = ql[0], ul = ql[2], al = ql[3], vl = ql[4], bl = ql[5], wl = ql[6], cl = ql[7];
    double rr = qr[0], ur = qr[2], ar = qr[3], vr = qr[4], br = qr[5], wr = qr[6], cr = qr[7];
    double entho = one / (gamma - one);
    double a = half * (al + ar);
    double fl[8], cL[8], fr[8], cR[8];
    phys_flux(rl, ploc, ul, vl, wl, a, bl, cl, entho, fl, cL);
    phys_flux(rr, proc, ur, vr, wr, a, br, cr, entho, fr, cR);
    if (solver == DMX_HLL) {
        double vleft = cfast(gamma, ploc, rl, a, bl, cl);
        double vright = cfast(gamma, proc, rr, a, br, cr);
        double sl = fmn(fmn(ul, ur) - fmx(vleft, vright), zero);
        double sr = fmx(fmx(ul, ur) + fmx(vleft, vright), zero);
        for (int n = 0; n < 8; n++)
            fg[n] = ((sr * fl[n] - sl * fr[n]) + sr * sl * (cR[n] - cL[n])) / (sr - sl);
    } else {
        double vel;
        if (solver == DMX_LLF) {
            double vleft = cfast(gamma, ploc, rl, a, bl, cl) + fabs(ul);
            double vright = cfast(gamma, proc, rr, a, br, cr) + fabs(ur);
            vel = fmx(vleft, vright);
        } else {
            vel = fabs(half * (ul + ur));
        }
        for (int n = 0; n < 8; n++)
            fg[n] = half * (fl[n] + fr[n]) - vel * half * (cR[n] - cL[n]);
    }
}

void dmxo_riemann1d(int solver, const double* ql, const double* qr,
                    double gamma, double ciso, int iso, double* flux,
                    double* star)
{
    double ploc = iso ? ql[0] * (ciso * ciso) : ql[1];
    double proc = iso ? qr[0] * (ciso * ciso) : qr[1];
    for (int n = 0; n < 7; n++) star[n] = zero;
    if (solver == DMX_HLLD) hlld1d(ql, qr, gamma, ploc, proc, flux, star);
    else simple1d(solver, ql, qr, gamma, ploc, proc, flux);
}

/* ------------------------------------------------------------------------ */
/* 2D (edge) Riemann solvers, solver_magnetic.f90:241-727.                   */
/* corner q = [r, p, u, v, A, B, C, E] (solver_magnetic_utils, :32-49)       */
/* ------------------------------------------------------------------------ */
static inline double cf2(double gamma, const double* q, int nrm)
{
    /* fast speed along x (nrm=0: normal field A) or y (nrm=1: B),
     * solver_magnetic.f90:391-431 */
    double r = q[0], pr = q[1], A = q[4], B = q[5], C = q[6];
    double c2 = gamma * pr / r;
    double b2 = nrm == 0 ? (A * A + B * B) + C * C : (B * B + A * A) + C * C;
    double bn = nrm == 0 ? A : B;
    double d2 = half * (c2 + b2 / r);
    return sqrt(d2 + sqrt(d2 * d2 - c2 * bn * bn / r));
}

static double hlld2d(const double* LL, const double* RL, const double* LR,
                     const double* RR, double dt, double gamma)
{
    /* hlld, solver_magnetic.f90:362-548 */
    double rLL = LL[0], pLL = LL[1], uLL = LL[2], vLL = LL[3], ALL = LL[4], BLL = LL[5], CLL = LL[6], ELL = LL[7];
    double rRL = RL[0], pRL = RL[1], uRL = RL[2], vRL = RL[3], ARL = RL[4], BRL = RL[5], CRL = RL[6], ERL = RL[7];
    double rLR = LR[0], pLR = LR[1], uLR = LR[2], vLR = LR[3], ALR = LR[4], BLR = LR[5], CLR = LR[6], ELR = LR[7];
    double rRR = RR[0], pRR = RR[1], uRR = RR[2], vRR = RR[3], ARR = RR[4], BRR = RR[5], CRR = RR[6], ERR = RR[7];
    double cfastLLx = cf2(gamma, LL, 0), cfastLRx = cf2(gamma, LR, 0);
    double cfastRLx = cf2(gamma, RL, 0), cfastRRx = cf2(gamma, RR, 0);
    double cfastLLy = cf2(gamma, LL, 1), cfastLRy = cf2(gamma, LR, 1);
    double cfastRLy = cf2(gamma, RL, 1), cfastRRy = cf2(gamma, RR, 1);
    double SL = fmn(fmn(fmn(uLL, uLR), uRL), uRR) - fmx(fmx(fmx(cfastLLx, cfastLRx), cfastRLx), cfastRRx);
    double SR = fmx(fmx(fmx(uLL, uLR), uRL), uRR) + fmx(fmx(fmx(cfastLLx, cfastLRx), cfastRLx), cfastRRx);
    double SB = fmn(fmn(fmn(vLL, vLR), vRL), vRR) - fmx(fmx(fmx(cfastLLy, cfastLRy), cfastRLy), cfastRRy);
    double ST = fmx(fmx(fmx(vLL, vLR), vRL), vRR) + fmx(fmx(fmx(cfastLLy, cfastLRy), cfastRLy), cfastRRy);
    double PtotLL = pLL + half * ((ALL * ALL + BLL * BLL) + CLL * CLL);
    double PtotLR = pLR + half * ((ALR * ALR + BLR * BLR) + CLR * CLR);
    double PtotRL = pRL + half * ((ARL * ARL + BRL * BRL) + CRL * CRL);
    double PtotRR = pRR + half * ((ARR * ARR + BRR * BRR) + CRR * CRR);
    double rcLLx = rLL * (uLL - SL), rcRLx = rRL * (SR - uRL);
    double rcLRx = rLR * (uLR - SL), rcRRx = rRR * (SR - uRR);
    double rcLLy = rLL * (vLL - SB), rcLRy = rLR * (ST - vLR);
    double rcRLy = rRL * (vRL - SB), rcRRy = rRR * (ST - vRR);
    double ustar = ((((rcLLx * uLL + rcLRx * uLR) + rcRLx * uRL) + rcRRx * uRR)
                    + (((PtotLL - PtotRL) + PtotLR) - PtotRR)) / (((rcLLx + rcLRx) + rcRLx) + rcRRx);
    double vstar = ((((rcLLy * vLL + rcLRy * vLR) + rcRLy * vRL) + rcRRy * vRR)
                    + (((PtotLL - PtotLR) + PtotRL) - PtotRR)) / (((rcLLy + rcLRy) + rcRLy) + rcRRy);
    double SLustar = one / (SL - ustar), SRustar = one / (SR - ustar);
    double SBvstar = one / (SB - vstar), STvstar = one / (ST - vstar);

    double rstarLLx = rLL * (SL - uLL) * SLustar, BstarLL = BLL * (SL - uLL) * SLustar;
    double rstarLLy = rLL * (SB - vLL) * SBvstar, AstarLL = ALL * (SB - vLL) * SBvstar;
    double rstarLL = rLL * (SL - uLL) * SLustar * (SB - vLL) * SBvstar;
    double EstarLLx = ustar * BstarLL - vLL * ALL, EstarLLy = uLL * BLL - vstar * AstarLL;
    double EstarLL = ustar * BstarLL - vstar * AstarLL;

    double rstarLRx = rLR * (SL - uLR) * SLustar, BstarLR = BLR * (SL - uLR) * SLustar;
    double rstarLRy = rLR * (ST - vLR) * STvstar, AstarLR = ALR * (ST - vLR) * STvstar;
    double rstarLR = rLR * (SL - uLR) * SLustar * (ST - vLR) * STvstar;
    double EstarLRx = ustar * BstarLR - vLR * ALR, EstarLRy = uLR * BLR - vstar * AstarLR;
    double EstarLR = ustar * BstarLR - vstar * AstarLR;

    double rstarRLx = rRL * (SR - uRL) * SRustar, BstarRL = BRL * (SR - uRL) * SRustar;
    double rstarRLy = rRL * (SB - vRL) * SBvstar, AstarRL = ARL * (SB - vRL) * SBvstar;
    double rstarRL = rRL * (SR - uRL) * SRustar * (SB - vRL) * SBvstar;
    double EstarRLx = ustar * BstarRL - vRL * ARL, EstarRLy = uRL * BRL - vstar * AstarRL;
    double EstarRL = ustar * BstarRL - vstar * AstarRL;

    double rstarRRx = rRR * (SR - uRR) * SRustar, BstarRR = BRR * (SR - uRR) * SRustar;
    double rstarRRy = rRR * (ST - vRR) * STvstar, AstarRR = ARR * (ST - vRR) * STvstar;
    double rstarRR = rRR * (SR - uRR) * SRustar * (ST - vRR) * STvstar;
    double EstarRRx = ustar * BstarRR - vRR * ARR, EstarRRy = uRR * BRR - vstar * AstarRR;
    double EstarRR = ustar * BstarRR - vstar * AstarRR;

    double sqrtrstarLR = one / sqrt(rstarLR), sqrtrstarRL = one / sqrt(rstarRL);
    double sqrtrstarRR = one / sqrt(rstarRR), sqrtrstarLL = one / sqrt(rstarLL);

    double calfvenL = fmx(fmx(fmx(fmx(fabs(ALR) / sqrt(rstarLRx), fabs(AstarLR) * sqrtrstarLR),
                                  fabs(ALL) / sqrt(rstarLLx)), fabs(AstarLL) * sqrtrstarLL), smallc);
    double calfvenR = fmx(fmx(fmx(fmx(fabs(ARR) / sqrt(rstarRRx), fabs(AstarRR) * sqrtrstarRR),
                                  fabs(ARL) / sqrt(rstarRLx)), fabs(AstarRL) * sqrtrstarRL), smallc);
    double calfvenB = fmx(fmx(fmx(fmx(fabs(BLL) / sqrt(rstarLLy), fabs(BstarLL) * sqrtrstarLL),
                                  fabs(BRL) / sqrt(rstarRLy)), fabs(BstarRL) * sqrtrstarRL), smallc);
    double calfvenT = fmx(fmx(fmx(fmx(fabs(BLR) / sqrt(rstarLRy), fabs(BstarLR) * sqrtrstarLR),
                                  fabs(BRR) / sqrt(rstarRRy)), fabs(BstarRR) * sqrtrstarRR), smallc);
    double SAL = fmn(ustar - calfvenL, zero), SAR = fmx(ustar + calfvenR, zero);
    double SAB = fmn(vstar - calfvenB, zero), SAT = fmx(vstar + calfvenT, zero);
    double SARSAL = one / (SAR - SAL);
    double SATSAB = one / (SAT - SAB);
    double AstarT = (SAR * AstarRR - SAL * AstarLR) * SARSAL;
    double AstarB = (SAR * AstarRL - SAL * AstarLL) * SARSAL;
    double BstarR = (SAT * BstarRR - SAB * BstarRL) * SATSAB;
    double BstarL = (SAT * BstarLR - SAB * BstarLL) * SATSAB;
    double E;
    if (SB > zero) {
        if (SL > zero) E = ELL;
        else if (SR < zero) E = ERL;
        else E = ((SAR * EstarLLx - SAL * EstarRLx) + SAR * SAL * (BRL - BLL)) * SARSAL;
    } else if (ST < zero) {
        if (SL > zero) E = ELR;
        else if (SR < zero) E = ERR;
        else E = ((SAR * EstarLRx - SAL * EstarRRx) + SAR * SAL * (BRR - BLR)) * SARSAL;
    } else if (SL > zero) {
        E = ((SAT * EstarLLy - SAB * EstarLRy) - SAT * SAB * (ALR - ALL)) * SATSAB;
    } else if (SR < zero) {
        E = ((SAT * EstarRLy - SAB * EstarRRy) - SAT * SAB * (ARR - ARL)) * SATSAB;
    } else {
        E = (((((SAL * SAB * EstarRR - SAL * SAT * EstarRL) - SAR * SAB * EstarLR) + SAR * SAT * EstarLL) * SARSAL) * SATSAB
              - SAT * SAB * SATSAB * (AstarT - AstarB))
            + SAR * SAL * SARSAL * (BstarR - BstarL);
    }
    return E * dt;
}

static double llf2d_dir(double gamma, double rl, double pl, double ul, double al,
                        double bl, double cl, double rr, double pr, double ur,
                        double ar, double br, double cr, int iso, double ciso)
{
    /* one direction of the 2D LLF, solver_magnetic.f90:267-297; the ISO
     * build replaces the averaged pressures by rho*ciso**2 (:283-289) */
    if (iso) {
        pl = rl * (ciso * ciso);
        pr = rr * (ciso * ciso);
    }
    al = half * (al + ar);
    ar = al;
    double c2 = gamma * pl / rl;
    double b2 = (al * al + bl * bl) + cl * cl;
    double d2 = half * (c2 + b2 / rl);
    double cf = sqrt(d2 + sqrt(d2 * d2 - c2 * al * al / rl));
    double vleft = cf + fabs(ul);
    c2 = gamma * pr / rr;
    b2 = (ar * ar + br * br) + cr * cr;
    d2 = half * (c2 + b2 / rr);
    cf = sqrt(d2 + sqrt(d2 * d2 - c2 * ar * ar / rr));
    double vright = cf + fabs(ur);
    double vel = fmx(vleft, vright);
    return -vel * half * (br - bl);
}

double dmxo_riemann2d(int solver, const double* LL, const double* RL,
                      const double* LR, const double* RR, double dt,
                      double gamma, double ciso, int iso)
{
    if (solver == DMX_HLLD) return hlld2d(LL, RL, LR, RR, dt, gamma);
    double rLL = LL[0], pLL = LL[1], uLL = LL[2], vLL = LL[3], ALL = LL[4], BLL = LL[5], CLL = LL[6], ELL = LL[7];
    double rRL = RL[0], pRL = RL[1], uRL = RL[2], vRL = RL[3], ARL = RL[4], BRL = RL[5], CRL = RL[6], ERL = RL[7];
    double rLR = LR[0], pLR = LR[1], uLR = LR[2], vLR = LR[3], ALR = LR[4], BLR = LR[5], CLR = LR[6], ELR = LR[7];
    double rRR = RR[0], pRR = RR[1], uRR = RR[2], vRR = RR[3], ARR = RR[4], BRR = RR[5], CRR = RR[6], ERR = RR[7];
    if (solver == DMX_LLF) {
        /* llf, solver_magnetic.f90:241-356 */
        double E = forth * (((ELL + ERL) + ELR) + ERR);
        double fx = llf2d_dir(gamma, half * (rLL + rLR), half * (pLL + pLR), half * (uLL + uLR),
                              half * (ALL + ALR), half * (BLL + BLR), half * (CLL + CLR),
                              half * (rRR + rRL), half * (pRR + pRL), half * (uRR + uRL),
                              half * (ARR + ARL), half * (BRR + BRL), half * (CRR + CRL), iso, ciso);
        double fy = llf2d_dir(gamma, half * (rLL + rRL), half * (pLL + pRL), half * (vLL + vRL),
                              half * (BLL + BRL), half * (ALL + ARL), half * (CLL + CRL),
                              half * (rRR + rLR), half * (pRR + pLR), half * (vRR + vLR),
                              half * (BRR + BLR), half * (ARR + ALR), half * (CRR + CLR), iso, ciso);
        double e = E + (fx - fy);
        return e * dt;
    }
    if (solver == DMX_UPWIND) {
        /* upwind, solver_magnetic.f90:554-600 */
        double E = forth * (((ELL + ERL) + ELR) + ERR);
        double ul = half * (uLL + uLR), bl = half * (BLL + BLR);
        double ur = half * (uRR + uRL), br = half * (BRR + BRL);
        double vel = half * (ul + ur);
        double fx = -fabs(vel) * half * (br - bl);
        ul = half * (vLL + vRL); bl = half * (ALL + ARL);
        ur = half * (vRR + vLR); br = half * (ARR + ALR);
        vel = half * (ul + ur);
        double fy = -fabs(vel) * half * (br - bl);
        double e = E + (fx - fy);
        return e * dt;
    }
    /* hlla, solver_magnetic.f90:688-727 */
    double cLLx = sqrt(ALL * ALL / rLL), cLRx = sqrt(ALR * ALR / rLR);
    double cRLx = sqrt(ARL * ARL / rRL), cRRx = sqrt(ARR * ARR / rRR);
    double cLLy = sqrt(BLL * BLL / rLL), cLRy = sqrt(BLR * BLR / rLR);
    double cRLy = sqrt(BRL * BRL / rRL), cRRy = sqrt(BRR * BRR / rRR);
    double SL = fmn(fmn(fmn(fmn(uLL, uLR), uRL), uRR) - fmx(fmx(fmx(fmx(cLLx, cLRx), cRLx), cRRx), smallc), zero);
    double SR = fmx(fmx(fmx(fmx(uLL, uLR), uRL), uRR) + fmx(fmx(fmx(fmx(cLLx, cLRx), cRLx), cRRx), smallc), zero);
    double SB = fmn(fmn(fmn(fmn(vLL, vLR), vRL), vRR) - fmx(fmx(fmx(fmx(cLLy, cLRy), cRLy), cRRy), smallc), zero);
    double ST = fmx(fmx(fmx(fmx(vLL, vLR), vRL), vRR) + fmx(fmx(fmx(fmx(cLLy, cLRy), cRLy), cRRy), smallc), zero);
    double e = ((((SL * SB * ERR - SL * ST * ERL) - SR * SB * ELR) + SR * ST * ELL) / (SR - SL) / (ST - SB)
                - ST * SB / (ST - SB) * (ARR - ALL))
               + SR * SL / (SR - SL) * (BRR - BLL);
    return e * dt;
}

/* riemann_solver, solver.f90:65-387 (Cartesian) */
static void riemann_faces(dmxo* s, double dt)
{
    const dmx_params* p = &s->p;
    int nx = s->nx, ny = s->ny, nz = s->nz;
    double gamma = p->gamma, ciso = p->ciso, Omega0 = p->omega0;
    memset(s->gstar, 0, sizeof(double) * 7 * s->n3);
    static const int ln_[3] = {IU, IV, IW}, lt1_[3] = {IV, IU, IU}, lt2_[3] = {IW, IW, IV};
    static const int bn_[3] = {IA, IB, IC}, bt1_[3] = {IB, IA, IA}, bt2_[3] = {IC, IC, IB};
    double ds[3] = {s->dy * s->dz, s->dx * s->dz, s->dx * s->dy};
    for (int d = 0; d < 3; d++) {
        int ln = ln_[d], lt1 = lt1_[d], lt2 = lt2_[d], bn = bn_[d], bt1 = bt1_[d], bt2 = bt2_[d];
        int ilo = 1, ihi = d == 0 ? nx + 1 : nx;
        int jlo = 1, jhi = d == 1 ? ny + 1 : ny;
        int klo = 1, khi = d == 2 ? nz + 1 : nz;
        int io = d == 0, jo = d == 1, ko = d == 2;
#pragma omp parallel for schedule(static)
        for (int k = klo; k <= khi; k++)
            for (int j = jlo; j <= jhi; j++)
                for (int i = ilo; i <= ihi; i++) {
                    int im = i - io, jm = j - jo, km = k - ko;
                    double bnm = half * (QM(im, jm, km, bn, d) + QP(i, j, k, bn, d));
                    double ql[8] = {QM(im, jm, km, IR, d), QM(im, jm, km, IP, d), QM(im, jm, km, ln, d), bnm,
                                    QM(im, jm, km, lt1, d), QM(im, jm, km, bt1, d),
                                    QM(im, jm, km, lt2, d), QM(im, jm, km, bt2, d)};
                    double qr[8] = {QP(i, j, k, IR, d), QP(i, j, k, IP, d), QP(i, j, k, ln, d), bnm,
                                    QP(i, j, k, lt1, d), QP(i, j, k, bt1, d),
                                    QP(i, j, k, lt2, d), QP(i, j, k, bt2, d)};
                    double fg[8], st[7];
                    dmxo_riemann1d(p->riemann, ql, qr, gamma, ciso, p->iso, fg, st);
                    if (p->riemann == DMX_HLLD)
                        for (int n = 0; n < 7; n++) GST(i, j, k, n) = st[n];
                    /* shearing-box upwind term on y faces, solver.f90:209-242 */
                    if (Omega0 > zero && d == 1 && !p->fargo) {
                        double shear = -1.5 * Omega0 * XC(i);
                        const double* q = shear > zero ? ql : qr;
                        double r = q[0], pq = q[1], u = q[2], a = q[3], v = q[4], b = q[5], w = q[6], c = q[7];
                        double Emag = half * ((a * a + b * b) + c * c);
                        double Ekin = half * ((u * u + v * v) + w * w);
                        double Etot = (Ekin + Emag) + pq / (gamma - one);
                        fg[0] = fg[0] + shear * r;
                        fg[1] = fg[1] + shear * ((Etot + Emag) - bnm * bnm);
                        fg[2] = fg[2] + shear * r * u;
                        fg[4] = fg[4] + shear * r * v;
                        fg[6] = fg[6] + shear * r * w;
                    }
                    for (int n = 0; n < 8; n++) FG(i, j, k, n, d) = fg[n];
                }
        /* pressure term, solver.f90:250-263 */
#pragma omp parallel for schedule(static)
        for (int k = klo; k <= khi - ko; k++)
            for (int j = jlo; j <= jhi - jo; j++)
                for (int i = ilo; i <= ihi - io; i++)
                    FPRE(i, j, k, d) = -(GST(i + io, j + jo, k + ko, 6) * ds[d] - GST(i, j, k, 6) * ds[d]) * dt;
    }
}

/* riemann_solver_magnetic, solver_magnetic.f90:55-235, called as in
 * godunov (scheme.f90:160-168): emfz(qRT,qRB,qLT,qLB), emfy(qRT,qLT,qRB,qLB),
 * emfx(qRT,qRB,qLT,qLB). */
static void riemann_edges(dmxo* s, double dt)
{
    const dmx_params* p = &s->p;
    int nx = s->nx, ny = s->ny, nz = s->nz;
    double gamma = p->gamma, ciso = p->ciso, Omega0 = p->omega0;
    for (int d = 0; d < 3; d++) {
        int ilo = 1, ihi = d == 0 ? nx : nx + 1;
        int jlo = 1, jhi = d == 1 ? ny : ny + 1;
        int klo = 1, khi = d == 2 ? nz : nz + 1;
        int lt1, lt2, bn, bt1, bt2;
        int ort[3], orb[3], olt[3], olb[3]; /* offsets (i,j,k) of the four states */
        double *aRT, *aRB, *aLT, *aLB;      /* arrays bound to the dummies */
        double* emf;
        if (d == 0) {
            lt1 = IV; lt2 = IW; bn = IA; bt1 = IB; bt2 = IC;
            int a[3] = {0, -1, -1}, b[3] = {0, -1, 0}, c[3] = {0, 0, -1}, e[3] = {0, 0, 0};
            memcpy(ort, a, sizeof a); memcpy(orb, b, sizeof b); memcpy(olt, c, sizeof c); memcpy(olb, e, sizeof e);
            aRT = s->qrt; aRB = s->qrb; aLT = s->qlt; aLB = s->qlb; emf = s->emfx;
        } else if (d == 1) {
            lt1 = IW; lt2 = IU; bn = IB; bt1 = IC; bt2 = IA;
            int a[3] = {-1, 0, -1}, b[3] = {0, 0, -1}, c[3] = {-1, 0, 0}, e[3] = {0, 0, 0};
            memcpy(ort, a, sizeof a); memcpy(orb, b, sizeof b); memcpy(olt, c, sizeof c); memcpy(olb, e, sizeof e);
            aRT = s->qrt; aRB = s->qlt; aLT = s->qrb; aLB = s->qlb; emf = s->emfy;
        } else {
            lt1 = IU; lt2 = IV; bn = IC; bt1 = IA; bt2 = IB;
            int a[3] = {-1, -1, 0}, b[3] = {-1, 0, 0}, c[3] = {0, -1, 0}, e[3] = {0, 0, 0};
            memcpy(ort, a, sizeof a); memcpy(orb, b, sizeof b); memcpy(olt, c, sizeof c); memcpy(olb, e, sizeof e);
            aRT = s->qrt; aRB = s->qrb; aLT = s->qlt; aLB = s->qlb; emf = s->emfz;
        }
#define QS(arr, o, v) F3(s, arr, d * 8 + (v), i + o[0], j + o[1], k + o[2])
#pragma omp parallel for schedule(static)
        for (int k = klo; k <= khi; k++)
            for (int j = jlo; j <= jhi; j++)
                for (int i = ilo; i <= ihi; i++) {
                    double LL[8], RL[8], LR[8], RR[8];
                    LL[0] = QS(aRT, ort, IR); RL[0] = QS(aLT, olt, IR);
                    LR[0] = QS(aRB, orb, IR); RR[0] = QS(aLB, olb, IR);
                    if (p->iso) {
                        double c2 = ciso * ciso;
                        LL[1] = LL[0] * c2; RL[1] = RL[0] * c2; LR[1] = LR[0] * c2; RR[1] = RR[0] * c2;
                    } else {
                        LL[1] = QS(aRT, ort, IP); RL[1] = QS(aLT, olt, IP);
                        LR[1] = QS(aRB, orb, IP); RR[1] = QS(aLB, olb, IP);
                    }
                    LL[2] = QS(aRT, ort, lt1); RL[2] = QS(aLT, olt, lt1);
                    LR[2] = QS(aRB, orb, lt1); RR[2] = QS(aLB, olb, lt1);
                    LL[3] = QS(aRT, ort, lt2); RL[3] = QS(aLT, olt, lt2);
                    LR[3] = QS(aRB, orb, lt2); RR[3] = QS(aLB, olb, lt2);
                    LL[4] = half * (QS(aRT, ort, bt1) + QS(aLT, olt, bt1));
                    RL[4] = LL[4];
                    LR[4] = half * (QS(aRB, orb, bt1) + QS(aLB, olb, bt1));
                    RR[4] = LR[4];
                    LL[5] = half * (QS(aRT, ort, bt2) + QS(aRB, orb, bt2));
                    RL[5] = half * (QS(aLT, olt, bt2) + QS(aLB, olb, bt2));
                    LR[5] = LL[5];
                    RR[5] = RL[5];
                    LL[6] = QS(aRT, ort, bn); RL[6] = QS(aLT, olt, bn);
                    LR[6] = QS(aRB, orb, bn); RR[6] = QS(aLB, olb, bn);
                    LL[7] = LL[2] * LL[5] - LL[3] * LL[4];
                    RL[7] = RL[2] * RL[5] - RL[3] * RL[4];
                    LR[7] = LR[2] * LR[5] - LR[3] * LR[4];
                    RR[7] = RR[2] * RR[5] - RR[3] * RR[4];
                    double e = dmxo_riemann2d(p->riemann2d, LL, RL, LR, RR, dt, gamma, ciso, p->iso);
                    /* shearing-box upwind term, solver_magnetic.f90:208-227 */
                    if (Omega0 > zero && !p->fargo) {
                        if (d == 0) {
                            double shear = -1.5 * Omega0 * XC(i);
                            e = shear > zero ? e + shear * LL[5] * dt : e + shear * RR[5] * dt;
                        }
                        if (d == 2) {
                            double shear = -1.5 * Omega0 * half * (XC(i) + XC(i - 1));
                            e = shear > zero ? e - shear * LL[4] * dt : e - shear * RR[4] * dt;
                        }
                    }
                    F3(s, emf, 0, i, j, k) = e;
                }
#undef QS
    }
}

/* compflux, scheme.f90:1716-1760 */
static void compflux(dmxo* s, double dt)
{
    static const int map[3][8] = {
        {IR, IP, IU, IA, IV, IB, IW, IC},
        {IR, IP, IV, IB, IU, IA, IW, IC},
        {IR, IP, IW, IC, IU, IA, IV, IB}};
    double ds[3] = {s->dy * s->dz, s->dx * s->dz, s->dx * s->dy};
    for (int d = 0; d < 3; d++)
#pragma omp parallel for schedule(static)
        for (int k = 1; k <= s->nz + 1; k++)
            for (int j = 1; j <= s->ny + 1; j++)
                for (int i = 1; i <= s->nx + 1; i++)
                    for (int n = 0; n < 8; n++)
                        FLX(i, j, k, map[d][n], d) = FG(i, j, k, n, d) * ds[d] * dt;
}

/* shearing geometry of shearingbox.f90:37-40 and siblings */
static void shear_shift(const dmxo* s, double itime, double* epsi, int* jp)
{
    const dmx_params* p = &s->p;
    double deltay = 1.5 * p->omega0 * (p->xmax - p->xmin) * itime;
    deltay = fmod(deltay, (p->ymax - p->ymin));
    *epsi = fmod(deltay, s->dy);
    *jp = (int)(deltay / s->dy);
}

static inline void remap_in(int j, int jp, int ny, int* jr, int* jrp)
{
    int a = j - jp - 1, b = a + 1;
    if (a <= 0) a += ny;
    if (b <= 0) b += ny;
    *jr = a; *jrp = b;
}

static inline void remap_out(int j, int jp, int ny, int* jr, int* jrp)
{
    int a = j + jp, b = a + 1;
    if (a > ny) a -= ny;
    if (b > ny) b -= ny;
    *jr = a; *jrp = b;
}

/* shearing_flux + shearing_emf, shearingbox.f90:207-390 (single rank) */
static void shearing_flux(dmxo* s, double itime)
{
    int nx = s->nx, ny = s->ny, nz = s->nz;
    double epsi; int jp;
    shear_shift(s, itime, &epsi, &jp);
    long nyz = (long)(ny + 7) * (nz + 7);
    double* inner = zalloc(nyz);
    double* outer = zalloc(nyz);
#define P2(a, j, k) (a)[(long)((j) + 2) + (long)(ny + 7) * ((k) + 2)]
    for (int k = 1; k <= nz + 1; k++)
        for (int j = 1; j <= ny + 1; j++) {
            P2(inner, j, k) = FLX(1, j, k, IR, 0);
            P2(outer, j, k) = FLX(nx + 1, j, k, IR, 0);
        }
    for (int k = 1; k <= nz + 1; k++)
        for (int j = 1; j <= ny + 1; j++) {
            int jr, jrp;
            remap_in(j, jp, ny, &jr, &jrp);
            double eps = one - epsi / s->dy;
            double f = (FLX(1, j, k, IR, 0) + (one - eps) * P2(outer, jr, k)) + eps * P2(outer, jrp, k);
            FLX(1, j, k, IR, 0) = half * f;
            remap_out(j, jp, ny, &jr, &jrp);
            eps = epsi / s->dy;
            f = (FLX(nx + 1, j, k, IR, 0) + (one - eps) * P2(inner, jr, k)) + eps * P2(inner, jrp, k);
            FLX(nx + 1, j, k, IR, 0) = half * f;
        }
    /* shearing_emf, shearingbox.f90:298-390 */
    for (int k = -2; k <= nz + 3; k++)
        for (int j = -2; j <= ny + 3; j++) {
            P2(inner, j, k) = EMFY(1, j, k);
            P2(outer, j, k) = EMFY(nx + 1, j, k);
        }
    for (int k = -2; k <= nz + 3; k++)
        for (int j = -1; j <= ny + 3; j++) {
            int jr, jrp;
            remap_in(j, jp, ny, &jr, &jrp);
            double eps = one - epsi / s->dy;
            double e = (EMFY(1, j, k) + (one - eps) * P2(outer, jr, k)) + eps * P2(outer, jrp, k);
            EMFY(1, j, k) = half * e;
        }
    for (int k = -2; k <= nz + 3; k++)
        for (int j = -2; j <= ny + 2; j++) {
            int jr, jrp;
            remap_out(j, jp, ny, &jr, &jrp);
            double eps = epsi / s->dy;
            double e = (EMFY(nx + 1, j, k) + (one - eps) * P2(inner, jr, k)) + eps * P2(inner, jrp, k);
            EMFY(nx + 1, j, k) = half * e;
        }
#undef P2
    free(inner);
    free(outer);
}

/* constrained_transport, scheme.f90:1940-1993 (Cartesian) */
static void constrained_transport(dmxo* s)
{
    int nx = s->nx, ny = s->ny, nz = s->nz;
    double dx = s->dx, dy = s->dy, dz = s->dz;
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny + 1; j++)
            for (int i = 1; i <= nx + 1; i++) {
                U(i, j, k, IA) = U(i, j, k, IA) + (EMFZ(i, j + 1, k) - EMFZ(i, j, k)) / dy;
                U(i, j, k, IB) = U(i, j, k, IB) - (EMFZ(i + 1, j, k) - EMFZ(i, j, k)) / dx;
            }
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz + 1; k++)
        for (int j = 1; j <= ny + 1; j++)
            for (int i = 1; i <= nx + 1; i++) {
                U(i, j, k, IA) = U(i, j, k, IA) - (EMFY(i, j, k + 1) - EMFY(i, j, k)) / dz;
                U(i, j, k, IB) = U(i, j, k, IB) + (EMFX(i, j, k + 1) - EMFX(i, j, k)) / dz;
                U(i, j, k, IC) = (U(i, j, k, IC) + (EMFY(i + 1, j, k) - EMFY(i, j, k)) / dx)
                                 - (EMFX(i, j + 1, k) - EMFX(i, j, k)) / dy;
            }
}

/* right-face copy of update/boundary, scheme.f90:1928-1932, boundary.f90:82-89 */
static void right_faces(dmxo* s)
{
    int nx = s->nx, ny = s->ny, nz = s->nz;
#pragma omp parallel for schedule(static)
    for (int k = -2; k <= nz + 3; k++)
        for (int j = -2; j <= ny + 3; j++)
            for (int i = -2; i <= nx + 3; i++) {
                if (i <= nx + 2) U(i, j, k, IA + 3) = U(i + 1, j, k, IA);
                if (j <= ny + 2) U(i, j, k, IB + 3) = U(i, j + 1, k, IB);
                if (k <= nz + 2) U(i, j, k, IC + 3) = U(i, j, k + 1, IC);
            }
}

/* update, scheme.f90:1764-1936 (Cartesian, NDIM=3) */
static void update(dmxo* s, double dt)
{
    const dmx_params* p = &s->p;
    double Omega0 = p->omega0;
    double lambda = 0, ratio = 0;
    if (Omega0 > zero) {
        lambda = Omega0 * dt;
        lambda = 2.5e-1 * lambda * lambda;
        ratio = (one - lambda) / (one + lambda);
    }
    double dvloc = one / ((s->dx * s->dy) * s->dz);
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= s->nz; k++)
        for (int j = 1; j <= s->ny; j++)
            for (int i = 1; i <= s->nx; i++) {
                if (Omega0 > zero) {
                    double dsx = two * Omega0 * dt * U(i, j, k, IV) / (one + lambda);
                    double dsy = -half * Omega0 * dt * U(i, j, k, IU) / (one + lambda);
                    U(i, j, k, IU) = U(i, j, k, IU) * ratio + dsx;
                    U(i, j, k, IV) = U(i, j, k, IV) * ratio + dsy;
                }
                for (int d = 0; d < 3; d++) {
                    int i0 = d == 0, j0 = d == 1, k0 = d == 2;
                    U(i, j, k, IR) = U(i, j, k, IR) + (FLX(i, j, k, IR, d) - FLX(i + i0, j + j0, k + k0, IR, d)) * dvloc;
                    U(i, j, k, IW) = U(i, j, k, IW) + (FLX(i, j, k, IW, d) * one - FLX(i + i0, j + j0, k + k0, IW, d) * one) * dvloc / one;
                    U(i, j, k, IP) = U(i, j, k, IP) + (FLX(i, j, k, IP, d) - FLX(i + i0, j + j0, k + k0, IP, d)) * dvloc;
                    if (Omega0 > zero) {
                        double dfn2 = (FLX(i, j, k, IU, d) - FLX(i + i0, j + j0, k + k0, IU, d)) * dvloc;
                        if (d == 0) dfn2 = dfn2 + FPRE(i, j, k, 0) * dvloc;
                        double dfn3 = (FLX(i, j, k, IV, d) - FLX(i + i0, j + j0, k + k0, IV, d)) * dvloc;
                        if (d == 1) dfn3 = dfn3 + FPRE(i, j, k, 1) * dvloc;
                        U(i, j, k, IU) = (U(i, j, k, IU) + dfn2 / (one + lambda)) + Omega0 * dt / (one + lambda) * dfn3;
                        U(i, j, k, IV) = (U(i, j, k, IV) + dfn3 / (one + lambda)) - 2.5e-1 * Omega0 * dt / (one + lambda) * dfn2;
                    } else {
                        U(i, j, k, IU) = U(i, j, k, IU) + (FLX(i, j, k, IU, d) * one - FLX(i + i0, j + j0, k + k0, IU, d) * one) * dvloc / one;
                        U(i, j, k, IV) = U(i, j, k, IV) + (FLX(i, j, k, IV, d) * one - FLX(i + i0, j + j0, k + k0, IV, d) * one) * dvloc / one;
                    }
                }
                if (Omega0 == 0) {
                    U(i, j, k, IU) = U(i, j, k, IU) + FPRE(i, j, k, 0) * dvloc;
                    U(i, j, k, IV) = U(i, j, k, IV) + FPRE(i, j, k, 1) * dvloc;
                }
                U(i, j, k, IW) = U(i, j, k, IW) + FPRE(i, j, k, 2) * dvloc;
            }
    constrained_transport(s);
    right_faces(s);
}

/* godunov, scheme.f90:104-189 */
void dmxo_godunov(dmxo* s, double time, double dt)
{
    fc_magnetic(s);
    primitive(s, dt);
    slope(s);
    trace3d(s, dt);
    riemann_faces(s, dt);
    riemann_edges(s, dt);
    compflux(s, dt);
    if (s->p.bdtype[0] == DMX_BC_SHEARINGBOX) shearing_flux(s, time + half * dt);
    update(s, dt);
}

/* ------------------------------------------------------------------------ */
/* boundaries, boundary.f90:18-341                                           */
/* ------------------------------------------------------------------------ */
static void copy_x(dmxo* s, int idst, int isrc)
{
    for (int v = 0; v < 11; v++)
        for (int k = -2; k <= s->nz + 3; k++)
            for (int j = -2; j <= s->ny + 3; j++) U(idst, j, k, v) = U(isrc, j, k, v);
}
static void copy_y(dmxo* s, int jdst, int jsrc)
{
    for (int v = 0; v < 11; v++)
        for (int k = -2; k <= s->nz + 3; k++)
            for (int i = -2; i <= s->nx + 3; i++) U(i, jdst, k, v) = U(i, jsrc, k, v);
}
static void copy_z(dmxo* s, int kdst, int ksrc)
{
    for (int v = 0; v < 11; v++)
        for (int j = -2; j <= s->ny + 3; j++)
            for (int i = -2; i <= s->nx + 3; i++) U(i, j, kdst, v) = U(i, j, ksrc, v);
}

static void bnd_x(dmxo* s)
{
    int nx = s->nx, t = s->p.bdtype[0];
    if (t == DMX_BC_PERIODIC) {
        copy_x(s, -2, nx - 2); copy_x(s, -1, nx - 1); copy_x(s, 0, nx);
        copy_x(s, nx + 1, 1); copy_x(s, nx + 2, 2); copy_x(s, nx + 3, 3);
    } else if (t == DMX_BC_ZERO_GRADIENT) {
        copy_x(s, -2, 1); copy_x(s, -1, 1); copy_x(s, 0, 1);
        copy_x(s, nx + 1, nx); copy_x(s, nx + 2, nx); copy_x(s, nx + 3, nx);
    }
}
static void bnd_y(dmxo* s)
{
    int ny = s->ny, t = s->p.bdtype[1];
    if (t == DMX_BC_PERIODIC) {
        copy_y(s, -2, ny - 2); copy_y(s, -1, ny - 1); copy_y(s, 0, ny);
        copy_y(s, ny + 1, 1); copy_y(s, ny + 2, 2); copy_y(s, ny + 3, 3);
    } else if (t == DMX_BC_ZERO_GRADIENT) {
        copy_y(s, -2, 1); copy_y(s, -1, 1); copy_y(s, 0, 1);
        copy_y(s, ny + 1, ny); copy_y(s, ny + 2, ny); copy_y(s, ny + 3, ny);
    }
}
static void bnd_z(dmxo* s)
{
    int nz = s->nz, t = s->p.bdtype[2];
    if (t == DMX_BC_PERIODIC) {
        copy_z(s, -2, nz - 2); copy_z(s, -1, nz - 1); copy_z(s, 0, nz);
        copy_z(s, nz + 1, 1); copy_z(s, nz + 2, 2); copy_z(s, nz + 3, 3);
    } else if (t == DMX_BC_ZERO_GRADIENT) {
        copy_z(s, -2, 1); copy_z(s, -1, 1); copy_z(s, 0, 1);
        copy_z(s, nz + 1, nz); copy_z(s, nz + 2, nz); copy_z(s, nz + 3, nz);
    }
}

/* shearing_slope, shearingbox.f90:394-508: y-slopes of an x-slab */
static void shear_slope(const dmxo* s, const double* slab, double* dsl)
{
    int ny = s->ny, nz = s->nz;
    long sy = ny + 7, sz = nz + 7;
#define SL(a, g, j, k, v) (a)[(long)(g) + 3 * ((long)((j) + 2) + sy * ((long)((k) + 2) + sz * (v)))]
    int st = s->p.slope_type < 2 ? s->p.slope_type : 2;
    if (st == 0) { memset(dsl, 0, sizeof(double) * 3 * sy * sz * 11); return; }
    for (int g = 0; g < 3; g++)
        for (int k = -2; k <= nz + 3; k++)
            for (int j = -1; j <= ny + 2; j++) {
                for (int v = 0; v < 11; v++) {
                    if (v == IB || v == IB + 3)
                        SL(dsl, g, j, k, v) = SL(slab, g, j + 1, k, v) - SL(slab, g, j, k, v);
                    else
                        SL(dsl, g, j, k, v) = tvd_slope(SL(slab, g, j - 1, k, v), SL(slab, g, j, k, v),
                                                        SL(slab, g, j + 1, k, v), st);
                }
            }
}

/* shearing_boundary, shearingbox.f90:19-203 (single rank) */
static void shearing_boundary(dmxo* s, double itime)
{
    int nx = s->nx, ny = s->ny, nz = s->nz;
    long sy = ny + 7, sz = nz + 7;
    long nslab = 3 * sy * sz * 11;
    double* inner = zalloc(nslab);
    double* outer = zalloc(nslab);
    double* idq = zalloc(nslab);
    double* odq = zalloc(nslab);
    double epsi; int jp;
    shear_shift(s, itime, &epsi, &jp);
    for (int v = 0; v < 11; v++)
        for (int k = -2; k <= nz + 3; k++)
            for (int j = -2; j <= ny + 3; j++)
                for (int g = 0; g < 3; g++) {
                    SL(inner, g, j, k, v) = U(1 + g, j, k, v);
                    SL(outer, g, j, k, v) = U(nx - 2 + g, j, k, v);
                }
    shear_slope(s, inner, idq);
    shear_slope(s, outer, odq);
    for (int k = -2; k <= nz + 3; k++)
        for (int j = -1; j <= ny + 2; j++) {
            int jr, jrp;
            remap_in(j, jp, ny, &jr, &jrp);
            double eps = one - epsi / s->dy;
            double lam = half * eps * (eps - one);
            for (int g = 0; g < 3; g++) {
                for (int v = 0; v < 6; v++)
                    U(-2 + g, j, k, v) = ((one - eps) * SL(outer, g, jr, k, v) + eps * SL(outer, g, jrp, k, v))
                                         + lam * (SL(odq, g, jr, k, v) - SL(odq, g, jrp, k, v));
                U(-2 + g, j, k, IB) = SL(outer, g, jr, k, IB) + eps * SL(odq, g, jr, k, IB);
                U(-2 + g, j, k, IC) = ((one - eps) * SL(outer, g, jr, k, IC) + eps * SL(outer, g, jrp, k, IC))
                                      + lam * (SL(odq, g, jr, k, IC) - SL(odq, g, jrp, k, IC));
            }
            remap_out(j, jp, ny, &jr, &jrp);
            eps = epsi / s->dy;
            lam = half * eps * (eps - one);
            for (int g = 0; g < 3; g++) {
                for (int v = 0; v < 5; v++)
                    U(nx + 1 + g, j, k, v) = ((one - eps) * SL(inner, g, jr, k, v) + eps * SL(inner, g, jrp, k, v))
                                             + lam * (SL(idq, g, jrp, k, v) - SL(idq, g, jr, k, v));
                if (g >= 1)
                    U(nx + 1 + g, j, k, IA) = ((one - eps) * SL(inner, g, jr, k, IA) + eps * SL(inner, g, jrp, k, IA))
                                              + lam * (SL(idq, g, jrp, k, IA) - SL(idq, g, jr, k, IA));
                U(nx + 1 + g, j, k, IB) = SL(inner, g, jr, k, IB) + eps * SL(idq, g, jr, k, IB);
                U(nx + 1 + g, j, k, IC) = ((one - eps) * SL(inner, g, jr, k, IC) + eps * SL(inner, g, jrp, k, IC))
                                          + lam * (SL(idq, g, jrp, k, IC) - SL(idq, g, jr, k, IC));
            }
            U(nx + 3, j, k, IA + 3) = ((one - eps) * SL(inner, 2, jr, k, IA + 3) + eps * SL(inner, 2, jrp, k, IA + 3))
                                      + lam * (SL(idq, 2, jrp, k, IA + 3) - SL(idq, 2, jr, k, IA + 3));
        }
#undef SL
    /* right faces of the x-ghost slabs, shearingbox.f90:144-195 */
    for (int i = 0; i >= -2; i--)
        for (int k = -2; k <= nz + 3; k++)
            for (int j = -2; j <= ny + 3; j++) U(i, j, k, IA + 3) = U(i + 1, j, k, IA);
    for (int k = -2; k <= nz + 3; k++)
        for (int j = -2; j <= ny + 2; j++)
            for (int i = -2; i <= 0; i++) U(i, j, k, IB + 3) = U(i, j + 1, k, IB);
    for (int k = -2; k <= nz + 2; k++)
        for (int j = -2; j <= ny + 3; j++)
            for (int i = -2; i <= 0; i++) U(i, j, k, IC + 3) = U(i, j, k + 1, IC);
    for (int i = nx + 1; i <= nx + 2; i++)
        for (int k = -2; k <= nz + 3; k++)
            for (int j = -2; j <= ny + 3; j++) U(i, j, k, IA + 3) = U(i + 1, j, k, IA);
    for (int k = -2; k <= nz + 3; k++)
        for (int j = -2; j <= ny + 2; j++)
            for (int i = nx + 1; i <= nx + 3; i++) U(i, j, k, IB + 3) = U(i, j + 1, k, IB);
    for (int k = -2; k <= nz + 2; k++)
        for (int j = -2; j <= ny + 3; j++)
            for (int i = nx + 1; i <= nx + 3; i++) U(i, j, k, IC + 3) = U(i, j, k + 1, IC);
    free(inner); free(outer); free(idq); free(odq);
}

/* boundary, boundary.f90:18-92 */
void dmxo_boundary(dmxo* s, double time, double dt)
{
    bnd_x(s);
    if (s->p.bdtype[0] == DMX_BC_SHEARINGBOX) {
        bnd_y(s);
        shearing_boundary(s, time + dt);
    }
    bnd_z(s);
    bnd_y(s);
    right_faces(s);
}

/* ------------------------------------------------------------------------ */
/* FARGO, fargo.f90:17-288                                                   */
/* ------------------------------------------------------------------------ */
void dmxo_fargo(dmxo* s, double dt)
{
    const dmx_params* p = &s->p;
    int nx = s->nx, ny = s->ny, nz = s->nz;
    double dx = s->dx, dy = s->dy, dz = s->dz, Omega0 = p->omega0;
    memset(s->emfx, 0, sizeof(double) * s->n3);
    memset(s->emfz, 0, sizeof(double) * s->n3);
    bnd_y(s);
    bnd_z(s);
    /* uin_slope, fargo.f90:247-287 */
    int st = p->slope_type < 2 ? p->slope_type : 2;
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz + 1; k++)
        for (int j = 1; j <= ny + 1; j++)
            for (int i = 1; i <= nx + 1; i++)
                for (int v = 0; v < 8; v++)
                    DUIN(i, j, k, v) = st == 0 ? zero
                                               : tvd_slope(U(i, j - 1, k, v), U(i, j, k, v), U(i, j + 1, k, v), st);
    memcpy(s->uold, s->uin, sizeof(double) * 11 * s->n3);
    /* remap of rho, rho*v, E, fargo.f90:59-93 */
#pragma omp parallel for schedule(static)
    for (int i = 1; i <= nx; i++) {
        double vK = -1.5 * Omega0 * XC(i);
        int nshft = (int)(fabs(vK * dt) / dy);
        double eps = fmod(fabs(vK * dt), dy);
        eps = eps / dy;
        double hme2 = (half - eps) * (half - eps);
        for (int j = 1; j <= ny; j++) {
            int js, jsp;
            double wa, wb;
            if (vK > zero) {
                js = j - nshft - 1;
                if (js < 1) js += ny;
                jsp = js + 1;
                if (jsp > ny) jsp -= ny;
                wa = eps; wb = one - eps;
            } else {
                js = j + nshft;
                if (js > ny) js -= ny;
                jsp = js + 1;
                if (jsp > ny) jsp -= ny;
                wa = one - eps; wb = eps;
            }
            for (int k = -2; k <= nz + 3; k++)
                for (int v = IR; v <= IP; v++) {
                    double u1 = UOLD(i, js, k, v) * wa + UOLD(i, jsp, k, v) * wb;
                    U(i, j, k, v) = (u1 + half * DUIN(i, js, k, v) * (forth - hme2))
                                    + half * DUIN(i, jsp, k, v) * (hme2 - forth);
                }
        }
    }
    /* get_emfx, fargo.f90:128-180; get_emfz, :186-240 */
#pragma omp parallel for schedule(static)
    for (int i = 1; i <= nx + 1; i++) {
        for (int which = 0; which < 2; which++) {
            if (which == 0 && i > nx) continue;
            double vK = which == 0 ? -1.5 * Omega0 * XC(i) : -1.5 * Omega0 * half * (XC(i) + XC(i - 1));
            int nshft = (int)(fabs(vK * dt) / dy);
            double eps = fmod(fabs(vK * dt), dy);
            eps = eps / dy;
            int var = which == 0 ? IC : IA;
            for (int j = 1; j <= ny + 1; j++)
                for (int k = -2; k <= nz + 3; k++) {
                    double acc = which == 0 ? EMFX(i, j, k) : EMFZ(i, j, k);
                    if (vK > zero) {
                        double lambda = half * eps * (one - eps);
                        for (int jj = 1; jj <= nshft; jj++) {
                            int js = j - jj;
                            if (js < 1) js += ny;
                            if (which == 0) acc = acc + U(i, js, k, var) * dy;
                            else acc = acc - U(i, js, k, var) * dy;
                        }
                        int js = j - nshft - 1;
                        if (js < 1) js += ny;
                        double demf = eps * U(i, js, k, var) + lambda * DUIN(i, js, k, var);
                        acc = which == 0 ? acc + demf * dy : acc - demf * dy;
                    } else {
                        double lambda = half * eps * (eps - one);
                        for (int jj = 0; jj <= nshft - 1; jj++) {
                            int js = j + jj;
                            if (js > ny) js -= ny;
                            if (which == 0) acc = acc - U(i, js, k, var) * dy;
                            else acc = acc + U(i, js, k, var) * dy;
                        }
                        int js = j + nshft;
                        if (js > ny) js -= ny;
                        double demf = -eps * U(i, js, k, var) - lambda * DUIN(i, js, k, var);
                        acc = which == 0 ? acc + demf * dy : acc - demf * dy;
                    }
                    if (which == 0) EMFX(i, j, k) = acc; else EMFZ(i, j, k) = acc;
                }
        }
    }
    /* constrained transport, fargo.f90:99-119 (left faces; the right faces
     * written there are overwritten by the following boundary call) */
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= nz + 1; k++)
        for (int j = 1; j <= ny + 1; j++)
            for (int i = 1; i <= nx + 1; i++) {
                U(i, j, k, IA) = U(i, j, k, IA) + (EMFZ(i, j + 1, k) - EMFZ(i, j, k)) / dy;
                U(i, j, k, IB) = (U(i, j, k, IB) - (EMFZ(i + 1, j, k) - EMFZ(i, j, k)) / dx)
                                 + (EMFX(i, j, k + 1) - EMFX(i, j, k)) / dz;
                U(i, j, k, IC) = U(i, j, k, IC) - (EMFX(i, j + 1, k) - EMFX(i, j, k)) / dy;
                U(i, j, k, IA + 3) = U(i, j, k, IA) + (EMFZ(i + 1, j + 1, k) - EMFZ(i, j, k)) / dy;
                U(i, j, k, IB + 3) = (U(i, j, k, IB) - (EMFZ(i + 1, j + 1, k) - EMFZ(i, j, k)) / dx)
                                     + (EMFX(i, j + 1, k + 1) - EMFX(i, j, k)) / dz;
                U(i, j, k, IC + 3) = U(i, j, k, IC) - (EMFX(i, j + 1, k + 1) - EMFX(i, j, k)) / dy;
            }
}

/* resistivity, physics.f90:296-459: edge EMFs -eta*J*dt from the face
 * fields on [1, n+1]^3 (emfx/y/z zeroed first, :313), then
 * constrained_transport (:457).  The ISO=0 energy fluxes of :363-456 go to
 * flux(:,:,:,ip,:), which nothing reads before viscosity zeroes it or the
 * next godunov overwrites it, so they are not formed here. */
static void resistivity(dmxo* s, double dt)
{
    const double eta = s->p.eta, dx = s->dx, dy = s->dy, dz = s->dz;
    const int nx = s->nx, ny = s->ny, nz = s->nz;
    memset(s->emfx, 0, sizeof(double) * s->n3);
    memset(s->emfy, 0, sizeof(double) * s->n3);
    memset(s->emfz, 0, sizeof(double) * s->n3);
    for (int k = 1; k <= nz + 1; k++)
        for (int j = 1; j <= ny + 1; j++)
            for (int i = 1; i <= nx + 1; i++) {
                double dBydx = (U(i, j, k, IB) - U(i - 1, j, k, IB)) / dx;
                double dBzdx = (U(i, j, k, IC) - U(i - 1, j, k, IC)) / dx;
                double dBxdy = (U(i, j, k, IA) - U(i, j - 1, k, IA)) / dy;
                double dBzdy = (U(i, j, k, IC) - U(i, j - 1, k, IC)) / dy;
                double dBxdz = (U(i, j, k, IA) - U(i, j, k - 1, IA)) / dz;
                double dBydz = (U(i, j, k, IB) - U(i, j, k - 1, IB)) / dz;
                double jx = dBzdy - dBydz;
                double jy = dBxdz - dBzdx, jz = dBydx - dBxdy;
                EMFX(i, j, k) = -(eta * jx * dt);
                EMFY(i, j, k) = -(eta * jy * dt);
                EMFZ(i, j, k) = -(eta * jz * dt);
            }
    constrained_transport(s);
}

/* momentum component v (IU..IW) of cell (i,j,k) over the density of (a,b,c) */
#define VEL(i, j, k, v, a, b, c) (U(i, j, k, v) / U(a, b, c, IR))

/* viscosity, physics.f90:463-663: fluxes on [1, n+1]^3 (flux zeroed first,
 * :284), then the conservative update of momentum and energy, x then y then
 * z.  The x-direction difference divides the momentum of cell i-1 by the
 * density of cell i (physics.f90:499), as the reference does. */
static void viscosity(dmxo* s, double dt)
{
    const double nu = s->p.nu, dx = s->dx, dy = s->dy, dz = s->dz;
    const double two3rd = 0.66666666666666667;
    const int nx = s->nx, ny = s->ny, nz = s->nz, iso = s->p.iso;
    memset(s->flux, 0, sizeof(double) * 24 * s->n3);
    for (int k = 1; k <= nz + 1; k++)
        for (int j = 1; j <= ny + 1; j++)
            for (int i = 1; i <= nx + 1; i++) {
                double dudx[3] = {0, 0, 0}, dudy[3] = {0, 0, 0}, dudz[3] = {0, 0, 0};
                double rho, u = 0, v = 0, w = 0, uR, uL, uRR, uRL, uLR, uLL;
                double txx, tyy, tzz, txy, txz, tyz;
                /* first direction */
                rho = half * (U(i, j, k, IR) + U(i - 1, j, k, IR));
                if (!iso) {
                    u = half * (VEL(i, j, k, IU, i, j, k) + VEL(i - 1, j, k, IU, i - 1, j, k));
                    v = half * (VEL(i, j, k, IV, i, j, k) + VEL(i - 1, j, k, IV, i - 1, j, k));
                    w = half * (VEL(i, j, k, IW, i, j, k) + VEL(i - 1, j, k, IW, i - 1, j, k));
                }
                for (int d = 0; d < 3; d++) {
                    uR = VEL(i, j, k, IU + d, i, j, k);
                    uL = VEL(i - 1, j, k, IU + d, i, j, k);
                    dudx[d] = (uR - uL) / dx;
                }
                for (int d = 0; d < 2; d++) {
                    uRR = VEL(i, j + 1, k, IU + d, i, j + 1, k);
                    uRL = VEL(i - 1, j + 1, k, IU + d, i - 1, j + 1, k);
                    uLR = VEL(i, j - 1, k, IU + d, i, j - 1, k);
                    uLL = VEL(i - 1, j - 1, k, IU + d, i - 1, j - 1, k);
                    uR = uRR + uRL; uL = uLR + uLL;
                    dudy[d] = forth * (uR - uL) / dy;
                }
                for (int d = 0; d < 3; d += 2) {
                    uRR = VEL(i, j, k + 1, IU + d, i, j, k + 1);
                    uRL = VEL(i - 1, j, k + 1, IU + d, i - 1, j, k + 1);
                    uLR = VEL(i, j, k - 1, IU + d, i, j, k - 1);
                    uLL = VEL(i - 1, j, k - 1, IU + d, i - 1, j, k - 1);
                    uR = uRR + uRL; uL = uLR + uLL;
                    dudz[d] = forth * (uR - uL) / dz;
                }
                txx = -(two3rd * nu * rho * (two * dudx[0] - dudy[1] - dudz[2]));
                txy = -(nu * rho * (dudy[0] + dudx[1]));
                txz = -(nu * rho * (dudz[0] + dudx[2]));
                FLX(i, j, k, IU, 0) = txx * dt / dx;
                FLX(i, j, k, IV, 0) = txy * dt / dx;
                FLX(i, j, k, IW, 0) = txz * dt / dx;
                if (!iso) {
                    FLX(i, j, k, IP, 0) = u * txx + v * txy + w * txz;
                    FLX(i, j, k, IP, 0) = FLX(i, j, k, IP, 0) * dt / dx;
                }
                /* second direction */
                rho = half * (U(i, j, k, IR) + U(i, j - 1, k, IR));
                if (!iso) {
                    u = half * (VEL(i, j, k, IU, i, j, k) + VEL(i, j - 1, k, IU, i, j - 1, k));
                    v = half * (VEL(i, j, k, IV, i, j, k) + VEL(i, j - 1, k, IV, i, j - 1, k));
                    w = half * (VEL(i, j, k, IW, i, j, k) + VEL(i, j - 1, k, IW, i, j - 1, k));
                }
                for (int d = 0; d < 3; d++) {
                    uR = VEL(i, j, k, IU + d, i, j, k);
                    uL = VEL(i, j - 1, k, IU + d, i, j - 1, k);
                    dudy[d] = (uR - uL) / dy;
                }
                for (int d = 0; d < 2; d++) {
                    uRR = VEL(i + 1, j, k, IU + d, i + 1, j, k);
                    uRL = VEL(i + 1, j - 1, k, IU + d, i + 1, j - 1, k);
                    uLR = VEL(i - 1, j, k, IU + d, i - 1, j, k);
                    uLL = VEL(i - 1, j - 1, k, IU + d, i - 1, j - 1, k);
                    uR = uRR + uRL; uL = uLR + uLL;
                    dudx[d] = forth * (uR - uL) / dx;
                }
                for (int d = 1; d < 3; d++) {
                    uRR = VEL(i, j, k + 1, IU + d, i, j, k + 1);
                    uRL = VEL(i, j - 1, k + 1, IU + d, i, j - 1, k + 1);
                    uLR = VEL(i, j, k - 1, IU + d, i, j, k - 1);
                    uLL = VEL(i, j - 1, k - 1, IU + d, i, j - 1, k - 1);
                    uR = uRR + uRL; uL = uLR + uLL;
                    dudz[d] = forth * (uR - uL) / dz;
                }
                tyy = -(two3rd * nu * rho * (two * dudy[1] - dudx[0] - dudz[2]));
                txy = -(nu * rho * (dudy[0] + dudx[1]));
                tyz = -(nu * rho * (dudz[1] + dudy[2]));
                FLX(i, j, k, IU, 1) = txy * dt / dy;
                FLX(i, j, k, IV, 1) = tyy * dt / dy;
                FLX(i, j, k, IW, 1) = tyz * dt / dy;
                if (!iso) {
                    FLX(i, j, k, IP, 1) = u * txy + v * tyy + w * tyz;
                    FLX(i, j, k, IP, 1) = FLX(i, j, k, IP, 1) * dt / dy;
                }
                /* third direction */
                rho = half * (U(i, j, k, IR) + U(i, j, k - 1, IR));
                if (!iso) {
                    u = half * (VEL(i, j, k, IU, i, j, k) + VEL(i, j, k - 1, IU, i, j, k - 1));
                    v = half * (VEL(i, j, k, IV, i, j, k) + VEL(i, j, k - 1, IV, i, j, k - 1));
                    w = half * (VEL(i, j, k, IW, i, j, k) + VEL(i, j, k - 1, IW, i, j, k - 1));
                }
                for (int d = 0; d < 3; d++) {
                    uR = VEL(i, j, k, IU + d, i, j, k);
                    uL = VEL(i, j, k - 1, IU + d, i, j, k - 1);
                    dudz[d] = (uR - uL) / dz;
                }
                for (int d = 0; d < 3; d += 2) {
                    uRR = VEL(i + 1, j, k, IU + d, i + 1, j, k);
                    uRL = VEL(i + 1, j, k - 1, IU + d, i + 1, j, k - 1);
                    uLR = VEL(i - 1, j, k, IU + d, i - 1, j, k);
                    uLL = VEL(i - 1, j, k - 1, IU + d, i - 1, j, k - 1);
                    uR = uRR + uRL; uL = uLR + uLL;
                    dudx[d] = forth * (uR - uL) / dx;
                }
                for (int d = 1; d < 3; d++) {
                    uRR = VEL(i, j + 1, k, IU + d, i, j + 1, k);
                    uRL = VEL(i, j + 1, k - 1, IU + d, i, j + 1, k - 1);
                    uLR = VEL(i, j - 1, k, IU + d, i, j - 1, k);
                    uLL = VEL(i, j - 1, k - 1, IU + d, i, j - 1, k - 1);
                    uR = uRR + uRL; uL = uLR + uLL;
                    dudy[d] = forth * (uR - uL) / dy;
                }
                tzz = -(two3rd * nu * rho * (two * dudz[2] - dudx[0] - dudy[1]));
                txz = -(nu * rho * (dudz[0] + dudx[2]));
                tyz = -(nu * rho * (dudz[1] + dudy[2]));
                FLX(i, j, k, IU, 2) = txz * dt / dz;
                FLX(i, j, k, IV, 2) = tyz * dt / dz;
                FLX(i, j, k, IW, 2) = tzz * dt / dz;
                if (!iso) {
                    FLX(i, j, k, IP, 2) = u * txz + v * tyz + w * tzz;
                    FLX(i, j, k, IP, 2) = FLX(i, j, k, IP, 2) * dt / dz;
                }
            }
    for (int d = 0; d < 3; d++) {
        int i0 = d == 0, j0 = d == 1, k0 = d == 2;
        for (int k = 1; k <= nz; k++)
            for (int j = 1; j <= ny; j++)
                for (int i = 1; i <= nx; i++)
                    for (int v = IU; v <= IP; v++)
                        U(i, j, k, v) = U(i, j, k, v) + (FLX(i, j, k, v, d) - FLX(i + i0, j + j0, k + k0, v, d));
    }
}
#undef VEL

/* dissipation, physics.f90:272-290 */
void dmxo_dissipation(dmxo* s, double time, double dt)
{
    if (s->p.eta > zero) resistivity(s, dt);
    if (s->p.nu > zero) viscosity(s, dt);
    dmxo_boundary(s, time, dt);
}

/* reference loop body, dumses.f90:80-120 */
double dmxo_advance(dmxo* s, int nsteps, double time, double* dts)
{
    for (int n = 0; n < nsteps; n++) {
        double dt = dmxo_compute_dt(s);
        dmxo_godunov(s, time, dt);
        if (s->p.fargo) dmxo_fargo(s, dt);
        dmxo_boundary(s, time, dt);
        if (s->p.nu > zero || s->p.eta > zero) dmxo_dissipation(s, time, dt);
        time = time + dt;
        if (dts) dts[n] = dt;
    }
    return time;
}

/* history sums, history.f90:62-118 (single rank); shear = 1 gives the
 * shearing-box stresses of the MRI problem's history (Maxwell -<Bx By>,
 * Reynolds <rho dvx dvy>, src/problem/mri/history.f90:54-90) */
void dmxo_history_variant(dmxo* s, int shear, double out[9]);
void dmxo_history(dmxo* s, double out[9]) { dmxo_history_variant(s, 0, out); }

void dmxo_history_variant(dmxo* s, int shear, double out[9])
{
    const int IT = shear ? IV : IW;   /* second velocity of the Reynolds stress */
    const dmx_params* p = &s->p;
    int nx = s->nx, ny = s->ny, nz = s->nz;
    double dtau = s->dx * s->dy * s->dz / (p->xmax - p->xmin) / (p->ymax - p->ymin) / (p->zmax - p->zmin);
    /* compute_yz_mean of vx and vz */
    double* mu = zalloc(nx + 7);
    double* mw = zalloc(nx + 7);
    for (int i = 1; i <= nx; i++) {
        double su = 0, sw = 0;
        for (int k = 1; k <= nz; k++)
            for (int j = 1; j <= ny; j++) {
                su = su + U(i, j, k, IU) / U(i, j, k, IR);
                sw = sw + U(i, j, k, IT) / U(i, j, k, IR);
            }
        mu[i + 2] = su / (double)ny / (double)nz;
        mw[i + 2] = sw / (double)ny / (double)nz;
    }
    double mass = 0, magp = 0, maxwell = 0, reynolds = 0, divB = 0, mb[3] = {0, 0, 0};
    for (int k = 1; k <= nz; k++)
        for (int j = 1; j <= ny; j++)
            for (int i = 1; i <= nx; i++) {
                double bx = U(i, j, k, IA) + U(i + 1, j, k, IA);
                double by = U(i, j, k, IB) + U(i, j + 1, k, IB);
                double bz = U(i, j, k, IC) + U(i, j, k + 1, IC);
                mass = mass + U(i, j, k, IR);
                magp = ((magp + forth * (bx * bx)) + forth * (by * by)) + forth * (bz * bz);
                maxwell = maxwell - forth * bx * (shear ? by : bz);
                mb[0] = mb[0] + U(i, j, k, IA);
                mb[1] = mb[1] + U(i, j, k, IB);
                mb[2] = mb[2] + U(i, j, k, IC);
                reynolds = reynolds + U(i, j, k, IR) * dtau * (U(i, j, k, IU) / U(i, j, k, IR) - mu[i + 2])
                                      * (U(i, j, k, IT) / U(i, j, k, IR) - mw[i + 2]);
                divB = ((divB + (U(i + 1, j, k, IA) - U(i, j, k, IA)) / s->dx)
                        + (U(i, j + 1, k, IB) - U(i, j, k, IB)) / s->dy)
                       + (U(i, j, k + 1, IC) - U(i, j, k, IC)) / s->dz;
            }
    free(mu);
    free(mw);
    magp = half * magp * dtau;
    mass = mass * dtau;
    maxwell = maxwell * dtau;
    out[0] = mass;
    out[1] = maxwell;
    out[2] = reynolds;
    out[3] = maxwell + reynolds;
    out[4] = magp;
    out[5] = mb[0] * dtau;
    out[6] = mb[1] * dtau;
    out[7] = mb[2] * dtau;
    out[8] = divB;
}
