/*
 * fvm_oracle.c -- CPU restatement of cpraveen/fvm2d's hot path, kept in the
 * reference's own evaluation order so that, built with -ffp-contract=off, it
 * reproduces the reference arithmetic operation for operation.
 *
 * TEST INFRASTRUCTURE ONLY (the checker): imported by tests/, by the smoke
 * check and by the CPU-baseline leg of bench.py.  Nothing in the product
 * library links or calls it.
 *
 * Pinning: every routine below is compared bit-for-bit against the flang
 * build of the reference's own sources (oracle/_ref, golden vectors under
 * tests/golden/) and against closed forms and invariants in
 * tests/test_oracle*.py.
 *
 * Layout: Fortran column-major, component first, 1-based connectivity.
 */
#include "fvm_oracle.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define NVAR 4
#define SOLID 3
#define FARFIELD 5
#define INVISCID 1
#define LAMINAR 2
#define TURBULENT 3
#define EPSILON 1.0e-16

struct orc_ctx {
    orc_input in;
    /* /dims/ (param.h:34-37) */
    int np, nt, ne, netot, nsp, nfp, nbp, nbe, nswe1, nswe2, nffe1, nffe2;
    /* /gdata/ (gdata.h:1-6) */
    int *ptype, *elem, *edge, *spts, *iedge, *edneigh, *beindx;
    int *esup1, *esup2, *psup1, *psup2;
    double *coord, *elarea, *cvarea, *mcarea, *ds, *dsb, *drmin, *wd, *wd1, *tcoord;
    double wdmin, wdmax;
    /* state */
    double *prim, *prim_old, *qx, *qy, *res, *dt, *mu, *mul, *nut, *divf, *con_new;
    /* /const/ /viscparam/ /samodel/ /lim/ (math.f:8-40) */
    double GAMMA, GAMMA1, GAS_CONST, M_PI_, prandtl, prandtl_turb;
    double ALBADA11, ALBADA12, ALBADA21, ALBADA22;
    double Cb1, Cb2, sigma_sa, kolm, Cw1, Cw2, Cw3, Cv1, Cv2, Cv11, Cw31, Cw32,
           kolm2, Cb2Sig1, Cb2Sig2;
    /* /inf/ (initialize.f:14-34) */
    double q_inf, r_inf, p_inf, T_inf, T_infd, aoa, u_inf, v_inf, ent_inf,
           a_inf, H_inf, SCONST, nut_inf, prim_inf[NVAR];
    double airk[3], birk[3];
    /* /itparam/ /resparam/ /liftdrag/ */
    int iter, started;
    double dtglobal, fres, fres1, fresi, cl, cd;
    int iresi;
    int err_index;
};

#define X(p)      c->coord[2 * ((p) - 1)]
#define Y(p)      c->coord[2 * ((p) - 1) + 1]
#define ELEM(k, t) c->elem[3 * ((t) - 1) + (k) - 1]
#define EDGE(k, e) c->edge[2 * ((e) - 1) + (k) - 1]
#define EDN(k, e)  c->edneigh[2 * ((e) - 1) + (k) - 1]
#define PRIM(k, p) c->prim[NVAR * ((p) - 1) + (k) - 1]
#define POLD(k, p) c->prim_old[NVAR * ((p) - 1) + (k) - 1]
#define QX(k, p)  c->qx[NVAR * ((p) - 1) + (k) - 1]
#define QY(k, p)  c->qy[NVAR * ((p) - 1) + (k) - 1]
#define RES(k, p) c->res[NVAR * ((p) - 1) + (k) - 1]
#define DS(k, e)  c->ds[2 * ((e) - 1) + (k) - 1]
#define TC(k, t)  c->tcoord[2 * ((t) - 1) + (k) - 1]
#define V1(a, p)  (a)[(p) - 1]

static double sq(double x) { return x * x; }
/* x**3 and x**6 as integer powers: binary exponentiation,
 * x**3 = x*(x*x), x**6 = (x*x)*((x*x)*(x*x)) */
static double pow3(double x) { double x2 = x * x; return x * x2; }
static double pow6(double x) { double x2 = x * x; double x4 = x2 * x2; return x2 * x4; }

/* ------------------------------------------------------------------ math.f:4-43 */
static void math_constants(orc_ctx *c)
{
    c->GAMMA = 1.4;
    c->GAMMA1 = c->GAMMA - 1.0;
    c->GAS_CONST = 1.0;
    c->M_PI_ = 4.0 * atan2(1.0, 1.0);
    c->prandtl = 0.72;
    c->prandtl_turb = 0.9;
    c->ALBADA11 = 2.0 / 3.0;
    c->ALBADA12 = 1.0 - c->ALBADA11;
    c->ALBADA21 = 4.0 / 3.0;
    c->ALBADA22 = 1.0 - c->ALBADA21;
    c->Cb1 = 0.1355;
    c->Cb2 = 0.622;
    c->sigma_sa = 2.0 / 3.0;
    c->kolm = 0.41;
    c->Cw1 = c->Cb1 / sq(c->kolm) + (1.0 + c->Cb2) / c->sigma_sa;
    c->Cw2 = 0.3;
    c->Cw3 = 2.0;
    c->Cv1 = 7.1;
    c->Cv2 = 5.0;
    c->Cv11 = pow3(c->Cv1);
    c->Cw31 = 1.0 + pow6(c->Cw3);
    c->Cw32 = pow6(c->Cw3);
    c->kolm2 = sq(c->kolm);
    c->Cb2Sig1 = (1.0 + c->Cb2) / c->sigma_sa;
    c->Cb2Sig2 = c->Cb2 / c->sigma_sa;
}

/* ---------------------------------------------------------- viscosity.f:2-71 */
static double sutherland_viscosity(const orc_ctx *c, double density, double pressure)
{
    double temp = pressure / density / c->GAS_CONST;
    double num = c->T_inf + c->SCONST;
    double den = temp + c->SCONST;
    return pow(temp / c->T_inf, 1.5) * (num / den) / c->in.rey;
}

static double mu_turb(const orc_ctx *c, double mu_lam, double density, double nu_tilde)
{
    double chi = nu_tilde * density / mu_lam;
    double chi3 = pow3(chi);
    double fv1 = chi3 / (chi3 + c->Cv11);
    return density * nu_tilde * fv1;
}

static void sutherland(orc_ctx *c)
{
    for (int i = 1; i <= c->np; i++)
        V1(c->mul, i) = sutherland_viscosity(c, PRIM(1, i), PRIM(4, i));
}

static void viscosity(orc_ctx *c)
{
    if (c->in.iflow == LAMINAR) {
        for (int i = 1; i <= c->np; i++) V1(c->mu, i) = V1(c->mul, i);
    } else if (c->in.iflow == TURBULENT) {
        for (int i = 1; i <= c->np; i++) {
            double mut = mu_turb(c, V1(c->mul, i), PRIM(1, i), V1(c->nut, i));
            V1(c->mu, i) = V1(c->mul, i) + mut;
        }
    }
}

/* ------------------------------------------------------ geometry (geometric.f) */
static int tri_orient(orc_ctx *c)                 /* tri_orient.f:15-41 */
{
    for (int i = 1; i <= c->nt; i++) {
        int p1 = ELEM(1, i), p2 = ELEM(2, i), p3 = ELEM(3, i);
        double dx1 = X(p2) - X(p1), dy1 = Y(p2) - Y(p1);
        double dx2 = X(p3) - X(p2), dy2 = Y(p3) - Y(p2);
        double cross = dx1 * dy2 - dx2 * dy1;
        if (cross == 0.0) { c->err_index = i; return ORC_E_DEGENERATE; }
        if (cross < 0.0) { int t = ELEM(2, i); ELEM(2, i) = ELEM(3, i); ELEM(3, i) = t; }
    }
    return ORC_OK;
}

static void el_surr_point(orc_ctx *c)              /* el_surr_point.f:12-40 */
{
    int *esup2 = c->esup2, *esup1 = c->esup1;      /* esup2[0..np], 1-based use */
    for (int i = 1; i <= c->np + 1; i++) esup2[i - 1] = 0;
    for (int ie = 1; ie <= c->nt; ie++)
        for (int k = 1; k <= 3; k++) esup2[ELEM(k, ie) + 1 - 1]++;
    for (int ip = 2; ip <= c->np + 1; ip++) esup2[ip - 1] += esup2[ip - 2];
    for (int ie = 1; ie <= c->nt; ie++)
        for (int k = 1; k <= 3; k++) {
            int ip = ELEM(k, ie);
            int istor = esup2[ip - 1] + 1;
            esup2[ip - 1] = istor;
            esup1[istor - 1] = ie;
        }
    for (int ip = c->np + 1; ip >= 2; ip--) esup2[ip - 1] = esup2[ip - 2];
    esup2[0] = 0;
}

static void pt_surr_pt(orc_ctx *c)                 /* pt_surr_pt.f:20-33 */
{
    int *lpoin = calloc((size_t)c->np + 1, sizeof(int));
    int istor = 0;
    c->psup2[0] = 0;
    for (int ip = 1; ip <= c->np; ip++) {
        for (int s = c->esup2[ip - 1] + 1; s <= c->esup2[ip]; s++) {
            int ie = c->esup1[s - 1];
            for (int k = 1; k <= 3; k++) {
                int jp = ELEM(k, ie);
                if (jp != ip && lpoin[jp] != ip) {
                    istor++;
                    c->psup1[istor - 1] = jp;
                    lpoin[jp] = ip;
                }
            }
        }
        c->psup2[ip] = istor;
    }
    free(lpoin);
}

static void create_edge(orc_ctx *c)                /* create_edge.f:9-18 */
{
    c->ne = 0;
    for (int ip = 1; ip <= c->np; ip++)
        for (int k = c->psup2[ip - 1] + 1; k <= c->psup2[ip]; k++) {
            int nb = c->psup1[k - 1];
            if (nb > ip) { c->ne++; EDGE(1, c->ne) = ip; EDGE(2, c->ne) = nb; }
        }
}

static void el_surr_edge(orc_ctx *c)               /* el_surr_edge.f:9-33 */
{
    for (int i = 1; i <= c->ne; i++) {
        EDN(1, i) = 0; EDN(2, i) = 0;
        int n1 = EDGE(1, i), n2 = EDGE(2, i);
        for (int jt = c->esup2[n1 - 1] + 1; jt <= c->esup2[n1]; jt++) {
            int el = c->esup1[jt - 1];
            if ((n1 == ELEM(1, el) && n2 == ELEM(2, el)) ||
                (n1 == ELEM(2, el) && n2 == ELEM(3, el)) ||
                (n1 == ELEM(3, el) && n2 == ELEM(1, el))) EDN(1, i) = el;
            if ((n2 == ELEM(1, el) && n1 == ELEM(2, el)) ||
                (n2 == ELEM(2, el) && n1 == ELEM(3, el)) ||
                (n2 == ELEM(3, el) && n1 == ELEM(1, el))) EDN(2, i) = el;
        }
        if (EDN(1, i) == 0) {
            EDN(1, i) = EDN(2, i); EDN(2, i) = 0;
            int t = EDGE(1, i); EDGE(1, i) = EDGE(2, i); EDGE(2, i) = t;
        }
    }
}

static void circumcenter(orc_ctx *c, int el)       /* tri_coord.f:70-136 */
{
    int n1 = ELEM(1, el), n2 = ELEM(2, el), n3 = ELEM(3, el);
    double x1 = X(n1), y1 = Y(n1), x2 = X(n2), y2 = Y(n2), x3 = X(n3), y3 = Y(n3);
    double dx1 = x2 - x3, dy1 = y2 - y3, l1 = sq(dx1) + sq(dy1);
    double dx2 = x3 - x1, dy2 = y3 - y1, l2 = sq(dx2) + sq(dy2);
    double dx3 = x1 - x2, dy3 = y1 - y2, l3 = sq(dx3) + sq(dy3);
    double beta1 = fmax(0.0, l2 + l3 - l1);
    double beta2 = fmax(0.0, l3 + l1 - l2);
    double beta3 = fmax(0.0, l1 + l2 - l3);
    double det = (x2 - x1) * (y3 - y1) - (x3 - x1) * (y2 - y1);
    double b1 = 0.5 * ((x2 - x1) * (x2 + x1) + (y2 - y1) * (y2 + y1));
    double b2 = 0.5 * ((x3 - x1) * (x3 + x1) + (y3 - y1) * (y3 + y1));
    double xc = ((y3 - y1) * b1 - (y2 - y1) * b2) / det;
    double yc = (-(x3 - x1) * b1 + (x2 - x1) * b2) / det;
    if (beta1 == 0.0)      { TC(1, el) = 0.5 * (x2 + x3); TC(2, el) = 0.5 * (y2 + y3); }
    else if (beta2 == 0.0) { TC(1, el) = 0.5 * (x3 + x1); TC(2, el) = 0.5 * (y3 + y1); }
    else if (beta3 == 0.0) { TC(1, el) = 0.5 * (x1 + x2); TC(2, el) = 0.5 * (y1 + y2); }
    else                   { TC(1, el) = xc; TC(2, el) = yc; }
}

static void tri_coord(orc_ctx *c)                  /* tri_coord.f:4-41 */
{
    for (int i = 1; i <= c->nt; i++) {
        if (c->in.cell_type == 1) {
            int n1 = ELEM(1, i), n2 = ELEM(2, i), n3 = ELEM(3, i);
            TC(1, i) = (X(n1) + X(n2) + X(n3)) / 3.0;
            TC(2, i) = (Y(n1) + Y(n2) + Y(n3)) / 3.0;
        } else {
            circumcenter(c, i);
        }
    }
}

static int areas(orc_ctx *c)                       /* areas.f:2-157 */
{
    for (int i = 1; i <= c->nt; i++) {             /* tri_area */
        int n1 = ELEM(1, i), n2 = ELEM(2, i), n3 = ELEM(3, i);
        double dx1 = X(n2) - X(n1), dy1 = Y(n2) - Y(n1);
        double dx2 = X(n3) - X(n1), dy2 = Y(n3) - Y(n1);
        V1(c->elarea, i) = 0.5 * (dx1 * dy2 - dx2 * dy1);
    }
    for (int i = 1; i <= c->np; i++) V1(c->mcarea, i) = 0.0;   /* static zero storage */
    for (int i = 1; i <= c->nt; i++) {             /* mc_area */
        double a3 = V1(c->elarea, i) / 3.0;
        for (int k = 1; k <= 3; k++) V1(c->mcarea, ELEM(k, i)) = V1(c->mcarea, ELEM(k, i)) + a3;
    }
    if (c->in.cell_type == 1) {
        for (int i = 1; i <= c->np; i++) V1(c->cvarea, i) = V1(c->mcarea, i);
        return ORC_OK;
    }
    for (int i = 1; i <= c->np; i++) V1(c->cvarea, i) = 0.0;   /* cvareas_bc */
    for (int i = 1; i <= c->nt; i++) {
        for (int j = 1; j <= 3; j++) {
            int n1 = ELEM(j, i);
            int n2 = (j == 1) ? ELEM(3, i) : ELEM(j - 1, i);
            int n3 = (j == 3) ? ELEM(1, i) : ELEM(j + 1, i);
            double x[5], y[5];
            x[0] = X(n1); y[0] = Y(n1);
            x[1] = 0.5 * (X(n1) + X(n3)); y[1] = 0.5 * (Y(n1) + Y(n3));
            x[2] = TC(1, i); y[2] = TC(2, i);
            x[3] = 0.5 * (X(n1) + X(n2)); y[3] = 0.5 * (Y(n1) + Y(n2));
            x[4] = x[0]; y[4] = y[0];
            double area = 0.0;
            for (int k = 0; k < 4; k++) area = area + x[k] * y[k + 1] - x[k + 1] * y[k];
            if (area < 0.0) { c->err_index = i; return ORC_E_AREA; }
            area = 0.5 * area;
            V1(c->cvarea, n1) = V1(c->cvarea, n1) + area;
        }
    }
    return ORC_OK;
}

static void wall_dist(orc_ctx *c)                  /* wall_dist.f:17-94 */
{
    int nse = 0;
    int *sedge = malloc(sizeof(int) * (size_t)(c->ne + 1));
    for (int i = 1; i <= c->ne; i++)
        if (V1(c->ptype, EDGE(1, i)) == SOLID && V1(c->ptype, EDGE(2, i)) == SOLID) sedge[nse++] = i;
    for (int i = 1; i <= c->np; i++) {
        double w = 1.0e10;
        for (int j = 0; j < nse; j++) {
            int je = sedge[j], n1 = EDGE(1, je), n2 = EDGE(2, je);
            double dx1 = X(n1) - X(i), dy1 = Y(n1) - Y(i);
            double ds1 = sqrt(sq(dx1) + sq(dy1));
            w = fmin(w, ds1);
            double dx2 = X(n2) - X(i), dy2 = Y(n2) - Y(i);
            double ds2 = sqrt(sq(dx2) + sq(dy2));
            w = fmin(w, ds2);
            double dx3 = X(n2) - X(n1);
            double dy3 = c->in.sa_compat ? X(n2) - X(n1) : Y(n2) - Y(n1);
            double dl2 = sq(dx3) + sq(dy3);
            double xi = -(dx1 * dx3 + dy1 * dy3) / dl2;
            if (xi >= 0.0 && xi <= 1.0) {
                double xp = X(n1) + xi * dx3, yp = Y(n1) + xi * dy3;
                double dsperp = sqrt(sq(X(i) - xp) + sq(Y(i) - yp));
                w = fmin(w, dsperp);
            }
        }
        if (V1(c->ptype, i) == SOLID) w = 0.0;
        V1(c->wd, i) = w;
    }
    for (int i = 1; i <= c->nsp; i++) {
        int n1 = V1(c->spts, i);
        double w1 = 1.0e10;
        for (int j = c->psup2[n1 - 1] + 1; j <= c->psup2[n1]; j++) {
            int n2 = c->psup1[j - 1];
            if (V1(c->ptype, n2) != SOLID) w1 = fmin(w1, V1(c->wd, n2));
        }
        V1(c->wd1, i) = w1;
    }
    c->wdmin = 1.0e10; c->wdmax = 0.0;
    for (int i = 1; i <= c->nsp; i++) {
        c->wdmin = fmin(c->wdmin, V1(c->wd1, i));
        c->wdmax = fmax(c->wdmax, V1(c->wd1, i));
    }
    for (int i = 1; i <= c->np; i++) if (V1(c->ptype, i) == SOLID) V1(c->wd, i) = 1.0;
    free(sedge);
}

static int flength(orc_ctx *c)                     /* flength.f:13-61 */
{
    for (int i = 1; i <= c->ne; i++) {
        int n1 = EDGE(1, i), n2 = EDGE(2, i), p1 = EDN(1, i), p2 = EDN(2, i);
        double x1, y1, x2, y2;
        if (p1 == 0) { c->err_index = i; return ORC_E_ORPHAN_EDGE; }
        x1 = TC(1, p1); y1 = TC(2, p1);
        if (p2 != 0) { x2 = TC(1, p2); y2 = TC(2, p2); }
        else { x2 = 0.5 * (X(n1) + X(n2)); y2 = 0.5 * (Y(n1) + Y(n2)); }
        DS(1, i) = -(y2 - y1);
        DS(2, i) = (x2 - x1);
    }
    for (int i = 1; i <= 2 * c->np; i++) c->dsb[i - 1] = 0.0;
    for (int i = 1; i <= c->ne; i++) {
        int n1 = EDGE(1, i), n2 = EDGE(2, i);
        if (EDN(1, i) * EDN(2, i) == 0) {
            double nx = Y(n2) - Y(n1), ny = -(X(n2) - X(n1));
            c->dsb[2 * (n1 - 1)] += 0.5 * nx; c->dsb[2 * (n1 - 1) + 1] += 0.5 * ny;
            c->dsb[2 * (n2 - 1)] += 0.5 * nx; c->dsb[2 * (n2 - 1) + 1] += 0.5 * ny;
        }
    }
    return ORC_OK;
}

static void reorder_edges(orc_ctx *c)              /* flength.f:71-108 */
{
    c->netot = c->ne;
    if (c->in.cell_type == 2) {
        int ec = 0;
        for (int i = 1; i <= c->ne; i++) {
            double length = sqrt(sq(DS(1, i)) + sq(DS(2, i)));
            if (length != 0.0) { ec++; V1(c->iedge, ec) = i; }
        }
        c->ne = ec;
    } else {
        for (int i = 1; i <= c->ne; i++) V1(c->iedge, i) = i;
    }
}

static void dtlength(orc_ctx *c)                   /* dtlength.f:17-44 */
{
    for (int ip = 1; ip <= c->np; ip++) V1(c->drmin, ip) = 1.0e15;
    for (int it = 1; it <= c->nt; it++) {
        int n1 = ELEM(1, it), n2 = ELEM(2, it), n3 = ELEM(3, it);
        double a = V1(c->elarea, it);
        double dx1 = X(n2) - X(n3), dy1 = Y(n2) - Y(n3), dr1 = sqrt(sq(dx1) + sq(dy1)), h1 = 2.0 * a / dr1;
        double dx2 = X(n3) - X(n1), dy2 = Y(n3) - Y(n1), dr2 = sqrt(sq(dx2) + sq(dy2)), h2 = 2.0 * a / dr2;
        double dx3 = X(n1) - X(n2), dy3 = Y(n1) - Y(n2), dr3 = sqrt(sq(dx3) + sq(dy3)), h3 = 2.0 * a / dr3;
        V1(c->drmin, n1) = fmin(h1, V1(c->drmin, n1));
        V1(c->drmin, n2) = fmin(h2, V1(c->drmin, n2));
        V1(c->drmin, n3) = fmin(h3, V1(c->drmin, n3));
    }
}

static void make_bdedges(orc_ctx *c)               /* make_bdedges.f:11-44 */
{
    int nswe = 0, nffe = 0;
    int *sw = malloc(sizeof(int) * (size_t)(c->netot + 1));
    int *ff = malloc(sizeof(int) * (size_t)(c->netot + 1));
    for (int i = 1; i <= c->netot; i++) {
        if (EDN(1, i) * EDN(2, i) != 0) continue;
        int n1 = EDGE(1, i), n2 = EDGE(2, i);
        if (V1(c->ptype, n1) == SOLID && V1(c->ptype, n2) == SOLID) sw[nswe++] = i;
        if (V1(c->ptype, n1) == FARFIELD && V1(c->ptype, n2) == FARFIELD) ff[nffe++] = i;
    }
    int ec = 0;
    c->nswe1 = 1;
    for (int i = 0; i < nswe; i++) c->beindx[ec++] = sw[i];
    c->nswe2 = ec;
    c->nffe1 = ec + 1;
    for (int i = 0; i < nffe; i++) c->beindx[ec++] = ff[i];
    c->nffe2 = ec;
    c->nbe = ec;
    free(sw); free(ff);
}

static int geom_stat(orc_ctx *c)                   /* geom_stat.f:11-37 */
{
    double minel = 1.0e8, mincv = 1.0e8;
    for (int i = 1; i <= c->nt; i++) minel = fmin(minel, V1(c->elarea, i));
    if (minel <= 0.0) return ORC_E_AREA;
    for (int i = 1; i <= c->np; i++) mincv = fmin(mincv, V1(c->cvarea, i));
    if (mincv <= 0.0) return ORC_E_AREA;
    return ORC_OK;
}

/* --------------------------------------------------------- initialize.f:14-58 */
static void freestream(orc_ctx *c)
{
    const orc_input *in = &c->in;
    c->q_inf = 1.0;
    c->r_inf = 1.0;
    c->p_inf = 1.0 / (c->GAMMA * sq(in->mach_inf));
    c->T_inf = c->p_inf / (c->r_inf * c->GAS_CONST);
    c->aoa = in->aoa_deg * c->M_PI_ / 180.0;
    c->u_inf = c->q_inf * cos(c->aoa);
    c->v_inf = c->q_inf * sin(c->aoa);
    c->ent_inf = c->p_inf / pow(c->r_inf, c->GAMMA);
    c->a_inf = sqrt(c->GAMMA * c->p_inf / c->r_inf);
    c->H_inf = sq(c->a_inf) / c->GAMMA1 + 0.5 * c->q_inf;
    c->T_infd = 300.0;
    c->SCONST = 110.4 * c->T_inf / c->T_infd;
    c->prim_inf[0] = c->r_inf; c->prim_inf[1] = c->u_inf;
    c->prim_inf[2] = c->v_inf; c->prim_inf[3] = c->p_inf;
    c->nut_inf = 0.0;
    c->airk[0] = 0.0; c->airk[1] = 3.0 / 4.0; c->airk[2] = 1.0 / 3.0;
    c->birk[0] = 1.0; c->birk[1] = 1.0 / 4.0; c->birk[2] = 2.0 / 3.0;
}

static void init_scratch(orc_ctx *c)               /* initialize.f:60-90 */
{
    for (int j = 1; j <= c->np; j++)
        for (int i = 1; i <= NVAR; i++) PRIM(i, j) = c->prim_inf[i - 1];
    if (c->in.iflow == TURBULENT) {
        sutherland(c);
        for (int i = 1; i <= c->np; i++) V1(c->nut, i) = c->in.nut_factor * V1(c->mul, i) / PRIM(1, i);
        viscosity(c);
        c->nut_inf = V1(c->nut, 1);
    } else if (c->in.iflow == LAMINAR) {
        sutherland(c);
        for (int i = 1; i <= c->np; i++) V1(c->nut, i) = 0.0;
        viscosity(c);
        c->nut_inf = 0.0;
    } else {
        for (int i = 1; i <= c->np; i++) { V1(c->mul, i) = 0.0; V1(c->mu, i) = 0.0; V1(c->nut, i) = 0.0; }
        c->nut_inf = 0.0;
    }
}

/* ------------------------------------------------------------ green_gauss.f */
static void green_gauss(orc_ctx *c, int visc)      /* green_gauss.f:2-92,97-194 */
{
    const double n1b6 = 1.0 / 6.0, n2b3 = 2.0 / 3.0;
    for (int is = 1; is <= c->np; is++)
        for (int iv = 1; iv <= NVAR; iv++) { QX(iv, is) = 0.0; QY(iv, is) = 0.0; }
    for (int jt = 1; jt <= c->nt; jt++) {
        int n[3] = {ELEM(1, jt), ELEM(2, jt), ELEM(3, jt)};
        double dbxx[3], dbyy[3], dxt[NVAR], dyt[NVAR];
        dbxx[0] = Y(n[1]) - Y(n[2]);
        dbxx[1] = Y(n[2]) - Y(n[0]);
        dbxx[2] = Y(n[0]) - Y(n[1]);
        dbyy[0] = X(n[2]) - X(n[1]);
        dbyy[1] = X(n[0]) - X(n[2]);
        dbyy[2] = X(n[1]) - X(n[0]);
        for (int k = 1; k <= NVAR; k++) {
            dxt[k - 1] = PRIM(k, n[0]) * dbxx[0] + PRIM(k, n[1]) * dbxx[1] + PRIM(k, n[2]) * dbxx[2];
            dyt[k - 1] = PRIM(k, n[0]) * dbyy[0] + PRIM(k, n[1]) * dbyy[1] + PRIM(k, n[2]) * dbyy[2];
            for (int m = 0; m < 3; m++) {
                QX(k, n[m]) = QX(k, n[m]) + dxt[k - 1];
                QY(k, n[m]) = QY(k, n[m]) + dyt[k - 1];
            }
        }
        if (!visc) continue;
        /* tri_visc, green_gauss.f:132-194 */
        double en[3], ext, eyt, um, vm, mut, mult, prim5t, rhot, chi, chi3, fv1;
        double txx, txy, tyy, gpr, efx, efy, areat;
        for (int m = 0; m < 3; m++) en[m] = PRIM(4, n[m]) / PRIM(1, n[m]);
        ext = en[0] * dbxx[0] + en[1] * dbxx[1] + en[2] * dbxx[2];
        eyt = en[0] * dbyy[0] + en[1] * dbyy[1] + en[2] * dbyy[2];
        um = (PRIM(2, n[0]) + PRIM(2, n[1]) + PRIM(2, n[2])) / 3.0;
        vm = (PRIM(3, n[0]) + PRIM(3, n[1]) + PRIM(3, n[2])) / 3.0;
        mut = (V1(c->mu, n[0]) + V1(c->mu, n[1]) + V1(c->mu, n[2])) / 3.0;
        mult = (V1(c->mul, n[0]) + V1(c->mul, n[1]) + V1(c->mul, n[2])) / 3.0;
        prim5t = (V1(c->nut, n[0]) + V1(c->nut, n[1]) + V1(c->nut, n[2])) / 3.0;
        rhot = (PRIM(1, n[0]) + PRIM(1, n[1]) + PRIM(1, n[2])) / 3.0;
        chi = prim5t * rhot / mult;
        chi3 = pow3(chi);
        fv1 = chi3 / (chi3 + c->Cv11);
        txx = n2b3 * mut * (2.0 * dxt[1] - dyt[2]);
        txy = mut * (dyt[1] + dxt[2]);
        tyy = n2b3 * mut * (2.0 * dyt[2] - dxt[1]);
        gpr = c->GAMMA * (mult / c->prandtl + rhot * prim5t * fv1 / c->prandtl_turb) / c->GAMMA1;
        efx = txx * um + txy * vm + gpr * ext;
        efy = txy * um + tyy * vm + gpr * eyt;
        areat = 0.25 / V1(c->elarea, jt);
        for (int m = 0; m < 3; m++) {
            RES(2, n[m]) = RES(2, n[m]) + (txx * dbxx[m] + txy * dbyy[m]) * areat;
            RES(3, n[m]) = RES(3, n[m]) + (txy * dbxx[m] + tyy * dbyy[m]) * areat;
            RES(4, n[m]) = RES(4, n[m]) + (efx * dbxx[m] + efy * dbyy[m]) * areat;
        }
    }
    for (int is = 1; is <= c->np; is++) {
        double ais = n1b6 / V1(c->mcarea, is);
        for (int k = 1; k <= NVAR; k++) { QX(k, is) = QX(k, is) * ais; QY(k, is) = QY(k, is) * ais; }
    }
}

/* ---------------------------------------------------------------- recon.f */
static double limit_albada(double ul, double ur)   /* recon.f:55-71 */
{
    if (ul * ur <= 0.0) return 0.0;
    double top = (sq(ul) + EPSILON) * ur + (sq(ur) + EPSILON) * ul;
    double bot = sq(ul) + sq(ur) + 2.0 * EPSILON;
    return top / bot;
}

static void recon(const orc_ctx *c, int e1, int e2, double *priml, double *primr)  /* recon.f:15-44 */
{
    double dx = X(e2) - X(e1), dy = Y(e2) - Y(e1);
    for (int i = 1; i <= NVAR; i++) {
        double p1 = PRIM(i, e1), p2 = PRIM(i, e2);
        double dqp = p2 - p1;
        if (c->in.ilimit == 1) {
            double dql = c->ALBADA21 * (dx * QX(i, e1) + dy * QY(i, e1)) + c->ALBADA22 * dqp;
            double si = limit_albada(dql, dqp);
            priml[i - 1] = p1 + 0.5 * si;
            double dqr = c->ALBADA21 * (dx * QX(i, e2) + dy * QY(i, e2)) + c->ALBADA22 * dqp;
            double sj = limit_albada(dqr, dqp);
            primr[i - 1] = p2 - 0.5 * sj;
        } else {
            double dql = dx * QX(i, e1) + dy * QY(i, e1);
            double si = c->ALBADA11 * dql + c->ALBADA12 * dqp;
            priml[i - 1] = p1 + 0.5 * si;
            double dqr = dx * QX(i, e2) + dy * QY(i, e2);
            double sj = c->ALBADA11 * dqr + c->ALBADA12 * dqp;
            primr[i - 1] = p2 - 0.5 * sj;
        }
    }
}

/* ------------------------------------------------------------- roe_flux.f:2-115 */
static int roe_flux(const orc_ctx *c, const double *ds, const double *pl_, const double *pr_, double *flux)
{
    const double G = c->GAMMA, G1 = c->GAMMA1;
    double length = sqrt(sq(ds[0]) + sq(ds[1]));
    double ct = ds[0] / length, st = ds[1] / length;
    double rl = pl_[0], ul = pl_[1], vl = pl_[2], pl = pl_[3];
    double al2 = G * pl / rl, ql2 = sq(ul) + sq(vl), hl = al2 / G1 + 0.5 * ql2;
    double rr = pr_[0], ur = pr_[1], vr = pr_[2], pr = pr_[3];
    double ar2 = G * pr / rr, qr2 = sq(ur) + sq(vr), hr = ar2 / G1 + 0.5 * qr2;
    double unl = ul * ct + vl * st;
    double Fc[4], Fd[4];
    Fc[0] = rl * unl;
    Fc[1] = pl * ct + rl * ul * unl;
    Fc[2] = pl * st + rl * vl * unl;
    Fc[3] = rl * hl * unl;
    double rl12 = sqrt(rl), rr12 = sqrt(rr), rd = 1.0 / (rl12 + rr12);
    double ua = (ul * rl12 + ur * rr12) * rd;
    double va = (vl * rl12 + vr * rr12) * rd;
    double ha = (hl * rl12 + hr * rr12) * rd;
    double qa2 = sq(ua) + sq(va);
    double aa2 = G1 * (ha - 0.5 * qa2);
    if (!(aa2 > 0.0)) return ORC_E_NEG_SOUND;
    double aa = sqrt(aa2);
    double una = ua * ct + va * st;
    double vna = -ua * st + va * ct;
    double l1 = fmin(una - aa, 0.0), l2 = fmin(una, 0.0), l3 = l2, l4 = fmin(una + aa, 0.0);
    double du1 = rr - rl, du2 = rr * ur - rl * ul, du3 = rr * vr - rl * vl;
    double du4 = (pr - pl) / G1 + 0.5 * (rr * qr2 - rl * ql2);
    double m1 = (ct * du2 + st * du3 - una * du1) / aa;
    double m2 = G1 * (du4 - ua * du2 - va * du3 + qa2 * du1) / sq(aa);
    double a4 = 0.5 * (m1 + m2), a1 = 0.5 * (m2 - m1), a3 = du1 - a1 - a4;
    double a2 = (st * du2 - ct * du3 + vna * du1) / aa;
    double a1l1 = a1 * l1, a2l2 = a2 * l2, a3l3 = a3 * l3, a4l4 = a4 * l4;
    double aact = aa * ct, aast = aa * st;
    Fd[0] = a1l1 + a3l3 + a4l4;
    Fd[1] = a1l1 * (ua - aact) + a2l2 * aa * st + a3l3 * ua + a4l4 * (ua + aact);
    Fd[2] = a1l1 * (va - aast) - a2l2 * aa * ct + a3l3 * va + a4l4 * (va + aast);
    Fd[3] = a1l1 * (ha - una * aa) + a2l2 * aa * vna + a3l3 * 0.5 * qa2 + a4l4 * (ha + una * aa);
    for (int i = 0; i < 4; i++) flux[i] = (Fc[i] + Fd[i]) * length;
    return ORC_OK;
}

/* ------------------------------------------------------------ kfvs_flux.f:2-81 */
static void kfvs_flux(const orc_ctx *c, const double *ds, const double *pl_, const double *pr_, double *flux)
{
    const double G1 = c->GAMMA1;
    double length = sqrt(sq(ds[0]) + sq(ds[1]));
    double ct = ds[0] / length, st = ds[1] / length;
    double rl = pl_[0], ul = pl_[1], vl = pl_[2], pl = pl_[3];
    double ql2 = sq(ul) + sq(vl), el = pl / G1 + 0.5 * rl * ql2;
    double rr = pr_[0], ur = pr_[1], vr = pr_[2], pr = pr_[3];
    double qr2 = sq(ur) + sq(vr), er = pr / G1 + 0.5 * rr * qr2;
    double unl = ul * ct + vl * st, unr = ur * ct + vr * st;
    double vnl = -ul * st + vl * ct, vnr = -ur * st + vr * ct;
    double Fp[4], Fm[4], Ff[4];
    double betal = 0.5 * rl / pl, sl = unl * sqrt(betal);
    double Al = 0.5 * (1.0 + erf(sl));
    double Bl = 0.5 * exp(-sq(sl)) / sqrt(c->M_PI_ * betal);
    Fp[0] = rl * (unl * Al + Bl);
    Fp[1] = (pl + rl * sq(unl)) * Al + rl * unl * Bl;
    Fp[2] = rl * (unl * Al + Bl) * vnl;
    Fp[3] = (el + pl) * unl * Al + (el + 0.5 * pl) * Bl;
    double betar = 0.5 * rr / pr, sr = unr * sqrt(betar);
    double Ar = 0.5 * (1.0 - erf(sr));
    double Br = 0.5 * exp(-sq(sr)) / sqrt(c->M_PI_ * betar);
    Fm[0] = rr * (unr * Ar - Br);
    Fm[1] = (pr + rr * sq(unr)) * Ar - rr * unr * Br;
    Fm[2] = rr * (unr * Ar - Br) * vnr;
    Fm[3] = (er + pr) * unr * Ar - (er + 0.5 * pr) * Br;
    for (int i = 0; i < 4; i++) Ff[i] = Fp[i] + Fm[i];
    flux[0] = Ff[0] * length;
    flux[1] = (ct * Ff[1] - st * Ff[2]) * length;
    flux[2] = (st * Ff[1] + ct * Ff[2]) * length;
    flux[3] = Ff[3] * length;
}

/* ----------------------------------------------------------- hllc_flux.f:4-157 */
static void hllc_flux(const orc_ctx *c, const double *ds, const double *pl_, const double *pr_, double *flux)
{
    const double gamma = c->GAMMA, G1 = c->GAMMA1;
    double length = sqrt(sq(ds[0]) + sq(ds[1]));
    double ct = ds[0] / length, st = ds[1] / length;
    double rl = pl_[0], ul = pl_[1] * ct + pl_[2] * st, vl = -pl_[1] * st + pl_[2] * ct, pl = pl_[3];
    double al = sqrt(gamma * pl / rl), El = pl / G1 + 0.5 * rl * (sq(ul) + sq(vl));
    double rr = pr_[0], ur = pr_[1] * ct + pr_[2] * st, vr = -pr_[1] * st + pr_[2] * ct, pr = pr_[3];
    double ar = sqrt(gamma * pr / rr), Er = pr / G1 + 0.5 * rr * (sq(ur) + sq(vr));
    double rbar = 0.5 * (rl + rr), abar = 0.5 * (al + ar);
    double Ppv = 0.5 * (pl + pr - (ur - ul) * rbar * abar);
    double Upv = 0.5 * (ul + ur - (pr - pl) / (rbar * abar));
    double prmax, prmin, pstar, ustar;
    if (pl > pr) { prmax = pl; prmin = pr; } else { prmax = pr; prmin = pl; }
    double Q = prmax / prmin;
    if (Q < 2.0 && prmin < Ppv && Ppv < prmax) {
        pstar = Ppv; ustar = Upv;
    } else if (Ppv < prmin) {
        double z = 0.5 * (gamma - 1.0) / gamma;
        double a1 = al + ar - 0.5 * (gamma - 1.0) * (ur - ul);
        double a2 = al / pow(pl, z) + ar / pow(pr, z);
        pstar = pow(a1 / a2, 1.0 / z);
        double Plr = pow(pl / pr, z);
        double b1 = Plr * ul / al + ur / ar + 2.0 * (Plr - 1.0) / (gamma - 1.0);
        double b2 = Plr / al + 1.0 / ar;
        ustar = b1 / b2;
    } else {
        double po = fmax(0.0, Ppv);
        double Ml = 2.0 / (gamma + 1.0) / rl, Nl = (gamma - 1.0) * pl / (gamma + 1.0);
        double gl = sqrt(Ml / (po + Nl));
        double Mr = 2.0 / (gamma + 1.0) / rr, Nr = (gamma - 1.0) * pr / (gamma + 1.0);
        double gr = sqrt(Mr / (po + Nr));
        double a1 = gl * pl + gr * pr - (ur - ul), a2 = gl + gr;
        pstar = a1 / a2;
        ustar = 0.5 * (ul + ur + (pstar - pr) * gr - (pstar - pl) * gl);
    }
    double ql = (pstar <= pl) ? 1.0 : sqrt(1.0 + 0.5 * (gamma + 1.0) * (pstar / pl - 1.0) / gamma);
    double qr = (pstar <= pr) ? 1.0 : sqrt(1.0 + 0.5 * (gamma + 1.0) * (pstar / pr - 1.0) / gamma);
    double Sl = ul - al * ql, Ss = ustar, Sr = ur + ar * qr;
    if (0.0 <= Sl) {
        flux[0] = rl * ul; flux[1] = pl + rl * sq(ul); flux[2] = rl * ul * vl; flux[3] = (El + pl) * ul;
    } else if (Sl <= 0.0 && 0.0 <= Ss) {
        double fl[4] = {rl * ul, pl + rl * sq(ul), rl * ul * vl, (El + pl) * ul}, du[4];
        double fact = rl * (Sl - ul) / (Sl - Ss);
        du[0] = fact - rl; du[1] = fact * Ss - rl * ul; du[2] = fact * vl - rl * vl;
        double t1 = El / rl + (Ss - ul) * (Ss + pl / rl / (Sl - ul));
        du[3] = fact * t1 - El;
        for (int i = 0; i < 4; i++) flux[i] = fl[i] + Sl * du[i];
    } else if (Ss <= 0.0 && 0.0 <= Sr) {
        double fr[4] = {rr * ur, pr + rr * sq(ur), rr * ur * vr, (Er + pr) * ur}, du[4];
        double fact = rr * (Sr - ur) / (Sr - Ss);
        du[0] = fact - rr; du[1] = fact * Ss - rr * ur; du[2] = fact * vr - rr * vr;
        double t1 = Er / rr + (Ss - ur) * (Ss + pr / rr / (Sr - ur));
        du[3] = fact * t1 - Er;
        for (int i = 0; i < 4; i++) flux[i] = fr[i] + Sr * du[i];
    } else {
        flux[0] = rr * ur; flux[1] = pr + rr * sq(ur); flux[2] = rr * ur * vr; flux[3] = (Er + pr) * ur;
    }
    /* hllc_rotate (no counterpart in the reference, which leaves the flux
     * in the edge frame, hllc_flux.f:105-154): momentum back to x-y with the
     * transpose of the rotation applied to the velocities above */
    if (c->in.hllc_rotate) {
        double fx = ct * flux[1] - st * flux[2], fy = st * flux[1] + ct * flux[2];
        flux[1] = fx;
        flux[2] = fy;
    }
    for (int i = 0; i < 4; i++) flux[i] = flux[i] * length;
}

/* One numerical flux on given states (test hook): iflux 1 Roe, 2 KFVS,
 * 3 HLLC (with the context's hllc_rotate).  Returns ORC_E_NEG_SOUND where
 * roe_flux.f:60-65 stops. */
int orc_flux(const orc_ctx *c, int iflux, const double *ds, const double *pl, const double *pr, double *flux)
{
    if (iflux == 1) return roe_flux(c, ds, pl, pr, flux);
    if (iflux == 2) { kfvs_flux(c, ds, pl, pr, flux); return ORC_OK; }
    if (iflux == 3) { hllc_flux(c, ds, pl, pr, flux); return ORC_OK; }
    return ORC_E_INPUT;
}

/* ----------------------------------------------------------- solid_flux.f:2-20 */
static void solid_flux(orc_ctx *c, int n1, int n2)
{
    double sx = 0.5 * (Y(n2) - Y(n1));
    double sy = -0.5 * (X(n2) - X(n1));
    double p1 = 0.75 * PRIM(4, n1) + 0.25 * PRIM(4, n2);
    double p2 = 0.25 * PRIM(4, n1) + 0.75 * PRIM(4, n2);
    RES(2, n1) = RES(2, n1) + p1 * sx;
    RES(3, n1) = RES(3, n1) + p1 * sy;
    RES(2, n2) = RES(2, n2) + p2 * sx;
    RES(3, n2) = RES(3, n2) + p2 * sy;
}

/* -------------------------------------------------------- farfield_flux.f:9-203 */
static void farfield_flux(orc_ctx *c, int n1, int n2)
{
    const double G = c->GAMMA, G1 = c->GAMMA1;
    double uinf, vinf, q2inf, pinf, rinf;
    if (c->in.mach_inf < 1.0 && c->in.vortex == 1) {
        double dx = 0.5 * (X(n1) + X(n2)) - c->in.xref;
        double dy = 0.5 * (Y(n1) + Y(n2)) - c->in.yref;
        double dref = sqrt(sq(dx) + sq(dy));
        double theta = atan2(dy, dx);
        double circ = 0.5 * c->q_inf * c->cl;
        double fact1 = circ * sqrt(1.0 - sq(c->in.mach_inf));
        double fact2 = 2.0 * c->M_PI_ * dref;
        double fact3 = 1.0 - sq(c->in.mach_inf * sin(theta - c->aoa));
        double fact = fact1 / (fact2 * fact3);
        uinf = c->u_inf + fact * sin(theta);
        vinf = c->v_inf - fact * cos(theta);
        q2inf = sq(uinf) + sq(vinf);
        double fact4 = 1.0 + 0.5 * G1 * (sq(c->q_inf) - q2inf) / sq(c->a_inf);
        rinf = c->r_inf * pow(fact4, 1.0 / G1);
        pinf = c->p_inf * pow(rinf / c->r_inf, G);
    } else {
        uinf = c->u_inf; vinf = c->v_inf; q2inf = sq(uinf) + sq(vinf);
        pinf = c->p_inf; rinf = c->r_inf;
    }
    double Uout[4] = {rinf, rinf * uinf, rinf * vinf, pinf / G1 + 0.5 * rinf * q2inf};
    double ra = 0.5 * (PRIM(1, n1) + PRIM(1, n2));
    double ua = 0.5 * (PRIM(2, n1) + PRIM(2, n2));
    double va = 0.5 * (PRIM(3, n1) + PRIM(3, n2));
    double pa = 0.5 * (PRIM(4, n1) + PRIM(4, n2));
    double qa2 = sq(ua) + sq(va);
    double aa2 = G * pa / ra;
    double aa = sqrt(G * pa / ra);
    double ha = aa2 / G1 + 0.5 * qa2;
    double Uin1[4], Uin2[4];
    Uin1[0] = PRIM(1, n1); Uin1[1] = PRIM(1, n1) * PRIM(2, n1); Uin1[2] = PRIM(1, n1) * PRIM(3, n1);
    Uin1[3] = PRIM(4, n1) / G1 + 0.5 * PRIM(1, n1) * (sq(PRIM(2, n1)) + sq(PRIM(3, n1)));
    Uin2[0] = PRIM(1, n2); Uin2[1] = PRIM(1, n2) * PRIM(2, n2); Uin2[2] = PRIM(1, n2) * PRIM(3, n2);
    Uin2[3] = PRIM(4, n2) / G1 + 0.5 * PRIM(1, n2) * (sq(PRIM(2, n2)) + sq(PRIM(3, n2)));
    double ct = (Y(n2) - Y(n1)), st = -(X(n2) - X(n1));
    double lent = sqrt(sq(ct) + sq(st));
    ct = ct / lent; st = st / lent;
    double una = ua * ct + va * st, vna = -ua * st + va * ct;
    double l[4] = {una, una, una + aa, una - aa}, lp[4], lm[4];
    lp[0] = fmax(l[0], 0.0); lp[1] = lp[0]; lp[2] = fmax(l[2], 0.0); lp[3] = fmax(l[3], 0.0);
    for (int k = 0; k < 4; k++) lm[k] = l[k] - lp[k];
    double T[4][4], S[4][4];          /* T[i][j] = T(i+1,j+1) */
    double t1 = 0.5 * ra / aa, m2 = qa2 / aa2;
    T[0][0] = 1.0; T[1][0] = ua; T[2][0] = va; T[3][0] = 0.5 * qa2;
    T[0][1] = 0.0; T[1][1] = ra * st; T[2][1] = -ra * ct; T[3][1] = -ra * vna;
    T[0][2] = t1; T[1][2] = t1 * (ua + aa * ct); T[2][2] = t1 * (va + aa * st); T[3][2] = t1 * (ha + aa * una);
    T[0][3] = t1; T[1][3] = t1 * (ua - aa * ct); T[2][3] = t1 * (va - aa * st); T[3][3] = t1 * (ha - aa * una);
    t1 = 0.5 * G1 * m2 * aa / ra;
    double t2 = una / ra, t3 = G1 * ua / aa, t4 = G1 * va / aa, t5 = G1 / (ra * aa);
    S[0][0] = 1.0 - 0.5 * G1 * m2; S[1][0] = vna / ra; S[2][0] = t1 - t2; S[3][0] = t1 + t2;
    S[0][1] = t3 / aa; S[1][1] = st / ra; S[2][1] = (ct - t3) / ra; S[3][1] = -(ct + t3) / ra;
    S[0][2] = t4 / aa; S[1][2] = -ct / ra; S[2][2] = (st - t4) / ra; S[3][2] = -(st + t4) / ra;
    S[0][3] = -G1 / aa2; S[1][3] = 0.0; S[2][3] = t5; S[3][3] = t5;
    double Tp[4][4], Tm[4][4], jp[4][4], jm[4][4];
    for (int i = 0; i < 4; i++)
        for (int k = 0; k < 4; k++) { Tp[i][k] = lp[k] * T[i][k]; Tm[i][k] = lm[k] * T[i][k]; }
    for (int i = 0; i < 4; i++)
        for (int j = 0; j < 4; j++) {
            jp[i][j] = 0.0; jm[i][j] = 0.0;
            for (int k = 0; k < 4; k++) {
                jp[i][j] = jp[i][j] + Tp[i][k] * S[k][j];
                jm[i][j] = jm[i][j] + Tm[i][k] * S[k][j];
            }
        }
    for (int i = 0; i < 4; i++) {
        double f1 = 0.0, f2 = 0.0;
        for (int j = 0; j < 4; j++) {
            f1 = f1 + jp[i][j] * Uin1[j] + jm[i][j] * Uout[j];
            f2 = f2 + jp[i][j] * Uin2[j] + jm[i][j] * Uout[j];
        }
        f1 = 0.5 * lent * f1;
        f2 = 0.5 * lent * f2;
        RES(i + 1, n1) = RES(i + 1, n1) + f1;
        RES(i + 1, n2) = RES(i + 1, n2) + f2;
    }
}

/* ------------------------------------------------------------ residu.f:2-91 */
int orc_residu(orc_ctx *c)
{
    for (int ip = 1; ip <= c->np; ip++)
        for (int iv = 1; iv <= NVAR; iv++) RES(iv, ip) = 0.0;
    if (c->in.iflow != INVISCID) { sutherland(c); viscosity(c); }
    green_gauss(c, c->in.iflow != INVISCID);
    for (int i = 1; i <= c->ne; i++) {
        int ie = V1(c->iedge, i), e1 = EDGE(1, ie), e2 = EDGE(2, ie);
        double priml[4], primr[4], flux[4];
        recon(c, e1, e2, priml, primr);
        const double *ds = &c->ds[2 * (ie - 1)];
        if (c->in.iflux == 1) {
            if (roe_flux(c, ds, priml, primr, flux) != ORC_OK) { c->err_index = ie; return ORC_E_NEG_SOUND; }
        } else if (c->in.iflux == 2) {
            kfvs_flux(c, ds, priml, primr, flux);
        } else if (c->in.iflux == 3) {
            hllc_flux(c, ds, priml, primr, flux);
        } else {
            return ORC_E_INPUT;
        }
        for (int k = 1; k <= 4; k++) {
            RES(k, e1) = RES(k, e1) + flux[k - 1];
            RES(k, e2) = RES(k, e2) - flux[k - 1];
        }
    }
    for (int i = c->nswe1; i <= c->nswe2; i++) {
        int ie = V1(c->beindx, i);
        solid_flux(c, EDGE(1, ie), EDGE(2, ie));
    }
    for (int i = c->nffe1; i <= c->nffe2; i++) {
        int ie = V1(c->beindx, i);
        farfield_flux(c, EDGE(1, ie), EDGE(2, ie));
    }
    return ORC_OK;
}

/* ------------------------------------------------------------ update.f:17-35 */
void orc_update(orc_ctx *c, int irk)
{
    const double G1 = c->GAMMA1;
    double a = c->airk[irk - 1], b = c->birk[irk - 1];
    for (int i = 1; i <= c->np; i++) {
        double c0[4], c1[4];
        c0[0] = POLD(1, i); c0[1] = POLD(1, i) * POLD(2, i); c0[2] = POLD(1, i) * POLD(3, i);
        c0[3] = POLD(4, i) / G1 + 0.5 * POLD(1, i) * (sq(POLD(2, i)) + sq(POLD(3, i)));
        c1[0] = PRIM(1, i); c1[1] = PRIM(1, i) * PRIM(2, i); c1[2] = PRIM(1, i) * PRIM(3, i);
        c1[3] = PRIM(4, i) / G1 + 0.5 * PRIM(1, i) * (sq(PRIM(2, i)) + sq(PRIM(3, i)));
        for (int j = 1; j <= 4; j++)
            c->con_new[4 * (i - 1) + j - 1] =
                a * c0[j - 1] + b * (c1[j - 1] - (V1(c->dt, i) / V1(c->cvarea, i)) * RES(j, i));
    }
    if (c->in.iflow != INVISCID)
        for (int i = 1; i <= c->nsp; i++) {
            int j = V1(c->spts, i);
            c->con_new[4 * (j - 1) + 1] = 0.0;
            c->con_new[4 * (j - 1) + 2] = 0.0;
        }
    for (int p = 1; p <= c->np; p++) {              /* con_to_prim, convert.f:31-47 */
        const double *u = &c->con_new[4 * (p - 1)];
        PRIM(1, p) = u[0];
        PRIM(2, p) = u[1] / u[0];
        PRIM(3, p) = u[2] / u[0];
        PRIM(4, p) = G1 * (u[3] - 0.5 * (sq(u[1]) + sq(u[2])) / u[0]);
    }
}

void orc_save_old(orc_ctx *c)                      /* save_old.f:11-15 */
{
    memcpy(c->prim_old, c->prim, sizeof(double) * 4 * (size_t)c->np);
}

void orc_time_step(orc_ctx *c)                     /* time_step.f:13-34 */
{
    const double G = c->GAMMA;
    for (int i = 1; i <= c->np; i++) {
        double q = sqrt(sq(PRIM(2, i)) + sq(PRIM(3, i)));
        double a = sqrt(G * PRIM(4, i) / PRIM(1, i));
        double dr = V1(c->drmin, i);
        if (c->in.iflow != INVISCID) {
            double dtv = 2.0 * G * V1(c->mu, i) / (PRIM(1, i) * c->prandtl);
            V1(c->dt, i) = c->in.cfl * sq(dr) / (dr * (q + a) + dtv);
        } else {
            V1(c->dt, i) = c->in.cfl * dr / (q + a);
        }
    }
    c->dtglobal = 1.0e10;
    for (int i = 1; i <= c->np; i++) c->dtglobal = fmin(c->dtglobal, V1(c->dt, i));
}

/* --------------------------------------------------------- sa_model.f:3-201
 * Called with nut as the 7th argument (the shipped call site main.f:54-55
 * omits it).  sa_compat keeps nut(1..3) in viscous_sa (sa_model.f:57-62). */
void orc_sa_model(orc_ctx *c)
{
    const double n1b6 = 1.0 / 6.0;
    double *divf = c->divf, *nut = c->nut, *mul = c->mul;
    for (int is = 1; is <= c->np; is++) V1(divf, is) = 0.0;
    for (int jt = 1; jt <= c->nt; jt++) {            /* viscous_sa */
        double areat = 0.25 / V1(c->elarea, jt);
        int n1 = ELEM(1, jt), n2 = ELEM(2, jt), n3 = ELEM(3, jt);
        double dbxx[3] = {Y(n2) - Y(n3), Y(n3) - Y(n1), Y(n1) - Y(n2)};
        double dbyy[3] = {X(n3) - X(n2), X(n1) - X(n3), X(n2) - X(n1)};
        double g1, g2, g3;
        if (c->in.sa_compat) { g1 = V1(nut, 1); g2 = V1(nut, 2); g3 = V1(nut, 3); }
        else { g1 = V1(nut, n1); g2 = V1(nut, n2); g3 = V1(nut, n3); }
        double nux = g1 * dbxx[0] + g2 * dbxx[1] + g3 * dbxx[2];
        double nuy = g1 * dbyy[0] + g2 * dbyy[1] + g3 * dbyy[2];
        double mult = (V1(mul, n1) + V1(mul, n2) + V1(mul, n3)) / 3.0;
        double prim5t = (V1(nut, n1) + V1(nut, n2) + V1(nut, n3)) / 3.0;
        double rhot = (PRIM(1, n1) + PRIM(1, n2) + PRIM(1, n3)) / 3.0;
        double mutb = mult / rhot + prim5t;
        int n[3] = {n1, n2, n3};
        double d[3];
        for (int k = 0; k < 3; k++)
            d[k] = (c->Cb2Sig1 * mutb - c->Cb2Sig2 * (V1(mul, n[k]) / PRIM(1, n[k]) + V1(nut, n[k])))
                   * (nux * dbxx[k] + nuy * dbyy[k]) * areat;
        for (int k = 0; k < 3; k++) V1(divf, n[k]) = V1(divf, n[k]) + d[k];
    }
    for (int ii = 1; ii <= c->ne; ii++) {            /* divergence_sa */
        int i = V1(c->iedge, ii), n1 = EDGE(1, i), n2 = EDGE(2, i);
        double unl = PRIM(2, n1) * DS(1, i) + PRIM(3, n1) * DS(2, i);
        double unr = PRIM(2, n2) * DS(1, i) + PRIM(3, n2) * DS(2, i);
        double una = 0.5 * (unl + unr);
        double flux = (una > 0.0) ? una * V1(nut, n1) : una * V1(nut, n2);
        V1(divf, n1) = V1(divf, n1) + flux;
        V1(divf, n2) = V1(divf, n2) - flux;
    }
    for (int i = c->nffe1; i <= c->nffe2; i++) {     /* outer_flux_sa */
        int ie = V1(c->beindx, i), n1 = EDGE(1, ie), n2 = EDGE(2, ie);
        double dx = X(n2) - X(n1), dy = Y(n2) - Y(n1), sx = dy, sy = -dx;
        double u = 0.5 * (PRIM(2, n1) + PRIM(2, n2)), v = 0.5 * (PRIM(3, n1) + PRIM(3, n2));
        double un = u * sx + v * sy, f1, f2;
        if (un > 0.0) { f1 = un * V1(nut, n1); f2 = un * V1(nut, n2); }
        else { f1 = un * c->nut_inf; f2 = un * c->nut_inf; }
        V1(divf, n1) = V1(divf, n1) + 0.5 * f1;
        V1(divf, n2) = V1(divf, n2) + 0.5 * f2;
    }
    for (int i = 1; i <= c->np; i++) {               /* update_sa */
        double nu = V1(nut, i), w = V1(c->wd, i), dti = V1(c->dt, i);
        double chi = nu * PRIM(1, i) / V1(mul, i);
        double chi3 = pow3(chi);
        double fv1 = chi3 / (chi3 + c->Cv11);
        double fv2 = 1.0 / pow3(1.0 + chi / c->Cv2);
        double r = nu / w / (w * c->kolm2);
        double Omega = QY(2, i) - QX(3, i);
        double fv3 = (1.0 + chi * fv1) * (1.0 - fv2) / fmax(chi, 0.001);
        double S = fv3 * fabs(Omega) + r * fv2;
        r = r / S;
        r = fmin(r, 2.0);
        double g = r + c->Cw2 * (pow6(r) - r);
        double fw = g * pow(c->Cw31 / (pow6(g) + c->Cw32), n1b6);
        double source = c->Cb1 * S * nu;
        double fact = 1.0 + c->Cw1 * fw * (nu / w) * (dti / w);
        V1(nut, i) = (nu - (dti / V1(c->cvarea, i)) * V1(divf, i) + dti * source) / fact;
    }
    for (int i = 1; i <= c->nsp; i++) V1(nut, V1(c->spts, i)) = 0.0;
}

int orc_check_positivity(orc_ctx *c)               /* check_positivity.f:11-26 */
{
    for (int pt = 1; pt <= c->np; pt++)
        if (PRIM(1, pt) <= 0.0 || PRIM(4, pt) <= 0.0 || V1(c->nut, pt) < 0.0) {
            c->err_index = pt;
            return ORC_E_NONPOSITIVE;
        }
    return ORC_OK;
}

void orc_clcd(orc_ctx *c)                          /* clcd.f:15-46 */
{
    double xf = 0.0, yf = 0.0;
    for (int i = c->nswe1; i <= c->nswe2; i++) {
        int ie = V1(c->beindx, i), ip1 = EDGE(1, ie), ip2 = EDGE(2, ie);
        double dx = X(ip2) - X(ip1), dy = Y(ip2) - Y(ip1), sx = dy, sy = -dx;
        double mu1 = V1(c->mu, ip1), mu2 = V1(c->mu, ip2);
        double p1 = PRIM(4, ip1);
        double txx1 = 2.0 / 3.0 * mu1 * (2.0 * QX(2, ip1) - QY(3, ip1));
        double txy1 = mu1 * (QY(2, ip1) + QX(3, ip1));
        double tyy1 = 2.0 / 3.0 * mu1 * (2.0 * QY(3, ip1) - QX(2, ip1));
        double p2 = PRIM(4, ip2);
        double txx2 = 2.0 / 3.0 * mu2 * (2.0 * QX(2, ip2) - QY(3, ip2));
        double txy2 = mu2 * (QY(2, ip2) + QX(3, ip2));
        double tyy2 = 2.0 / 3.0 * mu2 * (2.0 * QY(3, ip2) - QX(2, ip2));
        double p = 0.5 * (p1 + p2), txx = 0.5 * (txx1 + txx2), txy = 0.5 * (txy1 + txy2),
               tyy = 0.5 * (tyy1 + tyy2);
        xf = xf + p * sx - txx * sx - txy * sy;
        yf = yf + p * sy - txy * sx - tyy * sy;
    }
    double qdyn = 0.5 * c->r_inf * sq(c->q_inf);
    c->cl = (-sin(c->aoa) * xf + cos(c->aoa) * yf) / qdyn;
    c->cd = (cos(c->aoa) * xf + sin(c->aoa) * yf) / qdyn;
}

int orc_conres(orc_ctx *c)                         /* conres.f:11-31 */
{
    c->fres = 0.0; c->fresi = 0.0; c->iresi = 0;
    for (int i = 1; i <= c->np; i++) {
        double gra = PRIM(1, i) - POLD(1, i);
        c->fres = c->fres + sq(gra);
        double dif = fabs(gra);
        if (dif > c->fresi) { c->fresi = dif; c->iresi = i; }
    }
    c->fres = sqrt(c->fres / c->np);
    if (c->iter == c->in.iterlast) {
        c->fres1 = c->fres;
        if (c->fres1 == 0.0) return ORC_E_ZERO_RES1;
    }
    c->fres = c->fres / c->fres1;
    return ORC_OK;
}

/* ------------------------------------------------------------ main.f:42-76 */
int orc_iterate(orc_ctx *c, int n, double *hist, int *status)
{
    int done = 0, rc;
    *status = 0;
    if (!c->started) {
        c->iter = c->in.iterlast;
        c->fres = c->in.minres + 1.0;
        c->fres1 = 0.0;
        c->started = 1;
    }
    while (done < n) {
        if (!(c->fres > c->in.minres && c->iter < c->in.iterlast + c->in.maxiter)) {
            *status = (c->fres > c->in.minres) ? 2 : 1;
            break;
        }
        orc_save_old(c);
        orc_time_step(c);
        for (int irk = 1; irk <= 3; irk++) {
            if ((rc = orc_residu(c)) != ORC_OK) { *status = rc; return done; }
            orc_update(c, irk);
        }
        if (c->in.iflow == TURBULENT) orc_sa_model(c);
        if ((rc = orc_check_positivity(c)) != ORC_OK) { *status = rc; return done; }
        orc_clcd(c);
        if ((rc = orc_conres(c)) != ORC_OK) { *status = rc; return done; }
        if (hist) {
            double *h = &hist[6 * done];
            h[0] = c->iter; h[1] = c->fres; h[2] = c->iresi; h[3] = c->fresi; h[4] = c->cl; h[5] = c->cd;
        }
        c->iter++;
        done++;
        if (c->in.saveinterval > 0 && c->iter % c->in.saveinterval == 0 && c->iter >= c->in.maxiter) {
            *status = 2;
            break;
        }
    }
    if (*status == 0 && !(c->fres > c->in.minres)) *status = 1;
    return done;
}

/* ------------------------------------------------------------ setup */
orc_ctx *orc_create(int np, int nt, const double *coord, const int *elem,
                    const int *ptype, const orc_input *in, int *err)
{
    orc_ctx *c = calloc(1, sizeof(orc_ctx));
    c->in = *in;
    c->np = np; c->nt = nt;
    size_t P = (size_t)np, T = (size_t)nt, E = 3 * (size_t)nt + 3;
    c->coord = malloc(sizeof(double) * 2 * P); memcpy(c->coord, coord, sizeof(double) * 2 * P);
    c->elem = malloc(sizeof(int) * 3 * T); memcpy(c->elem, elem, sizeof(int) * 3 * T);
    c->ptype = malloc(sizeof(int) * P); memcpy(c->ptype, ptype, sizeof(int) * P);
    c->spts = calloc(P + 1, sizeof(int));
    c->edge = calloc(2 * E, sizeof(int)); c->edneigh = calloc(2 * E, sizeof(int));
    c->iedge = calloc(E, sizeof(int)); c->beindx = calloc(E, sizeof(int));
    c->esup1 = calloc(3 * T + 1, sizeof(int)); c->esup2 = calloc(P + 2, sizeof(int));
    c->psup1 = calloc(6 * T + 1, sizeof(int)); c->psup2 = calloc(P + 2, sizeof(int));
    c->elarea = calloc(T, sizeof(double)); c->tcoord = calloc(2 * T, sizeof(double));
    c->cvarea = calloc(P, sizeof(double)); c->mcarea = calloc(P, sizeof(double));
    c->ds = calloc(2 * E, sizeof(double)); c->dsb = calloc(2 * P, sizeof(double));
    c->drmin = calloc(P, sizeof(double)); c->wd = calloc(P, sizeof(double));
    c->wd1 = calloc(P + 1, sizeof(double));
    c->prim = calloc(4 * P, sizeof(double)); c->prim_old = calloc(4 * P, sizeof(double));
    c->qx = calloc(4 * P, sizeof(double)); c->qy = calloc(4 * P, sizeof(double));
    c->res = calloc(4 * P, sizeof(double)); c->dt = calloc(P, sizeof(double));
    c->mu = calloc(P, sizeof(double)); c->mul = calloc(P, sizeof(double));
    c->nut = calloc(P, sizeof(double)); c->divf = calloc(P, sizeof(double));
    c->con_new = calloc(4 * P, sizeof(double));

    int rc = ORC_OK;
    math_constants(c);
    /* read_grid.f:62-94: point classification */
    for (int ip = 1; ip <= np; ip++) {
        int t = V1(c->ptype, ip);
        if (t == SOLID) V1(c->spts, ++c->nsp) = ip;
        if (t == FARFIELD) c->nfp++;
        if (t != 0) c->nbp++;
    }
    if ((rc = tri_orient(c)) != ORC_OK) goto fail;
    el_surr_point(c);
    pt_surr_pt(c);
    create_edge(c);
    el_surr_edge(c);
    tri_coord(c);
    if ((rc = areas(c)) != ORC_OK) goto fail;
    if (in->iflow == TURBULENT) wall_dist(c);
    if ((rc = flength(c)) != ORC_OK) goto fail;
    reorder_edges(c);
    dtlength(c);
    make_bdedges(c);
    if ((rc = geom_stat(c)) != ORC_OK) goto fail;
    freestream(c);
    init_scratch(c);
    if (err) *err = ORC_OK;
    return c;
fail:
    if (err) *err = rc;
    orc_destroy(c);
    return NULL;
}

void orc_destroy(orc_ctx *c)
{
    if (!c) return;
    double *d[] = {c->coord, c->elarea, c->tcoord, c->cvarea, c->mcarea, c->ds, c->dsb, c->drmin,
                   c->wd, c->wd1, c->prim, c->prim_old, c->qx, c->qy, c->res, c->dt, c->mu,
                   c->mul, c->nut, c->divf, c->con_new};
    int *ii[] = {c->elem, c->ptype, c->spts, c->edge, c->edneigh, c->iedge, c->beindx, c->esup1,
                 c->esup2, c->psup1, c->psup2};
    for (size_t k = 0; k < sizeof d / sizeof d[0]; k++) free(d[k]);
    for (size_t k = 0; k < sizeof ii / sizeof ii[0]; k++) free(ii[k]);
    free(c);
}

void orc_dims(const orc_ctx *c, int *d)
{
    d[0] = c->np; d[1] = c->nt; d[2] = c->ne; d[3] = c->netot; d[4] = c->nsp; d[5] = c->nfp;
    d[6] = c->nbe; d[7] = c->nswe1; d[8] = c->nswe2; d[9] = c->nffe1; d[10] = c->nffe2;
}

typedef struct { const char *name; int is_int; size_t off; int per; int which; } field_t;

static void *field_ptr(const orc_ctx *c, const char *name, int *is_int, long *count)
{
    long P = c->np, T = c->nt, E = c->netot;
#define F(nm, ptr, isint, cnt) if (!strcmp(name, nm)) { *is_int = isint; *count = (cnt); return (void *)(ptr); }
    F("coord", c->coord, 0, 2 * P) F("elem", c->elem, 1, 3 * T) F("ptype", c->ptype, 1, P)
    F("spts", c->spts, 1, c->nsp) F("edge", c->edge, 1, 2 * E) F("edneigh", c->edneigh, 1, 2 * E)
    F("iedge", c->iedge, 1, c->ne) F("beindx", c->beindx, 1, c->nbe)
    F("ds", c->ds, 0, 2 * E) F("dsb", c->dsb, 0, 2 * P) F("elarea", c->elarea, 0, T)
    F("cvarea", c->cvarea, 0, P) F("mcarea", c->mcarea, 0, P) F("drmin", c->drmin, 0, P)
    F("wd", c->wd, 0, P) F("wd1", c->wd1, 0, c->nsp) F("tcoord", c->tcoord, 0, 2 * T)
    F("prim", c->prim, 0, 4 * P) F("prim_old", c->prim_old, 0, 4 * P) F("nut", c->nut, 0, P)
    F("qx", c->qx, 0, 4 * P) F("qy", c->qy, 0, 4 * P) F("mu", c->mu, 0, P) F("mul", c->mul, 0, P)
    F("res", c->res, 0, 4 * P) F("dt", c->dt, 0, P) F("divf", c->divf, 0, P)
#undef F
    return NULL;
}

long orc_get(const orc_ctx *c, const char *name, void *out)
{
    int is_int; long n;
    void *p = field_ptr(c, name, &is_int, &n);
    if (!p) return -1;
    memcpy(out, p, (size_t)n * (is_int ? sizeof(int) : sizeof(double)));
    return n;
}

int orc_set(orc_ctx *c, const char *name, const void *in)
{
    int is_int; long n;
    void *p = field_ptr(c, name, &is_int, &n);
    if (!p) return -1;
    memcpy(p, in, (size_t)n * (is_int ? sizeof(int) : sizeof(double)));
    return 0;
}

double orc_scalar(const orc_ctx *c, const char *name)
{
#define S(nm, v) if (!strcmp(name, nm)) return (double)(v);
    S("cl", c->cl) S("cd", c->cd) S("fres", c->fres) S("fres1", c->fres1) S("fresi", c->fresi)
    S("iresi", c->iresi) S("iter", c->iter) S("p_inf", c->p_inf) S("u_inf", c->u_inf)
    S("v_inf", c->v_inf) S("a_inf", c->a_inf) S("T_inf", c->T_inf) S("SCONST", c->SCONST)
    S("nut_inf", c->nut_inf) S("aoa", c->aoa) S("dtglobal", c->dtglobal) S("GAMMA1", c->GAMMA1)
    S("M_PI", c->M_PI_) S("Cw1", c->Cw1) S("Cv11", c->Cv11) S("Cw31", c->Cw31) S("Cw32", c->Cw32)
    S("kolm2", c->kolm2) S("Cb2Sig1", c->Cb2Sig1) S("Cb2Sig2", c->Cb2Sig2) S("wdmin", c->wdmin)
    S("wdmax", c->wdmax) S("ALBADA12", c->ALBADA12) S("ALBADA22", c->ALBADA22)
#undef S
    return NAN;
}

int orc_error_index(const orc_ctx *c) { return c->err_index; }
