/*
 * fvm_oracle.h -- CPU restatement of fvm2d's cell-vertex residual and RK3
 * update.  TEST INFRASTRUCTURE ONLY: tests/, the smoke check and the CPU
 * baseline leg of bench.py use it as the checker.  The product library never
 * links it.
 *
 * Arrays follow the reference's Fortran layout: column-major with the
 * component index first (prim(4,np) -> prim[4*(p-1)+(k-1)]), connectivity
 * 1-based.  Every function cites the reference file:line it restates.
 */
#ifndef FVM_ORACLE_H
#define FVM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Deck values (read_input.f:20-66) plus options the deck does not hold. */
typedef struct {
    int istart;          /* 1 scratch (2 restart is handled by the caller)   */
    int iflow;           /* 1 inviscid, 2 laminar, 3 turbulent               */
    int iterlast, maxiter, saveinterval;
    int iflux;           /* 1 Roe, 2 KFVS, 3 HLLC                            */
    int ilimit;          /* 0/1                                              */
    int vortex;          /* FarBC flag                                       */
    int cell_type;       /* 1 median, 2 Barth                                */
    int sa_compat;       /* 1: keep sa_model.f:57-62 and wall_dist.f:46 as shipped */
    double mach_inf, aoa_deg, rey, cfl, minres, xref, yref;
    double nut_factor;   /* initial nu_tilde = nut_factor*mul/rho (initialize.f:71 uses 0.1) */
    int hllc_rotate;     /* 0: HLLC flux left in the edge frame as shipped
                            (hllc_flux.f:105-154); 1: momentum rotated to x-y */
} orc_input;

typedef struct orc_ctx orc_ctx;

/* geometric.f:15-60 + math.f:8-40 + initialize.f:14-103 (scratch start).
 * coord(2,np), elem(3,nt) 1-based, ptype(np).  Returns NULL and sets *err on
 * a fatal check (degenerate triangle, non-positive area, orphan edge). */
orc_ctx *orc_create(int np, int nt, const double *coord, const int *elem,
                    const int *ptype, const orc_input *in, int *err);
void orc_destroy(orc_ctx *c);

/* dims: np nt ne netot nsp nfp nbe nswe1 nswe2 nffe1 nffe2 (11 ints) */
void orc_dims(const orc_ctx *c, int *dims);

/* Copy out a named array: "coord","elem","edge","edneigh","iedge","beindx",
 * "ds","dsb","elarea","cvarea","mcarea","drmin","wd","wd1","spts","ptype",
 * "prim","prim_old","nut","qx","qy","mu","mul","res","dt".  Returns the
 * number of elements copied (or -1). */
long orc_get(const orc_ctx *c, const char *name, void *out);
int orc_set(orc_ctx *c, const char *name, const void *in);

/* Scalars: cl, cd, fres, fres1, fresi, iresi, iter, and freestream values. */
double orc_scalar(const orc_ctx *c, const char *name);

/* One numerical flux on given states (a test hook; no reference
 * counterpart): iflux 1 Roe (roe_flux.f), 2 KFVS (kfvs_flux.f), 3 HLLC
 * (hllc_flux.f, rotated per hllc_rotate); ds(2), prim(4) each side. */
int orc_flux(const orc_ctx *c, int iflux, const double *ds, const double *pl, const double *pr,
             double *flux);

/* One call of residu (residu.f:2-91) on the current prim/nut. */
int orc_residu(orc_ctx *c);
/* update(irk) (update.f:17-35) with the current res. */
void orc_update(orc_ctx *c, int irk);
void orc_save_old(orc_ctx *c);
void orc_time_step(orc_ctx *c);
void orc_sa_model(orc_ctx *c);
int orc_check_positivity(orc_ctx *c);
void orc_clcd(orc_ctx *c);
int orc_conres(orc_ctx *c);

/* The loop body of main.f:42-76, at most n iterations.  hist gets 6 doubles
 * per iteration done: iter, fres, iresi, fresi, cl, cd.  Returns the number
 * of iterations done; *status: 0 running, 1 converged (fres <= MinRes),
 * 2 iteration limit, <0 a reference STOP (see ORC_E*). */
int orc_iterate(orc_ctx *c, int n, double *hist, int *status);

enum {
    ORC_OK = 0,
    ORC_E_DEGENERATE = -1,     /* tri_orient.f:28-31      */
    ORC_E_AREA = -2,           /* geom_stat.f:18-21,34-37, areas.f:145-150 */
    ORC_E_ORPHAN_EDGE = -3,    /* flength.f:20-27         */
    ORC_E_NEG_SOUND = -4,      /* roe_flux.f:60-65        */
    ORC_E_NONPOSITIVE = -5,    /* check_positivity.f:13-24 */
    ORC_E_ZERO_RES1 = -6,      /* conres.f:28             */
    ORC_E_INPUT = -7           /* read_input.f:72-110     */
};

/* Index (1-based) of the vertex/edge that raised the last error. */
int orc_error_index(const orc_ctx *c);

#ifdef __cplusplus
}
#endif
#endif
