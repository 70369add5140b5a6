/*
 * fvm2d_gpu.h -- C-ABI of the MI355X (gfx950) implementation of fvm2d's
 * cell-vertex residual and RK3 update.
 *
 * Conventions (those of the reference's Fortran, so that a Fortran host can
 * bind every entry point with iso_c_binding and pass its own arrays):
 *   - arrays are column-major with the component first: prim(4,np) is
 *     prim[4*(p-1) + k-1]; coord(2,np); elem(3,nt); edge(2,netot); ds(2,netot)
 *   - connectivity is 1-based
 *   - every pointer is borrowed for the duration of the call only
 *   - every entry point returns an fvm_status (0 = OK)
 * The reference has no plugin API: the interfaces below replace subroutine
 * calls of its driver loop (src/main.f:42-76) and of geometric()
 * (src/main.f:25-26).  The file:line each one replaces is given with it.
 */
#ifndef FVM2D_GPU_H
#define FVM2D_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

typedef enum {
    FVM_OK = 0,
    FVM_CONVERGED = 1,          /* fres <= MinRes (src/main.f:42)                  */
    FVM_ITER_LIMIT = 2,         /* iter reached LastIter+MaxIter (src/main.f:42,70) */
    FVM_E_DEGENERATE = -1,      /* src/tri_orient.f:28-31 ("triangle is degenerate") */
    FVM_E_AREA = -2,            /* src/geom_stat.f:18-21,34-37, src/areas.f:145-150  */
    FVM_E_ORPHAN_EDGE = -3,     /* src/flength.f:20-27                              */
    FVM_E_NEG_SOUND = -4,       /* src/roe_flux.f:60-65 ("Sonic speed is negative") */
    FVM_E_NONPOSITIVE = -5,     /* src/check_positivity.f:13-24                     */
    FVM_E_ZERO_RES1 = -6,       /* src/conres.f:28                                  */
    FVM_E_INPUT = -7,           /* src/read_input.f:72-110, src/residu.f:67-70      */
    FVM_E_SIZE = -8,            /* a caller-sized array is too small                */
    FVM_E_HIP = -10,            /* a HIP runtime call failed                        */
    FVM_E_NOMEM = -11           /* device allocation failed                         */
} fvm_status;

/* Values of the COMMON blocks the hot path reads (src/param.h:56-142,
 * src/common.h:10-19), set from the input deck (src/read_input.f:20-66). */
typedef struct {
    int iflow;          /* 1 inviscid, 2 laminar, 3 turbulent (FlowType)        */
    int iflux;          /* 1 Roe, 2 KFVS, 3 HLLC (Flux)                          */
    int ilimit;         /* 0/1 van Albada limiter (Limiter)                      */
    int vortex;         /* farfield point-vortex correction (FarBC)              */
    int cell_type;      /* 1 median, 2 Barth (CellType)                          */
    int iterlast;       /* LastIter                                              */
    int maxiter;        /* MaxIter                                               */
    int saveinterval;   /* SaveInt: with iter >= MaxIter, stops at a multiple    */
    double mach_inf, aoa_deg, rey, cfl, minres, xref, yref;
    /* options the deck does not hold */
    int sa_compat;      /* 1: nu_tilde gradient from vertices 1..3 as in
                           src/sa_model.f:57-62 (and wall_dist.f:46's dy3);
                           0: the vertices of the triangle                      */
    int hllc_rotate;    /* 0: HLLC flux left in the edge frame as in
                           src/hllc_flux.f:105-154; 1: rotated to x-y          */
    int precision;      /* 0 FP64 throughout; 1 split: FP32 reconstruction and
                           flux, FP64 geometry, accumulation, update, SA, norms */
    int patch_size;     /* vertices per patch of the LDS-tiled stage kernel
                           (<= 256; 0 = automatic: whole rounds of resident
                           workgroups, about 64 vertices a patch)               */
    double nut_factor;  /* initial nu_tilde = nut_factor*mu_l/rho
                           (src/initialize.f:71 uses 0.1)                       */
    int block_threads;  /* workgroup size of the stage kernel: 256, or 512
                           (Euler with the limiter only, any flux; patches up
                           to 512 vertices); 0 = automatic: 512 for limited
                           Euler solves of at least 32768 vertices with the
                           automatic patch size, else 256.  The same results
                           either way; batch members share one size          */
    int reserved_params; /* 0                                                   */
} fvm_params;

/* Mesh arrays as geometric() leaves them (src/geometric.f:15-60). */
typedef struct {
    int np, nt, ne, netot, nsp, nswe1, nswe2, nffe1, nffe2;
    int reserved;
    const double *coord;   /* (2,np)                                           */
    const int *elem;       /* (3,nt), counter-clockwise                        */
    const int *edge;       /* (2,netot), interior triangle on the left         */
    const int *iedge;      /* (ne) live edges (src/flength.f:84-105)           */
    const int *beindx;     /* (nffe2) solid then farfield edges                */
    const int *ptype;      /* (np) 0 interior, 3 solid, 5 farfield             */
    const int *spts;       /* (nsp) solid points                               */
    const double *ds;      /* (2,netot) dual-face normals                      */
    const double *elarea;  /* (nt)                                             */
    const double *cvarea;  /* (np)                                             */
    const double *mcarea;  /* (np) median-cell areas (gradient scaling)        */
    const double *drmin;   /* (np) time-step length                            */
    const double *wd;      /* (np) wall distance (turbulent only, else NULL)   */
} fvm_mesh;

typedef struct fvm_ctx fvm_ctx;

/* Replaces geometric(coord,elarea,cvarea,ds,dsb,mcarea,wd,drmin)
 * (src/geometric.f:2-63) and read_grid's point classification
 * (src/read_grid.f:62-94).  Host code.  elem is re-oriented in place
 * (src/tri_orient.f).  Edge arrays hold max_edges columns (3*nt suffices).
 * wd/wd1 are written only for iflow == 3.
 * dims: np nt ne netot nsp nfp nbe nswe1 nswe2 nffe1 nffe2 err_index. */
int fvm_geometric(int np, int nt, const double *coord, int *elem, const int *ptype,
                  int cell_type, int iflow, int sa_compat, int max_edges,
                  int *edge, int *edneigh, int *iedge, double *ds, double *dsb,
                  double *elarea, double *cvarea, double *mcarea, double *drmin,
                  double *wd, double *wd1, int *beindx, int *spts, int *dims);

/* Wall distance on the current GPU, replacing wall_dist(...)
 * (src/wall_dist.f:3-95), bit for bit: wd(np) for every point (1 on solid
 * points, as the reference leaves it) and, if wd1 is not NULL, wd1(nsp) the
 * distance of the nearest non-solid neighbour of each solid point.  edge is
 * the (2,netot) array fvm_geometric returns; sa_compat as in fvm_params. */
int fvm_gpu_wall_dist(int np, const double *coord, const int *ptype, int netot, const int *edge,
                      int sa_compat, double *wd, double *wd1);

/* fvm_gpu_wall_dist for callers that cannot allocate (code between graph
 * captures, tight serving loops, harnesses that forbid hipMalloc).  No
 * counterpart in the reference.  It allocates nothing: coord(2,np),
 * ptype(np), edge(2,netot) (1-based) and wd(np) are DEVICE pointers of the
 * caller, `stream` a hipStream_t (NULL: the null stream).  Its list of solid
 * edges goes to a static device array of at most 65536 edges, which each GPU
 * holds until the process ends; more gives FVM_E_SIZE, an edge end outside
 * 1..np FVM_E_INPUT.  wd is bit for bit fvm_gpu_wall_dist's (wd1 is not
 * computed).  The call waits for its stream.  All calls on one GPU share the
 * array: calls that overlap on one GPU (several host threads or streams)
 * race and must be serialised by the caller. */
int fvm_gpu_wall_dist_noalloc(int np, const double *coord, const int *ptype, int netot, const int *edge,
                              int sa_compat, double *wd, void *stream);

/* Uploads the mesh, builds the device layout and sets the free-stream state
 * (src/initialize.f:14-90 with Start 1).  One context = one solve, on the
 * current HIP device, with its own stream. */
int fvm_gpu_create(const fvm_params *par, const fvm_mesh *mesh, fvm_ctx **out);
int fvm_gpu_destroy(fvm_ctx *ctx);

/* State in the reference layout: prim(4,np) = (rho,u,v,p), nut(np).
 * set_state is the restart path (src/initialize.f:91-103) and also
 * restarts the iteration count at LastIter. */
int fvm_gpu_set_state(fvm_ctx *ctx, const double *prim, const double *nut);
/* Any output pointer may be NULL.  qx, qy, mu are those of the last residual. */
int fvm_gpu_get_state(fvm_ctx *ctx, double *prim, double *nut, double *qx, double *qy,
                      double *mu);

/* Sets the lift coefficient in COMMON /liftdrag/ (src/param.h:113-114) that
 * the farfield point-vortex correction reads (src/farfield_flux.f:33);
 * fvm_gpu_iterate updates it every iteration (src/clcd.f:46). */
int fvm_gpu_set_cl(fvm_ctx *ctx, double cl);

/* One residual on the given state: replaces
 * residu(coord,elarea,ds,prim,qx,qy,mu,mul,nut,mcarea,res) (src/residu.f:2-3).
 * res(4,np) in the reference's summation order (bitwise for the FP64 Roe and
 * HLLC paths).  Leaves the iteration state untouched. */
int fvm_gpu_residu(fvm_ctx *ctx, const double *prim, const double *nut, double *res);

/* Up to n iterations of the loop body src/main.f:44-63 (save_old, time_step,
 * 3 x [residu, update], sa_model, check_positivity, clcd, conres), replayed
 * from HIP graphs.  hist (6 x n doubles, may be NULL) receives one row
 * (iter, fres, iresi, fresi, cl, cd) per iteration done, as main.f:63
 * writes to RES.DAT.  *ndone = iterations done; *err_index = 1-based
 * vertex/edge of an error.  Returns FVM_OK (more iterations allowed),
 * FVM_CONVERGED, FVM_ITER_LIMIT or an error. */
int fvm_gpu_iterate(fvm_ctx *ctx, int n, double *hist, int *ndone, int *err_index);

/* Measurement: runs n iterations launching each kernel on its own and
 * timing every launch with HIP events on the solver's stream.  ms[k] =
 * average duration of kernel kind k, names[k] its name (up to nk kinds).
 * Returns the number of kinds. */
int fvm_gpu_time_kernels(fvm_ctx *ctx, int n, double *ms, const char **names, int nk);

/* Mean duration of the patch stage kernel over n back-to-back launches of
 * the first RK stage on the context's stream (HIP events around the n
 * launches, so per-launch event overhead is not counted).  Leaves the
 * solution untouched.  No counterpart in the reference; for the roofline. */
int fvm_gpu_time_stage(fvm_ctx *ctx, int n, double *ms);

/* Achievable HBM bandwidth on the current device: a plain copy kernel
 * (16-byte loads and stores, four workgroups per CU striding over the
 * buffers) between two fresh buffers of `bytes` each, timed over `reps`
 * launches between HIP events; *gbs = 2 * bytes per launch / mean duration,
 * in GB/s.  Allocates and frees the two buffers.  No counterpart in the
 * reference; the figure the stage kernel's bandwidth is read against beside
 * the 8 TB/s peak. */
int fvm_gpu_stream_copy(long long bytes, int reps, double *gbs);

/* Tuning knobs (no counterpart in the reference; used by the autotuner):
 * variant 3 (default) = the LDS-tiled patch stage kernel, one launch per RK
 * stage; variant 1 = edge-parallel flux into a per-edge buffer + vertex
 * gathers (three kernels per stage).  Both keep the reference's summation
 * order.  vertices_per_block in [16, 256] applies to variant 1; xcd_remap
 * gives each XCD a contiguous range of patches / vertex blocks. */
int fvm_gpu_tune(fvm_ctx *ctx, int variant, int vertices_per_block, int xcd_remap);

/* Host only (no GPU needed): plans the device layout of a mesh for the
 * LDS-tiled stage kernel and checks every index the kernels will use.
 * lds_limit: dynamic LDS bytes a workgroup may use (0 = 58 KiB; contexts use
 * the device's own limit); precision as in fvm_params.  stats[8]: patches, max patch vertices, max local
 * vertices (patch and its two rings), max local triangles, max local edges,
 * LDS bytes per workgroup, total halo vertices, failed check (0 = none). */
int fvm_patch_plan(const fvm_mesh *mesh, int iflow, int precision, int patch_size, int lds_limit,
                   long long *stats);

/* Bytes of device memory the context holds (admission by HBM footprint). */
/* ---------------------------------------------------------------- batches
 * No counterpart in the reference (it runs one solve per process).  A batch
 * advances many contexts of one kind (same iflow, iflux, ilimit, precision,
 * hllc_rotate, sa_compat; meshes, Mach numbers, angles and CFL may differ)
 * with one launch per RK stage over all their patches, iterations captured
 * in HIP graphs.  Each member computes exactly what fvm_gpu_iterate would:
 * the same operations in the same order, bit for bit.  A context belongs to
 * at most one batch; destroying it removes it. */
typedef struct fvm_batch fvm_batch;

int fvm_batch_create(fvm_batch **batch);
int fvm_batch_add(fvm_batch *batch, fvm_ctx *ctx);     /* FVM_E_INPUT: other kind */
int fvm_batch_remove(fvm_batch *batch, fvm_ctx *ctx);
int fvm_batch_size(const fvm_batch *batch);
/* Up to n iterations of every member that has not stopped (members stop on
 * their own: MinRes, MaxIter, errors).  status[i] (member i in the order
 * added, after removals): FVM_OK while running, FVM_CONVERGED,
 * FVM_ITER_LIMIT or an error code. */
int fvm_batch_iterate(fvm_batch *batch, int n, int *status);
/* The same in two halves, so one host thread can keep several batches (one
 * stream each) busy at once: launch enqueues n <= 4096 iterations and
 * returns; wait blocks until they are done and fills status as above.  Add,
 * remove and iterate wait for a launched chunk first. */
int fvm_batch_launch(fvm_batch *batch, int n);
int fvm_batch_wait(fvm_batch *batch, int *status);
/* 1 when the launched chunk has finished (or none is pending), 0 while it
 * runs; never blocks.  No counterpart in the reference: lets a serving loop
 * fill the time a chunk runs with host work. */
int fvm_batch_query(fvm_batch *batch);
int fvm_batch_destroy(fvm_batch *batch);

/* Monitors of the last iteration (src/main.f:63 RES.DAT row): mon[6] =
 * iter, fres, iresi, fresi, cl, cd.  Returns the context's status as above. */
int fvm_gpu_monitor(const fvm_ctx *ctx, double *mon);

/* Device layout of a context: info[8] = patches, max patch vertices, LDS
 * bytes per workgroup, resident workgroups per CU, CUs, patch size asked of
 * the partition (patch_size, or the automatic choice when it is 0), halo
 * vertices over all patches, kernel variant.  No counterpart in the
 * reference (its loops have no launch shape); for tuning and reports. */
int fvm_gpu_layout_info(const fvm_ctx *ctx, long long *info);

/* Threads per workgroup of the context's stage kernel (256 or 512: the
 * caller's block_threads, or the automatic choice when it is 0), or
 * FVM_E_INPUT for a handle that is not a live context.  No counterpart in
 * the reference; for tuning and reports. */
int fvm_gpu_workgroup_threads(const fvm_ctx *ctx);

long long fvm_gpu_footprint(const fvm_ctx *ctx);
/* Predicted footprint before creating a context. */
long long fvm_gpu_footprint_estimate(int np, int nt, int netot, int iflow, int precision);

const char *fvm_status_string(int status);

#ifdef __cplusplus
}
#endif
#endif
