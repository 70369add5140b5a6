// fvm2d_gpu.hip -- gfx950 kernels and the C-ABI (include/fvm2d_gpu.h) for
// fvm2d's cell-vertex residual and SSP-RK3 update.
//
// Design
//   * Vertex state is AoS double4 (rho,u,v,p): one 32-byte record per
//     gathered neighbour.  Gradients are two double4 arrays (qx, qy).
//   * One RK stage = three kernels:
//       k_grad   vertex-parallel Green-Gauss gradient (and, for viscous
//                flow, the Galerkin viscous residual), each vertex summing
//                its triangles in ascending triangle order;
//       k_flux   edge-parallel MUSCL reconstruction + Riemann flux into a
//                per-edge flux buffer; boundary edges get their own threads;
//       k_update vertex-parallel gather of the edge and boundary fluxes in
//                ascending edge order, then the RK stage in conserved
//                variables.
//     The gather order is exactly the order in which the CPU solver
//     accumulates res (src/residu.f:16-88), so the residual is deterministic,
//     needs no atomics, and is bit-identical to the CPU solver's.
//   * An iteration adds k_tstep (save_old + time_step), the SA kernel for
//     turbulent flow, and k_final (residual norms, CL/CD, stop test) in one
//     workgroup.  Iterations are replayed from HIP graphs; every kernel
//     leaves at once when the device-side stop flag is set, so a replay past
//     convergence or an error does nothing.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <numeric>
#include <vector>

#include "../../include/fvm2d_gpu.h"
#include "kernels.h"

using namespace fvm;

namespace {

constexpr int kBlock = 256;
constexpr int kHistCap = 4096;      // history rows per fvm_gpu_iterate chunk
constexpr int kGraphIters = 8;      // iterations captured per graph

struct DevState {
    int iter, done, status, err_code, err_index, nhist, iresi, pad;
    double fres, fres1, fresi, cl, cd;
};

struct Dev {
    int np, nt, ne, nsw, nff, nsp, nparts, sa1, sa2, sa3;
    const double2 *coord;
    double4 *prim, *prim_old, *qx, *qy, *resv;
    double *dt, *mu, *mul, *nut, *nut_new;
    const double *cvarea, *ais, *drmin, *wd, *elarea;
    const int *v2t_ptr, *v2t;
    const int4 *tri;
    const int2 *enode;
    const double2 *eds;
    double4 *eflux;
    const int *v2e_ptr, *v2e;
    const int2 *bnode;
    double4 *bflux;
    const int *v2b_ptr, *v2b;
    const unsigned char *solid;
    double *partial;
    DevState *S;
    double *hist;
};

// ----------------------------------------------------------------- kernels

// save_old + time_step, src/save_old.f:11-15, src/time_step.f:13-28
template <bool VISC>
__global__ void __launch_bounds__(kBlock) k_tstep(Dev d, Consts c)
{
    if (d.S->done) return;
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= d.np) return;
    double4 p = d.prim[v];
    d.prim_old[v] = p;
    double q = sqrt(sq(p.y) + sq(p.z));
    double a = sqrt(c.GAMMA * p.w / p.x);
    double dr = d.drmin[v];
    if (VISC) {
        double dtv = 2.0 * c.GAMMA * d.mu[v] / (p.x * c.prandtl);
        d.dt[v] = c.cfl * sq(dr) / (dr * (q + a) + dtv);
    } else {
        d.dt[v] = c.cfl * dr / (q + a);
    }
}

// Sutherland viscosity and total viscosity, src/residu.f:23-26
template <bool TURB>
__global__ void __launch_bounds__(kBlock) k_visc(Dev d, Consts c, const DevState *S)
{
    if (S->done) return;
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= d.np) return;
    double4 p = d.prim[v];
    double ml = sutherland_viscosity(c, p.x, p.w);
    d.mul[v] = ml;
    d.mu[v] = TURB ? ml + mu_turb(c, ml, p.x, d.nut[v]) : ml;
}

// Green-Gauss gradient, src/green_gauss.f:2-92 (tri_gg :188-218), and for
// viscous flow the Galerkin stress/heat-flux residual (tri_visc :223-285).
template <bool VISC>
__global__ void __launch_bounds__(kBlock) k_grad(Dev d, Consts c, const DevState *S)
{
    if (S->done) return;
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= d.np) return;
    double gx[4] = {0.0, 0.0, 0.0, 0.0}, gy[4] = {0.0, 0.0, 0.0, 0.0};
    double rv1 = 0.0, rv2 = 0.0, rv3 = 0.0;
    const int k1 = d.v2t_ptr[v + 1];
    for (int k = d.v2t_ptr[v]; k < k1; ++k) {
        int t = d.v2t[k];
        int4 n = d.tri[t];
        double2 x1 = d.coord[n.x], x2 = d.coord[n.y], x3 = d.coord[n.z];
        V4 p1 = ld4(d.prim, n.x), p2 = ld4(d.prim, n.y), p3 = ld4(d.prim, n.z);
        double bx[3] = {x2.y - x3.y, x3.y - x1.y, x1.y - x2.y};
        double by[3] = {x3.x - x2.x, x1.x - x3.x, x2.x - x1.x};
        double dxt[4], dyt[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            dxt[i] = p1.v[i] * bx[0] + p2.v[i] * bx[1] + p3.v[i] * bx[2];
            dyt[i] = p1.v[i] * by[0] + p2.v[i] * by[1] + p3.v[i] * by[2];
            gx[i] = gx[i] + dxt[i];
            gy[i] = gy[i] + dyt[i];
        }
        if (VISC) {
            const double n2b3 = 2.0 / 3.0;
            double e1 = p1.v[3] / p1.v[0], e2 = p2.v[3] / p2.v[0], e3 = p3.v[3] / p3.v[0];
            double ext = e1 * bx[0] + e2 * bx[1] + e3 * bx[2];
            double eyt = e1 * by[0] + e2 * by[1] + e3 * by[2];
            double um = (p1.v[1] + p2.v[1] + p3.v[1]) / 3.0;
            double vm = (p1.v[2] + p2.v[2] + p3.v[2]) / 3.0;
            double mut = (d.mu[n.x] + d.mu[n.y] + d.mu[n.z]) / 3.0;
            double mult = (d.mul[n.x] + d.mul[n.y] + d.mul[n.z]) / 3.0;
            double prim5t = (d.nut[n.x] + d.nut[n.y] + d.nut[n.z]) / 3.0;
            double rhot = (p1.v[0] + p2.v[0] + p3.v[0]) / 3.0;
            double chi = prim5t * rhot / mult;
            double chi3 = pow3(chi);
            double fv1 = chi3 / (chi3 + c.Cv11);
            double txx = n2b3 * mut * (2.0 * dxt[1] - dyt[2]);
            double txy = mut * (dyt[1] + dxt[2]);
            double tyy = n2b3 * mut * (2.0 * dyt[2] - dxt[1]);
            double gpr = c.GAMMA * (mult / c.prandtl + rhot * prim5t * fv1 / c.prandtl_turb) / c.GAMMA1;
            double efx = txx * um + txy * vm + gpr * ext;
            double efy = txy * um + tyy * vm + gpr * eyt;
            double areat = 0.25 / d.elarea[t];
            int m = (n.x == v) ? 0 : ((n.y == v) ? 1 : 2);
            double bxm = bx[m], bym = by[m];
            rv1 = rv1 + (txx * bxm + txy * bym) * areat;
            rv2 = rv2 + (txy * bxm + tyy * bym) * areat;
            rv3 = rv3 + (efx * bxm + efy * bym) * areat;
        }
    }
    double ais = d.ais[v];
    d.qx[v] = make_double4(gx[0] * ais, gx[1] * ais, gx[2] * ais, gx[3] * ais);
    d.qy[v] = make_double4(gy[0] * ais, gy[1] * ais, gy[2] * ais, gy[3] * ais);
    if (VISC) d.resv[v] = make_double4(0.0, rv1, rv2, rv3);
}

// Interior-edge flux, src/residu.f:37-66, one thread per live edge; the
// next nsw+nff threads do the solid (src/solid_flux.f) and farfield
// (src/farfield_flux.f) edges, writing one contribution per end point.
template <int FLUX, int LIM>
__global__ void __launch_bounds__(kBlock) k_flux(Dev d, Consts c, DevState *S)
{
    if (S->done) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < d.ne) {
        int2 e = d.enode[i];
        double2 ds = d.eds[i];
        double2 x1 = d.coord[e.x], x2 = d.coord[e.y];
        V4 p1 = ld4(d.prim, e.x), p2 = ld4(d.prim, e.y);
        V4 qx1 = ld4(d.qx, e.x), qy1 = ld4(d.qy, e.x), qx2 = ld4(d.qx, e.y), qy2 = ld4(d.qy, e.y);
        V4 pl, pr, f;
        recon<LIM>(c, x2.x - x1.x, x2.y - x1.y, p1, p2, qx1, qy1, qx2, qy2, pl, pr);
        if (FLUX == 1) {
            if (!roe_flux(c, ds.x, ds.y, pl, pr, f)) {
                atomicMin(&S->err_index, i + 1);       // position in the live-edge loop
                S->err_code = FVM_E_NEG_SOUND;
            }
        } else if (FLUX == 2) {
            kfvs_flux(c, ds.x, ds.y, pl, pr, f);
        } else if (FLUX == 3) {
            hllc_flux<0>(c, ds.x, ds.y, pl, pr, f);
        } else {
            hllc_flux<1>(c, ds.x, ds.y, pl, pr, f);
        }
        st4(d.eflux, i, f);
        return;
    }
    int b = i - d.ne;
    if (b >= d.nsw + d.nff) return;
    int2 e = d.bnode[b];
    double2 x1 = d.coord[e.x], x2 = d.coord[e.y];
    V4 r1, r2;
    if (b < d.nsw) {
        solid_flux(x1, x2, d.prim[e.x].w, d.prim[e.y].w, r1, r2);
    } else {
        V4 p1 = ld4(d.prim, e.x), p2 = ld4(d.prim, e.y);
        farfield_flux(c, S->cl, x1, x2, p1, p2, r1, r2);
    }
    st4(d.bflux, 2 * b, r1);
    st4(d.bflux, 2 * b + 1, r2);
}

// res(v) in the order src/residu.f builds it: viscous part, edge fluxes in
// edge order (+ on the first node, - on the second), boundary edges in
// beindx order.
template <bool VISC>
__device__ __forceinline__ V4 gather_res(const Dev &d, int v)
{
    V4 r;
    if (VISC) {
        r = ld4(d.resv, v);
    } else {
        r.v[0] = r.v[1] = r.v[2] = r.v[3] = 0.0;
    }
    const int k1 = d.v2e_ptr[v + 1];
    for (int k = d.v2e_ptr[v]; k < k1; ++k) {
        int s = d.v2e[k];
        if (s > 0) {
            double4 f = d.eflux[s - 1];
            r.v[0] = r.v[0] + f.x; r.v[1] = r.v[1] + f.y; r.v[2] = r.v[2] + f.z; r.v[3] = r.v[3] + f.w;
        } else {
            double4 f = d.eflux[-s - 1];
            r.v[0] = r.v[0] - f.x; r.v[1] = r.v[1] - f.y; r.v[2] = r.v[2] - f.z; r.v[3] = r.v[3] - f.w;
        }
    }
    const int b1 = d.v2b_ptr[v + 1];
    for (int k = d.v2b_ptr[v]; k < b1; ++k) {
        double4 f = d.bflux[d.v2b[k]];
        r.v[0] = r.v[0] + f.x; r.v[1] = r.v[1] + f.y; r.v[2] = r.v[2] + f.z; r.v[3] = r.v[3] + f.w;
    }
    return r;
}

// block-wide deterministic reduction of (sum, max, argmax) for conres
struct Red { double s, m; int i; };
__device__ __forceinline__ Red red_join(Red a, Red b)
{
    Red r;
    r.s = a.s + b.s;
    if (b.m > a.m || (b.m == a.m && b.i < a.i)) { r.m = b.m; r.i = b.i; }
    else { r.m = a.m; r.i = a.i; }
    return r;
}

__device__ Red block_reduce(Red x)
{
    __shared__ double ss[kBlock], sm[kBlock];
    __shared__ int si[kBlock];
    int t = threadIdx.x;
    ss[t] = x.s; sm[t] = x.m; si[t] = x.i;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (t < w) {
            Red a{ss[t], sm[t], si[t]}, b{ss[t + w], sm[t + w], si[t + w]};
            Red r = red_join(a, b);
            ss[t] = r.s; sm[t] = r.m; si[t] = r.i;
        }
        __syncthreads();
    }
    return Red{ss[0], sm[0], si[0]};
}

// RK stage, src/update.f:17-35 (+ src/convert.f).  LAST: the third stage,
// which also produces the conres partials and, for non-SA flow, the
// positivity check (src/check_positivity.f:11-26).
template <int IRK, bool VISC, bool LAST, bool TURB>
__global__ void __launch_bounds__(kBlock) k_update(Dev d, Consts c, DevState *S)
{
    if (S->done || S->err_code) return;
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    Red red{0.0, 0.0, 0};
    if (v < d.np) {
        V4 r = gather_res<VISC>(d, v);
        V4 po = ld4(d.prim_old, v), p = ld4(d.prim, v), u0, u1, un;
        prim2con(c, po, u0);
        prim2con(c, p, u1);
        const double a = c.airk[IRK - 1], b = c.birk[IRK - 1];
        double f = d.dt[v] / d.cvarea[v];
#pragma unroll
        for (int j = 0; j < 4; ++j) un.v[j] = a * u0.v[j] + b * (u1.v[j] - f * r.v[j]);
        if (VISC && d.solid[v]) { un.v[1] = 0.0; un.v[2] = 0.0; }
        double4 q;
        q.x = un.v[0];
        q.y = un.v[1] / un.v[0];
        q.z = un.v[2] / un.v[0];
        q.w = c.GAMMA1 * (un.v[3] - 0.5 * (sq(un.v[1]) + sq(un.v[2])) / un.v[0]);
        d.prim[v] = q;
        if (LAST) {
            double gra = q.x - po.v[0];
            double dif = fabs(gra);
            red.s = sq(gra);
            if (dif > 0.0) { red.m = dif; red.i = v + 1; }
            if (!TURB && (q.x <= 0.0 || q.w <= 0.0)) {
                atomicMin(&S->err_index, v + 1);
                S->err_code = FVM_E_NONPOSITIVE;
            }
        }
    }
    if (LAST) {
        Red r = block_reduce(red);
        if (threadIdx.x == 0) {
            d.partial[3 * blockIdx.x] = r.s;
            d.partial[3 * blockIdx.x + 1] = r.m;
            d.partial[3 * blockIdx.x + 2] = (double)r.i;
        }
    }
}

// residual only (the parity hook behind fvm_gpu_residu)
template <bool VISC>
__global__ void __launch_bounds__(kBlock) k_res(Dev d, double4 *res)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= d.np) return;
    st4(res, v, gather_res<VISC>(d, v));
}

// Spalart-Allmaras step after the RK stages, src/sa_model.f:3-201: the
// Galerkin diffusion (viscous_sa, triangle order), first-order upwind
// convection (divergence_sa, edge order), farfield inflow (outer_flux_sa)
// and the semi-implicit update (update_sa), then nu_tilde = 0 on the wall
// and the positivity check.  Reads nut, writes nut_new.
template <bool COMPAT>
__global__ void __launch_bounds__(kBlock) k_sa(Dev d, Consts c, DevState *S)
{
    if (S->done || S->err_code) return;
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < d.np) {
        const double *nut = d.nut;
        double divf = 0.0;
        double g1c = 0.0, g2c = 0.0, g3c = 0.0;
        if (COMPAT) { g1c = nut[d.sa1]; g2c = nut[d.sa2]; g3c = nut[d.sa3]; }
        for (int k = d.v2t_ptr[v]; k < d.v2t_ptr[v + 1]; ++k) {
            int t = d.v2t[k];
            int4 n = d.tri[t];
            double2 x1 = d.coord[n.x], x2 = d.coord[n.y], x3 = d.coord[n.z];
            double areat = 0.25 / d.elarea[t];
            double bx[3] = {x2.y - x3.y, x3.y - x1.y, x1.y - x2.y};
            double by[3] = {x3.x - x2.x, x1.x - x3.x, x2.x - x1.x};
            double g1 = COMPAT ? g1c : nut[n.x], g2 = COMPAT ? g2c : nut[n.y], g3 = COMPAT ? g3c : nut[n.z];
            double nux = g1 * bx[0] + g2 * bx[1] + g3 * bx[2];
            double nuy = g1 * by[0] + g2 * by[1] + g3 * by[2];
            double mult = (d.mul[n.x] + d.mul[n.y] + d.mul[n.z]) / 3.0;
            double prim5t = (nut[n.x] + nut[n.y] + nut[n.z]) / 3.0;
            double rhot = (d.prim[n.x].x + d.prim[n.y].x + d.prim[n.z].x) / 3.0;
            double mutb = mult / rhot + prim5t;
            int m = (n.x == v) ? 0 : ((n.y == v) ? 1 : 2);
            divf = divf + (c.Cb2Sig1 * mutb - c.Cb2Sig2 * (d.mul[v] / d.prim[v].x + nut[v]))
                              * (nux * bx[m] + nuy * by[m]) * areat;
        }
        for (int k = d.v2e_ptr[v]; k < d.v2e_ptr[v + 1]; ++k) {
            int s = d.v2e[k];
            int e = (s > 0 ? s : -s) - 1;
            int2 nn = d.enode[e];
            double2 ds = d.eds[e];
            double4 pa = d.prim[nn.x], pb = d.prim[nn.y];
            double unl = pa.y * ds.x + pa.z * ds.y;
            double unr = pb.y * ds.x + pb.z * ds.y;
            double una = 0.5 * (unl + unr);
            double flux = (una > 0.0) ? una * nut[nn.x] : una * nut[nn.y];
            divf = (s > 0) ? divf + flux : divf - flux;
        }
        for (int k = d.v2b_ptr[v]; k < d.v2b_ptr[v + 1]; ++k) {
            int b = d.v2b[k] >> 1;
            if (b < d.nsw) continue;
            int2 nn = d.bnode[b];
            double2 x1 = d.coord[nn.x], x2 = d.coord[nn.y];
            double sx = x2.y - x1.y, sy = -(x2.x - x1.x);
            double u = 0.5 * (d.prim[nn.x].y + d.prim[nn.y].y), w = 0.5 * (d.prim[nn.x].z + d.prim[nn.y].z);
            double un = u * sx + w * sy;
            double fl = (un > 0.0) ? un * nut[(d.v2b[k] & 1) ? nn.y : nn.x] : un * c.nut_inf;
            divf = divf + 0.5 * fl;
        }
        double4 p = d.prim[v];
        double nu = nut[v], wdv = d.wd[v], dti = d.dt[v];
        double chi = nu * p.x / d.mul[v];
        double chi3 = pow3(chi);
        double fv1 = chi3 / (chi3 + c.Cv11);
        double fv2 = 1.0 / pow3(1.0 + chi / c.Cv2);
        double r = nu / wdv / (wdv * c.kolm2);
        double Omega = d.qy[v].y - d.qx[v].z;
        double fv3 = (1.0 + chi * fv1) * (1.0 - fv2) / fmax(chi, 0.001);
        double Sv = fv3 * fabs(Omega) + r * fv2;
        r = r / Sv;
        r = fmin(r, 2.0);
        double g = r + c.Cw2 * (pow6(r) - r);
        double fw = g * pow(c.Cw31 / (pow6(g) + c.Cw32), 1.0 / 6.0);
        double source = c.Cb1 * Sv * nu;
        double fact = 1.0 + c.Cw1 * fw * (nu / wdv) * (dti / wdv);
        double nn = (nu - (dti / d.cvarea[v]) * divf + dti * source) / fact;
        if (d.solid[v]) nn = 0.0;
        d.nut_new[v] = nn;
        if (p.x <= 0.0 || p.w <= 0.0 || nn < 0.0) {
            atomicMin(&S->err_index, v + 1);
            S->err_code = FVM_E_NONPOSITIVE;
        }
    }
}

__global__ void k_copy(const double *__restrict__ a, double *__restrict__ b, int n, const DevState *S)
{
    if (S->done) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

// Monitors and the stop test of one iteration: conres (src/conres.f:11-31),
// clcd (src/clcd.f:15-46), the RES.DAT row (src/main.f:63) and the loop
// condition (src/main.f:42, :67-75).  One workgroup; fixed reduction order.
__global__ void __launch_bounds__(kBlock) k_final(Dev d, Consts c, DevState *S)
{
    if (S->done) return;
    __shared__ int s_err;
    if (threadIdx.x == 0) s_err = S->err_code;
    __syncthreads();
    if (s_err) {
        if (threadIdx.x == 0) { S->status = s_err; S->done = 1; }
        return;
    }
    int t = threadIdx.x;
    Red acc{0.0, 0.0, 0};
    for (int k = t; k < d.nparts; k += kBlock)
        acc = red_join(acc, Red{d.partial[3 * k], d.partial[3 * k + 1], (int)d.partial[3 * k + 2]});
    Red tot = block_reduce(acc);
    __syncthreads();
    // forces on the solid edges
    double xf = 0.0, yf = 0.0;
    for (int b = t; b < d.nsw; b += kBlock) {
        int2 e = d.bnode[b];
        double2 x1 = d.coord[e.x], x2 = d.coord[e.y];
        double sx = x2.y - x1.y, sy = -(x2.x - x1.x);
        double mu1 = d.mu[e.x], mu2 = d.mu[e.y];
        double4 qx1 = d.qx[e.x], qy1 = d.qy[e.x], qx2 = d.qx[e.y], qy2 = d.qy[e.y];
        double p1 = d.prim[e.x].w, p2 = d.prim[e.y].w;
        double txx1 = 2.0 / 3.0 * mu1 * (2.0 * qx1.y - qy1.z);
        double txy1 = mu1 * (qy1.y + qx1.z);
        double tyy1 = 2.0 / 3.0 * mu1 * (2.0 * qy1.z - qx1.y);
        double txx2 = 2.0 / 3.0 * mu2 * (2.0 * qx2.y - qy2.z);
        double txy2 = mu2 * (qy2.y + qx2.z);
        double tyy2 = 2.0 / 3.0 * mu2 * (2.0 * qy2.z - qx2.y);
        double p = 0.5 * (p1 + p2), txx = 0.5 * (txx1 + txx2), txy = 0.5 * (txy1 + txy2),
               tyy = 0.5 * (tyy1 + tyy2);
        xf = xf + p * sx - txx * sx - txy * sy;
        yf = yf + p * sy - txy * sx - tyy * sy;
    }
    Red fx = block_reduce(Red{xf, 0.0, 0});
    __syncthreads();
    Red fy = block_reduce(Red{yf, 0.0, 0});
    if (t != 0) return;
    double qdyn = 0.5 * c.r_inf * sq(c.q_inf);
    double cl = (-sin(c.aoa) * fx.s + cos(c.aoa) * fy.s) / qdyn;
    double cd = (cos(c.aoa) * fx.s + sin(c.aoa) * fy.s) / qdyn;
    double fres = sqrt(tot.s / c.np);
    if (S->iter == c.iterlast) {
        S->fres1 = fres;
        if (fres == 0.0) { S->status = FVM_E_ZERO_RES1; S->err_code = FVM_E_ZERO_RES1; S->done = 1; return; }
    }
    fres = fres / S->fres1;
    S->fres = fres; S->fresi = tot.m; S->iresi = tot.i; S->cl = cl; S->cd = cd;
    if (S->nhist < kHistCap) {
        double *h = d.hist + 6 * S->nhist;
        h[0] = (double)S->iter; h[1] = fres; h[2] = (double)tot.i; h[3] = tot.m; h[4] = cl; h[5] = cd;
        S->nhist++;
    }
    int it = ++S->iter;
    if (!(fres > c.minres)) { S->done = 1; S->status = FVM_CONVERGED; }
    else if (it >= c.iterlast + c.maxiter) { S->done = 1; S->status = FVM_ITER_LIMIT; }
    else if (c.saveinterval > 0 && it % c.saveinterval == 0 && it >= c.maxiter) {
        S->done = 1; S->status = FVM_ITER_LIMIT;
    }
}

__global__ void k_begin_chunk(DevState *S) { S->nhist = 0; }

}  // namespace

// ----------------------------------------------------------------- host side

struct fvm_ctx {
    fvm_params par;
    Consts c;
    Dev d;
    int np, nt, ne, netot, nsw, nff, nsp;
    hipStream_t stream = nullptr;
    std::vector<void *> allocs;
    long long bytes = 0;
    DevState *S = nullptr, *S_hook = nullptr;
    DevState hS{};
    hipGraphExec_t graph_n = nullptr, graph_1 = nullptr;
    bool graphs_ok = true;
    double4 *res_tmp = nullptr;
};

namespace {

#define HIPCHK(x)                                                                   \
    do {                                                                            \
        hipError_t e_ = (x);                                                        \
        if (e_ != hipSuccess) {                                                     \
            std::fprintf(stderr, "fvm2d_gpu: %s failed: %s (%s:%d)\n", #x,          \
                         hipGetErrorString(e_), __FILE__, __LINE__);                \
            return FVM_E_HIP;                                                       \
        }                                                                           \
    } while (0)

template <class T>
int dalloc(fvm_ctx *x, T **p, size_t n)
{
    size_t b = std::max<size_t>(n, 1) * sizeof(T);
    if (hipMalloc((void **)p, b) != hipSuccess) { *p = nullptr; return FVM_E_NOMEM; }
    x->allocs.push_back((void *)*p);
    x->bytes += (long long)b;
    return FVM_OK;
}

template <class T>
int upload(fvm_ctx *x, T **p, const std::vector<T> &h)
{
    int rc = dalloc(x, p, h.size());
    if (rc) return rc;
    if (!h.empty()) HIPCHK(hipMemcpy(*p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
    return FVM_OK;
}

void set_consts(fvm_ctx *x)
{
    Consts &c = x->c;
    const fvm_params &p = x->par;
    std::memset(&c, 0, sizeof c);
    // src/math.f:8-40
    c.GAMMA = 1.4;
    c.GAMMA1 = c.GAMMA - 1.0;
    c.M_PI_ = 4.0 * std::atan2(1.0, 1.0);
    c.prandtl = 0.72;
    c.prandtl_turb = 0.9;
    c.ALBADA11 = 2.0 / 3.0;
    c.ALBADA12 = 1.0 - c.ALBADA11;
    c.ALBADA21 = 4.0 / 3.0;
    c.ALBADA22 = 1.0 - c.ALBADA21;
    const double Cb2 = 0.622, sigma = 2.0 / 3.0, kolm = 0.41, Cw3 = 2.0, Cv1 = 7.1;
    c.Cb1 = 0.1355;
    c.Cw1 = c.Cb1 / (kolm * kolm) + (1.0 + Cb2) / sigma;
    c.Cw2 = 0.3;
    c.Cv2 = 5.0;
    { double v2 = Cv1 * Cv1; c.Cv11 = Cv1 * v2; }
    { double w2 = Cw3 * Cw3, w4 = w2 * w2; c.Cw32 = w2 * w4; c.Cw31 = 1.0 + c.Cw32; }
    c.kolm2 = kolm * kolm;
    c.Cb2Sig1 = (1.0 + Cb2) / sigma;
    c.Cb2Sig2 = Cb2 / sigma;
    // src/initialize.f:14-58
    c.q_inf = 1.0;
    c.r_inf = 1.0;
    c.mach_inf = p.mach_inf;
    c.p_inf = 1.0 / (c.GAMMA * (p.mach_inf * p.mach_inf));
    c.T_inf = c.p_inf / (c.r_inf * 1.0);
    c.aoa = p.aoa_deg * c.M_PI_ / 180.0;
    c.u_inf = c.q_inf * std::cos(c.aoa);
    c.v_inf = c.q_inf * std::sin(c.aoa);
    c.a_inf = std::sqrt(c.GAMMA * c.p_inf / c.r_inf);
    c.SCONST = 110.4 * c.T_inf / 300.0;
    c.rey = p.rey;
    c.cfl = p.cfl;
    c.xref = p.xref;
    c.yref = p.yref;
    c.minres = p.minres;
    c.airk[0] = 0.0; c.airk[1] = 3.0 / 4.0; c.airk[2] = 1.0 / 3.0;
    c.birk[0] = 1.0; c.birk[1] = 1.0 / 4.0; c.birk[2] = 2.0 / 3.0;
    c.iflow = p.iflow; c.iflux = p.iflux; c.ilimit = p.ilimit; c.vortex = p.vortex;
    c.sa_compat = p.sa_compat; c.hllc_rotate = p.hllc_rotate;
    c.iterlast = p.iterlast; c.maxiter = p.maxiter; c.saveinterval = p.saveinterval;
    c.np = x->np;
}

inline unsigned nblk(int n) { return (unsigned)((n + kBlock - 1) / kBlock); }

// host copy of the free-stream Sutherland viscosity (src/viscosity.f:245-258)
double sutherland_host(const Consts &c, double rho, double p)
{
    double temp = p / rho / 1.0;
    return std::pow(temp / c.T_inf, 1.5) * ((c.T_inf + c.SCONST) / (temp + c.SCONST)) / c.rey;
}

// launches one RK3 residual stage sequence; `S` selects the state block
typedef void (*launch_hook)(void *, int, hipEvent_t, hipEvent_t);

template <int FLUX, int LIM>
void launch_flux(fvm_ctx *x, DevState *S)
{
    unsigned g = nblk(x->ne + x->nsw + x->nff);
    hipLaunchKernelGGL((k_flux<FLUX, LIM>), dim3(g), dim3(kBlock), 0, x->stream, x->d, x->c, S);
}

void flux(fvm_ctx *x, DevState *S)
{
    int f = x->par.iflux == 3 && x->par.hllc_rotate ? 4 : x->par.iflux;
    bool lim = x->par.ilimit == 1;
    switch (f * 2 + (lim ? 1 : 0)) {
    case 2: launch_flux<1, 0>(x, S); break;
    case 3: launch_flux<1, 1>(x, S); break;
    case 4: launch_flux<2, 0>(x, S); break;
    case 5: launch_flux<2, 1>(x, S); break;
    case 6: launch_flux<3, 0>(x, S); break;
    case 7: launch_flux<3, 1>(x, S); break;
    case 8: launch_flux<4, 0>(x, S); break;
    default: launch_flux<4, 1>(x, S); break;
    }
}

// the residual kernels of one stage, up to (not including) the update
void residual_front(fvm_ctx *x, DevState *S)
{
    const int np = x->np;
    bool visc = x->par.iflow != 1, turb = x->par.iflow == 3;
    if (visc) {
        if (turb) hipLaunchKernelGGL(k_visc<true>, dim3(nblk(np)), dim3(kBlock), 0, x->stream, x->d, x->c, S);
        else hipLaunchKernelGGL(k_visc<false>, dim3(nblk(np)), dim3(kBlock), 0, x->stream, x->d, x->c, S);
        hipLaunchKernelGGL(k_grad<true>, dim3(nblk(np)), dim3(kBlock), 0, x->stream, x->d, x->c, S);
    } else {
        hipLaunchKernelGGL(k_grad<false>, dim3(nblk(np)), dim3(kBlock), 0, x->stream, x->d, x->c, S);
    }
    flux(x, S);
}

template <int IRK, bool LAST>
void launch_update(fvm_ctx *x)
{
    bool visc = x->par.iflow != 1, turb = x->par.iflow == 3;
    dim3 g(nblk(x->np)), b(kBlock);
    if (!visc) hipLaunchKernelGGL((k_update<IRK, false, LAST, false>), g, b, 0, x->stream, x->d, x->c, x->S);
    else if (!turb) hipLaunchKernelGGL((k_update<IRK, true, LAST, false>), g, b, 0, x->stream, x->d, x->c, x->S);
    else hipLaunchKernelGGL((k_update<IRK, true, LAST, true>), g, b, 0, x->stream, x->d, x->c, x->S);
}

// Kernel sequence of one iteration (src/main.f:44-63).  If `timer` is set,
// each launch is bracketed by events and its duration added to the tally.
struct Tally { std::vector<double> ms; std::vector<int> cnt; std::vector<const char *> names; };

void one_iteration(fvm_ctx *x, Tally *tally)
{
    const int np = x->np;
    bool visc = x->par.iflow != 1, turb = x->par.iflow == 3;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    if (tally) { (void)hipEventCreate(&e0); (void)hipEventCreate(&e1); }
    auto timed = [&](const char *name, auto &&fn) {
        if (!tally) { fn(); return; }
        (void)hipEventRecord(e0, x->stream);
        fn();
        (void)hipEventRecord(e1, x->stream);
        (void)hipEventSynchronize(e1);
        float ms = 0.f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        size_t k = 0;
        while (k < tally->names.size() && std::strcmp(tally->names[k], name) != 0) ++k;
        if (k == tally->names.size()) { tally->ms.push_back(0.0); tally->cnt.push_back(0); tally->names.push_back(name); }
        tally->ms[k] += ms;
        tally->cnt[k] += 1;
    };
    timed("k_tstep", [&] {
        if (visc) hipLaunchKernelGGL(k_tstep<true>, dim3(nblk(np)), dim3(kBlock), 0, x->stream, x->d, x->c);
        else hipLaunchKernelGGL(k_tstep<false>, dim3(nblk(np)), dim3(kBlock), 0, x->stream, x->d, x->c);
    });
    for (int irk = 1; irk <= 3; ++irk) {
        if (visc) {
            timed("k_visc", [&] {
                if (turb) hipLaunchKernelGGL(k_visc<true>, dim3(nblk(np)), dim3(kBlock), 0, x->stream, x->d, x->c, x->S);
                else hipLaunchKernelGGL(k_visc<false>, dim3(nblk(np)), dim3(kBlock), 0, x->stream, x->d, x->c, x->S);
            });
            timed("k_grad", [&] {
                hipLaunchKernelGGL(k_grad<true>, dim3(nblk(np)), dim3(kBlock), 0, x->stream, x->d, x->c, x->S);
            });
        } else {
            timed("k_grad", [&] {
                hipLaunchKernelGGL(k_grad<false>, dim3(nblk(np)), dim3(kBlock), 0, x->stream, x->d, x->c, x->S);
            });
        }
        timed("k_flux", [&] { flux(x, x->S); });
        timed("k_update", [&] {
            if (irk == 1) launch_update<1, false>(x);
            else if (irk == 2) launch_update<2, false>(x);
            else launch_update<3, true>(x);
        });
    }
    if (turb) {
        timed("k_sa", [&] {
            if (x->par.sa_compat) hipLaunchKernelGGL(k_sa<true>, dim3(nblk(np)), dim3(kBlock), 0, x->stream, x->d, x->c, x->S);
            else hipLaunchKernelGGL(k_sa<false>, dim3(nblk(np)), dim3(kBlock), 0, x->stream, x->d, x->c, x->S);
            hipLaunchKernelGGL(k_copy, dim3(nblk(np)), dim3(kBlock), 0, x->stream, x->d.nut_new, x->d.nut, np, x->S);
        });
    }
    timed("k_final", [&] {
        hipLaunchKernelGGL(k_final, dim3(1), dim3(kBlock), 0, x->stream, x->d, x->c, x->S);
    });
    if (tally) { (void)hipEventDestroy(e0); (void)hipEventDestroy(e1); }
}

int build_graph(fvm_ctx *x, int iters, hipGraphExec_t *out)
{
    hipGraph_t g = nullptr;
    if (hipStreamBeginCapture(x->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) return FVM_E_HIP;
    for (int i = 0; i < iters; ++i) one_iteration(x, nullptr);
    if (hipStreamEndCapture(x->stream, &g) != hipSuccess) return FVM_E_HIP;
    hipError_t e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    return e == hipSuccess ? FVM_OK : FVM_E_HIP;
}

int reset_state(fvm_ctx *x)
{
    DevState s{};
    s.iter = x->par.iterlast;
    s.fres = x->par.minres + 1.0;
    s.err_index = INT_MAX;
    if (!(x->par.maxiter > 0)) { s.done = 1; s.status = FVM_ITER_LIMIT; }
    x->hS = s;
    HIPCHK(hipMemcpyAsync(x->S, &s, sizeof s, hipMemcpyHostToDevice, x->stream));
    DevState h{};
    h.err_index = INT_MAX;
    HIPCHK(hipMemcpyAsync(x->S_hook, &h, sizeof h, hipMemcpyHostToDevice, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    return FVM_OK;
}

int check_params(const fvm_params *p)
{
    if (p->iflow < 1 || p->iflow > 3) return FVM_E_INPUT;
    if (p->iflux < 1 || p->iflux > 3) return FVM_E_INPUT;
    if (p->ilimit != 0 && p->ilimit != 1) return FVM_E_INPUT;
    if (p->vortex != 0 && p->vortex != 1) return FVM_E_INPUT;
    if (p->cell_type != 1 && p->cell_type != 2) return FVM_E_INPUT;
    if (p->precision != 0) return FVM_E_INPUT;
    return FVM_OK;
}

}  // namespace

namespace {
void drop_graphs(fvm_ctx *x)
{
    if (x->graph_n) (void)hipGraphExecDestroy(x->graph_n);
    if (x->graph_1) (void)hipGraphExecDestroy(x->graph_1);
    x->graph_n = x->graph_1 = nullptr;
}


// prim, nut in the reference layout; mu/mul follow (src/initialize.f:187-207)
int load_state(fvm_ctx *x, const double *prim, const double *nut)
{
    const int np = x->np;
    const Consts &c = x->c;
    HIPCHK(hipMemcpyAsync(x->d.prim, prim, sizeof(double) * 4 * np, hipMemcpyHostToDevice, x->stream));
    std::vector<double> nu(np, 0.0), mul(np, 0.0), mu(np, 0.0);
    if (x->par.iflow != 1) {
        for (int p = 0; p < np; ++p) {
            if (nut && x->par.iflow == 3) nu[p] = nut[p];
            double ml = sutherland_host(c, prim[4 * p], prim[4 * p + 3]);
            mul[p] = ml;
            if (x->par.iflow == 3) {
                double chi = nu[p] * prim[4 * p] / ml;
                double chi2 = chi * chi, chi3 = chi * chi2;
                double fv1 = chi3 / (chi3 + c.Cv11);
                mu[p] = ml + prim[4 * p] * nu[p] * fv1;
            } else {
                mu[p] = ml;
            }
        }
    }
    HIPCHK(hipMemcpyAsync(x->d.nut, nu.data(), sizeof(double) * np, hipMemcpyHostToDevice, x->stream));
    HIPCHK(hipMemcpyAsync(x->d.mul, mul.data(), sizeof(double) * np, hipMemcpyHostToDevice, x->stream));
    HIPCHK(hipMemcpyAsync(x->d.mu, mu.data(), sizeof(double) * np, hipMemcpyHostToDevice, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    return reset_state(x);
}

}  // namespace

extern "C" {

// Restart (src/initialize.f:91-103): that path leaves nut_inf at 0
// (src/initialize.f:153), which the farfield inflow of nu_tilde uses.
int fvm_gpu_set_state(fvm_ctx *x, const double *prim, const double *nut)
{
    x->c.nut_inf = 0.0;
    drop_graphs(x);
    return load_state(x, prim, nut);
}


const char *fvm_status_string(int s)
{
    switch (s) {
    case FVM_OK: return "ok";
    case FVM_CONVERGED: return "Residue has been reduced below MINRES";
    case FVM_ITER_LIMIT: return "Maximum number of iterations reached";
    case FVM_E_DEGENERATE: return "Fatal: triangle is degenerate";
    case FVM_E_AREA: return "Fatal: Element area is zero/negative";
    case FVM_E_ORPHAN_EDGE: return "flength: Fatal error at edge";
    case FVM_E_NEG_SOUND: return "Sonic speed is negative";
    case FVM_E_NONPOSITIVE: return "Density/pressure is negative at point";
    case FVM_E_ZERO_RES1: return "Residue in first iteration is zero";
    case FVM_E_INPUT: return "invalid input";
    case FVM_E_SIZE: return "array too small";
    case FVM_E_HIP: return "HIP runtime error";
    case FVM_E_NOMEM: return "out of device memory";
    }
    return "unknown status";
}

long long fvm_gpu_footprint_estimate(int np, int nt, int netot, int iflow, int precision)
{
    (void)precision;
    long long P = np, T = nt, E = netot;
    long long b = 0;
    b += P * (16 + 32 * 4 + 8 * 4);        // coord, prim, prim_old, qx, qy, dt, cvarea, ais, drmin
    b += P * 4 + 3 * T * 4 + T * 16;        // v2t_ptr, v2t, tri
    b += E * (8 + 16 + 32) + P * 4 + 2 * E * 4;   // enode, eds, eflux, v2e
    b += P * 5 + 4096;
    if (iflow != 1) b += P * (32 + 8 * 3) + T * 8;
    if (iflow == 3) b += P * 16;
    b += 6LL * 8 * kHistCap + 3LL * 8 * ((np + kBlock - 1) / kBlock);
    return b;
}

long long fvm_gpu_footprint(const fvm_ctx *x) { return x ? x->bytes : 0; }

int fvm_gpu_create(const fvm_params *par, const fvm_mesh *m, fvm_ctx **out)
{
    *out = nullptr;
    int rc = check_params(par);
    if (rc) return rc;
    if (m->np <= 0 || m->nt <= 0 || m->ne <= 0) return FVM_E_INPUT;
    if (par->iflow == 3 && !m->wd) return FVM_E_INPUT;
    fvm_ctx *x = new fvm_ctx();
    x->par = *par;
    const int np = m->np, nt = m->nt, ne = m->ne;
    x->np = np; x->nt = nt; x->ne = ne; x->netot = m->netot; x->nsp = m->nsp;
    x->nsw = m->nswe2 - m->nswe1 + 1;
    x->nff = m->nffe2 - m->nffe1 + 1;
    if (x->nsw < 0) x->nsw = 0;
    if (x->nff < 0) x->nff = 0;
    set_consts(x);
    if (hipStreamCreateWithFlags(&x->stream, hipStreamNonBlocking) != hipSuccess) { delete x; return FVM_E_HIP; }

    // ---- host-side layout (0-based, vertex order of the caller) ----
    std::vector<double2> coord(np);
    for (int p = 0; p < np; ++p) coord[p] = make_double2(m->coord[2 * p], m->coord[2 * p + 1]);
    std::vector<int4> tri(nt);
    for (int t = 0; t < nt; ++t)
        tri[t] = make_int4(m->elem[3 * t] - 1, m->elem[3 * t + 1] - 1, m->elem[3 * t + 2] - 1, 0);
    // triangles of each vertex, ascending
    std::vector<int> v2t_ptr(np + 1, 0), v2t(3 * (size_t)nt);
    for (int t = 0; t < nt; ++t) { ++v2t_ptr[tri[t].x + 1]; ++v2t_ptr[tri[t].y + 1]; ++v2t_ptr[tri[t].z + 1]; }
    std::partial_sum(v2t_ptr.begin(), v2t_ptr.end(), v2t_ptr.begin());
    {
        std::vector<int> f(v2t_ptr.begin(), v2t_ptr.end() - 1);
        for (int t = 0; t < nt; ++t) { v2t[f[tri[t].x]++] = t; v2t[f[tri[t].y]++] = t; v2t[f[tri[t].z]++] = t; }
    }
    // live edges in the order of the residual loop (src/residu.f:38-46)
    std::vector<int2> enode(ne);
    std::vector<double2> eds(ne);
    for (int i = 0; i < ne; ++i) {
        int ie = m->iedge[i] - 1;
        enode[i] = make_int2(m->edge[2 * ie] - 1, m->edge[2 * ie + 1] - 1);
        eds[i] = make_double2(m->ds[2 * ie], m->ds[2 * ie + 1]);
    }
    // edges of each vertex, ascending, signed (+ first node, - second node)
    std::vector<int> v2e_ptr(np + 1, 0), v2e(2 * (size_t)ne);
    for (int i = 0; i < ne; ++i) { ++v2e_ptr[enode[i].x + 1]; ++v2e_ptr[enode[i].y + 1]; }
    std::partial_sum(v2e_ptr.begin(), v2e_ptr.end(), v2e_ptr.begin());
    {
        std::vector<int> f(v2e_ptr.begin(), v2e_ptr.end() - 1);
        for (int i = 0; i < ne; ++i) { v2e[f[enode[i].x]++] = i + 1; v2e[f[enode[i].y]++] = -(i + 1); }
    }
    // boundary edges: solid then farfield, as beindx; per-vertex entries
    // 2*b (first node) or 2*b+1 (second node), in beindx order
    int nb = x->nsw + x->nff;
    std::vector<int2> bnode(std::max(nb, 1));
    for (int b = 0; b < nb; ++b) {
        int ie = m->beindx[m->nswe1 - 1 + b] - 1;
        bnode[b] = make_int2(m->edge[2 * ie] - 1, m->edge[2 * ie + 1] - 1);
    }
    std::vector<int> v2b_ptr(np + 1, 0), v2b(2 * (size_t)std::max(nb, 1));
    for (int b = 0; b < nb; ++b) { ++v2b_ptr[bnode[b].x + 1]; ++v2b_ptr[bnode[b].y + 1]; }
    std::partial_sum(v2b_ptr.begin(), v2b_ptr.end(), v2b_ptr.begin());
    {
        std::vector<int> f(v2b_ptr.begin(), v2b_ptr.end() - 1);
        for (int b = 0; b < nb; ++b) { v2b[f[bnode[b].x]++] = 2 * b; v2b[f[bnode[b].y]++] = 2 * b + 1; }
    }
    std::vector<double> ais(np), cvarea(m->cvarea, m->cvarea + np), drmin(m->drmin, m->drmin + np);
    const double n1b6 = 1.0 / 6.0;
    for (int p = 0; p < np; ++p) ais[p] = n1b6 / m->mcarea[p];
    std::vector<unsigned char> solid(np, 0);
    for (int k = 0; k < m->nsp; ++k) solid[m->spts[k] - 1] = 1;

    // ---- device buffers ----
    Dev &d = x->d;
    d.np = np; d.nt = nt; d.ne = ne; d.nsw = x->nsw; d.nff = x->nff; d.nsp = m->nsp;
    d.nparts = (int)nblk(np);
    d.sa1 = 0; d.sa2 = std::min(1, np - 1); d.sa3 = std::min(2, np - 1);
#define TRY(e) do { rc = (e); if (rc) { fvm_gpu_destroy(x); return rc; } } while (0)
    TRY(upload(x, (double2 **)&d.coord, coord));
    TRY(upload(x, (int4 **)&d.tri, tri));
    TRY(upload(x, (int **)&d.v2t_ptr, v2t_ptr));
    TRY(upload(x, (int **)&d.v2t, v2t));
    TRY(upload(x, (int2 **)&d.enode, enode));
    TRY(upload(x, (double2 **)&d.eds, eds));
    TRY(upload(x, (int **)&d.v2e_ptr, v2e_ptr));
    TRY(upload(x, (int **)&d.v2e, v2e));
    TRY(upload(x, (int2 **)&d.bnode, bnode));
    TRY(upload(x, (int **)&d.v2b_ptr, v2b_ptr));
    TRY(upload(x, (int **)&d.v2b, v2b));
    TRY(upload(x, (double **)&d.ais, ais));
    TRY(upload(x, (double **)&d.cvarea, cvarea));
    TRY(upload(x, (double **)&d.drmin, drmin));
    TRY(upload(x, (unsigned char **)&d.solid, solid));
    TRY(dalloc(x, &d.prim, np));
    TRY(dalloc(x, &d.prim_old, np));
    TRY(dalloc(x, &d.qx, np));
    TRY(dalloc(x, &d.qy, np));
    TRY(dalloc(x, &d.dt, np));
    TRY(dalloc(x, &d.eflux, ne));
    TRY(dalloc(x, &d.bflux, 2 * (size_t)std::max(nb, 1)));
    TRY(dalloc(x, &d.partial, 3 * (size_t)d.nparts));
    TRY(dalloc(x, &d.hist, 6 * (size_t)kHistCap));
    TRY(dalloc(x, &x->S, 1));
    TRY(dalloc(x, &x->S_hook, 1));
    TRY(dalloc(x, &x->res_tmp, np));
    d.S = x->S;
    TRY(dalloc(x, &d.mu, np));
    TRY(dalloc(x, &d.mul, np));
    TRY(dalloc(x, &d.nut, np));
    if (par->iflow != 1) {
        std::vector<double> elarea(m->elarea, m->elarea + nt);
        TRY(upload(x, (double **)&d.elarea, elarea));
        TRY(dalloc(x, &d.resv, np));
    }
    if (par->iflow == 3) {
        std::vector<double> wd(m->wd, m->wd + np);
        TRY(upload(x, (double **)&d.wd, wd));
        TRY(dalloc(x, &d.nut_new, np));
    }
#undef TRY

    // ---- free-stream start (src/initialize.f:60-90) ----
    const Consts &c = x->c;
    std::vector<double> prim(4 * (size_t)np), nut(np, 0.0);
    for (int p = 0; p < np; ++p) {
        prim[4 * p] = c.r_inf; prim[4 * p + 1] = c.u_inf; prim[4 * p + 2] = c.v_inf; prim[4 * p + 3] = c.p_inf;
    }
    if (par->iflow == 3) {
        double ml = sutherland_host(c, c.r_inf, c.p_inf);
        for (int p = 0; p < np; ++p) nut[p] = par->nut_factor * ml / c.r_inf;
        x->c.nut_inf = nut[0];
    }
    rc = load_state(x, prim.data(), nut.data());
    if (rc) { fvm_gpu_destroy(x); return rc; }
    *out = x;
    return FVM_OK;
}

int fvm_gpu_destroy(fvm_ctx *x)
{
    if (!x) return FVM_OK;
    if (x->stream) (void)hipStreamSynchronize(x->stream);
    drop_graphs(x);
    for (void *p : x->allocs) (void)hipFree(p);
    if (x->stream) (void)hipStreamDestroy(x->stream);
    delete x;
    return FVM_OK;
}

int fvm_gpu_set_cl(fvm_ctx *x, double cl)
{
    x->hS.cl = cl;
    DevState h = x->hS;
    HIPCHK(hipMemcpyAsync(&x->S->cl, &h.cl, sizeof(double), hipMemcpyHostToDevice, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    return FVM_OK;
}

int fvm_gpu_get_state(fvm_ctx *x, double *prim, double *nut, double *qx, double *qy, double *mu)
{
    const size_t np = (size_t)x->np;
    if (prim) HIPCHK(hipMemcpyAsync(prim, x->d.prim, 32 * np, hipMemcpyDeviceToHost, x->stream));
    if (nut) HIPCHK(hipMemcpyAsync(nut, x->d.nut, 8 * np, hipMemcpyDeviceToHost, x->stream));
    if (qx) HIPCHK(hipMemcpyAsync(qx, x->d.qx, 32 * np, hipMemcpyDeviceToHost, x->stream));
    if (qy) HIPCHK(hipMemcpyAsync(qy, x->d.qy, 32 * np, hipMemcpyDeviceToHost, x->stream));
    if (mu) HIPCHK(hipMemcpyAsync(mu, x->d.mu, 8 * np, hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    return FVM_OK;
}

int fvm_gpu_residu(fvm_ctx *x, const double *prim, const double *nut, double *res)
{
    const int np = x->np;
    HIPCHK(hipMemcpyAsync(x->d.prim, prim, sizeof(double) * 4 * np, hipMemcpyHostToDevice, x->stream));
    if (nut && x->par.iflow == 3)
        HIPCHK(hipMemcpyAsync(x->d.nut, nut, sizeof(double) * np, hipMemcpyHostToDevice, x->stream));
    // the hook keeps the iteration's error state apart, but shares cl for
    // the farfield vortex correction
    DevState h{};
    h.err_index = INT_MAX;
    h.cl = x->hS.cl;
    HIPCHK(hipMemcpyAsync(x->S_hook, &h, sizeof h, hipMemcpyHostToDevice, x->stream));
    residual_front(x, x->S_hook);
    if (x->par.iflow != 1)
        hipLaunchKernelGGL(k_res<true>, dim3(nblk(np)), dim3(kBlock), 0, x->stream, x->d, x->res_tmp);
    else
        hipLaunchKernelGGL(k_res<false>, dim3(nblk(np)), dim3(kBlock), 0, x->stream, x->d, x->res_tmp);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(res, x->res_tmp, sizeof(double) * 4 * np, hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipMemcpyAsync(&h, x->S_hook, sizeof h, hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    if (h.err_code) return h.err_code;
    return FVM_OK;
}

int fvm_gpu_iterate(fvm_ctx *x, int n, double *hist, int *ndone, int *err_index)
{
    if (ndone) *ndone = 0;
    if (err_index) *err_index = 0;
    int done_total = 0;
    while (done_total < n && !x->hS.done) {
        int chunk = std::min(n - done_total, kHistCap);
        hipLaunchKernelGGL(k_begin_chunk, dim3(1), dim3(1), 0, x->stream, x->S);
        if (x->graphs_ok && !x->graph_n) {
            if (build_graph(x, kGraphIters, &x->graph_n) || build_graph(x, 1, &x->graph_1)) {
                x->graphs_ok = false;
                (void)hipGetLastError();
            }
        }
        if (x->graphs_ok) {
            int q = chunk / kGraphIters, r = chunk % kGraphIters;
            for (int i = 0; i < q; ++i) HIPCHK(hipGraphLaunch(x->graph_n, x->stream));
            for (int i = 0; i < r; ++i) HIPCHK(hipGraphLaunch(x->graph_1, x->stream));
        } else {
            for (int i = 0; i < chunk; ++i) one_iteration(x, nullptr);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipMemcpyAsync(&x->hS, x->S, sizeof(DevState), hipMemcpyDeviceToHost, x->stream));
        HIPCHK(hipStreamSynchronize(x->stream));
        int got = x->hS.nhist;
        if (hist && got > 0)
            HIPCHK(hipMemcpy(hist + 6 * (size_t)done_total, x->d.hist, sizeof(double) * 6 * got,
                             hipMemcpyDeviceToHost));
        done_total += got;
        if (got < chunk) break;
    }
    if (ndone) *ndone = done_total;
    if (x->hS.err_code) {
        if (err_index) *err_index = x->hS.err_index;
        return x->hS.err_code;
    }
    return x->hS.done ? x->hS.status : FVM_OK;
}

int fvm_gpu_time_kernels(fvm_ctx *x, int n, double *ms, const char **names, int nk)
{
    Tally t;
    for (int i = 0; i < n && !x->hS.done; ++i) {
        hipLaunchKernelGGL(k_begin_chunk, dim3(1), dim3(1), 0, x->stream, x->S);
        one_iteration(x, &t);
        HIPCHK(hipMemcpyAsync(&x->hS, x->S, sizeof(DevState), hipMemcpyDeviceToHost, x->stream));
        HIPCHK(hipStreamSynchronize(x->stream));
    }
    int k = (int)t.ms.size();
    for (int i = 0; i < k && i < nk; ++i) {
        if (ms) ms[i] = t.cnt[i] ? t.ms[i] / t.cnt[i] : 0.0;
        if (names) names[i] = t.names[i];
    }
    return k;
}

}  // extern "C"
