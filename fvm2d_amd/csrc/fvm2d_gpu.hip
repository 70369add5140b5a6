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
#include <mutex>
#include <numeric>
#include <thread>
#include <type_traits>
#include <unordered_set>
#include <vector>

#include "../../include/fvm2d_gpu.h"
#include "kernels.h"
#include "partition.h"
#include "patch_kernel.h"

using namespace fvm;

namespace {

constexpr int kBlock = 256;
// waves per SIMD the stage kernel is compiled for (one workgroup of
// kBlock threads puts one wave on each SIMD of a CU): 4 (128 VGPRs) but for
// SA, whose instance keeps 3 (168 VGPRs; at 4 it spills and its patches shrink
// to fit four in the LDS: 0.81e9 -> 0.68e9 vertex-iterations/s on the
// 2049x513 batch, where laminar 1025x257 gains 1.00e9 -> 1.10e9)
constexpr int kPatchWavesEuler = 4, kPatchWavesLaminar = 4, kPatchWavesSA = 3, kPatchWavesSplit = 4;
template <bool VISC, bool TURB, int PREC>
constexpr int patch_waves()
{
    return TURB ? kPatchWavesSA : (VISC ? kPatchWavesLaminar : (PREC ? kPatchWavesSplit : kPatchWavesEuler));
}
constexpr int kHistCap = 4096;      // history rows per fvm_gpu_iterate chunk
constexpr int kGraphIters = 32;     // iterations captured per graph (8: 0.9 us more per
                                     // 513x129 iteration, the graph's last monitors and launch;
                                     // 64: 0.1 us less, but a batch whose members change
                                     // recaptures twice the nodes: 5% off 200-iteration requests)

struct DevState {
    int iter, done, status, err_code, err_index, nhist, iresi;
    int err_stage;              // stage kernel (1-3; 4 the SA step) that raised err_code
    double fres, fres1, fresi, cl, cd;
    // first-stage errors, kept apart from err_code until the second stage
    // (the monitors of the previous iteration may run beside the first stage)
    int err1_code, err1_index;
    // lowest vertex failing the positivity check in a last stage; the
    // monitors turn it into err_code unless an edge of the same stage failed
    // first (the reference stops in the flux, before the update runs)
    int errv_index;
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
    int vb, nblocks, xcd;             // (layout 1) vertices per workgroup; XCD remap
    const int *orig;                  // 1-based vertex number in the caller's order
    double4 *uout;                    // far-field state of each farfield edge (conserved)
    double4 *wall;                    // per solid-edge end point: p, txx, txy, tyy (last stage)
    const double2 *wall_s;            // per solid edge: (dy, -dx) of src/clcd.f:23-24
};



// ----------------------------------------------------------------- kernels

// save_old + time_step, src/save_old.f:11-15, src/time_step.f:13-28.
// SAVE: the edge-flux layout keeps a separate prim_old; the stage kernels
// read the iteration-start buffer itself and need no copy.  COPY: first the
// previous iteration's nu_tilde update (nut = nut_new), which captured SA
// iterations fold in here instead of a kernel of its own.
template <bool VISC, bool SAVE = true, bool COPY = false>
__device__ __forceinline__ void tstep_body(const Dev &d, const Consts &c, int v)
{
    if (d.S->done) return;
    if (v >= d.np) return;
    if (COPY) d.nut[v] = d.nut_new[v];
    double4 p = d.prim[v];
    if (SAVE) d.prim_old[v] = p;
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

template <bool VISC, bool SAVE = true, bool COPY = false>
__global__ void __launch_bounds__(kBlock) k_tstep(Dev d, Consts c)
{
    tstep_body<VISC, SAVE, COPY>(d, c, blockIdx.x * blockDim.x + threadIdx.x);
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
        const EdgeFrame<double> fr = edge_frame(ds.x, ds.y);
        if (FLUX == 1) {
            if (!roe_flux(c, fr, pl, pr, f)) {
                atomicMin(&S->err_index, i + 1);       // position in the live-edge loop
                S->err_code = FVM_E_NEG_SOUND;
            }
        } else if (FLUX == 2) {
            kfvs_flux(c, fr, pl, pr, f);
        } else if (FLUX == 3) {
            hllc_flux<0>(c, fr, pl, pr, f);
        } else {
            hllc_flux<1>(c, fr, pl, pr, f);
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
        farfield_flux(c, d.uout[b - d.nsw], x1, x2, p1, p2, r1, r2);
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

// Block reductions: the 64 lanes of each wavefront joined in a fixed pattern
// (every lane ends with the wavefront's total), then the wavefront totals in
// order.
// The pattern is fixed, so results repeat bit for bit; every thread returns
// the total.
__device__ __forceinline__ Red shfl_red(Red a, int o)
{
    return Red{__shfl_xor(a.s, o), __shfl_xor(a.m, o), __shfl_xor(a.i, o)};
}

// A lane's partner value under a DPP pattern (all lanes active): the xor 1
// and xor 2 quad permutations, the half-row and the row mirror.
template <int CTRL>
__device__ __forceinline__ int dpp_int(int x)
{
    return __builtin_amdgcn_update_dpp(0, x, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ double dpp_double(double x)
{
    const unsigned long long b = __builtin_bit_cast(unsigned long long, x);
    const unsigned lo = (unsigned)dpp_int<CTRL>((int)(unsigned)b);
    const unsigned hi = (unsigned)dpp_int<CTRL>((int)(unsigned)(b >> 32));
    return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
}
template <int CTRL>
__device__ __forceinline__ Red dpp_red(Red a)
{
    return Red{dpp_double<CTRL>(a.s), dpp_double<CTRL>(a.m), dpp_int<CTRL>(a.i)};
}
__device__ __forceinline__ Red lane_red(Red a, int l)
{
    const unsigned long long bs = __builtin_bit_cast(unsigned long long, a.s);
    const unsigned long long bm = __builtin_bit_cast(unsigned long long, a.m);
    auto rd = [&](unsigned long long b) {
        const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)b, l);
        const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(b >> 32), l);
        return __builtin_bit_cast(double, ((unsigned long long)hi << 32) | lo);
    };
    return Red{rd(bs), rd(bm), __builtin_amdgcn_readlane(a.i, l)};
}

// The total of a wavefront in every lane, without the LDS unit: lanes joined
// in quads (xor 1, xor 2), then half rows and rows (mirrors; each join pairs
// two groups that already hold equal values, and the join is commutative,
// so all lanes of a group stay equal), then the four rows in order.
__device__ __forceinline__ Red wave_reduce(Red x)
{
    x = red_join(x, dpp_red<0xB1>(x));      // quad_perm [1, 0, 3, 2]
    x = red_join(x, dpp_red<0x4E>(x));      // quad_perm [2, 3, 0, 1]
    x = red_join(x, dpp_red<0x141>(x));     // row_half_mirror
    x = red_join(x, dpp_red<0x140>(x));     // row_mirror
    return red_join(red_join(lane_red(x, 0), lane_red(x, 16)), red_join(lane_red(x, 32), lane_red(x, 48)));
}

// (w, lane: the thread's wave and its lane in the wave; LAST: the kernel's
// last use of the wavefront slots, which then need no closing barrier)
template <int NT, bool LAST = false>
__device__ Red block_reduce(Red x, const int w, const int lane)
{
    __shared__ double ss[NT / 64], sm[NT / 64];
    __shared__ int si[NT / 64];
    x = wave_reduce(x);
    if (lane == 0) { ss[w] = x.s; sm[w] = x.m; si[w] = x.i; }
    __syncthreads();
    Red r{ss[0], sm[0], si[0]};
#pragma unroll
    for (int k = 1; k < NT / 64; ++k) r = red_join(r, Red{ss[k], sm[k], si[k]});
    if (!LAST) __syncthreads();
    return r;
}

template <int NT>
__device__ Red block_reduce(Red x)
{
    return block_reduce<NT>(x, threadIdx.x >> 6, threadIdx.x & 63);
}

// The lane's position in its wave, computed where it is written (the
// compiler can neither hoist nor share it), so that a kernel need not keep
// its thread index live in a VGPR across a phase that needs every register.
__device__ __forceinline__ int lane_here()
{
    int l;
    asm volatile("v_mbcnt_lo_u32_b32 %0, -1, 0\n\tv_mbcnt_hi_u32_b32 %0, -1, %0" : "=v"(l));
    return l;
}

// What src/clcd.f:26-34 reads at a solid point, left by the last stage for
// every solid edge the point ends: the new pressure and, with viscosity, the
// stress of the last residual's gradients.  k_final then adds the edges in
// the reference's order without gathering anything else.
// [k0, k1) is the point's range of boundary-edge ends (v2b_ptr[v],
// v2b_ptr[v + 1]), which the caller loads early; the entries are read two at
// a time before either is written, so a point on a closed wall (two ends)
// waits for one load, not two in a row.
template <bool VISC>
__device__ __forceinline__ void wall_terms(const Dev &d, int k0, int k1, double p, double mu,
                                           double4 qx, double4 qy)
{
    double4 w = make_double4(p, 0.0, 0.0, 0.0);
    if (VISC) {
        w.y = 2.0 / 3.0 * mu * (2.0 * qx.y - qy.z);
        w.z = mu * (qy.y + qx.z);
        w.w = 2.0 / 3.0 * mu * (2.0 * qy.z - qx.y);
    }
    int k = k0;
    for (; k + 1 < k1; k += 2) {
        const int qa = d.v2b[k], qb = d.v2b[k + 1];
        if ((qa >> 1) < d.nsw) d.wall[qa] = w;
        if ((qb >> 1) < d.nsw) d.wall[qb] = w;
    }
    if (k < k1) {
        const int q = d.v2b[k];
        if ((q >> 1) < d.nsw) d.wall[q] = w;
    }
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
            if (d.solid[v])
                wall_terms<VISC>(d, d.v2b_ptr[v], d.v2b_ptr[v + 1], q.w, VISC ? d.mu[v] : 0.0, d.qx[v],
                                 d.qy[v]);
            double gra = q.x - po.v[0];
            double dif = fabs(gra);
            red.s = sq(gra);
            if (dif > 0.0) { red.m = dif; red.i = d.orig[v]; }
            if (!TURB && (q.x <= 0.0 || q.w <= 0.0)) {
                atomicMin(&S->err_index, d.orig[v]);
                S->err_code = FVM_E_NONPOSITIVE;
            }
        }
    }
    if (LAST) {
        Red r = block_reduce<kBlock>(red);
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
__device__ __forceinline__ void sa_body(const Dev &d, const Consts &c, DevState *S, int v)
{
    if (S->done || S->err_code) return;
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
            atomicMin(&S->err_index, d.orig[v]);
            S->err_stage = 4;
            S->err_code = FVM_E_NONPOSITIVE;
        }
    }
}

template <bool COMPAT>
__global__ void __launch_bounds__(kBlock) k_sa(Dev d, Consts c, DevState *S)
{
    sa_body<COMPAT>(d, c, S, blockIdx.x * blockDim.x + threadIdx.x);
}

__global__ void k_copy(const double *__restrict__ a, double *__restrict__ b, int n, const DevState *S)
{
    if (S->done) return;
    int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

// Monitors and the stop test of one iteration: conres (src/conres.f:11-31),
// clcd (src/clcd.f:15-46), the RES.DAT row (src/main.f:63) and the loop
// condition (src/main.f:42, :67-75).  One workgroup of 256 threads; fixed
// reduction order.  Without viscosity the stress terms are zero and are left
// out.  Inside captured iterations the same body also runs as one extra
// workgroup of the next iteration's first stage (a work item with patch
// index -1), so both places use the same width and the same sums.
constexpr int kFinal = 256;

// force of solid edge b on (xf, yf) from the wall terms its two end points
// left (src/clcd.f:18-43)
template <bool VISC>
__device__ __forceinline__ void wall_force(double4 w1, double4 w2, double2 sn, double &xf, double &yf)
{
    const double sx = sn.x, sy = sn.y;
    double p = 0.5 * (w1.x + w2.x);
    if (VISC) {
        double txx = 0.5 * (w1.y + w2.y), txy = 0.5 * (w1.z + w2.z), tyy = 0.5 * (w1.w + w2.w);
        xf = xf + p * sx - txx * sx - txy * sy;
        yf = yf + p * sy - txy * sx - tyy * sy;
    } else {
        xf = xf + p * sx;
        yf = yf + p * sy;
    }
}

// (sum, max, argmax) and a force pair over the first NT threads of the
// workgroup in one pass of barriers; each component follows the butterfly
// and the ordered sum of wavefront totals of block_reduce.  Every thread of
// the workgroup calls it (a larger workgroup's other waves only meet the
// barrier), so the result is the same bits in a workgroup of any size.
template <int NT>
__device__ void block_reduce_final(Red &x, double2 &f)
{
    __shared__ double ss[NT / 64], sm[NT / 64];
    __shared__ int si[NT / 64];
    __shared__ double2 s2[NT / 64];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        x = red_join(x, shfl_red(x, o));
        f.x = f.x + __shfl_xor(f.x, o);
        f.y = f.y + __shfl_xor(f.y, o);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0 && w < NT / 64) { ss[w] = x.s; sm[w] = x.m; si[w] = x.i; s2[w] = f; }
    __syncthreads();
    Red r{ss[0], sm[0], si[0]};
    double2 g = s2[0];
#pragma unroll
    for (int k = 1; k < NT / 64; ++k) {
        r = red_join(r, Red{ss[k], sm[k], si[k]});
        g.x = g.x + s2[k].x; g.y = g.y + s2[k].y;
    }
    x = r;
    f = g;
}

template <bool VISC, int NT = kFinal>
__device__ __forceinline__ void final_body(const Dev &d, const Consts &c, DevState *S)
{
    // Every load the monitors need (the solve's state words, the first
    // partial and the first solid edge of each thread) is issued before any
    // is used, so the workgroup waits for memory once; only the writes
    // depend on whether the solve is still running.
    const int done = S->done, err0 = S->err_code, errv = S->errv_index;
    const int iter = S->iter, nhist = S->nhist;
    const double fres1 = S->fres1;
    // the first kFinal threads do the work in a workgroup of NT threads
    const int t = threadIdx.x;
    const bool act = NT == kFinal || t < kFinal;
    Red acc{0.0, 0.0, 0};
    double xf = 0.0, yf = 0.0;
    {
        const bool hp = act && t < d.nparts, hw = act && t < d.nsw;
        double ps = 0.0, pm = 0.0, pi = 0.0;
        double4 w1 = make_double4(0.0, 0.0, 0.0, 0.0), w2 = w1;
        double2 sn = make_double2(0.0, 0.0);
        if (hp) { ps = d.partial[3 * t]; pm = d.partial[3 * t + 1]; pi = d.partial[3 * t + 2]; }
        if (hw) { w1 = d.wall[2 * t]; w2 = d.wall[2 * t + 1]; sn = d.wall_s[t]; }
        if (hp) acc = red_join(acc, Red{ps, pm, (int)pi});
        if (hw) wall_force<VISC>(w1, w2, sn, xf, yf);
    }
#pragma unroll 1
    for (int k = t + kFinal; act && k < d.nparts; k += kFinal)
        acc = red_join(acc, Red{d.partial[3 * k], d.partial[3 * k + 1], (int)d.partial[3 * k + 2]});
#pragma unroll 1
    for (int b = t + kFinal; act && b < d.nsw; b += kFinal)
        wall_force<VISC>(d.wall[2 * b], d.wall[2 * b + 1], d.wall_s[b], xf, yf);
    double2 f = make_double2(xf, yf);
    block_reduce_final<kFinal>(acc, f);
    if (done || t != 0) return;
    int err = err0;
    if (!err && errv != INT_MAX) {
        err = FVM_E_NONPOSITIVE;
        S->err_index = errv;
        S->err_code = err;
    }
    if (err) { S->status = err; S->done = 1; return; }
    const Red tot = acc;
    double qdyn = 0.5 * c.r_inf * sq(c.q_inf);
    double cl = (-c.sin_aoa * f.x + c.cos_aoa * f.y) / qdyn;
    double cd = (c.cos_aoa * f.x + c.sin_aoa * f.y) / qdyn;
    double fres = sqrt(tot.s / c.np);
    double f1 = fres1;
    if (iter == c.iterlast) {
        S->fres1 = f1 = fres;
        if (fres == 0.0) { S->status = FVM_E_ZERO_RES1; S->err_code = FVM_E_ZERO_RES1; S->done = 1; return; }
    }
    fres = fres / f1;
    S->fres = fres; S->fresi = tot.m; S->iresi = tot.i; S->cl = cl; S->cd = cd;
    if (nhist < kHistCap) {
        double *h = d.hist + 6 * nhist;
        h[0] = (double)iter; h[1] = fres; h[2] = (double)tot.i; h[3] = tot.m; h[4] = cl; h[5] = cd;
        S->nhist = nhist + 1;
    }
    int it = iter + 1;
    S->iter = it;
    if (!(fres > c.minres)) { S->done = 1; S->status = FVM_CONVERGED; }
    else if (it >= c.iterlast + c.maxiter) { S->done = 1; S->status = FVM_ITER_LIMIT; }
    else if (c.saveinterval > 0 && it % c.saveinterval == 0 && it >= c.maxiter) {
        S->done = 1; S->status = FVM_ITER_LIMIT;
    }
}

template <bool VISC>
__global__ void __launch_bounds__(kFinal) k_final(Dev d, Consts c, DevState *S)
{
    final_body<VISC>(d, c, S);
}


// ------------------------------------------------- launch helpers

// Workgroups are handed to the 8 XCDs round robin; with `xcd` set, each XCD
// gets a contiguous range of vertex blocks so neighbouring blocks share its L2.
__device__ __forceinline__ int logical_block(int nb, int xcd)
{
    int b = blockIdx.x;
    if (!xcd) return b;
    int q = nb >> 3, r = nb & 7, x = b & 7, k = b >> 3;
    return x * q + min(x, r) + k;
}

// ------------------------------------------------- patch stage kernel

// Green-Gauss term of one triangle, src/green_gauss.f:97-127: for each
// variable, sum over the corners of q_k (dy, -dx) of the opposite side.
// Descriptors, patch headers and work lists are read through the constant address space:
// they do not change during a launch, so the compiler treats them like
// kernel arguments (scalar loads, reloaded where needed instead of held in
// vector registers across the kernel's stores).
template <class T>
__device__ __forceinline__ const T &uniform_at(const T *p, int i)
{
    using CT = __attribute__((address_space(4))) const T;
    return *(const T *)((const CT *)p + i);
}

__device__ __forceinline__ void tri_gg(const double2 *s_coord, const double4 *s_prim, short4 tr,
                                       V4 &p1, V4 &p2, V4 &p3, double bx[3], double by[3],
                                       double dxt[4], double dyt[4])
{
    double2 x1 = s_coord[tr.x], x2 = s_coord[tr.y], x3 = s_coord[tr.z];
    double4 a1 = s_prim[tr.x], a2 = s_prim[tr.y], a3 = s_prim[tr.z];
    p1 = V4{{a1.x, a1.y, a1.z, a1.w}}; p2 = V4{{a2.x, a2.y, a2.z, a2.w}}; p3 = V4{{a3.x, a3.y, a3.z, a3.w}};
    bx[0] = x2.y - x3.y; bx[1] = x3.y - x1.y; bx[2] = x1.y - x2.y;
    by[0] = x3.x - x2.x; by[1] = x1.x - x3.x; by[2] = x2.x - x1.x;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        dxt[i] = p1.v[i] * bx[0] + p2.v[i] * bx[1] + p3.v[i] * bx[2];
        dyt[i] = p1.v[i] * by[0] + p2.v[i] * by[1] + p3.v[i] * by[2];
    }
}

template <int FLUX, int LIM, bool VISC, int MODE, bool TURB, int PREC, int WG = kBlock>
__device__ __forceinline__ void patch_body(const Dev &d, const Consts &c, DevState *S,
                                           const PatchArgs &pa, const int b, const int irk,
                                           const double4 *__restrict__ pin,
                                           double4 *__restrict__ pout, double4 *res_out)
{
    // The solve's status words are loaded with the patch header and tested
    // after the gather, which writes only LDS: their latency hides under it.
    // The header is read through the constant address space (scalar loads).
    //
    // A workgroup leaves when the solve is done or an earlier stage kernel
    // raised an error; an error of this same stage does not stop the others,
    // so every workgroup reports its lowest failing edge or vertex and the
    // index is the reference's (the first in its loop order), whatever the
    // order in which workgroups run.  The second stage first moves a
    // first-stage error to err_code.
    const int es = S->err_stage;
    const int stop = S->done | (es != 0 && es != irk);
    const int e1 = (MODE == 0 && irk == 2) ? S->err1_code : 0;
    const int hb = kPatchHdr * b;
    auto hdr = [&](int k) { return uniform_at(pa.hdr, hb + k); };
    const int v0 = hdr(PH_V0), NP = hdr(PH_NP), NL = hdr(PH_NL), NG = hdr(PH_NG);
    const int NT = hdr(PH_NTRI), NE = hdr(PH_NEDGE), NB = hdr(PH_NBND);
    const int NGTL = hdr(PH_NGTL), NTP = VISC ? hdr(PH_NTP) : 0;
    extern __shared__ double4 smem4[];
    double *smem = reinterpret_cast<double *>(smem4);
    const int NLa = (NL + 1) & ~1;               // keeps every LDS array 16-byte aligned
    const int NGa = (NG + 1) & ~1;
    const int *halo = pa.halo + hdr(PH_HALO);
    const short4 *ltri = pa.ltri + hdr(PH_TRI);
    const double *lqinv = pa.lqinv + hdr(PH_TRI);
    const short2 *ledge = pa.ledge + hdr(PH_EDGE);
    const int *ledge_id = pa.ledge_id + hdr(PH_EDGE);
    const double4 *ledge_fr = pa.ledge_fr + hdr(PH_EDGE);
    const short2 *lbnd = pa.lbnd + hdr(PH_BND);
    const int *lbnd_id = pa.lbnd_id + hdr(PH_BND);
    const int t = threadIdx.x;
    const int wave = __builtin_amdgcn_readfirstlane(t >> 6);   // a scalar
    // Issue priority: a CU's workgroups are issued oldest first, so the
    // youngest ones wait behind the others' edge fluxes.  The phases bound
    // by load and dependency latency (gather, gradients, ordered sums and
    // update) run at priority 1 and the edge fluxes, which have the VALU
    // work to fill every gap, at 0 (measured on 513x129 Euler/Roe: FP64
    // 49.6 -> 49.0 us per iteration, split precision 38.8 -> 37.4).
    __builtin_amdgcn_s_setprio(1);

    // LDS carve-up for this patch
    double4 *s_prim = reinterpret_cast<double4 *>(smem);
    double2 *s_coord = reinterpret_cast<double2 *>(s_prim + NLa);
    // per local vertex: viscosity and p/rho (which the stress terms would
    // otherwise divide out per triangle); SA adds the laminar viscosity
    // (s_mul) and nu_tilde (s_nut); laminar: s_mul is s_mu
    double *s_mu = reinterpret_cast<double *>(s_coord + NLa);
    double *s_e = s_mu + (VISC ? NLa : 0);
    double *s_mul = TURB ? s_e + NLa : s_mu;
    double *s_nut = s_mul + (TURB ? NLa : 0);
    // gradients in LDS: FP64, or rounded to FP32 in the split mode, whose
    // reconstruction reads them in FP32 (the FP64 values go to qx, qy)
    using GT = typename std::conditional<PREC == 1, float4, double4>::type;
    // inviscid split precision: an FP32 copy of the state and FP32
    // coordinates relative to the patch's first vertex, for the FP32
    // gradients and reconstruction (in the bytes the viscous arrays take)
    constexpr bool F32G = PREC == 1 && !VISC;
    float4 *s_p32 = reinterpret_cast<float4 *>(s_mu);
    float2 *s_c32 = reinterpret_cast<float2 *>(s_p32 + NLa);
    GT *s_gx = reinterpret_cast<GT *>(s_mu + NLa * (TURB ? 4 : (VISC ? 2 : (F32G ? 3 : 0))));
    GT *s_gy = s_gx + NG;
    // (with viscosity the gradients' bytes, at least 40 per triangle touching
    // P, first stage the stress terms); then the gradients' area factor ais
    // of G, gathered with the state so that the gradient phase reads no
    // global memory
    const int GRB = ((VISC && NTP * 40 > NG * (int)(2 * sizeof(GT))) ? NTP * 40 : NG * (int)(2 * sizeof(GT)));
    double *s_ais = reinterpret_cast<double *>(reinterpret_cast<char *>(s_gx) + ((GRB + 15) & ~15));
    double *s_work = s_ais + NGa;
    // work area: first the local triangles (and, with viscosity, 0.25 over
    // their areas), whose terms each vertex recomputes from LDS in its ordered sum
    // (same operands in the same order, so the same bits as terms computed
    // once); then the edge fluxes
    constexpr int TW = VISC ? 2 : 1;           // doubles per local triangle
    short4 *s_ltri = reinterpret_cast<short4 *>(s_work);
    double *s_lqinv = s_work + NT;
    const int WB = (max(NT * TW * 8, (NE + 2 * NB) * 32) + 15) & ~15;
    int *s_gptr = reinterpret_cast<int *>(reinterpret_cast<char *>(s_work) + WB);
    int *s_pptr = s_gptr + NG + 1;
    short *s_gtl = reinterpret_cast<short *>(s_pptr + NP + 1);
    short *s_psl = s_gtl + NGTL;

    // 1. stage state of P, H1, H2, and the patch's tables.  Every load that
    // needs only the header is issued before any waits: the index tables
    // (one 16-byte block a thread: gptr, pptr, gtl, pslot as laid out in
    // LDS) and, inviscid, the local triangles (two a load; each patch's
    // start 16-byte aligned); then the state, whose halo part waits for the
    // halo ids.
    const int4 *blob = pa.blob + hdr(PH_BLOB);
    const int NBLOB = hdr(PH_NBLOB), NT2 = (NT + 1) >> 1;
    const int4 *ltri4 = reinterpret_cast<const int4 *>(ltri);
    int4 *s_blob = reinterpret_cast<int4 *>(s_gptr);
    int4 *s_ltri4 = reinterpret_cast<int4 *>(s_ltri);
    // (clamped indices: plain global loads; a choice between the table and
    // a zero vector compiles to a flat load from either address, the zero
    // vector stored to scratch first: 0.6 us more per stage on 513x129)
    const int4 bl0 = blob[min(t, NBLOB - 1)];
    const int4 tr0 = VISC ? make_int4(0, 0, 0, 0) : ltri4[min(t, NT2 - 1)];
    // inviscid split precision: the FP32 state is relative to the patch's
    // first vertex (as the coordinates are), so the FP32 Green-Gauss sums
    // cancel no free-stream pressure
    const double2 xref = F32G ? uniform_at(d.coord, v0) : make_double2(0.0, 0.0);
    const double4 pref = F32G ? pin[v0] : make_double4(0.0, 0.0, 0.0, 0.0);
    for (int i = t; i < NL; i += WG) {
        const int gv = i < NP ? v0 + i : halo[i - NP];
        double4 p = pin[gv];
        const double2 xy = d.coord[gv];
        s_prim[i] = p;
        s_coord[i] = xy;
        if (F32G) {
            s_p32[i] = make_float4((float)(p.x - pref.x), (float)(p.y - pref.y), (float)(p.z - pref.z),
                                   (float)(p.w - pref.w));
            s_c32[i] = make_float2((float)(xy.x - xref.x), (float)(xy.y - xref.y));
        }
        if (i < NG) s_ais[i] = d.ais[gv];
        if (VISC) {
            double ml = sutherland_viscosity(c, p.x, p.w);
            s_e[i] = p.w / p.x;
            if (TURB) {
                double nu = d.nut[gv];
                s_mul[i] = ml; s_mu[i] = ml + mu_turb(c, ml, p.x, nu); s_nut[i] = nu;
            } else {
                s_mu[i] = ml;
            }
        }
    }
    if (t < NBLOB) s_blob[t] = bl0;
    for (int i = t + WG; i < NBLOB; i += WG) s_blob[i] = blob[i];
    if (VISC) {
        for (int i = t; i < NT; i += WG) {
            s_ltri[i] = ltri[i];
            s_lqinv[i] = lqinv[i];
        }
    } else {
        if (t < NT2) s_ltri4[t] = tr0;
        for (int i = t + WG; i < NT2; i += WG) s_ltri4[i] = ltri4[i];
    }
    __syncthreads();
    if (stop) return;
    if (e1) {
        if (t == 0) {
            atomicMin(&S->err_index, S->err1_index);
            S->err_stage = 1;
            S->err_code = e1;
        }
        return;
    }

    // Solid and farfield fluxes need no gradient: they are evaluated during
    // the gradient phase by threads from the top of the workgroup (which the
    // gradients, one vertex of G per thread, leave idle), unless their LDS
    // slots, after the NE edge terms, still hold local triangles then.
    auto bnd_flux = [&](const int kb) {
        short2 e = lbnd[kb];
        const int gb = lbnd_id[kb];
        double2 x1 = s_coord[e.x], x2 = s_coord[e.y];
        V4 r1, r2;
        if (gb < d.nsw) {
            solid_flux(x1, x2, s_prim[e.x].w, s_prim[e.y].w, r1, r2);
        } else {
            double4 a1 = s_prim[e.x], a2 = s_prim[e.y];
            V4 p1{{a1.x, a1.y, a1.z, a1.w}}, p2{{a2.x, a2.y, a2.z, a2.w}};
            farfield_flux(c, d.uout[gb - d.nsw], x1, x2, p1, p2, r1, r2);
        }
        double4 *fw = reinterpret_cast<double4 *>(s_work) + NE + 2 * kb;
        fw[0] = make_double4(r1.v[0], r1.v[1], r1.v[2], r1.v[3]);
        fw[1] = make_double4(r2.v[0], r2.v[1], r2.v[2], r2.v[3]);
    };
    const bool bnd_early = NE * 32 >= NT * TW * 8;
    if (bnd_early)
        for (int kb = WG - 1 - t; kb < NB; kb += WG) bnd_flux(kb);

    // 2. gradients of G (src/green_gauss.f): each vertex sums its triangles'
    // terms in ascending triangle order, each term recomputed from LDS.
    //
    // With viscosity, first the Galerkin stress terms (src/green_gauss.f:
    // 132-194): (a) each triangle touching P (local numbers below NTP, as the
    // triangles are numbered while P's vertices are scanned) evaluates its
    // stresses and heat flux once and stages them in the gradients' LDS;
    // (b) each vertex of P adds, in ascending triangle order, each triangle's
    // stresses against its own side (bx[m], by[m]) times 0.25/area: the
    // operations of the per-triangle loop of the reference, term for term.
    if (VISC) {
        double *s_vt = reinterpret_cast<double *>(s_gx);          // 5 per triangle
        for (int k = t; k < NTP; k += WG) {
            const short4 tr = s_ltri[k];
            V4 p1, p2, p3;
            double bx[3], by[3], dxt[4], dyt[4];
            tri_gg(s_coord, s_prim, tr, p1, p2, p3, bx, by, dxt, dyt);
            const double n2b3 = 2.0 / 3.0;
            const double e1 = s_e[tr.x], e2 = s_e[tr.y], e3 = s_e[tr.z];
            double ext = e1 * bx[0] + e2 * bx[1] + e3 * bx[2];
            double eyt = e1 * by[0] + e2 * by[1] + e3 * by[2];
            double um = (p1.v[1] + p2.v[1] + p3.v[1]) / 3.0;
            double vm = (p1.v[2] + p2.v[2] + p3.v[2]) / 3.0;
            double mut = (s_mu[tr.x] + s_mu[tr.y] + s_mu[tr.z]) / 3.0;
            double txx = n2b3 * mut * (2.0 * dxt[1] - dyt[2]);
            double txy = mut * (dyt[1] + dxt[2]);
            double tyy = n2b3 * mut * (2.0 * dyt[2] - dxt[1]);
            double gpr;
            if (TURB) {
                double mult = (s_mul[tr.x] + s_mul[tr.y] + s_mul[tr.z]) / 3.0;
                double prim5t = (s_nut[tr.x] + s_nut[tr.y] + s_nut[tr.z]) / 3.0;
                double rhot = (p1.v[0] + p2.v[0] + p3.v[0]) / 3.0;
                double chi = prim5t * rhot / mult;
                double chi3 = pow3(chi);
                double fv1 = chi3 / (chi3 + c.Cv11);
                gpr = c.GAMMA * (mult / c.prandtl + rhot * prim5t * fv1 / c.prandtl_turb) / c.GAMMA1;
            } else {
                // nu_tilde is zero, so the eddy term is +0 and the laminar
                // viscosity is the total one: the same value, fewer operations
                gpr = c.GAMMA * (mut / c.prandtl) / c.GAMMA1;
            }
            double *o = s_vt + 5 * k;
            o[0] = txx; o[1] = txy; o[2] = tyy;
            o[3] = txx * um + txy * vm + gpr * ext;
            o[4] = txy * um + tyy * vm + gpr * eyt;
        }
        __syncthreads();
        // each P vertex's viscous residual waits for the update in LDS
        // (after the index tables), not in registers through the edge fluxes
        double *s_rv = reinterpret_cast<double *>(reinterpret_cast<char *>(s_gptr) + 16 * NBLOB);
        for (int g = t; g < NP; g += WG) {
            double rv1 = 0.0, rv2 = 0.0, rv3 = 0.0;
            const int k1 = s_gptr[g + 1];
            for (int k = s_gptr[g]; k < k1; ++k) {
                const int e = s_gtl[k], tri = e >> 2, m = e & 3;
                const short4 tr = s_ltri[tri];
                // the side opposite corner m: bx[m] = y(m+1) - y(m+2),
                // by[m] = x(m+2) - x(m+1)
                const int ia = m == 0 ? tr.y : (m == 1 ? tr.z : tr.x);
                const int ib = m == 0 ? tr.z : (m == 1 ? tr.x : tr.y);
                const double2 xa = s_coord[ia], xb = s_coord[ib];
                const double bxm = xa.y - xb.y, bym = xb.x - xa.x;
                const double areat = s_lqinv[tri];       // 0.25 / elarea
                const double *o = s_vt + 5 * tri;
                rv1 = rv1 + (o[0] * bxm + o[1] * bym) * areat;
                rv2 = rv2 + (o[1] * bxm + o[2] * bym) * areat;
                rv3 = rv3 + (o[3] * bxm + o[4] * bym) * areat;
            }
            s_rv[3 * g] = rv1;
            s_rv[3 * g + 1] = rv2;
            s_rv[3 * g + 2] = rv3;
        }
        __syncthreads();                 // the gradients overwrite the staged terms
    }
    if (F32G) {
        // FP32 gradients for the FP32 reconstruction: per vertex its
        // triangles' Green-Gauss terms in ascending order, from the FP32
        // state and patch-relative coordinates (differences of nearby
        // coordinates, so the normals keep their relative accuracy)
        for (int g = t; g < NG; g += WG) {
            const float ais = (float)s_ais[g];
            float4 gx = make_float4(0.f, 0.f, 0.f, 0.f), gy = gx;
            const int k1 = s_gptr[g + 1];
            for (int k = s_gptr[g]; k < k1; ++k) {
                const short4 tr = s_ltri[s_gtl[k] >> 2];
                const float2 c1 = s_c32[tr.x], c2 = s_c32[tr.y], c3 = s_c32[tr.z];
                const float4 a1 = s_p32[tr.x], a2 = s_p32[tr.y], a3 = s_p32[tr.z];
                const float bx0 = c2.y - c3.y, bx1 = c3.y - c1.y, bx2 = c1.y - c2.y;
                const float by0 = c3.x - c2.x, by1 = c1.x - c3.x, by2 = c2.x - c1.x;
                gx.x = fmaf(a3.x, bx2, fmaf(a2.x, bx1, fmaf(a1.x, bx0, gx.x)));
                gx.y = fmaf(a3.y, bx2, fmaf(a2.y, bx1, fmaf(a1.y, bx0, gx.y)));
                gx.z = fmaf(a3.z, bx2, fmaf(a2.z, bx1, fmaf(a1.z, bx0, gx.z)));
                gx.w = fmaf(a3.w, bx2, fmaf(a2.w, bx1, fmaf(a1.w, bx0, gx.w)));
                gy.x = fmaf(a3.x, by2, fmaf(a2.x, by1, fmaf(a1.x, by0, gy.x)));
                gy.y = fmaf(a3.y, by2, fmaf(a2.y, by1, fmaf(a1.y, by0, gy.y)));
                gy.z = fmaf(a3.z, by2, fmaf(a2.z, by1, fmaf(a1.z, by0, gy.z)));
                gy.w = fmaf(a3.w, by2, fmaf(a2.w, by1, fmaf(a1.w, by0, gy.w)));
            }
            const GT qx{gx.x * ais, gx.y * ais, gx.z * ais, gx.w * ais};
            const GT qy{gy.x * ais, gy.y * ais, gy.z * ais, gy.w * ais};
            s_gx[g] = qx;
            s_gy[g] = qy;
            if (MODE == 2 && g < NP) {
                d.qx[v0 + g] = make_double4(qx.x, qx.y, qx.z, qx.w);
                d.qy[v0 + g] = make_double4(qy.x, qy.y, qy.z, qy.w);
            }
        }
    } else
    for (int g = t; g < NG; g += WG) {
        const double ais = s_ais[g];
        double gx[4] = {0.0, 0.0, 0.0, 0.0}, gy[4] = {0.0, 0.0, 0.0, 0.0};
        const int k1 = s_gptr[g + 1];
        for (int k = s_gptr[g]; k < k1; ++k) {
            const int e = s_gtl[k];
            const short4 tr = s_ltri[e >> 2];
            V4 p1, p2, p3;
            double bx[3], by[3], dxt[4], dyt[4];
            tri_gg(s_coord, s_prim, tr, p1, p2, p3, bx, by, dxt, dyt);
#pragma unroll
            for (int i = 0; i < 4; ++i) { gx[i] = gx[i] + dxt[i]; gy[i] = gy[i] + dyt[i]; }
        }
        double4 qx = make_double4(gx[0] * ais, gx[1] * ais, gx[2] * ais, gx[3] * ais);
        double4 qy = make_double4(gy[0] * ais, gy[1] * ais, gy[2] * ais, gy[3] * ais);
        if (PREC == 1) {
            s_gx[g] = GT{(float)qx.x, (float)qx.y, (float)qx.z, (float)qx.w};
            s_gy[g] = GT{(float)qy.x, (float)qy.y, (float)qy.z, (float)qy.w};
        } else {
            s_gx[g] = GT{qx.x, qx.y, qx.z, qx.w};
            s_gy[g] = GT{qy.x, qy.y, qy.z, qy.w};
        }
        // the gradients a caller may read: the residual hook's, and the last
        // stage's where the SA step or the wall stresses read them; without
        // viscosity the host recomputes them on request (same operations)
        if ((MODE == 2 || (MODE == 1 && VISC)) && g < NP) { d.qx[v0 + g] = qx; d.qy[v0 + g] = qy; }
    }
    __syncthreads();

    // 3. fluxes of the edges and boundary edges touching P
    __builtin_amdgcn_s_setprio(0);
    int bad = INT_MAX;
    for (int k = t; k < NE; k += WG) {
        short2 e = ledge[k];
        const double4 fr = ledge_fr[k];
        double2 x1 = s_coord[e.x], x2 = s_coord[e.y];
        double4 a1 = s_prim[e.x], a2 = s_prim[e.y];
        const GT g1 = s_gx[e.x], h1 = s_gy[e.x], g2 = s_gx[e.y], h2 = s_gy[e.y];
        V4 p1{{a1.x, a1.y, a1.z, a1.w}}, p2{{a2.x, a2.y, a2.z, a2.w}};
        V4 f;
        bool ok = true;
        if (PREC == 0) {
            V4 qx1{{g1.x, g1.y, g1.z, g1.w}}, qy1{{h1.x, h1.y, h1.z, h1.w}};
            V4 qx2{{g2.x, g2.y, g2.z, g2.w}}, qy2{{h2.x, h2.y, h2.z, h2.w}};
            V4 pl, pr;
            recon<LIM, double, FLUX == 1 && !VISC>(c, x2.x - x1.x, x2.y - x1.y, p1, p2, qx1, qy1, qx2, qy2, pl, pr);
            const EdgeFrame<double> efr{fr.x, fr.y, fr.z};
            if (FLUX == 1) ok = roe_flux(c, efr, pl, pr, f);
            else if (FLUX == 2) kfvs_flux(c, efr, pl, pr, f);
            else if (FLUX == 3) hllc_flux<0>(c, efr, pl, pr, f);
            else hllc_flux<1>(c, efr, pl, pr, f);
        } else {
            const V4f qx1{{(float)g1.x, (float)g1.y, (float)g1.z, (float)g1.w}};
            const V4f qy1{{(float)h1.x, (float)h1.y, (float)h1.z, (float)h1.w}};
            const V4f qx2{{(float)g2.x, (float)g2.y, (float)g2.z, (float)g2.w}};
            const V4f qy2{{(float)h2.x, (float)h2.y, (float)h2.z, (float)h2.w}};
            float dxf, dyf;
            if (F32G) {
                const float2 c1 = s_c32[e.x], c2 = s_c32[e.y];
                dxf = c2.x - c1.x; dyf = c2.y - c1.y;
            } else {
                dxf = (float)(x2.x - x1.x); dyf = (float)(x2.y - x1.y);
            }
            if (FLUX == 1) {
                // split precision, Roe: FP32 increments and dissipation on
                // FP64 states and consistent flux (recon_incr, roe_flux_split)
                const V4f dq{{(float)(p2.v[0] - p1.v[0]), (float)(p2.v[1] - p1.v[1]),
                              (float)(p2.v[2] - p1.v[2]), (float)(p2.v[3] - p1.v[3])}};
                V4f dl, dr;
                recon_incr<LIM>(c, dxf, dyf, dq, qx1, qy1, qx2, qy2, dl, dr);
                V4 pl, pr;
#pragma unroll
                for (int i = 0; i < 4; ++i) {
                    pl.v[i] = p1.v[i] + (double)dl.v[i];
                    pr.v[i] = p2.v[i] - (double)dr.v[i];
                }
                ok = roe_flux_split(c, EdgeFrame<double>{fr.x, fr.y, fr.z}, pl, pr, f);
                if (!ok) bad = min(bad, ledge_id[k] + 1);
                reinterpret_cast<double4 *>(s_work)[k] = make_double4(f.v[0], f.v[1], f.v[2], f.v[3]);
                continue;
            }
            // split precision, KFVS and HLLC: reconstruction and flux in FP32
            // on the FP64 states rounded (edge vector from FP64 coordinates,
            // edge frame rounded from FP64); the flux returns to FP64 for the
            // accumulation
            V4f pl, pr, ff;
            const V4f f1 = to_f(p1), f2 = to_f(p2);
            recon<LIM, float>(c, dxf, dyf, f1, f2, qx1, qy1, qx2, qy2, pl, pr);
            const EdgeFrame<float> efr{(float)fr.x, (float)fr.y, (float)fr.z};
            if (FLUX == 2) kfvs_flux<float>(c, efr, pl, pr, ff);
            else if (FLUX == 3) hllc_flux<0, float>(c, efr, pl, pr, ff);
            else hllc_flux<1, float>(c, efr, pl, pr, ff);
            f = to_d(ff);
        }
        if (!ok) bad = min(bad, ledge_id[k] + 1);
        reinterpret_cast<double4 *>(s_work)[k] = make_double4(f.v[0], f.v[1], f.v[2], f.v[3]);
    }
    // (in the last stage the thread index is rebuilt past the edge loop from
    // the wave and the lane: kept live across the edge fluxes it was the one
    // VGPR that stage spilled to scratch; in the other stages, where nothing
    // spills, rebuilding it costs more registers than it saves)
    if (!bnd_early)
        for (int kb = MODE == 1 ? (wave << 6) + lane_here() : t; kb < NB; kb += WG) bnd_flux(kb);
    if (bad != INT_MAX) {
        if (MODE == 0 && irk == 1) {
            atomicMin(&S->err1_index, bad);
            S->err1_code = FVM_E_NEG_SOUND;
        } else {
            atomicMin(&S->err_index, bad);
            S->err_stage = irk;
            S->err_code = FVM_E_NEG_SOUND;
        }
    }
    __syncthreads();

    // 4. residual of P in the reference's order, then the RK stage
    __builtin_amdgcn_s_setprio(1);
    const int lane = MODE == 1 ? lane_here() : (t & 63);
    const int tl = MODE == 1 ? (wave << 6) + lane : t;   // local vertex this thread updates
    Red red{0.0, 0.0, 0};
    if (tl < NP) {
        const int v = v0 + tl;
        // loads of the update, issued before the ordered sum
        double4 o = make_double4(0.0, 0.0, 0.0, 0.0);
        double dtv = 0.0, cva = 1.0;
        if (MODE != 2) { o = d.prim[v]; dtv = d.dt[v]; cva = d.cvarea[v]; }
        // and the last stage's per-point inputs: issued here, they arrive
        // during the sum; after the state's store the compiler may not hoist
        // them (the arrays may alias it), and they cost a round trip each
        double drv = 0.0;
        int orv = 0, wk0 = 0, wk1 = 0;
        bool sol = false;
        if (MODE == 1) {
            if (!TURB) drv = d.drmin[v];
            orv = d.orig[v];
            sol = d.solid[v] != 0;
            wk0 = d.v2b_ptr[v];
            wk1 = d.v2b_ptr[v + 1];
        }
        V4 r;
        r.v[0] = 0.0;
        if (VISC) {
            const double *s_rv = reinterpret_cast<const double *>(reinterpret_cast<const char *>(s_gptr) + 16 * NBLOB);
            r.v[1] = s_rv[3 * tl]; r.v[2] = s_rv[3 * tl + 1]; r.v[3] = s_rv[3 * tl + 2];
        } else {
            r.v[1] = 0.0; r.v[2] = 0.0; r.v[3] = 0.0;
        }
        const int k1 = s_pptr[tl + 1];
        const double4 *fl = reinterpret_cast<const double4 *>(s_work);
        // a term enters with the edge's orientation: r - f is r + (-f) (IEEE
        // 754 defines the difference as the sum with the negated operand:
        // the same value, rounding and sign of zero), so the flux's sign bit
        // is flipped by the orientation bit and the term added, without a
        // branch that the lanes of a wave would take both ways
        auto apply = [&](const int q, const double4 f) {
            const unsigned long long m = (unsigned long long)(q & 1) << 63;
            auto sg = [m](double x) {
                return __builtin_bit_cast(double, __builtin_bit_cast(unsigned long long, x) ^ m);
            };
            r.v[0] = r.v[0] + sg(f.x); r.v[1] = r.v[1] + sg(f.y); r.v[2] = r.v[2] + sg(f.z); r.v[3] = r.v[3] + sg(f.w);
        };
        // two terms a step, read before either is added, added in order
        int k = s_pptr[tl];
        for (; k + 1 < k1; k += 2) {
            const int qa = s_psl[k], qb = s_psl[k + 1];
            const double4 fa = fl[qa >> 1], fb = fl[qb >> 1];
            apply(qa, fa);
            apply(qb, fb);
        }
        if (k < k1) {
            const int q = s_psl[k];
            apply(q, fl[q >> 1]);
        }
        // mu, mul, qx, qy of the last residual of an iteration are what the
        // monitors, the SA step and the caller read
        if (VISC && MODE != 0) { d.mul[v] = s_mul[tl]; d.mu[v] = s_mu[tl]; }
        if (MODE == 2) {
            st4(res_out, v, r);
        } else {
            const double4 pc = s_prim[tl];           // the stage state
            V4 po{{o.x, o.y, o.z, o.w}}, p{{pc.x, pc.y, pc.z, pc.w}}, u0, u1, un;
            prim2con(c, po, u0);
            prim2con(c, p, u1);
            const double a = c.airk[irk - 1], bb = c.birk[irk - 1];
            const double f = dtv / cva;
#pragma unroll
            for (int j = 0; j < 4; ++j) un.v[j] = a * u0.v[j] + bb * (u1.v[j] - f * r.v[j]);
            if (VISC && (MODE == 1 ? sol : d.solid[v] != 0)) { un.v[1] = 0.0; un.v[2] = 0.0; }
            double4 q;
            q.x = un.v[0];
            q.y = un.v[1] / un.v[0];
            q.z = un.v[2] / un.v[0];
            q.w = c.GAMMA1 * (un.v[3] - 0.5 * (sq(un.v[1]) + sq(un.v[2])) / un.v[0]);
            pout[v] = q;
            if (MODE == 1) {
                if (sol)
                    wall_terms<VISC>(d, wk0, wk1, q.w, VISC ? s_mu[tl] : 0.0, d.qx[v], d.qy[v]);
                const double gra = q.x - o.x;
                const double dif = fabs(gra);
                red.s = sq(gra);
                if (dif > 0.0) { red.m = dif; red.i = orv; }
                if (!TURB) {
                    if (q.x <= 0.0 || q.w <= 0.0) {
                        atomicMin(&S->errv_index, orv);
                        S->err_stage = irk;
                    }
                    // time_step of the next iteration (src/time_step.f:13-28)
                    double qq = sqrt(sq(q.y) + sq(q.z));
                    double aa = sqrt(c.GAMMA * q.w / q.x);
                    double dr = drv;
                    if (VISC) {
                        double dtv = 2.0 * c.GAMMA * s_mu[tl] / (q.x * c.prandtl);
                        d.dt[v] = c.cfl * sq(dr) / (dr * (qq + aa) + dtv);
                    } else {
                        d.dt[v] = c.cfl * dr / (qq + aa);
                    }
                }
            }
        }
    }
    if (MODE == 1) {
        Red rr = block_reduce<WG, true>(red, wave, lane);
        if (tl == 0) {
            d.partial[3 * b] = rr.s;
            d.partial[3 * b + 1] = rr.m;
            d.partial[3 * b + 2] = (double)rr.i;
        }
    }
}

// waves per SIMD an instance of the stage kernel is compiled for: NT = 512
// puts two waves per SIMD per workgroup, two workgroups a CU
template <bool VISC, bool TURB, int PREC, int NT, bool WIDE = false>
constexpr int stage_waves()
{
    return NT != kBlock ? 4 : (WIDE ? 4 : patch_waves<VISC, TURB, PREC>());
}

template <int FLUX, int LIM, bool VISC, int MODE, bool TURB, int PREC, int NT = kBlock>
__global__ void __launch_bounds__(NT, (stage_waves<VISC, TURB, PREC, NT>()))
k_patch(Dev d, Consts c, DevState *S, PatchArgs pa, int irk, const double4 *__restrict__ pin,
        double4 *__restrict__ pout, double4 *res_out)
{
    patch_body<FLUX, LIM, VISC, MODE, TURB, PREC, NT>(d, c, S, pa, logical_block(pa.npatch, d.xcd), irk,
                                                      pin, pout, res_out);
}

// Wall distance, src/wall_dist.f:28-61: for every point the smallest
// distance to the end points of the solid edges and, where the foot of the
// perpendicular falls inside an edge, to the edge.  The solid edges are
// staged through LDS in tiles; fmin over the same candidates gives the same
// value in any order.  COMPAT keeps the reference's dy3 = x2 - x1
// (src/wall_dist.f:46).
// distance of x to the n solid edges a[k]-b[k] staged in LDS, folded into w
template <bool COMPAT>
__device__ __forceinline__ double wall_candidates(const double2 x, const double2 *ea, const double2 *eb, int n,
                                                  double w)
{
    for (int k = 0; k < n; ++k) {
        const double2 a = ea[k], b = eb[k];
        const double dx1 = a.x - x.x, dy1 = a.y - x.y;
        w = fmin(w, sqrt(dx1 * dx1 + dy1 * dy1));
        const double dx2 = b.x - x.x, dy2 = b.y - x.y;
        w = fmin(w, sqrt(dx2 * dx2 + dy2 * dy2));
        const double dx3 = b.x - a.x;
        const double dy3 = COMPAT ? b.x - a.x : b.y - a.y;
        const double xi = -(dx1 * dx3 + dy1 * dy3) / (dx3 * dx3 + dy3 * dy3);
        if (xi >= 0.0 && xi <= 1.0) {
            const double xp = a.x + xi * dx3, yp = a.y + xi * dy3;
            w = fmin(w, sqrt((x.x - xp) * (x.x - xp) + (x.y - yp) * (x.y - yp)));
        }
    }
    return w;
}

template <bool COMPAT>
__global__ void __launch_bounds__(kBlock) k_wall_dist(int np, const double2 *__restrict__ xy,
                                                     const int2 *__restrict__ sedge, int nse,
                                                     const unsigned char *__restrict__ solid,
                                                     double *__restrict__ wd)
{
    __shared__ double2 ea[kBlock], eb[kBlock];
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const double2 x = i < np ? xy[i] : make_double2(0.0, 0.0);
    double w = 1.0e10;
    for (int base = 0; base < nse; base += kBlock) {
        const int j = base + threadIdx.x;
        if (j < nse) { int2 e = sedge[j]; ea[threadIdx.x] = xy[e.x]; eb[threadIdx.x] = xy[e.y]; }
        __syncthreads();
        w = wall_candidates<COMPAT>(x, ea, eb, min(kBlock, nse - base), w);
        __syncthreads();
    }
    if (i < np) wd[i] = solid[i] ? 0.0 : w;
}

// ---- wall distance without allocation (fvm_gpu_wall_dist_noalloc) ----
// The solid-edge list lives in a static device array: words 0-3 are a
// counter, an overflow flag and a bad-edge flag, then up to kScratchEdges
// int2 end-point pairs.  Every call on a device shares its copy.
constexpr int kScratchEdges = 1 << 16;
constexpr int kScratchInts = 4 + 2 * kScratchEdges;
static __device__ __align__(256) int scratch[kScratchInts];

__global__ void k_scratch_clear(int *sc)
{
    if (threadIdx.x < 4) sc[threadIdx.x] = 0;
}

// the edges with both ends solid (ptype 3), in any order (the distance is a
// minimum over them, which the order does not change)
__global__ void __launch_bounds__(kBlock) k_collect_solid_edges(int np, int netot, const int *__restrict__ edge,
                                                               const int *__restrict__ ptype, int *sc)
{
    const int e = blockIdx.x * kBlock + threadIdx.x;
    if (e >= netot) return;
    const int a = edge[2 * e] - 1, b = edge[2 * e + 1] - 1;
    if (a < 0 || a >= np || b < 0 || b >= np) { atomicOr(&sc[2], 1); return; }
    if (ptype[a] != 3 || ptype[b] != 3) return;
    const int k = atomicAdd(&sc[0], 1);
    if (k < kScratchEdges) reinterpret_cast<int2 *>(sc + 4)[k] = make_int2(a, b);
    else atomicOr(&sc[1], 1);
}

// k_wall_dist over the collected list; solid points get 1 directly (the
// value src/wall_dist.f:90-92 leaves there)
template <bool COMPAT>
__global__ void __launch_bounds__(kBlock) k_wall_dist_list(int np, const double2 *__restrict__ xy,
                                                          const int *__restrict__ ptype, const int *sc,
                                                          double *__restrict__ wd)
{
    __shared__ double2 ea[kBlock], eb[kBlock];
    const int nse = (sc[1] | sc[2]) ? 0 : min(sc[0], kScratchEdges);
    const int2 *sedge = reinterpret_cast<const int2 *>(sc + 4);
    const int i = blockIdx.x * kBlock + threadIdx.x;
    const double2 x = i < np ? xy[i] : make_double2(0.0, 0.0);
    double w = 1.0e10;
    for (int base = 0; base < nse; base += kBlock) {
        const int j = base + threadIdx.x;
        if (j < nse) { int2 e = sedge[j]; ea[threadIdx.x] = xy[e.x]; eb[threadIdx.x] = xy[e.y]; }
        __syncthreads();
        w = wall_candidates<COMPAT>(x, ea, eb, min(kBlock, nse - base), w);
        __syncthreads();
    }
    if (i < np) wd[i] = ptype[i] == 3 ? 1.0 : w;
}

// Start of an iterate call: history rows from zero, and no first-stage error
// pending (one is moved to err_code within its own iteration, so what is
// left here came from a stage that ran without its iteration: the stage
// timing, or a first stage beside the monitors that stopped the solve).
__global__ void k_begin_chunk(DevState *S)
{
    S->nhist = 0;
    S->err1_code = 0;
    S->err1_index = INT_MAX;
}

// plain copy for the achievable-bandwidth figure: 16 bytes per thread and
// step, grid-stride
__global__ void __launch_bounds__(kBlock) k_stream_copy(const double2 *__restrict__ src,
                                                        double2 *__restrict__ dst, long long n)
{
    const long long stride = (long long)gridDim.x * blockDim.x;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride)
        dst[i] = src[i];
}

// far-field states of all farfield edges for a given cl
__global__ void k_farfield_state(Dev d, Consts c, double cl)
{
    int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= d.nff) return;
    int2 e = d.bnode[d.nsw + b];
    double u[4];
    farfield_state(c, cl, d.coord[e.x], d.coord[e.y], u);
    d.uout[b] = make_double4(u[0], u[1], u[2], u[3]);
}

// with the point vortex: far-field states of the next iteration from this
// iteration's cl (COMMON cl, src/farfield_flux.f:28-44)
__device__ __forceinline__ void vortex_body(const Dev &d, const Consts &c, const DevState *S, int b)
{
    if (S->err_code) return;
    if (b >= d.nff) return;
    int2 e = d.bnode[d.nsw + b];
    double u[4];
    farfield_state(c, S->cl, d.coord[e.x], d.coord[e.y], u);
    d.uout[b] = make_double4(u[0], u[1], u[2], u[3]);
}

__global__ void k_vortex_state(Dev d, Consts c, const DevState *S)
{
    vortex_body(d, c, S, blockIdx.x * blockDim.x + threadIdx.x);
}

// ------------------------------------------------- batches of solves
// A batch runs many solves of one kind (flow type, flux, limiter, precision)
// with one launch per RK stage: the workgroups of all members' patches form
// one work list of (member, patch) pairs, and each workgroup reads its
// member's descriptor.  A single solve runs as a batch of one.  Members that
// have stopped (converged, iteration limit, error) leave at once.
struct BatchMember {
    Dev d;
    Consts c;
    DevState *S;
    PatchArgs pa;
    double4 *pbuf[3];      // stage buffers O, B, C
};

// WIDE: an SA instance compiled for 4 waves/SIMD (128 VGPRs, a few spilled),
// launched when four workgroups fit a CU's LDS (small patches of small
// meshes, as served solves have); the same operations, so the same bits
template <int FLUX, int LIM, bool VISC, int MODE, bool TURB, int PREC, bool WIDE = false, int NT = kBlock>
__global__ void __launch_bounds__(NT, (stage_waves<VISC, TURB, PREC, NT, WIDE>()))
k_patch_batch(const BatchMember *__restrict__ mem, const int2 *__restrict__ work, int nwork, int xcd,
              int irk)
{
    const int2 w = uniform_at(work, logical_block(nwork, xcd));
    const BatchMember &M = uniform_at(mem, w.x);
    if (MODE == 0 && w.y < 0) {
        // the monitors of the member's previous iteration: they read only
        // what that iteration's last stage left, and this stage writes only
        // its own buffer, so they run beside it (work lists hold such items
        // for workgroups of 256 or 512 threads)
        final_body<VISC, NT>(M.d, M.c, M.S);
        return;
    }
    patch_body<FLUX, LIM, VISC, MODE, TURB, PREC, NT>(M.d, M.c, M.S, M.pa, w.y, irk, M.pbuf[irk - 1],
                                                      M.pbuf[irk % 3], nullptr);
}

// vertex kernels over a work list of (member, block of kBlock vertices)
template <bool VISC, bool COPY>
__global__ void __launch_bounds__(kBlock) k_tstep_batch(const BatchMember *__restrict__ mem,
                                                        const int2 *__restrict__ vwork)
{
    const int2 w = uniform_at(vwork, (int)blockIdx.x);
    const BatchMember &M = uniform_at(mem, w.x);
    tstep_body<VISC, false, COPY>(M.d, M.c, w.y * kBlock + threadIdx.x);
}

__global__ void __launch_bounds__(kBlock) k_copy_batch(const BatchMember *__restrict__ mem,
                                                       const int2 *__restrict__ vwork)
{
    const int2 w = uniform_at(vwork, (int)blockIdx.x);
    const BatchMember &M = uniform_at(mem, w.x);
    if (M.S->done) return;
    const int v = w.y * kBlock + threadIdx.x;
    if (v < M.d.np) M.d.nut[v] = M.d.nut_new[v];
}

// one wide workgroup per member
// SA step over patches (src/sa_model.f): sa_body with its gathers served
// from LDS.  A workgroup loads, for its patch P and 1-ring (G), the state
// after the RK stages, nu_tilde, the laminar viscosity and the coordinates,
// and the local triangle, edge and boundary-edge tables; each vertex of P
// then makes sa_body's sums in sa_body's order (triangles ascending, edges
// in edge order, farfield edges in boundary order), so nu_tilde comes out
// bit for bit the same.
template <bool COMPAT>
__device__ __forceinline__ void sa_patch_body(const Dev &d, const Consts &c, DevState *S,
                                              const PatchArgs &pa, const int b)
{
    // as in the stage kernel, an error of this step does not stop the others
    const int es = S->err_stage;
    const int stop = S->done | (es != 0 && es != 4);
    const int hb = kPatchHdr * b;
    auto hdr = [&](int k) { return uniform_at(pa.hdr, hb + k); };
    const int v0 = hdr(PH_V0), NP = hdr(PH_NP), NG = hdr(PH_NG);
    const int NT = hdr(PH_NTRI), NE = hdr(PH_NEDGE), NB = hdr(PH_NBND);
    const int NGTL = hdr(PH_NGTL), NPSL = hdr(PH_NPSL);
    const int *halo = pa.halo + hdr(PH_HALO);
    const short4 *ltri = pa.ltri + hdr(PH_TRI);
    const double *lqinv = pa.lqinv + hdr(PH_TRI);
    const int *gptr = pa.gptr + hdr(PH_GPTR);
    const short *gtl = pa.gtl + hdr(PH_GTL);
    const short2 *ledge = pa.ledge + hdr(PH_EDGE);
    const double2 *ledge_ds = pa.ledge_ds + hdr(PH_EDGE);
    const short2 *lbnd = pa.lbnd + hdr(PH_BND);
    const int *lbnd_id = pa.lbnd_id + hdr(PH_BND);
    const int *pptr = pa.pptr + hdr(PH_PPTR);
    const short *pslot = pa.pslot + hdr(PH_PSL);
    const int t = threadIdx.x;
    const double nut_inf = c.nut_inf;
    extern __shared__ double4 smem4[];
    const int NGa = (NG + 1) & ~1, NTa = (NT + 1) & ~1;
    double4 *s_prim = smem4;
    double2 *s_coord = reinterpret_cast<double2 *>(s_prim + NG);
    double *s_nut = reinterpret_cast<double *>(s_coord + NG);
    double *s_mul = s_nut + NGa;
    double *s_lqinv = s_mul + NGa;
    double2 *s_eds = reinterpret_cast<double2 *>(s_lqinv + NTa);
    short4 *s_ltri = reinterpret_cast<short4 *>(s_eds + NE);
    short2 *s_ledge = reinterpret_cast<short2 *>(s_ltri + NT);
    short2 *s_lbnd = s_ledge + NE;
    int *s_lbid = reinterpret_cast<int *>(s_lbnd + NB);
    int *s_gptr = s_lbid + NB;
    int *s_pptr = s_gptr + NP + 1;
    short *s_gtl = reinterpret_cast<short *>(s_pptr + NP + 1);
    short *s_psl = s_gtl + NGTL;

    for (int i = t; i < NG; i += kBlock) {
        const int gv = i < NP ? v0 + i : halo[i - NP];
        s_prim[i] = d.prim[gv];
        s_coord[i] = d.coord[gv];
        s_nut[i] = d.nut[gv];
        s_mul[i] = d.mul[gv];
    }
    for (int i = t; i < NT; i += kBlock) { s_ltri[i] = ltri[i]; s_lqinv[i] = lqinv[i]; }
    for (int i = t; i < NE; i += kBlock) { s_ledge[i] = ledge[i]; s_eds[i] = ledge_ds[i]; }
    for (int i = t; i < NB; i += kBlock) { s_lbnd[i] = lbnd[i]; s_lbid[i] = lbnd_id[i]; }
    for (int i = t; i <= NP; i += kBlock) { s_gptr[i] = gptr[i]; s_pptr[i] = pptr[i]; }
    for (int i = t; i < NGTL; i += kBlock) s_gtl[i] = gtl[i];
    for (int i = t; i < NPSL; i += kBlock) s_psl[i] = pslot[i];
    __syncthreads();
    if (stop || t >= NP) return;

    const int v = v0 + t;
    // loads of the update, issued before the sums
    const double wdv = d.wd[v], dti = d.dt[v], cva = d.cvarea[v];
    const double Omega = d.qy[v].y - d.qx[v].z;
    const bool solid = d.solid[v] != 0;
    const double4 p = s_prim[t];
    const double nu = s_nut[t], mulv = s_mul[t];
    double g1c = 0.0, g2c = 0.0, g3c = 0.0;
    if (COMPAT) { g1c = d.nut[d.sa1]; g2c = d.nut[d.sa2]; g3c = d.nut[d.sa3]; }
    double divf = 0.0;
    for (int k = s_gptr[t]; k < s_gptr[t + 1]; ++k) {
        const int e = s_gtl[k];
        const short4 n = s_ltri[e >> 2];
        const double2 x1 = s_coord[n.x], x2 = s_coord[n.y], x3 = s_coord[n.z];
        const double areat = s_lqinv[e >> 2];     // 0.25 / elarea
        const double bx[3] = {x2.y - x3.y, x3.y - x1.y, x1.y - x2.y};
        const double by[3] = {x3.x - x2.x, x1.x - x3.x, x2.x - x1.x};
        const double n1 = s_nut[n.x], n2 = s_nut[n.y], n3 = s_nut[n.z];
        const double g1 = COMPAT ? g1c : n1, g2 = COMPAT ? g2c : n2, g3 = COMPAT ? g3c : n3;
        const double nux = g1 * bx[0] + g2 * bx[1] + g3 * bx[2];
        const double nuy = g1 * by[0] + g2 * by[1] + g3 * by[2];
        const double mult = (s_mul[n.x] + s_mul[n.y] + s_mul[n.z]) / 3.0;
        const double prim5t = (n1 + n2 + n3) / 3.0;
        const double rhot = (s_prim[n.x].x + s_prim[n.y].x + s_prim[n.z].x) / 3.0;
        const double mutb = mult / rhot + prim5t;
        const int m = e & 3;
        divf = divf + (c.Cb2Sig1 * mutb - c.Cb2Sig2 * (mulv / p.x + nu)) * (nux * bx[m] + nuy * by[m]) * areat;
    }
    for (int k = s_pptr[t]; k < s_pptr[t + 1]; ++k) {
        const int q = s_psl[k], term = q >> 1;
        if (term < NE) {
            const short2 nn = s_ledge[term];
            const double2 ds = s_eds[term];
            const double4 pa_ = s_prim[nn.x], pb = s_prim[nn.y];
            const double unl = pa_.y * ds.x + pa_.z * ds.y;
            const double unr = pb.y * ds.x + pb.z * ds.y;
            const double una = 0.5 * (unl + unr);
            const double flux = (una > 0.0) ? una * s_nut[nn.x] : una * s_nut[nn.y];
            divf = (q & 1) ? divf - flux : divf + flux;
        } else {
            const int kb = (term - NE) >> 1, side = (term - NE) & 1;
            if (s_lbid[kb] < d.nsw) continue;
            const short2 nn = s_lbnd[kb];
            const double2 x1 = s_coord[nn.x], x2 = s_coord[nn.y];
            const double sx = x2.y - x1.y, sy = -(x2.x - x1.x);
            const double u = 0.5 * (s_prim[nn.x].y + s_prim[nn.y].y), w = 0.5 * (s_prim[nn.x].z + s_prim[nn.y].z);
            const double un = u * sx + w * sy;
            // (the choice between the two values, not their addresses: an
            // LDS and a global address make a flat load)
            const double nin = s_nut[side ? nn.y : nn.x], nout = nut_inf;
            const double fl = un * ((un > 0.0) ? nin : nout);
            divf = divf + 0.5 * fl;
        }
    }
    const double chi = nu * p.x / mulv;
    const double chi3 = pow3(chi);
    const double fv1 = chi3 / (chi3 + c.Cv11);
    const double fv2 = 1.0 / pow3(1.0 + chi / c.Cv2);
    double r = nu / wdv / (wdv * c.kolm2);
    const double fv3 = (1.0 + chi * fv1) * (1.0 - fv2) / fmax(chi, 0.001);
    const double Sv = fv3 * fabs(Omega) + r * fv2;
    r = r / Sv;
    r = fmin(r, 2.0);
    const double g = r + c.Cw2 * (pow6(r) - r);
    const double fw = g * pow(c.Cw31 / (pow6(g) + c.Cw32), 1.0 / 6.0);
    const double source = c.Cb1 * Sv * nu;
    const double fact = 1.0 + c.Cw1 * fw * (nu / wdv) * (dti / wdv);
    double nn = (nu - (dti / cva) * divf + dti * source) / fact;
    if (solid) nn = 0.0;
    d.nut_new[v] = nn;
    if (p.x <= 0.0 || p.w <= 0.0 || nn < 0.0) {
        atomicMin(&S->err_index, d.orig[v]);
        S->err_stage = 4;
        S->err_code = FVM_E_NONPOSITIVE;
    }
}

template <bool COMPAT>
__global__ void __launch_bounds__(kBlock) k_sa_patch_batch(const BatchMember *__restrict__ mem,
                                                           const int2 *__restrict__ work, int nwork, int xcd)
{
    const int2 w = uniform_at(work, logical_block(nwork, xcd));
    const BatchMember &M = uniform_at(mem, w.x);
    sa_patch_body<COMPAT>(M.d, M.c, M.S, M.pa, w.y);
}

template <bool VISC>
__global__ void __launch_bounds__(kFinal) k_final_batch(const BatchMember *__restrict__ mem)
{
    const BatchMember &M = uniform_at(mem, (int)blockIdx.x);
    final_body<VISC>(M.d, M.c, M.S);
}

__global__ void __launch_bounds__(kBlock) k_vortex_batch(const BatchMember *__restrict__ mem)
{
    const BatchMember &M = uniform_at(mem, (int)blockIdx.y);
    if (M.c.vortex == 1 && M.c.mach_inf < 1.0)
        vortex_body(M.d, M.c, M.S, blockIdx.x * blockDim.x + threadIdx.x);
}

__global__ void k_begin_batch(const BatchMember *__restrict__ mem, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {                        // as k_begin_chunk, per member
        mem[i].S->nhist = 0;
        mem[i].S->err1_code = 0;
        mem[i].S->err1_index = INT_MAX;
    }
}

__global__ void k_gather_batch(const BatchMember *__restrict__ mem, int n, DevState *out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = *mem[i].S;
}

}  // namespace

// ----------------------------------------------------------------- host side

// Handles are registered when fvm_gpu_create / fvm_batch_create return them
// and unregistered by destroy; every entry point looks a handle up before it
// reads it, so a destroyed or foreign handle is refused (FVM_E_INPUT) without
// touching freed memory.  The tag tells a context from a batch.
constexpr unsigned kCtxTag = 0x46564d43u, kBatchTag = 0x46564d42u;
namespace {
std::mutex g_live_mu;
std::unordered_set<const void *> g_live;
void live_add(const void *h)
{
    std::lock_guard<std::mutex> l(g_live_mu);
    g_live.insert(h);
}
void live_drop(const void *h)
{
    std::lock_guard<std::mutex> l(g_live_mu);
    g_live.erase(h);
}
bool live(const void *h)
{
    std::lock_guard<std::mutex> l(g_live_mu);
    return g_live.count(h) != 0;
}
}  // namespace

struct fvm_ctx {
    unsigned tag = kCtxTag;
    int cus = 256, per_cu = 4, lds_per_cu = 160 * 1024;   // device and occupancy
    int threads = 256;                                      // stage kernel workgroup size
    int target = 0;                                         // patch size asked of the partition
    int max_patch = 0;
    long long halo_total = 0;
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
    double4 *prim_bak = nullptr;
    double *nut_bak = nullptr;
    int variant = 3;                  // 1: edge-flux buffer + gathers; 3: patch stage kernel
    double4 *pbuf[3] = {nullptr, nullptr, nullptr};   // stage buffers O, B, C
    PatchArgs pa{};
    std::vector<int> perm;            // device vertex -> caller's vertex (0-based)
    BatchMember *self_mem = nullptr;  // this solve as a batch of one
    int2 *self_work = nullptr;
    int2 *self_work_mon = nullptr;    // first-stage list with the monitor item
    // inviscid stage kernels leave qx, qy unwritten; true once iterations ran
    // since they were last computed (fvm_gpu_get_state then recomputes them)
    bool qxy_stale = false;
    std::vector<int> patch_cost;      // per patch: edges + local vertices (work estimate)
    struct fvm_batch *batch = nullptr;   // the batch it belongs to, if any
};

// Many solves of one kind sharing each launch (see BatchMember).  Membership
// changes rebuild the work lists and the captured iterations.
struct fvm_batch {
    unsigned tag = kBatchTag;
    hipStream_t stream = nullptr;
    std::vector<fvm_ctx *> mem;
    BatchMember *d_mem = nullptr;
    int2 *d_work = nullptr, *d_vwork = nullptr;
    int2 *d_work_mon = nullptr;       // first-stage list with one monitor item per member
    DevState *d_states = nullptr;
    int nwork = 0, nvwork = 0, lds = 0, sa_lds = 0, max_nff = 0;
    bool dirty = true, any_vortex = false;
    hipGraphExec_t graph_n = nullptr, graph_1 = nullptr;
    bool graphs_ok = true;
    bool pending = false;             // a launched chunk not waited for yet
    // members' monitors after a launched chunk: pinned host memory, so the
    // copy is ordered on the batch's stream and fvm_batch_launch returns at
    // once (a copy into pageable memory may complete only with the chunk)
    DevState *hs = nullptr;
    int hs_cap = 0;
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

// As HIPCHK, where copies between host buffers of the caller's frame and the
// device may still be queued on `st`: the stream drains before the return,
// so no copy outlives the buffers it reads or writes.
#define HIPCHK_S(st, x)                                                             \
    do {                                                                            \
        hipError_t e_ = (x);                                                        \
        if (e_ != hipSuccess) {                                                     \
            std::fprintf(stderr, "fvm2d_gpu: %s failed: %s (%s:%d)\n", #x,          \
                         hipGetErrorString(e_), __FILE__, __LINE__);                \
            (void)hipStreamSynchronize(st);                                         \
            return FVM_E_HIP;                                                       \
        }                                                                           \
    } while (0)

inline bool ctx_ok(const fvm_ctx *x) { return x && live(x) && x->tag == kCtxTag; }
}  // namespace
int settle(fvm_ctx *x);
namespace {
inline bool batch_ok(const fvm_batch *b) { return b && live(b) && b->tag == kBatchTag; }

// monitor state with no error pending
inline DevState no_errors()
{
    DevState s{};
    s.err_index = INT_MAX;
    s.err1_index = INT_MAX;
    s.errv_index = INT_MAX;
    return s;
}

// Device memory of contexts and batches is stream-ordered (hipMallocAsync /
// hipFreeAsync on their own streams, from the device's default pool, which
// keeps freed memory for reuse): hipFree waits for the work of every stream
// on the device, so a solve created or destroyed while other solves' chunks
// run would stall the host until they end.
void keep_pool_memory()
{
    static std::mutex mu;
    static bool done[64] = {false};
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= 64) { (void)hipGetLastError(); return; }
    std::lock_guard<std::mutex> lock(mu);
    if (done[dev]) return;
    done[dev] = true;
    hipMemPool_t pool = nullptr;
    if (hipDeviceGetDefaultMemPool(&pool, dev) == hipSuccess) {
        uint64_t keep = UINT64_MAX;
        (void)hipMemPoolSetAttribute(pool, hipMemPoolAttrReleaseThreshold, &keep);
    }
    (void)hipGetLastError();
}

template <class T>
int dalloc(fvm_ctx *x, T **p, size_t n)
{
    size_t b = std::max<size_t>(n, 1) * sizeof(T);
    if (hipMallocAsync((void **)p, b, x->stream) != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return FVM_E_NOMEM;
    }
    x->allocs.push_back((void *)*p);
    x->bytes += (long long)b;
    return FVM_OK;
}

// copy on the context's stream, drained before the host vector can go
template <class T>
int upload(fvm_ctx *x, T **p, const std::vector<T> &h)
{
    int rc = dalloc(x, p, h.size());
    if (rc) return rc;
    if (!h.empty()) {
        HIPCHK_S(x->stream, hipMemcpyAsync(*p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, x->stream));
        HIPCHK(hipStreamSynchronize(x->stream));
    }
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
    c.sin_aoa = std::sin(c.aoa);
    c.cos_aoa = std::cos(c.aoa);
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

// host copy of the free-stream Sutherland viscosity (src/viscosity.f:42-55)
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


// Work estimate of one patch in a stage launch, for the balance of the
// workgroups over the CUs: its edge fluxes plus its boundary edges.  A
// patch with boundary edges evaluates them on otherwise idle threads beside
// the gradients, but the chain of dependent loads and divisions of a
// farfield (Steger-Warming) or wall edge lengthens the workgroup's critical
// path far beyond its share of the instructions.  Weights in units of
// 0.1 ns, fitted to per-CU end times measured on MI355X (s_memtime at each
// phase boundary of every workgroup, 513x129 Euler/Roe, one round of 1017
// workgroups): FP64 7.7 ns per edge, 116 per farfield and 70 per wall edge
// (R^2 0.82 over 768 CU samples, against 0.16 for edges + local vertices);
// split precision 4.6, 170 and 55 (R^2 0.87).
int patch_work(int nedge, int nfarfield, int nsolid, int precision)
{
    if (precision == 1) return 46 * nedge + 1700 * nfarfield + 550 * nsolid;
    return 77 * nedge + 1160 * nfarfield + 700 * nsolid;
}

// Order of a stage launch's work list (a performance choice only: every
// result is the same for every order).  The XCD-aware mapping hands the
// list to the 8 XCDs as contiguous ranges.  When a range fits one round of
// resident workgroups, the k-th workgroup an XCD receives runs on CU
// k mod C of it (C CUs per XCD; the dispatcher deals a round to the CUs in
// turn; measured), and a CU's workgroups, issued oldest first, end when the
// CU has done their summed work: the items are dealt heaviest first to the
// least loaded CU, then swaps between the most loaded CU and the others
// lower the largest load while they can; each CU's items sit at its
// positions j, j + C, ... heaviest first.  With more rounds, the range is
// sorted heaviest first, so that the last round holds the light ones.
void order_work(std::vector<int2> &work, const std::vector<int> &cost, int cus, int per_cu)
{
    const int n = (int)work.size();
    const int C = std::max(1, cus / 8);
    const int q = n >> 3, r = n & 7;
    std::vector<int2> out(work.size());
    for (int x = 0; x < 8; ++x) {
        const int lo = x * q + std::min(x, r), m = q + (x < r ? 1 : 0);
        if (m == 0) continue;
        std::vector<int> idx(m);
        std::iota(idx.begin(), idx.end(), lo);
        std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return cost[a] > cost[b]; });
        if (m > C * per_cu) {
            for (int i = 0; i < m; ++i) out[lo + i] = work[idx[i]];
            continue;
        }
        // CU j owns local positions j, j + C, ... below m
        std::vector<long long> load(C, 0);
        std::vector<int> cap(C);
        std::vector<std::vector<int>> items(C);       // indices into work
        for (int j = 0; j < C; ++j) cap[j] = j < m ? (m - 1 - j) / C + 1 : 0;
        for (int i = 0; i < m; ++i) {
            int best = -1;
            for (int j = 0; j < C; ++j)
                if ((int)items[j].size() < cap[j] && (best < 0 || load[j] < load[best])) best = j;
            items[best].push_back(idx[i]);
            load[best] += cost[idx[i]];
        }
        // swaps that lower the most loaded CU's load without raising the
        // partner's above it
        for (int it = 0; it < 4 * m; ++it) {
            int hi = 0;
            for (int j = 1; j < C; ++j) if (load[j] > load[hi]) hi = j;
            long long best_gain = 0;
            int bj = -1, ba = -1, bb = -1;
            for (int j = 0; j < C; ++j) {
                if (j == hi) continue;
                for (int a = 0; a < (int)items[hi].size(); ++a)
                    for (int q = 0; q < (int)items[j].size(); ++q) {
                        const long long dc = (long long)cost[items[hi][a]] - cost[items[j][q]];
                        if (dc <= 0 || load[j] + dc >= load[hi]) continue;
                        const long long gain = std::min(dc, load[hi] - (load[j] + dc));
                        if (gain > best_gain) { best_gain = gain; bj = j; ba = a; bb = q; }
                    }
            }
            if (bj < 0) break;
            const long long dc = (long long)cost[items[hi][ba]] - cost[items[bj][bb]];
            std::swap(items[hi][ba], items[bj][bb]);
            load[hi] -= dc;
            load[bj] += dc;
        }
        for (int j = 0; j < C; ++j) {
            std::stable_sort(items[j].begin(), items[j].end(), [&](int a, int b) { return cost[a] > cost[b]; });
            for (int f = 0; f < (int)items[j].size(); ++f) out[lo + j + C * f] = work[items[j][f]];
        }
    }
    work.swap(out);
}

// The first-stage work list of a captured iteration that follows another:
// the patches plus one item (member, -1) per member, which runs the monitors
// of the member's previous iteration beside the stage; costed at a quarter of
// a mean patch and ordered like the patches.
std::vector<int2> with_monitors(std::vector<int2> work, std::vector<int> cost, int nm, int cus,
                                int per_cu)
{
    long long sum = 0;
    for (int c : cost) sum += c;
    const int mc = std::max(1, (int)(sum / (long long)std::max<size_t>(cost.size(), 1) / 4));
    for (int i = 0; i < nm; ++i) {
        work.push_back(make_int2(i, -1));
        cost.push_back(mc);
    }
    order_work(work, cost, cus, per_cu);
    return work;
}

// LDS a stage-kernel workgroup takes on a CU: its dynamic LDS plus the
// kernel's 512 bytes of static LDS, in 1 KiB units
inline int lds_alloc(int lds_bytes) { return (lds_bytes + 512 + 1023) & ~1023; }

// workgroups of the stage kernel a CU holds: per_cu waves per SIMD allow
// per_cu workgroups of 256 threads (half as many of 512), fewer where the
// LDS runs out
int effective_per_cu(const fvm_ctx *x, int lds_bytes)
{
    return std::max(1, std::min(x->per_cu * kBlock / x->threads, x->lds_per_cu / lds_alloc(lds_bytes)));
}

// SA solves (Roe, limited) whose patches leave room for four workgroups a CU
// run the 4-wave instance of the stage kernel
bool wide_sa(const fvm_ctx *x, int lds_bytes)
{
    const fvm_params &q = x->par;
    return q.iflow == 3 && q.iflux == 1 && q.ilimit == 1 && lds_alloc(lds_bytes) * 4 <= x->lds_per_cu;
}

// workgroups a CU holds at once for the stage kernel of this solve
int stage_per_cu(const fvm_ctx *x, int lds_bytes)
{
    return wide_sa(x, lds_bytes) ? 4 : effective_per_cu(x, lds_bytes);
}

// Stage kernels read the stage state `pin` and write `pout`; the buffers
// rotate O -> B -> C -> O over the three stages, O (d.prim) being the state
// at the start of the iteration, which the RK combination also reads.

template <int V> using IC = std::integral_constant<int, V>;

// Calls f(FLUX, LIM, VISC, TURB, PREC) with compile-time constants for one
// kind of solve: flux 1 Roe, 2 KFVS, 3 HLLC, 4 HLLC rotated; limiter 0/1;
// flow 1 Euler, 2 laminar, 3 SA; precision 0 FP64, 1 split.
template <class F>
void with_kind(int iflow, int iflux, int hllc_rotate, int ilimit, int precision, F &&f)
{
    auto prec = [&](auto fl, auto lm, auto vi, auto tu) {
        if (precision == 1) f(fl, lm, vi, tu, IC<1>{});
        else f(fl, lm, vi, tu, IC<0>{});
    };
    auto flow = [&](auto fl, auto lm) {
        if (iflow == 1) prec(fl, lm, IC<0>{}, IC<0>{});
        else if (iflow == 2) prec(fl, lm, IC<1>{}, IC<0>{});
        else prec(fl, lm, IC<1>{}, IC<1>{});
    };
    auto lim = [&](auto fl) {
        if (ilimit == 1) flow(fl, IC<1>{});
        else flow(fl, IC<0>{});
    };
    switch (iflux == 3 && hllc_rotate ? 4 : iflux) {
    case 1: lim(IC<1>{}); break;
    case 2: lim(IC<2>{}); break;
    case 3: lim(IC<3>{}); break;
    default: lim(IC<4>{}); break;
    }
}

// what one launch of the batched stage kernel needs
struct StageLaunch {
    const BatchMember *mem;
    const int2 *work;
    int nwork, xcd, lds;
    hipStream_t stream;
    int iflow, iflux, hllc_rotate, ilimit, precision;
    bool wide;                  // SA, Roe, limited: the 4-wave instance
    int threads;                // workgroup size: 256, or 512 (Euler, limited)
};

// kinds with a 512-thread instance of the stage kernel: Euler with the
// limiter, every flux (the viscous and SA instances and the unlimited ones
// keep 256 threads)
bool has_wide_block(const fvm_params &q)
{
    return q.iflow == 1 && q.ilimit == 1;
}

void launch_stage(const StageLaunch &b, int irk, int mode)
{
    with_kind(b.iflow, b.iflux, b.hllc_rotate, b.ilimit, b.precision,
              [&](auto fl, auto lm, auto vi, auto tu, auto pr) {
        constexpr int F = decltype(fl)::value, L = decltype(lm)::value, P = decltype(pr)::value;
        constexpr bool V = decltype(vi)::value != 0, T = decltype(tu)::value != 0;
        if constexpr (L == 1 && !V && !T) {
            if (b.threads == 512) {
                if (mode == 0)
                    hipLaunchKernelGGL((k_patch_batch<F, L, V, 0, T, P, false, 512>), dim3(b.nwork), dim3(512),
                                       b.lds, b.stream, b.mem, b.work, b.nwork, b.xcd, irk);
                else
                    hipLaunchKernelGGL((k_patch_batch<F, L, V, 1, T, P, false, 512>), dim3(b.nwork), dim3(512),
                                       b.lds, b.stream, b.mem, b.work, b.nwork, b.xcd, irk);
                return;
            }
        }
        if constexpr (T && F == 1 && L == 1) {
            if (b.wide) {
                if (mode == 0)
                    hipLaunchKernelGGL((k_patch_batch<F, L, V, 0, T, P, true>), dim3(b.nwork), dim3(kBlock),
                                       b.lds, b.stream, b.mem, b.work, b.nwork, b.xcd, irk);
                else
                    hipLaunchKernelGGL((k_patch_batch<F, L, V, 1, T, P, true>), dim3(b.nwork), dim3(kBlock),
                                       b.lds, b.stream, b.mem, b.work, b.nwork, b.xcd, irk);
                return;
            }
        }
        if (mode == 0)
            hipLaunchKernelGGL((k_patch_batch<F, L, V, 0, T, P>), dim3(b.nwork), dim3(kBlock), b.lds,
                               b.stream, b.mem, b.work, b.nwork, b.xcd, irk);
        else
            hipLaunchKernelGGL((k_patch_batch<F, L, V, 1, T, P>), dim3(b.nwork), dim3(kBlock), b.lds,
                               b.stream, b.mem, b.work, b.nwork, b.xcd, irk);
    });
}

// residual of one state through the stage kernel (MODE 2, no update)
void launch_residual(fvm_ctx *x, DevState *S, const double4 *pin, double4 *res)
{
    const fvm_params &q = x->par;
    with_kind(q.iflow, q.iflux, q.hllc_rotate, q.ilimit, q.precision,
              [&](auto fl, auto lm, auto vi, auto tu, auto pr) {
        constexpr int F = decltype(fl)::value, L = decltype(lm)::value, P = decltype(pr)::value;
        constexpr bool V = decltype(vi)::value != 0, T = decltype(tu)::value != 0;
        if constexpr (L == 1 && !V && !T) {
            if (x->threads == 512) {
                hipLaunchKernelGGL((k_patch<F, L, V, 2, T, P, 512>), dim3(x->pa.npatch), dim3(512),
                                   x->pa.lds_bytes, x->stream, x->d, x->c, S, x->pa, 1, pin, x->pbuf[1], res);
                return;
            }
        }
        hipLaunchKernelGGL((k_patch<F, L, V, 2, T, P>), dim3(x->pa.npatch), dim3(kBlock),
                           x->pa.lds_bytes, x->stream, x->d, x->c, S, x->pa, 1, pin, x->pbuf[1], res);
    });
}

StageLaunch self_stage(const fvm_ctx *x, bool monitors = false)
{
    const fvm_params &q = x->par;
    return StageLaunch{x->self_mem, monitors ? x->self_work_mon : x->self_work,
                       x->pa.npatch + (monitors ? 1 : 0), x->d.xcd, x->pa.lds_bytes, x->stream,
                       q.iflow, q.iflux, q.hllc_rotate, q.ilimit, q.precision,
                       wide_sa(x, x->pa.lds_bytes), x->threads};
}

// Captured iterations after the first may run the monitors inside the next
// first stage, except with the point-vortex farfield (the next stage reads
// the monitors' cl).  With SA the next iteration's time step comes first:
// it writes only dt, which a stopped solve no longer uses.
bool monitors_merge(const fvm_params &q, bool vortex)
{
    (void)q;
    return !vortex;
}

BatchMember member_of(const fvm_ctx *x)
{
    BatchMember m;
    m.d = x->d;
    m.c = x->c;
    m.S = x->S;
    m.pa = x->pa;
    for (int i = 0; i < 3; ++i) m.pbuf[i] = x->pbuf[i];
    return m;
}

// the context's own descriptor, after anything in it changed
int sync_self(fvm_ctx *x)
{
    BatchMember m = member_of(x);
    HIPCHK_S(x->stream, hipMemcpyAsync(x->self_mem, &m, sizeof m, hipMemcpyHostToDevice, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    return FVM_OK;
}

// time step (and save_old for the edge-flux layout); copy: also the
// previous iteration's nu_tilde update (SA, stage kernels only)
void launch_tstep(fvm_ctx *x, bool copy = false)
{
    const int np = x->np;
    const dim3 g(nblk(np)), blk(kBlock);
    if (x->variant != 3) {
        if (x->par.iflow != 1) hipLaunchKernelGGL((k_tstep<true>), g, blk, 0, x->stream, x->d, x->c);
        else hipLaunchKernelGGL((k_tstep<false>), g, blk, 0, x->stream, x->d, x->c);
    } else if (copy) {
        hipLaunchKernelGGL((k_tstep<true, false, true>), g, blk, 0, x->stream, x->d, x->c);
    } else if (x->par.iflow != 1) {
        hipLaunchKernelGGL((k_tstep<true, false>), g, blk, 0, x->stream, x->d, x->c);
    } else {
        hipLaunchKernelGGL((k_tstep<false, false>), g, blk, 0, x->stream, x->d, x->c);
    }
}

// Kernel sequence of one iteration (src/main.f:44-63).  If `timer` is set,
// each launch is bracketed by events and its duration added to the tally.
struct Tally { std::vector<double> ms; std::vector<int> cnt; std::vector<const char *> names; };

// One iteration.  mon_prev: the first stage also runs the monitors of the
// previous iteration (which then had none of its own); final: the iteration
// ends with its own monitors.
void one_iteration(fvm_ctx *x, Tally *tally, bool mon_prev = false, bool final = true)
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
    if (x->variant == 3) {
        // time_step runs at the end of the previous iteration's last stage,
        // except when the SA step (which reads this iteration's dt) follows
        // with SA, iterations after the first of a captured graph also make
        // the previous iteration's nu_tilde update here (mon_prev marks them)
        if (turb) timed("k_tstep", [&] { launch_tstep(x, mon_prev); });
        for (int irk = 1; irk <= 3; ++irk) {
            const double4 *pin = x->pbuf[irk - 1];
            double4 *pout = x->pbuf[irk % 3];
            const int mode = irk == 3 ? 1 : 0;
            (void)pin; (void)pout;
            timed("k_patch", [&] { launch_stage(self_stage(x, mon_prev && irk == 1), irk, mode); });
        }
    } else {
    timed("k_tstep", [&] { launch_tstep(x); });
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
    }
    if (turb) {
        timed("k_sa", [&] {
            if (x->variant == 3) {
                const int n = x->pa.npatch, lds = x->pa.sa_lds_bytes;
                if (x->par.sa_compat)
                    hipLaunchKernelGGL(k_sa_patch_batch<true>, dim3(n), dim3(kBlock), lds, x->stream,
                                       x->self_mem, x->self_work, n, x->d.xcd);
                else
                    hipLaunchKernelGGL(k_sa_patch_batch<false>, dim3(n), dim3(kBlock), lds, x->stream,
                                       x->self_mem, x->self_work, n, x->d.xcd);
            } else if (x->par.sa_compat) {
                hipLaunchKernelGGL(k_sa<true>, dim3(nblk(np)), dim3(kBlock), 0, x->stream, x->d, x->c, x->S);
            } else {
                hipLaunchKernelGGL(k_sa<false>, dim3(nblk(np)), dim3(kBlock), 0, x->stream, x->d, x->c, x->S);
            }
            if (final || x->variant != 3)
                hipLaunchKernelGGL(k_copy, dim3(nblk(np)), dim3(kBlock), 0, x->stream, x->d.nut_new, x->d.nut,
                                   np, x->S);
        });
    }
    if (final) timed("k_final", [&] {
        if (x->par.iflow == 1)
            hipLaunchKernelGGL(k_final<false>, dim3(1), dim3(kFinal), 0, x->stream, x->d, x->c, x->S);
        else
            hipLaunchKernelGGL(k_final<true>, dim3(1), dim3(kFinal), 0, x->stream, x->d, x->c, x->S);
        if (x->c.vortex == 1 && x->c.mach_inf < 1.0 && x->nff > 0)
            hipLaunchKernelGGL(k_vortex_state, dim3(nblk(x->nff)), dim3(kBlock), 0, x->stream, x->d,
                               x->c, x->S);
    });
    if (tally) { (void)hipEventDestroy(e0); (void)hipEventDestroy(e1); }
}

// Applies the tuning knobs: kernel variant, vertices per workgroup, XCD
// remap.  Re-establishes the iteration-start invariant (prim_old = prim,
// dt = time_step(prim)) that the variant-2 last stage maintains.
int configure(fvm_ctx *x, int variant, int vb, int xcd)
{
    if (variant != 1 && variant != 3) return FVM_E_INPUT;
    if (variant == 1 && x->par.precision != 0) return FVM_E_INPUT;   // split precision: stage kernel only
    if (vb < 16 || vb > kBlock) return FVM_E_INPUT;
    x->variant = variant;
    x->d.vb = vb;
    x->d.nblocks = (x->np + vb - 1) / vb;
    x->d.xcd = xcd ? 1 : 0;
    x->d.nparts = variant == 3 ? x->pa.npatch : (int)nblk(x->np);
    if (x->self_mem) {
        int rc = sync_self(x);
        if (rc) return rc;
    }
    if (x->batch) x->batch->dirty = true;
    HIPCHK(hipStreamSynchronize(x->stream));      // no replay of the old graphs in flight
    if (x->graph_n) (void)hipGraphExecDestroy(x->graph_n);
    if (x->graph_1) (void)hipGraphExecDestroy(x->graph_1);
    x->graph_n = x->graph_1 = nullptr;
    x->graphs_ok = true;
    launch_tstep(x);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(x->stream));
    return FVM_OK;
}

// Ends a capture on `st` and instantiates it into *out.  On any failure the
// capture is ended all the same (the stream is left out of capture mode),
// what was built is destroyed, *out stays null and the sticky error is
// cleared, so the caller can fall back to launching the kernels directly.
int end_capture(hipStream_t st, hipGraphExec_t *out)
{
    *out = nullptr;
    hipGraph_t g = nullptr;
    hipError_t e = hipStreamEndCapture(st, &g);
    hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
    if (hipStreamIsCapturing(st, &cs) == hipSuccess && cs != hipStreamCaptureStatusNone) {
        hipGraph_t g2 = nullptr;
        (void)hipStreamEndCapture(st, &g2);
        if (g2) (void)hipGraphDestroy(g2);
    }
    hipGraphExec_t ex = nullptr;
    if (e == hipSuccess && g) e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    else if (e == hipSuccess) e = hipErrorInvalidValue;
    if (g) (void)hipGraphDestroy(g);
    if (e != hipSuccess) {
        if (ex) (void)hipGraphExecDestroy(ex);
        (void)hipGetLastError();
        return FVM_E_HIP;
    }
    *out = ex;
    return FVM_OK;
}

int build_graph(fvm_ctx *x, int iters, hipGraphExec_t *out)
{
    *out = nullptr;
    if (hipStreamBeginCapture(x->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        (void)hipGetLastError();
        return FVM_E_HIP;
    }
    const bool merge = x->variant == 3 && monitors_merge(x->par, x->c.vortex == 1 && x->c.mach_inf < 1.0 && x->nff > 0);
    for (int i = 0; i < iters; ++i) one_iteration(x, nullptr, merge && i > 0, !merge || i == iters - 1);
    return end_capture(x->stream, out);
}

int reset_state(fvm_ctx *x)
{
    DevState s = no_errors();
    s.iter = x->par.iterlast;
    s.fres = x->par.minres + 1.0;
    if (!(x->par.maxiter > 0)) { s.done = 1; s.status = FVM_ITER_LIMIT; }
    x->hS = s;
    const DevState h = no_errors();
    HIPCHK_S(x->stream, hipMemcpyAsync(x->S, &s, sizeof s, hipMemcpyHostToDevice, x->stream));
    HIPCHK_S(x->stream, hipMemcpyAsync(x->S_hook, &h, sizeof h, hipMemcpyHostToDevice, x->stream));
    if (x->nff > 0)       // COMMON cl starts at zero
        hipLaunchKernelGGL(k_farfield_state, dim3(nblk(x->nff)), dim3(kBlock), 0, x->stream, x->d, x->c, 0.0);
    HIPCHK_S(x->stream, hipGetLastError());
    HIPCHK(hipStreamSynchronize(x->stream));
    return FVM_OK;
}

int check_params(const fvm_params *p)
{
    if (p->block_threads != 0 && p->block_threads != 256 && p->block_threads != 512) return FVM_E_INPUT;
    if (p->block_threads == 512 && !has_wide_block(*p)) return FVM_E_INPUT;
    if (p->iflow < 1 || p->iflow > 3) return FVM_E_INPUT;
    if (p->iflux < 1 || p->iflux > 3) return FVM_E_INPUT;
    if (p->ilimit != 0 && p->ilimit != 1) return FVM_E_INPUT;
    if (p->vortex != 0 && p->vortex != 1) return FVM_E_INPUT;
    if (p->cell_type != 1 && p->cell_type != 2) return FVM_E_INPUT;
    if (p->precision != 0 && p->precision != 1) return FVM_E_INPUT;
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


// prim, nut in the reference layout; mu/mul follow (src/initialize.f:60-103)
// caller order <-> device order (vertices renumbered by patch)
std::vector<double> to_device(const fvm_ctx *x, const double *a, int comps)
{
    std::vector<double> out((size_t)x->np * comps);
    for (int n = 0; n < x->np; ++n)
        for (int k = 0; k < comps; ++k) out[(size_t)comps * n + k] = a[(size_t)comps * x->perm[n] + k];
    return out;
}

void to_caller(const fvm_ctx *x, const std::vector<double> &a, int comps, double *out)
{
    for (int n = 0; n < x->np; ++n)
        for (int k = 0; k < comps; ++k) out[(size_t)comps * x->perm[n] + k] = a[(size_t)comps * n + k];
}

int load_state(fvm_ctx *x, const double *prim_caller, const double *nut_caller)
{
    const int np = x->np;
    const Consts &c = x->c;
    std::vector<double> primv = to_device(x, prim_caller, 4), nutv;
    if (nut_caller) nutv = to_device(x, nut_caller, 1);
    const double *prim = primv.data();
    const double *nut = nut_caller ? nutv.data() : nullptr;
    HIPCHK_S(x->stream, hipMemcpyAsync(x->d.prim, prim, sizeof(double) * 4 * np, hipMemcpyHostToDevice, x->stream));
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
    HIPCHK_S(x->stream, hipMemcpyAsync(x->d.nut, nu.data(), sizeof(double) * np, hipMemcpyHostToDevice, x->stream));
    HIPCHK_S(x->stream, hipMemcpyAsync(x->d.mul, mul.data(), sizeof(double) * np, hipMemcpyHostToDevice, x->stream));
    HIPCHK_S(x->stream, hipMemcpyAsync(x->d.mu, mu.data(), sizeof(double) * np, hipMemcpyHostToDevice, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    int rc = reset_state(x);
    if (rc) return rc;
    if (x->d.nblocks > 0) {                 // iteration-start invariant
        launch_tstep(x);
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(x->stream));
    }
    return FVM_OK;
}


// ---- tables of the patch stage kernel ----
struct PatchHost {
    std::vector<int> hdr, halo, gtri, gptr, ledge_id, lbnd_id, pptr;
    std::vector<int4> blob;
    std::vector<short4> ltri;
    std::vector<short> gtl, pslot;
    std::vector<short2> ledge, lbnd;
    std::vector<double2> ledge_ds;
    int lds_max = 0;
    int sa_lds_max = 0;
};

// dynamic LDS per workgroup when the device is not asked: 64 KiB minus the
// static LDS of the block reduction (5 KiB) and a margin
constexpr int kLdsDefault = 58 * 1024;

// Tables of patches [k0, k1) into ph (offsets in the header relative to ph's
// arrays)
int build_patch_range(int np, const std::vector<int> &patch_ptr, const std::vector<int> &aptr,
                      const std::vector<int> &adj, const std::vector<int> &v2t_ptr,
                      const std::vector<int> &v2t, const std::vector<int4> &tri,
                      const std::vector<int> &v2e_ptr, const std::vector<int> &v2e,
                      const std::vector<int2> &enode, const std::vector<int> &v2b_ptr,
                      const std::vector<int> &v2b, const std::vector<int2> &bnode, int iflow,
                      bool split, int threads, int k0, int k1, PatchHost &ph)
{
    const int ntri = (int)tri.size(), nedge = (int)enode.size(), nbnd = (int)bnode.size();
    std::vector<int> loc(np, -1), tloc(ntri, -1), eloc(nedge, -1), bloc(nbnd, -1);
    std::vector<int> lv, tl, el, bl;
    for (int k = k0; k < k1; ++k) {
        const int v0 = patch_ptr[k], v1 = patch_ptr[k + 1], NP = v1 - v0;
        if (NP > threads) return FVM_E_SIZE;
        lv.clear(); tl.clear(); el.clear(); bl.clear();
        for (int v = v0; v < v1; ++v) { loc[v] = (int)lv.size(); lv.push_back(v); }
        for (int i = 0; i < NP; ++i)                       // 1-ring
            for (int j = aptr[lv[i]]; j < aptr[lv[i] + 1]; ++j)
                if (loc[adj[j]] < 0) { loc[adj[j]] = (int)lv.size(); lv.push_back(adj[j]); }
        const int NG = (int)lv.size();
        for (int i = NP; i < NG; ++i)                      // 2-ring
            for (int j = aptr[lv[i]]; j < aptr[lv[i] + 1]; ++j)
                if (loc[adj[j]] < 0) { loc[adj[j]] = (int)lv.size(); lv.push_back(adj[j]); }
        const int NL = (int)lv.size();
        int h[kPatchHdr] = {0};
        h[PH_V0] = v0; h[PH_NP] = NP; h[PH_NL] = NL; h[PH_NG] = NG;
        h[PH_HALO] = (int)ph.halo.size();
        for (int i = NP; i < NL; ++i) ph.halo.push_back(lv[i]);
        // triangles of G, each once; per G-vertex its triangles in order
        h[PH_TRI] = (int)ph.ltri.size();
        h[PH_GPTR] = (int)ph.gptr.size();
        h[PH_GTL] = (int)ph.gtl.size();
        ph.gptr.push_back(0);
        for (int g = 0; g < NG; ++g) {
            const int gv = lv[g];
            for (int j = v2t_ptr[gv]; j < v2t_ptr[gv + 1]; ++j) {
                const int t = v2t[j];
                const int4 n = tri[t];
                if (tloc[t] < 0) {
                    if (loc[n.x] < 0 || loc[n.y] < 0 || loc[n.z] < 0) return FVM_E_INPUT;
                    tloc[t] = (int)tl.size();
                    tl.push_back(t);
                    ph.ltri.push_back(make_short4((short)loc[n.x], (short)loc[n.y], (short)loc[n.z], 0));
                    ph.gtri.push_back(t);
                }
                const int m = (n.x == gv) ? 0 : ((n.y == gv) ? 1 : 2);
                ph.gtl.push_back((short)(4 * tloc[t] + m));
            }
            ph.gptr.push_back((int)ph.gtl.size() - h[PH_GTL]);
            if (g == NP - 1) h[PH_NTP] = (int)tl.size();     // P's triangles come first
        }
        h[PH_NTRI] = (int)tl.size();
        // each patch's triangles start 16-byte aligned (the stage kernel
        // copies them two at a time): an odd count gets a copy of the last
        if (tl.size() & 1) { ph.ltri.push_back(ph.ltri.back()); ph.gtri.push_back(ph.gtri.back()); }
        // edges and boundary edges touching P, each once
        h[PH_EDGE] = (int)ph.ledge.size();
        h[PH_BND] = (int)ph.lbnd.size();
        for (int i = 0; i < NP; ++i) {
            const int v = lv[i];
            for (int j = v2e_ptr[v]; j < v2e_ptr[v + 1]; ++j) {
                const int e = std::abs(v2e[j]) - 1;
                if (eloc[e] < 0) {
                    eloc[e] = (int)el.size();
                    el.push_back(e);
                    ph.ledge.push_back(make_short2((short)loc[enode[e].x], (short)loc[enode[e].y]));
                    ph.ledge_id.push_back(e);
                }
            }
        }
        for (int i = 0; i < NP; ++i) {
            const int v = lv[i];
            for (int j = v2b_ptr[v]; j < v2b_ptr[v + 1]; ++j) {
                const int b = v2b[j] >> 1;
                if (bloc[b] < 0) {
                    bloc[b] = (int)bl.size();
                    bl.push_back(b);
                    ph.lbnd.push_back(make_short2((short)loc[bnode[b].x], (short)loc[bnode[b].y]));
                    ph.lbnd_id.push_back(b);
                }
            }
        }
        const int NE = (int)el.size(), NB = (int)bl.size();
        h[PH_NEDGE] = NE;
        h[PH_NBND] = NB;
        // per P-vertex terms: edges in edge order (+/-), then boundary edges
        h[PH_PPTR] = (int)ph.pptr.size();
        h[PH_PSL] = (int)ph.pslot.size();
        ph.pptr.push_back(0);
        for (int i = 0; i < NP; ++i) {
            const int v = lv[i];
            for (int j = v2e_ptr[v]; j < v2e_ptr[v + 1]; ++j) {
                const int sgn = v2e[j];
                ph.pslot.push_back((short)(2 * eloc[std::abs(sgn) - 1] + (sgn < 0 ? 1 : 0)));
            }
            for (int j = v2b_ptr[v]; j < v2b_ptr[v + 1]; ++j) {
                const int q = v2b[j];
                ph.pslot.push_back((short)(2 * (NE + 2 * bloc[q >> 1] + (q & 1))));
            }
            ph.pptr.push_back((int)ph.pslot.size() - h[PH_PSL]);
        }
        if (NL > 32767 || 4 * (int)tl.size() > 32767 || 2 * (NE + 2 * NB) > 32767) return FVM_E_SIZE;
        h[PH_NGTL] = (int)ph.gtl.size() - h[PH_GTL];
        h[PH_NPSL] = (int)ph.pslot.size() - h[PH_PSL];
        {
            // gptr, pptr, gtl, pslot of the patch back to back, as the stage
            // kernel lays them out in LDS, in whole 16-byte units
            const int nb = 4 * (NG + 1) + 4 * (NP + 1) + 2 * (h[PH_NGTL] + h[PH_NPSL]);
            std::vector<char> img((nb + 15) & ~15, 0);
            char *q = img.data();
            std::memcpy(q, &ph.gptr[h[PH_GPTR]], 4 * (NG + 1)); q += 4 * (NG + 1);
            std::memcpy(q, &ph.pptr[h[PH_PPTR]], 4 * (NP + 1)); q += 4 * (NP + 1);
            std::memcpy(q, &ph.gtl[h[PH_GTL]], 2 * h[PH_NGTL]); q += 2 * h[PH_NGTL];
            std::memcpy(q, &ph.pslot[h[PH_PSL]], 2 * h[PH_NPSL]);
            h[PH_BLOB] = (int)ph.blob.size();
            h[PH_NBLOB] = (int)img.size() / 16;
            const size_t b0 = ph.blob.size();
            ph.blob.resize(b0 + img.size() / 16);
            std::memcpy(&ph.blob[b0], img.data(), img.size());
        }
        const int lds = patch_lds_bytes(NP, NL, NG, (int)tl.size(), NE, NB, h[PH_NGTL], h[PH_NPSL],
                                        h[PH_NTP], iflow, split);
        ph.lds_max = std::max(ph.lds_max, lds);
        ph.sa_lds_max = std::max(ph.sa_lds_max, sa_lds_bytes(NP, NG, (int)tl.size(), NE, NB, h[PH_NGTL],
                                                               h[PH_NPSL]));
        for (int i = 0; i < kPatchHdr; ++i) ph.hdr.push_back(h[i]);
        for (int v : lv) loc[v] = -1;
        for (int t : tl) tloc[t] = -1;
        for (int e : el) eloc[e] = -1;
        for (int b : bl) bloc[b] = -1;
    }
    return FVM_OK;
}

// Tables of every patch; large meshes build ranges of patches on host
// threads and join them in patch order (the same tables as one thread).
int build_patches(int np, const std::vector<int> &patch_ptr, const std::vector<int> &aptr,
                  const std::vector<int> &adj, const std::vector<int> &v2t_ptr,
                  const std::vector<int> &v2t, const std::vector<int4> &tri,
                  const std::vector<int> &v2e_ptr, const std::vector<int> &v2e,
                  const std::vector<int2> &enode, const std::vector<int> &v2b_ptr,
                  const std::vector<int> &v2b, const std::vector<int2> &bnode, int iflow,
                  bool split, int threads, int lds_limit, PatchHost &ph)
{
    const int npatch = (int)patch_ptr.size() - 1;
    const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
    const int nth = np < 200000 ? 1 : (int)std::min<unsigned>(8u, hw);
    std::vector<PatchHost> part(nth);
    std::vector<int> rcs(nth, FVM_OK);
    auto work = [&](int t) {
        const int k0 = (int)((long long)npatch * t / nth), k1 = (int)((long long)npatch * (t + 1) / nth);
        rcs[t] = build_patch_range(np, patch_ptr, aptr, adj, v2t_ptr, v2t, tri, v2e_ptr, v2e, enode, v2b_ptr,
                                   v2b, bnode, iflow, split, threads, k0, k1, part[t]);
    };
    if (nth == 1) {
        work(0);
    } else {
        std::vector<std::thread> th;
        for (int t = 0; t < nth; ++t) th.emplace_back(work, t);
        for (auto &x : th) x.join();
    }
    for (int t = 0; t < nth; ++t)
        if (rcs[t]) return rcs[t];
    ph = std::move(part[0]);
    for (int t = 1; t < nth; ++t) {
        PatchHost &q = part[t];
        const int off[][2] = {{PH_HALO, (int)ph.halo.size()}, {PH_TRI, (int)ph.ltri.size()},
                              {PH_GPTR, (int)ph.gptr.size()}, {PH_GTL, (int)ph.gtl.size()},
                              {PH_EDGE, (int)ph.ledge.size()}, {PH_BND, (int)ph.lbnd.size()},
                              {PH_PPTR, (int)ph.pptr.size()}, {PH_PSL, (int)ph.pslot.size()},
                              {PH_BLOB, (int)ph.blob.size()}};
        for (size_t h = 0; h < q.hdr.size(); h += kPatchHdr)
            for (const auto &o : off) q.hdr[h + o[0]] += o[1];
        auto cat = [](auto &a, const auto &b) { a.insert(a.end(), b.begin(), b.end()); };
        cat(ph.hdr, q.hdr); cat(ph.halo, q.halo); cat(ph.gtri, q.gtri); cat(ph.gptr, q.gptr);
        cat(ph.ledge_id, q.ledge_id); cat(ph.lbnd_id, q.lbnd_id); cat(ph.pptr, q.pptr);
        cat(ph.ltri, q.ltri); cat(ph.gtl, q.gtl); cat(ph.pslot, q.pslot); cat(ph.ledge, q.ledge);
        cat(ph.lbnd, q.lbnd); cat(ph.ledge_ds, q.ledge_ds); cat(ph.blob, q.blob);
        ph.lds_max = std::max(ph.lds_max, q.lds_max);
        ph.sa_lds_max = std::max(ph.sa_lds_max, q.sa_lds_max);
        q = PatchHost();                          // frees the fragment
    }
    return ph.lds_max > lds_limit ? FVM_E_SIZE : FVM_OK;
}


// Host layout of one mesh on the device: vertices renumbered patch by patch
// (triangle and edge numbers keep the reference's order, which fixes the
// summation order; only storage moves), gather tables of every kernel
// variant, and the patch tables of the stage kernel.  No HIP call.
struct Layout {
    std::vector<int> perm, iperm, patch_ptr, nadj_ptr, nadj;
    std::vector<double2> coord, eds;
    std::vector<int4> tri;
    std::vector<int> v2t_ptr, v2t, v2e_ptr, v2e, v2b_ptr, v2b, orig;
    std::vector<int2> enode, bnode;
    std::vector<double> ais, cvarea, drmin;
    std::vector<unsigned char> solid;
    PatchHost ph;
    int target = 0;                                // patch size asked of the partition
};

int build_layout(const fvm_mesh *m, int iflow, int precision, int patch_size, int lds_limit,
                 Layout &L, int threads = kBlock)
{
    const int np = m->np, nt = m->nt, ne = m->ne;
    const int nsw = std::max(0, m->nswe2 - m->nswe1 + 1), nff = std::max(0, m->nffe2 - m->nffe1 + 1);
    int target = patch_size > 0 ? std::min(patch_size, threads) : 128;
    // Vertices are renumbered patch by patch (compact patches, contiguous
    // ids).  Triangle and edge numbers keep the reference's order, which is
    // what fixes the summation order; only storage moves.
    std::vector<int> adj_ptr(np + 1, 0), adj;
    for (int e = 0; e < m->netot; ++e) { ++adj_ptr[m->edge[2 * e]]; ++adj_ptr[m->edge[2 * e + 1]]; }
    std::partial_sum(adj_ptr.begin(), adj_ptr.end(), adj_ptr.begin());
    adj.resize(adj_ptr[np]);
    {
        std::vector<int> f(adj_ptr.begin(), adj_ptr.end() - 1);
        for (int e = 0; e < m->netot; ++e) {
            int a0 = m->edge[2 * e] - 1, a1 = m->edge[2 * e + 1] - 1;
            adj[f[a0]++] = a1;
            adj[f[a1]++] = a0;
        }
    }
    fvm::Partition part = fvm::partition_vertices(np, adj_ptr, adj, target);
    L.target = target;
    std::vector<int> &perm = part.perm;            // new -> old
    std::vector<int> iperm(np);                    // old -> new
    for (int n = 0; n < np; ++n) iperm[perm[n]] = n;
    std::vector<int> nadj_ptr(np + 1, 0), nadj(adj.size());
    for (int n = 0; n < np; ++n) {
        int o = perm[n];
        nadj_ptr[n + 1] = nadj_ptr[n] + (adj_ptr[o + 1] - adj_ptr[o]);
        for (int k = adj_ptr[o]; k < adj_ptr[o + 1]; ++k) nadj[nadj_ptr[n] + k - adj_ptr[o]] = iperm[adj[k]];
    }

    // ---- host-side layout (0-based, new vertex numbers) ----
    std::vector<double2> coord(np);
    for (int n = 0; n < np; ++n) coord[n] = make_double2(m->coord[2 * perm[n]], m->coord[2 * perm[n] + 1]);
    std::vector<int4> tri(nt);
    for (int t = 0; t < nt; ++t)
        tri[t] = make_int4(iperm[m->elem[3 * t] - 1], iperm[m->elem[3 * t + 1] - 1],
                           iperm[m->elem[3 * t + 2] - 1], 0);
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
        enode[i] = make_int2(iperm[m->edge[2 * ie] - 1], iperm[m->edge[2 * ie + 1] - 1]);
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
    int nb = nsw + nff;
    std::vector<int2> bnode(std::max(nb, 1));
    for (int b = 0; b < nb; ++b) {
        int ie = m->beindx[m->nswe1 - 1 + b] - 1;
        bnode[b] = make_int2(iperm[m->edge[2 * ie] - 1], iperm[m->edge[2 * ie + 1] - 1]);
    }
    std::vector<int> v2b_ptr(np + 1, 0), v2b(2 * (size_t)std::max(nb, 1));
    for (int b = 0; b < nb; ++b) { ++v2b_ptr[bnode[b].x + 1]; ++v2b_ptr[bnode[b].y + 1]; }
    std::partial_sum(v2b_ptr.begin(), v2b_ptr.end(), v2b_ptr.begin());
    {
        std::vector<int> f(v2b_ptr.begin(), v2b_ptr.end() - 1);
        for (int b = 0; b < nb; ++b) { v2b[f[bnode[b].x]++] = 2 * b; v2b[f[bnode[b].y]++] = 2 * b + 1; }
    }
    std::vector<double> ais(np), cvarea(np), drmin(np);
    const double n1b6 = 1.0 / 6.0;
    for (int n = 0; n < np; ++n) {
        ais[n] = n1b6 / m->mcarea[perm[n]];
        cvarea[n] = m->cvarea[perm[n]];
        drmin[n] = m->drmin[perm[n]];
    }
    std::vector<unsigned char> solid(np, 0);
    for (int k = 0; k < m->nsp; ++k) solid[iperm[m->spts[k] - 1]] = 1;
    std::vector<int> orig(np);
    for (int n = 0; n < np; ++n) orig[n] = perm[n] + 1;


    // patch tables; when a patch does not fit the LDS budget, partition and
    // renumber again for patches of half the size
    L.ph = PatchHost();
    const int prc = build_patches(np, part.patch_ptr, nadj_ptr, nadj, v2t_ptr, v2t, tri, v2e_ptr, v2e,
                                  enode, v2b_ptr, v2b, bnode, iflow, precision == 1, threads, lds_limit,
                                  L.ph);
    if (prc == FVM_E_SIZE && target > 16)
        return build_layout(m, iflow, precision, target / 2, lds_limit, L, threads);
    if (prc != FVM_OK) return prc;
    L.perm = std::move(perm); L.iperm = std::move(iperm); L.patch_ptr = std::move(part.patch_ptr);
    L.nadj_ptr = std::move(nadj_ptr); L.nadj = std::move(nadj);
    L.coord = std::move(coord); L.eds = std::move(eds); L.tri = std::move(tri);
    L.v2t_ptr = std::move(v2t_ptr); L.v2t = std::move(v2t); L.v2e_ptr = std::move(v2e_ptr);
    L.v2e = std::move(v2e); L.v2b_ptr = std::move(v2b_ptr); L.v2b = std::move(v2b);
    L.orig = std::move(orig);
    L.enode = std::move(enode); L.bnode = std::move(bnode); L.ais = std::move(ais);
    L.cvarea = std::move(cvarea); L.drmin = std::move(drmin); L.solid = std::move(solid);
    return FVM_OK;
}


// Checks every index the kernels dereference against the array it indexes.
// Returns 0 or the number of the first failed check.
int validate_layout(const fvm_mesh *m, const Layout &L, int threads = kBlock)
{
    const int np = m->np, nt = m->nt, ne = m->ne;
    const int nb = (int)L.bnode.size();
    std::vector<char> seen(np, 0);
    if ((int)L.perm.size() != np) return 1;
    for (int n = 0; n < np; ++n) {
        int o = L.perm[n];
        if (o < 0 || o >= np || seen[o]) return 2;
        seen[o] = 1;
        if (L.iperm[o] != n) return 3;
    }
    for (const int4 &t : L.tri)
        if (t.x < 0 || t.x >= np || t.y < 0 || t.y >= np || t.z < 0 || t.z >= np) return 4;
    for (const int2 &e : L.enode) if (e.x < 0 || e.x >= np || e.y < 0 || e.y >= np) return 5;
    for (int b = 0; b < std::max(0, std::max(0, m->nffe2 - m->nswe1 + 1)); ++b)
        if (L.bnode[b].x < 0 || L.bnode[b].x >= np || L.bnode[b].y < 0 || L.bnode[b].y >= np) return 6;
    const PatchHost &ph = L.ph;
    const int npatch = (int)L.patch_ptr.size() - 1;
    if ((int)ph.hdr.size() != kPatchHdr * npatch || L.patch_ptr.back() != np) return 7;
    for (int k = 0; k < npatch; ++k) {
        const int *h = &ph.hdr[kPatchHdr * k];
        const int v0 = h[PH_V0], NP = h[PH_NP], NL = h[PH_NL], NG = h[PH_NG];
        const int NT = h[PH_NTRI], NE = h[PH_NEDGE], NB = h[PH_NBND];
        if (h[PH_NGTL] != ph.gptr[h[PH_GPTR] + NG] || h[PH_NPSL] != ph.pptr[h[PH_PPTR] + NP]) return 25;
        if (v0 != L.patch_ptr[k] || NP != L.patch_ptr[k + 1] - v0 || NP < 1 || NP > threads) return 8;
        if (NG < NP || NL < NG || h[PH_HALO] + (NL - NP) > (int)ph.halo.size()) return 9;
        for (int i = 0; i < NL - NP; ++i) {
            int gv = ph.halo[h[PH_HALO] + i];
            if (gv < 0 || gv >= np || (gv >= v0 && gv < v0 + NP)) return 10;
        }
        if (h[PH_TRI] + NT > (int)ph.ltri.size() || h[PH_GPTR] + NG + 1 > (int)ph.gptr.size()) return 11;
        for (int i = 0; i < NT; ++i) {
            short4 t = ph.ltri[h[PH_TRI] + i];
            if (t.x < 0 || t.x >= NL || t.y < 0 || t.y >= NL || t.z < 0 || t.z >= NL) return 12;
            if (ph.gtri[h[PH_TRI] + i] < 0 || ph.gtri[h[PH_TRI] + i] >= nt) return 13;
        }
        const int *gp = &ph.gptr[h[PH_GPTR]];
        if (gp[0] != 0) return 14;
        for (int g = 0; g < NG; ++g) {
            if (gp[g + 1] < gp[g] || h[PH_GTL] + gp[g + 1] > (int)ph.gtl.size()) return 15;
            int prev_t = -1;
            for (int j = gp[g]; j < gp[g + 1]; ++j) {
                int e = ph.gtl[h[PH_GTL] + j], lt = e >> 2, mm = e & 3;
                if (lt < 0 || lt >= NT || mm > 2) return 16;
                short4 t = ph.ltri[h[PH_TRI] + lt];
                int node = mm == 0 ? t.x : (mm == 1 ? t.y : t.z);
                if (node != g) return 17;
                int gt = ph.gtri[h[PH_TRI] + lt];
                if (gt <= prev_t) return 18;                 // ascending triangle order
                prev_t = gt;
            }
        }
        if (h[PH_EDGE] + NE > (int)ph.ledge.size() || h[PH_BND] + NB > (int)ph.lbnd.size()) return 19;
        for (int i = 0; i < NE; ++i) {
            short2 e = ph.ledge[h[PH_EDGE] + i];
            int id = ph.ledge_id[h[PH_EDGE] + i];
            if (e.x < 0 || e.x >= NG || e.y < 0 || e.y >= NG || id < 0 || id >= ne) return 20;
        }
        for (int i = 0; i < NB; ++i) {
            short2 e = ph.lbnd[h[PH_BND] + i];
            int id = ph.lbnd_id[h[PH_BND] + i];
            if (e.x < 0 || e.x >= NL || e.y < 0 || e.y >= NL || id < 0 || id >= nb) return 21;
        }
        if (h[PH_PPTR] + NP + 1 > (int)ph.pptr.size()) return 22;
        {
            // the blob is the LDS image of gptr, pptr, gtl, pslot
            const int nb = 4 * (NG + 1) + 4 * (NP + 1) + 2 * (h[PH_NGTL] + h[PH_NPSL]);
            if (h[PH_NBLOB] != (nb + 15) / 16 || h[PH_BLOB] < 0 ||
                h[PH_BLOB] + h[PH_NBLOB] > (int)ph.blob.size() || (h[PH_TRI] & 1))
                return 26;
            const char *q = reinterpret_cast<const char *>(&ph.blob[h[PH_BLOB]]);
            if (std::memcmp(q, &ph.gptr[h[PH_GPTR]], 4 * (NG + 1)) ||
                std::memcmp(q + 4 * (NG + 1), &ph.pptr[h[PH_PPTR]], 4 * (NP + 1)) ||
                std::memcmp(q + 4 * (NG + 1) + 4 * (NP + 1), &ph.gtl[h[PH_GTL]], 2 * h[PH_NGTL]) ||
                std::memcmp(q + 4 * (NG + 1) + 4 * (NP + 1) + 2 * h[PH_NGTL], &ph.pslot[h[PH_PSL]],
                            2 * h[PH_NPSL]))
                return 27;
        }
        const int *pp = &ph.pptr[h[PH_PPTR]];
        for (int i = 0; i < NP; ++i) {
            const int v = v0 + i;
            int want = (L.v2e_ptr[v + 1] - L.v2e_ptr[v]) + (L.v2b_ptr[v + 1] - L.v2b_ptr[v]);
            if (pp[i + 1] - pp[i] != want || h[PH_PSL] + pp[i + 1] > (int)ph.pslot.size()) return 23;
            for (int j = pp[i]; j < pp[i + 1]; ++j) {
                int q = ph.pslot[h[PH_PSL] + j];
                if (q < 0 || (q >> 1) >= NE + 2 * NB) return 24;
            }
        }
    }
    return 0;
}

}  // namespace

extern "C" {

// Restart (src/initialize.f:91-103): that path never sets nut_inf, which
// keeps the zero of its COMMON block; the farfield inflow of nu_tilde uses it.
int fvm_gpu_set_state(fvm_ctx *x, const double *prim, const double *nut)
{
    if (!ctx_ok(x) || !prim) return FVM_E_INPUT;
    if (int rc = settle(x)) return rc;
    x->c.nut_inf = 0.0;
    HIPCHK(hipStreamSynchronize(x->stream));
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
    (void)precision;                       // the same buffers in both precisions
    const long long P = np, T = nt, E = netot;
    long long b = 0;
    // per vertex: coord, stage buffers O/B/C, prim_old, qx, qy, residual and
    // hook backups, dt, cvarea, ais, drmin, mu, mul, nut, nut_new, wd,
    // solid, orig, adjacency offsets
    b += P * (16 + 3 * 32 + 32 + 2 * 32 + 2 * 32 + 8 * 11 + 1 + 4 + 3 * 4);
    b += T * (12 + 16 + 8);                // v2t, tri, elarea
    b += E * (8 + 16 + 32 + 8);            // enode, eds, eflux, v2e
    // patch tables: local triangles (shorts, number, area), Green-Gauss and
    // residual term lists, local edges (ends, number, normal, frame), halo,
    // headers; bounded for patches down to 24 vertices
    b += P * 900;
    b += 6LL * 8 * kHistCap + 64 * 1024;
    (void)iflow;
    return b;
}

int fvm_patch_plan(const fvm_mesh *m, int iflow, int precision, int patch_size, int lds_limit,
                   long long *st)
{
    for (int i = 0; i < 8; ++i) st[i] = 0;
    if (m->np <= 0 || m->nt <= 0 || m->ne <= 0) return FVM_E_INPUT;
    Layout L;
    int rc = build_layout(m, iflow, precision, patch_size, lds_limit > 0 ? lds_limit : kLdsDefault, L);
    if (rc) return rc;
    const PatchHost &ph = L.ph;
    const int npatch = (int)L.patch_ptr.size() - 1;
    st[0] = npatch;
    for (int k = 0; k < npatch; ++k) {
        const int *h = &ph.hdr[kPatchHdr * k];
        st[1] = std::max<long long>(st[1], h[PH_NP]);
        st[2] = std::max<long long>(st[2], h[PH_NL]);
        st[3] = std::max<long long>(st[3], h[PH_NTRI]);
        st[4] = std::max<long long>(st[4], h[PH_NEDGE]);
        st[6] += h[PH_NL] - h[PH_NP];
    }
    st[5] = ph.lds_max;
    st[7] = validate_layout(m, L);
    return st[7] ? FVM_E_INPUT : FVM_OK;
}

int fvm_gpu_layout_info(const fvm_ctx *x, long long *info)
{
    if (!ctx_ok(x) || !info) return FVM_E_INPUT;
    info[0] = x->pa.npatch;
    info[1] = x->max_patch;
    info[2] = x->pa.lds_bytes;
    info[3] = stage_per_cu(x, x->pa.lds_bytes);
    info[4] = x->cus;
    info[5] = x->target;
    info[6] = x->halo_total;
    info[7] = x->variant;
    return FVM_OK;
}

long long fvm_gpu_footprint(const fvm_ctx *x) { return ctx_ok(x) ? x->bytes : 0; }

int fvm_gpu_workgroup_threads(const fvm_ctx *x) { return ctx_ok(x) ? x->threads : FVM_E_INPUT; }

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
    keep_pool_memory();

    // LDS a workgroup may use on this device (less the static LDS), CUs and
    // LDS per CU for the automatic patch size
    int dev = 0, lds_dev = 64 * 1024, cus = 256, lds_cu = 160 * 1024;
    if (hipGetDevice(&dev) == hipSuccess) {
        (void)hipDeviceGetAttribute(&lds_dev, hipDeviceAttributeMaxSharedMemoryPerBlock, dev);
        (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
        (void)hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, dev);
    }
    x->cus = cus;
    x->per_cu = par->iflow == 3 ? kPatchWavesSA
                                : (par->iflow == 2 ? kPatchWavesLaminar
                                                   : (par->precision == 1 ? kPatchWavesSplit : kPatchWavesEuler));
    // Workgroup size: the caller's, or with the automatic launch shape 512
    // threads for limited Euler solves of at least 32 768 vertices (two
    // workgroups a CU on patches of about 130 vertices, fewer halo vertices
    // per patch vertex: 513x129 FP64 with Roe 41.8 against 45.0 us per
    // iteration; KFVS and HLLC stages 4-8% shorter), else 256.
    x->threads = par->block_threads ? par->block_threads
                                    : ((par->patch_size == 0 && has_wide_block(*par) && np >= 32768) ? 512 : kBlock);
    const int scale = x->threads / kBlock;          // 1, or 2 for 512-thread workgroups
    Layout L;
    if (par->patch_size > 0) {
        rc = build_layout(m, par->iflow, par->precision, par->patch_size, lds_dev - 6 * 1024, L, x->threads);
    } else {
        // Whole rounds of resident workgroups, about 64 vertices a patch
        // (128 with 512 threads): k rounds of cus * per_cu / scale
        // workgroups less one, the slot where the first stage runs the
        // previous iteration's monitors beside the patches (with every slot
        // taken they would wait for a second round); fewer vertices per patch
        // while the LDS would keep fewer workgroups on a CU.
        const long wg = x->per_cu / scale, slots = (long)cus * wg;
        long k = std::max(1L, std::lround(np / (64.0 * scale * slots)));
        bool done = false;
        if (scale == 2 && np / (192L * slots) >= 4) {
            // large meshes: patches of about 190 vertices where two still
            // fit a CU's LDS (1 050 624 vertices, FP64: 490 against 590 us
            // per iteration at 129)
            const long room = std::lround(np / (192.0 * slots)) * slots - 1;
            L = Layout();
            rc = build_layout(m, par->iflow, par->precision, (int)((np + room - 1) / room), lds_dev - 6 * 1024,
                              L, x->threads);
            done = rc != FVM_OK || lds_alloc(L.ph.lds_max) * wg <= lds_cu;
        }
        for (int tries = 0; !done; ++tries, ++k) {
            const long room = k * slots - 1;
            const int ps = (int)std::min(128L * scale, std::max(24L * scale, (np + room - 1) / room));
            L = Layout();
            rc = build_layout(m, par->iflow, par->precision, ps, lds_dev - 6 * 1024, L, x->threads);
            if (rc || tries == 3 || ps == 24 * scale || lds_alloc(L.ph.lds_max) * wg <= lds_cu) done = true;
        }
    }
    if (!rc) x->lds_per_cu = lds_cu;
    if (!rc && validate_layout(m, L, x->threads)) rc = FVM_E_INPUT;
    if (rc) { fvm_gpu_destroy(x); return rc; }
    x->perm = L.perm;
    const std::vector<int> &iperm = L.iperm;
    const int nb = x->nsw + x->nff;
    // ---- device buffers ----
    Dev &d = x->d;
    d.np = np; d.nt = nt; d.ne = ne; d.nsw = x->nsw; d.nff = x->nff; d.nsp = m->nsp;
    d.nparts = (int)nblk(np);
    d.sa1 = iperm[0]; d.sa2 = iperm[std::min(1, np - 1)]; d.sa3 = iperm[std::min(2, np - 1)];
#define TRY(e) do { rc = (e); if (rc) { fvm_gpu_destroy(x); return rc; } } while (0)
    TRY(upload(x, (double2 **)&d.coord, L.coord));
    TRY(upload(x, (int4 **)&d.tri, L.tri));
    TRY(upload(x, (int **)&d.v2t_ptr, L.v2t_ptr));
    TRY(upload(x, (int **)&d.v2t, L.v2t));
    TRY(upload(x, (int2 **)&d.enode, L.enode));
    TRY(upload(x, (double2 **)&d.eds, L.eds));
    TRY(upload(x, (int **)&d.v2e_ptr, L.v2e_ptr));
    TRY(upload(x, (int **)&d.v2e, L.v2e));
    TRY(upload(x, (int2 **)&d.bnode, L.bnode));
    TRY(upload(x, (int **)&d.v2b_ptr, L.v2b_ptr));
    TRY(upload(x, (int **)&d.v2b, L.v2b));
    TRY(upload(x, (double **)&d.ais, L.ais));
    TRY(upload(x, (double **)&d.cvarea, L.cvarea));
    TRY(upload(x, (double **)&d.drmin, L.drmin));
    TRY(upload(x, (unsigned char **)&d.solid, L.solid));
    TRY(upload(x, (int **)&d.orig, L.orig));
    TRY(upload(x, (int **)&x->pa.hdr, L.ph.hdr));
    TRY(upload(x, (int **)&x->pa.halo, L.ph.halo));
    TRY(upload(x, (short4 **)&x->pa.ltri, L.ph.ltri));
    TRY(upload(x, (int **)&x->pa.gtri, L.ph.gtri));
    {
        // 0.25 / elarea per local triangle: the quotient the viscous and SA
        // terms use, divided once here (IEEE division, the same bits as on
        // the device) instead of once per vertex and triangle
        std::vector<double> lqinv(L.ph.gtri.size());
        for (size_t i = 0; i < lqinv.size(); ++i) lqinv[i] = 0.25 / m->elarea[L.ph.gtri[i]];
        TRY(upload(x, (double **)&x->pa.lqinv, lqinv));
        std::vector<double2> lds(L.ph.ledge_id.size());
        for (size_t i = 0; i < lds.size(); ++i) lds[i] = L.eds[L.ph.ledge_id[i]];
        TRY(upload(x, (double2 **)&x->pa.ledge_ds, lds));
        // the frame of each local edge's ds (length, then the two IEEE
        // quotients: the fluxes' own first operations, done once here)
        std::vector<double4> lfr(lds.size());
        for (size_t i = 0; i < lfr.size(); ++i) {
            const double length = std::sqrt(lds[i].x * lds[i].x + lds[i].y * lds[i].y);
            lfr[i] = make_double4(lds[i].x / length, lds[i].y / length, length, 0.0);
        }
        TRY(upload(x, (double4 **)&x->pa.ledge_fr, lfr));
    }
    TRY(upload(x, (int **)&x->pa.gptr, L.ph.gptr));
    TRY(upload(x, (short **)&x->pa.gtl, L.ph.gtl));
    TRY(upload(x, (short2 **)&x->pa.ledge, L.ph.ledge));
    TRY(upload(x, (int **)&x->pa.ledge_id, L.ph.ledge_id));
    TRY(upload(x, (short2 **)&x->pa.lbnd, L.ph.lbnd));
    TRY(upload(x, (int **)&x->pa.lbnd_id, L.ph.lbnd_id));
    TRY(upload(x, (int **)&x->pa.pptr, L.ph.pptr));
    TRY(upload(x, (short **)&x->pa.pslot, L.ph.pslot));
    TRY(upload(x, (int4 **)&x->pa.blob, L.ph.blob));
    x->pa.npatch = (int)L.patch_ptr.size() - 1;
    x->pa.lds_bytes = L.ph.lds_max;
    x->pa.sa_lds_bytes = L.ph.sa_lds_max;
    x->target = L.target;
    {
        long long halo = 0; int maxp = 0;
        x->patch_cost.resize(x->pa.npatch);
        for (int k = 0; k < x->pa.npatch; ++k) {
            const int *hk = &L.ph.hdr[kPatchHdr * k];
            halo += hk[PH_NL] - hk[PH_NP];
            maxp = std::max(maxp, hk[PH_NP]);
            int nff = 0;
            for (int q = 0; q < hk[PH_NBND]; ++q) nff += L.ph.lbnd_id[hk[PH_BND] + q] >= x->nsw;
            x->patch_cost[k] = patch_work(hk[PH_NEDGE], nff, hk[PH_NBND] - nff, par->precision);
        }
        x->halo_total = halo; x->max_patch = maxp;
    }
    TRY(dalloc(x, &d.prim, np));
    TRY(dalloc(x, &x->pbuf[1], np));
    TRY(dalloc(x, &x->pbuf[2], np));
    TRY(dalloc(x, &d.prim_old, np));
    TRY(dalloc(x, &d.qx, np));
    TRY(dalloc(x, &d.qy, np));
    TRY(dalloc(x, &d.dt, np));
    TRY(dalloc(x, &d.eflux, ne));
    TRY(dalloc(x, &d.bflux, 2 * (size_t)std::max(nb, 1)));
    TRY(dalloc(x, &d.uout, (size_t)std::max(x->nff, 1)));
    TRY(dalloc(x, &d.wall, 2 * (size_t)std::max(x->nsw, 1)));
    TRY(hipMemsetAsync(d.wall, 0, sizeof(double4) * 2 * std::max(x->nsw, 1), x->stream) == hipSuccess ? FVM_OK
                                                                                                   : FVM_E_HIP);
    {
        std::vector<double2> ws(std::max(x->nsw, 1), make_double2(0.0, 0.0));
        for (int b = 0; b < x->nsw; ++b) {
            const double2 x1 = L.coord[L.bnode[b].x], x2 = L.coord[L.bnode[b].y];
            ws[b] = make_double2(x2.y - x1.y, -(x2.x - x1.x));
        }
        TRY(upload(x, (double2 **)&d.wall_s, ws));
    }
    TRY(dalloc(x, &d.partial, 3 * (size_t)(np / 16 + 2 + (int)L.patch_ptr.size())));
    TRY(dalloc(x, &d.hist, 6 * (size_t)kHistCap));
    TRY(dalloc(x, &x->S, 1));
    TRY(dalloc(x, &x->S_hook, 1));
    TRY(dalloc(x, &x->res_tmp, np));
    TRY(dalloc(x, &x->prim_bak, np));
    TRY(dalloc(x, &x->nut_bak, np));
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
        std::vector<double> wd(np);
        for (int n = 0; n < np; ++n) wd[n] = m->wd[L.perm[n]];
        TRY(upload(x, (double **)&d.wd, wd));
        TRY(dalloc(x, &d.nut_new, np));
    }
    TRY(dalloc(x, &x->self_mem, 1));
    {
        std::vector<int2> w(x->pa.npatch);
        for (int k = 0; k < x->pa.npatch; ++k) w[k] = make_int2(0, k);
        const int pcu = stage_per_cu(x, x->pa.lds_bytes);
        const std::vector<int2> wm = with_monitors(w, x->patch_cost, 1, x->cus, pcu);
        order_work(w, x->patch_cost, x->cus, pcu);
        TRY(upload(x, (int2 **)&x->self_work, w));
        TRY(upload(x, (int2 **)&x->self_work_mon, wm));
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
    x->d.vb = 64;
    x->d.nblocks = (np + 63) / 64;
    x->d.xcd = 1;
    x->d.nparts = x->d.nblocks;
    x->pbuf[0] = d.prim;
    rc = load_state(x, prim.data(), nut.data());
    if (!rc) rc = configure(x, 3, 64, 1);
    if (rc) { fvm_gpu_destroy(x); return rc; }
    live_add(x);
    *out = x;
    return FVM_OK;
}

int fvm_gpu_destroy(fvm_ctx *x)
{
    if (!x) return FVM_OK;
    if (!ctx_ok(x)) return FVM_E_INPUT;
    if (x->batch) (void)fvm_batch_remove(x->batch, x);
    if (x->stream) (void)hipStreamSynchronize(x->stream);
    drop_graphs(x);
    if (x->stream) {
        for (void *p : x->allocs) (void)hipFreeAsync(p, x->stream);
        (void)hipStreamSynchronize(x->stream);
        (void)hipStreamDestroy(x->stream);
    }
    live_drop(x);
    x->tag = 0;
    delete x;
    return FVM_OK;
}


int fvm_gpu_wall_dist(int np, const double *coord, const int *ptype, int netot, const int *edge,
                      int sa_compat, double *wd, double *wd1)
{
    if (np <= 0 || netot <= 0) return FVM_E_INPUT;
    std::vector<int2> se;
    for (int e = 0; e < netot; ++e) {
        int a = edge[2 * e] - 1, b = edge[2 * e + 1] - 1;
        if (a < 0 || a >= np || b < 0 || b >= np) return FVM_E_INPUT;
        if (ptype[a] == 3 && ptype[b] == 3) se.push_back(make_int2(a, b));
    }
    std::vector<unsigned char> solid(np);
    for (int p = 0; p < np; ++p) solid[p] = ptype[p] == 3;
    double2 *dxy = nullptr; int2 *dse = nullptr; unsigned char *dsol = nullptr; double *dwd = nullptr;
    hipStream_t st = nullptr;
    int rc = FVM_OK;
    auto cleanup = [&] {
        if (!st) return;
        (void)hipStreamSynchronize(st);
        for (void *p : {(void *)dxy, (void *)dse, (void *)dsol, (void *)dwd})
            if (p) (void)hipFreeAsync(p, st);
        (void)hipStreamSynchronize(st);
        (void)hipStreamDestroy(st);
    };
    if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return FVM_E_HIP;
    keep_pool_memory();
    if (hipMallocAsync((void **)&dxy, sizeof(double2) * np, st) != hipSuccess ||
        hipMallocAsync((void **)&dse, sizeof(int2) * std::max<size_t>(se.size(), 1), st) != hipSuccess ||
        hipMallocAsync((void **)&dsol, np, st) != hipSuccess ||
        hipMallocAsync((void **)&dwd, sizeof(double) * np, st) != hipSuccess) {
        (void)hipGetLastError();
        cleanup();
        return FVM_E_NOMEM;
    }
    if (hipMemcpyAsync(dxy, coord, sizeof(double2) * np, hipMemcpyHostToDevice, st) != hipSuccess ||
        (!se.empty() && hipMemcpyAsync(dse, se.data(), sizeof(int2) * se.size(), hipMemcpyHostToDevice, st) != hipSuccess) ||
        hipMemcpyAsync(dsol, solid.data(), np, hipMemcpyHostToDevice, st) != hipSuccess) {
        cleanup();
        return FVM_E_HIP;
    }
    if (sa_compat)
        hipLaunchKernelGGL(k_wall_dist<true>, dim3(nblk(np)), dim3(kBlock), 0, st, np, dxy, dse, (int)se.size(), dsol, dwd);
    else
        hipLaunchKernelGGL(k_wall_dist<false>, dim3(nblk(np)), dim3(kBlock), 0, st, np, dxy, dse, (int)se.size(), dsol, dwd);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(wd, dwd, sizeof(double) * np, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        rc = FVM_E_HIP;
    cleanup();
    if (rc) return rc;
    // second layer (src/wall_dist.f:64-73): nearest non-solid neighbour of
    // every solid point, in the solid points' order; then solid points -> 1
    if (wd1) {
        std::vector<double> w1(np, 1.0e10);
        for (int e = 0; e < netot; ++e) {
            int a = edge[2 * e] - 1, b = edge[2 * e + 1] - 1;
            if (solid[a] && !solid[b]) w1[a] = std::fmin(w1[a], wd[b]);
            if (solid[b] && !solid[a]) w1[b] = std::fmin(w1[b], wd[a]);
        }
        int k = 0;
        for (int p = 0; p < np; ++p)
            if (solid[p]) wd1[k++] = w1[p];
    }
    for (int p = 0; p < np; ++p)
        if (solid[p]) wd[p] = 1.0;
    return FVM_OK;
}

// Device address of `scratch` on each device, looked up once per device
// with that device current (the array's name is not a device address in
// host code).
static int *scratch_ptr(int dev)
{
    static std::mutex mu;
    static int *ptr[64] = {nullptr};
    if (dev < 0 || dev >= 64) return nullptr;
    std::lock_guard<std::mutex> lock(mu);
    if (!ptr[dev]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess) {
            (void)hipGetLastError();
            return nullptr;
        }
        ptr[dev] = static_cast<int *>(p);
    }
    return ptr[dev];
}

int fvm_gpu_wall_dist_noalloc(int np, const double *coord, const int *ptype, int netot, const int *edge,
                              int sa_compat, double *wd, void *stream)
{
    if (np <= 0 || netot <= 0 || !coord || !ptype || !edge || !wd) return FVM_E_INPUT;
    int dev = 0;
    HIPCHK(hipGetDevice(&dev));
    int *sc = scratch_ptr(dev);
    if (!sc) return FVM_E_HIP;
    hipStream_t st = static_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(k_scratch_clear, dim3(1), dim3(64), 0, st, sc);
    hipLaunchKernelGGL(k_collect_solid_edges, dim3(nblk(netot)), dim3(kBlock), 0, st, np, netot, edge, ptype, sc);
    if (sa_compat)
        hipLaunchKernelGGL(k_wall_dist_list<true>, dim3(nblk(np)), dim3(kBlock), 0, st, np,
                           reinterpret_cast<const double2 *>(coord), ptype, sc, wd);
    else
        hipLaunchKernelGGL(k_wall_dist_list<false>, dim3(nblk(np)), dim3(kBlock), 0, st, np,
                           reinterpret_cast<const double2 *>(coord), ptype, sc, wd);
    HIPCHK_S(st, hipGetLastError());
    int flags[4] = {0, 0, 0, 0};
    HIPCHK_S(st, hipMemcpyAsync(flags, sc, sizeof flags, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (flags[2]) return FVM_E_INPUT;
    if (flags[1]) {
        std::fprintf(stderr, "fvm_gpu_wall_dist_noalloc: %d solid edges exceed the static scratch of %d\n",
                     flags[0], kScratchEdges);
        return FVM_E_SIZE;
    }
    return FVM_OK;
}

int fvm_gpu_tune(fvm_ctx *x, int variant, int vertices_per_block, int xcd_remap)
{
    if (!ctx_ok(x)) return FVM_E_INPUT;
    if (int rc = settle(x)) return rc;
    return configure(x, variant, vertices_per_block, xcd_remap);
}

int fvm_gpu_set_cl(fvm_ctx *x, double cl)
{
    if (!ctx_ok(x)) return FVM_E_INPUT;
    if (int rc = settle(x)) return rc;
    x->hS.cl = cl;
    const double h = cl;
    HIPCHK_S(x->stream, hipMemcpyAsync(&x->S->cl, &h, sizeof(double), hipMemcpyHostToDevice, x->stream));
    if (x->nff > 0)
        hipLaunchKernelGGL(k_farfield_state, dim3(nblk(x->nff)), dim3(kBlock), 0, x->stream, x->d, x->c, cl);
    HIPCHK_S(x->stream, hipGetLastError());
    HIPCHK(hipStreamSynchronize(x->stream));
    return FVM_OK;
}

int fvm_gpu_get_state(fvm_ctx *x, double *prim, double *nut, double *qx, double *qy, double *mu)
{
    if (!ctx_ok(x)) return FVM_E_INPUT;
    if (int rc = settle(x)) return rc;
    const size_t np = (size_t)x->np;
    const DevState h = no_errors();        // lives until the stream drains below
    if (x->qxy_stale && (qx || qy)) {
        // the gradients of the last residual: those of the last stage's
        // input (buffer C), through the residual kernel, which writes them
        HIPCHK_S(x->stream, hipMemcpyAsync(x->S_hook, &h, sizeof h, hipMemcpyHostToDevice, x->stream));
        launch_residual(x, x->S_hook, x->pbuf[2], x->res_tmp);
        HIPCHK_S(x->stream, hipGetLastError());
        x->qxy_stale = false;
    }
    std::vector<double> a(4 * np), b(4 * np), q(4 * np), r(np), m(np);
    HIPCHK_S(x->stream, hipMemcpyAsync(a.data(), x->d.prim, 32 * np, hipMemcpyDeviceToHost, x->stream));
    HIPCHK_S(x->stream, hipMemcpyAsync(r.data(), x->d.nut, 8 * np, hipMemcpyDeviceToHost, x->stream));
    HIPCHK_S(x->stream, hipMemcpyAsync(b.data(), x->d.qx, 32 * np, hipMemcpyDeviceToHost, x->stream));
    HIPCHK_S(x->stream, hipMemcpyAsync(q.data(), x->d.qy, 32 * np, hipMemcpyDeviceToHost, x->stream));
    HIPCHK_S(x->stream, hipMemcpyAsync(m.data(), x->d.mu, 8 * np, hipMemcpyDeviceToHost, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    if (prim) to_caller(x, a, 4, prim);
    if (nut) to_caller(x, r, 1, nut);
    if (qx) to_caller(x, b, 4, qx);
    if (qy) to_caller(x, q, 4, qy);
    if (mu) to_caller(x, m, 1, mu);
    return FVM_OK;
}

int fvm_gpu_residu(fvm_ctx *x, const double *prim, const double *nut, double *res)
{
    if (!ctx_ok(x) || !prim || !res) return FVM_E_INPUT;
    if (int rc = settle(x)) return rc;
    const int np = x->np;
    // the solver's own state is put aside and restored afterwards
    HIPCHK(hipMemcpyAsync(x->prim_bak, x->d.prim, sizeof(double) * 4 * np, hipMemcpyDeviceToDevice, x->stream));
    HIPCHK(hipMemcpyAsync(x->nut_bak, x->d.nut, sizeof(double) * np, hipMemcpyDeviceToDevice, x->stream));
    std::vector<double> primv = to_device(x, prim, 4), nutv;
    HIPCHK_S(x->stream, hipMemcpyAsync(x->d.prim, primv.data(), sizeof(double) * 4 * np, hipMemcpyHostToDevice, x->stream));
    if (nut && x->par.iflow == 3) {
        nutv = to_device(x, nut, 1);
        HIPCHK_S(x->stream, hipMemcpyAsync(x->d.nut, nutv.data(), sizeof(double) * np, hipMemcpyHostToDevice, x->stream));
    }
    // the hook keeps the iteration's error state apart, but shares cl for
    // the farfield vortex correction
    DevState h = no_errors();
    h.cl = x->hS.cl;
    HIPCHK_S(x->stream, hipMemcpyAsync(x->S_hook, &h, sizeof h, hipMemcpyHostToDevice, x->stream));
    if (x->variant == 3) {
        launch_residual(x, x->S_hook, x->d.prim, x->res_tmp);
        x->qxy_stale = false;                   // the hook's gradients, as in the reference
    } else {
        residual_front(x, x->S_hook);
        x->qxy_stale = false;
        if (x->par.iflow != 1)
            hipLaunchKernelGGL(k_res<true>, dim3(nblk(np)), dim3(kBlock), 0, x->stream, x->d, x->res_tmp);
        else
            hipLaunchKernelGGL(k_res<false>, dim3(nblk(np)), dim3(kBlock), 0, x->stream, x->d, x->res_tmp);
    }
    HIPCHK_S(x->stream, hipGetLastError());
    std::vector<double> resv(4 * (size_t)np);
    HIPCHK_S(x->stream, hipMemcpyAsync(resv.data(), x->res_tmp, sizeof(double) * 4 * np, hipMemcpyDeviceToHost, x->stream));
    HIPCHK_S(x->stream, hipMemcpyAsync(&h, x->S_hook, sizeof h, hipMemcpyDeviceToHost, x->stream));
    HIPCHK_S(x->stream, hipMemcpyAsync(x->d.prim, x->prim_bak, sizeof(double) * 4 * np, hipMemcpyDeviceToDevice, x->stream));
    HIPCHK_S(x->stream, hipMemcpyAsync(x->d.nut, x->nut_bak, sizeof(double) * np, hipMemcpyDeviceToDevice, x->stream));
    HIPCHK(hipStreamSynchronize(x->stream));
    to_caller(x, resv, 4, res);
    if (h.err_code) return h.err_code;
    return FVM_OK;
}

int fvm_gpu_iterate(fvm_ctx *x, int n, double *hist, int *ndone, int *err_index)
{
    if (ndone) *ndone = 0;
    if (err_index) *err_index = 0;
    if (!ctx_ok(x) || n < 0) return FVM_E_INPUT;
    if (x->batch) return FVM_E_INPUT;           // a member advances with its batch
    int done_total = 0;
    while (done_total < n && !x->hS.done) {
        int chunk = std::min(n - done_total, kHistCap);
        hipLaunchKernelGGL(k_begin_chunk, dim3(1), dim3(1), 0, x->stream, x->S);
        x->qxy_stale = x->variant == 3 && x->par.iflow == 1;
        if (x->graphs_ok && !x->graph_n) {
            if (build_graph(x, kGraphIters, &x->graph_n) || build_graph(x, 1, &x->graph_1)) {
                drop_graphs(x);                 // none or both
                x->graphs_ok = false;
            }
        }
        if (x->graphs_ok) {
            int q = chunk / kGraphIters, r = chunk % kGraphIters;
            for (int i = 0; i < q; ++i) HIPCHK_S(x->stream, hipGraphLaunch(x->graph_n, x->stream));
            for (int i = 0; i < r; ++i) HIPCHK_S(x->stream, hipGraphLaunch(x->graph_1, x->stream));
        } else {
            for (int i = 0; i < chunk; ++i) one_iteration(x, nullptr);
        }
        HIPCHK_S(x->stream, hipGetLastError());
        HIPCHK_S(x->stream, hipMemcpyAsync(&x->hS, x->S, sizeof(DevState), hipMemcpyDeviceToHost, x->stream));
        HIPCHK(hipStreamSynchronize(x->stream));
        const int got = std::min(x->hS.nhist, chunk);
        if (hist && got > 0) {
            HIPCHK_S(x->stream, hipMemcpyAsync(hist + 6 * (size_t)done_total, x->d.hist, sizeof(double) * 6 * got,
                                               hipMemcpyDeviceToHost, x->stream));
            HIPCHK(hipStreamSynchronize(x->stream));
        }
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
    if (!ctx_ok(x) || x->batch) return FVM_E_INPUT;
    Tally t;
    for (int i = 0; i < n && !x->hS.done; ++i) {
        hipLaunchKernelGGL(k_begin_chunk, dim3(1), dim3(1), 0, x->stream, x->S);
        x->qxy_stale = x->variant == 3 && x->par.iflow == 1;
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

int fvm_gpu_stream_copy(long long bytes, int reps, double *gbs)
{
    *gbs = 0.0;
    if (bytes < 16 || reps < 1) return FVM_E_INPUT;
    const long long n = bytes / 16;
    double2 *a = nullptr, *b = nullptr;
    if (hipMalloc((void **)&a, n * 16) != hipSuccess) return FVM_E_NOMEM;
    if (hipMalloc((void **)&b, n * 16) != hipSuccess) { (void)hipFree(a); return FVM_E_NOMEM; }
    int rc = FVM_OK;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    // four workgroups of 256 threads per CU, each striding over the buffer:
    // the fastest of the shapes measured on MI355X (1-8 items in flight per
    // thread, 4-64 workgroups per CU)
    int dev = 0, cus = 256;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
        cus = 256;
    const int grid = 4 * cus;
    float t = 0.f;
    if (hipMemset(a, 0, n * 16) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
        hipEventCreate(&e1) != hipSuccess) {
        rc = FVM_E_HIP;
    } else {
        hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(kBlock), 0, nullptr, a, b, n);   // warm
        (void)hipEventRecord(e0, nullptr);
        for (int r = 0; r < reps; ++r)
            hipLaunchKernelGGL(k_stream_copy, dim3(grid), dim3(kBlock), 0, nullptr, a, b, n);
        (void)hipEventRecord(e1, nullptr);
        if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&t, e0, e1) != hipSuccess ||
            hipGetLastError() != hipSuccess)
            rc = FVM_E_HIP;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    (void)hipFree(a);
    (void)hipFree(b);
    if (rc == FVM_OK && t > 0.f) *gbs = 2.0 * (double)(n * 16) * reps / ((double)t * 1e-3) / 1e9;
    return rc;
}

int fvm_gpu_time_stage(fvm_ctx *x, int n, double *ms)
{
    *ms = 0.0;
    if (!ctx_ok(x) || x->batch || x->variant != 3 || n <= 0) return FVM_E_INPUT;
    if (x->hS.done) return x->hS.status;
    // stage 1 reads the iteration-start buffer and writes buffer B only, so
    // repeating it changes nothing the next iteration reads
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    HIPCHK(hipEventCreate(&e1));
    launch_stage(self_stage(x), 1, 0);                  // warm
    HIPCHK(hipEventRecord(e0, x->stream));
    for (int i = 0; i < n; ++i) launch_stage(self_stage(x), 1, 0);
    HIPCHK(hipEventRecord(e1, x->stream));
    HIPCHK(hipEventSynchronize(e1));
    float t = 0.f;
    HIPCHK(hipEventElapsedTime(&t, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *ms = (double)t / n;
    return FVM_OK;
}

}  // extern "C"


// ----------------------------------------------------------------- batches

namespace {

// (the batch's stream has drained: callers wait or synchronise first)
void batch_free(fvm_batch *b)
{
    if (b->graph_n) (void)hipGraphExecDestroy(b->graph_n);
    if (b->graph_1) (void)hipGraphExecDestroy(b->graph_1);
    b->graph_n = b->graph_1 = nullptr;
    for (void *p : {(void *)b->d_mem, (void *)b->d_work, (void *)b->d_vwork, (void *)b->d_work_mon,
                    (void *)b->d_states})
        if (p) (void)hipFreeAsync(p, b->stream);
    b->d_mem = nullptr; b->d_work = b->d_vwork = b->d_work_mon = nullptr; b->d_states = nullptr;
    b->nwork = b->nvwork = 0;
}

template <class T>
int batch_upload(fvm_batch *b, T **p, const std::vector<T> &h)
{
    if (hipMallocAsync((void **)p, std::max<size_t>(h.size(), 1) * sizeof(T), b->stream) != hipSuccess) {
        (void)hipGetLastError();
        *p = nullptr;
        return FVM_E_NOMEM;
    }
    if (!h.empty()) {
        HIPCHK_S(b->stream, hipMemcpyAsync(*p, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, b->stream));
        HIPCHK(hipStreamSynchronize(b->stream));
    }
    return FVM_OK;
}

// descriptors and work lists of the current members
int batch_rebuild(fvm_batch *b)
{
    HIPCHK(hipStreamSynchronize(b->stream));
    batch_free(b);
    std::vector<BatchMember> m;
    std::vector<int2> work, vwork;
    b->lds = 0; b->sa_lds = 0; b->max_nff = 0; b->any_vortex = false;
    std::vector<int> cost;
    for (int i = 0; i < (int)b->mem.size(); ++i) {
        const fvm_ctx *x = b->mem[i];
        m.push_back(member_of(x));
        for (int k = 0; k < x->pa.npatch; ++k) {
            work.push_back(make_int2(i, k));
            cost.push_back(x->patch_cost[k]);
        }
        for (int k = 0; k < (int)nblk(x->np); ++k) vwork.push_back(make_int2(i, k));
        b->lds = std::max(b->lds, x->pa.lds_bytes);
        b->sa_lds = std::max(b->sa_lds, x->pa.sa_lds_bytes);
        b->max_nff = std::max(b->max_nff, x->nff);
        if (x->c.vortex == 1 && x->c.mach_inf < 1.0) b->any_vortex = true;
    }
    std::vector<int2> work_mon;
    if (!b->mem.empty()) {
        const int cus = b->mem[0]->cus, pcu = stage_per_cu(b->mem[0], b->lds);
        work_mon = with_monitors(work, cost, (int)b->mem.size(), cus, pcu);
        order_work(work, cost, cus, pcu);
    }
    int rc = batch_upload(b, &b->d_mem, m);
    if (!rc) rc = batch_upload(b, &b->d_work, work);
    if (!rc) rc = batch_upload(b, &b->d_work_mon, work_mon);
    if (!rc) rc = batch_upload(b, &b->d_vwork, vwork);
    if (!rc && hipMallocAsync((void **)&b->d_states, std::max<size_t>(m.size(), 1) * sizeof(DevState),
                              b->stream) != hipSuccess) {
        (void)hipGetLastError();
        rc = FVM_E_NOMEM;
    }
    if (!rc && b->hs_cap < (int)m.size()) {
        // grows rarely (hipHostFree waits for the whole device)
        if (b->hs) (void)hipHostFree(b->hs);
        b->hs = nullptr;
        b->hs_cap = 0;
        const int cap = std::max<int>(1024, (int)m.size() * 2);
        if (hipHostMalloc((void **)&b->hs, sizeof(DevState) * cap, hipHostMallocDefault) != hipSuccess) {
            b->hs = nullptr;
            rc = FVM_E_NOMEM;
        } else {
            b->hs_cap = cap;
        }
    }
    if (rc) { batch_free(b); return rc; }
    b->nwork = (int)work.size();
    b->nvwork = (int)vwork.size();
    b->graphs_ok = true;
    b->dirty = false;
    return FVM_OK;
}

// one iteration of every member; mon_prev and final as in one_iteration
void batch_iteration(fvm_batch *b, bool mon_prev = false, bool final = true)
{
    const fvm_params &q = b->mem[0]->par;
    const int nm = (int)b->mem.size();
    const bool turb = q.iflow == 3;
    if (turb && mon_prev)      // the previous iteration's nu_tilde update folded in
        hipLaunchKernelGGL((k_tstep_batch<true, true>), dim3(b->nvwork), dim3(kBlock), 0, b->stream, b->d_mem,
                           b->d_vwork);
    else if (turb)
        hipLaunchKernelGGL((k_tstep_batch<true, false>), dim3(b->nvwork), dim3(kBlock), 0, b->stream, b->d_mem,
                           b->d_vwork);
    StageLaunch L{b->d_mem, b->d_work, b->nwork, 1, b->lds, b->stream,
                  q.iflow, q.iflux, q.hllc_rotate, q.ilimit, q.precision, wide_sa(b->mem[0], b->lds),
                  b->mem[0]->threads};
    StageLaunch LM = L;
    LM.work = b->d_work_mon;
    LM.nwork = b->nwork + nm;
    for (int irk = 1; irk <= 3; ++irk) launch_stage(mon_prev && irk == 1 ? LM : L, irk, irk == 3 ? 1 : 0);
    if (turb) {
        if (q.sa_compat)
            hipLaunchKernelGGL(k_sa_patch_batch<true>, dim3(b->nwork), dim3(kBlock), b->sa_lds, b->stream,
                               b->d_mem, b->d_work, b->nwork, 1);
        else
            hipLaunchKernelGGL(k_sa_patch_batch<false>, dim3(b->nwork), dim3(kBlock), b->sa_lds, b->stream,
                               b->d_mem, b->d_work, b->nwork, 1);
        if (final)             // otherwise the next iteration's time step copies
            hipLaunchKernelGGL(k_copy_batch, dim3(b->nvwork), dim3(kBlock), 0, b->stream, b->d_mem, b->d_vwork);
    }
    if (!final) return;
    if (q.iflow == 1)
        hipLaunchKernelGGL(k_final_batch<false>, dim3(nm), dim3(kFinal), 0, b->stream, b->d_mem);
    else
        hipLaunchKernelGGL(k_final_batch<true>, dim3(nm), dim3(kFinal), 0, b->stream, b->d_mem);
    if (b->any_vortex && b->max_nff > 0)
        hipLaunchKernelGGL(k_vortex_batch, dim3(nblk(b->max_nff), nm), dim3(kBlock), 0, b->stream, b->d_mem);
}

int batch_graph(fvm_batch *b, int iters, hipGraphExec_t *out)
{
    *out = nullptr;
    if (hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal) != hipSuccess) {
        (void)hipGetLastError();
        return FVM_E_HIP;
    }
    const bool merge = monitors_merge(b->mem[0]->par, b->any_vortex && b->max_nff > 0);
    for (int i = 0; i < iters; ++i) batch_iteration(b, merge && i > 0, !merge || i == iters - 1);
    return end_capture(b->stream, out);
}

// members share the kernel instances: the same kind of solve and workgroup
bool same_kind(const fvm_ctx *x, const fvm_ctx *y)
{
    const fvm_params &a = x->par, &b = y->par;
    return a.iflow == b.iflow && a.iflux == b.iflux && a.ilimit == b.ilimit && a.precision == b.precision &&
           a.hllc_rotate == b.hllc_rotate && a.sa_compat == b.sa_compat && x->threads == y->threads;
}

int member_status(const DevState &h)
{
    if (h.err_code) return h.err_code;
    return h.done ? h.status : FVM_OK;
}

}  // namespace

int fvm_batch_create(fvm_batch **out)
{
    if (!out) return FVM_E_INPUT;
    *out = nullptr;
    fvm_batch *b = new fvm_batch();
    if (hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess) { delete b; return FVM_E_HIP; }
    keep_pool_memory();
    live_add(b);
    *out = b;
    return FVM_OK;
}

namespace {

// waits for the launched chunk and hands each member its monitors
int batch_wait(fvm_batch *b)
{
    if (!b->pending) return FVM_OK;
    b->pending = false;
    HIPCHK(hipStreamSynchronize(b->stream));
    for (size_t i = 0; i < b->mem.size(); ++i) b->mem[i]->hS = b->hs[i];
    return FVM_OK;
}

}  // namespace

// a batch member's own calls first wait for its batch's launched chunk
int settle(fvm_ctx *x)
{
    return x->batch ? batch_wait(x->batch) : FVM_OK;
}

namespace {

}  // namespace

int fvm_batch_add(fvm_batch *b, fvm_ctx *x)
{
    if (!batch_ok(b) || !ctx_ok(x) || x->batch || x->variant != 3) return FVM_E_INPUT;
    if (!b->mem.empty() && !same_kind(b->mem[0], x)) return FVM_E_INPUT;
    int rc = batch_wait(b);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(x->stream));
    b->mem.push_back(x);
    x->batch = b;
    b->dirty = true;
    return FVM_OK;
}

int fvm_batch_remove(fvm_batch *b, fvm_ctx *x)
{
    if (!batch_ok(b) || !ctx_ok(x) || x->batch != b) return FVM_E_INPUT;
    int rc = batch_wait(b);
    if (rc) return rc;
    b->mem.erase(std::find(b->mem.begin(), b->mem.end(), x));
    x->batch = nullptr;
    b->dirty = true;
    return FVM_OK;
}

int fvm_batch_size(const fvm_batch *b) { return batch_ok(b) ? (int)b->mem.size() : 0; }

int fvm_batch_launch(fvm_batch *b, int n)
{
    if (!batch_ok(b)) return FVM_E_INPUT;
    const int nm = (int)b->mem.size();
    if (n < 0 || n > kHistCap || b->pending) return FVM_E_INPUT;
    if (nm == 0 || n == 0) return FVM_OK;
    if (b->dirty) {
        int rc = batch_rebuild(b);
        if (rc) return rc;
    }
    bool running = false;
    for (const fvm_ctx *x : b->mem) running = running || !x->hS.done;
    if (!running) return FVM_OK;
    for (fvm_ctx *x : b->mem) x->qxy_stale = x->par.iflow == 1;      // stage kernels only
    hipLaunchKernelGGL(k_begin_batch, dim3(nblk(nm)), dim3(kBlock), 0, b->stream, b->d_mem, nm);
    if (b->graphs_ok && !b->graph_n) {
        if (batch_graph(b, kGraphIters, &b->graph_n) || batch_graph(b, 1, &b->graph_1)) {
            if (b->graph_n) (void)hipGraphExecDestroy(b->graph_n);
            if (b->graph_1) (void)hipGraphExecDestroy(b->graph_1);
            b->graph_n = b->graph_1 = nullptr;          // none or both
            b->graphs_ok = false;
        }
    }
    if (b->graphs_ok) {
        const int q = n / kGraphIters, r = n % kGraphIters;
        for (int i = 0; i < q; ++i) HIPCHK_S(b->stream, hipGraphLaunch(b->graph_n, b->stream));
        for (int i = 0; i < r; ++i) HIPCHK_S(b->stream, hipGraphLaunch(b->graph_1, b->stream));
    } else {
        for (int i = 0; i < n; ++i) batch_iteration(b);
    }
    hipLaunchKernelGGL(k_gather_batch, dim3(nblk(nm)), dim3(kBlock), 0, b->stream, b->d_mem, nm, b->d_states);
    HIPCHK_S(b->stream, hipGetLastError());
    HIPCHK_S(b->stream, hipMemcpyAsync(b->hs, b->d_states, sizeof(DevState) * nm, hipMemcpyDeviceToHost, b->stream));
    b->pending = true;
    return FVM_OK;
}

int fvm_batch_query(fvm_batch *b)
{
    if (!batch_ok(b)) return FVM_E_INPUT;
    if (!b->pending) return 1;
    const hipError_t e = hipStreamQuery(b->stream);
    if (e == hipSuccess) return 1;
    if (e == hipErrorNotReady) return 0;
    return FVM_E_HIP;
}

int fvm_batch_wait(fvm_batch *b, int *status)
{
    if (!batch_ok(b)) return FVM_E_INPUT;
    int rc = batch_wait(b);
    if (rc) return rc;
    if (status)
        for (size_t i = 0; i < b->mem.size(); ++i) status[i] = member_status(b->mem[i]->hS);
    return FVM_OK;
}

int fvm_batch_iterate(fvm_batch *b, int n, int *status)
{
    if (!batch_ok(b)) return FVM_E_INPUT;
    if (b->mem.empty()) return n > 0 ? FVM_E_INPUT : FVM_OK;
    int rc = batch_wait(b);
    for (int left = n; !rc && left > 0; left -= kHistCap) {
        bool running = false;
        for (const fvm_ctx *x : b->mem) running = running || !x->hS.done;
        if (!running) break;
        rc = fvm_batch_launch(b, std::min(left, kHistCap));
        if (!rc) rc = batch_wait(b);
    }
    if (!rc && status)
        for (size_t i = 0; i < b->mem.size(); ++i) status[i] = member_status(b->mem[i]->hS);
    return rc;
}

int fvm_batch_destroy(fvm_batch *b)
{
    if (!b) return FVM_OK;
    if (!batch_ok(b)) return FVM_E_INPUT;
    (void)batch_wait(b);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (fvm_ctx *x : b->mem) x->batch = nullptr;
    b->mem.clear();
    batch_free(b);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    if (b->hs) (void)hipHostFree(b->hs);
    b->hs = nullptr;
    if (b->stream) (void)hipStreamDestroy(b->stream);
    live_drop(b);
    b->tag = 0;
    delete b;
    return FVM_OK;
}

int fvm_gpu_monitor(const fvm_ctx *x, double *mon)
{
    if (!ctx_ok(x) || !mon) return FVM_E_INPUT;
    const DevState &h = x->hS;
    mon[0] = (double)h.iter;
    mon[1] = h.fres;
    mon[2] = (double)h.iresi;
    mon[3] = h.fresi;
    mon[4] = h.cl;
    mon[5] = h.cd;
    return member_status(h);
}
