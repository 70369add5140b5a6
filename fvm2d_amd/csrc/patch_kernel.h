// patch_kernel.h -- one RK stage of the residual + update in a single
// kernel, tiled over vertex patches through LDS.
//
// A workgroup owns a patch P of consecutive vertices.  It loads the stage
// state of P, its 1-ring H1 and its 2-ring H2 into LDS, then
//   1. (with the gather) copies the local triangles into LDS;
//   2. sums, for each vertex of G = P + H1, its triangles' Green-Gauss
//      terms in ascending triangle order -> gradients of G
//      (src/green_gauss.f); with viscosity each vertex of P also sums the
//      triangles' Galerkin stress terms; every term is evaluated in the sum
//      from LDS (same operands, same order: the same bits as terms
//      evaluated once per triangle);
//   3. evaluates every edge touching P once (MUSCL + Riemann flux) and every
//      solid/farfield edge touching P (src/residu.f:37-88);
//   4. sums, for each vertex of P, its edge and boundary terms in the
//      reference's order and applies the RK stage (src/update.f).
// The result goes to a different state buffer than the one read (stage
// buffers rotate O -> B -> C -> O), so no workgroup reads a vertex another
// one is writing.  Halo work (H1 gradients, cut edges) is recomputed by the
// neighbouring patches; nothing crosses workgroups within a stage.
#pragma once

namespace fvm {

// per-patch header
constexpr int kPatchHdr = 24;
enum {
    PH_V0 = 0, PH_NP, PH_NL, PH_NG, PH_HALO, PH_NTRI, PH_TRI, PH_GPTR, PH_GTL,
    PH_NEDGE, PH_EDGE, PH_NBND, PH_BND, PH_PPTR, PH_PSL, PH_NGTL, PH_NPSL,
    PH_NTP,                 // local triangles touching P (numbered first)
    PH_BLOB, PH_NBLOB       // the patch's index tables as one 16-byte-aligned block
};

struct PatchArgs {
    const int *hdr;         // kPatchHdr ints per patch
    const int *halo;        // new vertex ids of H1 then H2
    const short4 *ltri;     // local nodes of each local triangle
    const int *gtri;        // its triangle number
    const double *lqinv;    // 0.25 / its area (elarea)
    const int *gptr;        // per G-vertex: range in gtl (patch-relative)
    const short *gtl;       // 4 * local triangle + position of the vertex in it
    const short2 *ledge;    // local end points of each local edge
    const int *ledge_id;    // its position in the live-edge loop
    const double2 *ledge_ds;  // its dual-face normal (ds)
    const double4 *ledge_fr;  // the frame of ds: ct, st, |ds| (and a pad)
    const short2 *lbnd;     // local end points of each local boundary edge
    const int *lbnd_id;     // its index among the boundary edges (solid first)
    const int *pptr;        // per P-vertex: range in pslot (patch-relative)
    const int4 *blob;       // per patch: gptr, pptr, gtl, pslot back to back (the
                            // LDS image of the index tables), padded to 16 bytes
    const short *pslot;     // 2 * (term index) + (1 if subtracted)
    int npatch;
    int lds_bytes;
    int sa_lds_bytes;       // of the SA-step kernel over the same patches
};

// bytes of the LDS carve-up of one patch; the kernel and the planner both use it
// flow: 1 Euler, 2 laminar (viscosity and p/rho per local vertex), 3 SA
// (also the laminar viscosity and nu_tilde per local vertex)
__host__ __device__ inline int patch_lds_bytes(int NP, int NL, int NG, int NT, int NE, int NB,
                                               int NGTL, int NPSL, int NTP, int flow, bool split)
{
    const bool visc = flow != 1;
    const int NLa = (NL + 1) & ~1;
    const int work = (NT * (visc ? 2 : 1) * 8) > ((NE + 2 * NB) * 32) ? NT * (visc ? 2 : 1) * 8
                                                                      : (NE + 2 * NB) * 32;
    const int index = 4 * (NG + 1) + 4 * (NP + 1) + 2 * (NGTL + NPSL);   // the blob
    // ais of G, gathered with the state
    const int pre = ((NG + 1) & ~1) * 8;
    // gradients: FP64, or FP32 in the split mode (its reconstruction reads
    // FP32); with viscosity the same bytes first stage 5 doubles of stress
    // terms per triangle touching P
    const int grad = NG * (split ? 32 : 64);
    const int grad_region = (((visc && NTP * 40 > grad) ? NTP * 40 : grad) + 15) & ~15;
    // per local vertex: state and coordinates; with viscosity 2 (SA: 4)
    // doubles; inviscid split precision an FP32 state and coordinates (24 B)
    // with viscosity the viscous residual of each P vertex (3 doubles),
    // after the blob
    return NLa * 48 + NLa * 8 * (flow == 3 ? 4 : (visc ? 2 : (split ? 3 : 0))) + grad_region + pre +
           ((work + 15) & ~15) + ((index + 15) & ~15) + (visc ? NP * 24 : 0);
}

// bytes of the LDS carve-up of the SA step over one patch: state,
// coordinates, nu_tilde and laminar viscosity of G; the local triangles with
// their areas; the edge and boundary-edge tables; the per-vertex lists
__host__ __device__ inline int sa_lds_bytes(int NP, int NG, int NT, int NE, int NB, int NGTL, int NPSL)
{
    const int NGa = (NG + 1) & ~1, NTa = (NT + 1) & ~1;
    return NG * 48 + NGa * 16 + NTa * 8 + NE * 16 + NT * 8 + NE * 4 + NB * 8 + 8 * (NP + 1) +
           2 * (NGTL + NPSL);
}

}  // namespace fvm
