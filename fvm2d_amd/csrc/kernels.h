// kernels.h -- point functions of the hot path, shared by every kernel.
//
// Each one evaluates its formula in the reference's order of operations, so
// that, compiled with -ffp-contract=off, the FP64 path reproduces the CPU
// solver bit for bit wherever no libm transcendental is involved.
#pragma once
#include <hip/hip_runtime.h>

namespace fvm {

constexpr double kEps = 1.0e-16;             // EPSILON, src/param.h:105

// COMMON-block values the kernels read (src/param.h:56-142, src/common.h)
struct Consts {
    double GAMMA, GAMMA1, M_PI_, prandtl, prandtl_turb;
    double ALBADA11, ALBADA12, ALBADA21, ALBADA22;
    double Cb1, Cw1, Cw2, Cv2, Cv11, Cw31, Cw32, kolm2, Cb2Sig1, Cb2Sig2;
    double q_inf, r_inf, p_inf, T_inf, aoa, u_inf, v_inf, a_inf, SCONST, nut_inf;
    double rey, cfl, mach_inf, xref, yref, minres;
    double sin_aoa, cos_aoa;              // host libm, as the reference's dsin/dcos
    double airk[3], birk[3];
    int iflow, iflux, ilimit, vortex, sa_compat, hllc_rotate;
    int iterlast, maxiter, saveinterval, np;
};

__device__ __forceinline__ double sq(double x) { return x * x; }
// integer powers as binary exponentiation: x**3 = x*(x*x), x**6 = x^2*x^4
__device__ __forceinline__ double pow3(double x) { double x2 = x * x; return x * x2; }
__device__ __forceinline__ double pow6(double x) { double x2 = x * x; double x4 = x2 * x2; return x2 * x4; }

template <class T> struct VT { T v[4]; };
using V4 = VT<double>;
using V4f = VT<float>;
__device__ __forceinline__ V4 ld4(const double4 *p, int i)
{
    double4 a = p[i];
    return V4{{a.x, a.y, a.z, a.w}};
}
__device__ __forceinline__ void st4(double4 *p, int i, const V4 &a)
{
    p[i] = make_double4(a.v[0], a.v[1], a.v[2], a.v[3]);
}
__device__ __forceinline__ V4f to_f(const V4 &a)
{
    return V4f{{(float)a.v[0], (float)a.v[1], (float)a.v[2], (float)a.v[3]}};
}
__device__ __forceinline__ V4 to_d(const V4f &a)
{
    return V4{{(double)a.v[0], (double)a.v[1], (double)a.v[2], (double)a.v[3]}};
}

// The reconstruction and the Riemann fluxes are templates on the arithmetic
// type: T = double is the reference's arithmetic, T = float the split-
// precision mode (geometry, accumulation, update and norms stay FP64).

// Division, square root, exp and erf of the reconstruction and fluxes.  In
// FP64 they are the IEEE operations (and libm), as in the reference.  In the
// split-precision mode (T = float) they are the hardware approximations
// (v_rcp_f32 times the numerator, v_sqrt_f32, fast exp): about 1 ulp of
// FP32, well inside what the FP32 arithmetic around them already loses, at
// a fraction of the instructions of correctly rounded FP32 division.
__device__ __forceinline__ double fdiv(double a, double b) { return a / b; }
__device__ __forceinline__ float fdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double fsqrt(double a) { return sqrt(a); }
__device__ __forceinline__ float fsqrt(float a) { return __builtin_amdgcn_sqrtf(a); }
__device__ __forceinline__ double fexp(double a) { return exp(a); }
__device__ __forceinline__ float fexp(float a) { return __expf(a); }
__device__ __forceinline__ double ferf(double a) { return erf(a); }
__device__ __forceinline__ float ferf(float a) { return erff(a); }

// The frame of a dual face ds = (dsx, dsy): its length and unit normal
// (ct, st), the first lines of src/roe_flux.f:18-20, src/kfvs_flux.f:16-18
// and src/hllc_flux.f:18-20.  It depends on the mesh only, so the stage
// kernel reads it from a per-edge table made once (the same IEEE square root
// and divisions, so the same bits).
template <class T> struct EdgeFrame { T ct, st, len; };

template <class T>
__device__ __forceinline__ EdgeFrame<T> edge_frame(T dsx, T dsy)
{
    const T length = fsqrt(dsx * dsx + dsy * dsy);
    return EdgeFrame<T>{fdiv(dsx, length), fdiv(dsy, length), length};
}

// van Albada limiter, src/recon.f:55-71.  SEL: the quotient is formed
// whether or not the slopes agree in sign and then selected (the same value:
// bot is at least 2 eps), without a branch around each of the eight
// divisions of an edge, so the compiler can interleave their dependency
// chains (inviscid FP64 Roe stage: 2% faster, 1.3% per iteration; the FP32
// quotient is cheap, the KFVS instance would spill and the SA-RANS one ran
// 1% slower, so they keep the branch).
template <class T, bool SEL = false>
__device__ __forceinline__ T limit_albada(T ul, T ur)
{
    if (!SEL && ul * ur <= T(0)) return T(0);
    T top = (ul * ul + T(kEps)) * ur + (ur * ur + T(kEps)) * ul;
    T bot = ul * ul + ur * ur + T(2.0 * kEps);
    const T q = fdiv(top, bot);
    return (SEL && ul * ur <= T(0)) ? T(0) : q;
}

// MUSCL states at the edge midpoint, src/recon.f:15-44
template <int LIM, class T = double, bool SEL = false>
__device__ __forceinline__ void recon(const Consts &c, T dx, T dy, const VT<T> &p1,
                                      const VT<T> &p2, const VT<T> &qx1, const VT<T> &qy1,
                                      const VT<T> &qx2, const VT<T> &qy2, VT<T> &pl, VT<T> &pr)
{
    const T A11 = T(c.ALBADA11), A12 = T(c.ALBADA12), A21 = T(c.ALBADA21), A22 = T(c.ALBADA22);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        T dqp = p2.v[i] - p1.v[i];
        if (LIM) {
            T dql = A21 * (dx * qx1.v[i] + dy * qy1.v[i]) + A22 * dqp;
            pl.v[i] = p1.v[i] + T(0.5) * limit_albada<T, SEL>(dql, dqp);
            T dqr = A21 * (dx * qx2.v[i] + dy * qy2.v[i]) + A22 * dqp;
            pr.v[i] = p2.v[i] - T(0.5) * limit_albada<T, SEL>(dqr, dqp);
        } else {
            T dql = dx * qx1.v[i] + dy * qy1.v[i];
            pl.v[i] = p1.v[i] + T(0.5) * (A11 * dql + A12 * dqp);
            T dqr = dx * qx2.v[i] + dy * qy2.v[i];
            pr.v[i] = p2.v[i] - T(0.5) * (A11 * dqr + A12 * dqp);
        }
    }
}

// Split precision, MUSCL increments (src/recon.f:15-44 in increment form):
// from the jump dq = p2 - p1 (taken in FP64, rounded) and FP32 gradients,
// the increments dl, dr with pl = p1 + dl and pr = p2 - dr, which the caller
// forms in FP64.  Increments are small, so FP32 keeps their relative
// accuracy where FP32 states (a pressure near p_inf = 1/(gamma M^2), about
// 32 at M 0.15) would not.
template <int LIM>
__device__ __forceinline__ void recon_incr(const Consts &c, float dx, float dy, const V4f &dq,
                                           const V4f &qx1, const V4f &qy1, const V4f &qx2,
                                           const V4f &qy2, V4f &dl, V4f &dr)
{
    const float A11 = float(c.ALBADA11), A12 = float(c.ALBADA12), A21 = float(c.ALBADA21),
                A22 = float(c.ALBADA22);
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        const float dqp = dq.v[i];
        if (LIM) {
            dl.v[i] = 0.5f * limit_albada<float>(A21 * (dx * qx1.v[i] + dy * qy1.v[i]) + A22 * dqp, dqp);
            dr.v[i] = 0.5f * limit_albada<float>(A21 * (dx * qx2.v[i] + dy * qy2.v[i]) + A22 * dqp, dqp);
        } else {
            dl.v[i] = 0.5f * (A11 * (dx * qx1.v[i] + dy * qy1.v[i]) + A12 * dqp);
            dr.v[i] = 0.5f * (A11 * (dx * qx2.v[i] + dy * qy2.v[i]) + A12 * dqp);
        }
    }
}

// Split precision, Roe flux (src/roe_flux.f:18-112 regrouped): the
// consistent part Fc(L), whose momentum and energy carry the free-stream
// pressure, in FP64 from the FP64 states; the jumps of the conserved
// variables in FP64; the Roe averages, eigenvalues, wave strengths and the
// upwind dissipation they make, all small or scaled by small jumps, in FP32.
// false: a^2 <= 0 (src/roe_flux.f:60-65).
__device__ __forceinline__ bool roe_flux_split(const Consts &c, const EdgeFrame<double> &fr, const V4 &L,
                                               const V4 &R, V4 &f)
{
    const double G = c.GAMMA, G1 = c.GAMMA1;
    const double ct = fr.ct, st = fr.st;
    const double rl = L.v[0], ul = L.v[1], vl = L.v[2], pl = L.v[3];
    const double rr = R.v[0], ur = R.v[1], vr = R.v[2], pr = R.v[3];
    const double unl = ul * ct + vl * st, ql2 = ul * ul + vl * vl, qr2 = ur * ur + vr * vr;
    // rho h = gamma/(gamma-1) p + rho q^2 / 2
    const double Fc0 = rl * unl, Fc1 = pl * ct + rl * ul * unl, Fc2 = pl * st + rl * vl * unl;
    const double Fc3 = (G / G1 * pl + 0.5 * rl * ql2) * unl;
    const float du1 = (float)(rr - rl), du2 = (float)(rr * ur - rl * ul), du3 = (float)(rr * vr - rl * vl);
    const float du4 = (float)((pr - pl) / G1 + 0.5 * (rr * qr2 - rl * ql2));
    const float Gf = (float)G, G1f = (float)G1, h = 0.5f, ctf = (float)ct, stf = (float)st;
    const float rlf = (float)rl, ulf = (float)ul, vlf = (float)vl, plf = (float)pl;
    const float rrf = (float)rr, urf = (float)ur, vrf = (float)vr, prf = (float)pr;
    const float hl = fdiv(fdiv(Gf * plf, rlf), G1f) + h * (float)ql2;
    const float hr = fdiv(fdiv(Gf * prf, rrf), G1f) + h * (float)qr2;
    const float rl12 = fsqrt(rlf), rr12 = fsqrt(rrf), rd = fdiv(1.0f, rl12 + rr12);
    const float ua = (ulf * rl12 + urf * rr12) * rd;
    const float va = (vlf * rl12 + vrf * rr12) * rd;
    const float ha = (hl * rl12 + hr * rr12) * rd;
    const float qa2 = ua * ua + va * va;
    const float aa2 = G1f * (ha - h * qa2);
    if (!(aa2 > 0.0f)) return false;
    const float aa = fsqrt(aa2);
    const float una = ua * ctf + va * stf;
    const float vna = -ua * stf + va * ctf;
    const float l1 = fminf(una - aa, 0.0f), l2 = fminf(una, 0.0f), l4 = fminf(una + aa, 0.0f);
    const float m1 = fdiv(ctf * du2 + stf * du3 - una * du1, aa);
    const float m2 = fdiv(G1f * (du4 - ua * du2 - va * du3 + qa2 * du1), aa * aa);
    const float a4 = h * (m1 + m2), a1 = h * (m2 - m1), a3 = du1 - a1 - a4;
    const float a2 = fdiv(stf * du2 - ctf * du3 + vna * du1, aa);
    const float a1l1 = a1 * l1, a2l2 = a2 * l2, a3l3 = a3 * l2, a4l4 = a4 * l4;
    const float aact = aa * ctf, aast = aa * stf;
    const float Fd0 = a1l1 + a3l3 + a4l4;
    const float Fd1 = a1l1 * (ua - aact) + a2l2 * aa * stf + a3l3 * ua + a4l4 * (ua + aact);
    const float Fd2 = a1l1 * (va - aast) - a2l2 * aa * ctf + a3l3 * va + a4l4 * (va + aast);
    const float Fd3 = a1l1 * (ha - una * aa) + a2l2 * aa * vna + a3l3 * h * qa2 + a4l4 * (ha + una * aa);
    const double length = fr.len;
    f.v[0] = (Fc0 + (double)Fd0) * length;
    f.v[1] = (Fc1 + (double)Fd1) * length;
    f.v[2] = (Fc2 + (double)Fd2) * length;
    f.v[3] = (Fc3 + (double)Fd3) * length;
    return true;
}

// Roe flux without entropy fix, src/roe_flux.f:18-112.  false: a^2 <= 0.
template <class T = double>
__device__ __forceinline__ bool roe_flux(const Consts &c, const EdgeFrame<T> &fr, const VT<T> &L,
                                         const VT<T> &R, VT<T> &f)
{
    const T G = T(c.GAMMA), G1 = T(c.GAMMA1), h = T(0.5);
    const T length = fr.len, ct = fr.ct, st = fr.st;
    T rl = L.v[0], ul = L.v[1], vl = L.v[2], pl = L.v[3];
    T al2 = fdiv(G * pl, rl), ql2 = ul * ul + vl * vl, hl = fdiv(al2, G1) + h * ql2;
    T rr = R.v[0], ur = R.v[1], vr = R.v[2], pr = R.v[3];
    T ar2 = fdiv(G * pr, rr), qr2 = ur * ur + vr * vr, hr = fdiv(ar2, G1) + h * qr2;
    T unl = ul * ct + vl * st;
    T rl12 = fsqrt(rl), rr12 = fsqrt(rr), rd = fdiv(T(1), rl12 + rr12);
    T ua = (ul * rl12 + ur * rr12) * rd;
    T va = (vl * rl12 + vr * rr12) * rd;
    T ha = (hl * rl12 + hr * rr12) * rd;
    T qa2 = ua * ua + va * va;
    T aa2 = G1 * (ha - h * qa2);
    if (!(aa2 > T(0))) return false;
    T aa = fsqrt(aa2);
    T una = ua * ct + va * st;
    T vna = -ua * st + va * ct;
    T l1 = fmin(una - aa, T(0)), l2 = fmin(una, T(0)), l4 = fmin(una + aa, T(0));
    T du1 = rr - rl, du2 = rr * ur - rl * ul, du3 = rr * vr - rl * vl;
    T du4 = fdiv(pr - pl, G1) + h * (rr * qr2 - rl * ql2);
    T m1 = fdiv(ct * du2 + st * du3 - una * du1, aa);
    T m2 = fdiv(G1 * (du4 - ua * du2 - va * du3 + qa2 * du1), aa * aa);
    T a4 = h * (m1 + m2), a1 = h * (m2 - m1), a3 = du1 - a1 - a4;
    T a2 = fdiv(st * du2 - ct * du3 + vna * du1, aa);
    T a1l1 = a1 * l1, a2l2 = a2 * l2, a3l3 = a3 * l2, a4l4 = a4 * l4;
    T aact = aa * ct, aast = aa * st;
    T Fc0 = rl * unl, Fc1 = pl * ct + rl * ul * unl, Fc2 = pl * st + rl * vl * unl,
      Fc3 = rl * hl * unl;
    T Fd0 = a1l1 + a3l3 + a4l4;
    T Fd1 = a1l1 * (ua - aact) + a2l2 * aa * st + a3l3 * ua + a4l4 * (ua + aact);
    T Fd2 = a1l1 * (va - aast) - a2l2 * aa * ct + a3l3 * va + a4l4 * (va + aast);
    T Fd3 = a1l1 * (ha - una * aa) + a2l2 * aa * vna + a3l3 * h * qa2 + a4l4 * (ha + una * aa);
    f.v[0] = (Fc0 + Fd0) * length;
    f.v[1] = (Fc1 + Fd1) * length;
    f.v[2] = (Fc2 + Fd2) * length;
    f.v[3] = (Fc3 + Fd3) * length;
    return true;
}

// kinetic flux-vector splitting, src/kfvs_flux.f:16-78
template <class T = double>
__device__ __forceinline__ void kfvs_flux(const Consts &c, const EdgeFrame<T> &fr, const VT<T> &L,
                                          const VT<T> &R, VT<T> &f)
{
    const T G1 = T(c.GAMMA1), h = T(0.5), PI = T(c.M_PI_);
    const T length = fr.len, ct = fr.ct, st = fr.st;
    T rl = L.v[0], ul = L.v[1], vl = L.v[2], pl = L.v[3];
    T el = fdiv(pl, G1) + h * rl * (ul * ul + vl * vl);
    T rr = R.v[0], ur = R.v[1], vr = R.v[2], pr = R.v[3];
    T er = fdiv(pr, G1) + h * rr * (ur * ur + vr * vr);
    T unl = ul * ct + vl * st, unr = ur * ct + vr * st;
    T vnl = -ul * st + vl * ct, vnr = -ur * st + vr * ct;
    T betal = fdiv(h * rl, pl), sl = unl * fsqrt(betal);
    T Al = h * (T(1) + ferf(sl));
    T Bl = fdiv(h * fexp(-(sl * sl)), fsqrt(PI * betal));
    T betar = fdiv(h * rr, pr), sr = unr * fsqrt(betar);
    T Ar = h * (T(1) - ferf(sr));
    T Br = fdiv(h * fexp(-(sr * sr)), fsqrt(PI * betar));
    T F0 = rl * (unl * Al + Bl) + rr * (unr * Ar - Br);
    T F1 = ((pl + rl * (unl * unl)) * Al + rl * unl * Bl) + ((pr + rr * (unr * unr)) * Ar - rr * unr * Br);
    T F2 = rl * (unl * Al + Bl) * vnl + rr * (unr * Ar - Br) * vnr;
    T F3 = ((el + pl) * unl * Al + (el + h * pl) * Bl) + ((er + pr) * unr * Ar - (er + h * pr) * Br);
    f.v[0] = F0 * length;
    f.v[1] = (ct * F1 - st * F2) * length;
    f.v[2] = (st * F1 + ct * F2) * length;
    f.v[3] = F3 * length;
}

// HLLC (Toro ch. 10), src/hllc_flux.f:18-154.  The reference adds the
// edge-frame flux to the x-y momentum residual; ROT=1 rotates it back.
template <int ROT, class T = double>
__device__ __forceinline__ void hllc_flux(const Consts &c, const EdgeFrame<T> &fr, const VT<T> &L,
                                          const VT<T> &R, VT<T> &f)
{
    const T gamma = T(c.GAMMA), G1 = T(c.GAMMA1), h = T(0.5), one = T(1), two = T(2);
    const T length = fr.len, ct = fr.ct, st = fr.st;
    T rl = L.v[0], ul = L.v[1] * ct + L.v[2] * st, vl = -L.v[1] * st + L.v[2] * ct, pl = L.v[3];
    T al = sqrt(gamma * pl / rl), El = pl / G1 + h * rl * (ul * ul + vl * vl);
    T rr = R.v[0], ur = R.v[1] * ct + R.v[2] * st, vr = -R.v[1] * st + R.v[2] * ct, pr = R.v[3];
    T ar = sqrt(gamma * pr / rr), Er = pr / G1 + h * rr * (ur * ur + vr * vr);
    T rbar = h * (rl + rr), abar = h * (al + ar);
    T Ppv = h * (pl + pr - (ur - ul) * rbar * abar);
    T Upv = h * (ul + ur - (pr - pl) / (rbar * abar));
    T prmax = (pl > pr) ? pl : pr, prmin = (pl > pr) ? pr : pl;
    T Q = prmax / prmin, pstar, ustar;
    if (Q < two && prmin < Ppv && Ppv < prmax) {
        pstar = Ppv; ustar = Upv;
    } else if (Ppv < prmin) {
        T z = h * (gamma - one) / gamma;
        T a1 = al + ar - h * (gamma - one) * (ur - ul);
        T a2 = al / pow(pl, z) + ar / pow(pr, z);
        pstar = pow(a1 / a2, one / z);
        T Plr = pow(pl / pr, z);
        T b1 = Plr * ul / al + ur / ar + two * (Plr - one) / (gamma - one);
        T b2 = Plr / al + one / ar;
        ustar = b1 / b2;
    } else {
        T po = fmax(T(0), Ppv);
        T Ml = two / (gamma + one) / rl, Nl = (gamma - one) * pl / (gamma + one);
        T gl = sqrt(Ml / (po + Nl));
        T Mr = two / (gamma + one) / rr, Nr = (gamma - one) * pr / (gamma + one);
        T gr = sqrt(Mr / (po + Nr));
        pstar = (gl * pl + gr * pr - (ur - ul)) / (gl + gr);
        ustar = h * (ul + ur + (pstar - pr) * gr - (pstar - pl) * gl);
    }
    T ql = (pstar <= pl) ? one : sqrt(one + h * (gamma + one) * (pstar / pl - one) / gamma);
    T qr = (pstar <= pr) ? one : sqrt(one + h * (gamma + one) * (pstar / pr - one) / gamma);
    T Sl = ul - al * ql, Ss = ustar, Sr = ur + ar * qr;
    T F[4];
    if (T(0) <= Sl) {
        F[0] = rl * ul; F[1] = pl + rl * (ul * ul); F[2] = rl * ul * vl; F[3] = (El + pl) * ul;
    } else if (Sl <= T(0) && T(0) <= Ss) {
        T fact = rl * (Sl - ul) / (Sl - Ss);
        T t1 = El / rl + (Ss - ul) * (Ss + pl / rl / (Sl - ul));
        F[0] = rl * ul + Sl * (fact - rl);
        F[1] = (pl + rl * (ul * ul)) + Sl * (fact * Ss - rl * ul);
        F[2] = rl * ul * vl + Sl * (fact * vl - rl * vl);
        F[3] = (El + pl) * ul + Sl * (fact * t1 - El);
    } else if (Ss <= T(0) && T(0) <= Sr) {
        T fact = rr * (Sr - ur) / (Sr - Ss);
        T t1 = Er / rr + (Ss - ur) * (Ss + pr / rr / (Sr - ur));
        F[0] = rr * ur + Sr * (fact - rr);
        F[1] = (pr + rr * (ur * ur)) + Sr * (fact * Ss - rr * ur);
        F[2] = rr * ur * vr + Sr * (fact * vr - rr * vr);
        F[3] = (Er + pr) * ur + Sr * (fact * t1 - Er);
    } else {
        F[0] = rr * ur; F[1] = pr + rr * (ur * ur); F[2] = rr * ur * vr; F[3] = (Er + pr) * ur;
    }
    if (ROT) {
        T fx = ct * F[1] - st * F[2], fy = st * F[1] + ct * F[2];
        F[1] = fx; F[2] = fy;
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) f.v[i] = F[i] * length;
}

// slip wall: pressure force with 3/4-1/4 weights, src/solid_flux.f:10-17
__device__ __forceinline__ void solid_flux(double2 x1, double2 x2, double p1_, double p2_, V4 &r1,
                                           V4 &r2)
{
    double sx = 0.5 * (x2.y - x1.y);
    double sy = -0.5 * (x2.x - x1.x);
    double p1 = 0.75 * p1_ + 0.25 * p2_;
    double p2 = 0.25 * p1_ + 0.75 * p2_;
    r1.v[0] = 0.0; r1.v[1] = p1 * sx; r1.v[2] = p1 * sy; r1.v[3] = 0.0;
    r2.v[0] = 0.0; r2.v[1] = p2 * sx; r2.v[2] = p2 * sy; r2.v[3] = 0.0;
}

__device__ __forceinline__ void prim2con(const Consts &c, const V4 &p, V4 &u)
{
    u.v[0] = p.v[0];
    u.v[1] = p.v[0] * p.v[1];
    u.v[2] = p.v[0] * p.v[2];
    u.v[3] = p.v[3] / c.GAMMA1 + 0.5 * p.v[0] * (sq(p.v[1]) + sq(p.v[2]));
}

// Far-field state seen by a farfield edge, src/farfield_flux.f:28-58: the
// free stream, or with FarBC 1 (subsonic) the free stream plus a point
// vortex of circulation 0.5*q_inf*cl at (xref, yref).  cl is the value of
// the previous iteration (COMMON /liftdrag/), so this is evaluated once per
// iteration and per edge, outside the stage kernels.  Conserved variables.
__device__ inline void farfield_state(const Consts &c, double cl, double2 x1, double2 x2,
                                      double *Uout)
{
    const double G = c.GAMMA, G1 = c.GAMMA1;
    double uinf, vinf, q2inf, pinf, rinf;
    if (c.mach_inf < 1.0 && c.vortex == 1) {
        double dx = 0.5 * (x1.x + x2.x) - c.xref;
        double dy = 0.5 * (x1.y + x2.y) - c.yref;
        double dref = sqrt(sq(dx) + sq(dy));
        double theta = atan2(dy, dx);
        double circ = 0.5 * c.q_inf * cl;
        double fact1 = circ * sqrt(1.0 - sq(c.mach_inf));
        double fact2 = 2.0 * c.M_PI_ * dref;
        double fact3 = 1.0 - sq(c.mach_inf * sin(theta - c.aoa));
        double fact = fact1 / (fact2 * fact3);
        uinf = c.u_inf + fact * sin(theta);
        vinf = c.v_inf - fact * cos(theta);
        q2inf = sq(uinf) + sq(vinf);
        double fact4 = 1.0 + 0.5 * G1 * (sq(c.q_inf) - q2inf) / sq(c.a_inf);
        rinf = c.r_inf * pow(fact4, 1.0 / G1);
        pinf = c.p_inf * pow(rinf / c.r_inf, G);
    } else {
        uinf = c.u_inf; vinf = c.v_inf; q2inf = sq(uinf) + sq(vinf);
        pinf = c.p_inf; rinf = c.r_inf;
    }
    Uout[0] = rinf;
    Uout[1] = rinf * uinf;
    Uout[2] = rinf * vinf;
    Uout[3] = pinf / G1 + 0.5 * rinf * q2inf;
}

// Steger-Warming farfield flux, src/farfield_flux.f:61-200, with the far-field
// state from farfield_state.  r1/r2 are the contributions to the two end
// points.
__device__ inline void farfield_flux(const Consts &c, const double4 uo, double2 x1, double2 x2,
                                     const V4 &P1, const V4 &P2, V4 &r1, V4 &r2)
{
    const double G = c.GAMMA, G1 = c.GAMMA1;
    const double Uout[4] = {uo.x, uo.y, uo.z, uo.w};
    double ra = 0.5 * (P1.v[0] + P2.v[0]);
    double ua = 0.5 * (P1.v[1] + P2.v[1]);
    double va = 0.5 * (P1.v[2] + P2.v[2]);
    double pa = 0.5 * (P1.v[3] + P2.v[3]);
    double qa2 = sq(ua) + sq(va);
    double aa2 = G * pa / ra;
    double aa = sqrt(G * pa / ra);
    double ha = aa2 / G1 + 0.5 * qa2;
    V4 U1, U2;
    prim2con(c, P1, U1);
    prim2con(c, P2, U2);
    double ct = (x2.y - x1.y), st = -(x2.x - x1.x);
    double lent = sqrt(sq(ct) + sq(st));
    ct = ct / lent; st = st / lent;
    double una = ua * ct + va * st, vna = -ua * st + va * ct;
    double lam[4] = {una, una, una + aa, una - aa}, lp[4], lm[4];
    lp[0] = fmax(lam[0], 0.0); lp[1] = lp[0]; lp[2] = fmax(lam[2], 0.0); lp[3] = fmax(lam[3], 0.0);
#pragma unroll
    for (int k = 0; k < 4; ++k) lm[k] = lam[k] - lp[k];
    // S (the inverse of the right eigenvectors T) is evaluated once, entry by
    // entry with the reference's operations; T row by row.  jac(i,j) =
    // sum_k (l_k T(i,k)) S(k,j) in the reference's order.  Both loops are
    // unrolled, so the divisions are independent of one another.
    const double t1 = 0.5 * ra / aa, m2 = qa2 / aa2;
    const double s1 = 0.5 * G1 * m2 * aa / ra, s2 = una / ra, s3 = G1 * ua / aa, s4 = G1 * va / aa,
                 s5 = G1 / (ra * aa);
    double S[4][4];                         // S[j][k] = S(k, j)
    S[0][0] = 1.0 - 0.5 * G1 * m2; S[0][1] = vna / ra;  S[0][2] = s1 - s2;        S[0][3] = s1 + s2;
    S[1][0] = s3 / aa;             S[1][1] = st / ra;   S[1][2] = (ct - s3) / ra; S[1][3] = -(ct + s3) / ra;
    S[2][0] = s4 / aa;             S[2][1] = -ct / ra;  S[2][2] = (st - s4) / ra; S[2][3] = -(st + s4) / ra;
    S[3][0] = -G1 / aa2;           S[3][1] = 0.0;       S[3][2] = s5;             S[3][3] = s5;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        double T0, T1, T2, T3;
        if (i == 0)      { T0 = 1.0;       T1 = 0.0;        T2 = t1;                 T3 = t1; }
        else if (i == 1) { T0 = ua;        T1 = ra * st;    T2 = t1 * (ua + aa * ct); T3 = t1 * (ua - aa * ct); }
        else if (i == 2) { T0 = va;        T1 = -ra * ct;   T2 = t1 * (va + aa * st); T3 = t1 * (va - aa * st); }
        else             { T0 = 0.5 * qa2; T1 = -ra * vna;  T2 = t1 * (ha + aa * una); T3 = t1 * (ha - aa * una); }
        const double pa0 = lp[0] * T0, pa1 = lp[1] * T1, pa2 = lp[2] * T2, pa3 = lp[3] * T3;
        const double pb0 = lm[0] * T0, pb1 = lm[1] * T1, pb2 = lm[2] * T2, pb3 = lm[3] * T3;
        double f1 = 0.0, f2 = 0.0;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            double a = 0.0, b = 0.0;
            a = a + pa0 * S[j][0]; b = b + pb0 * S[j][0];
            a = a + pa1 * S[j][1]; b = b + pb1 * S[j][1];
            a = a + pa2 * S[j][2]; b = b + pb2 * S[j][2];
            a = a + pa3 * S[j][3]; b = b + pb3 * S[j][3];
            f1 = f1 + a * U1.v[j] + b * Uout[j];
            f2 = f2 + a * U2.v[j] + b * Uout[j];
        }
        r1.v[i] = 0.5 * lent * f1;
        r2.v[i] = 0.5 * lent * f2;
    }
}

// x**1.5 rounded correctly: x * sqrt(x) carried in double-double.  s =
// sqrt(x) is correctly rounded and x - s*s is exact by FMA, so sqrt(x) = s + e
// with e = (x - s*s) / (2 s) to full relative precision (an approximate
// reciprocal suffices for this correction); x*s is split exactly by FMA, and
// hi + (lo + x*e) is rounded once.  Checked against quad-precision x**1.5 on
// 2e7 samples in [0.05, 20]: no difference (libm pow differed in 0.08% of
// them by one ulp).  About a tenth of the instructions of a general pow.
__device__ __forceinline__ double pow_1_5(double x)
{
    const double s = sqrt(x);
    const double r = fma(-s, s, x);
    const double e = (r == 0.0) ? 0.0 : r * (0.5 * __builtin_amdgcn_rcp(s));
    const double hi = x * s;
    const double lo = fma(x, s, -hi);
    return hi + (lo + x * e);
}

// Sutherland viscosity, src/viscosity.f:42-55 (the reference's
// (temp/T_inf)**1.5, correctly rounded)
__device__ __forceinline__ double sutherland_viscosity(const Consts &c, double rho, double p)
{
    double temp = p / rho / 1.0;
    double num = c.T_inf + c.SCONST;
    double den = temp + c.SCONST;
    return pow_1_5(temp / c.T_inf) * (num / den) / c.rey;
}

// SA eddy viscosity, src/viscosity.f:58-71
__device__ __forceinline__ double mu_turb(const Consts &c, double mu_lam, double rho, double nu)
{
    double chi = nu * rho / mu_lam;
    double chi3 = pow3(chi);
    double fv1 = chi3 / (chi3 + c.Cv11);
    return rho * nu * fv1;
}

}  // namespace fvm
