// Device-side definitions shared by the hfem_amd kernels (gfx950).
#pragma once
#include <hip/hip_runtime.h>

namespace hfem {

// Derived sizes of one element family.
struct Dims {
    int nd, ncu, nch, nc, npv, ngv, npf, ngf, nfe, ndf, na, nb, nd1, np;
};

struct Master {
    const double *shapvt;   // (ngv, npv, nd+1)
    const double *shapvg;   // (npv, ngv, nd+1)
    const double *shapft;   // (ngf, npf, nd)
    const double *shapfg;   // (npf, ngf, nd)
    const int *perm;        // (npf, nfe)
    const int *invperm;     // (npv, nfe): k with perm(k,a) = n, or -1
};

struct Phys {
    double param[6];        // gamma, eps_lm, Re, Pr, Minf, tau
    double ui[8];           // far-field state
    double fc_u, fc_q;
    int tdep;
};

// Element sources: which items (elements, or rounds of elements) of a range
// [0, n) a workgroup takes.  A kernel body runs
//   for (t = src.first(); t < n; t = src.next(t)) { ... }
// with every thread of the workgroup calling next() together.
//
// Workgroup b takes items b, b + grid, b + 2 grid, ...
struct StrideSource {
    __device__ int first() const { return blockIdx.x; }
    __device__ int next(int t) const { return t + gridDim.x; }
};

// Tickets from a device counter shared by the workgroups of a ragged launch: the
// ranges of several buckets are laid end to end, [lo, lo + n) being the current
// one.  next() draws the next ticket (thread 0 adds one to the counter and
// passes the value on through LDS); a ticket outside the current range ends the
// run and is kept in carry for the kernel to dispatch.  Tickets grow
// monotonically for each workgroup, so a workgroup visits the buckets in order.
struct TicketSource {
    unsigned *ctr;
    int *slot;      // one int of LDS
    int lo, n, carry;
    __device__ int draw()
    {
        if (threadIdx.x == 0) *slot = (int)atomicAdd(ctr, 1u);
        __syncthreads();
        const int v = __builtin_amdgcn_readfirstlane(*slot);
        __syncthreads();
        return v;
    }
    __device__ int first() const { return carry - lo; }
    __device__ int next(int)
    {
        const int v = draw();
        if (v >= lo && v < lo + n) return v - lo;
        carry = v;
        return n;
    }
};

// Forward-mode dual number with one derivative direction.  The flux Jacobian at
// a quadrature point is formed one column per thread: the thread seeds its
// direction and evaluates the flux once.
struct dual {
    double v, d;
};
__device__ __forceinline__ dual mk(double v, double d = 0.0) { return dual{v, d}; }
__device__ __forceinline__ dual operator+(dual a, dual b) { return {a.v + b.v, a.d + b.d}; }
__device__ __forceinline__ dual operator-(dual a, dual b) { return {a.v - b.v, a.d - b.d}; }
__device__ __forceinline__ dual operator-(dual a) { return {-a.v, -a.d}; }
__device__ __forceinline__ dual operator*(dual a, dual b) { return {a.v * b.v, a.d * b.v + a.v * b.d}; }
__device__ __forceinline__ dual operator*(double s, dual a) { return {s * a.v, s * a.d}; }
__device__ __forceinline__ dual inv(dual a) { double r = 1.0 / a.v; return {r, -r * r * a.d}; }
__device__ __forceinline__ dual operator/(dual a, dual b) { return a * inv(b); }
__device__ __forceinline__ dual dsqrt(dual a) { double r = sqrt(a.v); return {r, 0.5 / r * a.d}; }
__device__ __forceinline__ dual dfabs(dual a) { return a.v < 0 ? -a : a; }

// Compressible Navier-Stokes flux f[j + ncu*x] of w[nc] = (u, q) with mu = 1:
// inviscid part + tau(q)/Re + gamma/(Re Pr (gamma-1)) d(p/rho)
// (application/FM/nsNew/flux_ns2dNEW.c:78-85, flux_ns3dNEW.c; q columns are the
// gradient components as in Appendix A of the element layout).
template <int ND>
__device__ __forceinline__ void ns_flux(const double *param, const dual *w, dual *f)
{
    constexpr int NCU = ND + 2;
    const double gam = param[0], Re = param[2], Pr = param[3];
    const double gm1 = gam - 1.0;
    dual r = w[0], ri = inv(r), rE = w[ND + 1];
    dual v[ND], g[ND][ND], th[ND];
    dual ke = mk(0.0);
#pragma unroll
    for (int i = 0; i < ND; i++) { v[i] = w[1 + i] * ri; ke = ke + v[i] * v[i]; }
    ke = 0.5 * ke;
    dual p = gm1 * (rE - r * ke);
#pragma unroll
    for (int d = 0; d < ND; d++) {
        const dual *q = &w[NCU + d * NCU];
#pragma unroll
        for (int i = 0; i < ND; i++) g[i][d] = (q[1 + i] - v[i] * q[0]) * ri;
        dual t = (q[ND + 1] - (rE * ri) * q[0]) * ri;
#pragma unroll
        for (int i = 0; i < ND; i++) t = t - v[i] * g[i][d];
        th[d] = gm1 * t;
    }
    dual div = mk(0.0);
#pragma unroll
    for (int i = 0; i < ND; i++) div = div + g[i][i];
    const double kappa = gam / (Re * Pr * gm1);
    dual H = (rE + p) * ri;
#pragma unroll
    for (int d = 0; d < ND; d++) {
        f[NCU * d] = w[1 + d];
        dual e = w[1 + d] * H;
#pragma unroll
        for (int i = 0; i < ND; i++) {
            dual t = g[i][d] + g[d][i];
            if (i == d) t = t - (2.0 / 3.0) * div;
            dual tau = (1.0 / Re) * t;
            dual m = w[1 + i] * v[d] + tau;
            if (i == d) m = m + p;
            f[1 + i + NCU * d] = m;
            e = e + v[i] * tau;
        }
        f[ND + 1 + NCU * d] = e + kappa * th[d];
    }
}

// |lambda| with the smooth fix 0.5 (lambda^2 / (eps c) + eps c) below eps c (getan.c).
__device__ __forceinline__ dual abs_eig(dual l, double eps, dual cs)
{
    dual a = dfabs(l);
    if (eps > 0) {
        dual ec = eps * cs;
        if (a.v < ec.v) a = 0.5 * (a * a / ec + ec);
    }
    return a;
}

// Normal flux Jacobian A_n(uh) (absflag 0) or |A_n|(uh) (absflag 1) in the
// eigen-split form of application/FM/getan.c: column k of M (M[k*nch+row]) is
// lam3 e_k + cc1_k (1, v, h) + cc2_k (0, n, un).
template <int ND>
__device__ __forceinline__ void ns_normal_jac(double gam, double eps, const dual *uh, const double *nrm,
                                              int absflag, dual *M)
{
    constexpr int NCH = ND + 2;
    const double gm1 = gam - 1.0;
    dual r = uh[0], ri = inv(r), v[ND], af = mk(0.0), un = mk(0.0);
#pragma unroll
    for (int i = 0; i < ND; i++) { v[i] = uh[1 + i] * ri; af = af + v[i] * v[i]; un = un + nrm[i] * v[i]; }
    af = 0.5 * af;
    dual rE = uh[ND + 1], E = rE * ri;
    dual p = gm1 * (rE - r * af);
    dual h = E + p * ri;
    dual c2 = gam * (p * ri), cs = dsqrt(c2);
    dual l1 = un + cs, l2 = un - cs, l3 = un;
    if (absflag) {
        l1 = abs_eig(l1, eps, cs);
        l2 = abs_eig(l2, eps, cs);
        l3 = abs_eig(l3, eps, cs);
    }
    dual s1 = 0.5 * (l1 + l2), s2 = 0.5 * (l1 - l2), s13 = s1 - l3;
#pragma unroll
    for (int k = 0; k < NCH; k++) {
        dual cc1, cc2;
        if (k == 0) {
            cc1 = gm1 * (s13 * af) / c2 - (s2 * un) / cs;
            cc2 = gm1 * (s2 * af) / cs - s13 * un;
        } else if (k <= ND) {
            dual vi = v[k - 1];
            double ni = nrm[k - 1];
            cc1 = (-gm1) * ((s13 * vi) / c2) + ni * (s2 / cs);
            cc2 = (-gm1) * ((s2 * vi) / cs) + ni * s13;
        } else {
            cc1 = gm1 * (s13 / c2);
            cc2 = gm1 * (s2 / cs);
        }
        M[k * NCH] = cc1;
#pragma unroll
        for (int m = 0; m < ND; m++) M[k * NCH + 1 + m] = cc1 * v[m] + nrm[m] * cc2;
        M[k * NCH + ND + 1] = cc1 * h + cc2 * un;
        M[k * NCH + k] = M[k * NCH + k] + l3;
    }
}

// Per-element pointwise flux data written by the flux kernel, read by assembly.
// Volume quantities come folded with the Jacobian adjugate Xx(g,x,r) = jac d xi_r/d x_x:
//   fX[g + ngv*(j + ncu*r)]                  sum_x f(g,j,x) Xx(g,x,r)
//   row block j at J + j*sj (one contiguous record per row of the flux Jacobians):
//     [g + ngv*(r + nd*c)]                   sum_x d f(g,j,x)/d udg_c Xx(g,x,r)
//     [ngv*nd*nc + pp + np*c]                d fhat(pp,j)/d udg_c            (fhu)
//     [ngv*nd*nc + np*nc + pp + np*k]        d fhat(pp,j)/d uh_k             (fhuh)
//   fh[pp + np*j], fb[pp + np*j]            interior / far-field face flux
//   fbu[pp + np*(c + nc*j)], fbuh[pp + np*(k + nch*j)]   far-field Jacobians
struct FluxLayout {
    long fX, J, sj, fh, fb, fbu, fbuh, total;
};

__host__ __device__ inline FluxLayout flux_layout(const Dims &d)
{
    FluxLayout s;
    long o = 0;
    auto take = [&](long n) { long r = o; o += (n + 31) / 32 * 32; return r; };
    s.fX = take((long)d.ngv * d.ncu * d.nd);
    s.sj = (long)d.ngv * d.nd * d.nc + (long)d.np * d.nc + (long)d.np * d.nch;
    s.J = take((long)d.ncu * s.sj);
    s.fh = take((long)d.np * d.nch);
    s.fb = take((long)d.np * d.nch);
    s.fbu = take((long)d.np * d.nch * d.nc);
    s.fbuh = take((long)d.np * d.nch * d.nch);
    s.total = o;
    return s;
}

}  // namespace hfem
