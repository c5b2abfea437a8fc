// Discrete adjoint of the DG hydro step (gfx950, FP64): the reverse of the stage RHS with respect
// to the state, the per-node Manning n and the per-node depth DP, plus the gauge-misfit loss.
//
// The reference has no adjoint; this differentiates the operations of its explicit step as this
// library evaluates them (OCEAN/LAND/FLOW/INTERNAL_EDGE_HYDRO with the LLF flux
// src/numerical_flux.F:750-1022, RHS_DG_HYDRO src/rhs_dg_hydro.F:18-473, the SSP-RK update
// src/DG_hydro_timestep.F:238-290).  Non-smooth operations take the branch the forward values
// select: the LLF wave-speed bound differentiates through the largest of its six candidates, the
// Manning friction floor CF and the 0.9/DT friction cap pass no derivative when active, and
// Eirik's dry-ocean-edge fix differentiates |h|.  SLOPELIMITER5 is differentiated with its active
// set frozen (limiter_adj_kernel and the node passes after it).  WETDRY after the limiter is
// differentiated with its branches frozen too (wetdry_adj_element, run by the limiter kernel's
// adjoint variant on the limited values it forms again): the wet/dry flags at the start
// of each stage come from the forward run, a dry element's RHS is zero, and an edge with a dry
// side takes the reverse of the dry-side flux case the forward values select (dry_side_adj).
// Without the limiter WETDRY is differentiated only where it is the identity (every element wet
// and staying wet): the forward sweeps flag any element WETDRY changes, and the gradient call then
// fails.
//
// Kernel layout follows the forward stage kernel: one thread per element, gathers only.  The
// cotangent of an internal edge flux is the sum of what both neighbours' RHS cotangents ask of it;
// each element forms that sum from its own and its neighbour's cotangents and keeps the part of
// the flux derivative that belongs to its own trace, so no scatter or atomics are needed.
#include "dg_device.hpp"

namespace dg {

// Reverse of llf_flux: given fb = d(loss)/d(f0,f1,f2), returns the cotangents of
// (ze_in, ze_ex, qx_in, qx_ex, qy_in, qy_ex, hb_in, hb_ex) in bar[0..7].  NL: the nonlinear
// equations (NLEQ = 1, LEQ = 0) with the vanishing terms left out.
template <bool NL>
__device__ __forceinline__ void llf_flux_adj(const Phys& P, double zi, double ze, double xi, double xe,
                                             double yi, double ye, double hi, double he, double sfi,
                                             double sfe, double nx, double ny, double G,
                                             const double fb[3], double bar[8]) {
    // derivatives need not round like the forward step: let products and sums fuse
#pragma clang fp contract(fast)
    const double NLEQ = NL ? 1.0 : P.nleq, LEQ = NL ? 0.0 : P.leq, FG_L = NL ? 0.0 : P.fg_l;
    const double NLEQG = G * NLEQ;
    // ---- forward values needed by the reverse sweep
    const double j1 = ze - zi, j2 = xe - xi, j3 = ye - yi;
    const double side_z[2] = {zi, ze}, side_x[2] = {xi, xe}, side_y[2] = {yi, ye}, side_h[2] = {hi, he};
    const double side_s[2] = {sfi, sfe};
    double ht[2], fh[2], u[2], v[2], un[2], c[2], K[2];
#pragma unroll
    for (int s = 0; s < 2; s++) {
        ht[s] = side_z[s] * NLEQ + side_h[s];
        fh[s] = NL ? 1.0 / ht[s] : 1.0 / (NLEQ * ht[s] + LEQ);
        u[s] = side_x[s] * fh[s];
        v[s] = side_y[s] * fh[s];
        un[s] = (u[s] * nx + v[s] * ny) * NLEQ;
        const double nxs = nx * side_s[s];
        K[s] = ny * ny + nxs * nxs;
        c[s] = sqrt(G * ht[s] * K[s]);
    }
    // largest wave-speed candidate, in the forward's comparison order
    const double cand[6] = {fabs(un[0] + c[0]), fabs(un[0]), fabs(un[0] - c[0]),
                            fabs(un[1] + c[1]), fabs(un[1]), fabs(un[1] - c[1])};
    int arg = 0;
    double em = cand[0];
#pragma unroll
    for (int q = 1; q < 6; q++)
        if (cand[q] > em) { em = cand[q]; arg = q; }
    // ---- reverse
    const double a1b = fb[0], a2b = fb[1], a3b = fb[2];
    double emb, j1b, j2b, j3b;
    if constexpr (NL) {
        emb = -0.5 * (j1 * fb[0] + j2 * fb[1] + j3 * fb[2]);
        j1b = -0.5 * em * fb[0];
        j2b = -0.5 * em * fb[1];
        j3b = -0.5 * em * fb[2];
    } else {
        const double nx2 = NLEQ + LEQ * nx * nx, ny2 = NLEQ + LEQ * ny * ny, nxy = nx * ny;
        emb = -0.5 * j1 * fb[0] - 0.5 * (j2 * nx2 + LEQ * j3 * nxy) * fb[1] -
              0.5 * (j3 * ny2 + LEQ * j2 * nxy) * fb[2];
        j1b = -0.5 * em * fb[0];
        j2b = -0.5 * em * (nx2 * fb[1] + LEQ * nxy * fb[2]);
        j3b = -0.5 * em * (LEQ * nxy * fb[1] + ny2 * fb[2]);
    }
    double unb[2] = {0.0, 0.0}, cb[2] = {0.0, 0.0};
    {
        const int s = arg / 3, q = arg % 3;
        const double val = q == 0 ? un[s] + c[s] : (q == 1 ? un[s] : un[s] - c[s]);
        const double sg = val > 0.0 ? 1.0 : (val < 0.0 ? -1.0 : 0.0);
        unb[s] += sg * emb;
        if (q == 0) cb[s] += sg * emb;
        else if (q == 2) cb[s] -= sg * emb;
    }
    // flux averages: a = 0.5*((f_in + f_ex).n)
    const double fx1b = 0.5 * nx * a1b, fy1b = 0.5 * ny * a1b;
    const double fx2b = 0.5 * nx * a2b, fy2b = 0.5 * ny * a2b;
    const double fx3b = 0.5 * nx * a3b, fy3b = 0.5 * ny * a3b;
    double zb[2], xb[2], yb[2], hbb[2];
#pragma unroll
    for (int s = 0; s < 2; s++) {
        const double z = side_z[s], x = side_x[s], y = side_y[s], h = side_h[s], sf = side_s[s];
        const double fnl = NL ? 1.0 : NLEQ + LEQ * h;
        const double fu = NLEQ * x, fv = NLEQ * y, fg = NLEQG * z;
        double zbar = 0.0, xbar = 0.0, ybar = 0.0, hbar = 0.0;
        // fy3 = hvv + gh2; fx3 = huv*sf; fy2 = huv; fx2 = (huu + gh2)*sf
        const double hvvb = fy3b;
        double gh2b = fy3b + sf * fx2b;
        const double huvb = sf * fx3b + fy2b;
        const double huub = sf * fx2b;
        // gh2 = fg*(0.5*z + h) + FG_L*z
        double fgb = (0.5 * z + h) * gh2b;
        zbar += (NL ? 0.5 * fg : 0.5 * fg + FG_L) * gh2b;
        hbar += fg * gh2b;
        // huu = fu*u, hvv = fv*v, huv = fu*v
        double fub = u[s] * huub + v[s] * huvb;
        double ub = fu * huub;
        const double fvb = v[s] * hvvb;
        double vb = fv * hvvb + fu * huvb;
        // un = (u*nx + v*ny)*NLEQ
        ub += nx * NLEQ * unb[s];
        vb += ny * NLEQ * unb[s];
        // c = sqrt(G*ht*K): dc/dht = G*K/(2c) = c/(2*ht), and fh = 1/ht for the nonlinear equations
        double htb;
        if constexpr (NL) htb = c[s] > 0.0 ? 0.5 * c[s] * fh[s] * cb[s] : 0.0;
        else htb = c[s] > 0.0 ? G * K[s] / (2.0 * c[s]) * cb[s] : 0.0;
        // u = x*fh, v = y*fh
        xbar += fh[s] * ub;
        ybar += fh[s] * vb;
        const double fhb = x * ub + y * vb;
        // fh = 1/(NLEQ*ht + LEQ)
        htb += -NLEQ * fh[s] * fh[s] * fhb;
        // fg = NLEQG*z, fu = NLEQ*x, fv = NLEQ*y
        zbar += NLEQG * fgb;
        xbar += NLEQ * fub;
        ybar += NLEQ * fvb;
        // fx1 = x*fnl*sf, fy1 = y*fnl
        xbar += fnl * sf * fx1b;
        ybar += fnl * fy1b;
        if constexpr (!NL) hbar += LEQ * (x * sf * fx1b + y * fy1b);
        // ht = z*NLEQ + h
        zbar += NLEQ * htb;
        hbar += htb;
        zb[s] = zbar; xb[s] = xbar; yb[s] = ybar; hbb[s] = hbar;
    }
    // jumps
    bar[0] = zb[0] - j1b;
    bar[1] = zb[1] + j1b;
    bar[2] = xb[0] - j2b;
    bar[3] = xb[1] + j2b;
    bar[4] = yb[0] - j3b;
    bar[5] = yb[1] + j3b;
    bar[6] = hbb[0];
    bar[7] = hbb[1];
}

// The part of llf_flux_adj that one side of an internal edge keeps: the cotangents of side `own`'s
// arguments (ze, qx, qy, hb) only.  Both elements of an edge run it for their own side, with the
// same canonical (in, ex) argument order, so no lane evaluates the other side's reverse.
template <bool NL>
__device__ __forceinline__ void llf_flux_adj_side(const Phys& P, double zi, double ze, double xi, double xe,
                                                  double yi, double ye, double hi, double he, double sfi,
                                                  double sfe, double nx, double ny, double G,
                                                  const double fb[3], int own, double& zbar,
                                                  double& xbar, double& ybar, double& hbar) {
#pragma clang fp contract(fast)
    const double NLEQ = NL ? 1.0 : P.nleq, LEQ = NL ? 0.0 : P.leq, FG_L = NL ? 0.0 : P.fg_l;
    const double NLEQG = G * NLEQ;
    const double j1 = ze - zi, j2 = xe - xi, j3 = ye - yi;
    // forward values of both sides (the wave-speed bound needs both)
    const double side_z[2] = {zi, ze}, side_x[2] = {xi, xe}, side_y[2] = {yi, ye}, side_h[2] = {hi, he};
    const double side_s[2] = {sfi, sfe};
    double fh[2], u[2], v[2], un[2], c[2], K[2];
#pragma unroll
    for (int q = 0; q < 2; q++) {
        const double ht = side_z[q] * NLEQ + side_h[q];
        fh[q] = NL ? 1.0 / ht : 1.0 / (NLEQ * ht + LEQ);
        u[q] = side_x[q] * fh[q];
        v[q] = side_y[q] * fh[q];
        un[q] = (u[q] * nx + v[q] * ny) * NLEQ;
        const double nxs = nx * side_s[q];
        K[q] = ny * ny + nxs * nxs;
        c[q] = sqrt(G * ht * K[q]);
    }
    const double cand[6] = {fabs(un[0] + c[0]), fabs(un[0]), fabs(un[0] - c[0]),
                            fabs(un[1] + c[1]), fabs(un[1]), fabs(un[1] - c[1])};
    int arg = 0;
    double em = cand[0];
#pragma unroll
    for (int q = 1; q < 6; q++)
        if (cand[q] > em) { em = cand[q]; arg = q; }
    double emb, j1b, j2b, j3b;
    if constexpr (NL) {
        emb = -0.5 * (j1 * fb[0] + j2 * fb[1] + j3 * fb[2]);
        j1b = -0.5 * em * fb[0];
        j2b = -0.5 * em * fb[1];
        j3b = -0.5 * em * fb[2];
    } else {
        const double nx2 = NLEQ + LEQ * nx * nx, ny2 = NLEQ + LEQ * ny * ny, nxy = nx * ny;
        emb = -0.5 * j1 * fb[0] - 0.5 * (j2 * nx2 + LEQ * j3 * nxy) * fb[1] -
              0.5 * (j3 * ny2 + LEQ * j2 * nxy) * fb[2];
        j1b = -0.5 * em * fb[0];
        j2b = -0.5 * em * (nx2 * fb[1] + LEQ * nxy * fb[2]);
        j3b = -0.5 * em * (LEQ * nxy * fb[1] + ny2 * fb[2]);
    }
    // this side's forward values
    const double z = own ? ze : zi, x = own ? xe : xi, y = own ? ye : yi, h = own ? he : hi;
    const double sf = own ? sfe : sfi;
    const double fhs = own ? fh[1] : fh[0], us = own ? u[1] : u[0], vs = own ? v[1] : v[0];
    const double cs = own ? c[1] : c[0], Ks = own ? K[1] : K[0];
    double unb = 0.0, cb = 0.0;
    if (arg / 3 == own) {
        const int q = arg % 3;
        const double uns = own ? un[1] : un[0];
        const double val = q == 0 ? uns + cs : (q == 1 ? uns : uns - cs);
        const double sg = val > 0.0 ? 1.0 : (val < 0.0 ? -1.0 : 0.0);
        unb = sg * emb;
        if (q == 0) cb = sg * emb;
        else if (q == 2) cb = -sg * emb;
    }
    const double fx1b = 0.5 * nx * fb[0], fy1b = 0.5 * ny * fb[0];
    const double fx2b = 0.5 * nx * fb[1], fy2b = 0.5 * ny * fb[1];
    const double fx3b = 0.5 * nx * fb[2], fy3b = 0.5 * ny * fb[2];
    const double fnl = NL ? 1.0 : NLEQ + LEQ * h;
    const double fu = NLEQ * x, fv = NLEQ * y, fg = NLEQG * z;
    double zb = 0.0, xb = 0.0, yb = 0.0, hb = 0.0;
    const double hvvb = fy3b;
    const double gh2b = fy3b + sf * fx2b;
    const double huvb = sf * fx3b + fy2b;
    const double huub = sf * fx2b;
    const double fgb = (0.5 * z + h) * gh2b;
    zb += (NL ? 0.5 * fg : 0.5 * fg + FG_L) * gh2b;
    hb += fg * gh2b;
    const double fub = us * huub + vs * huvb;
    double ub = fu * huub;
    const double fvb = vs * hvvb;
    double vb = fv * hvvb + fu * huvb;
    ub += nx * NLEQ * unb;
    vb += ny * NLEQ * unb;
    double htb;
    if constexpr (NL) htb = cs > 0.0 ? 0.5 * cs * fhs * cb : 0.0;
    else htb = cs > 0.0 ? G * Ks / (2.0 * cs) * cb : 0.0;
    xb += fhs * ub;
    yb += fhs * vb;
    const double fhb = x * ub + y * vb;
    htb += -NLEQ * fhs * fhs * fhb;
    zb += NLEQG * fgb;
    xb += NLEQ * fub;
    yb += NLEQ * fvb;
    xb += fnl * sf * fx1b;
    yb += fnl * fy1b;
    if constexpr (!NL) hb += LEQ * (x * sf * fx1b + y * fy1b);
    zb += NLEQ * htb;
    hb += htb;
    zbar = own ? zb + j1b : zb - j1b;
    xbar = own ? xb + j2b : xb - j2b;
    ybar = own ? yb + j3b : yb - j3b;
    hbar = hb;
}

// The mass flux F0 of an internal edge at one point as the forward stage kernel forms it: its
// traces in the forward's operation order (no contraction) and the LLF flux, so that the dry-side
// case (src/internal_edge_hydro.F:212-302) is taken as the forward took it.
template <int DOF, bool NL>
__device__ __forceinline__ double edge_f0_exact(const Phys& ph, const double (&Vi)[3][DOF], const double (&Vx)[3][DOF],
                                                const double (&hbi)[3], const double (&hbx)[3], const double* pi_,
                                                const double* px_, double sfi, double sfx, double nx, double ny) {
    const double ze_in = trace1<DOF>(Vi[0], pi_), qx_in = trace1<DOF>(Vi[1], pi_), qy_in = trace1<DOF>(Vi[2], pi_);
    const double ze_ex = trace1<DOF>(Vx[0], px_), qx_ex = trace1<DOF>(Vx[1], px_), qy_ex = trace1<DOF>(Vx[2], px_);
    const double hb_in = trace1<3>(hbi, pi_), hb_ex = trace1<3>(hbx, px_);
    double f0, f1, f2;
    llf<NL>(ph, ze_in, ze_ex, qx_in, qx_ex, qy_in, qy_ex, hb_in, hb_ex, sfi, sfx, nx, ny, ph.g, f0, f1, f2);
    return f0;
}

// The P1 bathymetry modes of element g as the forward stage kernel stages them (no contraction).
__device__ __forceinline__ void elem_hb_exact(const ElemData& E, int ne, int g, double hb[3]) {
    const double c13 = 1.0 / 3.0, c16 = 1.0 / 6.0;
    const double d0 = E.dpv[g], d1 = E.dpv[(size_t)ne + g], d2 = E.dpv[2 * (size_t)ne + g];
    hb[0] = c13 * (d0 + d1 + d2);
    hb[1] = -(c16 * (d0 + d1)) + c13 * d2;
    hb[2] = -(0.5 * d0) + 0.5 * d1;
}

// Side geometry as the forward forms it (no contraction).
__device__ __forceinline__ void side_geometry_exact(const ElemData& E, int ne, int e, int L, int role, double& xlen,
                                                    double& nx, double& ny) {
    int va = (L + 1) % 3, vb = (L + 2) % 3;
    if (role) { const int t = va; va = vb; vb = t; }
    const double dy = E.yv[(size_t)vb * ne + e] - E.yv[(size_t)va * ne + e];
    const double dx = E.xv[(size_t)vb * ne + e] - E.xv[(size_t)va * ne + e];
    xlen = sqrt(dy * dy + dx * dx);
    nx = dy / xlen;
    ny = -dx / xlen;
}

template <bool NL>
__device__ void dry_side_adj_cases(const Phys& ph, int cas, int role, double zi, double ze, double xi, double xe,
                                   double yi, double ye, double hi, double he, double sfi, double sfx, double nx,
                                   double ny, const double cin[3], const double cex[3], double& zb, double& xb,
                                   double& yb, double& hbar);

// Reverse of an internal edge point where one side is dry (src/internal_edge_hydro.F:212-302),
// for this element's own side: the case is taken from the forward's mass flux F0 (recomputed as
// the forward formed it); cin/cex are d(loss)/d(flux the "in"/"ex" side received).  Cases:
// 1 reflect into a dry "in" (both sides take the reflected flux's F0, "ex" its momentum: a function
// of the "ex" traces only), 2 zero-gravity momentum for a dry "in", 3 reflect into a dry "ex",
// 4 zero-gravity momentum for a dry "ex", 0 the plain flux.  Outputs this side's trace cotangents.
template <int P, bool NL>
__device__ void dry_side_adj(const Phys& ph, const ElemData& E, int ne, int e, int nb, int L, int role, int nled,
                             int gi, int gx, const double (&U)[3][(P + 1) * (P + 2) / 2],
                             const double (&V)[3][(P + 1) * (P + 2) / 2], int wd, int wdn,
                             double zi, double ze, double xi, double xe, double yi, double ye, double hi, double he,
                             double sfi, double sfx, double nx, double ny, const double cin[3], const double cex[3],
                             double& zb, double& xb, double& yb, double& hbar) {
    constexpr int DOF = (P + 1) * (P + 2) / 2;
    const DevTables& T = tab<P>();
    int cas = 0;
    {
        // the forward's F0, operation by operation
        double hbo[3], hbn[3], xlen_x, nx_x, ny_x;
        elem_hb_exact(E, ne, e, hbo);
        elem_hb_exact(E, ne, nb, hbn);
        side_geometry_exact(E, ne, e, L, role, xlen_x, nx_x, ny_x);
        const int Li = role ? nled : L, Lx = role ? L : nled;
        const double f0 = role ? edge_f0_exact<DOF, NL>(ph, V, U, hbn, hbo, T.phi_edge[Li][gi], T.phi_edge[Lx][gx], sfi, sfx, nx_x, ny_x)
                               : edge_f0_exact<DOF, NL>(ph, U, V, hbo, hbn, T.phi_edge[Li][gi], T.phi_edge[Lx][gx], sfi, sfx, nx_x, ny_x);
        const int wd_in_el = role ? wdn : wd, wd_ex_el = role ? wd : wdn;
        if (fabs(f0) > 1.0e-8) {
            if (wd_in_el == 0) cas = f0 > 0.0 ? 1 : 2;
            else if (wd_ex_el == 0) cas = f0 < 0.0 ? 3 : 4;
        }
    }
    dry_side_adj_cases<NL>(ph, cas, role, zi, ze, xi, xe, yi, ye, hi, he, sfi, sfx, nx, ny, cin, cex, zb, xb, yb, hbar);
}

// The derivative part of dry_side_adj, for case `cas`.
template <bool NL>
__device__ void dry_side_adj_cases(const Phys& ph, int cas, int role, double zi, double ze, double xi, double xe,
                                   double yi, double ye, double hi, double he, double sfi, double sfx, double nx,
                                   double ny, const double cin[3], const double cex[3], double& zb, double& xb,
                                   double& yb, double& hbar) {
#pragma clang fp contract(fast)
    zb = xb = yb = hbar = 0.0;
    if (cas == 0 || cas == 2 || cas == 4) {
        double fbn[3] = {cin[0] + cex[0], cin[1] + cex[1], cin[2] + cex[2]};
        double fbg[3] = {0.0, 0.0, 0.0};
        if (cas == 2) { fbn[1] = cex[1]; fbn[2] = cex[2]; fbg[1] = cin[1]; fbg[2] = cin[2]; }
        if (cas == 4) { fbn[1] = cin[1]; fbn[2] = cin[2]; fbg[1] = cex[1]; fbg[2] = cex[2]; }
        llf_flux_adj_side<NL>(ph, zi, ze, xi, xe, yi, ye, hi, he, sfi, sfx, nx, ny, ph.g, fbn, role, zb, xb, yb, hbar);
        if (cas != 0) {
            double z2, x2, y2, h2;
            llf_flux_adj_side<NL>(ph, zi, ze, xi, xe, yi, ye, hi, he, sfi, sfx, nx, ny, 0.0, fbg, role, z2, x2, y2, h2);
            zb += z2; xb += x2; yb += y2; hbar += h2;
        }
        return;
    }
    // reflection: the flux is a function of one side's traces only (the "ex" side in case 1, the
    // "in" side in case 3); the other side keeps nothing
    const int src = cas == 1 ? 1 : 0;
    if (role != src) return;
    const double tx = -ny, ty = nx, den = nx * ty - ny * tx;
    const double sx = src ? xe : xi, sy = src ? ye : yi;
    const double qn = sx * nx + sy * ny, qt = sx * tx + sy * ty;
    const double mqx = (ty * -qn - ny * qt) / den, mqy = (-tx * -qn + nx * qt) / den;
    // d(mqx, mqy)/d(sx, sy)
    const double mxx = (-ty * nx - ny * tx) / den, mxy = (-ty * ny - ny * ty) / den;
    const double myx = (tx * nx + nx * tx) / den, myy = (tx * ny + nx * ty) / den;
    double bar[8];
    if (src) {
        const double fbr[3] = {cin[0] + cex[0], cex[1], cex[2]};
        llf_flux_adj<NL>(ph, ze, ze, mqx, xe, mqy, ye, he, he, sfx, sfx, nx, ny, ph.g, fbr, bar);
        zb = bar[0] + bar[1];
        hbar = bar[6] + bar[7];
        xb = bar[3] + mxx * bar[2] + myx * bar[4];
        yb = bar[5] + mxy * bar[2] + myy * bar[4];
    } else {
        const double fbr[3] = {cin[0] + cex[0], cin[1], cin[2]};
        llf_flux_adj<NL>(ph, zi, zi, xi, mqx, yi, mqy, hi, hi, sfi, sfi, nx, ny, ph.g, fbr, bar);
        zb = bar[0] + bar[1];
        hbar = bar[6] + bar[7];
        xb = bar[2] + mxx * bar[3] + myx * bar[5];
        yb = bar[4] + mxy * bar[3] + myy * bar[5];
    }
}

struct AdjArgs {
    ElemData ed;
    Phys ph;
    Forcing fo;
    Met mt;                              // NWS=8 node forcing of the step (when met != 0)
    int met;
    const double* __restrict__ u;        // stage state U(IRK) [3][dof][ne]
    int irk;                             // stage index
    int nsrc;                            // later-stage cotangents entering this stage
    const double* src[KMAXRK];           // cotangent of U(I'+1), I' >= IRK
    double acoef[KMAXRK];                // ATVD(I',IRK)
    double bcoef[KMAXRK];                // BTVD(I',IRK) (times the scenario's DT in the kernel)
    double* __restrict__ out;            // cotangent of U(IRK)
    double* __restrict__ gdp;            // [3][ne] accumulated cotangent of the vertex depths
    double* __restrict__ gnbar;          // [ne] accumulated cotangent of the element mean Manning n
    const uint8_t* __restrict__ wd;      // WDFLG during the stage, or nullptr (every element wet)
    // block-local staging, as in the forward stage kernel: tile t stages its elements and its halo
    // hidx[hoff[t]..hoff[t+1]); snbl packs each side's local neighbour index (10 bits each)
    const int* __restrict__ hoff;
    const int* __restrict__ hidx;
    const int* __restrict__ snbl;
    int ns;                              // staged elements per block (LDS row length)
};

// RHS cotangent of element `el` (its own stage contribution): sum_j b_j DT src_j.
template <int DOF>
__device__ __forceinline__ void load_lambda(const AdjArgs& a, int ne, int el, double dtdp, double (&L)[3][DOF]) {
#pragma unroll
    for (int f = 0; f < 3; f++)
#pragma unroll
        for (int k = 0; k < DOF; k++) L[f][k] = 0.0;
    for (int j = 0; j < a.nsrc; j++) {
        const double b = a.bcoef[j] * dtdp;
        if (b == 0.0) continue;
        const double* s = a.src[j];
#pragma unroll
        for (int f = 0; f < 3; f++)
#pragma unroll
            for (int k = 0; k < DOF; k++) L[f][k] = L[f][k] + b * s[(size_t)(f * DOF + k) * ne + el];
    }
}

// Element g's P1 bathymetry modes and AREAS (twice the area), from its vertex depths and
// coordinates, with the products and sums of the reference's set-up (src/read_input.F:1408,
// src/prep_DG.F:959-1163) and of the forward stage kernel.
__device__ __forceinline__ void elem_hb_areas(const ElemData& E, int ne, int g, double hb[3], double& areas) {
#pragma clang fp contract(fast)
    const double c13 = 1.0 / 3.0, c16 = 1.0 / 6.0;
    const double d0 = E.dpv[g], d1 = E.dpv[(size_t)ne + g], d2 = E.dpv[2 * (size_t)ne + g];
    hb[0] = c13 * (d0 + d1 + d2);
    hb[1] = -(c16 * (d0 + d1)) + c13 * d2;
    hb[2] = -(0.5 * d0) + 0.5 * d1;
    const double X0 = E.xv[g], X1 = E.xv[(size_t)ne + g], X2 = E.xv[2 * (size_t)ne + g];
    const double Y0 = E.yv[g], Y1 = E.yv[(size_t)ne + g], Y2 = E.yv[2 * (size_t)ne + g];
    areas = (X0 - X2) * (Y1 - Y2) + (X2 - X1) * (Y0 - Y2);
}

// QF: quadratic bottom friction without the hybrid factor (NOLIBF=1), fixed at compile time.
// WD: the stage's wet/dry flags are read (a.wd) and dry elements and dry-side edge cases are
// reversed; without it every element is wet and that code is compiled out (it would otherwise
// raise the kernel's register use past what the all-wet runs need).
template <int P, int NT, bool SPH, bool NL, bool QF, bool WD>
__global__ void __launch_bounds__(NT, P == 1 ? 2 : 1) adj_stage_kernel(AdjArgs a) {
#pragma clang fp contract(fast)
    constexpr int DOF = (P + 1) * (P + 2) / 2;
    constexpr int NAGP = (P == 1) ? 3 : 6;
    constexpr int NEGP = P + 2;
    __shared__ double ub_lds[3 * DOF][NT];     // cotangent of this element's modes
    __shared__ double tpe[3][NEGP][DOF], teq[3][NEGP][DOF];   // edge tables (lane-divergent index)
    // staged rows (ns entries each): state [3*DOF], RHS cotangent [3*DOF].  The bathymetry modes
    // and AREAS are formed from the vertex depths and coordinates where they are needed: staging
    // them too would take the p=2 block past a quarter of the CU's LDS.
    extern __shared__ double stg[];
    constexpr int RL = 3 * DOF;
    const DevTables& T = tab<P>();
    const ElemData& E = a.ed;
    const Phys& ph = a.ph;
    const int ne = E.ne;
    const int tid = threadIdx.x;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int e = tile * NT + tid;
    const double c13 = 1.0 / 3.0, c16 = 1.0 / 6.0;
    const int ns = a.ns;
    for (int j = tid; j < 3 * NEGP * DOF; j += NT) {
        const int l = j / (NEGP * DOF), g = (j / DOF) % NEGP, k = j % DOF;
        tpe[l][g][k] = T.phi_edge[l][g][k];
        teq[l][g][k] = T.edgeq[l][g][k];
    }
    {
        // this tile's elements and its halo, all loads in flight together
        const int h0 = a.hoff[tile], nh = a.hoff[tile + 1] - h0;
        for (int j = tid; j < NT + nh; j += NT) {
            const int g = j < NT ? tile * NT + j : a.hidx[h0 + j - NT];
            if (g < ne) {
#pragma unroll
                for (int f = 0; f < 3 * DOF; f++) stg[f * ns + j] = a.u[(size_t)f * ne + g];
                double L0[3][DOF];
                const int sg = E.sid ? (int)E.sid[g] : 0;
                load_lambda<DOF>(a, ne, g, E.scen[sg].dtdp, L0);
#pragma unroll
                for (int f = 0; f < 3; f++)
#pragma unroll
                    for (int k = 0; k < DOF; k++) stg[(RL + f * DOF + k) * ns + j] = L0[f][k];
            }
        }
    }
    __syncthreads();
    if (e >= ne) return;
    const int sid = E.sid ? (int)E.sid[e] : 0;
    const Scen sc = E.scen[sid];
    const int row = table_step(a.fo) * a.fo.nrk + a.irk;
    const bool quad = QF || ph.iflinbf == 0;
    const bool hybrid = !QF && ph.ifhybf == 1;

    double U[3][DOF];
#pragma unroll
    for (int f = 0; f < 3; f++)
#pragma unroll
        for (int k = 0; k < DOF; k++) U[f][k] = stg[(f * DOF + k) * ns + tid];
#define Lm(f, k) stg[(RL + (f) * DOF + (k)) * ns + tid]
#pragma unroll
    for (int j = 0; j < 3 * DOF; j++) ub_lds[j][tid] = 0.0;
    double hbb[3] = {0.0, 0.0, 0.0};
    double AREAS, hb[3];
    elem_hb_areas(E, ne, e, hb, AREAS);
    const int info = E.sinfo[e];
    const int nb0 = E.snbr[e], nb1 = E.snbr[(size_t)ne + e], nb2 = E.snbr[2 * (size_t)ne + e];
    const int sl = a.snbl[e];
    const int wd = WD ? (int)a.wd[e] : 1;
#pragma unroll 1
    for (int slot = 0; slot < 3; slot++) {
        const int bits = side_bits(info, slot);
        const int L = bits & 3, role = (bits >> 2) & 1, kind = (bits >> 3) & 3, nled = (bits >> 5) & 3;
        const int nb = slot == 0 ? nb0 : (slot == 1 ? nb1 : nb2);
        int va = (L + 1) % 3, vb = (L + 2) % 3;
        if (role) { const int t = va; va = vb; vb = t; }
        const double dy = E.yv[(size_t)vb * ne + e] - E.yv[(size_t)va * ne + e];
        const double dx = E.xv[(size_t)vb * ne + e] - E.xv[(size_t)va * ne + e];
        const double xlen = sqrt(dy * dy + dx * dx);
        const double nx = dy / xlen, ny = -dx / xlen;
        const double tx = -ny, ty = nx;
        if (kind == SIDE_INTERNAL) {
            const int ln = (sl >> (10 * slot)) & 1023;
            const int wdn = WD ? (int)a.wd[nb] : 1;
            if (wd == 0 && wdn == 0) continue;          // the forward adds no flux between dry elements
            double V[3][DOF], Ln[3][DOF];
#pragma unroll
            for (int f = 0; f < 3; f++)
#pragma unroll
                for (int k = 0; k < DOF; k++) {
                    V[f][k] = stg[(f * DOF + k) * ns + ln];
                    Ln[f][k] = stg[(RL + f * DOF + k) * ns + ln];
                }
            double hbn[3], areas_n;
            elem_hb_areas(E, ne, nb, hbn, areas_n);
            const double edfac = xlen / AREAS, edfac_n = xlen / areas_n;
#pragma unroll 1
            for (int I = 0; I < NEGP; I++) {
                const int gi = I, gx = NEGP - 1 - I;
                const int gown = role ? gx : gi, gnb = role ? gi : gx;
                const double* po = tpe[L][gown];
                const double* pn = tpe[nled][gnb];
                const double zo = trace<DOF>(U[0], po), xo = trace<DOF>(U[1], po), yo = trace<DOF>(U[2], po);
                const double zn = trace<DOF>(V[0], pn), xn = trace<DOF>(V[1], pn), yn = trace<DOF>(V[2], pn);
                const double ho = trace<3>(hb, po), hn = trace<3>(hbn, pn);
                double so = 1.0, sn = 1.0;
                if (SPH) {
                    so = E.sfaced[((size_t)L * NEGP + gown) * ne + e];
                    sn = E.sfaced[((size_t)nled * NEGP + gnb) * ne + nb];
                }
                // flux cotangent: the "in" element's RHS took -w_in*F, the "ex" element's +w_ex*F
                const double* eqo = teq[L][gown];
                const double* eqn = teq[nled][gnb];
                double fb[3], cin[3], cex[3];
#pragma unroll
                for (int f = 0; f < 3; f++) {
                    double so_ = 0.0, sn_ = 0.0;
#pragma unroll
                    for (int k = 0; k < DOF; k++) {
                        so_ = so_ + edfac * eqo[k] * Lm(f, k);
                        sn_ = sn_ + edfac_n * eqn[k] * Ln[f][k];
                    }
                    fb[f] = role ? (so_ - sn_) : (sn_ - so_);
                    cin[f] = role ? -sn_ : -so_;
                    cex[f] = role ? so_ : sn_;
                }
                // canonical (in, ex) order; this element is side `role`
                double zb, xb, yb, hbar;
                if (wd != 0 && wdn != 0) {
                    llf_flux_adj_side<NL>(ph, role ? zn : zo, role ? zo : zn, role ? xn : xo, role ? xo : xn,
                                          role ? yn : yo, role ? yo : yn, role ? hn : ho, role ? ho : hn,
                                          role ? sn : so, role ? so : sn, nx, ny, ph.g, fb, role, zb, xb, yb, hbar);
                } else {
                    dry_side_adj<P, NL>(ph, E, ne, e, nb, L, role, nled, gi, gx, U, V, wd, wdn,
                                        role ? zn : zo, role ? zo : zn, role ? xn : xo, role ? xo : xn,
                                        role ? yn : yo, role ? yo : yn, role ? hn : ho, role ? ho : hn,
                                        role ? sn : so, role ? so : sn, nx, ny, cin, cex, zb, xb, yb, hbar);
                }
#pragma unroll
                for (int k = 0; k < DOF; k++) {
                    ub_lds[k][tid] += zb * po[k];
                    ub_lds[DOF + k][tid] += xb * po[k];
                    ub_lds[2 * DOF + k][tid] += yb * po[k];
                }
#pragma unroll
                for (int k = 0; k < 3; k++) hbb[k] += hbar * po[k];
            }
        } else {
            if (wd == 0) continue;                      // the forward skips a dry element's boundary edges
#pragma unroll 1
            for (int I = 0; I < NEGP; I++) {
                const double* pv = tpe[L][I];
                const double ze_in = trace<DOF>(U[0], pv), qx_in = trace<DOF>(U[1], pv), qy_in = trace<DOF>(U[2], pv);
                const double hb_in = trace<3>(hb, pv);
                const double sf = SPH ? E.sfaced[((size_t)L * NEGP + I) * ne + e] : 1.0;
                // flux cotangent: RHS(K) -= w_K*F with w_K = 2 M_INV(K)/AREAS*XLEN*PHI*WEGP
                double fb[3] = {0.0, 0.0, 0.0};
                const double wq = 2.0 * xlen / AREAS * T.wegp[I];
#pragma unroll
                for (int k = 0; k < DOF; k++) {
                    const double w = wq * T.m_inv[k] * pv[k];
                    fb[0] = fb[0] - w * Lm(0, k);
                    fb[1] = fb[1] - w * Lm(1, k);
                    fb[2] = fb[2] - w * Lm(2, k);
                }
                double bar[8];
                double zb = 0.0, xb = 0.0, yb = 0.0, hbar = 0.0;
                if (kind == SIDE_LAND) {
                    const double qn_int = qx_in * nx + qy_in * ny, qt_int = qx_in * tx + qy_in * ty;
                    const double qn_ext = -qn_int, qt_ext = qt_int;
                    const double den = 1.0 / (nx * ty - ny * tx);
                    const double qx_ex = (ty * qn_ext - ny * qt_ext) * den;
                    const double qy_ex = (-tx * qn_ext + nx * qt_ext) * den;
                    llf_flux_adj<NL>(ph, ze_in, ze_in, qx_in, qx_ex, qy_in, qy_ex, hb_in, hb_in, sf, sf, nx, ny, ph.g, fb, bar);
                    zb = bar[0] + bar[1];
                    hbar = bar[6] + bar[7];
                    // mirror: (qx_ex, qy_ex) from (qx_in, qy_in)
                    const double qnb = (ty * bar[3] - tx * bar[5]) * den;
                    const double qtb = (-ny * bar[3] + nx * bar[5]) * den;
                    const double qnib = -qnb, qtib = qtb;
                    xb = bar[2] + nx * qnib + tx * qtib;
                    yb = bar[4] + ny * qnib + ty * qtib;
                } else if (kind == SIDE_OCEAN) {
                    double ze_ex = 0.0;
                    const double xe = T.xegp[I];
                    // the constituents' endpoint terms EMO*RFF*COS(ARGJ-EFA) of this stage (tide_kernel)
                    for (int jj = 0; jj < a.fo.nbfr; jj++) {
                        const size_t o = ((size_t)nb * a.fo.nbfr + jj) * 2;
                        const double z1 = a.fo.tz[(size_t)a.irk * a.fo.tzn + o];
                        const double z2 = a.fo.tz[(size_t)a.irk * a.fo.tzn + o + 1];
                        ze_ex = ze_ex + (0.5 * (z1 + z2) + 0.5 * (z2 - z1) * xe);
                    }
                    const bool eirik = (ze_ex * ph.ifnlfa + hb_in) <= 0;
                    if (eirik) ze_ex = fabs(hb_in) + sc.h0;
                    llf_flux_adj<NL>(ph, ze_in, ze_ex, qx_in, qx_in, qy_in, qy_in, hb_in, hb_in, sf, sf, nx, ny, ph.g, fb, bar);
                    zb = bar[0];
                    xb = bar[2] + bar[3];
                    yb = bar[4] + bar[5];
                    hbar = bar[6] + bar[7];
                    if (eirik) hbar += (hb_in > 0.0 ? 1.0 : (hb_in < 0.0 ? -1.0 : 0.0)) * bar[1];
                } else {
                    // the step's specified normal flux at this point (flow_kernel)
                    const double qn_ext = a.fo.fz[(size_t)a.irk * a.fo.fzn + (size_t)nb * NEGP + I];
                    const double qt_ext = 0.0;
                    const double qx_ex = -(ty * qn_ext - ny * qt_ext) / (nx * ty - ny * tx);
                    const double qy_ex = -(-tx * qn_ext + nx * qt_ext) / (nx * ty - ny * tx);
                    llf_flux_adj<NL>(ph, ze_in, ze_in, qx_in, qx_ex, qy_in, qy_ex, hb_in, hb_in, sf, sf, nx, ny, ph.g, fb, bar);
                    zb = bar[0] + bar[1];
                    xb = bar[2];
                    yb = bar[4];
                    hbar = bar[6] + bar[7];
                }
#pragma unroll
                for (int k = 0; k < DOF; k++) {
                    ub_lds[k][tid] += zb * pv[k];
                    ub_lds[DOF + k][tid] += xb * pv[k];
                    ub_lds[2 * DOF + k][tid] += yb * pv[k];
                }
#pragma unroll
                for (int k = 0; k < 3; k++) hbb[k] += hbar * pv[k];
            }
        }
    }

    // ---- volume quadrature (wet elements only, as in the forward)
    double nbarb = 0.0;
    if (wd != 0) {
        double drdx, dsdx, drdy, dsdy;
        {
            const double X0 = E.xv[e], X1 = E.xv[(size_t)ne + e], X2 = E.xv[2 * (size_t)ne + e];
            const double Y0 = E.yv[e], Y1 = E.yv[(size_t)ne + e], Y2 = E.yv[2 * (size_t)ne + e];
            double AREA = (X0 - X2) * (Y1 - Y2) + (X2 - X1) * (Y0 - Y2);
            AREA = AREA * 0.5;
            drdx = 1.0 / AREA * (Y2 - Y0);
            dsdx = 1.0 / AREA * (Y0 - Y1);
            drdy = 1.0 / AREA * (X0 - X2);
            dsdy = 1.0 / AREA * (X1 - X0);
        }
        const double depth_c = hb[0] + U[0][0];
        const double fh_nl_c = NL ? 1.0 / depth_c : 1.0 / (ph.nleq * depth_c + ph.leq);
        const double cori = E.cori_el ? E.cori_el[e] : sc.cori;
        const double nbar = ph.load_manning ? E.nbar[e] : 0.0;
        double depth_cb = 0.0;
        // NWS=8: the stage's vertex wind stresses and the pressure-gradient sums, as the forward
        // forms them; the pressure term -g*SFAC*H*dP/dx depends on the state through H
        double mw[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0}, pdx = 0.0, pdy = 0.0;
        if (a.met) {
            const double ramp = a.fo.ramp[(size_t)sid * a.fo.rows + row];
            const double drpsi[3] = {-1.0 / 2.0, 1.0 / 2.0, 0.0}, dspsi[3] = {-1.0 / 2.0, 0.0, 1.0 / 2.0};
#pragma unroll
            for (int v = 0; v < 3; v++) {
                const int n = E.nm[(size_t)v * ne + e];
                const double cd = a.mt.cd[n], wm = a.mt.wm[n];
                mw[v] = ramp * 0.001293 * cd * a.mt.wx[n] * wm;
                mw[3 + v] = ramp * 0.001293 * cd * a.mt.wy[n] * wm;
                const double pr = ramp * a.mt.prn[n];
                pdx += pr * (drpsi[v] * drdx + dspsi[v] * dsdx);
                pdy += pr * (drpsi[v] * drdy + dspsi[v] * dsdy);
            }
        }
#pragma unroll 1
        for (int I = 0; I < NAGP; I++) {
            const double* pa = T.phi_area[I];
            double ze = U[0][0], qx = U[1][0], qy = U[2][0];
#pragma unroll
            for (int k = 1; k < DOF; k++) {
                ze = ze + U[0][k] * pa[k];
                qx = qx + U[1][k] * pa[k];
                qy = qy + U[2][k] * pa[k];
            }
            double hbq = 0.0, dhbx = 0.0, dhby = 0.0;
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const double dpxk = T.drphi[I][k] * drdx + T.dsphi[I][k] * dsdx;
                const double dpyk = T.drphi[I][k] * drdy + T.dsphi[I][k] * dsdy;
                hbq = hbq + hb[k] * pa[k];
                dhbx = dhbx + hb[k] * dpxk;
                dhby = dhby + hb[k] * dpyk;
            }
            const double sfq = SPH ? E.sfacq[(size_t)I * ne + e] : 1.0;
            const double depth = ze + hbq;
            const double f1 = NL ? 1.0 : ph.nleq + ph.leq * hbq;
            const double fu = NL ? qx : ph.nleq * qx, fv = NL ? qy : ph.nleq * qy;
            const double fg = ph.nleqg * ze;
            const double fh = NL ? 1.0 / depth : 1.0 / (ph.nleq * depth + ph.leq);
            const double u = qx * fh, v = qy * fh;
            double fric = sc.fric;
            bool fric_active = false;
            double pw = 0.0;
            if (ph.load_manning) {
                pw = crm::cr_pow_third(depth);
                fric = ph.g * (nbar * nbar) / pw;
                if (fric < sc.cf) fric = sc.cf;
                else fric_active = true;
            }
            double tau = fric, umag = 0.0, hyb = 1.0, hyb_in = 0.0;
            bool capped = false;
            if (quad) {
                umag = sqrt(u * u + v * v);
                tau = tau * umag * fh;
                if (hybrid) {
                    hyb_in = 1.0 + pow(ph.hbreak * fh_nl_c, ph.ftheta);
                    hyb = pow(hyb_in, ph.fgamma / ph.ftheta);
                    tau = tau * hyb;
                }
                const double cap = 0.9 * sc.dtdph;
                if (cap < tau) { tau = cap; capped = true; }
            }
            // cotangents of the quadrature-point fluxes and sources from RHS(K)
            double fxb = 0.0, fyb = 0.0, gxb = 0.0, gyb = 0.0, hxb = 0.0, hyb_ = 0.0;
            double sxb = T.srfac[I][0] * Lm(1, 0), syb = T.srfac[I][0] * Lm(2, 0);
#pragma unroll
            for (int k = 1; k < DOF; k++) {
                const double dpxk = T.drphi[I][k] * drdx + T.dsphi[I][k] * dsdx;
                const double dpyk = T.drphi[I][k] * drdy + T.dsphi[I][k] * dsdy;
                const double xf = T.mw[I][k] * dpxk, yf = T.mw[I][k] * dpyk, sr = T.srfac[I][k];
                fxb += xf * Lm(0, k); fyb += yf * Lm(0, k);
                gxb += xf * Lm(1, k); gyb += yf * Lm(1, k); sxb += sr * Lm(1, k);
                hxb += xf * Lm(2, k); hyb_ += yf * Lm(2, k); syb += sr * Lm(2, k);
            }
            double zeb = 0.0, qxb = 0.0, qyb = 0.0, hbqb = 0.0, dhbxb = 0.0, dhbyb = 0.0, depthb = 0.0;
            // sx = -(tau*qx) + fg*dhbx*sfq + cori*qy ; sy = -(tau*qy) + fg*dhby - cori*qx
            double taub = -qx * sxb - qy * syb;
            double f1b_met = 0.0;
            if (a.met) {
                // sx += wsx/F1 - g*SFAC*H*pdx ; sy += wsy/F1 - g*H*pdy (F1 = 1 for the nonlinear equations)
                depthb += -ph.g * sfq * pdx * sxb - ph.g * pdy * syb;
                if constexpr (!NL) {
                    const double* ps = T.psi[I];
                    const double wsx = mw[0] * ps[0] + mw[1] * ps[1] + mw[2] * ps[2];
                    const double wsy = mw[3] * ps[0] + mw[4] * ps[1] + mw[5] * ps[2];
                    const double fw = 1.0 / f1;
                    f1b_met = -fw * fw * (wsx * sxb + wsy * syb);
                }
            }
            qxb += -tau * sxb - cori * syb;
            qyb += cori * sxb - tau * syb;
            double fgb = dhbx * sfq * sxb + dhby * syb;
            dhbxb += fg * sfq * sxb;
            dhbyb += fg * syb;
            // friction
            double fricb = 0.0, ub = 0.0, vb = 0.0, fhb = 0.0;
            if (quad) {
                if (!capped) {
                    // tau = fric*umag*fh*hyb
                    fricb += umag * fh * hyb * taub;
                    const double umagb = fric * fh * hyb * taub;
                    fhb += fric * umag * hyb * taub;
                    if (hybrid) {
                        const double hybb = fric * umag * fh * taub;
                        // hyb = (1 + (hbreak*fh_nl_c)^ftheta)^(fgamma/ftheta)
                        const double dh_in = (ph.fgamma / ph.ftheta) * pow(hyb_in, ph.fgamma / ph.ftheta - 1.0);
                        const double x = ph.hbreak * fh_nl_c;
                        const double dx_ = x > 0.0 ? ph.ftheta * pow(x, ph.ftheta - 1.0) : 0.0;
                        const double fhcb = hybb * dh_in * dx_ * ph.hbreak;
                        depth_cb += -ph.nleq * fh_nl_c * fh_nl_c * fhcb;
                    }
                    if (umag > 0.0) {
                        const double r = umagb / umag;
                        ub += u * r; vb += v * r;
                    }
                }
            } else {
                fricb += taub;
            }
            if (ph.load_manning && fric_active) {
                // fric = g*nbar^2/depth^(1/3): d/dnbar = 2*fric/nbar, d/ddepth = -fric/(3*depth)
                const double gp = ph.g / pw;
                nbarb += 2.0 * gp * nbar * fricb;
                depthb += -(1.0 / 3.0) * gp * (nbar * nbar) * (NL ? fh : 1.0 / depth) * fricb;
            }
            // hy = hvv + gh2 ; hx = huv*sfq ; gy = huv ; gx = (huu + gh2)*sfq ; fx = qx*f1*sfq ; fy = qy*f1
            const double hvvb = hyb_;
            const double gh2b = hyb_ + sfq * gxb;
            const double huvb = sfq * hxb + gyb;
            const double huub = sfq * gxb;
            fgb += (0.5 * ze + hbq) * gh2b;
            zeb += (0.5 * fg + ph.fg_l) * gh2b;
            hbqb += fg * gh2b;
            const double fub = u * huub + v * huvb;
            ub += fu * huub;
            const double fvb = v * hvvb;
            vb += fv * hvvb + fu * huvb;
            qxb += fh * ub; qyb += fh * vb;
            fhb += qx * ub + qy * vb;
            depthb += -ph.nleq * fh * fh * fhb;
            zeb += ph.nleqg * fgb;
            qxb += ph.nleq * fub;
            qyb += ph.nleq * fvb;
            qxb += f1 * sfq * fxb;
            qyb += f1 * fyb;
            const double f1b = qx * sfq * fxb + qy * fyb + f1b_met;
            hbqb += ph.leq * f1b;
            zeb += depthb;
            hbqb += depthb;
            // traces
            ub_lds[0][tid] += zeb;
            ub_lds[DOF][tid] += qxb;
            ub_lds[2 * DOF][tid] += qyb;
#pragma unroll
            for (int k = 1; k < DOF; k++) {
                ub_lds[k][tid] += zeb * pa[k];
                ub_lds[DOF + k][tid] += qxb * pa[k];
                ub_lds[2 * DOF + k][tid] += qyb * pa[k];
            }
#pragma unroll
            for (int k = 0; k < 3; k++) {
                const double dpxk = T.drphi[I][k] * drdx + T.dsphi[I][k] * dsdx;
                const double dpyk = T.drphi[I][k] * drdy + T.dsphi[I][k] * dsdy;
                hbb[k] += hbqb * pa[k] + dhbxb * dpxk + dhbyb * dpyk;
            }
        }
        // depth_c = hb[0] + U[0][0]
        hbb[0] += depth_cb;
        ub_lds[0][tid] += depth_cb;
    }
    // ---- cotangent of U(IRK): later stages' ATVD terms plus this stage's RHS reverse
    for (int j = 0; j < a.nsrc; j++) {
        const double ac = a.acoef[j];
        if (ac == 0.0) continue;
        const double* s = a.src[j];
#pragma unroll
        for (int f = 0; f < 3 * DOF; f++) ub_lds[f][tid] += ac * s[(size_t)f * ne + e];
    }
#pragma unroll
    for (int f = 0; f < 3 * DOF; f++) a.out[(size_t)f * ne + e] = ub_lds[f][tid];
    // parameters: element-mean Manning n and the vertex depths through the P1 bathymetry modes
    if (a.gnbar) a.gnbar[e] += nbarb;
    if (a.gdp) {
        const double D0b = c13 * hbb[0] - c16 * hbb[1] - 0.5 * hbb[2];
        const double D1b = c13 * hbb[0] - c16 * hbb[1] + 0.5 * hbb[2];
        const double D2b = c13 * hbb[0] + c13 * hbb[1];
        a.gdp[e] += D0b;
        a.gdp[(size_t)ne + e] += D1b;
        a.gdp[2 * (size_t)ne + e] += D2b;
    }
#undef Lm
}

// ------------------------------------------------------------------------------------------
// WETDRY (src/wetdry.F:58-548) of element e reversed with its branches frozen: from the state it
// was applied to (U: the limited stage output's modes 1-3, formed by limit_wetdry_kernel's adjoint
// variant), the vertex depths D and the element's flag before it, the element retakes the
// forward's decisions operation by operation (no contraction, so that they are the forward's) --
// which vertices are dry, the branch, the vertices' depth order, the two fmax(H0, .) choices, the
// vertices dry after the redistribution and the check -- and applies the transpose of the linear
// map that branch is.  ub (in place): cotangent of WETDRY's output -> of its input; gdp += the
// cotangents of the vertex depths.
// ------------------------------------------------------------------------------------------
template <int P>
__device__ void wetdry_adj_element(const DevTables& T, const double (&U)[3][3], const double (&D)[3], int wdflg,
                                   double H0, int e, int ne, double* __restrict__ ub, double* __restrict__ gdp) {
    constexpr int DOF = (P + 1) * (P + 2) / 2;
    const double ZERO = 1.0e-6, C13 = 1.0 / 3.0, C16 = 1.0 / 6.0;
    const double dpe_min = fmin(fmin(D[0], D[1]), D[2]);
    // the forward's vertex values, in its operation order
    int ndry = 0, ht_max_i = 0;
    double ht_max = ZERO, ze_node[3], ht_node[3];
#pragma unroll
    for (int i = 0; i < 3; i++) {
        double z = U[0][0];
#pragma unroll
        for (int k = 1; k < 3; k++) z = z + T.phi_corner[i][k] * U[0][k];
        ze_node[i] = z;
        ht_node[i] = D[i] + z;
        if (ht_node[i] > ht_max) { ht_max_i = i; ht_max = ht_node[i]; }
        if (ht_node[i] <= (H0 + ZERO)) ndry++;
    }
    if (ndry == 0 && wdflg == 1) return;                       // identity
    const double dp_avg = C13 * (D[0] + D[1] + D[2]);
    double G[3][DOF];                                           // output cotangent
#pragma unroll
    for (int f = 0; f < 3; f++)
#pragma unroll
        for (int k = 0; k < DOF; k++) G[f][k] = ub[(size_t)(f * DOF + k) * ne + e];
    double I[3][DOF];                                           // input cotangent
#pragma unroll
    for (int f = 0; f < 3; f++)
#pragma unroll
        for (int k = 0; k < DOF; k++) I[f][k] = 0.0;
    double Db[3] = {0.0, 0.0, 0.0};
    if (ndry == 0) {
        // the check alone: the element keeps its values, its momentum too when the highest
        // vertex is wet enough
        const bool keep = ze_node[ht_max_i] > (-dpe_min + H0);
#pragma unroll
        for (int k = 0; k < DOF; k++) {
            I[0][k] = G[0][k];
            I[1][k] = keep ? G[1][k] : 0.0;
            I[2][k] = keep ? G[2][k] : 0.0;
        }
    } else {
        const double ht_avg = U[0][0] + dp_avg;
        if (ndry == 3 || ht_avg <= (H0 + ZERO)) {
            // flat: ZE = (ht_avg - dp_avg | ZERO - dp_avg, -HB2, -HB3), momentum zero
            if (ht_avg < ZERO) {
#pragma unroll
                for (int i = 0; i < 3; i++) Db[i] -= C13 * G[0][0];
            } else {
                I[0][0] = G[0][0];
            }
            // -HB2 = C16*(D1+D2) - C13*D3, -HB3 = 0.5*D1 - 0.5*D2
            Db[0] += C16 * G[0][1] + 0.5 * G[0][2];
            Db[1] += C16 * G[0][1] - 0.5 * G[0][2];
            Db[2] += -C13 * G[0][1];
        } else {
            // redistribution, as the forward orders the vertices and takes its branches
            int l2h[3] = {0, 1, 2};
#pragma unroll
            for (int k1 = 0; k1 < 3; k1++)
#pragma unroll
                for (int k2 = k1; k2 < 3; k2++)
                    if (ht_node[l2h[k1]] > ht_node[l2h[k2]]) { const int t = l2h[k1]; l2h[k1] = l2h[k2]; l2h[k2] = t; }
            double dh[3];
            const bool b0 = H0 > ht_node[l2h[0]];
            dh[l2h[0]] = fmax(H0, ht_node[l2h[0]]) - ht_node[l2h[0]];
            const bool b1 = H0 > ht_node[l2h[1]] - 0.5 * dh[l2h[0]];
            dh[l2h[1]] = fmax(H0, ht_node[l2h[1]] - 0.5 * dh[l2h[0]]) - ht_node[l2h[1]];
            dh[l2h[2]] = -dh[l2h[0]] - dh[l2h[1]];
            double ht_hat[3];
            int nd = 0, dry[3];
            ht_max = 0;
#pragma unroll
            for (int i = 0; i < 3; i++) {
                ht_hat[i] = ht_node[i] + dh[i];
                if (ht_hat[i] > ht_max) { ht_max_i = i; ht_max = ht_hat[i]; }
                dry[i] = ht_hat[i] <= (H0 + 1.0e-10);
                nd += dry[i];
            }
            double ze_hat[3];
#pragma unroll
            for (int i = 0; i < 3; i++) ze_hat[i] = ht_hat[i] - D[i];
            const bool keep = ze_hat[ht_max_i] > (-dpe_min + H0);
            // momentum: QX(1..3) from the vertex values qxh, qxh = qxn + qxs at wet vertices,
            // qxs = (sum of the dry vertices' qxn)/(3 - nd); zero when nd = 3 or the check fails
            if (keep && nd < 3) {
#pragma unroll
                for (int f = 1; f < 3; f++) {
                    double hb_[3];
#pragma unroll
                    for (int i = 0; i < 3; i++)
                        hb_[i] = C13 * G[f][0] + (i < 2 ? -C16 : C13) * G[f][1] + (i == 0 ? -0.5 : (i == 1 ? 0.5 : 0.0)) * G[f][2];
                    double sb = 0.0, nb_[3];
#pragma unroll
                    for (int i = 0; i < 3; i++) { nb_[i] = dry[i] ? 0.0 : hb_[i]; if (!dry[i]) sb += hb_[i]; }
                    const double w = sb / (3.0 - (double)nd);
#pragma unroll
                    for (int i = 0; i < 3; i++) if (dry[i]) nb_[i] += w;
                    // qxn_i = QX1 + PHI(i,2)*QX2 + PHI(i,3)*QX3
#pragma unroll
                    for (int i = 0; i < 3; i++) {
                        I[f][0] += nb_[i];
                        I[f][1] += T.phi_corner[i][1] * nb_[i];
                        I[f][2] += T.phi_corner[i][2] * nb_[i];
                    }
                }
            }
            // elevation: ZE1 kept; ZE(2..3) from ze_hat = ht_hat - D = ze_node + dh
            I[0][0] += G[0][0];
            double zb[3], dhb[3], htb[3] = {0.0, 0.0, 0.0};
#pragma unroll
            for (int i = 0; i < 3; i++) {
                zb[i] = (i < 2 ? -C16 : C13) * G[0][1] + (i == 0 ? -0.5 : (i == 1 ? 0.5 : 0.0)) * G[0][2];
                dhb[i] = zb[i];
            }
            // dh(l2) = -dh(l0) - dh(l1)
            dhb[l2h[0]] -= dhb[l2h[2]];
            dhb[l2h[1]] -= dhb[l2h[2]];
            // dh(l1) = H0 - ht(l1), or (ht(l1) - 0.5*dh(l0)) - ht(l1) = -0.5*dh(l0)
            if (b1) htb[l2h[1]] -= dhb[l2h[1]];
            else dhb[l2h[0]] += -0.5 * dhb[l2h[1]];
            // dh(l0) = H0 - ht(l0), or zero
            if (b0) htb[l2h[0]] -= dhb[l2h[0]];
            // ht = D + ze_node
#pragma unroll
            for (int i = 0; i < 3; i++) {
                Db[i] += htb[i];
                zb[i] += htb[i];
                I[0][0] += zb[i];
                I[0][1] += T.phi_corner[i][1] * zb[i];
                I[0][2] += T.phi_corner[i][2] * zb[i];
            }
        }
    }
#pragma unroll
    for (int f = 0; f < 3; f++)
#pragma unroll
        for (int k = 0; k < DOF; k++) ub[(size_t)(f * DOF + k) * ne + e] = I[f][k];
    gdp[e] += Db[0];
    gdp[(size_t)ne + e] += Db[1];
    gdp[2 * (size_t)ne + e] += Db[2];
}

// Elevation at the gauges: eta_g = sum_K ZE(K, elem_g) * phi_K(gauge), written to out[g].
template <int P>
__global__ void gauge_kernel(int ng, int ne, const int* __restrict__ gel, const double* __restrict__ gphi,
                             const double* __restrict__ u, double* __restrict__ out) {
    constexpr int DOF = (P + 1) * (P + 2) / 2;
    const int g = blockIdx.x * blockDim.x + threadIdx.x;
    if (g >= ng) return;
    const int e = gel[g];
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < DOF; k++) s = s + u[(size_t)k * ne + e] * gphi[(size_t)g * DOF + k];
    out[g] = s;
}

// Cotangent seed of the gauge misfit sum_g w (eta_g - d_g)^2 into lam.  One thread per gauge
// element; gord lists the gauges grouped by element (ascending gauge index within a group, group u
// at gord[goff[u]..goff[u+1])), so gauges sharing an element add up in gauge order.
template <int P>
__global__ void __launch_bounds__(64) gauge_seed_kernel(int nu, int ne, const int* __restrict__ goff,
                                                        const int* __restrict__ gord, const int* __restrict__ gel,
                                                        const double* __restrict__ gphi,
                                                        const double* __restrict__ eta,
                                                        const double* __restrict__ data, double w,
                                                        double* __restrict__ lam) {
    constexpr int DOF = (P + 1) * (P + 2) / 2;
    const int u = blockIdx.x * blockDim.x + threadIdx.x;
    if (u >= nu) return;
    const int j0 = goff[u], j1 = goff[u + 1];
    const int e = gel[gord[j0]];
    double t[DOF];
#pragma unroll
    for (int k = 0; k < DOF; k++) t[k] = lam[(size_t)k * ne + e];
    for (int j = j0; j < j1; j++) {
        const int g = gord[j];
        const double r = 2.0 * w * (eta[g] - data[g]);
#pragma unroll
        for (int k = 0; k < DOF; k++) t[k] += r * gphi[(size_t)g * DOF + k];
    }
#pragma unroll
    for (int k = 0; k < DOF; k++) lam[(size_t)k * ne + e] = t[k];
}

// ------------------------------------------------------------------------------------------
// Reverse of SLOPELIMITER5 (src/slopelimiter.F:3434-3676) with its active set frozen: the
// forward limiter of each element is re-run on the stage's pre-limiter state W to record which
// branch every clamp, every redistribution test and every min(redfac, redmax) took; given those,
// the limiter is linear in (the element's linear modes, the node bounds) and its transpose maps
// the cotangent of the limited modes (in ub, in place) to the cotangent of W's modes, plus the
// cotangents of the vertex bounds (nbc: a record of six doubles [field][min|max] per element and
// vertex, record 3e+v, so that a node reads each incident element's six values from one place).
// Modes >= 3 pass through.
// ------------------------------------------------------------------------------------------
template <int P, int NB = 256>
__global__ void __launch_bounds__(NB) limiter_adj_kernel(int np, ElemData E, const double* __restrict__ W,
                                                          const double* __restrict__ mm, double* __restrict__ ub,
                                                          double* __restrict__ nbc) {
    // no contraction: the recorded branches must be those of the forward limiter, operation by
    // operation
    constexpr int DOF = (P + 1) * (P + 2) / 2;
    const DevTables& T = tab<P>();
    const int ne = E.ne;
    const int tid = threadIdx.x, e0 = blockIdx.x * NB;
    const int e = e0 + tid;
    // the block's bound-cotangent records, written out together after the element loop: the
    // block's elements hold one contiguous range of records
    __shared__ double rec[NB * 18];
    if (e < ne) {
        const int nn[3] = {E.nm[e], E.nm[(size_t)ne + e], E.nm[2 * (size_t)ne + e]};
        // every load of the three fields first (node bounds, linear modes, their cotangents), so
        // that their latencies overlap
        double bmn[3][3], bmx[3][3], wm[3][3], wb[3][3];
#pragma unroll
        for (int iv = 0; iv < 3; iv++) {
            const double* mn = mm + (size_t)(2 * iv) * np;
            const double* mx = mm + (size_t)(2 * iv + 1) * np;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                bmn[iv][c] = mn[nn[c]];
                bmx[iv][c] = mx[nn[c]];
                wm[iv][c] = W[(size_t)(iv * DOF + c) * ne + e];
                wb[iv][c] = ub[(size_t)(iv * DOF + c) * ne + e];
            }
        }
#pragma unroll
        for (int iv = 0; iv < 3; iv++) {
            const double vmin[3] = {bmn[iv][0], bmn[iv][1], bmn[iv][2]}, vmax[3] = {bmx[iv][0], bmx[iv][1], bmx[iv][2]};
            const double c0 = wm[iv][0], w1 = wm[iv][1], w2 = wm[iv][2];
            // forward, recording the branches
            int sel[3];
            double v[3];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                double t = c0;
                t = t + T.phi_corner[c][1] * w1;
                t = t + T.phi_corner[c][2] * w2;
                double x = t;
                sel[c] = 0;
                if (!(t <= vmax[c])) { x = vmax[c]; sel[c] = 1; }
                if (x < vmin[c]) { x = vmin[c]; sel[c] = 2; }
                v[c] = x;
            }
            double sgs[3], divs[3][3];
            int pos = 0, usemax = 0;
#pragma unroll
            for (int ll = 0; ll < 3; ll++) {
                const double sumloc = (v[0] + v[1] + v[2]) / 3.0;
                double sumdif = (sumloc - c0) * 3.0;
                const double sg = copysign(1.0, sumdif);
                sgs[ll] = sg;
                double dif[3];
#pragma unroll
                for (int c = 0; c < 3; c++) dif[c] = (v[c] - c0) * sg;
                int kdp = (dif[0] > 0) + (dif[1] > 0) + (dif[2] > 0);
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    const double div = fmax(1.0, (double)kdp);
                    divs[ll][k] = div;
                    double redfac = 0.0;
                    if (dif[k] > 0) { redfac = sumdif * sg / div; kdp = kdp - 1; pos |= 1 << (3 * ll + k); }
                    const double redmax = (sg > 0) ? v[k] - vmin[k] : vmax[k] - v[k];
                    if (!(redfac <= redmax)) { redfac = redmax; usemax |= 1 << (3 * ll + k); }
                    sumdif = sumdif - redfac * sg;
                    v[k] = v[k] - redfac * sg;
                }
            }
            // reverse: U1 = -(v0+v1)/6 + v2/3, U2 = (v1-v0)/2, U0 = c0
            const double u1b = wb[iv][1], u2b = wb[iv][2];
            double vb[3] = {-(1.0 / 6.0) * u1b - 0.5 * u2b, -(1.0 / 6.0) * u1b + 0.5 * u2b, (1.0 / 3.0) * u1b};
            double c0b = wb[iv][0];
            double vminb[3] = {0.0, 0.0, 0.0}, vmaxb[3] = {0.0, 0.0, 0.0};
#pragma unroll
            for (int ll = 2; ll >= 0; ll--) {
                const double sg = sgs[ll];
                double sb = 0.0;
#pragma unroll
                for (int k = 2; k >= 0; k--) {
                    // v_k -= redfac*sg; sumdif -= redfac*sg
                    const double rb = -sg * vb[k] - sg * sb;
                    if (usemax & (1 << (3 * ll + k))) {
                        // redfac = redmax = sg > 0 ? v_k - vmin_k : vmax_k - v_k
                        if (sg > 0) { vb[k] += rb; vminb[k] -= rb; }
                        else { vmaxb[k] += rb; vb[k] -= rb; }
                    } else if (pos & (1 << (3 * ll + k))) {
                        sb += rb * sg / divs[ll][k];      // redfac = sumdif*sg/div
                    }
                }
                // sumdif = ((v0+v1+v2)/3 - c0)*3
                vb[0] += sb; vb[1] += sb; vb[2] += sb;
                c0b += -3.0 * sb;
            }
            double w1b = 0.0, w2b = 0.0;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                if (sel[c] == 0) {
                    c0b += vb[c];
                    w1b += T.phi_corner[c][1] * vb[c];
                    w2b += T.phi_corner[c][2] * vb[c];
                } else if (sel[c] == 1) {
                    vmaxb[c] += vb[c];
                } else {
                    vminb[c] += vb[c];
                }
            }
            ub[(size_t)(iv * DOF) * ne + e] = c0b;
            ub[(size_t)(iv * DOF + 1) * ne + e] = w1b;
            ub[(size_t)(iv * DOF + 2) * ne + e] = w2b;
#pragma unroll
            for (int c = 0; c < 3; c++) {
                rec[tid * 18 + c * 6 + iv * 2] = vminb[c];
                rec[tid * 18 + c * 6 + iv * 2 + 1] = vmaxb[c];
            }
        }
    }
    __syncthreads();
    const int nvalid = min(NB, ne - e0) * 18;
    double* dst = nbc + (size_t)e0 * 18;
    for (int k = tid; k < nvalid; k += NB) dst[k] = rec[k];
}

// Per node: the sums of the incident elements' vertex-bound cotangents, and the element whose
// cell mean attains each bound (the first in the node's element list on ties; -1 when the
// +-99999 starting bound stands).  gsum, arg: a record [field][min|max] of six per node.  The three
// fields go through the node's list together: every incident element's cell means and its bound
// cotangent record (nbc) are loaded in one pass, and each field's sums still run in list order.
template <int P>
__global__ void __launch_bounds__(256) limiter_adj_node_kernel(int np, int ne, const int* __restrict__ off,
                                                               const int* __restrict__ n2e,
                                                               const double* __restrict__ W,
                                                               const double* __restrict__ nbc,
                                                               double* __restrict__ gsum, int* __restrict__ arg) {
    constexpr int DOF = (P + 1) * (P + 2) / 2;
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= np) return;
    double lo[3], hi[3], slo[3], shi[3];
    int alo[3], ahi[3];
#pragma unroll
    for (int f = 0; f < 3; f++) {
        lo[f] = 99999.; hi[f] = -99999.; slo[f] = 0.0; shi[f] = 0.0;
        alo[f] = -1; ahi[f] = -1;
    }
    const int j1 = off[n + 1];
    for (int j = off[n]; j < j1; j++) {
        const int ent = n2e[j];
        const int e = ent >> 2, v = ent & 3;
        const double2* rec = reinterpret_cast<const double2*>(nbc + ((size_t)e * 3 + v) * 6);
        double c[3];
        double2 b[3];
#pragma unroll
        for (int f = 0; f < 3; f++) {
            c[f] = W[(size_t)(f * DOF) * ne + e];
            b[f] = rec[f];
        }
#pragma unroll
        for (int f = 0; f < 3; f++) {
            if (c[f] < lo[f]) { lo[f] = c[f]; alo[f] = e; }
            if (c[f] > hi[f]) { hi[f] = c[f]; ahi[f] = e; }
            slo[f] += b[f].x;
            shi[f] += b[f].y;
        }
    }
    double2* g = reinterpret_cast<double2*>(gsum + (size_t)n * 6);
    int2* a = reinterpret_cast<int2*>(arg + (size_t)n * 6);
#pragma unroll
    for (int f = 0; f < 3; f++) {
        g[f] = make_double2(slo[f], shi[f]);
        a[f] = make_int2(alo[f], ahi[f]);
    }
}

// Per element: the bound cotangents of the nodes where this element's cell mean is the bound.
template <int P>
__global__ void __launch_bounds__(256) limiter_adj_gather_kernel(int np, ElemData E, const double* __restrict__ gsum,
                                                                 const int* __restrict__ arg, double* __restrict__ ub) {
    constexpr int DOF = (P + 1) * (P + 2) / 2;
    const int ne = E.ne;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    // the node indices, then every node's bound owners and cotangent sums together (a sum that is
    // not this element's adds zero, which leaves the running sum as it is)
    int2 a[3][3];
    double2 g[3][3];
#pragma unroll
    for (int v = 0; v < 3; v++) {
        const int n = E.nm[(size_t)v * ne + e];
#pragma unroll
        for (int f = 0; f < 3; f++) {
            a[v][f] = reinterpret_cast<const int2*>(arg + (size_t)n * 6)[f];
            g[v][f] = reinterpret_cast<const double2*>(gsum + (size_t)n * 6)[f];
        }
    }
#pragma unroll
    for (int f = 0; f < 3; f++) {
        double add = 0.0;
#pragma unroll
        for (int v = 0; v < 3; v++) {
            add += a[v][f].x == e ? g[v][f].x : 0.0;
            add += a[v][f].y == e ? g[v][f].y : 0.0;
        }
        ub[(size_t)(f * DOF) * ne + e] += add;
    }
}

// Node gradients from the element accumulators: dJ/dDP(n) = sum over incident elements of the
// vertex-depth cotangent; dJ/dn(n) = sum of the element-mean cotangents / 3.
__global__ void __launch_bounds__(256) param_gather_kernel(int np, int ne, const int* __restrict__ off,
                                                           const int* __restrict__ lst,
                                                           const double* __restrict__ gdp,
                                                           const double* __restrict__ gnbar,
                                                           double* __restrict__ out_dp, double* __restrict__ out_n) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= np) return;
    double sdp = 0.0, sn = 0.0;
    for (int j = off[n]; j < off[n + 1]; j++) {
        const int ent = lst[j];
        const int e = ent >> 2, v = ent & 3;
        sdp += gdp[(size_t)v * ne + e];
        sn += gnbar[e];
    }
    out_dp[n] = sdp;
    out_n[n] = sn / 3.0;
}

}  // namespace dg
