// Stage kernel, three lanes per element ("triplet" variant of stage_kernel).
//
// A block holds NB elements and 3*NB threads; thread t works on element t % NB and on part
// j = t / NB of it (j is uniform across a wavefront when NB is a multiple of 64):
//   edges     lane j evaluates the flux on the element's side slot j (boundary slots, which are
//             rare, are all evaluated in order by lane 0 so their sum keeps the reference's order)
//   volume    lane j evaluates the area quadrature points I = j, j+3, ... (one for p=1, two for p=2)
//   update    lane j accumulates the RHS of field j over the points in order and updates field j
// Partial results meet in LDS.  The arithmetic of every quantity and the order of every sum are
// those of stage_kernel (hence of the reference), so the two variants agree bit for bit; the
// triplet variant trades a few LDS round trips for a third of the registers per lane and three
// times the lanes in flight.
#include "dg_device.hpp"

namespace dg {

// Flux contribution of one internal side slot to the element's RHS, accumulated from zero over the
// edge's Lobatto points (src/internal_edge_hydro.F:37-383 with the dry-side branches).
// The element's own modes are read from LDS (su[(f*DOF+k)*NB]) and the neighbour's from global
// memory at each point, so that neither stays in registers across the edge loop.
template <int P, int NB>
__device__ __forceinline__ double lds_trace(const double* su, int f, const double* ph) {
    constexpr int DOF = (P + 1) * (P + 2) / 2;
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < DOF; k++) s = s + su[(f * DOF + k) * NB] * ph[k];
    return s;
}

template <int P>
__device__ __forceinline__ double glb_trace(const double* __restrict__ u, int ne, int el, int f, const double* ph) {
    constexpr int DOF = (P + 1) * (P + 2) / 2;
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < DOF; k++) s = s + u[(size_t)(f * DOF + k) * ne + el] * ph[k];
    return s;
}

template <int P, int NB, bool SPH>
__device__ __forceinline__ void internal_slot(const StageArgs& a, const DevTables& T, int e, int nb, int L,
                                              int role, int nled, double nx, double ny, double xlen, double AREAS,
                                              const double* su, const double (&hb)[3],
                                              int wd, int wdn, double (&acc)[3][(P + 1) * (P + 2) / 2]) {
    constexpr int DOF = (P + 1) * (P + 2) / 2;
    constexpr int NEGP = P + 2;
    const ElemData& E = a.ed;
    const Phys& ph = a.ph;
    const int ne = E.ne;
    const double c13 = 1.0 / 3.0, c16 = 1.0 / 6.0;
    const double tx = -ny, ty = nx;
    double hbn[3];
    {
        const double d0 = E.dpv[nb], d1 = E.dpv[(size_t)ne + nb], d2 = E.dpv[2 * (size_t)ne + nb];
        hbn[0] = c13 * (d0 + d1 + d2); hbn[1] = -(c16 * (d0 + d1)) + c13 * d2; hbn[2] = -(0.5 * d0) + 0.5 * d1;
    }
    const int wd_in_el = role ? wdn : wd, wd_ex_el = role ? wd : wdn;
    const double edfac = xlen / AREAS;
#pragma unroll
    for (int f = 0; f < 3; f++)
#pragma unroll
        for (int k = 0; k < DOF; k++) acc[f][k] = 0.0;
#pragma unroll 1
    for (int I = 0; I < NEGP; I++) {
        const int gi = I, gx = NEGP - 1 - I;
        const int gown = role ? gx : gi, gnb = role ? gi : gx;
        const double* po = T.phi_edge[L][gown];
        const double* pn = T.phi_edge[nled][gnb];
        const double zo = lds_trace<P, NB>(su, 0, po), xo = lds_trace<P, NB>(su, 1, po), yo = lds_trace<P, NB>(su, 2, po);
        const double zn = glb_trace<P>(a.uin, ne, nb, 0, pn), xn = glb_trace<P>(a.uin, ne, nb, 1, pn);
        const double yn = glb_trace<P>(a.uin, ne, nb, 2, pn);
        const double ho = trace<3>(hb, po), hn = trace<3>(hbn, pn);
        double so = 1.0, sn = 1.0;
        if (SPH) {
            so = E.sfaced[((size_t)L * NEGP + gown) * ne + e];
            sn = E.sfaced[((size_t)nled * NEGP + gnb) * ne + nb];
        }
        const double ze_in = role ? zn : zo, qx_in = role ? xn : xo, qy_in = role ? yn : yo;
        const double ze_ex = role ? zo : zn, qx_ex = role ? xo : xn, qy_ex = role ? yo : yn;
        const double hb_in = role ? hn : ho, hb_ex = role ? ho : hn;
        const double sf_in = role ? sn : so, sf_ex = role ? so : sn;
        double f0, f1, f2;
        llf_flux(ph, ze_in, ze_ex, qx_in, qx_ex, qy_in, qy_ex, hb_in, hb_ex, sf_in, sf_ex, nx, ny, ph.g, f0, f1, f2);
        double fo = f0, gm = f1, hm = f2;
        int mode = 0;
        if (fabs(f0) > 1.0e-8) {
            if (wd_in_el == 0) mode = (f0 > 0.0) ? 1 : (role ? 0 : 2);
            else if (wd_ex_el == 0) mode = (f0 < 0.0) ? 3 : (role ? 4 : 0);
        }
        if (mode) {
            double a_zi = ze_in, a_xi = qx_in, a_yi = qy_in, a_hi = hb_in, a_si = sf_in;
            double a_ze = ze_ex, a_xe = qx_ex, a_ye = qy_ex, a_he = hb_ex, a_se = sf_ex;
            double G = ph.g;
            if (mode == 1 || mode == 3) {
                const double sx = (mode == 1) ? qx_ex : qx_in, sy = (mode == 1) ? qy_ex : qy_in;
                const double qn = sx * nx + sy * ny, qt = sx * tx + sy * ty;
                const double qnm = -qn, qtm = qt;
                const double den = (nx * ty - ny * tx);
                const double mqx = (ty * qnm - ny * qtm) / den;
                const double mqy = (-tx * qnm + nx * qtm) / den;
                if (mode == 1) { a_zi = ze_ex; a_xi = mqx; a_yi = mqy; a_hi = hb_ex; a_si = sf_ex; }
                else { a_ze = ze_in; a_xe = mqx; a_ye = mqy; a_he = hb_in; a_se = sf_in; }
            } else {
                G = 0.0;
            }
            double r0, r1, r2;
            llf_flux(ph, a_zi, a_ze, a_xi, a_xe, a_yi, a_ye, a_hi, a_he, a_si, a_se, nx, ny, G, r0, r1, r2);
            if (mode == 1) { fo = r0; gm = role ? r1 : 0.0; hm = role ? r2 : 0.0; }
            else if (mode == 3) { fo = r0; gm = role ? 0.0 : r1; hm = role ? 0.0 : r2; }
            else { gm = r1; hm = r2; }
        }
        const double* eq = T.edgeq[L][gown];
#pragma unroll
        for (int k = 0; k < DOF; k++) {
            const double w = edfac * eq[k];
            if (!role) {
                acc[0][k] = acc[0][k] - w * fo; acc[1][k] = acc[1][k] - w * gm; acc[2][k] = acc[2][k] - w * hm;
            } else {
                acc[0][k] = acc[0][k] + w * fo; acc[1][k] = acc[1][k] + w * gm; acc[2][k] = acc[2][k] + w * hm;
            }
        }
    }
}

// Flux contribution of one boundary side slot (ocean, land or flux-specified), accumulated into acc
// in the order of the reference's boundary routines.
template <int P, int NB, bool SPH>
__device__ __forceinline__ void boundary_slot(const StageArgs& a, const DevTables& T, int e, int nb, int kind, int L,
                                              double nx, double ny, double xlen, double AREAS, int sid, int row,
                                              double h0, const double* su,
                                              const double (&hb)[3], double (&acc)[3][(P + 1) * (P + 2) / 2]) {
    constexpr int DOF = (P + 1) * (P + 2) / 2;
    constexpr int NEGP = P + 2;
    const ElemData& E = a.ed;
    const Phys& ph = a.ph;
    const int ne = E.ne;
    const double tx = -ny, ty = nx;
#pragma unroll 1
    for (int I = 0; I < NEGP; I++) {
        const double* pv = T.phi_edge[L][I];
        const double ze_in = lds_trace<P, NB>(su, 0, pv), qx_in = lds_trace<P, NB>(su, 1, pv);
        const double qy_in = lds_trace<P, NB>(su, 2, pv);
        const double hb_in = trace<3>(hb, pv);
        const double sf = SPH ? E.sfaced[((size_t)L * NEGP + I) * ne + e] : 1.0;
        double f0, f1, f2;
        if (kind == SIDE_LAND) {
            const double qn_int = qx_in * nx + qy_in * ny, qt_int = qx_in * tx + qy_in * ty;
            const double qn_ext = -qn_int, qt_ext = qt_int;
            const double den = 1.0 / (nx * ty - ny * tx);
            const double qx_ex = (ty * qn_ext - ny * qt_ext) * den;
            const double qy_ex = (-tx * qn_ext + nx * qt_ext) * den;
            llf_flux(ph, ze_in, ze_in, qx_in, qx_ex, qy_in, qy_ex, hb_in, hb_in, sf, sf, nx, ny, ph.g, f0, f1, f2);
        } else if (kind == SIDE_OCEAN) {
            double ze_ex = 0.0;
            const double xe = T.xegp[I];
            for (int jj = 0; jj < a.fo.nbfr; jj++) {
                const size_t t = ((size_t)sid * a.fo.rows + row) * a.fo.nbfr + jj;
                const double argj = a.fo.argj[t];
                const double rff = a.fo.rff[t];
                const size_t o = ((size_t)nb * a.fo.nbfr + jj) * 2;
                const double z1 = a.fo.emo[o] * rff * cos(argj - a.fo.efa[o]);
                const double z2 = a.fo.emo[o + 1] * rff * cos(argj - a.fo.efa[o + 1]);
                ze_ex = ze_ex + (0.5 * (z1 + z2) + 0.5 * (z2 - z1) * xe);
            }
            if ((ze_ex * ph.ifnlfa + hb_in) <= 0) ze_ex = fabs(hb_in) + h0;
            llf_flux(ph, ze_in, ze_ex, qx_in, qx_in, qy_in, qy_in, hb_in, hb_in, sf, sf, nx, ny, ph.g, f0, f1, f2);
        } else {
            double qn_ext = 0.0, qx_ex = 0.0, qy_ex = 0.0;
            const double xe = T.xegp[I];
            for (int jj = 0; jj < a.fo.nffr; jj++) {
                const size_t t = ((size_t)sid * a.fo.rows + row) * a.fo.nffr + jj;
                const double argj = a.fo.fargj[t];
                const double rff = a.fo.frff[t];
                const size_t o = ((size_t)nb * a.fo.nffr + jj) * 2;
                const double q1 = a.fo.qnam[o], q2 = a.fo.qnam[o + 1];
                const double p1 = a.fo.qnph[o], p2 = a.fo.qnph[o + 1];
                const double qnam_gp = 0.5 * (q1 + q2) + 0.5 * (q2 - q1) * xe;
                const double qnph_gp = 0.5 * (p1 + p2) + 0.5 * (p2 - p1) * xe;
                qn_ext = qn_ext + qnam_gp * rff * cos(argj - qnph_gp);
                const double qt_ext = 0.0;
                qx_ex = -(ty * qn_ext - ny * qt_ext) / (nx * ty - ny * tx);
                qy_ex = -(-tx * qn_ext + nx * qt_ext) / (nx * ty - ny * tx);
            }
            llf_flux(ph, ze_in, ze_in, qx_in, qx_ex, qy_in, qy_ex, hb_in, hb_in, sf, sf, nx, ny, ph.g, f0, f1, f2);
        }
#pragma unroll
        for (int k = 0; k < DOF; k++) {
            const double w = 2.0 * T.m_inv[k] / AREAS * xlen * pv[k] * T.wegp[I];
            acc[0][k] = acc[0][k] - w * f0;
            acc[1][k] = acc[1][k] - w * f1;
            acc[2][k] = acc[2][k] - w * f2;
        }
    }
}

template <int P, int NB, bool SPH, bool WDFUSED, bool MET>
__global__ void __launch_bounds__(3 * NB) stage3_kernel(StageArgs a) {
    constexpr int DOF = (P + 1) * (P + 2) / 2;
    constexpr int NAGP = (P == 1) ? 3 : 6;
    constexpr int NF = 3 * DOF;
    __shared__ double s_u[NF][NB];            // the element's modes
    __shared__ double s_b[NF][NB];            // boundary part, then the RHS after the edges
    __shared__ double s_x[3][NF][NB];         // per-slot internal parts, then quadrature-point values
    __shared__ double s_met[MET ? 8 : 1][NB];
    static_assert(3 * NF >= 8 * NAGP, "quadrature-point values must fit the slot buffer");
    const DevTables& T = tab<P>();
    const ElemData& E = a.ed;
    const Phys& ph = a.ph;
    const int ne = E.ne;
    const int j = threadIdx.x / NB, le = threadIdx.x % NB;
    const int e = xcd_tile(blockIdx.x, gridDim.x) * NB + le;
    const bool valid = e < ne;
    const double c13 = 1.0 / 3.0, c16 = 1.0 / 6.0;
    int sid = 0, wd = 0, info = 0;
    Scen sc{};
    if (valid) {
        sid = E.sid ? (int)E.sid[e] : 0;
        sc = E.scen[sid];
        wd = a.wd_in[e];
        info = E.sinfo[e];
#pragma unroll
        for (int k = 0; k < DOF; k++) s_u[j * DOF + k][le] = a.uin[(size_t)(j * DOF + k) * ne + e];
    }
    const int row = *a.fo.step * a.fo.nrk + a.irk;
    __syncthreads();

    const double* su = &s_u[0][le];
    double AREAS = 1.0, hb[3] = {0.0, 0.0, 0.0};
    if (valid) {
        const double X0 = E.xv[e], X1 = E.xv[(size_t)ne + e], X2 = E.xv[2 * (size_t)ne + e];
        const double Y0 = E.yv[e], Y1 = E.yv[(size_t)ne + e], Y2 = E.yv[2 * (size_t)ne + e];
        const double D0 = E.dpv[e], D1 = E.dpv[(size_t)ne + e], D2 = E.dpv[2 * (size_t)ne + e];
        AREAS = (X0 - X2) * (Y1 - Y2) + (X2 - X1) * (Y0 - Y2);
        hb[0] = c13 * (D0 + D1 + D2); hb[1] = -(c16 * (D0 + D1)) + c13 * D2; hb[2] = -(0.5 * D0) + 0.5 * D1;

        // ---- edges
        auto side = [&](int slot, int& L, int& role, int& kind, int& nled, int& nb, double& nx, double& ny,
                        double& xlen) {
            const int bits = side_bits(info, slot);
            L = bits & 3; role = (bits >> 2) & 1; kind = (bits >> 3) & 3; nled = (bits >> 5) & 3;
            nb = E.snbr[(size_t)slot * ne + e];
            int va = (L + 1) % 3, vb = (L + 2) % 3;
            if (role) { const int t = va; va = vb; vb = t; }
            const double dy = E.yv[(size_t)vb * ne + e] - E.yv[(size_t)va * ne + e];
            const double dx = E.xv[(size_t)vb * ne + e] - E.xv[(size_t)va * ne + e];
            xlen = sqrt(dy * dy + dx * dx);
            nx = dy / xlen; ny = -dx / xlen;
        };
        double acc[3][DOF];
        // boundary slots come first in slot order: lane 0 evaluates all of them in order
        if (j == 0) {
#pragma unroll
            for (int f = 0; f < 3; f++)
#pragma unroll
                for (int k = 0; k < DOF; k++) acc[f][k] = 0.0;
            if (wd != 0) {
#pragma unroll 1
                for (int slot = 0; slot < 3; slot++) {
                    if (((side_bits(info, slot) >> 3) & 3) == SIDE_INTERNAL) break;
                    int L, role, kind, nled, nb;
                    double nx, ny, xlen;
                    side(slot, L, role, kind, nled, nb, nx, ny, xlen);
                    boundary_slot<P, NB, SPH>(a, T, e, nb, kind, L, nx, ny, xlen, AREAS, sid, row, sc.h0, su, hb, acc);
                }
            }
#pragma unroll
            for (int f = 0; f < 3; f++)
#pragma unroll
                for (int k = 0; k < DOF; k++) s_b[f * DOF + k][le] = acc[f][k];
        }
        // internal slot j
        {
            int L, role, kind, nled, nb;
            double nx, ny, xlen;
            side(j, L, role, kind, nled, nb, nx, ny, xlen);
            bool any = false;
            if (kind == SIDE_INTERNAL) {
                const int wdn = a.wd_in[nb];
                if (wd != 0 || wdn != 0) {
                    internal_slot<P, NB, SPH>(a, T, e, nb, L, role, nled, nx, ny, xlen, AREAS, su, hb, wd, wdn, acc);
                    any = true;
                }
            }
#pragma unroll
            for (int f = 0; f < 3; f++)
#pragma unroll
                for (int k = 0; k < DOF; k++) s_x[j][f * DOF + k][le] = any ? acc[f][k] : 0.0;
        }
    }
    __syncthreads();
    // ---- RHS after the edges: boundary part + internal slot sum (add_internal_flux), then rain
    double R[DOF];
    if (valid) {
#pragma unroll
        for (int k = 0; k < DOF; k++) {
            const int q = j * DOF + k;
            double s = 0.0;
            s = s + s_x[0][q][le];
            s = s + s_x[1][q][le];
            s = s + s_x[2][q][le];
            R[k] = s_b[q][le] + s;
        }
        if (j == 0 && sc.rain > 0.0) {
#pragma unroll 1
            for (int I = 0; I < NAGP; I++)
#pragma unroll
                for (int k = 0; k < DOF; k++) R[k] = R[k] + T.srfac[I][k] * sc.rain;
        }
    }
    __syncthreads();   // s_x is reused below

    // ---- volume quadrature points I = j, j+3, ...
    double drdx = 0.0, dsdx = 0.0, drdy = 0.0, dsdy = 0.0;
    if (valid && wd != 0) {
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
        if (MET && j == 0) {
            const double ramp = a.fo.ramp[(size_t)sid * a.fo.rows + row];
            const double drpsi[3] = {-1.0 / 2.0, 1.0 / 2.0, 0.0}, dspsi[3] = {-1.0 / 2.0, 0.0, 1.0 / 2.0};
            double pdx = 0.0, pdy = 0.0;
#pragma unroll
            for (int v = 0; v < 3; v++) {
                const int n = E.nm[(size_t)v * ne + e];
                const double cd = a.mt.cd[n], wm = a.mt.wm[n];
                s_met[v][le] = ramp * 0.001293 * cd * a.mt.wx[n] * wm;
                s_met[3 + v][le] = ramp * 0.001293 * cd * a.mt.wy[n] * wm;
                const double pr = ramp * a.mt.prn[n];
                const double dpsidx = drpsi[v] * drdx + dspsi[v] * dsdx;
                const double dpsidy = drpsi[v] * drdy + dspsi[v] * dsdy;
                if (v == 0) { pdx = pr * dpsidx; pdy = pr * dpsidy; }
                else { pdx = pdx + pr * dpsidx; pdy = pdy + pr * dpsidy; }
            }
            s_met[6][le] = pdx;
            s_met[7][le] = pdy;
        }
    }
    if (MET) __syncthreads();
    if (valid && wd != 0) {
        const double depth_c = hb[0] + su[0];
        const double fh_nl_c = 1.0 / (ph.nleq * depth_c + ph.leq);
        const double cori = E.cori_el ? E.cori_el[e] : sc.cori;
        const double nbar = ph.load_manning ? E.nbar[e] : 0.0;
        double* xq = &s_x[0][0][0];
#pragma unroll 1
        for (int I = j; I < NAGP; I += 3) {
            const double* pa = T.phi_area[I];
            double ze = su[0], qx = su[DOF * NB], qy = su[2 * DOF * NB];
#pragma unroll
            for (int k = 1; k < DOF; k++) {
                ze = ze + su[k * NB] * pa[k];
                qx = qx + su[(DOF + k) * NB] * pa[k];
                qy = qy + su[(2 * DOF + k) * NB] * pa[k];
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
            const double f1 = ph.nleq + ph.leq * hbq;
            const double fx = qx * f1 * sfq, fy = qy * f1;
            const double fu = ph.nleq * qx, fv = ph.nleq * qy;
            const double fg = ph.nleqg * ze * wd;
            const double fh = 1.0 / (ph.nleq * depth + ph.leq);
            const double u = qx * fh, v = qy * fh;
            const double huu = fu * u, hvv = fv * v, huv = fu * v;
            const double gh2 = fg * (0.5 * ze + hbq) + ph.fg_l * ze;
            const double gx = (huu + gh2) * sfq, gy = huv;
            const double hx = huv * sfq, hy = hvv + gh2;
            double fric = sc.fric;
            if (ph.load_manning) {
                fric = ph.g * (nbar * nbar) / pow(depth, 1.0 / 3.0);
                if (fric < sc.cf) fric = sc.cf;
            }
            double tau = fric;
            if (ph.iflinbf == 0) {
                const double umag = sqrt(u * u + v * v);
                tau = tau * umag * fh;
                if (ph.ifhybf == 1) tau = tau * pow(1.0 + pow(ph.hbreak * fh_nl_c, ph.ftheta), ph.fgamma / ph.ftheta);
                const double cap = 0.9 * sc.dtdph;
                if (cap < tau) tau = cap;
            }
            double sx = -(tau * qx) + fg * dhbx * sfq + cori * qy;
            double sy = -(tau * qy) + fg * dhby - cori * qx;
            if (MET) {
                const double fw = 1.0 / f1;
                const double* ps = T.psi[I];
                sx = sx + fw * (s_met[0][le] * ps[0] + s_met[1][le] * ps[1] + s_met[2][le] * ps[2]) -
                     ph.g * sfq * depth * s_met[6][le];
                sy = sy + fw * (s_met[3][le] * ps[0] + s_met[4][le] * ps[1] + s_met[5][le] * ps[2]) -
                     ph.g * depth * s_met[7][le];
            }
            const double vals[8] = {fx, fy, gx, gy, hx, hy, sx, sy};
#pragma unroll
            for (int m = 0; m < 8; m++) xq[((size_t)(I * 8 + m)) * NB + le] = vals[m];
        }
    }
    __syncthreads();
    // ---- field j: RHS over the quadrature points in order, then the SSP-RK update
    if (valid) {
        if (wd != 0) {
            const double* xq = &s_x[0][0][0];
#pragma unroll 1
            for (int I = 0; I < NAGP; I++) {
                const double vx = xq[((size_t)(I * 8 + 2 * j)) * NB + le];
                const double vy = xq[((size_t)(I * 8 + 2 * j + 1)) * NB + le];
                const double sv = xq[((size_t)(I * 8 + 6 + (j == 2 ? 1 : 0))) * NB + le];
                if (j != 0) R[0] = R[0] + T.srfac[I][0] * sv;
#pragma unroll
                for (int k = 1; k < DOF; k++) {
                    const double dpxk = T.drphi[I][k] * drdx + T.dsphi[I][k] * dsdx;
                    const double dpyk = T.drphi[I][k] * drdy + T.dsphi[I][k] * dsdy;
                    const double xf = T.mw[I][k] * dpxk, yf = T.mw[I][k] * dpyk;
                    if (j == 0) R[k] = R[k] + xf * vx + yf * vy;
                    else R[k] = R[k] + xf * vx + yf * vy + T.srfac[I][k] * sv;
                }
            }
        }
        if (a.rout) {
#pragma unroll
            for (int k = 0; k < DOF; k++) a.rout[(size_t)(j * DOF + k) * ne + e] = R[k];
        }
        double W[DOF];
        const double bcur = a.bcur * sc.dtdp;
#pragma unroll
        for (int k = 0; k < DOF; k++) W[k] = 0.0;
#pragma unroll 1
        for (int i = 0; i < a.nprev; i++) {
            const double ai = a.aprev[i], bi = a.bprev[i] * sc.dtdp;
            if (ai != 0.0) {
#pragma unroll
                for (int k = 0; k < DOF; k++) W[k] = W[k] + ai * a.uprev[i][(size_t)(j * DOF + k) * ne + e];
            }
            if (bi != 0.0) {
#pragma unroll
                for (int k = 0; k < DOF; k++) W[k] = W[k] + bi * a.rprev[i][(size_t)(j * DOF + k) * ne + e];
            }
        }
#pragma unroll
        for (int k = 0; k < DOF; k++) W[k] = W[k] + a.acur * su[(j * DOF + k) * NB] + bcur * R[k];
        if (WDFUSED) {
            // WETDRY needs the whole element: lane 0 gathers the three fields through LDS
#pragma unroll
            for (int k = 0; k < DOF; k++) s_u[j * DOF + k][le] = W[k];
        } else {
#pragma unroll
            for (int k = 0; k < DOF; k++) a.uout[(size_t)(j * DOF + k) * ne + e] = W[k];
        }
    }
    if (WDFUSED) {
        __syncthreads();
        if (valid && j == 0) {
            double Wa[3][DOF];
#pragma unroll
            for (int f = 0; f < 3; f++)
#pragma unroll
                for (int k = 0; k < DOF; k++) Wa[f][k] = s_u[f * DOF + k][le];
            const double D[3] = {E.dpv[e], E.dpv[(size_t)ne + e], E.dpv[2 * (size_t)ne + e]};
            const double dpe_min = fmin(fmin(D[0], D[1]), D[2]);
            a.wd_out[e] = (uint8_t)wetdry_element<P>(T, Wa, D, hb, dpe_min, wd, sc.h0);
#pragma unroll
            for (int f = 0; f < 3; f++)
#pragma unroll
                for (int k = 0; k < DOF; k++) a.uout[(size_t)(f * DOF + k) * ne + e] = Wa[f][k];
        }
    }
}

}  // namespace dg
