// HIP kernels of the DG hydro step for gfx950 (MI355X), FP64.
//
// One time step = NRK stages; each stage is
//   stage_kernel     one thread per element: fluxes on its three edges (ocean / land / flux /
//                    internal, LLF), volume quadrature, SSP-RK update of the element's modes, and
//                    (when the slope limiter is off) the element's wet/dry correction.  Internal
//                    edge fluxes are evaluated by both neighbours with the same operands in the same
//                    order, so both see bit-identical values and no scatter or atomics are needed.
//   node_minmax      one thread per node: min/max of the updated cell means over incident elements
//   limit_wetdry     one thread per element: SLOPELIMITER5 on modes 2-3, then WETDRY
// preceded once per step by tide_kernel (the tidal boundary terms of every stage) and, with NWS=8,
// met_kernel (the Holland profile at the nodes); plus modal2nodal (one thread per node) on output
// steps.
//
// Arithmetic follows the reference statement by statement (listed per kernel) and is compiled with
// -ffp-contract=off, so each product and sum is rounded as in dgswem_serial; the cosine of the
// tidal terms and Manning's DEPTH**(1/3) are correctly rounded (dg_crmath.hpp).
#include "dg_device.hpp"
#include "dg_setup.hpp"

namespace dg {

__constant__ DevTables c_tab1;   // p = 1
__constant__ DevTables c_tab2;   // p = 2

template <int P> __device__ __forceinline__ const DevTables& tab();
template <> __device__ __forceinline__ const DevTables& tab<1>() { return c_tab1; }
template <> __device__ __forceinline__ const DevTables& tab<2>() { return c_tab2; }

// Blocks are dispatched round-robin over the 8 XCDs; give each XCD a contiguous range of element
// tiles so that a tile's neighbours (close in the Hilbert order) sit in the same XCD's L2.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
    const int q = nb >> 3, r = nb & 7, x = b & 7;
    return x * q + min(x, r) + (b >> 3);
}

struct StageArgs {
    ElemData ed;
    Phys ph;
    Forcing fo;
    Met mt;
    const double* __restrict__ uin;      // U(IRK) [3][dof][ne]
    double* __restrict__ uout;           // U(IRK+1)
    const uint8_t* __restrict__ wd_in;   // WDFLG during the stage
    uint8_t* __restrict__ wd_out;        // WDFLG after WETDRY (fused variant)
    int nprev;                           // stages I < IRK entering the update
    const double* uprev[KMAXRK];
    double aprev[KMAXRK];                // ATVD(IRK,I), 0 = term absent
    const double* rprev[KMAXRK];
    double bprev[KMAXRK];                // BTVD(IRK,I) (times the scenario's DT in the kernel)
    double acur, bcur;                   // ATVD(IRK,IRK), BTVD(IRK,IRK)
    double* __restrict__ rout;           // RHS(IRK) kept for later stages, or nullptr
    int irk;                             // stage index
    // block-local staging: tile t (elements [t*NT, t*NT+NT)) stages its elements and its halo
    // hidx[hoff[t]..hoff[t+1]) in LDS; snbl packs each side's local neighbour index (10 bits each)
    const int* __restrict__ hoff;
    const int* __restrict__ hidx;
    const int* __restrict__ snbl;
    int ns;                              // staged elements per block (LDS row length)
    int* __restrict__ wd_active;         // set to 1 when WETDRY changes an element, or nullptr
    const double* __restrict__ tip;      // TIP2 of this stage at the nodes (tidal potential), or nullptr
    double* __restrict__ cmout;          // copy of the output cell means [3][ne] for the limiter, or nullptr
};

// ------------------------------------------------------------------------------------------
// WETDRY for one element (src/wetdry.F:58-548, p=1 vertex logic; modes >= 4 zeroed).
// Updates ze/qx/qy in place and returns the new WDFLG.
// ------------------------------------------------------------------------------------------
template <int P>
__device__ __forceinline__ int wetdry_element(const DevTables& T, double (&U)[3][(P + 1) * (P + 2) / 2],
                                              const double dpn[3], const double hb[3], double dpe_min,
                                              int wdflg, double H0) {
    constexpr int DOF = (P + 1) * (P + 2) / 2;
    const double ZERO = 1.0e-6, C13 = 1.0 / 3.0, C16 = 1.0 / 6.0;
    int ndry = 0, ht_max_i = 0;
    double ht_max = ZERO;
    double ze_node[3], ht_node[3], ze_hat[3] = {0.0, 0.0, 0.0};
#pragma unroll
    for (int i = 0; i < 3; i++) {
        double z = U[0][0];
#pragma unroll
        for (int k = 1; k < 3; k++) z = z + T.phi_corner[i][k] * U[0][k];
        ze_node[i] = z;
        ht_node[i] = dpn[i] + z;
        if (ht_node[i] > ht_max) { ht_max_i = i; ht_max = ht_node[i]; }
        if (ht_node[i] <= (H0 + ZERO)) ndry++;
    }
    const double dp_avg = C13 * (dpn[0] + dpn[1] + dpn[2]);
    int check = 0;
    if (ndry == 0) {
        if (wdflg == 1) return 1;
        ze_hat[0] = ze_node[0]; ze_hat[1] = ze_node[1]; ze_hat[2] = ze_node[2];
        check = 1;
    } else if (ndry == 3) {
        const double ht_avg = U[0][0] + dp_avg;
#pragma unroll
        for (int k = 0; k < DOF; k++) { U[0][k] = 0.0; U[1][k] = 0.0; U[2][k] = 0.0; }
        U[0][0] = ht_avg - dp_avg;
        if (ht_avg < ZERO) U[0][0] = ZERO - dp_avg;
        U[0][1] = -hb[1]; U[0][2] = -hb[2];
        return 0;
    } else {
        const double ht_avg = U[0][0] + dp_avg;
        if (ht_avg <= (H0 + ZERO)) {
#pragma unroll
            for (int k = 0; k < DOF; k++) { U[0][k] = 0.0; U[1][k] = 0.0; U[2][k] = 0.0; }
            U[0][0] = ht_avg - dp_avg;
            if (ht_avg < ZERO) U[0][0] = ZERO - dp_avg;
            U[0][1] = -hb[1]; U[0][2] = -hb[2];
            ze_hat[0] = ze_hat[1] = ze_hat[2] = U[0][0];
        } else {
            // order the vertices by depth (the reference's exchange sort)
            int l2h[3] = {0, 1, 2};
#pragma unroll
            for (int k1 = 0; k1 < 3; k1++)
#pragma unroll
                for (int k2 = k1; k2 < 3; k2++)
                    if (ht_node[l2h[k1]] > ht_node[l2h[k2]]) { const int t = l2h[k1]; l2h[k1] = l2h[k2]; l2h[k2] = t; }
            double dh[3];
            dh[l2h[0]] = fmax(H0, ht_node[l2h[0]]) - ht_node[l2h[0]];
            dh[l2h[1]] = fmax(H0, ht_node[l2h[1]] - 0.5 * dh[l2h[0]]) - ht_node[l2h[1]];
            dh[l2h[2]] = -dh[l2h[0]] - dh[l2h[1]];
            int nd = 0;
            double qxs = 0, qys = 0;
            ht_max = 0;
            double ht_hat[3], qxn[3], qyn[3];
#pragma unroll
            for (int i = 0; i < 3; i++) {
                ht_hat[i] = ht_node[i] + dh[i];
                double a = U[1][0], b = U[2][0];
                if (ht_hat[i] > ht_max) { ht_max_i = i; ht_max = ht_hat[i]; }
#pragma unroll
                for (int k = 1; k < 3; k++) { a = a + T.phi_corner[i][k] * U[1][k]; b = b + T.phi_corner[i][k] * U[2][k]; }
                qxn[i] = a; qyn[i] = b;
                if (ht_hat[i] <= (H0 + 1.0e-10)) { nd++; qxs = qxs + a; qys = qys + b; }
            }
            if (nd == 3) {
#pragma unroll
                for (int k = 0; k < DOF; k++) { U[1][k] = 0.0; U[2][k] = 0.0; }
            } else {
                qxs = qxs / (3.0 - (double)nd);
                qys = qys / (3.0 - (double)nd);
                double qxh[3], qyh[3];
#pragma unroll
                for (int i = 0; i < 3; i++) {
                    if (ht_hat[i] <= (H0 + 1.0e-10)) { qxh[i] = 0.0; qyh[i] = 0.0; }
                    else { qxh[i] = qxn[i] + qxs; qyh[i] = qyn[i] + qys; }
                }
                U[1][0] = C13 * (qxh[0] + qxh[1] + qxh[2]);
                U[1][1] = -(C16 * (qxh[0] + qxh[1])) + C13 * qxh[2];
                U[1][2] = -(0.5 * qxh[0]) + 0.5 * qxh[1];
                U[2][0] = C13 * (qyh[0] + qyh[1] + qyh[2]);
                U[2][1] = -(C16 * (qyh[0] + qyh[1])) + C13 * qyh[2];
                U[2][2] = -(0.5 * qyh[0]) + 0.5 * qyh[1];
            }
#pragma unroll
            for (int i = 0; i < 3; i++) ze_hat[i] = ht_hat[i] - dpn[i];
            U[0][1] = -(C16 * (ze_hat[0] + ze_hat[1])) + C13 * ze_hat[2];
            U[0][2] = -(0.5 * ze_hat[0]) + 0.5 * ze_hat[1];
        }
#pragma unroll
        for (int k = 3; k < DOF; k++) { U[0][k] = 0.0; U[1][k] = 0.0; U[2][k] = 0.0; }
        check = 1;
    }
    if (check) {
        if (ze_hat[ht_max_i] > (-dpe_min + H0)) return 1;
#pragma unroll
        for (int k = 0; k < DOF; k++) { U[1][k] = 0.0; U[2][k] = 0.0; }
        return 0;
    }
    return wdflg;
}

// True when WETDRY leaves the element as it is (every vertex wet and the element already wet: the
// identity branch of wetdry_element).
template <int P>
__device__ __forceinline__ bool wetdry_keeps(const DevTables& T, const double (&U)[3][(P + 1) * (P + 2) / 2],
                                             const double dpn[3], int wdflg, double H0) {
    if (wdflg != 1) return false;
#pragma unroll
    for (int i = 0; i < 3; i++) {
        double z = U[0][0];
#pragma unroll
        for (int k = 1; k < 3; k++) z = z + T.phi_corner[i][k] * U[0][k];
        if (dpn[i] + z <= (H0 + 1.0e-6)) return false;
    }
    return true;
}

// ------------------------------------------------------------------------------------------
// Element geometry, once per mesh (src/calc_normal.F:28-49, src/prep_DG.F:959-1163): per slot the
// side's unit normal in the reference's node order (the counter-clockwise order of the edge's
// NEDEL(1) element, so both elements of an internal edge form the same normal from the same two
// vertices) and XLEN/AREAS, and the Jacobian terms DRDX, DSDX, DRDY, DSDY.  The stage kernel
// reads them instead of forming them every stage; the expressions are the ones it would evaluate,
// so the values are the same bits.
// ------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) geom_kernel(int ne, const double* __restrict__ xv, const double* __restrict__ yv,
                                                   const int* __restrict__ sinfo, double* __restrict__ geo) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= ne) return;
    const double X[3] = {xv[e], xv[(size_t)ne + e], xv[2 * (size_t)ne + e]};
    const double Y[3] = {yv[e], yv[(size_t)ne + e], yv[2 * (size_t)ne + e]};
    const double AREAS = (X[0] - X[2]) * (Y[1] - Y[2]) + (X[2] - X[1]) * (Y[0] - Y[2]);
    const int info = sinfo[e];
    for (int slot = 0; slot < 3; slot++) {
        const int bits = side_bits(info, slot);
        const int L = bits & 3, role = (bits >> 2) & 1;
        int va = (L + 1) % 3, vb = (L + 2) % 3;
        if (role) { const int t = va; va = vb; vb = t; }
        const double dy = Y[vb] - Y[va];
        const double dx = X[vb] - X[va];
        const double xlen = sqrt(dy * dy + dx * dx);
        geo[(size_t)(3 * slot) * ne + e] = dy / xlen;
        geo[(size_t)(3 * slot + 1) * ne + e] = -dx / xlen;
        geo[(size_t)(3 * slot + 2) * ne + e] = xlen / AREAS;
    }
    double AREA = (X[0] - X[2]) * (Y[1] - Y[2]) + (X[2] - X[1]) * (Y[0] - Y[2]);
    AREA = AREA * 0.5;
    geo[(size_t)9 * ne + e] = 1.0 / AREA * (Y[2] - Y[0]);
    geo[(size_t)10 * ne + e] = 1.0 / AREA * (Y[0] - Y[1]);
    geo[(size_t)11 * ne + e] = 1.0 / AREA * (X[0] - X[2]);
    geo[(size_t)12 * ne + e] = 1.0 / AREA * (X[1] - X[0]);
}

// XLEN of a side and AREAS (twice the element's area), formed from the vertex coordinates as the
// geometry kernel forms them: the boundary-edge terms need them apart.
__device__ __forceinline__ void side_xlen_areas(const ElemData& E, int ne, int e, int L, int role, double& xlen,
                                                double& AREAS) {
    const double X[3] = {E.xv[e], E.xv[(size_t)ne + e], E.xv[2 * (size_t)ne + e]};
    const double Y[3] = {E.yv[e], E.yv[(size_t)ne + e], E.yv[2 * (size_t)ne + e]};
    AREAS = (X[0] - X[2]) * (Y[1] - Y[2]) + (X[2] - X[1]) * (Y[0] - Y[2]);
    int va = (L + 1) % 3, vb = (L + 2) % 3;
    if (role) { const int t = va; va = vb; vb = t; }
    const double dy = Y[vb] - Y[va];
    const double dx = X[vb] - X[va];
    xlen = sqrt(dy * dy + dx * dx);
}

// ------------------------------------------------------------------------------------------
// Stage kernel.  Reference statements restated:
//   OCEAN_EDGE_HYDRO   src/ocean_edge_hydro.F:19-355
//   LAND_EDGE_HYDRO    src/land_edge_hydro.F:26-311
//   FLOW_EDGE_HYDRO    src/flow_edge_hydro.F:20-333
//   INTERNAL_EDGE_HYDRO + add_internal_flux  src/internal_edge_hydro.F:37-383
//   RHS_DG_HYDRO       src/rhs_dg_hydro.F:18-473
//   SSP-RK update      src/DG_hydro_timestep.F:238-290
//   element geometry   src/read_input.F:1408, src/prep_DG.F:959-1163, src/calc_normal.F:28-49
// ------------------------------------------------------------------------------------------
// Trace of the element's modes at a point with basis values ph[] (sum from zero, mode order).
template <int DOF>
__device__ __forceinline__ double trace(const double (&u)[DOF], const double* ph) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < DOF; k++) s = s + u[k] * ph[k];
    return s;
}

// The same sum where ph[0] == 1: 0 + u0*1 = u0 for every u0 but -0, so the sum starts at u0.
template <int DOF>
__device__ __forceinline__ double trace1(const double (&u)[DOF], const double* ph) {
    double s = u[0];
#pragma unroll
    for (int k = 1; k < DOF; k++) s = s + u[k] * ph[k];
    return s;
}

template <int P, int NT, bool SPH, bool WDFUSED, bool MET, bool NL>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(P == 1 ? 2 : 1))) stage_kernel(StageArgs a) {
    constexpr int DOF = (P + 1) * (P + 2) / 2;
    constexpr int NAGP = (P == 1) ? 3 : 6;
    constexpr int NEGP = P + 2;
    // per-thread accumulators kept in LDS to bound register use: [0] boundary-edge part of the
    // RHS (accumulated first, like the reference's boundary edge routines), [1] internal-edge
    // slot sum (add_internal_flux)
    __shared__ double acc_lds[2][3 * DOF][NT];
    // edge basis and edge quadrature tables: indexed by each lane's own local edge number, so
    // read from LDS (short latency) rather than as divergent constant-memory loads
    __shared__ double tpe[3][NEGP][DOF], teq[3][NEGP][DOF];
    // staged states: rows [field*DOF + mode] then the three bathymetry modes, ns entries each, then
    // the wet/dry flags (bytes)
    extern __shared__ double stg[];
    const DevTables& T = tab<P>();
    const ElemData& E = a.ed;
    const Phys& ph = a.ph;
    const int ne = E.ne;
    const int tid = threadIdx.x;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int e = tile * NT + tid;
    const double c13 = 1.0 / 3.0, c16 = 1.0 / 6.0;
    const int ns = a.ns;
    uint8_t* const stwd = reinterpret_cast<uint8_t*>(stg + (size_t)(3 * DOF + 3) * ns);
    for (int j = tid; j < 3 * NEGP * DOF; j += NT) {
        const int l = j / (NEGP * DOF), g = (j / DOF) % NEGP, k = j % DOF;
        tpe[l][g][k] = T.phi_edge[l][g][k];
        teq[l][g][k] = T.edgeq[l][g][k];
    }
    // per-element loads that do not depend on the staged data, issued before the staging loads so
    // that the latencies overlap (lanes past the last element read element ne-1 and exit after
    // the barrier)
    const int ec = e < ne ? e : ne - 1;
    double G[9];                         // per slot NX, NY, XLEN/AREAS
#pragma unroll
    for (int j = 0; j < 9; j++) G[j] = E.geo[(size_t)j * ne + ec];
    const int info = E.sinfo[ec];
    const int nb0 = E.snbr[ec], nb1 = E.snbr[(size_t)ne + ec], nb2 = E.snbr[2 * (size_t)ne + ec];
    const int sl = a.snbl[ec];
    const int sid = E.sid ? (int)E.sid[ec] : 0;
    const Scen sc = E.scen[sid];
    const double cori = E.cori_el ? E.cori_el[ec] : sc.cori;
    const double nbar = ph.load_manning ? E.nbar[ec] : 0.0;
    {
        // this tile's elements and its halo, two entries per lane per pass with all their loads
        // issued before any is used
        const int h0 = a.hoff[tile], nh = a.hoff[tile + 1] - h0;
        for (int j = tid; j < NT + nh; j += 2 * NT) {
            const int j2 = j + NT;
            const bool has2 = j2 < NT + nh;
            int g = j < NT ? tile * NT + j : a.hidx[h0 + j - NT];
            int g2 = has2 ? a.hidx[h0 + j2 - NT] : ne;
            const bool ok = g < ne, ok2 = g2 < ne;
            g = ok ? g : 0;
            g2 = ok2 ? g2 : 0;
            double v[3 * DOF + 3], v2[3 * DOF + 3];
#pragma unroll
            for (int f = 0; f < 3 * DOF; f++) { v[f] = a.uin[(size_t)f * ne + g]; v2[f] = a.uin[(size_t)f * ne + g2]; }
#pragma unroll
            for (int m = 0; m < 3; m++) { v[3 * DOF + m] = E.dpv[(size_t)m * ne + g]; v2[3 * DOF + m] = E.dpv[(size_t)m * ne + g2]; }
            const uint8_t w = a.wd_in[g], w2 = a.wd_in[g2];
            if (ok) {
#pragma unroll
                for (int f = 0; f < 3 * DOF; f++) stg[f * ns + j] = v[f];
                const double d0 = v[3 * DOF], d1 = v[3 * DOF + 1], d2 = v[3 * DOF + 2];
                stg[(3 * DOF) * ns + j] = c13 * (d0 + d1 + d2);
                stg[(3 * DOF + 1) * ns + j] = -(c16 * (d0 + d1)) + c13 * d2;
                stg[(3 * DOF + 2) * ns + j] = -(0.5 * d0) + 0.5 * d1;
                stwd[j] = w;
            }
            if (ok2) {
#pragma unroll
                for (int f = 0; f < 3 * DOF; f++) stg[f * ns + j2] = v2[f];
                const double d0 = v2[3 * DOF], d1 = v2[3 * DOF + 1], d2 = v2[3 * DOF + 2];
                stg[(3 * DOF) * ns + j2] = c13 * (d0 + d1 + d2);
                stg[(3 * DOF + 1) * ns + j2] = -(c16 * (d0 + d1)) + c13 * d2;
                stg[(3 * DOF + 2) * ns + j2] = -(0.5 * d0) + 0.5 * d1;
                stwd[j2] = w2;
            }
        }
    }
    __syncthreads();
    if (e >= ne) return;
    const int row = table_step(a.fo) * a.fo.nrk + a.irk;

    const int wd = stwd[tid];
#pragma unroll
    for (int j = 0; j < 3 * DOF; j++) { acc_lds[0][j][tid] = 0.0; acc_lds[1][j][tid] = 0.0; }

#pragma unroll 1
    for (int slot = 0; slot < 3; slot++) {
        const int bits = side_bits(info, slot);
        const int L = bits & 3, role = (bits >> 2) & 1, kind = (bits >> 3) & 3, nled = (bits >> 5) & 3;
        const int nb = slot == 0 ? nb0 : (slot == 1 ? nb1 : nb2);
        // edge geometry in the reference's node order (the counter-clockwise order of NEDEL(1))
        const double nx = slot == 0 ? G[0] : (slot == 1 ? G[3] : G[6]);
        const double ny = slot == 0 ? G[1] : (slot == 1 ? G[4] : G[7]);
        const double edfac = slot == 0 ? G[2] : (slot == 1 ? G[5] : G[8]);
        const double tx = -ny, ty = nx;
        if (kind == SIDE_INTERNAL) {
            const int ln = (sl >> (10 * slot)) & 1023;
            const int wdn = stwd[ln];
            if (wd == 0 && wdn == 0) continue;
            // the edge's "in" (NEDEL(1)) and "ex" (NEDEL(2)) elements as staged rows, their local
            // edge numbers and wet/dry flags: the traces are formed in in/ex order directly
            const int si = role ? ln : tid, sx = role ? tid : ln;
            const int Li = role ? nled : L, Lx = role ? L : nled;
            double Vi[3][DOF], Vx[3][DOF], hbi[3], hbx[3];
#pragma unroll
            for (int f = 0; f < 3; f++)
#pragma unroll
                for (int k = 0; k < DOF; k++) {
                    Vi[f][k] = stg[(f * DOF + k) * ns + si];
                    Vx[f][k] = stg[(f * DOF + k) * ns + sx];
                }
#pragma unroll
            for (int m = 0; m < 3; m++) {
                hbi[m] = stg[(3 * DOF + m) * ns + si];
                hbx[m] = stg[(3 * DOF + m) * ns + sx];
            }
            const int wd_in_el = role ? wdn : wd, wd_ex_el = role ? wd : wdn;
            // spherical factors of both sides at every point, loaded together
            double sfi[NEGP], sfx[NEGP];
#pragma unroll
            for (int I = 0; I < NEGP; I++) {
                sfi[I] = 1.0;
                sfx[I] = 1.0;
                if (SPH) {
                    sfi[I] = E.sfs[((size_t)((role ? 3 : 0) + slot) * NEGP + I) * ne + e];
                    sfx[I] = E.sfs[((size_t)((role ? 0 : 3) + slot) * NEGP + I) * ne + e];
                }
            }
            // the "in" side subtracts w*F and the "ex" side adds it; -(a*b) == (-a)*b exactly, so
            // the sign goes on the edge factor once and one update serves both roles
            const double edfs = role ? edfac : -edfac;
            double acc[3][DOF];
#pragma unroll
            for (int I = 0; I < NEGP; I++) {
                const int gi = I, gx = NEGP - 1 - I;
                const int gown = role ? gx : gi;
                const double* pi_ = tpe[Li][gi];
                const double* px_ = tpe[Lx][gx];
                // PHI_EDGE mode 1 is 1 (checked at set-up), so the sum from zero starts at mode 1's
                // coefficient with the same value
                const double ze_in = trace1<DOF>(Vi[0], pi_), qx_in = trace1<DOF>(Vi[1], pi_), qy_in = trace1<DOF>(Vi[2], pi_);
                const double ze_ex = trace1<DOF>(Vx[0], px_), qx_ex = trace1<DOF>(Vx[1], px_), qy_ex = trace1<DOF>(Vx[2], px_);
                const double hb_in = trace1<3>(hbi, pi_), hb_ex = trace1<3>(hbx, px_);
                const double sf_in = sfi[I], sf_ex = sfx[I];
                double f0, f1, f2;
                llf<NL>(ph, ze_in, ze_ex, qx_in, qx_ex, qy_in, qy_ex, hb_in, hb_ex, sf_in, sf_ex, nx, ny, ph.g, f0, f1, f2);
                double fo = f0, gm = f1, hm = f2;
                // dry-side treatment (src/internal_edge_hydro.F:212-302): 1 reflect into a dry
                // "in" element, 2 zero-gravity flux for a dry "in", 3 reflect into a dry "ex",
                // 4 zero-gravity flux for a dry "ex"; only what this element's side receives
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
                    llf<NL>(ph, a_zi, a_ze, a_xi, a_xe, a_yi, a_ye, a_hi, a_he, a_si, a_se, nx, ny, G, r0, r1, r2);
                    // mode 1: fo, ex momentum from the reflected flux, zero for the dry "in"
                    // mode 3: fo, in momentum from the reflected flux, zero for the dry "ex"
                    if (mode == 1) { fo = r0; gm = role ? r1 : 0.0; hm = role ? r2 : 0.0; }
                    else if (mode == 3) { fo = r0; gm = role ? 0.0 : r1; hm = role ? 0.0 : r2; }
                    else { gm = r1; hm = r2; }
                }
                // the slot sums start from zero: 0 + x == x for every x but -0, so the first
                // point's terms start them
                const double* eq = teq[L][gown];
#pragma unroll
                for (int k = 0; k < DOF; k++) {
                    const double ws = edfs * eq[k];
                    if (I == 0) {
                        acc[0][k] = ws * fo; acc[1][k] = ws * gm; acc[2][k] = ws * hm;
                    } else {
                        acc[0][k] = acc[0][k] + ws * fo; acc[1][k] = acc[1][k] + ws * gm; acc[2][k] = acc[2][k] + ws * hm;
                    }
                }
            }
#pragma unroll
            for (int f = 0; f < 3; f++)
#pragma unroll
                for (int k = 0; k < DOF; k++) acc_lds[1][f * DOF + k][tid] = acc_lds[1][f * DOF + k][tid] + acc[f][k];
        } else {
            if (wd == 0) continue;
            double xlen, AREAS;
            side_xlen_areas(E, ne, e, L, role, xlen, AREAS);
            double U[3][DOF], hb[3];
#pragma unroll
            for (int f = 0; f < 3; f++)
#pragma unroll
                for (int k = 0; k < DOF; k++) U[f][k] = stg[(f * DOF + k) * ns + tid];
#pragma unroll
            for (int m = 0; m < 3; m++) hb[m] = stg[(3 * DOF + m) * ns + tid];
#pragma unroll 1
            for (int I = 0; I < NEGP; I++) {
                const double* pv = tpe[L][I];
                const double ze_in = trace<DOF>(U[0], pv), qx_in = trace<DOF>(U[1], pv), qy_in = trace<DOF>(U[2], pv);
                const double hb_in = trace<3>(hb, pv);
                const double sf = SPH ? E.sfs[((size_t)slot * NEGP + I) * ne + e] : 1.0;
                double f0, f1, f2;
                if (kind == SIDE_LAND) {
                    const double qn_int = qx_in * nx + qy_in * ny, qt_int = qx_in * tx + qy_in * ty;
                    const double qn_ext = -qn_int, qt_ext = qt_int;
                    const double den = 1.0 / (nx * ty - ny * tx);
                    const double qx_ex = (ty * qn_ext - ny * qt_ext) * den;
                    const double qy_ex = (-tx * qn_ext + nx * qt_ext) * den;
                    llf<NL>(ph, ze_in, ze_in, qx_in, qx_ex, qy_in, qy_ex, hb_in, hb_in, sf, sf, nx, ny, ph.g, f0, f1, f2);
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
                    if ((ze_ex * ph.ifnlfa + hb_in) <= 0) ze_ex = fabs(hb_in) + sc.h0;
                    llf<NL>(ph, ze_in, ze_ex, qx_in, qx_in, qy_in, qy_in, hb_in, hb_in, sf, sf, nx, ny, ph.g, f0, f1, f2);
                } else {
                    // the specified normal flux at this point (flow_kernel); the reference forms the
                    // exterior momentum after every constituent, so the last, full sum is what counts
                    const double qn_ext = a.fo.fz[(size_t)a.irk * a.fo.fzn + (size_t)nb * NEGP + I];
                    const double qt_ext = 0.0;
                    const double qx_ex = -(ty * qn_ext - ny * qt_ext) / (nx * ty - ny * tx);
                    const double qy_ex = -(-tx * qn_ext + nx * qt_ext) / (nx * ty - ny * tx);
                    llf<NL>(ph, ze_in, ze_in, qx_in, qx_ex, qy_in, qy_ex, hb_in, hb_in, sf, sf, nx, ny, ph.g, f0, f1, f2);
                }
#pragma unroll
                for (int k = 0; k < DOF; k++) {
                    const double w = 2.0 * T.m_inv[k] / AREAS * xlen * pv[k] * T.wegp[I];
                    acc_lds[0][k][tid] = acc_lds[0][k][tid] - w * f0;
                    acc_lds[0][DOF + k][tid] = acc_lds[0][DOF + k][tid] - w * f1;
                    acc_lds[0][2 * DOF + k][tid] = acc_lds[0][2 * DOF + k][tid] - w * f2;
                }
            }
        }
    }
    // the element's own state and bathymetry modes, staged
    double U[3][DOF], hb[3];
#pragma unroll
    for (int f = 0; f < 3; f++)
#pragma unroll
        for (int k = 0; k < DOF; k++) U[f][k] = stg[(f * DOF + k) * ns + tid];
#pragma unroll
    for (int m = 0; m < 3; m++) hb[m] = stg[(3 * DOF + m) * ns + tid];
    // add_internal_flux: RHS = RHS(boundary edges) + slot sum
    double R[3][DOF];
#pragma unroll
    for (int f = 0; f < 3; f++)
#pragma unroll
        for (int k = 0; k < DOF; k++) R[f][k] = acc_lds[0][f * DOF + k][tid] + acc_lds[1][f * DOF + k][tid];

    // loads of the volume quadrature and the RK update, issued together before the quadrature's
    // arithmetic: the nodal wind and pressure (MET), the area-point spherical factors (SPH) and
    // the first RK base state
    double mcd[3], mwm[3], mwx[3], mwy[3], mpr[3];
    if (MET) {
#pragma unroll
        for (int v = 0; v < 3; v++) {
            const int n = E.nm[(size_t)v * ne + e];
            mcd[v] = a.mt.cd[n]; mwm[v] = a.mt.wm[n]; mwx[v] = a.mt.wx[n]; mwy[v] = a.mt.wy[n]; mpr[v] = a.mt.prn[n];
        }
    }
    double tipv[3] = {0.0, 0.0, 0.0};
    if (a.tip) {
#pragma unroll
        for (int v = 0; v < 3; v++) tipv[v] = a.tip[E.nm[(size_t)v * ne + e]];
    }
    double sfqv[NAGP];
#pragma unroll
    for (int I = 0; I < NAGP; I++) sfqv[I] = SPH ? E.sfacq[(size_t)I * ne + e] : 1.0;
    double jac[4];                       // DRDX, DSDX, DRDY, DSDY
#pragma unroll
    for (int j = 0; j < 4; j++) jac[j] = E.geo[(size_t)(9 + j) * ne + e];
    const bool pre0 = a.nprev > 0 && a.aprev[0] != 0.0;
    double B0[3][DOF];
    if (pre0) {
        const double* up = a.uprev[0];
#pragma unroll
        for (int f = 0; f < 3; f++)
#pragma unroll
            for (int k = 0; k < DOF; k++) B0[f][k] = up[(size_t)(f * DOF + k) * ne + e];
    }

    // ---- volume quadrature (RHS_DG_HYDRO)
    if (sc.rain > 0.0) {
#pragma unroll 1
        for (int I = 0; I < NAGP; I++)
#pragma unroll
            for (int k = 0; k < DOF; k++) R[0][k] = R[0][k] + T.srfac[I][k] * sc.rain;
    }
    if (wd != 0) {
        const double drdx = jac[0], dsdx = jac[1], drdy = jac[2], dsdy = jac[3];
        // wind stress at the element's vertices and the pressure gradient sums (MET)
        double mw[8];
        if (MET) {
            // nodal WSX2/WSY2/PR2 of this stage (src/met_forcing.F:507-509) and the
            // QP-independent sums PR2(Ni)*DPSIDX(i) of src/rhs_dg_hydro.F:117-122,366-372
            const double ramp = a.fo.ramp[(size_t)sid * a.fo.rows + row];
            const double drpsi[3] = {-1.0 / 2.0, 1.0 / 2.0, 0.0}, dspsi[3] = {-1.0 / 2.0, 0.0, 1.0 / 2.0};
            double pdx = 0.0, pdy = 0.0;
#pragma unroll
            for (int v = 0; v < 3; v++) {
                const double cd = mcd[v], wm = mwm[v];
                mw[v] = ramp * 0.001293 * cd * mwx[v] * wm;
                mw[3 + v] = ramp * 0.001293 * cd * mwy[v] * wm;
                const double pr = ramp * mpr[v];
                const double dpsidx = drpsi[v] * drdx + dspsi[v] * dsdx;
                const double dpsidy = drpsi[v] * drdy + dspsi[v] * dsdy;
                if (v == 0) { pdx = pr * dpsidx; pdy = pr * dpsidy; }
                else { pdx = pdx + pr * dpsidx; pdy = pdy + pr * dpsidy; }
            }
            mw[6] = pdx;
            mw[7] = pdy;
        }
        // tidal potential (src/rhs_dg_hydro.F:117-122,378-386): the sums DPSIDX(i)*TIP2(Ni) are
        // the same at every area point; RAMPDG multiplies them once more at each point
        double tdx = 0.0, tdy = 0.0, tramp = 0.0;
        if (a.tip) {
            const double drpsi[3] = {-1.0 / 2.0, 1.0 / 2.0, 0.0}, dspsi[3] = {-1.0 / 2.0, 0.0, 1.0 / 2.0};
#pragma unroll
            for (int v = 0; v < 3; v++) {
                const double dpsidx = drpsi[v] * drdx + dspsi[v] * dsdx;
                const double dpsidy = drpsi[v] * drdy + dspsi[v] * dsdy;
                if (v == 0) { tdx = dpsidx * tipv[v]; tdy = dpsidy * tipv[v]; }
                else { tdx = tdx + dpsidx * tipv[v]; tdy = tdy + dpsidy * tipv[v]; }
            }
            tramp = a.fo.ramp[(size_t)sid * a.fo.rows + row];
        }
        // hybrid friction factor (1 + (HBREAK/H_c)^FTHETA)^(FGAMMA/FTHETA) from the element-mean
        // depth (src/rhs_dg_hydro.F:306-309): the same at every area point, evaluated once and only
        // when the hybrid law is on
        double hybf = 1.0;
        if (ph.ifhybf == 1) {
            const double depth_c = hb[0] + U[0][0];
            const double fh_nl_c = NL ? 1.0 / depth_c : 1.0 / (ph.nleq * depth_c + ph.leq);
            hybf = pow(1.0 + pow(ph.hbreak * fh_nl_c, ph.ftheta), ph.fgamma / ph.ftheta);
        }
        // the P1 bathymetry's gradient: the linear modes' derivative tables are the same at every
        // area point (checked at set-up), so the sums are those of every point
        double dhbx = 0.0, dhby = 0.0;
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const double dpxk = T.drphi[0][k] * drdx + T.dsphi[0][k] * dsdx;
            const double dpyk = T.drphi[0][k] * drdy + T.dsphi[0][k] * dsdy;
            dhbx = dhbx + hb[k] * dpxk;
            dhby = dhby + hb[k] * dpyk;
        }
        // p=1: the three area points unrolled, so that their long dependent chains overlap (at p=2
        // the six points unrolled run slower)
#pragma unroll(P == 1 ? NAGP : 1)
        for (int I = 0; I < NAGP; I++) {
            const double* pa = T.phi_area[I];
            double ze = U[0][0], qx = U[1][0], qy = U[2][0];
#pragma unroll
            for (int k = 1; k < DOF; k++) {
                ze = ze + U[0][k] * pa[k];
                qx = qx + U[1][k] * pa[k];
                qy = qy + U[2][k] * pa[k];
            }
            // PHI_AREA of mode 1 is 1 (checked at set-up): the sum from zero starts at hb[0]
            double hbq = hb[0];
#pragma unroll
            for (int k = 1; k < 3; k++) hbq = hbq + hb[k] * pa[k];
            const double sfq = sfqv[I];
            const double depth = ze + hbq;
            // with NL (NLEQ = 1, LEQ = 0) the constant factors fold exactly
            const double f1 = NL ? 1.0 : ph.nleq + ph.leq * hbq;
            const double fx = NL ? qx * sfq : qx * f1 * sfq, fy = NL ? qy : qy * f1;
            const double fu = NL ? qx : ph.nleq * qx, fv = NL ? qy : ph.nleq * qy;
            const double fg = ph.nleqg * ze * wd;
            const double fh = NL ? 1.0 / depth : 1.0 / (ph.nleq * depth + ph.leq);
            const double u = qx * fh, v = qy * fh;
            const double huu = fu * u, hvv = fv * v, huv = fu * v;
            const double gh2 = NL ? fg * (0.5 * ze + hbq) : fg * (0.5 * ze + hbq) + ph.fg_l * ze;
            const double gx = (huu + gh2) * sfq, gy = huv;
            const double hx = huv * sfq, hy = hvv + gh2;
            double fric = sc.fric;
            if (ph.load_manning) {
                // DEPTH**(1.D0/3.D0) (src/rhs_dg_hydro.F:302), correctly rounded
                fric = ph.g * (nbar * nbar) / crm::cr_pow_third(depth);
                if (fric < sc.cf) fric = sc.cf;
            }
            double tau = fric;
            if (ph.iflinbf == 0) {
                const double umag = sqrt(u * u + v * v);
                tau = tau * umag * fh;
                if (ph.ifhybf == 1) tau = tau * hybf;
                const double cap = 0.9 * sc.dtdph;
                if (cap < tau) tau = cap;
            }
            double sx = -(tau * qx) + fg * dhbx * sfq + cori * qy;
            double sy = -(tau * qy) + fg * dhby - cori * qx;
            if (MET) {
                // wind stress and atmospheric pressure (src/rhs_dg_hydro.F:362-373)
                const double* ps = T.psi[I];
                const double wsx = mw[0] * ps[0] + mw[1] * ps[1] + mw[2] * ps[2];
                const double wsy = mw[3] * ps[0] + mw[4] * ps[1] + mw[5] * ps[2];
                const double fw = 1.0 / f1;
                sx = sx + (NL ? wsx : fw * wsx) - ph.g * sfq * depth * mw[6];
                sy = sy + (NL ? wsy : fw * wsy) - ph.g * depth * mw[7];
            }
            if (a.tip) {
                sx = sx + tramp * ph.g * depth * sfq * tdx;
                sy = sy + tramp * ph.g * depth * tdy;
            }
            R[1][0] = R[1][0] + T.srfac[I][0] * sx;
            R[2][0] = R[2][0] + T.srfac[I][0] * sy;
#pragma unroll
            for (int k = 1; k < DOF; k++) {
                const double dpxk = T.drphi[I][k] * drdx + T.dsphi[I][k] * dsdx;
                const double dpyk = T.drphi[I][k] * drdy + T.dsphi[I][k] * dsdy;
                const double xf = T.mw[I][k] * dpxk, yf = T.mw[I][k] * dpyk, sr = T.srfac[I][k];
                R[0][k] = R[0][k] + xf * fx + yf * fy;
                R[1][k] = R[1][k] + xf * gx + yf * gy + sr * sx;
                R[2][k] = R[2][k] + xf * hx + yf * hy + sr * sy;
            }
        }
    }
    // ---- SSP-RK update U(IRK+1) = sum_I ATVD(IRK,I) U(I) + BTVD(IRK,I) DT RHS(I)
    if (a.rout) {
#pragma unroll
        for (int f = 0; f < 3; f++)
#pragma unroll
            for (int k = 0; k < DOF; k++) a.rout[(size_t)(f * DOF + k) * ne + e] = R[f][k];
    }
    double W[3][DOF];
    const double bcur = a.bcur * sc.dtdp;
#pragma unroll
    for (int f = 0; f < 3; f++)
#pragma unroll
        for (int k = 0; k < DOF; k++) W[f][k] = 0.0;
#pragma unroll 1
    for (int i = 0; i < a.nprev; i++) {
        const double ai = a.aprev[i], bi = a.bprev[i] * sc.dtdp;
        if (i == 0 && pre0) {
#pragma unroll
            for (int f = 0; f < 3; f++)
#pragma unroll
                for (int k = 0; k < DOF; k++) W[f][k] = W[f][k] + ai * B0[f][k];
        } else if (ai != 0.0) {
            const double* up = a.uprev[i];
#pragma unroll
            for (int f = 0; f < 3; f++)
#pragma unroll
                for (int k = 0; k < DOF; k++) W[f][k] = W[f][k] + ai * up[(size_t)(f * DOF + k) * ne + e];
        }
        if (bi != 0.0) {
            const double* rp = a.rprev[i];
#pragma unroll
            for (int f = 0; f < 3; f++)
#pragma unroll
                for (int k = 0; k < DOF; k++) W[f][k] = W[f][k] + bi * rp[(size_t)(f * DOF + k) * ne + e];
        }
    }
#pragma unroll
    for (int f = 0; f < 3; f++)
#pragma unroll
        for (int k = 0; k < DOF; k++) W[f][k] = W[f][k] + a.acur * U[f][k] + bcur * R[f][k];

    if (WDFUSED) {
        const double D[3] = {E.dpv[e], E.dpv[(size_t)ne + e], E.dpv[2 * (size_t)ne + e]};
        const double dpe_min = fmin(fmin(D[0], D[1]), D[2]);
        if (a.wd_active && !wetdry_keeps<P>(T, W, D, wd, sc.h0)) a.wd_active[0] = 1;
        const int nw = wetdry_element<P>(T, W, D, hb, dpe_min, wd, sc.h0);
        a.wd_out[e] = (uint8_t)nw;
    }
#pragma unroll
    for (int f = 0; f < 3; f++)
#pragma unroll
        for (int k = 0; k < DOF; k++) a.uout[(size_t)(f * DOF + k) * ne + e] = W[f][k];
    if (a.cmout) {
#pragma unroll
        for (int f = 0; f < 3; f++) a.cmout[(size_t)f * ne + e] = W[f][0];
    }
}

// ------------------------------------------------------------------------------------------
// Tidal forcing of one step at the ocean boundary (src/ocean_edge_hydro.F:459-488): for every
// stage, ocean edge, constituent and end node the term EMO*RFF*COS(ARGJ - EFA) that the edge's
// elevation sums, with the cosine correctly rounded.  One thread per term; the stage kernels
// read the stage's terms.
// ------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) tide_kernel(Forcing fo, const uint16_t* __restrict__ osid,
                                                   double* __restrict__ tz) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= fo.nrk * fo.tzn) return;
    const int irk = i / fo.tzn, o = i - irk * fo.tzn;       // o = (edge*nbfr + jj)*2 + end
    const int jj = (o >> 1) % fo.nbfr, edge = (o >> 1) / fo.nbfr;
    const int sid = osid ? (int)osid[edge] : 0;
    const int row = table_step(fo) * fo.nrk + irk;
    if (fo.ets) {   // non-periodic: the stage's ramped end-node values (nbfr is 1)
        tz[i] = fo.ets[((size_t)sid * fo.rows + row) * fo.tzn + o];
        return;
    }
    const size_t t = ((size_t)sid * fo.rows + row) * fo.nbfr + jj;
    tz[i] = fo.emo[o] * fo.rff[t] * crm::cr_cos(fo.argj[t] - fo.efa[o]);
}

// ------------------------------------------------------------------------------------------
// Specified normal flux of one step (src/flow_edge_hydro.F:143-173): for every stage, flux edge
// and edge quadrature point Q_N_EXT = sum over constituents of QNAM_GP*RFF*COS(ARGJ - QNPH_GP),
// with the cosine correctly rounded.  For NFFR = 0 the one term has ARGJ = 0, RFF = RAMPDG and the
// end values of the time series (already ramped once, src/DG_hydro_timestep.F:188-195).
// ------------------------------------------------------------------------------------------
template <int P>
__global__ void __launch_bounds__(256) flow_kernel(Forcing fo, const uint16_t* __restrict__ fsid,
                                                   double* __restrict__ fz) {
    constexpr int NEGP = P + 2;
    const DevTables& T = tab<P>();
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= fo.nrk * fo.fzn) return;
    const int irk = i / fo.fzn, o = i - irk * fo.fzn;       // o = edge*NEGP + point
    const int edge = o / NEGP, I = o - edge * NEGP;
    const int sid = fsid ? (int)fsid[edge] : 0;
    const int row = table_step(fo) * fo.nrk + irk;
    const double xe = T.xegp[I];
    double qn = 0.0;
    for (int jj = 0; jj < fo.nffr; jj++) {
        const size_t t = ((size_t)sid * fo.rows + row) * fo.nffr + jj;
        const double argj = fo.fargj[t];
        const double rff = fo.frff[t];
        double q1, q2;
        if (fo.qts) {
            const size_t q = (((size_t)sid * fo.rows + row) * fo.nfeds + edge) * 2;
            q1 = fo.qts[q]; q2 = fo.qts[q + 1];
        } else {
            const size_t q = ((size_t)edge * fo.nffr + jj) * 2;
            q1 = fo.qnam[q]; q2 = fo.qnam[q + 1];
        }
        const size_t q = ((size_t)edge * fo.nffr + jj) * 2;
        const double p1 = fo.qnph[q], p2 = fo.qnph[q + 1];
        const double qnam_gp = 0.5 * (q1 + q2) + 0.5 * (q2 - q1) * xe;
        const double qnph_gp = 0.5 * (p1 + p2) + 0.5 * (p2 - p1) * xe;
        qn = qn + qnam_gp * rff * crm::cr_cos(argj - qnph_gp);
    }
    fz[i] = qn;
}

// ------------------------------------------------------------------------------------------
// TIDAL_POTENTIAL (src/tidal_potential.F:15-80) for every stage of one step: TIP2 at every node,
// with the cosines of the time-dependent arguments correctly rounded (COS(SFEA)**2 and
// SIN(2*SFEA) come from the host's libm).
// ------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) tip_kernel(Tip tp, Forcing fo) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= fo.nrk * tp.np) return;
    const int irk = i / tp.np, n = i - irk * tp.np;
    const int row = table_step(fo) * fo.nrk + irk;
    const double* r = tp.row + (size_t)row * tp.ntif * 3;
    double tip = 0.0;
    for (int j = 0; j < tp.ntif; j++) {
        const double argt = r[3 * j], tpmul = r[3 * j + 1], saltmul = r[3 * j + 2];
        const double sa = tp.salt_amp ? tp.salt_amp[(size_t)j * tp.np + n] : 0.0;
        const double sp = tp.salt_pha ? tp.salt_pha[(size_t)j * tp.np + n] : 0.0;
        const int na = tp.na[j];
        if (na == 1) {
            const double argtp = argt + 2.0 * tp.slam[n];
            const double argsalt = argt - sp;
            tip = tip + tpmul * tp.ccsfea[n] * crm::cr_cos(argtp) + saltmul * sa * crm::cr_cos(argsalt);
        }
        if (na == 2) {
            const double argtp = argt + tp.slam[n];
            const double argsalt = argt - sp;
            tip = tip + tpmul * tp.s2sfea[n] * crm::cr_cos(argtp) + saltmul * sa * crm::cr_cos(argsalt);
        }
    }
    tp.tip2[(size_t)irk * tp.np + n] = tip;
}

// ------------------------------------------------------------------------------------------
// SLOPELIMITER5 part 1 (src/slopelimiter.F:3290-3404): node min/max of the updated cell means.
// nminmax layout [6][np]: zmin zmax qxmin qxmax qymin qymax.
// ------------------------------------------------------------------------------------------
template <int P>
__global__ void __launch_bounds__(256) node_minmax_kernel(int np, int ne, const int* __restrict__ off,
                                                          const int* __restrict__ n2e,
                                                          const double* __restrict__ u,
                                                          double* __restrict__ mm) {
    constexpr int DOF = (P + 1) * (P + 2) / 2;
    const int n = xcd_tile(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
    if (n >= np) return;
    double zmin = 99999., zmax = -99999., xmin = 99999., xmax = -99999., ymin = 99999., ymax = -99999.;
    const int j1 = off[n + 1];
    // up to eight incidences per pass: all their indices, then all their cell means, are loaded
    // before any is used; past the node's last incidence the pass repeats that incidence (min and
    // max are exact and idempotent, so neither the order nor a repeat changes them)
    constexpr int B = 8;
    for (int j = off[n]; j < j1; j += B) {
        int ei[B];
#pragma unroll
        for (int q = 0; q < B; q++) ei[q] = n2e[min(j + q, j1 - 1)] >> 2;
        double z[B], x[B], y[B];
#pragma unroll
        for (int q = 0; q < B; q++) {
            z[q] = u[ei[q]];
            x[q] = u[(size_t)DOF * ne + ei[q]];
            y[q] = u[(size_t)2 * DOF * ne + ei[q]];
        }
#pragma unroll
        for (int q = 0; q < B; q++) {
            zmin = fmin(zmin, z[q]); xmin = fmin(xmin, x[q]); ymin = fmin(ymin, y[q]);
            zmax = fmax(zmax, z[q]); xmax = fmax(xmax, x[q]); ymax = fmax(ymax, y[q]);
        }
    }
    mm[n] = zmin; mm[(size_t)np + n] = zmax; mm[2 * (size_t)np + n] = xmin;
    mm[3 * (size_t)np + n] = xmax; mm[4 * (size_t)np + n] = ymin; mm[5 * (size_t)np + n] = ymax;
}

// ------------------------------------------------------------------------------------------
// SLOPELIMITER5 (src/slopelimiter.F:3290-3676) then WETDRY (src/wetdry.F), one thread per element.
// The node bounds of part 1 are formed per tile in LDS (LimTiles): the tile stages the updated cell
// means of its elements and of its vertex halo, each node of the tile takes the min/max over its
// incident elements in the node->element order (as node_minmax_kernel does, so the bounds are the
// same bit for bit), and each element then limits its modes 2-3 against the bounds of its nodes.
// With WETDRY the cell means come from the stage kernel's copy `cmean` [3][ne] (field stride
// cms = ne): WETDRY rewrites some cell means of u in place, which other tiles' halos must not see.
// Without it they are read from u itself (cms = DOF*ne), whose cell means the limiter keeps.
// Blocks of LIM_NT threads; the launch passes 8*(3*nsl + 6*nql) bytes of dynamic LDS.
// ------------------------------------------------------------------------------------------
// With WDADJ (the adjoint's reverse sweep; WETDRY false) the kernel writes nothing to u: each
// element's limited modes, formed in registers exactly as the forward formed them, go to the
// frozen-branch WETDRY reverse of that element (wetdry_adj_element), which turns the cotangent ub
// of WETDRY's output into that of its input and adds the vertex-depth cotangents to gdp.
constexpr int LIM_NT = 256;

template <int P>
__device__ void wetdry_adj_element(const DevTables& T, const double (&U)[3][3], const double (&D)[3], int wdflg,
                                   double H0, int e, int ne, double* __restrict__ ub, double* __restrict__ gdp);

template <int P, bool WETDRY, bool WDADJ>
__global__ void __launch_bounds__(LIM_NT) limit_wetdry_kernel(ElemData E, double* __restrict__ u,
                                                              const double* __restrict__ cmean, size_t cms,
                                                              LimTiles lt,
                                                              const uint8_t* __restrict__ wd_in,
                                                              uint8_t* __restrict__ wd_out,
                                                              int* __restrict__ wd_active,
                                                              double* __restrict__ ub, double* __restrict__ gdp) {
    static_assert(!(WETDRY && WDADJ), "the WETDRY reverse starts from the limiter's output alone");
    constexpr int DOF = (P + 1) * (P + 2) / 2;
    const DevTables& T = tab<P>();
    const int ne = E.ne;
    const int tid = threadIdx.x;
    const int tile = xcd_tile(blockIdx.x, gridDim.x);
    const int e = tile * LIM_NT + tid;
    extern __shared__ double lds[];
    const int nsl = lt.nsl, nql = lt.nql;
    double* const cm = lds;                  // [3][nsl] staged cell means
    double* const bnd = lds + 3 * nsl;       // [6][nql] zmin zmax qxmin qxmax qymin qymax
    // the element's own loads first, so that their latency overlaps the staging below (lanes past
    // the last element read element ne-1 and exit after the second barrier)
    const int ec = e < ne ? e : ne - 1;
    double U[3][DOF];
#pragma unroll
    for (int f = 0; f < 3; f++)
#pragma unroll
        for (int k = 1; k < DOF; k++) U[f][k] = u[(size_t)(f * DOF + k) * ne + ec];
    const int lq = lt.elq[ec];
    double D[3] = {0.0, 0.0, 0.0}, h0 = 0.0;
    int wdi = 0;
    if (WETDRY || WDADJ) {
#pragma unroll
        for (int v = 0; v < 3; v++) D[v] = E.dpv[(size_t)v * ne + ec];
        const int sid = E.sid ? (int)E.sid[ec] : 0;
        h0 = E.scen[sid].h0;
        wdi = wd_in[ec];
    }
    {
        // the tile's cell means and its vertex halo's, two entries per lane per pass with all their
        // loads issued before any is stored
        const int h0t = lt.vhoff[tile], nh = lt.vhoff[tile + 1] - h0t;
        for (int j = tid; j < LIM_NT + nh; j += 2 * LIM_NT) {
            const int j2 = j + LIM_NT;
            const bool has2 = j2 < LIM_NT + nh;
            int g = j < LIM_NT ? tile * LIM_NT + j : lt.vhidx[h0t + j - LIM_NT];
            int g2 = has2 ? lt.vhidx[h0t + j2 - LIM_NT] : ne;
            const bool ok = g < ne, ok2 = g2 < ne;
            g = ok ? g : 0;
            g2 = ok2 ? g2 : 0;
            const double z = cmean[g], x = cmean[cms + g], y = cmean[2 * cms + g];
            const double z2 = cmean[g2], x2 = cmean[cms + g2], y2 = cmean[2 * cms + g2];
            if (ok) { cm[j] = z; cm[nsl + j] = x; cm[2 * nsl + j] = y; }
            if (ok2) { cm[j2] = z2; cm[nsl + j2] = x2; cm[2 * nsl + j2] = y2; }
        }
    }
    __syncthreads();
    {
        const int q0 = lt.qoff[tile], nq = lt.qoff[tile + 1] - q0;
        for (int q = tid; q < nq; q += LIM_NT) {
            double zmin = 99999., zmax = -99999., xmin = 99999., xmax = -99999., ymin = 99999., ymax = -99999.;
            const int i1 = lt.qinc_off[q0 + q + 1];
            // up to eight incidences per pass, their rows loaded together; past the node's last
            // incidence the pass repeats it (min and max are idempotent, so a repeat changes nothing)
            constexpr int B = 8;
            for (int i = lt.qinc_off[q0 + q]; i < i1; i += B) {
                int r[B];
#pragma unroll
                for (int b = 0; b < B; b++) r[b] = lt.qinc[min(i + b, i1 - 1)];
#pragma unroll
                for (int b = 0; b < B; b++) {
                    const double z = cm[r[b]], x = cm[nsl + r[b]], y = cm[2 * nsl + r[b]];
                    zmin = fmin(zmin, z); xmin = fmin(xmin, x); ymin = fmin(ymin, y);
                    zmax = fmax(zmax, z); xmax = fmax(xmax, x); ymax = fmax(ymax, y);
                }
            }
            bnd[q] = zmin; bnd[nql + q] = zmax; bnd[2 * nql + q] = xmin;
            bnd[3 * nql + q] = xmax; bnd[4 * nql + q] = ymin; bnd[5 * nql + q] = ymax;
        }
    }
    __syncthreads();
    if (e >= ne) return;
#pragma unroll
    for (int f = 0; f < 3; f++) U[f][0] = cm[f * nsl + tid];
    // the limiter keeps the cell means; wetting and drying changes them only for some elements,
    // so a cell mean is written back only when its bits changed
    const double c0in[3] = {U[0][0], U[1][0], U[2][0]};
    double bmin[3][3], bmax[3][3];
    {
        const int q[3] = {lq & 1023, (lq >> 10) & 1023, (lq >> 20) & 1023};
#pragma unroll
        for (int iv = 0; iv < 3; iv++)
#pragma unroll
            for (int c = 0; c < 3; c++) {
                bmin[iv][c] = bnd[(2 * iv) * nql + q[c]];
                bmax[iv][c] = bnd[(2 * iv + 1) * nql + q[c]];
            }
    }
    {
#pragma unroll
        for (int iv = 0; iv < 3; iv++) {
            const double* vmin = bmin[iv];
            const double* vmax = bmax[iv];
            const double c0 = U[iv][0];
            double v[3];
#pragma unroll
            for (int c = 0; c < 3; c++) {
                double t = c0;
#pragma unroll
                for (int kk = 1; kk < 3; kk++) t = t + T.phi_corner[c][kk] * U[iv][kk];
                v[c] = fmax(fmin(t, vmax[c]), vmin[c]);
            }
#pragma unroll
            for (int ll = 0; ll < 3; ll++) {
                const double sumloc = (v[0] + v[1] + v[2]) / 3.0;
                double sumdif = (sumloc - c0) * 3.0;
                const double signdif = copysign(1.0, sumdif);
                double dif[3];
#pragma unroll
                for (int c = 0; c < 3; c++) dif[c] = (v[c] - c0) * signdif;
                int kdp = (dif[0] > 0) + (dif[1] > 0) + (dif[2] > 0);
#pragma unroll
                for (int k = 0; k < 3; k++) {
                    double redfac;
                    if (dif[k] > 0) {
                        // (sumdif*signdif)/DIV with DIV = kdp, which is 1, 2 or 3 here: x/1 and
                        // x/2 are x and x*0.5 exactly, so only DIV = 3 needs a division
                        const double a = sumdif * signdif;
                        redfac = kdp == 1 ? a : (kdp == 2 ? a * 0.5 : a / 3.0);
                        kdp = kdp - 1;
                    } else {
                        redfac = 0;
                    }
                    const double redmax = (signdif > 0) ? v[k] - vmin[k] : vmax[k] - v[k];
                    redfac = fmin(redfac, redmax);
                    sumdif = sumdif - redfac * signdif;
                    v[k] = v[k] - redfac * signdif;
                }
            }
            U[iv][1] = -(1.0 / 6.0 * (v[0] + v[1])) + 1.0 / 3.0 * v[2];
            U[iv][2] = -(0.5 * v[0]) + 0.5 * v[1];
        }
    }
    if (WDADJ) {
        const double U3[3][3] = {{U[0][0], U[0][1], U[0][2]}, {U[1][0], U[1][1], U[1][2]}, {U[2][0], U[2][1], U[2][2]}};
        wetdry_adj_element<P>(T, U3, D, wdi, h0, e, ne, ub, gdp);
        return;
    }
    if (WETDRY) {
        const double c13 = 1.0 / 3.0, c16 = 1.0 / 6.0;
        const double hb[3] = {c13 * (D[0] + D[1] + D[2]), -(c16 * (D[0] + D[1])) + c13 * D[2], -(0.5 * D[0]) + 0.5 * D[1]};
        const double dpe_min = fmin(fmin(D[0], D[1]), D[2]);
        if (wd_active && !wetdry_keeps<P>(T, U, D, wdi, h0)) wd_active[0] = 1;
        wd_out[e] = (uint8_t)wetdry_element<P>(T, U, D, hb, dpe_min, wdi, h0);
    }
#pragma unroll
    for (int f = 0; f < 3; f++) {
        if (WETDRY && __double_as_longlong(U[f][0]) != __double_as_longlong(c0in[f]))
            u[(size_t)(f * DOF) * ne + e] = U[f][0];
#pragma unroll
        for (int k = 1; k < DOF; k++) u[(size_t)(f * DOF + k) * ne + e] = U[f][k];
    }
}

// ------------------------------------------------------------------------------------------
// modal2nodal (src/write_results.F:14-174) + NODECODE from the wet flags.  Each node's incident
// elements are summed in the order of its CSR row (the host's ELETAB order when the host gave it).
// With etamax set, ETAMAX is raised as the reference raises it: after every wet element's
// contribution to ETA2, so it also sees the partial sums (src/write_results.F:164).
// ------------------------------------------------------------------------------------------
template <int P>
__global__ void __launch_bounds__(256) modal2nodal_kernel(int np, ElemData E, const int* __restrict__ off,
                                                          const int* __restrict__ n2e,
                                                          const double* __restrict__ u,
                                                          const uint8_t* __restrict__ wd, double nleq,
                                                          double leq, double* __restrict__ out,
                                                          int* __restrict__ nodecode,
                                                          double* __restrict__ etamax) {
    constexpr int DOF = (P + 1) * (P + 2) / 2;
    const DevTables& T = tab<P>();
    const int ne = E.ne;
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= np) return;
    double asum = 0;
    int nc = 0;
    for (int j = off[n]; j < off[n + 1]; j++) {
        const int e = n2e[j] >> 2;
        if (wd[e] == 0) continue;
        nc = 1;
        const double X0 = E.xv[e], X1 = E.xv[(size_t)ne + e], X2 = E.xv[2 * (size_t)ne + e];
        const double Y0 = E.yv[e], Y1 = E.yv[(size_t)ne + e], Y2 = E.yv[2 * (size_t)ne + e];
        const double areas = (X0 - X2) * (Y1 - Y2) + (X2 - X1) * (Y0 - Y2);
        asum = asum + 0.5 * areas;
    }
    double e2 = 0.0, u2 = 0.0, v2 = 0.0;
    double emax = etamax ? etamax[n] : 0.0;
    for (int j = off[n]; j < off[n + 1]; j++) {
        const int e = n2e[j] >> 2, c = n2e[j] & 3;
        if (wd[e] == 0) continue;
        const double X0 = E.xv[e], X1 = E.xv[(size_t)ne + e], X2 = E.xv[2 * (size_t)ne + e];
        const double Y0 = E.yv[e], Y1 = E.yv[(size_t)ne + e], Y2 = E.yv[2 * (size_t)ne + e];
        const double areas = (X0 - X2) * (Y1 - Y2) + (X2 - X1) * (Y0 - Y2);
        const double D0 = E.dpv[e], D1 = E.dpv[(size_t)ne + e], D2 = E.dpv[2 * (size_t)ne + e];
        const double hb[3] = {1.0 / 3.0 * (D0 + D1 + D2), -(1.0 / 6.0 * (D0 + D1)) + 1.0 / 3.0 * D2, -(0.5 * D0) + 0.5 * D1};
        double zd = u[e], xd = u[(size_t)DOF * ne + e], yd = u[(size_t)2 * DOF * ne + e], hd = hb[0];
#pragma unroll
        for (int kk = 1; kk < DOF; kk++) {
            const double pc = T.phi_corner[c][kk];
            zd = zd + pc * u[(size_t)kk * ne + e];
            xd = xd + pc * u[(size_t)(DOF + kk) * ne + e];
            yd = yd + pc * u[(size_t)(2 * DOF + kk) * ne + e];
            hd = hd + pc * (kk < 3 ? hb[kk] : 0.0);
        }
        const double area = 0.50 * areas / asum;
        const double depth = zd + hd;
        const double fh = 1.0 / (nleq * depth + leq);
        e2 = e2 + area * zd;
        if (emax < e2) emax = e2;
        u2 = u2 + area * xd * fh;
        v2 = v2 + area * yd * fh;
    }
    out[n] = e2; out[(size_t)np + n] = u2; out[2 * (size_t)np + n] = v2;
    nodecode[n] = nc;
    if (etamax) etamax[n] = emax;
}

// ------------------------------------------------------------------------------------------
// NWS=8 node kernel: HollandGet's per-node profile (src/wind.F:2062-2131) and the wind speed and
// Garratt drag of MET_FORCING (src/met_forcing.F:488-500, src/wind.F:135-141), at TIME_A of the
// current step.  The stage's ramp is applied in the stage kernel.
// ------------------------------------------------------------------------------------------
__global__ void __launch_bounds__(256) met_kernel(Met m, const int* __restrict__ step, int step_imm) {
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= m.np) return;
    const int sid = m.nsid ? (int)m.nsid[n] : 0;
    const double* r = m.row + ((size_t)sid * m.rows + (step ? *step : step_imm)) * MET_ROW;
    const double re = 6378206.4;
    const double deg2rad = 3.141592653589793 / 180.0;
    const double lat = r[MR_LAT], lon = r[MR_LON], spd = r[MR_SPD], B = r[MR_B];
    const double dx = (m.xc[n] - lon) * deg2rad, dy = (m.yc[n] - lat) * deg2rad;
    const double theta = atan2(dy, dx);
    const double s1 = sin(dy / 2.0), s2 = sin(dx / 2.0);
    const double rad = re * (2.0 * asin(sqrt(s1 * s1 + r[MR_COSLAT] * m.cosy[n] * (s2 * s2))));
    const double cor = m.cor[n];
    const double rb = pow(r[MR_RMW] * 1000.0 / rad, B);
    const double press = (r[MR_CPRESS] + r[MR_DEFICIT] * exp(-rb)) / r[MR_RHOG];
    double vr = sqrt(rb * exp(1.0 - rb) * (spd * spd) + (rad * rad) * (cor * cor) / 4.0) - rad * cor / 2.0;
    const double tsx = (fabs(vr) / spd) * r[MR_TVX], tsy = (fabs(vr) / spd) * r[MR_TVY];
    vr = vr * r[MR_WMULT];
    double wx = -(vr * sin(theta)), wy = vr * cos(theta);
    wx = wx * r[MR_BLADJ]; wy = wy * r[MR_BLADJ];
    wx = wx * 0.88; wy = wy * 0.88;
    wx = wx + tsx; wy = wy + tsy;
    const double wmag = sqrt(wx * wx + wy * wy);
    double cd = 0.001 * (0.75 + 0.067 * wmag);
    if (cd > 0.0035) cd = 0.0035;
    m.wx[n] = wx; m.wy[n] = wy; m.wm[n] = wmag; m.cd[n] = cd; m.prn[n] = press;
}

// Device step counter used by the forcing-table lookups of captured steps (a kernel, not a memset,
// so that the sequence can be captured into a HIP graph).
__global__ void step_counter_kernel(int* __restrict__ step, int set, int value) {
    if (threadIdx.x == 0 && blockIdx.x == 0) *step = set ? value : *step + 1;
}

// Water volume per element: 0.5*AREAS*(ze_1 + hb_1), reduced per block.
template <int P>
__global__ void __launch_bounds__(256) mass_kernel(ElemData E, const double* __restrict__ u,
                                                   double* __restrict__ partial) {
    __shared__ double sh[256];
    const int ne = E.ne;
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    double m = 0.0;
    if (e < ne) {
        const double X0 = E.xv[e], X1 = E.xv[(size_t)ne + e], X2 = E.xv[2 * (size_t)ne + e];
        const double Y0 = E.yv[e], Y1 = E.yv[(size_t)ne + e], Y2 = E.yv[2 * (size_t)ne + e];
        const double D0 = E.dpv[e], D1 = E.dpv[(size_t)ne + e], D2 = E.dpv[2 * (size_t)ne + e];
        const double areas = (X0 - X2) * (Y1 - Y2) + (X2 - X1) * (Y0 - Y2);
        m = 0.5 * areas * (u[e] + 1.0 / 3.0 * (D0 + D1 + D2));
    }
    sh[threadIdx.x] = m;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) sh[threadIdx.x] += sh[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = sh[0];
}

// out = ((0 + v[0]) + v[1]) + ... + v[n-1], the left-to-right order of a host loop.
__global__ void __launch_bounds__(64) sum_in_order_kernel(const double* __restrict__ v, int n,
                                                          double* __restrict__ out) {
    if (threadIdx.x != 0) return;
    double s = 0.0;
    for (int i = 0; i < n; i++) s += v[i];
    *out = s;
}

}  // namespace dg
