// Device-side data structures and the numerical flux of the DG hydro step (gfx950, FP64).
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>

#include "dg_crmath.hpp"

namespace dg {

constexpr int KMAXDOF = 6;
constexpr int KMAXAGP = 6;
constexpr int KMAXEGP = 4;
constexpr int KMAXRK = 8;

// Per-slot side descriptor packed 7 bits per slot into one int32 per element:
//   bits 0-1 local edge L (edge L runs from vertex L+1 to L+2, mod 3)
//   bit  2   role: 0 = this element is NEDEL(1) ("in"), 1 = NEDEL(2) ("ex")
//   bits 3-4 kind: 0 internal, 1 elevation-specified (ocean), 2 no-normal-flow (land), 3 flux
//   bits 5-6 neighbour's local edge (internal edges)
// Slots are stored in the reference's accumulation order: boundary edges first (ocean, land,
// flux, each in its list order), then internal edges by global edge number.
enum SideKind { SIDE_INTERNAL = 0, SIDE_OCEAN = 1, SIDE_LAND = 2, SIDE_FLOW = 3 };

__host__ __device__ inline int side_bits(int info, int slot) { return (info >> (7 * slot)) & 127; }

// Reference-element tables in constant memory (filled from RefTables at init).
struct DevTables {
    double phi_area[KMAXAGP][KMAXDOF];
    double drphi[KMAXAGP][KMAXDOF];
    double dsphi[KMAXAGP][KMAXDOF];
    double mw[KMAXAGP][KMAXDOF];
    double srfac[KMAXAGP][KMAXDOF];
    double phi_edge[3][KMAXEGP][KMAXDOF];
    double edgeq[3][KMAXEGP][KMAXDOF];
    double phi_corner[3][KMAXDOF];
    double m_inv[KMAXDOF];
    double wegp[KMAXEGP];
    double xegp[KMAXEGP];
    double psi[KMAXAGP][3];   // linear nodal basis at the area points (wind stress interpolation)
};

// Physics switches and constants shared by every scenario of a batch
// (src/dgswem.F:131-170 derivations).
struct Phys {
    double g, nleq, leq, nleqg, fg_l;
    double hbreak, ftheta, fgamma;
    int ifnlfa, iflinbf, ifhybf, load_manning;
};

// Per-scenario scalars (one entry per ensemble member; one entry for a single run).
struct Scen {
    double dtdp;       // DTDP
    double dtdph;      // 1/DTDP (friction cap)
    double h0;         // H0
    double cf;         // CF
    double fric;       // uniform FRIC_EL (without Manning)
    double cori;       // uniform Coriolis (when cori_el == nullptr)
    double rain;       // rain source rate (rainfall = 1) or 0
    double pad;
};

// Element-major structure-of-arrays mesh data on the device.
struct ElemData {
    int ne;
    const uint16_t* __restrict__ sid;   // [ne] scenario of each element, nullptr for one scenario
    const Scen* __restrict__ scen;      // [nscen]
    const double* __restrict__ xv;      // [3][ne] vertex x
    const double* __restrict__ yv;      // [3][ne] vertex y
    const double* __restrict__ dpv;     // [3][ne] vertex depth DP
    const double* __restrict__ nbar;    // [ne] mean nodal Manning n, or nullptr
    const double* __restrict__ cori_el; // [ne] or nullptr
    const double* __restrict__ sfacq;   // [nagp][ne] spherical factor at area points, or nullptr
    const double* __restrict__ sfaced;  // [3][negp][ne] spherical factor at edge points, or nullptr
    const double* __restrict__ sfs;     // [2][3 slots][negp][ne] the same in slot order: [0] own side
                                        // at the edge's point I, [1] the internal neighbour's side
    const int* __restrict__ sinfo;      // [ne] packed side descriptors
    const int* __restrict__ snbr;       // [3][ne] neighbour element / boundary-list index per slot
    const int* __restrict__ nm;         // [3][ne] 0-based vertex nodes
    const double* __restrict__ geo;     // [13][ne] per slot s: rows 3s..3s+2 the side's unit normal
                                        // (NX, NY) and XLEN/AREAS; rows 9..12 DRDX, DSDX, DRDY, DSDY
};

// Boundary forcing.  Per-stage scalars come from host-built tables [scenario][row][constituent];
// the row of the current stage is (*step)*nrk + irk, with the step counter kept on the device so
// that a captured graph of several steps can be replayed.
struct Forcing {
    int nbfr, nffr, rows, nrk;
    const int* __restrict__ step;       // device step counter within the current tables (graph
                                        // replays), or nullptr: the step is step_imm
    int step_imm;
    const double* __restrict__ emo;     // [needs][nbfr][2]
    const double* __restrict__ efa;     // [needs][nbfr][2]
    const double* __restrict__ qnam;    // [nfeds][nffr][2]
    const double* __restrict__ qnph;    // [nfeds][nffr][2]
    const double* __restrict__ argj;    // [nstage][nbfr]
    const double* __restrict__ rff;     // [nstage][nbfr]
    const double* __restrict__ fargj;   // [nstage][nffr]
    const double* __restrict__ frff;    // [nstage][nffr]
    const double* __restrict__ ramp;    // [scenario][row] RAMPDG of each stage (meteorological forcing)
    // tidal elevation terms EMO*RFF*COS(ARGJ-EFA) of the current step, [stage][needs][nbfr][2]
    // (tzn = needs*nbfr*2 per stage), written by tide_kernel before the step's stages
    const double* __restrict__ tz;
    int tzn;
    // specified normal flux Q_N_EXT at every flux-edge quadrature point of the current step,
    // [stage][nfeds][negp] (fzn = nfeds*negp per stage), written by flow_kernel
    const double* __restrict__ fz;
    int fzn;
    // non-periodic boundaries (NBFR = 0, NFFR = 0): the ramped end-node values of every stage row,
    // [scenario*rows + row][needs][2] (elevation; nbfr is 1) and [scenario*rows + row][nfeds][2]
    // (normal flux; nffr is 1), or nullptr
    const double* __restrict__ ets;
    const double* __restrict__ qts;
    int nfeds;
};

// Tidal potential (NTIP != 0): per-stage-row constituent terms and node data for tip_kernel, which
// writes TIP2 of every stage at the nodes once per step.
struct Tip {
    int np, ntif;
    const double* __restrict__ row;      // [scenario*rows + row][ntif][3]: ARGT, TPMUL, SALTMUL
    const int* __restrict__ na;          // [ntif]: 1 semidiurnal, 2 diurnal, else no term
    const double* __restrict__ slam;     // [np] longitude (rad)
    const double* __restrict__ ccsfea;   // [np] COS(SFEA)**2
    const double* __restrict__ s2sfea;   // [np] SIN(2*SFEA)
    const double* __restrict__ salt_amp; // [ntif][np], or nullptr (NTIP = 1: zero)
    const double* __restrict__ salt_pha; // [ntif][np] radians
    double* __restrict__ tip2;           // [stage][np]
};

// Step index within the current forcing tables.
__device__ __forceinline__ int table_step(const Forcing& f) { return f.step ? *f.step : f.step_imm; }


// NWS=8 meteorological forcing at the nodes.  The node kernel evaluates the Holland profile once
// per step at TIME_A (the reference evaluates it at every stage with the same TIME_A) and keeps the
// ramp-free factors; the stage kernel forms WSX2 = RampMete*0.001293*Cd*Wx*|W| and
// PR2 = RampMete*P with the stage's ramp, in the reference's operation order.
struct Met {
    int np, rows;                        // rows: steps per scenario in the storm table
    const double* __restrict__ row;     // [scenario][rows][MET_ROW] storm scalars per step
    const uint16_t* __restrict__ nsid;  // [np] scenario of each node, nullptr for one scenario
    const double* __restrict__ xc;      // [np] longitude (deg)
    const double* __restrict__ yc;      // [np] latitude (deg)
    const double* __restrict__ cosy;    // [np] cos(latitude)
    const double* __restrict__ cor;     // [np] 2 omega sin(latitude)
    double* __restrict__ wx;            // [np] 10-m wind (m/s), x
    double* __restrict__ wy;            // [np] y
    double* __restrict__ wm;            // [np] |W|
    double* __restrict__ cd;            // [np] drag coefficient
    double* __restrict__ prn;           // [np] surface pressure (m of water)
};

// Block-local node bounds of the slope limiter.  Tile t (elements [t*NT, t*NT+NT)) stages the
// cell means of its own elements (rows 0..NT-1) and of its vertex halo -- the elements outside
// the tile that share a node with it, vhidx[vhoff[t]..vhoff[t+1]) (rows NT..) -- and forms the
// bounds of its nodes qnode[qoff[t]..qoff[t+1]) from their incident elements, listed in the
// node->element order as staged rows qinc[qinc_off[q]..qinc_off[q+1]).  elq packs each element's
// three tile-local node numbers (10 bits each).
struct LimTiles {
    const int* __restrict__ vhoff;
    const int* __restrict__ vhidx;
    const int* __restrict__ qoff;
    const int* __restrict__ qinc_off;
    const uint16_t* __restrict__ qinc;
    const int* __restrict__ elq;
    int nsl, nql;                       // LDS rows: staged elements, nodes (largest tile)
};

// Local Lax-Friedrichs flux (src/numerical_flux.F:750-1022), operation order kept.
__device__ __forceinline__ void llf_flux(const Phys& P, double ze_in, double ze_ex, double qx_in,
                                         double qx_ex, double qy_in, double qy_ex, double hb_in,
                                         double hb_ex, double sf_in, double sf_ex, double nx,
                                         double ny, double G, double& f0, double& f1, double& f2) {
    const double NLEQ = P.nleq, LEQ = P.leq, FG_L = P.fg_l;
    const double NLEQG = G * NLEQ;
    const double j1 = ze_ex - ze_in, j2 = qx_ex - qx_in, j3 = qy_ex - qy_in;
    const double ht_in = ze_in * NLEQ + hb_in;
    double fnl = NLEQ + LEQ * hb_in;
    const double fx1_in = qx_in * fnl * sf_in, fy1_in = qy_in * fnl;
    double fu = NLEQ * qx_in, fv = NLEQ * qy_in, fg = NLEQG * ze_in;
    double fh = 1.0 / (NLEQ * ht_in + LEQ);
    const double u_in = qx_in * fh, v_in = qy_in * fh;
    double huu = fu * u_in, hvv = fv * v_in, huv = fu * v_in;
    double gh2 = fg * (0.5 * ze_in + hb_in) + FG_L * ze_in;
    const double fx2_in = (huu + gh2) * sf_in, fy2_in = huv;
    const double fx3_in = huv * sf_in, fy3_in = hvv + gh2;
    const double ht_ex = ze_ex * NLEQ + hb_ex;
    fnl = NLEQ + LEQ * hb_ex;
    const double fx1_ex = qx_ex * fnl * sf_ex, fy1_ex = qy_ex * fnl;
    fu = NLEQ * qx_ex; fv = NLEQ * qy_ex; fg = NLEQG * ze_ex;
    fh = 1.0 / (NLEQ * ht_ex + LEQ);
    const double u_ex = qx_ex * fh, v_ex = qy_ex * fh;
    huu = fu * u_ex; hvv = fv * v_ex; huv = fu * v_ex;
    gh2 = fg * (0.5 * ze_ex + hb_ex) + FG_L * ze_ex;
    const double fx2_ex = (huu + gh2) * sf_ex, fy2_ex = huv;
    const double fx3_ex = huv * sf_ex, fy3_ex = hvv + gh2;
    const double a1 = 0.5 * ((fx1_in + fx1_ex) * nx + (fy1_in + fy1_ex) * ny);
    const double a2 = 0.5 * ((fx2_in + fx2_ex) * nx + (fy2_in + fy2_ex) * ny);
    const double a3 = 0.5 * ((fx3_in + fx3_ex) * nx + (fy3_in + fy3_ex) * ny);
    const double un_in = (u_in * nx + v_in * ny) * NLEQ;
    double nxs = nx * sf_in;
    const double c_in = sqrt(G * ht_in * (ny * ny + nxs * nxs));
    const double un_ex = (u_ex * nx + v_ex * ny) * NLEQ;
    nxs = nx * sf_ex;
    const double c_ex = sqrt(G * ht_ex * (ny * ny + nxs * nxs));
    double em = fabs(un_in + c_in);
    em = fmax(em, fabs(un_in));
    em = fmax(em, fabs(un_in - c_in));
    em = fmax(em, fabs(un_ex + c_ex));
    em = fmax(em, fabs(un_ex));
    em = fmax(em, fabs(un_ex - c_ex));
    const double nx2 = NLEQ + LEQ * nx * nx, ny2 = NLEQ + LEQ * ny * ny, nxy = nx * ny;
    f0 = a1 - 0.5 * em * (j1);
    f1 = a2 - 0.5 * em * (j2 * nx2 + LEQ * j3 * nxy);
    f2 = a3 - 0.5 * em * (j3 * ny2 + LEQ * j2 * nxy);
}

// llf_flux for the nonlinear equations (NOLICAT != 0: NLEQ = 1, LEQ = 0) with those constants
// folded by hand: x*NLEQ = x, NLEQ + LEQ*h = 1 and y + LEQ*z = y hold exactly for finite operands,
// so the values equal llf_flux's bit for bit (up to the sign of an exact zero).
__device__ __forceinline__ void llf_flux_nl(double ze_in, double ze_ex, double qx_in, double qx_ex,
                                            double qy_in, double qy_ex, double hb_in, double hb_ex,
                                            double sf_in, double sf_ex, double nx, double ny, double G,
                                            double& f0, double& f1, double& f2) {
    const double j1 = ze_ex - ze_in, j2 = qx_ex - qx_in, j3 = qy_ex - qy_in;
    const double ht_in = ze_in + hb_in;
    const double fx1_in = qx_in * sf_in, fy1_in = qy_in;
    double fg = G * ze_in;
    double fh = 1.0 / ht_in;
    const double u_in = qx_in * fh, v_in = qy_in * fh;
    double huu = qx_in * u_in, hvv = qy_in * v_in, huv = qx_in * v_in;
    double gh2 = fg * (0.5 * ze_in + hb_in);
    const double fx2_in = (huu + gh2) * sf_in, fy2_in = huv;
    const double fx3_in = huv * sf_in, fy3_in = hvv + gh2;
    const double ht_ex = ze_ex + hb_ex;
    const double fx1_ex = qx_ex * sf_ex, fy1_ex = qy_ex;
    fg = G * ze_ex;
    fh = 1.0 / ht_ex;
    const double u_ex = qx_ex * fh, v_ex = qy_ex * fh;
    huu = qx_ex * u_ex; hvv = qy_ex * v_ex; huv = qx_ex * v_ex;
    gh2 = fg * (0.5 * ze_ex + hb_ex);
    const double fx2_ex = (huu + gh2) * sf_ex, fy2_ex = huv;
    const double fx3_ex = huv * sf_ex, fy3_ex = hvv + gh2;
    const double a1 = 0.5 * ((fx1_in + fx1_ex) * nx + (fy1_in + fy1_ex) * ny);
    const double a2 = 0.5 * ((fx2_in + fx2_ex) * nx + (fy2_in + fy2_ex) * ny);
    const double a3 = 0.5 * ((fx3_in + fx3_ex) * nx + (fy3_in + fy3_ex) * ny);
    const double un_in = u_in * nx + v_in * ny;
    double nxs = nx * sf_in;
    const double c_in = sqrt(G * ht_in * (ny * ny + nxs * nxs));
    const double un_ex = u_ex * nx + v_ex * ny;
    nxs = nx * sf_ex;
    const double c_ex = sqrt(G * ht_ex * (ny * ny + nxs * nxs));
    double em = fabs(un_in + c_in);
    em = fmax(em, fabs(un_in));
    em = fmax(em, fabs(un_in - c_in));
    em = fmax(em, fabs(un_ex + c_ex));
    em = fmax(em, fabs(un_ex));
    em = fmax(em, fabs(un_ex - c_ex));
    f0 = a1 - 0.5 * em * (j1);
    f1 = a2 - 0.5 * em * (j2);
    f2 = a3 - 0.5 * em * (j3);
}

template <bool NL>
__device__ __forceinline__ void llf(const Phys& P, double ze_in, double ze_ex, double qx_in, double qx_ex,
                                    double qy_in, double qy_ex, double hb_in, double hb_ex, double sf_in,
                                    double sf_ex, double nx, double ny, double G, double& f0, double& f1,
                                    double& f2) {
    if constexpr (NL)
        llf_flux_nl(ze_in, ze_ex, qx_in, qx_ex, qy_in, qy_ex, hb_in, hb_ex, sf_in, sf_ex, nx, ny, G, f0, f1, f2);
    else
        llf_flux(P, ze_in, ze_ex, qx_in, qx_ex, qy_in, qy_ex, hb_in, hb_ex, sf_in, sf_ex, nx, ny, G, f0, f1, f2);
}

}  // namespace dg
