// Host-side set-up of the DG hydro step: reference-element tables, SSP-RK coefficients and the
// edge tables.  Plain C++ (built by hipcc into libdgswem_hip.so), no device code.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace dg {

constexpr int MAXDOF = 6;   // p <= 2
constexpr int MAXAGP = 6;
constexpr int MAXEGP = 4;
constexpr int MAXRK = 8;

// Reference-element quantities for one polynomial order (src/prep_DG.F:479-870).
struct RefTables {
    int p = 1, dof = 3, nagp = 3, negp = 3;
    double xagp[MAXAGP]{}, yagp[MAXAGP]{}, wagp[MAXAGP]{};
    double xegp[MAXEGP]{}, wegp[MAXEGP]{};
    double phi_area[MAXAGP][MAXDOF]{}, drphi[MAXAGP][MAXDOF]{}, dsphi[MAXAGP][MAXDOF]{};
    double phi_edge[3][MAXEGP][MAXDOF]{};   // local edge, Lobatto point, mode
    double edgeq[3][MAXEGP][MAXDOF]{};      // 2*M_INV(K)*PHI_EDGE*WEGP
    double phi_corner[3][MAXDOF]{};
    double m_inv[MAXDOF]{};
    double mw[MAXAGP][MAXDOF]{};            // M_INV(K)*WAGP(I): XFAC = mw*dphi/dx
    double srfac[MAXAGP][MAXDOF]{};         // M_INV(K)*WAGP(I)*PHI_AREA(K,I)
    double psi[MAXAGP][3]{};                // linear nodal basis PSI1-3 at the area points
};
bool build_ref_tables(int p, RefTables& t, std::string& err);
// Values of the p<=2 modes at reference coordinates (r, s) of the triangle (-1,-1),(1,-1),(-1,1).
void basis_at(int p, double r, double s, double* phi);

// SSP Runge-Kutta scheme in Shu-Osher form (src/prep_DG.F:1551-2078).
struct RKScheme {
    int nrk = 2;
    double a[MAXRK][MAXRK]{}, b[MAXRK][MAXRK]{}, dtvd[MAXRK]{};
};
bool build_rk(int stages, int order, RKScheme& s, std::string& err);

// Edge tables in the reference's conventions, 0-based internally.
struct EdgeData {
    int nedges = 0;
    std::vector<int> nedno;   // [ned][2] nodes
    std::vector<int> nedel;   // [ned][2] elements, -1 if none
    std::vector<int> nedsd;   // [ned][2] local edge 0..2, -1 if none
    std::vector<int> neled;   // [ne][3] edge of each local edge
    std::vector<int> niedn, nledn, needn, nfedn;
};
bool create_edge_data(int np, int ne, const int* nm_colmajor_1based, int nope, const int* nvdll,
                      const int* nbdv, int nbou, const int* ibtype, const int* nvell,
                      const int* nbvv, EdgeData& ed, std::string& err);

// Node-to-element table in PREP_DG's ELETAB order (src/prep_DG.F:292-330): around each node the
// elements between consecutive neighbours, the neighbours sorted by NEIGHB by clockwise angle from
// east (src/dgswem.F:478-570).  row[off[i]..off[i+1]) holds element*4 + local vertex; a node whose
// neighbour fan does not yield all its elements keeps them in ascending element order.
// nm: [3][ne] 0-based (vertex-major), x/y: the projected node coordinates.
void eletab_order(int np, int ne, const int* nm, const double* x, const double* y,
                  std::vector<int>& off, std::vector<int>& row);

// NWS=8 storm track after the first-call processing of GetHollandStormData
// (src/wind.F:2185-2410): repeated lines dropped, forecasts given the nowcast's pressure and radii,
// SI units, cast times relative to the wind reference time, translation velocities.
struct HollandTrack {
    int n = 0, storm_number = 0;
    double bladj = 1.0;
    std::vector<double> time, lat, lon, cpress, spd, rrp, rmw, tvx, tvy;
};
// rec: [nrec][12] year month day hour cast(0 BEST, 1 OFCL) fcstinc lat(0.1N) lon(0.1W) spd(kt)
// cpress(mbar) rrp(nm) rmw(nm); ref: IREFYR IREFMO IREFDAY IREFHR.
bool build_holland_track(int nrec, const int* rec, const int* ref, int storm_number, double bladj,
                         HollandTrack& t, std::string& err);

// Per-step storm scalars for the node kernel: the track interpolated at TIME_A and the
// node-independent part of HollandGet (src/wind.F:2020-2061).
constexpr int MET_ROW = 16;
enum {
    MR_LAT, MR_LON, MR_CPRESS, MR_DEFICIT, MR_SPD, MR_RMW, MR_TVX, MR_TVY, MR_B, MR_COSLAT,
    MR_WMULT, MR_BLADJ, MR_RHOG
};
bool holland_row(const HollandTrack& t, double time, double g, double* row, std::string& err);

}  // namespace dg
