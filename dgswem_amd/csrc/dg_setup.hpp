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
};
bool build_ref_tables(int p, RefTables& t, std::string& err);

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

}  // namespace dg
