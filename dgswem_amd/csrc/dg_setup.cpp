// Host-side set-up: reference-element tables, SSP-RK coefficients, edge tables.
//
// The numbers produced here must equal the reference's to the last bit for p=1, because the
// step's arithmetic is compared with dgswem_serial bit for bit.  Each formula therefore keeps the
// reference's operation order:
//   Dubiner basis      src/orthogonal_basis_v1.F:114-239 (modes ordered by total degree, the
//                      p=2 modes expanded from the same three-term recurrence)
//   area rules         src/quadrature_v1.F:245-275 (degree 2 and 4 symmetric triangle rules)
//   Lobatto rule       src/quad_rules_general.F:185-249 (GLL nodes; weights 2/(n(n-1)P_{n-1}^2))
//   mass inverse       src/prep_DG.F:712-721
//   RK_TIME            src/prep_DG.F:1551-2078
//   CREATE_EDGE_DATA   src/create_edge_data.F:31-635
//   Holland track      src/wind.F:315-348 (TIMECONV), 2153-2450 (GetHollandStormData),
//                      1964-2061 (HollandGet's storm-wide scalars)
#include "dg_setup.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <unordered_map>

namespace dg {

namespace {

// Basis value and reference-coordinate derivatives of the p<=2 Dubiner modes at (x, y).
// Mode order: 1, phi(0,1), phi(1,0), phi(0,2), phi(1,1), phi(2,0).
void dubiner(int p, double x, double y, double* v, double* dr, double* ds) {
    // degree 0 and 1 (PHI(0,0), PHI(1,0), PHI(0,1) of the recurrence)
    const double p10 = (1.0 + 2.0 * x + y) / 2.0;
    const double d10r = 1.0, d10s = 1.0 / 2.0;
    const double f01 = (1.0 + 2.0 * 0 + (3.0 + 2.0 * 0) * y) / 2.0;   // (1+3y)/2
    const double p01 = 1.0 * f01;
    const double d01r = 0.0 * f01;
    const double d01s = 0.0 * f01 + 1.0 * (3.0 + 2.0 * 0) / 2.0;
    double val[6] = {1.0, p01, p10, 0, 0, 0};
    double vr[6] = {0.0, d01r, d10r, 0, 0, 0};
    double vs[6] = {0.0, d01s, d10s, 0, 0, 0};
    if (p >= 2) {
        // PHI(2,0): (3/2) PHI10 PHI10 - (1/2) ((1-y)/2)^2
        const double a = (2.0 * 1 + 1.0) / (1 + 1.0), b = 1 / (1 + 1.0);
        const double om = (1.0 - y) / 2.0, om2 = om * om;
        const double p20 = a * p10 * p10 - b * om2 * 1.0;
        const double d20r = a * (d10r * p10 + p10 * d10r) - b * om2 * 0.0;
        const double d20s = a * (d10s * p10 + p10 * d10s) - b * ((y - 1) / 2 * 1.0 + om2 * 0.0);
        // PHI(1,1) = PHI(1,0) (1 + 2 + 5y)/2
        const double g = (1.0 + 2.0 * 1 + (3.0 + 2.0 * 1) * y) / 2.0;
        const double p11 = p10 * g;
        const double d11r = d10r * g;
        const double d11s = d10s * (1.0 + 2.0 * 1 + (3.0 + 2.0 * 1) * y) / 2.0 + p10 * (3.0 + 2.0 * 1) / 2.0;
        // PHI(0,2) = (AQ y + BQ) PHI(0,1) - CQ PHI(0,0), q=1, p=0
        const double AQ = (2.0 * 1 + 2.0 + 0.0) * (2.0 * 1 + 3.0 + 0.0) / (2.0 * 1 + 2.0) / (1 + 2.0 + 0.0);
        const double t = (2.0 * 0 + 1.0);
        const double BQ = t * t * (2.0 * 1 + 2.0 + 0.0) / (2.0 * 1 + 2.0) / (2.0 * 1 + 0.0 + 1.0) / (1 + 2.0 + 0.0);
        const double CQ = (1 + 0.0 + 1.0) * 1 * (2.0 * 1 + 3.0 + 0.0) / (1 + 1.0) / (1 + 2.0 + 0.0) / (2.0 * 1 + 0.0 + 1.0);
        const double p02 = (AQ * y + BQ) * p01 - CQ * 1.0;
        const double d02r = (AQ * y + BQ) * d01r - CQ * 0.0;
        const double d02s = (AQ * y + BQ) * d01s + AQ * p01 - CQ * 0.0;
        val[3] = p02; val[4] = p11; val[5] = p20;
        vr[3] = d02r; vr[4] = d11r; vr[5] = d20r;
        vs[3] = d02s; vs[4] = d11s; vs[5] = d20s;
    }
    const int dof = (p + 1) * (p + 2) / 2;
    for (int k = 0; k < dof; k++) {
        // the reference zeroes |value| < 1e-15 (src/orthogonal_basis_v1.F:229-231)
        v[k] = std::fabs(val[k]) < 1.0e-15 ? 0.0 : val[k];
        dr[k] = std::fabs(vr[k]) < 1.0e-15 ? 0.0 : vr[k];
        ds[k] = std::fabs(vs[k]) < 1.0e-15 ? 0.0 : vs[k];
    }
}

// Legendre P_n(x) by the three-term recurrence.
double legendre(int n, double x) {
    double pm = 1.0, pc = x;
    if (n == 0) return pm;
    for (int k = 1; k < n; k++) {
        const double kk = k;
        const double pn = ((2.0 * kk + 1.0) * x * pc - kk * pm) / (kk + 1.0);
        pm = pc; pc = pn;
    }
    return pc;
}

}  // namespace

void basis_at(int p, double r, double s, double* phi) {
    double dr[MAXDOF], ds[MAXDOF];
    dubiner(p, r, s, phi, dr, ds);
}

bool build_ref_tables(int p, RefTables& t, std::string& err) {
    if (p < 1 || p > 2) { err = "polynomial order must be 1 or 2"; return false; }
    t = RefTables{};
    t.p = p;
    t.dof = (p + 1) * (p + 2) / 2;
    t.negp = p + 2;                      // NEGP = ceil((2p+3)/2)
    // --- area rule of degree 2p: orbits of multiplicity 3 at barycentric (a, a, 1-2a)
    double A[2], W[2];
    int norb;
    if (p == 1) { norb = 1; A[0] = 0.1666666666666666666666666666666666; W[0] = 0.3333333333333333333333333333333333; }
    else {
        norb = 2;
        A[0] = 0.4459484909159648863183292538830519; A[1] = 0.0915762135097707434595714634022015;
        W[0] = 0.2233815896780114656950070084331228; W[1] = 0.1099517436553218676383263249002105;
    }
    t.nagp = 3 * norb;
    const double v1x = -1.0, v1y = -1.0, v2x = 1.0, v2y = -1.0, v3x = -1.0, v3y = 1.0, area = 2.0;
    for (int o = 0; o < norb; o++) {
        const double a = A[o], b = a, c = 1.0 - (a + b);
        const double bc[3][3] = {{a, b, c}, {b, c, a}, {c, a, b}};
        for (int r = 0; r < 3; r++) {
            const int q = 3 * o + r;
            t.xagp[q] = v1x * bc[r][0] + v2x * bc[r][1] + v3x * bc[r][2];
            t.yagp[q] = v1y * bc[r][0] + v2y * bc[r][1] + v3y * bc[r][2];
            t.wagp[q] = area * W[o];
        }
    }
    // --- Gauss-Lobatto-Legendre rule on [-1,1]
    const int n = t.negp;
    if (n == 3) { t.xegp[0] = -1.0; t.xegp[1] = 0.0; t.xegp[2] = 1.0; }
    else { const double r = 1.0 / std::sqrt(5.0); t.xegp[0] = -1.0; t.xegp[1] = -r; t.xegp[2] = r; t.xegp[3] = 1.0; }
    {
        const double m = n - 2;
        for (int q = 0; q < n; q++) {
            if (q == 0 || q == n - 1) t.wegp[q] = 2.0 / ((m + 1.0) * (m + 2.0));
            else {
                const double pl = legendre(n - 1, t.xegp[q]);
                t.wegp[q] = 2.0 / ((m + 1.0) * (m + 2.0) * pl * pl);
            }
        }
    }
    // --- basis tables
    double v[MAXDOF], dr[MAXDOF], ds[MAXDOF];
    for (int q = 0; q < t.nagp; q++) {
        dubiner(p, t.xagp[q], t.yagp[q], v, dr, ds);
        for (int k = 0; k < t.dof; k++) { t.phi_area[q][k] = v[k]; t.drphi[q][k] = dr[k]; t.dsphi[q][k] = ds[k]; }
    }
    const double cx[3] = {-1.0, 1.0, -1.0}, cy[3] = {-1.0, -1.0, 1.0};
    for (int c = 0; c < 3; c++) {
        dubiner(p, cx[c], cy[c], v, dr, ds);
        for (int k = 0; k < t.dof; k++) t.phi_corner[c][k] = v[k];
    }
    // local edge l runs from vertex l+1 to vertex l+2 (mod 3): points parametrised by XEGP
    for (int l = 0; l < 3; l++)
        for (int q = 0; q < n; q++) {
            const double s = t.xegp[q];
            const double px = (l == 0) ? -s : (l == 1 ? -1.0 : s);
            const double py = (l == 0) ? s : (l == 1 ? -s : -1.0);
            dubiner(p, px, py, v, dr, ds);
            for (int k = 0; k < t.dof; k++) t.phi_edge[l][q][k] = v[k];
        }
    {
        int m = 0;
        for (int j = 0; j <= p; j++)
            for (int i = 0; i <= j; i++) {
                const int jj = j - i;
                t.m_inv[m++] = ((2.0 * i + 1.0) * (2.0 * jj + 2.0 * i + 2.0) / 4.0);
            }
    }
    for (int l = 0; l < 3; l++)
        for (int q = 0; q < n; q++)
            for (int k = 0; k < t.dof; k++)
                t.edgeq[l][q][k] = 2.0 * t.m_inv[k] * t.phi_edge[l][q][k] * t.wegp[q];
    for (int q = 0; q < t.nagp; q++)
        for (int k = 0; k < t.dof; k++) {
            t.mw[q][k] = t.m_inv[k] * t.wagp[q];
            t.srfac[q][k] = t.m_inv[k] * t.wagp[q] * t.phi_area[q][k];
        }
    // linear nodal basis at the area points (src/prep_DG.F:940-942)
    for (int q = 0; q < t.nagp; q++) {
        t.psi[q][0] = -1.0 / 2.0 * (t.xagp[q] + t.yagp[q]);
        t.psi[q][1] = 1.0 / 2.0 * (t.xagp[q] + 1.0);
        t.psi[q][2] = 1.0 / 2.0 * (t.yagp[q] + 1.0);
    }
    return true;
}

bool build_rk(int s, int r, RKScheme& k, std::string& err) {
    k = RKScheme{};
    if (s < 1 || s > MAXRK) { err = "unsupported number of RK stages"; return false; }
    k.nrk = s;
    auto A = [&](int i, int j, double v) { k.a[i - 1][j - 1] = v; };
    auto B = [&](int i, int j, double v) { k.b[i - 1][j - 1] = v; };
    if (s == 1 && r == 1) { A(1, 1, 1.0); B(1, 1, 1.0); }
    else if (r == 2 && s == 3) {
        A(1,1,1.0); A(2,1,0.087353119859156); A(3,1,0.344956917166841); A(2,2,0.912646880140844);
        A(3,3,0.655043082833159); B(1,1,0.528005024856522); B(3,1,0.022826837460491);
        B(2,2,0.481882138633993); B(3,3,0.345866039233415);
    } else if (r == 2 && s == 4) {
        A(1,1,1.0); A(2,1,0.394806441339829); A(3,1,0.002797307087390); A(4,1,0.252860909354373);
        A(2,2,0.605193558660171); A(3,3,0.997202692912610); A(4,4,0.747139090645627);
        B(1,1,0.406584463657504); B(3,1,0.013637216641451); B(4,1,0.016453567333598);
        B(2,2,0.246062298456822); B(3,3,0.405447122055692); B(4,4,0.303775146447707);
    } else if (r == 2 && s == 5) {
        A(1,1,1.0); A(2,1,0.235593265061659); A(3,1,0.174017972351526); A(4,1,0.235264368870758);
        A(5,1,0.141720372339803); A(2,2,0.764406734938341); A(4,2,0.000058643383967);
        A(5,2,0.095374613155521); A(3,3,0.825982027648475); A(5,3,0.000311763705780);
        A(4,4,0.764676987745275); A(5,5,0.762593250798895);
        B(1,1,0.324840618151514); B(3,1,0.108822380501601); B(4,1,0.054392262422093);
        B(5,1,0.000000180291569); B(2,2,0.248310356296551); B(4,2,0.000019049753098);
        B(5,2,0.030981548293401); B(3,3,0.268312512443371); B(5,3,0.000101273514903);
        B(4,4,0.248398145385413); B(5,5,0.247721262987686);
    } else if (r == 2) {
        // optimal SSP(s,2): Euler steps of dt/(s-1), last stage averaged with the start
        for (int i = 1; i <= s; i++) {
            if (i < s) { A(i, i, 1.0); B(i, i, 1.0 / (s - 1)); }
            else { A(s, 1, 1.0 / s); A(s, s, (s - 1.0) / s); B(s, s, 1.0 / s); }
        }
    } else if (s == 3 && r == 3) {
        A(1,1,1.0); A(2,1,3.0/4.0); A(2,2,1.0/4.0); A(3,1,1.0/3.0); A(3,3,2.0/3.0);
        B(1,1,1.0); B(2,2,1.0/4.0); B(3,3,2.0/3.0);
    } else { err = "unsupported RK scheme (rk_stage, rk_order)"; return false; }
    // stage times: c_k = sum_{l<k} C(k-1,l) with C(i,j) = B(i,j) + sum_l C(l,j) A(i,l+1)
    double C[MAXRK][MAXRK]{};
    for (int j = 0; j < s; j++)
        for (int i = 0; i < s; i++) {
            double acc = 0.0;
            for (int l = j + 1; l <= i; l++) acc = acc + C[l - 1][j] * k.a[i][l];   // L = K+1..I-1 (1-based I = i+1)
            C[i][j] = k.b[i][j] + acc;
        }
    k.dtvd[0] = 0.0;
    for (int st = 1; st < s; st++) {
        double acc = 0.0;
        for (int l = 0; l < st; l++) acc = acc + C[st - 1][l];
        k.dtvd[st] = acc;
    }
    return true;
}

bool create_edge_data(int np, int ne, const int* nm, int nope, const int* nvdll, const int* nbdv,
                      int nbou, const int* ibtype, const int* nvell, const int* nbvv, EdgeData& ed,
                      std::string& err) {
    auto node = [&](int e, int k) { return nm[(size_t)k * ne + e] - 1; };
    ed = EdgeData{};
    ed.neled.assign((size_t)ne * 3, -1);
    ed.nedno.reserve((size_t)ne * 3); ed.nedel.reserve((size_t)ne * 3);
    std::unordered_map<uint64_t, int> key2edge;
    key2edge.reserve((size_t)ne * 2);
    auto key = [](int a, int b) { if (a > b) std::swap(a, b); return ((uint64_t)(uint32_t)a << 32) | (uint32_t)b; };
    // edges numbered in order of first appearance scanning elements then local edges 1..3;
    // the first element is EL_IN and fixes the node order (its counter-clockwise order)
    for (int e = 0; e < ne; e++) {
        const int n[3] = {node(e, 0), node(e, 1), node(e, 2)};
        for (int l = 0; l < 3; l++) {
            const int a = n[(l + 1) % 3], b = n[(l + 2) % 3];
            if (a < 0 || a >= np || b < 0 || b >= np) { err = "node index out of range"; return false; }
            auto it = key2edge.find(key(a, b));
            if (it == key2edge.end()) {
                const int id = ed.nedges++;
                key2edge.emplace(key(a, b), id);
                ed.nedno.push_back(a); ed.nedno.push_back(b);
                ed.nedel.push_back(e); ed.nedel.push_back(-1);
                ed.neled[(size_t)e * 3 + l] = id;
            } else {
                const int id = it->second;
                if (ed.nedel[2 * (size_t)id + 1] >= 0) {
                    err = "possible duplicate element " + std::to_string(e + 1);
                    return false;
                }
                ed.nedel[2 * (size_t)id + 1] = e;
                ed.neled[(size_t)e * 3 + l] = id;
            }
        }
    }
    const int ned = ed.nedges;
    std::vector<int> cls(ned, -1);
    for (int i = 0; i < ned; i++)
        if (ed.nedel[2 * (size_t)i] >= 0 && ed.nedel[2 * (size_t)i + 1] >= 0) { ed.niedn.push_back(i); cls[i] = 0; }
    auto find = [&](int a, int b) -> int {
        auto it = key2edge.find(key(a, b));
        return it == key2edge.end() ? -1 : it->second;
    };
    size_t off = 0;
    for (int k = 0; k < nope; k++) {
        for (int i = 0; i + 1 < nvdll[k]; i++) {
            const int e = find(nbdv[off + i] - 1, nbdv[off + i + 1] - 1);
            if (e < 0) { err = "error in processing open ocean boundary conditions"; return false; }
            ed.needn.push_back(e); cls[e] = 4;
        }
        off += nvdll[k];
    }
    off = 0;
    for (int k = 0; k < nbou; k++) {
        const int t = ibtype[k];
        const bool land = (t == 0 || t == 10 || t == 20 || t == 1 || t == 11 || t == 21);
        const bool flow = (t == 2 || t == 12 || t == 22);
        if (!land && !flow) { err = "boundary type " + std::to_string(t) + " is not supported"; return false; }
        for (int i = 0; i + 1 < nvell[k]; i++) {
            const int e = find(nbvv[off + i] - 1, nbvv[off + i + 1] - 1);
            if (e < 0) continue;                 // the reference reports and skips such pairs
            cls[e] = 1;
            (land ? ed.nledn : ed.nfedn).push_back(e);
        }
        off += nvell[k];
    }
    for (int i = 0; i < ned; i++)
        if (cls[i] < 0) ed.nledn.push_back(i);   // unclassified boundary edges are land
    ed.nedsd.assign((size_t)ned * 2, -1);
    for (int i = 0; i < ned; i++)
        for (int s = 0; s < 2; s++) {
            const int e = ed.nedel[2 * (size_t)i + s];
            if (e < 0) continue;
            for (int l = 0; l < 3; l++)
                if (ed.neled[(size_t)e * 3 + l] == i) ed.nedsd[2 * (size_t)i + s] = l;
        }
    return true;
}

namespace {

// seconds since the start of the year, minutes and seconds zero
double timeconv(int iyr, int imo, int iday, int ihr) {
    double t = (double)((iday - 1) * 86400 + ihr * 3600);
    const int ileap = (iyr / 4) * 4;
    static const int add[12] = {0, 31, 0, 31, 30, 31, 30, 31, 31, 30, 31, 30};
    for (int m = 2; m <= 12; m++) {
        if (imo < m) break;
        if (m == 3) t = t + (ileap == iyr ? 29 : 28) * 86400;
        else t = t + add[m - 1] * 86400;
    }
    return t;
}

bool same_date(const int* a, const int* b) { return a[0] == b[0] && a[1] == b[1] && a[2] == b[2] && a[3] == b[3]; }

}  // namespace

bool build_holland_track(int nrec, const int* rec, const int* ref, int storm_number, double bladj,
                         HollandTrack& t, std::string& err) {
    const double re = 6378206.4;
    const double pi = 4.0 * std::atan(1.0);
    const double deg2rad = pi / 180.0;
    if (nrec < 2) { err = "the storm track needs at least two records"; return false; }
    const double ref_time = timeconv(ref[0], ref[1], ref[2], ref[3]);
    t = HollandTrack{};
    t.storm_number = storm_number;
    t.bladj = bladj;
    // kept records: the last non-repeated line is compared with the next
    int prev_date[4] = {0, 0, 0, 0}, prev_inc = 0;
    double prev_fcst = 0.0, prev_time = 0.0;
    int now_cp = 0, now_rmw = 0, now_rrp = 0;
    for (int r = 0; r < nrec; r++) {
        const int* R = rec + 12 * r;
        const int kept = (int)t.time.size();
        int cp = R[9], rrp = R[10], rmw = R[11];
        double time, fcst;
        if (R[4] == 0) {                    // BEST: hindcast/nowcast
            if (kept > 0 && same_date(R, prev_date)) continue;
            now_cp = cp; now_rmw = rmw; now_rrp = rrp;
            time = timeconv(R[0], R[1], R[2], R[3]) - ref_time;
            fcst = R[5];
        } else if (R[4] == 1) {             // OFCL: forecast from the last nowcast
            if (kept == 0) { err = "the first storm track record must be a BEST (hindcast) line"; return false; }
            if ((R[5] == 0 && same_date(R, prev_date)) || (R[5] != 0 && R[5] == prev_inc)) continue;
            fcst = R[5];
            if (R[5] == 0) {
                time = timeconv(R[0], R[1], R[2], R[3]) - ref_time;
            } else {
                fcst = fcst * 3600.0;
                time = prev_time + (fcst - prev_fcst);
            }
            cp = now_cp; rmw = now_rmw; rrp = now_rrp;
        } else {
            err = "invalid storm track record type (valid: BEST hindcast, OFCL forecast)";
            return false;
        }
        double lat = R[6], lon = R[7], spd = R[8];
        double cpd = cp, rrpd = rrp, rmwd = rmw;
        lat = lat / 10.0;
        lon = lon / 10.0;
        lon = -1.0 * lon;
        cpd = cpd * 100.0;
        rrpd = rrpd * 1.852000003180799 * 1000.0;
        rmwd = rmwd * 1.852000003180799;
        spd = spd * 0.51444444;
        t.time.push_back(time); t.lat.push_back(lat); t.lon.push_back(lon); t.cpress.push_back(cpd);
        t.spd.push_back(spd); t.rrp.push_back(rrpd); t.rmw.push_back(rmwd);
        for (int k = 0; k < 4; k++) prev_date[k] = R[k];
        prev_inc = R[5];
        prev_fcst = fcst;
        prev_time = time;
    }
    t.n = (int)t.time.size();
    if (t.n < 2) { err = "the storm track needs at least two distinct times"; return false; }
    t.tvx.assign(t.n, 0.0);
    t.tvy.assign(t.n, 0.0);
    for (int k = 1; k < t.n; k++) {
        const double dt = t.time[k] - t.time[k - 1];
        const double s = std::sin((t.lon[k] * deg2rad - t.lon[k - 1] * deg2rad) / 2.0);
        t.tvx[k] = std::copysign((re * (2.0 * std::asin(std::sqrt(std::cos(t.lat[k] * deg2rad) *
                                                                 std::cos(t.lat[k - 1] * deg2rad) * (s * s))))) / dt,
                                 t.lon[k] - t.lon[k - 1]);
        const double q = std::sin((t.lat[k] * deg2rad - t.lat[k - 1] * deg2rad) / 2.0);
        t.tvy[k] = std::copysign((re * (2.0 * std::asin(std::sqrt(q * q)))) / dt, t.lat[k] - t.lat[k - 1]);
    }
    t.tvx[0] = t.tvx[1];
    t.tvy[0] = t.tvy[1];
    return true;
}

bool holland_row(const HollandTrack& t, double time, double g, double* row, std::string& err) {
    // the interval [CastTime(i-1), CastTime(i)) holding the time; the reference finds it once and
    // then advances one record per call, which is the same interval while a time step is shorter
    // than the record spacing
    int i = 1;
    while (i < t.n && !(time >= t.time[i - 1] && time < t.time[i])) i++;
    if (i >= t.n) {
        err = "the storm track (fort.22) does not cover model time " + std::to_string(time) + " s";
        return false;
    }
    const double w = (time - t.time[i - 1]) / (t.time[i] - t.time[i - 1]);
    const double cpress = t.cpress[i - 1] + w * (t.cpress[i] - t.cpress[i - 1]);
    const double lon = t.lon[i - 1] + w * (t.lon[i] - t.lon[i - 1]);
    const double lat = t.lat[i - 1] + w * (t.lat[i] - t.lat[i - 1]);
    double spd = t.spd[i - 1] + w * (t.spd[i] - t.spd[i - 1]);
    const double rmw = t.rmw[i - 1] + w * (t.rmw[i] - t.rmw[i - 1]);
    const double tvx = t.tvx[i - 1] + w * (t.tvx[i] - t.tvx[i - 1]);
    const double tvy = t.tvy[i - 1] + w * (t.tvy[i] - t.tvy[i - 1]);
    double deficit = 101300.0 - cpress;
    if (deficit < 100.0) deficit = 100.0;
    const double ts = std::sqrt(tvx * tvx + tvy * tvy);
    spd = spd - ts;
    spd = spd / t.bladj;
    double B = 1.15 * 2.718281828459045 * (spd * spd) / deficit;
    if (B < 1.0) B = 1.0;
    if (B > 2.5) B = 2.5;
    const double pi = 4.0 * std::atan(1.0);
    const double deg2rad = pi / 180.0;
    for (int k = 0; k < MET_ROW; k++) row[k] = 0.0;
    row[MR_LAT] = lat; row[MR_LON] = lon; row[MR_CPRESS] = cpress; row[MR_DEFICIT] = deficit;
    row[MR_SPD] = spd; row[MR_RMW] = rmw; row[MR_TVX] = tvx; row[MR_TVY] = tvy; row[MR_B] = B;
    row[MR_COSLAT] = std::cos(lat * deg2rad);
    row[MR_WMULT] = t.storm_number == 2 ? 1.2 : 1.0;
    row[MR_BLADJ] = t.bladj;
    row[MR_RHOG] = 1000.0 * g;
    return true;
}

}  // namespace dg

namespace dg {

void eletab_order(int np, int ne, const int* nm, const double* x, const double* y,
                  std::vector<int>& off, std::vector<int>& row) {
    // incident elements in element order (NNOEL's content; its order does not matter below, an
    // element being fixed by its three nodes)
    off.assign(np + 1, 0);
    for (int e = 0; e < ne; e++) for (int v = 0; v < 3; v++) off[nm[(size_t)v * ne + e] + 1]++;
    for (int i = 0; i < np; i++) off[i + 1] += off[i];
    row.assign((size_t)3 * ne, 0);
    {
        std::vector<int> fill(np, 0);
        for (int e = 0; e < ne; e++)
            for (int v = 0; v < 3; v++) {
                const int n = nm[(size_t)v * ne + e];
                row[off[n] + fill[n]++] = (e << 2) | v;
            }
    }
    // NEIGHB: neighbours in insertion order (pairs 1-2, 1-3, 2-3 of every element, each once)
    std::vector<std::vector<int>> nb(np);
    for (int e = 0; e < ne; e++) {
        const int en[3] = {nm[e], nm[(size_t)ne + e], nm[2 * (size_t)ne + e]};
        const int pr[3][2] = {{0, 1}, {0, 2}, {1, 2}};
        for (auto& q : pr) {
            const int a = en[q[0]], b = en[q[1]];
            bool found = false;
            for (int m : nb[a]) if (m == b) { found = true; break; }
            if (found) continue;
            nb[a].push_back(b);
            nb[b].push_back(a);
        }
    }
    const double rad2deg = 45.0 / std::atan(1.0);
    std::vector<double> ang;
    std::vector<int> sorted, tmp;
    for (int i = 0; i < np; i++) {
        const int nn = (int)nb[i].size();
        ang.assign(nn, 0.0);
        for (int j = 0; j < nn; j++) {
            const int m = nb[i][j];
            const double delx = x[m] - x[i], dely = y[m] - y[i];
            const double dist = std::sqrt(delx * delx + dely * dely);
            if (dely != 0.0) {
                ang[j] = rad2deg * std::acos(delx / dist);
                if (dely > 0.0) ang[j] = 360.0 - ang[j];
            } else {
                if (delx > 0.0) ang[j] = 0.0;
                if (delx < 0.0) ang[j] = 180.0;
            }
        }
        // selection by increasing angle, as NEIGHB does it
        sorted.assign(nn + 1, i);
        double more = -1.0;
        int jlow = 0;
        for (int jj = 0; jj < nn; jj++) {
            double low = 400.0;
            for (int j = 0; j < nn; j++)
                if (ang[j] < low && ang[j] > more) { low = ang[j]; jlow = j; }
            sorted[1 + jj] = nb[i][jlow];
            more = low;
        }
        // ELETAB: the element of each consecutive (cyclic) neighbour pair
        const int cnt = off[i + 1] - off[i], nneigh = nn + 1;
        tmp.clear();
        for (int k = 1; k <= nneigh - 1; k++) {
            const int j2 = sorted[k], j3 = k < nneigh - 1 ? sorted[k + 1] : sorted[1];
            for (int j = off[i]; j < off[i + 1]; j++) {
                const int e = row[j] >> 2;
                const int v0 = nm[e], v1 = nm[(size_t)ne + e], v2 = nm[2 * (size_t)ne + e];
                const bool h2 = v0 == j2 || v1 == j2 || v2 == j2;
                const bool h3 = v0 == j3 || v1 == j3 || v2 == j3;
                if (h2 && h3) { tmp.push_back(row[j]); break; }
            }
        }
        if ((int)tmp.size() == cnt) std::copy(tmp.begin(), tmp.end(), row.begin() + off[i]);
    }
}

}  // namespace dg
