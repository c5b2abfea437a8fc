/*
 * dgswem_oracle.c -- CPU restatement of DG-SWEM's explicit hydro step, FP64.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker for the HIP implementation in dgswem_amd/:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The product path
 * never calls it.
 *
 * It restates, loop by loop and in the reference's evaluation order, the serial solver's time
 * step so that its results can be compared bit-for-bit with dgswem_serial (oracle/_ref) and with
 * the reference's golden fort.63 fixtures.  Array layouts mirror the Fortran ones (column-major,
 * modes fastest) with 0-based indices.
 *
 * Reference units restated here (paths relative to the reference root):
 *   PREP_DG tables ........ src/prep_DG.F:25-1539          RK_TIME ....... src/prep_DG.F:1551-2078
 *   ORTHOGONAL_BASIS ...... src/orthogonal_basis_v1.F:50-280
 *   triangle rules ........ src/quadrature_v1.F:245-275,804-862
 *   Gauss-Lobatto rule .... src/quad_rules_general.F:40-260 (closed form for 3 and 4 points)
 *   CREATE_EDGE_DATA ...... src/create_edge_data.F:31-635 (+ src/count_edges.cpp get_edge_no)
 *   CALC_NORMAL ........... src/calc_normal.F:13-52
 *   DG_HYDRO_TIMESTEP ..... src/DG_hydro_timestep.F:24-744
 *   OCEAN_EDGE_HYDRO ...... src/ocean_edge_hydro.F:19-355
 *   LAND_EDGE_HYDRO ....... src/land_edge_hydro.F:26-311
 *   FLOW_EDGE_HYDRO ....... src/flow_edge_hydro.F:20-333
 *   INTERNAL_EDGE_HYDRO ... src/internal_edge_hydro.F:37-383 (incl. add_internal_flux)
 *   numerical_flux (LLF) .. src/numerical_flux.F:750-1022
 *   RHS_DG_HYDRO .......... src/rhs_dg_hydro.F:18-473
 *   SLOPELIMITER5 ......... src/slopelimiter.F:3247-3783
 *   WETDRY ................ src/wetdry.F:17-553
 *   modal2nodal ........... src/write_results.F:14-174 (with ETAMAX, :164)
 *   TIDAL_POTENTIAL ....... src/tidal_potential.F:15-80 (+ its source, src/rhs_dg_hydro.F:378-386)
 *   non-periodic elevation and flux boundaries (fort.19, fort.20)
 *                           src/DG_hydro_timestep.F:153-200, src/cstart.F:80-124,
 *                           src/ocean_edge_hydro.F:180-184, src/flow_edge_hydro.F:143-173
 *   NWS=8 Holland vortex .. src/wind.F:116-141,1964-2450, src/met_forcing.F:479-520
 *
 * Deliberate differences from the reference (all only where the reference is broken for p=2):
 *   - internal edges trace and accumulate all DOFH modes (the reference hard-codes 3,
 *     src/internal_edge_hydro.F:59-64,306), so p=2 is a correct DG scheme here;
 *   - the element y used for the spherical factor is the P1 projection for every p (the reference
 *     fills YDUB only for chi=1, src/prep_DG.F:870-872).
 * For p=1 both are identical to the reference.
 *
 * Parity status: p=1 is pinned against dgswem_serial (FP64 hot-start dumps and the reference's
 * fort.63.true fixtures, see tests/), including the tidal potential (NTIP 1 and 2), both time-series
 * boundaries and NWS=8, all bit for bit.  p=2 has no working reference and is pinned only by
 * invariants (lake at rest, mass conservation, symmetry); see tests/test_oracle_invariants.py.
 * modal2nodal visits each node's elements in the reference's ELETAB order (NEIGHB's angular sort,
 * src/dgswem.F:478-570, and src/prep_DG.F:292-330) and raises ETAMAX after every wet contribution,
 * so ETA2 and ETAMAX are pinned against dgswem_serial's fort.63 and maxele.63 as well.
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#define MAXDOF 6
#define MAXAGP 6
#define MAXEGP 4
#define MAXRK 8

typedef struct oracle {
    /* sizes */
    int np, ne, p, dofh, nagp, negp, nrk, rk_order;
    /* integer switches */
    int ics, nolibf, nolifa, nolica, nolicat, nramp, slopeflag, rainfall, load_manning;
    /* real parameters */
    double g, dtdp, statim, dramp, h0, cf, hbreak, ftheta, fgamma, sfea0;
    /* derived flags (src/dgswem.F:131-170) */
    double nleq, leq, nleqg, fg_l;
    int ifnlfa, iflinbf, ifhybf;
    /* mesh (node arrays copied) */
    double *x, *y, *dp, *sfac, *corif, *manning;
    int *nm; /* [ne][3] 0-based */
    /* boundary segments */
    int nope, *open_len, *open_nodes;
    int nbou, *land_type, *land_len, *land_nodes;
    /* tides */
    int nbfr; double *amig, *ff, *face, *per, *emo, *efa; /* emo/efa [nbfr][neta] */
    int neta;
    /* flow */
    int nffr; double *famig, *fff, *fface, *fper, *qnam, *qnph; int nflownodes;
    /* edges */
    int nedges, *nedno, *nedel, *nedsd, *neled; /* nedno[2*ned] 0-based node, nedel -1 if none */
    int nieds, *niedn, nleds, *nledn, needs, *needn, nfeds, *nfedn;
    int *elem_edge; /* [ne][3], -1 if unused */
    double *xlen, *cosnx, *sinnx;
    double *emo_dg, *efa_dg; /* [needs][nbfr][2] */
    double *qnam_dg, *qnph_dg; /* [nfeds][nffr][2] */
    /* reference-element tables */
    double xagp[MAXAGP], yagp[MAXAGP], wagp[MAXAGP], xegp[MAXEGP], wegp[MAXEGP];
    double phi_area[MAXAGP][MAXDOF], drphi[MAXAGP][MAXDOF], dsphi[MAXAGP][MAXDOF];
    double phi_corner[3][MAXDOF], phi_edge[3][MAXEGP][MAXDOF], m_inv[MAXDOF];
    double psi1[MAXAGP], psi2[MAXAGP], psi3[MAXAGP];
    double atvd[MAXRK][MAXRK], btvd[MAXRK][MAXRK], dtvd[MAXRK];
    /* per-element tables */
    double *areas, *drdx, *dsdx, *drdy, *dsdy, *cori_el, *fric_el, *dpe_min;
    double *hb;                       /* [ne][dofh] */
    double *bath, *dbathdx, *dbathdy, *sfac_elem; /* [ne][nagp] */
    double *xfac, *yfac, *srfac;      /* [ne][nagp][dofh] */
    double *bathed, *sfaced;          /* [ne][3][negp] */
    double *edgeq;                    /* [3][negp][dofh] (element independent) */
    double *dp_node;                  /* [ne][3] */
    /* node -> element table */
    int *el_count, *eletab_off, *eletab;
    double *etamax;          /* ETAMAX, the running maximum modal2nodal keeps (write_results.F:164) */
    /* state */
    double *ze, *qx, *qy;             /* [(nrk+1)][ne][dofh] */
    double *rhs_ze, *rhs_qx, *rhs_qy; /* [nrk][ne][dofh] */
    double *ze_edge, *qx_edge, *qy_edge; /* [ne][3 slots][3] -> generalised to dofh */
    int *wdflg, *nodecode;
    int nthreads;
    /* stage context */
    int irk;
    double timedg, rampdg;
    /* limiter scratch */
    double *zemin, *zemax, *qxmin, *qxmax, *qymin, *qymax;
    /* NWS=8 Holland vortex: processed track, node coordinates and the stage's node forcing */
    int nws, hl_n, hl_ptr, hl_first, storm_number;
    double bladj, wind_ref_time;
    double *hl_time, *hl_lat, *hl_lon, *hl_cpress, *hl_spd, *hl_rrp, *hl_rmw, *hl_tvx, *hl_tvy;
    double *slam, *sfea;
    double *wsx2, *wsy2, *pr2;
    double drpsi[3], dspsi[3];
    int met_error;
    /* tidal potential (NTIP = 1, 2): constituents, the SAL terms [ntif][np] and the stage's TIP2 */
    int ntip, ntif;
    double reftim, *tpk, *amigt, *etrf, *fft, *facet, *pert, *saltamp, *saltpha, *tip2;
    double *tslam, *tsfea;
    /* non-periodic elevation boundary (fort.19): records [nrec][neta], the two in use, their times,
       and ETA2 at the boundary nodes (the reference writes these into its nodal ETA2 array) */
    int eser, eser_nrec, eser_next, *nbd;
    double etiminc, etime1, etime2, *esbin, *esbin1, *esbin2, *eta2bc;
    /* non-periodic normal flux (fort.20): boundary codes of the NVEL land nodes, records of the
       flux-node values, the two in use (QNIN1/QNIN2 over all NVEL nodes) and their times */
    int fser, fser_nrec, fser_next, nvel, nfluxnodes, *lbcodei;
    double ftiminc, qtime1, qtime2, *qnin, *qnin1, *qnin2;
    int series_error;
    /* frozen-branch SLOPELIMITER5, WETDRY and internal-edge dry-side treatment (a check of the
       adjoint): mode 1 records the branches every stage takes, mode 2 replays recorded branches
       instead of taking them */
    int lim_mode, lim_cap, lim_pos;
    int *lim_node;   /* [call][np][6]: elements attaining each field's min and max bound, -1 none */
    int *lim_elem;   /* [call][ne][3]: per field, the clamp (2 bits per vertex), pass signs, positive
                        differences and min(redfac, redmax) choices, packed */
    int wd_pos;
    int *wd_rec;     /* [call][ne]: WETDRY's branches (wetdry_code), recorded with the limiter's */
    int ie_pos;
    unsigned char *ie_rec;  /* [call][nieds][negp]: each internal edge point's dry-side treatment
                               (0 none; a dry "in": 1 reflection, 2 zero-gravity momentum flux; a
                               dry "ex": 3 reflection, 4 zero-gravity momentum flux) */
} oracle_t;

#define ZE(s, l, k) o->ze[((size_t)(s) * o->ne + (l)) * o->dofh + (k)]
#define QX(s, l, k) o->qx[((size_t)(s) * o->ne + (l)) * o->dofh + (k)]
#define QY(s, l, k) o->qy[((size_t)(s) * o->ne + (l)) * o->dofh + (k)]
#define RZE(s, l, k) o->rhs_ze[((size_t)(s) * o->ne + (l)) * o->dofh + (k)]
#define RQX(s, l, k) o->rhs_qx[((size_t)(s) * o->ne + (l)) * o->dofh + (k)]
#define RQY(s, l, k) o->rhs_qy[((size_t)(s) * o->ne + (l)) * o->dofh + (k)]
#define HB(l, k) o->hb[(size_t)(l) * o->dofh + (k)]
#define BATHED(l, led, gp) o->bathed[((size_t)(l) * 3 + (led)) * o->negp + (gp)]
#define SFACED(l, led, gp) o->sfaced[((size_t)(l) * 3 + (led)) * o->negp + (gp)]
#define EDGEQ(led, gp, k) o->edgeq[((led) * o->negp + (gp)) * o->dofh + (k)]

static void *xcalloc(size_t n, size_t s) {
    void *p = calloc(n ? n : 1, s);
    if (!p) { fprintf(stderr, "oracle: out of memory\n"); abort(); }
    return p;
}
static double *dup_d(const double *a, size_t n) {
    double *b = (double *)xcalloc(n, sizeof(double));
    if (a) memcpy(b, a, n * sizeof(double));
    return b;
}

/* ------------------------------------------------------------------------------------------ */
/* construction                                                                                */
/* ------------------------------------------------------------------------------------------ */
oracle_t *oracle_new(int np, int ne, int p, int rk_stage, int rk_order) {
    oracle_t *o = (oracle_t *)xcalloc(1, sizeof(oracle_t));
    o->np = np; o->ne = ne; o->p = p; o->dofh = (p + 1) * (p + 2) / 2;
    o->nrk = rk_stage; o->rk_order = rk_order;
    o->nagp = (p == 1) ? 3 : 6;
    o->negp = p + 2;
    o->nthreads = 1;
    return o;
}

void oracle_set_threads(oracle_t *o, int n) { o->nthreads = n > 0 ? n : 1; }

/* dparams: g dtdp statim dramp h0 cf hbreak ftheta fgamma sfea0
 * iparams: ics nolibf nolifa nolica nolicat nramp slopeflag rainfall */
void oracle_set_params(oracle_t *o, const double *d, const int *i) {
    o->g = d[0]; o->dtdp = d[1]; o->statim = d[2]; o->dramp = d[3]; o->h0 = d[4];
    o->cf = d[5]; o->hbreak = d[6]; o->ftheta = d[7]; o->fgamma = d[8]; o->sfea0 = d[9];
    o->ics = i[0]; o->nolibf = i[1]; o->nolifa = i[2]; o->nolica = i[3]; o->nolicat = i[4];
    o->nramp = i[5]; o->slopeflag = i[6]; o->rainfall = i[7];
}

/* x,y projected node coordinates; nm is [ne][3] 1-based (Fortran NM(:,1..3) per element). */
void oracle_set_mesh(oracle_t *o, const double *x, const double *y, const double *dp,
                     const double *sfac, const double *corif, const double *manning,
                     const int *nm) {
    o->x = dup_d(x, o->np); o->y = dup_d(y, o->np); o->dp = dup_d(dp, o->np);
    o->sfac = dup_d(sfac, o->np); o->corif = dup_d(corif, o->np);
    o->load_manning = manning != NULL;
    o->manning = manning ? dup_d(manning, o->np) : NULL;
    o->nm = (int *)xcalloc((size_t)o->ne * 3, sizeof(int));
    for (size_t k = 0; k < (size_t)o->ne * 3; k++) o->nm[k] = nm[k] - 1;
}

void oracle_set_open(oracle_t *o, int nseg, const int *len, const int *nodes) {
    o->nope = nseg;
    o->open_len = (int *)xcalloc(nseg, sizeof(int));
    int tot = 0;
    for (int k = 0; k < nseg; k++) { o->open_len[k] = len[k]; tot += len[k]; }
    o->neta = tot;
    o->open_nodes = (int *)xcalloc(tot, sizeof(int));
    for (int k = 0; k < tot; k++) o->open_nodes[k] = nodes[k] - 1;
}

void oracle_set_land(oracle_t *o, int nseg, const int *type, const int *len, const int *nodes) {
    o->nbou = nseg;
    o->land_type = (int *)xcalloc(nseg, sizeof(int));
    o->land_len = (int *)xcalloc(nseg, sizeof(int));
    int tot = 0;
    for (int k = 0; k < nseg; k++) { o->land_type[k] = type[k]; o->land_len[k] = len[k]; tot += len[k]; }
    o->land_nodes = (int *)xcalloc(tot, sizeof(int));
    for (int k = 0; k < tot; k++) o->land_nodes[k] = nodes[k] - 1;
}

void oracle_set_tide(oracle_t *o, int nbfr, const double *amig, const double *ff,
                     const double *face, const double *per, const double *emo, const double *efa) {
    o->nbfr = nbfr;
    o->amig = dup_d(amig, nbfr); o->ff = dup_d(ff, nbfr); o->face = dup_d(face, nbfr);
    o->per = dup_d(per, nbfr);
    o->emo = dup_d(emo, (size_t)nbfr * o->neta); o->efa = dup_d(efa, (size_t)nbfr * o->neta);
}

void oracle_set_flow(oracle_t *o, int nffr, int nflownodes, const double *famig,
                     const double *fff, const double *fface, const double *fper,
                     const double *qnam, const double *qnph) {
    o->nffr = nffr; o->nflownodes = nflownodes;
    o->famig = dup_d(famig, nffr); o->fff = dup_d(fff, nffr); o->fface = dup_d(fface, nffr);
    o->fper = dup_d(fper, nffr);
    o->qnam = dup_d(qnam, (size_t)nffr * nflownodes); o->qnph = dup_d(qnph, (size_t)nffr * nflownodes);
}

/* Tidal potential constituents (src/read_input.F:1856-1923: FACET in radians, PERT = 2*PI/AMIGT)
   and, for NTIP = 2, the self-attraction/load terms SALTAMP, SALTPHA (radians) as [ntif][np]. */
void oracle_set_tidal_potential(oracle_t *o, int ntip, int ntif, const double *tpk, const double *amigt,
                                const double *etrf, const double *fft, const double *facet,
                                const double *pert, const double *saltamp, const double *saltpha,
                                const double *slam, const double *sfea, double reftim) {
    o->ntip = ntip; o->ntif = ntif; o->reftim = reftim;
    o->tpk = dup_d(tpk, ntif); o->amigt = dup_d(amigt, ntif); o->etrf = dup_d(etrf, ntif);
    o->fft = dup_d(fft, ntif); o->facet = dup_d(facet, ntif); o->pert = dup_d(pert, ntif);
    o->saltamp = (double *)xcalloc((size_t)ntif * o->np + 1, sizeof(double));
    o->saltpha = (double *)xcalloc((size_t)ntif * o->np + 1, sizeof(double));
    if (saltamp) memcpy(o->saltamp, saltamp, (size_t)ntif * o->np * sizeof(double));
    if (saltpha) memcpy(o->saltpha, saltpha, (size_t)ntif * o->np * sizeof(double));
    o->tslam = dup_d(slam, o->np); o->tsfea = dup_d(sfea, o->np);
    o->tip2 = (double *)xcalloc(o->np, sizeof(double));
}

/* fort.19 (src/cstart.F:80-97): ETIMINC and nrec records of the NETA open-boundary node values;
   nbd: the open-boundary nodes in NBD order (0-based). */
int oracle_set_elev_series(oracle_t *o, int neta, const int *nbd, double etiminc, int nrec,
                           const double *esbin) {
    if (nrec < 2) return 1;
    o->eser = 1; o->eser_nrec = nrec; o->etiminc = etiminc;
    o->nbd = (int *)xcalloc(neta, sizeof(int));
    memcpy(o->nbd, nbd, neta * sizeof(int));
    o->esbin = dup_d(esbin, (size_t)nrec * neta);
    o->esbin1 = dup_d(esbin, neta);
    o->esbin2 = dup_d(esbin + neta, neta);
    o->eser_next = 2;
    o->etime1 = o->statim * 86400.0;
    o->etime2 = o->etime1 + etiminc;
    o->eta2bc = (double *)xcalloc(o->np, sizeof(double));
    o->neta = neta;
    return 0;
}

/* fort.20 (src/cstart.F:101-124): FTIMINC, the NVEL land-boundary node codes LBCODEI and nrec
   records of one value per flux node (codes 2, 12, 22) in NVEL order. */
int oracle_set_flow_series(oracle_t *o, int nvel, const int *lbcodei, double ftiminc, int nrec,
                           const double *qnin) {
    if (nrec < 2) return 1;
    o->fser = 1; o->fser_nrec = nrec; o->ftiminc = ftiminc; o->nvel = nvel;
    o->lbcodei = (int *)xcalloc(nvel, sizeof(int));
    memcpy(o->lbcodei, lbcodei, nvel * sizeof(int));
    int nf = 0;
    for (int j = 0; j < nvel; j++) if (lbcodei[j] == 2 || lbcodei[j] == 12 || lbcodei[j] == 22) nf++;
    o->nfluxnodes = nf;
    o->qnin = dup_d(qnin, (size_t)nrec * nf);
    /* QNIN has one slot past NVEL: the edge of the last flux node reads QNIN(I+1) */
    o->qnin1 = (double *)xcalloc(nvel + 1, sizeof(double));
    o->qnin2 = (double *)xcalloc(nvel + 1, sizeof(double));
    for (int j = 0, k = 0; j < nvel; j++)
        if (lbcodei[j] == 2 || lbcodei[j] == 12 || lbcodei[j] == 22) { o->qnin1[j] = qnin[k]; o->qnin2[j] = qnin[nf + k]; k++; }
    o->fser_next = 2;
    o->qtime1 = o->statim * 86400.0;
    o->qtime2 = o->qtime1 + ftiminc;
    return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* basis and quadrature                                                                        */
/* ------------------------------------------------------------------------------------------ */
/* Triangle Dubiner basis at (X,Y), degree d <= 2; ordering r = (q, p) with q total degree
 * (src/orthogonal_basis_v1.F:114-239). */
static void ortho_basis(double X, double Y, int d, double *basis, double *dbx, double *dby) {
    double PHI[3][3], D1[3][3], D2[3][3];
    memset(PHI, 0, sizeof PHI); memset(D1, 0, sizeof D1); memset(D2, 0, sizeof D2);
    PHI[0][0] = 1.0; D1[0][0] = 0.0; D2[0][0] = 0.0;
    PHI[1][0] = (1.0 + 2.0 * X + Y) / 2.0; D1[1][0] = 1.0; D2[1][0] = 1.0 / 2.0;
    for (int P = 1; P <= d - 1; P++) {
        double a = (2.0 * P + 1.0) / (P + 1.0);
        double b = P / (P + 1.0);
        double om = (1.0 - Y) / 2.0; double om2 = om * om;
        PHI[P + 1][0] = a * PHI[1][0] * PHI[P][0] - b * om2 * PHI[P - 1][0];
        D1[P + 1][0] = a * (D1[1][0] * PHI[P][0] + PHI[1][0] * D1[P][0]) - b * om2 * D1[P - 1][0];
        D2[P + 1][0] = a * (D2[1][0] * PHI[P][0] + PHI[1][0] * D2[P][0])
                     - b * ((Y - 1) / 2 * PHI[P - 1][0] + om2 * D2[P - 1][0]);
    }
    for (int P = 0; P <= d - 1; P++) {
        double f = (1.0 + 2.0 * P + (3.0 + 2.0 * P) * Y) / 2.0;
        PHI[P][1] = PHI[P][0] * f;
        D1[P][1] = D1[P][0] * f;
        D2[P][1] = D2[P][0] * (1.0 + 2.0 * P + (3.0 + 2.0 * P) * Y) / 2.0 + PHI[P][0] * (3.0 + 2.0 * P) / 2.0;
    }
    for (int Q = 1; Q <= d - 1; Q++) {
        for (int P = 0; P <= d - Q - 1; P++) {
            double AQ = (2.0 * Q + 2.0 + 2.0 * P) * (2.0 * Q + 3.0 + 2.0 * P) / (2.0 * Q + 2.0) / (Q + 2.0 + 2.0 * P);
            double t = (2.0 * P + 1.0);
            double BQ = t * t * (2.0 * Q + 2.0 + 2.0 * P) / (2.0 * Q + 2.0) / (2.0 * Q + 2.0 * P + 1.0) / (Q + 2.0 + 2.0 * P);
            double CQ = (Q + 2.0 * P + 1.0) * Q * (2.0 * Q + 3.0 + 2.0 * P) / (Q + 1.0) / (Q + 2.0 + 2.0 * P) / (2.0 * Q + 2.0 * P + 1.0);
            PHI[P][Q + 1] = (AQ * Y + BQ) * PHI[P][Q] - CQ * PHI[P][Q - 1];
            D1[P][Q + 1] = (AQ * Y + BQ) * D1[P][Q] - CQ * D1[P][Q - 1];
            D2[P][Q + 1] = (AQ * Y + BQ) * D2[P][Q] + AQ * PHI[P][Q] - CQ * D2[P][Q - 1];
        }
    }
    int r = 0;
    for (int Q = 0; Q <= d; Q++)
        for (int P = 0; P <= Q; P++) {
            double v = PHI[P][Q - P], a = D1[P][Q - P], b = D2[P][Q - P];
            if (fabs(v) < 1.0e-15) v = 0.0;
            if (fabs(a) < 1.0e-15) a = 0.0;
            if (fabs(b) < 1.0e-15) b = 0.0;
            basis[r] = v; dbx[r] = a; dby[r] = b; r++;
        }
}

/* The modal basis at reference coordinates (r, s) (gauge points of the adjoint tests). */
void oracle_basis_at(int p, double r, double s, double *phi) {
    double dbx[6], dby[6];
    ortho_basis(r, s, p, phi, dbx, dby);
}

/* Symmetric triangle rule: multiplicity-3 orbits (src/quadrature_v1.F:245-275,804-862). */
static void tri_rule(int p, double *px, double *py, double *w) {
    double A[2], W[2]; int L;
    if (p == 1) { L = 1; A[0] = 1.0 / 6.0; W[0] = 1.0 / 3.0; }
    else {
        L = 2;
        A[0] = 0.4459484909159648863183292538830519; A[1] = 0.0915762135097707434595714634022015;
        W[0] = 0.2233815896780114656950070084331228; W[1] = 0.1099517436553218676383263249002105;
    }
    const double V1[2] = {-1.0, -1.0}, V2[2] = {1.0, -1.0}, V3[2] = {-1.0, 1.0};
    const double AREA = 2.0;
    int j = 0;
    for (int i = 0; i < L; i++) {
        double a = A[i], b = a, c = 1.0 - (a + b);
        px[j] = V1[0] * a + V2[0] * b + V3[0] * c; py[j] = V1[1] * a + V2[1] * b + V3[1] * c; w[j] = AREA * W[i];
        px[j + 1] = V1[0] * b + V2[0] * c + V3[0] * a; py[j + 1] = V1[1] * b + V2[1] * c + V3[1] * a; w[j + 1] = AREA * W[i];
        px[j + 2] = V1[0] * c + V2[0] * a + V3[0] * b; py[j + 2] = V1[1] * c + V2[1] * a + V3[1] * b; w[j + 2] = AREA * W[i];
        j += 3;
    }
}

/* Gauss-Lobatto-Legendre rule with n = 3 or 4 points, ascending (src/quad_rules_general.F:40-260).
 * Interior points are the roots of P'_{n-1}; weights 2/(n(n-1) P_{n-1}(x)^2) computed with the
 * same Legendre recurrence as the reference. */
static void lobatto(int n, double *x, double *w) {
    if (n == 3) { x[0] = -1.0; x[1] = 0.0; x[2] = 1.0; }
    else { double a = 1.0 / sqrt(5.0); x[0] = -1.0; x[1] = -a; x[2] = a; x[3] = 1.0; }
    int nin = n - 2; double nrl = (double)nin;
    for (int m = 0; m < n; m++) {
        if (m == 0) w[m] = 2.0 / ((nrl + 1.0) * (nrl + 2.0));
        else if (m == n - 1) w[m] = w[0];
        else {
            double po[6]; po[0] = 1.0; po[1] = x[m];
            for (int p = 1; p <= nin; p++) { double prl = p; po[p + 1] = ((2.0 * prl + 1.0) * x[m] * po[p] - prl * po[p - 1]) / (prl + 1.0); }
            w[m] = 2.0 / ((nrl + 1.0) * (nrl + 2.0) * po[nin + 1] * po[nin + 1]);
        }
    }
}

/* RK_TIME (src/prep_DG.F:1551-2078) for the schemes used by DG-SWEM inputs. */
static int rk_time(oracle_t *o) {
    int s = o->nrk, r = o->rk_order;
    double A[MAXRK][MAXRK], B[MAXRK][MAXRK], C[MAXRK][MAXRK];
    memset(A, 0, sizeof A); memset(B, 0, sizeof B); memset(C, 0, sizeof C);
    if (s > MAXRK) return -1;
#define SA(i, j, v) A[(i) - 1][(j) - 1] = (v)
#define SB(i, j, v) B[(i) - 1][(j) - 1] = (v)
    if (s == 1 && r == 1) { SA(1, 1, 1.0); SB(1, 1, 1.0); }
    else if (r == 2) {
        if (s == 3) {
            SA(1,1,1.0); SA(2,1,0.087353119859156); SA(3,1,0.344956917166841); SA(2,2,0.912646880140844);
            SA(3,3,0.655043082833159); SB(1,1,0.528005024856522); SB(3,1,0.022826837460491);
            SB(2,2,0.481882138633993); SB(3,3,0.345866039233415);
        } else if (s == 4) {
            SA(1,1,1.0); SA(2,1,0.394806441339829); SA(3,1,0.002797307087390); SA(4,1,0.252860909354373);
            SA(2,2,0.605193558660171); SA(3,3,0.997202692912610); SA(4,4,0.747139090645627);
            SB(1,1,0.406584463657504); SB(3,1,0.013637216641451); SB(4,1,0.016453567333598);
            SB(2,2,0.246062298456822); SB(3,3,0.405447122055692); SB(4,4,0.303775146447707);
        } else if (s == 5) {
            SA(1,1,1.0); SA(2,1,0.235593265061659); SA(3,1,0.174017972351526); SA(4,1,0.235264368870758);
            SA(5,1,0.141720372339803); SA(2,2,0.764406734938341); SA(4,2,0.000058643383967);
            SA(5,2,0.095374613155521); SA(3,3,0.825982027648475); SA(5,3,0.000311763705780);
            SA(4,4,0.764676987745275); SA(5,5,0.762593250798895);
            SB(1,1,0.324840618151514); SB(3,1,0.108822380501601); SB(4,1,0.054392262422093);
            SB(5,1,0.000000180291569); SB(2,2,0.248310356296551); SB(4,2,0.000019049753098);
            SB(5,2,0.030981548293401); SB(3,3,0.268312512443371); SB(5,3,0.000101273514903);
            SB(4,4,0.248398145385413); SB(5,5,0.247721262987686);
        } else {
            for (int I = 1; I <= s; I++)
                for (int J = 0; J <= s - 1; J++) {
                    if (J == I - 1 && I < s) { A[I - 1][J] = 1.0; B[I - 1][J] = 1.0 / (s - 1); }
                    else if (J == 0 && I == s) { A[I - 1][J] = 1.0 / s; }
                    else if (J == s - 1 && I == s) { A[I - 1][J] = (s - 1.0) / s; B[I - 1][J] = 1.0 / s; }
                }
        }
    } else if (s == 3 && r == 3) {
        SA(1,1,1.0); SA(2,1,3.0/4.0); SA(2,2,1.0/4.0); SA(3,1,1.0/3.0); SA(3,3,2.0/3.0);
        SB(1,1,1.0); SB(2,2,1.0/4.0); SB(3,3,2.0/3.0);
    } else return -1;
#undef SA
#undef SB
    for (int K = 0; K <= s - 1; K++)
        for (int I = 1; I <= s; I++) {
            double casum = 0.0;
            for (int L = K + 1; L <= I - 1; L++) casum = casum + C[L - 1][K] * A[I - 1][L];
            C[I - 1][K] = B[I - 1][K] + casum;
        }
    memset(o->dtvd, 0, sizeof o->dtvd);
    for (int K = 1; K <= s - 1; K++) {
        o->dtvd[K] = 0.0;
        for (int L = 0; L <= K - 1; L++) o->dtvd[K] = o->dtvd[K] + C[K - 1][L];
    }
    memcpy(o->atvd, A, sizeof A); memcpy(o->btvd, B, sizeof B);
    return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* CREATE_EDGE_DATA                                                                            */
/* ------------------------------------------------------------------------------------------ */
static int find_edge(oracle_t *o, const int *adj_off, const int *adj, int n1, int n2) {
    /* get_edge_no (src/count_edges.cpp:26-37): edge joining n1,n2 in either orientation */
    for (int k = adj_off[n1]; k < adj_off[n1 + 1]; k++) {
        int e = adj[k];
        int a = o->nedno[2 * e], b = o->nedno[2 * e + 1];
        if ((a == n1 && b == n2) || (a == n2 && b == n1)) return e;
    }
    return -1;
}

static int create_edges(oracle_t *o) {
    int ne = o->ne, np = o->np;
    int maxed = 3 * ne;
    o->nedno = (int *)xcalloc(2 * (size_t)maxed, sizeof(int));
    o->nedel = (int *)xcalloc(2 * (size_t)maxed, sizeof(int));
    o->nedsd = (int *)xcalloc(2 * (size_t)maxed, sizeof(int));
    o->neled = (int *)xcalloc(3 * (size_t)ne, sizeof(int));
    for (int k = 0; k < 2 * maxed; k++) o->nedel[k] = -1;
    char *edflg = (char *)xcalloc(3 * (size_t)ne, 1);
    /* node -> element lists (NDEL) */
    int *cnt = (int *)xcalloc(np + 1, sizeof(int));
    for (int e = 0; e < ne; e++) for (int k = 0; k < 3; k++) cnt[o->nm[3 * e + k] + 1]++;
    for (int i = 0; i < np; i++) cnt[i + 1] += cnt[i];
    int *ndel = (int *)xcalloc(cnt[np] + 1, sizeof(int));
    int *fill = (int *)xcalloc(np, sizeof(int));
    for (int e = 0; e < ne; e++) for (int k = 0; k < 3; k++) { int n = o->nm[3 * e + k]; ndel[cnt[n] + fill[n]++] = e; }
    int ned = 0;
    for (int iel = 0; iel < ne; iel++) {
        int n1 = o->nm[3 * iel], n2 = o->nm[3 * iel + 1], n3 = o->nm[3 * iel + 2];
        int led[3][2] = {{n2, n3}, {n3, n1}, {n1, n2}};
        for (int ied = 0; ied < 3; ied++) {
            if (edflg[3 * iel + ied]) continue;
            int i1 = led[ied][0], i2 = led[ied][1];
            int id = ned++;
            o->neled[3 * iel + ied] = id;
            o->nedno[2 * id] = i1; o->nedno[2 * id + 1] = i2;
            o->nedel[2 * id] = iel;
            edflg[3 * iel + ied] = 1;
            for (int jj = cnt[i1]; jj < cnt[i1 + 1]; jj++) {
                int jel = ndel[jj];
                if (jel == iel) continue;
                int found = 0;
                for (int jed = 0; jed < 3; jed++) {
                    int j1 = o->nm[3 * jel + (jed + 1) % 3], j2 = o->nm[3 * jel + (jed + 2) % 3];
                    if ((j1 == i1 && j2 == i2) || (j1 == i2 && j2 == i1)) {
                        if (edflg[3 * jel + jed]) { fprintf(stderr, "oracle: possible duplicate element %d\n", jel + 1); return -1; }
                        o->neled[3 * jel + jed] = id;
                        o->nedel[2 * id + 1] = jel;
                        edflg[3 * jel + jed] = 1;
                        found = 1; break;
                    }
                }
                if (found) break;
            }
        }
    }
    o->nedges = ned;
    free(edflg); free(ndel); free(fill); free(cnt);
    /* adjacency node -> edges for get_edge_no */
    int *aoff = (int *)xcalloc(np + 1, sizeof(int));
    for (int e = 0; e < ned; e++) { aoff[o->nedno[2 * e] + 1]++; aoff[o->nedno[2 * e + 1] + 1]++; }
    for (int i = 0; i < np; i++) aoff[i + 1] += aoff[i];
    int *adj = (int *)xcalloc(aoff[np] + 1, sizeof(int));
    int *af = (int *)xcalloc(np, sizeof(int));
    for (int e = 0; e < ned; e++) {
        int a = o->nedno[2 * e], b = o->nedno[2 * e + 1];
        adj[aoff[a] + af[a]++] = e; adj[aoff[b] + af[b]++] = e;
    }
    int *ncount = (int *)xcalloc(ned, sizeof(int));
    for (int e = 0; e < ned; e++) ncount[e] = -1;
    o->niedn = (int *)xcalloc(ned, sizeof(int));
    o->nledn = (int *)xcalloc(ned, sizeof(int));
    o->needn = (int *)xcalloc(ned, sizeof(int));
    o->nfedn = (int *)xcalloc(ned, sizeof(int));
    o->nieds = o->nleds = o->needs = o->nfeds = 0;
    for (int e = 0; e < ned; e++)
        if (o->nedel[2 * e] >= 0 && o->nedel[2 * e + 1] >= 0) { o->niedn[o->nieds++] = e; ncount[e] = 0; }
    /* open (elevation specified) edges */
    int off = 0;
    for (int k = 0; k < o->nope; k++) {
        for (int i = 0; i < o->open_len[k] - 1; i++) {
            int e = find_edge(o, aoff, adj, o->open_nodes[off + i], o->open_nodes[off + i + 1]);
            if (e < 0) { fprintf(stderr, "oracle: error in processing open ocean boundary\n"); return -1; }
            o->needn[o->needs++] = e; ncount[e] = 4;
        }
        off += o->open_len[k];
    }
    /* land / flow segments */
    off = 0;
    for (int k = 0; k < o->nbou; k++) {
        int t = o->land_type[k];
        for (int i = 0; i < o->land_len[k] - 1; i++) {
            int e = find_edge(o, aoff, adj, o->land_nodes[off + i], o->land_nodes[off + i + 1]);
            if (e < 0) continue;
            ncount[e] = 1;
            if (t == 0 || t == 10 || t == 20 || t == 1 || t == 11 || t == 21) o->nledn[o->nleds++] = e;
            else if (t == 2 || t == 12 || t == 22) o->nfedn[o->nfeds++] = e;
            else { fprintf(stderr, "oracle: boundary type %d not supported\n", t); return -1; }
        }
        off += o->land_len[k];
    }
    for (int e = 0; e < ned; e++) if (ncount[e] < 0) o->nledn[o->nleds++] = e;
    free(ncount); free(aoff); free(adj); free(af);
    /* NEDSD: local edge number of each side */
    for (int e = 0; e < ned; e++)
        for (int k = 0; k < 2; k++) {
            int el = o->nedel[2 * e + k];
            o->nedsd[2 * e + k] = -1;
            if (el < 0) continue;
            for (int s = 0; s < 3; s++) if (o->neled[3 * el + s] == e) o->nedsd[2 * e + k] = s;
        }
    /* elem_edge slots, in internal-edge order */
    o->elem_edge = (int *)xcalloc(3 * (size_t)ne, sizeof(int));
    int *slot = (int *)xcalloc(ne, sizeof(int));
    for (int k = 0; k < 3 * ne; k++) o->elem_edge[k] = -1;
    for (int i = 0; i < o->nieds; i++) {
        int ged = o->niedn[i];
        int a = o->nedel[2 * ged], b = o->nedel[2 * ged + 1];
        o->elem_edge[3 * a + slot[a]++] = ged;
        o->elem_edge[3 * b + slot[b]++] = ged;
    }
    free(slot);
    /* CALC_NORMAL */
    o->xlen = (double *)xcalloc(ned, sizeof(double));
    o->cosnx = (double *)xcalloc(ned, sizeof(double));
    o->sinnx = (double *)xcalloc(ned, sizeof(double));
    for (int e = 0; e < ned; e++) {
        int n1 = o->nedno[2 * e], n2 = o->nedno[2 * e + 1];
        double dy = o->y[n2] - o->y[n1], dx = o->x[n2] - o->x[n1];
        o->xlen[e] = sqrt(dy * dy + dx * dx);
        o->cosnx[e] = dy / o->xlen[e];
        o->sinnx[e] = -dx / o->xlen[e];
    }
    return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* PREP_DG                                                                                     */
/* ------------------------------------------------------------------------------------------ */
/* NEIGHB (src/dgswem.F:478-570): each node's neighbours in insertion order, then sorted by
   increasing clockwise angle from east; PREP_DG's ELETAB (src/prep_DG.F:292-330): for consecutive
   neighbour pairs (J2, J3) around node I, cyclically, the element holding I, J2 and J3.  The
   node's segment of o->eletab (filled in element order) is rewritten in that order; a node whose
   fan does not give all its elements keeps the element order. */
static void eletab_reference_order(oracle_t *o) {
    int np = o->np, ne = o->ne;
    int *cnt = (int *)xcalloc(np, sizeof(int));
    int cap = 0;
    for (int i = 0; i < np; i++) if (o->el_count[i] + 2 > cap) cap = o->el_count[i] + 2;
    int *nb = (int *)xcalloc((size_t)np * cap, sizeof(int));
    for (int n = 0; n < ne; n++) {
        int en[3] = {o->nm[3 * n], o->nm[3 * n + 1], o->nm[3 * n + 2]};
        const int pr[3][2] = {{0, 1}, {0, 2}, {1, 2}};
        for (int q = 0; q < 3; q++) {
            int a = en[pr[q][0]], b = en[pr[q][1]], found = 0;
            for (int j = 0; j < cnt[a]; j++) if (nb[(size_t)a * cap + j] == b) { found = 1; break; }
            if (found) continue;
            if (cnt[a] >= cap || cnt[b] >= cap) continue;
            nb[(size_t)a * cap + cnt[a]++] = b;
            nb[(size_t)b * cap + cnt[b]++] = a;
        }
    }
    const double rad2deg = 45.0 / atan(1.0);
    double *ang = (double *)xcalloc(cap, sizeof(double));
    int *sorted = (int *)xcalloc(cap + 1, sizeof(int));
    int *tmp = (int *)xcalloc(cap + 1, sizeof(int));
    for (int i = 0; i < np; i++) {
        int nn = cnt[i];
        for (int j = 0; j < nn; j++) {
            int m = nb[(size_t)i * cap + j];
            double delx = o->x[m] - o->x[i], dely = o->y[m] - o->y[i];
            double dist = sqrt(delx * delx + dely * dely);
            ang[j] = 0.0;
            if (dely != 0.0) {
                ang[j] = rad2deg * acos(delx / dist);
                if (dely > 0.0) ang[j] = 360.0 - ang[j];
            }
            if (dely == 0.0) {
                if (delx > 0.0) ang[j] = 0.0;
                if (delx < 0.0) ang[j] = 180.0;
            }
        }
        double more = -1.0;
        int jlow = 0;
        for (int jj = 0; jj < nn; jj++) {
            double low = 400.0;
            for (int j = 0; j < nn; j++)
                if (ang[j] < low && ang[j] > more) { low = ang[j]; jlow = j; }
            sorted[1 + jj] = nb[(size_t)i * cap + jlow];
            more = low;
        }
        sorted[0] = i;
        int nneigh = nn + 1, kk = 0;
        const int *seg = o->eletab + o->eletab_off[i];
        for (int k = 1; k <= nneigh - 1; k++) {
            int j2 = sorted[k], j3 = (k < nneigh - 1) ? sorted[k + 1] : sorted[1];
            for (int j = 0; j < o->el_count[i]; j++) {
                int el = seg[j];
                const int *v = &o->nm[3 * el];
                int h1 = v[0] == i || v[1] == i || v[2] == i;
                int h2 = v[0] == j2 || v[1] == j2 || v[2] == j2;
                int h3 = v[0] == j3 || v[1] == j3 || v[2] == j3;
                if (h1 && h2 && h3) { if (kk < cap) tmp[kk] = el; kk++; break; }
            }
        }
        if (kk == o->el_count[i]) memcpy(o->eletab + o->eletab_off[i], tmp, kk * sizeof(int));
    }
    free(cnt); free(nb); free(ang); free(sorted); free(tmp);
}

int oracle_prep(oracle_t *o) {
    int ne = o->ne, dofh = o->dofh, nagp = o->nagp, negp = o->negp, p = o->p;
    if (p < 1 || p > 2) return -1;
    if (rk_time(o)) { fprintf(stderr, "oracle: RK scheme not supported\n"); return -1; }
    /* derived flags (src/dgswem.F:131-170) */
    o->ifnlfa = o->nolifa == 0 ? 0 : 1;
    o->nleq = (o->nolicat == 0) ? 0.0 : 1.0; o->leq = (o->nolicat == 0) ? 1.0 : 0.0;
    if (o->nolica == 0 && o->nolicat != 0) { /* NOLICAT read last wins (src/dgswem.F:150-165) */ }
    o->nleqg = o->nleq * o->g; o->fg_l = o->leq * o->g;
    o->iflinbf = (o->nolibf == 0) ? 1 : 0; o->ifhybf = (o->nolibf == 2) ? 1 : 0;

    if (create_edges(o)) return -1;

    /* tides on ocean edges (src/prep_DG.F:183-210) */
    o->emo_dg = (double *)xcalloc((size_t)o->needs * (o->nbfr ? o->nbfr : 1) * 2, sizeof(double));
    o->efa_dg = (double *)xcalloc((size_t)o->needs * (o->nbfr ? o->nbfr : 1) * 2, sizeof(double));
    for (int i = 0; i < o->nbfr; i++) {
        int ii = 0, jj = 0;
        for (int j = 0; j < o->nope; j++) {
            for (int k = 0; k < o->open_len[j] - 1; k++) {
                o->emo_dg[((size_t)ii * o->nbfr + i) * 2 + 0] = o->emo[(size_t)i * o->neta + jj];
                o->emo_dg[((size_t)ii * o->nbfr + i) * 2 + 1] = o->emo[(size_t)i * o->neta + jj + 1];
                o->efa_dg[((size_t)ii * o->nbfr + i) * 2 + 0] = o->efa[(size_t)i * o->neta + jj];
                o->efa_dg[((size_t)ii * o->nbfr + i) * 2 + 1] = o->efa[(size_t)i * o->neta + jj + 1];
                ii++; jj++;
            }
            jj++;
        }
    }
    /* flow edges (src/prep_DG.F:211-250) */
    int nffr = o->nffr;
    o->qnam_dg = (double *)xcalloc((size_t)o->nfeds * (nffr ? nffr : 1) * 2, sizeof(double));
    o->qnph_dg = (double *)xcalloc((size_t)o->nfeds * (nffr ? nffr : 1) * 2, sizeof(double));
    for (int i = 0; i < nffr; i++) {
        int ii = 0, jj = 0;
        for (int j = 0; j < o->nbou; j++) {
            int t = o->land_type[j];
            if (!(t == 2 || t == 12 || t == 22)) continue;
            for (int k = 0; k < o->land_len[j] - 1; k++) {
                o->qnam_dg[((size_t)ii * nffr + i) * 2 + 0] = o->qnam[(size_t)i * o->nflownodes + jj];
                o->qnam_dg[((size_t)ii * nffr + i) * 2 + 1] = o->qnam[(size_t)i * o->nflownodes + jj + 1];
                o->qnph_dg[((size_t)ii * nffr + i) * 2 + 0] = o->qnph[(size_t)i * o->nflownodes + jj];
                o->qnph_dg[((size_t)ii * nffr + i) * 2 + 1] = o->qnph[(size_t)i * o->nflownodes + jj + 1];
                ii++; jj++;
            }
            jj++;
        }
    }

    /* node -> element table: ELETAB, in PREP_DG's angular order (modal2nodal's summation order) */
    o->etamax = (double *)xcalloc(o->np, sizeof(double));
    for (int i = 0; i < o->np; i++) o->etamax[i] = -99999.0;      /* src/prep_DG.F:295 */
    o->el_count = (int *)xcalloc(o->np, sizeof(int));
    o->eletab_off = (int *)xcalloc(o->np + 1, sizeof(int));
    for (int e = 0; e < ne; e++) for (int k = 0; k < 3; k++) o->eletab_off[o->nm[3 * e + k] + 1]++;
    for (int i = 0; i < o->np; i++) { o->el_count[i] = o->eletab_off[i + 1]; o->eletab_off[i + 1] += o->eletab_off[i]; }
    o->eletab = (int *)xcalloc(o->eletab_off[o->np] + 1, sizeof(int));
    {
        int *f = (int *)xcalloc(o->np, sizeof(int));
        for (int e = 0; e < ne; e++) for (int k = 0; k < 3; k++) { int n = o->nm[3 * e + k]; o->eletab[o->eletab_off[n] + f[n]++] = e; }
        free(f);
    }
    eletab_reference_order(o);

    /* quadrature and basis tables */
    tri_rule(p, o->xagp, o->yagp, o->wagp);
    lobatto(negp, o->xegp, o->wegp);
    double bas[MAXDOF], dbx[MAXDOF], dby[MAXDOF];
    for (int q = 0; q < nagp; q++) {
        ortho_basis(o->xagp[q], o->yagp[q], p, bas, dbx, dby);
        for (int k = 0; k < dofh; k++) { o->phi_area[q][k] = bas[k]; o->drphi[q][k] = dbx[k]; o->dsphi[q][k] = dby[k]; }
    }
    const double VX[3] = {-1.0, 1.0, -1.0}, VY[3] = {-1.0, -1.0, 1.0};
    for (int v = 0; v < 3; v++) {
        ortho_basis(VX[v], VY[v], p, bas, dbx, dby);
        for (int k = 0; k < dofh; k++) o->phi_corner[v][k] = bas[k];
    }
    for (int m = 0; m < dofh; m++) o->m_inv[m] = 0.0;
    {
        int M = 0;
        for (int J = 0; J <= p; J++) for (int I = 0; I <= J; I++) {
            int JJ = J - I;
            o->m_inv[M++] = ((2.0 * I + 1.0) * (2.0 * JJ + 2.0 * I + 2.0) / 4.0);
        }
    }
    for (int ii = 0; ii < 3; ii++)
        for (int q = 0; q < negp; q++) {
            double X, Y, xe = o->xegp[q];
            if (ii == 0) { X = -xe; Y = xe; }
            else if (ii == 1) { X = -1.0; Y = -xe; }
            else { X = xe; Y = -1.0; }
            ortho_basis(X, Y, p, bas, dbx, dby);
            for (int k = 0; k < dofh; k++) o->phi_edge[ii][q][k] = bas[k];
        }
    for (int q = 0; q < nagp; q++) {
        o->psi1[q] = -1.0 / 2.0 * (o->xagp[q] + o->yagp[q]);
        o->psi2[q] = 1.0 / 2.0 * (o->xagp[q] + 1.0);
        o->psi3[q] = 1.0 / 2.0 * (o->yagp[q] + 1.0);
        o->drpsi[0] = -1.0 / 2.0; o->drpsi[1] = 1.0 / 2.0; o->drpsi[2] = 0.0;   /* src/prep_DG.F:948-953 */
        o->dspsi[0] = -1.0 / 2.0; o->dspsi[1] = 0.0; o->dspsi[2] = 1.0 / 2.0;
    }
    o->edgeq = (double *)xcalloc(3 * (size_t)negp * dofh, sizeof(double));
    for (int l = 0; l < 3; l++) for (int q = 0; q < negp; q++) for (int k = 0; k < dofh; k++)
        EDGEQ(l, q, k) = 2.0 * o->m_inv[k] * o->phi_edge[l][q][k] * o->wegp[q];

    /* per-element geometry and tables */
    o->areas = (double *)xcalloc(ne, sizeof(double));
    o->drdx = (double *)xcalloc(ne, sizeof(double)); o->dsdx = (double *)xcalloc(ne, sizeof(double));
    o->drdy = (double *)xcalloc(ne, sizeof(double)); o->dsdy = (double *)xcalloc(ne, sizeof(double));
    o->cori_el = (double *)xcalloc(ne, sizeof(double)); o->fric_el = (double *)xcalloc(ne, sizeof(double));
    o->dpe_min = (double *)xcalloc(ne, sizeof(double));
    o->hb = (double *)xcalloc((size_t)ne * dofh, sizeof(double));
    o->bath = (double *)xcalloc((size_t)ne * nagp, sizeof(double));
    o->dbathdx = (double *)xcalloc((size_t)ne * nagp, sizeof(double));
    o->dbathdy = (double *)xcalloc((size_t)ne * nagp, sizeof(double));
    o->sfac_elem = (double *)xcalloc((size_t)ne * nagp, sizeof(double));
    o->xfac = (double *)xcalloc((size_t)ne * nagp * dofh, sizeof(double));
    o->yfac = (double *)xcalloc((size_t)ne * nagp * dofh, sizeof(double));
    o->srfac = (double *)xcalloc((size_t)ne * nagp * dofh, sizeof(double));
    o->bathed = (double *)xcalloc((size_t)ne * 3 * negp, sizeof(double));
    o->sfaced = (double *)xcalloc((size_t)ne * 3 * negp, sizeof(double));
    o->dp_node = (double *)xcalloc((size_t)ne * 3, sizeof(double));
    const double R = 6378206.4;
    for (int J = 0; J < ne; J++) {
        int n1 = o->nm[3 * J], n2 = o->nm[3 * J + 1], n3 = o->nm[3 * J + 2];
        double X1 = o->x[n1], X2 = o->x[n2], X3 = o->x[n3];
        double Y1 = o->y[n1], Y2 = o->y[n2], Y3 = o->y[n3];
        /* AREAS = 2A (src/read_input.F:1408) */
        o->areas[J] = (X1 - X3) * (Y2 - Y3) + (X3 - X2) * (Y1 - Y3);
        /* P1 bathymetry modes (src/prep_DG.F:371-384) */
        double d1 = o->dp[n1], d2 = o->dp[n2], d3 = o->dp[n3];
        HB(J, 0) = 1.0 / 3.0 * (d1 + d2 + d3);
        HB(J, 1) = -(1.0 / 6.0 * (d1 + d2)) + 1.0 / 3.0 * d3;
        HB(J, 2) = -(0.5 * d1) + 0.5 * d2;
        double ydub[MAXDOF] = {0};
        ydub[0] = 1.0 / 3.0 * (Y1 + Y2 + Y3);
        ydub[1] = -(1.0 / 6.0 * (Y1 + Y2)) + 1.0 / 3.0 * Y3;
        ydub[2] = -(0.5 * Y1) + 0.5 * Y2;
        /* geometry (src/prep_DG.F:959-981) */
        double AREA = (X1 - X3) * (Y2 - Y3) + (X3 - X2) * (Y1 - Y3);
        AREA = AREA * 0.5;
        o->drdx[J] = 1.0 / AREA * (Y3 - Y1);
        o->dsdx[J] = 1.0 / AREA * (Y1 - Y2);
        o->drdy[J] = 1.0 / AREA * (X1 - X3);
        o->dsdy[J] = 1.0 / AREA * (X2 - X1);
        o->cori_el[J] = (o->corif[n1] + o->corif[n2] + o->corif[n3]) / 3.0;
        o->fric_el[J] = (o->cf + o->cf + o->cf) / 3.0;
        for (int I = 0; I < nagp; I++) {
            double bath = 0.0, dbx_ = 0.0, dby_ = 0.0, yel = 0.0;
            for (int K = 0; K < dofh; K++) {
                double dphidx = o->drphi[I][K] * o->drdx[J] + o->dsphi[I][K] * o->dsdx[J];
                double dphidy = o->drphi[I][K] * o->drdy[J] + o->dsphi[I][K] * o->dsdy[J];
                size_t t = ((size_t)J * nagp + I) * dofh + K;
                o->xfac[t] = o->m_inv[K] * o->wagp[I] * dphidx;
                o->yfac[t] = o->m_inv[K] * o->wagp[I] * dphidy;
                o->srfac[t] = o->m_inv[K] * o->wagp[I] * o->phi_area[I][K];
                bath = bath + HB(J, K) * o->phi_area[I][K];
                yel = yel + ydub[K] * o->phi_area[I][K];
                dbx_ = dbx_ + HB(J, K) * dphidx;
                dby_ = dby_ + HB(J, K) * dphidy;
            }
            o->bath[(size_t)J * nagp + I] = bath;
            o->dbathdx[(size_t)J * nagp + I] = dbx_;
            o->dbathdy[(size_t)J * nagp + I] = dby_;
            o->sfac_elem[(size_t)J * nagp + I] = (o->ics == 1) ? 1.0 : cos(o->sfea0) / cos(yel / R);
        }
        for (int L = 0; L < 3; L++)
            for (int I = 0; I < negp; I++) {
                double be = 0.0, yed = 0.0;
                for (int K = 0; K < dofh; K++) {
                    be = be + HB(J, K) * o->phi_edge[L][I][K];
                    yed = yed + ydub[K] * o->phi_edge[L][I][K];
                }
                BATHED(J, L, I) = be;
                SFACED(J, L, I) = (o->ics == 1) ? 1.0 : cos(o->sfea0) / cos(yed / R);
            }
        o->dp_node[3 * J] = d1; o->dp_node[3 * J + 1] = d2; o->dp_node[3 * J + 2] = d3;
        double m = d1 < d2 ? d1 : d2; o->dpe_min[J] = m < d3 ? m : d3;
    }

    /* state */
    size_t nst = (size_t)(o->nrk + 1) * ne * dofh;
    o->ze = (double *)xcalloc(nst, sizeof(double)); o->qx = (double *)xcalloc(nst, sizeof(double));
    o->qy = (double *)xcalloc(nst, sizeof(double));
    size_t nr = (size_t)o->nrk * ne * dofh;
    o->rhs_ze = (double *)xcalloc(nr, sizeof(double)); o->rhs_qx = (double *)xcalloc(nr, sizeof(double));
    o->rhs_qy = (double *)xcalloc(nr, sizeof(double));
    o->ze_edge = (double *)xcalloc((size_t)ne * 3 * dofh, sizeof(double));
    o->qx_edge = (double *)xcalloc((size_t)ne * 3 * dofh, sizeof(double));
    o->qy_edge = (double *)xcalloc((size_t)ne * 3 * dofh, sizeof(double));
    o->wdflg = (int *)xcalloc(ne, sizeof(int));
    o->nodecode = (int *)xcalloc(o->np, sizeof(int));
    for (int i = 0; i < o->np; i++) o->nodecode[i] = 1;
    o->zemin = (double *)xcalloc(o->np, sizeof(double)); o->zemax = (double *)xcalloc(o->np, sizeof(double));
    o->qxmin = (double *)xcalloc(o->np, sizeof(double)); o->qxmax = (double *)xcalloc(o->np, sizeof(double));
    o->qymin = (double *)xcalloc(o->np, sizeof(double)); o->qymax = (double *)xcalloc(o->np, sizeof(double));

    /* initial wet/dry state (src/prep_DG.F:1214-1300), cold start ZE=QX=QY=0 */
    if (o->nolifa == 0 || o->nolifa == 1) {
        for (int J = 0; J < ne; J++) o->wdflg[J] = 1;
    } else {
        double H0 = o->h0;
        for (int J = 0; J < ne; J++) {
            double z0 = ZE(0, J, 0);
            double z1 = z0, z2 = z0, z3 = z0;
            double dpa = o->dp[o->nm[3 * J]], dpb = o->dp[o->nm[3 * J + 1]], dpc = o->dp[o->nm[3 * J + 2]];
            if (dpa < H0) z1 = fmax(z1, H0 - dpa);
            if (dpb < H0) z2 = fmax(z2, H0 - dpb);
            if (dpc < H0) z3 = fmax(z3, H0 - dpc);
            if ((z1 + z2 + z3) / 3.0 != ZE(0, J, 0)) {
                if (ZE(0, J, 0) > (z1 + z2 + z3) / 3.0) {
                    for (int k = 1; k < dofh; k++) ZE(0, J, k) = 0.0;
                } else {
                    ZE(0, J, 0) = (z1 + z2 + z3) / 3.0;
                    ZE(0, J, 1) = -(1.0 / 6.0 * (z1 + z2)) + 1.0 / 3.0 * z3;
                    ZE(0, J, 2) = -(0.5 * z1) + 0.5 * z2;
                    for (int k = 3; k < dofh; k++) ZE(0, J, k) = 0.0;
                }
                o->wdflg[J] = 0;
            } else o->wdflg[J] = 1;
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------------------------ */
/* numerical flux (LLF)                                                                        */
/* ------------------------------------------------------------------------------------------ */
static void llf(const oracle_t *o, double ze_in, double ze_ex, double qx_in, double qx_ex,
                double qy_in, double qy_ex, double hb_in, double hb_ex, double sfac_in,
                double sfac_ex, double nx, double ny, double G, double *out) {
    double NLEQ = o->nleq, LEQ = o->leq, FG_L = o->fg_l;
    double NLEQG = G * NLEQ;
    double j1 = ze_ex - ze_in, j2 = qx_ex - qx_in, j3 = qy_ex - qy_in;
    double ht_in = ze_in * NLEQ + hb_in;
    double f1 = NLEQ + LEQ * hb_in;
    double fx1_in = qx_in * f1 * sfac_in, fy1_in = qy_in * f1;
    double fu = NLEQ * qx_in, fv = NLEQ * qy_in, fg = NLEQG * ze_in;
    double fh = 1.0 / (NLEQ * ht_in + LEQ);
    double u_in = qx_in * fh, v_in = qy_in * fh;
    double huu = fu * u_in, hvv = fv * v_in, huv = fu * v_in;
    double gh2 = fg * (0.5 * ze_in + hb_in) + FG_L * ze_in;
    double fx2_in = (huu + gh2) * sfac_in, fy2_in = huv;
    double fx3_in = huv * sfac_in, fy3_in = hvv + gh2;
    double ht_ex = ze_ex * NLEQ + hb_ex;
    f1 = NLEQ + LEQ * hb_ex;
    double fx1_ex = qx_ex * f1 * sfac_ex, fy1_ex = qy_ex * f1;
    fu = NLEQ * qx_ex; fv = NLEQ * qy_ex; fg = NLEQG * ze_ex;
    fh = 1.0 / (NLEQ * ht_ex + LEQ);
    double u_ex = qx_ex * fh, v_ex = qy_ex * fh;
    huu = fu * u_ex; hvv = fv * v_ex; huv = fu * v_ex;
    gh2 = fg * (0.5 * ze_ex + hb_ex) + FG_L * ze_ex;
    double fx2_ex = (huu + gh2) * sfac_ex, fy2_ex = huv;
    double fx3_ex = huv * sfac_ex, fy3_ex = hvv + gh2;
    double f1a = 0.5 * ((fx1_in + fx1_ex) * nx + (fy1_in + fy1_ex) * ny);
    double f2a = 0.5 * ((fx2_in + fx2_ex) * nx + (fy2_in + fy2_ex) * ny);
    double f3a = 0.5 * ((fx3_in + fx3_ex) * nx + (fy3_in + fy3_ex) * ny);
    double un_in = (u_in * nx + v_in * ny) * NLEQ;
    double nxs = nx * sfac_in;
    double c_in = sqrt(G * ht_in * (ny * ny + nxs * nxs));
    double e1 = fabs(un_in + c_in), e2 = fabs(un_in), e3 = fabs(un_in - c_in);
    double un_ex = (u_ex * nx + v_ex * ny) * NLEQ;
    nxs = nx * sfac_ex;
    double c_ex = sqrt(G * ht_ex * (ny * ny + nxs * nxs));
    double e4 = fabs(un_ex + c_ex), e5 = fabs(un_ex), e6 = fabs(un_ex - c_ex);
    double em = e1;
    if (e2 > em) em = e2; if (e3 > em) em = e3; if (e4 > em) em = e4; if (e5 > em) em = e5; if (e6 > em) em = e6;
    double nx2 = NLEQ + LEQ * nx * nx, ny2 = NLEQ + LEQ * ny * ny, nxy = nx * ny;
    out[0] = f1a - 0.5 * em * (j1);
    out[1] = f2a - 0.5 * em * (j2 * nx2 + LEQ * j3 * nxy);
    out[2] = f3a - 0.5 * em * (j3 * ny2 + LEQ * j2 * nxy);
}

/* ------------------------------------------------------------------------------------------ */
/* edge kernels                                                                                */
/* ------------------------------------------------------------------------------------------ */
static void ocean_edges(oracle_t *o) {
    int irk = o->irk, dofh = o->dofh, negp = o->negp;
    for (int L = 0; L < o->needs; L++) {
        int ged = o->needn[L], led = o->nedsd[2 * ged], el = o->nedel[2 * ged];
        if (o->wdflg[el] == 0) continue;
        double nx = o->cosnx[ged], ny = o->sinnx[ged];
        for (int I = 0; I < negp; I++) {
            double ze_in = ZE(irk, el, 0), qx_in = QX(irk, el, 0), qy_in = QY(irk, el, 0);
            double ze_ex = 0.0;
            double hb_in = BATHED(el, led, I), sfac_in = SFACED(el, led, I);
            for (int jj = 0; jj < o->nbfr; jj++) {
                int ncyc = (o->per[jj] == 0.0) ? 0 : (int)(o->timedg / o->per[jj]);
                double argj = o->amig[jj] * (o->timedg - ncyc * o->per[jj]) + o->face[jj];
                double rff = o->ff[jj] * o->rampdg;
                double zg[2];
                for (int ipt = 0; ipt < 2; ipt++) {
                    double efa = o->efa_dg[((size_t)L * o->nbfr + jj) * 2 + ipt];
                    double emo = o->emo_dg[((size_t)L * o->nbfr + jj) * 2 + ipt];
                    zg[ipt] = emo * rff * cos(argj - efa);
                }
                double zf = 0.5 * (zg[0] + zg[1]) + 0.5 * (zg[1] - zg[0]) * o->xegp[I];
                ze_ex = ze_ex + zf;
            }
            if (o->nbfr == 0 && o->eser) {
                /* non-periodic elevation: the boundary nodes' values of the stage */
                int n1 = o->nedno[2 * ged], n2 = o->nedno[2 * ged + 1];
                ze_ex = 0.5 * (o->eta2bc[n1] + o->eta2bc[n2]) + 0.5 * (o->eta2bc[n2] - o->eta2bc[n1]) * o->xegp[I];
            }
            for (int K = 1; K < dofh; K++) {
                double ph = o->phi_edge[led][I][K];
                ze_in = ze_in + ZE(irk, el, K) * ph;
                qx_in = qx_in + QX(irk, el, K) * ph;
                qy_in = qy_in + QY(irk, el, K) * ph;
            }
            double hb_ex = hb_in, sfac_ex = sfac_in;
            if ((ze_ex * o->ifnlfa + hb_ex) <= 0) ze_ex = fabs(hb_ex) + o->h0;
            double f[3];
            llf(o, ze_in, ze_ex, qx_in, qx_in, qy_in, qy_in, hb_in, hb_ex, sfac_in, sfac_ex, nx, ny, o->g, f);
            for (int K = 0; K < dofh; K++) {
                double w = 2.0 * o->m_inv[K] / o->areas[el] * o->xlen[ged] * o->phi_edge[led][I][K] * o->wegp[I];
                RZE(irk, el, K) = RZE(irk, el, K) - w * f[0];
                RQX(irk, el, K) = RQX(irk, el, K) - w * f[1];
                RQY(irk, el, K) = RQY(irk, el, K) - w * f[2];
            }
        }
    }
}

static void land_edges(oracle_t *o) {
    int irk = o->irk, dofh = o->dofh, negp = o->negp;
    for (int L = 0; L < o->nleds; L++) {
        int ged = o->nledn[L], led = o->nedsd[2 * ged], el = o->nedel[2 * ged];
        if (o->wdflg[el] == 0) continue;
        double nx = o->cosnx[ged], ny = o->sinnx[ged], tx = -ny, ty = nx;
        for (int I = 0; I < negp; I++) {
            double ze_in = ZE(irk, el, 0), qx_in = QX(irk, el, 0), qy_in = QY(irk, el, 0);
            double hb_in = BATHED(el, led, I), sfac_in = SFACED(el, led, I);
            for (int K = 1; K < dofh; K++) {
                double ph = o->phi_edge[led][I][K];
                ze_in = ze_in + ZE(irk, el, K) * ph;
                qx_in = qx_in + QX(irk, el, K) * ph;
                qy_in = qy_in + QY(irk, el, K) * ph;
            }
            double qn_int = qx_in * nx + qy_in * ny, qt_int = qx_in * tx + qy_in * ty;
            double qn_ext = -qn_int, qt_ext = qt_int;
            double den = 1.0 / (nx * ty - ny * tx);
            double qx_ex = (ty * qn_ext - ny * qt_ext) * den;
            double qy_ex = (-tx * qn_ext + nx * qt_ext) * den;
            double f[3];
            llf(o, ze_in, ze_in, qx_in, qx_ex, qy_in, qy_ex, hb_in, hb_in, sfac_in, sfac_in, nx, ny, o->g, f);
            for (int K = 0; K < dofh; K++) {
                double w = 2.0 * o->m_inv[K] / o->areas[el] * o->xlen[ged] * o->phi_edge[led][I][K] * o->wegp[I];
                RZE(irk, el, K) = RZE(irk, el, K) - w * f[0];
                RQX(irk, el, K) = RQX(irk, el, K) - w * f[1];
                RQY(irk, el, K) = RQY(irk, el, K) - w * f[2];
            }
        }
    }
}

static void flow_edges(oracle_t *o) {
    int irk = o->irk, dofh = o->dofh, negp = o->negp, nffr = o->nffr;
    for (int L = 0; L < o->nfeds; L++) {
        int ged = o->nfedn[L], led = o->nedsd[2 * ged], el = o->nedel[2 * ged];
        if (o->wdflg[el] == 0) continue;
        double nx = o->cosnx[ged], ny = o->sinnx[ged], tx = -ny, ty = nx;
        for (int I = 0; I < negp; I++) {
            double ze_in = 0.0, qx_in = 0.0, qy_in = 0.0, qx_ex = 0.0, qy_ex = 0.0, hb_in = 0.0;
            double sfac_in = SFACED(el, led, I);
            double qn_ext = 0.0, qt_ext = 0.0;
            /* NFFR = 0: one term from the time series, ARGJ = 0 and RFF = RAMPDG (the values are
               ramped once more where they are interpolated) */
            int nterm = o->fser ? 1 : nffr, nf1 = o->fser ? 1 : nffr;
            for (int jj = 0; jj < nterm; jj++) {
                double argj = 0.0, rff = o->rampdg;
                if (!o->fser) {
                    int ncyc = (o->fper[jj] == 0.0) ? 0 : (int)(o->timedg / o->fper[jj]);
                    argj = o->famig[jj] * (o->timedg - ncyc * o->fper[jj]) + o->fface[jj];
                    rff = o->fff[jj] * o->rampdg;
                }
                const double *qa = &o->qnam_dg[((size_t)L * nf1 + jj) * 2];
                const double *qp = &o->qnph_dg[((size_t)L * nf1 + jj) * 2];
                double qnam_gp = 0.5 * (qa[0] + qa[1]) + 0.5 * (qa[1] - qa[0]) * o->xegp[I];
                double qnph_gp = 0.5 * (qp[0] + qp[1]) + 0.5 * (qp[1] - qp[0]) * o->xegp[I];
                double arg = argj - qnph_gp;
                qn_ext = qn_ext + qnam_gp * rff * cos(arg);
                qt_ext = 0.0;
                qx_ex = -(ty * qn_ext - ny * qt_ext) / (nx * ty - ny * tx);
                qy_ex = -(-tx * qn_ext + nx * qt_ext) / (nx * ty - ny * tx);
            }
            for (int K = 0; K < dofh; K++) {
                double ph = o->phi_edge[led][I][K];
                ze_in = ze_in + ZE(irk, el, K) * ph;
                qx_in = qx_in + QX(irk, el, K) * ph;
                qy_in = qy_in + QY(irk, el, K) * ph;
                hb_in = hb_in + HB(el, K) * ph;
            }
            double f[3];
            llf(o, ze_in, ze_in, qx_in, qx_ex, qy_in, qy_ex, hb_in, hb_in, sfac_in, sfac_in, nx, ny, o->g, f);
            for (int K = 0; K < dofh; K++) {
                double w = 2.0 * o->m_inv[K] / o->areas[el] * o->xlen[ged] * o->phi_edge[led][I][K] * o->wegp[I];
                RZE(irk, el, K) = RZE(irk, el, K) - w * f[0];
                RQX(irk, el, K) = RQX(irk, el, K) - w * f[1];
                RQY(irk, el, K) = RQY(irk, el, K) - w * f[2];
            }
        }
    }
}

static void internal_edges(oracle_t *o) {
    int irk = o->irk, dofh = o->dofh, negp = o->negp;
    /* frozen-branch runs: this stage's dry-side records (one internal_edges call per stage, as
       one limiter call) */
    unsigned char *ie_calls = NULL;
    if (o->lim_mode != 0 && o->ie_rec) {
        int call = o->ie_pos++;
        if (call < o->lim_cap) ie_calls = o->ie_rec + (size_t)call * o->nieds * negp;
    }
#pragma omp parallel for num_threads(o->nthreads) schedule(static)
    for (int L = 0; L < o->nieds; L++) {
        double rzi[MAXDOF], rxi[MAXDOF], ryi[MAXDOF], rze[MAXDOF], rxe[MAXDOF], rye[MAXDOF];
        for (int k = 0; k < dofh; k++) rzi[k] = rxi[k] = ryi[k] = rze[k] = rxe[k] = rye[k] = 0.0;
        int ged = o->niedn[L];
        int led_in = o->nedsd[2 * ged], led_ex = o->nedsd[2 * ged + 1];
        int el_in = o->nedel[2 * ged], el_ex = o->nedel[2 * ged + 1];
        if (o->wdflg[el_in] == 0 && o->wdflg[el_ex] == 0) continue;
        double nx = o->cosnx[ged], ny = o->sinnx[ged], tx = -ny, ty = nx;
        double edfac_in = o->xlen[ged] / o->areas[el_in];
        double edfac_ex = o->xlen[ged] / o->areas[el_ex];
        unsigned char *side_rec = ie_calls ? ie_calls + (size_t)L * negp : NULL;
        for (int I = 0; I < negp; I++) {
            int gi = I, gx = negp - I - 1;
            double ze_in = 0.0, qx_in = 0.0, qy_in = 0.0, ze_ex = 0.0, qx_ex = 0.0, qy_ex = 0.0;
            double hb_in = BATHED(el_in, led_in, gi), sfac_in = SFACED(el_in, led_in, gi);
            double hb_ex = BATHED(el_ex, led_ex, gx), sfac_ex = SFACED(el_ex, led_ex, gx);
            for (int K = 0; K < dofh; K++) {
                double pi = o->phi_edge[led_in][gi][K], px = o->phi_edge[led_ex][gx][K];
                ze_in = ze_in + ZE(irk, el_in, K) * pi;
                qx_in = qx_in + QX(irk, el_in, K) * pi;
                qy_in = qy_in + QY(irk, el_in, K) * pi;
                ze_ex = ze_ex + ZE(irk, el_ex, K) * px;
                qx_ex = qx_ex + QX(irk, el_ex, K) * px;
                qy_ex = qy_ex + QY(irk, el_ex, K) * px;
            }
            double f[3];
            llf(o, ze_in, ze_ex, qx_in, qx_ex, qy_in, qy_ex, hb_in, hb_ex, sfac_in, sfac_ex, nx, ny, o->g, f);
            double f_o = f[0], g_in = f[1], h_in = f[2], g_ex = f[1], h_ex = f[2];
            /* the dry-side treatment (src/internal_edge_hydro.F:212-302), taken from f[0] and the
               flags, or replayed */
            int side = 0;
            if (side_rec && o->lim_mode == 2) {
                side = side_rec[I];
            } else if (fabs(f[0]) > 1.0e-8) {
                if (o->wdflg[el_in] == 0) side = f[0] > 0.0 ? 1 : 2;
                else if (o->wdflg[el_ex] == 0) side = f[0] < 0.0 ? 3 : 4;
            }
            if (side_rec && o->lim_mode == 1) side_rec[I] = (unsigned char)side;
            if (side != 0) {
                if (side <= 2) {
                    if (side == 1) {
                        double qn_ext = qx_ex * nx + qy_ex * ny, qt_ext = qx_ex * tx + qy_ex * ty;
                        double qn_int = -qn_ext, qt_int = qt_ext;
                        double den = (nx * ty - ny * tx);
                        qx_in = (ty * qn_int - ny * qt_int) / den;
                        qy_in = (-tx * qn_int + nx * qt_int) / den;
                        ze_in = ze_ex; hb_in = hb_ex; sfac_in = sfac_ex;
                        llf(o, ze_in, ze_ex, qx_in, qx_ex, qy_in, qy_ex, hb_in, hb_ex, sfac_in, sfac_ex, nx, ny, o->g, f);
                        f_o = f[0]; g_ex = f[1]; h_ex = f[2];
                        g_in = 0.0; h_in = 0.0;
                    } else {
                        llf(o, ze_in, ze_ex, qx_in, qx_ex, qy_in, qy_ex, hb_in, hb_ex, sfac_in, sfac_ex, nx, ny, 0.0, f);
                        g_in = f[1]; h_in = f[2];
                    }
                } else {
                    if (side == 3) {
                        double qn_int = qx_in * nx + qy_in * ny, qt_int = qx_in * tx + qy_in * ty;
                        double qn_ext = -qn_int, qt_ext = qt_int;
                        double den = (nx * ty - ny * tx);
                        qx_ex = (ty * qn_ext - ny * qt_ext) / den;
                        qy_ex = (-tx * qn_ext + nx * qt_ext) / den;
                        ze_ex = ze_in; hb_ex = hb_in; sfac_ex = sfac_in;
                        llf(o, ze_in, ze_ex, qx_in, qx_ex, qy_in, qy_ex, hb_in, hb_ex, sfac_in, sfac_ex, nx, ny, o->g, f);
                        f_o = f[0]; g_in = f[1]; h_in = f[2];
                        g_ex = 0.0; h_ex = 0.0;
                    } else {
                        llf(o, ze_in, ze_ex, qx_in, qx_ex, qy_in, qy_ex, hb_in, hb_ex, sfac_in, sfac_ex, nx, ny, 0.0, f);
                        g_ex = f[1]; h_ex = f[2];
                    }
                }
            }
            for (int K = 0; K < dofh; K++) {
                double w_in = edfac_in * EDGEQ(led_in, gi, K);
                double w_ex = edfac_ex * EDGEQ(led_ex, gx, K);
                rzi[K] = rzi[K] - w_in * f_o; rxi[K] = rxi[K] - w_in * g_in; ryi[K] = ryi[K] - w_in * h_in;
                rze[K] = rze[K] + w_ex * f_o; rxe[K] = rxe[K] + w_ex * g_ex; rye[K] = rye[K] + w_ex * h_ex;
            }
        }
        int ind_in = -1, ind_ex = -1;
        for (int s = 0; s < 3; s++) {
            if (o->elem_edge[3 * el_ex + s] == ged) ind_ex = s;
            if (o->elem_edge[3 * el_in + s] == ged) ind_in = s;
        }
        for (int K = 0; K < dofh; K++) {
            size_t a = ((size_t)el_ex * 3 + ind_ex) * dofh + K, b = ((size_t)el_in * 3 + ind_in) * dofh + K;
            o->ze_edge[a] = rze[K]; o->qx_edge[a] = rxe[K]; o->qy_edge[a] = rye[K];
            o->ze_edge[b] = rzi[K]; o->qx_edge[b] = rxi[K]; o->qy_edge[b] = ryi[K];
        }
    }
    /* add_internal_flux */
#pragma omp parallel for num_threads(o->nthreads) schedule(static)
    for (int l = 0; l < o->ne; l++) {
        for (int k = 0; k < dofh; k++) {
            double zs = 0.0, xs = 0.0, ys = 0.0;
            for (int i = 0; i < 3; i++) {
                size_t a = ((size_t)l * 3 + i) * dofh + k;
                zs = zs + o->ze_edge[a]; xs = xs + o->qx_edge[a]; ys = ys + o->qy_edge[a];
            }
            RZE(irk, l, k) = RZE(irk, l, k) + zs;
            RQX(irk, l, k) = RQX(irk, l, k) + xs;
            RQY(irk, l, k) = RQY(irk, l, k) + ys;
        }
        for (int i = 0; i < 3 * dofh; i++) {
            size_t a = (size_t)l * 3 * dofh + i;
            o->ze_edge[a] = 0.0; o->qx_edge[a] = 0.0; o->qy_edge[a] = 0.0;
        }
    }
}

/* ------------------------------------------------------------------------------------------ */
/* NWS=8 Holland vortex                                                                        */
/* ------------------------------------------------------------------------------------------ */
/* TIMECONV with IMIN=0, SEC=0: seconds since the start of the year (src/wind.F:315-348). */
static double timeconv(int iyr, int imo, int iday, int ihr) {
    double t = (double)((iday - 1) * 86400 + ihr * 3600 + 0 * 60) + 0.0;
    int ileap = (iyr / 4) * 4;
    if (imo >= 2) t = t + 31 * 86400;
    if (ileap == iyr && imo >= 3) t = t + 29 * 86400;
    if (ileap != iyr && imo >= 3) t = t + 28 * 86400;
    if (imo >= 4) t = t + 31 * 86400;
    if (imo >= 5) t = t + 30 * 86400;
    if (imo >= 6) t = t + 31 * 86400;
    if (imo >= 7) t = t + 30 * 86400;
    if (imo >= 8) t = t + 31 * 86400;
    if (imo >= 9) t = t + 31 * 86400;
    if (imo >= 10) t = t + 30 * 86400;
    if (imo >= 11) t = t + 31 * 86400;
    if (imo == 12) t = t + 30 * 86400;
    return t;
}

/* The track part of GetHollandStormData's first call (src/wind.F:2185-2410): drop repeated
 * hindcast lines and repeated forecast increments, give forecasts the nowcast's pressure and radii,
 * convert units and compute the translation velocities.  rec is [nrec][12] (see holland.py);
 * ref = IREFYR IREFMO IREFDAY IREFHR.  Returns 0, or -1 for an invalid record sequence. */
int oracle_set_holland(oracle_t *o, int nrec, const int *rec, const int *ref, int storm_number,
                       double bladj, const double *slam, const double *sfea) {
    const double re = 6378206.4;
    double pi = 4.0 * atan(1.0);
    double deg2rad = pi / 180.0;
    o->nws = 8;
    o->storm_number = storm_number;
    o->bladj = bladj;
    o->wind_ref_time = timeconv(ref[0], ref[1], ref[2], ref[3]);
    o->slam = dup_d(slam, o->np);
    o->sfea = dup_d(sfea, o->np);
    int n = nrec > 0 ? nrec : 1;
    int *iy = (int *)xcalloc(n, sizeof(int)), *im = (int *)xcalloc(n, sizeof(int));
    int *id = (int *)xcalloc(n, sizeof(int)), *ih = (int *)xcalloc(n, sizeof(int));
    int *ifc = (int *)xcalloc(n, sizeof(int));
    int *icp = (int *)xcalloc(n, sizeof(int)), *irrp = (int *)xcalloc(n, sizeof(int));
    int *irmw = (int *)xcalloc(n, sizeof(int));
    double *fcst = (double *)xcalloc(n, sizeof(double));
    o->hl_time = (double *)xcalloc(n, sizeof(double));
    o->hl_lat = (double *)xcalloc(n, sizeof(double)); o->hl_lon = (double *)xcalloc(n, sizeof(double));
    o->hl_cpress = (double *)xcalloc(n, sizeof(double)); o->hl_spd = (double *)xcalloc(n, sizeof(double));
    o->hl_rrp = (double *)xcalloc(n, sizeof(double)); o->hl_rmw = (double *)xcalloc(n, sizeof(double));
    o->hl_tvx = (double *)xcalloc(n, sizeof(double)); o->hl_tvy = (double *)xcalloc(n, sizeof(double));
    int i = 0, pl = 0, now_cp = 0, now_rmw = 0, now_rrp = 0, bad = 0;
    for (int r = 0; r < nrec && !bad; r++) {
        const int *R = rec + 12 * r;
        iy[i] = R[0]; im[i] = R[1]; id[i] = R[2]; ih[i] = R[3];
        ifc[i] = R[5]; icp[i] = R[9]; irrp[i] = R[10]; irmw[i] = R[11];
        int lat = R[6], lon = R[7], spd = R[8];
        if (R[4] == 0) {            /* BEST */
            if (i > 0 && iy[i] == iy[i - 1] && im[i] == im[i - 1] && id[i] == id[i - 1] && ih[i] == ih[i - 1])
                continue;
            now_cp = icp[i]; now_rmw = irmw[i]; now_rrp = irrp[i];
            o->hl_time[i] = timeconv(iy[i], im[i], id[i], ih[i]) - o->wind_ref_time;
            fcst[i] = ifc[i];
        } else if (R[4] == 1) {     /* OFCL */
            if (i == 0) { bad = 1; break; }
            if ((ifc[i] == 0 && iy[i] == iy[i - 1] && im[i] == im[i - 1] && id[i] == id[i - 1] && ih[i] == ih[i - 1]) ||
                (ifc[i] != 0 && ifc[i] == ifc[i - 1]))
                continue;
            fcst[i] = ifc[i];
            if (ifc[i] == 0) {
                o->hl_time[i] = timeconv(iy[i], im[i], id[i], ih[i]) - o->wind_ref_time;
            } else {
                fcst[i] = fcst[i] * 3600.0;
                o->hl_time[i] = o->hl_time[i - 1] + (fcst[i] - fcst[i - 1]);
            }
            icp[i] = now_cp; irmw[i] = now_rmw; irrp[i] = now_rrp;
        } else {
            bad = 1; break;
        }
        o->hl_lat[i] = lat; o->hl_lon[i] = lon; o->hl_spd[i] = spd;
        o->hl_cpress[i] = icp[i]; o->hl_rrp[i] = irrp[i]; o->hl_rmw[i] = irmw[i];
        o->hl_lat[i] = o->hl_lat[i] / 10.0;
        o->hl_lon[i] = o->hl_lon[i] / 10.0;
        o->hl_lon[i] = -1.0 * o->hl_lon[i];
        o->hl_cpress[i] = o->hl_cpress[i] * 100.0;
        o->hl_rrp[i] = o->hl_rrp[i] * 1.852000003180799 * 1000.0;
        o->hl_rmw[i] = o->hl_rmw[i] * 1.852000003180799;
        o->hl_spd[i] = o->hl_spd[i] * 0.51444444;
        pl = i + 1;
        i++;
    }
    free(iy); free(im); free(id); free(ih); free(ifc); free(icp); free(irrp); free(irmw); free(fcst);
    if (bad || pl < 2) return -1;
    double *LAT = o->hl_lat, *LON = o->hl_lon, *T = o->hl_time;
    for (int k = 1; k < pl; k++) {
        double s = sin((LON[k] * deg2rad - LON[k - 1] * deg2rad) / 2.0);
        o->hl_tvx[k] = copysign((re * (2.0 * asin(sqrt(cos(LAT[k] * deg2rad) * cos(LAT[k - 1] * deg2rad) * (s * s))))) /
                                    (T[k] - T[k - 1]),
                                LON[k] - LON[k - 1]);
        double q = sin((LAT[k] * deg2rad - LAT[k - 1] * deg2rad) / 2.0);
        o->hl_tvy[k] = copysign((re * (2.0 * asin(sqrt(q * q)))) / (T[k] - T[k - 1]), LAT[k] - LAT[k - 1]);
    }
    o->hl_tvx[0] = o->hl_tvx[1];
    o->hl_tvy[0] = o->hl_tvy[1];
    o->hl_n = pl;
    o->hl_first = 1;
    o->wsx2 = (double *)xcalloc(o->np, sizeof(double));
    o->wsy2 = (double *)xcalloc(o->np, sizeof(double));
    o->pr2 = (double *)xcalloc(o->np, sizeof(double));
    return 0;
}

/* MET_FORCING for NWS=8 at TIME_A with RampMete = RAMPDG (src/met_forcing.F:479-520): the track
 * interpolation of GetHollandStormData (src/wind.F:2363-2431, including its one-record-per-call
 * pointer advance), HollandGet (src/wind.F:1964-2137) and the Garratt drag law WindDrag
 * (src/wind.F:116-141). */
/* Storm scalars at `time`: s = {lat, lon, cpress, deficit, spd (adjusted), rmw, tvx, tvy, B}.
 * Returns nonzero when the track does not cover the time. */
static int storm_at(oracle_t *o, double time, double *s) {
    const double ambient = 101300.0, rho_air = 1.15, e = 2.718281828459045;
    double *T = o->hl_time;
    if (o->hl_first) {
        int i = 1;
        while (!(time >= T[i - 1] && time < T[i])) {
            i++;
            if (i >= o->hl_n) return 1;
        }
        o->hl_ptr = i;
        o->hl_first = 0;
    }
    int i = o->hl_ptr;
    if (time >= T[i]) i++;
    if (i >= o->hl_n) return 1;
    o->hl_ptr = i;
    double w = (time - T[i - 1]) / (T[i] - T[i - 1]);
    double cpress = o->hl_cpress[i - 1] + w * (o->hl_cpress[i] - o->hl_cpress[i - 1]);
    double lon = o->hl_lon[i - 1] + w * (o->hl_lon[i] - o->hl_lon[i - 1]);
    double lat = o->hl_lat[i - 1] + w * (o->hl_lat[i] - o->hl_lat[i - 1]);
    double spd = o->hl_spd[i - 1] + w * (o->hl_spd[i] - o->hl_spd[i - 1]);
    double rmw = o->hl_rmw[i - 1] + w * (o->hl_rmw[i] - o->hl_rmw[i - 1]);
    double tvx = o->hl_tvx[i - 1] + w * (o->hl_tvx[i] - o->hl_tvx[i - 1]);
    double tvy = o->hl_tvy[i - 1] + w * (o->hl_tvy[i] - o->hl_tvy[i - 1]);
    double deficit = ambient - cpress;
    if (deficit < 100.0) deficit = 100.0;
    double ts = sqrt(tvx * tvx + tvy * tvy);
    spd = spd - ts;
    spd = spd / o->bladj;
    double B = rho_air * e * (spd * spd) / deficit;
    if (B < 1.0) B = 1.0;
    if (B > 2.5) B = 2.5;
    s[0] = lat; s[1] = lon; s[2] = cpress; s[3] = deficit; s[4] = spd; s[5] = rmw; s[6] = tvx;
    s[7] = tvy; s[8] = B;
    return 0;
}

/* Storm scalars at increasing times (test access to the track interpolation). */
int oracle_storm_at(oracle_t *o, double time, double *s) { return storm_at(o, time, s); }

static void met_forcing(oracle_t *o, double time) {
    const double re = 6378206.4;
    const double sampling = 0.88, rad2deg = 180.0 / 3.141592653589793;
    double pi = 4.0 * atan(1.0);
    double deg2rad = pi / 180.0;
    double omega = 2.0 * pi / 86164.2;
    double S[9];
    if (storm_at(o, time, S)) { o->met_error = 1; return; }
    double lat = S[0], lon = S[1], cpress = S[2], deficit = S[3], spd = S[4], rmw = S[5];
    double tvx = S[6], tvy = S[7], B = S[8];
    double wmult = o->storm_number == 2 ? 1.2 : 1.0;
    double rampm = o->rampdg;
    double rhog = 1000.0 * o->g;
#pragma omp parallel for num_threads(o->nthreads) schedule(static)
    for (int n = 0; n < o->np; n++) {
        double xc = o->slam[n] * rad2deg, yc = o->sfea[n] * rad2deg;
        double dx = (xc - lon) * deg2rad, dy = (yc - lat) * deg2rad;
        double theta = atan2(dy, dx);
        double s1 = sin(dy / 2.0), s2 = sin(dx / 2.0);
        double rad = re * (2.0 * asin(sqrt(s1 * s1 + cos(lat * deg2rad) * cos(yc * deg2rad) * (s2 * s2))));
        double cor = 2.0 * omega * sin(yc * deg2rad);
        double rb = pow(rmw * 1000.0 / rad, B);
        double press = (cpress + deficit * exp(-rb)) / rhog;
        double vr = sqrt(rb * exp(1.0 - rb) * (spd * spd) + (rad * rad) * (cor * cor) / 4.0) - rad * cor / 2.0;
        double tsx = (fabs(vr) / spd) * tvx, tsy = (fabs(vr) / spd) * tvy;
        vr = vr * wmult;
        double wx = -(vr * sin(theta)), wy = vr * cos(theta);
        wx = wx * o->bladj; wy = wy * o->bladj;
        wx = wx * sampling; wy = wy * sampling;
        wx = wx + tsx; wy = wy + tsy;
        double wmag = sqrt(wx * wx + wy * wy);
        double cd = 0.001 * (0.75 + 0.067 * wmag);
        if (cd > 0.0035) cd = 0.0035;
        o->wsx2[n] = rampm * 0.001293 * cd * wx * wmag;
        o->wsy2[n] = rampm * 0.001293 * cd * wy * wmag;
        o->pr2[n] = rampm * press;
    }
}

/* ------------------------------------------------------------------------------------------ */
/* tidal potential and non-periodic boundary forcing of a stage                                */
/* ------------------------------------------------------------------------------------------ */
/* TIDAL_POTENTIAL (src/tidal_potential.F:15-80): TIP2 at every node for the stage. */
static void tidal_potential(oracle_t *o, int it) {
    double timeh = it * o->dtdp + (o->statim - o->reftim) * 86400.0;
    double timeh_dg = timeh - o->dtdp + o->dtvd[o->irk] * o->dtdp;
#pragma omp parallel for num_threads(o->nthreads) schedule(static)
    for (int i = 0; i < o->np; i++) {
        double tip = 0.0;
        for (int j = 0; j < o->ntif; j++) {
            int ncyc = (o->pert[j] == 0.0) ? 0 : (int)(timeh_dg / o->pert[j]);
            double argt = o->amigt[j] * (timeh_dg - ncyc * o->pert[j]) + o->facet[j];
            double tpmul = o->rampdg * o->etrf[j] * o->tpk[j] * o->fft[j];
            double saltmul = o->rampdg * o->fft[j];
            long na = lround(0.00014 / o->amigt[j]);
            double sa = o->saltamp[(size_t)j * o->np + i], sp = o->saltpha[(size_t)j * o->np + i];
            if (na == 1) {
                double argtp = argt + 2.0 * o->tslam[i];
                double argsalt = argt - sp;
                double ccsfea = cos(o->tsfea[i]);
                ccsfea = ccsfea * ccsfea;
                tip = tip + tpmul * ccsfea * cos(argtp) + saltmul * sa * cos(argsalt);
            }
            if (na == 2) {
                double argtp = argt + o->tslam[i];
                double argsalt = argt - sp;
                double s2sfea = sin(2.0 * o->tsfea[i]);
                tip = tip + tpmul * s2sfea * cos(argtp) + saltmul * sa * cos(argsalt);
            }
        }
        o->tip2[i] = tip;
    }
}

/* Non-periodic elevation (src/DG_hydro_timestep.F:153-169) and normal flux
   (src/DG_hydro_timestep.F:172-200) boundary values of the stage. */
static void boundary_series(oracle_t *o, int it) {
    double time_a = it * o->dtdp + o->statim * 86400.0;
    if (o->eser) {
        if (time_a > o->etime2) {
            o->etime1 = o->etime2;
            o->etime2 = o->etime1 + o->etiminc;
            if (o->eser_next >= o->eser_nrec) { o->series_error = 1; return; }
            memcpy(o->esbin1, o->esbin2, o->neta * sizeof(double));
            memcpy(o->esbin2, o->esbin + (size_t)o->eser_next * o->neta, o->neta * sizeof(double));
            o->eser_next++;
        }
        double etratio = (o->timedg - o->etime1) / o->etiminc;
        for (int i = 0; i < o->neta; i++)
            o->eta2bc[o->nbd[i]] = o->rampdg * (o->esbin1[i] + etratio * (o->esbin2[i] - o->esbin1[i]));
    }
    if (o->fser) {
        if (time_a > o->qtime2) {
            o->qtime1 = o->qtime2;
            o->qtime2 = o->qtime1 + o->ftiminc;
            if (o->fser_next >= o->fser_nrec) { o->series_error = 1; return; }
            const double *rec = o->qnin + (size_t)o->fser_next * o->nfluxnodes;
            for (int j = 0, k = 0; j < o->nvel; j++) {
                o->qnin1[j] = o->qnin2[j];
                o->qnin2[j] = 0.0;
                int c = o->lbcodei[j];
                if (c == 2 || c == 12 || c == 22) o->qnin2[j] = rec[k++];
            }
            o->fser_next++;
        }
        double qtratio = (o->timedg - o->qtime1) / o->ftiminc;
        int nqeds = 0;
        for (int i = 0; i < o->nvel; i++) {
            int c = o->lbcodei[i];
            if (c == 2 || c == 12 || c == 22) {
                nqeds++;
                if (nqeds <= o->nfeds) {
                    double *qa = &o->qnam_dg[(size_t)(nqeds - 1) * 2];
                    double *qp = &o->qnph_dg[(size_t)(nqeds - 1) * 2];
                    qa[0] = o->rampdg * (o->qnin1[i] + qtratio * (o->qnin2[i] - o->qnin1[i]));
                    qp[0] = 0.0;
                    qa[1] = o->rampdg * (o->qnin1[i + 1] + qtratio * (o->qnin2[i + 1] - o->qnin1[i + 1]));
                    qp[1] = 0.0;
                }
            }
        }
    }
}

/* ------------------------------------------------------------------------------------------ */
/* volume RHS                                                                                  */
/* ------------------------------------------------------------------------------------------ */
static void rhs_volume(oracle_t *o) {
    int irk = o->irk, dofh = o->dofh, nagp = o->nagp;
    double dtdph = 1.0 / o->dtdp;
#pragma omp parallel for num_threads(o->nthreads) schedule(static)
    for (int L = 0; L < o->ne; L++) {
        if (o->rainfall == 1) {
            double src = 7.0556e-6;
            for (int I = 0; I < nagp; I++)
                for (int K = 0; K < dofh; K++)
                    RZE(irk, L, K) = RZE(irk, L, K) + o->srfac[((size_t)L * nagp + I) * dofh + K] * src;
        }
        if (o->wdflg[L] == 0) continue;
        int n1 = o->nm[3 * L], n2 = o->nm[3 * L + 1], n3 = o->nm[3 * L + 2];
        double depth_c = HB(L, 0) + ZE(irk, L, 0);
        double fh_nl_c = 1.0 / (o->nleq * depth_c + o->leq);
        double fric = o->fric_el[L];
        for (int I = 0; I < nagp; I++) {
            double ze_in = ZE(irk, L, 0), qx_in = QX(irk, L, 0), qy_in = QY(irk, L, 0);
            double hb_in = o->bath[(size_t)L * nagp + I];
            double dhb_x = o->dbathdx[(size_t)L * nagp + I], dhb_y = o->dbathdy[(size_t)L * nagp + I];
            double sfq = o->sfac_elem[(size_t)L * nagp + I];
            double depth = 0.0;
            for (int K = 1; K < dofh; K++) {
                double ph = o->phi_area[I][K];
                ze_in = ze_in + ZE(irk, L, K) * ph;
                qx_in = qx_in + QX(irk, L, K) * ph;
                qy_in = qy_in + QY(irk, L, K) * ph;
                depth = ze_in + hb_in;
            }
            double f1 = o->nleq + o->leq * hb_in;
            double fx = qx_in * f1 * sfq, fy = qy_in * f1;
            double fu = o->nleq * qx_in, fv = o->nleq * qy_in;
            double fg = o->nleqg * ze_in * o->wdflg[L];
            double fh = 1.0 / (o->nleq * depth + o->leq);
            double u = qx_in * fh, v = qy_in * fh;
            double huu = fu * u, hvv = fv * v, huv = fu * v;
            double gh2 = fg * (0.5 * ze_in + hb_in) + o->fg_l * ze_in;
            double gx = (huu + gh2 + 0.0) * sfq, gy = huv + 0.0;
            double hx = (huv + 0.0) * sfq, hy = hvv + gh2 + 0.0;
            if (o->load_manning) {
                double nbar = (o->manning[n1] + o->manning[n2] + o->manning[n3]) / 3.0;
                fric = o->g * (nbar * nbar) / pow(depth, 1.0 / 3.0);
                if (fric < o->cf) fric = o->cf;
            }
            double tau = fric;
            if (o->iflinbf == 0) {
                double umag = sqrt(u * u + v * v);
                tau = tau * umag * fh;
                if (o->ifhybf == 1)
                    tau = tau * pow(1.0 + pow(o->hbreak * fh_nl_c, o->ftheta), o->fgamma / o->ftheta);
                double cap = 0.9 * dtdph;
                if (cap < tau) tau = cap;
            }
            double sx = -(tau * qx_in) + fg * dhb_x * sfq + o->cori_el[L] * qy_in;
            double sy = -(tau * qy_in) + fg * dhb_y - o->cori_el[L] * qx_in;
            if (o->nws != 0) {
                /* wind stress and pressure gradient (src/rhs_dg_hydro.F:362-373) */
                double dpsidx[3], dpsidy[3];
                for (int k = 0; k < 3; k++) {
                    dpsidx[k] = o->drpsi[k] * o->drdx[L] + o->dspsi[k] * o->dsdx[L];
                    dpsidy[k] = o->drpsi[k] * o->drdy[L] + o->dspsi[k] * o->dsdy[L];
                }
                double fw = 1.0 / f1;
                sx = sx + fw * (o->wsx2[n1] * o->psi1[I] + o->wsx2[n2] * o->psi2[I] + o->wsx2[n3] * o->psi3[I]) -
                     o->g * sfq * depth * (o->pr2[n1] * dpsidx[0] + o->pr2[n2] * dpsidx[1] + o->pr2[n3] * dpsidx[2]);
                sy = sy + fw * (o->wsy2[n1] * o->psi1[I] + o->wsy2[n2] * o->psi2[I] + o->wsy2[n3] * o->psi3[I]) -
                     o->g * depth * (o->pr2[n1] * dpsidy[0] + o->pr2[n2] * dpsidy[1] + o->pr2[n3] * dpsidy[2]);
            }
            if (o->ntip != 0) {
                /* tidal potential (src/rhs_dg_hydro.F:378-386) */
                double dpsidx[3], dpsidy[3];
                for (int k = 0; k < 3; k++) {
                    dpsidx[k] = o->drpsi[k] * o->drdx[L] + o->dspsi[k] * o->dsdx[L];
                    dpsidy[k] = o->drpsi[k] * o->drdy[L] + o->dspsi[k] * o->dsdy[L];
                }
                sx = sx + o->rampdg * o->g * depth * sfq *
                              (dpsidx[0] * o->tip2[n1] + dpsidx[1] * o->tip2[n2] + dpsidx[2] * o->tip2[n3]);
                sy = sy + o->rampdg * o->g * depth *
                              (dpsidy[0] * o->tip2[n1] + dpsidy[1] * o->tip2[n2] + dpsidy[2] * o->tip2[n3]);
            }
            size_t t0 = ((size_t)L * nagp + I) * dofh;
            RQX(irk, L, 0) = RQX(irk, L, 0) + o->srfac[t0] * sx;
            RQY(irk, L, 0) = RQY(irk, L, 0) + o->srfac[t0] * sy;
            for (int K = 1; K < dofh; K++) {
                double xf = o->xfac[t0 + K], yf = o->yfac[t0 + K], sr = o->srfac[t0 + K];
                RZE(irk, L, K) = RZE(irk, L, K) + xf * fx + yf * fy;
                RQX(irk, L, K) = RQX(irk, L, K) + xf * gx + yf * gy + sr * sx;
                RQY(irk, L, K) = RQY(irk, L, K) + xf * hx + yf * hy + sr * sy;
            }
        }
    }
}

/* ------------------------------------------------------------------------------------------ */
/* SSP-RK update, limiter, wet/dry                                                             */
/* ------------------------------------------------------------------------------------------ */
static void rk_update(oracle_t *o) {
    int irk = o->irk, dofh = o->dofh;
    double dt = o->dtdp;
#pragma omp parallel for num_threads(o->nthreads) schedule(static)
    for (int J = 0; J < o->ne; J++)
        for (int I = 0; I <= irk; I++) {
            double ark = o->atvd[irk][I], brk = o->btvd[irk][I] * dt;
            for (int K = 0; K < dofh; K++) {
                ZE(irk + 1, J, K) = ZE(irk + 1, J, K) + ark * ZE(I, J, K) + brk * RZE(I, J, K);
                QX(irk + 1, J, K) = QX(irk + 1, J, K) + ark * QX(I, J, K) + brk * RQX(I, J, K);
                QY(irk + 1, J, K) = QY(irk + 1, J, K) + ark * QY(I, J, K) + brk * RQY(I, J, K);
            }
        }
}

/* Branch bits of one field of one element (the order the adjoint kernel records them in):
   bits 0-5 vertex clamp (0 none, 1 to the max bound, 2 to the min bound; two bits per vertex),
   bits 6-8 the sign of SUMDIF of passes 1-3 (1 for +1), bits 9-17 DIF(K) > 0 of pass ll at bit
   9+3ll+k, bits 18-26 REDMAX chosen over REDFAC at bit 18+3ll+k. */
static void limiter_element(const oracle_t *o, const double zec[3], const double vmin[3],
                            const double vmax[3], int mode, int *bits, double out[2]) {
    double v[3] = {zec[0], zec[0], zec[0]};
    for (int kk = 1; kk < 3; kk++)
        for (int c = 0; c < 3; c++) v[c] = v[c] + o->phi_corner[c][kk] * zec[kk];
    int b = mode == 2 ? *bits : 0;
    for (int c = 0; c < 3; c++) {
        int sel;
        if (mode == 2) sel = (b >> (2 * c)) & 3;
        else sel = !(v[c] <= vmax[c]) ? 1 : (fmin(v[c], vmax[c]) < vmin[c] ? 2 : 0);
        v[c] = sel == 1 ? vmax[c] : (sel == 2 ? vmin[c] : v[c]);
        /* the free forward: fmax(fmin(v, vmax), vmin), the same value */
        if (mode != 2) b |= sel << (2 * c);
    }
    for (int ll = 0; ll < 3; ll++) {
        double sumloc = (v[0] + v[1] + v[2]) / 3.0;
        double sumdif = (sumloc - zec[0]) * 3.0;
        double signdif;
        if (mode == 2) signdif = (b >> (6 + ll)) & 1 ? 1.0 : -1.0;
        else { signdif = copysign(1.0, sumdif); if (signdif > 0) b |= 1 << (6 + ll); }
        double dif[3];
        int posb = 0;
        for (int c = 0; c < 3; c++) {
            dif[c] = (v[c] - zec[0]) * signdif;
            if (mode == 2) posb |= ((b >> (9 + 3 * ll + c)) & 1) << c;
            else if (dif[c] > 0) { posb |= 1 << c; b |= 1 << (9 + 3 * ll + c); }
        }
        int kdp = (posb & 1) + ((posb >> 1) & 1) + ((posb >> 2) & 1);
        for (int k = 0; k < 3; k++) {
            double div = fmax(1.0, (double)kdp);
            double redfac = 0.0;
            if ((posb >> k) & 1) { redfac = sumdif * signdif / div; kdp = kdp - 1; }
            double redmax = (signdif > 0) ? v[k] - vmin[k] : vmax[k] - v[k];
            int usemax;
            if (mode == 2) usemax = (b >> (18 + 3 * ll + k)) & 1;
            else { usemax = !(redfac <= redmax); if (usemax) b |= 1 << (18 + 3 * ll + k); }
            if (usemax) redfac = redmax;
            sumdif = sumdif - redfac * signdif;
            v[k] = v[k] - redfac * signdif;
        }
    }
    if (mode != 2) *bits = b;
    out[0] = -(1.0 / 6.0 * (v[0] + v[1])) + 1.0 / 3.0 * v[2];
    out[1] = -(0.5 * v[0]) + 0.5 * v[1];
}

/* SLOPELIMITER5 with its branches recorded (mode 1) or replayed (mode 2): the node bounds come
   from the recorded bound-attaining elements, the element pass from limiter_element.  The
   recorded map is what the adjoint differentiates; replayed at the recording point it gives the
   free limiter's result bit for bit. */
static void slopelimiter5_frozen(oracle_t *o) {
    int s = o->irk + 1, call = o->lim_pos++;
    if (call >= o->lim_cap) { o->series_error = 1; return; }
    int *nrec = o->lim_node + (size_t)call * o->np * 6;
    int *erec = o->lim_elem + (size_t)call * o->ne * 3;
    double *fld[3] = {o->ze, o->qx, o->qy};
    double *mnv[3] = {o->zemin, o->qxmin, o->qymin}, *mxv[3] = {o->zemax, o->qxmax, o->qymax};
    for (int i = 0; i < o->np; i++) {
        for (int f = 0; f < 3; f++) {
            int alo = -1, ahi = -1;
            double lo = 99999., hi = -99999.;
            if (o->lim_mode == 2) {
                alo = nrec[(size_t)i * 6 + 2 * f]; ahi = nrec[(size_t)i * 6 + 2 * f + 1];
                if (alo >= 0) lo = fld[f][((size_t)s * o->ne + alo) * o->dofh];
                if (ahi >= 0) hi = fld[f][((size_t)s * o->ne + ahi) * o->dofh];
            } else {
                for (int j = o->eletab_off[i]; j < o->eletab_off[i + 1]; j++) {
                    int e = o->eletab[j];
                    double c = fld[f][((size_t)s * o->ne + e) * o->dofh];
                    if (c < lo) { lo = c; alo = e; }
                    if (c > hi) { hi = c; ahi = e; }
                }
                nrec[(size_t)i * 6 + 2 * f] = alo; nrec[(size_t)i * 6 + 2 * f + 1] = ahi;
            }
            mnv[f][i] = lo; mxv[f][i] = hi;
        }
    }
    for (int I = 0; I < o->ne; I++) {
        int nn[3] = {o->nm[3 * I], o->nm[3 * I + 1], o->nm[3 * I + 2]};
        for (int f = 0; f < 3; f++) {
            double *u = &fld[f][((size_t)s * o->ne + I) * o->dofh];
            double zec[3] = {u[0], u[1], u[2]};
            double vmax[3] = {mxv[f][nn[0]], mxv[f][nn[1]], mxv[f][nn[2]]};
            double vmin[3] = {mnv[f][nn[0]], mnv[f][nn[1]], mnv[f][nn[2]]};
            double out[2];
            limiter_element(o, zec, vmin, vmax, o->lim_mode, &erec[(size_t)I * 3 + f], out);
            u[1] = out[0];
            u[2] = out[1];
        }
    }
}

/* Record the limiter branches of the next ncalls limiter calls (NRK per step). */
void oracle_limiter_record(oracle_t *o, int ncalls) {
    free(o->lim_node); free(o->lim_elem); free(o->wd_rec); free(o->ie_rec);
    o->lim_node = (int *)xcalloc((size_t)ncalls * o->np * 6 + 1, sizeof(int));
    o->lim_elem = (int *)xcalloc((size_t)ncalls * o->ne * 3 + 1, sizeof(int));
    o->wd_rec = (int *)xcalloc((size_t)ncalls * o->ne + 1, sizeof(int));
    o->ie_rec = (unsigned char *)xcalloc((size_t)ncalls * o->nieds * o->negp + 1, 1);
    o->lim_cap = ncalls; o->lim_pos = 0; o->wd_pos = 0; o->ie_pos = 0; o->lim_mode = 1;
}
/* Copy out / in the records of ncalls calls ([call][np][6], [call][ne][3], [call][ne] and
   [call][nieds][negp]). */
int oracle_limiter_get(const oracle_t *o, int *node, int *elem, int *wd, unsigned char *ie) {
    if (!o->lim_node) return 1;
    memcpy(node, o->lim_node, (size_t)o->lim_cap * o->np * 6 * sizeof(int));
    memcpy(elem, o->lim_elem, (size_t)o->lim_cap * o->ne * 3 * sizeof(int));
    memcpy(wd, o->wd_rec, (size_t)o->lim_cap * o->ne * sizeof(int));
    memcpy(ie, o->ie_rec, (size_t)o->lim_cap * o->nieds * o->negp);
    return 0;
}
/* Replay: the next ncalls limiter, WETDRY and internal-edge dry-side decisions take the given
   branches. */
void oracle_limiter_replay(oracle_t *o, int ncalls, const int *node, const int *elem, const int *wd,
                           const unsigned char *ie) {
    oracle_limiter_record(o, ncalls);
    memcpy(o->lim_node, node, (size_t)ncalls * o->np * 6 * sizeof(int));
    memcpy(o->lim_elem, elem, (size_t)ncalls * o->ne * 3 * sizeof(int));
    memcpy(o->wd_rec, wd, (size_t)ncalls * o->ne * sizeof(int));
    memcpy(o->ie_rec, ie, (size_t)ncalls * o->nieds * o->negp);
    o->lim_mode = 2;
}
/* Internal edges (the size of a dry-side record's [nieds] dimension). */
int oracle_num_internal_edges(const oracle_t *o) { return o->nieds; }

static void slopelimiter5(oracle_t *o) {
    if (o->lim_mode != 0) { slopelimiter5_frozen(o); return; }
    int s = o->irk + 1;
#pragma omp parallel for num_threads(o->nthreads) schedule(static)
    for (int i = 0; i < o->np; i++) {
        double zmin = 99999., zmax = -99999., xmin = 99999., xmax = -99999., ymin = 99999., ymax = -99999.;
        for (int j = o->eletab_off[i]; j < o->eletab_off[i + 1]; j++) {
            int e = o->eletab[j];
            double z = ZE(s, e, 0), x = QX(s, e, 0), y = QY(s, e, 0);
            zmin = fmin(zmin, z); xmin = fmin(xmin, x); ymin = fmin(ymin, y);
            zmax = fmax(zmax, z); xmax = fmax(xmax, x); ymax = fmax(ymax, y);
        }
        o->zemin[i] = zmin; o->zemax[i] = zmax; o->qxmin[i] = xmin; o->qxmax[i] = xmax;
        o->qymin[i] = ymin; o->qymax[i] = ymax;
    }
#pragma omp parallel for num_threads(o->nthreads) schedule(static)
    for (int I = 0; I < o->ne; I++) {
        int nn[3] = {o->nm[3 * I], o->nm[3 * I + 1], o->nm[3 * I + 2]};
        for (int ivar = 0; ivar < 3; ivar++) {
            double *U = ivar == 0 ? o->ze : (ivar == 1 ? o->qx : o->qy);
            const double *mn = ivar == 0 ? o->zemin : (ivar == 1 ? o->qxmin : o->qymin);
            const double *mx = ivar == 0 ? o->zemax : (ivar == 1 ? o->qxmax : o->qymax);
            double *u = &U[((size_t)s * o->ne + I) * o->dofh];
            double zec[3] = {u[0], u[1], u[2]};
            double vmax[3] = {mx[nn[0]], mx[nn[1]], mx[nn[2]]}, vmin[3] = {mn[nn[0]], mn[nn[1]], mn[nn[2]]};
            double v[3] = {zec[0], zec[0], zec[0]};
            for (int kk = 1; kk < 3; kk++)
                for (int c = 0; c < 3; c++) v[c] = v[c] + o->phi_corner[c][kk] * zec[kk];
            for (int c = 0; c < 3; c++) v[c] = fmax(fmin(v[c], vmax[c]), vmin[c]);
            for (int ll = 0; ll < 3; ll++) {
                double sumloc = (v[0] + v[1] + v[2]) / 3.0;
                double sumdif = (sumloc - zec[0]) * 3.0;
                double signdif = copysign(1.0, sumdif);
                double dif[3];
                for (int c = 0; c < 3; c++) dif[c] = (v[c] - zec[0]) * signdif;
                int kdp = (dif[0] > 0) + (dif[1] > 0) + (dif[2] > 0);
                for (int k = 0; k < 3; k++) {
                    double div = fmax(1.0, (double)kdp);
                    double redfac;
                    if (dif[k] > 0) { redfac = sumdif * signdif / div; kdp = kdp - 1; }
                    else redfac = 0;
                    double redmax = (signdif > 0) ? v[k] - vmin[k] : vmax[k] - v[k];
                    redfac = fmin(redfac, redmax);
                    sumdif = sumdif - redfac * signdif;
                    v[k] = v[k] - redfac * signdif;
                }
            }
            u[1] = -(1.0 / 6.0 * (v[0] + v[1])) + 1.0 / 3.0 * v[2];
            u[2] = -(0.5 * v[0]) + 0.5 * v[1];
        }
    }
}

static void wetdry(oracle_t *o) {
    int s = o->irk + 1, dofh = o->dofh;
    const double ZERO = 1.0e-6, HUGEV = 1.0e6, H0 = o->h0;
    const double C13 = 1.0 / 3.0, C16 = 1.0 / 6.0;
    for (int i = 0; i < o->np; i++) o->nodecode[i] = 0;
    /* elements are independent; NODECODE only ever receives the value 1 (atomic writes) */
#pragma omp parallel for num_threads(o->nthreads) schedule(static)
    for (int J = 0; J < o->ne; J++) {
        int no[3] = {o->nm[3 * J], o->nm[3 * J + 1], o->nm[3 * J + 2]};
        double *ze = &o->ze[((size_t)s * o->ne + J) * dofh];
        double *qx = &o->qx[((size_t)s * o->ne + J) * dofh];
        double *qy = &o->qy[((size_t)s * o->ne + J) * dofh];
        const double *dpn = &o->dp_node[3 * J];
        int ndry = 0, ht_max_i = 0;
        double ht_min = HUGEV, ht_max = ZERO;
        double ze_node[3], ht_node[3], ze_hat[3];
        for (int I = 0; I < 3; I++) {
            ze_node[I] = ze[0];
            for (int K = 1; K < 3; K++) ze_node[I] = ze_node[I] + o->phi_corner[I][K] * ze[K];
            ht_node[I] = dpn[I] + ze_node[I];
            if (ht_node[I] < ht_min) ht_min = ht_node[I];
            if (ht_node[I] > ht_max) { ht_max_i = I; ht_max = ht_node[I]; }
            if (ht_node[I] <= (H0 + ZERO)) ndry++;
        }
        double dp_avg = C13 * (dpn[0] + dpn[1] + dpn[2]);
        int check = 0;
        if (ndry == 0) {
            if (o->wdflg[J] == 1) {
                check = 0;
                for (int c = 0; c < 3; c++) {
#pragma omp atomic write
                    o->nodecode[no[c]] = 1;
                }
            } else {
                for (int c = 0; c < 3; c++) ze_hat[c] = ze_node[c];
                check = 1;
            }
        } else if (ndry == 3) {
            double ht_avg = ze[0] + dp_avg;
            for (int k = 0; k < dofh; k++) ze[k] = 0.0;
            ze[0] = ht_avg - dp_avg;
            if (ht_avg < ZERO) ze[0] = ZERO - dp_avg;
            ze[1] = -HB(J, 1); ze[2] = -HB(J, 2);
            for (int k = 0; k < dofh; k++) { qx[k] = 0.0; qy[k] = 0.0; }
            o->wdflg[J] = 0;
            check = 0;
        } else {
            double ht_avg = ze[0] + dp_avg;
            if (ht_avg <= (H0 + ZERO)) {
                for (int k = 0; k < dofh; k++) ze[k] = 0.0;
                ze[0] = ht_avg - dp_avg;
                if (ht_avg < ZERO) ze[0] = ZERO - dp_avg;
                ze[1] = -HB(J, 1); ze[2] = -HB(J, 2);
                for (int c = 0; c < 3; c++) ze_hat[c] = ze[0];
                for (int k = 0; k < dofh; k++) { qx[k] = 0.0; qy[k] = 0.0; }
            } else {
                int l2h[3] = {0, 1, 2};
                for (int k1 = 0; k1 < 3; k1++)
                    for (int k2 = k1; k2 < 3; k2++)
                        if (ht_node[l2h[k1]] > ht_node[l2h[k2]]) { int t = l2h[k1]; l2h[k1] = l2h[k2]; l2h[k2] = t; }
                double dh[3];
                dh[l2h[0]] = fmax(H0, ht_node[l2h[0]]) - ht_node[l2h[0]];
                dh[l2h[1]] = fmax(H0, ht_node[l2h[1]] - 0.5 * dh[l2h[0]]) - ht_node[l2h[1]];
                dh[l2h[2]] = -dh[l2h[0]] - dh[l2h[1]];
                int nd = 0; double qxs = 0, qys = 0; ht_max = 0;
                double ht_hat[3], qxn[3], qyn[3];
                for (int I = 0; I < 3; I++) {
                    ht_hat[I] = ht_node[I] + dh[I];
                    qxn[I] = qx[0]; qyn[I] = qy[0];
                    if (ht_hat[I] > ht_max) { ht_max_i = I; ht_max = ht_hat[I]; }
                    for (int K = 1; K < 3; K++) {
                        qxn[I] = qxn[I] + o->phi_corner[I][K] * qx[K];
                        qyn[I] = qyn[I] + o->phi_corner[I][K] * qy[K];
                    }
                    if (ht_hat[I] <= (H0 + 1.0e-10)) { nd++; qxs = qxs + qxn[I]; qys = qys + qyn[I]; }
                }
                if (nd == 3) {
                    for (int k = 0; k < dofh; k++) { qx[k] = 0.0; qy[k] = 0.0; }
                } else {
                    qxs = qxs / (3.0 - (double)nd);
                    qys = qys / (3.0 - (double)nd);
                    double qxh[3], qyh[3];
                    for (int I = 0; I < 3; I++) {
                        if (ht_hat[I] <= (H0 + 1.0e-10)) { qxh[I] = 0.0; qyh[I] = 0.0; }
                        else { qxh[I] = qxn[I] + qxs; qyh[I] = qyn[I] + qys; }
                    }
                    qx[0] = C13 * (qxh[0] + qxh[1] + qxh[2]);
                    qx[1] = -(C16 * (qxh[0] + qxh[1])) + C13 * qxh[2];
                    qx[2] = -(0.5 * qxh[0]) + 0.5 * qxh[1];
                    qy[0] = C13 * (qyh[0] + qyh[1] + qyh[2]);
                    qy[1] = -(C16 * (qyh[0] + qyh[1])) + C13 * qyh[2];
                    qy[2] = -(0.5 * qyh[0]) + 0.5 * qyh[1];
                }
                for (int c = 0; c < 3; c++) ze_hat[c] = ht_hat[c] - dpn[c];
                ze[1] = -(C16 * (ze_hat[0] + ze_hat[1])) + C13 * ze_hat[2];
                ze[2] = -(0.5 * ze_hat[0]) + 0.5 * ze_hat[1];
            }
            for (int k = 3; k < dofh; k++) { ze[k] = 0.0; qx[k] = 0.0; qy[k] = 0.0; }
            check = 1;
        }
        if (check) {
            if (ze_hat[ht_max_i] > (-o->dpe_min[J] + H0)) {
                o->wdflg[J] = 1;
                for (int c = 0; c < 3; c++) {
#pragma omp atomic write
                    o->nodecode[no[c]] = 1;
                }
            } else {
                for (int k = 0; k < dofh; k++) { qx[k] = 0.0; qy[k] = 0.0; }
                o->wdflg[J] = 0;
            }
        }
        (void)ht_min;
    }
}

/* WETDRY with its branches recorded (lim_mode 1) or replayed (lim_mode 2), the frozen-branch map
   the adjoint differentiates.  The code of one element (wetdry_code):
     bits 0-1  branch: 0 identity (every vertex wet, element wet), 1 every vertex wet but the element
               dry (the check alone), 2 flat (every vertex dry, or the mean depth at most H0+ZERO),
               3 redistribution
     bit  2    flat: the mean depth floored at ZERO
     bit  3    flat: the element had a dry vertex but not all three (the check follows)
     bits 4-6  redistribution: the vertices in increasing depth, as a permutation index
     bits 7-8  redistribution: the two fmax(H0, ...) took H0
     bits 9-11 redistribution: the vertices dry after the redistribution
     bit  12   the check kept the momentum (element wet)
     bit  13   the element's flag afterwards
   Replayed at the recording point every branch gives the free WETDRY's values bit for bit. */
static const int PERM3[6][3] = {{0, 1, 2}, {0, 2, 1}, {1, 0, 2}, {1, 2, 0}, {2, 0, 1}, {2, 1, 0}};

static void wetdry_frozen(oracle_t *o) {
    int s = o->irk + 1, dofh = o->dofh, call = o->wd_pos++;
    if (call >= o->lim_cap) { o->series_error = 1; return; }
    int *rec = o->wd_rec + (size_t)call * o->ne;
    const int mode = o->lim_mode;
    const double ZERO = 1.0e-6, H0 = o->h0;
    const double C13 = 1.0 / 3.0, C16 = 1.0 / 6.0;
    for (int i = 0; i < o->np; i++) o->nodecode[i] = 0;
#pragma omp parallel for num_threads(o->nthreads) schedule(static)
    for (int J = 0; J < o->ne; J++) {
        int no[3] = {o->nm[3 * J], o->nm[3 * J + 1], o->nm[3 * J + 2]};
        double *ze = &o->ze[((size_t)s * o->ne + J) * dofh];
        double *qx = &o->qx[((size_t)s * o->ne + J) * dofh];
        double *qy = &o->qy[((size_t)s * o->ne + J) * dofh];
        const double *dpn = &o->dp_node[3 * J];
        double ze_node[3], ht_node[3], ze_hat[3];
        int ndry = 0, ht_max_i = 0;
        double ht_max = ZERO;
        for (int I = 0; I < 3; I++) {
            ze_node[I] = ze[0];
            for (int K = 1; K < 3; K++) ze_node[I] = ze_node[I] + o->phi_corner[I][K] * ze[K];
            ht_node[I] = dpn[I] + ze_node[I];
            if (ht_node[I] > ht_max) { ht_max_i = I; ht_max = ht_node[I]; }
            if (ht_node[I] <= (H0 + ZERO)) ndry++;
        }
        double dp_avg = C13 * (dpn[0] + dpn[1] + dpn[2]);
        int code = 0, branch, check = 0;
        if (mode == 2) {
            code = rec[J];
            branch = code & 3;
        } else {
            if (ndry == 0) branch = o->wdflg[J] == 1 ? 0 : 1;
            else if (ndry == 3) branch = 2;
            else branch = (ze[0] + dp_avg <= (H0 + ZERO)) ? 2 : 3;
            code = branch;
        }
        if (branch == 0) {
            code |= 1 << 13;
        } else if (branch == 1) {
            for (int c = 0; c < 3; c++) ze_hat[c] = ze_node[c];
            check = 1;
        } else if (branch == 2) {
            double ht_avg = ze[0] + dp_avg;
            int floor_, partial;
            if (mode == 2) { floor_ = (code >> 2) & 1; partial = (code >> 3) & 1; }
            else {
                floor_ = ht_avg < ZERO; partial = ndry != 3;
                code |= floor_ << 2 | partial << 3;
            }
            for (int k = 0; k < dofh; k++) ze[k] = 0.0;
            ze[0] = floor_ ? ZERO - dp_avg : ht_avg - dp_avg;
            ze[1] = -HB(J, 1); ze[2] = -HB(J, 2);
            for (int k = 0; k < dofh; k++) { qx[k] = 0.0; qy[k] = 0.0; }
            if (partial) { for (int c = 0; c < 3; c++) ze_hat[c] = ze[0]; check = 1; }
        } else {
            int pi, b0, b1, dmask;
            if (mode == 2) { pi = (code >> 4) & 7; b0 = (code >> 7) & 1; b1 = (code >> 8) & 1; dmask = (code >> 9) & 7; }
            else {
                int l2h[3] = {0, 1, 2};
                for (int k1 = 0; k1 < 3; k1++)
                    for (int k2 = k1; k2 < 3; k2++)
                        if (ht_node[l2h[k1]] > ht_node[l2h[k2]]) { int t = l2h[k1]; l2h[k1] = l2h[k2]; l2h[k2] = t; }
                pi = 0;
                for (int q = 0; q < 6; q++)
                    if (PERM3[q][0] == l2h[0] && PERM3[q][1] == l2h[1]) pi = q;
                b0 = H0 > ht_node[l2h[0]];
                b1 = 0; dmask = 0;   /* set below */
            }
            const int *l2h = PERM3[pi];
            double dh[3];
            dh[l2h[0]] = (b0 ? H0 : ht_node[l2h[0]]) - ht_node[l2h[0]];
            if (mode != 2) b1 = H0 > ht_node[l2h[1]] - 0.5 * dh[l2h[0]];
            dh[l2h[1]] = (b1 ? H0 : ht_node[l2h[1]] - 0.5 * dh[l2h[0]]) - ht_node[l2h[1]];
            dh[l2h[2]] = -dh[l2h[0]] - dh[l2h[1]];
            double ht_hat[3], qxn[3], qyn[3];
            int nd = 0; double qxs = 0, qys = 0; ht_max = 0;
            for (int I = 0; I < 3; I++) {
                ht_hat[I] = ht_node[I] + dh[I];
                qxn[I] = qx[0]; qyn[I] = qy[0];
                if (ht_hat[I] > ht_max) { ht_max_i = I; ht_max = ht_hat[I]; }
                for (int K = 1; K < 3; K++) {
                    qxn[I] = qxn[I] + o->phi_corner[I][K] * qx[K];
                    qyn[I] = qyn[I] + o->phi_corner[I][K] * qy[K];
                }
                if (mode != 2 && ht_hat[I] <= (H0 + 1.0e-10)) dmask |= 1 << I;
                if ((dmask >> I) & 1) { nd++; qxs = qxs + qxn[I]; qys = qys + qyn[I]; }
            }
            if (mode != 2) code |= pi << 4 | b0 << 7 | b1 << 8 | dmask << 9;
            if (nd == 3) {
                for (int k = 0; k < dofh; k++) { qx[k] = 0.0; qy[k] = 0.0; }
            } else {
                qxs = qxs / (3.0 - (double)nd);
                qys = qys / (3.0 - (double)nd);
                double qxh[3], qyh[3];
                for (int I = 0; I < 3; I++) {
                    if ((dmask >> I) & 1) { qxh[I] = 0.0; qyh[I] = 0.0; }
                    else { qxh[I] = qxn[I] + qxs; qyh[I] = qyn[I] + qys; }
                }
                qx[0] = C13 * (qxh[0] + qxh[1] + qxh[2]);
                qx[1] = -(C16 * (qxh[0] + qxh[1])) + C13 * qxh[2];
                qx[2] = -(0.5 * qxh[0]) + 0.5 * qxh[1];
                qy[0] = C13 * (qyh[0] + qyh[1] + qyh[2]);
                qy[1] = -(C16 * (qyh[0] + qyh[1])) + C13 * qyh[2];
                qy[2] = -(0.5 * qyh[0]) + 0.5 * qyh[1];
            }
            for (int c = 0; c < 3; c++) ze_hat[c] = ht_hat[c] - dpn[c];
            ze[1] = -(C16 * (ze_hat[0] + ze_hat[1])) + C13 * ze_hat[2];
            ze[2] = -(0.5 * ze_hat[0]) + 0.5 * ze_hat[1];
            check = 1;
        }
        if (branch >= 2) for (int k = 3; k < dofh; k++) { ze[k] = 0.0; qx[k] = 0.0; qy[k] = 0.0; }
        if (check) {
            int keep;
            if (mode == 2) keep = (code >> 12) & 1;
            else { keep = ze_hat[ht_max_i] > (-o->dpe_min[J] + H0); code |= keep << 12 | keep << 13; }
            if (!keep) for (int k = 0; k < dofh; k++) { qx[k] = 0.0; qy[k] = 0.0; }
        }
        if (mode != 2) rec[J] = code;
        o->wdflg[J] = (code >> 13) & 1;
        if (o->wdflg[J]) {
            for (int c = 0; c < 3; c++) {
#pragma omp atomic write
                o->nodecode[no[c]] = 1;
            }
        }
    }
}

/* ------------------------------------------------------------------------------------------ */
/* time step                                                                                   */
/* ------------------------------------------------------------------------------------------ */
static void stage_time(oracle_t *o, int it, int irk) {
    double itdt = it * o->dtdp;
    double time_a = itdt + o->statim * 86400.0;
    o->timedg = time_a - o->dtdp + o->dtvd[irk] * o->dtdp;
    o->rampdg = 1.0;
    if (o->nramp == 1) o->rampdg = tanh((2.0 * ((it - 1) + o->dtvd[irk]) * o->dtdp / 86400.0) / o->dramp);
}

/* RHS of stage irk (0-based) at step it, accumulated into RHS(irk) (assumed zero). */
void oracle_rhs(oracle_t *o, int it, int irk) {
    o->irk = irk;
    stage_time(o, it, irk);
    if (o->nws == 8) met_forcing(o, it * o->dtdp + o->statim * 86400.0);
    if (o->ntip != 0) tidal_potential(o, it);
    /* NBFR = 0 with open boundaries, NFFR = 0 with flux boundaries: the time series must be set */
    if ((o->needs > 0 && o->nbfr == 0 && !o->eser) || (o->nfeds > 0 && o->nffr == 0 && !o->fser)) {
        o->series_error = 1;
        return;
    }
    if (o->eser || o->fser) boundary_series(o, it);
    if (o->needs > 0) ocean_edges(o);
    if (o->nleds > 0) land_edges(o);
    if (o->nfeds > 0) flow_edges(o);
    internal_edges(o);
    rhs_volume(o);
}

/* One DG_HYDRO_TIMESTEP(it) (src/DG_hydro_timestep.F:24-744). */
void oracle_step(oracle_t *o, int it) {
    for (int irk = 0; irk < o->nrk; irk++) {
        oracle_rhs(o, it, irk);
        rk_update(o);
        if (o->slopeflag != 0) slopelimiter5(o);
        if (o->nolifa >= 2) {
            if (o->lim_mode != 0) wetdry_frozen(o);
            else wetdry(o);
        }
    }
    size_t n = (size_t)o->ne * o->dofh;
    memcpy(o->ze, o->ze + (size_t)o->nrk * n, n * sizeof(double));
    memcpy(o->qx, o->qx + (size_t)o->nrk * n, n * sizeof(double));
    memcpy(o->qy, o->qy + (size_t)o->nrk * n, n * sizeof(double));
    memset(o->ze + n, 0, (size_t)o->nrk * n * sizeof(double));
    memset(o->qx + n, 0, (size_t)o->nrk * n * sizeof(double));
    memset(o->qy + n, 0, (size_t)o->nrk * n * sizeof(double));
    memset(o->rhs_ze, 0, (size_t)o->nrk * n * sizeof(double));
    memset(o->rhs_qx, 0, (size_t)o->nrk * n * sizeof(double));
    memset(o->rhs_qy, 0, (size_t)o->nrk * n * sizeof(double));
}

void oracle_run(oracle_t *o, int it0, int nsteps) {
    for (int it = it0; it < it0 + nsteps; it++) oracle_step(o, it);
}

/* modal2nodal (src/write_results.F:14-174) */
void oracle_modal2nodal(oracle_t *o, double *eta2, double *uu2, double *vv2) {
    int dofh = o->dofh;
#pragma omp parallel for num_threads(o->nthreads) schedule(static)
    for (int i = 0; i < o->np; i++) {
        double asum = 0;
        for (int j = o->eletab_off[i]; j < o->eletab_off[i + 1]; j++) {
            int e = o->eletab[j];
            if (o->wdflg[e] == 0) continue;
            asum = asum + 0.5 * o->areas[e];
        }
        double e2 = 0.0, u2 = 0.0, v2 = 0.0;
        for (int j = o->eletab_off[i]; j < o->eletab_off[i + 1]; j++) {
            int e = o->eletab[j];
            if (o->wdflg[e] == 0) continue;
            int k = 0;
            for (k = 0; k < 3; k++) if (o->nm[3 * e + k] == i) break;
            double zd = ZE(0, e, 0), xd = QX(0, e, 0), yd = QY(0, e, 0), hd = HB(e, 0);
            for (int kk = 1; kk < dofh; kk++) {
                double ph = o->phi_corner[k][kk];
                zd = zd + ph * ZE(0, e, kk); xd = xd + ph * QX(0, e, kk);
                yd = yd + ph * QY(0, e, kk); hd = hd + ph * HB(e, kk);
            }
            double area = 0.50 * o->areas[e] / asum;
            double depth = zd + hd;
            double fh = 1.0 / (o->nleq * depth + o->leq);
            e2 = e2 + area * zd;
            /* the reference compares after every wet element's contribution (write_results.F:164),
               so ETAMAX also sees the partial sums */
            if (o->etamax[i] < e2) o->etamax[i] = e2;
            u2 = u2 + area * xd * fh;
            v2 = v2 + area * yd * fh;
        }
        eta2[i] = e2; uu2[i] = u2; vv2[i] = v2;
    }
}

/* ------------------------------------------------------------------------------------------ */
/* accessors                                                                                   */
/* ------------------------------------------------------------------------------------------ */
void oracle_get_state(const oracle_t *o, double *ze, double *qx, double *qy, int *wdflg) {
    size_t n = (size_t)o->ne * o->dofh;
    if (ze) memcpy(ze, o->ze, n * sizeof(double));
    if (qx) memcpy(qx, o->qx, n * sizeof(double));
    if (qy) memcpy(qy, o->qy, n * sizeof(double));
    if (wdflg) memcpy(wdflg, o->wdflg, o->ne * sizeof(int));
}
void oracle_set_state(oracle_t *o, const double *ze, const double *qx, const double *qy, const int *wdflg) {
    size_t n = (size_t)o->ne * o->dofh;
    if (ze) memcpy(o->ze, ze, n * sizeof(double));
    if (qx) memcpy(o->qx, qx, n * sizeof(double));
    if (qy) memcpy(o->qy, qy, n * sizeof(double));
    if (wdflg) memcpy(o->wdflg, wdflg, o->ne * sizeof(int));
}
void oracle_get_rhs(const oracle_t *o, int irk, double *rze, double *rqx, double *rqy) {
    size_t n = (size_t)o->ne * o->dofh;
    memcpy(rze, o->rhs_ze + (size_t)irk * n, n * sizeof(double));
    memcpy(rqx, o->rhs_qx + (size_t)irk * n, n * sizeof(double));
    memcpy(rqy, o->rhs_qy + (size_t)irk * n, n * sizeof(double));
}
void oracle_clear_rhs(oracle_t *o) {
    size_t n = (size_t)o->nrk * o->ne * o->dofh;
    memset(o->rhs_ze, 0, n * sizeof(double)); memset(o->rhs_qx, 0, n * sizeof(double));
    memset(o->rhs_qy, 0, n * sizeof(double));
}
void oracle_get_nodecode(const oracle_t *o, int *nc) { memcpy(nc, o->nodecode, o->np * sizeof(int)); }
void oracle_get_etamax(const oracle_t *o, double *etamax) { memcpy(etamax, o->etamax, o->np * sizeof(double)); }
void oracle_set_etamax(oracle_t *o, const double *etamax) { memcpy(o->etamax, etamax, o->np * sizeof(double)); }
int oracle_num_edges(const oracle_t *o, int *counts) {
    counts[0] = o->nedges; counts[1] = o->nieds; counts[2] = o->nleds; counts[3] = o->needs; counts[4] = o->nfeds;
    return o->nedges;
}
/* Edge arrays, 1-based like the reference: nedno/nedel/nedsd are [ned][2]; nedel 0 if none. */
void oracle_get_edges(const oracle_t *o, int *nedno, int *nedel, int *nedsd, int *niedn, int *nledn,
                      int *needn, int *nfedn, double *xlen, double *cosnx, double *sinnx) {
    for (int e = 0; e < o->nedges; e++)
        for (int k = 0; k < 2; k++) {
            nedno[2 * e + k] = o->nedno[2 * e + k] + 1;
            nedel[2 * e + k] = o->nedel[2 * e + k] + 1;
            nedsd[2 * e + k] = o->nedsd[2 * e + k] + 1;
        }
    for (int i = 0; i < o->nieds; i++) niedn[i] = o->niedn[i] + 1;
    for (int i = 0; i < o->nleds; i++) nledn[i] = o->nledn[i] + 1;
    for (int i = 0; i < o->needs; i++) needn[i] = o->needn[i] + 1;
    for (int i = 0; i < o->nfeds; i++) nfedn[i] = o->nfedn[i] + 1;
    memcpy(xlen, o->xlen, o->nedges * sizeof(double));
    memcpy(cosnx, o->cosnx, o->nedges * sizeof(double));
    memcpy(sinnx, o->sinnx, o->nedges * sizeof(double));
}
/* Reference-element tables, for checks against closed forms. */
void oracle_get_tables(const oracle_t *o, double *phi_area, double *phi_edge, double *phi_corner,
                       double *m_inv, double *xegp, double *wegp, double *xagp, double *yagp,
                       double *wagp, double *atvd, double *btvd, double *dtvd) {
    int d = o->dofh;
    for (int q = 0; q < o->nagp; q++) for (int k = 0; k < d; k++) phi_area[q * d + k] = o->phi_area[q][k];
    for (int l = 0; l < 3; l++) for (int q = 0; q < o->negp; q++) for (int k = 0; k < d; k++)
        phi_edge[(l * o->negp + q) * d + k] = o->phi_edge[l][q][k];
    for (int v = 0; v < 3; v++) for (int k = 0; k < d; k++) phi_corner[v * d + k] = o->phi_corner[v][k];
    for (int k = 0; k < d; k++) m_inv[k] = o->m_inv[k];
    for (int q = 0; q < o->negp; q++) { xegp[q] = o->xegp[q]; wegp[q] = o->wegp[q]; }
    for (int q = 0; q < o->nagp; q++) { xagp[q] = o->xagp[q]; yagp[q] = o->yagp[q]; wagp[q] = o->wagp[q]; }
    for (int i = 0; i < o->nrk; i++) { dtvd[i] = o->dtvd[i]; for (int j = 0; j < o->nrk; j++) { atvd[i * o->nrk + j] = o->atvd[i][j]; btvd[i * o->nrk + j] = o->btvd[i][j]; } }
}
/* Nonzero once a met-forcing time fell outside the track (the reference STOPs there). */
int oracle_met_error(const oracle_t *o) { return o->met_error; }
int oracle_series_error(const oracle_t *o) { return o->series_error; }
void oracle_get_tip2(const oracle_t *o, double *tip2) { memcpy(tip2, o->tip2, o->np * sizeof(double)); }

/* The stage's node forcing WSX2, WSY2, PR2 (after the last oracle_rhs). */
void oracle_get_met(const oracle_t *o, double *wsx2, double *wsy2, double *pr2) {
    memcpy(wsx2, o->wsx2, o->np * sizeof(double));
    memcpy(wsy2, o->wsy2, o->np * sizeof(double));
    memcpy(pr2, o->pr2, o->np * sizeof(double));
}

double oracle_mass(const oracle_t *o) {
    /* sum_e A_e * (ze_1 + hb_1): total water volume (mode 1 is the cell mean) */
    double m = 0.0;
    for (int e = 0; e < o->ne; e++) m += 0.5 * o->areas[e] * (ZE(0, e, 0) + HB(e, 0));
    return m;
}

void oracle_free(oracle_t *o) {
    if (!o) return;
    void *ptrs[] = {o->hl_time, o->hl_lat, o->hl_lon, o->hl_cpress, o->hl_spd, o->hl_rrp, o->hl_rmw,
                    o->hl_tvx, o->hl_tvy, o->slam, o->sfea, o->wsx2, o->wsy2, o->pr2,
                    o->x, o->y, o->dp, o->sfac, o->corif, o->manning, o->nm, o->open_len, o->open_nodes,
                    o->land_type, o->land_len, o->land_nodes, o->amig, o->ff, o->face, o->per, o->emo, o->efa,
                    o->famig, o->fff, o->fface, o->fper, o->qnam, o->qnph, o->nedno, o->nedel, o->nedsd,
                    o->neled, o->niedn, o->nledn, o->needn, o->nfedn, o->elem_edge, o->xlen, o->cosnx,
                    o->sinnx, o->emo_dg, o->efa_dg, o->qnam_dg, o->qnph_dg, o->areas, o->drdx, o->dsdx,
                    o->drdy, o->dsdy, o->cori_el, o->fric_el, o->dpe_min, o->hb, o->bath, o->dbathdx,
                    o->dbathdy, o->sfac_elem, o->xfac, o->yfac, o->srfac, o->bathed, o->sfaced, o->edgeq,
                    o->dp_node, o->el_count, o->eletab_off, o->eletab, o->ze, o->qx, o->qy, o->rhs_ze,
                    o->rhs_qx, o->rhs_qy, o->ze_edge, o->qx_edge, o->qy_edge, o->wdflg, o->nodecode,
                    o->zemin, o->zemax, o->qxmin, o->qxmax, o->qymin, o->qymax, o->etamax,
                    o->tpk, o->amigt, o->etrf, o->fft, o->facet, o->pert, o->saltamp, o->saltpha,
                    o->tip2, o->tslam, o->tsfea, o->nbd, o->esbin, o->esbin1, o->esbin2, o->eta2bc,
                    o->lbcodei, o->qnin, o->qnin1, o->qnin2, o->lim_node, o->lim_elem, o->wd_rec, o->ie_rec};
    for (size_t k = 0; k < sizeof ptrs / sizeof ptrs[0]; k++) free(ptrs[k]);
    free(o);
}
