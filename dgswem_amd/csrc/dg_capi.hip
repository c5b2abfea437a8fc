// C ABI of libdgswem_hip.so (declared in include/dgswem_gpu.h): device contexts, the per-step
// launch sequence, ensembles of uneven meshes batched into one launch per stage, HIP-graph replay
// and the block-size autotuner.
//
// A context is a batch of members (one member for a single run).  Members are concatenated
// element-major: element e of member m is global element eoff[m] + e; neighbour, node and
// boundary-list indices are offset the same way, and a per-element scenario id selects the
// member's time step, H0, CF and forcing rows.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/dgswem_gpu.h"
#include "dg_device.hpp"
#include "dg_setup.hpp"

// the kernels are compiled in this translation unit
#include "dg_kernels.hip"
#include "dg_adjoint.hip"

using namespace dg;

// Working memory of dg_gpu_mass_noalloc (see scratch_ptr below).
constexpr int kScratchDoubles = 65536;                   // 512 KiB
static __device__ __align__(256) double scratch[kScratchDoubles];

namespace {

thread_local std::string g_err;

int fail(const std::string& m) { g_err = m; return 1; }

#define HIPCHK(x)                                                                                  \
    do {                                                                                           \
        hipError_t e_ = (x);                                                                       \
        if (e_ != hipSuccess) return fail(std::string(#x) + ": " + hipGetErrorString(e_));          \
    } while (0)

enum { K_STAGE = 0, K_MINMAX = 1, K_LIMIT = 2, K_M2N = 3, K_MET = 4, K_COUNT = 5 };
const char* kKernelNames[K_COUNT] = {"stage_kernel", "node_minmax_kernel", "limit_wetdry_kernel",
                                     "modal2nodal_kernel", "met_kernel"};
constexpr int kTableChunk = 2048;   // steps of forcing rows per table upload

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0, cap = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { if (p) (void)hipFree(p); }
    hipError_t alloc(size_t count) {
        n = count;
        if (p && count <= cap) return hipSuccess;
        if (p) { (void)hipFree(p); p = nullptr; }
        cap = std::max<size_t>(count, 1);
        return hipMalloc(&p, cap * sizeof(T));
    }
    hipError_t upload(const T* h, size_t count, hipStream_t s) {
        hipError_t e = alloc(count);
        if (e != hipSuccess || count == 0) return e;
        return hipMemcpyAsync(p, h, count * sizeof(T), hipMemcpyHostToDevice, s);
    }
};

// Host-side data of one scenario (one mesh and its forcing), in local indices.
struct Member {
    dg_params prm{};
    int np = 0, ne = 0;
    std::vector<double> xv, yv, dpv, nbar, cori, sfacq, sfaced;   // element SoA, [k][ne]
    double cori_val = 0.0;                                        // uniform Coriolis
    std::vector<int> sinfo, snbr, nm, n2e_off, n2e;
    int needs = 0, nfeds = 0;
    int nbfr = 0, nffr = 0;
    std::vector<double> amig, ff, face, per, famig, fff, fface, fper;
    std::vector<double> emo, efa, qnam, qnph;   // [edge][constituent][2]
    std::vector<double> ze0;                    // [dof][ne] cold-start elevation modes
    std::vector<uint8_t> wd0;                   // cold-start wet/dry flags
    // NWS=8 Holland vortex
    bool met = false;
    HollandTrack track;
    std::vector<double> mxc, myc, mcosy, mcor;  // [np] node lon/lat (deg), cos(lat), Coriolis
    // tidal potential (NTIP != 0): constituents, node data, SAL terms [ntif][np] (empty: zero)
    int ntif = 0;
    double reftim = 0.0;
    std::vector<double> tpk, amigt, etrf, fft, facet, pert, tslam, tccs, ts2s, tsamp, tspha;
    std::vector<int> tna;
    // non-periodic boundaries: fort.19 / fort.20 records, the ocean edges' end nodes (NEDNO order),
    // the NVEL boundary codes, and the replay state of the reference's record reading
    std::vector<int> onode;                     // [needs][2] 0-based
    bool eser = false, fser = false;
    int neta = 0, nvel = 0, nflux = 0, e_nrec = 0, f_nrec = 0;
    double etiminc = 0.0, ftiminc = 0.0;
    std::vector<int> nbd, lbcodei;
    std::vector<double> esbin, qnin;            // [nrec][neta], [nrec][nflux]
    struct Series {
        int it = 0;                             // steps whose stages have been applied
        double t1 = 0.0, t2 = 0.0;
        int next = 2;
        std::vector<double> v1, v2;
    } es, fs;
};

struct Ctx {
    std::vector<Member> mem;
    int S = 1, np = 0, ne = 0, dof = 0, nagp = 0, negp = 0;
    std::vector<int> eoff, noff, ooff, foff;     // member offsets: elements, nodes, ocean, flux edges
    int nbfr = 0, nffr = 0;                      // constituent counts, maxima over members
    dg_params prm{};                             // switches shared by all members
    RefTables T;
    RKScheme rk;
    Phys phys{};
    hipStream_t stream = nullptr;
    // state: NRK+1 rotating stage slots, kept RHS of earlier stages, double-buffered wet/dry flags
    DevBuf<double> u[KMAXRK + 1];
    int ucur = 0;                                // slot holding the stage-1 state
    DevBuf<double> r[KMAXRK];
    bool keep_r[KMAXRK] = {};
    DevBuf<uint8_t> wd[2];
    int wdcur = 0;
    // mesh
    DevBuf<uint16_t> sid;
    DevBuf<Scen> scen;
    DevBuf<double> xv, yv, dpv, nbar, cori_el, sfacq, sfaced, sfs, geo;
    DevBuf<int> sinfo, snbr, nm, n2e_off, n2e;
    std::vector<int> h_sinfo, h_snbr;            // host copies (staging lists)
    bool has_cori_el = false;
    // block-local staging of the stage kernel, per block size: each block copies the states of its
    // own elements and of its halo (out-of-block internal neighbours) into LDS
    struct Staging {
        int ns = 0;                              // staged elements per block: block size + largest halo
        size_t lds = 0;                          // dynamic LDS bytes per block
        DevBuf<int> hoff, hidx, snbl;            // halo offsets per tile, halo elements, packed local indices
    };
    std::map<int, Staging> stg;
    int max_lds = 65536;                         // LDS bytes per block the device allows
    DevBuf<double> nmm, outbuf, partial;
    // slope limiter: the stage kernel's copy of the updated cell means [3][ne] (the limiter's node
    // bounds read them there, never from the state it limits in place) and the per-tile node lists
    DevBuf<double> cmean;
    struct LimBufs {
        DevBuf<int> vhoff, vhidx, qoff, qinc_off, elq;
        DevBuf<uint16_t> qinc;
        int nsl = 0, nql = 0;
    } lim;
    DevBuf<int> nodecode;
    // modal2nodal: node rows in the host's ELETAB order (default: the node->element CSR) and the
    // running maximum ETAMAX, when the host has handed it over
    DevBuf<int> m2n_off, m2n;
    bool m2n_set = false;
    DevBuf<double> etamax;
    bool has_etamax = false;
    // forcing
    DevBuf<double> emo, efa, qnam, qnph, argj, rff, fargj, frff;
    DevBuf<int> step;                            // device step counter within the current tables
    DevBuf<double> tz;                           // tidal terms of the current step [stage][needs][nbfr][2]
    DevBuf<uint16_t> osid;                       // scenario of each ocean edge (ensembles)
    int tzn = 0;                                 // tidal terms per stage
    int hstep = 0;                               // the same index for eager (uncaptured) launches
    int rows = 0;                                // table rows per scenario
    // specified normal flux of the current step at every flux-edge point, and each flux edge's
    // scenario (flow_kernel)
    DevBuf<double> fz;
    DevBuf<uint16_t> fsid;
    int fzn = 0;
    // non-periodic boundaries (single runs): the ramped end-node values of every stage row
    bool eser = false, fser = false;
    DevBuf<double> ets, qts;
    // tidal potential (single runs): per-row constituent terms, node data and TIP2 of the step
    bool tip = false;
    int ntif = 0;
    DevBuf<double> tiprow, tslam, tccs, ts2s, tsamp, tspha, tip2;
    DevBuf<int> tna;
    // NWS=8 meteorological forcing
    bool met = false;
    DevBuf<double> ramp, mrow, mxc, myc, mcosy, mcor, mwx, mwy, mwm, mcd, mprn;
    DevBuf<uint16_t> nsid;
    int next_it = 1;                             // IT of the next step (for the autotuner's tables)
    // set while the adjoint's forward sweeps run: WETDRY raises it when it changes an element
    int* wd_active = nullptr;
    DevBuf<int> wdact;
    DevBuf<uint8_t> wdsave;                      // wet/dry flags at the start of a gradient call
    // launch configuration
    int block = 256;
    int dev = 0;                                 // device current at dg_gpu_init
    // adjoint: gauges, checkpoints, segment stage states, cotangents, parameter accumulators
    int ng = 0, ngu = 0;                         // gauges, distinct gauge elements
    DevBuf<int> gel, goff, gord;
    DevBuf<double> gphi, geta, gdata, ckpt, seg, lam, gdp, gnb, pout;
    DevBuf<double> ubar[KMAXRK];
    // frozen-limiter reverse: pre-limiter stage outputs of a segment, vertex-bound cotangents,
    // node sums and bound-attaining elements
    DevBuf<double> wseg, lnbc, lgsum;
    DevBuf<int> larg;
    // active wetting/drying under the limiter: the flags at each checkpoint and at the start of every
    // stage of a segment
    DevBuf<uint8_t> ckptwd, segwd;
    bool adj_wd = false;      // WETDRY after the limiter: flags kept at the checkpoints
    bool adj_wdrev = false;   // ... and WETDRY changed an element in the forward sweep: reversed
    // graphs of one slot-rotation period, keyed by (ucur, wdcur) at capture
    bool use_graph = false;
    std::map<int, hipGraphExec_t> graphs;
    // profiling
    bool prof = false;
    int prof_every = 1;                          // time the launches of every prof_every-th step
    long long prof_step = 0;
    bool prof_this = true;                       // the current step is sampled
    double ktime[K_COUNT] = {};
    long long kcount[K_COUNT] = {};
    std::vector<std::pair<hipEvent_t, hipEvent_t>> evpool;
    std::vector<std::pair<int, int>> pending;    // (kernel id, event pair index)
    void drop_graphs() {
        for (auto& g : graphs) (void)hipGraphExecDestroy(g.second);
        graphs.clear();
    }
    ~Ctx() {
        if (stream) (void)hipStreamSynchronize(stream);
        drop_graphs();
        for (auto& e : evpool) { (void)hipEventDestroy(e.first); (void)hipEventDestroy(e.second); }
        if (stream) (void)hipStreamDestroy(stream);
    }
};

std::mutex g_mu;
std::vector<std::unique_ptr<Ctx>> g_ctx;
bool g_tables_set[3] = {false, false, false};

Ctx* get(int h) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (h < 0 || h >= (int)g_ctx.size() || !g_ctx[h]) { g_err = "invalid handle"; return nullptr; }
    return g_ctx[h].get();
}

DevTables to_dev(const RefTables& t) {
    DevTables d{};
    std::memcpy(d.phi_area, t.phi_area, sizeof d.phi_area);
    std::memcpy(d.drphi, t.drphi, sizeof d.drphi);
    std::memcpy(d.dsphi, t.dsphi, sizeof d.dsphi);
    std::memcpy(d.mw, t.mw, sizeof d.mw);
    std::memcpy(d.srfac, t.srfac, sizeof d.srfac);
    std::memcpy(d.phi_edge, t.phi_edge, sizeof d.phi_edge);
    std::memcpy(d.edgeq, t.edgeq, sizeof d.edgeq);
    std::memcpy(d.phi_corner, t.phi_corner, sizeof d.phi_corner);
    std::memcpy(d.m_inv, t.m_inv, sizeof d.m_inv);
    std::memcpy(d.wegp, t.wegp, sizeof d.wegp);
    std::memcpy(d.xegp, t.xegp, sizeof d.xegp);
    static_assert(sizeof d.psi == sizeof t.psi, "psi table layout");
    std::memcpy(d.psi, t.psi, sizeof d.psi);
    return d;
}

// ---- profiling helpers (no events are recorded into a captured graph)
int prof_begin(Ctx& c, int kid, bool capturing) {
    if (!c.prof || capturing || !c.prof_this) return -1;
    const int idx = (int)c.pending.size();
    if (idx >= (int)c.evpool.size()) {
        hipEvent_t a, b;
        if (hipEventCreate(&a) != hipSuccess || hipEventCreate(&b) != hipSuccess) return -1;
        c.evpool.push_back({a, b});
    }
    (void)hipEventRecord(c.evpool[idx].first, c.stream);
    c.pending.push_back({kid, idx});
    return idx;
}
void prof_end(Ctx& c, int idx) {
    if (idx >= 0) (void)hipEventRecord(c.evpool[idx].second, c.stream);
}
int prof_flush(Ctx& c) {
    if (c.pending.empty()) return 0;
    HIPCHK(hipStreamSynchronize(c.stream));
    for (auto& pr : c.pending) {
        float ms = 0.f;
        HIPCHK(hipEventElapsedTime(&ms, c.evpool[pr.second].first, c.evpool[pr.second].second));
        c.ktime[pr.first] += ms;
        c.kcount[pr.first] += 1;
    }
    c.pending.clear();
    return 0;
}

ElemData elem_data(Ctx& c) {
    ElemData E{};
    E.ne = c.ne;
    E.sid = c.S > 1 ? c.sid.p : nullptr;
    E.scen = c.scen.p;
    E.xv = c.xv.p; E.yv = c.yv.p; E.dpv = c.dpv.p;
    E.nbar = c.phys.load_manning ? c.nbar.p : nullptr;
    E.cori_el = c.has_cori_el ? c.cori_el.p : nullptr;
    E.sfacq = c.prm.ics == 2 ? c.sfacq.p : nullptr;
    E.sfaced = c.prm.ics == 2 ? c.sfaced.p : nullptr;
    E.sfs = c.prm.ics == 2 ? c.sfs.p : nullptr;
    E.geo = c.geo.p;
    E.sinfo = c.sinfo.p; E.snbr = c.snbr.p; E.nm = c.nm.p;
    return E;
}

// Forcing-table arguments.  Captured launches read the step index from the device counter (graph
// replays advance it); eager launches take it as an argument.
Forcing forcing(Ctx& c, bool captured = false) {
    Forcing F{};
    F.nbfr = c.nbfr; F.nffr = c.nffr; F.rows = c.rows; F.nrk = c.rk.nrk;
    F.step = captured ? c.step.p : nullptr;
    F.step_imm = c.hstep;
    F.emo = c.emo.p; F.efa = c.efa.p; F.qnam = c.qnam.p; F.qnph = c.qnph.p;
    F.argj = c.argj.p; F.rff = c.rff.p; F.fargj = c.fargj.p; F.frff = c.frff.p;
    F.ramp = c.ramp.p;
    F.tz = c.tz.p; F.tzn = c.tzn;
    F.fz = c.fz.p; F.fzn = c.fzn;
    F.ets = c.eser ? c.ets.p : nullptr;
    F.qts = c.fser ? c.qts.p : nullptr;
    F.nfeds = c.foff[c.S];
    return F;
}

Tip tip_args(Ctx& c) {
    Tip t{};
    t.np = c.np; t.ntif = c.ntif;
    t.row = c.tiprow.p; t.na = c.tna.p;
    t.slam = c.tslam.p; t.ccsfea = c.tccs.p; t.s2sfea = c.ts2s.p;
    t.salt_amp = c.tsamp.n ? c.tsamp.p : nullptr;
    t.salt_pha = c.tspha.n ? c.tspha.p : nullptr;
    t.tip2 = c.tip2.p;
    return t;
}

Met met_args(Ctx& c) {
    Met m{};
    m.np = c.np; m.rows = kTableChunk;
    m.row = c.mrow.p;
    m.nsid = c.S > 1 ? c.nsid.p : nullptr;
    m.xc = c.mxc.p; m.yc = c.myc.p; m.cosy = c.mcosy.p; m.cor = c.mcor.p;
    m.wx = c.mwx.p; m.wy = c.mwy.p; m.wm = c.mwm.p; m.cd = c.mcd.p; m.prn = c.mprn.p;
    return m;
}

// --------------------------------------------------------------------------------------------
// Member construction on the host: side slots in the reference's accumulation order, element
// SoA arrays, the node->element CSR and the cold start.
// --------------------------------------------------------------------------------------------
int build_member(Member& m, const dg_params* prm, const RefTables& T, int np, int ne, const double* x,
                 const double* y, const double* dp, const double* corif, const double* manning,
                 const int* nm, int nedges, const int* nedel, const int* nedsd, int nieds,
                 const int* niedn, int nleds, const int* nledn, int needs, const int* needn,
                 int nfeds, const int* nfedn) {
    m.prm = *prm;
    m.np = np; m.ne = ne;
    const int nagp = T.nagp, negp = T.negp, dof = T.dof;
    std::vector<int> kind_of(nedges, -1), list_pos(nedges, 0);
    auto mark = [&](int n, const int* lst, int kind) -> int {
        for (int i = 0; i < n; i++) {
            const int ed = lst[i] - 1;
            if (ed < 0 || ed >= nedges) return fail("edge list refers to an edge out of range");
            kind_of[ed] = kind; list_pos[ed] = i;
        }
        return 0;
    };
    if (mark(nleds, nledn, SIDE_LAND) || mark(nfeds, nfedn, SIDE_FLOW) || mark(needs, needn, SIDE_OCEAN) ||
        mark(nieds, niedn, SIDE_INTERNAL))
        return 1;
    m.sinfo.assign(ne, 0);
    m.snbr.assign((size_t)3 * ne, -1);
    {
        // per element: its three edges with (kind order, position) sort keys
        std::vector<std::array<long long, 3>> keys(ne);
        std::vector<std::array<int, 3>> bitsv(ne), nbv(ne);
        std::vector<int> cnt(ne, 0);
        for (int ed = 0; ed < nedges; ed++) {
            for (int s = 0; s < 2; s++) {
                const int el = nedel[2 * ed + s] - 1;
                if (el < 0) continue;
                const int L = nedsd[2 * ed + s] - 1;
                if (el >= ne || L < 0 || L > 2) return fail("corrupt edge tables");
                const int k = kind_of[ed];
                if (k < 0) return fail("edge " + std::to_string(ed + 1) + " is in no edge list");
                int bits = L | (s << 2) | (k << 3);
                int nb = -1;
                if (k == SIDE_INTERNAL) {
                    const int other = nedel[2 * ed + (1 - s)] - 1;
                    const int oled = nedsd[2 * ed + (1 - s)] - 1;
                    if (other < 0 || other >= ne || oled < 0 || oled > 2)
                        return fail("internal edge without a valid second element");
                    bits |= oled << 5;
                    nb = other;
                } else {
                    if (s != 0) return fail("boundary edge on the second side");
                    nb = list_pos[ed];
                }
                // boundary sides first in the order ocean, land, flux and their list order;
                // internal sides after, by global edge number
                const int korder = (k == SIDE_OCEAN) ? 0 : (k == SIDE_LAND) ? 1 : (k == SIDE_FLOW) ? 2 : 3;
                const long long key = (long long)korder * (1LL << 40) + (k == SIDE_INTERNAL ? ed : list_pos[ed]);
                const int j = cnt[el]++;
                if (j >= 3) return fail("element with more than three edges");
                keys[el][j] = key; bitsv[el][j] = bits; nbv[el][j] = nb;
            }
        }
        for (int e = 0; e < ne; e++) {
            if (cnt[e] != 3) return fail("element " + std::to_string(e + 1) + " does not have three edges");
            int ord[3] = {0, 1, 2};
            std::sort(ord, ord + 3, [&](int a, int b) { return keys[e][a] < keys[e][b]; });
            int info = 0;
            for (int j = 0; j < 3; j++) {
                info |= bitsv[e][ord[j]] << (7 * j);
                m.snbr[(size_t)j * ne + e] = nbv[e][ord[j]];
            }
            m.sinfo[e] = info;
        }
    }
    m.needs = needs; m.nfeds = nfeds;
    // ---- element SoA arrays (src/prep_DG.F:982-983 nodal means)
    m.xv.resize((size_t)3 * ne); m.yv.resize((size_t)3 * ne); m.dpv.resize((size_t)3 * ne);
    m.nm.resize((size_t)3 * ne);
    const bool man = prm->load_manning && manning;
    if (man) m.nbar.resize(ne);
    bool cori_uniform = true;
    for (int i = 1; i < np; i++) if (corif[i] != corif[0]) { cori_uniform = false; break; }
    if (!cori_uniform) m.cori.resize(ne);
    for (int e = 0; e < ne; e++) {
        int n[3];
        for (int v = 0; v < 3; v++) {
            n[v] = nm[(size_t)v * ne + e] - 1;
            if (n[v] < 0 || n[v] >= np) return fail("NM refers to a node out of range");
            m.xv[(size_t)v * ne + e] = x[n[v]];
            m.yv[(size_t)v * ne + e] = y[n[v]];
            m.dpv[(size_t)v * ne + e] = dp[n[v]];
            m.nm[(size_t)v * ne + e] = n[v];
        }
        const double AREAS = (x[n[0]] - x[n[2]]) * (y[n[1]] - y[n[2]]) + (x[n[2]] - x[n[1]]) * (y[n[0]] - y[n[2]]);
        if (!(AREAS > 0.0)) return fail("element " + std::to_string(e + 1) + " is not counter-clockwise");
        if (man) m.nbar[e] = (manning[n[0]] + manning[n[1]] + manning[n[2]]) / 3.0;
        if (!cori_uniform) m.cori[e] = (corif[n[0]] + corif[n[1]] + corif[n[2]]) / 3.0;
    }
    m.cori_val = (corif[0] + corif[0] + corif[0]) / 3.0;
    // spherical factors at the quadrature points (src/prep_DG.F:1010-1016,1085-1091)
    if (prm->ics == 2) {
        const double R = 6378206.4;
        m.sfacq.resize((size_t)nagp * ne);
        m.sfaced.resize((size_t)3 * negp * ne);
        for (int e = 0; e < ne; e++) {
            const double Y1 = m.yv[e], Y2 = m.yv[(size_t)ne + e], Y3 = m.yv[2 * (size_t)ne + e];
            const double yd[3] = {1.0 / 3.0 * (Y1 + Y2 + Y3), -(1.0 / 6.0 * (Y1 + Y2)) + 1.0 / 3.0 * Y3,
                                  -(0.5 * Y1) + 0.5 * Y2};
            for (int I = 0; I < nagp; I++) {
                double ye = 0.0;
                for (int k = 0; k < 3; k++) ye = ye + yd[k] * T.phi_area[I][k];
                m.sfacq[(size_t)I * ne + e] = std::cos(prm->sfea0) / std::cos(ye / R);
            }
            for (int L = 0; L < 3; L++)
                for (int I = 0; I < negp; I++) {
                    double ye = 0.0;
                    for (int k = 0; k < 3; k++) ye = ye + yd[k] * T.phi_edge[L][I][k];
                    m.sfaced[((size_t)L * negp + I) * ne + e] = std::cos(prm->sfea0) / std::cos(ye / R);
                }
        }
    }
    // ---- node -> element CSR (entries: element*4 + local vertex) in the reference's ELETAB order,
    // the order modal2nodal sums in; the limiter's node min/max does not depend on it
    eletab_order(np, ne, m.nm.data(), x, y, m.n2e_off, m.n2e);
    // ---- cold start: ZE=QX=QY=0 except the initial dry check (src/prep_DG.F:1214-1300)
    m.ze0.assign((size_t)dof * ne, 0.0);
    m.wd0.assign(ne, 1);
    if (prm->nolifa >= 2) {
        const double H0 = prm->h0;
        for (int e = 0; e < ne; e++) {
            const double d[3] = {m.dpv[e], m.dpv[(size_t)ne + e], m.dpv[2 * (size_t)ne + e]};
            double z[3] = {0.0, 0.0, 0.0};
            for (int v = 0; v < 3; v++) if (d[v] < H0) z[v] = std::max(z[v], H0 - d[v]);
            if ((z[0] + z[1] + z[2]) / 3.0 != 0.0) {
                if (!(0.0 > (z[0] + z[1] + z[2]) / 3.0)) {
                    m.ze0[(size_t)0 * ne + e] = (z[0] + z[1] + z[2]) / 3.0;
                    m.ze0[(size_t)1 * ne + e] = -(1.0 / 6.0 * (z[0] + z[1])) + 1.0 / 3.0 * z[2];
                    m.ze0[(size_t)2 * ne + e] = -(0.5 * z[0]) + 0.5 * z[1];
                }
                m.wd0[e] = 0;
            }
        }
    }
    return 0;
}

// Per-scenario scalars (src/dgswem.F:131-170, src/rhs_dg_hydro.F:46, src/prep_DG.F:982-983).
Scen scen_of(const Member& m) {
    Scen s{};
    s.dtdp = m.prm.dtdp;
    s.dtdph = 1.0 / m.prm.dtdp;
    s.h0 = m.prm.h0;
    s.cf = m.prm.cf;
    s.fric = (m.prm.cf + m.prm.cf + m.prm.cf) / 3.0;
    s.cori = m.cori_val;
    s.rain = m.prm.rainfall == 1 ? 7.0556e-6 : 0.0;
    return s;
}

// Boundary forcing coefficients of all members, padded to the largest constituent counts, and
// the per-stage tables, allocated once per context so that pointers captured in a graph stay valid.
int upload_forcing(Ctx& c) {
    c.nbfr = 0; c.nffr = 0;
    for (auto& m : c.mem) { c.nbfr = std::max(c.nbfr, m.nbfr); c.nffr = std::max(c.nffr, m.nffr); }
    const int tot_o = c.ooff[c.S], tot_f = c.foff[c.S];
    std::vector<double> em((size_t)tot_o * std::max(c.nbfr, 1) * 2, 0.0), ef(em.size(), 0.0);
    std::vector<double> qa((size_t)tot_f * std::max(c.nffr, 1) * 2, 0.0), qp(qa.size(), 0.0);
    for (int s = 0; s < c.S; s++) {
        const Member& m = c.mem[s];
        for (int L = 0; L < m.needs; L++)
            for (int j = 0; j < m.nbfr; j++)
                for (int k = 0; k < 2; k++) {
                    const size_t d = ((size_t)(c.ooff[s] + L) * c.nbfr + j) * 2 + k, o = ((size_t)L * m.nbfr + j) * 2 + k;
                    em[d] = m.emo[o];
                    ef[d] = m.efa[o];
                }
        for (int L = 0; L < m.nfeds; L++)
            for (int j = 0; j < m.nffr; j++)
                for (int k = 0; k < 2; k++) {
                    const size_t d = ((size_t)(c.foff[s] + L) * c.nffr + j) * 2 + k, o = ((size_t)L * m.nffr + j) * 2 + k;
                    qa[d] = m.qnam[o];
                    qp[d] = m.qnph[o];
                }
    }
    c.drop_graphs();
    HIPCHK(hipStreamSynchronize(c.stream));
    HIPCHK(c.emo.upload(em.data(), em.size(), c.stream));
    HIPCHK(c.efa.upload(ef.data(), ef.size(), c.stream));
    HIPCHK(c.qnam.upload(qa.data(), qa.size(), c.stream));
    HIPCHK(c.qnph.upload(qp.data(), qp.size(), c.stream));
    c.tzn = tot_o * c.nbfr * 2;
    if (c.tzn > 0) {
        std::vector<uint16_t> os(tot_o);
        for (int s = 0; s < c.S; s++)
            for (int L = c.ooff[s]; L < c.ooff[s + 1]; L++) os[L] = (uint16_t)s;
        HIPCHK(c.osid.upload(os.data(), os.size(), c.stream));
        HIPCHK(c.tz.alloc((size_t)c.rk.nrk * c.tzn));
    }
    c.fzn = tot_f * c.negp;
    if (c.fzn > 0) {
        std::vector<uint16_t> fsd(tot_f);
        for (int s = 0; s < c.S; s++)
            for (int L = c.foff[s]; L < c.foff[s + 1]; L++) fsd[L] = (uint16_t)s;
        HIPCHK(c.fsid.upload(fsd.data(), fsd.size(), c.stream));
        HIPCHK(c.fz.alloc((size_t)c.rk.nrk * c.fzn));
    }
    const size_t rows = (size_t)kTableChunk * c.rk.nrk;
    c.eser = c.fser = false;
    for (auto& m : c.mem) { c.eser = c.eser || m.eser; c.fser = c.fser || m.fser; }
    if (c.eser) HIPCHK(c.ets.alloc((size_t)c.S * rows * std::max(c.tzn, 1)));
    if (c.fser) HIPCHK(c.qts.alloc((size_t)c.S * rows * tot_f * 2));
    HIPCHK(c.argj.alloc((size_t)c.S * rows * std::max(c.nbfr, 1)));
    HIPCHK(c.rff.alloc((size_t)c.S * rows * std::max(c.nbfr, 1)));
    HIPCHK(c.fargj.alloc((size_t)c.S * rows * std::max(c.nffr, 1)));
    HIPCHK(c.frff.alloc((size_t)c.S * rows * std::max(c.nffr, 1)));
    HIPCHK(c.ramp.alloc((size_t)c.S * rows));
    c.rows = (int)rows;
    HIPCHK(hipStreamSynchronize(c.stream));
    return 0;
}

// Node arrays of the Holland forcing of all members (all members have it, or none).
int upload_met(Ctx& c) {
    c.met = c.mem[0].met;
    for (auto& m : c.mem)
        if (m.met != c.met) return fail("ensemble members must all have NWS=8 forcing, or none");
    if (!c.met) return 0;
    const int np = c.np;
    std::vector<double> xc(np), yc(np), cy(np), co(np);
    std::vector<uint16_t> ns(np);
    for (int s = 0; s < c.S; s++) {
        const Member& m = c.mem[s];
        for (int i = 0; i < m.np; i++) {
            const int g = c.noff[s] + i;
            xc[g] = m.mxc[i]; yc[g] = m.myc[i]; cy[g] = m.mcosy[i]; co[g] = m.mcor[i];
            ns[g] = (uint16_t)s;
        }
    }
    c.drop_graphs();
    HIPCHK(hipStreamSynchronize(c.stream));
    HIPCHK(c.mxc.upload(xc.data(), np, c.stream));
    HIPCHK(c.myc.upload(yc.data(), np, c.stream));
    HIPCHK(c.mcosy.upload(cy.data(), np, c.stream));
    HIPCHK(c.mcor.upload(co.data(), np, c.stream));
    if (c.S > 1) HIPCHK(c.nsid.upload(ns.data(), np, c.stream));
    for (DevBuf<double>* b : {&c.mwx, &c.mwy, &c.mwm, &c.mcd, &c.mprn}) HIPCHK(b->alloc(np));
    HIPCHK(c.mrow.alloc((size_t)c.S * kTableChunk * MET_ROW));
    HIPCHK(hipStreamSynchronize(c.stream));
    return 0;
}

// The slope limiter's tiles of LIM_NT elements (LimTiles): each tile's vertex halo (ascending),
// its nodes in order of first appearance, every node's incident elements as staged rows in the
// node->element order `n2e` (entries elem<<2 | vertex), and each element's tile-local nodes.
int build_limtiles(Ctx& c, const std::vector<int>& nm, const std::vector<int>& off, const std::vector<int>& n2e) {
    const int ne = c.ne, nt = LIM_NT, ntiles = (ne + nt - 1) / nt;
    std::vector<int> vhoff(ntiles + 1, 0), vhidx, qoff(ntiles + 1, 0), qinc_off(1, 0), elq(ne, 0), halo, nodes;
    std::vector<uint16_t> qinc;
    std::vector<int> qlocal(c.np, -1);
    int nsl = nt, nql = 1;
    for (int t = 0; t < ntiles; t++) {
        const int b = t * nt, en = std::min(ne, b + nt);
        nodes.clear();
        halo.clear();
        for (int e = b; e < en; e++) {
            int packed = 0;
            for (int v = 0; v < 3; v++) {
                const int n = nm[(size_t)v * ne + e];
                if (qlocal[n] < 0) { qlocal[n] = (int)nodes.size(); nodes.push_back(n); }
                packed |= qlocal[n] << (10 * v);
            }
            elq[e] = packed;
        }
        for (int n : nodes)
            for (int j = off[n]; j < off[n + 1]; j++) {
                const int g = n2e[j] >> 2;
                if (g < b || g >= en) halo.push_back(g);
            }
        std::sort(halo.begin(), halo.end());
        halo.erase(std::unique(halo.begin(), halo.end()), halo.end());
        if (nt + (int)halo.size() > 65535 || (int)nodes.size() > 1023)
            return fail("slope limiter tiles: a tile touches too many nodes or elements");
        for (int n : nodes) {
            for (int j = off[n]; j < off[n + 1]; j++) {
                const int g = n2e[j] >> 2;
                const int r = (g >= b && g < en) ? g - b
                                                 : nt + (int)(std::lower_bound(halo.begin(), halo.end(), g) - halo.begin());
                qinc.push_back((uint16_t)r);
            }
            qinc_off.push_back((int)qinc.size());
            qlocal[n] = -1;
        }
        nsl = std::max(nsl, nt + (int)halo.size());
        nql = std::max(nql, (int)nodes.size());
        vhidx.insert(vhidx.end(), halo.begin(), halo.end());
        vhoff[t + 1] = (int)vhidx.size();
        qoff[t + 1] = qoff[t] + (int)nodes.size();
    }
    if ((size_t)8 * (3 * nsl + 6 * nql) > (size_t)c.max_lds)
        return fail("slope limiter tiles: a tile needs more LDS than a block may use");
    if (vhidx.empty()) vhidx.push_back(0);
    if (qinc.empty()) qinc.push_back(0);
    Ctx::LimBufs& L = c.lim;
    L.nsl = nsl;
    L.nql = nql;
    HIPCHK(L.vhoff.upload(vhoff.data(), vhoff.size(), c.stream));
    HIPCHK(L.vhidx.upload(vhidx.data(), vhidx.size(), c.stream));
    HIPCHK(L.qoff.upload(qoff.data(), qoff.size(), c.stream));
    HIPCHK(L.qinc_off.upload(qinc_off.data(), qinc_off.size(), c.stream));
    HIPCHK(L.qinc.upload(qinc.data(), qinc.size(), c.stream));
    HIPCHK(L.elq.upload(elq.data(), elq.size(), c.stream));
    HIPCHK(c.cmean.alloc((size_t)3 * ne));
    HIPCHK(hipStreamSynchronize(c.stream));
    return 0;
}

// Concatenate the members and upload everything except the state.
int upload_mesh(Ctx& c) {
    const int S = c.S;
    c.eoff.assign(S + 1, 0); c.noff.assign(S + 1, 0); c.ooff.assign(S + 1, 0); c.foff.assign(S + 1, 0);
    for (int s = 0; s < S; s++) {
        c.eoff[s + 1] = c.eoff[s] + c.mem[s].ne;
        c.noff[s + 1] = c.noff[s] + c.mem[s].np;
        c.ooff[s + 1] = c.ooff[s] + c.mem[s].needs;
        c.foff[s + 1] = c.foff[s] + c.mem[s].nfeds;
    }
    if ((long long)c.eoff[S] * 4 >= (1LL << 31) || (long long)c.noff[S] >= (1LL << 31))
        return fail("batch too large for 32-bit element indices");
    const int ne = c.ne = c.eoff[S], np = c.np = c.noff[S];
    const int nagp = c.nagp, negp = c.negp;
    std::vector<double> xv((size_t)3 * ne), yv((size_t)3 * ne), dpv((size_t)3 * ne);
    std::vector<int> sinfo(ne), snbr((size_t)3 * ne), nm((size_t)3 * ne), off(np + 1, 0), n2e((size_t)3 * ne);
    std::vector<uint16_t> sid(ne);
    std::vector<Scen> scen(S);
    std::vector<double> nbar, cori, sfq, sfe;
    c.has_cori_el = false;
    for (auto& m : c.mem) if (!m.cori.empty()) c.has_cori_el = true;
    if (c.phys.load_manning) nbar.resize(ne);
    if (c.has_cori_el) cori.resize(ne);
    if (c.prm.ics == 2) { sfq.resize((size_t)nagp * ne); sfe.resize((size_t)3 * negp * ne); }
    for (int s = 0; s < S; s++) {
        const Member& m = c.mem[s];
        const int e0 = c.eoff[s], n0 = c.noff[s], me = m.ne;
        scen[s] = scen_of(m);
        for (int e = 0; e < me; e++) {
            const int g = e0 + e;
            sid[g] = (uint16_t)s;
            sinfo[g] = m.sinfo[e];
            for (int v = 0; v < 3; v++) {
                xv[(size_t)v * ne + g] = m.xv[(size_t)v * me + e];
                yv[(size_t)v * ne + g] = m.yv[(size_t)v * me + e];
                dpv[(size_t)v * ne + g] = m.dpv[(size_t)v * me + e];
                nm[(size_t)v * ne + g] = m.nm[(size_t)v * me + e] + n0;
                const int kind = (side_bits(m.sinfo[e], v) >> 3) & 3;
                const int nb = m.snbr[(size_t)v * me + e];
                const int shift = kind == SIDE_INTERNAL ? e0 : kind == SIDE_OCEAN ? c.ooff[s] : kind == SIDE_FLOW ? c.foff[s] : 0;
                snbr[(size_t)v * ne + g] = nb + shift;
            }
            if (c.phys.load_manning) nbar[g] = m.nbar[e];
            if (c.has_cori_el) cori[g] = m.cori.empty() ? m.cori_val : m.cori[e];
            if (c.prm.ics == 2) {
                for (int I = 0; I < nagp; I++) sfq[(size_t)I * ne + g] = m.sfacq[(size_t)I * me + e];
                for (int j = 0; j < 3 * negp; j++) sfe[(size_t)j * ne + g] = m.sfaced[(size_t)j * me + e];
            }
        }
        for (int i = 0; i < m.np; i++) off[n0 + i + 1] = m.n2e_off[i + 1] - m.n2e_off[i];
        for (size_t j = 0; j < m.n2e.size(); j++) {
            const int ent = m.n2e[j];
            n2e[(size_t)3 * e0 + j] = (((ent >> 2) + e0) << 2) | (ent & 3);
        }
    }
    for (int i = 0; i < np; i++) off[i + 1] += off[i];
    HIPCHK(c.xv.upload(xv.data(), xv.size(), c.stream));
    HIPCHK(c.yv.upload(yv.data(), yv.size(), c.stream));
    HIPCHK(c.dpv.upload(dpv.data(), dpv.size(), c.stream));
    HIPCHK(c.nm.upload(nm.data(), nm.size(), c.stream));
    HIPCHK(c.sinfo.upload(sinfo.data(), sinfo.size(), c.stream));
    HIPCHK(c.snbr.upload(snbr.data(), snbr.size(), c.stream));
    HIPCHK(c.geo.alloc((size_t)13 * ne));
    hipLaunchKernelGGL(geom_kernel, dim3((ne + 255) / 256), dim3(256), 0, c.stream, ne, (const double*)c.xv.p,
                       (const double*)c.yv.p, (const int*)c.sinfo.p, c.geo.p);
    HIPCHK(hipGetLastError());
    c.stg.clear();
    HIPCHK(c.n2e_off.upload(off.data(), off.size(), c.stream));
    HIPCHK(c.n2e.upload(n2e.data(), n2e.size(), c.stream));
    HIPCHK(c.scen.upload(scen.data(), scen.size(), c.stream));
    if (S > 1) HIPCHK(c.sid.upload(sid.data(), sid.size(), c.stream));
    if (c.phys.load_manning) HIPCHK(c.nbar.upload(nbar.data(), nbar.size(), c.stream));
    if (c.has_cori_el) HIPCHK(c.cori_el.upload(cori.data(), cori.size(), c.stream));
    if (c.prm.ics == 2) {
        HIPCHK(c.sfacq.upload(sfq.data(), sfq.size(), c.stream));
        HIPCHK(c.sfaced.upload(sfe.data(), sfe.size(), c.stream));
        // slot order, so that each lane reads its own element's column (coalesced) instead of a
        // row chosen by its local edge number or a neighbour's column
        std::vector<double> sfs((size_t)2 * 3 * negp * ne, 1.0);
        for (int e = 0; e < ne; e++)
            for (int sl = 0; sl < 3; sl++) {
                const int bits = side_bits(sinfo[e], sl);
                const int L = bits & 3, role = (bits >> 2) & 1, kind = (bits >> 3) & 3, nled = (bits >> 5) & 3;
                for (int I = 0; I < negp; I++) {
                    if (kind == SIDE_INTERNAL) {
                        const int gown = role ? negp - 1 - I : I, gnb = role ? I : negp - 1 - I;
                        const int nb = snbr[(size_t)sl * ne + e];
                        sfs[((size_t)sl * negp + I) * ne + e] = sfe[((size_t)L * negp + gown) * ne + e];
                        sfs[((size_t)(3 + sl) * negp + I) * ne + e] = sfe[((size_t)nled * negp + gnb) * ne + nb];
                    } else {
                        sfs[((size_t)sl * negp + I) * ne + e] = sfe[((size_t)L * negp + I) * ne + e];
                    }
                }
            }
        HIPCHK(c.sfs.upload(sfs.data(), sfs.size(), c.stream));
        HIPCHK(hipStreamSynchronize(c.stream));
    }
    HIPCHK(c.nmm.alloc((size_t)6 * np));
    if (c.prm.slopeflag != 0 && build_limtiles(c, nm, off, n2e)) return 1;
    HIPCHK(c.outbuf.alloc((size_t)3 * np));
    HIPCHK(c.nodecode.alloc(np));
    HIPCHK(c.partial.alloc((ne + 255) / 256));
    HIPCHK(c.step.alloc(1));
    const int nrk = c.rk.nrk, dof = c.dof;
    for (int i = 0; i <= nrk; i++) {
        HIPCHK(c.u[i].alloc((size_t)3 * dof * ne));
        HIPCHK(hipMemsetAsync(c.u[i].p, 0, (size_t)3 * dof * ne * sizeof(double), c.stream));
    }
    for (int i = 0; i < nrk; i++) {
        c.keep_r[i] = false;
        for (int s = i + 1; s < nrk; s++) if (c.rk.b[s][i] != 0.0) c.keep_r[i] = true;
        if (c.keep_r[i]) HIPCHK(c.r[i].alloc((size_t)3 * dof * ne));
    }
    for (int i = 0; i < 2; i++) HIPCHK(c.wd[i].alloc(ne));
    // the host vectors go out of scope: wait for the copies
    HIPCHK(hipStreamSynchronize(c.stream));
    c.h_sinfo = std::move(sinfo);
    c.h_snbr = std::move(snbr);
    if (upload_met(c)) return 1;
    return upload_forcing(c);
}

// Reference tables, RK scheme, stream and the physics constants shared by all members
// (src/dgswem.F:131-170).
int init_ctx_common(Ctx& c, const dg_params* prm, bool manning) {
    c.prm = *prm;
    std::string err;
    if (!build_ref_tables(prm->p, c.T, err)) return fail(err);
    if (!build_rk(prm->rk_stage, prm->rk_order, c.rk, err)) return fail(err);
    // the kernels start edge traces at mode 1's coefficient, which needs PHI_EDGE(1, .) == 1
    for (int l = 0; l < 3; l++)
        for (int q = 0; q < c.T.negp; q++)
            if (c.T.phi_edge[l][q][0] != 1.0) return fail("basis tables: PHI_EDGE of mode 1 is not 1");
    // the stage kernel forms the P1 bathymetry gradient once per element: the linear modes'
    // derivative tables must be the same at every area point
    for (int q = 0; q < c.T.nagp; q++)
        if (c.T.phi_area[q][0] != 1.0) return fail("basis tables: PHI_AREA of mode 1 is not 1");
    for (int q = 1; q < c.T.nagp; q++)
        for (int k = 0; k < 3; k++)
            if (c.T.drphi[q][k] != c.T.drphi[0][k] || c.T.dsphi[q][k] != c.T.dsphi[0][k])
                return fail("basis tables: derivatives of the linear modes vary between area points");
    c.dof = c.T.dof; c.nagp = c.T.nagp; c.negp = c.T.negp;
    c.block = 128;
    {
        int dev = 0, lds = 0;
        HIPCHK(hipGetDevice(&dev));
        HIPCHK(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        c.max_lds = lds;
        c.dev = dev;
    }
    HIPCHK(hipStreamCreateWithFlags(&c.stream, hipStreamNonBlocking));
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (!g_tables_set[prm->p]) {
            DevTables d = to_dev(c.T);
            if (prm->p == 1) HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_tab1), &d, sizeof d));
            else HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_tab2), &d, sizeof d));
            g_tables_set[prm->p] = true;
        }
    }
    Phys& P = c.phys;
    P.g = prm->g;
    P.nleq = prm->nolicat == 0 ? 0.0 : 1.0;
    P.leq = prm->nolicat == 0 ? 1.0 : 0.0;
    P.nleqg = P.nleq * P.g;
    P.fg_l = P.leq * P.g;
    P.hbreak = prm->hbreak; P.ftheta = prm->ftheta; P.fgamma = prm->fgamma;
    P.ifnlfa = prm->nolifa == 0 ? 0 : 1;
    P.iflinbf = prm->nolibf == 0 ? 1 : 0;
    P.ifhybf = prm->nolibf == 2 ? 1 : 0;
    P.load_manning = manning ? 1 : 0;
    return 0;
}

// Members' cold-start state into the stage-1 slot.
int upload_initial_state(Ctx& c) {
    const int ne = c.ne, dof = c.dof;
    std::vector<double> u((size_t)3 * dof * ne, 0.0);
    std::vector<uint8_t> w(ne, 1);
    for (int s = 0; s < c.S; s++) {
        const Member& m = c.mem[s];
        for (int k = 0; k < dof; k++)
            for (int e = 0; e < m.ne; e++) u[(size_t)k * ne + c.eoff[s] + e] = m.ze0[(size_t)k * m.ne + e];
        for (int e = 0; e < m.ne; e++) w[c.eoff[s] + e] = m.wd0[e];
    }
    c.ucur = 0; c.wdcur = 0;
    HIPCHK(hipMemcpyAsync(c.u[0].p, u.data(), u.size() * sizeof(double), hipMemcpyHostToDevice, c.stream));
    HIPCHK(hipMemcpyAsync(c.wd[0].p, w.data(), ne, hipMemcpyHostToDevice, c.stream));
    HIPCHK(hipStreamSynchronize(c.stream));
    return 0;
}

// The non-periodic boundary records the reference reads as time goes on: its state (the two
// records in use and their times) before step 1, and the advance of one stage.  An advance
// happens at a stage whose TIME_A is past the second record's time (src/DG_hydro_timestep.F:153-160
// for fort.19, 174-185 for fort.20); at most one per stage, as in the reference.
void series_reset(const Member& m, Member::Series& S, bool elev) {
    const int n = elev ? m.neta : m.nflux;
    const double* rec = elev ? m.esbin.data() : m.qnin.data();
    S.it = 0;
    S.t1 = m.prm.statim * 86400.0;
    S.t2 = S.t1 + (elev ? m.etiminc : m.ftiminc);
    S.next = 2;
    if (elev) {
        S.v1.assign(rec, rec + n);
        S.v2.assign(rec + n, rec + 2 * n);
    } else {
        // QNIN1/QNIN2 over the NVEL nodes (zero where the code is not a flux code), one slot past
        // NVEL for the edge after the last flux node
        S.v1.assign(m.nvel + 1, 0.0);
        S.v2.assign(m.nvel + 1, 0.0);
        for (int j = 0, k = 0; j < m.nvel; j++)
            if (m.lbcodei[j] == 2 || m.lbcodei[j] == 12 || m.lbcodei[j] == 22) { S.v1[j] = rec[k]; S.v2[j] = rec[n + k]; k++; }
    }
}

int series_advance(const Member& m, Member::Series& S, bool elev, double time_a) {
    if (!(time_a > S.t2)) return 0;
    const int nrec = elev ? m.e_nrec : m.f_nrec;
    S.t1 = S.t2;
    S.t2 = S.t1 + (elev ? m.etiminc : m.ftiminc);
    if (S.next >= nrec)
        return fail(std::string(elev ? "fort.19" : "fort.20") + ": the time series ends before the run");
    if (elev) {
        S.v1 = S.v2;
        const double* r = m.esbin.data() + (size_t)S.next * m.neta;
        S.v2.assign(r, r + m.neta);
    } else {
        const double* r = m.qnin.data() + (size_t)S.next * m.nflux;
        for (int j = 0, k = 0; j < m.nvel; j++) {
            S.v1[j] = S.v2[j];
            S.v2[j] = 0.0;
            if (m.lbcodei[j] == 2 || m.lbcodei[j] == 12 || m.lbcodei[j] == 22) S.v2[j] = r[k++];
        }
    }
    S.next++;
    return 0;
}

// Bring a series to the state after step `it` (replayed from the start when an earlier step is
// asked for, as the adjoint's recomputation and the autotuner do).
int series_seek(Member& m, Member::Series& S, bool elev, int it, int nrk) {
    if (S.it > it || S.v1.empty()) series_reset(m, S, elev);
    while (S.it < it) {
        const double time_a = (S.it + 1) * m.prm.dtdp + m.prm.statim * 86400.0;
        for (int irk = 0; irk < nrk; irk++)
            if (series_advance(m, S, elev, time_a)) return 1;
        S.it++;
    }
    return 0;
}

// Per-stage forcing rows for steps it..it+n-1 of every member (host arithmetic of
// src/DG_hydro_timestep.F:79-200, src/ocean_edge_hydro.F:59-75, src/flow_edge_hydro.F:45-62,
// src/tidal_potential.F:26-48).
int build_stage_tables(Ctx& c, int it, int n) {
    if (c.nbfr == 0 && c.nffr == 0 && !c.met && !c.tip) return 0;
    if (n > kTableChunk) return fail("internal: table chunk overflow");
    const int nrk = c.rk.nrk, rows = c.rows;
    const int tot_f = c.foff[c.S];
    const bool ramped = c.met || c.tip;
    std::vector<double> aj((size_t)c.S * rows * std::max(c.nbfr, 1), 0.0), rf(aj.size(), 0.0);
    std::vector<double> fa((size_t)c.S * rows * std::max(c.nffr, 1), 0.0), fr(fa.size(), 0.0);
    std::vector<double> rp, mr, et, qt, tr;
    if (ramped) rp.assign((size_t)c.S * rows, 0.0);
    if (c.met) mr.assign((size_t)c.S * kTableChunk * MET_ROW, 0.0);
    if (c.eser) et.assign((size_t)c.S * rows * std::max(c.tzn, 1), 0.0);
    if (c.fser) qt.assign((size_t)c.S * rows * tot_f * 2, 0.0);
    if (c.tip) tr.assign((size_t)c.S * rows * c.ntif * 3, 0.0);
    std::vector<double> eta_bc;
    std::string err;
    for (int sc = 0; sc < c.S; sc++) {
        Member& m = c.mem[sc];
        if (m.eser && series_seek(m, m.es, true, it - 1, nrk)) return 1;
        if (m.fser && series_seek(m, m.fs, false, it - 1, nrk)) return 1;
        if (m.eser) eta_bc.assign(m.np, 0.0);
        for (int s = 0; s < n; s++) {
            const int IT = it + s;
            const double itdt = IT * m.prm.dtdp;
            const double time_a = itdt + m.prm.statim * 86400.0;
            const double timeh = itdt + (m.prm.statim - m.reftim) * 86400.0;
            // storm scalars at TIME_A (MET_FORCING is called with TIME_A at every stage)
            if (c.met && !holland_row(m.track, time_a, m.prm.g, &mr[((size_t)sc * kTableChunk + s) * MET_ROW], err))
                return fail(err);
            for (int irk = 0; irk < nrk; irk++) {
                const size_t row = (size_t)sc * rows + (size_t)s * nrk + irk;
                const double timedg = time_a - m.prm.dtdp + c.rk.dtvd[irk] * m.prm.dtdp;
                double ramp = 1.0;
                if (m.prm.nramp == 1)
                    ramp = std::tanh((2.0 * ((IT - 1) + c.rk.dtvd[irk]) * m.prm.dtdp / 86400.0) / m.prm.dramp);
                if (ramped) rp[row] = ramp;
                if (m.ntif > 0) {
                    const double timeh_dg = timeh - m.prm.dtdp + c.rk.dtvd[irk] * m.prm.dtdp;
                    for (int j = 0; j < m.ntif; j++) {
                        const int ncyc = (m.pert[j] == 0.0) ? 0 : (int)(timeh_dg / m.pert[j]);
                        double* t = &tr[(row * c.ntif + j) * 3];
                        t[0] = m.amigt[j] * (timeh_dg - ncyc * m.pert[j]) + m.facet[j];
                        t[1] = ramp * m.etrf[j] * m.tpk[j] * m.fft[j];
                        t[2] = ramp * m.fft[j];
                    }
                }
                if (m.eser) {
                    if (series_advance(m, m.es, true, time_a)) return 1;
                    const double etratio = (timedg - m.es.t1) / m.etiminc;
                    for (int i = 0; i < m.neta; i++)
                        eta_bc[m.nbd[i]] = ramp * (m.es.v1[i] + etratio * (m.es.v2[i] - m.es.v1[i]));
                    for (int L = 0; L < m.needs; L++)
                        for (int k = 0; k < 2; k++)
                            et[row * c.tzn + (size_t)(c.ooff[sc] + L) * 2 + k] = eta_bc[m.onode[2 * L + k]];
                } else {
                    for (int jj = 0; jj < m.nbfr; jj++) {
                        const int ncyc = (m.per[jj] == 0.0) ? 0 : (int)(timedg / m.per[jj]);
                        aj[row * c.nbfr + jj] = m.amig[jj] * (timedg - ncyc * m.per[jj]) + m.face[jj];
                        rf[row * c.nbfr + jj] = m.ff[jj] * ramp;
                    }
                }
                if (m.fser) {
                    // one term: ARGJ = 0, RFF = RAMPDG (src/flow_edge_hydro.F:147-149)
                    fa[row * c.nffr] = 0.0;
                    fr[row * c.nffr] = ramp;
                    if (series_advance(m, m.fs, false, time_a)) return 1;
                    const double qtratio = (timedg - m.fs.t1) / m.ftiminc;
                    int nqeds = 0;
                    for (int i = 0; i < m.nvel; i++) {
                        const int code = m.lbcodei[i];
                        if (!(code == 2 || code == 12 || code == 22)) continue;
                        nqeds++;
                        if (nqeds > m.nfeds) continue;
                        const double* v1 = m.fs.v1.data();
                        const double* v2 = m.fs.v2.data();
                        double* q = &qt[(row * tot_f + c.foff[sc] + nqeds - 1) * 2];
                        q[0] = ramp * (v1[i] + qtratio * (v2[i] - v1[i]));
                        q[1] = ramp * (v1[i + 1] + qtratio * (v2[i + 1] - v1[i + 1]));
                    }
                } else {
                    for (int jj = 0; jj < m.nffr; jj++) {
                        const int ncyc = (m.fper[jj] == 0.0) ? 0 : (int)(timedg / m.fper[jj]);
                        fa[row * c.nffr + jj] = m.famig[jj] * (timedg - ncyc * m.fper[jj]) + m.fface[jj];
                        fr[row * c.nffr + jj] = m.fff[jj] * ramp;
                    }
                }
            }
            if (m.eser) m.es.it = IT;
            if (m.fser) m.fs.it = IT;
        }
    }
    // the previous chunk's launches may still read the tables
    HIPCHK(hipStreamSynchronize(c.stream));
    if (c.nbfr) {
        HIPCHK(hipMemcpyAsync(c.argj.p, aj.data(), aj.size() * sizeof(double), hipMemcpyHostToDevice, c.stream));
        HIPCHK(hipMemcpyAsync(c.rff.p, rf.data(), rf.size() * sizeof(double), hipMemcpyHostToDevice, c.stream));
    }
    if (c.nffr) {
        HIPCHK(hipMemcpyAsync(c.fargj.p, fa.data(), fa.size() * sizeof(double), hipMemcpyHostToDevice, c.stream));
        HIPCHK(hipMemcpyAsync(c.frff.p, fr.data(), fr.size() * sizeof(double), hipMemcpyHostToDevice, c.stream));
    }
    if (ramped) HIPCHK(hipMemcpyAsync(c.ramp.p, rp.data(), rp.size() * sizeof(double), hipMemcpyHostToDevice, c.stream));
    if (c.met) HIPCHK(hipMemcpyAsync(c.mrow.p, mr.data(), mr.size() * sizeof(double), hipMemcpyHostToDevice, c.stream));
    if (c.eser) HIPCHK(hipMemcpyAsync(c.ets.p, et.data(), et.size() * sizeof(double), hipMemcpyHostToDevice, c.stream));
    if (c.fser) HIPCHK(hipMemcpyAsync(c.qts.p, qt.data(), qt.size() * sizeof(double), hipMemcpyHostToDevice, c.stream));
    if (c.tip) HIPCHK(hipMemcpyAsync(c.tiprow.p, tr.data(), tr.size() * sizeof(double), hipMemcpyHostToDevice, c.stream));
    HIPCHK(hipStreamSynchronize(c.stream));
    return 0;
}

// Static LDS bytes of stage_kernel<P, NT, ..., MET, ...> (its __shared__ arrays).
size_t stage_static_lds(int dof, int negp, int nt, bool met) {
    (void)met;
    return (size_t)8 * (2 * 3 * dof * nt + 2 * 3 * negp * dof);
}

// Dynamic LDS bytes of adj_stage_kernel: 6*dof staged rows of ns doubles.
size_t adj_dynamic_lds(int dof, int ns) { return (size_t)8 * (6 * dof) * ns; }

// Block size the stage kernel runs with: the chosen one for nonlinear runs, 64 for linearised runs.
int stage_nt(const Ctx& c, int nt) { return c.prm.nolicat != 0 ? nt : 64; }

// Halo lists of the stage kernel's blocks of `nt` elements (tile t = elements [t*nt, t*nt+nt)):
// the out-of-tile internal neighbours of each tile in ascending order, and per element the local
// staging index of each internal side's neighbour (tile element j -> j, halo entry i -> nt + i),
// packed 10 bits per side.  Built once per block size; rejected when it does not fit in LDS.
int ensure_staging(Ctx& c, int nt) {
    if (c.stg.count(nt)) return 0;
    const int ne = c.ne, ntiles = (ne + nt - 1) / nt;
    std::vector<int> hoff(ntiles + 1, 0), hidx, snbl(ne, 0), halo;
    int hmax = 0;
    for (int t = 0; t < ntiles; t++) {
        const int b = t * nt, en = std::min(ne, b + nt);
        halo.clear();
        for (int e = b; e < en; e++)
            for (int sl = 0; sl < 3; sl++) {
                if (((side_bits(c.h_sinfo[e], sl) >> 3) & 3) != SIDE_INTERNAL) continue;
                const int nb = c.h_snbr[(size_t)sl * ne + e];
                if (nb < b || nb >= en) halo.push_back(nb);
            }
        std::sort(halo.begin(), halo.end());
        halo.erase(std::unique(halo.begin(), halo.end()), halo.end());
        hmax = std::max(hmax, (int)halo.size());
        for (int e = b; e < en; e++) {
            int packed = 0;
            for (int sl = 0; sl < 3; sl++) {
                if (((side_bits(c.h_sinfo[e], sl) >> 3) & 3) != SIDE_INTERNAL) continue;
                const int nb = c.h_snbr[(size_t)sl * ne + e];
                int ln;
                if (nb >= b && nb < en) ln = nb - b;
                else ln = nt + (int)(std::lower_bound(halo.begin(), halo.end(), nb) - halo.begin());
                packed |= ln << (10 * sl);
            }
            snbl[e] = packed;
        }
        hidx.insert(hidx.end(), halo.begin(), halo.end());
        hoff[t + 1] = (int)hidx.size();
    }
    const int ns = nt + hmax;
    if (ns > 1023) return fail("stage kernel staging: halo too large for block size " + std::to_string(nt));
    const size_t dyn = ((size_t)8 * (3 * c.dof + 3) * ns + ns + 7) / 8 * 8;
    if (dyn + stage_static_lds(c.dof, c.negp, nt, c.met) > (size_t)c.max_lds)
        return fail("stage kernel staging: block size " + std::to_string(nt) + " needs more LDS than a block may use");
    Ctx::Staging& S = c.stg[nt];
    S.ns = ns;
    S.lds = dyn;
    if (hidx.empty()) hidx.push_back(0);
    HIPCHK(S.hoff.upload(hoff.data(), hoff.size(), c.stream));
    HIPCHK(S.hidx.upload(hidx.data(), hidx.size(), c.stream));
    HIPCHK(S.snbl.upload(snbl.data(), snbl.size(), c.stream));
    HIPCHK(hipStreamSynchronize(c.stream));
    return 0;
}

template <int P, int NT, bool MET, bool NL>
int launch_stage_v(Ctx& c, const StageArgs& a0, bool sph, bool fuse) {
    const int grid = (c.ne + NT - 1) / NT;
    auto it = c.stg.find(NT);
    if (it == c.stg.end()) return fail("internal: stage kernel staging not built for this block size");
    StageArgs a = a0;
    a.hoff = it->second.hoff.p; a.hidx = it->second.hidx.p; a.snbl = it->second.snbl.p; a.ns = it->second.ns;
    const size_t lds = it->second.lds;
    // every pointer the variant dereferences must be set (a missing one is an error, not a launch)
    if (MET && (!a.mt.row || !a.mt.cd || !a.mt.wm || !a.mt.wx || !a.mt.wy || !a.mt.prn || !a.fo.ramp))
        return fail("internal: NWS=8 stage launch without its forcing arguments");
    if (!a.uin || !a.uout || !a.wd_in || !a.hoff || !a.hidx || !a.snbl)
        return fail("internal: stage launch with a missing buffer");
    if (sph) {
        if (fuse) hipLaunchKernelGGL((stage_kernel<P, NT, true, true, MET, NL>), dim3(grid), dim3(NT), lds, c.stream, a);
        else hipLaunchKernelGGL((stage_kernel<P, NT, true, false, MET, NL>), dim3(grid), dim3(NT), lds, c.stream, a);
    } else {
        if (fuse) hipLaunchKernelGGL((stage_kernel<P, NT, false, true, MET, NL>), dim3(grid), dim3(NT), lds, c.stream, a);
        else hipLaunchKernelGGL((stage_kernel<P, NT, false, false, MET, NL>), dim3(grid), dim3(NT), lds, c.stream, a);
    }
    return 0;
}

// The step's boundary and potential forcing for all stages, before the step's stage launches (the
// forcing rows of the step must be in the tables): the open-boundary elevation terms (tide_kernel),
// the specified normal flux (flow_kernel) and the tidal potential at the nodes (tip_kernel).
int launch_forcing(Ctx& c, bool capturing) {
    const Forcing F = forcing(c, capturing);
    if (c.tzn > 0) {
        const int n = c.rk.nrk * c.tzn;
        hipLaunchKernelGGL(tide_kernel, dim3((n + 255) / 256), dim3(256), 0, c.stream, F,
                           (const uint16_t*)(c.S > 1 ? c.osid.p : nullptr), c.tz.p);
        HIPCHK(hipGetLastError());
    }
    if (c.fzn > 0) {
        const int n = c.rk.nrk * c.fzn;
        const uint16_t* fs = c.S > 1 ? c.fsid.p : nullptr;
        if (c.prm.p == 1)
            hipLaunchKernelGGL(flow_kernel<1>, dim3((n + 255) / 256), dim3(256), 0, c.stream, F, fs, c.fz.p);
        else
            hipLaunchKernelGGL(flow_kernel<2>, dim3((n + 255) / 256), dim3(256), 0, c.stream, F, fs, c.fz.p);
        HIPCHK(hipGetLastError());
    }
    if (c.tip) {
        const int n = c.rk.nrk * c.np;
        hipLaunchKernelGGL(tip_kernel, dim3((n + 255) / 256), dim3(256), 0, c.stream, tip_args(c), F);
        HIPCHK(hipGetLastError());
    }
    return 0;
}

// Nonlinear runs (NOLICAT != 0) use the kernels with NLEQ/LEQ folded at compile time; linearised
// runs use the general kernels with 64-thread blocks.
template <int P, int NT>
int launch_stage(Ctx& c, const StageArgs& a, bool sph, bool fuse) {
    if (c.prm.nolicat != 0) {
        if (c.met) return launch_stage_v<P, NT, true, true>(c, a, sph, fuse);
        return launch_stage_v<P, NT, false, true>(c, a, sph, fuse);
    }
    if (c.met) return launch_stage_v<P, 64, true, false>(c, a, sph, fuse);
    return launch_stage_v<P, 64, false, false>(c, a, sph, fuse);
}

// Launch shapes, by threads per block (one thread per element): 64, 128, and 256 for p=1.
bool valid_block(int p, int nt) {
    if (p == 1) return nt == 64 || nt == 128 || nt == 256;
    return nt == 64 || nt == 128;
}

template <int P>
int launch_stage_nt(Ctx& c, const StageArgs& a, bool sph, bool fuse, int nt) {
    if (nt == 64) return launch_stage<P, 64>(c, a, sph, fuse);
    if constexpr (P == 1) {
        if (nt == 128) return launch_stage<P, 128>(c, a, sph, fuse);
        return launch_stage<P, 256>(c, a, sph, fuse);
    }
    return launch_stage<P, 128>(c, a, sph, fuse);
}

// The stage kernel's copy of its output cell means: needed by the limiter when WETDRY follows it
// (WETDRY rewrites cell means in place).
double* cmean_out(Ctx& c) { return c.prm.slopeflag != 0 && c.prm.nolifa >= 2 ? c.cmean.p : nullptr; }

StageArgs stage_args(Ctx& c, int irk, const int slot[], bool captured = false) {
    StageArgs a{};
    a.ed = elem_data(c); a.ph = c.phys; a.fo = forcing(c, captured);
    if (c.met) a.mt = met_args(c);
    a.uin = c.u[slot[irk]].p;
    a.uout = c.u[slot[irk + 1]].p;
    a.wd_in = c.wd[c.wdcur].p;
    a.wd_out = c.wd[1 - c.wdcur].p;
    a.nprev = irk;
    for (int i = 0; i < irk; i++) {
        a.uprev[i] = c.u[slot[i]].p; a.aprev[i] = c.rk.a[irk][i];
        a.rprev[i] = c.r[i].p; a.bprev[i] = c.rk.b[irk][i];
    }
    a.acur = c.rk.a[irk][irk];
    a.bcur = c.rk.b[irk][irk];
    a.rout = c.keep_r[irk] ? c.r[irk].p : nullptr;
    a.irk = irk;
    a.wd_active = c.wd_active;
    a.tip = c.tip ? c.tip2.p + (size_t)irk * c.np : nullptr;
    a.cmout = cmean_out(c);
    return a;
}

// SLOPELIMITER5 (+ WETDRY) on the stage output u, in place; with WETDRY the stage kernel has left
// u's cell means in c.cmean.
template <int P>
int launch_limiter(Ctx& c, double* u, const ElemData& ed, bool capturing) {
    const int ne = c.ne;
    const int ge = (ne + LIM_NT - 1) / LIM_NT;
    if (!c.lim.elq.p || !c.cmean.p) return fail("internal: slope limiter tiles not built");
    const bool wdon = c.prm.nolifa >= 2;
    const double* cm = wdon ? c.cmean.p : u;
    const size_t cms = wdon ? (size_t)ne : (size_t)c.dof * ne;
    LimTiles lt{c.lim.vhoff.p, c.lim.vhidx.p, c.lim.qoff.p, c.lim.qinc_off.p, c.lim.qinc.p, c.lim.elq.p,
                c.lim.nsl, c.lim.nql};
    const size_t lds = (size_t)8 * (3 * lt.nsl + 6 * lt.nql);
    int ev = prof_begin(c, K_LIMIT, capturing);
    if (c.prm.nolifa >= 2) {
        hipLaunchKernelGGL((limit_wetdry_kernel<P, true, false>), dim3(ge), dim3(LIM_NT), lds, c.stream, ed, u, cm,
                           cms, lt, (const uint8_t*)c.wd[c.wdcur].p, c.wd[1 - c.wdcur].p, c.wd_active,
                           (double*)nullptr, (double*)nullptr);
        c.wdcur = 1 - c.wdcur;
    } else {
        hipLaunchKernelGGL((limit_wetdry_kernel<P, false, false>), dim3(ge), dim3(LIM_NT), lds, c.stream, ed, u, cm,
                           cms, lt, (const uint8_t*)c.wd[c.wdcur].p, c.wd[1 - c.wdcur].p, (int*)nullptr,
                           (double*)nullptr, (double*)nullptr);
    }
    prof_end(c, ev);
    HIPCHK(hipGetLastError());
    return 0;
}

// Frozen-branch WETDRY reverse of one stage: SLOPELIMITER5 formed again from the stage's
// pre-limiter output wpre (not written), each element's limited modes going straight to its
// WETDRY reverse with the stage's flags `flags`: cot (in place) becomes the cotangent of the
// limiter's output, gdp gains the vertex-depth terms.
template <int P>
int launch_wetdry_rev(Ctx& c, const double* wpre, const uint8_t* flags, double* cot) {
    const int ne = c.ne;
    const int ge = (ne + LIM_NT - 1) / LIM_NT;
    if (!c.lim.elq.p) return fail("internal: slope limiter tiles not built");
    LimTiles lt{c.lim.vhoff.p, c.lim.vhidx.p, c.lim.qoff.p, c.lim.qinc_off.p, c.lim.qinc.p, c.lim.elq.p,
                c.lim.nsl, c.lim.nql};
    const size_t lds = (size_t)8 * (3 * lt.nsl + 6 * lt.nql);
    const ElemData ed = elem_data(c);
    hipLaunchKernelGGL((limit_wetdry_kernel<P, false, true>), dim3(ge), dim3(LIM_NT), lds, c.stream, ed,
                       const_cast<double*>(wpre), wpre, (size_t)c.dof * ne, lt, flags, (uint8_t*)nullptr,
                       (int*)nullptr, cot, c.gdp.p);
    HIPCHK(hipGetLastError());
    return 0;
}

// One time step: NRK stages (stage kernel, then node min/max + limiter/wet-dry when the slope
// limiter is on) and the device step-counter increment.
template <int P>
int launch_step(Ctx& c, bool capturing) {
    if (launch_forcing(c, capturing)) return 1;
    const int np = c.np;
    const int B = 256;
    const int gn = (np + B - 1) / B;
    const bool wdon = c.prm.nolifa >= 2, lim = c.prm.slopeflag != 0, sph = c.prm.ics == 2;
    const int nrk = c.rk.nrk;
    // stage slots: slot(i) holds U(i+1); slot(0) = current stage-1 state
    int slot[KMAXRK + 1];
    for (int i = 0; i <= nrk; i++) slot[i] = (c.ucur + i) % (nrk + 1);
    if (c.prof && !capturing) c.prof_this = (c.prof_step++ % c.prof_every) == 0;
    if (c.met) {
        int ev = prof_begin(c, K_MET, capturing);
        hipLaunchKernelGGL(met_kernel, dim3(gn), dim3(B), 0, c.stream, met_args(c),
                           capturing ? (const int*)c.step.p : nullptr, c.hstep);
        prof_end(c, ev);
        HIPCHK(hipGetLastError());
    }
    for (int irk = 0; irk < nrk; irk++) {
        StageArgs a = stage_args(c, irk, slot, capturing);
        for (int i = 0; i < irk; i++)
            if (a.bprev[i] != 0.0 && !c.keep_r[i]) return fail("internal: RHS of an earlier stage was not kept");
        const bool fuse = wdon && !lim;
        int ev = prof_begin(c, K_STAGE, capturing);
        if (launch_stage_nt<P>(c, a, sph, fuse, c.block)) return 1;
        prof_end(c, ev);
        HIPCHK(hipGetLastError());
        if (fuse) c.wdcur = 1 - c.wdcur;
        if (lim && launch_limiter<P>(c, a.uout, a.ed, capturing)) return 1;
    }
    if (capturing) {
        hipLaunchKernelGGL(step_counter_kernel, dim3(1), dim3(1), 0, c.stream, c.step.p, 0, 0);
        HIPCHK(hipGetLastError());
    } else {
        c.hstep++;
    }
    c.prof_this = true;
    c.ucur = slot[nrk];
    return 0;
}

int launch_steps(Ctx& c, int n, bool capturing) {
    for (int s = 0; s < n; s++) {
        const int rc = c.prm.p == 1 ? launch_step<1>(c, capturing) : launch_step<2>(c, capturing);
        if (rc) return rc;
    }
    return 0;
}

// One step with explicit state pointers: st[0] input, st[1..NRK-1] the intermediate stage states,
// st[NRK] output.  With the slope limiter on, each stage's output before limiting is also copied to
// wsave[irk] (when wsave is given) for the adjoint's frozen-limiter reverse, and the wet/dry flags
// at the start of each stage to fsave[irk] (when given).
template <int P>
int launch_step_explicit(Ctx& c, double* const* st, double* const* wsave = nullptr, uint8_t* const* fsave = nullptr) {
    const int nrk = c.rk.nrk;
    const bool sph = c.prm.ics == 2;
    const bool wdon = c.prm.nolifa >= 2, lim = c.prm.slopeflag != 0, fuse = wdon && !lim;
    const size_t ssz = (size_t)3 * c.dof * c.ne;
    if (launch_forcing(c, false)) return 1;
    if (c.met) {
        hipLaunchKernelGGL(met_kernel, dim3((c.np + 255) / 256), dim3(256), 0, c.stream, met_args(c), nullptr, c.hstep);
        HIPCHK(hipGetLastError());
    }
    for (int irk = 0; irk < nrk; irk++) {
        StageArgs a{};
        a.ed = elem_data(c); a.ph = c.phys; a.fo = forcing(c);
        if (c.met) a.mt = met_args(c);
        a.uin = st[irk];
        a.uout = st[irk + 1];
        a.wd_in = c.wd[c.wdcur].p;
        a.wd_out = c.wd[1 - c.wdcur].p;
        a.nprev = irk;
        for (int i = 0; i < irk; i++) {
            a.uprev[i] = st[i]; a.aprev[i] = c.rk.a[irk][i];
            a.rprev[i] = c.r[i].p; a.bprev[i] = c.rk.b[irk][i];
        }
        a.acur = c.rk.a[irk][irk];
        a.bcur = c.rk.b[irk][irk];
        a.rout = c.keep_r[irk] ? c.r[irk].p : nullptr;
        a.irk = irk;
        a.wd_active = c.wd_active;
        a.tip = c.tip ? c.tip2.p + (size_t)irk * c.np : nullptr;
        a.cmout = cmean_out(c);
        if (fsave) HIPCHK(hipMemcpyAsync(fsave[irk], c.wd[c.wdcur].p, c.ne, hipMemcpyDeviceToDevice, c.stream));
        if (launch_stage_nt<P>(c, a, sph, fuse, c.block)) return 1;
        HIPCHK(hipGetLastError());
        if (fuse) c.wdcur = 1 - c.wdcur;
        if (lim) {
            if (wsave)
                HIPCHK(hipMemcpyAsync(wsave[irk], st[irk + 1], ssz * sizeof(double), hipMemcpyDeviceToDevice,
                                      c.stream));
            if (launch_limiter<P>(c, st[irk + 1], a.ed, false)) return 1;
        }
    }
    c.hstep++;
    return 0;
}

// The steps of one table chunk: replays of a captured graph of one slot-rotation period (NRK+1
// steps, after which the stage slots and wet/dry buffers are back in place), then eager launches
// for the remainder.
int run_chunk(Ctx& c, int n) {
    c.hstep = 0;
    const int period = c.rk.nrk + 1;
    int done = 0;
    if (c.use_graph && n >= period) {
        hipLaunchKernelGGL(step_counter_kernel, dim3(1), dim3(1), 0, c.stream, c.step.p, 1, 0);
        HIPCHK(hipGetLastError());
        const int key = c.ucur * 2 + c.wdcur;
        auto it = c.graphs.find(key);
        if (it == c.graphs.end()) {
            const int u0 = c.ucur, w0 = c.wdcur;
            HIPCHK(hipStreamBeginCapture(c.stream, hipStreamCaptureModeThreadLocal));
            const int rc = launch_steps(c, period, true);
            hipGraph_t g = nullptr;
            const hipError_t ee = hipStreamEndCapture(c.stream, &g);
            if (rc) { if (g) (void)hipGraphDestroy(g); return rc; }
            if (ee != hipSuccess) return fail(std::string("hipStreamEndCapture: ") + hipGetErrorString(ee));
            if (c.ucur != u0 || c.wdcur != w0) { (void)hipGraphDestroy(g); return fail("internal: graph period mismatch"); }
            hipGraphExec_t ex = nullptr;
            const hipError_t ie = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ie != hipSuccess) return fail(std::string("hipGraphInstantiate: ") + hipGetErrorString(ie));
            it = c.graphs.emplace(key, ex).first;
        }
        const int reps = n / period;
        for (int r = 0; r < reps; r++) HIPCHK(hipGraphLaunch(it->second, c.stream));
        done = reps * period;
        c.hstep = done;
    }
    return launch_steps(c, n - done, false);
}

// ---- discrete adjoint driver ------------------------------------------------------------
int adj_check(Ctx& c) {
    if (c.ng == 0) return fail("no gauges set (dg_adj_set_gauges)");
    if (c.tip || c.eser || c.fser)
        return fail("the adjoint does not cover the tidal potential or non-periodic boundary forcing");
    return 0;
}

template <int P, int NT, bool NL, bool WD>
int launch_adj_v(Ctx& c, const AdjArgs& a, int grid, size_t lds, bool qf) {
    if (c.prm.ics == 2) {
        if (qf) hipLaunchKernelGGL((adj_stage_kernel<P, NT, true, NL, NL, WD>), dim3(grid), dim3(NT), lds, c.stream, a);
        else hipLaunchKernelGGL((adj_stage_kernel<P, NT, true, NL, false, WD>), dim3(grid), dim3(NT), lds, c.stream, a);
    } else {
        if (qf) hipLaunchKernelGGL((adj_stage_kernel<P, NT, false, NL, NL, WD>), dim3(grid), dim3(NT), lds, c.stream, a);
        else hipLaunchKernelGGL((adj_stage_kernel<P, NT, false, NL, false, WD>), dim3(grid), dim3(NT), lds, c.stream, a);
    }
    return 0;
}

template <int P, int NT, bool NL>
int launch_adj_nl(Ctx& c, const AdjArgs& a0) {
    const int grid = (c.ne + NT - 1) / NT;
    auto it = c.stg.find(NT);
    if (it == c.stg.end()) return fail("internal: adjoint staging not built for this block size");
    AdjArgs a = a0;
    a.hoff = it->second.hoff.p; a.hidx = it->second.hidx.p; a.snbl = it->second.snbl.p; a.ns = it->second.ns;
    const size_t lds = adj_dynamic_lds(c.dof, it->second.ns);
    if (a.met && (!a.mt.cd || !a.mt.wm || !a.mt.wx || !a.mt.wy || !a.mt.prn || !a.fo.ramp))
        return fail("internal: NWS=8 adjoint launch without its forcing arguments");
    if (!a.u || !a.out) return fail("internal: adjoint launch with a missing buffer");
    // NOLIBF=1 (quadratic friction, no hybrid factor) with the nonlinear equations: friction fixed
    // at compile time
    const bool qf = NL && c.phys.iflinbf == 0 && c.phys.ifhybf == 0;
    if (a.wd) return launch_adj_v<P, NT, NL, true>(c, a, grid, lds, qf);
    return launch_adj_v<P, NT, NL, false>(c, a, grid, lds, qf);
}

template <int P, int NT>
int launch_adj(Ctx& c, const AdjArgs& a) {
    if (c.prm.nolicat != 0) return launch_adj_nl<P, NT, true>(c, a);
    return launch_adj_nl<P, NT, false>(c, a);
}

// Block size of the adjoint stage kernel.
int adj_nt(int p) { return p == 1 ? 128 : 64; }

// The adjoint kernel's block-local staging: the forward's halo lists for its block size, with its
// own row count (state and RHS cotangent per staged element).
int ensure_adj_staging(Ctx& c) {
    const int nt = adj_nt(c.prm.p);
    if (ensure_staging(c, nt)) return 1;
    const size_t stat = (size_t)8 * (3 * c.dof * nt + 2 * 3 * c.negp * c.dof);
    if (adj_dynamic_lds(c.dof, c.stg[nt].ns) + stat > (size_t)c.max_lds)
        return fail("adjoint staging needs more LDS than a block may use");
    return 0;
}

// Forward N steps from the current state recording the gauges after every step (geta[k-1] after
// step k) and saving the state every `interval` steps into ckpt; the handle's state ends at U^N.
template <int P>
int adj_forward(Ctx& c, int it0, int nsteps, int interval, bool save) {
    const size_t ssz = (size_t)3 * c.dof * c.ne;
    const int B = 256, gg = (c.ng + B - 1) / B;
    for (int s0 = 0; s0 < nsteps; s0 += kTableChunk) {
        const int n = std::min(kTableChunk, nsteps - s0);
        if (build_stage_tables(c, it0 + s0, n)) return 1;
        c.hstep = 0;
        for (int s = 0; s < n; s++) {
            const int k = s0 + s;   // steps done before this one
            if (save && k % interval == 0) {
                HIPCHK(hipMemcpyAsync(c.ckpt.p + (size_t)(k / interval) * ssz, c.u[c.ucur].p, ssz * sizeof(double),
                                      hipMemcpyDeviceToDevice, c.stream));
                if (c.adj_wd)
                    HIPCHK(hipMemcpyAsync(c.ckptwd.p + (size_t)(k / interval) * c.ne, c.wd[c.wdcur].p, c.ne,
                                          hipMemcpyDeviceToDevice, c.stream));
            }
            if (launch_step<P>(c, false)) return 1;
            hipLaunchKernelGGL((gauge_kernel<P>), dim3(gg), dim3(B), 0, c.stream, c.ng, c.ne, (const int*)c.gel.p,
                               (const double*)c.gphi.p, (const double*)c.u[c.ucur].p, c.geta.p + (size_t)k * c.ng);
            HIPCHK(hipGetLastError());
        }
    }
    return 0;
}

template <int P>
int adj_backward(Ctx& c, int it0, int nsteps, int interval, double weight) {
    const int nrk = c.rk.nrk, ne = c.ne;
    const size_t ssz = (size_t)3 * c.dof * ne;
    const int nseg = (nsteps + interval - 1) / interval;
    HIPCHK(hipMemsetAsync(c.lam.p, 0, ssz * sizeof(double), c.stream));
    HIPCHK(hipMemsetAsync(c.gdp.p, 0, (size_t)3 * ne * sizeof(double), c.stream));
    HIPCHK(hipMemsetAsync(c.gnb.p, 0, (size_t)ne * sizeof(double), c.stream));
    // seed: misfit at step N
    const int gs = (c.ngu + 63) / 64;
    auto seed = [&](int k, double* l) {   // misfit of the state after step k (1-based)
        hipLaunchKernelGGL((gauge_seed_kernel<P>), dim3(gs), dim3(64), 0, c.stream, c.ngu, ne, (const int*)c.goff.p,
                           (const int*)c.gord.p, (const int*)c.gel.p, (const double*)c.gphi.p,
                           (const double*)(c.geta.p + (size_t)(k - 1) * c.ng),
                           (const double*)(c.gdata.p + (size_t)(k - 1) * c.ng), weight, l);
    };
    seed(nsteps, c.lam.p);
    // frozen-limiter reverse of one stage: the cotangent `cot` of the limited stage output becomes
    // the cotangent of the pre-limiter output Wpre (node bounds from Wpre's cell means)
    const bool lim = c.prm.slopeflag != 0;
    const int np = c.np, B = 256, ge = (ne + B - 1) / B, gn = (np + B - 1) / B;
    const ElemData ed = elem_data(c);
    // frozen-branch WETDRY reverse of one stage (active wetting/drying): from the stage's limited
    // output, formed again from its pre-limiter output, and the flags at the start of the stage
    const bool wdrev_on = c.adj_wdrev;
    auto wdrev = [&](double* cot, const double* wpre, const uint8_t* flags) -> int {
        return launch_wetdry_rev<P>(c, wpre, flags, cot);
    };
    auto limrev = [&](double* cot, const double* wpre) -> int {
        hipLaunchKernelGGL((node_minmax_kernel<P>), dim3(gn), dim3(B), 0, c.stream, np, ne, (const int*)c.n2e_off.p,
                           (const int*)c.n2e.p, wpre, c.nmm.p);
        hipLaunchKernelGGL((limiter_adj_kernel<P>), dim3(ge), dim3(B), 0, c.stream, np, ed, wpre,
                           (const double*)c.nmm.p, cot, c.lnbc.p);
        hipLaunchKernelGGL((limiter_adj_node_kernel<P>), dim3(gn), dim3(B), 0, c.stream, np, ne,
                           (const int*)c.n2e_off.p, (const int*)c.n2e.p, wpre, (const double*)c.lnbc.p,
                           c.lgsum.p, c.larg.p);
        hipLaunchKernelGGL((limiter_adj_gather_kernel<P>), dim3(ge), dim3(B), 0, c.stream, np, ed,
                           (const double*)c.lgsum.p, (const int*)c.larg.p, cot);
        HIPCHK(hipGetLastError());
        return 0;
    };
    double* lam = c.lam.p;
    double* ub[KMAXRK + 1];
    for (int i = 0; i < nrk; i++) ub[i] = c.ubar[i].p;
    for (int sg = nseg - 1; sg >= 0; sg--) {
        const int k0 = sg * interval, k1 = std::min(k0 + interval, nsteps);
        const int n = k1 - k0;
        // stage states of steps k0+1..k1: seg[(k-k0-1)*nrk + i] holds stage i of step k
        auto stage = [&](int k, int i) { return c.seg.p + ((size_t)(k - k0 - 1) * nrk + i) * ssz; };
        // pre-limiter outputs: wstage(k, i) is stage i's output of step k before limiting
        auto wstage = [&](int k, int i) { return c.wseg.p + ((size_t)(k - k0 - 1) * nrk + i) * ssz; };
        // wet/dry flags at the start of stage i of step k (active wetting/drying)
        auto fstage = [&](int k, int i) { return c.segwd.p + ((size_t)(k - k0 - 1) * nrk + i) * ne; };
        if (build_stage_tables(c, it0 + k0, n)) return 1;
        HIPCHK(hipMemcpyAsync(stage(k0 + 1, 0), c.ckpt.p + (size_t)sg * ssz, ssz * sizeof(double),
                              hipMemcpyDeviceToDevice, c.stream));
        if (wdrev_on)
            HIPCHK(hipMemcpyAsync(c.wd[c.wdcur].p, c.ckptwd.p + (size_t)sg * ne, ne, hipMemcpyDeviceToDevice, c.stream));
        c.hstep = 0;
        for (int k = k0 + 1; k <= k1; k++) {
            // the last step runs for its intermediate stages, pre-limiter outputs and flags only
            if (k == k1 && !(nrk > 1 || lim)) break;
            double* st[KMAXRK + 1];
            double* ws[KMAXRK];
            uint8_t* fs[KMAXRK];
            for (int i = 0; i < nrk; i++) {
                st[i] = stage(k, i);
                ws[i] = lim ? wstage(k, i) : nullptr;
                fs[i] = wdrev_on ? fstage(k, i) : nullptr;
            }
            st[nrk] = k < k1 ? stage(k + 1, 0) : c.pout.p;   // the last step's output is unused
            if (launch_step_explicit<P>(c, st, lim ? ws : nullptr, wdrev_on ? fs : nullptr)) return 1;
        }
        for (int k = k1; k > k0; k--) {
            c.hstep = k - k0 - 1;
            if (launch_forcing(c, false)) return 1;   // the step's tidal terms for the adjoint stages
            if (c.met) {   // the step's node forcing for the adjoint stages
                hipLaunchKernelGGL(met_kernel, dim3((c.np + 255) / 256), dim3(256), 0, c.stream, met_args(c), nullptr,
                                   c.hstep);
                HIPCHK(hipGetLastError());
            }
            ub[nrk] = lam;
            if (wdrev_on && wdrev(ub[nrk], wstage(k, nrk - 1), fstage(k, nrk - 1))) return 1;
            if (lim && limrev(ub[nrk], wstage(k, nrk - 1))) return 1;
            for (int i = nrk - 1; i >= 0; i--) {
                AdjArgs a{};
                a.ed = elem_data(c); a.ph = c.phys; a.fo = forcing(c);
                if (c.met) { a.mt = met_args(c); a.met = 1; }
                a.u = stage(k, i);
                a.irk = i;
                a.nsrc = nrk - i;
                for (int j = 0; j < a.nsrc; j++) {
                    const int ip = i + j;   // later stage I' whose output U(I'+1) carries a cotangent
                    a.src[j] = ub[ip + 1];
                    a.acoef[j] = c.rk.a[ip][i];
                    a.bcoef[j] = c.rk.b[ip][i];
                }
                a.out = ub[i];
                a.gdp = c.gdp.p;
                a.gnbar = c.phys.load_manning ? c.gnb.p : nullptr;
                a.wd = wdrev_on ? fstage(k, i) : nullptr;
                int rc;
                if constexpr (P == 1) rc = launch_adj<P, 128>(c, a);
                else rc = launch_adj<P, 64>(c, a);
                if (rc) return rc;
                HIPCHK(hipGetLastError());
                // ub[i] is now the whole cotangent of U(i), the output of stage i-1
                if (wdrev_on && i >= 1 && wdrev(ub[i], wstage(k, i - 1), fstage(k, i - 1))) return 1;
                if (lim && i >= 1 && limrev(ub[i], wstage(k, i - 1))) return 1;
            }
            std::swap(lam, ub[0]);   // cotangent of U^(k-1)
            if (k - 1 >= 1) seed(k - 1, lam);
        }
    }
    // keep the buffers' identities for the next call
    if (lam != c.lam.p) {
        for (int i = 0; i < nrk; i++)
            if (ub[i] == c.lam.p) { std::swap(c.lam.p, c.ubar[i].p); break; }
    }
    return 0;
}

int validate(const dg_params* p) {
    if (!p) return fail("null parameters");
    if (p->p < 1 || p->p > 2) return fail("p must be 1 or 2");
    if (!(p->dtdp > 0)) return fail("dtdp must be positive");
    if (p->ics != 1 && p->ics != 2) return fail("ics must be 1 or 2");
    if (p->nramp < 0 || p->nramp > 1) return fail("only NRAMP 0 or 1 is supported");
    if (p->rainfall < 0 || p->rainfall > 1) return fail("only rainfall 0 or 1 is supported");
    if (p->nolibf < 0 || p->nolibf > 2) return fail("nolibf must be 0, 1 or 2");
    // inputs whose terms the step does not have: refused, never run without them
    if (p->nws != 0 && p->nws != 8)
        return fail("NWS=" + std::to_string(p->nws) + ": only NWS=0 and NWS=8 (Holland vortex) are supported");
    if (p->ntip < 0 || p->ntip > 2) return fail("NTIP=" + std::to_string(p->ntip) + ": only NTIP 0, 1 and 2 are supported");
    if (p->artdif != 0) return fail("artdif=" + std::to_string(p->artdif) + ": LDG artificial diffusion is not supported");
    if (p->padapt != 0) return fail("padapt=" + std::to_string(p->padapt) + ": p-adaptivity is not supported");
    if (p->pl != p->p)
        return fail("fort.dg pl=" + std::to_string(p->pl) + " differs from ph=" + std::to_string(p->p) +
                    ": mixed orders are not supported");
    if (p->load_quadratic_fric != 0)
        return fail("a per-node quadratic friction attribute (FRIC) is not supported: CF must be uniform");
    return 0;
}

int register_ctx(std::unique_ptr<Ctx> c, int* handle) {
    std::lock_guard<std::mutex> lk(g_mu);
    g_ctx.push_back(std::move(c));
    *handle = (int)g_ctx.size() - 1;
    return 0;
}

// EMO_DG(NBFR,NEEDS,2)-style column-major (constituent, edge, component) -> [edge][constituent][2]
void unpack_cols(int nc, int nedge, const double* src, std::vector<double>& dst) {
    dst.assign((size_t)nedge * nc * 2, 0.0);
    for (int L = 0; L < nedge; L++)
        for (int j = 0; j < nc; j++)
            for (int k = 0; k < 2; k++)
                dst[((size_t)L * nc + j) * 2 + k] = src[(size_t)j + (size_t)nc * (L + (size_t)nedge * k)];
}

// ---- dg_gpu_mass_noalloc's working memory.  A file-scope device array: every GPU the library runs
// on holds its own copy until the process ends, and every call on that GPU shares it (calls that
// overlap on one GPU race).  Host code never passes the array's name to a kernel or a copy (on the
// host that name is a stand-in, not a device address); it uses the address hipGetSymbolAddress
// returns for the current device, looked up once per device.
constexpr int kScratchMaxBlocks = kScratchDoubles - 1;   // partial sums; the last slot holds the total
constexpr int kMaxDevices = 64;
double* g_scratch[kMaxDevices] = {};

int scratch_ptr(int dev, double** out) {
    if (dev < 0 || dev >= kMaxDevices) return fail("device index out of range");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[dev]) {
        void* ptr = nullptr;
        HIPCHK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        if (!ptr) return fail("hipGetSymbolAddress returned no address for the scratch array");
        g_scratch[dev] = static_cast<double*>(ptr);
    }
    *out = g_scratch[dev];
    return 0;
}

}  // namespace

extern "C" {

const char* dg_gpu_last_error(void) { return g_err.c_str(); }

int dg_create_edge_data(int np, int ne, const int* nm, int nope, const int* nvdll, const int* nbdv,
                        int nbou, const int* ibtype, const int* nvell, const int* nbvv, int ned_cap,
                        int* nedno, int* nedel, int* nedsd, int* niedn, int* nledn, int* needn,
                        int* nfedn, int* counts) {
    EdgeData ed;
    std::string err;
    if (!create_edge_data(np, ne, nm, nope, nvdll, nbdv, nbou, ibtype, nvell, nbvv, ed, err)) return fail(err);
    if (ed.nedges > ned_cap) return fail("edge capacity too small");
    for (int i = 0; i < ed.nedges; i++)
        for (int s = 0; s < 2; s++) {
            nedno[2 * i + s] = ed.nedno[2 * i + s] + 1;
            nedel[2 * i + s] = ed.nedel[2 * i + s] + 1;
            nedsd[2 * i + s] = ed.nedsd[2 * i + s] + 1;
        }
    auto put = [](const std::vector<int>& v, int* o) { for (size_t i = 0; i < v.size(); i++) o[i] = v[i] + 1; };
    put(ed.niedn, niedn); put(ed.nledn, nledn); put(ed.needn, needn); put(ed.nfedn, nfedn);
    counts[0] = ed.nedges; counts[1] = (int)ed.niedn.size(); counts[2] = (int)ed.nledn.size();
    counts[3] = (int)ed.needn.size(); counts[4] = (int)ed.nfedn.size();
    return 0;
}

int dg_gpu_init(const dg_params* prm, int np, int ne, const double* x, const double* y, const double* dp,
                const double* sfac, const double* corif, const double* manning, const int* nm, int nedges,
                const int* nedno, const int* nedel, const int* nedsd, int nieds, const int* niedn,
                int nleds, const int* nledn, int needs, const int* needn, int nfeds, const int* nfedn,
                int* handle) {
    (void)sfac;
    if (validate(prm)) return 1;
    if (np <= 0 || ne <= 0) return fail("empty mesh");
    if ((long long)ne * 4 >= (1LL << 31)) return fail("mesh too large for 32-bit indices");
    auto c = std::make_unique<Ctx>();
    if (init_ctx_common(*c, prm, prm->load_manning && manning)) return 1;
    c->S = 1;
    c->mem.resize(1);
    if (build_member(c->mem[0], prm, c->T, np, ne, x, y, dp, corif, manning, nm, nedges, nedel, nedsd,
                     nieds, niedn, nleds, nledn, needs, needn, nfeds, nfedn))
        return 1;
    {
        // end nodes of the ocean edges in NEDNO order (non-periodic elevation boundaries)
        Member& m = c->mem[0];
        m.onode.resize((size_t)2 * needs);
        for (int L = 0; L < needs; L++) {
            const int ged = needn[L] - 1;
            for (int k = 0; k < 2; k++) {
                const int nd = nedno[2 * ged + k] - 1;
                if (nd < 0 || nd >= np) return fail("ocean edge with a node out of range");
                m.onode[2 * L + k] = nd;
            }
        }
    }
    if (upload_mesh(*c) || upload_initial_state(*c)) return 1;
    return register_ctx(std::move(c), handle);
}

int dg_gpu_set_tide(int h, int nbfr, const double* amig, const double* ff, const double* face,
                    const double* per, const double* emo_dg, const double* efa_dg) {
    Ctx* c = get(h);
    if (!c) return 1;
    if (c->S != 1) return fail("the forcing of an ensemble is set on its members before dg_ens_create");
    Member& m = c->mem[0];
    if (nbfr < 0) return fail("nbfr < 0");
    if (m.needs > 0 && nbfr == 0) return fail("NBFR=0: a non-periodic elevation boundary is set with dg_gpu_set_elev_series");
    m.nbfr = nbfr;
    m.amig.assign(amig, amig + nbfr); m.ff.assign(ff, ff + nbfr);
    m.face.assign(face, face + nbfr); m.per.assign(per, per + nbfr);
    unpack_cols(nbfr, m.needs, emo_dg, m.emo);
    unpack_cols(nbfr, m.needs, efa_dg, m.efa);
    return upload_forcing(*c);
}

int dg_gpu_set_flow(int h, int nffr, const double* famig, const double* fff, const double* fface,
                    const double* fper, const double* qnam_dg, const double* qnph_dg) {
    Ctx* c = get(h);
    if (!c) return 1;
    if (c->S != 1) return fail("the forcing of an ensemble is set on its members before dg_ens_create");
    Member& m = c->mem[0];
    if (nffr < 0) return fail("nffr < 0");
    if (m.nfeds > 0 && nffr == 0) return fail("NFFR=0: a non-periodic normal flux is set with dg_gpu_set_flow_series");
    m.nffr = nffr;
    m.famig.assign(famig, famig + nffr); m.fff.assign(fff, fff + nffr);
    m.fface.assign(fface, fface + nffr); m.fper.assign(fper, fper + nffr);
    unpack_cols(nffr, m.nfeds, qnam_dg, m.qnam);
    unpack_cols(nffr, m.nfeds, qnph_dg, m.qnph);
    return upload_forcing(*c);
}

int dg_gpu_set_tidal_potential(int h, int ntif, const double* tpk, const double* amigt, const double* etrf,
                               const double* fft, const double* facet, const double* saltamp,
                               const double* saltpha, const double* slam, const double* sfea, double reftim) {
    Ctx* c = get(h);
    if (!c) return 1;
    if (c->S != 1) return fail("the forcing of an ensemble is set on its members before dg_ens_create");
    Member& m = c->mem[0];
    if (m.prm.ntip == 0) return fail("NTIP=0: the run was set up without tidal potential");
    if (ntif <= 0) return fail("NTIF must be positive");
    if (m.prm.ntip == 2 && (!saltamp || !saltpha)) return fail("NTIP=2 needs SALTAMP and SALTPHA");
    const double pi = 3.141592653589793;     // PI of src/global.F:72
    m.ntif = ntif;
    m.reftim = reftim;
    m.tpk.assign(tpk, tpk + ntif); m.amigt.assign(amigt, amigt + ntif); m.etrf.assign(etrf, etrf + ntif);
    m.fft.assign(fft, fft + ntif); m.facet.assign(facet, facet + ntif);
    m.pert.resize(ntif); m.tna.resize(ntif);
    for (int j = 0; j < ntif; j++) {
        m.pert[j] = amigt[j] == 0.0 ? 0.0 : 2.0 * pi / amigt[j];               // src/read_input.F:1870-1874
        m.tna[j] = amigt[j] == 0.0 ? 0 : (int)std::lround(0.00014 / amigt[j]); // NINT(0.00014/AMIGT)
    }
    const int np = m.np;
    m.tslam.assign(slam, slam + np);
    m.tccs.resize(np); m.ts2s.resize(np);
    for (int i = 0; i < np; i++) {
        const double cc = std::cos(sfea[i]);
        m.tccs[i] = cc * cc;
        m.ts2s[i] = std::sin(2.0 * sfea[i]);
    }
    m.tsamp.clear(); m.tspha.clear();
    if (m.prm.ntip == 2) {   // SALTAMP(MNTIF,MNP), column-major -> [ntif][np]
        m.tsamp.resize((size_t)ntif * np); m.tspha.resize((size_t)ntif * np);
        for (int i = 0; i < np; i++)
            for (int j = 0; j < ntif; j++) {
                m.tsamp[(size_t)j * np + i] = saltamp[(size_t)j + (size_t)ntif * i];
                m.tspha[(size_t)j * np + i] = saltpha[(size_t)j + (size_t)ntif * i];
            }
    }
    c->drop_graphs();
    HIPCHK(hipStreamSynchronize(c->stream));
    c->ntif = ntif;
    HIPCHK(c->tna.upload(m.tna.data(), ntif, c->stream));
    HIPCHK(c->tslam.upload(m.tslam.data(), np, c->stream));
    HIPCHK(c->tccs.upload(m.tccs.data(), np, c->stream));
    HIPCHK(c->ts2s.upload(m.ts2s.data(), np, c->stream));
    if (m.prm.ntip == 2) {
        HIPCHK(c->tsamp.upload(m.tsamp.data(), m.tsamp.size(), c->stream));
        HIPCHK(c->tspha.upload(m.tspha.data(), m.tspha.size(), c->stream));
    }
    HIPCHK(c->tiprow.alloc((size_t)c->rows * ntif * 3));
    HIPCHK(c->tip2.alloc((size_t)c->rk.nrk * np));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->tip = true;
    return 0;
}

int dg_gpu_set_elev_series(int h, int neta, const int* nbd, double etiminc, int nrec, const double* esbin) {
    Ctx* c = get(h);
    if (!c) return 1;
    if (c->S != 1) return fail("the forcing of an ensemble is set on its members before dg_ens_create");
    Member& m = c->mem[0];
    if (m.needs == 0) return fail("the mesh has no elevation-specified (ocean) edges");
    if (neta <= 0 || nrec < 2) return fail("fort.19 needs NETA > 0 and at least two records");
    if (!(etiminc > 0.0)) return fail("ETIMINC must be positive");
    for (int i = 0; i < neta; i++)
        if (nbd[i] < 1 || nbd[i] > m.np) return fail("NBD refers to a node out of range");
    m.neta = neta;
    m.nbd.resize(neta);
    for (int i = 0; i < neta; i++) m.nbd[i] = nbd[i] - 1;
    m.etiminc = etiminc;
    m.e_nrec = nrec;
    m.esbin.assign(esbin, esbin + (size_t)nrec * neta);
    m.eser = true;
    m.es = Member::Series{};
    // one pseudo-constituent whose end-node values come from the series (tide_kernel copies them)
    m.nbfr = 1;
    m.amig.assign(1, 0.0); m.ff.assign(1, 0.0); m.face.assign(1, 0.0); m.per.assign(1, 0.0);
    m.emo.assign((size_t)m.needs * 2, 0.0); m.efa.assign((size_t)m.needs * 2, 0.0);
    return upload_forcing(*c);
}

int dg_gpu_set_flow_series(int h, int nvel, const int* lbcodei, double ftiminc, int nrec, const double* qnin) {
    Ctx* c = get(h);
    if (!c) return 1;
    if (c->S != 1) return fail("the forcing of an ensemble is set on its members before dg_ens_create");
    Member& m = c->mem[0];
    if (m.nfeds == 0) return fail("the mesh has no flux-specified edges");
    if (nvel <= 0 || nrec < 2) return fail("fort.20 needs NVEL > 0 and at least two records");
    if (!(ftiminc > 0.0)) return fail("FTIMINC must be positive");
    int nflux = 0;
    for (int j = 0; j < nvel; j++) if (lbcodei[j] == 2 || lbcodei[j] == 12 || lbcodei[j] == 22) nflux++;
    if (nflux == 0) return fail("LBCODEI has no flux-specified node");
    m.nvel = nvel;
    m.lbcodei.assign(lbcodei, lbcodei + nvel);
    m.nflux = nflux;
    m.ftiminc = ftiminc;
    m.f_nrec = nrec;
    m.qnin.assign(qnin, qnin + (size_t)nrec * nflux);
    m.fser = true;
    m.fs = Member::Series{};
    // one term with ARGJ = 0 and RFF = RAMPDG, its end values from the series each stage
    m.nffr = 1;
    m.famig.assign(1, 0.0); m.fff.assign(1, 1.0); m.fface.assign(1, 0.0); m.fper.assign(1, 0.0);
    m.qnam.assign((size_t)m.nfeds * 2, 0.0); m.qnph.assign((size_t)m.nfeds * 2, 0.0);
    return upload_forcing(*c);
}

int dg_gpu_set_holland(int h, int nrec, const int* rec, const int* ref, int storm_number, double bladj,
                       const double* slam, const double* sfea) {
    Ctx* c = get(h);
    if (!c) return 1;
    if (c->S != 1) return fail("the forcing of an ensemble is set on its members before dg_ens_create");
    Member& m = c->mem[0];
    std::string err;
    HollandTrack t;
    if (!build_holland_track(nrec, rec, ref, storm_number, bladj, t, err)) return fail(err);
    if (!(bladj > 0.0)) return fail("BLAdj must be positive");
    const double rad2deg = 180.0 / 3.141592653589793;
    const double pi = 4.0 * std::atan(1.0), deg2rad = pi / 180.0, omega = 2.0 * pi / 86164.2;
    m.mxc.resize(m.np); m.myc.resize(m.np); m.mcosy.resize(m.np); m.mcor.resize(m.np);
    for (int i = 0; i < m.np; i++) {
        m.mxc[i] = slam[i] * rad2deg;
        m.myc[i] = sfea[i] * rad2deg;
        m.mcosy[i] = std::cos(m.myc[i] * deg2rad);
        m.mcor[i] = 2.0 * omega * std::sin(m.myc[i] * deg2rad);
    }
    m.track = t;
    m.met = true;
    if (upload_met(*c) || upload_forcing(*c)) return 1;
    return 0;
}

int dg_holland_storm_at(int nrec, const int* rec, const int* ref, int storm_number, double bladj,
                        double g, int ntimes, const double* times, double* rows) {
    std::string err;
    HollandTrack t;
    if (!build_holland_track(nrec, rec, ref, storm_number, bladj, t, err)) return fail(err);
    for (int k = 0; k < ntimes; k++)
        if (!holland_row(t, times[k], g, rows + (size_t)k * MET_ROW, err)) return fail(err);
    return 0;
}

int dg_gpu_set_state(int h, const double* ze, const double* qx, const double* qy, const int* wdflg) {
    Ctx* c = get(h);
    if (!c) return 1;
    const int ne = c->ne, dof = c->dof;
    std::vector<double> buf((size_t)3 * dof * ne);
    const double* src[3] = {ze, qx, qy};
    for (int f = 0; f < 3; f++)
        for (int e = 0; e < ne; e++)
            for (int k = 0; k < dof; k++) buf[((size_t)f * dof + k) * ne + e] = src[f][(size_t)e * dof + k];
    HIPCHK(hipMemcpyAsync(c->u[c->ucur].p, buf.data(), buf.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    std::vector<uint8_t> w;
    if (wdflg) {
        w.resize(ne);
        for (int e = 0; e < ne; e++) w[e] = wdflg[e] ? 1 : 0;
        HIPCHK(hipMemcpyAsync(c->wd[c->wdcur].p, w.data(), ne, hipMemcpyHostToDevice, c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int dg_gpu_get_state(int h, double* ze, double* qx, double* qy, int* wdflg) {
    Ctx* c = get(h);
    if (!c) return 1;
    const int ne = c->ne, dof = c->dof;
    std::vector<double> buf((size_t)3 * dof * ne);
    HIPCHK(hipMemcpyAsync(buf.data(), c->u[c->ucur].p, buf.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    std::vector<uint8_t> w(ne);
    HIPCHK(hipMemcpyAsync(w.data(), c->wd[c->wdcur].p, ne, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    double* dst[3] = {ze, qx, qy};
    for (int f = 0; f < 3; f++) {
        if (!dst[f]) continue;
        for (int e = 0; e < ne; e++)
            for (int k = 0; k < dof; k++) dst[f][(size_t)e * dof + k] = buf[((size_t)f * dof + k) * ne + e];
    }
    if (wdflg) for (int e = 0; e < ne; e++) wdflg[e] = w[e];
    return 0;
}

int dg_gpu_step(int h, int it, int nsteps) {
    Ctx* c = get(h);
    if (!c) return 1;
    if (nsteps <= 0) return 0;
    for (auto& m : c->mem) {
        if (m.needs > 0 && m.nbfr == 0) return fail("tidal forcing not set (dg_gpu_set_tide)");
        if (m.nfeds > 0 && m.nffr == 0) return fail("flux forcing not set (dg_gpu_set_flow)");
        if (m.prm.nws == 8 && !m.met) return fail("NWS=8 forcing not set (dg_gpu_set_holland)");
        if (m.prm.ntip != 0 && m.ntif == 0) return fail("tidal potential not set (dg_gpu_set_tidal_potential)");
    }
    if (ensure_staging(*c, stage_nt(*c, c->block))) return 1;
    for (int s0 = 0; s0 < nsteps; s0 += kTableChunk) {
        const int n = std::min(kTableChunk, nsteps - s0);
        if (build_stage_tables(*c, it + s0, n)) return 1;
        if (run_chunk(*c, n)) return 1;
    }
    c->next_it = it + nsteps;
    if (c->prof) return prof_flush(*c);
    return 0;
}

int dg_gpu_synchronize(int h) {
    Ctx* c = get(h);
    if (!c) return 1;
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int dg_gpu_nodal_update(int h) {
    Ctx* c = get(h);
    if (!c) return 1;
    const int np = c->np;
    ElemData E = elem_data(*c);
    const int B = 256, gn = (np + B - 1) / B;
    const int* off = c->m2n_set ? c->m2n_off.p : c->n2e_off.p;
    const int* row = c->m2n_set ? c->m2n.p : c->n2e.p;
    double* em = c->has_etamax ? c->etamax.p : nullptr;
    int ev = prof_begin(*c, K_M2N, false);
    if (c->prm.p == 1)
        hipLaunchKernelGGL((modal2nodal_kernel<1>), dim3(gn), dim3(B), 0, c->stream, np, E, off, row,
                           (const double*)c->u[c->ucur].p, (const uint8_t*)c->wd[c->wdcur].p,
                           c->phys.nleq, c->phys.leq, c->outbuf.p, c->nodecode.p, em);
    else
        hipLaunchKernelGGL((modal2nodal_kernel<2>), dim3(gn), dim3(B), 0, c->stream, np, E, off, row,
                           (const double*)c->u[c->ucur].p, (const uint8_t*)c->wd[c->wdcur].p,
                           c->phys.nleq, c->phys.leq, c->outbuf.p, c->nodecode.p, em);
    prof_end(*c, ev);
    HIPCHK(hipGetLastError());
    return 0;
}

int dg_gpu_get_nodal(int h, double* eta2, double* uu2, double* vv2, int* nodecode, double* etamax) {
    Ctx* c = get(h);
    if (!c) return 1;
    const size_t np = (size_t)c->np;
    if (etamax && !c->has_etamax) return fail("ETAMAX was not set (dg_gpu_set_etamax)");
    double* dst[3] = {eta2, uu2, vv2};
    for (int k = 0; k < 3; k++)
        if (dst[k]) HIPCHK(hipMemcpyAsync(dst[k], c->outbuf.p + k * np, np * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (nodecode) HIPCHK(hipMemcpyAsync(nodecode, c->nodecode.p, np * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    if (etamax) HIPCHK(hipMemcpyAsync(etamax, c->etamax.p, np * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->prof && prof_flush(*c)) return 1;
    return 0;
}

int dg_gpu_modal2nodal(int h, double* eta2, double* uu2, double* vv2, int* nodecode) {
    if (dg_gpu_nodal_update(h)) return 1;
    return dg_gpu_get_nodal(h, eta2, uu2, vv2, nodecode, nullptr);
}

int dg_gpu_set_etamax(int h, const double* etamax) {
    Ctx* c = get(h);
    if (!c) return 1;
    if (!etamax) return fail("null ETAMAX");
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(c->etamax.upload(etamax, (size_t)c->np, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->has_etamax = true;
    return 0;
}

int dg_gpu_set_eletab(int h, int mnp, const int* el_count, const int* eletab) {
    Ctx* c = get(h);
    if (!c) return 1;
    if (c->S != 1) return fail("dg_gpu_set_eletab applies to single runs");
    const Member& m = c->mem[0];
    const int np = m.np, ne = m.ne;
    if (mnp < np) return fail("ELETAB has fewer rows than nodes");
    std::vector<int> off(np + 1, 0), row;
    row.reserve(m.n2e.size());
    for (int i = 0; i < np; i++) {
        const int cnt = el_count[i];
        if (cnt != m.n2e_off[i + 1] - m.n2e_off[i])
            return fail("ELETAB: node " + std::to_string(i + 1) + " lists " + std::to_string(cnt) +
                        " elements, the mesh has " + std::to_string(m.n2e_off[i + 1] - m.n2e_off[i]));
        for (int j = 0; j < cnt; j++) {
            const int e = eletab[(size_t)(1 + j) * mnp + i] - 1;   // ELETAB(I, 1+J)
            if (e < 0 || e >= ne) return fail("ELETAB: element index out of range");
            int k = 0;
            while (k < 3 && m.nm[(size_t)k * ne + e] != i) k++;
            if (k == 3) return fail("ELETAB: node " + std::to_string(i + 1) + " is not a corner of element " + std::to_string(e + 1));
            for (int q = off[i]; q < (int)row.size(); q++)
                if ((row[q] >> 2) == e) return fail("ELETAB: element listed twice for node " + std::to_string(i + 1));
            row.push_back(e << 2 | k);
        }
        off[i + 1] = (int)row.size();
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(c->m2n_off.upload(off.data(), off.size(), c->stream));
    HIPCHK(c->m2n.upload(row.data(), row.size(), c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->m2n_set = true;
    return 0;
}

int dg_gpu_mass(int h, double* mass) {
    Ctx* c = get(h);
    if (!c) return 1;
    ElemData E = elem_data(*c);
    const int nb = (c->ne + 255) / 256;
    if (c->prm.p == 1)
        hipLaunchKernelGGL((mass_kernel<1>), dim3(nb), dim3(256), 0, c->stream, E, (const double*)c->u[c->ucur].p, c->partial.p);
    else
        hipLaunchKernelGGL((mass_kernel<2>), dim3(nb), dim3(256), 0, c->stream, E, (const double*)c->u[c->ucur].p, c->partial.p);
    HIPCHK(hipGetLastError());
    std::vector<double> part(nb);
    HIPCHK(hipMemcpyAsync(part.data(), c->partial.p, nb * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    double m = 0.0;
    for (double v : part) m += v;
    *mass = m;
    return 0;
}

int dg_gpu_mass_noalloc(int h, double* mass) {
    Ctx* c = get(h);
    if (!c) return 1;
    if (!mass) return fail("null output pointer");
    const int nb = (c->ne + 255) / 256;
    if (nb > kScratchMaxBlocks)
        return fail("dg_gpu_mass_noalloc: " + std::to_string(c->ne) + " elements exceed its limit of " +
                    std::to_string((long long)kScratchMaxBlocks * 256));
    int dev = -1;
    HIPCHK(hipGetDevice(&dev));
    if (dev != c->dev) return fail("dg_gpu_mass_noalloc: the handle's device is not the current device");
    double* ws = nullptr;
    if (scratch_ptr(dev, &ws)) return 1;
    ElemData E = elem_data(*c);
    if (c->prm.p == 1)
        hipLaunchKernelGGL((mass_kernel<1>), dim3(nb), dim3(256), 0, c->stream, E, (const double*)c->u[c->ucur].p, ws);
    else
        hipLaunchKernelGGL((mass_kernel<2>), dim3(nb), dim3(256), 0, c->stream, E, (const double*)c->u[c->ucur].p, ws);
    HIPCHK(hipGetLastError());
    // the partial sums added in block order, as dg_gpu_mass adds them on the host
    hipLaunchKernelGGL(sum_in_order_kernel, dim3(1), dim3(64), 0, c->stream, (const double*)ws, nb,
                       ws + kScratchMaxBlocks);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(mass, ws + kScratchMaxBlocks, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int dg_gpu_profile(int h, int enable) {
    Ctx* c = get(h);
    if (!c) return 1;
    if (prof_flush(*c)) return 1;
    c->prof = enable > 0;
    c->prof_every = std::max(1, enable);
    c->prof_step = 0;
    c->prof_this = true;
    for (int k = 0; k < K_COUNT; k++) { c->ktime[k] = 0.0; c->kcount[k] = 0; }
    return 0;
}

int dg_gpu_kernel_stats(int h, int kid, double* total_ms, long long* launches, char* name, int name_len) {
    Ctx* c = get(h);
    if (!c) return 1;
    if (kid < 0 || kid >= K_COUNT) return fail("kernel id out of range");
    if (prof_flush(*c)) return 1;
    if (total_ms) *total_ms = c->ktime[kid];
    if (launches) *launches = c->kcount[kid];
    if (name && name_len > 0) { std::strncpy(name, kKernelNames[kid], name_len - 1); name[name_len - 1] = 0; }
    return 0;
}

int dg_gpu_set_manning(int h, const double* manning) {
    Ctx* c = get(h);
    if (!c) return 1;
    if (!c->phys.load_manning) return fail("the run was set up without Manning's n");
    // per node of the run (of all members, in member order, for an ensemble)
    std::vector<double> nb((size_t)c->ne);
    for (int s = 0; s < c->S; s++) {
        Member& m = c->mem[s];
        const int ne = m.ne;
        const double* mn = manning + c->noff[s];
        for (int e = 0; e < ne; e++) {
            const int n0 = m.nm[e], n1 = m.nm[(size_t)ne + e], n2 = m.nm[2 * (size_t)ne + e];
            m.nbar[e] = (mn[n0] + mn[n1] + mn[n2]) / 3.0;
            nb[(size_t)c->eoff[s] + e] = m.nbar[e];
        }
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(c->nbar.upload(nb.data(), nb.size(), c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int dg_gpu_set_depth(int h, const double* dp) {
    Ctx* c = get(h);
    if (!c) return 1;
    // per node of the run (of all members, in member order, for an ensemble)
    const size_t NE = (size_t)c->ne;
    std::vector<double> dv(3 * NE);
    for (int s = 0; s < c->S; s++) {
        Member& m = c->mem[s];
        const int ne = m.ne;
        const double* md = dp + c->noff[s];
        for (int v = 0; v < 3; v++)
            for (int e = 0; e < ne; e++) {
                m.dpv[(size_t)v * ne + e] = md[m.nm[(size_t)v * ne + e]];
                dv[(size_t)v * NE + c->eoff[s] + e] = m.dpv[(size_t)v * ne + e];
            }
    }
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(c->dpv.upload(dv.data(), dv.size(), c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int dg_adj_set_gauges(int h, int ng, const int* elem, const double* r, const double* s) {
    Ctx* c = get(h);
    if (!c) return 1;
    if (ng <= 0) return fail("ng must be positive");
    const int dof = c->dof;
    std::vector<int> ge(ng);
    std::vector<double> ph((size_t)ng * dof);
    for (int g = 0; g < ng; g++) {
        if (elem[g] < 1 || elem[g] > c->ne) return fail("gauge element out of range");
        if (!(r[g] >= -1.0 && s[g] >= -1.0 && r[g] + s[g] <= 0.0)) return fail("gauge outside its element");
        ge[g] = elem[g] - 1;
        basis_at(c->prm.p, r[g], s[g], &ph[(size_t)g * dof]);
    }
    // gauges grouped by element, ascending gauge index within a group
    std::vector<int> ord(ng), off;
    for (int g = 0; g < ng; g++) ord[g] = g;
    std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return ge[a] < ge[b]; });
    for (int j = 0; j < ng; j++)
        if (j == 0 || ge[ord[j]] != ge[ord[j - 1]]) off.push_back(j);
    off.push_back(ng);
    c->ng = ng;
    c->ngu = (int)off.size() - 1;
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(c->gel.upload(ge.data(), ng, c->stream));
    HIPCHK(c->gord.upload(ord.data(), ng, c->stream));
    HIPCHK(c->goff.upload(off.data(), off.size(), c->stream));
    HIPCHK(c->gphi.upload(ph.data(), ph.size(), c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int dg_adj_gauges(int h, int it, int nsteps, double* eta) {
    Ctx* c = get(h);
    if (!c) return 1;
    if (c->ng == 0) return fail("no gauges set (dg_adj_set_gauges)");
    if (nsteps <= 0) return 0;
    if (ensure_staging(*c, stage_nt(*c, c->block))) return 1;
    HIPCHK(c->geta.alloc((size_t)nsteps * c->ng));
    const int rc = c->prm.p == 1 ? adj_forward<1>(*c, it, nsteps, 1, false) : adj_forward<2>(*c, it, nsteps, 1, false);
    if (rc) return rc;
    c->next_it = it + nsteps;
    HIPCHK(hipMemcpyAsync(eta, c->geta.p, (size_t)nsteps * c->ng * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int dg_adj_gradient(int h, int it, int nsteps, const double* data, double weight, int interval, double* loss,
                    double* grad_manning, double* grad_dp) {
    Ctx* c = get(h);
    if (!c) return 1;
    if (adj_check(*c)) return 1;
    if (nsteps <= 0) return fail("nsteps must be positive");
    if (ensure_staging(*c, stage_nt(*c, c->block)) || ensure_adj_staging(*c)) return 1;
    if (interval <= 0) interval = std::max(1, (int)std::lround(std::sqrt((double)nsteps)));
    interval = std::min(interval, nsteps);
    const int nrk = c->rk.nrk, ne = c->ne, np = c->np;
    const size_t ssz = (size_t)3 * c->dof * ne;
    const int ncp = (nsteps + interval - 1) / interval;
    HIPCHK(c->geta.alloc((size_t)nsteps * c->ng));
    HIPCHK(c->gdata.upload(data, (size_t)nsteps * c->ng, c->stream));
    HIPCHK(c->ckpt.alloc((size_t)ncp * ssz));
    HIPCHK(c->seg.alloc((size_t)interval * nrk * ssz));
    HIPCHK(c->lam.alloc(ssz));
    for (int i = 0; i < nrk; i++) HIPCHK(c->ubar[i].alloc(ssz));
    HIPCHK(c->gdp.alloc((size_t)3 * ne));
    HIPCHK(c->gnb.alloc(ne));
    HIPCHK(c->pout.alloc(std::max(ssz, (size_t)2 * np)));
    if (c->prm.slopeflag != 0) {
        HIPCHK(c->wseg.alloc((size_t)interval * nrk * ssz));
        HIPCHK(c->lnbc.alloc((size_t)18 * ne));
        HIPCHK(c->lgsum.alloc((size_t)6 * np));
        HIPCHK(c->larg.alloc((size_t)6 * np));
    }
    // wetting/drying after the slope limiter: differentiated with its branches frozen (flags kept
    // per checkpoint and per stage)
    c->adj_wd = c->prm.nolifa >= 2 && c->prm.slopeflag != 0;
    c->adj_wdrev = false;
    if (c->adj_wd) HIPCHK(c->ckptwd.alloc((size_t)ncp * ne));
    // the state and wet/dry flags at entry: every exit path puts them back (the sweeps rotate the
    // stage slots and flip the flag buffers)
    HIPCHK(c->wdsave.alloc(ne));
    HIPCHK(hipMemcpyAsync(c->ckpt.p, c->u[c->ucur].p, ssz * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
    HIPCHK(hipMemcpyAsync(c->wdsave.p, c->wd[c->wdcur].p, ne, hipMemcpyDeviceToDevice, c->stream));
    auto restore = [&]() -> int {
        HIPCHK(hipMemcpyAsync(c->u[c->ucur].p, c->ckpt.p, ssz * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(c->wd[c->wdcur].p, c->wdsave.p, ne, hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        return 0;
    };
    // The forward sweeps flag any element WETDRY changes (wets, dries, redistributes or keeps
    // flat).  Where WETDRY leaves every element as it is, it is the identity and the reverse skips
    // it.  Without the slope limiter WETDRY runs fused into the stage kernel and has no reverse:
    // runs it changes are refused.  With the limiter a run it changes is reversed with WETDRY's
    // branches frozen (the flags at the start of every stage kept per segment).
    HIPCHK(c->wdact.alloc(1));
    HIPCHK(hipMemsetAsync(c->wdact.p, 0, sizeof(int), c->stream));
    c->wd_active = c->prm.nolifa >= 2 ? c->wdact.p : nullptr;
    const bool prof = c->prof;
    c->prof = false;
    int rc = c->prm.p == 1 ? adj_forward<1>(*c, it, nsteps, interval, true) : adj_forward<2>(*c, it, nsteps, interval, true);
    if (!rc && c->adj_wd) {
        int changed = 0;
        if (hipMemcpyAsync(&changed, c->wdact.p, sizeof(int), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess)
            rc = fail("reading the wet/dry activity flag failed");
        else if (changed) {
            c->adj_wdrev = true;
            c->wd_active = nullptr;
            if (c->segwd.alloc((size_t)interval * nrk * ne) != hipSuccess)
                rc = fail("out of device memory for the wet/dry reverse");
        }
    }
    if (!rc) rc = c->prm.p == 1 ? adj_backward<1>(*c, it, nsteps, interval, weight) : adj_backward<2>(*c, it, nsteps, interval, weight);
    c->prof = prof;
    c->wd_active = nullptr;
    if (rc) {
        const std::string msg = g_err;
        (void)restore();
        return fail(msg);
    }
    int wdflag = 0;
    HIPCHK(hipMemcpyAsync(&wdflag, c->wdact.p, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (wdflag && !c->adj_wd) {
        if (restore()) return 1;
        return fail("wetting/drying changed elements during the run: without the slope limiter the adjoint "
                    "differentiates WETDRY only where it leaves every element as it is");
    }
    // node gradients
    double* gdp_n = c->pout.p;
    double* gn_n = c->pout.p + np;
    if (!c->phys.load_manning) HIPCHK(hipMemsetAsync(c->gnb.p, 0, (size_t)ne * sizeof(double), c->stream));
    hipLaunchKernelGGL(param_gather_kernel, dim3((np + 255) / 256), dim3(256), 0, c->stream, np, ne,
                       (const int*)c->n2e_off.p, (const int*)c->n2e.p, (const double*)c->gdp.p,
                       (const double*)c->gnb.p, gdp_n, gn_n);
    HIPCHK(hipGetLastError());
    std::vector<double> eta((size_t)nsteps * c->ng);
    HIPCHK(hipMemcpyAsync(eta.data(), c->geta.p, eta.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (grad_dp) HIPCHK(hipMemcpyAsync(grad_dp, gdp_n, (size_t)np * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (grad_manning) HIPCHK(hipMemcpyAsync(grad_manning, gn_n, (size_t)np * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    // the handle's state returns to where the call started
    if (restore()) return 1;
    if (loss) {
        double J = 0.0;
        for (size_t k = 0; k < eta.size(); k++) {
            const double d = eta[k] - data[k];
            J += d * d;
        }
        *loss = weight * J;
    }
    return 0;
}

int dg_ens_create(int nmembers, const int* handles, int* ens_handle) {
    if (nmembers < 1 || nmembers > 65535) return fail("an ensemble has 1 to 65535 members");
    std::vector<Ctx*> src(nmembers);
    for (int i = 0; i < nmembers; i++) {
        src[i] = get(handles[i]);
        if (!src[i]) return 1;
        if (src[i]->S != 1) return fail("ensemble members must be single runs");
    }
    const dg_params& p0 = src[0]->prm;
    for (int i = 1; i < nmembers; i++) {
        const dg_params& q = src[i]->prm;
        if (q.p != p0.p || q.rk_stage != p0.rk_stage || q.rk_order != p0.rk_order || q.ics != p0.ics ||
            q.nolibf != p0.nolibf || q.nolifa != p0.nolifa || q.nolicat != p0.nolicat ||
            q.slopeflag != p0.slopeflag || q.g != p0.g || q.hbreak != p0.hbreak || q.ftheta != p0.ftheta ||
            q.fgamma != p0.fgamma || src[i]->phys.load_manning != src[0]->phys.load_manning)
            return fail("ensemble members must share p, the RK scheme, ICS, NOLIBF, NOLIFA, NOLICAT, the "
                        "slope limiter, G, the hybrid friction constants and whether Manning n is loaded");
    }
    for (int i = 0; i < nmembers; i++)
        for (const Member& m : src[i]->mem) {
            if (m.ntif > 0 || m.eser || m.fser)
                return fail("ensemble members with tidal potential or non-periodic boundary forcing are not supported");
            if (m.needs > 0 && m.nbfr == 0) return fail("member without tidal forcing (dg_gpu_set_tide)");
            if (m.nfeds > 0 && m.nffr == 0) return fail("member without flux forcing (dg_gpu_set_flow)");
        }
    auto c = std::make_unique<Ctx>();
    if (init_ctx_common(*c, &p0, src[0]->phys.load_manning != 0)) return 1;
    c->S = nmembers;
    c->mem.resize(nmembers);
    for (int i = 0; i < nmembers; i++) c->mem[i] = src[i]->mem[0];
    if (upload_mesh(*c)) return 1;
    // each member's current state into the batch
    const int dof = c->dof;
    for (int i = 0; i < nmembers; i++) {
        Ctx& s = *src[i];
        HIPCHK(hipStreamSynchronize(s.stream));
        for (int fk = 0; fk < 3 * dof; fk++)
            HIPCHK(hipMemcpyAsync(c->u[0].p + (size_t)fk * c->ne + c->eoff[i], s.u[s.ucur].p + (size_t)fk * s.ne,
                                  (size_t)s.ne * sizeof(double), hipMemcpyDeviceToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(c->wd[0].p + c->eoff[i], s.wd[s.wdcur].p, s.ne, hipMemcpyDeviceToDevice, c->stream));
    }
    c->ucur = 0; c->wdcur = 0;
    HIPCHK(hipStreamSynchronize(c->stream));
    return register_ctx(std::move(c), ens_handle);
}

int dg_ens_layout(int h, int* nmembers, int* elem_offsets, int* node_offsets) {
    Ctx* c = get(h);
    if (!c) return 1;
    if (nmembers) *nmembers = c->S;
    for (int i = 0; i <= c->S; i++) {
        if (elem_offsets) elem_offsets[i] = c->eoff[i];
        if (node_offsets) node_offsets[i] = c->noff[i];
    }
    return 0;
}

int dg_gpu_use_graph(int h, int enable) {
    Ctx* c = get(h);
    if (!c) return 1;
    c->use_graph = enable != 0;
    if (!c->use_graph) {
        HIPCHK(hipStreamSynchronize(c->stream));
        c->drop_graphs();
    }
    return 0;
}

int dg_gpu_autotune(int h, int* block) {
    Ctx* c = get(h);
    if (!c) return 1;
    // time the first-stage kernel at each block size; it reads the stage-1 slot and writes only
    // the stage-2 slot and the spare wet/dry buffer, both rewritten by the next step
    const int nrk = c->rk.nrk;
    int slot[KMAXRK + 1];
    for (int i = 0; i <= nrk; i++) slot[i] = (c->ucur + i) % (nrk + 1);
    c->hstep = 0;
    StageArgs a = stage_args(*c, 0, slot);
    a.rout = nullptr;
    const bool sph = c->prm.ics == 2, fuse = c->prm.nolifa >= 2 && c->prm.slopeflag == 0;
    const std::vector<int> cand = c->prm.p == 1 ? std::vector<int>{64, 128, 256} : std::vector<int>{64, 128};
    if (build_stage_tables(*c, c->next_it, 1)) return 1;
    if (launch_forcing(*c, false)) return 1;
    if (c->met) {
        hipLaunchKernelGGL(met_kernel, dim3((c->np + 255) / 256), dim3(256), 0, c->stream, met_args(*c),
                           nullptr, 0);
        HIPCHK(hipGetLastError());
    }
    hipEvent_t e0, e1;
    HIPCHK(hipEventCreate(&e0));
    if (hipEventCreate(&e1) != hipSuccess) { (void)hipEventDestroy(e0); return fail("hipEventCreate failed"); }
    float best = 1e30f;
    int best_nt = c->block, rc = 0;
    for (int nt : cand) {
        if (ensure_staging(*c, stage_nt(*c, nt))) continue;   // does not fit: not a candidate
        float ms = 0.f;
        for (int rep = 0; rep < 2 && !rc; rep++) {   // the first pass warms caches and code
            if (hipEventRecord(e0, c->stream) != hipSuccess) { rc = fail("hipEventRecord failed"); break; }
            for (int k = 0; k < 3 && !rc; k++) {
                if (c->prm.p == 1) rc = launch_stage_nt<1>(*c, a, sph, fuse, nt);
                else rc = launch_stage_nt<2>(*c, a, sph, fuse, nt);
            }
            if (rc) break;
            hipError_t e = hipGetLastError();
            if (e == hipSuccess) e = hipEventRecord(e1, c->stream);
            if (e == hipSuccess) e = hipEventSynchronize(e1);
            if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
            if (e != hipSuccess) rc = fail(std::string("autotune: ") + hipGetErrorString(e));
        }
        if (rc) break;
        if (ms < best) { best = ms; best_nt = nt; }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    if (best_nt != c->block) {
        c->block = best_nt;
        c->drop_graphs();
    }
    if (block) *block = best_nt;
    return 0;
}

int dg_gpu_set_block(int h, int block) {
    Ctx* c = get(h);
    if (!c) return 1;
    if (!valid_block(c->prm.p, block))
        return fail("block size must be 64, 128 or, for p=1, 256 threads");
    HIPCHK(hipStreamSynchronize(c->stream));
    if (ensure_staging(*c, stage_nt(*c, block))) return 1;
    if (block != c->block) { c->block = block; c->drop_graphs(); }
    return 0;
}

int dg_gpu_finalize(int h) {
    std::unique_ptr<Ctx> c;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (h < 0 || h >= (int)g_ctx.size() || !g_ctx[h]) { g_err = "invalid handle"; return 1; }
        c = std::move(g_ctx[h]);
    }
    return 0;   // the context's destructor waits for its stream and frees its memory
}

}  // extern "C"
