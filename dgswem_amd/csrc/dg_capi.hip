// C ABI of libdgswem_hip.so (declared in include/dgswem_gpu.h): device context management, the
// per-step launch sequence and host<->device layout conversion.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/dgswem_gpu.h"
#include "dg_device.hpp"
#include "dg_setup.hpp"

// the kernels are compiled in this translation unit
#include "dg_kernels.hip"

using namespace dg;

namespace {

thread_local std::string g_err;

int fail(const std::string& m) { g_err = m; return 1; }

#define HIPCHK(x)                                                                                  \
    do {                                                                                           \
        hipError_t e_ = (x);                                                                       \
        if (e_ != hipSuccess) return fail(std::string(#x) + ": " + hipGetErrorString(e_));          \
    } while (0)

enum { K_STAGE = 0, K_MINMAX = 1, K_LIMIT = 2, K_M2N = 3, K_COUNT = 4 };
const char* kKernelNames[K_COUNT] = {"stage_kernel", "node_minmax_kernel", "limit_wetdry_kernel",
                                     "modal2nodal_kernel"};

template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0, cap = 0;
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

struct Ctx {
    dg_params prm{};
    int np = 0, ne = 0, dof = 0, nagp = 0, negp = 0;
    RefTables T;
    RKScheme rk;
    Phys phys{};
    hipStream_t stream = nullptr;
    // state
    DevBuf<double> u[KMAXRK + 1];
    int ucur = 0;                      // slot holding the stage-1 state
    DevBuf<double> r[KMAXRK];
    bool keep_r[KMAXRK] = {};
    DevBuf<uint8_t> wd[2];
    int wdcur = 0;
    // mesh
    DevBuf<double> xv, yv, dpv, nbar, cori_el, sfacq, sfaced;
    DevBuf<int> sinfo, snbr, nm, n2e_off, n2e;
    DevBuf<double> nmm;
    DevBuf<double> outbuf;
    DevBuf<int> nodecode;
    DevBuf<double> partial;
    // forcing
    int needs = 0, nfeds = 0, nbfr = 0, nffr = 0;
    std::vector<double> amig, ff, face, per, famig, fff, fface, fper;
    DevBuf<double> emo, efa, qnam, qnph;
    DevBuf<double> argj, rff, fargj, frff;
    std::vector<double> h_argj, h_rff, h_fargj, h_frff;
    // profiling
    bool prof = false;
    double ktime[K_COUNT] = {};
    long long kcount[K_COUNT] = {};
    std::vector<std::pair<hipEvent_t, hipEvent_t>> evpool;
    std::vector<std::pair<int, int>> pending;   // (kernel id, event pair index)
    ~Ctx() {
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
    return d;
}

// ---- profiling helpers
int prof_begin(Ctx& c, int kid) {
    if (!c.prof) return -1;
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
    E.xv = c.xv.p; E.yv = c.yv.p; E.dpv = c.dpv.p;
    E.nbar = c.prm.load_manning ? c.nbar.p : nullptr;
    E.cori_el = c.cori_el.n ? c.cori_el.p : nullptr;
    E.sfacq = c.prm.ics == 2 ? c.sfacq.p : nullptr;
    E.sfaced = c.prm.ics == 2 ? c.sfaced.p : nullptr;
    E.sinfo = c.sinfo.p; E.snbr = c.snbr.p; E.nm = c.nm.p;
    return E;
}

// Per-stage boundary forcing scalars for steps it..it+n-1 (host arithmetic of
// src/DG_hydro_timestep.F:79-100, src/ocean_edge_hydro.F:59-75, src/flow_edge_hydro.F:45-62).
int build_stage_tables(Ctx& c, int it, int n) {
    const int ns = n * c.rk.nrk;
    c.h_argj.assign((size_t)ns * std::max(c.nbfr, 1), 0.0);
    c.h_rff.assign((size_t)ns * std::max(c.nbfr, 1), 0.0);
    c.h_fargj.assign((size_t)ns * std::max(c.nffr, 1), 0.0);
    c.h_frff.assign((size_t)ns * std::max(c.nffr, 1), 0.0);
    for (int s = 0; s < n; s++) {
        const int IT = it + s;
        const double itdt = IT * c.prm.dtdp;
        const double time_a = itdt + c.prm.statim * 86400.0;
        for (int irk = 0; irk < c.rk.nrk; irk++) {
            const int row = s * c.rk.nrk + irk;
            const double timedg = time_a - c.prm.dtdp + c.rk.dtvd[irk] * c.prm.dtdp;
            double ramp = 1.0;
            if (c.prm.nramp == 1) ramp = std::tanh((2.0 * ((IT - 1) + c.rk.dtvd[irk]) * c.prm.dtdp / 86400.0) / c.prm.dramp);
            for (int jj = 0; jj < c.nbfr; jj++) {
                const int ncyc = (c.per[jj] == 0.0) ? 0 : (int)(timedg / c.per[jj]);
                c.h_argj[(size_t)row * c.nbfr + jj] = c.amig[jj] * (timedg - ncyc * c.per[jj]) + c.face[jj];
                c.h_rff[(size_t)row * c.nbfr + jj] = c.ff[jj] * ramp;
            }
            for (int jj = 0; jj < c.nffr; jj++) {
                const int ncyc = (c.fper[jj] == 0.0) ? 0 : (int)(timedg / c.fper[jj]);
                c.h_fargj[(size_t)row * c.nffr + jj] = c.famig[jj] * (timedg - ncyc * c.fper[jj]) + c.fface[jj];
                c.h_frff[(size_t)row * c.nffr + jj] = c.fff[jj] * ramp;
            }
        }
    }
    HIPCHK(c.argj.upload(c.h_argj.data(), c.h_argj.size(), c.stream));
    HIPCHK(c.rff.upload(c.h_rff.data(), c.h_rff.size(), c.stream));
    HIPCHK(c.fargj.upload(c.h_fargj.data(), c.h_fargj.size(), c.stream));
    HIPCHK(c.frff.upload(c.h_frff.data(), c.h_frff.size(), c.stream));
    return 0;
}

template <int P>
int launch_step(Ctx& c, int row0) {
    const int ne = c.ne, np = c.np;
    const int B = 256, BS = StageBlock<P>::N;
    const int ge = (ne + B - 1) / B, gn = (np + B - 1) / B, gs = (ne + BS - 1) / BS;
    const bool wdon = c.prm.nolifa >= 2;
    const bool lim = c.prm.slopeflag != 0;
    const bool sph = c.prm.ics == 2;
    ElemData E = elem_data(c);
    Forcing F{};
    F.nbfr = c.nbfr; F.nffr = c.nffr;
    F.emo = c.emo.p; F.efa = c.efa.p; F.qnam = c.qnam.p; F.qnph = c.qnph.p;
    F.argj = c.argj.p; F.rff = c.rff.p; F.fargj = c.fargj.p; F.frff = c.frff.p;
    const int nrk = c.rk.nrk;
    // stage slots: slot(i) holds U(i+1); slot(0) = current stage-1 state
    int slot[KMAXRK + 1];
    for (int i = 0; i <= nrk; i++) slot[i] = (c.ucur + i) % (nrk + 1);
    for (int irk = 0; irk < nrk; irk++) {
        StageArgs a{};
        a.ed = E; a.ph = c.phys; a.fo = F;
        a.uin = c.u[slot[irk]].p;
        a.uout = c.u[slot[irk + 1]].p;
        a.wd_in = c.wd[c.wdcur].p;
        a.wd_out = c.wd[1 - c.wdcur].p;
        a.nprev = irk;
        for (int i = 0; i < irk; i++) {
            a.uprev[i] = c.u[slot[i]].p; a.aprev[i] = c.rk.a[irk][i];
            a.rprev[i] = c.r[i].p; a.bprev[i] = c.rk.b[irk][i] * c.prm.dtdp;
            if (a.bprev[i] != 0.0 && !c.keep_r[i]) return fail("internal: RHS of an earlier stage was not kept");
        }
        a.acur = c.rk.a[irk][irk];
        a.bcur = c.rk.b[irk][irk] * c.prm.dtdp;
        a.rout = c.keep_r[irk] ? c.r[irk].p : nullptr;
        a.stage = row0 + irk;
        a.rain = c.prm.rainfall == 1 ? 7.0556e-6 : 0.0;
        const bool fuse = wdon && !lim;
        int ev = prof_begin(c, K_STAGE);
        if (sph) {
            if (fuse) hipLaunchKernelGGL((stage_kernel<P, true, true>), dim3(gs), dim3(BS), 0, c.stream, a);
            else hipLaunchKernelGGL((stage_kernel<P, true, false>), dim3(gs), dim3(BS), 0, c.stream, a);
        } else {
            if (fuse) hipLaunchKernelGGL((stage_kernel<P, false, true>), dim3(gs), dim3(BS), 0, c.stream, a);
            else hipLaunchKernelGGL((stage_kernel<P, false, false>), dim3(gs), dim3(BS), 0, c.stream, a);
        }
        prof_end(c, ev);
        HIPCHK(hipGetLastError());
        if (fuse) c.wdcur = 1 - c.wdcur;
        if (lim) {
            ev = prof_begin(c, K_MINMAX);
            hipLaunchKernelGGL((node_minmax_kernel<P>), dim3(gn), dim3(B), 0, c.stream, np, ne,
                               (const int*)c.n2e_off.p, (const int*)c.n2e.p, (const double*)a.uout, c.nmm.p);
            prof_end(c, ev);
            HIPCHK(hipGetLastError());
            ev = prof_begin(c, K_LIMIT);
            if (wdon) {
                hipLaunchKernelGGL((limit_wetdry_kernel<P, true, true>), dim3(ge), dim3(B), 0, c.stream, np, E,
                                   a.uout, (const double*)c.nmm.p, (const uint8_t*)c.wd[c.wdcur].p,
                                   c.wd[1 - c.wdcur].p, c.prm.h0);
                c.wdcur = 1 - c.wdcur;
            } else {
                hipLaunchKernelGGL((limit_wetdry_kernel<P, true, false>), dim3(ge), dim3(B), 0, c.stream, np, E,
                                   a.uout, (const double*)c.nmm.p, (const uint8_t*)c.wd[c.wdcur].p,
                                   c.wd[1 - c.wdcur].p, c.prm.h0);
            }
            prof_end(c, ev);
            HIPCHK(hipGetLastError());
        }
    }
    c.ucur = slot[nrk];
    return 0;
}

int validate(const dg_params* p) {
    if (!p) return fail("null parameters");
    if (p->p < 1 || p->p > 2) return fail("p must be 1 or 2");
    if (p->dtdp <= 0) return fail("dtdp must be positive");
    if (p->ics != 1 && p->ics != 2) return fail("ics must be 1 or 2");
    if (p->nramp < 0 || p->nramp > 1) return fail("only NRAMP 0 or 1 is supported");
    if (p->rainfall < 0 || p->rainfall > 1) return fail("only rainfall 0 or 1 is supported");
    if (p->nolibf < 0 || p->nolibf > 2) return fail("nolibf must be 0, 1 or 2");
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
    if (validate(prm)) return 1;
    if (np <= 0 || ne <= 0) return fail("empty mesh");
    if ((long long)ne * 3 >= (1LL << 31)) return fail("mesh too large for 32-bit indices");
    auto c = std::make_unique<Ctx>();
    c->prm = *prm;
    c->np = np; c->ne = ne;
    std::string err;
    if (!build_ref_tables(prm->p, c->T, err)) return fail(err);
    if (!build_rk(prm->rk_stage, prm->rk_order, c->rk, err)) return fail(err);
    c->dof = c->T.dof; c->nagp = c->T.nagp; c->negp = c->T.negp;
    const int dof = c->dof, nagp = c->nagp, negp = c->negp;
    HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (!g_tables_set[prm->p]) {
            DevTables d = to_dev(c->T);
            if (prm->p == 1) HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_tab1), &d, sizeof d));
            else HIPCHK(hipMemcpyToSymbol(HIP_SYMBOL(c_tab2), &d, sizeof d));
            g_tables_set[prm->p] = true;
        }
    }
    // physics constants (src/dgswem.F:131-170, src/rhs_dg_hydro.F:46, src/prep_DG.F:982-983)
    Phys& P = c->phys;
    P.g = prm->g;
    P.nleq = prm->nolicat == 0 ? 0.0 : 1.0;
    P.leq = prm->nolicat == 0 ? 1.0 : 0.0;
    P.nleqg = P.nleq * P.g;
    P.fg_l = P.leq * P.g;
    P.cf = prm->cf; P.hbreak = prm->hbreak; P.ftheta = prm->ftheta; P.fgamma = prm->fgamma;
    P.dtdph = 1.0 / prm->dtdp;
    P.h0 = prm->h0;
    P.ifnlfa = prm->nolifa == 0 ? 0 : 1;
    P.iflinbf = prm->nolibf == 0 ? 1 : 0;
    P.ifhybf = prm->nolibf == 2 ? 1 : 0;
    P.load_manning = (prm->load_manning && manning) ? 1 : 0;
    P.rainfall = prm->rainfall;
    P.fric = (prm->cf + prm->cf + prm->cf) / 3.0;
    // ---- edges: which slot each side goes to, in the reference's accumulation order
    std::vector<int> elem_slots((size_t)ne * 3, -1);   // edge id per local edge
    std::vector<int> kind_of(nedges, 0), list_pos(nedges, 0);
    for (int i = 0; i < nleds; i++) { kind_of[nledn[i] - 1] = SIDE_LAND; list_pos[nledn[i] - 1] = i; }
    for (int i = 0; i < nfeds; i++) { kind_of[nfedn[i] - 1] = SIDE_FLOW; list_pos[nfedn[i] - 1] = i; }
    for (int i = 0; i < needs; i++) { kind_of[needn[i] - 1] = SIDE_OCEAN; list_pos[needn[i] - 1] = i; }
    for (int i = 0; i < nieds; i++) { kind_of[niedn[i] - 1] = SIDE_INTERNAL; list_pos[niedn[i] - 1] = i; }
    std::vector<int> sinfo(ne, 0), snbr((size_t)3 * ne, -1);
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
                int bits = L | (s << 2) | (k << 3);
                int nb = -1;
                if (k == SIDE_INTERNAL) {
                    const int other = nedel[2 * ed + (1 - s)] - 1;
                    const int oled = nedsd[2 * ed + (1 - s)] - 1;
                    if (other < 0) return fail("internal edge without a second element");
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
                snbr[(size_t)j * ne + e] = nbv[e][ord[j]];
            }
            sinfo[e] = info;
        }
    }
    c->needs = needs; c->nfeds = nfeds;
    // ---- element SoA arrays
    std::vector<double> xv((size_t)3 * ne), yv((size_t)3 * ne), dpv((size_t)3 * ne), nb((size_t)ne);
    std::vector<int> nmv((size_t)3 * ne);
    bool cori_uniform = true;
    for (int i = 1; i < np; i++) if (corif[i] != corif[0]) { cori_uniform = false; break; }
    std::vector<double> cori;
    if (!cori_uniform) cori.resize(ne);
    for (int e = 0; e < ne; e++) {
        int n[3];
        for (int v = 0; v < 3; v++) {
            n[v] = nm[(size_t)v * ne + e] - 1;
            if (n[v] < 0 || n[v] >= np) return fail("NM refers to a node out of range");
            xv[(size_t)v * ne + e] = x[n[v]]; yv[(size_t)v * ne + e] = y[n[v]]; dpv[(size_t)v * ne + e] = dp[n[v]];
            nmv[(size_t)v * ne + e] = n[v];
        }
        const double AREAS = (x[n[0]] - x[n[2]]) * (y[n[1]] - y[n[2]]) + (x[n[2]] - x[n[1]]) * (y[n[0]] - y[n[2]]);
        if (!(AREAS > 0.0)) return fail("element " + std::to_string(e + 1) + " is not counter-clockwise");
        if (P.load_manning) nb[e] = (manning[n[0]] + manning[n[1]] + manning[n[2]]) / 3.0;
        if (!cori_uniform) cori[e] = (corif[n[0]] + corif[n[1]] + corif[n[2]]) / 3.0;
    }
    P.cori = (corif[0] + corif[0] + corif[0]) / 3.0;
    HIPCHK(c->xv.upload(xv.data(), xv.size(), c->stream));
    HIPCHK(c->yv.upload(yv.data(), yv.size(), c->stream));
    HIPCHK(c->dpv.upload(dpv.data(), dpv.size(), c->stream));
    HIPCHK(c->nm.upload(nmv.data(), nmv.size(), c->stream));
    if (P.load_manning) HIPCHK(c->nbar.upload(nb.data(), nb.size(), c->stream));
    if (!cori_uniform) HIPCHK(c->cori_el.upload(cori.data(), cori.size(), c->stream));
    HIPCHK(c->sinfo.upload(sinfo.data(), sinfo.size(), c->stream));
    HIPCHK(c->snbr.upload(snbr.data(), snbr.size(), c->stream));
    // spherical factors at quadrature points (src/prep_DG.F:1010-1016,1085-1091)
    if (prm->ics == 2) {
        const double R = 6378206.4;
        std::vector<double> sq((size_t)nagp * ne), se((size_t)3 * negp * ne);
        for (int e = 0; e < ne; e++) {
            const double Y1 = yv[e], Y2 = yv[(size_t)ne + e], Y3 = yv[2 * (size_t)ne + e];
            const double yd[3] = {1.0 / 3.0 * (Y1 + Y2 + Y3), -(1.0 / 6.0 * (Y1 + Y2)) + 1.0 / 3.0 * Y3, -(0.5 * Y1) + 0.5 * Y2};
            for (int I = 0; I < nagp; I++) {
                double ye = 0.0;
                for (int k = 0; k < 3; k++) ye = ye + yd[k] * c->T.phi_area[I][k];
                sq[(size_t)I * ne + e] = std::cos(prm->sfea0) / std::cos(ye / R);
            }
            for (int L = 0; L < 3; L++)
                for (int I = 0; I < negp; I++) {
                    double ye = 0.0;
                    for (int k = 0; k < 3; k++) ye = ye + yd[k] * c->T.phi_edge[L][I][k];
                    se[((size_t)L * negp + I) * ne + e] = std::cos(prm->sfea0) / std::cos(ye / R);
                }
        }
        HIPCHK(c->sfacq.upload(sq.data(), sq.size(), c->stream));
        HIPCHK(c->sfaced.upload(se.data(), se.size(), c->stream));
    }
    (void)sfac;
    // ---- node -> element (CSR, element index order) for the limiter and modal2nodal
    {
        std::vector<int> off(np + 1, 0), lst((size_t)3 * ne);
        for (int e = 0; e < ne; e++) for (int v = 0; v < 3; v++) off[nmv[(size_t)v * ne + e] + 1]++;
        for (int i = 0; i < np; i++) off[i + 1] += off[i];
        std::vector<int> fill(np, 0);
        for (int e = 0; e < ne; e++)
            for (int v = 0; v < 3; v++) {
                const int n = nmv[(size_t)v * ne + e];
                lst[off[n] + fill[n]++] = (e << 2) | v;
            }
        HIPCHK(c->n2e_off.upload(off.data(), off.size(), c->stream));
        HIPCHK(c->n2e.upload(lst.data(), lst.size(), c->stream));
    }
    HIPCHK(c->nmm.alloc((size_t)6 * np));
    HIPCHK(c->outbuf.alloc((size_t)3 * np));
    HIPCHK(c->nodecode.alloc(np));
    HIPCHK(c->partial.alloc((ne + 255) / 256));
    // ---- state slots and kept RHS
    const int nrk = c->rk.nrk;
    for (int i = 0; i <= nrk; i++) {
        HIPCHK(c->u[i].alloc((size_t)3 * dof * ne));
        HIPCHK(hipMemsetAsync(c->u[i].p, 0, (size_t)3 * dof * ne * sizeof(double), c->stream));
    }
    for (int i = 0; i < nrk; i++) {
        c->keep_r[i] = false;
        for (int s = i + 1; s < nrk; s++) if (c->rk.b[s][i] != 0.0) c->keep_r[i] = true;
        if (c->keep_r[i]) HIPCHK(c->r[i].alloc((size_t)3 * dof * ne));
    }
    for (int i = 0; i < 2; i++) HIPCHK(c->wd[i].alloc(ne));
    // cold start: ZE=QX=QY=0 except the initial dry check (src/prep_DG.F:1214-1300)
    {
        std::vector<double> ze((size_t)dof * ne, 0.0);
        std::vector<uint8_t> wdf(ne, 1);
        if (prm->nolifa >= 2) {
            const double H0 = prm->h0;
            for (int e = 0; e < ne; e++) {
                const double d[3] = {dpv[e], dpv[(size_t)ne + e], dpv[2 * (size_t)ne + e]};
                double z[3] = {0.0, 0.0, 0.0};
                for (int v = 0; v < 3; v++) if (d[v] < H0) z[v] = std::max(z[v], H0 - d[v]);
                if ((z[0] + z[1] + z[2]) / 3.0 != 0.0) {
                    if (!(0.0 > (z[0] + z[1] + z[2]) / 3.0)) {
                        ze[(size_t)0 * ne + e] = (z[0] + z[1] + z[2]) / 3.0;
                        ze[(size_t)1 * ne + e] = -(1.0 / 6.0 * (z[0] + z[1])) + 1.0 / 3.0 * z[2];
                        ze[(size_t)2 * ne + e] = -(0.5 * z[0]) + 0.5 * z[1];
                    }
                    wdf[e] = 0;
                }
            }
        }
        HIPCHK(hipMemcpyAsync(c->u[0].p, ze.data(), ze.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipMemcpyAsync(c->wd[0].p, wdf.data(), ne, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
    }
    c->ucur = 0; c->wdcur = 0;
    std::lock_guard<std::mutex> lk(g_mu);
    g_ctx.push_back(std::move(c));
    *handle = (int)g_ctx.size() - 1;
    return 0;
}

int dg_gpu_set_tide(int h, int nbfr, const double* amig, const double* ff, const double* face,
                    const double* per, const double* emo_dg, const double* efa_dg) {
    Ctx* c = get(h);
    if (!c) return 1;
    if (nbfr < 0) return fail("nbfr < 0");
    if (c->needs > 0 && nbfr == 0) return fail("non-periodic elevation boundary forcing is not supported");
    c->nbfr = nbfr;
    c->amig.assign(amig, amig + nbfr); c->ff.assign(ff, ff + nbfr);
    c->face.assign(face, face + nbfr); c->per.assign(per, per + nbfr);
    // EMO_DG(NBFR,NEEDS,2) column-major -> [needs][nbfr][2]
    std::vector<double> em((size_t)c->needs * nbfr * 2), ef(em.size());
    for (int L = 0; L < c->needs; L++)
        for (int j = 0; j < nbfr; j++)
            for (int k = 0; k < 2; k++) {
                const size_t src = (size_t)j + (size_t)nbfr * (L + (size_t)c->needs * k);
                em[((size_t)L * nbfr + j) * 2 + k] = emo_dg[src];
                ef[((size_t)L * nbfr + j) * 2 + k] = efa_dg[src];
            }
    HIPCHK(c->emo.upload(em.data(), em.size(), c->stream));
    HIPCHK(c->efa.upload(ef.data(), ef.size(), c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int dg_gpu_set_flow(int h, int nffr, const double* famig, const double* fff, const double* fface,
                    const double* fper, const double* qnam_dg, const double* qnph_dg) {
    Ctx* c = get(h);
    if (!c) return 1;
    if (c->nfeds > 0 && nffr == 0) return fail("time-series flux boundary forcing is not supported");
    c->nffr = nffr;
    c->famig.assign(famig, famig + nffr); c->fff.assign(fff, fff + nffr);
    c->fface.assign(fface, fface + nffr); c->fper.assign(fper, fper + nffr);
    std::vector<double> qa((size_t)c->nfeds * nffr * 2), qp(qa.size());
    for (int L = 0; L < c->nfeds; L++)
        for (int j = 0; j < nffr; j++)
            for (int k = 0; k < 2; k++) {
                const size_t src = (size_t)j + (size_t)nffr * (L + (size_t)c->nfeds * k);
                qa[((size_t)L * nffr + j) * 2 + k] = qnam_dg[src];
                qp[((size_t)L * nffr + j) * 2 + k] = qnph_dg[src];
            }
    HIPCHK(c->qnam.upload(qa.data(), qa.size(), c->stream));
    HIPCHK(c->qnph.upload(qp.data(), qp.size(), c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
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
    if (wdflg) {
        std::vector<uint8_t> w(ne);
        for (int e = 0; e < ne; e++) w[e] = wdflg[e] ? 1 : 0;
        HIPCHK(hipMemcpyAsync(c->wd[c->wdcur].p, w.data(), ne, hipMemcpyHostToDevice, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
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
    if (c->needs > 0 && c->amig.empty() && c->nbfr == 0) return fail("tidal forcing not set (dg_gpu_set_tide)");
    if (c->nfeds > 0 && c->nffr == 0) return fail("flux forcing not set (dg_gpu_set_flow)");
    // forcing tables are built per chunk of steps so the host work overlaps nothing but stays small
    const int chunk = 4096;
    for (int s0 = 0; s0 < nsteps; s0 += chunk) {
        const int n = std::min(chunk, nsteps - s0);
        if (s0 > 0 || c->nbfr || c->nffr) {
            HIPCHK(hipStreamSynchronize(c->stream));   // previous tables may still be read
            if (build_stage_tables(*c, it + s0, n)) return 1;
        }
        for (int s = 0; s < n; s++) {
            const int rc = c->prm.p == 1 ? launch_step<1>(*c, s * c->rk.nrk) : launch_step<2>(*c, s * c->rk.nrk);
            if (rc) return rc;
        }
    }
    if (c->prof) return prof_flush(*c);
    return 0;
}

int dg_gpu_synchronize(int h) {
    Ctx* c = get(h);
    if (!c) return 1;
    HIPCHK(hipStreamSynchronize(c->stream));
    return 0;
}

int dg_gpu_modal2nodal(int h, double* eta2, double* uu2, double* vv2, int* nodecode) {
    Ctx* c = get(h);
    if (!c) return 1;
    const int np = c->np;
    ElemData E = elem_data(*c);
    const int B = 256, gn = (np + B - 1) / B;
    int ev = prof_begin(*c, K_M2N);
    if (c->prm.p == 1)
        hipLaunchKernelGGL((modal2nodal_kernel<1>), dim3(gn), dim3(B), 0, c->stream, np, E, (const int*)c->n2e_off.p,
                           (const int*)c->n2e.p, (const double*)c->u[c->ucur].p, (const uint8_t*)c->wd[c->wdcur].p,
                           c->phys.nleq, c->phys.leq, c->outbuf.p, c->nodecode.p);
    else
        hipLaunchKernelGGL((modal2nodal_kernel<2>), dim3(gn), dim3(B), 0, c->stream, np, E, (const int*)c->n2e_off.p,
                           (const int*)c->n2e.p, (const double*)c->u[c->ucur].p, (const uint8_t*)c->wd[c->wdcur].p,
                           c->phys.nleq, c->phys.leq, c->outbuf.p, c->nodecode.p);
    prof_end(*c, ev);
    HIPCHK(hipGetLastError());
    std::vector<double> out((size_t)3 * np);
    HIPCHK(hipMemcpyAsync(out.data(), c->outbuf.p, out.size() * sizeof(double), hipMemcpyDeviceToHost, c->stream));
    if (nodecode) HIPCHK(hipMemcpyAsync(nodecode, c->nodecode.p, (size_t)np * sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->prof && prof_flush(*c)) return 1;
    if (eta2) std::memcpy(eta2, out.data(), (size_t)np * sizeof(double));
    if (uu2) std::memcpy(uu2, out.data() + np, (size_t)np * sizeof(double));
    if (vv2) std::memcpy(vv2, out.data() + 2 * (size_t)np, (size_t)np * sizeof(double));
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

int dg_gpu_profile(int h, int enable) {
    Ctx* c = get(h);
    if (!c) return 1;
    if (prof_flush(*c)) return 1;
    c->prof = enable != 0;
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

int dg_gpu_finalize(int h) {
    std::lock_guard<std::mutex> lk(g_mu);
    if (h < 0 || h >= (int)g_ctx.size() || !g_ctx[h]) { g_err = "invalid handle"; return 1; }
    (void)hipStreamSynchronize(g_ctx[h]->stream);
    g_ctx[h].reset();
    return 0;
}

}  // extern "C"
