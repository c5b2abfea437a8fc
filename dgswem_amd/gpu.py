"""Host-side driver of libdgswem_hip.so through its C ABI (include/dgswem_gpu.h).

`GpuSolver(case)` does for the GPU library what the reference's Fortran host does around
DG_TIMESTEP: it hands over the mesh, the edge tables of CREATE_EDGE_DATA, the boundary forcing
assembled the way PREP_DG assembles EMO_DG/EFA_DG and QNAM_DG/QNPH_DG, and then advances the state
with dg_gpu_step.  There is no CPU fallback: without the built library every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from .fortio import Case

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("DGSWEM_HIP_LIB") or os.path.join(_HERE, "libdgswem_hip.so")
_LIB = None

_dp = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
_ip = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")

# kernel ids of dg_gpu_kernel_stats
K_STAGE, K_MINMAX, K_LIMIT, K_M2N, K_MET = 0, 1, 2, 3, 4

# every symbol include/dgswem_gpu.h declares
EXPORTED = ["dg_create_edge_data", "dg_gpu_init", "dg_gpu_set_tide", "dg_gpu_set_flow",
            "dg_gpu_set_state", "dg_gpu_get_state", "dg_gpu_step", "dg_gpu_synchronize",
            "dg_gpu_modal2nodal", "dg_gpu_nodal_update", "dg_gpu_get_nodal",
            "dg_gpu_set_etamax", "dg_gpu_set_eletab", "dg_gpu_mass", "dg_gpu_mass_noalloc", "dg_gpu_profile", "dg_gpu_kernel_stats",
            "dg_gpu_set_holland", "dg_holland_storm_at", "dg_gpu_set_tidal_potential",
            "dg_gpu_set_elev_series", "dg_gpu_set_flow_series", "dg_gpu_set_manning", "dg_gpu_set_depth",
            "dg_adj_set_gauges", "dg_adj_gauges", "dg_adj_gradient", "dg_ens_create", "dg_ens_layout", "dg_gpu_use_graph", "dg_gpu_autotune",
            "dg_gpu_set_block", "dg_gpu_last_error", "dg_gpu_finalize"]


class DgParams(C.Structure):
    _fields_ = [("p", C.c_int), ("rk_stage", C.c_int), ("rk_order", C.c_int), ("ics", C.c_int),
                ("nolibf", C.c_int), ("nolifa", C.c_int), ("nolica", C.c_int),
                ("nolicat", C.c_int), ("nramp", C.c_int), ("slopeflag", C.c_int),
                ("rainfall", C.c_int), ("load_manning", C.c_int), ("nws", C.c_int),
                ("ntip", C.c_int), ("artdif", C.c_int), ("pl", C.c_int), ("padapt", C.c_int),
                ("load_quadratic_fric", C.c_int), ("g", C.c_double),
                ("dtdp", C.c_double), ("statim", C.c_double), ("dramp", C.c_double),
                ("h0", C.c_double), ("cf", C.c_double), ("hbreak", C.c_double),
                ("ftheta", C.c_double), ("fgamma", C.c_double), ("sfea0", C.c_double)]


class GpuError(RuntimeError):
    pass


def load_library(path: str = LIB_PATH):
    """Load libdgswem_hip.so and declare its signatures (no device call happens here)."""
    global _LIB
    if _LIB is not None and path == LIB_PATH:
        return _LIB
    if not os.path.exists(path):
        raise GpuError(f"{path} is not built: run `python -c 'import __graft_entry__ as g; g.build()'`")
    L = C.CDLL(path)
    L.dg_gpu_last_error.restype = C.c_char_p
    L.dg_create_edge_data.argtypes = [C.c_int, C.c_int, _ip, C.c_int, _ip, _ip, C.c_int, _ip, _ip,
                                      _ip, C.c_int, _ip, _ip, _ip, _ip, _ip, _ip, _ip, _ip]
    L.dg_gpu_init.argtypes = [C.POINTER(DgParams), C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp,
                              C.c_void_p, _ip, C.c_int, _ip, _ip, _ip, C.c_int, _ip, C.c_int, _ip,
                              C.c_int, _ip, C.c_int, _ip, C.POINTER(C.c_int)]
    L.dg_gpu_set_tide.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
    L.dg_gpu_set_flow.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, _dp]
    L.dg_gpu_set_state.argtypes = [C.c_int, _dp, _dp, _dp, C.c_void_p]
    L.dg_gpu_get_state.argtypes = [C.c_int, _dp, _dp, _dp, _ip]
    L.dg_gpu_step.argtypes = [C.c_int, C.c_int, C.c_int]
    L.dg_gpu_synchronize.argtypes = [C.c_int]
    L.dg_gpu_modal2nodal.argtypes = [C.c_int, _dp, _dp, _dp, _ip]
    L.dg_gpu_nodal_update.argtypes = [C.c_int]
    L.dg_gpu_get_nodal.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                   C.c_void_p]
    L.dg_gpu_set_etamax.argtypes = [C.c_int, _dp]
    L.dg_gpu_set_eletab.argtypes = [C.c_int, C.c_int, _ip, _ip]
    L.dg_gpu_mass.argtypes = [C.c_int, C.POINTER(C.c_double)]
    L.dg_gpu_mass_noalloc.argtypes = [C.c_int, C.POINTER(C.c_double)]
    L.dg_gpu_profile.argtypes = [C.c_int, C.c_int]
    L.dg_gpu_kernel_stats.argtypes = [C.c_int, C.c_int, C.POINTER(C.c_double),
                                      C.POINTER(C.c_longlong), C.c_char_p, C.c_int]
    L.dg_gpu_set_holland.argtypes = [C.c_int, C.c_int, _ip, _ip, C.c_int, C.c_double, _dp, _dp]
    L.dg_holland_storm_at.argtypes = [C.c_int, _ip, _ip, C.c_int, C.c_double, C.c_double, C.c_int,
                                      _dp, _dp]
    L.dg_gpu_set_tidal_potential.argtypes = [C.c_int, C.c_int, _dp, _dp, _dp, _dp, _dp, C.c_void_p,
                                             C.c_void_p, _dp, _dp, C.c_double]
    L.dg_gpu_set_elev_series.argtypes = [C.c_int, C.c_int, _ip, C.c_double, C.c_int, _dp]
    L.dg_gpu_set_flow_series.argtypes = [C.c_int, C.c_int, _ip, C.c_double, C.c_int, _dp]
    L.dg_gpu_set_manning.argtypes = [C.c_int, _dp]
    L.dg_gpu_set_depth.argtypes = [C.c_int, _dp]
    L.dg_adj_set_gauges.argtypes = [C.c_int, C.c_int, _ip, _dp, _dp]
    L.dg_adj_gauges.argtypes = [C.c_int, C.c_int, C.c_int, _dp]
    L.dg_adj_gradient.argtypes = [C.c_int, C.c_int, C.c_int, _dp, C.c_double, C.c_int,
                                  C.POINTER(C.c_double), _dp, _dp]
    L.dg_ens_create.argtypes = [C.c_int, _ip, C.POINTER(C.c_int)]
    L.dg_ens_layout.argtypes = [C.c_int, C.POINTER(C.c_int), C.c_void_p, C.c_void_p]
    L.dg_gpu_use_graph.argtypes = [C.c_int, C.c_int]
    L.dg_gpu_autotune.argtypes = [C.c_int, C.POINTER(C.c_int)]
    L.dg_gpu_set_block.argtypes = [C.c_int, C.c_int]
    L.dg_gpu_finalize.argtypes = [C.c_int]
    for name in EXPORTED:
        getattr(L, name).restype = C.c_char_p if name == "dg_gpu_last_error" else C.c_int
    if path == LIB_PATH:
        _LIB = L
    return L


def _check(rc: int) -> None:
    if rc != 0:
        raise GpuError(load_library().dg_gpu_last_error().decode())


def _f(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def params_of(case: Case) -> DgParams:
    p = case.params
    return DgParams(p=case.dg.ph, rk_stage=case.dg.rk_stage, rk_order=case.dg.rk_order, ics=p.ics,
                    nolibf=p.nolibf, nolifa=p.nolifa, nolica=p.nolica, nolicat=p.nolicat,
                    nramp=p.nramp, slopeflag=case.dg.slopeflag, rainfall=case.dg.rainfall,
                    load_manning=1 if case.manning is not None else 0, nws=p.nws, ntip=p.ntip,
                    artdif=case.dg.artdif, pl=case.dg.pl, padapt=case.dg.padapt,
                    load_quadratic_fric=0, g=p.g, dtdp=p.dtdp,
                    statim=p.statim, dramp=p.dramp, h0=p.h0, cf=p.cf, hbreak=p.hbreak,
                    ftheta=p.ftheta, fgamma=p.fgamma, sfea0=p.sfea0)


def create_edge_data(case: Case) -> dict:
    """CREATE_EDGE_DATA through the library (1-based arrays, like the reference)."""
    L = load_library()
    ne, n_nodes = case.ne, case.np_
    nm_cm = _i(case.nm.T.reshape(-1))            # NM(NE,3) column-major
    ov = _i([len(s) for s in case.open_segs] or [0])
    on = _i(np.concatenate(case.open_segs) if case.open_segs else [0])
    lt = _i([t for t, _ in case.land_segs] or [0])
    lv = _i([len(s) for _, s in case.land_segs] or [0])
    ln = _i(np.concatenate([s for _, s in case.land_segs]) if case.land_segs else [0])
    cap = 3 * ne
    nedno = np.zeros(2 * cap, np.int32); nedel = np.zeros(2 * cap, np.int32)
    nedsd = np.zeros(2 * cap, np.int32)
    niedn = np.zeros(cap, np.int32); nledn = np.zeros(cap, np.int32)
    needn = np.zeros(cap, np.int32); nfedn = np.zeros(cap, np.int32)
    counts = np.zeros(5, np.int32)
    _check(L.dg_create_edge_data(n_nodes, ne, nm_cm, len(case.open_segs), ov, on,
                                 len(case.land_segs), lt, lv, ln, cap, nedno, nedel, nedsd, niedn,
                                 nledn, needn, nfedn, counts))
    ned, ni, nl, no, nf = [int(c) for c in counts]
    return dict(nedges=ned, nedno=nedno[:2 * ned].reshape(ned, 2), nedel=nedel[:2 * ned].reshape(ned, 2),
                nedsd=nedsd[:2 * ned].reshape(ned, 2), niedn=niedn[:ni], nledn=nledn[:nl],
                needn=needn[:no], nfedn=nfedn[:nf])


def emo_dg(case: Case) -> tuple:
    """EMO_DG/EFA_DG(NBFR,NEEDS,2) from per-node EMO/EFA (src/prep_DG.F:183-210), column-major."""
    p = case.params
    pairs = []
    jj = 0
    for seg in case.open_segs:
        for k in range(len(seg) - 1):
            pairs.append((jj, jj + 1)); jj += 1
        jj += 1
    needs = len(pairs)
    emo = np.zeros((p.nbfr, needs, 2)); efa = np.zeros((p.nbfr, needs, 2))
    for i in range(p.nbfr):
        for L, (a, b) in enumerate(pairs):
            emo[i, L] = (p.emo[i, a], p.emo[i, b]); efa[i, L] = (p.efa[i, a], p.efa[i, b])
    return emo.reshape(-1, order="F"), efa.reshape(-1, order="F")


def qnam_dg(case: Case) -> tuple:
    """QNAM_DG/QNPH_DG(NFFR,NFEDS,2) from per-node QNAM/QNPH (src/prep_DG.F:211-250)."""
    p = case.params
    pairs = []
    jj = 0
    for t, seg in case.land_segs:
        if t not in (2, 12, 22):
            continue
        for k in range(len(seg) - 1):
            pairs.append((jj, jj + 1)); jj += 1
        jj += 1
    qa = np.zeros((p.nffr, len(pairs), 2)); qp = np.zeros((p.nffr, len(pairs), 2))
    for i in range(p.nffr):
        for L, (a, b) in enumerate(pairs):
            qa[i, L] = (p.qnam[i, a], p.qnam[i, b]); qp[i, L] = (p.qnph[i, a], p.qnph[i, b])
    return qa.reshape(-1, order="F"), qp.reshape(-1, order="F")


def holland_storm_at(storm, g: float, times) -> np.ndarray:
    """The library's storm scalars at model times (host code only; see dg_holland_storm_at)."""
    t = _f(times)
    rows = np.zeros((len(t), 16))
    _check(load_library().dg_holland_storm_at(len(storm.records), _i(storm.records.reshape(-1)),
                                              _i(storm.ref()), storm.storm_number,
                                              float(storm.bladj), float(g), len(t), t, rows))
    return rows


class _Handle:
    """Operations common to a single run and an ensemble (one library handle each)."""

    h = -1
    ne = np_ = dofh = 0
    it = 0

    def close(self):
        if getattr(self, "h", -1) >= 0:
            load_library().dg_gpu_finalize(self.h)
            self.h = -1

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def step(self, n: int = 1, sync: bool = True) -> None:
        """Advance n time steps (DG_HYDRO_TIMESTEP(IT) for IT = it+1 .. it+n)."""
        L = load_library()
        _check(L.dg_gpu_step(self.h, self.it + 1, n))
        self.it += n
        if sync:
            _check(L.dg_gpu_synchronize(self.h))

    def synchronize(self) -> None:
        _check(load_library().dg_gpu_synchronize(self.h))

    def state(self):
        """(ZE, QX, QY) as [element][mode] arrays and WDFLG."""
        n = self.ne * self.dofh
        ze, qx, qy = np.empty(n), np.empty(n), np.empty(n)
        wd = np.empty(self.ne, np.int32)
        _check(load_library().dg_gpu_get_state(self.h, ze, qx, qy, wd))
        s = (self.ne, self.dofh)
        return ze.reshape(s), qx.reshape(s), qy.reshape(s), wd

    def set_state(self, ze, qx, qy, wdflg=None):
        w = _i(wdflg) if wdflg is not None else None
        _check(load_library().dg_gpu_set_state(self.h, _f(ze).reshape(-1), _f(qx).reshape(-1),
                                                _f(qy).reshape(-1),
                                                w.ctypes.data if w is not None else None))

    def nodal(self):
        """ETA2, UU2, VV2 and NODECODE of modal2nodal."""
        e, u, v = np.empty(self.np_), np.empty(self.np_), np.empty(self.np_)
        nc = np.empty(self.np_, np.int32)
        _check(load_library().dg_gpu_modal2nodal(self.h, e, u, v, nc))
        return e, u, v, nc

    def nodal_update(self) -> None:
        """modal2nodal on the device only (raises ETAMAX there once set); nothing is copied."""
        _check(load_library().dg_gpu_nodal_update(self.h))

    def get_nodal(self):
        """ETA2, UU2, VV2, NODECODE and ETAMAX of the latest nodal_update (ETAMAX None if unset)."""
        e, u, v = np.empty(self.np_), np.empty(self.np_), np.empty(self.np_)
        nc = np.empty(self.np_, np.int32)
        em = np.empty(self.np_) if getattr(self, "_etamax_set", False) else None
        ptr = lambda a: None if a is None else a.ctypes.data_as(C.c_void_p)
        _check(load_library().dg_gpu_get_nodal(self.h, ptr(e), ptr(u), ptr(v), ptr(nc), ptr(em)))
        return e, u, v, nc, em

    def set_etamax(self, etamax) -> None:
        _check(load_library().dg_gpu_set_etamax(self.h, _f(etamax)))
        self._etamax_set = True

    def mass(self) -> float:
        m = C.c_double()
        _check(load_library().dg_gpu_mass(self.h, C.byref(m)))
        return m.value

    def profile(self, enable: bool, every: int = 1) -> None:
        """Time kernel launches with HIP events: those of every `every`-th step."""
        _check(load_library().dg_gpu_profile(self.h, max(1, int(every)) if enable else 0))

    def kernel_stats(self, kid: int):
        t = C.c_double(); n = C.c_longlong(); name = C.create_string_buffer(64)
        _check(load_library().dg_gpu_kernel_stats(self.h, kid, C.byref(t), C.byref(n), name, 64))
        return name.value.decode(), t.value, n.value

    def use_graph(self, enable: bool = True) -> None:
        _check(load_library().dg_gpu_use_graph(self.h, 1 if enable else 0))

    def autotune(self) -> int:
        b = C.c_int()
        _check(load_library().dg_gpu_autotune(self.h, C.byref(b)))
        return b.value

    def set_block(self, block: int) -> None:
        _check(load_library().dg_gpu_set_block(self.h, block))

    # ---- parameters and the discrete adjoint (single runs)
    def set_manning(self, manning) -> None:
        """Per-node Manning n (ManningsN); element means are rebuilt as in the forward set-up."""
        _check(load_library().dg_gpu_set_manning(self.h, _f(manning)))

    def set_depth(self, dp) -> None:
        """Per-node depth DP (after any H0 floor)."""
        _check(load_library().dg_gpu_set_depth(self.h, _f(dp)))

    def set_gauges(self, elem, r, s) -> None:
        """Gauges at reference coordinates (r, s) of 1-based elements."""
        self.ng = len(elem)
        _check(load_library().dg_adj_set_gauges(self.h, self.ng, _i(elem), _f(r), _f(s)))

    def gauges(self, n: int) -> np.ndarray:
        """Advance n steps recording the gauge elevations after each: array [n, ng]."""
        eta = np.empty((n, self.ng))
        _check(load_library().dg_adj_gauges(self.h, self.it + 1, n, eta.reshape(-1)))
        self.it += n
        return eta

    def gradient(self, data, weight: float = 1.0, interval: int = 0):
        """Loss J = weight * sum_k sum_g (eta_g(U^k) - data[k-1, g])^2 over the next len(data)
        steps and its gradient with respect to the per-node Manning n and depth DP, by the
        discrete adjoint with checkpoints every `interval` steps (0: about sqrt(N)).  The state
        is left where it was."""
        d = _f(data)
        n = d.shape[0]
        gn, gdp = np.empty(self.np_), np.empty(self.np_)
        J = C.c_double()
        _check(load_library().dg_adj_gradient(self.h, self.it + 1, n, d.reshape(-1), float(weight),
                                              int(interval), C.byref(J), gn, gdp))
        return J.value, gn, gdp

    def layout(self):
        """Element and node offsets of the members (two entries for a single run)."""
        L = load_library()
        n = C.c_int()
        _check(L.dg_ens_layout(self.h, C.byref(n), None, None))
        eo = np.empty(n.value + 1, np.int32); no = np.empty(n.value + 1, np.int32)
        _check(L.dg_ens_layout(self.h, C.byref(n), eo.ctypes.data, no.ctypes.data))
        return eo, no


class GpuSolver(_Handle):
    """One DG-SWEM run on the GPU (one library handle)."""

    def __init__(self, case: Case):
        L = load_library()
        self.case = case
        self.ne, self.np_, self.dofh = case.ne, case.np_, case.dofh
        ed = create_edge_data(case)
        self.edges = ed
        prm = params_of(case)
        man = _f(case.manning) if case.manning is not None else None
        h = C.c_int(-1)
        _check(L.dg_gpu_init(C.byref(prm), case.np_, case.ne, _f(case.x), _f(case.y), _f(case.dp),
                             _f(case.sfac), _f(case.corif),
                             man.ctypes.data if man is not None else None,
                             _i(case.nm.T.reshape(-1)), ed["nedges"], _i(ed["nedno"].reshape(-1)),
                             _i(ed["nedel"].reshape(-1)), _i(ed["nedsd"].reshape(-1)),
                             len(ed["niedn"]), _i(ed["niedn"]), len(ed["nledn"]), _i(ed["nledn"]),
                             len(ed["needn"]), _i(ed["needn"]), len(ed["nfedn"]), _i(ed["nfedn"]),
                             C.byref(h)))
        self.h = h.value
        p = case.params
        if p.nbfr > 0:
            emo, efa = emo_dg(case)
            _check(L.dg_gpu_set_tide(self.h, p.nbfr, _f(p.amig), _f(p.ff), _f(p.face), _f(p.per),
                                     _f(emo), _f(efa)))
        if p.nffr > 0:
            qa, qp = qnam_dg(case)
            _check(L.dg_gpu_set_flow(self.h, p.nffr, _f(p.famig), _f(p.fff), _f(p.fface),
                                     _f(p.fper), _f(qa), _f(qp)))
        if case.elev_series is not None:
            inc, vals = case.elev_series
            _check(L.dg_gpu_set_elev_series(self.h, vals.shape[1], _i(case.nbd()), float(inc),
                                            vals.shape[0], _f(vals.reshape(-1))))
        if case.flow_series is not None:
            inc, vals = case.flow_series
            lb = _i(case.lbcodei())
            _check(L.dg_gpu_set_flow_series(self.h, len(lb), lb, float(inc), vals.shape[0],
                                            _f(vals.reshape(-1))))
        if p.ntip != 0:
            # SALTAMP(MNTIF,MNP) column-major = [node][constituent] row-major
            sa = _f(p.saltamp.T.reshape(-1)) if p.saltamp is not None else None
            sp = _f(p.saltpha.T.reshape(-1)) if p.saltpha is not None else None
            _check(L.dg_gpu_set_tidal_potential(
                self.h, p.ntif, _f(p.tpk), _f(p.amigt), _f(p.etrf), _f(p.fft), _f(p.facet),
                sa.ctypes.data if sa is not None else None, sp.ctypes.data if sp is not None else None,
                _f(case.slam), _f(case.sfea), float(p.reftim)))
        if p.nws == 8:
            st = case.storm
            _check(L.dg_gpu_set_holland(self.h, len(st.records), _i(st.records.reshape(-1)),
                                        _i(st.ref()), st.storm_number, float(st.bladj),
                                        _f(case.slam), _f(case.sfea)))
        elif p.nws != 0:
            raise GpuError(f"NWS={p.nws} meteorological forcing is not supported")
        self.it = 0


class Ensemble(_Handle):
    """Members (GpuSolver runs, possibly of different meshes and time steps) advanced together:
    one stage-kernel launch per RK stage covers every member.  The members keep their own
    handles; the ensemble starts from their current states and all must be at the same step."""

    def __init__(self, members: list):
        if not members:
            raise GpuError("an ensemble needs at least one member")
        its = {m.it for m in members}
        if len(its) != 1:
            raise GpuError("ensemble members must be at the same time step")
        L = load_library()
        hs = np.array([m.h for m in members], np.int32)
        h = C.c_int(-1)
        _check(L.dg_ens_create(len(members), hs, C.byref(h)))
        self.h = h.value
        self.nmembers = len(members)
        self.dofh = members[0].dofh
        self.eoff, self.noff = self.layout()
        self.ne, self.np_ = int(self.eoff[-1]), int(self.noff[-1])
        self.it = its.pop()

    def split(self, arrays, nodes: bool = False):
        """Split per-element (or per-node) arrays of the batch into per-member lists."""
        off = self.noff if nodes else self.eoff
        return [[a[off[i]:off[i + 1]] for a in arrays] for i in range(self.nmembers)]
