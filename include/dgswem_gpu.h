/*
 * dgswem_gpu.h -- C ABI of libdgswem_hip.so, the MI355X (gfx950) explicit hydro step of DG-SWEM.
 *
 * Every entry point is plain C: scalars by value, arrays as pointers to host memory, 1-based
 * indices and Fortran column-major layouts exactly as the reference's modules hold them, so that
 * the Fortran host binds them with iso_c_binding beside src/c_interface.F90 (the reference's only
 * C-ABI precedent, src/c_interface.F90:3-20 + src/count_edges.cpp:12-40).  All entry points return
 * 0 on success and a nonzero code on failure; dg_gpu_last_error() gives the message (the Fortran
 * side STOPs with it, like the reference's fatal-error style).
 *
 * Ownership: the library copies every input at the call; it never keeps host pointers.  Device
 * buffers belong to the handle and are freed by dg_gpu_finalize.  One host thread per handle.
 */
#ifndef DGSWEM_GPU_H
#define DGSWEM_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

/* Run-control scalars after READ_INPUT / PREP_DG (names as in src/global.F, src/dg.F). */
typedef struct dg_params {
    int p;            /* PL = PH: polynomial order, 1 or 2 (src/read_fort_dg.F90:455-457)      */
    int rk_stage;     /* RK_STAGE (NRK)                                                          */
    int rk_order;     /* RK_ORDER                                                                */
    int ics;          /* 1 Cartesian, 2 spherical via CPP (SFAC != 1)                             */
    int nolibf;       /* 0 linear, 1 quadratic/Manning, 2 hybrid bottom friction                 */
    int nolifa;       /* >=2 switches wetting and drying on                                       */
    int nolica;       /* advective terms flags (NLEQ/LEQ, src/dgswem.F:150-170)                   */
    int nolicat;
    int nramp;        /* 0 or 1: tanh ramp over DRAMP days                                        */
    int slopeflag;    /* 0 off, otherwise SLOPELIMITER5                                           */
    int rainfall;     /* 0 off, 1 constant rain rate                                              */
    int load_manning; /* LoadManningsN: per-node Manning's n is given                            */
    /* Inputs the step does not implement, passed so that the library can refuse them instead of
     * running without their terms (each must be as stated, else dg_gpu_init fails):             */
    int nws;          /* NWS: 0, or 8 (then dg_gpu_set_holland before the first step)             */
    int ntip;         /* NTIP: 0, 1 or 2 (then dg_gpu_set_tidal_potential before the first step) */
    int artdif;       /* fort.dg artdif: 0 (no LDG artificial diffusion)                           */
    int pl, padapt;   /* fort.dg pl and padapt: pl == p and padapt == 0 (no p-adaptivity)          */
    int load_quadratic_fric; /* LoadQuadraticFric: 0 (FRIC must be the uniform CF)                */
    double g;         /* G                                                                       */
    double dtdp;      /* DTDP, the time step (s)                                                  */
    double statim;    /* STATIM (days)                                                            */
    double dramp;     /* DRAMP (days)                                                             */
    double h0;        /* H0, minimum depth for wetting and drying                                 */
    double cf;        /* CF (or TAU for NOLIBF=0)                                                 */
    double hbreak, ftheta, fgamma; /* hybrid friction parameters                                 */
    double sfea0;     /* SFEA0 (radians), CPP projection centre latitude                          */
} dg_params;

/* CREATE_EDGE_DATA + CALC_NORMAL (src/create_edge_data.F:31-635, src/calc_normal.F:13-52):
 * build the edge tables from NM and the boundary node strings.  Hosts that already hold the
 * reference's tables (the Fortran driver after PREP_DG) skip this and pass theirs to dg_gpu_init.
 *   nm          NM(NE,3), 1-based, counter-clockwise (Fortran column-major: nm[k*ne + e])
 *   nope, nvdll, nbdv      open-boundary segments: NVDLL(k) nodes each, NBDV concatenated
 *   nbou, ibtype, nvell, nbvv  land/flux segments with their types (0,1,10,11,20,21 land;
 *                          2,12,22 flux-specified)
 * Output arrays are caller-allocated with capacity 3*NE edges (ned_cap); counts returned in
 * counts[5] = {NEDGES, NIEDS, NLEDS, NEEDS, NFEDS}.  nedno/nedel/nedsd are the reference's
 * (2,MNED) arrays (column-major: [edge][2]); nedel is 0 where an edge has one element. */
int dg_create_edge_data(int np, int ne, const int* nm,
                        int nope, const int* nvdll, const int* nbdv,
                        int nbou, const int* ibtype, const int* nvell, const int* nbvv,
                        int ned_cap, int* nedno, int* nedel, int* nedsd,
                        int* niedn, int* nledn, int* needn, int* nfedn, int* counts);

/* Upload mesh, parameters and edge tables; cold-start the state (ZE=QX=QY=0) and set the initial
 * wet/dry flags the way PREP_DG does (src/prep_DG.F:1214-1300).  Replaces, per step, the
 * reference's PREP_DG-side tables that DG_HYDRO_TIMESTEP reads (src/dgswem.F:196 call site).
 *   x, y        projected node coordinates X(NP), Y(NP)
 *   dp          DP(NP) after READ_INPUT's H0 floor
 *   sfac        SFAC(NP) (1 for ICS=1)
 *   corif       CORIF(NP)
 *   manning     ManningsN(NP), or NULL when not loaded
 *   edge arrays as produced by dg_create_edge_data / CREATE_EDGE_DATA
 * *handle receives the new context id. */
int dg_gpu_init(const dg_params* prm, int np, int ne, const double* x, const double* y,
                const double* dp, const double* sfac, const double* corif, const double* manning,
                const int* nm, int nedges, const int* nedno, const int* nedel, const int* nedsd,
                int nieds, const int* niedn, int nleds, const int* nledn, int needs,
                const int* needn, int nfeds, const int* nfedn, int* handle);

/* Periodic elevation forcing on ocean edges (replaces the OCEAN_EDGE_HYDRO inputs,
 * src/ocean_edge_hydro.F:59-90; src/prep_DG.F:183-210 builds EMO_DG/EFA_DG).
 *   amig, ff, face, per   AMIG(NBFR), FF, FACE (rad), PER
 *   emo_dg, efa_dg        EMO_DG(NBFR,NEEDS,2), EFA_DG(NBFR,NEEDS,2), column-major */
int dg_gpu_set_tide(int handle, int nbfr, const double* amig, const double* ff,
                    const double* face, const double* per, const double* emo_dg,
                    const double* efa_dg);

/* Periodic normal-flux forcing on flux-specified edges (FLOW_EDGE_HYDRO inputs,
 * src/flow_edge_hydro.F:45-70; QNAM_DG/QNPH_DG(NFFR,NFEDS,2) from src/prep_DG.F:211-250). */
int dg_gpu_set_flow(int handle, int nffr, const double* famig, const double* fff,
                    const double* fface, const double* fper, const double* qnam_dg,
                    const double* qnph_dg);

/* Non-periodic elevation boundary (NBFR = 0 with open boundaries): the fort.19 time series
 * (src/cstart.F:80-97), in place of dg_gpu_set_tide.  Each stage sets the open-boundary nodes'
 * ETA2 to RAMPDG*(ESBIN1 + ETRATIO*(ESBIN2-ESBIN1)) and the ocean edges interpolate it between their
 * end nodes (src/DG_hydro_timestep.F:153-169, src/ocean_edge_hydro.F:180-184); a record is consumed
 * whenever TIME_A passes the later record's time, as the reference reads unit 19.
 *   neta, nbd   NETA and NBD(NETA), the open-boundary nodes (1-based)
 *   etiminc     ETIMINC (s); record k holds the values at STATIM*86400 + k*ETIMINC
 *   esbin       nrec records of NETA values, record-major (the file's order); nrec >= 2, and the
 *               run fails with an error when it needs a record past the last */
int dg_gpu_set_elev_series(int handle, int neta, const int* nbd, double etiminc, int nrec,
                           const double* esbin);

/* Non-periodic normal flux (NFFR = 0 with flux-specified boundaries): the fort.20 time series
 * (src/cstart.F:101-124), in place of dg_gpu_set_flow.  Each stage forms QNAM_DG(1,L,1:2) from
 * QNIN1/QNIN2 the way the reference does (src/DG_hydro_timestep.F:172-200, nodes I and I+1 of the
 * NVEL list for the L-th flux node) and FLOW_EDGE_HYDRO applies it with RFF = RAMPDG.
 *   nvel, lbcodei  NVEL and LBCODEI(NVEL): the boundary code of every land-boundary node
 *   ftiminc        FTIMINC (s)
 *   qnin           nrec records of one value per flux node (codes 2, 12, 22) in NVEL order */
int dg_gpu_set_flow_series(int handle, int nvel, const int* lbcodei, double ftiminc, int nrec,
                           const double* qnin);

/* Tidal potential forcing (NTIP = 1 or 2; TIDAL_POTENTIAL, src/tidal_potential.F:15-80, and its
 * source term in RHS_DG_HYDRO, src/rhs_dg_hydro.F:117-122,378-386).  The constituents as READ_INPUT
 * holds them (src/read_input.F:1856-1923): TPK, AMIGT (rad/s), ETRF, FFT, FACET (radians), NTIF
 * each; saltamp/saltpha SALTAMP(MNTIF,MNP) and SALTPHA (radians, column-major) for NTIP = 2, NULL
 * for NTIP = 1; slam/sfea SLAM, SFEA (radians); reftim REFTIM (days).  TIP2 is evaluated at every
 * node once per step for all stages; single runs only. */
int dg_gpu_set_tidal_potential(int handle, int ntif, const double* tpk, const double* amigt,
                               const double* etrf, const double* fft, const double* facet,
                               const double* saltamp, const double* saltpha, const double* slam,
                               const double* sfea, double reftim);

/* Replace / read the stage-1 state ZE(DOFH,NE,1), QX, QY (column-major: [elem][mode]) and
 * WDFLG(NE) -- the arrays DG_HYDRO_TIMESTEP advances (src/dg.F:286-313). */
int dg_gpu_set_state(int handle, const double* ze, const double* qx, const double* qy,
                     const int* wdflg);
int dg_gpu_get_state(int handle, double* ze, double* qx, double* qy, int* wdflg);

/* Advance nsteps time steps it, it+1, ..., it+nsteps-1: each is one DG_HYDRO_TIMESTEP(IT)
 * (src/DG_hydro_timestep.F:24-744: all RK stages, edge fluxes, volume quadrature, SSP-RK update,
 * SLOPELIMITER5, WETDRY).  Replaces the call at src/DG_timestep.F:30.  Asynchronous on the
 * handle's stream; dg_gpu_synchronize waits. */
int dg_gpu_step(int handle, int it, int nsteps);
int dg_gpu_synchronize(int handle);

/* modal2nodal (src/write_results.F:14-174): ETA2, UU2, VV2 at the nodes and NODECODE (1 where an
 * incident element is wet; WETDRY's rule, src/wetdry.F:142-169,525-548). Synchronous.  The same
 * as dg_gpu_nodal_update followed by dg_gpu_get_nodal without ETAMAX. */
int dg_gpu_modal2nodal(int handle, double* eta2, double* uu2, double* vv2, int* nodecode);

/* Device-resident modal2nodal, for hosts that read the nodal values only on output steps.
 * dg_gpu_nodal_update replaces `call modal2nodal()` at src/DG_timestep.F:31: it computes ETA2,
 * UU2, VV2, NODECODE into device arrays (and raises ETAMAX there, when set) and copies nothing;
 * asynchronous on the handle's stream.  dg_gpu_get_nodal copies the latest values to the host
 * (any pointer may be NULL; etamax only after dg_gpu_set_etamax).  Synchronous. */
int dg_gpu_nodal_update(int handle);
int dg_gpu_get_nodal(int handle, double* eta2, double* uu2, double* vv2, int* nodecode,
                     double* etamax);

/* ETAMAX(NP), the running maximum of ETA2 that modal2nodal keeps for maxele.63
 * (src/write_results.F:164, set to -99999 by src/prep_DG.F:295): upload the host's values.  From
 * then on every modal2nodal on the device raises it as the reference does, after each wet
 * element's contribution is added to ETA2 (so partial sums count too). */
int dg_gpu_set_etamax(int handle, const double* etamax);

/* The order in which modal2nodal visits each node's elements: EL_COUNT(MNP) and ELETAB(MNP,*)
 * as PREP_DG builds them (src/prep_DG.F:292-330; column-major, ELETAB(I,1+J) is the J-th element
 * of node I).  The order fixes the rounding of ETA2 and which partial sums ETAMAX sees; without
 * this call the elements are visited in ascending element number.  Every node must list exactly
 * its incident elements.  Single runs only. */
int dg_gpu_set_eletab(int handle, int mnp, const int* el_count, const int* eletab);

/* Total water volume sum_e A_e*(ze_1 + hb_1) of the stage-1 state (mode 1 is the cell mean). */
int dg_gpu_mass(int handle, double* mass);

/* The same total as dg_gpu_mass, bit for bit, from an entry point that allocates nothing: no
 * hipMalloc, hipFree or allocating library routine.  No counterpart in the reference; it is for
 * callers that cannot allocate device memory (tight serving loops, harnesses that forbid it).
 * Its partial sums and total live in a fixed 512 KiB device array that the library holds on each
 * GPU until the process ends, so the mesh may have at most 65535*256 = 16,776,960 elements (more
 * is refused with an error).  Every call on one GPU shares that array: calls that overlap on one
 * GPU (two host threads, two handles) race and must be serialised by the caller.  The handle's
 * device must be the current device. */
int dg_gpu_mass_noalloc(int handle, double* mass);

/* Kernel timing with HIP events on the handle's stream.  With enable=n>0 every kernel launch of
 * every n-th step of dg_gpu_step is bracketed by events (n=1: all steps; a larger n samples a
 * long run with little event overhead); enable=0 switches timing off.  dg_gpu_kernel_stats
 * returns the summed milliseconds and
 * launch count of kernel `kernel_id` (0 stage kernel, 1 node min/max, 2 limiter+wet/dry,
 * 3 modal2nodal, 4 NWS=8 node forcing) since the last reset.  No counterpart in the reference
 * (it has no timers). */
int dg_gpu_profile(int handle, int enable);
int dg_gpu_kernel_stats(int handle, int kernel_id, double* total_ms, long long* launches,
                        char* name, int name_len);

/* NWS=8 Holland (1980) vortex wind stress and atmospheric pressure.  Replaces GetHollandStormData
 * (src/wind.F:2153-2450), HollandGet (src/wind.F:1964-2137), WindDrag (src/wind.F:116-141), the
 * NWS=8 branch of MET_FORCING (src/met_forcing.F:479-520) and the wind/pressure source terms of
 * RHS_DG_HYDRO (src/rhs_dg_hydro.F:362-373).  rec: nrec fort.22 records of 12 ints each, the
 * FORMAT 14 fields in order (year, month, day, hour, cast type 0=BEST 1=OFCL, forecast increment
 * h, latitude 0.1 deg N, longitude 0.1 deg W, max wind kt, central pressure mbar, RRP nm, RMW nm);
 * ref: IREFYR IREFMO IREFDAY IREFHR; storm_number and bladj as on the fort.15 NWS=8 line; slam and
 * sfea: node longitude and latitude in radians (SLAM, SFEA).  The profile is evaluated on the GPU
 * once per step at TIME_A; a step whose TIME_A the track does not cover fails with an error. */
int dg_gpu_set_holland(int handle, int nrec, const int* rec, const int* ref, int storm_number,
                       double bladj, const double* slam, const double* sfea);

/* Host-only view of the same track processing, for checking: the storm scalars used at each of
 * ntimes model times (TIME_A, s) -- latitude, longitude (deg), central pressure and pressure
 * deficit (Pa), boundary-layer maximum wind speed (m/s), radius to maximum wind (km), translation
 * velocity (m/s, x and y), Holland B, cos(latitude), wind multiplier, BLAdj, RHOWAT0*G -- as
 * rows[ntimes][16].  Touches no GPU. */
int dg_holland_storm_at(int nrec, const int* rec, const int* ref, int storm_number, double bladj,
                        double g, int ntimes, const double* times, double* rows);

/* ---- Parameters and the discrete adjoint.  None of these has a counterpart in the reference. */

/* Replace the per-node Manning n (ManningsN; element means rebuilt as at set-up) or the per-node
 * depth DP: NP values, for an ensemble the members' nodes in member order (node_offsets). */
int dg_gpu_set_manning(int handle, const double* manning);
int dg_gpu_set_depth(int handle, const double* dp);

/* Gauges for the misfit loss: ng points at reference coordinates (r, s) of 1-based elements
 * (the reference triangle is (-1,-1), (1,-1), (-1,1)); the gauge elevation is the element's
 * modal expansion of ZE there.  For an ensemble the elements are numbered across the members
 * (elem_offsets), so each member's gauges sit in its own elements. */
int dg_adj_set_gauges(int handle, int ng, const int* elem, const double* r, const double* s);

/* Advance nsteps steps (IT = it .. it+nsteps-1) and write eta[k][g], the gauge elevations after
 * step it+k. */
int dg_adj_gauges(int handle, int it, int nsteps, double* eta);

/* Loss J = weight * sum_{k<nsteps} sum_g (eta_g after step it+k - data[k][g])^2 from the current
 * state, and its gradient with respect to the per-node Manning n and depth DP, by the discrete
 * adjoint of the stage RHS and the SSP-RK stages.  States are checkpointed every `interval` steps
 * (0: round(sqrt(nsteps))) and each segment is recomputed once on the way back.  The slope
 * limiter (SLOPELIMITER5) is differentiated with its active set frozen as the forward run
 * recomputes it (the selected node bounds, the redistribution counts and signs are those of the
 * forward run; the reverse of each branch is exact), so the gradient is that of the piecewise
 * smooth map on the forward run's branch.  With the slope limiter on, wetting/drying (WETDRY and
 * the dry-side edge fluxes) is differentiated the same way, with the wet/dry flags and WETDRY's
 * branches those of the forward run.  Without the limiter it is taken as the identity it is when
 * no element changes state; if any element wets, dries or is redistributed during such a run the
 * call fails (nonzero, dg_gpu_last_error() says so) and leaves the state as it was.  NWS=8 forcing
 * and ensembles are supported.  For an ensemble
 * the members are independent, so the gradient of the summed loss restricted to a member's nodes
 * is that member's own gradient (all members in one batched forward and one batched reverse
 * sweep).  The handle's state is left as it was.  grad_manning / grad_dp: NP values each (either
 * may be NULL). */
int dg_adj_gradient(int handle, int it, int nsteps, const double* data, double weight,
                    int interval, double* loss, double* grad_manning, double* grad_dp);

/* ---- Batched ensembles, graph replay and launch tuning.  None of these has a counterpart in the
 * reference, which runs one scenario per process: an ensemble replaces N separate dgswem_serial
 * runs (one DG_HYDRO_TIMESTEP per member per step) with one stage kernel launch per RK stage over
 * all members. */

/* Batch `nmembers` single-run handles (each set up with dg_gpu_init, dg_gpu_set_tide/flow and
 * optionally dg_gpu_set_state) into a new handle.  Meshes may differ in size; members must share
 * p, the RK scheme, ICS, NOLIBF, NOLIFA, NOLICAT, the slope limiter, G, the hybrid friction
 * constants and whether Manning n is loaded.  DTDP, H0, CF, STATIM, DRAMP, RAINFALL, the Coriolis
 * parameter and the boundary forcing are per member.  The member handles stay valid and
 * independent; the ensemble copies their mesh, forcing and current state.  All ensemble entry
 * points (step, set/get_state, modal2nodal, mass) work on the concatenation of the members:
 * element e of member m is element elem_offsets[m] + e, node n is node node_offsets[m] + n. */
int dg_ens_create(int nmembers, const int* member_handles, int* ens_handle);

/* Member count and the nmembers+1 element and node offsets of a handle (0, NE for one run). */
int dg_ens_layout(int handle, int* nmembers, int* elem_offsets, int* node_offsets);

/* enable=1: dg_gpu_step replays HIP graphs of NRK+1 steps (after which the rotating stage buffers
 * are back in place) instead of launching kernels one by one.  Graphs are captured on first use
 * and dropped when the forcing or the block size changes.  Kernel timing (dg_gpu_profile) sees
 * only the eager remainder of each call. */
int dg_gpu_use_graph(int handle, int enable);

/* Time the stage kernel at each supported launch shape (see dg_gpu_set_block) on the handle's
 * current state and keep the fastest; writes it to *block.  Touches only scratch stage buffers:
 * the state is unchanged. */
int dg_gpu_autotune(int handle, int* block);

/* Set the stage-kernel block size explicitly (one thread per element): 64, 128 or (p=1) 256.
 * All sizes give bit-identical results. */
int dg_gpu_set_block(int handle, int block);

/* Message of the last failure on this thread ("" if none). */
const char* dg_gpu_last_error(void);

/* Free the handle's device buffers (src/dgswem.F:429-443 end of run). */
int dg_gpu_finalize(int handle);

#ifdef __cplusplus
}
#endif
#endif /* DGSWEM_GPU_H */
