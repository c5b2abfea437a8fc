! dg_timestep_gpu.F90 -- DG_TIMESTEP for the reference's Fortran host, driving the MI355X step.
!
! Drop-in for the reference's SUBROUTINE DG_TIMESTEP (src/DG_timestep.F:11-50): linked in place of
! DG_timestep.o, it replaces the call DG_HYDRO_TIMESTEP(IT) with dg_gpu_step and the host
! modal2nodal with dg_gpu_nodal_update, then calls the reference's own WRITE_RESULTS, so the run
! still reads fort.14/fort.15/fort.dg(/fort.13/fort.22) and writes fort.6x, maxele.63 and the DG
! hot start.  The first call hands the module arrays built by READ_INPUT and PREP_DG to the
! library, with the boundary time series of fort.19/fort.20 and the tidal potential constituents
! when the run has them; inputs whose terms the library does not have (NWS other than 0 and 8,
! artdif, p-adaptivity, a per-node quadratic friction attribute) make dg_gpu_init fail and the run
! STOP.
!
! The nodal values (ETA2, UU2, VV2, NODECODE), the running maximum ETAMAX and the modal state stay
! on the GPU; they are copied to the host only on the steps at which WRITE_RESULTS reads them: an
! output whose spool counter fires (src/write_results.F:205-208,285-288,343-346,402-403,458-461,
! 660-663,716-717,790-793), harmonic analysis, a hot start, the DG hot start, and the last step.
subroutine DG_TIMESTEP(IT)
  use sizes, only : sz, MNP
  use global, only : NP, NE, X, Y, DP, SFAC, CORIF, NM, G, DTDP, STATIM, DRAMP, H0, SFEA0, ICS, &
       NOLIFA, NOLICA, NOLICAT, NRAMP, NT, NBFR, AMIG, FF, FACE, PER, ETA2, UU2, VV2, NODECODE, &
       ETAMAX, NWS, NTIP, IREFYR, IREFMO, IREFDAY, IREFHR, SLAM, SFEA, EL_COUNT, ELETAB
  use global, only : NOUTE, NTCYSE, NTCYFE, NSCOUE, NSPOOLE, NOUTV, NTCYSV, NTCYFV, NSCOUV, &
       NSPOOLV, NOUTC, NTCYSC, NTCYFC, NSCOUC, NSPOOLC, NOUTM, NTCYSM, NTCYFM, NSCOUM, NSPOOLM, &
       NOUTGE, NTCYSGE, NTCYFGE, NSCOUGE, NSPOOLGE, NOUTGV, NTCYSGV, NTCYFGV, NSCOUGV, NSPOOLGV, &
       NOUTGC, NTCYSGC, NTCYFGC, NSCOUGC, NSPOOLGC, NOUTGW, NTCYSGW, NTCYFGW, NSCOUGW, NSPOOLGW, &
       IHARIND, ITHAS, NHSTAR, NHSINC
  use dg, only : PH, PL, PADAPT, ARTDIF, RK_STAGE, RK_ORDER, SLOPEFLAG, RAINFALL, DGHOT, &
       DGHOTSPOOL, NEDGES, NEDNO, NEDEL, NEDSD, NIEDS, NIEDN, NLEDS, NLEDN, NEEDS, NEEDN, NFEDS, &
       NFEDN, EMO_DG, EFA_DG, QNAM_DG, QNPH_DG, ZE, QX, QY, WDFLG
  use global, only : NFFR, FAMIG, FFF, FFACE, FPER
  use global, only : NTIF, TPK, AMIGT, ETRF, FFT, FACET, SALTAMP, SALTPHA, REFTIM, NETA, NBD, &
       ESBIN1, ESBIN2, ETIMINC, NVEL, LBCODEI, QNIN1, QNIN2, FTIMINC
  use NodalAttributes, only : NoLiBF, CF, HBreak, FTheta, FGamma, LoadManningsN, ManningsN, &
       LoadQuadraticFric
  use wind, only : StormNumber, BLAdj
  use dg_gpu_interface
  implicit none
  integer, intent(in) :: IT
  integer(c_int), save :: handle = -1
  type(dg_params) :: prm
  real(c_double) :: nodummy(1)
  integer(c_int), allocatable :: eletab_i(:, :)
  integer :: ncol, i, j
  logical :: host

  if (handle < 0) then
     prm%p = PH
     prm%rk_stage = RK_STAGE
     prm%rk_order = RK_ORDER
     prm%ics = ICS
     prm%nolibf = NoLiBF
     prm%nolifa = NOLIFA
     prm%nolica = NOLICA
     prm%nolicat = NOLICAT
     prm%nramp = NRAMP
     prm%slopeflag = SLOPEFLAG
     prm%rainfall = RAINFALL
     prm%load_manning = merge(1, 0, LoadManningsN)
     prm%nws = NWS
     prm%ntip = NTIP
     prm%artdif = ARTDIF
     prm%pl = PL
     prm%padapt = PADAPT
     prm%load_quadratic_fric = merge(1, 0, LoadQuadraticFric)
     prm%g = G
     prm%dtdp = DTDP
     prm%statim = STATIM
     prm%dramp = DRAMP
     prm%h0 = H0
     prm%cf = CF
     prm%hbreak = HBreak
     prm%ftheta = FTheta
     prm%fgamma = FGamma
     prm%sfea0 = SFEA0
     nodummy = 0.0d0
     if (LoadManningsN) then
        call dg_gpu_check(dg_gpu_init(prm, NP, NE, X, Y, DP, SFAC, CORIF, ManningsN, NM, NEDGES, &
             NEDNO, NEDEL, NEDSD, NIEDS, NIEDN, NLEDS, NLEDN, NEEDS, NEEDN, NFEDS, NFEDN, handle), &
             'dg_gpu_init')
     else
        call dg_gpu_check(dg_gpu_init(prm, NP, NE, X, Y, DP, SFAC, CORIF, nodummy, NM, NEDGES, &
             NEDNO, NEDEL, NEDSD, NIEDS, NIEDN, NLEDS, NLEDN, NEEDS, NEEDN, NFEDS, NFEDN, handle), &
             'dg_gpu_init')
     end if
     if (NEEDS > 0 .and. NBFR > 0) then
        call dg_gpu_check(dg_gpu_set_tide(handle, NBFR, AMIG, FF, FACE, PER, &
             EMO_DG(1:NBFR, 1:NEEDS, 1:2), EFA_DG(1:NBFR, 1:NEEDS, 1:2)), 'dg_gpu_set_tide')
     end if
     if (NEEDS > 0 .and. NBFR == 0) call set_elevation_series()
     if (NFEDS > 0 .and. NFFR > 0) then
        call dg_gpu_check(dg_gpu_set_flow(handle, NFFR, FAMIG, FFF, FFACE, FPER, &
             QNAM_DG(1:NFFR, 1:NFEDS, 1:2), QNPH_DG(1:NFFR, 1:NFEDS, 1:2)), 'dg_gpu_set_flow')
     end if
     if (NFEDS > 0 .and. NFFR == 0) call set_flow_series()
     if (NTIP /= 0) then
        call dg_gpu_check(dg_gpu_set_tidal_potential(handle, NTIF, TPK, AMIGT, ETRF, FFT, FACET, &
             SALTAMP, SALTPHA, SLAM, SFEA, REFTIM), 'dg_gpu_set_tidal_potential')
     end if
     if (NWS == 8) call set_holland_track()
     ! modal2nodal visits each node's elements in PREP_DG's order (ELETAB holds them as reals)
     ncol = maxval(EL_COUNT(1:NP)) + 1
     allocate(eletab_i(MNP, ncol))
     eletab_i = 0
     do i = 1, NP
        do j = 1, EL_COUNT(i) + 1
           eletab_i(i, j) = nint(ELETAB(i, j))
        end do
     end do
     call dg_gpu_check(dg_gpu_set_eletab(handle, MNP, EL_COUNT, eletab_i), 'dg_gpu_set_eletab')
     deallocate(eletab_i)
     call dg_gpu_check(dg_gpu_set_etamax(handle, ETAMAX), 'dg_gpu_set_etamax')
     ! the state PREP_DG prepared (cold or DG hot start)
     call dg_gpu_check(dg_gpu_set_state(handle, ZE(:, 1:NE, 1), QX(:, 1:NE, 1), QY(:, 1:NE, 1), &
          WDFLG), 'dg_gpu_set_state')
  end if

  ! DG_HYDRO_TIMESTEP(IT) and modal2nodal on the GPU (both every step, as the reference does)
  call dg_gpu_check(dg_gpu_step(handle, IT, 1), 'dg_gpu_step')
  call dg_gpu_check(dg_gpu_nodal_update(handle), 'dg_gpu_nodal_update')

  ! copy back only where WRITE_RESULTS(IT) reads the nodal values or the modal state
  host = (IT == NT) &
       .or. fires(NOUTE, NTCYSE, NTCYFE, NSCOUE, NSPOOLE) &
       .or. fires(NOUTV, NTCYSV, NTCYFV, NSCOUV, NSPOOLV) &
       .or. fires(NOUTC, NTCYSC, NTCYFC, NSCOUC, NSPOOLC) &
       .or. fires(NOUTM, NTCYSM, NTCYFM, NSCOUM, NSPOOLM) &
       .or. fires(NOUTGE, NTCYSGE, NTCYFGE, NSCOUGE, NSPOOLGE) &
       .or. fires(NOUTGV, NTCYSGV, NTCYFGV, NSCOUGV, NSPOOLGV) &
       .or. fires(NOUTGC, NTCYSGC, NTCYFGC, NSCOUGC, NSPOOLGC) &
       .or. fires(NOUTGW, NTCYSGW, NTCYFGW, NSCOUGW, NSPOOLGW) &
       .or. ((IHARIND == 1) .and. (IT > ITHAS))
  if (NHSTAR == 1) host = host .or. (mod(IT, NHSINC) == 0)
  if (DGHOT == 1) host = host .or. (mod(IT, DGHOTSPOOL) == 0)
  if (host) then
     call dg_gpu_check(dg_gpu_get_nodal(handle, ETA2, UU2, VV2, NODECODE, ETAMAX), 'dg_gpu_get_nodal')
     call dg_gpu_check(dg_gpu_get_state(handle, ZE(:, 1:NE, 1), QX(:, 1:NE, 1), QY(:, 1:NE, 1), &
          WDFLG), 'dg_gpu_get_state')
  end if
  call WRITE_RESULTS(IT, .FALSE.)

contains

  ! WRITE_RESULTS' test for one output: inside its window, and the spool counter it is about to
  ! increment reaches the spool interval
  logical function fires(nout, ntcys, ntcyf, nscou, nspool)
    integer, intent(in) :: nout, ntcys, ntcyf, nscou, nspool
    fires = (nout /= 0) .and. (IT > ntcys) .and. (IT <= ntcyf) .and. (nscou + 1 == nspool)
  end function fires

  ! fort.19 for NBFR = 0: cold start read the first two records into ESBIN1/ESBIN2 (src/cstart.F:
  ! 88-97) and left unit 19 open; the stages would read the rest one record at a time
  ! (src/DG_hydro_timestep.F:158), so the whole remainder goes to the library now
  subroutine set_elevation_series()
    real(c_double), allocatable :: rec(:), tmp(:)
    real(c_double) :: v
    integer :: n, ios, nrec
    allocate(rec(2 * NETA + 1024))
    rec(1:NETA) = ESBIN1(1:NETA)
    rec(NETA + 1:2 * NETA) = ESBIN2(1:NETA)
    n = 2 * NETA
    do
       read(19, *, iostat=ios) v
       if (ios /= 0) exit
       if (n == size(rec)) then
          allocate(tmp(2 * size(rec)))
          tmp(1:n) = rec(1:n)
          call move_alloc(tmp, rec)
       end if
       n = n + 1
       rec(n) = v
    end do
    nrec = n / NETA
    call dg_gpu_check(dg_gpu_set_elev_series(handle, NETA, NBD, ETIMINC, nrec, rec), &
         'dg_gpu_set_elev_series')
    deallocate(rec)
  end subroutine set_elevation_series

  ! fort.20 for NFFR = 0: QNIN1/QNIN2 hold the first two records over the NVEL nodes (zero where
  ! the code is not a flux code, src/cstart.F:109-121); the rest is read from unit 20 the same way
  subroutine set_flow_series()
    real(c_double), allocatable :: rec(:), tmp(:)
    real(c_double) :: v
    integer :: n, ios, nrec, nflux, j
    nflux = 0
    do j = 1, NVEL
       if (LBCODEI(j) == 2 .or. LBCODEI(j) == 12 .or. LBCODEI(j) == 22) nflux = nflux + 1
    end do
    allocate(rec(2 * nflux + 1024))
    n = 0
    do j = 1, NVEL
       if (LBCODEI(j) == 2 .or. LBCODEI(j) == 12 .or. LBCODEI(j) == 22) then
          n = n + 1
          rec(n) = QNIN1(j)
       end if
    end do
    do j = 1, NVEL
       if (LBCODEI(j) == 2 .or. LBCODEI(j) == 12 .or. LBCODEI(j) == 22) then
          n = n + 1
          rec(n) = QNIN2(j)
       end if
    end do
    do
       read(20, *, iostat=ios) v
       if (ios /= 0) exit
       if (n == size(rec)) then
          allocate(tmp(2 * size(rec)))
          tmp(1:n) = rec(1:n)
          call move_alloc(tmp, rec)
       end if
       n = n + 1
       rec(n) = v
    end do
    nrec = n / nflux
    call dg_gpu_check(dg_gpu_set_flow_series(handle, NVEL, LBCODEI, FTIMINC, nrec, rec), &
         'dg_gpu_set_flow_series')
    deallocate(rec)
  end subroutine set_flow_series

  ! fort.22 (ATCF track) for NWS=8, read with the fixed columns GetHollandStormData reads
  ! (src/wind.F:2224-2227, FORMAT 14 at src/wind.F:2440): 12 integers per record for the library
  subroutine set_holland_track()
    integer :: u, ios, n, k
    integer(c_int), allocatable :: rec(:, :)
    integer(c_int) :: ref(4)
    integer :: iyr, imo, idy, ihr, ifc, ila, ilo, isp, icp, irr, irm
    character(len=4) :: cast
    open(newunit=u, file='fort.22', status='old', action='read', iostat=ios)
    if (ios /= 0) then
       print *, 'DG GPU ERROR: NWS=8 needs fort.22'
       stop
    end if
    n = 0
    do
       read(u, '(A)', iostat=ios)
       if (ios /= 0) exit
       n = n + 1
    end do
    rewind(u)
    allocate(rec(12, n))
    do k = 1, n
       read(u, 14, iostat=ios) iyr, imo, idy, ihr, cast, ifc, ila, ilo, isp, icp, irr, irm
       if (ios /= 0) then
          print *, 'DG GPU ERROR: cannot read fort.22 line ', k
          stop
       end if
       rec(:, k) = [iyr, imo, idy, ihr, 2, ifc, ila, ilo, isp, icp, irr, irm]
       if (cast == 'BEST') rec(5, k) = 0
       if (cast == 'OFCL') rec(5, k) = 1
       if (rec(5, k) == 2) then
          print *, 'DG GPU ERROR: fort.22 line ', k, ' has cast type ', cast, &
               ' (BEST or OFCL expected)'
          stop
       end if
    end do
    close(u)
    ref = [IREFYR, IREFMO, IREFDAY, IREFHR]
    call dg_gpu_check(dg_gpu_set_holland(handle, n, rec, ref, StormNumber, BLAdj, SLAM, SFEA), &
         'dg_gpu_set_holland')
    deallocate(rec)
14  format(8x, i4, i2, i2, i2, 6x, a4, 2x, i3, 1x, i4, 3x, i4, 3x, i3, 2x, i4, 47x, i3, 2x, i3)
  end subroutine set_holland_track

end subroutine DG_TIMESTEP
