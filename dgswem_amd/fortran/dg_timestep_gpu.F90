! dg_timestep_gpu.F90 -- DG_TIMESTEP for the reference's Fortran host, driving the MI355X step.
!
! Drop-in for the reference's SUBROUTINE DG_TIMESTEP (src/DG_timestep.F:11-50): linked in place of
! DG_timestep.o, it replaces the call DG_HYDRO_TIMESTEP(IT) with dg_gpu_step and the host
! modal2nodal with dg_gpu_modal2nodal, then calls the reference's own WRITE_RESULTS, so the run
! still reads fort.14/fort.15/fort.dg(/fort.13) and writes fort.6x and the DG hot start.
! The first call hands the module arrays built by READ_INPUT and PREP_DG to the library.
subroutine DG_TIMESTEP(IT)
  use sizes, only : sz
  use global, only : NP, NE, X, Y, DP, SFAC, CORIF, NM, G, DTDP, STATIM, DRAMP, H0, SFEA0, ICS, &
       NOLIFA, NOLICA, NOLICAT, NRAMP, NT, NBFR, AMIG, FF, FACE, PER, ETA2, UU2, VV2, NODECODE, &
       ETAMAX
  use dg, only : PH, RK_STAGE, RK_ORDER, SLOPEFLAG, RAINFALL, DGHOT, DGHOTSPOOL, NEDGES, NEDNO, &
       NEDEL, NEDSD, NIEDS, NIEDN, NLEDS, NLEDN, NEEDS, NEEDN, NFEDS, NFEDN, EMO_DG, EFA_DG, &
       QNAM_DG, QNPH_DG, ZE, QX, QY, WDFLG
  use global, only : NFFR, FAMIG, FFF, FFACE, FPER
  use NodalAttributes, only : NoLiBF, CF, HBreak, FTheta, FGamma, LoadManningsN, ManningsN
  use dg_gpu_interface
  implicit none
  integer, intent(in) :: IT
  integer(c_int), save :: handle = -1
  type(dg_params) :: prm
  real(c_double) :: nodummy(1)
  integer :: i

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
     if (NEEDS > 0) then
        call dg_gpu_check(dg_gpu_set_tide(handle, NBFR, AMIG, FF, FACE, PER, &
             EMO_DG(1:NBFR, 1:NEEDS, 1:2), EFA_DG(1:NBFR, 1:NEEDS, 1:2)), 'dg_gpu_set_tide')
     end if
     if (NFEDS > 0) then
        call dg_gpu_check(dg_gpu_set_flow(handle, NFFR, FAMIG, FFF, FFACE, FPER, &
             QNAM_DG(1:NFFR, 1:NFEDS, 1:2), QNPH_DG(1:NFFR, 1:NFEDS, 1:2)), 'dg_gpu_set_flow')
     end if
     ! the state PREP_DG prepared (cold or DG hot start)
     call dg_gpu_check(dg_gpu_set_state(handle, ZE(:, 1:NE, 1), QX(:, 1:NE, 1), QY(:, 1:NE, 1), &
          WDFLG), 'dg_gpu_set_state')
  end if

  ! DG_HYDRO_TIMESTEP(IT) on the GPU, then modal2nodal (both every step, as the reference does)
  call dg_gpu_check(dg_gpu_step(handle, IT, 1), 'dg_gpu_step')
  call dg_gpu_check(dg_gpu_modal2nodal(handle, ETA2, UU2, VV2, NODECODE), 'dg_gpu_modal2nodal')
  do i = 1, NP
     if (ETAMAX(i) < ETA2(i)) ETAMAX(i) = ETA2(i)
  end do
  ! the modal state is only read on the host when the DG hot start is written
  if (DGHOT == 1) then
     if ((mod(IT, DGHOTSPOOL) == 0) .or. (IT == NT)) then
        call dg_gpu_check(dg_gpu_get_state(handle, ZE(:, 1:NE, 1), QX(:, 1:NE, 1), QY(:, 1:NE, 1), &
             WDFLG), 'dg_gpu_get_state')
     end if
  end if
  call WRITE_RESULTS(IT, .FALSE.)
end subroutine DG_TIMESTEP
