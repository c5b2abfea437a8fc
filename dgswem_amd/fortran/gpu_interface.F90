! gpu_interface.F90 -- iso_c_binding interfaces to libdgswem_hip.so (include/dgswem_gpu.h).
!
! The Fortran side of the drop-in boundary, placed beside the reference's src/c_interface.F90
! (the reference's own bind(C) precedent, src/c_interface.F90:3-20).  Every argument keeps the
! reference's layout: 1-based indices, column-major arrays, REAL(SZ)=REAL(8).
module dg_gpu_interface
  use, intrinsic :: iso_c_binding
  implicit none

  type, bind(C) :: dg_params
     integer(c_int) :: p, rk_stage, rk_order, ics, nolibf, nolifa, nolica, nolicat, nramp
     integer(c_int) :: slopeflag, rainfall, load_manning
     integer(c_int) :: nws, ntip, artdif, pl, padapt, load_quadratic_fric
     real(c_double) :: g, dtdp, statim, dramp, h0, cf, hbreak, ftheta, fgamma, sfea0
  end type dg_params

  interface
     integer(c_int) function dg_gpu_init(prm, np, ne, x, y, dp, sfac, corif, manning, nm, &
          nedges, nedno, nedel, nedsd, nieds, niedn, nleds, nledn, needs, needn, nfeds, nfedn, &
          handle) bind(C, name='dg_gpu_init')
       import :: c_int, c_double, dg_params
       type(dg_params), intent(in) :: prm
       integer(c_int), value :: np, ne, nedges, nieds, nleds, needs, nfeds
       real(c_double), intent(in) :: x(*), y(*), dp(*), sfac(*), corif(*), manning(*)
       integer(c_int), intent(in) :: nm(*), nedno(*), nedel(*), nedsd(*)
       integer(c_int), intent(in) :: niedn(*), nledn(*), needn(*), nfedn(*)
       integer(c_int), intent(out) :: handle
     end function dg_gpu_init

     integer(c_int) function dg_gpu_set_tide(handle, nbfr, amig, ff, face, per, emo_dg, efa_dg) &
          bind(C, name='dg_gpu_set_tide')
       import :: c_int, c_double
       integer(c_int), value :: handle, nbfr
       real(c_double), intent(in) :: amig(*), ff(*), face(*), per(*), emo_dg(*), efa_dg(*)
     end function dg_gpu_set_tide

     integer(c_int) function dg_gpu_set_flow(handle, nffr, famig, fff, fface, fper, qnam_dg, qnph_dg) &
          bind(C, name='dg_gpu_set_flow')
       import :: c_int, c_double
       integer(c_int), value :: handle, nffr
       real(c_double), intent(in) :: famig(*), fff(*), fface(*), fper(*), qnam_dg(*), qnph_dg(*)
     end function dg_gpu_set_flow

     integer(c_int) function dg_gpu_set_state(handle, ze, qx, qy, wdflg) bind(C, name='dg_gpu_set_state')
       import :: c_int, c_double
       integer(c_int), value :: handle
       real(c_double), intent(in) :: ze(*), qx(*), qy(*)
       integer(c_int), intent(in) :: wdflg(*)
     end function dg_gpu_set_state

     integer(c_int) function dg_gpu_get_state(handle, ze, qx, qy, wdflg) bind(C, name='dg_gpu_get_state')
       import :: c_int, c_double
       integer(c_int), value :: handle
       real(c_double), intent(out) :: ze(*), qx(*), qy(*)
       integer(c_int), intent(out) :: wdflg(*)
     end function dg_gpu_get_state

     integer(c_int) function dg_gpu_step(handle, it, nsteps) bind(C, name='dg_gpu_step')
       import :: c_int
       integer(c_int), value :: handle, it, nsteps
     end function dg_gpu_step

     integer(c_int) function dg_gpu_synchronize(handle) bind(C, name='dg_gpu_synchronize')
       import :: c_int
       integer(c_int), value :: handle
     end function dg_gpu_synchronize

     integer(c_int) function dg_gpu_modal2nodal(handle, eta2, uu2, vv2, nodecode) &
          bind(C, name='dg_gpu_modal2nodal')
       import :: c_int, c_double
       integer(c_int), value :: handle
       real(c_double), intent(out) :: eta2(*), uu2(*), vv2(*)
       integer(c_int), intent(out) :: nodecode(*)
     end function dg_gpu_modal2nodal

     integer(c_int) function dg_gpu_nodal_update(handle) bind(C, name='dg_gpu_nodal_update')
       import :: c_int
       integer(c_int), value :: handle
     end function dg_gpu_nodal_update

     integer(c_int) function dg_gpu_get_nodal(handle, eta2, uu2, vv2, nodecode, etamax) &
          bind(C, name='dg_gpu_get_nodal')
       import :: c_int, c_double
       integer(c_int), value :: handle
       real(c_double), intent(out) :: eta2(*), uu2(*), vv2(*), etamax(*)
       integer(c_int), intent(out) :: nodecode(*)
     end function dg_gpu_get_nodal

     integer(c_int) function dg_gpu_set_etamax(handle, etamax) bind(C, name='dg_gpu_set_etamax')
       import :: c_int, c_double
       integer(c_int), value :: handle
       real(c_double), intent(in) :: etamax(*)
     end function dg_gpu_set_etamax

     integer(c_int) function dg_gpu_set_eletab(handle, mnp, el_count, eletab) &
          bind(C, name='dg_gpu_set_eletab')
       import :: c_int
       integer(c_int), value :: handle, mnp
       integer(c_int), intent(in) :: el_count(*), eletab(*)
     end function dg_gpu_set_eletab

     integer(c_int) function dg_gpu_set_holland(handle, nrec, rec, ref, storm_number, bladj, &
          slam, sfea) bind(C, name='dg_gpu_set_holland')
       import :: c_int, c_double
       integer(c_int), value :: handle, nrec, storm_number
       integer(c_int), intent(in) :: rec(*), ref(*)
       real(c_double), value :: bladj
       real(c_double), intent(in) :: slam(*), sfea(*)
     end function dg_gpu_set_holland

     integer(c_int) function dg_gpu_set_elev_series(handle, neta, nbd, etiminc, nrec, esbin) &
          bind(C, name='dg_gpu_set_elev_series')
       import :: c_int, c_double
       integer(c_int), value :: handle, neta, nrec
       integer(c_int), intent(in) :: nbd(*)
       real(c_double), value :: etiminc
       real(c_double), intent(in) :: esbin(*)
     end function dg_gpu_set_elev_series

     integer(c_int) function dg_gpu_set_flow_series(handle, nvel, lbcodei, ftiminc, nrec, qnin) &
          bind(C, name='dg_gpu_set_flow_series')
       import :: c_int, c_double
       integer(c_int), value :: handle, nvel, nrec
       integer(c_int), intent(in) :: lbcodei(*)
       real(c_double), value :: ftiminc
       real(c_double), intent(in) :: qnin(*)
     end function dg_gpu_set_flow_series

     integer(c_int) function dg_gpu_set_tidal_potential(handle, ntif, tpk, amigt, etrf, fft, facet, &
          saltamp, saltpha, slam, sfea, reftim) bind(C, name='dg_gpu_set_tidal_potential')
       import :: c_int, c_double
       integer(c_int), value :: handle, ntif
       real(c_double), intent(in) :: tpk(*), amigt(*), etrf(*), fft(*), facet(*)
       real(c_double), intent(in) :: saltamp(*), saltpha(*), slam(*), sfea(*)
       real(c_double), value :: reftim
     end function dg_gpu_set_tidal_potential

     type(c_ptr) function dg_gpu_last_error() bind(C, name='dg_gpu_last_error')
       import :: c_ptr
     end function dg_gpu_last_error

     integer(c_int) function dg_gpu_finalize(handle) bind(C, name='dg_gpu_finalize')
       import :: c_int
       integer(c_int), value :: handle
     end function dg_gpu_finalize

     integer(c_size_t) function c_strlen(s) bind(C, name='strlen')
       import :: c_size_t, c_ptr
       type(c_ptr), value :: s
     end function c_strlen
  end interface

contains

  ! STOP with the library's message, like the reference's fatal-error handling.
  subroutine dg_gpu_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    type(c_ptr) :: p
    integer :: n, i
    character(len=512) :: text
    if (rc == 0) return
    p = dg_gpu_last_error()
    n = int(c_strlen(p))
    call c_f_pointer(p, msg, [n])
    text = ''
    do i = 1, min(n, 512)
       text(i:i) = msg(i)
    end do
    write(*, '(4A)') 'DG GPU ERROR in ', where, ': ', trim(text)
    stop
  end subroutine dg_gpu_check

end module dg_gpu_interface
