! ISO_C_BINDING interface to the lag2d C ABI (include/lag2d.h).
!
! Self-contained: no dependency on the reference's modules.  A Fortran caller
! uses it directly, or through the drop-in getsolvdg_lag of
! getsolvdg_lag_shim.f90.
module lag2d_c
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: LAG2D_OK = 0, LAG2D_EINVAL = 1, LAG2D_EUNSUPPORTED = 2, &
                               LAG2D_EDEVICE = 3, LAG2D_ENONFINITE = 4, LAG2D_EVOLUME = 5, &
                               LAG2D_ETOPOLOGY = 6

  ! struct lag2d_config
  type, bind(c) :: lag2d_config
    integer(c_int) :: ndegr, npoly, ncurv, nstag, ndens, nlimi, ncase, nrz, nfint, nriem, ndt
    real(c_double) :: gamlg, cfl, cimpd, cdrho, tend, dtfix, tacru
    integer(c_int) :: strict
  end type lag2d_config

  interface
    integer(c_int) function lag2d_create(cfg, nquad, npoin, nbfac, ipqua, bface, esuv1, esuv2, ctx) &
        bind(c, name='lag2d_create')
      import :: c_int, c_ptr, lag2d_config
      type(lag2d_config), intent(in) :: cfg
      integer(c_int), value :: nquad, npoin, nbfac
      integer(c_int), intent(in) :: ipqua(*), bface(*), esuv1(*), esuv2(*)
      type(c_ptr), intent(out) :: ctx
    end function lag2d_create

    integer(c_int) function lag2d_create_hybrid(cfg, ntria, nquad, npoin, nbfac, iptri, ipqua, bface, &
        esuv1, esuv2, ctx) bind(c, name='lag2d_create_hybrid')
      import :: c_int, c_ptr, lag2d_config
      type(lag2d_config), intent(in) :: cfg
      integer(c_int), value :: ntria, nquad, npoin, nbfac
      integer(c_int), intent(in) :: iptri(*), ipqua(*), bface(*), esuv1(*), esuv2(*)
      type(c_ptr), intent(out) :: ctx
    end function lag2d_create_hybrid

    integer(c_int) function lag2d_upload(ctx, coord, coold, geoel, amatr, unkno, ustar) &
        bind(c, name='lag2d_upload')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: coord(*), coold(*), geoel(*), amatr(*), unkno(*), ustar(*)
    end function lag2d_upload

    integer(c_int) function lag2d_init_state(ctx, coord, rho, u, v, p, nblast, blast, energy) &
        bind(c, name='lag2d_init_state')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: coord(*), rho(*), u(*), v(*), p(*), energy(*)
      integer(c_int), value :: nblast
      integer(c_int), intent(in) :: blast(*)
    end function lag2d_init_state

    integer(c_int) function lag2d_step(ctx, itime) bind(c, name='lag2d_step')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: itime
    end function lag2d_step

    integer(c_int) function lag2d_run(ctx, itime0, nsteps, taken) bind(c, name='lag2d_run')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: itime0, nsteps
      integer(c_int), intent(out) :: taken
    end function lag2d_run

    integer(c_int) function lag2d_rhs(ctx, rhsel) bind(c, name='lag2d_rhs')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: rhsel(*)
    end function lag2d_rhs

    integer(c_int) function lag2d_download(ctx, coord, geoel, unkno, ustar) bind(c, name='lag2d_download')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: coord(*), geoel(*), unkno(*), ustar(*)
    end function lag2d_download

    integer(c_int) function lag2d_get_time(ctx, dtfix, tacru) bind(c, name='lag2d_get_time')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: dtfix, tacru
    end function lag2d_get_time

    integer(c_int) function lag2d_get_mass_matrix(ctx, amatr) bind(c, name='lag2d_get_mass_matrix')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: amatr(*)
    end function lag2d_get_mass_matrix

    integer(c_int) function lag2d_use_graphs(ctx, enable) bind(c, name='lag2d_use_graphs')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: enable
    end function lag2d_use_graphs

    integer(c_int) function lag2d_adj_record(ctx, itime0, nsteps, every) bind(c, name='lag2d_adj_record')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: itime0, nsteps, every
    end function lag2d_adj_record

    integer(c_int) function lag2d_adj_replay(ctx, nsteps) bind(c, name='lag2d_adj_replay')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nsteps
    end function lag2d_adj_replay

    integer(c_int) function lag2d_adj_gradient(ctx, lam_coord, grad) bind(c, name='lag2d_adj_gradient')
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: lam_coord(*)
      real(c_double), intent(out) :: grad(*)
    end function lag2d_adj_gradient

    integer(c_int) function lag2d_check(ctx) bind(c, name='lag2d_check')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function lag2d_check

    integer(c_int) function lag2d_sync(ctx) bind(c, name='lag2d_sync')
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function lag2d_sync

    subroutine lag2d_destroy(ctx) bind(c, name='lag2d_destroy')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine lag2d_destroy

    type(c_ptr) function lag2d_last_error() bind(c, name='lag2d_last_error')
      import :: c_ptr
    end function lag2d_last_error
  end interface

contains

  ! Message of the last failure as a Fortran string.
  function lag2d_error_message() result(msg)
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: s(:)
    type(c_ptr) :: p
    integer :: n
    interface
      integer(c_size_t) function c_strlen(str) bind(c, name='strlen')
        import :: c_size_t, c_ptr
        type(c_ptr), value :: str
      end function c_strlen
    end interface
    p = lag2d_last_error()
    n = int(c_strlen(p))
    call c_f_pointer(p, s, [n])
    allocate(character(len=n) :: msg)
    msg = transfer(s(1:n), msg)
  end function lag2d_error_message

  ! Stop with the library's message on a non-zero status (the reference has
  ! no error returns; a failure ends the run, as its own `stop`s do).
  subroutine lag2d_require(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    if (rc /= LAG2D_OK) then
      write(*, '(a, a, a, i0, a, a)') 'lag2d: ', where, ' failed (', rc, '): ', lag2d_error_message()
      error stop 1
    end if
  end subroutine lag2d_require

end module lag2d_c
