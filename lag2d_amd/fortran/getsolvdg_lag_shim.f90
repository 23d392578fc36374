! Drop-in replacement of the reference's getsolvdg_lag (Getrhslag.f90:1553-2035)
! with its exact argument list, running the step on the GPU through the
! lag2d C ABI.  Compiled inside the reference's build (it uses the
! reference's module `constant`, dgflo2d.f90:4-57) in place of the
! reference's own getsolvdg_lag; the driver (dgflo2d.f90:372-373) is
! unchanged.
!
! State handling: the mesh and the state are uploaded on the first call (or
! when the driver restarts at itime = 1) and stay resident on the GPU; each
! call advances one step and copies the results back into the arrays the
! driver reads afterwards -- unold (new DOFs), cooro (new coordinates),
! ustar, geoel rows 3, 5, 6 (on the RZ path nrz=2 also its mass-centre,
! radius and area rows, and amatr, which the reference's RZ stages rewrite
! through this routine's amatr argument) -- and the module scalars dtfix,
! tacru and epsaw, as the reference routine leaves them.  The driver then
! copies unold into unkno and cooro into coord (dgflo2d.f90:408-422) as
! before.
subroutine getsolvdg_lag(intfac,inpoel,iptri,ipqua,bface,cooro,coord,coora,geofa,geoel,gesgt0,gesgq0,uchar,&
                         unkno,unold,unkaw,unkgd,unogd,strnq_devtp,&
                         ltelem,estri,esqua,itime,amatr,amagd,ustar,coold, esuv1, esuv2)
use constant
use lag2d_c
implicit none
integer*4,dimension(1:nvtri,1:ntria), intent(in)::inpoel
integer,  dimension(1:nvtri,1:ntria), intent(in):: iptri
integer,  dimension(1:nvqua,1:nquad), intent(in):: ipqua
integer*4,dimension(1:nifai,1:nafac), intent(in)::intfac
real*8,dimension(1:ndimn,1:npoin),  intent(in)::coold
real*8,dimension(1:ndimn,1:npoin),  intent(in)::coord
real*8,dimension(1:ndimn,1:npoin)::cooro
real*8,dimension(1:ngefa,1:nafac)::geofa
real*8,dimension(1:ngeel,1:nsize)::geoel
real*8,dimension(1:3,1:ngesgt,1:ntria+nbfac),intent(in)::gesgt0
real*8,dimension(1:3,1:ngesgq,1:nquad),      intent(in)::gesgq0
real*8,dimension(1:nq)::uchar
real*8,dimension(1:ndegr,1:nq,1:nsize)::unold
real*8,dimension(1:ndegr,1:nq,1:nsize),intent(in)::unkno
integer*4,dimension(1:3,1:ncell)::ltelem
integer, dimension(1:nftri,1:ntria), intent(in)::estri
integer, dimension(1:nfqua,1:nquad), intent(in)::esqua
! declared intent(in) by the reference, whose RZ stages nevertheless
! rewrite it (Getrhslagmc.f90:6724); no intent here so it can be returned
real*8,dimension(nmatr,ncell)::amatr
real*8,dimension(1:ndimn,1:npoin)     ::ustar
integer*4,dimension(1:nbfai,1:nbfac)::bface
integer*4, intent(in)::esuv1(npoin1),esuv2(npoin2)
integer::itime
real*8,dimension(1:nmatr,1:ncell),     intent(in)::amagd
real*8,dimension(1:ndegr,1:4,1:nsize), intent(in)::unkgd
real*8,dimension(1:ndegr,1:4,1:nsize)::unogd
real*8,dimension(1:3,1:3,ngstrnf+ngausdq, nquad)::strnq_devtp
real*8,dimension(1:ndimn,1:npoin)::coora
real*8,dimension(1:ndegr,1:nq,1:nsize)::unkaw

type(c_ptr), save :: ctx = c_null_ptr
type(lag2d_config) :: cfg
real(c_double) :: dt, t

if (itime == 1 .and. c_associated(ctx)) then
  call lag2d_destroy(ctx)
  ctx = c_null_ptr
end if
if (.not. c_associated(ctx)) then
  cfg%ndegr = ndegr; cfg%npoly = npoly; cfg%ncurv = ncurv; cfg%nstag = nstag
  cfg%ndens = ndens; cfg%nlimi = nlimi; cfg%ncase = ncase; cfg%nrz = nrz
  cfg%nfint = nfint; cfg%nriem = nriem; cfg%ndt = ndt
  cfg%gamlg = gamlg; cfg%cfl = cfl; cfg%cimpd = cimpd; cfg%cdrho = cdrho; cfg%tend = tend
  cfg%dtfix = dtfix; cfg%tacru = tacru; cfg%strict = 1
  ! the library's nfint=5 branch implements the 5-point Lobatto face rule,
  ! the one that branch's tables are written for
  if (nfint == 5 .and. ngausf /= 5) error stop 'lag2d: nfint=5 needs ngausf=5'
  if (ntria > 0) then
    call lag2d_require(lag2d_create_hybrid(cfg, ntria, nquad, npoin, nbfac, iptri, ipqua, bface, esuv1, esuv2, &
                                           ctx), 'create')
  else
    call lag2d_require(lag2d_create(cfg, nquad, npoin, nbfac, ipqua, bface, esuv1, esuv2, ctx), 'create')
  end if
  call lag2d_require(lag2d_upload(ctx, coord, coold, geoel, amatr, unkno, ustar), 'upload')
end if
call lag2d_require(lag2d_step(ctx, itime), 'step')
call lag2d_require(lag2d_download(ctx, cooro, geoel, unold, ustar), 'download')
call lag2d_require(lag2d_get_time(ctx, dt, t), 'get_time')
dtfix = dt
tacru = t
if (nrz == 2) then
  call lag2d_require(lag2d_get_mass_matrix(ctx, amatr), 'get_mass_matrix')
  ! the step's epsaw (Getrhslag.f90:1663-1692)
  epsaw = 1.d-1
  if (ncase == 7) then
    if (itime <= 2000) then
      epsaw = 1.d0
    elseif (itime <= 6000) then
      epsaw = 4.d-1
    elseif (itime <= 10000) then
      epsaw = 2.d-1
    else
      epsaw = 2.d-2
    end if
  elseif (ncase == 11) then
    if (itime <= 7500) then
      epsaw = 1.d-3
    else
      epsaw = 2.d-2
    end if
  end if
end if
unogd = unkgd
end subroutine getsolvdg_lag
