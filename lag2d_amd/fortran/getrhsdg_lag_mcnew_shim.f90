! Drop-in replacement of the reference's getrhsdg_lag_mcnew
! (Getrhslagmc.f90:6693-8731), the right-hand side of one RK stage, with its
! exact argument list, computed on the GPU through the lag2d C ABI.
! Compiled inside the reference's build (it uses the reference's module
! `constant`, dgflo2d.f90:4-57) in place of the reference's own routine; the
! reference's getsolvdg_lag (Getrhslag.f90:1553-2035) keeps the RK loop on
! the host and calls this once per stage (Getrhslag.f90:1747-1750).
!
! Each call uploads the stage state it is given -- coord (the stage
! coordinates), coold, geoel, amatr, unkno (the stage DOFs) and ustar (the
! previous nodal velocity) -- computes the right-hand side and copies back
! what the reference's routine writes: rhsel, the DOFs limited in place
! (Getrhslagmc.f90:7283-7296, WENO.f90:8086) and the new nodal velocity.
! The mesh is uploaded on the first call.  Covers the planar paths
! (nrz = 0): the straight P1 branch and the curved sub-cell branch.  The
! deformation-gradient right-hand side rhsgd, which no result of the
! Lagrangian gas path reads, is returned as zero.
subroutine getrhsdg_lag_mcnew(ustar,uchar,bface,unkno,unkaw,unkgd,strnq_devtp,rhsel,rhsgd,intfac,inpoel,iptri,ipqua,&
                              geofa,geoel,gesgt0,gesgq0,&
                              coord,coold,coora,&
                              esuv1, esuv2,ltelem,estri,esqua,amatr,itime)
use constant
use lag2d_c
implicit none
integer*4,dimension(1:nbfai,nbfac)::bface
integer*4,dimension(1:nvtri,1:ntria), intent(in)::inpoel
integer,  dimension(1:nvtri,1:ntria), intent(in):: iptri
integer,  dimension(1:nvqua,1:nquad), intent(in):: ipqua
integer*4, intent(in)::esuv1(npoin1),esuv2(npoin2)
integer::ltelem(3,ncell)
integer, dimension(1:nftri,1:ntria), intent(in)::estri
integer, dimension(1:nfqua,1:nquad), intent(in)::esqua
integer*4,dimension(1:nifai,1:nafac), intent(in)::intfac
real*8::uchar(1:nq)
real*8,dimension(1:ngeel,1:nsize)::geoel
real*8,dimension(1:3,1:ngesgt,1:ntria+nbfac),intent(in)::gesgt0
real*8,dimension(1:3,1:ngesgq,1:nquad),      intent(in)::gesgq0
real*8,dimension(1:ndegr,1:nq,1:nsize),intent(inout)::unkno
real*8,dimension(1:ndegr,1:nq,1:nsize),intent(inout)::unkaw
real*8,dimension(1:ndegr,1:4,1:nsize),intent(inout)::unkgd
real*8,dimension(1:3,1:3,ngstrnf+ngausdq, nquad)::strnq_devtp
real*8,dimension(1:ndegr,1:nq,1:ncell),intent(out)::rhsel
real*8,dimension(1:ndegr,1:4,1:ncell),intent(out)::rhsgd
real*8,dimension(1:ngefa,1:nafac)::geofa
real*8,dimension(1:ndimn,1:npoin),  intent(in)::coold
real*8,dimension(1:ndimn,1:npoin),  intent(inout)::coora
real*8,dimension(1:ndimn,1:npoin)::coord
real*8,dimension(nmatr,ncell),intent(inout)::amatr
real*8,dimension(1:ndimn,1:npoin)   ::ustar
integer::itime

type(c_ptr), save :: ctx = c_null_ptr
integer, save :: itime_ctx = -1
type(lag2d_config) :: cfg
real(c_double), allocatable :: xdum(:,:)

if (nrz /= 0) then
  write(*, '(a)') 'lag2d: the right-hand-side drop-in covers the planar paths (nrz = 0)'
  error stop 1
end if
! a new run (the driver restarting at step 1) gets a new context
if (itime == 1 .and. itime_ctx > 1 .and. c_associated(ctx)) then
  call lag2d_destroy(ctx)
  ctx = c_null_ptr
end if
itime_ctx = itime
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
end if
call lag2d_require(lag2d_upload(ctx, coord, coold, geoel, amatr, unkno, ustar), 'upload')
call lag2d_require(lag2d_rhs(ctx, rhsel), 'rhs')
allocate(xdum(ndimn, npoin))
call lag2d_require(lag2d_download(ctx, xdum, geoel, unkno, ustar), 'download')
deallocate(xdum)
rhsgd = 0.d0
end subroutine getrhsdg_lag_mcnew
