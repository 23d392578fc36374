! Dump driver for the reference solver (checker only, never shipped).
!
! Written for this project.  It performs the reference's own Lagrangian
! initialisation sequence (dgflo2d.f90:165-335, nmeth=2, ncurv<=1) by calling
! the reference subroutines, optionally overrides the initial DOFs with values
! read from inject.bin, and then writes raw little-endian stream files:
!   init_<name>.bin   state after initialisation
!   rhs_<name>.bin    one call of getrhsdg_lag_mcnew on the initial state
!                     (Getrhslag.f90:1747), after getcellvol_lag (:1620)
!   s<k>_<name>.bin   state after k calls of getsolvdg_lag (dgflo2d.f90:372),
!                     with the driver's copy-back (dgflo2d.f90:408-422)
! Control file refdump.in (list-directed): nsteps  every  inject
program refdump
use constant
implicit none
integer*4,allocatable::inpoel(:,:),intfac(:,:),bface(:,:)
integer,allocatable::ipqua(:,:),iptri(:,:),ltelem(:,:),esqua(:,:),estri(:,:)
integer,allocatable::esuv1(:),esuv2(:)
real*8,allocatable::coord(:,:),cooro(:,:),coora(:,:),coold(:,:),geofa(:,:),geoel(:,:)
real*8,allocatable::amatr(:,:),amagd(:,:),ustar(:,:),uchar(:)
real*8,allocatable::unkno(:,:,:),unold(:,:,:),unkaw(:,:,:),rhsel(:,:,:),rhsgd(:,:,:)
real*8,allocatable::unkgd(:,:,:),unogd(:,:,:),gesgq0(:,:,:),gesgt0(:,:,:),gflag(:,:)
real*8,allocatable::strnq_devtp(:,:,:,:)
real*8,allocatable::w_geoel(:,:),w_unkno(:,:,:),w_ustar(:,:),w_unkgd(:,:,:),w_coora(:,:),w_amatr(:,:)
real*8,allocatable::w_unkaw(:,:,:)
real*8::sc(4)
integer::nsteps,every,inject,itime,ie,iq
character(len=16)::tag

open(21,file='refdump.in')
read(21,*) nsteps, every, inject
close(21)

call preprocess1
if(nmeth.ne.2 .or. ncurv.gt.1) stop 'refdump: Lagrangian ncurv<=1 only'
if(ncurv.eq.1)then
  ngesgq = 16; nsubq = 4; nsubt = 4; ngstrnf = 9
  if(nfint.eq.5.or.nfint.eq.10) ngstrnf = 17
else
  ngstrnf = 4
endif
allocate(intfac(nifai,nafac),inpoel(nvtri,ntria),iptri(nvtri,ntria),ipqua(nvqua,nquad),bface(nbfai,nbfac))
allocate(coord(ndimn,npoin),coold(ndimn,npoin),coora(ndimn,npoin),cooro(ndimn,npoin),ustar(ndimn,npoin))
allocate(geoel(ngeel,nsize),geofa(ngefa,nafac))
allocate(unkno(ndegr,nq,nsize),unold(ndegr,nq,nsize),unkaw(ndegr,nq,nsize),rhsel(ndegr,nq,ncell))
allocate(uchar(nq),ltelem(3,ncell),estri(nftri,ntria),esqua(nfqua,nquad),esuv1(npoin1),esuv2(npoin2))
allocate(unkgd(ndegr,4,nsize),unogd(ndegr,4,nsize),rhsgd(ndegr,4,ncell))
allocate(gesgq0(3,ngesgq,nquad),gesgt0(3,ngesgt,ntria+nbfac),gflag(ngflg,nbfac))
allocate(strnq_devtp(3,3,ngstrnf+ngausdq,nquad))
slpdu = (gamlg+1.d0)/2.d0
if(nrz.eq.1.or.nrz.eq.2)then
  alfrz = 1.d0; tfcus = 6.726842539779d-3
else
  alfrz = 0.d0; tfcus = 7.26483157d-3
endif
ngsft = ngausf*3; ngsfq = ngausf*4
tacru = 0.d0; nofile = -1
geoel = 0.d0; unkno = 0.d0

call preprocess2(intfac,inpoel,iptri,ipqua,geofa,geoel,bface,coord,esuv1,esuv2)
call getgeofa(intfac, geofa, coord)
call getgeoel_initial_lag(iptri, ipqua, geoel, coord)
if(npoly==2) call getgeoelho_initial_lag(iptri, ipqua, geoel, coord)
call inifield_mcnew(unkno,uchar,geoel,coord,inpoel,iptri,ipqua)
ustar = 0.d0
coold = coord
if(npoly==0)then
  nmatr = 1
elseif(npoly==1)then
  nmatr = 4
else
  nmatr = 16
endif
allocate(amatr(nmatr,ncell),amagd(nmatr,ncell))
call getamatr_initial_lag(unkno,amatr,geoel,coord,iptri,ipqua)
call getamatr_lag_dg(amagd,geoel,coord,iptri,ipqua)
if(npoly.ge.1) call getunkno_initial_lag(unkno,amatr,geoel,coord,inpoel,iptri,ipqua)
if(inject.eq.1)then
  open(22,file='inject.bin',access='stream',form='unformatted',status='old')
  read(22) unkno(1:ndegr,1:nq,1:ncell)
  close(22)
endif
unkgd = 0.d0
unkgd(1,1,:) = 1.d0; unkgd(1,4,:) = 1.d0
strnq_devtp = 0.d0
if(ncurv.eq.0)then
  call getfnds_lag_hybridgd(geoel,gflag,gesgq0,gesgt0,intfac,iptri,ipqua,coord,unkgd)
else
  call getfnds_laghybrid_gd(geoel,gflag,gesgq0,gesgt0,intfac,iptri,ipqua,coord,unkgd)
endif
call getesuel(intfac,ipqua,iptri,estri,esqua)
rklag(1,1:3) = [0.d0, 0.5d0, 0.d0]
rklag(2,1:3) = [0.d0, 0.75d0, 0.3333333333333333d0]
crklg(1:2) = [0.d0, 1.d0]
crklg2(1:3) = [0.d0, 1.d0, 0.5d0]

tag = 'init'
call dumpi(tag,'ipqua',ipqua,size(ipqua))
if(ntria.gt.0) call dumpi(tag,'iptri',iptri,size(iptri))
call dumpi(tag,'bface',bface,size(bface))
call dumpi(tag,'intfac',intfac,size(intfac))
call dumpi(tag,'esuv1',esuv1,size(esuv1))
call dumpi(tag,'esuv2',esuv2,size(esuv2))
call dumpi(tag,'esqua',esqua,size(esqua))
call dumpr(tag,'coord',coord,size(coord))
call dumpr(tag,'geoel',geoel,size(geoel))
call dumpr(tag,'amatr',amatr,size(amatr))
call dumpr(tag,'unkno',unkno,size(unkno))
sc = [dble(ncell), dble(npoin), dble(nbfac), dble(nafac)]
call dumpr(tag,'sizes',sc,4)

! One right-hand side on the initial state, on copies.
allocate(w_geoel(ngeel,nsize),w_unkno(ndegr,nq,nsize),w_ustar(ndimn,npoin),w_unkgd(ndegr,4,nsize))
allocate(w_coora(ndimn,npoin),w_amatr(nmatr,ncell),w_unkaw(ndegr,nq,nsize))
w_geoel = geoel; w_unkno = unkno; w_ustar = ustar; w_unkgd = unkgd; w_coora = coord
w_amatr = amatr; w_unkaw = unkno
call getcellvol_lag(iptri, ipqua, w_geoel, coord)
rkstg = 1
! epsaw of the first step (Getrhslag.f90:1663-1737): the RZ area weighting
! (nrz = 2) reads it in the right-hand side
epsaw = 1.d-1
if(nrz.ge.1.and.ncase.eq.7) epsaw = 1.d0
if(nrz.ge.1.and.ncase.eq.11) epsaw = 1.d-3
if(ncase.eq.-14) epsaw = 1.d-2
cooro = coord
call getrhsdg_lag_mcnew(w_ustar,uchar,bface,w_unkno,w_unkaw,w_unkgd,strnq_devtp,rhsel,rhsgd,intfac,inpoel,iptri,ipqua,&
                        geofa,w_geoel,gesgt0,gesgq0,cooro,coold,w_coora,esuv1,esuv2,ltelem,estri,esqua,w_amatr,1)
tag = 'rhs'
call dumpr(tag,'rhsel',rhsel,size(rhsel))
call dumpr(tag,'ustar',w_ustar,size(w_ustar))
call dumpr(tag,'unkno',w_unkno,size(w_unkno))
call dumpr(tag,'geoel',w_geoel,size(w_geoel))

! Time loop as in the reference driver.
unold = unkno
unogd = unkgd
cooro = coord
do itime = 1, nsteps
  call getsolvdg_lag(intfac,inpoel,iptri,ipqua,bface,cooro,coord,coora,geofa,geoel,gesgt0,gesgq0,uchar,&
                     unkno,unold,unkaw,unkgd,unogd,strnq_devtp,ltelem,estri,esqua,itime,amatr,amagd,ustar,coold,esuv1,esuv2)
  do ie = 1, ncell
    do iq = 1, ndegr
      unkno(iq,1:nq,ie) = unold(iq,1:nq,ie)
    enddo
  enddo
  do ie = 1, ncell
    do iq = 1, ndegr
      unkgd(iq,1:4,ie) = unogd(iq,1:4,ie)
    enddo
  enddo
  coord = cooro
  if(mod(itime,every).eq.0 .or. itime.eq.nsteps .or. abs(tacru-tend)/tend.le.1e-11)then
    write(tag,'(a,i0)') 's', itime
    call dumpr(tag,'unkno',unkno,size(unkno))
    call dumpr(tag,'coord',coord,size(coord))
    call dumpr(tag,'ustar',ustar,size(ustar))
    call dumpr(tag,'geoel',geoel,size(geoel))
    sc = [dtfix, tacru, 0.d0, 0.d0]
    call dumpr(tag,'time',sc,4)
  endif
  if(abs(tacru-tend)/tend.le.1e-11) exit
enddo

contains
  subroutine dumpr(t, nm, a, n)
    character(len=*), intent(in)::t, nm
    integer, intent(in)::n
    real*8, intent(in)::a(n)
    open(23,file=trim(t)//'_'//trim(nm)//'.bin',access='stream',form='unformatted',status='replace')
    write(23) a
    close(23)
  end subroutine dumpr
  subroutine dumpi(t, nm, a, n)
    character(len=*), intent(in)::t, nm
    integer, intent(in)::n
    integer, intent(in)::a(n)
    open(23,file=trim(t)//'_'//trim(nm)//'.bin',access='stream',form='unformatted',status='replace')
    write(23) a
    close(23)
  end subroutine dumpi
end program refdump
