! var_inc.f90 -- test stand-in for the driver's module var_inc, declaring the
! names the drop-in wrappers and the test driver use, with the same types and
! array shapes as the reference module (BOX/var_inc.f90:1-121) at a small box:
! 32^3 cells, 3 spheres of radius 4, release at step 2.  Build with default
! REAL of 8 bytes, as the reference is built.
      module var_inc
      implicit none

      integer,parameter:: nx = 32, ny = nx, nz = nx
      integer,parameter:: lx = nx
      integer,parameter:: npop = 19, npart = 3
      integer,parameter:: NTAB = 32
      integer,parameter:: ndiag = 5, nstat = 3, npartout = 4, nsij = 6
      integer,parameter:: irelease = 2

      real,parameter:: rho0 = 1.0, rhopart = 1.5
      real,parameter:: rad = 4.0, mingap = 2.0

      integer istep, istep0, istep00, nsteps
      integer lz, ly
      integer iseedp, msize, nps, iyp
      logical ipart

      real pi2, vscale, tscale
      real omegepsl, omegepslj, omegxx
      real s1, s2, s4, s9, s10, s13, s16
      real ww0, ww1, ww2
      real rhoerr, rhoepsl
      real volp, amp, aip, gscale, rhog
      real stf0, stf1

      integer,dimension(NTAB):: ivp
      integer                :: iseedf, iyf
      integer,dimension(NTAB):: ivf
      real,dimension(6,5,5)  :: a1r,a2r,a3r,b1r,b2r,b3r
      real,allocatable,dimension(:,:,:):: force_realx,force_realy,     &
                                          force_realz

      integer,allocatable,dimension(:,:,:):: ibnodes, ibnodes0
      integer,allocatable,dimension(:,:,:):: isnodes, isnodes0

      real,allocatable,dimension(:,:,:,:):: f
      real,allocatable,dimension(:,:,:):: rho, rhop
      real,allocatable,dimension(:,:,:):: ux, uy, uz
      real,allocatable,dimension(:,:):: yp, ypglb, ypglbp
      real,allocatable,dimension(:,:):: wp, wpp
      real,allocatable,dimension(:,:):: omgp, omgpp
      real,allocatable,dimension(:,:):: dwdt, domgdt
      real,allocatable,dimension(:,:):: forcepp
      real,allocatable,dimension(:,:):: torqp, torqpp
      real,allocatable,dimension(:,:):: thetap
      real,allocatable,dimension(:,:):: fHIp
      real,allocatable,dimension(:,:):: flubp

      end module var_inc
