! ref_driver.f90 -- test-infrastructure driver for the reference's own
! var_inc / para / collision sources (3DBox_Forced_HI_Turbulence_Moving_
! Particles_LBM/var_inc.f90, para.f90, collision.f90), compiled unchanged
! with amdflang -fdefault-real-8 by oracle/build_ref.sh.  It runs one of the
! reference routines on a state read from a raw binary file and writes the
! result, so the C oracle can be checked against the reference itself.
!
! Input file (stream, native endianness):
!   int  mode      1 = collision_MRT, 2 = macrovar, 3 = rhoupdat, 4 = para only
!   int  ly, lz    local slab extents (lx = nx of var_inc)
!   int  np        number of particles carried for macrovar (<= npart)
!   real f(0:18,lx,ly,lz), rho, ux, uy, uz, force_realx/y/z (lx,ly,lz)
!   int  ibnodes(lx,ly,lz), isnodes(lx,ly,lz)
!   real ypglb(3,np), wp(3,np), omgp(3,np)
! Output file: f, rho, ux, uy, uz, then the para scalars
!   tau, s1, s2, s4, s9, s10, s13, s16, omegepsl, omegepslj, omegxx,
!   ww0, ww1, ww2, visc, vscale, tscale, pi2.
      program ref_driver
      use var_inc
      implicit none
      integer :: mode, nly, nlz, np
      character(len=512) :: fin, fout

      call get_command_argument(1, fin)
      call get_command_argument(2, fout)

      myid = 0
      nproc = 8
      call para

      open(unit=11, file=trim(fin), access='stream', form='unformatted', &
           status='old')
      read(11) mode, nly, nlz, np
      ly = nly
      lz = nlz
      lyext = 0
      indy = 0
      indz = 0
      ipart = (mode == 2)
      msize = npart
      call allocarray

      read(11) f, rho, ux, uy, uz, force_realx, force_realy, force_realz
      read(11) ibnodes
      if (ipart) then
        read(11) isnodes
        ypglb = 0.0
        wp = 0.0
        omgp = 0.0
        if (np > 0) then
          read(11) ypglb(:, 1:np), wp(:, 1:np), omgp(:, 1:np)
        end if
      end if
      close(11)

      select case (mode)
      case (1)
        call collision_MRT
      case (2)
        call macrovar
      case (3)
        call rhoupdat
      case default
      end select

      open(unit=12, file=trim(fout), access='stream', form='unformatted', &
           status='replace')
      write(12) f, rho, ux, uy, uz
      write(12) tau, s1, s2, s4, s9, s10, s13, s16, omegepsl, omegepslj, &
                omegxx, ww0, ww1, ww2, visc, vscale, tscale, pi2
      close(12)
      end program ref_driver
