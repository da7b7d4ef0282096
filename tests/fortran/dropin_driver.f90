! dropin_driver.f90 -- runs the driver's call sequence (BOX/main.f90:73-188:
! initpop, rhoupdat, the density pre-relaxation loop, macrovar, then the time
! loop with the particle release at step irelease) through the drop-in
! wrappers, exactly as the unchanged main program calls them.
!
! Reads, from the working directory, dropin_in.nml (namelist /run/ with the
! constants para would set, plus nsteps and a cap on pre-relaxation
! iterations) and dropin_u.bin (ux, uy, uz as raw doubles, x fastest), and
! writes dropin_out.bin: f, rho, ux, uy, uz, then ypglb, wp, omgp, fHIp,
! torqp, flubp, then the forcing seed state (iseedf, iyf) and the number of
! pre-relaxation iterations, all raw (stream access).
      program dropin_driver
      use var_inc
      implicit none
      integer :: nrelax, it, u
      namelist /run/ s1, s2, s4, s9, s10, s13, s16, omegepsl, omegepslj, &
                     omegxx, ww0, ww1, ww2, volp, amp, aip, gscale, rhog,  &
                     stf0, stf1, vscale, tscale, pi2, iseedf, iseedp,      &
                     rhoepsl, nsteps, nrelax

      ly = ny
      lz = nz
      ipart = .true.
      nrelax = 5
      open(newunit=u, file='dropin_in.nml', status='old', action='read')
      read(u, nml=run)
      close(u)
      ! para.f90:281-286: the forcing seed is negated, its shuffle table cleared
      iseedf = -iseedf
      ivf = 0
      iyf = 0
      b1r = 0.0
      b2r = 0.0
      b3r = 0.0
      a1r = 0.0
      a2r = 0.0
      a3r = 0.0

      allocate(f(0:npop-1, lx, ly, lz))
      allocate(rho(lx, ly, lz), rhop(lx, ly, lz))
      allocate(ux(lx, ly, lz), uy(lx, ly, lz), uz(lx, ly, lz))
      allocate(force_realx(lx, ly, lz), force_realy(lx, ly, lz),          &
               force_realz(lx, ly, lz))
      allocate(yp(3, npart), ypglb(3, npart), ypglbp(3, npart))
      allocate(wp(3, npart), wpp(3, npart), omgp(3, npart), omgpp(3, npart))
      allocate(torqp(3, npart), fHIp(3, npart), flubp(3, npart))
      f = 0.0
      rho = 0.0
      ypglb = 0.0
      ypglbp = 0.0
      wp = 0.0
      wpp = 0.0
      omgp = 0.0
      omgpp = 0.0

      open(newunit=u, file='dropin_u.bin', access='stream', status='old', &
           action='read')
      read(u) ux, uy, uz
      close(u)

      call initpop
      call rhoupdat
      istep = 0
      do
        call collision_MRT
        call streaming
        rhop = rho
        call rhoupdat
        rhoerr = maxval(abs(rho - rhop))
        if (rhoerr <= rhoepsl .or. istep + 1 >= nrelax) exit
        istep = istep + 1
      end do
      it = istep + 1
      call macrovar

      istep0 = 0
      do istep = istep0 + 1, istep0 + nsteps
        if (ipart .and. istep == irelease) then
          istep00 = 0
          call initpartpos
          call initpartvel
          call initpartomg
          call beads_links
          istep00 = 1
        end if
        call FORCING
        call collision_MRT
        if (ipart .and. istep >= irelease) call beads_collision
        call streaming
        if (ipart .and. istep >= irelease) then
          call beads_lubforce
          call beads_move
          call beads_redistribute
          call beads_links
          call beads_filling
        end if
        call macrovar
      end do
      ! the loop variable is one past the last step here; macrovar saw the last
      open(newunit=u, file='dropin_out.bin', access='stream',             &
           status='replace', action='write')
      write(u) f, rho, ux, uy, uz
      write(u) ypglb, wp, omgp, fHIp, torqp, flubp
      write(u) iseedf, iyf, it
      close(u)
      write(*,*) 'dropin_driver: done, pre-relaxation iterations ', it
      end program dropin_driver
