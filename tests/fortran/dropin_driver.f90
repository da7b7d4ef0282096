! dropin_driver.f90 -- runs the driver's call sequence (BOX/main.f90:73-188:
! initpop, rhoupdat, the density pre-relaxation loop, macrovar, then the time
! loop with the particle release at step irelease) through the drop-in
! wrappers, exactly as the unchanged main program calls them.
!
! With nsplit > 0 the run stops after step nsplit, keeps only what the
! reference's restart files hold (savecntdflow: f; input_outputf: the
! forcing-process state; savecntdpart: nps, ibnodes0, isnodes0, yp, thetap,
! forcepp, torqpp, dwdt, domgdt, ypglb, wp, omgp; BOX/saveload.f90:181-274,
! 582-618), clears every other array, ends the device context and continues
! to step nsteps through the restart sequence of BOX/main.f90:121-135
! (loadcntdflow, input_outputf, loadcntdpart, beads_links, macrovar), with
! the step counter carried on instead of reset.
!
! Reads, from the working directory, dropin_in.nml (namelist /run/ with the
! constants para would set, plus nsteps, a cap on pre-relaxation iterations
! and nsplit) and dropin_u.bin (ux, uy, uz as raw doubles, x fastest), and
! writes dropin_out.bin: f, rho, ux, uy, uz, then ypglb, wp, omgp, fHIp,
! torqp, flubp, forcepp, torqpp, dwdt, domgdt, thetap, then the forcing seed
! state (iseedf, iyf), the number of pre-relaxation iterations and ibnodes,
! isnodes, ibnodes0, isnodes0, all raw (stream access).
      program dropin_driver
      use var_inc
      implicit none
      integer :: nrelax, it, u, nsplit, nlast
      real, allocatable :: keep_f(:,:,:,:), keep_p(:,:,:)
      integer, allocatable :: keep_m(:,:,:,:)
      integer :: keep_seed(2), keep_iv(NTAB)
      real :: keep_modes(6,5,5,6)
      namelist /run/ s1, s2, s4, s9, s10, s13, s16, omegepsl, omegepslj, &
                     omegxx, ww0, ww1, ww2, volp, amp, aip, gscale, rhog,  &
                     stf0, stf1, vscale, tscale, pi2, iseedf, iseedp,      &
                     rhoepsl, nsteps, nrelax, nsplit

      ly = ny
      lz = nz
      ipart = .true.
      nrelax = 5
      nsplit = 0
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

      ! rank-local particle arrays are (3, msize) as allocarray makes them;
      ! one rank uses the first npart slots
      msize = npart + 2
      allocate(f(0:npop-1, lx, ly, lz))
      allocate(rho(lx, ly, lz), rhop(lx, ly, lz))
      allocate(ux(lx, ly, lz), uy(lx, ly, lz), uz(lx, ly, lz))
      allocate(force_realx(lx, ly, lz), force_realy(lx, ly, lz),          &
               force_realz(lx, ly, lz))
      allocate(ibnodes(lx, ly, lz), isnodes(lx, ly, lz))
      allocate(ibnodes0(lx, ly, lz), isnodes0(lx, ly, lz))
      allocate(yp(3, msize), ypglb(3, npart), ypglbp(3, npart))
      allocate(wp(3, npart), wpp(3, npart), omgp(3, npart), omgpp(3, npart))
      allocate(torqp(3, npart), fHIp(3, npart), flubp(3, npart))
      allocate(forcepp(3, msize), torqpp(3, msize), dwdt(3, msize),      &
               domgdt(3, msize), thetap(3, msize))
      f = 0.0
      rho = 0.0
      ypglb = 0.0
      ypglbp = 0.0
      wp = 0.0
      wpp = 0.0
      omgp = 0.0
      omgpp = 0.0
      ibnodes = -1
      isnodes = -1
      ibnodes0 = -1
      isnodes0 = -1
      call clear_history

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

      nlast = nsteps
      istep0 = 0
      if (nsplit > 0 .and. nsplit < nlast) then
        nsteps = nsplit
        call time_loop
        ! what the restart files keep
        allocate(keep_f(0:npop-1, lx, ly, lz), keep_p(3, npart, 12),      &
                 keep_m(lx, ly, lz, 2))
        keep_f = f
        keep_p(:, :, 1) = yp(:, 1:npart)
        keep_p(:, :, 2) = thetap(:, 1:npart)
        keep_p(:, :, 3) = forcepp(:, 1:npart)
        keep_p(:, :, 4) = torqpp(:, 1:npart)
        keep_p(:, :, 5) = dwdt(:, 1:npart)
        keep_p(:, :, 6) = domgdt(:, 1:npart)
        keep_p(:, :, 7) = ypglb
        keep_p(:, :, 8) = wp
        keep_p(:, :, 9) = omgp
        keep_m(:, :, :, 1) = ibnodes0
        keep_m(:, :, :, 2) = isnodes0
        keep_seed = (/ iseedf, iyf /)
        keep_iv = ivf
        keep_modes(:, :, :, 1) = b1r
        keep_modes(:, :, :, 2) = b2r
        keep_modes(:, :, :, 3) = b3r
        keep_modes(:, :, :, 4) = a1r
        keep_modes(:, :, :, 5) = a2r
        keep_modes(:, :, :, 6) = a3r
        call mfc_dropin_finish
        ! a new process: nothing else survives
        f = 0.0
        rho = 0.0
        ux = 0.0
        uy = 0.0
        uz = 0.0
        force_realx = 0.0
        force_realy = 0.0
        force_realz = 0.0
        ypglb = 0.0
        ypglbp = 0.0
        wp = 0.0
        wpp = 0.0
        omgp = 0.0
        omgpp = 0.0
        fHIp = 0.0
        flubp = 0.0
        torqp = 0.0
        yp = 0.0
        ibnodes = -1
        isnodes = -1
        call clear_history
        ! loadcntdflow, input_outputf(1), loadcntdpart
        f = keep_f
        iseedf = keep_seed(1)
        iyf = keep_seed(2)
        ivf = keep_iv
        b1r = keep_modes(:, :, :, 1)
        b2r = keep_modes(:, :, :, 2)
        b3r = keep_modes(:, :, :, 3)
        a1r = keep_modes(:, :, :, 4)
        a2r = keep_modes(:, :, :, 5)
        a3r = keep_modes(:, :, :, 6)
        nps = npart
        yp(:, 1:npart) = keep_p(:, :, 1)
        thetap(:, 1:npart) = keep_p(:, :, 2)
        forcepp(:, 1:npart) = keep_p(:, :, 3)
        torqpp(:, 1:npart) = keep_p(:, :, 4)
        dwdt(:, 1:npart) = keep_p(:, :, 5)
        domgdt(:, 1:npart) = keep_p(:, :, 6)
        ypglb = keep_p(:, :, 7)
        wp = keep_p(:, :, 8)
        omgp = keep_p(:, :, 9)
        ibnodes0 = keep_m(:, :, :, 1)
        isnodes0 = keep_m(:, :, :, 2)
        call beads_links
        call macrovar
        istep0 = nsplit
        nsteps = nlast - nsplit
      end if
      call time_loop
      open(newunit=u, file='dropin_out.bin', access='stream',             &
           status='replace', action='write')
      write(u) f, rho, ux, uy, uz
      write(u) ypglb, wp, omgp, fHIp, torqp, flubp
      write(u) forcepp(:, 1:npart), torqpp(:, 1:npart), dwdt(:, 1:npart), &
               domgdt(:, 1:npart), thetap(:, 1:npart)
      write(u) iseedf, iyf, it
      write(u) ibnodes, isnodes, ibnodes0, isnodes0
      close(u)
      write(*,*) 'dropin_driver: done, pre-relaxation iterations ', it

      contains

      subroutine clear_history
        forcepp = 0.0
        torqpp = 0.0
        dwdt = 0.0
        domgdt = 0.0
        thetap = 0.0
      end subroutine clear_history

      ! steps istep0 + 1 .. istep0 + nsteps of the time loop
      subroutine time_loop
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
        ! the loop variable is one past the last step; macrovar saw the last
        istep = istep0 + nsteps
      end subroutine time_loop
      end program dropin_driver
