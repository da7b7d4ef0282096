! lbm_dropin.f90 -- drop-in replacements for the hot-path subroutines of the
! forced-turbulence particle code (3DBox_Forced_HI_Turbulence_Moving_Particles_
! LBM/, "BOX/"): same names, same argument-less interfaces, same module state
! (use var_inc).  Link this file instead of collision.f90, streaming.f90 and
! the bead routines of partlib.f90 / FORCING of gaussian.f90; main.f90,
! para.f90 and var_inc.f90 stay as they are.  Build with the reference's
! flags (default REAL of 8 bytes, e.g. amdflang -fdefault-real-8) and link
! libmfc_lbm.so.
!
! The whole box lives on one GPU (nprocY = nprocZ = 1, so ly = ny, lz = nz).
! Device state is authoritative; the host arrays of var_inc are refreshed
! exactly when the unchanged driver reads them (BOX/main.f90:79-236):
!   initpop ............ f (and rho = 0) after the equilibrium is built
!   rhoupdat ........... rho every call (the driver forms rhoerr on the host);
!                        f as well when that residual ends the pre-relaxation
!   macrovar ........... rho, ux, uy, uz before the time loop and at the
!                        diag / statistc / rmsstat / sijstat03 cadences;
!                        particle arrays at the outputpart cadence; f,
!                        particles and the forcing-process state at the
!                        last step (savecntdflow / savecntdpart / input_outputf)
! and pushed to the device when the context is created (restart: f, ux..uz,
! particles with the integrator history that savecntdpart / loadcntdpart
! carry) and at the first FORCING (the forcing-process state).  After the
! release the node maps ibnodes / isnodes follow the device at the macro
! cadences above, and ibnodes0 / isnodes0 at the last step (diag, rmsstat,
! sijstat03 mask by them; savecntdpart writes the previous pair).
! Precision and storage are read from the environment when the context is
! created: MFC_PRECISION = fp64 (default) | fp32 | fp32x (FP32 storage, FP64
! collision); MFC_STORAGE = ab (default, two buffers) | aa (one buffer
! updated in place); MFC_BB_ORDER = reference (default: the bounce-back
! links in this loop's order, each value stored at once) | snapshot (every
! link against one post-collision field; the two differ only for links whose
! node is solid for another sphere).
! A library error prints 'fault: <routine> (<code>) <message>' as the
! reference's own fatal checks do, then ends the run with a nonzero status.
      module mfc_dropin_state
      use, intrinsic :: iso_c_binding
      use mfc_lbm_c
      implicit none
      type(c_ptr), save :: ctx = c_null_ptr
      logical, save :: forcing_started = .false.
      logical, save :: created = .false.
      ! the device holds node maps (beads_links has run on this context)
      logical, save :: maps_built = .false.
      contains

      ! value of environment variable `name`, or `dflt` when it is unset
      function env_choice(name, dflt) result(v)
        character(len=*), intent(in) :: name, dflt
        character(len=16) :: v
        integer :: n, st
        call get_environment_variable(name, v, n, st)
        if (st /= 0 .or. n == 0) v = dflt
      end function env_choice

      subroutine mfc_check(rc, where)
        integer(c_int), intent(in) :: rc
        character(len=*), intent(in) :: where
        character(kind=c_char), pointer :: msg(:)
        character(len=256) :: text
        integer :: k
        if (rc == MFC_OK) return
        text = ' '
        call c_f_pointer(mfc_last_error(ctx), msg, [256])
        do k = 1, 256
          if (msg(k) == c_null_char) exit
          text(k:k) = msg(k)
        end do
        write(*,*) 'fault: ', where, ' (', rc, ') ', trim(text)
        error stop 1
      end subroutine mfc_check

      ! mfc_upload_macro takes nullable pointers; the dummies are TARGETs so
      ! that C_LOC is valid for the duration of the call
      subroutine mfc_upload_fields(rho, ux, uy, uz, with_rho, where)
        real(c_double), target, intent(in) :: rho(*), ux(*), uy(*), uz(*)
        logical, intent(in) :: with_rho
        character(len=*), intent(in) :: where
        type(c_ptr) :: prho
        prho = c_null_ptr
        if (with_rho) prho = c_loc(rho)
        call mfc_check(mfc_upload_macro(ctx, prho, c_loc(ux), c_loc(uy),   &
                       c_loc(uz)), where)
      end subroutine mfc_upload_fields

      end module mfc_dropin_state
!===========================================================================

      subroutine mfc_dropin_create
      use mfc_dropin_state
      use var_inc
      implicit none
      type(mfc_params) :: p
      integer(c_int) :: rc
      if (created) return
      if (ly /= ny .or. lz /= nz) then
        write(*,*) 'fault: the GPU drop-in holds the whole box on one rank',  &
                   ' (nprocY = nprocZ = 1)'
        error stop 1
      end if
      p%nx = nx
      p%ny = ny
      p%nz = nz
      select case (trim(env_choice('MFC_PRECISION', 'fp64')))
      case ('fp64')
        p%precision = MFC_PREC_FP64
      case ('fp32')
        p%precision = MFC_PREC_FP32
      case ('fp32x')
        p%precision = MFC_PREC_FP32_F64MATH
      case default
        write(*,*) 'fault: MFC_PRECISION must be fp64, fp32 or fp32x'
        error stop 1
      end select
      p%npart_max = 0
      if (ipart) p%npart_max = npart
      p%walls_y = 0
      p%iseedf = iseedf
      p%irelease = irelease
      select case (trim(env_choice('MFC_STORAGE', 'ab')))
      case ('ab')
        p%storage = MFC_STORE_AB
      case ('aa')
        p%storage = MFC_STORE_AA
      case default
        write(*,*) 'fault: MFC_STORAGE must be ab or aa'
        error stop 1
      end select
      p%s1 = s1
      p%s2 = s2
      p%s4 = s4
      p%s9 = s9
      p%s10 = s10
      p%s13 = s13
      p%s16 = s16
      p%omegepsl = omegepsl
      p%omegepslj = omegepslj
      p%omegxx = omegxx
      p%ww0 = ww0
      p%ww1 = ww1
      p%ww2 = ww2
      p%rho0 = rho0
      p%rad = rad
      p%rhopart = rhopart
      p%mingap = mingap
      p%volp = volp
      p%amp = amp
      p%aip = aip
      p%rhog = rhog
      p%stf0 = stf0
      p%stf1 = stf1
      p%gscale = gscale
      p%vscale = vscale
      p%tscale = tscale
      p%xkf = sqrt(8.0d0)
      p%pi2 = pi2
      p%body_force = 0.0d0
      p%particle_force = 0.0d0
      ! the periodic box has no walls; the channel's defaults
      p%mingap_w = mingap
      p%stf0_w = stf0
      p%stf1_w = stf1
      rc = mfc_create(p, ctx)
      if (rc /= MFC_OK) then
        write(*,*) 'fault: mfc_create failed (', rc, ')'
        error stop 1
      end if
      created = .true.
      select case (trim(env_choice('MFC_BB_ORDER', 'reference')))
      case ('reference')
        call mfc_check(mfc_set_bb_order(ctx, MFC_BB_REFERENCE), 'set_bb_order')
      case ('snapshot')
        call mfc_check(mfc_set_bb_order(ctx, MFC_BB_SNAPSHOT), 'set_bb_order')
      case default
        write(*,*) 'fault: MFC_BB_ORDER must be reference or snapshot'
        error stop 1
      end select
      ! whatever the host holds now is the state (restart: loaded f, ux..uz)
      call mfc_check(mfc_upload_f(ctx, f), 'upload f')
      call mfc_upload_fields(rho, ux, uy, uz, .true., 'upload rho/u')
      if (ipart) then
        call mfc_check(mfc_prt_init(ctx, int(npart, c_int32_t)), 'prt_init')
        call mfc_push_particles
      end if
      end subroutine mfc_dropin_create
!===========================================================================

      subroutine mfc_push_particles
      use mfc_dropin_state
      use var_inc
      implicit none
      call mfc_check(mfc_prt_set(ctx, MFC_PF_YPGLB, ypglb), 'set ypglb')
      call mfc_check(mfc_prt_set(ctx, MFC_PF_YPGLBP, ypglbp), 'set ypglbp')
      call mfc_check(mfc_prt_set(ctx, MFC_PF_WP, wp), 'set wp')
      call mfc_check(mfc_prt_set(ctx, MFC_PF_WPP, wpp), 'set wpp')
      call mfc_check(mfc_prt_set(ctx, MFC_PF_OMGP, omgp), 'set omgp')
      call mfc_check(mfc_prt_set(ctx, MFC_PF_OMGPP, omgpp), 'set omgpp')
      ! integrator history of beads_move (BOX/partlib.f90:2332-2365), held
      ! per rank-local slot; one rank keeps slot ip = particle ip
      call mfc_check(mfc_prt_set(ctx, MFC_PF_FORCEPP, forcepp), 'set forcepp')
      call mfc_check(mfc_prt_set(ctx, MFC_PF_TORQPP, torqpp), 'set torqpp')
      call mfc_check(mfc_prt_set(ctx, MFC_PF_DWDT, dwdt), 'set dwdt')
      call mfc_check(mfc_prt_set(ctx, MFC_PF_DOMGDT, domgdt), 'set domgdt')
      call mfc_check(mfc_prt_set(ctx, MFC_PF_THETAP, thetap), 'set thetap')
      end subroutine mfc_push_particles
!===========================================================================

      subroutine mfc_pull_particles
      use mfc_dropin_state
      use var_inc
      implicit none
      call mfc_check(mfc_prt_get(ctx, MFC_PF_YPGLB, ypglb), 'get ypglb')
      call mfc_check(mfc_prt_get(ctx, MFC_PF_YPGLBP, ypglbp), 'get ypglbp')
      call mfc_check(mfc_prt_get(ctx, MFC_PF_WP, wp), 'get wp')
      call mfc_check(mfc_prt_get(ctx, MFC_PF_WPP, wpp), 'get wpp')
      call mfc_check(mfc_prt_get(ctx, MFC_PF_OMGP, omgp), 'get omgp')
      call mfc_check(mfc_prt_get(ctx, MFC_PF_OMGPP, omgpp), 'get omgpp')
      call mfc_check(mfc_prt_get(ctx, MFC_PF_FHIP, fHIp), 'get fHIp')
      call mfc_check(mfc_prt_get(ctx, MFC_PF_FLUBP, flubp), 'get flubp')
      call mfc_check(mfc_prt_get(ctx, MFC_PF_TORQP, torqp), 'get torqp')
      call mfc_check(mfc_prt_get(ctx, MFC_PF_FORCEPP, forcepp), 'get forcepp')
      call mfc_check(mfc_prt_get(ctx, MFC_PF_TORQPP, torqpp), 'get torqpp')
      call mfc_check(mfc_prt_get(ctx, MFC_PF_DWDT, dwdt), 'get dwdt')
      call mfc_check(mfc_prt_get(ctx, MFC_PF_DOMGDT, domgdt), 'get domgdt')
      call mfc_check(mfc_prt_get(ctx, MFC_PF_THETAP, thetap), 'get thetap')
      yp(:, 1:npart) = ypglb(:, 1:npart)
      end subroutine mfc_pull_particles
!===========================================================================

! Ends the device context (frees the GPU state); the next wrapper call
! creates a new one from the host arrays of var_inc, as a restarted run
! does.  No counterpart in the reference, where the process exit does this.
      subroutine mfc_dropin_finish
      use mfc_dropin_state
      implicit none
      integer(c_int) :: rc
      if (.not. created) return
      rc = mfc_destroy(ctx)
      ctx = c_null_ptr
      created = .false.
      forcing_started = .false.
      maps_built = .false.
      if (rc /= MFC_OK) then
        write(*,*) 'fault: mfc_destroy failed (', rc, ')'
        error stop 1
      end if
      end subroutine mfc_dropin_finish
!===========================================================================

      SUBROUTINE initpop
      use mfc_dropin_state
      use var_inc
      implicit none
      call mfc_dropin_create
      call mfc_upload_fields(rho, ux, uy, uz, .false., 'upload u')
      call mfc_check(mfc_initpop(ctx), 'initpop')
      rho = 0.0
      call mfc_check(mfc_download_f(ctx, f), 'download f')
      END SUBROUTINE initpop
!===========================================================================

      subroutine rhoupdat
      use mfc_dropin_state
      use var_inc
      implicit none
      real(c_double) :: dmax
      call mfc_dropin_create
      call mfc_check(mfc_rhoupdat(ctx, dmax), 'rhoupdat')
      call mfc_check(mfc_download_macro(ctx, rho, ux, uy, uz), 'download rho')
      ! the pre-relaxation ends on this residual (main.f90:93-104): the
      ! driver then saves f, so hand it over now
      if (maxval(abs(rho - rhop)) <= rhoepsl .or. istep > 15000) then
        call mfc_check(mfc_download_f(ctx, f), 'download f')
      end if
      end subroutine rhoupdat
!===========================================================================

      SUBROUTINE FORCING
      use mfc_dropin_state
      use var_inc
      implicit none
      call mfc_dropin_create
      if (.not. forcing_started) then
        ! the forcing-process state held by var_inc (para or input_outputf)
        call mfc_check(mfc_set_forcing_state(ctx, int(iseedf, c_int32_t),  &
                       int(iyf, c_int32_t), ivf, b1r, b2r, b3r, a1r, a2r, a3r), &
                       'forcing state')
        forcing_started = .true.
      end if
      call mfc_check(mfc_forcing(ctx), 'FORCING')
      END SUBROUTINE FORCING
!===========================================================================

      subroutine collision_MRT
      use mfc_dropin_state
      implicit none
      call mfc_dropin_create
      call mfc_check(mfc_collision_mrt(ctx), 'collision_MRT')
      end subroutine collision_MRT
!===========================================================================

      subroutine streaming
      use mfc_dropin_state
      implicit none
      call mfc_dropin_create
      call mfc_check(mfc_streaming(ctx), 'streaming')
      end subroutine streaming
!===========================================================================

      subroutine macrovar
      use mfc_dropin_state
      use var_inc
      implicit none
      logical :: want_macro, want_part, last
      integer(c_int32_t) :: i1, i2
      call mfc_dropin_create
      call mfc_check(mfc_macrovar(ctx), 'macrovar')
      last = forcing_started .and. istep == istep0 + nsteps
      want_macro = (.not. forcing_started) .or. last                      &
                   .or. mod(istep - 1, ndiag) == 0 .or. mod(istep - 1, nstat) == 0 &
                   .or. mod(istep, nstat) == 0 .or. mod(istep, nsij) == 0
      want_part = ipart .and. (last .or. mod(istep, npartout) == 0)
      if (want_macro) then
        call mfc_check(mfc_download_macro(ctx, rho, ux, uy, uz), 'download macro')
        call mfc_check(mfc_download_force(ctx, force_realx, force_realy,   &
                       force_realz), 'download force')
      end if
      if (want_part) call mfc_pull_particles
      if (ipart .and. maps_built .and. want_macro) then
        call mfc_check(mfc_get_maps(ctx, 0_c_int32_t, ibnodes, isnodes),   &
                       'get maps')
        if (last) call mfc_check(mfc_get_maps(ctx, 1_c_int32_t, ibnodes0, &
                                 isnodes0), 'get previous maps')
      end if
      if (last) then
        call mfc_check(mfc_download_f(ctx, f), 'download f')
        call mfc_check(mfc_get_forcing_state(ctx, i1, i2, ivf, b1r, b2r, b3r, &
                       a1r, a2r, a3r), 'forcing state')
        iseedf = i1
        iyf = i2
      end if
      end subroutine macrovar
!===========================================================================

      subroutine initpartpos
      use mfc_dropin_state
      use var_inc
      implicit none
      integer(c_int32_t) :: seed
      call mfc_dropin_create
      seed = iseedp
      call mfc_check(mfc_initpartpos(ctx, seed, 1_c_int32_t), 'initpartpos')
      iseedp = seed
      call mfc_check(mfc_prt_get(ctx, MFC_PF_YPGLB, ypglb), 'get ypglb')
      yp(:, 1:npart) = ypglb(:, 1:npart)
      nps = npart
      end subroutine initpartpos
!===========================================================================

      subroutine initpartvel
      use mfc_dropin_state
      use var_inc
      implicit none
      call mfc_dropin_create
      call mfc_check(mfc_initpartvel(ctx), 'initpartvel')
      call mfc_check(mfc_prt_get(ctx, MFC_PF_WP, wp), 'get wp')
      end subroutine initpartvel
!===========================================================================

      subroutine initpartomg
      use mfc_dropin_state
      use var_inc
      implicit none
      call mfc_dropin_create
      call mfc_check(mfc_initpartomg(ctx), 'initpartomg')
      omgp = 0.0
      end subroutine initpartomg
!===========================================================================

      subroutine beads_links
      use mfc_dropin_state
      implicit none
      call mfc_dropin_create
      call mfc_check(mfc_beads_links(ctx), 'beads_links')
      maps_built = .true.
      end subroutine beads_links
!===========================================================================

      subroutine beads_collision
      use mfc_dropin_state
      implicit none
      call mfc_check(mfc_beads_collision(ctx), 'beads_collision')
      end subroutine beads_collision
!===========================================================================

      subroutine beads_lubforce
      use mfc_dropin_state
      implicit none
      call mfc_check(mfc_beads_lubforce(ctx), 'beads_lubforce')
      end subroutine beads_lubforce
!===========================================================================

      subroutine beads_move
      use mfc_dropin_state
      use var_inc
      implicit none
      call mfc_check(mfc_beads_move(ctx, int(istep, c_int32_t)), 'beads_move')
      end subroutine beads_move
!===========================================================================

      subroutine beads_redistribute
      use mfc_dropin_state
      use var_inc
      implicit none
      ! the overlap check is fatal, as in the reference (partlib.f90:2641-2646)
      call mfc_check(mfc_beads_redistribute(ctx, int(istep, c_int32_t)),  &
                     'in beads_redis')
      end subroutine beads_redistribute
!===========================================================================

      subroutine beads_filling
      use mfc_dropin_state
      implicit none
      call mfc_check(mfc_beads_filling(ctx), 'beads_filling')
      end subroutine beads_filling
!===========================================================================
