! mfc_lbm_c.f90 -- ISO_C_BINDING interface to libmfc_lbm (include/mfc_lbm.h).
!
! One interface per C entry point; the derived type mfc_params mirrors the
! C struct field for field (eight int32 then doubles, so no padding).
      module mfc_lbm_c
      use, intrinsic :: iso_c_binding
      implicit none

      integer(c_int), parameter :: MFC_OK = 0, MFC_ERR_OVERLAP = 4,     &
                                   MFC_ERR_CAPACITY = 5
      integer(c_int), parameter :: MFC_PREC_FP64 = 0, MFC_PREC_FP32 = 1, &
                                   MFC_PREC_FP32_F64MATH = 2
      integer(c_int), parameter :: MFC_STORE_AB = 0, MFC_STORE_AA = 1
      integer(c_int), parameter :: MFC_BB_SNAPSHOT = 0, MFC_BB_REFERENCE = 1
      integer(c_int), parameter :: MFC_PF_YPGLB = 0, MFC_PF_YPGLBP = 1, &
        MFC_PF_WP = 2, MFC_PF_WPP = 3, MFC_PF_OMGP = 4, MFC_PF_OMGPP = 5,&
        MFC_PF_FHIP = 6, MFC_PF_FLUBP = 7, MFC_PF_TORQP = 8,            &
        MFC_PF_FORCEPP = 9, MFC_PF_TORQPP = 10, MFC_PF_DWDT = 11,       &
        MFC_PF_DOMGDT = 12, MFC_PF_THETAP = 13

      type, bind(C) :: mfc_params
        integer(c_int32_t) :: nx, ny, nz, precision, npart_max, walls_y
        integer(c_int32_t) :: iseedf, irelease, storage
        real(c_double) :: s1, s2, s4, s9, s10, s13, s16
        real(c_double) :: omegepsl, omegepslj, omegxx
        real(c_double) :: ww0, ww1, ww2, rho0
        real(c_double) :: rad, rhopart, mingap
        real(c_double) :: volp, amp, aip, rhog
        real(c_double) :: stf0, stf1, gscale
        real(c_double) :: vscale, tscale, xkf, pi2
        real(c_double) :: body_force(3)
        real(c_double) :: particle_force(3)
        real(c_double) :: mingap_w, stf0_w, stf1_w
      end type mfc_params

      interface
        integer(c_int) function mfc_create(p, ctx) bind(C, name="mfc_create")
          import :: c_int, c_ptr, mfc_params
          type(mfc_params), intent(in) :: p
          type(c_ptr), intent(out) :: ctx
        end function
        integer(c_int) function mfc_destroy(ctx) bind(C, name="mfc_destroy")
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
        end function
        type(c_ptr) function mfc_last_error(ctx) bind(C, name="mfc_last_error")
          import :: c_ptr
          type(c_ptr), value :: ctx
        end function
        integer(c_int) function mfc_sync(ctx) bind(C, name="mfc_sync")
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
        end function
        integer(c_int) function mfc_upload_f(ctx, f) bind(C, name="mfc_upload_f")
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: ctx
          real(c_double), intent(in) :: f(*)
        end function
        integer(c_int) function mfc_download_f(ctx, f) bind(C, name="mfc_download_f")
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: ctx
          real(c_double), intent(out) :: f(*)
        end function
        integer(c_int) function mfc_upload_macro(ctx, rho, ux, uy, uz)   &
            bind(C, name="mfc_upload_macro")
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx, rho, ux, uy, uz
        end function
        integer(c_int) function mfc_download_macro(ctx, rho, ux, uy, uz) &
            bind(C, name="mfc_download_macro")
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: ctx
          real(c_double), intent(out) :: rho(*), ux(*), uy(*), uz(*)
        end function
        integer(c_int) function mfc_download_force(ctx, fx, fy, fz)      &
            bind(C, name="mfc_download_force")
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: ctx
          real(c_double), intent(out) :: fx(*), fy(*), fz(*)
        end function
        integer(c_int) function mfc_get_forcing_state(ctx, iseedf, iyf, ivf, &
            b1r, b2r, b3r, a1r, a2r, a3r) bind(C, name="mfc_get_forcing_state")
          import :: c_int, c_ptr, c_int32_t, c_double
          type(c_ptr), value :: ctx
          integer(c_int32_t), intent(out) :: iseedf, iyf, ivf(32)
          real(c_double), intent(out) :: b1r(*), b2r(*), b3r(*), a1r(*), a2r(*), a3r(*)
        end function
        integer(c_int) function mfc_set_forcing_state(ctx, iseedf, iyf, ivf, &
            b1r, b2r, b3r, a1r, a2r, a3r) bind(C, name="mfc_set_forcing_state")
          import :: c_int, c_ptr, c_int32_t, c_double
          type(c_ptr), value :: ctx
          integer(c_int32_t), value :: iseedf, iyf
          integer(c_int32_t), intent(in) :: ivf(32)
          real(c_double), intent(in) :: b1r(*), b2r(*), b3r(*), a1r(*), a2r(*), a3r(*)
        end function
        integer(c_int) function mfc_initpop(ctx) bind(C, name="mfc_initpop")
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
        end function
        integer(c_int) function mfc_forcing(ctx) bind(C, name="mfc_forcing")
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
        end function
        integer(c_int) function mfc_collision_mrt(ctx) bind(C, name="mfc_collision_mrt")
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
        end function
        integer(c_int) function mfc_streaming(ctx) bind(C, name="mfc_streaming")
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
        end function
        integer(c_int) function mfc_rhoupdat(ctx, maxdiff) bind(C, name="mfc_rhoupdat")
          import :: c_int, c_ptr, c_double
          type(c_ptr), value :: ctx
          real(c_double), intent(out) :: maxdiff
        end function
        integer(c_int) function mfc_macrovar(ctx) bind(C, name="mfc_macrovar")
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
        end function
        integer(c_int) function mfc_prt_init(ctx, npart) bind(C, name="mfc_prt_init")
          import :: c_int, c_ptr, c_int32_t
          type(c_ptr), value :: ctx
          integer(c_int32_t), value :: npart
        end function
        integer(c_int) function mfc_prt_set(ctx, field, a) bind(C, name="mfc_prt_set")
          import :: c_int, c_ptr, c_int32_t, c_double
          type(c_ptr), value :: ctx
          integer(c_int32_t), value :: field
          real(c_double), intent(in) :: a(*)
        end function
        integer(c_int) function mfc_prt_get(ctx, field, a) bind(C, name="mfc_prt_get")
          import :: c_int, c_ptr, c_int32_t, c_double
          type(c_ptr), value :: ctx
          integer(c_int32_t), value :: field
          real(c_double), intent(out) :: a(*)
        end function
        integer(c_int) function mfc_initpartpos(ctx, iseedp, mingap_check) &
            bind(C, name="mfc_initpartpos")
          import :: c_int, c_ptr, c_int32_t
          type(c_ptr), value :: ctx
          integer(c_int32_t), intent(inout) :: iseedp
          integer(c_int32_t), value :: mingap_check
        end function
        integer(c_int) function mfc_initpartvel(ctx) bind(C, name="mfc_initpartvel")
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
        end function
        integer(c_int) function mfc_initpartomg(ctx) bind(C, name="mfc_initpartomg")
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
        end function
        integer(c_int) function mfc_beads_links(ctx) bind(C, name="mfc_beads_links")
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
        end function
        integer(c_int) function mfc_beads_collision(ctx) bind(C, name="mfc_beads_collision")
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
        end function
        integer(c_int) function mfc_set_bb_order(ctx, order) bind(C, name="mfc_set_bb_order")
          import :: c_int, c_ptr, c_int32_t
          type(c_ptr), value :: ctx
          integer(c_int32_t), value :: order
        end function
        integer(c_int) function mfc_beads_lubforce(ctx) bind(C, name="mfc_beads_lubforce")
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
        end function
        integer(c_int) function mfc_beads_move(ctx, istep) bind(C, name="mfc_beads_move")
          import :: c_int, c_ptr, c_int32_t
          type(c_ptr), value :: ctx
          integer(c_int32_t), value :: istep
        end function
        integer(c_int) function mfc_beads_redistribute(ctx, istep)         &
            bind(C, name="mfc_beads_redistribute")
          import :: c_int, c_ptr, c_int32_t
          type(c_ptr), value :: ctx
          integer(c_int32_t), value :: istep
        end function
        integer(c_int) function mfc_beads_filling(ctx) bind(C, name="mfc_beads_filling")
          import :: c_int, c_ptr
          type(c_ptr), value :: ctx
        end function
        integer(c_int) function mfc_get_maps(ctx, which, ibn, isn)         &
            bind(C, name="mfc_get_maps")
          import :: c_int, c_ptr, c_int32_t
          type(c_ptr), value :: ctx
          integer(c_int32_t), value :: which
          integer(c_int32_t), intent(out) :: ibn(*), isn(*)
        end function
        integer(c_int) function mfc_step(ctx, istep, nsteps) bind(C, name="mfc_step")
          import :: c_int, c_ptr, c_int32_t
          type(c_ptr), value :: ctx
          integer(c_int32_t), value :: istep, nsteps
        end function
      end interface

      end module mfc_lbm_c
