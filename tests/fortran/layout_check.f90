! layout_check.f90 -- prints the size of the Fortran mirror of struct
! mfc_params and the byte offset of every field, one "name offset" pair per
! line, for comparison with the C layout.
      program layout_check
      use, intrinsic :: iso_c_binding
      use mfc_lbm_c
      implicit none
      type(mfc_params), target :: p
      integer(c_intptr_t) :: b
      b = transfer(c_loc(p), b)
      write(*,'(a,1x,i0)') 'sizeof', int(c_sizeof(p))
      call show('nx', c_loc(p%nx))
      call show('ny', c_loc(p%ny))
      call show('nz', c_loc(p%nz))
      call show('precision', c_loc(p%precision))
      call show('npart_max', c_loc(p%npart_max))
      call show('walls_y', c_loc(p%walls_y))
      call show('iseedf', c_loc(p%iseedf))
      call show('irelease', c_loc(p%irelease))
      call show('storage', c_loc(p%storage))
      call show('s1', c_loc(p%s1))
      call show('s2', c_loc(p%s2))
      call show('s4', c_loc(p%s4))
      call show('s9', c_loc(p%s9))
      call show('s10', c_loc(p%s10))
      call show('s13', c_loc(p%s13))
      call show('s16', c_loc(p%s16))
      call show('omegepsl', c_loc(p%omegepsl))
      call show('omegepslj', c_loc(p%omegepslj))
      call show('omegxx', c_loc(p%omegxx))
      call show('ww0', c_loc(p%ww0))
      call show('ww1', c_loc(p%ww1))
      call show('ww2', c_loc(p%ww2))
      call show('rho0', c_loc(p%rho0))
      call show('rad', c_loc(p%rad))
      call show('rhopart', c_loc(p%rhopart))
      call show('mingap', c_loc(p%mingap))
      call show('volp', c_loc(p%volp))
      call show('amp', c_loc(p%amp))
      call show('aip', c_loc(p%aip))
      call show('rhog', c_loc(p%rhog))
      call show('stf0', c_loc(p%stf0))
      call show('stf1', c_loc(p%stf1))
      call show('gscale', c_loc(p%gscale))
      call show('vscale', c_loc(p%vscale))
      call show('tscale', c_loc(p%tscale))
      call show('xkf', c_loc(p%xkf))
      call show('pi2', c_loc(p%pi2))
      call show('body_force', c_loc(p%body_force))
      call show('particle_force', c_loc(p%particle_force))
      call show('mingap_w', c_loc(p%mingap_w))
      call show('stf0_w', c_loc(p%stf0_w))
      call show('stf1_w', c_loc(p%stf1_w))
      contains
      subroutine show(name, a)
        character(len=*), intent(in) :: name
        type(c_ptr), intent(in) :: a
        integer(c_intptr_t) :: v
        v = transfer(a, v)
        write(*,'(a,1x,i0)') name, int(v - b)
      end subroutine show
      end program layout_check
