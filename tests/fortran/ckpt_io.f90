! ckpt_io: writes and reads files with the record structure of the
! reference's restart files (flow, particle and forcing-process files of one
! rank) through the Fortran runtime, to check the Python reader and writer
! against a Fortran compiler's sequential unformatted format.
!   ckpt_io write <dir>   writes the three files with values set below
!   ckpt_io read  <dir>   reads the step-42 files and writes them back as step 43
      program ckpt_io
      implicit none
      integer, parameter :: lx = 5, ly = 4, lz = 3, npart = 2, msize = 10
      real(8) :: f(0:18, lx, ly, lz)
      integer :: ibnodes0(lx, ly, lz), isnodes0(lx, ly, lz), ipglb(msize)
      real(8), dimension(3, msize) :: yp, thetap, forcepp, torqpp, dwdt, domgdt
      real(8), dimension(3, npart) :: ypglb, wp, omgp
      real(8), dimension(6, 5, 5) :: b1r, b2r, b3r
      integer :: ivf(32), iseedf, iyf, nsteps, istat, imovie, nps
      integer :: i, j, k, ip
      character(len=16) :: mode
      character(len=512) :: dir

      call get_command_argument(1, mode)
      call get_command_argument(2, dir)
      if (trim(mode) == 'write') then
        do k = 1, lz
          do j = 1, ly
            do i = 1, lx
              do ip = 0, 18
                f(ip, i, j, k) = ip + 100.0d0 * i + 1.0d4 * j + 1.0d6 * k + 0.125d0
              end do
              ibnodes0(i, j, k) = merge(1, -1, mod(i + j + k, 3) == 0)
              isnodes0(i, j, k) = merge(2, -1, mod(i + j + k, 3) == 0)
            end do
          end do
        end do
        ipglb = 0
        ipglb(1) = 1
        ipglb(2) = 2
        do i = 1, msize
          do k = 1, 3
            yp(k, i) = 10.0d0 * i + k
            thetap(k, i) = -1.0d0 * i - 0.5d0 * k
            forcepp(k, i) = 1.0d-3 * i * k
            torqpp(k, i) = -2.0d-3 * i * k
            dwdt(k, i) = 3.0d-4 * i + k
            domgdt(k, i) = 4.0d-4 * i - k
          end do
        end do
        do i = 1, npart
          do k = 1, 3
            ypglb(k, i) = yp(k, i)
            wp(k, i) = 0.01d0 * i * k
            omgp(k, i) = -0.02d0 * i * k
          end do
        end do
        do k = 1, 5
          do j = 1, 5
            do i = 1, 6
              b1r(i, j, k) = i + 10 * j + 100 * k
              b2r(i, j, k) = -b1r(i, j, k)
              b3r(i, j, k) = 0.5d0 * b1r(i, j, k)
            end do
          end do
        end do
        do i = 1, 32
          ivf(i) = 1000 + i
        end do
        open(12, file=trim(dir)//'/endrunflow2D8x8S1.000042.000', form='unformatted', status='replace')
        write(12) 42, 7, 3
        write(12) f
        close(12)
        open(14, file=trim(dir)//'/endrunpart2D.000042.000', form='unformatted', status='replace')
        write(14) npart, ipglb
        write(14) ibnodes0, isnodes0
        write(14) yp, thetap
        write(14) forcepp, torqpp
        write(14) dwdt, domgdt
        write(14) ypglb, wp, omgp
        close(14)
        open(10, file=trim(dir)//'/force.000042', form='unformatted', status='replace')
        write(10) -12345
        write(10) ivf
        write(10) 77
        write(10) b1r, b2r, b3r
        close(10)
      else
        open(12, file=trim(dir)//'/endrunflow2D8x8S1.000042.000', form='unformatted', status='old')
        read(12) nsteps, istat, imovie
        read(12) f
        close(12)
        open(14, file=trim(dir)//'/endrunpart2D.000042.000', form='unformatted', status='old')
        read(14) nps, ipglb
        read(14) ibnodes0, isnodes0
        read(14) yp, thetap
        read(14) forcepp, torqpp
        read(14) dwdt, domgdt
        read(14) ypglb, wp, omgp
        close(14)
        open(10, file=trim(dir)//'/force.000042', form='unformatted', status='old')
        read(10) iseedf
        read(10) ivf
        read(10) iyf
        read(10) b1r, b2r, b3r
        close(10)
        ! write back what was read, under step 43
        open(12, file=trim(dir)//'/endrunflow2D8x8S1.000043.000', form='unformatted', status='replace')
        write(12) nsteps, istat, imovie
        write(12) f
        close(12)
        open(14, file=trim(dir)//'/endrunpart2D.000043.000', form='unformatted', status='replace')
        write(14) nps, ipglb
        write(14) ibnodes0, isnodes0
        write(14) yp, thetap
        write(14) forcepp, torqpp
        write(14) dwdt, domgdt
        write(14) ypglb, wp, omgp
        close(14)
        open(10, file=trim(dir)//'/force.000043', form='unformatted', status='replace')
        write(10) iseedf
        write(10) ivf
        write(10) iyf
        write(10) b1r, b2r, b3r
        close(10)
      end if
      end program ckpt_io
