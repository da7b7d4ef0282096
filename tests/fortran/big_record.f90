! big_record: writes one sequential unformatted record of n REAL(8) values
! a(i) = i (n from the command line) to the file named second, as the
! reference's savecntdflow writes its population record, to check how this
! compiler's runtime frames records over 2 GB.
      program big_record
      implicit none
      integer(8) :: n, i
      real(8), allocatable :: a(:)
      character(len=32) :: arg
      character(len=512) :: path
      call get_command_argument(1, arg)
      call get_command_argument(2, path)
      read(arg, *) n
      allocate(a(n))
      do i = 1, n
        a(i) = real(i, 8)
      end do
      open(11, file=trim(path), form='unformatted', status='replace')
      write(11) a
      close(11)
      end program big_record
