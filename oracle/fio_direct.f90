! Test infrastructure (not reference code): one direct-access record of N
! doubles with RECL = the array's IOLENGTH, written and read by this
! compiler's Fortran runtime, to check the library's direct-access format.
logical function fio_save_dir(n, a, fname)
  implicit none
  integer(kind=8), intent(in) :: n
  double precision, intent(in) :: a(n)
  character(len=*), intent(in) :: fname
  integer(kind=8) :: rl
  integer :: ios
  fio_save_dir = .false.
  inquire(iolength=rl) a
  open(77, file=fname, status='replace', form='unformatted', access='direct', recl=rl, iostat=ios)
  if (ios /= 0) return
  write(77, rec=1) a
  close(77)
  fio_save_dir = .true.
end function

logical function fio_load_dir(n, a, fname)
  implicit none
  integer(kind=8), intent(in) :: n
  double precision, intent(out) :: a(n)
  character(len=*), intent(in) :: fname
  integer(kind=8) :: rl
  integer :: ios
  fio_load_dir = .false.
  inquire(iolength=rl) a
  open(77, file=fname, status='old', form='unformatted', access='direct', recl=rl, iostat=ios)
  if (ios /= 0) return
  read(77, rec=1, iostat=ios) a
  if (ios /= 0) return
  close(77)
  fio_load_dir = .true.
end function
