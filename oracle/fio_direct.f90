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

! Per-point records (the layout P2SaveDataGrid writes,
! src/libutil/parallel/p_io/p2io_grid.f90:62-83, here from one rank):
! RECL = IOLENGTH of data(1,:,:), record j holds data(j,1:NG,1:NG).
logical function fio_save_grid(n, ng, a, fname)
  implicit none
  integer(kind=8), intent(in) :: n, ng
  double precision, intent(in) :: a(n, ng, ng)
  character(len=*), intent(in) :: fname
  integer(kind=8) :: rl, j
  integer :: ios
  fio_save_grid = .false.
  inquire(iolength=rl) a(1, :, :)
  open(78, file=fname, status='replace', form='unformatted', access='direct', recl=rl, iostat=ios)
  if (ios /= 0) return
  do j = 1, n
    write(78, rec=j) a(j, 1:ng, 1:ng)
  end do
  close(78)
  fio_save_grid = .true.
end function
