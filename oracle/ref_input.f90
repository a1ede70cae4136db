! Test infrastructure (not reference code): reads an iterate input file with the
! reference's own reader (module osb, readOSBFile -> myReadOSB) and writes
! what it read, one value per line, to a dump file for comparison.
subroutine refin_dump(fin, fout)
  use osb
  implicit none
  character(len=*), intent(in) :: fin, fout
  integer :: i
  call readOSBFile(fin)
  open(98, file=fout, status='replace')
  write(98, '(I12)') sF
  do i = 1, sF
    write(98, '(I12)') sN(i)
  end do
  do i = 1, sF
    write(98, '(L2)') sDEP(i)
  end do
  write(98, '(I12)') sJOB, sOSB
  write(98, '(L2)') sCX, sNDVR, sST, sAP
  write(98, '(I12)') sBJ%mMax
  write(98, '(ES26.17E3)') sBJ%mTol
  write(98, '(I12)') sQMR%mMax
  write(98, '(ES26.17E3)') sQMR%mTol
  write(98, '(ES26.17E3)') sConv%mE0, sConv%mTol
  write(98, '(I12)') sConv%mStart, sConv%mStep, sConv%mMax, sConv%mNum, sConv%mGap
  write(98, '(ES26.17E3)') sOSBW%mE0, sOSBW%mDE, sOSBW%mBeta
  write(98, '(I12)') sOSBW%mCnt
  write(98, '(I12)') sHOSB, sVOSB, sHW, sVX, sPT
  write(98, '(A)') trim(H0File)
  if (sNDVR) then
    write(98, '(A)') trim(OUTHFile)
  else
    write(98, '(A)') trim(RESFile)
  end if
  if (any(sDEP(1:sF))) write(98, '(A)') trim(DEPFile)
  if (sHOSB /= 0) write(98, '(A)') trim(HOSBFile)
  if (sVOSB /= 0) then
    write(98, '(A)') trim(VOSBFile)
    write(98, '(A)') trim(EigFile)
  end if
  close(98)
end subroutine
