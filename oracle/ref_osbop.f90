! TEST INFRASTRUCTURE ONLY (our own glue, not reference code).  Fills the
! operator state of the reference's serial module osb (sF, sN, sDEP, sNDVR,
! H0, DEP, RES, OUTH, VOSB, EIG0, sOSBW) from caller arrays and calls the
! module's own MYHX (src/iterate/module/osb.hx.h:71-110) and MYPX
! (src/iterate/module/osb.px.h:125-219), so that the oracle's level loops and
! every preconditioner kernel are checked against the reference as compiled.
! Arrays are the module's concatenated layouts (calOSB, osb.init.h:2-49).

subroutine refop_set(f, an, aflag, ndvr, ah0, adep, ares, aouth) bind(C, name='refop_set')
  use iso_c_binding
  use osb
  implicit none
  integer(c_int), value :: f, ndvr
  integer(c_int), intent(in) :: an(f), aflag(f)
  real(c_double), intent(in) :: ah0(*), adep(*), ares(*), aouth(*)
  sF = f
  sN(1:f) = an(1:f)
  sDEP(1:f) = aflag(1:f) /= 0
  sNDVR = ndvr /= 0
  sST = .false.
  call calOSB()
  if (allocated(H0)) deallocate(H0)
  allocate(H0(max(1, myH0%mLen)))
  if (myH0%mLen > 0) H0(1:myH0%mLen) = ah0(1:myH0%mLen)
  if (allocated(DEP)) deallocate(DEP)
  allocate(DEP(max(1, myDep%mLen)))
  if (myDep%mLen > 0) DEP(1:myDep%mLen) = adep(1:myDep%mLen)
  if (allocated(RES)) deallocate(RES)
  allocate(RES(myLen))
  RES(1:myLen) = ares(1:myLen)
  if (allocated(OUTH)) deallocate(OUTH)
  allocate(OUTH(max(1, outLen)))
  if (outLen > 0) OUTH(1:outLen) = aouth(1:outLen)
end subroutine

subroutine refop_set_shift(e0, de, beta) bind(C, name='refop_set_shift')
  use iso_c_binding
  use osb
  implicit none
  real(c_double), value :: e0, de, beta
  sOSBW%mE0 = e0
  sOSBW%mDE = de
  sOSBW%mBeta = beta
end subroutine

subroutine refop_set_precond(avosb, aeig0) bind(C, name='refop_set_precond')
  use iso_c_binding
  use osb
  implicit none
  real(c_double), intent(in) :: avosb(*), aeig0(*)
  if (allocated(VOSB)) deallocate(VOSB)
  allocate(VOSB(myVOSB%mLen))
  VOSB(1:myVOSB%mLen) = avosb(1:myVOSB%mLen)
  if (allocated(EIG0)) deallocate(EIG0)
  allocate(EIG0(myLen))
  EIG0(1:myLen) = aeig0(1:myLen)
end subroutine

! nType as osb's TA0 = 0 / TA1 = 1
subroutine refop_hx(ntype, x, y) bind(C, name='refop_hx')
  use iso_c_binding
  use osb
  implicit none
  integer(c_int), value :: ntype
  real(c_double), intent(in) :: x(*)
  real(c_double), intent(out) :: y(*)
  call MYHX(ntype, myLen, x, y)
end subroutine

! nOSB as osb's TOSB = 1 / TOSBD1 = 2 / TOSBD2 = 3; nPC TA0 = 0 / TA1 = 1
subroutine refop_px(nosbk, npc, x, y) bind(C, name='refop_px')
  use iso_c_binding
  use osb
  implicit none
  integer(c_int), value :: nosbk, npc
  real(c_double), intent(in) :: x(*)
  real(c_double), intent(out) :: y(*)
  call MYPX(nosbk, npc, myLen, x, y)
end subroutine

! The window set-up of TOSBW as the module does it (initOSBW,
! src/iterate/module/osb.osbw.h:2-50, without the LAPACK factorisation):
! adjustWinSize, getOSBWSize, getOSBWIndex and getVOSBH0
! (src/iterate/module/osb.hij.h:133-160).  ind gets 1-based grid indices,
! hw the window's H (EIG0 on the diagonal); both sized by the caller for cnt.
subroutine refop_osbw(cnt, e0, de, wlen, wix, hmat) bind(C, name='refop_osbw')
  use iso_c_binding
  use osb
  implicit none
  integer(c_int), value :: cnt
  real(c_double), value :: e0
  real(c_double), intent(inout) :: de
  integer(c_int), intent(out) :: wlen
  integer(c_int), intent(out) :: wix(cnt)
  real(c_double), intent(out) :: hmat(cnt, cnt)
  integer(kind=8) :: n8
  call adjustWinSize(cnt, e0, de)
  n8 = getOSBWSize(e0, de)
  wlen = int(n8)
  if (n8 /= cnt) return
  call getOSBWIndex(e0, de, n8, wix)
  call getVOSBH0(n8, wix, hmat)
end subroutine

! One OSB state's vector (getVi, src/iterate/module/osb.vi.h:74-80) for the
! 1-based grid index k.
subroutine refop_vi(k, vi) bind(C, name='refop_vi')
  use iso_c_binding
  use osb
  implicit none
  integer(c_int), value :: k
  real(c_double), intent(out) :: vi(*)
  integer(kind=8) :: one
  integer :: kix(1), bk(sF, 1), col(sF, 1)
  one = 1
  kix(1) = k
  call getLevelIndex(one, kix, bk, col)
  call getVi(bk(1:sF, 1), col(1:sF, 1), vi)
end subroutine

! The absorbing potentials of module osb (AP, APP, APR; osb.data.h:61,
! loaded by osb.io.h:20-30): which 0 = AP, 1 = APP, 2 = APR.
subroutine refop_set_ap(which, a) bind(C, name='refop_set_ap')
  use iso_c_binding
  use osb
  implicit none
  integer(c_int), value :: which
  real(c_double), intent(in) :: a(*)
  ! (re)sized for the current grid: an earlier operator may have had another
  if (allocated(AP)) then
    if (size(AP) /= myLen) deallocate(AP, APP, APR)
  end if
  if (.not. allocated(AP)) allocate(AP(myLen), APP(myLen), APR(myLen))
  select case (which)
  case (0)
    AP(1:myLen) = a(1:myLen)
  case (1)
    APP(1:myLen) = a(1:myLen)
  case default
    APR(1:myLen) = a(1:myLen)
  end select
end subroutine

! Module osb's MYHX_DX (src/iterate/module/osb.hxdx.h:177-240) and MYPX_DX
! (osb.mypxdx.h) on complex vectors; nType / nPC as osb.data.h:18-22, nOSB
! as refop_px.
subroutine refop_hx_dx(ntype, x, y) bind(C, name='refop_hx_dx')
  use iso_c_binding
  use osb
  implicit none
  integer(c_int), value :: ntype
  complex(c_double_complex), intent(in) :: x(*)
  complex(c_double_complex), intent(out) :: y(*)
  call MYHX_DX(ntype, myLen, x, y)
end subroutine

subroutine refop_px_dx(nosbk, npc, x, y) bind(C, name='refop_px_dx')
  use iso_c_binding
  use osb
  implicit none
  integer(c_int), value :: nosbk, npc
  complex(c_double_complex), intent(in) :: x(*)
  complex(c_double_complex), intent(out) :: y(*)
  call MYPX_DX(nosbk, npc, myLen, x, y)
end subroutine

! OSB_QMRDX (src/iterate/module/osbw.qmrdx.h:16-27): QMR_CX with HX_DX =
! MYHX_DX(sHC) and PX_DX = MYPX_DX(sOSB, sPC); nOSB as refop_px.
function refop_qmr_dx(nosbk, hc, pc, ntype, qmax, tol, b, x, qres) bind(C, name='refop_qmr_dx') result(r)
  use iso_c_binding
  use osb
  implicit none
  integer(c_int), value :: nosbk, hc, pc, ntype, qmax
  real(c_double), value :: tol
  complex(c_double_complex), intent(in) :: b(*)
  complex(c_double_complex), intent(inout) :: x(*)
  real(c_double), intent(out) :: qres
  integer(c_int) :: r
  integer :: QMR_CX
  sOSB = nosbk
  sHC = hc
  sPC = pc
  r = QMR_CX(ntype, qmax, tol, myLen, b, HX_DX, PX_DX, x, qres)
end function

! sCX = T (complex level matrices): H0CX, DEPCX (dependent levels) and
! OUTHCX (sNDVR) of module osb, in the module's concatenated layouts, after
! refop_set has sized them (calOSB).
subroutine refop_set_cx(ah0, adep, aouth) bind(C, name='refop_set_cx')
  use iso_c_binding
  use osb
  implicit none
  complex(c_double_complex), intent(in) :: ah0(*), adep(*), aouth(*)
  sCX = .true.
  if (allocated(H0CX)) deallocate(H0CX)
  allocate(H0CX(max(1, myH0%mLen)))
  if (myH0%mLen > 0) H0CX(1:myH0%mLen) = ah0(1:myH0%mLen)
  if (allocated(DEPCX)) deallocate(DEPCX)
  allocate(DEPCX(max(1, myDep%mLen)))
  if (myDep%mLen > 0) DEPCX(1:myDep%mLen) = adep(1:myDep%mLen)
  if (allocated(OUTHCX)) deallocate(OUTHCX)
  allocate(OUTHCX(max(1, outLen)))
  if (outLen > 0) OUTHCX(1:outLen) = aouth(1:outLen)
end subroutine

! Module osb's MYHX_CX (src/iterate/module/osb.hxcx.h:191-245).
subroutine refop_hx_cx(ntype, x, y) bind(C, name='refop_hx_cx')
  use iso_c_binding
  use osb
  implicit none
  integer(c_int), value :: ntype
  complex(c_double_complex), intent(in) :: x(*)
  complex(c_double_complex), intent(out) :: y(*)
  call MYHX_CX(ntype, myLen, x, y)
end subroutine

! The sCX branch of myDiag (src/iterate/module/osb.diag.h:76-101) without
! its HOSB file: EIG0 = RES, then for levels sF..1 the level's HOSB from
! myDiagInitCX (osb.diag.h:36-49: OUTHCX, HOSBINIT_DEP_CX or
! HOSBINIT_XYZ_CX) and HosbDiagDX, which leaves the real rotations in VOSB
! and the real parts of the rotated diagonal in EIG0.
subroutine refop_diag_cx(bjmax, bjtol, avosb, aeig0) bind(C, name='refop_diag_cx')
  use iso_c_binding
  use osb
  implicit none
  integer(c_int), value :: bjmax
  real(c_double), value :: bjtol
  real(c_double), intent(out) :: avosb(*), aeig0(*)
  complex(c_double), allocatable :: hb(:)
  integer :: level, mx
  double precision :: tol
  mx = bjmax
  tol = bjtol
  if (allocated(VOSB)) deallocate(VOSB)
  allocate(VOSB(myVOSB%mLen))
  if (allocated(EIG0)) deallocate(EIG0)
  allocate(EIG0(myLen))
  EIG0(1:myLen) = RES(1:myLen)
  do level = sF, 1, -1
    allocate(hb(sN(level) * myLen))
    if (sNDVR .and. level == sF) then
      hb(1:sN(sF) * myLen) = OUTHCX(1:sN(sF) * myLen)
    else if (sDEP(level)) then
      call HOSBINIT_DEP_CX(myBlk(level), sN(level), myDim(level), H0CX(myH0%mStart(level)), &
                           DEPCX(myDEP%mStart(level)), EIG0, hb)
    else
      call HOSBINIT_XYZ_CX(myBlk(level), sN(level), myDim(level), H0CX(myH0%mStart(level)), EIG0, hb)
    end if
    call HosbDiagDX(mx, tol, myBlk(level), sN(level), myDim(level), hb, VOSB(myVOSB%mStart(level)), EIG0)
    deallocate(hb)
  end do
  avosb(1:myVOSB%mLen) = VOSB(1:myVOSB%mLen)
  aeig0(1:myLen) = EIG0(1:myLen)
end subroutine
