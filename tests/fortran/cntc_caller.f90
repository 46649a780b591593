! A Fortran caller of libcntc_amd.so through ISO_C_BINDING interface blocks: the binding a
! maintainer of the reference's Fortran host would write for the module-3 routines
! (src/contact_addon.f90 declares the same names with bind(c); every argument by reference).
! It sets up a Hertzian steady-rolling case, reads settings back through the getters and calls
! cntc_calculate: on a machine without a GPU the library reports CNTC_err_other (-99), on one
! with a GPU it solves and the forces are printed.
module cntc_ifc
   use iso_c_binding, only: c_int, c_double
   implicit none
   interface
      subroutine cntc_setflags(ire, icp, lenflg, params, values) bind(c, name='cntc_setflags')
         import :: c_int
         integer(c_int), intent(in) :: ire, icp, lenflg, params(*), values(*)
      end subroutine
      subroutine cntc_getflags(ire, icp, lenflg, params, values) bind(c, name='cntc_getflags')
         import :: c_int
         integer(c_int), intent(in)  :: ire, icp, lenflg, params(*)
         integer(c_int), intent(out) :: values(*)
      end subroutine
      subroutine cntc_setmaterialparameters(ire, icp, m_digit, nparam, rparam) &
            bind(c, name='cntc_setmaterialparameters')
         import :: c_int, c_double
         integer(c_int), intent(in) :: ire, icp, m_digit, nparam
         real(c_double), intent(in) :: rparam(*)
      end subroutine
      subroutine cntc_getparameters(ire, icp, itask, lenarr, values) bind(c, name='cntc_getparameters')
         import :: c_int, c_double
         integer(c_int), intent(in)  :: ire, icp, itask, lenarr
         real(c_double), intent(out) :: values(*)
      end subroutine
      subroutine cntc_setfrictionmethod(ire, icp, imeth, nparam, params) bind(c, name='cntc_setfrictionmethod')
         import :: c_int, c_double
         integer(c_int), intent(in) :: ire, icp, imeth, nparam
         real(c_double), intent(in) :: params(*)
      end subroutine
      subroutine cntc_setsolverflags(ire, icp, gdigit, nints, iparam, nreals, rparam) &
            bind(c, name='cntc_setsolverflags')
         import :: c_int, c_double
         integer(c_int), intent(in) :: ire, icp, gdigit, nints, iparam(*), nreals
         real(c_double), intent(in) :: rparam(*)
      end subroutine
      subroutine cntc_setreferencevelocity(ire, icp, veloc) bind(c, name='cntc_setreferencevelocity')
         import :: c_int, c_double
         integer(c_int), intent(in) :: ire, icp
         real(c_double), intent(in) :: veloc
      end subroutine
      subroutine cntc_getreferencevelocity(ire, icp, veloc) bind(c, name='cntc_getreferencevelocity')
         import :: c_int, c_double
         integer(c_int), intent(in)  :: ire, icp
         real(c_double), intent(out) :: veloc
      end subroutine
      subroutine cntc_sethertzcontact(ire, icp, ipotcn, nparam, params) bind(c, name='cntc_sethertzcontact')
         import :: c_int, c_double
         integer(c_int), intent(in) :: ire, icp, ipotcn, nparam
         real(c_double), intent(in) :: params(*)
      end subroutine
      subroutine cntc_gethertzcontact(ire, icp, lenarr, values) bind(c, name='cntc_gethertzcontact')
         import :: c_int, c_double
         integer(c_int), intent(in)  :: ire, icp, lenarr
         real(c_double), intent(out) :: values(*)
      end subroutine
      subroutine cntc_setnormalforce(ire, icp, fn) bind(c, name='cntc_setnormalforce')
         import :: c_int, c_double
         integer(c_int), intent(in) :: ire, icp
         real(c_double), intent(in) :: fn
      end subroutine
      subroutine cntc_setcreepages(ire, icp, vx, vy, phi) bind(c, name='cntc_setcreepages')
         import :: c_int, c_double
         integer(c_int), intent(in) :: ire, icp
         real(c_double), intent(in) :: vx, vy, phi
      end subroutine
      subroutine cntc_setmetadata(ire, icp, lenmta, params, values) bind(c, name='cntc_setmetadata')
         import :: c_int, c_double
         integer(c_int), intent(in) :: ire, icp, lenmta, params(*)
         real(c_double), intent(in) :: values(*)
      end subroutine
      subroutine cntc_calculate(ire, icp, ierror) bind(c, name='cntc_calculate')
         import :: c_int
         integer(c_int), intent(in)  :: ire, icp
         integer(c_int), intent(out) :: ierror
      end subroutine
      subroutine cntc_getcontactforces(ire, icp, fn, tx, ty, mz) bind(c, name='cntc_getcontactforces')
         import :: c_int, c_double
         integer(c_int), intent(in)  :: ire, icp
         real(c_double), intent(out) :: fn, tx, ty, mz
      end subroutine
      subroutine cntc_finalize(ire) bind(c, name='cntc_finalize')
         import :: c_int
         integer(c_int), intent(in) :: ire
      end subroutine
   end interface
end module cntc_ifc

program cntc_caller
   use cntc_ifc
   implicit none
   integer(c_int), parameter :: ire = 1, icp = 1
   integer(c_int) :: ierror, flgs(3), vals(3), iparam(5), meta(2)
   real(c_double) :: rparam(8), hz(10), v, prm(7), fn, fx, fy, mz, mvals(2)

   flgs = (/ 1972, 1973, 1983 /)          ! CNTC_ic_tang, CNTC_ic_norm, CNTC_ic_iestim
   vals = (/ 3, 1, 0 /)
   call cntc_setflags(ire, icp, 3, flgs, vals)
   call cntc_setmaterialparameters(ire, icp, 0, 4, (/ 0.28d0, 0.28d0, 82000d0, 82000d0 /))
   call cntc_setfrictionmethod(ire, icp, 0, 2, (/ 0.3d0, 0.3d0 /))
   call cntc_setreferencevelocity(ire, icp, 10000d0)
   iparam = (/ 999, 20, 30, 1, 1 /)
   rparam = (/ 1d-8, 1d0, 0.05d0, 2d0, -1d0, 1d0, 2.6d0, 1d0 /)
   call cntc_setsolverflags(ire, icp, 5, 5, iparam, 8, rparam)
   call cntc_sethertzcontact(ire, icp, -3, 5, (/ 30d0, 30d0, 7.19d0, 5.41d0, 1.3d0 /))
   call cntc_setnormalforce(ire, icp, 1d5)
   call cntc_setcreepages(ire, icp, -0.001d0, 0d0, 0d0)
   meta = (/ 2011, 2021 /)               ! CNTC_mt_tim, CNTC_mt_run
   mvals = (/ 0.5d0, 7d0 /)
   call cntc_setmetadata(ire, icp, 2, meta, mvals)

   vals = 0
   call cntc_getflags(ire, icp, 3, flgs, vals)
   write(*,'(a,3i3)') 'flags', vals
   call cntc_getreferencevelocity(ire, icp, v)
   write(*,'(a,f10.1)') 'veloc', v
   call cntc_getparameters(ire, icp, 2, 7, prm)
   write(*,'(a,3f10.1)') 'moduli', prm(1:3)
   call cntc_gethertzcontact(ire, icp, 10, hz)
   write(*,'(a,f6.2)') 'scale', hz(7)

   call cntc_calculate(ire, icp, ierror)
   write(*,'(a,i5)') 'ierror', ierror
   if (ierror.eq.0) then
      call cntc_getcontactforces(ire, icp, fn, fx, fy, mz)
      write(*,'(a,3es16.8)') 'forces', fn, fx, fy
   endif
   call cntc_finalize(ire)
   write(*,'(a)') 'FORTRAN_CALLER_DONE'
end program cntc_caller
