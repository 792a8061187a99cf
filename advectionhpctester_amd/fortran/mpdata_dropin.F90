! Literal drop-ins for the reference's stage modules: modules with the
! reference's names (module_upwind3d_gpubc, module_antidiff3d_gauge_gpubc,
! module_antidiff3d_standard_gpubc) whose subroutines have the reference's
! abstract interfaces exactly -- REAL(euwp) explicit-shape array dummies with
! lower bound 1-ih, the same intents and optional arguments
! (advection/src_algorithms/upwind3d-gpubc.f90:20-78,
! antidiff3d-gauge-gpubc.f90:14-49, antidiff3d-standard-gpubc.f90:14-46) --
! so that the reference's advec_dwarf binds its procedure pointers to them
! unchanged (advec.f90:143-144, 196-203) when these modules replace its
! stage sources at link time.
!
! Each call runs the HIP library's stage through module mpdata_hip: arrays
! in device or managed memory are passed as they are; host arrays (the
! reference's CPU build) are copied to device buffers kept for the domain,
! and the arrays the stage writes are copied back.  The library context
! (the stage's scratch) is created for the domain of the first call and
! re-created when the domain changes.
MODULE mpdata_dropin_support
   USE iso_c_binding
   USE mpdata_hip, ONLY: euwp, mpdata_alloc, mpdata_free, mpdata_put, mpdata_get, &
                         mpdata_ensure_domain
   IMPLICIT NONE
   PRIVATE
   PUBLIC :: euwp, ensure_domain, dev_of, copy_back, nscal, nfx, nfy, nfz

   INTERFACE
      FUNCTION c_addr(a) BIND(C, name="mpdata_addr")
         IMPORT :: c_ptr, euwp
         REAL(euwp), INTENT(IN) :: a(*)
         TYPE(c_ptr) :: c_addr
      END FUNCTION
      FUNCTION c_is_device(p) BIND(C, name="mpdata_is_device_ptr")
         IMPORT :: c_ptr, c_int
         TYPE(c_ptr), VALUE :: p
         INTEGER(c_int) :: c_is_device
      END FUNCTION
   END INTERFACE

   INTEGER, PARAMETER :: nslot = 12
   TYPE(c_ptr), SAVE :: buf(nslot) = c_null_ptr
   INTEGER(c_long_long), SAVE :: bufn(nslot) = 0

CONTAINS

   SUBROUTINE ensure_domain(np, mp, lp, ih, ibcx, ibcy, ibcz)
      INTEGER, INTENT(IN) :: np, mp, lp, ih, ibcx, ibcy, ibcz
      CALL mpdata_ensure_domain(np, mp, lp, ih, ibcx, ibcy, ibcz)
   END SUBROUTINE ensure_domain

   INTEGER(c_long_long) FUNCTION nscal(np, mp, lp, ih)
      INTEGER, INTENT(IN) :: np, mp, lp, ih
      nscal = INT(np + 2*ih, c_long_long)*(mp + 2*ih)*(lp + 2*ih)
   END FUNCTION nscal
   INTEGER(c_long_long) FUNCTION nfx(np, mp, lp, ih)
      INTEGER, INTENT(IN) :: np, mp, lp, ih
      nfx = INT(np + 1 + 2*ih, c_long_long)*(mp + 2*ih)*(lp + 2*ih)
   END FUNCTION nfx
   INTEGER(c_long_long) FUNCTION nfy(np, mp, lp, ih)
      INTEGER, INTENT(IN) :: np, mp, lp, ih
      nfy = INT(np + 2*ih, c_long_long)*(mp + 1 + 2*ih)*(lp + 2*ih)
   END FUNCTION nfy
   INTEGER(c_long_long) FUNCTION nfz(np, mp, lp, ih)
      INTEGER, INTENT(IN) :: np, mp, lp, ih
      nfz = INT(np + 2*ih, c_long_long)*(mp + 2*ih)*(lp + 1 + 2*ih)
   END FUNCTION nfz

   ! The device address of array a (n elements): a itself in device or
   ! managed memory, else staging buffer `slot` filled with a.
   FUNCTION dev_of(a, n, slot) RESULT(p)
      REAL(euwp), INTENT(IN) :: a(*)
      INTEGER(c_long_long), INTENT(IN) :: n
      INTEGER, INTENT(IN) :: slot
      TYPE(c_ptr) :: p
      p = c_addr(a)
      IF (c_is_device(p) /= 0) RETURN
      IF (bufn(slot) /= n) THEN
         IF (c_associated(buf(slot))) CALL mpdata_free(buf(slot))
         buf(slot) = mpdata_alloc(n)
         bufn(slot) = n
      END IF
      CALL mpdata_put(buf(slot), a, n)
      p = buf(slot)
   END FUNCTION dev_of

   ! After the stage: a's values back from the buffer it was staged in (no-op
   ! when a was passed as it is).
   SUBROUTINE copy_back(a, n, p)
      REAL(euwp), INTENT(INOUT) :: a(*)
      INTEGER(c_long_long), INTENT(IN) :: n
      TYPE(c_ptr), INTENT(IN) :: p
      IF (c_associated(p, c_addr(a))) RETURN
      CALL mpdata_get(a, p, n)
   END SUBROUTINE copy_back
END MODULE mpdata_dropin_support

MODULE module_upwind3d_gpubc
   USE iso_c_binding
   USE mpdata_dropin_support
   USE mpdata_hip, ONLY: upwind_dev => upwind3d_gpubc
   IMPLICIT NONE
CONTAINS
   SUBROUTINE upwind3d_gpubc(u1, u2, u3, &
                             x_in, xant, &
                             xforc_impl, xforc_expl, &
                             rhr, h, &
                             bcx, bcy, &
                             ipoles0, &
                             ibcx0, ibcy0, ibcz0, &
                             np, mp, lp, ih, dt, &
                             do_serialization_in, do_serialization_out)
      LOGICAL, OPTIONAL :: do_serialization_in
      LOGICAL, OPTIONAL :: do_serialization_out
      INTEGER, INTENT(IN) :: ipoles0, ibcx0, ibcy0, ibcz0, np, mp, lp, ih
      REAL(KIND=euwp), INTENT(IN) :: dt
      REAL(KIND=euwp), INTENT(INOUT) :: &
         x_in(1 - ih:np + ih, 1 - ih:mp + ih, 1 - ih:lp + ih)
      REAL(KIND=euwp), INTENT(IN) :: &
         u1(1 - ih:np + 1 + ih, 1 - ih:mp + ih, 1 - ih:lp + ih), &
         u2(1 - ih:np + ih, 1 - ih:mp + 1 + ih, 1 - ih:lp + ih), &
         u3(1 - ih:np + ih, 1 - ih:mp + ih, 1 - ih:lp + 1 + ih)
      REAL(KIND=euwp), INTENT(OUT) :: &
         xant(1 - ih:np + ih, 1 - ih:mp + ih, 1 - ih:lp + ih)
      REAL(KIND=euwp), INTENT(IN) :: &
         h(1 - ih:np + ih, 1 - ih:mp + ih, 1 - ih:lp + ih), &
         rhr(1 - ih:np + ih, 1 - ih:mp + ih, 1 - ih:lp + ih), &
         xforc_expl(1 - ih:np + ih, 1 - ih:mp + ih, 1 - ih:lp + ih)
      REAL(KIND=euwp), INTENT(IN), OPTIONAL :: &
         xforc_impl(1 - ih:np + ih, 1 - ih:mp + ih, 1 - ih:lp + ih)
      REAL(KIND=euwp), INTENT(IN) :: &
         bcx(mp, lp, 2), &
         bcy(np, lp, 2)
      TYPE(c_ptr) :: d(11)
      INTEGER(c_long_long) :: ns
      CALL ensure_domain(np, mp, lp, ih, ibcx0, ibcy0, ibcz0)
      ns = nscal(np, mp, lp, ih)
      d(1) = dev_of(u1, nfx(np, mp, lp, ih), 1)
      d(2) = dev_of(u2, nfy(np, mp, lp, ih), 2)
      d(3) = dev_of(u3, nfz(np, mp, lp, ih), 3)
      d(4) = dev_of(x_in, ns, 4)
      d(5) = dev_of(xant, ns, 5)
      d(6) = c_null_ptr
      IF (PRESENT(xforc_impl)) d(6) = dev_of(xforc_impl, ns, 6)
      d(7) = dev_of(xforc_expl, ns, 7)
      d(8) = dev_of(rhr, ns, 8)
      d(9) = dev_of(h, ns, 9)
      d(10) = dev_of(bcx, INT(mp, c_long_long)*lp*2, 10)
      d(11) = dev_of(bcy, INT(np, c_long_long)*lp*2, 11)
      CALL upwind_dev(d(1), d(2), d(3), d(4), d(5), d(6), d(7), d(8), d(9), d(10), d(11), &
                      ipoles0, ibcx0, ibcy0, ibcz0, np, mp, lp, ih, dt, &
                      do_serialization_in, do_serialization_out)
      CALL copy_back(x_in, ns, d(4))
      CALL copy_back(xant, ns, d(5))
   END SUBROUTINE upwind3d_gpubc
END MODULE module_upwind3d_gpubc

MODULE module_antidiff3d_gauge_gpubc
   USE iso_c_binding
   USE mpdata_dropin_support
   USE mpdata_hip, ONLY: gauge_dev => antidiff3d_gauge_gpubc
   IMPLICIT NONE
CONTAINS
   SUBROUTINE antidiff3d_gauge_gpubc(lupdatemulti, u1, u2, u3, xant, x_in, rhr, h, &
                                     iflip, ipoles0, ibcx0, ibcy0, ibcz0, np, mp, lp, ih, &
                                     do_serialization_out)
      LOGICAL, INTENT(IN) :: lupdatemulti
      LOGICAL, OPTIONAL :: do_serialization_out
      INTEGER, INTENT(IN) :: iflip, ipoles0, ibcx0, ibcy0, ibcz0, np, mp, lp, ih
      REAL(KIND=euwp), INTENT(IN) :: &
         u1(1 - ih:np + 1 + ih, 1 - ih:mp + ih, 1 - ih:lp + ih), &
         u2(1 - ih:np + ih, 1 - ih:mp + 1 + ih, 1 - ih:lp + ih), &
         u3(1 - ih:np + ih, 1 - ih:mp + ih, 1 - ih:lp + 1 + ih), &
         h(1 - ih:np + ih, 1 - ih:mp + ih, 1 - ih:lp + ih), &
         rhr(1 - ih:np + ih, 1 - ih:mp + ih, 1 - ih:lp + ih)
      REAL(KIND=euwp), INTENT(INOUT) :: &
         x_in(1 - ih:np + ih, 1 - ih:mp + ih, 1 - ih:lp + ih), &
         xant(1 - ih:np + ih, 1 - ih:mp + ih, 1 - ih:lp + ih)
      TYPE(c_ptr) :: d(7)
      INTEGER(c_long_long) :: ns
      CALL ensure_domain(np, mp, lp, ih, ibcx0, ibcy0, ibcz0)
      ns = nscal(np, mp, lp, ih)
      d(1) = dev_of(u1, nfx(np, mp, lp, ih), 1)
      d(2) = dev_of(u2, nfy(np, mp, lp, ih), 2)
      d(3) = dev_of(u3, nfz(np, mp, lp, ih), 3)
      d(4) = dev_of(xant, ns, 5)
      d(5) = dev_of(x_in, ns, 4)
      d(6) = dev_of(rhr, ns, 8)
      d(7) = dev_of(h, ns, 9)
      CALL gauge_dev(lupdatemulti, d(1), d(2), d(3), d(4), d(5), d(6), d(7), iflip, ipoles0, &
                     ibcx0, ibcy0, ibcz0, np, mp, lp, ih, do_serialization_out)
      CALL copy_back(xant, ns, d(4))
      CALL copy_back(x_in, ns, d(5))
   END SUBROUTINE antidiff3d_gauge_gpubc
END MODULE module_antidiff3d_gauge_gpubc

MODULE module_antidiff3d_standard_gpubc
   USE iso_c_binding
   USE mpdata_dropin_support
   USE mpdata_hip, ONLY: standard_dev => antidiff3d_standard_gpubc
   IMPLICIT NONE
CONTAINS
   SUBROUTINE antidiff3d_standard_gpubc(lupdatemulti, u1, u2, u3, xant, x_in, rhr, h, &
                                        iflip, ipoles0, ibcx0, ibcy0, ibcz0, np, mp, lp, ih, &
                                        do_serialization_out)
      LOGICAL, INTENT(IN) :: lupdatemulti
      LOGICAL, OPTIONAL :: do_serialization_out
      INTEGER, INTENT(IN) :: iflip, ipoles0, ibcx0, ibcy0, ibcz0, np, mp, lp, ih
      REAL(KIND=euwp), INTENT(IN) :: &
         u1(1 - ih:np + 1 + ih, 1 - ih:mp + ih, 1 - ih:lp + ih), &
         u2(1 - ih:np + ih, 1 - ih:mp + 1 + ih, 1 - ih:lp + ih), &
         u3(1 - ih:np + ih, 1 - ih:mp + ih, 1 - ih:lp + 1 + ih), &
         h(1 - ih:np + ih, 1 - ih:mp + ih, 1 - ih:lp + ih), &
         rhr(1 - ih:np + ih, 1 - ih:mp + ih, 1 - ih:lp + ih)
      REAL(KIND=euwp), INTENT(INOUT) :: &
         x_in(1 - ih:np + ih, 1 - ih:mp + ih, 1 - ih:lp + ih)
      REAL(KIND=euwp), INTENT(INOUT) :: &
         xant(1 - ih:np + ih, 1 - ih:mp + ih, 1 - ih:lp + ih)
      TYPE(c_ptr) :: d(7)
      INTEGER(c_long_long) :: ns
      CALL ensure_domain(np, mp, lp, ih, ibcx0, ibcy0, ibcz0)
      ns = nscal(np, mp, lp, ih)
      d(1) = dev_of(u1, nfx(np, mp, lp, ih), 1)
      d(2) = dev_of(u2, nfy(np, mp, lp, ih), 2)
      d(3) = dev_of(u3, nfz(np, mp, lp, ih), 3)
      d(4) = dev_of(xant, ns, 5)
      d(5) = dev_of(x_in, ns, 4)
      d(6) = dev_of(rhr, ns, 8)
      d(7) = dev_of(h, ns, 9)
      CALL standard_dev(lupdatemulti, d(1), d(2), d(3), d(4), d(5), d(6), d(7), iflip, &
                        ipoles0, ibcx0, ibcy0, ibcz0, np, mp, lp, ih, do_serialization_out)
      CALL copy_back(xant, ns, d(4))
      CALL copy_back(x_in, ns, d(5))
   END SUBROUTINE antidiff3d_standard_gpubc
END MODULE module_antidiff3d_standard_gpubc
