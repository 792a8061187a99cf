! The density step through the Fortran drop-in mpdata3d_rho_gauge_gpubc
! (advection/src_algorithms/mpdata3d-rho-gauge-gpubc.f90:23-34) on a
! 26x21x12 open domain: three calls with substepping at Courant about 1.5,
! carrying nsubsteps, nsubsteps_o, dtn and dtnold from call to call as the
! reference's INOUT arguments.  Usage: mpdata_rho_check <dump>.  Writes every
! array before the first call to <dump>.in and after the last call to
! <dump>.out (stream files of REAL(euwp), in the order ox, oy, oz, u, v, w,
! u1, u2, u3, x_in, x_incomp, h, rhoadv, rhr, pexe, pexr1, pexr2, bcx, bcy;
! ox..w with their three time levels) and prints the scalar state per call.
PROGRAM mpdata_rho_check
   USE iso_c_binding
   USE mpdata_hip
   IMPLICIT NONE
   INTEGER, PARAMETER :: np = 26, mp = 21, lp = 12, ih = 1, nfield = 19, ncall = 3
   REAL(euwp), PARAMETER :: dxi = 0.01_euwp, dt = 40.0_euwp
   TYPE(c_ptr) :: d(nfield)
   INTEGER(c_long_long) :: n(nfield), ns, nu1, nu2, nu3
   REAL(euwp), ALLOCATABLE :: buf(:)
   INTEGER :: f, c, nsub, nsubo, liner, q, unit_in, unit_out
   REAL(euwp) :: dtn, dtnold
   CHARACTER(len=512) :: dump

   IF (COMMAND_ARGUMENT_COUNT() < 1) THEN
      WRITE (*, '(A)') 'usage: mpdata_rho_check <dump>'
      ERROR STOP 2
   END IF
   CALL GET_COMMAND_ARGUMENT(1, dump)
   CALL mpdata_init(np, mp, lp, ih, 0, 0, 0)
   ns = INT(np + 2 * ih, c_long_long) * (mp + 2 * ih) * (lp + 2 * ih)
   nu1 = INT(np + 1 + 2 * ih, c_long_long) * (mp + 2 * ih) * (lp + 2 * ih)
   nu2 = INT(np + 2 * ih, c_long_long) * (mp + 1 + 2 * ih) * (lp + 2 * ih)
   nu3 = INT(np + 2 * ih, c_long_long) * (mp + 2 * ih) * (lp + 1 + 2 * ih)
   n(1:6) = 3 * ns
   n(7) = nu1
   n(8) = nu2
   n(9) = nu3
   n(10:17) = ns
   n(18) = INT(mp, c_long_long) * lp * 2
   n(19) = INT(np, c_long_long) * lp * 2
   ALLOCATE (buf(MAXVAL(n)))

   ! inputs: smooth fields of the flat index, velocities in m/s (Courant
   ! |u| dt dxi up to about 0.5 per direction), level n-1 momentum = level n
   ! times h, density positive
   DO f = 1, nfield
      DO q = 1, INT(n(f))
         buf(q) = field_value(f, q)
      END DO
      d(f) = mpdata_alloc(n(f))
      CALL mpdata_put(d(f), buf, n(f))
   END DO
   ! x_in = x_incomp / h and the momentum level ox(.,2) = ox(.,0) * h, on the host
   BLOCK
      REAL(euwp), ALLOCATABLE :: xinc(:), hh(:), lev(:)
      ALLOCATE (xinc(ns), hh(ns), lev(3 * ns))
      CALL mpdata_get(xinc, d(11), ns)
      CALL mpdata_get(hh, d(12), ns)
      buf(1:ns) = xinc / hh
      CALL mpdata_put(d(10), buf, ns)
      DO f = 1, 3
         CALL mpdata_get(lev, d(f), 3 * ns)
         lev(2 * ns + 1:3 * ns) = lev(1:ns) * hh
         CALL mpdata_put(d(f), lev, 3 * ns)
      END DO
   END BLOCK

   OPEN (newunit=unit_in, file=TRIM(dump)//'.in', access='stream', form='unformatted', &
         status='replace')
   DO f = 1, nfield
      CALL mpdata_get(buf, d(f), n(f))
      WRITE (unit_in) buf(1:n(f))
   END DO
   CLOSE (unit_in)

   nsub = 1
   nsubo = 1
   dtn = dt
   dtnold = dt
   liner = 0
   DO c = 1, ncall
      CALL mpdata3d_rho_gauge_gpubc(.FALSE., nsub, nsubo, .TRUE., dtn, dtnold, .FALSE., &
                                    .FALSE., .FALSE., 0.9_euwp, 2 * dt, &
                                    d(1), d(2), d(3), d(4), d(5), d(6), d(7), d(8), d(9), &
                                    d(10), d(11), d(12), d(13), d(14), d(15), d(16), d(17), &
                                    d(18), d(19), 0, 0, 0, 0, np, mp, lp, ih, &
                                    dxi, dxi, dxi, dt, liner)
      WRITE (*, '(A,I0,A,I0,1X,I0,1X,ES25.17E3,1X,ES25.17E3)') 'call ', c, ' state ', nsub, &
         nsubo, dtn, dtnold
   END DO
   CALL mpdata_sync()

   OPEN (newunit=unit_out, file=TRIM(dump)//'.out', access='stream', form='unformatted', &
         status='replace')
   DO f = 1, nfield
      CALL mpdata_get(buf, d(f), n(f))
      WRITE (unit_out) buf(1:n(f))
   END DO
   CLOSE (unit_out)
   DO f = 1, nfield
      CALL mpdata_free(d(f))
   END DO
   CALL mpdata_finalize()
   WRITE (*, '(A)') 'rho check: done'

CONTAINS

   ! value q of field f: level-n velocities of amplitude 1.2 m/s and more
   ! random-looking values elsewhere
   REAL(euwp) FUNCTION field_value(f, q) RESULT(v)
      INTEGER, INTENT(IN) :: f, q
      REAL(euwp) :: t
      t = REAL(q, euwp)
      SELECT CASE (f)
      CASE (1:3)        ! ox, oy, oz
         v = 1.2_euwp * SIN(0.013_euwp * t * f + 0.7_euwp * f)
      CASE (4:6)        ! u, v, w
         v = 3.0_euwp * COS(0.007_euwp * t + f)
      CASE (7:9)        ! u1..u3 (overwritten on the faces, kept in the halo)
         v = 0.05_euwp * SIN(0.011_euwp * t + f)
      CASE (11)         ! x_incomp
         v = 1.0_euwp + 0.5_euwp * SIN(0.0037_euwp * t) * COS(0.019_euwp * t)
      CASE (12)         ! h
         v = 1.0_euwp + 0.1_euwp * COS(0.029_euwp * t)
      CASE (15)         ! pexe
         v = 1.0_euwp + 0.2_euwp * SIN(0.017_euwp * t)
      CASE (18, 19)     ! inflow values
         v = 0.9_euwp + 0.3_euwp * SIN(0.05_euwp * t + f)
      CASE DEFAULT      ! x_in (set from x_incomp), outputs
         v = 0.5_euwp + 0.25_euwp * SIN(0.023_euwp * t)
      END SELECT
   END FUNCTION field_value

END PROGRAM mpdata_rho_check
