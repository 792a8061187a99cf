! The rotating-sphere advection dwarf (DWARF3) as a Fortran host over the
! HIP library: the set-up and time loop of the reference's test program
! (advection/src_testapps/advection_dwarf_test_modular.f90:89-232, initial
! state set_initial15 and init_data of
! advection/src_testapps/advection_test_setup.f90:174-255, 619-637) with each
! step the drop-in advec_dwarf of module mpdata_hip, timed by the reference's
! timer ids and printed as its ttprt table.
!
! Usage: mpdata_dwarf [nt [dumpfile [t_adv [imin:imax]]]]
!   nt        number of time steps (default 556, the reference's)
!   dumpfile  raw file receiving the interior of x (n**3 reals, Fortran
!             order) after steps 1, 50, 100 (when <= nt) and nt; '-' = none
!             (the first size of the sweep only)
!   t_adv     1 gauge (default), 2 standard
!   imin:imax the reference's size sweep (its testloop, :89-96): for iscale
!             = imin..imax the domain is n = 59*iscale cubed and the time
!             step dt/iscale (same Courant number); default 1:1
! Prints, per size, the Energy and Error L2 norms against the initial field
! (advection_test_setup.f90:472-492, 595-617), the step throughput and the
! timer table (ids 300 = each advec_dwarf call, 55/57 = its two-step driver).
PROGRAM mpdata_dwarf
   USE iso_c_binding
   USE mpdata_hip
   IMPLICIT NONE
   INTEGER :: nt, t_adv, nargs, imin, imax, iscale, colon
   CHARACTER(LEN=512) :: arg, dumpfile

   nt = 556
   t_adv = 1
   dumpfile = '-'
   imin = 1
   imax = 1
   nargs = command_argument_count()
   IF (nargs >= 1) THEN
      CALL get_command_argument(1, arg)
      READ (arg, *) nt
   END IF
   IF (nargs >= 2) CALL get_command_argument(2, dumpfile)
   IF (nargs >= 3) THEN
      CALL get_command_argument(3, arg)
      READ (arg, *) t_adv
   END IF
   IF (nargs >= 4) THEN
      CALL get_command_argument(4, arg)
      colon = INDEX(arg, ':')
      IF (colon > 0) THEN
         READ (arg(1:colon - 1), *) imin
         READ (arg(colon + 1:), *) imax
      ELSE
         READ (arg, *) imin
         imax = imin
      END IF
   END IF
   DO iscale = imin, imax
      IF (iscale == imin) THEN
         CALL run_case(iscale, nt, t_adv, TRIM(dumpfile))
      ELSE
         CALL run_case(iscale, nt, t_adv, '-')
      END IF
   END DO

CONTAINS

   SUBROUTINE run_case(iscale, nt, t_adv, dumpfile)
      INTEGER, INTENT(IN) :: iscale, nt, t_adv
      CHARACTER(*), INTENT(IN) :: dumpfile
      INTEGER, PARAMETER :: ih = 1
      REAL(euwp), PARAMETER :: dx00 = 100._euwp
      REAL(euwp), ALLOCATABLE :: x(:, :, :), x0(:, :, :), ua(:, :, :), va(:, :, :), wa(:, :, :)
      REAL(euwp), ALLOCATABLE :: ones(:, :, :)
      TYPE(c_ptr) :: x_d, x0_d, ua_d, va_d, wa_d, u1_d, u2_d, u3_d, h_d, rhr_d, xlr_d, xbt_d
      INTEGER(c_long_long) :: nsc, nfx, nfy, nfz, nbc
      REAL(euwp) :: dx, dy, dz, dt, sqrt3i, x0c, y0c, z0c, xc, yc, zc, r_squared, h_sphere
      REAL(euwp) :: omega, omegax, omegay, omegaz, xp, yp, zp, tmp
      REAL(c_double) :: s(4), nmli, timetot, energy, error
      INTEGER :: n, i, j, k, it, unit_dump
      LOGICAL :: dumping
      INTEGER(8) :: c0, c1, rate

      n = 59*iscale
      ! grid and time step: eulag_lib_parameters.f90:488-511, the sweep's
      ! dt_orig/iscale (advection_dwarf_test_modular.f90:93-95)
      dx = dx00/REAL(n - 1, euwp)
      dy = dx
      dz = dx
      dt = (0.018_euwp*2._euwp*ACOS(-1._euwp))/REAL(iscale, euwp)

      ! sphere of radius 15 and value 4, solid-body rotation about (1,1,1)
      ALLOCATE (x(1 - ih:n + ih, 1 - ih:n + ih, 1 - ih:n + ih), source=0._euwp)
      ALLOCATE (ua, va, wa, ones, mold=x)
      ua = 0._euwp
      va = 0._euwp
      wa = 0._euwp
      ones = 1._euwp
      sqrt3i = 1._euwp/SQRT(3._euwp)
      x0c = 0.5_euwp*dx00
      y0c = 0.5_euwp*dx00
      z0c = 0.5_euwp*dx00
      xc = x0c - 0.25_euwp*dx00*sqrt3i
      yc = y0c + 0.25_euwp*dx00*sqrt3i
      zc = z0c + 0.25_euwp*dx00*sqrt3i
      r_squared = 15._euwp*15._euwp
      h_sphere = 4._euwp
      omega = 0.1_euwp/SQRT(3._euwp)
      omegax = omega
      omegay = omega
      omegaz = omega
      DO k = 1, n
         DO j = 1, n
            DO i = 1, n
               xp = REAL(i, euwp)*dx
               yp = REAL(j, euwp)*dy
               zp = REAL(k, euwp)*dz
               tmp = (xp - xc)*(xp - xc) + (yp - yc)*(yp - yc) + (zp - zc)*(zp - zc)
               IF (tmp <= r_squared) x(i, j, k) = h_sphere
               ua(i, j, k) = (-(omegaz*(REAL(j, euwp)*dy - y0c)) + omegay*(REAL(k, euwp)*dz - z0c))*dt/dx
               va(i, j, k) = (omegaz*(REAL(i, euwp)*dx - x0c) - omegax*(REAL(k, euwp)*dz - z0c))*dt/dy
               wa(i, j, k) = (-(omegay*(REAL(i, euwp)*dx - x0c)) + omegax*(REAL(j, euwp)*dy - y0c))*dt/dz
            END DO
         END DO
      END DO
      ALLOCATE (x0, source=x)

      ! device fields
      CALL mpdata_init(n, n, n, ih, 0, 0, 0)
      nsc = INT(n + 2*ih, c_long_long)**3
      nfx = INT(n + 1 + 2*ih, c_long_long)*INT(n + 2*ih, c_long_long)**2
      nfy = nfx
      nfz = nfx
      nbc = 2_c_long_long*n*n
      x_d = mpdata_alloc(nsc)
      x0_d = mpdata_alloc(nsc)
      ua_d = mpdata_alloc(nsc)
      va_d = mpdata_alloc(nsc)
      wa_d = mpdata_alloc(nsc)
      h_d = mpdata_alloc(nsc)
      rhr_d = mpdata_alloc(nsc)
      u1_d = mpdata_alloc(nfx)
      u2_d = mpdata_alloc(nfy)
      u3_d = mpdata_alloc(nfz)
      xlr_d = mpdata_alloc(nbc)       ! zero inflow (init_data)
      xbt_d = mpdata_alloc(nbc)
      CALL mpdata_put(x_d, x, nsc)
      CALL mpdata_put(x0_d, x0, nsc)
      CALL mpdata_put(ua_d, ua, nsc)
      CALL mpdata_put(va_d, va, nsc)
      CALL mpdata_put(wa_d, wa, nsc)
      CALL mpdata_put(h_d, ones, nsc)
      CALL mpdata_put(rhr_d, ones, nsc)
      ! C-grid advective velocities from the A-grid ones (open walls)
      CALL velprdA_to_C(ua_d, va_d, wa_d, u1_d, u2_d, u3_d)

      WRITE (*, '(A,I0,A,I0,A,I0,A,I0,A,I0,A,A)') 'mpdata_dwarf: iscale ', iscale, ', ', nt, &
         ' steps on ', n, ' x ', n, ' x ', n, ', ', mpdata_version()
      dumping = TRIM(dumpfile) /= '-'
      IF (dumping) &
         OPEN (NEWUNIT=unit_dump, FILE=TRIM(dumpfile), ACCESS='stream', FORM='unformatted', &
               STATUS='replace')

      CALL mpdata_sync()
      CALL ttini()
      CALL SYSTEM_CLOCK(c0, rate)
      DO it = 1, nt
         CALL ttbeg(300)
         CALL advec_dwarf(4, t_adv, .FALSE., .FALSE., n, n, n, ih, 0, 0, 0, 0, &
                          x_d, xlr_d, xbt_d, u1_d, u2_d, u3_d, rhr_d, h_d)
         CALL ttend(300)
         IF (dumping .AND. (it == 1 .OR. it == 50 .OR. it == 100 .OR. it == nt)) THEN
            CALL mpdata_get(x, x_d, nsc)
            WRITE (unit_dump) x(1:n, 1:n, 1:n)
         END IF
      END DO
      CALL mpdata_sync()
      CALL SYSTEM_CLOCK(c1)
      IF (dumping) CLOSE (unit_dump)

      ! norms: advection_test_setup.f90:472-492, 595-617
      CALL norm_sums(x_d, x0_d, s)
      nmli = 1._c_double/REAL(n, c_double)**3
      timetot = REAL(nt, c_double)*REAL(dt, c_double)
      energy = SQRT(ABS(s(1) - s(2))*nmli)/timetot
      error = SQRT(s(3)*nmli)/timetot
      WRITE (*, '(A,I0,1X,ES16.9)') 'Energy L2 norm: ', nt, energy
      WRITE (*, '(A,I0,1X,ES16.9)') 'Error  L2 norm: ', nt, error
      IF (n == 59 .AND. nt == 556) WRITE (*, '(A)') &
         'Published reference result for this case (Error L2): 0.0028'
      IF (c1 > c0) WRITE (*, '(A,ES12.5,A)') 'throughput: ', &
         REAL(n, 8)**3*REAL(nt, 8)/(REAL(c1 - c0, 8)/REAL(rate, 8)), ' cell-updates/s'
      WRITE (*, '(A)') ' Dwarf statistics summary'
      CALL ttprt()

      CALL mpdata_free(x_d)
      CALL mpdata_free(x0_d)
      CALL mpdata_free(ua_d)
      CALL mpdata_free(va_d)
      CALL mpdata_free(wa_d)
      CALL mpdata_free(h_d)
      CALL mpdata_free(rhr_d)
      CALL mpdata_free(u1_d)
      CALL mpdata_free(u2_d)
      CALL mpdata_free(u3_d)
      CALL mpdata_free(xlr_d)
      CALL mpdata_free(xbt_d)
      CALL mpdata_finalize()
   END SUBROUTINE run_case
END PROGRAM mpdata_dwarf
