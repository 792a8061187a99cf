! Many fields on one grid through the Fortran drop-in CE_advection
! (advection/src_drivers/advection_CEtimestep_twostep_GPUBC.f90:27-42):
! th, u, v, w, tht, pstr, qv and qc on a 30x22x14 open domain with one
! velocity, rhr and rhoadv, then each field again through mpdata_step with
! the variant CE_advection gives it.  Prints the largest difference per
! field and "CE check: PASS" when every field is identical.  With a
! directory as the first argument, CE_advection runs with the reference's
! serialization flags (savepoints mysnapshot.run-in / run-out in that
! directory), and the program also writes the inputs and results of its own
! per-field steps at savepoints check.in / check.out.
PROGRAM mpdata_ce_check
   USE iso_c_binding
   USE mpdata_hip
   IMPLICIT NONE
   INTEGER, PARAMETER :: np = 30, mp = 22, lp = 14, ih = 1, nf = 8
   INTEGER, PARAMETER :: roles(nf) = (/ 0, 1, 1, 1, 1, 1, 0, 0 /)
   INTEGER(c_long_long) :: ns, nu1, nu2, nu3, nbx, nby
   REAL(euwp), ALLOCATABLE :: hs(:, :, :), hu1(:, :, :), hu2(:, :, :), hu3(:, :, :)
   REAL(euwp), ALLOCATABLE :: hbx(:, :, :), hby(:, :, :), a(:, :, :), b(:, :, :)
   TYPE(c_ptr) :: x(nf), y(nf), bx(nf), by(nf), u1, u2, u3, rhr, h
   INTEGER :: f, i, j, k, variant
   REAL(euwp) :: d, worst
   CHARACTER(len=512) :: serdir
   LOGICAL :: ser
   CHARACTER(len=4), PARAMETER :: names(nf) = (/ 'th  ', 'u   ', 'v   ', 'w   ', 'tht ', 'pstr', &
                                                 'qv  ', 'qc  ' /)

   serdir = ''
   IF (COMMAND_ARGUMENT_COUNT() >= 1) CALL GET_COMMAND_ARGUMENT(1, serdir)
   ser = LEN_TRIM(serdir) > 0
   IF (ser) CALL mpdata_serializer_open(TRIM(serdir))

   CALL mpdata_init(np, mp, lp, ih, 0, 0, 0)
   ns = INT(np + 2 * ih, c_long_long) * (mp + 2 * ih) * (lp + 2 * ih)
   nu1 = INT(np + 1 + 2 * ih, c_long_long) * (mp + 2 * ih) * (lp + 2 * ih)
   nu2 = INT(np + 2 * ih, c_long_long) * (mp + 1 + 2 * ih) * (lp + 2 * ih)
   nu3 = INT(np + 2 * ih, c_long_long) * (mp + 2 * ih) * (lp + 1 + 2 * ih)
   nbx = INT(mp, c_long_long) * lp * 2
   nby = INT(np, c_long_long) * lp * 2
   ALLOCATE (hs(1 - ih:np + ih, 1 - ih:mp + ih, 1 - ih:lp + ih))
   ALLOCATE (hu1(1 - ih:np + 1 + ih, 1 - ih:mp + ih, 1 - ih:lp + ih))
   ALLOCATE (hu2(1 - ih:np + ih, 1 - ih:mp + 1 + ih, 1 - ih:lp + ih))
   ALLOCATE (hu3(1 - ih:np + ih, 1 - ih:mp + ih, 1 - ih:lp + 1 + ih))
   ALLOCATE (hbx(mp, lp, 2), hby(np, lp, 2), a(1 - ih:np + ih, 1 - ih:mp + ih, 1 - ih:lp + ih))
   ALLOCATE (b, mold=a)

   ! shared velocity, rhr, rhoadv (smooth, Courant below 0.4)
   DO k = 1 - ih, lp + ih
      DO j = 1 - ih, mp + ih
         DO i = 1 - ih, np + 1 + ih
            hu1(i, j, k) = 0.25_euwp * SIN(0.31_euwp * j + 0.17_euwp * k) + 0.05_euwp
         END DO
      END DO
   END DO
   DO k = 1 - ih, lp + ih
      DO j = 1 - ih, mp + 1 + ih
         DO i = 1 - ih, np + ih
            hu2(i, j, k) = 0.2_euwp * COS(0.23_euwp * i - 0.11_euwp * k)
         END DO
      END DO
   END DO
   DO k = 1 - ih, lp + 1 + ih
      DO j = 1 - ih, mp + ih
         DO i = 1 - ih, np + ih
            hu3(i, j, k) = 0.1_euwp * SIN(0.19_euwp * i + 0.29_euwp * j)
         END DO
      END DO
   END DO
   u1 = mpdata_alloc(nu1)
   u2 = mpdata_alloc(nu2)
   u3 = mpdata_alloc(nu3)
   CALL mpdata_put(u1, hu1, nu1)
   CALL mpdata_put(u2, hu2, nu2)
   CALL mpdata_put(u3, hu3, nu3)
   hs = 1.0_euwp
   FORALL (i = 1 - ih:np + ih, j = 1 - ih:mp + ih, k = 1 - ih:lp + ih) &
      hs(i, j, k) = 1.0_euwp + 0.1_euwp * COS(0.2_euwp * i + 0.3_euwp * j + 0.4_euwp * k)
   h = mpdata_alloc(ns)
   CALL mpdata_put(h, hs, ns)
   FORALL (i = 1 - ih:np + ih, j = 1 - ih:mp + ih, k = 1 - ih:lp + ih) &
      hs(i, j, k) = 1.0_euwp + 0.1_euwp * SIN(0.3_euwp * i - 0.2_euwp * j + 0.1_euwp * k)
   rhr = mpdata_alloc(ns)
   CALL mpdata_put(rhr, hs, ns)

   ! fields: two copies each; inflow values per field
   DO f = 1, nf
      FORALL (i = 1 - ih:np + ih, j = 1 - ih:mp + ih, k = 1 - ih:lp + ih) &
         hs(i, j, k) = REAL(roles(f), euwp) * (-0.5_euwp) + 1.0_euwp + 0.8_euwp * &
                       SIN(0.21_euwp * f * i + 0.37_euwp * j - 0.13_euwp * k * f)
      x(f) = mpdata_alloc(ns)
      y(f) = mpdata_alloc(ns)
      CALL mpdata_put(x(f), hs, ns)
      CALL mpdata_put(y(f), hs, ns)
      hbx = 0.5_euwp + 0.1_euwp * f
      hby = 0.7_euwp - 0.05_euwp * f
      bx(f) = mpdata_alloc(nbx)
      by(f) = mpdata_alloc(nby)
      CALL mpdata_put(bx(f), hbx, nbx)
      CALL mpdata_put(by(f), hby, nby)
   END DO

   IF (ser) THEN
      DO f = 1, nf
         CALL mpdata_serialize('check.in', TRIM(names(f)), y(f), np + 2*ih, mp + 2*ih, lp + 2*ih)
      END DO
   END IF
   CALL CE_advection(.FALSE., 0, 0, 0, 0, 0, 0, np, mp, lp, ih, ser, ser, &
                     rhr, h, u1, u2, u3, &
                     x(1), bx(1), by(1), x(2), bx(2), by(2), x(3), bx(3), by(3), &
                     x(4), bx(4), by(4), x(5), bx(5), by(5), x(6), bx(6), by(6), &
                     qv=x(7), qv_lr=bx(7), qv_bt=by(7), qc=x(8), qc_lr=bx(8), qc_bt=by(8))
   DO f = 1, nf
      variant = MERGE(MPDATA_GAUGE, MPDATA_STANDARD, roles(f) == 1)
      CALL mpdata_step(variant, 2, .TRUE., y(f), u1, u2, u3, rhr, h, bx(f), by(f))
   END DO
   CALL mpdata_sync()
   IF (ser) THEN
      DO f = 1, nf
         CALL mpdata_serialize('check.out', TRIM(names(f)), y(f), np + 2*ih, mp + 2*ih, lp + 2*ih)
      END DO
      CALL mpdata_serializer_close()
   END IF
   worst = 0.0_euwp
   DO f = 1, nf
      CALL mpdata_get(a, x(f), ns)
      CALL mpdata_get(b, y(f), ns)
      d = MAXVAL(ABS(a(1:np, 1:mp, 1:lp) - b(1:np, 1:mp, 1:lp)))
      WRITE (*, '(A,I0,A,ES12.4,A,ES12.4)') 'field ', f, ': max diff ', d, '  max |x| ', &
         MAXVAL(ABS(a(1:np, 1:mp, 1:lp)))
      worst = MAX(worst, d)
   END DO
   IF (worst == 0.0_euwp) THEN
      WRITE (*, '(A)') 'CE check: PASS'
   ELSE
      WRITE (*, '(A)') 'CE check: FAIL'
      ERROR STOP 1
   END IF
   CALL mpdata_finalize()
END PROGRAM mpdata_ce_check
