! Fortran host binding of libmpdata_hip.so (C ABI: include/mpdata_hip.h).
!
! The reference's advection driver dispatches through procedure pointers to
! upwind3d_gpubc and antidiff3d_{gauge,standard}_gpubc
! (advection/src_drivers/advec.f90:143-144, 221-240) and keeps its scratch in
! module globals (common/scratch_datafields_modular.f90).  This module keeps
! the same shape: mpdata_init creates one module-global library context for
! the domain (the scratch), and the drop-in procedures below carry the
! reference's names and argument lists.  Fields live in GPU memory; where the
! reference declares them MANAGED (advection/include/defines.inc:1-11) a host
! that links this module holds them as TYPE(C_PTR) device handles from
! mpdata_alloc, filled and read back with mpdata_put / mpdata_get.
!
! Errors: every library call returns a status; a non-zero status prints the
! library's message and ends the run with ERROR STOP, as the reference ends
! with STOP on its own errors.
!
! Precision follows the reference's switch (common/precisions.f90:9-17):
! compiled with -DFLOAT_PRECISION=4 the module binds the _f32 entry points,
! otherwise the _f64 ones.
#if FLOAT_PRECISION == 4
#define MP_SFX "_f32"
#else
#define MP_SFX "_f64"
#endif
MODULE mpdata_hip
   USE iso_c_binding
   IMPLICIT NONE
   PRIVATE

#if FLOAT_PRECISION == 4
   INTEGER, PARAMETER, PUBLIC :: euwp = c_float
#else
   INTEGER, PARAMETER, PUBLIC :: euwp = c_double
#endif
   INTEGER(c_int), PARAMETER, PUBLIC :: MPDATA_OK = 0, MPDATA_EINVAL = 1, MPDATA_EHIP = 2, &
                                        MPDATA_ENOMEM = 3, MPDATA_EUNSUPPORTED = 4, &
                                        MPDATA_ESTOP = 5
   INTEGER(c_int), PARAMETER, PUBLIC :: MPDATA_UPWIND = 0, MPDATA_GAUGE = 1, MPDATA_STANDARD = 2
   INTEGER(c_int), PARAMETER, PUBLIC :: MPDATA_PATH_AUTO = 0, MPDATA_PATH_STAGES = 1, &
                                        MPDATA_PATH_FUSED = 2

   PUBLIC :: mpdata_init, mpdata_finalize, mpdata_set_path, mpdata_context
   PUBLIC :: mpdata_alloc, mpdata_free, mpdata_put, mpdata_get, mpdata_sync, mpdata_check
   PUBLIC :: upwind3d_gpubc, antidiff3d_gauge_gpubc, antidiff3d_standard_gpubc, advec_dwarf
   PUBLIC :: mpdata_step, mpdata_timeloop, velprdA_to_C, compute_courants, norm_sums
   PUBLIC :: mpdata_version, CE_advection, mpdata3d_rho_gauge_gpubc, velprd_driver
   PUBLIC :: mpdata_serializer_open, mpdata_serializer_close, mpdata_serialize
   PUBLIC :: ttini, ttbeg, ttend, ttprt, mpdata_ensure_domain

   ! the module-global context (the reference's module-global scratch) and
   ! the domain it was created for (np, mp, lp, ih, ibcx, ibcy, ibcz)
   TYPE(c_ptr), SAVE :: ctx = c_null_ptr
   INTEGER, SAVE :: ctx_dims(7) = 0

   INTERFACE
      FUNCTION c_serializer_open(dir) BIND(C, name="mpdata_serializer_open")
         IMPORT :: c_int, c_char
         CHARACTER(KIND=c_char), DIMENSION(*), INTENT(IN) :: dir
         INTEGER(c_int) :: c_serializer_open
      END FUNCTION
      FUNCTION c_serializer_close() BIND(C, name="mpdata_serializer_close")
         IMPORT :: c_int
         INTEGER(c_int) :: c_serializer_close
      END FUNCTION
      FUNCTION c_serialize_field(savepoint, name, field, n1, n2, n3, stream) &
            BIND(C, name="mpdata_serialize_field"//MP_SFX)
         IMPORT :: c_int, c_char, c_ptr
         CHARACTER(KIND=c_char), DIMENSION(*), INTENT(IN) :: savepoint, name
         TYPE(c_ptr), VALUE :: field, stream
         INTEGER(c_int), VALUE :: n1, n2, n3
         INTEGER(c_int) :: c_serialize_field
      END FUNCTION
      FUNCTION c_ttini() BIND(C, name="mpdata_ttini")
         IMPORT :: c_int
         INTEGER(c_int) :: c_ttini
      END FUNCTION
      FUNCTION c_ttbeg(id, stream) BIND(C, name="mpdata_ttbeg")
         IMPORT :: c_int, c_ptr
         INTEGER(c_int), VALUE :: id
         TYPE(c_ptr), VALUE :: stream
         INTEGER(c_int) :: c_ttbeg
      END FUNCTION
      FUNCTION c_ttend(id, stream) BIND(C, name="mpdata_ttend")
         IMPORT :: c_int, c_ptr
         INTEGER(c_int), VALUE :: id
         TYPE(c_ptr), VALUE :: stream
         INTEGER(c_int) :: c_ttend
      END FUNCTION
      FUNCTION c_ttprt() BIND(C, name="mpdata_ttprt")
         IMPORT :: c_int
         INTEGER(c_int) :: c_ttprt
      END FUNCTION
      FUNCTION c_version() BIND(C, name="mpdata_version")
         IMPORT :: c_ptr
         TYPE(c_ptr) :: c_version
      END FUNCTION
      FUNCTION c_last_error() BIND(C, name="mpdata_last_error")
         IMPORT :: c_ptr
         TYPE(c_ptr) :: c_last_error
      END FUNCTION
      FUNCTION c_strlen(s) BIND(C, name="strlen")
         IMPORT :: c_ptr, c_size_t
         TYPE(c_ptr), VALUE :: s
         INTEGER(c_size_t) :: c_strlen
      END FUNCTION
      FUNCTION c_ctx_create(np, mp, lp, ih, ibcx, ibcy, ibcz, real_bytes, out) &
            BIND(C, name="mpdata_ctx_create")
         IMPORT :: c_int, c_ptr
         INTEGER(c_int), VALUE :: np, mp, lp, ih, ibcx, ibcy, ibcz, real_bytes
         TYPE(c_ptr), INTENT(OUT) :: out
         INTEGER(c_int) :: c_ctx_create
      END FUNCTION
      FUNCTION c_ctx_destroy(c) BIND(C, name="mpdata_ctx_destroy")
         IMPORT :: c_int, c_ptr
         TYPE(c_ptr), VALUE :: c
         INTEGER(c_int) :: c_ctx_destroy
      END FUNCTION
      FUNCTION c_ctx_set_path(c, path) BIND(C, name="mpdata_ctx_set_path")
         IMPORT :: c_int, c_ptr
         TYPE(c_ptr), VALUE :: c
         INTEGER(c_int), VALUE :: path
         INTEGER(c_int) :: c_ctx_set_path
      END FUNCTION
      FUNCTION c_dev_alloc(bytes, p) BIND(C, name="mpdata_dev_alloc")
         IMPORT :: c_int, c_ptr, c_long_long
         INTEGER(c_long_long), VALUE :: bytes
         TYPE(c_ptr), INTENT(OUT) :: p
         INTEGER(c_int) :: c_dev_alloc
      END FUNCTION
      FUNCTION c_dev_free(p) BIND(C, name="mpdata_dev_free")
         IMPORT :: c_int, c_ptr
         TYPE(c_ptr), VALUE :: p
         INTEGER(c_int) :: c_dev_free
      END FUNCTION
      FUNCTION c_copy(dst, src, bytes, kind) BIND(C, name="mpdata_copy")
         IMPORT :: c_int, c_ptr, c_long_long
         TYPE(c_ptr), VALUE :: dst, src
         INTEGER(c_long_long), VALUE :: bytes
         INTEGER(c_int), VALUE :: kind
         INTEGER(c_int) :: c_copy
      END FUNCTION
      FUNCTION c_sync(stream) BIND(C, name="mpdata_sync")
         IMPORT :: c_int, c_ptr
         TYPE(c_ptr), VALUE :: stream
         INTEGER(c_int) :: c_sync
      END FUNCTION
      FUNCTION c_upwind(c, u1, u2, u3, x_in, xant, xforc_impl, xforc_expl, rhr, h, bcx, bcy, &
                        ipoles0, ibcx0, ibcy0, ibcz0, np, mp, lp, ih, dt, stream) &
            BIND(C, name="mpdata_upwind3d_gpubc"//MP_SFX)
         IMPORT :: c_int, c_ptr, euwp
         TYPE(c_ptr), VALUE :: c, u1, u2, u3, x_in, xant, xforc_impl, xforc_expl, rhr, h, bcx, bcy
         INTEGER(c_int), VALUE :: ipoles0, ibcx0, ibcy0, ibcz0, np, mp, lp, ih
         REAL(euwp), VALUE :: dt
         TYPE(c_ptr), VALUE :: stream
         INTEGER(c_int) :: c_upwind
      END FUNCTION
      FUNCTION c_antidiff_gauge(c, lupdatemulti, u1, u2, u3, xant, x_in, rhr, h, iflip, &
                                ipoles0, ibcx0, ibcy0, ibcz0, np, mp, lp, ih, stream) &
            BIND(C, name="mpdata_antidiff3d_gauge_gpubc"//MP_SFX)
         IMPORT :: c_int, c_ptr
         TYPE(c_ptr), VALUE :: c
         INTEGER(c_int), VALUE :: lupdatemulti
         TYPE(c_ptr), VALUE :: u1, u2, u3, xant, x_in, rhr, h
         INTEGER(c_int), VALUE :: iflip, ipoles0, ibcx0, ibcy0, ibcz0, np, mp, lp, ih
         TYPE(c_ptr), VALUE :: stream
         INTEGER(c_int) :: c_antidiff_gauge
      END FUNCTION
      FUNCTION c_antidiff_standard(c, lupdatemulti, u1, u2, u3, xant, x_in, rhr, h, iflip, &
                                   ipoles0, ibcx0, ibcy0, ibcz0, np, mp, lp, ih, stream) &
            BIND(C, name="mpdata_antidiff3d_standard_gpubc"//MP_SFX)
         IMPORT :: c_int, c_ptr
         TYPE(c_ptr), VALUE :: c
         INTEGER(c_int), VALUE :: lupdatemulti
         TYPE(c_ptr), VALUE :: u1, u2, u3, xant, x_in, rhr, h
         INTEGER(c_int), VALUE :: iflip, ipoles0, ibcx0, ibcy0, ibcz0, np, mp, lp, ih
         TYPE(c_ptr), VALUE :: stream
         INTEGER(c_int) :: c_antidiff_standard
      END FUNCTION
      FUNCTION c_advec_dwarf(c, legacyoptmode, t_adv, lupdatemulti, lvertsplit, np, mp, lp, ih, &
                             ipoles, ibcx, ibcy, ibcz, x, x_lr, x_bt, uadv, vadv, wadv, rhr, &
                             rhoadv, stream) BIND(C, name="mpdata_advec_dwarf"//MP_SFX)
         IMPORT :: c_int, c_ptr
         TYPE(c_ptr), VALUE :: c
         INTEGER(c_int), VALUE :: legacyoptmode, t_adv, lupdatemulti, lvertsplit
         INTEGER(c_int), VALUE :: np, mp, lp, ih, ipoles, ibcx, ibcy, ibcz
         TYPE(c_ptr), VALUE :: x, x_lr, x_bt, uadv, vadv, wadv, rhr, rhoadv, stream
         INTEGER(c_int) :: c_advec_dwarf
      END FUNCTION
      FUNCTION c_step(c, variant, iters, nonos, x, u1, u2, u3, rhr, h, bcx, bcy, stream) &
            BIND(C, name="mpdata_step"//MP_SFX)
         IMPORT :: c_int, c_ptr
         TYPE(c_ptr), VALUE :: c
         INTEGER(c_int), VALUE :: variant, iters, nonos
         TYPE(c_ptr), VALUE :: x, u1, u2, u3, rhr, h, bcx, bcy, stream
         INTEGER(c_int) :: c_step
      END FUNCTION
      FUNCTION c_timeloop(c, variant, iters, nonos, nt, x, u1, u2, u3, rhr, h, bcx, bcy, stream) &
            BIND(C, name="mpdata_c_timeloop"//MP_SFX)
         IMPORT :: c_int, c_ptr
         TYPE(c_ptr), VALUE :: c
         INTEGER(c_int), VALUE :: variant, iters, nonos, nt
         TYPE(c_ptr), VALUE :: x, u1, u2, u3, rhr, h, bcx, bcy, stream
         INTEGER(c_int) :: c_timeloop
      END FUNCTION
      FUNCTION c_a_to_c(c, ox, oy, oz, u1, u2, u3, stream) BIND(C, name="mpdata_velprdA_to_C"//MP_SFX)
         IMPORT :: c_int, c_ptr
         TYPE(c_ptr), VALUE :: c, ox, oy, oz, u1, u2, u3, stream
         INTEGER(c_int) :: c_a_to_c
      END FUNCTION
      FUNCTION c_courants(c, ox, oy, oz, h, ntimelevel, dxi, dyi, dzi, dt, out, stream) &
            BIND(C, name="mpdata_compute_courants"//MP_SFX)
         IMPORT :: c_int, c_ptr, c_double, euwp
         TYPE(c_ptr), VALUE :: c, ox, oy, oz, h
         INTEGER(c_int), VALUE :: ntimelevel
         REAL(euwp), VALUE :: dxi, dyi, dzi, dt
         REAL(c_double), INTENT(OUT) :: out(7)
         TYPE(c_ptr), VALUE :: stream
         INTEGER(c_int) :: c_courants
      END FUNCTION
      FUNCTION c_ce_advection(c, alg_choice, liner, nfields, role, x, bcx, bcy, u1, u2, u3, &
                              rhr, h, stream) BIND(C, name="mpdata_ce_advection"//MP_SFX)
         IMPORT :: c_ptr, c_int
         TYPE(c_ptr), VALUE :: c, u1, u2, u3, rhr, h, stream
         INTEGER(c_int), VALUE :: alg_choice, liner, nfields
         INTEGER(c_int), INTENT(IN) :: role(*)
         TYPE(c_ptr), INTENT(IN) :: x(*), bcx(*), bcy(*)
         INTEGER(c_int) :: c_ce_advection
      END FUNCTION
      FUNCTION c_velprd_driver(c, nsubsteps, nsubsteps_o, lsubstepping, dtn, dtnold, &
                               ltimeadapt, lvertsplit, lcrdiag_nphlf, crmaxlim, dtmax, ox, oy, &
                               oz, u, v, w, uadv, vadv, wadv, h, ipoles0, ibcx0, ibcy0, ibcz0, &
                               np, mp, lp, ih, dxi, dyi, dzi, dt, stream) &
            BIND(C, name="mpdata_velprd_driver"//MP_SFX)
         IMPORT :: c_int, c_ptr, euwp
         TYPE(c_ptr), VALUE :: c
         INTEGER(c_int), INTENT(INOUT) :: nsubsteps, nsubsteps_o
         REAL(euwp), INTENT(INOUT) :: dtn, dtnold
         INTEGER(c_int), VALUE :: lsubstepping, ltimeadapt, lvertsplit, lcrdiag_nphlf
         REAL(euwp), VALUE :: crmaxlim, dtmax, dxi, dyi, dzi, dt
         TYPE(c_ptr), VALUE :: ox, oy, oz, u, v, w, uadv, vadv, wadv, h, stream
         INTEGER(c_int), VALUE :: ipoles0, ibcx0, ibcy0, ibcz0, np, mp, lp, ih
         INTEGER(c_int) :: c_velprd_driver
      END FUNCTION
      FUNCTION c_rho_gauge(c, lupdatemulti, nsubsteps, nsubsteps_o, lsubstepping, dtn, dtnold, &
                           ltimeadapt, lvertsplit, lcrdiag_nphlf, crmaxlim, dtmax, ox, oy, oz, &
                           u, v, w, u1, u2, u3, x_in, x_incomp, h, rhoadv, rhr, pexe, pexr1, &
                           pexr2, bcx, bcy, ipoles0, ibcx0, ibcy0, ibcz0, np, mp, lp, ih, dxi, &
                           dyi, dzi, dt, liner, stream) &
            BIND(C, name="mpdata3d_rho_gauge_gpubc"//MP_SFX)
         IMPORT :: c_int, c_ptr, euwp
         TYPE(c_ptr), VALUE :: c
         INTEGER(c_int), VALUE :: lupdatemulti, lsubstepping, ltimeadapt, lvertsplit, &
                                  lcrdiag_nphlf, liner
         INTEGER(c_int), INTENT(INOUT) :: nsubsteps, nsubsteps_o
         REAL(euwp), INTENT(INOUT) :: dtn, dtnold
         REAL(euwp), VALUE :: crmaxlim, dtmax, dxi, dyi, dzi, dt
         TYPE(c_ptr), VALUE :: ox, oy, oz, u, v, w, u1, u2, u3, x_in, x_incomp, h, rhoadv, rhr
         TYPE(c_ptr), VALUE :: pexe, pexr1, pexr2, bcx, bcy, stream
         INTEGER(c_int), VALUE :: ipoles0, ibcx0, ibcy0, ibcz0, np, mp, lp, ih
         INTEGER(c_int) :: c_rho_gauge
      END FUNCTION
      FUNCTION c_norm_sums(c, x, x0, out, stream) BIND(C, name="mpdata_norm_sums"//MP_SFX)
         IMPORT :: c_int, c_ptr, c_double
         TYPE(c_ptr), VALUE :: c, x, x0
         REAL(c_double), INTENT(OUT) :: out(4)
         TYPE(c_ptr), VALUE :: stream
         INTEGER(c_int) :: c_norm_sums
      END FUNCTION
   END INTERFACE

CONTAINS

   FUNCTION cstr(p) RESULT(s)
      TYPE(c_ptr), INTENT(IN) :: p
      CHARACTER(LEN=:), ALLOCATABLE :: s
      CHARACTER(KIND=c_char), POINTER :: chars(:)
      INTEGER :: n, i
      IF (.NOT. c_associated(p)) THEN
         s = ''
         RETURN
      END IF
      n = INT(c_strlen(p))
      CALL c_f_pointer(p, chars, [n])
      ALLOCATE (CHARACTER(LEN=n) :: s)
      DO i = 1, n
         s(i:i) = chars(i)
      END DO
   END FUNCTION cstr

   FUNCTION mpdata_version() RESULT(s)
      CHARACTER(LEN=:), ALLOCATABLE :: s
      s = cstr(c_version())
   END FUNCTION mpdata_version

   ! Non-zero status: print the library's message and stop.
   SUBROUTINE mpdata_check(rc, what)
      INTEGER(c_int), INTENT(IN) :: rc
      CHARACTER(LEN=*), INTENT(IN) :: what
      IF (rc /= MPDATA_OK) THEN
         WRITE (*, '(A,A,I0,A,A)') what, ': mpdata error ', rc, ': ', cstr(c_last_error())
         ERROR STOP 1
      END IF
   END SUBROUTINE mpdata_check

   INTEGER(c_int) FUNCTION flag(l)
      LOGICAL, INTENT(IN) :: l
      flag = MERGE(1_c_int, 0_c_int, l)
   END FUNCTION flag

   ! Domain set-up: the reference's geomset/allocate_scratch_* for one
   ! process (common/mpi_parallel.f90:88-501, scratch_datafields_modular.f90).
   SUBROUTINE mpdata_init(np, mp, lp, ih, ibcx, ibcy, ibcz)
      INTEGER, INTENT(IN) :: np, mp, lp, ih, ibcx, ibcy, ibcz
      IF (c_associated(ctx)) CALL mpdata_finalize()
      CALL mpdata_check(c_ctx_create(INT(np, c_int), INT(mp, c_int), INT(lp, c_int), &
                                     INT(ih, c_int), INT(ibcx, c_int), INT(ibcy, c_int), &
                                     INT(ibcz, c_int), INT(euwp, c_int), ctx), 'mpdata_init')
      ctx_dims = (/ np, mp, lp, ih, ibcx, ibcy, ibcz /)
   END SUBROUTINE mpdata_init

   ! mpdata_init unless the context already holds this domain
   SUBROUTINE mpdata_ensure_domain(np, mp, lp, ih, ibcx, ibcy, ibcz)
      INTEGER, INTENT(IN) :: np, mp, lp, ih, ibcx, ibcy, ibcz
      IF (c_associated(ctx) .AND. ALL(ctx_dims == (/ np, mp, lp, ih, ibcx, ibcy, ibcz /))) RETURN
      CALL mpdata_init(np, mp, lp, ih, ibcx, ibcy, ibcz)
   END SUBROUTINE mpdata_ensure_domain

   SUBROUTINE mpdata_finalize()
      IF (c_associated(ctx)) CALL mpdata_check(c_ctx_destroy(ctx), 'mpdata_finalize')
      ctx = c_null_ptr
   END SUBROUTINE mpdata_finalize

   SUBROUTINE mpdata_set_path(path)
      INTEGER, INTENT(IN) :: path
      CALL mpdata_check(c_ctx_set_path(ctx, INT(path, c_int)), 'mpdata_set_path')
   END SUBROUTINE mpdata_set_path

   FUNCTION mpdata_context() RESULT(c)
      TYPE(c_ptr) :: c
      c = ctx
   END FUNCTION mpdata_context

   ! ---- device fields ------------------------------------------------------
   FUNCTION mpdata_alloc(nelem) RESULT(p)
      INTEGER(c_long_long), INTENT(IN) :: nelem
      TYPE(c_ptr) :: p
      CALL mpdata_check(c_dev_alloc(nelem*INT(euwp, c_long_long), p), 'mpdata_alloc')
   END FUNCTION mpdata_alloc

   SUBROUTINE mpdata_free(p)
      TYPE(c_ptr), INTENT(INOUT) :: p
      CALL mpdata_check(c_dev_free(p), 'mpdata_free')
      p = c_null_ptr
   END SUBROUTINE mpdata_free

   ! host array (any rank, contiguous) -> device field of nelem elements
   SUBROUTINE mpdata_put(dev, host, nelem)
      TYPE(c_ptr), INTENT(IN) :: dev
      INTEGER(c_long_long), INTENT(IN) :: nelem
      REAL(euwp), INTENT(IN), TARGET :: host(*)
      CALL mpdata_check(c_copy(dev, c_loc(host(1)), nelem*INT(euwp, c_long_long), 0_c_int), &
                        'mpdata_put')
   END SUBROUTINE mpdata_put

   SUBROUTINE mpdata_get(host, dev, nelem)
      REAL(euwp), INTENT(INOUT), TARGET :: host(*)
      TYPE(c_ptr), INTENT(IN) :: dev
      INTEGER(c_long_long), INTENT(IN) :: nelem
      CALL mpdata_check(c_copy(c_loc(host(1)), dev, nelem*INT(euwp, c_long_long), 1_c_int), &
                        'mpdata_get')
   END SUBROUTINE mpdata_get

   SUBROUTINE mpdata_sync()
      CALL mpdata_check(c_sync(c_null_ptr), 'mpdata_sync')
   END SUBROUTINE mpdata_sync

   ! ---- drop-ins with the reference's argument lists -----------------------
   ! upwind3d_gpubc: advection/src_algorithms/upwind3d-gpubc.f90:20-28
   SUBROUTINE upwind3d_gpubc(u1, u2, u3, x_in, xant, xforc_impl, xforc_expl, rhr, h, bcx, bcy, &
                             ipoles0, ibcx0, ibcy0, ibcz0, np, mp, lp, ih, dt, &
                             do_serialization_in, do_serialization_out)
      TYPE(c_ptr), INTENT(IN) :: u1, u2, u3, x_in, xant, xforc_impl, xforc_expl, rhr, h, bcx, bcy
      INTEGER, INTENT(IN) :: ipoles0, ibcx0, ibcy0, ibcz0, np, mp, lp, ih
      REAL(euwp), INTENT(IN) :: dt
      LOGICAL, INTENT(IN), OPTIONAL :: do_serialization_in, do_serialization_out
      CALL mpdata_check(c_upwind(ctx, u1, u2, u3, x_in, xant, xforc_impl, xforc_expl, rhr, h, &
                                 bcx, bcy, INT(ipoles0, c_int), INT(ibcx0, c_int), &
                                 INT(ibcy0, c_int), INT(ibcz0, c_int), INT(np, c_int), &
                                 INT(mp, c_int), INT(lp, c_int), INT(ih, c_int), dt, c_null_ptr), &
                        'upwind3d_gpubc')
   END SUBROUTINE upwind3d_gpubc

   ! antidiff3d_gauge_gpubc: advection/src_algorithms/antidiff3d-gauge-gpubc.f90:14-16
   SUBROUTINE antidiff3d_gauge_gpubc(lupdatemulti, u1, u2, u3, xant, x_in, rhr, h, iflip, &
                                     ipoles0, ibcx0, ibcy0, ibcz0, np, mp, lp, ih, &
                                     do_serialization_out)
      LOGICAL, INTENT(IN) :: lupdatemulti
      TYPE(c_ptr), INTENT(IN) :: u1, u2, u3, xant, x_in, rhr, h
      INTEGER, INTENT(IN) :: iflip, ipoles0, ibcx0, ibcy0, ibcz0, np, mp, lp, ih
      LOGICAL, INTENT(IN), OPTIONAL :: do_serialization_out
      CALL mpdata_check(c_antidiff_gauge(ctx, flag(lupdatemulti), u1, u2, u3, xant, x_in, rhr, &
                                         h, INT(iflip, c_int), INT(ipoles0, c_int), &
                                         INT(ibcx0, c_int), INT(ibcy0, c_int), &
                                         INT(ibcz0, c_int), INT(np, c_int), INT(mp, c_int), &
                                         INT(lp, c_int), INT(ih, c_int), c_null_ptr), &
                        'antidiff3d_gauge_gpubc')
   END SUBROUTINE antidiff3d_gauge_gpubc

   ! antidiff3d_standard_gpubc: advection/src_algorithms/antidiff3d-standard-gpubc.f90:14-16
   SUBROUTINE antidiff3d_standard_gpubc(lupdatemulti, u1, u2, u3, xant, x_in, rhr, h, iflip, &
                                        ipoles0, ibcx0, ibcy0, ibcz0, np, mp, lp, ih, &
                                        do_serialization_out)
      LOGICAL, INTENT(IN) :: lupdatemulti
      TYPE(c_ptr), INTENT(IN) :: u1, u2, u3, xant, x_in, rhr, h
      INTEGER, INTENT(IN) :: iflip, ipoles0, ibcx0, ibcy0, ibcz0, np, mp, lp, ih
      LOGICAL, INTENT(IN), OPTIONAL :: do_serialization_out
      CALL mpdata_check(c_antidiff_standard(ctx, flag(lupdatemulti), u1, u2, u3, xant, x_in, &
                                            rhr, h, INT(iflip, c_int), INT(ipoles0, c_int), &
                                            INT(ibcx0, c_int), INT(ibcy0, c_int), &
                                            INT(ibcz0, c_int), INT(np, c_int), INT(mp, c_int), &
                                            INT(lp, c_int), INT(ih, c_int), c_null_ptr), &
                        'antidiff3d_standard_gpubc')
   END SUBROUTINE antidiff3d_standard_gpubc

   ! advec_dwarf: advection/src_drivers/advec.f90:116-122 (the optional
   ! trailing arrays of the reference are unused there and omitted here)
   SUBROUTINE advec_dwarf(legacyoptmode, t_adv, lupdatemulti, lvertsplit, np, mp, lp, ih, &
                          ipoles, ibcx, ibcy, ibcz, x, x_lr, x_bt, uadv, vadv, wadv, rhr, rhoadv)
      INTEGER, INTENT(IN) :: legacyoptmode, t_adv
      LOGICAL, INTENT(IN) :: lupdatemulti, lvertsplit
      INTEGER, INTENT(IN) :: np, mp, lp, ih, ipoles, ibcx, ibcy, ibcz
      TYPE(c_ptr), INTENT(IN) :: x, x_lr, x_bt, uadv, vadv, wadv, rhr, rhoadv
      CALL mpdata_check(c_advec_dwarf(ctx, INT(legacyoptmode, c_int), INT(t_adv, c_int), &
                                      flag(lupdatemulti), flag(lvertsplit), INT(np, c_int), &
                                      INT(mp, c_int), INT(lp, c_int), INT(ih, c_int), &
                                      INT(ipoles, c_int), INT(ibcx, c_int), INT(ibcy, c_int), &
                                      INT(ibcz, c_int), x, x_lr, x_bt, uadv, vadv, wadv, rhr, &
                                      rhoadv, c_null_ptr), 'advec_dwarf')
   END SUBROUTINE advec_dwarf

   ! CE_advection: advection/src_drivers/advection_CEtimestep_twostep_GPUBC.f90:27-42.
   ! One step of th, u, v, w, tht, pstr and the moist fields present, with
   ! one velocity (uadv, vadv, wadv), rhr and rhoadv; u, v, w, tht and pstr
   ! take the antidiffusive pass the reference calls mpdatm3d_antidiff, th
   ! and the moist fields mpdata3d_antidiff (:145-155, :270-304).  lupdatemulti
   ! and iflip have no effect on one process without poles.
   ! Timers keyed by the reference's ids (common/mpi_parallel.f90:7088-7447),
   ! HIP events on the default stream; ttini turns them on and resets them.
   SUBROUTINE ttini()
      CALL mpdata_check(c_ttini(), 'ttini')
   END SUBROUTINE ttini

   SUBROUTINE ttbeg(id)
      INTEGER, INTENT(IN) :: id
      CALL mpdata_check(c_ttbeg(INT(id, c_int), c_null_ptr), 'ttbeg')
   END SUBROUTINE ttbeg

   SUBROUTINE ttend(id)
      INTEGER, INTENT(IN) :: id
      CALL mpdata_check(c_ttend(INT(id, c_int), c_null_ptr), 'ttend')
   END SUBROUTINE ttend

   SUBROUTINE ttprt()
      FLUSH (6)
      CALL mpdata_check(c_ttprt(), 'ttprt')
   END SUBROUTINE ttprt

   ! Serialization savepoints (the reference's serialbox calls under
   ! SERIALIZE): fields written as <dir>/<savepoint>__<name>.npy with an index
   ! line in <dir>/savepoints.txt.
   SUBROUTINE mpdata_serializer_open(dir)
      CHARACTER(*), INTENT(IN) :: dir
      CALL mpdata_check(c_serializer_open(TRIM(dir)//c_null_char), 'mpdata_serializer_open')
   END SUBROUTINE mpdata_serializer_open

   SUBROUTINE mpdata_serializer_close()
      CALL mpdata_check(c_serializer_close(), 'mpdata_serializer_close')
   END SUBROUTINE mpdata_serializer_close

   ! one field (device handle or host C_LOC) of the Fortran shape n1 x n2 x n3
   SUBROUTINE mpdata_serialize(savepoint, name, field, n1, n2, n3)
      CHARACTER(*), INTENT(IN) :: savepoint, name
      TYPE(c_ptr), INTENT(IN) :: field
      INTEGER, INTENT(IN) :: n1, n2, n3
      CALL mpdata_check(c_serialize_field(TRIM(savepoint)//c_null_char, TRIM(name)//c_null_char, &
                                          field, INT(n1, c_int), INT(n2, c_int), INT(n3, c_int), &
                                          c_null_ptr), 'mpdata_serialize')
   END SUBROUTINE mpdata_serialize

   SUBROUTINE CE_advection(lupdatemulti, mpdata_alg_choice, liner, ipoles, ibcx, ibcy, ibcz, &
                           np, mp, lp, ih, do_serialization_in, do_serialization_out, &
                           rhr, rhoadv, uadv, vadv, wadv, &
                           th, th_lr, th_bt, u, u_lr, u_bt, v, v_lr, v_bt, w, w_lr, w_bt, &
                           tht, tht_lr, tht_bt, pstr, pstr_lr, pstr_bt, &
                           qv, qv_lr, qv_bt, qc, qc_lr, qc_bt, qr, qr_lr, qr_bt, &
                           qi, qi_lr, qi_bt, qs, qs_lr, qs_bt, qg, qg_lr, qg_bt)
      LOGICAL, INTENT(IN) :: lupdatemulti
      INTEGER, INTENT(IN) :: mpdata_alg_choice, liner, ipoles, ibcx, ibcy, ibcz, np, mp, lp, ih
      LOGICAL, INTENT(IN), OPTIONAL :: do_serialization_in, do_serialization_out
      TYPE(c_ptr), INTENT(IN) :: rhr, rhoadv, uadv, vadv, wadv
      TYPE(c_ptr), INTENT(IN) :: th, th_lr, th_bt, u, u_lr, u_bt, v, v_lr, v_bt, &
                                 w, w_lr, w_bt, tht, tht_lr, tht_bt, pstr, pstr_lr, pstr_bt
      TYPE(c_ptr), INTENT(IN), OPTIONAL :: qv, qv_lr, qv_bt, qc, qc_lr, qc_bt, qr, qr_lr, qr_bt, &
                                           qi, qi_lr, qi_bt, qs, qs_lr, qs_bt, qg, qg_lr, qg_bt
      TYPE(c_ptr) :: x(12), xl(12), xb(12)
      INTEGER(c_int) :: role(12)
      INTEGER :: n
      IF (ipoles /= 0) THEN
         WRITE (*, '(A)') 'CE_advection: only the Cartesian case (ipoles = 0) exists'
         ERROR STOP 1
      END IF
      IF (ANY(ctx_dims /= (/ np, mp, lp, ih, ibcx, ibcy, ibcz /))) THEN
         WRITE (*, '(A)') 'CE_advection: domain or boundary flags differ from mpdata_init'
         ERROR STOP 1
      END IF
      n = 0
      CALL add(th, th_lr, th_bt, 0)
      CALL add(u, u_lr, u_bt, 1)
      CALL add(v, v_lr, v_bt, 1)
      CALL add(w, w_lr, w_bt, 1)
      CALL add(tht, tht_lr, tht_bt, 1)
      CALL add(pstr, pstr_lr, pstr_bt, 1)
      IF (PRESENT(qv)) CALL add(qv, qv_lr, qv_bt, 0)
      IF (PRESENT(qc)) CALL add(qc, qc_lr, qc_bt, 0)
      IF (PRESENT(qr)) CALL add(qr, qr_lr, qr_bt, 0)
      IF (PRESENT(qi)) CALL add(qi, qi_lr, qi_bt, 0)
      IF (PRESENT(qs)) CALL add(qs, qs_lr, qs_bt, 0)
      IF (PRESENT(qg)) CALL add(qg, qg_lr, qg_bt, 0)
      IF (lupdatemulti) THEN
         ! no effect here (one process)
      ENDIF
      ! the reference's savepoints (advection_CEtimestep_twostep_GPUBC.f90:
      ! 122-142, 355-375): the inputs at mysnapshot.run-in, the results at
      ! mysnapshot.run-out, in the open serializer (mpdata_serializer_open)
      IF (PRESENT(do_serialization_in)) THEN
         IF (do_serialization_in) CALL savepoint('mysnapshot.run-in', 'uadv', 'vadv', 'wadv')
      END IF
      CALL mpdata_check(c_ce_advection(ctx, INT(mpdata_alg_choice, c_int), INT(liner, c_int), &
                                       INT(n, c_int), role, x, xl, xb, uadv, vadv, wadv, rhr, &
                                       rhoadv, c_null_ptr), 'CE_advection')
      IF (PRESENT(do_serialization_out)) THEN
         IF (do_serialization_out) CALL savepoint('mysnapshot.run-out', 'u1', 'u2', 'u3')
      END IF
   CONTAINS
      SUBROUTINE savepoint(sp, n1, n2, n3)
         CHARACTER(*), INTENT(IN) :: sp, n1, n2, n3
         CALL mpdata_serialize(sp, n1, uadv, np + 1 + 2*ih, mp + 2*ih, lp + 2*ih)
         CALL mpdata_serialize(sp, n2, vadv, np + 2*ih, mp + 1 + 2*ih, lp + 2*ih)
         CALL mpdata_serialize(sp, n3, wadv, np + 2*ih, mp + 2*ih, lp + 1 + 2*ih)
         IF (sp == 'mysnapshot.run-in') &
            CALL mpdata_serialize(sp, 'rhoadv', rhoadv, np + 2*ih, mp + 2*ih, lp + 2*ih)
         CALL field3(sp, 'u', u, u_lr, u_bt)
         CALL field3(sp, 'v', v, v_lr, v_bt)
         CALL field3(sp, 'w', w, w_lr, w_bt)
         CALL field3(sp, 'th', th, th_lr, th_bt)
         CALL field3(sp, 'tht', tht, tht_lr, tht_bt)
         CALL field3(sp, 'pstr', pstr, pstr_lr, pstr_bt)
         IF (PRESENT(qv)) CALL field3(sp, 'qv', qv, qv_lr, qv_bt)
         IF (PRESENT(qc)) CALL field3(sp, 'qc', qc, qc_lr, qc_bt)
         IF (PRESENT(qr)) CALL field3(sp, 'qr', qr, qr_lr, qr_bt)
         IF (PRESENT(qi)) CALL field3(sp, 'qi', qi, qi_lr, qi_bt)
         IF (PRESENT(qs)) CALL field3(sp, 'qs', qs, qs_lr, qs_bt)
         IF (PRESENT(qg)) CALL field3(sp, 'qg', qg, qg_lr, qg_bt)
      END SUBROUTINE savepoint
      SUBROUTINE field3(sp, nm, f, f_lr, f_bt)
         CHARACTER(*), INTENT(IN) :: sp, nm
         TYPE(c_ptr), INTENT(IN) :: f, f_lr, f_bt
         CALL mpdata_serialize(sp, nm, f, np + 2*ih, mp + 2*ih, lp + 2*ih)
         CALL mpdata_serialize(sp, nm//'_lr', f_lr, mp, lp, 2)
         CALL mpdata_serialize(sp, nm//'_bt', f_bt, np, lp, 2)
      END SUBROUTINE field3
      SUBROUTINE add(f, f_lr, f_bt, r)
         TYPE(c_ptr), INTENT(IN) :: f, f_lr, f_bt
         INTEGER, INTENT(IN) :: r
         n = n + 1
         x(n) = f
         xl(n) = f_lr
         xb(n) = f_bt
         role(n) = INT(r, c_int)
      END SUBROUTINE add
   END SUBROUTINE CE_advection

   ! mpdata3d_rho_gauge_gpubc: advection/src_algorithms/mpdata3d-rho-gauge-gpubc.f90:23-34.
   ! The density step with the predictor driver; ox..w carry three time
   ! levels (0:2) each.  lupdatemulti has no effect on one process.
   SUBROUTINE mpdata3d_rho_gauge_gpubc(lupdatemulti, nsubsteps, nsubsteps_o, lsubstepping, &
                                       dtn, dtnold, ltimeadapt, lvertsplit, lcrdiag_nphlf, &
                                       crmaxlim, dtmax, ox, oy, oz, u, v, w, u1, u2, u3, &
                                       x_in, x_incomp, h, rhoadv, rhr, pexe, pexr1, pexr2, &
                                       bcx, bcy, ipoles0, ibcx0, ibcy0, ibcz0, np, mp, lp, ih, &
                                       dxi, dyi, dzi, dt, liner, do_serialization_in, &
                                       do_serialization_out)
      LOGICAL, INTENT(IN) :: lupdatemulti, lsubstepping, ltimeadapt, lvertsplit, lcrdiag_nphlf
      INTEGER, INTENT(INOUT) :: nsubsteps, nsubsteps_o, liner
      REAL(euwp), INTENT(INOUT) :: dtn, dtnold
      REAL(euwp), INTENT(IN) :: crmaxlim, dtmax, dxi, dyi, dzi, dt
      TYPE(c_ptr), INTENT(IN) :: ox, oy, oz, u, v, w, u1, u2, u3, x_in, x_incomp, h, rhoadv, rhr
      TYPE(c_ptr), INTENT(IN) :: pexe, pexr1, pexr2, bcx, bcy
      INTEGER, INTENT(IN) :: ipoles0, ibcx0, ibcy0, ibcz0, np, mp, lp, ih
      LOGICAL, INTENT(IN), OPTIONAL :: do_serialization_in, do_serialization_out
      INTEGER(c_int) :: ns, nso
      ns = INT(nsubsteps, c_int)
      nso = INT(nsubsteps_o, c_int)
      CALL mpdata_check(c_rho_gauge(ctx, flag(lupdatemulti), ns, nso, flag(lsubstepping), dtn, &
                                    dtnold, flag(ltimeadapt), flag(lvertsplit), &
                                    flag(lcrdiag_nphlf), crmaxlim, dtmax, ox, oy, oz, u, v, w, &
                                    u1, u2, u3, x_in, x_incomp, h, rhoadv, rhr, pexe, pexr1, &
                                    pexr2, bcx, bcy, INT(ipoles0, c_int), INT(ibcx0, c_int), &
                                    INT(ibcy0, c_int), INT(ibcz0, c_int), INT(np, c_int), &
                                    INT(mp, c_int), INT(lp, c_int), INT(ih, c_int), dxi, dyi, &
                                    dzi, dt, INT(liner, c_int), c_null_ptr), &
                        'mpdata3d_rho_gauge_gpubc')
      nsubsteps = ns
      nsubsteps_o = nso
      IF (PRESENT(do_serialization_in) .OR. PRESENT(do_serialization_out)) THEN
         ! no serialization here
      ENDIF
   END SUBROUTINE mpdata3d_rho_gauge_gpubc

   ! velprd_driver: advection/src_algorithms/velprd_driver.f90:14-19
   SUBROUTINE velprd_driver(nsubsteps, nsubsteps_o, lsubstepping, dtn, dtnold, ltimeadapt, &
                            lvertsplit, lcrdiag_nphlf, crmaxlim, dtmax, ox, oy, oz, u, v, w, &
                            uadv, vadv, wadv, h, ipoles0, ibcx0, ibcy0, ibcz0, np, mp, lp, ih, &
                            dxi, dyi, dzi, dt)
      INTEGER, INTENT(INOUT) :: nsubsteps, nsubsteps_o
      LOGICAL, INTENT(IN) :: lsubstepping, ltimeadapt, lvertsplit, lcrdiag_nphlf
      REAL(euwp), INTENT(INOUT) :: dtn, dtnold
      REAL(euwp), INTENT(IN) :: crmaxlim, dtmax, dxi, dyi, dzi, dt
      TYPE(c_ptr), INTENT(IN) :: ox, oy, oz, u, v, w, uadv, vadv, wadv, h
      INTEGER, INTENT(IN) :: ipoles0, ibcx0, ibcy0, ibcz0, np, mp, lp, ih
      INTEGER(c_int) :: ns, nso
      ns = INT(nsubsteps, c_int)
      nso = INT(nsubsteps_o, c_int)
      CALL mpdata_check(c_velprd_driver(ctx, ns, nso, flag(lsubstepping), dtn, dtnold, &
                                        flag(ltimeadapt), flag(lvertsplit), flag(lcrdiag_nphlf), &
                                        crmaxlim, dtmax, ox, oy, oz, u, v, w, uadv, vadv, wadv, &
                                        h, INT(ipoles0, c_int), INT(ibcx0, c_int), &
                                        INT(ibcy0, c_int), INT(ibcz0, c_int), INT(np, c_int), &
                                        INT(mp, c_int), INT(lp, c_int), INT(ih, c_int), dxi, &
                                        dyi, dzi, dt, c_null_ptr), 'velprd_driver')
      nsubsteps = ns
      nsubsteps_o = nso
   END SUBROUTINE velprd_driver

   ! ---- generalised step, time loop, predictor, diagnostics ---------------
   SUBROUTINE mpdata_step(variant, iters, nonos, x, u1, u2, u3, rhr, h, bcx, bcy)
      INTEGER, INTENT(IN) :: variant, iters
      LOGICAL, INTENT(IN) :: nonos
      TYPE(c_ptr), INTENT(IN) :: x, u1, u2, u3, rhr, h, bcx, bcy
      CALL mpdata_check(c_step(ctx, INT(variant, c_int), INT(iters, c_int), flag(nonos), x, u1, &
                               u2, u3, rhr, h, bcx, bcy, c_null_ptr), 'mpdata_step')
   END SUBROUTINE mpdata_step

   SUBROUTINE mpdata_timeloop(variant, iters, nonos, nt, x, u1, u2, u3, rhr, h, bcx, bcy)
      INTEGER, INTENT(IN) :: variant, iters, nt
      LOGICAL, INTENT(IN) :: nonos
      TYPE(c_ptr), INTENT(IN) :: x, u1, u2, u3, rhr, h, bcx, bcy
      CALL mpdata_check(c_timeloop(ctx, INT(variant, c_int), INT(iters, c_int), flag(nonos), &
                                   INT(nt, c_int), x, u1, u2, u3, rhr, h, bcx, bcy, c_null_ptr), &
                        'mpdata_timeloop')
   END SUBROUTINE mpdata_timeloop

   ! velprdA_to_C: advection/src_algorithms/module_velprd.f90:166-364
   SUBROUTINE velprdA_to_C(ox, oy, oz, u1, u2, u3)
      TYPE(c_ptr), INTENT(IN) :: ox, oy, oz, u1, u2, u3
      CALL mpdata_check(c_a_to_c(ctx, ox, oy, oz, u1, u2, u3, c_null_ptr), 'velprdA_to_C')
   END SUBROUTINE velprdA_to_C

   ! compute_courants: common/diagutils.f90:588-672; cr = (cr3d, crxy, crxz,
   ! cryz, crx, cry, crz)
   SUBROUTINE compute_courants(ox, oy, oz, h, ntimelevel, dxi, dyi, dzi, dt, cr)
      TYPE(c_ptr), INTENT(IN) :: ox, oy, oz, h
      INTEGER, INTENT(IN) :: ntimelevel
      REAL(euwp), INTENT(IN) :: dxi, dyi, dzi, dt
      REAL(c_double), INTENT(OUT) :: cr(7)
      CALL mpdata_check(c_courants(ctx, ox, oy, oz, h, INT(ntimelevel, c_int), dxi, dyi, dzi, &
                                   dt, cr, c_null_ptr), 'compute_courants')
   END SUBROUTINE compute_courants

   ! interior sums of x^2, x0^2, (x0-x)^2 and x, accumulated in FP64
   SUBROUTINE norm_sums(x, x0, s)
      TYPE(c_ptr), INTENT(IN) :: x, x0
      REAL(c_double), INTENT(OUT) :: s(4)
      CALL mpdata_check(c_norm_sums(ctx, x, x0, s, c_null_ptr), 'norm_sums')
   END SUBROUTINE norm_sums

END MODULE mpdata_hip
