! ---------------------------------------------------------------------------
! TEST INFRASTRUCTURE ONLY.
! bind(C) shim over the reference's own Fortran routines.  oracle/Makefile
! compiles the reference sources where they lie (advectionHPCtester tree,
! CPU path: -DNOCUDACODE -DTIMERSCPU, amdflang -O2, no fast-math) together
! with this file into oracle/_ref/libmpdata_ref.so.  Tests load it through
! ctypes to pin the C oracle bit for bit; nothing in the product loads it.
!
! Arrays are passed as C pointers to Fortran-ordered storage with lower
! bound 1-ih, exactly as the reference declares them.
! ---------------------------------------------------------------------------
MODULE ref_shim
   USE iso_c_binding
   USE precisions, ONLY: euwp
   IMPLICIT NONE
   INTEGER, SAVE :: s_np = 0, s_mp = 0, s_lp = 0, s_ih = 0
   INTEGER, SAVE :: s_ibcx = 0, s_ibcy = 0, s_ibcz = 0
   LOGICAL, SAVE :: s_alloc = .FALSE.
CONTAINS

   ! Domain set-up as advection_dwarf_test_modular.f90:106-141 does it for one
   ! process: parameters, geomset (edge flags), subdomains, scratch fields.
   SUBROUTINE ref_setup(n, m, l, ih, ibcx, ibcy, ibcz, dx00, nt, dt) BIND(C, name='ref_setup')
      USE parameters, ONLY: set_eulagcommon_mpi_parameters, set_eulagcommon_lib_domainsize
      USE parameters, ONLY: np, mp, lp
      USE mpi_parallel, ONLY: geomset, init_subdomains, ttini
      USE scratch_datafields, ONLY: allocate_scratch_mpdata_datafields, &
                                    allocate_scratch_advdrv_datafields, &
                                    allocate_scratch_advtst_datafields, &
                                    allocate_scratch_aux_datafields, &
                                    deallocate_scratch_mpdata_datafields, &
                                    deallocate_scratch_advdrv_datafields, &
                                    deallocate_scratch_advtst_datafields, &
                                    deallocate_scratch_aux_datafields
      INTEGER(c_int), VALUE :: n, m, l, ih, ibcx, ibcy, ibcz, nt
      REAL(c_double), VALUE :: dx00, dt
      IF (s_alloc) THEN
         CALL deallocate_scratch_mpdata_datafields
         CALL deallocate_scratch_advdrv_datafields
         CALL deallocate_scratch_advtst_datafields
         CALL deallocate_scratch_aux_datafields
      END IF
      CALL set_eulagcommon_mpi_parameters(n_in=n, nprocx_in=1, nprocy_in=1, nprocz_in=1, &
                                          ibcx_in=ibcx, ibcy_in=ibcy, ibcz_in=ibcz, &
                                          isphere_in=0, ipoles_in=0)
      CALL set_eulagcommon_lib_domainsize(n_in=n, m_in=m, l_in=l, ih_in=ih, &
                                          dx00_in=REAL(dx00, euwp), dy00_in=REAL(dx00, euwp), &
                                          dz00_in=REAL(dx00, euwp), nt_in=nt, dt_in=REAL(dt, euwp))
      CALL geomset(.FALSE.)
      CALL init_subdomains
      CALL allocate_scratch_mpdata_datafields
      CALL allocate_scratch_advdrv_datafields
      CALL allocate_scratch_advtst_datafields
      CALL allocate_scratch_aux_datafields
      CALL ttini
      s_np = np; s_mp = mp; s_lp = lp; s_ih = ih
      s_ibcx = ibcx; s_ibcy = ibcy; s_ibcz = ibcz
      s_alloc = .TRUE.
   END SUBROUTINE ref_setup

   INTEGER(c_int) FUNCTION ref_precision() BIND(C, name='ref_precision')
      ref_precision = euwp
   END FUNCTION ref_precision

   SUBROUTINE ref_upwind(pu1, pu2, pu3, px, pxant, prhr, ph, pbcx, pbcy) BIND(C, name='ref_upwind')
      USE module_upwind3d_gpubc, ONLY: upwind3d_gpubc
      USE scratch_datafields, ONLY: scratch_zeros
      USE parameters, ONLY: dt
      TYPE(c_ptr), VALUE :: pu1, pu2, pu3, px, pxant, prhr, ph, pbcx, pbcy
      REAL(euwp), POINTER :: u1(:, :, :), u2(:, :, :), u3(:, :, :), x(:, :, :), xant(:, :, :)
      REAL(euwp), POINTER :: rhr(:, :, :), h(:, :, :), bcx(:, :, :), bcy(:, :, :)
      CALL bind_fields(pu1, pu2, pu3, u1, u2, u3)
      CALL bind_scalar(px, x); CALL bind_scalar(pxant, xant)
      CALL bind_scalar(prhr, rhr); CALL bind_scalar(ph, h)
      CALL c_f_pointer(pbcx, bcx, [s_mp, s_lp, 2])
      CALL c_f_pointer(pbcy, bcy, [s_np, s_lp, 2])
      CALL upwind3d_gpubc(u1, u2, u3, x, xant, scratch_zeros, scratch_zeros, rhr, h, bcx, bcy, &
                          0, s_ibcx, s_ibcy, s_ibcz, s_np, s_mp, s_lp, s_ih, dt, .FALSE., .FALSE.)
   END SUBROUTINE ref_upwind

   ! variant 1 = gauge, 2 = standard.  After the call the module scratch
   ! (v, cp, cn and for the standard variant f, fns) is copied out.
   SUBROUTINE ref_antidiff(variant, pu1, pu2, pu3, pxant, px, prhr, ph, &
                           pv1, pv2, pv3, pcp, pcn, pf1, pf2, pf3, pf1ns, pf2ns, pf3ns) &
      BIND(C, name='ref_antidiff')
      USE module_antidiff3d_gauge_gpubc, ONLY: antidiff3d_gauge_gpubc
      USE module_antidiff3d_standard_gpubc, ONLY: antidiff3d_standard_gpubc
      USE scratch_datafields, ONLY: v1, v2, v3, cp, cn, f1, f2, f3, f1ns, f2ns, f3ns
      INTEGER(c_int), VALUE :: variant
      TYPE(c_ptr), VALUE :: pu1, pu2, pu3, pxant, px, prhr, ph
      TYPE(c_ptr), VALUE :: pv1, pv2, pv3, pcp, pcn, pf1, pf2, pf3, pf1ns, pf2ns, pf3ns
      REAL(euwp), POINTER :: u1(:, :, :), u2(:, :, :), u3(:, :, :), x(:, :, :), xant(:, :, :)
      REAL(euwp), POINTER :: rhr(:, :, :), h(:, :, :), o1(:, :, :), o2(:, :, :), o3(:, :, :)
      REAL(euwp), POINTER :: oc(:, :, :)
      CALL bind_fields(pu1, pu2, pu3, u1, u2, u3)
      CALL bind_scalar(px, x); CALL bind_scalar(pxant, xant)
      CALL bind_scalar(prhr, rhr); CALL bind_scalar(ph, h)
      IF (variant == 1) THEN
         CALL antidiff3d_gauge_gpubc(.FALSE., u1, u2, u3, xant, x, rhr, h, 0, 0, &
                                     s_ibcx, s_ibcy, s_ibcz, s_np, s_mp, s_lp, s_ih, .FALSE.)
      ELSE
         CALL antidiff3d_standard_gpubc(.FALSE., u1, u2, u3, xant, x, rhr, h, 0, 0, &
                                        s_ibcx, s_ibcy, s_ibcz, s_np, s_mp, s_lp, s_ih, .FALSE.)
      END IF
      IF (c_associated(pv1)) THEN
         CALL bind_fields(pv1, pv2, pv3, o1, o2, o3)
         o1 = v1; o2 = v2; o3 = v3
      END IF
      IF (c_associated(pcp)) THEN
         CALL bind_scalar(pcp, oc); oc = cp
         CALL bind_scalar(pcn, oc); oc = cn
      END IF
      IF (variant == 2 .AND. c_associated(pf1)) THEN
         CALL bind_fields(pf1, pf2, pf3, o1, o2, o3)
         o1 = f1; o2 = f2; o3 = f3
         CALL bind_fields(pf1ns, pf2ns, pf3ns, o1, o2, o3)
         o1 = f1ns; o2 = f2ns; o3 = f3ns
      END IF
   END SUBROUTINE ref_antidiff

   ! One time step of the dwarf driver: advec.f90:116-244.
   SUBROUTINE ref_advec_dwarf(legacyoptmode, t_adv, px, pbcx, pbcy, pu1, pu2, pu3, prhr, ph) &
      BIND(C, name='ref_advec_dwarf')
      USE advec_driver, ONLY: advec_dwarf
      INTEGER(c_int), VALUE :: legacyoptmode, t_adv
      TYPE(c_ptr), VALUE :: px, pbcx, pbcy, pu1, pu2, pu3, prhr, ph
      REAL(euwp), POINTER :: u1(:, :, :), u2(:, :, :), u3(:, :, :), x(:, :, :)
      REAL(euwp), POINTER :: rhr(:, :, :), h(:, :, :), bcx(:, :, :), bcy(:, :, :)
      REAL(euwp), POINTER :: w3(:, :, :)
      CALL bind_fields(pu1, pu2, pu3, u1, u2, u3)
      CALL bind_scalar(px, x)
      CALL bind_scalar(prhr, rhr); CALL bind_scalar(ph, h)
      CALL c_f_pointer(pbcx, bcx, [s_mp, s_lp, 2])
      CALL c_f_pointer(pbcy, bcy, [s_np, s_lp, 2])
      ! advec.f90:136 declares wadv with an x-face shape; the storage is the
      ! z-face array (sequence association), as the test app passes it.
      CALL c_f_pointer(pu3, w3, [s_np + 1 + 2*s_ih, s_mp + 2*s_ih, s_lp + 1 + 2*s_ih])
      CALL advec_dwarf(legacyoptmode, t_adv, .FALSE., .FALSE., s_np, s_mp, s_lp, s_ih, &
                       0, s_ibcx, s_ibcy, s_ibcz, x, bcx, bcy, u1, u2, w3, rhr, h)
   END SUBROUTINE ref_advec_dwarf

   ! Initial condition of the dwarf test: advection_test_setup.f90:174-255
   ! (sphere of radius 15, solid-body rotation, A->C averaging ver2).
   SUBROUTINE ref_set_initial15(pu1, pu2, pu3, px, prho) BIND(C, name='ref_set_initial15')
      USE testing, ONLY: set_initial15
      TYPE(c_ptr), VALUE :: pu1, pu2, pu3, px, prho
      REAL(euwp), POINTER :: u1(:, :, :), u2(:, :, :), u3(:, :, :), x(:, :, :), rho(:, :, :)
      CALL bind_fields(pu1, pu2, pu3, u1, u2, u3)
      CALL bind_scalar(px, x); CALL bind_scalar(prho, rho)
      CALL set_initial15(u1, u2, u3, x, rho, s_np, s_mp, s_lp, s_ih)
   END SUBROUTINE ref_set_initial15

   ! Norms printed by the dwarf test: advection_test_setup.f90:472-492,595-617
   SUBROUTINE ref_norms(px, px0, icnt) BIND(C, name='ref_norms')
      USE testing, ONLY: test_solution_energy, test_solution_error
      TYPE(c_ptr), VALUE :: px, px0
      INTEGER(c_int), VALUE :: icnt
      REAL(euwp), POINTER :: x(:, :, :), x0(:, :, :)
      CALL bind_scalar(px, x); CALL bind_scalar(px0, x0)
      CALL test_solution_energy(x, x0, icnt, 'a', s_np, s_mp, s_lp, s_ih)
      CALL test_solution_error(x, x0, icnt, 'c', s_np, s_mp, s_lp, s_ih)
      FLUSH (6)
   END SUBROUTINE ref_norms

   SUBROUTINE ref_velprdA_to_C(ver, pox, poy, poz, pu1, pu2, pu3) BIND(C, name='ref_velprdA_to_C')
      USE module_velprd, ONLY: velprdA_to_C, velprdA_to_C_ver2
      INTEGER(c_int), VALUE :: ver
      TYPE(c_ptr), VALUE :: pox, poy, poz, pu1, pu2, pu3
      REAL(euwp), POINTER :: u1(:, :, :), u2(:, :, :), u3(:, :, :)
      REAL(euwp), POINTER :: ox(:, :, :), oy(:, :, :), oz(:, :, :)
      CALL bind_fields(pu1, pu2, pu3, u1, u2, u3)
      CALL bind_scalar(pox, ox); CALL bind_scalar(poy, oy); CALL bind_scalar(poz, oz)
      IF (ver == 2) THEN
         CALL velprdA_to_C_ver2(ox, oy, oz, u1, u2, u3, 0, s_ibcx, s_ibcy, s_ibcz, &
                                s_np, s_mp, s_lp, s_ih)
      ELSE
         CALL velprdA_to_C(ox, oy, oz, u1, u2, u3, 0, s_ibcx, s_ibcy, s_ibcz, &
                           s_np, s_mp, s_lp, s_ih)
      END IF
   END SUBROUTINE ref_velprdA_to_C

   ! ox/oy/oz: three time levels (0:2) of scalar fields, contiguous.
   SUBROUTINE ref_velprd_traj0(pox, poy, poz, pu1, pu2, pu3, ph, dxi, dyi, dzi, dt, dt_ratio) &
      BIND(C, name='ref_velprd_traj0')
      USE module_velprd, ONLY: velprd_traj0
      TYPE(c_ptr), VALUE :: pox, poy, poz, pu1, pu2, pu3, ph
      REAL(c_double), VALUE :: dxi, dyi, dzi, dt, dt_ratio
      REAL(euwp), POINTER :: u1(:, :, :), u2(:, :, :), u3(:, :, :), h(:, :, :)
      REAL(euwp), POINTER :: ox(:, :, :, :), oy(:, :, :, :), oz(:, :, :, :)
      INTEGER :: e(4)
      CALL bind_fields(pu1, pu2, pu3, u1, u2, u3)
      CALL bind_scalar(ph, h)
      e = [s_np + 2*s_ih, s_mp + 2*s_ih, s_lp + 2*s_ih, 3]
      CALL c_f_pointer(pox, ox, e); CALL c_f_pointer(poy, oy, e); CALL c_f_pointer(poz, oz, e)
      CALL velprd_traj0(ox, oy, oz, u1, u2, u3, h, REAL(dxi, euwp), REAL(dyi, euwp), &
                        REAL(dzi, euwp), REAL(dt, euwp), REAL(dt_ratio, euwp), &
                        0, s_ibcx, s_ibcy, s_ibcz, s_np, s_mp, s_lp, s_ih)
   END SUBROUTINE ref_velprd_traj0

   SUBROUTINE ref_compute_courants(pox, poy, poz, ph, ntimelevel, dxi, dyi, dzi, dt, out) &
      BIND(C, name='ref_compute_courants')
      USE eulag_diagutils, ONLY: compute_courants
      TYPE(c_ptr), VALUE :: pox, poy, poz, ph
      INTEGER(c_int), VALUE :: ntimelevel
      REAL(c_double), VALUE :: dxi, dyi, dzi, dt
      REAL(c_double), INTENT(OUT) :: out(7)
      REAL(euwp), POINTER :: h(:, :, :)
      REAL(euwp), POINTER :: ox(:, :, :, :), oy(:, :, :, :), oz(:, :, :, :)
      REAL(euwp) :: c(7)
      INTEGER :: e(4)
      CALL bind_scalar(ph, h)
      e = [s_np + 2*s_ih, s_mp + 2*s_ih, s_lp + 2*s_ih, 3]
      CALL c_f_pointer(pox, ox, e); CALL c_f_pointer(poy, oy, e); CALL c_f_pointer(poz, oz, e)
      CALL compute_courants(c(1), c(2), c(3), c(4), c(5), c(6), c(7), ox, oy, oz, h, ntimelevel, &
                            REAL(dxi, euwp), REAL(dyi, euwp), REAL(dzi, euwp), REAL(dt, euwp), &
                            s_np, s_mp, s_lp, s_ih)
      out = REAL(c, c_double)
   END SUBROUTINE ref_compute_courants

   ! mpdata3d-rho-gauge-gpubc.f90:23-622 (density step with the predictor
   ! driver, velprd_driver.f90:14-295).  flags = lupdatemulti, lsubstepping,
   ! ltimeadapt, lvertsplit, lcrdiag_nphlf; isub = nsubsteps, nsubsteps_o
   ! (in/out); rdt = dtn, dtnold (in/out), crmaxlim, dtmax; sp = dxi, dyi,
   ! dzi, dt; liner in/out.  ox..w hold three time levels (0:2).  When
   ! only_driver /= 0 the predictor driver alone runs.
   SUBROUTINE ref_rho_gauge(only_driver, flags, isub, rdt, pox, poy, poz, pu, pv, pw, &
                            pu1, pu2, pu3, px, pxinc, ph, prhoadv, prhr, ppexe, ppexr1, &
                            ppexr2, pbcx, pbcy, sp, liner) BIND(C, name='ref_rho_gauge')
      USE module_mpdata3d_rho_gauge_gpubc, ONLY: mpdata3d_rho_gauge_gpubc
      USE velprd_driver_module, ONLY: velprd_driver
      INTEGER(c_int), VALUE :: only_driver
      INTEGER(c_int), INTENT(IN) :: flags(5)
      INTEGER(c_int), INTENT(INOUT) :: isub(2), liner
      REAL(c_double), INTENT(INOUT) :: rdt(4)
      REAL(c_double), INTENT(IN) :: sp(4)
      TYPE(c_ptr), VALUE :: pox, poy, poz, pu, pv, pw, pu1, pu2, pu3, px, pxinc, ph
      TYPE(c_ptr), VALUE :: prhoadv, prhr, ppexe, ppexr1, ppexr2, pbcx, pbcy
      REAL(euwp), POINTER :: u1(:, :, :), u2(:, :, :), u3(:, :, :), x(:, :, :), xinc(:, :, :)
      REAL(euwp), POINTER :: h(:, :, :), rhoadv(:, :, :), rhr(:, :, :), pexe(:, :, :)
      REAL(euwp), POINTER :: pexr1(:, :, :), pexr2(:, :, :), bcx(:, :, :), bcy(:, :, :)
      REAL(euwp), POINTER :: ox(:, :, :, :), oy(:, :, :, :), oz(:, :, :, :)
      REAL(euwp), POINTER :: u(:, :, :, :), v(:, :, :, :), w(:, :, :, :)
      LOGICAL :: lf(5)
      INTEGER :: e(4), nsub, nsubo, lin
      REAL(euwp) :: dtn, dtnold
      lf = flags /= 0
      CALL bind_fields(pu1, pu2, pu3, u1, u2, u3)
      CALL bind_scalar(ph, h)
      e = [s_np + 2*s_ih, s_mp + 2*s_ih, s_lp + 2*s_ih, 3]
      CALL c_f_pointer(pox, ox, e); CALL c_f_pointer(poy, oy, e); CALL c_f_pointer(poz, oz, e)
      CALL c_f_pointer(pu, u, e); CALL c_f_pointer(pv, v, e); CALL c_f_pointer(pw, w, e)
      nsub = isub(1); nsubo = isub(2); lin = liner
      dtn = REAL(rdt(1), euwp); dtnold = REAL(rdt(2), euwp)
      IF (only_driver /= 0) THEN
         CALL velprd_driver(nsub, nsubo, lf(2), dtn, dtnold, lf(3), lf(4), lf(5), &
                            REAL(rdt(3), euwp), REAL(rdt(4), euwp), ox, oy, oz, u, v, w, &
                            u1, u2, u3, h, 0, s_ibcx, s_ibcy, s_ibcz, s_np, s_mp, s_lp, s_ih, &
                            REAL(sp(1), euwp), REAL(sp(2), euwp), REAL(sp(3), euwp), &
                            REAL(sp(4), euwp))
      ELSE
         CALL bind_scalar(px, x); CALL bind_scalar(pxinc, xinc)
         CALL bind_scalar(prhoadv, rhoadv); CALL bind_scalar(prhr, rhr)
         CALL bind_scalar(ppexe, pexe); CALL bind_scalar(ppexr1, pexr1)
         CALL bind_scalar(ppexr2, pexr2)
         CALL c_f_pointer(pbcx, bcx, [s_mp, s_lp, 2])
         CALL c_f_pointer(pbcy, bcy, [s_np, s_lp, 2])
         CALL mpdata3d_rho_gauge_gpubc(lf(1), nsub, nsubo, lf(2), dtn, dtnold, lf(3), lf(4), &
                                       lf(5), REAL(rdt(3), euwp), REAL(rdt(4), euwp), &
                                       ox, oy, oz, u, v, w, u1, u2, u3, x, xinc, h, rhoadv, rhr, &
                                       pexe, pexr1, pexr2, bcx, bcy, 0, s_ibcx, s_ibcy, s_ibcz, &
                                       s_np, s_mp, s_lp, s_ih, REAL(sp(1), euwp), &
                                       REAL(sp(2), euwp), REAL(sp(3), euwp), REAL(sp(4), euwp), &
                                       lin, .FALSE., .FALSE.)
      END IF
      isub(1) = nsub; isub(2) = nsubo; liner = lin
      rdt(1) = REAL(dtn, c_double); rdt(2) = REAL(dtnold, c_double)
      FLUSH (6)
   END SUBROUTINE ref_rho_gauge

   ! -- helpers: bind C storage to Fortran pointers with lower bound 1-ih --
   SUBROUTINE bind_scalar(p, a)
      TYPE(c_ptr), VALUE :: p
      REAL(euwp), POINTER :: a(:, :, :)
      REAL(euwp), POINTER :: t(:, :, :)
      CALL c_f_pointer(p, t, [s_np + 2*s_ih, s_mp + 2*s_ih, s_lp + 2*s_ih])
      a(1 - s_ih:, 1 - s_ih:, 1 - s_ih:) => t
   END SUBROUTINE bind_scalar

   SUBROUTINE bind_fields(p1, p2, p3, a1, a2, a3)
      TYPE(c_ptr), VALUE :: p1, p2, p3
      REAL(euwp), POINTER :: a1(:, :, :), a2(:, :, :), a3(:, :, :)
      REAL(euwp), POINTER :: t(:, :, :)
      CALL c_f_pointer(p1, t, [s_np + 1 + 2*s_ih, s_mp + 2*s_ih, s_lp + 2*s_ih])
      a1(1 - s_ih:, 1 - s_ih:, 1 - s_ih:) => t
      CALL c_f_pointer(p2, t, [s_np + 2*s_ih, s_mp + 1 + 2*s_ih, s_lp + 2*s_ih])
      a2(1 - s_ih:, 1 - s_ih:, 1 - s_ih:) => t
      CALL c_f_pointer(p3, t, [s_np + 2*s_ih, s_mp + 2*s_ih, s_lp + 1 + 2*s_ih])
      a3(1 - s_ih:, 1 - s_ih:, 1 - s_ih:) => t
   END SUBROUTINE bind_fields
END MODULE ref_shim
