!------------------------------------------------------------------------------
! pf_gpu: ParaFEM's EBE-PCG call surface on an MI355X, through ISO_C_BINDING
! over libparafem_amd (include/parafem_amd.h).
!
! The routine names and argument meaning follow the reference's library so a
! p121/p123-style driver keeps its loop:
!   GATHER(p_pp,pmul_pp), SCATTER(u_pp,utemp_pp), SCATTER_NOADD(utemp_pp,u_pp)
!     src/modules/mpi/gather_scatter.f90:547-993
!   DOT_PRODUCT_P(a,b), SUM_P(a), checon_par(loads,tol,converged,oldlds)
!     src/modules/mpi/maths.f90:168-216, 271-315, 999-1069
! Each of these is generic over two forms: TYPE(pf_dvec) handles of device
! vectors (the fast form: the vectors stay in HBM), and REAL(iwp) host arrays
! with the reference's own shapes -- p_pp(neq_pp), pmul_pp(ntot,nels_pp),
! u_pp(neq_pp), utemp_pp(ntot,nels_pp) -- so that p121.f90:89-104 compiles
! unchanged against this module (the host forms stage the arrays through
! device buffers and return the same bits as the device forms).  The one
! device context is module state, as the reference keeps its halo tables as
! module state.
!------------------------------------------------------------------------------
MODULE pf_gpu
  USE, INTRINSIC :: ISO_C_BINDING
  IMPLICIT NONE
  PRIVATE

  INTEGER, PARAMETER, PUBLIC :: iwp = SELECTED_REAL_KIND(15, 300)
  INTEGER(c_int), PARAMETER, PUBLIC :: PF_OK = 0, PF_NOT_CONVERGED = 5
  INTEGER(c_int), PARAMETER, PUBLIC :: PF_STORE_FP64 = 0, PF_STORE_FP32 = 1
  INTEGER(c_int), PARAMETER, PUBLIC :: PF_STOP_CHECON = 0, PF_STOP_RELRES = 1
  INTEGER(c_int), PARAMETER, PUBLIC :: PF_PREC_FP64 = 0, PF_PREC_MIXED = 1
  INTEGER(c_int), PARAMETER, PUBLIC :: PF_DOT_TREE = 0, PF_DOT_SEQ = 1

  TYPE, BIND(C), PUBLIC :: pf_pcg_opts
    INTEGER(c_int) :: limit
    REAL(c_double) :: tol
    INTEGER(c_int) :: stop_mode
    INTEGER(c_int) :: precision
    INTEGER(c_int) :: check_every
    INTEGER(c_int) :: refine_every
    REAL(c_double) :: refine_delta
  END TYPE pf_pcg_opts

  TYPE, BIND(C), PUBLIC :: pf_pcg_result
    INTEGER(c_int) :: iters
    INTEGER(c_int) :: converged
    REAL(c_double) :: crit
    REAL(c_double) :: rr0
    REAL(c_double) :: rr
    INTEGER(c_int) :: refinements
    INTEGER(c_int) :: status
  END TYPE pf_pcg_result

  ! A device vector of n REAL(iwp).
  TYPE, PUBLIC :: pf_dvec
    TYPE(c_ptr) :: ptr = c_null_ptr
    INTEGER(c_int64_t) :: n = 0
  END TYPE pf_dvec

  TYPE(c_ptr), SAVE :: ctx = c_null_ptr
  ! Sizes of the steering on the device (pf_steer): the host-array forms
  ! check their arguments against them before any kernel runs.
  INTEGER(c_int64_t), SAVE :: m_neq = 0, m_nslot = 0
  ! Device staging buffers of the host-array forms (grown on demand).
  TYPE(pf_dvec), SAVE :: st_a, st_b

  PUBLIC :: pf_init, pf_finalize, pf_alloc, pf_release, pf_put, pf_get
  PUBLIC :: pf_mesh, pf_steer, pf_assemble_elastic, pf_assemble_laplace
  PUBLIC :: pf_storkm, pf_storkm_get, pf_diag, pf_fixed, pf_pcg, pf_check, pcg_km
  PUBLIC :: gather, scatter, scatter_noadd, elem_matvec
  PUBLIC :: dot_product_p, sum_p, checon_par, pf_dot_order, pf_matvec

  INTERFACE gather
    MODULE PROCEDURE gather_dev, gather_host
  END INTERFACE
  INTERFACE scatter
    MODULE PROCEDURE scatter_dev, scatter_host
  END INTERFACE
  INTERFACE scatter_noadd
    MODULE PROCEDURE scatter_noadd_dev, scatter_noadd_host
  END INTERFACE
  INTERFACE dot_product_p
    MODULE PROCEDURE dot_product_p_dev, dot_product_p_host
  END INTERFACE
  INTERFACE sum_p
    MODULE PROCEDURE sum_p_dev, sum_p_host
  END INTERFACE
  INTERFACE checon_par
    MODULE PROCEDURE checon_par_dev, checon_par_host
  END INTERFACE
  PUBLIC :: vec_fill, vec_waxpy, vec_mul, vec_recip
  PUBLIC :: pf_stress, calc_nodes_pp, scatter_nodes, dismsh_ensi_p
  PUBLIC :: pf_assemble_transient, pf_apply_raw, pf_plastic_update, pf_fixed_rows, pf_cg_warm

  INTERFACE
    INTEGER(c_int) FUNCTION c_create(device, c) BIND(C, name="pf_create")
      IMPORT :: c_int, c_ptr
      INTEGER(c_int), VALUE :: device
      TYPE(c_ptr) :: c
    END FUNCTION
    INTEGER(c_int) FUNCTION c_destroy(c) BIND(C, name="pf_destroy")
      IMPORT :: c_int, c_ptr
      TYPE(c_ptr), VALUE :: c
    END FUNCTION
    FUNCTION c_status_string(s) BIND(C, name="pf_status_string")
      IMPORT :: c_int, c_ptr
      INTEGER(c_int), VALUE :: s
      TYPE(c_ptr) :: c_status_string
    END FUNCTION
    INTEGER(c_int) FUNCTION c_dev_alloc(c, bytes, p) BIND(C, name="pf_dev_alloc")
      IMPORT :: c_int, c_ptr, c_size_t
      TYPE(c_ptr), VALUE :: c
      INTEGER(c_size_t), VALUE :: bytes
      TYPE(c_ptr) :: p
    END FUNCTION
    INTEGER(c_int) FUNCTION c_dev_free(c, p) BIND(C, name="pf_dev_free")
      IMPORT :: c_int, c_ptr
      TYPE(c_ptr), VALUE :: c, p
    END FUNCTION
    INTEGER(c_int) FUNCTION c_h2d(c, dst, src, bytes) BIND(C, name="pf_memcpy_h2d")
      IMPORT :: c_int, c_ptr, c_size_t
      TYPE(c_ptr), VALUE :: c, dst, src
      INTEGER(c_size_t), VALUE :: bytes
    END FUNCTION
    INTEGER(c_int) FUNCTION c_d2h(c, dst, src, bytes) BIND(C, name="pf_memcpy_d2h")
      IMPORT :: c_int, c_ptr, c_size_t
      TYPE(c_ptr), VALUE :: c, dst, src
      INTEGER(c_size_t), VALUE :: bytes
    END FUNCTION
    INTEGER(c_int) FUNCTION c_mesh_set(c, nod, nels, g_coord) BIND(C, name="pf_mesh_set")
      IMPORT :: c_int, c_ptr, c_double
      TYPE(c_ptr), VALUE :: c
      INTEGER(c_int), VALUE :: nod, nels
      REAL(c_double), INTENT(IN) :: g_coord(*)
    END FUNCTION
    INTEGER(c_int) FUNCTION c_steer_set(c, ntot, nels, neq, g_g) BIND(C, name="pf_steer_set")
      IMPORT :: c_int, c_ptr
      TYPE(c_ptr), VALUE :: c
      INTEGER(c_int), VALUE :: ntot, nels, neq
      INTEGER(c_int), INTENT(IN) :: g_g(*)
    END FUNCTION
    INTEGER(c_int) FUNCTION c_assemble_elastic(c, e, v, e_elem, store) BIND(C, name="pf_assemble_elastic")
      IMPORT :: c_int, c_ptr, c_double
      TYPE(c_ptr), VALUE :: c
      REAL(c_double), VALUE :: e, v
      TYPE(c_ptr), VALUE :: e_elem
      INTEGER(c_int), VALUE :: store
    END FUNCTION
    INTEGER(c_int) FUNCTION c_assemble_laplace(c, kx, ky, kz, k_elem, store) BIND(C, name="pf_assemble_laplace")
      IMPORT :: c_int, c_ptr, c_double
      TYPE(c_ptr), VALUE :: c
      REAL(c_double), VALUE :: kx, ky, kz
      TYPE(c_ptr), VALUE :: k_elem
      INTEGER(c_int), VALUE :: store
    END FUNCTION
    INTEGER(c_int) FUNCTION c_storkm_set(c, storkm, store) BIND(C, name="pf_storkm_set")
      IMPORT :: c_int, c_ptr, c_double
      TYPE(c_ptr), VALUE :: c
      REAL(c_double), INTENT(IN) :: storkm(*)
      INTEGER(c_int), VALUE :: store
    END FUNCTION
    INTEGER(c_int) FUNCTION c_storkm_get(c, storkm) BIND(C, name="pf_storkm_get")
      IMPORT :: c_int, c_ptr, c_double
      TYPE(c_ptr), VALUE :: c
      REAL(c_double), INTENT(OUT) :: storkm(*)
    END FUNCTION
    INTEGER(c_int) FUNCTION c_storkm_shared(c, km) BIND(C, name="pf_storkm_shared")
      IMPORT :: c_int, c_ptr, c_double
      TYPE(c_ptr), VALUE :: c
      REAL(c_double), INTENT(IN) :: km(*)
    END FUNCTION
    INTEGER(c_int) FUNCTION c_precon_set(c, dp) BIND(C, name="pf_precon_set")
      IMPORT :: c_int, c_ptr, c_double
      TYPE(c_ptr), VALUE :: c
      REAL(c_double), INTENT(IN) :: dp(*)
    END FUNCTION
    INTEGER(c_int) FUNCTION c_pcg_r(c, r, on_device) BIND(C, name="pf_pcg_r")
      IMPORT :: c_int, c_ptr, c_double
      TYPE(c_ptr), VALUE :: c
      REAL(c_double), INTENT(OUT) :: r(*)
      INTEGER(c_int), VALUE :: on_device
    END FUNCTION
    INTEGER(c_int) FUNCTION c_diag_get(c, diag) BIND(C, name="pf_diag_get")
      IMPORT :: c_int, c_ptr, c_double
      TYPE(c_ptr), VALUE :: c
      REAL(c_double), INTENT(OUT) :: diag(*)
    END FUNCTION
    INTEGER(c_int) FUNCTION c_fixed_set(c, nfix, no_f, penalty, store) BIND(C, name="pf_fixed_set")
      IMPORT :: c_int, c_ptr, c_double
      TYPE(c_ptr), VALUE :: c
      INTEGER(c_int), VALUE :: nfix
      INTEGER(c_int), INTENT(IN) :: no_f(*)
      REAL(c_double), VALUE :: penalty
      REAL(c_double), INTENT(OUT) :: store(*)
    END FUNCTION
    INTEGER(c_int) FUNCTION c_gather(c, p, pmul) BIND(C, name="pf_gather")
      IMPORT :: c_int, c_ptr
      TYPE(c_ptr), VALUE :: c, p, pmul
    END FUNCTION
    INTEGER(c_int) FUNCTION c_elem_matvec(c, pmul, utemp) BIND(C, name="pf_elem_matvec")
      IMPORT :: c_int, c_ptr
      TYPE(c_ptr), VALUE :: c, pmul, utemp
    END FUNCTION
    INTEGER(c_int) FUNCTION c_scatter(c, u, utemp) BIND(C, name="pf_scatter")
      IMPORT :: c_int, c_ptr
      TYPE(c_ptr), VALUE :: c, u, utemp
    END FUNCTION
    INTEGER(c_int) FUNCTION c_scatter_noadd(c, utemp, u) BIND(C, name="pf_scatter_noadd")
      IMPORT :: c_int, c_ptr
      TYPE(c_ptr), VALUE :: c, utemp, u
    END FUNCTION
    INTEGER(c_int) FUNCTION c_dot(c, a, b, n, out) BIND(C, name="pf_dot_product_p")
      IMPORT :: c_int, c_ptr, c_int64_t, c_double
      TYPE(c_ptr), VALUE :: c, a, b
      INTEGER(c_int64_t), VALUE :: n
      REAL(c_double) :: out
    END FUNCTION
    INTEGER(c_int) FUNCTION c_sum(c, a, n, out) BIND(C, name="pf_sum_p")
      IMPORT :: c_int, c_ptr, c_int64_t, c_double
      TYPE(c_ptr), VALUE :: c, a
      INTEGER(c_int64_t), VALUE :: n
      REAL(c_double) :: out
    END FUNCTION
    INTEGER(c_int) FUNCTION c_checon(c, xnew, tol, conv, x, n) BIND(C, name="pf_checon_par")
      IMPORT :: c_int, c_ptr, c_int64_t, c_double
      TYPE(c_ptr), VALUE :: c, xnew, x
      REAL(c_double), VALUE :: tol
      INTEGER(c_int) :: conv
      INTEGER(c_int64_t), VALUE :: n
    END FUNCTION
    INTEGER(c_int) FUNCTION c_dot_order_set(c, order) BIND(C, name="pf_dot_order_set")
      IMPORT :: c_int, c_ptr
      TYPE(c_ptr), VALUE :: c
      INTEGER(c_int), VALUE :: order
    END FUNCTION
    INTEGER(c_int) FUNCTION c_vec_fill(c, x, val, n) BIND(C, name="pf_vec_fill")
      IMPORT :: c_int, c_ptr, c_int64_t, c_double
      TYPE(c_ptr), VALUE :: c, x
      REAL(c_double), VALUE :: val
      INTEGER(c_int64_t), VALUE :: n
    END FUNCTION
    INTEGER(c_int) FUNCTION c_vec_waxpy(c, w, x, y, a, n) BIND(C, name="pf_vec_waxpy")
      IMPORT :: c_int, c_ptr, c_int64_t, c_double
      TYPE(c_ptr), VALUE :: c, w, x, y
      REAL(c_double), VALUE :: a
      INTEGER(c_int64_t), VALUE :: n
    END FUNCTION
    INTEGER(c_int) FUNCTION c_vec_mul(c, z, x, y, n) BIND(C, name="pf_vec_mul")
      IMPORT :: c_int, c_ptr, c_int64_t
      TYPE(c_ptr), VALUE :: c, z, x, y
      INTEGER(c_int64_t), VALUE :: n
    END FUNCTION
    INTEGER(c_int) FUNCTION c_vec_recip(c, y, x, n) BIND(C, name="pf_vec_recip")
      IMPORT :: c_int, c_ptr, c_int64_t
      TYPE(c_ptr), VALUE :: c, y, x
      INTEGER(c_int64_t), VALUE :: n
    END FUNCTION
    INTEGER(c_int) FUNCTION c_pcg_solve(c, r0, x, opts, res) BIND(C, name="pf_pcg_solve")
      IMPORT :: c_int, c_ptr, c_double, pf_pcg_opts, pf_pcg_result
      TYPE(c_ptr), VALUE :: c
      REAL(c_double), INTENT(IN) :: r0(*)
      REAL(c_double), INTENT(OUT) :: x(*)
      TYPE(pf_pcg_opts), INTENT(IN) :: opts
      TYPE(pf_pcg_result), INTENT(OUT) :: res
    END FUNCTION
    INTEGER(c_int) FUNCTION c_stress_centroid(c, x, e, v, e_elem, sigma) BIND(C, name="pf_stress_centroid")
      IMPORT :: c_int, c_ptr, c_double
      TYPE(c_ptr), VALUE :: c
      REAL(c_double), INTENT(IN) :: x(*)
      REAL(c_double), VALUE :: e, v
      TYPE(c_ptr), VALUE :: e_elem
      REAL(c_double), INTENT(OUT) :: sigma(*)
    END FUNCTION
    INTEGER(c_int) FUNCTION c_assemble_transient(c, theta, dtim, np_types, prop, etype, store) &
        BIND(C, name="pf_assemble_transient")
      IMPORT :: c_int, c_ptr, c_double
      TYPE(c_ptr), VALUE :: c
      REAL(c_double), VALUE :: theta, dtim
      INTEGER(c_int), VALUE :: np_types, store
      REAL(c_double), INTENT(IN) :: prop(*)
      INTEGER(c_int), INTENT(IN) :: etype(*)
    END FUNCTION
    INTEGER(c_int) FUNCTION c_plastic_update(c, x, evpt, tensor, utemp, e, v, phi, coh, psi, dt, final) &
        BIND(C, name="pf_plastic_update")
      IMPORT :: c_int, c_ptr, c_double
      TYPE(c_ptr), VALUE :: c, x, evpt, tensor, utemp
      REAL(c_double), VALUE :: e, v, phi, coh, psi, dt
      INTEGER(c_int), VALUE :: final
    END FUNCTION
    INTEGER(c_int) FUNCTION c_pcg_solve_warm(c, r0, x0, xnew, mode, opts, res) BIND(C, name="pf_pcg_solve_warm")
      IMPORT :: c_int, c_ptr, pf_pcg_opts, pf_pcg_result
      TYPE(c_ptr), VALUE :: c, r0, x0, xnew
      INTEGER(c_int), VALUE :: mode
      TYPE(pf_pcg_opts), INTENT(IN) :: opts
      TYPE(pf_pcg_result), INTENT(OUT) :: res
    END FUNCTION c_pcg_solve_warm
    INTEGER(c_int) FUNCTION c_fixed_rows(c, u, p, mode) BIND(C, name="pf_fixed_rows")
      IMPORT :: c_int, c_ptr
      TYPE(c_ptr), VALUE :: c, u, p
      INTEGER(c_int), VALUE :: mode
    END FUNCTION
    INTEGER(c_int) FUNCTION c_apply_raw(c, which, x, y) BIND(C, name="pf_apply_raw")
      IMPORT :: c_int, c_ptr, c_double
      TYPE(c_ptr), VALUE :: c
      INTEGER(c_int), VALUE :: which
      REAL(c_double), INTENT(IN) :: x(*)
      REAL(c_double), INTENT(OUT) :: y(*)
    END FUNCTION
    INTEGER(c_int) FUNCTION c_scatter_nodes(c, nn, nod, nels, g_num, numvar, contrib, nodal, flag) &
        BIND(C, name="pf_scatter_nodes")
      IMPORT :: c_int, c_ptr, c_double
      TYPE(c_ptr), VALUE :: c
      INTEGER(c_int), VALUE :: nn, nod, nels, numvar, flag
      INTEGER(c_int), INTENT(IN) :: g_num(*)
      REAL(c_double), INTENT(IN) :: contrib(*)
      REAL(c_double), INTENT(OUT) :: nodal(*)
    END FUNCTION
  END INTERFACE

CONTAINS

  !> Stop with the library's message on a failed call (the reference's fatal
  !> paths STOP as well, e.g. gather_scatter.f90:804-816).
  SUBROUTINE pf_check(rc, what)
    INTEGER(c_int), INTENT(IN) :: rc
    CHARACTER(*), INTENT(IN) :: what
    CHARACTER(KIND=c_char), POINTER :: msg(:)
    INTEGER :: i
    CHARACTER(LEN=256) :: text
    IF (rc == PF_OK) RETURN
    CALL c_f_pointer(c_status_string(rc), msg, [256])
    text = ' '
    DO i = 1, 256
      IF (msg(i) == c_null_char) EXIT
      text(i:i) = msg(i)
    END DO
    WRITE(*, '(A,A,A,I0,A,A)') 'pf_gpu: ', what, ' failed, status ', rc, ': ', TRIM(text)
    STOP 1
  END SUBROUTINE pf_check

  !> FIND_PE_PROCS counterpart: one context on one GPU.
  SUBROUTINE pf_init(device)
    INTEGER, INTENT(IN) :: device
    CALL pf_check(c_create(INT(device, c_int), ctx), 'pf_create')
  END SUBROUTINE pf_init

  !> SHUTDOWN counterpart.
  SUBROUTINE pf_finalize()
    INTEGER(c_int) :: rc
    rc = c_destroy(ctx)            ! frees every device buffer, the staging ones too
    ctx = c_null_ptr
    st_a = pf_dvec()
    st_b = pf_dvec()
    m_neq = 0
    m_nslot = 0
  END SUBROUTINE pf_finalize

  SUBROUTINE pf_alloc(v, n)
    TYPE(pf_dvec), INTENT(INOUT) :: v
    INTEGER, INTENT(IN) :: n
    v%n = INT(n, c_int64_t)
    CALL pf_check(c_dev_alloc(ctx, INT(MAX(n, 1), c_size_t) * 8_c_size_t, v%ptr), 'pf_dev_alloc')
  END SUBROUTINE pf_alloc

  SUBROUTINE pf_release(v)
    TYPE(pf_dvec), INTENT(INOUT) :: v
    INTEGER(c_int) :: rc
    IF (c_associated(v%ptr)) rc = c_dev_free(ctx, v%ptr)
    v%ptr = c_null_ptr
    v%n = 0
  END SUBROUTINE pf_release

  !> Host array -> device vector (sizes must match).
  SUBROUTINE pf_put(v, a)
    TYPE(pf_dvec), INTENT(INOUT) :: v
    REAL(iwp), INTENT(IN), TARGET, CONTIGUOUS :: a(:)
    IF (INT(SIZE(a), c_int64_t) /= v%n) CALL pf_check(1_c_int, 'pf_put size')
    CALL pf_check(c_h2d(ctx, v%ptr, c_loc(a), INT(v%n, c_size_t) * 8_c_size_t), 'pf_memcpy_h2d')
  END SUBROUTINE pf_put

  !> Device vector -> host array.
  SUBROUTINE pf_get(v, a)
    TYPE(pf_dvec), INTENT(IN) :: v
    REAL(iwp), INTENT(INOUT), TARGET, CONTIGUOUS :: a(:)
    IF (INT(SIZE(a), c_int64_t) /= v%n) CALL pf_check(1_c_int, 'pf_get size')
    CALL pf_check(c_d2h(ctx, c_loc(a), v%ptr, INT(v%n, c_size_t) * 8_c_size_t), 'pf_memcpy_d2h')
  END SUBROUTINE pf_get

  !> g_coord_pp(nod,3,nels) (READ_G_COORD_PP's array) to the device.
  SUBROUTINE pf_mesh(g_coord_pp)
    REAL(iwp), INTENT(IN) :: g_coord_pp(:, :, :)
    CALL pf_check(c_mesh_set(ctx, INT(SIZE(g_coord_pp, 1), c_int), INT(SIZE(g_coord_pp, 3), c_int), &
                             g_coord_pp), 'pf_mesh_set')
  END SUBROUTINE pf_mesh

  !> g_g_pp(ntot,nels) (find_g3/find_g4's array) and neq to the device.
  SUBROUTINE pf_steer(g_g_pp, neq)
    INTEGER, INTENT(IN) :: g_g_pp(:, :), neq
    CALL pf_check(c_steer_set(ctx, INT(SIZE(g_g_pp, 1), c_int), INT(SIZE(g_g_pp, 2), c_int), &
                              INT(neq, c_int), g_g_pp), 'pf_steer_set')
    m_neq = INT(neq, c_int64_t)
    m_nslot = INT(SIZE(g_g_pp), c_int64_t)
  END SUBROUTINE pf_steer

  !> p121's element loop (p121.f90:54-64) on the device; e_elem optional.
  SUBROUTINE pf_assemble_elastic(e, v, store, e_elem)
    REAL(iwp), INTENT(IN) :: e, v
    INTEGER, INTENT(IN) :: store
    REAL(iwp), INTENT(IN), TARGET, OPTIONAL :: e_elem(:)
    TYPE(c_ptr) :: pe
    pe = c_null_ptr
    IF (PRESENT(e_elem)) pe = c_loc(e_elem)
    CALL pf_check(c_assemble_elastic(ctx, e, v, pe, INT(store, c_int)), 'pf_assemble_elastic')
  END SUBROUTINE pf_assemble_elastic

  !> p123's element loop (p123.f90:70-84) on the device.
  SUBROUTINE pf_assemble_laplace(kx, ky, kz, store, k_elem)
    REAL(iwp), INTENT(IN) :: kx, ky, kz
    INTEGER, INTENT(IN) :: store
    REAL(iwp), INTENT(IN), TARGET, OPTIONAL :: k_elem(:)
    TYPE(c_ptr) :: pk
    pk = c_null_ptr
    IF (PRESENT(k_elem)) pk = c_loc(k_elem)
    CALL pf_check(c_assemble_laplace(ctx, kx, ky, kz, pk, INT(store, c_int)), 'pf_assemble_laplace')
  END SUBROUTINE pf_assemble_laplace

  !> A driver's own storkm_pp(ntot,ntot,nels) to the device.
  SUBROUTINE pf_storkm(storkm_pp, store)
    REAL(iwp), INTENT(IN) :: storkm_pp(:, :, :)
    INTEGER, INTENT(IN) :: store
    CALL pf_check(c_storkm_set(ctx, storkm_pp, INT(store, c_int)), 'pf_storkm_set')
  END SUBROUTINE pf_storkm

  !> storkm_pp(ntot,ntot,nels) from the device (FP64, or hi+lo).
  SUBROUTINE pf_storkm_get(storkm_pp)
    REAL(iwp), INTENT(OUT) :: storkm_pp(:, :, :)
    CALL pf_check(c_storkm_get(ctx, storkm_pp), 'pf_storkm_get')
  END SUBROUTINE pf_storkm_get

  !> PCG_KM(diag_precon_pp, km, limit, nels_pp, r_pp, timest, tol, xnew_pp,
  !> iters) (src/modules/mpi/maths.f90:1152-1323): every element has the
  !> matrix km; diag_precon_pp is the inverted Jacobi preconditioner; stops
  !> by checon_par; r_pp returns the final recurrence residual.  The steering
  !> must be set (pf_steer) with nels_pp elements; timest is not touched.
  SUBROUTINE pcg_km(diag_precon_pp, km, limit, nels_pp, r_pp, timest, tol, xnew_pp, iters)
    REAL(iwp), INTENT(IN) :: diag_precon_pp(:), km(:, :), tol
    INTEGER, INTENT(IN) :: limit, nels_pp
    REAL(iwp), INTENT(INOUT) :: r_pp(:), xnew_pp(:), timest(:)
    INTEGER, INTENT(OUT) :: iters
    TYPE(pf_pcg_opts) :: o
    TYPE(pf_pcg_result) :: res
    INTEGER(c_int) :: rc
    IF (nels_pp < 1 .OR. SIZE(km, 1) /= SIZE(km, 2)) CALL pf_check(1_c_int, 'PCG_KM arguments')
    CALL pf_check(c_storkm_shared(ctx, km), 'PCG_KM km')
    CALL pf_check(c_precon_set(ctx, diag_precon_pp), 'PCG_KM diag_precon_pp')
    o%limit = limit
    o%tol = tol
    o%stop_mode = PF_STOP_CHECON
    o%precision = PF_PREC_FP64
    o%check_every = 0
    o%refine_every = 0
    o%refine_delta = 0.0_iwp
    rc = c_pcg_solve(ctx, r_pp, xnew_pp, o, res)
    IF (rc /= PF_OK .AND. rc /= PF_NOT_CONVERGED) CALL pf_check(rc, 'PCG_KM')
    CALL pf_check(c_pcg_r(ctx, r_pp, 0_c_int), 'PCG_KM r_pp')
    iters = res%iters
  END SUBROUTINE pcg_km

  !> Assembled diagonal (diag_precon before inversion, p121.f90:65-69).
  SUBROUTINE pf_diag(diag)
    REAL(iwp), INTENT(OUT) :: diag(:)
    CALL pf_check(c_diag_get(ctx, diag), 'pf_diag_get')
  END SUBROUTINE pf_diag

  !> p123 penalty rows (p123.f90:120-125); returns store_pp.
  SUBROUTINE pf_fixed(no_f, penalty, store)
    INTEGER, INTENT(IN) :: no_f(:)
    REAL(iwp), INTENT(IN) :: penalty
    REAL(iwp), INTENT(OUT) :: store(:)
    CALL pf_check(c_fixed_set(ctx, INT(SIZE(no_f), c_int), no_f, penalty, store), 'pf_fixed_set')
  END SUBROUTINE pf_fixed

  !> GATHER(p_pp, pmul_pp): pmul(:,e) = p(g_g(:,e)), 0 for restrained.
  SUBROUTINE gather_dev(p_pp, pmul_pp)
    TYPE(pf_dvec), INTENT(IN) :: p_pp
    TYPE(pf_dvec), INTENT(INOUT) :: pmul_pp
    CALL pf_check(c_gather(ctx, p_pp%ptr, pmul_pp%ptr), 'GATHER')
  END SUBROUTINE gather_dev

  !> utemp(:,e) = MATMUL(storkm(:,:,e), pmul(:,e)) for every element.
  SUBROUTINE elem_matvec(pmul_pp, utemp_pp)
    TYPE(pf_dvec), INTENT(IN) :: pmul_pp
    TYPE(pf_dvec), INTENT(INOUT) :: utemp_pp
    CALL pf_check(c_elem_matvec(ctx, pmul_pp%ptr, utemp_pp%ptr), 'element products')
  END SUBROUTINE elem_matvec

  !> SCATTER(u_pp, utemp_pp): u accumulates, ascending element order.
  SUBROUTINE scatter_dev(u_pp, utemp_pp)
    TYPE(pf_dvec), INTENT(INOUT) :: u_pp
    TYPE(pf_dvec), INTENT(IN) :: utemp_pp
    CALL pf_check(c_scatter(ctx, u_pp%ptr, utemp_pp%ptr), 'SCATTER')
  END SUBROUTINE scatter_dev

  !> SCATTER_NOADD(utemp_pp, u_pp): assignment, the last element wins.
  SUBROUTINE scatter_noadd_dev(utemp_pp, u_pp)
    TYPE(pf_dvec), INTENT(IN) :: utemp_pp
    TYPE(pf_dvec), INTENT(INOUT) :: u_pp
    CALL pf_check(c_scatter_noadd(ctx, utemp_pp%ptr, u_pp%ptr), 'SCATTER_NOADD')
  END SUBROUTINE scatter_noadd_dev

  FUNCTION dot_product_p_dev(vectora, vectorb) RESULT(s)
    TYPE(pf_dvec), INTENT(IN) :: vectora, vectorb
    REAL(iwp) :: s
    REAL(c_double) :: out
    CALL pf_check(c_dot(ctx, vectora%ptr, vectorb%ptr, vectora%n, out), 'DOT_PRODUCT_P')
    s = out
  END FUNCTION dot_product_p_dev

  FUNCTION sum_p_dev(vectora) RESULT(s)
    TYPE(pf_dvec), INTENT(IN) :: vectora
    REAL(iwp) :: s
    REAL(c_double) :: out
    CALL pf_check(c_sum(ctx, vectora%ptr, vectora%n, out), 'SUM_P')
    s = out
  END FUNCTION sum_p_dev

  !> checon_par(loads, tol, converged, oldlds): oldlds = loads on return.
  SUBROUTINE checon_par_dev(loads, tol, converged, oldlds)
    TYPE(pf_dvec), INTENT(IN) :: loads
    REAL(iwp), INTENT(IN) :: tol
    LOGICAL, INTENT(OUT) :: converged
    TYPE(pf_dvec), INTENT(INOUT) :: oldlds
    INTEGER(c_int) :: conv
    CALL pf_check(c_checon(ctx, loads%ptr, tol, conv, oldlds%ptr, loads%n), 'checon_par')
    converged = conv /= 0
  END SUBROUTINE checon_par_dev

  !------------------------------------------------ host-array forms
  ! The reference's signatures on REAL(iwp) arrays: GATHER(p_pp,pmul_pp)
  ! (gather_scatter.f90:580-581), SCATTER(u_pp,utemp_pp) (:726-727),
  ! SCATTER_NOADD(utemp_pp,u_pp) (:856,913), DOT_PRODUCT_P(vectora,vectorb)
  ! (maths.f90:206-208), SUM_P(vectora) (:271-315) and
  ! checon_par(loads,tol,converged,oldlds) (:1047-1050).  Each copies its
  ! inputs to device staging buffers, runs the device form and copies the
  ! results back: the same bits as the device form on the same values.

  !> Device staging buffer holding at least n REAL(iwp).
  SUBROUTINE stage(buf, n)
    TYPE(pf_dvec), INTENT(INOUT) :: buf
    INTEGER(c_int64_t), INTENT(IN) :: n
    IF (buf%n >= n .AND. c_associated(buf%ptr)) RETURN
    CALL pf_release(buf)
    buf%n = MAX(n, 1_c_int64_t)
    CALL pf_check(c_dev_alloc(ctx, INT(buf%n, c_size_t) * 8_c_size_t, buf%ptr), 'pf_dev_alloc (staging)')
  END SUBROUTINE stage

  SUBROUTINE put_n(buf, a, n)
    TYPE(pf_dvec), INTENT(INOUT) :: buf
    REAL(iwp), INTENT(IN), TARGET, CONTIGUOUS :: a(:)
    INTEGER(c_int64_t), INTENT(IN) :: n
    CALL stage(buf, n)
    IF (n > 0) CALL pf_check(c_h2d(ctx, buf%ptr, c_loc(a), INT(n, c_size_t) * 8_c_size_t), 'pf_memcpy_h2d')
  END SUBROUTINE put_n

  SUBROUTINE get_n(buf, a, n)
    TYPE(pf_dvec), INTENT(IN) :: buf
    REAL(iwp), INTENT(INOUT), TARGET, CONTIGUOUS :: a(:)
    INTEGER(c_int64_t), INTENT(IN) :: n
    IF (n > 0) CALL pf_check(c_d2h(ctx, c_loc(a), buf%ptr, INT(n, c_size_t) * 8_c_size_t), 'pf_memcpy_d2h')
  END SUBROUTINE get_n

  !> The sizes the steering fixes: neq_pp for vectors, ntot*nels_pp for
  !> element arrays (the reference's arrays have these shapes at one rank).
  SUBROUTINE need_sizes(nvec, nelem, what)
    INTEGER(c_int64_t), INTENT(IN) :: nvec, nelem
    CHARACTER(*), INTENT(IN) :: what
    IF (m_neq == 0 .OR. nvec /= m_neq .OR. (nelem >= 0 .AND. nelem /= m_nslot)) THEN
      WRITE(*, '(A,A,A,I0,A,I0,A)') 'pf_gpu: ', what, ': arrays must hold neq_pp = ', m_neq, &
        ' and ntot*nels_pp = ', m_nslot, ' entries (pf_steer sets them)'
      STOP 1
    END IF
  END SUBROUTINE need_sizes

  SUBROUTINE gather_host(p_pp, pmul_pp)
    REAL(iwp), INTENT(IN), CONTIGUOUS :: p_pp(:)
    REAL(iwp), INTENT(INOUT), CONTIGUOUS, TARGET :: pmul_pp(:, :)
    REAL(iwp), POINTER, CONTIGUOUS :: flat(:)
    CALL need_sizes(INT(SIZE(p_pp), c_int64_t), INT(SIZE(pmul_pp), c_int64_t), 'GATHER')
    CALL put_n(st_a, p_pp, m_neq)
    CALL stage(st_b, m_nslot)
    CALL pf_check(c_gather(ctx, st_a%ptr, st_b%ptr), 'GATHER')
    flat(1:SIZE(pmul_pp)) => pmul_pp
    CALL get_n(st_b, flat, m_nslot)
  END SUBROUTINE gather_host

  SUBROUTINE scatter_host(u_pp, utemp_pp)
    REAL(iwp), INTENT(INOUT), CONTIGUOUS :: u_pp(:)
    REAL(iwp), INTENT(IN), CONTIGUOUS, TARGET :: utemp_pp(:, :)
    REAL(iwp), POINTER, CONTIGUOUS :: flat(:)
    CALL need_sizes(INT(SIZE(u_pp), c_int64_t), INT(SIZE(utemp_pp), c_int64_t), 'SCATTER')
    flat(1:SIZE(utemp_pp)) => utemp_pp
    CALL put_n(st_b, flat, m_nslot)
    CALL put_n(st_a, u_pp, m_neq)
    CALL pf_check(c_scatter(ctx, st_a%ptr, st_b%ptr), 'SCATTER')
    CALL get_n(st_a, u_pp, m_neq)
  END SUBROUTINE scatter_host

  SUBROUTINE scatter_noadd_host(utemp_pp, u_pp)
    REAL(iwp), INTENT(IN), CONTIGUOUS, TARGET :: utemp_pp(:, :)
    REAL(iwp), INTENT(INOUT), CONTIGUOUS :: u_pp(:)
    REAL(iwp), POINTER, CONTIGUOUS :: flat(:)
    CALL need_sizes(INT(SIZE(u_pp), c_int64_t), INT(SIZE(utemp_pp), c_int64_t), 'SCATTER_NOADD')
    flat(1:SIZE(utemp_pp)) => utemp_pp
    CALL put_n(st_b, flat, m_nslot)
    CALL put_n(st_a, u_pp, m_neq)      ! equations no element touches keep their value
    CALL pf_check(c_scatter_noadd(ctx, st_b%ptr, st_a%ptr), 'SCATTER_NOADD')
    CALL get_n(st_a, u_pp, m_neq)
  END SUBROUTINE scatter_noadd_host

  FUNCTION dot_product_p_host(vectora, vectorb) RESULT(s)
    REAL(iwp), INTENT(IN), CONTIGUOUS :: vectora(:), vectorb(:)
    REAL(iwp) :: s
    REAL(c_double) :: out
    INTEGER(c_int64_t) :: n
    n = INT(SIZE(vectora), c_int64_t)
    IF (INT(SIZE(vectorb), c_int64_t) /= n) CALL pf_check(1_c_int, 'DOT_PRODUCT_P sizes')
    CALL put_n(st_a, vectora, n)
    CALL put_n(st_b, vectorb, n)
    CALL pf_check(c_dot(ctx, st_a%ptr, st_b%ptr, n, out), 'DOT_PRODUCT_P')
    s = out
  END FUNCTION dot_product_p_host

  FUNCTION sum_p_host(vectora) RESULT(s)
    REAL(iwp), INTENT(IN), CONTIGUOUS :: vectora(:)
    REAL(iwp) :: s
    REAL(c_double) :: out
    INTEGER(c_int64_t) :: n
    n = INT(SIZE(vectora), c_int64_t)
    CALL put_n(st_a, vectora, n)
    CALL pf_check(c_sum(ctx, st_a%ptr, n, out), 'SUM_P')
    s = out
  END FUNCTION sum_p_host

  SUBROUTINE checon_par_host(loads, tol, converged, oldlds)
    REAL(iwp), INTENT(IN), CONTIGUOUS :: loads(:)
    REAL(iwp), INTENT(IN) :: tol
    LOGICAL, INTENT(OUT) :: converged
    REAL(iwp), INTENT(INOUT), CONTIGUOUS :: oldlds(:)
    INTEGER(c_int) :: conv
    INTEGER(c_int64_t) :: n
    n = INT(SIZE(loads), c_int64_t)
    IF (INT(SIZE(oldlds), c_int64_t) /= n) CALL pf_check(1_c_int, 'checon_par sizes')
    CALL put_n(st_a, loads, n)
    CALL put_n(st_b, oldlds, n)
    CALL pf_check(c_checon(ctx, st_a%ptr, tol, conv, st_b%ptr, n), 'checon_par')
    converged = conv /= 0
    oldlds = loads
  END SUBROUTINE checon_par_host

  !> Summation order of DOT_PRODUCT_P, SUM_P and the PCG's dots:
  !> PF_DOT_TREE (the device's fixed tree, default) or PF_DOT_SEQ (the
  !> reference's sequential DOT_PRODUCT order at one rank).
  SUBROUTINE pf_dot_order(order)
    INTEGER, INTENT(IN) :: order
    CALL pf_check(c_dot_order_set(ctx, INT(order, c_int)), 'pf_dot_order_set')
  END SUBROUTINE pf_dot_order

  !> One element product u = km p as the library forms it: a chain of fused
  !> multiply-adds in ascending column order, s = fma(km(i,j), p(j), s) from
  !> s = 0 (MATMUL fixes no order; this is the device kernel's), so that a
  !> driver's host element loop returns the device's bits.
  FUNCTION pf_matvec(km, p) RESULT(u)
    USE, INTRINSIC :: IEEE_ARITHMETIC, ONLY: IEEE_FMA
    REAL(iwp), INTENT(IN) :: km(:, :), p(:)
    REAL(iwp) :: u(SIZE(km, 1))
    INTEGER :: i, j
    REAL(iwp) :: s
    DO i = 1, SIZE(km, 1)
      s = 0.0_iwp
      DO j = 1, SIZE(km, 2)
        s = IEEE_FMA(km(i, j), p(j), s)
      END DO
      u(i) = s
    END DO
  END FUNCTION pf_matvec

  SUBROUTINE vec_fill(x, val)
    TYPE(pf_dvec), INTENT(INOUT) :: x
    REAL(iwp), INTENT(IN) :: val
    CALL pf_check(c_vec_fill(ctx, x%ptr, val, x%n), 'vec_fill')
  END SUBROUTINE vec_fill

  !> w = x + y*a
  SUBROUTINE vec_waxpy(w, x, y, a)
    TYPE(pf_dvec), INTENT(INOUT) :: w
    TYPE(pf_dvec), INTENT(IN) :: x, y
    REAL(iwp), INTENT(IN) :: a
    CALL pf_check(c_vec_waxpy(ctx, w%ptr, x%ptr, y%ptr, a, w%n), 'vec_waxpy')
  END SUBROUTINE vec_waxpy

  !> z = x*y
  SUBROUTINE vec_mul(z, x, y)
    TYPE(pf_dvec), INTENT(INOUT) :: z
    TYPE(pf_dvec), INTENT(IN) :: x, y
    CALL pf_check(c_vec_mul(ctx, z%ptr, x%ptr, y%ptr, z%n), 'vec_mul')
  END SUBROUTINE vec_mul

  !> y = 1/x
  SUBROUTINE vec_recip(y, x)
    TYPE(pf_dvec), INTENT(INOUT) :: y
    TYPE(pf_dvec), INTENT(IN) :: x
    CALL pf_check(c_vec_recip(ctx, y%ptr, x%ptr, y%n), 'vec_recip')
  END SUBROUTINE vec_recip

  !> The whole PCG loop on the device (pf_pcg_solve).
  SUBROUTINE pf_pcg(r0, x, limit, tol, stop_mode, precision, iters, converged)
    REAL(iwp), INTENT(IN) :: r0(:), tol
    REAL(iwp), INTENT(OUT) :: x(:)
    INTEGER, INTENT(IN) :: limit, stop_mode, precision
    INTEGER, INTENT(OUT) :: iters
    LOGICAL, INTENT(OUT) :: converged
    TYPE(pf_pcg_opts) :: o
    TYPE(pf_pcg_result) :: res
    INTEGER(c_int) :: rc
    o%limit = limit
    o%tol = tol
    o%stop_mode = stop_mode
    o%precision = precision
    o%check_every = 0
    o%refine_every = 0
    o%refine_delta = 0.0_iwp
    rc = c_pcg_solve(ctx, r0, x, o, res)
    IF (rc /= PF_OK .AND. rc /= PF_NOT_CONVERGED) CALL pf_check(rc, 'pf_pcg_solve')
    iters = res%iters
    converged = res%converged /= 0
  END SUBROUTINE pf_pcg

  !> p124's element loop (p124.f90:73-86) on the device: storka_pp (the
  !> PCG's matrices, stored as `store`) and storkb_pp from prop(5,np_types)
  !> (kx, ky, kz, rho, cp) and the element types etype_pp.
  SUBROUTINE pf_assemble_transient(theta, dtim, prop, etype_pp, store)
    REAL(iwp), INTENT(IN) :: theta, dtim, prop(:, :)
    INTEGER, INTENT(IN) :: etype_pp(:), store
    CALL pf_check(c_assemble_transient(ctx, theta, dtim, INT(SIZE(prop, 2), c_int), prop, etype_pp, &
                                       INT(store, c_int)), 'pf_assemble_transient')
  END SUBROUTINE pf_assemble_transient

  !> u = SCATTER(storka_pp . GATHER(x)) (which = 0) or storkb_pp (which = 1),
  !> without fixed-freedom rows (p124.f90:158-162, 171-174).
  SUBROUTINE pf_apply_raw(which, x, u)
    INTEGER, INTENT(IN) :: which
    REAL(iwp), INTENT(IN) :: x(:)
    REAL(iwp), INTENT(OUT) :: u(:)
    CALL pf_check(c_apply_raw(ctx, INT(which, c_int), x, u), 'pf_apply_raw')
  END SUBROUTINE pf_apply_raw

  !> p122's stress update and body loads (p122.f90:164-199) on the device:
  !> from loads_pp (the displacement increment), evpt_pp and tensor_pp
  !> (nst*nip*nels_pp) are updated and each element's body loads written to
  !> utemp_pp for SCATTER.  final: plastic_converged .OR. plasiters==plasits.
  SUBROUTINE pf_plastic_update(loads_pp, evpt_pp, tensor_pp, utemp_pp, e, v, phi, c, psi, dt, final)
    TYPE(pf_dvec), INTENT(IN) :: loads_pp
    TYPE(pf_dvec), INTENT(INOUT) :: evpt_pp, tensor_pp, utemp_pp
    REAL(iwp), INTENT(IN) :: e, v, phi, c, psi, dt
    LOGICAL, INTENT(IN) :: final
    INTEGER(c_int) :: f
    f = 0
    IF (final) f = 1
    CALL pf_check(c_plastic_update(ctx, loads_pp%ptr, evpt_pp%ptr, tensor_pp%ptr, utemp_pp%ptr, e, v, phi, c, &
                                   psi, dt, f), 'pf_plastic_update')
  END SUBROUTINE pf_plastic_update

  !> The fixed-freedom rows of p122's CG product (p122.f90:141-146):
  !> u(j) = p(j)*store (first plastic iteration) or 0.
  SUBROUTINE pf_fixed_rows(u_pp, p_pp, first)
    TYPE(pf_dvec), INTENT(INOUT) :: u_pp
    TYPE(pf_dvec), INTENT(IN) :: p_pp
    LOGICAL, INTENT(IN) :: first
    INTEGER(c_int) :: mode
    mode = 0
    IF (first) mode = 1
    CALL pf_check(c_fixed_rows(ctx, u_pp%ptr, p_pp%ptr, mode), 'pf_fixed_rows')
  END SUBROUTINE pf_fixed_rows

  !> p122's conjugate_gradients loop (p122.f90:147-162) as one call: from
  !> x_pp and r_pp (= loads - K x_pp, formed by the caller as p122.f90:141-145
  !> does), with the fixed rows of pf_fixed as pf_fixed_rows(first) sets
  !> them, to checon_par at cjtol or cjits iterations; xnew_pp receives the
  !> last iterate (and x_pp, as checon_par leaves it), cjiters the count.
  SUBROUTINE pf_cg_warm(r_pp, x_pp, xnew_pp, first, cjits, cjtol, cjiters)
    TYPE(pf_dvec), INTENT(IN) :: r_pp
    TYPE(pf_dvec), INTENT(INOUT) :: x_pp, xnew_pp
    LOGICAL, INTENT(IN) :: first
    INTEGER, INTENT(IN) :: cjits
    REAL(iwp), INTENT(IN) :: cjtol
    INTEGER, INTENT(OUT) :: cjiters
    TYPE(pf_pcg_opts) :: o
    TYPE(pf_pcg_result) :: res
    INTEGER(c_int) :: rc, mode
    mode = 0
    IF (first) mode = 1
    o%limit = cjits
    o%tol = cjtol
    o%stop_mode = PF_STOP_CHECON
    o%precision = PF_PREC_FP64
    o%check_every = 0
    o%refine_every = 0
    o%refine_delta = 0.0_iwp
    rc = c_pcg_solve_warm(ctx, r_pp%ptr, x_pp%ptr, xnew_pp%ptr, mode, o, res)
    IF (rc /= PF_OK .AND. rc /= PF_NOT_CONVERGED) CALL pf_check(rc, 'pf_pcg_solve_warm')
    CALL vec_waxpy(x_pp, xnew_pp, xnew_pp, 0.0_iwp)
    cjiters = res%iters
  END SUBROUTINE pf_cg_warm

  !> Stresses at every element's centroid (p121.f90:112-123 for each element):
  !> sigma(6,nels) from the displacements x(neq) and DEEMAT(e, v).
  SUBROUTINE pf_stress(x, e, v, sigma)
    REAL(iwp), INTENT(IN) :: x(:), e, v
    REAL(iwp), INTENT(OUT) :: sigma(:, :)
    CALL pf_check(c_stress_centroid(ctx, x, e, v, c_null_ptr, sigma), 'pf_stress_centroid')
  END SUBROUTINE pf_stress

  !> CALC_NODES_PP (src/modules/mpi/gather_scatter.f90) at one rank: every
  !> node is this rank's.
  SUBROUTINE calc_nodes_pp(nn, npes, numpe, node_end, node_start, nodes_pp)
    INTEGER, INTENT(IN) :: nn, npes, numpe
    INTEGER, INTENT(OUT) :: node_end, node_start, nodes_pp
    IF (npes /= 1 .OR. numpe /= 1) THEN
      WRITE(*, '(A)') 'pf_gpu: calc_nodes_pp runs on one process'
      STOP 1
    END IF
    node_start = 1
    node_end = nn
    nodes_pp = nn
  END SUBROUTINE calc_nodes_pp

  !> SCATTER_NODES (src/modules/mpi/gather_scatter.f90:1786-1929) with the
  !> reference's arguments, at one rank: nodal sums, in ascending element
  !> order, of element_contribution(numvar*nod, nels_pp); flag 1 averages.
  SUBROUTINE scatter_nodes(npes, nn, nels_pp, g_num_pp, nod, numvar, nodes_pp, node_start, node_end, &
                           element_contribution, nodal_value, flag)
    INTEGER, INTENT(IN) :: npes, nn, nels_pp, nod, numvar, nodes_pp, node_start, node_end, flag
    INTEGER, INTENT(IN) :: g_num_pp(:, :)
    REAL(iwp), INTENT(IN) :: element_contribution(:, :)
    REAL(iwp), INTENT(OUT) :: nodal_value(:)
    IF (npes /= 1 .OR. node_start /= 1 .OR. node_end /= nn .OR. nodes_pp /= nn) THEN
      WRITE(*, '(A)') 'pf_gpu: scatter_nodes runs on one process (all nodes local)'
      STOP 1
    END IF
    CALL pf_check(c_scatter_nodes(ctx, INT(nn, c_int), INT(nod, c_int), INT(nels_pp, c_int), g_num_pp, &
                                  INT(numvar, c_int), element_contribution, nodal_value, INT(flag, c_int)), &
                  'pf_scatter_nodes')
  END SUBROUTINE scatter_nodes

  !> DISMSH_ENSI_P (src/modules/mpi/output.f90:2983-3111) at one rank: one
  !> value per line, edit descriptor e12.5.
  SUBROUTINE dismsh_ensi_p(filnum, iload, nodes_pp, npes, numpe, numvar, stress)
    INTEGER, INTENT(IN) :: filnum, iload, nodes_pp, npes, numpe, numvar
    REAL(iwp), INTENT(IN) :: stress(:)
    INTEGER :: i
    IF (npes /= 1) THEN
      WRITE(*, '(A)') 'pf_gpu: dismsh_ensi_p runs on one process'
      STOP 1
    END IF
    IF (numpe == 1) THEN
      DO i = 1, nodes_pp
        WRITE(filnum, '(e12.5)') stress(i)
      END DO
    END IF
    IF (iload < 0 .OR. numvar < 0) RETURN   ! unused at one rank, as in the reference
  END SUBROUTINE dismsh_ensi_p

END MODULE pf_gpu
