!------------------------------------------------------------------------------
! pf_p121: the p121 program flow (src/programs/5th_ed/p121/p121.f90) on one
! MI355X through the pf_gpu module -- a 3-D elastic solid of 20-node bricks
! solved by Jacobi-preconditioned EBE-PCG.
!
!   pf_p121 <deck> [dropin|host|fused|hostseq|fusedseq|mixed|km] [relres-tol]
!
! Reads <deck>.dat/.d/.bnd/.lds in the formats READ_P121, READ_G_NUM_PP,
! READ_G_COORD_PP, READ_REST and READ_LOADS read (src/modules/mpi/
! input.f90:3234-3396, 935-1080, 288-445, 2570-2632, 2350-2411) and writes
! <deck>.pf.res (the lines p121.f90:72-74,83,106-109,114-121,138-139
! write), <deck>.pf.x (xnew_pp as neq raw REAL(8) values, stream access) and
! <deck>.pf.ensi.DISPL-000001 (the nodal displacements as p121.f90:125-137
! writes them: SCATTER_NODES averages of the gathered element displacements,
! DISMSH_ENSI_P per component).
!
!   dropin : p121's own loop, statement for statement (p121.f90:86-104), with
!            GATHER / element products / SCATTER / DOT_PRODUCT_P / checon_par
!            and the whole-array statements on device vectors.
!   host   : p121.f90:89-104 as the reference writes it, on host arrays
!            p_pp(neq), pmul_pp(ntot,nels), ... : the module's GATHER,
!            SCATTER, DOT_PRODUCT_P and checon_par generics on REAL(iwp)
!            arrays, the whole-array statements in Fortran, and the element
!            loop with pf_matvec (the device's order of the products) for
!            MATMUL.
!   fused  : the same iteration as one library call (pf_pcg, FP64, checon).
!   hostseq, fusedseq : host and fused with DOT_PRODUCT_P and SUM_P in the
!            reference's sequential order at one rank (PF_DOT_SEQ).
!   mixed  : FP32-stored element matrices with FP64 vectors and reliable
!            updates, stopped on the relative residual (default 1e-10).
!   km     : PCG_KM (maths.f90:1152-1323) with element 1's matrix for all
!            elements, for uniform meshes and materials.
!------------------------------------------------------------------------------
PROGRAM pf_p121
  USE pf_gpu
  IMPLICIT NONE
  INTEGER, PARAMETER :: nodof = 3, ndim = 3
  INTEGER :: nels, nn, nr, nip, nod, loaded_nodes, meshgen, partitioner, limit
  INTEGER :: ntot, neq, iel, i, j, k, iters, ios, nlen
  REAL(iwp) :: e, v, tol, up, alpha, beta, q, rtol
  LOGICAL :: converged
  CHARACTER(LEN=256) :: argv, mode, word, line
  CHARACTER(LEN=15) :: element
  INTEGER, ALLOCATABLE :: g_num(:, :), rest(:, :), nf(:, :), g_g_pp(:, :), node(:)
  REAL(iwp), ALLOCATABLE :: coord(:, :), g_coord_pp(:, :, :), val(:, :), r(:), diag(:), x(:)
  REAL(iwp), ALLOCATABLE :: storkm(:, :, :), timest(:), sigma(:, :), eld(:, :), disp(:), temp(:)
  INTEGER :: node_start, node_end, nodes_pp
  INTEGER(8) :: ta, tst
  INTEGER, ALLOCATABLE :: row(:)
  TYPE(pf_dvec) :: p_pp, r_pp, x_pp, xnew_pp, u_pp, d_pp, diag_precon_pp, pmul_pp, utemp_pp
  ! host mode: the reference's own arrays (p121.f90:20-25)
  REAL(iwp), ALLOCATABLE :: ph_pp(:), rh_pp(:), xh_pp(:), xnewh_pp(:), uh_pp(:), dh_pp(:), dph_pp(:)
  REAL(iwp), ALLOCATABLE :: pmulh_pp(:, :), utemph_pp(:, :), storkm_pp(:, :, :)
  INTEGER(8) :: t0, t1, rate

  CALL SYSTEM_CLOCK(tst, rate)
  CALL GET_COMMAND_ARGUMENT(1, argv, nlen)
  IF (nlen == 0) THEN
    WRITE(*, '(A)') 'usage: pf_p121 <deck> [dropin|host|fused|hostseq|fusedseq|mixed|km] [relres-tol]'
    STOP 2
  END IF
  mode = 'dropin'
  IF (COMMAND_ARGUMENT_COUNT() >= 2) CALL GET_COMMAND_ARGUMENT(2, mode)
  rtol = 1.0E-10_iwp
  IF (COMMAND_ARGUMENT_COUNT() >= 3) THEN
    CALL GET_COMMAND_ARGUMENT(3, word)
    READ(word, *) rtol
  END IF
  IF (mode /= 'dropin' .AND. mode /= 'fused' .AND. mode /= 'mixed' .AND. mode /= 'km' .AND. &
      mode /= 'host' .AND. mode /= 'hostseq' .AND. mode /= 'fusedseq') THEN
    WRITE(*, '(A,A)') 'pf_p121: unknown mode ', TRIM(mode)
    STOP 2
  END IF

  !---------------------------------------------------------------- .dat
  ! READ_P121: an optional 'gpu'/'cpu' line, element, meshgen, partitioner,
  ! nels nn nr nip nod loaded_nodes, e v tol limit.
  OPEN(10, FILE=argv(1:nlen)//'.dat', STATUS='OLD', ACTION='READ')
  READ(10, *) word
  IF (word == 'gpu' .OR. word == 'cpu') READ(10, *) word
  element = word
  READ(10, *) meshgen
  READ(10, *) partitioner
  READ(10, *) nels, nn, nr, nip, nod, loaded_nodes
  READ(10, *) e, v, tol, limit
  CLOSE(10)
  IF (nod /= 20 .AND. nod /= 8) THEN
    WRITE(*, '(A,I0)') 'pf_p121: unsupported nod ', nod
    STOP 2
  END IF
  ntot = nod * nodof

  !---------------------------------------------------------------- .d
  ALLOCATE(coord(ndim, nn), g_num(nod, nels), row(nod + 5))
  OPEN(10, FILE=argv(1:nlen)//'.d', STATUS='OLD', ACTION='READ')
  DO
    READ(10, '(A)') line
    IF (INDEX(ADJUSTL(line), '*NODES') == 1) EXIT
  END DO
  DO i = 1, nn
    READ(10, *) k, q, up, alpha
    coord(:, k) = [q, up, alpha]
  END DO
  DO
    READ(10, '(A)') line
    IF (INDEX(ADJUSTL(line), '*ELEMENTS') == 1) EXIT
  END DO
  DO iel = 1, nels
    READ(10, *) row
    g_num(:, iel) = row(5:4 + nod)
  END DO
  CLOSE(10)
  IF (meshgen == 2) CALL abaqus2sg(g_num)
  ALLOCATE(g_coord_pp(nod, ndim, nels))
  DO iel = 1, nels
    DO i = 1, nod
      g_coord_pp(i, :, iel) = coord(:, g_num(i, iel))
    END DO
  END DO

  !---------------------------------------------------------------- .bnd
  ALLOCATE(rest(nodof + 1, nr))
  IF (nr > 0) THEN
    OPEN(10, FILE=argv(1:nlen)//'.bnd', STATUS='OLD', ACTION='READ')
    DO i = 1, nr
      READ(10, *) rest(:, i)
    END DO
    CLOSE(10)
  END IF

  !------------------------------------------------ steering and equations
  ! Equations numbered by ascending node, then freedom, restrained (0)
  ! freedoms skipped: the numbering REARRANGE + FIND_G3 give
  ! (src/modules/shared/new_library.f90:3059-3212).
  ALLOCATE(nf(nodof, nn))
  nf = 1
  DO i = 1, nr
    nf(:, rest(1, i)) = rest(2:, i)
  END DO
  neq = 0
  DO i = 1, nn
    DO j = 1, nodof
      IF (nf(j, i) /= 0) THEN
        neq = neq + 1
        nf(j, i) = neq
      END IF
    END DO
  END DO
  ALLOCATE(g_g_pp(ntot, nels))
  DO iel = 1, nels
    DO i = 1, nod
      g_g_pp((i - 1) * nodof + 1:i * nodof, iel) = nf(:, g_num(i, iel))
    END DO
  END DO

  !------------------------------------------------ device setup, assembly
  CALL pf_init(0)
  CALL pf_mesh(g_coord_pp)
  CALL pf_steer(g_g_pp, neq)
  IF (mode == 'mixed') THEN
    CALL pf_assemble_elastic(e, v, PF_STORE_FP32)
  ELSE
    CALL pf_assemble_elastic(e, v, PF_STORE_FP64)
  END IF
  ALLOCATE(diag(neq), r(neq), x(neq))
  CALL pf_diag(diag)

  OPEN(11, FILE=argv(1:nlen)//'.pf.res', STATUS='REPLACE', ACTION='WRITE')
  WRITE(11, '(A,I7,A)') "This job ran on ", 1, " processes"
  WRITE(11, '(A,3(I12,A))') "There are ", nn, " nodes", nr, " restrained and ", neq, " equations"

  !---------------------------------------------------------------- loads
  ! LOAD (src/modules/mpi/loading.f90:36-142): r(eq) = val, assignment.
  r = 0.0_iwp
  IF (loaded_nodes > 0) THEN
    ALLOCATE(node(loaded_nodes), val(ndim, loaded_nodes))
    OPEN(10, FILE=argv(1:nlen)//'.lds', STATUS='OLD', ACTION='READ')
    DO i = 1, loaded_nodes
      READ(10, *) node(i), val(:, i)
    END DO
    CLOSE(10)
    DO i = 1, loaded_nodes
      DO j = 1, nodof
        IF (nf(j, node(i)) /= 0) r(nf(j, node(i))) = val(j, i)
      END DO
    END DO
  END IF
  q = SUM(r)
  WRITE(11, '(A,E12.4)') "The total load is:", q

  IF (mode == 'hostseq' .OR. mode == 'fusedseq') CALL pf_dot_order(PF_DOT_SEQ)
  CALL SYSTEM_CLOCK(t0, rate)
  IF (mode == 'host' .OR. mode == 'hostseq') THEN
    !---------------------------- p121.f90:86-104 on host arrays, as written
    ALLOCATE(ph_pp(neq), rh_pp(neq), xh_pp(neq), xnewh_pp(neq), uh_pp(neq), dh_pp(neq), dph_pp(neq))
    ALLOCATE(pmulh_pp(ntot, nels), utemph_pp(ntot, nels), storkm_pp(ntot, ntot, nels))
    CALL pf_storkm_get(storkm_pp)
    rh_pp = r; dph_pp = diag; xnewh_pp = 0.0_iwp
    dph_pp = 1._iwp / dph_pp
    dh_pp = dph_pp * rh_pp; ph_pp = dh_pp; xh_pp = 0.0_iwp
    iters = 0
    iterations: DO
      iters = iters + 1; uh_pp = 0.0_iwp; pmulh_pp = 0.0_iwp; utemph_pp = 0.0_iwp
      CALL gather(ph_pp, pmulh_pp)
      elements_3: DO iel = 1, nels
        utemph_pp(:, iel) = pf_matvec(storkm_pp(:, :, iel), pmulh_pp(:, iel))
      END DO elements_3; CALL scatter(uh_pp, utemph_pp)
      up = DOT_PRODUCT_P(rh_pp, dh_pp); alpha = up / DOT_PRODUCT_P(ph_pp, uh_pp)
      xnewh_pp = xh_pp + ph_pp * alpha; rh_pp = rh_pp - uh_pp * alpha
      dh_pp = dph_pp * rh_pp; beta = DOT_PRODUCT_P(rh_pp, dh_pp) / up
      ph_pp = dh_pp + ph_pp * beta; CALL checon_par(xnewh_pp, tol, converged, xh_pp)
      IF (converged .OR. iters == limit) EXIT
    END DO iterations
    x = xnewh_pp
    DEALLOCATE(ph_pp, rh_pp, xh_pp, xnewh_pp, uh_pp, dh_pp, dph_pp, pmulh_pp, utemph_pp, storkm_pp)
  ELSE IF (mode == 'dropin') THEN
    !------------------------------------------- p121.f90:86-104 verbatim flow
    CALL pf_alloc(p_pp, neq); CALL pf_alloc(r_pp, neq); CALL pf_alloc(x_pp, neq)
    CALL pf_alloc(xnew_pp, neq); CALL pf_alloc(u_pp, neq); CALL pf_alloc(d_pp, neq)
    CALL pf_alloc(diag_precon_pp, neq)
    CALL pf_alloc(pmul_pp, ntot * nels); CALL pf_alloc(utemp_pp, ntot * nels)
    CALL pf_put(diag_precon_pp, diag)
    CALL pf_put(r_pp, r)
    CALL vec_fill(xnew_pp, 0.0_iwp)
    CALL vec_recip(diag_precon_pp, diag_precon_pp)
    CALL vec_mul(d_pp, diag_precon_pp, r_pp)
    CALL vec_waxpy(p_pp, d_pp, d_pp, 0.0_iwp)   ! p = d (d + d*0 is d exactly)
    CALL vec_fill(x_pp, 0.0_iwp)
    iters = 0
    DO
      iters = iters + 1
      CALL vec_fill(u_pp, 0.0_iwp)
      CALL gather(p_pp, pmul_pp)
      CALL elem_matvec(pmul_pp, utemp_pp)
      CALL scatter(u_pp, utemp_pp)
      up = dot_product_p(r_pp, d_pp)
      alpha = up / dot_product_p(p_pp, u_pp)
      CALL vec_waxpy(xnew_pp, x_pp, p_pp, alpha)
      CALL vec_waxpy(r_pp, r_pp, u_pp, -alpha)
      CALL vec_mul(d_pp, diag_precon_pp, r_pp)
      beta = dot_product_p(r_pp, d_pp) / up
      CALL vec_waxpy(p_pp, d_pp, p_pp, beta)
      CALL checon_par(xnew_pp, tol, converged, x_pp)
      IF (converged .OR. iters == limit) EXIT
    END DO
    CALL pf_get(xnew_pp, x)
  ELSE IF (mode == 'fused' .OR. mode == 'fusedseq') THEN
    CALL pf_pcg(r, x, limit, tol, PF_STOP_CHECON, PF_PREC_FP64, iters, converged)
  ELSE IF (mode == 'km') THEN
    ! PCG_KM with element 1's matrix for every element (a uniform mesh and
    ! material, as PCG_KM assumes) and the inverted assembled diagonal.
    ALLOCATE(storkm(ntot, ntot, nels), timest(20))
    CALL pf_storkm_get(storkm)
    diag = 1.0_iwp / diag
    x = 0.0_iwp
    CALL pcg_km(diag, storkm(:, :, 1), limit, nels, r, timest, tol, x, iters)
    converged = iters < limit
    DEALLOCATE(storkm, timest)
  ELSE
    CALL pf_pcg(r, x, 100000, rtol, PF_STOP_RELRES, PF_PREC_MIXED, iters, converged)
  END IF
  CALL SYSTEM_CLOCK(t1)

  WRITE(11, '(A,I6)') "The number of iterations to convergence was ", iters
  WRITE(11, '(A,F10.4)') "Time to solve equations was  :", REAL(t1 - t0, iwp) / REAL(rate, iwp)
  WRITE(11, '(A,E12.4)') "The central nodal displacement is :", x(1)

  OPEN(12, FILE=argv(1:nlen)//'.pf.x', STATUS='REPLACE', ACTION='WRITE', ACCESS='STREAM', &
       FORM='UNFORMATTED', IOSTAT=ios)
  WRITE(12) x
  CLOSE(12)

  !------------------------------- stresses at the centroid (p121.f90:112-123)
  ALLOCATE(sigma(6, nels))
  CALL pf_stress(x, e, v, sigma)
  WRITE(11, '(A)') "The Centroid point stresses for element 1 are"
  WRITE(11, '(A,I5)') "Point ", 1
  WRITE(11, '(6E12.4)') sigma(:, 1)
  DEALLOCATE(sigma)

  !------------------------------------ displacements (p121.f90:125-137)
  ! eld_pp = GATHER(xnew_pp) with xnew(0) = 0, averaged over the elements of
  ! every node and written one component after the other.
  ALLOCATE(eld(ntot, nels))
  DO iel = 1, nels
    DO i = 1, ntot
      eld(i, iel) = 0.0_iwp
      IF (g_g_pp(i, iel) /= 0) eld(i, iel) = x(g_g_pp(i, iel))
    END DO
  END DO
  CALL calc_nodes_pp(nn, 1, 1, node_end, node_start, nodes_pp)
  OPEN(12, FILE=argv(1:nlen)//'.pf.ensi.DISPL-000001', STATUS='REPLACE', ACTION='WRITE')
  WRITE(12, '(A)') "Alya Ensight Gold --- Vector per-node variable file"
  WRITE(12, '(A/A/A)') "part", "     1", "coordinates"
  ALLOCATE(disp(nodes_pp * ndim), temp(nodes_pp))
  CALL scatter_nodes(1, nn, nels, g_num, nod, ndim, nodes_pp, node_start, node_end, eld, disp, 1)
  DO i = 1, ndim
    DO j = 1, nodes_pp
      temp(j) = disp(i + ndim * (j - 1))
    END DO
    CALL dismsh_ensi_p(12, 1, nodes_pp, 1, 1, 1, temp)
  END DO
  CLOSE(12)
  DEALLOCATE(eld, disp, temp)
  CALL SYSTEM_CLOCK(ta)
  WRITE(11, '(A,F10.4)') "This analysis took  :", REAL(ta - tst, iwp) / REAL(rate, iwp)
  CLOSE(11)
  WRITE(*, '(A,A,A,I0,A,L1)') 'pf_p121 ', TRIM(mode), ': iterations ', iters, ' converged ', converged

  IF (mode == 'dropin') THEN
    CALL pf_release(p_pp); CALL pf_release(r_pp); CALL pf_release(x_pp)
    CALL pf_release(xnew_pp); CALL pf_release(u_pp); CALL pf_release(d_pp)
    CALL pf_release(diag_precon_pp); CALL pf_release(pmul_pp); CALL pf_release(utemp_pp)
  END IF
  CALL pf_finalize()

CONTAINS

  !> ABAQUS2SG (src/modules/shared/new_library.f90:3578-3619): Abaqus node
  !> order of each element to Smith & Griffiths order.
  SUBROUTINE abaqus2sg(g)
    INTEGER, INTENT(INOUT) :: g(:, :)
    INTEGER, PARAMETER :: perm20(20) = [4, 12, 1, 9, 2, 10, 3, 11, 20, 17, 18, 19, 8, 16, 5, 13, 6, 14, 7, 15]
    INTEGER, PARAMETER :: perm8(8) = [1, 5, 6, 2, 4, 8, 7, 3]
    INTEGER :: ie
    INTEGER :: tmp(SIZE(g, 1))
    DO ie = 1, SIZE(g, 2)
      tmp = g(:, ie)
      IF (SIZE(g, 1) == 20) THEN
        g(:, ie) = tmp(perm20)
      ELSE
        g(:, ie) = tmp(perm8)
      END IF
    END DO
  END SUBROUTINE abaqus2sg

END PROGRAM pf_p121
