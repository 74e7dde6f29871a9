!------------------------------------------------------------------------------
! pf_p123: the p123 program flow (src/programs/5th_ed/p123/p123.f90) on one
! MI355X through the pf_gpu module -- three-dimensional steady conduction
! (Laplace's equation) on 8-node bricks, element matrices integrated on the
! device, fixed freedoms by penalty, Jacobi-preconditioned EBE-PCG.
!
!   pf_p123 <deck> [fused|mixed] [relres-tol]
!
! fused (default): the reference's loop and stopping rule (checon_par at the
! deck's tol) as one device call; mixed: element matrices stored in FP32,
! FP64 vectors and reliable updates, stopped at relres-tol (default 1e-10).
! Reads <deck>.dat/.d (and .bnd, .fix, .lds when the deck has restraints,
! fixed or loaded freedoms) in the formats READ_P123, READ_G_NUM_PP,
! READ_G_COORD_PP, READ_REST, READ_FIXED and READ_LOADS read
! (src/modules/mpi/input.f90:3620-3806, 935-1080, 288-445, 2570-2632,
! 2483-2564, 2350-2411) and writes, as p123 does, <deck>.pf.res (job data,
! iterations, total load, the potentials of freedoms nres..nres+3) and
! <deck>.pf.ensi.NDPTL-000001 (nodal potentials through SCATTER_NODES and
! DISMSH_ENSI_P); <deck>.pf.x holds xnew (neq raw REAL(8) values).
!------------------------------------------------------------------------------
PROGRAM pf_p123
  USE pf_gpu
  IMPLICIT NONE
  INTEGER, PARAMETER :: nodof = 1, ndim = 3
  REAL(iwp), PARAMETER :: penalty = 1.0E20_iwp
  INTEGER :: nels, nn, nr, nip, nod, loaded_freedoms, fixed_freedoms, meshgen, partitioner
  INTEGER :: limit, nres, ntot, neq, iel, i, k, m, iters, nlen, mlen, ios
  INTEGER :: node_start, node_end, nodes_pp, store_mode, prec, stop_mode
  REAL(iwp) :: kx, ky, kz, tol, q, xyz(3)
  LOGICAL :: converged
  CHARACTER(LEN=256) :: argv, mode, arg3, line
  CHARACTER(LEN=15) :: element
  INTEGER, ALLOCATABLE :: g_num(:, :), rest(:, :), nf(:), g_g_pp(:, :), row(:)
  INTEGER, ALLOCATABLE :: fix_node(:), fix_sense(:), no_f(:), load_no(:)
  REAL(iwp), ALLOCATABLE :: coord(:, :), g_coord_pp(:, :, :), val_f(:), store(:), load_val(:)
  REAL(iwp), ALLOCATABLE :: r(:), xnew(:), eld(:, :), ptl(:)
  INTEGER(8) :: t0, t1, t2, t3, rate

  CALL SYSTEM_CLOCK(t0, rate)
  CALL GET_COMMAND_ARGUMENT(1, argv, nlen)
  IF (nlen == 0) THEN
    WRITE(*, '(A)') 'usage: pf_p123 <deck> [fused|mixed] [relres-tol]'
    STOP 2
  END IF
  mode = 'fused'
  CALL GET_COMMAND_ARGUMENT(2, mode, mlen)
  IF (mlen == 0) mode = 'fused'
  IF (TRIM(mode) /= 'fused' .AND. TRIM(mode) /= 'mixed') THEN
    WRITE(*, '(A,A)') 'pf_p123: unknown mode ', TRIM(mode)
    STOP 2
  END IF

  !---------------------------------------------------------------- .dat
  ! READ_P123: element, meshgen, partitioner, nels, nn, nr, nip, nod,
  ! loaded_freedoms, fixed_freedoms, kx, ky, kz, tol, limit, nres.
  OPEN(10, FILE=argv(1:nlen)//'.dat', STATUS='OLD', ACTION='READ')
  READ(10, *) element, meshgen, partitioner, nels, nn, nr, nip, nod, loaded_freedoms, fixed_freedoms, &
              kx, ky, kz, tol, limit, nres
  CLOSE(10)
  IF (nod /= 8) THEN
    WRITE(*, '(A,I0)') 'pf_p123: the 8-node brick only, nod = ', nod
    STOP 2
  END IF
  ntot = nod * nodof
  stop_mode = PF_STOP_CHECON
  prec = PF_PREC_FP64
  store_mode = PF_STORE_FP64
  IF (TRIM(mode) == 'mixed') THEN
    stop_mode = PF_STOP_RELRES
    prec = PF_PREC_MIXED
    store_mode = PF_STORE_FP32
    tol = 1.0E-10_iwp
    CALL GET_COMMAND_ARGUMENT(3, arg3, mlen)
    IF (mlen > 0) READ(arg3, *) tol
  END IF

  !---------------------------------------------------------------- .d
  ALLOCATE(coord(ndim, nn), g_num(nod, nels), row(nod + 5))
  OPEN(10, FILE=argv(1:nlen)//'.d', STATUS='OLD', ACTION='READ')
  DO
    READ(10, '(A)') line
    IF (INDEX(ADJUSTL(line), '*NODES') == 1) EXIT
  END DO
  DO i = 1, nn
    READ(10, *) k, xyz
    coord(:, k) = xyz
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
  IF (meshgen == 2) CALL abaqus2sg8(g_num)
  ALLOCATE(g_coord_pp(nod, ndim, nels))
  DO iel = 1, nels
    DO i = 1, nod
      g_coord_pp(i, :, iel) = coord(:, g_num(i, iel))
    END DO
  END DO

  !------------------------------------------------------- .bnd and steering
  ! nr > 0: REARRANGE_2 + FIND_G4, equations by ascending node over the free
  ! nodes; nr = 0: g_g_pp = g_num_pp (p123.f90:49-55).
  ALLOCATE(rest(nodof + 1, MAX(nr, 1)), nf(nn))
  nf = 1
  IF (nr > 0) THEN
    OPEN(10, FILE=argv(1:nlen)//'.bnd', STATUS='OLD', ACTION='READ')
    DO i = 1, nr
      READ(10, *) rest(:, i)
      nf(rest(1, i)) = rest(2, i)
    END DO
    CLOSE(10)
  END IF
  neq = 0
  DO i = 1, nn
    IF (nf(i) /= 0) THEN
      neq = neq + 1
      nf(i) = neq
    END IF
  END DO
  ALLOCATE(g_g_pp(ntot, nels))
  DO iel = 1, nels
    g_g_pp(:, iel) = nf(g_num(:, iel))
  END DO

  OPEN(11, FILE=argv(1:nlen)//'.pf.res', STATUS='REPLACE', ACTION='WRITE')
  WRITE(11, '(A,I5,A)') "This job ran on ", 1, "  processes"
  WRITE(11, '(A,3(I12,A))') "There are ", nn, " nodes", nr, " restrained and   ", neq, " equations"
  CALL SYSTEM_CLOCK(t1)
  WRITE(11, '(A,F10.4)') "Time after setup is ", REAL(t1 - t0, iwp) / REAL(rate, iwp)

  !------------------------------ element matrices and diagonal on the GPU
  ! p123.f90:70-92: storkc = kcx kx + kcy ky + kcz kz per element, the
  ! diagonal scattered (the library keeps it for the preconditioner).
  CALL pf_init(0)
  CALL pf_mesh(g_coord_pp)
  CALL pf_steer(g_g_pp, neq)
  CALL pf_assemble_laplace(kx, ky, kz, store_mode)

  !------------------------------------------------- fixed freedoms (.fix)
  ! READ_FIXED node, sense, value; FIND_NO2: the equation of the node's
  ! freedom; penalty on the diagonal, store = the new diagonal
  ! (p123.f90:94-105, 120-125).
  ALLOCATE(no_f(fixed_freedoms), store(fixed_freedoms), val_f(fixed_freedoms))
  IF (fixed_freedoms > 0) THEN
    ALLOCATE(fix_node(fixed_freedoms), fix_sense(fixed_freedoms))
    OPEN(10, FILE=argv(1:nlen)//'.fix', STATUS='OLD', ACTION='READ')
    DO i = 1, fixed_freedoms
      READ(10, *) fix_node(i), fix_sense(i), val_f(i)
      no_f(i) = nf(fix_node(i))
    END DO
    CLOSE(10)
    CALL pf_fixed(no_f, penalty, store)
  END IF

  !------------------------------------------------ loaded freedoms (.lds)
  ! READ_LOADS node, value; p123 takes the numbers as equation numbers
  ! (p123.f90:108-118), then q = SUM_P(r).
  ALLOCATE(r(neq), xnew(neq))
  r = 0.0_iwp
  q = 0.0_iwp
  IF (loaded_freedoms > 0) THEN
    ALLOCATE(load_no(loaded_freedoms), load_val(loaded_freedoms))
    OPEN(10, FILE=argv(1:nlen)//'.lds', STATUS='OLD', ACTION='READ')
    DO i = 1, loaded_freedoms
      READ(10, *) load_no(i), load_val(i)
      r(load_no(i)) = load_val(i)
    END DO
    CLOSE(10)
    q = SUM(r)
  END IF
  ! the fixed rows of the starting residual (p123.f90:127-132)
  DO i = 1, fixed_freedoms
    r(no_f(i)) = store(i) * val_f(i)
  END DO

  !-------------------------------------------------------------- the PCG
  CALL SYSTEM_CLOCK(t2)
  CALL pf_pcg(r, xnew, limit, tol, stop_mode, prec, iters, converged)
  CALL SYSTEM_CLOCK(t3)
  WRITE(11, '(A,I5)') "The number of iterations to convergence was ", iters
  WRITE(11, '(A,E12.4)') "The total load is ", q
  WRITE(11, '(A)') "The potentials are:"
  WRITE(11, '(A)') " Freedom       Potential"
  DO i = 1, 4
    IF (nres + i - 1 >= 1 .AND. nres + i - 1 <= neq) WRITE(11, '(I8,A,E12.4)') nres + i - 1, "     ", xnew(nres + i - 1)
  END DO
  WRITE(11, '(A,F10.4)') "Integration, preconditioning and loading took ", REAL(t2 - t1, iwp) / REAL(rate, iwp)
  WRITE(11, '(A,F10.4)') "Time spent in the solver was ", REAL(t3 - t2, iwp) / REAL(rate, iwp)

  !------------------------------------- nodal potentials (p123.f90:164-176)
  OPEN(12, FILE=argv(1:nlen)//'.pf.ensi.NDPTL-000001', STATUS='REPLACE', ACTION='WRITE')
  WRITE(12, '(A)') "Alya Ensight Gold --- Scalar per-node variable file"
  WRITE(12, '(A/A/A)') "part", "    1", "coordinates"
  CALL calc_nodes_pp(nn, 1, 1, node_end, node_start, nodes_pp)
  ALLOCATE(eld(ntot, nels), ptl(nodes_pp * nodof))
  DO iel = 1, nels
    DO m = 1, ntot
      IF (g_g_pp(m, iel) > 0) THEN
        eld(m, iel) = xnew(g_g_pp(m, iel))
      ELSE
        eld(m, iel) = 0.0_iwp
      END IF
    END DO
  END DO
  ptl = 0.0_iwp
  CALL scatter_nodes(1, nn, nels, g_num, nod, nodof, nodes_pp, node_start, node_end, eld, ptl, 1)
  CALL dismsh_ensi_p(12, 1, nodes_pp, 1, 1, 1, ptl)
  CLOSE(12)
  CALL SYSTEM_CLOCK(t3)
  WRITE(11, '(A,F10.4)') "This analysis took ", REAL(t3 - t0, iwp) / REAL(rate, iwp)
  CLOSE(11)

  OPEN(13, FILE=argv(1:nlen)//'.pf.x', STATUS='REPLACE', ACTION='WRITE', ACCESS='STREAM', &
       FORM='UNFORMATTED', IOSTAT=ios)
  WRITE(13) xnew
  CLOSE(13)
  WRITE(*, '(A,I0,A,L1)') 'pf_p123: ', iters, ' iterations, converged ', converged
  CALL pf_finalize()

CONTAINS

  !> ABAQUS2SG for the 8-node brick (src/modules/shared/new_library.f90:
  !> 3609-3619).
  SUBROUTINE abaqus2sg8(g)
    INTEGER, INTENT(INOUT) :: g(:, :)
    INTEGER, PARAMETER :: perm8(8) = [1, 5, 6, 2, 4, 8, 7, 3]
    INTEGER :: ie
    INTEGER :: tmp(8)
    DO ie = 1, SIZE(g, 2)
      tmp = g(:, ie)
      g(:, ie) = tmp(perm8)
    END DO
  END SUBROUTINE abaqus2sg8

END PROGRAM pf_p123
