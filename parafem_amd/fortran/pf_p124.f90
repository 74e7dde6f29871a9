!------------------------------------------------------------------------------
! pf_p124: the p124 program flow (src/programs/5th_ed/p124/p124.f90) on one
! MI355X through the pf_gpu module -- three-dimensional transient heat
! conduction on 8-node bricks, theta-method in time, every time step solved
! by Jacobi-preconditioned EBE-PCG stopped by checon.
!
!   pf_p124 <deck>
!
! Reads <deck>.dat/.d/.bnd/.mat (and .fix, .lds when the deck has fixed or
! loaded freedoms) in the formats READ_P124, READ_ELEMENTS, READ_G_COORD_PP,
! READ_REST, READ_MATERIAL, READ_FIXED and READ_LOADS read
! (src/modules/mpi/input.f90:3997-4169, 1434-1583, 288-445, 2570-2632,
! 3067-3102, 2483-2564, 2350-2411) and writes, as p124 does,
! <deck>.pf.res (job data, the temperature of equation nres and the PCG
! iterations every npri steps), <deck>.pf.ensi.NDTTR-<step> (nodal
! temperatures through SCATTER_NODES and DISMSH_ENSI_P) and <deck>.pf.npp;
! <deck>.pf.x holds the last step's temperatures (neq raw REAL(8) values).
!------------------------------------------------------------------------------
PROGRAM pf_p124
  USE pf_gpu
  IMPLICIT NONE
  INTEGER, PARAMETER :: nodof = 1, ndim = 3, nprops = 5
  REAL(iwp), PARAMETER :: penalty = 1.0E20_iwp
  INTEGER :: nels, nn, nr, nip, nod, loaded_freedoms, fixed_freedoms, meshgen, partitioner, np_types
  INTEGER :: nstep, npri, limit, nres, ntot, neq, iel, i, j, k, l, iters, nlen, ios, tz
  INTEGER :: node_start, node_end, nodes_pp
  REAL(iwp) :: val0, dtim, theta, tol, real_time, q, xyz(3)
  LOGICAL :: converged
  CHARACTER(LEN=256) :: argv, line
  CHARACTER(LEN=15) :: element
  CHARACTER(LEN=6) :: ch
  INTEGER, ALLOCATABLE :: g_num(:, :), etype(:), rest(:, :), nf(:), g_g_pp(:, :), row(:)
  INTEGER, ALLOCATABLE :: fix_node(:), fix_sense(:), no_f(:), load_node(:)
  REAL(iwp), ALLOCATABLE :: coord(:, :), g_coord_pp(:, :, :), prop(:, :), val_f(:), store(:), load_val(:)
  REAL(iwp), ALLOCATABLE :: x(:), xnew(:), loads(:), u(:), r(:), eld(:, :), ttr(:)
  INTEGER(8) :: t0, t1, t2, t3, rate, tsolve, twrite

  CALL SYSTEM_CLOCK(t0, rate)
  CALL GET_COMMAND_ARGUMENT(1, argv, nlen)
  IF (nlen == 0) THEN
    WRITE(*, '(A)') 'usage: pf_p124 <deck>'
    STOP 2
  END IF

  !---------------------------------------------------------------- .dat
  ! READ_P124: element, meshgen, partitioner, np_types, nels, nn, nr, nip,
  ! nod, loaded_freedoms, fixed_freedoms, val0, dtim, nstep, npri, theta,
  ! tol, limit, nres (list-directed).
  OPEN(10, FILE=argv(1:nlen)//'.dat', STATUS='OLD', ACTION='READ')
  READ(10, *) element, meshgen, partitioner, np_types, nels, nn, nr, nip, nod, loaded_freedoms, &
              fixed_freedoms, val0, dtim, nstep, npri, theta, tol, limit, nres
  CLOSE(10)
  IF (nod /= 8) THEN
    WRITE(*, '(A,I0)') 'pf_p124: the 8-node brick only, nod = ', nod
    STOP 2
  END IF
  ntot = nod * nodof

  !---------------------------------------------------------------- .d
  ! READ_ELEMENTS keeps each element's type (the last column).
  ALLOCATE(coord(ndim, nn), g_num(nod, nels), etype(nels), row(nod + 5))
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
    etype(iel) = row(5 + nod)
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
  ! REARRANGE_2 + FIND_G4: equations by ascending node over the free nodes.
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

  !---------------------------------------------------------------- .mat
  ALLOCATE(prop(nprops, np_types))
  OPEN(10, FILE=argv(1:nlen)//'.mat', STATUS='OLD', ACTION='READ')
  READ(10, *)
  READ(10, *)
  DO i = 1, np_types
    READ(10, *) k, prop(:, i)
  END DO
  CLOSE(10)

  !-------------------------------------------- element matrices on the GPU
  CALL pf_init(0)
  CALL pf_mesh(g_coord_pp)
  CALL pf_steer(g_g_pp, neq)
  CALL pf_assemble_transient(theta, dtim, prop, etype, PF_STORE_FP64)

  !------------------------------------------------- fixed freedoms (.fix)
  ! READ_FIXED node, sense, value; FIND_NO2 maps them to equations; the
  ! penalty is added to the diagonal and kept in store (p124.f90:96-114).
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
  ! READ_LOADS node, value; p124 takes the node numbers as equation
  ! numbers (REINDEX of `node`, p124.f90:116-123).
  ALLOCATE(load_node(loaded_freedoms), load_val(loaded_freedoms))
  IF (loaded_freedoms > 0) THEN
    OPEN(10, FILE=argv(1:nlen)//'.lds', STATUS='OLD', ACTION='READ')
    DO i = 1, loaded_freedoms
      READ(10, *) load_node(i), load_val(i)
    END DO
    CLOSE(10)
  END IF

  OPEN(11, FILE=argv(1:nlen)//'.pf.res', STATUS='REPLACE', ACTION='WRITE')
  WRITE(11, '(A,I5,A)') "This job ran on ", 1, "  processes"
  WRITE(11, '(A,3(I12,A))') "There are ", nn, " nodes", nr, " restrained and   ", neq, " equations"
  CALL SYSTEM_CLOCK(t1)
  WRITE(11, '(A,F10.4)') "Time after setup is ", REAL(t1 - t0, iwp) / REAL(rate, iwp)

  ALLOCATE(x(neq), xnew(neq), loads(neq), u(neq), r(neq), eld(ntot, nels), ttr(nn))
  xnew = 0.0_iwp
  q = 0.0_iwp
  tsolve = 0
  twrite = 0
  CALL calc_nodes_pp(nn, 1, 1, node_end, node_start, nodes_pp)
  WRITE(11, '(A)') "  Time       Temperature  Iterations "
  OPEN(13, FILE=argv(1:nlen)//'.pf.npp', STATUS='REPLACE', ACTION='WRITE')
  WRITE(13, *) nn
  WRITE(13, *) nstep / npri
  WRITE(13, *) 1
  CLOSE(13)

  !------------------------------------------------------------- time steps
  DO j = 1, nstep
    real_time = j * dtim
    CALL SYSTEM_CLOCK(t2)
    loads = 0.0_iwp
    DO i = 1, loaded_freedoms
      loads(load_node(i)) = load_val(i) * dtim
    END DO
    IF (loaded_freedoms > 0) q = q + SUM(loads)
    IF (j /= 1) THEN
      ! u = storkb . xnew, fixed rows u(l) = store * val_f
      CALL pf_apply_raw(1, xnew, u)
      DO i = 1, fixed_freedoms
        u(no_f(i)) = store(i) * val_f(i)
      END DO
      loads = loads + u
    ELSE
      ! the initial temperatures, written as step 0, and u = storka . x0
      x = val0
      DO i = 1, fixed_freedoms
        x(no_f(i)) = val_f(i)
      END DO
      WRITE(11, '(2e12.4)') 0.0_iwp, x(nres)
      tz = 0
      CALL write_ndttr(tz, x)
      CALL pf_apply_raw(0, x, u)
      loads = loads + u
    END IF
    ! x = 0, so r = loads - storka . 0 with the fixed rows r(l) = store * val_f
    r = 0.0_iwp
    DO i = 1, fixed_freedoms
      r(no_f(i)) = store(i) * val_f(i)
    END DO
    r = loads - r
    CALL pf_pcg(r, xnew, limit, tol, PF_STOP_CHECON, PF_PREC_FP64, iters, converged)
    CALL SYSTEM_CLOCK(t3)
    tsolve = tsolve + (t3 - t2)
    IF (j / npri * npri == j) THEN
      CALL SYSTEM_CLOCK(t2)
      CALL write_ndttr(j, xnew)
      WRITE(11, '(2E12.4,I10)') real_time, xnew(nres), iters
      CALL SYSTEM_CLOCK(t3)
      twrite = twrite + (t3 - t2)
    END IF
  END DO
  CALL SYSTEM_CLOCK(t3)
  WRITE(11, '(A,F10.4)') "The solution phase took ", REAL(tsolve, iwp) / REAL(rate, iwp)
  WRITE(11, '(A,F10.4)') "Writing the output took ", REAL(twrite, iwp) / REAL(rate, iwp)
  WRITE(11, '(A,F10.4)') "This analysis took      ", REAL(t3 - t0, iwp) / REAL(rate, iwp)
  CLOSE(11)

  OPEN(12, FILE=argv(1:nlen)//'.pf.x', STATUS='REPLACE', ACTION='WRITE', ACCESS='STREAM', &
       FORM='UNFORMATTED', IOSTAT=ios)
  WRITE(12) xnew
  CLOSE(12)
  WRITE(*, '(A,I0,A)') 'pf_p124: ', nstep, ' steps'
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

  !> The nodal temperatures of step `step` as p124 writes them
  !> (p124.f90:177-186, 218-226): GATHER, SCATTER_NODES averaging,
  !> DISMSH_ENSI_P.
  SUBROUTINE write_ndttr(step, t)
    INTEGER, INTENT(IN) :: step
    REAL(iwp), INTENT(IN) :: t(:)
    INTEGER :: ie, m
    WRITE(ch, '(I6.6)') step
    OPEN(12, FILE=argv(1:nlen)//'.pf.ensi.NDTTR-'//ch, STATUS='REPLACE', ACTION='WRITE')
    WRITE(12, '(A)') "Alya Ensight Gold --- Scalar per-node variable file"
    WRITE(12, '(A/A/A)') "part", "    1", "coordinates"
    DO ie = 1, nels
      DO m = 1, ntot
        eld(m, ie) = 0.0_iwp
        IF (g_g_pp(m, ie) /= 0) eld(m, ie) = t(g_g_pp(m, ie))
      END DO
    END DO
    CALL scatter_nodes(1, nn, nels, g_num, nod, nodof, nodes_pp, node_start, node_end, eld, ttr, 1)
    CALL dismsh_ensi_p(12, 1, nodes_pp, 1, 1, 1, ttr)
    CLOSE(12)
  END SUBROUTINE write_ndttr

END PROGRAM pf_p124
