!------------------------------------------------------------------------------
! pf_p122: the p122 program flow (src/programs/5th_ed/p122/p122.f90) on one
! MI355X through the pf_gpu module -- three-dimensional strain of an
! elastic-viscoplastic (Mohr-Coulomb) solid of 8-node bricks, load
! increments, plastic iterations, each a warm-started Jacobi PCG (r formed
! with GATHER / element products / SCATTER on device vectors, the CG loop as
! one device call doing p122's statements in their order), and the
! Gauss-point stress update on the device.
!
!   pf_p122 <deck>
!
! Reads <deck>.dat/.d/.bnd (and .fix, .lds when the deck has fixed freedoms
! or loaded nodes) in the formats READ_P122, READ_QINC, READ_G_NUM_PP,
! READ_G_COORD_PP, READ_REST, READ_FIXED and READ_LOADS read
! (src/modules/mpi/input.f90:3402-3611, 935-1080, 288-445, 2570-2632,
! 2483-2564, 2350-2411) and writes, as p122 does, <deck>.pf.res and
! <deck>.pf.ensi.DISPL-<increment>; <deck>.pf.x holds the total
! displacements (neq raw REAL(8) values).
!------------------------------------------------------------------------------
PROGRAM pf_p122
  USE pf_gpu
  IMPLICIT NONE
  INTEGER, PARAMETER :: nodof = 3, ndim = 3, nst = 6
  REAL(iwp), PARAMETER :: penalty = 1.0E20_iwp
  INTEGER :: nels, nn, nr, nip, nod, fixed_freedoms, loaded_nodes, meshgen, partitioner, incs, plasits, cjits
  INTEGER :: ntot, neq, iel, i, j, k, iy, plasiters, cjiters, cjtot, nlen, ios
  INTEGER :: node_start, node_end, nodes_pp
  REAL(iwp) :: phi, c, psi, e, v, plastol, cjtol, dt, snph, pi, xyz(3)
  LOGICAL :: plastic_converged, final
  CHARACTER(LEN=256) :: argv, line
  CHARACTER(LEN=15) :: element
  CHARACTER(LEN=6) :: ch
  INTEGER, ALLOCATABLE :: g_num(:, :), rest(:, :), nf(:, :), g_g_pp(:, :), row(:), nodef(:), sense(:), no_f(:)
  INTEGER, ALLOCATABLE :: node(:)
  REAL(iwp), ALLOCATABLE :: coord(:, :), g_coord_pp(:, :, :), qinc(:), valf(:), store(:), val(:, :)
  REAL(iwp), ALLOCATABLE :: h(:), ld0(:), totd(:), tensor(:), eld(:, :), disp(:), temp(:)
  TYPE(pf_dvec) :: loads_pp, bdylds_pp, oldis_pp, x_pp, xnew_pp, r_pp
  TYPE(pf_dvec) :: ld0_pp, totd_pp, pmul_pp, utemp_pp, evpt_pp, tensor_pp
  INTEGER(8) :: t0, t1, rate

  CALL SYSTEM_CLOCK(t0, rate)
  CALL GET_COMMAND_ARGUMENT(1, argv, nlen)
  IF (nlen == 0) THEN
    WRITE(*, '(A)') 'usage: pf_p122 <deck>'
    STOP 2
  END IF

  !---------------------------------------------------------------- .dat
  ! READ_P122 (list-directed), then READ_QINC's incs values.
  OPEN(10, FILE=argv(1:nlen)//'.dat', STATUS='OLD', ACTION='READ')
  READ(10, *) element, meshgen, partitioner, nels, nn, nr, nip, nod, fixed_freedoms, loaded_nodes, phi, c, &
              psi, e, v, incs, plasits, cjits, plastol, cjtol
  ALLOCATE(qinc(incs))
  READ(10, *) qinc
  CLOSE(10)
  IF (nod /= 8) THEN
    WRITE(*, '(A,I0)') 'pf_p122: the 8-node brick only, nod = ', nod
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
  ! REARRANGE + FIND_G: equations by ascending node, then freedom.
  ALLOCATE(rest(nodof + 1, MAX(nr, 1)), nf(nodof, nn))
  nf = 1
  IF (nr > 0) THEN
    OPEN(10, FILE=argv(1:nlen)//'.bnd', STATUS='OLD', ACTION='READ')
    DO i = 1, nr
      READ(10, *) rest(:, i)
      nf(:, rest(1, i)) = rest(2:, i)
    END DO
    CLOSE(10)
  END IF
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

  OPEN(11, FILE=argv(1:nlen)//'.pf.res', STATUS='REPLACE', ACTION='WRITE')
  WRITE(11, '(A,I5,A)') "This job ran on ", 1, "  processes"
  WRITE(11, '(A,3(I7,A))') "There are ", nn, " nodes", nr, " restrained and   ", neq, " equations"
  CALL SYSTEM_CLOCK(t1)
  WRITE(11, '(A,F10.4)') "Time after setup is:", REAL(t1 - t0, iwp) / REAL(rate, iwp)

  !------------------------------------------------------ elastic matrices
  pi = ACOS(-1.0_iwp)
  snph = SIN(phi * pi / 180.0_iwp)
  dt = 4.0_iwp * (1.0_iwp + v) * (1.0_iwp - 2.0_iwp * v) / (e * (1.0_iwp - 2.0_iwp * v + snph * snph))
  WRITE(11, '(A,E12.4)') "The critical timestep is   ", dt
  CALL pf_init(0)
  CALL pf_mesh(g_coord_pp)
  CALL pf_steer(g_g_pp, neq)
  CALL pf_assemble_elastic(e, v, PF_STORE_FP64)
  ALLOCATE(h(neq), ld0(neq), totd(neq))

  !---------------------------------------------------------------- loads
  ld0 = 0.0_iwp
  IF (loaded_nodes > 0) THEN
    ALLOCATE(node(loaded_nodes), val(ndim, loaded_nodes))
    OPEN(10, FILE=argv(1:nlen)//'.lds', STATUS='OLD', ACTION='READ')
    DO i = 1, loaded_nodes
      READ(10, *) node(i), val(:, i)
    END DO
    CLOSE(10)
    DO i = 1, loaded_nodes
      DO j = 1, nodof
        IF (nf(j, node(i)) /= 0) ld0(nf(j, node(i))) = val(j, i)
      END DO
    END DO
  END IF

  !------------------------------------------------------- fixed freedoms
  ! READ_FIXED node, sense, value; FIND_NO: equation of freedom `sense` of
  ! the node; penalty on the diagonal, store = the new diagonal (the
  ! preconditioner is 1/diagonal, p122.f90:112-118).
  ALLOCATE(nodef(fixed_freedoms), sense(fixed_freedoms), valf(fixed_freedoms), no_f(fixed_freedoms), &
           store(fixed_freedoms))
  IF (fixed_freedoms > 0) THEN
    OPEN(10, FILE=argv(1:nlen)//'.fix', STATUS='OLD', ACTION='READ')
    DO i = 1, fixed_freedoms
      READ(10, *) nodef(i), sense(i), valf(i)
      no_f(i) = nf(sense(i), nodef(i))
    END DO
    CLOSE(10)
    CALL pf_fixed(no_f, penalty, store)
  END IF

  !------------------------------------------------------ device vectors
  CALL pf_alloc(loads_pp, neq); CALL pf_alloc(bdylds_pp, neq); CALL pf_alloc(oldis_pp, neq)
  CALL pf_alloc(x_pp, neq); CALL pf_alloc(xnew_pp, neq); CALL pf_alloc(r_pp, neq)
  CALL pf_alloc(ld0_pp, neq); CALL pf_alloc(totd_pp, neq)
  CALL pf_alloc(pmul_pp, ntot * nels); CALL pf_alloc(utemp_pp, ntot * nels)
  CALL pf_alloc(evpt_pp, nst * nip * nels); CALL pf_alloc(tensor_pp, nst * nip * nels)
  CALL pf_put(ld0_pp, ld0)
  CALL vec_fill(oldis_pp, 0.0_iwp); CALL vec_fill(x_pp, 0.0_iwp); CALL vec_fill(totd_pp, 0.0_iwp)
  CALL vec_fill(tensor_pp, 0.0_iwp)
  ALLOCATE(tensor(nst * nip * nels), eld(ntot, nels))
  CALL calc_nodes_pp(nn, 1, 1, node_end, node_start, nodes_pp)
  ALLOCATE(disp(nodes_pp * ndim), temp(nodes_pp))

  !------------------------------------------------------ load increments
  DO iy = 1, incs
    plasiters = 0
    CALL vec_fill(bdylds_pp, 0.0_iwp)
    CALL vec_fill(evpt_pp, 0.0_iwp)
    cjtot = 0
    WRITE(11, '(/,A,I5)') "Load Increment   ", iy
    DO
      plasiters = plasiters + 1
      IF (plasiters == 1) THEN
        IF (loaded_nodes > 0) THEN
          CALL vec_waxpy(loads_pp, bdylds_pp, ld0_pp, qinc(iy))
        ELSE
          h = 0.0_iwp
          DO i = 1, fixed_freedoms
            h(no_f(i)) = store(i) * valf(i) * qinc(iy)
          END DO
          CALL pf_put(loads_pp, h)
        END IF
      ELSE
        IF (loaded_nodes > 0) THEN
          CALL vec_waxpy(loads_pp, bdylds_pp, ld0_pp, qinc(iy))
        ELSE
          CALL vec_waxpy(loads_pp, bdylds_pp, bdylds_pp, 0.0_iwp)
        END IF
        CALL pf_fixed_rows(loads_pp, loads_pp, .FALSE.)
      END IF
      ! r = loads - K x (x warm from the previous solve)
      CALL vec_fill(r_pp, 0.0_iwp)
      CALL gather(x_pp, pmul_pp)
      CALL elem_matvec(pmul_pp, utemp_pp)
      CALL scatter(r_pp, utemp_pp)
      CALL vec_waxpy(r_pp, loads_pp, r_pp, -1.0_iwp)
      ! d = diag_precon r, p = d and the conjugate_gradients loop
      ! (p122.f90:147-162: u = K p with the fixed rows, alpha, xnew, r, d,
      ! beta, p, checon_par) on the device as one call, statement for
      ! statement the same arithmetic
      CALL pf_cg_warm(r_pp, x_pp, xnew_pp, plasiters == 1, cjits, cjtol, cjiters)
      cjtot = cjtot + cjiters
      CALL vec_waxpy(loads_pp, xnew_pp, xnew_pp, 0.0_iwp)
      CALL checon_par(loads_pp, plastol, plastic_converged, oldis_pp)
      IF (plasiters == 1) plastic_converged = .FALSE.
      final = plastic_converged .OR. plasiters == plasits
      IF (final) CALL vec_fill(bdylds_pp, 0.0_iwp)
      CALL pf_plastic_update(loads_pp, evpt_pp, tensor_pp, utemp_pp, e, v, phi, c, psi, dt, final)
      CALL scatter(bdylds_pp, utemp_pp)
      IF (final) EXIT
    END DO
    CALL vec_waxpy(totd_pp, totd_pp, loads_pp, 1.0_iwp)
    CALL pf_get(totd_pp, totd)
    CALL pf_get(tensor_pp, tensor)
    WRITE(11, '(A,E12.4)') "The displacement is  ", totd(1)
    WRITE(11, '(A)') "  sigma z    sigma x     sigma y"
    WRITE(11, '(3E12.4)') tensor(3), tensor(1), tensor(2)
    WRITE(11, '(A,I12)') "The total number of cj iterations was  ", cjtot
    WRITE(11, '(A,I12)') "The number of plastic iterations was  ", plasiters
    WRITE(11, '(A,F11.2)') "cj iterations per plastic iteration were ", REAL(cjtot) / REAL(plasiters)
    !---------------------------- displacements for ParaView (p122.f90:213-225)
    WRITE(ch, '(I6.6)') iy
    OPEN(12, FILE=argv(1:nlen)//'.pf.ensi.DISPL-'//ch, STATUS='REPLACE', ACTION='WRITE')
    WRITE(12, '(A)') "Alya Ensight Gold --- Vector per-node variable file"
    WRITE(12, '(A/A/A)') "part", "    1", "coordinates"
    DO iel = 1, nels
      DO i = 1, ntot
        eld(i, iel) = 0.0_iwp
        IF (g_g_pp(i, iel) /= 0) eld(i, iel) = totd(g_g_pp(i, iel))
      END DO
    END DO
    CALL scatter_nodes(1, nn, nels, g_num, nod, ndim, nodes_pp, node_start, node_end, eld, disp, 1)
    DO i = 1, ndim
      DO j = 1, nodes_pp
        temp(j) = disp(i + ndim * (j - 1))
      END DO
      CALL dismsh_ensi_p(12, 1, nodes_pp, 1, 1, 1, temp)
    END DO
    CLOSE(12)
    IF (plasiters == plasits) EXIT
  END DO
  CALL SYSTEM_CLOCK(t1)
  WRITE(11, '(A,F10.4)') "This analysis took: ", REAL(t1 - t0, iwp) / REAL(rate, iwp)
  CLOSE(11)
  OPEN(12, FILE=argv(1:nlen)//'.pf.x', STATUS='REPLACE', ACTION='WRITE', ACCESS='STREAM', &
       FORM='UNFORMATTED', IOSTAT=ios)
  WRITE(12) totd
  CLOSE(12)
  WRITE(*, '(A,I0,A)') 'pf_p122: ', iy - 1, ' increments'
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

END PROGRAM pf_p122
