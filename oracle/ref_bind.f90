!------------------------------------------------------------------------------
! C-callable wrappers around the reference's own element routines -- TEST
! INFRASTRUCTURE ONLY.
!
! Compiled (oracle/Makefile, target `ref`) together with the reference's
! src/modules/shared/{precision,geometry,new_library}.f90, exactly as they lie
! in the reference tree, into oracle/_ref/libparafem_ref.so.  The tests use it
! to pin the oracle's restatements (SAMPLE, SHAPE_DER, SHAPE_FUN, DEEMAT,
! BEEMAT, REARRANGE + FIND_G3, ABAQUS2SG) and the mesh generators
! (GEOMETRY_20BXZ, GEOMETRY_8BXZ, CUBE_BC20) against the reference's code.
! Arrays are Fortran column-major with the reference's shapes.
!------------------------------------------------------------------------------
MODULE ref_bind
  USE, INTRINSIC :: ISO_C_BINDING
  USE precision
  USE geometry
  USE new_library
  IMPLICIT NONE
CONTAINS

  ! SAMPLE('hexahedron', s, wt)            new_library.f90:1006-1530
  SUBROUTINE ref_sample_hex(nip, s, wt) BIND(C, name="ref_sample_hex")
    INTEGER(c_int), VALUE :: nip
    REAL(c_double), INTENT(OUT) :: s(nip, 3), wt(nip)
    CALL sample('hexahedron', s, wt)
  END SUBROUTINE

  ! SHAPE_DER(der, points, i)               new_library.f90:490-912
  SUBROUTINE ref_shape_der(nod, nip, points, i, der) BIND(C, name="ref_shape_der")
    INTEGER(c_int), VALUE :: nod, nip, i
    REAL(c_double), INTENT(IN) :: points(nip, 3)
    REAL(c_double), INTENT(OUT) :: der(3, nod)
    CALL shape_der(der, points, INT(i))
  END SUBROUTINE

  ! SHAPE_FUN(fun, points, i)               new_library.f90:204-484
  SUBROUTINE ref_shape_fun(nod, nip, points, i, fun) BIND(C, name="ref_shape_fun")
    INTEGER(c_int), VALUE :: nod, nip, i
    REAL(c_double), INTENT(IN) :: points(nip, 3)
    REAL(c_double), INTENT(OUT) :: fun(nod)
    CALL shape_fun(fun, points, INT(i))
  END SUBROUTINE

  ! DEEMAT(dee, e, v), nst = 6              new_library.f90:1604-1691
  SUBROUTINE ref_deemat(e, v, dee) BIND(C, name="ref_deemat")
    REAL(c_double), VALUE :: e, v
    REAL(c_double), INTENT(OUT) :: dee(6, 6)
    dee = 0.0_iwp
    CALL deemat(dee, e, v)
  END SUBROUTINE

  ! BEEMAT(bee, deriv), ih = 6              new_library.f90:918-1000
  SUBROUTINE ref_beemat(nod, deriv, bee) BIND(C, name="ref_beemat")
    INTEGER(c_int), VALUE :: nod
    REAL(c_double), INTENT(IN) :: deriv(3, nod)
    REAL(c_double), INTENT(OUT) :: bee(6, 3 * nod)
    bee = 0.0_iwp
    CALL beemat(bee, deriv)
  END SUBROUTINE

  ! GEOMETRY_20BXZ(iel,nxe,nze,aa,bb,cc,coord,num)  geometry.f90:172-286
  SUBROUTINE ref_geometry_20bxz(iel, nxe, nze, aa, bb, cc, coord, num) BIND(C, name="ref_geometry_20bxz")
    INTEGER(c_int), VALUE :: iel, nxe, nze
    REAL(c_double), VALUE :: aa, bb, cc
    REAL(c_double), INTENT(OUT) :: coord(20, 3)
    INTEGER(c_int), INTENT(OUT) :: num(20)
    INTEGER :: n(20)
    CALL geometry_20bxz(INT(iel), INT(nxe), INT(nze), aa, bb, cc, coord, n)
    num = n
  END SUBROUTINE

  ! GEOMETRY_8BXZ(iel,nxe,nze,aa,bb,cc,coord,num)   geometry.f90:70-166
  SUBROUTINE ref_geometry_8bxz(iel, nxe, nze, aa, bb, cc, coord, num) BIND(C, name="ref_geometry_8bxz")
    INTEGER(c_int), VALUE :: iel, nxe, nze
    REAL(c_double), VALUE :: aa, bb, cc
    REAL(c_double), INTENT(OUT) :: coord(8, 3)
    INTEGER(c_int), INTENT(OUT) :: num(8)
    INTEGER :: n(8)
    CALL geometry_8bxz(INT(iel), INT(nxe), INT(nze), aa, bb, cc, coord, n)
    num = n
  END SUBROUTINE

  ! cube_bc20(rest,nxe,nye,nze), rest(nr,4)          geometry.f90:425-500
  SUBROUTINE ref_cube_bc20(nxe, nye, nze, nr, rest) BIND(C, name="ref_cube_bc20")
    INTEGER(c_int), VALUE :: nxe, nye, nze, nr
    INTEGER(c_int), INTENT(OUT) :: rest(nr, 4)
    INTEGER :: r(nr, 4)
    r = 0
    CALL cube_bc20(r, INT(nxe), INT(nye), INT(nze))
    rest = r
  END SUBROUTINE

  ! REARRANGE(rest) then FIND_G3(num, g, rest) for each element
  !                                         new_library.f90:3059-3212
  SUBROUTINE ref_steering(nod, nodof, nels, nr, g_num, rest, g_g) BIND(C, name="ref_steering")
    INTEGER(c_int), VALUE :: nod, nodof, nels, nr
    INTEGER(c_int), INTENT(IN) :: g_num(nod, nels), rest(nr, nodof + 1)
    INTEGER(c_int), INTENT(OUT) :: g_g(nod * nodof, nels)
    INTEGER :: r(nr, nodof + 1), num(nod), g(nod * nodof), iel
    r = rest
    CALL rearrange(r)
    DO iel = 1, nels
      num = g_num(:, iel)
      g = 0
      CALL find_g3(num, g, r)
      g_g(:, iel) = g
    END DO
  END SUBROUTINE

  ! ABAQUS2SG('hexahedron', g_num)          new_library.f90:3515-3682
  SUBROUTINE ref_abaqus2sg(nod, nels, g_num) BIND(C, name="ref_abaqus2sg")
    INTEGER(c_int), VALUE :: nod, nels
    INTEGER(c_int), INTENT(INOUT) :: g_num(nod, nels)
    INTEGER :: g(nod, nels)
    CHARACTER(LEN=15) :: element
    element = 'hexahedron'
    g = g_num
    CALL abaqus2sg(element, g)
    g_num = g
  END SUBROUTINE

END MODULE ref_bind
