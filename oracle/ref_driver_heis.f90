! ref_driver_heis.f90 -- test harness (not product code) that exposes the
! reference's own Heisenberg HPROD (pyhub/libs/src/hprod_heis.f90:1-61) and
! its J_SPIN (static_rq_heis.f90:58-116), compiled from their sources, through
! a C ABI for the oracle checks.
MODULE PHREF_HEIS
  USE ISO_C_BINDING
  USE HUBMOD
  IMPLICIT NONE
CONTAINS
  SUBROUTINE PHREF_HEIS_SETUP(N, NSU, BU, JM) BIND(C, NAME="phref_heis_setup")
    INTEGER(C_INT), VALUE :: N, NSU
    INTEGER(C_INT), INTENT(IN) :: BU(NSU)
    REAL(C_DOUBLE), INTENT(IN) :: JM(N,N)
    NORB = N; NUP = N/2; NDOWN = 0; NSUP = NSU; NSDOWN = 1
    NSTATES = NSU
    IF (ALLOCATED(BUP)) DEALLOCATE(BUP)
    IF (ALLOCATED(J_MATRIX)) DEALLOCATE(J_MATRIX)
    ALLOCATE(BUP(NSU), J_MATRIX(N,N))
    BUP = BU
    J_MATRIX = TRANSPOSE(JM)
  END SUBROUTINE
  SUBROUTINE PHREF_HEIS_HPROD(V, W) BIND(C, NAME="phref_heis_hprod")
    REAL(C_DOUBLE), INTENT(IN) :: V(*)
    REAL(C_DOUBLE), INTENT(OUT) :: W(*)
    CALL HPROD(V(1:NSTATES), W(1:NSTATES))
  END SUBROUTINE
  ! SMATRIX(J,K) of the Heisenberg J_SPIN (Fortran order)
  SUBROUTINE PHREF_HEIS_J_SPIN(BRA, KET, SM) BIND(C, NAME="phref_heis_j_spin")
    REAL(C_DOUBLE), INTENT(IN) :: BRA(*), KET(*)
    REAL(C_DOUBLE), INTENT(OUT) :: SM(*)
    REAL*8, ALLOCATABLE :: S(:,:)
    ALLOCATE(S(NORB,NORB))
    CALL J_SPIN(BRA(1:NSTATES), KET(1:NSTATES), S)
    SM(1:NORB*NORB) = RESHAPE(S, (/NORB*NORB/))
    DEALLOCATE(S)
  END SUBROUTINE
END MODULE
