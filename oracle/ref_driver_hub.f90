! ref_driver_hub.f90 -- test harness (not product code) that exposes the
! reference's own Hubbard HPROD (pyhub/libs/src/hprod_hubbard.f90:17-118),
! compiled from its sources, through a C ABI for the oracle checks.
! It only fills the HUBMOD globals (hubmod.f90:1-29) that HPROD reads.
MODULE PHREF_HUB
  USE ISO_C_BINDING
  USE HUBMOD
  IMPLICIT NONE
CONTAINS
  SUBROUTINE PHREF_HUB_SETUP(N, NU, ND, NSU, NSD, BU, BD, TM, UD) BIND(C, NAME="phref_hub_setup")
    INTEGER(C_INT), VALUE :: N, NU, ND, NSU, NSD
    INTEGER(C_INT), INTENT(IN) :: BU(NSU), BD(NSD)
    REAL(C_DOUBLE), INTENT(IN) :: TM(N,N), UD(N)
    NORB = N; NUP = NU; NDOWN = ND; NSUP = NSU; NSDOWN = NSD
    NSTATES = NSU*NSD
    IF (ALLOCATED(BUP)) DEALLOCATE(BUP)
    IF (ALLOCATED(BDOWN)) DEALLOCATE(BDOWN)
    IF (ALLOCATED(T)) DEALLOCATE(T)
    IF (ALLOCATED(ULDIAG)) DEALLOCATE(ULDIAG)
    IF (ALLOCATED(UPARTIAL)) DEALLOCATE(UPARTIAL)
    ALLOCATE(BUP(NSU), BDOWN(NSD), T(N,N), ULDIAG(N), UPARTIAL(N,N))
    BUP = BU; BDOWN = BD; ULDIAG = UD; UPARTIAL = 0.D0
    ! TM arrives row-major T[J][K]; Fortran sees it transposed
    T = TRANSPOSE(TM)
    IS_ULOC = 2
  END SUBROUTINE
  SUBROUTINE PHREF_HUB_HPROD(V, W) BIND(C, NAME="phref_hub_hprod")
    REAL(C_DOUBLE), INTENT(IN) :: V(*)
    REAL(C_DOUBLE), INTENT(OUT) :: W(*)
    CALL HPROD(V(1:NSTATES), W(1:NSTATES))
  END SUBROUTINE
END MODULE
