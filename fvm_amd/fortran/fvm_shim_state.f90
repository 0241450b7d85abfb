! fvm_shim_state.f90 -- what the reference-interface shims share: the context, the binding of
! host arrays (the reference's module arrays) to device fields, the module scalars MAIN reads
! (UIN, RESIP, NITER, the CALCCOEF results), and MAIN's own array statements on the device.
!
! The shims themselves are fvm_shims.f90 (the 3D LES solver's subroutines) and fvm_shims2d.f90
! (the 2D solver's); a driver links this file and one of them.  Each array argument of a shim
! stands for the device array it was bound to with FVM_BIND (once, after FVM_SHIM_OPEN).  The
! host arrays are handles; their contents cross only where the driver asks (FVM_PULL /
! FVM_PUSH, e.g. for restart I/O).  MAIN's array statements become the FVM_ helpers at the end
! of this file: U_N = U (FVM_ASSIGN), NLX = NLX + SGS_X*DJAC (FVM_ADD_SGS),
! U = U -+ DTIME*S_U/DJAC (FVM_UPDATE) and the pressure reference block (FVM_PREF).  UIN is the
! module variable MAIN sets before each ITER_U / BCU (1 for U, 0 for V and W); the shims check it.
MODULE FVM_SHIM_STATE
  USE ISO_C_BINDING
  USE FVM_CAPI
  IMPLICIT NONE
  TYPE(C_PTR) :: SHIM_CTX = C_NULL_PTR
  DOUBLE PRECISION :: UIN = 1.D0
  ! the scalars the reference keeps in COMDAT_SHARED after a step
  DOUBLE PRECISION :: RESIP = 0.D0, C_AVG = 0.D0, PREF = 0.D0
  INTEGER :: NITER = 0
  TYPE(FVM_STEP_INFO) :: SHIM_INFO
  INTEGER, PARAMETER :: MAXREG = 64
  INTEGER :: NREG = 0
  INTEGER(C_INTPTR_T) :: REG_ADDR(MAXREG)
  INTEGER(C_INT) :: REG_SEL(MAXREG)
  INTEGER(C_INTPTR_T) :: REG_LEN(MAXREG)

CONTAINS

  SUBROUTINE FVM_SHIM_OPEN(CTX)
    TYPE(C_PTR), INTENT(IN) :: CTX
    SHIM_CTX = CTX
    NREG = 0
  END SUBROUTINE FVM_SHIM_OPEN

  ! bind host array A (its first element's address) to device field SEL
  SUBROUTINE FVM_BIND(A, N, SEL)
    INTEGER, INTENT(IN) :: N
    DOUBLE PRECISION, TARGET, INTENT(IN) :: A(N)
    INTEGER(C_INT), INTENT(IN) :: SEL
    IF (NREG .GE. MAXREG) ERROR STOP 'FVM_BIND: too many arrays'
    NREG = NREG + 1
    REG_ADDR(NREG) = TRANSFER(C_LOC(A(1)), REG_ADDR(NREG))
    REG_SEL(NREG) = SEL
    REG_LEN(NREG) = N
  END SUBROUTINE FVM_BIND

  ! device field of a bound host array, from the address of its first element
  INTEGER(C_INT) FUNCTION SEL_OF(ADDR, WHAT)
    TYPE(C_PTR), INTENT(IN) :: ADDR
    CHARACTER(*), INTENT(IN) :: WHAT
    INTEGER :: Q
    INTEGER(C_INTPTR_T) :: A
    A = TRANSFER(ADDR, A)
    DO Q = 1, NREG
      IF (REG_ADDR(Q) .EQ. A) THEN
        SEL_OF = REG_SEL(Q)
        RETURN
      END IF
    END DO
    WRITE (*, '(A,A,A)') ' fvm_amd: ', WHAT, ': an array argument is not bound (FVM_BIND)'
    ERROR STOP 1
  END FUNCTION SEL_OF

  ! host array A <- its device field, or device field <- host array A
  SUBROUTINE FVM_PULL(A, N)
    INTEGER, INTENT(IN) :: N
    DOUBLE PRECISION, TARGET, INTENT(INOUT) :: A(N)
    CALL FVM_CHECK(FVM_GET_FIELD(SHIM_CTX, SEL_OF(C_LOC(A(1)), 'FVM_PULL'), A), 'FVM_PULL')
  END SUBROUTINE FVM_PULL

  SUBROUTINE FVM_PUSH(A, N)
    INTEGER, INTENT(IN) :: N
    DOUBLE PRECISION, TARGET, INTENT(IN) :: A(N)
    CALL FVM_CHECK(FVM_SET_FIELD(SHIM_CTX, SEL_OF(C_LOC(A(1)), 'FVM_PUSH'), A), 'FVM_PUSH')
  END SUBROUTINE FVM_PUSH

  ! UIN as MAIN sets it for the component of selector SEL
  SUBROUTINE CHECK_UIN(SEL, WHAT)
    INTEGER(C_INT), INTENT(IN) :: SEL
    CHARACTER(*), INTENT(IN) :: WHAT
    DOUBLE PRECISION :: WANT
    WANT = 0.D0
    IF (SEL .EQ. FVM_U) WANT = 1.D0
    IF (UIN .NE. WANT) THEN
      WRITE (*, '(A,A,A,F4.1,A,I3)') ' fvm_amd: ', WHAT, ' with UIN =', UIN, &
           ' for field', SEL
      ERROR STOP 1
    END IF
  END SUBROUTINE CHECK_UIN

END MODULE FVM_SHIM_STATE

! ---- MAIN's array statements ----

SUBROUTINE FVM_ASSIGN(DST, SRC)                              ! DST = SRC (main.f90:42-45, 51-54)
  USE FVM_SHIM_STATE
  DOUBLE PRECISION, TARGET :: DST(*), SRC(*)
  CALL FVM_CHECK(FVM_COPY_FIELD(SHIM_CTX, SEL_OF(C_LOC(DST(1)), '='), &
       SEL_OF(C_LOC(SRC(1)), '=')), 'array assignment')
END SUBROUTINE FVM_ASSIGN

SUBROUTINE FVM_ADD_SGS(NLA, SGS)                             ! NL = NL + SGS*DJAC (main.f90:61-66)
  USE FVM_SHIM_STATE
  DOUBLE PRECISION, TARGET :: NLA(*), SGS(*)
  CALL FVM_CHECK(FVM_ADD_SGS_C(SHIM_CTX, SEL_OF(C_LOC(NLA(1)), 'NL + SGS'), &
       SEL_OF(C_LOC(SGS(1)), 'NL + SGS')), 'NL + SGS')
END SUBROUTINE FVM_ADD_SGS

SUBROUTINE FVM_UPDATE(F, S, SGN)               ! F = F + SGN*DTIME*S/DJAC (main.f90:90-92,108-111)
  USE FVM_SHIM_STATE
  DOUBLE PRECISION, TARGET :: F(*), S(*)
  DOUBLE PRECISION, INTENT(IN) :: SGN
  CALL FVM_CHECK(FVM_CORRECT_C(SHIM_CTX, SEL_OF(C_LOC(F(1)), 'U = U + DT S / J'), &
       SEL_OF(C_LOC(S(1)), 'U = U + DT S / J'), SGN), 'U = U + DT S / J')
END SUBROUTINE FVM_UPDATE

SUBROUTINE FVM_PREF()                                        ! main.f90:125-135
  USE FVM_SHIM_STATE
  REAL(C_DOUBLE) :: R
  CALL FVM_CHECK(FVM_PRESSURE_REFERENCE(SHIM_CTX, R), 'pressure reference')
  PREF = R
END SUBROUTINE FVM_PREF
