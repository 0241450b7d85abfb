! fvm_shims.f90 -- the reference 3D solver's subroutines, with its names and argument lists, on
! libfvm_amd.so (state and binding: fvm_shim_state.f90).
!
! A MAIN written against LES/v2.3_LES-Reynolds stresses (main.f90:35-161) calls external
! subroutines NL(IN,OUT), DIF(IN,OUT), ITER_U(IN,INM,INA,INB,INC,IND), BCU(A,B), P_SOURCE(P),
! CONTRA_VEL, STRESS_TENSOR_DY, ITER_P, BCP and CALCCOEF on the arrays of its module
! COMDAT_SHARED.  These shims keep those interfaces: NL(V_N, NLY) acts on the device arrays bound
! to V_N and NLY.

! ---- the reference's subroutines (RLES/*.f90), external as the reference's are ----

SUBROUTINE CONTRA_VEL()                                      ! contra_vel.f90:1-58
  USE FVM_SHIM_STATE
  CALL FVM_CHECK(FVM_CONTRA_VEL(SHIM_CTX), 'CONTRA_VEL')
END SUBROUTINE CONTRA_VEL

SUBROUTINE NL(IN, OUT)                                       ! nl.f90:1-49
  USE FVM_SHIM_STATE
  DOUBLE PRECISION, TARGET :: IN(*), OUT(*)
  CALL FVM_CHECK(FVM_NL_C(SHIM_CTX, SEL_OF(C_LOC(IN(1)), 'NL'), SEL_OF(C_LOC(OUT(1)), 'NL')), &
       'NL')
END SUBROUTINE NL

SUBROUTINE STRESS_TENSOR_DY()                                ! stress_tensor_dy.f90
  USE FVM_SHIM_STATE
  CALL FVM_CHECK(FVM_STRESS_TENSOR_DY_C(SHIM_CTX, C_AVG), 'STRESS_TENSOR_DY')
END SUBROUTINE STRESS_TENSOR_DY

SUBROUTINE DIF(IN, OUT)                                      ! dif.f90:1-40
  USE FVM_SHIM_STATE
  DOUBLE PRECISION, TARGET :: IN(*), OUT(*)
  CALL FVM_CHECK(FVM_DIF_C(SHIM_CTX, SEL_OF(C_LOC(IN(1)), 'DIF'), SEL_OF(C_LOC(OUT(1)), 'DIF')), &
       'DIF')
END SUBROUTINE DIF

SUBROUTINE P_SOURCE(P)                                       ! p_source.f90:1-40
  USE FVM_SHIM_STATE
  DOUBLE PRECISION, TARGET :: P(*)
  IF (SEL_OF(C_LOC(P(1)), 'P_SOURCE') .NE. FVM_P) ERROR STOP 'P_SOURCE: argument is not P'
  CALL FVM_CHECK(FVM_P_SOURCE(SHIM_CTX), 'P_SOURCE')
END SUBROUTINE P_SOURCE

SUBROUTINE ITER_U(IN, INM, INA, INB, INC, IND)               ! iter_u.f90:1-262
  USE FVM_SHIM_STATE
  DOUBLE PRECISION, TARGET :: IN(*), INM(*), INA(*), INB(*), INC(*), IND(*)
  INTEGER(C_INT) :: F
  F = SEL_OF(C_LOC(IN(1)), 'ITER_U')
  CALL CHECK_UIN(F, 'ITER_U')
  CALL FVM_CHECK(FVM_ITER_U_C(SHIM_CTX, F, SEL_OF(C_LOC(INM(1)), 'ITER_U'), &
       SEL_OF(C_LOC(INA(1)), 'ITER_U'), SEL_OF(C_LOC(INB(1)), 'ITER_U'), &
       SEL_OF(C_LOC(INC(1)), 'ITER_U'), SEL_OF(C_LOC(IND(1)), 'ITER_U')), 'ITER_U')
END SUBROUTINE ITER_U

SUBROUTINE BCU(A, B)                                         ! bcu.f90:1-50
  USE FVM_SHIM_STATE
  DOUBLE PRECISION, TARGET :: A(*), B(*)
  INTEGER(C_INT) :: F
  F = SEL_OF(C_LOC(A(1)), 'BCU')
  CALL CHECK_UIN(F, 'BCU')
  CALL FVM_CHECK(FVM_BCU(SHIM_CTX, F, SEL_OF(C_LOC(B(1)), 'BCU')), 'BCU')
END SUBROUTINE BCU

SUBROUTINE ITER_P()                                          ! iter_p.f90:1-66
  USE FVM_SHIM_STATE
  INTEGER(C_INT) :: N
  REAL(C_DOUBLE) :: R
  CALL FVM_CHECK(FVM_ITER_P(SHIM_CTX, R, N), 'ITER_P')
  RESIP = R
  NITER = N
END SUBROUTINE ITER_P

SUBROUTINE BCP()                                             ! bcp.f90:1-38
  USE FVM_SHIM_STATE
  CALL FVM_CHECK(FVM_BCP(SHIM_CTX), 'BCP')
END SUBROUTINE BCP

SUBROUTINE CALCCOEF()                                        ! calccoef.f90:1-201
  USE FVM_SHIM_STATE
  CALL FVM_CHECK(FVM_CALCCOEF(SHIM_CTX, SHIM_INFO), 'CALCCOEF')
END SUBROUTINE CALCCOEF

