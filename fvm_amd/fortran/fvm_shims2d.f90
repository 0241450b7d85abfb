! fvm_shims2d.f90 -- the reference 2D solver's subroutines (cylinder/cylinder_v3.0.3_AF_original
! 4step_face_nl/), with its names and argument lists, on libfvm_amd.so's 2D context (state and
! binding: fvm_shim_state.f90).
!
! The 2D MAIN (Main.f90:48-118) calls NL(IN,OUT), DIF(IN,OUT) and the argument-free CONTRA_VEL,
! P_SOURCE, ITER_U, ITER_V, BCU, BCV, ITER_P, BCP and CALCCOEF on its module arrays; ITER_U /
! ITER_V (iter_u.f90, iter_v.f90) advance U / V from U_N, NLX, NLX_NM1, DIFX, S_U (V likewise),
! BCU / BCV set the ghosts of U / V from U_N / V_N.  The array arguments stand for the device
! arrays bound to them (FVM_BIND); the argument-free ones act on the fields the 2D module names.

SUBROUTINE CONTRA_VEL()                                      ! contra_vel.f90
  USE FVM_SHIM_STATE
  CALL FVM_CHECK(FVM_CONTRA_VEL(SHIM_CTX), 'CONTRA_VEL')
END SUBROUTINE CONTRA_VEL

SUBROUTINE NL(IN, OUT)                                       ! nl.f90:1-29
  USE FVM_SHIM_STATE
  DOUBLE PRECISION, TARGET :: IN(*), OUT(*)
  CALL FVM_CHECK(FVM_NL_C(SHIM_CTX, SEL_OF(C_LOC(IN(1)), 'NL'), SEL_OF(C_LOC(OUT(1)), 'NL')), &
       'NL')
END SUBROUTINE NL

SUBROUTINE DIF(IN, OUT)                                      ! dif.f90:1-25
  USE FVM_SHIM_STATE
  DOUBLE PRECISION, TARGET :: IN(*), OUT(*)
  CALL FVM_CHECK(FVM_DIF_C(SHIM_CTX, SEL_OF(C_LOC(IN(1)), 'DIF'), SEL_OF(C_LOC(OUT(1)), 'DIF')), &
       'DIF')
END SUBROUTINE DIF

SUBROUTINE P_SOURCE()                                        ! p_source.f90:1-38
  USE FVM_SHIM_STATE
  CALL FVM_CHECK(FVM_P_SOURCE(SHIM_CTX), 'P_SOURCE')
END SUBROUTINE P_SOURCE

SUBROUTINE ITER_U()                                          ! iter_u.f90:1-121
  USE FVM_SHIM_STATE
  CALL FVM_CHECK(FVM_ITER_U_C(SHIM_CTX, FVM_U, FVM_UN, FVM_NLX, FVM_NLX_NM1, FVM_DIFX, FVM_SU), &
       'ITER_U')
END SUBROUTINE ITER_U

SUBROUTINE ITER_V()                                          ! iter_v.f90
  USE FVM_SHIM_STATE
  CALL FVM_CHECK(FVM_ITER_U_C(SHIM_CTX, FVM_V, FVM_VN, FVM_NLY, FVM_NLY_NM1, FVM_DIFY, FVM_SV), &
       'ITER_V')
END SUBROUTINE ITER_V

SUBROUTINE BCU()                                             ! bcu.f90:1-53
  USE FVM_SHIM_STATE
  CALL FVM_CHECK(FVM_BCU(SHIM_CTX, FVM_U, FVM_UN), 'BCU')
END SUBROUTINE BCU

SUBROUTINE BCV()                                             ! bcv.f90
  USE FVM_SHIM_STATE
  CALL FVM_CHECK(FVM_BCU(SHIM_CTX, FVM_V, FVM_VN), 'BCV')
END SUBROUTINE BCV

SUBROUTINE ITER_P()                                          ! iter_p.f90:1-54
  USE FVM_SHIM_STATE
  INTEGER(C_INT) :: N
  REAL(C_DOUBLE) :: R
  CALL FVM_CHECK(FVM_ITER_P(SHIM_CTX, R, N), 'ITER_P')
  RESIP = R
  NITER = N
END SUBROUTINE ITER_P

SUBROUTINE BCP()                                             ! bcp.f90:1-45
  USE FVM_SHIM_STATE
  CALL FVM_CHECK(FVM_BCP(SHIM_CTX), 'BCP')
END SUBROUTINE BCP

SUBROUTINE CALCCOEF()                                        ! calccoef.f90:1-154
  USE FVM_SHIM_STATE
  CALL FVM_CHECK(FVM_CALCCOEF(SHIM_CTX, SHIM_INFO), 'CALCCOEF')
END SUBROUTINE CALCCOEF
