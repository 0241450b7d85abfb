! fvm_rles_shim_main.f90 -- MAIN of the reference 3D LES (LES/v2.3_LES-Reynolds stresses/
! main.f90:35-161) written out statement for statement against the reference's own subroutine
! interfaces, which fvm_shims.f90 binds to libfvm_amd.so.  The arrays of COMDAT_SHARED are the
! host arrays of module RLES_ARRAYS, each bound once to its device field; MAIN's array statements
! are the FVM_ helpers of fvm_shims.f90.  Inputs and outputs as fvm_rles (namelist fvm_rles.nml,
! the grid file, RESTART000.DAT; NU_CD_CL.PLT, RESTART000.DAT).  No passive temperature (the
! restart's T is carried), no field output and no POST_PROCESS: the call sequence of the step.
MODULE RLES_ARRAYS
  IMPLICIT NONE
  DOUBLE PRECISION, ALLOCATABLE, TARGET, DIMENSION(:,:,:) :: U, V, W, P, U_N, V_N, W_N, &
       NLX, NLY, NLZ, NLX_NM1, NLY_NM1, NLZ_NM1, DIFX, DIFY, DIFZ, S_U, S_V, S_W, &
       SGS_X, SGS_Y, SGS_Z, T, ADV
END MODULE RLES_ARRAYS

PROGRAM FVM_RLES_SHIM
  USE ISO_C_BINDING
  USE FVM_CAPI
  USE FVM_SHIM_STATE
  USE RLES_ARRAYS
  IMPLICIT NONE
  EXTERNAL :: CONTRA_VEL, NL, STRESS_TENSOR_DY, DIF, P_SOURCE, ITER_U, BCU, ITER_P, BCP, &
       CALCCOEF, FVM_ASSIGN, FVM_ADD_SGS, FVM_UPDATE, FVM_PREF

  INTEGER :: NI = 0, NJ = 0, NK = 0, NSWPP = 50, USE_LES = 1, POISSON_MODE = 2
  INTEGER :: NSTEP = 1
  DOUBLE PRECISION :: RE = 3900.D0, DTIME = 2.D-3, RESIMAX = 1.D-8, REF_A = -1.D0
  NAMELIST /FVM_RUN/ NI, NJ, NK, RE, DTIME, RESIMAX, NSWPP, USE_LES, REF_A, POISSON_MODE, NSTEP

  TYPE(FVM_PARAMS) :: PRM
  TYPE(C_PTR) :: CTX
  DOUBLE PRECISION, ALLOCATABLE :: XG(:), YG(:), ZG(:)
  DOUBLE PRECISION :: RE_F, PE_F, DT_F, RTIME
  INTEGER :: NITER_TIME, ISTEP, NNODE, NCELL, Q, IU, I, J, K
  CHARACTER(LEN=15) :: GRIDFILE

  OPEN (NEWUNIT=IU, FILE='fvm_rles.nml', STATUS='OLD')
  READ (IU, NML=FVM_RUN)
  CLOSE (IU)
  IF (REF_A .LE. 0.D0) REF_A = 4.D0 * ATAN(1.D0)

  NNODE = NI * NJ * NK
  NCELL = (NI + 1) * (NJ + 1) * (NK + 1)
  ALLOCATE (XG(NNODE), YG(NNODE), ZG(NNODE))
  GRIDFILE = '000x000x000.dat'
  WRITE (GRIDFILE(1:3), '(I3.3)') NI
  WRITE (GRIDFILE(5:7), '(I3.3)') NJ
  WRITE (GRIDFILE(9:11), '(I3.3)') NK
  OPEN (NEWUNIT=IU, FILE=GRIDFILE, STATUS='OLD')
  DO Q = 1, NNODE
    READ (IU, *) XG(Q), YG(Q), ZG(Q)
  END DO
  CLOSE (IU)

  ALLOCATE (U(0:NI,0:NJ,0:NK), V(0:NI,0:NJ,0:NK), W(0:NI,0:NJ,0:NK), P(0:NI,0:NJ,0:NK), &
       U_N(0:NI,0:NJ,0:NK), V_N(0:NI,0:NJ,0:NK), W_N(0:NI,0:NJ,0:NK), &
       NLX(0:NI,0:NJ,0:NK), NLY(0:NI,0:NJ,0:NK), NLZ(0:NI,0:NJ,0:NK), &
       NLX_NM1(0:NI,0:NJ,0:NK), NLY_NM1(0:NI,0:NJ,0:NK), NLZ_NM1(0:NI,0:NJ,0:NK), &
       DIFX(0:NI,0:NJ,0:NK), DIFY(0:NI,0:NJ,0:NK), DIFZ(0:NI,0:NJ,0:NK), &
       S_U(0:NI,0:NJ,0:NK), S_V(0:NI,0:NJ,0:NK), S_W(0:NI,0:NJ,0:NK), &
       SGS_X(0:NI,0:NJ,0:NK), SGS_Y(0:NI,0:NJ,0:NK), SGS_Z(0:NI,0:NJ,0:NK), &
       T(0:NI,0:NJ,0:NK), ADV(0:NI,0:NJ,0:NK))
  ! RESTART000.DAT (init.f90:188-203)
  OPEN (NEWUNIT=IU, FILE='RESTART000.DAT', STATUS='OLD')
  READ (IU, 122) RE_F, PE_F, DT_F, RTIME, NITER_TIME
  DO K = 0, NK
    DO J = 0, NJ
      DO I = 0, NI
        READ (IU, 128) NLX(I,J,K), NLY(I,J,K), NLZ(I,J,K), ADV(I,J,K)
        READ (IU, 129) U(I,J,K), V(I,J,K), W(I,J,K), P(I,J,K), T(I,J,K)
      END DO
    END DO
  END DO
  CLOSE (IU)
  NITER_TIME = 0

  PRM%NI = NI; PRM%NJ = NJ; PRM%NK = NK
  PRM%RE = RE; PRM%DTIME = DTIME; PRM%REF_A = REF_A
  PRM%USE_LES = USE_LES; PRM%NSWPP = NSWPP; PRM%RESIMAX = RESIMAX
  PRM%POISSON_MODE = POISSON_MODE; PRM%CHECK_EVERY = 8; PRM%AF_MODE = 0
  PRM%AF_TOL = 0.D0
  CALL FVM_CHECK(FVM_CREATE(PRM, XG, YG, ZG, CTX), 'fvm_create')   ! DEFINE/GRID/METRIC/INIT
  CALL FVM_SHIM_OPEN(CTX)
  CALL FVM_BIND(U, NCELL, FVM_U); CALL FVM_BIND(V, NCELL, FVM_V)
  CALL FVM_BIND(W, NCELL, FVM_W); CALL FVM_BIND(P, NCELL, FVM_P)
  CALL FVM_BIND(U_N, NCELL, FVM_UN); CALL FVM_BIND(V_N, NCELL, FVM_VN)
  CALL FVM_BIND(W_N, NCELL, FVM_WN)
  CALL FVM_BIND(NLX, NCELL, FVM_NLX); CALL FVM_BIND(NLY, NCELL, FVM_NLY)
  CALL FVM_BIND(NLZ, NCELL, FVM_NLZ)
  CALL FVM_BIND(NLX_NM1, NCELL, FVM_NLX_NM1); CALL FVM_BIND(NLY_NM1, NCELL, FVM_NLY_NM1)
  CALL FVM_BIND(NLZ_NM1, NCELL, FVM_NLZ_NM1)
  CALL FVM_BIND(DIFX, NCELL, FVM_DIFX); CALL FVM_BIND(DIFY, NCELL, FVM_DIFY)
  CALL FVM_BIND(DIFZ, NCELL, FVM_DIFZ)
  CALL FVM_BIND(S_U, NCELL, FVM_SU); CALL FVM_BIND(S_V, NCELL, FVM_SV)
  CALL FVM_BIND(S_W, NCELL, FVM_SW)
  CALL FVM_BIND(SGS_X, NCELL, FVM_SGSX); CALL FVM_BIND(SGS_Y, NCELL, FVM_SGSY)
  CALL FVM_BIND(SGS_Z, NCELL, FVM_SGSZ)
  CALL FVM_PUSH(U, NCELL); CALL FVM_PUSH(V, NCELL); CALL FVM_PUSH(W, NCELL)
  CALL FVM_PUSH(P, NCELL)
  CALL FVM_PUSH(NLX, NCELL); CALL FVM_PUSH(NLY, NCELL); CALL FVM_PUSH(NLZ, NCELL)

  OPEN (NEWUNIT=IU, FILE='NU_CD_CL.PLT', STATUS='REPLACE')
  WRITE (IU, *) 'variables=TIME,NU,CD,CL,CPB,CD_P,CD_V,CL_P,CL_V,CFL'
  CLOSE (IU)

  DO ISTEP = 1, NSTEP
    NITER_TIME = NITER_TIME + 1                                          ! main.f90:38-39
    RTIME = RTIME + DTIME
    CALL FVM_ASSIGN(U_N, U)                                              ! main.f90:41-44
    CALL FVM_ASSIGN(V_N, V)
    CALL FVM_ASSIGN(W_N, W)
    CALL CONTRA_VEL                                                      ! main.f90:48
    CALL FVM_ASSIGN(NLX_NM1, NLX)                                        ! main.f90:51-54
    CALL FVM_ASSIGN(NLY_NM1, NLY)
    CALL FVM_ASSIGN(NLZ_NM1, NLZ)
    CALL NL(U_N, NLX)                                                    ! main.f90:55-57
    CALL NL(V_N, NLY)
    CALL NL(W_N, NLZ)
    IF (USE_LES .EQ. 1) THEN                                             ! main.f90:59-69
      CALL STRESS_TENSOR_DY
      CALL FVM_ADD_SGS(NLX, SGS_X)
      CALL FVM_ADD_SGS(NLY, SGS_Y)
      CALL FVM_ADD_SGS(NLZ, SGS_Z)
    END IF
    CALL DIF(U_N, DIFX)                                                  ! main.f90:72-74
    CALL DIF(V_N, DIFY)
    CALL DIF(W_N, DIFZ)
    CALL P_SOURCE(P)                                                     ! main.f90:78
    UIN = 1.D0                                                           ! main.f90:82-88
    CALL ITER_U(U, U_N, NLX, NLX_NM1, DIFX, S_U)
    UIN = 0.D0
    CALL ITER_U(V, V_N, NLY, NLY_NM1, DIFY, S_V)
    CALL ITER_U(W, W_N, NLZ, NLZ_NM1, DIFZ, S_W)
    CALL FVM_UPDATE(U, S_U, -1.D0)                                       ! main.f90:90-92
    CALL FVM_UPDATE(V, S_V, -1.D0)
    CALL FVM_UPDATE(W, S_W, -1.D0)
    UIN = 1.D0                                                           ! main.f90:94-98
    CALL BCU(U, U_N)
    UIN = 0.D0
    CALL BCU(V, V_N)
    CALL BCU(W, W_N)
    CALL CONTRA_VEL                                                      ! main.f90:101
    CALL ITER_P                                                          ! main.f90:103
    CALL BCP                                                             ! main.f90:105
    CALL P_SOURCE(P)                                                     ! main.f90:108
    CALL FVM_UPDATE(U, S_U, 1.D0)                                        ! main.f90:109-111
    CALL FVM_UPDATE(V, S_V, 1.D0)
    CALL FVM_UPDATE(W, S_W, 1.D0)
    CALL FVM_PREF                                                        ! main.f90:125-135
    UIN = 1.D0                                                           ! main.f90:137-142
    CALL BCU(U, U_N)
    UIN = 0.D0
    CALL BCU(V, V_N)
    CALL BCU(W, W_N)
    CALL CALCCOEF                                                        ! main.f90:150
    OPEN (NEWUNIT=IU, FILE='NU_CD_CL.PLT', STATUS='OLD', ACCESS='APPEND')
    WRITE (IU, 106) RTIME, 0.D0, SHIM_INFO%CD, SHIM_INFO%CL, SHIM_INFO%CPB, SHIM_INFO%CD_P, &
         SHIM_INFO%CD_V, SHIM_INFO%CL_P, SHIM_INFO%CL_V, SHIM_INFO%CFL
    CLOSE (IU)
    WRITE (*, '(A,I6,A,ES12.4,A,I8,A,F12.6)') ' step', ISTEP, ' RESIP', RESIP, ' NITER', NITER, &
         ' CD', SHIM_INFO%CD
  END DO

  ! RESTART000.DAT in the reference's layout (writedata.f90:193-216)
  CALL FVM_PULL(U, NCELL); CALL FVM_PULL(V, NCELL); CALL FVM_PULL(W, NCELL)
  CALL FVM_PULL(P, NCELL)
  CALL FVM_PULL(NLX, NCELL); CALL FVM_PULL(NLY, NCELL); CALL FVM_PULL(NLZ, NCELL)
  OPEN (NEWUNIT=IU, FILE='RESTART000.DAT', STATUS='REPLACE')
  WRITE (IU, 122) RE, PE_F, DTIME, RTIME, NITER_TIME
  DO K = 0, NK
    DO J = 0, NJ
      DO I = 0, NI
        WRITE (IU, 128) NLX(I,J,K), NLY(I,J,K), NLZ(I,J,K), ADV(I,J,K)
        WRITE (IU, 129) U(I,J,K), V(I,J,K), W(I,J,K), P(I,J,K), T(I,J,K)
      END DO
    END DO
  END DO
  CLOSE (IU)
  CALL FVM_CHECK(FVM_DESTROY(CTX), 'fvm_destroy')

106 FORMAT(10(1X,ES24.16E3))
122 FORMAT(TR1,E24.16,TR1,E24.16,TR1,E24.16,TR1,E24.16,TR1,I16)
128 FORMAT(TR1,E24.16,TR1,E24.16,TR1,E24.16,TR1,E24.16)
129 FORMAT(TR1,E24.16,TR1,E24.16,TR1,E24.16,TR1,E24.16,TR1,E24.16)
END PROGRAM FVM_RLES_SHIM
