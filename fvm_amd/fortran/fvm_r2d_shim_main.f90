! fvm_r2d_shim_main.f90 -- MAIN of the reference 2D solver (cylinder/cylinder_v3.0.3_AF_original
! 4step_face_nl/Main.f90:48-118) written out statement for statement against the reference's own
! subroutine interfaces, which fvm_shims2d.f90 binds to libfvm_amd.so's 2D context.  The module
! arrays are the host arrays of R2D_ARRAYS, bound once to their device fields; MAIN's array
! statements are the FVM_ helpers of fvm_shim_state.f90.  Inputs and outputs as fvm_r2d
! (namelist fvm_r2d.nml, the grid file, RESTART.DAT; NU_CD_CL.PLT, RESTART.DAT).  The passive
! thermal scalar is carried through the restart (ITER_T / BCT not called); VEL_ROT = 0.
MODULE R2D_ARRAYS
  IMPLICIT NONE
  DOUBLE PRECISION, ALLOCATABLE, TARGET, DIMENSION(:,:) :: U, V, P, T, U_N, V_N, NLX, NLY, &
       NLX_NM1, NLY_NM1, DIFX, DIFY, S_U, S_V, ADV
END MODULE R2D_ARRAYS

PROGRAM FVM_R2D_SHIM
  USE ISO_C_BINDING
  USE FVM_CAPI
  USE FVM_SHIM_STATE
  USE R2D_ARRAYS
  IMPLICIT NONE
  EXTERNAL :: CONTRA_VEL, NL, DIF, P_SOURCE, ITER_U, ITER_V, BCU, BCV, ITER_P, BCP, CALCCOEF, &
       FVM_ASSIGN, FVM_UPDATE

  INTEGER :: NI = 0, NJ = 0, NSWPP = 50, POISSON_MODE = 2, NSTEP = 1
  DOUBLE PRECISION :: RESIMAX = 1.D-8, RTIME_MAX = HUGE(1.D0)
  NAMELIST /FVM_RUN/ NI, NJ, NSWPP, RESIMAX, POISSON_MODE, NSTEP, RTIME_MAX

  TYPE(FVM_PARAMS) :: PRM
  TYPE(C_PTR) :: CTX
  DOUBLE PRECISION, ALLOCATABLE :: XG(:), YG(:)
  DOUBLE PRECISION :: RE, PR, DTIME, VEL_ROT, RTIME
  INTEGER :: NITER_TIME, ISTEP, NNODE, NCELL, Q, IU, I, J
  CHARACTER(LEN=11) :: GRIDFILE

  OPEN (NEWUNIT=IU, FILE='fvm_r2d.nml', STATUS='OLD')
  READ (IU, NML=FVM_RUN)
  CLOSE (IU)
  NNODE = NI * NJ
  NCELL = (NI + 1) * (NJ + 1)
  ALLOCATE (XG(NNODE), YG(NNODE))
  GRIDFILE = '000x000.dat'
  WRITE (GRIDFILE(1:3), '(I3.3)') NI
  WRITE (GRIDFILE(5:7), '(I3.3)') NJ
  OPEN (NEWUNIT=IU, FILE=GRIDFILE, STATUS='OLD')
  DO Q = 1, NNODE
    READ (IU, *) XG(Q), YG(Q)
  END DO
  CLOSE (IU)

  ALLOCATE (U(0:NI,0:NJ), V(0:NI,0:NJ), P(0:NI,0:NJ), T(0:NI,0:NJ), U_N(0:NI,0:NJ), &
       V_N(0:NI,0:NJ), NLX(0:NI,0:NJ), NLY(0:NI,0:NJ), NLX_NM1(0:NI,0:NJ), &
       NLY_NM1(0:NI,0:NJ), DIFX(0:NI,0:NJ), DIFY(0:NI,0:NJ), S_U(0:NI,0:NJ), S_V(0:NI,0:NJ), &
       ADV(0:NI,0:NJ))
  OPEN (NEWUNIT=IU, FILE='RESTART.DAT', STATUS='OLD')                 ! init.f90:47-64
  READ (IU, *) RE, PR, DTIME, VEL_ROT
  READ (IU, *) RTIME, NITER_TIME
  DO J = 0, NJ
    DO I = 0, NI
      READ (IU, *) NLX(I,J), NLY(I,J), ADV(I,J)
      READ (IU, *) U(I,J), V(I,J), P(I,J), T(I,J)
    END DO
  END DO
  CLOSE (IU)
  NITER_TIME = 0
  IF (VEL_ROT .NE. 0.D0) THEN
    WRITE (*, *) 'fvm_r2d_shim: a rotating wall (VEL_ROT /= 0) is not supported'
    ERROR STOP 1
  END IF

  PRM%NI = NI; PRM%NJ = NJ; PRM%NK = 2
  PRM%RE = RE; PRM%DTIME = DTIME; PRM%REF_A = 1.D0
  PRM%USE_LES = 0; PRM%NSWPP = NSWPP; PRM%RESIMAX = RESIMAX
  PRM%POISSON_MODE = POISSON_MODE; PRM%CHECK_EVERY = 8; PRM%AF_MODE = 0
  PRM%AF_TOL = 0.D0
  CALL FVM_CHECK(FVM_CREATE2D(PRM, XG, YG, CTX), 'fvm_create2d')    ! GRID/METRIC/INIT
  CALL FVM_SHIM_OPEN(CTX)
  CALL FVM_BIND(U, NCELL, FVM_U); CALL FVM_BIND(V, NCELL, FVM_V); CALL FVM_BIND(P, NCELL, FVM_P)
  CALL FVM_BIND(U_N, NCELL, FVM_UN); CALL FVM_BIND(V_N, NCELL, FVM_VN)
  CALL FVM_BIND(NLX, NCELL, FVM_NLX); CALL FVM_BIND(NLY, NCELL, FVM_NLY)
  CALL FVM_BIND(NLX_NM1, NCELL, FVM_NLX_NM1); CALL FVM_BIND(NLY_NM1, NCELL, FVM_NLY_NM1)
  CALL FVM_BIND(DIFX, NCELL, FVM_DIFX); CALL FVM_BIND(DIFY, NCELL, FVM_DIFY)
  CALL FVM_BIND(S_U, NCELL, FVM_SU); CALL FVM_BIND(S_V, NCELL, FVM_SV)
  CALL FVM_PUSH(U, NCELL); CALL FVM_PUSH(V, NCELL); CALL FVM_PUSH(P, NCELL)
  CALL FVM_PUSH(NLX, NCELL); CALL FVM_PUSH(NLY, NCELL)

  OPEN (NEWUNIT=IU, FILE='NU_CD_CL.PLT', STATUS='REPLACE')
  WRITE (IU, *) 'variables=TIME,NU,CD,CL,CPB,CD_P,CD_V,CL_P,CL_V,cfl_max'
  CLOSE (IU)

  ISTEP = 0
  DO WHILE (RTIME .LT. RTIME_MAX .AND. ISTEP .LT. NSTEP)                ! Main.f90:48,118
    ISTEP = ISTEP + 1
    NITER_TIME = NITER_TIME + 1                                         ! Main.f90:48-50
    RTIME = RTIME + DTIME
    CALL FVM_ASSIGN(U_N, U)                                             ! Main.f90:52-53
    CALL FVM_ASSIGN(V_N, V)
    CALL CONTRA_VEL                                                     ! Main.f90:58
    CALL FVM_ASSIGN(NLX_NM1, NLX)                                       ! Main.f90:61-62
    CALL FVM_ASSIGN(NLY_NM1, NLY)
    CALL NL(U_N, NLX)                                                   ! Main.f90:64-65
    CALL NL(V_N, NLY)
    CALL DIF(U_N, DIFX)                                                 ! Main.f90:69-70
    CALL DIF(V_N, DIFY)
    CALL P_SOURCE                                                       ! Main.f90:74
    CALL ITER_U                                                         ! Main.f90:77
    CALL ITER_V                                                         ! Main.f90:78
    CALL FVM_UPDATE(U, S_U, -1.D0)                                      ! Main.f90:80-84
    CALL FVM_UPDATE(V, S_V, -1.D0)
    CALL BCU                                                            ! Main.f90:86
    CALL BCV                                                            ! Main.f90:87
    CALL CONTRA_VEL                                                     ! Main.f90:91
    CALL ITER_P                                                         ! Main.f90:92
    CALL BCP                                                            ! Main.f90:93
    CALL P_SOURCE                                                       ! Main.f90:97
    CALL FVM_UPDATE(U, S_U, 1.D0)                                       ! Main.f90:99-103
    CALL FVM_UPDATE(V, S_V, 1.D0)
    CALL BCU                                                            ! Main.f90:106
    CALL BCV                                                            ! Main.f90:107
    CALL CALCCOEF                                                       ! Main.f90:114
    OPEN (NEWUNIT=IU, FILE='NU_CD_CL.PLT', STATUS='OLD', ACCESS='APPEND')
    WRITE (IU, 106) RTIME, 0.D0, SHIM_INFO%CD, SHIM_INFO%CL, SHIM_INFO%CPB, SHIM_INFO%CD_P, &
         SHIM_INFO%CD_V, SHIM_INFO%CL_P, SHIM_INFO%CL_V, SHIM_INFO%CFL
    CLOSE (IU)
    WRITE (*, '(A,I6,A,ES12.4,A,I8,A,F12.6)') ' step', ISTEP, ' RESIP', RESIP, ' NITER', NITER, &
         ' CD', SHIM_INFO%CD
  END DO

  ! RESTART.DAT in the reference's layout (writedata.f90:77-90)
  CALL FVM_PULL(U, NCELL); CALL FVM_PULL(V, NCELL); CALL FVM_PULL(P, NCELL)
  CALL FVM_PULL(NLX, NCELL); CALL FVM_PULL(NLY, NCELL)
  OPEN (NEWUNIT=IU, FILE='RESTART.DAT', STATUS='REPLACE')
  WRITE (IU, *) RE, PR, DTIME, VEL_ROT
  WRITE (IU, *) RTIME, NITER_TIME
  DO J = 0, NJ
    DO I = 0, NI
      WRITE (IU, *) NLX(I,J), NLY(I,J), ADV(I,J)
      WRITE (IU, *) U(I,J), V(I,J), P(I,J), T(I,J)
    END DO
  END DO
  CLOSE (IU)
  CALL FVM_CHECK(FVM_DESTROY(CTX), 'fvm_destroy')
106 FORMAT(10(1X,ES24.16E3))
END PROGRAM FVM_R2D_SHIM
