! fvm_r2d_main.f90 -- time loop of the 2D cylinder solver on libfvm_amd.so.
!
! Mirrors MAIN of cylinder/cylinder_v3.0.3_AF_original 4step_face_nl/Main.f90:48-110 call for
! call through the C ABI (fvm_create2d: the 2D step on a one-cell span slab of the kernels).
! Inputs are the reference's own files: the grid file NNNxNNN.dat (grid.f90:28-43) and
! RESTART.DAT (init.f90:47-64, list-directed).  Outputs: NU_CD_CL.PLT (calccoef.f90:114-116),
! the console lines of calccoef.f90:118-125 and RESTART.DAT at the end (writedata.f90:77-90).
! With RTIME_END > RTIME_MAX the run continues past RTIME_MAX in the post-processing phase
! (Main.f90:121-123, post_process.f90:9-18): every step adds to the running means of CALCCOEF's
! coefficients, the Strouhal number from the sign changes of dC_L and the C_L amplitude, and
! the rms of C_D and C_L about their running means, written to POST_NU_CD_CL.PLT
! (calccoef.f90:156-195); at RTIME_END (or NSTEP) the final values go to
! TIME_AVERAGED_RESULTS.DAT (post_process.f90:19-60; the time-averaged fields are not written).
! Run constants come from the namelist file fvm_r2d.nml.  The passive thermal scalar is carried
! through the restart unchanged and NU is written as 0; the wall does not rotate (VEL_ROT = 0).
PROGRAM FVM_R2D
  USE ISO_C_BINDING
  USE FVM_CAPI
  IMPLICIT NONE

  INTEGER :: NI = 0, NJ = 0, NSWPP = 50, POISSON_MODE = 2, NSTEP = HUGE(1)
  DOUBLE PRECISION :: RESIMAX = 1.D-8, RTIME_MAX = HUGE(1.D0), RTIME_END = -1.D0
  NAMELIST /FVM_RUN/ NI, NJ, NSWPP, RESIMAX, POISSON_MODE, NSTEP, RTIME_MAX, RTIME_END

  TYPE(FVM_PARAMS) :: PRM
  TYPE(FVM_STEP_INFO) :: INFO
  TYPE(C_PTR) :: CTX
  DOUBLE PRECISION, ALLOCATABLE :: XG(:), YG(:)
  DOUBLE PRECISION, ALLOCATABLE :: U(:), V(:), P(:), T(:), NLX(:), NLY(:), ADV(:)
  DOUBLE PRECISION :: RE, PR, DTIME, VEL_ROT, RTIME, RESIP, NU
  INTEGER :: NITER_TIME, NITER, ISTEP, NNODE, NCELL, Q, IU
  ! the post-processing phase (calccoef.f90:156-195): CL of the previous step, the toggle of the
  ! dC_L sign, the times and C_L of the last extrema, the running sums
  INTEGER :: IPOST = 0, TOGGLE = 1, NITER_POS = 0
  DOUBLE PRECISION :: CL_NM1 = 0.D0, TIME1 = 0.D0, TIME2 = 0.D0, CL1 = 0.D0, ST = 0.D0, &
       CL_AMP = 0.D0, NU_TA = 0.D0, CD_TA = 0.D0, CL_TA = 0.D0, ST_TA = 0.D0, CPB_TA = 0.D0, &
       CL_AMP_TA = 0.D0, RMS_CD = 0.D0, RMS_CL = 0.D0, DUM_A, DUM_B
  CHARACTER(LEN=256) :: NMLFILE
  CHARACTER(LEN=11) :: GRIDFILE
  LOGICAL :: EXISTS

  NMLFILE = 'fvm_r2d.nml'
  IF (COMMAND_ARGUMENT_COUNT() .GE. 1) CALL GET_COMMAND_ARGUMENT(1, NMLFILE)
  OPEN (NEWUNIT=IU, FILE=TRIM(NMLFILE), STATUS='OLD')
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

  ALLOCATE (U(NCELL), V(NCELL), P(NCELL), T(NCELL), NLX(NCELL), NLY(NCELL), ADV(NCELL))
  OPEN (NEWUNIT=IU, FILE='RESTART.DAT', STATUS='OLD')
  READ (IU, *) RE, PR, DTIME, VEL_ROT
  READ (IU, *) RTIME, NITER_TIME
  DO Q = 1, NCELL
    READ (IU, *) NLX(Q), NLY(Q), ADV(Q)
    READ (IU, *) U(Q), V(Q), P(Q), T(Q)
  END DO
  CLOSE (IU)
  NITER_TIME = 0
  IF (VEL_ROT .NE. 0.D0) THEN
    WRITE (*, *) 'fvm_r2d: a rotating wall (VEL_ROT /= 0) is not supported'
    ERROR STOP 1
  END IF

  PRM%NI = NI; PRM%NJ = NJ; PRM%NK = 2
  PRM%RE = RE; PRM%DTIME = DTIME; PRM%REF_A = 1.D0
  PRM%USE_LES = 0; PRM%NSWPP = NSWPP; PRM%RESIMAX = RESIMAX
  PRM%POISSON_MODE = POISSON_MODE; PRM%CHECK_EVERY = 8; PRM%AF_MODE = 0
  PRM%AF_TOL = 0.D0
  CALL FVM_CHECK(FVM_CREATE2D(PRM, XG, YG, CTX), 'fvm_create2d')    ! GRID/METRIC/INIT
  CALL FVM_CHECK(FVM_SET_FIELD(CTX, FVM_U, U), 'fvm_set_field')
  CALL FVM_CHECK(FVM_SET_FIELD(CTX, FVM_V, V), 'fvm_set_field')
  CALL FVM_CHECK(FVM_SET_FIELD(CTX, FVM_P, P), 'fvm_set_field')
  CALL FVM_CHECK(FVM_SET_FIELD(CTX, FVM_NLX, NLX), 'fvm_set_field')
  CALL FVM_CHECK(FVM_SET_FIELD(CTX, FVM_NLY, NLY), 'fvm_set_field')

  INQUIRE (FILE='NU_CD_CL.PLT', EXIST=EXISTS)
  IF (.NOT. EXISTS) THEN
    OPEN (NEWUNIT=IU, FILE='NU_CD_CL.PLT', STATUS='NEW')
    WRITE (IU, *) 'variables=TIME,NU,CD,CL,CPB,CD_P,CD_V,CL_P,CL_V,cfl_max'
    CLOSE (IU)
  END IF
  NU = 0.D0

  ISTEP = 0
  DO                                                                   ! Main.f90:48,118-123
  IF (IPOST .EQ. 0 .AND. RTIME .GE. RTIME_MAX .AND. RTIME_END .GT. RTIME_MAX) THEN
    ! POST_PROCESS's first call (post_process.f90:9-18)
    OPEN (NEWUNIT=IU, FILE='POST_NU_CD_CL.PLT', STATUS='REPLACE')
    WRITE (IU, *) 'variables=TIME,NU,CD,CL,CL_AMP,CPB,ST,,RMS_CD,RMS_CL'
    CLOSE (IU)
    RTIME_MAX = RTIME_END
    NITER_POS = 0
    TOGGLE = 1
    IPOST = 1
  END IF
  DO WHILE (RTIME .LT. RTIME_MAX .AND. ISTEP .LT. NSTEP)
    ISTEP = ISTEP + 1
    NITER_TIME = NITER_TIME + 1
    RTIME = RTIME + DTIME
    CALL FVM_CHECK(FVM_BEGIN_STEP(CTX), 'U_N = U')                     ! Main.f90:54-56
    CALL FVM_CHECK(FVM_CONTRA_VEL(CTX), 'CONTRA_VEL')                  ! Main.f90:60
    CALL FVM_CHECK(FVM_NL(CTX), 'NL')                                  ! Main.f90:63-68
    CALL FVM_CHECK(FVM_DIF(CTX), 'DIF')                                ! Main.f90:71-73
    CALL FVM_CHECK(FVM_P_SOURCE(CTX), 'P_SOURCE')                      ! Main.f90:76
    CALL FVM_CHECK(FVM_ITER_U(CTX), 'ITER_U / ITER_V')                 ! Main.f90:79-80
    CALL FVM_CHECK(FVM_CORRECT(CTX, -1.D0), 'correction + BCU/BCV')    ! Main.f90:82-89
    CALL FVM_CHECK(FVM_CONTRA_VEL(CTX), 'CONTRA_VEL')                  ! Main.f90:93
    CALL FVM_CHECK(FVM_ITER_P(CTX, RESIP, NITER), 'ITER_P + BCP')      ! Main.f90:94-95
    CALL FVM_CHECK(FVM_P_SOURCE(CTX), 'P_SOURCE')                      ! Main.f90:99
    CALL FVM_CHECK(FVM_CORRECT(CTX, 1.D0), 'projection + BCU/BCV')     ! Main.f90:101-109
    CALL FVM_CHECK(FVM_CALCCOEF(CTX, INFO), 'CALCCOEF')                ! Main.f90:116

    OPEN (NEWUNIT=IU, FILE='NU_CD_CL.PLT', STATUS='OLD', ACCESS='APPEND')
    WRITE (IU, 106) RTIME, NU, INFO%CD, INFO%CL, INFO%CPB, INFO%CD_P, INFO%CD_V, INFO%CL_P, &
         INFO%CL_V, INFO%CFL
    CLOSE (IU)
    PRINT *, ''
    WRITE (*, 101) RTIME, NITER_TIME, INFO%CFL
    WRITE (*, 102) NU
    WRITE (*, 103) INFO%CD, INFO%CD_P, INFO%CD_V
    WRITE (*, 104) INFO%CL, INFO%CL_P, INFO%CL_V, INFO%CPB
    WRITE (*, 105) NITER, RESIP
    IF (IPOST .EQ. 1) CALL POST_STEP()
    CL_NM1 = INFO%CL
  END DO
  IF (IPOST .EQ. 1 .OR. ISTEP .GE. NSTEP .OR. RTIME_END .LE. RTIME_MAX) EXIT
  END DO
  IF (IPOST .EQ. 1 .AND. NITER_POS .GT. 0) THEN
    ! POST_PROCESS's second call: the time-averaged coefficients (post_process.f90:19-60)
    OPEN (NEWUNIT=IU, FILE='TIME_AVERAGED_RESULTS.DAT', STATUS='REPLACE')
    WRITE (IU, *) 'variables=TIME,NU,CD,CL,CL_AMP,CPB,ST,RMS_CD,RMS_CL'
    WRITE (IU, 107) RTIME, NU_TA / DBLE(NITER_POS), CD_TA / DBLE(NITER_POS), &
         CL_TA / DBLE(NITER_POS), CL_AMP_TA / DBLE(NITER_POS), CPB_TA / DBLE(NITER_POS), &
         ST_TA / DBLE(NITER_POS), SQRT(RMS_CD / DBLE(NITER_POS)), SQRT(RMS_CL / DBLE(NITER_POS))
    CLOSE (IU)
  END IF

  ! RESTART.DAT in the reference's layout (writedata.f90:77-90)
  CALL FVM_CHECK(FVM_GET_FIELD(CTX, FVM_U, U), 'fvm_get_field')
  CALL FVM_CHECK(FVM_GET_FIELD(CTX, FVM_V, V), 'fvm_get_field')
  CALL FVM_CHECK(FVM_GET_FIELD(CTX, FVM_P, P), 'fvm_get_field')
  CALL FVM_CHECK(FVM_GET_FIELD(CTX, FVM_NLX, NLX), 'fvm_get_field')
  CALL FVM_CHECK(FVM_GET_FIELD(CTX, FVM_NLY, NLY), 'fvm_get_field')
  OPEN (NEWUNIT=IU, FILE='RESTART.DAT', STATUS='REPLACE')
  WRITE (IU, *) RE, PR, DTIME, VEL_ROT
  WRITE (IU, *) RTIME, NITER_TIME
  DO Q = 1, NCELL
    WRITE (IU, *) NLX(Q), NLY(Q), ADV(Q)
    WRITE (IU, *) U(Q), V(Q), P(Q), T(Q)
  END DO
  CLOSE (IU)
  CALL FVM_CHECK(FVM_DESTROY(CTX), 'fvm_destroy')

101 FORMAT(TR6,'Time',F10.3,TR12,'ITER',I10,TR12,'CFL',F10.3)
102 FORMAT(TR1,'NU =',F10.6)
103 FORMAT(TR1,'CD =',F10.6,TR4,'CD_P =',F10.6,TR4,'CD_V =',F10.6)
104 FORMAT(TR1,'CL =',F10.6,TR4,'CL_P =',F10.6,TR4,'CL_V =',F10.6,TR4,'CPB =',F10.6)
105 FORMAT(TR1,'NITER =',I8,TR4,'RESIP =',ES12.4)
106 FORMAT(10(1X,ES24.16E3))
107 FORMAT(9(1X,ES24.16E3))

CONTAINS

  ! calccoef.f90:156-195 after CALCCOEF's coefficients of this step (NU = 0: no temperature)
  SUBROUTINE POST_STEP()
    INTEGER :: JU
    NITER_POS = NITER_POS + 1
    IF (DBLE(TOGGLE) * (INFO%CL - CL_NM1) .LT. 0.D0) THEN   ! an extremum of C_L
      TOGGLE = -1 * TOGGLE
      ST = 1.D0 / (RTIME - TIME2)
      TIME2 = TIME1
      TIME1 = RTIME
      CL_AMP = ABS(INFO%CL - CL1) / 2.D0
      CL1 = INFO%CL
    END IF
    NU_TA = NU_TA + NU
    CD_TA = CD_TA + INFO%CD
    CL_TA = CL_TA + INFO%CL
    ST_TA = ST_TA + ST
    CPB_TA = CPB_TA + INFO%CPB
    CL_AMP_TA = CL_AMP_TA + CL_AMP
    DUM_A = CD_TA / DBLE(NITER_POS)
    DUM_B = CL_TA / DBLE(NITER_POS)
    RMS_CD = RMS_CD + (INFO%CD - DUM_A)**2
    RMS_CL = RMS_CL + (INFO%CL - DUM_B)**2
    OPEN (NEWUNIT=JU, FILE='POST_NU_CD_CL.PLT', STATUS='OLD', ACCESS='APPEND')
    WRITE (JU, 207) RTIME, NU_TA / DBLE(NITER_POS), DUM_A, DUM_B, CL_AMP_TA / DBLE(NITER_POS), &
         CPB_TA / DBLE(NITER_POS), ST_TA / DBLE(NITER_POS), SQRT(RMS_CD / DBLE(NITER_POS)), &
         SQRT(RMS_CL / DBLE(NITER_POS))
    CLOSE (JU)
207 FORMAT(9(1X,ES24.16E3))
  END SUBROUTINE POST_STEP

END PROGRAM FVM_R2D
