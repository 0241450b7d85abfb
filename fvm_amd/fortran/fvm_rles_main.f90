! fvm_rles_main.f90 -- time loop of the 3D dynamic-Smagorinsky LES on libfvm_amd.so.
!
! Mirrors MAIN (LES/v2.3_LES-Reynolds stresses/main.f90:35-161) call for call; every reference
! subroutine of the step is one entry point of the C ABI acting on the device-resident state.
! Inputs are the reference's own files: the grid file NNNxNNNxNNN.dat (grid.f90:46-67) and
! RESTART000.DAT (init.f90:183-208).  Outputs: NU_CD_CL.PLT (calccoef.f90:186-187), the console
! lines of calccoef.f90:190-198, CS.PLT / CT.PLT (stress_tensor_dy.f90:519-522, 707-711), the
! nodal Tecplot fields PVT_nnn_000.PLT (x, y, z, p, u, v, w, T) and VOR_nnn_000.PLT (x, y, z,
! |omega|, lambda_2) every NITER_W steps (writedata.f90:8-189) and RESTART000.DAT every NITER_R
! steps and at the end (writedata.f90:193-216).  With RTIME_END > RTIME_MAX the run continues
! after RTIME_MAX in the time-averaging phase of POST_PROCESS (main.f90:164-166,
! post_process.f90:9-276): DTIME = DTIME_POST up to RTIME_END (or NSTEP), then the means, rms,
! Reynolds stresses and wall coefficients to Avrg000.PLT, Urms000.PLT, Strs000.PLT and
! Surf000.PLT.  Run constants
! come from the namelist file fvm_rles.nml (the reference compiles them in,
! comdat_param.f90:6-48).  The passive temperature is advanced as in the reference
! (THERMAL = 1, Peclet number PE) or carried through the restart unchanged with NU = 0
! (THERMAL = 0).
PROGRAM FVM_RLES
  USE ISO_C_BINDING
  USE FVM_CAPI
  IMPLICIT NONE

  INTEGER :: NI = 0, NJ = 0, NK = 0, NSWPP = 50, USE_LES = 1, POISSON_MODE = 2
  INTEGER :: NITER_R = HUGE(1), NSTEP = HUGE(1), AF_MODE = 0, THERMAL = 1, NITER_W = HUGE(1)
  DOUBLE PRECISION :: RE = 3900.D0, DTIME = 2.D-3, RESIMAX = 1.D-8, REF_A = -1.D0
  DOUBLE PRECISION :: RTIME_MAX = HUGE(1.D0), PE = 2769.D0, PR = 0.71D0
  DOUBLE PRECISION :: RTIME_END = -1.D0, DTIME_POST = 1.D-3   ! comdat_param.f90:26, 44
  NAMELIST /FVM_RUN/ NI, NJ, NK, RE, DTIME, RESIMAX, NSWPP, USE_LES, REF_A, POISSON_MODE, &
       RTIME_MAX, NITER_R, NSTEP, AF_MODE, THERMAL, PE, PR, NITER_W, RTIME_END, DTIME_POST

  TYPE(FVM_PARAMS) :: PRM
  TYPE(FVM_STEP_INFO) :: INFO
  TYPE(C_PTR) :: CTX
  DOUBLE PRECISION, ALLOCATABLE :: XG(:), YG(:), ZG(:)
  DOUBLE PRECISION, ALLOCATABLE :: U(:), V(:), W(:), P(:), NLX(:), NLY(:), NLZ(:), ADV(:)
  DOUBLE PRECISION, ALLOCATABLE, TARGET :: T(:)
  DOUBLE PRECISION :: RE_F, PE_F, DT_F, RTIME, RESIP, C_AVG, PREF, NU
  INTEGER :: NITER_TIME, NITER, ISTEP, NNODE, NCELL, Q, IU, IOS, NUM_WRITE = 0, IPOST = 0
  CHARACTER(LEN=256) :: NMLFILE
  CHARACTER(LEN=15) :: GRIDFILE
  LOGICAL :: EXISTS

  NMLFILE = 'fvm_rles.nml'
  IF (COMMAND_ARGUMENT_COUNT() .GE. 1) CALL GET_COMMAND_ARGUMENT(1, NMLFILE)
  OPEN (NEWUNIT=IU, FILE=TRIM(NMLFILE), STATUS='OLD')
  READ (IU, NML=FVM_RUN)
  CLOSE (IU)
  IF (REF_A .LE. 0.D0) REF_A = 4.D0 * ATAN(1.D0)   ! D * L_z with L_z = pi (calccoef.f90:175-179)

  ! grid file: NI*NJ*NK records X Y Z, I fastest (grid.f90:46-67)
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

  ! restart: header, then per cell NLX NLY NLZ ADV / U V W P T (init.f90:188-203)
  ALLOCATE (U(NCELL), V(NCELL), W(NCELL), P(NCELL), T(NCELL), NLX(NCELL), NLY(NCELL), &
       NLZ(NCELL), ADV(NCELL))
  OPEN (NEWUNIT=IU, FILE='RESTART000.DAT', STATUS='OLD')
  READ (IU, 122) RE_F, PE_F, DT_F, RTIME, NITER_TIME
  DO Q = 1, NCELL
    READ (IU, 128) NLX(Q), NLY(Q), NLZ(Q), ADV(Q)
    READ (IU, 129) U(Q), V(Q), W(Q), P(Q), T(Q)
  END DO
  CLOSE (IU)
  NITER_TIME = 0

  PRM%NI = NI; PRM%NJ = NJ; PRM%NK = NK
  PRM%RE = RE; PRM%DTIME = DTIME; PRM%REF_A = REF_A
  PRM%USE_LES = USE_LES; PRM%NSWPP = NSWPP; PRM%RESIMAX = RESIMAX
  PRM%POISSON_MODE = POISSON_MODE; PRM%CHECK_EVERY = 8; PRM%AF_MODE = AF_MODE
  PRM%AF_TOL = 0.D0
  CALL FVM_CHECK(FVM_CREATE(PRM, XG, YG, ZG, CTX), 'fvm_create')   ! DEFINE/GRID/METRIC/INIT
  CALL FVM_CHECK(FVM_SET_FIELD(CTX, FVM_U, U), 'fvm_set_field')
  CALL FVM_CHECK(FVM_SET_FIELD(CTX, FVM_V, V), 'fvm_set_field')
  CALL FVM_CHECK(FVM_SET_FIELD(CTX, FVM_W, W), 'fvm_set_field')
  CALL FVM_CHECK(FVM_SET_FIELD(CTX, FVM_P, P), 'fvm_set_field')
  CALL FVM_CHECK(FVM_SET_FIELD(CTX, FVM_NLX, NLX), 'fvm_set_field')
  CALL FVM_CHECK(FVM_SET_FIELD(CTX, FVM_NLY, NLY), 'fvm_set_field')
  CALL FVM_CHECK(FVM_SET_FIELD(CTX, FVM_NLZ, NLZ), 'fvm_set_field')
  IF (THERMAL .EQ. 1) THEN
    CALL FVM_CHECK(FVM_SET_THERMAL(CTX, PE), 'fvm_set_thermal')
    CALL FVM_CHECK(FVM_SET_FIELD(CTX, FVM_T, T), 'fvm_set_field')
    CALL FVM_CHECK(FVM_SET_FIELD(CTX, FVM_ADV, ADV), 'fvm_set_field')
  END IF

  INQUIRE (FILE='NU_CD_CL.PLT', EXIST=EXISTS)
  IF (.NOT. EXISTS) THEN
    OPEN (NEWUNIT=IU, FILE='NU_CD_CL.PLT', STATUS='NEW')
    WRITE (IU, *) 'variables=TIME,NU,CD,CL,CPB,CD_P,CD_V,CL_P,CL_V,CFL'
    CLOSE (IU)
    OPEN (NEWUNIT=IU, FILE='CS.PLT', STATUS='REPLACE')
    WRITE (IU, *) 'variables=TIME,Cs'
    CLOSE (IU)
    OPEN (NEWUNIT=IU, FILE='CT.PLT', STATUS='REPLACE')
    WRITE (IU, *) 'variables=TIME,Ct'
    CLOSE (IU)
  END IF
  NU = 0.D0

  ISTEP = 0
  DO                                                                     ! main.f90:35,161-166
  DO WHILE (RTIME .LT. RTIME_MAX .AND. ISTEP .LT. NSTEP)
    ISTEP = ISTEP + 1
    NITER_TIME = NITER_TIME + 1
    RTIME = RTIME + DTIME
    CALL FVM_CHECK(FVM_BEGIN_STEP(CTX), 'U_N = U')                        ! main.f90:42-45
    CALL FVM_CHECK(FVM_CONTRA_VEL(CTX), 'CONTRA_VEL')                     ! main.f90:48
    CALL FVM_CHECK(FVM_NL(CTX), 'NL')                                     ! main.f90:51-57
    C_AVG = 0.D0
    IF (USE_LES .EQ. 1) CALL FVM_CHECK(FVM_STRESS_TENSOR_DY(CTX, C_AVG), 'STRESS_TENSOR_DY')
    CALL FVM_CHECK(FVM_DIF(CTX), 'DIF')                                   ! main.f90:72-75
    CALL FVM_CHECK(FVM_P_SOURCE(CTX), 'P_SOURCE')                         ! main.f90:78
    CALL FVM_CHECK(FVM_ITER_U(CTX), 'ITER_U')                             ! main.f90:82-88
    CALL FVM_CHECK(FVM_CORRECT(CTX, -1.D0), 'predictor correction + BCU') ! main.f90:90-98
    CALL FVM_CHECK(FVM_CONTRA_VEL(CTX), 'CONTRA_VEL')                     ! main.f90:101
    CALL FVM_CHECK(FVM_ITER_P(CTX, RESIP, NITER), 'ITER_P + BCP')         ! main.f90:103-105
    CALL FVM_CHECK(FVM_P_SOURCE(CTX), 'P_SOURCE')                         ! main.f90:108
    CALL FVM_CHECK(FVM_CORRECT(CTX, 1.D0), 'projection + BCU')            ! main.f90:109-111,137-142
    CALL FVM_CHECK(FVM_PRESSURE_REFERENCE(CTX, PREF), 'pressure reference') ! main.f90:125-135
    IF (THERMAL .EQ. 1) CALL FVM_CHECK(FVM_ITER_T(CTX), 'ITER_T + BCT')  ! main.f90:145-147
    CALL FVM_CHECK(FVM_CALCCOEF(CTX, INFO), 'CALCCOEF')                   ! main.f90:150
    IF (THERMAL .EQ. 1) NU = INFO%NU
    IF (USE_LES .EQ. 1) THEN   ! stress_tensor_dy.f90:519-522, 707-711
      OPEN (NEWUNIT=IU, FILE='CS.PLT', STATUS='OLD', ACCESS='APPEND')
      WRITE (IU, 107) RTIME, SQRT(C_AVG)
      CLOSE (IU)
      IF (THERMAL .EQ. 1) THEN
        OPEN (NEWUNIT=IU, FILE='CT.PLT', STATUS='OLD', ACCESS='APPEND')
        WRITE (IU, 107) RTIME, SQRT(INFO%C_T)
        CLOSE (IU)
      END IF
    END IF

    OPEN (NEWUNIT=IU, FILE='NU_CD_CL.PLT', STATUS='OLD', ACCESS='APPEND')
    WRITE (IU, 106) RTIME, NU, INFO%CD, INFO%CL, INFO%CPB, INFO%CD_P, INFO%CD_V, INFO%CL_P, &
         INFO%CL_V, INFO%CFL
    CLOSE (IU)
    PRINT *, ''
    WRITE (*, 101) RTIME, NITER_TIME, INFO%CFL
    WRITE (*, 102) NU
    WRITE (*, 103) INFO%CD, INFO%CD_P, INFO%CD_V
    WRITE (*, 104) INFO%CL, INFO%CL_P, INFO%CL_V, INFO%CPB
    WRITE (*, 105) NITER, RESIP, C_AVG, PREF
    IF (MOD(NITER_TIME, NITER_W) .EQ. 0) CALL WRITE_FIELDS()              ! main.f90:154
    IF (MOD(NITER_TIME, NITER_R) .EQ. 0) CALL WRITE_RESTART()
  END DO
  ! POST_PROCESS (main.f90:164-166): its first call starts the averaging phase, its second
  ! writes the averages
  IF (IPOST .EQ. 1) THEN
    CALL POST_END()
    EXIT
  END IF
  IF (RTIME_END .LE. RTIME_MAX .OR. ISTEP .GE. NSTEP) EXIT
  CALL POST_BEGIN()
  END DO
  CALL WRITE_RESTART()
  CALL FVM_CHECK(FVM_DESTROY(CTX), 'fvm_destroy')

101 FORMAT(TR6,'Time',F10.3,TR12,'ITER',I10,TR12,'CFL',F10.3)
102 FORMAT(TR1,'NU =',F10.6)
103 FORMAT(TR1,'CD =',F10.6,TR4,'CD_P =',F10.6,TR4,'CD_V =',F10.6)
104 FORMAT(TR1,'CL =',F10.6,TR4,'CL_P =',F10.6,TR4,'CL_V =',F10.6,TR4,'CPB =',F10.6)
105 FORMAT(TR1,'NITER =',I8,TR4,'RESIP =',ES12.4,TR4,'C =',ES12.4,TR4,'PREF =',ES12.4)
106 FORMAT(10(1X,ES24.16E3))
107 FORMAT(2(1X,ES24.16E3))
122 FORMAT(TR1,E24.16,TR1,E24.16,TR1,E24.16,TR1,E24.16,TR1,I16)
128 FORMAT(TR1,E24.16,TR1,E24.16,TR1,E24.16,TR1,E24.16)
129 FORMAT(TR1,E24.16,TR1,E24.16,TR1,E24.16,TR1,E24.16,TR1,E24.16)

CONTAINS

  ! Tecplot output of WRITEDATA (writedata.f90:8-189), one rank (the rank field of each name
  ! is 000): the wall y+ on the wall cells J = 1..NJ-1, K = 1..NK-1 to Ypl_nnn_000.PLT (the
  ! reference composes this name but opens its CONTOUR variable, writedata.f90:14-20), and the
  ! nodal fields on the nodes I = 1..NI, J = 1..NJ, K = 1..NK to PVT_nnn_000.PLT and
  ! VOR_nnn_000.PLT
  SUBROUTINE WRITE_FIELDS()
    INTEGER :: JU, I, J, K, C, N
    DOUBLE PRECISION, ALLOCATABLE :: F(:,:)
    CHARACTER(LEN=15) :: FNAME
    INTEGER(C_INT) :: SEL(11)
    SEL = (/ FVM_PG, FVM_UG, FVM_VG, FVM_WG, FVM_TG, FVM_VORG, FVM_LAMDAG, FVM_X, FVM_Y, &
             FVM_Z, FVM_YPLUS /)
    ALLOCATE (F(NCELL, 11))
    CALL FVM_CHECK(FVM_WRITEDATA_FIELDS(CTX), 'WRITEDATA fields')
    DO N = 1, 11
      CALL FVM_CHECK(FVM_GET_FIELD(CTX, SEL(N), F(:, N)), 'fvm_get_field')
    END DO
    NUM_WRITE = NUM_WRITE + 1
    FNAME = 'Ypl_000_000.PLT'
    WRITE (FNAME(5:7), '(I3.3)') NUM_WRITE
    OPEN (NEWUNIT=JU, FILE=FNAME, STATUS='REPLACE')
    WRITE (JU, *) 'TITLE="YPLUS"'
    WRITE (JU, *) 'variables="x","y","z","Yplus"'
    WRITE (JU, 132) 1, NJ - 1, NK - 1, RTIME
    DO K = 1, NK - 1
      DO J = 1, NJ - 1
        C = 2 + (NI + 1) * (J + (NJ + 1) * K)
        WRITE (JU, 137) F(C, 8), F(C, 9), F(C, 10), F(C, 11)
      END DO
    END DO
    CLOSE (JU)
    FNAME = 'PVT_000_000.PLT'
    WRITE (FNAME(5:7), '(I3.3)') NUM_WRITE
    OPEN (NEWUNIT=JU, FILE=FNAME, STATUS='REPLACE')
    WRITE (JU, *) 'TITLE="', 'RE', RE, 'PR', PR, '"'
    WRITE (JU, *) 'variables=x,y,z,p,u,v,w,T'
    WRITE (JU, 156) NI, NJ, NK, RTIME
    DO K = 1, NK
      DO J = 1, NJ
        DO I = 1, NI
          C = 1 + I + (NI + 1) * (J + (NJ + 1) * K)
          Q = I + NI * ((J - 1) + NJ * (K - 1))
          WRITE (JU, 161) XG(Q), YG(Q), ZG(Q), F(C, 1), F(C, 2), F(C, 3), F(C, 4), F(C, 5)
        END DO
      END DO
    END DO
    CLOSE (JU)
    FNAME = 'VOR_000_000.PLT'
    WRITE (FNAME(5:7), '(I3.3)') NUM_WRITE
    OPEN (NEWUNIT=JU, FILE=FNAME, STATUS='REPLACE')
    WRITE (JU, *) 'TITLE="', 'RE', RE, 'PR', PR, '"'
    WRITE (JU, *) 'variables=x,y,z,vorM,lamda'
    WRITE (JU, 156) NI, NJ, NK, RTIME
    DO K = 1, NK
      DO J = 1, NJ
        DO I = 1, NI
          C = 1 + I + (NI + 1) * (J + (NJ + 1) * K)
          Q = I + NI * ((J - 1) + NJ * (K - 1))
          WRITE (JU, 183) XG(Q), YG(Q), ZG(Q), F(C, 6), F(C, 7)
        END DO
      END DO
    END DO
    CLOSE (JU)
    DEALLOCATE (F)
132 FORMAT(TR1,'ZONE T="POST", I=',I6,', j=',I6,', k=',I6,', SOLUTIONTIME=',F10.4)
137 FORMAT(TR1,D16.9,TR1,D16.9,TR1,D16.9,TR1,D16.9)
156 FORMAT(TR1,'ZONE T="WHOLE", I=',I6,', J=',I6,', K=',I6,', SOLUTIONTIME=',F10.4)
161 FORMAT(TR1,D16.9,TR1,D16.9,TR1,D16.9,TR1,D16.9,TR1,D16.9,TR1,D16.9,TR1,D16.9,TR1,D16.9)
183 FORMAT(TR1,D16.9,TR1,D16.9,TR1,D16.9,TR1,D16.9,TR1,D16.9)
  END SUBROUTINE WRITE_FIELDS

  ! POST_PROCESS's first call (post_process.f90:9-61): zeroed sums, CALCCOEF accumulates from
  ! the next step on, RTIME_MAX = RTIME_END, DTIME = DTIME_POST.  Without the passive
  ! temperature the restart's T is the one averaged.
  SUBROUTINE POST_BEGIN()
    IF (THERMAL .EQ. 1) THEN
      CALL FVM_CHECK(FVM_POST_BEGIN(CTX, C_NULL_PTR), 'POST_PROCESS')
    ELSE
      CALL FVM_CHECK(FVM_POST_BEGIN(CTX, C_LOC(T)), 'POST_PROCESS')
    END IF
    RTIME_MAX = RTIME_END
    DTIME = DTIME_POST
    CALL FVM_CHECK(FVM_SET_DTIME(CTX, DTIME), 'DTIME = DTIME_POST')
    IPOST = 1
  END SUBROUTINE POST_BEGIN

  ! POST_PROCESS's second call (post_process.f90:62-276): the node averages of the means
  ! (Avrg000.PLT), of the rms (Urms000.PLT) and of the Reynolds stresses with the SGS part
  ! (Strs000.PLT) on the nodes I = 1..NI, J = 1..NJ, K = 1..NK, and the wall coefficients
  ! on the wall cells J = 1..NJ-1, K = 1..NK-1 (Surf000.PLT; THETA in radians, as the
  ! reference's degree conversion at post_process.f90:259 addresses no wall cell)
  SUBROUTINE POST_END()
    INTEGER :: JU, I, J, K, C, N
    INTEGER(C_INT) :: NPOS, SEL(21)
    DOUBLE PRECISION, ALLOCATABLE :: F(:,:)
    SEL = (/ FVM_AVG_PG, FVM_AVG_UG, FVM_AVG_VG, FVM_AVG_WG, FVM_AVG_TG, FVM_RMS_UG, &
             FVM_RMS_VG, FVM_RMS_WG, FVM_RS_UU, FVM_RS_UV, FVM_RS_UW, FVM_RS_VV, FVM_RS_VW, &
             FVM_RS_WW, FVM_SURF_CF, FVM_SURF_CP, FVM_SURF_NU, FVM_X, FVM_Y, FVM_Z, FVM_THETA /)
    CALL FVM_CHECK(FVM_POST_FINALIZE(CTX, NPOS), 'POST_PROCESS')
    ALLOCATE (F(NCELL, 21))
    DO N = 1, 21
      CALL FVM_CHECK(FVM_GET_FIELD(CTX, SEL(N), F(:, N)), 'fvm_get_field')
    END DO
    OPEN (NEWUNIT=JU, FILE='Avrg000.PLT', STATUS='REPLACE')
    WRITE (JU, *) 'TITLE="', 'RE', RE, 'PR', PR, '"'
    WRITE (JU, *) 'variables=x,y,z,p,u,v,w,T'
    WRITE (JU, 68) NI, NJ, NK, RTIME
    DO K = 1, NK
      DO J = 1, NJ
        DO I = 1, NI
          C = 1 + I + (NI + 1) * (J + (NJ + 1) * K)
          Q = I + NI * ((J - 1) + NJ * (K - 1))
          WRITE (JU, 73) XG(Q), YG(Q), ZG(Q), (F(C, N), N = 1, 5)
        END DO
      END DO
    END DO
    CLOSE (JU)
    OPEN (NEWUNIT=JU, FILE='Urms000.PLT', STATUS='REPLACE')
    WRITE (JU, *) 'TITLE="', 'RE', RE, 'PR', PR, '"'
    WRITE (JU, *) 'variables=x,y,z,u_rms,v_rms,w_rms'
    WRITE (JU, 68) NI, NJ, NK, RTIME
    DO K = 1, NK
      DO J = 1, NJ
        DO I = 1, NI
          C = 1 + I + (NI + 1) * (J + (NJ + 1) * K)
          Q = I + NI * ((J - 1) + NJ * (K - 1))
          WRITE (JU, 109) XG(Q), YG(Q), ZG(Q), (F(C, N), N = 6, 8)
        END DO
      END DO
    END DO
    CLOSE (JU)
    OPEN (NEWUNIT=JU, FILE='Strs000.PLT', STATUS='REPLACE')
    WRITE (JU, *) 'TITLE="', 'RE', RE, 'PR', PR, '"'
    WRITE (JU, *) 'variables=x,y,z,uu,uv,uw,vv,vw,ww'
    WRITE (JU, 68) NI, NJ, NK, RTIME
    DO K = 1, NK
      DO J = 1, NJ
        DO I = 1, NI
          C = 1 + I + (NI + 1) * (J + (NJ + 1) * K)
          Q = I + NI * ((J - 1) + NJ * (K - 1))
          WRITE (JU, 249) XG(Q), YG(Q), ZG(Q), (F(C, N), N = 9, 14)
        END DO
      END DO
    END DO
    CLOSE (JU)
    OPEN (NEWUNIT=JU, FILE='Surf000.PLT', STATUS='REPLACE')
    WRITE (JU, *) 'TITLE="SURFACE COEF."'
    WRITE (JU, *) 'variables="x","y","z","Theta","Cf","Cp","Nu"'
    WRITE (JU, 132) 1, NJ - 1, NK - 1
    DO K = 1, NK - 1
      DO J = 1, NJ - 1
        C = 2 + (NI + 1) * (J + (NJ + 1) * K)
        WRITE (JU, 137) F(C, 18), F(C, 19), F(C, 20), F(C, 21), F(C, 15), F(C, 16), F(C, 17)
      END DO
    END DO
    CLOSE (JU)
    DEALLOCATE (F)
68  FORMAT(TR1,'ZONE T="WHOLE", I=',I6,', J=',I6,', K=',I6,', SOLUTIONTIME=',F10.4)
73  FORMAT(TR1,D16.9,TR1,D16.9,TR1,D16.9,TR1,D16.9,TR1,D16.9,TR1,D16.9,TR1,D16.9,TR1,D16.9)
109 FORMAT(TR1,D16.9,TR1,D16.9,TR1,D16.9,TR1,D16.9,TR1,D16.9,TR1,D16.9)
132 FORMAT(TR1,'ZONE T="POST", I=',I6,', j=',I6,', k=',I6)
137 FORMAT(TR1,D16.9,TR1,D16.9,TR1,D16.9,TR1,D16.9,TR1,D16.9,TR1,D16.9,TR1,D16.9)
249 FORMAT(TR1,D16.9,TR1,D16.9,TR1,D16.9,TR1,D16.9,TR1,D16.9,TR1,D16.9,TR1,D16.9,TR1,D16.9, &
           TR1,D16.9)
  END SUBROUTINE POST_END

  ! RESTART000.DAT in the reference's layout (writedata.f90:193-216)
  SUBROUTINE WRITE_RESTART()
    INTEGER :: JU, C
    CALL FVM_CHECK(FVM_GET_FIELD(CTX, FVM_U, U), 'fvm_get_field')
    CALL FVM_CHECK(FVM_GET_FIELD(CTX, FVM_V, V), 'fvm_get_field')
    CALL FVM_CHECK(FVM_GET_FIELD(CTX, FVM_W, W), 'fvm_get_field')
    CALL FVM_CHECK(FVM_GET_FIELD(CTX, FVM_P, P), 'fvm_get_field')
    CALL FVM_CHECK(FVM_GET_FIELD(CTX, FVM_NLX, NLX), 'fvm_get_field')
    CALL FVM_CHECK(FVM_GET_FIELD(CTX, FVM_NLY, NLY), 'fvm_get_field')
    CALL FVM_CHECK(FVM_GET_FIELD(CTX, FVM_NLZ, NLZ), 'fvm_get_field')
    IF (THERMAL .EQ. 1) THEN
      CALL FVM_CHECK(FVM_GET_FIELD(CTX, FVM_T, T), 'fvm_get_field')
      CALL FVM_CHECK(FVM_GET_FIELD(CTX, FVM_ADV, ADV), 'fvm_get_field')
    END IF
    OPEN (NEWUNIT=JU, FILE='RESTART000.DAT', STATUS='REPLACE')
    WRITE (JU, 203) RE, PE_F, DTIME, RTIME, NITER_TIME
    DO C = 1, NCELL
      WRITE (JU, 209) NLX(C), NLY(C), NLZ(C), ADV(C)
      WRITE (JU, 210) U(C), V(C), W(C), P(C), T(C)
    END DO
    CLOSE (JU)
203 FORMAT(TR1,E24.16,TR1,E24.16,TR1,E24.16,TR1,E24.16,TR1,I16)
209 FORMAT(TR1,E24.16,TR1,E24.16,TR1,E24.16,TR1,E24.16)
210 FORMAT(TR1,E24.16,TR1,E24.16,TR1,E24.16,TR1,E24.16,TR1,E24.16)
  END SUBROUTINE WRITE_RESTART

END PROGRAM FVM_RLES
