!  Trajectory harness for the reference 2D solver (cylinder/cylinder_v3.0.3_AF_original
!  4step_face_nl), test infrastructure built by build_ref2d.sh into oracle/_ref/.
!
!  From RESTART.DAT it runs the reference time loop (Main.f90:48-110, the passive thermal scalar
!  left out) for NSTEP steps with the post-processing phase on from the first step: POST_PROCESS's
!  first call opens POST_NU_CD_CL.PLT and sets IPOST, so CALCCOEF writes NU_CD_CL.PLT and the
!  running means, Strouhal number and rms values of POST_NU_CD_CL.PLT every step
!  (calccoef.f90:114-195); POST_PROCESS's second call writes TIME_AVERAGED_RESULTS.DAT.
!
!  usage: harness2d_post NSTEP   (run directory holds NNNxNNN.dat and RESTART.DAT)
PROGRAM HARNESS2D_POST
  USE COMDAT_SHARED
  IMPLICIT NONE
  INTEGER :: NSTEP, ISTEP
  CHARACTER(64) :: ARG

  CALL GET_COMMAND_ARGUMENT(1, ARG)
  READ(ARG, *) NSTEP
  OPEN(UNIT=61, FILE='scalars.txt', STATUS='REPLACE')
  IRESTART = 1
  IPOST = 0
  CALL GRID
  CALL METRIC
  CALL INIT
  CALL POST_PROCESS

  DO ISTEP = 1, NSTEP
    NITER_TIME = NITER_TIME + 1
    RTIME = RTIME + DTIME
    U_N(:,:,:) = U(:,:,:)
    V_N(:,:,:) = V(:,:,:)
    T_N(:,:,:) = T(:,:,:)
    FLAG = 0
    CALL CONTRA_VEL
    NLX_NM1 = NLX
    NLY_NM1 = NLY
    ADV_NM1 = ADV
    CALL NL(U_N, NLX)
    CALL NL(V_N, NLY)
    CALL DIF(U_N, DIFX)
    CALL DIF(V_N, DIFY)
    CALL P_SOURCE
    CALL ITER_U
    CALL ITER_V
    U(SX1:EX1,SY1:EY1,2:NKM1) = U(SX1:EX1,SY1:EY1,2:NKM1) &
         - S_U(SX1:EX1,SY1:EY1,2:NKM1)*DTIME/DJAC(SX1:EX1,SY1:EY1,2:NKM1)
    V(SX1:EX1,SY1:EY1,2:NKM1) = V(SX1:EX1,SY1:EY1,2:NKM1) &
         - S_V(SX1:EX1,SY1:EY1,2:NKM1)*DTIME/DJAC(SX1:EX1,SY1:EY1,2:NKM1)
    CALL BCU
    CALL BCV
    FLAG = 0
    CALL CONTRA_VEL
    CALL ITER_P
    CALL BCP
    CALL P_SOURCE
    U(SX1:EX1,SY1:EY1,2:NKM1) = U(SX1:EX1,SY1:EY1,2:NKM1) &
         + S_U(SX1:EX1,SY1:EY1,2:NKM1)*DTIME/DJAC(SX1:EX1,SY1:EY1,2:NKM1)
    V(SX1:EX1,SY1:EY1,2:NKM1) = V(SX1:EX1,SY1:EY1,2:NKM1) &
         + S_V(SX1:EX1,SY1:EY1,2:NKM1)*DTIME/DJAC(SX1:EX1,SY1:EY1,2:NKM1)
    CALL BCU
    CALL BCV
    CALL CALCCOEF
    WRITE(61, '(A,I6,I9,9ES26.17E3)') 'step ', ISTEP, NITER, RESIP, CFL_MAX, CPB, CD, CL, &
         CD_P, CD_V, CL_P, CL_V
    FLUSH(61)
  ENDDO
  CLOSE(61)
  CALL POST_PROCESS
END PROGRAM HARNESS2D_POST
