!  Test harness for the reference 2D solver (cylinder/cylinder_v3.0.3_AF_original 4step_face_nl).
!
!  This is test infrastructure: it is linked against the reference's own subroutines (built
!  from /root/reference by build_ref2d.sh into oracle/_ref/) and calls them in the order of the
!  reference time loop (Main.f90:48-110; the passive thermal scalar and file output left out).
!  After each step it dumps U, V, P of the K = 2 plane in raw binary (stream access,
!  column-major, ghosts included) and the CALCCOEF scalars.
!
!  usage: harness2d NSTEP   (run directory holds NNNxNNN.dat and RESTART.DAT; stdin answers
!                            the PAUSE statements of INIT and CALCCOEF)
PROGRAM HARNESS2D
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

  CALL DUMP('setup_x', X); CALL DUMP('setup_y', Y); CALL DUMP('setup_djac', DJAC)
  CALL DUMP('setup_xix', XIX); CALL DUMP('setup_xiy', XIY)
  CALL DUMP('setup_etx', ETX); CALL DUMP('setup_ety', ETY)
  CALL DUMP('setup_theta', THETA)

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
    CALL DUMPS(ISTEP, 'uhat', U)
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
    CALL DUMPS(ISTEP, 'u', U); CALL DUMPS(ISTEP, 'v', V); CALL DUMPS(ISTEP, 'p', P)
    CALL DUMPS(ISTEP, 'nlx', NLX); CALL DUMPS(ISTEP, 'difx', DIFX)
    WRITE(61, '(A,I6,I9,9ES26.17E3)') 'step ', ISTEP, NITER, RESIP, CFL_MAX, CPB, CD, CL, &
         CD_P, CD_V, CL_P, CL_V
  ENDDO
  CLOSE(61)

CONTAINS

  SUBROUTINE DUMP(NAME, A)
    CHARACTER(*), INTENT(IN) :: NAME
    DOUBLE PRECISION, INTENT(IN) :: A(:,:,:)
    INTEGER :: IU
    OPEN(NEWUNIT=IU, FILE=NAME//'.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
    WRITE(IU) A
    CLOSE(IU)
  END SUBROUTINE DUMP

  SUBROUTINE DUMPS(ISTEP, NAME, A)
    INTEGER, INTENT(IN) :: ISTEP
    CHARACTER(*), INTENT(IN) :: NAME
    DOUBLE PRECISION, INTENT(IN) :: A(:,:,:)
    CHARACTER(16) :: PFX
    WRITE(PFX, '(A,I3.3,A)') 's', ISTEP, '_'
    CALL DUMP(TRIM(PFX)//NAME, A)
  END SUBROUTINE DUMPS

END PROGRAM HARNESS2D
