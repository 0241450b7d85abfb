!  Test harness for the reference 3D LES solver (LES/v2.3_LES-Reynolds stresses).
!
!  This is test infrastructure: it is linked against the reference's own subroutines (built
!  from /root/reference by build_ref.sh into oracle/_ref/) and calls them in the order of the
!  reference time loop (LES/v2.3_LES-Reynolds stresses/main.f90:39-150, the passive
!  temperature included; file output left out).  After each sub-stage it dumps the
!  module arrays in raw binary (stream access, column-major, ghosts included) so the
!  project's CPU restatement and GPU kernels can be compared with the reference stage by stage.
!
!  usage: harness3d NSTEP [NPOST [NPRE]]   (run directory holds NNNxNNNxNNN.dat and
!         RESTART000.DAT)
!
!  With NPRE > 0 the first NPRE steps run at DTIME/10 (a start-up from the impulsive start; the
!  reference reassigns DTIME at run time the same way, post_process.f90:59), the rest at DTIME.
!
!  With NPOST > 0 the last NPOST steps run in the time-averaging phase of POST_PROCESS (its
!  first call before step NSTEP-NPOST+1: sums zeroed, DTIME = DTIME_POST; CALCCOEF accumulates;
!  its second call after the last step writes Avrg000/Urms000/Strs000/Surf000.PLT).  The sums
!  are dumped before that second call and its in-place results after it; WRITEDATA is skipped
!  in this mode (post_process.f90:259 scales THETA at the loop indices one past the array).
PROGRAM HARNESS3D
  USE COMDAT_SHARED
  IMPLICIT NONE
  INCLUDE 'mpif.h'
  INTEGER :: NSTEP, ISTEP, NPOST, IST, NPRE
  DOUBLE PRECISION :: PREF, DT_RUN
  CHARACTER(64) :: ARG

  CALL GET_COMMAND_ARGUMENT(1, ARG)
  READ(ARG, *) NSTEP
  NPOST = 0
  CALL GET_COMMAND_ARGUMENT(2, ARG, STATUS=IST)
  IF (IST .EQ. 0 .AND. LEN_TRIM(ARG) .GT. 0) READ(ARG, *) NPOST
  NPRE = 0
  CALL GET_COMMAND_ARGUMENT(3, ARG, STATUS=IST)
  IF (IST .EQ. 0 .AND. LEN_TRIM(ARG) .GT. 0) READ(ARG, *) NPRE

  ! files the reference appends to with STATUS='OLD'
  OPEN(UNIT=11, FILE='NU_CD_CL.PLT', STATUS='REPLACE'); CLOSE(11)
  OPEN(UNIT=12, FILE='CS.PLT', STATUS='REPLACE'); CLOSE(12)
  OPEN(UNIT=13, FILE='CT.PLT', STATUS='REPLACE'); CLOSE(13)
  OPEN(UNIT=61, FILE='scalars.txt', STATUS='REPLACE')

  CALL DEFINE
  CALL GRID
  CALL METRIC
  CALL INIT

  CALL DUMP('setup_x', X); CALL DUMP('setup_y', Y); CALL DUMP('setup_z', Z)
  CALL DUMP('setup_djac', DJAC)
  CALL DUMP('setup_xix', XIX); CALL DUMP('setup_xiy', XIY); CALL DUMP('setup_xiz', XIZ)
  CALL DUMP('setup_etx', ETX); CALL DUMP('setup_ety', ETY); CALL DUMP('setup_etz', ETZ)
  CALL DUMP('setup_ztx', ZTX); CALL DUMP('setup_zty', ZTY); CALL DUMP('setup_ztz', ZTZ)
  CALL DUMP('setup_q12', Q12); CALL DUMP('setup_q13', Q13); CALL DUMP('setup_q23', Q23)
  CALL DUMP('setup_de', DE); CALL DUMP('setup_dn', DN); CALL DUMP('setup_dt', DT)
  CALL DUMP('setup_theta', THETA)
  WRITE(61, '(A,2ES26.17E3)') 'area ', SUM(AREA1), SUM(AREA2)

  DT_RUN = DTIME
  DO ISTEP = 1, NSTEP
    IF (NPRE .GT. 0) THEN
      DTIME = DT_RUN
      IF (ISTEP .LE. NPRE) DTIME = 0.1D0*DT_RUN
    ENDIF
    IF (NPOST .GT. 0 .AND. ISTEP .EQ. NSTEP - NPOST + 1) CALL POST_PROCESS
    NITER_TIME = NITER_TIME + 1
    RTIME = RTIME + DTIME
    U_N(IS:IE,JS:JE,KS:KE) = U(IS:IE,JS:JE,KS:KE)
    V_N(IS:IE,JS:JE,KS:KE) = V(IS:IE,JS:JE,KS:KE)
    W_N(IS:IE,JS:JE,KS:KE) = W(IS:IE,JS:JE,KS:KE)
    T_N(IS:IE,JS:JE,KS:KE) = T(IS:IE,JS:JE,KS:KE)

    CALL CONTRA_VEL
    CALL DUMPS(ISTEP, 'ue', UE); CALL DUMPS(ISTEP, 'vn', VN); CALL DUMPS(ISTEP, 'wt', WT)

    NLX_NM1(IS1:IE1,JS1:JE1,KS1:KE1) = NLX(IS1:IE1,JS1:JE1,KS1:KE1)
    NLY_NM1(IS1:IE1,JS1:JE1,KS1:KE1) = NLY(IS1:IE1,JS1:JE1,KS1:KE1)
    NLZ_NM1(IS1:IE1,JS1:JE1,KS1:KE1) = NLZ(IS1:IE1,JS1:JE1,KS1:KE1)
    ADV_NM1(IS1:IE1,JS1:JE1,KS1:KE1) = ADV(IS1:IE1,JS1:JE1,KS1:KE1)
    CALL NL(U_N, NLX)
    CALL NL(V_N, NLY)
    CALL NL(W_N, NLZ)
    CALL NL(T_N, ADV)
    CALL DUMPS(ISTEP, 'nl0x', NLX); CALL DUMPS(ISTEP, 'nl0y', NLY); CALL DUMPS(ISTEP, 'nl0z', NLZ)
    CALL DUMPS(ISTEP, 'adv0', ADV)
    IF (USE_LES .EQ. 1) THEN
      CALL STRESS_TENSOR_DY
      CALL DUMPS(ISTEP, 'sgsx', SGS_X); CALL DUMPS(ISTEP, 'sgsy', SGS_Y)
      CALL DUMPS(ISTEP, 'sgsz', SGS_Z); CALL DUMPS(ISTEP, 'nut', EDDY_VISCOS)
      CALL DUMPS(ISTEP, 'smag', MAGNITUDE_LSS)
      NLX(IS1:IE1,JS1:JE1,KS1:KE1) = NLX(IS1:IE1,JS1:JE1,KS1:KE1) &
           + SGS_X(IS1:IE1,JS1:JE1,KS1:KE1)*DJAC(IS1:IE1,JS1:JE1,KS1:KE1)
      NLY(IS1:IE1,JS1:JE1,KS1:KE1) = NLY(IS1:IE1,JS1:JE1,KS1:KE1) &
           + SGS_Y(IS1:IE1,JS1:JE1,KS1:KE1)*DJAC(IS1:IE1,JS1:JE1,KS1:KE1)
      NLZ(IS1:IE1,JS1:JE1,KS1:KE1) = NLZ(IS1:IE1,JS1:JE1,KS1:KE1) &
           + SGS_Z(IS1:IE1,JS1:JE1,KS1:KE1)*DJAC(IS1:IE1,JS1:JE1,KS1:KE1)
      ADV(IS1:IE1,JS1:JE1,KS1:KE1) = ADV(IS1:IE1,JS1:JE1,KS1:KE1) &
           + SGS_T(IS1:IE1,JS1:JE1,KS1:KE1)*DJAC(IS1:IE1,JS1:JE1,KS1:KE1)
      CALL DUMPS(ISTEP, 'sgst', SGS_T); CALL DUMPS(ISTEP, 'kappa', EDDY_CONDUC)
      CALL DUMPS(ISTEP, 'that', T_HAT); CALL DUMPS(ISTEP, 'dtdx', DTDX)
      CALL DUMPS(ISTEP, 'dtdy', DTDY); CALL DUMPS(ISTEP, 'dtdz', DTDZ)
    ENDIF
    CALL DUMPS(ISTEP, 'adv', ADV)
    CALL DUMPS(ISTEP, 'nlx', NLX); CALL DUMPS(ISTEP, 'nly', NLY); CALL DUMPS(ISTEP, 'nlz', NLZ)

    CALL DIF(U_N, DIFX)
    CALL DIF(V_N, DIFY)
    CALL DIF(W_N, DIFZ)
    CALL DIF(T_N, DIFT)
    CALL DUMPS(ISTEP, 'difx', DIFX); CALL DUMPS(ISTEP, 'dify', DIFY); CALL DUMPS(ISTEP, 'difz', DIFZ)
    CALL DUMPS(ISTEP, 'dift', DIFT)

    CALL P_SOURCE(P)
    CALL DUMPS(ISTEP, 'su0', S_U); CALL DUMPS(ISTEP, 'sv0', S_V); CALL DUMPS(ISTEP, 'sw0', S_W)

    UIN = 1.D0
    CALL ITER_U(U, U_N, NLX, NLX_NM1, DIFX, S_U)
    UIN = 0.D0
    CALL ITER_U(V, V_N, NLY, NLY_NM1, DIFY, S_V)
    CALL ITER_U(W, W_N, NLZ, NLZ_NM1, DIFZ, S_W)
    CALL DUMPS(ISTEP, 'ustar', U); CALL DUMPS(ISTEP, 'vstar', V); CALL DUMPS(ISTEP, 'wstar', W)

    U(IS1:IE1,JS1:JE1,KS1:KE1) = U(IS1:IE1,JS1:JE1,KS1:KE1)-DTIME*S_U(IS1:IE1,JS1:JE1,KS1:KE1)/DJAC(IS1:IE1,JS1:JE1,KS1:KE1)
    V(IS1:IE1,JS1:JE1,KS1:KE1) = V(IS1:IE1,JS1:JE1,KS1:KE1)-DTIME*S_V(IS1:IE1,JS1:JE1,KS1:KE1)/DJAC(IS1:IE1,JS1:JE1,KS1:KE1)
    W(IS1:IE1,JS1:JE1,KS1:KE1) = W(IS1:IE1,JS1:JE1,KS1:KE1)-DTIME*S_W(IS1:IE1,JS1:JE1,KS1:KE1)/DJAC(IS1:IE1,JS1:JE1,KS1:KE1)
    UIN = 1.D0
    CALL BCU(U, U_N)
    UIN = 0.D0
    CALL BCU(V, V_N)
    CALL BCU(W, W_N)
    CALL DUMPS(ISTEP, 'uhat', U); CALL DUMPS(ISTEP, 'vhat', V); CALL DUMPS(ISTEP, 'what', W)

    CALL CONTRA_VEL
    CALL ITER_P
    CALL BCP
    CALL DUMPS(ISTEP, 'p', P)

    CALL P_SOURCE(P)
    U(IS1:IE1,JS1:JE1,KS1:KE1) = U(IS1:IE1,JS1:JE1,KS1:KE1)+DTIME*S_U(IS1:IE1,JS1:JE1,KS1:KE1)/DJAC(IS1:IE1,JS1:JE1,KS1:KE1)
    V(IS1:IE1,JS1:JE1,KS1:KE1) = V(IS1:IE1,JS1:JE1,KS1:KE1)+DTIME*S_V(IS1:IE1,JS1:JE1,KS1:KE1)/DJAC(IS1:IE1,JS1:JE1,KS1:KE1)
    W(IS1:IE1,JS1:JE1,KS1:KE1) = W(IS1:IE1,JS1:JE1,KS1:KE1)+DTIME*S_W(IS1:IE1,JS1:JE1,KS1:KE1)/DJAC(IS1:IE1,JS1:JE1,KS1:KE1)

    DUM_B = 0.D0
    IF (JS .LT. NJ/4 .AND. JE .GT. NJ/4) THEN
      DO K = KS1, KE1
        DUM_B = DUM_B + P(IE, NJ/4, K)
      ENDDO
      DUM_B = DUM_B/DBLE(KE1-KS1+1)
    ENDIF
    CALL MPI_ALLREDUCE(DUM_B, DUM_A, 1, MPI_DOUBLE_PRECISION, MPI_SUM, NMCW, ERR)
    P(IS:IE,JS:JE,KS:KE) = P(IS:IE,JS:JE,KS:KE) - DUM_A/NDIMK
    PREF = DUM_A

    UIN = 1.D0
    CALL BCU(U, U_N)
    UIN = 0.D0
    CALL BCU(V, V_N)
    CALL BCU(W, W_N)

    CALL ITER_T
    CALL BCT(T, T_N)

    CALL CALCCOEF
    CALL DUMPS(ISTEP, 'u', U); CALL DUMPS(ISTEP, 'v', V); CALL DUMPS(ISTEP, 'w', W)
    CALL DUMPS(ISTEP, 'pref', P); CALL DUMPS(ISTEP, 't', T)
    WRITE(61, '(A,I6,I8,10ES26.17E3)') 'step ', ISTEP, NITER, RESIP, PREF, CFL_MAX, CPB, &
         CD, CL, CD_P, CD_V, NU
  ENDDO
  CLOSE(61)

  IF (NPOST .GT. 0) THEN
    CALL DUMP('post_sum_u', U_AVG); CALL DUMP('post_sum_v', V_AVG)
    CALL DUMP('post_sum_w', W_AVG); CALL DUMP('post_sum_p', P_AVG)
    CALL DUMP('post_sum_t', T_AVG)
    CALL DUMP('post_sum_uu', UU_AVG); CALL DUMP('post_sum_uv', UV_AVG)
    CALL DUMP('post_sum_uw', UW_AVG); CALL DUMP('post_sum_vv', VV_AVG)
    CALL DUMP('post_sum_vw', VW_AVG); CALL DUMP('post_sum_ww', WW_AVG)
    CALL DUMP('post_sum_sgs11', SGS11); CALL DUMP('post_sum_sgs12', SGS12)
    CALL DUMP('post_sum_sgs13', SGS13); CALL DUMP('post_sum_sgs22', SGS22)
    CALL DUMP('post_sum_sgs23', SGS23); CALL DUMP('post_sum_sgs33', SGS33)
    OPEN(UNIT=63, FILE='post_niter.txt', STATUS='REPLACE'); WRITE(63, *) NITER_POS; CLOSE(63)
    CALL POST_PROCESS
    ! in-place results: the means, and the Reynolds stresses left in UU_AVG..WW_AVG
    CALL DUMP('post_fin_u', U_AVG); CALL DUMP('post_fin_v', V_AVG)
    CALL DUMP('post_fin_w', W_AVG); CALL DUMP('post_fin_p', P_AVG)
    CALL DUMP('post_fin_t', T_AVG)
    CALL DUMP('post_fin_uu', UU_AVG); CALL DUMP('post_fin_uv', UV_AVG)
    CALL DUMP('post_fin_uw', UW_AVG); CALL DUMP('post_fin_vv', VV_AVG)
    CALL DUMP('post_fin_vw', VW_AVG); CALL DUMP('post_fin_ww', WW_AVG)
    ! node averages of the Strs000.PLT pass (uu vv ww in UG VG WG, uv uw vw in U..W_RMS) and
    ! of the means of p and T (Avrg000.PLT); the other node averages are in the .PLT files
    CALL DUMPN('post_node_uu', UG); CALL DUMPN('post_node_vv', VG)
    CALL DUMPN('post_node_ww', WG); CALL DUMPN('post_node_pg', PG)
    CALL DUMPN('post_node_tg', TG)
    CALL DUMP('post_cell_uv', U_RMS); CALL DUMP('post_cell_uw', V_RMS)
    CALL DUMP('post_cell_vw', W_RMS)
    OPEN(UNIT=62, FILE='post_surf.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
    WRITE(62) SH_S, CP_S, NU_S
    CLOSE(62)
    CALL MPI_FINALIZE(ERR)
    STOP
  ENDIF

  ! field output of the last step (WRITEDATA, writedata.f90:8-189: CALC_VOR, lambda_2 by
  ! FIND_ROOT, nodal averages, Tecplot files).  VOR_M's I ghost planes are never assigned by the
  ! reference (calc_vor.f90:100-104 covers IS1:IE1 only) and LAMDA2 keeps its value where no
  ! root is strictly between the other two: both start from zero here.
  VOR_M = 0.D0
  LAMDA2 = 0.D0
  CONTOUR = 'YPLUS.PLT'
  NITER_W = 1
  NITER_R = HUGE(1)
  CALL WRITEDATA
  CALL DUMP('out_lamda2', LAMDA2); CALL DUMP('out_vorm', VOR_M)
  CALL DUMPN('out_pg', PG); CALL DUMPN('out_ug', UG); CALL DUMPN('out_vg', VG)
  CALL DUMPN('out_wg', WG); CALL DUMPN('out_tg', TG); CALL DUMPN('out_lamdag', LAMDAG)
  CALL DUMP('out_dum3d', DUM_3D)
  ! YPLUS of the last CALCCOEF: the wall plane only, YPLUS(IS1, JS1:JE1, KS1:KE1) (init.f90:81)
  OPEN(UNIT=62, FILE='out_yplus.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
  WRITE(62) YPLUS
  CLOSE(62)
  CALL MPI_FINALIZE(ERR)

CONTAINS

  SUBROUTINE DUMP(TAG, A)
    CHARACTER(*), INTENT(IN) :: TAG
    DOUBLE PRECISION, INTENT(IN) :: A(IS:IE,JS:JE,KS:KE)
    OPEN(UNIT=62, FILE=TAG//'.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
    WRITE(62) A
    CLOSE(62)
  END SUBROUTINE DUMP

  SUBROUTINE DUMPN(TAG, A)   ! node arrays (IS1:IE, JS1:JE, KS1:KE)
    CHARACTER(*), INTENT(IN) :: TAG
    DOUBLE PRECISION, INTENT(IN) :: A(IS1:IE,JS1:JE,KS1:KE)
    OPEN(UNIT=62, FILE=TAG//'.bin', ACCESS='STREAM', FORM='UNFORMATTED', STATUS='REPLACE')
    WRITE(62) A
    CLOSE(62)
  END SUBROUTINE DUMPN

  SUBROUTINE DUMPS(ISTEP, TAG, A)
    INTEGER, INTENT(IN) :: ISTEP
    CHARACTER(*), INTENT(IN) :: TAG
    DOUBLE PRECISION, INTENT(IN) :: A(IS:IE,JS:JE,KS:KE)
    CHARACTER(16) :: PFX
    WRITE(PFX, '(A,I3.3,A)') 's', ISTEP, '_'
    CALL DUMP(TRIM(PFX)//TAG, A)
  END SUBROUTINE DUMPS

END PROGRAM HARNESS3D
