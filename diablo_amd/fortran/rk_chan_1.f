C----*|--.---------.---------.---------.---------.---------.---------.-|-------|
C Drop-in RK_CHAN_1 (replaces channel.f:62-1073) running the substep on the
C GPU through libdiablo_hip.so.  Compiled in the reference's build tree with
C its header; link this object ahead of channel.o (whose RK_CHAN_1 symbol is
C weakened: objcopy --weaken-symbol=rk_chan_1_) together with diablo_hip.o
C and -ldiablo_hip.  The device keeps the state between substeps; it is copied
C back after RK_STEP=3 whenever the driver is about to read or rewrite it
C (SAVE_STATS / SAVE_FLOW / the last step / END_RUN, diablo.f:87-119), and
C copied in again before the next RK_STEP=1.
C----*|--.---------.---------.---------.---------.---------.---------.-|-------|
      SUBROUTINE RK_CHAN_1
      USE diablo_hip
      INCLUDE 'header'

      TYPE(C_PTR), SAVE :: CTX = C_NULL_PTR
      LOGICAL, SAVE :: HOST_DIRTY = .TRUE.
      INTEGER, SAVE :: TS_FIRST = 0
      TYPE(dh_params) :: PRM
      INTEGER(C_INT) :: RC
      INTEGER(C_INT32_T) :: TS
      REAL*8 :: TDEV, WT
      LOGICAL :: STOPFILE, NEEDED
      INTEGER :: N

      IF (.NOT. C_ASSOCIATED(CTX)) THEN
        PRM%NX = NX
        PRM%NY = NY
        PRM%NZ = NZ
        PRM%N_TH = N_TH
        PRM%NU = NU
        PRM%BETA = BETA
        PRM%LX = LX
        PRM%LZ = LZ
        PRM%NU_V_SCALE = NU_V_SCALE
        PRM%DELTA_T = DELTA_T
        PRM%CFL = CFL
        PRM%VARIABLE_DT = MERGE(1, 0, VARIABLE_DT)
        PRM%UPDATE_DT = UPDATE_DT
        PRM%F_TYPE = F_TYPE
        PRM%PX0 = PX0
        PRM%OMEGA0 = OMEGA0
        PRM%AMP_OMEGA0 = AMP_OMEGA0
        PRM%I_RO = I_RO
        PRM%GRAV_X = GRAV_X
        PRM%GRAV_Y = GRAV_Y
        PRM%GRAV_Z = GRAV_Z
        PRM%U_BC_YMIN = U_BC_YMIN
        PRM%V_BC_YMIN = V_BC_YMIN
        PRM%W_BC_YMIN = W_BC_YMIN
        PRM%U_BC_YMAX = U_BC_YMAX
        PRM%V_BC_YMAX = V_BC_YMAX
        PRM%W_BC_YMAX = W_BC_YMAX
        PRM%U_BC_YMIN_C1 = U_BC_YMIN_C1
        PRM%V_BC_YMIN_C1 = V_BC_YMIN_C1
        PRM%W_BC_YMIN_C1 = W_BC_YMIN_C1
        PRM%U_BC_YMAX_C1 = U_BC_YMAX_C1
        PRM%V_BC_YMAX_C1 = V_BC_YMAX_C1
        PRM%W_BC_YMAX_C1 = W_BC_YMAX_C1
        PRM%TH_BC_YMIN = 0
        PRM%TH_BC_YMAX = 0
        PRM%TH_BC_YMIN_C1 = 0.D0
        PRM%TH_BC_YMAX_C1 = 0.D0
        PRM%RI = 0.D0
        PRM%PR = 1.D0
        PRM%DRHODX = 0.D0
        PRM%DRHODZ = 0.D0
        DO N=1,N_TH
          PRM%TH_BC_YMIN(N) = TH_BC_YMIN(N)
          PRM%TH_BC_YMAX(N) = TH_BC_YMAX(N)
          PRM%TH_BC_YMIN_C1(N) = TH_BC_YMIN_C1(N)
          PRM%TH_BC_YMAX_C1(N) = TH_BC_YMAX_C1(N)
          PRM%RI(N) = RI(N)
          PRM%PR(N) = PR(N)
          PRM%DRHODX(N) = DRHODX(N)
          PRM%DRHODZ(N) = DRHODZ(N)
        END DO
        PRM%CREATE_NEW_FLOW = MERGE(1, 0, CREATE_NEW_FLOW)
C The device carries the constant-Smagorinsky model only; les.f:269-272
C stops on the dynamic models too
        IF (LES .AND. LES_MODEL_TYPE .NE. 1)
     &    STOP 'RK_CHAN_1 on the GPU: LES_MODEL_TYPE must be 1'
        TS_FIRST = TIME_STEP
        RC = dh_create(PRM, CTX)
        IF (RC .NE. 0) STOP 'dh_create failed'
        RC = dh_set_grid(CTX, GY, GYF, DY, DYF)
        CALL DH_CHECK(CTX, RC, 'dh_set_grid')
        IF (LES) THEN
          RC = dh_set_les(CTX, 1)
          CALL DH_CHECK(CTX, RC, 'dh_set_les')
        END IF
        HOST_DIRTY = .TRUE.
      END IF

      IF (RK_STEP .EQ. 1) THEN
C TIME advances on the host (diablo.f:109); DELTA_T may have been changed
        RC = dh_set_clock(CTX, TIME, DELTA_T, TIME_STEP)
        CALL DH_CHECK(CTX, RC, 'dh_set_clock')
        IF (HOST_DIRTY) THEN
          RC = dh_upload_arrays(CTX, CU1, CU2, CU3, CP, CF1, CF2, CF3,
     &                          CTH, CFTH, TH)
          CALL DH_CHECK(CTX, RC, 'dh_upload')
          HOST_DIRTY = .FALSE.
        END IF
      END IF

      RC = dh_rk_substep(CTX, RK_STEP)
      CALL DH_CHECK(CTX, RC, 'dh_rk_substep')

      IF (RK_STEP .EQ. 3) THEN
C COURANT ran on the device: DELTA_T and H_BAR for diablo.f:109
        IF (VARIABLE_DT) THEN
          RC = dh_get_clock(CTX, TDEV, DELTA_T, H_BAR, TS)
          CALL DH_CHECK(CTX, RC, 'dh_get_clock')
        END IF
C Will the driver read or rewrite the state before the next substep?
        NEEDED = (MOD(TIME_STEP, ABS(SAVE_STATS_INT)) .EQ. 0)
     &      .OR. (MOD(TIME_STEP, ABS(SAVE_FLOW_INT)) .EQ. 0)
     &      .OR. (SAVE_STATS_INT .LT. 0) .OR. (SAVE_FLOW_INT .LT. 0)
     &      .OR. (TIME_STEP .GE. TS_FIRST - 1 + N_TIME_STEPS)
        INQUIRE(FILE='stop.now', EXIST=STOPFILE)
        CALL WALL_TIME(WT)
        NEEDED = NEEDED .OR. STOPFILE
     &      .OR. (WT-START_TIME .GT. TIME_LIMIT)
        IF (NEEDED) THEN
          RC = dh_download_arrays(CTX, CU1, CU2, CU3, CP, CF1, CF2, CF3,
     &                            CTH, CFTH, TH)
          CALL DH_CHECK(CTX, RC, 'dh_download')
          HOST_DIRTY = .TRUE.
        END IF
      END IF

      RETURN
      END

      SUBROUTINE DH_CHECK(CTX, RC, WHAT)
      USE diablo_hip
      TYPE(C_PTR) :: CTX
      INTEGER(C_INT) :: RC
      CHARACTER(*) :: WHAT
      IF (RC .NE. 0) THEN
        WRITE(*,*) 'libdiablo_hip: ', WHAT, ' failed: ',
     &             TRIM(dh_error_text(CTX))
        STOP
      END IF
      END
