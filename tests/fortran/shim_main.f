! Test driver that only links the drop-in: one call of RK_CHAN_1 would need
! a GPU, so the program stops before it unless asked to run.
      PROGRAM SHIM_MAIN
      INCLUDE 'header'
      CHARACTER(8) ARG
      CALL GET_COMMAND_ARGUMENT(1, ARG)
      IF (ARG .NE. 'run') STOP
      RK_STEP = 1
      CALL RK_CHAN_1
      END

! the driver's wall clock (diablo_io.f), which the drop-in reads
      SUBROUTINE WALL_TIME(WT)
      REAL*8 WT
      WT = 0.D0
      END
