C     Test driver of libsdegpu.so from Fortran (TEST INFRASTRUCTURE).
C
C     Calls GRDCF, GRDH2D, FFTR2D and GRDDEH2D as the reference's own
C     callers do (vis/visgrid.f:97-112,160-165, arr/arrfftr.f:139-141,
C     vis/visdegri.f:93,135-140), declares the reference's COMMON /SYSI/
C     (source/include/stdinc.h:64-67) and its own ERRREPOR, so the
C     library's sticky ERROR flag and error reports are the program's.
C     Reads its inputs from in.bin and writes its outputs to out.bin
C     (stream access, native byte order), in the working directory.
C
      PROGRAM DROPIN
      LOGICAL ERROR, SYSDEBUG, SYSMON, SYSINTRP, SYSMSG, SYSPAR
      INTEGER MSGWIDTH
      COMMON /SYSI/ ERROR, MSGWIDTH, SYSDEBUG, SYSMON,
     1   SYSINTRP, SYSMSG, SYSPAR
      CHARACTER*80 LASTMSG
      INTEGER NREP
      COMMON /TMSG/ NREP, LASTMSG
C
      INTEGER MAXVIS, MAXNU, MAXNV, MAXPIX
      PARAMETER (MAXVIS = 20000, MAXNU = 65, MAXNV = 128)
      PARAMETER (MAXPIX = 128)
      COMPLEX VIS(MAXVIS), GVIS(MAXNU, MAXNV), MGRID(MAXNU, MAXNV)
      COMPLEX MVIS(MAXVIS), WORK(1), WORK2(1), XGRID(MAXNU, MAXNV)
      REAL WT(MAXVIS), U(MAXVIS), V(MAXVIS), W(MAXVIS), MWT(MAXVIS)
      REAL BADWT(MAXVIS)
      REAL CF(1025), SUMWT, SUMWT1, USCALE, VSCALE, RWORK(1)
C     FFTR2D writes RARRAY(NX,NY) contiguously from the start of IMG
      REAL IMG(MAXPIX * MAXPIX)
      DOUBLE PRECISION SHIFT(3, 3)
      INTEGER NVIS, NU, NV, NX, NY, FNU, UORIG, VORIG, SUPP, OSAMP
      INTEGER I, J, KEPT, NOWGT
      LOGICAL PSF
C
      NREP = 0
      LASTMSG = ' '
      OPEN (10, FILE='in.bin', ACCESS='STREAM', FORM='UNFORMATTED',
     1   STATUS='OLD')
      READ (10) NVIS, NU, NV, NX, NY, UORIG, VORIG
      READ (10) USCALE, VSCALE
      READ (10) (VIS(I), I = 1, NVIS)
      READ (10) (WT(I), I = 1, NVIS)
      READ (10) (U(I), I = 1, NVIS)
      READ (10) (V(I), I = 1, NVIS)
      READ (10) (W(I), I = 1, NVIS)
      READ (10) ((SHIFT(I, J), I = 1, 3), J = 1, 3)
      READ (10) ((MGRID(I, J), I = 1, NU), J = 1, NV)
      CLOSE (10)
      FNU = 2 * (NU - 1)
C
C     The convolution function and the dirty grid (vis/visgrid.f)
C
      ERROR = .FALSE.
      PSF = .FALSE.
      CALL GRDCF ('SF', CF, SUPP, OSAMP)
      CALL GRDH2D (VIS, WT, U, V, W, NVIS, USCALE, 0.0, VSCALE, 0.0,
     1   UORIG, VORIG, GVIS, PSF, NU, NV, CF, SUPP, OSAMP, CF,
     2   SUPP, OSAMP, SHIFT, SUMWT)
      SUMWT1 = SUMWT
C
C     The image of that grid (arr/arrfftr.f: DIR < 0), on a copy: FFTR2D
C     overwrites its XARRAY with the column transform
C
      DO 8 J = 1, NV
         DO 7 I = 1, NU
            XGRID(I, J) = GVIS(I, J)
   7     CONTINUE
   8  CONTINUE
      CALL FFTR2D (IMG, XGRID, WORK, WORK2, RWORK, NX, NY, FNU, NV, -1)
C
C     Degridding of the given grid (vis/visdegri.f)
C
      DO 10 I = 1, NVIS
         MVIS(I) = (0.0, 0.0)
         MWT(I) = WT(I)
  10  CONTINUE
      CALL GRDDEH2D (MVIS, MWT, U, V, W, NVIS, USCALE, 0.0, VSCALE,
     1   0.0, UORIG, VORIG, MGRID, NU, NV, CF, SUPP, OSAMP, CF,
     2   SUPP, OSAMP, SHIFT)
C
C     With ERROR set on entry a routine returns at once, its outputs
C     untouched
C
      ERROR = .TRUE.
      DO 30 J = 1, NV
         DO 20 I = 1, NU
            MGRID(I, J) = (7.0, -7.0)
  20     CONTINUE
  30  CONTINUE
      SUMWT = -5.0
      CALL GRDH2D (VIS, WT, U, V, W, NVIS, USCALE, 0.0, VSCALE, 0.0,
     1   UORIG, VORIG, MGRID, PSF, NU, NV, CF, SUPP, OSAMP, CF,
     2   SUPP, OSAMP, SHIFT, SUMWT)
      KEPT = 1
      IF (SUMWT .NE. -5.0) KEPT = 0
      DO 50 J = 1, NV
         DO 40 I = 1, NU
            IF (MGRID(I, J) .NE. (7.0, -7.0)) KEPT = 0
  40     CONTINUE
  50  CONTINUE
      IF (.NOT. ERROR) KEPT = 0
C
C     No weight: the library reports through this program's ERRREPOR
C
      ERROR = .FALSE.
      DO 60 I = 1, NVIS
         BADWT(I) = -1.0
  60  CONTINUE
      CALL GRDH2D (VIS, BADWT, U, V, W, NVIS, USCALE, 0.0, VSCALE,
     1   0.0, UORIG, VORIG, MGRID, PSF, NU, NV, CF, SUPP, OSAMP, CF,
     2   SUPP, OSAMP, SHIFT, SUMWT)
      NOWGT = 0
      IF (ERROR) NOWGT = 1
      ERROR = .FALSE.
C
      OPEN (11, FILE='out.bin', ACCESS='STREAM', FORM='UNFORMATTED',
     1   STATUS='REPLACE')
      WRITE (11) SUPP, OSAMP, KEPT, NOWGT, NREP
      WRITE (11) SUMWT1
      WRITE (11) (CF(I), I = 1, 1025)
      WRITE (11) ((GVIS(I, J), I = 1, NU), J = 1, NV)
      WRITE (11) (IMG(I), I = 1, NX * NY)
      WRITE (11) (MVIS(I), I = 1, NVIS)
      WRITE (11) (MWT(I), I = 1, NVIS)
      WRITE (11) LASTMSG
      CLOSE (11)
      END
C
C     The host program's error report (the reference's ERRREPOR sets
C     ERROR and records the reason, main/err.f): keeps the last text.
C
      SUBROUTINE ERRREPOR (REASON, ROUTINE, TEXT)
      CHARACTER*(*) REASON, ROUTINE, TEXT
      LOGICAL ERROR, SYSDEBUG, SYSMON, SYSINTRP, SYSMSG, SYSPAR
      INTEGER MSGWIDTH
      COMMON /SYSI/ ERROR, MSGWIDTH, SYSDEBUG, SYSMON,
     1   SYSINTRP, SYSMSG, SYSPAR
      CHARACTER*80 LASTMSG
      INTEGER NREP
      COMMON /TMSG/ NREP, LASTMSG
      ERROR = .TRUE.
      NREP = NREP + 1
      LASTMSG = ROUTINE // ': ' // TEXT
      END
