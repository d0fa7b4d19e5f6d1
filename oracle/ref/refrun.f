      PROGRAM REFRUN
C     -----------------------------------------------------------------
C     ORACLE HARNESS (TEST INFRASTRUCTURE ONLY)
C     Drives the reference SENGA2 routines on one rank from a restart
C     dump (input/cont.dat, cold start switch 1) and writes raw FP64
C     arrays (stream access, Fortran order) for the golden fixtures.
C     The time loop follows src/senga2.f:91-190 with the turbulent
C     inflow generator, output and termination calls left out.
C
C     refrun.ctl line 1: MODE
C       1 = one right-hand-side evaluation (BOUNDT,PARFER,RHSCAL,
C           RHSVEL,BOUNDS at RK substage 1), no register update
C       2 = NTIME full RK steps, final solution register
C       3 = the nine derivative operators applied to opin.bin
C       4 = one substage (mode 1 plus LINCOM at substage 1)
C       5 = mode 2 with OUTPUT called after every step as
C           src/senga2.f:186 does (report, statistics, out*.res
C           post-processing files and dumps in output/)
C     refrun.ctl line 2: output file name
C     -----------------------------------------------------------------
      INCLUDE 'com_senga2.h'

      INTEGER MODE,JTIME,JRKSTP,IC,JC,KC,ISPEC,NX,NY,NZ
      CHARACTER*256 FNOUT

      NX = NXSIZE
      NY = NYSIZE
      NZ = NZSIZE

      OPEN(UNIT=91,FILE='refrun.ctl',STATUS='OLD',FORM='FORMATTED')
      READ(91,*)MODE
      READ(91,'(A)')FNOUT
      CLOSE(91)

      CALL PARDOM
      CALL INDATA

      OPEN(UNIT=92,FILE=TRIM(FNOUT),ACCESS='STREAM',
     +     FORM='UNFORMATTED',STATUS='REPLACE')
      WRITE(92)NX,NY,NZ,NSPEC,MODE

C     TEMPIN RESULTS (INITIAL TEMPERATURE AND INTERVAL INDEX)
      WRITE(92)TRUN(1:NX,1:NY,1:NZ)
      WRITE(92)ITNDEX(1:NX,1:NY,1:NZ,1)

      IF(MODE.EQ.1.OR.MODE.EQ.4)THEN

        ITIME = NTIME1
        IRKSTP = 1
        CALL BOUNDT
        CALL PARFER
        CALL RHSCAL
        CALL RHSVEL
        CALL BOUNDS
        WRITE(92)DRHS(1:NX,1:NY,1:NZ),URHS(1:NX,1:NY,1:NZ),
     +           VRHS(1:NX,1:NY,1:NZ),WRHS(1:NX,1:NY,1:NZ),
     +           ERHS(1:NX,1:NY,1:NZ),
     +           YRHS(1:NX,1:NY,1:NZ,1:NSPEC)
        WRITE(92)TRUN(1:NX,1:NY,1:NZ),PRUN(1:NX,1:NY,1:NZ),
     +           RRTE(1:NX,1:NY,1:NZ,1:NSPEC)
        IF(MODE.EQ.4)THEN
          CALL LINCOM
          WRITE(92)DRUN(1:NX,1:NY,1:NZ),URUN(1:NX,1:NY,1:NZ),
     +             VRUN(1:NX,1:NY,1:NZ),WRUN(1:NX,1:NY,1:NZ),
     +             ERUN(1:NX,1:NY,1:NZ),
     +             YRUN(1:NX,1:NY,1:NZ,1:NSPEC)
          WRITE(92)DRHS(1:NX,1:NY,1:NZ),URHS(1:NX,1:NY,1:NZ),
     +             VRHS(1:NX,1:NY,1:NZ),WRHS(1:NX,1:NY,1:NZ),
     +             ERHS(1:NX,1:NY,1:NZ),
     +             YRHS(1:NX,1:NY,1:NZ,1:NSPEC)
        ENDIF

      ELSEIF(MODE.EQ.2.OR.MODE.EQ.5)THEN

        DO JTIME = NTIME1,NTIME2
          ITIME = JTIME
          DO JRKSTP = 1,NRKSM1
            IRKSTP = JRKSTP
            CALL BOUNDT
            CALL PARFER
            CALL RHSCAL
            CALL RHSVEL
            CALL BOUNDS
            CALL LINCOM
          ENDDO
          IRKSTP = NRKSTP
          CALL BOUNDT
          CALL PARFER
          CALL RHSCAL
          CALL RHSVEL
          CALL BOUNDS
          CALL FINCOM
          ETIME = ETIME + TSTEP
          CALL BOUNTT
          CALL ADAPTT
          IF(MODE.EQ.5)CALL OUTPUT
        ENDDO
        WRITE(92)DRUN(1:NX,1:NY,1:NZ),URUN(1:NX,1:NY,1:NZ),
     +           VRUN(1:NX,1:NY,1:NZ),WRUN(1:NX,1:NY,1:NZ),
     +           ERUN(1:NX,1:NY,1:NZ),
     +           YRUN(1:NX,1:NY,1:NZ,1:NSPEC)
        WRITE(92)TRUN(1:NX,1:NY,1:NZ),PRUN(1:NX,1:NY,1:NZ),
     +           RRTE(1:NX,1:NY,1:NZ,1:NSPEC)
        WRITE(92)ETIME,TSTEP

      ELSEIF(MODE.EQ.3)THEN

        CALL OPTEST(NX,NY,NZ)

      ENDIF

      CLOSE(92)
      STOP
      END


      SUBROUTINE OPTEST(NX,NY,NZ)
C     APPLY THE NINE DERIVATIVE OPERATORS TO THE FIELD IN opin.bin.
C     HALOS IN PERIODIC DIRECTIONS ARE FILLED BY WRAP-AROUND COPIES IN
C     THE ORDER X, Y (X-EXTENDED), Z (X,Y-EXTENDED) AS src/parfer.f
C     DOES ON A SINGLE RANK.
      INCLUDE 'com_senga2.h'
      INTEGER NX,NY,NZ,IC,JC,KC,IH
      DOUBLE PRECISION FIELD(NXSIZE,NYSIZE,NZSIZE)

      OPEN(UNIT=93,FILE='opin.bin',ACCESS='STREAM',
     +     FORM='UNFORMATTED',STATUS='OLD')
      READ(93)FIELD
      CLOSE(93)

      DO KC = NZBIGL,NZBIGR
        DO JC = NYBIGL,NYBIGR
          DO IC = NXBIGL,NXBIGR
            STORE7(IC,JC,KC) = 0.0D0
          ENDDO
        ENDDO
      ENDDO
      DO KC = 1,NZ
        DO JC = 1,NY
          DO IC = 1,NX
            STORE7(IC,JC,KC) = FIELD(IC,JC,KC)
          ENDDO
        ENDDO
      ENDDO
      IF(NENDXL.EQ.NPERI)THEN
        DO KC = 1,NZ
          DO JC = 1,NY
            DO IH = 1,NHALOX
              STORE7(1-IH,JC,KC) = STORE7(NX+1-IH,JC,KC)
              STORE7(NX+IH,JC,KC) = STORE7(IH,JC,KC)
            ENDDO
          ENDDO
        ENDDO
      ENDIF
      IF(NENDYL.EQ.NPERI)THEN
        DO KC = 1,NZ
          DO IH = 1,NHALOY
            DO IC = NXBIGL,NXBIGR
              STORE7(IC,1-IH,KC) = STORE7(IC,NY+1-IH,KC)
              STORE7(IC,NY+IH,KC) = STORE7(IC,IH,KC)
            ENDDO
          ENDDO
        ENDDO
      ENDIF
      IF(NENDZL.EQ.NPERI)THEN
        DO IH = 1,NHALOZ
          DO JC = NYBIGL,NYBIGR
            DO IC = NXBIGL,NXBIGR
              STORE7(IC,JC,1-IH) = STORE7(IC,JC,NZ+1-IH)
              STORE7(IC,JC,NZ+IH) = STORE7(IC,JC,IH)
            ENDDO
          ENDDO
        ENDDO
      ENDIF

      CALL DFBYDX(STORE7,STORE1)
      WRITE(92)STORE1
      CALL DFBYDY(STORE7,STORE1)
      WRITE(92)STORE1
      CALL DFBYDZ(STORE7,STORE1)
      WRITE(92)STORE1
      CALL D2FDX2(STORE7,STORE1)
      WRITE(92)STORE1
      CALL D2FDY2(STORE7,STORE1)
      WRITE(92)STORE1
      CALL D2FDZ2(STORE7,STORE1)
      WRITE(92)STORE1
      CALL D2FDXY(STORE7,STORE1)
      WRITE(92)STORE1
      CALL D2FDXZ(STORE7,STORE1)
      WRITE(92)STORE1
      CALL D2FDYZ(STORE7,STORE1)
      WRITE(92)STORE1

      RETURN
      END
