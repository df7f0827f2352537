C     Legacy (Fortran 77 style) FFTW caller: the dfftw_* / dfftwf_* /
C     sfftw_* subroutines with every argument by reference and the plan in
C     an INTEGER*8.  Each case transforms a closed-form input whose DFT is
C     known exactly (complex tones, cosines), so a dimension-order or sign
C     mistake moves the peak.  Prints PASS or FAIL <case>.
      PROGRAM LEGACY
      IMPLICIT NONE
      INCLUDE 'fftw_wrapper.f'
      INTEGER N, NX, NY, HM
      PARAMETER (N=64, NX=8, NY=6, HM=3)
      INTEGER*8 PLAN
      INTEGER I, J, K, NFAIL, NN(1), HN(1), HIS(1), HOS(1)
      INTEGER ONE
      DOUBLE COMPLEX X(N), Y(N), A(NX,NY), B(NX,NY), EXPECT
      DOUBLE COMPLEX XM(N,HM), YM(N,HM)
      COMPLEX XS(N), YS(N)
      DOUBLE PRECISION XR(N), PI, ERR, TWOPI
      DOUBLE COMPLEX XC(N/2+1)
      NFAIL = 0
      ONE = 1
      PI = 4.0D0 * ATAN(1.0D0)
      TWOPI = 2.0D0 * PI
C     1-D double, forward: tone at f = 5 -> N at bin 5 (index 6)
      DO I = 1, N
         X(I) = EXP(DCMPLX(0.0D0, TWOPI * 5.0D0 * (I - 1) / N))
      END DO
      CALL DFFTW_PLAN_DFT_1D(PLAN, N, X, Y, FFTW_FORWARD,
     &                       FFTW_ESTIMATE)
      CALL DFFTW_EXECUTE_DFT(PLAN, X, Y)
      CALL DFFTW_DESTROY_PLAN(PLAN)
      ERR = 0.0D0
      DO K = 1, N
         EXPECT = (0.0D0, 0.0D0)
         IF (K .EQ. 6) EXPECT = DCMPLX(DBLE(N), 0.0D0)
         ERR = MAX(ERR, ABS(Y(K) - EXPECT))
      END DO
      IF (ERR .GT. 1.0D-10) THEN
         PRINT *, 'FAIL 1d-double', ERR
         NFAIL = NFAIL + 1
      END IF
C     2-D double, A(NX, NY) column-major: tone (fx, fy) = (3, 2)
      DO J = 1, NY
         DO I = 1, NX
            A(I, J) = EXP(DCMPLX(0.0D0, TWOPI * (3.0D0 * (I - 1) / NX
     &                + 2.0D0 * (J - 1) / NY)))
         END DO
      END DO
      CALL DFFTW_PLAN_DFT_2D(PLAN, NX, NY, A, B, FFTW_FORWARD,
     &                       FFTW_ESTIMATE)
      CALL DFFTW_EXECUTE(PLAN)
      CALL DFFTW_DESTROY_PLAN(PLAN)
      ERR = 0.0D0
      DO J = 1, NY
         DO I = 1, NX
            EXPECT = (0.0D0, 0.0D0)
            IF (I .EQ. 4 .AND. J .EQ. 3)
     &         EXPECT = DCMPLX(DBLE(NX * NY), 0.0D0)
            ERR = MAX(ERR, ABS(B(I, J) - EXPECT))
         END DO
      END DO
      IF (ERR .GT. 1.0D-10) THEN
         PRINT *, 'FAIL 2d-double', ERR
         NFAIL = NFAIL + 1
      END IF
C     1-D single (dfftwf_ and sfftw_ names), backward: tone at f = 7
C     -> N at bin N - 7 (index N - 6)
      DO I = 1, N
         XS(I) = CMPLX(EXP(DCMPLX(0.0D0, TWOPI * 7.0D0 * (I - 1) / N)))
      END DO
      CALL DFFTWF_PLAN_DFT_1D(PLAN, N, XS, YS, FFTW_BACKWARD,
     &                        FFTW_ESTIMATE)
      CALL DFFTWF_EXECUTE_DFT(PLAN, XS, YS)
      CALL DFFTWF_DESTROY_PLAN(PLAN)
      ERR = 0.0D0
      DO K = 1, N
         EXPECT = (0.0D0, 0.0D0)
         IF (K .EQ. N - 6) EXPECT = DCMPLX(DBLE(N), 0.0D0)
         ERR = MAX(ERR, ABS(DCMPLX(YS(K)) - EXPECT))
      END DO
      CALL SFFTW_PLAN_DFT_1D(PLAN, N, XS, YS, FFTW_BACKWARD,
     &                       FFTW_ESTIMATE)
      CALL SFFTW_EXECUTE(PLAN)
      CALL SFFTW_DESTROY_PLAN(PLAN)
      DO K = 1, N
         EXPECT = (0.0D0, 0.0D0)
         IF (K .EQ. N - 6) EXPECT = DCMPLX(DBLE(N), 0.0D0)
         ERR = MAX(ERR, ABS(DCMPLX(YS(K)) - EXPECT))
      END DO
      IF (ERR .GT. 1.0D-3) THEN
         PRINT *, 'FAIL 1d-single', ERR
         NFAIL = NFAIL + 1
      END IF
C     R2C double: cos(2 pi 4 j / N) -> N/2 at bin 4 (index 5)
      DO I = 1, N
         XR(I) = COS(TWOPI * 4.0D0 * (I - 1) / N)
      END DO
      CALL DFFTW_PLAN_DFT_R2C_1D(PLAN, N, XR, XC, FFTW_ESTIMATE)
      CALL DFFTW_EXECUTE_DFT_R2C(PLAN, XR, XC)
      CALL DFFTW_DESTROY_PLAN(PLAN)
      ERR = 0.0D0
      DO K = 1, N / 2 + 1
         EXPECT = (0.0D0, 0.0D0)
         IF (K .EQ. 5) EXPECT = DCMPLX(DBLE(N) / 2.0D0, 0.0D0)
         ERR = MAX(ERR, ABS(XC(K) - EXPECT))
      END DO
      IF (ERR .GT. 1.0D-10) THEN
         PRINT *, 'FAIL r2c-double', ERR
         NFAIL = NFAIL + 1
      END IF
C     plan_many: HM transforms of length N, column m holds tone f = m
      DO J = 1, HM
         DO I = 1, N
            XM(I, J) = EXP(DCMPLX(0.0D0, TWOPI * J * (I - 1) / N))
         END DO
      END DO
      NN(1) = N
      CALL DFFTW_PLAN_MANY_DFT(PLAN, ONE, NN, HM, XM, NN, ONE, N,
     &     YM, NN, ONE, N, FFTW_FORWARD, FFTW_ESTIMATE)
      CALL DFFTW_EXECUTE_DFT(PLAN, XM, YM)
      CALL DFFTW_DESTROY_PLAN(PLAN)
      ERR = 0.0D0
      DO J = 1, HM
         DO K = 1, N
            EXPECT = (0.0D0, 0.0D0)
            IF (K .EQ. J + 1) EXPECT = DCMPLX(DBLE(N), 0.0D0)
            ERR = MAX(ERR, ABS(YM(K, J) - EXPECT))
         END DO
      END DO
      IF (ERR .GT. 1.0D-10) THEN
         PRINT *, 'FAIL many-double', ERR
         NFAIL = NFAIL + 1
      END IF
C     guru: the same HM transforms through (n, is, os) and one howmany axis
      HN(1) = HM
      HIS(1) = N
      HOS(1) = N
      CALL DFFTW_PLAN_GURU_DFT(PLAN, ONE, NN, ONE, ONE, ONE, HN, HIS,
     &     HOS, XM, YM, FFTW_FORWARD, FFTW_ESTIMATE)
      CALL DFFTW_EXECUTE_DFT(PLAN, XM, YM)
      CALL DFFTW_DESTROY_PLAN(PLAN)
      ERR = 0.0D0
      DO J = 1, HM
         DO K = 1, N
            EXPECT = (0.0D0, 0.0D0)
            IF (K .EQ. J + 1) EXPECT = DCMPLX(DBLE(N), 0.0D0)
            ERR = MAX(ERR, ABS(YM(K, J) - EXPECT))
         END DO
      END DO
      IF (ERR .GT. 1.0D-10) THEN
         PRINT *, 'FAIL guru-double', ERR
         NFAIL = NFAIL + 1
      END IF
      IF (NFAIL .EQ. 0) PRINT *, 'PASS legacy'
      IF (NFAIL .NE. 0) STOP 1
      END
