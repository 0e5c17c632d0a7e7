C  ref_bilan.F -- TEST INFRASTRUCTURE ONLY (oracle harness, compiled by oracle/build_ref.py).
C
C  Drives the reference's own energy bookkeeping routines on one element group of a state,
C  as SFORC3 and CFORC3 call them on output cycles (sforc3.F:1135-1144, cforc3 -> CBILAN):
C  SBILAN and CBILAN add the group's internal energy, EI = FILL*EINT*VOL for solids
C  (sbilan.F:140) and EINT(I,1) + EINT(I,2) for shells (cbilan.F:168), into PARTSAV(1,part),
C  element after element.  One part, no additional outputs (IPARG(80) = 0), no group
C  outputs (IGRE = 0), no thermal expansion (IEXPAN = 0), no XFEM, Parith/OFF sensor path
C  with no energy sensor.
      SUBROUTINE REF_SBILAN(NEL, VEL, EINT, RHO, VOL, VNEW, OFF, PSAV)
     .   BIND(C, NAME='ref_sbilan')
      USE ISO_C_BINDING
      USE SENSOR_MOD
#include "implicit_f.inc"
#include "mvsiz_p.inc"
#include "param_c.inc"
#include "vect01_c.inc"
#include "inter22.inc"
      INTEGER(C_INT), VALUE :: NEL
      REAL(C_DOUBLE) :: VEL(3,8,NEL), EINT(NEL), RHO(NEL), VOL(NEL), VNEW(NEL), OFF(NEL)
      REAL(C_DOUBLE) :: PSAV(NPSAV)
      my_real VX(MVSIZ,8), VY(MVSIZ,8), VZ(MVSIZ,8), XD(MVSIZ,8), RK(MVSIZ), FILL(MVSIZ),
     .   EINTTH(MVSIZ), MOM(MVSIZ,3), GRESAV(1), WPLA(1)
      INTEGER IPARTS(MVSIZ), GRTH(1), IGRTH(1), IPARG(100), I, K
      TYPE(SENSORS_) :: SENS
      INT22 = 0
      IPARTSPH = 0
      IGRE = 0
      IPARG = 0
      DO I = 1, NEL
        DO K = 1, 8
          VX(I,K) = VEL(1,K,I)
          VY(I,K) = VEL(2,K,I)
          VZ(I,K) = VEL(3,K,I)
          XD(I,K) = ZERO
        ENDDO
        IPARTS(I) = 1
        FILL(I) = ONE
        RK(I) = ZERO
        EINTTH(I) = ZERO
        MOM(I,1:3) = ZERO
      ENDDO
      CALL SBILAN(PSAV, EINT, RHO, RK, VOL,
     .   VX(1,1), VX(1,2), VX(1,3), VX(1,4), VX(1,5), VX(1,6), VX(1,7), VX(1,8),
     .   VY(1,1), VY(1,2), VY(1,3), VY(1,4), VY(1,5), VY(1,6), VY(1,7), VY(1,8),
     .   VZ(1,1), VZ(1,2), VZ(1,3), VZ(1,4), VZ(1,5), VZ(1,6), VZ(1,7), VZ(1,8),
     .   VNEW, IPARTS, GRESAV, GRTH, IGRTH, OFF,
     .   0, EINTTH, FILL, MOM, NEL, 0,
     .   XD(1,1), XD(1,2), XD(1,3), XD(1,4), XD(1,5), XD(1,6), XD(1,7), XD(1,8),
     .   XD(1,1), XD(1,2), XD(1,3), XD(1,4), XD(1,5), XD(1,6), XD(1,7), XD(1,8),
     .   XD(1,1), XD(1,2), XD(1,3), XD(1,4), XD(1,5), XD(1,6), XD(1,7), XD(1,8),
     .   0, IPARG, SENS, 0, WPLA)
      END

      SUBROUTINE REF_CBILAN(NEL, NNOD, X, V, VR, IXC4, EINT, RHO, GVOL,
     .   OFF, PSAV) BIND(C, NAME='ref_cbilan')
      USE ISO_C_BINDING
      USE SENSOR_MOD
      use element_mod , only : nixc
#include "implicit_f.inc"
#include "mvsiz_p.inc"
#include "param_c.inc"
#include "com_xfem1.inc"
#include "parit_c.inc"
      INTEGER(C_INT), VALUE :: NEL, NNOD
      REAL(C_DOUBLE) :: X(3,NNOD), V(3,NNOD), VR(3,NNOD), EINT(NEL,2), GVOL(NEL), OFF(NEL)
      REAL(C_DOUBLE), VALUE :: RHO
      INTEGER(C_INT) :: IXC4(4,NEL)
      REAL(C_DOUBLE) :: PSAV(NPSAV)
      INTEGER IXC(NIXC,MVSIZ), MAT(MVSIZ), IPARTC(MVSIZ), GRTH(1), IGRTH(1), I
      my_real PM(500,1), THK(MVSIZ), AREA(MVSIZ), VOL0(MVSIZ), VOL00(MVSIZ), THK0(MVSIZ),
     .   THK02(MVSIZ), GRESAV(1), EINTTH(MVSIZ), WPLA(1), GV(MVSIZ),
     .   X1G(MVSIZ), X2G(MVSIZ), X3G(MVSIZ), X4G(MVSIZ),
     .   Y1G(MVSIZ), Y2G(MVSIZ), Y3G(MVSIZ), Y4G(MVSIZ),
     .   Z1G(MVSIZ), Z2G(MVSIZ), Z3G(MVSIZ), Z4G(MVSIZ),
     .   VL1(MVSIZ,3), VL2(MVSIZ,3), VL3(MVSIZ,3), VL4(MVSIZ,3),
     .   VRL1(MVSIZ,3), VRL2(MVSIZ,3), VRL3(MVSIZ,3), VRL4(MVSIZ,3)
      TYPE(SENSORS_) :: SENS
      ICRACK3D = 0
      IPARIT = 0
      ALLOCATE(SENS%NGR_SENSOR(1))
      SENS%NGR_SENSOR = 0
      PM = ZERO
      PM(1,1) = RHO
      DO I = 1, NEL
        IXC(1,I) = 1
        IXC(2:5,I) = IXC4(1:4,I)
        IXC(6,I) = 1
        IXC(7,I) = I
        MAT(I) = 1
        IPARTC(I) = 1
        GV(I) = GVOL(I)
        THK(I) = ZERO
        AREA(I) = ZERO
        VOL0(I) = ZERO
        VOL00(I) = ZERO
        THK0(I) = ZERO
        THK02(I) = ZERO
        EINTTH(I) = ZERO
      ENDDO
C     IFLA = 2: CBILAN gathers X, V, VR of the element nodes itself (cbilan.F:116-160)
      CALL CBILAN(1, NEL, PM, V,
     .   IXC, THK, EINT, PSAV,
     .   AREA, MAT, IPARTC, X,
     .   VR, VOL0, VOL00, THK0,
     .   THK02, 2, OFF, 0,
     .   GRESAV, GRTH, IGRTH, VL1,
     .   VL2, VL3, VL4, VRL1,
     .   VRL2, VRL3, VRL4, X1G,
     .   X2G, X3G, X4G, Y1G,
     .   Y2G, Y3G, Y4G, Z1G,
     .   Z2G, Z3G, Z4G, 0,
     .   0, EINTTH, 0, GV,
     .   0, 0, SENS, NEL,
     .   0, WPLA)
      DEALLOCATE(SENS%NGR_SENSOR)
      END
