C  ref_shell.F -- harness that drives the reference's own compiled Q4 routines.
C
C  TEST INFRASTRUCTURE ONLY (built by oracle/build_ref.py into oracle/_ref).
C  For one group of NEL <= 128 Q4 shells it sets the common-block switches of
C  the restated branch (Ishell=1, ISMSTR=2, ISHFRAM=0, IGTYP=1, law 2, Parith/ON,
C  explicit, element time step) and calls the reference routines in the order
C  CFORC3 calls them (cforc3.F:403-751): CCOOR3, CNVEC3, CDERI3, CCOEF3, CDLEN3,
C  CDEFO3, CCURV3, CSTRA3, SIGEPS02C (with M2CPLR inside) per integration point,
C  CHVIS3, CDT3, CFINT3, CUPDT3P.  The pieces it does itself are the element
C  strain rate of CFORC3 (cforc3.F:537-546) and the law-2 glue of MULAWC around
C  SIGEPS02C (mulawc.F90:542-579, 745-760, 845-852, 1102-1151, 2628-2660,
C  2898-2912, 3013-3062): layer positions/weights from COQINI_WM and the Z0/WF
C  tables, resultant sums, rupture test, thickness clamp and energy update.
      SUBROUTINE REF_Q4_FORCES(NEL, NNOD, X, V, VR, CONN, OFFG, THK,
     .   THKE, FOR, MOM, GSTR, HOUR, EINT, EPSDG, SIGL, PLAL, EPSDL,
     .   NSKY, FSKY, IADC, PRM, IPRM, DTIN1, DT2T, EHOU_OUT)
     .   BIND(C, NAME='ref_q4_forces')
      USE ISO_C_BINDING
      USE MATPARAM_DEF_MOD
      USE ELBUFDEF_MOD
#include "implicit_f.inc"
#include "mvsiz_p.inc"
#include "param_c.inc"
#include "com01_c.inc"
#include "com04_c.inc"
#include "com06_c.inc"
#include "com08_c.inc"
#include "impl1_c.inc"
#include "parit_c.inc"
#include "remesh_c.inc"
#include "scr02_c.inc"
#include "scr06_c.inc"
#include "scr14_c.inc"
#include "scr17_c.inc"
#include "scr18_c.inc"
#include "sms_c.inc"
      INTEGER(C_INT), VALUE :: NEL, NNOD, NSKY
      REAL(C_DOUBLE) :: X(3,NNOD), V(3,NNOD), VR(3,NNOD)
      INTEGER(C_INT) :: CONN(4,NEL), IADC(4,NEL), IPRM(8)
      REAL(C_DOUBLE) :: OFFG(NEL), THK(NEL), THKE(NEL), FOR(NEL,5),
     .   MOM(NEL,3), GSTR(NEL,8), HOUR(NEL,5), EINT(NEL,2), EPSDG(NEL),
     .   SIGL(NEL,5,*), PLAL(NEL,*), EPSDL(NEL,*), FSKY(8,NSKY),
     .   PRM(24), DT2T, EHOU_OUT
      REAL(C_DOUBLE), VALUE :: DTIN1
C
      TYPE(MATPARAM_STRUCT_) :: MATP
      TYPE(ELBUF_STRUCT_) :: ELB
      INTEGER I, IP, NPT, IPLA, ITHK, ISMS, IHB, NPTTOT
      INTEGER IXC(7,MVSIZ), MAT(MVSIZ), PID(MVSIZ), NGL(MVSIZ),
     .   IPARTC(MVSIZ), INDXOF(MVSIZ), NDT(MVSIZ),
     .   IOFF_DUCT(MVSIZ), NELTST, ITYPTST, IOFC
      DOUBLE PRECISION SMSTR(6*MVSIZ)
      my_real
     .   PM_STACK(20,1),
     .   OFF(MVSIZ), SIGY(MVSIZ), DT1C(MVSIZ), THK0(MVSIZ),
     .   VL1(MVSIZ,3), VL2(MVSIZ,3), VL3(MVSIZ,3), VL4(MVSIZ,3),
     .   VRL1(MVSIZ,3), VRL2(MVSIZ,3), VRL3(MVSIZ,3), VRL4(MVSIZ,3),
     .   X1G(MVSIZ), X2G(MVSIZ), X3G(MVSIZ), X4G(MVSIZ),
     .   Y1G(MVSIZ), Y2G(MVSIZ), Y3G(MVSIZ), Y4G(MVSIZ),
     .   Z1G(MVSIZ), Z2G(MVSIZ), Z3G(MVSIZ), Z4G(MVSIZ),
     .   E1X(MVSIZ), E1Y(MVSIZ), E1Z(MVSIZ), E2X(MVSIZ), E2Y(MVSIZ),
     .   E2Z(MVSIZ), E3X(MVSIZ), E3Y(MVSIZ), E3Z(MVSIZ),
     .   DIRA(1), DIRB(1), STI(MVSIZ), STIR(MVSIZ), AREA(MVSIZ),
     .   PX1(MVSIZ), PX2(MVSIZ), PY1(MVSIZ), PY2(MVSIZ),
     .   X2(MVSIZ), X3(MVSIZ), X4(MVSIZ), Y2(MVSIZ), Y3(MVSIZ),
     .   Y4(MVSIZ), Z2(MVSIZ), VHX(MVSIZ), VHY(MVSIZ), A_I(MVSIZ),
     .   UX1(MVSIZ), UX2(MVSIZ), UX3(MVSIZ), UX4(MVSIZ),
     .   UY1(MVSIZ), UY2(MVSIZ), UY3(MVSIZ), UY4(MVSIZ)
      my_real
     .   SHF(MVSIZ), THK02(MVSIZ), NU(MVSIZ), G(MVSIZ), YM(MVSIZ),
     .   A11(MVSIZ), A12(MVSIZ), SSP(MVSIZ), RHO(MVSIZ), H1(MVSIZ),
     .   H2(MVSIZ), H3(MVSIZ), VOL0(MVSIZ), VOL00(MVSIZ), ALPE(MVSIZ),
     .   GS(MVSIZ), SRH1(MVSIZ), SRH2(MVSIZ), SRH3(MVSIZ), A11R(MVSIZ),
     .   ALDT(MVSIZ), EXX(MVSIZ), EYY(MVSIZ), EXY(MVSIZ), EXZ(MVSIZ),
     .   EYZ(MVSIZ), KXX(MVSIZ), KYY(MVSIZ), KXY(MVSIZ),
     .   VX1(MVSIZ), VX2(MVSIZ), VX3(MVSIZ), VX4(MVSIZ),
     .   VY1(MVSIZ), VY2(MVSIZ), VY3(MVSIZ), VY4(MVSIZ),
     .   VZ1(MVSIZ), VZ2(MVSIZ), VZ3(MVSIZ), VZ4(MVSIZ),
     .   RX1(MVSIZ), RX2(MVSIZ), RX3(MVSIZ), RX4(MVSIZ),
     .   RY1(MVSIZ), RY2(MVSIZ), RY3(MVSIZ), RY4(MVSIZ),
     .   WXY(MVSIZ), TANI(6,MVSIZ), GSTRW(MVSIZ), F_DEF(MVSIZ,8)
      my_real
     .   EPSD_PG(MVSIZ), DEGMB(MVSIZ), DEGFX(MVSIZ), THKN(MVSIZ),
     .   OFF_OLD(MVSIZ), EPCHK(MVSIZ), ETSE(MVSIZ), G_IMP(MVSIZ),
     .   DPLA(MVSIZ), TSTAR(MVSIZ), HARDM(MVSIZ), SIGKSI(MVSIZ,5),
     .   TEMPEL(MVSIZ), FHEAT(MVSIZ), VOLIP(MVSIZ), THKLYL(MVSIZ),
     .   LOFF(MVSIZ), DPLANL(MVSIZ), SIGB(MVSIZ,3),
     .   SO1(MVSIZ), SO2(MVSIZ), SO3(MVSIZ), SO4(MVSIZ), SO5(MVSIZ),
     .   SN1(MVSIZ), SN2(MVSIZ), SN3(MVSIZ), SN4(MVSIZ), SN5(MVSIZ),
     .   DEXX(MVSIZ), DEYY(MVSIZ), DEXY(MVSIZ), DEYZ(MVSIZ),
     .   DEZX(MVSIZ), EPX(MVSIZ), EPY(MVSIZ), EPXY(MVSIZ),
     .   PLA1(MVSIZ), EPS1(MVSIZ), VISCMX(MVSIZ), WM(11,11)
      my_real
     .   H11(MVSIZ), H12(MVSIZ), H13(MVSIZ), H21(MVSIZ), H22(MVSIZ),
     .   H23(MVSIZ), H31(MVSIZ), H32(MVSIZ), H33(MVSIZ),
     .   B11(MVSIZ), B12(MVSIZ), B13(MVSIZ), B14(MVSIZ),
     .   B21(MVSIZ), B22(MVSIZ), B23(MVSIZ), B24(MVSIZ),
     .   F11(MVSIZ), F12(MVSIZ), F13(MVSIZ), F14(MVSIZ),
     .   F21(MVSIZ), F22(MVSIZ), F23(MVSIZ), F24(MVSIZ),
     .   F31(MVSIZ), F32(MVSIZ), F33(MVSIZ), F34(MVSIZ),
     .   M11(MVSIZ), M12(MVSIZ), M13(MVSIZ), M14(MVSIZ),
     .   M21(MVSIZ), M22(MVSIZ), M23(MVSIZ), M24(MVSIZ),
     .   M31(MVSIZ), M32(MVSIZ), M33(MVSIZ), M34(MVSIZ),
     .   EANI(MVSIZ), DTCSUB(MVSIZ), DT2C(MVSIZ),
     .   MSC(MVSIZ), DMELC(MVSIZ), DTEL(MVSIZ), FZERO(3,4,1),
     .   THEM(MVSIZ,4), FTHESKY(1), CONDNSKY(1), CONDE(MVSIZ),
     .   Z0T(11,11), WFT(11,11)
      my_real DTINV, THKLY, WMC, ZT, VOL2, EPS_K2, EPS_M2, ASRT
      my_real, ALLOCATABLE :: PM(:,:), GEO(:,:), PARTSAV(:,:)
      INTEGER, ALLOCATABLE :: IPM(:,:), IGEO(:,:)
C
C-----array sizes of the restart (rdcomm.F:508-526, 762-764): any value above the
C     largest index the called routines use
      NPROPM = 500
      NPROPG = 1000
      NPROPMI = 400
      NPROPGI = 1000
      NPSAV = 29
      ALLOCATE(PM(NPROPM,1), GEO(NPROPG,1), PARTSAV(NPSAV,1))
      ALLOCATE(IPM(NPROPMI,1), IGEO(NPROPGI,1))
C-----switches of the restated branch (common blocks)
      DT1 = DTIN1
      TT = ONE
      IMPL_S = 0
      IKT = 0
      ISHFRAM = 0
      IEPSDOT = 0
      NADMESH = 0
      HELAS = HALF
      HVISC = HALF
      HVLIN = ZERO
      NODADT = 0
      IDT1SH = 0
      IDTMINS = 0
      IDTMIN(1:102) = 0
      IDTMIN(3) = 2
      DTFAC1(3) = PRM(21)
      DTMIN1(3) = PRM(22)
      ISIGI = 0
      IRUN = 1
      NCYCLE = 1
      IVECTOR = 0
      LSKY = NSKY
      INVSTR = 100
      NUMELS = 0
      NUMNOD = NNOD
      IRODDL = 1
      IDEL7NOK = 0
      NPT = IPRM(5)
      IPLA = IPRM(6)
      ITHK = IPRM(7)
      ISMS = 2
      IHB = 1
      NPTTOT = NPT
C-----material (hm_read_mat02_jc.F90 storage) and property
      PM = ZERO
      GEO = ZERO
      IGEO = 0
      PM_STACK = ZERO
      PM(1,1) = PRM(1)
      PM(20,1) = PRM(2)
      PM(21,1) = PRM(3)
      PM(22,1) = PRM(4)
      PM(24,1) = PRM(5)
      PM(25,1) = PRM(6)
      PM(27,1) = PRM(7)
      GEO(13,1) = PRM(17)
      GEO(14,1) = PRM(18)
      GEO(15,1) = PRM(19)
      GEO(18,1) = SQRT(GEO(13,1))
      GEO(19,1) = SQRT(GEO(14,1))
      GEO(20,1) = SQRT(GEO(15,1))
      GEO(37,1) = IPRM(8)
      GEO(38,1) = PRM(20)
      IGEO(11,1) = 1
      IGEO(98,1) = 0
      MATP%YOUNG = PRM(2)
      MATP%NU = PRM(3)
      MATP%SHEAR = PRM(4)
      ALLOCATE(MATP%UPARAM(12), MATP%IPARAM(4))
      MATP%UPARAM = ZERO
      MATP%UPARAM(1) = PRM(8)
      MATP%UPARAM(2) = PRM(9)
      MATP%UPARAM(3) = PRM(10)
      MATP%UPARAM(4) = PRM(11)
      MATP%UPARAM(5) = PRM(12)
      MATP%UPARAM(6) = PRM(13)
      MATP%UPARAM(7) = PRM(14)
      MATP%UPARAM(8) = ZERO
      MATP%UPARAM(10) = PRM(16)
      MATP%IPARAM(1) = IPRM(4)
      MATP%IPARAM(2) = IPRM(1)
      MATP%IPARAM(3) = IPRM(2)
      MATP%IPARAM(4) = IPRM(3)
      MATP%THERM%TREF = 300.0D0
      MATP%THERM%TMELT = 1.0D20
      MATP%THERM%RHOCP = ZERO
      MATP%THERM%EFRAC = ZERO
      ELB%IDRAPE = 0
      DO I = 1, NEL
        IXC(1,I) = 1
        IXC(2,I) = CONN(1,I)
        IXC(3,I) = CONN(2,I)
        IXC(4,I) = CONN(3,I)
        IXC(5,I) = CONN(4,I)
        IXC(6,I) = 1
        IXC(7,I) = I
        IPARTC(I) = 1
      ENDDO
      PARTSAV = ZERO
      EANI = ZERO
C-----CFORC3 sequence
      CALL CCOOR3(1, NEL, X, IXC, GEO, OFFG, OFF, SIGY, PID, V, VR,
     .   VL1, VL2, VL3, VL4, VRL1, VRL2, VRL3, VRL4, MAT, DT1C, THKE,
     .   THK0, NGL, X1G, X2G, X3G, X4G, Y1G, Y2G, Y3G, Y4G,
     .   Z1G, Z2G, Z3G, Z4G)
      CALL CNVEC3(ELB, DIRA, DIRB, 1, NEL, 0, 1, 1,
     .   X1G, X2G, X3G, X4G, Y1G, Y2G, Y3G, Y4G, Z1G, Z2G, Z3G, Z4G,
     .   E1X, E1Y, E1Z, E2X, E2Y, E2Z, E3X, E3Y, E3Z, NEL)
      CALL CDERI3(1, NEL, SMSTR, OFFG, STI, STIR, AREA, PX1, PX2,
     .   PY1, PY2, X2, X3, X4, Y2, Y3, Y4, Z2, X1G, X2G, X3G, X4G,
     .   Y1G, Y2G, Y3G, Y4G, Z1G, Z2G, Z3G, Z4G, E1X, E1Y, E1Z,
     .   E2X, E2Y, E2Z, E3X, E3Y, E3Z, VHX, VHY, A_I, UX1, UX2, UX3,
     .   UX4, UY1, UY2, UY3, UY4, NEL, ISMS)
      CALL CCOEF3(1, NEL, PM, MAT, GEO, PID, OFF, AREA, STI, STIR,
     .   SHF, THK0, THK02, NU, G, YM, A11, A12, THK, SSP, RHO, H1, H2,
     .   H3, VOL0, VOL00, ALPE, GS, 2, ITHK, ISMS, NPTTOT, 0,
     .   SRH1, SRH2, SRH3, IGEO, A11R, 1, PM_STACK)
      CALL CDLEN3(1, NEL, PM, OFF, AREA, X2, X3, X4, Y2, Y3, Y4, ALDT,
     .   MAT, GEO, PID, IHB)
      CALL CDEFO3(1, NEL, VL1, VL2, VL3, VL4, DT1C, PX1, PX2, PY1, PY2,
     .   AREA, EXX, EYY, EXY, EXZ, EYZ, X2, X3, X4, Y2, Y3, Y4, Z2,
     .   VX1, VX2, VX3, VX4, VY1, VY2, VY3, VY4, VZ1, VZ2, VZ3, VZ4,
     .   E1X, E1Y, E1Z, E2X, E2Y, E2Z, E3X, E3Y, E3Z, IHB)
      CALL CCURV3(1, NEL, VRL1, VRL2, VRL3, VRL4, PX1, PX2, PY1, PY2,
     .   AREA, RX1, RX2, RX3, RX4, RY1, RY2, RY3, RY4, E1X, E1Y, E1Z,
     .   E2X, E2Y, E2Z, E3X, E3Y, E3Z, KXX, KYY, KXY, EXZ, EYZ, WXY,
     .   ISMS)
      CALL CSTRA3(1, NEL, GSTR, SHF, AREA, EXX, EYY, EXY, EXZ, EYZ,
     .   KXX, KYY, KXY, DT1C, TANI, FOR, MOM, ISMS, 2, IHB, 0, 1,
     .   UX1, UX2, UX3, UX4, UY1, UY2, UY3, UY4, PX1, PX2, PY1, PY2,
     .   WXY, GSTRW, F_DEF, NEL)
C-----element strain rate of CFORC3 (cforc3.F:537-546)
      DTINV = DT1 / MAX(DT1**2, EM20)
      DO I = 1, NEL
        EPS_K2 = (KXX(I)**2+KYY(I)**2+KXX(I)*KYY(I)+FOURTH*KXY(I)**2)
     .         * ONE_OVER_9*THK(I)**2
        EPS_M2 = FOUR_OVER_3*(EXX(I)**2+EYY(I)**2+EXX(I)*EYY(I)
     .         + FOURTH*EXY(I)**2)
        EPSD_PG(I) = SQRT(EPS_K2 + EPS_M2)*DTINV
      ENDDO
      EPSDG(1:NEL) = ONE*EPSD_PG(1:NEL) + (ONE-ONE)*EPSDG(1:NEL)
C-----MULAWC law-2 glue
      DO I = 1, NEL
        DEGMB(I) = FOR(I,1)*EXX(I)+FOR(I,2)*EYY(I)+FOR(I,3)*EXY(I)
     .           + FOR(I,4)*EYZ(I)+FOR(I,5)*EXZ(I)
        DEGFX(I) = MOM(I,1)*KXX(I)+MOM(I,2)*KYY(I)+MOM(I,3)*KXY(I)
        VOL0(I) = AREA(I)*THK0(I)
        THKN(I) = THK(I)
        FOR(I,1:5) = ZERO
        MOM(I,1:3) = ZERO
        SIGY(I) = ZERO
        ETSE(I) = ONE
        OFF_OLD(I) = OFF(I)
        IOFF_DUCT(I) = 0
        EPCHK(I) = ZERO
        VISCMX(I) = ZERO
        TEMPEL(I) = ZERO
        FHEAT(I) = ZERO
        LOFF(I) = ONE
        DPLANL(I) = ZERO
        SIGB(I,1:3) = ZERO
        VOLIP(I) = ZERO
      ENDDO
      IF (IPRM(3) > 0) THEN
        ASRT = MIN(ONE, PRM(15)*DT1)
      ELSE
        ASRT = ONE
      ENDIF
      CALL COQINI_WM(WM)
      CALL COQINI_Z0WF(Z0T, WFT)
      DO IP = 1, NPT
        THKLY = WFT(IP,NPT)
        WMC = WM(IP,NPT)
        DO I = 1, NEL
          THKLYL(I) = THKLY*THK0(I)
          ZT = (Z0T(IP,NPT) + ZERO)*THK0(I)
          DEXX(I) = EXX(I)+ZT*KXX(I)
          DEYY(I) = EYY(I)+ZT*KYY(I)
          DEXY(I) = EXY(I)+ZT*KXY(I)
          DEYZ(I) = EYZ(I)
          DEZX(I) = EXZ(I)
          EPX(I) = DEXX(I)*DTINV
          EPY(I) = DEYY(I)*DTINV
          EPXY(I) = DEXY(I)*DTINV
          SO1(I) = SIGL(I,1,IP)
          SO2(I) = SIGL(I,2,IP)
          SO3(I) = SIGL(I,3,IP)
          SO4(I) = SIGL(I,4,IP)
          SO5(I) = SIGL(I,5,IP)
          PLA1(I) = PLAL(I,IP)
          EPS1(I) = EPSDL(I,IP)
          TSTAR(I) = ZERO
        ENDDO
        CALL SIGEPS02C(MATP, NEL, EINT, THKN, TEMPEL, FHEAT, OFF, SIGY,
     .     DT1, IPLA, SIGKSI, VOLIP, GS, THKLYL, ETSE, G_IMP, DPLA,
     .     TSTAR, 0, HARDM, EPCHK, NPT, PLA1, OFF_OLD, LOFF, IOFF_DUCT,
     .     SO1, SO2, SO3, SO4, SO5, SN1, SN2, SN3, SN4, SN5,
     .     DEXX, DEYY, DEXY, DEYZ, DEZX, EPX, EPY, EPXY, EPSD_PG, EPS1,
     .     ASRT, SIGB(1,1), SIGB(1,2), SIGB(1,3), 0, DPLANL)
        DO I = 1, NEL
          SIGL(I,1,IP) = SN1(I)*ONE
          SIGL(I,2,IP) = SN2(I)*ONE
          SIGL(I,3,IP) = SN3(I)*ONE
          SIGL(I,4,IP) = SN4(I)*ONE
          SIGL(I,5,IP) = SN5(I)*ONE
          PLAL(I,IP) = PLA1(I)
          EPSDL(I,IP) = EPS1(I)
          FOR(I,1) = FOR(I,1) + THKLY*SN1(I)
          FOR(I,2) = FOR(I,2) + THKLY*SN2(I)
          FOR(I,3) = FOR(I,3) + THKLY*SN3(I)
          FOR(I,4) = FOR(I,4) + THKLY*SN4(I)
          FOR(I,5) = FOR(I,5) + THKLY*SN5(I)
          MOM(I,1) = MOM(I,1) + WMC*SN1(I)
          MOM(I,2) = MOM(I,2) + WMC*SN2(I)
          MOM(I,3) = MOM(I,3) + WMC*SN3(I)
        ENDDO
      ENDDO
      DO I = 1, NEL
        IF (OFF(I) == FOUR_OVER_5 .AND. IOFF_DUCT(I) == 0 .OR.
     .      OFF(I) > ZERO .AND. OFF_OLD(I) < EM01) OFF(I) = ZERO
        THK(I) = MAX(THKN(I), EM30)
        FOR(I,1:5) = FOR(I,1:5)*OFF(I)*ONE
        MOM(I,1:3) = MOM(I,1:3)*OFF(I)*ONE
        DEGMB(I) = DEGMB(I)+FOR(I,1)*EXX(I)+FOR(I,2)*EYY(I)
     .           + FOR(I,3)*EXY(I)+FOR(I,4)*EYZ(I)+FOR(I,5)*EXZ(I)
        DEGFX(I) = DEGFX(I)+MOM(I,1)*KXX(I)+MOM(I,2)*KYY(I)
     .           + MOM(I,3)*KXY(I)
        VOL2 = HALF*VOL0(I)
        EINT(I,1) = EINT(I,1) + DEGMB(I)*VOL2
        EINT(I,2) = EINT(I,2) + DEGFX(I)*THK0(I)*VOL2
        VISCMX(I) = SQRT(ONE + VISCMX(I)*VISCMX(I)) - VISCMX(I)
      ENDDO
C     EHOUR starts the group from zero, so that it ends as the group's EHOURT
      EHOUR = ZERO
      CALL CHVIS3(1, NEL, THK, HOUR, OFF, PX1, PX2, PY1, PY2,
     .   IXC, DT1C, SSP, RHO, STI, VX1, VX2, VX3, VX4, VY1,
     .   VY2, VY3, VY4, VZ1, VZ2, VZ3, VZ4, AREA, THK0, VHX,
     .   VHY, SHF, Z2, EANI, STIR, VISCMX, G, A11,
     .   H1, H2, H3, YM, NU, THK02, ALPE, H11,
     .   H12, H13, H21, H22, H23, H31, H32, H33,
     .   B11, B12, B13, B14, B21, B22, B23, B24,
     .   RX1, RX2, RX3, RX4, RY1, RY2, RY3, RY4,
     .   IPARTC, PARTSAV, IHB, 0, ISMS, 0,
     .   SRH1, SRH2, SRH3, 1, 0, A11R, NEL)
      EHOU_OUT = EHOUR
      IOFC = 0
      CALL CDT3(1, NEL, YM, OFF, DT2T, NELTST, ITYPTST, STI, STIR, OFFG,
     .   DTCSUB, NDT, DT2C, IXC, SSP, VISCMX, PX1, PX2, PY1, PY2,
     .   VOL0, VOL00, RHO, ALDT, ALPE, INDXOF, NGL, ISMS, IOFC, MSC,
     .   DMELC, 0, 0, DTEL)
      CALL CFINT3(1, NEL, FOR, MOM, THK0, THK02, PX1, PX2, PY1, PY2,
     .   AREA, Z2, F11, F12, F13, F14, F21, F22, F23, F24, F31, F32,
     .   F33, F34, H11, H12, H13, H21, H22, H23, H31, H32, H33,
     .   B11, B12, B13, B14, B21, B22, B23, B24, NEL,
     .   M11, M12, M13, M14, M21, M22, M23, M24, M31, M32, M33, M34,
     .   E1X, E1Y, E1Z, E2X, E2Y, E2Z, E3X, E3Y, E3Z, IHB, NPTTOT,
     .   FZERO)
      CALL CUPDT3P(1, NEL, OFFG, OFF, STI, STIR, FSKY, FSKY, IADC, IXC,
     .   F11, F12, F13, F14, F21, F22, F23, F24, F31, F32, F33, F34,
     .   M11, M12, M13, M14, M21, M22, M23, M24, M31, M32, M33, M34,
     .   EINT, PARTSAV, MAT, IPARTC, PM, AREA, THK, 0, THEM, FTHESKY,
     .   CONDNSKY, CONDE, 0)
      DEALLOCATE(MATP%UPARAM, MATP%IPARAM)
      DEALLOCATE(PM, GEO, PARTSAV, IPM, IGEO)
      RETURN
      END
C
C  Z0 and WF tables of COQINI (coqini.F:46-152), read through COQINI itself.
      SUBROUTINE COQINI_Z0WF(Z0T, WFT)
#include "implicit_f.inc"
#include "com20_c.inc"
      my_real Z0T(11,11), WFT(11,11)
      CALL COQINI
      Z0T = Z0
      WFT = WF
      RETURN
      END
C
C  ACCELE with rotational DOF (accele.F), IRODDL=1.
      SUBROUTINE REF_ACCELE6(NNOD, A, AR, V, MS, IN)
     .   BIND(C, NAME='ref_accele6')
      USE ISO_C_BINDING
#include "implicit_f.inc"
#include "com01_c.inc"
#include "com04_c.inc"
#include "inter22.inc"
      INTEGER(C_INT), VALUE :: NNOD
      REAL(C_DOUBLE) :: A(3,NNOD), AR(3,NNOD), V(3,NNOD), MS(NNOD),
     .   IN(NNOD)
      INTEGER NALE(1), NPBY(1,1)
      my_real MS2D(1)
      NUMNOD = NNOD
      N2D = 0
      IRODDL = 1
      NMULT = 0
      INT22 = 0
      CALL ACCELE(A, AR, V, MS, IN, 1, NALE, MS2D, 1, NPBY)
      RETURN
      END
C
C  VELOCITY (with VR) then DEPLA of the reference (velocity.F, displacement.F).
      SUBROUTINE REF_VELDEP6(NNOD, A, AR, V, VR, X, D, DT2IN, DT12IN)
     .   BIND(C, NAME='ref_veldep6')
      USE ISO_C_BINDING
#include "implicit_f.inc"
#include "com01_c.inc"
#include "com04_c.inc"
#include "com08_c.inc"
#include "scr03_c.inc"
#include "scr05_c.inc"
#include "scr16_c.inc"
#include "inter22.inc"
      INTEGER(C_INT), VALUE :: NNOD
      REAL(C_DOUBLE) :: A(3,NNOD), AR(3,NNOD), V(3,NNOD), VR(3,NNOD),
     .   X(3,NNOD), D(3,NNOD)
      REAL(C_DOUBLE), VALUE :: DT2IN, DT12IN
      INTEGER NALE(1), ITAB(1)
      my_real DR(3,1), FZERO(3,1)
      DOUBLE PRECISION XDP(3,1), DDP(3,1)
      NUMNOD = NNOD
      N2D = 0
      IRODDL = 1
      ISIGI = 0
      INT22 = 0
      IRESP = 0
      ISECUT = 0
      IISROT = 0
      IMPOSE_DR = 0
      IDROT = 0
      DT2 = DT2IN
      DT12 = DT12IN
      CALL VELOCITY(A, AR, V, VR, FZERO, ITAB, NALE)
      CALL DEPLA(V, D, X, VR, DR, XDP, DDP, NNOD)
      RETURN
      END
C
C  Parameter constants of constant_mod.F used by the shell path.
      SUBROUTINE REF_SHELL_CONSTS(OUT) BIND(C, NAME='ref_shell_consts')
      USE ISO_C_BINDING
#include "implicit_f.inc"
      REAL(C_DOUBLE) :: OUT(4)
      OUT(1) = ZEP81
      OUT(2) = ZEP072169
      OUT(3) = SQRT(TWO)*HALF
      OUT(4) = EM30
      RETURN
      END
