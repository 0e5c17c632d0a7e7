C  ref_driver.F -- harness that drives the reference's own compiled routines.
C
C  TEST INFRASTRUCTURE ONLY (built by oracle/build_ref.py into oracle/_ref).
C  It sets the common-block switches of the restated branch (Lagrangian hex,
C  Isolid=1, ISMSTR=4, law 2, Parith/ON, explicit) and calls, for one group
C  of NEL <= 128 elements, the reference routines in the order SFORC3 calls
C  them (sforc3.F:669-1764): SDERI3, SDLEN3, SDEFO3, SRHO3, SROTA3, M2LAW
C  (with MQVISCB and MSTRAIN_RATE inside), SMALLB3, SHVIS3, SFINT3, SCUMU3P.
C  The two pieces it does itself are the nodal gather of SCOOR3 (plain copies)
C  and the law-2 glue of MMAIN (mmain.F90:668-745: VOL_AVG, AMU, RHOREF).
      SUBROUTINE REF_HEX_FORCES(NEL, NNOD, X, V, CONN, SIG, EINT, RHO,
     .   VOL, PLA, EPSD, OFFG, QVIS, NSKY, FSKY, IADS, PRM, IPRM,
     .   DTIN1, DT2T, MSTOP_OUT, EHOU_OUT) BIND(C, NAME='ref_hex_forces')
      USE ISO_C_BINDING
      USE MATPARAM_DEF_MOD
      USE GLOB_THERM_MOD
      USE ALEANIM_MOD, ONLY : FANI_CELL_
      USE MESSAGE_MOD
#include "implicit_f.inc"
#include "mvsiz_p.inc"
#include "param_c.inc"
#include "com01_c.inc"
#include "com04_c.inc"
#include "com06_c.inc"
#include "com08_c.inc"
#include "impl1_c.inc"
#include "parit_c.inc"
#include "scr05_c.inc"
#include "scr07_c.inc"
#include "scr17_c.inc"
#include "scr18_c.inc"
#include "vect01_c.inc"
      INTEGER(C_INT), VALUE :: NEL, NNOD, NSKY
      REAL(C_DOUBLE) :: X(3,NNOD), V(3,NNOD)
      INTEGER(C_INT) :: CONN(8,NEL), IADS(8,NEL), IPRM(4), MSTOP_OUT
      REAL(C_DOUBLE) :: SIG(NEL,6), EINT(NEL), RHO(NEL), VOL(NEL),
     .   PLA(NEL), EPSD(NEL), OFFG(NEL), QVIS(NEL), FSKY(8,NSKY),
     .   PRM(20), DT2T, EHOU_OUT
      REAL(C_DOUBLE), VALUE :: DTIN1
C
      TYPE(MATPARAM_STRUCT_) :: MATP
      TYPE(GLOB_THERM_) :: GTH
      TYPE(FANI_CELL_) :: FANI
      INTEGER I, K, NC
      INTEGER NGL(MVSIZ), MAT(MVSIZ), NGEO(MVSIZ), IPARTS(MVSIZ),
     .   IPARG1(200), IXS(11,1),
     .   NELTST, ITYPTST, NC1(MVSIZ), NC2(MVSIZ), NC3(MVSIZ),
     .   NC4(MVSIZ), NC5(MVSIZ), NC6(MVSIZ), NC7(MVSIZ), NC8(MVSIZ)
      DOUBLE PRECISION XD(MVSIZ,8), YD(MVSIZ,8), ZD(MVSIZ,8),
     .   VOLDP(MVSIZ), VOL0DP(MVSIZ), SAV(MVSIZ,21)
      my_real
     .   OFF(MVSIZ), VOLN(MVSIZ),
     .   XS(MVSIZ,8), YS(MVSIZ,8), ZS(MVSIZ,8),
     .   VX(MVSIZ,8), VY(MVSIZ,8), VZ(MVSIZ,8),
     .   PX(MVSIZ,8), PY(MVSIZ,8), PZ(MVSIZ,8), PXH(MVSIZ,12),
     .   JAC(MVSIZ,6), DELTAX(MVSIZ), NN(MVSIZ,18),
     .   DXX(MVSIZ), DXY(MVSIZ), DXZ(MVSIZ), DYX(MVSIZ), DYY(MVSIZ),
     .   DYZ(MVSIZ), DZX(MVSIZ), DZY(MVSIZ), DZZ(MVSIZ),
     .   D4(MVSIZ), D5(MVSIZ), D6(MVSIZ), WXX(MVSIZ), WYY(MVSIZ),
     .   WZZ(MVSIZ), VXR(MVSIZ), VYR(MVSIZ), VZR(MVSIZ),
     .   DIVDE(MVSIZ), FLUX(1,6), FLU1(MVSIZ), DVOL(MVSIZ),
     .   TAG22(MVSIZ), AMU(MVSIZ), S(MVSIZ,6), VOL_AVG(MVSIZ),
     .   RHOREF(MVSIZ), RHOSP(MVSIZ), STI(MVSIZ), MUMAX(MVSIZ),
     .   SSP(MVSIZ), AIRE(MVSIZ), VD2(MVSIZ), VIS(MVSIZ),
     .   PNEW(MVSIZ), PSH(MVSIZ), QNEW(MVSIZ), SSP_EQ(MVSIZ),
     .   SIGY(MVSIZ), DEFP(MVSIZ), DPLA(MVSIZ), EPSP(MVSIZ),
     .   TSTAR(MVSIZ), ET(MVSIZ), MSSA(MVSIZ), DMELS(MVSIZ),
     .   TEMPEL(MVSIZ), SIGB(MVSIZ,6), AL_IMP(MVSIZ),
     .   SIGNOR(MVSIZ,6), CONDE(MVSIZ), DTEL(MVSIZ), DMG(MVSIZ),
     .   DPDM(MVSIZ), FHEAT(MVSIZ), F(MVSIZ,24), EANI(MVSIZ),
     .   DFE(MVSIZ,3), PBAK(MVSIZ), SVIS(MVSIZ,6),
     .   AR(3,1), FR_WAVE(1), FR_WAV(MVSIZ), MXYZ(MVSIZ,24),
     .   THEM(MVSIZ,8), FTHESKY(1), CONDNSKY(1), EHOUR0
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
      N2D = 0
      INTEG8 = 0
      IRESP = 0
      IMPL_S = 0
      IMCONV = 1
      INCONV = 1
      INEG_V = 0
      IDTS6 = 0
      ISCAU = 0
      IMP_LR = 0
      IDYNA = 0
      ISMDISP = 0
      IVECTOR = 0
C-----message table as build_msg.F allocates it (2 types x message ids), without the texts
C     of the Engine's message file: ANCMSG (SCHKJABT3 on a negative Jacobian) prints the id only
      IF (.NOT. ALLOCATED(MESSAGES)) ALLOCATE(MESSAGES(2, 5000))
      INVSTR = 100
      NUMMAT = 1
      NUMGEO = 1
      NUMNOD = NNOD
      LSKY = NSKY
      IDTMIN(1:102) = 0
      DTFAC1(1) = PRM(19)
      JLAG = 1
      JALE = 0
      JEUL = 0
      JHBE = 1
      ISMSTR = 4
      ISORTH = 0
      ISROT = 0
      MTN = 2
      MSTOP = 0
      FANI%IS_F18_FVM_REQUESTED = .FALSE.
      FANI%IS_VORT_X_REQUESTED = .FALSE.
      FANI%IS_VORT_Y_REQUESTED = .FALSE.
      FANI%IS_VORT_Z_REQUESTED = .FALSE.
C-----material (hm_read_mat02_jc.F90 storage) and property
      PM = ZERO
      GEO = ZERO
      IPM = 0
      IGEO = 0
      IPARG1 = 0
      PM(1,1) = PRM(1)
      GEO(13,1) = PRM(16)
      GEO(14,1) = PRM(14)
      GEO(15,1) = PRM(15)
      GEO(16,1) = PRM(17)
      GEO(17,1) = PRM(18)
      MATP%RHO = PRM(1)
      MATP%RHO0 = PRM(1)
      MATP%BULK = PRM(2)
      MATP%SHEAR = PRM(3)
      MATP%IEOS = 0
      ALLOCATE(MATP%UPARAM(12), MATP%IPARAM(4))
      MATP%UPARAM = ZERO
      MATP%UPARAM(1) = PRM(4)
      MATP%UPARAM(2) = PRM(5)
      MATP%UPARAM(3) = PRM(6)
      MATP%UPARAM(4) = PRM(7)
      MATP%UPARAM(5) = PRM(8)
      MATP%UPARAM(6) = PRM(9)
      MATP%UPARAM(7) = PRM(10)
      MATP%UPARAM(8) = PRM(11)
      MATP%UPARAM(9) = PRM(12)
      MATP%UPARAM(10) = PRM(13)
      MATP%IPARAM(1) = IPRM(4)
      MATP%IPARAM(2) = IPRM(1)
      MATP%IPARAM(3) = IPRM(2)
      MATP%IPARAM(4) = IPRM(3)
      MATP%THERM%TREF = 300.0D0
      MATP%THERM%TMELT = 1.0D20
      MATP%THERM%RHOCP = ZERO
C-----SCOOR3 gather (scoor3.F:109-118, 149-172, 233-258, 286-340)
      DO I = 1, NEL
        NGL(I) = I
        MAT(I) = 1
        NGEO(I) = 1
        IPARTS(I) = 1
        VIS(I) = ZERO
        VD2(I) = ZERO
        OFF(I) = ABS(OFFG(I))
        DO K = 1, 8
          NC = CONN(K,I)
          XD(I,K) = X(1,NC)
          YD(I,K) = X(2,NC)
          ZD(I,K) = X(3,NC)
          XS(I,K) = XD(I,K)
          YS(I,K) = YD(I,K)
          ZS(I,K) = ZD(I,K)
          VX(I,K) = V(1,NC)
          VY(I,K) = V(2,NC)
          VZ(I,K) = V(3,NC)
          IF (OFFG(I) < ZERO) THEN
            VX(I,K) = ZERO
            VY(I,K) = ZERO
            VZ(I,K) = ZERO
          ENDIF
        ENDDO
        NC1(I) = CONN(1,I)
        NC2(I) = CONN(2,I)
        NC3(I) = CONN(3,I)
        NC4(I) = CONN(4,I)
        NC5(I) = CONN(5,I)
        NC6(I) = CONN(6,I)
        NC7(I) = CONN(7,I)
        NC8(I) = CONN(8,I)
      ENDDO
C-----SFORC3 sequence
      CALL SDERI3(OFF, VOLN, NGL,
     .   XD(1,1), XD(1,2), XD(1,3), XD(1,4), XD(1,5), XD(1,6), XD(1,7), XD(1,8),
     .   YD(1,1), YD(1,2), YD(1,3), YD(1,4), YD(1,5), YD(1,6), YD(1,7), YD(1,8),
     .   ZD(1,1), ZD(1,2), ZD(1,3), ZD(1,4), ZD(1,5), ZD(1,6), ZD(1,7), ZD(1,8),
     .   PX(1,1), PX(1,2), PX(1,3), PX(1,4), PY(1,1), PY(1,2), PY(1,3), PY(1,4),
     .   PZ(1,1), PZ(1,2), PZ(1,3), PZ(1,4),
     .   PXH(1,1), PXH(1,2), PXH(1,3), PXH(1,4), PXH(1,5), PXH(1,6),
     .   PXH(1,7), PXH(1,8), PXH(1,9), PXH(1,10), PXH(1,11), PXH(1,12),
     .   JAC(1,1), JAC(1,2), JAC(1,3), JAC(1,4), JAC(1,5), JAC(1,6),
     .   SAV, OFFG, NEL, VOLDP, 1, 4, 1)
      CALL SDLEN3(VOLN, DELTAX,
     .   XS(1,1), XS(1,2), XS(1,3), XS(1,4), XS(1,5), XS(1,6), XS(1,7), XS(1,8),
     .   YS(1,1), YS(1,2), YS(1,3), YS(1,4), YS(1,5), YS(1,6), YS(1,7), YS(1,8),
     .   ZS(1,1), ZS(1,2), ZS(1,3), ZS(1,4), ZS(1,5), ZS(1,6), ZS(1,7), ZS(1,8),
     .   NN(1,1), NN(1,2), NN(1,3), NN(1,4), NN(1,5), NN(1,6),
     .   NN(1,7), NN(1,8), NN(1,9), NN(1,10), NN(1,11), NN(1,12),
     .   NN(1,13), NN(1,14), NN(1,15), NN(1,16), NN(1,17), NN(1,18),
     .   NEL, 2, 0, 0)
      CALL SDEFO3(FANI,
     .   PX(1,1), PX(1,2), PX(1,3), PX(1,4), PY(1,1), PY(1,2), PY(1,3), PY(1,4),
     .   PZ(1,1), PZ(1,2), PZ(1,3), PZ(1,4),
     .   PX(1,5), PX(1,6), PX(1,7), PX(1,8), PY(1,5), PY(1,6), PY(1,7), PY(1,8),
     .   PZ(1,5), PZ(1,6), PZ(1,7), PZ(1,8),
     .   VX(1,1), VX(1,2), VX(1,3), VX(1,4), VX(1,5), VX(1,6), VX(1,7), VX(1,8),
     .   VY(1,1), VY(1,2), VY(1,3), VY(1,4), VY(1,5), VY(1,6), VY(1,7), VY(1,8),
     .   VZ(1,1), VZ(1,2), VZ(1,3), VZ(1,4), VZ(1,5), VZ(1,6), VZ(1,7), VZ(1,8),
     .   DXX, DXY, DXZ, DYX, DYY, DYZ, DZX, DZY, DZZ, D4, D5, D6,
     .   WXX, WYY, WZZ, VXR, VYR, VZR, NEL, 0, 4, 0, 1, 0, 0)
      DO I = 1, NEL
        DIVDE(I) = DT1*(DXX(I)+DYY(I)+DZZ(I))
      ENDDO
      CALL SRHO3(PM, VOL, RHO, EINT, DIVDE, FLUX, FLU1, VOLN, DVOL,
     .   NGL, MAT, OFF, 0, TAG22, VOLDP, VOL0DP, AMU, OFFG, NEL, 2, 0,
     .   4, 0, 1, 1, 6, 0)
      CALL SROTA3(SIG, S(1,1), S(1,2), S(1,3), S(1,4), S(1,5), S(1,6),
     .   WXX, WYY, WZZ, NEL, 2, 4)
C-----MMAIN law-2 glue (mmain.F90:668-745)
      DO I = 1, NEL
        VOL_AVG(I) = VOLN(I) - HALF*DVOL(I)
        AMU(I) = RHO(I)/MATP%RHO - ONE
        RHOREF(I) = RHO(I)
        RHOSP(I) = RHO(I)
        TSTAR(I) = ZERO
        DPLA(I) = ZERO
        ET(I) = ONE
        STI(I) = ZERO
        DMG(I) = ZERO
        SVIS(I,1:6) = ZERO
      ENDDO
      CALL M2LAW(MATP, PM, OFF, SIG, EINT,
     .   RHO, QVIS, PLA, EPSD,
     .   VOL, STI, DT2T, NELTST,
     .   ITYPTST, OFFG, GEO, NGEO,
     .   AMU, VOL_AVG, MUMAX, MAT,
     .   NGL, SSP, DVOL, AIRE,
     .   VOLN, VD2, DELTAX, VIS,
     .   DXX, DYY, DZZ, D4,
     .   D5, D6, PNEW, PSH,
     .   QNEW, SSP_EQ, S(1,1), S(1,2),
     .   S(1,3), S(1,4), S(1,5), S(1,6),
     .   SIGY, DEFP, DPLA, 1,
     .   EPSP, TSTAR, ET, MSSA,
     .   DMELS, TEMPEL, SIGB, AL_IMP,
     .   SIGNOR, CONDE, DTEL, 0,
     .   NEL, IPM, RHOREF, RHOSP,
     .   1, DMG, 1, 0,
     .   0, 0, 4, 0,
     .   1, DPDM, FHEAT, GTH,
     .   1, IGEO)
      CALL SMALLB3(OFFG, OFF, NEL, 4)
C     EHOUR starts the group from zero, so that it ends as the group's EHOURT
      EHOUR0 = ZERO
      EHOUR = ZERO
      PARTSAV = ZERO
      EANI = ZERO
      CALL SHVIS3(PM, RHO, OFF,
     .   VX(1,1), VX(1,2), VX(1,3), VX(1,4), VX(1,5), VX(1,6), VX(1,7), VX(1,8),
     .   VY(1,1), VY(1,2), VY(1,3), VY(1,4), VY(1,5), VY(1,6), VY(1,7), VY(1,8),
     .   VZ(1,1), VZ(1,2), VZ(1,3), VZ(1,4), VZ(1,5), VZ(1,6), VZ(1,7), VZ(1,8),
     .   F(1,1), F(1,2), F(1,3), F(1,4), F(1,5), F(1,6), F(1,7), F(1,8),
     .   F(1,9), F(1,10), F(1,11), F(1,12), F(1,13), F(1,14), F(1,15),
     .   F(1,16), F(1,17), F(1,18), F(1,19), F(1,20), F(1,21), F(1,22),
     .   F(1,23), F(1,24),
     .   PXH(1,1), PXH(1,2), PXH(1,3), PXH(1,4), PXH(1,5), PXH(1,6),
     .   PXH(1,7), PXH(1,8), PXH(1,9), PXH(1,10), PXH(1,11), PXH(1,12),
     .   VOLN, MAT, SSP, VIS, VD2, DELTAX, EANI, NGEO, GEO,
     .   PARTSAV, IPARTS, OFFG, VOL, IPARG1, 0, NEL, 0, 2, 4, 1, 1)
      EHOU_OUT = EHOUR - EHOUR0
      CALL SFINT3(SIG,
     .   PX(1,1), PX(1,2), PX(1,3), PX(1,4), PY(1,1), PY(1,2), PY(1,3), PY(1,4),
     .   PZ(1,1), PZ(1,2), PZ(1,3), PZ(1,4),
     .   PX(1,5), PX(1,6), PX(1,7), PX(1,8), PY(1,5), PY(1,6), PY(1,7), PY(1,8),
     .   PZ(1,5), PZ(1,6), PZ(1,7), PZ(1,8),
     .   F(1,1), F(1,2), F(1,3), F(1,4), F(1,5), F(1,6), F(1,7), F(1,8),
     .   F(1,9), F(1,10), F(1,11), F(1,12), F(1,13), F(1,14), F(1,15),
     .   F(1,16), F(1,17), F(1,18), F(1,19), F(1,20), F(1,21), F(1,22),
     .   F(1,23), F(1,24), VOLN, QNEW,
     .   NN(1,1), NN(1,2), NN(1,3), NN(1,4), NN(1,5), NN(1,6),
     .   NN(1,7), NN(1,8), NN(1,9), NN(1,10), NN(1,11), NN(1,12),
     .   NN(1,13), NN(1,14), NN(1,15), NN(1,16), NN(1,17), NN(1,18),
     .   DFE, PBAK, IXS, NEL, 0, 0, 0, SVIS)
      CALL SCUMU3P(OFFG, STI, FSKY, FSKY, IADS,
     .   F(1,1), F(1,2), F(1,3), F(1,4), F(1,5), F(1,6), F(1,7), F(1,8),
     .   F(1,9), F(1,10), F(1,11), F(1,12), F(1,13), F(1,14), F(1,15),
     .   F(1,16), F(1,17), F(1,18), F(1,19), F(1,20), F(1,21), F(1,22),
     .   F(1,23), F(1,24),
     .   NC1, NC2, NC3, NC4, NC5, NC6, NC7, NC8, AR, FR_WAVE, FR_WAV,
     .   MXYZ(1,1), MXYZ(1,2), MXYZ(1,3), MXYZ(1,4), MXYZ(1,5),
     .   MXYZ(1,6), MXYZ(1,7), MXYZ(1,8), MXYZ(1,9), MXYZ(1,10),
     .   MXYZ(1,11), MXYZ(1,12), MXYZ(1,13), MXYZ(1,14), MXYZ(1,15),
     .   MXYZ(1,16), MXYZ(1,17), MXYZ(1,18), MXYZ(1,19), MXYZ(1,20),
     .   MXYZ(1,21), MXYZ(1,22), MXYZ(1,23), MXYZ(1,24),
     .   THEM, FTHESKY, CONDNSKY, CONDE, NEL, 0, 0, 0, 0, 0)
      MSTOP_OUT = MSTOP
      DEALLOCATE(MATP%UPARAM, MATP%IPARAM)
      DEALLOCATE(PM, GEO, PARTSAV, IPM, IGEO)
      RETURN
      END
C
C  ACCELE of the reference on the whole node set (accele.F).
      SUBROUTINE REF_ACCELE(NNOD, A, V, MS) BIND(C, NAME='ref_accele')
      USE ISO_C_BINDING
#include "implicit_f.inc"
#include "com01_c.inc"
#include "com04_c.inc"
#include "inter22.inc"
      INTEGER(C_INT), VALUE :: NNOD
      REAL(C_DOUBLE) :: A(3,NNOD), V(3,NNOD), MS(NNOD)
      INTEGER NALE(1), NPBY(1,1)
      my_real AR(3,1), IN(1), MS2D(1)
      NUMNOD = NNOD
      N2D = 0
      IRODDL = 0
      NMULT = 0
      INT22 = 0
      CALL ACCELE(A, AR, V, MS, IN, 1, NALE, MS2D, 1, NPBY)
      RETURN
      END
C
C  VELOCITY then DEPLA of the reference (velocity.F, displacement.F).
      SUBROUTINE REF_VELDEP(NNOD, A, V, X, D, DT2IN, DT12IN)
     .   BIND(C, NAME='ref_veldep')
      USE ISO_C_BINDING
#include "implicit_f.inc"
#include "com01_c.inc"
#include "com04_c.inc"
#include "com08_c.inc"
#include "scr05_c.inc"
#include "inter22.inc"
      INTEGER(C_INT), VALUE :: NNOD
      REAL(C_DOUBLE) :: A(3,NNOD), V(3,NNOD), X(3,NNOD), D(3,NNOD)
      REAL(C_DOUBLE), VALUE :: DT2IN, DT12IN
      INTEGER NALE(1), ITAB(1)
      my_real AR(3,1), VR(3,1), DR(3,1), FZERO(3,1)
      DOUBLE PRECISION XDP(3,1), DDP(3,1)
      NUMNOD = NNOD
      N2D = 0
      IRODDL = 0
      ISIGI = 0
      INT22 = 0
      IRESP = 0
      DT2 = DT2IN
      DT12 = DT12IN
      CALL VELOCITY(A, AR, V, VR, FZERO, ITAB, NALE)
      CALL DEPLA(V, D, X, VR, DR, XDP, DDP, NNOD)
      RETURN
      END
