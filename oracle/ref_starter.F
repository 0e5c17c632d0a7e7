C     Harness around the reference Starter's own element initialisation routines, compiled from
C     the reference sources into oracle/_ref/libref_starter.so by oracle/build_ref.py.
C     TEST INFRASTRUCTURE ONLY: it writes the Engine inputs of this path the way the Starter
C     writes them, to check the model builder against.
C
C     REF_ST_VOLUME  the solid volume of SDERI3 (VOLDP = DET = GBUF%VOL, starter sderi3.F:108-155),
C                    as SINIT3 calls it (sinit3.F:322) for groups of at most 128 solids;
C     REF_ST_MASS    the nodal masses and inertias of a model of solids and shells: SDERI3 and
C                    SMASS3 (sinit3.F:322, 434) for the solids' corner masses, CEVECI and CINMAS
C                    (cinit3.F:278, 302) for the shells' masses and inertias, then SPMD_MSIN's
C                    sums over the elements in user-id order (initia.F:1757).
C
C     Every Starter setting these routines read is set to the value a deck on this path gives it
C     (Lagrangian, no ALE/Euler, no thermal, no mass scaling or added mass, no drape, no adaptive
C     meshing, Parith/ON by element ids, one property and one material).
      SUBROUTINE REF_ST_SETUP(NN, NELS, NELC)
      USE DRAPE_MOD, ONLY : NUMELC_DRAPE, NUMELTG_DRAPE, SCDRAPE
#include "implicit_f.inc"
#include "com01_c.inc"
#include "com04_c.inc"
#include "param_c.inc"
#include "scr03_c.inc"
#include "scr12_c.inc"
#include "scr17_c.inc"
#include "scr22_c.inc"
#include "remesh_c.inc"
#include "vect01_c.inc"
#include "drape_c.inc"
      INTEGER NN, NELS, NELC
      NPROPG = 1000
      NPROPM = 250
      LTITR = 40
      NPROPGI = 750 + LTITR
      LVEUL = 32
      KSH4TREE = 4
      KSH3TREE = 4
      NIXIG3D = 15
      NUMNOD = NN
      NUMELS = NELS
      NUMELS8 = NELS
      NUMELS10 = 0
      NUMELS20 = 0
      NUMELS16 = 0
      NUMELQ = 0
      NUMELC = NELC
      NUMELT = 0
      NUMELP = 0
      NUMELR = 0
      NUMELTG = 0
      NUMELIG3D = 0
      NUMGEO = 1
      NUMMAT = 1
      NPERTURB = 0
      IRODDL = 0
      IF (NELC > 0) IRODDL = 1
      IALE = 0
      IEULER = 0
      IALELAG = 0
      ISHFRAM = 0
      IREST_MSELT = 0
      I7STIFS = 0
      NADMESH = 0
      ISTATCND = 0
      INER_9_12 = ZERO
      NXREF = 0
      NDRAPE = 0
      NUMELC_DRAPE = 0
      NUMELTG_DRAPE = 0
      SCDRAPE = 0
      JALE = 0
      JEUL = 0
      JLAG = 1
      JTHE = 0
      ISROT = 0
      ISHXFEM_PLY = 0
      RETURN
      END

C     The Starter's SDERI3 volume of NEL solids, IXS(8,NEL) 1-based node ids, X(3,*).
      SUBROUTINE REF_ST_VOLUME(NEL, IXS, X, VOL)
#include "implicit_f.inc"
#include "mvsiz_p.inc"
#include "com04_c.inc"
#include "param_c.inc"
      INTEGER NEL, IXS(8,*)
      DOUBLE PRECISION X(3,*), VOL(*)
      INTEGER G0, N, I, K, NGL(MVSIZ), NGEO(MVSIZ)
      INTEGER, ALLOCATABLE :: IGEO(:,:)
      DOUBLE PRECISION, ALLOCATABLE :: GEO(:,:)
      DOUBLE PRECISION VEUL(32,1)
      DOUBLE PRECISION XD(MVSIZ,8), YD(MVSIZ,8), ZD(MVSIZ,8)
      DOUBLE PRECISION J1(MVSIZ),J2(MVSIZ),J3(MVSIZ),J4(MVSIZ),J5(MVSIZ),J6(MVSIZ),
     .   P(MVSIZ,12), DET(MVSIZ), VDP(MVSIZ), GV(MVSIZ)
      CALL REF_ST_SETUP(0, NEL, 0)
      ALLOCATE(IGEO(NPROPGI,1), GEO(NPROPG,1))
      IGEO = 0
      GEO = ZERO
      DO G0 = 0, NEL - 1, 128
        N = MIN(128, NEL - G0)
        DO I = 1, N
          NGL(I) = G0 + I
          NGEO(I) = 1
          DO K = 1, 8
            XD(I,K) = X(1,IXS(K,G0+I))
            YD(I,K) = X(2,IXS(K,G0+I))
            ZD(I,K) = X(3,IXS(K,G0+I))
          ENDDO
        ENDDO
        CALL SDERI3(GV, VEUL, GEO, IGEO,
     .    XD(1,1),XD(1,2),XD(1,3),XD(1,4),XD(1,5),XD(1,6),XD(1,7),XD(1,8),
     .    YD(1,1),YD(1,2),YD(1,3),YD(1,4),YD(1,5),YD(1,6),YD(1,7),YD(1,8),
     .    ZD(1,1),ZD(1,2),ZD(1,3),ZD(1,4),ZD(1,5),ZD(1,6),ZD(1,7),ZD(1,8),
     .    J1, J2, J3, J4, J5, J6, NGL, NGEO,
     .    P(1,1),P(1,2),P(1,3),P(1,4),P(1,5),P(1,6),P(1,7),P(1,8),
     .    P(1,9),P(1,10),P(1,11),P(1,12), DET, VDP, N, 0, 0, 0)
        DO I = 1, N
          VOL(G0+I) = GV(I)
        ENDDO
      ENDDO
      DEALLOCATE(IGEO, GEO)
      RETURN
      END

C     Nodal masses MS and inertias IN of NELS solids (IXS(11,NELS): row 1 material, rows 2-9
C     1-based nodes, row 10 property, row 11 user id; density RHOS) and NELC shells (IXC(7,NELC):
C     row 1 material, rows 2-5 nodes, row 6 property, row 7 user id; thickness THKE, density
C     RHOC, shell formulation IHBE), as the Starter computes them.  Also returned: the solids'
C     volumes VOLS, corner masses MSS(8,NELS), and the shells' areas AREAC, masses MSC and
C     inertias INC.
      SUBROUTINE REF_ST_MASS(NN, X, NELS, IXS, RHOS, NELC, IXC, THKE, RHOC, IHBE,
     .                       MS, IN, VOLS, MSS, AREAC, MSC, INC)
      USE ELBUFDEF_MOD
      USE STACK_MOD
      USE DRAPE_MOD
#include "implicit_f.inc"
#include "mvsiz_p.inc"
#include "com04_c.inc"
#include "param_c.inc"
#include "vect01_c.inc"
      INTEGER NN, NELS, NELC, IHBE, IXS(11,*), IXC(7,*)
      DOUBLE PRECISION X(3,*), RHOS, RHOC, THKE(*), MS(*), IN(*), VOLS(*), MSS(8,*),
     .   AREAC(*), MSC(*), INC(*)
      INTEGER G0, N, I, K, IPART(MVSIZ), NC(MVSIZ,8), IX1(MVSIZ), IX2(MVSIZ), IX3(MVSIZ),
     .   IX4(MVSIZ), INDX(1), PERTURB(1), ISUBSTACK, NLAY, IDRAPE, NINTEMP, DUMI(1)
      INTEGER, ALLOCATABLE :: IGEO(:,:), INDEX(:), ITRI(:), SH4TREE(:,:), NSHNOD(:), IXSS(:,:),
     .   NOD8(:,:)
      INTEGER*8 I8MI(6,1)
      DOUBLE PRECISION, ALLOCATABLE :: GEO(:,:), PM(:,:), V(:,:), VR(:,:), INS(:,:)
      DOUBLE PRECISION PARTSAV(20,2), RHO(MVSIZ), FILL(MVSIZ), FRAC(MVSIZ), RHOF(MVSIZ),
     .   RHOCP(MVSIZ), VOLU(MVSIZ), MSNF(1), MSSF(8,MVSIZ), WMA(1), MCP(1), MCPS(8,MVSIZ),
     .   MSSA(1), AREA(MVSIZ), XC(MVSIZ,4), YC(MVSIZ,4), ZC(MVSIZ,4), E(MVSIZ,9), THK(MVSIZ),
     .   THKI(MVSIZ), XREFC(4,3,MVSIZ), ETNOD(1), STC(MVSIZ), TEMP(1), BID(1), RNOISE(1,1),
     .   DUM(1)
      TYPE(ELBUF_STRUCT_), TARGET :: EB
      TYPE(STACK_PLY) :: STK
      TYPE(DRAPE_), DIMENSION(1), TARGET :: DRP
      CALL REF_ST_SETUP(NN, NELS, NELC)
      ALLOCATE(IGEO(NPROPGI,1), GEO(NPROPG,1), PM(NPROPM,1), V(3,NN), VR(3,NN))
      ALLOCATE(INS(8,MAX(NELS,1)), INDEX(2*MAX(NELS,NELC,1)), ITRI(MAX(NELS,NELC,1)))
      ALLOCATE(SH4TREE(4,MAX(NELC,1)), NSHNOD(NN), IXSS(11,MAX(NELS,1)), NOD8(8,MAX(NELS,1)))
      IGEO = 0
      GEO = ZERO
      PM = ZERO
      V = ZERO
      VR = ZERO
      INS = ZERO
      SH4TREE = 0
      NSHNOD = 0
      PARTSAV = ZERO
      I8MI = 0
      DO I = 1, NN
        MS(I) = ZERO
        IN(I) = ZERO
      ENDDO
C     solids: SINIT3 per group (volume, then SMASS3 with FILL = 1)
      IF (NELS > 0) THEN
        DO I = 1, NELS
          DO K = 1, 8
            NOD8(K,I) = IXS(K+1,I)
          ENDDO
          DO K = 1, 11
            IXSS(K,I) = IXS(K,I)
          ENDDO
        ENDDO
        CALL REF_ST_VOLUME(NELS, NOD8, X, VOLS)
        CALL REF_ST_SETUP(NN, NELS, NELC)
      ENDIF
      DO G0 = 0, NELS - 1, 128
        N = MIN(128, NELS - G0)
        LFT = 1
        LLT = N
        NFT = G0
        DO I = 1, N
          RHO(I) = RHOS
          FILL(I) = ONE
          FRAC(I) = ZERO
          RHOF(I) = ZERO
          RHOCP(I) = ZERO
          VOLU(I) = VOLS(G0+I)
          IPART(I) = 1
          DO K = 1, 8
            NC(I,K) = IXS(K+1,G0+I)
          ENDDO
        ENDDO
        CALL SMASS3(RHO, MS, PARTSAV, X, V, IPART, MSS(1,G0+1), VOLU,
     .              MSNF, MSSF, IN, VR, INS(1,G0+1), WMA, RHOCP, MCP,
     .              MCPS, MSSA, RHOF, FRAC, FILL,
     .              NC(1,1), NC(1,2), NC(1,3), NC(1,4), NC(1,5), NC(1,6), NC(1,7), NC(1,8))
      ENDDO
C     shells: CINIT3 per group (CEVECI area, then CINMAS), property type 1, material 1
      GEO(12,1) = ONE
      PM(89,1) = RHOC
      ISUBSTACK = 0
      NLAY = 1
      IDRAPE = 0
      NINTEMP = 0
      INDX(1) = 0
      PERTURB(1) = 0
      DO G0 = 0, NELC - 1, 128
        N = MIN(128, NELC - G0)
        LFT = 1
        LLT = N
        NFT = G0
        NPT = 1
        DO I = 1, N
          IPART(I) = 2
          IX1(I) = IXC(2,G0+I)
          IX2(I) = IXC(3,G0+I)
          IX3(I) = IXC(4,G0+I)
          IX4(I) = IXC(5,G0+I)
          XC(I,1) = X(1,IX1(I))
          XC(I,2) = X(1,IX2(I))
          XC(I,3) = X(1,IX3(I))
          XC(I,4) = X(1,IX4(I))
          YC(I,1) = X(2,IX1(I))
          YC(I,2) = X(2,IX2(I))
          YC(I,3) = X(2,IX3(I))
          YC(I,4) = X(2,IX4(I))
          ZC(I,1) = X(3,IX1(I))
          ZC(I,2) = X(3,IX2(I))
          ZC(I,3) = X(3,IX3(I))
          ZC(I,4) = X(3,IX4(I))
          THK(I) = THKE(G0+I)
        ENDDO
        CALL CEVECI(LFT, LLT, AREA,
     .              XC(1,1), XC(1,2), XC(1,3), XC(1,4), YC(1,1), YC(1,2),
     .              YC(1,3), YC(1,4), ZC(1,1), ZC(1,2), ZC(1,3), ZC(1,4),
     .              E(1,1), E(1,2), E(1,3), E(1,4), E(1,5), E(1,6), E(1,7), E(1,8), E(1,9))
        DO I = 1, N
          AREAC(G0+I) = AREA(I)
        ENDDO
        CALL CINMAS(X, XREFC, IXC, GEO, PM,
     .              MS, IN, THK, IHBE, PARTSAV,
     .              V, IPART, MSC(G0+1), INC(G0+1), AREA,
     .              I8MI, IGEO, ETNOD, 1, 1,
     .              NSHNOD, STC, SH4TREE, MCP, MCPS,
     .              TEMP, BID, BID, BID, BID,
     .              BID, BID, ISUBSTACK, NLAY, EB,
     .              STK, THKI, RNOISE, DRP, NINTEMP,
     .              PERTURB, IX1, IX2, IX3, IX4,
     .              IDRAPE, INDX)
      ENDDO
C     nodal sums in element-id order (SPMD_MSIN); the arrays of other element types are empty
      CALL SPMD_MSIN(
     1     IXSS, DUMI, IXC, DUMI, DUMI,
     2     DUMI, DUMI, MSS, DUM,
     3     MSC, DUM, DUM, DUM, DUM,
     4     INC, DUM, DUM, DUM,
     5     INDEX, ITRI, MS, IN,
     6     DUM, GEO, DUMI, DUMI,
     7     DUMI, DUM, DUM, DUM, DUM,
     8     DUM, DUM, DUM, DUM, DUM,
     9     DUM, DUM, DUM, DUM, DUM,
     A     DUM, DUM, DUM, INS,
     B     DUM, DUM, DUM, DUM,
     C     DUM, SH4TREE, DUMI, DUM, DUM,
     D     DUM, DUM, DUM, DUM, DUM,
     E     DUMI, DUMI, DUM, 1, DUM,
     F     DUM, DUM, DUM, DUM, NSHNOD, DUM, DUM,
     G     DUM, 0)
      DEALLOCATE(IGEO, GEO, PM, V, VR, INS, INDEX, ITRI, SH4TREE, NSHNOD, IXSS, NOD8)
      RETURN
      END
