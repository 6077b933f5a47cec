************************************************************************
*     ref_driver.f -- test harness (TEST INFRASTRUCTURE ONLY).
*
*     Runs the reference's own density routines, compiled from the
*     reference sources (grids.f, particles.f, num.f, haloes_grids.f)
*     without OpenMP, on one input file and writes the fields they fill.
*       mode 1: SUBROUTINE DENSITY            (grids.f:710-994)
*       mode 2: SUBROUTINE INTERPOLATE_DENSITY_TSC (grids.f:1070-1398)
*       mode 3: SUBROUTINE SORT_DM_PARTICLES_LOCALDENSITY (reader.f:140-336),
*               with RODO = SCAL(2) and RE0 = SCAL(3); writes DX, DY, DZ,
*               NPART_ESP(0:8) and ORIPA(1:N) (the original index of the
*               particle at each position after the partition)
*       mode 4: SUBROUTINE SORT_CELLS_AMR (haloes_grids.f:1383-1425) on
*               K = hdr(6) candidate cells: input real*4 ddd(K), int32
*               dddx(K), dddy(K), dddz(K), dddp(K) after the header;
*               output the same five arrays after the sort
*       mode 5: as mode 2, then the grid masks and the density dump of
*               asohf.f:515-545: VEINSGRID for IR = 1..NL (grids.f:
*               2229-2506), COMPUTE_CR0AMR (grids.f:2076-2226) and the
*               FW2 records (U1, CR0AMR, then per patch U11, CR0AMR11,
*               SOLAP) written to the output file as an unformatted
*               sequential file.  Extra input after the particles: int32
*               patchx(P), patchy(P), patchz(P), pare(P).  Needs a large
*               stack (VEINSGRID's VECINO is NPALEV x NPALEV).
*       mode 6: SUBROUTINE CREATE_MESH (grids.f:2-707) with NL_MESH =
*               hdr(5) on the particles (P = 0, species from npart_esp).
*               Extra input after the particles: int32 refine_thr,
*               min_patchsize, bor, boramr, bor_ovlp; real*4
*               frac_refinable; int32 mpapolev(NLEVELS).  Output: int32
*               nl_mesh, npatch(0:NLEVELS), then for the P patches int32
*               pnx, pny, pnz, px, py, pz, pare and real*4 prx, pry, prz.
*     The grid constants come from the reference's MALLA (grids.f:2509)
*     called with LADO = LADO0 - LADO0/NX as asohf.f:323 does.
*
*     Usage: asohf_ref_density <input file> <output file>
*     Input (stream, native endian):
*       int32 hdr(8) = mode, nx, ny, nz, nl, n, interp_degree, P
*       int32 npatch(0:nl), int32 npart_esp(0:8)
*       real*4 lado0, rote, rete, fdm
*       int32 pnx(P), pny(P), pnz(P); real*4 prx(P), pry(P), prz(P)
*       real*4 x(n), y(n), z(n), m(n)
*     Output (stream):
*       real*4 dx, dy, dz; real*4 radx(0:nx+1)
*       real*4 u1(1:nx,1:ny,1:nz); then for p = 1..P u11(1:n1,1:n2,1:n3,p)
*     Fields are divided by FDM as RENORM_DENSITY (grids.f:2668) does.
************************************************************************
      PROGRAM REF_DRIVER
      USE PARTICLES
      IMPLICIT NONE
      INCLUDE 'input_files/asohf_parameters.dat'

      REAL*4 U1(NMAX,NMAY,NMAZ)
      REAL*4 U11(NAMRX,NAMRY,NAMRZ,NPALEV)
      COMMON /VARIA/ U1,U11
      REAL*4 DX,DY,DZ
      COMMON /ESPACIADO/ DX,DY,DZ
      REAL*4 RETE,HTE,ROTE
      COMMON /BACK/ RETE,HTE,ROTE
      REAL*4 RADX(0:NMAX+1),RADY(0:NMAY+1),RADZ(0:NMAZ+1)
      COMMON /GRID/ RADX,RADY,RADZ

      INTEGER HDR(8),MODE,NX,NY,NZ,NL,N,IDEG,NPT,I,IP,N1,N2,N3
      INTEGER IX,JY,KZ
      INTEGER NPATCH(0:NLEVELS),PARE(NPALEV)
      INTEGER PATCHNX(NPALEV),PATCHNY(NPALEV),PATCHNZ(NPALEV)
      INTEGER PATCHX(NPALEV),PATCHY(NPALEV),PATCHZ(NPALEV)
      REAL*4 PATCHRX(NPALEV),PATCHRY(NPALEV),PATCHRZ(NPALEV)
      INTEGER NPART_ESP(0:N_ESP-1)
      REAL*4 SCAL(4),LADO0,LADO,FDM,T,ZETA
      CHARACTER*512 FIN,FOUT
      REAL*4, ALLOCATABLE :: DDD(:)
      INTEGER SOLAP(NAMRX,NAMRY,NAMRZ,NPALEV)
      INTEGER CR0AMR(NMAX,NMAY,NMAZ)
      INTEGER CR0AMR11(NAMRX,NAMRY,NAMRZ,NPALEV)
      INTEGER IR
      INTEGER RTHR,MINPS,BOR,BORAMR,BOVLP,MPAPOLEV(NLEVELS),NLM
      REAL*4 FRACR
      INTEGER, ALLOCATABLE :: DDDX(:),DDDY(:),DDDZ(:),DDDP(:)

      CALL GET_COMMAND_ARGUMENT(1,FIN)
      CALL GET_COMMAND_ARGUMENT(2,FOUT)
      OPEN(10,FILE=FIN,ACCESS='STREAM',FORM='UNFORMATTED',
     &     STATUS='OLD',ACTION='READ')
      READ(10) HDR
      MODE=HDR(1)
      IF (MODE.EQ.4) THEN
       N=HDR(6)
       ALLOCATE(DDD(N),DDDX(N),DDDY(N),DDDZ(N),DDDP(N))
       READ(10) (DDD(I),I=1,N)
       READ(10) (DDDX(I),I=1,N)
       READ(10) (DDDY(I),I=1,N)
       READ(10) (DDDZ(I),I=1,N)
       READ(10) (DDDP(I),I=1,N)
       CLOSE(10)
       CALL SORT_CELLS_AMR(N,DDD,DDDX,DDDY,DDDZ,DDDP)
       OPEN(11,FILE=FOUT,ACCESS='STREAM',FORM='UNFORMATTED',
     &      STATUS='REPLACE',ACTION='WRITE')
       WRITE(11) (DDD(I),I=1,N)
       WRITE(11) (DDDX(I),I=1,N)
       WRITE(11) (DDDY(I),I=1,N)
       WRITE(11) (DDDZ(I),I=1,N)
       WRITE(11) (DDDP(I),I=1,N)
       CLOSE(11)
       STOP
      END IF
      NX=HDR(2)
      NY=HDR(3)
      NZ=HDR(4)
      NL=HDR(5)
      N=HDR(6)
      IDEG=HDR(7)
      NPT=HDR(8)
      IF (NX.GT.NMAX.OR.NY.GT.NMAY.OR.NZ.GT.NMAZ.OR.NL.GT.NLEVELS
     &    .OR.NPT.GT.NPALEV) THEN
       WRITE(*,*) 'ref_driver: input exceeds compile-time bounds'
       STOP 2
      END IF
      NPATCH=0
      READ(10) (NPATCH(I),I=0,NL)
      READ(10) NPART_ESP
      READ(10) SCAL
      LADO0=SCAL(1)
      ROTE=SCAL(2)
      RETE=SCAL(3)
      FDM=SCAL(4)
      HTE=0.0
      PARE=0
      PATCHX=0
      PATCHY=0
      PATCHZ=0
      IF (NPT.GT.0) THEN
       READ(10) (PATCHNX(I),I=1,NPT)
       READ(10) (PATCHNY(I),I=1,NPT)
       READ(10) (PATCHNZ(I),I=1,NPT)
       READ(10) (PATCHRX(I),I=1,NPT)
       READ(10) (PATCHRY(I),I=1,NPT)
       READ(10) (PATCHRZ(I),I=1,NPT)
      END IF
      DO I=1,NPT
       IF (PATCHNX(I).GT.NAMRX.OR.PATCHNY(I).GT.NAMRY.OR.
     &     PATCHNZ(I).GT.NAMRZ) THEN
        WRITE(*,*) 'ref_driver: patch exceeds NAMRX/Y/Z'
        STOP 2
       END IF
      END DO
      PARTI=MAX(N,1)
      ALLOCATE(RXPA(PARTI),RYPA(PARTI),RZPA(PARTI),MASAP(PARTI))
      IF (N.GT.0) THEN
       READ(10) (RXPA(I),I=1,N)
       READ(10) (RYPA(I),I=1,N)
       READ(10) (RZPA(I),I=1,N)
       READ(10) (MASAP(I),I=1,N)
       IF (MODE.EQ.6) THEN
        READ(10) RTHR,MINPS,BOR,BORAMR,BOVLP
        READ(10) FRACR
        READ(10) MPAPOLEV
       END IF
       IF (MODE.EQ.5) THEN
        READ(10) (PATCHX(I),I=1,NPT)
        READ(10) (PATCHY(I),I=1,NPT)
        READ(10) (PATCHZ(I),I=1,NPT)
        READ(10) (PARE(I),I=1,NPT)
       END IF
      END IF
      CLOSE(10)

      LADO=LADO0-(LADO0/NX)
      CALL MALLA(NX,NY,NZ,LADO)
      T=0.0
      ZETA=0.0

      IF (MODE.EQ.6) THEN
       NLM=NL
       NPATCH=0
       CALL CREATE_MESH(0,NX,NY,NZ,NLM,NPATCH,PARE,
     &           PATCHNX,PATCHNY,PATCHNZ,PATCHX,PATCHY,PATCHZ,
     &           PATCHRX,PATCHRY,PATCHRZ,N,N,0,LADO0,T,
     &           ZETA,RTHR,MINPS,FRACR,BOR,
     &           BORAMR,BOVLP,NPART_ESP,0,MPAPOLEV)
       NPT=SUM(NPATCH(0:NLEVELS))
       OPEN(11,FILE=FOUT,ACCESS='STREAM',FORM='UNFORMATTED',
     &      STATUS='REPLACE',ACTION='WRITE')
       WRITE(11) NLM
       WRITE(11) NPATCH
       WRITE(11) (PATCHNX(I),I=1,NPT)
       WRITE(11) (PATCHNY(I),I=1,NPT)
       WRITE(11) (PATCHNZ(I),I=1,NPT)
       WRITE(11) (PATCHX(I),I=1,NPT)
       WRITE(11) (PATCHY(I),I=1,NPT)
       WRITE(11) (PATCHZ(I),I=1,NPT)
       WRITE(11) (PARE(I),I=1,NPT)
       WRITE(11) (PATCHRX(I),I=1,NPT)
       WRITE(11) (PATCHRY(I),I=1,NPT)
       WRITE(11) (PATCHRZ(I),I=1,NPT)
       CLOSE(11)
       STOP
      END IF
      IF (MODE.EQ.3) THEN
       ALLOCATE(U2DM(PARTI),U3DM(PARTI),U4DM(PARTI),ORIPA(PARTI))
       DO I=1,N
        U2DM(I)=0.0
        U3DM(I)=0.0
        U4DM(I)=0.0
        ORIPA(I)=I
       END DO
       CALL SORT_DM_PARTICLES_LOCALDENSITY(N,NPART_ESP,0,0,
     &                                     SCAL(2),SCAL(3))
       OPEN(11,FILE=FOUT,ACCESS='STREAM',FORM='UNFORMATTED',
     &      STATUS='REPLACE',ACTION='WRITE')
       WRITE(11) DX,DY,DZ
       WRITE(11) NPART_ESP
       WRITE(11) (ORIPA(I),I=1,N)
       CLOSE(11)
       STOP
      END IF
      IF (MODE.EQ.1) THEN
       CALL DENSITY(0,NX,NY,NZ,NL,NPATCH,PARE,PATCHNX,PATCHNY,
     &              PATCHNZ,PATCHX,PATCHY,PATCHZ,PATCHRX,PATCHRY,
     &              PATCHRZ,N,N,0,LADO0,T,ZETA,NPART_ESP,IDEG)
      ELSE
       CALL INTERPOLATE_DENSITY_TSC(NX,NY,NZ,NL,NPATCH,
     &           PARE,PATCHNX,PATCHNY,PATCHNZ,PATCHX,PATCHY,PATCHZ,
     &           PATCHRX,PATCHRY,PATCHRZ,N,N,LADO0,U1,U11)
      END IF
      CALL RENORM_DENSITY(NL,NX,NY,NZ,NPATCH,PATCHNX,PATCHNY,
     &                    PATCHNZ,U1,U11,LADO0,0.0,0.0,FDM)

      IF (MODE.EQ.5) THEN
       DO IR=1,NL
        CALL VEINSGRID(IR,NL,NPATCH,PARE,PATCHNX,PATCHNY,PATCHNZ,
     &                 PATCHX,PATCHY,PATCHZ,PATCHRX,PATCHRY,PATCHRZ,
     &                 SOLAP)
       END DO
       CALL COMPUTE_CR0AMR(NL,NX,NY,NZ,NPATCH,PARE,PATCHNX,PATCHNY,
     &                     PATCHNZ,PATCHX,PATCHY,PATCHZ,PATCHRX,
     &                     PATCHRY,PATCHRZ,CR0AMR,CR0AMR11,LADO0)
       OPEN(99,FILE=FOUT,STATUS='UNKNOWN',FORM='UNFORMATTED')
       WRITE(99) (((U1(IX,JY,KZ),IX=1,NX),JY=1,NY),KZ=1,NZ)
       WRITE(99) (((CR0AMR(IX,JY,KZ),IX=1,NX),JY=1,NY),KZ=1,NZ)
       DO IP=1,NPT
        N1=PATCHNX(IP)
        N2=PATCHNY(IP)
        N3=PATCHNZ(IP)
        WRITE(99) (((U11(IX,JY,KZ,IP),IX=1,N1),JY=1,N2),KZ=1,N3)
        WRITE(99) (((CR0AMR11(IX,JY,KZ,IP),IX=1,N1),JY=1,N2),KZ=1,N3)
        WRITE(99) (((SOLAP(IX,JY,KZ,IP),IX=1,N1),JY=1,N2),KZ=1,N3)
       END DO
       CLOSE(99)
       STOP
      END IF

      OPEN(11,FILE=FOUT,ACCESS='STREAM',FORM='UNFORMATTED',
     &     STATUS='REPLACE',ACTION='WRITE')
      WRITE(11) DX,DY,DZ
      WRITE(11) (RADX(I),I=0,NX+1)
      WRITE(11) (((U1(IX,JY,KZ),IX=1,NX),JY=1,NY),KZ=1,NZ)
      DO IP=1,NPT
       N1=PATCHNX(IP)
       N2=PATCHNY(IP)
       N3=PATCHNZ(IP)
       WRITE(11) (((U11(IX,JY,KZ,IP),IX=1,N1),JY=1,N2),KZ=1,N3)
      END DO
      CLOSE(11)
      END
