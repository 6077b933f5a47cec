************************************************************************
*     drive_density.f -- test harness for the Fortran drop-in DENSITY.
*     Fills MODULE PARTICLES and the COMMON blocks the way asohf.f does
*     before its CALL DENSITY (asohf.f:467-470), pre-sets U1/U11 to -1 as
*     INIT_GRIDVARS does (alloc.f:162-186), calls DENSITY and writes the
*     fields plus the number of padding cells that changed.
*     Usage: drive_density <input> <output>
*     Input (stream): int32 hdr(8) = nx, ny, nz, nl, n, interp_degree, P, 0
*       int32 npatch(0:nl), int32 npart_esp(0:8)
*       real*4 lado0, rote, rete, dx, dy, dz
*       int32 pnx(P), pny(P), pnz(P); real*4 prx(P), pry(P), prz(P)
*       real*4 x(n), y(n), z(n), m(n)
*     Output: real*4 u1(1:nx,1:ny,1:nz), u11 interiors patch by patch,
*       int32 number of padding cells that are no longer -1
************************************************************************
      PROGRAM DRIVE_DENSITY
      USE PARTICLES
      IMPLICIT NONE
      INCLUDE 'input_files/asohf_parameters.dat'
      REAL*4 U1(NMAX,NMAY,NMAZ)
      REAL*4 U11(NAMRX,NAMRY,NAMRZ,NPALEV)
      COMMON /VARIA/ U1,U11
      REAL DX,DY,DZ
      COMMON /ESPACIADO/ DX,DY,DZ
      REAL*4 RETE,HTE,ROTE
      COMMON /BACK/ RETE,HTE,ROTE
      INTEGER HDR(8),NX,NY,NZ,NL,N,IDEG,NPT,I,IP,N1,N2,N3,IX,JY,KZ
      INTEGER NPATCH(0:NLEVELS),PARE(NPALEV)
      INTEGER PATCHNX(NPALEV),PATCHNY(NPALEV),PATCHNZ(NPALEV)
      INTEGER PATCHX(NPALEV),PATCHY(NPALEV),PATCHZ(NPALEV)
      REAL*4 PATCHRX(NPALEV),PATCHRY(NPALEV),PATCHRZ(NPALEV)
      INTEGER NPART_ESP(0:N_ESP-1),CHANGED
      REAL*4 SCAL(6),LADO0,T,ZETA
      CHARACTER*512 FIN,FOUT

      CALL GET_COMMAND_ARGUMENT(1,FIN)
      CALL GET_COMMAND_ARGUMENT(2,FOUT)
      OPEN(10,FILE=FIN,ACCESS='STREAM',FORM='UNFORMATTED',
     &     STATUS='OLD',ACTION='READ')
      READ(10) HDR
      NX=HDR(1)
      NY=HDR(2)
      NZ=HDR(3)
      NL=HDR(4)
      N=HDR(5)
      IDEG=HDR(6)
      NPT=HDR(7)
      IF (NX.GT.NMAX.OR.NY.GT.NMAY.OR.NZ.GT.NMAZ.OR.NPT.GT.NPALEV)
     &   STOP 2
      NPATCH=0
      READ(10) (NPATCH(I),I=0,NL)
      READ(10) NPART_ESP
      READ(10) SCAL
      LADO0=SCAL(1)
      ROTE=SCAL(2)
      RETE=SCAL(3)
      DX=SCAL(4)
      DY=SCAL(5)
      DZ=SCAL(6)
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
      PARTI=MAX(N,1)
      ALLOCATE(RXPA(PARTI),RYPA(PARTI),RZPA(PARTI),MASAP(PARTI))
      IF (N.GT.0) THEN
       READ(10) (RXPA(I),I=1,N)
       READ(10) (RYPA(I),I=1,N)
       READ(10) (RZPA(I),I=1,N)
       READ(10) (MASAP(I),I=1,N)
      END IF
      CLOSE(10)

      U1=-1.0
      U11=-1.0
      T=0.0
      ZETA=0.0
      CALL DENSITY(0,NX,NY,NZ,NL,NPATCH,PARE,PATCHNX,PATCHNY,
     &             PATCHNZ,PATCHX,PATCHY,PATCHZ,PATCHRX,PATCHRY,
     &             PATCHRZ,N,N,0,LADO0,T,ZETA,NPART_ESP,IDEG)

      CHANGED=0
      DO KZ=1,NMAZ
      DO JY=1,NMAY
      DO IX=1,NMAX
       IF (IX.GT.NX.OR.JY.GT.NY.OR.KZ.GT.NZ) THEN
        IF (U1(IX,JY,KZ).NE.-1.0) CHANGED=CHANGED+1
       END IF
      END DO
      END DO
      END DO
      DO IP=1,NPALEV
       DO KZ=1,NAMRZ
       DO JY=1,NAMRY
       DO IX=1,NAMRX
        IF (IP.GT.NPT.OR.IX.GT.PATCHNX(IP).OR.JY.GT.PATCHNY(IP).OR.
     &      KZ.GT.PATCHNZ(IP)) THEN
         IF (U11(IX,JY,KZ,IP).NE.-1.0) CHANGED=CHANGED+1
        END IF
       END DO
       END DO
       END DO
      END DO

      OPEN(11,FILE=FOUT,ACCESS='STREAM',FORM='UNFORMATTED',
     &     STATUS='REPLACE',ACTION='WRITE')
      WRITE(11) (((U1(IX,JY,KZ),IX=1,NX),JY=1,NY),KZ=1,NZ)
      DO IP=1,NPT
       N1=PATCHNX(IP)
       N2=PATCHNY(IP)
       N3=PATCHNZ(IP)
       WRITE(11) (((U11(IX,JY,KZ,IP),IX=1,N1),JY=1,N2),KZ=1,N3)
      END DO
      WRITE(11) CHANGED
      CLOSE(11)
      END
