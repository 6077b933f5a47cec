************************************************************************
*     density_hip.f -- drop-in SUBROUTINE DENSITY for asohf.f backed by
*     libasohf_hip (MI355X).  Same name and dummy arguments as the
*     reference routine (grids.f:710-713), same implicit inputs
*     (MODULE PARTICLES, COMMON /ESPACIADO/, /BACK/) and outputs
*     (COMMON /VARIA/ U1, U11 interiors; padding untouched), same
*     per-level min/max report (grids.f:835, 987).  On any library error
*     it prints the message and STOPs 1, as the reference does on bad
*     input.  FDM is not applied here: RENORM_DENSITY (grids.f:2668)
*     still divides by it afterwards, as in the reference pipeline.
*
*     Build it in the ASOHF tree (it INCLUDEs input_files/
*     asohf_parameters.dat and USEs PARTICLES) and link libasohf_hip.
************************************************************************
      SUBROUTINE DENSITY(ITER,NX,NY,NZ,NL,NPATCH,PARE,PATCHNX,PATCHNY,
     &              PATCHNZ,PATCHX,PATCHY,PATCHZ,PATCHRX,PATCHRY,
     &              PATCHRZ,N_PARTICLES,N_DM,N_GAS,LADO0,T,ZETA,
     &              NPART_ESP,INTERP_DEGREE)
      USE PARTICLES
      USE ISO_C_BINDING
      IMPLICIT NONE
      INCLUDE 'input_files/asohf_parameters.dat'

      INTEGER ITER,NX,NY,NZ,NL,N_PARTICLES,N_DM,N_GAS
      INTEGER, TARGET :: NPATCH(0:NLEVELS)
      INTEGER PARE(NPALEV)
      INTEGER, TARGET :: PATCHNX(NPALEV),PATCHNY(NPALEV),
     &                   PATCHNZ(NPALEV)
      INTEGER PATCHX(NPALEV),PATCHY(NPALEV),PATCHZ(NPALEV)
      REAL, TARGET :: PATCHRX(NPALEV),PATCHRY(NPALEV),PATCHRZ(NPALEV)
      REAL LADO0,T,ZETA
      INTEGER, TARGET :: NPART_ESP(0:N_ESP-1)
      INTEGER INTERP_DEGREE

      REAL*4 U1(NMAX,NMAY,NMAZ)
      REAL*4 U11(NAMRX,NAMRY,NAMRZ,NPALEV)
      COMMON /VARIA/ U1,U11
      REAL DX,DY,DZ
      COMMON /ESPACIADO/ DX,DY,DZ
      REAL*4 RETE,HTE,ROTE
      COMMON /BACK/ RETE,HTE,ROTE

*     C ABI of include/asohf_hip.h
      TYPE, BIND(C) :: ASOHF_GEOM
       INTEGER(C_INT32_T) :: NX,NY,NZ,NL
       TYPE(C_PTR) :: NPATCH,PNX,PNY,PNZ,PRX,PRY,PRZ
       REAL(C_FLOAT) :: DX,DY,DZ,LADO0,ROTE,RETE,FDM
       INTEGER(C_INT32_T) :: INTERP_DEGREE
       TYPE(C_PTR) :: RADX,RADY,RADZ
      END TYPE
      TYPE, BIND(C) :: ASOHF_OPTS
       INTEGER(C_INT32_T) :: MODE,REDUCTION,PATH,USE_GRAPH
       INTEGER(C_INT32_T) :: U1_LD(2),U11_LD(3)
       TYPE(C_PTR) :: STREAM
      END TYPE
      INTERFACE
       FUNCTION ASOHF_HIP_CREATE(CTX,DEVICE)
     &          BIND(C,NAME='asohf_hip_create')
        IMPORT :: C_PTR,C_INT32_T
        TYPE(C_PTR) :: CTX
        INTEGER(C_INT32_T), VALUE :: DEVICE
        INTEGER(C_INT32_T) :: ASOHF_HIP_CREATE
       END FUNCTION
       FUNCTION ASOHF_HIP_DENSITY_HOST(CTX,GEOM,N,RX,RY,RZ,M,ESP,
     &          V1,V11,GMASS,OPTS) BIND(C,NAME='asohf_hip_density_host')
        IMPORT :: C_PTR,C_INT32_T,C_INT64_T,C_FLOAT,C_DOUBLE,
     &            ASOHF_GEOM,ASOHF_OPTS
        TYPE(C_PTR), VALUE :: CTX
        TYPE(ASOHF_GEOM) :: GEOM
        INTEGER(C_INT64_T), VALUE :: N
        REAL(C_FLOAT) :: RX(*),RY(*),RZ(*),M(*)
        INTEGER(C_INT32_T) :: ESP(*)
        REAL(C_FLOAT) :: V1(*),V11(*)
        REAL(C_DOUBLE) :: GMASS(*)
        TYPE(ASOHF_OPTS) :: OPTS
        INTEGER(C_INT32_T) :: ASOHF_HIP_DENSITY_HOST
       END FUNCTION
       FUNCTION ASOHF_HIP_LAST_ERROR(CTX)
     &          BIND(C,NAME='asohf_hip_last_error')
        IMPORT :: C_PTR
        TYPE(C_PTR), VALUE :: CTX
        TYPE(C_PTR) :: ASOHF_HIP_LAST_ERROR
       END FUNCTION
      END INTERFACE

      TYPE(C_PTR), SAVE :: CTX = C_NULL_PTR
      TYPE(ASOHF_GEOM) :: G
      TYPE(ASOHF_OPTS) :: O
      INTEGER(C_INT32_T) :: ST
      INTEGER(C_INT64_T) :: NPART
      REAL(C_DOUBLE), ALLOCATABLE :: GMASS(:)
      CHARACTER(KIND=C_CHAR), POINTER :: MSG(:)
      INTEGER IR,LOW1,LOW2,I,N1,N2,N3,L
      REAL BAS,BAS2

      WRITE(*,*) '==== Density interpolation (libasohf_hip)'
      IF (.NOT.C_ASSOCIATED(CTX)) THEN
       ST=ASOHF_HIP_CREATE(CTX,-1)
       IF (ST.NE.0) THEN
        WRITE(*,*) 'DENSITY: no HIP device for libasohf_hip, status',ST
        STOP 1
       END IF
      END IF

      G%NX=NX
      G%NY=NY
      G%NZ=NZ
      G%NL=NL
      G%NPATCH=C_LOC(NPATCH(0))
      G%PNX=C_LOC(PATCHNX(1))
      G%PNY=C_LOC(PATCHNY(1))
      G%PNZ=C_LOC(PATCHNZ(1))
      G%PRX=C_LOC(PATCHRX(1))
      G%PRY=C_LOC(PATCHRY(1))
      G%PRZ=C_LOC(PATCHRZ(1))
      G%DX=DX
      G%DY=DY
      G%DZ=DZ
      G%LADO0=LADO0
      G%ROTE=ROTE
      G%RETE=RETE
      G%FDM=1.0
      G%INTERP_DEGREE=INTERP_DEGREE
      G%RADX=C_NULL_PTR
      G%RADY=C_NULL_PTR
      G%RADZ=C_NULL_PTR
      O%MODE=1
      O%REDUCTION=0
      O%PATH=0
      O%USE_GRAPH=0
      O%U1_LD(1)=NMAX
      O%U1_LD(2)=NMAY
      O%U11_LD(1)=NAMRX
      O%U11_LD(2)=NAMRY
      O%U11_LD(3)=NAMRZ
      O%STREAM=C_NULL_PTR
      NPART=SUM(NPART_ESP)
      ALLOCATE(GMASS(1+SUM(NPATCH(0:NL))))
      ST=ASOHF_HIP_DENSITY_HOST(CTX,G,NPART,RXPA,RYPA,RZPA,MASAP,
     &                          NPART_ESP,U1,U11,GMASS,O)
      IF (ST.NE.0) THEN
       CALL C_F_POINTER(ASOHF_HIP_LAST_ERROR(CTX),MSG,[512])
       L=0
       DO I=1,512
        IF (MSG(I).EQ.C_NULL_CHAR) EXIT
        L=I
       END DO
       WRITE(*,*) 'DENSITY (libasohf_hip) failed, status',ST,': ',
     &            (MSG(I),I=1,L)
       STOP 1
      END IF
      DEALLOCATE(GMASS)

      WRITE(*,*) 'At level',0,MINVAL(U1(1:NX,1:NY,1:NZ)),
     &           MAXVAL(U1(1:NX,1:NY,1:NZ))
      DO IR=1,NL
       LOW1=SUM(NPATCH(0:IR-1))+1
       LOW2=SUM(NPATCH(0:IR))
       BAS=1000000.0
       BAS2=-1000000.0
       DO I=LOW1,LOW2
        N1=PATCHNX(I)
        N2=PATCHNY(I)
        N3=PATCHNZ(I)
        BAS=MIN(BAS,MINVAL(U11(1:N1,1:N2,1:N3,I)))
        BAS2=MAX(BAS2,MAXVAL(U11(1:N1,1:N2,1:N3,I)))
       END DO
       WRITE(*,*) 'At level',IR,BAS,BAS2
      END DO
      WRITE(*,*) '==== End density interpolation'

      RETURN
      END
