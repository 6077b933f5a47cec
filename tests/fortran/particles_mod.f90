! Test harness: the particle store the density build reads, with the names
! and types of ASOHF's MODULE PARTICLES (FP32 SoA positions and masses).
module particles
  implicit none
  save
  integer :: parti, parti_pph
  real, allocatable :: rxpa(:), rypa(:), rzpa(:)
  real, allocatable :: u2dm(:), u3dm(:), u4dm(:)
  real, allocatable :: masap(:)
  integer, allocatable :: oripa(:), particles_per_halo(:)
end module particles
