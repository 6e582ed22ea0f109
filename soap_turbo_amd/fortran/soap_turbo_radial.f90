! The public flag of the reference's soap_turbo_radial module
! (src/soap_turbo_radial.f90:67): which filter recursion seeds the radial
! expansion. .true. (default) reproduces the reference's legacy seed; .false.
! the finite-difference-consistent one. get_soap reads it on every call.
!
! extended_precision_w (no reference counterpart): .false. (default) builds
! W = S^{-1/2} by the reference's own LAPACK sequence (dgesvd, dgemm,
! dpotrf/dpotri on the host's LAPACK), so W is the reference's; .true. uses the
! library's extended-precision route instead, closer to the exact S^{-1/2}
! where S is ill-conditioned (poly3 at alpha_max 8).
module soap_turbo_radial
  implicit none
  logical, public :: legacy_filter_seed = .true.
  logical, public :: extended_precision_w = .false.
end module soap_turbo_radial
