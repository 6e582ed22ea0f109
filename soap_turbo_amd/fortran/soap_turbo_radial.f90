! The public flag of the reference's soap_turbo_radial module
! (src/soap_turbo_radial.f90:67): which filter recursion seeds the radial
! expansion. .true. (default) reproduces the reference's legacy seed; .false.
! the finite-difference-consistent one. get_soap reads it on every call.
module soap_turbo_radial
  implicit none
  logical, public :: legacy_filter_seed = .true.
end module soap_turbo_radial
