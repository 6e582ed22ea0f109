! Flat C-callable wrappers of the Fortran drop-in (get_soap, soap_contract,
! get_soap_adjoint_forces, get_compress_indices, get_soap_central_hessian), so that tests can drive the Fortran layer itself
! from a C-ABI caller. Arrays arrive as plain pointers with explicit sizes and
! are handed to the drop-in routines as Fortran arrays.
module soap_turbo_flat
  use, intrinsic :: iso_c_binding
  use soap_turbo_desc
  use soap_turbo_radial, only: legacy_filter_seed, extended_precision_w
  use soap_turbo_compress_module
  implicit none

contains

  subroutine st_f90_get_soap(n_sites, n_species, n_pairs, max_mult, n_neigh, species, species_mult, mask, rjs, &
                             thetas, phis, alpha_max, l_max, rcut_hard, rcut_soft, nf, global_scaling, &
                             atom_sigma_r, atom_sigma_r_scaling, atom_sigma_t, atom_sigma_t_scaling, &
                             amplitude_scaling, radial_enhancement, central_weight, basis_code, scaling_code, &
                             do_derivatives, compress, P_nonzero, P_i, P_j, P_el, n_soap, soap, soap_cart_der, &
                             contract, legacy_seed) bind(c, name="st_f90_get_soap")
    integer(c_int32_t), value :: n_sites, n_species, n_pairs, max_mult, l_max, radial_enhancement, basis_code, &
                                 scaling_code, do_derivatives, compress, P_nonzero, n_soap, contract, legacy_seed
    integer(c_int32_t), intent(in) :: n_neigh(n_sites), species(max_mult, n_sites), species_mult(n_sites), &
                                      mask(n_pairs, n_species), alpha_max(n_species), P_i(*), P_j(*)
    real(c_double), intent(in) :: rjs(n_pairs), thetas(n_pairs), phis(n_pairs), rcut_hard(n_species), &
                                  rcut_soft(n_species), nf(n_species), global_scaling(n_species), &
                                  atom_sigma_r(n_species), atom_sigma_r_scaling(n_species), atom_sigma_t(n_species), &
                                  atom_sigma_t_scaling(n_species), amplitude_scaling(n_species), &
                                  central_weight(n_species), P_el(*)
    real(c_double), intent(inout) :: soap(n_soap, n_sites), soap_cart_der(3, n_soap, *)
    logical, allocatable :: lmask(:,:)
    character(len=16) :: basis, scal
    integer :: nz, ncart
    integer, allocatable :: pi(:), pj(:)
    real(c_double), allocatable :: pel(:)
    allocate( lmask(n_pairs, n_species) )
    lmask = mask /= 0
    basis = "poly3"
    if( basis_code == 1 ) basis = "poly3gauss"
    if( basis_code == 2 ) basis = "poly3operator"
    scal = "polynomial"
    if( scaling_code /= 0 ) scal = "none"
    legacy_filter_seed = legacy_seed /= 0
    nz = max(int(P_nonzero), 1)
    allocate( pi(nz), pj(nz), pel(nz) )
    pi = 1; pj = 1; pel = 0.d0
    if( compress /= 0 )then
      pi = P_i(1:nz); pj = P_j(1:nz); pel = P_el(1:nz)
    end if
    ncart = 1
    if( do_derivatives /= 0 .and. contract == 0 ) ncart = n_pairs
    call get_soap(int(n_sites), n_neigh, int(n_species), species, species_mult, int(n_pairs), lmask, rjs, thetas, &
                  phis, alpha_max, int(l_max), rcut_hard, rcut_soft, nf, global_scaling, atom_sigma_r, &
                  atom_sigma_r_scaling, atom_sigma_t, atom_sigma_t_scaling, amplitude_scaling, &
                  int(radial_enhancement), central_weight, basis, scal, .false., do_derivatives /= 0, &
                  compress /= 0, int(P_nonzero), pi, pj, pel, soap, soap_cart_der(:, :, 1:ncart), contract /= 0)
  end subroutine

  subroutine st_f90_get_soap_adjoint_forces(n_sites, n_species, n_pairs, max_mult, n_neigh, species, species_mult, &
                                            mask, rjs, thetas, phis, alpha_max, l_max, rcut_hard, rcut_soft, nf, &
                                            global_scaling, atom_sigma_r, atom_sigma_r_scaling, atom_sigma_t, &
                                            atom_sigma_t_scaling, amplitude_scaling, radial_enhancement, &
                                            central_weight, basis_code, scaling_code, compress, P_nonzero, P_i, P_j, &
                                            P_el, n_soap, soap, G, sgn, neighbors_list, n_atoms, forces, virial, &
                                            with_xyz, xyz, legacy_seed) bind(c, name="st_f90_get_soap_adjoint_forces")
    integer(c_int32_t), value :: n_sites, n_species, n_pairs, max_mult, l_max, radial_enhancement, basis_code, &
                                 scaling_code, compress, P_nonzero, n_soap, n_atoms, with_xyz, legacy_seed
    integer(c_int32_t), intent(in) :: n_neigh(n_sites), species(max_mult, n_sites), species_mult(n_sites), &
                                      mask(n_pairs, n_species), alpha_max(n_species), P_i(*), P_j(*), &
                                      neighbors_list(n_pairs)
    real(c_double), intent(in) :: rjs(n_pairs), thetas(n_pairs), phis(n_pairs), rcut_hard(n_species), &
                                  rcut_soft(n_species), nf(n_species), global_scaling(n_species), &
                                  atom_sigma_r(n_species), atom_sigma_r_scaling(n_species), atom_sigma_t(n_species), &
                                  atom_sigma_t_scaling(n_species), amplitude_scaling(n_species), &
                                  central_weight(n_species), P_el(*), G(n_sites, n_soap), xyz(3, *)
    real(c_double), value :: sgn
    real(c_double), intent(out) :: soap(n_soap, n_sites), forces(3, n_atoms), virial(3, 3)
    logical, allocatable :: lmask(:,:)
    character(len=16) :: basis, scal
    integer :: nz
    integer, allocatable :: pi(:), pj(:), nbr(:)
    real(c_double), allocatable :: pel(:)
    allocate( lmask(n_pairs, n_species) )
    lmask = mask /= 0
    basis = "poly3"
    if( basis_code == 1 ) basis = "poly3gauss"
    if( basis_code == 2 ) basis = "poly3operator"
    scal = "polynomial"
    if( scaling_code /= 0 ) scal = "none"
    legacy_filter_seed = legacy_seed /= 0
    nz = max(int(P_nonzero), 1)
    allocate( pi(nz), pj(nz), pel(nz) )
    pi = 1; pj = 1; pel = 0.d0
    if( compress /= 0 )then
      pi = P_i(1:nz); pj = P_j(1:nz); pel = P_el(1:nz)
    end if
    nbr = neighbors_list
    if( with_xyz /= 0 )then
      call get_soap_adjoint_forces(int(n_sites), n_neigh, int(n_species), species, species_mult, int(n_pairs), &
                                   lmask, rjs, thetas, phis, alpha_max, int(l_max), rcut_hard, rcut_soft, nf, &
                                   global_scaling, atom_sigma_r, atom_sigma_r_scaling, atom_sigma_t, &
                                   atom_sigma_t_scaling, amplitude_scaling, int(radial_enhancement), &
                                   central_weight, basis, scal, compress /= 0, int(P_nonzero), pi, pj, pel, soap, &
                                   G, sgn, nbr, int(n_atoms), forces, virial, xyz(:, 1:n_pairs))
    else
      call get_soap_adjoint_forces(int(n_sites), n_neigh, int(n_species), species, species_mult, int(n_pairs), &
                                   lmask, rjs, thetas, phis, alpha_max, int(l_max), rcut_hard, rcut_soft, nf, &
                                   global_scaling, atom_sigma_r, atom_sigma_r_scaling, atom_sigma_t, &
                                   atom_sigma_t_scaling, amplitude_scaling, int(radial_enhancement), &
                                   central_weight, basis, scal, compress /= 0, int(P_nonzero), pi, pj, pel, soap, &
                                   G, sgn, nbr, int(n_atoms), forces, virial)
    end if
  end subroutine

  subroutine st_f90_soap_contract(n_sites, n_soap, n_pairs, soap, G, sgn, pair) bind(c, name="st_f90_soap_contract")
    integer(c_int32_t), value :: n_sites, n_soap, n_pairs
    real(c_double), intent(in) :: soap(n_soap, n_sites), G(n_sites, n_soap)
    real(c_double), value :: sgn
    real(c_double), intent(out) :: pair(3, n_pairs)
    call soap_contract(soap, G, sgn, pair)
  end subroutine

  subroutine st_f90_compress_indices(n_species, alpha_max, l_max, dim, P_nonzero, P_i, P_j, P_el, set) &
      bind(c, name="st_f90_compress_indices")
    integer(c_int32_t), value :: n_species, l_max, set
    integer(c_int32_t), intent(in) :: alpha_max(n_species)
    integer(c_int32_t), intent(inout) :: dim, P_nonzero, P_i(*), P_j(*)
    real(c_double), intent(inout) :: P_el(*)
    integer :: d, nz
    integer, allocatable :: pi(:), pj(:)
    real(c_double), allocatable :: pel(:)
    if( set == 0 )then
      allocate( pi(1), pj(1), pel(1) )
      call get_compress_indices("trivial", alpha_max, int(l_max), d, nz, pi, pj, pel, "get_dim")
    else
      nz = int(P_nonzero)
      allocate( pi(nz), pj(nz), pel(nz) )
      call get_compress_indices("trivial", alpha_max, int(l_max), d, nz, pi, pj, pel, "set_indices")
      P_i(1:nz) = pi; P_j(1:nz) = pj; P_el(1:nz) = pel
    end if
    dim = d
    P_nonzero = nz
  end subroutine

  subroutine st_f90_set_extended_w(on) bind(c, name="st_f90_set_extended_w")
    integer(c_int32_t), value :: on
    extended_precision_w = on /= 0
  end subroutine

  subroutine st_f90_set_hessian(on) bind(c, name="st_f90_set_hessian")
    integer(c_int32_t), value :: on
    soap_hessian_enabled = on /= 0
  end subroutine

  subroutine st_f90_set_hessian_selftest(on) bind(c, name="st_f90_set_hessian_selftest")
    integer(c_int32_t), value :: on
    soap_hessian_selftest = on /= 0
  end subroutine

  function st_f90_hessian_selftest_err() bind(c, name="st_f90_hessian_selftest_err")
    real(c_double) :: st_f90_hessian_selftest_err
    st_f90_hessian_selftest_err = soap_hessian_selftest_err
  end function

  subroutine st_f90_central_hessian(n_sites, n_species, n_pairs, n_neigh, mask, rjs, thetas, phis, l_max, rcut_hard, &
                                    atom_sigma_t, atom_sigma_t_scaling, compress, P_nonzero, P_i, P_j, P_el, n_soap, &
                                    soap, n_vec, vecs, hess) bind(c, name="st_f90_central_hessian")
    integer(c_int32_t), value :: n_sites, n_species, n_pairs, l_max, compress, P_nonzero, n_soap, n_vec
    integer(c_int32_t), intent(in) :: n_neigh(n_sites), mask(n_pairs, n_species), P_i(*), P_j(*)
    real(c_double), intent(in) :: rjs(n_pairs), thetas(n_pairs), phis(n_pairs), rcut_hard(n_species), &
                                  atom_sigma_t(n_species), atom_sigma_t_scaling(n_species), P_el(*)
    real(c_double), intent(in) :: soap(n_soap, n_sites), vecs(n_soap, n_vec, n_sites)
    real(c_double), intent(out) :: hess(3, 3, n_vec, n_pairs)
    logical, allocatable :: lmask(:,:)
    integer :: nz
    nz = max(int(P_nonzero), 1)
    allocate( lmask(n_pairs, n_species) )
    lmask = mask /= 0
    call get_soap_central_hessian(int(n_sites), n_neigh, int(n_pairs), int(n_species), lmask, rjs, thetas, phis, &
                                  int(l_max), rcut_hard, atom_sigma_t, atom_sigma_t_scaling, compress /= 0, &
                                  int(P_nonzero), P_i(1:nz), P_j(1:nz), P_el(1:nz), soap, int(n_vec), vecs, hess)
  end subroutine

end module soap_turbo_flat
