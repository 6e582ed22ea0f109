! Drop-in module soap_turbo_desc: get_soap and soap_contract with the
! reference's argument lists (src/soap_turbo.f90:139-169, :1723-1730), computed
! by the HIP library on one GPU through its C-ABI.
!
! Caller contract, as for the reference: soap and soap_cart_der are zero on
! entry; get_soap(..., contract=.true.) must precede soap_contract. The
! library keeps the state between them (not reentrant). Errors on which the
! reference stops write the reference's message and stop.
!
! The GPU is device 0 unless SOAP_TURBO_DEVICE names another.
module soap_turbo_desc
  use, intrinsic :: iso_fortran_env, only: real64
  use, intrinsic :: iso_c_binding
  use soap_turbo_c_api
  use soap_turbo_radial, only: legacy_filter_seed
  implicit none
  private
  integer, parameter :: dp = real64

  public :: get_soap, soap_contract, get_soap_central_hessian, assign_species_multiplicity
  ! the reference's public Hessian flags (src/soap_turbo.f90:76-88)
  logical, public, save :: soap_hessian_enabled = .false.
  logical, public, save :: soap_hessian_selftest = .false.
  real(dp), public, save :: soap_hessian_selftest_err = 0.d0

  type(c_ptr), save :: ctx = c_null_ptr
  integer, save :: ct_n_sites = 0, ct_n_pairs = 0, ct_n_soap = 0
  ! what get_soap_central_hessian needs from the get_soap before it (that call's parameters; the
  ! neighbour list stays resident in the library): the reference's hs_* state, src/soap_turbo.f90:948-970
  logical, save :: hs_ready = .false.
  integer, save :: hs_n_sites = 0, hs_n_pairs = 0, hs_l_max = -1, hs_n_soap = 0
  type(st_params), target, save :: hs_prm
  type(st_compression), target, save :: hs_cmp
  integer(c_int32_t), allocatable, target, save :: hs_am(:), hs_pi(:), hs_pj(:)
  real(c_double), allocatable, target, save :: hs_rch(:), hs_rcs(:), hs_nf(:), hs_gs(:), hs_sr(:), hs_srs(:), &
                                               hs_st(:), hs_sts(:), hs_as(:), hs_cw(:), hs_pel(:)

contains

  subroutine ensure_context()
    character(len=32) :: env
    integer :: n, dev, rc, ios
    if( c_associated(ctx) ) return
    dev = 0
    call get_environment_variable("SOAP_TURBO_DEVICE", env, n)
    if( n > 0 )then
      read(env, *, iostat=ios) dev
      if( ios /= 0 ) dev = 0
    end if
    rc = st_create(ctx, int(dev, c_int))
    if( rc /= ST_OK )then
      write(*,*) "ERROR: soap_turbo: no usable GPU (device ", dev, ")"
      stop
    end if
  end subroutine

  subroutine check(rc)
    integer(c_int), intent(in) :: rc
    if( rc /= ST_OK )then
      write(*,*) st_message(ctx)
      stop
    end if
  end subroutine

  subroutine get_soap(n_sites, n_neigh, n_species, species, species_multiplicity, n_atom_pairs, mask, rjs, &
                      thetas, phis, alpha_max, l_max, rcut_hard, rcut_soft, nf, global_scaling, atom_sigma_r, &
                      atom_sigma_r_scaling, atom_sigma_t, atom_sigma_t_scaling, &
                      amplitude_scaling, radial_enhancement, central_weight, basis, scaling_mode, do_timing, &
                      do_derivatives, compress_soap, compress_P_nonzero, compress_P_i, compress_P_j, &
                      compress_P_el, soap, soap_cart_der, contract)
    implicit none
    real(dp), intent(in) :: rjs(:), thetas(:), phis(:)
    real(dp), intent(in) :: amplitude_scaling(:), atom_sigma_r_scaling(:), atom_sigma_t(:), atom_sigma_t_scaling(:)
    real(dp), intent(in) :: central_weight(:), atom_sigma_r(:), global_scaling(:)
    real(dp), intent(in) :: nf(:), rcut_hard(:), rcut_soft(:)
    real(dp), intent(in) :: compress_P_el(:)
    integer, intent(in) :: n_species, radial_enhancement, species(:,:), species_multiplicity(:)
    integer, intent(in) :: n_sites, n_neigh(:), l_max, n_atom_pairs, alpha_max(:), compress_P_nonzero, &
                           compress_P_i(:), compress_P_j(:)
    logical, intent(in) :: do_derivatives, do_timing, mask(:,:), compress_soap
    character(*), intent(in) :: basis, scaling_mode
    logical, intent(in), optional :: contract
    real(dp), intent(inout), target, contiguous :: soap(:,:), soap_cart_der(:,:,:)

    ! contiguous C-interoperable copies of the inputs
    integer(c_int32_t), allocatable, target :: c_am(:), c_nn(:), c_sp(:,:), c_sm(:), c_mask(:,:), c_pi(:), c_pj(:)
    real(c_double), allocatable, target :: c_r(:), c_t(:), c_f(:), c_pel(:)
    real(c_double), allocatable, target :: c_rch(:), c_rcs(:), c_nf(:), c_gs(:), c_sr(:), c_srs(:), c_st(:), &
                                           c_sts(:), c_as(:), c_cw(:)
    type(st_params), target :: prm
    type(st_neighbours), target :: nb
    type(st_compression), target :: cmp
    integer(c_int32_t) :: ibasis, iscal, do_ctr, n_soap
    integer(c_int) :: rc
    integer(8) :: t0, t1, rate
    character(len=:), allocatable :: b

    if( do_timing ) call system_clock(t0, rate)
    if( soap_hessian_enabled .and. do_derivatives .and. legacy_filter_seed )then
      ! the reference refuses second radial derivatives with the legacy seed (src/soap_turbo_radial.f90:318)
      write(*,*) "ERROR: second radial derivatives need soap_turbo_radial's legacy_filter_seed = .false. ."
      stop
    end if
    hs_ready = .false.
    do_ctr = 0
    if( present(contract) )then
      if( contract .and. do_derivatives ) do_ctr = 1
    end if
    b = trim(adjustl(basis))
    if( b == "poly3" )then
      ibasis = 0
    else if( b == "poly3gauss" )then
      ibasis = 1
    else if( b == "poly3operator" )then
      ibasis = 2
    else
      write(*,*) "ERROR: soap_turbo: basis ", b, " is not supported by this build (poly3, poly3gauss, poly3operator)"
      stop
    end if
    iscal = 0
    if( scaling_mode /= "polynomial" ) iscal = 1
    call ensure_context()

    c_am = alpha_max(1:n_species)
    c_rch = rcut_hard(1:n_species); c_rcs = rcut_soft(1:n_species); c_nf = nf(1:n_species)
    c_gs = global_scaling(1:n_species); c_sr = atom_sigma_r(1:n_species)
    c_srs = atom_sigma_r_scaling(1:n_species); c_st = atom_sigma_t(1:n_species)
    c_sts = atom_sigma_t_scaling(1:n_species); c_as = amplitude_scaling(1:n_species)
    c_cw = central_weight(1:n_species)
    prm = st_params(int(n_species, c_int32_t), int(l_max, c_int32_t), int(radial_enhancement, c_int32_t), ibasis, &
                    iscal, merge(1_c_int32_t, 0_c_int32_t, legacy_filter_seed), c_loc(c_am), c_loc(c_rch), &
                    c_loc(c_rcs), c_loc(c_nf), c_loc(c_gs), c_loc(c_sr), c_loc(c_srs), c_loc(c_st), c_loc(c_sts), &
                    c_loc(c_as), c_loc(c_cw))

    c_nn = n_neigh(1:n_sites)
    c_sp = species(:, 1:n_sites)
    c_sm = species_multiplicity(1:n_sites)
    allocate( c_mask(n_atom_pairs, n_species) )
    c_mask = merge(1_c_int32_t, 0_c_int32_t, mask(1:n_atom_pairs, 1:n_species))
    c_r = rjs(1:n_atom_pairs); c_t = thetas(1:n_atom_pairs); c_f = phis(1:n_atom_pairs)
    nb = st_neighbours(int(n_sites, c_int32_t), int(n_atom_pairs, c_int32_t), int(n_species, c_int32_t), &
                       int(size(species, 1), c_int32_t), c_loc(c_nn), c_loc(c_sp), c_loc(c_sm), c_loc(c_mask), &
                       c_loc(c_r), c_loc(c_t), c_loc(c_f))

    if( compress_soap )then
      c_pi = compress_P_i(1:compress_P_nonzero)
      c_pj = compress_P_j(1:compress_P_nonzero)
      c_pel = compress_P_el(1:compress_P_nonzero)
      cmp = st_compression(int(compress_P_nonzero, c_int32_t), c_loc(c_pi), c_loc(c_pj), c_loc(c_pel))
    else
      cmp = st_compression(0_c_int32_t, c_null_ptr, c_null_ptr, c_null_ptr)
    end if

    n_soap = int(size(soap, 1), c_int32_t)
    if( do_derivatives .and. do_ctr == 0 )then
      rc = st_get_soap(ctx, c_loc(prm), c_loc(nb), c_loc(cmp), 1_c_int32_t, 0_c_int32_t, n_soap, c_loc(soap), &
                       c_loc(soap_cart_der))
    else
      rc = st_get_soap(ctx, c_loc(prm), c_loc(nb), c_loc(cmp), merge(1_c_int32_t, 0_c_int32_t, do_derivatives), &
                       do_ctr, n_soap, c_loc(soap), c_null_ptr)
    end if
    call check(rc)
    if( soap_hessian_enabled .and. do_derivatives )then
      hs_am = c_am; hs_rch = c_rch; hs_rcs = c_rcs; hs_nf = c_nf; hs_gs = c_gs; hs_sr = c_sr; hs_srs = c_srs
      hs_st = c_st; hs_sts = c_sts; hs_as = c_as; hs_cw = c_cw
      hs_prm = st_params(prm%n_species, prm%l_max, prm%radial_enhancement, prm%basis, prm%scaling_mode, &
                         prm%legacy_filter_seed, c_loc(hs_am), c_loc(hs_rch), c_loc(hs_rcs), c_loc(hs_nf), &
                         c_loc(hs_gs), c_loc(hs_sr), c_loc(hs_srs), c_loc(hs_st), c_loc(hs_sts), c_loc(hs_as), &
                         c_loc(hs_cw))
      if( compress_soap )then
        hs_pi = c_pi; hs_pj = c_pj; hs_pel = c_pel
        hs_cmp = st_compression(cmp%P_nonzero, c_loc(hs_pi), c_loc(hs_pj), c_loc(hs_pel))
      else
        hs_cmp = st_compression(0_c_int32_t, c_null_ptr, c_null_ptr, c_null_ptr)
      end if
      hs_n_sites = n_sites
      hs_n_pairs = n_atom_pairs
      hs_l_max = l_max
      hs_n_soap = n_soap
      hs_ready = .true.
    end if
    if( do_ctr == 1 )then
      ct_n_sites = n_sites
      ct_n_pairs = n_atom_pairs
      ct_n_soap = n_soap
    else
      ct_n_sites = 0
    end if

    if( do_timing )then
      call system_clock(t1)
      write(*,*)'                                       |'
      write(*,*)'SOAP timings:                          |'
      write(*,'(A, F8.3, A)') '  *) get_soap on the GPU: ', real(t1 - t0, dp) / real(rate, dp), ' seconds |'
      write(*,*)'.......................................|'
    end if
  end subroutine get_soap

  subroutine soap_contract(soap, G, sgn, pair)
    implicit none
    real(dp), intent(in), target, contiguous :: soap(:,:)
    real(dp), intent(in), target, contiguous :: G(:,:)
    real(dp), intent(in) :: sgn
    real(dp), intent(out), target, contiguous :: pair(:,:)
    integer(c_int) :: rc
    if( ct_n_sites == 0 .or. .not. c_associated(ctx) )then
      write(*,*) "ERROR: soap_contract needs get_soap called with contract = .true. first"
      stop
    end if
    if( size(soap, 1) /= ct_n_soap .or. size(soap, 2) /= ct_n_sites .or. size(G, 1) /= ct_n_sites .or. &
        size(G, 2) /= ct_n_soap .or. size(pair, 1) /= 3 .or. size(pair, 2) /= ct_n_pairs )then
      write(*,*) "ERROR: soap_contract: array shapes do not match the last get_soap"
      stop
    end if
    rc = st_soap_contract(ctx, c_loc(soap), c_loc(G), real(sgn, c_double), c_loc(pair))
    call check(rc)
  end subroutine soap_contract

  ! get_soap_central_hessian with the reference's argument list (src/soap_turbo.f90:1151-1172), after
  ! a get_soap with soap_hessian_enabled and do_derivatives on the same batch. The library recomputes
  ! what it needs on the resident neighbour list; hess(a, b, v, k) is sum_d vecs(d, v, i)
  ! d^2 q_d / dr_ia dr_jb. The self-test of the reference compares a per-pair derivative this build
  ! does not form: soap_hessian_selftest_err is left at zero.
  subroutine get_soap_central_hessian(n_sites, n_neigh, n_atom_pairs, n_species, mask, rjs, thetas, phis, l_max, &
                                      rcut_hard, atom_sigma_t, atom_sigma_t_scaling, compress_soap, &
                                      compress_P_nonzero, compress_P_i, compress_P_j, compress_P_el, soap, n_vec, &
                                      vecs, hess)
    implicit none
    integer, intent(in) :: n_sites, n_neigh(:), n_atom_pairs, n_species, l_max, n_vec
    integer, intent(in) :: compress_P_nonzero, compress_P_i(:), compress_P_j(:)
    logical, intent(in) :: mask(:,:), compress_soap
    real(dp), intent(in) :: rjs(:), thetas(:), phis(:), rcut_hard(:)
    real(dp), intent(in) :: atom_sigma_t(:), atom_sigma_t_scaling(:), compress_P_el(:)
    real(dp), intent(in) :: soap(:,:)
    real(dp), intent(in), target, contiguous :: vecs(:,:,:)
    real(dp), intent(out), target, contiguous :: hess(:,:,:,:)
    integer(c_int) :: rc
    if( .not. hs_ready .or. .not. c_associated(ctx) )then
      write(*,*) "ERROR: get_soap_central_hessian was called without a matching get_soap."
      write(*,*) "       Set soap_hessian_enabled = .true. and call get_soap with"
      write(*,*) "       do_derivatives = .true. first."
      stop
    end if
    if( n_sites /= hs_n_sites .or. n_atom_pairs /= hs_n_pairs .or. l_max /= hs_l_max .or. &
        size(soap, 1) /= hs_n_soap )then
      write(*,*) "ERROR: get_soap_central_hessian was called on a different batch than the"
      write(*,*) "       get_soap that filled the buffers it reads."
      stop
    end if
    if( size(vecs, 1) /= hs_n_soap .or. size(vecs, 2) < n_vec .or. size(vecs, 3) < n_sites .or. &
        size(hess, 1) /= 3 .or. size(hess, 2) /= 3 .or. size(hess, 3) /= n_vec .or. size(hess, 4) < n_atom_pairs )then
      write(*,*) "ERROR: get_soap_central_hessian: vecs(n_soap, n_vec, n_sites) and hess(3, 3, n_vec, n_atom_pairs)"
      stop
    end if
    if( size(vecs, 2) /= n_vec .or. size(vecs, 3) /= n_sites )then
      write(*,*) "ERROR: get_soap_central_hessian: vecs must be exactly (n_soap, n_vec, n_sites)"
      stop
    end if
    rc = st_central_hessian(ctx, c_loc(hs_prm), c_loc(hs_cmp), int(n_vec, c_int32_t), c_loc(vecs), c_loc(hess))
    call check(rc)
    soap_hessian_selftest_err = 0.d0
  end subroutine

  ! Species bookkeeping of the caller side (src/soap_turbo.f90:2099-2221): the
  ! species indices of every site, and the per-pair species mask, where a
  ! neighbour j counts for species s when its species name matches; the
  ! central atom's multiplicity sets how many species a neighbour may take.
  subroutine assign_species_multiplicity(max_species_multiplicity, species_types_soap, xyz_species, &
                                         xyz_species_supercell, n_species_soap, all_atoms, which_atom, &
                                         indices, n_neigh, n_atom_pairs, neighbors_list, mask_species, &
                                         species, species_multiplicity, species_multiplicity_supercell )
    implicit none
    integer, intent(in) :: max_species_multiplicity, n_species_soap, which_atom, indices(1:3), &
                           n_neigh(:), n_atom_pairs, neighbors_list(:)
    character*8, intent(in) :: species_types_soap(:), xyz_species(:), xyz_species_supercell(:)
    logical, intent(in) :: all_atoms
    integer, allocatable, intent(out) :: species(:,:), species_multiplicity(:), species_multiplicity_supercell(:)
    logical, allocatable, intent(out) :: mask_species(:,:)
    integer, allocatable :: sp_super(:,:)
    integer :: n_sites, n_super, i, j, k, k2, i2, centre, mult_keep

    n_sites = size(xyz_species)
    n_super = size(xyz_species_supercell)
    allocate( species(1:max_species_multiplicity, 1:n_sites), species_multiplicity(1:n_sites) )
    species = 0
    species_multiplicity = 0
    do i = 1, n_sites
      do j = 1, n_species_soap
        if( xyz_species(i) == species_types_soap(j) )then
          species_multiplicity(i) = species_multiplicity(i) + 1
          species(species_multiplicity(i), i) = j
        end if
      end do
    end do
    allocate( sp_super(1:max_species_multiplicity, 1:n_super), species_multiplicity_supercell(1:n_super) )
    if( n_super > n_sites )then
      sp_super = 0
      species_multiplicity_supercell = 0
      do i = 1, n_super
        do j = 1, n_species_soap
          if( xyz_species_supercell(i) == species_types_soap(j) )then
            species_multiplicity_supercell(i) = species_multiplicity_supercell(i) + 1
            sp_super(species_multiplicity_supercell(i), i) = j
          end if
        end do
      end do
    else
      sp_super = species
      species_multiplicity_supercell = species_multiplicity
    end if
    if( .not. all_atoms )then
      ! one site only: which_atom's species go first, swapped with atom 1 in the supercell list
      n_sites = 1
      deallocate( species )
      allocate( species(1:max_species_multiplicity, 1) )
      species(:, 1) = sp_super(:, which_atom)
      sp_super(:, which_atom) = sp_super(:, 1)
      sp_super(:, 1) = species(:, 1)
      mult_keep = species_multiplicity(which_atom)
      deallocate( species_multiplicity )
      allocate( species_multiplicity(1) )
      species_multiplicity(1) = mult_keep
    end if
    allocate( mask_species(1:n_atom_pairs, 1:n_species_soap) )
    mask_species = .false.
    k2 = 0
    centre = 0
    do i = 1, n_sites
      do k = 1, n_neigh(i)
        k2 = k2 + 1
        j = neighbors_list(k2)
        if( k == 1 ) centre = j
        do i2 = 1, species_multiplicity_supercell(centre)
          if( sp_super(i2, j) > 0 ) mask_species(k2, sp_super(i2, j)) = .true.
        end do
      end do
    end do
    deallocate( sp_super )
  end subroutine assign_species_multiplicity

end module soap_turbo_desc
