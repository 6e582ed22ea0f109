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
  use soap_turbo_radial, only: legacy_filter_seed, extended_precision_w
  implicit none
  private
  integer, parameter :: dp = real64

  public :: get_soap, soap_contract, get_soap_central_hessian, get_soap_adjoint_forces
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

  ! contiguous C-interoperable copies of one call's get_soap inputs, at module level so that the
  ! structures pointing into them stay valid while the library uses them
  integer(c_int32_t), allocatable, target, save :: pk_am(:), pk_nn(:), pk_sp(:,:), pk_sm(:), pk_mask(:,:), &
                                                   pk_pi(:), pk_pj(:)
  real(c_double), allocatable, target, save :: pk_r(:), pk_t(:), pk_f(:), pk_pel(:), pk_rch(:), pk_rcs(:), &
                                               pk_nf(:), pk_gs(:), pk_sr(:), pk_srs(:), pk_st(:), pk_sts(:), &
                                               pk_as(:), pk_cw(:)
  type(st_params), target, save :: pk_prm
  type(st_neighbours), target, save :: pk_nb
  type(st_compression), target, save :: pk_cmp

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

    integer(c_int32_t) :: do_ctr, n_soap
    integer(c_int) :: rc
    integer(8) :: t0, t1, rate

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
    call pack_inputs(n_sites, n_neigh, n_species, species, species_multiplicity, n_atom_pairs, mask, rjs, thetas, &
                     phis, alpha_max, l_max, rcut_hard, rcut_soft, nf, global_scaling, atom_sigma_r, &
                     atom_sigma_r_scaling, atom_sigma_t, atom_sigma_t_scaling, amplitude_scaling, &
                     radial_enhancement, central_weight, basis, scaling_mode, compress_soap, compress_P_nonzero, &
                     compress_P_i, compress_P_j, compress_P_el)

    n_soap = int(size(soap, 1), c_int32_t)
    if( do_derivatives .and. do_ctr == 0 )then
      rc = st_get_soap(ctx, c_loc(pk_prm), c_loc(pk_nb), c_loc(pk_cmp), 1_c_int32_t, 0_c_int32_t, n_soap, &
                       c_loc(soap), c_loc(soap_cart_der))
    else
      rc = st_get_soap(ctx, c_loc(pk_prm), c_loc(pk_nb), c_loc(pk_cmp), &
                       merge(1_c_int32_t, 0_c_int32_t, do_derivatives), do_ctr, n_soap, c_loc(soap), c_null_ptr)
    end if
    call check(rc)
    if( soap_hessian_enabled .and. do_derivatives )then
      hs_am = pk_am; hs_rch = pk_rch; hs_rcs = pk_rcs; hs_nf = pk_nf; hs_gs = pk_gs; hs_sr = pk_sr
      hs_srs = pk_srs; hs_st = pk_st; hs_sts = pk_sts; hs_as = pk_as; hs_cw = pk_cw
      hs_prm = st_params(pk_prm%n_species, pk_prm%l_max, pk_prm%radial_enhancement, pk_prm%basis, &
                         pk_prm%scaling_mode, pk_prm%legacy_filter_seed, c_loc(hs_am), c_loc(hs_rch), &
                         c_loc(hs_rcs), c_loc(hs_nf), c_loc(hs_gs), c_loc(hs_sr), c_loc(hs_srs), c_loc(hs_st), &
                         c_loc(hs_sts), c_loc(hs_as), c_loc(hs_cw))
      if( compress_soap )then
        hs_pi = pk_pi; hs_pj = pk_pj; hs_pel = pk_pel
        hs_cmp = st_compression(pk_cmp%P_nonzero, c_loc(hs_pi), c_loc(hs_pj), c_loc(hs_pel))
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

  ! The get_soap inputs as the library's structures (pk_prm, pk_nb, pk_cmp), over module-level copies.
  subroutine pack_inputs(n_sites, n_neigh, n_species, species, species_multiplicity, n_atom_pairs, mask, rjs, &
                         thetas, phis, alpha_max, l_max, rcut_hard, rcut_soft, nf, global_scaling, atom_sigma_r, &
                         atom_sigma_r_scaling, atom_sigma_t, atom_sigma_t_scaling, amplitude_scaling, &
                         radial_enhancement, central_weight, basis, scaling_mode, compress_soap, compress_P_nonzero, &
                         compress_P_i, compress_P_j, compress_P_el)
    implicit none
    real(dp), intent(in) :: rjs(:), thetas(:), phis(:)
    real(dp), intent(in) :: amplitude_scaling(:), atom_sigma_r_scaling(:), atom_sigma_t(:), atom_sigma_t_scaling(:)
    real(dp), intent(in) :: central_weight(:), atom_sigma_r(:), global_scaling(:)
    real(dp), intent(in) :: nf(:), rcut_hard(:), rcut_soft(:)
    real(dp), intent(in) :: compress_P_el(:)
    integer, intent(in) :: n_species, radial_enhancement, species(:,:), species_multiplicity(:)
    integer, intent(in) :: n_sites, n_neigh(:), l_max, n_atom_pairs, alpha_max(:), compress_P_nonzero, &
                           compress_P_i(:), compress_P_j(:)
    logical, intent(in) :: mask(:,:), compress_soap
    character(*), intent(in) :: basis, scaling_mode
    integer(c_int32_t) :: ibasis, iscal
    character(len=:), allocatable :: b

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
    call check(st_set_basis_route(ctx, merge(ST_BASIS_ROUTE_EXTENDED, ST_BASIS_ROUTE_LAPACK, extended_precision_w)))

    pk_am = alpha_max(1:n_species)
    pk_rch = rcut_hard(1:n_species); pk_rcs = rcut_soft(1:n_species); pk_nf = nf(1:n_species)
    pk_gs = global_scaling(1:n_species); pk_sr = atom_sigma_r(1:n_species)
    pk_srs = atom_sigma_r_scaling(1:n_species); pk_st = atom_sigma_t(1:n_species)
    pk_sts = atom_sigma_t_scaling(1:n_species); pk_as = amplitude_scaling(1:n_species)
    pk_cw = central_weight(1:n_species)
    pk_prm = st_params(int(n_species, c_int32_t), int(l_max, c_int32_t), int(radial_enhancement, c_int32_t), &
                       ibasis, iscal, merge(1_c_int32_t, 0_c_int32_t, legacy_filter_seed), c_loc(pk_am), &
                       c_loc(pk_rch), c_loc(pk_rcs), c_loc(pk_nf), c_loc(pk_gs), c_loc(pk_sr), c_loc(pk_srs), &
                       c_loc(pk_st), c_loc(pk_sts), c_loc(pk_as), c_loc(pk_cw))

    pk_nn = n_neigh(1:n_sites)
    pk_sp = species(:, 1:n_sites)
    pk_sm = species_multiplicity(1:n_sites)
    if( allocated(pk_mask) ) deallocate( pk_mask )
    allocate( pk_mask(n_atom_pairs, n_species) )
    pk_mask = merge(1_c_int32_t, 0_c_int32_t, mask(1:n_atom_pairs, 1:n_species))
    pk_r = rjs(1:n_atom_pairs); pk_t = thetas(1:n_atom_pairs); pk_f = phis(1:n_atom_pairs)
    pk_nb = st_neighbours(int(n_sites, c_int32_t), int(n_atom_pairs, c_int32_t), int(n_species, c_int32_t), &
                          int(size(species, 1), c_int32_t), c_loc(pk_nn), c_loc(pk_sp), c_loc(pk_sm), &
                          c_loc(pk_mask), c_loc(pk_r), c_loc(pk_t), c_loc(pk_f))

    if( compress_soap )then
      pk_pi = compress_P_i(1:compress_P_nonzero)
      pk_pj = compress_P_j(1:compress_P_nonzero)
      pk_pel = compress_P_el(1:compress_P_nonzero)
      pk_cmp = st_compression(int(compress_P_nonzero, c_int32_t), c_loc(pk_pi), c_loc(pk_pj), c_loc(pk_pel))
    else
      pk_cmp = st_compression(0_c_int32_t, c_null_ptr, c_null_ptr, c_null_ptr)
    end if
  end subroutine pack_inputs

  ! The FP32 adjoint, callable from Fortran (no reference counterpart: the FP32 form of
  ! get_soap(..., contract=.true.) followed by soap_contract(soap, G, sgn, pair), src/soap_turbo.f90:1703-1730,
  ! with the caller's assembly of forces from the pair gradients). It takes get_soap's inputs, dE/dq as
  ! G(n_sites, n_soap) (soap_contract's layout), sgn, and each pair's neighbour atom
  ! (neighbors_list(n_atom_pairs), 1-based, the site's own atom for its first pair), and returns
  !   soap(n_soap, n_sites)  the normalised descriptors (computed in FP32),
  !   forces(3, n_atoms)     forces(:, a) = -sgn sum of the pair gradients of the pairs whose neighbour is a,
  !   virial(3, 3)           -sgn sum_k r_k (x) pair_k over the pairs, r_k = xyz(:, k) (the pair vector
  !                          r_j - r_i) if given, else rebuilt from rjs, thetas, phis,
  ! without forming soap_cart_der. The pair gradient is soap_contract's with G. The library's resident
  ! state is replaced: soap_contract and get_soap_central_hessian need a new get_soap afterwards.
  subroutine get_soap_adjoint_forces(n_sites, n_neigh, n_species, species, species_multiplicity, n_atom_pairs, &
                                     mask, rjs, thetas, phis, alpha_max, l_max, rcut_hard, rcut_soft, nf, &
                                     global_scaling, atom_sigma_r, atom_sigma_r_scaling, atom_sigma_t, &
                                     atom_sigma_t_scaling, amplitude_scaling, radial_enhancement, central_weight, &
                                     basis, scaling_mode, compress_soap, compress_P_nonzero, compress_P_i, &
                                     compress_P_j, compress_P_el, soap, G, sgn, neighbors_list, n_atoms, forces, &
                                     virial, xyz)
    implicit none
    real(dp), intent(in) :: rjs(:), thetas(:), phis(:)
    real(dp), intent(in) :: amplitude_scaling(:), atom_sigma_r_scaling(:), atom_sigma_t(:), atom_sigma_t_scaling(:)
    real(dp), intent(in) :: central_weight(:), atom_sigma_r(:), global_scaling(:)
    real(dp), intent(in) :: nf(:), rcut_hard(:), rcut_soft(:)
    real(dp), intent(in) :: compress_P_el(:)
    integer, intent(in) :: n_species, radial_enhancement, species(:,:), species_multiplicity(:)
    integer, intent(in) :: n_sites, n_neigh(:), l_max, n_atom_pairs, alpha_max(:), compress_P_nonzero, &
                           compress_P_i(:), compress_P_j(:)
    logical, intent(in) :: mask(:,:), compress_soap
    character(*), intent(in) :: basis, scaling_mode
    real(dp), intent(out), target, contiguous :: soap(:,:)
    real(dp), intent(in), target, contiguous :: G(:,:)
    real(dp), intent(in) :: sgn
    integer, intent(in) :: neighbors_list(:), n_atoms
    real(dp), intent(out), target, contiguous :: forces(:,:), virial(:,:)
    real(dp), intent(in), target, contiguous, optional :: xyz(:,:)
    integer(c_int32_t), allocatable, target :: pair_atom(:)
    integer(c_int32_t) :: n_soap
    integer(c_int) :: rc
    type(c_ptr) :: vec

    hs_ready = .false.
    ct_n_sites = 0
    call pack_inputs(n_sites, n_neigh, n_species, species, species_multiplicity, n_atom_pairs, mask, rjs, thetas, &
                     phis, alpha_max, l_max, rcut_hard, rcut_soft, nf, global_scaling, atom_sigma_r, &
                     atom_sigma_r_scaling, atom_sigma_t, atom_sigma_t_scaling, amplitude_scaling, &
                     radial_enhancement, central_weight, basis, scaling_mode, compress_soap, compress_P_nonzero, &
                     compress_P_i, compress_P_j, compress_P_el)
    n_soap = int(size(soap, 1), c_int32_t)
    if( size(soap, 2) /= n_sites .or. size(G, 1) /= n_sites .or. size(G, 2) /= n_soap .or. &
        size(neighbors_list) < n_atom_pairs .or. size(forces, 1) /= 3 .or. size(forces, 2) < n_atoms .or. &
        size(virial, 1) /= 3 .or. size(virial, 2) /= 3 )then
      write(*,*) "ERROR: get_soap_adjoint_forces: soap(n_soap, n_sites), G(n_sites, n_soap), ", &
                 "neighbors_list(n_atom_pairs), forces(3, n_atoms), virial(3, 3)"
      stop
    end if
    vec = c_null_ptr
    if( present(xyz) )then
      if( size(xyz, 1) /= 3 .or. size(xyz, 2) < n_atom_pairs )then
        write(*,*) "ERROR: get_soap_adjoint_forces: xyz(3, n_atom_pairs)"
        stop
      end if
      vec = c_loc(xyz)
    end if
    pair_atom = int(neighbors_list(1:n_atom_pairs) - 1, c_int32_t)
    rc = st_upload_neighbours(ctx, c_loc(pk_nb))
    call check(rc)
    rc = st_upload_atoms(ctx, int(n_atoms, c_int32_t), c_loc(pair_atom), vec)
    call check(rc)
    rc = st_upload_dedq(ctx, n_soap, c_loc(G))
    call check(rc)
    rc = st_adjoint_f32(ctx, c_loc(pk_prm), c_loc(pk_cmp))
    call check(rc)
    rc = st_download_soap_f32(ctx, c_loc(soap))
    call check(rc)
    rc = st_download_forces(ctx, 1_c_int32_t, c_loc(forces), c_loc(virial))
    call check(rc)
    if( sgn /= 1.d0 )then
      forces(:, 1:n_atoms) = sgn * forces(:, 1:n_atoms)
      virial = sgn * virial
    end if
  end subroutine get_soap_adjoint_forces

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
  ! d^2 q_d / dr_ia dr_jb. With soap_hessian_selftest, soap_hessian_selftest_err gets the largest
  ! relative difference between the first derivatives the Hessian's contraction also produces and
  ! those of the independent FP64 Jacobian (st_hessian_selftest).
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
    real(c_double), target :: selftest_err
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
    if( soap_hessian_selftest )then
      ! the Hessian route's first derivatives against the independent Jacobian's (st_hessian_selftest)
      rc = st_hessian_selftest(ctx, c_loc(hs_prm), c_loc(hs_cmp), int(n_vec, c_int32_t), c_loc(vecs), &
                               c_loc(selftest_err))
      call check(rc)
      soap_hessian_selftest_err = selftest_err
    end if
  end subroutine

end module soap_turbo_desc
