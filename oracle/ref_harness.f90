! Test harness around the reference soap_turbo library (compiled from
! /root/reference/src by oracle/Makefile into oracle/_ref; never shipped).
!
! It reads one problem in the binary layout written by oracle/refio.py, calls
! the reference routines the get_soap hot path is made of, and writes their
! outputs back, so that the project's restatement and kernels can be compared
! with the reference on identical inputs:
!   * W and S per species (get_orthonormalization_matrix_poly3[gauss],
!     src/soap_turbo_radial.f90:1023, :1113), assembled block-diagonally as
!     get_soap does (src/soap_turbo.f90:364-386);
!   * get_compress_indices "trivial" and the nu_R_nu_S modes "0_0" .. "2_2"
!     (src/soap_turbo_compress.f90:44);
!   * get_soap without derivatives, with derivatives (soap_cart_der), and with
!     contract = .true. followed by soap_contract (src/soap_turbo.f90:139, :1723);
!   * the radial and angular expansions alone (src/soap_turbo.f90:491-588);
!   * the central-atom Hessian: get_soap with soap_hessian_enabled (and the
!     self-test on), then get_soap_central_hessian for n_vec weight vectors
!     (src/soap_turbo.f90:1151), plus the radial second derivatives alone;
!   * wall-clock timings (omp_get_wtime) of the full-Jacobian path (get_soap
!     with derivatives, then pair = sum_s G(i,s) soap_cart_der(:,s,k)) and of
!     the contract path.
program ref_harness
  use, intrinsic :: iso_fortran_env, only: real64, int32
  use omp_lib
  use soap_turbo_functions
  use soap_turbo_radial
  use soap_turbo_angular
  use soap_turbo_desc
  use soap_turbo_compress_module
  implicit none
  integer, parameter :: dp = real64
  character(len=1024) :: fin, fout
  integer :: hdr(16), n_sites, n_pairs, n_species, l_max, max_mult, radial_enhancement, ibasis, iscal
  integer :: icompress, ilegacy, modes, n_reps, with_G, n_vec
  real(dp), allocatable :: vecs(:,:,:), hess(:,:,:,:), rad2(:,:)
  integer, allocatable :: alpha_max(:), n_neigh(:), species(:,:), species_multiplicity(:), imask(:)
  real(dp), allocatable :: rcut_hard(:), rcut_soft(:), nf(:), global_scaling(:), atom_sigma_r(:), &
                           atom_sigma_r_scaling(:), atom_sigma_t(:), atom_sigma_t_scaling(:), &
                           amplitude_scaling(:), central_weight(:)
  logical, allocatable :: mask(:,:)
  real(dp), allocatable :: rjs(:), thetas(:), phis(:), G(:,:)
  real(dp), allocatable :: W(:,:), S(:,:), Wt(:,:), St(:,:)
  real(dp), allocatable :: soap(:,:), soap_cart_der(:,:,:), pair(:,:), P_el(:), cart_dummy(:,:,:)
  integer, allocatable :: P_i(:), P_j(:), i_beg(:), i_end(:)
  integer :: n_max, k_max, n_soap, n_soap_unc, P_nonzero, dim, i, rep
  character(len=16) :: basis, scaling_mode
  character(len=8) :: cmode
  logical :: compress, do_central
  real(dp) :: t0, t1, t_full, t_contract, rcut_max
  ! expansion dump
  real(dp), allocatable :: rad(:,:), rad_der(:,:), preflm(:), plm_array(:), prefl(:), fact_array(:), &
                           prefl_rad_der(:)
  complex(dp), allocatable :: ang(:,:), ang_r(:,:), ang_a(:,:), ang_p(:,:), eimphi(:), prefm(:), &
                              eimphi_rad_der(:)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open(unit=11, file=trim(fin), access="stream", form="unformatted", status="old")
  read(11) hdr
  n_sites = hdr(2); n_pairs = hdr(3); n_species = hdr(4); l_max = hdr(5); max_mult = hdr(6)
  radial_enhancement = hdr(7); ibasis = hdr(8); iscal = hdr(9); icompress = hdr(10)
  ilegacy = hdr(11); modes = hdr(12); n_reps = max(hdr(13), 1); with_G = hdr(15); n_vec = hdr(14)
  allocate( alpha_max(n_species), rcut_hard(n_species), rcut_soft(n_species), nf(n_species), &
            global_scaling(n_species), atom_sigma_r(n_species), atom_sigma_r_scaling(n_species), &
            atom_sigma_t(n_species), atom_sigma_t_scaling(n_species), amplitude_scaling(n_species), &
            central_weight(n_species) )
  read(11) alpha_max
  read(11) rcut_hard, rcut_soft, nf, global_scaling, atom_sigma_r, atom_sigma_r_scaling, atom_sigma_t, &
           atom_sigma_t_scaling, amplitude_scaling, central_weight
  allocate( n_neigh(n_sites), species(max_mult, n_sites), species_multiplicity(n_sites) )
  read(11) n_neigh
  read(11) species
  read(11) species_multiplicity
  allocate( imask(n_pairs*n_species), mask(n_pairs, n_species) )
  read(11) imask
  mask = reshape( imask /= 0, [n_pairs, n_species] )
  allocate( rjs(n_pairs), thetas(n_pairs), phis(n_pairs) )
  read(11) rjs
  read(11) thetas
  read(11) phis

  basis = "poly3"
  if( ibasis == 1 ) basis = "poly3gauss"
  if( ibasis == 2 ) basis = "poly3operator"
  scaling_mode = "polynomial"
  if( iscal /= 0 ) scaling_mode = "none"
  legacy_filter_seed = ( ilegacy /= 0 )
  compress = ( icompress /= 0 )
  ! header code: 1 "trivial", 2 + 3 nu_R + nu_S the mode "nu_R_nu_S"
  cmode = "trivial"
  if( icompress >= 2 ) write(cmode, '(I1,A,I1)') (icompress - 2) / 3, "_", mod(icompress - 2, 3)

  n_max = sum(alpha_max)
  k_max = 1 + l_max*(l_max+1)/2 + l_max
  n_soap_unc = n_max*(n_max+1)/2*(l_max+1)
  if( compress )then
    allocate( P_i(1), P_j(1), P_el(1) )
    call get_compress_indices(trim(cmode), alpha_max, l_max, dim, P_nonzero, P_i, P_j, P_el, "get_dim")
    deallocate( P_i, P_j, P_el )
    allocate( P_i(P_nonzero), P_j(P_nonzero), P_el(P_nonzero) )
    call get_compress_indices(trim(cmode), alpha_max, l_max, dim, P_nonzero, P_i, P_j, P_el, "set_indices")
    n_soap = maxval(P_i)
  else
    P_nonzero = 1
    allocate( P_i(1), P_j(1), P_el(1) )
    P_i = 1; P_j = 1; P_el = 0.d0
    n_soap = n_soap_unc
  end if
  if( with_G /= 0 )then
    allocate( G(n_sites, n_soap) )
    read(11) G
  end if
  if( n_vec > 0 )then
    allocate( vecs(n_soap, n_vec, n_sites) )
    read(11) vecs
  end if
  close(11)

  open(unit=12, file=trim(fout), access="stream", form="unformatted", status="replace")
  write(12) n_max, k_max, n_soap, P_nonzero
  if( compress )then
    write(12) P_i, P_j, P_el
  end if

  ! block-diagonal W and S, as get_soap assembles them
  allocate( i_beg(n_species), i_end(n_species) )
  i_beg(1) = 1; i_end(1) = alpha_max(1)
  do i = 2, n_species
    i_beg(i) = i_end(i-1) + 1
    i_end(i) = i_beg(i) + alpha_max(i) - 1
  end do
  allocate( W(n_max, n_max), S(n_max, n_max) )
  W = 0.d0; S = 0.d0
  do i = 1, n_species
    allocate( Wt(alpha_max(i), alpha_max(i)), St(alpha_max(i), alpha_max(i)) )
    Wt = 0.d0; St = 0.d0
    if( ibasis == 1 )then
      call get_orthonormalization_matrix_poly3gauss(alpha_max(i), atom_sigma_r(i), rcut_hard(i), St, Wt)
    else if( ibasis == 2 )then
      call get_orthonormalization_matrix_poly3_tabulated(alpha_max(i), St, Wt)
    else
      call get_orthonormalization_matrix_poly3(alpha_max(i), St, Wt)
    end if
    W(i_beg(i):i_end(i), i_beg(i):i_end(i)) = Wt
    S(i_beg(i):i_end(i), i_beg(i):i_end(i)) = St
    deallocate( Wt, St )
  end do
  write(12) W, S

  allocate( soap(n_soap, n_sites) )
  allocate( cart_dummy(3, n_soap, 1) )
  cart_dummy = 0.d0
  if( iand(modes, 1) /= 0 )then
    soap = 0.d0
    call run_get_soap(.false., .false., cart_dummy)
    write(12) soap
  end if

  if( iand(modes, 2) /= 0 .or. iand(modes, 16) /= 0 )then
    allocate( soap_cart_der(3, n_soap, n_pairs) )
  end if

  if( iand(modes, 2) /= 0 )then
    soap = 0.d0; soap_cart_der = 0.d0
    call run_get_soap(.true., .false., soap_cart_der)
    write(12) soap
    write(12) soap_cart_der
  end if

  if( iand(modes, 4) /= 0 )then
    allocate( pair(3, n_pairs) )
    soap = 0.d0
    call run_get_soap(.true., .true., cart_dummy)
    call soap_contract(soap, G, 1.d0, pair)
    write(12) soap
    write(12) pair
  end if

  if( iand(modes, 8) /= 0 )then
    allocate( rad(n_max, n_pairs), rad_der(n_max, n_pairs) )
    allocate( ang(k_max, n_pairs), ang_r(k_max, n_pairs), ang_a(k_max, n_pairs), ang_p(k_max, n_pairs) )
    do i = 1, n_species
      do_central = ( central_weight(i) /= 0.d0 )
      if( ibasis == 1 )then
        call get_radial_expansion_coefficients_poly3gauss(n_sites, n_neigh, rjs, alpha_max(i), rcut_soft(i), &
               rcut_hard(i), atom_sigma_r(i), atom_sigma_r_scaling(i), amplitude_scaling(i), nf(i), &
               W(i_beg(i):i_end(i), i_beg(i):i_end(i)), scaling_mode, mask(:,i), radial_enhancement, &
               .true., rad(i_beg(i):i_end(i), :), rad_der(i_beg(i):i_end(i), :))
      else if( ibasis == 2 )then
        call get_radial_expansion_coefficients_poly3operator(n_sites, n_neigh, rjs, alpha_max(i), rcut_soft(i), &
               rcut_hard(i), atom_sigma_r(i), atom_sigma_r_scaling(i), amplitude_scaling(i), &
               W(i_beg(i):i_end(i), i_beg(i):i_end(i)), scaling_mode, mask(:,i), radial_enhancement, &
               .true., .true., central_weight(i), rad(i_beg(i):i_end(i), :), rad_der(i_beg(i):i_end(i), :))
      else
        call get_radial_expansion_coefficients_poly3(n_sites, n_neigh, rjs, alpha_max(i), rcut_soft(i), &
               rcut_hard(i), atom_sigma_r(i), atom_sigma_r_scaling(i), amplitude_scaling(i), nf(i), &
               W(i_beg(i):i_end(i), i_beg(i):i_end(i)), scaling_mode, mask(:,i), radial_enhancement, &
               .true., do_central, central_weight(i), rad(i_beg(i):i_end(i), :), rad_der(i_beg(i):i_end(i), :))
      end if
      rad(i_beg(i):i_end(i), :) = rad(i_beg(i):i_end(i), :) * global_scaling(i)
      rad_der(i_beg(i):i_end(i), :) = rad_der(i_beg(i):i_end(i), :) * global_scaling(i)
    end do
    allocate( preflm(k_max), plm_array(k_max), eimphi(k_max), prefl(0:l_max), prefm(0:l_max), &
              fact_array(max(l_max,1)), prefl_rad_der(0:l_max), eimphi_rad_der(k_max) )
    call get_preflm(preflm, l_max)
    rcut_max = maxval(rcut_hard)
    call get_angular_expansion_coefficients(n_sites, n_neigh, thetas, phis, rjs, atom_sigma_t, &
           atom_sigma_t_scaling, rcut_max, l_max, eimphi, preflm, plm_array, prefl, prefm, fact_array, &
           mask, n_species, eimphi_rad_der, .true., prefl_rad_der, ang, ang_r, ang_a, ang_p)
    write(12) rad, rad_der
    write(12) ang, ang_r, ang_a, ang_p
  end if

  if( iand(modes, 32) /= 0 )then
    ! central-atom Hessian (needs legacy_filter_seed = .false.)
    if( .not. allocated(soap_cart_der) ) allocate( soap_cart_der(3, n_soap, n_pairs) )
    allocate( hess(3, 3, n_vec, n_pairs) )
    soap_hessian_enabled = .true.
    soap_hessian_selftest = .true.
    soap = 0.d0; soap_cart_der = 0.d0
    call run_get_soap(.true., .false., soap_cart_der)
    call get_soap_central_hessian(n_sites, n_neigh, n_pairs, n_species, mask, rjs, thetas, phis, l_max, &
                                  rcut_hard, atom_sigma_t, atom_sigma_t_scaling, compress, P_nonzero, P_i, &
                                  P_j, P_el, soap, n_vec, vecs, hess)
    soap_hessian_enabled = .false.
    write(12) soap
    write(12) hess
    write(12) soap_hessian_selftest_err
    ! the radial expansion's second derivative alone, species by species as get_soap does it
    allocate( rad2(n_max, n_pairs) )
    if( .not. allocated(rad) ) allocate( rad(n_max, n_pairs), rad_der(n_max, n_pairs) )
    do i = 1, n_species
      do_central = ( central_weight(i) /= 0.d0 )
      if( ibasis == 1 )then
        call get_radial_expansion_coefficients_poly3gauss(n_sites, n_neigh, rjs, alpha_max(i), rcut_soft(i), &
               rcut_hard(i), atom_sigma_r(i), atom_sigma_r_scaling(i), amplitude_scaling(i), nf(i), &
               W(i_beg(i):i_end(i), i_beg(i):i_end(i)), scaling_mode, mask(:,i), radial_enhancement, &
               .true., rad(i_beg(i):i_end(i), :), rad_der(i_beg(i):i_end(i), :), rad2(i_beg(i):i_end(i), :))
      else
        call get_radial_expansion_coefficients_poly3(n_sites, n_neigh, rjs, alpha_max(i), rcut_soft(i), &
               rcut_hard(i), atom_sigma_r(i), atom_sigma_r_scaling(i), amplitude_scaling(i), nf(i), &
               W(i_beg(i):i_end(i), i_beg(i):i_end(i)), scaling_mode, mask(:,i), radial_enhancement, &
               .true., do_central, central_weight(i), rad(i_beg(i):i_end(i), :), rad_der(i_beg(i):i_end(i), :), &
               rad2(i_beg(i):i_end(i), :))
      end if
      rad2(i_beg(i):i_end(i), :) = rad2(i_beg(i):i_end(i), :) * global_scaling(i)
    end do
    write(12) rad2
  end if

  if( iand(modes, 16) /= 0 )then
    if( .not. allocated(pair) ) allocate( pair(3, n_pairs) )
    ! full-Jacobian path: get_soap with derivatives, then contract soap_cart_der with G
    t_full = huge(1.d0)
    do rep = 1, n_reps
      soap = 0.d0
      t0 = omp_get_wtime()
      call zero_cart()
      call run_get_soap(.true., .false., soap_cart_der)
      call contract_full()
      t1 = omp_get_wtime()
      t_full = min(t_full, t1 - t0)
    end do
    t_contract = huge(1.d0)
    do rep = 1, n_reps
      soap = 0.d0
      t0 = omp_get_wtime()
      call run_get_soap(.true., .true., cart_dummy)
      call soap_contract(soap, G, 1.d0, pair)
      t1 = omp_get_wtime()
      t_contract = min(t_contract, t1 - t0)
    end do
    write(12) t_full, t_contract
    write(12) int(omp_get_max_threads(), int32)
  end if
  close(12)

contains

  subroutine zero_cart()
    integer :: kk
    !$omp parallel do schedule(static)
    do kk = 1, n_pairs
      soap_cart_der(:, :, kk) = 0.d0
    end do
    !$omp end parallel do
  end subroutine

  subroutine run_get_soap(do_der, do_contract, cart)
    logical, intent(in) :: do_der, do_contract
    real(dp), intent(inout) :: cart(:,:,:)
    call get_soap(n_sites, n_neigh, n_species, species, species_multiplicity, n_pairs, mask, rjs, &
                  thetas, phis, alpha_max, l_max, rcut_hard, rcut_soft, nf, global_scaling, atom_sigma_r, &
                  atom_sigma_r_scaling, atom_sigma_t, atom_sigma_t_scaling, amplitude_scaling, &
                  radial_enhancement, central_weight, basis, scaling_mode, .false., do_der, compress, &
                  P_nonzero, P_i, P_j, P_el, soap, cart, do_contract)
  end subroutine

  ! pair(:,k) = sum_s G(i,s) soap_cart_der(:,s,k): what a caller does with the Jacobian
  subroutine contract_full()
    integer :: ii, jj, kk, ss, kb
    integer, allocatable :: kf(:)
    allocate( kf(n_sites) )
    kb = 0
    do ii = 1, n_sites
      kf(ii) = kb
      kb = kb + n_neigh(ii)
    end do
    !$omp parallel do schedule(dynamic, 16) private(ii, jj, kk, ss)
    do ii = 1, n_sites
      do jj = 1, n_neigh(ii)
        kk = kf(ii) + jj
        pair(:, kk) = 0.d0
        do ss = 1, n_soap
          pair(:, kk) = pair(:, kk) + G(ii, ss) * soap_cart_der(:, ss, kk)
        end do
      end do
    end do
    !$omp end parallel do
    deallocate( kf )
  end subroutine

end program ref_harness
