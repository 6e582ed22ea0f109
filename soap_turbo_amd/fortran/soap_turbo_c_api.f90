! iso_c_binding interfaces of libsoapturbo_hip.so (include/soap_turbo_hip.h).
module soap_turbo_c_api
  use, intrinsic :: iso_c_binding
  implicit none
  private :: c_int32_t

  integer(c_int), parameter :: ST_OK = 0, ST_ERR_ARG = 1, ST_ERR_HIP = 2, ST_ERR_UNSUPPORTED = 3, &
                               ST_ERR_STATE = 4, ST_ERR_STOP = 5
  integer(c_int32_t), parameter :: ST_BASIS_ROUTE_EXTENDED = 0, ST_BASIS_ROUTE_LAPACK = 1

  type, bind(c) :: st_params
    integer(c_int32_t) :: n_species, l_max, radial_enhancement, basis, scaling_mode, legacy_filter_seed
    type(c_ptr) :: alpha_max, rcut_hard, rcut_soft, nf, global_scaling, atom_sigma_r, atom_sigma_r_scaling, &
                   atom_sigma_t, atom_sigma_t_scaling, amplitude_scaling, central_weight
  end type

  type, bind(c) :: st_neighbours
    integer(c_int32_t) :: n_sites, n_pairs, n_species, max_mult
    type(c_ptr) :: n_neigh, species, species_multiplicity, mask, rjs, thetas, phis
  end type

  type, bind(c) :: st_compression
    integer(c_int32_t) :: P_nonzero
    type(c_ptr) :: P_i, P_j, P_el
  end type

  interface
    function st_create(ctx, device) bind(c, name="st_create")
      import :: c_ptr, c_int
      type(c_ptr) :: ctx
      integer(c_int), value :: device
      integer(c_int) :: st_create
    end function

    function st_error(ctx) bind(c, name="st_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
      type(c_ptr) :: st_error
    end function

    function st_get_soap(ctx, p, nb, cmp, do_derivatives, contract, n_soap, soap, cart) bind(c, name="st_get_soap")
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: ctx, p, nb, cmp, soap, cart
      integer(c_int32_t), value :: do_derivatives, contract, n_soap
      integer(c_int) :: st_get_soap
    end function

    function st_soap_contract(ctx, soap, G, sgn, pair) bind(c, name="st_soap_contract")
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx, soap, G, pair
      real(c_double), value :: sgn
      integer(c_int) :: st_soap_contract
    end function

    function st_get_compress_indices(mode, n_species, alpha_max, l_max, dim, P_nonzero, P_i, P_j, P_el, &
                                     set_indices) bind(c, name="st_get_compress_indices")
      import :: c_ptr, c_int, c_int32_t, c_char
      character(kind=c_char), dimension(*) :: mode
      integer(c_int32_t), value :: n_species, l_max, set_indices
      type(c_ptr), value :: alpha_max, dim, P_nonzero, P_i, P_j, P_el
      integer(c_int) :: st_get_compress_indices
    end function

    function c_strlen(s) bind(c, name="strlen")
      import :: c_ptr, c_size_t
      type(c_ptr), value :: s
      integer(c_size_t) :: c_strlen
    end function
    function st_central_hessian(ctx, p, cmp, n_vec, vecs, hess) bind(c, name="st_central_hessian")
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: ctx, p, cmp, vecs, hess
      integer(c_int32_t), value :: n_vec
      integer(c_int) :: st_central_hessian
    end function
    function st_set_basis_route(ctx, route) bind(c, name="st_set_basis_route")
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: route
      integer(c_int) :: st_set_basis_route
    end function
    function st_upload_neighbours(ctx, nb) bind(c, name="st_upload_neighbours")
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx, nb
      integer(c_int) :: st_upload_neighbours
    end function
    function st_upload_atoms(ctx, n_atoms, pair_atom, vec) bind(c, name="st_upload_atoms")
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: ctx, pair_atom, vec
      integer(c_int32_t), value :: n_atoms
      integer(c_int) :: st_upload_atoms
    end function
    function st_upload_dedq(ctx, n_soap, G) bind(c, name="st_upload_dedq")
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: ctx, G
      integer(c_int32_t), value :: n_soap
      integer(c_int) :: st_upload_dedq
    end function
    function st_adjoint_f32(ctx, p, cmp) bind(c, name="st_adjoint_f32")
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx, p, cmp
      integer(c_int) :: st_adjoint_f32
    end function
    function st_download_soap_f32(ctx, soap) bind(c, name="st_download_soap_f32")
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx, soap
      integer(c_int) :: st_download_soap_f32
    end function
    function st_download_forces(ctx, which, forces, virial) bind(c, name="st_download_forces")
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: ctx, forces, virial
      integer(c_int32_t), value :: which
      integer(c_int) :: st_download_forces
    end function
    function st_hessian_selftest(ctx, p, cmp, n_vec, vecs, err) bind(c, name="st_hessian_selftest")
      import :: c_ptr, c_int, c_int32_t
      type(c_ptr), value :: ctx, p, cmp, vecs, err
      integer(c_int32_t), value :: n_vec
      integer(c_int) :: st_hessian_selftest
    end function
  end interface

contains

  ! The library's message for the last failure on ctx, as a Fortran string.
  function st_message(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: n, i
    p = st_error(ctx)
    if( .not. c_associated(p) )then
      msg = ""
      return
    end if
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate( character(len=n) :: msg )
    do i = 1, n
      msg(i:i) = chars(i)
    end do
  end function

end module soap_turbo_c_api
