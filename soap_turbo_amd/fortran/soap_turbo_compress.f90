! get_compress_indices with the reference's interface
! (src/soap_turbo_compress.f90:44-54), over the library's C-ABI: the
! "trivial"/"Trivial" mode and the nu_R_nu_S modes "0_0" .. "2_2", with the
! reference's messages where it stops.
module soap_turbo_compress_module
  use, intrinsic :: iso_fortran_env, only: real64
  use, intrinsic :: iso_c_binding
  use soap_turbo_c_api
  implicit none
  integer, parameter, private :: dp = real64

contains

  subroutine get_compress_indices( compress_mode, alpha_max, l_max, dim, P_nonzero, P_i, P_j, P_el, what_to_do )
    implicit none
    integer, intent(in) :: alpha_max(:), l_max
    character(*), intent(in) :: compress_mode, what_to_do
    real(dp), intent(inout) :: P_el(:)
    integer, intent(inout) :: dim, P_nonzero, P_i(:), P_j(:)

    integer(c_int32_t), allocatable, target :: am(:), pi(:), pj(:)
    real(c_double), allocatable, target :: pel(:)
    integer(c_int32_t), target :: cdim, cnz
    integer :: rc, set, n
    character(len=:), allocatable :: mode

    if( what_to_do == "get_dim" )then
      set = 0
    else if( what_to_do == "set_indices" )then
      set = 1
    else
      write(*,*) "ERROR: incorrect what_to_do in get_compress_indices() subroutine!"
      stop
    end if
    mode = trim(compress_mode)
    allocate( am(size(alpha_max)) )
    am = alpha_max
    ! sizes first, so that set_indices writes only what the caller allocated
    rc = st_get_compress_indices(mode // c_null_char, int(size(am), c_int32_t), c_loc(am), int(l_max, c_int32_t), &
                                 c_loc(cdim), c_loc(cnz), c_null_ptr, c_null_ptr, c_null_ptr, 0_c_int32_t)
    if( rc == ST_ERR_STOP )then
      if( len(mode) == 3 .and. verify(mode(1:1), "012") == 0 .and. mode(2:2) == "_" .and. &
          verify(mode(3:3), "012") == 0 )then
        write(*,'(A,A,A)') "ERROR: compress_mode=", trim(adjustl(compress_mode)), &
                           " requires all species to have the same number of radial basis functions!"
      else
        write(*,*) "ERROR: I don't understand compress_mode =", compress_mode
      end if
      stop
    end if
    if( rc /= ST_OK )then
      write(*,*) "ERROR: get_compress_indices failed"
      stop
    end if
    if( set == 1 )then
      n = int(cnz)
      if( size(P_i) < n .or. size(P_j) < n .or. size(P_el) < n )then
        write(*,*) "ERROR: get_compress_indices: P_i, P_j and P_el need P_nonzero entries"
        stop
      end if
      allocate( pi(n), pj(n), pel(n) )
      P_i = 0
      P_j = 0
      P_el = 0.d0
      rc = st_get_compress_indices(mode // c_null_char, int(size(am), c_int32_t), c_loc(am), int(l_max, c_int32_t), &
                                   c_loc(cdim), c_loc(cnz), c_loc(pi), c_loc(pj), c_loc(pel), 1_c_int32_t)
      P_i(1:n) = pi
      P_j(1:n) = pj
      P_el(1:n) = pel
    end if
    dim = int(cdim)
    P_nonzero = int(cnz)
  end subroutine get_compress_indices

end module soap_turbo_compress_module
