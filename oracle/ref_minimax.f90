! Checker driver, test infrastructure only: runs the reference's vendored greenX
! minimax generator gx_minimax_grid (thirdparty/greenX/GX-TimeFrequency/src/
! minimax_grids.F90:53-237), compiled from its sources, so that tests can hold
! the oracle's grids and transform matrices against it for any size and energy
! range. LibRPA calls it through src/interface/gx_minimax_wrp.f90:94-180 from
! TFGrids::generate_minimax (src/core/timefreq.cpp:193-219): bare weights off,
! the given regularization, no sine omega->tau matrix.
!
! Input and output are little-endian unformatted streams (int32, float64):
!   in : count, then per case {n (int32), e_min, e_max, regularization}
!   out: per case {ierr (int32); when ierr == 0: tau_points(n), tau_weights(n),
!        omega_points(n), omega_weights(n), cosft_wt(n,n), cosft_tw(n,n),
!        sinft_wt(n,n) in Fortran (column-major) order, max_errors(3),
!        cosft_duality_error}
!
! Usage: ref_minimax <in> <out>

program ref_minimax
  use kinds, only: dp
  use minimax_grids, only: gx_minimax_grid
  use iso_fortran_env, only: int32
  implicit none

  character(len=4096) :: in_path, out_path
  integer :: u_in, u_out, icase
  integer(int32) :: count, n32, ierr32
  integer :: n, ierr
  real(dp) :: e_min, e_max, reg, max_errors(3), duality
  real(dp), allocatable :: tau_p(:), tau_w(:), om_p(:), om_w(:)
  real(dp), allocatable :: cos_wt(:, :), cos_tw(:, :), sin_wt(:, :)

  if (command_argument_count() /= 2) then
    write (*, '(a)') 'usage: ref_minimax <in> <out>'
    stop 2
  end if
  call get_command_argument(1, in_path)
  call get_command_argument(2, out_path)
  open (newunit=u_in, file=trim(in_path), access='stream', form='unformatted', status='old')
  open (newunit=u_out, file=trim(out_path), access='stream', form='unformatted', status='replace')

  read (u_in) count
  do icase = 1, count
    read (u_in) n32, e_min, e_max, reg
    n = n32
    call gx_minimax_grid(n, e_min, e_max, tau_p, tau_w, om_p, om_w, &
                         cos_wt, cos_tw, sin_wt, max_errors, duality, ierr, &
                         .false., reg)
    ierr32 = ierr
    write (u_out) ierr32
    if (ierr == 0) then
      write (u_out) tau_p, tau_w, om_p, om_w, cos_wt, cos_tw, sin_wt, max_errors, duality
    end if
    if (allocated(tau_p)) deallocate (tau_p)
    if (allocated(tau_w)) deallocate (tau_w)
    if (allocated(om_p)) deallocate (om_p)
    if (allocated(om_w)) deallocate (om_w)
    if (allocated(cos_wt)) deallocate (cos_wt)
    if (allocated(cos_tw)) deallocate (cos_tw)
    if (allocated(sin_wt)) deallocate (sin_wt)
  end do
  close (u_in)
  close (u_out)
end program ref_minimax
