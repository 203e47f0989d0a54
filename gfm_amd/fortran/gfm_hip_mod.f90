! gfm_hip_mod.f90 -- Fortran interfaces to the C-ABI of include/gfm_hip.h.
module gfm_hip_mod
  use iso_c_binding
  use gfm_types
  implicit none
  integer(c_int), parameter :: GFM_ORDER_REFERENCE = 0, GFM_ORDER_ZEBRA = 1
  integer(c_int), parameter :: GFM_PREC_FP64 = 0, GFM_PREC_MIXED = 1
  integer(c_int), parameter :: GFM_PCORR_REFERENCE = 0, GFM_PCORR_TOLERANCE = 1
  integer(c_int), parameter :: GFM_PGUESS_ZERO = 0, GFM_PGUESS_LEARNED = 1
  interface
    integer(c_int) function gfm_hip_init(ctrl, h, device, ctx) bind(C, name="gfm_hip_init")
      import :: c_int, c_ptr, gfm_ctrl, gfm_host_arrays
      type(gfm_ctrl), intent(in) :: ctrl
      type(gfm_host_arrays), intent(in) :: h
      integer(c_int), value :: device
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function gfm_hip_upload(ctx, ctrl, h) bind(C, name="gfm_hip_upload")
      import :: c_int, c_ptr, gfm_ctrl, gfm_host_arrays
      type(c_ptr), value :: ctx
      type(gfm_ctrl), intent(in) :: ctrl
      type(gfm_host_arrays), intent(in) :: h
    end function
    integer(c_int) function gfm_hip_download(ctx, ctrl, h) bind(C, name="gfm_hip_download")
      import :: c_int, c_ptr, gfm_ctrl, gfm_host_arrays
      type(c_ptr), value :: ctx
      type(gfm_ctrl), intent(inout) :: ctrl
      type(gfm_host_arrays), intent(in) :: h
    end function
    integer(c_int) function gfm_hip_initsv_tail(ctx, ctrl) bind(C, name="gfm_hip_initsv_tail")
      import :: c_int, c_ptr, gfm_ctrl
      type(c_ptr), value :: ctx
      type(gfm_ctrl), intent(inout) :: ctrl
    end function
    integer(c_int) function gfm_hip_gflow(ctx, ctrl) bind(C, name="gfm_hip_gflow")
      import :: c_int, c_ptr, gfm_ctrl
      type(c_ptr), value :: ctx
      type(gfm_ctrl), intent(inout) :: ctrl
    end function
    integer(c_int) function gfm_hip_sphase(ctx, ctrl, max_iters, res_hist, n_done) bind(C, name="gfm_hip_sphase")
      import :: c_int, c_ptr, c_double, gfm_ctrl
      type(c_ptr), value :: ctx
      type(gfm_ctrl), intent(inout) :: ctrl
      integer(c_int), value :: max_iters
      real(c_double), intent(out) :: res_hist(*)
      integer(c_int), intent(out) :: n_done
    end function
    integer(c_int) function gfm_hip_adlbl3(ctx, ctrl, nelg, ap, as, bs, xt, ib, ie, jb, je, kb, ke, resg) &
        bind(C, name="gfm_hip_adlbl3")
      import :: c_int, c_ptr, c_double, gfm_ctrl
      type(c_ptr), value :: ctx
      type(gfm_ctrl), intent(in) :: ctrl
      integer(c_int), value :: nelg, ib, ie, jb, je, kb, ke
      real(c_double), intent(in) :: ap(*), as(*), bs(*)
      real(c_double), intent(inout) :: xt(*)
      real(c_double), intent(out) :: resg(2)
    end function
    integer(c_int) function gfm_hip_calc_residual(ctx, ap, as, bs, xt, ib, ie, jb, je, kb, ke, resid) &
        bind(C, name="gfm_hip_calc_residual")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: ap(*), as(*), bs(*), xt(*)
      integer(c_int), value :: ib, ie, jb, je, kb, ke
      real(c_double), intent(out) :: resid
    end function
    integer(c_int) function gfm_hip_finalize(ctx) bind(C, name="gfm_hip_finalize")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function gfm_hip_set_pguess(ctx, net) bind(C, name="gfm_hip_set_pguess")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, net
    end function
    integer(c_int) function gfm_hip_pcorr_pair(ctx, x, p) bind(C, name="gfm_hip_pcorr_pair")
      import :: c_int, c_ptr, c_float
      type(c_ptr), value :: ctx
      real(c_float), intent(out) :: x(*), p(*)
    end function
    ! FLX1D's flow rates of the last iteration run with want_plane_flows set:
    ! flx(MP,2), fly(NP,2), flyh(NP,2), flux_z(LP,2), fluxh_z(LP,2)
    integer(c_int) function gfm_hip_plane_flows(ctx, flx, fly, flyh, flux_z, fluxh_z) &
        bind(C, name="gfm_hip_plane_flows")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: flx(*), fly(*), flyh(*), flux_z(*), fluxh_z(*)
    end function
    ! minor species (INITSVM, GSLV3M, MSFLOW on the device)
    integer(c_int) function gfm_hip_ms_upload(ctx, ms, a) bind(C, name="gfm_hip_ms_upload")
      import :: c_int, c_ptr, gfm_ms_ctrl, gfm_ms_arrays
      type(c_ptr), value :: ctx
      type(gfm_ms_ctrl), intent(in) :: ms
      type(gfm_ms_arrays), intent(in) :: a
    end function
    integer(c_int) function gfm_hip_ms_download(ctx, ms, a) bind(C, name="gfm_hip_ms_download")
      import :: c_int, c_ptr, gfm_ms_ctrl, gfm_ms_arrays
      type(c_ptr), value :: ctx
      type(gfm_ms_ctrl), intent(inout) :: ms
      type(gfm_ms_arrays), intent(in) :: a
    end function
    integer(c_int) function gfm_hip_initsvm(ctx, ms, restart) bind(C, name="gfm_hip_initsvm")
      import :: c_int, c_ptr, gfm_ms_ctrl
      type(c_ptr), value :: ctx
      type(gfm_ms_ctrl), intent(inout) :: ms
      integer(c_int), value :: restart
    end function
    integer(c_int) function gfm_hip_gslv3m(ctx, ms) bind(C, name="gfm_hip_gslv3m")
      import :: c_int, c_ptr, gfm_ms_ctrl
      type(c_ptr), value :: ctx
      type(gfm_ms_ctrl), intent(inout) :: ms
    end function
    integer(c_int) function gfm_hip_msflow(ctx, ms) bind(C, name="gfm_hip_msflow")
      import :: c_int, c_ptr, gfm_ms_ctrl
      type(c_ptr), value :: ctx
      type(gfm_ms_ctrl), intent(inout) :: ms
    end function
    ! radiation transport (RAD_EMIS's volume part, rad_qe_only) and the wall exchange (QRNF's wall part)
    integer(c_int) function gfm_hip_rad_emis(ctx, rc) bind(C, name="gfm_hip_rad_emis")
      import :: c_int, c_ptr, gfm_rad_ctrl
      type(c_ptr), value :: ctx
      type(gfm_rad_ctrl), intent(inout) :: rc
    end function
    integer(c_int) function gfm_hip_rad_qe_only(ctx, rc) bind(C, name="gfm_hip_rad_qe_only")
      import :: c_int, c_ptr, gfm_rad_ctrl
      type(c_ptr), value :: ctx
      type(gfm_rad_ctrl), intent(inout) :: rc
    end function
    integer(c_int) function gfm_hip_rad_download(ctx, rc, a) bind(C, name="gfm_hip_rad_download")
      import :: c_int, c_ptr, gfm_rad_ctrl, gfm_rad_arrays
      type(c_ptr), value :: ctx
      type(gfm_rad_ctrl), intent(in) :: rc
      type(gfm_rad_arrays), intent(in) :: a
    end function
    integer(c_int) function gfm_hip_rad_upload(ctx, a) bind(C, name="gfm_hip_rad_upload")
      import :: c_int, c_ptr, gfm_rad_arrays
      type(c_ptr), value :: ctx
      type(gfm_rad_arrays), intent(in) :: a
    end function
    integer(c_int) function gfm_hip_wall_setup(ctx, wc, iwall) bind(C, name="gfm_hip_wall_setup")
      import :: c_int, c_ptr, gfm_wall_ctrl
      type(c_ptr), value :: ctx
      type(gfm_wall_ctrl), intent(inout) :: wc
      type(c_ptr), value :: iwall
    end function
    integer(c_int) function gfm_hip_wallq(ctx, wc) bind(C, name="gfm_hip_wallq")
      import :: c_int, c_ptr, gfm_wall_ctrl
      type(c_ptr), value :: ctx
      type(gfm_wall_ctrl), intent(inout) :: wc
    end function
    integer(c_int) function gfm_hip_wall_download(ctx, wc, a) bind(C, name="gfm_hip_wall_download")
      import :: c_int, c_ptr, gfm_wall_ctrl, gfm_wall_arrays
      type(c_ptr), value :: ctx
      type(gfm_wall_ctrl), intent(in) :: wc
      type(gfm_wall_arrays), intent(in) :: a
    end function
    integer(c_int) function gfm_hip_wall_upload(ctx, wc, a) bind(C, name="gfm_hip_wall_upload")
      import :: c_int, c_ptr, gfm_wall_ctrl, gfm_wall_arrays
      type(c_ptr), value :: ctx
      type(gfm_wall_ctrl), intent(in) :: wc
      type(gfm_wall_arrays), intent(in) :: a
    end function
    ! include/gfm_pnet.h: the network of the learned p' guess
    integer(c_int) function gfm_pnet_create(device, ch, max_batch, max_cells, seed, net) bind(C, name="gfm_pnet_create")
      import :: c_int, c_ptr, c_long_long, c_int64_t
      integer(c_int), value :: device, max_batch
      integer(c_int), intent(in) :: ch(4)
      integer(c_long_long), value :: max_cells
      integer(c_int64_t), value :: seed
      type(c_ptr), intent(out) :: net
    end function
    integer(c_int) function gfm_pnet_set_params(net, p, n) bind(C, name="gfm_pnet_set_params")
      import :: c_int, c_ptr, c_float, c_size_t
      type(c_ptr), value :: net
      real(c_float), intent(in) :: p(*)
      integer(c_size_t), value :: n
    end function
    integer(c_int) function gfm_pnet_param_count(net, n_total, n_effective) bind(C, name="gfm_pnet_param_count")
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: net
      integer(c_size_t), intent(out) :: n_total, n_effective
    end function
    integer(c_int) function gfm_pnet_destroy(net) bind(C, name="gfm_pnet_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: net
    end function
    type(c_ptr) function gfm_hip_last_error() bind(C, name="gfm_hip_last_error")
      import :: c_ptr
    end function
  end interface
contains
  function gfm_error_text() result(msg)
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: p(:)
    integer :: n
    call c_f_pointer(gfm_hip_last_error(), p, [512])
    n = 0
    do while (n < 512)
      if (p(n + 1) == c_null_char) exit
      n = n + 1
    end do
    allocate (character(len=n) :: msg)
    msg = transfer(p(1:n), msg)
  end function
end module gfm_hip_mod
