! gfm_comb.f90 -- Fortran driver of the MI355X combustion-space loop.
!
! Reads a case state (reference layout) from a stream file, runs SPHASE on
! the GPU for the requested iterations, writes the state back and prints
! the res_mass history (the convg log line of C/Sphase.f90:53).  With a
! fifth argument it also writes the restart record rg####c.d there (SAV3F),
! with a sixth the convg log of every iteration (convg####c.plt, header of
! C/Utility.f90:1066-1080, record of C/Sphase.f90:53).
!   usage: gfm_comb <state-in> <state-out> <iterations> [sweep_order [rg-out [convg-out]]]
! The staged form runs one ADLBL3 on a system read from a file, called as
! the reference calls it (C/Gslv3v.f90:91-94: calc_residual, ADLBL3,
! calc_residual, implicit interfaces, NELG from module state):
!   gfm_comb adlbl3 <state-in> <system-in> <out>
! and the minor-species solve (INITSVM, MSFLOW, SAV3FM):
!   gfm_comb msflow <state-in> <ms-ctrl-in> <rs-out> [rs-in]
! system-in (stream): int32 NELG, int32 IB,IE,JB,JE,KB,KE, then AP, AS, BS
! and XT in the reference layout; out: the two residuals, then XT.
program gfm_comb
  use iso_c_binding
  use gfm_types
  use gfm_state
  use gfm_solver
  implicit none
  character(len=512) :: fin, fout, arg
  integer :: niter, order, i
  call get_command_argument(1, arg)
  if (trim(arg) == 'adlbl3') then
    call staged_adlbl3()
    stop
  end if
  if (trim(arg) == 'msflow') then
    call minor_species()
    stop
  end if
  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  call get_command_argument(3, arg)
  read (arg, *) niter
  order = -1
  if (command_argument_count() >= 4) then
    call get_command_argument(4, arg)
    read (arg, *) order
  end if
  call read_state(trim(fin))
  if (order >= 0) ctrl%sweep_order = order
  ctrl%itr_gend = ctrl%itr_gas + niter
  poll_every = max(1, min(100, niter))
  if (command_argument_count() >= 6) then
    call get_command_argument(6, arg)
    open (newunit=nug, file=trim(arg), status='replace')
    log_convg = .true.
    write (nug, *) '# Combustion space mean and max mass residual - log plot'
    write (nug, *) '# PlotData'
    write (nug, *) '# iteration  average mass residual  max mass residual'
    poll_every = 1
  end if
  call SPHASE
  if (log_convg) close (nug)
  do i = 1, n_hist
    write (*, '(i7,es26.17e3)') ctrl%itr_gas - n_hist + i, res_hist(i)
  end do
  call write_state(trim(fout))
  if (command_argument_count() >= 5) then
    call get_command_argument(5, arg)
    if (len_trim(arg) > 0) call SAV3F(trim(arg))
  end if
  call gpu_close()
contains
  subroutine staged_adlbl3()
    character(len=512) :: fstate, fsys, fo
    integer(c_int32_t) :: hdr(7)
    integer :: u, ib, ie, jb, je, kb, ke
    real(c_double) :: r0, r1
    real(c_double), allocatable :: xt(:, :, :)
    call get_command_argument(2, fstate)
    call get_command_argument(3, fsys)
    call get_command_argument(4, fo)
    call read_state(trim(fstate))
    allocate (AP(MP, NP, LP), AS(MP, NP, LP, 3, 2), BS(MP, NP, LP), xt(MP, NP, LP))
    open (newunit=u, file=trim(fsys), access='stream', form='unformatted', status='old', action='read')
    read (u) hdr
    read (u) AP, AS, BS, xt
    close (u)
    NELG = hdr(1)
    ib = hdr(2); ie = hdr(3); jb = hdr(4); je = hdr(5); kb = hdr(6); ke = hdr(7)
    call calc_residual(xt, r0, ib, ie, jb, je, kb, ke)
    call ADLBL3(xt, ib, ie, jb, je, kb, ke)
    call calc_residual(xt, r1, ib, ie, jb, je, kb, ke)
    open (newunit=u, file=trim(fo), access='stream', form='unformatted', status='replace', action='write')
    write (u) r0, r1, xt
    close (u)
    call gpu_close()
  end subroutine

  ! the minor-species solve as the main program runs it after SPHASE
  ! (C/A0main.f90:219-222): INITSVM, MSFLOW, SAV3FM
  !   gfm_comb msflow <state-in> <ms-ctrl-in> <rs-out> [rs-in]
  ! ms-ctrl-in (stream): the gfm_ms_ctrl block; with rs-in INITSVM restarts
  ! from that record (IRSTYPM = 1).  Prints itr_ms, the residuals and bal_ms.
  subroutine minor_species()
    character(len=512) :: fstate, fms, fo
    integer :: u, l
    call get_command_argument(2, fstate)
    call get_command_argument(3, fms)
    call get_command_argument(4, fo)
    call read_state(trim(fstate))
    open (newunit=u, file=trim(fms), access='stream', form='unformatted', status='old', action='read')
    read (u) msc
    close (u)
    if (command_argument_count() >= 5) then
      call get_command_argument(5, rs_file)
      IRSTYPM = 1
    end if
    call INITSVM
    call MSFLOW
    call SAV3FM(trim(fo))
    write (*, '(a,i8,a,i5)') 'itr_ms', msc%itr_ms, ' iterations', msc%iterations
    write (*, '(a,8es26.17e3)') 'resid_ms', (msc%resid_ms(l), l = 1, 8)
    write (*, '(a,es26.17e3)') 'bal_ms', msc%bal_ms
    call gpu_close()
  end subroutine

  subroutine read_state(path)
    character(len=*), intent(in) :: path
    character(len=8) :: magic
    integer(c_int32_t) :: ver, csize, m, n, l, nr
    integer :: u
    open (newunit=u, file=path, access='stream', form='unformatted', status='old', action='read')
    read (u) magic, ver, csize
    if (magic /= 'GFMSTATE' .or. csize /= storage_size(ctrl) / 8) call stop_run('bad state file '//path)
    read (u) ctrl
    m = ctrl%mp; n = ctrl%np; l = ctrl%lp; nr = ctrl%nr0
    call allocate_state(m, n, l, nr)
    read (u) X, DX, R, DR, RS, Z, DZ, IBCELL
    read (u) P, T, DNST, THETA, TMU, GCP, GLAM, GAMA, GTE, CRATE, HRATE, BS_HPREV, UG, GF
    read (u) RTIM, DELRT, HRT, DELR, TIMED, HRD
    close (u)
  end subroutine
  subroutine write_state(path)
    character(len=*), intent(in) :: path
    integer :: u
    open (newunit=u, file=path, access='stream', form='unformatted', status='replace', action='write')
    write (u) 'GFMSTATE', int(1, c_int32_t), int(storage_size(ctrl) / 8, c_int32_t)
    write (u) ctrl
    write (u) X, DX, R, DR, RS, Z, DZ, IBCELL
    write (u) P, T, DNST, THETA, TMU, GCP, GLAM, GAMA, GTE, CRATE, HRATE, BS_HPREV, UG, GF
    write (u) RTIM, DELRT, HRT, DELR, TIMED, HRD
    close (u)
  end subroutine
end program gfm_comb
