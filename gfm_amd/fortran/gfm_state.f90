! gfm_state.f90 -- Fortran host layer over the MI355X SIMPLER library.
!
! Module gfm_state holds the subset of the reference's MODULE GBL_VAR
! (C/Var_mod.f90:3-889) that the combustion-space loop reads and writes,
! in the reference layout (1-based, column-major, doubled lattice).
! Module gfm_solver runs the loop on the GPU through ISO_C_BINDING
! (gfm_hip_mod); gfm_external.f90 puts the reference's subroutine interface
! on it as external subroutines, so the reference's implicit-interface call
! sites link unchanged:
!   SPHASE                               C/Sphase.f90:10   (called C/A0main.f90:204)
!   GFLOW                                C/Gflow.f90:6     (called C/Sphase.f90:51)
!   ADLBL3(XT,IB,IE,JB,JE,KB,KE)         C/Adlbl3.f90:10   (called C/Gslv3v.f90:93,133,163, C/Gslv3g.f90:96)
!   calc_residual(xt,resid,ib,ie,jb,je,kb,ke) C/Adlbl3.f90:239 (called C/Gslv3v.f90:91,94,132,135,164)
!   INITSVM, MSFLOW, GSLV3M              C/Initsv.f90:329, C/Msflow.f90:10, C/Gslv3m.f90:8
!                                        (called C/A0main.f90:220-221, C/Sphase.f90:216-217)
! A non-zero return from the library ends the run through stop_run, as the
! reference's fatal paths do (C/A0main.f90:415-424).
module gfm_state
  use iso_c_binding
  use gfm_types
  implicit none
  integer :: MP, NP, LP, NR0
  real(c_double), allocatable, target :: X(:), DX(:), R(:), DR(:), RS(:), Z(:), DZ(:)
  integer(c_int16_t), allocatable, target :: IBCELL(:, :, :)
  real(c_double), allocatable, target :: P(:, :, :), T(:, :, :), DNST(:, :, :), THETA(:, :, :), TMU(:, :, :)
  real(c_double), allocatable, target :: GCP(:, :, :), GLAM(:, :, :), GAMA(:, :, :), GTE(:, :, :)
  real(c_double), allocatable, target :: CRATE(:, :, :), HRATE(:, :, :), BS_HPREV(:, :, :)
  real(c_double), allocatable, target :: UG(:, :, :, :), GF(:, :, :, :)
  real(c_double), allocatable, target :: RTIM(:), DELRT(:), HRT(:), DELR(:), TIMED(:), HRD(:)
  ! ADLBL3's module inputs (C/Adlbl3.f90 reads AP, AS, BS and the equation
  ! index NELG from GBL_VAR, C/Var_mod.f90:567)
  real(c_double), allocatable, target :: AP(:, :, :), AS(:, :, :, :, :), BS(:, :, :)
  integer :: NELG = 0
  type(gfm_ctrl), target :: ctrl
  type(gfm_host_arrays), target :: harr
  character(len=256) :: casedir = '.'
  ! restart-record state the combustion-space loop does not use
  ! (C/Var_mod.f90:455,468,812,866; C/Setup.f90:755-756)
  real(c_double), allocatable :: T_SURF_PREV(:, :)
  integer :: i1st_rad_qe_done = 0, itr_rad = 0, interval_rad = 100
  integer, parameter :: LEND = 7
  ! convg log (C/Sphase.f90:53), written by SPHASE on unit nug when log_convg
  integer :: nug = 0
  logical :: log_convg = .false.
  ! minor species (MS = 1; C/Var_mod.f90:218-235,824-848): GFM(:,:,:,1..7) =
  ! CH4, O2, CO, H2O, CO2, N2, NO, the soot mass fraction, the soot-source
  ! history and the reaction rates; msc carries MAXMS, BMCON, the relaxation
  ! factors and the soot constants in and itr_ms and the residuals out
  integer, parameter :: NSP0 = 7, NRM0 = 5
  real(c_double), allocatable, target :: GFM(:, :, :, :), SMF(:, :, :), BS_SOOT_PREV(:, :, :), SFP_SOOT_PREV(:, :, :)
  real(c_double), allocatable, target :: RRMS(:, :, :, :)
  type(gfm_ms_ctrl), target :: msc
  type(gfm_ms_arrays), target :: marr
  integer :: IRSTYPM = 0                  ! 1: INITSVM restarts from rs_file when it exists
  character(len=512) :: rs_file = 'rs.d'  ! the rs####c.d record INITSVM reads and SAV3FM writes
contains
  subroutine allocate_state(m, n, l, nr)
    integer, intent(in) :: m, n, l, nr
    MP = m; NP = n; LP = l; NR0 = nr
    allocate (X(MP), DX(MP), R(NP), DR(NP), RS(NP), Z(LP), DZ(LP))
    allocate (IBCELL(MP, NP, LP))
    allocate (P(MP, NP, LP), T(MP, NP, LP), DNST(MP, NP, LP), THETA(MP, NP, LP), TMU(MP, NP, LP))
    allocate (GCP(MP, NP, LP), GLAM(MP, NP, LP), GAMA(MP, NP, LP), GTE(MP, NP, LP))
    allocate (CRATE(MP, NP, LP), HRATE(MP, NP, LP), BS_HPREV(MP, NP, LP))
    allocate (UG(MP, NP, LP, 3), GF(MP, NP, LP, 9))
    allocate (RTIM(NR0), DELRT(NR0), HRT(NR0), DELR(NR0), TIMED(NR0), HRD(NR0))
    allocate (T_SURF_PREV(MP / 2, NP / 2))
    T_SURF_PREV = 1500.0d0
  end subroutine

  ! SAV3F: the restart record rg####c.d of the combustion space
  ! (C/Sav3f.f90:21-88, single phase): per cell centre (K innermost) the
  ! record P, U(I+1), V(J+1), W(K+1), T, THETA (face index MP-1, NP-1, LP-1
  ! on the last layer) and the record GF(:,:,:,1:LEND); then bs_hprev,
  ! T_surf_prev, (FACTOR, FADJ), the radiation flags, itr_gas,
  ! (itr_rad, itr_gend), interval_rad.
  subroutine SAV3F(filename)
    character(len=*), intent(in) :: filename
    integer :: u, i, j, k, iu, ju, ku, l
    integer(c_int32_t) :: qe, rad, itg, itr2(2), irad
    real(c_double) :: rec6(6), ff(2)
    open (newunit=u, file=filename, form='unformatted', access='sequential', status='replace')
    do i = 2, MP, 2
      iu = merge(MP - 1, i + 1, i == MP)
      do j = 2, NP, 2
        ju = merge(NP - 1, j + 1, j == NP)
        do k = 2, LP, 2
          ku = merge(LP - 1, k + 1, k == LP)
          rec6 = [P(i, j, k), UG(iu, j, k, 1), UG(i, ju, k, 2), UG(i, j, ku, 3), T(i, j, k), THETA(i, j, k)]
          write (u) rec6
          write (u) (GF(i, j, k, l), l = 1, LEND)
        end do
      end do
    end do
    write (u) BS_HPREV
    write (u) T_SURF_PREV
    ff = [ctrl%factor, ctrl%fadj]
    write (u) ff
    qe = i1st_rad_qe_done; rad = ctrl%i1st_rad_done; itg = ctrl%itr_gas
    itr2 = [int(itr_rad, c_int32_t), int(ctrl%itr_gend, c_int32_t)]
    irad = interval_rad
    write (u) qe
    write (u) rad
    write (u) itg
    write (u) itr2
    write (u) irad
    close (u)
  end subroutine

  subroutine allocate_ms_state()
    if (allocated(GFM)) return
    allocate (GFM(MP, NP, LP, NSP0), SMF(MP, NP, LP), BS_SOOT_PREV(MP, NP, LP), SFP_SOOT_PREV(MP, NP, LP))
    allocate (RRMS(MP / 2, NP / 2, LP / 2, NRM0))
    GFM = 0.0d0; SMF = 0.0d0; BS_SOOT_PREV = 0.0d0; SFP_SOOT_PREV = 0.0d0; RRMS = 0.0d0
    marr%gfm = c_loc(GFM); marr%smf = c_loc(SMF); marr%bs_soot_prev = c_loc(BS_SOOT_PREV)
    marr%sfp_soot_prev = c_loc(SFP_SOOT_PREV); marr%rrms = c_loc(RRMS)
  end subroutine

  ! SAV3FM: the minor-species restart record rs####c.d (C/Sav3f.f90:96-116):
  ! per cell centre (K innermost) GFM(1:NSP0) and smf, then bs_soot_prev,
  ! sfp_soot_prev, RRMS and itr_ms
  subroutine SAV3FM(filename)
    character(len=*), intent(in) :: filename
    integer :: u, i, j, k, l
    integer(c_int32_t) :: itr
    open (newunit=u, file=filename, form='unformatted', access='sequential', status='replace')
    do i = 2, MP, 2
      do j = 2, NP, 2
        do k = 2, LP, 2
          write (u) (GFM(i, j, k, l), l = 1, NSP0), SMF(i, j, k)
        end do
      end do
    end do
    write (u) BS_SOOT_PREV
    write (u) SFP_SOOT_PREV
    write (u) RRMS
    itr = msc%itr_ms
    write (u) itr
    close (u)
  end subroutine

  ! the reads of INITSVM's restart branch (C/Initsv.f90:352-372)
  subroutine read_rs(filename)
    character(len=*), intent(in) :: filename
    integer :: u, i, j, k, l
    integer(c_int32_t) :: itr
    open (newunit=u, file=filename, form='unformatted', access='sequential', status='old', action='read')
    do i = 2, MP, 2
      do j = 2, NP, 2
        do k = 2, LP, 2
          read (u) (GFM(i, j, k, l), l = 1, NSP0), SMF(i, j, k)
        end do
      end do
    end do
    read (u) BS_SOOT_PREV
    read (u) SFP_SOOT_PREV
    read (u) RRMS
    read (u) itr
    msc%itr_ms = itr
    close (u)
  end subroutine

  subroutine bind_host_arrays()
    harr%x = c_loc(X); harr%dx = c_loc(DX); harr%r = c_loc(R); harr%dr = c_loc(DR); harr%rs = c_loc(RS)
    harr%z = c_loc(Z); harr%dz = c_loc(DZ); harr%ibcell = c_loc(IBCELL)
    harr%p = c_loc(P); harr%t = c_loc(T); harr%dnst = c_loc(DNST); harr%theta = c_loc(THETA)
    harr%tmu = c_loc(TMU); harr%gcp = c_loc(GCP); harr%glam = c_loc(GLAM); harr%gama = c_loc(GAMA)
    harr%gte = c_loc(GTE); harr%crate = c_loc(CRATE); harr%hrate = c_loc(HRATE); harr%bs_hprev = c_loc(BS_HPREV)
    harr%ug = c_loc(UG); harr%gf = c_loc(GF)
    harr%rtim = c_loc(RTIM); harr%delrt = c_loc(DELRT); harr%hrt = c_loc(HRT)
    harr%delr = c_loc(DELR); harr%timed = c_loc(TIMED); harr%hrd = c_loc(HRD)
  end subroutine

  ! stop_run: record why the run ended and stop (C/A0main.f90:415-424)
  subroutine stop_run(reason)
    character(len=*), intent(in) :: reason
    integer :: u
    open (newunit=u, file=trim(casedir)//'/runend.txt', status='replace')
    write (u, '(a)') trim(reason)
    close (u)
    write (*, '(a)') trim(reason)
    stop 1
  end subroutine
end module gfm_state

module gfm_solver
  use iso_c_binding
  use gfm_types
  use gfm_hip_mod
  use gfm_state
  implicit none
  type(c_ptr) :: gpu_ctx = c_null_ptr
  integer :: poll_every = 100              ! iterations between host polls (runstop.dat)
  real(c_double), allocatable :: res_hist(:)
  integer :: n_hist = 0
contains
  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= 0) call stop_run('GPU: '//what//' failed: '//gfm_error_text())
  end subroutine

  ! device state for the module arrays; INITSV's tail runs on the device
  subroutine gpu_open(device)
    integer, intent(in) :: device
    call bind_host_arrays()
    call check(gfm_hip_init(ctrl, harr, int(device, c_int), gpu_ctx), 'gfm_hip_init')
    call check(gfm_hip_initsv_tail(gpu_ctx, ctrl), 'gfm_hip_initsv_tail')
  end subroutine

  subroutine gpu_sync_to_host()
    call check(gfm_hip_download(gpu_ctx, ctrl, harr), 'gfm_hip_download')
  end subroutine

  subroutine gpu_close()
    integer(c_int) :: rc
    if (c_associated(gpu_ctx)) rc = gfm_hip_finalize(gpu_ctx)
    gpu_ctx = c_null_ptr
  end subroutine

  ! the combustion-space loop (C/Sphase.f90:37-235) on the device,
  ! returning to the host every poll_every iterations to honour runstop.dat
  ! (C/Sphase.f90:125-131).  The module arrays are current on return.
  subroutine gpu_sphase
    integer(c_int) :: n_done
    real(c_double), allocatable :: chunk(:)
    logical :: stop_req
    if (.not. c_associated(gpu_ctx)) call gpu_open(0)
    allocate (chunk(poll_every))
    if (.not. allocated(res_hist)) allocate (res_hist(max(ctrl%maxgi, 1)))
    do while (ctrl%itr_gas < ctrl%itr_gend .and. ctrl%itr_gas < ctrl%maxgi)
      call check(gfm_hip_sphase(gpu_ctx, ctrl, int(poll_every, c_int), chunk, n_done), 'gfm_hip_sphase')
      if (n_done > 0 .and. n_hist + n_done <= size(res_hist)) then
        res_hist(n_hist + 1:n_hist + n_done) = chunk(1:n_done)
        n_hist = n_hist + n_done
      end if
      ! the convg line of every iteration (needs poll_every = 1 for the
      ! per-iteration maximum)
      if (log_convg .and. n_done == 1) write (nug, '(i7,4e25.16)') ctrl%itr_gas, ctrl%res_mass, ctrl%res_mass_max
      if (n_done == 0) exit
      inquire (file=trim(casedir)//'/runstop.dat', exist=stop_req)
      if (stop_req) exit
    end do
    call gpu_sync_to_host()
  end subroutine

  ! one SIMPLER iteration (C/Gflow.f90:6-408) at itr_gas
  subroutine gpu_gflow
    if (.not. c_associated(gpu_ctx)) call gpu_open(0)
    call check(gfm_hip_gflow(gpu_ctx, ctrl), 'gfm_hip_gflow')
  end subroutine

  ! line relaxation of XT (C/Adlbl3.f90:10) with the module's AP, AS, BS
  ! and NTIMESG(NELG) sweeps; RESG(NELG,1:2) back in ctrl%resg
  subroutine gpu_adlbl3(XT, IB, IE, JB, JE, KB, KE)
    real(c_double), intent(inout) :: XT(MP, NP, LP)
    integer, intent(in) :: IB, IE, JB, JE, KB, KE
    real(c_double) :: resg(2)
    if (.not. c_associated(gpu_ctx)) call gpu_open(0)
    call check(gfm_hip_adlbl3(gpu_ctx, ctrl, int(NELG, c_int), AP, AS, BS, XT, int(IB, c_int), int(IE, c_int), &
                              int(JB, c_int), int(JE, c_int), int(KB, c_int), int(KE, c_int), resg), 'gfm_hip_adlbl3')
    if (ctrl%want_irc /= 0) then
      ctrl%resg(NELG) = resg(1)
      ctrl%resg(20 + NELG) = resg(2)
    end if
  end subroutine

  ! INITSVM (C/Initsv.f90:329-380) on the device: a restart from rs_file
  ! when IRSTYPM = 1 and the file exists, else a new start from the gas field
  subroutine gpu_initsvm
    logical :: exst
    integer(c_int) :: restart
    if (.not. c_associated(gpu_ctx)) call gpu_open(0)
    call allocate_ms_state()
    inquire (file=trim(rs_file), exist=exst)
    restart = 0
    if (IRSTYPM /= 0 .and. exst) then
      call read_rs(trim(rs_file))
      restart = 1
    end if
    call check(gfm_hip_ms_upload(gpu_ctx, msc, marr), 'gfm_hip_ms_upload')
    call check(gfm_hip_initsvm(gpu_ctx, msc, restart), 'gfm_hip_initsvm')
    call check(gfm_hip_ms_download(gpu_ctx, msc, marr), 'gfm_hip_ms_download')
  end subroutine

  ! GSLV3M (C/Gslv3m.f90:8-286) and MSFLOW (C/Msflow.f90:10-169) on the
  ! device state INITSVM set up; the module arrays are current on return
  subroutine gpu_gslv3m
    if (.not. allocated(GFM)) call gpu_initsvm()
    call check(gfm_hip_gslv3m(gpu_ctx, msc), 'gfm_hip_gslv3m')
    call check(gfm_hip_ms_download(gpu_ctx, msc, marr), 'gfm_hip_ms_download')
  end subroutine
  subroutine gpu_msflow
    if (.not. allocated(GFM)) call gpu_initsvm()
    call check(gfm_hip_msflow(gpu_ctx, msc), 'gfm_hip_msflow')
    call check(gfm_hip_ms_download(gpu_ctx, msc, marr), 'gfm_hip_ms_download')
  end subroutine

  ! calc_residual (C/Adlbl3.f90:239-271) of XT with the module's AP, AS, BS
  subroutine gpu_calc_residual(XT, RESID, IB, IE, JB, JE, KB, KE)
    real(c_double), intent(in) :: XT(MP, NP, LP)
    real(c_double), intent(out) :: RESID
    integer, intent(in) :: IB, IE, JB, JE, KB, KE
    if (.not. c_associated(gpu_ctx)) call gpu_open(0)
    call check(gfm_hip_calc_residual(gpu_ctx, AP, AS, BS, XT, int(IB, c_int), int(IE, c_int), int(JB, c_int), &
                                     int(JE, c_int), int(KB, c_int), int(KE, c_int), RESID), 'gfm_hip_calc_residual')
  end subroutine
end module gfm_solver
