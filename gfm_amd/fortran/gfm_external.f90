! gfm_external.f90 -- the reference's subroutine interface on the GPU path.
!
! External subroutines (outside any module) with the reference's names and
! argument lists, so that its implicit-interface call sites link to them
! unchanged; each forwards to gfm_solver.
!   SPHASE                                    C/Sphase.f90:10, called C/A0main.f90:204
!   GFLOW                                     C/Gflow.f90:6, called C/Sphase.f90:51
!   ADLBL3(XT,IB,IE,JB,JE,KB,KE)              C/Adlbl3.f90:10, called C/Gslv3v.f90:93,133,163, C/Gslv3g.f90:96;
!                                             the equation index NELG comes from module state as in GBL_VAR
!   calc_residual(xt,resid,ib,ie,jb,je,kb,ke) C/Adlbl3.f90:239, called C/Gslv3v.f90:91,94,132,135,164
!   INITSVM                                   C/Initsv.f90:329, called C/A0main.f90:191,220, C/Sphase.f90:216
!   MSFLOW                                    C/Msflow.f90:10, called C/A0main.f90:221, C/Sphase.f90:217
!   GSLV3M                                    C/Gslv3m.f90:8, called C/Msflow.f90:51
subroutine SPHASE
  use gfm_solver, only: gpu_sphase
  implicit none
  call gpu_sphase()
end subroutine SPHASE

subroutine GFLOW
  use gfm_solver, only: gpu_gflow
  implicit none
  call gpu_gflow()
end subroutine GFLOW

subroutine ADLBL3(XT, IB, IE, JB, JE, KB, KE)
  use gfm_state, only: MP, NP, LP
  use gfm_solver, only: gpu_adlbl3
  implicit none
  integer :: IB, IE, JB, JE, KB, KE
  double precision :: XT(MP, NP, LP)
  call gpu_adlbl3(XT, IB, IE, JB, JE, KB, KE)
end subroutine ADLBL3

subroutine calc_residual(xt, resid, ib, ie, jb, je, kb, ke)
  use gfm_state, only: MP, NP, LP
  use gfm_solver, only: gpu_calc_residual
  implicit none
  integer :: ib, ie, jb, je, kb, ke
  double precision :: xt(MP, NP, LP), resid
  call gpu_calc_residual(xt, resid, ib, ie, jb, je, kb, ke)
end subroutine calc_residual

subroutine INITSVM
  use gfm_solver, only: gpu_initsvm
  implicit none
  call gpu_initsvm()
end subroutine INITSVM

subroutine MSFLOW
  use gfm_solver, only: gpu_msflow
  implicit none
  call gpu_msflow()
end subroutine MSFLOW

subroutine GSLV3M
  use gfm_solver, only: gpu_gslv3m
  implicit none
  call gpu_gslv3m()
end subroutine GSLV3M
