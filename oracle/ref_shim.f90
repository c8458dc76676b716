! C-callable entry points into the reference's own pair routines, compiled
! from the reference sources where they lie (oracle/Makefile, target ref):
! module globvars (src/globvars.f90) and module matelem (src/matelem.f90).
! Test infrastructure only: the tests compare the oracle's restatement and
! the device with what the reference's own code computes.
!
! ref_setup fills the globals MatrixElements and MatrixElementsForExpcVals
! read, the way ProgramDataInit does (src/workproc.f90:76, 604-621):
! Glob_n, Glob_np, Glob_npt, Glob_2raised3n2, Glob_Mass, Glob_MassMatrix,
! Glob_bvc, Glob_PseudoCharge0, Glob_PseudoCharge, Glob_CurrEnergy.
module ref_shim
  use, intrinsic :: iso_c_binding
  use globvars
  use matelem
  implicit none
contains

  subroutine ref_setup(n, masses, q0, q, curr_energy) bind(C, name="ref_setup")
    integer(c_int), value :: n
    real(c_double), intent(in) :: masses(n + 1), q(n)
    real(c_double), value :: q0, curr_energy
    integer :: i, npart
    real(dprec) :: mtot
    npart = n + 1
    Glob_n = n
    Glob_np = n * (n + 1) / 2
    Glob_npt = 2 * Glob_np
    Glob_2raised3n2 = 2.0_dprec**((3 * Glob_n) / 2.0_dprec)
    if (allocated(Glob_Mass)) deallocate(Glob_Mass)
    if (allocated(Glob_MassMatrix)) deallocate(Glob_MassMatrix)
    if (allocated(Glob_bvc)) deallocate(Glob_bvc)
    if (allocated(Glob_PseudoCharge)) deallocate(Glob_PseudoCharge)
    allocate(Glob_Mass(npart), Glob_MassMatrix(n, n), Glob_bvc(n, npart), Glob_PseudoCharge(n))
    Glob_Mass(1:npart) = masses(1:npart)
    Glob_MassMatrix(1:n, 1:n) = 0.5_dprec / Glob_Mass(1)
    do i = 1, n
      Glob_MassMatrix(i, i) = Glob_MassMatrix(i, i) + 0.5_dprec / Glob_Mass(i + 1)
    end do
    mtot = sum(Glob_Mass(1:npart))
    do i = 1, npart
      Glob_bvc(1:n, i) = -Glob_Mass(2:n + 1) / mtot
    end do
    do i = 2, npart
      Glob_bvc(i - 1, i) = Glob_bvc(i - 1, i) + 1.0_dprec
    end do
    Glob_PseudoCharge0 = q0
    Glob_PseudoCharge(1:n) = q(1:n)
    Glob_CurrEnergy = curr_energy
  end subroutine

  ! MatrixElements (src/matelem.f90:12) for one pair and one symmetry matrix
  ! P (n x n, column-major).  Dk, Dl: 4 np complex each.
  subroutine ref_matrix_elements(pk, pl, P, gradflag, Hkl, Skl, Dk, Dl) bind(C, name="ref_matrix_elements")
    real(c_double), intent(in) :: pk(*), pl(*)
    integer(c_int), intent(in) :: P(Glob_n, Glob_n)
    integer(c_int), value :: gradflag
    complex(c_double_complex), intent(out) :: Hkl, Skl, Dk(4 * Glob_np), Dl(4 * Glob_np)
    integer :: Pf(Glob_n, Glob_n)
    Pf = P
    Dk = (0.0_dprec, 0.0_dprec)
    Dl = (0.0_dprec, 0.0_dprec)
    call MatrixElements(pk(1:2 * Glob_np), pl(1:2 * Glob_np), Pf, Hkl, Skl, Dk, Dl, int(gradflag))
  end subroutine

  ! MatrixElementsForExpcVals (src/matelem.f90:949) without correlation
  ! functions and densities.  out (complex): H, S, T, V, MV, drach_MV,
  ! Darwin, drach_Darwin, OO, then n x n blocks (column-major) of 1/r^2,
  ! 1/r, r, r^2, delta, drach_delta: 9 + 6 n^2 values.
  subroutine ref_expc_pair(pk, pl, Pbra, Pket, out) bind(C, name="ref_expc_pair")
    real(c_double), intent(in) :: pk(*), pl(*)
    integer(c_int), intent(in) :: Pbra(Glob_n, Glob_n), Pket(Glob_n, Glob_n)
    complex(c_double_complex), intent(out) :: out(*)
    integer :: Pbm(Glob_n, Glob_n), Pkm(Glob_n, Glob_n), n, nn2
    complex(dprec) :: H, S, T, V, MV, dMV, Dar, dDar, OO
    complex(dprec), allocatable :: rm2(:, :), rm(:, :), r(:, :), r2(:, :), dl(:, :), ddl(:, :)
    complex(dprec) :: CF(Glob_n * (Glob_n + 1) / 2, 1), Dens(Glob_n + 1, 1)
    real(dprec) :: grid(1)
    n = Glob_n
    nn2 = n * n
    allocate(rm2(n, n), rm(n, n), r(n, n), r2(n, n), dl(n, n), ddl(n, n))
    rm2 = 0; rm = 0; r = 0; r2 = 0; dl = 0; ddl = 0
    Pbm = Pbra
    Pkm = Pket
    grid = 0.0_dprec
    call MatrixElementsForExpcVals(pk(1:2 * Glob_np), pl(1:2 * Glob_np), Pbm, Pkm, H, S, T, V, rm2, rm, r, r2, dl, &
                                   ddl, MV, dMV, Dar, dDar, OO, 1, grid, CF, 1, grid, Dens, .false., .false.)
    out(1:9) = (/ H, S, T, V, MV, dMV, Dar, dDar, OO /)
    out(10:9 + nn2) = reshape(rm2, (/ nn2 /))
    out(10 + nn2:9 + 2 * nn2) = reshape(rm, (/ nn2 /))
    out(10 + 2 * nn2:9 + 3 * nn2) = reshape(r, (/ nn2 /))
    out(10 + 3 * nn2:9 + 4 * nn2) = reshape(r2, (/ nn2 /))
    out(10 + 4 * nn2:9 + 5 * nn2) = reshape(dl, (/ nn2 /))
    out(10 + 5 * nn2:9 + 6 * nn2) = reshape(ddl, (/ nn2 /))
  end subroutine

end module ref_shim
