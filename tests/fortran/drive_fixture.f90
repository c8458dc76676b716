! Test driver: runs the Fortran host layer (module ecg_hostpath) the way the
! reference's method 'I' does — ComputeMatElem + GHEPIIS on host arrays —
! and the device-resident EnergyIA / EnergyIB, on a system and basis read
! from a plain text file written by the Python test:
!   n T / T matrices (n*n ints, column-major) / T coefficients /
!   mass matrix (n*n, column-major) / q0 / q(1..n) / K which approx_energy /
!   K columns of 2*n(n+1)/2 parameters
! Prints "GHEPIIS <E> <err> <iters>", "ENERGYIA <E> <err>",
! "ENERGYIB <E> <err> <|grad|> <grad(1)> <grad(npt*K)>", then method 'G'
! (Glob_EigvalTol as read, which ZHEGVX does not use): "ENERGYGA <E> <err>",
! "ENERGYGAM <E> <err> <c^H S c>", "ENERGYGB <E> <err> <|grad|> <c^H S c>",
! and "ENERGYIAM <E> <err> <x^H S x> <max |Re|,|Im| of Glob_c>".
program drive_fixture
  use ecg_globals
  use ecg_hostpath
  implicit none
  character(len=512) :: path
  integer :: u, T, K, i, err, iters, ld
  real(dprec) :: approx, lambda, relacc, E, tol
  complex(dprec), allocatable :: v(:), w(:), x(:)
  real(dprec), allocatable :: grad(:)

  call get_command_argument(1, path)
  open(newunit=u, file=trim(path), status='old', action='read')
  read(u, *) Glob_n, T
  Glob_NumYHYTerms = T
  Glob_np = Glob_n * (Glob_n + 1) / 2
  Glob_npt = 2 * Glob_np
  allocate(Glob_YHYMatr(Glob_n, Glob_n, T), Glob_YHYCoeff(T), Glob_MassMatrix(Glob_n, Glob_n), &
           Glob_PseudoCharge(Glob_n))
  read(u, *) Glob_YHYMatr
  read(u, *) Glob_YHYCoeff
  read(u, *) Glob_MassMatrix
  read(u, *) Glob_PseudoCharge0
  read(u, *) Glob_PseudoCharge
  read(u, *) K, Glob_WhichEigenvalue, approx
  read(u, *) tol
  Glob_EigvalTol = tol
  allocate(Glob_NonlinParam(Glob_npt, K))
  read(u, *) Glob_NonlinParam
  close(u)

  Glob_CurrBasisSize = K
  ld = ((K + 63) / 64) * 64
  Glob_HSLeadDim = ld
  Glob_nfru = 0
  Glob_nfo = K
  allocate(Glob_H(ld, ld), Glob_S(ld, ld), Glob_diagH(ld), Glob_diagS(ld), Glob_c(K), Glob_invD(ld))
  allocate(v(K), w(K), x(K), grad(Glob_npt * K))
  Glob_H = (0.0_dprec, 0.0_dprec)
  Glob_S = (0.0_dprec, 0.0_dprec)

  call EcgHostInit(err)
  if (err /= 0) then
    print '(a,i0)', 'INIT ', err
    stop 1
  end if

  ! method 'I' on host arrays: H - approx S in the lower triangle, S full
  Glob_GSEPSolutionMethod = 'I'
  Glob_ApproxEnergy = approx
  call ComputeMatElem(1, K)
  v = (1.0_dprec, 0.0_dprec)
  call GHEPIIS(1, K, Glob_H, ld, Glob_invD, Glob_S, ld, approx, v, w, Glob_EigvalTol, lambda, x, relacc, 20, 0, &
               iters, err)
  print '(a,es25.16e3,1x,i0,1x,i0)', 'GHEPIIS ', lambda, err, iters

  E = EnergyIA(1, K, .true., err)
  print '(a,es25.16e3,1x,i0)', 'ENERGYIA ', E, err

  call EnergyIB(E, grad, .true., err)
  print '(a,es25.16e3,1x,i0,3(1x,es25.16e3))', 'ENERGYIB ', E, err, sqrt(sum(grad**2)), grad(1), grad(Glob_npt * K)

  ! method 'G': the reference's ZHEGVX has no tolerance; Glob_EigvalTol
  ! keeps the value read above and must not make these calls fail
  Glob_GSEPSolutionMethod = 'G'
  E = EnergyGA(1, K, .true., err)
  print '(a,es25.16e3,1x,i0)', 'ENERGYGA ', E, err
  Glob_c = (0.0_dprec, 0.0_dprec)
  E = EnergyGAM(1, K, .true., err)
  print '(a,es25.16e3,1x,i0,1x,es25.16e3)', 'ENERGYGAM ', E, err, snorm(Glob_c)
  Glob_c = (0.0_dprec, 0.0_dprec)
  call EnergyGB(E, grad, .true., err)
  print '(a,es25.16e3,1x,i0,2(1x,es25.16e3))', 'ENERGYGB ', E, err, sqrt(sum(grad**2)), snorm(Glob_c)
  E = EnergyIAM(1, K, .true., err)
  print '(a,es25.16e3,1x,i0,2(1x,es25.16e3))', 'ENERGYIAM ', E, err, snorm(Glob_LastEigvector), &
        maxval(max(abs(real(Glob_c)), abs(aimag(Glob_c))))

  call EcgHostFinish()

contains

  ! c^H S c with the normalised S of the 'I' assembly above (both triangles)
  real(dprec) function snorm(c)
    complex(dprec), intent(in) :: c(:)
    snorm = real(dot_product(c(1:K), matmul(Glob_S(1:K, 1:K), c(1:K))))
  end function
end program drive_fixture
