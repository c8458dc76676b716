! Fortran host layer of the MI355X ECG path (libecg_hip.so).
!
! Three modules:
!   ecg_capi     ISO_C_BINDING interface blocks for every entry point of
!                include/ecg.h (plain pointers, sizes and scalars).
!   ecg_globals  the part of the reference's global state (src/globvars.f90)
!                that the path reads and writes, with the reference's names
!                and types.  Inside ATOM-MOL-nonBO this module is replaced by
!                "use globvars" (see the line marked GLOBALS below).
!   ecg_hostpath the routines the path replaces, with the reference's names
!                and argument lists:
!                  MatrixElements          src/matelem.f90:12
!                  ComputeMatElem          src/matform.f90:119
!                  ComputeMatElemAndDeriv  src/matform.f90:234
!                  GHEPIIS                 src/linalg.f90:1781
!                  EnergyGA / EnergyIA     src/workproc.f90:1697, 1954
!                  EnergyGAM / EnergyIAM   src/workproc.f90:1775, 2014
!                  EnergyGB / EnergyIB     src/workproc.f90:1841, 2063
!                  ComputeOverlapPenalty   src/workproc.f90:1491
!                  ComputeOverlapPenaltyAndAddGradient  src/workproc.f90:1571
!                  EcgExpectationValues               src/workproc.f90:8158-8300 (ExpectationValuesG's loop)
!                  EcgCorrelationFunctions            src/workproc.f90:8234-8243 (its correlation functions)
!                  EcgParticleDensities               src/workproc.f90:8245-8254 (its particle densities)
!                  EcgRelativisticCorrections         src/matelem.f90:1479-1667, 1866-1884 (MV, Darwin)
!                  EcgDrachmanizedCorrections         src/matelem.f90:1276-1478, 1669-1933 (drach delta, OO,
!                                                     drach MV, drach Darwin at Glob_CurrEnergy)
!                plus EcgHostInit / EcgHostFinish (context set-up, no
!                counterpart in the reference: ProgramDataInit ends by
!                calling EcgHostInit).
! Errors come back in ErrorCode with the reference's meaning (0 = fine,
! 1 = singular, 2 = not converged) or the library's negative codes; nothing
! here stops the program.

module ecg_capi
  use, intrinsic :: iso_c_binding
  implicit none
  integer(c_int), parameter :: ECG_OK = 0, ECG_ERR_ARG = -1, ECG_ERR_HIP = -2, ECG_ERR_NOMEM = -3, &
                               ECG_ERR_NONFINITE = -4, ECG_ERR_STATE = -5, ECG_ERR_SYNC = -6, ECG_SINGULAR = 1, &
                               ECG_NOT_CONVERGED = 2, ECG_ROOT_NOT_FOUND = 3
  interface
    integer(c_int) function ecg_ctx_create(n, T, yhy_mats, yhy_coef, mass, q0, q, ctx) &
        bind(C, name="ecg_ctx_create")
      import :: c_int, c_double, c_ptr
      integer(c_int), value :: n, T
      integer(c_int), intent(in) :: yhy_mats(*), yhy_coef(*)
      real(c_double), intent(in) :: mass(*)
      real(c_double), value :: q0
      real(c_double), intent(in) :: q(*)
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function ecg_ctx_destroy(ctx) bind(C, name="ecg_ctx_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function ecg_set_params(ctx, K, params) bind(C, name="ecg_set_params")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: K
      real(c_double), intent(in) :: params(*)
    end function
    integer(c_int) function ecg_set_params_device(ctx, K, dparams) bind(C, name="ecg_set_params_device")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: K
      type(c_ptr), value :: dparams   ! device address
    end function
    integer(c_int) function ecg_update_params(ctx, first, count, params) bind(C, name="ecg_update_params")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: first, count
      real(c_double), intent(in) :: params(*)
    end function
    integer(c_int) function ecg_matrix_elements(ctx, paramk, paraml, P, gradflag, Hkl, Skl, Dk, Dl) &
        bind(C, name="ecg_matrix_elements")
      import :: c_int, c_double, c_double_complex, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: paramk(*), paraml(*)
      integer(c_int), intent(in) :: P(*)
      integer(c_int), value :: gradflag
      complex(c_double_complex), intent(out) :: Hkl, Skl
      complex(c_double_complex), intent(inout) :: Dk(*), Dl(*)
    end function
    integer(c_int) function ecg_matrix_elements_noalloc(ctx, paramk, paraml, P, gradflag, Hkl, Skl, Dk, Dl) &
        bind(C, name="ecg_matrix_elements_noalloc")
      import :: c_int, c_double, c_double_complex, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: paramk(*), paraml(*)
      integer(c_int), intent(in) :: P(*)
      integer(c_int), value :: gradflag
      complex(c_double_complex), intent(out) :: Hkl, Skl
      complex(c_double_complex), intent(inout) :: Dk(*), Dl(*)
    end function
    integer(c_int) function ecg_compute_mat_elem(ctx, nmin, nmax, method, approx_energy, H, S, ld, &
        diagH, diagS) bind(C, name="ecg_compute_mat_elem")
      import :: c_int, c_double, c_double_complex, c_ptr, c_char
      type(c_ptr), value :: ctx
      integer(c_int), value :: nmin, nmax, ld
      character(kind=c_char), value :: method
      real(c_double), value :: approx_energy
      complex(c_double_complex), intent(inout) :: H(*), S(*)
      real(c_double), intent(inout) :: diagH(*), diagS(*)
    end function
    integer(c_int) function ecg_compute_mat_elem_and_deriv(ctx, nfru, method, approx_energy, H, S, ld, &
        diagH, diagS, D, ldD) bind(C, name="ecg_compute_mat_elem_and_deriv")
      import :: c_int, c_double, c_double_complex, c_ptr, c_char
      type(c_ptr), value :: ctx
      integer(c_int), value :: nfru, ld, ldD
      character(kind=c_char), value :: method
      real(c_double), value :: approx_energy
      complex(c_double_complex), intent(inout) :: H(*), S(*), D(*)
      real(c_double), intent(inout) :: diagH(*), diagS(*)
    end function
    integer(c_int) function ecg_ghepiis(ctx, k, n, M, nM, invD, B, nB, apprlambda, v, tol, maxiter, &
        specif_norm, lambda, x, relacc, numiter) bind(C, name="ecg_ghepiis")
      import :: c_int, c_double, c_double_complex, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: k, n, nM, nB, maxiter, specif_norm
      complex(c_double_complex), intent(inout) :: M(*), invD(*)
      complex(c_double_complex), intent(in) :: B(*), v(*)
      real(c_double), value :: apprlambda, tol
      real(c_double), intent(out) :: lambda, relacc
      complex(c_double_complex), intent(out) :: x(*)
      integer(c_int), intent(out) :: numiter
    end function
    integer(c_int) function ecg_energy(ctx, which, sigma, tol, maxiter, E, c, info) bind(C, name="ecg_energy")
      import :: c_int, c_double, c_double_complex, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: which, maxiter
      real(c_double), value :: sigma, tol
      real(c_double), intent(out) :: E
      complex(c_double_complex), intent(out) :: c(*)
      real(c_double), intent(out) :: info(*)
    end function
    integer(c_int) function ecg_energy_grad(ctx, nfru, which, sigma, tol, maxiter, E, grad, info) &
        bind(C, name="ecg_energy_grad")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nfru, which, maxiter
      real(c_double), value :: sigma, tol
      real(c_double), intent(out) :: E
      real(c_double), intent(out) :: grad(*), info(*)
    end function
    integer(c_int) function ecg_energy_grad_stored(ctx, nfru, which, sigma, tol, maxiter, E, grad, info) &
        bind(C, name="ecg_energy_grad_stored")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nfru, which, maxiter
      real(c_double), value :: sigma, tol
      real(c_double), intent(out) :: E
      real(c_double), intent(out) :: grad(*), info(*)
    end function
    integer(c_int) function ecg_screen_set_shapes(ctx, row_tile, cand_per_block) &
        bind(C, name="ecg_screen_set_shapes")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: row_tile, cand_per_block
    end function
    integer(c_int) function ecg_batch_create(nprob, ctxs, batch) bind(C, name="ecg_batch_create")
      import :: c_int, c_ptr
      integer(c_int), value :: nprob
      type(c_ptr), intent(in) :: ctxs(*)
      type(c_ptr), intent(out) :: batch
    end function
    integer(c_int) function ecg_batch_destroy(batch) bind(C, name="ecg_batch_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: batch
    end function
    integer(c_int) function ecg_batch_set_config(batch, packed, nstreams) bind(C, name="ecg_batch_set_config")
      import :: c_int, c_ptr
      type(c_ptr), value :: batch
      integer(c_int), value :: packed, nstreams
    end function
    integer(c_int) function ecg_batch_energy_grad(batch, sigma, tol, maxiter, E, grad, lat_ms) &
        bind(C, name="ecg_batch_energy_grad")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: batch
      real(c_double), intent(in) :: sigma(*)
      real(c_double), value :: tol
      integer(c_int), value :: maxiter
      real(c_double), intent(out) :: E(*), grad(*), lat_ms(*)
    end function
    integer(c_int) function ecg_trace(ctx, enable) bind(C, name="ecg_trace")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: enable
    end function
    integer(c_int) function ecg_trace_read(ctx, count, total_ms, nmax) bind(C, name="ecg_trace_read")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: count(*), total_ms(*)
      integer(c_int), value :: nmax
    end function
    integer(c_int) function ecg_get_params(ctx, params) bind(C, name="ecg_get_params")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: params(*)
    end function
    integer(c_int) function ecg_screen_prepare(ctx, sigma, tol, maxiter, info) bind(C, name="ecg_screen_prepare")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), value :: sigma, tol
      integer(c_int), value :: maxiter
      real(c_double), intent(out) :: info(*)
    end function
    integer(c_int) function ecg_screen(ctx, ncand, cand, iters, E, info) bind(C, name="ecg_screen")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: ncand, iters
      real(c_double), intent(in) :: cand(*)
      real(c_double), intent(out) :: E(*), info(*)
    end function
    integer(c_int) function ecg_screen_loop(ctx, nsteps, ncand, cand, iters, Ebest, ibest) &
        bind(C, name="ecg_screen_loop")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nsteps, ncand, iters
      real(c_double), intent(in) :: cand(*)
      real(c_double), intent(out) :: Ebest(*)
      integer(c_int), intent(out) :: ibest(*)
    end function
    integer(c_int) function ecg_set_overlap_penalty(ctx, allowed, max_pair_penalty, threshold2, nfru) &
        bind(C, name="ecg_set_overlap_penalty")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: allowed, nfru
      real(c_double), value :: max_pair_penalty, threshold2
    end function
    integer(c_int) function ecg_overlap_penalty(ctx, nfru, max_pair_penalty, threshold2, total, grad) &
        bind(C, name="ecg_overlap_penalty")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int), value :: nfru
      real(c_double), value :: max_pair_penalty, threshold2
      real(c_double), intent(out) :: total
      type(c_ptr), value :: grad          ! npt*(K-nfru) doubles or c_null_ptr
    end function
    integer(c_int) function ecg_expectation_values(ctx, c, values, errors) &
        bind(C, name="ecg_expectation_values")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      type(c_ptr), value :: c              ! K complex or c_null_ptr (last eigenvector)
      real(c_double), intent(out) :: values(*)
      real(c_double), intent(out) :: errors(*)
    end function
    integer(c_int) function ecg_correlation_functions(ctx, c, ngrid, grid, values) &
        bind(C, name="ecg_correlation_functions")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      type(c_ptr), value :: c              ! K complex or c_null_ptr (last eigenvector)
      integer(c_int), value :: ngrid
      real(c_double), intent(in) :: grid(*)
      real(c_double), intent(out) :: values(*)
    end function
    integer(c_int) function ecg_relativistic_corrections(ctx, c, masses, values, errors) &
        bind(C, name="ecg_relativistic_corrections")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      type(c_ptr), value :: c              ! K complex or c_null_ptr (last eigenvector)
      real(c_double), intent(in) :: masses(*)
      real(c_double), intent(out) :: values(*)
      real(c_double), intent(out) :: errors(*)
    end function
    integer(c_int) function ecg_drachmanized_corrections(ctx, c, masses, E, values, errors) &
        bind(C, name="ecg_drachmanized_corrections")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      type(c_ptr), value :: c              ! K complex or c_null_ptr (last eigenvector)
      real(c_double), intent(in) :: masses(*)
      real(c_double), value :: E           ! Glob_CurrEnergy
      real(c_double), intent(out) :: values(*)
      real(c_double), intent(out) :: errors(*)
    end function
    integer(c_int) function ecg_particle_densities(ctx, c, masses, ngrid, grid, values) &
        bind(C, name="ecg_particle_densities")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      type(c_ptr), value :: c              ! K complex or c_null_ptr (last eigenvector)
      real(c_double), intent(in) :: masses(*)
      integer(c_int), value :: ngrid
      real(c_double), intent(in) :: grid(*)
      real(c_double), intent(out) :: values(*)
    end function
    integer(c_int) function ecg_eigenvector(ctx, c) bind(C, name="ecg_eigenvector")
      import :: c_int, c_double_complex, c_ptr
      type(c_ptr), value :: ctx
      complex(c_double_complex), intent(out) :: c(*)
    end function
    integer(c_int) function ecg_last_timings(ctx, ms, nmax) bind(C, name="ecg_last_timings")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: ms(*)
      integer(c_int), value :: nmax
    end function
  end interface
end module ecg_capi

! The reference's global state used by the path (names and types of
! src/globvars.f90).  Replaced by the reference's own module on integration.
module ecg_globals
  implicit none
  integer, parameter :: dprec = kind(1.0d0)
  integer :: Glob_n = 0, Glob_np = 0, Glob_npt = 0
  integer :: Glob_NumYHYTerms = 0
  integer, allocatable :: Glob_YHYMatr(:,:,:), Glob_YHYCoeff(:)
  real(dprec), allocatable :: Glob_MassMatrix(:,:), Glob_PseudoCharge(:)
  real(dprec) :: Glob_PseudoCharge0 = 0
  integer :: Glob_CurrBasisSize = 0, Glob_HSLeadDim = 0, Glob_nfru = 0, Glob_nfo = 0
  real(dprec), allocatable :: Glob_NonlinParam(:,:)
  complex(dprec), allocatable :: Glob_H(:,:), Glob_S(:,:), Glob_D(:,:,:), Glob_c(:), Glob_invD(:)
  real(dprec), allocatable :: Glob_diagH(:), Glob_diagS(:)
  character :: Glob_GSEPSolutionMethod = 'G'
  real(dprec) :: Glob_ApproxEnergy = 0
  real(dprec) :: Glob_CurrEnergy = 0
  integer :: Glob_WhichEigenvalue = 1
  real(dprec) :: Glob_EigvalTol = 1.0d-12
  integer :: Glob_InvItMaxIter = 20
  complex(dprec), allocatable :: Glob_LastEigvector(:)
  ! Method 'G' (ZHEGVX in the reference) has no tolerance or iteration
  ! limit of its own.  On the device its root is refined by inverse
  ! iteration run to its rounding floor: the call succeeds when GHEPIIS's
  ! measure there is <= Ecg_GFloorBound (an eigenvector error of 1e-7 gives
  ! an energy error of order 1e-14 relative), within Ecg_GIterLimit
  ! iterations.  These two are this layer's own settings, not reference
  ! globals.
  real(dprec) :: Ecg_GFloorBound = 1.0d-7
  integer :: Ecg_GIterLimit = 60
  logical :: Glob_OverlapPenaltyAllowed = .false.
  real(dprec) :: Glob_OverlapPenaltyThreshold2 = 1
  real(dprec) :: Glob_MaxOverlapPenalty = 1
  real(dprec) :: Glob_TotalOverlapPenalty = 0
end module ecg_globals

module ecg_hostpath
  use, intrinsic :: iso_c_binding
  use, intrinsic :: ieee_arithmetic
  use ecg_capi
  use ecg_globals   ! GLOBALS: "use globvars" inside ATOM-MOL-nonBO
  implicit none
  private
  public :: EcgHostInit, EcgHostFinish, MatrixElements, ComputeMatElem, ComputeMatElemAndDeriv, GHEPIIS, &
            EnergyGA, EnergyIA, EnergyGAM, EnergyIAM, EnergyGB, EnergyIB, ComputeOverlapPenalty, ComputeOverlapPenaltyAndAddGradient, &
            EcgExpectationValues, EcgCorrelationFunctions, EcgParticleDensities, &
            EcgRelativisticCorrections, EcgDrachmanizedCorrections
  type(c_ptr), save :: ctx = c_null_ptr

contains

  ! Create the device context from the system globals (after the reference's
  ! ProgramDataInit has filled Glob_MassMatrix, Glob_PseudoCharge*, Glob_YHYMatr).
  subroutine EcgHostInit(ErrorCode)
    integer, intent(out) :: ErrorCode
    if (c_associated(ctx)) ErrorCode = ecg_ctx_destroy(ctx)
    ErrorCode = ecg_ctx_create(Glob_n, Glob_NumYHYTerms, Glob_YHYMatr, Glob_YHYCoeff, Glob_MassMatrix, &
                               Glob_PseudoCharge0, Glob_PseudoCharge, ctx)
  end subroutine

  subroutine EcgHostFinish()
    integer :: rc
    if (c_associated(ctx)) rc = ecg_ctx_destroy(ctx)
    ctx = c_null_ptr
  end subroutine

  ! parameters of functions 1..K to the device, with the overlap penalty
  ! settings the energy routines apply
  integer function sync_params(K)
    integer, intent(in) :: K
    integer(c_int) :: allowed
    allowed = 0
    if (Glob_OverlapPenaltyAllowed) allowed = 1
    sync_params = ecg_set_params(ctx, K, Glob_NonlinParam)
    if (sync_params == ECG_OK) sync_params = ecg_set_overlap_penalty(ctx, allowed, Glob_MaxOverlapPenalty, &
                                                                     Glob_OverlapPenaltyThreshold2, Glob_nfru)
  end function

  ! src/matelem.f90:12 — one pair, one symmetry matrix, normalised elements
  subroutine MatrixElements(Paramk, Paraml, P, Hkl, Skl, Dk, Dl, gradflag)
    real(dprec), intent(in) :: Paramk(2*Glob_np), Paraml(2*Glob_np)
    integer, intent(in) :: P(Glob_n, Glob_n)
    complex(dprec), intent(out) :: Skl, Hkl
    complex(dprec), intent(out) :: Dk(4*Glob_np), Dl(4*Glob_np)
    integer, intent(in) :: gradflag
    integer :: rc
    rc = ecg_matrix_elements(ctx, Paramk, Paraml, P, gradflag, Hkl, Skl, Dk, Dl)
    if (rc /= ECG_OK) then
      Hkl = cmplx(ieee_value(1.0_dprec, ieee_quiet_nan), 0.0_dprec, dprec)
      Skl = Hkl
    end if
  end subroutine

  ! src/matform.f90:119 — rows Nmin..Nmax into Glob_H, Glob_S, Glob_diagH,
  ! Glob_diagS with StoreHS's normalisation for Glob_GSEPSolutionMethod
  subroutine ComputeMatElem(Nmin, Nmax)
    integer, intent(in) :: Nmin, Nmax
    integer :: rc
    rc = sync_params(Nmax)
    if (rc == ECG_OK) rc = ecg_compute_mat_elem(ctx, Nmin, Nmax, Glob_GSEPSolutionMethod, Glob_ApproxEnergy, &
                                                Glob_H, Glob_S, Glob_HSLeadDim, Glob_diagH, Glob_diagS)
    if (rc /= ECG_OK) Glob_diagS(Nmin) = ieee_value(1.0_dprec, ieee_quiet_nan)
  end subroutine

  ! src/matform.f90:234 — as ComputeMatElem plus Glob_D in StoreHSD's
  ! layout for the optimised functions Glob_nfru+1..Nmax (Nmin = Glob_nfru+1)
  subroutine ComputeMatElemAndDeriv(Nmin, Nmax)
    integer, intent(in) :: Nmin, Nmax
    integer :: rc
    rc = sync_params(Nmax)
    if (rc == ECG_OK) rc = ecg_compute_mat_elem_and_deriv(ctx, Nmin - 1, Glob_GSEPSolutionMethod, &
                                                          Glob_ApproxEnergy, Glob_H, Glob_S, Glob_HSLeadDim, &
                                                          Glob_diagH, Glob_diagS, Glob_D, size(Glob_D, 2))
    if (rc /= ECG_OK) Glob_diagS(Nmin) = ieee_value(1.0_dprec, ieee_quiet_nan)
  end subroutine

  ! src/linalg.f90:1781 — inverse iteration on host arrays; the LDL^H
  ! factor (L^H in the upper triangle of M, 1/D in invD) is recomputed on
  ! the device from row 1, so rows < k need not hold a factor; w (a work
  ! array of the reference routine) is not used.
  subroutine GHEPIIS(k, n, M, nM, invD, B, nB, apprlambda, v, w, Tol, lambda, x, RelAcc, MaxIter, SpecifNorm, &
                     NumIter, ErrorCode)
    integer :: k, n, nM, nB, MaxIter, SpecifNorm, NumIter, ErrorCode
    complex(dprec) :: M(nM, n), invD(n), B(nB, n), v(n), w(n), x(n)
    real(dprec) :: apprlambda, Tol, lambda, RelAcc
    ErrorCode = ecg_ghepiis(ctx, k, n, M, nM, invD, B, nB, apprlambda, v, Tol, MaxIter, SpecifNorm, lambda, x, &
                            RelAcc, NumIter)
  end subroutine

  ! src/workproc.f90:1697 — energy of root Glob_WhichEigenvalue of functions
  ! 1..Nmax (root selected by inertia counts instead of ZHEGVX RANGE='I');
  ! eigenvector to Glob_c.  H and S are assembled and kept on the device
  ! whatever Nmin and AreMatElemNeeded say (a full assembly costs less
  ! than the eigen step), and Glob_H, Glob_S are not written.
  real(dprec) function EnergyGA(Nmin, Nmax, AreMatElemNeeded, ErrorCode)
    integer :: Nmin, Nmax, ErrorCode
    logical :: AreMatElemNeeded
    real(dprec) :: info(8)
    complex(dprec), allocatable :: cv(:)
    info = 0
    EnergyGA = ieee_value(1.0_dprec, ieee_quiet_nan)
    allocate(cv(max(Nmax, 1)))   ! ZHEGVX(JOBZ='N'): Glob_c is not written
    ErrorCode = sync_params(Nmax)
    if (ErrorCode == ECG_OK) ErrorCode = ecg_energy(ctx, Glob_WhichEigenvalue, ieee_value(1.0_dprec, ieee_quiet_nan), &
                                                    -Ecg_GFloorBound, Ecg_GIterLimit, EnergyGA, cv, info)
    if (ErrorCode == ECG_OK) Glob_TotalOverlapPenalty = info(7)
  end function

  ! src/workproc.f90:1775 — as EnergyGA, with the eigenvector (c^H S c = 1,
  ! ZHEGVX's normalisation for ITYPE=1) in Glob_c
  real(dprec) function EnergyGAM(Nmin, Nmax, AreMatElemNeeded, ErrorCode)
    integer :: Nmin, Nmax, ErrorCode
    logical :: AreMatElemNeeded
    real(dprec) :: info(8)
    info = 0
    EnergyGAM = ieee_value(1.0_dprec, ieee_quiet_nan)
    ErrorCode = sync_params(Nmax)
    if (ErrorCode == ECG_OK) ErrorCode = ecg_energy(ctx, Glob_WhichEigenvalue, ieee_value(1.0_dprec, ieee_quiet_nan), &
                                                    -Ecg_GFloorBound, Ecg_GIterLimit, EnergyGAM, Glob_c, info)
    if (ErrorCode == ECG_OK) Glob_TotalOverlapPenalty = info(7)
  end function

  ! The vectors GHEPIIS leaves behind in EnergyIA/EnergyIAM
  ! (src/workproc.f90:1989-1991, 2038-2040): x, normalised as SpecifNorm
  ! asks, goes to Glob_LastEigvector; Glob_c is GHEPIIS's work vector v,
  ! i.e. the iterate before the last one scaled to max |Re|,|Im| = 1.  The
  ! device returns the final iterate (c^H S c = 1); its max-abs scaling
  ! stands in for v, from which it differs by less than the convergence
  ! measure.
  subroutine iteration_vectors(n, c, specif_norm)
    integer, intent(in) :: n, specif_norm
    complex(dprec), intent(in) :: c(:)
    real(dprec) :: t
    if (.not. allocated(Glob_LastEigvector)) allocate(Glob_LastEigvector(size(Glob_c)))
    t = maxval(max(abs(real(c(1:n))), abs(aimag(c(1:n)))))
    Glob_c(1:n) = c(1:n) / t
    if (specif_norm == 0) then
      Glob_LastEigvector(1:n) = c(1:n)
    else
      Glob_LastEigvector(1:n) = Glob_c(1:n)
    end if
  end subroutine

  ! src/workproc.f90:1954 — as EnergyGA with the shift Glob_ApproxEnergy
  ! (method 'I')
  real(dprec) function EnergyIA(Nmin, Nmax, AreMatElemNeeded, ErrorCode)
    integer :: Nmin, Nmax, ErrorCode
    logical :: AreMatElemNeeded
    real(dprec) :: info(8)
    complex(dprec), allocatable :: cv(:)
    info = 0
    EnergyIA = ieee_value(1.0_dprec, ieee_quiet_nan)
    allocate(cv(max(Nmax, 1)))
    ErrorCode = sync_params(Nmax)
    if (ErrorCode == ECG_OK) ErrorCode = ecg_energy(ctx, Glob_WhichEigenvalue, Glob_ApproxEnergy, Glob_EigvalTol, &
                                                    Glob_InvItMaxIter, EnergyIA, cv, info)
    if (ErrorCode == ECG_OK .or. ErrorCode == ECG_NOT_CONVERGED) call iteration_vectors(Nmax, cv, -1)
    if (ErrorCode == ECG_OK) Glob_TotalOverlapPenalty = info(7)
  end function

  ! src/workproc.f90:2014 — as EnergyIA with SpecifNorm = 0: the
  ! eigenvector normalised to x^H S x = 1 in Glob_LastEigvector
  real(dprec) function EnergyIAM(Nmin, Nmax, AreMatElemNeeded, ErrorCode)
    integer :: Nmin, Nmax, ErrorCode
    logical :: AreMatElemNeeded
    real(dprec) :: info(8)
    complex(dprec), allocatable :: cv(:)
    info = 0
    EnergyIAM = ieee_value(1.0_dprec, ieee_quiet_nan)
    allocate(cv(max(Nmax, 1)))
    ErrorCode = sync_params(Nmax)
    if (ErrorCode == ECG_OK) ErrorCode = ecg_energy(ctx, Glob_WhichEigenvalue, Glob_ApproxEnergy, Glob_EigvalTol, &
                                                    Glob_InvItMaxIter, EnergyIAM, cv, info)
    if (ErrorCode == ECG_OK .or. ErrorCode == ECG_NOT_CONVERGED) call iteration_vectors(Nmax, cv, 0)
    if (ErrorCode == ECG_OK) Glob_TotalOverlapPenalty = info(7)
  end function

  ! src/workproc.f90:1841 — energy and gradient w.r.t. the parameters of
  ! functions Glob_nfru+1..Glob_CurrBasisSize, (dE/dvechL_k, dE/dvechB_k)
  ! per function
  subroutine EnergyGB(Evalue, Gradient, AreMatElemNeeded, ErrorCode)
    real(dprec) :: Evalue
    real(dprec) :: Gradient(Glob_npt*Glob_nfo)
    logical :: AreMatElemNeeded
    integer :: ErrorCode
    real(dprec) :: info(8)
    info = 0
    ErrorCode = sync_params(Glob_CurrBasisSize)
    if (ErrorCode == ECG_OK) ErrorCode = ecg_energy_grad(ctx, Glob_nfru, Glob_WhichEigenvalue, &
                                                         ieee_value(1.0_dprec, ieee_quiet_nan), -Ecg_GFloorBound, &
                                                         Ecg_GIterLimit, Evalue, Gradient, info)
    if (ErrorCode == ECG_OK) Glob_TotalOverlapPenalty = info(7)
    ! ZHEGVX(JOBZ='V') leaves the eigenvector in Glob_c
    if (ErrorCode == ECG_OK) ErrorCode = ecg_eigenvector(ctx, Glob_c)
  end subroutine

  ! src/workproc.f90:2063 — as EnergyGB with the shift Glob_ApproxEnergy
  subroutine EnergyIB(Evalue, Gradient, AreMatElemNeeded, ErrorCode)
    real(dprec) :: Evalue
    real(dprec) :: Gradient(Glob_npt*Glob_nfo)
    logical :: AreMatElemNeeded
    integer :: ErrorCode
    real(dprec) :: info(8)
    info = 0
    ErrorCode = sync_params(Glob_CurrBasisSize)
    if (ErrorCode == ECG_OK) ErrorCode = ecg_energy_grad(ctx, Glob_nfru, Glob_WhichEigenvalue, Glob_ApproxEnergy, &
                                                         Glob_EigvalTol, Glob_InvItMaxIter, Evalue, Gradient, info)
    if (ErrorCode == ECG_OK) Glob_TotalOverlapPenalty = info(7)
    ! GHEPIIS(SpecifNorm = 0): x^H S x = 1 in Glob_LastEigvector, copied to Glob_c
    if (ErrorCode == ECG_OK .or. ErrorCode == ECG_NOT_CONVERGED) then
      if (.not. allocated(Glob_LastEigvector)) allocate(Glob_LastEigvector(size(Glob_c)))
      if (ecg_eigenvector(ctx, Glob_c) == ECG_OK) Glob_LastEigvector(1:Glob_CurrBasisSize) = Glob_c(1:Glob_CurrBasisSize)
    end if
  end subroutine

  ! src/workproc.f90:1491 — penalty of the overlaps assembled on the device
  ! by the last energy call or ComputeMatElem (pairs j > Glob_nfru)
  subroutine ComputeOverlapPenalty(MaxPairOverlapPenalty, OverlapThreshold2, TotalPenalty)
    real(dprec) :: MaxPairOverlapPenalty, OverlapThreshold2, TotalPenalty
    integer :: rc
    rc = ecg_overlap_penalty(ctx, Glob_nfru, MaxPairOverlapPenalty, OverlapThreshold2, TotalPenalty, c_null_ptr)
    if (rc /= ECG_OK) TotalPenalty = ieee_value(1.0_dprec, ieee_quiet_nan)
  end subroutine

  ! src/workproc.f90:8158-8300 — the main loop of ExpectationValuesG on the
  ! device, with the eigenvector of the last energy call: Values and Errors
  ! (5 m + 4 each, m = Glob_n (Glob_n + 1) / 2) hold <1/r^2>, <1/r>, <r>,
  ! <r^2>, <delta(r)> for the pairs (a, b >= a), then <H>, <S>, <T>, <V>,
  ! and the driver's rounding-error estimates.
  subroutine EcgExpectationValues(Values, Errors, ErrorCode)
    real(dprec) :: Values(:), Errors(:)
    integer :: ErrorCode
    ErrorCode = ecg_expectation_values(ctx, c_null_ptr, Values, Errors)
  end subroutine

  ! src/workproc.f90:8234-8243 with src/matelem.f90:1934-1946 — the pair
  ! correlation functions CF(b, a) at the radii CFGrid(a), b over the
  ! pairs (i, j >= i), in the eigenvector of the last energy call
  subroutine EcgCorrelationFunctions(NumCFGridPoints, CFGrid, CF, ErrorCode)
    integer :: NumCFGridPoints, ErrorCode
    real(dprec) :: CFGrid(:), CF(:, :)
    ErrorCode = ecg_correlation_functions(ctx, c_null_ptr, NumCFGridPoints, CFGrid, CF)
  end subroutine

  ! src/workproc.f90:8245-8254 with src/matelem.f90:1948-1963 — the particle
  ! densities Dens(i, a) at the radii DensGrid(a), i = 1..n+1, from the
  ! particle masses Mass(1:n+1) (Glob_Mass), in the last eigenvector
  ! src/workproc.f90:8158-8300 with src/matelem.f90:1479-1667, 1866-1884 —
  ! MV and Darwin (Values(1:2)) with their rounding estimates, from the
  ! particle masses Mass(1:n+1), in the last eigenvector
  subroutine EcgRelativisticCorrections(Mass, Values, Errors, ErrorCode)
    real(dprec) :: Mass(:), Values(:), Errors(:)
    integer :: ErrorCode
    ErrorCode = ecg_relativistic_corrections(ctx, c_null_ptr, Mass, Values, Errors)
  end subroutine

  ! Values: Glob_n*(Glob_n+1)/2 Drachmanized delta functions, then OO,
  ! drach_MV, drach_Darwin, in the last eigenvector at Glob_CurrEnergy
  subroutine EcgDrachmanizedCorrections(Mass, Values, Errors, ErrorCode)
    real(dprec) :: Mass(:), Values(:), Errors(:)
    integer :: ErrorCode
    ErrorCode = ecg_drachmanized_corrections(ctx, c_null_ptr, Mass, Glob_CurrEnergy, Values, Errors)
  end subroutine

  subroutine EcgParticleDensities(Mass, NumDensGridPoints, DensGrid, Dens, ErrorCode)
    integer :: NumDensGridPoints, ErrorCode
    real(dprec) :: Mass(:), DensGrid(:), Dens(:, :)
    ErrorCode = ecg_particle_densities(ctx, c_null_ptr, Mass, NumDensGridPoints, DensGrid, Dens)
  end subroutine

  ! src/workproc.f90:1571 — as ComputeOverlapPenalty, adding the penalty's
  ! gradient to WkGR (functions Glob_nfru+1.., npt per function)
  subroutine ComputeOverlapPenaltyAndAddGradient(MaxPairOverlapPenalty, OverlapThreshold2, TotalPenalty, WkGR)
    real(dprec) :: MaxPairOverlapPenalty, OverlapThreshold2, TotalPenalty, WkGR(:)
    real(c_double), allocatable, target :: g(:)
    integer :: rc
    allocate(g(size(WkGR)))
    g = WkGR
    rc = ecg_overlap_penalty(ctx, Glob_nfru, MaxPairOverlapPenalty, OverlapThreshold2, TotalPenalty, c_loc(g))
    if (rc == ECG_OK) then
      WkGR = g
    else
      TotalPenalty = ieee_value(1.0_dprec, ieee_quiet_nan)
    end if
  end subroutine

end module ecg_hostpath
