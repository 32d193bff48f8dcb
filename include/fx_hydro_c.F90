!! fx_hydro_c.F90 -- ISO_C_BINDING module over the C ABI of include/fx_hydro.h,
!! for a FLASH5 build whose unsplit Hydro (source/physics/Hydro/HydroMain/
!! unsplit/Hydro.F90) hands its tiles to libfx_hydro.so.  Added beside
!! Hydro.F90; the derived types mirror the C structs field for field (the
!! layouts are checked against the C compiler's by tests/test_fortran_binding.py).
module fx_hydro_c
  use iso_c_binding
  implicit none

  ! slope limiters (UHD.h:15-19), Riemann solvers (UHD.h:59-66), precision modes
  integer(c_int), parameter :: FX_MINMOD=1, FX_MC=2, FX_HYBRID=3, FX_VANLEER=4, FX_LIMITED=5
  integer(c_int), parameter :: FX_HLL=3, FX_HLLC=4, FX_HYBR=8, FX_EXACT=20
  integer(c_int), parameter :: FX_PREC_FP64=0, FX_PREC_MIXED=2
  integer(c_int), parameter :: FX_NFLUX=8, FX_NVAR_USED=10

  type, bind(C) :: fx_params
     integer(c_int) :: ndim, nxb, nyb, nzb, ng, nunk
     integer(c_int) :: var(10)          ! 0-based slots of DENS..EINT (Flash.h *_VAR - 1)
     real(c_double) :: gamma
     integer(c_int) :: limiter
     real(c_double) :: limitedSlopeBeta
     integer(c_int) :: charLimiting, flattening, steepening, use3dFullCTU, useAuxEint, riemann
     real(c_double) :: fP, smlrho, smallp, smallE, eintSwitch, tiny, gasConstant, singleSpeciesA
     integer(c_int) :: exactMaxIter
     real(c_double) :: exactTol
     integer(c_int) :: precision, learnedSeed
     integer(c_int) :: shockDetect, shockLowerCFL
     integer(c_int) :: shokVar          ! SHOK_VAR - 1 (the unsplit Config's "shok")
     integer(c_int) :: fallbackLowerCFL ! hy_fallbackLowerCFL
     real(c_double) :: cflFallbackFactor ! hy_cflFallbackFactor
  end type

  type, bind(C) :: fx_tile
     type(c_ptr)    :: U                ! U(i,j,k,var) over the guard-inclusive box
     type(c_ptr)    :: flux(3)          ! flux densities F(i,j,k,flux) per direction
     type(c_ptr)    :: pavg             ! face-averaged pressure per interior cell
     real(c_double) :: del(3)           ! tileDesc%deltas
     integer(c_int) :: level, block_class
  end type

  type, bind(C) :: fx_amr_block
     integer(c_int) :: level, pos(3)
  end type

  type, bind(C) :: fx_status
     integer(c_int)       :: code, first_tile, first_i, first_j, first_k, first_dir
     integer(c_long_long) :: fallbacks, newton_iters, newton_fallbacks
     real(c_double)       :: min_cell_cfl, fallback_cfl
  end type

  type, bind(C) :: fx_fluxreg_face
     integer(c_int) :: coarse_tile, fine_tile, dir
     integer(c_int) :: ci, cj, ck       ! 0-based face in the coarse tile's flux(dir)
     integer(c_int) :: fi, fj, fk       ! first covering face in the fine tile's flux(dir)
  end type

  interface
     integer(c_int) function fx_hydro_init(p) bind(C, name="fx_hydro_init")
       import :: c_int, fx_params
       type(fx_params), intent(in) :: p
     end function
     subroutine fx_hydro_finalize() bind(C, name="fx_hydro_finalize")
     end subroutine
     function fx_hydro_version() bind(C, name="fx_hydro_version")
       import :: c_ptr
       type(c_ptr) :: fx_hydro_version
     end function
     integer(c_int) function fx_hydro_compute_fluxes(tiles, n, dt, dtOld, inPlace, st) &
          bind(C, name="fx_hydro_compute_fluxes")
       import :: c_int, c_double, c_ptr, fx_tile, fx_status
       type(fx_tile), intent(in) :: tiles(*)
       integer(c_int), value     :: n
       real(c_double), value     :: dt, dtOld
       type(c_ptr), value        :: inPlace          ! unsigned char[n] or c_null_ptr
       type(fx_status)           :: st
     end function
     integer(c_int) function fx_hydro_update_solution(tiles, n, dt, dtOld, st) &
          bind(C, name="fx_hydro_update_solution")
       import :: c_int, c_double, fx_tile, fx_status
       type(fx_tile), intent(in) :: tiles(*)
       integer(c_int), value     :: n
       real(c_double), value     :: dt, dtOld
       type(fx_status)           :: st
     end function
     integer(c_int) function fx_hydro_advance_batch(U, nBlocks, stride, del, dt, st) &
          bind(C, name="fx_hydro_advance_batch")
       import :: c_int, c_long, c_double, c_ptr, fx_status
       type(c_ptr), value        :: U
       integer(c_long), value    :: nBlocks, stride
       real(c_double), intent(in):: del(3)
       real(c_double), value     :: dt
       type(fx_status)           :: st
     end function
     ! the same without any allocation (a fixed scratch array per GPU)
     integer(c_int) function fx_hydro_advance_batch_noalloc(U, nBlocks, stride, del, dt, st) &
          bind(C, name="fx_hydro_advance_batch_noalloc")
       import :: c_int, c_long, c_double, c_ptr, fx_status
       type(c_ptr), value        :: U
       integer(c_long), value    :: nBlocks, stride
       real(c_double), intent(in):: del(3)
       real(c_double), value     :: dt
       type(fx_status)           :: st
     end function
     integer(c_int) function fx_hydro_compute_dt(U, nBlocks, stride, del, cfl, dt) &
          bind(C, name="fx_hydro_compute_dt")
       import :: c_int, c_long, c_double, c_ptr
       type(c_ptr), value        :: U
       integer(c_long), value    :: nBlocks, stride
       real(c_double), intent(in):: del(3)
       real(c_double), value     :: cfl
       real(c_double)            :: dt
     end function
     integer(c_int) function fx_fluxreg_put(fine, nFine, faces, nFaces, reg, st) &
          bind(C, name="fx_fluxreg_put")
       import :: c_int, c_ptr, fx_tile, fx_fluxreg_face, fx_status
       type(fx_tile), intent(in)         :: fine(*)
       integer(c_int), value             :: nFine, nFaces
       type(fx_fluxreg_face), intent(in) :: faces(*)
       type(c_ptr), value                :: reg            ! device double[FX_NFLUX*nFaces]
       type(fx_status)                   :: st
     end function
     integer(c_int) function fx_fluxreg_conserve(coarse, nCoarse, faces, nFaces, reg, st) &
          bind(C, name="fx_fluxreg_conserve")
       import :: c_int, c_ptr, fx_tile, fx_fluxreg_face, fx_status
       type(fx_tile), intent(in)         :: coarse(*)
       integer(c_int), value             :: nCoarse, nFaces
       type(fx_fluxreg_face), intent(in) :: faces(*)
       type(c_ptr), value                :: reg
       type(fx_status)                   :: st
     end function
     integer(c_int) function fx_hydro_shock_detect(tiles, n, st) &
          bind(C, name="fx_hydro_shock_detect")
       import :: c_int, fx_tile, fx_status
       type(fx_tile), intent(in) :: tiles(*)
       integer(c_int), value     :: n
       type(fx_status)           :: st
     end function
     integer(c_int) function fx_hydro_shock_detect_batch(U, nBlocks, stride, del, st) &
          bind(C, name="fx_hydro_shock_detect_batch")
       import :: c_int, c_long, c_double, c_ptr, fx_status
       type(c_ptr), value        :: U
       integer(c_long), value    :: nBlocks, stride
       real(c_double), intent(in):: del(3)
       type(fx_status)           :: st
     end function
     subroutine fx_hydro_set_stream(stream) bind(C, name="fx_hydro_set_stream")
       import :: c_ptr
       type(c_ptr), value :: stream                      ! hipStream_t
     end subroutine
     ! levels of different block sizes in one call (no reference counterpart)
     integer(c_int) function fx_hydro_set_block_classes(nb, nClasses) &
          bind(C, name="fx_hydro_set_block_classes")
       import :: c_int
       integer(c_int), intent(in) :: nb(3, *)
       integer(c_int), value      :: nClasses
     end function
     ! launch shapes timed on the caller's own tiles (no reference counterpart)
     integer(c_int) function fx_hydro_autotune(tiles, n, dt, reps, out, cap, nOut) &
          bind(C, name="fx_hydro_autotune")
       import :: c_int, c_double, c_ptr, fx_tile
       type(fx_tile), intent(in) :: tiles(*)
       integer(c_int), value     :: n, reps, cap
       real(c_double), value     :: dt
       type(c_ptr), value        :: out              ! fx_tune_result[cap] or c_null_ptr
       integer(c_int)            :: nOut
     end function
     ! a multi-level grid held on the device
     integer(c_int) function fx_amr_setup(tiles, blocks, n, nLevels, levelExtent, periodic, st) &
          bind(C, name="fx_amr_setup")
       import :: c_int, fx_tile, fx_amr_block, fx_status
       type(fx_tile), intent(in)      :: tiles(*)
       type(fx_amr_block), intent(in) :: blocks(*)
       integer(c_int), value          :: n, nLevels
       integer(c_int), intent(in)     :: levelExtent(3, *), periodic(3)
       type(fx_status)                :: st
     end function
     integer(c_int) function fx_amr_fill_guards(st) bind(C, name="fx_amr_fill_guards")
       import :: c_int, fx_status
       type(fx_status) :: st
     end function
     integer(c_int) function fx_amr_restrict(st) bind(C, name="fx_amr_restrict")
       import :: c_int, fx_status
       type(fx_status) :: st
     end function
     integer(c_int) function fx_amr_compute_dt(cfl, dt) bind(C, name="fx_amr_compute_dt")
       import :: c_int, c_double
       real(c_double), value :: cfl
       real(c_double)        :: dt
     end function
     integer(c_int) function fx_amr_step(dt, st) bind(C, name="fx_amr_step")
       import :: c_int, c_double, fx_status
       real(c_double), value :: dt
       type(fx_status)       :: st
     end function
  end interface
end module fx_hydro_c
