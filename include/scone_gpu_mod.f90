!!
!! scone_gpu_mod -- iso_c_binding interface to libscone_amd.so (include/scone_gpu.h) for SCONE.
!!
!! A SCONE physics package uses these interfaces in place of its per-history OpenMP loop
!! (PhysicsPackages/eigenPhysicsPackage_class.f90:211-252).  Derived types mirror the C structs
!! field for field; logical flags travel as integer(c_int32_t).
!!
module scone_gpu_mod
  use iso_c_binding
  implicit none
  private

  integer(c_int32_t), parameter, public :: SCN_OK = 0, SCN_TRANSPORT_ST = 1, SCN_TRANSPORT_DT = 2

  !! particleStateData (ParticleObjects/particle_class.f90:50-65) as bind(C)
  type, public, bind(C) :: scn_particle_state
    real(c_double)     :: wgt
    real(c_double)     :: r(3)
    real(c_double)     :: dir(3)
    real(c_double)     :: E
    integer(c_int32_t) :: G
    integer(c_int32_t) :: isMG
    integer(c_int32_t) :: type
    integer(c_int32_t) :: pad0
    real(c_double)     :: time
    real(c_double)     :: lambda
    integer(c_int32_t) :: matIdx
    integer(c_int32_t) :: cellIdx
    integer(c_int32_t) :: uniqueID
    integer(c_int32_t) :: collisionN
    integer(c_int32_t) :: broodID
    integer(c_int32_t) :: pad1
  end type scn_particle_state

  type, public, bind(C) :: scn_eigen_settings
    integer(c_int32_t) :: pop
    integer(c_int32_t) :: n_inactive
    integer(c_int32_t) :: n_active
    integer(c_int32_t) :: inactive_tally
    integer(c_int64_t) :: seed
    real(c_double)     :: keff0
    integer(c_int32_t) :: events_per_launch
    integer(c_int32_t) :: block_size
  end type scn_eigen_settings

  interface
    !! engine lifetime (tables built host side from the same dictionaries SCONE reads)
    integer(c_int) function scn_create(problem, capacity, engine) bind(C, name="scn_create")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value        :: problem
      integer(c_int32_t), value :: capacity
      type(c_ptr)               :: engine
    end function scn_create

    subroutine scn_destroy(engine) bind(C, name="scn_destroy")
      import :: c_ptr
      type(c_ptr), value :: engine
    end subroutine scn_destroy

    type(c_ptr) function scn_last_error(engine) bind(C, name="scn_last_error")
      import :: c_ptr
      type(c_ptr), value :: engine
    end function scn_last_error

    !! one cycle of transport: replaces the !$omp parallel do over thisCycle
    integer(c_int) function scn_eigen_cycle(engine, bank_in, n_in, rng_state, k_eff, bank_out, &
                                            bank_out_cap, n_out, bin_scores, keff_scores) &
                                            bind(C, name="scn_eigen_cycle")
      import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr, scn_particle_state
      type(c_ptr), value                    :: engine
      type(scn_particle_state), intent(in)  :: bank_in(*)
      integer(c_int32_t), value             :: n_in
      integer(c_int64_t), value             :: rng_state
      real(c_double), value                 :: k_eff
      type(scn_particle_state), intent(out) :: bank_out(*)
      integer(c_int32_t), value             :: bank_out_cap
      integer(c_int32_t), intent(out)       :: n_out
      real(c_double), intent(out)           :: bin_scores(*)
      real(c_double), intent(out)           :: keff_scores(4)
    end function scn_eigen_cycle

    !! particleDungeon%normSize_Repr on the device
    integer(c_int) function scn_norm_size_repr(engine, bank, n, cap, tot_pop, rng_state) &
                                               bind(C, name="scn_norm_size_repr")
      import :: c_int, c_int32_t, c_int64_t, c_ptr, scn_particle_state
      type(c_ptr), value                      :: engine
      type(scn_particle_state), intent(inout) :: bank(*)
      integer(c_int32_t), intent(inout)       :: n
      integer(c_int32_t), value               :: cap
      integer(c_int32_t), value               :: tot_pop
      integer(c_int64_t), value               :: rng_state
    end function scn_norm_size_repr

    !! whole power iteration resident in HBM
    integer(c_int) function scn_eigen_run(engine, settings, initial_bank, result) &
                                          bind(C, name="scn_eigen_run")
      import :: c_int, c_ptr, scn_eigen_settings, scn_particle_state
      type(c_ptr), value                   :: engine
      type(scn_eigen_settings), intent(in) :: settings
      type(scn_particle_state), intent(in) :: initial_bank(*)
      type(c_ptr), value                   :: result
    end function scn_eigen_run
  end interface

  public :: scn_create, scn_destroy, scn_last_error, scn_eigen_cycle, scn_norm_size_repr, scn_eigen_run

end module scone_gpu_mod
