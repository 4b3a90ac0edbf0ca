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
  integer(c_int32_t), parameter, public :: SCN_ERR_BUSY = 11, SCN_SRC_POINT = 1, SCN_SRC_MATERIAL = 2
  integer(c_int32_t), parameter, public :: SCN_SLOT_FREE = 0, SCN_SLOT_RUNNING = 1, SCN_SLOT_DONE = 2

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

  !! pointSource / materialSource parameters (ParticleObjects/Source/*.f90 init)
  type, public, bind(C) :: scn_source
    integer(c_int32_t) :: type
    integer(c_int32_t) :: isotropic
    integer(c_int32_t) :: G
    integer(c_int32_t) :: mat
    integer(c_int32_t) :: n_probG
    integer(c_int32_t) :: pad
    real(c_double)     :: r(3)
    real(c_double)     :: dir(3)
    real(c_double)     :: bottom(3)
    real(c_double)     :: top(3)
    real(c_double)     :: t_low, t_high
    type(c_ptr)        :: probG
  end type scn_source

  !! one fixedSourcePhysicsPackage solve
  type, public, bind(C) :: scn_fixed_request
    type(c_ptr)        :: problem
    type(c_ptr)        :: source
    integer(c_int32_t) :: pop
    integer(c_int32_t) :: n_batches
    integer(c_int64_t) :: seed
    integer(c_int64_t) :: stop_bin
    real(c_double)     :: stop_rel
    integer(c_int32_t) :: record
    integer(c_int32_t) :: buffer
    integer(c_int32_t) :: window
    integer(c_int32_t) :: pad0
  end type scn_fixed_request

  type, public, bind(C) :: scn_fixed_result
    integer(c_int32_t) :: state
    integer(c_int32_t) :: n_batches
    integer(c_int64_t) :: n_bins
    integer(c_int64_t) :: n_histories
    integer(c_int64_t) :: n_collisions
    integer(c_int64_t) :: n_segments
  end type scn_fixed_result

  type, public, bind(C) :: scn_batch_settings
    integer(c_int32_t) :: transport
    integer(c_int32_t) :: max_slots
    integer(c_int64_t) :: hbm_budget
    integer(c_int32_t) :: tile
    integer(c_int32_t) :: k0
    integer(c_int32_t) :: n_launch
    integer(c_int32_t) :: use_graph
  end type scn_batch_settings

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

    !! fixed-source engine with resident slots: replaces the history loop of
    !! fixedSourcePhysicsPackage%cycles (PhysicsPackages/fixedSourcePhysicsPackage_class.f90:187-275)
    integer(c_int) function scn_batch_create(settings, batch) bind(C, name="scn_batch_create")
      import :: c_int, c_ptr, scn_batch_settings
      type(scn_batch_settings), intent(in) :: settings
      type(c_ptr)                          :: batch
    end function scn_batch_create

    subroutine scn_batch_destroy(batch) bind(C, name="scn_batch_destroy")
      import :: c_ptr
      type(c_ptr), value :: batch
    end subroutine scn_batch_destroy

    type(c_ptr) function scn_batch_last_error(batch) bind(C, name="scn_batch_last_error")
      import :: c_ptr
      type(c_ptr), value :: batch
    end function scn_batch_last_error

    integer(c_int) function scn_batch_add(batch, request, slot) bind(C, name="scn_batch_add")
      import :: c_int, c_int32_t, c_ptr, scn_fixed_request
      type(c_ptr), value                  :: batch
      type(scn_fixed_request), intent(in) :: request
      integer(c_int32_t), intent(out)     :: slot
    end function scn_batch_add

    integer(c_int) function scn_batch_cycle(batch, n_cycles, n_running) bind(C, name="scn_batch_cycle")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value              :: batch
      integer(c_int32_t), value       :: n_cycles
      integer(c_int32_t), intent(out) :: n_running
    end function scn_batch_cycle

    integer(c_int) function scn_batch_query(batch, slot, result, bins, hist_rng, hist_coll) &
                                            bind(C, name="scn_batch_query")
      import :: c_int, c_int32_t, c_ptr, scn_fixed_result
      type(c_ptr), value                   :: batch
      integer(c_int32_t), value            :: slot
      type(scn_fixed_result), intent(out)  :: result
      type(c_ptr), value                   :: bins, hist_rng, hist_coll
    end function scn_batch_query

    integer(c_int) function scn_batch_release(batch, slot) bind(C, name="scn_batch_release")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value        :: batch
      integer(c_int32_t), value :: slot
    end function scn_batch_release
  end interface

  public :: scn_create, scn_destroy, scn_last_error, scn_eigen_cycle, scn_norm_size_repr, scn_eigen_run
  public :: scn_batch_create, scn_batch_destroy, scn_batch_last_error, scn_batch_add, scn_batch_cycle, &
            scn_batch_query, scn_batch_release

end module scone_gpu_mod
