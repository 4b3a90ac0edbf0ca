!!
!! scone_gpu_mod -- iso_c_binding interface to libscone_amd.so (include/scone_gpu.h) for SCONE.
!!
!! A SCONE physics package uses these interfaces in place of its per-history OpenMP loop
!! (PhysicsPackages/eigenPhysicsPackage_class.f90:211-252, fixedSourcePhysicsPackage_class.f90:
!! 187-275).  Every derived type mirrors its C struct field for field (bind(C) gives the C layout);
!! logical flags travel as integer(c_int32_t), table arrays as type(c_ptr) (c_loc of a target array).
!! Conversions between SCONE's own types and these records live in scone_gpu_adapter.
!!
module scone_gpu_mod
  use iso_c_binding
  implicit none
  private

  !! error codes (SCN_ERR_*)
  integer(c_int), parameter, public :: SCN_OK = 0, SCN_ERR_ARG = 1, SCN_ERR_HIP = 2, SCN_ERR_OVERFLOW = 3, &
                                       SCN_ERR_MATERIAL = 4, SCN_ERR_NAN = 5, SCN_ERR_MAJORANT = 6, &
                                       SCN_ERR_NORM = 7, SCN_ERR_SAMPLING = 8, SCN_ERR_UNSUPPORTED = 9, &
                                       SCN_ERR_NO_DEVICE = 10, SCN_ERR_BUSY = 11
  !! reference constants (SharedModules/universalVariables.f90)
  integer(c_int32_t), parameter, public :: SCN_MAX_NEST = 8, SCN_OUTSIDE_MAT = 0, SCN_VOID_MAT = 2147483647, &
                                           SCN_UNDEF_MAT = 2147483646, SCN_OVERLAP_MAT = 2147483645
  integer(c_int32_t), parameter, public :: SCN_BC_VACUUM = 0, SCN_BC_REFLECTIVE = 1, SCN_BC_PERIODIC = 2
  integer(c_int32_t), parameter, public :: SCN_TRANSPORT_ST = 1, SCN_TRANSPORT_DT = 2, SCN_TRANSPORT_HT = 3
  integer(c_int32_t), parameter, public :: SCN_COLL_STD = 0, SCN_COLL_IMP = 1
  integer(c_int32_t), parameter, public :: SCN_UNI_ROOT = 1, SCN_UNI_LAT = 2, SCN_UNI_PIN = 3, SCN_UNI_CELL = 4
  integer(c_int32_t), parameter, public :: SCN_UCELL = -1, SCN_UCELL_UNION = 1073741824, &
                                           SCN_UCELL_INTERSECT = 1073741825, SCN_UCELL_OPEN = 1073741826, &
                                           SCN_UCELL_CLOSE = 1073741827, SCN_UCELL_COMPLEM = 1073741828
  integer(c_int32_t), parameter, public :: SCN_SURF_APLANE = 1, SCN_SURF_PLANE = 2, SCN_SURF_SPHERE = 3, &
                                           SCN_SURF_CYLINDER = 4, SCN_SURF_QUADRIC = 5, SCN_SURF_BOX = 6, &
                                           SCN_SURF_TRUNC_CYLINDER = 7, SCN_SURF_TRUNC_CONE = 8, SCN_SURF_WEDGE = 9
  integer(c_int32_t), parameter, public :: SCN_MAP_SPACE = 1, SCN_MAP_MATERIAL = 2, SCN_MAP_ENERGY = 3, &
                                           SCN_MAP_CELL = 4, SCN_GRID_LIN = 1, SCN_GRID_UNSTRUCT = 3
  integer(c_int32_t), parameter, public :: SCN_MT_FLUX = 0, SCN_MT_TOTAL = -1, SCN_MT_DISAPPEARANCE = -2, &
                                           SCN_MT_ESCATTER = -3, SCN_MT_IESCATTER = -4, SCN_MT_FISSION = -6, &
                                           SCN_MT_NUFISSION = -7, SCN_MT_KAPPA = -80, SCN_MT_ALLSCATTER = -20, &
                                           SCN_MT_ABSORPTION = -21, SCN_MT_NONELASTIC = -22
  integer(c_int32_t), parameter, public :: SCN_CLERK_COLLISION = 0, SCN_CLERK_TRACK = 1
  integer(c_int32_t), parameter, public :: SCN_SRC_POINT = 1, SCN_SRC_MATERIAL = 2
  integer(c_int32_t), parameter, public :: SCN_SLOT_FREE = 0, SCN_SLOT_RUNNING = 1, SCN_SLOT_DONE = 2
  integer(c_int32_t), parameter, public :: SCN_NOALLOC_MAX_SITES = 1048576

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

  !! box / squareCylinder border (Geometry/Surfaces/CompositeSurfaces)
  type, public, bind(C) :: scn_box
    integer(c_int32_t) :: n_axes
    integer(c_int32_t) :: axes(3)
    integer(c_int32_t) :: bc(6)
    real(c_double)     :: origin(3)
    real(c_double)     :: halfwidth(3)
  end type scn_box

  !! cell and border surfaces (Geometry/Surfaces)
  type, public, bind(C) :: scn_surface
    integer(c_int32_t) :: type
    integer(c_int32_t) :: axis
    real(c_double)     :: tol
    real(c_double)     :: c(12)
  end type scn_surface

  !! root / lattice / pin / cell universes (Geometry/Universes)
  type, public, bind(C) :: scn_universe
    integer(c_int32_t) :: type
    integer(c_int32_t) :: n_local
    real(c_double)     :: origin(3)
    integer(c_int32_t) :: border
    integer(c_int32_t) :: size(3)
    integer(c_int32_t) :: out_local
    integer(c_int32_t) :: offset_mode
    integer(c_int32_t) :: offset_map
    real(c_double)     :: pitch(3)
    real(c_double)     :: corner(3)
    real(c_double)     :: a_bar(3)
    real(c_double)     :: outline_hw(3)
    integer(c_int32_t) :: n_annuli
    integer(c_int32_t) :: annuli
    integer(c_int32_t) :: n_cells
    integer(c_int32_t) :: cells
    integer(c_int32_t) :: border_surf
    integer(c_int32_t) :: rotated
    integer(c_int32_t) :: global
    integer(c_int32_t) :: check_overlap
    real(c_double)     :: rot(9)
  end type scn_universe

  !! multigroup data (NuclearData/mgNeutronData/baseMgNeutron): pointers to [mat][G][6] xs,
  !! [mat][G_in][G_out] P0 / prod / P1, [mat][G] chi, [mat] fissile, [G] majorant
  type, public, bind(C) :: scn_mg_data
    integer(c_int32_t) :: n_mat
    integer(c_int32_t) :: n_groups
    type(c_ptr)        :: xs
    type(c_ptr)        :: P0
    type(c_ptr)        :: prod
    type(c_ptr)        :: chi
    type(c_ptr)        :: fissile
    type(c_ptr)        :: majorant
    real(c_double)     :: collision_xs
    type(c_ptr)        :: P1
  end type scn_mg_data

  !! one tally map (Tallies/TallyMaps)
  type, public, bind(C) :: scn_map1d
    integer(c_int32_t) :: type
    integer(c_int32_t) :: axis
    integer(c_int32_t) :: nbins
    integer(c_int32_t) :: grid_type
    integer(c_int32_t) :: data
  end type scn_map1d

  !! collisionClerk / trackClerk (Tallies/TallyClerks) at its scoreMemory address
  type, public, bind(C) :: scn_clerk
    integer(c_int32_t) :: gtop, glow
    integer(c_int32_t) :: n_maps
    integer(c_int32_t) :: handle_virtual
    type(scn_map1d)    :: maps(3)
    integer(c_int32_t) :: multi(3)
    integer(c_int32_t) :: width
    integer(c_int32_t) :: resp(8)
    integer(c_int32_t) :: kind
    integer(c_int64_t) :: mem_addr
    integer(c_int64_t) :: size
  end type scn_clerk

  !! every table the loop reads
  type, public, bind(C) :: scn_problem
    integer(c_int32_t) :: transport
    integer(c_int32_t) :: st_cache
    integer(c_int32_t) :: n_universes
    integer(c_int32_t) :: root_uni
    type(c_ptr)        :: universes
    integer(c_int32_t) :: n_boxes
    integer(c_int32_t) :: border_box
    type(c_ptr)        :: boxes
    integer(c_int32_t) :: graph_size
    type(c_ptr)        :: graph_fill
    type(c_ptr)        :: graph_id
    integer(c_int32_t) :: n_real
    type(c_ptr)        :: real_data
    integer(c_int32_t) :: n_int
    type(c_ptr)        :: int_data
    type(scn_mg_data)  :: mg
    integer(c_int32_t) :: n_clerks
    type(c_ptr)        :: clerks
    integer(c_int64_t) :: n_bins
    integer(c_int64_t) :: norm_bin
    real(c_double)     :: norm_value
    real(c_double)     :: ht_cutoff
    integer(c_int32_t) :: n_surfaces
    type(c_ptr)        :: surfaces
    integer(c_int32_t) :: collision
    integer(c_int32_t) :: max_split
    integer(c_int32_t) :: ww_n
    integer(c_int32_t) :: fixed_point_tally
    type(c_ptr)        :: ww_map
    type(c_ptr)        :: ww_lower
    type(c_ptr)        :: ww_upper
    real(c_double)     :: ww_survival
  end type scn_problem

  type, public, bind(C) :: scn_eigen_settings
    integer(c_int32_t) :: pop
    integer(c_int32_t) :: n_inactive
    integer(c_int32_t) :: n_active
    integer(c_int32_t) :: inactive_tally
    integer(c_int64_t) :: seed
    real(c_double)     :: keff0
    integer(c_int32_t) :: events_per_launch
    integer(c_int32_t) :: block_size
    integer(c_int32_t) :: cycle_graph
    integer(c_int32_t) :: refill
    real(c_double)     :: select_window
  end type scn_eigen_settings

  type, public, bind(C) :: scn_eigen_result
    real(c_double)     :: k_analog(2)
    real(c_double)     :: k_implicit(2)
    real(c_double)     :: imp(2, 4)        !! C imp[4][2]
    type(c_ptr)        :: k_history
    type(c_ptr)        :: cycle_time
    type(c_ptr)        :: bins_active
    type(c_ptr)        :: bins_inactive
    real(c_double)     :: transport_time
    integer(c_int64_t) :: n_collisions
    integer(c_int64_t) :: n_events
    integer(c_int64_t) :: n_histories
  end type scn_eigen_result

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
    integer(c_int32_t) :: adapt
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
    integer(c_int32_t) :: refill
    integer(c_int32_t) :: pad0
  end type scn_batch_settings

  type, public, bind(C) :: scn_stats
    real(c_double)     :: transport_kernel_ms
    integer(c_int64_t) :: transport_launches
    integer(c_int64_t) :: n_histories
    integer(c_int64_t) :: n_segments
    integer(c_int64_t) :: n_collisions
    integer(c_int64_t) :: n_sites
    integer(c_int64_t) :: n_scores
    integer(c_int64_t) :: n_cycles
    real(c_double)     :: cycle_ms
  end type scn_stats

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

    integer(c_int) function scn_device_info(name, name_len, n_cu) bind(C, name="scn_device_info")
      import :: c_int, c_int32_t, c_char
      character(kind=c_char), intent(out) :: name(*)
      integer(c_int32_t), value           :: name_len
      integer(c_int32_t), intent(out)     :: n_cu
    end function scn_device_info

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

    !! whole power iteration resident in HBM, in one call or split into start / step / finish
    integer(c_int) function scn_eigen_run(engine, settings, initial_bank, result) &
                                          bind(C, name="scn_eigen_run")
      import :: c_int, c_ptr, scn_eigen_settings, scn_particle_state, scn_eigen_result
      type(c_ptr), value                    :: engine
      type(scn_eigen_settings), intent(in)  :: settings
      type(scn_particle_state), intent(in)  :: initial_bank(*)
      type(scn_eigen_result), intent(inout) :: result
    end function scn_eigen_run

    integer(c_int) function scn_eigen_start(engine, settings, initial_bank) bind(C, name="scn_eigen_start")
      import :: c_int, c_ptr, scn_eigen_settings, scn_particle_state
      type(c_ptr), value                   :: engine
      type(scn_eigen_settings), intent(in) :: settings
      type(scn_particle_state), intent(in) :: initial_bank(*)
    end function scn_eigen_start

    integer(c_int) function scn_eigen_step(engine, n_cycles, done) bind(C, name="scn_eigen_step")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value              :: engine
      integer(c_int32_t), value       :: n_cycles
      integer(c_int32_t), intent(out) :: done
    end function scn_eigen_step

    integer(c_int) function scn_eigen_finish(engine, result) bind(C, name="scn_eigen_finish")
      import :: c_int, c_ptr, scn_eigen_result
      type(c_ptr), value                    :: engine
      type(scn_eigen_result), intent(inout) :: result
    end function scn_eigen_finish

    !! fissionSource%sampleParticle for particles 1..n
    integer(c_int) function scn_fission_source(engine, n, rng_state, bottom, top, g_in, attempts, bank) &
                                               bind(C, name="scn_fission_source")
      import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr, scn_particle_state
      type(c_ptr), value                    :: engine
      integer(c_int32_t), value             :: n
      integer(c_int64_t), value             :: rng_state
      real(c_double), intent(in)            :: bottom(3), top(3)
      integer(c_int32_t), value             :: g_in
      integer(c_int32_t), value             :: attempts
      type(scn_particle_state), intent(out) :: bank(*)
    end function scn_fission_source

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

    integer(c_int) function scn_batch_states(batch, states, n) bind(C, name="scn_batch_states")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value              :: batch
      integer(c_int32_t), intent(out) :: states(*)
      integer(c_int32_t), value       :: n
    end function scn_batch_states

    integer(c_int) function scn_batch_release(batch, slot) bind(C, name="scn_batch_release")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value        :: batch
      integer(c_int32_t), value :: slot
    end function scn_batch_release

    integer(c_int) function scn_batch_set_window(batch, slot, window) bind(C, name="scn_batch_set_window")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value        :: batch
      integer(c_int32_t), value :: slot
      integer(c_int32_t), value :: window
    end function scn_batch_set_window

    integer(c_int64_t) function scn_batch_secondary_bytes(transport, buffer) bind(C, name="scn_batch_secondary_bytes")
      import :: c_int32_t, c_int64_t
      integer(c_int32_t), value :: transport
      integer(c_int32_t), value :: buffer
    end function scn_batch_secondary_bytes

    integer(c_int) function scn_batch_bin(batch, slot, bin, result, n_batches) bind(C, name="scn_batch_bin")
      import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
      type(c_ptr), value              :: batch
      integer(c_int32_t), value       :: slot
      integer(c_int64_t), value       :: bin
      real(c_double), intent(out)     :: result(2)
      integer(c_int32_t), intent(out) :: n_batches
    end function scn_batch_bin

    integer(c_int) function scn_batch_stats(batch, out) bind(C, name="scn_batch_stats")
      import :: c_int, c_ptr, scn_stats
      type(c_ptr), value           :: batch
      type(scn_stats), intent(out) :: out
    end function scn_batch_stats

    !! population control that allocates nothing: device banks, scratch in a static device array
    integer(c_int) function scn_norm_size_repr_noalloc(bank_in, n, bank_out, out_cap, tot_pop, rng_state, &
                                                       stream, n_out) bind(C, name="scn_norm_size_repr_noalloc")
      import :: c_int, c_int32_t, c_int64_t, c_ptr
      type(c_ptr), value              :: bank_in
      integer(c_int32_t), value       :: n
      type(c_ptr), value              :: bank_out
      integer(c_int32_t), value       :: out_cap
      integer(c_int32_t), value       :: tot_pop
      integer(c_int64_t), value       :: rng_state
      type(c_ptr), value              :: stream
      integer(c_int32_t), intent(out) :: n_out
    end function scn_norm_size_repr_noalloc

    integer(c_int) function scn_get_stats(engine, out) bind(C, name="scn_get_stats")
      import :: c_int, c_ptr, scn_stats
      type(c_ptr), value           :: engine
      type(scn_stats), intent(out) :: out
    end function scn_get_stats

    integer(c_int) function scn_reset_stats(engine) bind(C, name="scn_reset_stats")
      import :: c_int, c_ptr
      type(c_ptr), value :: engine
    end function scn_reset_stats
  end interface

  public :: scn_create, scn_destroy, scn_last_error, scn_device_info, scn_eigen_cycle, scn_norm_size_repr
  public :: scn_eigen_run, scn_eigen_start, scn_eigen_step, scn_eigen_finish, scn_fission_source
  public :: scn_batch_create, scn_batch_destroy, scn_batch_last_error, scn_batch_add, scn_batch_cycle, &
            scn_batch_query, scn_batch_states, scn_batch_release, scn_batch_stats, scn_batch_set_window, &
            scn_batch_bin, scn_batch_secondary_bytes
  public :: scn_norm_size_repr_noalloc, scn_get_stats, scn_reset_stats

end module scone_gpu_mod
