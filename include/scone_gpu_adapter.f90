!!
!! scone_gpu_adapter -- what a SCONE checkout needs around scone_gpu_mod to hand its loop to the
!! engine:
!!   state_toC / state_fromC    particleState <-> scn_particle_state.  particleState
!!                              (ParticleObjects/particle_class.f90:50-65) holds isMG as a logical;
!!                              the C record holds it as int32 (1 / 0).  The routines take the
!!                              components, so this file compiles without SCONE's modules:
!!                                call state_toC(p % wgt, p % r, p % dir, p % E, p % G, p % isMG, &
!!                                               p % type, p % time, p % lambda, p % matIdx, &
!!                                               p % cellIdx, p % uniqueID, p % collisionN, &
!!                                               p % broodID, cbank(n))
!!   gpu_tracking               the `tracking` entry of a transportOperatorGPU dictionary (or the
!!                              name of the CPU operator it stands in for) -> SCN_TRANSPORT_*; the
!!                              case new_transportOperator gains
!!                              (TransportOperator/transportOperatorFactory_func.f90:36-67)
!!   gpu_message                scn_last_error / scn_batch_last_error text for fatalError
!!
module scone_gpu_adapter
  use iso_c_binding
  use scone_gpu_mod
  implicit none
  private

  public :: state_toC, state_fromC, gpu_tracking, gpu_message

contains

  !! particleState components -> bind(C) record (pad fields zeroed)
  pure subroutine state_toC(wgt, r, dir, E, G, isMG, type, time, lambda, matIdx, cellIdx, uniqueID, &
                            collisionN, broodID, c)
    real(c_double), intent(in)            :: wgt, r(3), dir(3), E, time, lambda
    integer(c_int32_t), intent(in)        :: G, type, matIdx, cellIdx, uniqueID, collisionN, broodID
    logical, intent(in)                   :: isMG
    type(scn_particle_state), intent(out) :: c
    c % wgt = wgt
    c % r = r
    c % dir = dir
    c % E = E
    c % G = G
    c % isMG = merge(1_c_int32_t, 0_c_int32_t, isMG)
    c % type = type
    c % pad0 = 0_c_int32_t
    c % time = time
    c % lambda = lambda
    c % matIdx = matIdx
    c % cellIdx = cellIdx
    c % uniqueID = uniqueID
    c % collisionN = collisionN
    c % broodID = broodID
    c % pad1 = 0_c_int32_t
  end subroutine state_toC

  !! bind(C) record -> particleState components (isMG: any non-zero value is .true.)
  pure subroutine state_fromC(c, wgt, r, dir, E, G, isMG, type, time, lambda, matIdx, cellIdx, uniqueID, &
                              collisionN, broodID)
    type(scn_particle_state), intent(in) :: c
    real(c_double), intent(out)          :: wgt, r(3), dir(3), E, time, lambda
    integer(c_int32_t), intent(out)      :: G, type, matIdx, cellIdx, uniqueID, collisionN, broodID
    logical, intent(out)                 :: isMG
    wgt = c % wgt
    r = c % r
    dir = c % dir
    E = c % E
    G = c % G
    isMG = c % isMG /= 0_c_int32_t
    type = c % type
    time = c % time
    lambda = c % lambda
    matIdx = c % matIdx
    cellIdx = c % cellIdx
    uniqueID = c % uniqueID
    collisionN = c % collisionN
    broodID = c % broodID
  end subroutine state_fromC

  !! transportOperatorGPU: 'tracking' = ST | DT | HT, or the name of the CPU operator
  !! (transportOperatorST / DT / HT); 0 if the name is none of these
  pure integer(c_int32_t) function gpu_tracking(name) result(kind)
    character(*), intent(in) :: name
    select case (trim(adjustl(name)))
    case ('ST', 'transportOperatorST')
      kind = SCN_TRANSPORT_ST
    case ('DT', 'transportOperatorDT')
      kind = SCN_TRANSPORT_DT
    case ('HT', 'transportOperatorHT')
      kind = SCN_TRANSPORT_HT
    case default
      kind = 0_c_int32_t
    end select
  end function gpu_tracking

  !! the NUL-terminated message behind a const char* (an empty string for a null pointer)
  function gpu_message(p) result(msg)
    type(c_ptr), intent(in)                :: p
    character(:), allocatable              :: msg
    character(kind=c_char), pointer        :: s(:)
    integer                                :: n
    interface
      integer(c_size_t) function c_strlen(str) bind(C, name="strlen")
        import :: c_size_t, c_ptr
        type(c_ptr), value :: str
      end function c_strlen
    end interface
    if (.not. c_associated(p)) then
      msg = ''
      return
    end if
    n = int(c_strlen(p))
    call c_f_pointer(p, s, [n])
    allocate(character(n) :: msg)
    msg = transfer(s(1:n), msg)
  end function gpu_message

end module scone_gpu_adapter
