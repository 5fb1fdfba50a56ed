! ISO_C_BINDING layer over libopentbl_amd (include/opentbl.h).
!
! Module otbl_gpu binds the C entry points; the external subroutines after it
! keep the reference's names and argument lists (rhsp.F90:22-27,
! pois.f90:11, boun.f90:63, coeft.f90:10) so a driver written like bl_1.F90
! reaches the GPU kernels.  The fields live in device memory inside the
! context `otbl_box`; the array arguments of the drop-in subroutines are the
! caller's host buffers and are not read or written (state moves only through
! otbl_put / otbl_fetch).
module otbl_gpu
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: OTBL_FP64 = 0, OTBL_MIXED = 1
  integer(c_int), parameter :: OTBL_DT_EVERY_STEP = 0, OTBL_DT_FROZEN = 1
  integer(c_int), parameter :: OTBL_U = 0, OTBL_V = 1, OTBL_W = 2, OTBL_P = 3, &
       OTBL_RHSUPA = 4, OTBL_RHSVPA = 5, OTBL_RHSWPA = 6

  type, bind(C) :: otbl_params
     integer(c_int) :: nx, ny, nz, xout
     real(c_double) :: re, ax, az, cfl
     integer(c_int) :: precision, dt_mode
  end type otbl_params

  type(c_ptr), save :: otbl_box = c_null_ptr

  interface
     integer(c_int) function otbl_create(prm, y, ctx) bind(C, name="otbl_create")
       import :: c_int, c_double, c_ptr, otbl_params
       type(otbl_params), intent(in) :: prm
       real(c_double), intent(in) :: y(*)
       type(c_ptr), intent(out) :: ctx
     end function otbl_create
     integer(c_int) function otbl_destroy(ctx) bind(C, name="otbl_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function otbl_destroy
     type(c_ptr) function otbl_last_error() bind(C, name="otbl_last_error")
       import :: c_ptr
     end function otbl_last_error
     integer(c_int) function otbl_set_field(ctx, id, host) bind(C, name="otbl_set_field")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: id
       real(c_double), intent(in) :: host(*)
     end function otbl_set_field
     integer(c_int) function otbl_get_field(ctx, id, host) bind(C, name="otbl_get_field")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: id
       real(c_double), intent(out) :: host(*)
     end function otbl_get_field
     integer(c_int) function otbl_set_inflow(ctx, u0, v0, um) bind(C, name="otbl_set_inflow")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: u0(*), v0(*), um(*)
     end function otbl_set_inflow
     integer(c_int) function otbl_set_time(ctx, tiempo, timeinit, dt) bind(C, name="otbl_set_time")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), value :: tiempo, timeinit, dt
     end function otbl_set_time
     integer(c_int) function otbl_get_time(ctx, tiempo, dt) bind(C, name="otbl_get_time")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: tiempo, dt
     end function otbl_get_time
     integer(c_int) function otbl_rhsp(ctx, m) bind(C, name="otbl_rhsp")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: m
     end function otbl_rhsp
     integer(c_int) function otbl_outflow_correction(ctx) bind(C, name="otbl_outflow_correction")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function otbl_outflow_correction
     integer(c_int) function otbl_pois(ctx, varstep) bind(C, name="otbl_pois")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), value :: varstep
     end function otbl_pois
     integer(c_int) function otbl_step(ctx, nsteps) bind(C, name="otbl_step")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: nsteps
     end function otbl_step
     integer(c_int) function otbl_use_graphs(ctx, on) bind(C, name="otbl_use_graphs")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: on
     end function otbl_use_graphs
     integer(c_int) function otbl_sync(ctx) bind(C, name="otbl_sync")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function otbl_sync
     integer(c_int) function otbl_diag_enable(ctx, capacity) bind(C, name="otbl_diag_enable")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: capacity
     end function otbl_diag_enable
     integer(c_int) function otbl_diag_read(ctx, out, max_records, n_records, dropped) &
          bind(C, name="otbl_diag_read")
       import :: c_int, c_ptr, c_double, c_long_long
       type(c_ptr), value :: ctx
       real(c_double) :: out(*)
       integer(c_int), value :: max_records
       integer(c_int) :: n_records
       integer(c_long_long) :: dropped
     end function otbl_diag_read
  end interface

contains

  ! status check: the reference stops on errors (bl_1.F90:294-297)
  subroutine otbl_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    character(len=512) :: text
    integer :: i
    if (rc == 0) return
    call c_f_pointer(otbl_last_error(), msg, [512])
    text = ' '
    do i = 1, 512
       if (msg(i) == c_null_char) exit
       text(i:i) = msg(i)
    end do
    write(*,'(a,a,a,i0,a,a)') 'otbl: ', what, ' failed (status ', rc, '): ', trim(text)
    error stop 1
  end subroutine otbl_check

  ! box setup: iniciap_1 + coef + alloavar/alloabuff in device memory
  subroutine otbl_open(nx, ny, nz, xout, re, ax, az, cfl, y, dt_mode)
    integer, intent(in) :: nx, ny, nz, xout, dt_mode
    real(8), intent(in) :: re, ax, az, cfl, y(0:ny+1)
    type(otbl_params) :: prm
    prm = otbl_params(nx, ny, nz, xout, re, ax, az, cfl, OTBL_FP64, dt_mode)
    call otbl_check(otbl_create(prm, y, otbl_box), 'otbl_create')
  end subroutine otbl_open

  subroutine otbl_put(id, a)
    integer(c_int), intent(in) :: id
    real(8), intent(in) :: a(*)
    call otbl_check(otbl_set_field(otbl_box, id, a), 'otbl_set_field')
  end subroutine otbl_put

  subroutine otbl_fetch(id, a)
    integer(c_int), intent(in) :: id
    real(8), intent(out) :: a(*)
    call otbl_check(otbl_get_field(otbl_box, id, a), 'otbl_get_field')
  end subroutine otbl_fetch

  ! info_per_step (genflu.F90:247-295): the records the substeps appended to
  ! the device ring since the last call, one line each in the format of the
  ! reference's unit 62 (genflu.F90:287-288); capacity from otbl_diag_enable
  subroutine otbl_write_info(unit, capacity)
    integer, intent(in) :: unit, capacity
    real(8), allocatable :: rec(:)
    integer(c_int) :: n
    integer(c_long_long) :: lost
    integer :: r
    allocate(rec(49 * capacity))
    call otbl_check(otbl_diag_read(otbl_box, rec, int(capacity, c_int), n, lost), 'otbl_diag_read')
    do r = 0, n - 1
       write(unit, '(49(d22.14))') rec(49 * r + 1:49 * r + 49)
    end do
    call flush(unit)
    deallocate(rec)
  end subroutine otbl_write_info

  subroutine otbl_close()
    integer(c_int) :: rc
    rc = otbl_destroy(otbl_box)
    otbl_box = c_null_ptr
  end subroutine otbl_close

end module otbl_gpu

! ---- drop-in subroutines with the reference's interfaces ----

! rhsp.F90:22-27.  dt comes back from the GPU (set at substep 1).
subroutine rhsp(ut,vt,wt,pt,rhsupat,rhsvpat,rhswpat, resu,resut,resv,resvt,resw,reswt, &
     rhsut,rhsvt,rhswt, wki1,wki1t,wki2,wki2t,wki3,wki3t, wkp,wkf,wkpo,wkfo,bufuphy, &
     buf_corr,buf_corr2, dt,m,ical,istep,mpiid,communicator)
  use otbl_gpu
  implicit none
  real(8) :: ut(*),vt(*),wt(*),pt(*),rhsupat(*),rhsvpat(*),rhswpat(*)
  real(8) :: resu(*),resut(*),resv(*),resvt(*),resw(*),reswt(*),rhsut(*),rhsvt(*),rhswt(*)
  real(8) :: wki1(*),wki1t(*),wki2(*),wki2t(*),wki3(*),wki3t(*),wkp(*),wkf(*),wkpo(*),wkfo(*)
  real(8) :: bufuphy(*),buf_corr(*),buf_corr2(*)
  real(8) :: dt, t
  integer :: m, ical, istep, mpiid, communicator
  call otbl_check(otbl_rhsp(otbl_box, int(m, c_int)), 'rhsp')
  call otbl_check(otbl_get_time(otbl_box, t, dt), 'rhsp dt')
end subroutine rhsp

! boun.f90:63
subroutine outflow_correction(ut,vt,rhsupat,communicator)
  use otbl_gpu
  implicit none
  real(8) :: ut(*), vt(*), rhsupat(*)
  integer :: communicator
  call otbl_check(otbl_outflow_correction(otbl_box), 'outflow_correction')
end subroutine outflow_correction

! pois.f90:11
subroutine pois(ut,vt,wt,pt,res,rest,rt,varstep,mpiid,communicator)
  use otbl_gpu
  implicit none
  real(8) :: ut(*),vt(*),wt(*),pt(*),res(*),rest(*),rt(*)
  real(8) :: varstep
  integer :: mpiid, communicator
  call otbl_check(otbl_pois(otbl_box, varstep), 'pois')
end subroutine pois
