! ISO_C_BINDING layer over libopentbl_amd (include/opentbl.h).
!
! Module otbl_gpu binds the C entry points; the external subroutines after it
! keep the reference's names and argument lists (rhsp.F90:22-27,
! pois.f90:11, boun.f90:63, lip.f90:52-597, ffts.f90:27-181,
! change.f90:224-275, coeft.f90:10,1244,1338, pois.f90:211) so code written
! like the reference reaches the GPU kernels.
!
! The box state lives in device memory inside the context `otbl_box`.  With
! otbl_mirror = .true. (the default) rhsp, outflow_correction and pois copy
! the caller's field arrays to the device before the call and the fields the
! reference routine updates back after it, so the caller's arrays hold what
! the reference's callees leave in them; a driver that keeps the state on the
! device (otbl_put / otbl_fetch around the time loop) sets it to .false.
! The operator drop-ins (lip, ffts, soltriy, coefinv) take the caller's host
! arrays in the reference's layout, run on the GPU and write their results
! back; the coefficient arrays they are passed are the ones coef built, and
! the context's own device copy of the same set is used, selected by the
! flag that selects it in the reference (loguv, vel, chv, jee).
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
  ! box sizes (set by otbl_open) and the field mirroring of the drop-ins
  integer, save :: otbl_nx = 0, otbl_ny = 0, otbl_nz = 0, otbl_nz1 = 0
  real(8), save :: otbl_re = 0d0
  real(8), allocatable, save :: otbl_y(:)
  logical, save :: otbl_mirror = .true.
  ! device scratch of the operator drop-ins: three buffers, grown on demand
  type(c_ptr), save :: otbl_dbuf(3) = c_null_ptr
  integer(c_size_t), save :: otbl_dcap(3) = 0

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
     integer(c_int) function otbl_interpxx(ctx, gen, geni, R, which_u) bind(C, name="otbl_interpxx")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, gen, geni
       integer(c_int), value :: R, which_u
     end function otbl_interpxx
     integer(c_int) function otbl_difvisxx(ctx, prod, vel, out, R, which_u) bind(C, name="otbl_difvisxx")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, prod, vel, out
       integer(c_int), value :: R, which_u
     end function otbl_difvisxx
     integer(c_int) function otbl_interpyy(ctx, gen, geni, loguv, nplanes) bind(C, name="otbl_interpyy")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, gen, geni
       integer(c_int), value :: loguv, nplanes
     end function otbl_interpyy
     integer(c_int) function otbl_differzy(ctx, geni, res, which, nplanes) bind(C, name="otbl_differzy")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, geni, res
       integer(c_int), value :: which, nplanes
     end function otbl_differzy
     integer(c_int) function otbl_vistyy(ctx, u, res, R, nplanes) bind(C, name="otbl_vistyy")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, u, res
       integer(c_int), value :: R, nplanes
     end function otbl_vistyy
     integer(c_int) function otbl_implzy(ctx, u, res, R, rkk, nplanes) bind(C, name="otbl_implzy")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, u, res
       integer(c_int), value :: R, nplanes
       real(c_double), value :: rkk
     end function otbl_implzy
     integer(c_int) function otbl_fourxz(ctx, fou, phys, iopt, nrows) bind(C, name="otbl_fourxz")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, fou, phys
       integer(c_int), value :: iopt, nrows
     end function otbl_fourxz
     integer(c_int) function otbl_cosft_lines(ctx, lines, nlines, iopt) bind(C, name="otbl_cosft_lines")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, lines
       integer(c_int), value :: nlines, iopt
     end function otbl_cosft_lines
     integer(c_int) function otbl_soltriy(ctx, u, us, ii, a, ny1) bind(C, name="otbl_soltriy")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx, u, us
       integer(c_int), value :: ii, ny1
       real(c_double) :: a(*)
     end function otbl_soltriy
     integer(c_int) function otbl_coefinv(ctx, a, npter1, neq) bind(C, name="otbl_coefinv")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double) :: a(*)
       integer(c_int), value :: npter1, neq
     end function otbl_coefinv
     integer(c_int) function otbl_soltrix(ctx, u, nstr, nlen, neq, us, npter1, npter2, a, iflag) &
          bind(C, name="otbl_soltrix")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double) :: u(*), us(*), a(*)
       integer(c_int), value :: nstr, nlen, neq, npter1, npter2
       integer(c_int) :: iflag
     end function otbl_soltrix
     integer(c_int) function otbl_coefb(n, flagv, a, b, c, d, bcc) bind(C, name="otbl_coefb")
       import :: c_int, c_double
       integer(c_int), value :: n, flagv
       real(c_double), value :: a, b, c, d
       real(c_double) :: bcc(*)
     end function otbl_coefb
     integer(c_int) function otbl_get_coef(ctx, name, out, rows, cols) bind(C, name="otbl_get_coef")
       import :: c_int, c_ptr, c_double, c_char
       type(c_ptr), value :: ctx
       character(kind=c_char) :: name(*)
       real(c_double) :: out(*)
       integer(c_int) :: rows, cols
     end function otbl_get_coef
     integer(c_int) function otbl_dev_alloc(bytes, ptr) bind(C, name="otbl_dev_alloc")
       import :: c_int, c_ptr, c_size_t
       integer(c_size_t), value :: bytes
       type(c_ptr) :: ptr
     end function otbl_dev_alloc
     integer(c_int) function otbl_dev_free(ptr) bind(C, name="otbl_dev_free")
       import :: c_int, c_ptr
       type(c_ptr), value :: ptr
     end function otbl_dev_free
     integer(c_int) function otbl_h2d(dst, src, bytes) bind(C, name="otbl_h2d")
       import :: c_int, c_ptr, c_double, c_size_t
       type(c_ptr), value :: dst
       real(c_double), intent(in) :: src(*)
       integer(c_size_t), value :: bytes
     end function otbl_h2d
     integer(c_int) function otbl_d2h(dst, src, bytes) bind(C, name="otbl_d2h")
       import :: c_int, c_ptr, c_double, c_size_t
       real(c_double) :: dst(*)
       type(c_ptr), value :: src
       integer(c_size_t), value :: bytes
     end function otbl_d2h
     integer(c_int) function otbl_stats_enable(ctx, on) bind(C, name="otbl_stats_enable")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: on
     end function otbl_stats_enable
     integer(c_int) function otbl_stats_schedule(ctx, avcrt, stest) bind(C, name="otbl_stats_schedule")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: avcrt, stest
     end function otbl_stats_schedule
     integer(c_int) function otbl_stats_request(ctx) bind(C, name="otbl_stats_request")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function otbl_stats_request
     integer(c_int) function otbl_statsp(ctx) bind(C, name="otbl_statsp")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function otbl_statsp
     integer(c_int) function otbl_stats_get(ctx, id, out) bind(C, name="otbl_stats_get")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: id
       real(c_double) :: out(*)
     end function otbl_stats_get
     integer(c_int) function otbl_stats_ical(ctx, ical) bind(C, name="otbl_stats_ical")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int) :: ical
     end function otbl_stats_ical
     integer(c_int) function otbl_stats_reset(ctx) bind(C, name="otbl_stats_reset")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function otbl_stats_reset
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
    otbl_nx = nx
    otbl_ny = ny
    otbl_nz = nz
    otbl_nz1 = 2 * (nz / 3)
    otbl_re = re
    if (allocated(otbl_y)) deallocate(otbl_y)
    allocate(otbl_y(0:ny + 1))
    otbl_y = y
  end subroutine otbl_open

  subroutine otbl_need_box(what)
    character(*), intent(in) :: what
    if (.not. c_associated(otbl_box)) then
       write(*,'(a,a,a)') 'otbl: ', what, ' called before otbl_open'
       error stop 1
    end if
  end subroutine otbl_need_box

  subroutine otbl_fail(what)
    character(*), intent(in) :: what
    write(*,'(a,a)') 'otbl: ', what
    error stop 1
  end subroutine otbl_fail

  ! device scratch buffer k (1..3) of at least n doubles
  type(c_ptr) function otbl_scratch(k, n)
    integer, intent(in) :: k
    integer(c_size_t), intent(in) :: n
    integer(c_size_t) :: bytes
    bytes = 8_c_size_t * max(n, 1_c_size_t)
    if (otbl_dcap(k) < bytes) then
       if (c_associated(otbl_dbuf(k))) call otbl_check(otbl_dev_free(otbl_dbuf(k)), 'otbl_dev_free')
       call otbl_check(otbl_dev_alloc(bytes, otbl_dbuf(k)), 'otbl_dev_alloc')
       otbl_dcap(k) = bytes
    end if
    otbl_scratch = otbl_dbuf(k)
  end function otbl_scratch

  ! host array a (n doubles) to scratch buffer k, and back
  type(c_ptr) function otbl_up(k, a, n)
    integer, intent(in) :: k
    integer(c_size_t), intent(in) :: n
    real(8), intent(in) :: a(*)
    otbl_up = otbl_scratch(k, n)
    call otbl_check(otbl_h2d(otbl_up, a, 8_c_size_t * n), 'otbl_h2d')
  end function otbl_up

  subroutine otbl_down(k, a, n)
    integer, intent(in) :: k
    integer(c_size_t), intent(in) :: n
    real(8), intent(out) :: a(*)
    call otbl_check(otbl_d2h(a, otbl_dbuf(k), 8_c_size_t * n), 'otbl_d2h')
  end subroutine otbl_down

  ! a coefficient array of the context by its reference name (coef's
  ! output, e.g. 'ayp', 'inxu'), in the reference's shape
  subroutine otbl_coef(name, out)
    character(*), intent(in) :: name
    real(8), intent(out) :: out(*)
    integer(c_int) :: r, c
    call otbl_check(otbl_get_coef(otbl_box, trim(name) // c_null_char, out, r, c), 'otbl_get_coef ' // name)
  end subroutine otbl_coef

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

  ! C-callable forms of otbl_open / otbl_close and the mirror switch, for
  ! callers that are not Fortran (the tests load the drop-ins from Python)
  subroutine otbl_f_open(nx, ny, nz, xout, re, ax, az, cfl, y, dt_mode) bind(C, name="otbl_f_open")
    integer(c_int), value :: nx, ny, nz, xout, dt_mode
    real(c_double), value :: re, ax, az, cfl
    real(c_double), intent(in) :: y(0:ny + 1)
    call otbl_open(int(nx), int(ny), int(nz), int(xout), re, ax, az, cfl, y, int(dt_mode))
  end subroutine otbl_f_open

  subroutine otbl_f_close() bind(C, name="otbl_f_close")
    call otbl_close()
  end subroutine otbl_f_close

  subroutine otbl_f_mirror(on) bind(C, name="otbl_f_mirror")
    integer(c_int), value :: on
    otbl_mirror = on /= 0
  end subroutine otbl_f_mirror

  ! the context, for C entry points not wrapped here (set_field, set_inflow)
  type(c_ptr) function otbl_f_box() bind(C, name="otbl_f_box")
    otbl_f_box = otbl_box
  end function otbl_f_box

  subroutine otbl_close()
    integer(c_int) :: rc
    integer :: k
    do k = 1, 3
       if (c_associated(otbl_dbuf(k))) rc = otbl_dev_free(otbl_dbuf(k))
       otbl_dbuf(k) = c_null_ptr
       otbl_dcap(k) = 0
    end do
    rc = otbl_destroy(otbl_box)
    otbl_box = c_null_ptr
  end subroutine otbl_close

end module otbl_gpu


! ---- drop-in subroutines with the reference's interfaces ----

! stotf.F90:51-60: one statistics sample of the box's state (the u, v, w, p
! the next rhsp starts from), its x interpolation and x derivatives formed
! on the GPU (rhsp.F90:72-90); the reference's intermediate and buffer
! arguments are not read.  Statistics are switched on with
! otbl_stats_enable (a sample before that stops with the library's
! message); otbl_stats_request sets dostat for the next substep 1 instead.
subroutine statsp(u_x,v,w,p,dudx,dvdx,dwdx,buf1,buf2,buf3,buf4,buf5,buf6,buf7,buf8, &
     buf_cor,buf_corp,buf_big,bphy1,bphy2,bphy3,mpiid,communicator)
  use otbl_gpu
  implicit none
  real(8) :: u_x(*),v(*),w(*),p(*),dudx(*),dvdx(*),dwdx(*),buf1(*),buf2(*),buf3(*),buf4(*)
  real(8) :: buf5(*),buf6(*),buf7(*),buf8(*),buf_cor(*),buf_corp(*),buf_big(*),bphy1(*),bphy2(*),bphy3(*)
  integer :: mpiid, communicator
  call otbl_need_box('statsp')
  call otbl_check(otbl_statsp(otbl_box), 'statsp')
end subroutine statsp

! rhsp.F90:22-27.  dt comes back from the GPU (computed at substep 1).
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
  call otbl_need_box('rhsp')
  if (otbl_mirror) then
     call otbl_put(OTBL_U, ut)
     call otbl_put(OTBL_V, vt)
     call otbl_put(OTBL_W, wt)
     call otbl_put(OTBL_P, pt)
     call otbl_put(OTBL_RHSUPA, rhsupat)
     call otbl_put(OTBL_RHSVPA, rhsvpat)
     call otbl_put(OTBL_RHSWPA, rhswpat)
  end if
  call otbl_check(otbl_rhsp(otbl_box, int(m, c_int)), 'rhsp')
  call otbl_check(otbl_get_time(otbl_box, t, dt), 'rhsp dt')
  if (otbl_mirror) then
     call otbl_fetch(OTBL_U, ut)
     call otbl_fetch(OTBL_V, vt)
     call otbl_fetch(OTBL_W, wt)
     call otbl_fetch(OTBL_RHSUPA, rhsupat)
     call otbl_fetch(OTBL_RHSVPA, rhsvpat)
     call otbl_fetch(OTBL_RHSWPA, rhswpat)
  end if
end subroutine rhsp

! boun.f90:63
subroutine outflow_correction(ut,vt,rhsupat,communicator)
  use otbl_gpu
  implicit none
  real(8) :: ut(*), vt(*), rhsupat(*)
  integer :: communicator
  call otbl_need_box('outflow_correction')
  if (otbl_mirror) then
     call otbl_put(OTBL_U, ut)
     call otbl_put(OTBL_V, vt)
     call otbl_put(OTBL_RHSUPA, rhsupat)
  end if
  call otbl_check(otbl_outflow_correction(otbl_box), 'outflow_correction')
  if (otbl_mirror) then
     call otbl_fetch(OTBL_U, ut)
     call otbl_fetch(OTBL_RHSUPA, rhsupat)
  end if
end subroutine outflow_correction

! pois.f90:11 (res, rest, rt are the reference's work arrays)
subroutine pois(ut,vt,wt,pt,res,rest,rt,varstep,mpiid,communicator)
  use otbl_gpu
  implicit none
  real(8) :: ut(*),vt(*),wt(*),pt(*),res(*),rest(*),rt(*)
  real(8) :: varstep
  integer :: mpiid, communicator
  call otbl_need_box('pois')
  if (otbl_mirror) then
     call otbl_put(OTBL_U, ut)
     call otbl_put(OTBL_V, vt)
     call otbl_put(OTBL_W, wt)
     call otbl_put(OTBL_P, pt)
  end if
  call otbl_check(otbl_pois(otbl_box, varstep), 'pois')
  if (otbl_mirror) then
     call otbl_fetch(OTBL_U, ut)
     call otbl_fetch(OTBL_V, vt)
     call otbl_fetch(OTBL_W, wt)
     call otbl_fetch(OTBL_P, pt)
  end if
end subroutine pois

! change.f90:224-255: planes plan(0:nz1-1, jee, nx) to pencils pen(nx, nz1 jee),
! line p = kr + nz1 (j - 1) + 1 (one process: every line in one block).  The
! GPU kernels read the planes layout directly and never call this; it is
! for reference code that works on pencils.  FP64 (the reference's
! transport rounds through FP32, change.f90:565-686).
subroutine chp2x(pen,plan,buf,rank,jee,communicator)
  use otbl_gpu
  implicit none
  integer :: rank, jee, communicator
  real(8) :: plan(otbl_nz1 * jee, otbl_nx), pen(otbl_nx, otbl_nz1 * jee), buf(*)
  real(8), allocatable :: t(:,:)
  call otbl_need_box('chp2x')
  allocate(t(otbl_nx, otbl_nz1 * jee))
  t = transpose(plan)
  pen = t
end subroutine chp2x

! change.f90:257-275: pencils to planes.  The first argument receives the
! planes and the second holds the pencils: the reference's chx2p passes its
! (pen, plan) on to chx2pu, whose dummies are (plan, pen) (change.f90:411);
! every reference call site is in place, where the order does not matter.
subroutine chx2p(pen,plan,buf,rank,jee,communicator)
  use otbl_gpu
  implicit none
  integer :: rank, jee, communicator
  real(8) :: pen(otbl_nz1 * jee, otbl_nx), plan(otbl_nx, otbl_nz1 * jee), buf(*)
  real(8), allocatable :: t(:,:)
  call otbl_need_box('chx2p')
  allocate(t(otbl_nz1 * jee, otbl_nx))
  t = transpose(plan)
  pen = t
end subroutine chx2p

! lip.f90:248-300 on pencils gen(nx, nlin), all lines of a field
subroutine interpxx(gen,geni,icon,econ,inbx,chi,nlin,loguv)
  use otbl_gpu
  implicit none
  integer :: chi, nlin, loguv
  real(8) :: gen(otbl_nx, nlin), geni(otbl_nx, nlin), icon(3,*), econ(2,*), inbx(4,4)
  real(8), allocatable :: pl(:,:)
  type(c_ptr) :: din, dout
  integer :: R
  integer(c_size_t) :: n
  call otbl_need_box('interpxx')
  R = nlin / otbl_nz1
  if (R * otbl_nz1 /= nlin .or. (R /= otbl_ny .and. R /= otbl_ny + 1)) &
       call otbl_fail('interpxx: nlin must be nz1*ny or nz1*(ny+1), every line of a field')
  if (.not. ((loguv == 1 .and. chi == 2) .or. (loguv == 0 .and. chi == 1))) &
       call otbl_fail('interpxx: (chi, loguv) must be (2, 1), the u set, or (1, 0), the v/w set')
  n = int(otbl_nx, c_size_t) * nlin
  allocate(pl(nlin, otbl_nx))
  pl = transpose(gen)
  din = otbl_up(1, pl, n)
  dout = otbl_scratch(2, n)
  call otbl_check(otbl_interpxx(otbl_box, din, dout, int(R, c_int), int(loguv, c_int)), 'interpxx')
  call otbl_down(2, pl, n)
  geni = transpose(pl)
end subroutine interpxx

! lip.f90:52-142 on pencils; result may be upen_int (rhsp.F90:182-192)
subroutine difvisxx(upen_int,upencil,result,icon,iconv,dcbx,econ,econv,cofvbx,chv,nlin,vel,rex)
  use otbl_gpu
  implicit none
  integer :: chv, nlin, vel
  real(8) :: upen_int(otbl_nx, nlin), upencil(otbl_nx, nlin), result(otbl_nx, nlin)
  real(8) :: icon(3,*), iconv(3,*), dcbx(5,4), econ(2,*), econv(3,*), cofvbx(4,4), rex
  real(8), allocatable :: pl(:,:)
  type(c_ptr) :: dp, dv, dout
  integer :: R
  integer(c_size_t) :: n
  call otbl_need_box('difvisxx')
  R = nlin / otbl_nz1
  if (R * otbl_nz1 /= nlin .or. (R /= otbl_ny .and. R /= otbl_ny + 1)) &
       call otbl_fail('difvisxx: nlin must be nz1*ny or nz1*(ny+1), every line of a field')
  if (.not. ((vel == 1 .and. chv == 1) .or. (vel == 0 .and. chv == 2))) &
       call otbl_fail('difvisxx: (chv, vel) must be (1, 1), the u set, or (2, 0), the v/w set')
  if (abs(rex + 1d0 / otbl_re) > 1d-14 * abs(rex)) call otbl_fail('difvisxx: rex is not -1/Re of the box')
  n = int(otbl_nx, c_size_t) * nlin
  allocate(pl(nlin, otbl_nx))
  pl = transpose(upen_int)
  dp = otbl_up(1, pl, n)
  pl = transpose(upencil)
  dv = otbl_up(2, pl, n)
  dout = otbl_scratch(3, n)
  call otbl_check(otbl_difvisxx(otbl_box, dp, dv, dout, int(R, c_int), int(vel, c_int)), 'difvisxx')
  call otbl_down(3, pl, n)
  result = transpose(pl)
end subroutine difvisxx

! lip.f90:359-440: m columns of the rows gen(dim, 1:jee) to geni(dim, 1:ny)
! (loguv = 0: u/w rows to v rows; 1: v rows to u rows), in place allowed;
! the columns go to the GPU in runs of nz1
subroutine interpyy(gen,geni,icon,econ,inby,jee,loguv,dim,m)
  use otbl_gpu
  implicit none
  integer :: jee, loguv, dim, m
  real(8) :: gen(dim, jee), geni(dim, otbl_ny + 1), icon(3,*), econ(2,*), inby(4,4)
  real(8), allocatable :: a(:,:,:), o(:,:,:)
  type(c_ptr) :: din, dout
  integer :: nb, q, c0, nc, Rin
  integer(c_size_t) :: nin, nout
  call otbl_need_box('interpyy')
  Rin = otbl_ny + 1
  if (loguv == 1) Rin = otbl_ny
  if (jee /= Rin) call otbl_fail('interpyy: jee must be ny+1 (loguv = 0) or ny (loguv = 1)')
  nb = (m + otbl_nz1 - 1) / otbl_nz1
  if (nb > otbl_nx) call otbl_fail('interpyy: too many columns')
  allocate(a(otbl_nz1, Rin, nb), o(otbl_nz1, otbl_ny + 1, nb))
  a = 0d0
  do q = 1, nb
     c0 = (q - 1) * otbl_nz1
     nc = min(otbl_nz1, m - c0)
     a(1:nc, :, q) = gen(c0 + 1:c0 + nc, 1:jee)
  end do
  nin = int(otbl_nz1, c_size_t) * Rin * nb
  nout = int(otbl_nz1, c_size_t) * (otbl_ny + 1) * nb
  din = otbl_up(1, a, nin)
  dout = otbl_scratch(2, nout)
  call otbl_check(otbl_interpyy(otbl_box, din, dout, int(loguv, c_int), int(nb, c_int)), 'interpyy')
  call otbl_down(2, o, nout)
  do q = 1, nb
     c0 = (q - 1) * otbl_nz1
     nc = min(otbl_nz1, m - c0)
     geni(c0 + 1:c0 + nc, 1:otbl_ny) = o(1:nc, 1:otbl_ny, q)
  end do
end subroutine interpyy

! lip.f90:152-232 on one plane: (chv, jee) = (3, ny+1): d/dy from v rows to
! u rows (dcyu), (1, ny): from u rows to v rows (dcyv), the two forms rhsp
! calls (rhsp.F90:164,385,406); in place allowed
subroutine differzy(geni,result,icon,dcby,econ,chv,npter1,fwb1,fwb2,jee)
  use otbl_gpu
  implicit none
  integer :: chv, npter1, fwb1, fwb2, jee
  real(8) :: geni(otbl_nz1, jee), result(otbl_nz1, otbl_ny + 1), icon(3,*), dcby(4,4), econ(2,*)
  real(8), allocatable :: a(:,:)
  type(c_ptr) :: din, dout
  integer :: which
  integer(c_size_t) :: n
  call otbl_need_box('differzy')
  if (chv == 3 .and. fwb1 == -1 .and. fwb2 == 0 .and. jee == otbl_ny + 1) then
     which = 0
  else if (chv == 1 .and. fwb1 == 0 .and. fwb2 == 1 .and. jee == otbl_ny) then
     which = 1
  else
     call otbl_fail('differzy: (chv, fwb1, fwb2, jee) must be (3, -1, 0, ny+1) or (1, 0, 1, ny)')
  end if
  n = int(otbl_nz1, c_size_t) * (otbl_ny + 1)
  allocate(a(otbl_nz1, otbl_ny + 1))
  a = 0d0
  a(:, 1:jee) = geni(:, 1:jee)
  din = otbl_up(1, a, n)
  dout = otbl_scratch(2, n)
  call otbl_check(otbl_differzy(otbl_box, din, dout, int(which, c_int), 1_c_int), 'differzy')
  call otbl_down(2, a, n)
  result(:, 1:jee) = a(:, 1:jee)
end subroutine differzy

! lip.f90:447-513 on one plane u(nz1, jee)
subroutine vistyy(u,result,icon,econ,cofvby,jee)
  use otbl_gpu
  implicit none
  integer :: jee
  real(8) :: u(otbl_nz1, jee), result(otbl_nz1, otbl_ny + 1), icon(3,*), econ(3,*), cofvby(4,4)
  real(8), allocatable :: a(:,:)
  type(c_ptr) :: din, dout
  integer(c_size_t) :: n
  call otbl_need_box('vistyy')
  if (jee /= otbl_ny .and. jee /= otbl_ny + 1) call otbl_fail('vistyy: jee must be ny or ny+1')
  n = int(otbl_nz1, c_size_t) * jee
  din = otbl_up(1, u, n)
  dout = otbl_scratch(2, n)
  call otbl_check(otbl_vistyy(otbl_box, din, dout, int(jee, c_int), 1_c_int), 'vistyy')
  allocate(a(otbl_nz1, jee))
  call otbl_down(2, a, n)
  result(:, 1:jee) = a
end subroutine vistyy

! lip.f90:528-597 on one plane, u in place; var1 = rkk (rhsp.F90:479)
subroutine implzy(u,res,icon,econ,jee,var1)
  use otbl_gpu
  implicit none
  integer :: jee
  real(8) :: u(otbl_nz1, jee), res(otbl_nz1, jee), icon(3,*), econ(3,*), var1
  type(c_ptr) :: du, dr
  integer(c_size_t) :: n
  call otbl_need_box('implzy')
  if (jee /= otbl_ny .and. jee /= otbl_ny + 1) call otbl_fail('implzy: jee must be ny or ny+1')
  n = int(otbl_nz1, c_size_t) * jee
  du = otbl_up(1, u, n)
  dr = otbl_up(2, res, n)
  call otbl_check(otbl_implzy(otbl_box, du, dr, int(jee, c_int), var1, 1_c_int), 'implzy')
  call otbl_down(1, u, n)
end subroutine implzy

! ffts.f90:27-67: rows jbf..jef of fou (complex 0:nz2) and phys (complex
! 0:nz/2, its first nz reals the physical row); iopt >= 0 c2r, < 0 r2c
subroutine fourxz(fou,phys,iopt,jee,jbf,jef)
  use otbl_gpu
  implicit none
  integer :: iopt, jee, jbf, jef
  real(8) :: fou(otbl_nz1, jbf:jef), phys(otbl_nz + 2, jbf:jef)
  real(8), allocatable :: a(:,:)
  type(c_ptr) :: df, dp
  integer :: nr
  integer(c_size_t) :: nf, np
  call otbl_need_box('fourxz')
  nr = jef - jbf + 1
  if (nr < 1) return
  nf = int(otbl_nz1, c_size_t) * nr
  np = int(otbl_nz, c_size_t) * nr
  allocate(a(otbl_nz, nr))
  if (iopt >= 0) then
     df = otbl_up(1, fou, nf)
     dp = otbl_scratch(2, np)
     call otbl_check(otbl_fourxz(otbl_box, df, dp, int(iopt, c_int), int(nr, c_int)), 'fourxz')
     call otbl_down(2, a, np)
     phys(1:otbl_nz, :) = a
  else
     a = phys(1:otbl_nz, :)
     dp = otbl_up(2, a, np)
     df = otbl_scratch(1, nf)
     call otbl_check(otbl_fourxz(otbl_box, df, dp, int(iopt, c_int), int(nr, c_int)), 'fourxz')
     call otbl_down(1, fou, nf)
  end if
end subroutine fourxz

! ffts.f90:79 and :124: the transforms are planned when the box is created
! (otbl_open); these check the length the reference plans for
subroutine rfti(n)
  use otbl_gpu
  implicit none
  integer :: n
  call otbl_need_box('rfti')
  if (n /= otbl_nz) call otbl_fail('rfti: the z transform length is nz')
end subroutine rfti

subroutine cosfti(n)
  use otbl_gpu
  implicit none
  integer :: n
  call otbl_need_box('cosfti')
  if (n /= otbl_nx - 1) call otbl_fail('cosfti: the x transform length is nx-1')
end subroutine cosfti

! ffts.f90:147-181: lines c(1:nx-1, j), j = mb..me, of stride isa, in place
subroutine cosftx(c,isa,mb,me,iopt)
  use otbl_gpu
  implicit none
  integer :: isa, mb, me, iopt
  real(8) :: c(isa, mb:me)
  real(8), allocatable :: a(:,:)
  type(c_ptr) :: d
  integer :: nl, n
  integer(c_size_t) :: cnt
  call otbl_need_box('cosftx')
  n = otbl_nx - 1
  nl = me - mb + 1
  if (nl < 1) return
  if (isa < n) call otbl_fail('cosftx: line stride shorter than nx-1')
  allocate(a(n, nl))
  a = c(1:n, :)
  cnt = int(n, c_size_t) * nl
  d = otbl_up(1, a, cnt)
  call otbl_check(otbl_cosft_lines(otbl_box, d, int(nl, c_int), int(iopt, c_int)), 'cosftx')
  call otbl_down(1, a, cnt)
  c(1:n, :) = a
end subroutine cosftx

! pois.f90:211-239: column ii of the complex plane u(0:nz2, ny1) into us,
! a(3, ny1) LU-factored in place as the reference does
subroutine soltriy(u,us,ii,a)
  use otbl_gpu
  implicit none
  integer :: ii
  real(8) :: u(otbl_nz1, otbl_ny - 1), us(otbl_nz1, otbl_ny - 1), a(3, otbl_ny - 1)
  type(c_ptr) :: du, dus
  integer(c_size_t) :: n
  call otbl_need_box('soltriy')
  n = int(otbl_nz1, c_size_t) * (otbl_ny - 1)
  du = otbl_up(1, u, n)
  dus = otbl_up(2, us, n)
  call otbl_check(otbl_soltriy(otbl_box, du, dus, int(ii, c_int), a, int(otbl_ny - 1, c_int)), 'soltriy')
  call otbl_down(2, us, n)
end subroutine soltriy

! coeft.f90:1338-1353: LU of a(3, npter1:neq) in place, on the GPU
subroutine coefinv(a,neq,npter1)
  use otbl_gpu
  implicit none
  integer :: neq, npter1
  real(8) :: a(3, npter1:neq)
  call otbl_need_box('coefinv')
  call otbl_check(otbl_coefinv(otbl_box, a, int(npter1, c_int), int(neq, c_int)), 'coefinv')
end subroutine coefinv

! coeft.f90:10: every coefficient of the step is built (and factored on the
! GPU) when the box is created from its y grid (otbl_open); the reference
! calls coef after reading the grid (bl_1.F90:56).  A caller that needs a
! coefficient array on the host reads it with otbl_coef(name, array).
subroutine coef(mpiid)
  use otbl_gpu
  implicit none
  integer :: mpiid
  call otbl_need_box('coef')
end subroutine coef

! coeft.f90:1244-1332: the y Laplacian of the pressure solve, ayp(3, 1:ny-1),
! from the box's grid (the one given to otbl_open, which must be yy)
subroutine mtr(ayp,cyclx,cycly,yy)
  use otbl_gpu
  implicit none
  integer :: cyclx, cycly
  real(8) :: ayp(3, otbl_ny), yy(0:otbl_ny + 1)
  call otbl_need_box('mtr')
  if (any(yy /= otbl_y)) call otbl_fail('mtr: yy is not the grid of the box')
  call otbl_coef('ayp', ayp)
end subroutine mtr

! coeft.f90:1926-1967: columns npter2 .. npter2+nlen-1 of u(nstr, *), rows
! npter1 .. npter1+neq-1, solved with the tridiagonal a(3, *) into us (us may
! be u); a is LU-factored in place first when iflag is 0, which then
! becomes 1 (on the GPU)
subroutine soltrix(u,nstr,nlen,neq,us,npter1,npter2,a,iflag)
  use otbl_gpu
  implicit none
  integer :: nstr, nlen, neq, npter1, npter2, iflag
  real(8) :: u(nstr, *), us(nstr, *), a(3, *)
  integer(c_int) :: fl
  call otbl_need_box('soltrix')
  fl = int(iflag, c_int)
  call otbl_check(otbl_soltrix(otbl_box, u, int(nstr, c_int), int(nlen, c_int), int(neq, c_int), us, &
       int(npter1, c_int), int(npter2, c_int), a, fl), 'soltrix')
  iflag = fl
end subroutine soltrix

! coeft.f90:1360-1789: the closure weights bcc(1:n) for the spacings a, b,
! c, d, for the flags the reference calls coefb with (0, 1, 3 to 7); np is
! the reference's declared size of its work matrix and is not needed
subroutine coefb(n,np,a,b,c,d,bcc,flagv)
  use otbl_gpu
  implicit none
  integer :: n, np, flagv
  real(8) :: a, b, c, d, bcc(5)
  call otbl_check(otbl_coefb(int(n, c_int), int(flagv, c_int), a, b, c, d, bcc), 'coefb')
end subroutine coefb

! ffts.f90:188-213: each OpenMP thread's range of y rows and pencils for the
! FFT loops.  The GPU grids cover every row and line and the host side is
! one thread, so there is nothing to set.
subroutine pointersfft(mpiid)
  implicit none
  integer :: mpiid
end subroutine pointersfft

! change.f90:150-222: the transpose bookkeeping of the x-slab decomposition
! (comm_setup for the u and v line sets).  One process holds the whole box
! and the kernels read the planes layout directly (chp2x / chx2p above), so
! there is nothing to set.
subroutine pointers_p2p(rank)
  implicit none
  integer :: rank
end subroutine pointers_p2p

! change.f90:29-148: the per-rank block sizes of the MPI transposes; node
! and domain are the reference's shared_mem derived types, received here as
! untyped storage and not touched (one process, no transposes between
! ranks).
subroutine comm_setup(nx,ny,nz,rank,size,node,domain,pblock)
  implicit none
  integer :: nx, ny, nz, rank, size, pblock
  integer :: node(*), domain(*)
end subroutine comm_setup
