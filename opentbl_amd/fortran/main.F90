! Driver of the GPU RK3 step written like the reference's bl_1
! (bl_1.F90:85-128): reads a box, runs nsteps steps of three substeps
! through the drop-in rhsp / outflow_correction / pois, writes the fields.
!
!   otbl_driver <input.bin> <output.bin> <nsteps> [mirror]
!
! Without "mirror" the state stays on the device between the calls (put
! once, fetched at the end; otbl_mirror = .false.).  With it, the drop-ins
! copy the fields in and out of the arrays passed to every call, as a
! reference caller that works on its own arrays between calls needs.
! input.bin (stream, little endian): int32 nx, ny, nz, xout, dt_mode;
! float64 re, ax, az, cfl; y(0:ny+1); u0(ny+1); v0(ny); um(ny+1);
! then u, v, w, p in the planes layout.  output.bin: tiempo, dt, u, v, w, p.
program otbl_driver
  use otbl_gpu
  implicit none
  integer :: nx, ny, nz, xout, dt_mode, nsteps, istep, isubstp, nz1, ical, mpiid, comm
  real(8) :: re, ax, az, cfl, dt, vardt, tiempo
  real(8), allocatable :: y(:), u0(:), v0(:), um(:), u(:), v(:), w(:), p(:), dum(:)
  real(8), allocatable :: rhsupa(:), rhsvpa(:), rhswpa(:)
  real(8), parameter :: rkdv(3) = [4d0/15d0, 1d0/15d0, 1d0/6d0]   ! coeft.f90:506-508
  character(len=512) :: fin, fout, arg, mode
  integer :: lu, lv

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  call get_command_argument(3, arg)
  read(arg, *) nsteps
  mode = ' '
  if (command_argument_count() >= 4) call get_command_argument(4, mode)
  otbl_mirror = trim(mode) == 'mirror'

  open(10, file=trim(fin), access='stream', form='unformatted', status='old')
  read(10) nx, ny, nz, xout, dt_mode
  read(10) re, ax, az, cfl
  nz1 = 2*(nz/3)
  lu = nz1*(ny+1)*nx
  lv = nz1*ny*nx
  allocate(y(0:ny+1), u0(ny+1), v0(ny), um(ny+1), u(lu), v(lv), w(lu), p(lv), dum(1))
  allocate(rhsupa(lu), rhsvpa(lv), rhswpa(lu))
  rhsupa = 0d0
  rhsvpa = 0d0
  rhswpa = 0d0
  read(10) y, u0, v0, um, u, v, w, p
  close(10)

  call otbl_open(nx, ny, nz, xout, re, ax, az, cfl, y, dt_mode)
  call otbl_put(OTBL_U, u)
  call otbl_put(OTBL_V, v)
  call otbl_put(OTBL_W, w)
  call otbl_put(OTBL_P, p)
  call otbl_check(otbl_set_inflow(otbl_box, u0, v0, um), 'set_inflow')
  call otbl_check(otbl_set_time(otbl_box, 0d0, 0d0, 0d0), 'set_time')

  dt = 0d0
  tiempo = 0d0
  ical = 0
  mpiid = 0
  comm = 0
  do istep = 1, nsteps
     do isubstp = 1, 3
        call rhsp(u,v,w,p,rhsupa,rhsvpa,rhswpa, dum,dum,dum,dum,dum,dum, dum,dum,dum, &
             dum,dum,dum,dum,dum,dum, dum,dum,dum,dum,dum,dum,dum, dt,isubstp,ical,istep,mpiid,comm)
        call outflow_correction(u,v,rhsupa,comm)
        vardt = 5d-1/dt/rkdv(isubstp)
        call pois(u,v,w,p,dum,dum,dum,vardt,mpiid,comm)
     end do
     ! end of step (bl_1.F90:128); genflu reads tiempo on the device
     tiempo = tiempo + dt
     call otbl_check(otbl_set_time(otbl_box, tiempo, 0d0, dt), 'set_time')
  end do
  if (.not. otbl_mirror) then
     call otbl_fetch(OTBL_U, u)
     call otbl_fetch(OTBL_V, v)
     call otbl_fetch(OTBL_W, w)
     call otbl_fetch(OTBL_P, p)
  end if
  call otbl_close()

  open(11, file=trim(fout), access='stream', form='unformatted', status='replace')
  write(11) tiempo, dt, u, v, w, p
  close(11)
  write(*,'(a,i0,a,es12.5)') 'otbl_driver: ', nsteps, ' steps done, dt = ', dt
end program otbl_driver
