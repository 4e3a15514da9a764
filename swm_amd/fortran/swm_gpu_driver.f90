!> A driver in the shape of the reference's src/model.f90 (initModel,
!! timestepModel, finishModel; model.f90:45-60, 153-197) running the time loop
!! on the device through swm_gpu_shim.  The host initialisation the reference
!! performs (initDomain, damping, forcing, initial conditions, tracers) is read
!! from a raw stream file instead of the namelists and netCDF inputs; output
!! steps copy the registered fields back, as a due diag task would.
!!
!! Usage: swm_gpu_driver <input.bin> <output.bin>
!! Input (native byte order, no record markers): int32 Nx, Ny, variant, flags,
!! precision, Nt, out_every, ntracers, has_taux; float64 G, Ah, dt, A, dLambda,
!! dTheta, RHO0, minD, lbc, ab_chi, tau_scale; per grid (H, u, v, eta): int8
!! ocean(Nx,Ny), float64 H(Nx,Ny), cos_lat(Ny), tan_lat(Ny), f(Ny); float64
!! impl_u, impl_v, impl_eta, F_x_const, F_y_const, F_eta_const (Nx,Ny each);
!! tau_x(Nx,Ny) if has_taux; u, v, eta (Nx,Ny); per tracer float64 kappa_h,
!! CH, C, C0, gamma_C, impl (Nx,Ny each).
!! Output: per output step int64 itt, float64 u, v, eta (Nx,Ny), then CH of
!! every tracer.
program swm_gpu_driver
  use, intrinsic :: iso_c_binding
  use swm_gpu_shim
  implicit none
  character(len=1024) :: fin, fout
  integer :: uin, uout, g, t, itt
  integer(c_int32_t) :: hdr(9), tid
  real(c_double) :: sc(11)
  integer :: Nx, Ny, Nt, nout, ntrc
  type(swm_desc), target :: d
  type(swm_tracer_desc), target :: td
  type(c_ptr) :: gpu = c_null_ptr
  integer(c_int8_t), allocatable, target :: ocean(:, :, :)
  real(c_double), allocatable, target :: H(:, :, :), cosl(:, :), tanl(:, :), f(:, :)
  real(c_double), allocatable, target :: impl(:, :, :), Fc(:, :, :), taux(:, :)
  real(c_double), allocatable, target :: u(:, :), v(:, :), eta(:, :), field(:, :)
  real(c_double), allocatable, target :: tkappa(:), tCH(:, :, :), tC(:, :, :), tC0(:, :, :), tgC(:, :, :), &
                                         timpl(:, :, :)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open (newunit=uin, file=trim(fin), access='stream', form='unformatted', status='old', action='read')
  read (uin) hdr
  Nx = hdr(1); Ny = hdr(2); Nt = hdr(6); nout = hdr(7); ntrc = hdr(8)
  read (uin) sc
  allocate (ocean(Nx, Ny, 4), H(Nx, Ny, 4), cosl(Ny, 4), tanl(Ny, 4), f(Ny, 4))
  do g = 1, 4
    read (uin) ocean(:, :, g), H(:, :, g), cosl(:, g), tanl(:, g), f(:, g)
  end do
  allocate (impl(Nx, Ny, 3), Fc(Nx, Ny, 3), taux(Nx, Ny), u(Nx, Ny), v(Nx, Ny), eta(Nx, Ny), field(Nx, Ny))
  read (uin) impl, Fc
  if (hdr(9) /= 0) read (uin) taux
  read (uin) u, v, eta
  allocate (tkappa(max(ntrc, 1)), tCH(Nx, Ny, max(ntrc, 1)), tC(Nx, Ny, max(ntrc, 1)), tC0(Nx, Ny, max(ntrc, 1)), &
            tgC(Nx, Ny, max(ntrc, 1)), timpl(Nx, Ny, max(ntrc, 1)))
  do t = 1, ntrc
    read (uin) tkappa(t), tCH(:, :, t), tC(:, :, t), tC0(:, :, t), tgC(:, :, t), timpl(:, :, t)
  end do
  close (uin)

  ! initModel (model.f90:91-131): the device copy of the initialised modules
  d%Nx = hdr(1); d%Ny = hdr(2); d%variant = hdr(3); d%flags = hdr(4); d%precision = hdr(5)
  d%G = sc(1); d%Ah = sc(2); d%dt = sc(3); d%A = sc(4); d%dLambda = sc(5); d%dTheta = sc(6)
  d%RHO0 = sc(7); d%minD = sc(8); d%lbc = sc(9); d%ab_chi = sc(10); d%tau_scale = sc(11)
  do g = 1, 4
    d%grid(g) = swm_grid_desc(c_loc(ocean(1, 1, g)), c_loc(H(1, 1, g)), c_loc(cosl(1, g)), c_loc(tanl(1, g)), &
                              c_loc(f(1, g)))
  end do
  d%impl_u = c_loc(impl(1, 1, 1)); d%impl_v = c_loc(impl(1, 1, 2)); d%impl_eta = c_loc(impl(1, 1, 3))
  d%F_x_const = c_loc(Fc(1, 1, 1)); d%F_y_const = c_loc(Fc(1, 1, 2)); d%F_eta_const = c_loc(Fc(1, 1, 3))
  if (hdr(9) /= 0) d%tau_x = c_loc(taux)
  call gpu_check(gpu, swm_gpu_create(d, gpu), "swm_gpu_create")
  call gpu_check(gpu, swm_gpu_upload_state(gpu, c_loc(u), c_loc(v), c_loc(eta), c_null_ptr, c_null_ptr, &
                 c_null_ptr, 0_c_int64_t), "swm_gpu_upload_state")
  do t = 1, ntrc        ! TRC_initTracer (model.f90:128)
    td = swm_tracer_desc(tkappa(t), c_loc(tCH(1, 1, t)), c_loc(tC(1, 1, t)), c_null_ptr, c_loc(tC0(1, 1, t)), &
                         c_loc(tgC(1, 1, t)), c_loc(timpl(1, 1, t)))
    call gpu_check(gpu, swm_gpu_add_tracer(gpu, td, tid), "swm_gpu_add_tracer")
  end do

  ! timestepModel (model.f90:153-197)
  open (newunit=uout, file=trim(fout), access='stream', form='unformatted', status='replace', action='write')
  do itt = 1, Nt
    ! SWM_timestep, TRC_timestep, model_advance
    call gpu_check(gpu, swm_gpu_step(gpu, 1_c_int64_t), "swm_gpu_step")
    ! Diag: the registered fields a due task reads
    if (mod(itt, nout) == 0) then
      write (uout) int(swm_gpu_itt(gpu), c_int64_t)
      call gpu_sync_field(gpu, "SWM_U", -1, field)
      write (uout) field
      call gpu_sync_field(gpu, "SWM_V", -1, field)
      write (uout) field
      call gpu_sync_field(gpu, "SWM_ETA", -1, field)
      write (uout) field
      do t = 1, ntrc
        call gpu_sync_field(gpu, "CH", t - 1, field)
        write (uout) field
      end do
    end if
  end do
  close (uout)

  ! finishModel (model.f90:206-225)
  call swm_gpu_destroy(gpu)
end program swm_gpu_driver
