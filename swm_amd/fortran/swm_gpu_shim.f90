!> ISO_C_BINDING interface of the MI355X SWM time-stepping library.
!!
!! A build of the reference model uses this module in place of the CPU bodies
!! of SWM_timestep / SWM_advance (src/swm_module.f90:71-139), TRC_timestep /
!! TRC_advance (src/tracer_module.f90:109-146) and main_ElSolv_SOR
!! (src/ElSolv_SOR.f90:150-222): model.f90 keeps its call order, fields stay on
!! the device, and a diag hook copies registered fields back when a task is due.
!! The C declarations are in include/swm_gpu.h.
module swm_gpu_shim
  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer(c_int32_t), parameter, public :: SWM_GPU_ABI_VERSION = 2
  integer(c_int32_t), parameter, public :: SWM_FULLY_NONLINEAR = 1, SWM_LINEARISED_MEAN_STATE = 2, &
                                           SWM_LINEARISED_STATE_OF_REST = 3
  integer(c_int32_t), parameter, public :: SWM_LATERAL_MIXING = 1, SWM_QUADRATIC_BOTTOM_FRICTION = 2, &
                                           SWM_BAROTROPIC = 4, SWM_TAU_SCALE = 8, SWM_ELLIPTIC_SOLVER = 16, &
                                           SWM_FP32_LOW_PART = 32
  integer(c_int32_t), parameter, public :: SWM_FP64 = 0, SWM_FP32 = 1
  integer(c_int32_t), parameter, public :: SWM_COMP_X = 0, SWM_COMP_Y = 1, SWM_COMP_ETA = 2
  integer(c_int32_t), parameter, public :: SWM_STREAM_WIND = 0, SWM_STREAM_CUSTOM = 1, SWM_STREAM_HEATING = 2

  !> swm_grid_desc
  type, bind(C), public :: swm_grid_desc
    type(c_ptr) :: ocean = c_null_ptr
    type(c_ptr) :: H = c_null_ptr
    type(c_ptr) :: cos_lat = c_null_ptr
    type(c_ptr) :: tan_lat = c_null_ptr
    type(c_ptr) :: f = c_null_ptr
  end type swm_grid_desc

  !> swm_desc
  type, bind(C), public :: swm_desc
    integer(c_int32_t) :: abi_version = SWM_GPU_ABI_VERSION
    integer(c_int32_t) :: Nx = 0, Ny = 0, variant = 0, flags = 0, precision = SWM_FP64
    real(c_double) :: G = 0, Ah = 0, dt = 0, A = 0, dLambda = 0, dTheta = 0, RHO0 = 0, minD = 1, &
                      lbc = 0, ab_chi = 0.1_c_double, tau_scale = 0
    type(swm_grid_desc) :: grid(4)
    type(c_ptr) :: impl_u = c_null_ptr, impl_v = c_null_ptr, impl_eta = c_null_ptr
    type(c_ptr) :: gamma_sq_u = c_null_ptr, gamma_sq_v = c_null_ptr
    type(c_ptr) :: F_x_const = c_null_ptr, F_y_const = c_null_ptr, F_eta_const = c_null_ptr
    type(c_ptr) :: tau_x = c_null_ptr, tau_y = c_null_ptr
    type(c_ptr) :: osc_re(3) = c_null_ptr, osc_im(3) = c_null_ptr
    real(c_double) :: osc_omega(3) = 0
    type(c_ptr) :: u_bs = c_null_ptr, v_bs = c_null_ptr, zeta_bs = c_null_ptr
  end type swm_desc

  !> swm_tracer_desc
  type, bind(C), public :: swm_tracer_desc
    real(c_double) :: kappa_h = 0
    type(c_ptr) :: CH = c_null_ptr, C = c_null_ptr, G_CH = c_null_ptr
    type(c_ptr) :: C0 = c_null_ptr, gamma_C = c_null_ptr, impl = c_null_ptr
  end type swm_tracer_desc

  interface
    integer(c_int) function swm_gpu_create(d, ctx) bind(C, name="swm_gpu_create")
      import :: c_int, c_ptr, swm_desc
      type(swm_desc), intent(in) :: d
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function swm_gpu_upload_state(ctx, u, v, eta, G_u, G_v, G_eta, itt) &
        bind(C, name="swm_gpu_upload_state")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx, u, v, eta, G_u, G_v, G_eta
      integer(c_int64_t), value :: itt
    end function
    integer(c_int) function swm_gpu_add_tracer(ctx, t, id) bind(C, name="swm_gpu_add_tracer")
      import :: c_int, c_ptr, c_int32_t, swm_tracer_desc
      type(c_ptr), value :: ctx
      type(swm_tracer_desc), intent(in) :: t
      integer(c_int32_t), intent(out) :: id
    end function
    ! SWM_forcing_initStream for a dataset of nrec records (memchunk); records(Nx,Ny,nrec)
    integer(c_int) function swm_gpu_add_forcing_stream(ctx, kind, component, nrec, rec_dt, chunk_size, &
                                                        records, id) bind(C, name="swm_gpu_add_forcing_stream")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: kind, component, nrec, chunk_size
      real(c_double), value :: rec_dt
      real(c_double), intent(in) :: records(*)
      integer(c_int32_t), intent(out) :: id
    end function
    ! DFF_initDynFromFile (src/dynFromFile_module.f90:56-95): one dataset of u, v or eta
    integer(c_int) function swm_gpu_add_dyn_stream(ctx, field, nrec, rec_dt, chunk_size, records) &
        bind(C, name="swm_gpu_add_dyn_stream")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: field, nrec, chunk_size
      real(c_double), value :: rec_dt
      real(c_double), intent(in) :: records(*)
    end function
    integer(c_int) function swm_gpu_step(ctx, nsteps) bind(C, name="swm_gpu_step")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), value :: nsteps
    end function
    integer(c_int64_t) function swm_gpu_itt(ctx) bind(C, name="swm_gpu_itt")
      import :: c_int64_t, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function swm_gpu_sync(ctx, name, tracer_id, out) bind(C, name="swm_gpu_sync")
      import :: c_int, c_ptr, c_char, c_int32_t
      type(c_ptr), value :: ctx
      character(kind=c_char), dimension(*), intent(in) :: name
      integer(c_int32_t), value :: tracer_id
      type(c_ptr), value :: out
    end function
    integer(c_int) function swm_gpu_sor(ctx, B, chi, epsilon, first_guess, iters, anorm) &
        bind(C, name="swm_gpu_sor")
      import :: c_int, c_ptr, c_double, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx, B, chi
      real(c_double), value :: epsilon
      integer(c_int32_t), value :: first_guess
      integer(c_int64_t), intent(out) :: iters
      real(c_double), intent(out) :: anorm
    end function
    integer(c_int) function swm_gpu_velocity_potential(ctx, omega, rel_tol, max_sweeps, fp32_sweeps, reset, &
        iters, rel_res) bind(C, name="swm_gpu_velocity_potential")
      import :: c_int, c_ptr, c_double, c_int32_t, c_int64_t
      type(c_ptr), value :: ctx
      real(c_double), value :: omega, rel_tol
      integer(c_int64_t), value :: max_sweeps
      integer(c_int32_t), value :: fp32_sweeps, reset
      integer(c_int64_t), intent(out) :: iters
      real(c_double), intent(out) :: rel_res
    end function
    integer(c_int) function swm_gpu_avg_add(ctx, name, tracer_id, square, avg_id) bind(C, name="swm_gpu_avg_add")
      import :: c_int, c_ptr, c_char, c_int32_t
      type(c_ptr), value :: ctx
      character(kind=c_char), dimension(*), intent(in) :: name
      integer(c_int32_t), value :: tracer_id, square
      integer(c_int32_t), intent(out) :: avg_id
    end function
    integer(c_int) function swm_gpu_avg_read(ctx, avg_id, out, weight, reset) bind(C, name="swm_gpu_avg_read")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx, out
      integer(c_int32_t), value :: avg_id, reset
      real(c_double), intent(out) :: weight
    end function
    integer(c_int) function swm_gpu_sums(ctx, tracer_id, out) bind(C, name="swm_gpu_sums")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: tracer_id
      real(c_double), dimension(3), intent(out) :: out
    end function
    integer(c_int) function swm_gpu_sums_noalloc(ctx, tracer_id, out) bind(C, name="swm_gpu_sums_noalloc")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: tracer_id
      real(c_double), dimension(3), intent(out) :: out
    end function
    integer(c_int) function swm_gpu_elliptic_stats(ctx, iters, anorm) bind(C, name="swm_gpu_elliptic_stats")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int64_t), intent(out) :: iters
      real(c_double), intent(out) :: anorm
    end function
    integer(c_int) function swm_gpu_tune(ctx, steps, best_shape, times_ms) bind(C, name="swm_gpu_tune")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx, times_ms
      integer(c_int32_t), value :: steps
      integer(c_int32_t), intent(out) :: best_shape
    end function
    type(c_ptr) function swm_gpu_last_error(ctx) bind(C, name="swm_gpu_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function
    subroutine swm_gpu_destroy(ctx) bind(C, name="swm_gpu_destroy")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine
  end interface

  public :: swm_gpu_create, swm_gpu_upload_state, swm_gpu_add_tracer, swm_gpu_add_forcing_stream, &
            swm_gpu_add_dyn_stream, &
            swm_gpu_step, swm_gpu_itt, &
            swm_gpu_sync, swm_gpu_sor, swm_gpu_velocity_potential, swm_gpu_avg_add, swm_gpu_avg_read, &
            swm_gpu_sums, swm_gpu_sums_noalloc, swm_gpu_elliptic_stats, swm_gpu_tune, swm_gpu_last_error, &
            swm_gpu_destroy
  public :: gpu_check, gpu_sync_field, gpu_main_ElSolv_SOR, gpu_avg_register, gpu_avg_write

contains

  !> Turn a non-zero status into the reference's fatal error path (log_fatal -> STOP 1).
  subroutine gpu_check(ctx, rc, what)
    type(c_ptr), intent(in) :: ctx
    integer(c_int), intent(in) :: rc
    character(*), intent(in) :: what
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (rc == 0) return
    call c_f_pointer(swm_gpu_last_error(ctx), msg, [512])
    n = 0
    do while (n < 512)
      if (msg(n + 1) == c_null_char) exit
      n = n + 1
    end do
    write (*, '(A,": status ",I0,": ",512A)') what, rc, msg(1:n)
    stop 1
  end subroutine gpu_check

  !> Register read of a due diag task: copy device field `name` into `field`.
  subroutine gpu_sync_field(ctx, name, tracer_id, field)
    type(c_ptr), intent(in) :: ctx
    character(*), intent(in) :: name
    integer, intent(in) :: tracer_id
    real(c_double), dimension(:, :), target, contiguous, intent(inout) :: field
    call gpu_check(ctx, swm_gpu_sync(ctx, trim(name)//c_null_char, int(tracer_id, c_int32_t), c_loc(field)), &
                   "swm_gpu_sync "//trim(name))
  end subroutine gpu_sync_field

  !> main_ElSolv_SOR(B, chi, epsilon, first_guess) with the reference's contract.
  subroutine gpu_main_ElSolv_SOR(ctx, B, chi, epsilon, first_guess)
    type(c_ptr), intent(in) :: ctx
    real(c_double), dimension(:, :), target, contiguous, intent(in) :: B
    real(c_double), dimension(:, :), target, contiguous, intent(inout) :: chi
    real(c_double), intent(in) :: epsilon
    logical, intent(in) :: first_guess
    integer(c_int64_t) :: iters
    real(c_double) :: anorm
    integer(c_int) :: rc
    rc = swm_gpu_sor(ctx, c_loc(B), c_loc(chi), epsilon, merge(1_c_int32_t, 0_c_int32_t, first_guess), iters, anorm)
    if (rc == 1) then
      write (*, '(A,X,I6,X,A,X,E11.4)') "Elliptic solver did not converge after", iters, "iterations. Residual:", anorm
    else
      call gpu_check(ctx, rc, "swm_gpu_sor")
    end if
  end subroutine gpu_main_ElSolv_SOR

  !> Averaging diag task kept on the device (diagTask.f90 type "A"): register once,
  !! then read (and reset) at each output step instead of adding every step on the host.
  subroutine gpu_avg_register(ctx, name, tracer_id, square, avg_id)
    type(c_ptr), intent(in) :: ctx
    character(*), intent(in) :: name
    integer, intent(in) :: tracer_id
    logical, intent(in) :: square
    integer(c_int32_t), intent(out) :: avg_id
    call gpu_check(ctx, swm_gpu_avg_add(ctx, trim(name)//c_null_char, int(tracer_id, c_int32_t), &
                   merge(1_c_int32_t, 0_c_int32_t, square), avg_id), "swm_gpu_avg_add "//trim(name))
  end subroutine gpu_avg_register

  !> task%buffer / task%bufferCount of an averaging task, buffer reset afterwards.
  subroutine gpu_avg_write(ctx, avg_id, field)
    type(c_ptr), intent(in) :: ctx
    integer(c_int32_t), intent(in) :: avg_id
    real(c_double), dimension(:, :), target, contiguous, intent(inout) :: field
    real(c_double) :: weight
    call gpu_check(ctx, swm_gpu_avg_read(ctx, avg_id, c_loc(field), weight, 1_c_int32_t), "swm_gpu_avg_read")
  end subroutine gpu_avg_write

end module swm_gpu_shim
