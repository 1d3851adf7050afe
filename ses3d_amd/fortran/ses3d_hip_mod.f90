!==============================================================================
! ses3d_hip_mod -- iso_c_binding interface to libses3d_hip.so (include/ses3d_hip.h)
!
! A Fortran driver structured like SOURCE/ses3d_main.f90 calls these instead of
! ses3d_evolution / record_seismograms / ses3d_store / ses3d_grad; all sizes are
! run-time values, arrays are passed in the reference's layout.
!==============================================================================
module ses3d_hip_mod
  use iso_c_binding
  implicit none

  type, bind(c) :: ses3d_hip_config
     integer(c_int32_t) :: nx, ny, nz, lpd, pml, is_diss, adjoint_flag, samp_ad, nt
     real(c_float) :: dt, xmin, xmax, ymin, ymax, zmin, zmax, zmax_global
     real(c_float) :: tau_p(3), D_p(3)
  end type ses3d_hip_config

  interface
     integer(c_int) function ses3d_hip_create(cfg, n_events, ctx) bind(c, name='ses3d_hip_create')
       import :: c_int, c_int32_t, c_ptr, ses3d_hip_config
       type(ses3d_hip_config), intent(in) :: cfg
       integer(c_int32_t), value :: n_events
       type(c_ptr), intent(out) :: ctx
     end function

     subroutine ses3d_hip_destroy(ctx) bind(c, name='ses3d_hip_destroy')
       import :: c_ptr
       type(c_ptr), value :: ctx
     end subroutine

     type(c_ptr) function ses3d_hip_last_error() bind(c, name='ses3d_hip_last_error')
       import :: c_ptr
     end function

     integer(c_int) function ses3d_hip_set_model(ctx, rhoinv, mu, lambda, A, B, C, Q) &
          bind(c, name='ses3d_hip_set_model')
       import :: c_int, c_ptr, c_float
       type(c_ptr), value :: ctx
       real(c_float), intent(in) :: rhoinv(*), mu(*), lambda(*), A(*), B(*), C(*)
       type(c_ptr), value :: Q
     end function

     integer(c_int) function ses3d_hip_set_source(ctx, ev, source_type, xxs, yys, zzs, mom, so) &
          bind(c, name='ses3d_hip_set_source')
       import :: c_int, c_ptr, c_float, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: ev, source_type
       real(c_float), value :: xxs, yys, zzs
       real(c_float), intent(in) :: mom(6), so(*)
     end function

     integer(c_int) function ses3d_hip_set_receivers(ctx, ev, nr, rec) bind(c, name='ses3d_hip_set_receivers')
       import :: c_int, c_ptr, c_float, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: ev, nr
       real(c_float), intent(in) :: rec(*)
     end function

     integer(c_int) function ses3d_hip_set_adjoint_sources(ctx, ev, n, loc, stf) &
          bind(c, name='ses3d_hip_set_adjoint_sources')
       import :: c_int, c_ptr, c_float, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: ev, n
       real(c_float), intent(in) :: loc(*), stf(*)
     end function

     integer(c_int) function ses3d_hip_init(ctx) bind(c, name='ses3d_hip_init')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function

     integer(c_int) function ses3d_hip_run(ctx, it_begin, n_steps) bind(c, name='ses3d_hip_run')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: it_begin, n_steps
     end function

     integer(c_int) function ses3d_hip_synchronize(ctx) bind(c, name='ses3d_hip_synchronize')
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function

     integer(c_int) function ses3d_hip_get_nr(ctx, ev, nr) bind(c, name='ses3d_hip_get_nr')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: ev
       integer(c_int32_t), intent(out) :: nr
     end function

     integer(c_int) function ses3d_hip_get_receiver_index(ctx, ev, idx) bind(c, name='ses3d_hip_get_receiver_index')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: ev
       integer(c_int32_t), intent(out) :: idx(*)
     end function

     integer(c_int) function ses3d_hip_get_seismograms(ctx, ev, out) bind(c, name='ses3d_hip_get_seismograms')
       import :: c_int, c_ptr, c_float, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: ev
       real(c_float), intent(out) :: out(*)
     end function

     integer(c_int) function ses3d_hip_get_field(ctx, ev, name, out) bind(c, name='ses3d_hip_get_field')
       import :: c_int, c_ptr, c_float, c_int32_t, c_char
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: ev
       character(kind=c_char), intent(in) :: name(*)
       real(c_float), intent(out) :: out(*)
     end function

     integer(c_int) function ses3d_hip_get_kernels(ctx, ev, g_rho, g_cp, g_csh, g_csv, g_q) &
          bind(c, name='ses3d_hip_get_kernels')
       import :: c_int, c_ptr, c_double, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: ev
       real(c_double), intent(out) :: g_rho(*), g_cp(*), g_csh(*), g_csv(*)
       type(c_ptr), value :: g_q
     end function

     integer(c_int) function ses3d_hip_get_history(ctx, ev, step, out9) bind(c, name='ses3d_hip_get_history')
       import :: c_int, c_ptr, c_float, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: ev, step
       real(c_float), intent(out) :: out9(*)
     end function

     integer(c_int) function ses3d_hip_put_history(ctx, ev, step, in9) bind(c, name='ses3d_hip_put_history')
       import :: c_int, c_ptr, c_float, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: ev, step
       real(c_float), intent(in) :: in9(*)
     end function

     integer(c_int) function ses3d_hip_get_saving_vector(ctx, sv) bind(c, name='ses3d_hip_get_saving_vector')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), intent(out) :: sv(*)
     end function

     ! degree of the stored forward fields (ses3d_store's deg; 4 as the reference)
     integer(c_int) function ses3d_hip_set_history_degree(ctx, deg) bind(c, name='ses3d_hip_set_history_degree')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: deg
     end function

     ! type of the kernel accumulators: 32 (float32, the reference's grad_* arrays) or 64
     integer(c_int) function ses3d_hip_set_kernel_precision(ctx, bits) bind(c, name='ses3d_hip_set_kernel_precision')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: bits
     end function

     ! gradient without disk history: restart states in HBM every k steps, the
     ! sampled steps of the last `segments` segments kept, the rest replayed
     integer(c_int) function ses3d_hip_set_checkpoint_keep(fwd, segments) bind(c, name='ses3d_hip_set_checkpoint_keep')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: fwd
       integer(c_int32_t), value :: segments
     end function

     ! optional: allocate the checkpoint storage before the timed forward pass
     integer(c_int) function ses3d_hip_prepare_checkpoints(fwd, k) bind(c, name='ses3d_hip_prepare_checkpoints')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: fwd
       integer(c_int32_t), value :: k
     end function

     integer(c_int) function ses3d_hip_forward_checkpointed(fwd, k) bind(c, name='ses3d_hip_forward_checkpointed')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: fwd
       integer(c_int32_t), value :: k
     end function

     integer(c_int) function ses3d_hip_adjoint_checkpointed(adj, fwd) bind(c, name='ses3d_hip_adjoint_checkpointed')
       import :: c_int, c_ptr
       type(c_ptr), value :: adj, fwd
     end function

     integer(c_int) function ses3d_hip_history_size(ctx, n) bind(c, name='ses3d_hip_history_size')
       import :: c_int, c_ptr, c_int64_t
       type(c_ptr), value :: ctx
       integer(c_int64_t), intent(out) :: n
     end function
  end interface

contains

  ! Stop with the library's message when a call fails (the reference aborts
  ! through the Fortran runtime on its own errors).
  subroutine hip_check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    type(c_ptr) :: p
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (rc == 0) return
    p = ses3d_hip_last_error()
    call c_f_pointer(p, msg, [4096])
    n = 0
    do while (n < 4096)
       if (msg(n + 1) == c_null_char) exit
       n = n + 1
    end do
    write(*, *) 'ses3d_hip error in ', what, ': ', msg(1:n)
    stop 1
  end subroutine hip_check

end module ses3d_hip_mod
