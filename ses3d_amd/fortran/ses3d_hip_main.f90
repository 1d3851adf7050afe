!==============================================================================
! ses3d_hip_main -- drop-in replacement of the SES3D executable (one GPU).
!
! Launched from MAIN/ like the reference (MAIN/job_run_ses3d); reads the same
! files with the same relative paths and writes the same outputs:
!   ../INPUT/event_list, setup, event_<ev>, relax, stf, recfile_<ev>
!   ../MODELS/MODELS/boxfile and rhoinv<r> mu<r> lambda<r> A<r> B<r> C<r> [Q<r>]
!   ../ADJOINT/<ev>/ad_srcfile, ad_src_<k>              (adjoint_flag=2)
!   <ofd>/<station>.x/.y/.z, vx_<r>_<it> ...            (forward)
!   <ffd>/<ev>/vx_<r> ... eyz_<r>, saving_vector_<r>    (adjoint_flag=1)
!   <ofd>/grad_rho_<r>, grad_cp_<r>, grad_csh_<r>, grad_csv_<r> (adjoint_flag=2)
! for every rank box r of the boxfile (px*py*pz of them, as the reference's
! MPI ranks would read and write them): the boxes' model records are placed
! into one section (SOURCE/ses3d_input.f90:49-170 gives each rank its box,
! :472-501 its records), the section is solved on the GPU, and each box's part
! of the outputs goes to that box's files, in the box's own extent.
! The absorbing-layer width is the reference's compile-time pml
! (SOURCE/ses3d_modules.f90:21, default 2); here it is read at run time from
! the command line (ses3d_hip.exe --pml N) or SES3D_PML.
! Structure follows SOURCE/ses3d_main.f90:77-182 with ses3d_input
! (SOURCE/ses3d_input.f90) restated; the time step, seismograms,
! forward-field storage and kernels run in libses3d_hip.so.
!==============================================================================
program ses3d_hip_main
  use iso_c_binding
  use ses3d_hip_mod
  implicit none

  real, parameter :: pi = 3.1415926535898       ! SOURCE/ses3d_modules.f90:25
  integer, parameter :: lpd = 4

  character(len=1000) :: junk, ofd, ffd, line
  character(len=1000), allocatable :: event_indices(:)
  character(len=12), allocatable :: station_name(:)
  character(len=40) :: s_idx, s_it, s_rank
  integer :: n_events, i_events, pp, px, py, pz, k, i, it, ios, ib, pml
  ! rank boxes: multi index, element index ranges, physical bounds; offsets and
  ! extents of each box's elements in the section
  integer, allocatable :: bm(:,:), blo(:,:), bhi(:,:), boff(:,:), bn(:,:)
  real, allocatable :: bmin(:,:), bmax(:,:)
  integer :: nt, source_type, is_diss, adjoint_flag, samp_ad, ssamp, output_displacement
  integer :: nr_global, nr, nad, npts, n_stored, nsnap
  real :: xmin, xmax, ymin, ymax, zmin, zmax, dt
  real :: xmin_g, xmax_g, ymin_g, ymax_g, zmin_g, zmax_g, xxs, yys, zzs
  real :: mom(6), tau_p(3), D_p(3)
  real, allocatable :: so(:), rec(:,:), seis(:), field(:), hist(:), adloc(:,:), adstf(:,:,:)
  real, allocatable, target :: rhoinv(:), mu(:), lambda(:), A(:), B(:), C(:), Q(:)
  real, allocatable :: vxs(:), vys(:), vzs(:)
  real(c_double), allocatable :: g1(:), g2(:), g3(:), g4(:)
  integer(c_int32_t), allocatable :: ridx(:), sv(:)
  integer :: nx, ny, nz
  type(ses3d_hip_config) :: cfg
  type(c_ptr) :: ctx
  character(len=3), parameter :: hname(9) = ['vx ', 'vy ', 'vz ', 'exx', 'eyy', 'ezz', 'exy', 'exz', 'eyz']

  pml = absorbing_width()
  call system('mkdir -p ../DATA/LOGFILES')
  open(unit=98, file='../INPUT/event_list', status='old', action='read')
  read(98, *) n_events
  allocate(event_indices(n_events))
  do k = 1, n_events
     read(98, *) event_indices(k)
  end do
  close(98)
  open(unit=99, file='../DATA/LOGFILES/logfile0', action='write')
  write(99, *) 'ses3d on one GPU (libses3d_hip), events: ', n_events

  do i_events = 1, n_events
     !---------------------------------------------------------- boxfile
     ! SOURCE/ses3d_input.f90:49-99: every rank box of the section
     open(unit=15, file='../MODELS/MODELS/boxfile', status='old', action='read')
     do k = 1, 14
        read(15, *) junk
     end do
     read(15, *) pp
     read(15, *) px
     read(15, *) py
     read(15, *) pz
     read(15, *) junk
     if (pp /= px * py * pz .or. pp < 1) then
        write(*, *) 'boxfile: ', pp, ' boxes for ', px, 'x', py, 'x', pz, ' processes'
        stop 1
     end if
     allocate(bm(3, pp), blo(3, pp), bhi(3, pp), boff(3, pp), bn(3, pp), bmin(3, pp), bmax(3, pp))
     do ib = 1, pp
        read(15, *) k
        read(15, *) bm(1, ib), bm(2, ib), bm(3, ib)
        read(15, *) blo(1, ib), bhi(1, ib)
        read(15, *) blo(2, ib), bhi(2, ib)
        read(15, *) blo(3, ib), bhi(3, ib)
        read(15, *) bmin(1, ib), bmax(1, ib)
        read(15, *) bmin(2, ib), bmax(2, ib)
        read(15, *) bmin(3, ib), bmax(3, ib)
        read(15, *) junk
        if (k /= ib) then
           write(*, *) 'boxfile: box ', ib, ' is numbered ', k
           stop 1
        end if
     end do
     close(15)
     call place_boxes()
     npts = (nx + 1) * (ny + 1) * (nz + 1) * (lpd + 1)**3
     !---------------------------------------------------------- event file
     open(unit=15, file='../INPUT/event_'//trim(event_indices(i_events)), status='old', action='read')
     read(15, *) junk
     read(15, *) nt
     read(15, *) dt
     read(15, *) junk
     read(15, *) xxs
     read(15, *) yys
     read(15, *) zzs
     read(15, *) source_type
     do k = 1, 6
        read(15, *) mom(k)
     end do
     read(15, *) junk
     read(15, '(A1000)') ofd
     ofd = trim(ofd)//'/'
     call system('mkdir -p '//trim(ofd))
     read(15, *) junk
     read(15, *) ssamp
     read(15, *) output_displacement
     close(15)
     !---------------------------------------------------------- setup file
     open(unit=15, file='../INPUT/setup', status='old', action='read')
     read(15, *) junk
     read(15, *) xmin_g
     read(15, *) xmax_g
     read(15, *) ymin_g
     read(15, *) ymax_g
     read(15, *) zmin_g
     read(15, *) zmax_g
     read(15, *) is_diss
     read(15, *) k
     read(15, *) junk
     do k = 1, 7
        read(15, *) i
     end do
     read(15, *) junk
     read(15, *) adjoint_flag
     read(15, *) samp_ad
     read(15, '(A1000)') ffd
     ffd = trim(ffd)//'/'//trim(event_indices(i_events))//'/'
     call system('mkdir -p '//trim(ffd))
     close(15)
     xxs = xxs * pi / 180
     yys = yys * pi / 180
     !---------------------------------------------------------- relax file
     open(unit=15, file='../INPUT/relax', status='old', action='read')
     read(15, *) junk
     read(15, *) tau_p(1)
     read(15, *) tau_p(2)
     read(15, *) tau_p(3)
     read(15, *) junk
     read(15, *) D_p(1)
     read(15, *) D_p(2)
     read(15, *) D_p(3)
     close(15)
     !---------------------------------------------------------- model
     ! each box's records in its part of the section (SOURCE/ses3d_input.f90:472-501)
     allocate(rhoinv(npts), mu(npts), lambda(npts), A(npts), B(npts), C(npts), Q(npts))
     Q = 1.0
     call read_boxes('../MODELS/MODELS/rhoinv', rhoinv)
     call read_boxes('../MODELS/MODELS/mu', mu)
     call read_boxes('../MODELS/MODELS/lambda', lambda)
     call read_boxes('../MODELS/MODELS/A', A)
     call read_boxes('../MODELS/MODELS/B', B)
     call read_boxes('../MODELS/MODELS/C', C)
     if (is_diss == 1) call read_boxes('../MODELS/MODELS/Q', Q)
     !---------------------------------------------------------- context
     cfg%nx = nx; cfg%ny = ny; cfg%nz = nz; cfg%lpd = lpd; cfg%pml = pml
     cfg%is_diss = is_diss; cfg%adjoint_flag = adjoint_flag; cfg%samp_ad = samp_ad; cfg%nt = nt
     cfg%dt = dt; cfg%xmin = xmin; cfg%xmax = xmax; cfg%ymin = ymin; cfg%ymax = ymax
     cfg%zmin = zmin; cfg%zmax = zmax; cfg%zmax_global = zmax_g
     cfg%tau_p = tau_p; cfg%D_p = D_p
     call hip_check(ses3d_hip_create(cfg, 1_c_int32_t, ctx), 'create')
     if (is_diss == 1) then
        call hip_check(ses3d_hip_set_model(ctx, rhoinv, mu, lambda, A, B, C, c_loc(Q)), 'set_model')
     else
        call hip_check(ses3d_hip_set_model(ctx, rhoinv, mu, lambda, A, B, C, c_null_ptr), 'set_model')
     end if
     deallocate(rhoinv, mu, lambda, A, B, C, Q)

     if (adjoint_flag == 0 .or. adjoint_flag == 1) then
        ! source time function after 4 header lines (SOURCE/ses3d_input.f90:419-428)
        allocate(so(nt))
        open(unit=15, file='../INPUT/stf', status='old', action='read')
        do k = 1, 4
           read(15, *) junk
        end do
        do k = 1, nt
           read(15, *) so(k)
        end do
        close(15)
        call hip_check(ses3d_hip_set_source(ctx, 0_c_int32_t, int(source_type, c_int32_t), xxs, yys, zzs, mom, so), &
             'set_source')
        ! receivers (SOURCE/ses3d_input.f90:531-546)
        open(unit=10, file='../INPUT/recfile_'//trim(event_indices(i_events)), status='old', action='read')
        read(10, *) nr_global
        allocate(station_name(max(nr_global, 1)), rec(3, max(nr_global, 1)))
        do k = 1, nr_global
           read(10, *) station_name(k)
           read(10, *) rec(1, k), rec(2, k), rec(3, k)
           rec(1, k) = rec(1, k) * pi / 180
           rec(2, k) = rec(2, k) * pi / 180
        end do
        close(10)
        call hip_check(ses3d_hip_set_receivers(ctx, 0_c_int32_t, int(nr_global, c_int32_t), rec), 'set_receivers')
     else
        ! adjoint sources (SOURCE/ses3d_input.f90:592-611, ses3d_init.f90:475-492)
        open(unit=10, file='../ADJOINT/'//trim(event_indices(i_events))//'/ad_srcfile', action='read')
        read(10, *) nad
        allocate(adloc(3, max(nad, 1)), adstf(3, nt, max(nad, 1)))
        do k = 1, nad
           read(10, *) adloc(1, k), adloc(2, k), adloc(3, k)
           adloc(1, k) = adloc(1, k) * pi / 180
           adloc(2, k) = adloc(2, k) * pi / 180
        end do
        close(10)
        do k = 1, nad
           write(s_idx, '(I0)') k
           open(unit=10, file='../ADJOINT/'//trim(event_indices(i_events))//'/ad_src_'//trim(s_idx), action='read')
           do i = 1, 4
              read(10, *) junk
           end do
           do i = 1, nt
              read(10, *) adstf(1, i, k), adstf(2, i, k), adstf(3, i, k)
           end do
           close(10)
        end do
        call hip_check(ses3d_hip_set_adjoint_sources(ctx, 0_c_int32_t, int(nad, c_int32_t), adloc, adstf), &
             'set_adjoint_sources')
        ! float32 kernel sums, as the reference's grad_* arrays: its kernel files bit for bit
        call hip_check(ses3d_hip_set_kernel_precision(ctx, 32_c_int32_t), 'set_kernel_precision')
     end if
     call hip_check(ses3d_hip_init(ctx), 'init')

     if (adjoint_flag == 2) then
        ! forward history written by an adjoint_flag=1 run, in time order
        allocate(sv(nt))
        open(unit=10, file=trim(ffd)//'saving_vector_0', action='read')
        do k = 1, nt
           read(10, *) sv(k)
        end do
        close(10)
        n_stored = count(sv > 0)
        allocate(hist(9 * npts))
        do ib = 1, pp
           write(s_rank, '(I0)') ib - 1
           do k = 1, 9
              open(unit=100 + 9 * (ib - 1) + k, file=trim(ffd)//trim(hname(k))//'_'//trim(s_rank), &
                   action='read', form='unformatted')
           end do
        end do
        do it = 1, nt
           if (sv(it) <= 0) cycle
           do ib = 1, pp
              allocate(field(box_points(ib)))
              do k = 1, 9
                 read(100 + 9 * (ib - 1) + k) field
                 call place(hist((k - 1) * npts + 1:k * npts), field, ib)
              end do
              deallocate(field)
           end do
           call hip_check(ses3d_hip_put_history(ctx, 0_c_int32_t, int(it, c_int32_t), hist), 'put_history')
        end do
        do ib = 1, pp
           do k = 1, 9
              close(100 + 9 * (ib - 1) + k)
           end do
        end do
        deallocate(hist, sv)
     end if

     !---------------------------------------------------------- time loop
     allocate(vxs(npts), vys(npts), vzs(npts))
     it = 1
     do while (it <= nt)
        ! the next step whose snapshot is due (SOURCE/ses3d_main.f90:132-136, :180)
        nsnap = nt + 1
        if (output_displacement == 1) then
           do k = it, nt
              if (mod(k, ssamp) == 0 .or. k == nt) then
                 nsnap = k
                 exit
              end if
           end do
        end if
        if (nsnap <= nt) then
           if (nsnap > it) call hip_check(ses3d_hip_run(ctx, int(it, c_int32_t), int(nsnap - it, c_int32_t)), 'run')
           ! the snapshot of step nsnap holds the velocity gathered at its start
           call hip_check(ses3d_hip_get_field(ctx, 0_c_int32_t, 'vx'//c_null_char, vxs), 'get_field')
           call hip_check(ses3d_hip_get_field(ctx, 0_c_int32_t, 'vy'//c_null_char, vys), 'get_field')
           call hip_check(ses3d_hip_get_field(ctx, 0_c_int32_t, 'vz'//c_null_char, vzs), 'get_field')
           call hip_check(ses3d_hip_run(ctx, int(nsnap, c_int32_t), 1_c_int32_t), 'run')
           write(s_it, '(I0)') nsnap
           call write_boxes(trim(ofd)//'vx_', '_'//trim(s_it), vxs)
           call write_boxes(trim(ofd)//'vy_', '_'//trim(s_it), vys)
           call write_boxes(trim(ofd)//'vz_', '_'//trim(s_it), vzs)
           it = nsnap + 1
        else
           call hip_check(ses3d_hip_run(ctx, int(it, c_int32_t), int(nt - it + 1, c_int32_t)), 'run')
           it = nt + 1
        end if
        write(99, *) 'iteration', it - 1
     end do
     call hip_check(ses3d_hip_synchronize(ctx), 'synchronize')
     deallocate(vxs, vys, vzs)

     !---------------------------------------------------------- outputs
     if (adjoint_flag == 1) then
        ! forward field records and saving vector (SOURCE/ses3d_output.f90:138-156, :379-392)
        allocate(sv(nt), hist(9 * npts))
        call hip_check(ses3d_hip_get_saving_vector(ctx, sv), 'get_saving_vector')
        do ib = 1, pp
           write(s_rank, '(I0)') ib - 1
           do k = 1, 9
              open(unit=100 + 9 * (ib - 1) + k, file=trim(ffd)//trim(hname(k))//'_'//trim(s_rank), &
                   action='write', form='unformatted')
           end do
        end do
        do it = 1, nt
           if (sv(it) <= 0) cycle
           call hip_check(ses3d_hip_get_history(ctx, 0_c_int32_t, int(it, c_int32_t), hist), 'get_history')
           do ib = 1, pp
              allocate(field(box_points(ib)))
              do k = 1, 9
                 call take(field, hist((k - 1) * npts + 1:k * npts), ib)
                 write(100 + 9 * (ib - 1) + k) field
              end do
              deallocate(field)
           end do
        end do
        do ib = 1, pp
           do k = 1, 9
              close(100 + 9 * (ib - 1) + k)
           end do
           write(s_rank, '(I0)') ib - 1
           open(unit=10, file=trim(ffd)//'saving_vector_'//trim(s_rank), action='write')
           do k = 1, nt
              write(10, *) sv(k)
           end do
           close(10)
        end do
        deallocate(sv, hist)
     end if
     if (adjoint_flag == 2) then
        ! kernels (SOURCE/ses3d_output.f90:318-343): float32 sums, returned widened, written as float32
        allocate(g1(npts), g2(npts), g3(npts), g4(npts), field(npts))
        call hip_check(ses3d_hip_get_kernels(ctx, 0_c_int32_t, g1, g2, g3, g4, c_null_ptr), 'get_kernels')
        field = real(g1); call write_boxes(trim(ofd)//'grad_rho_', '', field)
        field = real(g2); call write_boxes(trim(ofd)//'grad_cp_', '', field)
        field = real(g3); call write_boxes(trim(ofd)//'grad_csh_', '', field)
        field = real(g4); call write_boxes(trim(ofd)//'grad_csv_', '', field)
        deallocate(g1, g2, g3, g4, field)
        deallocate(adloc, adstf)
     else
        call hip_check(ses3d_hip_get_nr(ctx, 0_c_int32_t, nr), 'get_nr')
        allocate(seis(max(3 * nr * nt, 1)), ridx(max(nr, 1)))
        call hip_check(ses3d_hip_get_seismograms(ctx, 0_c_int32_t, seis), 'get_seismograms')
        call hip_check(ses3d_hip_get_receiver_index(ctx, 0_c_int32_t, ridx), 'get_receiver_index')
        call write_seismograms()
        deallocate(seis, ridx, so, station_name, rec)
     end if
     call ses3d_hip_destroy(ctx)
     deallocate(bm, blo, bhi, boff, bn, bmin, bmax)
  end do
  close(99)

contains

  ! Absorbing-layer width in elements: --pml N on the command line, else
  ! SES3D_PML, else the reference's default of 2.
  integer function absorbing_width()
    character(len=64) :: arg
    integer :: n, st
    absorbing_width = 2
    call get_environment_variable('SES3D_PML', arg, status=st)
    if (st == 0 .and. len_trim(arg) > 0) read(arg, *) absorbing_width
    do n = 1, command_argument_count() - 1
       call get_command_argument(n, arg)
       if (trim(arg) == '--pml') then
          call get_command_argument(n + 1, arg)
          read(arg, *) absorbing_width
       end if
    end do
  end function absorbing_width

  ! The section the boxes tile.  Box b (multi index (i,j,k)) holds elements
  ! nx_min+i-1 .. nx_max+i-1 along theta, likewise along phi, counted from the
  ! section's minimum; radially nz_min+k-1 .. nz_max+k-1 counted from zmin,
  ! while element index k of the arrays runs down from zmax
  ! (MODELS/SOURCE/generate_models.f90:200-282, SOURCE/ses3d_init.f90:239).
  subroutine place_boxes()
    integer :: a, ntot(3), nsum
    do a = 1, 3
       ntot(a) = maxval(bhi(a, :) + bm(a, :))
    end do
    nsum = 0
    do ib = 1, pp
       do a = 1, 3
          bn(a, ib) = bhi(a, ib) - blo(a, ib) + 1
          boff(a, ib) = blo(a, ib) + bm(a, ib) - 1
       end do
       boff(3, ib) = ntot(3) - (bhi(3, ib) + bm(3, ib))
       nsum = nsum + bn(1, ib) * bn(2, ib) * bn(3, ib)
       if (any(boff(:, ib) < 0) .or. any(boff(:, ib) + bn(:, ib) > ntot)) then
          write(*, *) 'boxfile: box ', ib, ' lies outside the section'
          stop 1
       end if
    end do
    if (nsum /= ntot(1) * ntot(2) * ntot(3)) then
       write(*, *) 'boxfile: the boxes do not tile a section of ', ntot(1), 'x', ntot(2), 'x', ntot(3), ' elements'
       stop 1
    end if
    nx = ntot(1) - 1; ny = ntot(2) - 1; nz = ntot(3) - 1
    do ib = 1, pp
       if (bm(1, ib) == 1) xmin = bmin(1, ib)
       if (bm(1, ib) == px) xmax = bmax(1, ib)
       if (bm(2, ib) == 1) ymin = bmin(2, ib)
       if (bm(2, ib) == py) ymax = bmax(2, ib)
       if (bm(3, ib) == 1) zmin = bmin(3, ib)
       if (bm(3, ib) == pz) zmax = bmax(3, ib)
    end do
  end subroutine place_boxes

  integer function box_points(bb)
    integer, intent(in) :: bb
    box_points = bn(1, bb) * bn(2, bb) * bn(3, bb) * (lpd + 1)**3
  end function box_points

  ! copy box bb's field (its own extent, reference layout) into the section's
  subroutine place(g, a, bb)
    real, intent(inout) :: g(:)
    real, intent(in) :: a(:)
    integer, intent(in) :: bb
    integer :: ii, jj, kk, q
    integer(8) :: ia, ig
    ia = 0
    do q = 0, (lpd + 1)**3 - 1
       do kk = 0, bn(3, bb) - 1
          do jj = 0, bn(2, bb) - 1
             do ii = 0, bn(1, bb) - 1
                ia = ia + 1
                ig = 1 + (ii + boff(1, bb)) + int(nx + 1, 8) * ((jj + boff(2, bb)) + int(ny + 1, 8) * &
                     ((kk + boff(3, bb)) + int(nz + 1, 8) * q))
                g(ig) = a(ia)
             end do
          end do
       end do
    end do
  end subroutine place

  ! box bb's part of a section field
  subroutine take(a, g, bb)
    real, intent(out) :: a(:)
    real, intent(in) :: g(:)
    integer, intent(in) :: bb
    integer :: ii, jj, kk, q
    integer(8) :: ia, ig
    ia = 0
    do q = 0, (lpd + 1)**3 - 1
       do kk = 0, bn(3, bb) - 1
          do jj = 0, bn(2, bb) - 1
             do ii = 0, bn(1, bb) - 1
                ia = ia + 1
                ig = 1 + (ii + boff(1, bb)) + int(nx + 1, 8) * ((jj + boff(2, bb)) + int(ny + 1, 8) * &
                     ((kk + boff(3, bb)) + int(nz + 1, 8) * q))
                a(ia) = g(ig)
             end do
          end do
       end do
    end do
  end subroutine take

  ! <prefix><rank> of every box into the section field
  subroutine read_boxes(prefix, g)
    character(len=*), intent(in) :: prefix
    real, intent(inout) :: g(:)
    real, allocatable :: a(:)
    character(len=40) :: sr
    integer :: bb
    do bb = 1, pp
       write(sr, '(I0)') bb - 1
       allocate(a(box_points(bb)))
       call read_field(prefix//trim(sr), a)
       call place(g, a, bb)
       deallocate(a)
    end do
  end subroutine read_boxes

  ! each box's part of a section field to <prefix><rank><suffix>
  subroutine write_boxes(prefix, suffix, g)
    character(len=*), intent(in) :: prefix, suffix
    real, intent(in) :: g(:)
    real, allocatable :: a(:)
    character(len=40) :: sr
    integer :: bb
    do bb = 1, pp
       write(sr, '(I0)') bb - 1
       allocate(a(box_points(bb)))
       call take(a, g, bb)
       call write_field(prefix//trim(sr)//suffix, a)
       deallocate(a)
    end do
  end subroutine write_boxes

  subroutine read_field(fn, a)
    character(len=*), intent(in) :: fn
    real, intent(out) :: a(:)
    open(unit=17, file=fn, status='old', action='read', form='unformatted')
    read(17) a
    close(17)
  end subroutine read_field

  subroutine write_field(fn, a)
    character(len=*), intent(in) :: fn
    real, intent(in) :: a(:)
    open(unit=18, file=fn, action='write', form='unformatted')
    write(18) a
    close(18)
  end subroutine write_field

  ! SOURCE/ses3d_output.f90:398-451: same list-directed records per file
  subroutine write_seismograms()
    integer :: r, j, g
    character(len=12) :: sname
    do r = 1, nr
       g = ridx(r) + 1
       sname = station_name(g)
       open(unit=51, file=trim(ofd)//sname//'.x', action='write')
       open(unit=52, file=trim(ofd)//sname//'.y', action='write')
       open(unit=53, file=trim(ofd)//sname//'.z', action='write')
       write(51, *) 'theta component seismograms'
       write(52, *) 'phi component seismograms'
       write(53, *) 'r component seismograms'
       write(51, *) 'nt=', nt
       write(51, *) 'dt=', dt
       write(52, *) 'nt=', nt
       write(52, *) 'dt=', dt
       write(53, *) 'nt=', nt
       write(53, *) 'dt=', dt
       write(51, *) 'receiver location (colat [deg],lon [deg],depth [m])'
       write(52, *) 'receiver location (colat [deg],lon [deg],depth [m])'
       write(53, *) 'receiver location (colat [deg],lon [deg],depth [m])'
       write(51, *) 'x=', rec(1, g) * 180 / pi, 'y=', rec(2, g) * 180 / pi, 'z=', rec(3, g)
       write(52, *) 'x=', rec(1, g) * 180 / pi, 'y=', rec(2, g) * 180 / pi, 'z=', rec(3, g)
       write(53, *) 'x=', rec(1, g) * 180 / pi, 'y=', rec(2, g) * 180 / pi, 'z=', rec(3, g)
       write(51, *) 'source location (colat [deg],lon [deg],depth [m])'
       write(52, *) 'source location (colat [deg],lon [deg],depth [m])'
       write(53, *) 'source location (colat [deg],lon [deg],depth [m])'
       write(51, *) 'x=', xxs * 180 / pi, 'y=', yys * 180 / pi, 'z=', zzs
       write(52, *) 'x=', xxs * 180 / pi, 'y=', yys * 180 / pi, 'z=', zzs
       write(53, *) 'x=', xxs * 180 / pi, 'y=', yys * 180 / pi, 'z=', zzs
       do j = 1, nt
          write(51, *) seis((0 * nr + (r - 1)) * nt + j)
          write(52, *) seis((1 * nr + (r - 1)) * nt + j)
          write(53, *) seis((2 * nr + (r - 1)) * nt + j)
       end do
       close(51)
       close(52)
       close(53)
    end do
  end subroutine write_seismograms

end program ses3d_hip_main
