! TEST INFRASTRUCTURE ONLY: a main program for the reference solver's own
! subroutines, compiled unmodified from SOURCE/ with ses3d_grad_Q.f90 linked in
! place of ses3d_grad.f90 (oracle/Makefile, target ref).  It runs the adjoint
! pass of one event list through the reference's step order
! (SOURCE/ses3d_main.f90:77-182 with adjoint_flag=2: ses3d_evolution, then
! ses3d_grad, snapshots every ssamp steps, ses3d_output(nt) after the loop) and
! then writes grad_Q_mu, which ses3d_grad_Q.f90:111-128,151-153 computes but the
! reference never writes (its writer is commented out at
! SOURCE/ses3d_output.f90:346-370), in the record format of the other kernel
! files: one Fortran unformatted record over the full (0:nx_max,...) array.
!
! Run from <run tree>/MAIN like ses3d.exe; the forward pass (adjoint_flag=1)
! is left to ses3d.exe itself.
program ref_grad_q_harness
use parameters
use variables
implicit none
include 'mpif.h'

    character(len=20) :: rank_str
    character(len=1000) :: fn

    call mpi_init(ierr)
    call mpi_comm_rank(mpi_comm_world, my_rank, ierr)
    call mpi_comm_size(mpi_comm_world, p, ierr)

    if (my_rank == 0) then
        open(unit=98, file='../INPUT/event_list', action='read')
        read(98, *) n_events
        do i_events = 1, n_events
            read(98, *) event_indices(i_events)
        enddo
        close(unit=98)
    endif
    call mpi_bcast(n_events, 1, mpi_integer, 0, mpi_comm_world, ierr)
    call mpi_bcast(event_indices, 1000 * 1000, mpi_character, 0, mpi_comm_world, ierr)

    call int2str(my_rank, rank_str)
    ! ses3d_input / ses3d_init log to unit 99
    open(unit=99, file='harness_log_'//trim(rank_str), action='write')

    do i_events = 1, n_events
        call ses3d_input
        if (adjoint_flag /= 2) then
            write(*, *) 'ref_grad_q_harness runs adjoint_flag=2 passes only'
            call mpi_abort(mpi_comm_world, 1, ierr)
        endif
        call ses3d_init
        do it = 1, nt
            call ses3d_evolution
            if (mod(it, ssamp) == 0) call ses3d_output(it)
            call ses3d_grad
        enddo
        call ses3d_output(nt)
        fn = trim(ofd)//'grad_Q_mu_'//trim(rank_str)
        open(unit=10, file=fn, action='write', form='unformatted')
        write(10) grad_Q_mu
        close(unit=10)
    enddo

    close(unit=99)
    call mpi_finalize(ierr)
end program ref_grad_q_harness
