!******************************************************************************
! pic2d_gpu: the time loop of xiaocanli/pbpic's `program pic2d`
! (src/programs/pic2d.f90:148-192) driving the MI355X step through the
! drop-in subroutines of module pbpic_gpu.  The initial state is the seeded
! thermal plasma of the CPU-runnable benchmark case (128x128 cells,
! dx = 0.5 d_e, dt = 0.2/w_pe, 16 particles per cell and species,
! vth_e = 0.1c, vth_i = 0.02c) instead of the /dev/urandom-seeded Harris
! sheet of initparticles (:357-578).
!
! Usage: pic2d_gpu [ntime] [precision 0|1]; prints itime and
! totene(1:4) every ntdiag steps like energyinfo (:1996-2115).
!******************************************************************************
program pic2d_gpu
    use iso_c_binding
    use pbpic_gpu
    implicit none
    integer, parameter :: nx = 128, ny = 128, ppg = 16, ntdiag = 10
    real(c_double), parameter :: mime = 25.0d0, cva = 15.0d0
    real(c_double) :: dx, dt, in0
    integer :: itime, ntime, prec, nargs, mode
    character(len=32) :: arg
    real(c_double) :: d(16)

    ntime = 100
    prec = PBPIC_PREC_FP64
    nargs = command_argument_count()
    if (nargs >= 1) then
        call get_command_argument(1, arg)
        read(arg, *) ntime
    end if
    if (nargs >= 2) then
        call get_command_argument(2, arg)
        read(arg, *) prec
    end if
    ! mode 0: the reference's loop body; mode 1: one pbpic_step call per step
    mode = 0
    if (nargs >= 3) then
        call get_command_argument(3, arg)
        read(arg, *) mode
    end if

    ! electron units mapped onto the reference normalization: dx = 0.5 d_e,
    ! dt = 0.2/w_pe with w_pe = cva*sqrt(mime) for in0 = 1/ppg
    dx = 0.5d0 / sqrt(mime)
    dt = 0.2d0 / (cva * sqrt(mime))
    in0 = 1.0d0 / ppg

    call pbpic_init(nx, ny, dx, dx, dt, cva, mime, in0, prec, 50)
    call pbpic_check(c_add_species(ctx, 1_c_int32_t, int(nx*ny*ppg, c_int64_t), sp_ion), 'species')
    call pbpic_check(c_add_species(ctx, -1_c_int32_t, int(nx*ny*ppg, c_int64_t), sp_ele), 'species')
    call pbpic_check(c_init_maxwellian(ctx, sp_ion, ppg, 0.02d0, 0.0d0, 0, 20260101_c_int64_t, 1), 'init')
    call pbpic_check(c_init_maxwellian(ctx, sp_ele, ppg, 0.1d0, 0.0d0, 0, 20260101_c_int64_t, 0), 'init')

    call countingsort                                   ! pic2d.f90:161
    do itime = 1, ntime                                 ! :171
        if (mode == 1) then
            call pbpic_step(1)                          ! :173-182 in one call
        else
            call zerocurrent                            ! :173-176
            call pusher                                 ! :177
            call fieldsolver                            ! :178
            if (mod(itime, 50) == 0) call countingsort  ! :180-182
        end if
        if (mod(itime, ntdiag) == 0) then               ! :183-187
            call fieldsenergy
            call pbpic_check(c_diag(ctx, d), 'diag')
            write(*, '(a,i6,4es16.8,a,es12.4)') 'itime', itime, totene, '  gauss_dlinf', d(10)
        end if
    end do
    call releasememory                                  ! :189
end program pic2d_gpu
