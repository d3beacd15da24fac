!******************************************************************************
! pbpic_gpu: ISO_C_BINDING interface to libpbpic.so (include/pbpic.h) and
! drop-in replacements for the step subroutines of xiaocanli/pbpic
! (src/programs/pic2d.f90): pusher (:912), currentdeposit (:1226),
! fieldsolver (:678), fieldsenergy (:653) and countingsort (:1303).  As in the
! reference they take no arguments (currentdeposit aside) and work on
! module-level state -- here a device context instead of module arrays.
! A failing call prints the library's message and stops, as the reference
! stops through MPI's default error handler.
!******************************************************************************
module pbpic_gpu
    use iso_c_binding
    implicit none

    integer(c_int32_t), parameter :: PBPIC_PREC_FP64 = 0, PBPIC_PREC_SPLIT32 = 1
    integer(c_int32_t), parameter :: PBPIC_COMPAT_REFERENCE = 7
    integer(c_int32_t), parameter :: PBPIC_FLAG_DEPOSIT_RHO = 1, PBPIC_FLAG_TRACK_PID = 2
    integer(c_int32_t), parameter :: PBPIC_FLAG_PUSH_ORDER = 4
    integer(c_int32_t), parameter :: PBPIC_FLAG_MARDER = 8, PBPIC_FLAG_CHERENKOV = 16
    integer(c_int32_t), parameter :: PBPIC_YBOUND_PERIODIC = 0, PBPIC_YBOUND_CONDUCTING = 1
    integer(c_int32_t), parameter :: PBPIC_HARRIS_DOUBLE_X = 0, PBPIC_HARRIS_SINGLE_Y = 1
    integer, parameter :: PBPIC_NDIAG = 16

    type, bind(c) :: pbpic_params
        integer(c_int32_t) :: nx, ny, precision, compat
        real(c_double) :: dx, dy, dt, cva, mime, in0, rho_background
        integer(c_int32_t) :: sort_every, flags
        integer(c_int32_t) :: ybound
        integer(c_int32_t) :: reserved(7)
    end type

    ! pbpic_harris: parameters of the Harris-sheet loader that replaces
    ! initparticles/initfields (pic2d.f90:357-648; constants :22-30)
    type, bind(c) :: pbpic_harris
        integer(c_int32_t) :: geometry = PBPIC_HARRIS_DOUBLE_X, ppc = 64
        real(c_double) :: half_width = 0.5d0, nb = 0.2d0, b0 = 1.0d0, psi0 = 0.0d0
        real(c_double) :: vthe_cs = 0.092d0, vthi_cs = 0.0411d0
        real(c_double) :: vthe_bg = 0.0291d0, vthi_bg = 0.0058d0
        real(c_double) :: vde_z = -0.0667d0, vdi_z = 0.0667d0
        integer(c_int32_t) :: thermal_z = 0, reserved0 = 0
        integer(c_int64_t) :: seed = 1
        integer(c_int32_t) :: stream = 0
        integer(c_int32_t) :: reserved(7) = 0
    end type

    interface
        subroutine c_params_default(p) bind(c, name="pbpic_params_default")
            import :: pbpic_params
            type(pbpic_params), intent(out) :: p
        end subroutine
        integer(c_int32_t) function c_ctx_create(p, ctx) bind(c, name="pbpic_ctx_create")
            import :: pbpic_params, c_int32_t, c_ptr
            type(pbpic_params), intent(in) :: p
            type(c_ptr), intent(out) :: ctx
        end function
        integer(c_int32_t) function c_ctx_destroy(ctx) bind(c, name="pbpic_ctx_destroy")
            import :: c_ptr, c_int32_t
            type(c_ptr), value :: ctx
        end function
        type(c_ptr) function c_last_error(ctx) bind(c, name="pbpic_last_error")
            import :: c_ptr
            type(c_ptr), value :: ctx
        end function
        integer(c_int32_t) function c_add_species(ctx, q, cap, sp) bind(c, name="pbpic_add_species")
            import :: c_ptr, c_int32_t, c_int64_t
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: q
            integer(c_int64_t), value :: cap
            integer(c_int32_t), intent(out) :: sp
        end function
        integer(c_int32_t) function c_init_maxwellian(ctx, sp, ppc, vth, vd, cs, seed, stream) &
                bind(c, name="pbpic_init_maxwellian")
            import :: c_ptr, c_int32_t, c_double, c_int64_t
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: sp, ppc, cs, stream
            real(c_double), value :: vth, vd
            integer(c_int64_t), value :: seed
        end function
        integer(c_int32_t) function c_init_harris(ctx, si, se, h) bind(c, name="pbpic_init_harris")
            import :: c_ptr, c_int32_t, pbpic_harris
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: si, se
            type(pbpic_harris), intent(in) :: h
        end function
        integer(c_int32_t) function c_zero_current(ctx) bind(c, name="pbpic_zero_current")
            import :: c_ptr, c_int32_t
            type(c_ptr), value :: ctx
        end function
        integer(c_int32_t) function c_push(ctx) bind(c, name="pbpic_push")
            import :: c_ptr, c_int32_t
            type(c_ptr), value :: ctx
        end function
        integer(c_int32_t) function c_deposit(ctx, n, i1, i2, j1, j2, dx1, dx2, dy1, dy2, q, vz) &
                bind(c, name="pbpic_deposit")
            import :: c_ptr, c_int32_t, c_int64_t, c_double
            type(c_ptr), value :: ctx
            integer(c_int64_t), value :: n
            integer(c_int32_t), intent(in) :: i1(*), i2(*), j1(*), j2(*), q(*)
            real(c_double), intent(in) :: dx1(*), dx2(*), dy1(*), dy2(*), vz(*)
        end function
        integer(c_int32_t) function c_fieldsolve(ctx) bind(c, name="pbpic_fieldsolve")
            import :: c_ptr, c_int32_t
            type(c_ptr), value :: ctx
        end function
        integer(c_int32_t) function c_fieldsenergy(ctx, out) bind(c, name="pbpic_fieldsenergy")
            import :: c_ptr, c_int32_t, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: out(2)
        end function
        integer(c_int32_t) function c_kineticenergy(ctx, out) bind(c, name="pbpic_kineticenergy")
            import :: c_ptr, c_int32_t, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: out(2)
        end function
        integer(c_int32_t) function c_countingsort(ctx) bind(c, name="pbpic_countingsort")
            import :: c_ptr, c_int32_t
            type(c_ptr), value :: ctx
        end function
        integer(c_int32_t) function c_step(ctx, nsteps) bind(c, name="pbpic_step")
            import :: c_ptr, c_int32_t
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: nsteps
        end function
        integer(c_int32_t) function c_diag(ctx, out) bind(c, name="pbpic_diag")
            import :: c_ptr, c_int32_t, c_double
            type(c_ptr), value :: ctx
            real(c_double), intent(out) :: out(16)
        end function
        integer(c_int32_t) function c_get_field(ctx, which, plane) bind(c, name="pbpic_get_field")
            import :: c_ptr, c_int32_t, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: which
            real(c_double), intent(out) :: plane(*)
        end function
        integer(c_size_t) function c_strlen(s) bind(c, name="strlen")
            import :: c_ptr, c_size_t
            type(c_ptr), value :: s
        end function
    end interface

    ! module state: the device context replaces the reference's module arrays
    type(c_ptr), save :: ctx = c_null_ptr
    integer(c_int32_t), save :: sp_ion = -1, sp_ele = -1
    real(c_double), save :: totene(4) = 0.0d0

contains

    subroutine pbpic_check(rc, where)
        integer(c_int32_t), intent(in) :: rc
        character(*), intent(in) :: where
        character(kind=c_char), pointer :: msg(:)
        integer(c_size_t) :: n
        type(c_ptr) :: p
        if (rc == 0) return
        p = c_last_error(ctx)
        n = c_strlen(p)
        call c_f_pointer(p, msg, [n])
        write(*, '(a,a,i0,a,*(a))') where, ': pbpic error ', rc, ': ', msg
        error stop 1
    end subroutine

    ! Context creation: the allocation part of init (pic2d.f90:197-308).
    ! push_order = .false.: sorts give the reference's exact permutation
    ! (pic2d.f90:1303-1329) instead of the push kernel's tile order
    ! ybound = PBPIC_YBOUND_CONDUCTING: conducting walls in y; ny then counts
    ! two guard rows above the domain (see include/pbpic.h)
    subroutine pbpic_init(nx, ny, dx, dy, dt, cva, mime, in0, precision, sort_every, push_order, &
                          ybound)
        integer, intent(in) :: nx, ny, precision, sort_every
        real(c_double), intent(in) :: dx, dy, dt, cva, mime, in0
        logical, intent(in), optional :: push_order
        integer, intent(in), optional :: ybound
        type(pbpic_params) :: p
        call c_params_default(p)
        p%nx = nx; p%ny = ny; p%dx = dx; p%dy = dy; p%dt = dt
        p%cva = cva; p%mime = mime; p%in0 = in0
        p%precision = precision; p%sort_every = sort_every
        if (present(push_order)) then
            if (.not. push_order) p%flags = iand(p%flags, not(PBPIC_FLAG_PUSH_ORDER))
        end if
        if (present(ybound)) p%ybound = ybound
        call pbpic_check(c_ctx_create(p, ctx), 'pbpic_ctx_create')
    end subroutine

    ! initparticles + initfields (pic2d.f90:357-648) as the seeded device
    ! loader: fills the module's ion and electron species and sets E, B
    subroutine initharris(h)
        type(pbpic_harris), intent(in) :: h
        call pbpic_check(c_init_harris(ctx, sp_ion, sp_ele, h), 'initharris')
    end subroutine

    ! releasememory (pic2d.f90:2557-2582)
    subroutine releasememory
        if (c_associated(ctx)) call pbpic_check(c_ctx_destroy(ctx), 'releasememory')
        ctx = c_null_ptr
    end subroutine

    ! denjr = 0 (pic2d.f90:173-176)
    subroutine zerocurrent
        call pbpic_check(c_zero_current(ctx), 'zerocurrent')
    end subroutine

    ! pusher (pic2d.f90:912-1217): gather, Boris push, move, deposit, wrap, KE
    ! The kinetic-energy tally totene(1:2) (:1061-1062) is taken with the field
    ! energies in fieldsenergy, from the pushed momenta: the pusher itself never
    ! waits for the device.
    subroutine pusher
        call pbpic_check(c_push(ctx), 'pusher')
    end subroutine

    ! currentdeposit (pic2d.f90:1226-1296) for one particle; uz is the velocity
    ! vz passed by the reference's pusher (:1028-1030).
    subroutine currentdeposit(i1, i2, j1, j2, dx1, dx2, dy1, dy2, q0, uz)
        integer, intent(in) :: i1, i2, j1, j2, q0
        real(c_double), intent(in) :: dx1, dx2, dy1, dy2, uz
        call pbpic_check(c_deposit(ctx, 1_c_int64_t, [int(i1, c_int32_t)], [int(i2, c_int32_t)], &
            [int(j1, c_int32_t)], [int(j2, c_int32_t)], [dx1], [dx2], [dy1], [dy2], &
            [int(q0, c_int32_t)], [uz]), 'currentdeposit')
    end subroutine

    ! fieldsolver (pic2d.f90:678-860) incl. fieldsenergy (:848)
    subroutine fieldsolver
        call pbpic_check(c_fieldsolve(ctx), 'fieldsolver')
    end subroutine

    ! fieldsenergy (pic2d.f90:653-673), with the kinetic energies of the last
    ! push (:1061-1062): all of totene(1:4)
    subroutine fieldsenergy
        real(c_double) :: e(2), k(2)
        call pbpic_check(c_fieldsenergy(ctx, e), 'fieldsenergy')
        call pbpic_check(c_kineticenergy(ctx, k), 'fieldsenergy')
        totene(1) = k(1)
        totene(2) = k(2)
        totene(3) = e(1)
        totene(4) = e(2)
    end subroutine

    ! countingsort (pic2d.f90:1303-1329)
    subroutine countingsort
        call pbpic_check(c_countingsort(ctx), 'countingsort')
    end subroutine

    ! nsteps whole time-loop bodies (zerocurrent, pusher, fieldsolver and the
    ! sort at the context's cadence, pic2d.f90:171-188) in one call
    subroutine pbpic_step(nsteps)
        integer, intent(in) :: nsteps
        call pbpic_check(c_step(ctx, int(nsteps, c_int32_t)), 'pbpic_step')
    end subroutine

end module pbpic_gpu
