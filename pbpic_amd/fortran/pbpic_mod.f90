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
    ! field planes (include/pbpic.h PBPIC_EX ...): emf%ex .. emf%bz, denjr%jx ..
    integer(c_int32_t), parameter :: PBPIC_EX = 0, PBPIC_EY = 1, PBPIC_EZ = 2
    integer(c_int32_t), parameter :: PBPIC_BX = 3, PBPIC_BY = 4, PBPIC_BZ = 5
    integer(c_int32_t), parameter :: PBPIC_JX = 6, PBPIC_JY = 7, PBPIC_JZ = 8, PBPIC_RHO = 9

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
        integer(c_int32_t) function c_set_field(ctx, which, plane) bind(c, name="pbpic_set_field")
            import :: c_ptr, c_int32_t, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: which
            real(c_double), intent(in) :: plane(*)
        end function
        integer(c_int32_t) function c_sync(ctx) bind(c, name="pbpic_sync")
            import :: c_ptr, c_int32_t
            type(c_ptr), value :: ctx
        end function
        integer(c_int32_t) function c_num_particles(ctx, sp, n) bind(c, name="pbpic_num_particles")
            import :: c_ptr, c_int32_t, c_int64_t
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: sp
            integer(c_int64_t), intent(out) :: n
        end function
        integer(c_int32_t) function c_countingsort_species(ctx, sp, counts, perm) &
                bind(c, name="pbpic_countingsort_species")
            import :: c_ptr, c_int32_t
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: sp
            type(c_ptr), value :: counts, perm
        end function
        integer(c_int32_t) function c_set_particles(ctx, sp, n, grid, dx, dy, ux, uy, uz, pid) &
                bind(c, name="pbpic_set_particles")
            import :: c_ptr, c_int32_t, c_int64_t, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: sp
            integer(c_int64_t), value :: n
            integer(c_int32_t), intent(in) :: grid(*), pid(*)
            real(c_double), intent(in) :: dx(*), dy(*), ux(*), uy(*), uz(*)
        end function
        integer(c_int32_t) function c_set_particles_f32(ctx, sp, n, grid, dx, dy, ux, uy, uz, pid) &
                bind(c, name="pbpic_set_particles_f32")
            import :: c_ptr, c_int32_t, c_int64_t, c_float
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: sp
            integer(c_int64_t), value :: n
            integer(c_int32_t), intent(in) :: grid(*), pid(*)
            real(c_float), intent(in) :: dx(*), dy(*), ux(*), uy(*), uz(*)
        end function
        integer(c_int32_t) function c_get_particles(ctx, sp, cap, n, grid, dx, dy, ux, uy, uz, pid) &
                bind(c, name="pbpic_get_particles")
            import :: c_ptr, c_int32_t, c_int64_t, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: sp
            integer(c_int64_t), value :: cap
            integer(c_int64_t), intent(out) :: n
            integer(c_int32_t), intent(out) :: grid(*), pid(*)
            real(c_double), intent(out) :: dx(*), dy(*), ux(*), uy(*), uz(*)
        end function
        integer(c_int32_t) function c_get_particles_f32(ctx, sp, cap, n, grid, dx, dy, ux, uy, uz, &
                pid) bind(c, name="pbpic_get_particles_f32")
            import :: c_ptr, c_int32_t, c_int64_t, c_float
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: sp
            integer(c_int64_t), value :: cap
            integer(c_int64_t), intent(out) :: n
            integer(c_int32_t), intent(out) :: grid(*), pid(*)
            real(c_float), intent(out) :: dx(*), dy(*), ux(*), uy(*), uz(*)
        end function
        integer(c_size_t) function c_strlen(s) bind(c, name="strlen")
            import :: c_ptr, c_size_t
            type(c_ptr), value :: s
        end function
    end interface

    ! Host state of the reference (pic2d.f90:47-57, :391, :593) crosses in
    ! either real kind: the reference's default `real` (4 bytes) or the
    ! -fdefault-real-8 build (8 bytes); generic resolution picks the entry.
    interface pbpic_put_ptl
        module procedure put_ptl_r4, put_ptl_r8
    end interface
    interface pbpic_take_ptl
        module procedure take_ptl_r4, take_ptl_r8
    end interface
    interface pbpic_put_field
        module procedure put_field_r4, put_field_r8
    end interface
    interface pbpic_take_field
        module procedure take_field_r4, take_field_r8
    end interface
    interface currentdeposit
        module procedure currentdeposit_r4, currentdeposit_r8
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
    ! track_pid = .true.: keep ptl%pid on the device (pbpic_take_ptl returns it)
    subroutine pbpic_init(nx, ny, dx, dy, dt, cva, mime, in0, precision, sort_every, push_order, &
                          ybound, track_pid)
        integer, intent(in) :: nx, ny, precision, sort_every
        real(c_double), intent(in) :: dx, dy, dt, cva, mime, in0
        logical, intent(in), optional :: push_order, track_pid
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
        if (present(track_pid)) then
            if (track_pid) p%flags = ior(p%flags, PBPIC_FLAG_TRACK_PID)
        end if
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
    ! vz passed by the reference's pusher (:1028-1030).  The reference declares
    ! the offsets and uz as default `real` (:1230-1231): currentdeposit_r4
    ! takes those actuals, currentdeposit_r8 those of a -fdefault-real-8 build.
    subroutine currentdeposit_r8(i1, i2, j1, j2, dx1, dx2, dy1, dy2, q0, uz)
        integer, intent(in) :: i1, i2, j1, j2, q0
        real(c_double), intent(in) :: dx1, dx2, dy1, dy2, uz
        call pbpic_check(c_deposit(ctx, 1_c_int64_t, [int(i1, c_int32_t)], [int(i2, c_int32_t)], &
            [int(j1, c_int32_t)], [int(j2, c_int32_t)], [dx1], [dx2], [dy1], [dy2], &
            [int(q0, c_int32_t)], [uz]), 'currentdeposit')
    end subroutine

    subroutine currentdeposit_r4(i1, i2, j1, j2, dx1, dx2, dy1, dy2, q0, uz)
        integer, intent(in) :: i1, i2, j1, j2, q0
        real(c_float), intent(in) :: dx1, dx2, dy1, dy2, uz
        call currentdeposit_r8(i1, i2, j1, j2, real(dx1, c_double), real(dx2, c_double), &
            real(dy1, c_double), real(dy2, c_double), q0, real(uz, c_double))
    end subroutine

    ! The reference's particle array ptl (type particle, pic2d.f90:47-50, :391)
    ! to the device: the component arrays of ptl(1:n) (e.g. ptl(1:n)%dx) are
    ! split by charge into the module's ion and electron species, each in its
    ! original order.  Cells are the reference's 1-based grid.
    subroutine put_ptl_r8(grid, q, pid, dx, dy, ux, uy, uz)
        integer, intent(in) :: grid(:), q(:), pid(:)
        real(c_double), intent(in) :: dx(:), dy(:), ux(:), uy(:), uz(:)
        logical, allocatable :: m(:)
        integer :: s
        integer(c_int32_t) :: sp
        do s = 1, 2
            m = merge(q > 0, q < 0, s == 1)
            sp = merge(sp_ion, sp_ele, s == 1)
            call pbpic_check(c_set_particles(ctx, sp, int(count(m), c_int64_t), &
                int(pack(grid, m), c_int32_t), pack(dx, m), pack(dy, m), pack(ux, m), &
                pack(uy, m), pack(uz, m), int(pack(pid, m), c_int32_t)), 'pbpic_put_ptl')
        end do
    end subroutine

    subroutine put_ptl_r4(grid, q, pid, dx, dy, ux, uy, uz)
        integer, intent(in) :: grid(:), q(:), pid(:)
        real(c_float), intent(in) :: dx(:), dy(:), ux(:), uy(:), uz(:)
        logical, allocatable :: m(:)
        integer :: s
        integer(c_int32_t) :: sp
        do s = 1, 2
            m = merge(q > 0, q < 0, s == 1)
            sp = merge(sp_ion, sp_ele, s == 1)
            call pbpic_check(c_set_particles_f32(ctx, sp, int(count(m), c_int64_t), &
                int(pack(grid, m), c_int32_t), pack(dx, m), pack(dy, m), pack(ux, m), &
                pack(uy, m), pack(uz, m), int(pack(pid, m), c_int32_t)), 'pbpic_put_ptl')
        end do
    end subroutine

    ! The device particles back into the reference's array (for particlesinfo,
    ! pic2d.f90:1359-1621): ions first, then electrons, in device order, with
    ! q = +1 / -1; n returns the count (the arrays must hold all particles).
    subroutine take_ptl_r8(n, grid, q, pid, dx, dy, ux, uy, uz)
        integer, intent(out) :: n
        integer, intent(inout) :: grid(:), q(:), pid(:)
        real(c_double), intent(inout) :: dx(:), dy(:), ux(:), uy(:), uz(:)
        integer(c_int64_t) :: ni, ne, m
        integer(c_int32_t), allocatable :: g(:), p(:)
        real(c_double), allocatable :: a(:, :)
        call pbpic_check(c_num_particles(ctx, sp_ion, ni), 'pbpic_take_ptl')
        call pbpic_check(c_num_particles(ctx, sp_ele, ne), 'pbpic_take_ptl')
        if (ni + ne > size(grid)) call pbpic_check(-3_c_int32_t, 'pbpic_take_ptl: array too small')
        allocate(g(max(ni, ne)), p(max(ni, ne)), a(max(ni, ne), 5))
        call pbpic_check(c_get_particles(ctx, sp_ion, int(size(g), c_int64_t), m, g, a(:, 1), &
            a(:, 2), a(:, 3), a(:, 4), a(:, 5), p), 'pbpic_take_ptl')
        grid(1:ni) = g(1:ni); pid(1:ni) = p(1:ni); q(1:ni) = 1
        dx(1:ni) = a(1:ni, 1); dy(1:ni) = a(1:ni, 2)
        ux(1:ni) = a(1:ni, 3); uy(1:ni) = a(1:ni, 4); uz(1:ni) = a(1:ni, 5)
        call pbpic_check(c_get_particles(ctx, sp_ele, int(size(g), c_int64_t), m, g, a(:, 1), &
            a(:, 2), a(:, 3), a(:, 4), a(:, 5), p), 'pbpic_take_ptl')
        grid(ni+1:ni+ne) = g(1:ne); pid(ni+1:ni+ne) = p(1:ne); q(ni+1:ni+ne) = -1
        dx(ni+1:ni+ne) = a(1:ne, 1); dy(ni+1:ni+ne) = a(1:ne, 2)
        ux(ni+1:ni+ne) = a(1:ne, 3); uy(ni+1:ni+ne) = a(1:ne, 4); uz(ni+1:ni+ne) = a(1:ne, 5)
        n = int(ni + ne)
    end subroutine

    subroutine take_ptl_r4(n, grid, q, pid, dx, dy, ux, uy, uz)
        integer, intent(out) :: n
        integer, intent(inout) :: grid(:), q(:), pid(:)
        real(c_float), intent(inout) :: dx(:), dy(:), ux(:), uy(:), uz(:)
        integer(c_int64_t) :: ni, ne, m
        integer(c_int32_t), allocatable :: g(:), p(:)
        real(c_float), allocatable :: a(:, :)
        call pbpic_check(c_num_particles(ctx, sp_ion, ni), 'pbpic_take_ptl')
        call pbpic_check(c_num_particles(ctx, sp_ele, ne), 'pbpic_take_ptl')
        if (ni + ne > size(grid)) call pbpic_check(-3_c_int32_t, 'pbpic_take_ptl: array too small')
        allocate(g(max(ni, ne)), p(max(ni, ne)), a(max(ni, ne), 5))
        call pbpic_check(c_get_particles_f32(ctx, sp_ion, int(size(g), c_int64_t), m, g, a(:, 1), &
            a(:, 2), a(:, 3), a(:, 4), a(:, 5), p), 'pbpic_take_ptl')
        grid(1:ni) = g(1:ni); pid(1:ni) = p(1:ni); q(1:ni) = 1
        dx(1:ni) = a(1:ni, 1); dy(1:ni) = a(1:ni, 2)
        ux(1:ni) = a(1:ni, 3); uy(1:ni) = a(1:ni, 4); uz(1:ni) = a(1:ni, 5)
        call pbpic_check(c_get_particles_f32(ctx, sp_ele, int(size(g), c_int64_t), m, g, a(:, 1), &
            a(:, 2), a(:, 3), a(:, 4), a(:, 5), p), 'pbpic_take_ptl')
        grid(ni+1:ni+ne) = g(1:ne); pid(ni+1:ni+ne) = p(1:ne); q(ni+1:ni+ne) = -1
        dx(ni+1:ni+ne) = a(1:ne, 1); dy(ni+1:ni+ne) = a(1:ne, 2)
        ux(ni+1:ni+ne) = a(1:ne, 3); uy(ni+1:ni+ne) = a(1:ne, 4); uz(ni+1:ni+ne) = a(1:ne, 5)
        n = int(ni + ne)
    end subroutine

    ! One plane of the reference's field or current arrays to the device: the
    ! owned region, e.g. emf(1:mx,1:my)%ex (pic2d.f90:593) or
    ! denjr(1:mx,1:my)%jx (:221, normalized as after :1066), x fastest.
    subroutine put_field_r8(which, plane)
        integer(c_int32_t), intent(in) :: which
        real(c_double), intent(in) :: plane(:, :)
        call pbpic_check(c_set_field(ctx, which, reshape(plane, [size(plane)])), 'pbpic_put_field')
    end subroutine

    subroutine put_field_r4(which, plane)
        integer(c_int32_t), intent(in) :: which
        real(c_float), intent(in) :: plane(:, :)
        call pbpic_check(c_set_field(ctx, which, real(reshape(plane, [size(plane)]), c_double)), &
                         'pbpic_put_field')
    end subroutine

    ! One device plane back into the owned region of a reference array.
    subroutine take_field_r8(which, plane)
        integer(c_int32_t), intent(in) :: which
        real(c_double), intent(inout) :: plane(:, :)
        real(c_double), allocatable :: t(:)
        allocate(t(size(plane)))
        call pbpic_check(c_get_field(ctx, which, t), 'pbpic_take_field')
        plane = reshape(t, shape(plane))
    end subroutine

    subroutine take_field_r4(which, plane)
        integer(c_int32_t), intent(in) :: which
        real(c_float), intent(inout) :: plane(:, :)
        real(c_double), allocatable :: t(:)
        allocate(t(size(plane)))
        call pbpic_check(c_get_field(ctx, which, t), 'pbpic_take_field')
        plane = real(reshape(t, shape(plane)), c_float)
    end subroutine

    ! Waits for the device and reports its error flags (CFL, non-finite,
    ! cell range); the reference has no counterpart (it is synchronous).
    subroutine pbpic_sync
        call pbpic_check(c_sync(ctx), 'pbpic_sync')
    end subroutine

    ! countingsort (pic2d.f90:1303-1329) of one species, returning the
    ! per-cell counts (ptlalloc after :1310-1313, nx*ny entries indexed by
    ! grid) and the permutation (perm(k) = 0-based index before the sort of
    ! the particle now at position k); use push_order = .false. in pbpic_init
    ! for the reference's exact order.
    subroutine countingsort_species(sp, counts, perm)
        integer(c_int32_t), intent(in) :: sp
        integer(c_int32_t), intent(inout), target :: counts(:), perm(:)
        call pbpic_check(c_countingsort_species(ctx, sp, c_loc(counts), c_loc(perm)), &
                         'countingsort_species')
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
