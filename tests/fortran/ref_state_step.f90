!******************************************************************************
! ref_state_step: the reference's own host state through module pbpic_gpu
! (TEST PROGRAM).  It declares pic2d.f90's types as the reference does --
! type particle (:47-50), type fields (:52-57), type density (:65-67) -- in
! default `real`, so that a plain build passes 4-byte reals (the reference's
! CMake build) and a -fdefault-real-8 build 8-byte reals, and it moves
! ptl(:) (:391), emf(0:mx+1,0:my+1) (:593) and denjr(0:mx+2,0:my+2) (:221)
! to and from the device with the module's generic entry points.  It then
! runs one time-loop body (zerocurrent, pusher, fieldsolver; :173-178) and a
! batch of currentdeposit calls with default-real actuals (:1029-1030).
!
! Usage: ref_state_step <input file> <output file>
!   input  (stream, native endian): int32 mx, my, n, ndep; float64 dx, dy,
!          dt, cva, mime, in0; int32 grid(n), q(n), pid(n); float64 dx(n),
!          dy(n), ux(n), uy(n), uz(n); float64 ex, ey, ez, bx, by, bz planes
!          (mx*my each, x fastest); int32 i1, i2, j1, j2, q of ndep entries;
!          float64 dx1, dx2, dy1, dy2, vz of ndep entries
!   output int32 n, grid(n), q(n), pid(n); float64 dx .. uz (n each); float64
!          planes ex, ey, ez, bx, by, bz, jx, jy, jz after the step, then jx,
!          jy, jz after the currentdeposit calls
! Precision of the device form: FP64 for 8-byte default reals, the split
! form for 4-byte ones.
!******************************************************************************
program ref_state_step
    use iso_c_binding
    use pbpic_gpu
    implicit none
    type particle
        real dx, dy, ux, uy, uz
        integer grid, q, pid
    end type
    type fields
        real ex, dex_dy, ey, dey_dx
        real ez, dez_dx, dez_dy, dive
        real bx, dbx_dy, by, dby_dx
        real bz, dbz_dx, dbz_dy, pad1
    end type
    type density
        real jx, jy, jz, rho
    end type
    type(particle), allocatable :: ptl(:)
    type(fields), allocatable :: emf(:, :)
    type(density), allocatable :: denjr(:, :)
    integer(c_int32_t) :: hdr(4)
    real(c_double) :: prm(6)
    integer(c_int32_t), allocatable :: ib(:), dep_i(:, :)
    real(c_double), allocatable :: rb(:, :), planes(:, :), dep_r(:, :)
    integer :: mx, my, n, ndep, k, nout, prec, u
    character(len=512) :: fin, fout

    call get_command_argument(1, fin)
    call get_command_argument(2, fout)
    open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old')
    read(u) hdr, prm
    mx = hdr(1); my = hdr(2); n = hdr(3); ndep = hdr(4)
    allocate(ptl(n), emf(0:mx+1, 0:my+1), denjr(0:mx+2, 0:my+2))
    allocate(ib(n), rb(n, 5), planes(mx*my, 6), dep_i(ndep, 5), dep_r(ndep, 5))
    read(u) ib
    ptl%grid = ib
    read(u) ib
    ptl%q = ib
    read(u) ib
    ptl%pid = ib
    read(u) rb, planes, dep_i, dep_r
    close(u)
    ptl%dx = real(rb(:, 1)); ptl%dy = real(rb(:, 2))
    ptl%ux = real(rb(:, 3)); ptl%uy = real(rb(:, 4)); ptl%uz = real(rb(:, 5))
    emf(1:mx, 1:my)%ex = reshape(real(planes(:, 1)), [mx, my])
    emf(1:mx, 1:my)%ey = reshape(real(planes(:, 2)), [mx, my])
    emf(1:mx, 1:my)%ez = reshape(real(planes(:, 3)), [mx, my])
    emf(1:mx, 1:my)%bx = reshape(real(planes(:, 4)), [mx, my])
    emf(1:mx, 1:my)%by = reshape(real(planes(:, 5)), [mx, my])
    emf(1:mx, 1:my)%bz = reshape(real(planes(:, 6)), [mx, my])

    prec = PBPIC_PREC_SPLIT32
    if (kind(1.0) == 8) prec = PBPIC_PREC_FP64
    call pbpic_init(mx, my, prm(1), prm(2), prm(3), prm(4), prm(5), prm(6), prec, 0, &
                    push_order=.false., track_pid=.true.)
    call pbpic_check(c_add_species(ctx, 1_c_int32_t, int(count(ptl%q > 0), c_int64_t), sp_ion), &
                     'species')
    call pbpic_check(c_add_species(ctx, -1_c_int32_t, int(count(ptl%q < 0), c_int64_t), sp_ele), &
                     'species')
    call pbpic_put_ptl(ptl%grid, ptl%q, ptl%pid, ptl%dx, ptl%dy, ptl%ux, ptl%uy, ptl%uz)
    call pbpic_put_field(PBPIC_EX, emf(1:mx, 1:my)%ex)
    call pbpic_put_field(PBPIC_EY, emf(1:mx, 1:my)%ey)
    call pbpic_put_field(PBPIC_EZ, emf(1:mx, 1:my)%ez)
    call pbpic_put_field(PBPIC_BX, emf(1:mx, 1:my)%bx)
    call pbpic_put_field(PBPIC_BY, emf(1:mx, 1:my)%by)
    call pbpic_put_field(PBPIC_BZ, emf(1:mx, 1:my)%bz)

    call zerocurrent                                    ! pic2d.f90:173-176
    call pusher                                         ! :177
    call fieldsolver                                    ! :178
    call pbpic_sync

    call pbpic_take_ptl(nout, ptl%grid, ptl%q, ptl%pid, ptl%dx, ptl%dy, ptl%ux, ptl%uy, ptl%uz)
    call pbpic_take_field(PBPIC_EX, emf(1:mx, 1:my)%ex)
    call pbpic_take_field(PBPIC_EY, emf(1:mx, 1:my)%ey)
    call pbpic_take_field(PBPIC_EZ, emf(1:mx, 1:my)%ez)
    call pbpic_take_field(PBPIC_BX, emf(1:mx, 1:my)%bx)
    call pbpic_take_field(PBPIC_BY, emf(1:mx, 1:my)%by)
    call pbpic_take_field(PBPIC_BZ, emf(1:mx, 1:my)%bz)
    call pbpic_take_field(PBPIC_JX, denjr(1:mx, 1:my)%jx)
    call pbpic_take_field(PBPIC_JY, denjr(1:mx, 1:my)%jy)
    call pbpic_take_field(PBPIC_JZ, denjr(1:mx, 1:my)%jz)

    open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace')
    write(u) int(nout, c_int32_t)
    write(u) int(ptl(1:nout)%grid, c_int32_t), int(ptl(1:nout)%q, c_int32_t), &
             int(ptl(1:nout)%pid, c_int32_t)
    write(u) real(ptl(1:nout)%dx, c_double), real(ptl(1:nout)%dy, c_double), &
             real(ptl(1:nout)%ux, c_double), real(ptl(1:nout)%uy, c_double), &
             real(ptl(1:nout)%uz, c_double)
    write(u) real(emf(1:mx, 1:my)%ex, c_double), real(emf(1:mx, 1:my)%ey, c_double), &
             real(emf(1:mx, 1:my)%ez, c_double), real(emf(1:mx, 1:my)%bx, c_double), &
             real(emf(1:mx, 1:my)%by, c_double), real(emf(1:mx, 1:my)%bz, c_double)
    write(u) real(denjr(1:mx, 1:my)%jx, c_double), real(denjr(1:mx, 1:my)%jy, c_double), &
             real(denjr(1:mx, 1:my)%jz, c_double)

    ! currentdeposit (pic2d.f90:1226) called as the reference's pusher calls
    ! it (:1029-1030): integer cells and charge, default-real offsets and vz
    call zerocurrent
    do k = 1, ndep
        call currentdeposit(dep_i(k, 1), dep_i(k, 2), dep_i(k, 3), dep_i(k, 4), &
                            real(dep_r(k, 1)), real(dep_r(k, 2)), real(dep_r(k, 3)), &
                            real(dep_r(k, 4)), dep_i(k, 5), real(dep_r(k, 5)))
    end do
    call pbpic_take_field(PBPIC_JX, denjr(1:mx, 1:my)%jx)
    call pbpic_take_field(PBPIC_JY, denjr(1:mx, 1:my)%jy)
    call pbpic_take_field(PBPIC_JZ, denjr(1:mx, 1:my)%jz)
    write(u) real(denjr(1:mx, 1:my)%jx, c_double), real(denjr(1:mx, 1:my)%jy, c_double), &
             real(denjr(1:mx, 1:my)%jz, c_double)
    close(u)
    call releasememory                                  ! :189
end program ref_state_step
