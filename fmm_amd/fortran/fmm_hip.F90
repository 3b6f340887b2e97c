! Fortran drop-in for the field engine of peter-reinholdt/fmm.
!
! Module `fmm` keeps the names, argument order, kinds and intents of the
! reference's public procedures (src/fmm.F90:8-12: field_direct, field_fmm,
! get_multipoles, get_boxes, interpolate_electron_fields) and forwards them
! through iso_c_binding to the C ABI of libfmm_hip.so (include/fmm_hip.h),
! which runs on the GPU.
! Module `octree` re-exports them under the name the README uses
! (README.md:26).  Module `precision` provides the reference's kind names
! (src/precision.F90:8-26, default build).
!
! Semantics kept from the reference: field is accumulated into, exclusions
! are 1-based and must contain the site itself, damping applies to the near
! field only, unknown damp types stop with "Unknown damp type: <damp_type>".
! comm is accepted and ignored (one rank).  Periodic boxes follow the
! reference (27 image cells, lattice sum, dipole correction).
module precision
    use, intrinsic :: iso_fortran_env, only: int32, real64
    implicit none
    integer(int32), parameter :: ip = int32
    integer(ip), parameter :: rp = real64
    integer(ip), parameter :: lp = int32
end module precision

module fmm
    use, intrinsic :: iso_c_binding
    use precision
    implicit none
    private
    public field_direct, field_fmm, get_multipoles, get_boxes, interpolate_electron_fields

    interface
        integer(c_int32_t) function c_parse_damp(s, len, code) bind(C, name="fmmh_parse_damp_type")
            import :: c_char, c_int32_t, c_int64_t
            character(kind=c_char), intent(in) :: s(*)
            integer(c_int64_t), value :: len
            integer(c_int32_t), intent(out) :: code
        end function c_parse_damp

        integer(c_int32_t) function c_field_direct(ctx, comm, coords, ldc, mult, ldm, n, dim, excl, lde, nexcl, &
                                                   field_order, field, ldf, damp, df, ndf) &
                bind(C, name="fmmh_field_direct")
            import :: c_ptr, c_int32_t, c_int64_t, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: comm, n, dim, nexcl, field_order, damp
            integer(c_int64_t), value :: ldc, ldm, lde, ldf, ndf
            real(c_double), intent(in) :: coords(*), mult(*), df(*)
            integer(c_int32_t), intent(in) :: excl(*)
            real(c_double), intent(inout) :: field(*)
        end function c_field_direct

        integer(c_int32_t) function c_get_boxes(ctx, coords, ldc, n, tg, ldt, nt, theta, ncrit, p, sizes) &
                bind(C, name="fmmh_get_boxes")
            import :: c_ptr, c_int32_t, c_int64_t, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: n, nt, ncrit, p
            integer(c_int64_t), value :: ldc, ldt
            real(c_double), value :: theta
            real(c_double), intent(in) :: coords(*), tg(*)
            integer(c_int64_t), intent(out) :: sizes(2)
        end function c_get_boxes

        integer(c_int32_t) function c_get_boxes_export(ctx, near, bc, bi) bind(C, name="fmmh_get_boxes_export")
            import :: c_ptr, c_int32_t, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), intent(out) :: near(*), bi(*)
            real(c_double), intent(out) :: bc(*)
        end function c_get_boxes_export

        integer(c_int32_t) function c_get_multipoles(ctx, comm, coords, ldc, n, mult, ldm, dim, ndens, tg, ldt, nt, &
                                                     theta, ncrit, p, sizes) bind(C, name="fmmh_get_multipoles")
            import :: c_ptr, c_int32_t, c_int64_t, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: comm, n, dim, ndens, nt, ncrit, p
            integer(c_int64_t), value :: ldc, ldm, ldt
            real(c_double), value :: theta
            real(c_double), intent(in) :: coords(*), mult(*), tg(*)
            integer(c_int64_t), intent(out) :: sizes(2)
        end function c_get_multipoles

        integer(c_int32_t) function c_get_multipoles_export(ctx, near, bc, bm) &
                bind(C, name="fmmh_get_multipoles_export")
            import :: c_ptr, c_int32_t, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), intent(out) :: near(*)
            real(c_double), intent(out) :: bc(*), bm(*)
        end function c_get_multipoles_export

        integer(c_int32_t) function c_interpolate(ctx, comm, coords, ldc, n, le, ldle, nbox, ndens, bi, theta, &
                                                  ncrit, p, field_order, field, ldf) &
                bind(C, name="fmmh_interpolate_electron_fields")
            import :: c_ptr, c_int32_t, c_int64_t, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: comm, n, nbox, ndens, ncrit, p, field_order
            integer(c_int64_t), value :: ldc, ldle, ldf
            real(c_double), value :: theta
            real(c_double), intent(in) :: coords(*), le(*)
            integer(c_int32_t), intent(in) :: bi(*)
            real(c_double), intent(inout) :: field(*)
        end function c_interpolate

        integer(c_int32_t) function c_field_fmm(ctx, comm, coords, ldc, mult, ldm, n, dim, excl, lde, nexcl, &
                                                theta, ncrit, expansion_order, field_order, field, ldf, damp, df, &
                                                ndf, is_periodic, box) bind(C, name="fmmh_field_fmm")
            import :: c_ptr, c_int32_t, c_int64_t, c_double
            type(c_ptr), value :: ctx
            integer(c_int32_t), value :: comm, n, dim, nexcl, ncrit, expansion_order, field_order, damp, is_periodic
            integer(c_int64_t), value :: ldc, ldm, lde, ldf, ndf
            real(c_double), value :: theta
            real(c_double), intent(in) :: coords(*), mult(*), df(*), box(*)
            integer(c_int32_t), intent(in) :: excl(*)
            real(c_double), intent(inout) :: field(*)
        end function c_field_fmm

        type(c_ptr) function c_last_error() bind(C, name="fmmh_last_error")
            import :: c_ptr
        end function c_last_error
    end interface

contains

    subroutine stop_on_error(rc)
        integer(c_int32_t), intent(in) :: rc
        character(kind=c_char), pointer :: msg(:)
        character(len=:), allocatable :: text
        integer :: i, l
        if (rc == 0) return
        call c_f_pointer(c_last_error(), msg, [4096])
        l = 0
        do i = 1, 4096
            if (msg(i) == c_null_char) exit
            l = i
        end do
        allocate (character(len=l) :: text)
        do i = 1, l
            text(i:i) = msg(i)
        end do
        error stop text
    end subroutine stop_on_error

    function damp_code(damp_type) result(code)
        character(len=*), intent(in) :: damp_type
        integer(c_int32_t) :: code
        character(kind=c_char) :: blank(1)
        if (len(damp_type) == 0) then
            blank(1) = ' '
            call stop_on_error(c_parse_damp(blank, 0_c_int64_t, code))
        else
            call stop_on_error(c_parse_damp(damp_type, int(len(damp_type), c_int64_t), code))
        end if
    end function damp_code

    ! src/fmm.F90:1338-1352
    subroutine field_direct(comm, coordinates, multipoles, exclusions, field_order, field, damp_type, damping_factors)
        integer(ip), intent(in) :: comm
        real(rp), intent(in), target :: coordinates(:, :)
        real(rp), intent(in), target :: multipoles(:, :)
        integer(ip), intent(in) :: exclusions(:, :)
        integer(ip), intent(in) :: field_order
        character(len=*), intent(in) :: damp_type
        real(rp), intent(in) :: damping_factors(:)
        real(rp), intent(out) :: field(size(coordinates, 1), (field_order + 1) * (field_order + 2) * (field_order + 3) / 6)
        integer(c_int32_t) :: code, n
        real(c_double) :: dummy(1)
        n = int(size(coordinates, 1), c_int32_t)
        code = damp_code(damp_type)
        dummy = 0.0_c_double
        if (size(damping_factors) > 0) then
            call stop_on_error(c_field_direct(c_null_ptr, int(comm, c_int32_t), coordinates, int(n, c_int64_t), &
                multipoles, int(n, c_int64_t), n, int(size(multipoles, 2), c_int32_t), exclusions, &
                int(n, c_int64_t), int(size(exclusions, 2), c_int32_t), int(field_order, c_int32_t), field, &
                int(n, c_int64_t), code, damping_factors, int(size(damping_factors), c_int64_t)))
        else
            call stop_on_error(c_field_direct(c_null_ptr, int(comm, c_int32_t), coordinates, int(n, c_int64_t), &
                multipoles, int(n, c_int64_t), n, int(size(multipoles, 2), c_int32_t), exclusions, &
                int(n, c_int64_t), int(size(exclusions, 2), c_int32_t), int(field_order, c_int32_t), field, &
                int(n, c_int64_t), code, dummy, 0_c_int64_t))
        end if
    end subroutine field_direct

    ! src/fmm.F90:1354-1400
    subroutine field_fmm(comm, coordinates, multipoles, exclusions, theta, ncrit, expansion_order, field_order, &
                         field, damp_type, damping_factors, is_periodic, box_matrix)
        integer(ip), intent(in) :: comm
        real(rp), intent(in), target :: coordinates(:, :)
        real(rp), intent(in), target :: multipoles(:, :)
        integer(ip), intent(in) :: exclusions(:, :)
        real(rp), intent(in) :: theta
        integer(ip), intent(in) :: ncrit
        integer(ip), intent(in) :: expansion_order
        integer(ip), intent(in) :: field_order
        character(len=*), intent(in) :: damp_type
        real(rp), intent(in) :: damping_factors(:)
        logical(lp), intent(in) :: is_periodic
        real(rp), intent(in), dimension(3, 3) :: box_matrix
        real(rp), intent(out) :: field(size(coordinates, 1), (field_order + 1) * (field_order + 2) * (field_order + 3) / 6)
        integer(c_int32_t) :: code, n, per
        real(c_double) :: dummy(1)
        n = int(size(coordinates, 1), c_int32_t)
        code = damp_code(damp_type)
        per = 0
        if (is_periodic) per = 1
        dummy = 0.0_c_double
        if (size(damping_factors) > 0) then
            call stop_on_error(c_field_fmm(c_null_ptr, int(comm, c_int32_t), coordinates, int(n, c_int64_t), &
                multipoles, int(n, c_int64_t), n, int(size(multipoles, 2), c_int32_t), exclusions, &
                int(n, c_int64_t), int(size(exclusions, 2), c_int32_t), real(theta, c_double), &
                int(ncrit, c_int32_t), int(expansion_order, c_int32_t), int(field_order, c_int32_t), field, &
                int(n, c_int64_t), code, damping_factors, int(size(damping_factors), c_int64_t), per, box_matrix))
        else
            call stop_on_error(c_field_fmm(c_null_ptr, int(comm, c_int32_t), coordinates, int(n, c_int64_t), &
                multipoles, int(n, c_int64_t), n, int(size(multipoles, 2), c_int32_t), exclusions, &
                int(n, c_int64_t), int(size(exclusions, 2), c_int32_t), real(theta, c_double), &
                int(ncrit, c_int32_t), int(expansion_order, c_int32_t), int(field_order, c_int32_t), field, &
                int(n, c_int64_t), code, dummy, 0_c_int64_t, per, box_matrix))
        end if
    end subroutine field_fmm

    ! src/fmm.F90:1171-1229
    subroutine get_multipoles(comm, coordinates, multipoles, target_coordinates, theta, ncrit, expansion_order, &
                              idx_near_field, box_coordinates, box_multipoles)
        integer(ip), intent(in) :: comm
        real(rp), intent(in), target :: coordinates(:, :)
        real(rp), intent(in), target :: multipoles(:, :, :)
        real(rp), intent(in) :: target_coordinates(:, :)
        real(rp), intent(in) :: theta
        integer(ip), intent(in) :: ncrit
        integer(ip), intent(in) :: expansion_order
        real(rp), intent(out), allocatable :: box_coordinates(:, :)
        real(rp), intent(out), allocatable :: box_multipoles(:, :, :)
        integer(ip), intent(out), allocatable :: idx_near_field(:)
        integer(c_int64_t) :: sizes(2)
        integer(c_int32_t) :: n, nt
        n = int(size(coordinates, 1), c_int32_t)
        nt = int(size(target_coordinates, 1), c_int32_t)
        call stop_on_error(c_get_multipoles(c_null_ptr, int(comm, c_int32_t), coordinates, int(n, c_int64_t), n, &
            multipoles, int(n, c_int64_t), int(size(multipoles, 2), c_int32_t), &
            int(size(multipoles, 3), c_int32_t), target_coordinates, int(max(nt, 1), c_int64_t), nt, &
            real(theta, c_double), int(ncrit, c_int32_t), int(expansion_order, c_int32_t), sizes))
        allocate (idx_near_field(sizes(1)))
        allocate (box_coordinates(sizes(2), 3))
        allocate (box_multipoles(sizes(2), (expansion_order + 1) * (expansion_order + 2) * (expansion_order + 3) / 6, &
                                 size(multipoles, 3)))
        call stop_on_error(c_get_multipoles_export(c_null_ptr, idx_near_field, box_coordinates, box_multipoles))
    end subroutine get_multipoles

    ! src/fmm.F90:1231-1291
    subroutine get_boxes(coordinates, target_coordinates, theta, ncrit, expansion_order, idx_near_field, &
                         box_coordinates, box_indices)
        real(rp), intent(in), target :: coordinates(:, :)
        real(rp), intent(in) :: target_coordinates(:, :)
        real(rp), intent(in) :: theta
        integer(ip), intent(in) :: ncrit
        integer(ip), intent(in) :: expansion_order
        real(rp), intent(out), allocatable :: box_coordinates(:, :)
        integer(ip), intent(out), allocatable :: idx_near_field(:), box_indices(:)
        integer(c_int64_t) :: sizes(2)
        integer(c_int32_t) :: n, nt
        n = int(size(coordinates, 1), c_int32_t)
        nt = int(size(target_coordinates, 1), c_int32_t)
        call stop_on_error(c_get_boxes(c_null_ptr, coordinates, int(n, c_int64_t), n, target_coordinates, &
            int(max(nt, 1), c_int64_t), nt, real(theta, c_double), int(ncrit, c_int32_t), &
            int(expansion_order, c_int32_t), sizes))
        allocate (idx_near_field(sizes(1)))
        allocate (box_coordinates(sizes(2), 3))
        allocate (box_indices(sizes(2)))
        call stop_on_error(c_get_boxes_export(c_null_ptr, idx_near_field, box_coordinates, box_indices))
    end subroutine get_boxes

    ! src/fmm.F90:1293-1336
    subroutine interpolate_electron_fields(comm, coordinates, local_expansion, box_indices, theta, ncrit, &
                                           expansion_order, field_order, field)
        integer(ip), intent(in) :: comm
        real(rp), intent(in), target :: coordinates(:, :)
        real(rp), intent(in) :: local_expansion(:, :, :)
        integer(ip), intent(in) :: box_indices(:)
        real(rp), intent(in) :: theta
        integer(ip), intent(in) :: ncrit
        integer(ip), intent(in) :: expansion_order
        integer(ip), intent(in) :: field_order
        real(rp), intent(out) :: field(size(coordinates, 1), (field_order + 1) * (field_order + 2) * &
                                       (field_order + 3) / 6, size(local_expansion, 3))
        integer(c_int32_t) :: n, nb
        n = int(size(coordinates, 1), c_int32_t)
        nb = int(size(local_expansion, 1), c_int32_t)
        call stop_on_error(c_interpolate(c_null_ptr, int(comm, c_int32_t), coordinates, int(n, c_int64_t), n, &
            local_expansion, int(max(nb, 1), c_int64_t), nb, int(size(local_expansion, 3), c_int32_t), box_indices, &
            real(theta, c_double), int(ncrit, c_int32_t), int(expansion_order, c_int32_t), &
            int(field_order, c_int32_t), field, int(n, c_int64_t)))
    end subroutine interpolate_electron_fields
end module fmm

module octree
    use fmm, only: field_direct, field_fmm, get_multipoles, get_boxes, interpolate_electron_fields
    implicit none
    public
end module octree
