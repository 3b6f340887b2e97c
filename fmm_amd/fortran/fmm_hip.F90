! Fortran drop-in for the field engine of peter-reinholdt/fmm.
!
! Module `fmm` keeps the names, argument order, kinds and intents of the
! reference's public field calls (src/fmm.F90:1338-1346 field_direct,
! src/fmm.F90:1354-1368 field_fmm) and forwards them through iso_c_binding
! to the C ABI of libfmm_hip.so (include/fmm_hip.h), which runs on the GPU.
! Module `octree` re-exports them under the name the README uses
! (README.md:26).  Module `precision` provides the reference's kind names
! (src/precision.F90:8-26, default build).
!
! Semantics kept from the reference: field is accumulated into, exclusions
! are 1-based and must contain the site itself, damping applies to the near
! field only, unknown damp types stop with "Unknown damp type: <damp_type>".
! comm is accepted and ignored (one rank).  Periodic boxes stop with an
! error: this engine does not implement them.
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
    public field_direct, field_fmm

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
end module fmm

module octree
    use fmm, only: field_direct, field_fmm
    implicit none
    public
end module octree
