! ORACLE -- test infrastructure only.
!
! bind(C) entry points over the UNMODIFIED reference module `fmm`
! (peter-reinholdt/fmm src/fmm.F90:1338-1400), compiled from the reference's
! own sources by oracle/Makefile into oracle/_ref/.  Used to write golden
! fixtures and, in bench.py, to time the reference on the host as the CPU
! baseline.  Nothing in the product links this.
!
! Arrays are column-major, as the reference expects: coordinates n x 3,
! multipoles n x dim, exclusions n x nexcl (1-based, 0-padded), field
! n x ncomp(field_order) (the reference accumulates into it; pass zeros).
! damp: 0 '', 1 'THOLE', 2 'AMOEBA', 3 'ERF'.
module fmmref_bindc
    use iso_c_binding
    use precision, only: ip, rp, lp
    use fmm, only: field_direct, field_fmm, get_multipoles, get_boxes, interpolate_electron_fields
    implicit none
contains
    function damp_name(damp) result(dt)
        integer(c_int), intent(in) :: damp
        character(len=6) :: dt
        select case (damp)
        case (1); dt = 'THOLE '
        case (2); dt = 'AMOEBA'
        case (3); dt = 'ERF   '
        case default; dt = '      '
        end select
    end function damp_name

    subroutine ref_field_direct(n, dim, nexcl, coords, mults, excl, field_order, damp, df, field) &
            bind(C, name="ref_field_direct")
        integer(c_int), value :: n, dim, nexcl, field_order, damp
        real(c_double), intent(in) :: coords(n, 3), mults(n, dim), df(*)
        integer(c_int), intent(in) :: excl(n, nexcl)
        real(c_double), intent(inout) :: field(n, (field_order + 1) * (field_order + 2) * (field_order + 3) / 6)
        integer(ip) :: comm
        comm = 0
        if (damp == 0) then
            call field_direct(comm, coords, mults, excl, field_order, field, damp_name(damp), df(1:0))
        else
            call field_direct(comm, coords, mults, excl, field_order, field, damp_name(damp), df(1:n))
        end if
    end subroutine ref_field_direct

    subroutine ref_field_fmm(n, dim, nexcl, coords, mults, excl, theta, ncrit, p, field_order, damp, df, field) &
            bind(C, name="ref_field_fmm")
        integer(c_int), value :: n, dim, nexcl, ncrit, p, field_order, damp
        real(c_double), value :: theta
        real(c_double), intent(in) :: coords(n, 3), mults(n, dim), df(*)
        integer(c_int), intent(in) :: excl(n, nexcl)
        real(c_double), intent(inout) :: field(n, (field_order + 1) * (field_order + 2) * (field_order + 3) / 6)
        integer(ip) :: comm
        logical(lp) :: periodic
        real(rp) :: box(3, 3)
        comm = 0
        periodic = .false.
        box = 0.0_rp
        if (damp == 0) then
            call field_fmm(comm, coords, mults, excl, theta, ncrit, p, field_order, field, damp_name(damp), &
                           df(1:0), periodic, box)
        else
            call field_fmm(comm, coords, mults, excl, theta, ncrit, p, field_order, field, damp_name(damp), &
                           df(1:n), periodic, box)
        end if
    end subroutine ref_field_fmm

    ! periodic field_fmm: box is the 3 x 3 box_matrix (column-major)
    subroutine ref_field_fmm_periodic(n, dim, nexcl, coords, mults, excl, theta, ncrit, p, field_order, damp, df, &
                                      box, field) bind(C, name="ref_field_fmm_periodic")
        integer(c_int), value :: n, dim, nexcl, ncrit, p, field_order, damp
        real(c_double), value :: theta
        real(c_double), intent(in) :: coords(n, 3), mults(n, dim), df(*), box(3, 3)
        integer(c_int), intent(in) :: excl(n, nexcl)
        real(c_double), intent(inout) :: field(n, (field_order + 1) * (field_order + 2) * (field_order + 3) / 6)
        integer(ip) :: comm
        logical(lp) :: periodic
        comm = 0
        periodic = .true.
        if (damp == 0) then
            call field_fmm(comm, coords, mults, excl, theta, ncrit, p, field_order, field, damp_name(damp), &
                           df(1:0), periodic, box)
        else
            call field_fmm(comm, coords, mults, excl, theta, ncrit, p, field_order, field, damp_name(damp), &
                           df(1:n), periodic, box)
        end if
    end subroutine ref_field_fmm_periodic

    ! get_boxes: copies at most cap_near / cap_far entries; sizes = (num_near, num_far)
    subroutine ref_get_boxes(n, coords, nt, targets, theta, ncrit, p, cap_near, cap_far, sizes, idx_near, &
                             box_coords, box_idx) bind(C, name="ref_get_boxes")
        integer(c_int), value :: n, nt, ncrit, p, cap_near, cap_far
        real(c_double), value :: theta
        real(c_double), intent(in) :: coords(n, 3), targets(nt, 3)
        integer(c_int), intent(out) :: sizes(2), idx_near(cap_near), box_idx(cap_far)
        real(c_double), intent(out) :: box_coords(cap_far, 3)
        real(rp), allocatable :: bc(:, :)
        integer(ip), allocatable :: inf(:), bi(:)
        call get_boxes(coords, targets, theta, ncrit, p, inf, bc, bi)
        sizes(1) = size(inf)
        sizes(2) = size(bi)
        idx_near(1:min(cap_near, size(inf))) = inf(1:min(cap_near, size(inf)))
        box_idx(1:min(cap_far, size(bi))) = bi(1:min(cap_far, size(bi)))
        box_coords(1:min(cap_far, size(bi)), :) = bc(1:min(cap_far, size(bi)), :)
    end subroutine ref_get_boxes

    subroutine ref_get_multipoles(n, dim, ndens, coords, mults, nt, targets, theta, ncrit, p, cap_near, cap_far, &
                                  sizes, idx_near, box_coords, box_mults) bind(C, name="ref_get_multipoles")
        integer(c_int), value :: n, dim, ndens, nt, ncrit, p, cap_near, cap_far
        real(c_double), value :: theta
        real(c_double), intent(in) :: coords(n, 3), mults(n, dim, ndens), targets(nt, 3)
        integer(c_int), intent(out) :: sizes(2), idx_near(cap_near)
        real(c_double), intent(out) :: box_coords(cap_far, 3)
        real(c_double), intent(out) :: box_mults(cap_far, (p + 1) * (p + 2) * (p + 3) / 6, ndens)
        real(rp), allocatable :: bc(:, :), bm(:, :, :)
        integer(ip), allocatable :: inf(:)
        integer(ip) :: comm, nf
        comm = 0
        call get_multipoles(comm, coords, mults, targets, theta, ncrit, p, inf, bc, bm)
        sizes(1) = size(inf)
        sizes(2) = size(bc, 1)
        nf = min(cap_far, size(bc, 1))
        idx_near(1:min(cap_near, size(inf))) = inf(1:min(cap_near, size(inf)))
        box_coords(1:nf, :) = bc(1:nf, :)
        box_mults(1:nf, :, :) = bm(1:nf, :, :)
    end subroutine ref_get_multipoles

    subroutine ref_interpolate_electron_fields(n, coords, nbox, nm, ndens, le, box_idx, theta, ncrit, p, &
                                               field_order, field) bind(C, name="ref_interpolate_electron_fields")
        integer(c_int), value :: n, nbox, nm, ndens, ncrit, p, field_order
        real(c_double), value :: theta
        real(c_double), intent(in) :: coords(n, 3), le(nbox, nm, ndens)
        integer(c_int), intent(in) :: box_idx(nbox)
        real(c_double), intent(inout) :: field(n, (field_order + 1) * (field_order + 2) * (field_order + 3) / 6, ndens)
        integer(ip) :: comm
        comm = 0
        call interpolate_electron_fields(comm, coords, le, box_idx, theta, ncrit, p, field_order, field)
    end subroutine ref_interpolate_electron_fields
end module fmmref_bindc
