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
    use fmm, only: field_direct, field_fmm
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
end module fmmref_bindc
