! Example caller of the Fortran drop-in, as PElib-style host code would use
! it: `use octree, only: field_direct, field_fmm` (README.md:26), plain
! Fortran arrays, no HIP in sight.  Reads a raw binary case file written by
! tests/test_gpu_fortran.py and writes the field back.
!
!   drive.x <in.bin> <out.bin>
! in.bin: int32 mode(0 direct, 1 fmm), n, dim, nexcl, field_order, ncrit,
!         expansion_order, damp_len; real64 theta; damp_type chars;
!         coordinates(n,3), multipoles(n,dim), int32 exclusions(n,nexcl),
!         real64 damping_factors(n)
program drive
    use precision, only: ip, rp, lp
    use octree, only: field_direct, field_fmm
    implicit none
    character(len=512) :: fin, fout
    integer(ip) :: hdr(8), mode, n, dim, nexcl, fo, ncrit, p, dl
    real(rp) :: theta, box(3, 3)
    character(len=:), allocatable :: damp
    real(rp), allocatable :: x(:, :), m(:, :), df(:), field(:, :)
    integer(ip), allocatable :: e(:, :)
    logical(lp) :: periodic
    call get_command_argument(1, fin)
    call get_command_argument(2, fout)
    open (10, file=trim(fin), access='stream', form='unformatted', status='old')
    read (10) hdr
    mode = hdr(1); n = hdr(2); dim = hdr(3); nexcl = hdr(4); fo = hdr(5); ncrit = hdr(6); p = hdr(7); dl = hdr(8)
    read (10) theta
    allocate (character(len=dl) :: damp)
    if (dl > 0) read (10) damp
    allocate (x(n, 3), m(n, dim), e(n, nexcl), df(n))
    allocate (field(n, (fo + 1) * (fo + 2) * (fo + 3) / 6))
    read (10) x, m, e, df
    close (10)
    field = 0.0_rp
    periodic = .false.
    box = 0.0_rp
    if (mode == 0) then
        call field_direct(0, x, m, e, fo, field, damp, df)
    else
        call field_fmm(0, x, m, e, theta, ncrit, p, fo, field, damp, df, periodic, box)
    end if
    open (11, file=trim(fout), access='stream', form='unformatted', status='replace')
    write (11) field
    close (11)
end program drive
