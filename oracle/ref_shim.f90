! ref_shim.f90 — bind(C) wrappers around the reference's own Fortran routines.
!
! TEST INFRASTRUCTURE ONLY.  oracle/Makefile compiles this file together with the reference's
! src/constants.f90, src/vector_class.f90, src/random_mod.f90, src/surfaces.f90 and src/grid.f90,
! read in place from the reference checkout, into oracle/_ref/libsmcrt_ref.so.  Those five files
! build with flang alone; the rest of the reference's hot path needs un-vendored fpm dependencies
! (utils, tomlf, stdlib, ...) and is not built.  tests/test_oracle.py uses this library to pin
! the oracle's Fresnel/reflect/refract, the 4x4 row-vector transform and the grid faces against
! the reference's compiled code.
module smcrt_ref_shim
    use iso_c_binding
    use constants,    only : wp
    use vector_class, only : vector
    use surfaces,     only : reflect_refract
    use random,       only : init_rng
    use gridMod,      only : cart_grid, init_grid
    implicit none
contains

    ! reflect_refract (src/surfaces.f90:14-40) on n rows; returns the Fresnel coefficient, the
    ! branch taken (1 = reflected) and the new direction.  The branch is drawn from the
    ! reference's own generator, seeded by init_rng(seed).
    subroutine ref_reflect_refract(n, inc, nrm, n1, n2, seed, ri, rflag, outdir) bind(C, name="ref_reflect_refract")
        integer(c_int64_t), value :: n
        real(c_double), intent(in) :: inc(3, n), nrm(3, n), n1(n), n2(n)
        integer(c_int), value :: seed
        real(c_double), intent(out) :: ri(n), outdir(3, n)
        integer(c_int), intent(out) :: rflag(n)
        type(vector) :: I, NN
        logical :: fl
        real(wp) :: r
        integer(c_int64_t) :: j
        call init_rng(spread(int(seed), 1, 8), fwd=.true.)
        do j = 1, n
            I = vector(inc(1, j), inc(2, j), inc(3, j))
            NN = vector(nrm(1, j), nrm(2, j), nrm(3, j))
            call reflect_refract(I, NN, n1(j), n2(j), fl, r)
            ri(j) = r
            rflag(j) = merge(1, 0, fl)
            outdir(:, j) = [I%x, I%y, I%z]
        end do
    end subroutine ref_reflect_refract

    ! the row-vector transform [p,1].T of src/vector_class.f90:292-304 on n points
    subroutine ref_transform(n, t, pts, outp) bind(C, name="ref_transform")
        integer(c_int64_t), value :: n
        real(c_double), intent(in) :: t(4, 4), pts(3, n)
        real(c_double), intent(out) :: outp(3, n)
        type(vector) :: p, q
        integer(c_int64_t) :: j
        do j = 1, n
            p = vector(pts(1, j), pts(2, j), pts(3, j))
            q = p .dot. t
            outp(:, j) = [q%x, q%y, q%z]
        end do
    end subroutine ref_transform

    ! init_grid (src/grid.f90:53-93): voxel wall positions (xface nx+1, yface ny+1, zface nz+2)
    subroutine ref_grid_faces(nx, ny, nz, xmax, ymax, zmax, xf, yf, zf) bind(C, name="ref_grid_faces")
        integer(c_int), value :: nx, ny, nz
        real(c_double), value :: xmax, ymax, zmax
        real(c_double), intent(out) :: xf(nx + 1), yf(ny + 1), zf(nz + 2)
        type(cart_grid) :: g
        g = init_grid(nx, ny, nz, xmax, ymax, zmax)
        xf = g%xface
        yf = g%yface
        zf = g%zface
    end subroutine ref_grid_faces
end module smcrt_ref_shim
