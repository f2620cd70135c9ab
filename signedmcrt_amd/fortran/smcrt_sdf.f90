!> The SDF model API for Fortran hosts, with the reference's constructor names and argument
!! order: primitives (src/sdfs/sdfs.f90:158-492), `model` and the fold operators
!! (src/sdfs/sdf_base.f90:100-143, src/sdfs/sdfModifiers.f90:412-475), the modifiers
!! (src/sdfs/sdfModifiers.f90:125-271), the 4x4 transform helpers (src/sdfs/sdfHelpers.f90) and
!! `mono` (src/opticalProps/opticalProperties.f90:107-125).
!!
!! An `sdf` here is a description: its postfix node program for libsmcrt (include/smcrt.h),
!! its optical properties and its layer.  `scene_create` flattens an array of them, in array
!! order, into a `smcrt_scene_desc` and creates the library's scene.  The fold operators are
!! integer codes (the reference passes procedures): `model([a, b], SmoothUnion, 0.01_dp)` reads
!! the same.  Matrices use the reference's row-vector convention, q = [x y z 1] . T.
module smcrt_sdf
    use, intrinsic :: iso_c_binding
    use smcrt
    implicit none

    private
    public :: dp, mono, sdf, sphere, box, torus, cylinder, triprism, segment, capsule, cone, egg, &
              plane, model, extrude, revolution, onion, elongate, twist, bend, identity, &
              translate, rotate_x, rotate_y, rotate_z, invert, scene_create
    public :: opunion, SmoothUnion, subtraction, intersection

    integer, parameter :: dp = c_double
    integer, parameter :: opunion = SMCRT_UNION, SmoothUnion = SMCRT_SMOOTH_UNION, &
                          subtraction = SMCRT_SUBTRACTION, intersection = SMCRT_INTERSECTION

    !> single-wavelength optical properties
    type :: mono
        real(dp) :: mus = 0._dp, mua = 0._dp, hgg = 0._dp, n = 1._dp
    end type mono

    type :: sdf
        type(smcrt_node), allocatable :: nodes(:)
        type(mono) :: optProps
        integer :: layer = 1
    end type sdf

contains

    ! ---- 4x4 helpers -------------------------------------------------------------------------
    pure function identity() result(t)
        real(dp) :: t(4, 4)
        integer :: i
        t = 0._dp
        do i = 1, 4
            t(i, i) = 1._dp
        end do
    end function identity

    !> offset in row 4 (src/sdfs/sdfHelpers.f90:170-182)
    pure function translate(o) result(t)
        real(dp), intent(in) :: o(3)
        real(dp) :: t(4, 4)
        t = identity()
        t(4, 1:3) = o
    end function translate

    pure function rotate_x(angle) result(t)
        real(dp), intent(in) :: angle
        real(dp) :: t(4, 4), c, s
        c = cos(angle * acos(-1._dp) / 180._dp)
        s = sin(angle * acos(-1._dp) / 180._dp)
        t = identity()
        t(2, 2) = c; t(3, 2) = -s; t(2, 3) = s; t(3, 3) = c
    end function rotate_x

    pure function rotate_y(angle) result(t)
        real(dp), intent(in) :: angle
        real(dp) :: t(4, 4), c, s
        c = cos(angle * acos(-1._dp) / 180._dp)
        s = sin(angle * acos(-1._dp) / 180._dp)
        t = identity()
        t(1, 1) = c; t(3, 1) = s; t(1, 3) = -s; t(3, 3) = c
    end function rotate_y

    pure function rotate_z(angle) result(t)
        real(dp), intent(in) :: angle
        real(dp) :: t(4, 4), c, s
        c = cos(angle * acos(-1._dp) / 180._dp)
        s = sin(angle * acos(-1._dp) / 180._dp)
        t = identity()
        t(1, 1) = c; t(2, 1) = -s; t(1, 2) = s; t(2, 2) = c
    end function rotate_z

    !> inverse of a 4x4 matrix (Gauss-Jordan with partial pivoting)
    function invert(m) result(r)
        real(dp), intent(in) :: m(4, 4)
        real(dp) :: r(4, 4), a(4, 8), tmp(8), f
        integer :: i, j, p
        a(:, 1:4) = m
        a(:, 5:8) = identity()
        do i = 1, 4
            p = i - 1 + maxloc(abs(a(i:4, i)), dim=1)
            if (p /= i) then
                tmp = a(i, :); a(i, :) = a(p, :); a(p, :) = tmp
            end if
            a(i, :) = a(i, :) / a(i, i)
            do j = 1, 4
                if (j /= i) then
                    f = a(j, i)
                    a(j, :) = a(j, :) - f * a(i, :)
                end if
            end do
        end do
        r = a(:, 5:8)
    end function invert

    ! ---- primitives --------------------------------------------------------------------------
    function prim(kind, params, optProp, layer, transform) result(out)
        integer(c_int32_t), intent(in) :: kind
        real(dp), intent(in) :: params(:)
        type(mono), intent(in) :: optProp
        integer, intent(in) :: layer
        real(dp), optional, intent(in) :: transform(4, 4)
        type(sdf) :: out
        real(dp) :: t(4, 4)
        integer :: r
        allocate(out%nodes(1))
        out%nodes(1)%op = SMCRT_OP_PRIM
        out%nodes(1)%kind = kind
        out%nodes(1)%p(1:size(params)) = params
        t = identity()
        if (present(transform)) t = transform
        do r = 1, 4
            out%nodes(1)%t(3 * (r - 1) + 1:3 * r) = t(r, 1:3)
        end do
        out%optProps = optProp
        out%layer = layer
    end function prim

    function sphere(radius, optProp, layer, transform) result(out)
        real(dp), intent(in) :: radius
        type(mono), intent(in) :: optProp
        integer, intent(in) :: layer
        real(dp), optional, intent(in) :: transform(4, 4)
        type(sdf) :: out
        out = prim(SMCRT_SPHERE, [radius], optProp, layer, transform)
    end function sphere

    !> box(lengths, ...): full edge lengths, halved as the reference's constructor does
    function box(lengths, optProp, layer, transform) result(out)
        real(dp), intent(in) :: lengths(3)
        type(mono), intent(in) :: optProp
        integer, intent(in) :: layer
        real(dp), optional, intent(in) :: transform(4, 4)
        type(sdf) :: out
        out = prim(SMCRT_BOX, lengths / 2._dp, optProp, layer, transform)
    end function box

    function torus(oradius, iradius, optProp, layer, transform) result(out)
        real(dp), intent(in) :: oradius, iradius
        type(mono), intent(in) :: optProp
        integer, intent(in) :: layer
        real(dp), optional, intent(in) :: transform(4, 4)
        type(sdf) :: out
        out = prim(SMCRT_TORUS, [oradius, iradius], optProp, layer, transform)
    end function torus

    function cylinder(a, b, radius, optProp, layer, transform) result(out)
        real(dp), intent(in) :: a(3), b(3), radius
        type(mono), intent(in) :: optProp
        integer, intent(in) :: layer
        real(dp), optional, intent(in) :: transform(4, 4)
        type(sdf) :: out
        out = prim(SMCRT_CYLINDER, [a, b, radius], optProp, layer, transform)
    end function cylinder

    function triprism(h1, h2, optProp, layer, transform) result(out)
        real(dp), intent(in) :: h1, h2
        type(mono), intent(in) :: optProp
        integer, intent(in) :: layer
        real(dp), optional, intent(in) :: transform(4, 4)
        type(sdf) :: out
        out = prim(SMCRT_TRIPRISM, [h1, h2], optProp, layer, transform)
    end function triprism

    function segment(a, b, optProp, layer, transform) result(out)
        real(dp), intent(in) :: a(3), b(3)
        type(mono), intent(in) :: optProp
        integer, intent(in) :: layer
        real(dp), optional, intent(in) :: transform(4, 4)
        type(sdf) :: out
        out = prim(SMCRT_SEGMENT, [a, b], optProp, layer, transform)
    end function segment

    function capsule(a, b, r, optProp, layer, transform) result(out)
        real(dp), intent(in) :: a(3), b(3), r
        type(mono), intent(in) :: optProp
        integer, intent(in) :: layer
        real(dp), optional, intent(in) :: transform(4, 4)
        type(sdf) :: out
        out = prim(SMCRT_CAPSULE, [a, b, r], optProp, layer, transform)
    end function capsule

    function cone(a, b, ra, rb, optProp, layer, transform) result(out)
        real(dp), intent(in) :: a(3), b(3), ra, rb
        type(mono), intent(in) :: optProp
        integer, intent(in) :: layer
        real(dp), optional, intent(in) :: transform(4, 4)
        type(sdf) :: out
        out = prim(SMCRT_CONE, [a, b, ra, rb], optProp, layer, transform)
    end function cone

    function egg(r1, r2, h, optProp, layer, transform) result(out)
        real(dp), intent(in) :: r1, r2, h
        type(mono), intent(in) :: optProp
        integer, intent(in) :: layer
        real(dp), optional, intent(in) :: transform(4, 4)
        type(sdf) :: out
        out = prim(SMCRT_EGG, [r1, r2, h], optProp, layer, transform)
    end function egg

    function plane(a, optProp, layer, transform) result(out)
        real(dp), intent(in) :: a(3)
        type(mono), intent(in) :: optProp
        integer, intent(in) :: layer
        real(dp), optional, intent(in) :: transform(4, 4)
        type(sdf) :: out
        out = prim(SMCRT_PLANE, a, optProp, layer, transform)
    end function plane

    ! ---- model: fold op over the children left to right; properties/layer of the first -------
    function model(array, func, kopt) result(out)
        type(sdf), intent(in) :: array(:)
        integer, intent(in) :: func
        real(dp), optional, intent(in) :: kopt
        type(sdf) :: out
        type(smcrt_node) :: op
        integer :: i
        op%op = SMCRT_OP_CSG
        op%kind = func
        if (present(kopt)) op%k = kopt
        out%nodes = array(1)%nodes
        do i = 2, size(array)
            out%nodes = [out%nodes, array(i)%nodes, op]
        end do
        out%optProps = array(1)%optProps
        out%layer = array(1)%layer
    end function model

    ! ---- modifiers ---------------------------------------------------------------------------
    function modifier(kind, prim_sdf, params) result(out)
        integer(c_int32_t), intent(in) :: kind
        type(sdf), intent(in) :: prim_sdf
        real(dp), intent(in) :: params(:)
        type(sdf) :: out
        type(smcrt_node) :: b, e
        b%op = SMCRT_OP_MOD_BEGIN
        b%kind = kind
        b%p(1:size(params)) = params
        e = b
        e%op = SMCRT_OP_MOD_END
        out%nodes = [b, prim_sdf%nodes, e]
        out%optProps = prim_sdf%optProps
        out%layer = prim_sdf%layer
    end function modifier

    function extrude(prim_sdf, h) result(out)
        type(sdf), intent(in) :: prim_sdf
        real(dp), intent(in) :: h
        type(sdf) :: out
        out = modifier(SMCRT_EXTRUDE, prim_sdf, [h])
    end function extrude

    function revolution(prim_sdf, o) result(out)
        type(sdf), intent(in) :: prim_sdf
        real(dp), intent(in) :: o
        type(sdf) :: out
        out = modifier(SMCRT_REVOLUTION, prim_sdf, [o])
    end function revolution

    function onion(prim_sdf, thickness) result(out)
        type(sdf), intent(in) :: prim_sdf
        real(dp), intent(in) :: thickness
        type(sdf) :: out
        out = modifier(SMCRT_ONION, prim_sdf, [thickness])
    end function onion

    function elongate(prim_sdf, size3) result(out)
        type(sdf), intent(in) :: prim_sdf
        real(dp), intent(in) :: size3(3)
        type(sdf) :: out
        out = modifier(SMCRT_ELONGATE, prim_sdf, size3)
    end function elongate

    function twist(prim_sdf, k) result(out)
        type(sdf), intent(in) :: prim_sdf
        real(dp), intent(in) :: k
        type(sdf) :: out
        out = modifier(SMCRT_TWIST, prim_sdf, [k])
    end function twist

    function bend(prim_sdf, k) result(out)
        type(sdf), intent(in) :: prim_sdf
        real(dp), intent(in) :: k
        type(sdf) :: out
        out = modifier(SMCRT_BEND, prim_sdf, [k])
    end function bend

    ! ---- scene -------------------------------------------------------------------------------
    !> Flatten `array` (one region per element, in order), create the library scene.  Materials
    !! with identical properties share one material index.  Returns the library status.
    function scene_create(array, grid, source, dets, scene) result(status)
        type(sdf), intent(in) :: array(:)
        type(smcrt_grid), intent(in) :: grid
        type(smcrt_source), intent(in) :: source
        type(smcrt_detector), intent(in) :: dets(:)
        type(c_ptr), intent(out) :: scene
        integer(c_int) :: status
        type(smcrt_node), allocatable, target :: nodes(:)
        type(smcrt_region), allocatable, target :: regions(:)
        type(smcrt_material), allocatable, target :: mats(:)
        type(smcrt_detector), allocatable, target :: dcopy(:)
        type(smcrt_scene_desc) :: desc
        integer :: i, j, m, nn

        nn = 0
        do i = 1, size(array)
            nn = nn + size(array(i)%nodes)
        end do
        allocate(nodes(nn), regions(size(array)), mats(0))
        nn = 0
        do i = 1, size(array)
            m = -1
            do j = 1, size(mats)
                if (mats(j)%mus == array(i)%optProps%mus .and. mats(j)%mua == array(i)%optProps%mua .and. &
                    mats(j)%hgg == array(i)%optProps%hgg .and. mats(j)%n == array(i)%optProps%n) m = j - 1
            end do
            if (m < 0) then
                mats = [mats, smcrt_material(array(i)%optProps%mus, array(i)%optProps%mua, &
                                             array(i)%optProps%hgg, array(i)%optProps%n)]
                m = size(mats) - 1
            end if
            regions(i) = smcrt_region(nn, size(array(i)%nodes), m, array(i)%layer)
            nodes(nn + 1:nn + size(array(i)%nodes)) = array(i)%nodes
            nn = nn + size(array(i)%nodes)
        end do
        allocate(dcopy(max(1, size(dets))))
        if (size(dets) > 0) dcopy(1:size(dets)) = dets
        desc%abi_version = SMCRT_ABI
        desc%n_nodes = nn
        desc%n_regions = size(array)
        desc%n_materials = size(mats)
        desc%nodes = c_loc(nodes)
        desc%regions = c_loc(regions)
        desc%materials = c_loc(mats)
        desc%grid = grid
        desc%source = source
        desc%n_detectors = size(dets)
        desc%detectors = c_loc(dcopy)
        status = smcrt_scene_create(desc, scene)
    end function scene_create

end module smcrt_sdf
