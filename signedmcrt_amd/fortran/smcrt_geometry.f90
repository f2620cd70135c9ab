!> Geometry registry of the Fortran host: `setup_geometry(cfg)` dispatches on the TOML
!! `geom_name` like the reference's `setup_simulation` (src/setup.f90:33-52).  Kept from the
!! reference: scat_test (src/setupGeometry.f90:221-247), scat_test2 (:249-274), aptran
!! (:147-175).  Added: slab, refracting_sphere, skin_vessels — the same scenes as
!! signedmcrt_amd/scenes.py (skin_vessels with its seeded layer and vessel draws written out).
module smcrt_geometry
    use, intrinsic :: iso_c_binding
    use smcrt
    use smcrt_sdf
    implicit none

    private
    public :: setup_geometry

contains

    !> the SDF array for cfg%geom_name, or an unallocated array ("no such routine")
    subroutine setup_geometry(cfg, array)
        type(smcrt_config), intent(in) :: cfg
        type(sdf), allocatable, intent(out) :: array(:)
        character(len=:), allocatable :: name

        name = c_string(cfg%geom_name)
        select case (name)
        case ("scat_test")
            array = [sphere(1._dp, mono(cfg%tau, 0._dp, 0._dp, 1._dp), 1), &
                     box([2._dp, 2._dp, 2._dp], mono(0._dp, 0._dp, 0._dp, 1._dp), 2)]
        case ("scat_test2")
            array = [box([200._dp, 200._dp, 200._dp], mono(cfg%tau, 1e-17_dp, cfg%hgg, 1._dp), 2)]
        case ("aptran")
            array = [sphere(0.5_dp, mono(0._dp, 1e-17_dp, 0._dp, 1.33_dp), 1, &
                            transform=invert(translate([0._dp, 0._dp, 0._dp]))), &
                     box([2._dp, 2._dp, 2._dp], mono(0._dp, 1e-17_dp, 0._dp, 1._dp), 2), &
                     box([2.01_dp, 2.01_dp, 2.01_dp], mono(0._dp, 10000000._dp, 0._dp, 1._dp), 3)]
        case ("slab")
            array = [box([1._dp, 1._dp, 0.2_dp], mono(100._dp, 0.1_dp, 0.9_dp, 1.37_dp), 1), &
                     box([2._dp, 2._dp, 2._dp], mono(0._dp, 0._dp, 0._dp, 1._dp), 2)]
        case ("refracting_sphere")
            array = [sphere(0.5_dp, mono(0._dp, 0.01_dp, 0._dp, 1.5_dp), 1), &
                     box([2._dp, 2._dp, 2._dp], mono(50._dp, 0.05_dp, 0.8_dp, 1.33_dp), 2), &
                     box([4._dp, 4._dp, 4._dp], mono(0._dp, 0._dp, 0._dp, 1._dp), 3)]
        case ("skin_vessels")
            call skin_vessels(array)
        case default
            return
        end select
    end subroutine setup_geometry

    !> five stacked tissue layers + eight blood vessels (cylinders along x) merged by
    !! SmoothUnion (k = 0.01), in air; vessels first, then the layers top to bottom, then air
    subroutine skin_vessels(array)
        type(sdf), allocatable, intent(out) :: array(:)
        real(dp), parameter :: thick(5) = [0.023579242248433302_dp, 0.45428302718717112_dp, &
            0.4418828418028698_dp, 0.31619966495624524_dp, 0.39745025743799078_dp]
        real(dp), parameter :: mua(5) = [1.9120879597547533_dp, 2.0863640715779939_dp, &
            0.27052369750314281_dp, 3.9482730941525719_dp, 3.723057164173925_dp]
        real(dp), parameter :: mus(5) = [168.36072086882908_dp, 43.419236817104419_dp, &
            95.556391939302046_dp, 35.063954398708162_dp, 178.6531771964363_dp]
        real(dp), parameter :: g(5) = [0.80283766060048745_dp, 0.83080604313662665_dp, &
            0.83740322445205639_dp, 0.93481670254287319_dp, 0.81616959398168698_dp]
        real(dp), parameter :: n(5) = [1.4298942490153093_dp, 1.4271762892606397_dp, &
            1.4292623265572941_dp, 1.4456093983025182_dp, 1.4067568697162327_dp]
        real(dp), parameter :: vr(8) = [0.019265398770678815_dp, 0.048621995064390966_dp, &
            0.034586892698687109_dp, 0.012990889808798839_dp, 0.015402647671141323_dp, &
            0.04057027136774688_dp, 0.025742250950993494_dp, 0.047178995209683973_dp]
        real(dp), parameter :: vdepth(8) = [0.46847639461365892_dp, 0.9473739604668574_dp, &
            0.38889706939273966_dp, 0.83638536833264499_dp, 0.28203963145301619_dp, &
            0.8589946876254273_dp, 0.49861982879682853_dp, 0.21914649869539587_dp]
        real(dp), parameter :: vy(8) = [0.20913682236984199_dp, 0.2428482111650152_dp, &
            -0.29450386028166753_dp, -0.67421282899869128_dp, 0.78608884211674446_dp, &
            -0.23875005154913698_dp, -0.73296025142503396_dp, -0.18306599949059055_dp]
        type(sdf) :: cyls(8), layers(5)
        type(mono) :: blood
        real(dp) :: top, z, zc, total
        integer :: i

        total = sum(thick)
        top = total / 2._dp
        z = top
        do i = 1, 5
            zc = z - thick(i) / 2._dp
            layers(i) = box([2._dp, 2._dp, thick(i)], mono(mus(i), mua(i), g(i), n(i)), i + 1, &
                            transform=invert(translate([0._dp, 0._dp, zc])))
            z = z - thick(i)
        end do
        blood = mono(150._dp, 20._dp, 0.95_dp, 1.40_dp)
        do i = 1, 8
            zc = top - vdepth(i)
            cyls(i) = cylinder([-1._dp, vy(i), zc], [1._dp, vy(i), zc], vr(i), blood, 1)
        end do
        array = [model(cyls, SmoothUnion, 0.01_dp), layers, &
                 box([4._dp, 4._dp, 4._dp + total], mono(0._dp, 0._dp, 0._dp, 1._dp), 7)]
    end subroutine skin_vessels

end module smcrt_geometry
