!> Fortran host driver: the reference's `mcpolar` program running `pathlength_scatter`
!! (app/main.f90, src/kernelsMod.f90:165-333) with the packet loop in libsmcrt.so.
!!
!!     smcrt_run [input.toml] [--weight]   (default scat_test.toml, as app/main.f90)
!!
!! --weight runs the `weight_scatter` estimator (src/kernelsMod.f90:12-163) instead of
!! `pathlength_scatter`.
!! Reads the reference's TOML input (smcrt_config_load), builds the geometry named by
!! [geometry] geom_name (smcrt_geometry), runs `nphotons` packets with seed `iseed`, prints
!! Photons/s and the mean number of scatters, writes the normalised fluence to
!! data/jmean/<fluence> as NRRD and each detector to data/detectors/detector_<i>.dat
!! (src/kernelsMod.f90:548-633, src/writer.f90).  Exit status 1 on any error.
program smcrt_main
    use, intrinsic :: iso_c_binding
    use, intrinsic :: iso_fortran_env, only : int64, real64, real32
    use smcrt
    use smcrt_sdf
    use smcrt_geometry
    implicit none

    character(len=512) :: file
    type(smcrt_config) :: cfg
    type(sdf), allocatable :: array(:)
    type(smcrt_grid) :: grid
    type(smcrt_source) :: src
    type(smcrt_run_opts) :: opts
    type(smcrt_tallies) :: tal
    type(c_ptr) :: scene
    real(c_double), allocatable, target :: jmean(:, :, :)
    real(c_float), allocatable :: fluence(:, :, :), image(:, :, :)
    integer(c_int64_t), allocatable, target :: counts(:)
    integer(c_int64_t) :: nb, off
    integer(int64) :: c0, c1, rate
    integer :: i, status
    character(len=:), allocatable :: name, meta
    character(len=32) :: buf
    character(len=512) :: arg
    logical :: weight

    file = "scat_test.toml"
    weight = .false.
    do i = 1, command_argument_count()
        call get_command_argument(i, arg)
        if (trim(arg) == "--weight") then
            weight = .true.
        else
            file = arg
        end if
    end do

    status = smcrt_config_load(trim(file)//c_null_char, cfg)
    call check(status)

    call setup_geometry(cfg, array)
    if (.not. allocated(array)) then
        print '(a)', "no such routine: "//c_string(cfg%geom_name)
        stop 1
    end if

    grid%nx = cfg%nxg; grid%ny = cfg%nyg; grid%nz = cfg%nzg
    grid%xmax = cfg%xmax; grid%ymax = cfg%ymax; grid%zmax = cfg%zmax

    name = c_string(cfg%source_name)
    select case (name)
    case ("pencil"); src%kind = SMCRT_SRC_PENCIL
    case ("uniform"); src%kind = SMCRT_SRC_UNIFORM
    case ("circular"); src%kind = SMCRT_SRC_CIRCULAR
    case ("point"); src%kind = SMCRT_SRC_POINT
    case default
        print '(a)', "No such source! ("//name//")"
        stop 1
    end select
    src%pos = cfg%position
    src%dir = [0._c_double, 0._c_double, -1._c_double]
    if (cfg%has_direction /= 0) src%dir = cfg%direction
    src%p1 = cfg%point1; src%p2 = cfg%point2; src%p3 = cfg%point3
    src%radius = cfg%radius

    status = scene_create(array, grid, src, cfg%detectors(1:cfg%n_detectors), scene)
    call check(status)

    print '(a,a,a,i0,a)', "Running ", c_string(cfg%geom_name), " with ", cfg%nphotons, " photons"
    allocate(jmean(cfg%nxg, cfg%nyg, cfg%nzg), fluence(cfg%nxg, cfg%nyg, cfg%nzg))
    jmean = 0._c_double
    nb = 0
    do i = 1, cfg%n_detectors
        nb = nb + smcrt_detector_bins(scene, int(i - 1, c_int32_t))
    end do
    allocate(counts(max(1_c_int64_t, nb)))
    counts = 0
    tal%jmean = c_loc(jmean)
    if (cfg%n_detectors > 0) tal%detector_counts = c_loc(counts)
    opts%n_packets = cfg%nphotons
    opts%seed = cfg%iseed
    if (weight) opts%flags = SMCRT_FLAG_WEIGHT

    call system_clock(c0, rate)
    status = smcrt_run(scene, opts, tal)
    call check(status)
    call system_clock(c1)
    print *, "Photons/s: ", real(cfg%nphotons, real64) / max(real(c1 - c0, real64) / real(rate, real64), 1e-9_real64)
    print *, "Average # of scatters per photon:", real(tal%stats%scatters, real64) / real(cfg%nphotons, real64)

    ! outputs (src/kernelsMod.f90:601-625)
    call execute_command_line("mkdir -p data/jmean data/detectors")
    status = smcrt_normalise_fluence(grid, jmean, cfg%nphotons, fluence)
    call check(status)
    write (buf, '(i0)') cfg%nphotons
    meta = "grid_data: fluence map"//new_line("a")//"nphotons: "//trim(buf)//new_line("a")// &
           "source: "//name//new_line("a")//"experiment: "//c_string(cfg%geom_name)
    status = smcrt_write_nrrd("data/jmean/"//c_string(cfg%fluence)//c_null_char, fluence, &
                              cfg%nxg, cfg%nyg, cfg%nzg, meta//c_null_char)
    call check(status)
    off = 0
    do i = 1, cfg%n_detectors
        write (buf, '(i0)') i
        status = smcrt_write_detector("data/detectors/detector_"//trim(buf)//".dat"//c_null_char, &
                                      cfg%detectors(i), counts(off + 1:))
        call check(status)
        off = off + smcrt_detector_bins(scene, int(i - 1, c_int32_t))
    end do
    ! geometry render (src/kernelsMod.f90:532-536, render_sub in src/sdfs/sdf_base.f90:302-346)
    if (cfg%render_geom /= 0) then
        allocate(image(cfg%render_size(1), cfg%render_size(2), cfg%render_size(3)))
        status = smcrt_render(scene, cfg%render_size, image)
        call check(status)
        status = smcrt_write_nrrd("data/"//c_string(cfg%render)//c_null_char, image, cfg%render_size(1), &
                                  cfg%render_size(2), cfg%render_size(3), c_null_char)
        call check(status)
    end if
    status = smcrt_scene_destroy(scene)

contains

    subroutine check(rc)
        integer(c_int), intent(in) :: rc
        if (rc /= SMCRT_OK) then
            print '(a)', "smcrt error: "//smcrt_error_message()
            stop 1
        end if
    end subroutine check

end program smcrt_main
