!> iso_c_binding interface to libsmcrt.so (include/smcrt.h): the MI355X packet engine that
!! replaces the OpenMP region of `pathlength_scatter` (src/kernelsMod.f90:239-329).
!!
!! Every derived type here is `bind(C)` and mirrors a struct of include/smcrt.h field for field
!! (checked against smcrt_sizeof by tests/test_fortran.py); every interface is one `extern "C"`
!! entry point.  Strings passed in must be NUL-terminated (`trim(s)//c_null_char`).
module smcrt
    use, intrinsic :: iso_c_binding
    implicit none

    integer(c_int), parameter :: SMCRT_ABI = 3  !< SMCRT_ABI_VERSION of include/smcrt.h

    ! status codes
    integer(c_int), parameter :: SMCRT_OK = 0, SMCRT_EINVAL = -1, SMCRT_ENOMEM = -2, &
                                 SMCRT_EDEVICE = -3, SMCRT_ERANGE = -4, SMCRT_EIO = -5
    ! node ops
    integer(c_int32_t), parameter :: SMCRT_OP_PRIM = 1, SMCRT_OP_CSG = 2, SMCRT_OP_MOD_BEGIN = 3, &
                                     SMCRT_OP_MOD_END = 4
    ! primitives (src/sdfs/sdfs.f90:494-735)
    integer(c_int32_t), parameter :: SMCRT_SPHERE = 1, SMCRT_BOX = 2, SMCRT_TORUS = 3, &
                                     SMCRT_CYLINDER = 4, SMCRT_TRIPRISM = 5, SMCRT_SEGMENT = 6, &
                                     SMCRT_CAPSULE = 7, SMCRT_CONE = 8, SMCRT_EGG = 9, SMCRT_PLANE = 10, &
                                     SMCRT_POLYGON = 11
    ! CSG operators (src/sdfs/sdfModifiers.f90:412-475)
    integer(c_int32_t), parameter :: SMCRT_UNION = 1, SMCRT_SMOOTH_UNION = 2, &
                                     SMCRT_SUBTRACTION = 3, SMCRT_INTERSECTION = 4
    ! modifiers (src/sdfs/sdfModifiers.f90:273-392)
    integer(c_int32_t), parameter :: SMCRT_EXTRUDE = 1, SMCRT_REVOLUTION = 2, SMCRT_ONION = 3, &
                                     SMCRT_ELONGATE = 4, SMCRT_TWIST = 5, SMCRT_BEND = 6
    ! sources (src/photon.f90)
    integer(c_int32_t), parameter :: SMCRT_SRC_PENCIL = 1, SMCRT_SRC_UNIFORM = 2, &
                                     SMCRT_SRC_CIRCULAR = 3, SMCRT_SRC_POINT = 4
    ! detectors (src/detectors/detectors.f90)
    integer(c_int32_t), parameter :: SMCRT_DET_CIRCLE = 1, SMCRT_DET_ANNULUS = 2, SMCRT_DET_CAMERA = 3
    integer, parameter :: SMCRT_MAX_DETECTORS = 16
    ! run flags
    integer(c_int32_t), parameter :: SMCRT_FLAG_MOMENTS = 1, SMCRT_FLAG_END_EARLY = 2, &
                                     SMCRT_FLAG_NO_TALLY = 4, SMCRT_FLAG_NO_LDS_TILE = 8, &
                                     SMCRT_FLAG_WEIGHT = 16, SMCRT_FLAG_SHARED_GRID = 32, &
                                     SMCRT_FLAG_DIRECT_TALLY = 64, SMCRT_FLAG_BINNED_TALLY = 128, &
                                     SMCRT_FLAG_GRID_COPIES_SHIFT = 8

    type, bind(C) :: smcrt_node
        integer(c_int32_t) :: op = 0, kind = 0
        real(c_double)     :: k = 0._c_double
        real(c_double)     :: p(8) = 0._c_double
        real(c_double)     :: t(12) = 0._c_double
    end type smcrt_node

    type, bind(C) :: smcrt_region
        integer(c_int32_t) :: first_node = 0, n_nodes = 0, material = 0, layer = 0
    end type smcrt_region

    type, bind(C) :: smcrt_material
        real(c_double) :: mus = 0._c_double, mua = 0._c_double, hgg = 0._c_double, n = 1._c_double
    end type smcrt_material

    type, bind(C) :: smcrt_grid
        integer(c_int32_t) :: nx = 0, ny = 0, nz = 0, pad = 0
        real(c_double)     :: xmax = 0._c_double, ymax = 0._c_double, zmax = 0._c_double
    end type smcrt_grid

    type, bind(C) :: smcrt_source
        integer(c_int32_t) :: kind = 0, pad = 0
        real(c_double)     :: pos(3) = 0._c_double, dir(3) = 0._c_double
        real(c_double)     :: p1(3) = 0._c_double, p2(3) = 0._c_double, p3(3) = 0._c_double
        real(c_double)     :: radius = 0._c_double
    end type smcrt_source

    type, bind(C) :: smcrt_detector
        integer(c_int32_t) :: kind = 0, layer = 1, nbins = 100, track_history = 0
        real(c_double)     :: pos(3) = 0._c_double, dir(3) = 0._c_double
        real(c_double)     :: p2(3) = 0._c_double, p3(3) = 0._c_double
        real(c_double)     :: radius1 = 0._c_double, radius2 = 0._c_double, maxval = 100._c_double
    end type smcrt_detector

    type, bind(C) :: smcrt_scene_desc
        integer(c_int32_t) :: abi_version = SMCRT_ABI
        integer(c_int32_t) :: n_nodes = 0, n_regions = 0, n_materials = 0
        type(c_ptr)        :: nodes = c_null_ptr, regions = c_null_ptr, materials = c_null_ptr
        type(smcrt_grid)   :: grid
        type(smcrt_source) :: source
        integer(c_int32_t) :: n_detectors = 0, n_poly_vertices = 0
        type(c_ptr)        :: detectors = c_null_ptr
        type(c_ptr)        :: poly_vertices = c_null_ptr
    end type smcrt_scene_desc

    type, bind(C) :: smcrt_run_opts
        integer(c_int64_t) :: n_packets = 0, seed = 123456789, first_packet = 0
        integer(c_int32_t) :: flags = 0, max_bounces = 0, block_size = 0, blocks_per_cu = 0
        integer(c_int32_t) :: chunk = 0, trace_packets = 0, trace_events = 0, walk_cap = 0
        integer(c_int32_t) :: refill_min = 0, grad_mask_lo = 0, grad_mask_hi = 0, waves = 0
    end type smcrt_run_opts

    type, bind(C) :: smcrt_stats
        integer(c_int64_t) :: launched, absorbed, escaped_r, escaped_t, killed_surface, killed_bounce
        integer(c_int64_t) :: scatters, ev_sweeps, ev_region_evals, ev_normals, ev_steps
        integer(c_int64_t) :: ev_crossings, ev_interfaces, ev_fresnel, err_unreachable, err_nan
        integer(c_int64_t) :: ev_atomics
        real(c_double)     :: moments(3, 4, 2)
        real(c_double)     :: kernel_ms, tally_sum
    end type smcrt_stats

    type, bind(C) :: smcrt_tallies
        type(c_ptr)       :: jmean = c_null_ptr, jmean_f32 = c_null_ptr, trace = c_null_ptr
        type(c_ptr)       :: detector_counts = c_null_ptr
        type(smcrt_stats) :: stats
    end type smcrt_tallies

    !> the TOML input (src/parse/parse.f90, parse_source.f90, parse_detectors.f90)
    type, bind(C) :: smcrt_config
        character(kind=c_char) :: source_name(32)
        integer(c_int64_t)     :: nphotons
        integer(c_int32_t)     :: has_position, has_direction
        real(c_double)         :: position(3), direction(3), point1(3), point2(3), point3(3)
        real(c_double)         :: radius, beta, radius_hi
        character(kind=c_char) :: annulus_type(32)
        integer(c_int32_t)     :: nxg, nyg, nzg, pad0
        real(c_double)         :: xmax, ymax, zmax
        character(kind=c_char) :: units(16)
        character(kind=c_char) :: geom_name(64)
        real(c_double)         :: tau, musb, muab, musc, muac, hgg
        integer(c_int32_t)     :: num_spheres, pad1
        character(kind=c_char) :: fluence(256), absorb(256), render(256)
        integer(c_int32_t)     :: render_geom, overwrite
        integer(c_int32_t)     :: render_size(3), pad2
        integer(c_int64_t)     :: iseed
        integer(c_int32_t)     :: tev, absorb_flag, load_checkpoint, checkpoint_every_n
        character(kind=c_char) :: checkpoint_file(256)
        integer(c_int32_t)     :: n_detectors, pad3
        type(smcrt_detector)   :: detectors(SMCRT_MAX_DETECTORS)
        character(kind=c_char) :: history_file(256)
    end type smcrt_config

    interface
        integer(c_int) function smcrt_abi_version() bind(C, name="smcrt_abi_version")
            import :: c_int
        end function smcrt_abi_version

        integer(c_int) function smcrt_last_error(buf, n) bind(C, name="smcrt_last_error")
            import :: c_int, c_char, c_size_t
            character(kind=c_char), intent(out) :: buf(*)
            integer(c_size_t), value :: n
        end function smcrt_last_error

        integer(c_int64_t) function smcrt_sizeof(name) bind(C, name="smcrt_sizeof")
            import :: c_int64_t, c_char
            character(kind=c_char), intent(in) :: name(*)
        end function smcrt_sizeof

        integer(c_int) function smcrt_scene_create(desc, scene) bind(C, name="smcrt_scene_create")
            import :: c_int, c_ptr, smcrt_scene_desc
            type(smcrt_scene_desc), intent(in) :: desc
            type(c_ptr), intent(out) :: scene
        end function smcrt_scene_create

        integer(c_int) function smcrt_scene_destroy(scene) bind(C, name="smcrt_scene_destroy")
            import :: c_int, c_ptr
            type(c_ptr), value :: scene
        end function smcrt_scene_destroy

        integer(c_int) function smcrt_scene_set_materials(scene, mats, n) &
                bind(C, name="smcrt_scene_set_materials")
            import :: c_int, c_ptr, c_int32_t, smcrt_material
            type(c_ptr), value :: scene
            type(smcrt_material), intent(in) :: mats(*)
            integer(c_int32_t), value :: n
        end function smcrt_scene_set_materials

        integer(c_int) function smcrt_run(scene, opts, out) bind(C, name="smcrt_run")
            import :: c_int, c_ptr, smcrt_run_opts, smcrt_tallies
            type(c_ptr), value :: scene
            type(smcrt_run_opts), intent(in) :: opts
            type(smcrt_tallies), intent(inout) :: out
        end function smcrt_run

        integer(c_int) function smcrt_run_device(scene, opts, d_jmean, stats, stream) &
                bind(C, name="smcrt_run_device")
            import :: c_int, c_ptr, smcrt_run_opts, smcrt_stats
            type(c_ptr), value :: scene
            type(smcrt_run_opts), intent(in) :: opts
            type(c_ptr), value :: d_jmean
            type(smcrt_stats), intent(out) :: stats
            type(c_ptr), value :: stream
        end function smcrt_run_device

        integer(c_int) function smcrt_evaluate(scene, region, n, pts, dist, normals) &
                bind(C, name="smcrt_evaluate")
            import :: c_int, c_ptr, c_int32_t, c_int64_t
            type(c_ptr), value :: scene
            integer(c_int32_t), value :: region
            integer(c_int64_t), value :: n
            type(c_ptr), value :: pts, dist, normals
        end function smcrt_evaluate

        integer(c_int) function smcrt_render(scene, samples, image) bind(C, name="smcrt_render")
            import :: c_int, c_ptr, c_int32_t, c_float
            type(c_ptr), value :: scene
            integer(c_int32_t), intent(in) :: samples(3)
            real(c_float), intent(out) :: image(*)
        end function smcrt_render

        integer(c_int64_t) function smcrt_detector_bins(scene, i) bind(C, name="smcrt_detector_bins")
            import :: c_int64_t, c_ptr, c_int32_t
            type(c_ptr), value :: scene
            integer(c_int32_t), value :: i
        end function smcrt_detector_bins

        integer(c_int) function smcrt_config_load(path, cfg) bind(C, name="smcrt_config_load")
            import :: c_int, c_char, smcrt_config
            character(kind=c_char), intent(in) :: path(*)
            type(smcrt_config), intent(out) :: cfg
        end function smcrt_config_load

        integer(c_int) function smcrt_normalise_fluence(grid, raw, n_packets, out) &
                bind(C, name="smcrt_normalise_fluence")
            import :: c_int, c_double, c_float, c_int64_t, smcrt_grid
            type(smcrt_grid), intent(in) :: grid
            real(c_double), intent(in) :: raw(*)
            integer(c_int64_t), value :: n_packets
            real(c_float), intent(out) :: out(*)
        end function smcrt_normalise_fluence

        integer(c_int) function smcrt_write_nrrd(path, data, nx, ny, nz, meta) &
                bind(C, name="smcrt_write_nrrd")
            import :: c_int, c_char, c_float, c_int32_t
            character(kind=c_char), intent(in) :: path(*)
            real(c_float), intent(in) :: data(*)
            integer(c_int32_t), value :: nx, ny, nz
            character(kind=c_char), intent(in) :: meta(*)
        end function smcrt_write_nrrd

        integer(c_int) function smcrt_write_detector(path, det, counts) &
                bind(C, name="smcrt_write_detector")
            import :: c_int, c_char, c_int64_t, smcrt_detector
            character(kind=c_char), intent(in) :: path(*)
            type(smcrt_detector), intent(in) :: det
            integer(c_int64_t), intent(in) :: counts(*)
        end function smcrt_write_detector
    end interface

contains

    !> smcrt_last_error as a Fortran string
    function smcrt_error_message() result(msg)
        character(len=:), allocatable :: msg
        character(kind=c_char) :: buf(1024)
        integer :: i, n, rc
        rc = smcrt_last_error(buf, int(size(buf), c_size_t))
        n = 0
        do i = 1, size(buf)
            if (buf(i) == c_null_char) exit
            n = i
        end do
        allocate(character(len=n) :: msg)
        do i = 1, n
            msg(i:i) = buf(i)
        end do
    end function smcrt_error_message

    !> a NUL-terminated C char array (from smcrt_config) as a Fortran string
    function c_string(a) result(s)
        character(kind=c_char), intent(in) :: a(:)
        character(len=:), allocatable :: s
        integer :: i, n
        n = 0
        do i = 1, size(a)
            if (a(i) == c_null_char) exit
            n = i
        end do
        allocate(character(len=n) :: s)
        do i = 1, n
            s(i:i) = a(i)
        end do
    end function c_string

end module smcrt
