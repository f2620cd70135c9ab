!> Layout check of the iso_c_binding mirrors: c_sizeof of every bind(C) type against the
!! library's smcrt_sizeof.  Prints one line per struct; exit status 1 on any mismatch.
program smcrt_abi_check
    use, intrinsic :: iso_c_binding
    use smcrt
    implicit none
    type(smcrt_node) :: node
    type(smcrt_region) :: region
    type(smcrt_material) :: material
    type(smcrt_grid) :: grid
    type(smcrt_source) :: source
    type(smcrt_scene_desc) :: desc
    type(smcrt_run_opts) :: opts
    type(smcrt_stats) :: stats
    type(smcrt_tallies) :: tallies
    type(smcrt_detector) :: det
    type(smcrt_config) :: cfg
    integer :: bad

    bad = 0
    call cmp("node", c_sizeof(node))
    call cmp("region", c_sizeof(region))
    call cmp("material", c_sizeof(material))
    call cmp("grid", c_sizeof(grid))
    call cmp("source", c_sizeof(source))
    call cmp("scene_desc", c_sizeof(desc))
    call cmp("run_opts", c_sizeof(opts))
    call cmp("stats", c_sizeof(stats))
    call cmp("tallies", c_sizeof(tallies))
    call cmp("detector", c_sizeof(det))
    call cmp("config", c_sizeof(cfg))
    if (smcrt_abi_version() /= SMCRT_ABI) then
        print '(a)', "abi version mismatch"
        bad = bad + 1
    end if
    if (bad > 0) stop 1

contains

    subroutine cmp(name, fsize)
        character(*), intent(in) :: name
        integer(c_size_t), intent(in) :: fsize
        integer(c_int64_t) :: csize
        csize = smcrt_sizeof(name//c_null_char)
        print '(a,1x,i0,1x,i0)', name, fsize, csize
        if (csize /= int(fsize, c_int64_t)) bad = bad + 1
    end subroutine cmp

end program smcrt_abi_check
