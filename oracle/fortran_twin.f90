! Driver for the reference's Fortran twin (fortran/*.f90), compiled with its
! sources where they lie (oracle/Makefile, target `twin`) into oracle/_ref/.
! Test infrastructure only: it lets the tests compare the Fortran program's
! generator and colloid stencil with the C++ reference and the oracle.
!
!   fortran_twin ran1 <seed> <n> <out>   n draws of ran1 (fortran/ran1.f90) from
!                                        <seed>, as raw float64
!   fortran_twin create_box <out>        nbox (int32), then box_neigh(1:nbox, cell)
!                                        for every cell of the twin's 30^3 box
!                                        (fortran/create_box.f90), raw int32,
!                                        1-based, cells in the twin's order
!   fortran_twin stream_bin <in> <out>   reads pos_fl, vel_fl (3 per fluid
!                                        particle) and pos_colloid (3 per
!                                        colloid) as raw float64, streams one
!                                        substep (fortran/update_pos_mpcd.f90)
!                                        and bins (fortran/neighbour_list_mpcd.f90);
!                                        writes pos_fl (float64), then no_neigh
!                                        and neigh_fl(1:10000, colloid) (int32)
!   fortran_twin rotate <in> <out>       reads pos_fl and vel_fl, applies one SRD
!                                        rotation with its thermostat
!                                        (fortran/rotation_mpcd.f90; the shift and
!                                        axes drawn by ran1 from the twin's seed
!                                        77777), writes vel_fl
!   fortran_twin force <in> <out>        reads pos_colloid, builds the colloid
!                                        pair list and the WCA forces
!                                        (fortran/neighbour_list_md.f90,
!                                        fortran/compute_force_md.f90), writes f
!                                        (3 per colloid), then potential_colloid
!   fortran_twin collide <in> <out>      reads pos_fl, vel_fl, pos_colloid,
!                                        vel_colloid and ang_vel_colloid, bins
!                                        the fluid and applies the stochastic
!                                        reflections (fortran/fluid_colloid_collision.f90,
!                                        stochastic_reflection.f90, gauss.f90;
!                                        draws by ran1 from seed 77777); writes
!                                        pos_fl, vel_fl, vel_colloid, ang_vel_colloid
!   fortran_twin colloid <in> <out>      reads pos_fl, vel_fl, pos_colloid,
!                                        vel_colloid, ang_vel_colloid, ra, f and
!                                        old_force, bins the fluid, then run
!                                        (the active kick, fortran/run.f90),
!                                        update_activity_direction, update_pos_md
!                                        and update_velocity_colloid; writes
!                                        vel_fl, pos_colloid, vel_colloid, ra
!   fortran_twin startup <out>           the twin's start-up from its seed 77777:
!                                        initialize_colloid, initialize_fluid,
!                                        tumble; writes pos_fl, vel_fl,
!                                        pos_colloid, vel_colloid,
!                                        ang_vel_colloid, ra
!   fortran_twin steps <k> <out>         the start-up, the initial lists, forces
!                                        and orientations, then k MPCD steps in
!                                        the order of fortran/active_brownian.f90:7-30
!                                        (rotation, run, n substeps); writes the
!                                        same arrays as startup
program fortran_twin
    use all_parameters
    implicit none
    double precision :: ran1
    character(len=256) :: mode, a1, a2, path
    integer :: seed, ndraw, i, u, ksteps, l
    double precision, allocatable :: draws(:)

    call get_command_argument(1, mode)
    if (trim(mode) == 'ran1') then
        call get_command_argument(2, a1)
        call get_command_argument(3, a2)
        call get_command_argument(4, path)
        read (a1, *) seed
        read (a2, *) ndraw
        allocate (draws(ndraw))
        do i = 1, ndraw
            draws(i) = ran1(seed)
        end do
        open (newunit=u, file=trim(path), access='stream', form='unformatted', status='replace')
        write (u) draws
        close (u)
    else if (trim(mode) == 'create_box') then
        call get_command_argument(2, path)
        call create_box
        open (newunit=u, file=trim(path), access='stream', form='unformatted', status='replace')
        write (u) nbox
        write (u) box_neigh(1:nbox, 1:lx*ly*lz)
        close (u)
    else if (trim(mode) == 'stream_bin') then
        call get_command_argument(2, a1)
        call get_command_argument(3, path)
        open (newunit=u, file=trim(a1), access='stream', form='unformatted', status='old')
        read (u) pos_fl
        read (u) vel_fl
        read (u) pos_colloid
        close (u)
        call create_box
        call update_pos_mpcd
        call neighbour_list_mpcd
        open (newunit=u, file=trim(path), access='stream', form='unformatted', status='replace')
        write (u) pos_fl
        write (u) no_neigh
        write (u) neigh_fl
        close (u)
    else if (trim(mode) == 'rotate') then
        call get_command_argument(2, a1)
        call get_command_argument(3, path)
        open (newunit=u, file=trim(a1), access='stream', form='unformatted', status='old')
        read (u) pos_fl
        read (u) vel_fl
        close (u)
        nn = 1
        call rotation_mpcd
        open (newunit=u, file=trim(path), access='stream', form='unformatted', status='replace')
        write (u) vel_fl
        close (u)
    else if (trim(mode) == 'force') then
        call get_command_argument(2, a1)
        call get_command_argument(3, path)
        open (newunit=u, file=trim(a1), access='stream', form='unformatted', status='old')
        read (u) pos_colloid
        close (u)
        call neighbour_list_md
        call compute_force_md
        open (newunit=u, file=trim(path), access='stream', form='unformatted', status='replace')
        write (u) f
        write (u) potential_colloid
        close (u)
    else if (trim(mode) == 'collide') then
        call get_command_argument(2, a1)
        call get_command_argument(3, path)
        open (newunit=u, file=trim(a1), access='stream', form='unformatted', status='old')
        read (u) pos_fl
        read (u) vel_fl
        read (u) pos_colloid
        read (u) vel_colloid
        read (u) ang_vel_colloid
        close (u)
        call create_box
        call neighbour_list_mpcd
        call fluid_colloid_collision
        open (newunit=u, file=trim(path), access='stream', form='unformatted', status='replace')
        write (u) pos_fl
        write (u) vel_fl
        write (u) vel_colloid
        write (u) ang_vel_colloid
        close (u)
    else if (trim(mode) == 'colloid') then
        call get_command_argument(2, a1)
        call get_command_argument(3, path)
        open (newunit=u, file=trim(a1), access='stream', form='unformatted', status='old')
        read (u) pos_fl
        read (u) vel_fl
        read (u) pos_colloid
        read (u) vel_colloid
        read (u) ang_vel_colloid
        read (u) ra
        read (u) f
        read (u) old_force
        close (u)
        call create_box
        call neighbour_list_mpcd
        call run
        call update_activity_direction
        call update_pos_md
        call update_velocity_colloid
        open (newunit=u, file=trim(path), access='stream', form='unformatted', status='replace')
        write (u) vel_fl
        write (u) pos_colloid
        write (u) vel_colloid
        write (u) ra
        close (u)
    else if (trim(mode) == 'startup') then
        call get_command_argument(2, path)
        call initialize_colloid
        call initialize_fluid
        call tumble
        open (newunit=u, file=trim(path), access='stream', form='unformatted', status='replace')
        write (u) pos_fl
        write (u) vel_fl
        write (u) pos_colloid
        write (u) vel_colloid
        write (u) ang_vel_colloid
        write (u) ra
        close (u)
    else if (trim(mode) == 'steps') then
        call get_command_argument(2, a1)
        call get_command_argument(3, path)
        read (a1, *) ksteps
        call initialize_colloid
        call initialize_fluid
        call create_box
        call neighbour_list_mpcd
        call neighbour_list_md
        call compute_force_md
        call tumble
        do nn = 1, ksteps
            call rotation_mpcd
            call run
            do l = 1, n
                old_force = f
                call update_pos_md
                call neighbour_list_md
                call update_pos_mpcd
                call neighbour_list_mpcd
                call fluid_colloid_collision
                call update_activity_direction
                call compute_force_md
                call update_velocity_colloid
            end do
        end do
        open (newunit=u, file=trim(path), access='stream', form='unformatted', status='replace')
        write (u) pos_fl
        write (u) vel_fl
        write (u) pos_colloid
        write (u) vel_colloid
        write (u) ang_vel_colloid
        write (u) ra
        close (u)
    else
        write (*, *) 'usage: fortran_twin ran1 <seed> <n> <out> | create_box <out> | stream_bin <in> <out>'
        write (*, *) '       | rotate <in> <out> | force <in> <out> | collide <in> <out> | colloid <in> <out>'
        write (*, *) '       | startup <out> | steps <k> <out>'
        stop 2
    end if
end program fortran_twin
