! ref_harness.F90 — drives the compiled Fortran reference on given inputs.
!
! TEST INFRASTRUCTURE ONLY.  This program is the project's own driver; the
! reference subroutines it calls are compiled from their sources where they lie
! under the reference checkout (`Two phase flows/newOpenMP/{2D,3D+IBM}/src`,
! pulled in through the reference's own `included.f90`).  It writes golden
! vectors for the CPU oracle; it is never part of the product.
!
! usage: ref_harness <op> <config> <indir> <outdir> [nsteps]
!   (ops plot and output write the reference's files under ./out/)
!   config  reference-format input file (`.../lib/data/tree.f90:48-106`)
!   indir   raw little-endian float64 files named after the fields (phi, u ...)
!           holding the full ghosted arrays; missing files leave zeros
!   outdir  the same field files after the operation, plus scalars.txt
! Build with -DHARNESS_3D for the 3D+IBM reference.
module all
use tree
implicit none
save
type(manager) :: p
end module all

program fcfd_ref_harness
use all
implicit none
character(len=512) :: op, cfg, indir, outdir
real(8) :: imass_in, time_in
integer :: iter_in, ios, u, nsteps, istep
integer(8) :: clk0, clk1, clk_rate
character(len=32) :: nstr

call get_command_argument(1, op)
call get_command_argument(2, cfg)
call get_command_argument(3, indir)
call get_command_argument(4, outdir)

call p%init(trim(cfg))

call zero_all()
call load_all()

imass_in = 0.0d0; time_in = 0.0d0; iter_in = 0
open(newunit=u, file=trim(indir)//'/scalars.txt', status='old', iostat=ios)
if (ios == 0) then
    read(u, *) imass_in, time_in, iter_in
    close(u)
endif
p%glb%imass = imass_in
p%glb%time = time_in
p%glb%iter = iter_in
#ifdef HARNESS_3D
p%loc%ibm%z = 0.0d0
p%loc%ibm%w = 0.0d0
#endif

select case (trim(op))
case ('none')
case ('bc_phi')
    call bc(p%loc%phi%now)
case ('velbc')
#ifdef HARNESS_3D
    call velbc(p%loc%vel%x%now, p%loc%vel%y%now, p%loc%vel%z%now)
#else
    call velbc(p%loc%vel%x%now, p%loc%vel%y%now)
#endif
case ('ls_funs')
    call ls_funs
case ('rho_mu')
    call rho_mu
case ('ls_mv')
    call ls_mv
case ('ls_adv')
    call p%loc%nvel%switch
    call level_set_rk3_solver
case ('uccd')
    call p%loc%nvel%switch
    call harness_uccd
case ('redis0')
    call level_set_rk3_redis(0)
case ('redis1')
    call level_set_rk3_redis(1)
case ('mpls')
    call mass_preserving_level_set
case ('surface_norms')
    call surface_norms
case ('curv')
    call curv
case ('ppe_init')
    call ppe_sor_init
case ('ppe_solve')
    call ppe_sor_init
    p%glb%piter = 0
    call ppe_sor_solver(p%glb%p_tol)
case ('projection')
    call ns_momentum_correction
case ('node_vel')
    call node_vel
case ('ns_setup')
    call ns_linearize
    call ns_ab_setup
    call ns_ab_predictor
case ('ns_solver')
    call harness_ns_solver
case ('interface')
    call interface_solver
case ('step')
    p%glb%time = p%glb%time + p%glb%dt
    p%glb%iter = p%glb%iter + 1
    call p%switch
    call interface_solver
    call harness_ns_solver
    call ls_mv
case ('steps_timed')
    ! argument 5 main-loop steps, timed inside the program (no file I/O in
    ! the timed region) with system_clock; timing.txt gets the seconds, the
    ! step count and the summed pressure relaxation sweeps
    call get_command_argument(5, nstr)
    read(nstr, *) nsteps
    call system_clock(clk0, clk_rate)
    do istep = 1, nsteps
        p%glb%time = p%glb%time + p%glb%dt
        p%glb%iter = p%glb%iter + 1
        call p%switch
        call interface_solver
        call harness_ns_solver
        call ls_mv
    end do
    call system_clock(clk1)
    open(newunit=u, file=trim(outdir)//'/timing.txt', status='replace')
    write(u, '(ES24.16, 2I12)') real(clk1 - clk0, 8) / real(clk_rate, 8), nsteps, p%glb%piter
    close(u)
case ('plot')
    ! plot snapshot 0 (plot.f90:10 writes when |time - pid*t2p| <= dt)
    p%glb%pid = 0
    p%glb%time = 0.0d0
    call plot
case ('output')
    ! one _MVloss.plt row: ls_mv of the given phi against imass from the
    ! scalars and ivol = 1.01 * vol (so both loss columns are non-zero)
    p%glb%time = time_in
    call ls_mv
    p%glb%ivol = 1.01d0 * p%glb%vol
    call output
case ('init')
    call harness_problem_init(.true.)
case ('init_nore')
    call harness_problem_init(.false.)
case default
    write(*,*) 'unknown op ', trim(op)
    stop 2
end select

call dump_all()

contains

include 'included.f90'

! root_solve("uccd") of phi on every interior line with nvel%old as the
! upwind selector (the call of the reference's commented UCCD level-set
! source, `3D+IBM/src/interface_solver/level_set_rk3_solver.f90:109,123,137`,
! in the current ccd_root interface); fx -> normals x/y/z, fxx -> xx/yy/zz.
subroutine harness_uccd()
    integer :: i, j
#ifdef HARNESS_3D
    integer :: k
    do k = p%loc%ks, p%loc%ke
    do j = p%loc%js, p%loc%je
        call p%loc%ccdsolvers%x%solve("uccd", p%loc%phi%now(:,j,k), p%loc%normals%x%now(:,j,k), &
            & p%loc%normals%xx%now(:,j,k), p%loc%nvel%x%old(:,j,k))
    end do
    end do
    do k = p%loc%ks, p%loc%ke
    do i = p%loc%is, p%loc%ie
        call p%loc%ccdsolvers%y%solve("uccd", p%loc%phi%now(i,:,k), p%loc%normals%y%now(i,:,k), &
            & p%loc%normals%yy%now(i,:,k), p%loc%nvel%y%old(i,:,k))
    end do
    end do
    do j = p%loc%js, p%loc%je
    do i = p%loc%is, p%loc%ie
        call p%loc%ccdsolvers%z%solve("uccd", p%loc%phi%now(i,j,:), p%loc%normals%z%now(i,j,:), &
            & p%loc%normals%zz%now(i,j,:), p%loc%nvel%z%old(i,j,:))
    end do
    end do
#else
    do j = p%loc%js, p%loc%je
        call p%loc%ccdsolvers%x%solve("uccd", p%loc%phi%now(:,j), p%loc%normals%x%now(:,j), &
            & p%loc%normals%xx%now(:,j), p%loc%nvel%x%old(:,j))
    end do
    do i = p%loc%is, p%loc%ie
        call p%loc%ccdsolvers%y%solve("uccd", p%loc%phi%now(i,:), p%loc%normals%y%now(i,:), &
            & p%loc%normals%yy%now(i,:), p%loc%nvel%y%old(i,:))
    end do
#endif
end subroutine

! ns_solver without the 3D IBM gate (`3D+IBM/src/ns_solver/ns_solvers.f90:1-20`
! minus ibm_bc, which is out of scope); the 2D reference has no ibm_bc.
subroutine harness_ns_solver()
    call ns_init
    call ns_ab_solver_SOR
    call ns_check_convergence_div
    call node_vel
end subroutine

! problem_init after the initial condition (`.../setup/problem_init.f90`):
! bc, velbc, redis(0), node_vel, ns_init, p%loc%init, ns_ab_setup, ls_mv.
subroutine harness_problem_init(redis0)
    logical :: redis0
    call bc(p%loc%phi%now)
#ifdef HARNESS_3D
    call velbc(p%loc%vel%x%now, p%loc%vel%y%now, p%loc%vel%z%now)
#else
    call velbc(p%loc%vel%x%now, p%loc%vel%y%now)
#endif
    if (redis0) call level_set_rk3_redis(0)
    call node_vel
    call ns_init
    call p%loc%init
    call ns_ab_setup
    call ls_mv
    p%glb%imass = p%glb%mass
    p%glb%time = 0.0d0
    p%glb%iter = 0
end subroutine

subroutine do_fields(mode)
    integer :: mode
#ifdef HARNESS_3D
    call fld(mode, 'phi', p%loc%phi%now)
    call fld(mode, 'phi_old', p%loc%phi%old)
    call fld(mode, 'u', p%loc%vel%x%now)
    call fld(mode, 'v', p%loc%vel%y%now)
    call fld(mode, 'w', p%loc%vel%z%now)
    call fld(mode, 'u_old', p%loc%vel%x%old)
    call fld(mode, 'v_old', p%loc%vel%y%old)
    call fld(mode, 'w_old', p%loc%vel%z%old)
    call fld(mode, 'nu', p%loc%nvel%x%now)
    call fld(mode, 'nv', p%loc%nvel%y%now)
    call fld(mode, 'nw', p%loc%nvel%z%now)
    call fld(mode, 'src_u', p%loc%velsrc%x%now)
    call fld(mode, 'src_v', p%loc%velsrc%y%now)
    call fld(mode, 'src_w', p%loc%velsrc%z%now)
    call fld(mode, 'src_u_old', p%loc%velsrc%x%old)
    call fld(mode, 'src_v_old', p%loc%velsrc%y%old)
    call fld(mode, 'src_w_old', p%loc%velsrc%z%old)
    call fld(mode, 'p', p%loc%p%now)
    call fld(mode, 'heavy', p%loc%heavy%now)
    call fld(mode, 'delta', p%loc%delta%now)
    call fld(mode, 'sign', p%loc%sign%now)
    call fld(mode, 'rho', p%loc%rho%now)
    call fld(mode, 'mu', p%loc%mu%now)
    call fld(mode, 'grad', p%loc%grad%now)
    call fld(mode, 'nx', p%loc%normals%x%now)
    call fld(mode, 'ny', p%loc%normals%y%now)
    call fld(mode, 'nz', p%loc%normals%z%now)
    call fld(mode, 'nxx', p%loc%normals%xx%now)
    call fld(mode, 'nyy', p%loc%normals%yy%now)
    call fld(mode, 'nzz', p%loc%normals%zz%now)
    call fld(mode, 'nxy', p%loc%normals%xy%now)
    call fld(mode, 'nxz', p%loc%normals%xz%now)
    call fld(mode, 'nyz', p%loc%normals%yz%now)
    call fld(mode, 'curv', p%loc%normals%curv%now)
    call fld(mode, 'ppe_src', p%loc%coe%src)
    call fld(mode, 'rho_old', p%loc%rho%old)
    call fld(mode, 'mu_old', p%loc%mu%old)
    call fld(mode, 'delta_old', p%loc%delta%old)
    call fld(mode, 'curv_old', p%loc%normals%curv%old)
    call fld(mode, 'nx_old', p%loc%normals%x%old)
    call fld(mode, 'ny_old', p%loc%normals%y%old)
    call fld(mode, 'nz_old', p%loc%normals%z%old)
    call fld(mode, 'coe_r', p%loc%coe%r)
    call fld(mode, 'coe_l', p%loc%coe%l)
    call fld(mode, 'coe_f', p%loc%coe%f)
    call fld(mode, 'coe_b', p%loc%coe%b)
    call fld(mode, 'coe_u', p%loc%coe%u)
    call fld(mode, 'coe_d', p%loc%coe%d)
    call fld(mode, 'coe_c', p%loc%coe%c)
#else
    call fld(mode, 'phi', p%loc%phi%now)
    call fld(mode, 'phi_old', p%loc%phi%old)
    call fld(mode, 'u', p%loc%vel%x%now)
    call fld(mode, 'v', p%loc%vel%y%now)
    call fld(mode, 'u_old', p%loc%vel%x%old)
    call fld(mode, 'v_old', p%loc%vel%y%old)
    call fld(mode, 'nu', p%loc%nvel%x%now)
    call fld(mode, 'nv', p%loc%nvel%y%now)
    call fld(mode, 'src_u', p%loc%velsrc%x%now)
    call fld(mode, 'src_v', p%loc%velsrc%y%now)
    call fld(mode, 'src_u_old', p%loc%velsrc%x%old)
    call fld(mode, 'src_v_old', p%loc%velsrc%y%old)
    call fld(mode, 'p', p%loc%p%now)
    call fld(mode, 'heavy', p%loc%heavy%now)
    call fld(mode, 'delta', p%loc%delta%now)
    call fld(mode, 'sign', p%loc%sign%now)
    call fld(mode, 'rho', p%loc%rho%now)
    call fld(mode, 'mu', p%loc%mu%now)
    call fld(mode, 'grad', p%loc%grad%now)
    call fld(mode, 'nx', p%loc%normals%x%now)
    call fld(mode, 'ny', p%loc%normals%y%now)
    call fld(mode, 'nxx', p%loc%normals%xx%now)
    call fld(mode, 'nyy', p%loc%normals%yy%now)
    call fld(mode, 'nxy', p%loc%normals%xy%now)
    call fld(mode, 'curv', p%loc%normals%curv%now)
    call fld(mode, 'ppe_src', p%loc%coe%src)
    call fld(mode, 'rho_old', p%loc%rho%old)
    call fld(mode, 'mu_old', p%loc%mu%old)
    call fld(mode, 'delta_old', p%loc%delta%old)
    call fld(mode, 'curv_old', p%loc%normals%curv%old)
    call fld(mode, 'nx_old', p%loc%normals%x%old)
    call fld(mode, 'ny_old', p%loc%normals%y%old)
    call fld(mode, 'coe_r', p%loc%coe%r)
    call fld(mode, 'coe_l', p%loc%coe%l)
    call fld(mode, 'coe_f', p%loc%coe%f)
    call fld(mode, 'coe_b', p%loc%coe%b)
    call fld(mode, 'coe_c', p%loc%coe%c)
#endif
end subroutine

subroutine fld(mode, nm, arr)
    integer :: mode
    character(len=*) :: nm
    real(8), dimension(*) :: arr
    if (mode == 0) then
        arr(1:field_size()) = 0.0d0
    else if (mode == 1) then
        call load_field(nm, arr)
    else
        call dump_field(nm, arr)
    endif
end subroutine

subroutine zero_all()
    call do_fields(0)
#ifdef HARNESS_3D
    p%loc%ibm%solid%now = 0.0d0
#endif
end subroutine

subroutine load_all()
    call do_fields(1)
end subroutine

subroutine dump_all()
    integer :: uu
    call do_fields(2)
    open(newunit=uu, file=trim(outdir)//'/scalars.txt', status='replace')
    write(uu, '(ES25.17,1X,ES25.17,1X,ES25.17,1X,ES25.17,1X,I10,1X,ES25.17)') p%glb%mass, p%glb%vol, &
        p%glb%imass, p%glb%time, p%glb%piter, p%glb%ppe_linf
    close(uu)
end subroutine

subroutine load_field(nm, arr)
    character(len=*) :: nm
    real(8), dimension(*) :: arr
    integer :: uu, ios2, n
    n = field_size()
    open(newunit=uu, file=trim(indir)//'/'//nm, access='stream', form='unformatted', status='old', iostat=ios2)
    if (ios2 /= 0) return
    read(uu) arr(1:n)
    close(uu)
end subroutine

subroutine dump_field(nm, arr)
    character(len=*) :: nm
    real(8), dimension(*) :: arr
    integer :: uu, n
    n = field_size()
    open(newunit=uu, file=trim(outdir)//'/'//nm, access='stream', form='unformatted', status='replace')
    write(uu) arr(1:n)
    close(uu)
end subroutine

integer function field_size()
    field_size = (p%glb%node_x + 2*p%glb%ghc) * (p%glb%node_y + 2*p%glb%ghc)
#ifdef HARNESS_3D
    field_size = field_size * (p%glb%node_z + 2*p%glb%ghc)
#endif
end function

end program fcfd_ref_harness
