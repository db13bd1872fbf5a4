! n2_dropin.F90 — the newOpenMP 2D main loop with its hot path in libfcfd.
!
! TEST INFRASTRUCTURE: the Fortran drop-in, end to end.  It is built against
! the reference's own data modules (`Two phase flows/newOpenMP/2D/lib`,
! compiled where they lie by build_ref.sh) and the fcfd_c binding module
! (the ISO_C_BINDING interfaces of include/fcfd.h), and linked with
! libfcfd.so.  The
! loop is the reference's (`.../2D/src/main.f90:13-46`) with the switch and
! the hot-path subroutines answered by shims of the reference's names
! (interface_solver, ns_solver, ls_mv) that call the library; the fields stay
! on the device and are pulled once at the end.  Plotting and the screen
! report are left out.
!
! usage: n2_dropin <config> <indir> <outdir> <nsteps> <precision 0|1> <ppe_tol>
!   config   reference-format input file (`.../lib/data/tree.f90:48-106`)
!   indir    raw float64 phi (full ghosted array) of the initial condition;
!            the fluid starts at rest
!   outdir   phi, u, v, p, nu, nv after the last step and mass.txt (one line
!            per step: iter, mass, volume)
module all
use tree
implicit none
save
type(manager) :: p
end module all

module dropin_ctx
use iso_c_binding
implicit none
type(c_ptr) :: fcfd_ctx_main
end module dropin_ctx

program n2_dropin
use all
use fcfd_c
use dropin_ctx
implicit none
character(len=512) :: cfg, indir, outdir, arg
integer :: nsteps, prec, n, u, um
real(8) :: ppe_tol
type(fcfd_grid) :: g
type(fcfd_params) :: prm

call get_command_argument(1, cfg)
call get_command_argument(2, indir)
call get_command_argument(3, outdir)
call get_command_argument(4, arg); read(arg, *) nsteps
call get_command_argument(5, arg); read(arg, *) prec
call get_command_argument(6, arg); read(arg, *) ppe_tol

call p%init(trim(cfg))
open(newunit=u, file=trim(indir)//'/phi', access='stream', form='unformatted', status='old')
read(u) p%loc%phi%now
close(u)
p%loc%vel%x%now = 0.0d0
p%loc%vel%y%now = 0.0d0

! the context from p%glb after manager_init (dt, rdt, ls_wid already times dx)
g%dim = 2; g%nx = p%glb%node_x; g%ny = p%glb%node_y; g%nz = 1; g%ghc = p%glb%ghc; g%pad_ = 0
g%xstart = p%glb%xstart; g%xend = p%glb%xend; g%ystart = p%glb%ystart; g%yend = p%glb%yend
g%zstart = 0.0d0; g%zend = 1.0d0
prm%dt = p%glb%dt; prm%rdt = p%glb%rdt; prm%ls_wid = p%glb%ls_wid
prm%rho_12 = p%glb%rho_12; prm%mu_12 = p%glb%mu_12
prm%re = p%glb%re; prm%we = p%glb%we; prm%fr = p%glb%fr
prm%gx = p%glb%gx; prm%gy = p%glb%gy; prm%gz = 0.0d0
prm%btn_sf = p%glb%btn_sf; prm%btn_g = p%glb%btn_g; prm%imass = 0.0d0
prm%ppe_tol = ppe_tol; prm%p_w1 = p%glb%p_w1; prm%p_tol = p%glb%p_tol
prm%ubc = p%glb%ubc; prm%vbc = p%glb%vbc; prm%wbc = 1
prm%redis_interval = 5          ! interface_solver.f90:21 reinitialises every fifth step
prm%picard_iters = 4; prm%ppe_precision = prec; prm%ppe_max_cycles = 200; prm%ppe_krylov = 1
prm%ppe_inner = 1; prm%ls_scheme = 0; prm%ppe_nu1 = 2; prm%ppe_nu2 = 2; prm%ppe_nu_coarse = 6
prm%ppe_fused = 0; prm%ppe_rho_half = 0

call fcfd_check(fcfd_create(g, prm, fcfd_ctx_main), 'fcfd_create')
call fcfd_check(fcfd_push(fcfd_ctx_main, FCFD_PHI, p%loc%phi%now), 'fcfd_push phi')
call fcfd_check(fcfd_push(fcfd_ctx_main, FCFD_U, p%loc%vel%x%now), 'fcfd_push u')
call fcfd_check(fcfd_push(fcfd_ctx_main, FCFD_V, p%loc%vel%y%now), 'fcfd_push v')
call fcfd_check(fcfd_problem_init(fcfd_ctx_main, 0), 'problem_init')

open(newunit=um, file=trim(outdir)//'/mass.txt', status='replace')
do n = 1, nsteps
    p%glb%time = p%glb%time + p%glb%dt
    p%glb%iter = p%glb%iter + 1
    call switch
    call interface_solver
    call ns_solver
    call ls_mv
    write(um, '(I8, 2ES26.17E3)') p%glb%iter, p%glb%mass, p%glb%vol
end do
close(um)

call fcfd_check(fcfd_pull(fcfd_ctx_main, FCFD_PHI, p%loc%phi%now), 'fcfd_pull phi')
call fcfd_check(fcfd_pull(fcfd_ctx_main, FCFD_U, p%loc%vel%x%now), 'fcfd_pull u')
call fcfd_check(fcfd_pull(fcfd_ctx_main, FCFD_V, p%loc%vel%y%now), 'fcfd_pull v')
call fcfd_check(fcfd_pull(fcfd_ctx_main, FCFD_P, p%loc%p%now), 'fcfd_pull p')
call fcfd_check(fcfd_pull(fcfd_ctx_main, FCFD_NU, p%loc%nvel%x%now), 'fcfd_pull nu')
call fcfd_check(fcfd_pull(fcfd_ctx_main, FCFD_NV, p%loc%nvel%y%now), 'fcfd_pull nv')
call save('phi', p%loc%phi%now)
call save('u', p%loc%vel%x%now)
call save('v', p%loc%vel%y%now)
call save('p', p%loc%p%now)
call save('nu', p%loc%nvel%x%now)
call save('nv', p%loc%nvel%y%now)
call fcfd_check(fcfd_destroy(fcfd_ctx_main), 'fcfd_destroy')

contains

subroutine save(name, a)
    character(*), intent(in) :: name
    real(8), intent(in) :: a(:, :)
    integer :: uu
    open(newunit=uu, file=trim(outdir)//'/'//name, access='stream', form='unformatted', status='replace')
    write(uu) a
    close(uu)
end subroutine

! the shims: the reference's names, the library's work (time and step
! counter follow p%glb, as the reference's interface_solver reads them)
subroutine switch()
    call fcfd_check(fcfd_set_time(fcfd_ctx_main, p%glb%time, int(p%glb%iter, c_int)), 'fcfd_set_time')
    call fcfd_check(fcfd_switch(fcfd_ctx_main), 'switch')
end subroutine

subroutine interface_solver()
    call fcfd_check(fcfd_interface_solver(fcfd_ctx_main), 'interface_solver')
end subroutine

subroutine ns_solver()
    call fcfd_check(fcfd_ns_solver(fcfd_ctx_main), 'ns_solver')
end subroutine

subroutine ls_mv()
    call fcfd_check(fcfd_ls_mv(fcfd_ctx_main, p%glb%mass, p%glb%vol), 'ls_mv')
end subroutine

end program n2_dropin
