! m3_harness.F90 — drives the compiled MPLS_RisingBubble reference on given
! inputs.
!
! TEST INFRASTRUCTURE ONLY.  This program is the project's own driver; the
! reference subroutines it calls are the internal subroutines of the
! reference's `MPLS_RisingBubble/20190612/MAIN.F90`, compiled from their
! sources where they lie (pulled in here with the same INCLUDE lines, and the
! modules of its DATA_MOD.F90).  It writes golden vectors for the CPU oracle;
! it is never part of the product.  FREE_ARRAY is never called (it frees two
! arrays that are never allocated); OPENMP_SETTINGS.F90 is left out (only the
! reference's main program calls it, and it passes an integer where
! omp_set_dynamic takes a logical, which flang rejects).
!
! usage: m3_harness <op> <indir> <outdir>
!   indir/params.txt  one line per value, in the order read below
!   indir/<field>     raw little-endian float64 full ghosted arrays
!                     (-2:NODE_X+3, -2:NODE_Y+3, -2:NODE_Z+3); missing files
!                     leave zeros
!   outdir            the same field files after the operation, plus
!                     scalars.txt
program fcfd_m3_harness
use precision
use problem_def
use ls_data, only: B1_iM, B2_iM, IMASS_LS, IVOL_LS, MASS_LS, VOL_LS, INTERFACE_WIDTH, LS_SOLVER, PHI
use vof_data, only: VOF_SOLVER
use fluid_properties, only: U, V, W
use dummy_data, only: S0
implicit none
character(len=512) :: op, indir, outdir
integer :: iu

call get_command_argument(1, op)
call get_command_argument(2, indir)
call get_command_argument(3, outdir)

open(newunit=iu, file=trim(indir)//'/params.txt', status='old')
read(iu, *) NODE_X, NODE_Y, NODE_Z
read(iu, *) XSTART, XEND, YSTART, YEND, ZSTART, ZEND
read(iu, *) DX, DY, DZ
read(iu, *) DT, RDT, INTERFACE_WIDTH
read(iu, *) RATIO_RHO, RATIO_AMU, RE, WE, FR
read(iu, *) ST_FORCE, G_FORCE, VEL_BC
read(iu, *) TOL_P, TOL_M, PRESS_OMEGA, PRESS_MAX_IT
read(iu, *) B1_iM, B2_iM, IMASS_LS, IVOL_LS, TIME, ITER, REC_MASS
close(iu)
XL = XEND - XSTART
YL = YEND - YSTART
ZL = ZEND - ZSTART
INTERFACE_METHOD = 0
IBM_SOLVER = 0
LS_SOLVER = 10
VOF_SOLVER = 1
TIME_TO_STOP = 1.0d30
METHOD_CNT = 0
GRID_CNT = 0

call ALLOCATE_ARRAY
call zero_all()
call load_all()

select case (trim(op))
case ('none')
case ('bc3d_phi')
    call BC3D(PHI)
case ('amurho')
    call AMURHO
case ('heavy_f')
    call HEAVY_F(PHI)
case ('ls_ccd')
    call LS_CCD(PHI)
case ('cell_face_velocity')
    call CELL_FACE_VELOCITY
case ('ls_source')
    call CELL_FACE_VELOCITY
    call LS_sorce(S0)
case ('ls_rk3')
    call LS_RK3_SOLVER
case ('gradfi')
    call SGNF(PHI)
    call GRADFI(PHI)
case ('redistance')
    call LS_REDISTANCE(PHI, 1)
case ('mpls1')
    call MPLS(B1_iM)
case ('solve_ls')
    call SOLVE_LS
case ('rhsp')
    call RHSP
case ('press_poisson')
    call PRESS_POISSON
case ('second_order_upwind')
    call SECOND_ORDER_UPWIND
case ('momentum')
    call SOLVE_MOMENTUM
    call VELOCITY_BC
case ('solve_ns')
    call SOLVE_NS
case ('find_mass')
    call FIND_MASS
case ('setup_tail')
    ! the part of SETUP after the initial level sets are set (SETUP.F90:210-222)
    call BC3D(PHI)
    call BC3D(U)
    call BC3D(V)
    call BC3D(W)
    call LS_REDISTANCE(PHI, 1)
    PLOT_INT = 0
    TIME = ITER * DT
    call AMURHO
    call STORE_DATA
    call SOLVE_NS
    call STORE_DATA
    call FIND_MASS
case ('step')
    ! one pass of the main loop (MAIN.F90:10-26) without output
    ITER = ITER + 1
    TIME = TIME + DT
    call SOLVE_VOF
    call SOLVE_LS
    call SOLVE_NS
    call FIND_MASS
    call STORE_DATA
case default
    write(*, *) 'unknown op ', trim(op)
    stop 2
end select

call save_all()

open(newunit=iu, file=trim(outdir)//'/scalars.txt', status='replace')
write(iu, '(9ES26.17E3, 2I12)') MASS_LS, VOL_LS, IMASS_LS, IVOL_LS, B1_iM, B2_iM, TIME, DIV, &
    L2_U + L2_V + L2_W, ITER, REC_MASS
close(iu)

contains

subroutine zero_all()
    use fluid_properties
    use ls_data
    use vof_data
    use ibm_data
    use dummy_data
    U = 0; UH = 0; U_OLD = 0; U_TMP = 0; U1 = 0; US = 0; US_OLD = 0
    V = 0; VH = 0; V_OLD = 0; V_TMP = 0; V1 = 0; VS = 0; VS_OLD = 0
    W = 0; WH = 0; W_OLD = 0; W_TMP = 0; W1 = 0; WS = 0; WS_OLD = 0
    RHO = 0; RHO_OLD = 0; AMU = 0; AMU_OLD = 0; P = 0; P_OLD = 0; RHS = 0
    PHI = 0; PHI_OLD = 0; PHI_TMP = 0; PHI_V = 0; PHI1 = 0; PHI2 = 0
    NORMAL_X = 0; NORMAL_Y = 0; NORMAL_Z = 0
    CURV = 0; HEAVY = 0; DELTA = 0; GRAD = 0; SGN = 0
    VOF_SX = 0; VOF_SY = 0; VOF_SZ = 0; VVOF = 0; VOF = 0; VOF_OLD = 0; VOF_X = 0; VOF_Y = 0; VOF_Z = 0
    S0 = 0; S1 = 0; S2 = 0; SS0 = 0; SS1 = 0; SS2 = 0
    UP = 0; UM = 0; VP = 0; VM = 0; WP = 0; WM = 0; A = 0; B = 0; C = 0; F = 0; G = 0; H = 0
    SOLID_PHI = 0; SOLID_VOF = 0
    FP = 0; FM = 0; HP = 0; HM = 0; GP = 0; GM = 0
    X = 0; Y = 0; Z = 0
end subroutine

subroutine rw(name, a, doread)
    character(len=*), intent(in) :: name
    real(8), intent(inout) :: a(:, :, :)
    logical, intent(in) :: doread
    integer :: uu, ios
    if (doread) then
        open(newunit=uu, file=trim(indir)//'/'//name, access='stream', form='unformatted', status='old', &
             iostat=ios)
        if (ios /= 0) return
        read(uu) a
    else
        open(newunit=uu, file=trim(outdir)//'/'//name, access='stream', form='unformatted', status='replace')
        write(uu) a
    endif
    close(uu)
end subroutine

subroutine io_all(doread)
    use fluid_properties
    use ls_data
    use dummy_data, only: S0
    logical, intent(in) :: doread
    call rw('phi', PHI, doread); call rw('phi1', PHI1, doread); call rw('phi2', PHI2, doread)
    call rw('phi_old', PHI_OLD, doread)
    call rw('u', U, doread); call rw('v', V, doread); call rw('w', W, doread)
    call rw('u_old', U_OLD, doread); call rw('v_old', V_OLD, doread); call rw('w_old', W_OLD, doread)
    call rw('uh', UH, doread); call rw('vh', VH, doread); call rw('wh', WH, doread)
    call rw('us', US, doread); call rw('vs', VS, doread); call rw('ws', WS, doread)
    call rw('us_old', US_OLD, doread); call rw('vs_old', VS_OLD, doread); call rw('ws_old', WS_OLD, doread)
    call rw('u1', U1, doread); call rw('v1', V1, doread); call rw('w1', W1, doread)
    call rw('rho', RHO, doread); call rw('rho_old', RHO_OLD, doread)
    call rw('amu', AMU, doread); call rw('amu_old', AMU_OLD, doread)
    call rw('p', P, doread); call rw('rhs', RHS, doread)
    call rw('heavy', HEAVY, doread); call rw('delta', DELTA, doread); call rw('grad', GRAD, doread)
    call rw('sgn', SGN, doread)
    call rw('normal_x', NORMAL_X, doread); call rw('normal_y', NORMAL_Y, doread)
    call rw('normal_z', NORMAL_Z, doread); call rw('curv', CURV, doread)
    call rw('s0', S0, doread)
end subroutine

subroutine load_all()
    call io_all(.true.)
end subroutine

subroutine save_all()
    call io_all(.false.)
end subroutine

INCLUDE "STORE_DATA.F90"
INCLUDE "SOLVE_PRESSURE_POISSON.F90"
INCLUDE "SOLVE_NS.F90"
INCLUDE "SETUP.F90"
INCLUDE "ALLOCATE_ARRAY.F90"
INCLUDE "PLOT.F90"
INCLUDE "LS_REDISTANCE.F90"
INCLUDE "SCHEME_WENO.F90"
INCLUDE "SCHEME_COMPACT.F90"
INCLUDE "FIND_ERROR.F90"
INCLUDE "BC.F90"
INCLUDE "FIND_CURVATURE.F90"
INCLUDE "AMURHO.F90"
INCLUDE "LS_SOLVER.F90"
INCLUDE "VORTEX_VELOCITY.F90"
INCLUDE "FIND_MASS.F90"
INCLUDE "PRINT_SCREEN.F90"
INCLUDE "TDMA.F90"
INCLUDE "LS_MAINTAIN.F90"
INCLUDE "VOF_SOLVER.F90"
INCLUDE "TO_FILE.F90"
INCLUDE "READ_DATA.F90"
INCLUDE "ERROR_ESTIMATION.F90"
INCLUDE "FIND_NORMALS.F90"
INCLUDE "VORTEX_IDENTIFICATION.F90"
INCLUDE "CPU_REC.F90"
INCLUDE "SECOND_ORDER_UPWIND.F90"
INCLUDE "QUICK.F90"

end program fcfd_m3_harness
