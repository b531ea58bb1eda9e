! pm_mefdrive.f90 -- a Fortran host that runs the mef Biot u-p time march on
! the MI355X through the pmhip module, as the reference's Main.f would
! after the mesh macro: structure set-up, materials and loads, then `nsteps`
! solv macro steps.  Each step appends the u, v, w, p history of node `pnode`
! to <out>_up_node.txt in printnode's format (i9,10e18.10) (Pdata.f:733-790);
! the final u (4, nnode) and dp (nnode) go to <out>_state.bin.
!
! Input: a stream file (written by poromec_amd.fdeck) holding, in order,
!   numel, nnodev, numat, nsteps, pnode, reorder        (int32)
!   dt, gravity(3), gmod                                 (real64)
!   solver, precond, maxit, ngram, precision, maxnlit, stop_crit (int32)
!   tol, nltol                                           (real64)
!   ix(21, numel), idv(4, nnodev)                        (int32)
!   x(3, nnodev), e(15, numat)                           (real64)
!   eload(7, numel), eloadp(7, numel), load(3, 200)      (int32)
!   fload(1000, 5, 200)                                  (real64)
program pm_mefdrive
  use iso_c_binding
  use pmhip
  implicit none
  character(len=512) :: deck, outp
  integer(c_int) :: numel, nnodev, numat, nsteps, pnode, reorder
  integer(c_int) :: isolver, iprecond, imaxit, ingram, iprec, imaxnlit, icrit
  real(c_double) :: dt, gravity(3), gmod, tol, nltol, t
  integer(c_int), allocatable :: ix(:), idv(:), eload(:), eloadp(:), load(:)
  real(c_double), allocatable :: x(:), e(:), fload(:), u(:), u0(:), dp(:), poro(:)
  type(c_ptr) :: ctx
  type(pm_solver_cfg) :: cfg
  type(pm_step_info) :: info
  integer(c_int) :: nnode, neq, ierr
  integer(c_int64_t) :: nad
  integer :: istep, k

  call get_command_argument(1, deck)
  call get_command_argument(2, outp)
  open(10, file=trim(deck), access='stream', form='unformatted', status='old', action='read')
  read(10) numel, nnodev, numat, nsteps, pnode, reorder
  read(10) dt, gravity, gmod
  read(10) isolver, iprecond, imaxit, ingram, iprec, imaxnlit, icrit
  read(10) tol, nltol
  allocate(ix(21 * numel), idv(4 * nnodev), x(3 * nnodev), e(15 * numat))
  allocate(eload(7 * numel), eloadp(7 * numel), load(3 * 200), fload(1000 * 5 * 200))
  read(10) ix, idv, x, e, eload, eloadp, load, fload
  close(10)

  ierr = pm_ctx_create(ctx, 0_c_int)
  call check(ierr, 'pm_ctx_create')
  ierr = pm_mesh_setup(ctx, numel, nnodev, ix, x, idv, reorder, nnode, neq, nad)
  call check(ierr, 'pm_mesh_setup')
  ierr = pm_set_materials(ctx, numat, e, eload, eloadp, load, fload, gravity, gmod)
  call check(ierr, 'pm_set_materials')
  write(*, '(a,i10,a,i10,a,i14)') ' nnode', nnode, ' neq', neq, ' nad', nad

  cfg%solver = isolver
  cfg%precond = iprecond
  cfg%maxit = imaxit
  cfg%nkrylov = ingram
  cfg%tol = tol
  cfg%precision = iprec
  cfg%maxnlit = imaxnlit
  cfg%nltol = nltol
  cfg%stop_crit = icrit
  allocate(u(4 * nnode), u0(4 * nnode), dp(nnode), poro(nnode))
  open(20, file=trim(outp)//'_up_node.txt', status='replace', action='write')
  write(20, '(a,i9)') '# u v w p history of node', pnode
  do istep = 1, nsteps
    call step_pm_hip(ctx, dt, cfg, info, ierr)
    call check(ierr, 'pm_step')
    ierr = pm_get_state(ctx, u, u0, dp, poro, t)
    call check(ierr, 'pm_get_state')
    write(20, '(i9,10e18.10)') istep, t, (u(4 * (pnode - 1) + k), k = 1, 4)
    write(*, '(a,i6,a,i3,a,i8)') ' step', istep, ' nlit', info%nlit, ' iters', info%iters(1)
  end do
  close(20)
  open(30, file=trim(outp)//'_state.bin', access='stream', form='unformatted', status='replace')
  write(30) u, dp
  close(30)
  call pm_ctx_destroy(ctx)

contains
  subroutine check(code, what)
    integer(c_int), intent(in) :: code
    character(len=*), intent(in) :: what
    if (code /= PM_OK) then
      write(*, '(a,a,a,i4)') ' *** ', what, ' failed: code ', code
      stop 1
    end if
  end subroutine
end program pm_mefdrive
