! pmhip.f90 -- Fortran (iso_c_binding) interface to libpmhip, the MI355X
! implementation of the mef Biot u-p solve path, and drop-in wrappers with
! the reference's argument lists.
!
!   solv_pm_hip   replaces solv_pm      (Solv.f:27-35)  on the caller's CSRC
!   pform_pm_hip  replaces pform_pm     (Pform.f:1-19)  isw 2 / 4, device matrix
!   step_pm_hip   replaces the solv macro body          (Main.f:876-1123)
!
! All arrays keep the reference's layouts (column-major, 1-based numbers).
module pmhip
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: PM_OK = 0
  integer(c_int), parameter :: PM_PREC_FP64 = 0, PM_PREC_FP32IR = 1

  type, bind(C) :: pm_solver_cfg
    integer(c_int) :: solver = 11
    integer(c_int) :: precond = 2
    integer(c_int) :: maxit = 50000
    integer(c_int) :: nkrylov = 150
    real(c_double) :: tol = 1.0d-11
    integer(c_int) :: precision = 0
    real(c_double) :: ir_inner_tol = 1.0d-5
    integer(c_int) :: ir_maxouter = 10
    integer(c_int) :: maxnlit = 500
    real(c_double) :: nltol = 1.0d-4
    integer(c_int) :: stop_crit = 2
    integer(c_int) :: maxnlit_accept = 0
  end type

  type, bind(C) :: pm_solve_info
    integer(c_int) :: iters, outer, used_solver
    real(c_double) :: resid, true_rel, ms
    integer(c_int) :: spmv64, spmv32
  end type

  type, bind(C) :: pm_step_info
    integer(c_int) :: nlit, nsolve
    integer(c_int) :: iters(16), used(16)
    real(c_double) :: true_rel(16)
    real(c_double) :: ms_total, ms_solve, ms_element
  end type

  interface
    integer(c_int) function pm_ctx_create(ctx, device) bind(C, name='pm_ctx_create')
      import :: c_ptr, c_int
      type(c_ptr) :: ctx
      integer(c_int), value :: device
    end function
    subroutine pm_ctx_destroy(ctx) bind(C, name='pm_ctx_destroy')
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine
    integer(c_int) function pm_mesh_setup(ctx, numel, nnodev, ix, x, idv, reorder, nnode, neq, &
                                          nad) bind(C, name='pm_mesh_setup')
      import :: c_ptr, c_int, c_double, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: numel, nnodev, reorder
      integer(c_int) :: ix(*), idv(*)
      real(c_double) :: x(*)
      integer(c_int) :: nnode, neq
      integer(c_int64_t) :: nad
    end function
    integer(c_int) function pm_mesh_setup_nen(ctx, nen, numel, nnodev, ix, x, idv, reorder, &
                                              nnode, neq, nad) bind(C, name='pm_mesh_setup_nen')
      import :: c_ptr, c_int, c_double, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: nen, numel, nnodev, reorder
      integer(c_int) :: ix(*), idv(*)
      real(c_double) :: x(*)
      integer(c_int) :: nnode, neq
      integer(c_int64_t) :: nad
    end function
    integer(c_int) function pm_mesh_numbering(numel, nnodev, ix, idv, reorder, nnode_max, ix20, &
                                              id, perm, nnode, neq, nad) &
        bind(C, name='pm_mesh_numbering')
      import :: c_int, c_int64_t
      integer(c_int), value :: numel, nnodev, reorder, nnode_max
      integer(c_int) :: ix(*), idv(*), ix20(*), id(*), perm(*)
      integer(c_int) :: nnode, neq
      integer(c_int64_t) :: nad
    end function
    integer(c_int) function pm_get_numbering(ctx, ix20, id, perm) bind(C, name='pm_get_numbering')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int) :: ix20(*), id(*), perm(*)
    end function
    integer(c_int) function pm_set_materials(ctx, numat, e, eload, eloadp, load, fload, gravity, &
                                             gmod) bind(C, name='pm_set_materials')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: numat
      real(c_double) :: e(*), fload(*), gravity(3)
      integer(c_int) :: eload(*), eloadp(*), load(*)
      real(c_double), value :: gmod
    end function
    integer(c_int) function pm_pform(ctx, isw, lhs, rhs, dt, t, target) bind(C, name='pm_pform')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      integer(c_int), value :: isw, lhs, rhs, target
      real(c_double), value :: dt, t
    end function
    integer(c_int) function pm_solve(ctx, cfg, b, x, info) bind(C, name='pm_solve')
      import :: c_ptr, c_int, c_double, pm_solver_cfg, pm_solve_info
      type(c_ptr), value :: ctx
      type(pm_solver_cfg) :: cfg
      real(c_double) :: b(*), x(*)
      type(pm_solve_info) :: info
    end function
    integer(c_int) function pm_struct_upload(ctx, neq, nad, ia, ja) bind(C, name='pm_struct_upload')
      import :: c_ptr, c_int, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int), value :: neq
      integer(c_int64_t), value :: nad
      integer(c_int64_t) :: ia(*)
      integer(c_int) :: ja(*)
    end function
    integer(c_int) function pm_values_upload(ctx, ad, al) bind(C, name='pm_values_upload')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: ad(*), al(*)
    end function
    integer(c_int) function pm_set_state(ctx, u, u0, dp, poro, t) bind(C, name='pm_set_state')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: u(*), u0(*), dp(*), poro(*)
      real(c_double), value :: t
    end function
    integer(c_int) function pm_get_state(ctx, u, u0, dp, poro, t) bind(C, name='pm_get_state')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: u(*), u0(*), dp(*), poro(*)
      real(c_double) :: t
    end function
    integer(c_int) function pm_step(ctx, dt, cfg, info) bind(C, name='pm_step')
      import :: c_ptr, c_int, c_double, pm_solver_cfg, pm_step_info
      type(c_ptr), value :: ctx
      real(c_double), value :: dt
      type(pm_solver_cfg) :: cfg
      type(pm_step_info) :: info
    end function
    integer(c_int) function pm_tform(ctx, tx, txe, txb, flux) bind(C, name='pm_tform')
      import :: c_ptr, c_int, c_double
      type(c_ptr), value :: ctx
      real(c_double) :: tx(*), txe(*), txb(*), flux(*)
    end function
    integer(c_int) function pm_set_vprop(ctx, ie, vprop) bind(C, name='pm_set_vprop')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int) :: ie(*), vprop(4)
    end function
    integer(c_int) function pm_set_ordered(ctx, on) bind(C, name='pm_set_ordered')
      import :: c_ptr, c_int
      type(c_ptr), value :: ctx
      integer(c_int), value :: on
    end function
    integer(c_int) function pm_matrix_info(ctx, neq, nblocks, nnz, ncols) &
        bind(C, name='pm_matrix_info')
      import :: c_ptr, c_int, c_int64_t
      type(c_ptr), value :: ctx
      integer(c_int) :: neq, nblocks
      integer(c_int64_t) :: nnz, ncols
    end function
    type(c_ptr) function pm_error_string(code) bind(C, name='pm_error_string')
      import :: c_ptr, c_int
      integer(c_int), value :: code
    end function
  end interface

contains

  ! solv_pm with its reference argument meaning: CSRC (ip int64, ja, ad, al),
  ! preconditioner m (out), right-hand side b, solution x.  The matrix
  ! pattern is uploaded on the first call with a new ctx.
  subroutine solv_pm_hip(ctx, neq, nad, ip, ja, ad, al, m, b, x, tol, maxit, ngram, solver, &
                         precond, iters, ierr)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: neq, maxit, ngram, solver, precond
    integer(8), intent(in) :: nad
    integer(8), intent(in) :: ip(neq + 1)
    integer, intent(in) :: ja(*)
    real(8), intent(in) :: ad(neq), al(*), b(neq), tol
    real(8), intent(out) :: m(neq), x(neq)
    integer, intent(out) :: iters, ierr
    type(pm_solver_cfg) :: cfg
    type(pm_solve_info) :: info
    integer :: i
    logical, save :: uploaded = .false.
    if (.not. uploaded) then
      ierr = pm_struct_upload(ctx, int(neq, c_int), int(nad, c_int64_t), ip, ja)
      if (ierr /= PM_OK) return
      uploaded = .true.
    end if
    ierr = pm_values_upload(ctx, ad, al)
    if (ierr /= PM_OK) return
    cfg%solver = solver
    cfg%precond = precond
    cfg%maxit = maxit
    cfg%nkrylov = ngram
    cfg%tol = tol
    ierr = pm_solve(ctx, cfg, b, x, info)
    iters = info%iters
    do i = 1, neq
      m(i) = 1.0d0 / ad(i)
      if (precond == 5) m(i) = abs(m(i))
    end do
  end subroutine

  ! pform_pm for elmt17_pm on the device matrix: isw 2 (lhs/rhs) or 4 (rhs)
  subroutine pform_pm_hip(ctx, isw, lhs, rhs, dt, t, to_b, ierr)
    type(c_ptr), intent(in) :: ctx
    integer, intent(in) :: isw
    logical, intent(in) :: lhs, rhs, to_b
    real(8), intent(in) :: dt, t
    integer, intent(out) :: ierr
    ierr = pm_pform(ctx, int(isw, c_int), merge(1_c_int, 0_c_int, lhs), &
                    merge(1_c_int, 0_c_int, rhs), dt, t, merge(1_c_int, 0_c_int, to_b))
  end subroutine

  ! one solv macro step
  subroutine step_pm_hip(ctx, dt, cfg, info, ierr)
    type(c_ptr), intent(in) :: ctx
    real(8), intent(in) :: dt
    type(pm_solver_cfg), intent(inout) :: cfg
    type(pm_step_info), intent(out) :: info
    integer, intent(out) :: ierr
    ierr = pm_step(ctx, dt, cfg, info)
  end subroutine

end module pmhip
