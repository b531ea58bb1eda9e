! pm_mefdrive.f90 -- a Fortran host that runs the mef Biot u-p path on the
! MI355X through the pmhip module.
!
!   pm_mefdrive <deck> <prefix> [maxsteps] [fp32ir] [parse | vtkcheck]
!
! is the reference's `mef <input> <output>` (Main.f:336-402) for the macros
! module pmdeck reads: the deck's mesh block sets up the structure, the
! materials and the loads; its macro program then runs in order -- dt, solv
! (one backward-Euler step, Main.f:876-1123), pnup and pnsf (the printnode
! histories of the setpnode nodes, Main.f:2310-2481) and pres (the results
! file, Main.f:1322-1545).  maxsteps > 0 ends the program before its
! (maxsteps+1)-th solv; fp32ir solves with FP32 matrix values inside FP64
! iterative refinement.  Two modes make no device calls: parse writes what
! the reader read to <prefix>_deck.bin; vtkcheck writes results files of
! synthetic fields on the deck's mesh.  Files written, named as the
! reference's Filenames.f names them:
!   <prefix>_up_node_<n>.txt          u v w p per pnup, (i9,10e18.10)
!   <prefix>_stress_node_<n>.txt      total stress per pnsf
!   <prefix>_stressE_node_<n>.txt     Terzaghi effective stress
!   <prefix>_stressB_node_<n>.txt     Biot effective stress
!   <prefix>_flux_node_<n>.txt        Darcy flux
!   <prefix>_poro_node_<n>.txt        porosity
!   <prefix>_istep_<n>.vtu            per pres (.vtk with legacy_vtk)
!   <prefix>_log_1.txt                one line per linear solve
!   <prefix>_log_nonlinear_1.txt      step, time and nonlinear iterations
!   <prefix>_t_1.txt                  device times and system sizes
!
!   pm_mefdrive <file>.bin <prefix>
!
! runs a problem given as a stream file (written by poromec_amd.fdeck) for
! `nsteps` steps from rest, appending the history of node `pnode` to
! <prefix>_up_node.txt and writing the final u (4, nnode) and dp (nnode) to
! <prefix>_state.bin.  The stream holds, in order,
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
  use pmdeck
  use pmvtk
  implicit none
  character(len=1024) :: a1, a2, a3
  integer :: n1, maxsteps, ios, i
  logical :: fp32ir, parse, vcheck

  call get_command_argument(1, a1, n1)
  call get_command_argument(2, a2)
  if (n1 == 0 .or. len_trim(a2) == 0) then
    write(*, '(a)') ' usage: pm_mefdrive <deck> <prefix> [maxsteps] [fp32ir] [parse]'
    write(*, '(a)') '        pm_mefdrive <file>.bin <prefix>'
    stop 2
  end if
  if (n1 > 4) then
    if (a1(n1 - 3:n1) == '.bin') then
      call run_bin(a1, a2)
      stop
    end if
  end if
  maxsteps = 0
  fp32ir = .false.
  parse = .false.
  vcheck = .false.
  do i = 3, command_argument_count()
    call get_command_argument(i, a3)
    select case (trim(a3))
    case ('fp32ir')
      fp32ir = .true.
    case ('parse')
      parse = .true.
    case ('vtkcheck')
      vcheck = .true.
    case default
      read(a3, *, iostat=ios) maxsteps
      if (ios /= 0) then
        write(*, '(a)') ' *** unknown argument ' // trim(a3)
        stop 2
      end if
    end select
  end do
  if (parse) then
    call dump_deck(a1, a2)
  else if (vcheck) then
    call vtk_check(a1, a2)
  else
    call run_deck(a1, a2, maxsteps, fp32ir)
  end if

contains

  subroutine check(code, what)
    integer(c_int), intent(in) :: code
    character(len=*), intent(in) :: what
    if (code /= PM_OK) then
      write(*, '(a,a,a,i4)') ' *** ', what, ' failed: code ', code
      stop 1
    end if
  end subroutine

  function itoa(i) result(s)
    integer, intent(in) :: i
    character(len=:), allocatable :: s
    character(len=16) :: b
    write(b, '(i0)') i
    s = trim(b)
  end function

  function solver_name(id) result(s)
    integer(c_int), intent(in) :: id
    character(len=:), allocatable :: s
    select case (id)
    case (1)
      s = 'PCG'
    case (2)
      s = 'GMRES'
    case (4)
      s = 'BICGSTAB'
    case (7)
      s = 'MINRES'
    case (11)
      s = 'SQMR'
    case default
      s = ''
    end select
  end function

  ! one printnode line per print node of one history kind; the first write
  ! of a kind replaces the file and writes its header (printnode,
  ! Pdata.f:733-790)
  subroutine print_nodes(prefix, stem, hname, first, pnodes, istep, t, v, ld, nv)
    character(len=*), intent(in) :: prefix, stem, hname
    logical, intent(in) :: first
    integer, intent(in) :: pnodes(:), istep, ld, nv
    real(c_double), intent(in) :: t, v(*)
    integer :: i, k, u, n
    character(len=:), allocatable :: fn
    do i = 1, size(pnodes)
      n = pnodes(i)
      fn = trim(prefix) // '_' // stem // '_node_' // itoa(n) // '.txt'
      if (first) then
        open(newunit=u, file=fn, status='replace', action='write')
        write(u, '(a,a,a,i9)') '# Variacao ', hname, ' no tempo no ponto', n
      else
        open(newunit=u, file=fn, status='old', position='append', action='write')
      end if
      write(u, '(i9,10e18.10)') istep, t, (v(ld * (n - 1) + k), k = 1, nv)
      close(u)
    end do
  end subroutine

  ! what the deck reader made of a deck, as a stream file <prefix>_deck.bin
  ! (no device calls): numel, nnodev, numat, reorder, vprop(4), solver,
  ! precond, maxit, ngram, maxnlit, stop_crit, maxnlit_accept, npnodes,
  ! nprog (int32); tol, nltol, gravity(3), gmod (real64); pnodes, op (int32);
  ! opval (real64); ix, idv, ie, eload, eloadp, load (int32); x, e, u0,
  ! fload (real64); print_flag(13), legacy_vtk, bvtk, nen (int32)
  subroutine dump_deck(path, prefix)
    character(len=*), intent(in) :: path, prefix
    type(deck_t) :: d
    integer :: u
    call read_deck(path, d)
    open(newunit=u, file=trim(prefix) // '_deck.bin', access='stream', form='unformatted', &
         status='replace', action='write')
    write(u) d%numel, d%nnodev, d%numat, merge(1_c_int, 0_c_int, d%reorder), d%vprop, d%solver, &
      d%precond, d%maxit, d%ngram, d%maxnlit, d%stop_crit, d%maxnlit_accept, &
      int(size(d%pnodes), c_int), int(d%nprog, c_int)
    write(u) d%tol, d%nltol, d%gravity, d%gmod
    write(u) int(d%pnodes, c_int), int(d%op(1:d%nprog), c_int)
    write(u) d%opval(1:d%nprog)
    write(u) d%ix, d%idv, d%ie, d%eload, d%eloadp, d%load
    write(u) d%x, d%e, d%u0, d%fload
    write(u) merge(1_c_int, 0_c_int, d%fprint), merge(1_c_int, 0_c_int, d%vtk_legacy), &
      merge(1_c_int, 0_c_int, d%vtk_binary), d%nen
    close(u)
    write(*, '(a,i10,a,i10,a,i8)') ' parsed: numel', d%numel, ' nnodev', d%nnodev, ' ops', d%nprog
  end subroutine

  ! the results writer on a deck's mesh with exactly representable synthetic
  ! fields (no device calls): <prefix>_{a,b,x}{l,q}_istep_7.vtk|vtu for legacy
  ! ASCII, legacy binary and VTU, over the vertices (l) and every node (q)
  subroutine vtk_check(path, prefix)
    character(len=*), intent(in) :: path, prefix
    type(deck_t) :: d
    integer(c_int), allocatable :: ix20(:), idn(:), perm(:)
    real(c_double), allocatable :: f(:,:)
    integer(c_int) :: nmax, nnode, neq, ierr
    integer(c_int64_t) :: nad
    integer :: i, k, iq, ie, code
    logical :: fl(13), leg, bin
    character(len=1) :: enc(3) = ['a', 'b', 'x'], qs(2) = ['l', 'q']
    call read_deck(path, d)
    nmax = d%nnodev + 12 * d%numel
    allocate(ix20(21 * d%numel), idn(4 * nmax), perm(nmax))
    ierr = pm_mesh_numbering(d%numel, d%nnodev, d%ix, d%idv, merge(1_c_int, 0_c_int, d%reorder), &
                             nmax, ix20, idn, perm, nnode, neq, nad)
    call check(ierr, 'pm_mesh_numbering')
    allocate(f(27, nnode))
    do i = 1, nnode
      do k = 1, 27
        f(k, i) = real(mod(i * 37 + k * 101, 2001) - 1000, c_double) * 2.0d0**(2 * k - 20)
      end do
    end do
    do iq = 1, 2
      fl = .true.
      fl(1) = iq == 2
      fl(10:13) = .false.
      do ie = 1, 3
        leg = ie /= 3
        bin = ie == 2
        call write_res(res_name(trim(prefix) // '_' // enc(ie) // qs(iq), 7, leg), ix20, d%x, &
                       d%nnodev, f(1:4, :), f(5, :), f(6, :), f(7:12, :), f(13:18, :), &
                       f(19:24, :), f(25:27, :), nnode, d%numel, 7, 12345.678d0, fl, leg, bin, code)
        if (code /= 0) stop 3
      end do
    end do
    write(*, '(a,i10)') ' vtk files written, nnode', nnode
  end subroutine

  subroutine run_deck(path, prefix, maxsteps, fp32ir)
    character(len=*), intent(in) :: path, prefix
    integer, intent(in) :: maxsteps
    logical, intent(in) :: fp32ir
    type(deck_t) :: d
    type(c_ptr) :: ctx
    type(pm_solver_cfg) :: cfg
    type(pm_step_info) :: info
    integer(c_int) :: nnode, neq, nb, reorder, ierr
    integer(c_int64_t) :: nad, nnz, ncols
    real(c_double), allocatable :: u(:), u0(:), dp(:), poro(:), tx(:), txe(:), txb(:), flux(:)
    integer(c_int), allocatable :: ix20(:), idn(:), perm(:), ix9(:,:)
    logical :: fres(13)
    real(c_double) :: dt, t, tnow, ms_elem, ms_solve, ms_total
    integer :: iop, istep, k, flog, fnl, ftm, nsolv
    integer(8) :: c0, c1, crate
    logical :: first_up, first_sf
    character(len=:), allocatable :: nm
    character(len=12) :: key12
    character(len=10) :: key10

    call system_clock(c0, crate)
    call read_deck(path, d)
    reorder = merge(1_c_int, 0_c_int, d%reorder)
    ierr = pm_ctx_create(ctx, 0_c_int)
    call check(ierr, 'pm_ctx_create')
    if (d%nen == 8) then  ! maxno 8: ix (9, numel), the Q1/Q1 element on the mesh as read
      allocate(ix9(9, d%numel))
      ix9(1:8, :) = d%ix(1:8, :)
      ix9(9, :) = d%ix(21, :)
      ierr = pm_mesh_setup_nen(ctx, 8_c_int, d%numel, d%nnodev, ix9, d%x, d%idv, reorder, &
                               nnode, neq, nad)
      call check(ierr, 'pm_mesh_setup_nen')
    else
      ierr = pm_mesh_setup(ctx, d%numel, d%nnodev, d%ix, d%x, d%idv, reorder, nnode, neq, nad)
      call check(ierr, 'pm_mesh_setup')
    end if
    ierr = pm_set_materials(ctx, d%numat, d%e, d%eload, d%eloadp, d%load, d%fload, d%gravity, &
                            d%gmod)
    call check(ierr, 'pm_set_materials')
    if (d%vprop(1) /= 0) then
      ierr = pm_set_vprop(ctx, d%ie, d%vprop)
      call check(ierr, 'pm_set_vprop')
    end if
    write(*, '(a,i10,a,i10,a,i14)') ' nnode', nnode, ' neq', neq, ' nad', nad
    do k = 1, size(d%pnodes)
      if (d%pnodes(k) < 1 .or. d%pnodes(k) > nnode) then
        write(*, '(a,i10,a,i10)') ' *** print node', d%pnodes(k), ' outside 1 ..', nnode
        stop 2
      end if
    end do

    ! the vertex values the deck gives (initialpres, hydrostatic) are u and
    ! u0 of the first step; the mid-edge nodes start at zero
    allocate(u(4 * nnode), u0(4 * nnode), dp(nnode), poro(nnode))
    ierr = pm_get_state(ctx, u, u0, dp, poro, t)
    call check(ierr, 'pm_get_state')
    u = 0.0d0
    u(1:4 * d%nnodev) = reshape(d%u0, [4 * d%nnodev])
    u0 = u
    ierr = pm_set_state(ctx, u, u0, dp, poro, 0.0d0)
    call check(ierr, 'pm_set_state')

    cfg%solver = d%solver
    cfg%precond = d%precond
    cfg%maxit = d%maxit
    cfg%nkrylov = d%ngram
    cfg%tol = d%tol
    cfg%precision = merge(PM_PREC_FP32IR, PM_PREC_FP64, fp32ir)
    cfg%maxnlit = d%maxnlit
    cfg%nltol = d%nltol
    cfg%stop_crit = d%stop_crit
    cfg%maxnlit_accept = d%maxnlit_accept

    dt = 1.0d0
    t = 0.0d0
    istep = 0
    nsolv = 0
    first_up = .true.
    first_sf = .true.
    ms_elem = 0.0d0
    ms_solve = 0.0d0
    ms_total = 0.0d0
    open(newunit=flog, file=trim(prefix) // '_log_1.txt', status='replace', action='write')
    open(newunit=fnl, file=trim(prefix) // '_log_nonlinear_1.txt', status='replace', &
         action='write')
    write(flog, '(a)') 'Solver control flop.'
    write(fnl, '(a)') 'Non-linear control flop.'
    do iop = 1, d%nprog
      select case (d%op(iop))
      case (OP_DT)
        dt = d%opval(iop)
      case (OP_SOLV)
        if (maxsteps > 0 .and. nsolv == maxsteps) exit
        nsolv = nsolv + 1
        istep = istep + 1
        t = t + dt
        write(fnl, '(a,i9,e13.6)') 'step', istep, t
        call step_pm_hip(ctx, dt, cfg, info, ierr)
        call check(ierr, 'pm_step')
        ms_elem = ms_elem + info%ms_element
        ms_solve = ms_solve + info%ms_solve
        ms_total = ms_total + info%ms_total
        do k = 1, min(info%nsolve, 16)
          nm = solver_name(info%used(k))
          if (len(nm) == 0) nm = solver_name(cfg%solver)
          write(flog, '(a,a,i10,a,e17.10,a,e17.10)') nm, ':  it', info%iters(k), &
            ' true rel resid ', info%true_rel(k), ' tol ', cfg%tol
        end do
        write(fnl, '(a,i7)') 'nlit', info%nlit
        write(*, '(a,i6,a,e13.6,a,i4,a,i8)') ' step', istep, ' t', t, ' nlit', info%nlit, &
          ' iters', info%iters(1)
      case (OP_PNUP)
        if (size(d%pnodes) == 0) then
          write(*, '(a)') ' *** No node to print for PNUP!'
          stop 2
        end if
        ierr = pm_get_state(ctx, u, u0, dp, poro, tnow)
        call check(ierr, 'pm_get_state')
        call print_nodes(prefix, 'up', 'DeslocAndPress', first_up, d%pnodes, istep, t, u, 4, 4)
        first_up = .false.
      case (OP_PNSF)
        if (size(d%pnodes) == 0) then
          write(*, '(a)') ' *** No node to print for PNSF!'
          stop 2
        end if
        if (.not. allocated(tx)) allocate(tx(6 * nnode), txe(6 * nnode), txb(6 * nnode), &
                                          flux(3 * nnode))
        ierr = pm_tform(ctx, tx, txe, txb, flux)
        call check(ierr, 'pm_tform')
        ierr = pm_get_state(ctx, u, u0, dp, poro, tnow)
        call check(ierr, 'pm_get_state')
        call print_nodes(prefix, 'stress', 'totalStress', first_sf, d%pnodes, istep, t, tx, 6, 6)
        call print_nodes(prefix, 'stressE', 'eSstress', first_sf, d%pnodes, istep, t, txe, 6, 6)
        call print_nodes(prefix, 'stressB', 'biotStress', first_sf, d%pnodes, istep, t, txb, 6, 6)
        call print_nodes(prefix, 'flux', 'darcyFlux', first_sf, d%pnodes, istep, t, flux, 3, 3)
        call print_nodes(prefix, 'poro', 'poro', first_sf, d%pnodes, istep, t, poro, 1, 1)
        first_sf = .false.
      case (OP_PRES)
        if (.not. allocated(tx)) allocate(tx(6 * nnode), txe(6 * nnode), txb(6 * nnode), &
                                          flux(3 * nnode))
        if (any(d%fprint(5:8))) then
          ierr = pm_tform(ctx, tx, txe, txb, flux)
          call check(ierr, 'pm_tform')
        else
          tx = 0.0d0
          txe = 0.0d0
          txb = 0.0d0
          flux = 0.0d0
        end if
        ierr = pm_get_state(ctx, u, u0, dp, poro, tnow)
        call check(ierr, 'pm_get_state')
        if (.not. allocated(ix20)) then
          allocate(ix20(21 * d%numel), idn(4 * nnode), perm(nnode))
          ierr = pm_get_numbering(ctx, ix20, idn, perm)
          call check(ierr, 'pm_get_numbering')
        end if
        fres = d%fprint
        if (d%nen == 8) fres(1) = .false.  ! maxno 8: vertices only, no quadratic cells
        call write_res(res_name(prefix, istep, d%vtk_legacy), ix20, d%x, d%nnodev, u, dp, poro, &
                       tx, txb, txe, flux, nnode, d%numel, istep, t, fres, d%vtk_legacy, &
                       d%vtk_binary, k)
        if (k /= 0) then
          write(*, '(a,i3)') ' *** pres: results file not written, code', k
          stop 2
        end if
      end select
    end do
    close(flog)
    close(fnl)

    ierr = pm_matrix_info(ctx, neq, nb, nnz, ncols)
    call check(ierr, 'pm_matrix_info')
    call system_clock(c1)
    open(newunit=ftm, file=trim(prefix) // '_t_1.txt', status='replace', action='write')
    write(ftm, '(a)') '# Arquivo de log do poro mecanico (MI355X)'
    write(ftm, '(a)') 'Tempos (seg):'
    key12 = 'ELMT'
    write(ftm, '(a,f12.6)') key12, ms_elem / 1.0d3
    key12 = 'SOLVR'
    write(ftm, '(a,f12.6)') key12, ms_solve / 1.0d3
    key12 = 'STEPS'
    write(ftm, '(a,f12.6)') key12, ms_total / 1.0d3
    key12 = 'TOTAL'
    write(ftm, '(a,f12.6)') key12, real(c1 - c0, c_double) / real(crate, c_double)
    write(ftm, '(a)') 'Malha e sistema linear:'
    key10 = 'neq'
    write(ftm, '(a,i12)') key10, neq
    key10 = 'nad'
    write(ftm, '(a,i12)') key10, nad
    key10 = 'nnz'
    write(ftm, '(a,i12)') key10, nnz
    key10 = 'nnode'
    write(ftm, '(a,i12)') key10, nnode
    key10 = 'numel'
    write(ftm, '(a,i12)') key10, d%numel
    close(ftm)
    call pm_ctx_destroy(ctx)
    write(*, '(a,i8,a,i10,a,i10,a,i14)') ' done:', nsolv, ' steps, nnode', nnode, ' neq', neq, &
      ' nad', nad
  end subroutine

  subroutine run_bin(deck, outp)
    character(len=*), intent(in) :: deck, outp
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
  end subroutine
end program pm_mefdrive
