! pm_deck.f90 -- reads the reference's input decks (the part the mef Biot u-p
! path runs) into the arrays the pmhip calls take, and the macro program.
!
! The reference reads a deck with readmacro (Rdata.f:1945-2001): a line is
! split into blank-separated words, a line's first word selects a macro, data
! lines are read word by word.  The top level (Main.f:405-470) runs macros in
! order; loop n ... next repeats the macros between them; the mesh macro reads
! the mesh block (rdat_pm, Rdata.f:1-1037).  Read here:
!   top level: config, solver, gravity, nl, conseq, dt, mesh, setpnode, pnup,
!              pnsf, solv, loop/next, stop;
!   mesh block: the parameter line, materials, fmaterials, hexa8, coordinates,
!              constrainpmec, elmtloads, elmtpresloads, loads (types 1-3, 6-8
!              inline, 39-43 from a time-table file), initialpres,
!              hydrostatic, insert/return, end.
! The earlier forms of the decks the reference ships with its outputs are read
! too: config and gravity inline, pcg maxit tol, maxnlit n (stop criterion 1,
! the step kept at maxnlit), inline type 39-43 loads (constant in time),
! element type 7 for 17, and end <name> of a block this reader does not know
! closing that block rather than the mesh.  Post-processing macros (pgeo,
! pres, deltatc, ...) are skipped; macros that change the system this path
! solves stop the program, as stop_mef does.
module pmdeck
  use iso_c_binding
  implicit none
  private
  public :: deck_t, read_deck, OP_SOLV, OP_PNUP, OP_PNSF, OP_DT, OP_PRES

  integer, parameter :: MAXW = 96, LLEN = 2048, MAXINC = 16
  integer, parameter :: OP_SOLV = 1, OP_PNUP = 2, OP_PNSF = 3, OP_DT = 4, OP_PRES = 5
  ! setprint words (set_print_vtk_pm), in print_flag order
  character(len=14), parameter :: PRINT_WORDS(13) = [character(len=14) :: 'quadratic', &
    'desloc', 'pressure', 'dpressure', 'totalstress', 'biotstress', 'terzaghistress', &
    'darcyflux', 'porosity', 'pconsolidation', 'elplastic', 'temperature', 'heatFlux']
  real(c_double), parameter :: EPS = 2.0d0**(-53)

  type deck_t
    integer(c_int) :: numel = 0, nnodev = 0, numat = 0
    ! maxno: 20 (hexa8 promoted to the hexa20/8 element) or 8 (the mesh as
    ! read, the equal-order Q1/Q1 element, type 7)
    integer(c_int) :: nen = 20
    integer(c_int), allocatable :: ix(:,:), idv(:,:), ie(:), eload(:,:), eloadp(:,:)
    real(c_double), allocatable :: x(:,:), e(:,:), u0(:,:)
    integer(c_int) :: load(3, 200)
    real(c_double) :: fload(1000, 5, 200)
    real(c_double) :: gravity(3), gmod
    logical :: reorder = .true.
    integer(c_int) :: vprop(4) = 0
    integer(c_int) :: solver = 11, precond = 2, maxit = 50000, ngram = 150
    integer(c_int) :: maxnlit = 500, stop_crit = 2, maxnlit_accept = 0
    real(c_double) :: tol = 1.0d-11, nltol = 1.0d-4
    integer, allocatable :: pnodes(:)
    ! the pres macro: print_flag (Main.f:220-232 defaults), legacy_vtk, bvtk
    logical :: fprint(13) = [.false., .true., .true., .false., .false., .false., .false., &
                             .false., .false., .false., .false., .false., .false.]
    logical :: vtk_legacy = .false., vtk_binary = .false.
    integer :: nprog = 0
    integer, allocatable :: op(:)
    real(c_double), allocatable :: opval(:)
  end type

  ! the line source: the main file and the files it inserts
  integer :: nstack = 0
  integer :: units(MAXINC)
  character(len=1024) :: dirs(MAXINC)
  character(len=LLEN) :: line
  character(len=256) :: words(MAXW)
  integer :: nw = 0

contains

  subroutine fail(msg)
    character(len=*), intent(in) :: msg
    write(*, '(a,a)') ' *** deck: ', trim(msg)
    stop 2
  end subroutine

  function dirname(path) result(d)
    character(len=*), intent(in) :: path
    character(len=1024) :: d
    integer :: k
    k = index(path, '/', back=.true.)
    d = ''
    if (k > 0) d = path(1:k)
  end function

  ! a file named in the deck, relative to the including file's folder
  function resolve(name) result(p)
    character(len=*), intent(in) :: name
    character(len=1024) :: p
    if (name(1:1) == '/' .or. nstack == 0) then
      p = name
    else
      p = trim(dirs(nstack)) // trim(name)
    end if
  end function

  subroutine push(path)
    character(len=*), intent(in) :: path
    integer :: u, ios
    character(len=1024) :: p
    if (nstack >= MAXINC) call fail('insert nested too deep')
    p = resolve(path)
    open(newunit=u, file=trim(p), status='old', action='read', iostat=ios)
    if (ios /= 0) call fail('File ' // trim(path) // ' not found !')
    nstack = nstack + 1
    units(nstack) = u
    dirs(nstack) = dirname(p)
  end subroutine

  subroutine pop()
    if (nstack > 1) then
      close(units(nstack))
      nstack = nstack - 1
    end if
  end subroutine

  ! next line split into words; .false. at the end of the main file
  logical function next_line()
    integer :: ios, i, k, n
    next_line = .false.
    words(1:nw) = ' '
    nw = 0
    do
      read(units(nstack), '(a)', iostat=ios) line
      if (ios /= 0) then
        if (nstack == 1) return
        close(units(nstack))
        nstack = nstack - 1
        cycle
      end if
      exit
    end do
    next_line = .true.
    do i = 1, len(line)
      if (line(i:i) == char(13) .or. line(i:i) == char(9)) line(i:i) = ' '
    end do
    nw = 0
    n = len_trim(line)
    i = 1
    do while (i <= n)
      if (line(i:i) == ' ') then
        i = i + 1
        cycle
      end if
      k = i
      do while (k <= n)
        if (line(k:k) == ' ') exit
        k = k + 1
      end do
      if (nw < MAXW) then
        nw = nw + 1
        words(nw) = line(i:k - 1)
      end if
      i = k
    end do
  end function

  ! next non-empty line
  subroutine data_line(what)
    character(len=*), intent(in) :: what
    do
      if (.not. next_line()) call fail(what)
      if (nw > 0) return
    end do
  end subroutine

  integer function inum(w)
    character(len=*), intent(in) :: w
    integer :: ios
    read(w, *, iostat=ios) inum
    if (ios /= 0) call fail('not an integer: ' // trim(w))
  end function

  real(c_double) function fnum(w)
    character(len=*), intent(in) :: w
    integer :: ios
    read(w, *, iostat=ios) fnum
    if (ios /= 0) call fail('not a number: ' // trim(w))
  end function

  logical function lflag(w)
    character(len=*), intent(in) :: w
    select case (trim(w))
    case ('true', '.true.', 'T', 't')
      lflag = .true.
    case ('false', '.false.', 'F', 'f')
      lflag = .false.
    case default
      lflag = .false.
      call fail('not a logical: ' // trim(w))
    end select
  end function

  ! every blank-separated word of a file
  subroutine file_words(path, ws, n)
    character(len=*), intent(in) :: path
    character(len=256), allocatable, intent(out) :: ws(:)
    integer, intent(out) :: n
    integer :: u, ios, i, k, m
    character(len=LLEN) :: l
    character(len=256), allocatable :: tmp(:)
    open(newunit=u, file=trim(resolve(path)), status='old', action='read', iostat=ios)
    if (ios /= 0) call fail('File ' // trim(path) // ' not found !')
    allocate(ws(64))
    n = 0
    do
      read(u, '(a)', iostat=ios) l
      if (ios /= 0) exit
      do i = 1, len(l)
        if (l(i:i) == char(13) .or. l(i:i) == char(9)) l(i:i) = ' '
      end do
      m = len_trim(l)
      i = 1
      do while (i <= m)
        if (l(i:i) == ' ') then
          i = i + 1
          cycle
        end if
        k = i
        do while (k <= m)
          if (l(k:k) == ' ') exit
          k = k + 1
        end do
        if (n == size(ws)) then
          allocate(tmp(2 * n))
          tmp(1:n) = ws(1:n)
          call move_alloc(tmp, ws)
        end if
        n = n + 1
        ws(n) = l(i:k - 1)
        i = k
      end do
    end do
    close(u)
  end subroutine

  subroutine add_op(d, code, val)
    type(deck_t), intent(inout) :: d
    integer, intent(in) :: code
    real(c_double), intent(in) :: val
    integer, allocatable :: t1(:)
    real(c_double), allocatable :: t2(:)
    if (.not. allocated(d%op)) then
      allocate(d%op(64), d%opval(64))
    else if (d%nprog == size(d%op)) then
      allocate(t1(2 * d%nprog), t2(2 * d%nprog))
      t1(1:d%nprog) = d%op
      t2(1:d%nprog) = d%opval
      call move_alloc(t1, d%op)
      call move_alloc(t2, d%opval)
    end if
    d%nprog = d%nprog + 1
    d%op(d%nprog) = code
    d%opval(d%nprog) = val
  end subroutine

  ! ---- top level (Main.f:405-470) ----
  subroutine read_deck(path, d)
    character(len=*), intent(in) :: path
    type(deck_t), intent(inout) :: d
    character(len=8) :: mc
    integer :: loopn, loop0, k, j
    logical :: inloop, have_mesh
    real(c_double) :: g0(3), gm0
    d%gravity = [0.0d0, 0.0d0, -9.81d0]
    d%gmod = sqrt(d%gravity(3)**2)
    d%load = 0
    d%fload = 0.0d0
    nstack = 0
    call push(path)
    inloop = .false.
    have_mesh = .false.
    loopn = 0
    loop0 = 0
    do while (next_line())
      if (nw == 0) cycle
      mc = words(1)(1:8)
      select case (trim(mc))
      case ('loop')
        if (nw < 2) call fail('Error reading macro (LOOP) !')
        loopn = inum(words(2))
        loop0 = d%nprog
        inloop = .true.
      case ('next')
        if (inloop) then   ! the body ran once as read; repeat it loopn - 1 times
          k = d%nprog
          do j = 2, loopn
            call repeat_ops(d, loop0 + 1, k)
          end do
          if (loopn <= 0) d%nprog = loop0
          inloop = .false.
        end if
      case ('solv')
        call add_op(d, OP_SOLV, 0.0d0)
      case ('pnup')
        call add_op(d, OP_PNUP, 0.0d0)
      case ('pnsf')
        call add_op(d, OP_PNSF, 0.0d0)
      case ('pres')
        call add_op(d, OP_PRES, 0.0d0)
      case ('setprint')
        if (nw < 2) call fail('setprint without a file')
        call read_setprint(d, words(2), have_mesh)
      case ('dt')
        if (nw < 2) call fail('Error reading macro (DT) !')
        call add_op(d, OP_DT, fnum(words(2)))
      case ('stop')
        exit
      case ('config')
        call read_config(d)
      case ('gravity')   ! after the mesh it no longer reaches the loads
        g0 = d%gravity
        gm0 = d%gmod
        call read_gravity(d)
        if (have_mesh) then
          d%gravity = g0
          d%gmod = gm0
        end if
      case ('solver')
        if (have_mesh) call fail('This macro can only be used before macro mesh')
        if (nw < 2) call fail('solver without a file')
        call read_solver(d, words(2))
      case ('conseq')
        if (have_mesh) call fail('This macro can only be used before macro mesh')
        if (nw < 2) call fail('conseq without a file')
        call read_conseq(d, words(2))
      case ('pcg', 'cg', 'sqmr', 'minres', 'bicgstab')
        if (nw >= 3) then   ! the earlier inline form: <name> maxit tol
          select case (trim(mc))
          case ('pcg', 'cg')
            d%solver = 1
          case ('sqmr')
            d%solver = 11
          case ('minres')
            d%solver = 7
          case ('bicgstab')
            d%solver = 4
          end select
          d%maxit = inum(words(2))
          d%tol = fnum(words(3))
          if (d%tol == 0.0d0) d%tol = EPS
        end if
      case ('nl')
        if (nw < 4) call fail('Error reading macro (NL) !')
        d%maxnlit = inum(words(2))
        d%nltol = fnum(words(3))
        d%stop_crit = inum(words(4))
      case ('maxnlit')   ! the earlier loop: criterion 1, the step kept at maxnlit
        d%maxnlit = inum(words(2))
        d%stop_crit = 1
        d%maxnlit_accept = 1
      case ('setpnode')
        if (nw < 2) call fail('setpnode without a file')
        call read_pnode(d, words(2))
      case ('mesh')
        call read_mesh(d)
        have_mesh = .true.
        ! an elastic run prints neither consolidation pressure nor plastic flag
        d%fprint(10) = .false.
        d%fprint(11) = .false.
      case ('block_pu', 'hextotet', 'solvt', 'solvm', 'presolv', 'neqs', 'run')
        call fail('macro ' // trim(mc) // ' is outside the mef Biot u-p path')
      case default
        continue   ! unknown words are skipped, as the reference's macro loop does
      end select
    end do
    if (inloop) d%nprog = loop0   ! a loop without next adds nothing
    if (.not. allocated(d%pnodes)) allocate(d%pnodes(0))
    if (d%vtk_binary .and. .not. d%vtk_legacy .and. any(d%op(1:d%nprog) == OP_PRES)) &
      call fail('bvtk without legacy_vtk: the reference has no binary VTU writer')
    if (.not. allocated(d%op)) allocate(d%op(0), d%opval(0))
    do while (nstack > 0)
      close(units(nstack))
      nstack = nstack - 1
    end do
    if (.not. have_mesh) call fail('deck has no mesh macro')
    if (d%stop_crit /= 1 .and. d%stop_crit /= 2) &
      call fail('only nonlinear stop criteria 1 and 2 are in this path')
    if (d%nen == 8 .and. d%vprop(1) /= 0) call fail('maxno 8: the Q1/Q1 element has no conseq vprop')
    ! check_element (Rdata.f:3744-3758)
    if ((d%vprop(1) /= 0) .neqv. any(d%ie == 18)) &
      call fail('Erro: Invalid elements!! (elmt18 goes with conseq vprop, elmt17 without)')
  end subroutine

  subroutine repeat_ops(d, a, b)
    type(deck_t), intent(inout) :: d
    integer, intent(in) :: a, b
    integer :: i, code
    real(c_double) :: val
    do i = a, b   ! by value: add_op may move the arrays
      code = d%op(i)
      val = d%opval(i)
      call add_op(d, code, val)
    end do
  end subroutine

  ! config <file> (read_config, Rdata.f:2929-3130) or the earlier inline block;
  ! reord is the one key this path uses
  subroutine read_config(d)
    type(deck_t), intent(inout) :: d
    character(len=256), allocatable :: ws(:)
    integer :: n, i
    if (nw > 1) then
      call file_words(words(2), ws, n)
    else
      allocate(ws(256))
      n = 0
      do
        call data_line('config block without end')
        if (trim(words(1)) == 'end') exit
        do i = 1, nw
          if (n < size(ws)) then
            n = n + 1
            ws(n) = words(i)
          end if
        end do
      end do
    end if
    i = 1
    do while (i <= n)
      if (trim(ws(i)) == 'end') exit
      select case (trim(ws(i)))
      case ('reord')
        if (i + 1 > n) call fail('config: reord without a value')
        d%reorder = lflag(ws(i + 1))
        i = i + 2
      case ('bvtk')
        if (i + 1 > n) call fail('config: bvtk without a value')
        d%vtk_binary = lflag(ws(i + 1))
        i = i + 2
      case ('legacy_vtk')
        if (i + 1 > n) call fail('config: legacy_vtk without a value')
        d%vtk_legacy = lflag(ws(i + 1))
        i = i + 2
      case ('newton_raphson')
        if (i + 1 <= n) then
          if (lflag(ws(i + 1))) call fail('newton_raphson is outside this path')
        end if
        i = i + 2
      case ('memory', 'omp_elmt', 'omp_solver', 'nth_elmt', 'nth_solver')
        i = i + 2
      case default
        i = i + 1
      end select
    end do
  end subroutine

  ! gravity <file> with gx gy gz |g| (read_gravity, Rdata.f:3543-3581), or
  ! the earlier inline gravity gx gy gz with |g| the vector's norm
  subroutine read_gravity(d)
    type(deck_t), intent(inout) :: d
    character(len=256), allocatable :: ws(:)
    integer :: n
    if (nw >= 4) then
      d%gravity = [fnum(words(2)), fnum(words(3)), fnum(words(4))]
      d%gmod = sqrt(d%gravity(1) * d%gravity(1) + d%gravity(2) * d%gravity(2) &
                    + d%gravity(3) * d%gravity(3))
    else if (nw == 2) then
      call file_words(words(2), ws, n)
      if (n < 4) call fail('Erro na leitura da macro (GRAVITY)')
      d%gravity = [fnum(ws(1)), fnum(ws(2)), fnum(ws(3))]
      d%gmod = fnum(ws(4))
    else
      call fail('Erro na leitura da macro (GRAVITY)')
    end if
  end subroutine

  ! the solver config file (read_solver_config_pm, Rdata.f:3141-3305)
  subroutine read_solver(d, fname)
    type(deck_t), intent(inout) :: d
    character(len=*), intent(in) :: fname
    character(len=256), allocatable :: ws(:)
    integer :: n, i
    call file_words(fname, ws, n)
    i = 1
    do while (i <= n)
      if (trim(ws(i)) == 'end') exit
      if (i + 1 > n) call fail('*** Error reading macro: ' // trim(ws(i)) // ' in macro solver!')
      select case (trim(ws(i)))
      case ('name')
        select case (trim(ws(i + 1)))
        case ('cg', 'pcg')
          d%solver = 1
        case ('gmres')
          d%solver = 2
        case ('bicgs', 'bicgstab')
          d%solver = 4
        case ('minres')
          d%solver = 7
        case ('sqmr')
          d%solver = 11
        case default
          call fail('solver ' // trim(ws(i + 1)) // ' is outside this path')
        end select
      case ('it')
        d%maxit = inum(ws(i + 1))
      case ('tol')
        d%tol = fnum(ws(i + 1))
        if (d%tol == 0.0d0) d%tol = EPS
      case ('precond')
        select case (trim(ws(i + 1)))
        case ('none')
          d%precond = 1
        case ('diag')
          d%precond = 2
        case ('diagm')
          d%precond = 5
        case default
          call fail('preconditioner ' // trim(ws(i + 1)) // ' is outside this path')
        end select
      case ('nkrylov')
        d%ngram = inum(ws(i + 1))
      case ('histlog', 'alfap', 'alfau', 'ctol', 'cmaxit', 'fprint')
        continue
      case default
        i = i + 1
        cycle
      end select
      i = i + 2
    end do
  end subroutine

  ! setprint <file>: the first word of each line until end names a field the
  ! pres macro writes; the file clears every flag first (set_print_vtk_pm,
  ! Rdata.f:3383-3520).  After the mesh the printed nodes are fixed.
  subroutine read_setprint(d, fname, have_mesh)
    type(deck_t), intent(inout) :: d
    character(len=*), intent(in) :: fname
    logical, intent(in) :: have_mesh
    logical :: fl(13)
    integer :: u, ios, k, i, m
    character(len=LLEN) :: l
    character(len=256) :: w
    open(newunit=u, file=trim(resolve(fname)), status='old', action='read', iostat=ios)
    if (ios /= 0) call fail('File ' // trim(fname) // ' not found !')
    fl = .false.
    do
      read(u, '(a)', iostat=ios) l
      if (ios /= 0) exit
      do i = 1, len(l)
        if (l(i:i) == char(13) .or. l(i:i) == char(9)) l(i:i) = ' '
      end do
      l = adjustl(l)
      m = index(l, ' ')
      if (m <= 1) cycle
      w = l(1:m - 1)
      if (trim(w) == 'end') exit
      do k = 1, 13
        if (trim(w) == trim(PRINT_WORDS(k))) fl(k) = .true.
      end do
    end do
    close(u)
    if (have_mesh) then
      if (fl(1) .neqv. d%fprint(1)) &
        call fail('setprint quadratic after the mesh macro (the printed nodes are fixed)')
      if (fl(10) .or. fl(11)) call fail('pconsolidation / elplastic are fields of the plastic path')
    end if
    d%fprint = fl
  end subroutine

  ! conseq <file> (read_constitutive_equation, Rdata.f:2545-2640)
  subroutine read_conseq(d, fname)
    type(deck_t), intent(inout) :: d
    character(len=*), intent(in) :: fname
    character(len=256), allocatable :: ws(:)
    integer :: n, i
    call file_words(fname, ws, n)
    do i = 1, n
      select case (trim(ws(i)))
      case ('end')
        exit
      case ('plastic')
        call fail('plasticity (conseq plastic) is outside the mef Biot u-p path')
      case ('vprop')
        d%vprop(1) = 1
      case ('kozeny_carman')
        d%vprop(2) = 1
      case ('density')
        d%vprop(3) = 1
      case ('hashin_shtrikman')
        d%vprop(4) = 1
      end select
    end do
    if (any(d%vprop(2:4) /= 0)) d%vprop(1) = 1
  end subroutine

  ! setpnode file: the count, then the nodes (readpnode, Rdata.f:3916-3955)
  subroutine read_pnode(d, fname)
    type(deck_t), intent(inout) :: d
    character(len=*), intent(in) :: fname
    character(len=256), allocatable :: ws(:)
    integer :: n, k, i
    call file_words(fname, ws, n)
    if (n < 1) call fail('setpnode: empty file')
    k = inum(ws(1))
    if (k > 50) call fail('number max print node is 50 !!!')
    k = max(0, min(k, n - 1))
    if (allocated(d%pnodes)) deallocate(d%pnodes)
    allocate(d%pnodes(k))
    do i = 1, k
      d%pnodes(i) = inum(ws(1 + i))
    end do
  end subroutine

  ! ---- the mesh block (rdat_pm, Rdata.f:1-1037) ----
  subroutine read_mesh(d)
    type(deck_t), intent(inout) :: d
    integer :: nnode, numel, numat, maxno, ndf, ndim, i, j, k
    logical :: have_el, have_x
    character(len=15) :: mc
    ! the parameter line (parameters, Rdata.f:1863-1943)
    call data_line('mesh without parameters')
    nnode = -1; numel = -1; numat = -1; maxno = -1; ndf = -1; ndim = -1
    i = 1
    do while (i < nw)
      select case (trim(words(i)(1:8)))
      case ('nnode')
        nnode = inum(words(i + 1))
      case ('numel')
        numel = inum(words(i + 1))
      case ('numat')
        numat = inum(words(i + 1))
      case ('maxno')
        maxno = inum(words(i + 1))
      case ('ndf')
        ndf = inum(words(i + 1))
      case ('dim')
        ndim = inum(words(i + 1))
      case default
        i = i + 1
        cycle
      end select
      i = i + 2
    end do
    if (min(nnode, numel, numat, maxno, ndf, ndim) < 0) call fail('Missing macro in parameters')
    if (ndf /= 4 .or. ndim /= 3) call fail('this path is the 3-D u-p system (ndf 4, dim 3)')
    if (maxno /= 20 .and. maxno /= 8) &
      call fail('this path is the hexa20/8 element (maxno 20) or the Q1/Q1 element (maxno 8)')
    d%nen = maxno
    d%numel = numel
    d%nnodev = nnode
    d%numat = numat
    allocate(d%ix(21, numel), d%idv(4, nnode), d%ie(numat), d%eload(7, numel), &
             d%eloadp(7, numel), d%x(3, nnode), d%e(15, numat), d%u0(4, nnode))
    d%ix = 0; d%idv = 0; d%ie = 0; d%eload = 0; d%eloadp = 0
    d%x = 0.0d0; d%e = 0.0d0; d%u0 = 0.0d0
    have_el = .false.
    have_x = .false.
    do
      call data_line('mesh block without end')
      mc = words(1)(1:15)
      select case (trim(mc))
      case ('end')
        if (nw == 1) exit
        if (trim(words(2)) == 'mesh') exit
        cycle   ! closes a block this reader does not know
      case ('materials')
        call read_materials(d)
      case ('fmaterials')
        call read_fmaterials(d)
      case ('hexa8')
        do k = 1, numel
          call data_line('*** Erro na leitura dos elementos !')
          if (nw < 10) call fail('*** Erro na leitura dos elementos !')
          j = inum(words(1))
          if (j < 1 .or. j > numel) call fail('*** Erro na leitura dos elementos !')
          do i = 1, 8
            d%ix(i, j) = inum(words(1 + i))
          end do
          d%ix(21, j) = inum(words(10))
        end do
        call data_line('*** Erro na leitura dos elementos !')
        if (trim(words(1)) /= 'end') call fail('*** Erro na leitura dos elementos !')
        have_el = .true.
      case ('coordinates')
        do k = 1, nnode
          call data_line('*** Erro na leitura das coordenadas !')
          if (nw < 4) call fail('*** Erro na leitura das coordenadas !')
          j = inum(words(1))
          if (j < 1 .or. j > nnode) call fail('*** Erro na leitura das coordenadas !')
          d%x(:, j) = [fnum(words(2)), fnum(words(3)), fnum(words(4))]
        end do
        call data_line('*** Erro na leitura das coordenadas !')
        if (trim(words(1)) /= 'end') call fail('*** Erro na leitura das coordenadas !')
        have_x = .true.
      case ('constrainpmec')
        if (.not. have_el) call fail('MACRO: constrainpmec !! Unread Elements')
        call read_bound(d%idv, nnode, 4)
      case ('elmtloads')
        if (.not. have_el) call fail('MACRO: elmtloads !! Unread Elements')
        call read_bound(d%eload, numel, 7)
      case ('elmtpresloads')
        if (.not. have_el) call fail('MACRO: elmtpresloads !! Unread Elements')
        call read_bound(d%eloadp, numel, 7)
      case ('initialpres')
        if (.not. have_el) call fail('MACRO: initialpres !! Unread Elements')
        do
          call data_line('*** Erro na leitura das forcas nodais !')
          if (trim(words(1)) == 'end') exit
          j = inum(words(1))
          if (j < 1 .or. j > nnode .or. nw < 2) call fail('*** Erro na leitura das forcas nodais !')
          d%u0(4, j) = fnum(words(2))
        end do
      case ('loads')
        call read_loads(d)
      case ('hydrostatic')
        if (.not. have_x) call fail('hydrostatic before coordinates')
        call read_hydrostatic(d, words(2))
      case ('initialdisp')   ! read and dropped by the reference (Rdata.f:1600-1607)
        do
          call data_line('initialdisp without end')
          if (trim(words(1)) == 'end') exit
        end do
      case ('insert')
        if (nw < 2) call fail('insert without a file name')
        call push(words(2))
      case ('return')
        call pop()
      case ('hexa20', 'bar2', 'tria3', 'quad4', 'tetra4', 'tetra10', 'constraindisp', &
            'nodalforces', 'nodalloads', 'constraintemp', 'nodalthermloads', 'hydrostress', &
            'initialstress', 'parallel', 'tria3ov', 'quad4ov', 'tetra4ov', 'hexa8ov', &
            'tetra10ov', 'hexa20ov')
        call fail('mesh macro ' // trim(mc) // ' is outside the mef Biot u-p path')
      case default
        continue
      end select
    end do
    if (.not. (have_el .and. have_x)) call fail('mesh block without hexa8 elements or coordinates')
    if (minval(d%ix(21, :)) < 1 .or. maxval(d%ix(21, :)) > numat) &
      call fail('element material id out of range')
    if (d%nen == 20) then
      where (d%ie == 7) d%ie = 17
    else if (any(d%ie /= 7)) then
      call fail('maxno 8: every material must be element type 7 (Q1/Q1)')
    end if
    ! boundc (Rdata.f:2300): restrained dofs start at their prescribed values
    where (d%idv > 0) d%u0 = 0.0d0
  end subroutine

  ! materials: ma type e1 e2 ... per line until end (mate, Rdata.f:1132-1182)
  subroutine read_materials(d)
    type(deck_t), intent(inout) :: d
    integer :: ma, i
    do
      call data_line('*** Erro na leitura dos materiais !')
      if (trim(words(1)) == 'end') return
      ma = inum(words(1))
      if (ma < 1 .or. ma > d%numat .or. nw < 2 .or. nw - 2 > 15) &
        call fail('*** Erro na leitura dos materiais !')
      d%ie(ma) = inum(words(2))
      call check_type(d%ie(ma))
      do i = 3, nw
        d%e(i - 2, ma) = fnum(words(i))
      end do
    end do
  end subroutine

  ! fmaterials: ma file per line; key value pairs until end (file_prop,
  ! Rdata.f:2683-2850)
  subroutine read_fmaterials(d)
    type(deck_t), intent(inout) :: d
    character(len=12), parameter :: keys(12) = [character(len=12) :: 'modE', 'poisson', &
      'permeability', 'mbiot', 'cbiot', 'density', 'fdensity', 'porosity', 'l_plastic', &
      'k_plastic', 'mcs', 'pc0']
    character(len=256), allocatable :: ws(:)
    integer :: ma, n, i, k, ety
    do
      call data_line('*** Erro na leitura dos materiais !')
      if (trim(words(1)) == 'end') return
      ma = inum(words(1))
      if (ma < 1 .or. ma > d%numat .or. nw < 2) call fail('*** Erro na leitura dos materiais !')
      call file_words(words(2), ws, n)
      ety = -1
      i = 1
      do while (i + 1 <= n)
        if (trim(ws(i)) == 'end') exit
        if (trim(ws(i)) == 'elType') ety = inum(ws(i + 1))
        do k = 1, 12
          if (trim(ws(i)) == trim(keys(k))) d%e(k, ma) = fnum(ws(i + 1))
        end do
        i = i + 2
      end do
      if (ety < 0) call fail('fmaterials: elType missing')
      call check_type(ety)
      d%ie(ma) = ety
    end do
  end subroutine

  subroutine check_type(ty)
    integer(c_int), intent(in) :: ty
    if (ty /= 17 .and. ty /= 7 .and. ty /= 18) &
      call fail('element type outside this path (hexa20/8 Biot: 17, or 18)')
  end subroutine

  ! j v1 .. vk lines until end (bound, Rdata.f:1408-1460)
  subroutine read_bound(a, n, ncol)
    integer(c_int), intent(inout) :: a(:,:)
    integer, intent(in) :: n, ncol
    integer :: j, i
    do
      call data_line('*** Erro na leitura das restricoes !')
      if (trim(words(1)) == 'end') return
      j = inum(words(1))
      if (j < 1 .or. j > n .or. nw - 1 > ncol) call fail('*** Erro na leitura das restricoes !')
      do i = 2, nw
        a(i - 1, j) = inum(words(i))
      end do
    end do
  end subroutine

  ! hydrostatic <file>: density, p0, datum, scale (init_hydrostatic_pres,
  ! Rdata.f:1301-1345)
  subroutine read_hydrostatic(d, fname)
    type(deck_t), intent(inout) :: d
    character(len=*), intent(in) :: fname
    character(len=256), allocatable :: ws(:)
    integer :: n, i
    real(c_double) :: dens, p0, h, scale
    call file_words(fname, ws, n)
    if (n < 4) call fail('*** Erro na leitura da funcao init_hydrostatic_pres !')
    dens = fnum(ws(1)); p0 = fnum(ws(2)); h = fnum(ws(3)); scale = fnum(ws(4))
    do i = 1, d%nnodev
      d%u0(4, i) = (p0 + dens * d%gmod * (h - d%x(3, i))) * scale
    end do
  end subroutine

  ! loads (rload, Rdata.f:1487-1690); fload(parameter, term, load)
  subroutine read_loads(d)
    type(deck_t), intent(inout) :: d
    character(len=256), allocatable :: ws(:)
    integer :: nc, ty, n, k, l, nterm, ntime, m
    do
      call data_line('*** Load reading error !')
      if (trim(words(1)) == 'end') return
      nc = inum(words(1))
      if (nc < 1 .or. nc > 200 .or. nw < 2) call fail('*** Load reading error !')
      ty = inum(words(2))
      d%load(1, nc) = ty
      d%load(2, nc) = 1
      select case (ty)
      case (1)
        d%fload(1, 1, nc) = fnum(words(3))
      case (2, 3)
        d%fload(1, 1, nc) = fnum(words(3))
        d%fload(2, 1, nc) = fnum(words(4))
      case (6)
        n = inum(words(3))
        if (n < 0 .or. n > 5) call fail('*** Load reading error !')
        d%load(2, nc) = n
        do k = 1, n
          call data_line('*** Load reading error !')
          d%fload(1, k, nc) = fnum(words(1))
          d%fload(2, k, nc) = fnum(words(2))
          d%fload(3, k, nc) = fnum(words(3))
        end do
      case (7)
        d%fload(1, 1, nc) = fnum(words(3))
        d%fload(2, 1, nc) = fnum(words(4))
        d%fload(3, 1, nc) = fnum(words(5))
        n = inum(words(6))
        if (n < 1 .or. n > 1000) call fail('*** Load reading error !')
        d%load(2, nc) = n
        do k = 1, n
          call data_line('*** Load reading error !')
          d%fload(k, 2, nc) = fnum(words(1))
          d%fload(k, 3, nc) = fnum(words(2))
        end do
      case (8)
        n = inum(words(3))
        if (n < 1 .or. n > 1000) call fail('*** Load reading error !')
        d%load(2, nc) = n
        do k = 1, n
          call data_line('*** Load reading error !')
          d%fload(k, 1, nc) = fnum(words(1))
          d%fload(k, 2, nc) = fnum(words(2))
        end do
      case (39, 40, 41, 42, 43)
        if (nw == 3) then   ! <file>: nterm ntime, then ntime rows of nterm values
          call file_words(words(3), ws, m)
          nterm = inum(ws(1))
          ntime = inum(ws(2))
          if (nterm < 1 .or. nterm > 5 .or. ntime < 1 .or. ntime > 1000) &
            call fail('*** Load reading error !')
          if (m < 2 + nterm * ntime) call fail('*** Load reading error !')
          d%load(2, nc) = nterm
          d%load(3, nc) = ntime
          do l = 1, ntime
            do k = 1, nterm
              d%fload(l, k, nc) = fnum(ws(2 + (l - 1) * nterm + k))
            end do
          end do
        else   ! the earlier inline form: nc type n v1 .. vn, constant in time
          n = inum(words(3))
          if (n < 1 .or. n > 4 .or. nw < 3 + n) call fail('*** Load reading error !')
          d%load(2, nc) = n + 1
          d%load(3, nc) = 2
          d%fload(1, 1, nc) = 0.0d0
          d%fload(2, 1, nc) = 1.0d30
          do k = 1, n
            d%fload(1, 1 + k, nc) = fnum(words(3 + k))
            d%fload(2, 1 + k, nc) = fnum(words(3 + k))
          end do
        end if
      case default
        call fail('load type outside this path')
      end select
    end do
  end subroutine

end module pmdeck
