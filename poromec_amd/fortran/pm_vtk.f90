! pm_vtk.f90 -- the results file of the pres macro, written as
! write_mesh_res_pm (Write_vtk.f:912-1393) writes it with the Vtk.f writers:
! legacy VTK (ASCII, or big-endian BINARY with bvtk) when legacy_vtk is set,
! XML VTU (ASCII) otherwise.  Cells: hexa8 (type 12) over the vertices, or
! hexa20 (type 25, elm_vtk's corner and edge order) over every node when
! fprint(1) (quadratic) is set; then the edge nodes' coordinates, pressures
! and recovered fields are their two vertices' average.  Point fields by
! fprint: 2 desloc, 3 pressao, 4 deltaPressao, 9 porosity, 8 darcyFlux,
! 5 totalStress, 7 terzaghiStress, 6 BiotStress (tensors in full, 9 terms).
module pmvtk
  use iso_c_binding
  implicit none
  private
  public :: write_res, res_name

  character(len=1), parameter :: LF = char(10)
  ! hexa20 nodes in VTK order and the (vertex, vertex, edge node) triples
  integer, parameter :: VTK20(20) = [5, 6, 7, 8, 1, 2, 3, 4, 13, 14, 15, 16, 9, 10, 11, 12, &
                                     17, 18, 19, 20]
  integer, parameter :: EDGE(3, 12) = reshape([1, 2, 9, 2, 3, 10, 3, 4, 11, 4, 1, 12, &
                                               5, 6, 13, 6, 7, 14, 7, 8, 15, 8, 5, 16, &
                                               1, 5, 17, 2, 6, 18, 3, 7, 19, 4, 8, 20], [3, 12])

  type field_t
    integer :: kind = 0          ! 0 int scalar, 1 scalar, 2 vector, 3 tensor
    character(len=15) :: name = ''
    integer :: ncomp = 1
    real(c_double), allocatable :: v(:,:)
    integer, allocatable :: iv(:)
  end type

contains

  function res_name(prefix, istep, legacy) result(fn)
    character(len=*), intent(in) :: prefix
    integer, intent(in) :: istep
    logical, intent(in) :: legacy
    character(len=:), allocatable :: fn
    character(len=16) :: b
    write(b, '(i0)') istep
    if (legacy) then
      fn = trim(prefix) // '_istep_' // trim(b) // '.vtk'
    else
      fn = trim(prefix) // '_istep_' // trim(b) // '.vtu'
    end if
  end function

  ! the first np columns of f(nc, *), edge nodes averaged when quad
  subroutine quad_field(f, nc, ix, numel, np, quad, out)
    integer, intent(in) :: nc, numel, np
    real(c_double), intent(in) :: f(nc, *)
    integer(c_int), intent(in) :: ix(21, *)
    logical, intent(in) :: quad
    real(c_double), allocatable, intent(out) :: out(:,:)
    integer :: i, j, n1, n2, n3
    allocate(out(nc, np))
    out = f(:, 1:np)
    if (.not. quad) return
    do i = 1, numel
      do j = 1, 12
        n1 = ix(EDGE(1, j), i)
        n2 = ix(EDGE(2, j), i)
        n3 = ix(EDGE(3, j), i)
        out(:, n3) = 0.5d0 * (f(:, n1) + f(:, n2))
      end do
    end do
  end subroutine

  subroutine full_tensor(t6, n, t9)
    integer, intent(in) :: n
    real(c_double), intent(in) :: t6(6, *)
    real(c_double), allocatable, intent(out) :: t9(:,:)
    integer :: i
    allocate(t9(9, n))
    do i = 1, n
      t9(:, i) = [t6(1, i), t6(4, i), t6(6, i), t6(4, i), t6(2, i), t6(5, i), t6(6, i), &
                  t6(5, i), t6(3, i)]
    end do
  end subroutine

  subroutine write_res(fname, ix, x, nnodev, u, dp, poro, tx, txb, txe, flux, nnode, numel, &
                       istep, t, fprint, legacy, bvtk, ierr)
    character(len=*), intent(in) :: fname
    integer, intent(in) :: nnodev, nnode, numel, istep
    integer(c_int), intent(in) :: ix(21, numel)
    real(c_double), intent(in) :: x(3, nnodev), u(4, nnode), dp(nnode), poro(nnode)
    real(c_double), intent(in) :: tx(6, nnode), txb(6, nnode), txe(6, nnode), flux(3, nnode)
    real(c_double), intent(in) :: t
    logical, intent(in) :: fprint(13), legacy, bvtk
    integer, intent(out) :: ierr
    logical :: quad
    integer :: np, nf, i, nen, ctype
    real(c_double), allocatable :: xf(:,:), xq(:,:), t9(:,:)
    integer, allocatable :: conn(:,:)
    type(field_t) :: flds(12)

    ierr = 0
    if (bvtk .and. .not. legacy) then
      ierr = 1   ! the reference has no binary VTU writer
      return
    end if
    if (fprint(10) .or. fprint(11)) then
      ierr = 2   ! plastic-path fields
      return
    end if
    quad = fprint(1)
    np = nnodev
    if (quad) np = nnode
    allocate(xf(3, nnode))
    xf = 0.0d0
    xf(:, 1:nnodev) = x
    call quad_field(xf, 3, ix, numel, np, quad, xq)
    if (quad) then
      nen = 20
      ctype = 25
      allocate(conn(20, numel))
      do i = 1, numel
        conn(:, i) = ix(VTK20, i) - 1
      end do
    else
      nen = 8
      ctype = 12
      allocate(conn(8, numel))
      conn = ix(1:8, :) - 1
    end if

    nf = 1
    flds(1)%kind = 0
    flds(1)%name = 'noIdG'
    allocate(flds(1)%iv(np))
    flds(1)%iv = [(i, i = 1, np)]
    if (fprint(2)) then
      nf = nf + 1
      call set_field(flds(nf), 2, 'desloc', 3)
      flds(nf)%v = u(1:3, 1:np)
    end if
    if (fprint(3)) then
      nf = nf + 1
      call set_field(flds(nf), 1, 'pressao', 1)
      call quad_field(u(4:4, :), 1, ix, numel, np, quad, flds(nf)%v)
    end if
    if (fprint(4)) then
      nf = nf + 1
      call set_field(flds(nf), 1, 'deltaPressao', 1)
      call quad_field(dp, 1, ix, numel, np, quad, flds(nf)%v)
    end if
    if (fprint(9)) then
      nf = nf + 1
      call set_field(flds(nf), 1, 'porosity', 1)
      call quad_field(poro, 1, ix, numel, np, quad, flds(nf)%v)
    end if
    if (fprint(8)) then
      nf = nf + 1
      call set_field(flds(nf), 2, 'darcyFlux', 3)
      call quad_field(flux, 3, ix, numel, np, quad, flds(nf)%v)
    end if
    if (fprint(5)) then
      nf = nf + 1
      call set_field(flds(nf), 3, 'totalStress', 9)
      call full_tensor(tx, nnode, t9)
      call quad_field(t9, 9, ix, numel, np, quad, flds(nf)%v)
    end if
    if (fprint(7)) then
      nf = nf + 1
      call set_field(flds(nf), 3, 'terzaghiStress', 9)
      call full_tensor(txe, nnode, t9)
      call quad_field(t9, 9, ix, numel, np, quad, flds(nf)%v)
    end if
    if (fprint(6)) then
      nf = nf + 1
      call set_field(flds(nf), 3, 'BiotStress', 9)
      call full_tensor(txb, nnode, t9)
      call quad_field(t9, 9, ix, numel, np, quad, flds(nf)%v)
    end if

    if (legacy .and. bvtk) then
      call legacy_binary(fname, xq, np, conn, nen, ctype, ix, numel, flds, nf, istep, t)
    else if (legacy) then
      call legacy_ascii(fname, xq, np, conn, nen, ctype, ix, numel, flds, nf, istep, t)
    else
      call vtu_ascii(fname, xq, np, conn, nen, ctype, ix, numel, flds, nf, istep, t)
    end if
  end subroutine

  subroutine set_field(f, kind, name, nc)
    type(field_t), intent(inout) :: f
    integer, intent(in) :: kind, nc
    character(len=*), intent(in) :: name
    f%kind = kind
    f%name = name
    f%ncomp = nc
  end subroutine

  subroutine legacy_ascii(fname, xq, np, conn, nen, ctype, ix, numel, flds, nf, istep, t)
    character(len=*), intent(in) :: fname
    integer, intent(in) :: np, nen, ctype, numel, nf, istep
    real(c_double), intent(in) :: xq(3, np), t
    integer, intent(in) :: conn(nen, numel)
    integer(c_int), intent(in) :: ix(21, numel)
    type(field_t), intent(in) :: flds(:)
    character(len=30) :: head
    character(len=15) :: cname
    integer :: u, i, j, k
    head = 'Malha poro mec'
    open(newunit=u, file=fname, status='replace', action='write')
    write(u, '(a)') '# vtk DataFile Version 3.0'
    write(u, '(a)') head
    write(u, '(a)') 'ASCII'
    write(u, '(a)') 'DATASET UNSTRUCTURED_GRID'
    write(u, '(a)') 'FIELD FieldData 4'
    write(u, '(a)') 'TIME_S 1 1 double'
    write(u, '(f16.6)') t
    write(u, '(a)') 'TIME_H 1 1 double'
    write(u, '(f16.6)') t / 3600.0d0
    write(u, '(a)') 'TIME_D 1 1 double'
    write(u, '(f16.6)') t / (24.0d0 * 3600.0d0)
    write(u, '(a)') 'CYCLE 1 1 int'
    write(u, '(i9)') istep
    write(u, '(a,i10,a)') 'POINTS ', np, ' double'
    do i = 1, np
      write(u, '(3e15.5)') xq(:, i)
    end do
    write(u, '(a,i10,i10)') 'CELLS ', numel, numel * (nen + 1)
    do i = 1, numel
      write(u, '(21i10)') nen, conn(:, i)
    end do
    write(u, '(a,i10)') 'CELL_TYPES ', numel
    do i = 1, numel
      write(u, '(i3)') ctype
    end do
    write(u, '(a,1x,i9)') 'CELL_DATA', numel
    cname = 'mat'
    write(u, *) 'SCALARS ', cname, ' int ', 1
    write(u, '(a)') 'LOOKUP_TABLE default'
    do i = 1, numel
      write(u, '(7i10)') ix(21, i)
    end do
    cname = 'elIdG'
    write(u, *) 'SCALARS ', cname, ' int ', 1
    write(u, '(a)') 'LOOKUP_TABLE default'
    do i = 1, numel
      write(u, '(7i10)') i
    end do
    write(u, '(a,1x,i9)') 'POINT_DATA', np
    do k = 1, nf
      select case (flds(k)%kind)
      case (0)
        write(u, '(a,1x,a15,1x,a8,1x,i3)') 'SCALARS', flds(k)%name, 'int    ', 1
        write(u, '(a)') 'LOOKUP_TABLE default'
        do i = 1, np
          write(u, '(99i10)') flds(k)%iv(i)
        end do
      case (1)
        write(u, '(a,1x,a15,1x,a8,1x,i3)') 'SCALARS', flds(k)%name, 'double ', 1
        write(u, '(a)') 'LOOKUP_TABLE default'
        do i = 1, np
          write(u, '(99e15.5e3)') flds(k)%v(1, i)
        end do
      case (2)
        write(u, '(a,15a,a)') 'VECTORS ', flds(k)%name, ' double'
        do i = 1, np
          write(u, '(3e15.5e3)') (flds(k)%v(j, i), j = 1, 3)
        end do
      case (3)
        write(u, '(a,15a,a)') 'TENSORS ', flds(k)%name, ' double'
        do i = 1, np
          write(u, '(9e15.5e3)') (flds(k)%v(j, i), j = 1, 9)
        end do
      end select
    end do
    close(u)
  end subroutine

  subroutine legacy_binary(fname, xq, np, conn, nen, ctype, ix, numel, flds, nf, istep, t)
    character(len=*), intent(in) :: fname
    integer, intent(in) :: np, nen, ctype, numel, nf, istep
    real(c_double), intent(in) :: xq(3, np), t
    integer, intent(in) :: conn(nen, numel)
    integer(c_int), intent(in) :: ix(21, numel)
    type(field_t), intent(in) :: flds(:)
    character(len=30) :: head
    character(len=15) :: s15a, s15b, cname
    character(len=8) :: s8
    character(len=6) :: s6
    integer :: u, i, k
    head = 'Malha poro mec'
    open(newunit=u, file=fname, access='stream', form='unformatted', convert='big_endian', &
         status='replace', action='write')
    write(u) '# vtk DataFile Version 3.0' // LF
    write(u) head
    write(u) LF // 'BINARY' // LF
    write(u) 'DATASET UNSTRUCTURED_GRID' // LF
    write(u) 'FIELD FieldData 4' // LF
    write(u) 'TIME_S 1 1 double' // LF
    write(u) t
    write(u) 'TIME_H 1 1 double' // LF
    write(u) t / 3600.0d0
    write(u) 'TIME_D 1 1 double' // LF
    write(u) t / (24.0d0 * 3600.0d0)
    write(u) 'CYCLE 1 1 int' // LF
    write(u) int(istep, c_int)
    write(s15a, '(i15)') np
    write(u) 'POINTS ' // s15a // ' double' // LF
    do i = 1, np
      write(u) xq(:, i)
    end do
    write(s15a, '(i15)') numel
    write(s15b, '(i15)') numel * (nen + 1)
    write(u) LF // 'CELLS ' // s15a // s15b // LF
    do i = 1, numel
      write(u) int(nen, c_int), int(conn(:, i), c_int)
    end do
    write(u) LF // 'CELL_TYPES ' // s15a // LF
    do i = 1, numel
      write(u) int(ctype, c_int)
    end do
    write(s8, '(i8)') numel
    write(u) LF // 'CELL_DATA ' // s8 // LF
    write(s8, '(i8)') 1
    cname = 'mat'
    write(u) 'SCALARS ' // cname // ' int' // s8 // LF
    write(u) 'LOOKUP_TABLE default' // LF
    do i = 1, numel
      write(u) ix(21, i)
    end do
    cname = 'elIdG'
    write(u) 'SCALARS ' // cname // ' int' // s8 // LF
    write(u) 'LOOKUP_TABLE default' // LF
    do i = 1, numel
      write(u) int(i, c_int)
    end do
    write(s8, '(i8)') np
    write(u) LF // ' POINT_DATA ' // s8 // LF
    write(s6, '(i6)') 1
    do k = 1, nf
      select case (flds(k)%kind)
      case (0)
        write(u) ' SCALARS ' // flds(k)%name // ' int ' // s6 // LF
        write(u) ' LOOKUP_TABLE default ' // LF
        do i = 1, np
          write(u) int(flds(k)%iv(i), c_int)
        end do
      case (1)
        write(u) LF // 'SCALARS ' // flds(k)%name // ' double' // s6 // LF
        write(u) LF // 'LOOKUP_TABLE default ' // LF
        do i = 1, np
          write(u) flds(k)%v(1, i)
        end do
      case (2)
        write(u) ' VECTORS ' // flds(k)%name // ' double' // LF
        do i = 1, np
          write(u) flds(k)%v(1:3, i)
        end do
      case (3)
        write(u) ' TENSORS ' // flds(k)%name // ' double' // LF
        do i = 1, np
          write(u) flds(k)%v(1:9, i)
        end do
      end select
    end do
    close(u)
  end subroutine

  subroutine vtu_ascii(fname, xq, np, conn, nen, ctype, ix, numel, flds, nf, istep, t)
    character(len=*), intent(in) :: fname
    integer, intent(in) :: np, nen, ctype, numel, nf, istep
    real(c_double), intent(in) :: xq(3, np), t
    integer, intent(in) :: conn(nen, numel)
    integer(c_int), intent(in) :: ix(21, numel)
    type(field_t), intent(in) :: flds(:)
    character(len=15) :: s15a, s15b
    character(len=:), allocatable :: nc
    integer :: u, i, j, k
    integer(8) :: off
    open(newunit=u, file=fname, status='replace', action='write')
    write(u, '(a)') '<?xml version="1.0"?>'
    write(u, '(a)') '<VTKFile type="UnstructuredGrid" version="0.1" byte_order="BigEndian">'
    write(u, '(a)') '<UnstructuredGrid>'
    write(u, '(a)') '<FieldData>'
    write(u, '(a)') '<DataArray type="Int32" Name="CYCLE" NumberOfTuples="1" format="ascii">'
    write(u, *) istep
    write(u, '(a)') '</DataArray>'
    write(u, '(a)') '<DataArray type="Float64" Name="TIME_S" NumberOfTuples="1" format="ascii">'
    write(u, *) t
    write(u, '(a)') '</DataArray>'
    write(u, '(a)') '<DataArray type="Float64" Name="TIME_H" NumberOfTuples="1" format="ascii">'
    write(u, *) t / 3600.0d0
    write(u, '(a)') '</DataArray>'
    write(u, '(a)') '<DataArray type="Float64" Name="TIME_D" NumberOfTuples="1" format="ascii">'
    write(u, *) t / (24.0d0 * 3600.0d0)
    write(u, '(a)') '</DataArray>'
    write(u, '(a)') '</FieldData>'
    write(s15a, '(i15)') np
    write(s15b, '(i15)') numel
    write(u, '(a)') '<Piece NumberOfPoints="' // s15a // '" NumberOfCells="' // s15b // '">'
    write(u, '(a)') '<Points>'
    write(u, '(a)') '<DataArray type="Float64" NumberOfComponents ="3" format="ascii">'
    do i = 1, np
      write(u, '(3e15.5)') xq(:, i)
    end do
    write(u, '(a)') '</DataArray>'
    write(u, '(a)') '</Points>'
    write(u, '(a)') '<Cells> '
    write(u, '(a)') '<DataArray type="Int32" Name="connectivity" format="ascii">'
    do i = 1, numel
      write(u, '(21i10)') conn(:, i)
    end do
    write(u, '(a)') '</DataArray>'
    write(u, '(a)') '<DataArray type="Int64" Name="offsets" format="ascii">'
    off = nen
    do i = 1, numel
      write(u, '(i20)') off
      off = off + nen
    end do
    write(u, '(a)') '</DataArray>'
    write(u, '(a)') '<DataArray type="UInt8" Name="types" format="ascii">'
    do i = 1, numel
      write(u, '(i3)') ctype
    end do
    write(u, '(a)') '</DataArray>'
    write(u, '(a)') '</Cells>'
    write(u, '(a)') '<CellData Scalars="scalars">'
    write(u, '(a)') '<DataArray type="Int32" NumberOfComponents="1" Name="mat" format="ascii">'
    do i = 1, numel
      write(u, '(15i9)') ix(21, i)
    end do
    write(u, '(a)') '</DataArray>'
    write(u, '(a)') '<DataArray type="Int32" NumberOfComponents="1" Name="elIdG" format="ascii">'
    do i = 1, numel
      write(u, '(15i9)') i
    end do
    write(u, '(a)') '</DataArray>'
    write(u, '(a)') '</CellData>'
    write(u, '(a)') '<PointData>'
    do k = 1, nf
      select case (flds(k)%kind)
      case (0)
        write(u, '(a)') '<DataArray type="Int32" NumberOfComponents=" 1" Name="' // &
          trim(flds(k)%name) // '" format="ascii">'
        do i = 1, np
          write(u, '(10i10)') flds(k)%iv(i)
        end do
      case (1)
        write(u, '(a)') '<DataArray type="Float64" NumberOfComponents=" 1" Name="' // &
          trim(flds(k)%name) // '" format="ascii">'
        do i = 1, np
          write(u, '(e15.5e3)') flds(k)%v(1, i)
        end do
      case (2, 3)
        nc = ' 3'
        if (flds(k)%kind == 3) nc = '9'
        write(u, '(a)') '<DataArray type="Float64" NumberOfComponents="' // nc // '" Name="' // &
          trim(flds(k)%name) // '" format="ascii">'
        do i = 1, np
          write(u, '(10e15.5e3)') (flds(k)%v(j, i), j = 1, flds(k)%ncomp)
        end do
      end select
      write(u, '(a)') '</DataArray>'
    end do
    write(u, '(a)') '</PointData>'
    write(u, '(a)') '</Piece>'
    write(u, '(a)') '</UnstructuredGrid>'
    write(u, '(a)') '</VTKFile>'
    close(u)
  end subroutine

end module pmvtk
