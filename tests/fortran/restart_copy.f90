! Reads a restart file record by record with Fortran unformatted sequential input, in the
! order and with the kinds the restart format has (8-byte reals, default integers), and
! writes the same records to a second file with Fortran output.  Used by the tests to show
! that the library's restart files are what a Fortran program reads and writes.
! Usage: restart_copy <in> <out>
program restart_copy
  implicit none
  character(len=4096) :: fin, fout
  real(8) :: lb(3), time, vbkg(3), area
  integer :: lt, nrbc, nwall, i, n0(2), n(2), ctype, nv(2)
  real(8), allocatable :: x(:, :, :), wx(:, :), wf(:, :)
  integer, allocatable :: e2v(:, :)

  call get_command_argument(1, fin)
  call get_command_argument(2, fout)
  open (11, file=trim(fin), form='unformatted', action='read', status='old')
  open (12, file=trim(fout), form='unformatted', action='write')

  read (11) lb
  read (11) lt
  read (11) time
  read (11) vbkg
  write (12) lb
  write (12) lt
  write (12) time
  write (12) vbkg

  read (11) nrbc
  write (12) nrbc
  do i = 1, nrbc
    read (11) n0(1), n0(2)
    read (11) n(1), n(2)
    read (11) ctype
    read (11) area
    allocate (x(n(1), n(2), 3))
    read (11) x
    write (12) n0(1), n0(2)
    write (12) n(1), n(2)
    write (12) ctype
    write (12) area
    write (12) x
    deallocate (x)
  end do

  read (11) nwall
  write (12) nwall
  do i = 1, nwall
    read (11) nv(1), nv(2)
    allocate (wx(nv(1), 3), wf(nv(1), 3), e2v(nv(2), 3))
    read (11) wx
    read (11) wf
    read (11) e2v
    write (12) nv(1), nv(2)
    write (12) wx
    write (12) wf
    write (12) e2v
    deallocate (wx, wf, e2v)
  end do

  close (11)
  close (12)
  print '(A,I0,A,I0,A,I0)', 'restart lt=', lt, ' nrbc=', nrbc, ' nwall=', nwall
end program restart_copy
