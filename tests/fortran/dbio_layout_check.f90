! Prints the size of mck_dbfile_t and the byte offsets of its fields as the Fortran module
! mck_dbio lays them out, and the library's last error after opening a missing file.
program dbio_layout_check
  use, intrinsic :: iso_c_binding
  use mck_dbio
  implicit none
  type(mck_dbfile_t), target :: f
  integer(c_int) :: rc
  write(*, '(a, 1x, i0)') 'size', c_sizeof(f)
  write(*, '(a, 1x, i0)') 'nt', off(c_loc(f%nt), c_loc(f))
  write(*, '(a, 1x, i0)') 'dt', off(c_loc(f%dt), c_loc(f))
  write(*, '(a, 1x, i0)') 'file_version', off(c_loc(f%file_version), c_loc(f))
  write(*, '(a, 1x, i0)') 'stf_dump', off(c_loc(f%stf_dump), c_loc(f))
  write(*, '(a, 1x, i0)') 'mesh_eta', off(c_loc(f%mesh_eta), c_loc(f))
  write(*, '(a, 1x, i0)') 'mapped', off(c_loc(f%mapped), c_loc(f))
  write(*, '(a, 1x, i0)') 'priv', off(c_loc(f%priv), c_loc(f))
  rc = mck_db_open('no_such_file.nc4'//c_null_char, 0_c_int32_t, 1_c_int32_t, f)
  write(*, '(a, 1x, i0, 1x, a)') 'open', rc, mck_dbio_message()
contains
  integer function off(a, b)
    type(c_ptr), intent(in) :: a, b
    off = int(transfer(a, 0_c_intptr_t) - transfer(b, 0_c_intptr_t))
  end function off
end program dbio_layout_check
