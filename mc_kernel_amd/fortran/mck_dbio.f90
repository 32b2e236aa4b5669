! mck_dbio.f90 - iso_c_binding module for include/mck_dbio.h (libmck_dbio.so): merged
! database files opened for mck_setup_t, replacing sem_data%open_files / read_meshes for
! a merged database (readfields.f90:499-658, 2925-3078).  Paths are passed as
! trim(path)//c_null_char.
module mck_dbio
  use, intrinsic :: iso_c_binding
  use mck_gpu, only : mck_db_t
  implicit none
  private

  type, bind(c), public :: mck_dbfile_t
     type(mck_db_t) :: db
     integer(c_int32_t) :: nt
     real(c_double) :: dt, planet_radius, source_shift, source_depth_km
     integer(c_int32_t) :: file_version
     type(c_ptr) :: stf_dump, stf_d_dump, sem_mesh
     type(c_ptr) :: mesh_vp, mesh_vs, mesh_rho, mesh_phi, mesh_xi, mesh_eta
     integer(c_int32_t) :: mapped
     type(c_ptr) :: priv
  end type mck_dbfile_t

  public :: mck_db_open, mck_db_close, mck_db_write, mck_dbio_last_error, mck_dbio_message

  interface
     ! nvar_expected 10 (fwd), 5 (bwd) or 0; map_snapshots 1 maps contiguous snapshots
     integer(c_int) function mck_db_open(path, nvar_expected, map_snapshots, f) bind(c, name='mck_db_open')
       import :: c_int, c_int32_t, c_char, mck_dbfile_t
       character(kind=c_char), intent(in) :: path(*)
       integer(c_int32_t), value :: nvar_expected, map_snapshots
       type(mck_dbfile_t), intent(out) :: f
     end function mck_db_open

     subroutine mck_db_close(f) bind(c, name='mck_db_close')
       import :: mck_dbfile_t
       type(mck_dbfile_t), intent(inout) :: f
     end subroutine mck_db_close

     integer(c_int) function mck_db_write(path, f) bind(c, name='mck_db_write')
       import :: c_int, c_char, mck_dbfile_t
       character(kind=c_char), intent(in) :: path(*)
       type(mck_dbfile_t), intent(in) :: f
     end function mck_db_write

     type(c_ptr) function mck_dbio_last_error() bind(c, name='mck_dbio_last_error')
       import :: c_ptr
     end function mck_dbio_last_error

     integer(c_size_t) function dbio_strlen(s) bind(c, name='strlen')
       import :: c_size_t, c_ptr
       type(c_ptr), value :: s
     end function dbio_strlen
  end interface

contains

  ! mck_dbio_last_error() as a Fortran string
  function mck_dbio_message() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: n, i
    p = mck_dbio_last_error()
    n = int(dbio_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
       msg(i:i) = chars(i)
    end do
  end function mck_dbio_message
end module mck_dbio
