! End-to-end drive of the library from Fortran through the iso_c_binding module, as the
! reference's do_slave would drive it (src/slave.f90:94-219): a set-up written as raw
! arrays is read back, mck_setup_t is filled with c_loc of those arrays, mck_init runs,
! one task goes through mck_slave_work (batched form, the library's stream positions) and
! mck_slave_work_sequential (the reference's own Halton stream), and the results are
! written back as raw arrays.
!
!   slave_e2e <dir>
!
! <dir>/setup.bin     the bytes of a mck_setup_t (its pointer fields are zero)
! <dir>/manifest.txt  one line per array: <field> <bytes>, read from <dir>/<field>.bin
! <dir>/task.txt      nvertices nelems nvpe nbasis
! <dir>/vertices.bin  vertices(3, nvertices) in m;  <dir>/conn.bin  conn(nvpe, nelems)
! out (in <dir>): values.bin, variance.bin, niterations.bin, halton.bin, evaluations.bin
! (batched form) and seq_values.bin, seq_niterations.bin, seq_offsets.bin, seq_halton.bin
program slave_e2e
  use, intrinsic :: iso_c_binding
  use mck_gpu
  implicit none
  type :: buffer
     integer(c_int8_t), allocatable :: b(:)
  end type buffer
  type(mck_setup_t), target :: setup
  type(mck_stats_t) :: stats
  type(buffer), target, allocatable :: bufs(:)
  integer(c_int8_t), allocatable :: raw(:)
  character(len=512) :: dir
  character(len=64) :: field
  integer :: u, n, i, ios, nbytes, nvert, nel, nvpe, nbasis, nk
  integer(c_int32_t) :: elem_type
  real(c_double), allocatable :: vertices(:, :), values(:, :, :), variance(:, :, :), ctime(:)
  integer(c_int32_t), allocatable :: conn(:, :), niter(:, :)
  integer(c_int64_t), allocatable, target :: offsets(:)
  integer(c_int64_t) :: halton_step

  call get_command_argument(1, dir)
  ! the set-up's scalars
  allocate(raw(c_sizeof(setup)))
  call read_raw(trim(dir)//'/setup.bin', raw)
  setup = transfer(raw, setup)
  ! its arrays
  open(newunit=u, file=trim(dir)//'/manifest.txt', status='old', action='read')
  n = 0
  do
     read(u, *, iostat=ios) field, nbytes
     if (ios /= 0) exit
     n = n + 1
  end do
  rewind(u)
  allocate(bufs(n))
  do i = 1, n
     read(u, *) field, nbytes
     allocate(bufs(i)%b(nbytes))
     call read_raw(trim(dir)//'/'//trim(field)//'.bin', bufs(i)%b)
     call attach(trim(field), c_loc(bufs(i)%b))
  end do
  close(u)
  nk = setup%nkernel

  ! the task
  open(newunit=u, file=trim(dir)//'/task.txt', status='old', action='read')
  read(u, *) nvert, nel, nvpe, nbasis
  close(u)
  allocate(vertices(3, nvert), conn(nvpe, nel))
  open(newunit=u, file=trim(dir)//'/vertices.bin', access='stream', form='unformatted', status='old')
  read(u) vertices
  close(u)
  open(newunit=u, file=trim(dir)//'/conn.bin', access='stream', form='unformatted', status='old')
  read(u) conn
  close(u)
  elem_type = merge(4_c_int32_t, 5_c_int32_t, nvpe == 4)
  allocate(values(nk, nbasis, nel), variance(nk, nbasis, nel), niter(nk, nel), ctime(nel), offsets(nel))

  call mck_check(mck_init(setup), 'mck_init')

  ! the reference's slave_work call, replaced (slave.f90:205-219)
  halton_step = 0
  call mck_check(mck_slave_work(elem_type, int(size(conn, 1), c_int32_t), int(size(conn, 2), c_int32_t), &
                                int(nbasis, c_int32_t), int(size(vertices, 2), c_int32_t), vertices, conn, &
                                values, variance, niter, ctime, halton_step, c_null_ptr), 'mck_slave_work')
  call mck_check(mck_get_stats(stats), 'mck_get_stats')
  call write_raw(trim(dir)//'/values.bin', values)
  call write_raw(trim(dir)//'/variance.bin', variance)
  call write_raw_i(trim(dir)//'/niterations.bin', niter)
  call write_raw_l(trim(dir)//'/halton.bin', [halton_step, stats%evaluations])

  ! the reference's own stream
  halton_step = 0
  call mck_check(mck_slave_work_sequential(elem_type, int(size(conn, 1), c_int32_t), &
                                           int(size(conn, 2), c_int32_t), int(nbasis, c_int32_t), &
                                           int(size(vertices, 2), c_int32_t), vertices, conn, values, &
                                           variance, niter, ctime, halton_step, c_loc(offsets)), &
                 'mck_slave_work_sequential')
  call write_raw(trim(dir)//'/seq_values.bin', values)
  call write_raw_i(trim(dir)//'/seq_niterations.bin', niter)
  call write_raw_l(trim(dir)//'/seq_offsets.bin', offsets)
  call write_raw_l(trim(dir)//'/seq_halton.bin', [halton_step])
  call mck_finalize()
  write(*, '(a)') 'slave_e2e done'

contains

  subroutine read_raw(path, b)
    character(len=*), intent(in) :: path
    integer(c_int8_t), intent(out) :: b(:)
    integer :: v
    open(newunit=v, file=path, access='stream', form='unformatted', status='old', action='read')
    read(v) b
    close(v)
  end subroutine read_raw

  subroutine write_raw(path, a)
    character(len=*), intent(in) :: path
    real(c_double), intent(in) :: a(..)
    integer :: v
    open(newunit=v, file=path, access='stream', form='unformatted', status='replace')
    select rank (a)
    rank (3)
       write(v) a
    end select
    close(v)
  end subroutine write_raw

  subroutine write_raw_i(path, a)
    character(len=*), intent(in) :: path
    integer(c_int32_t), intent(in) :: a(:, :)
    integer :: v
    open(newunit=v, file=path, access='stream', form='unformatted', status='replace')
    write(v) a
    close(v)
  end subroutine write_raw_i

  subroutine write_raw_l(path, a)
    character(len=*), intent(in) :: path
    integer(c_int64_t), intent(in) :: a(:)
    integer :: v
    open(newunit=v, file=path, access='stream', form='unformatted', status='replace')
    write(v) a
    close(v)
  end subroutine write_raw_l

  ! points one pointer field of the set-up at the array read for it
  subroutine attach(name, p)
    character(len=*), intent(in) :: name
    type(c_ptr), intent(in) :: p
    select case (name)
    case ('fwd.mesh_s');        setup%fwd%mesh_s = p
    case ('fwd.mesh_z');        setup%fwd%mesh_z = p
    case ('fwd.mp_s');          setup%fwd%mp_s = p
    case ('fwd.mp_z');          setup%fwd%mp_z = p
    case ('fwd.fem_mesh');      setup%fwd%fem_mesh = p
    case ('fwd.eltype');        setup%fwd%eltype = p
    case ('fwd.axis');          setup%fwd%axis = p
    case ('fwd.snapshots');     setup%fwd%snapshots = p
    case ('bwd.mesh_s');        setup%bwd%mesh_s = p
    case ('bwd.mesh_z');        setup%bwd%mesh_z = p
    case ('bwd.mp_s');          setup%bwd%mp_s = p
    case ('bwd.mp_z');          setup%bwd%mp_z = p
    case ('bwd.fem_mesh');      setup%bwd%fem_mesh = p
    case ('bwd.eltype');        setup%bwd%eltype = p
    case ('bwd.axis');          setup%bwd%axis = p
    case ('bwd.snapshots');     setup%bwd%snapshots = p
    case ('stf_fwd');           setup%stf_fwd = p
    case ('stf_src');           setup%stf_src = p
    case ('rec_lat');           setup%rec_lat = p
    case ('rec_lon');           setup%rec_lon = p
    case ('seis');              setup%seis = p
    case ('filters');           setup%filters = p
    case ('kernel_receiver');   setup%kernel_receiver = p
    case ('kernel_filter');     setup%kernel_filter = p
    case ('kernel_misfit');     setup%kernel_misfit = p
    case ('kernel_param');      setup%kernel_param = p
    case ('kernel_window');     setup%kernel_window = p
    case ('model_vp');          setup%model_vp = p
    case ('model_vs');          setup%model_vs = p
    case ('model_rho');         setup%model_rho = p
    case ('model_phi');         setup%model_phi = p
    case ('model_xi');          setup%model_xi = p
    case ('model_eta');         setup%model_eta = p
    case default
       write(*, '(2a)') 'unknown set-up field: ', name
       stop 1
    end select
  end subroutine attach

end program slave_e2e
