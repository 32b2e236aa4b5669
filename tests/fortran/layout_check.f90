! Prints the sizes of the bind(c) types and the byte offsets of selected fields, and the
! library's mck_last_error() string, for comparison with the C header (no GPU call).
program layout_check
  use, intrinsic :: iso_c_binding
  use mck_gpu
  implicit none
  type(mck_db_t), target :: db
  type(mck_filter_t), target :: flt
  type(mck_setup_t), target :: s
  type(mck_stats_t), target :: st
  type(mck_tuning_t), target :: tu
  write(*, '(a, 1x, i0)') 'mck_db_t size', c_sizeof(db)
  write(*, '(a, 1x, i0)') 'mck_db_t snapshots', off(c_loc(db%snapshots), c_loc(db))
  write(*, '(a, 1x, i0)') 'mck_db_t amplitude', off(c_loc(db%amplitude), c_loc(db))
  write(*, '(a, 1x, i0)') 'mck_filter_t size', c_sizeof(flt)
  write(*, '(a, 1x, i0)') 'mck_setup_t size', c_sizeof(s)
  write(*, '(a, 1x, i0)') 'mck_setup_t nt', off(c_loc(s%nt), c_loc(s))
  write(*, '(a, 1x, i0)') 'mck_setup_t component', off(c_loc(s%component), c_loc(s))
  write(*, '(a, 1x, i0)') 'mck_setup_t kernel_window', off(c_loc(s%kernel_window), c_loc(s))
  write(*, '(a, 1x, i0)') 'mck_setup_t model_dr', off(c_loc(s%model_dr), c_loc(s))
  write(*, '(a, 1x, i0)') 'mck_setup_t damp_radius', off(c_loc(s%damp_radius), c_loc(s))
  write(*, '(a, 1x, i0)') 'mck_setup_t mask_src_rec', off(c_loc(s%mask_src_rec), c_loc(s))
  write(*, '(a, 1x, i0)') 'mck_setup_t db_page_elems', off(c_loc(s%db_page_elems), c_loc(s))
  write(*, '(a, 1x, i0)') 'mck_setup_t db_device_pages', off(c_loc(s%db_device_pages), c_loc(s))
  write(*, '(a, 1x, i0)') 'mck_stats_t size', c_sizeof(st)
  write(*, '(a, 1x, i0)') 'mck_stats_t device_ms', off(c_loc(st%device_ms), c_loc(st))
  write(*, '(a, 1x, i0)') 'mck_stats_t paged_bytes', off(c_loc(st%paged_bytes), c_loc(st))
  write(*, '(a, 1x, i0)') 'mck_stats_t alg_bytes', off(c_loc(st%alg_bytes), c_loc(st))
  write(*, '(a, 1x, i0)') 'mck_stats_t spectra_ms', off(c_loc(st%spectra_ms), c_loc(st))
  write(*, '(a, 1x, i0)') 'mck_stats_t mfma_groups', off(c_loc(st%mfma_groups), c_loc(st))
  write(*, '(a, 1x, i0)') 'mck_tuning_t size', c_sizeof(tu)
  write(*, '(a, 1x, i0)') 'mck_tuning_t default_ms', off(c_loc(tu%default_ms), c_loc(tu))
  write(*, '(a, 1x, i0)') 'mck_tuning_t auto_ms', off(c_loc(tu%auto_ms), c_loc(tu))
  write(*, '(a, 1x, i0)') 'mck_stats_t pair_launches', off(c_loc(st%pair_launches), c_loc(st))
  write(*, '(a, 1x, a)') 'last_error', mck_error_message()
contains
  integer function off(a, b)
    type(c_ptr), intent(in) :: a, b
    off = int(transfer(a, 0_c_intptr_t) - transfer(b, 0_c_intptr_t))
  end function off
end program layout_check
