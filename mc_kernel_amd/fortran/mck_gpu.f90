! mck_gpu.f90 - iso_c_binding interface of the MC Kernel MI355X library (include/mck.h)
! for a Fortran slave process.
!
! The derived types are bind(c) mirrors of mck_db_t, mck_filter_t, mck_setup_t and
! mck_stats_t; the interfaces bind the C entry points one to one.  Arrays are passed as
! assumed-size dummies (no copies): the reference's Fortran arrays vertices(3, nvert),
! connectivity(nvpe, nelems), kernel_values(nkernel, nbasis, nelems), ... have exactly the
! memory order the C side documents.  mck_check reproduces the reference's error
! behaviour (pabort, src/commpi.F90:515-556): print the message and stop the process.
module mck_gpu
  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: MCK_OK = 0, MCK_EINVAL = 1, MCK_ENOTINIT = 2, &
       MCK_ENOMEM = 3, MCK_EHIP = 4, MCK_ENOTFOUND = 5, MCK_ENAN = 6, MCK_EUNSUPPORTED = 7
  integer(c_int32_t), parameter, public :: MCK_PARAM_LAM = 0, MCK_PARAM_VP = 1, &
       MCK_PARAM_VS = 2, MCK_PARAM_MU = 3, MCK_PARAM_VSH = 4, MCK_PARAM_VSV = 5, &
       MCK_PARAM_VPH = 6, MCK_PARAM_VPV = 7, MCK_PARAM_ETA = 8, MCK_PARAM_RHO = 9
  integer(c_int32_t), parameter, public :: MCK_MISFIT_CC = 0, MCK_MISFIT_AM = 1
  integer(c_int32_t), parameter, public :: MCK_RES_KERNEL = 0, MCK_RES_VARIANCE = 1, &
       MCK_RES_ERROR = 2, MCK_RES_MODEL = 3, MCK_RES_NITER = 4, MCK_RES_HETERO = 5
  integer(c_int32_t), parameter, public :: MCK_INT_PARSEVAL = 0, MCK_INT_TRAPEZOIDAL = 1
  integer(c_int32_t), parameter, public :: MCK_FILTER_IDENT = 0, MCK_FILTER_GABOR = 1, &
       MCK_FILTER_BUTTER_LP = 2, MCK_FILTER_BUTTER_HP = 3, MCK_FILTER_BUTTER_BP = 4, &
       MCK_FILTER_BUTTER_LP_ZP = 5, MCK_FILTER_BUTTER_HP_ZP = 6, MCK_FILTER_BUTTER_BP_ZP = 7

  type, bind(c), public :: mck_db_t
     integer(c_int32_t) :: npol, nelem, npoints
     type(c_ptr) :: mesh_s, mesh_z, mp_s, mp_z, fem_mesh, eltype, axis
     integer(c_int32_t) :: nvar
     type(c_ptr) :: snapshots
     real(c_double) :: amplitude
  end type mck_db_t

  type, bind(c), public :: mck_filter_t
     integer(c_int32_t) :: filterclass
     real(c_double) :: frequencies(4)
  end type mck_filter_t

  type, bind(c), public :: mck_setup_t
     type(mck_db_t) :: fwd, bwd
     integer(c_int32_t) :: nt
     real(c_double) :: dt, planet_radius, timeshift_fwd, timeshift_bwd
     type(c_ptr) :: stf_fwd
     real(c_double) :: src_lat, src_lon, src_depth_km
     real(c_double) :: src_mij(6)
     type(c_ptr) :: stf_src
     real(c_double) :: stf_src_shift
     integer(c_int32_t) :: nrec
     type(c_ptr) :: rec_lat, rec_lon
     character(kind=c_char) :: component
     type(c_ptr) :: seis
     integer(c_int32_t) :: nfilter
     type(c_ptr) :: filters
     integer(c_int32_t) :: nkernel
     type(c_ptr) :: kernel_receiver, kernel_filter, kernel_misfit, kernel_param, kernel_window
     integer(c_int32_t) :: model_ndepth
     real(c_float) :: model_dr
     type(c_ptr) :: model_vp, model_vs, model_rho, model_phi, model_xi, model_eta
     real(c_float) :: vp_max, vs_max
     integer(c_int32_t) :: npoints_per_step, max_iter
     real(c_double) :: allowed_error, allowed_relative_error
     integer(c_int32_t) :: int_over_volume, relative_kernel, quasirandom, int_scheme, deconv_stf
     real(c_double) :: damp_radius
     integer(c_int32_t) :: mask_src_rec
     integer(c_int32_t) :: db_page_elems, db_device_pages
  end type mck_setup_t

  type, bind(c), public :: mck_stats_t
     integer(c_int64_t) :: evaluations, points, fwd_locations, bwd_locations, fwd_bytes, bwd_bytes
     integer(c_int32_t) :: mc_steps, strain_launches
     real(c_double) :: device_ms, strain_fwd_ms, strain_bwd_ms, wall_ms
     integer(c_int64_t) :: paged_bytes
     integer(c_int32_t) :: spt, nd_f, nd_b, bwd_path
     integer(c_int64_t) :: fwd_elements, bwd_elem_receivers, alg_bytes
     real(c_double) :: spectra_ms, mfma_ms
     integer(c_int32_t) :: mfma_launches, mfma_groups
     integer(c_int64_t) :: mfma_runs
     real(c_double) :: pair_ms
     integer(c_int64_t) :: pair_pairs
     integer(c_int32_t) :: pair_launches, reserved
  end type mck_stats_t

  type, bind(c), public :: mck_tuning_t
     integer(c_int32_t) :: bwd_path, mfma_waves, mfma_xcd, ncandidates
     real(c_double) :: best_ms, default_ms, candidate_ms(5)
     real(c_double) :: cost_group_us, cost_run_us, cost_pair_us, auto_ms
  end type mck_tuning_t

  public :: mck_tune, mck_set_tuning
  public :: mck_init, mck_slave_work, mck_slave_work_plain, mck_slave_work_sequential, mck_locate_points, &
            mck_load_strain_points, mck_base_kernels, mck_seismograms, mck_integrate_model, &
            mck_set_hetero, mck_integrate_hetero, &
            mck_results_init, mck_results_add, mck_results_add_noalloc, mck_results_get, mck_results_load, &
            mck_results_export, &
            mck_get_stats, mck_finalize, &
            mck_last_error, mck_error_message, mck_check

  interface
     integer(c_int) function mck_init(setup) bind(c, name='mck_init')
       import :: c_int, mck_setup_t
       type(mck_setup_t), intent(in) :: setup
     end function mck_init

     integer(c_int) function mck_slave_work(elem_type, nvpe, nelems, nbasis, nvertices, &
          vertices, conn, kernel_values, kernel_variance, niterations, computation_time, &
          halton_step, halton_offsets) bind(c, name='mck_slave_work')
       import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
       integer(c_int32_t), value :: elem_type, nvpe, nelems, nbasis, nvertices
       real(c_double), intent(in) :: vertices(*)
       integer(c_int32_t), intent(in) :: conn(*)
       real(c_double), intent(out) :: kernel_values(*), kernel_variance(*)
       integer(c_int32_t), intent(out) :: niterations(*)
       real(c_double), intent(out) :: computation_time(*)
       integer(c_int64_t), intent(inout) :: halton_step
       type(c_ptr), value :: halton_offsets
     end function mck_slave_work

     integer(c_int) function mck_slave_work_plain(elem_type, nvpe, nelems, nbasis, nvertices, &
          vertices, conn, kernel_values, kernel_variance, niterations, computation_time, &
          halton_step, halton_offsets) bind(c, name='mck_slave_work_plain')
       import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
       integer(c_int32_t), value :: elem_type, nvpe, nelems, nbasis, nvertices
       real(c_double), intent(in) :: vertices(*)
       integer(c_int32_t), intent(in) :: conn(*)
       real(c_double), intent(out) :: kernel_values(*), kernel_variance(*)
       integer(c_int32_t), intent(out) :: niterations(*)
       real(c_double), intent(out) :: computation_time(*)
       integer(c_int64_t), intent(inout) :: halton_step
       type(c_ptr), value :: halton_offsets
     end function mck_slave_work_plain

     ! the reference's own Halton stream: element e starts where element e-1's draws
     ! ended; offsets_used (c_null_ptr or c_loc of an int64 array(nelems)) receives the
     ! starts, halton_step the position after the task
     integer(c_int) function mck_slave_work_sequential(elem_type, nvpe, nelems, nbasis, nvertices, &
          vertices, conn, kernel_values, kernel_variance, niterations, computation_time, &
          halton_step, offsets_used) bind(c, name='mck_slave_work_sequential')
       import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
       integer(c_int32_t), value :: elem_type, nvpe, nelems, nbasis, nvertices
       real(c_double), intent(in) :: vertices(*)
       integer(c_int32_t), intent(in) :: conn(*)
       real(c_double), intent(out) :: kernel_values(*), kernel_variance(*)
       integer(c_int32_t), intent(out) :: niterations(*)
       real(c_double), intent(out) :: computation_time(*)
       integer(c_int64_t), intent(inout) :: halton_step
       type(c_ptr), value :: offsets_used
     end function mck_slave_work_sequential

     integer(c_int) function mck_locate_points(frame, npts, xyz, elem, xi, eta, ties) &
          bind(c, name='mck_locate_points')
       import :: c_int, c_int32_t, c_double, c_ptr
       integer(c_int32_t), value :: frame, npts
       real(c_double), intent(in) :: xyz(*)
       integer(c_int32_t), intent(out) :: elem(*)
       real(c_double), intent(out) :: xi(*), eta(*)
       type(c_ptr), value :: ties
     end function mck_locate_points

     integer(c_int) function mck_load_strain_points(frame, npts, xyz, traces) &
          bind(c, name='mck_load_strain_points')
       import :: c_int, c_int32_t, c_double, c_float
       integer(c_int32_t), value :: frame, npts
       real(c_double), intent(in) :: xyz(*)
       real(c_float), intent(out) :: traces(*)
     end function mck_load_strain_points

     integer(c_int) function mck_base_kernels(npts, xyz, base) bind(c, name='mck_base_kernels')
       import :: c_int, c_int32_t, c_double
       integer(c_int32_t), value :: npts
       real(c_double), intent(in) :: xyz(*)
       real(c_double), intent(out) :: base(*)
     end function mck_base_kernels

     ! seis(nt, nrec): the seismograms the kernels use (mck_setup_t%seis or, when that was
     ! c_null_ptr, the ones computed by reciprocity at mck_init)
     integer(c_int) function mck_seismograms(seis) bind(c, name='mck_seismograms')
       import :: c_int, c_double
       real(c_double), intent(out) :: seis(*)
     end function mck_seismograms

     ! model(12, nbasis, nelems): integrate_1d_model for the task (slave.f90:815-869)
     integer(c_int) function mck_integrate_model(elem_type, nvpe, nelems, nbasis, nvertices, &
                                                 vertices, conn, model, nsteps, halton_step, &
                                                 halton_offsets) bind(c, name='mck_integrate_model')
       import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
       integer(c_int32_t), value :: elem_type, nvpe, nelems, nbasis, nvertices
       real(c_double), intent(in) :: vertices(*)
       integer(c_int32_t), intent(in) :: conn(*)
       real(c_double), intent(out) :: model(*)
       integer(c_int32_t), intent(out) :: nsteps(*)
       integer(c_int64_t), intent(inout) :: halton_step
       type(c_ptr), value :: halton_offsets
     end function mck_integrate_model

     ! heterogeneity model (load_het_rtpv): rtp(3, nhet) r km, lat, lon; dlnhet(3, nhet)
     integer(c_int) function mck_set_hetero(nhet, rtp, dlnhet) bind(c, name='mck_set_hetero')
       import :: c_int, c_int32_t, c_double, c_float
       integer(c_int32_t), value :: nhet
       real(c_double), intent(in) :: rtp(*)
       real(c_float), intent(in) :: dlnhet(*)
     end function mck_set_hetero

     ! integrate_3d_model for a task: hetero(3, nbasis, nelems)
     integer(c_int) function mck_integrate_hetero(elem_type, nvpe, nelems, nbasis, nvertices, &
                                                  vertices, conn, hetero, nsteps, halton_step, &
                                                  halton_offsets) bind(c, name='mck_integrate_hetero')
       import :: c_int, c_int32_t, c_int64_t, c_double, c_ptr
       integer(c_int32_t), value :: elem_type, nvpe, nelems, nbasis, nvertices
       real(c_double), intent(in) :: vertices(*)
       integer(c_int32_t), intent(in) :: conn(*)
       real(c_double), intent(out) :: hetero(*)
       integer(c_int32_t), intent(out) :: nsteps(*)
       integer(c_int64_t), intent(inout) :: halton_step
       type(c_ptr), value :: halton_offsets
     end function mck_integrate_hetero

     ! the master's accumulators in device memory (extract_receive_buffer,
     ! master_queue.f90:253-323); connectivity(nbasis, nelements), c_null_ptr when volumetric
     integer(c_int) function mck_results_init(int_type, npoints, nelements, nbasis, nmodel, &
                                              nhet, connectivity) bind(c, name='mck_results_init')
       import :: c_int, c_int32_t, c_ptr
       integer(c_int32_t), value :: int_type, npoints, nelements, nbasis, nmodel, nhet
       type(c_ptr), value :: connectivity
     end function mck_results_init

     ! elems(nelems) global element ids (-1 = empty slot); the result arrays as
     ! mck_slave_work returns them (host or device addresses), all c_null_ptr = the
     ! last mck_slave_work's results in device memory
     integer(c_int) function mck_results_add(nelems, elems, kernel_values, kernel_variance, &
                                             niterations, computation_time, model, hetero, rank) &
                                             bind(c, name='mck_results_add')
       import :: c_int, c_int32_t, c_ptr
       integer(c_int32_t), value :: nelems, rank
       integer(c_int32_t), intent(in) :: elems(*)
       type(c_ptr), value :: kernel_values, kernel_variance, niterations, computation_time, model, hetero
     end function mck_results_add

     ! mck_results_add without any allocation, for callers that cannot allocate (no
     ! counterpart in the reference): the four result arrays, model and hetero in device
     ! memory or c_null_ptr, at most 131072 (element, basis function) pairs per call
     integer(c_int) function mck_results_add_noalloc(nelems, elems, kernel_values, kernel_variance, &
                                                     niterations, computation_time, model, hetero, rank) &
                                                     bind(c, name='mck_results_add_noalloc')
       import :: c_int, c_int32_t, c_ptr
       integer(c_int32_t), value :: nelems, rank
       integer(c_int32_t), intent(in) :: elems(*)
       type(c_ptr), value :: kernel_values, kernel_variance, niterations, computation_time, model, hetero
     end function mck_results_add_noalloc

     ! K_x, Var (nkernel, npoints), Bg_Model (nmodel, npoints), niterations
     ! (nkernel, nelements), computation_time, element_proc (nelements): c_null_ptr skips
     integer(c_int) function mck_results_get(kernel, variance, model, hetero, niterations, &
                                             computation_time, element_proc) &
                                             bind(c, name='mck_results_get')
       import :: c_int, c_ptr
       type(c_ptr), value :: kernel, variance, model, hetero, niterations, computation_time, element_proc
     end function mck_results_get

     integer(c_int) function mck_results_load(kernel, variance, model, hetero, niterations, &
                                              computation_time, element_proc) &
                                              bind(c, name='mck_results_load')
       import :: c_int, c_ptr
       type(c_ptr), value :: kernel, variance, model, hetero, niterations, computation_time, element_proc
     end function mck_results_load

     ! finalize's single-precision variables, out(npoints, nentries) as the files hold them
     integer(c_int) function mck_results_export(which, out) bind(c, name='mck_results_export')
       import :: c_int, c_int32_t, c_float
       integer(c_int32_t), value :: which
       real(c_float), intent(out) :: out(*)
     end function mck_results_export

     ! the measured choice of the backward evaluation (FFTW_PLAN's counterpart)
     integer(c_int) function mck_tune(npts, xyz, reps, tuning) bind(c, name='mck_tune')
       import :: c_int, c_int32_t, c_double, mck_tuning_t
       integer(c_int32_t), value :: npts, reps
       real(c_double), intent(in) :: xyz(*)
       type(mck_tuning_t), intent(out) :: tuning
     end function mck_tune

     integer(c_int) function mck_set_tuning(tuning) bind(c, name='mck_set_tuning')
       import :: c_int, mck_tuning_t
       type(mck_tuning_t), intent(in) :: tuning
     end function mck_set_tuning

     integer(c_int) function mck_get_stats(stats) bind(c, name='mck_get_stats')
       import :: c_int, mck_stats_t
       type(mck_stats_t), intent(out) :: stats
     end function mck_get_stats

     subroutine mck_finalize() bind(c, name='mck_finalize')
     end subroutine mck_finalize

     type(c_ptr) function mck_last_error() bind(c, name='mck_last_error')
       import :: c_ptr
     end function mck_last_error

     integer(c_size_t) function c_strlen(s) bind(c, name='strlen')
       import :: c_size_t, c_ptr
       type(c_ptr), value :: s
     end function c_strlen
  end interface

contains

  ! mck_last_error() as a Fortran string
  function mck_error_message() result(msg)
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: n, i
    p = mck_last_error()
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: msg)
    do i = 1, n
       msg(i:i) = chars(i)
    end do
  end function mck_error_message

  ! nonzero return code: report and stop, as pabort does for the reference
  subroutine mck_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    if (rc /= MCK_OK) then
       write(*, '(a, a, i0, a, a)') where, ': mck error ', rc, ': ', mck_error_message()
       error stop 1
    end if
  end subroutine mck_check

end module mck_gpu
