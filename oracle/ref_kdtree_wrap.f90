! ref_kdtree_wrap.f90 - bind(c) entry points around the reference's own kd-tree
! (src/kdtree2.f90, compiled unchanged from the reference tree into oracle/_ref).
! TEST INFRASTRUCTURE: lets the tests compare the oracle's 12-nearest-midpoint order with
! kdtree2_n_nearest + kdtree2_sort_results exactly as find_element calls them
! (readfields.f90:1065-1073, build_kdtree readfields.f90:2772-2850), and the oracle's
! nearest heterogeneity-model point with the 3-D tree of build_hetero_kdtree /
! load_model_coeffs (heterogeneities.f90:114-188).
module ref_kdtree_wrap
  use iso_c_binding
  use kdtree2_module
  implicit none
  type(kdtree2), pointer, save :: tree => null(), tree3 => null()
  real(kdkind), allocatable, target, save :: pts(:,:), pts3(:,:)
contains
  subroutine ref_kd_build(n, s, z) bind(c, name="ref_kd_build")
    integer(c_int), value :: n
    real(c_float), intent(in) :: s(n), z(n)
    if (associated(tree)) call kdtree2_destroy(tree)
    if (allocated(pts)) deallocate(pts)
    allocate(pts(2, n))
    pts(1, :) = s
    pts(2, :) = z
    tree => kdtree2_create(pts, dim=2, sort=.true., rearrange=.true.)
  end subroutine ref_kd_build

  subroutine ref_kd_nearest(s, z, nn, idx, dis) bind(c, name="ref_kd_nearest")
    real(c_double), value :: s, z
    integer(c_int), value :: nn
    integer(c_int), intent(out) :: idx(nn)
    real(c_float), intent(out) :: dis(nn)
    type(kdtree2_result), allocatable :: res(:)
    allocate(res(nn))
    call kdtree2_n_nearest(tree, real([s, z], kind=kdkind), nn, res)
    idx = res%idx
    dis = res%dis
    deallocate(res)
  end subroutine ref_kd_nearest

  subroutine ref_kd3_build(n, xyz) bind(c, name="ref_kd3_build")
    integer(c_int), value :: n
    real(c_float), intent(in) :: xyz(3, n)
    if (associated(tree3)) call kdtree2_destroy(tree3)
    if (allocated(pts3)) deallocate(pts3)
    allocate(pts3(3, n))
    pts3 = xyz
    tree3 => kdtree2_create(pts3, dim=3, sort=.true., rearrange=.true.)
  end subroutine ref_kd3_build

  subroutine ref_kd3_nearest(q, idx, dis) bind(c, name="ref_kd3_nearest")
    real(c_float), intent(in) :: q(3)
    integer(c_int), intent(out) :: idx
    real(c_float), intent(out) :: dis
    type(kdtree2_result), allocatable :: res(:)
    allocate(res(1))
    call kdtree2_n_nearest(tree3, q, 1, res)
    idx = res(1)%idx
    dis = res(1)%dis
    deallocate(res)
  end subroutine ref_kd3_nearest
end module ref_kdtree_wrap
