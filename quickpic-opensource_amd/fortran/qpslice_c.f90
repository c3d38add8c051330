! ISO_C_BINDING interfaces of libqpslice.so (include/qpslice.h).
!
! One interface per C entry point; the comment names the reference method it
! stands for.  Handles are type(c_ptr); arrays cross in the reference's own
! column-major layouts, always as real(c_double).
module qpslice_c
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: QP_FP64 = 0, QP_FP32 = 1

  type, bind(C) :: qp_species_params
     integer(c_int) :: ppc(2)
     real(c_double) :: density, q, m
  end type qp_species_params

  type, bind(C) :: qp_beam_params
     real(c_double) :: center(3), sigma(3), sigma_v(3), centroid_x(3), centroid_y(3)
     real(c_double) :: gamma, peak, q, m
     integer(c_int) :: np(3)
     integer(c_int) :: quiet, evolution
     integer(c_int64_t) :: npmax
  end type qp_beam_params

  interface
     ! ---- context (spect2d/spect3d, simulation%new) ----
     integer(c_int) function qp_ctx_create(indx, indy, indz, box, prec, dev, out) &
          bind(C, name="qp_ctx_create")
       import :: c_int, c_double, c_ptr
       integer(c_int), value :: indx, indy, indz, prec, dev
       real(c_double), intent(in) :: box(6)
       type(c_ptr), intent(out) :: out
     end function qp_ctx_create
     integer(c_int) function qp_ctx_destroy(c) bind(C, name="qp_ctx_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: c
     end function qp_ctx_destroy
     integer(c_int) function qp_ctx_sync(c) bind(C, name="qp_ctx_sync")
       import :: c_int, c_ptr
       type(c_ptr), value :: c
     end function qp_ctx_sync
     integer(c_int) function qp_ctx_stream(c, stream) bind(C, name="qp_ctx_stream")
       import :: c_int, c_ptr
       type(c_ptr), value :: c
       type(c_ptr), intent(out) :: stream
     end function qp_ctx_stream
     integer(c_int) function qp_ctx_info(c, nx, ny, nz, prec, dex, dxi) &
          bind(C, name="qp_ctx_info")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: c
       integer(c_int), intent(out) :: nx, ny, nz, prec
       real(c_double), intent(out) :: dex, dxi
     end function qp_ctx_info
     type(c_ptr) function qp_last_error() bind(C, name="qp_last_error")
       import :: c_ptr
     end function qp_last_error

     ! ---- field2d (field2d_class.f03:38-62) ----
     integer(c_int) function qp_field2d_create(c, dim, state, out) &
          bind(C, name="qp_field2d_create")
       import :: c_int, c_ptr
       type(c_ptr), value :: c
       integer(c_int), value :: dim, state
       type(c_ptr), intent(out) :: out
     end function qp_field2d_create
     integer(c_int) function qp_field2d_destroy(f) bind(C, name="qp_field2d_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: f
     end function qp_field2d_destroy
     integer(c_int) function qp_field2d_download(f, space, host, ld1, ld2) &
          bind(C, name="qp_field2d_download")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: f
       integer(c_int), value :: space, ld1, ld2
       real(c_double), intent(out) :: host(*)
     end function qp_field2d_download
     integer(c_int) function qp_field2d_upload(f, space, host, ld1, ld2) &
          bind(C, name="qp_field2d_upload")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: f
       integer(c_int), value :: space, ld1, ld2
       real(c_double), intent(in) :: host(*)
     end function qp_field2d_upload
     integer(c_int) function qp_field2d_as(f, v) bind(C, name="qp_field2d_as")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: f
       real(c_double), value :: v
     end function qp_field2d_as
     integer(c_int) function qp_field2d_fftrk(f, kind) bind(C, name="qp_field2d_fftrk")
       import :: c_int, c_ptr
       type(c_ptr), value :: f
       integer(c_int), value :: kind
     end function qp_field2d_fftrk
     integer(c_int) function qp_field2d_fftkr(f, kind) bind(C, name="qp_field2d_fftkr")
       import :: c_int, c_ptr
       type(c_ptr), value :: f
       integer(c_int), value :: kind
     end function qp_field2d_fftkr
     integer(c_int) function qp_field2d_div(f, out) bind(C, name="qp_field2d_div")
       import :: c_int, c_ptr
       type(c_ptr), value :: f, out
     end function qp_field2d_div
     integer(c_int) function qp_field2d_grad(f, out) bind(C, name="qp_field2d_grad")
       import :: c_int, c_ptr
       type(c_ptr), value :: f, out
     end function qp_field2d_grad
     integer(c_int) function qp_field2d_pot(f, out) bind(C, name="qp_field2d_pot")
       import :: c_int, c_ptr
       type(c_ptr), value :: f, out
     end function qp_field2d_pot
     integer(c_int) function qp_field2d_smooth(f, out) bind(C, name="qp_field2d_smooth")
       import :: c_int, c_ptr
       type(c_ptr), value :: f, out
     end function qp_field2d_smooth
     integer(c_int) function qp_field2d_elf(f, out) bind(C, name="qp_field2d_elf")
       import :: c_int, c_ptr
       type(c_ptr), value :: f, out
     end function qp_field2d_elf
     integer(c_int) function qp_field2d_bf(f, out) bind(C, name="qp_field2d_bf")
       import :: c_int, c_ptr
       type(c_ptr), value :: f, out
     end function qp_field2d_bf
     integer(c_int) function qp_field2d_bfqp(cu, dcu, amu, bxy, c, dex) &
          bind(C, name="qp_field2d_bfqp")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: cu, dcu, amu, bxy
       real(c_double), value :: c, dex
     end function qp_field2d_bfqp
     integer(c_int) function qp_field2d_add(f, a1, a2, n, dim, dim1, dim2) &
          bind(C, name="qp_field2d_add")
       import :: c_int, c_ptr
       type(c_ptr), value :: f, a1, a2
       integer(c_int), value :: n
       integer(c_int), intent(in) :: dim(*), dim1(*), dim2(*)
     end function qp_field2d_add
     integer(c_int) function qp_field2d_sub(f, a1, a2, n, dim, dim1, dim2) &
          bind(C, name="qp_field2d_sub")
       import :: c_int, c_ptr
       type(c_ptr), value :: f, a1, a2
       integer(c_int), value :: n
       integer(c_int), intent(in) :: dim(*), dim1(*), dim2(*)
     end function qp_field2d_sub
     integer(c_int) function qp_field2d_mult(f, a1, n, dim, dim1, val) &
          bind(C, name="qp_field2d_mult")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: f, a1
       integer(c_int), value :: n
       integer(c_int), intent(in) :: dim(*), dim1(*)
       real(c_double), intent(in) :: val(*)
     end function qp_field2d_mult
     integer(c_int) function qp_field2d_cb(f, f3, lpos, n, sdim, ddim) &
          bind(C, name="qp_field2d_cb")
       import :: c_int, c_ptr
       type(c_ptr), value :: f, f3
       integer(c_int), value :: lpos, n
       integer(c_int), intent(in) :: sdim(*), ddim(*)
     end function qp_field2d_cb

     ! ---- field3d (ufield3d_class.f03:396-490) ----
     integer(c_int) function qp_field3d_create(c, dim, out) bind(C, name="qp_field3d_create")
       import :: c_int, c_ptr
       type(c_ptr), value :: c
       integer(c_int), value :: dim
       type(c_ptr), intent(out) :: out
     end function qp_field3d_create
     integer(c_int) function qp_field3d_destroy(f) bind(C, name="qp_field3d_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: f
     end function qp_field3d_destroy
     integer(c_int) function qp_field3d_as(f, v) bind(C, name="qp_field3d_as")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: f
       real(c_double), value :: v
     end function qp_field3d_as

     ! ---- species2d (species2d_class.f03:37-51) ----
     integer(c_int) function qp_species_create(c, p, dt, s, out) &
          bind(C, name="qp_species_create")
       import :: c_int, c_double, c_ptr, qp_species_params
       type(c_ptr), value :: c
       type(qp_species_params), intent(in) :: p
       real(c_double), value :: dt, s
       type(c_ptr), intent(out) :: out
     end function qp_species_create
     integer(c_int) function qp_species_destroy(sp) bind(C, name="qp_species_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: sp
     end function qp_species_destroy
     integer(c_int) function qp_species_renew(sp, s) bind(C, name="qp_species_renew")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: sp
       real(c_double), value :: s
     end function qp_species_renew
     integer(c_int) function qp_species_qdp(sp, q) bind(C, name="qp_species_qdp")
       import :: c_int, c_ptr
       type(c_ptr), value :: sp, q
     end function qp_species_qdp
     integer(c_int) function qp_species_extpsi(sp, psi, dex) bind(C, name="qp_species_extpsi")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: sp, psi
       real(c_double), value :: dex
     end function qp_species_extpsi
     integer(c_int) function qp_species_amjdp(sp, ef, bf, psit, cu, amu, dcu, dex) &
          bind(C, name="qp_species_amjdp")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: sp, ef, bf, psit, cu, amu, dcu
       real(c_double), value :: dex
     end function qp_species_amjdp
     integer(c_int) function qp_species_push(sp, ef, bf, psit, dex) &
          bind(C, name="qp_species_push")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: sp, ef, bf, psit
       real(c_double), value :: dex
     end function qp_species_push
     integer(c_int) function qp_species_cbq(sp, q3, pos) bind(C, name="qp_species_cbq")
       import :: c_int, c_ptr
       type(c_ptr), value :: sp, q3
       integer(c_int), value :: pos
     end function qp_species_cbq

     ! ---- beam3d (beam3d_class.f03:35-43) ----
     integer(c_int) function qp_beam_create(c, p, dt, seed, out) bind(C, name="qp_beam_create")
       import :: c_int, c_double, c_ptr, c_int64_t, qp_beam_params
       type(c_ptr), value :: c
       type(qp_beam_params), intent(in) :: p
       real(c_double), value :: dt
       integer(c_int64_t), value :: seed
       type(c_ptr), intent(out) :: out
     end function qp_beam_create
     integer(c_int) function qp_beam_destroy(b) bind(C, name="qp_beam_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: b
     end function qp_beam_destroy
     integer(c_int) function qp_beam_qdp(b) bind(C, name="qp_beam_qdp")
       import :: c_int, c_ptr
       type(c_ptr), value :: b
     end function qp_beam_qdp
     integer(c_int) function qp_beam_qdpcopy(b, q, slice) bind(C, name="qp_beam_qdpcopy")
       import :: c_int, c_ptr
       type(c_ptr), value :: b, q
       integer(c_int), value :: slice
     end function qp_beam_qdpcopy
     integer(c_int) function qp_beam_push(b, ef, bf, dex, dez) bind(C, name="qp_beam_push")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: b, ef, bf
       real(c_double), value :: dex, dez
     end function qp_beam_push

     ! ---- the slice loop of go_simulation (simulation_class.f03:643-736) ----
     integer(c_int) function qp_sim_create(c, sp, nsp, bm, nb, iter, out) &
          bind(C, name="qp_sim_create")
       import :: c_int, c_ptr
       type(c_ptr), value :: c
       type(c_ptr), intent(in) :: sp(*), bm(*)
       integer(c_int), value :: nsp, nb, iter
       type(c_ptr), intent(out) :: out
     end function qp_sim_create
     integer(c_int) function qp_sim_destroy(s) bind(C, name="qp_sim_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: s
     end function qp_sim_destroy
     integer(c_int) function qp_sim_field2d(s, which, out) bind(C, name="qp_sim_field2d")
       import :: c_int, c_ptr
       type(c_ptr), value :: s
       integer(c_int), value :: which
       type(c_ptr), intent(out) :: out
     end function qp_sim_field2d
     integer(c_int) function qp_sim_begin_step(s) bind(C, name="qp_sim_begin_step")
       import :: c_int, c_ptr
       type(c_ptr), value :: s
     end function qp_sim_begin_step
     integer(c_int) function qp_slice_run(s, j, iter_max, tol, iters, res) &
          bind(C, name="qp_slice_run")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: s
       integer(c_int), value :: j, iter_max
       real(c_double), value :: tol
       integer(c_int), intent(out) :: iters
       real(c_double), intent(out) :: res
     end function qp_slice_run
     ! the slice in two parts, for a caller-provided corrector seed
     integer(c_int) function qp_slice_prepare(s, j) bind(C, name="qp_slice_prepare")
       import :: c_int, c_ptr
       type(c_ptr), value :: s
       integer(c_int), value :: j
     end function qp_slice_prepare
     integer(c_int) function qp_slice_finish(s, j, iter_max, tol, iters, res) &
          bind(C, name="qp_slice_finish")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: s
       integer(c_int), value :: j, iter_max
       real(c_double), value :: tol
       integer(c_int), intent(out) :: iters
       real(c_double), intent(out) :: res
     end function qp_slice_finish
     integer(c_int) function qp_sim_end_step(s, snew) bind(C, name="qp_sim_end_step")
       import :: c_int, c_double, c_ptr
       type(c_ptr), value :: s
       real(c_double), value :: snew
     end function qp_sim_end_step
  end interface

contains

  ! qp_last_error() as a Fortran string
  function qp_error_string() result(msg)
    character(len=:), allocatable :: msg
    character(kind=c_char), pointer :: cs(:)
    type(c_ptr) :: p
    integer :: n
    p = qp_last_error()
    if (.not. c_associated(p)) then
       msg = ''
       return
    end if
    call c_f_pointer(p, cs, [4096])
    n = 0
    do while (n < 4096)
       if (cs(n + 1) == c_null_char) exit
       n = n + 1
    end do
    allocate(character(len=n) :: msg)
    msg = transfer(cs(1:n), msg)
  end function qp_error_string

end module qpslice_c
