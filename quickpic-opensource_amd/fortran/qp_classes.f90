! Fortran-2003 classes over libqpslice.so with the reference's method names.
!
! field2d  : new, del, fftrk, fftkr, div, grad, pot, smooth, elf, bf, bfqp,
!            as, add, sub, mult, cb, get        (field2d_class.f03:38-62)
! field3d  : new, del, as                       (field3d_class.f03)
! species2d: new, del, renew, qdp, extpsi, amjdp, push, cbq
!                                               (species2d_class.f03:37-51)
! beam3d   : new, del, qdp (3-D and slice copy), push  (beam3d_class.f03:35-43)
!
! Each method is one call of the C ABI; a nonzero status stops the program
! with the library's message, as perrors%equit does (perrors_class.f03:96-118).
! The objects own only their handles: all data stays on the GPU.
module qp_classes
  use iso_c_binding
  use qpslice_c
  implicit none
  private
  public :: qp_context, field2d, field3d, species2d, beam3d, qp_check

  type :: qp_context
     type(c_ptr) :: h = c_null_ptr
     integer :: nx = 0, ny = 0, nz = 0, precision = 0
     real(8) :: dex = 0.0d0, dxi = 0.0d0
   contains
     procedure :: new => ctx_new
     procedure :: del => ctx_del
     procedure :: sync => ctx_sync
  end type qp_context

  type :: field2d
     type(c_ptr) :: h = c_null_ptr
     integer :: dim = 0
     class(qp_context), pointer :: ctx => null()
   contains
     procedure :: new => f2_new
     procedure :: del => f2_del
     procedure :: fftrk => f2_fftrk
     procedure :: fftkr => f2_fftkr
     procedure :: div => f2_div
     procedure :: grad => f2_grad
     procedure :: pot => f2_pot
     procedure :: smooth => f2_smooth
     procedure :: elf => f2_elf
     procedure :: bf => f2_bf
     procedure :: bfqp => f2_bfqp
     procedure :: as => f2_as
     procedure :: add => f2_add
     procedure :: sub => f2_sub
     generic :: mult => f2_mult_all, f2_mult_map
     procedure, private :: f2_mult_all, f2_mult_map
     procedure :: cb => f2_cb
     procedure :: get => f2_get
  end type field2d

  type :: field3d
     type(c_ptr) :: h = c_null_ptr
     integer :: dim = 0
   contains
     procedure :: new => f3_new
     procedure :: del => f3_del
     procedure :: as => f3_as
  end type field3d

  type :: species2d
     type(c_ptr) :: h = c_null_ptr
   contains
     procedure :: new => sp_new
     procedure :: del => sp_del
     procedure :: renew => sp_renew
     procedure :: qdp => sp_qdp
     procedure :: extpsi => sp_extpsi
     procedure :: amjdp => sp_amjdp
     procedure :: push => sp_push
     procedure :: cbq => sp_cbq
  end type species2d

  type :: beam3d
     type(c_ptr) :: h = c_null_ptr
   contains
     procedure :: new => bm_new
     procedure :: del => bm_del
     generic :: qdp => bm_qdeposit, bm_qdpcopy
     procedure, private :: bm_qdeposit, bm_qdpcopy
     procedure :: push => bm_push
  end type beam3d

contains

  subroutine qp_check(rc, where)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: where
    if (rc /= 0) then
       write(*, '(a,a,a,i0,a,a)') 'qpslice: ', where, ' failed (', rc, '): ', qp_error_string()
       error stop 1
    end if
  end subroutine qp_check

  ! ---- context ----
  subroutine ctx_new(this, indx, indy, indz, box, precision, device)
    class(qp_context), intent(inout) :: this
    integer, intent(in) :: indx, indy, indz, precision, device
    real(8), intent(in) :: box(6)
    integer(c_int) :: nx, ny, nz, pr
    real(c_double) :: dex, dxi
    call qp_check(qp_ctx_create(indx, indy, indz, box, precision, device, this%h), 'ctx%new')
    call qp_check(qp_ctx_info(this%h, nx, ny, nz, pr, dex, dxi), 'ctx%info')
    this%nx = nx; this%ny = ny; this%nz = nz; this%precision = pr
    this%dex = dex; this%dxi = dxi
  end subroutine ctx_new

  subroutine ctx_del(this)
    class(qp_context), intent(inout) :: this
    call qp_check(qp_ctx_destroy(this%h), 'ctx%del')
    this%h = c_null_ptr
  end subroutine ctx_del

  subroutine ctx_sync(this)
    class(qp_context), intent(inout) :: this
    call qp_check(qp_ctx_sync(this%h), 'ctx%sync')
  end subroutine ctx_sync

  ! ---- field2d ----
  subroutine f2_new(this, ctx, dim, state)
    class(field2d), intent(inout) :: this
    class(qp_context), target, intent(in) :: ctx
    integer, intent(in) :: dim, state
    this%ctx => ctx
    this%dim = dim
    call qp_check(qp_field2d_create(ctx%h, dim, state, this%h), 'field2d%new')
  end subroutine f2_new

  subroutine f2_del(this)
    class(field2d), intent(inout) :: this
    call qp_check(qp_field2d_destroy(this%h), 'field2d%del')
    this%h = c_null_ptr
  end subroutine f2_del

  subroutine f2_fftrk(this, kind)
    class(field2d), intent(inout) :: this
    integer, intent(in) :: kind
    call qp_check(qp_field2d_fftrk(this%h, kind), 'field2d%fftrk')
  end subroutine f2_fftrk

  subroutine f2_fftkr(this, kind)
    class(field2d), intent(inout) :: this
    integer, intent(in) :: kind
    call qp_check(qp_field2d_fftkr(this%h, kind), 'field2d%fftkr')
  end subroutine f2_fftkr

  subroutine f2_div(this, that)
    class(field2d), intent(inout) :: this, that
    call qp_check(qp_field2d_div(this%h, that%h), 'field2d%div')
  end subroutine f2_div

  subroutine f2_grad(this, that)
    class(field2d), intent(inout) :: this, that
    call qp_check(qp_field2d_grad(this%h, that%h), 'field2d%grad')
  end subroutine f2_grad

  subroutine f2_pot(this, that)
    class(field2d), intent(inout) :: this, that
    call qp_check(qp_field2d_pot(this%h, that%h), 'field2d%pot')
  end subroutine f2_pot

  subroutine f2_smooth(this, that)
    class(field2d), intent(inout) :: this, that
    call qp_check(qp_field2d_smooth(this%h, that%h), 'field2d%smooth')
  end subroutine f2_smooth

  subroutine f2_elf(this, that)
    class(field2d), intent(inout) :: this, that
    call qp_check(qp_field2d_elf(this%h, that%h), 'field2d%elf')
  end subroutine f2_elf

  subroutine f2_bf(this, that)
    class(field2d), intent(inout) :: this, that
    call qp_check(qp_field2d_bf(this%h, that%h), 'field2d%bf')
  end subroutine f2_bf

  ! bfqp(cu, dcu, amu, bxy, c, dex), called as cu%bfqp(dcu, amu, bxy, c, dex)
  subroutine f2_bfqp(cu, dcu, amu, bxy, c, dex)
    class(field2d), intent(inout) :: cu, dcu, amu, bxy
    real(8), intent(in) :: c, dex
    call qp_check(qp_field2d_bfqp(cu%h, dcu%h, amu%h, bxy%h, c, dex), 'field2d%bfqp')
  end subroutine f2_bfqp

  subroutine f2_as(this, value)
    class(field2d), intent(inout) :: this
    real(8), intent(in) :: value
    call qp_check(qp_field2d_as(this%h, value), 'field2d%as')
  end subroutine f2_as

  subroutine f2_add(this, a1, a2, dim, dim1, dim2)
    class(field2d), intent(inout) :: this
    class(field2d), intent(in) :: a1, a2
    integer, intent(in), optional :: dim(:), dim1(:), dim2(:)
    integer(c_int) :: z(1) = 0
    if (present(dim)) then
       call qp_check(qp_field2d_add(this%h, a1%h, a2%h, size(dim), dim, dim1, dim2), 'field2d%add')
    else
       call qp_check(qp_field2d_add(this%h, a1%h, a2%h, 0, z, z, z), 'field2d%add')
    end if
  end subroutine f2_add

  subroutine f2_sub(this, a1, a2, dim, dim1, dim2)
    class(field2d), intent(inout) :: this
    class(field2d), intent(in) :: a1, a2
    integer, intent(in), optional :: dim(:), dim1(:), dim2(:)
    integer(c_int) :: z(1) = 0
    if (present(dim)) then
       call qp_check(qp_field2d_sub(this%h, a1%h, a2%h, size(dim), dim, dim1, dim2), 'field2d%sub')
    else
       call qp_check(qp_field2d_sub(this%h, a1%h, a2%h, 0, z, z, z), 'field2d%sub')
    end if
  end subroutine f2_sub

  ! multiply1: this = a1 * value (all components)
  subroutine f2_mult_all(this, a1, value)
    class(field2d), intent(inout) :: this
    class(field2d), intent(in) :: a1
    real(8), intent(in) :: value
    integer(c_int) :: z(1) = 0
    real(c_double) :: v(1)
    v(1) = value
    call qp_check(qp_field2d_mult(this%h, a1%h, 0, z, z, v), 'field2d%mult')
  end subroutine f2_mult_all

  ! multiply2: this(dim(i)) = a1(dim1(i)) * value(i)
  subroutine f2_mult_map(this, a1, dim, dim1, value)
    class(field2d), intent(inout) :: this
    class(field2d), intent(in) :: a1
    integer, intent(in) :: dim(:), dim1(:)
    real(8), intent(in) :: value(:)
    call qp_check(qp_field2d_mult(this%h, a1%h, size(dim), dim, dim1, value), 'field2d%mult')
  end subroutine f2_mult_map

  ! copyto (field2d_class.f03:538-558)
  subroutine f2_cb(this, that, lpos, sdim, ddim)
    class(field2d), intent(inout) :: this
    class(field3d), intent(in) :: that
    integer, intent(in) :: lpos, sdim(:), ddim(:)
    call qp_check(qp_field2d_cb(this%h, that%h, lpos, size(sdim), sdim, ddim), 'field2d%cb')
  end subroutine f2_cb

  ! download, space 0 real f(dim, nx+1, ny+1) / 1 k f(dim, ny+1, nx+1)
  subroutine f2_get(this, space, a)
    class(field2d), intent(in) :: this
    integer, intent(in) :: space
    real(8), intent(out) :: a(:, :, :)
    call qp_check(qp_field2d_download(this%h, space, a, size(a, 2), size(a, 3)), 'field2d%get')
  end subroutine f2_get

  ! ---- field3d ----
  subroutine f3_new(this, ctx, dim)
    class(field3d), intent(inout) :: this
    class(qp_context), intent(in) :: ctx
    integer, intent(in) :: dim
    this%dim = dim
    call qp_check(qp_field3d_create(ctx%h, dim, this%h), 'field3d%new')
  end subroutine f3_new

  subroutine f3_del(this)
    class(field3d), intent(inout) :: this
    call qp_check(qp_field3d_destroy(this%h), 'field3d%del')
    this%h = c_null_ptr
  end subroutine f3_del

  subroutine f3_as(this, value)
    class(field3d), intent(inout) :: this
    real(8), intent(in) :: value
    call qp_check(qp_field3d_as(this%h, value), 'field3d%as')
  end subroutine f3_as

  ! ---- species2d ----
  subroutine sp_new(this, ctx, ppc, density, q, m, dt, s)
    class(species2d), intent(inout) :: this
    class(qp_context), intent(in) :: ctx
    integer, intent(in) :: ppc(2)
    real(8), intent(in) :: density, q, m, dt, s
    type(qp_species_params) :: p
    p%ppc = ppc
    p%density = density
    p%q = q
    p%m = m
    call qp_check(qp_species_create(ctx%h, p, dt, s, this%h), 'species2d%new')
  end subroutine sp_new

  subroutine sp_del(this)
    class(species2d), intent(inout) :: this
    call qp_check(qp_species_destroy(this%h), 'species2d%del')
    this%h = c_null_ptr
  end subroutine sp_del

  subroutine sp_renew(this, s)
    class(species2d), intent(inout) :: this
    real(8), intent(in) :: s
    call qp_check(qp_species_renew(this%h, s), 'species2d%renew')
  end subroutine sp_renew

  subroutine sp_qdp(this, q)
    class(species2d), intent(inout) :: this
    class(field2d), intent(inout) :: q
    call qp_check(qp_species_qdp(this%h, q%h), 'species2d%qdp')
  end subroutine sp_qdp

  subroutine sp_extpsi(this, psi, dex)
    class(species2d), intent(inout) :: this
    class(field2d), intent(in) :: psi
    real(8), intent(in) :: dex
    call qp_check(qp_species_extpsi(this%h, psi%h, dex), 'species2d%extpsi')
  end subroutine sp_extpsi

  subroutine sp_amjdp(this, ef, bf, psit, cu, amu, dcu, dex)
    class(species2d), intent(inout) :: this
    class(field2d), intent(in) :: ef, bf, psit
    class(field2d), intent(inout) :: cu, amu, dcu
    real(8), intent(in) :: dex
    call qp_check(qp_species_amjdp(this%h, ef%h, bf%h, psit%h, cu%h, amu%h, dcu%h, dex), &
                  'species2d%amjdp')
  end subroutine sp_amjdp

  subroutine sp_push(this, ef, bf, psit, dex)
    class(species2d), intent(inout) :: this
    class(field2d), intent(in) :: ef, bf, psit
    real(8), intent(in) :: dex
    call qp_check(qp_species_push(this%h, ef%h, bf%h, psit%h, dex), 'species2d%push')
  end subroutine sp_push

  subroutine sp_cbq(this, pos)
    class(species2d), intent(inout) :: this
    integer, intent(in) :: pos
    call qp_check(qp_species_cbq(this%h, c_null_ptr, pos), 'species2d%cbq')
  end subroutine sp_cbq

  ! ---- beam3d ----
  subroutine bm_new(this, ctx, center, sigma, sigma_v, np, peak, gamma, dt)
    class(beam3d), intent(inout) :: this
    class(qp_context), intent(in) :: ctx
    real(8), intent(in) :: center(3), sigma(3), sigma_v(3), peak, gamma, dt
    integer, intent(in) :: np(3)
    type(qp_beam_params) :: p
    p%center = center
    p%sigma = sigma
    p%sigma_v = sigma_v
    p%centroid_x = 0.0d0
    p%centroid_y = 0.0d0
    p%gamma = gamma
    p%peak = peak
    p%q = -1.0d0
    p%m = 1.0d0
    p%np = np
    p%quiet = 1
    p%evolution = 1
    p%npmax = 0
    call qp_check(qp_beam_create(ctx%h, p, dt, 0_c_int64_t, this%h), 'beam3d%new')
  end subroutine bm_new

  subroutine bm_del(this)
    class(beam3d), intent(inout) :: this
    call qp_check(qp_beam_destroy(this%h), 'beam3d%del')
    this%h = c_null_ptr
  end subroutine bm_del

  subroutine bm_qdeposit(this)
    class(beam3d), intent(inout) :: this
    call qp_check(qp_beam_qdp(this%h), 'beam3d%qdp')
  end subroutine bm_qdeposit

  subroutine bm_qdpcopy(this, q, slice)
    class(beam3d), intent(inout) :: this
    class(field2d), intent(inout) :: q
    integer, intent(in) :: slice
    call qp_check(qp_beam_qdpcopy(this%h, q%h, slice), 'beam3d%qdp')
  end subroutine bm_qdpcopy

  subroutine bm_push(this, ef, bf, dex, dez)
    class(beam3d), intent(inout) :: this
    class(field3d), intent(in) :: ef, bf
    real(8), intent(in) :: dex, dez
    call qp_check(qp_beam_push(this%h, ef%h, bf%h, dex, dez), 'beam3d%push')
  end subroutine bm_push

end module qp_classes
