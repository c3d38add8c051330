! A Fortran host driving the GPU slice solver through the class methods.
!
! It restates the 3-D step's slice loop of go_simulation
! (simulation_class.f03:611-736) for one uniform plasma species and one
! bi-Gaussian beam, calling the field2d / species2d / beam3d methods in the
! reference's order (mode 0), or the library's fused slice qp_slice_run
! (mode 1).  After the last slice it writes psit, fxy, bxyz (real space) and
! cu, epw (k space) as raw real(8) in that order.
!
! usage: qp_slice_driver <indx> <indz> <nslices> <precision 0|1> <mode 0|1> <out>
program qp_slice_driver
  use iso_c_binding
  use qpslice_c
  use qp_classes
  implicit none
  type(qp_context), target :: ctx
  type(field2d) :: qb, qe, psit, psi, dv, fxy, bxyz, cu, dcu, amu, epw, epwb
  type(field3d) :: bexyz, bbxyz
  type(species2d) :: sp
  type(beam3d) :: bm
  integer :: ind, indz, nsl, prec, mode, j, l, iter, iters, u
  character(len=256) :: arg, outfile
  real(8) :: box(6), dex, dxi, res
  real(8), allocatable :: a3(:, :, :), a2(:, :, :), a1(:, :, :), k3(:, :, :), k2(:, :, :)
  type(c_ptr) :: sim, sph(1), bmh(1)

  call get_command_argument(1, arg); read(arg, *) ind
  call get_command_argument(2, arg); read(arg, *) indz
  call get_command_argument(3, arg); read(arg, *) nsl
  call get_command_argument(4, arg); read(arg, *) prec
  call get_command_argument(5, arg); read(arg, *) mode
  call get_command_argument(6, outfile)
  box = [-10.0d0, 10.0d0, -10.0d0, 10.0d0, 0.0d0, 10.0d0]
  iter = 2
  call ctx%new(ind, ind, indz, box, prec, 0)
  dex = ctx%dex
  dxi = ctx%dxi
  ! init_sim_fields (simulation_class.f03:267-287)
  call qb%new(ctx, 1, 0); call qe%new(ctx, 1, 0); call psit%new(ctx, 1, 0)
  call psi%new(ctx, 1, 0); call dv%new(ctx, 1, 0); call fxy%new(ctx, 2, 0)
  call bxyz%new(ctx, 3, 0); call cu%new(ctx, 3, 1); call dcu%new(ctx, 2, 0)
  call amu%new(ctx, 3, 0); call epw%new(ctx, 2, 1); call epwb%new(ctx, 2, 0)
  call bexyz%new(ctx, 3); call bbxyz%new(ctx, 3)
  ! the species advances by the xi step dxi per slice (simulation_class.f03:578)
  call sp%new(ctx, [2, 2], 1.0d0, -1.0d0, 1.0d0, dxi, 0.0d0)
  call bm%new(ctx, [0.0d0, 0.0d0, 1.0d0], [1.0d0, 1.0d0, 0.6d0], [1.0d0, 1.0d0, 0.0d0], &
              [16, 16, 32], 4.0d0, 2.0d4, 10.0d0)

  if (mode == 1) then
     sph(1) = sp%h
     bmh(1) = bm%h
     call qp_check(qp_sim_create(ctx%h, sph, 1, bmh, 1, iter, sim), 'sim%new')
     call qp_check(qp_sim_begin_step(sim), 'sim%begin_step')
     do j = 1, nsl
        call qp_check(qp_slice_run(sim, j, iter, 0.0d0, iters, res), 'qp_slice_run')
     end do
     call sim_field(sim, 2, psit); call sim_field(sim, 5, fxy); call sim_field(sim, 6, bxyz)
     call sim_field(sim, 7, cu); call sim_field(sim, 10, epw)
  else
     ! start of the 3-D step (simulation_class.f03:628-641)
     call bm%qdp()
     call fxy%as(0.0d0); call psit%as(0.0d0); call bxyz%as(0.0d0)
     call fxy%cb(bexyz, 1, [1, 2], [1, 2])
     call psit%cb(bexyz, 1, [1], [3])
     call bxyz%cb(bbxyz, 1, [1, 2, 3], [1, 2, 3])
     do j = 1, nsl
        ! beam charge and its transverse field (:650-656)
        call qb%as(0.0d0)
        call bm%qdp(qb, j + 1)
        call qb%fftrk(1)
        call qb%elf(epwb)
        ! plasma charge, psi, E_perp (:657-670)
        call qe%as(0.0d0)
        call sp%qdp(qe)
        call qe%fftrk(1)
        call qe%pot(psi)
        call psi%grad(fxy)
        call fxy%fftkr(1)
        call psi%fftkr(1)
        call sp%extpsi(psi, dex)
        call psi%mult(psi, dex * dex)
        ! predictor fields (:671-675)
        call psit_and_b()
        ! corrector passes (:676-710)
        do l = 1, iter
           call assemble_b()
           call bxyz%fftkr(2)
           call cu%as(0.0d0); call dcu%as(0.0d0); call amu%as(0.0d0)
           call sp%amjdp(fxy, bxyz, psit, cu, amu, dcu, dex)
           call cu%mult(cu, dex); call amu%mult(amu, dex); call dcu%mult(dcu, dex)
           call cu%fftrk(1); call dcu%fftrk(1); call amu%fftrk(3)
           call cu%bfqp(dcu, amu, epw, dex * dex, dex)
           call psit_and_b()
        end do
        call sp%cbq(j + 1)
        ! final fields and predictor current (:717-726)
        call assemble_b()
        call bxyz%fftkr(2)
        call fxy%mult(fxy, -1.0d0)
        call fxy%add(fxy, bxyz, [1], [1], [2])
        call fxy%sub(fxy, bxyz, [2], [2], [1])
        call dcu%mult(dcu, dxi)
        call cu%sub(cu, dcu, [1, 2], [1, 2], [1, 2])
        ! plasma push (:727-730)
        call sp%push(fxy, bxyz, psit, dex)
        ! slice j+1 of the 3-D wake (:731-735)
        call fxy%mult(fxy, dex)
        call fxy%cb(bexyz, j + 1, [1, 2], [1, 2])
        call psit%cb(bexyz, j + 1, [1], [3])
        call bxyz%mult(bxyz, [1, 2], [1, 2], [dex, dex])
        call bxyz%cb(bbxyz, j + 1, [1, 2, 3], [1, 2, 3])
     end do
  end if
  call ctx%sync()

  allocate(a1(1, ctx%nx + 1, ctx%ny + 1), a2(2, ctx%nx + 1, ctx%ny + 1))
  allocate(a3(3, ctx%nx + 1, ctx%ny + 1))
  allocate(k3(3, ctx%ny + 1, ctx%nx + 1), k2(2, ctx%ny + 1, ctx%nx + 1))
  open(newunit=u, file=trim(outfile), access='stream', form='unformatted', status='replace')
  call psit%get(0, a1); write(u) a1
  call fxy%get(0, a2); write(u) a2
  call bxyz%get(0, a3); write(u) a3
  call cu%get(1, k3); write(u) k3
  call epw%get(1, k2); write(u) k2
  close(u)
  if (mode == 1) then
     ! these five handles belong to the fused driver
     psit%h = c_null_ptr; fxy%h = c_null_ptr; bxyz%h = c_null_ptr
     cu%h = c_null_ptr; epw%h = c_null_ptr
     call qp_check(qp_sim_destroy(sim), 'sim%del')
  end if
  call bm%del(); call sp%del()
  call bexyz%del(); call bbxyz%del()
  call qb%del(); call qe%del(); call psit%del(); call psi%del(); call dv%del()
  call fxy%del(); call bxyz%del(); call cu%del(); call dcu%del(); call amu%del()
  call epw%del(); call epwb%del()
  call ctx%del()
  write(*, '(a,i0,a)') 'qp_slice_driver: ', nsl, ' slices done'

contains

  ! cu -> div -> pot -> psit * (-dex) -> real; cu -> bxyz (k)  (:705-709)
  subroutine psit_and_b()
    call cu%div(dv)
    call dv%pot(psit)
    call psit%mult(psit, -dex)
    call psit%fftkr(1)
    call cu%bf(bxyz)
  end subroutine psit_and_b

  ! (:677-680)
  subroutine assemble_b()
    call bxyz%mult(epw, [1, 2], [2, 1], [-dex, dex])
    call bxyz%mult(bxyz, [3], [3], [dex])
    call bxyz%sub(bxyz, epwb, [1], [1], [2])
    call bxyz%add(bxyz, epwb, [2], [2], [1])
  end subroutine assemble_b

  ! a field of the fused driver, wrapped as a field2d without ownership
  subroutine sim_field(s, which, f)
    type(c_ptr), intent(in) :: s
    integer, intent(in) :: which
    type(field2d), intent(inout) :: f
    call qp_check(qp_field2d_destroy(f%h), 'field2d%del')
    call qp_check(qp_sim_field2d(s, which, f%h), 'sim%field2d')
  end subroutine sim_field

end program qp_slice_driver
