! gsis_cavity_host.f90 -- a Fortran host driving the GSIS kinetic loop on the
! GPU through the gsis_capi module, in the shape of Hybrid2D's main loop
! (src/opencfd_ec2d.f90:45-56: one Hybrid_Time_advance_LU_SGS per step, the
! residual monitor every Kstep_show steps).
!
! Case: lid-driven cavity on [0,1]^2, n x n uniform cells, diffuse walls at
! T=1, lid velocity `lid` on the top face, gas at rest initially; velocity
! grid v = 6 sinh(2s)/sinh(2) on nv uniform s in [-1,1] with composite
! Newton-Cotes weights times dv/ds; Kn given, cfc = sqrt(pi)/(2 Kn), Pr=2/3,
! omega=0.81, Shakhov model, NND flux.
!
! usage: gsis_cavity_host n nv kn lid dt nsteps kstep_show [restart]
! prints "res <kstep> r1 r2 r3 r4" every kstep_show steps and finally
! "sum <sum of S_GKUA(1:4) over cells 0..n> <sum of CF>".
! restart = save1 / save2: after the run, back up into restart1 / restart2
! (backup_Macro_write, src/Backup_write_read.f90:114-209; the directories
! restart?/star00 must exist, as for the reference); restart = resume: before
! the run, restore from the backups as init_flow_read does (:4-112).
program gsis_cavity_host
  use iso_c_binding
  use gsis_capi
  implicit none
  integer :: n, nv, nsteps, nshow, nx, ny, i, j, k, iv, jv, done
  real(c_double) :: kn, lid, dt, hs, p0, g, ma, cfc, sres(4), tt
  integer(c_int32_t) :: kstep, nsit
  character(len=64) :: arg, rmode
  real(c_double), allocatable, target :: x(:,:), y(:,:), vol(:,:), si(:,:), sj(:,:), ni1(:,:), ni2(:,:), &
                                         nj1(:,:), nj2(:,:), vi(:), wi(:), frd(:,:,:,:,:), s(:,:,:), &
                                         cf(:,:), u(:,:,:)
  type(gsis_bc_t), allocatable, target :: bc(:)
  type(gsis_params_t) :: p
  type(c_ptr) :: h
  real(c_double), parameter :: pi = 3.141592653589793d0

  call get_command_argument(1, arg); read(arg, *) n
  call get_command_argument(2, arg); read(arg, *) nv
  call get_command_argument(3, arg); read(arg, *) kn
  call get_command_argument(4, arg); read(arg, *) lid
  call get_command_argument(5, arg); read(arg, *) dt
  call get_command_argument(6, arg); read(arg, *) nsteps
  call get_command_argument(7, arg); read(arg, *) nshow
  rmode = 'none'
  if (command_argument_count() >= 8) call get_command_argument(8, rmode)
  nx = n + 1
  ny = n + 1

  ! nodes x(-1:nx+2,-1:ny+2) (Update_coordinate_buffer_onemesh layout)
  allocate(x(-1:nx+2, -1:ny+2), y(-1:nx+2, -1:ny+2))
  x = 0d0; y = 0d0
  hs = 1d0 / n
  do j = 1, ny
    do i = 1, nx
      x(i, j) = (i - 1) * hs
      y(i, j) = (j - 1) * hs
    end do
  end do
  allocate(bc(4))
  bc(1) = gsis_bc_t(2, 1, nx, 1, 1, GSIS_BC_WALL, 0d0, 0d0)
  bc(2) = gsis_bc_t(3, nx, nx, 1, ny, GSIS_BC_WALL, 0d0, 0d0)
  bc(3) = gsis_bc_t(4, 1, nx, ny, ny, GSIS_BC_WALL, lid, 0d0)
  bc(4) = gsis_bc_t(1, 1, 1, 1, ny, GSIS_BC_WALL, 0d0, 0d0)
  call ghost_nodes(x)
  call ghost_nodes(y)
  call geometry()

  ! velocity grid
  allocate(vi(nv), wi(nv))
  call sinh_grid(nv, 6d0, 2d0, vi, wi)

  ! constants of a case given by Kn
  g = 5d0 / 3d0
  ma = 1d0
  cfc = sqrt(pi) / (2d0 * kn)
  p%nx = nx; p%ny = ny; p%nvi = nv; p%nvj = nv
  p%vi = c_loc(vi); p%vj = c_loc(vi); p%wi = c_loc(wi); p%wj = c_loc(wi)
  p%x = c_loc(x); p%y = c_loc(y)
  p%vol = c_loc(vol); p%si = c_loc(si); p%sj = c_loc(sj)
  p%ni1 = c_loc(ni1); p%ni2 = c_loc(ni2); p%nj1 = c_loc(nj1); p%nj2 = c_loc(nj2)
  p%nbc = 4; p%bc = c_loc(bc)
  p%cfc = cfc; p%x1s = 1d0 - 0.81d0; p%cpr = (1d0 - 2d0 / 3d0) / 2.5d0; p%twall = 1d0; p%dt = dt
  p%ma = ma; p%gamma = g; p%re = cfc * ma * sqrt(2d0 * g); p%pr = 2d0 / 3d0
  p%cv = 1d0 / (g * (g - 1d0) * ma * ma); p%cp = p%cv * g; p%aoa = 0d0
  p%scheme = GSIS_SCHEME_NND; p%mode = GSIS_MODE_M0; p%precision = GSIS_PREC_FP64; p%if_viscous = 1
  p%local_dt = 1; p%cfl = 20d0; p%dtmax = 10d0; p%dtmin = 1d-7
  p%ns_tol = real(1e-6, c_double); p%ns_maxit = 0
  p%kstep_show = nshow; p%split_kernels = 0; p%walk_tiles = 0
  call gsis_check(gsis_create(p, h), 'gsis_create')

  ! gas at rest, n = T = 1: init_flow_zero (src/sub_init.f90:400-457) and the
  ! Maxwellian fill of GKUA_sudufenbu (src/GKUA_sudufenbu.f90:22-46)
  allocate(frd(2, -3:nx+3, -3:ny+3, nv, nv), s(11, -3:nx+3, -3:ny+3), cf(-1:nx+1, -1:ny+1), &
           u(4, -1:nx+1, -1:ny+1))
  frd = 0d0; s = 0d0
  do jv = 1, nv
    do iv = 1, nv
      do j = -1, ny + 1
        do i = -1, nx + 1
          frd(1, i, j, iv, jv) = exp(-(vi(iv)**2 + vi(jv)**2)) / pi
          frd(2, i, j, iv, jv) = frd(1, i, j, iv, jv) / 2d0
        end do
      end do
    end do
  end do
  s(1, -1:nx+1, -1:ny+1) = 1d0
  s(4, -1:nx+1, -1:ny+1) = 1d0
  cf = cfc
  p0 = 1d0 / (g * ma * ma)
  u(1, :, :) = 1d0; u(2, :, :) = 0d0; u(3, :, :) = 0d0; u(4, :, :) = p0 / (g - 1d0)
  call gsis_check(gsis_upload_frd(h, frd), 'gsis_upload_frd')
  call gsis_check(gsis_upload_macro(h, c_loc(s), c_loc(cf), c_loc(u), c_null_ptr), 'gsis_upload_macro')
  call gsis_check(gsis_refresh_moments(h), 'gsis_refresh_moments')
  if (rmode == 'resume') call read_backup()

  ! main loop
  done = 0
  do while (done < nsteps)
    k = min(nshow, nsteps - done)
    call gsis_check(gsis_step(h, int(k, c_int32_t)), 'gsis_step')
    done = done + k
    call gsis_check(gsis_residual(h, sres), 'gsis_residual')
    call gsis_check(gsis_counters(h, kstep, tt, nsit), 'gsis_counters')
    write(*, '(a,1x,i0,4(1x,es24.16))') 'res', kstep, sres
  end do
  call gsis_check(gsis_download_macro(h, c_loc(s), c_null_ptr, c_loc(cf), c_null_ptr, c_null_ptr, &
                                      c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr), 'gsis_download_macro')
  write(*, '(a,2(1x,es24.16))') 'sum', sum(s(1:4, 0:nx, 0:ny)), sum(cf(0:nx, 0:ny))
  if (rmode == 'save1') call write_backup(1)
  if (rmode == 'save2') call write_backup(2)
  call gsis_check(gsis_destroy(h), 'gsis_destroy')

contains

  ! backup_Macro_write(flag) with backup_Mesco_write: B%FRD as one unformatted
  ! record, then flag.dat = 1, tt, Kstep (one rank: star00/restart00000.dat)
  subroutine write_backup(flag)
    integer, intent(in) :: flag
    character(len=64) :: fname
    integer :: un
    call gsis_check(gsis_download_frd(h, frd), 'gsis_download_frd')
    call gsis_check(gsis_counters(h, kstep, tt, nsit), 'gsis_counters')
    write(fname, '(a,i1,a)') 'restart', flag, '/star00/restart00000.dat'
    open(newunit=un, file=trim(fname), form='unformatted', status='replace')
    write(un) frd
    close(un)
    write(fname, '(a,i1,a)') 'restart', flag, '/flag.dat'
    open(newunit=un, file=trim(fname), status='replace')
    write(un, *) 1
    write(un, *) tt
    write(un, *) kstep
    close(un)
  end subroutine write_backup

  ! init_flow_read + backup_Mesco_read: restart2 when its flag is 1, else
  ! restart1; then the kinetic BCs and moments of the restored distribution
  ! and the saved counters
  subroutine read_backup()
    character(len=64) :: fname
    integer :: un, flag, slot, ios
    integer(c_int32_t) :: ks
    real(c_double) :: t
    slot = 0
    do k = 2, 1, -1
      write(fname, '(a,i1,a)') 'restart', k, '/flag.dat'
      open(newunit=un, file=trim(fname), action='read', status='old', iostat=ios)
      if (ios /= 0) cycle
      read(un, *) flag
      read(un, *) t
      read(un, *) ks
      close(un)
      if (flag == 1) then
        slot = k
        exit
      end if
    end do
    if (slot == 0) then
      write(*, '(a)') 'No backup data is available'
      stop 1
    end if
    write(fname, '(a,i1,a)') 'restart', slot, '/star00/restart00000.dat'
    open(newunit=un, file=trim(fname), form='unformatted', action='read', status='old')
    read(un) frd
    close(un)
    call gsis_check(gsis_upload_frd(h, frd), 'gsis_upload_frd')
    call gsis_check(gsis_refresh_moments(h), 'gsis_refresh_moments')
    call gsis_check(gsis_set_counters(h, ks, t), 'gsis_set_counters')
  end subroutine read_backup

  ! ghost nodes of the physical boundaries by linear extrapolation, in BC
  ! order, then the corners (src/sub_init.f90:777-810)
  subroutine ghost_nodes(a)
    real(c_double), intent(inout) :: a(-1:, -1:)
    integer :: b, m, ib, ie, jb, je, i1, j1, i2, j2, ii, jj
    do b = 1, 4
      ib = bc(b)%ist; ie = bc(b)%iend; jb = bc(b)%jst; je = bc(b)%jend
      do m = 1, 2
        select case (bc(b)%face)
        case (1)
          a(ib - m, jb:je) = 2d0 * a(ib, jb:je) - a(ib + m, jb:je)
        case (2)
          a(ib:ie, jb - m) = 2d0 * a(ib:ie, jb) - a(ib:ie, jb + m)
        case (3)
          a(ie + m, jb:je) = 2d0 * a(ie, jb:je) - a(ie - m, jb:je)
        case default
          a(ib:ie, je + m) = 2d0 * a(ib:ie, je) - a(ib:ie, je - m)
        end select
      end do
    end do
    do jj = 1, 2
      do ii = 1, 2
        i1 = 1 - ii; j1 = 1 - jj; i2 = nx + ii; j2 = ny + jj
        a(i1, j1) = a(i1, 1) + a(1, j1) - a(1, 1)
        a(i2, j2) = a(i2, ny) + a(nx, j2) - a(nx, ny)
        a(i1, j2) = a(i1, ny) + a(1, j2) - a(1, ny)
        a(i2, j1) = a(i2, 1) + a(nx, j1) - a(nx, 1)
      end do
    end do
  end subroutine ghost_nodes

  ! cell volumes, face lengths and normals of Creat_Mesh (src/sub_init.f90:183-208)
  subroutine geometry()
    real(c_double) :: dx, dy, sl
    allocate(vol(nx, ny), si(nx, ny), sj(nx, ny), ni1(nx, ny), ni2(nx, ny), nj1(nx, ny), nj2(nx, ny))
    vol = 0d0; si = 0d0; sj = 0d0; ni1 = 0d0; ni2 = 0d0; nj1 = 0d0; nj2 = 0d0
    do j = 1, ny - 1
      do i = 1, nx - 1
        vol(i, j) = abs((x(i, j) - x(i + 1, j + 1)) * (y(i + 1, j) - y(i, j + 1)) &
                        - (x(i + 1, j) - x(i, j + 1)) * (y(i, j) - y(i + 1, j + 1))) * 0.5d0
      end do
    end do
    do j = 1, ny - 1
      do i = 1, nx
        dx = x(i, j + 1) - x(i, j); dy = y(i, j + 1) - y(i, j)
        sl = sqrt(dx * dx + dy * dy)
        si(i, j) = sl; ni1(i, j) = dy / sl; ni2(i, j) = -dx / sl
      end do
    end do
    do j = 1, ny
      do i = 1, nx - 1
        dx = x(i + 1, j) - x(i, j); dy = y(i + 1, j) - y(i, j)
        sl = sqrt(dx * dx + dy * dy)
        sj(i, j) = sl; nj1(i, j) = -dy / sl; nj2(i, j) = dx / sl
      end do
    end do
  end subroutine geometry

  ! v = vmax sinh(a s)/sinh(a) on a uniform s-grid of [-1,1]; weights =
  ! composite Newton-Cotes weight in s (Simpson 1/3, plus 3/8 on the last
  ! three intervals when the interval count is odd) times dv/ds
  subroutine sinh_grid(m, vmax, a, v, w)
    integer, intent(in) :: m
    real(c_double), intent(in) :: vmax, a
    real(c_double), intent(out) :: v(m), w(m)
    real(c_double) :: step, sk
    integer :: q, n13, intervals
    step = 2d0 / (m - 1)
    w = 0d0
    intervals = m - 1
    n13 = intervals
    if (mod(intervals, 2) == 1) n13 = intervals - 3
    do q = 0, n13 - 1, 2
      w(q + 1) = w(q + 1) + step / 3d0
      w(q + 2) = w(q + 2) + 4d0 * step / 3d0
      w(q + 3) = w(q + 3) + step / 3d0
    end do
    if (mod(intervals, 2) == 1) then
      q = n13
      w(q + 1) = w(q + 1) + 3d0 * step / 8d0
      w(q + 2) = w(q + 2) + 9d0 * step / 8d0
      w(q + 3) = w(q + 3) + 9d0 * step / 8d0
      w(q + 4) = w(q + 4) + 3d0 * step / 8d0
    end if
    do q = 1, m
      sk = -1d0 + (q - 1) * step
      if (q == m) sk = 1d0
      v(q) = vmax * sinh(a * sk) / sinh(a)
      w(q) = w(q) * vmax * a * cosh(a * sk) / sinh(a)
    end do
  end subroutine sinh_grid

end program gsis_cavity_host
