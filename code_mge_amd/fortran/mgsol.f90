! mgsol.f90 -- the single-rank mgsol driver on the MI355X library.
!
! Follows srcf90/main.f90: read the job (make_fsijob: jobsheet, .pin,
! .p3d surface), build the cut-cell geometry and the multigrid blocks,
! initialise the flow, run the solver and report.  The solver is the
! control of srcf90/solver.f90:110-290, restated here on the library's
! per-block calls (timestep, euler, movco, movfin), exactly as the reference
! sequences them: an FMG ramp over the levels with mcyc cycles each, V or W
! cycles (mgtype), the residual log <job>_res.dat (level, cycle, residual in
! (2i6,e18.10)) and the drop test; for unsteady jobs (runtype = 1) that ramp
! once per physical time step (period x step of them, time level itime
! from 0), each followed by the time-level shift and the step's "(I6)" line
! in the log.  The restart
! file <job>_restart.dat is read when the jobsheet sets iflow = 1 and
! written at the end (main.f90:78,107), with a full-precision .bin beside it.
!
! usage: mgsol [jobsheet [precision]]    precision 64 (default) or 32
program mgsol
  use iso_c_binding
  use mgeapi
  implicit none
  character(len=512) :: jobfile, arg
  type(mge_job_t) :: job
  type(mge_params_t) :: prm
  type(c_ptr) :: h
  integer(c_int) :: lvl, n(3), ratio(3), ntri, nmir, prec, mxc, nlog, i, itime, nsteps, nval
  real(c_double), allocatable, target :: tri(:), rlog(:, :), solut(:, :)
  real(c_double), allocatable :: xv(:), yv(:), zv(:)
  integer(c_int8_t), allocatable, target :: blank(:)
  type(mge_cutrec_t), allocatable, target :: recs(:)
  real(c_double) :: hx, hy, hz, F(3), aref, q, ca, sa, cb, sb
  character(len=64) :: jobname
  real :: t0, t1

  jobfile = 'jobsheet'
  prec = 64
  if (command_argument_count() >= 1) call get_command_argument(1, jobfile)
  if (command_argument_count() >= 2) then
    call get_command_argument(2, arg)
    read (arg, *) prec
  end if

  ! ---- make_fsijob: job, grid and surface
  call check_in(mge_read_job(cstr(jobfile), job), 'reading the job')
  call check_in(mge_job_params(job, prec, prm), 'deriving the parameters')
  jobname = fstr(job%job)
  ntri = 0
  if (len_trim(fstr(job%surface)) > 0) then
    call check_in(mge_read_p3d(cstr(fstr(job%surface)), 0, c_null_ptr, ntri), 'reading the surface')
    allocate (tri(9*max(ntri, 1)))
    call check_in(mge_read_p3d(cstr(fstr(job%surface)), ntri, c_loc(tri), ntri), 'reading the surface')
  else
    allocate (tri(9))
  end if
  print '(a,a)', ' job      : ', trim(jobname)
  print '(a,3i6,a,i3,a,i7,a)', ' grid     : ', job%n, ' cells,', job%nlevel, ' levels,', ntri, ' surface triangles'
  print '(a,f8.4,a,f9.2,a,i3)', ' mach     : ', job%mach, '  reynolds: ', job%reynolds, '  precision: ', prec
  if (prm%execmode == 1) print '(a,i5,a,i5,a,es12.5)', ' unsteady : ', job%nperiod, ' periods x', job%nstep, &
    ' steps, dtstep ', prm%dtstep

  ! ---- make_flow / init_metric / solid_ipolcoef per level
  call check(mge_case_create(prm, job%nlevel, h), 'creating the case')
  mxc = maxval(job%mcyc(1:job%nlevel)) + 2
  do lvl = 1, job%nlevel
    call check_in(mge_job_level(job, lvl, n, ratio), 'level dimensions')
    allocate (xv(0:n(1) + 3), yv(0:n(2) + 3), zv(0:n(3) + 3))
    call vertices(job%lo(1), job%hi(1), n(1), xv, hx)
    call vertices(job%lo(2), job%hi(2), n(2), yv, hy)
    call vertices(job%lo(3), job%hi(3), n(3), zv, hz)
    allocate (blank(n(1)*n(2)*n(3)))
    call check(mge_geom_level_tri(tri, ntri, n(1), n(2), n(3), xv, yv, zv, c_loc(blank), c_null_ptr, 0, nmir), &
               'cut-cell geometry')
    allocate (recs(max(nmir, 1)))
    call check(mge_geom_level_tri(tri, ntri, n(1), n(2), n(3), xv, yv, zv, c_loc(blank), c_loc(recs), nmir, nmir), &
               'cut-cell geometry')
    call check(mge_level_define(h, lvl, n(1), n(2), n(3), xv, yv, zv, hx*hy*hz, job%bctype, ratio, c_loc(blank), &
                                nmir, c_loc(recs), mxc), 'defining a level')
    print '(a,i2,a,3i6,a,i7,a)', ' level ', lvl, ':', n, ' cells,', nmir, ' wall records'
    deallocate (xv, yv, zv, blank, recs)
  end do
  ! init_flow (main.f90:78): freestream, or the restart file when iflow = 1
  call check(mge_init_flow(h), 'init_flow')
  if (job%iflow == 1) then
    call check(mge_read_solut(h, cstr(trim(jobname)//'_restart.dat'), 0), 'read_solut')
    print '(a,a)', ' restart from ', trim(jobname)//'_restart.dat'
  end if

  ! ---- solver: the physical time loop (solver.f90:110-112, 263-290); one
  ! pass for steady flow
  nsteps = 1
  if (prm%execmode == 1) nsteps = max(job%nperiod, 1)*max(job%nstep, 1)
  allocate (rlog(3, (sum(job%mcyc(1:job%nlevel)) + 1)*nsteps))
  call cpu_time(t0)
  nlog = 0
  do itime = 0, nsteps - 1
    call solver(h, job%nlevel, job%mcyc, prm%mgtype, prm%drop, itime, rlog, nlog)
    if (prm%execmode == 1) then
      call check(mge_shift_time_levels(h), 'shift_time_levels')
      nlog = nlog + 1
      rlog(:, nlog) = [dble(job%nlevel), -1.0d0, 0.0d0]
    end if
  end do
  call check(mge_synchronize(h), 'synchronising')
  call cpu_time(t1)
  print '("Time = ",f16.8," seconds  SOLVER ")', t1 - t0

  open (unit=11, file=trim(jobname)//'_res.dat', status='replace')
  do i = 1, nlog
    if (rlog(2, i) < 0.0d0) then
      write (11, '(I6)') int(rlog(1, i))
    else
      write (11, '(2i6,e18.10)') int(rlog(1, i)), int(rlog(2, i)), rlog(3, i)
    end if
  end do
  close (11)
  open (unit=12, file=trim(jobname)//'_res_full.dat', status='replace')
  do i = 1, nlog
    if (rlog(2, i) < 0.0d0) then
      write (12, '(I6)') int(rlog(1, i))
    else
      write (12, '(2i6,es26.17)') int(rlog(1, i)), int(rlog(2, i)), rlog(3, i)
    end if
  end do

  ! ---- forces on the finest level (FP64 wall-face momentum flux)
  call check(mge_wall_force(h, job%nlevel, F), 'wall force')
  aref = frontal_area(tri, ntri)
  q = 0.5d0*prm%rho0*(prm%u0*prm%u0 + prm%v0*prm%v0 + prm%w0*prm%w0)
  ca = cos(job%alphad*(acos(-1.0d0)/180.0d0))
  sa = sin(job%alphad*(acos(-1.0d0)/180.0d0))
  cb = cos(job%betad*(acos(-1.0d0)/180.0d0))
  sb = sin(job%betad*(acos(-1.0d0)/180.0d0))
  write (12, '(a,3es26.17)') ' force ', F
  close (12)
  print '(a,3es16.8)', ' wall force   : ', F
  if (aref > 0.0d0) then
    print '(a,f12.8,a,f12.8,a,es10.3)', ' Cd = ', (F(1)*ca*cb + F(2)*sa*cb + F(3)*sb)/(q*aref), &
      '   Cl = ', (-F(1)*sa + F(2)*ca)/(q*aref), '   reference area ', aref
  end if
  ! upload_solut (solver.f90:279): the output fields of the finest level,
  ! reported as their ranges
  call check(mge_upload_solut(h, job%nlevel, 1, 0.0d0, c_null_ptr, nval), 'upload_solut')
  call check_in(mge_job_level(job, job%nlevel, n, ratio), 'level dimensions')
  allocate (solut(n(1)*n(2)*n(3), nval))
  call check(mge_upload_solut(h, job%nlevel, 1, 0.0d0, c_loc(solut), nval), 'upload_solut')
  print '(a,2f10.5,a,f8.5,a,es11.3)', ' cp range     : ', minval(solut(:, 5)), maxval(solut(:, 5)), &
    '   max Mach ', maxval(solut(:, 6)), '   max pt-loss ', maxval(abs(solut(:, 7)))
  ! write_solut (main.f90:107), and the full-precision variant
  call check(mge_write_solut(h, cstr(trim(jobname)//'_restart.dat'), 0), 'write_solut')
  call check(mge_write_solut(h, cstr(trim(jobname)//'_restart.bin'), 1), 'write_solut')
  call check(mge_case_destroy(h), 'destroying the case')

contains

  ! x(i) = h (i - 2) + lo for i >= 1 (index 0 unused): vertex i is the low
  ! face of cell i, interior cells 2..n+1
  subroutine vertices(lo, hi, n, v, h)
    real(c_double), intent(in) :: lo, hi
    integer(c_int), intent(in) :: n
    real(c_double), intent(out) :: v(0:), h
    integer :: i
    h = (hi - lo)/dble(n)
    v(0) = 0.0d0
    do i = 1, n + 3
      v(i) = h*(dble(i) - 2.0d0) + lo
    end do
  end subroutine

  ! projected bounding-box area of the surface normal to x
  function frontal_area(t, nt) result(a)
    real(c_double), intent(in) :: t(:)
    integer(c_int), intent(in) :: nt
    real(c_double) :: a, ymin, ymax, zmin, zmax
    integer :: m, c
    a = 0.0d0
    if (nt < 1) return
    ymin = huge(1.0d0); ymax = -huge(1.0d0); zmin = huge(1.0d0); zmax = -huge(1.0d0)
    do m = 0, nt - 1
      do c = 0, 2
        ymin = min(ymin, t(9*m + 3*c + 2)); ymax = max(ymax, t(9*m + 3*c + 2))
        zmin = min(zmin, t(9*m + 3*c + 3)); zmax = max(zmax, t(9*m + 3*c + 3))
      end do
    end do
    a = (ymax - ymin)*(zmax - zmin)
  end function

  ! solver.f90:119-261, one physical time step at time level itime: FMG
  ! ramp, the residual of the finest level after each cycle, max_res reset
  ! to 1000 per cycle as :166-176 do, the NaN stop (:180) and the drop test
  ! (:238); log rows are appended after nlog
  subroutine solver(h, nlevel, mcyc, mgtype, drop, itime, rlog, nlog)
    type(c_ptr), intent(in) :: h
    integer(c_int), intent(in) :: nlevel, mcyc(:), mgtype, itime
    real(c_double), intent(in) :: drop
    real(c_double), intent(inout) :: rlog(:, :)
    integer(c_int), intent(inout) :: nlog
    integer :: mlvl, ncyc
    real(c_double) :: r, ini_res, max_res
    ini_res = 0.0d0
    do mlvl = 1, nlevel
      do ncyc = 1, mcyc(mlvl)
        call cycle(h, nlevel, mlvl, ncyc, mgtype, itime, r)
        max_res = 1000.0d0
        if (ncyc == 1) ini_res = r
        if (r < max_res) max_res = r
        if (r /= r) then
          print *, 'solver: residual is NaN'
          stop 1
        end if
        nlog = nlog + 1
        rlog(:, nlog) = [dble(mlvl), dble(ncyc), max_res]
        if (mod(ncyc, 50) == 0 .or. ncyc == 1) print '(a,i3,a,i6,a,es14.6)', ' Level=', mlvl, ' NCYC=', ncyc, ' RESID=', max_res
        if (mlvl > 1 .and. ini_res/max_res > drop) exit
      end do
      if (mlvl < nlevel) call check(mge_movfin(h, mlvl + 1, mlvl), 'movfin')
    end do
  end subroutine

  ! one multigrid cycle with finest level mlvl (solver.f90:187-233): relax
  ! and residual pass going down with movco, relax going up with movfin;
  ! W cycles revisit a level mgtype times
  subroutine cycle(h, nlevel, mlvl, ncyc, mgtype, itime, r)
    type(c_ptr), intent(in) :: h
    integer(c_int), intent(in) :: nlevel, mlvl, ncyc, mgtype, itime
    real(c_double), intent(out), target :: r
    integer :: icyc(nlevel), mgl, lfir
    icyc = 0
    mgl = mlvl
    do
      do
        icyc(mgl) = icyc(mgl) + 1
        lfir = 1
        if (mgl == mlvl) lfir = 0
        call check(mge_timestep(h, mgl, 1), 'timestep')
        if (mgl == mlvl) then
          call check(mge_euler(h, mgl, 1, itime, ncyc, mgl, mlvl, 0, lfir, c_loc(r)), 'euler')
        else
          call check(mge_euler(h, mgl, 1, itime, ncyc, mgl, mlvl, 0, lfir, c_null_ptr), 'euler')
        end if
        if (mlvl == 1) return
        call check(mge_euler(h, mgl, 1, itime, ncyc, mgl, mlvl, 1, 0, c_null_ptr), 'euler (residual)')
        mgl = mgl - 1
        call check(mge_movco(h, mgl), 'movco')
        if (mgl == 1) exit
      end do
      lfir = 1
      do
        call check(mge_timestep(h, mgl, 1), 'timestep')
        call check(mge_euler(h, mgl, 1, itime, ncyc, mgl, mlvl, 0, lfir, c_null_ptr), 'euler')
        lfir = 0
        mgl = mgl + 1
        call check(mge_movfin(h, mgl, mlvl), 'movfin')
        if (mgl == mlvl) return
        if (icyc(mgl) /= mgtype) exit
        icyc(mgl) = 0
      end do
    end do
  end subroutine

  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc < 0) then
      print '(a,a,a,i4,a,a)', ' ERROR in ', what, ': code ', rc, ': ', trim(cmsg(mge_last_error()))
      stop 2
    else if (rc > 0) then
      print '(a,a,a,i4)', ' warning in ', what, ': code ', rc
    end if
  end subroutine

  subroutine check_in(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= 0) then
      print '(a,a,a,a)', ' ERROR ', what, ': ', trim(cmsg(mge_input_error()))
      stop 2
    end if
  end subroutine

end program mgsol
