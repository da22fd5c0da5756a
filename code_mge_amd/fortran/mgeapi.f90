! mgeapi.f90 -- iso_c_binding interface of libmge (include/mge.h) for a
! Fortran driver.  Every entry point keeps the granularity of the code_mge
! routine it replaces (timestep, euler, bcond, bc_cutcell, movco, movfin,
! solver), so the driver's per-block calls bind one to one.
module mgeapi
  use iso_c_binding
  implicit none

  ! scheme and freestream parameters (mge_params)
  type, bind(C) :: mge_params_t
    integer(c_int)  :: mstage
    real(c_double)  :: cstp(8), cdis(8)
    real(c_double)  :: cfl
    integer(c_int)  :: vt
    real(c_double)  :: vis2, vis4, smoo
    integer(c_int)  :: mgtype
    real(c_double)  :: gamma, rho0, p0, c0, ei0, u0, v0, w0, h0
    integer(c_int)  :: eqntype, dissip
    real(c_double)  :: viscl, prlmean, gas_const, drop
    integer(c_int)  :: distype, precision
    real(c_double)  :: epsentr
    integer(c_int)  :: execmode
    real(c_double)  :: dtstep
    integer(c_int)  :: smoother
    real(c_double)  :: ffreeze
    real(c_double)  :: lfreeze
  end type

  ! what a jobsheet defines (mge_job)
  type, bind(C) :: mge_job_t
    character(kind=c_char) :: job(64), model(64), mesh(512), surface(512)
    integer(c_int)  :: nlevel, mcyc(16)
    real(c_double)  :: mach, reynolds, gamma, prlmean, alphad, betad
    real(c_double)  :: cflf, vis2, vis4, smoo, epsentr
    integer(c_int)  :: mgtype, mstage
    real(c_double)  :: cstp(8), cdis(8)
    integer(c_int)  :: eqntype, distype, dissip
    integer(c_int)  :: n(3)
    real(c_double)  :: lo(3), hi(3), dom_lo(3), dom_hi(3)
    integer(c_int)  :: bctype(6)
    integer(c_int)  :: iflow
    integer(c_int)  :: execmode, nperiod, nstep
    real(c_double)  :: dt, pmean, tmean, rgasmean
  end type

  ! one wall record (mge_cutrec): the txcut fields bc_cutcell/efluxb/dfluxb read
  type, bind(C) :: mge_cutrec_t
    integer(c_int)  :: faceid, zz, ccid(3), scid(3)
    real(c_double)  :: psurf(3), nvec(3), radcur(2), rr
    integer(c_int)  :: npoint
    integer(c_int)  :: ijk(3, 8)
    real(c_double)  :: coef(8)
    real(c_double)  :: psol, wsol(5)
  end type

  interface
    ! ---- input files (make_fsijob's readers)
    integer(c_int) function mge_read_job(path, job) bind(C, name="mge_read_job")
      import :: c_int, c_char, mge_job_t
      character(kind=c_char), intent(in) :: path(*)
      type(mge_job_t), intent(out) :: job
    end function
    integer(c_int) function mge_job_params(job, precision, prm) bind(C, name="mge_job_params")
      import :: c_int, mge_job_t, mge_params_t
      type(mge_job_t), intent(in) :: job
      integer(c_int), value :: precision
      type(mge_params_t), intent(out) :: prm
    end function
    integer(c_int) function mge_job_level(job, lvl, n, ratio) bind(C, name="mge_job_level")
      import :: c_int, mge_job_t
      type(mge_job_t), intent(in) :: job
      integer(c_int), value :: lvl
      integer(c_int), intent(out) :: n(3), ratio(3)
    end function
    integer(c_int) function mge_read_p3d(path, cap, tri, ntri) bind(C, name="mge_read_p3d")
      import :: c_int, c_char, c_ptr
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: cap
      type(c_ptr), value :: tri
      integer(c_int), intent(out) :: ntri
    end function
    ! ---- geometry preprocessor (make_fsijob + solid_ipolcoef)
    integer(c_int) function mge_geom_level_tri(tri, ntri, ni, nj, nk, xv, yv, zv, blank, recs, cap, nmir) &
        bind(C, name="mge_geom_level_tri")
      import :: c_int, c_double, c_ptr
      real(c_double), intent(in) :: tri(*)
      integer(c_int), value :: ntri, ni, nj, nk
      real(c_double), intent(in) :: xv(*), yv(*), zv(*)
      type(c_ptr), value :: blank, recs
      integer(c_int), value :: cap
      integer(c_int), intent(out) :: nmir
    end function
    ! ---- case lifetime (make_flow / init_metric / init_flow)
    integer(c_int) function mge_case_create(prm, nlevel, h) bind(C, name="mge_case_create")
      import :: c_int, c_ptr, mge_params_t
      type(mge_params_t), intent(in) :: prm
      integer(c_int), value :: nlevel
      type(c_ptr), intent(out) :: h
    end function
    integer(c_int) function mge_case_destroy(h) bind(C, name="mge_case_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function mge_level_define(h, lvl, ni, nj, nk, xv, yv, zv, cellvolume, bctype, ratio, blank, &
                                             nmir, recs, mxcyc) bind(C, name="mge_level_define")
      import :: c_int, c_double, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: lvl, ni, nj, nk
      real(c_double), intent(in) :: xv(*), yv(*), zv(*)
      real(c_double), value :: cellvolume
      integer(c_int), intent(in) :: bctype(6), ratio(3)
      type(c_ptr), value :: blank
      integer(c_int), value :: nmir
      type(c_ptr), value :: recs
      integer(c_int), value :: mxcyc
    end function
    integer(c_int) function mge_init_flow(h) bind(C, name="mge_init_flow")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    ! ---- the solve step, per block
    integer(c_int) function mge_timestep(h, lvl, blk) bind(C, name="mge_timestep")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: lvl, blk
    end function
    integer(c_int) function mge_euler(h, lvl, blk, itime, ncyc, mgl, mlvl, lres, lfir, rtrms) &
        bind(C, name="mge_euler")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: lvl, blk, itime, ncyc, mgl, mlvl, lres, lfir
      type(c_ptr), value :: rtrms
    end function
    integer(c_int) function mge_bcond(h, lvl, blk) bind(C, name="mge_bcond")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: lvl, blk
    end function
    integer(c_int) function mge_bc_cutcell(h, lvl, blk) bind(C, name="mge_bc_cutcell")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: lvl, blk
    end function
    integer(c_int) function mge_bc_overlap(h, lvl) bind(C, name="mge_bc_overlap")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: lvl
    end function
    integer(c_int) function mge_movco(h, mgl) bind(C, name="mge_movco")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: mgl
    end function
    integer(c_int) function mge_movfin(h, mgl, mlvl) bind(C, name="mge_movfin")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
      integer(c_int), value :: mgl, mlvl
    end function
    integer(c_int) function mge_shift_time_levels(h) bind(C, name="mge_shift_time_levels")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    integer(c_int) function mge_solve(h, mcyc, reslog, cap, nlog, cell_stages) bind(C, name="mge_solve")
      import :: c_int, c_ptr, c_double, c_long_long
      type(c_ptr), value :: h
      integer(c_int), intent(in) :: mcyc(*)
      real(c_double) :: reslog(*)
      integer(c_int), value :: cap
      integer(c_int), intent(out) :: nlog
      integer(c_long_long), intent(out) :: cell_stages
    end function
    ! ---- results
    integer(c_int) function mge_download(h, lvl, blk, field, out) bind(C, name="mge_download")
      import :: c_int, c_ptr, c_char, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: lvl, blk
      character(kind=c_char), intent(in) :: field(*)
      real(c_double) :: out(*)
    end function
    integer(c_int) function mge_write_solut(h, path, format) bind(C, name="mge_write_solut")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: h
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: format
    end function
    integer(c_int) function mge_read_solut(h, path, format) bind(C, name="mge_read_solut")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: h
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: format
    end function
    integer(c_int) function mge_upload_solut(h, lvl, blk, p_tot, solut, nval) bind(C, name="mge_upload_solut")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: lvl, blk
      real(c_double), value :: p_tot
      type(c_ptr), value :: solut
      integer(c_int), intent(out) :: nval
    end function
    integer(c_int) function mge_wall_force(h, lvl, out) bind(C, name="mge_wall_force")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: h
      integer(c_int), value :: lvl
      real(c_double), intent(out) :: out(3)
    end function
    integer(c_int) function mge_synchronize(h) bind(C, name="mge_synchronize")
      import :: c_int, c_ptr
      type(c_ptr), value :: h
    end function
    type(c_ptr) function mge_last_error() bind(C, name="mge_last_error")
      import :: c_ptr
    end function
    type(c_ptr) function mge_input_error() bind(C, name="mge_input_error")
      import :: c_ptr
    end function
  end interface

contains

  ! a NUL-terminated copy of a Fortran string
  function cstr(s) result(c)
    character(len=*), intent(in) :: s
    character(kind=c_char) :: c(len_trim(s) + 1)
    integer :: i
    do i = 1, len_trim(s)
      c(i) = s(i:i)
    end do
    c(len_trim(s) + 1) = c_null_char
  end function

  ! a Fortran string from a NUL-terminated char array
  function fstr(a) result(s)
    character(kind=c_char), intent(in) :: a(:)
    character(len=size(a)) :: s
    integer :: i
    s = ' '
    do i = 1, size(a)
      if (a(i) == c_null_char) exit
      s(i:i) = a(i)
    end do
  end function

  ! the message behind a C pointer (mge_last_error / mge_input_error)
  function cmsg(p) result(s)
    type(c_ptr), intent(in) :: p
    character(len=256) :: s
    character(kind=c_char), pointer :: a(:)
    integer :: i
    s = ' '
    if (.not. c_associated(p)) return
    call c_f_pointer(p, a, [256])
    do i = 1, 256
      if (a(i) == c_null_char) exit
      s(i:i) = a(i)
    end do
  end function

end module mgeapi
