!===============================================================================
!
! amun_host.F90 -- Fortran 2003 host driver of the MI355X block update
!
! The host keeps what the reference keeps on the CPU -- parameters, the block
! list with its u(nv,im,jm,km) arrays, problem set-up and I/O -- and hands the
! whole time step to the device through the C-ABI (module amun_dev_mod).  It
! mirrors the reference driver: read the parameter file given with -i
! (parameters.F90:99-195: "key = value", # comments, quotes, last occurrence
! wins), set up the problem on every block, then loop
!   advance(dtnext); time = time + dt        (driver.F90:571-609)
! for nmax steps or until tmax, and report the time per step
! (driver.F90:730-793).
!
! Problems: "blast" (the seeded 2D/3D hydrodynamic blast: rho = 1, v = 0,
! p = 10 inside r < 0.1 else 0.1, density * (1 + 0.01 (2 xi - 1))),
! "orszag_tang" (2D MHD vortex with the same density noise) and "file"
! (conserved state read from "input" as raw FP64 U(nx,ny,nz,nv)).  The noise
! xi comes from the counter-based SplitMix64 generator, bit-identical to the
! Python side (amun_amd/problems.py).
!
! Outputs: "output" (raw FP64 U(nx,ny,nz,nv) after the run, optional) and a
! summary on stdout.
!
!===============================================================================
!
program amun_host

  use iso_c_binding
  use iso_fortran_env, only : int64, real64, error_unit
  use amun_dev_mod

  implicit none

  integer, parameter :: mpar = 256
  character(len=64)  :: pnames(mpar)
  character(len=256) :: pvalues(mpar)
  integer            :: npar = 0

  type(amun_cfg) :: cfg
  type(c_ptr)    :: ctx
  integer        :: iret, nv, nc, ng, im, km, nblk, nx, ny, nz, kn
  integer        :: nmax, s, b, i, j, k, p
  integer(int64) :: seed, c0, c1, crate, c2p(3)
  real(real64)   :: dt, cmax, time, tmax, dtnext, elapsed, ints(8), stats(21)
  character(len=256) :: fin, fout, problem, str
  real(c_double), allocatable, target :: u(:,:,:,:,:)
  integer(c_int32_t), allocatable :: coords(:,:)
  real(real64), allocatable :: ug(:,:,:,:)

  call read_params()

  call amun_dev_default_cfg(cfg)
  cfg%ndims = get_int("ndims", 3)
  str = get_str("equation_system", "hd")
  select case(trim(str))
  case("hd", "HD", "hydrodynamic", "HYDRODYNAMIC")
    cfg%eqsys = AMUN_EQSYS_HD
  case("mhd", "MHD", "magnetohydrodynamic", "MAGNETOHYDRODYNAMIC")
    cfg%eqsys = AMUN_EQSYS_MHD
  case("srhd", "SRHD")
    cfg%eqsys = AMUN_EQSYS_SRHD
  case("srmhd", "SRMHD")
    cfg%eqsys = AMUN_EQSYS_SRMHD
  case default
    cfg%eqsys = -1
  end select
  str = get_str("equation_of_state", "adi")
  select case(trim(str))
  case("adi", "ADI", "adiabatic", "ADIABATIC")
    cfg%eos = AMUN_EOS_ADI
  case("iso", "ISO", "isothermal", "ISOTHERMAL")
    cfg%eos = AMUN_EOS_ISO
  case default
    cfg%eos = -1
  end select
  cfg%csnd = get_real("csnd", 1.0d0)
  str = get_str("reconstruction", "tvd")
  select case(trim(str))
  case("tvd", "TVD")
    cfg%reconstruction = AMUN_REC_TVD
  case("weno5z", "weno5-z", "WENO5Z", "WENO5-Z")
    cfg%reconstruction = AMUN_REC_WENO5Z
  case("gp", "GP")
    cfg%reconstruction = AMUN_REC_GP
  case("weno5yc", "weno5-yc", "WENO5YC", "WENO5-YC")
    cfg%reconstruction = AMUN_REC_WENO5YC
  case("weno5ns", "weno5-ns", "WENO5NS", "WENO5-NS")
    cfg%reconstruction = AMUN_REC_WENO5NS
  case("mp5", "MP5")
    cfg%reconstruction = AMUN_REC_MP5
  case("mp7", "MP7")
    cfg%reconstruction = AMUN_REC_MP7
  case("weno3", "WENO3")
    cfg%reconstruction = AMUN_REC_WENO3
  case("limo3", "LIMO3", "LimO3")
    cfg%reconstruction = AMUN_REC_LIMO3
  case("ppm", "PPM")
    cfg%reconstruction = AMUN_REC_PPM
  case("crweno5z", "crweno5-z", "CRWENO5Z", "CRWENO5-Z")
    cfg%reconstruction = AMUN_REC_CRWENO5Z
  case("crweno5yc", "crweno5-yc", "CRWENO5YC", "CRWENO5-YC")
    cfg%reconstruction = AMUN_REC_CRWENO5YC
  case("crweno5ns", "crweno5-ns", "CRWENO5NS", "CRWENO5-NS")
    cfg%reconstruction = AMUN_REC_CRWENO5NS
  case("crmp5", "CRMP5")
    cfg%reconstruction = AMUN_REC_CRMP5
  case("crmp5l", "crmp5ld", "CRMP5L", "CRMP5LD")
    cfg%reconstruction = AMUN_REC_CRMP5LD
  case("crmp7", "CRMP7")
    cfg%reconstruction = AMUN_REC_CRMP7
  case("mgp", "MGP")
    cfg%reconstruction = AMUN_REC_MGP
  case default
    cfg%reconstruction = -1
  end select
  str = get_str("limiter", "mm")
  select case(trim(str))
  case("mm", "minmod")
    cfg%limiter = AMUN_LIM_MINMOD
  case("mc", "monotonized_central")
    cfg%limiter = AMUN_LIM_MC
  case("sb", "superbee")
    cfg%limiter = AMUN_LIM_SUPERBEE
  case("vl", "vanleer")
    cfg%limiter = AMUN_LIM_VANLEER
  case("va", "vanalbada")
    cfg%limiter = AMUN_LIM_VANALBADA
  case default
    cfg%limiter = AMUN_LIM_ZERO
  end select
  str = get_str("prolongation_limiter", "mm")
  select case(trim(str))
  case("mm", "minmod")
    cfg%prol_limiter = AMUN_LIM_MINMOD
  case("mc", "monotonized_central")
    cfg%prol_limiter = AMUN_LIM_MC
  case("sb", "superbee")
    cfg%prol_limiter = AMUN_LIM_SUPERBEE
  case("vl", "vanleer")
    cfg%prol_limiter = AMUN_LIM_VANLEER
  case default
    cfg%prol_limiter = AMUN_LIM_ZERO
  end select
  cfg%bnd_type(1) = bnd_of(get_str("xlbndry", "periodic"))
  cfg%bnd_type(2) = bnd_of(get_str("xubndry", "periodic"))
  cfg%bnd_type(3) = bnd_of(get_str("ylbndry", "periodic"))
  cfg%bnd_type(4) = bnd_of(get_str("yubndry", "periodic"))
  cfg%bnd_type(5) = bnd_of(get_str("zlbndry", "periodic"))
  cfg%bnd_type(6) = bnd_of(get_str("zubndry", "periodic"))
  str = get_str("primitive_solver", "1Dw")
  select case(trim(str))
  case("1Dw", "1dw", "1DW", "1D(w)", "1D(W)")
    cfg%primitive_solver = AMUN_C2P_1DW
  case("2dwv", "2Dwv", "2D(w,v)", "2D(W,v)")
    cfg%primitive_solver = AMUN_C2P_2DWV
  case("2dwu", "2Dwu", "2D(w,u)", "2D(W,u)")
    cfg%primitive_solver = AMUN_C2P_2DWU
  case default
    cfg%primitive_solver = -1
  end select
  str = get_str("riemann_solver", "hlld")
  select case(trim(str))
  case("hll", "HLL")
    cfg%riemann_solver = AMUN_RS_HLL
  case("hllc", "HLLC")
    cfg%riemann_solver = AMUN_RS_HLLC
  case("hlld", "HLLD")
    cfg%riemann_solver = AMUN_RS_HLLD
  case("roe", "ROE")
    cfg%riemann_solver = AMUN_RS_ROE
  case("hlldm", "hlld-m", "HLLDM", "HLLD-M")
    cfg%riemann_solver = AMUN_RS_HLLDM
  case default
    cfg%riemann_solver = -1
  end select
  str = get_str("time_advance", "rk2")
  select case(trim(str))
  case("euler", "EULER")
    cfg%time_advance = AMUN_TA_EULER
  case("rk2", "RK2")
    cfg%time_advance = AMUN_TA_RK2
  case("rk3", "RK3")
    cfg%time_advance = AMUN_TA_RK3
  case("ssprk(m,2)", "SSPRK(m,2)")
    cfg%time_advance = AMUN_TA_SSPRK_M2
  case("rk3.4", "RK3.4", "ssprk(4,3)", "SSPRK(4,3)")
    cfg%time_advance = AMUN_TA_SSPRK43
  case("rk3.5", "RK3.5", "ssprk(5,3)", "SSPRK(5,3)")
    cfg%time_advance = AMUN_TA_SSPRK53
  case default
    cfg%time_advance = -1
  end select
  str = get_str("precision", "fp64")
  cfg%precision = merge(AMUN_PREC_MIXED, AMUN_PREC_FP64, trim(str) == "mixed")
  cfg%gamma   = get_real("gamma", 5.0d0 / 3.0d0)
  cfg%cfl     = get_real("cfl", 0.5d0)
  cfg%eps     = get_real("eps", epsilon(1.0d0))
  cfg%alpha   = get_real("alpha", 2.0d0)
  cfg%ncells  = get_int("ncells", 8)
  cfg%nghosts = get_int("nghosts", 2)
  cfg%xblocks = get_int("xblocks", 1)
  cfg%yblocks = get_int("yblocks", 1)
  cfg%zblocks = get_int("zblocks", 1)
  cfg%xmin = get_real("xmin", 0.0d0)
  cfg%xmax = get_real("xmax", 1.0d0)
  cfg%ymin = get_real("ymin", 0.0d0)
  cfg%ymax = get_real("ymax", 1.0d0)
  cfg%zmin = get_real("zmin", 0.0d0)
  cfg%zmax = get_real("zmax", 1.0d0)
  cfg%tolerance = get_real("tolerance", 1.0d-10)
  cfg%pmin      = get_real("pmin", 1.0d-16)
  cfg%nr_maxit  = get_int("nr_maxit", 100)
  cfg%nr_extra  = get_int("nr_extra", 2)
  cfg%sgp       = get_real("sgp", 9.0d0)
  cfg%kappa     = get_real("kappa", 1.0d0)
  cfg%kbeta     = get_real("kbeta", 1.0d0)
  cfg%ngp       = get_int("ngp", 5)
  cfg%stages    = get_int("stages", 2)
  cfg%limo3_rad = get_real("limo3_rad", 0.5d0)
  cfg%ppm_const = get_real("ppm_const", 1.25d0)
  cfg%fix_positivity       = on_of(get_str("fix_positivity", "off"))
  cfg%clip_extrema         = on_of(get_str("clip_extrema", "off"))
  cfg%mlp_limiting         = on_of(get_str("mlp_limiting", "off"))
  cfg%fix_unphysical_cells = on_of(get_str("fix_unphysical_cells", "off"))
  str = get_str("glm_source_terms", "none")
  select case(trim(str))
  case("eglm", "EGLM")
    cfg%glm_source_terms = 1
  case("heglm", "HEGLM")
    cfg%glm_source_terms = 2
  case default
    cfg%glm_source_terms = 0
  end select
  cfg%viscosity   = get_real("viscosity", 0.0d0)
  cfg%resistivity = get_real("resistivity", 0.0d0)
  nmax    = get_int("nmax", 10)
  tmax    = get_real("tmax", 1.0d30)
  seed    = int(get_int("seed", 20261202), int64)
  problem = get_str("problem", "blast")
  ! gravity.F90:141-150: the Rayleigh-Taylor problem turns gravity on, with
  ! the acceleration (0, gacc, 0)
  select case(trim(problem))
  case("rt", "rayleightaylor", "rayleigh-taylor")
    cfg%gravity = 1
  case default
    cfg%gravity = 0
  end select
  cfg%gacc = get_real("gacc", -1.0d0)
  fin     = get_str("input", "")
  fout    = get_str("output", "")

  nv   = merge(5, 9, cfg%eqsys == AMUN_EQSYS_HD .or. cfg%eqsys == AMUN_EQSYS_SRHD)
  if (cfg%eos == AMUN_EOS_ISO) nv = nv - 1
  nc   = cfg%ncells
  ng   = cfg%nghosts
  im   = nc + 2 * ng
  km   = merge(im, 1, cfg%ndims == 3)
  kn   = merge(nc, 1, cfg%ndims == 3)
  nblk = cfg%xblocks * cfg%yblocks * merge(cfg%zblocks, 1_c_int32_t, cfg%ndims == 3)
  nx   = cfg%xblocks * nc
  ny   = cfg%yblocks * nc
  nz   = merge(cfg%zblocks * nc, 1, cfg%ndims == 3)
  allocate(u(nv,im,im,km,nblk), coords(3,nblk), ug(nx,ny,nz,nv))

! problem set-up on the global grid (host side, as setup_problem does per
! block in the reference)
  select case(trim(problem))
  case("file")
    call read_state(trim(fin))
  case("orszag_tang")
    call setup_orszag_tang()
  case("srblast")
    call setup_srblast()
  case default
    call setup_blast()
  end select
  str = get_str("dump_initial", "")
  if (len_trim(str) > 0) then
    call write_state(trim(str))
    stop
  end if

  iret = amun_dev_init(cfg, ctx)
  if (iret /= 0) then
    write(error_unit, "(a,i0,2a)") "amun_dev_init failed with code ", iret, &
                                   ": ", trim(amun_dev_last_error())
    stop 2
  end if
  ! one launch per stage (bitwise the per-direction launches' results) where
  ! the library builds it; the configurations it does not cover return 600
  ! and keep the per-direction launches
  str = get_str("stage_fusion", "on")
  if (trim(str) == "on" .or. trim(str) == "yes" .or. trim(str) == "true") then
    iret = amun_dev_set_stage_fusion(ctx, 1)
    if (iret /= 0 .and. iret /= 600) call die("set_stage_fusion")
  end if
  iret = amun_dev_block_coords(ctx, coords)
  call scatter()
  iret = amun_dev_upload_blocks(ctx, u)
  if (iret /= 0) call die("upload")

  time   = 0.0d0
  dtnext = 2.0d0 * tmax
  iret = amun_dev_new_time_step(ctx, dtnext, dt, cmax)
  call system_clock(c0, crate)
  s = 0
  do while (s < nmax .and. time < tmax)
    iret = amun_dev_step(ctx, dtnext, dt, cmax)
    if (iret /= 0) call die("step")
    time = time + dt
    s = s + 1
  end do
  call system_clock(c1)
  elapsed = real(c1 - c0, real64) / real(crate, real64)

  iret = amun_dev_integrals(ctx, ints)
  iret = amun_dev_statistics(ctx, stats)
  iret = amun_dev_download_blocks(ctx, u, c_null_ptr)
  if (iret /= 0) call die("download")
  call gather()
  if (len_trim(fout) > 0) call write_state(trim(fout))

  write(*,"(a,i0,a,es22.15)") "steps = ", s, "  time = ", time
  write(*,"(a,es22.15)") "last dt = ", dt
  write(*,"(a,8es22.14)") "integrals = ", ints
  write(*,"(a,21es16.8)") "statistics = ", stats
  if (cfg%eqsys == AMUN_EQSYS_SRMHD .or. cfg%eqsys == AMUN_EQSYS_SRHD) then
    iret = amun_dev_c2p_events(ctx, c2p)
    write(*,"(a,3(1x,i0))") "primitive recovery failed / p<=0 / not converged =", c2p
  end if
  if (s > 0) then
    write(*,"(a,es12.5,a)") "TIME PER STEP = ", elapsed / s, " s"
    write(*,"(a,es12.5)") "zone-updates/s = ", real(nx, real64) * ny * nz * s / elapsed
  end if

  call amun_dev_finalize(ctx)

contains

  subroutine die(what)
    character(len=*), intent(in) :: what
    write(error_unit, "(3a)") what, " failed: ", trim(amun_dev_last_error())
    call amun_dev_finalize(ctx)
    stop 3
  end subroutine die

! ---------------------------------------------------------------- parameters
  subroutine read_params()
    character(len=512) :: line, fname
    integer :: l, u, ios, e, h
    fname = "params.in"
    do l = 1, command_argument_count()
      call get_command_argument(l, line)
      if (trim(line) == "-i" .or. trim(line) == "--input") then
        call get_command_argument(l + 1, fname)
      end if
    end do
    open(newunit=u, file=trim(fname), status="old", action="read", iostat=ios)
    if (ios /= 0) then
      write(error_unit, "(3a)") "cannot open parameter file '", trim(fname), "'"
      stop 1
    end if
    do
      read(u, "(a)", iostat=ios) line
      if (ios /= 0) exit
      h = index(line, "#")
      if (h > 0) line = line(1:h-1)
      e = index(line, "=")
      if (e == 0) cycle
      call set_param(adjustl(line(1:e-1)), adjustl(line(e+1:)))
    end do
    close(u)
  end subroutine read_params

  subroutine set_param(name, value)
    character(len=*), intent(in) :: name, value
    character(len=256) :: v
    integer :: n
    v = trim(value)
    n = len_trim(v)
    if (n >= 2) then
      if ((v(1:1) == '"' .and. v(n:n) == '"') .or. &
          (v(1:1) == "'" .and. v(n:n) == "'")) v = v(2:n-1)
    end if
    do n = 1, npar
      if (trim(pnames(n)) == trim(name)) then
        pvalues(n) = v
        return
      end if
    end do
    if (npar < mpar) then
      npar = npar + 1
      pnames(npar)  = trim(name)
      pvalues(npar) = v
    end if
  end subroutine set_param

  function get_str(name, def) result(r)
    character(len=*), intent(in) :: name, def
    character(len=256) :: r
    integer :: n
    r = def
    do n = 1, npar
      if (trim(pnames(n)) == name) r = pvalues(n)
    end do
  end function get_str

! an on/off switch of the parameter file
  function on_of(name) result(t)
    character(len=*), intent(in) :: name
    integer(c_int32_t)           :: t
    select case(trim(name))
    case("on", "ON", "On", "yes", "YES", "y", "Y", "t", "T", "true", ".true.", "1")
      t = 1
    case default
      t = 0
    end select
  end function on_of

! the boundary type names of boundaries.F90:170-260
  function bnd_of(name) result(t)
    character(len=*), intent(in) :: name
    integer(c_int32_t)           :: t
    select case(trim(name))
    case("periodic")
      t = AMUN_BND_PERIODIC
    case("open")
      t = AMUN_BND_OPEN
    case("outflow", "out")
      t = AMUN_BND_OUTFLOW
    case("reflective", "reflecting", "reflect")
      t = AMUN_BND_REFLECTIVE
    case default
      t = -1
    end select
  end function bnd_of

  integer function get_int(name, def)
    character(len=*), intent(in) :: name
    integer, intent(in)          :: def
    character(len=256) :: r
    integer :: ios
    r = get_str(name, "")
    get_int = def
    if (len_trim(r) > 0) read(r, *, iostat=ios) get_int
  end function get_int

  real(real64) function get_real(name, def)
    character(len=*), intent(in) :: name
    real(real64), intent(in)     :: def
    character(len=256) :: r
    integer :: ios
    r = get_str(name, "")
    get_real = def
    if (len_trim(r) > 0) read(r, *, iostat=ios) get_real
  end function get_real

! ------------------------------------------------------------ random numbers
! SplitMix64 (Steele, Lea & Flood 2014) in wrapping 64-bit arithmetic
  integer(int64) function splitmix64(x)
    integer(int64), intent(in) :: x
    integer(int64) :: z
    z = x + transfer(int(z'9E3779B97F4A7C15', int64), 0_int64)
    z = ieor(z, ishft(z, -30)) * transfer(int(z'BF58476D1CE4E5B9', int64), 0_int64)
    z = ieor(z, ishft(z, -27)) * transfer(int(z'94D049BB133111EB', int64), 0_int64)
    splitmix64 = ieor(z, ishft(z, -31))
  end function splitmix64

! n uniform doubles in [0,1): the 53 high bits of SplitMix64(key + i)
  subroutine uniform(stream, n, xi)
    integer, intent(in)       :: stream, n
    real(real64), intent(out) :: xi(n)
    integer(int64) :: key
    integer :: m
    key = splitmix64(ieor(seed, ishft(int(stream, int64), 40)))
    do m = 1, n
      xi(m) = real(ishft(splitmix64(key + int(m - 1, int64)), -11), real64) &
              * (1.0d0 / 9007199254740992.0d0)
    end do
  end subroutine uniform

! ------------------------------------------------------------------ problems
  subroutine cons_from_prim(q)
    real(real64), intent(in) :: q(nx,ny,nz,nv)
    real(real64) :: ek(nx,ny,nz), g
    g = cfg%gamma
    if (cfg%eqsys == AMUN_EQSYS_SRMHD .or. cfg%eqsys == AMUN_EQSYS_SRHD) then
      call cons_from_prim_sr(q)
      return
    end if
    ug(:,:,:,1) = q(:,:,:,1)
    do p = 2, 4
      ug(:,:,:,p) = q(:,:,:,1) * q(:,:,:,p)
    end do
    if (cfg%eos == AMUN_EOS_ISO) then
      ! prim2cons_hd_iso / prim2cons_mhd_iso (equations.F90:1360, 2220)
      if (nv == 8) ug(:,:,:,5:8) = q(:,:,:,5:8)
      return
    end if
    ek = 0.5d0 * (ug(:,:,:,2) * q(:,:,:,2) + ug(:,:,:,3) * q(:,:,:,3) &
                + ug(:,:,:,4) * q(:,:,:,4))
    if (nv == 5) then
      ug(:,:,:,5) = q(:,:,:,5) / (g - 1.0d0) + ek
    else
      ug(:,:,:,5:8) = q(:,:,:,5:8)
      ug(:,:,:,9) = (q(:,:,:,9) / (g - 1.0d0) + ek) + 0.5d0 * (q(:,:,:,5)**2 &
                    + q(:,:,:,6)**2 + q(:,:,:,7)**2)
    end if
  end subroutine cons_from_prim

! prim2cons_srmhd_adi (equations.F90:4590-4653) and prim2cons_srhd_adi
! (3478-3532) on the global grid
  subroutine cons_from_prim_sr(q)
    real(real64), intent(in) :: q(nx,ny,nz,nv)
    real(real64) :: vv, bb, vb, vm, wt, gx
    gx = (cfg%gamma - 1.0d0) / cfg%gamma
    if (nv == 5) then
      do k = 1, nz
        do j = 1, ny
          do i = 1, nx
            vv = (q(i,j,k,2)**2 + q(i,j,k,3)**2) + q(i,j,k,4)**2
            vm = 1.0d0 - vv
            wt = (q(i,j,k,1) + q(i,j,k,5) / gx) / vm
            ug(i,j,k,1) = q(i,j,k,1) / sqrt(vm)
            do p = 2, 4
              ug(i,j,k,p) = wt * q(i,j,k,p)
            end do
            ug(i,j,k,5) = (wt - q(i,j,k,5)) - ug(i,j,k,1)
          end do
        end do
      end do
      return
    end if
    do k = 1, nz
      do j = 1, ny
        do i = 1, nx
          vv = (q(i,j,k,2)**2 + q(i,j,k,3)**2) + q(i,j,k,4)**2
          bb = (q(i,j,k,5)**2 + q(i,j,k,6)**2) + q(i,j,k,7)**2
          vb = (q(i,j,k,2) * q(i,j,k,5) + q(i,j,k,3) * q(i,j,k,6)) &
             + q(i,j,k,4) * q(i,j,k,7)
          vm = 1.0d0 - vv
          wt = (q(i,j,k,1) + q(i,j,k,9) / gx) / vm + bb
          ug(i,j,k,1) = q(i,j,k,1) / sqrt(vm)
          do p = 2, 4
            ug(i,j,k,p) = wt * q(i,j,k,p) - vb * q(i,j,k,p+3)
          end do
          ug(i,j,k,5:8) = q(i,j,k,5:8)
          ug(i,j,k,9) = ((wt - q(i,j,k,9)) - ug(i,j,k,1)) &
                        - 0.5d0 * (vm * bb + vb * vb)
        end do
      end do
    end do
  end subroutine cons_from_prim_sr

! relativistic (magnetised) blast: rho = 1, v = 0, p = 1 for r < 0.1 else
! 0.01, Bx = 0.5 and psi = 0 with a field
  subroutine setup_srblast()
    real(real64), allocatable :: q(:,:,:,:)
    real(real64) :: x, y, dx, dy
    allocate(q(nx,ny,nz,nv))
    dx = (cfg%xmax - cfg%xmin) / nx
    dy = (cfg%ymax - cfg%ymin) / ny
    q = 0.0d0
    do k = 1, nz
      do j = 1, ny
        do i = 1, nx
          x = cfg%xmin + (i - 0.5d0) * dx
          y = cfg%ymin + (j - 0.5d0) * dy
          q(i,j,k,1) = 1.0d0
          if (nv == 9) q(i,j,k,5) = 0.5d0
          q(i,j,k,nv) = merge(1.0d0, 0.01d0, sqrt(x * x + y * y) < 0.1d0)
        end do
      end do
    end do
    call cons_from_prim(q)
  end subroutine setup_srblast

  subroutine setup_blast()
    real(real64), allocatable :: q(:,:,:,:), xi(:)
    real(real64) :: x, y, z, r, dx, dy, dz
    allocate(q(nx,ny,nz,nv), xi(nx*ny*nz))
    call uniform(0, nx*ny*nz, xi)
    dx = (cfg%xmax - cfg%xmin) / nx
    dy = (cfg%ymax - cfg%ymin) / ny
    dz = (cfg%zmax - cfg%zmin) / max(nz, 1)
    q = 0.0d0
    do k = 1, nz
      do j = 1, ny
        do i = 1, nx
          x = cfg%xmin + (i - 0.5d0) * dx
          y = cfg%ymin + (j - 0.5d0) * dy
          z = merge(cfg%zmin + (k - 0.5d0) * dz, 0.0d0, cfg%ndims == 3)
          r = sqrt(x * x + y * y + merge(z * z, 0.0d0, cfg%ndims == 3))
          q(i,j,k,1) = 1.0d0 * (1.0d0 + 0.01d0 * (2.0d0 * xi(i + nx * (j - 1 + ny * (k - 1))) - 1.0d0))
          q(i,j,k,nv) = merge(10.0d0, 0.1d0, r < 0.1d0)
        end do
      end do
    end do
    call cons_from_prim(q)
  end subroutine setup_blast

  subroutine setup_orszag_tang()
    real(real64), allocatable :: q(:,:,:,:), xi(:)
    real(real64) :: x, y, dx, dy, pi, b0
    pi = acos(-1.0d0)
    b0 = 1.0d0 / sqrt(4.0d0 * pi)
    allocate(q(nx,ny,nz,nv), xi(nx*ny*nz))
    call uniform(5, nx*ny*nz, xi)
    dx = (cfg%xmax - cfg%xmin) / nx
    dy = (cfg%ymax - cfg%ymin) / ny
    q = 0.0d0
    do k = 1, nz
      do j = 1, ny
        do i = 1, nx
          x = cfg%xmin + (i - 0.5d0) * dx
          y = cfg%ymin + (j - 0.5d0) * dy
          q(i,j,k,1) = 25.0d0 / (36.0d0 * pi) * (1.0d0 + 0.01d0 * (2.0d0 &
                       * xi(i + nx * (j - 1 + ny * (k - 1))) - 1.0d0))
          q(i,j,k,2) = -sin(2.0d0 * pi * y)
          q(i,j,k,3) =  sin(2.0d0 * pi * x)
          if (nv == 9) then
            q(i,j,k,5) = -b0 * sin(2.0d0 * pi * y)
            q(i,j,k,6) =  b0 * sin(4.0d0 * pi * x)
          end if
          q(i,j,k,nv) = 5.0d0 / (12.0d0 * pi)
        end do
      end do
    end do
    call cons_from_prim(q)
  end subroutine setup_orszag_tang

! --------------------------------------------------------------- global I/O
  subroutine read_state(fname)
    character(len=*), intent(in) :: fname
    integer :: un, ios
    open(newunit=un, file=fname, access="stream", form="unformatted", &
         status="old", iostat=ios)
    if (ios /= 0) then
      write(error_unit, "(3a)") "cannot read state '", fname, "'"
      stop 4
    end if
    read(un) ug
    close(un)
  end subroutine read_state

  subroutine write_state(fname)
    character(len=*), intent(in) :: fname
    integer :: un
    open(newunit=un, file=fname, access="stream", form="unformatted", &
         status="replace")
    write(un) ug
    close(un)
  end subroutine write_state

! global grid <-> block interiors (ghost layers are filled on the device)
  subroutine scatter()
    integer :: i0, j0, k0
    u = 0.0d0
    do b = 1, nblk
      i0 = coords(1,b) * nc
      j0 = coords(2,b) * nc
      k0 = coords(3,b) * kn
      do p = 1, nv
        u(p,ng+1:ng+nc,ng+1:ng+nc,merge(ng+1,1,cfg%ndims==3):merge(ng+nc,1,cfg%ndims==3),b) = &
          ug(i0+1:i0+nc,j0+1:j0+nc,k0+1:k0+kn,p)
      end do
    end do
  end subroutine scatter

  subroutine gather()
    integer :: i0, j0, k0
    do b = 1, nblk
      i0 = coords(1,b) * nc
      j0 = coords(2,b) * nc
      k0 = coords(3,b) * kn
      do p = 1, nv
        ug(i0+1:i0+nc,j0+1:j0+nc,k0+1:k0+kn,p) = &
          u(p,ng+1:ng+nc,ng+1:ng+nc,merge(ng+1,1,cfg%ndims==3):merge(ng+nc,1,cfg%ndims==3),b)
      end do
    end do
  end subroutine gather

end program amun_host
