!===============================================================================
!
! amun_dev_mod.F90 -- Fortran 2003 binding of the C-ABI in include/amun_dev.h
!
! The derived type amun_cfg is interoperable with the C struct of the same
! name; every entry point is declared with bind(C) and takes plain values,
! arrays and the opaque context pointer.  Block arrays are passed in the
! reference layout u(nv,im,jm,km) per block (blocks.F90:1120-1191).
!
!===============================================================================
!
module amun_dev_mod

  use iso_c_binding

  implicit none

  integer(c_int32_t), parameter :: AMUN_EQSYS_HD = 1, AMUN_EQSYS_MHD = 2
  integer(c_int32_t), parameter :: AMUN_EQSYS_SRHD = 3, AMUN_EQSYS_SRMHD = 4
  integer(c_int32_t), parameter :: AMUN_EOS_ADI = 1
  integer(c_int32_t), parameter :: AMUN_EOS_ISO = 2
  integer(c_int32_t), parameter :: AMUN_REC_TVD = 1, AMUN_REC_WENO5Z = 5
  integer(c_int32_t), parameter :: AMUN_REC_WENO3 = 2, AMUN_REC_LIMO3 = 3
  integer(c_int32_t), parameter :: AMUN_REC_PPM = 4
  integer(c_int32_t), parameter :: AMUN_REC_WENO5YC = 6
  integer(c_int32_t), parameter :: AMUN_REC_WENO5NS = 7
  integer(c_int32_t), parameter :: AMUN_REC_MP5 = 11
  integer(c_int32_t), parameter :: AMUN_REC_MP7 = 12
  integer(c_int32_t), parameter :: AMUN_REC_GP = 20
  integer(c_int32_t), parameter :: AMUN_REC_CRWENO5Z = 8, AMUN_REC_CRWENO5YC = 9
  integer(c_int32_t), parameter :: AMUN_REC_CRWENO5NS = 10
  integer(c_int32_t), parameter :: AMUN_REC_CRMP5 = 13, AMUN_REC_CRMP5LD = 14
  integer(c_int32_t), parameter :: AMUN_REC_CRMP7 = 15, AMUN_REC_MGP = 21
  integer(c_int32_t), parameter :: AMUN_LIM_ZERO = 0, AMUN_LIM_MINMOD = 1
  integer(c_int32_t), parameter :: AMUN_LIM_MC = 2, AMUN_LIM_SUPERBEE = 3
  integer(c_int32_t), parameter :: AMUN_LIM_VANLEER = 4, AMUN_LIM_VANALBADA = 5
  integer(c_int32_t), parameter :: AMUN_RS_HLL = 1, AMUN_RS_HLLC = 2
  integer(c_int32_t), parameter :: AMUN_RS_HLLD = 3, AMUN_RS_ROE = 4, AMUN_RS_HLLDM = 5
  integer(c_int32_t), parameter :: AMUN_TA_EULER = 0, AMUN_TA_RK2 = 1
  integer(c_int32_t), parameter :: AMUN_TA_RK3 = 2, AMUN_TA_SSPRK_M2 = 3
  integer(c_int32_t), parameter :: AMUN_TA_SSPRK43 = 4, AMUN_TA_SSPRK53 = 5
  integer(c_int32_t), parameter :: AMUN_PREC_FP64 = 0, AMUN_PREC_MIXED = 1
  integer(c_int32_t), parameter :: AMUN_C2P_1DW = 0, AMUN_C2P_2DWV = 1
  integer(c_int32_t), parameter :: AMUN_C2P_2DWU = 2
  integer(c_int32_t), parameter :: AMUN_BND_PERIODIC = 0, AMUN_BND_OPEN = 1
  integer(c_int32_t), parameter :: AMUN_BND_OUTFLOW = 2, AMUN_BND_REFLECTIVE = 3

  type, bind(C) :: amun_cfg
    integer(c_int32_t) :: ndims, eqsys, eos, nv, ncells, nghosts
    integer(c_int32_t) :: xblocks, yblocks, zblocks
    integer(c_int32_t) :: reconstruction, limiter, riemann_solver
    integer(c_int32_t) :: time_advance, precision, fix_positivity
    integer(c_int32_t) :: clip_extrema
    real(c_double)     :: gamma, cfl, eps, alpha
    real(c_double)     :: xmin, xmax, ymin, ymax, zmin, zmax
    real(c_double)     :: tolerance, pmin
    integer(c_int32_t) :: nr_maxit, nr_extra
    real(c_double)     :: sgp, kappa, kbeta
    integer(c_int32_t) :: ngp, stages
    integer(c_int32_t) :: prol_limiter, primitive_solver
    real(c_double)     :: limo3_rad, ppm_const
    integer(c_int32_t) :: bnd_type(6)
    integer(c_int32_t) :: mlp_limiting, fix_unphysical_cells
    integer(c_int32_t) :: glm_source_terms, reserved_cfg
    real(c_double)     :: viscosity, resistivity
    real(c_double)     :: csnd
    integer(c_int32_t) :: states_4vec, reserved_cfg2
    integer(c_int32_t) :: gravity, reserved_cfg3
    real(c_double)     :: gacc
  end type amun_cfg

  interface

    subroutine amun_dev_default_cfg(cfg) bind(C, name="amun_dev_default_cfg")
      import :: amun_cfg
      type(amun_cfg), intent(out) :: cfg
    end subroutine

    integer(c_int) function amun_dev_init(cfg, ctx) bind(C, name="amun_dev_init")
      import :: amun_cfg, c_int, c_ptr
      type(amun_cfg), intent(in) :: cfg
      type(c_ptr), intent(out)   :: ctx
    end function

    ! a frozen multi-level mesh: meta(4, nblk) = level, i0, j0, k0 per leaf
    ! (the mesh module's leaf list after generate_mesh / update_mesh)
    integer(c_int) function amun_dev_init_amr(cfg, nblk, meta, ctx) &
                                               bind(C, name="amun_dev_init_amr")
      import :: amun_cfg, c_int, c_int32_t, c_ptr
      type(amun_cfg), intent(in)        :: cfg
      integer(c_int32_t), value         :: nblk
      integer(c_int32_t), intent(in)    :: meta(4, *)
      type(c_ptr), intent(out)          :: ctx
    end function

    subroutine amun_dev_finalize(ctx) bind(C, name="amun_dev_finalize")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine

    integer(c_int) function amun_dev_upload_blocks(ctx, u) &
                                          bind(C, name="amun_dev_upload_blocks")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value                      :: ctx
      real(c_double), dimension(*), intent(in) :: u
    end function

    integer(c_int) function amun_dev_download_blocks(ctx, u, q) &
                                        bind(C, name="amun_dev_download_blocks")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value                        :: ctx
      real(c_double), dimension(*), intent(out) :: u
      type(c_ptr), value                        :: q
    end function

    integer(c_int) function amun_dev_new_time_step(ctx, dtnext, dt, cmax) &
                                          bind(C, name="amun_dev_new_time_step")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value         :: ctx
      real(c_double), value      :: dtnext
      real(c_double), intent(out) :: dt, cmax
    end function

    integer(c_int) function amun_dev_step(ctx, dtnext, dt, cmax) &
                                                   bind(C, name="amun_dev_step")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value         :: ctx
      real(c_double), value      :: dtnext
      real(c_double), intent(out) :: dt, cmax
    end function

    integer(c_int) function amun_dev_run(ctx, nsteps, dtnext, dts) &
                                                    bind(C, name="amun_dev_run")
      import :: c_int, c_ptr, c_double, c_int32_t
      type(c_ptr), value                        :: ctx
      integer(c_int32_t), value                 :: nsteps
      real(c_double), value                     :: dtnext
      real(c_double), dimension(*), intent(out) :: dts
    end function

    integer(c_int) function amun_dev_integrals(ctx, out) &
                                              bind(C, name="amun_dev_integrals")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value                        :: ctx
      real(c_double), dimension(8), intent(out) :: out
    end function

    integer(c_int) function amun_dev_statistics(ctx, out) &
                                             bind(C, name="amun_dev_statistics")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value                         :: ctx
      real(c_double), dimension(21), intent(out) :: out
    end function

    integer(c_int) function amun_dev_get_time(ctx, time, step) &
                                               bind(C, name="amun_dev_get_time")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value              :: ctx
      real(c_double), intent(out)     :: time
      integer(c_int64_t), intent(out) :: step
    end function

    integer(c_int) function amun_dev_block_coords(ctx, coords) &
                                           bind(C, name="amun_dev_block_coords")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value                            :: ctx
      integer(c_int32_t), dimension(*), intent(out) :: coords
    end function

    integer(c_int) function amun_dev_c2p_events(ctx, n) &
                                             bind(C, name="amun_dev_c2p_events")
      import :: c_int, c_ptr, c_int64_t
      type(c_ptr), value                             :: ctx
      integer(c_int64_t), dimension(3), intent(out)  :: n
    end function

    integer(c_int) function amun_dev_set_stage_fusion(ctx, enable) &
                                       bind(C, name="amun_dev_set_stage_fusion")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value        :: ctx
      integer(c_int32_t), value :: enable
    end function

    type(c_ptr) function amun_dev_last_error_c() &
                                             bind(C, name="amun_dev_last_error")
      import :: c_ptr
    end function

  end interface

contains

! the message of the last failing C-ABI call as a Fortran string
  function amun_dev_last_error() result(msg)
    character(len=256)              :: msg
    character(kind=c_char), pointer :: p(:)
    type(c_ptr)                     :: cp
    integer                         :: i
    msg = ""
    cp = amun_dev_last_error_c()
    if (.not. c_associated(cp)) return
    call c_f_pointer(cp, p, [256])
    do i = 1, 256
      if (p(i) == c_null_char) exit
      msg(i:i) = p(i)
    end do
  end function amun_dev_last_error

end module amun_dev_mod
