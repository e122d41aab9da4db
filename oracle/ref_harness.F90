!===============================================================================
!
! ref_harness.F90 -- TEST INFRASTRUCTURE ONLY (never part of the product path)
!
! A driver program written for this project that links the UNMODIFIED AMUN
! reference modules (compiled straight from the reference sources by
! oracle/build_ref.sh into oracle/_ref/) and runs them on seeded inputs read
! from raw little-endian FP64 files.  Its outputs are the golden vectors that
! pin the C oracle (oracle/amun_oracle.c) and, through it, the HIP kernels.
!
! The initialisation order mirrors the reference driver (driver.F90:157-477)
! and the time loop mirrors driver.F90:571-581 (advance(dtnext), then
! time = time + dt).  Parameters are read by the reference parameter parser
! (parameters.F90:99-195) from the file given with "-i".  Extra keys used only
! by this harness (ignored by every reference module):
!
!   harness_mode   = run | ghosts | amr | flux | recon | eqs | c2p | algebra
!   harness_input  = input file (raw FP64)
!   harness_output = output file prefix
!   harness_steps  = number of steps (run)
!   harness_count  = number of blocks / pencils (flux, recon)
!   harness_n      = pencil length (recon) or vector length (eqs)
!   harness_cmax   = value of cmax used by the GLM fluxes (flux, eqs)
!   harness_criterion = 1: also write check_refinement_criterion of every
!                    block after the steps (amr)
!
! File layouts (all raw FP64, C order written as Fortran order):
!   run   in : U(nx,ny,nz,nv)            conserved variables of the whole grid
!         out: <prefix>_u.bin  U(nx,ny,nz,nv) after harness_steps steps
!              <prefix>_dt.bin (dt, cmax) for each step
!              <prefix>_time.txt wall-clock seconds of the step loop
!   ghosts in: as run; out: <prefix>_q.bin every block's q with ghost layers
!              after the initial boundary_variables, preceded by its offsets
!   flux  in : count x [ q(nv,im,jm,km), dx(NDIMS) ]
!         out: <prefix>_f.bin  count x f(NDIMS,nv,im,jm,km)
!   recon in : f(n,count), h
!         out: <prefix>_fl.bin fl(n,count), <prefix>_fr.bin fr(n,count)
!   eqs   in : q(nv,n)
!         out: <prefix>_eqs.bin u(nv,n), q2(nv,n), f(nv,n), cm(n), cp(n)
!   c2p   in : u(nv,n), q(nv,n) (initial content of the output)
!         out: <prefix>_c2p.bin q(nv,n) after cons2prim
!   algebra in: a(5,n); out: <prefix>_alg.bin r(19,n) (see mode_algebra)
!   amr   in : none (the reference's generate_mesh refines the reference
!              problem named by "problem" up to maxlev)
!         out: <prefix>_mesh0.bin per data block: level, i0, j0, k0 (cell
!              offsets at the block's level, as doubles), u(nv,nc,nc,nc)
!              interior, for the initial state (q recomputed from u)
!              <prefix>_q0.bin per data block q(nv,im,jm,km) with ghosts
!              after the initial boundary_variables
!              <prefix>_mesh1.bin the same as mesh0 after harness_steps steps
!              <prefix>_dt.bin (dt, cmax) for each step
!              <prefix>_crit.bin with harness_criterion = 1: per data block
!              check_refinement_criterion (refinement.F90:257) after the
!              steps, as doubles, in the order of mesh1
!              <prefix>_mesh2.bin with harness_update = 1: the blocks after
!              one update_mesh (mesh.F90:756-833) that follows the steps,
!              in the layout of mesh0 and the order of the new data list
!         With harness_amr_input = 1 the block interiors are replaced by
!         harness_input, written in the layout of <prefix>_mesh0.bin without
!         the four leading numbers, before q is recomputed.
!         The mesh is frozen after generation (toplev = 1 skips update_mesh
!         in advance, evolution.F90:329), so the steps run on a fixed
!         multi-level mesh with the reference's restriction, prolongation
!         and flux correction.
!
!===============================================================================
!
program amun_ref_harness

  use algebra       , only : quadratic, quadratic_normalized
  use algebra       , only : cubic, cubic_normalized, quartic
  use blocks        , only : initialize_blocks, block_meta, block_data
  use blocks        , only : list_data
  use boundaries    , only : initialize_boundaries, boundary_variables
  use coordinates   , only : initialize_coordinates
  use coordinates   , only : im, jm, km, ib, ie, jb, je, kb, ke, nc
  use coordinates   , only : adx, ady, adz, xmin, ymin, zmin, ir, jr, kr
  use coordinates   , only : toplev, maxlev
  use equations     , only : initialize_equations, nv, cmax, cmax2
  use equations     , only : prim2cons, cons2prim, fluxspeed
  use equations     , only : update_primitive_variables
  use evolution     , only : initialize_evolution, advance, new_time_step
  use evolution     , only : dt, time, step
  use gravity       , only : initialize_gravity
  use integrals     , only : initialize_integrals, store_integrals
  use integrals     , only : finalize_integrals
  use interpolations, only : initialize_interpolations, reconstruct
  use io            , only : initialize_io
  use mesh          , only : initialize_mesh, generate_mesh, update_mesh
  use mpitools      , only : initialize_mpitools
  use operators     , only : initialize_operators
  use parameters    , only : read_parameters, get_parameter_integer
  use parameters    , only : get_parameter_real, get_parameter_string
  use problems      , only : initialize_problems
  use random        , only : initialize_random
  use refinement    , only : initialize_refinement, check_refinement_criterion
  use schemes       , only : initialize_schemes, update_flux
  use shapes        , only : initialize_shapes
  use sources       , only : initialize_sources
  use user_problem  , only : initialize_user_problem

  implicit none

  character(len=255) :: mode = "run", fin = "in.bin", fout = "out"
  integer            :: iret = 0, nsteps = 1, ncount = 1, n = 1
  real(kind=8)       :: hcmax = 1.0d+00, dtnext, tmax = 1.0d+30
  integer            :: nrun = 1, amr_input = 0, crit_out = 0, mesh_update = 0

  call initialize_mpitools()
  call read_parameters(iret)
  if (iret > 0) stop 1

  call get_parameter_string ("harness_mode"  , mode  )
  call get_parameter_string ("harness_input" , fin   )
  call get_parameter_string ("harness_output", fout  )
  call get_parameter_integer("harness_steps" , nsteps)
  call get_parameter_integer("harness_count" , ncount)
  call get_parameter_integer("harness_n"     , n     )
  call get_parameter_real   ("harness_cmax"  , hcmax )
  call get_parameter_integer("harness_amr_input", amr_input)
  call get_parameter_integer("harness_criterion", crit_out)
  call get_parameter_integer("harness_update", mesh_update)
  call get_parameter_real   ("tmax"          , tmax  )
  dtnext = 2.0d+00 * tmax

  call initialize_random(1, 0)
  call initialize_coordinates(.false., iret)
  if (iret > 0) stop 2
  call initialize_equations(.false., iret)
  if (iret > 0) stop 3
  call initialize_user_problem(.false., iret)
  call initialize_refinement(.false., iret)
  call initialize_evolution(.false., iret)
  if (iret > 0) stop 4
  call initialize_schemes(.false., iret)
  if (iret > 0) stop 5
  call initialize_interpolations(.false., iret)
  if (iret > 0) stop 6
  call initialize_operators(.false., iret)
  call initialize_blocks(.false., iret)
  call initialize_boundaries(.false., iret)
  call initialize_problems(.false., iret)
  call initialize_shapes(.false., iret)
  call initialize_gravity(.false., iret)
  call initialize_sources(.false., iret)
  call initialize_io(.false., nrun, iret)

  select case(trim(mode))
  case("run", "ghosts")
    call mode_run()
  case("amr")
    call mode_amr()
  case("flux")
    call mode_flux()
  case("recon")
    call mode_recon()
  case("eqs")
    call mode_eqs()
  case("c2p")
    call mode_c2p()
  case("algebra")
    call mode_algebra()
  case default
    write(*,*) "unknown harness_mode: ", trim(mode)
    stop 9
  end select

contains

!-------------------------------------------------------------------------------
! full runs: the reference's own mesh, ghost fill and integrator
!-------------------------------------------------------------------------------
  subroutine mode_run()

    type(block_data), pointer :: pdata
    type(block_meta), pointer :: pmeta
    real(kind=8), dimension(:,:,:,:), allocatable :: ug
    real(kind=8), dimension(:,:)    , allocatable :: trace
    integer :: nx, ny, nz, i0, j0, k0, p, s, u, hint
    integer(kind=8) :: c0, c1, crate

    nx = ir * nc
    ny = jr * nc
#if NDIMS == 3
    nz = kr * nc
#else
    nz = 1
#endif
    allocate(ug(nx,ny,nz,nv), trace(2,nsteps))

    open(newunit=u, file=trim(fin), access="stream", form="unformatted",     &
         status="old")
    read(u) ug
    close(u)

    call initialize_mesh(nrun, .false., iret)
    call generate_mesh()

! overwrite the problem set-up with the seeded input state
    pdata => list_data
    do while (associated(pdata))
      pmeta => pdata%meta
      i0 = nint((pmeta%xmin - xmin) / adx(pmeta%level))
      j0 = nint((pmeta%ymin - ymin) / ady(pmeta%level))
#if NDIMS == 3
      k0 = nint((pmeta%zmin - zmin) / adz(pmeta%level))
#else
      k0 = 0
#endif
      do p = 1, nv
        pdata%u0(p,ib:ie,jb:je,kb:ke) = ug(i0+1:i0+ie-ib+1,j0+1:j0+je-jb+1,   &
                                           k0+1:k0+ke-kb+1,p)
      end do
      pdata%u => pdata%u0
      call update_primitive_variables(pdata%u0, pdata%q)
      pdata => pdata%next
    end do

! the reference driver: ghost fill, then the first time step estimate
    call boundary_variables(0.0d+00, dtnext)
    call new_time_step(dtnext)

! "ghosts": dump every block's primitive array with its ghost layers, each
! preceded by the block's global cell offsets (as doubles)
    if (trim(mode) == "ghosts") then
      open(newunit=u, file=trim(fout)//"_q.bin", access="stream",            &
           form="unformatted", status="replace")
      pdata => list_data
      do while (associated(pdata))
        pmeta => pdata%meta
        write(u) (pmeta%xmin - xmin) / adx(pmeta%level),                      &
                 (pmeta%ymin - ymin) / ady(pmeta%level),                      &
                 (pmeta%zmin - zmin) / adz(pmeta%level), pdata%q
        pdata => pdata%next
      end do
      close(u)
      return
    end if

    call system_clock(c0, crate)
    do s = 1, nsteps
      call advance(dtnext)
      trace(1,s) = dt
      trace(2,s) = cmax
      time = time + dt
      step = step + 1
    end do
    call system_clock(c1)

! the reference's integrals and statistics files of the final state
! (integrals_interval = 1 makes store_integrals write at any step)
    call get_parameter_integer("harness_integrals", hint)
    if (hint > 0) then
      call initialize_integrals(.false., nrun, iret)
      call store_integrals()
      call finalize_integrals()
    end if

! wall-clock seconds of the step loop only (CPU baseline timing)
    open(newunit=u, file=trim(fout)//"_time.txt", status="replace")
    write(u,*) real(c1 - c0, kind=8) / real(crate, kind=8)
    close(u)

    pdata => list_data
    do while (associated(pdata))
      pmeta => pdata%meta
      i0 = nint((pmeta%xmin - xmin) / adx(pmeta%level))
      j0 = nint((pmeta%ymin - ymin) / ady(pmeta%level))
#if NDIMS == 3
      k0 = nint((pmeta%zmin - zmin) / adz(pmeta%level))
#else
      k0 = 0
#endif
      do p = 1, nv
        ug(i0+1:i0+ie-ib+1,j0+1:j0+je-jb+1,k0+1:k0+ke-kb+1,p) =                &
                                         pdata%u(p,ib:ie,jb:je,kb:ke)
      end do
      pdata => pdata%next
    end do

    open(newunit=u, file=trim(fout)//"_u.bin", access="stream",              &
         form="unformatted", status="replace")
    write(u) ug
    close(u)
    open(newunit=u, file=trim(fout)//"_dt.bin", access="stream",             &
         form="unformatted", status="replace")
    write(u) trace
    close(u)

  end subroutine mode_run

!-------------------------------------------------------------------------------
! multi-level runs on the reference's own refined mesh
!-------------------------------------------------------------------------------
  subroutine mode_amr()

    type(block_data), pointer :: pdata
    real(kind=8), dimension(:,:), allocatable :: trace
    integer :: s, u

    allocate(trace(2,max(nsteps,1)))
    trace(:,:) = 0.0d+00
    call initialize_mesh(nrun, .false., iret)
    call generate_mesh()

! a well-defined start: q = cons2prim(u) in every block (u optionally read)
    if (amr_input == 1) open(newunit=u, file=trim(fin), access="stream",     &
                             form="unformatted", status="old")
    pdata => list_data
    do while (associated(pdata))
      if (amr_input == 1) read(u) pdata%u0(1:nv,ib:ie,jb:je,kb:ke)
      pdata%u => pdata%u0
      call update_primitive_variables(pdata%u0, pdata%q)
      pdata => pdata%next
    end do
    if (amr_input == 1) close(u)
    call dump_blocks(trim(fout)//"_mesh0.bin")

    call boundary_variables(0.0d+00, dtnext)
    open(newunit=u, file=trim(fout)//"_q0.bin", access="stream",              &
         form="unformatted", status="replace")
    pdata => list_data
    do while (associated(pdata))
      write(u) pdata%q
      pdata => pdata%next
    end do
    close(u)
    call new_time_step(dtnext)

! freeze the mesh: advance() calls update_mesh only while toplev > 1
    toplev = 1
    do s = 1, nsteps
      call advance(dtnext)
      trace(1,s) = dt
      trace(2,s) = cmax
      time = time + dt
      step = step + 1
    end do
    call dump_blocks(trim(fout)//"_mesh1.bin")
    open(newunit=u, file=trim(fout)//"_dt.bin", access="stream",             &
         form="unformatted", status="replace")
    write(u) trace(:,1:nsteps)
    close(u)
    if (crit_out == 1) then
      open(newunit=u, file=trim(fout)//"_crit.bin", access="stream",         &
           form="unformatted", status="replace")
      pdata => list_data
      do while (associated(pdata))
        write(u) real(check_refinement_criterion(pdata), kind=8)
        pdata => pdata%next
      end do
      close(u)
    end if
    if (mesh_update == 1) then
! the ghost cells hold the state after the last stage's boundary_variables,
! as in advance (evolution.F90:321-340)
      toplev = maxlev
      call update_mesh()
      call dump_blocks(trim(fout)//"_mesh2.bin")
    end if

  end subroutine mode_amr

  subroutine dump_blocks(fname)

    character(len=*), intent(in) :: fname
    type(block_data), pointer :: pdata
    type(block_meta), pointer :: pmeta
    integer :: u

    open(newunit=u, file=fname, access="stream", form="unformatted",         &
         status="replace")
    pdata => list_data
    do while (associated(pdata))
      pmeta => pdata%meta
      write(u) real(pmeta%level, kind=8),                                     &
               (pmeta%xmin - xmin) / adx(pmeta%level),                        &
               (pmeta%ymin - ymin) / ady(pmeta%level),                        &
#if NDIMS == 3
               (pmeta%zmin - zmin) / adz(pmeta%level),                        &
#else
               0.0d+00,                                                       &
#endif
               pdata%u(1:nv,ib:ie,jb:je,kb:ke)
      pdata => pdata%next
    end do
    close(u)

  end subroutine dump_blocks

!-------------------------------------------------------------------------------
! one block: reconstruction + Riemann fluxes (update_flux) on given q
!-------------------------------------------------------------------------------
  subroutine mode_flux()

    real(kind=8), dimension(:,:,:,:)  , allocatable :: q
    real(kind=8), dimension(:,:,:,:,:), allocatable :: f
    real(kind=8), dimension(NDIMS) :: dx
    integer :: b, ui, uo

    allocate(q(nv,im,jm,km), f(NDIMS,nv,im,jm,km))
    cmax  = hcmax
    cmax2 = hcmax * hcmax

    open(newunit=ui, file=trim(fin), access="stream", form="unformatted",    &
         status="old")
    open(newunit=uo, file=trim(fout)//"_f.bin", access="stream",             &
         form="unformatted", status="replace")
    do b = 1, ncount
      read(ui) q, dx
      call update_flux(dx, q, f)
      write(uo) f
    end do
    close(ui)
    close(uo)

  end subroutine mode_flux

!-------------------------------------------------------------------------------
! 1D reconstruction of pencils (reconstruct -> reconstruct_states [+ clip])
!-------------------------------------------------------------------------------
  subroutine mode_recon()

    real(kind=8), dimension(:,:), allocatable :: fp, fl, fr
    real(kind=8) :: h
    integer :: p, u

    allocate(fp(n,ncount), fl(n,ncount), fr(n,ncount))
    open(newunit=u, file=trim(fin), access="stream", form="unformatted",     &
         status="old")
    read(u) fp, h
    close(u)
    do p = 1, ncount
      call reconstruct(n, h, fp(:,p), fl(:,p), fr(:,p))
    end do
    open(newunit=u, file=trim(fout)//"_fl.bin", access="stream",             &
         form="unformatted", status="replace")
    write(u) fl
    close(u)
    open(newunit=u, file=trim(fout)//"_fr.bin", access="stream",             &
         form="unformatted", status="replace")
    write(u) fr
    close(u)

  end subroutine mode_recon

!-------------------------------------------------------------------------------
! pointwise equation-of-state routines
!-------------------------------------------------------------------------------
  subroutine mode_eqs()

    real(kind=8), dimension(:,:), allocatable :: q, uu, q2, f
    real(kind=8), dimension(:)  , allocatable :: cm, cp
    integer :: u

    allocate(q(nv,n), uu(nv,n), q2(nv,n), f(nv,n), cm(n), cp(n))
    cmax  = hcmax
    cmax2 = hcmax * hcmax
    open(newunit=u, file=trim(fin), access="stream", form="unformatted",     &
         status="old")
    read(u) q
    close(u)
    call prim2cons(n, q, uu)
    call cons2prim(n, uu, q2)
    call fluxspeed(n, q, uu, f, cm, cp)
    open(newunit=u, file=trim(fout)//"_eqs.bin", access="stream",            &
         form="unformatted", status="replace")
    write(u) uu, q2, f, cm, cp
    close(u)

  end subroutine mode_eqs

!-------------------------------------------------------------------------------
! cons2prim alone on given conserved states; q starts as the input's second
! half so that entries a failing cell leaves untouched are known
!-------------------------------------------------------------------------------
  subroutine mode_c2p()

    real(kind=8), dimension(:,:), allocatable :: uu, q
    integer :: u

    allocate(uu(nv,n), q(nv,n))
    open(newunit=u, file=trim(fin), access="stream", form="unformatted",     &
         status="old")
    read(u) uu, q
    close(u)
    call cons2prim(n, uu, q)
    open(newunit=u, file=trim(fout)//"_c2p.bin", access="stream",            &
         form="unformatted", status="replace")
    write(u) q
    close(u)

  end subroutine mode_c2p

!-------------------------------------------------------------------------------
! the polynomial root finders on coefficient sets a(5) (increasing powers);
! per set 19 values: quadratic(a(1:3)) [nr,x(2)], quadratic_normalized(a(1:2))
! [nr,x(2)], cubic(a(1:4)) [nr,x(3)], cubic_normalized(a(1:3)) [nr,x(3)],
! quartic(a(1:5)) [nr,x(4)]; root arrays start at zero
!-------------------------------------------------------------------------------
  subroutine mode_algebra()

    real(kind=8), dimension(:,:), allocatable :: a, r
    real(kind=8), dimension(4) :: x
    integer :: i, u, nr

    allocate(a(5,n), r(19,n))
    open(newunit=u, file=trim(fin), access="stream", form="unformatted",     &
         status="old")
    read(u) a
    close(u)
    do i = 1, n
      x(:) = 0.0d+00
      nr = quadratic(a(1:3,i), x(1:2))
      r(1,i) = nr
      r(2:3,i) = x(1:2)
      x(:) = 0.0d+00
      nr = quadratic_normalized(a(1:2,i), x(1:2))
      r(4,i) = nr
      r(5:6,i) = x(1:2)
      x(:) = 0.0d+00
      nr = cubic(a(1:4,i), x(1:3))
      r(7,i) = nr
      r(8:10,i) = x(1:3)
      x(:) = 0.0d+00
      nr = cubic_normalized(a(1:3,i), x(1:3))
      r(11,i) = nr
      r(12:14,i) = x(1:3)
      x(:) = 0.0d+00
      nr = quartic(a(1:5,i), x(1:4))
      r(15,i) = nr
      r(16:19,i) = x(1:4)
    end do
    open(newunit=u, file=trim(fout)//"_alg.bin", access="stream",            &
         form="unformatted", status="replace")
    write(u) r
    close(u)

  end subroutine mode_algebra

end program amun_ref_harness
