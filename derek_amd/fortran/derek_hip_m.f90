!-------------------------------------------------------------------------------
! derek_hip_m: iso_c_binding interfaces to libderek_hip.so (include/derek_hip.h)
!
! This is the binding the reference's Fortran driver uses in place of the
! inner loops of Source/fft.f90 (fft_3d), Source/pot.f90 (pot_apply_wfn),
! Source/wave.f90 (wave_dot_*), Source/density.f90 (density_calculate),
! Source/xc.f90 and the SCF body of Source/electronic.f90.  It follows the
! pattern of the reference's own FFTW interface blocks (Source/fft.f90:57-132):
! scalars by value, arrays by reference as assumed-size dummies, an opaque
! type(c_ptr) handle.  Every function returns 0 on success; derek_check turns
! a non-zero status into an error stop with the library's message, like the
! reference's io_errors (Source/io.f90:885-930).
!-------------------------------------------------------------------------------
module derek_hip_m
  use iso_c_binding
  implicit none
  private

  integer(c_int32_t), parameter, public :: DEREK_OK = 0, DEREK_FP64 = 0, DEREK_MIXED = 1

  type, bind(C), public :: derek_dims
     integer(c_int32_t) :: ngrid(3)
     integer(c_int32_t) :: nk
     integer(c_int32_t) :: nbands
     integer(c_int32_t) :: precision
     real(c_double)     :: volume
     integer(c_int64_t) :: max_npw
  end type derek_dims

  type, bind(C), public :: derek_scf_params
     real(c_double)     :: nelec
     real(c_double)     :: kT
     real(c_double)     :: mix_beta
     real(c_double)     :: mix_q0
     integer(c_int32_t) :: n_david_first
     integer(c_int32_t) :: n_david
     integer(c_int32_t) :: history
     integer(c_int32_t) :: reserved
     integer(c_int64_t) :: seed
     real(c_double)     :: fp32_switch_tol
     integer(c_int32_t) :: refine_steps
     integer(c_int32_t) :: kbatch
     integer(c_int32_t) :: use_graph
     integer(c_int32_t) :: autotune
  end type derek_scf_params

  type, bind(C), public :: derek_scf_report
     integer(c_int32_t) :: iscf
     integer(c_int32_t) :: ndav
     real(c_double)     :: etot, e_band, e_hartree, e_xc, e_dc, minus_ts, fermi, drho, max_res
     real(c_double)     :: t_dense, t_step
     integer(c_int32_t) :: fp32, refine_left
  end type derek_scf_report

  public :: derek_check

  interface
     integer(c_int32_t) function derek_hip_version() bind(C, name='derek_hip_version')
       import :: c_int32_t
     end function derek_hip_version

     integer(c_int32_t) function derek_hip_last_error(buf, len) bind(C, name='derek_hip_last_error')
       import :: c_int32_t, c_int64_t, c_char
       character(kind=c_char), dimension(*) :: buf
       integer(c_int64_t), value :: len
     end function derek_hip_last_error

     ! basis.f90:87-105 with basis_prime_fact
     integer(c_int32_t) function derek_basis_grid_dims(cell, ecut, g_fine_scale, std_dims, fine_dims) &
          bind(C, name='derek_basis_grid_dims')
       import :: c_int32_t, c_double
       real(c_double), dimension(*) :: cell
       real(c_double), value :: ecut, g_fine_scale
       integer(c_int32_t), dimension(*) :: std_dims, fine_dims
     end function derek_basis_grid_dims

     integer(c_int32_t) function derek_basis_inv_cell(cell, inv_cell) bind(C, name='derek_basis_inv_cell')
       import :: c_int32_t, c_double
       real(c_double), dimension(*) :: cell, inv_cell
     end function derek_basis_inv_cell

     integer(c_int32_t) function derek_basis_g_map(inv_cell, dims, g, m) bind(C, name='derek_basis_g_map')
       import :: c_int32_t, c_double
       real(c_double), dimension(*) :: inv_cell, g
       integer(c_int32_t), dimension(*) :: dims, m
     end function derek_basis_g_map

     integer(c_int32_t) function derek_basis_real_map(cell, dims, r, frac) bind(C, name='derek_basis_real_map')
       import :: c_int32_t, c_double
       real(c_double), dimension(*) :: cell, r, frac
       integer(c_int32_t), dimension(*) :: dims
     end function derek_basis_real_map

     integer(c_int32_t) function derek_basis_kpoints(mp, kfrac, nk) bind(C, name='derek_basis_kpoints')
       import :: c_int32_t, c_double
       integer(c_int32_t), dimension(*) :: mp
       real(c_double), dimension(*) :: kfrac
       integer(c_int32_t) :: nk
     end function derek_basis_kpoints

     integer(c_int32_t) function derek_basis_sphere(inv_cell, dims, kfrac, ecut, cap, npw, fft_index, kin) &
          bind(C, name='derek_basis_sphere')
       import :: c_int32_t, c_int64_t, c_double
       real(c_double), dimension(*) :: inv_cell, kfrac, kin
       integer(c_int32_t), dimension(*) :: dims, fft_index
       real(c_double), value :: ecut
       integer(c_int64_t), value :: cap
       integer(c_int64_t) :: npw
     end function derek_basis_sphere

     integer(c_int32_t) function derek_basis_std_index(fine_dims, std_dims, npw, fft_index, std_index) &
          bind(C, name='derek_basis_std_index')
       import :: c_int32_t, c_int64_t
       integer(c_int32_t), dimension(*) :: fine_dims, std_dims, fft_index
       integer(c_int64_t), value :: npw
       integer(c_int64_t), dimension(*) :: std_index
     end function derek_basis_std_index

     ! .info input (io.f90:360-831) and command-line help texts
     integer(c_int32_t) function derek_info_read(path, safety, info) bind(C, name='derek_info_read')
       import :: c_int32_t, c_char, c_ptr
       character(kind=c_char), dimension(*) :: path, safety
       type(c_ptr) :: info
     end function derek_info_read

     integer(c_int32_t) function derek_info_free(info) bind(C, name='derek_info_free')
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: info
     end function derek_info_free

     integer(c_int32_t) function derek_info_num(info, key, v) bind(C, name='derek_info_num')
       import :: c_int32_t, c_char, c_ptr, c_double
       type(c_ptr), value :: info
       character(kind=c_char), dimension(*) :: key
       real(c_double), dimension(*) :: v
     end function derek_info_num

     integer(c_int32_t) function derek_info_str(info, key, buf, len) bind(C, name='derek_info_str')
       import :: c_int32_t, c_int64_t, c_char, c_ptr
       type(c_ptr), value :: info
       character(kind=c_char), dimension(*) :: key, buf
       integer(c_int64_t), value :: len
     end function derek_info_str

     integer(c_int32_t) function derek_info_present(info, key, present) bind(C, name='derek_info_present')
       import :: c_int32_t, c_char, c_ptr
       type(c_ptr), value :: info
       character(kind=c_char), dimension(*) :: key
       integer(c_int32_t) :: present
     end function derek_info_present

     integer(c_int32_t) function derek_info_cell(info, cell) bind(C, name='derek_info_cell')
       import :: c_int32_t, c_ptr, c_double
       type(c_ptr), value :: info
       real(c_double), dimension(*) :: cell
     end function derek_info_cell

     integer(c_int32_t) function derek_info_wells(info, n, centres, heights, widths) bind(C, name='derek_info_wells')
       import :: c_int32_t, c_ptr, c_double
       type(c_ptr), value :: info
       integer(c_int32_t) :: n
       real(c_double), dimension(*) :: centres, heights, widths
     end function derek_info_wells

     integer(c_int32_t) function derek_info_log(info, buf, len) bind(C, name='derek_info_log')
       import :: c_int32_t, c_int64_t, c_char, c_ptr
       type(c_ptr), value :: info
       character(kind=c_char), dimension(*) :: buf
       integer(c_int64_t), value :: len
     end function derek_info_log

     integer(c_int32_t) function derek_info_preset_pot(info, fine_dims, v) bind(C, name='derek_info_preset_pot')
       import :: c_int32_t, c_ptr, c_double
       type(c_ptr), value :: info
       integer(c_int32_t), dimension(*) :: fine_dims
       real(c_double), dimension(*) :: v
     end function derek_info_preset_pot

     integer(c_int32_t) function derek_info_help(mode, term, buf, len) bind(C, name='derek_info_help')
       import :: c_int32_t, c_int64_t, c_char
       character(kind=c_char), dimension(*) :: mode, term, buf
       integer(c_int64_t), value :: len
     end function derek_info_help

     ! interchange files (pot.f90:194-224,303-340; density.f90:133-163;
     ! wave.f90:896-926); path is a NUL-terminated name
     integer(c_int32_t) function derek_io_pot_write(path, std_dims, nfine, nc_pot, allocated) &
          bind(C, name='derek_io_pot_write')
       import :: c_int32_t, c_int64_t, c_char, c_double_complex
       character(kind=c_char), dimension(*) :: path
       integer(c_int32_t), dimension(*) :: std_dims
       integer(c_int64_t), value :: nfine
       complex(c_double_complex), dimension(*) :: nc_pot
       integer(c_int32_t), value :: allocated
     end function derek_io_pot_write

     integer(c_int32_t) function derek_io_pot_read(path, std_dims, nfine, nc_pot, allocated) &
          bind(C, name='derek_io_pot_read')
       import :: c_int32_t, c_int64_t, c_char, c_double_complex
       character(kind=c_char), dimension(*) :: path
       integer(c_int32_t), dimension(*) :: std_dims
       integer(c_int64_t), value :: nfine
       complex(c_double_complex), dimension(*) :: nc_pot
       integer(c_int32_t) :: allocated
     end function derek_io_pot_read

     integer(c_int32_t) function derek_io_den_write(path, std_dims, nfine, charge, spin, allocated) &
          bind(C, name='derek_io_den_write')
       import :: c_int32_t, c_int64_t, c_char, c_double_complex
       character(kind=c_char), dimension(*) :: path
       integer(c_int32_t), dimension(*) :: std_dims
       integer(c_int64_t), value :: nfine
       complex(c_double_complex), dimension(*) :: charge, spin
       integer(c_int32_t), value :: allocated
     end function derek_io_den_write

     integer(c_int32_t) function derek_io_den_read(path, std_dims, nfine, charge, spin, allocated) &
          bind(C, name='derek_io_den_read')
       import :: c_int32_t, c_int64_t, c_char, c_double_complex
       character(kind=c_char), dimension(*) :: path
       integer(c_int32_t), dimension(*) :: std_dims
       integer(c_int64_t), value :: nfine
       complex(c_double_complex), dimension(*) :: charge, spin
       integer(c_int32_t) :: allocated
     end function derek_io_den_read

     integer(c_int32_t) function derek_io_wvfn_write(path, std_dims, nk, nb, npw, std_index, psi, allocated) &
          bind(C, name='derek_io_wvfn_write')
       import :: c_int32_t, c_int64_t, c_char, c_double_complex
       character(kind=c_char), dimension(*) :: path
       integer(c_int32_t), dimension(*) :: std_dims
       integer(c_int32_t), value :: nk, nb, allocated
       integer(c_int64_t), dimension(*) :: npw, std_index
       complex(c_double_complex), dimension(*) :: psi
     end function derek_io_wvfn_write

     integer(c_int32_t) function derek_io_wvfn_read(path, std_dims, nk, nb, npw, std_index, psi, allocated) &
          bind(C, name='derek_io_wvfn_read')
       import :: c_int32_t, c_int64_t, c_char, c_double_complex
       character(kind=c_char), dimension(*) :: path
       integer(c_int32_t), dimension(*) :: std_dims
       integer(c_int32_t), value :: nk, nb
       integer(c_int64_t), dimension(*) :: npw, std_index
       complex(c_double_complex), dimension(*) :: psi
       integer(c_int32_t) :: allocated
     end function derek_io_wvfn_read

     integer(c_int32_t) function derek_basis_gsq(inv_cell, dims, gsq) bind(C, name='derek_basis_gsq')
       import :: c_int32_t, c_double
       real(c_double), dimension(*) :: inv_cell, gsq
       integer(c_int32_t), dimension(*) :: dims
     end function derek_basis_gsq

     integer(c_int32_t) function derek_gaussian_pot(cell, dims, nwells, centres, heights, widths, cos_amp, v) &
          bind(C, name='derek_gaussian_pot')
       import :: c_int32_t, c_double
       real(c_double), dimension(*) :: cell, centres, heights, widths, v
       integer(c_int32_t), dimension(*) :: dims
       integer(c_int32_t), value :: nwells
       real(c_double), value :: cos_amp
     end function derek_gaussian_pot

     ! fft_init (fft.f90:139-219): device state for one system
     integer(c_int32_t) function derek_hip_init(ctx, dims) bind(C, name='derek_hip_init')
       import :: c_int32_t, c_ptr, derek_dims
       type(c_ptr) :: ctx
       type(derek_dims) :: dims
     end function derek_hip_init

     integer(c_int32_t) function derek_hip_destroy(ctx) bind(C, name='derek_hip_destroy')
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: ctx
     end function derek_hip_destroy

     integer(c_int32_t) function derek_hip_set_basis(ctx, k, npw, fft_index, kin, weight) &
          bind(C, name='derek_hip_set_basis')
       import :: c_int32_t, c_int64_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: k
       integer(c_int64_t), value :: npw
       integer(c_int32_t), dimension(*) :: fft_index
       real(c_double), dimension(*) :: kin
       real(c_double), value :: weight
     end function derek_hip_set_basis

     integer(c_int32_t) function derek_hip_set_gsq(ctx, gsq) bind(C, name='derek_hip_set_gsq')
       import :: c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), dimension(*) :: gsq
     end function derek_hip_set_gsq

     integer(c_int32_t) function derek_hip_set_vloc(ctx, v) bind(C, name='derek_hip_set_vloc')
       import :: c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), dimension(*) :: v
     end function derek_hip_set_vloc

     integer(c_int32_t) function derek_hip_set_vext(ctx, v) bind(C, name='derek_hip_set_vext')
       import :: c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), dimension(*) :: v
     end function derek_hip_set_vext

     ! fft_3d (fft.f90:222-303): in place, dir = +1 forward / -1 backward
     integer(c_int32_t) function derek_hip_fft3d(ctx, grid, howmany, dir) bind(C, name='derek_hip_fft3d')
       import :: c_int32_t, c_double_complex, c_ptr
       type(c_ptr), value :: ctx
       complex(c_double_complex), dimension(*) :: grid
       integer(c_int32_t), value :: howmany, dir
     end function derek_hip_fft3d

     ! fft_1d (fft.f90:305-376): howmany lines of length n, in place
     integer(c_int32_t) function derek_hip_fft1d(ctx, lines, n, howmany, dir) bind(C, name='derek_hip_fft1d')
       import :: c_int32_t, c_int64_t, c_double_complex, c_ptr
       type(c_ptr), value :: ctx
       complex(c_double_complex), dimension(*) :: lines
       integer(c_int32_t), value :: n, dir
       integer(c_int64_t), value :: howmany
     end function derek_hip_fft1d

     ! wavefunction + - * and copies (wave.f90:226-855): z = a x + b y
     integer(c_int32_t) function derek_hip_wave_axpby(ctx, n, a, x, b, y, z) bind(C, name='derek_hip_wave_axpby')
       import :: c_int32_t, c_int64_t, c_double, c_double_complex, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int64_t), value :: n
       real(c_double), dimension(2) :: a, b
       complex(c_double_complex), dimension(*) :: x, y, z
     end function derek_hip_wave_axpby

     ! S = X^H Y on device memory with no allocation (no reference counterpart;
     ! for callers that cannot allocate: HIP graph capture, serving loops);
     ! X, Y, S are device addresses, m * n <= 2**21
     integer(c_int32_t) function derek_hip_overlap_noalloc(ctx, npw, m, n, x, y, s) &
          bind(C, name='derek_hip_overlap_noalloc')
       import :: c_int32_t, c_int64_t, c_ptr
       type(c_ptr), value :: ctx, x, y, s
       integer(c_int64_t), value :: npw
       integer(c_int32_t), value :: m, n
     end function derek_hip_overlap_noalloc

     ! pot_apply_wfn (pot.f90:344-382) + kinetic term
     integer(c_int32_t) function derek_hip_apply_h(ctx, k, nb, psi, hpsi) bind(C, name='derek_hip_apply_h')
       import :: c_int32_t, c_double_complex, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: k, nb
       complex(c_double_complex), dimension(*) :: psi, hpsi
     end function derek_hip_apply_h

     ! wave_dot_* (wave.f90:929-1036) as the band-band overlap
     integer(c_int32_t) function derek_hip_overlap(ctx, npw, m, n, x, y, s) bind(C, name='derek_hip_overlap')
       import :: c_int32_t, c_int64_t, c_double_complex, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int64_t), value :: npw
       integer(c_int32_t), value :: m, n
       complex(c_double_complex), dimension(*) :: x, y, s
     end function derek_hip_overlap

     integer(c_int32_t) function derek_hip_rotate(ctx, npw, m, n, x, u, z) bind(C, name='derek_hip_rotate')
       import :: c_int32_t, c_int64_t, c_double_complex, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int64_t), value :: npw
       integer(c_int32_t), value :: m, n
       complex(c_double_complex), dimension(*) :: x, u, z
     end function derek_hip_rotate

     ! density_calculate (density.f90:101-131)
     integer(c_int32_t) function derek_hip_density(ctx, k, nb, psi, occ, rho, accumulate) &
          bind(C, name='derek_hip_density')
       import :: c_int32_t, c_double, c_double_complex, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: k, nb, accumulate
       complex(c_double_complex), dimension(*) :: psi
       real(c_double), dimension(*) :: occ, rho
     end function derek_hip_density

     ! same arguments as libxc's xc_f03_lda_exc_vxc (libxc_master.F90:1168-1176)
     integer(c_int32_t) function derek_hip_lda(ctx, np, rho, exc, vxc) bind(C, name='derek_hip_lda')
       import :: c_int32_t, c_int64_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int64_t), value :: np
       real(c_double), dimension(*) :: rho, exc, vxc
     end function derek_hip_lda

     integer(c_int32_t) function derek_hip_hartree(ctx, rho, vh, eh) bind(C, name='derek_hip_hartree')
       import :: c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), dimension(*) :: rho, vh
       real(c_double) :: eh
     end function derek_hip_hartree

     ! electronic_scf body (electronic.f90:86-175)
     integer(c_int32_t) function derek_hip_scf_init(ctx, params) bind(C, name='derek_hip_scf_init')
       import :: c_int32_t, c_ptr, derek_scf_params
       type(c_ptr), value :: ctx
       type(derek_scf_params) :: params
     end function derek_hip_scf_init

     integer(c_int32_t) function derek_hip_scf_step(ctx, report) bind(C, name='derek_hip_scf_step')
       import :: c_int32_t, c_ptr, derek_scf_report
       type(c_ptr), value :: ctx
       type(derek_scf_report) :: report
     end function derek_hip_scf_step

     integer(c_int32_t) function derek_hip_scf_polish(ctx, max_iter, res_tol, max_res) &
          bind(C, name='derek_hip_scf_polish')
       import :: c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: max_iter
       real(c_double), value :: res_tol
       real(c_double) :: max_res
     end function derek_hip_scf_polish

     ! non-self-consistent eigenpairs at the potential of derek_hip_set_vloc
     integer(c_int32_t) function derek_hip_nscf(ctx, max_iter, res_tol, max_res) bind(C, name='derek_hip_nscf')
       import :: c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: max_iter
       real(c_double), value :: res_tol
       real(c_double) :: max_res
     end function derek_hip_nscf

     integer(c_int32_t) function derek_hip_get_eigenvalues(ctx, eig) bind(C, name='derek_hip_get_eigenvalues')
       import :: c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), dimension(*) :: eig
     end function derek_hip_get_eigenvalues

     integer(c_int32_t) function derek_hip_get_residuals(ctx, res) bind(C, name='derek_hip_get_residuals')
       import :: c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), dimension(*) :: res
     end function derek_hip_get_residuals

     integer(c_int32_t) function derek_hip_get_occupations(ctx, occ) bind(C, name='derek_hip_get_occupations')
       import :: c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), dimension(*) :: occ
     end function derek_hip_get_occupations

     integer(c_int32_t) function derek_hip_get_density(ctx, rho) bind(C, name='derek_hip_get_density')
       import :: c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), dimension(*) :: rho
     end function derek_hip_get_density

     integer(c_int32_t) function derek_hip_get_potential(ctx, v) bind(C, name='derek_hip_get_potential')
       import :: c_int32_t, c_double, c_ptr
       type(c_ptr), value :: ctx
       real(c_double), dimension(*) :: v
     end function derek_hip_get_potential

     integer(c_int32_t) function derek_hip_get_psi(ctx, k, psi) bind(C, name='derek_hip_get_psi')
       import :: c_int32_t, c_double_complex, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: k
       complex(c_double_complex), dimension(*) :: psi
     end function derek_hip_get_psi

     integer(c_int32_t) function derek_hip_put_psi(ctx, k, psi) bind(C, name='derek_hip_put_psi')
       import :: c_int32_t, c_double_complex, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: k
       complex(c_double_complex), dimension(*) :: psi
     end function derek_hip_put_psi

     integer(c_int32_t) function derek_hip_sync(ctx) bind(C, name='derek_hip_sync')
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: ctx
     end function derek_hip_sync

     ! out(1:5): bands per workgroup, items per launch, k-points per batch, graph, fused y/x pass
     integer(c_int32_t) function derek_hip_get_tuning(ctx, out) bind(C, name='derek_hip_get_tuning')
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), dimension(8) :: out
     end function derek_hip_get_tuning

     ! z-pass sticks per workgroup (0, 16, 32), overlap chunk rounds (1, 2, 4),
     ! BF16 overlap tile edge (64, 128)
     integer(c_int32_t) function derek_hip_set_tiles(ctx, z_sticks, gemm_rounds, gemm_tile) &
          bind(C, name='derek_hip_set_tiles')
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: z_sticks, gemm_rounds, gemm_tile
     end function derek_hip_set_tiles

     integer(c_int32_t) function derek_hip_set_launch(ctx, bands_per_wg, items_per_launch, fused) &
          bind(C, name='derek_hip_set_launch')
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: bands_per_wg, items_per_launch, fused
     end function derek_hip_set_launch

     ! complex64 overlaps/rotations: 0 FP32 MFMA, 1..3 BF16 MFMA with that many bf16 terms per value
     integer(c_int32_t) function derek_hip_set_gemm(ctx, bf16_terms) bind(C, name='derek_hip_set_gemm')
       import :: c_int32_t, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int32_t), value :: bf16_terms
     end function derek_hip_set_gemm
  end interface

  public :: derek_hip_version, derek_hip_last_error, derek_basis_grid_dims, derek_basis_inv_cell, &
       derek_basis_g_map, derek_basis_real_map, derek_basis_kpoints, derek_basis_sphere, derek_basis_gsq, &
       derek_gaussian_pot, derek_hip_init, derek_hip_destroy, derek_hip_set_basis, derek_hip_set_gsq, &
       derek_hip_set_vloc, derek_hip_set_vext, derek_hip_fft3d, derek_hip_fft1d, derek_hip_wave_axpby, &
       derek_hip_apply_h, derek_hip_overlap, derek_hip_overlap_noalloc, &
       derek_hip_rotate, derek_hip_density, derek_hip_lda, derek_hip_hartree, derek_hip_scf_init, &
       derek_hip_scf_step, derek_hip_scf_polish, derek_hip_get_eigenvalues, derek_hip_get_occupations, &
       derek_hip_get_density, derek_hip_get_psi, derek_hip_put_psi, derek_hip_sync, derek_hip_get_potential, &
       derek_basis_std_index, derek_io_pot_write, derek_io_pot_read, derek_io_den_write, derek_io_den_read, &
       derek_io_wvfn_write, derek_io_wvfn_read, derek_hip_nscf, derek_info_read, derek_info_free, derek_info_num, &
       derek_info_str, derek_info_present, derek_info_cell, derek_info_wells, derek_info_log, derek_info_preset_pot, &
       derek_info_help, derek_hip_get_tuning, derek_hip_set_launch, derek_hip_set_gemm, derek_hip_set_tiles

contains

  ! Stop with the library's message on a non-zero status (io_errors style).
  subroutine derek_check(status, where)
    integer(c_int32_t), intent(in) :: status
    character(*), intent(in) :: where
    character(kind=c_char) :: buf(1024)
    character(len=1024) :: msg
    integer :: i, rc
    if (status == DEREK_OK) return
    rc = derek_hip_last_error(buf, int(size(buf), c_int64_t))
    msg = ''
    do i = 1, size(buf)
       if (buf(i) == c_null_char) exit
       msg(i:i) = buf(i)
    end do
    write(*, '(a,a,a,i0,a,a)') ' Error in ', where, ' (status ', status, '): ', trim(msg)
    error stop 1
  end subroutine derek_check

end module derek_hip_m
