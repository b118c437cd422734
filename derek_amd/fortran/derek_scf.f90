!-------------------------------------------------------------------------------
! derek_scf: Fortran SCF driver on the gfx950 library.
!
! Usage (io_cl_parser, io.f90:994-1138):
!   derek_scf <seed>             run <seed>.info
!   derek_scf -c|--check <seed>  parse, build the basis, report the grids, stop
!   derek_scf -h|--help [key]    usage, or the description of one key
!   derek_scf -s|--search <text> keys whose name or description contains text
!   derek_scf -l|--list          every key with its default
!   derek_scf -v                 version
!
! <seed>.info is read by the library's restatement of the reference's input
! language (derek_info_*: all 45 keys, spell check, $DATA lattice and
! gaussian_pot blocks, output_level presets, units), plus this build's keys
! nbands, cos_pot_amp, wfn_seed and precision.  The external potential is
! the gaussian_pot block (wells), a preset (jelly, finite_barrier,
! periodic_pot; pot.f90:130-163) or a <file>.potex (pot.f90:167-184).
! The loop, its table and the stop test |E_i - E_{i-1}| < energy_tol follow
! electronic_scf (electronic.f90:86-175); write_potential / write_density /
! write_wvfn write <seed>.pot, <seed>.den and <seed>.wvfn after the SCF
! (state.f90:350-507); task (singlepoint, dos, bandstructure; io.f90:518-527)
! with dos_kpt_mp_grid: after the SCF the converged total potential is held
! fixed and the bands are found without self-consistency on the DOS mesh
! (<seed>.spec eigenvalues, <seed>.dos Gaussian-broadened density of states)
! or along the simple-cubic path G-X-M-G-R (<seed>.bands); the reference's
! spec_dos / spec_bands are empty (spectral.f90:33-44).
!-------------------------------------------------------------------------------
program derek_scf
  use iso_c_binding
  use derek_hip_m
  implicit none

  real(c_double) :: dummy3(3)
  character(len=256) :: seed, arg, line, key, val
  real(c_double) :: cell(3, 3), inv_cell(3, 3), ecut = 200.0d0, g_fine = 2.0d0, temp_k = 0.0d0
  real(c_double) :: energy_tol = 1.0d-7, cos_amp = 0.0d0, nelec = 1.0d0
  integer(c_int32_t) :: mp(3) = (/1, 1, 1/), max_scf = 40, nbands = 0, prec = DEREK_FP64
  integer(c_int64_t) :: wfn_seed = 1
  real(c_double), allocatable :: centres(:, :), heights(:), widths(:), kfrac(:, :), kin(:), gsq(:), vext(:)
  real(c_double), allocatable :: eig(:, :)
  integer(c_int32_t), allocatable :: fft_index(:)
  integer(c_int32_t) :: std_dims(3), fine_dims(3), nk, k, nwells = 0, iscf
  integer(c_int64_t) :: npw, max_npw, ngrid
  logical :: check_only = .false.
  type(c_ptr) :: inf
  logical :: write_pot = .false., write_den = .false., write_wfn = .false.
  character(len=256) :: ext_pot = 'gaussian_pot', task = 'singlepoint'
  integer(c_int32_t) :: dos_mp(3) = (/11, 11, 11/)
  complex(c_double_complex), allocatable :: ncpot(:, :, :), charge(:), psi_all(:)
  real(c_double), allocatable :: rho(:)
  integer(c_int64_t), allocatable :: npw_k(:), std_idx(:)
  integer(c_int32_t), allocatable :: fidx(:)
  integer(c_int32_t) :: alloc_flag
  integer(c_int64_t) :: off
  integer :: u, ios, i, n
  type(c_ptr) :: ctx
  type(derek_dims) :: dims
  type(derek_scf_params) :: par
  type(derek_scf_report) :: rep
  real(c_double) :: e_old, t0, t1
  logical :: converged

  call parse_command_line()
  call derek_check(derek_info_read(trim(seed)//'.info'//c_null_char, 'Normal'//c_null_char, inf), 'info')
  call print_text_log()
  call derek_check(derek_info_cell(inf, cell), 'info')
  ecut = getr('cut_off_energy')
  g_fine = getr('g_fine_scale')
  mp = get3('kpt_mp_grid')
  dos_mp = get3('dos_kpt_mp_grid')
  nelec = getr('n_electrons')
  temp_k = getr('electronic_temp')        ! already in Ha (converted from Kelvin)
  energy_tol = getr('energy_tol')
  max_scf = nint(getr('max_scf'))
  cos_amp = getr('cos_pot_amp')
  wfn_seed = nint(getr('wfn_seed'), c_int64_t)
  if (trim(gets('precision')) == 'mixed') prec = DEREK_MIXED
  nbands = nint(getr('nbands'))
  ! scalar bands hold two electrons; the reference counts spinor bands,
  ! n_electrons + conduction_bands (io.f90:316)
  if (nbands <= 0) nbands = ceiling(nelec / 2.0d0) + nint(getr('conduction_bands'))
  ext_pot = gets('external_pot')
  task = gets('task')
  write_pot = getr('write_potential') /= 0.0d0
  write_den = getr('write_density') /= 0.0d0
  write_wfn = getr('write_wvfn') /= 0.0d0
  if (getr('check') /= 0.0d0) check_only = .true.
  n = 0
  call derek_check(derek_info_wells(inf, n, dummy3, dummy3, dummy3), 'info')
  deallocate(centres, heights, widths)
  allocate(centres(3, max(n, 1)), heights(max(n, 1)), widths(max(n, 1)))
  call derek_check(derek_info_wells(inf, n, centres, heights, widths), 'info')
  nwells = n

  call derek_check(derek_basis_grid_dims(cell, ecut, g_fine, std_dims, fine_dims), 'basis_grid_dims')
  call derek_check(derek_basis_inv_cell(cell, inv_cell), 'basis_inv_cell')
  nk = mp(1) * mp(2) * mp(3)
  if (nk <= 0) call fail(' invalid number of kpoints')
  allocate(kfrac(nk, 3))
  call derek_check(derek_basis_kpoints(mp, kfrac, nk), 'basis_kpoints')
  ngrid = int(fine_dims(1), c_int64_t) * fine_dims(2) * fine_dims(3)
  max_npw = 0
  do k = 1, nk
     call derek_check(derek_basis_sphere(inv_cell, fine_dims, kfrac(k, :), ecut, 0_c_int64_t, npw, fft_dummy(), &
          kin_dummy()), 'basis_sphere')
     max_npw = max(max_npw, npw)
  end do
  write(*, '(a,3i5)') ' Standard FFT grid      :', std_dims
  write(*, '(a,3i5)') ' Fine FFT grid          :', fine_dims
  write(*, '(a,i8)') ' k-points               :', nk
  write(*, '(a,i8)') ' max plane waves per k  :', max_npw
  write(*, '(a,i8)') ' bands                  :', nbands
  if (check_only) stop

  dims%ngrid = fine_dims
  dims%nk = nk
  dims%nbands = nbands
  dims%precision = prec
  dims%volume = cell(1, 1) * cell(2, 2) * cell(3, 3)
  dims%max_npw = max_npw
  call derek_check(derek_hip_init(ctx, dims), 'hip_init')
  allocate(fft_index(max_npw), kin(max_npw), gsq(ngrid), vext(ngrid), eig(nbands, nk))
  do k = 1, nk
     call derek_check(derek_basis_sphere(inv_cell, fine_dims, kfrac(k, :), ecut, max_npw, npw, fft_index, kin), &
          'basis_sphere')
     call derek_check(derek_hip_set_basis(ctx, k - 1, npw, fft_index, kin, 1.0d0 / nk), 'hip_set_basis')
  end do
  call derek_check(derek_basis_gsq(inv_cell, fine_dims, gsq), 'basis_gsq')
  call derek_check(derek_hip_set_gsq(ctx, gsq), 'hip_set_gsq')
  select case (trim(ext_pot))
  case ('gaussian_pot')
     call derek_check(derek_gaussian_pot(cell, fine_dims, nwells, centres, heights, widths, cos_amp, vext), &
          'gaussian_pot')
  case ('jelly', 'finite_barrier', 'periodic_pot')
     call derek_check(derek_info_preset_pot(inf, fine_dims, vext), 'preset potential')
  case ('lorentzian_pot')
     call fail('lorentzian_pot is not supported by this build')
  case default
     ! a potential file replaces the preset (pot.f90:167-184)
     allocate(ncpot(ngrid, 2, 2))
     call derek_check(derek_io_pot_read(trim(ext_pot)//c_null_char, std_dims, ngrid, ncpot, alloc_flag), &
          'io_pot_read '//trim(ext_pot))
     vext = 0.5d0 * real(ncpot(:, 1, 1) + ncpot(:, 2, 2), c_double)
     deallocate(ncpot)
  end select
  call derek_check(derek_hip_set_vext(ctx, vext), 'hip_set_vext')

  par%nelec = nelec
  par%kT = temp_k
  par%mix_beta = 0.5d0
  par%mix_q0 = 1.0d0
  par%n_david_first = 4
  par%n_david = 2
  par%history = 8
  par%reserved = 0
  par%seed = wfn_seed
  par%fp32_switch_tol = 0.0d0
  par%refine_steps = 0
  par%kbatch = 0
  par%use_graph = 0
  par%autotune = 0
  call derek_check(derek_hip_scf_init(ctx, par), 'hip_scf_init')

  write(*, *) '+--------+-------------------+-------------------+---------+ <-- SCF'
  write(*, *) '|  SCF   |      Energy       |   Energy Change   |  Time   | <-- SCF'
  write(*, *) '|  Step  |       (Ha)        |        (Ha)       |   (s)   | <-- SCF'
  write(*, *) '+--------+-------------------+-------------------+---------+ <-- SCF'
  call cpu_time(t0)
  converged = .false.
  e_old = 0.0d0
  do iscf = 1, max_scf
     call derek_check(derek_hip_scf_step(ctx, rep), 'hip_scf_step')
     call cpu_time(t1)
     if (iscf == 1) then
        write(*, 11) iscf, rep%etot, t1 - t0
     else
        write(*, 12) iscf, rep%etot, rep%etot - e_old, t1 - t0
        ! the FP64 refinement of DEREK_MIXED must be complete (refine_left = 0)
        if (abs(rep%etot - e_old) < energy_tol .and. rep%refine_left == 0) converged = .true.
     end if
     e_old = rep%etot
     if (converged) exit
  end do
  write(*, *) '+--------+-------------------+-------------------+---------+ <-- SCF'
  call derek_check(derek_hip_get_eigenvalues(ctx, eig), 'hip_get_eigenvalues')
  write(*, '(a,f20.12)') ' Final energy (Ha)      :', rep%etot
  write(*, '(a,f20.12)') ' Fermi level (Ha)       :', rep%fermi
  write(*, '(a,l2)') ' Converged              :', converged
  write(*, '(a,4f14.8)') ' lowest eigenvalues (Ha):', eig(1:min(4, nbands), 1)
  if (write_pot) then
     allocate(rho(ngrid), ncpot(ngrid, 2, 2))
     call derek_check(derek_hip_get_potential(ctx, rho), 'hip_get_potential')
     ncpot = (0.0d0, 0.0d0)
     ncpot(:, 1, 1) = cmplx(rho, 0.0d0, c_double_complex)
     ncpot(:, 2, 2) = ncpot(:, 1, 1)
     call derek_check(derek_io_pot_write(trim(seed)//'.pot'//c_null_char, std_dims, ngrid, ncpot, 1), 'io_pot_write')
     write(*, '(a)') ' Total potential to     : '//trim(seed)//'.pot'
     deallocate(rho, ncpot)
  end if
  if (write_den) then
     allocate(rho(ngrid), charge(ngrid))
     call derek_check(derek_hip_get_density(ctx, rho), 'hip_get_density')
     charge = cmplx(rho, 0.0d0, c_double_complex)
     allocate(psi_all(3 * ngrid))
     psi_all = (0.0d0, 0.0d0)
     call derek_check(derek_io_den_write(trim(seed)//'.den'//c_null_char, std_dims, ngrid, charge, psi_all, 1), &
          'io_den_write')
     write(*, '(a)') ' Density to             : '//trim(seed)//'.den'
     deallocate(rho, charge, psi_all)
  end if
  if (write_wfn) then
     allocate(npw_k(nk))
     do k = 1, nk
        call derek_check(derek_basis_sphere(inv_cell, fine_dims, kfrac(k, :), ecut, 0_c_int64_t, npw_k(k), &
             fft_dummy(), kin_dummy()), 'basis_sphere')
     end do
     allocate(std_idx(sum(npw_k)), psi_all(sum(npw_k) * nbands), fidx(max_npw))
     off = 0
     do k = 1, nk
        call derek_check(derek_basis_sphere(inv_cell, fine_dims, kfrac(k, :), ecut, max_npw, npw, fidx, kin), &
             'basis_sphere')
        call derek_check(derek_basis_std_index(fine_dims, std_dims, npw, fidx, std_idx(off + 1:off + npw)), &
             'basis_std_index')
        if (any(std_idx(off + 1:off + npw) < 0)) call fail('plane wave outside the standard grid')
        call derek_check(derek_hip_get_psi(ctx, k - 1, psi_all(off * nbands + 1:(off + npw) * nbands)), 'hip_get_psi')
        off = off + npw
     end do
     call derek_check(derek_io_wvfn_write(trim(seed)//'.wvfn'//c_null_char, std_dims, nk, nbands, npw_k, std_idx, &
          psi_all, 1), 'io_wvfn_write')
     write(*, '(a)') ' Wavefunction to        : '//trim(seed)//'.wvfn'
     deallocate(npw_k, std_idx, psi_all, fidx)
  end if
  if (trim(task) == 'dos' .or. trim(task) == 'bandstructure') call spectral_task()
  call derek_check(derek_hip_destroy(ctx), 'hip_destroy')
11 format(1x,'|',T6,I05,T11,"|",T15,ES16.8E3,T31,"|",T51,'|',T53,f8.2,T61,'| <-- SCF')
12 format(1x,'|',T6,I05,T11,"|",T15,ES16.8E3,T31,"|",T35,ES16.8E3,T51,'|',T53,f8.2,T61,'| <-- SCF')

contains

  function lower(s) result(t)
    character(*), intent(in) :: s
    character(len=len(s)) :: t
    integer :: j
    t = s
    do j = 1, len(s)
       if (t(j:j) >= 'A' .and. t(j:j) <= 'Z') t(j:j) = achar(iachar(t(j:j)) + 32)
    end do
  end function lower

  ! Bands at the converged total potential on the DOS mesh or the band path
  subroutine spectral_task()
    real(c_double), allocatable :: kf2(:, :), vtot(:), eig2(:, :), path_len(:), egrid(:), g(:)
    integer(c_int32_t) :: nk2, k2, b, ie
    integer(c_int64_t) :: npw2, maxnpw2
    type(c_ptr) :: ctx2
    type(derek_dims) :: dims2
    real(c_double) :: res, emin, emax, sigma, x
    real(c_double), parameter :: verts(3, 5) = reshape((/0.0d0, 0.0d0, 0.0d0, 0.5d0, 0.0d0, 0.0d0, &
         0.5d0, 0.5d0, 0.0d0, 0.0d0, 0.0d0, 0.0d0, 0.5d0, 0.5d0, 0.5d0/), (/3, 5/))
    integer, parameter :: nseg = 10
    integer :: sv, su, j
    allocate(vtot(ngrid))
    call derek_check(derek_hip_get_potential(ctx, vtot), 'hip_get_potential')
    if (trim(task) == 'dos') then
       nk2 = dos_mp(1) * dos_mp(2) * dos_mp(3)
       if (nk2 <= 0) call fail(' invalid dos_kpt_mp_grid')
       allocate(kf2(nk2, 3))
       call derek_check(derek_basis_kpoints(dos_mp, kf2, nk2), 'basis_kpoints')
    else
       nk2 = 4 * nseg + 1
       allocate(kf2(nk2, 3), path_len(nk2))
       kf2(1, :) = verts(:, 1)
       path_len(1) = 0.0d0
       k2 = 1
       do sv = 1, 4
          do j = 1, nseg
             k2 = k2 + 1
             kf2(k2, :) = verts(:, sv) + (verts(:, sv + 1) - verts(:, sv)) * (real(j, c_double) / nseg)
             path_len(k2) = path_len(k2 - 1) + norm2(kf2(k2, :) - kf2(k2 - 1, :))
          end do
       end do
    end if
    maxnpw2 = 0
    do k2 = 1, nk2
       call derek_check(derek_basis_sphere(inv_cell, fine_dims, kf2(k2, :), ecut, 0_c_int64_t, npw2, fft_dummy(), &
            kin_dummy()), 'basis_sphere')
       maxnpw2 = max(maxnpw2, npw2)
    end do
    dims2 = dims
    dims2%nk = nk2
    dims2%max_npw = maxnpw2
    call derek_check(derek_hip_init(ctx2, dims2), 'hip_init')
    deallocate(fft_index, kin)
    allocate(fft_index(maxnpw2), kin(maxnpw2), eig2(nbands, nk2))
    do k2 = 1, nk2
       call derek_check(derek_basis_sphere(inv_cell, fine_dims, kf2(k2, :), ecut, maxnpw2, npw2, fft_index, kin), &
            'basis_sphere')
       call derek_check(derek_hip_set_basis(ctx2, k2 - 1, npw2, fft_index, kin, 1.0d0 / nk2), 'hip_set_basis')
    end do
    call derek_check(derek_hip_set_gsq(ctx2, gsq), 'hip_set_gsq')
    call derek_check(derek_hip_set_vext(ctx2, vext), 'hip_set_vext')
    call derek_check(derek_hip_scf_init(ctx2, par), 'hip_scf_init')
    call derek_check(derek_hip_set_vloc(ctx2, vtot), 'hip_set_vloc')
    call derek_check(derek_hip_nscf(ctx2, 300, 1.0d-9, res), 'hip_nscf')
    call derek_check(derek_hip_get_eigenvalues(ctx2, eig2), 'hip_get_eigenvalues')
    call derek_check(derek_hip_destroy(ctx2), 'hip_destroy')
    write(*, '(a,i8,a,es10.2)') ' Non-SCF k-points       :', nk2, '   max residual', res
    if (trim(task) == 'dos') then
       open(newunit=su, file=trim(seed)//'.spec', status='replace', action='write')
       write(su, '(a)') '# non-self-consistent eigenvalues (Ha) at the converged potential'
       write(su, '(a,3i5,a,i6,a,i5)') '# dos_kpt_mp_grid', dos_mp, '   k-points', nk2, '   bands', nbands
       write(su, '(a,f20.12)') '# Fermi level (Ha)', rep%fermi
       do k2 = 1, nk2
          write(su, '(3f12.8,*(f16.10))') kf2(k2, :), eig2(:, k2)
       end do
       close(su)
       ! Gaussian broadening, states per Ha per cell, both spin channels
       sigma = 0.01d0
       emin = minval(eig2) - 5.0d0 * sigma
       emax = maxval(eig2) + 5.0d0 * sigma
       allocate(egrid(2001), g(2001))
       do ie = 1, 2001
          egrid(ie) = emin + (emax - emin) * (ie - 1) / 2000.0d0
          g(ie) = 0.0d0
          do k2 = 1, nk2
             do b = 1, nbands
                x = (egrid(ie) - eig2(b, k2)) / sigma
                g(ie) = g(ie) + 2.0d0 / nk2 * exp(-0.5d0 * x * x) / (sigma * sqrt(2.0d0 * acos(-1.0d0)))
             end do
          end do
       end do
       open(newunit=su, file=trim(seed)//'.dos', status='replace', action='write')
       write(su, '(a,f8.4,a)') '# energy (Ha)   DOS (states/Ha/cell, both spins; Gaussian sigma', sigma, ' Ha)'
       write(su, '(a,f20.12)') '# Fermi level (Ha)', rep%fermi
       do ie = 1, 2001
          write(su, '(f16.10,es20.10)') egrid(ie), g(ie)
       end do
       close(su)
       write(*, '(a)') ' Spectral files         : '//trim(seed)//'.spec, '//trim(seed)//'.dos'
    else
       open(newunit=su, file=trim(seed)//'.bands', status='replace', action='write')
       write(su, '(a)') '# band structure along G-X-M-G-R at the converged potential (Ha)'
       write(su, '(a,f20.12)') '# Fermi level (Ha)', rep%fermi
       do k2 = 1, nk2
          write(su, '(f12.8,3f12.8,*(f16.10))') path_len(k2), kf2(k2, :), eig2(:, k2)
       end do
       close(su)
       write(*, '(a)') ' Band structure         : '//trim(seed)//'.bands'
    end if
  end subroutine spectral_task

  real(c_double) function getr(key)
    character(*), intent(in) :: key
    real(c_double) :: v(3)
    call derek_check(derek_info_num(inf, trim(key)//c_null_char, v), 'info '//trim(key))
    getr = v(1)
  end function getr

  function get3(key) result(m)
    character(*), intent(in) :: key
    integer(c_int32_t) :: m(3)
    real(c_double) :: v(3)
    call derek_check(derek_info_num(inf, trim(key)//c_null_char, v), 'info '//trim(key))
    m = nint(v)
  end function get3

  function gets(key) result(t)
    character(*), intent(in) :: key
    character(len=256) :: t
    character(kind=c_char) :: buf(256)
    call derek_check(derek_info_str(inf, trim(key)//c_null_char, buf, int(size(buf), c_int64_t)), 'info '//trim(key))
    t = c_text(buf)
  end function gets

  function c_text(buf) result(t)
    character(kind=c_char), intent(in) :: buf(:)
    character(len=size(buf)) :: t
    integer :: j
    t = ''
    do j = 1, size(buf)
       if (buf(j) == c_null_char) exit
       t(j:j) = buf(j)
    end do
  end function c_text

  ! spell-check warnings of the reader, printed as the reference prints its
  ! autocorrect table (io.f90:454-470)
  subroutine print_text_log()
    character(kind=c_char) :: buf(8192)
    character(len=8192) :: t
    integer :: p0, p1
    call derek_check(derek_info_log(inf, buf, int(size(buf), c_int64_t)), 'info log')
    t = c_text(buf)
    if (len_trim(t) == 0) return
    write(*, '(a)') ' ********************************************************************'
    write(*, '(a)') ' *                         *** WARNING ***                          *'
    p0 = 1
    do
       p1 = index(t(p0:), achar(10))
       if (p1 == 0) exit
       write(*, '(a)') ' '//t(p0:p0 + p1 - 2)
       p0 = p0 + p1
    end do
    write(*, '(a)') ' ********************************************************************'
  end subroutine print_text_log

  subroutine print_help(mode, term)
    character(*), intent(in) :: mode, term
    character(kind=c_char) :: buf(65536)
    character(len=65536) :: t
    integer :: p0, p1
    call derek_check(derek_info_help(trim(mode)//c_null_char, trim(term)//c_null_char, buf, &
         int(size(buf), c_int64_t)), 'help')
    t = c_text(buf)
    p0 = 1
    do
       p1 = index(t(p0:), achar(10))
       if (p1 == 0) exit
       write(*, '(a)') t(p0:p0 + p1 - 2)
       p0 = p0 + p1
    end do
  end subroutine print_help

  subroutine usage()
    write(*, '(a)') ' Usage:'
    write(*, '(a)') '   derek_scf    <seed>                 : Run a calculation from <seed>.info'
    write(*, '(a)') '       "        -v                     : Print version information.'
    write(*, '(a)') '       "        -h,--help    <keyword> : Get help and commandline options. Optional keyword.'
    write(*, '(a)') '       "        -s,--search  <keyword> : Search list of available parameters.'
    write(*, '(a)') '       "        -l,--list              : Print list of parameters.'
    write(*, '(a)') '       "        -c,--check   <seed>    : Run calculation to check input files.'
  end subroutine usage

  ! command line of io_cl_parser (io.f90:994-1138)
  subroutine parse_command_line()
    integer :: na, ia
    character(len=256) :: a1, a2
    na = command_argument_count()
    seed = ''
    if (na == 0) then
       call usage()
       stop
    end if
    ia = 1
    do while (ia <= na)
       call get_command_argument(ia, a1)
       a2 = ''
       if (ia < na) call get_command_argument(ia + 1, a2)
       select case (trim(a1))
       case ('-h', '--help')
          if (len_trim(a2) > 0) then
             call print_help('help', a2)
          else
             call usage()
          end if
          stop
       case ('-s', '--search')
          call print_help('search', a2)
          stop
       case ('-l', '--list')
          call print_help('list', '')
          stop
       case ('-v')
          write(*, '(a,i0)') ' derek_scf on libderek_hip, ABI version ', derek_hip_version()
          stop
       case ('-c', '--check')
          check_only = .true.
       case default
          if (a1(1:1) == '-') then
             write(*, '(a)') ' Unknown argument: '//trim(a1)
             call usage()
             error stop 2
          end if
          seed = a1
       end select
       ia = ia + 1
    end do
    if (len_trim(seed) == 0) then
       call usage()
       error stop 2
    end if
  end subroutine parse_command_line

  subroutine fail(msg)
    character(*), intent(in) :: msg
    write(*, *) 'Error: '//trim(msg)
    error stop 1
  end subroutine fail

  function fft_dummy() result(p)
    integer(c_int32_t) :: p(1)
    p = 0
  end function fft_dummy

  function kin_dummy() result(p)
    real(c_double) :: p(1)
    p = 0.0d0
  end function kin_dummy



end program derek_scf
