!-------------------------------------------------------------------------------
! derek_scf: Fortran SCF driver on the gfx950 library.
!
! Usage: derek_scf <seed> [-c]     reads <seed>.info
!   -c   check mode (like the reference's `derek -c`, io.f90:1074-1080):
!        parse the input, build the basis, report the grids and stop (no GPU).
!
! Input: the reference's `key : value` lines (io.f90:360-831; `!`/`#`
! comments) and the $DATA lattice / gaussian_pot blocks (io.f90:1765-1935),
! for the keys the SCF needs: cut_off_energy, g_fine_scale, kpt_mp_grid,
! n_electrons, electronic_temp (Kelvin, io.f90:340), energy_tol, max_scf,
! unit_energy / unit_length (Ha and bohr only: their factors are exactly 1),
! plus nbands, cos_pot_amp, wfn_seed and precision; external_pot (a preset
! name, or a <file>.potex whose potential replaces the wells, pot.f90:167-184)
! and write_potential / write_density / write_wvfn, which write <seed>.pot,
! <seed>.den and <seed>.wvfn after the SCF (state.f90:350-507); task
! (singlepoint, dos, bandstructure; io.f90:518-527) with dos_kpt_mp_grid:
! after the SCF the converged total potential is held fixed and the bands are
! found without self-consistency on the DOS mesh (<seed>.spec eigenvalues,
! <seed>.dos Gaussian-broadened density of states) or along the simple-cubic
! path G-X-M-G-R (<seed>.bands); the reference's spec_dos / spec_bands are
! empty (spectral.f90:33-44).
! The loop, its table and the stop test |E_i - E_{i-1}| < energy_tol follow
! electronic_scf (electronic.f90:86-175).
!-------------------------------------------------------------------------------
program derek_scf
  use iso_c_binding
  use derek_hip_m
  implicit none

  real(c_double), parameter :: kelvin_ha = 3.1668115634556076d-6  ! k_B / E_h (CODATA 2018)
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
  logical :: check_only = .false., have_cell = .false.
  logical :: write_pot = .false., write_den = .false., write_wfn = .false., have_file
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

  call get_command_argument(1, seed)
  if (len_trim(seed) == 0) then
     write(*, *) 'usage: derek_scf <seed> [-c]'
     error stop 2
  end if
  do i = 2, command_argument_count()
     call get_command_argument(i, arg)
     if (trim(arg) == '-c') check_only = .true.
  end do
  cell = 0.0d0
  allocate(centres(3, 0), heights(0), widths(0))

  open(newunit=u, file=trim(seed)//'.info', status='old', action='read', iostat=ios)
  if (ios /= 0) then
     write(*, *) "Error: file '"//trim(seed)//".info' does not exist."
     error stop 1
  end if
  do
     read(u, '(a)', iostat=ios) line
     if (ios /= 0) exit
     line = adjustl(line)
     if (len_trim(line) == 0) cycle
     if (line(1:1) == '!' .or. line(1:1) == '#') cycle
     if (index(lower(line), '$data') == 1) then
        call read_block(u, lower(line))
        cycle
     end if
     i = scan(line, ':=')
     if (i == 0) cycle
     key = lower(adjustl(line(1:i-1)))
     val = adjustl(line(i+1:))
     select case (trim(key))
     case ('cut_off_energy'); read(val, *) ecut
     case ('g_fine_scale'); read(val, *) g_fine
     case ('kpt_mp_grid'); read(val, *) mp
     case ('n_electrons'); read(val, *) nelec
     case ('nbands'); read(val, *) nbands
     case ('electronic_temp'); read(val, *) temp_k
     case ('energy_tol'); read(val, *) energy_tol
     case ('max_scf'); read(val, *) max_scf
     case ('cos_pot_amp'); read(val, *) cos_amp
     case ('wfn_seed'); read(val, *) wfn_seed
     case ('external_pot'); ext_pot = val
     case ('task')
        task = lower(val)
        select case (trim(task))
        case ('singlepoint', 'bandstructure', 'dos')
           continue
        case default
           call fail('Unknown task: '//trim(val))
        end select
     case ('dos_kpt_mp_grid'); read(val, *) dos_mp
     case ('write_potential'); write_pot = truth(val)
     case ('write_density'); write_den = truth(val)
     case ('write_wvfn'); write_wfn = truth(val)
     case ('precision')
        if (trim(lower(val)) == 'mixed') prec = DEREK_MIXED
     case ('unit_energy')
        if (trim(val) /= 'Ha') call fail('only unit_energy : Ha is supported by this driver')
     case ('unit_length')
        if (trim(val) /= 'bohr') call fail('only unit_length : bohr is supported by this driver')
     case default
        continue
     end select
  end do
  close(u)
  if (.not. have_cell) call fail(' No lattice provided')
  ! external_pot other than a preset names a .potex file (pot.f90:167-176)
  if (trim(lower(ext_pot)) /= 'gaussian_pot') then
     if (index(trim(ext_pot), '.potex') == 0) call fail('potential naming convension not correct')
     inquire(file=trim(ext_pot), exist=have_file)
     if (.not. have_file) call fail('file '//trim(ext_pot)//' does not exist')
  end if
  if (nbands <= 0) nbands = int(nelec / 2.0d0) + 4

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
  n = size(heights)
  select case (trim(lower(ext_pot)))
  case ('gaussian_pot')
     call derek_check(derek_gaussian_pot(cell, fine_dims, n, centres, heights, widths, cos_amp, vext), &
          'gaussian_pot')
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
  par%kT = temp_k * kelvin_ha
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

  logical function truth(v)
    character(*), intent(in) :: v
    select case (trim(lower(adjustl(v))))
    case ('true', 't', '.true.', 'yes', 'on')
       truth = .true.
    case default
       truth = .false.
    end select
  end function truth

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

  ! $DATA <type> ... $ENDDATA <type>; lattice rows are cell(row, :)
  ! (io.f90:1859-1878); gaussian_pot rows are cx cy cz height width.
  subroutine read_block(unit, head)
    integer, intent(in) :: unit
    character(*), intent(in) :: head
    character(len=256) :: l, kind
    integer :: r, st
    real(c_double) :: row(5)
    real(c_double), allocatable :: tc(:, :), th(:), tw(:)
    read(head(6:), *) kind
    r = 0
    do
       read(unit, '(a)', iostat=st) l
       if (st /= 0) call fail('Block '//trim(kind)//' not closed')
       l = adjustl(l)
       if (index(lower(l), '$enddata') == 1) exit
       if (len_trim(l) == 0 .or. l(1:1) == '!' .or. l(1:1) == '#') cycle
       r = r + 1
       select case (trim(kind))
       case ('lattice')
          if (r > 3) call fail('LATTICE must have 3 lines')
          read(l, *) cell(r, 1), cell(r, 2), cell(r, 3)
          have_cell = .true.
       case ('gaussian_pot')
          read(l, *) row
          allocate(tc(3, r), th(r), tw(r))
          if (r > 1) then
             tc(:, 1:r-1) = centres
             th(1:r-1) = heights
             tw(1:r-1) = widths
          end if
          tc(:, r) = row(1:3)
          th(r) = row(4)
          tw(r) = row(5)
          call move_alloc(tc, centres)
          call move_alloc(th, heights)
          call move_alloc(tw, widths)
       case default
          call fail(' Unknown block type '//trim(kind))
       end select
    end do
    if (trim(kind) == 'lattice' .and. r /= 3) call fail('LATTICE must have 3 lines')
  end subroutine read_block

end program derek_scf
