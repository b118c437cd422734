/*
 * derek_hip.h - C-ABI of the MI355X (gfx950) Kohn-Sham SCF hot path for DEReK.
 *
 * Every entry point returns an int32 status: 0 (DEREK_OK) or one of the
 * DEREK_ERR_* codes; derek_hip_last_error() copies the message of the last
 * failure of the calling thread.  Arguments are plain pointers and sizes.
 * Host arrays are Fortran-compatible: column-major, contiguous, complex
 * values interleaved (re, im) as complex(C_DOUBLE_COMPLEX).  No device
 * pointer crosses the boundary: the context owns all GPU memory.
 *
 * Grids are flat with z fastest and x slowest, exactly the layout the
 * reference's fft_3d reshapes around FFTW (Source/fft.f90:258-273).
 * dir = +1 is FFTW_FORWARD (exp(-iGr), grids_real2recip), dir = -1 is
 * FFTW_BACKWARD (exp(+iGr), grids_recip2real); both unnormalised.
 *
 * Not thread-safe: one context per process and host thread, calls are
 * ordered on the context's HIP stream.
 */
#ifndef DEREK_HIP_H
#define DEREK_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DEREK_OK 0
#define DEREK_ERR_ARG 1     /* invalid argument (checked on the host before any launch) */
#define DEREK_ERR_HIP 2     /* HIP runtime / device failure */
#define DEREK_ERR_NUMERIC 3 /* e.g. overlap matrix not positive definite */
#define DEREK_ERR_STATE 4   /* call out of order (e.g. basis not set) */
#define DEREK_ERR_IO 5      /* file cannot be opened / read / written, or malformed record */

#define DEREK_FP64 0  /* complex128 wavefunctions, FP64 everywhere */
#define DEREK_MIXED 1 /* complex64 wavefunctions, FFTs, H|psi> and density; BF16-in / FP32-accumulate
                        MFMA overlaps and rotations (derek_hip_set_gemm); FP64 subspace solves,
                        reductions and energies; the last SCF steps refined in FP64 from the
                        FP32-converged state */

typedef struct derek_ctx derek_ctx;

typedef struct {
  int32_t ngrid[3];  /* fine FFT grid (fine_ngx, fine_ngy, fine_ngz), basis.f90:93-105 */
  int32_t nk;        /* number of k-points, io.f90:2008 */
  int32_t nbands;    /* bands per k-point */
  int32_t precision; /* DEREK_FP64 or DEREK_MIXED */
  double volume;     /* cell volume in bohr^3, io.f90:1976 */
  int64_t max_npw;   /* largest plane-wave count over the k-points */
} derek_dims;

typedef struct {
  double nelec;          /* electrons (spin-degenerate, occupation 2 per band) */
  double kT;             /* Fermi-Dirac smearing in Ha (0: fixed occupations) */
  double mix_beta;       /* Pulay/Kerker step, default 0.5 */
  double mix_q0;         /* Kerker wavevector in 1/bohr, default 1.0 */
  int32_t n_david_first; /* Davidson steps in the first SCF step, default 4 */
  int32_t n_david;       /* Davidson steps per later SCF step, default 2 */
  int32_t history;       /* Pulay history, default 8 (max 16) */
  int32_t reserved;
  uint64_t seed;         /* initial-wavefunction seed */
  double fp32_switch_tol; /* DEREK_MIXED: the FP32 phase ends once |E_i - E_{i-1}| < max(this, 2e-7 |E_i|)
                             (0: 1e-5 Ha), or once |dE| grows again after falling below 1e-3 Ha */
  int32_t refine_steps;   /* DEREK_MIXED: FP64 steps after the switch before convergence may be declared (0: 2,
                             i.e. the stop test compares two consecutive FP64 energies) */
  int32_t kbatch;         /* k-points per batched launch (0: all k-points in one launch) */
  int32_t use_graph;      /* 1: capture the SCF step in a hipGraph once and replay it */
  int32_t autotune;       /* 1: time the tunable launch shapes at scf_init and keep the fastest */
} derek_scf_params;

typedef struct {
  int32_t iscf;     /* step counter, starts at 1 */
  int32_t ndav;     /* Davidson steps done in this SCF step (first k) */
  double etot;      /* E_band - int V_Hxc,in rho_out + E_H + E_xc - TS (Ha) */
  double e_band;
  double e_hartree;
  double e_xc;
  double e_dc;      /* int (V_H,in + V_xc,in) rho_out */
  double minus_ts;
  double fermi;
  double drho;      /* || rho_out - rho_in ||_2 (bohr^-3/2) */
  double max_res;   /* largest Davidson residual norm */
  double t_dense;   /* GPU seconds of the FP64 subspace solves in this step (profiling enabled; else 0) */
  double t_step;    /* wall seconds of the step */
  int32_t fp32;     /* 1 if the step ran in the FP32 phase of DEREK_MIXED */
  int32_t refine_left; /* FP64 refinement steps still due before convergence may be declared */
} derek_scf_report;

/* ---- library ------------------------------------------------------------ */
int32_t derek_hip_version(void);
int32_t derek_hip_last_error(char* buf, int64_t len);

/* ---- discretisation (host; restates Source/basis.f90 and Source/io.f90) -- */
/* Standard and fine grid sizes: basis.f90:87-105 with basis_prime_fact (:257-309).
 * cell is the 3x3 lattice, column-major as current_structure%cell. */
int32_t derek_basis_grid_dims(const double cell[9], double ecut, double g_fine_scale, int32_t std_dims[3],
                              int32_t fine_dims[3]);
/* transpose(twopi*io_inv(cell)), io.f90:1954 (orthorhombic cells). */
int32_t derek_basis_inv_cell(const double cell[9], double inv_cell[9]);
/* grid_points / fine_grid_points (N,3) and the integer indices (N,3), basis.f90:124-149. */
int32_t derek_basis_g_map(const double inv_cell[9], const int32_t dims[3], double* g, int32_t* m);
/* real_*grid_points and *frac_points (N,3), basis.f90:155-188. */
int32_t derek_basis_real_map(const double cell[9], const int32_t dims[3], double* r, double* frac);
/* kpt_scf_list (nk,3), io.f90:1992-2035.  *nk receives the count. */
int32_t derek_basis_kpoints(const int32_t mp[3], double* kfrac, int32_t* nk);
/* Plane waves 1/2|G+k|^2 <= ecut in fine-map order: FFT-bin index in the flat
 * z-fastest layout and kinetic energy.  Outputs may be NULL (count only);
 * at most cap entries are written. */
int32_t derek_basis_sphere(const double inv_cell[9], const int32_t dims[3], const double kfrac[3], double ecut,
                           int64_t cap, int64_t* npw, int32_t* fft_index, double* kin);
/* |G|^2 of every FFT bin of the grid (flat layout). */
int32_t derek_basis_gsq(const double inv_cell[9], const int32_t dims[3], double* gsq);
/* Standard-grid position (0-based, basis order of basis.f90:124-135: x
 * outer, z inner, m from -n/2 to n-n/2-1) of each plane wave given by its
 * fine-grid FFT bin; -1 where it lies outside the standard grid.  Places
 * sphere coefficients in the reference's coeff(gvec, ...) layout. */
int32_t derek_basis_std_index(const int32_t fine_dims[3], const int32_t std_dims[3], int64_t npw,
                              const int32_t* fft_index, int64_t* std_index);
/* External potential of $DATA gaussian_pot wells (io.f90:1888-1922; the
 * reference's pot_external_pot sets it to zero, pot.f90:165-166) at the
 * real_fine_grid_points: V = sum_i h_i sum_R exp(-|r-c_i-R|^2/(2 w_i^2))
 * + cos_amp*cos(2 pi x/L).  centres are fractional (3,nwells) column-major. */
int32_t derek_gaussian_pot(const double cell[9], const int32_t dims[3], int32_t nwells, const double* centres,
                           const double* heights, const double* widths, double cos_amp, double* v);

/* ---- context -------------------------------------------------------------- */
/* Allocates the device state (replaces fft_init's plans, fft.f90:139-219). */
/* ---- .info input (host only): the reference's input language
 * io_read_param (io.f90:360-831): lines with '!' anywhere are skipped, '#' in
 * the first column is a comment, `key : value` or `key = value`, keys are
 * case-insensitive and values lower-cased except for unit keys; a key within
 * Levenshtein distance < 3 ("Normal"; "Risky" 5, "Safe" 0) of a known key is
 * autocorrected with a warning (derek_info_log), otherwise "Unknown parameter
 * - <key>"; duplicates, invalid values and units fail with the reference's
 * messages; $DATA lattice / gaussian_pot blocks (io.f90:1765-1935);
 * output_level presets (io.f90:778-812); values converted to atomic units
 * as io_convert_all does (io.f90:324-358).  All 45 reference keys plus this
 * build's nbands, cos_pot_amp, wfn_seed, precision.  Errors -> DEREK_ERR_ARG. */
typedef struct derek_info derek_info;
int32_t derek_info_read(const char* path, const char* safety, derek_info** info);
int32_t derek_info_free(derek_info* info);
/* numeric key (logicals as 0/1, triplets as 3 values) */
int32_t derek_info_num(const derek_info* info, const char* key, double* v);
int32_t derek_info_str(const derek_info* info, const char* key, char* buf, int64_t len);
int32_t derek_info_present(const derek_info* info, const char* key, int32_t* present);
/* cell(row, col) column-major in bohr; " No lattice provided" without a block */
int32_t derek_info_cell(const derek_info* info, double cell[9]);
/* gaussian_pot rows: *n in = capacity (0: count only), out = count */
int32_t derek_info_wells(const derek_info* info, int32_t* n, double* centres, double* heights, double* widths);
int32_t derek_info_log(const derek_info* info, char* buf, int64_t len);
/* jelly / finite_barrier / periodic_pot potentials of pot_external_pot
 * (pot.f90:130-163) on the fine grid */
int32_t derek_info_preset_pot(const derek_info* info, const int32_t fine_dims[3], double* v);
/* command-line help texts (io_help / io_search / io_list_params,
 * io.f90:1140-1673): mode "help" (one key), "search" (substring), "list" */
int32_t derek_info_help(const char* mode, const char* term, char* buf, int64_t len);

/* ---- interchange files (host only; no context needed)
 * Fortran unformatted sequential files as the reference's gfortran build
 * writes them (-fconvert=big-endian, Makefile:86): one record per file (the
 * reference writes each type through a defined unformatted I/O procedure,
 * whose child statements continue the parent record), 4-byte big-endian
 * record markers, records over 2147483639 bytes split into subrecords the
 * way gfortran does.  The header holds the STANDARD grid dimensions (as
 * pot_write / pot_read do, pot.f90:216-218,326-332) and the arrays are
 * fine-grid sized.  Readers fail with DEREK_ERR_ARG "fine grid mismatch"
 * when the header differs from std_dims, as pot_read does.
 * Complex arrays are interleaved (re, im) doubles, Fortran order. */
/* potential record (write(unit) pot -> pot_write, pot.f90:194-224): header,
 * nc_pot(nfine, 2, 2), allocated.  A scalar potential V is nc_pot(:,1,1) =
 * nc_pot(:,2,2) = V, off-diagonal blocks 0 (xsf.f90:43-45 reads back V as
 * the mean of the diagonal).  Used for .pot and .potex (external_pot, read
 * with pot_read at pot.f90:167-184). */
int32_t derek_io_pot_write(const char* path, const int32_t std_dims[3], int64_t nfine, const double* nc_pot,
                           int32_t allocated);
int32_t derek_io_pot_read(const char* path, const int32_t std_dims[3], int64_t nfine, double* nc_pot,
                          int32_t* allocated);
/* density record (density_write, density.f90:133-163): header,
 * charge(nfine), spin(nfine, 3) (spin may be NULL: written as zeros /
 * skipped on reading), allocated. */
int32_t derek_io_den_write(const char* path, const int32_t std_dims[3], int64_t nfine, const double* charge,
                           const double* spin, int32_t allocated);
int32_t derek_io_den_read(const char* path, const int32_t std_dims[3], int64_t nfine, double* charge, double* spin,
                          int32_t* allocated);
/* wavefunction record (wave_write, wave.f90:896-926): header,
 * coeff(nstd, nk, nb, 2), nbands, kpts, allocated.  The caller passes the
 * sphere coefficients (per k-point npw[k] x nb column-major, k-points one
 * after the other) and their standard-grid positions (derek_basis_std_index,
 * per k-point, concatenated); the writer scatters them into spinor 1 of
 * the standard-grid array (zeros elsewhere and in spinor 2), the reader
 * gathers them back and checks nbands and kpts. */
int32_t derek_io_wvfn_write(const char* path, const int32_t std_dims[3], int32_t nk, int32_t nb,
                            const int64_t* npw, const int64_t* std_index, const double* psi, int32_t allocated);
int32_t derek_io_wvfn_read(const char* path, const int32_t std_dims[3], int32_t nk, int32_t nb, const int64_t* npw,
                           const int64_t* std_index, double* psi, int32_t* allocated);

int32_t derek_hip_init(derek_ctx** ctx, const derek_dims* dims);
int32_t derek_hip_destroy(derek_ctx* ctx);
/* Plane-wave set of k-point k (0-based) and its weight. */
int32_t derek_hip_set_basis(derek_ctx* ctx, int32_t k, int64_t npw, const int32_t* fft_index, const double* kin,
                            double weight);
int32_t derek_hip_set_gsq(derek_ctx* ctx, const double* gsq);
/* Total local potential used by derek_hip_apply_h (N_g doubles). */
int32_t derek_hip_set_vloc(derek_ctx* ctx, const double* v);
/* External potential used by the SCF (N_g doubles). */
int32_t derek_hip_set_vext(derek_ctx* ctx, const double* v);

/* ---- per-call operations on host buffers ---------------------------------- */
/* In-place 3D FFT of howmany grids (complex, N_g each); replaces fft_3d,
 * fft.f90:222-303.  dir must be +1 or -1. */
int32_t derek_hip_fft3d(derek_ctx* ctx, double* grid, int32_t howmany, int32_t dir);
/* In-place 1D FFTs of howmany contiguous lines of length n (complex), n one
 * of the context's grid lengths; replaces fft_1d (fft.f90:305-376: the
 * length-nz transforms of the STD or FINE grid, dir +1 FFTW_FORWARD, -1
 * FFTW_BACKWARD, unnormalised). */
int32_t derek_hip_fft1d(derek_ctx* ctx, double* data, int32_t n, int64_t howmany, int32_t dir);
/* z = a x + b y over n complex values, a and b complex ({re, im}); y is not
 * read when b = 0, and z may alias x or y.  The wavefunction operators of
 * wave.f90: + (wave_add_*, :226-406, a = b = 1), - (wave_sub_*, :409-589,
 * a = 1, b = -1), * by a real or complex scalar (wave_scale_*, :591-736,
 * b = 0) and the copies (wave_copy_*, :740-855, a = 1, b = 0). */
int32_t derek_hip_wave_axpby(derek_ctx* ctx, int64_t n, const double a[2], const double* x, const double b[2],
                             const double* y, double* z);
/* H|psi> for nb bands of k-point k: hpsi = 1/2|G+k|^2 psi + FFT[V psi(r)]/N.
 * psi, hpsi: npw_k x nb complex.  Replaces pot_apply_wfn (pot.f90:344-382)
 * plus the kinetic term. */
int32_t derek_hip_apply_h(derek_ctx* ctx, int32_t k, int32_t nb, const double* psi, double* hpsi);
/* S = X^H Y (m x n complex); X: npw x m, Y: npw x n (X == Y with m == n
 * computes only the upper tiles of the Hermitian result).  Replaces the
 * wave_dot_* stubs (wave.f90:929-1036) with the full band-band overlap.
 * In a DEREK_MIXED context X and Y are rounded to complex64 and multiplied
 * on the f32 MFMA path (chunk sums in FP64), as in the SCF's FP32 phase;
 * derek_hip_rotate likewise. */
int32_t derek_hip_overlap(derek_ctx* ctx, int64_t npw, int32_t m, int32_t n, const double* X, const double* Y,
                          double* S);
/* S = X^H Y as derek_hip_overlap, with X, Y and S in device memory
 * (complex128; X == Y with m == n takes the Hermitian path), for callers
 * that cannot allocate: code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc.  No counterpart in the reference.  It
 * allocates nothing: the split-K partial sums go to a file-scope device
 * array of 2^21 complex128 values (32 MiB per GPU, held until the process
 * ends), whose address derek_hip_init resolves once per device; m * n above
 * 2^21 is refused with DEREK_ERR_ARG, and the row chunks are capped so their
 * partials fit.  Needs a DEREK_FP64 context.  Asynchronous on the context's
 * stream (derek_hip_stream; derek_hip_sync waits).  Every call on one GPU
 * shares the array: calls that overlap in time on one GPU race. */
int32_t derek_hip_overlap_noalloc(derek_ctx* ctx, int64_t npw, int32_t m, int32_t n, const double* X,
                                  const double* Y, double* S);
/* Z = X U; X: npw x m, U: m x n, Z: npw x n (complex). */
int32_t derek_hip_rotate(derek_ctx* ctx, int64_t npw, int32_t m, int32_t n, const double* X, const double* U,
                         double* Z);
/* rho(r) (accumulate ? += : =) weight_k * sum_b occ[b] |psi_b(r)|^2 / volume.
 * Replaces density_calculate (density.f90:101-131). */
int32_t derek_hip_density(derek_ctx* ctx, int32_t k, int32_t nb, const double* psi, const double* occ, double* rho,
                          int32_t accumulate);
/* LDA (Slater + Perdew-Zunger 81), spin-unpolarised, per point; the same
 * arguments as libxc's xc_f03_lda_exc_vxc(p, np, rho, zk, vrho)
 * (ExLib/libxc/src/libxc_master.F90:1168-1176). */
int32_t derek_hip_lda(derek_ctx* ctx, int64_t np, const double* rho, double* exc, double* vxc);
/* V_H(r) from rho(r) (V_H(G) = 4 pi rho(G)/|G|^2, G = 0 dropped) and
 * E_H = 1/2 int V_H rho.  Needs derek_hip_set_gsq. */
int32_t derek_hip_hartree(derek_ctx* ctx, const double* rho, double* vh, double* eh);

/* ---- FP64 subspace solvers on the GPU (host buffers) ----------------------
 * The reference calls LAPACK only for a 3x3 inverse (io.f90:2488,2494); the
 * SCF's subspace problems (absent there) run on the device. */
/* Lowest nev eigenpairs of a Hermitian n x n A (column-major complex):
 * w ascending, Z n x nev. */
int32_t derek_hip_heev(derek_ctx* ctx, int32_t n, int32_t nev, const double* A, double* w, double* Z);
/* Lowest nev solutions of H c = w S c, C^H S C = I (S Hermitian positive
 * semi-definite; columns of S numerically dependent on earlier ones are
 * dropped; C rows of dropped columns are zero).  C is m x nev. */
int32_t derek_hip_hegv(derek_ctx* ctx, int32_t m, int32_t nev, const double* S, const double* H, double* w, double* C);

/* ---- resident SCF (electronic_scf body, electronic.f90:86-175) ----------- */
/* Random orthonormal wavefunctions, uniform density nelec/volume. */
int32_t derek_hip_scf_init(derek_ctx* ctx, const derek_scf_params* params);
/* One SCF step: V_in from rho_in, Davidson for every k, occupations,
 * rho_out, energies, Pulay mixing.  The caller tests convergence
 * (|E_i - E_{i-1}| < energy_tol, electronic.f90:156-165), and in DEREK_MIXED
 * mode also requires report->refine_left == 0. */
int32_t derek_hip_scf_step(derek_ctx* ctx, derek_scf_report* report);
/* Davidson on the last step's potential, up to max_iter steps per k-point or
 * until every residual norm is below res_tol (eigenvalues for reporting). */
int32_t derek_hip_scf_polish(derek_ctx* ctx, int32_t max_iter, double res_tol, double* max_res);
/* eigenvalues (nbands, nk) column-major, Ha */
/* Non-self-consistent eigenpairs (task : bandstructure / dos, io.f90:522-527;
 * the reference's spec_bands / spec_dos are empty, spectral.f90:33-44): after
 * scf_init, Davidson on every k-point of the context at the total potential
 * given with derek_hip_set_vloc (e.g. the converged potential of an SCF run
 * on another k-point set), without density or potential updates, until all
 * residuals are below res_tol or max_iter iterations.  *max_res receives the
 * largest residual norm. */
int32_t derek_hip_nscf(derek_ctx* ctx, int32_t max_iter, double res_tol, double* max_res);
int32_t derek_hip_get_eigenvalues(derek_ctx* ctx, double* eig);
int32_t derek_hip_get_occupations(derek_ctx* ctx, double* occ);
/* Residual norms |H psi - eps psi| of every band at every k-point after the
 * last Davidson step (n_k x nbands doubles, band fastest), as the eigenvalues
 * are laid out.  No counterpart in the reference (its SCF reports no
 * residuals); a caller that converges only the lowest bands of a block reads
 * them here. */
int32_t derek_hip_get_residuals(derek_ctx* ctx, double* res);
/* rho_out of the last step (N_g doubles) */
int32_t derek_hip_get_density(derek_ctx* ctx, double* rho);
/* The total local potential V_ext + V_H + V_xc (fine grid, Ha) that the next
 * H|psi> uses: the reference's tot_pot (state.f90:350-367 writes it as
 * <seed>.pot). */
int32_t derek_hip_get_potential(derek_ctx* ctx, double* v);
/* resident wavefunctions of k (npw_k x nbands complex, FP64 on the host) */
int32_t derek_hip_get_psi(derek_ctx* ctx, int32_t k, double* psi);
int32_t derek_hip_put_psi(derek_ctx* ctx, int32_t k, const double* psi);
/* H applied to the first nb resident wavefunctions of k (for timing; result
 * kept on device); k = -1: of every k-point, all k-points in one batch. */
int32_t derek_hip_apply_h_resident(derek_ctx* ctx, int32_t k, int32_t nb);
/* Per-kernel-class timers: while enabled, HIP events on the context's
 * stream bracket every launch (not inside a captured graph); read returns
 * launches, total milliseconds and the algorithmic bytes / flops of those
 * launches.  Classes: 0 z pass on sticks with the plane-wave scatter,
 * 1 y pass on planes (inverse), 2 x pass fused with the V(r) multiply,
 * 3 y pass on planes (forward), 4 z pass on sticks with the gather and the
 * kinetic term, 5 x pass fused with the density accumulation, 6 overlap GEMM,
 * 7 rotation GEMM, 8/9/10 full-grid FFT z/y/x passes, 11 FP64 subspace
 * solves, 12 other SCF kernels (band statistics, TPA, potentials, mixing),
 * 13 fused y and x passes with the V(r) multiply (in place of 1, 2 and 3),
 * 14 fused y and x inverse passes with the density accumulation (in place of
 * 1 and 5).
 * (The reference's equivalent is its trace/profile module, trace.f90:515-667.) */
int32_t derek_hip_profile_enable(derek_ctx* ctx, int32_t on);
int32_t derek_hip_profile_reset(derek_ctx* ctx);
int32_t derek_hip_profile_read(derek_ctx* ctx, int32_t cls, int64_t* launches, double* total_ms, double* bytes,
                               double* flops);
/* Launch shapes in use: out[0] bands per x-pass workgroup (1, 2, 4, 8),
 * out[1] (k-point, band) items per pipeline launch, out[2] k-points per
 * batched launch, out[3] 1 if SCF steps are replayed from a hipGraph,
 * out[4] the fused y/x passes in use where the grid allows them: bit 0 for
 * H|psi>, bit 1 for the density, out[5] sticks per z-pass workgroup (16,
 * 32), out[6] rounds of resident workgroups the overlaps' row chunks fill
 * (1, 2, 4), out[7] the C tile edge of the BF16 overlaps (64, 128). */
int32_t derek_hip_get_tuning(derek_ctx* ctx, int32_t out[8]);
/* Set the FFT and GEMM tile choices (what scf_init's autotuner would
 * choose): sticks per workgroup of the z passes (16 or 32; 0 = 16 for
 * complex128, 32 for complex64), the rounds of resident workgroups that the
 * split-K overlaps' row chunks fill (1, 2 or 4) and the C tile edge of the
 * BF16 overlaps (64 or 128).  All of them regroup the same operations.  No
 * counterpart in the reference. */
int32_t derek_hip_set_tiles(derek_ctx* ctx, int32_t z_sticks, int32_t gemm_rounds, int32_t gemm_tile);
/* Set the H|psi> launch shapes (what scf_init's autotuner would choose):
 * bands per workgroup (1, 2, 4, 8), items per launch (>= 1), fused 0..3
 * (bit 0: fused y/x pass for H|psi>, bit 1: for the density).
 * The fused y/x pass is used for nx = ny in {32, 72, 108, 144} and z lengths
 * in the same set when its LDS working set fits; otherwise the separate
 * passes run.  No counterpart in the reference (launch-shape control of the
 * GPU pipeline). */
int32_t derek_hip_set_launch(derek_ctx* ctx, int32_t bands_per_wg, int32_t items_per_launch, int32_t fused);
/* Matrix cores for the complex64 overlaps and rotations of a DEREK_MIXED
 * context (derek_hip_overlap, derek_hip_rotate and the FP32 phase of the
 * SCF): 0 multiplies on the FP32 MFMAs; 1..3 on the BF16 MFMAs with FP32
 * accumulation, each FP32 operand split into that many bf16 terms (1: plain
 * BF16, 8 significant bits; 2 (the default of a DEREK_MIXED context): three
 * products per real product, about 16 bits; 3: six products, about FP32's
 * 24).  FP64 data is unaffected.
 * No counterpart in the reference (precision control of the GPU GEMMs). */
int32_t derek_hip_set_gemm(derek_ctx* ctx, int32_t bf16_terms);
/* Block until all work on the context's stream is done. */
int32_t derek_hip_sync(derek_ctx* ctx);
/* Host-visible stream handle (hipStream_t) for event timing. */
int32_t derek_hip_stream(derek_ctx* ctx, void** stream);

#ifdef __cplusplus
}
#endif
#endif /* DEREK_HIP_H */
