/*
 * hhdmrg.h -- C ABI of the MI355X-native SU(2) DMRG inner loop.
 *
 * Drop-in boundary for the inner loop of cfengno1/Hofstadter_Hubbard_TSC
 * (Fortran).  Every entry point names the reference subroutine it replaces
 * (paths relative to the reference tree).  The arguments are plain pointers
 * and sizes; complex arrays are interleaved (re, im) doubles, column-major,
 * as Fortran's double complex.  "dev" pointers are device (HBM) addresses on
 * the context's GPU; everything else is host memory borrowed for the call.
 *
 * Return codes: 0 ok, > 0 quantum-number / shape / argument error,
 * < 0 HIP or rocSOLVER error.  hh_last_error() gives the message.  No entry
 * point aborts the process (the reference stops on QN mismatches,
 * code/mat_vec.f90:987-990).
 *
 * Threading: one context per host thread; calls on a context serialise on
 * its HIP stream.  There is no global mutable state.
 */
#ifndef HHDMRG_H
#define HHDMRG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hh_ctx hh_ctx;
typedef struct hh_block hh_block;
typedef struct hh_superblock hh_superblock;
typedef struct hh_trunc hh_trunc;
typedef struct hh_batch hh_batch;

/* Model and lattice: replaces the hard-coded couplings and lattice setup of
 * program Quantum (code/quantum.f90:93-160), dmrg_setup
 * (code/warm_up.f90:347-666) and get_phase (code/quantum.f90:293-438).
 * Bond phases: gauge 0 is the Landau gauge (x bonds real, y bonds
 * t*exp(i 2 pi flux x), the triangular diagonal its straight-line phase);
 * gauge 1 is get_phase's own pattern with gauge = 1 (code/quantum.f90:299-326)
 * on the triangular lattice: for the bond from (x, y) to its neighbour j the
 * factor p multiplying t in the amplitude of c^dag_(x,y) c_j is
 * -(-1)^x on (x, y + 1), exp(i theta) on (x + 1, y) and (-1)^x on
 * (x + 1, y - 1) (y wraps when periodic); the reference runs theta = pi/2,
 * -pi/2 of flux in every triangle.  flux must be 0 with gauge 1. */
typedef struct {
  int32_t kind;        /* 0 Hubbard, 1 t-J */
  int32_t Lx, Ly;
  int32_t lattice;     /* 0 square, 1 triangular */
  int32_t periodic_y;
  int32_t n_elec;
  double t, U, J, flux;
  int32_t gauge;       /* 0 Landau, 1 get_phase gauge 1 (triangular only) */
  int32_t reserved;    /* must be 0 */
  double theta;        /* gauge 1: phase of the x bonds */
} hh_model_params;

/* Sweep schedule of Get_Ground_State (code/quantum.f90:179-288). */
typedef struct {
  int32_t kept_max;          /* finite-sweep kept_max */
  int32_t kept_max_warm;     /* warm-up / first sweep kept_max (0: kept_max) */
  int32_t n_sweeps;
  int32_t max_iter;          /* Lanczos vectors (reference: 100) */
  double tol;                /* lanerr of the finite sweeps */
  double warm_tol_factor;    /* reference: 1000 */
  double first_tol_factor;   /* reference: 15 */
  uint64_t seed;
  int32_t precision;         /* 0 FP64 MFMA (plain form) */
  int32_t solver;            /* ground-state solve of every step:
                                0 Lanczos with FP64 H.psi (lanczos_memo_new, the plain form);
                                1 mixed precision: Lanczos rounds with FP32-storage H.psi,
                                  each started by an FP64 H.psi of its start vector, until two
                                  successive FP64 Rayleigh quotients agree to the tolerance;
                                2 the chosen form: Davidson with FP32-storage H.psi, FP64
                                  restarts and the learned per-sector preconditioner (see
                                  hh_driver_train_precond), falling back to solver 1 when the
                                  FP64 residual stops falling; converged when the FP64 Rayleigh
                                  quotient moved by at most the tolerance since the previous
                                  restart, or confirms to the tolerance the last Ritz value of a
                                  round of FP32 expansions that ended by the stopping rule */
} hh_schedule;

/* One DMRG step (a line of Energy_All.dat + Num_State.dat + Entropy.dat,
 * code/lanczos.f90:170-174, code/get_den.f90:336-360). */
typedef struct {
  int32_t phase;             /* 0 warm-up, 1 first sweep, k+1 sweep k */
  int32_t sys_len, env_len;
  int32_t iters;
  int32_t num_state;
  int32_t converged;         /* 1: the Lanczos stopping rule held (code/lanczos.f90:77-80,
                                :102); 0: it stopped at the vector budget (loop_in = 100) */
  int64_t dim;               /* superblock dimension (multiplet space) */
  double energy;
  double trunc_err;
  double entropy;
  double seconds;            /* wall time of the step */
  double matvec_seconds;     /* wall time inside H.psi */
  double matvec_flops;       /* reference-order algorithmic flops of H.psi */
  /* wall time of the phases of the step (stream synchronised at each end) */
  double bind_seconds;       /* Get_basis + Get_super_basis + hmerge */
  double wave_seconds;       /* wave_trans_LR / wave_trans_RL (or the start vector) */
  double lanczos_seconds;    /* lanczos_memo_new, H.psi included */
  double trunc_seconds;      /* density matrix, eigendecomposition, kept-count rule */
  double grow_seconds;       /* Get_Hamiltonian_Sys/Env with update_trun_* */
  double ritz_change;        /* last change of the lowest Ritz value */
  double residual;           /* |H psi - E psi| of the unit ground-state vector (Lanczos:
                                beta_k |z_k|, the remainder of the Krylov relation) */
  int32_t hpsi_fp64;         /* H.psi applications with FP64 arithmetic in the solve */
  int32_t hpsi_fp32;         /* H.psi applications with FP32 arithmetic in the solve */
  int32_t solver;            /* hh_schedule.solver of the step */
  int32_t fallback;          /* 1: the chosen form handed the solve to the mixed Lanczos */
} hh_step_record;

/* ---- context ---------------------------------------------------------- */
int hh_ctx_create(int device, hh_ctx** out);
int hh_ctx_destroy(hh_ctx* ctx);
const char* hh_last_error(const hh_ctx* ctx);
/* couplings / lattice (code/quantum.f90:93-160, code/warm_up.f90:347-666) */
int hh_set_model(hh_ctx* ctx, const hh_model_params* p);
/* sites of a block of `length` sites that keep operators: those with a
 * lattice neighbour outside the block (dmrg_setup0, code/warm_up.f90:311-341;
 * the environment uses the same labels counted from its own end).  *n is set
 * to the count; sites may be null to query it. */
int hh_boundary_sites(hh_ctx* ctx, int length, int32_t* n, int32_t* sites);

/* ---- device memory helpers for callers without their own allocator ----- */
int hh_dev_alloc(hh_ctx* ctx, int64_t bytes, void** dev);
int hh_dev_free(hh_ctx* ctx, void* dev);
int hh_memcpy_h2d(hh_ctx* ctx, void* dev, const void* host, int64_t bytes);
int hh_memcpy_d2h(hh_ctx* ctx, void* host, const void* dev, int64_t bytes);

/* ---- blocks (Total_Model, code/pubdata.f90:167-181) --------------------- */
/* Single-site block with H = U*double (Get_hamiltonian_site,
 * code/warm_up.f90:1208-1244 and warm_up_point :1301-1325). side 0 sys, 1 env. */
int hh_block_site(hh_ctx* ctx, int side, hh_block** out);
/* Empty block with the given sectors (N, 2S, dim) and boundary sites (block
 * index 1..length); the caller uploads H and the operators
 * (model_from_disk, code/space_info.f90:1569-1619). */
int hh_block_create(hh_ctx* ctx, int side, int length, int nsec, const int32_t* N, const int32_t* S2,
                    const int32_t* dim, int nsites, const int32_t* sites, hh_block** out);
int hh_block_destroy(hh_block* b);
/* sizes: arena length (complex elements), sector count, boundary-site count */
int hh_block_info(const hh_block* b, int64_t* arena, int32_t* nsec, int32_t* nsites, int32_t* length);
int hh_block_sectors(const hh_block* b, int32_t* N, int32_t* S2, int32_t* dim, int64_t* h_off);
int hh_block_sites(const hh_block* b, int32_t* sites);
/* layout of operator kind (0 c^dag, 1 S, 2 n): sub-blocks (bra, ket, offset);
 * pass NULL arrays to query nsub.  op_off[site][kind] offsets via hh_block_op_offset. */
int hh_block_layout(const hh_block* b, int kind, int32_t* nsub, int32_t* bra, int32_t* ket, int64_t* off);
int hh_block_op_offset(const hh_block* b, int site_index, int kind, int64_t* off);
/* copy the whole arena in / out (complex128, interleaved); the other side may
 * be host or device memory */
int hh_block_upload(hh_ctx* ctx, hh_block* b, const double* src);
int hh_block_download(hh_ctx* ctx, const hh_block* b, double* dst);

/* ---- superblock: hmerge + task compilation (code/mat_vec.f90:3623-3852) -- */
/* Binds sys and env (borrowed until destroy) at total N and S = 0; builds the
 * enlarged bases (Get_basis, code/update.f90:11-112), the super basis
 * (Get_super_basis, :118-171), the merged boundary operators and the task
 * lists of H.psi, and uploads them.  Geometry = dmrg_setup(sys len, env len). */
int hh_superblock_bind(hh_ctx* ctx, const hh_block* sys, const hh_block* env, int N, hh_superblock** out);
int hh_superblock_destroy(hh_superblock* sb);
/* psi length (complex) and number of (sys, env) sector pairs */
int hh_superblock_info(const hh_superblock* sb, int64_t* dim, int32_t* npairs, double* flops_alg,
                       double* flops_exec);
/* per pair: sys enlarged sector, env enlarged sector, rows, cols, offset */
int hh_superblock_pairs(const hh_superblock* sb, int32_t* sys_sec, int32_t* env_sec, int32_t* rows,
                        int32_t* cols, int64_t* off);
/* enlarged basis of one side: sectors (N, 2S, dim) */
int hh_superblock_enlarged(const hh_superblock* sb, int side, int32_t* nsec, int32_t* N, int32_t* S2, int32_t* dim);

/* wave_to_disk / wave_from_disk (code/space_info.f90:1406-1468): psi of a
 * bound superblock in the reference's own wavefunction file, Fortran
 * sequential unformatted records with 4-byte length markers: sys_len,
 * env_len (enlarged blocks), num, dim, then per super sector an empty record,
 * (N_sys, 2S_sys), (N_env, 2S_env), (sys_dim, env_dim) and the complex*16
 * block.  Reading checks every header field against the superblock and
 * refuses a file that differs or is damaged, leaving psi_dev untouched. */
int hh_wave_to_disk(hh_ctx* ctx, const hh_superblock* sb, const double* psi_dev, const char* path);
int hh_wave_from_disk(hh_ctx* ctx, const hh_superblock* sb, const char* path, double* psi_dev);

/* H.psi: replaces model_wave(sys, env, sys_bs, env_bs, invec, outvec)
 * (code/mat_vec.f90:3476); psi_out is overwritten. */
int hh_model_wave(hh_ctx* ctx, hh_superblock* sb, const double* psi_in_dev, double* psi_out_dev);

/* Batched H.psi over independent superblocks (different flux, doping or
 * geometry): one grouped launch per H.psi stage for all of them.  No
 * counterpart in the reference, which solves one superblock at a time; it is
 * model_wave (code/mat_vec.f90:3476) applied to each member.  At most 32
 * superblocks; they must stay bound while the batch exists. */
int hh_batch_create(hh_ctx* ctx, int nsb, hh_superblock* const* sbs, hh_batch** out);
int hh_batch_destroy(hh_batch* b);
/* total psi length over the members and reference-order flops of one call */
int hh_batch_info(const hh_batch* b, int64_t* total_dim, double* flops_alg);
/* psi_out_dev[i] = H_i psi_in_dev[i] for every member i */
int hh_batch_model_wave(hh_ctx* ctx, hh_batch* b, const double* const* psi_in_dev, double* const* psi_out_dev);
/* lanczos_memo_new (code/lanczos.f90:4-176, levels = 1) of every member at
 * once, in lock step: one grouped FP64 H.psi per iteration for all members
 * (replayed from a HIP graph captured on the batch's first call unless
 * hh_batch_set_graph(b, 0)), the members' vector work in segmented launches,
 * one host round trip for all members' scalars.  Each member follows the
 * reference's stopping rule on its own and stops there; its later products
 * are discarded.  psi_dev[i] holds member i's start vector on entry and its
 * ground state on exit; e0[i], iters[i] and converged[i] (may be null) as
 * hh_lanczos and hh_lanczos_status report them.  No counterpart in the
 * reference, which solves one superblock at a time. */
int hh_batch_lanczos(hh_ctx* ctx, hh_batch* b, double* const* psi_dev, double lanzero, int max_iter, double* e0,
                     int32_t* iters, int32_t* converged);
/* 1: hh_batch_lanczos replays its grouped H.psi from a HIP graph (default);
 * 0: it launches the same kernels one by one */
int hh_batch_set_graph(hh_batch* b, int on);

/* Autotuning of the H.psi launches; no counterpart in the reference.  With
 * mode 1, hh_superblock_bind and hh_batch_create time each H.psi stage's task
 * list under a few ways of cutting it into work items (work items per launch,
 * term groups per output tile, shape classes for edge tiles, tile order) on
 * the superblock's own buffers, and keep the fastest; the choice is cached
 * under the flop-weighted histogram of the list's (M, N, K) shapes in factors
 * of two, so later superblocks of the same shape statistics reuse it without
 * timing.  Mode 0 (default) keeps the fixed cut.  Every cut gives the same
 * H.psi up to the order of the split partial sums (1e-13 relative); a tuned
 * run stays deterministic for a given choice, but close timings can choose
 * differently between runs. */
int hh_set_autotune(hh_ctx* ctx, int mode);
/* shape keys tuned so far, timing runs, cache hits, seconds spent timing */
int hh_autotune_stats(const hh_ctx* ctx, int32_t* keys, int32_t* tunes, int32_t* hits, double* seconds);
/* Force one cut on every H.psi task list uploaded from now on (superblocks
 * bound, batches created), in place of the fixed cut and of the autotuner:
 * about target_items work items per launch, at most max_groups term groups
 * per output tile, edge shape classes (1 on, 0 off, -1 by the 20 % rule),
 * tiles ordered in bands of `band` tile rows, wide 0 (32-column tiles), 1 (a
 * last chunk of at most 16 columns merged into a 48-column tile) or 2 (all
 * 48-column tiles).  target_items = 0 clears it.  No counterpart in the
 * reference; for tests and experiments. */
int hh_set_launch_config(hh_ctx* ctx, int target_items, int max_groups, int edge, int band, int wide);

/* Lanczos: replaces lanczos_memo_new (code/lanczos.f90:4-176) with levels=1:
 * psi_dev holds the start vector on entry and the ground state on exit. */
int hh_lanczos(hh_ctx* ctx, hh_superblock* sb, double* psi_dev, double lanzero, int max_iter, double* e0,
               int* iters);
/* lanczos_memo_new with levels > 1 (code/lanczos.f90:29-44): the `levels`
 * lowest states of the superblock one by one.  psi[l] (device, complex128,
 * super-basis order) holds the start vector of level l and receives its unit
 * eigenvector, e[l] its energy and iters[l] its Lanczos vector count.  Level
 * l runs the Lanczos of levels = 1 with its start vector and every new Lanczos
 * vector projected off the levels below (Gram_Schmidt_Lanczos,
 * code/space_info.f90:1222-1279).  The projection removes u and i u of each
 * lower level u (a complex projection); the reference's real-coefficient one
 * keeps i u, which has u's energy in the real inner-product space, so there
 * level 2 returns E_0 again.  Requires 2 (levels - 1) < max_iter. */
int hh_lanczos_levels(hh_ctx* ctx, hh_superblock* sb, double* const* psi_dev, int levels, double lanzero,
                      int max_iter, double* e, int32_t* iters);
/* Outcome of the last hh_lanczos call on ctx: whether the stopping rule of
 * lanczos_memo_new (code/lanczos.f90:77-80, :102) held before max_iter, the
 * last change of the lowest Ritz value and |H psi - E psi| of the returned
 * unit vector (beta_k |z_k|).  The reference has no such output: it writes
 * the energy only (Energy_All.dat, :170-174). */
int hh_lanczos_status(const hh_ctx* ctx, int32_t* converged, double* ritz_change, double* residual);

/* Get_density_matrix_sys_new / Get_density_matrix_env_new
 * (code/get_den.f90:106-177, levels = 1): the reduced density matrix of the
 * enlarged block of `side` (0 sys: psi psi^dag, 1 env: psi^T conj(psi)), one
 * block per psi sector in super-basis order.  hh_density_matrix_layout gives
 * the block count and the total size in complex elements; when sector, dim
 * and off are not null it also fills, per block, the enlarged sector index,
 * the dimension and the element offset.  hh_density_matrix writes the blocks
 * (complex128, column-major, dim x dim at off) to rho_dev. */
int hh_density_matrix_layout(const hh_superblock* sb, int side, int32_t* nblk, int64_t* size, int32_t* sector,
                             int32_t* dim, int64_t* off);
int hh_density_matrix(hh_ctx* ctx, hh_superblock* sb, int side, const double* psi_dev, double* rho_dev);

/* Reduced density matrix + Get_truncation_operator
 * (code/get_den.f90:4-99, :184-428): side 0 sys (psi psi^dag), 1 env. */
/* Get_density_matrix_sys_new / _env_new with levels >= 2
 * (code/get_den.f90:103-177): (0.1 rho(psi1) + 1.9 rho(psi2)) / 2 in the
 * layout of hh_density_matrix, the weights being the reference's
 * single-precision literals (trace 1 - 1.1e-8); hh_truncate_levels is
 * hh_truncate on that mixed density matrix. */
int hh_density_matrix_levels(hh_ctx* ctx, hh_superblock* sb, int side, const double* psi1_dev,
                             const double* psi2_dev, double* rho_dev);
int hh_truncate_levels(hh_ctx* ctx, hh_superblock* sb, int side, const double* psi1_dev, const double* psi2_dev,
                       int kept_min, int kept_max, double trun_err, hh_trunc** out, int* num_state, double* err,
                       double* entropy);
int hh_truncate(hh_ctx* ctx, hh_superblock* sb, int side, const double* psi_dev, int kept_min, int kept_max,
                double trun_err, hh_trunc** out, int* num_state, double* err, double* entropy);
/* The eigendecomposition inside Get_truncation_operator on its own (ZHEEV
 * per sector, code/get_den.f90:242): nsec Hermitian blocks of dim[i] x dim[i]
 * complex128, column-major, packed one after another at blocks_dev (the upper
 * triangle is read).  evals_dev receives every eigenvalue, ascending per
 * block, blocks concatenated.  On return block i holds the eigenvectors of
 * its top keep[i] eigenvalues as its last keep[i] columns, in ascending
 * order; with solver 0 (the hand-written eigensolver) every other entry of
 * the block is zero, with solver 1 (rocSOLVER zheevd) the other columns hold
 * the other eigenvectors.  Status 3: the hand-written solver handed the
 * problem back (a block above 3072, or kept vectors it could not
 * orthonormalise). */
int hh_eigh_sectors(hh_ctx* ctx, int nsec, const int32_t* dim, const int32_t* keep, double* blocks_dev,
                    double* evals_dev, int solver);
/* Eigensolver of hh_truncate and the drivers: 0 hand-written (default), 1
 * rocSOLVER zheevd per sector.  hh_eigensolver_status reports it, how many
 * truncations the hand-written solver handed back to zheevd so far, and the
 * orthonormalisation steps and max |U^dag U - I| of its last call.  No
 * counterpart in the reference (which calls ZHEEV). */
int hh_set_eigensolver(hh_ctx* ctx, int solver);
int hh_eigensolver_status(const hh_ctx* ctx, int32_t* solver, int64_t* handbacks, int32_t* ortho_steps,
                          double* ortho_error);
/* truns_to_disk / truns_from_disk (code/space_info.f90:1621-1661, the
 * block_to_disk / block_from_disk records of :1470-1521) of the truncation
 * operator of Get_truncation_operator (code/get_den.f90:362-421): len, num,
 * dim, up_dif, down_dif, then per kept sector an empty record, (N, 2S, 0),
 * (row_dim, sdim) and the kept eigenvector block.  length is the length of
 * the block the truncation produces.  Reading needs the block it was grown
 * from (its enlarged basis); sectors are matched by (N, 2S) and every field
 * is checked.  hh_grow_block then rebuilds the reference's block from it.
 * Environment truncations are stored conjugated, as the reference's
 * environment density matrix is psi^dag psi (code/get_den.f90:84-91). */
int hh_trunc_to_disk(hh_ctx* ctx, const hh_trunc* t, int length, const char* path);
int hh_trunc_from_disk(hh_ctx* ctx, const hh_block* prev, int side, const char* path, int32_t* length,
                       hh_trunc** out);
/* basis_to_disk / basis_from_disk (code/space_info.f90:1665-1762): the
 * enlarged basis of blk and one site (Get_basis, code/update.f90:11-112) as
 * the reference's formatted text file (len, num, dim; per sector an empty
 * record, idx, num, dim, new_num_up = N, new_num_down = 2S; per subsector
 * pos, dim and the (N, 2S) of its block and site multiplets).  Writing
 * refuses a value wider than the reference's I6 / I4 fields.  Reading checks
 * every field against Get_basis of blk (status 1 names the first line that
 * differs); *length receives the file's len (blk's length + 1). */
int hh_basis_to_disk(hh_ctx* ctx, const hh_block* blk, const char* path);
int hh_basis_from_disk(hh_ctx* ctx, const hh_block* blk, const char* path, int32_t* length);
/* identity truncation of the warm-up (code/warm_up.f90:1421-1432) */
int hh_trunc_identity(hh_ctx* ctx, hh_superblock* sb, int side, hh_trunc** out);
int hh_trunc_destroy(hh_trunc* t);
/* kept multiplets per enlarged sector (enlarged-basis order of that side) */
int hh_trunc_kept(const hh_trunc* t, int32_t* nsec, int32_t* sector, int32_t* keep);
/* kept eigenvectors U of kept sector i: columns of a (ld x keep[i]) block at
 * element offset u_off[i] of the truncation's arena (arena length in *size;
 * pass NULL arrays to query it); hh_trunc_download copies the arena out */
int hh_trunc_layout(const hh_trunc* t, int64_t* size, int64_t* u_off, int32_t* u_ld);
int hh_trunc_download(hh_ctx* ctx, const hh_trunc* t, double* dst);

/* update_trun_dia / update_trun_ndia (code/update.f90:1387-1605) for one
 * operator: eff = U^dag O U per sub-block with the kept eigenvectors U of the
 * truncation t.  The operator is given in the truncation's enlarged basis:
 * nsub sub-blocks (bra, ket enlarged-sector indices, column-major dim_bra x
 * dim_ket at element offset off of op_dev).  The result holds the
 * sub-blocks whose two sectors are both kept, in input order, indexed by
 * kept sector (the new block's sector order), keep_bra x keep_ket each,
 * packed at out_off of out_dev.  With out_dev null only *out_nsub,
 * *out_size (complex elements) and the out arrays are filled. */
int hh_renormalize(hh_ctx* ctx, const hh_trunc* t, int nsub, const int32_t* bra, const int32_t* ket,
                   const int64_t* off, const double* op_dev, int32_t* out_nsub, int64_t* out_size, int32_t* out_bra,
                   int32_t* out_ket, int64_t* out_off, double* out_dev);

/* Complex-conjugate a block's whole arena (H and every operator) in place.
 * The reference stores environment-side operators conjugated and applies
 * them as psi . B^dag (block_transfer_trans, code/space_info.f90:32-63;
 * code/mat_vec.f90:3563, :3590); this library stores the physical operator
 * on both sides and applies psi . B^T, so an environment block uploaded from
 * the reference's arrays is conjugated once with this call. */
int hh_block_conjugate(hh_ctx* ctx, hh_block* b);

/* Block growth + renormalisation: replaces Get_Hamiltonian_Sys/Env
 * (code/get_ham.f90:4-659) and update_trun_dia/ndia (code/update.f90:1387-1605):
 * the side's block of sb grown by its centre site and truncated with t. */
int hh_grow_block(hh_ctx* ctx, hh_superblock* sb, int side, const hh_trunc* t, hh_block** out);

/* psi prediction between steps: replaces wave_trans_LR / wave_trans_RL
 * (code/warm_up.f90:1855-2170).  dir 0 = LR, 1 = RL.  sys_t / env_t are the
 * truncations of the old sys (LR: this step's; RL: the one that built the old
 * sys block) and of the env (LR: the one that built the old env block; RL:
 * this step's).  psi_new_dev is overwritten and normalised. */
int hh_wave_trans(hh_ctx* ctx, int dir, hh_superblock* old_sb, const double* psi_old_dev, const hh_trunc* sys_t,
                  const hh_trunc* env_t, hh_superblock* new_sb, double* psi_new_dev);

/* Whole ground-state search with the reference schedule (Get_Ground_State,
 * code/quantum.f90:179-288): warm-up, centre-out half sweep, n_sweeps sweeps.
 * Fills up to max_rec records; *nrec = steps done. */
int hh_dmrg_run(hh_ctx* ctx, const hh_schedule* s, hh_step_record* rec, int max_rec, int* nrec);

/* ---- step-wise driver: the sweep routines one call at a time ----------- */
typedef struct hh_driver hh_driver;
int hh_driver_create(hh_ctx* ctx, const hh_schedule* s, hh_driver** out);
int hh_driver_destroy(hh_driver* d);
/* warm_up_point(1, Nsite/2) (code/warm_up.f90:1284-1491) */
int hh_driver_warmup(hh_driver* d, hh_step_record* rec, int max_rec, int* nrec);
/* left_to_right (dir 0, code/warm_up.f90:1495-1671) or right_to_left (dir 1,
 * :1675-1849) from label bg to end; fresh = 1 starts the first step from a
 * new random vector instead of the carried / transformed one.  Blocks and
 * truncations stay resident on the device between calls. */
int hh_driver_sweep(hh_driver* d, int dir, int bg, int end, double lanzero, double trun_err, int kept_max, int fresh,
                    int phase, hh_step_record* rec, int max_rec, int* nrec);
/* checkpoint / restart: write every stored block (H and operators), every
 * truncation, the carried wave function and the step counter of the driver
 * to one file, and read it back into a driver of a context with the same
 * model -- the content of model_to_disk / truns_to_disk / wave_to_disk
 * (code/space_info.f90:1406-1762) for restart.dat's resume
 * (code/quantum.f90:43-47, :207-253), in this library's own binary layout.
 * After a load the next step starts from the carried wave function (as after
 * hh_driver_reset_transform); a sweep resumed from a checkpoint written
 * between sweeps repeats the uninterrupted run bit for bit. */
int hh_driver_save(hh_driver* d, const char* path);
int hh_driver_load(hh_driver* d, const char* path);
/* forget the previous superblock (next first step does not transform) */
int hh_driver_reset_transform(hh_driver* d);
/* the ground-state solver of the steps that follow (hh_schedule.solver);
 * the reference always runs lanczos_memo_new (code/lanczos.f90:4-176) */
int hh_driver_set_solver(hh_driver* d, int solver);
/* An independent copy of a driver (blocks, truncations, carried wave
 * function, previous superblock, step counter and schedule), made on the
 * device; both continue on their own.  No counterpart in the reference: it
 * lets one state be swept by two solvers. */
int hh_driver_fork(const hh_driver* src, hh_driver** out);
/* hh_driver_fork holding only the blocks and truncations of sys_lo..sys_hi
 * sites on the sys side and env_lo..env_hi on the env side (the current
 * superblock's blocks must lie inside): the copy for a few steps that touch
 * only those, at a fraction of the memory.  A step that needs a block
 * outside the window fails with status 1.  No counterpart in the reference
 * (whose blocks live on disk, model_to_disk, code/space_info.f90:1523). */
int hh_driver_fork_window(const hh_driver* src, int sys_lo, int sys_hi, int env_lo, int env_hi, hh_driver** out);

/* ---- learned preconditioner of the chosen form (solver 2) ----------------
 * No counterpart in the reference, whose steps all use lanczos_memo_new
 * (code/lanczos.f90:4-176).  The Davidson expansion vector of sector block p
 * (element i) is t_i = alpha_p r_i / (max(d_i - theta, 0) + sigma_p), with r
 * the residual, theta the Ritz value and d the diagonal of the sys and env
 * block Hamiltonians in the enlarged bases; (sigma_p, alpha_p) come from an
 * MLP (8 -> 256 -> 256 -> 2, tanh, FP32) on features of block p, as a
 * correction of the fixed diagonal form (sigma = 0.1 x the spread of d,
 * alpha = 1) by factors within [1/4, 4].  Training
 * triples (diagonal, residual of the step's start vector, exact correction
 * toward the step's ground state) are captured from plain solves.
 *
 * hh_driver_capture: the steps that follow record training triples until
 * max_samples are kept on the device: a step solved by solver 0 records its
 * start vector (exact correction toward the step's ground state), a step
 * solved by solver 2 up to three of its FP64 restart points.
 * hh_driver_train_precond: epochs passes of Adam (learning rate lr) over the
 * captured triples with the library's own HIP forward and backward kernels;
 * losses (epochs doubles, may be null) receives the mean loss
 * 1 - Re<t, delta>^2 / (|t|^2 |delta|^2) per epoch, *nsamples the number of
 * triples.
 * hh_driver_copy_precond: dst takes a copy of src's network (parameters and
 * Adam state; both drivers on one context), so one trained network serves
 * several drivers.
 * hh_driver_set_precond: 0 the trained network (the fixed diagonal form while
 * none is trained), 1 the fixed diagonal form (sigma = 0.1 x the spread of
 * d, alpha = 1), 2 no preconditioner (t = r). */
int hh_driver_capture(hh_driver* d, int max_samples);
int hh_driver_train_precond(hh_driver* d, int epochs, double lr, double* losses, int32_t* nsamples);
int hh_driver_set_precond(hh_driver* d, int mode);
int hh_driver_copy_precond(hh_driver* dst, const hh_driver* src);
/* Fallback rule of solver 2: at every FP64 restart the exact residual must
 * be below factor times its value at the previous restart (default 0.9), or
 * the mixed-precision Lanczos (solver 1) finishes the solve from the current
 * vector; the step record's fallback flag is then set. */
int hh_driver_set_fallback(hh_driver* d, double factor);
/* Solvers 1 and 2 hand superblocks of dimension below min_dim to the plain
 * Lanczos (solver 0); default 262144, 0 = never.  Below that size the H.psi
 * is a small part of the step and the plain form is the faster one. */
int hh_driver_set_min_dim(hh_driver* d, int64_t min_dim);
/* Largest Davidson subspace of solver 2 before it restarts from its Ritz
 * vector (default 32, at most 128); the workspace holds twice that many
 * vectors.  No counterpart in the reference, whose lanczos_memo_new keeps up
 * to loop_in = 100 Lanczos vectors (code/lanczos.f90:14, :22). */
int hh_driver_set_subspace(hh_driver* d, int kmax);
/* Number of levels every later warm-up and sweep step targets (default 1, at
 * most 8; the levels argument of warm_up_point / left_to_right /
 * right_to_left, code/warm_up.f90:1284-1296, :1495-1508, carried from
 * Get_Ground_State, code/quantum.f90:179-184).  With levels > 1 each step
 * solves the levels as hh_lanczos_levels does (lanczos_memo_new with levels,
 * code/lanczos.f90:29-44), starting level 1 from the wave-transformed state
 * and levels 2.. from seeded vectors (wave_initialize after the wave
 * transformation, code/warm_up.f90:1616-1620), and truncates on the mixed
 * density matrix of levels 1 and 2 (Get_density_matrix_*_new,
 * code/get_den.f90:106-177; hh_density_matrix_levels).  The step record's
 * energy, iterations, convergence and residual are level 1's, its H.psi
 * count and time those of all levels; hh_driver_level_energies gives every
 * level's energy.  Several levels need solver 0.  hh_driver_save keeps the
 * setting and the upper levels' vectors (the reference writes one wave file
 * per level), hh_driver_load restores both.  A superblock of dimension n
 * solves at most n levels and reports NaN for the rest. */
int hh_driver_set_levels(hh_driver* d, int levels);
/* Start vectors of levels 2.. at a step with a wave transformation: 0 seeded
 * vectors (default, the reference's: it transforms every level,
 * code/warm_up.f90:1606-1611, then re-initialises levels 2.. with
 * wave_initialize, :1616-1620), 1 the transformed vectors themselves.  The
 * converged levels are the same; carried starts need fewer H.psi.  No
 * counterpart as a setting in the reference; not kept in a checkpoint. */
int hh_driver_set_level_start(hh_driver* d, int carry);
/* The levels setting of d (1 unless set, or restored by hh_driver_load). */
int hh_driver_levels(const hh_driver* d, int32_t* levels);
/* The level energies of the records the last warmup / sweep call returned,
 * levels per record (record-major, the levels set at that call) into e,
 * at most max_rec records; *nrec = that call's record count. (Save_result_to_disk
 * writes Energys(1:levels), code/quantum.f90:794-805.) */
int hh_driver_level_energies(const hh_driver* d, int max_rec, double* e, int* nrec);
/* The preconditioner's whole training pass on caller data, for checking the
 * kernels: np sector blocks (rows x cols, 2S and N of the sys sector), the
 * diagonal d(a, b) = hs[a] + he[b] per block (hs, he concatenated over the
 * blocks), residual r and target delta (complex128, interleaved, blocks in
 * order), Ritz value theta and network parameters w (hh_precond_size()
 * floats).  Returns the features (np x 8), network outputs (np x 2), sector
 * parameters (sigma_p, alpha_p), the loss and its gradient with respect to
 * w; any output pointer may be null. */
int hh_precond_eval(hh_ctx* ctx, int np, const int32_t* rows, const int32_t* cols, const int32_t* s2,
                    const int32_t* nq, int n_mid, const double* hs, const double* he, const double* r,
                    const double* delta, double theta, const float* w, float* feat, float* out, double* par,
                    double* loss, float* grad);
int hh_precond_size(void);
/* The ground state of the last sweep step (the eigvec that left_to_right /
 * right_to_left hand to wave_trans_LR / wave_trans_RL, code/warm_up.f90:1607,
 * :1786): the lengths of the sys and env blocks its superblock was bound from
 * and its dimension; when psi_dev is not null the vector (complex128, super-
 * basis order) is copied there.  Fails when no sweep step has run since the
 * driver was created, loaded or reset. */
int hh_driver_wave(const hh_driver* d, int32_t* sys_len, int32_t* env_len, int64_t* dim, double* psi_dev);
int hh_driver_block_count(const hh_driver* d, int side, int32_t* count);
/* borrowed handle of a stored block (side 0 sys / 1 env, by length); valid
 * until the driver replaces or frees it */
int hh_driver_block(const hh_driver* d, int side, int length, const hh_block** out);
/* borrowed handle of the stored truncation that made the block of `length` */
int hh_driver_trunc(const hh_driver* d, int side, int length, const hh_trunc** out);
/* measure_operator_dia (code/measure.f90:3-61) with the operator chain of
 * Get_sys_oper_dia / Get_env_oper_dia (the stored truncations of every
 * length): <n_i> for every lattice position i, n_out[i-1], in the ground
 * state psi_dev of the superblock bound from the driver's sys block of
 * sys_len sites and env block of Nsite - sys_len - 2 sites. */
int hh_driver_density(hh_driver* d, int sys_len, const double* psi_dev, double* n_out);
/* density-density correlations <n_i n_j> across the superblock's centre
 * (i on the sys side, j on the env side, centre sites included): the pairs of
 * Get_density_cor (code/correlation.f90:43-195) that straddle the two
 * blocks, i.e. sys_block_env_block_cor_ndia (:1666-1929),
 * sys_block_env_site_cor_ndia (:1932-2169), sys_site_env_site_cor_ndia
 * (:2172-2402) and sys_site_env_block_cor_ndia (:2405-2656) for rank-0
 * operators.  nn_out is Nsite x Nsite row-major by lattice position - 1,
 * symmetric, NaN for pairs on one side (those straddle another cut). */
int hh_driver_density_cross(hh_driver* d, int sys_len, const double* psi_dev, double* nn_out);
/* one row of Get_density_cor (code/correlation.f90:43-195) with
 * oper1 = oper2 = the density: <n_site n_j> for every lattice position j
 * (nn_out[j-1]; NaN at j = site), for a row site on the sys side of the cut
 * (1 .. sys_len + 1).  Pairs inside the sys block use the two-site operator
 * carried through the truncations (Get_sys2b, :1-40, measured as
 * measure_sys_block); the block site with the sys centre site is
 * sys_block_site_cor_ndia (:1335-1489); the env side as in
 * hh_driver_density_cross.  The reference evaluates this row for j on the
 * same leg as the row site; every j is returned here. */
int hh_driver_density_row(hh_driver* d, int sys_len, int site, const double* psi_dev, double* nn_out);
/* spin correlations <S_i . S_j> across the superblock's centre (i on the sys
 * side, j on the env side, centre sites included): Get_density_cor
 * (code/correlation.f90:43-195) with oper1 = oper2 = st_sd, its pairs
 * sys_block_env_block_cor_ndia (:1666-1929), sys_block_env_site_cor_ndia
 * (:1932-2169), sys_site_env_site_cor_ndia (:2172-2402) and
 * sys_site_env_block_cor_ndia (:2405-2656), with the spin of every block site
 * carried through the stored truncations (Get_sys_oper_ndia /
 * update_trun_ndia).  Returns the physical <S_i . S_j>.  ss_out as in
 * hh_driver_density_cross. */
int hh_driver_spin_cross(hh_driver* d, int sys_len, const double* psi_dev, double* ss_out);
/* one row of Get_density_cor (code/correlation.f90:43-195) with oper1 =
 * oper2 = st_sd: <S_site . S_j> for every lattice position j (ss_out[j-1];
 * NaN at j = site), row site on the sys side (1 .. sys_len + 1).  Block
 * pairs use the product S_r . S_j carried through the truncations (Get_sys2b,
 * :1-40); a block site with the sys centre site is sys_block_site_cor_ndia
 * (:1335-1489); the env side as in hh_driver_spin_cross. */
int hh_driver_spin_row(hh_driver* d, int sys_len, int site, const double* psi_dev, double* ss_out);
/* single-particle correlations sum_sigma <c^dag_i,sigma c_j,sigma> across the
 * superblock's centre: Get_density_cor (code/correlation.f90:43-195) with
 * st_elec_up / st_elec_down and flag 'F', its pairs across the centre, with
 * c^dag of every block site carried through the stored truncations (fermion
 * signs of update_block_ndia_* / update_site_ndia_*).  g_out is Nsite x Nsite
 * complex (re, im interleaved), row-major by lattice position - 1; the entry
 * (j, i) holds the conjugate; pairs on one side are NaN. */
int hh_driver_hopping_cross(hh_driver* d, int sys_len, const double* psi_dev, double* g_out);
/* one row of the same correlation: sum_sigma <c^dag_site,sigma c_j,sigma>
 * for every lattice position j (complex, g_out[2(j-1)], g_out[2(j-1)+1];
 * NaN at j = site), row site on the sys side (1 .. sys_len + 1); block pairs
 * through the carried two-site product (Get_sys2b), block-centre as
 * sys_block_site_cor_ndia, env side as hh_driver_hopping_cross. */
int hh_driver_hopping_row(hh_driver* d, int sys_len, int site, const double* psi_dev, double* g_out);
/* every pair of sites and the on-site values at once: Get_oper_cor_ndia1
 * (code/correlation.f90:2891-3107, the measbl path of Measurement,
 * code/quantum.f90:556-575) with oper = oper1 = n.  Pairs inside one block
 * as Get_sys_oper_cor_ndia1 / Get_env_oper_cor_ndia1 (:851-974,
 * :1097-1216): each product formed when its later site joins and carried
 * through the stored truncations; a block site with its side's centre site
 * as sys_block_site_cor_ndia / env_block_site_cor_ndia; pairs across the
 * centre as hh_driver_density_cross; on-site <n_i^2>.  nn_out is Nsite x
 * Nsite, symmetric, row-major by lattice position - 1. */
int hh_driver_density_all(hh_driver* d, int sys_len, const double* psi_dev, double* nn_out);
/* the same for oper = oper1 = the spin: <S_i . S_j> for every pair, on-site
 * <S_i . S_i> = 3/4 <n_i - 2 n_i,up n_i,down>. */
int hh_driver_spin_all(hh_driver* d, int sys_len, const double* psi_dev, double* ss_out);
/* the same for st_elec_up / st_elec_down, flag 'F': sum_sigma
 * <c^dag_i,sigma c_j,sigma> for every pair, on-site <n_i>.  g_out is Nsite x
 * Nsite complex (re, im interleaved), Hermitian. */
int hh_driver_hopping_all(hh_driver* d, int sys_len, const double* psi_dev, double* g_out);
/* singlet pair correlations <Delta^dag_a Delta_b>, Delta^dag_ij =
 * (c^dag_i,up c^dag_j,down - c^dag_i,down c^dag_j,up)/sqrt(2), for bonds a
 * inside the sys block and b inside the env block: the block-block part of
 * Get_mSCOP_Operator_cor0 (code/correlation.f90:198-327), with the pair
 * creators formed when their later site joins (Get_sys_two, :331-367) and
 * carried through the stored truncations, measured as
 * sys_block_env_block_cor_ndia (:1666-1929).  bonds holds nbond pairs of
 * lattice positions; p_out is nbond x nbond complex (re, im interleaved),
 * row-major; entry (b, a) holds the conjugate; pairs that are not
 * (sys block, env block) are NaN. */
int hh_driver_pair_cross(hh_driver* d, int sys_len, const double* psi_dev, int nbond, const int32_t* bonds,
                         double* p_out);
/* the same-block part of Get_mSCOP_Operator_cor0: <Delta^dag_a Delta_b>
 * for bonds a = (s1, s2) and b = (s3, s4) inside the sys block (lattice
 * positions 1 .. sys_len) with every site of a before every site of b, built
 * as Get_sys_two1 (code/correlation.f90:370-428) does: the pair creator of a
 * carried through the truncations, multiplied by the annihilator of s3 when
 * s3 joins and coupled with that of s4 to a scalar when s4 joins, then read
 * off the block (measure_sys_block, code/measure.f90:611-673).  quads holds
 * nquad x 4 positions (s1, s2, s3, s4); p_out receives nquad complex values
 * (re, im interleaved). */
int hh_driver_pair_block(hh_driver* d, int sys_len, const double* psi_dev, int nquad, const int32_t* quads,
                         double* p_out);

/* ---- timing of the H.psi kernels (HIP events on the context stream) ------ */
typedef struct {
  int64_t matvecs;           /* H.psi calls since the last reset */
  double matvec_ms;          /* wall time of those calls */
  double stage1_ms;          /* kernel time of the stage-1 (A.psi) tile launches */
  double stage2_ms;          /* kernel time of the stage-2 (all terms) tile launches */
  double flops_alg;          /* reference-order algorithmic flops of those calls */
  double flops_exec;         /* flops the engine executed */
  double bytes_alg;          /* compulsory bytes: psi in/out + operator sub-blocks */
  double stage1_flops_alg;   /* part of flops_alg done by the stage-1 launches */
  double stage2_flops_alg;   /* part of flops_alg done by the stage-2 launches */
  double reduce_ms;          /* kernel time of the split-sum reductions of both stages */
  double stage1_flops_exec;  /* part of flops_exec run by the stage-1 launches */
  double stage2_flops_exec;  /* part of flops_exec run by the stage-2 launches */
  int64_t matvecs_fp32;      /* of matvecs: those in the FP32-storage form */
  double stage2_ms_fp32;     /* of stage2_ms: the FP32-storage launches */
  double stage2_flops_alg_fp32;  /* of stage2_flops_alg: the FP32-storage launches */
} hh_stats;
/* Arithmetic of the H.psi tile kernels for the calls that follow on this
 * context: 0 FP64 MFMA (plain form), 1 FP32 MFMA on FP64-stored operands
 * with FP64 output, 2 FP32 storage: complex64 copies of the blocks, merged
 * operators, stage-1 intermediate and vectors, FP32 MFMA with FP32
 * accumulation (the vectors passed in and out stay complex128).  Lanczos
 * vectors and reductions stay FP64. */
int hh_set_precision(hh_ctx* ctx, int precision);
/* profile = 1 records HIP events around every H.psi launch (adds a sync).
 * While it is on, the H.psi tile kernels run as their own instantiations
 * (k_tiles<double, 3, ...> stage 1, k_tiles<double, 4, ...> stage 2, identical code),
 * so a kernel-trace profile separates the timed launches from all others. */
int hh_set_profiling(hh_ctx* ctx, int profile);
int hh_get_stats(hh_ctx* ctx, hh_stats* out);
int hh_reset_stats(hh_ctx* ctx);

/* ---- host-logic inspection (no GPU work) -------------------------------- */
/* Task lists of a bound superblock: stage 0 merges, 1 stage-1 products,
 * 2 H.psi terms.  Arrays as in the engine: tiles (c_off, c_ld, M, N, t0, t1,
 * c_buf, accumulate) and terms (see hh_core.hpp).  Pointers stay valid until
 * the superblock is destroyed. */
int hh_superblock_tasks(const hh_superblock* sb, int stage, const void** tiles, int32_t* ntiles,
                        const void** terms, int32_t* nterms, int64_t* scratch);
/* Host-only plans: the same task compilers without a device, for checking
 * the host logic.  Blocks are described by sectors (N, 2S, dim) and boundary
 * sites; which = 0 sys / 1 env / 2 grown block. */
typedef struct hh_plan hh_plan;
int hh_plan_matvec(const hh_model_params* m, int sys_len, int sys_nsec, const int32_t* sys_N, const int32_t* sys_S2,
                   const int32_t* sys_dim, int sys_nsites, const int32_t* sys_sites, int env_len, int env_nsec,
                   const int32_t* env_N, const int32_t* env_S2, const int32_t* env_dim, int env_nsites,
                   const int32_t* env_sites, int N, hh_plan** out);
/* Growth plan of one block (side 0/1) in the geometry (ns, ne) with kept
 * sectors (enlarged index, keep, U offset, U leading dimension). */
int hh_plan_grow(const hh_model_params* m, int side, int length, int nsec, const int32_t* Nq, const int32_t* S2q,
                 const int32_t* dimq, int nsites, const int32_t* sites, int ns, int ne, int nkept,
                 const int32_t* kept_a, const int32_t* kept_n, const int64_t* u_off, const int32_t* u_ld,
                 hh_plan** out);
int hh_plan_destroy(hh_plan* p);
int hh_plan_tasks(const hh_plan* p, int stage, const void** tiles, int32_t* ntiles, const void** terms,
                  int32_t* nterms, int64_t* scratch);
int hh_plan_block(const hh_plan* p, int which, int64_t* arena, int32_t* nsec, int32_t* N, int32_t* S2, int32_t* dim,
                  int64_t* h_off, int32_t* nsites, int32_t* sites);
int hh_plan_layout(const hh_plan* p, int which, int kind, int32_t* nsub, int32_t* bra, int32_t* ket, int64_t* off);
int hh_plan_op_offset(const hh_plan* p, int which, int site_index, int kind, int64_t* off);
/* Enlarged basis of side 0 (sys) / 1 (env) of a matvec plan, as Get_basis
 * (code/update.f90:11-112) orders it: sectors (N, 2S, dim) in QuickSort_Integer
 * order (2S descending, then N descending, code/update.f90:240-273), and per
 * subsector {enlarged sector, block sector, site sector, spos, sdim} in
 * block-major, site-minor order (subs: nsub x 5 int32; NULL to query). */
int hh_plan_enlarged(const hh_plan* p, int side, int32_t* nsec, int32_t* N, int32_t* S2, int32_t* dim, int32_t* nsub,
                     int32_t* subs);
int hh_plan_super(const hh_plan* p, int64_t* dim, int32_t* npairs, int32_t* sys_sec, int32_t* env_sec, int32_t* rows,
                  int32_t* cols, int64_t* off);

/* Run one tile list on the device: bufs_dev[8] are the buffer slots the
 * terms address (hh_core.hpp: Buf); precision 0 FP64, 1 FP32.  The engine
 * entry point behind every task list above, exposed for direct use. */
int hh_run_tasks(hh_ctx* ctx, const void* tiles, int32_t ntiles, const void* terms, int32_t nterms,
                 void* const* bufs_dev, int precision);

/* hh_run_tasks without any device allocation.  No counterpart in the
 * reference; it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The task tables, work items and split-sum partial buffer go
 * into a fixed 64 MiB device array of the library (held on every GPU the
 * library runs on until the process ends); a task list that needs more
 * returns status 1.  The caller still passes every input and output buffer.
 * All calls on one GPU share that array: calls that overlap on one GPU race. */
int hh_run_tasks_noalloc(hh_ctx* ctx, const void* tiles, int32_t ntiles, const void* terms, int32_t nterms,
                         void* const* bufs_dev, int precision);

/* Counter-based start vector (splitmix64 stream) written to psi_dev. */
int hh_start_vector(hh_ctx* ctx, double* psi_dev, int64_t n, uint64_t seed);

#ifdef __cplusplus
}
#endif
#endif /* HHDMRG_H */
