/*
 * nssdmrg.h -- C ABI of the MI355X (gfx950) superblock path of nssDMRG.
 *
 * Plain pointers, sizes and status codes only.  Every entry point returns 0 on
 * success and a nonzero NSS_E_* code on failure; nss_last_error() then holds a
 * message.  A context is single-caller; its kernels run on one HIP stream.
 *
 * Layout convention shared with the reference (src/DMRG_SUPERBLOCK.f90:72,
 * src/DMRG_SUPERBLOCK_SETUP.f90:770): the superblock vector is the
 * concatenation over QN sectors q of Psi_q, a Dl_q x Dr_q matrix stored with
 * the right index fastest (C row-major Psi_q[il][ir] == Fortran V_q(ir,il)),
 * starting at offset_q = sum_{p<q} Dl_p*Dr_p.
 */
#ifndef NSSDMRG_H
#define NSSDMRG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  NSS_OK = 0,
  NSS_E_ARG = 1,      /* invalid argument / shape mismatch */
  NSS_E_HIP = 2,      /* HIP runtime or device failure */
  NSS_E_STATE = 3,    /* call out of order (e.g. hv before layout) */
  NSS_E_NOCONV = 4,   /* eigen-solver did not converge */
  NSS_E_NUMERIC = 5,  /* numerical invariant violated (e.g. rho not symmetric) */
  NSS_E_ALLOC = 6     /* device allocation failed */
};

enum { NSS_FP64 = 0, NSS_FP32_MIXED = 1 };           /* H*v precision */
enum { NSS_ROW_MAJOR = 0, NSS_COL_MAJOR = 1 };       /* dense block layout */

typedef struct nss_ctx nss_ctx;
typedef struct nss_dmrg nss_dmrg;

/* Message of the last failing call on this thread ("" if none). */
const char* nss_last_error(void);

/* Context: one device, one stream, a caching device allocator.
 * Replaces nothing in the reference (it holds the module state that the
 * reference keeps in src/DMRG_GLOBAL.f90:41-124). */
int nss_ctx_create(int device, nss_ctx** out);
int nss_ctx_destroy(nss_ctx* ctx);
/* 1 if a gfx950 device with the built code object is usable, else 0. */
int nss_device_ok(int device);

/* ------------------------------------------------------------------------
 * Low-level superblock H*v: the drop-in for spMatVec_direct_main
 * (src/DMRG_SUPERBLOCK_SETUP.f90:740-865), i.e. for the procedure pointer
 * spHtimesV_p of abstract interface sparse_HxV (src/DMRG_GLOBAL.f90:41-53).
 * ---------------------------------------------------------------------- */

/* Sector layout: Dls/Drs/Offset of sb_build_dims (src/DMRG_SUPERBLOCK.f90:36-121).
 * offset may be NULL (then the contiguous prefix sum is used). */
int nss_sb_set_layout(nss_ctx* ctx, int nq, const int32_t* Dl, const int32_t* Dr,
                      const int64_t* offset);

/* Operator blocks of Setup_SuperBlock_Direct (SETUP:113-693), as dense host
 * arrays: HL[q] (Dl_q x Dl_q), HR[q] (Dr_q x Dr_q); for each term t,
 * A[t] (Dl_dst x Dl_src) and B[t] (Dr_dst x Dr_src) with
 *     Hv_dst += A_t . Psi_src . B_t^T
 * (dst/src play the roles of RowOffset/ColOffset, SETUP:289-359,601-637).
 * layout selects row- or column-major storage for every block (Fortran
 * callers pass NSS_COL_MAJOR and their own arrays).  precision: NSS_FP64 or
 * NSS_FP32_MIXED (FP32 MFMA copies are made as well).  Arrays are copied;
 * the device copies live until the next set_layout or ctx destroy. */
int nss_sb_upload_blocks(nss_ctx* ctx, const double* const* HL, const double* const* HR,
                         int nterms, const int32_t* dst, const int32_t* src,
                         const double* const* A, const double* const* B,
                         int layout, int precision);

/* Hv = H_sb . v for host arrays of length n = sum Dl*Dr (overwrites hv). */
int nss_sb_hv(nss_ctx* ctx, const double* v, double* hv, int64_t n);
/* Same on device pointers, enqueued on the context stream (no host sync). */
int nss_sb_hv_device(nss_ctx* ctx, const double* d_v, double* d_hv, int64_t n);
/* Largest term-product scratch (doubles, 128 MiB) of nss_sb_hv_device_noalloc. */
#define NSS_NOALLOC_W_MAX (16LL * 1024 * 1024)
/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops).  The result of
 * nss_sb_hv_device, bit for bit, with no hipMalloc/hipFree, no pool
 * allocation and no timing events: the term products W_t go to a file-scope
 * device array of NSS_NOALLOC_W_MAX doubles that each device holds until the
 * process ends.  A superblock whose term products need more is refused
 * (NSS_E_ARG).  Every call on one device shares that array, so calls that
 * overlap on one device (several contexts or streams) race. */
int nss_sb_hv_device_noalloc(nss_ctx* ctx, const double* d_v, double* d_hv, int64_t n);
/* H*v timing recorded with HIP events on the context stream, on every k-th
 * H*v of nss_sb_hv*, nss_lanczos and the engine (k from the environment
 * variable NSS_HV_TIMING when the context is created: default 4, 1 times
 * every call, 0 none).  out[12]: for FP64 (out[0..5]) then FP32
 * (out[6..11]): {timed calls, seconds in their W launches (Psi.B^T per
 * term), seconds in their Y launches (per-sector sums), algorithmic FLOPs of
 * their W launches, FLOPs of their Y launches, all calls}; reset clears the
 * counters. */
int nss_ctx_hv_stats(nss_ctx* ctx, int reset, double* out);
/* Phase timers collected when the environment sets NSS_PROFILE=1 (each
 * phase ends with a stream synchronisation): a JSON object of seconds per
 * phase name written to buf (truncated to len). */
int nss_ctx_timers(nss_ctx* ctx, char* buf, int len);
/* FP32-MFMA H*v (FP32 operands, FP32 accumulation) on host arrays. */
int nss_sb_hv_fp32(nss_ctx* ctx, const double* v, double* hv, int64_t n);
/* Algorithmic FLOPs of one H*v with dense sector blocks (reference order). */
int nss_sb_flops(nss_ctx* ctx, double* flops);

/* Lowest neigen eigenpairs of H_sb (the role of SciFortran sp_eigh at
 * src/DMRG_SUPERBLOCK.f90:412-425): Lanczos with full re-orthogonalisation
 * on the GPU, Krylov dimension <= kmax (restarted from the Ritz vectors),
 * residual ||H x - e x|| <= tol * max(1,|e|).  Start vector: SplitMix64
 * uniform[-1,1) from seed (v0_mode 0), constant 1/sqrt(n) (v0_mode 1,
 * LANC_V0_dble=T), or the first n entries of evecs on input (v0_mode 2, e.g.
 * a predicted wavefunction; a zero vector falls back to v0_mode 0).  precision NSS_FP32_MIXED runs FP32 H*v first and ends
 * with an FP64 restart.  evecs: host, n x neigen, vector-contiguous. */
int nss_lanczos(nss_ctx* ctx, int neigen, int kmax, int maxiter, double tol, uint64_t seed,
                int v0_mode, int precision, double* evals, double* evecs, int* numop);

/* Reduced density matrices of one superblock vector psi (host, length n):
 * side 0 -> rho_L,q = Psi_q Psi_q^T (Dl_q x Dl_q), side 1 -> rho_R,q =
 * Psi_q^T Psi_q (Dr_q x Dr_q), built on the GPU, then diagonalised per block
 * (sb_get_rdm + build_density_matrix + blocks_matrix%eigh,
 * src/DMRG_RDM.f90:18-162, src/MATRIX/MATRIX_BLOCKS.f90:589-657).
 * evals: ascending per block, negative values set to DBL_MIN (tiny(1d0));
 * evecs: per block, row-major n_q x n_q with eigenvectors as columns,
 * blocks concatenated in sector order. */
int nss_rdm_eigh(nss_ctx* ctx, const double* psi, int side, double* evals, double* evecs);

/* Kept-state rule of renormalize_block (src/DMRG_RDM.f90:204-220):
 * evals_desc: all rho eigenvalues sorted descending (length ne);
 * m_s = min(mstates, mdim, ne) (or the estates threshold when mstates==0),
 * extended over every eigenvalue within deg_thr relative of the m_s-th.
 * Returns m_s and the truncation error 1 - sum_{i<m_s} lambda_i. */
int nss_kept_count(const double* evals_desc, int ne, int mstates, double estates, int mdim,
                   double deg_thr, int* m_s, double* trunc_err);

/* Batched block rotation out_b = U_row_b^T . O_b . U_col_b
 * (operators_renormalization, src/DMRG_RDM.f90:355-404, per QN block).
 * O_b: rows_b x cols_b, U_row_b: rows_b x mr_b, U_col_b: cols_b x mc_b,
 * out_b: mr_b x mc_b; row-major host arrays. */
int nss_rotate_blocks(nss_ctx* ctx, int nblk, const int32_t* rows, const int32_t* cols,
                      const int32_t* mr, const int32_t* mc, const double* const* O,
                      const double* const* Urow, const double* const* Ucol, double* const* out);

/* Host-only checks of the small symmetric eigen-solvers (no device needed):
 * a (n x n row-major symmetric) -> eigenvectors as columns, w ascending;
 * tridiagonal (d[n], e[n-1]) -> w ascending, z (n x n row-major) columns. */
int nss_host_sym_eig(int n, double* a, double* w);
int nss_host_tridiag_eig(int n, const double* d, const double* e, double* w, double* z);
/* Lowest eigenvalue of the tridiagonal (d[n], e[n-1]) and the last
 * component of its unit eigenvector (sign arbitrary): the residual estimate
 * |beta * last component| of the Lanczos convergence test of nss_lanczos. */
int nss_host_tridiag_lowest(int n, const double* d, const double* e, double* theta, double* zlast);

/* Lane-layout check of the MFMA builtins (prec 0: f64 16x16x4, 1: f32
 * 16x16x4): one wave, lane l supplies a[l] (A operand) and b[l] (B operand);
 * d[4*l + r] receives accumulator item r of lane l.  a, b: 64, d: 256. */
int nss_probe_mfma(int prec, const double* a, const double* b, double* d);

/* The H*v tiling modes the autotuner has measured in this process, one row
 * of five ints per shape class: {phase (1 W, 2 Y), operand bytes (8 or 4),
 * ceil(4 log2) of its 32x32 tile count and of its largest output edge, mode (1:
 * 32x32 tiles, 2: 64x64 tiles, 3: large-output path of 128x128 tiles)}.  Up to maxrows rows go to rows (may be NULL); *count gets the
 * number of classes.  No counterpart in the reference: it reports how this
 * library tiles its grouped H*v launches (NSS_GEMM_AUTOTUNE=0 disables the
 * measurement, NSS_GEMM_TILE forces a tile). */
int nss_gemm_tune_table(int* rows, int maxrows, int* count);

/* Plain device copy of n doubles (n even), timed over reps launches on the
 * null stream: *seconds = average launch time.  A probe with no counterpart
 * in the reference: it calibrates HBM byte counters and measures the copy
 * bandwidth that bounds the streaming Lanczos kernels. */
int nss_probe_copy(int64_t n, int reps, double* seconds);

/* ------------------------------------------------------------------------
 * Device-resident DMRG engine: init_dmrg / run_DMRG / finalize_dmrg
 * (src/DMRG_MAIN.f90:22-241) with every block operator kept in HBM as
 * QN-blocked dense matrices across steps and sweeps.
 * ---------------------------------------------------------------------- */
enum { NSS_MODEL_SPIN12 = 0, NSS_MODEL_SPIN1 = 1, NSS_MODEL_HUBBARD = 2 };

typedef struct {
  int32_t model;        /* NSS_MODEL_* (spin_site(2|3), electron_site with Norb=1) */
  int32_t hfmode;       /* HFMODE: Hartree shift -U/2 (n_up+n_dw) */
  double hij[4];        /* Hij 2x2 row-major: spin diag(Jz, Jxy/2), fermion diag(t_up, t_dw) */
  double uloc;          /* ULOC */
  double hvec[2];       /* spin local field (Sx, Sz) = HVEC */
  double eloc[2];       /* fermion on-site energies (MH) per spin */
} nss_model;

typedef struct {
  int32_t ldmrg;              /* LDMRG */
  int32_t mdmrg;              /* MDMRG */
  double edmrg;               /* EDMRG (used when mdmrg == 0) */
  int32_t nsweep;             /* NSWEEP */
  const int32_t* msweep;      /* MSWEEP[nsweep] */
  const double* esweep;       /* ESWEEP[nsweep] (may be NULL) */
  double target_qn[2];        /* DMRG_QN, QNtype=local */
  int32_t neigen;             /* LANC_NEIGEN */
  int32_t lanc_dim_threshold; /* LANC_DIM_THRESHOLD */
  int32_t lanc_kmax;          /* Krylov dimension before restart */
  int32_t lanc_maxiter;       /* LANC_NITER (total H*v cap per diag) */
  double lanc_tol;            /* LANC_TOLERANCE */
  uint64_t seed;              /* start-vector seed */
  int32_t v0_mode;            /* 0 random, 1 constant (LANC_V0_dble) */
  double deg_thr;             /* DEG_EVALS_THRESHOLD */
  int32_t precision;          /* NSS_FP64 | NSS_FP32_MIXED */
  int32_t verbose;
  /* 1: with neigen == 1, a finite step that follows a step in the same sweep
   * direction starts its Lanczos from the previous ground state carried into
   * the new superblock basis (wavefunction transformation through the
   * kept-state matrices of both blocks); 0, or neigen > 1: every step starts
   * from v0_mode, as the reference does (src/DMRG_SUPERBLOCK.f90:412-425).
   * Converged energies, entropies and kept counts agree to the Lanczos
   * tolerance either way. */
  int32_t predict;
  /* PBCdmrg: periodic boundary conditions (src/INPUT_VARS.f90:171-183).  An
   * even ldmrg is shifted to the next odd value; every infinite iteration is
   * a link-'n' step followed by a link-'p' step that adds the sites at the
   * blocks' outer ends (src/DMRG_MAIN.f90:155-158), and the superblock
   * carries the connecting terms of both links.  The finite algorithm and
   * nss_dmrg_measure refuse PBC runs, as the reference's finite_DMRG does
   * (src/DMRG_MAIN.f90:179-180). */
  int32_t pbc;
  /* QNtype (src/INPUT_VARS.f90:223-226, sb_set_current_qn
   * src/DMRG_SUPERBLOCK.f90:717-726): 0 local, the target at L sites is
   * int(target_qn L) (Norb = 1); 1 global, it is min(L, int(target_qn)). */
  int32_t qn_type;
  /* QNtruncation_error / QNtruncation_dim (src/INPUT_VARS.f90:211-217,
   * eigh_blocks_matrix src/MATRIX/MATRIX_BLOCKS.f90:627-629): every rho
   * eigenvalue of a QN block whose weight sum is below qn_trunc_error, or
   * whose dimension is below qn_trunc_dim, becomes tiny(1d0) (so its states
   * are kept only after all others).  0 and 0 (or 1) leave rho untouched. */
  double qn_trunc_error;
  int32_t qn_trunc_dim;
} nss_params;

typedef struct {
  int32_t type;               /* 'i' or 'f' */
  int32_t label, sweep;       /* finite: 1 left->right, 2 right->left; sweep index */
  int32_t left_len, right_len, L;
  int32_t neigen;
  double energies[4];         /* gs_energy(1:neigen) */
  double entropy_left, entropy_right;
  int32_t m_left, m_right;    /* renormalized dims (or enlarged dims if not renormalized) */
  int32_t renormalized;
  double trunc_left, trunc_right;
  double cut_left, cut_right; /* rho eigenvalue of the last kept state (lambda_{m_s}) */
  int64_t m_sb;               /* superblock dimension */
  int32_t nsect, nterms;
  int32_t numop;              /* H*v calls (FP64 + FP32) */
  int32_t numop_fp32;
  double hv_flops;            /* algorithmic FLOPs of one H*v */
  double t_step, t_diag, t_hv, t_rdm, t_renorm, t_enlarge;  /* seconds */
  int32_t link;               /* 'n', or 'p' for the PBC_prev step of a PBC iteration */
  /* write_entanglement (src/DMRG_MAIN.f90:501-562): with P_q the weight of
   * rho's QN block q and S_q the entropy of its normalised spectrum, the
   * number entropy S_N = -sum_q P_q ln P_q and the configurational entropy
   * S_c = sum_q P_q S_q, for rho_L and rho_R (S = S_N + S_c). */
  double sn_left, sc_left, sn_right, sc_right;
} nss_step_info;

int nss_dmrg_create(nss_ctx* ctx, const nss_model* model, const nss_params* params,
                    nss_dmrg** out);
/* type 'i': infinite_DMRG; 'f': finite_DMRG (iDMRG to ldmrg, then sweeps). */
int nss_dmrg_run(nss_dmrg* d, int type);
/* finite_DMRG in parts: the iDMRG growth that fills blocks_list
 * (src/DMRG_MAIN.f90:196-203), then sweep im (0-based, Mstates =
 * msweep[im], :214-239) one call at a time. */
int nss_dmrg_finite_begin(nss_dmrg* d);
int nss_dmrg_sweep(nss_dmrg* d, int im);
/* One step of the finite sweeps after nss_dmrg_finite_begin: the next
 * step_dmrg of the three loops of sweep after sweep (src/DMRG_MAIN.f90:
 * 214-239), returning on the device-synchronised end of that step; *done = 1
 * (and no step run) once every MSWEEP sweep is complete.  Sweeps and single
 * steps may be mixed as long as nss_dmrg_sweep starts at a sweep boundary. */
int nss_dmrg_finite_step(nss_dmrg* d, int* done);
int nss_dmrg_nsteps(nss_dmrg* d);
/* Independent chains solved together (no counterpart in the reference,
 * whose runs are one chain per process): the n engines share one context
 * and LDMRG (and NSWEEP for 'f'), and step in lockstep; each step's ground
 * states come from one batched Lanczos whose H*v over all chains is one
 * grouped W launch and one grouped Y launch.  Per engine the results are
 * those of nss_dmrg_run on it alone (same tolerances; the Lanczos iterates
 * differ).  Steps on the dense path, with LANC_NEIGEN > 1 or in mixed
 * precision keep the engine's own solver inside the lockstep.  The step
 * records (nss_dmrg_step_info) stay per engine; the batch's H*v time is
 * booked on the first engine's steps. */
int nss_dmrg_batch_run(int n, nss_dmrg* const* engines, int type);
int nss_dmrg_batch_finite_begin(int n, nss_dmrg* const* engines);
int nss_dmrg_batch_sweep(int n, nss_dmrg* const* engines, int im);
int nss_dmrg_step_info(nss_dmrg* d, int i, nss_step_info* out);
/* Kept-state counts per QN of step i (side 0 left, 1 right), counting only
 * kept states whose rho eigenvalue exceeds min_weight (0: all kept states;
 * states at rounding level belong to rho's null space, whose basis is
 * arbitrary).  Writes up to max entries of (qn[2] as doubles, count);
 * *nout = number of kept sectors. */
int nss_dmrg_kept_qn(nss_dmrg* d, int i, int side, double min_weight, int max, double* qn,
                     int32_t* count, int32_t* nout);
/* rho%lambda() of step i (get_evals_blocks_matrix,
 * src/MATRIX/MATRIX_BLOCKS.f90:660-687, written as LambdaQ_left/right by
 * write_entanglement, src/DMRG_MAIN.f90:523-561): the eigenvalues of every
 * QN block of rho_L (side 0) or rho_R (side 1), ascending within a block,
 * negative ones as tiny(1d0).  Blocks in the superblock's sector order.
 * Writes up to maxb blocks (qn[2] as doubles, count) and up to maxv values
 * back to back; *nblocks and *nvals = the totals.  NULL buffers with 0
 * capacities query the sizes. */
int nss_dmrg_lambda(nss_dmrg* d, int i, int side, int maxb, double* qn, int32_t* count, int64_t maxv,
                    double* vals, int32_t* nblocks, int64_t* nvals);
/* Superblock of the most recent step (for checks and CPU baselines): sector
 * dims and dense host copies of its operator blocks in the layout of
 * nss_sb_upload_blocks (row-major).  Call with NULL buffers to get sizes. */
int nss_dmrg_last_sb_dims(nss_dmrg* d, int* nq, int* nterms, int64_t* n);
int nss_dmrg_last_sb_export(nss_dmrg* d, int32_t* Dl, int32_t* Dr, int32_t* dst, int32_t* src,
                            double* HL, double* HR, double* A, double* B, double* psi);
/* Measure_DMRG (src/DMRG_MEASURE.f90:148-287): <psi| O(pos) |psi> on the
 * ground state of the last step, for nops site operators (each dim x dim,
 * row-major, in the model's site basis: spin-1/2 |up>,|dn>; spin-1
 * |+1>,|0>,|-1>; electron |0>,|up>,|dw>,|up dw>) at npos chain positions
 * (1-based, 1..L_left+L_right of the last superblock).  Each operator is
 * carried through the stored blocks' kept-state matrices on the device
 * (Build_Op_dmrg / Advance_Op_dmrg / Average_Op_dmrg).  Operators must
 * conserve the quantum numbers.  out[iop*npos + ipos]. */
int nss_dmrg_measure(nss_dmrg* d, int nops, const double* ops, int npos, const int32_t* pos, double* out);
/* Block restart files in the reference's text format (write_block /
 * read_block, src/BLOCKS/BLOCKS.f90:418-537).  load: sets the left and right
 * blocks an infinite run continues from (infinite_DMRG loading
 * block_left/right.restart, src/DMRG_MAIN.f90:151-152); operator keys of
 * both generations are accepted (S_z or S_z_n, C_l1_s1 or C_l1_s1_n, ...).
 * save: writes the two blocks that entered the last step (SAVE_BLOCK at
 * length Ldmrg-1, src/DMRG_MAIN.f90:308-317) with the current keys.  Both
 * carry the rotation matrices of every length (omatrices "1".."length",
 * each over the previous length's basis in the order its file lists it);
 * with all of them a loaded block keeps its history and a continued run can
 * be measured, as in the reference.  A file with only "1", or with matrices
 * that do not fit the block, still continues a run (without measurement). */
int nss_dmrg_load_blocks(nss_dmrg* d, const char* left_path, const char* right_path);
int nss_dmrg_save_blocks(nss_dmrg* d, const char* left_path, const char* right_path);
int nss_dmrg_destroy(nss_dmrg* d);

#ifdef __cplusplus
}
#endif
#endif /* NSSDMRG_H */
