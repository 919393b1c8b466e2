/*
 * pot3d_hip.h -- C-ABI of libpot3d_hip.so, the gfx950 (MI355X) solver for
 * POT3D's preconditioned conjugate-gradient solve of the finite-difference
 * spherical Laplace equation.
 *
 * Three groups of entry points:
 *
 *  1. The legacy external-LU-solver ABI, symbol for symbol the one that
 *     POT3D binds under -DHIPSPARSE (src/pot3d.F90:581-609, implemented by
 *     src/lusol_hipsparse.c).  An unmodified pot3d.F90 built with
 *     -DHIPSPARSE links against this library instead of the hipSPARSE shim
 *     and runs ifprec=2 through the D-ILU factorisation and trisolves here.
 *
 *  2. A whole-solve ABI that replaces the do-concurrent solver loops of
 *     potfld/solve/cgsolve (src/pot3d.F90:3866-4485, 5102-5844): the
 *     operator, boundary values, phi seam, dots, preconditioners and the
 *     CG iteration all run on the GPU.  The driver keeps the mesh, Br input,
 *     I/O and diagnostics and passes host arrays.
 *
 *  3. A learned warm start (p3d_nn_*) that proposes the initial guess of
 *     group 2's solve; it has no counterpart in the reference.
 *
 * Conventions (group 2): every function returns 0 (P3D_OK) or an error
 * code; p3d_last_error() describes the last failure on the calling thread.
 * Arrays are host pointers unless the name ends in _dev.  1-D mesh arrays
 * are the reference's 1-based Fortran arrays passed as plain pointers
 * (element (1) first).  Packed solver vectors have N=(nr-2)(nt-2)(np-2)
 * entries, r fastest (the layout of x_cg/rhs_cg, src/pot3d.F90:3917-3924);
 * padded arrays are (nr,nt,np) in Fortran order; br0/br1 are (nt,np).
 * All calls are synchronous at return.  A solver handle must not be used
 * by two threads at once.
 */
#ifndef POT3D_HIP_H
#define POT3D_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------ */
/* 1. Legacy external LU solver ABI (drop-in for src/lusol_hipsparse.c) */

/* Replaces load_external_lu_solver (interface src/pot3d.F90:586-592,
 * implementation src/lusol_hipsparse.c:70-300).  CSR of the reference's
 * ILU matrix (diacsr, src/pot3d.F90:4870-5018): 1-based, FP32 values,
 * int32 indices, N rows, M = nnz+1 (the reference's allocated length).
 * Pointers may be device addresses (the -DCUSPARSE calling convention)
 * or host addresses (the -DHIPSPARSE convention; the arrays are read from
 * host memory).  The 7-point + phi-wrap structure is recognised from IA/JA
 * and factorised as D-ILU with FP64 pivots; the LU factors are written
 * back over csr_lu in FP32, as the reference's in-place csrilu02 does.
 * On an unsupported pattern or any error: message on stderr, exit(1)
 * (the reference shim's behaviour). */
void load_external_lu_solver(float* csr_lu, int* csr_i, int* csr_j, int N, int M);

/* Replaces external_lu_solver (src/pot3d.F90:594-599,
 * src/lusol_hipsparse.c:334-363): x <- (LU)^-1 x in place, FP32 inside,
 * x a device (or host) address of N doubles; synchronous on return. */
void external_lu_solver(double* x);

/* Replaces unload_external_lu_solver (src/pot3d.F90:601-603,
 * src/lusol_hipsparse.c:369-384): frees all state of the legacy solver. */
void unload_external_lu_solver(void);

/* ------------------------------------------------------------------ */
/* 2. Whole-solve ABI                                                   */

typedef struct p3d_solver p3d_solver;

enum {
  P3D_OK = 0,
  P3D_EINVAL = 1,     /* bad argument or shape */
  P3D_ENOMEM = 2,     /* device allocation failed */
  P3D_EHIP = 3,       /* HIP runtime error */
  P3D_ENOCONV = 4,    /* ncgmax reached (the reference's ierr=1) */
  P3D_ENONFINITE = 5, /* NaN/Inf in the residual (e.g. non-finite br0), or
                         a zero / non-finite ILU0 pivot (p3d_create,
                         block options) */
  P3D_ESTATE = 6      /* call out of order (e.g. solve before set_br0) */
};

/* Preconditioner choice: the reference's ifprec (src/pot3d.F90:428). */
enum { P3D_PREC_NONE = 0, P3D_PREC_DIAG = 1, P3D_PREC_ILU0 = 2 };

/* Options for p3d_set_option. */
enum {
  P3D_OPT_STOP = 1,       /* 0: epsn = sqrt(r.M^-1r / b.M^-1b) < epscg (the
                             reference's ernorm, default); 1: ||r||2/||b||2 */
  P3D_OPT_MIXED = 2,      /* FP32 inner PCG inside an FP64 defect-
                             correction loop to ||b-Ax||2/||b||2 <= epscg:
                             1 each outer step restarts the inner PCG on
                             the new residual (inner tolerance
                             P3D_OPT_INNER_TOL); 2 reliable updates: the
                             FP32 recurrence keeps its search direction and
                             its residual is replaced by the FP64 b-Ax
                             whenever it has fallen 1000-fold */
  P3D_OPT_INNER_TOL = 3,  /* inner relative tolerance (default 1e-5) */
  P3D_OPT_CHUNK = 4,      /* CG iterations per captured graph (default auto) */
  P3D_OPT_GRAPH = 5,      /* 1: replay iterations as a hipGraph (default) */
  P3D_OPT_SWEEP = 6,      /* ILU triangular-solve schedule: 0 tile diagonals
                             (8x8x32 tiles, one launch per anti-diagonal;
                             default), 1 one launch per i+j+k hyperplane */
  P3D_OPT_VERBOSE = 7,    /* 1: print the reference's residual lines */
  P3D_OPT_GUESS_CHECK = 8, /* before solving, replace an initial guess that
                             is worse than zero by zero (for guesses from
                             the warm-start network): 1 when ||b-Ax0||2 >=
                             ||b||2; 2 when its A-norm error is not smaller,
                             i.e. x0.Ax0 - 2 b.x0 >= 0 */
  P3D_OPT_TILE_K = 9,     /* planes per tile of the tile-diagonal ILU
                             schedule: 16, 32 (default) or 64 */
  P3D_OPT_TILE_WAVES = 10, /* waves per tile of that schedule: 1 (8x8
                             columns, default), 4 (16x16) or 8 (16x32, at
                             most 32 planes) */
  P3D_OPT_ILU_BLOCKS_R = 11, /* block-Jacobi ILU0: blocks along r, theta, */
  P3D_OPT_ILU_BLOCKS_T = 12, /* phi (default 1 each: the single-rank ILU0). */
  P3D_OPT_ILU_BLOCKS_P = 13, /* Couplings between blocks are dropped, as in
                                the reference's ILU0 on a grid decomposed
                                over that many ranks (src/pot3d.F90:4940-
                                4967); block extents are rounded up to 16,
                                32 and 64 points. */
  P3D_OPT_TILE_STAGE = 14,  /* 8x8-column tiles: 0 planes/8 waves load each
                                tile and one marches (default); 1 the
                                marching wave loads its tile alone (one wave
                                per workgroup, more tiles resident per CU);
                                2 the marching wave holds its tile in
                                registers (16 or 32 planes, one wave per
                                tile).  Same results bit for bit. */
  P3D_OPT_STENCIL = 15,     /* operator kernel: 0 one thread per (i,j)
                                column marching phi; 1 each thread holds 8
                                theta rows of one column, the neighbours in
                                i from the adjacent lanes; 2 64 x 4WJ (r,
                                theta) patches staged plane by plane in LDS,
                                WJ = 8 when that still makes 2048 or more
                                workgroups, else 4; 3 and 4 the same with
                                WJ = 8 and WJ = 4; -1 (default) 2.  The LDS
                                kernels apply A x; b - A x and outputs of
                                2^32 bytes or more take 1 (FP32 operator
                                whose 62-lane waves cover the r extent to
                                90% or more) or 0.  Same operator values
                                bit for bit. */
  P3D_OPT_APPLY_FP32 = 16,  /* p3d_apply_A: 1 applies the FP32 operator of
                                the mixed-precision recurrence (x rounded to
                                float, the result in float); 0 FP64 */
  P3D_OPT_FUSE_RESIDUAL = 17, /* FP32 recurrence with the register-tile ILU
                                sweeps (P3D_OPT_TILE_STAGE 2): 1 folds the
                                residual update r -= alpha Ap into the
                                forward sweep, 0 runs it on its own; -1
                                (default) 1 when N >= 2^26 unknowns (the
                                sweeps are then bandwidth-bound and the
                                fold saves a pass; below, they are
                                latency-bound and it costs).  Same values
                                bit for bit. */
  P3D_OPT_BLOCK_SWEEP = 18   /* block-Jacobi ILU on register tiles
                                (P3D_OPT_TILE_STAGE 2 with a cut): 1 runs
                                each sweep as one launch, a workgroup per
                                block marching the block's tile diagonals;
                                2 so only the small diagonals at either end
                                (fewer tiles than the device holds at
                                once), one launch per other diagonal; 0 one
                                launch per diagonal over all blocks; -1
                                (default) 1 below 2^26 unknowns, else 0.
                                Same values bit for bit. */
};

/* Creates a solver for one rank's grid (nr,nt,np include the boundary /
 * ghost points, as in the reference's local_dims).  Mesh arrays, 1-based
 * Fortran arrays of src/pot3d.F90:2730-2861:
 *   r, dr      (nr-1)   main r mesh and its spacing
 *   rh, drh    (nr)     half r mesh and its spacing
 *   st, dt     (nt-1)   sin(t) of the main theta mesh, spacing
 *   sth, dth   (nt)     sin(th) of the half theta mesh, spacing
 *   dp, dph    (np)     phi spacings of the main and half meshes
 * pm_r1: +1 for option 'potential', -1 for 'open'/'ss' (:1250-1285).
 * ifprec: P3D_PREC_*.  The operator is matrix-free: no coefficient arrays
 * are stored (the reference stores 7 FP64 planes, :4506). */
int p3d_create(p3d_solver** out, int nr, int nt, int np, const double* r, const double* dr,
               const double* rh, const double* drh, const double* st, const double* dt,
               const double* sth, const double* dth, const double* dp, const double* dph,
               double pm_r1, int ifprec);
int p3d_destroy(p3d_solver* s);
const char* p3d_last_error(void);
int p3d_set_option(p3d_solver* s, int key, double value);

/* Boundary data: br0(nt,np) at r0 (module fields, src/pot3d.F90:399);
 * optional br1(nt,np) selects the validation-run Neumann value at r1
 * (:5420-5423).  Pass br1=NULL to clear it. */
int p3d_set_br0(p3d_solver* s, const double* br0);
int p3d_set_br1(p3d_solver* s, const double* br1);

/* rhs(N) of the boundary trick (src/pot3d.F90:3941-3950). */
int p3d_rhs(p3d_solver* s, double* rhs);

/* The reference's solve(x,rhs,N,ierr) (src/pot3d.F90:4169-4229 ->
 * cgsolve :4231-4369): x(N) initial guess in, solution out; rhs(N) in,
 * residual b-Ax out.  rhs may be NULL: the boundary-trick rhs of the
 * current br0/br1 is used (and the residual is not returned).  ncghist>0
 * prints the reference's residual lines every ncghist iterations when
 * P3D_OPT_VERBOSE is set; the history is also kept for p3d_history.
 * Returns P3D_OK, P3D_ENOCONV (ierr=1), P3D_ENONFINITE, ...; *ncg and
 * *epsn as the reference's cgcom variables. */
int p3d_solve(p3d_solver* s, double* x, double* rhs, double epscg, int ncgmax, int ncghist,
              int* ncg, double* epsn);
/* Same with device pointers (N doubles each; rhs_dev NULL as above). */
int p3d_solve_dev(p3d_solver* s, double* x_dev, double* rhs_dev, double epscg, int ncgmax,
                  int* ncg, double* epsn);

/* After a solve: phi(nr,nt,np) = unpack(x) + boundary values (vmask=1) +
 * phi seam (src/pot3d.F90:3966-3969).  x NULL uses the last solution. */
int p3d_fill_phi(p3d_solver* s, const double* x, double* phi);

/* getb + magnetic_energy (src/pot3d.F90:6187-6321) of phi (host, NULL =
 * the phi of the last p3d_fill_phi): w[0..2] = energy in Br^2, Bt^2, Bp^2.
 * br/bt/bp may be NULL; else receive (nr-1,nt,np), (nr,nt-1,np),
 * (nr,nt,np-1). */
int p3d_energy(p3d_solver* s, const double* phi, double* w, double* br, double* bt, double* bp);

/* The arithmetic of the driver's readbr + set_flux on the GPU: replaces
 * intrp2d/interp (src/pot3d.F90:6830-6910, 7028-7072), readbr's seam, pole
 * and flux-balance steps (:6751-6829) and set_flux's |Br| (:3806-3833); the
 * driver keeps rdhdf_2d and passes the map.  Interpolates bn(ntn,npn)
 * (Fortran order, on the increasing scales tn(ntn), pn(npn)) bilinearly onto
 * the half-mesh interior th(2:nt-1) x ph(2:np-1), sets br0(:,1) =
 * br0(:,np-1) and br0(:,np) = br0(:,2), the m=0 pole ghosts br0(1,:) and
 * br0(nt,:), and subtracts br00err = (fluxp+fluxm)/area from every point
 * when balance_flux != 0 (the reference skips it for 'ss'/'open' with
 * do_not_balance_flux); abs_br != 0 then takes |br0| (option 'open').
 * Mesh arrays: th, sth, dth (nt) and ph, dph (np) of the global mesh
 * (src/pot3d.F90:2583-2706); pl = phi period (p1-p0).  br0(nt,np) receives
 * the result; flux (may be NULL) receives {fluxp, fluxm, area, br00err}.
 * P3D_EINVAL for scales that do not increase (intrp2d's ierr=1) or a mesh
 * point outside them (ierr=2).  Host arrays; synchronous. */
int p3d_readbr(int nt, int np, const double* th, const double* ph, const double* sth,
               const double* dth, const double* dph, double pl, int ntn, int npn,
               const double* tn, const double* pn, const double* bn, int balance_flux,
               int abs_br, double* br0, double* flux);

/* Single operations (host vectors of N), for parity checks:
 * y = A x (ax, :5102-5150); x <- M^-1 x (prec_inv, :5152-5225). */
int p3d_apply_A(p3d_solver* s, const double* x, double* y);
int p3d_apply_prec(p3d_solver* s, double* x);
/* ILU pivot inverses real(1/d, REAL32) of the D-ILU factorisation (N). */
int p3d_ilu_dinv(p3d_solver* s, float* dinv);

/* Statistics of the last solve. */
typedef struct {
  int ncg;              /* CG iterations (sum of inner iterations if mixed) */
  int n_outer;          /* defect-correction sweeps (mixed), else 0 */
  int status;           /* 1 converged, 2 ncgmax, 3 non-finite, 4 zero rhs */
  int guess;            /* P3D_OPT_GUESS_CHECK: 1 guess kept, 2 replaced by
                           zero, 0 not checked */
  double epsn;          /* final epsn (reference definition) */
  double relres;        /* final ||r||2/||b||2 of the recurrence residual */
  double true_relres;   /* ||b - A x||2/||b||2 recomputed in FP64 */
  double t_solve;       /* seconds, p3d_solve entry to return */
  double t_setup;       /* seconds spent in preconditioner set-up */
} p3d_stats;
int p3d_get_stats(p3d_solver* s, p3d_stats* st);
/* epsn per iteration 0..n-1 of the last (inner) solve; returns count. */
int p3d_history(p3d_solver* s, double* hist, int n);

/* Kernel timing with HIP events on the solver's stream: `which` =
 * 0 FP64 operator, 1 FP32 operator (output placed as in the FP32
 * recurrence), 2 ILU apply (FP64 in/out), 3 one FP64 PCG iteration
 * (replayed through the solve's captured graph), 4 FP64 CG vector update,
 * 5 FP64 p update, 6 plain copy of N doubles (bandwidth reference), 7 ILU
 * apply in FP32; with the single-wave ILU tile sweeps (tile-major vectors,
 * else P3D_ESTATE) the FP32 recurrence's 8 p update with the FP64 solution
 * update, 9 residual update, 10 r.z with r.r, and with the register tiles
 * (P3D_OPT_TILE_STAGE 2) 11 the FP32 ILU apply with the residual update
 * folded into its forward sweep (as the recurrence runs it).  The work
 * vectors' contents are left undefined (a later solve sets them up).
 * Average milliseconds per launch over reps launches. */
int p3d_time_kernel(p3d_solver* s, int which, int reps, double* avg_ms);

/* ---- Learned warm start ----------------------------------------------------
 * No counterpart in the reference, whose initial guess is always zero
 * (x_cg(:)=0, src/pot3d.F90:3922-3925; cgsolve's x0 at :4244-4248).  A
 * 4-level 3-D convolutional encoder-decoder (3x3x3 kernels, 16-128
 * channels, BF16 weights and activations, FP32 accumulation on the matrix
 * cores) maps br0 and the radial grid to a potential on the solver's
 * interior points, in the packed x_cg layout.  It is trained on the device
 * (hand-written forward, backward and Adam) on (br0, x) pairs that
 * p3d_solve produced, and its prediction is only ever an initial guess for
 * p3d_solve, which still runs to the same epscg.
 * Errors are described by p3d_last_error. */
typedef struct p3d_nn p3d_nn;
/* Network for the interior grid of (nr, nt, np); rh(nr) is the half r mesh
 * (for the radial input channels); seed initialises the weights. */
int p3d_nn_create(p3d_nn** out, int nr, int nt, int np, const double* rh, unsigned seed);
int p3d_nn_destroy(p3d_nn* n);
size_t p3d_nn_num_params(p3d_nn* n);
/* Input and output scales: features use br0/bscale, targets x/phiscale.
 * bscale = 0 selects per-sample normalisation: each br0 is divided by its
 * own rms s over the interior points and its x by s*phiscale (the solution
 * is linear in br0). */
int p3d_nn_set_scales(p3d_nn* n, double bscale, double phiscale);
/* FP32 master parameters (p3d_nn_num_params floats). */
int p3d_nn_get_params(p3d_nn* n, float* params);
int p3d_nn_set_params(p3d_nn* n, const float* params);
/* Stores a training pair br0(nt,np), x(N) on the device. */
int p3d_nn_add_sample(p3d_nn* n, const double* br0, const double* x);
/* One forward/backward of the mean-square loss on (br0, x): *loss and, if
 * grads != NULL, d loss / d params (p3d_nn_num_params floats). */
int p3d_nn_loss_grad(p3d_nn* n, const double* br0, const double* x, double* loss,
                     float* grads);
/* nsteps Adam steps (lr; betas 0.9, 0.999) over the stored pairs in a
 * seeded shuffled order; losses (nsteps floats, may be NULL) receive each
 * step's loss. */
int p3d_nn_train(p3d_nn* n, int nsteps, double lr, unsigned seed, float* losses);
/* Training augmentation for p3d_nn_train, by exact symmetries of the
 * discrete problem (all off by default): phi_roll rotates each sample by a
 * random number of interior phi points (valid for a uniform phi mesh),
 * theta_flip reflects it in theta at random (a theta mesh symmetric about
 * the equator), sign_flip negates br0 and x at random (the solution is
 * linear in br0).  The caller decides which hold for its mesh. */
int p3d_nn_set_augment(p3d_nn* n, int phi_roll, int theta_flip, int sign_flip);
/* x(N) predicted for br0(nt,np): host x, or device x_dev when non-NULL. */
int p3d_nn_predict(p3d_nn* n, const double* br0, double* x, double* x_dev);
/* Trilinear prolongation of a coarse padded field phic(nrc,ntc,npc) to the
 * af x bf x cf interior points of a finer grid: per fine interior index, the
 * lower coarse padded index (ir, it, ip) and the weight of the upper one
 * (wr, wt, wp).  xf(af*bf*cf) packed as x_cg. */
int p3d_nn_prolong(int nrc, int ntc, int npc, const double* phic, int af, int bf, int cf,
                   const int* ir, const double* wr, const int* it, const double* wt,
                   const int* ip, const double* wp, double* xf);
/* Same with phic_dev and xf_dev device addresses (tables on the host). */
int p3d_nn_prolong_dev(int nrc, int ntc, int npc, const double* phic_dev, int af, int bf, int cf,
                       const int* ir, const double* wr, const int* it, const double* wt,
                       const int* ip, const double* wp, double* xf_dev);
/* No counterpart in the reference.  For callers that cannot allocate device
 * memory (code under HIP graph capture, serving loops, harnesses that
 * forbid hipMalloc): p3d_nn_prolong_dev with the same arguments and results,
 * but the per-axis tables are staged in a fixed device array of the
 * library (96 KiB per GPU, held until the process ends) instead of buffers
 * allocated per call; no hipMalloc or hipFree.  af+bf+cf must be at most
 * 8192 (else P3D_EINVAL).  Calls that overlap on one GPU race on that
 * array. */
int p3d_nn_prolong_dev_noalloc(int nrc, int ntc, int npc, const double* phic_dev, int af, int bf,
                               int cf, const int* ir, const double* wr, const int* it,
                               const double* wt, const int* ip, const double* wp,
                               double* xf_dev);
/* One 16x16x16 BF16 matrix-core product D = A B (row-major BF16 bits in,
 * FP32 out), for checking the operand layout the network kernels use. */
int p3d_nn_mfma_check(const unsigned short* A, const unsigned short* B, float* D);

#ifdef __cplusplus
}
#endif
#endif /* POT3D_HIP_H */
