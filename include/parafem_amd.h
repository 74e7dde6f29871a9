/*
 * parafem_amd -- MI355X (gfx950) element-by-element PCG for ParaFEM's
 * p121 (hex20 elasticity) and p123 (hex8 Laplace) solves.
 *
 * C ABI.  Plain pointers and sizes only.  Every entry returns an int status
 * (PF_OK = 0).  Conventions follow ParaFEM's Fortran arrays:
 *   - integer arrays are int32, equation and node ids are 1-based and 0 in a
 *     steering array means "restrained freedom";
 *   - matrices are column-major: storkm_pp(ntot,ntot,nels) is passed as
 *     nels consecutive ntot x ntot column-major blocks, g_g_pp(ntot,nels) as
 *     nels consecutive ntot-vectors, g_coord_pp(nod,3,nels) as nels
 *     consecutive (nod,3) column-major blocks.
 * Host pointers ("host") are read or written only during the call.  Device
 * pointers ("dev") must be device memory of the context's GPU (from
 * pf_dev_alloc or any HIP allocator); work on them is ordered on the
 * context's stream.  One context per host thread; no global state other than
 * what each context owns.
 *
 * The reference has no such library: its drivers call module procedures of
 * the same names from Fortran (src/modules/readme.txt:26-38 describes the
 * same-name library swap).  Each declaration cites the reference routine it
 * replaces; paths are relative to ParaFEM's source tree.
 */
#ifndef PARAFEM_AMD_H
#define PARAFEM_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------ status */
#define PF_OK              0
#define PF_ERR_ARG         1  /* bad argument (null, size, order of calls)   */
#define PF_ERR_HIP         2  /* HIP runtime error                            */
#define PF_ERR_OOM         3  /* device allocation failed                     */
#define PF_ERR_STATE       4  /* required set-up call missing                 */
#define PF_NOT_CONVERGED   5  /* PCG hit its iteration limit                  */
#define PF_ERR_BREAKDOWN   6  /* non-finite or zero curvature p.u / up        */
#define PF_ERR_NODEVICE    7  /* no HIP device / device index out of range    */
#define PF_ERR_TOO_LARGE   8  /* the problem exceeds the fixed device scratch
                                 of an allocation-free entry point            */

/* Storage precision of element matrices. */
#define PF_STORE_FP64      0  /* storkm_pp as REAL(iwp) (precision.f90:19)    */
#define PF_STORE_FP32      1  /* FP32 hi + FP32 lo split (K ~= hi + lo) of the
                                 upper triangle, packed symmetric: ntot/2 + 1
                                 entries per row, 52% of the full matrix      */

/* PCG stopping rule. */
#define PF_STOP_CHECON     0  /* checon_par: max|dx|/max|x| <= tol
                                 (src/modules/mpi/maths.f90:999-1069)         */
#define PF_STOP_RELRES     1  /* ||r||_2 <= tol * ||r0||_2                    */

/* PCG arithmetic. */
#define PF_PREC_FP64       0  /* FP64 everywhere (the reference's arithmetic) */
#define PF_PREC_MIXED      1  /* element products from the FP32 hi matrices in
                                 FP32; vectors, dots and recurrences in FP64;
                                 reliable updates replace r by the true
                                 residual f - (hi+lo) x computed in FP64.
                                 Needs PF_STORE_FP32 and PF_STOP_RELRES.       */
#define PF_PREC_MIXED64    2  /* as PF_PREC_MIXED, but the element products
                                 widen hi to FP64 and use the FP64 p and FP64
                                 sums (the products see one fixed FP32
                                 rounding of K instead of a new one each
                                 iteration: fewer iterations on problems
                                 whose FP64 solve converges superlinearly).  */
#define PF_PREC_SPLIT      3  /* FP32 hi + lo storage, element products of
                                 (double)hi + (double)lo in FP64 every
                                 iteration: the FP64 operator to ~2^-46, the
                                 FP64 iteration, half the FP64 path's bytes. */

typedef struct pf_ctx pf_ctx;

typedef struct {
    int    limit;        /* iteration limit (p121.f90:103 "iters==limit")     */
    double tol;          /* tolerance for the stopping rule                   */
    int    stop_mode;    /* PF_STOP_*                                          */
    int    precision;    /* PF_PREC_*                                          */
    int    check_every;  /* host polls convergence every n iterations (0=16)  */
    int    refine_every; /* mixed: also replace r at least every n iterations
                            (0 = only on the residual-drop trigger)           */
    double refine_delta; /* mixed: replace r when ||r|| falls below delta
                            times its value at the last replacement
                            (0 = 1e-2)                                         */
} pf_pcg_opts;

typedef struct {
    int    iters;        /* completed iterations, as the reference counts     */
    int    converged;    /* 1 if the stopping rule was met                    */
    double crit;         /* last value of the stopping criterion              */
    double rr0;          /* ||r0||_2^2                                         */
    double rr;           /* ||r||_2^2 of the recurrence residual at exit      */
    int    refinements;  /* mixed: number of true-residual replacements        */
    int    status;       /* PF_OK, PF_NOT_CONVERGED or PF_ERR_BREAKDOWN        */
} pf_pcg_result;

/* ------------------------------------------------------ context, memory */
/* Library version (major*10000 + minor*100 + patch). */
int pf_version(void);
/* Human-readable text for a status code; never NULL. */
const char *pf_status_string(int status);
/* Replaces FIND_PE_PROCS (src/modules/mpi/mp_interface.f90:42-91): one
 * process drives one GPU, so the "process" is a context on a device. */
int pf_create(int device, pf_ctx **ctx);
/* Replaces SHUTDOWN (mp_interface.f90:97-106); frees all device state. */
int pf_destroy(pf_ctx *ctx);
/* Free the current problem's device buffers (mesh, maps, matrices,
 * vectors) and keep the context for another problem (the DEALLOCATE
 * statements at the end of p121.f90:111-123). */
int pf_reset(pf_ctx *ctx);
/* Order the context's work on a caller's hipStream_t (NULL = the context's
 * own stream).  The stream must outlive its use by the context. */
int pf_set_stream(pf_ctx *ctx, void *hip_stream);
/* The hipStream_t the context currently enqueues on. */
void *pf_get_stream(pf_ctx *ctx);
/* Wait for all work the context enqueued. */
int pf_sync(pf_ctx *ctx);
/* Device buffers for callers without an allocator of their own (Fortran). */
int pf_dev_alloc(pf_ctx *ctx, size_t bytes, void **dptr);
int pf_dev_free(pf_ctx *ctx, void *dptr);
int pf_memcpy_h2d(pf_ctx *ctx, void *dst_dev, const void *src_host, size_t bytes);
int pf_memcpy_d2h(pf_ctx *ctx, void *dst_host, const void *src_dev, size_t bytes);
/* Bytes of device memory the context holds (in use plus freed blocks it
 * keeps for reuse). */
int pf_mem_used(pf_ctx *ctx, size_t *bytes);
/* Return the freed blocks the context keeps for reuse to the device. */
int pf_trim(pf_ctx *ctx);
/* Device-to-device copy by a plain streaming kernel (k_copy): moves
 * `bytes` (a multiple of 16) once through HBM in each direction.  Also the
 * calibration kernel for the HBM byte counters.  No reference counterpart. */
int pf_dev_copy(pf_ctx *ctx, void *dst_dev, const void *src_dev, size_t bytes);

/* ------------------------------------------------------------- set-up */
/* Element coordinates g_coord_pp(nod,3,nels) (host), as read by
 * READ_G_COORD_PP (src/modules/mpi/input.f90:288-445).  nod = 8 or 20. */
int pf_mesh_set(pf_ctx *ctx, int nod, int nels, const double *g_coord_pp);
/* Element steering g_g_pp(ntot,nels) (host) and neq, as built by
 * REARRANGE + FIND_G3 / FIND_G4 (src/modules/shared/new_library.f90:
 * 3059-3271; p121.f90:44-48).  Builds the device gather map and the
 * transposed (equation -> element slots, ascending element) scatter map that
 * MAKE_GGL's 1-rank tables stand for (src/modules/mpi/gather_scatter.f90:
 * 1387-1780).  Integer maps are exact; no equation may be missing. */
int pf_steer_set(pf_ctx *ctx, int ntot, int nels, int neq, const int *g_g_pp);
/* REARRANGE + FIND_G3 / FIND_G4 (src/modules/shared/new_library.f90:
 * 3059-3271) and the gather of g_g_pp (p121.f90:44-47) with the element
 * steering built on the device: g_num_pp (nod x nels, node numbers 1..nn)
 * and rest as the driver holds it after READ_REST (nr rows of nodof+1
 * ints, row after row: the node, then per freedom 0 restrained / nonzero
 * free; nodes without a row are free).  Equations are numbered in ascending
 * node, then freedom, order; *neq receives their number and nf (nodof x nn,
 * nullable) the node freedom table.  Then as pf_steer_set. */
int pf_steer_build(pf_ctx *ctx, int nod, int nels, int nodof, const int *g_num_pp, int nn, int nr,
                   const int *rest, int *neq, int *nf);
/* The device's g_g_pp (ntot x nels) back to the host. */
int pf_steer_get(pf_ctx *ctx, int *g_g_pp);
/* Element stiffness of p121 (p121.f90:54-64): SAMPLE, SHAPE_DER, jacobian,
 * DETERMINANT, INVERT, BEEMAT, DEEMAT, K += B^T D B det w over 2x2x2 Gauss
 * points, fused in one kernel that writes storkm_pp to HBM.  e_elem (host,
 * nels, nullable) gives a per-element Young's modulus, else e. */
int pf_assemble_elastic(pf_ctx *ctx, double e, double v, const double *e_elem, int store);
/* Element conductivity of p123 (p123.f90:70-84):
 * storkc = kx*kcx + ky*kcy + kz*kcz; k_elem (host, nullable) scales all three
 * per element. */
int pf_assemble_laplace(pf_ctx *ctx, double kx, double ky, double kz, const double *k_elem, int store);
/* p124's element loop (src/programs/5th_ed/p124/p124.f90:73-86) on 8-node
 * bricks, one freedom per node: conductivity kc (kay = diag(kx,ky,kz)) and
 * consistent capacity pm (SHAPE_FUN, rho*cp) per Gauss point, then
 * storka_pp = pm + kc*theta*dtim, the PCG's matrices (stored as `store`),
 * and storkb_pp = pm - kc*(1-theta)*dtim (FP64, for pf_apply_raw).
 * prop: (5, np_types) column-major kx, ky, kz, rho, cp per type
 * (READ_MATERIAL); etype: nels type numbers 1..np_types. */
int pf_assemble_transient(pf_ctx *ctx, double theta, double dtim, int np_types, const double *prop,
                          const int *etype, int store);
/* Upload / download storkm_pp(ntot,ntot,nels) (host, FP64).  Upload lets a
 * caller keep its own assembly (the driver's loop A) and use the device PCG. */
int pf_storkm_set(pf_ctx *ctx, const double *storkm_pp, int store);
int pf_storkm_get(pf_ctx *ctx, double *storkm_pp);
/* One element matrix km(ntot,ntot) (host, FP64, column-major) used by every
 * element: the identical-K_e solve of PCG_KM (src/modules/mpi/maths.f90:
 * 1152-1323, utemp_pp = MATMUL(km, pmul_pp)).  FP64 only; storkm_get then
 * returns km for every element. */
int pf_storkm_shared(pf_ctx *ctx, const double *km);
/* Jacobi preconditioner given by the caller, already inverted
 * (diag_precon_pp, host, neq), as PCG_KM receives it (maths.f90:1152-1170):
 * the PCG uses it instead of 1/diag until the next assembly or upload. */
int pf_precon_set(pf_ctx *ctx, const double *diag_precon);
/* Assembled diagonal of the element matrices (host, neq): the
 * diag_precon_tmp -> SCATTER step of p121.f90:65-69 before inversion. */
int pf_diag_get(pf_ctx *ctx, double *diag);
/* p123 fixed freedoms by penalty (p123.f90:94-132): adds penalty to the
 * assembled diagonal of equations no_f (host, 1-based) and makes the PCG
 * replace u(no_f(i)) by p(no_f(i))*store(i) after each scatter
 * (p123.f90:141-145), store(i) = diag + penalty.  Call after assembly.
 * store_out (host, nullable, nfix) receives the store values. */
int pf_fixed_set(pf_ctx *ctx, int nfix, const int *no_f, double penalty, double *store_out);

/* ----------------------------------------- drop-in primitives (device) */
/* GATHER(p_pp, pmul_pp) (src/modules/mpi/gather_scatter.f90:547-688):
 * pmul(:,e) = p(g_g(:,e)), 0 for restrained freedoms.  p: dev neq,
 * pmul: dev ntot*nels. */
int pf_gather(pf_ctx *ctx, const double *p_dev, double *pmul_dev);
/* utemp(:,e) = MATMUL(storkm(:,:,e), pmul(:,e)) (p121.f90:93-97). */
int pf_elem_matvec(pf_ctx *ctx, const double *pmul_dev, double *utemp_dev);
/* SCATTER(u_pp, utemp_pp) (gather_scatter.f90:694-850): u(eq) += sum of
 * utemp over the element slots of eq in ascending element order. */
int pf_scatter(pf_ctx *ctx, double *u_dev, const double *utemp_dev);
/* SCATTER_NOADD(utemp_pp, u_pp) (gather_scatter.f90:856-993): assignment. */
int pf_scatter_noadd(pf_ctx *ctx, const double *utemp_dev, double *u_dev);
/* DOT_PRODUCT_P(a, b) (src/modules/mpi/maths.f90:168-216); result to host. */
int pf_dot_product_p(pf_ctx *ctx, const double *a_dev, const double *b_dev, int64_t n, double *out);
/* SUM_P(a) (maths.f90:271-315); result to host. */
int pf_sum_p(pf_ctx *ctx, const double *a_dev, int64_t n, double *out);
/* checon_par(loads, tol, converged, oldlds) (maths.f90:999-1069):
 * converged = max|xnew-x|/max|xnew| <= tol, then x = xnew. */
int pf_checon_par(pf_ctx *ctx, const double *xnew_dev, double tol, int *converged, double *x_dev, int64_t n);
/* The whole-array statements of the PCG loop on device vectors of length n
 * (p121.f90:86-87,91,100-102), each rounded as written:
 *   pf_vec_fill : x = value                    (u_pp = zero)
 *   pf_vec_waxpy: w = x + y*a                   (xnew = x + p*alpha,
 *                 r = r - u*alpha as a = -alpha, p = d + p*beta)
 *   pf_vec_mul  : z = x*y                       (d = diag_precon*r)
 *   pf_vec_recip: y = 1/x                       (diag_precon = 1/diag_precon)
 * w, z, y may alias the inputs. */
int pf_vec_fill(pf_ctx *ctx, double *x_dev, double value, int64_t n);
int pf_vec_waxpy(pf_ctx *ctx, double *w_dev, const double *x_dev, const double *y_dev, double a, int64_t n);
int pf_vec_mul(pf_ctx *ctx, double *z_dev, const double *x_dev, const double *y_dev, int64_t n);
int pf_vec_recip(pf_ctx *ctx, double *y_dev, const double *x_dev, int64_t n);

/* ----------------------------------------------------- the PCG solve */
/* The PCG loop of p121.f90:86-104 / p123.f90:133-151 on device-resident
 * state: d = dinv*r, p = d, x = 0, then per iteration gather -> element
 * products -> scatter, alpha, x/r/d updates, beta, p update and the stopping
 * rule.  r0 is the loaded right-hand side (host neq).  x receives xnew (host
 * neq).  Returns PF_NOT_CONVERGED if the limit was reached. */
int pf_pcg_solve(pf_ctx *ctx, const double *r0_host, double *x_host,
                 const pf_pcg_opts *opts, pf_pcg_result *res);
/* Split form for callers that keep vectors on the device: begin (r0 host or
 * dev), run n more iterations without host synchronisation (iterations after
 * convergence are no-ops on the device), poll (synchronises), read x. */
int pf_pcg_begin(pf_ctx *ctx, const double *r0, int r0_on_device, const pf_pcg_opts *opts);
int pf_pcg_run(pf_ctx *ctx, int n_iters);
int pf_pcg_poll(pf_ctx *ctx, pf_pcg_result *res);
int pf_pcg_x(pf_ctx *ctx, double *x, int x_on_device);
/* The recurrence residual r at the end of the solve (PCG_KM's r_pp is
 * INOUT, maths.f90:1164). */
int pf_pcg_r(pf_ctx *ctx, double *r, int r_on_device);
/* No counterpart in the reference (serving many solves from one host
 * thread).  pf_pcg_run_graph: as pf_pcg_run, but the n iterations are
 * captured once into a HIP graph and replayed (same kernels, same results);
 * the capture is redone when n, the precision or the stopping rule change.
 * pf_pcg_query: non-blocking poll -- the first call after work was enqueued
 * snapshots the PCG state at that point of the stream; *ready becomes 1 (and
 * res is filled) once the snapshot has arrived. */
int pf_pcg_run_graph(pf_ctx *ctx, int n_iters);
int pf_pcg_query(pf_ctx *ctx, pf_pcg_result *res, int *ready);
/* No counterpart in the reference (serving many solves at once).  A batch
 * advances up to PF_BATCH_MAX solves -- contexts each holding a begun PCG
 * (pf_pcg_begin) with the same ntot, precision and stopping rule and the
 * tree dot order -- with one launch per kernel of the iteration for all of
 * them: each solve's blocks run its own launch's work (same local blocks,
 * same reduction grid), so every solve's iterates are bit-identical to
 * pf_pcg_run on its context alone, and a solve that has stopped does
 * nothing.  The batch has its own stream; pf_batch_run orders it after
 * every context's stream and every context's stream after the run, so the
 * context calls that follow (pf_pcg_x, pf_pcg_query, pf_reset) see its
 * results.  The set of contexts may change from call to call (serving:
 * admit and retire solves between runs).
 *   pf_batch_run       : n_iters iterations as direct launches;
 *   pf_batch_run_graph : the same, captured once into a HIP graph for the
 *                        current set and replayed;
 *   pf_batch_query     : non-blocking poll.  The first call after a
 *                        snapshot was delivered takes a new one of the
 *                        solves of the last run (in the order given) at
 *                        that point of the batch stream; runs enqueued
 *                        after it do not change it.  *ready = 1 and res[k]
 *                        filled, one per solve of that run, once it has
 *                        arrived. */
#define PF_BATCH_MAX 64
typedef struct pf_batch pf_batch;
int pf_batch_create(int device, pf_batch **batch);
int pf_batch_destroy(pf_batch *batch);
int pf_batch_run(pf_batch *batch, int nctx, pf_ctx *const *ctxs, int n_iters);
int pf_batch_run_graph(pf_batch *batch, int nctx, pf_ctx *const *ctxs, int n_iters);
/*   pf_batch_run_persistent : the iterations of every solve given, in one
 *                        cooperative launch of resident workgroups that
 *                        runs until every solve has stopped or max_iters
 *                        iterations (grid-wide barriers between the
 *                        iteration's phases, the same work and bits as the
 *                        launches per kernel).  Needs cooperative-launch
 *                        support (PF_ERR_ARG otherwise).  A barrier that
 *                        waits for seconds gives up, and the next
 *                        pf_batch_query returns PF_ERR_HIP. */
int pf_batch_run_persistent(pf_batch *batch, int nctx, pf_ctx *const *ctxs, int max_iters);
int pf_batch_query(pf_batch *batch, pf_pcg_result *res, int *ready);
/*   pf_batch_solo_set : solves of at least min_elems elements (0: none, the
 *                        default) run their own iteration launches -- those
 *                        of pf_pcg_run -- in the batch's stream and graph
 *                        instead of a share of the batched launches (a large
 *                        solve fills the device by itself); the same bits
 *                        either way.  pf_batch_run_persistent ignores it. */
int pf_batch_solo_set(pf_batch *batch, long min_elems);
/* y = K x over the assembled operator in FP64 (host vectors, neq), with the
 * p123 fixed-freedom rows applied.  Used to report true residuals. */
int pf_apply(pf_ctx *ctx, const double *x_host, double *y_host);
/* No counterpart in the reference: pf_apply for callers that cannot
 * allocate device memory (tight serving loops, harnesses that forbid
 * hipMalloc).  Same operator, same bits as pf_apply, but the padded x, y
 * and the element products live in a fixed device scratch array of
 * PF_NOALLOC_SCRATCH_BYTES per GPU (held from library load until the process
 * ends) instead of buffers allocated per call.  Needs
 * 8*(ntot*nels) + 2*8*(neq+1) bytes (each part rounded up to 256) within
 * that size, else PF_ERR_TOO_LARGE.  Every call on one GPU shares that GPU's
 * scratch: calls that overlap on one GPU (two host threads, or two streams
 * with work in flight) race. */
#define PF_NOALLOC_SCRATCH_BYTES (64u << 20)
int pf_apply_noalloc(pf_ctx *ctx, const double *x_host, double *y_host);
/* The element products of p124's time loop without the fixed-freedom rows:
 * y = SCATTER(storka_pp . GATHER(x)) (which = 0, the PCG's matrices, in
 * FP64) or storkb_pp (which = 1) (p124.f90:158-162, 171-174, 191-194).
 * Host vectors, neq. */
int pf_apply_raw(pf_ctx *ctx, int which, const double *x_host, double *y_host);

/* ------------------------------------------------------ p122 plasticity */
/* p122's viscoplastic Mohr-Coulomb stress update and body loads
 * (src/programs/5th_ed/p122/p122.f90:164-199) on 8-node bricks (ntot 24),
 * one element per thread: from the displacement increment x_dev (device,
 * neq; p122's loads_pp after the solve), per Gauss point eps = B eld - evpt,
 * stress = D eps + tensor, INVAR / MOCOUF / MOCOUQ / FORMM
 * (src/modules/shared/new_library.f90:1813-1900, 2364-2460, 85-198);
 * updates evpt_dev and (final_pass != 0, p122's plastic_converged .OR.
 * plasiters == plasits) tensor_dev, both (6, 8, nels) device arrays, and
 * writes each element's body loads into utemp_dev (24, nels) for SCATTER.
 * e, v: DEEMAT; phi, c, psi in degrees / stress units; dt the viscoplastic
 * time step.  Trigonometric functions are the device library's. */
int pf_plastic_update(pf_ctx *ctx, const double *x_dev, double *evpt_dev, double *tensor_dev, double *utemp_dev,
                      double e, double v, double phi, double c, double psi, double dt, int final_pass);
/* p122's warm-started CG (p122.f90:141-162) as one call on device vectors
 * (neq): x = x0 and r = r0, which the caller forms as loads - K x0 with
 * GATHER / element products / SCATTER as p122.f90:141-145 does; d = dinv*r,
 * p = d; then the iterations of pf_pcg_solve (checon_par at opts->tol,
 * opts->limit = cjits) with the freedoms of pf_fixed_set giving
 * u(j) = p(j)*store(j) (fixed_mode 1, p122's first plastic iteration) or
 * u(j) = 0 (fixed_mode 0, the later ones).  xnew receives the last iterate
 * (p122's xnew_pp; checon_par leaves x_pp equal to it).  PF_PREC_FP64 or
 * PF_PREC_SPLIT; the same arithmetic, in the same order, as the statement
 * form on the primitives above. */
int pf_pcg_solve_warm(pf_ctx *ctx, const double *r0_dev, const double *x0_dev, double *xnew_dev, int fixed_mode,
                      const pf_pcg_opts *opts, pf_pcg_result *res);
/* p122's CG statement for the fixed freedoms set by pf_fixed_set
 * (p122.f90:141-146): u(j) = p(j)*store(j) (mode 1, first plastic
 * iteration) or u(j) = 0 (mode 0, later ones).  Device vectors, neq. */
int pf_fixed_rows(pf_ctx *ctx, double *u_dev, const double *p_dev, int mode);

/* ------------------------------------------------------- post-solve */
/* Stresses at the centroid of every element, p121.f90:112-123 (which
 * prints element 1's): eld = GATHER(xnew), SHAPE_DER at points = 0,
 * jac = MATMUL(der, coord), INVERT, deriv, BEEMAT, eps = MATMUL(bee, eld),
 * sigma = MATMUL(dee, eps) with DEEMAT(e_elem[iel] or e, v).  x: host neq;
 * e_elem: host nels or NULL; sigma: host (6, nels) column-major.  Needs the
 * mesh and an elastic steering (ntot = 3 nod). */
int pf_stress_centroid(pf_ctx *ctx, const double *x, double e, double v, const double *e_elem, double *sigma);
/* SCATTER_NODES(npes,nn,nels_pp,g_num_pp,nod,numvar,nodes_pp,node_start,
 * node_end,element_contribution,nodal_value,flag)
 * (src/modules/mpi/gather_scatter.f90:1786-1929) at one rank (npes = 1,
 * nodes_pp = nn, node_start = 1, node_end = nn): nodal_value((n-1)*numvar
 * + k) is the sum, in ascending element and then element-node order, of
 * element_contribution((j-1)*numvar + k, iel) over every element node j of
 * element iel that is node n; flag 1 divides each sum by its number of
 * terms (the nodal average; 0/0 for a node no element uses, as there).
 * Node numbers outside 1..nn are skipped.  p121 passes the gathered
 * displacements with numvar = ndim (p121.f90:132-133).  Host arrays:
 * g_num_pp (nod x nels), element_contribution (numvar*nod x nels),
 * nodal_value (numvar*nn). */
int pf_scatter_nodes(pf_ctx *ctx, int nn, int nod, int nels, const int *g_num_pp, int numvar,
                     const double *element_contribution, double *nodal_value, int flag);

/* ------------------------------------------------- multi-RHS block solve */
/* No counterpart in the reference: up to 64 load cases on one mesh solved
 * together.  The element product for all of them, K_e (ntot x ntot) times
 * the gathered (ntot x k) block, runs on the matrix cores (FP32 MFMA in the
 * iteration, FP64 MFMA for the true residuals); the iteration per column is
 * p121's Jacobi PCG (p121.f90:86-104), batched as PCG_KM batches
 * MATMUL(km, pmul_pp) (src/modules/mpi/maths.f90:1152-1323).  Solver: FP32
 * block PCG (per-column alpha/beta) with FP64 residual replacement: when a
 * column's recurrence residual has fallen by inner_tol since the last
 * replacement, x is accumulated in FP64, r = f - K x is recomputed in FP64
 * and the iteration continues from it (from p = dinv r again if a column's
 * true residual grew).  Needs FP64 element matrices
 * (PF_STORE_FP64, full or shared) and no p123 fixed freedoms. */
typedef struct {
    double tol;          /* per column: ||f - K x||_2 <= tol ||f||_2          */
    double inner_tol;    /* residual drop that triggers a replacement (0=0.1) */
    int    inner_limit;  /* block iterations (0 = 100000)                     */
    int    outer_limit;  /* residual replacements (0 = no limit)              */
    int    check_every;  /* host polls every n iterations (0 = 32)            */
} pf_block_opts;

typedef struct {
    int    outer;        /* residual replacements                             */
    int    inner_iters;  /* block iterations                                  */
    int    converged;    /* columns that met tol                               */
    double max_relres;   /* largest true relative residual over the columns   */
    int    status;       /* PF_OK or PF_NOT_CONVERGED                          */
} pf_block_result;

/* f, x: host, neq x k column-major (column c = load case c), 1 <= k <= 64. */
int pf_block_solve(pf_ctx *ctx, int k, const double *f, double *x, const pf_block_opts *opts,
                   pf_block_result *res);
/* y = K x for k columns through the block element product: precision
 * PF_PREC_FP64 (FP64 MFMA) or PF_PREC_MIXED (fl32(x), FP32 MFMA, FP32
 * scatter).  Host, neq x k column-major. */
int pf_block_apply(pf_ctx *ctx, int k, const double *x, double *y, int precision);
/* Measurement: average ms of the FP32 matrix-core element product over reps
 * launches, HIP events on the context's stream, applied to the FP32 block
 * the last pf_block_apply(PF_PREC_MIXED) converted or the last residual of
 * a block solve. */
int pf_block_time_ebe(pf_ctx *ctx, int reps, double *ms);

/* ------------------------------------------------------- launch shapes */
/* No counterpart in the reference (its element loop has no launch shape):
 * the shapes of the element-product kernels, picked per (ntot, precision,
 * k, device) by the autotuner in parafem_amd/tune.py and set per context.
 * 0 keeps the built-in choice.  None of them changes a result: every lane
 * forms its row in the same order whatever the shape.
 *   ebe_epl:   single-RHS products, elements a lane group loads before it
 *              multiplies: 1 or 2 (ntot 60), 1, 2 or 4 (ntot 24), 1, 2, 4
 *              or 8 (ntot 8);
 *   blk_waves: block products, elements (wavefronts) per workgroup: 1, 2, 4;
 *   blk_rows:  FP64 block products, 16-row output tiles per pass: 1, 2, 4
 *              (the FP32 block product forms every row tile in one pass).
 * Setting them drops a captured iteration chunk. */
typedef struct {
    int ebe_epl;
    int blk_waves;
    int blk_rows;
} pf_tune;
int pf_tune_set(pf_ctx *ctx, const pf_tune *t);
int pf_tune_get(pf_ctx *ctx, pf_tune *t);
/* The device's architecture name (gcnArchName, e.g. "gfx950:sramecc+:xnack-")
 * and compute-unit count: the autotuner's cache key. */
int pf_device_info(pf_ctx *ctx, char *arch, int len, int *cus);

/* ------------------------------------------------------------ timing */
/* No counterpart in the reference (its timing is CPU_TIME around phases,
 * src/modules/shared/timing.f90:35-62).  While enabled, pf_pcg_run records a
 * HIP event pair on the context's stream around every kernel it launches.
 * pf_pcg_timing_get synchronises and returns, per kernel group of an
 * iteration (0 element products, 1 scatter + p.u, 2 x/r/d update, 3 p
 * update, 4 the mixed solver's true-residual replacement, whose launches do
 * nothing on iterations without a replacement), the summed milliseconds and
 * the launch count since the last enable.  pf_block_solve records its FP32
 * iteration the same way (0 block element products, 1 block scatter + p.u,
 * 2 x/r update, 3 p update). */
#define PF_TIMING_SLOTS 5
int pf_pcg_timing_enable(pf_ctx *ctx, int on);
int pf_pcg_timing_get(pf_ctx *ctx, double ms[PF_TIMING_SLOTS], int counts[PF_TIMING_SLOTS]);

/* ------------------------------------------------- layout introspection */
/* Grid size G = clamp(ceil(n/256), 1, 16384) (blocks of 256 threads) the
 * FP64 reductions use for length n.  It fixes the summation order of every
 * dot product: thread g = 256*b + t adds a_i*b_i for i = g, g + 256*G, ...
 * in turn (s = s + a*b, no fused multiply-add); each 64-lane wavefront
 * folds its lanes with offsets 32,16,8,4,2,1 (v[l] = v[l] + v[l+off]); a
 * block adds its 4 wavefront sums left to right; a one-block kernel folds
 * the G block sums the same way (thread t sums entries t, t+256, ..., then
 * wavefronts, then the 4 wavefront sums left to right). */
int pf_reduce_grid(int64_t n);

/* Summation order of the FP64 dot products and sums (DOT_PRODUCT_P, SUM_P
 * and the dots of the PCG loop) on a context:
 *   PF_DOT_TREE (default): the fixed two-level tree above;
 *   PF_DOT_SEQ: the reference's order at one rank, s = s + a(i)*b(i) for
 *     i = 1..n ascending (DOT_PRODUCT inside DOT_PRODUCT_P,
 *     src/modules/mpi/maths.f90:168-216; SUM inside SUM_P, :271-315),
 *     every product and sum rounded as written.  One wavefront walks the
 *     chain, so a dot costs about 12 clocks per entry: meant for parity runs
 *     against the reference's FP64 trajectory, not for throughput.
 * Maxima (checon_par) are exact in any order.  The order applies to the
 * calls made after the change; a captured iteration chunk is re-captured. */
#define PF_DOT_TREE        0
#define PF_DOT_SEQ         1
int pf_dot_order_set(pf_ctx *ctx, int order);
int pf_dot_order_get(pf_ctx *ctx, int *order);

/* How the mixed solver (PF_PREC_MIXED) sums the element products into u --
 * SCATTER in p121.f90:89-97 (gather_scatter.f90:694-850):
 *   PF_SCATTER_SLOTS (default): the products go to utemp and every
 *     equation adds its element slots in ascending element order (the
 *     reference's order; every other precision always does this);
 *   PF_SCATTER_RANGES: one workgroup forms the products of a range of
 *     consecutive elements and adds them per equation in LDS, then into u
 *     once; ranges sharing an equation run in different launches (colours).
 *     No utemp round trip.  Deterministic, but the sums' order is not the
 *     reference's.  Built at pf_pcg_begin from the steering and fixed rows
 *     (PF_ERR_ARG if the steering does not allow it: an equation listed
 *     twice by one element, or more than 32 colours).  The batch
 *     (pf_batch_*) refuses contexts in this mode. */
#define PF_SCATTER_SLOTS   0
#define PF_SCATTER_RANGES  1
int pf_scatter_mode_set(pf_ctx *ctx, int mode);
int pf_scatter_mode_get(pf_ctx *ctx, int *mode);

#ifdef __cplusplus
}
#endif
#endif /* PARAFEM_AMD_H */
