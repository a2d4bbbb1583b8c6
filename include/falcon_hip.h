/* falcon_hip.h - C-ABI of the MI355X-native Falcon hot path (libfalcon_hip.so).
 *
 * Plain C types only.  Every function returns 0 on success and -1 on failure;
 * falcon_last_error() then holds the message (the Fortran shims map != 0 to
 * `stop 1`, the reference's only failure mode).  A handle owns one Yin-Yang
 * block tree with every block resident in HBM; host arrays are stale until
 * falcon_get_state().  One HIP stream per handle; a handle is not re-entrant.
 *
 * Host arrays use the reference's Fortran layout: a padded block array is
 * (-ng+1:ni+ng, -ng+1:nj+ng, -ng+1:nk+ng, 8) column-major (r fastest), with the
 * variables rho1, vr, vt, vp, br, bt, bp, s1 (3_6/ModBlock.f90:153-163, psi
 * dropped: no term of the reference writes or reads it when DivB_option=0).
 * Block ids are 0-based here; the reference's leaf labels are id+1.
 */
#ifndef FALCON_HIP_H
#define FALCON_HIP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct falcon_handle *falcon_handle_t;

/* ---------------------------------------------------------------- tree
 * Replaces YinYangTree_InitTree (3_6/ModYinYangTree.f90:91-110) plus
 * ModAMR_set_grid (3_6/ModAMR.f90:13-30) for drivers that build the tree in
 * the library.  r range in cm (3_6/ModReadParameters.f90:71). */
int falcon_create(falcon_handle_t *out, int ni, int nj, int nk, int ng, double r_lo_cgs, double r_hi_cgs);
int falcon_destroy(falcon_handle_t h);
/* Divide every leaf into nr x nt x np children with the split rule of
 * YinYangTree_Divide (3_6/ModYinYangTree.f90:210-269); {1,2} per axis is
 * YinYangTree_DivideAll (:129-136), larger N generalises the same formulas. */
int falcon_tree_divide_all(falcon_handle_t h, int nr, int nt, int np);
/* One AMR pass, ModAMR_Divide_OneBranch (3_6/ModAMR.f90:65-99): flags =
 * rtp_if_divide, rng_rsun = rtp_range_Rsun(3,2) column-major. */
int falcon_tree_amr(falcon_handle_t h, const int flags[3], int if_global, const double rng_rsun[6]);
/* Local refinement (no counterpart in the reference's control files, which
 * refine by rtp_range only): split the listed leaves, given as 0-based
 * ordinals in the current depth-first order (Yin then Yang), into 2x2x2
 * children with YinYangTree_Divide's rule. */
int falcon_tree_refine(falcon_handle_t h, int n, const int *leaf_ordinals);
/* Drop-in path: take the Fortran driver's own tree.  Nodes in depth-first
 * pre-order (Yin root first, then its subtree, then the Yang root); parent[n] = -1
 * for roots; rng6 = rtp_range(3,2) column-major per node. */
int falcon_tree_set_nodes(falcon_handle_t h, int n_nodes, const int *parent, const int *is_yin, const double *rng6);
/* Leaf count after labelling (YinYangTree_SetAll, 3_6/ModYinYangTree.f90:294-344). */
int falcon_nblocks(falcon_handle_t h, int *nblocks, int *nleaf_yin);
/* xijk_range(3,2), flags = (if_yin, if_top, if_bottom), padded cell centres. */
int falcon_block_info(falcon_handle_t h, int ib, double *rng6, int *flags3, double *xi, double *xj, double *xk);

/* ---------------------------------------------------------------- tables
 * ModCommunication_SetGCAll (3_6/ModCommunication.f90:22-64): donor search,
 * GC targets, interpolation bases/weights and Yin-Yang rotations, built on the
 * host and uploaded once.  Also allocates every block's state in HBM. */
int falcon_set_gc_all(falcon_handle_t h);
/* Build the ghost tables of falcon_set_gc_all on the device (on != 0) or on
 * the host (0, the default); before the tables are built.  The device build
 * runs the donor search of every ghost cell (YinYangTree_FindNode,
 * 3_6/ModYinYangTree.f90:449-505), the target grouping (ModGC_GetGC_Targets_
 * single, 3_6/ModGC.f90:53-147) and the interpolation data on the GPU; the
 * tables equal the host build's bit for bit, after any refinement. */
int falcon_set_device_tables(falcon_handle_t h, int on);
/* Host half of falcon_set_gc_all (no device calls): tables only. */
int falcon_build_tables(falcon_handle_t h);
/* GC_iBlocks_III of block ib (padded, Fortran order; -777 = outside the domain,
 * interior cells -777 as after ModGC_FindGC_Single, 3_6/ModGC.f90:34). */
int falcon_gc_iblocks(falcon_handle_t h, int ib, int32_t *out);
int falcon_gc_ntargets(falcon_handle_t h, int ib, int *ntargets);
/* GC_targets(it) of block ib (3_6/ModGC.f90:111-143): donor leaf label (1-based),
 * nGC, ijk_list(nGC,3) column-major.  ijk may be NULL to query ngc. */
int falcon_gc_target(falcon_handle_t h, int ib, int it, int *donor, int *ngc, int32_t *ijk);
/* All ghost entries in device order: n on query (arrays NULL), else per entry
 * recv block, donor block, ghost ijk[3], base ijk[3], weights[3], yy flag, rot[2]. */
int falcon_gc_entries(falcon_handle_t h, int64_t *n, int32_t *recv, int32_t *donor, int32_t *gijk, int32_t *bijk,
                      double *w, int32_t *yy, double *rot);
/* FP32 step's pull table (host only, after falcon_build_tables): pull (NULL or
 * nblocks*27 int32) receives, for block R and direction d of its ghosts (entry
 * R*27 + (d0+1)*9 + (d1+1)*3 + (d2+1)), the same-level block whose interior the
 * kernels read those ghosts from, or -1; n_copy: the FP32 copy entries, n_rest:
 * those the ghost fill still runs.  No counterpart in the reference. */
int falcon_ghost_pull_table(falcon_handle_t h, int32_t *pull, int64_t *n_copy, int64_t *n_rest);

/* ---------------------------------------------------------------- data
 * Background profiles per block, replacing ModBlock_InitSSM's calls to
 * ModStratification_new_get_vars (3_6/ModBlock.f90:232-249): bg is 8 x NI
 * column-major: g, rho0, p0, T0, gamma1, gamma3, heat (diffusion+cooling), xi_rsst. */
int falcon_set_background(falcon_handle_t h, int ib, const double *bg);
int falcon_get_background(falcon_handle_t h, int ib, double *bg);
/* n=1.5 polytrope for every block: rho(r_lo)/rho(r_hi) = contrast. */
int falcon_background_polytrope(falcon_handle_t h, double contrast, double rho_top);
int falcon_set_state(falcon_handle_t h, int ib, const double *prim);
int falcon_get_state(falcon_handle_t h, int ib, double *prim);

/* GLM divergence cleaning: DivB_option of 3_6/ModParameters.f90:70, 1 = GLM
 * (ModDivB_GLM, 3_6/ModDivB.f90:13-38, called from 3_6/ModAdvance.f90:82-84).
 * With option 1 psi, the ninth variable of the reference's MHD blocks
 * (3_6/ModBlock.f90:163), is stepped, exchanged and diffused with the state;
 * FP64 path only (falcon_set_precision(h, 1) is refused, and the reverse).
 * psi is set and read per block as one padded array in Fortran order
 * (-ng+1:ni+ng, -ng+1:nj+ng, -ng+1:nk+ng), i.e. primitive_IV(:,:,:,psi_);
 * falcon_set_psi writes the stage buffers too, as the reference's
 * primitive_rk_IV starts as a copy of primitive_IV. */
int falcon_set_divb(falcon_handle_t h, int option);
int falcon_set_psi(falcon_handle_t h, int ib, const double *psi);
int falcon_get_psi(falcon_handle_t h, int ib, double *psi);
/* Seeded synthetic initial state on the interior: rho1 = v = 0, s1 = s_amp*c_p*U(-1,1),
 * B = curl of a seeded smooth vector potential scaled to max|B|/sqrt(rho0) = b_ratio*c_s.
 * Ghosts are then filled by two exchanges (3_6/test1.f90:143-144). */
int falcon_init_synthetic(falcon_handle_t h, uint64_t seed, double s_amp, double b_ratio);

/* ---------------------------------------------------------------- stepping
 * precision: 0 = FP64 everywhere; 1 = FP32 stencil arithmetic on the deviation
 * state with FP64 background, stage accumulation, ghost weights and dt. */
int falcon_set_precision(falcon_handle_t h, int mode);
/* FP32 step: on (default), the stage and diffusion kernels read the ghosts
 * that are plain copies of a same-level neighbour's cells from that
 * neighbour's interior, and the ghost fill between them runs only the other
 * entries; off, the ghost fill runs every entry and the kernels read their own
 * ghost cells.  Both read the same values.  No counterpart in the reference
 * (its ghost fill is ModCommunication_SendRecvGC_new). */
int falcon_set_ghost_pull(falcon_handle_t h, int on);
/* Launch shapes of the stage and diffusion marches for one precision (0 FP64,
 * 1 FP32 deviation).  tile: the (theta, phi) tile edge, 16 (16x16 lanes, 256
 * per workgroup) or 8 (8x8, 64 lanes).  wgs: 0 launches one workgroup per
 * tile; n > 0 launches n workgroups that loop over the tiles (a persistent
 * launch).  16 and 0 for both when the handle is made.  Every result is the
 * same in any shape.  shape[4] of get: stage tile, stage wgs, diffusion tile,
 * diffusion wgs.  No counterpart in the reference (its loops run over whole
 * blocks). */
int falcon_set_tiles(falcon_handle_t h, int precision, int stage_tile, int stage_wgs, int diffusion_tile,
                     int diffusion_wgs);
int falcon_get_tiles(falcon_handle_t h, int precision, int *shape);
/* Autotuner for the current precision: times the stage kernel (a step's four
 * stage launches) and the diffusion kernel, reps times each (1..1000), twice
 * over four shapes on the current state: tile 16 or 8, one workgroup per tile
 * or the persistent grid (the workgroups the device holds resident at once).
 * Keeps the fastest of each.
 * shape_out[4] (optional): the shape chosen, as in falcon_get_tiles;
 * ms_out[8] (optional): ms per stage launch (a step's four averaged) in the
 * shapes (16, per tile), (16, persistent), (8, per tile), (8, persistent),
 * then ms per diffusion launch in the same order.  The state and later results are unchanged.
 * No counterpart in the reference. */
int falcon_autotune(falcon_handle_t h, int reps, int *shape_out, double *ms_out);
/* ModCommunication_SendRecvGC_new(Tree, if_rk) (3_6/ModCommunication.f90:655-795):
 * which = 0 fills the state's ghosts, 1 the current stage buffer's. */
int falcon_sendrecv_gc(falcon_handle_t h, int which);
/* ModAdvance_rk4(Tree, .true., dt) (3_6/ModAdvance.f90:18-112): one step, dt out. */
int falcon_advance_rk4(falcon_handle_t h, double cfl, double *dt_out);
/* nsteps steps without host synchronisation; use_graph captures two steps in a
 * hipGraph and replays it.  dt_last = dt of the last step. */
int falcon_advance(falcon_handle_t h, int nsteps, double cfl, int use_graph, double *dt_last);
/* ModCheck_primitive (3_6/ModCheck.f90:12-71) on the device: which = 0 scans
 * the state (primitive_IV), 1 the current stage buffer (primitive_rk_IV), every
 * padded cell and variable.  n_nan = NaN count; where[5] (optional) = block
 * (0-based), I, J, K (padded, 0-based: index - (-ng+1)) and variable (0-based;
 * 8 is psi, scanned with GLM divergence cleaning on) of the first NaN in the
 * reference's scan order (blocks in order, then phi, theta, r, variable
 * innermost), or -1s.  The Fortran shim turns n_nan > 0
 * into the reference's message and `stop 1` (its MPI_Abort). */
int falcon_check_state(falcon_handle_t h, int which, int64_t *n_nan, int32_t *where);
/* ModSave_AllCells (3_6/ModSavePlot.f90:520-650): writes
 * <outdir>/AllCells_<8-digit step>/info.dat and one Block_<8-digit label>.dat
 * per block, each line the corner's Cartesian position in solar radii and the
 * 9 variables averaged over its 8 surrounding cells (v, B in Cartesian
 * components, psi = 0), in Fortran ES16.8 fields, byte for byte as the
 * reference's writer.  The corner values are formed on the device from the
 * HBM state (ghosts refreshed first); the host formats and writes.  merged =
 * 1 writes merged.dat (the Tecplot FEPOINT BRICK file of
 * 3_6/merge_AllCells.py) and .merged in place of the block files. */
int falcon_save_allcells(falcon_handle_t h, int istep, const char *outdir, int merged);
/* n values as Fortran ES16.8 fields: out receives 16*n characters and a
 * terminating NUL.  The edit the AllCells writer uses; no counterpart in the
 * reference beyond Fortran's own formatted output. */
int falcon_format_es16(const double *x, int n, char *out);
/* dt > 0: every later step uses this dt in place of CFL * min(h / v_wave)
 * (ModTimeStep_Dynamo, 3_6/ModTimeStep.f90:10-30); 0 restores the CFL dt.  For
 * stepping two precisions with one dt sequence.  No counterpart in the
 * reference. */
int falcon_set_fixed_dt(falcon_handle_t h, double dt);
/* Interior difference of the current states of two handles with the same
 * layout on one device, per variable v (rho1, vr, ..., s1): max_diff[v] =
 * max|a - b|, max_ref[v] = max|b|, n_above[v] = cells with |a - b| >
 * rel_tol * max|b|, argmax[v] = smallest interior index ((blk*ni + i)*nj +
 * j)*nk + k (0-based) where |a - b| = max_diff[v].  Any output may be NULL.
 * No counterpart in the reference (comparison of two runs). */
int falcon_state_diff(falcon_handle_t a, falcon_handle_t b, double rel_tol, double *max_diff, double *max_ref,
                      int64_t *n_above, int64_t *argmax);
/* ModWaveSpeed_Dynamo + ModTimeStep_Dynamo (3_6/ModWaveSpeed.f90, ModTimeStep.f90). */
int falcon_compute_dt(falcon_handle_t h, double cfl, double *dt);
/* falcon_compute_dt without allocating: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight stepping loops).  Same result as falcon_compute_dt, bit for bit, but its
 * partial minima and dt live in a fixed static device array of the library (one
 * per GPU, held until the process ends) instead of the handle's buffers, and it
 * calls no hipMalloc/hipFree.  Limits: at most 16384 partial minima (nblocks x
 * the diffusion's tiles per block: the 128x256x768-per-panel shell in 16^3
 * blocks needs 12288 at tile edge 16), else an error; the handle's device
 * tables must already be on the GPU (after a first falcon_compute_dt,
 * falcon_sendrecv_gc or step), else an error; call it with the handle's GPU
 * current.  Calls on one GPU share the array, so calls that overlap in time on
 * one GPU (two host threads, two handles) race. */
int falcon_compute_dt_noalloc(falcon_handle_t h, double cfl, double *dt);
/* Single operators for parity: R of ModEquation_Dynamo after ModBoundary on the
 * state (3_6/ModEquation.f90:24-37), and of ModDiffusion_Aritificial_1 (h=2).
 * Out: per block (ni,nj,nk,8) Fortran order, blocks consecutive. */
int falcon_rhs(falcon_handle_t h, double *r_out);
int falcon_diffusion_rhs(falcon_handle_t h, double *r_out);
/* Average kernel durations (ms, HIP events on the handle's stream) over nsteps
 * steps: out[0] dt, [1] boundary, [2] stage, [3] exchange (stage buffers),
 * [4] exchange (state), [5] diffusion, [6] whole step. */
int falcon_profile_step(falcon_handle_t h, int nsteps, double cfl, double *out_ms);
/* Bytes of device memory held by the handle. */
int falcon_device_bytes(falcon_handle_t h, int64_t *bytes);
int falcon_synchronize(falcon_handle_t h);

/* ---------------------------------------------------------------- heat conduction
 * ModHeatConduction (3_6/ModHeatConduction.f90:1-16) is an empty module in the
 * reference; these entry points define its solve in the shape of
 * ModPFSS_setup/ModPFSS_solve (3_5/ModPFSS.f90:21,102):
 *   (V/dt) T + sum_f kappa_f S_f/D_f (T - T_nb) = (V/dt) T^n,
 *   kappa(r) = kappa0 (r / r_ref)^-3, insulated radial boundaries,
 * solved by flexible CG preconditioned with a per-block multigrid V-cycle
 * (3_5/ModLinearSolver.f90:201-427 as a method), FP64, for nsys systems at once.
 * T arrays: nsys x nblocks x (n,n,n) Fortran order per block, cubic blocks. */
int falcon_heat_setup(falcon_handle_t h, double kappa0, double r_ref_rsun, int nsys);
int falcon_heat_info(falcon_handle_t h, int *levels, int64_t *cells, double *dt_explicit);
int falcon_heat_set_dt(falcon_handle_t h, const double *dt);
/* right-hand side b = (V/dt) T^n from T^n (set dt first) */
int falcon_heat_set_T(falcon_handle_t h, const double *T);
/* seeded Gaussian T^n with standard deviation sigma */
int falcon_heat_seed_T(falcon_handle_t h, uint64_t seed, double sigma);
/* ModHeatConduction_solve: to ||b - A T||/||b|| <= tol; use_corrector adds the
 * learned correction to the V-cycle (falcon_cnn_*).  Per system iterations and
 * final relative residual; wall seconds of the solve. */
int falcon_heat_solve(falcon_handle_t h, double tol, int max_iter, int use_corrector, int32_t *iters, double *relres,
                      double *seconds);
int falcon_heat_get_T(falcon_handle_t h, int sys, double *T);
/* on (default): each FCG iteration is one replay of a captured hipGraph with the
 * convergence test on the device; off: the same kernels launched one by one */
int falcon_heat_set_graph(falcon_handle_t h, int on);
/* learned correction step: one per 16^3 block (on, default) or one per system
 * (off).  No counterpart in the reference (learned V-cycle correction). */
int falcon_heat_set_learned_blockstep(falcon_handle_t h, int on);
/* the learned correction builds the first n search directions of a learned
 * solve and the plain V-cycle the rest; 0 (default): every direction.  No
 * counterpart in the reference (learned V-cycle correction). */
int falcon_heat_set_learned_iters(falcon_handle_t h, int n);
/* training loss of the network: h1_alpha < 0 the mean squared error
 * (default); h1_alpha >= 0 the gradient-weighted norm
 * |v|_A^2 = h1_alpha |v|^2 + sum_a w_a |differences of v along axis a|^2 over
 * each 16^3 image (axes r, theta, phi with weights w_r, w_t, w_p) of
 * d = output - target, or with scale_free (h1_alpha > 0) the fraction
 * 1 - <c,Ae>^2/(|c|_A^2 |e|_A^2) of the target's norm that the best multiple
 * of the output c leaves.  No counterpart in the reference (learned V-cycle
 * correction). */
int falcon_cnn_set_loss(falcon_handle_t h, double h1_alpha, double w_r, double w_t, double w_p, int scale_free);
/* Stop of the coarsest-level CG: r.z <= tol^2 (r.z)_0 or max_iter iterations
 * (defaults 1e-2 and 50; the reference's JCB_CG stops at an absolute r.z). */
int falcon_heat_set_coarse(falcon_handle_t h, double tol, int max_iter);
/* y = A x for all systems (parity tests) */
int falcon_heat_apply(falcon_handle_t h, const double *x, double *y);
/* z = M r: one V-cycle of the preconditioner for all systems, zero initial
 * guess (ModMultigrid_Vcycle, 3_5/ModMultigrid.f90:338-376, as a method):
 * damped Jacobi on the per-block levels, summed restriction, linear
 * prolongation, and on the 1-cell-per-block level a Jacobi-preconditioned CG
 * over the whole system (ModLinearSolver_JCB_CG, 3_5/ModLinearSolver.f90:86-180) */
int falcon_heat_vcycle(falcon_handle_t h, const double *r, double *z);

/* ---------------------------------------------------------------- PFSS potential field
 * ModPFSS_setup / ModPFSS_solve (3_5/ModPFSS.f90:21-148) on the solver above:
 * the finite-volume Laplacian (uniform kappa = 1, no V/dt term: 1/dt = 0),
 * insulated bottom faces whose flux b = Br S is the right-hand side (:38-44),
 * and the value 0 on the top face, the source surface (:88-91: the face's
 * coefficient enters the diagonal twice and couples to nothing).  The solve is
 * the flexible CG with the multigrid V-cycle (:116), FP64, one system.  The
 * setup replaces any heat setup of the handle: falcon_heat_apply and
 * falcon_heat_vcycle then apply the PFSS operator.  Cubic power-of-two blocks.
 * Magnetograms: for each bottom block (falcon_pfss_bottom_blocks, block
 * order) Br at its nj x nk bottom faces, j fastest. */
int falcon_pfss_setup(falcon_handle_t h);
/* the blocks at the domain's inner radius, in block order, and their number */
int falcon_pfss_bottom_blocks(falcon_handle_t h, int32_t *blocks, int *n);
/* the areas r_F^2 (cos t_j - cos t_j+1) dphi of those blocks' bottom faces
 * (Si_FLL(1,:,:), 3_5/ModMultigrid.f90:150-233), in magnetogram order */
int falcon_pfss_bottom_areas(falcon_handle_t h, double *area);
/* right-hand side b = Br S from a magnetogram (3_5/ModPFSS.f90:38-44) */
int falcon_pfss_set_magnetogram(falcon_handle_t h, const double *br);
/* ModMagnetogram_dipole_magnetogram_ALL (3_5/ModMagnetogram.f90:13-80): Br of
 * n dipoles (centres xyz[3n], pole charges q[n], pole separations l[3n], cm)
 * at the bottom faces, written to br (a magnetogram) */
int falcon_pfss_dipole_magnetogram(falcon_handle_t h, int n, const double *xyz, const double *q, const double *l,
                                   double *br);
/* ModPFSS_solve with i_option = 1 (3_5/ModPFSS.f90:102-116): to ||b - A phi||/||b|| <= tol */
int falcon_pfss_solve(falcon_handle_t h, double tol, int max_iter, int32_t *iters, double *relres, double *seconds);
/* phi: nblocks x (n,n,n) Fortran order per block */
int falcon_pfss_get_phi(falcon_handle_t h, double *phi);
/* B = -grad(phi) (3_5/ModPFSS.f90:138-145 with ModSpherical_Grad_f_O2,
 * 3_6/ModSpherical.f90:161-199), second order, at every cell: nblocks x 3 x
 * (n,n,n) Fortran order (r, theta, phi components).  Neighbours across block
 * faces are the operator's face ghosts; at the bottom phi_0 = phi_1 + Br dr,
 * at the top phi_{n+1} = -phi_n. */
int falcon_pfss_field(falcon_handle_t h, double *B);

/* ---------------------------------------------------------------- learned correction
 * No counterpart in the reference.  A 1 -> 16 -> 16 -> 1 network of 3x3x3
 * convolutions over 16^3 block images (zero padding, ReLU after layers 1-2),
 * BF16 on MFMA with FP32 accumulation, FP32 master weights.  Parameters (7809
 * floats): W1[27][16], b1[16], W2[27][16][16] (tap, in, out), b2[16],
 * W3[27][16], b3; tap t = (di+1)*9 + (dj+1)*3 + (dk+1); images [4096] with
 * voxel (i*16 + j)*16 + k. */
int falcon_cnn_init(falcon_handle_t h, uint64_t seed);
int falcon_cnn_set_params(falcon_handle_t h, const float *params);
int falcon_cnn_get_params(falcon_handle_t h, float *params);
int falcon_cnn_forward(falcon_handle_t h, int64_t nimg, const float *x, float *out);
/* sum over the images of d(sum gout*out)/d(params) */
int falcon_cnn_gradient(falcon_handle_t h, int nimg, const float *x, const float *gout, double *grad);
/* Adam on the mean squared error, shuffled minibatches; last epoch's mean loss */
int falcon_cnn_train(falcon_handle_t h, int64_t nimg, const float *x, const float *y, int epochs, int batch, double lr,
                     uint64_t seed, double *loss);
/* Solver-generated training: batch solves of the handle's nsys systems with dt
 * drawn log-uniformly in [dt_lo, dt_hi] x the explicit limit and seeded Gaussian
 * T' (sigma 1e-3) record npairs snapshots (one system at one V-cycle
 * application: normalised residual left by the V-cycle, normalised error it
 * left), each as blocks_per_pair random block images; the network trains on
 * them (Adam, learning rate decaying linearly to lr/10) and is attached to the
 * solver (falcon_heat_solve with use_corrector = 1).  Leaves the last training
 * system's dt and right-hand side set. */
int falcon_heat_train_corrector(falcon_handle_t h, int npairs, int blocks_per_pair, double dt_lo, double dt_hi,
                                int max_solves, int epochs, int batch, double lr, uint64_t seed, double *loss,
                                int64_t *pairs_made, double *seconds);
/* Finest-level pre/post smoothing sweeps of the V-cycle that the learned
 * correction follows (training records and learned solves; default 2 / 2, the
 * plain method's).  The network can take the place of finest-level sweeps. */
int falcon_heat_set_learned_sweeps(falcon_handle_t h, int pre0, int post0);
/* attach the handle's network (e.g. after falcon_cnn_set_params) */
int falcon_heat_attach_corrector(falcon_handle_t h);

/* ---------------------------------------------------------------- errors */
int falcon_last_error(char *buf, int len);

#ifdef __cplusplus
}
#endif
#endif
