/*
 * hbem.h -- C-ABI of the MI355X-native HelmholtzBEM sigma_min(k) path.
 *
 * This is the drop-in boundary.  The reference exposes the path as a C++
 * library API (GalerkinBuilder / SolutionsOperator / randomized_svd); this ABI
 * is the layer underneath a C++ facade that keeps those names
 * (include/hbem/hbem.hpp), and is what ctypes / cgo / JNI style bindings call.
 *
 * Conventions
 *   - Every function returns 0 on success and a negative code on error; the
 *     message of the last error on the calling thread is returned by
 *     hbem_last_error().  The messages of the reference's runtime_errors are
 *     kept verbatim (galerkin_builder.cpp:901-906, 1050-1056).
 *   - Matrices are column-major arrays of complex double ("double _Complex" /
 *     std::complex<double>, interleaved re,im), leading dimension = rows, the
 *     layout of Eigen::MatrixXcd.  Host pointers unless the name says _dev.
 *   - Panels: 8 doubles each {type, a0..a6}:
 *       type 0  line        {x1, y1, x2, y2}            ParametrizedLine
 *       type 1  arc         {cx, cy, r, phi0, phi1}     ParametrizedCircularArc
 *       type 2  trig curve  {t0, t1, first, count}      z(t) = sum c_m e^{imt}
 *       type 3  spline      {tmin, tmax, first, pieces} ParametrizedPolynomial
 *     with a coefficient table of rows of 3 doubles: {m, re, im} for type 2;
 *     for type 3 three rows per cubic piece j of the spline, {t_j, x_j, y_j},
 *     {bx, cx, dx}, {by, cy, dy} (z = z_j + b d + c d^2 + e d^3, d = t - t_j),
 *     and a closing row {t_K, x_K, y_K}.
 *   - Spaces: test = trial = continuous P1 (ContinuousSpace<1>), the space of
 *     every sigma_min driver in the reference.
 *   - One GPU; all work of a geometry context runs on the context's stream.
 */
#ifndef HBEM_H
#define HBEM_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* error codes */
#define HBEM_OK 0
#define HBEM_EINVAL (-1)     /* invalid argument (reference: std::runtime_error / out_of_range) */
#define HBEM_ENOTSUP (-2)    /* not supported by this library (e.g. submatrix assembly) */
#define HBEM_EDEVICE (-3)    /* HIP runtime error */
#define HBEM_ENOMEM (-4)     /* allocation failed */
#define HBEM_ENUMERIC (-5)   /* numerical failure of a solve (reference: std::runtime_error) */

/* LayerType bits (galerkin_builder.hpp:100-106) */
#define HBEM_LAYER_DOUBLE 1
#define HBEM_LAYER_SINGLE 2
#define HBEM_LAYER_HYPERSINGULAR 4
#define HBEM_LAYER_ALL 7

/* geometry flags */
/* Reproduce the reference's coinciding single-layer symmetry copy for d >= 1
 * (galerkin_builder.cpp:744-746 copies the d = 0 entry into the d = 1, 2
 * entries).  Off by default: the default is the symmetric copy of the same
 * derivative order, which is what the k-derivative of V requires. */
#define HBEM_FLAG_QUIRK_SLP_COINC_DER 1
/* Lower-order quadrature on separated panel pairs.  Every assembly through the
 * context at real positive wavenumbers (matrices, T, sweeps) integrates the
 * general pairs of line and arc panels whose bounding discs lie at least one
 * panel diameter apart with a tensor Gauss rule of 5-10 points a side instead
 * of the quadrature order, the order chosen from the separation class (one to
 * two diameters, two to eight, eight to sixteen, more) and kappa times the
 * largest panel diameter -- per block of 32 wavenumbers in the sweeps, per
 * batch otherwise -- so that each pair's local matrices stay within 1e-13 of
 * their largest entry of the 11-point rule's.  Pairs on trigonometric-curve
 * or spline panels, complex wavenumbers, and kappa d > 5 keep the full rule.
 * Off by default: the reference integrates every general pair with the same
 * rule (galerkin_builder.cpp:308-327).  Not accepted by
 * hbem_sigma_sweep_multi. */
#define HBEM_FLAG_FAR_FIELD 2

typedef struct hbem_geometry* hbem_geometry_t;

/* Last error message of the calling thread; returns its length. */
int hbem_last_error(char* buf, size_t len);

/* Selects the HIP device used by contexts created afterwards. */
int hbem_init(int device);

/*
 * BuilderData(mesh, test_space, trial_space, order) + SolutionsOperator's ctor
 * (galerkin_builder.hpp:43-46, galerkin_builder.cpp:17-187;
 *  gen_sol_op.hpp:28, gen_sol_op.cpp:23-43).
 * Uploads the panels, builds the quadrature grids, classifies panel pairs
 * (coinciding / adjacent / general, galerkin_builder.cpp:189-201, 872-897),
 * precomputes the pair-node geometry on the device and, when proj != 0, the
 * mass matrix and the projection factor F = P^T L sqrt(D) of
 * blockdiag(M, M^T).
 */
int hbem_geometry_create(const double* panels, int npanels, const double* coefs, int ncoefs,
                         int order, int proj, int flags, hbem_geometry_t* out);
int hbem_geometry_destroy(hbem_geometry_t g);
/* Frees the context's per-call work buffers (local matrices, T, factor and
 * thin blocks; the geometry, node tables and projection factor stay); the
 * next call allocates them again.  For callers that keep many contexts alive.
 * (No counterpart in the reference, which allocates per call.) */
int hbem_geometry_release_work(hbem_geometry_t g);
/* getNumPanels, getGaussQROrder, getCGaussQROrder, getPanelsAreLines (galerkin_builder.hpp:94-97) */
int hbem_geometry_info(hbem_geometry_t g, int* npanels, int* order, int* corder, int* lines);

/*
 * GalerkinBuilder::assembleDense(k, c, layers, d) followed by the getters
 * getDoubleLayer(t) / getHypersingular(t) / getSingleLayer(t), t = 0..d
 * (galerkin_builder.hpp:221-276, galerkin_builder.cpp:918-1018, 1050-1068).
 * K, W, V: host arrays of (d+1) N x N matrices each (slice t at offset t*N*N),
 * or NULL to skip copying that layer.  Any k with Im k >= 0: real positive k
 * takes the real Hankel path, every other k the complex one
 * (galerkin_builder.cpp:224-235, 899-916).
 */
int hbem_assemble_dense(hbem_geometry_t g, double k_re, double k_im, double c, int layers, int d,
                        void* K, void* W, void* V);

/*
 * assembleDense over a grid of real wavenumbers in one pass: the same
 * matrices as nk calls of hbem_assemble_dense(g, ks[i], 0, c, layers, d, ...)
 * (galerkin_builder.cpp:918-1018), assembled batch-wise (all panel pairs x all
 * k per launch).  K, W, V: host arrays of nk (d+1) N x N matrices, slice t of
 * wavenumber i at offset (i (d+1) + t) N N, or NULL.  The reference drivers
 * make these calls one k at a time (examples/roots_brent_square_rsvd.cpp:167-171).
 */
int hbem_assemble_dense_batch(hbem_geometry_t g, const double* ks, int nk, double c, int layers, int d, void* K,
                              void* W, void* V);

/* mass_matrix::GalerkinMatrix as built by SolutionsOperator (mass_matrix.cpp:6-47,
 * gen_sol_op.cpp:30-33).  M: N x N real, column-major. */
int hbem_mass_matrix(hbem_geometry_t g, double* M);

/* The projection factor F (2N x 2N, real) and the LDLT transpositions. */
int hbem_projection_factor(hbem_geometry_t g, double* F, int* transpositions);

/*
 * SolutionsOperator::gen_sol_op / gen_sol_op_1st_der / gen_sol_op_2nd_der
 * (gen_sol_op.hpp:32-46, gen_sol_op.cpp:50-196).  Writes T and, for d >= 1,
 * T' (Td) and for d = 2 T'' (Tdd), each 2N x 2N, projected when the context
 * was created with proj != 0.  The derivative outputs follow the intended
 * semantics (the reference writes every assemble_operator call into T,
 * gen_sol_op.cpp:150-151, 193-195).
 */
int hbem_gen_sol_op(hbem_geometry_t g, double k_re, double k_im, double c_o, double c_i, int d,
                    void* T, void* Td, void* Tdd);

/* SolutionsOperator::projection(T) = F^{-1} T F^{-T} (gen_sol_op.cpp:45-48). */
int hbem_projection(hbem_geometry_t g, const void* T, void* Tp);

/*
 * tp::direct_second_kind::A_rhs (solvers.hpp:82-90, solvers.cpp:106-131):
 * A = gen_sol_op(k, c_o, c_i) (2N x 2N, projected when the context was
 * created with proj != 0, as the reference's sol_op argument decides) and
 * rhs = A_o u_inc_N with A_o = [[-K_o + M/2, V_o], [W_o, K_o^T + M/2]], the
 * operators K_o, W_o, V_o assembled at (k, c_o).  u_inc_N: 2N x nrhs, each
 * column the P1 interpolation coefficients of one incoming wave -- its values
 * at the panel start points (ContinuousSpace<1>::Interpolate_helmholtz,
 * continuous_space.hpp:67-82), then the normal derivatives there with the
 * normal (t_y, -t_x) / |t| of the panel tangent t (Interpolate_helmholtz_neu,
 * continuous_space.hpp:84-101).  The reference takes one column; nrhs
 * columns share one assembly.  A or rhs may be NULL.
 */
int hbem_tp_A_rhs(hbem_geometry_t g, double k_re, double k_im, double c_o, double c_i, const void* u_inc_N, int nrhs,
                  void* A, void* rhs);

/*
 * tp::direct_second_kind::solve (solvers.hpp:107-113, solvers.cpp:133-152):
 * the traces sol (2N x nrhs: Dirichlet, then Neumann coefficients of the
 * total field) solving A sol = rhs with A the unprojected gen_sol_op (the
 * reference builds its own SolutionsOperator(builder_data, false)),
 * whatever proj the context was created with.  LU with partial pivoting on
 * the device (the reference calls LAPACK zgesv, solvers.cpp:25-51); the
 * factorisation is shared by the nrhs columns.  Errors (HBEM_ENUMERIC) carry
 * the reference's messages: "Failed to solve for traces, right-hand side is
 * undefined" / "... solutions operator is undefined" (non-finite entries)
 * and "zgesv: LU is singular" (an exactly zero pivot); sol is then left
 * unwritten.
 */
int hbem_tp_solve(hbem_geometry_t g, double k_re, double k_im, double c_o, double c_i, const void* u_inc_N, int nrhs,
                  void* sol);

/*
 * randomized_svd::sv / sv_der / sv_der2 (randsvd.hpp:23-30,
 * randsvd.cpp:99-216) on host matrices; the factorisations run on the GPU.
 * out: sigma (sv), {sigma, sigma'} (sv_der), {sigma, sigma', sigma''} (sv_der2).
 * sv_der and sv_der2 take the eigenvector of the Wielandt matrix
 * [[0, T], [T^H, 0]] as the reference does, the last column of Q of the
 * column-pivoted QR of W - sigma I (randsvd.cpp:160-162): the pivot order of
 * that QR is computed on the GPU and the column itself, (W - sigma I)^{-1}
 * e_p for the last pivot p, by the blocked LU, so sigma' agrees with the
 * reference at any q, not only once the subspace has converged (the batched
 * sweeps use the subspace singular pair instead, see hbem_sigma_sweep).
 * sv_der2 then solves the reference's bordered system (randsvd.cpp:190-214).
 */
int hbem_rsvd_sv(const void* T, int n, const void* R, int nc, int q, double* sigma);
int hbem_rsvd_sv_der(const void* T, const void* Td, int n, const void* R, int nc, int q, double* out2);
int hbem_rsvd_sv_der2(const void* T, const void* Td, const void* Tdd, int n, const void* R, int nc, int q,
                      double* out3);

/*
 * hbem_rsvd_sv on device-resident inputs without any device allocation.  No
 * counterpart in the reference: it is for callers that cannot allocate (code
 * under HIP graph capture, tight serving loops).  T_dev (n x n, column-major,
 * not modified), R_dev (n x nc) and sigma_dev (one double) are device
 * pointers of the current device; stream is a hipStream_t (NULL: the null
 * stream).  The work enqueued on the stream (a copy of T, LU, the randomized
 * iteration) uses a fixed device array of ~66 MiB per GPU, held until the
 * process ends, so n <= 2048 and nc <= 8; larger sizes return HBEM_EINVAL
 * before anything is enqueued.  The call returns without waiting; calls that
 * overlap on one GPU share that array and race.  The array's device address
 * is looked up by hbem_init(device) or by the first call on a device, so a
 * caller that captures graphs calls either one beforehand.
 */
int hbem_rsvd_sv_noalloc(const void* T_dev, int n, const void* R_dev, int nc, int q, double* sigma_dev, void* stream);

/*
 * Batched sigma_min(k) sweep -- the caller pattern of the reference drivers
 * (std::transform(par_unseq, ks) { gen_sol_op(k, c_o, c_i, T); sv(T, R, q); },
 * examples/roots_brent_square_rsvd.cpp:167-171).  Everything stays on the
 * device; only the nk results come back.  want_der = 1 assembles the
 * operators with their first k-derivatives, want_der = 2 with the first and
 * second (gen_sol_op_2nd_der's operator set); in both cases
 * dsigma/dk = Re(u^H T' v) is returned in dsigma (may be NULL when want_der = 0),
 * (u, v) the singular pair of the randomized subspace: the same as
 * hbem_rsvd_sv_der's once the subspace has converged, within
 * O((sigma - sigma_min) / gap) of it at small q.
 * R: n x nc test matrix (n = 2N), uploaded once per call.  The wavenumbers run
 * in batches sized to the free device memory; a sweep of >= 32 real positive
 * k assembles every batch with the same kernel, so each result is the same
 * bit for bit whatever the batch size (and in any other such sweep).
 */
int hbem_sigma_sweep(hbem_geometry_t g, const double* ks, int nk, double c_o, double c_i,
                     const void* R, int nc, int q, int want_der, double* sigma, double* dsigma);

/*
 * The Newton refinement's evaluations batched on the device: for every k,
 * gen_sol_op_2nd_der's T, T', T'' (gen_sol_op.cpp:154-196) and
 * randomized_svd::sv_der2 (randsvd.cpp:171-216) -- sigma, dsigma/dk and
 * d2sigma/dk2, the last from the reference's bordered 2n x 2n system built
 * from the singular pair of the randomized subspace (as hbem_rsvd_sv_der2).
 * The reference's Newton drivers make these calls one k at a time
 * (examples/roots_newton_square.cpp:155-172).  Nothing but the 3 nk results
 * crosses to the host.
 */
int hbem_sigma_sweep_der2(hbem_geometry_t g, const double* ks, int nk, double c_o, double c_i, const void* R, int nc,
                          int q, double* sigma, double* dsigma, double* d2sigma);

/*
 * hbem_sigma_sweep (sigma only) with the dense stage of randomized_svd::sv
 * (randsvd.cpp:99-146) in mixed precision: T is formed in FP64 and kept, its
 * FP32 copy is LU-factorised on the FP32 matrix cores, and every solve of the
 * randomized iteration is refined against T in FP64 -- residual in FP64,
 * correction from the FP32 factors in FP64 arithmetic: the last solve
 * `refine` times (0..20), the ones before it `refine_inner` times (-1: as the
 * last).  The refinement of the last solve certifies each k: when the error it
 * leaves (its last correction times the contraction rate observed between its
 * corrections) exceeds 1e-10 of the solution, the rate exceeds 1/2, or refine
 * = 0, the k is recomputed with the FP64 factorisation and fp64_used[i] (may
 * be NULL) is set to 1; so is it when the refinement of an earlier solve does
 * not contract (its last correction above half the one before it, or above
 * half the unrefined solution after one step).  The assembly stays FP64 (on
 * a context created with HBEM_FLAG_FAR_FIELD, with its lower-order rules on
 * separated pairs).  No counterpart in the reference, which factorises in
 * FP64 only.
 */
int hbem_sigma_sweep_mixed(hbem_geometry_t g, const double* ks, int nk, double c_o, double c_i, const void* R, int nc,
                           int q, int refine, int refine_inner, double* sigma, int* fp64_used);

/*
 * sigma_min(k) for many scatterers at once -- the reference drivers' loops
 * over instances and wavenumbers (examples/roots_brent_square_rsvd.cpp:167-171,
 * find_roots.cpp:1615-1700) as one batched computation.  geoms[i] (contexts
 * with the same quadrature order, projection and flags) is swept over its nk
 * wavenumbers ks[i nk .. i nk + nk - 1] with inner index c_i[i], outer c_o
 * and test matrix R[i] (2 N_i x nc, host); sigma[i nk + j] receives the
 * result.  The instances are grouped by their order 2 N_i rounded up to a
 * multiple of `bucket` (a multiple of 32; 0 = 256): every panel-pair class of
 * a group is one assembly launch, and the group's T matrices, padded with the
 * identity to the common order, are one batched LU and randomized SVD.  The
 * padding leaves every result equal bit for bit to hbem_sigma_sweep of the
 * instance alone (nk >= 32, whose sweep uses the same assembly kernel).
 */
int hbem_sigma_sweep_multi(const hbem_geometry_t* geoms, int ninst, const double* ks, int nk, double c_o,
                           const double* c_i, const void* const* R, int nc, int q, int bucket, double* sigma);

/* Phase timing (HIP events on the context's stream), for the benchmark's
 * roofline figures.  enable != 0 clears and starts recording; the events are
 * only read by query (no host synchronisation inside the timed calls).  query
 * returns the accumulated milliseconds and launch count of phase `which`
 * (0 = general-pair assembly kernel, 1 = adjacent-pair kernel, 2 =
 * coinciding-pair kernel, 3 = LU factorisation, 4 = T formation, 5 = the
 * randomized SVD stage after the LU, 6 = the LU's matrix-core launches, its
 * trailing and sub-panel updates) and the algorithmic flop count (slots 0-2:
 * the frozen FP64 model of the assembly; 3: 8 n^3 / 3 per matrix; 6: the 3M
 * products' 6 m n k per update and matrix, in the precision of the
 * factorisation; 0 where none is defined). */
int hbem_profile_enable(hbem_geometry_t g, int enable);
int hbem_profile_query(hbem_geometry_t g, int which, double* ms, long long* launches, double* flops);

/* Wavenumbers per batch that hbem_sigma_sweep uses for a closed curve of
 * npanels panels at Gauss order `order`, derivative order d and nc random
 * columns when avail_bytes of device memory are usable and nk wavenumbers are
 * requested (at most 256, batches balanced; nk >= 32 real positive k use the
 * T-difference layout, half the local-matrix bytes).  Host-only: no device
 * is used.
 * (No counterpart in the reference, whose sweep runs one k per thread.) */
int hbem_sweep_batch_size(int npanels, int order, int d, int nc, unsigned long long avail_bytes, int nk, int* B);

/* randomized_svd::benchmarking, get_lu_time / get_qr_time / get_sub_iter_time
 * / get_svd_time, reset_timer (randsvd.hpp:23-35, randsvd.cpp:27-47).  While
 * benchmarking is on, hbem_rsvd_sv / sv_der / sv_der2 record device events
 * around the LU, the first solve + orthonormalisation, the subspace
 * iterations and the last solve + SVD; hbem_rsvd_timers returns the
 * accumulated seconds of the four phases in that order. */
int hbem_rsvd_benchmarking(int yes);
int hbem_rsvd_timers(double* seconds4);
int hbem_rsvd_reset_timer(void);

#ifdef __cplusplus
}
#endif

#endif /* HBEM_H */
