// hbem_internal.hpp -- declarations shared by the host and kernel translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <string>
#include <vector>

#include "hbem_device.hpp"

namespace hbem {

// Node table record: r, r^2, the double-layer factors of the direct and the
// transposed block, tau_i.tau_j, |tau_i||tau_j|, then 1/r, (2/pi) ln r, 1/sqrt(r)
// (all three 0 at r = 0): the r-dependent factors every wavenumber slot needs,
// computed once per geometry instead of once per staging workgroup.
constexpr int kNodeStride = 9;

struct AsmParams {
    const double* xg;            // Gauss rule on [0,1]
    const double* wg;
    const double* xc;            // composite rule
    const double* wc;
    const int* ent;              // class entries (i, j, swap)
    const double* ngeo;          // node geometry [entry][nodes][kNodeStride]
    const unsigned char* flags;  // per entry skip flags
    const long long* pidx;       // triangular pair index per entry
    int nent;                    // entries of this class
    int nk;                      // wavenumber slots in the batch
    int nchunk, chunk;           // node chunks per entry, nodes per chunk
    const double* kap;           // kappa = k sqrt(c) per slot (real part)
    const double* kr;            // k per slot (real part)
    const double* kapi;          // imaginary parts, or null: every slot real positive
    const double* kri;
    const double* cv;            // c per slot
    long long npairs_total;      // N (N + 1) / 2
    cdbl* out;                   // nchunk == 1: local matrices [npairs_total][16 (D+1)][nk];
                                 // else partials [nent][nchunk][16 (D+1)][nk]
    // several scatterers in one launch (paired-lane kernel only): entry e
    // belongs to instance i = einst[e], whose nk slots are kap/kr/cv[i nk ..]
    // and whose node table igeo[i] holds entry e at position e - ibase[i];
    // pidx is then the pair's index in the concatenation of the instances'
    // pair lists.  einst null: one instance.
    const int* einst;
    const double* const* igeo;
    const long long* ibase;
    int ns, nc;                  // Gauss / composite rule sizes
    int qflags;                  // HBEM_FLAG_* quirk bits
    // paired-lane kernel only: the slots come in blocks of 32 (nk a multiple
    // of 64), block 2w holding wavenumbers 32w .. 32w + 31 with the inner
    // refraction index and block 2w + 1 the same ones with the outer; out
    // receives per wavenumber the differences T is formed from (K_i - K_o,
    // W_o - W_i, V_o - V_i) with nkt (the wavenumbers) in place of nk in the
    // layouts above; kt0: the first wavenumber of the launch's slots in those
    // layouts (a launch over a range of slot blocks)
    int tdiff, nkt, kt0;
};

// assembly (hbem_assembly.hip)
void launch_setup_gauss_table(const double* panels, const double* coefs, int npan, const double* xg, int ns,
                              double* gtab, hipStream_t st);
void launch_setup_general(const int* ent, int nent, const double* gtab, int ns, double* ngeo, unsigned char* flags,
                          hipStream_t st);
void launch_setup_singular(int cls, const int* ent, int nent, const double* panels, const double* coefs,
                           const double* xc, int nc, double* ngeo, unsigned char* flags, hipStream_t st);
void launch_assemble(int cls, int D, const AsmParams& P, int G, hipStream_t st);
void launch_chunk_sum(const cdbl* part, int nent, int nchunk, int nk, int nslot, const long long* pidx, long long npairs,
                      cdbl* L, hipStream_t st);
// L: local matrices [pair][slot][wavenumber] of nk wavenumber slots
void launch_gather_layers(const cdbl* L, int nk, int nslot, int q, int N, int D, int layers, cdbl* K, cdbl* W,
                          cdbl* V, hipStream_t st);
// T_t for output b is written at T[t] + b * ld^2 (t = 0..D) with leading
// dimension ld (0: 2N); qo null: L is in the tdiff layout (AsmParams) and
// output b reads wavenumber qi[b] of it; T32 (may be null): T_0 again in
// FP32, each entry rounded to nearest
void launch_form_T(const cdbl* L, int nk, int nslot, const int* qi, const int* qo, int nb, int N, int D,
                   const double* M, cdbl* T0, cdbl* T1, cdbl* T2, hipStream_t st, int ld = 0, cflt* T32 = nullptr);
// identity in rows and columns n..ld-1 of nb matrices of order ld
void launch_pad_identity(cdbl* T, int n, int ld, int nb, hipStream_t st);
// A_o = [[-K_o + M/2, V_o], [W_o, K_o^T + M/2]] of wavenumber slot o (2N x 2N)
void launch_form_Ao(const cdbl* L, int nk, int nslot, int o, int N, const double* M, cdbl* A, hipStream_t st);

// dense (hbem_dense.hip); all batched over `batch` matrices of order n with
// stride n*n (column-major, lda = n) unless noted.
// Optional timing of the LU's matrix-core launches (the trailing and
// sub-panel updates C -= A B): begin before and end after each launch, with
// the real FLOP its 3M complex product performs (three real m x n x k
// products, 6 m n k per matrix).
struct LaunchTimer {
    virtual void begin(hipStream_t st) = 0;
    virtual void end(hipStream_t st, double flops) = 0;
    virtual ~LaunchTimer() = default;
};
void lu_factor_batched(cdbl* A, int n, int batch, int* ipiv, int* perm, hipStream_t st, LaunchTimer* gemm = nullptr);
// X (n x nc, stride n*nc per matrix) <- op(A)^{-1} X with A = P^T L U from
// lu_factor_batched (factors of matrix b at LU + b * ms); adj = false solves
// A X = B, adj = true solves A^H X = B.  work: n * nc * batch.
void lu_solve_thin_batched(const cdbl* LU, long long ms, const int* perm, int n, int batch, cdbl* X, int nc, bool adj,
                           cdbl* work, hipStream_t st);
// The same blocked LU in single precision (the mixed-precision sweep), its
// thin solves (FP32 factors read, FP64 arithmetic), and the FP32 copy of T.
void lu_factor_batched_f32(cflt* A, int n, int batch, int* ipiv, int* perm, hipStream_t st,
                           LaunchTimer* gemm = nullptr);
void lu_solve_thin_batched_f32(const cflt* LU, long long ms, const int* perm, int n, int batch, cdbl* X, int nc,
                               bool adj, cdbl* work, hipStream_t st);
// R = Bv - op(T) Z (op = T or T^H) for n x nc blocks of a batch, T (n x n,
// stride ts) in FP64; part: residual_part_count(n, nc, batch) complex scratch
void residual_thin_batched(const cdbl* T, long long ts, const cdbl* Z, const cdbl* Bv, cdbl* R, int n, int nc,
                           int batch, bool adj, cdbl* part, hipStream_t st);
size_t residual_part_count(int n, int nc, int batch);
// Z += D (cnt entries per matrix); prev[b]: max|D|^2 of the previous step
// (first: taken from Z before the update); check 1: flag[b] = 1 unless the
// estimated contraction is <= 1/2; check 2: also unless the estimated
// remaining error <= tol max|Z|; check 0: no flag
void refine_update_batched(cdbl* Z, const cdbl* D, int cnt, int batch, double tol, int check, bool first, double* prev,
                           int* flag, hipStream_t st);
void fill_int(int* p, int count, int v, hipStream_t st);
void add_offset_ll(long long* p, long long count, long long off, hipStream_t st);
void copy_bytes(void* dst, const void* src, size_t nbytes, hipStream_t st);
// *flag = 1 if any of cnt doubles is not finite / a diagonal entry of A is 0
void flag_nonfinite(const double* x, long long cnt, int* flag, hipStream_t st);
void flag_zero_diag(const cdbl* A, int n, int* flag, hipStream_t st);
// Y = S X for a real sparse n x n S in compressed-row form (rowptr[n + 1],
// column indices ascending within a row); X, Y: n x ncols, column-major, ld n.
void csr_apply_batched(const int* rowptr, const int* colidx, const double* val, int n, const cdbl* X, cdbl* Y,
                       long long ncols, hipStream_t st);
// A sparse real matrix in compressed-row form on the device, and the product
// Y = S_b X_b for the n_pad x ncols blocks of a batch whose matrix b has its
// own S_b = desc[b] of order desc[b].n <= n_pad (rows beyond it are zero).
struct CsrDesc {
    const int* rp;
    const int* ci;
    const double* v;
    int n;
};
void csr_apply_multi(const CsrDesc* desc, int n_pad, const cdbl* X, cdbl* Y, int ncols, int batch, hipStream_t st);
// X_b (n_pad x nc) = the n_b x nc block at src + off[b] (column-major, ld
// n_b) padded with zero rows, n_b = nrow[b]
void replicate_padded(const cdbl* src, const long long* off, const int* nrow, int n_pad, int nc, int batch, cdbl* X,
                      hipStream_t st);
// dst = batch copies of the count entries at src, back to back
void replicate_batched(const cdbl* src, long long count, int batch, cdbl* dst, hipStream_t st);
void orthonormalize_thin_batched(cdbl* X, int n, int batch, int nc, hipStream_t st);
// sigma = 1 / sigma_max(B); u = left singular vector of B for sigma_max,
// y = right singular vector (nc) -- v = Q y is formed by the caller.
void sigma_from_thin_batched(const cdbl* B, const cdbl* Q, int n, int batch, int nc, double* sigma, cdbl* u, cdbl* v,
                             hipStream_t st);
// out[b] = Re(a_b^H T_b b_b) for T_b column-major n x n (stride tstride);
// part: scratch of kBilinearParts * batch doubles
constexpr int kBilinearParts = 8;
void bilinear_re_batched(const cdbl* T, long long tstride, const cdbl* a, const cdbl* bvec, int n, int batch,
                         double* out, double* part, hipStream_t st);
// y_b = T_b x_b (adj = false) or T_b^H x_b (adj = true), T_b n x n with stride tstride
void gemv_batched(const cdbl* T, long long tstride, const cdbl* x, cdbl* y, int n, int batch, bool adj, hipStream_t st);
// W = [[0, T], [T^H, 0]] - s I  (2n x 2n) for a single matrix
void wielandt_build(const cdbl* T, int n, double s, cdbl* W, hipStream_t st);
// last pivot of the column-pivoted Householder QR (xGEQP3 rule) of each m x m
// matrix of A (stride ms, overwritten) into last[b]; nrm2: 2 m batch doubles,
// slot2: 2 m batch ints of scratch
void colpiv_last_pivot_batched(cdbl* A, long long ms, int m, int batch, double* nrm2, int* slot2, int* last,
                               hipStream_t st);
// The bordered system of randomized_svd::sv_der2, batched over nb matrices of
// order n (systems of order 2n, vectors of 2n entries back to back):
// X_b = [u_b; v_b] / x_m with m_b the first index of the largest |x_l|;
// B_b = [[0, Tp_b], [Tp_b^H, 0]] - sigma_b I with column m_b set to -X_b;
// negate_pick: Y <- -Y, t_b = Y_b[m_b], then Y_b[m_b] = 0;
// rhs: out = P2 + 2 (P - t Ud); final: d2_b = -Re(Y2_b[m_b]).
void border_prep_batched(const cdbl* U, const cdbl* V, int n, int nb, int* mv, cdbl* X, hipStream_t st);
void border_build_batched(const cdbl* Tp, int n, int nb, const double* sigma, const int* mv, const cdbl* X, cdbl* B,
                          hipStream_t st);
void border_negate_pick_batched(cdbl* Y, int n, int nb, const int* mv, cdbl* t, hipStream_t st);
void border_rhs_batched(const cdbl* P, const cdbl* P2, const cdbl* Ud, const cdbl* t, int n, int nb, cdbl* out,
                        hipStream_t st);
void border_final_batched(const cdbl* Y2, int n, int nb, const int* mv, double* d2, hipStream_t st);
// dst[b dstride + doff + i] = src[b sstride + soff + i] for i < len, b < nb
void copy_blocks_batched(const cdbl* src, long long sstride, long long soff, cdbl* dst, long long dstride,
                         long long doff, int len, int nb, hipStream_t st);
// C = alpha * A * op(B) + beta * C   (complex, column-major, batched with strides)
void zgemm_batched(int m, int n, int k, cdbl alpha, const cdbl* A, int lda, long long sa, const cdbl* B, int ldb,
                   long long sb, bool transB, cdbl beta, cdbl* C, int ldc, long long sc, int batch, hipStream_t st);

// host setup (hbem_setup.cpp)
void gauss_rule(double a, double b, int n, double* x, double* w);
void composite_rule(int n, double* x, double* w);
void mass_matrix_host(const double* panels, int N, const double* coefs, bool lines, int order, double* M);
// Eigen-style LDLT with diagonal pivoting; F = P^T L sqrt(D) (n x n, column-major)
void ldlt_projection_factor(const double* A, int n, double* F, int* transpositions);
void invert_projection_factor(const double* F, const int* tr, int n, double* Finv);
// start/end points of every panel as BuilderData::is_adjacent reads them
// (operator()(-1 / +1), galerkin_builder.cpp:189-201): 4 doubles per panel
void panel_endpoints(const double* panels, const double* coefs, int N, double* ends);

// A disc holding each panel: cr[3i], cr[3i + 1] its centre (the panel's point
// at s = 1/2), cr[3i + 2] its radius (half the length of a line panel; for a
// curved panel the largest distance of 65 equally spaced points from the
// centre, plus 2%)
void panel_bounds(const double* panels, const double* coefs, int N, double* cr);

}  // namespace hbem
