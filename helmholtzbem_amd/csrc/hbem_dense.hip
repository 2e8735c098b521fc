// hbem_dense.hip -- dense complex FP64 kernels of the randomized-SVD stage.
//
// Restates the Eigen calls of randomized_svd::sv (randsvd.cpp:99-146) for the
// GPU, batched over independent matrices (one per wavenumber):
//   * PartialPivLU (randsvd.cpp:104): right-looking blocked LU with partial
//     pivoting (largest modulus, first index on ties) in 64-column steps of two
//     32-column sub-panels; panel factorisation by one workgroup per matrix
//     with the panel rows in registers, row interchanges, triangular solve of
//     the U12 block, trailing update by a complex GEMM on the FP64 matrix cores;
//   * lu.solve / lu.adjoint().solve on thin right-hand sides (n x nc), reading
//     the factors in place: blocked substitution, the 64 x 64 diagonal blocks
//     by one wavefront with cross-lane broadcasts, the off-diagonal updates by
//     the whole workgroup;
//   * the projection factor's thin products in compressed-row form;
//   * HouseholderQR's role (an orthonormal basis of the span): modified
//     Gram-Schmidt applied twice;
//   * JacobiSVD's role (largest singular value of the thin n x nc matrix):
//     Gram matrix plus a Hermitian Jacobi eigensolver on one lane.
#include "hbem_internal.hpp"

namespace hbem {

// Reciprocal and product with every multiply-add an explicit fma: the LU's
// arithmetic per matrix entry is then fixed by the source, whichever kernel
// instantiation (rows per thread, chunk width) performs it, so a matrix
// factorised inside a larger zero-padded one gives the same bits.
__device__ __forceinline__ cdbl crecip(cdbl d) {
    const double s = 1.0 / fma(d.x, d.x, d.y * d.y);
    return mk(d.x * s, -d.y * s);
}
__device__ __forceinline__ cdbl cmulx(cdbl a, cdbl b) { return mk(fma(a.x, b.x, -(a.y * b.y)), fma(a.x, b.y, a.y * b.x)); }
__device__ __forceinline__ void cfms(cdbl& acc, cdbl a, cdbl b) {   // acc -= a * b
    acc.x = fma(-a.x, b.x, fma(a.y, b.y, acc.x));
    acc.y = fma(-a.x, b.y, fma(-a.y, b.x, acc.y));
}
__device__ __forceinline__ void cfma(cdbl& acc, cdbl a, cdbl b) {   // acc += a * b
    acc.x = fma(a.x, b.x, fma(-a.y, b.y, acc.x));
    acc.y = fma(a.x, b.y, fma(a.y, b.x, acc.y));
}
__device__ __forceinline__ cdbl shfl_c(cdbl v, int src) {
    return mk(__shfl(v.x, src, 64), __shfl(v.y, src, 64));
}
// lane src's value in every lane, src the same for the whole wavefront:
// v_readlane through a scalar register, no LDS round trip
__device__ __forceinline__ double bcast_lane(double v, int src) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), src);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), src);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ cdbl bcast_lane(cdbl v, int src) { return mk(bcast_lane(v.x, src), bcast_lane(v.y, src)); }

// The same operations on single-precision complex values (the factors of the
// mixed-precision LU).
__device__ __forceinline__ float cabs2(cflt a) { return a.x * a.x + a.y * a.y; }
__device__ __forceinline__ cflt operator*(cflt a, cflt b) { return cflt{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x}; }
__device__ __forceinline__ cflt crecip(cflt d) {
    const float s = 1.0f / fmaf(d.x, d.x, d.y * d.y);
    return cflt{d.x * s, -d.y * s};
}
__device__ __forceinline__ cflt cmulx(cflt a, cflt b) {
    return cflt{fmaf(a.x, b.x, -(a.y * b.y)), fmaf(a.x, b.y, a.y * b.x)};
}
__device__ __forceinline__ void cfms(cflt& acc, cflt a, cflt b) {   // acc -= a * b
    acc.x = fmaf(-a.x, b.x, fmaf(a.y, b.y, acc.x));
    acc.y = fmaf(-a.x, b.y, fmaf(-a.y, b.x, acc.y));
}
__device__ __forceinline__ cflt shfl_c(cflt v, int src) { return cflt{__shfl(v.x, src, 64), __shfl(v.y, src, 64)}; }
// widening load of a factor entry for double-precision arithmetic
__device__ __forceinline__ cdbl to_cdbl(cdbl v) { return v; }
__device__ __forceinline__ cdbl to_cdbl(cflt v) { return mk((double)v.x, (double)v.y); }

// ---------------------------------------------------------------------------
// LU with partial pivoting
// ---------------------------------------------------------------------------
// Panel factorisation of columns [j0, j0 + jb) (jb <= 32), one workgroup of
// 1024 threads per matrix: per column the pivot (largest |a|^2, first index on
// ties) by wavefront shuffles and one LDS exchange, the row interchange inside
// the panel, then the rank-1 update of the panel with one thread per row and
// the pivot row staged in LDS.
template <class C>
__global__ __launch_bounds__(1024) void k_lu_panel(C* __restrict__ A, int n, int j0, int jb, int* __restrict__ ipiv) {
    C* a = A + (size_t)blockIdx.x * n * n;
    int* ip = ipiv + (size_t)blockIdx.x * n;
    __shared__ double sv[16];
    __shared__ int si[16];
    __shared__ int spiv;
    __shared__ C prow[32];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    for (int c = 0; c < jb; ++c) {
        const int col = j0 + c;
        double best = -1.0;
        int bi = n;
        for (int i = col + tid; i < n; i += blockDim.x) {
            const double v = cabs2(a[i + (size_t)col * n]);
            if (v > best) { best = v; bi = i; }
        }
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
            const double v2 = __shfl_xor(best, o, 64);
            const int i2 = __shfl_xor(bi, o, 64);
            if (v2 > best || (v2 == best && i2 < bi)) { best = v2; bi = i2; }
        }
        if (lane == 0) { sv[wv] = best; si[wv] = bi; }
        __syncthreads();
        if (tid < 64) {
            const int nw = blockDim.x >> 6;
            best = lane < nw ? sv[lane] : -1.0;
            bi = lane < nw ? si[lane] : n;
#pragma unroll
            for (int o = 8; o > 0; o >>= 1) {
                const double v2 = __shfl_xor(best, o, 64);
                const int i2 = __shfl_xor(bi, o, 64);
                if (v2 > best || (v2 == best && i2 < bi)) { best = v2; bi = i2; }
            }
            // no candidate (a column of NaNs): keep the row, never index past n
            if (tid == 0) { const int pv = bi < n ? bi : col; spiv = pv; ip[col] = pv; }
        }
        __syncthreads();
        const int piv = spiv;
        if (tid < jb) {
            C* pc = a + (size_t)(j0 + tid) * n;
            const C u = pc[piv];
            if (piv != col) { pc[piv] = pc[col]; pc[col] = u; }
            prow[tid] = u;
        }
        __syncthreads();
        const C d = prow[c];
        const C inv = (d.x == 0.0 && d.y == 0.0) ? C{} : crecip(d);
        for (int i = col + 1 + tid; i < n; i += blockDim.x) {
            const C l = cmulx(a[i + (size_t)col * n], inv);
            a[i + (size_t)col * n] = l;
            // eight columns per round, loads issued together
            for (int cc0 = c + 1; cc0 < jb; cc0 += 8) {
                C v[8];
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    if (cc0 + u < jb) v[u] = a[i + (size_t)(j0 + cc0 + u) * n];
#pragma unroll
                for (int u = 0; u < 8; ++u)
                    if (cc0 + u < jb) {
                        cfms(v[u], l, prow[cc0 + u]);
                        a[i + (size_t)(j0 + cc0 + u) * n] = v[u];
                    }
            }
        }
        __syncthreads();
    }
}

// The interchanges of the panel [j0, j0 + jb) as a map over the rows they
// touch, by one wavefront (lane k): the distinct pivot rows below the panel are
// numbered in order of first appearance (erow), and the jb swaps are replayed
// on the slot -> source index with wave-uniform readlanes, slot t < jb being
// panel row j0 + t and slot jb + e pivot row erow[e]; src[t] is the slot whose
// row the swaps move into slot t, nes the number of pivot rows below.
__device__ __forceinline__ void row_map(const int* ip, int j0, int jb, int* src, int* erow, int* nes) {
    const int k = threadIdx.x & 63;
    const int p = k < jb ? ip[j0 + k] : 0;
    const bool ex = k < jb && p >= j0 + jb;
    int first = k;
    if (ex)
        for (int k2 = 0; k2 < k; ++k2)
            if (ip[j0 + k2] == p) { first = k2; break; }
    const unsigned long long firsts = __ballot(ex && first == k);
    const int e = __popcll(firsts & ((1ull << first) - 1ull));
    if (ex && first == k) erow[e] = p;
    const int partner = k < jb ? (ex ? jb + e : p - j0) : k;   // slot swapped with slot k
    int v = k;
    for (int sw = 0; sw < jb; ++sw) {
        const int b = __builtin_amdgcn_readlane(partner, sw);
        if (b != sw) {
            const int va = __builtin_amdgcn_readlane(v, sw);
            const int vb = __builtin_amdgcn_readlane(v, b);
            v = k == sw ? vb : (k == b ? va : v);
        }
    }
    src[k] = v;
    if (k == 0) *nes = __popcll(firsts);
}

// Map layout in memory (the caller's perm array doubles as its store while
// the factorisation runs): src[64], erow[32], nes.
constexpr int kRowMapInts = 97;

// The same panel factorisation with the rows held in registers, CW panel
// columns (a chunk) at a time, for panels of at most 1024 R rows: thread t owns
// rows j0 + t + 1024 u, u < R.  Inside a chunk every column step -- pivot
// search, exchange of the two rows' chunk values through LDS, scaling, rank-1
// update -- works on registers; the interchange of the other panel columns
// goes to memory, one thread per column.  After the chunk, its rows are
// stored, U12 = L11^{-1} A12 is solved for the panel columns right of it (one
// thread per column) and those columns receive A22 -= L21 U12 in one pass.
// The panel is read and written about jb / CW + 1 times instead of once per
// column step.  R x CW complex values stay in registers (at most 32 of the
// 128 VGPRs a thread of a 1024-thread workgroup has).
template <class C, int R, int CW>
__global__ __launch_bounds__(1024) void k_lu_panel_rg(C* __restrict__ A, int n, int j0, int jb,
                                                      int* __restrict__ ipiv, int* __restrict__ rmap) {
    constexpr int YB = R >= 2 ? 4 : 8;           // trailing columns updated per round
    C* a = A + (size_t)blockIdx.x * n * n;
    int* ip = ipiv + (size_t)blockIdx.x * n;
    __shared__ double sv[16];
    __shared__ int si[16];
    __shared__ int spiv;
    __shared__ C prow[CW];                   // pivot row (chunk columns)
    __shared__ C xrow[CW];                   // the row it displaces
    __shared__ C L11[CW][CW];
    __shared__ C U12[CW][32];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    int rowu[R];
#pragma unroll
    for (int u = 0; u < R; ++u) rowu[u] = j0 + tid + 1024 * u;
    for (int c0 = 0; c0 < jb; c0 += CW) {
        const int w = min(CW, jb - c0);
        const int cb = j0 + c0;                  // first column of the chunk = its first diagonal row
        C v[R][CW];
#pragma unroll
        for (int u = 0; u < R; ++u) {
            const bool own = rowu[u] < n && rowu[u] >= cb;
#pragma unroll
            for (int e = 0; e < CW; ++e) v[u][e] = (own && e < w) ? a[rowu[u] + (size_t)(cb + e) * n] : C{};
        }
#pragma unroll
        for (int c = 0; c < CW; ++c) {
            if (c >= w) break;
            const int col = cb + c;
            double best = -1.0;
            int bi = n;
            // rows ascending within a thread: strict > keeps the first index on ties
#pragma unroll
            for (int u = 0; u < R; ++u)
                if (rowu[u] < n && rowu[u] >= col) {
                    const double x = cabs2(v[u][c]);
                    if (x > best) { best = x; bi = rowu[u]; }
                }
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                const double v2 = __shfl_xor(best, o, 64);
                const int i2 = __shfl_xor(bi, o, 64);
                if (v2 > best || (v2 == best && i2 < bi)) { best = v2; bi = i2; }
            }
            if (lane == 0) { sv[wv] = best; si[wv] = bi; }
            __syncthreads();
            if (tid < 64) {
                best = lane < 16 ? sv[lane] : -1.0;
                bi = lane < 16 ? si[lane] : n;
#pragma unroll
                for (int o = 8; o > 0; o >>= 1) {
                    const double v2 = __shfl_xor(best, o, 64);
                    const int i2 = __shfl_xor(bi, o, 64);
                    if (v2 > best || (v2 == best && i2 < bi)) { best = v2; bi = i2; }
                }
                // no candidate (a column of NaNs): keep the row, never index past n
                if (tid == 0) { const int pv = bi < n ? bi : col; spiv = pv; ip[col] = pv; }
            }
            __syncthreads();
            const int piv = spiv;
#pragma unroll
            for (int u = 0; u < R; ++u) {
                if (rowu[u] == piv) {
#pragma unroll
                    for (int e = 0; e < CW; ++e) prow[e] = v[u][e];
                }
                if (rowu[u] == col) {
#pragma unroll
                    for (int e = 0; e < CW; ++e) xrow[e] = v[u][e];
                }
            }
            if (piv != col && tid < jb && (tid < c0 || tid >= c0 + w)) {
                C* pc = a + (size_t)(j0 + tid) * n;
                const C t = pc[piv];
                pc[piv] = pc[col];
                pc[col] = t;
            }
            __syncthreads();
            if (piv != col) {
#pragma unroll
                for (int u = 0; u < R; ++u) {
                    if (rowu[u] == piv) {
#pragma unroll
                        for (int e = 0; e < CW; ++e) v[u][e] = xrow[e];
                    } else if (rowu[u] == col) {
#pragma unroll
                        for (int e = 0; e < CW; ++e) v[u][e] = prow[e];
                    }
                }
            }
            const C d = prow[c];
            const C inv = (d.x == 0.0 && d.y == 0.0) ? C{} : crecip(d);
#pragma unroll
            for (int u = 0; u < R; ++u)
                if (rowu[u] < n && rowu[u] > col) {
                    const C l = cmulx(v[u][c], inv);
                    v[u][c] = l;
#pragma unroll
                    for (int e = c + 1; e < CW; ++e)
                        if (e < w) cfms(v[u][e], l, prow[e]);
                }
        }
        // chunk rows back to memory; L11 of the chunk for the U12 solve
#pragma unroll
        for (int u = 0; u < R; ++u) {
            if (rowu[u] < n && rowu[u] >= cb) {
#pragma unroll
                for (int e = 0; e < CW; ++e)
                    if (e < w) a[rowu[u] + (size_t)(cb + e) * n] = v[u][e];
                const int r = rowu[u] - cb;
                if (r < w) {
#pragma unroll
                    for (int e = 0; e < CW; ++e) L11[r][e] = v[u][e];
                }
            }
        }
        const int rem = jb - c0 - w;             // panel columns right of the chunk
        if (rem <= 0) break;
        __syncthreads();
        // (loads below take clamped, in-range addresses so that they issue
        // together; what lies outside the range is never used or stored)
        if (tid < rem) {
            C* pc = a + (size_t)(cb + w + tid) * n + cb;
            C x[CW];
#pragma unroll
            for (int r = 0; r < CW; ++r) x[r] = pc[min(r, w - 1)];
#pragma unroll
            for (int r = 0; r < CW; ++r) {
                if (r < w) {
#pragma unroll
                    for (int e = 0; e < CW; ++e)
                        if (e < r) cfms(x[r], L11[r][e], x[e]);
                    pc[r] = x[r];
                    U12[r][tid] = x[r];
                }
            }
        }
        __syncthreads();
#pragma unroll
        for (int u = 0; u < R; ++u) {
            const int row = rowu[u];
            if (row >= n) break;                 // rows ascend with u
            const bool upd = row >= cb + w;
            for (int cc0 = 0; cc0 < rem; cc0 += YB) {
                C y[YB];
#pragma unroll
                for (int e = 0; e < YB; ++e) y[e] = a[row + (size_t)(cb + w + min(cc0 + e, rem - 1)) * n];
#pragma unroll
                for (int e = 0; e < YB; ++e) {
                    if (cc0 + e < rem) {
#pragma unroll
                        for (int f = 0; f < CW; ++f)
                            if (f < w) cfms(y[e], v[u][f], U12[f][cc0 + e]);
                        if (upd) a[row + (size_t)(cb + w + cc0 + e) * n] = y[e];
                    }
                }
            }
        }
        __syncthreads();
    }
    // the interchange map for the row kernels (the pivots are in ip; every
    // thread has passed a barrier since the last was written)
    if (rmap) {
        __syncthreads();
        if (wv == 0) {
            int* m = rmap + (size_t)blockIdx.x * n;
            row_map(ip, j0, jb, m, m + 64, m + 96);
        }
    }
}

// Row interchanges of the panel [j0, j0 + jb) (jb <= 32) and the U12 solve.
// MODE bit 0: apply the panel's interchanges to every column outside the panel;
// MODE bit 1: for the columns right of the panel, U12 <- L11^{-1} A12 (L11 unit lower).
// A workgroup takes 32 columns of one matrix (50 KB of LDS: three per CU).
//   * Interchanges: wavefront 0 numbers the distinct pivot rows below the panel
//     in order of first appearance (ballot) and replays the jb swaps on the
//     slot -> source index, one slot per lane, with wave-uniform readlanes; the
//     rows involved (panel rows, then those pivot rows) are staged in LDS in
//     their permuted order and written back.
//   * Solve: a wavefront takes 8 columns, 8 lanes per column holding 4 rows
//     each in registers; step k broadcasts row k within its column by a lane
//     permute and updates the rows below it with L11 read from LDS.
template <class C, int MODE>
__global__ __launch_bounds__(256) void k_lu_rows(C* __restrict__ A, int n, int j0, int jb,
                                                 const int* __restrict__ ipiv, const int* __restrict__ rmap) {
    C* a = A + (size_t)blockIdx.y * n * n;
    const int* ip = ipiv + (size_t)blockIdx.y * n;
    __shared__ C rows[64][33];   // slot t < jb: panel row j0 + t; slot jb + e: pivot row erow[e]
    __shared__ C L[32][33];
    __shared__ int erow[32];
    __shared__ int src[64];
    __shared__ int nes;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int nout = n - jb;
    const int c0 = blockIdx.x * 32;
    const int ncol = min(32, nout - c0);
    if (MODE == 2 && c0 + ncol <= j0) return;             // only columns left of the panel: nothing to solve
    // column index of local column cc (skipping the panel)
    auto colof = [&](int cc) { const int q = c0 + cc; return q < j0 ? q : q + jb; };
    auto touched = [&](int cc) { return cc < ncol && (MODE != 2 || colof(cc) >= j0 + jb); };
    int ne = 0;
    if (MODE & 1) {
        if (rmap) {                               // left by the panel kernel
            const int* m = rmap + (size_t)blockIdx.y * n;
            if (tid < 64) src[tid] = m[tid];
            else if (tid < 96) erow[tid - 64] = m[tid];
            else if (tid == 96) nes = m[96];
        } else if (wv == 0) {
            row_map(ip, j0, jb, src, erow, &nes);
        }
        __syncthreads();
        ne = nes;
    }
    const int nrow = jb + ne;
    if (MODE & 2)
        for (int e2 = tid; e2 < jb * jb; e2 += blockDim.x) {
            const int r = e2 % jb, c = e2 / jb;
            L[r][c] = a[(j0 + r) + (size_t)(j0 + c) * n];
        }
    // staged already permuted: slot t receives the row that the swaps move into it
    auto row_of = [&](int t) { return t < jb ? j0 + t : erow[t - jb]; };
    for (int e2 = tid; e2 < nrow * 32; e2 += blockDim.x) {
        const int t = e2 % nrow, cc = e2 / nrow;
        if (touched(cc)) rows[t][cc] = a[row_of((MODE & 1) ? src[t] : t) + (size_t)colof(cc) * n];
    }
    __syncthreads();
    if (MODE & 2) {
        const int c8 = lane & 7, g = lane >> 3;
        const int cc = 8 * wv + c8;
        const bool act = cc < ncol && colof(cc) >= j0 + jb;
        C x[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int i = g + 8 * u;
            x[u] = (act && i < jb) ? rows[i][cc] : C{};
        }
        // row k = 8 uk + kk sits in register uk of lane kk * 8 + c8
#pragma unroll
        for (int uk = 0; uk < 4; ++uk) {
            for (int kk = 0; kk < 8; ++kk) {
                const int k = 8 * uk + kk;
                if (k + 1 >= jb) break;
                const C xk = shfl_c(x[uk], (kk << 3) | c8);
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    const int i = g + 8 * u;
                    if (i > k && i < jb) cfms(x[u], L[i][k], xk);
                }
            }
        }
        if (act) {
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const int i = g + 8 * u;
                if (i < jb) rows[i][cc] = x[u];
            }
        }
        __syncthreads();
    }
    for (int e2 = tid; e2 < nrow * 32; e2 += blockDim.x) {
        const int t = e2 % nrow, cc = e2 / nrow;
        if (touched(cc)) a[row_of(t) + (size_t)colof(cc) * n] = rows[t][cc];
    }
}

// final row permutation: (P A)(i, :) = A(perm[i], :)
__global__ void k_lu_perm(int n, const int* __restrict__ ipiv, int* __restrict__ perm) {
    if (threadIdx.x != 0) return;
    const int* ip = ipiv + (size_t)blockIdx.x * n;
    int* p = perm + (size_t)blockIdx.x * n;
    for (int i = 0; i < n; ++i) p[i] = i;
    for (int k = 0; k < n; ++k) {
        int j = ip[k];
        if (j != k) { int t = p[k]; p[k] = p[j]; p[j] = t; }
    }
}

// ---------------------------------------------------------------------------
// Complex GEMM on the matrix cores: C = alpha A op(B) + beta C
// ---------------------------------------------------------------------------
// v_mfma_f64_16x16x4f64 multiplies a 16 x 4 by a 4 x 16 FP64 block: lane l
// supplies row (l & 15), column (l >> 4) of the first operand and row (l >> 4),
// column (l & 15) of the second, and holds D(4 v + (l >> 4), l & 15) in element
// v of the accumulator.  The complex product takes three real ones (the 3M
// form): P1 = Re A Re B, P2 = Im A Im B, P3 = (Re A + Im A)(Re B + Im B),
// accumulated over k, then Re C = P1 - P2, Im C = P3 - P1 - P2 -- normwise as
// accurate as the four-product form, with three independent accumulators per
// block.  The product is formed transposed -- first operand B^T, second A^T --
// so element v of lane l is C(r0 + (l & 15), c0 + 4 v + (l >> 4)): sixteen
// consecutive rows of one column per lane group, for coalesced stores.
// Workgroup = 64 x 64 tile of C as four wavefronts of 32 x 32 (2 x 2 blocks of
// 16 x 16); k advances in steps of 16 through LDS while the next step's tile
// is fetched into registers.  Workgroups are dealt to the eight XCDs in turn
// (id % 8); the grid is renumbered so that each XCD gets a contiguous run of
// (matrix, tile) indices, i.e. whole matrices, whose A and B panels then stay
// in that XCD's L2 instead of being fetched by all eight.
typedef double dbl4 __attribute__((ext_vector_type(4)));

// Tile (row block tr, column block tc) of tile index t in a tm x tn grid of
// one matrix, numbered in groups of kTileGroup row blocks: consecutive indices
// run down a group's rows before moving one column to the right.  The ~64
// workgroups an XCD holds at a time then cover an 8 x 8 square of tiles and
// share its 8 row blocks of A and 8 column blocks of B in that XCD's L2 (down
// whole columns of tiles they shared only the B block, and A's panel was read
// again for every column).  A tile's arithmetic does not depend on its number.
constexpr int kTileGroup = 8;
__device__ __forceinline__ void grouped_tile(int t, int tm, int tn, int& tr, int& tc) {
    const int per_group = kTileGroup * tn;
    const int g = t / per_group, first = g * kTileGroup;
    const int rows = min(kTileGroup, tm - first);
    const int r = t - g * per_group;
    tr = first + r % rows;
    tc = r / rows;
}

template <bool TB>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2, 2))) void k_zgemm(int m, int n, int k, cdbl alpha, const cdbl* __restrict__ A, int lda,
                                               long long sa, const cdbl* __restrict__ B, int ldb, long long sb,
                                               cdbl beta, cdbl* __restrict__ C, int ldc, long long sc, int batch) {
    const int tm = (m + 63) >> 6, tn = (n + 63) >> 6;
    const long long total = (long long)tm * tn * batch;
    const long long per = (total + 7) >> 3;
    const long long li = (long long)(blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (li >= total) return;                 // whole workgroup, before any barrier
    const int bmat = (int)(li / ((long long)tm * tn));
    const int tile = (int)(li - (long long)bmat * tm * tn);
    A += bmat * sa;
    B += bmat * sb;
    C += bmat * sc;
    __shared__ cdbl As[16][65];
    __shared__ cdbl Bs[16][65];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int l16 = lane & 15, lk = lane >> 4;
    const int wr = (wv & 1) * 32, wc = (wv >> 1) * 32;
    int tr, tc;
    grouped_tile(tile, tm, tn, tr, tc);
    const int row0 = tr * 64, col0 = tc * 64;
    dbl4 p1[2][2], p2[2][2], p3[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            p1[i][j] = dbl4{0.0, 0.0, 0.0, 0.0};
            p2[i][j] = dbl4{0.0, 0.0, 0.0, 0.0};
            p3[i][j] = dbl4{0.0, 0.0, 0.0, 0.0};
        }
    cdbl pa[4], pb[4];
    auto fetch = [&](int kb) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int e = tid + 256 * u, r = e & 63, kk = e >> 6;
            pa[u] = (row0 + r < m && kb + kk < k) ? A[(row0 + r) + (size_t)(kb + kk) * lda] : mk(0, 0);
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int e = tid + 256 * u;
            const int kk = TB ? (e >> 6) : (e & 15), cc = TB ? (e & 63) : (e >> 4);
            cdbl v = mk(0, 0);
            if (col0 + cc < n && kb + kk < k)
                v = TB ? B[(col0 + cc) + (size_t)(kb + kk) * ldb] : B[(kb + kk) + (size_t)(col0 + cc) * ldb];
            pb[u] = v;
        }
    };
    fetch(0);
    for (int kb = 0; kb < k; kb += 16) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int e = tid + 256 * u;
            As[e >> 6][e & 63] = pa[u];
            const int kk = TB ? (e >> 6) : (e & 15), cc = TB ? (e & 63) : (e >> 4);
            Bs[kk][cc] = pb[u];
        }
        __syncthreads();
        if (kb + 16 < k) fetch(kb + 16);
#pragma unroll
        for (int k4 = 0; k4 < 16; k4 += 4) {
            cdbl a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) a[i] = As[k4 + lk][wr + 16 * i + l16];
#pragma unroll
            for (int j = 0; j < 2; ++j) b[j] = Bs[k4 + lk][wc + 16 * j + l16];
            const double as[2] = {a[0].x + a[0].y, a[1].x + a[1].y};
            const double bs[2] = {b[0].x + b[0].y, b[1].x + b[1].y};
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    p1[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[j].x, a[i].x, p1[i][j], 0, 0, 0);
                    p2[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(b[j].y, a[i].y, p2[i][j], 0, 0, 0);
                    p3[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(bs[j], as[i], p3[i][j], 0, 0, 0);
                }
        }
        __syncthreads();
    }
    // Epilogue per row block i: the eight C values (beta != 0) are loaded
    // together from clamped (in-range) addresses, then combined and stored.
    const bool rd = beta.x != 0.0 || beta.y != 0.0;
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int r = row0 + wr + 16 * i + l16, rc = min(r, m - 1);
        cdbl cv[2][4];
        if (rd) {
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int c = min(col0 + wc + 16 * j + 4 * v + lk, n - 1);
                    cv[j][v] = C[rc + (size_t)c * ldc];
                }
        }
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int c = col0 + wc + 16 * j + 4 * v + lk;
                if (r < m && c < n) {
                    const double q1 = p1[i][j][v], q2 = p2[i][j][v];
                    cdbl x = alpha * mk(q1 - q2, p3[i][j][v] - q1 - q2);
                    if (rd) x = x + beta * cv[j][v];
                    C[r + (size_t)c * ldc] = x;
                }
            }
    }
}

void zgemm_batched(int m, int n, int k, cdbl alpha, const cdbl* A, int lda, long long sa, const cdbl* B, int ldb,
                   long long sb, bool transB, cdbl beta, cdbl* C, int ldc, long long sc, int batch, hipStream_t st) {
    if (m <= 0 || n <= 0 || batch <= 0) return;
    const long long total = (long long)((m + 63) / 64) * ((n + 63) / 64) * batch;
    const dim3 grid((unsigned)(8 * ((total + 7) / 8)));
    if (transB)
        hipLaunchKernelGGL(k_zgemm<true>, grid, dim3(256), 0, st, m, n, k, alpha, A, lda, sa, B, ldb, sb, beta, C, ldc, sc,
                           batch);
    else
        hipLaunchKernelGGL(k_zgemm<false>, grid, dim3(256), 0, st, m, n, k, alpha, A, lda, sa, B, ldb, sb, beta, C, ldc, sc,
                           batch);
}

// C -= A B on the FP32 matrix cores for the mixed-precision LU (single-
// precision complex, column-major, all three with leading dimension ld and
// batch stride sm): the tiling of k_zgemm -- 64 x 64 tiles of four
// wavefronts of 32 x 32, k through LDS in steps of 16 with the next step
// prefetched, the 3M complex product formed transposed, XCD-grouped
// numbering -- with v_mfma_f32_16x16x4f32, FP32 operands and FP32
// accumulation, at four wavefronts per SIMD (108 VGPRs; against two, the
// polygon2000 FP32 LU took 241 instead of 258 ms, same bits).  Element v of lane l of that instruction's result is
// D(4 (l >> 4) + v, l & 15) (the transposed product's row, i.e. C's column).
typedef float flt4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(4, 4))) void k_cgemm32_sub(
    int m, int n, int k, const cflt* __restrict__ A, const cflt* __restrict__ B, cflt* __restrict__ C, int ld,
    long long sm, int batch) {
    const int tm = (m + 63) >> 6, tn = (n + 63) >> 6;
    const long long total = (long long)tm * tn * batch;
    const long long per = (total + 7) >> 3;
    const long long li = (long long)(blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (li >= total) return;                 // whole workgroup, before any barrier
    const int bmat = (int)(li / ((long long)tm * tn));
    const int tile = (int)(li - (long long)bmat * tm * tn);
    A += bmat * sm;
    B += bmat * sm;
    C += bmat * sm;
    __shared__ cflt As[16][65];
    __shared__ cflt Bs[16][65];
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int l16 = lane & 15, lk = lane >> 4;
    const int wr = (wv & 1) * 32, wc = (wv >> 1) * 32;
    int tr, tc;
    grouped_tile(tile, tm, tn, tr, tc);
    const int row0 = tr * 64, col0 = tc * 64;
    flt4 p1[2][2], p2[2][2], p3[2][2];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            p1[i][j] = flt4{0.f, 0.f, 0.f, 0.f};
            p2[i][j] = flt4{0.f, 0.f, 0.f, 0.f};
            p3[i][j] = flt4{0.f, 0.f, 0.f, 0.f};
        }
    cflt pa[4], pb[4];
    auto fetch = [&](int kb) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int e = tid + 256 * u, r = e & 63, kk = e >> 6;
            pa[u] = (row0 + r < m && kb + kk < k) ? A[(row0 + r) + (size_t)(kb + kk) * ld] : cflt{0.f, 0.f};
        }
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int e = tid + 256 * u, kk = e & 15, cc = e >> 4;
            pb[u] = (col0 + cc < n && kb + kk < k) ? B[(kb + kk) + (size_t)(col0 + cc) * ld] : cflt{0.f, 0.f};
        }
    };
    fetch(0);
    for (int kb = 0; kb < k; kb += 16) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int e = tid + 256 * u;
            As[e >> 6][e & 63] = pa[u];
            Bs[e & 15][e >> 4] = pb[u];
        }
        __syncthreads();
        if (kb + 16 < k) fetch(kb + 16);
#pragma unroll
        for (int k4 = 0; k4 < 16; k4 += 4) {
            cflt a[2], b[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) a[i] = As[k4 + lk][wr + 16 * i + l16];
#pragma unroll
            for (int j = 0; j < 2; ++j) b[j] = Bs[k4 + lk][wc + 16 * j + l16];
            const float as[2] = {a[0].x + a[0].y, a[1].x + a[1].y};
            const float bs[2] = {b[0].x + b[0].y, b[1].x + b[1].y};
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    p1[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(b[j].x, a[i].x, p1[i][j], 0, 0, 0);
                    p2[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(b[j].y, a[i].y, p2[i][j], 0, 0, 0);
                    p3[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(bs[j], as[i], p3[i][j], 0, 0, 0);
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
        const int r = row0 + wr + 16 * i + l16, rc = min(r, m - 1);
        cflt cv[2][4];
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int c = min(col0 + wc + 16 * j + 4 * lk + v, n - 1);
                cv[j][v] = C[rc + (size_t)c * ld];
            }
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int c = col0 + wc + 16 * j + 4 * lk + v;
                if (r < m && c < n) {
                    const float q1 = p1[i][j][v], q2 = p2[i][j][v];
                    C[r + (size_t)c * ld] = cflt{cv[j][v].x - (q1 - q2), cv[j][v].y - (p3[i][j][v] - q1 - q2)};
                }
            }
    }
}

// C -= A B for the blocked LU: FP64 on k_zgemm, FP32 on k_cgemm32_sub
static void gemm_sub(int m, int n, int k, const cdbl* A, const cdbl* B, cdbl* C, int ld, long long s, int batch,
                     hipStream_t st) {
    zgemm_batched(m, n, k, mk(-1, 0), A, ld, s, B, ld, s, false, mk(1, 0), C, ld, s, batch, st);
}
static void gemm_sub(int m, int n, int k, const cflt* A, const cflt* B, cflt* C, int ld, long long s, int batch,
                     hipStream_t st) {
    if (m <= 0 || n <= 0 || batch <= 0) return;
    const long long total = (long long)((m + 63) / 64) * ((n + 63) / 64) * batch;
    hipLaunchKernelGGL(k_cgemm32_sub, dim3((unsigned)(8 * ((total + 7) / 8))), dim3(256), 0, st, m, n, k, A, B, C, ld, s,
                       batch);
}

// Blocked right-looking LU in steps of S 32-column sub-panels (S = 2; above
// order 2048, 8 for FP64 factors and 4 for FP32 ones: at order 4000 the
// FP64 LU took 391 ms at S = 8 against 412 at S = 4 for 128 matrices, while
// the FP32 LU at S = 8 left the mixed-precision sigma ten times farther from
// FP64 -- 2e-9 against 2e-10 -- for 14 of its 230 ms).  Within a step,
// sub-panel s is first brought up to date with the step's earlier sub-panels
// (k = 32 s), factorised, its row interchanges applied to every other column,
// and its U rows right of it formed; the trailing matrix then receives one
// k = 32 S update per step.
// That update reads and writes the trailing matrix once per step -- the
// traffic that bounds it at these sizes -- so wider steps cut it, at the cost
// of more small sub-panel products:
//   s = 0:  panel; interchanges everywhere + U0 for every column right of it
//   s > 0:  A[js:, js:js+32] -= L[js:, j0:js] U[j0:js, js:js+32]; panel;
//           interchanges everywhere; A[js:js+32, js+32:] -= L[js:js+32, j0:js] U[j0:js, js+32:];
//           Us = Lss^{-1} A[js:js+32, js+32:]
//   then:   A[jE:, jE:] -= L[jE:, j0:jE] U[j0:jE, jE:]
// C = cdbl: the FP64 factorisation; C = cflt: the same steps in FP32.  The
// step width depends on n and the precision only, so a matrix gives the same
// bits in any batch.
template <class C>
static void lu_factor_impl(C* A, int n, int batch, int* ipiv, int* perm, hipStream_t st, LaunchTimer* tm) {
    const int sb = 32;
    const int S = n > 2048 ? (sizeof(C) == sizeof(cdbl) ? 8 : 4) : 2;
    const long long s = (long long)n * n;
    auto at = [&](int r, int c) { return A + r + (size_t)c * n; };
    auto update = [&](int m, int nn, int k, const C* a, const C* b, C* c) {
        if (tm) tm->begin(st);
        gemm_sub(m, nn, k, a, b, c, n, s, batch, st);
        if (tm) tm->end(st, 6.0 * m * nn * k * batch);
    };
    // the register-resident panel kernels leave the interchange map of their
    // panel in perm (free until the final permutation is written), so that the
    // row kernels' workgroups do not each replay the swaps
    const int* rmap = nullptr;
    auto rows = [&](int mode, int j0, int jb) {
        const int nout = n - jb;
        if (nout <= 0) return;
        const dim3 g((nout + 31) / 32, batch), t(256);
        if (mode == 3) hipLaunchKernelGGL((k_lu_rows<C, 3>), g, t, 0, st, A, n, j0, jb, ipiv, rmap);
        else if (mode == 1) hipLaunchKernelGGL((k_lu_rows<C, 1>), g, t, 0, st, A, n, j0, jb, ipiv, rmap);
        else hipLaunchKernelGGL((k_lu_rows<C, 2>), g, t, 0, st, A, n, j0, jb, ipiv, nullptr);
    };
    auto panel = [&](int j, int jb) {
        const int rows = n - j;
        const dim3 g(batch), t(1024);
        constexpr bool f32 = sizeof(C) == sizeof(cflt);   // half the registers per value: wider chunks
        int* m = n >= kRowMapInts ? perm : nullptr;
        rmap = rows <= 4096 ? m : nullptr;
        if (rows <= 1024) hipLaunchKernelGGL((k_lu_panel_rg<C, 1, 8>), g, t, 0, st, A, n, j, jb, ipiv, m);
        else if (rows <= 2048) hipLaunchKernelGGL((k_lu_panel_rg<C, 2, f32 ? 8 : 6>), g, t, 0, st, A, n, j, jb, ipiv, m);
        else if (rows <= 4096) hipLaunchKernelGGL((k_lu_panel_rg<C, 4, f32 ? 8 : 4>), g, t, 0, st, A, n, j, jb, ipiv, m);
        else hipLaunchKernelGGL((k_lu_panel<C>), g, t, 0, st, A, n, j, jb, ipiv);
    };
    for (int j0 = 0; j0 < n; j0 += S * sb) {
        const int jE = n < j0 + S * sb ? n : j0 + S * sb;
        for (int sp = 0; sp < S; ++sp) {
            const int js = j0 + sp * sb;
            if (js >= n) break;
            const int jb = n - js < sb ? n - js : sb;
            const int je = js + jb;
            if (sp == 0) {
                panel(js, jb);
                rows(3, js, jb);
                continue;
            }
            update(n - js, jb, js - j0, at(js, j0), at(j0, js), at(js, js));
            panel(js, jb);
            rows(1, js, jb);
            if (n - je > 0) {
                update(jb, n - je, js - j0, at(js, j0), at(j0, je), at(js, je));
                rows(2, js, jb);
            }
        }
        if (n - jE > 0) update(n - jE, n - jE, jE - j0, at(jE, j0), at(j0, jE), at(jE, jE));
    }
    hipLaunchKernelGGL(k_lu_perm, dim3(batch), dim3(64), 0, st, n, ipiv, perm);
}

void lu_factor_batched(cdbl* A, int n, int batch, int* ipiv, int* perm, hipStream_t st, LaunchTimer* gemm) {
    lu_factor_impl(A, n, batch, ipiv, perm, st, gemm);
}

void lu_factor_batched_f32(cflt* A, int n, int batch, int* ipiv, int* perm, hipStream_t st, LaunchTimer* gemm) {
    lu_factor_impl(A, n, batch, ipiv, perm, st, gemm);
}

// ---------------------------------------------------------------------------
// Thin triangular solves
// ---------------------------------------------------------------------------
// A = P^T L U with L (unit lower) and U in one column-major array LU.  All four
// substitutions read LU in place, always down its columns:
//   A x = b:    L y = P b forward, U x = y backward -- right-looking: once the
//               block's unknowns are known, columns of L (below) / U (above)
//               update the other rows;
//   A^H x = b:  U^H y = b forward, L^H z = y backward, x = P^T z -- left-
//               looking: (U^H y)_i = sum_j conj(U(j, i)) y_j runs down column
//               i of U above the diagonal (of L below it for L^H), so each
//               block first gathers its rows' sums from the finished unknowns.
// Blocks of 64 rows; the diagonal block is staged in LDS (as the operator of
// the pass) and solved by wavefront 0 with cross-lane broadcasts, while the
// other wavefronts stage the next block.
constexpr int kMaxNc = 8;
constexpr int kSolveThreads = 1024;
constexpr int kTS = 65;   // padded row stride of the 64 x 64 LDS diagonal block

struct SolveLds {
    cdbl tile[2][64 * kTS];   // the current diagonal block and the next one
    cdbl xb[64 * kMaxNc];     // solution of the last block
};

// Stage the diagonal block [r0, r0 + bs)^2 of the pass operator with threads
// t0, t0 + stride, ...: conj = false: op(i, j) = LU(i, j); conj = true:
// op(i, j) = conj(LU(j, i)).
template <class F>
__device__ __forceinline__ void stage_diag(const F* M, int n, int r0, int bs, bool conj, cdbl* tile, int t0,
                                           int stride) {
    for (int e = t0; e < 64 * 64; e += stride) {
        const int p = e & 63, q = e >> 6;   // p runs down a column of LU
        const bool in = p < bs && q < bs;
        const cdbl v = in ? to_cdbl(M[(r0 + p) + (size_t)(r0 + q) * n]) : mk(0, 0);
        if (conj) tile[q * kTS + p] = cconj(v);
        else tile[p * kTS + q] = v;
    }
}

// Wavefront 0: solve the staged block for w[r0, r0 + bs) in place and leave
// the solution in xb (zero beyond bs).  With a non-unit diagonal the block is
// solved as (M D^-1) y = w, x = D^-1 y: the reciprocal pivots are formed by
// all lanes at once and scale the tile entries as they are read, eight steps
// ahead, and x = D^-1 y is formed at the end, so a step's dependence chain is
// one broadcast (v_readlane: the step's lane is the same for the whole
// wavefront) and one complex multiply-add.  Every complex product is written
// as explicit FMAs, so the bits do not depend on how the compiler contracts
// the code around each call site.
template <int NC>
__device__ __forceinline__ void diag_solve(int n, int r0, int bs, bool lower, bool unit, cdbl* w,
                                           const cdbl* __restrict__ tile, cdbl* xb) {
    const int lane = threadIdx.x & 63;
    cdbl x[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) x[c] = lane < bs ? w[r0 + lane + (size_t)c * n] : mk(0, 0);
    const cdbl dv = tile[lane * kTS + lane];
    const cdbl dinv = (unit || lane >= bs) ? mk(1, 0) : crecip(dv);
    for (int s0 = 0; s0 < bs; s0 += 8) {
        cdbl lv[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int step = min(s0 + u, bs - 1);
            const int j = lower ? step : bs - 1 - step;
            lv[u] = tile[lane * kTS + j];
            if (!unit) lv[u] = cmulx(lv[u], bcast_lane(dinv, j));
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int step = s0 + u;
            if (step >= bs) break;
            const int j = lower ? step : bs - 1 - step;
            const bool other = lower ? (lane > j && lane < bs) : (lane < j);
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                const cdbl xj = bcast_lane(x[c], j);
                if (other) cfms(x[c], lv[u], xj);
            }
        }
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        if (!unit) x[c] = cmulx(x[c], dinv);
        if (lane < bs) w[r0 + lane + (size_t)c * n] = x[c];
        xb[lane * kMaxNc + c] = lane < bs ? x[c] : mk(0, 0);
    }
}

// Right-looking pass with the column-major factor: L (lower, unit) forward or
// U (upper) backward, over the rows [pb, pe) (the whole matrix, or one panel
// of the panelled form).  After block b, the rows of the range outside it
// (below / above) take w_i -= sum_j M(i, r0 + j) x_j: a wavefront covers RW
// rows x 64 / RW groups
// of columns (lane = row + RW group), each thread loading its columns in
// batches of eight.  RW = 16 for FP64 factors, 32 for FP32 ones, so that a
// load instruction reads runs of 256 bytes down the columns either way.
template <int NC, class F, int RWT = 0>
__device__ __forceinline__ void subst_right(const F* M, int n, bool lower, cdbl* w, SolveLds& S, int pb, int pe) {
    // rows per wavefront, columns per group (RWT: chosen by the caller)
    constexpr int RW = RWT ? RWT : (sizeof(F) == sizeof(cflt) ? 32 : 16), CW = RW;
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nwv = blockDim.x >> 6;
    const int l16 = lane % RW, grp = lane / RW;
    const int nblk = (pe - pb + 63) / 64;
    auto rb = [&](int s) { return pb + 64 * (lower ? s : nblk - 1 - s); };   // first row of the s-th block solved
    auto bsz = [&](int s) { return min(64, pe - rb(s)); };
    stage_diag(M, n, rb(0), bsz(0), false, S.tile[0], tid, blockDim.x);
    for (int s = 0; s < nblk; ++s) {
        const int r0 = rb(s), bs = bsz(s);
        __syncthreads();                          // block s staged, block s - 1's update done
        if (wv == 0) {
            diag_solve<NC>(n, r0, bs, lower, lower, w, S.tile[s & 1], S.xb);
        } else if (s + 1 < nblk) {
            stage_diag(M, n, rb(s + 1), bsz(s + 1), false, S.tile[(s + 1) & 1], tid - 64, blockDim.x - 64);
        }
        __syncthreads();
        const int ia = lower ? r0 + bs : pb, ib = lower ? pe : r0;
        for (int i0 = ia + RW * wv; i0 < ib; i0 += RW * nwv) {
            const int i = i0 + l16;
            cdbl acc[NC];
#pragma unroll
            for (int c = 0; c < NC; ++c) acc[c] = mk(0, 0);
            {
                // eight loads at a time from clamped (in-range) addresses, values
                // beyond the block zeroed afterwards: no branch between the loads
                const int ic = min(i, ib - 1);
#pragma unroll 1
                for (int u0 = 0; u0 < CW; u0 += 8) {
                    F t[8];
#pragma unroll
                    for (int u = 0; u < 8; ++u)
                        t[u] = M[ic + (size_t)(r0 + min(CW * grp + u0 + u, bs - 1)) * n];
#pragma unroll
                    for (int u = 0; u < 8; ++u) {
                        const int j = CW * grp + u0 + u;
                        const cdbl tv = j < bs ? to_cdbl(t[u]) : mk(0, 0);
#pragma unroll
                        for (int c = 0; c < NC; ++c) cfma(acc[c], tv, S.xb[j * kMaxNc + c]);
                    }
                }
            }
#pragma unroll
            for (int c = 0; c < NC; ++c)
#pragma unroll
                for (int o = RW; o < 64; o <<= 1) {
                    acc[c].x += __shfl_xor(acc[c].x, o, 64);
                    acc[c].y += __shfl_xor(acc[c].y, o, 64);
                }
            if (grp == 0 && i < ib) {
#pragma unroll
                for (int c = 0; c < NC; ++c) {
                    cdbl& y = w[i + (size_t)c * n];
                    y = y - acc[c];
                }
            }
        }
    }
    __syncthreads();
}

// Left-looking pass with the conjugate transpose of the factor: U^H (lower)
// forward or L^H (upper, unit) backward, over the rows [pb, pe).  Block b
// first takes w_i -= sum_j conj(M(j, i)) w_j over the finished unknowns j of
// the range (above the block for U^H, below it for L^H; those outside the
// range are gathered beforehand): wavefront v sums columns i = r0 + 4 v .. + 3
// (CP of them per sweep down the column, CP = 4 / 2 / 1 for up to 2 / 4 / 8
// right-hand sides) with its lanes down the column, then the diagonal block is
// solved.
template <int NC, class F, int TW = 1>
__device__ __forceinline__ void subst_left(const F* M, int n, bool lower, cdbl* w, SolveLds& S, int pb, int pe) {
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int nblk = (pe - pb + 63) / 64;
    auto rb = [&](int s) { return pb + 64 * (lower ? s : nblk - 1 - s); };
    auto bsz = [&](int s) { return min(64, pe - rb(s)); };
    stage_diag(M, n, rb(0), bsz(0), true, S.tile[0], tid, blockDim.x);
    for (int s = 0; s < nblk; ++s) {
        const int r0 = rb(s), bs = bsz(s);
        const int ja = lower ? pb : r0 + bs, jb = lower ? r0 : pe;
        constexpr int CP = NC <= 2 ? 4 : (NC <= 4 ? 2 : 1);
        for (int q0 = 0; q0 < 4; q0 += CP) {
            if (4 * wv + q0 >= bs) break;        // wave-uniform
            const int i0 = r0 + 4 * wv + q0;
            cdbl acc[CP][NC];
#pragma unroll
            for (int q = 0; q < CP; ++q)
#pragma unroll
                for (int c = 0; c < NC; ++c) acc[q][c] = mk(0, 0);
            // the same number of sweeps for every lane, TW sweeps' loads in
            // flight at once, at clamped addresses (values beyond the range
            // replaced by zero afterwards; sweeps past the last add zeros)
            const int T = (jb - ja + 63) >> 6;
            const int qmax = bs - 1 - 4 * wv - q0;    // last column of the block for this wavefront
            for (int tt = 0; tt < T; tt += TW) {
                cdbl xj[TW][NC];
                F mv[TW][CP];
#pragma unroll
                for (int h = 0; h < TW; ++h) {
                    const int jc = min(ja + lane + 64 * (tt + h), jb - 1);
#pragma unroll
                    for (int c = 0; c < NC; ++c) xj[h][c] = w[jc + (size_t)c * n];
#pragma unroll
                    for (int q = 0; q < CP; ++q) mv[h][q] = M[jc + (size_t)(i0 + min(q, qmax)) * n];
                }
#pragma unroll
                for (int h = 0; h < TW; ++h) {
                    const bool in = ja + lane + 64 * (tt + h) < jb;
#pragma unroll
                    for (int c = 0; c < NC; ++c)
                        if (!in) xj[h][c] = mk(0, 0);
#pragma unroll
                    for (int q = 0; q < CP; ++q) {
                        const cdbl t = q <= qmax ? cconj(to_cdbl(mv[h][q])) : mk(0, 0);
#pragma unroll
                        for (int c = 0; c < NC; ++c) cfma(acc[q][c], t, xj[h][c]);
                    }
                }
            }
#pragma unroll
            for (int q = 0; q < CP; ++q)
#pragma unroll
                for (int c = 0; c < NC; ++c)
#pragma unroll
                    for (int o = 32; o > 0; o >>= 1) {
                        acc[q][c].x += __shfl_xor(acc[q][c].x, o, 64);
                        acc[q][c].y += __shfl_xor(acc[q][c].y, o, 64);
                    }
            // lane q stores column i0 + q (register arrays are indexed with constants)
#pragma unroll
            for (int q = 0; q < CP; ++q)
                if (lane == q && 4 * wv + q0 + q < bs) {
#pragma unroll
                    for (int c = 0; c < NC; ++c) {
                        cdbl& y = w[i0 + q + (size_t)c * n];
                        y = y - acc[q][c];
                    }
                }
        }
        __syncthreads();                          // sums in, block s staged
        if (wv == 0) {
            diag_solve<NC>(n, r0, bs, lower, !lower, w, S.tile[s & 1], S.xb);
        } else if (s + 1 < nblk) {
            stage_diag(M, n, rb(s + 1), bsz(s + 1), true, S.tile[(s + 1) & 1], tid - 64, blockDim.x - 64);
        }
        __syncthreads();
    }
}

// one workgroup of kSolveThreads threads per matrix
template <class F, int NC>
__global__ __launch_bounds__(kSolveThreads) void k_solve_thin(const F* __restrict__ LU, long long ms,
                                                              const int* __restrict__ perm, int n,
                                                              cdbl* __restrict__ X, int adj, cdbl* __restrict__ work) {
    __shared__ SolveLds S;
    const int b = blockIdx.x;
    const int* p = perm + (size_t)b * n;
    cdbl* x = X + (size_t)b * n * NC;
    cdbl* w = work + (size_t)b * n * NC;
    const F* lu = LU + (size_t)b * ms;
    const int tid = threadIdx.x;
    if (!adj) {
        // A = P^T L U:  w = P x, L y = w, U z = y
        for (int e = tid; e < n * NC; e += blockDim.x) {
            const int i = e % n, c = e / n;
            w[e] = x[p[i] + (size_t)c * n];
        }
        __syncthreads();
        subst_right<NC>(lu, n, true, w, S, 0, n);
        subst_right<NC>(lu, n, false, w, S, 0, n);
        for (int e = tid; e < n * NC; e += blockDim.x) x[e] = w[e];
    } else {
        // A^H = U^H L^H P:  U^H y = x, L^H z = y, x = P^T z
        for (int e = tid; e < n * NC; e += blockDim.x) w[e] = x[e];
        __syncthreads();
        subst_left<NC>(lu, n, true, w, S, 0, n);
        subst_left<NC>(lu, n, false, w, S, 0, n);
        for (int e = tid; e < n * NC; e += blockDim.x) {
            const int i = e % n, c = e / n;
            x[p[i] + (size_t)c * n] = w[e];
        }
    }
}

// Panelled form of the thin solves for orders above 2 kSolvePanel: the
// passes run panel by panel of kSolvePanel rows.  A panel's triangle is solved
// by one workgroup per matrix (k_solve_panel: the blocked substitution above,
// restricted to the panel); the panel's coupling to the other rows runs on
// many workgroups per matrix: a right-looking pass then updates every row
// after the panel with the panel's columns (k_panel_update), a left-looking
// pass first gathers into the panel's rows the sums over the unknowns finished
// before it (k_panel_gather).  Each entry of the factors is still read once,
// but the bulk of the reads no longer waits on one workgroup per matrix.
constexpr int kSolvePanel = 512;

// w = P x (A x = b) or w = x (A^H x = b)
__global__ void k_thin_in(const int* __restrict__ perm, int n, int nc, const cdbl* __restrict__ X, int adj,
                          cdbl* __restrict__ work) {
    const int b = blockIdx.y;
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long long)n * nc) return;
    const int i = (int)(e % n), c = (int)(e / n);
    const cdbl* x = X + (size_t)b * n * nc;
    work[(size_t)b * n * nc + e] = adj ? x[e] : x[perm[(size_t)b * n + i] + (size_t)c * n];
}

// x = w (A x = b) or x = P^T w (A^H x = b)
__global__ void k_thin_out(const int* __restrict__ perm, int n, int nc, cdbl* __restrict__ X, int adj,
                           const cdbl* __restrict__ work) {
    const int b = blockIdx.y;
    const long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= (long long)n * nc) return;
    const int i = (int)(e % n), c = (int)(e / n);
    cdbl* x = X + (size_t)b * n * nc;
    const cdbl v = work[(size_t)b * n * nc + e];
    if (adj) x[perm[(size_t)b * n + i] + (size_t)c * n] = v;
    else x[e] = v;
}

// one workgroup of kSolveThreads threads per matrix: pass 0 (L, or U^H) or
// pass 1 (U, or L^H) over the panel [pb, pe)
// Right-looking coupling of the unknowns [cb, ce) (one panel, staged in xs)
// onto row i: w_i -= sum_j M(i, j) w_j, the sum in column order.  A thread per
// row (consecutive rows: the loads run down the columns), LR column loads in
// flight.
template <int NC>
__device__ __forceinline__ void stage_unknowns(const cdbl* w, int n, int cb, int ce, cdbl* xs) {
    const int pw = ce - cb;
    for (int e = threadIdx.x; e < pw * NC; e += blockDim.x) {
        const int j = e % pw, c = e / pw;
        xs[j * NC + c] = w[cb + j + (size_t)c * n];
    }
}
template <class F, int NC>
__device__ __forceinline__ void couple_right_row(const F* M, int n, cdbl* w, int cb, int ce, int i, const cdbl* xs) {
    constexpr int LR = sizeof(F) == sizeof(cflt) ? 16 : 8;
    const int pw = ce - cb;
    cdbl acc[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] = mk(0, 0);
    const F* mr = M + i + (size_t)cb * n;
    for (int j0 = 0; j0 < pw; j0 += LR) {
        F t[LR];
#pragma unroll
        for (int u = 0; u < LR; ++u) t[u] = mr[(size_t)min(j0 + u, pw - 1) * n];
#pragma unroll
        for (int u = 0; u < LR; ++u) {
            if (j0 + u < pw) {
                const cdbl tv = to_cdbl(t[u]);
#pragma unroll
                for (int c = 0; c < NC; ++c) cfma(acc[c], tv, xs[(j0 + u) * NC + c]);
            }
        }
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) {
        cdbl& y = w[i + (size_t)c * n];
        y = y - acc[c];
    }
}

// Left-looking coupling of the unknowns [ja, jb) onto rows i0 .. i0 + 3 (those
// below re): w_i -= sum_j conj(M(j, i)) w_j.  One wavefront, lanes down the
// columns (lane j mod 64: ja is a multiple of 64), two sweeps' loads in
// flight; the lane sums are reduced by shuffles in a fixed order.
template <class F, int NC>
__device__ __forceinline__ void couple_left_quad(const F* M, int n, cdbl* w, int ja, int jb, int i0, int re) {
    const int lane = threadIdx.x & 63;
    const int qmax = min(4, re - i0) - 1;
    cdbl acc[4][NC];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int c = 0; c < NC; ++c) acc[q][c] = mk(0, 0);
    const int T = (jb - ja + 63) >> 6;
    for (int tt = 0; tt < T; tt += 2) {
        cdbl xj[2][NC];
        F mv[2][4];
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const int jc = min(ja + lane + 64 * (tt + h), jb - 1);
#pragma unroll
            for (int c = 0; c < NC; ++c) xj[h][c] = w[jc + (size_t)c * n];
#pragma unroll
            for (int q = 0; q < 4; ++q) mv[h][q] = M[jc + (size_t)(i0 + min(q, qmax)) * n];
        }
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            const bool in = ja + lane + 64 * (tt + h) < jb;
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const cdbl t = (in && q <= qmax) ? cconj(to_cdbl(mv[h][q])) : mk(0, 0);
#pragma unroll
                for (int c = 0; c < NC; ++c) cfma(acc[q][c], t, xj[h][c]);
            }
        }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                acc[q][c].x += __shfl_xor(acc[q][c].x, o, 64);
                acc[q][c].y += __shfl_xor(acc[q][c].y, o, 64);
            }
#pragma unroll
    for (int q = 0; q < 4; ++q)
        if (lane == q && q <= qmax) {
#pragma unroll
            for (int c = 0; c < NC; ++c) {
                cdbl& y = w[i0 + q + (size_t)c * n];
                y = y - acc[q][c];
            }
        }
}

// One panel [pb, pe) of a pass, blocks 0 .. batch - 1 (one per matrix), and
// beside it, in blocks batch + cwg b + k, a coupling that does not touch the
// panel: the unknowns [cb, ce) onto the rows [rb, re) of matrix b,
// right-looking (a thread per row, 1024 rows a block) or, for adj, left-looking
// (a wavefront per four rows, 64 rows a block).  The coupling blocks take the
// CUs the batch's panel solves leave idle.  Up to kFuseNc right-hand sides:
// with more, the left-looking sums do not fit the registers of a
// 1024-thread block, and the couplings run as launches of their own.
constexpr int kFuseNc = 3;
template <class F, int NC>
__global__ __launch_bounds__(kSolveThreads) void k_solve_panel(const F* __restrict__ LU, long long ms, int n, int pass,
                                                               int adj, cdbl* __restrict__ work, int pb, int pe,
                                                               int batch, int cwg, int cb, int ce, int rb, int re) {
    __shared__ SolveLds S;
    if ((int)blockIdx.x < batch) {
        const int b = blockIdx.x;
        cdbl* w = work + (size_t)b * n * NC;
        const F* lu = LU + (size_t)b * ms;
        // inside a panel the updates cover at most kSolvePanel - 64 rows: 16 rows
        // per wavefront (more wavefronts busy, fewer load rounds each) and two
        // sweeps' loads in flight in the left-looking sums
        if (!adj) subst_right<NC, F, 16>(lu, n, pass == 0, w, S, pb, pe);
        else subst_left<NC, F, NC <= 4 ? 2 : 1>(lu, n, pass == 0, w, S, pb, pe);
        return;
    }
    if constexpr (NC <= kFuseNc) {
        const int q = blockIdx.x - batch, b = q / cwg, k = q - b * cwg;
        cdbl* w = work + (size_t)b * n * NC;
        const F* M = LU + (size_t)b * ms;
        if (!adj) {
            cdbl* xs = reinterpret_cast<cdbl*>(&S);   // the panel solve's LDS is not used by this block
            stage_unknowns<NC>(w, n, cb, ce, xs);
            __syncthreads();
            const int i = rb + k * blockDim.x + threadIdx.x;
            if (i < re) couple_right_row<F, NC>(M, n, w, cb, ce, i, xs);
        } else {
            const int i0 = rb + (k * (blockDim.x >> 6) + (threadIdx.x >> 6)) * 4;
            if (i0 < re) couple_left_quad<F, NC>(M, n, w, cb, ce, i0, re);
        }
    }
}

// The same couplings on their own.  Right-looking, 64 rows a block: wavefront
// v sums the unknowns cb + 128 v .. + 127 for the block's rows (lane = row,
// so the loads still run down the columns in 64-row segments), and the four
// partial sums are added in a fixed order, ((0 + 1) + (2 + 3)), by wavefront 0:
// each thread's chain of dependent load rounds is a quarter as long as with
// one thread per row over the whole panel.  The quarters are fixed column
// ranges from the panel's start, so a panel padded with zero columns gives the
// same sums.
template <class F, int NC>
__global__ __launch_bounds__(256) void k_panel_update(const F* __restrict__ LU, long long ms, int n,
                                                      cdbl* __restrict__ work, int cb, int ce, int rb, int re) {
    constexpr int LR = sizeof(F) == sizeof(cflt) ? 16 : 8;
    constexpr int QW = kSolvePanel / 4;
    __shared__ cdbl xs[kSolvePanel * NC];
    __shared__ cdbl part[4][64][NC];
    const int b = blockIdx.y, lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    cdbl* w = work + (size_t)b * n * NC;
    stage_unknowns<NC>(w, n, cb, ce, xs);
    __syncthreads();
    const int i = rb + blockIdx.x * 64 + lane;
    const int q0 = wv * QW, q1 = min(ce - cb, q0 + QW);
    cdbl acc[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] = mk(0, 0);
    if (i < re) {
        const F* mr = LU + (size_t)b * ms + i + (size_t)cb * n;
        for (int j0 = q0; j0 < q1; j0 += LR) {
            F t[LR];
#pragma unroll
            for (int u = 0; u < LR; ++u) t[u] = mr[(size_t)min(j0 + u, q1 - 1) * n];
#pragma unroll
            for (int u = 0; u < LR; ++u) {
                if (j0 + u < q1) {
                    const cdbl tv = to_cdbl(t[u]);
#pragma unroll
                    for (int c = 0; c < NC; ++c) cfma(acc[c], tv, xs[(j0 + u) * NC + c]);
                }
            }
        }
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) part[wv][lane][c] = acc[c];
    __syncthreads();
    if (wv == 0 && i < re) {
#pragma unroll
        for (int c = 0; c < NC; ++c) {
            cdbl& y = w[i + (size_t)c * n];
            y = y - ((part[0][lane][c] + part[1][lane][c]) + (part[2][lane][c] + part[3][lane][c]));
        }
    }
}

// Left-looking, 16 rows a block.
template <class F, int NC>
__global__ __launch_bounds__(256) void k_panel_gather(const F* __restrict__ LU, long long ms, int n,
                                                      cdbl* __restrict__ work, int ja, int jb, int rb, int re) {
    const int b = blockIdx.y;
    const int i0 = rb + (blockIdx.x * 4 + (threadIdx.x >> 6)) * 4;
    if (i0 >= re) return;                     // whole wavefront; no barrier in this kernel
    couple_left_quad<F, NC>(LU + (size_t)b * ms, n, work + (size_t)b * n * NC, ja, jb, i0, re);
}

// Panelled substitution (n > 2 kSolvePanel).  A pass solves its panels in
// order; every row must have taken the couplings of all panels solved before
// its own, in that order.  After panel p's solve, the next panel's rows take
// p's coupling (a small launch); the rows beyond the next panel take it in
// the launch that solves the next panel, beside that solve.  Left-looking
// passes (adj) gather instead: the next panel's rows take the unknowns solved
// before p beside p's solve, then p's own unknowns.  A row's sums are split
// only at panel boundaries and at fixed offsets inside a panel, and which
// launch serves a row depends only on its panel, so a matrix padded at the
// end gives the same bits.  Above
// kFuseNc right-hand sides: each panel's coupling as a launch of its own,
// all rows at once, after its solve (right-looking) or before it (left).
template <class F, int NC>
static void solve_thin_nc(const F* LU, long long ms, const int* perm, int n, int batch, cdbl* X, bool adj, cdbl* work,
                          hipStream_t st) {
    const int a = adj ? 1 : 0;
    if (n <= 2 * kSolvePanel) {
        hipLaunchKernelGGL((k_solve_thin<F, NC>), dim3(batch), dim3(kSolveThreads), 0, st, LU, ms, perm, n, X, a, work);
        return;
    }
    const dim3 gio((unsigned)(((long long)n * NC + 255) / 256), (unsigned)batch);
    hipLaunchKernelGGL(k_thin_in, gio, dim3(256), 0, st, perm, n, NC, X, a, work);
    const int np = (n + kSolvePanel - 1) / kSolvePanel;
    auto first = [&](int p) { return p * kSolvePanel; };
    auto last = [&](int p) { return min(n, (p + 1) * kSolvePanel); };
    if constexpr (NC > kFuseNc) {
        for (int pass = 0; pass < 2; ++pass) {
            const bool lower = pass == 0;
            for (int t = 0; t < np; ++t) {
                const int p = lower ? t : np - 1 - t;
                const int pb = first(p), pe = last(p);
                if (adj && (lower ? pb > 0 : pe < n))
                    hipLaunchKernelGGL((k_panel_gather<F, NC>), dim3((unsigned)((pe - pb + 15) / 16), (unsigned)batch),
                                       dim3(256), 0, st, LU, ms, n, work, lower ? 0 : pe, lower ? pb : n, pb, pe);
                hipLaunchKernelGGL((k_solve_panel<F, NC>), dim3((unsigned)batch), dim3(kSolveThreads), 0, st, LU, ms, n,
                                   pass, a, work, pb, pe, batch, 1, 0, 0, 0, 0);
                const int rb = lower ? pe : 0, re = lower ? n : pb;
                if (!adj && re > rb)
                    hipLaunchKernelGGL((k_panel_update<F, NC>), dim3((unsigned)((re - rb + 63) / 64), (unsigned)batch),
                                       dim3(256), 0, st, LU, ms, n, work, pb, pe, rb, re);
            }
        }
        hipLaunchKernelGGL(k_thin_out, gio, dim3(256), 0, st, perm, n, NC, X, a, work);
        return;
    }
    for (int pass = 0; pass < 2; ++pass) {
        const bool lower = pass == 0;
        for (int t = 0; t < np; ++t) {
            const int p = lower ? t : np - 1 - t;
            const int pb = first(p), pe = last(p);
            const bool next = t + 1 < np;
            const int np_ = lower ? p + 1 : p - 1;             // the next panel of the pass
            const int nb = next ? first(np_) : 0, ne = next ? last(np_) : 0;
            int cb = 0, ce = 0, rb = 0, re = 0;                 // the coupling beside this solve
            if (!adj && t > 0) {
                const int pp = lower ? p - 1 : p + 1;          // the panel solved before
                cb = first(pp);
                ce = last(pp);
                rb = lower ? pe : 0;                           // the rows beyond this panel
                re = lower ? n : pb;
            } else if (adj && next) {
                cb = lower ? 0 : pe;
                ce = lower ? pb : n;
                rb = nb;
                re = ne;
            }
            const bool cpl = ce > cb && re > rb;
            const int cwg = !cpl ? 0 : (adj ? (re - rb + 63) / 64 : (re - rb + kSolveThreads - 1) / kSolveThreads);
            hipLaunchKernelGGL((k_solve_panel<F, NC>), dim3((unsigned)(batch + batch * cwg)), dim3(kSolveThreads), 0, st,
                               LU, ms, n, pass, a, work, pb, pe, batch, cwg > 0 ? cwg : 1, cb, ce, rb, re);
            if (!next) continue;
            if (!adj)
                hipLaunchKernelGGL((k_panel_update<F, NC>), dim3((unsigned)((ne - nb + 63) / 64), (unsigned)batch),
                                   dim3(256), 0, st, LU, ms, n, work, pb, pe, nb, ne);
            else
                hipLaunchKernelGGL((k_panel_gather<F, NC>), dim3((unsigned)((ne - nb + 15) / 16), (unsigned)batch),
                                   dim3(256), 0, st, LU, ms, n, work, pb, pe, nb, ne);
        }
    }
    hipLaunchKernelGGL(k_thin_out, gio, dim3(256), 0, st, perm, n, NC, X, a, work);
}

template <class F>
static void solve_thin_impl(const F* LU, long long ms, const int* perm, int n, int batch, cdbl* X, int nc, bool adj,
                            cdbl* work, hipStream_t st) {
    switch (nc) {
    case 1: solve_thin_nc<F, 1>(LU, ms, perm, n, batch, X, adj, work, st); break;
    case 2: solve_thin_nc<F, 2>(LU, ms, perm, n, batch, X, adj, work, st); break;
    case 3: solve_thin_nc<F, 3>(LU, ms, perm, n, batch, X, adj, work, st); break;
    case 4: solve_thin_nc<F, 4>(LU, ms, perm, n, batch, X, adj, work, st); break;
    case 5: solve_thin_nc<F, 5>(LU, ms, perm, n, batch, X, adj, work, st); break;
    case 6: solve_thin_nc<F, 6>(LU, ms, perm, n, batch, X, adj, work, st); break;
    case 7: solve_thin_nc<F, 7>(LU, ms, perm, n, batch, X, adj, work, st); break;
    default: solve_thin_nc<F, 8>(LU, ms, perm, n, batch, X, adj, work, st); break;
    }
}

void lu_solve_thin_batched(const cdbl* LU, long long ms, const int* perm, int n, int batch, cdbl* X, int nc, bool adj,
                           cdbl* work, hipStream_t st) {
    solve_thin_impl(LU, ms, perm, n, batch, X, nc, adj, work, st);
}

void lu_solve_thin_batched_f32(const cflt* LU, long long ms, const int* perm, int n, int batch, cdbl* X, int nc,
                               bool adj, cdbl* work, hipStream_t st) {
    solve_thin_impl(LU, ms, perm, n, batch, X, nc, adj, work, st);
}

// ---------------------------------------------------------------------------
// FP64 residuals and updates of the mixed-precision iterative refinement
// ---------------------------------------------------------------------------
// R = Bv - op(T) Z for the thin blocks of a batch (T: n x n FP64 with stride
// ts; Z, Bv, R: n x NC with stride n NC).  op = T: a workgroup takes 256 rows
// and a slice of kResidSlice columns (the slice's Z rows staged in LDS), one
// thread per row streaming down the columns; the slices' partial sums are
// added in slice order by k_resid_sum.  op = T^H: a wavefront takes four
// columns i of T, (T^H z)_i = sum_j conj(T(j, i)) z_j, lanes down the columns
// (k_resid_h).
constexpr int kResidSlice = 256;

// an entry of T read once per residual pass: a non-temporal load, so that the
// 33 GB stream does not displace what the cache holds (k_resid_h 5.28 ->
// 4.91 ms per polygon2000 launch)
typedef double dbl2v __attribute__((ext_vector_type(2)));
__device__ __forceinline__ cdbl load_stream(const cdbl* p) {
    const dbl2v v = __builtin_nontemporal_load(reinterpret_cast<const dbl2v*>(p));
    return mk(v.x, v.y);
}

template <int NC>
__global__ __launch_bounds__(256) void k_resid_n(const cdbl* __restrict__ T, long long ts, const cdbl* __restrict__ Z,
                                                 int n, int nslice, cdbl* __restrict__ part) {
    __shared__ cdbl zs[kResidSlice][NC];
    const int b = blockIdx.y;
    const int rb = blockIdx.x / nslice, sl = blockIdx.x - rb * nslice;
    const int i = rb * 256 + threadIdx.x;
    const int j0 = sl * kResidSlice, j1 = min(n, j0 + kResidSlice);
    const cdbl* z = Z + (size_t)b * n * NC;
    for (int e = threadIdx.x; e < kResidSlice * NC; e += 256) {
        const int jj = e % kResidSlice, c = e / kResidSlice;
        zs[jj][c] = j0 + jj < j1 ? z[j0 + jj + (size_t)c * n] : mk(0, 0);
    }
    __syncthreads();
    cdbl acc[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] = mk(0, 0);
    if (i < n) {
        const cdbl* t = T + (size_t)b * ts + i + (size_t)j0 * n;
        int j = j0;
        for (; j + 4 <= j1; j += 4) {
            cdbl tv[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) tv[u] = load_stream(t + (size_t)(j - j0 + u) * n);
#pragma unroll
            for (int u = 0; u < 4; ++u)
#pragma unroll
                for (int c = 0; c < NC; ++c) cfma(acc[c], tv[u], zs[j - j0 + u][c]);
        }
        for (; j < j1; ++j) {
            const cdbl tv = t[(size_t)(j - j0) * n];
#pragma unroll
            for (int c = 0; c < NC; ++c) cfma(acc[c], tv, zs[j - j0][c]);
        }
        cdbl* p = part + (((size_t)sl * gridDim.y + b) * NC) * n + i;
#pragma unroll
        for (int c = 0; c < NC; ++c) p[(size_t)c * n] = acc[c];
    }
}

__global__ __launch_bounds__(256) void k_resid_sum(const cdbl* __restrict__ part, int nslice, int n, int ncb, int batch,
                                                   const cdbl* __restrict__ Bv, cdbl* __restrict__ R) {
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long per = (long long)n * ncb;
    if (id >= per * batch) return;
    cdbl acc = mk(0, 0);
    for (int sl = 0; sl < nslice; ++sl) acc = acc + part[(size_t)sl * per * batch + id];
    R[id] = Bv[id] - acc;
}

// R = Bv - T^H Z: a wavefront per four columns i of T (column i gives row i
// of T^H), lanes down the columns, two sweeps' loads in flight; every z value
// loaded serves the four columns.  Each column's lane sums run over j in
// sweep order and are reduced by shuffles in a fixed order.
template <int NC>
__global__ __launch_bounds__(256) void k_resid_h(const cdbl* __restrict__ T, long long ts, const cdbl* __restrict__ Z,
                                                 const cdbl* __restrict__ Bv, cdbl* __restrict__ R, int n) {
    const int b = blockIdx.y;
    const int i0 = (blockIdx.x * 4 + (threadIdx.x >> 6)) * 4, lane = threadIdx.x & 63;
    if (i0 >= n) return;                      // whole wavefront; no barrier in this kernel
    const int qmax = min(4, n - i0) - 1;
    const cdbl* T0 = T + (size_t)b * ts;
    const cdbl* z = Z + (size_t)b * n * NC;
    cdbl acc[4][NC];
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int c = 0; c < NC; ++c) acc[q][c] = mk(0, 0);
    int j = lane;
    for (; j + 64 < n; j += 128) {
        cdbl tv[2][4], zv[2][NC];
#pragma unroll
        for (int u = 0; u < 2; ++u) {
#pragma unroll
            for (int q = 0; q < 4; ++q) tv[u][q] = load_stream(T0 + j + 64 * u + (size_t)(i0 + min(q, qmax)) * n);
#pragma unroll
            for (int c = 0; c < NC; ++c) zv[u][c] = z[j + 64 * u + (size_t)c * n];
        }
#pragma unroll
        for (int u = 0; u < 2; ++u)
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const cdbl t = cconj(tv[u][q]);
#pragma unroll
                for (int c = 0; c < NC; ++c) cfma(acc[q][c], t, zv[u][c]);
            }
    }
    for (; j < n; j += 64) {
        cdbl zv[NC];
#pragma unroll
        for (int c = 0; c < NC; ++c) zv[c] = z[j + (size_t)c * n];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const cdbl t = cconj(T0[j + (size_t)(i0 + min(q, qmax)) * n]);
#pragma unroll
            for (int c = 0; c < NC; ++c) cfma(acc[q][c], t, zv[c]);
        }
    }
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
        for (int c = 0; c < NC; ++c)
#pragma unroll
            for (int o = 32; o > 0; o >>= 1) {
                acc[q][c].x += __shfl_xor(acc[q][c].x, o, 64);
                acc[q][c].y += __shfl_xor(acc[q][c].y, o, 64);
            }
#pragma unroll
    for (int q = 0; q < 4; ++q)
        if (lane == q && q <= qmax) {
            const size_t base = (size_t)b * n * NC + i0 + q;
#pragma unroll
            for (int c = 0; c < NC; ++c) R[base + (size_t)c * n] = Bv[base + (size_t)c * n] - acc[q][c];
        }
}

template <int NC>
static void resid_impl(const cdbl* T, long long ts, const cdbl* Z, const cdbl* Bv, cdbl* R, int n, int batch, bool adj,
                       cdbl* part, hipStream_t st) {
    if (adj) {
        hipLaunchKernelGGL(k_resid_h<NC>, dim3((n + 15) / 16, batch), dim3(256), 0, st, T, ts, Z, Bv, R, n);
        return;
    }
    const int nslice = (n + kResidSlice - 1) / kResidSlice, nrb = (n + 255) / 256;
    hipLaunchKernelGGL(k_resid_n<NC>, dim3(nrb * nslice, batch), dim3(256), 0, st, T, ts, Z, n, nslice, part);
    const long long tot = (long long)n * NC * batch;
    hipLaunchKernelGGL(k_resid_sum, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, part, nslice, n, NC, batch, Bv,
                       R);
}

void residual_thin_batched(const cdbl* T, long long ts, const cdbl* Z, const cdbl* Bv, cdbl* R, int n, int nc,
                           int batch, bool adj, cdbl* part, hipStream_t st) {
    if (batch <= 0 || n <= 0) return;
    switch (nc) {
    case 1: resid_impl<1>(T, ts, Z, Bv, R, n, batch, adj, part, st); break;
    case 2: resid_impl<2>(T, ts, Z, Bv, R, n, batch, adj, part, st); break;
    case 3: resid_impl<3>(T, ts, Z, Bv, R, n, batch, adj, part, st); break;
    case 4: resid_impl<4>(T, ts, Z, Bv, R, n, batch, adj, part, st); break;
    case 5: resid_impl<5>(T, ts, Z, Bv, R, n, batch, adj, part, st); break;
    case 6: resid_impl<6>(T, ts, Z, Bv, R, n, batch, adj, part, st); break;
    case 7: resid_impl<7>(T, ts, Z, Bv, R, n, batch, adj, part, st); break;
    default: resid_impl<8>(T, ts, Z, Bv, R, n, batch, adj, part, st); break;
    }
}

size_t residual_part_count(int n, int nc, int batch) {
    return (size_t)((n + kResidSlice - 1) / kResidSlice) * (size_t)n * nc * batch;
}

// Z += D for the cnt entries of each matrix's block.  prev[b] keeps max |D|^2
// of the step before (first: max |Z|^2 before the update, the size of the
// unrefined solution), so |D| / |D_prev| estimates the contraction rate and
// rate |D| the error left after this update.  check = 1: flag[b] = 1 unless
// the rate is at most 1/2 (the refinement contracts); check = 2: also unless
// rate max|D| <= tol max|Z|.  NaN fails both.
__global__ __launch_bounds__(256) void k_refine_update(cdbl* __restrict__ Z, const cdbl* __restrict__ Dv, int cnt,
                                                       double tol2, int check, int first, double* __restrict__ prev,
                                                       int* __restrict__ flag) {
    __shared__ double sh[3][4];
    const int b = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    cdbl* z = Z + (size_t)b * cnt;
    const cdbl* d = Dv + (size_t)b * cnt;
    double md = 0.0, mzo = 0.0, mz = 0.0;
    for (int e = tid; e < cnt; e += 256) {
        const cdbl dv = d[e], zo = z[e];
        const cdbl zv = zo + dv;
        z[e] = zv;
        md = fmax(md, cabs2(dv));
        mzo = fmax(mzo, cabs2(zo));
        mz = fmax(mz, cabs2(zv));
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        md = fmax(md, __shfl_xor(md, o, 64));
        mzo = fmax(mzo, __shfl_xor(mzo, o, 64));
        mz = fmax(mz, __shfl_xor(mz, o, 64));
    }
    if (lane == 0) { sh[0][wv] = md; sh[1][wv] = mzo; sh[2][wv] = mz; }
    __syncthreads();
    if (tid == 0) {
        for (int w = 1; w < 4; ++w) { md = fmax(md, sh[0][w]); mzo = fmax(mzo, sh[1][w]); mz = fmax(mz, sh[2][w]); }
        const double mp = first ? mzo : prev[b];
        if (check) {
            const double rate2 = mp > 0.0 ? md / mp : (md > 0.0 ? 1.0 : 0.0);
            const bool ok = rate2 <= 0.25 && (check == 1 || rate2 * md <= tol2 * mz);   // NaN fails too
            if (!ok) flag[b] = 1;
        }
        prev[b] = md;
    }
}

void refine_update_batched(cdbl* Z, const cdbl* D, int cnt, int batch, double tol, int check, bool first, double* prev,
                           int* flag, hipStream_t st) {
    if (batch <= 0) return;
    hipLaunchKernelGGL(k_refine_update, dim3(batch), dim3(256), 0, st, Z, D, cnt, tol * tol, check,
                       first ? 1 : 0, prev, flag);
}


// dst[i] = src[i] for nbytes bytes (a kernel, not a runtime copy, so that it
// can stand in a captured graph like the other steps)
__global__ __launch_bounds__(256) void k_copy_bytes(unsigned char* __restrict__ dst, const unsigned char* __restrict__ src,
                                                    long long nbytes) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nbytes) dst[i] = src[i];
}

void copy_bytes(void* dst, const void* src, size_t nbytes, hipStream_t st) {
    if (nbytes == 0) return;
    hipLaunchKernelGGL(k_copy_bytes, dim3((unsigned)((nbytes + 255) / 256)), dim3(256), 0, st, (unsigned char*)dst,
                       (const unsigned char*)src, (long long)nbytes);
}

__global__ void k_add_ll(long long* __restrict__ p, long long count, long long off) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] += off;
}

void add_offset_ll(long long* p, long long count, long long off, hipStream_t st) {
    if (count <= 0 || off == 0) return;
    hipLaunchKernelGGL(k_add_ll, dim3((unsigned)((count + 255) / 256)), dim3(256), 0, st, p, count, off);
}

__global__ void k_fill_int(int* __restrict__ p, int count, int v) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < count) p[i] = v;
}

void fill_int(int* p, int count, int v, hipStream_t st) {
    if (count <= 0) return;
    hipLaunchKernelGGL(k_fill_int, dim3((count + 255) / 256), dim3(256), 0, st, p, count, v);
}

// ---------------------------------------------------------------------------
// Checks of the dense solve: *flag = 1 when any of the cnt doubles at x is not
// finite / when a diagonal entry of the n x n complex matrix A is exactly zero
// (LAPACK zgesv's info > 0 after the LU).  Threads only ever store 1, so the
// concurrent stores need no atomics; the host clears the flag first.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_flag_nonfinite(const double* __restrict__ x, long long cnt, int* flag) {
    bool bad = false;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < cnt; i += (long long)gridDim.x * blockDim.x)
        bad = bad || !isfinite(x[i]);
    if (bad) *flag = 1;
}

__global__ __launch_bounds__(256) void k_flag_zero_diag(const cdbl* __restrict__ A, int n, int* flag) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const cdbl d = A[i + (size_t)i * n];
        if (d.x == 0.0 && d.y == 0.0) *flag = 1;
    }
}

void flag_nonfinite(const double* x, long long cnt, int* flag, hipStream_t st) {
    if (cnt <= 0) return;
    const long long blocks = (cnt + 255) / 256;
    hipLaunchKernelGGL(k_flag_nonfinite, dim3((unsigned)(blocks < 2048 ? blocks : 2048)), dim3(256), 0, st, x, cnt, flag);
}

void flag_zero_diag(const cdbl* A, int n, int* flag, hipStream_t st) {
    if (n <= 0) return;
    hipLaunchKernelGGL(k_flag_zero_diag, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, A, n, flag);
}

// ---------------------------------------------------------------------------
// Y = S X for a real sparse n x n S in compressed-row form; X and Y are
// n x ncols column-major (leading dimension n) -- every thin block of a batch
// at once, since the blocks of a batch are stored back to back.  One thread
// per output entry; consecutive threads take consecutive rows of one column,
// so the writes are coalesced and the gathered reads of a banded S stay in a
// few cache lines.  The row sum runs in stored (ascending column) order.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_csr_apply(const int* __restrict__ rowptr, const int* __restrict__ colidx,
                                                   const double* __restrict__ val, int n,
                                                   const cdbl* __restrict__ X, cdbl* __restrict__ Y, long long ncols) {
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long long)n * ncols) return;
    const int i = (int)(id % n);
    const cdbl* x = X + (id - i);
    cdbl s = mk(0, 0);
    const int p1 = rowptr[i + 1];
    for (int p = rowptr[i]; p < p1; ++p) axpy(s, val[p], x[colidx[p]]);
    Y[id] = s;
}

void csr_apply_batched(const int* rowptr, const int* colidx, const double* val, int n, const cdbl* X, cdbl* Y,
                       long long ncols, hipStream_t st) {
    const long long total = (long long)n * ncols;
    if (total <= 0) return;
    hipLaunchKernelGGL(k_csr_apply, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, rowptr, colidx, val, n, X,
                       Y, ncols);
}

__global__ __launch_bounds__(256) void k_csr_apply_multi(const CsrDesc* __restrict__ desc, int n_pad,
                                                         const cdbl* __restrict__ X, cdbl* __restrict__ Y, int ncols,
                                                         int batch) {
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long per = (long long)n_pad * ncols;
    if (id >= per * batch) return;
    const int b = (int)(id / per);
    const int i = (int)(id % n_pad);
    const CsrDesc d = desc[b];
    cdbl s = mk(0, 0);
    if (i < d.n) {
        const cdbl* x = X + (id - i);
        const int p1 = d.rp[i + 1];
        for (int p = d.rp[i]; p < p1; ++p) axpy(s, d.v[p], x[d.ci[p]]);
    }
    Y[id] = s;
}

void csr_apply_multi(const CsrDesc* desc, int n_pad, const cdbl* X, cdbl* Y, int ncols, int batch, hipStream_t st) {
    const long long total = (long long)n_pad * ncols * batch;
    if (total <= 0) return;
    hipLaunchKernelGGL(k_csr_apply_multi, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, desc, n_pad, X, Y,
                       ncols, batch);
}

__global__ __launch_bounds__(256) void k_replicate_padded(const cdbl* __restrict__ src, const long long* __restrict__ off,
                                                          const int* __restrict__ nrow, int n_pad, int nc, int batch,
                                                          cdbl* __restrict__ X) {
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long per = (long long)n_pad * nc;
    if (id >= per * batch) return;
    const int b = (int)(id / per);
    const long long r = id - (long long)b * per;
    const int c = (int)(r / n_pad), i = (int)(r % n_pad);
    const int nb = nrow[b];
    X[id] = i < nb ? src[off[b] + (long long)c * nb + i] : mk(0, 0);
}

void replicate_padded(const cdbl* src, const long long* off, const int* nrow, int n_pad, int nc, int batch, cdbl* X,
                      hipStream_t st) {
    const long long total = (long long)n_pad * nc * batch;
    if (total <= 0) return;
    hipLaunchKernelGGL(k_replicate_padded, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, src, off, nrow,
                       n_pad, nc, batch, X);
}

// dst[b * count + i] = src[i] for every matrix b of the batch
__global__ __launch_bounds__(256) void k_replicate(const cdbl* __restrict__ src, long long count, long long total,
                                                   cdbl* __restrict__ dst) {
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id < total) dst[id] = src[id % count];
}

void replicate_batched(const cdbl* src, long long count, int batch, cdbl* dst, hipStream_t st) {
    const long long total = count * batch;
    if (total <= 0) return;
    hipLaunchKernelGGL(k_replicate, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, src, count, total, dst);
}

// ---------------------------------------------------------------------------
// Block reductions
// ---------------------------------------------------------------------------
__device__ cdbl block_sum(cdbl v, cdbl* sh) {
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    for (int o = 32; o > 0; o >>= 1) { v.x += __shfl_xor(v.x, o, 64); v.y += __shfl_xor(v.y, o, 64); }
    __syncthreads();
    if (lane == 0) sh[wv] = v;
    __syncthreads();
    cdbl r = mk(0, 0);
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) r = r + sh[q];
    return r;
}

// Modified Gram-Schmidt, applied twice per column.
__global__ __launch_bounds__(256) void k_orth(cdbl* __restrict__ X, int n, int nc) {
    __shared__ cdbl sh[4];
    cdbl* x = X + (size_t)blockIdx.x * n * nc;
    const int tid = threadIdx.x;
    for (int c = 0; c < nc; ++c) {
        cdbl* xc = x + (size_t)c * n;
        for (int pass = 0; pass < 2; ++pass) {
            for (int c2 = 0; c2 < c; ++c2) {
                const cdbl* q = x + (size_t)c2 * n;
                cdbl part = mk(0, 0);
                for (int i = tid; i < n; i += blockDim.x) cfma(part, cconj(q[i]), xc[i]);
                const cdbl r = block_sum(part, sh);
                for (int i = tid; i < n; i += blockDim.x) cfms(xc[i], r, q[i]);
                __syncthreads();
            }
        }
        double part = 0.0;
        for (int i = tid; i < n; i += blockDim.x) part += cabs2(xc[i]);
        const cdbl nr = block_sum(mk(part, 0.0), sh);
        const double inv = nr.x > 0.0 ? 1.0 / sqrt(nr.x) : 0.0;
        for (int i = tid; i < n; i += blockDim.x) xc[i] = inv * xc[i];
        __syncthreads();
    }
}

void orthonormalize_thin_batched(cdbl* X, int n, int batch, int nc, hipStream_t st) {
    hipLaunchKernelGGL(k_orth, dim3(batch), dim3(256), 0, st, X, n, nc);
}

// Hermitian Jacobi eigensolver (cyclic, complex rotations) for nc x nc, one lane.
__device__ void herm_eig_max(cdbl* G, int m, double& lmax, cdbl* y) {
    cdbl Z[kMaxNc * kMaxNc];
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) Z[i + j * m] = mk(i == j ? 1.0 : 0.0, 0.0);
    for (int sweep = 0; sweep < 50; ++sweep) {
        double off = 0.0, diag = 0.0;
        for (int p = 0; p < m; ++p) {
            diag += G[p + p * m].x * G[p + p * m].x;
            for (int q = p + 1; q < m; ++q) off += cabs2(G[p + q * m]);
        }
        if (off <= 1e-32 * diag) break;
        for (int p = 0; p < m; ++p)
            for (int q = p + 1; q < m; ++q) {
                const cdbl apq = G[p + q * m];
                const double g = sqrt(cabs2(apq));
                if (g == 0.0) continue;
                const cdbl e = mk(apq.x / g, apq.y / g);
                const double app = G[p + p * m].x, aqq = G[q + q * m].x;
                const double tau = (aqq - app) / (2.0 * g);
                const double t = (tau >= 0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
                const double cs = 1.0 / sqrt(1.0 + t * t), sn = t * cs;
                // V = [[cs, sn e], [-sn conj(e), cs]] on columns p, q:  G <- V^H G V, Z <- Z V
                for (int k = 0; k < m; ++k) {   // G <- G V
                    const cdbl gp = G[k + p * m], gq = G[k + q * m];
                    G[k + p * m] = cs * gp - sn * (gq * cconj(e));
                    G[k + q * m] = sn * (gp * e) + cs * gq;
                    const cdbl zp = Z[k + p * m], zq = Z[k + q * m];
                    Z[k + p * m] = cs * zp - sn * (zq * cconj(e));
                    Z[k + q * m] = sn * (zp * e) + cs * zq;
                }
                for (int k = 0; k < m; ++k) {   // G <- V^H G
                    const cdbl gp = G[p + k * m], gq = G[q + k * m];
                    G[p + k * m] = cs * gp - sn * (e * gq);
                    G[q + k * m] = sn * (cconj(e) * gp) + cs * gq;
                }
                G[p + q * m] = mk(0, 0);
                G[q + p * m] = mk(0, 0);
            }
    }
    int best = 0;
    for (int i = 1; i < m; ++i) if (G[i + i * m].x > G[best + best * m].x) best = i;
    lmax = G[best + best * m].x;
    for (int i = 0; i < m; ++i) y[i] = Z[i + best * m];
}

// sigma = 1 / sigma_max(B); u = B y / sigma_max(B), v = Q y (y: top right singular vector of B)
__global__ __launch_bounds__(256) void k_sigma(const cdbl* __restrict__ Bm, const cdbl* __restrict__ Qm, int n, int nc,
                                               double* __restrict__ sigma, cdbl* __restrict__ U, cdbl* __restrict__ V) {
    __shared__ cdbl sh[4];
    __shared__ cdbl Gs[kMaxNc * kMaxNc];
    __shared__ cdbl ys[kMaxNc];
    __shared__ double smax;
    const int b = blockIdx.x, tid = threadIdx.x;
    const cdbl* B = Bm + (size_t)b * n * nc;
    for (int p = 0; p < nc; ++p)
        for (int q = p; q < nc; ++q) {
            cdbl part = mk(0, 0);
            for (int i = tid; i < n; i += blockDim.x) cfma(part, cconj(B[i + (size_t)p * n]), B[i + (size_t)q * n]);
            const cdbl r = block_sum(part, sh);
            if (tid == 0) { Gs[p + q * nc] = r; Gs[q + p * nc] = cconj(r); }
        }
    __syncthreads();
    if (tid == 0) {
        cdbl G[kMaxNc * kMaxNc], y[kMaxNc];
        for (int e = 0; e < nc * nc; ++e) G[e] = Gs[e];
        for (int p = 0; p < nc; ++p) G[p + p * nc].y = 0.0;
        double lmax;
        herm_eig_max(G, nc, lmax, y);
        const double s1 = sqrt(lmax);
        smax = s1;
        for (int i = 0; i < nc; ++i) ys[i] = y[i];
        sigma[b] = 1.0 / s1;
    }
    __syncthreads();
    if (U == nullptr) return;
    const cdbl* Q = Qm + (size_t)b * n * nc;
    const double inv = 1.0 / smax;
    for (int i = tid; i < n; i += blockDim.x) {
        cdbl u = mk(0, 0), v = mk(0, 0);
        for (int c = 0; c < nc; ++c) { cfma(u, B[i + (size_t)c * n], ys[c]); cfma(v, Q[i + (size_t)c * n], ys[c]); }
        U[(size_t)b * n + i] = inv * u;
        V[(size_t)b * n + i] = v;
    }
}

void sigma_from_thin_batched(const cdbl* B, const cdbl* Q, int n, int batch, int nc, double* sigma, cdbl* u, cdbl* v,
                             hipStream_t st) {
    hipLaunchKernelGGL(k_sigma, dim3(batch), dim3(256), 0, st, B, Q, n, nc, sigma, u, v);
}

// out = Re(a^H T b).  The columns of each matrix are split over kBilinearParts
// workgroups (enough loads in flight to stream T at the memory rate); one
// wavefront per column j forms (a^H T)_j b_j, each workgroup stores its
// partial sum and a second kernel adds the partials in a fixed order.

__global__ __launch_bounds__(256) void k_bilinear(const cdbl* __restrict__ T, long long tstride, const cdbl* __restrict__ a,
                                                  const cdbl* __restrict__ bv, int n, double* __restrict__ part) {
    __shared__ cdbl sh[4];
    const int b = blockIdx.y, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int cs = (n + kBilinearParts - 1) / kBilinearParts;
    const int jlo = blockIdx.x * cs, jhi = min(n, jlo + cs);
    const cdbl* Tm = T + (size_t)b * tstride;
    const cdbl* av = a + (size_t)b * n;
    const cdbl* bb = bv + (size_t)b * n;
    double acc = 0.0;
    for (int j = jlo + wv; j < jhi; j += 4) {
        cdbl s = mk(0, 0);
        for (int i = lane; i < n; i += 64) cfma(s, cconj(av[i]), Tm[i + (size_t)j * n]);
        for (int o = 32; o > 0; o >>= 1) { s.x += __shfl_xor(s.x, o, 64); s.y += __shfl_xor(s.y, o, 64); }
        const cdbl t = s * bb[j];
        acc += t.x;
    }
    const cdbl r = block_sum(mk(lane == 0 ? acc : 0.0, 0.0), sh);
    if (tid == 0) part[(size_t)b * kBilinearParts + blockIdx.x] = r.x;
}

__global__ void k_bilinear_sum(const double* __restrict__ part, int batch, double* __restrict__ out) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= batch) return;
    double s = 0.0;
    for (int q = 0; q < kBilinearParts; ++q) s += part[(size_t)b * kBilinearParts + q];
    out[b] = s;
}

void bilinear_re_batched(const cdbl* T, long long tstride, const cdbl* a, const cdbl* bvec, int n, int batch,
                         double* out, double* part, hipStream_t st) {
    if (batch <= 0) return;
    hipLaunchKernelGGL(k_bilinear, dim3(kBilinearParts, batch), dim3(256), 0, st, T, tstride, a, bvec, n, part);
    hipLaunchKernelGGL(k_bilinear_sum, dim3((batch + 255) / 256), dim3(256), 0, st, part, batch, out);
}

// y = T x or T^H x; one wavefront per output row (T^H: row i of T^H = column i of T)
__global__ __launch_bounds__(256) void k_gemv(const cdbl* __restrict__ T, long long tstride, const cdbl* __restrict__ x,
                                              cdbl* __restrict__ y, int n, int adj) {
    const int b = blockIdx.y;
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    const cdbl* Tm = T + (size_t)b * tstride;
    const cdbl* xv = x + (size_t)b * n;
    cdbl s = mk(0, 0);
    if (!adj) { for (int j = lane; j < n; j += 64) cfma(s, Tm[row + (size_t)j * n], xv[j]); }
    else { for (int j = lane; j < n; j += 64) cfma(s, cconj(Tm[j + (size_t)row * n]), xv[j]); }
    for (int o = 32; o > 0; o >>= 1) { s.x += __shfl_xor(s.x, o, 64); s.y += __shfl_xor(s.y, o, 64); }
    if (lane == 0) y[(size_t)b * n + row] = s;
}

void gemv_batched(const cdbl* T, long long tstride, const cdbl* x, cdbl* y, int n, int batch, bool adj, hipStream_t st) {
    hipLaunchKernelGGL(k_gemv, dim3((n + 3) / 4, batch), dim3(256), 0, st, T, tstride, x, y, n, adj ? 1 : 0);
}

__global__ void k_wielandt(const cdbl* __restrict__ T, int n, double s, cdbl* __restrict__ W) {
    const int m2 = 2 * n;
    long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long long)m2 * m2) return;
    const int i = (int)(id % m2), j = (int)(id / m2);
    cdbl v = mk(0, 0);
    if (i < n && j >= n) v = T[i + (size_t)(j - n) * n];
    else if (i >= n && j < n) v = cconj(T[j + (size_t)(i - n) * n]);
    if (i == j) v.x -= s;
    W[id] = v;
}

void wielandt_build(const cdbl* T, int n, double s, cdbl* W, hipStream_t st) {
    long long tot = 4LL * n * n;
    hipLaunchKernelGGL(k_wielandt, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, T, n, s, W);
}

// ---------------------------------------------------------------------------
// Pivot order of a column-pivoted Householder QR, as the reference's
// W.colPivHouseholderQr() in randomized_svd::sv_der / sv_der2
// (randsvd.cpp:160-162, 183-184), with the xGEQP3 rule: at step k the
// remaining column of largest squared norm (rows k..m-1 after the earlier
// reflections), the first slot on ties; the column norms are recomputed
// after every step.  Only the order is kept: the reference's vector, the last
// column of Q, is orthogonal to every column but the last pivot p, so for the
// Hermitian W - sigma I it is the direction of (W - sigma I)^{-1} e_p, which
// the caller forms with the LU.
// One launch per step, one workgroup per remaining slot of every matrix;
// slot order and norms are double-buffered by step parity, so no launch
// reads what it writes.  A (m x m per matrix, stride ms) is overwritten.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double block_sum(double v, double* red) {
    const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    __syncthreads();                       // red is free (earlier sums read)
    if (lane == 0) red[wv] = v;
    __syncthreads();
    return (red[0] + red[1]) + (red[2] + red[3]);
}

__global__ __launch_bounds__(256) void k_cp_norms(const cdbl* __restrict__ A, long long ms, int m,
                                                 double* __restrict__ nrm, int* __restrict__ slot) {
    __shared__ double red[4];
    const int c = blockIdx.x, b = blockIdx.y;
    const cdbl* col = A + (size_t)b * ms + (size_t)c * m;
    double v = 0.0;
    for (int i = threadIdx.x; i < m; i += 256) v += cabs2(col[i]);
    v = block_sum(v, red);
    if (threadIdx.x == 0) {
        nrm[(size_t)b * m + c] = v;
        slot[(size_t)b * m + c] = c;
    }
}

__global__ __launch_bounds__(256) void k_cp_step(cdbl* __restrict__ A, long long ms, int m, int k,
                                                const double* __restrict__ nin, double* __restrict__ nout,
                                                const int* __restrict__ sin_, int* __restrict__ sout,
                                                int* __restrict__ last) {
    __shared__ double red[4];
    __shared__ double bv[256];
    __shared__ int bs[256];
    const int tid = threadIdx.x, b = blockIdx.y;
    const int s = k + blockIdx.x;          // this workgroup's slot
    A += (size_t)b * ms;
    nin += (size_t)b * m;
    nout += (size_t)b * m;
    sin_ += (size_t)b * m;
    sout += (size_t)b * m;
    // pivot: the first slot of largest norm among k..m-1 (each thread scans
    // its slots in increasing order; the tree keeps the lower slot on ties)
    double best = -1.0;
    int bslot = m;
    for (int t = k + tid; t < m; t += 256) {
        const double v = nin[sin_[t]];
        if (v > best) { best = v; bslot = t; }
    }
    bv[tid] = best;
    bs[tid] = bslot;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if (tid < w) {
            const double v2 = bv[tid + w];
            const int s2 = bs[tid + w];
            if (v2 > bv[tid] || (v2 == bv[tid] && s2 < bs[tid])) { bv[tid] = v2; bs[tid] = s2; }
        }
        __syncthreads();
    }
    const int ps = bs[0] < m ? bs[0] : k;
    const int pc = sin_[ps];
    if (s == k) {
        // the pivot's workgroup: the next slot order (slots k and ps exchanged)
        for (int t = tid; t < m; t += 256) sout[t] = t == k ? pc : (t == ps ? sin_[k] : sin_[t]);
        if (k == m - 1 && tid == 0) last[b] = pc;
        return;
    }
    const int c = s == ps ? sin_[k] : sin_[s];
    const cdbl* pv = A + (size_t)pc * m;
    cdbl* cv = A + (size_t)c * m;
    // reflector of the pivot column, rows k..m-1: v = x - beta e_k,
    // beta = -(alpha / |alpha|) |x|, tau = 2 / |v|^2
    double xs = 0.0;
    for (int i = k + tid; i < m; i += 256) xs += cabs2(pv[i]);
    const double xn = sqrt(block_sum(xs, red));
    const cdbl alpha = pv[k];
    const double aa = sqrt(cabs2(alpha));
    const cdbl beta = aa == 0.0 ? mk(-xn, 0.0) : mk(-alpha.x / aa * xn, -alpha.y / aa * xn);
    const cdbl vk = alpha - beta;
    double vs = 0.0;
    for (int i = k + 1 + tid; i < m; i += 256) vs += cabs2(pv[i]);
    const double vn = cabs2(vk) + block_sum(vs, red);
    const double tau = vn == 0.0 ? 0.0 : 2.0 / vn;
    // s_c = tau v^H c; c -= s_c v over rows k..m-1
    double dr = 0.0, di = 0.0;
    for (int i = k + tid; i < m; i += 256) {
        const cdbl v = i == k ? vk : pv[i];
        const cdbl x = cv[i];
        dr += v.x * x.x + v.y * x.y;
        di += v.x * x.y - v.y * x.x;
    }
    const double sr = tau * block_sum(dr, red);
    const double si = tau * block_sum(di, red);
    double ns = 0.0;
    for (int i = k + tid; i < m; i += 256) {
        const cdbl v = i == k ? vk : pv[i];
        cdbl x = cv[i];
        x.x -= sr * v.x - si * v.y;
        x.y -= sr * v.y + si * v.x;
        cv[i] = x;
        if (i > k) ns += cabs2(x);
    }
    ns = block_sum(ns, red);
    if (tid == 0) nout[c] = ns;
}

void colpiv_last_pivot_batched(cdbl* A, long long ms, int m, int batch, double* nrm2, int* slot2, int* last,
                               hipStream_t st) {
    if (m <= 0 || batch <= 0) return;
    double* nb[2] = {nrm2, nrm2 + (size_t)m * batch};
    int* sb[2] = {slot2, slot2 + (size_t)m * batch};
    hipLaunchKernelGGL(k_cp_norms, dim3(m, batch), dim3(256), 0, st, (const cdbl*)A, ms, m, nb[0], sb[0]);
    for (int k = 0; k < m; ++k) {
        const int cur = k & 1;
        hipLaunchKernelGGL(k_cp_step, dim3(m - k, batch), dim3(256), 0, st, A, ms, m, k, (const double*)nb[cur],
                           nb[cur ^ 1], (const int*)sb[cur], sb[cur ^ 1], last);
    }
}

// ---------------------------------------------------------------------------
// The bordered Wielandt system of randomized_svd::sv_der2 (randsvd.cpp:171-216),
// batched over nb wavenumbers: x = [u; v] (the singular pair of T_p for sigma),
// m = the first index of the largest |x_l| (5 when none is positive, as the
// reference's search starts), u_n = x / x_m, and B = W - sigma I with column m
// replaced by -u_n, W = [[0, T_p], [T_p^H, 0]] (2n x 2n).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_border_prep(const cdbl* __restrict__ U, const cdbl* __restrict__ V, int n,
                                                     int* __restrict__ mv, cdbl* __restrict__ X) {
    __shared__ double sbest[4];
    __shared__ int sidx[4];
    __shared__ int sm;
    const int b = blockIdx.x, tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
    const int m2 = 2 * n;
    const cdbl* u = U + (size_t)b * n;
    const cdbl* v = V + (size_t)b * n;
    double best = 0.0;
    int bi = m2;
    for (int l = tid; l < m2; l += blockDim.x) {
        const double a = sqrt(cabs2(l < n ? u[l] : v[l - n]));
        if (a > best) { best = a; bi = l; }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        const double b2 = __shfl_xor(best, o, 64);
        const int i2 = __shfl_xor(bi, o, 64);
        if (b2 > best || (b2 == best && i2 < bi)) { best = b2; bi = i2; }
    }
    if (lane == 0) { sbest[wv] = best; sidx[wv] = bi; }
    __syncthreads();
    if (tid == 0) {
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w)
            if (sbest[w] > best || (sbest[w] == best && sidx[w] < bi)) { best = sbest[w]; bi = sidx[w]; }
        sm = bi < m2 ? bi : (5 < m2 ? 5 : 0);
        mv[b] = sm;
    }
    __syncthreads();
    const int m = sm;
    const cdbl xm = m < n ? u[m] : v[m - n];
    const double den = cabs2(xm);
    const cdbl inv = den > 0.0 ? mk(xm.x / den, -xm.y / den) : mk(0, 0);
    cdbl* x = X + (size_t)b * m2;
    for (int l = tid; l < m2; l += blockDim.x) x[l] = (l < n ? u[l] : v[l - n]) * inv;
}

// B of matrix b in 32 x 32 tiles; the T_p^H block is read through LDS so that
// both the reads of T_p and the writes of B run down columns.
__global__ __launch_bounds__(256) void k_border_build(const cdbl* __restrict__ Tp, int n, const double* __restrict__ sigma,
                                                      const int* __restrict__ mv, const cdbl* __restrict__ X,
                                                      cdbl* __restrict__ B) {
    __shared__ cdbl tile[32][33];
    const int m2 = 2 * n;
    const int nt = (m2 + 31) / 32;
    const int b = blockIdx.y;
    const int ti = blockIdx.x % nt, tj = blockIdx.x / nt;
    const int i0 = 32 * ti, j0 = 32 * tj;
    const cdbl* T = Tp + (size_t)b * n * n;
    cdbl* Bm = B + (size_t)b * m2 * m2;
    const int r = threadIdx.x & 31, c0 = threadIdx.x >> 5;   // 8 columns per pass
    const double s = sigma[b];
    const int m = mv[b];
    const cdbl* x = X + (size_t)b * m2;
    // lower-left block (i >= n, j < n): B(i, j) = conj(T_p(j, i - n)); stage T_p(j0 + r, i0 - n + c)
    // in the tile for every tile that meets the block (a tile may straddle row n or column n)
    const bool lowleft = i0 + 31 >= n && j0 < n;
    if (lowleft) {
        for (int c = c0; c < 32; c += 8) {
            const int jr = j0 + r, ic = i0 - n + c;
            tile[c][r] = (jr < n && ic >= 0 && ic < n) ? T[jr + (size_t)ic * n] : mk(0, 0);
        }
    }
    __syncthreads();
    for (int c = c0; c < 32; c += 8) {
        const int i = i0 + r, j = j0 + c;
        if (i >= m2 || j >= m2) continue;
        cdbl v;
        if (j == m) v = mk(-x[i].x, -x[i].y);
        else {
            if (i < n && j >= n) v = T[i + (size_t)(j - n) * n];
            else if (i >= n && j < n) v = cconj(tile[r][c]);
            else v = mk(0, 0);
            if (i == j) v.x -= s;
        }
        Bm[i + (size_t)j * m2] = v;
    }
}

// Y <- -Y (the eigenvector derivative u' = -B^{-1} p), t_b = u'_m, then u'_m = 0
__global__ __launch_bounds__(256) void k_border_negate_pick(cdbl* __restrict__ Y, int m2, int nb, const int* __restrict__ mv,
                                                            cdbl* __restrict__ t) {
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long long)m2 * nb) return;
    const int b = (int)(id / m2), l = (int)(id - (long long)b * m2);
    const cdbl y = Y[id];
    if (l == mv[b]) {
        t[b] = mk(-y.x, -y.y);
        Y[id] = mk(0, 0);
    } else {
        Y[id] = mk(-y.x, -y.y);
    }
}

// out = p2 + 2 (p - t u')
__global__ __launch_bounds__(256) void k_border_rhs(const cdbl* __restrict__ P, const cdbl* __restrict__ P2,
                                                   const cdbl* __restrict__ Ud, const cdbl* __restrict__ t, int m2, int nb,
                                                   cdbl* __restrict__ out) {
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long long)m2 * nb) return;
    const int b = (int)(id / m2);
    const cdbl q = P[id] - t[b] * Ud[id];
    out[id] = P2[id] + 2.0 * q;
}

// d2_b = -Re(y_m) of the second solve
__global__ void k_border_final(const cdbl* __restrict__ Y2, int m2, int nb, const int* __restrict__ mv,
                               double* __restrict__ d2) {
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= nb) return;
    d2[b] = -Y2[(size_t)b * m2 + mv[b]].x;
}

// dst[b dstride + doff + i] = src[b sstride + soff + i], i < len
__global__ __launch_bounds__(256) void k_copy_blocks(const cdbl* __restrict__ src, long long sstride, long long soff,
                                                     cdbl* __restrict__ dst, long long dstride, long long doff, int len,
                                                     int nb) {
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long long)len * nb) return;
    const int b = (int)(id / len), i = (int)(id - (long long)b * len);
    dst[b * dstride + doff + i] = src[b * sstride + soff + i];
}

static unsigned blocks_for(long long n) { return (unsigned)((n + 255) / 256); }

void border_prep_batched(const cdbl* U, const cdbl* V, int n, int nb, int* mv, cdbl* X, hipStream_t st) {
    if (nb <= 0) return;
    hipLaunchKernelGGL(k_border_prep, dim3(nb), dim3(256), 0, st, U, V, n, mv, X);
}

void border_build_batched(const cdbl* Tp, int n, int nb, const double* sigma, const int* mv, const cdbl* X, cdbl* B,
                          hipStream_t st) {
    if (nb <= 0) return;
    const unsigned nt = (unsigned)((2 * n + 31) / 32);
    hipLaunchKernelGGL(k_border_build, dim3(nt * nt, nb), dim3(256), 0, st, Tp, n, sigma, mv, X, B);
}

void border_negate_pick_batched(cdbl* Y, int n, int nb, const int* mv, cdbl* t, hipStream_t st) {
    const long long tot = 2LL * n * nb;
    if (tot <= 0) return;
    hipLaunchKernelGGL(k_border_negate_pick, dim3(blocks_for(tot)), dim3(256), 0, st, Y, 2 * n, nb, mv, t);
}

void border_rhs_batched(const cdbl* P, const cdbl* P2, const cdbl* Ud, const cdbl* t, int n, int nb, cdbl* out,
                        hipStream_t st) {
    const long long tot = 2LL * n * nb;
    if (tot <= 0) return;
    hipLaunchKernelGGL(k_border_rhs, dim3(blocks_for(tot)), dim3(256), 0, st, P, P2, Ud, t, 2 * n, nb, out);
}

void border_final_batched(const cdbl* Y2, int n, int nb, const int* mv, double* d2, hipStream_t st) {
    if (nb <= 0) return;
    hipLaunchKernelGGL(k_border_final, dim3(blocks_for(nb)), dim3(256), 0, st, Y2, 2 * n, nb, mv, d2);
}

void copy_blocks_batched(const cdbl* src, long long sstride, long long soff, cdbl* dst, long long dstride,
                         long long doff, int len, int nb, hipStream_t st) {
    const long long tot = (long long)len * nb;
    if (tot <= 0) return;
    hipLaunchKernelGGL(k_copy_blocks, dim3(blocks_for(tot)), dim3(256), 0, st, src, sstride, soff, dst, dstride, doff,
                       len, nb);
}

}  // namespace hbem
