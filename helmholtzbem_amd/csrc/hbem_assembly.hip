// hbem_assembly.hip -- Galerkin assembly of V, K, W (and d/dk, d2/dk2) on gfx950.
//
// Restates GalerkinBuilder (galerkin_builder.cpp:17-1018) for continuous P1
// test = trial spaces:
//   * every unordered panel pair (i >= j) is one of three classes
//     (coinciding i == j, adjacent = shared endpoint, general), assembled by
//     its own kernel launch;
//   * per pair and wavenumber the kernel produces the reference's local
//     interaction matrices (double layer direct + transposed block,
//     hypersingular, single layer; derivative orders 0..D) into a pair-indexed
//     buffer;
//   * a gather kernel sums, for each global entry, the (at most four) pair
//     blocks that touch it, in a fixed order -- deterministic, no atomics --
//     and applies the factor i/4 (galerkin_builder.cpp:1014-1017).
//
// Geometry is precomputed once per mesh into per-pair node tables
// {r, r^2, v.N_j |tau_i|, -v.N_i |tau_j|, tau_i.tau_j, |tau_i||tau_j|, 1/r,
// (2/pi) ln r, 1/sqrt(r)}, so the per-wavenumber kernels only evaluate the
// Hankel functions and accumulate.
//
// Two kernel families:
//   * k_assemble_x2 (batches of >= 64 wavenumber slots, the k-sweeps): a
//     wavefront takes one (pair, node chunk) and 32 slots, its two lane halves
//     evaluating the Hankel functions of two nodes and sharing them, with the
//     row-constant shape functions of general and adjacent pairs folded once
//     per row of the quadrature rule (see the comments at the kernel);
//   * k_assemble (single wavenumbers): work item = (pair, node chunk,
//     wavenumber); G lanes of a wavefront cooperate on one item, striding over
//     its nodes, and reduce with cross-lane shuffles; 64/G items share a
//     wavefront.
// The singular classes split their Nc^2 / 2 Nc^2 nodes into chunks whose
// partial local matrices are summed in chunk order by k_chunk_sum.
//
// Local slot layout per derivative order t (16 complex):
//   [0..3] K direct (test on panel i, trial on j), [4..7] K transposed block,
//   [8..11] W, [12..15] V; entry (m, n) at m + 2n, m/n = P1 local index
//   (b0(t) = t, b1(t) = 1 - t; continuous_space.hpp:115-120).
// Buffer layout: [pair][slot][wavenumber] -- the wavenumber index runs
// fastest, so the assembly stores (lanes over wavenumbers) and the gathers
// (threads over wavenumbers) are coalesced.  Chunk partials of the singular
// classes: [entry][chunk][slot][wavenumber].
#include "hbem_device.hpp"
#include "hbem_internal.hpp"

namespace hbem {

// ---------------------------------------------------------------------------
// Setup kernels (once per geometry)
// ---------------------------------------------------------------------------

// Per-panel point / tangent / |tangent| at the Gauss nodes (BuilderData's
// op_sg, op_tg, Derivative_01_sg/tg; galerkin_builder.cpp:167-186).
__global__ void k_setup_gauss_table(const double* __restrict__ panels, const double* __restrict__ coefs, int npan,
                                    const double* __restrict__ xg, int ns, double* __restrict__ gtab) {
    int id = blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= npan * ns) return;
    int p = id / ns, a = id % ns;
    cdbl z, tau;
    panel_eval(panels + 8 * p, coefs, xg[a], false, z, tau);
    double* o = gtab + (size_t)id * 5;
    o[0] = z.x; o[1] = z.y; o[2] = tau.x; o[3] = tau.y; o[4] = hypot(tau.x, tau.y);
}

// 1/r, (2/pi) ln r and 1/sqrt(r) of a node record (0 at r = 0)
__device__ __forceinline__ void node_r_factors(double* o) {
    const double r = o[0];
    o[6] = r > 0.0 ? 1.0 / r : 0.0;
    o[7] = r > 0.0 ? kTwoOverPi * log(r) : 0.0;
    o[8] = r > 0.0 ? 1.0 / sqrt(r) : 0.0;
}

// General pairs: node geometry on the Ns x Ns tensor grid (galerkin_builder.cpp:
// 308-327, 461-541) and the skip flags of the reference's isArrayZero tests
// (470-495, 852-870): bit0 all |v.N_j| < eps (direct K skipped), bit1 the
// transposed block, bit2 all |tau_i.tau_j| < eps.  One workgroup per pair.
__global__ void k_setup_general(const int* __restrict__ ent, int nent, const double* __restrict__ gtab, int ns,
                                double* __restrict__ ngeo, unsigned char* __restrict__ flags) {
    const int e = blockIdx.x;
    if (e >= nent) return;
    const int i = ent[3 * e], j = ent[3 * e + 1];
    const int nn = ns * ns;
    __shared__ int nz[3];
    if (threadIdx.x < 3) nz[threadIdx.x] = 0;
    __syncthreads();
    int loc[3] = {0, 0, 0};
    for (int nd = threadIdx.x; nd < nn; nd += blockDim.x) {
        const int a = nd / ns, b = nd % ns;
        const double* gi = gtab + ((size_t)i * ns + a) * 5;
        const double* gj = gtab + ((size_t)j * ns + b) * 5;
        const double vx = gi[0] - gj[0], vy = gi[1] - gj[1];
        const double r2 = vx * vx + vy * vy;
        const double vd = vx * gj[3] - vy * gj[2];
        const double vt = vy * gi[2] - vx * gi[3];
        const double td = gi[3] * gj[3] + gi[2] * gj[2];
        double* o = ngeo + ((size_t)e * nn + nd) * kNodeStride;
        o[0] = sqrt(r2); o[1] = r2; o[2] = vd * gi[4]; o[3] = vt * gj[4]; o[4] = td; o[5] = gi[4] * gj[4];
        node_r_factors(o);
        if (!(fabs(vd) < kDblEps)) loc[0] = 1;
        if (!(fabs(vt) < kDblEps)) loc[1] = 1;
        if (!(fabs(td) < kDblEps)) loc[2] = 1;
    }
    for (int q = 0; q < 3; ++q)
        if (loc[q]) atomicOr(&nz[q], 1);
    __syncthreads();
    if (threadIdx.x == 0) flags[e] = (nz[0] ? 0 : 1) | (nz[1] ? 0 : 2) | (nz[2] ? 0 : 4);
}

// Singular pairs: node geometry {r, r2, vd, vt, tdt, ttp, 1/r, (2/pi) ln r,
// 1/sqrt(r)} per node and skip flags.
// Coinciding (cls 0): one node set of nc^2 nodes; vd/vt hold the two
// triangles' double-layer factors (galerkin_builder.cpp:376-383).
// Adjacent (cls 1): two node sets (halves K = 0, 1) of nc^2 nodes each
// (galerkin_builder.cpp:267-306, 406-426).
// One workgroup per pair.
__global__ void k_setup_singular(int cls, const int* __restrict__ ent, int nent, const double* __restrict__ panels,
                                 const double* __restrict__ coefs, const double* __restrict__ xc, int nc,
                                 double* __restrict__ ngeo, unsigned char* __restrict__ flags) {
    const int e = blockIdx.x;
    if (e >= nent) return;
    const int i = ent[3 * e], j = ent[3 * e + 1];
    const bool swap = ent[3 * e + 2] != 0;
    const int nn = nc * nc;
    const int nhalf = cls == 0 ? 1 : 2;
    __shared__ int nz[6];
    if (threadIdx.x < 6) nz[threadIdx.x] = 0;
    __syncthreads();
    int loc[6] = {0, 0, 0, 0, 0, 0};
    for (int nd = threadIdx.x; nd < nhalf * nn; nd += blockDim.x) {
        const int h = nd / nn, g = nd % nn, a = g / nc, b = g % nc;
        cdbl zi, zj, ti, tj;
        if (cls == 0) {
            double t = xc[b], s = t * (1.0 - xc[a]);
            panel_eval(panels + 8 * i, coefs, s, false, zi, ti);
            panel_eval(panels + 8 * i, coefs, t, false, zj, tj);
        } else {
            double t = xc[a], s = t * xc[b];
            double si = h == 0 ? s : t, sj = h == 0 ? t : s;
            panel_eval(panels + 8 * i, coefs, si, !swap, zi, ti);
            panel_eval(panels + 8 * j, coefs, sj, swap, zj, tj);
        }
        const cdbl v = zi - zj;
        const double r2 = v.x * v.x + v.y * v.y;
        const double ni = hypot(ti.x, ti.y), nj = hypot(tj.x, tj.y);
        double vd = (v.x * tj.y - v.y * tj.x) * ni;
        double vt = (v.y * ti.x - v.x * ti.y) * nj;
        const double td = ti.y * tj.y + ti.x * tj.x;
        double* o = ngeo + ((size_t)e * nhalf * nn + nd) * kNodeStride;
        o[0] = sqrt(r2); o[1] = r2; o[2] = vd; o[3] = vt; o[4] = td; o[5] = ni * nj;
        node_r_factors(o);
        if (!(fabs(vd) < kDblEps)) loc[3 * h + 0] = 1;
        if (!(fabs(vt) < kDblEps)) loc[3 * h + 1] = 1;
        if (!(fabs(td) < kDblEps)) loc[3 * h + 2] = 1;
    }
    for (int q = 0; q < 6; ++q)
        if (loc[q]) atomicOr(&nz[q], 1);
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned char f = 0;
        if (cls == 0) {
            // coinciding: bit0 = triangle 0 K skipped, bit3 = triangle 1 K skipped, bit2 = tdottp zero
            if (!nz[0]) f |= 1;
            if (!nz[1]) f |= 8;
            if (!nz[2]) f |= 4;
        } else {
            for (int h = 0; h < 2; ++h) {
                if (!nz[3 * h + 0]) f |= (unsigned char)(1 << (3 * h));
                if (!nz[3 * h + 1]) f |= (unsigned char)(2 << (3 * h));
                if (!nz[3 * h + 2]) f |= (unsigned char)(4 << (3 * h));
            }
        }
        flags[e] = f;
    }
}

// ---------------------------------------------------------------------------
// Assembly kernels
// ---------------------------------------------------------------------------
template <int D>
struct Acc {
    cdbl kd[D + 1][4];   // sum cfd * {h1/r, h0, h1 r}
    cdbl kt[D + 1][4];   // sum cft * {h1/r, h0, h1 r}
    cdbl wa[D + 1];      // sum w * {h0, h1 r, h0 r^2}
    cdbl wb[D + 1][4];   // sum Bc * {h0, h1 r, h0 r^2}
    cdbl v[D + 1][4];    // sum g * {h0, h1 r, h0 r^2}
};

template <int D>
__device__ __forceinline__ void acc_zero(Acc<D>& A) {
#pragma unroll
    for (int t = 0; t <= D; ++t) {
        A.wa[t] = mk(0, 0);
#pragma unroll
        for (int q = 0; q < 4; ++q) { A.kd[t][q] = mk(0, 0); A.kt[t][q] = mk(0, 0); A.wb[t][q] = mk(0, 0); A.v[t][q] = mk(0, 0); }
    }
}

template <int G>
__device__ __forceinline__ double group_sum(double x) {
#pragma unroll
    for (int o = 1; o < G; o <<= 1) x += __shfl_xor(x, o, 64);
    return x;
}

template <int D, int G>
__device__ __forceinline__ void acc_reduce(Acc<D>& A) {
    if (G == 1) return;
#pragma unroll
    for (int t = 0; t <= D; ++t) {
        A.wa[t].x = group_sum<G>(A.wa[t].x); A.wa[t].y = group_sum<G>(A.wa[t].y);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            A.kd[t][q].x = group_sum<G>(A.kd[t][q].x); A.kd[t][q].y = group_sum<G>(A.kd[t][q].y);
            A.kt[t][q].x = group_sum<G>(A.kt[t][q].x); A.kt[t][q].y = group_sum<G>(A.kt[t][q].y);
            A.wb[t][q].x = group_sum<G>(A.wb[t][q].x); A.wb[t][q].y = group_sum<G>(A.wb[t][q].y);
            A.v[t][q].x = group_sum<G>(A.v[t][q].x);   A.v[t][q].y = group_sum<G>(A.v[t][q].y);
        }
    }
}

// One node's contribution.  Geometry: r, r2 (|v|, |v|^2), w (weight),
// bi/bj (P1 shape values of the test panel i / trial panel j), cd (K direct
// factor v.N_j |tau_i|, 0 when skipped), ct (transposed factor, 0 when
// skipped), bt (2 tau_i.tau_j, 0 when tdottp is zero), ttp (|tau_i||tau_j|).
// Per lane: kappa.  `sym` selects the coinciding form, where the two
// triangles of [0,1]^2 share the node (roles of s and t exchanged; cd2 is the
// second triangle's K factor) and only the entries m >= n of W and V are kept.
// H0, H1 at kappa r: real positive kappa (the reference's Hankel1Real01
// path) or any other kappa (its complex path)
template <bool TINY>
__device__ __forceinline__ void hankel_at(double kap, double r, cdbl& h0, cdbl& h1) {
    double j0, y0, j1, y1;
    bessel_jy01<TINY>(kap * r, j0, y0, j1, y1);
    h0 = mk(j0, y0);
    h1 = mk(j1, y1);
}
template <bool TINY>
__device__ __forceinline__ void hankel_at(cdbl kap, double r, cdbl& h0, cdbl& h1) { hankel01_cplx(r * kap, h0, h1); }

template <int D, typename KT, bool TINY>
__device__ __forceinline__ void node_accum(Acc<D>& A, KT kap, double r, double r2, double w, const double bi[2],
                                           const double bj[2], double cd, double ct, double bt, double ttp, bool sym,
                                           double cd2) {
    cdbl h0, h1;
    hankel_at<TINY>(kap, r, h0, h1);
    const cdbl hr = r * h1;
    const double rinv = r > 0.0 ? 1.0 / r : 0.0;
    const cdbl h1r = rinv * h1;
    const cdbl h0r2 = r2 * h0;
    double pbb[4], pbt[4];
#pragma unroll
    for (int n = 0; n < 2; ++n)
#pragma unroll
        for (int m = 0; m < 2; ++m) {
            pbb[m + 2 * n] = bi[m] * bj[n];   // test m at panel-i point, trial n at panel-j point
            pbt[m + 2 * n] = bj[m] * bi[n];   // exchanged roles
        }
    const double wcd = w * cd, wct = w * ct, wbt = w * bt, wtt = w * ttp;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        double ckd, ckt, cb, cg;
        if (sym) {
            ckd = wcd * pbb[q] + w * cd2 * pbt[q];
            ckt = 0.0;
            cb = wbt * (pbb[q] + pbt[q]);
            cg = wtt * (pbb[q] + pbt[q]);
        } else {
            ckd = wcd * pbb[q];
            ckt = wct * pbt[q];
            cb = wbt * pbb[q];
            cg = wtt * pbb[q];
        }
        const bool wv = !sym || q != 2;      // coinciding: entry (0,1) of W, V is copied afterwards
        axpy(A.kd[0][q], ckd, h1r);
        if (!sym) axpy(A.kt[0][q], ckt, h1r);
        if (wv) { axpy(A.wb[0][q], cb, h0); axpy(A.v[0][q], cg, h0); }
        if (D >= 1) {
            axpy(A.kd[1][q], ckd, h0);
            if (!sym) axpy(A.kt[1][q], ckt, h0);
            if (wv) { axpy(A.wb[1][q], cb, hr); axpy(A.v[1][q], cg, hr); }
        }
        if (D >= 2) {
            axpy(A.kd[2][q], ckd, hr);
            if (!sym) axpy(A.kt[2][q], ckt, hr);
            if (wv) { axpy(A.wb[2][q], cb, h0r2); axpy(A.v[2][q], cg, h0r2); }
        }
    }
    const double wa = sym ? 2.0 * w : w;
    axpy(A.wa[0], wa, h0);
    if (D >= 1) axpy(A.wa[1], wa, hr);
    if (D >= 2) axpy(A.wa[2], wa, h0r2);
}

// Accumulators -> the reference's local matrices (scaled as at the end of
// double_layer_*, hypersingular_*, single_layer_*; before the factor i/4).
// Linear in the accumulators, so partial sums over node chunks may be summed
// afterwards.
template <int D, typename KT>
__device__ __forceinline__ void acc_store(const Acc<D>& A, KT kap, KT kr, double c, bool coinc, int qflags,
                                          cdbl* __restrict__ dst, long long stride) {
    cdbl out[16 * (D + 1)];
    const double sqc = sqrt(c);
    const KT k2h = 0.5 * kap * kap, kinv = recip(kap);
    const double dd[4] = {1.0, -1.0, -1.0, 1.0};   // b'_m b'_n for P1 (+1, -1)
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        cdbl K0 = kap * A.kd[0][q], Kt0 = kap * A.kt[0][q];
        cdbl W0 = dd[q] * A.wa[0] - k2h * A.wb[0][q];
        cdbl V0 = A.v[0][q];
        out[q] = K0; out[4 + q] = Kt0; out[8 + q] = W0; out[12 + q] = V0;
        if (D >= 1) {
            cdbl K1 = (kr * c) * A.kd[1][q], Kt1 = (kr * c) * A.kt[1][q];
            cdbl W1 = (-sqc) * (dd[q] * A.wa[1] - k2h * A.wb[1][q] + kap * A.wb[0][q]);
            cdbl V1 = (-sqc) * A.v[1][q];
            out[16 + q] = K1; out[20 + q] = Kt1; out[24 + q] = W1; out[28 + q] = V1;
        }
        if (D >= 2) {
            cdbl K2 = c * (A.kd[1][q] - kap * A.kd[2][q]), Kt2 = c * (A.kt[1][q] - kap * A.kt[2][q]);
            cdbl W2 = (-c) * (dd[q] * (A.wa[2] - kinv * A.wa[1]) - k2h * (A.wb[2][q] - kinv * A.wb[1][q]) -
                              (2.0 * kap) * A.wb[1][q] + A.wb[0][q]);
            cdbl V2 = c * (kinv * A.v[1][q] - A.v[2][q]);
            out[32 + q] = K2; out[36 + q] = Kt2; out[40 + q] = W2; out[44 + q] = V2;
        }
    }
    if (coinc) {
        // entry (0,1) = entry (1,0) (galerkin_builder.cpp:571-573, 744-746)
#pragma unroll
        for (int t = 0; t <= D; ++t) {
            out[16 * t + 8 + 2] = out[16 * t + 8 + 1];
            out[16 * t + 12 + 2] = out[16 * t + 12 + 1];
            out[16 * t + 4 + 0] = mk(0, 0); out[16 * t + 4 + 1] = mk(0, 0);
            out[16 * t + 4 + 2] = mk(0, 0); out[16 * t + 4 + 3] = mk(0, 0);
        }
        if (qflags & 1) {
            // reference quirk: the d = 0 sum of entry (1,0), scaled for order t
            if (D >= 1) out[16 + 12 + 2] = (-sqc) * out[12 + 1];
            if (D >= 2) out[32 + 12 + 2] = c * out[12 + 1];
        }
    }
#pragma unroll
    for (int i = 0; i < 16 * (D + 1); ++i) dst[(size_t)i * stride] = out[i];
}

__device__ __forceinline__ int wave_uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// Inputs of one quadrature node: the six geometry values of the node table
// and the two rule points / weights it combines.
struct NodeIn {
    double o[6];
    double xa, xb, wa, wb;
};

__device__ __forceinline__ void fetch_node(NodeIn& n, const double* __restrict__ geo, int nd,
                                           const double* __restrict__ x, const double* __restrict__ w, int a, int b) {
    const double* p = geo + (size_t)nd * kNodeStride;
#pragma unroll
    for (int i = 0; i < 6; ++i) n.o[i] = p[i];
    n.xa = x[a]; n.xb = x[b]; n.wa = w[a]; n.wb = w[b];
}

__device__ __forceinline__ void load_k(const AsmParams& P, int q, double& kap, double& kr) {
    kap = P.kap[q];
    kr = P.kr[q];
}
__device__ __forceinline__ void load_k(const AsmParams& P, int q, cdbl& kap, cdbl& kr) {
    kap = mk(P.kap[q], P.kapi[q]);
    kr = mk(P.kr[q], P.kri[q]);
}

// cls: 0 coinciding, 1 adjacent, 2 general; KT: double (real positive
// wavenumbers) or cdbl (all others)
template <int CLS, int D, int G, typename KT>
__global__ __launch_bounds__(256) void k_assemble(AsmParams P) {
    const int lane = threadIdx.x & 63;
    const int g = lane % G;
    int e, chunk, q;
    bool active;
    if (G == 1) {
        // Waves enumerate (pair, chunk, block of 64 slots); lanes run over the
        // slots.  The wavefront shares (pair, chunk), so the node geometry is
        // read with wave-uniform addresses.
        const int wave = wave_uniform((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
        const int nqb = (P.nk + 63) >> 6;
        const int qb = wave % nqb;
        const int rest = wave / nqb;
        chunk = rest % P.nchunk;
        e = rest / P.nchunk;
        q = qb * 64 + lane;
        active = e < P.nent && q < P.nk;
        if (e >= P.nent) { e = P.nent - 1; chunk = 0; }
        if (q >= P.nk) q = P.nk - 1;
    } else {
        const long long wave = ((long long)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
        const long long nitems = (long long)P.nent * P.nchunk * P.nk;
        long long item = wave * (64 / G) + lane / G;
        active = item < nitems;
        if (!active) item = nitems - 1;
        q = (int)(item % P.nk);
        const long long rest = item / P.nk;
        chunk = (int)(rest % P.nchunk);
        e = (int)(rest / P.nchunk);
    }
    KT kap, kr;
    load_k(P, q, kap, kr);
    const unsigned f = P.flags[e];
    const int nn = CLS == 2 ? P.ns * P.ns : (CLS == 1 ? 2 * P.nc * P.nc : P.nc * P.nc);
    const int lo = chunk * P.chunk;
    const int hi = min(nn, lo + P.chunk);
    const double* __restrict__ geo = P.ngeo + (size_t)e * nn * kNodeStride;

    Acc<D> A;
    acc_zero<D>(A);

    // The node loop is software-pipelined: the tables of node nd + G are
    // loaded before node nd is evaluated, so their latency hides behind one
    // node's arithmetic.
    if (CLS == 2) {
        const double fd = (f & 1) ? 0.0 : 1.0, ft = (f & 2) ? 0.0 : 1.0, fb = (f & 4) ? 0.0 : 2.0;
        const int ns = P.ns;
        NodeIn cur, nxt;
        int nd = lo + g;
        if (nd < hi) fetch_node(cur, geo, nd, P.xg, P.wg, nd / ns, nd % ns);
        for (; nd < hi; nd += G) {
            if (nd + G < hi) fetch_node(nxt, geo, nd + G, P.xg, P.wg, (nd + G) / ns, (nd + G) % ns);
            const double s = cur.xa, t = cur.xb;
            const double w = cur.wa * cur.wb;
            const double bi[2] = {s, 1.0 - s}, bj[2] = {t, 1.0 - t};
            node_accum<D, KT, false>(A, kap, cur.o[0], cur.o[1], w, bi, bj, fd * cur.o[2], ft * cur.o[3],
                                     fb * cur.o[4], cur.o[5], false, 0.0);
            cur = nxt;
        }
    } else if (CLS == 1) {
        const bool swap = P.ent[3 * e + 2] != 0;
        const int nc = P.nc, nh = nc * nc;
        NodeIn cur, nxt;
        int nd = lo + g;
        if (nd < hi) { const int gg = nd % nh; fetch_node(cur, geo, nd, P.xc, P.wc, gg / nc, gg % nc); }
        for (; nd < hi; nd += G) {
            if (nd + G < hi) { const int gg = (nd + G) % nh; fetch_node(nxt, geo, nd + G, P.xc, P.wc, gg / nc, gg % nc); }
            const int h = nd >= nh ? 1 : 0;
            const double t = cur.xa, s = t * cur.xb;
            const double w = t * cur.wa * cur.wb;
            const double si = h == 0 ? s : t, sj = h == 0 ? t : s;
            // test panel i: forward when start(i) is the shared vertex (swap), else reversed; trial j opposite
            const double bi[2] = {swap ? si : 1.0 - si, swap ? 1.0 - si : si};
            const double bj[2] = {swap ? 1.0 - sj : sj, swap ? sj : 1.0 - sj};
            const unsigned fh = f >> (3 * h);
            const double cdv = (fh & 1) ? 0.0 : cur.o[2];
            const double ctv = (fh & 2) ? 0.0 : cur.o[3];
            const double bt = (fh & 4) ? 0.0 : 2.0 * cur.o[4];
            node_accum<D, KT, true>(A, kap, cur.o[0], cur.o[1], w, bi, bj, cdv, ctv, bt, cur.o[5], false, 0.0);
            cur = nxt;
        }
    } else {
        const int nc = P.nc;
        const double f0 = (f & 1) ? 0.0 : 1.0, f1 = (f & 8) ? 0.0 : 1.0, fb = (f & 4) ? 0.0 : 2.0;
        NodeIn cur, nxt;
        int nd = lo + g;
        // xa = xc[b], xb = xc[a] (t = xc[b], s = t (1 - xc[a]))
        if (nd < hi) fetch_node(cur, geo, nd, P.xc, P.wc, nd % nc, nd / nc);
        for (; nd < hi; nd += G) {
            if (nd + G < hi) fetch_node(nxt, geo, nd + G, P.xc, P.wc, (nd + G) % nc, (nd + G) / nc);
            const double t = cur.xa, s = t * (1.0 - cur.xb);
            const double w = t * cur.wa * cur.wb;
            const double bi[2] = {s, 1.0 - s}, bj[2] = {t, 1.0 - t};
            node_accum<D, KT, true>(A, kap, cur.o[0], cur.o[1], w, bi, bj, f0 * cur.o[2], 0.0, fb * cur.o[4],
                                    cur.o[5], true, f1 * cur.o[3]);
            cur = nxt;
        }
    }

    acc_reduce<D, G>(A);
    if (active && g == 0) {
        const size_t rec = P.nchunk == 1 ? (size_t)P.pidx[e] : (size_t)e * P.nchunk + chunk;
        cdbl* out = P.out + rec * (16 * (D + 1)) * P.nk + q;
        acc_store<D, KT>(A, kap, kr, P.cv[q], CLS == 0, P.qflags, out, P.nk);
    }
}

// ---------------------------------------------------------------------------
// Paired-lane assembly for batches of wavenumbers.
//
// A wavefront takes one (pair, node chunk) and 32 wavenumber slots; lane
// j + 32 h (h = 0, 1) serves slot j.  The workgroup's four wavefronts share
// the (pair, chunk) and stage its node records in LDS, 256 nodes at a time,
// with the kappa-free factors of the Hankel arguments (1/r, ln r, 1/sqrt r)
// and the quadrature weight products computed once per node there.  Nodes are
// taken two at a time: lane half h evaluates J0, Y0, J1, Y1 at node p + h,
// and v_permlane32_swap hands both halves both nodes' values.  Half h keeps
// the sums of the shape products with trial function n = h (entries q = m + 2h,
// m = 0, 1), so a lane holds 9 (D + 1) complex sums instead of 17 (D + 1).
// ---------------------------------------------------------------------------
constexpr int kRec = 22;     // doubles per staged node record
constexpr int kStage = 256;  // nodes per staging round
// general and adjacent pairs: 16-double records, 128 nodes per round (their
// LDS leaves room for the row-folded sums, see row_fold)
template <int CLS> constexpr int rec_len() { return CLS >= 1 ? 16 : kRec; }
template <int CLS> constexpr int stage_len() { return CLS >= 1 ? 128 : kStage; }

// record: 0 r, 1 r^2, 2 1/r, 3 (2/pi) ln r, 4 1/sqrt r, 5 weight of the curl sum,
//         6 + 8h + {0,1} K direct, {2,3} K transposed, {4,5} W, {6,7} V coefficient
//         of the entry q = m + 2h (m = 0, 1): quadrature weight x geometry factor x
//         shape product, the per-node factors of node_accum for lane half h.
// General pairs (tensor Gauss rule, node a * ns + b at s = x_a on panel i,
// t = x_b on panel j): the test shape b_m(s) of K direct, W and V is constant
// along a row a, so those records hold the coefficient without it --
//         6 + 5h + 0 K direct, 1 W, 2 V (weight x geometry x b_h(t)),
//         3 + {0,1} K transposed of the entry q = m + 2h as above --
// and the kernel applies b_m(s) once per row (row_fold).
// Adjacent pairs (node half hn, row a at t = xc_a): in half 1 the test point
// is t, so b_m is constant along a row as for general pairs; in half 0 the
// trial point is t, so there lane half h takes the test shape b_h(s) instead
// (it holds entries (h, n), exchanged into (m, h) at the end) and b_n(t) is
// folded per row.
template <int CLS>
__device__ __forceinline__ void stage_node(const AsmParams& P, const double* __restrict__ geo, unsigned f, bool swap,
                                           int nd, double* __restrict__ rec) {
    const double* o = geo + (size_t)nd * kNodeStride;
    double w, bi[2], bj[2], cd, ct, bt;
    int hn = 0;
    if (CLS == 2) {
        const int ns = P.ns, a = nd / ns, b = nd - a * ns;
        const double s = P.xg[a], t = P.xg[b];
        w = P.wg[a] * P.wg[b];
        bi[0] = s; bi[1] = 1.0 - s; bj[0] = t; bj[1] = 1.0 - t;
        cd = (f & 1) ? 0.0 : o[2];
        ct = (f & 2) ? 0.0 : o[3];
        bt = (f & 4) ? 0.0 : 2.0 * o[4];
    } else if (CLS == 1) {
        const int nc = P.nc, nh = nc * nc;
        hn = nd >= nh ? 1 : 0;
        const int gg = nd - hn * nh, a = gg / nc, b = gg - a * nc;
        const double t = P.xc[a], s = t * P.xc[b];
        w = t * P.wc[a] * P.wc[b];
        const double si = hn == 0 ? s : t, sj = hn == 0 ? t : s;
        bi[0] = swap ? si : 1.0 - si; bi[1] = swap ? 1.0 - si : si;
        bj[0] = swap ? 1.0 - sj : sj; bj[1] = swap ? sj : 1.0 - sj;
        const unsigned fh = f >> (3 * hn);
        cd = (fh & 1) ? 0.0 : o[2];
        ct = (fh & 2) ? 0.0 : o[3];
        bt = (fh & 4) ? 0.0 : 2.0 * o[4];
    } else {
        const int nc = P.nc, a = nd / nc, b = nd - a * nc;
        const double t = P.xc[b], s = t * (1.0 - P.xc[a]);
        w = t * P.wc[a] * P.wc[b];
        bi[0] = s; bi[1] = 1.0 - s; bj[0] = t; bj[1] = 1.0 - t;
        cd = (f & 1) ? 0.0 : o[2];
        ct = (f & 8) ? 0.0 : o[3];     // second triangle's K factor
        bt = (f & 4) ? 0.0 : 2.0 * o[4];
    }
    const double r = o[0];
    rec[0] = r;
    rec[1] = o[1];
    rec[2] = o[6];
    rec[3] = o[7];
    rec[4] = o[8];
    rec[5] = CLS == 0 ? 2.0 * w : w;
    const double wcd = w * cd, wct = w * ct, wbt = w * bt, wtt = w * o[5];
    if (CLS >= 1) {
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            double* c = rec + 6 + 5 * h;
            const double own = (CLS == 1 && hn == 0) ? bi[h] : bj[h];   // the shape not folded per row
            c[0] = wcd * own;
            c[1] = wbt * own;
            c[2] = wtt * own;
            c[3] = wct * (bj[0] * bi[h]);
            c[4] = wct * (bj[1] * bi[h]);
        }
        return;
    }
#pragma unroll
    for (int h = 0; h < 2; ++h)
#pragma unroll
        for (int m = 0; m < 2; ++m) {
            const double pbb = bi[m] * bj[h];   // test m at the panel-i point, trial h at the panel-j point
            const double pbt = bj[m] * bi[h];   // exchanged roles
            double* c = rec + 6 + 8 * h;
            if (CLS == 0) {
                c[m] = wcd * pbb + wct * pbt;
                c[2 + m] = 0.0;
                c[4 + m] = wbt * (pbb + pbt);
                c[6 + m] = wtt * (pbb + pbt);
            } else {
                c[m] = wcd * pbb;
                c[2 + m] = wct * pbt;
                c[4 + m] = wbt * pbb;
                c[6 + m] = wtt * pbb;
            }
        }
}

// padding record for an odd node count: finite argument, zero weights
template <int CLS>
__device__ __forceinline__ void stage_pad(double* rec) {
    rec[0] = 1.0; rec[1] = 1.0; rec[2] = 1.0; rec[3] = 0.0; rec[4] = 1.0;
#pragma unroll
    for (int i = 5; i < rec_len<CLS>(); ++i) rec[i] = 0.0;
}

// both lane halves get the value of the lower half (v0) and of the upper half (v1)
__device__ __forceinline__ void halves(double v, double& v0, double& v1) {
    const unsigned lo = (unsigned)__double2loint(v), hi = (unsigned)__double2hiint(v);
    const auto a = __builtin_amdgcn_permlane32_swap(lo, lo, false, false);
    const auto b = __builtin_amdgcn_permlane32_swap(hi, hi, false, false);
    v0 = __hiloint2double((int)b[0], (int)a[0]);
    v1 = __hiloint2double((int)b[1], (int)a[1]);
}

template <int D>
struct AccX {
    cdbl kd[D + 1][2], kt[D + 1][2], wb[D + 1][2], v[D + 1][2], wa[D + 1];
};

// J0, Y0, J1, Y1 of x = kappa r from the node's kappa-free factors.
//   y = x^2 / 64, L2 = (2/pi) ln(x/2) = lk + lr, 1/x = kinv / r,
//   1/sqrt(pi x) = ksq / sqrt(r)
// TINY >= 1 (touching panel pairs, whose arguments are small): x < HBEM_TINY_X
// takes the entire parts from the degree-HBEM_TINY_DEG table fitted on that
// range instead of the degree-HBEM_SMALL_DEG table of (0, 8); TINY = 2 takes
// x < HBEM_TINY1_X from the degree-HBEM_TINY1_DEG table of that range first.
// sin / cos beyond x = 1e6 (never reached by the benchmark's arguments): kept
// out of line so the library routine's constants do not occupy registers of
// the callers' loops
__device__ __noinline__ void sincos_far(double x, double& s, double& c) { sincos(x, &s, &c); }

template <int TINY>
__device__ __forceinline__ void bessel_jy01_f(double x, double y, double L2, double xinv, double b, double& j0,
                                              double& y0, double& j1, double& y1) {
    if (x < 8.0) {
        double A, B, C, Dd;
        if (TINY >= 2 && x < (double)HBEM_TINY1_X)
            horner4<HBEM_TINY1_DEG>(bessel_table() + HBEM_TAB_TINY1, y, A, B, C, Dd);
        else if (TINY >= 1 && x < (double)HBEM_TINY_X)
            horner4<HBEM_TINY_DEG>(bessel_table() + HBEM_TAB_TINY, y, A, B, C, Dd);
        else
            horner4<HBEM_SMALL_DEG>(bessel_table() + HBEM_TAB_SMALL, y, A, B, C, Dd);
        const bool pos = x > 0.0;
        const double J1 = x * B;
        j0 = pos ? A : 1.0;
        j1 = pos ? J1 : 0.0;
        y0 = pos ? fma(L2, A, C) : 0.0;
        y1 = pos ? fma(L2, J1, fma(x, Dd, -kTwoOverPi * xinv)) : 0.0;
    } else if (x <= 1e6) {
        double s, c;
        sincos_med(x, s, c);
        const double cs = c + s, sc = s - c;
        const double z = 64.0 * xinv * xinv;
        double P0, Q0, P1, Q1;
        horner4<HBEM_PQ_DEG>(bessel_table() + HBEM_TAB_PQ, z, P0, Q0, P1, Q1);
        Q0 *= xinv;
        Q1 *= xinv;
        j0 = b * fma(P0, cs, -Q0 * sc);
        y0 = b * fma(P0, sc, Q0 * cs);
        j1 = b * fma(P1, sc, Q1 * cs);
        y1 = b * fma(Q1, sc, -P1 * cs);
    } else {
        double s, c;
        sincos_far(x, s, c);
        const double cs = c + s, sc = s - c;
        j0 = b * cs; y0 = b * sc; j1 = b * sc; y1 = -b * cs;
    }
}

// one node's contribution to this lane's sums (entries q = m + 2h).  General
// and adjacent pairs: K direct, W and V go to the running row-free sums U in
// slot m = 1 (see row_fold); K transposed to both entries as for coinciding
// pairs.
template <int CLS, int D>
__device__ __forceinline__ void node_accum_x(AccX<D>& A, const double* __restrict__ R, int h, double j0, double y0,
                                             double j1, double y1) {
    const double r = R[0], r2 = R[1], rinv = R[2], wa = R[5];
    const cdbl h0 = mk(j0, y0);
    const cdbl h1r = mk(j1 * rinv, y1 * rinv);
    const cdbl hr = mk(j1 * r, y1 * r);
    const cdbl h0r2 = mk(j0 * r2, y0 * r2);
    const double* __restrict__ c = R + 6 + (CLS >= 1 ? 5 : 8) * h;
    if (CLS >= 1) {
        const double ukd = c[0], ub = c[1], ug = c[2];
        axpy(A.kd[0][1], ukd, h1r);
        axpy(A.wb[0][1], ub, h0);
        axpy(A.v[0][1], ug, h0);
        if (D >= 1) { axpy(A.kd[1][1], ukd, h0); axpy(A.wb[1][1], ub, hr); axpy(A.v[1][1], ug, hr); }
        if (D >= 2) { axpy(A.kd[2][1], ukd, hr); axpy(A.wb[2][1], ub, h0r2); axpy(A.v[2][1], ug, h0r2); }
#pragma unroll
        for (int m = 0; m < 2; ++m) {
            const double ckt = c[3 + m];
            axpy(A.kt[0][m], ckt, h1r);
            if (D >= 1) axpy(A.kt[1][m], ckt, h0);
            if (D >= 2) axpy(A.kt[2][m], ckt, hr);
        }
        axpy(A.wa[0], wa, h0);
        if (D >= 1) axpy(A.wa[1], wa, hr);
        if (D >= 2) axpy(A.wa[2], wa, h0r2);
        return;
    }
#pragma unroll
    for (int m = 0; m < 2; ++m) {
        const double ckd = c[m], ckt = c[2 + m], cb = c[4 + m], cg = c[6 + m];
        axpy(A.kd[0][m], ckd, h1r);
        if (CLS != 0) axpy(A.kt[0][m], ckt, h1r);
        axpy(A.wb[0][m], cb, h0);
        axpy(A.v[0][m], cg, h0);
        if (D >= 1) {
            axpy(A.kd[1][m], ckd, h0);
            if (CLS != 0) axpy(A.kt[1][m], ckt, h0);
            axpy(A.wb[1][m], cb, hr);
            axpy(A.v[1][m], cg, hr);
        }
        if (D >= 2) {
            axpy(A.kd[2][m], ckd, hr);
            if (CLS != 0) axpy(A.kt[2][m], ckt, hr);
            axpy(A.wb[2][m], cb, h0r2);
            axpy(A.v[2][m], cg, h0r2);
        }
    }
    axpy(A.wa[0], wa, h0);
    if (D >= 1) axpy(A.wa[1], wa, hr);
    if (D >= 2) axpy(A.wa[2], wa, h0r2);
}

// General and adjacent pairs, end of row a of the node chunk.  With U_a the
// running sum of a term after row a (from the start of the chunk) and
// T_a = U_a - U_{a-1} the chunk's part of row a, the row-constant shape
// f(a) = b_0 gives sum_a f(a) T_a = sum_{a < a1} (f(a) - f(a+1)) U_a + f(a1) U_a1
// (a1 the chunk's last row): one update per row, where coinciding pairs pay one
// per node.  These nine sums V (K direct, W, V for every order) are touched
// once per row, so they live in LDS (vacc, one column per thread) and leave
// the registers to the node loop.  b_1 = 1 - b_0 is then U - V (row_finish).
template <int D>
__device__ __forceinline__ void row_fold(const AccX<D>& A, double delta, cdbl* vacc) {
    const int t0 = threadIdx.x;
#pragma unroll
    for (int t = 0; t <= D; ++t) {
        cdbl* v = vacc + (3 * t) * 256 + t0;
        cdbl a = v[0], b = v[256], c = v[512];
        axpy(a, delta, A.kd[t][1]);
        axpy(b, delta, A.wb[t][1]);
        axpy(c, delta, A.v[t][1]);
        v[0] = a; v[256] = b; v[512] = c;
    }
}
template <int D>
__device__ __forceinline__ void row_finish(AccX<D>& A, const cdbl* vacc) {
    const int t0 = threadIdx.x;
#pragma unroll
    for (int t = 0; t <= D; ++t) {
        const cdbl* v = vacc + (3 * t) * 256 + t0;
        A.kd[t][0] = v[0];
        A.wb[t][0] = v[256];
        A.v[t][0] = v[512];
        A.kd[t][1] = A.kd[t][1] - A.kd[t][0];
        A.wb[t][1] = A.wb[t][1] - A.wb[t][0];
        A.v[t][1] = A.v[t][1] - A.v[t][0];
    }
}

// cls: 0 coinciding, 1 adjacent, 2 general; blockDim 256.  Three waves per
// SIMD (at most 168 VGPRs): at d = 2 the register allocator then spills a few
// loop-invariant values, and the third wave still hides more latency than the
// spill reloads cost (about 3% faster than two waves without spills).
template <int CLS, int D>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(3, 3))) void k_assemble_x2(AsmParams P) {
    constexpr int REC = rec_len<CLS>(), STG = stage_len<CLS>();
    __shared__ double lds[(STG + 1) * REC];
    __shared__ cdbl vacc[CLS >= 1 ? 3 * (D + 1) * 256 : 1];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    const int h = lane >> 5, j = lane & 31;
    const int nqb = (P.nk + 31) >> 5;        // blocks of 32 slots
    const int ngrp = (nqb + 3) >> 2;         // groups of 4 blocks = one workgroup
    // Workgroups are dealt to the eight XCDs in turn (id % 8); the grid is
    // renumbered so that each XCD takes a contiguous run of (pair, chunk, slot
    // group) indices: the four slot groups of a node chunk then share the
    // chunk's node table in that XCD's L2 instead of each fetching it from HBM.
    const long long total = (long long)P.nent * P.nchunk * ngrp;
    const long long per = (total + 7) >> 3;
    const long long li = (long long)(blockIdx.x & 7) * per + (blockIdx.x >> 3);
    if (li >= total) return;                 // whole workgroup, before any barrier
    const int grp = (int)(li % ngrp);
    const long long rest = li / ngrp;
    const int chunk = (int)(rest % P.nchunk);
    const int e = (int)(rest / P.nchunk);
    const int qb = grp * 4 + wv;
    int q = qb * 32 + j;
    const bool active = qb < nqb && q < P.nk;
    if (q >= P.nk) q = P.nk - 1;
    const int qi = (P.einst ? P.einst[e] * P.nk : 0) + q;   // the slot's entry in the kappa tables
    const double kap = P.kap[qi];
    const double lk = kTwoOverPi * log(0.5 * kap);
    const double kinv = 1.0 / kap;
    const double ksq = 1.0 / sqrt(kPi * kap);
    const unsigned f = P.flags[e];
    const bool swap = CLS == 1 ? P.ent[3 * e + 2] != 0 : false;
    const int nn = CLS == 2 ? P.ns * P.ns : (CLS == 1 ? 2 * P.nc * P.nc : P.nc * P.nc);
    // adjacent pairs: the chunks of each node half are numbered after each other
    // (a chunk never spans the two halves)
    int lo = chunk * P.chunk, hi = min(nn, lo + P.chunk), hn = 0;
    if (CLS == 1) {
        const int nh = P.nc * P.nc, cph = (nh + P.chunk - 1) / P.chunk;
        hn = chunk / cph;
        lo = hn * nh + (chunk - hn * cph) * P.chunk;
        hi = min(hn * nh + nh, lo + P.chunk);
    }
    const double* __restrict__ geo =
        P.einst ? P.igeo[P.einst[e]] + (size_t)(e - P.ibase[P.einst[e]]) * nn * kNodeStride : P.ngeo + (size_t)e * nn * kNodeStride;

    AccX<D> A;
#pragma unroll
    for (int t = 0; t <= D; ++t) {
        A.wa[t] = mk(0, 0);
#pragma unroll
        for (int m = 0; m < 2; ++m) { A.kd[t][m] = mk(0, 0); A.kt[t][m] = mk(0, 0); A.wb[t][m] = mk(0, 0); A.v[t][m] = mk(0, 0); }
    }

    if (CLS >= 1) {
#pragma unroll
        for (int t = 0; t < 3 * (D + 1); ++t) vacc[t * 256 + threadIdx.x] = mk(0, 0);
    }
    // Row-constant shape value b_0 of row a: general pairs s = x_a (their chunk
    // is the whole tensor rule, so the last row's fold is x_{ns-1} - 0, see the
    // node loop); adjacent pairs the test shape at t = xc_a in half 1, the trial
    // shape at t in half 0, folded after node nd when it ends a row or the chunk.
    const bool fx = CLS == 1 && (swap != (hn == 1));   // adjacent rows whose b_0 is 1 - t
    auto row_shape = [&](int a) {
        const double t = P.xc[a];
        return fx ? 1.0 - t : t;
    };
    auto fold_after = [&](int nd) {
        const int rl = P.nc;
        const int g = nd - hn * P.nc * P.nc, a = g / rl;
        if (nd == hi - 1) row_fold(A, row_shape(a), vacc);
        else if (g - a * rl == rl - 1) row_fold(A, row_shape(a) - row_shape(a + 1), vacc);
    };
    for (int base = lo; base < hi; base += STG) {
        const int cnt = min(STG, hi - base);
        __syncthreads();   // the previous round's records are no longer read
        for (int i = threadIdx.x; i < cnt; i += blockDim.x) stage_node<CLS>(P, geo, f, swap, base + i, lds + i * REC);
        if ((cnt & 1) && threadIdx.x == 0) stage_pad<CLS>(lds + cnt * REC);
        __syncthreads();
        if (qb < nqb) {
            for (int p = 0; p < cnt; p += 2) {
                const double* R0 = lds + p * REC;
                const double* Rown = R0 + h * REC;
                const double r = Rown[0];
                const double x = kap * r;
                double j0, y0, j1, y1;
                bessel_jy01_f<CLS == 2 ? 0 : 2>(x, x * x * (1.0 / 64.0), lk + Rown[3], kinv * Rown[2], ksq * Rown[4],
                                                j0, y0, j1, y1);
                double a0, a1, b0, b1, c0, c1, d0, d1;
                halves(j0, a0, a1);
                halves(y0, b0, b1);
                halves(j1, c0, c1);
                halves(y1, d0, d1);
                node_accum_x<CLS, D>(A, R0, h, a0, b0, c0, d0);
                if (CLS == 2) {
                    const int nd = base + p, a = nd / P.ns;
                    if (nd - a * P.ns == P.ns - 1) row_fold(A, P.xg[a] - (a + 1 < P.ns ? P.xg[a + 1] : 0.0), vacc);
                } else if (CLS == 1) {
                    fold_after(base + p);
                }
                node_accum_x<CLS, D>(A, R0 + REC, h, a1, b1, c1, d1);
                if (CLS == 2) {
                    const int nd = base + p + 1, a = nd / P.ns;
                    if (nd < hi && nd - a * P.ns == P.ns - 1)
                        row_fold(A, P.xg[a] - (a + 1 < P.ns ? P.xg[a + 1] : 0.0), vacc);
                } else if (CLS == 1 && base + p + 1 < hi) {
                    fold_after(base + p + 1);
                }
            }
        }
    }
    if (CLS >= 1) row_finish(A, vacc);
    if (CLS == 1 && hn == 0) {
        // adjacent half 0: lane half h holds P = entry (h, 0) and Q = entry (h, 1)
        // of K direct, W and V; lane half n takes entries (0, n), (1, n)
#pragma unroll
        for (int t = 0; t <= D; ++t) {
            cdbl* terms[3][2] = {{&A.kd[t][0], &A.kd[t][1]}, {&A.wb[t][0], &A.wb[t][1]}, {&A.v[t][0], &A.v[t][1]}};
#pragma unroll
            for (int u = 0; u < 3; ++u) {
                const cdbl Pv = *terms[u][0], Qv = *terms[u][1];
                double p0x, p1x, p0y, p1y, q0x, q1x, q0y, q1y;
                halves(Pv.x, p0x, p1x);
                halves(Pv.y, p0y, p1y);
                halves(Qv.x, q0x, q1x);
                halves(Qv.y, q0y, q1y);
                *terms[u][0] = h == 0 ? mk(p0x, p0y) : mk(q0x, q0y);
                *terms[u][1] = h == 0 ? mk(p1x, p1y) : mk(q1x, q1y);
            }
        }
    }

    // Sums -> this lane's entries q = m + 2h of the reference's local matrices
    // (same scalings as acc_store)
    const double kr = P.kr[qi], c = P.cv[qi];
    const double sqc = sqrt(c), k2h = 0.5 * kap * kap;
    cdbl Ko[D + 1][2], Kto[D + 1][2], Wo[D + 1][2], Vo[D + 1][2];
#pragma unroll
    for (int m = 0; m < 2; ++m) {
        const int qq = m + 2 * h;
        const double dd = (qq == 0 || qq == 3) ? 1.0 : -1.0;   // b'_m b'_n for P1 (+1, -1)
        Ko[0][m] = kap * A.kd[0][m]; Kto[0][m] = kap * A.kt[0][m];
        Wo[0][m] = dd * A.wa[0] - k2h * A.wb[0][m];
        Vo[0][m] = A.v[0][m];
        if (D >= 1) {
            Ko[1][m] = (kr * c) * A.kd[1][m]; Kto[1][m] = (kr * c) * A.kt[1][m];
            Wo[1][m] = (-sqc) * (dd * A.wa[1] - k2h * A.wb[1][m] + kap * A.wb[0][m]);
            Vo[1][m] = (-sqc) * A.v[1][m];
        }
        if (D >= 2) {
            Ko[2][m] = c * (A.kd[1][m] - kap * A.kd[2][m]); Kto[2][m] = c * (A.kt[1][m] - kap * A.kt[2][m]);
            Wo[2][m] = (-c) * (dd * (A.wa[2] - kinv * A.wa[1]) - k2h * (A.wb[2][m] - kinv * A.wb[1][m]) -
                              (2.0 * kap) * A.wb[1][m] + A.wb[0][m]);
            Vo[2][m] = c * (kinv * A.v[1][m] - A.v[2][m]);
        }
    }
    if (CLS == 0) {
        // entry (0,1) of W and V = entry (1,0) (galerkin_builder.cpp:571-573, 744-746):
        // q = 2 is (m = 0, h = 1), q = 1 is (m = 1, h = 0)
        cdbl W10[D + 1], V10[D + 1];
#pragma unroll
        for (int t = 0; t <= D; ++t) {
            double u0, u1;
            halves(Wo[t][1].x, u0, u1); W10[t].x = u0;
            halves(Wo[t][1].y, u0, u1); W10[t].y = u0;
            halves(Vo[t][1].x, u0, u1); V10[t].x = u0;
            halves(Vo[t][1].y, u0, u1); V10[t].y = u0;
        }
        if (h == 1) {
#pragma unroll
            for (int t = 0; t <= D; ++t) { Wo[t][0] = W10[t]; Vo[t][0] = V10[t]; }
            if (P.qflags & 1) {
                // reference quirk: the d = 0 sum of entry (1,0), scaled for order t
                if (D >= 1) Vo[1][0] = (-sqc) * V10[0];
                if (D >= 2) Vo[2][0] = c * V10[0];
            }
        }
    }
    const size_t rec = P.nchunk == 1 ? (size_t)P.pidx[e] : (size_t)e * P.nchunk + chunk;
    if (P.tdiff) {
        // Waves 2w (inner slot block) and 2w + 1 (outer) of the workgroup hold
        // the same 32 wavenumbers.  They form the differences T needs (K: inner
        // - outer, W and V: outer - inner): the inner wave entry m = 0 of its
        // lane half, the outer wave m = 1, each written once per wavenumber.
        // Both waves put their entries in the staging LDS (two layers per
        // round) and read their own and the partner's entry m back, so no
        // register array is indexed by the wave's parity (that would place
        // it in scratch memory).  Every thread reaches the barriers.
        static_assert((STG + 1) * REC >= 256 * 4 * 2, "exchange needs 16 KB of staging LDS");
        const int mine = wv & 1;
        const int kt = P.kt0 + ((qb >> 1) << 5) + j;
        const bool kact = qb < nqb && kt < P.nkt;
        const size_t nkt = (size_t)P.nkt;
        cdbl* out = P.out + rec * (16 * (D + 1)) * nkt + kt;
        cdbl* xch = reinterpret_cast<cdbl*>(lds);
        const int qq = mine + 2 * h;
#pragma unroll
        for (int t = 0; t <= D; ++t) {
#pragma unroll
            for (int half = 0; half < 2; ++half) {   // 0: K direct, K transposed; 1: W, V
                __syncthreads();   // the staging records (or the previous round) are no longer read
                cdbl* x = xch + threadIdx.x * 4;
                x[0] = half ? Wo[t][0] : Ko[t][0];
                x[1] = half ? Wo[t][1] : Ko[t][1];
                x[2] = half ? Vo[t][0] : Kto[t][0];
                x[3] = half ? Vo[t][1] : Kto[t][1];
                __syncthreads();
                const cdbl* own = xch + threadIdx.x * 4 + mine;
                const cdbl* par = xch + (threadIdx.x ^ 64) * 4 + mine;   // partner wave, same lane
                const cdbl in0 = mine ? par[0] : own[0], ou0 = mine ? own[0] : par[0];
                const cdbl in1 = mine ? par[2] : own[2], ou1 = mine ? own[2] : par[2];
                if (kact) {
                    if (half == 0) {
                        out[(16 * t + qq) * nkt] = in0 - ou0;
                        out[(16 * t + 4 + qq) * nkt] = CLS == 0 ? mk(0, 0) : in1 - ou1;
                    } else {
                        out[(16 * t + 8 + qq) * nkt] = ou0 - in0;
                        out[(16 * t + 12 + qq) * nkt] = ou1 - in1;
                    }
                }
            }
        }
    } else if (active) {
        const size_t nk = P.nk;
        cdbl* out = P.out + rec * (16 * (D + 1)) * nk + q;
#pragma unroll
        for (int t = 0; t <= D; ++t)
#pragma unroll
            for (int m = 0; m < 2; ++m) {
                const int qq = m + 2 * h;
                out[(16 * t + qq) * nk] = Ko[t][m];
                out[(16 * t + 4 + qq) * nk] = CLS == 0 ? mk(0, 0) : Kto[t][m];
                out[(16 * t + 8 + qq) * nk] = Wo[t][m];
                out[(16 * t + 12 + qq) * nk] = Vo[t][m];
            }
    }
}

// Sums the node-chunk partials of one class in chunk order into the pair buffer.
__global__ void k_chunk_sum(const cdbl* __restrict__ part, int nent, int nchunk, int nk, int nslot,
                            const long long* __restrict__ pidx, long long npairs, cdbl* __restrict__ L) {
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const long long total = (long long)nk * nent * nslot;
    if (id >= total) return;
    const int q = (int)(id % nk);
    const long long es = id / nk;
    const int s = (int)(es % nslot), e = (int)(es / nslot);
    const cdbl* p = part + ((size_t)e * nchunk * nslot + s) * nk + q;
    cdbl acc = mk(0, 0);
    for (int c = 0; c < nchunk; ++c) acc = acc + p[(size_t)c * nslot * nk];
    L[((size_t)pidx[e] * nslot + s) * nk + q] = acc;
}

// ---------------------------------------------------------------------------
// Gather: global entry (m, n) of an operator = sum of the local blocks of the
// panel pairs (P, Q), P in {m-1, m}, Q in {n-1, n} (P1 dofs: dof(0,P) = P+1,
// dof(1,P) = P; continuous_space.hpp:36-45), scatter rule of
// galerkin_builder.cpp:992-1009.
// ---------------------------------------------------------------------------
__device__ __forceinline__ long long tri(int i, int j) { return (long long)i * (i + 1) / 2 + j; }

// layer: 0 K, 1 W, 2 V
__device__ __forceinline__ cdbl op_entry(const cdbl* __restrict__ L, int nk, int nslot, int q, int layer, int t,
                                         int m, int n, int N) {
    cdbl e[2][2];
#pragma unroll
    for (int a = 0; a < 2; ++a) {
        const int P = a == 0 ? (m == 0 ? N - 1 : m - 1) : m;
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            const int Q = b == 0 ? (n == 0 ? N - 1 : n - 1) : n;
            long long p;
            int slot;
            if (P >= Q) {
                p = tri(P, Q);
                slot = (layer == 0 ? 0 : layer == 1 ? 8 : 12) + a + 2 * b;
            } else {
                p = tri(Q, P);
                slot = layer == 0 ? 4 + a + 2 * b : (layer == 1 ? 8 : 12) + b + 2 * a;
            }
            e[a][b] = L[((size_t)p * nslot + 16 * t + slot) * nk + q];
        }
    }
    // (e00 + e11) + (e01 + e10): under (m, n) -> (n, m) the symmetric layers
    // exchange e01 and e10 only, so V and W come out exactly symmetric
    const cdbl acc = (e[0][0] + e[1][1]) + (e[0][1] + e[1][0]);
    return mk(-0.25 * acc.y, 0.25 * acc.x);   // * i/4
}

// Builder matrices for wavenumber slot q: out[layer][t] (N x N column-major)
__global__ void k_gather_layers(const cdbl* __restrict__ L, int nk, int nslot, int q, int N, int D, int layers,
                                cdbl* __restrict__ Kout, cdbl* __restrict__ Wout, cdbl* __restrict__ Vout) {
    const long long NN = (long long)N * N;
    long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= NN * (D + 1)) return;
    const int t = (int)(id / NN);
    const long long e = id - t * NN;
    const int m = (int)(e % N), n = (int)(e / N);
    if ((layers & 1) && Kout) Kout[id] = op_entry(L, nk, nslot, q, 0, t, m, n, N);
    if ((layers & 4) && Wout) Wout[id] = op_entry(L, nk, nslot, q, 1, t, m, n, N);
    if ((layers & 2) && Vout) Vout[id] = op_entry(L, nk, nslot, q, 2, t, m, n, N);
}

// Transmission matrix (assemble_operator, gen_sol_op.cpp:59-74, unprojected):
// T = [[K_i - K_o + M, V_o - V_i], [W_o - W_i, K_o - K_i + M^T]], M only for t = 0.
// Output b uses wavenumber slots qi[b] (inner c_i), qo[b] (outer c_o); with qo
// null, L holds the differences per wavenumber (K_i - K_o, W_o - W_i, V_o - V_i
// at slot qi[b]; the paired-lane kernel's tdiff layout).  T_t of output b is
// written at Tt + b ld^2.
// Workgroup = one pair of 8 x 8 tiles (I, J), I >= J, of the N x N blocks for
// 32 outputs b; thread (i, jb) = (threadIdx & 7, threadIdx >> 3) forms the
// four block entries (m, n), (m + N, n + N), (m, n + N), (m + N, n) of row
// m = 8 I + i over the 8 columns of tile (I, J), then (for I > J) of row
// 8 J + i over the columns of tile (J, I).  op_entry sums the symmetric layers
// W and V bitwise symmetric, so tile (J, I) takes them from tile (I, J)
// through LDS: every local-matrix record is read once.  Reads run over 32
// wavenumbers per record slot, writes over 8 rows per column; the output
// chunks of one tile pair are neighbouring workgroups, so the chunks of a
// record slot are read at about the same time (long runs in each DRAM page).
__global__ __launch_bounds__(256) void k_form_T(const cdbl* __restrict__ L, int nk, int nslot,
                                                const int* __restrict__ qi, const int* __restrict__ qo, int nb, int N,
                                                int D, const double* __restrict__ M, cdbl* __restrict__ T0,
                                                cdbl* __restrict__ T1, cdbl* __restrict__ T2, int ld,
                                                cflt* __restrict__ T32) {
    __shared__ cdbl sW[8][8][32];   // [column in tile (I, J)][row in tile (I, J)][output]
    __shared__ cdbl sV[8][8][32];
    const int mt = (N + 7) >> 3, nbc = (nb + 31) >> 5;
    const int bc = blockIdx.x % nbc, tile = blockIdx.x / nbc;
    const int I = tile % mt, J = tile / mt;
    if (J > I) return;                       // whole workgroup: no barrier is skipped
    const int i = threadIdx.x & 7, jb = threadIdx.x >> 3;
    const int b = bc * 32 + jb;
    const bool act = b < nb;                 // idle threads still reach every barrier
    const int n2 = ld;                       // leading dimension (2N, or more for a zero-padded T)
    const size_t NN = (size_t)n2 * n2;
    const bool diff = qo == nullptr;
    const int a = act ? qi[b] : 0, o = act && !diff ? qo[b] : 0;
    const bool mirror = I != J;
    for (int t = 0; t <= D; ++t) {
        cdbl* Tt = (t == 0 ? T0 : (t == 1 ? T1 : T2)) + (size_t)(act ? b : 0) * NN;
        // T and, when asked, its FP32 copy for the mixed-precision LU (each
        // entry rounded to nearest) in the same pass
        cflt* Ts = t == 0 && T32 ? T32 + (size_t)(act ? b : 0) * NN : nullptr;
        auto put = [&](size_t idx, cdbl v) {
            Tt[idx] = v;
            if (Ts) Ts[idx] = cflt{(float)v.x, (float)v.y};
        };
        const int m = 8 * I + i;
        for (int c = 0; c < 8; ++c) {
            const int n = 8 * J + c;
            if (!act || m >= N || n >= N) continue;
            cdbl k00, k11, w, v;
            if (diff) {
                k00 = op_entry(L, nk, nslot, a, 0, t, m, n, N);
                k11 = mk(-k00.x, -k00.y);
                w = op_entry(L, nk, nslot, a, 1, t, m, n, N);
                v = op_entry(L, nk, nslot, a, 2, t, m, n, N);
            } else {
                const cdbl Ki = op_entry(L, nk, nslot, a, 0, t, m, n, N), Ko = op_entry(L, nk, nslot, o, 0, t, m, n, N);
                const cdbl Wi = op_entry(L, nk, nslot, a, 1, t, m, n, N), Wo = op_entry(L, nk, nslot, o, 1, t, m, n, N);
                const cdbl Vi = op_entry(L, nk, nslot, a, 2, t, m, n, N), Vo = op_entry(L, nk, nslot, o, 2, t, m, n, N);
                k00 = Ki - Ko; k11 = Ko - Ki;
                w = Wo - Wi; v = Vo - Vi;
            }
            if (t == 0) { k00.x += M[m + (size_t)n * N]; k11.x += M[n + (size_t)m * N]; }
            put(m + (size_t)n * n2, k00);
            put((m + N) + (size_t)(n + N) * n2, k11);
            put(m + (size_t)(n + N) * n2, v);
            put((m + N) + (size_t)n * n2, w);
            if (mirror) { sW[c][i][jb] = w; sV[c][i][jb] = v; }
        }
        if (!mirror) continue;
        __syncthreads();
        const int r = 8 * J + i;
        for (int c = 0; c < 8; ++c) {
            const int col = 8 * I + c;
            if (!act || r >= N || col >= N) continue;
            cdbl k00, k11;
            if (diff) {
                k00 = op_entry(L, nk, nslot, a, 0, t, r, col, N);
                k11 = mk(-k00.x, -k00.y);
            } else {
                const cdbl Ki = op_entry(L, nk, nslot, a, 0, t, r, col, N), Ko = op_entry(L, nk, nslot, o, 0, t, r, col, N);
                k00 = Ki - Ko; k11 = Ko - Ki;
            }
            if (t == 0) { k00.x += M[r + (size_t)col * N]; k11.x += M[col + (size_t)r * N]; }
            put(r + (size_t)col * n2, k00);
            put((r + N) + (size_t)(col + N) * n2, k11);
            put(r + (size_t)(col + N) * n2, sV[i][c][jb]);   // V(r, col) = V(col, r)
            put((r + N) + (size_t)col * n2, sW[i][c][jb]);
        }
        __syncthreads();                     // the next order overwrites sW / sV
    }
}

// Right-hand-side operator of the transmission solve (tp::direct_second_kind::
// A_rhs, solvers.cpp:116-122) from the outer operators of wavenumber slot o:
// A_o = [[-K_o + M/2, V_o], [W_o, K_o^T + M/2]] (2N x 2N, column-major).
// One thread per (m, n) writes the four block entries; K_o^T(m, n) = K_o(n, m)
// is gathered from the local matrices directly.
__global__ __launch_bounds__(256) void k_form_Ao(const cdbl* __restrict__ L, int nk, int nslot, int o, int N,
                                                 const double* __restrict__ M, cdbl* __restrict__ A) {
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (id >= (long long)N * N) return;
    const int m = (int)(id % N), n = (int)(id / N);
    const size_t n2 = 2 * (size_t)N;
    const double hm = 0.5 * M[m + (size_t)n * N];
    const cdbl Ko = op_entry(L, nk, nslot, o, 0, 0, m, n, N), KoT = op_entry(L, nk, nslot, o, 0, 0, n, m, N);
    A[m + (size_t)n * n2] = mk(-Ko.x + hm, -Ko.y);
    A[m + (size_t)(n + N) * n2] = op_entry(L, nk, nslot, o, 2, 0, m, n, N);
    A[(m + N) + (size_t)n * n2] = op_entry(L, nk, nslot, o, 1, 0, m, n, N);
    A[(m + N) + (size_t)(n + N) * n2] = mk(KoT.x + hm, KoT.y);
}

// ---------------------------------------------------------------------------
// Host launchers
// ---------------------------------------------------------------------------
void launch_setup_gauss_table(const double* panels, const double* coefs, int npan, const double* xg, int ns,
                              double* gtab, hipStream_t st) {
    int n = npan * ns;
    hipLaunchKernelGGL(k_setup_gauss_table, dim3((n + 255) / 256), dim3(256), 0, st, panels, coefs, npan, xg, ns, gtab);
}

void launch_setup_general(const int* ent, int nent, const double* gtab, int ns, double* ngeo, unsigned char* flags,
                          hipStream_t st) {
    if (nent <= 0) return;
    hipLaunchKernelGGL(k_setup_general, dim3(nent), dim3(256), 0, st, ent, nent, gtab, ns, ngeo, flags);
}

void launch_setup_singular(int cls, const int* ent, int nent, const double* panels, const double* coefs,
                           const double* xc, int nc, double* ngeo, unsigned char* flags, hipStream_t st) {
    if (nent <= 0) return;
    hipLaunchKernelGGL(k_setup_singular, dim3(nent), dim3(256), 0, st, cls, ent, nent, panels, coefs, xc, nc, ngeo, flags);
}

template <int CLS, int D>
static void launch_asm_g(const AsmParams& P, int G, hipStream_t st) {
    const long long nitems = (long long)P.nent * P.nchunk * P.nk;
    if (nitems <= 0) return;
    if (G == 0) {
        // paired-lane kernel: one workgroup per (pair, chunk, 4 blocks of 32 slots)
        const long long groups = ((P.nk + 31) / 32 + 3) / 4;
        const long long blocks = 8 * (((long long)P.nent * P.nchunk * groups + 7) / 8);   // a multiple of the XCD count
        hipLaunchKernelGGL((k_assemble_x2<CLS, D>), dim3((unsigned)blocks), dim3(256), 0, st, P);
        return;
    }
    // G = 1: one wave per (pair, chunk, block of 64 slots); else 64/G items per wave
    const long long waves = G == 1 ? (long long)P.nent * P.nchunk * ((P.nk + 63) / 64) : (nitems * G + 63) / 64;
    const int blocks = (int)((waves + 3) / 4);
    if (P.kapi) {
        switch (G) {
        case 4: hipLaunchKernelGGL((k_assemble<CLS, D, 4, cdbl>), dim3(blocks), dim3(256), 0, st, P); break;
        case 16: hipLaunchKernelGGL((k_assemble<CLS, D, 16, cdbl>), dim3(blocks), dim3(256), 0, st, P); break;
        default: hipLaunchKernelGGL((k_assemble<CLS, D, 64, cdbl>), dim3(blocks), dim3(256), 0, st, P); break;
        }
        return;
    }
    switch (G) {
    case 1: hipLaunchKernelGGL((k_assemble<CLS, D, 1, double>), dim3(blocks), dim3(256), 0, st, P); break;
    case 4: hipLaunchKernelGGL((k_assemble<CLS, D, 4, double>), dim3(blocks), dim3(256), 0, st, P); break;
    case 16: hipLaunchKernelGGL((k_assemble<CLS, D, 16, double>), dim3(blocks), dim3(256), 0, st, P); break;
    default: hipLaunchKernelGGL((k_assemble<CLS, D, 64, double>), dim3(blocks), dim3(256), 0, st, P); break;
    }
}

void launch_assemble(int cls, int D, const AsmParams& P, int G, hipStream_t st) {
    if (cls == 0) {
        if (D == 0) launch_asm_g<0, 0>(P, G, st); else if (D == 1) launch_asm_g<0, 1>(P, G, st); else launch_asm_g<0, 2>(P, G, st);
    } else if (cls == 1) {
        if (D == 0) launch_asm_g<1, 0>(P, G, st); else if (D == 1) launch_asm_g<1, 1>(P, G, st); else launch_asm_g<1, 2>(P, G, st);
    } else {
        if (D == 0) launch_asm_g<2, 0>(P, G, st); else if (D == 1) launch_asm_g<2, 1>(P, G, st); else launch_asm_g<2, 2>(P, G, st);
    }
}

void launch_chunk_sum(const cdbl* part, int nent, int nchunk, int nk, int nslot, const long long* pidx, long long npairs,
                      cdbl* L, hipStream_t st) {
    const long long total = (long long)nk * nent * nslot;
    if (total <= 0) return;
    hipLaunchKernelGGL(k_chunk_sum, dim3((unsigned)((total + 255) / 256)), dim3(256), 0, st, part, nent, nchunk, nk,
                       nslot, pidx, npairs, L);
}

void launch_gather_layers(const cdbl* L, int nk, int nslot, int q, int N, int D, int layers, cdbl* K, cdbl* W,
                          cdbl* V, hipStream_t st) {
    long long n = (long long)N * N * (D + 1);
    hipLaunchKernelGGL(k_gather_layers, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, L, nk, nslot, q, N, D,
                       layers, K, W, V);
}

void launch_form_T(const cdbl* L, int nk, int nslot, const int* qi, const int* qo, int nb, int N, int D,
                   const double* M, cdbl* T0, cdbl* T1, cdbl* T2, hipStream_t st, int ld, cflt* T32) {
    const unsigned mt = (unsigned)((N + 7) / 8);
    hipLaunchKernelGGL(k_form_T, dim3(mt * mt * (unsigned)((nb + 31) / 32)), dim3(256), 0, st, L, nk, nslot, qi, qo,
                       nb, N, D, M, T0, T1, T2, ld > 0 ? ld : 2 * N, T32);
}

// The rows and columns n..ld-1 of nb matrices of order ld (stride ld^2): the
// identity there, so that a matrix of order n padded to ld has the same LU
// factors on its leading n x n block and the identity below.
__global__ __launch_bounds__(256) void k_pad_identity(cdbl* __restrict__ T, int n, int ld) {
    const int b = blockIdx.y;
    const long long id = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const int pad = ld - n;
    // padded columns n..ld-1 (all rows), then the padded rows of columns 0..n-1
    const long long ncolpad = (long long)pad * ld, total = ncolpad + (long long)pad * n;
    if (id >= total) return;
    int i, j;
    if (id < ncolpad) { j = n + (int)(id / ld); i = (int)(id % ld); }
    else { const long long r = id - ncolpad; j = (int)(r / pad); i = n + (int)(r % pad); }
    T[(size_t)b * ld * ld + i + (size_t)j * ld] = mk(i == j ? 1.0 : 0.0, 0.0);
}

void launch_pad_identity(cdbl* T, int n, int ld, int nb, hipStream_t st) {
    const long long total = (long long)(ld - n) * ld + (long long)(ld - n) * n;
    if (total <= 0 || nb <= 0) return;
    hipLaunchKernelGGL(k_pad_identity, dim3((unsigned)((total + 255) / 256), nb), dim3(256), 0, st, T, n, ld);
}

void launch_form_Ao(const cdbl* L, int nk, int nslot, int o, int N, const double* M, cdbl* A, hipStream_t st) {
    const long long n = (long long)N * N;
    if (n <= 0) return;
    hipLaunchKernelGGL(k_form_Ao, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, L, nk, nslot, o, N, M, A);
}

}  // namespace hbem
