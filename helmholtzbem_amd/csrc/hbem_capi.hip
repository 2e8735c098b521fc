// hbem_capi.hip -- geometry contexts and the C-ABI of include/hbem.h.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/hbem.h"
#include "hbem_internal.hpp"

using hbem::cdbl;
using hbem::cflt;

namespace {

thread_local std::string g_err;
int g_device = 0;

struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void hip_check(hipError_t e, const char* what) {
    if (e != hipSuccess) throw Error(HBEM_EDEVICE, std::string(what) + ": " + hipGetErrorString(e));
}
#define HIPC(x) hip_check((x), #x)

// Growable device buffer.
struct DBuf {
    void* p = nullptr;
    size_t n = 0;
    void* get(size_t bytes) {
        if (bytes > n) {
            if (p) (void)hipFree(p);
            p = nullptr;
            n = 0;
            if (bytes > 0) {
                if (hipMalloc(&p, bytes) != hipSuccess) { p = nullptr; throw Error(HBEM_ENOMEM, "device allocation failed"); }
            }
            n = bytes;
        }
        return p;
    }
    template <class T> T* as(size_t count) { return (T*)get(count * sizeof(T)); }
    ~DBuf() { if (p) (void)hipFree(p); }
};

template <class T>
T* upload(DBuf& b, const std::vector<T>& v, hipStream_t st) {
    T* d = b.as<T>(std::max<size_t>(v.size(), 1));
    if (!v.empty()) HIPC(hipMemcpyAsync(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice, st));
    return d;
}

// Real sparse n x n matrix in compressed-row form on the device.
struct DCsr {
    DBuf rp, ci, v;
    // rows of op(A) for a dense column-major A (op = A or A^T); exact zeros are dropped
    void build(const std::vector<double>& A, int n, bool trans, hipStream_t st) {
        std::vector<int> r(n + 1, 0), c;
        std::vector<double> x;
        for (int i = 0; i < n; ++i) {
            for (int j = 0; j < n; ++j) {
                const double a = trans ? A[j + (size_t)i * n] : A[i + (size_t)j * n];
                if (a != 0.0) { c.push_back(j); x.push_back(a); }
            }
            r[i + 1] = (int)c.size();
        }
        upload(rp, r, st);
        upload(ci, c, st);
        upload(v, x, st);
    }
};

// Algorithmic FP64 flops per node (frozen model): Hankel node (13 geometry + 171
// Bessel) and accumulation for P1 x P1, three layers, derivative order d.
double flops_per_node(int d) { return 184.0 + (d == 0 ? 108.0 : d == 1 ? 192.0 : 314.0); }

// Timed sections on a stream.  begin/end record a pair of events around
// device work; nothing waits for them until resolve(), which the caller runs
// after the timed work (so timing does not add host synchronisations to it).
struct EventLog {
    static constexpr int kSlots = 8;
    struct Rec { int slot; hipEvent_t a, b; double flops; };
    std::vector<Rec> pending;
    std::vector<hipEvent_t> pool;
    double ms[kSlots] = {};
    long long n[kSlots] = {};
    double flops[kSlots] = {};
    hipEvent_t take() {
        if (!pool.empty()) { hipEvent_t e = pool.back(); pool.pop_back(); return e; }
        hipEvent_t e = nullptr;
        HIPC(hipEventCreate(&e));
        return e;
    }
    hipEvent_t begin(hipStream_t st) {
        hipEvent_t e = take();
        HIPC(hipEventRecord(e, st));
        return e;
    }
    void end(hipStream_t st, int slot, hipEvent_t a, double fl = 0.0) {
        hipEvent_t b = take();
        HIPC(hipEventRecord(b, st));
        pending.push_back({slot, a, b, fl});
    }
    void resolve() {
        for (auto& r : pending) {
            HIPC(hipEventSynchronize(r.b));
            float t = 0.f;
            HIPC(hipEventElapsedTime(&t, r.a, r.b));
            ms[r.slot] += t;
            n[r.slot] += 1;
            flops[r.slot] += r.flops;
            pool.push_back(r.a);
            pool.push_back(r.b);
        }
        pending.clear();
    }
    void reset() {
        resolve();
        for (int i = 0; i < kSlots; ++i) { ms[i] = 0; n[i] = 0; flops[i] = 0; }
    }
    ~EventLog() {
        for (auto& r : pending) { (void)hipEventDestroy(r.a); (void)hipEventDestroy(r.b); }
        for (auto e : pool) (void)hipEventDestroy(e);
    }
};

// The LU's matrix-core launches recorded into an event log slot
struct GemmTimer : hbem::LaunchTimer {
    EventLog* log;
    int slot;
    hipEvent_t e = nullptr;
    GemmTimer(EventLog* l, int s) : log(l), slot(s) {}
    void begin(hipStream_t st) override { e = log->begin(st); }
    void end(hipStream_t st, double flops) override { log->end(st, slot, e, flops); }
};

// Profile slots of a geometry context (hbem_profile_query)
enum { kProfGeneral = 0, kProfAdjacent = 1, kProfCoinciding = 2, kProfLU = 3, kProfFormT = 4, kProfRsvd = 5,
       kProfGemm = 6 };
// randomized_svd timer slots (randsvd.cpp:27-47): LU, first solve + QR, subspace iterations, last solve + SVD
enum { kTimLU = 0, kTimQR = 1, kTimIter = 2, kTimSVD = 3 };

// Number of node chunks of the paired-lane kernel per singular pair (chunks of
// 2048 nodes; adjacent pairs chunk each of their two node halves on its own).
constexpr int kChunkNodes = 2048;
long long adj_chunks(int nc) { return 2 * (((long long)nc * nc + kChunkNodes - 1) / kChunkNodes); }
long long coi_chunks(int nc) { return ((long long)nc * nc + kChunkNodes - 1) / kChunkNodes; }

// Device bytes a sigma sweep needs per wavenumber of a batch: local matrices
// and chunk partials (lcols columns per k: 1 in the T-difference layout, 2
// for the two refraction indices otherwise), T (factorised in place), T',
// T'', pivots and the thin blocks of the randomized SVD.
size_t sweep_bytes_per_k(long long npairs, long long nadj, long long ncoi, int nc_rule, int n, int D, int ncols,
                         int lcols) {
    const size_t nslot = 16 * (size_t)(D + 1);
    const size_t part = (size_t)std::max(nadj * adj_chunks(nc_rule), ncoi * coi_chunks(nc_rule));
    const size_t nn = (size_t)n * n;
    return (size_t)lcols * ((size_t)npairs + part) * nslot * sizeof(cdbl) + (size_t)(D + 1) * nn * sizeof(cdbl) +
           2 * (size_t)n * sizeof(int) + (4 * (size_t)ncols + 4) * n * sizeof(cdbl) + 64;
}

// Sweeps of >= 32 real positive wavenumbers run every batch on the paired-lane
// kernel in the T-difference layout (AsmParams::tdiff), so a k's bits do not
// depend on the batch size the free memory allows.
bool sweep_tdiff(const double* ks, int nk) {
    bool allpos = nk >= 32;
    for (int j = 0; j < nk && allpos; ++j) allpos = ks[j] > 0.0;
    return allpos;
}

// Extra device bytes per wavenumber of the sigma'' stage: T_p, the bordered
// 2n x 2n matrix (which is also the projection's scratch), its pivots and
// the 2n-vectors of the two solves.
size_t der2_bytes_per_k(int n) {
    const size_t nn = (size_t)n * n;
    return 5 * nn * sizeof(cdbl) + 4 * (size_t)n * sizeof(int) + 16 * (size_t)n * sizeof(cdbl) + 64;
}

// Extra device bytes per wavenumber of the mixed-precision dense stage: the
// FP32 copy of T (factorised in place) and the refinement's thin blocks.
size_t mixed_bytes_per_k(int n, int ncols) {
    return (size_t)n * n * sizeof(cflt) + (2 * (size_t)ncols * n + hbem::residual_part_count(n, ncols, 1)) * sizeof(cdbl) +
           64;
}

// Refinement tolerance of the mixed-precision sweep: a k is recomputed in FP64
// when the estimated error left after a refined solve (the last correction
// times the observed contraction rate) exceeds this fraction of the solution.
constexpr double kMixedTol = 1e-10;

// Wavenumbers per sweep batch: as many as fit in `avail` bytes, at most 256
// (the dense stage then has about one matrix per CU; more only adds latency),
// and balanced so the batches of nk wavenumbers have equal sizes.
int sweep_batch(size_t per_k, size_t avail, int nk) {
    long long B = (long long)(avail / std::max<size_t>(per_k, 1));
    if (B > 256) B = 256;
    if (B < 1) B = 1;
    if (nk > 0) {
        const long long nbatch = (nk + B - 1) / B;
        B = (nk + nbatch - 1) / nbatch;
    }
    return (int)B;
}

// Work space of hbem_rsvd_sv_noalloc: a fixed device array, one per GPU (the
// runtime instantiates file-scope device variables on every device the
// library runs on, each at its own address).  Layout for order n <= kNaMaxN
// and nc <= kNaMaxCols columns, each part 256-byte aligned: the LU copy of T
// (n^2 complex), pivots, permutation (n ints each), X, Y (n nc complex each),
// thin-solve work (2 n nc complex).
constexpr int kNaMaxN = 2048;
constexpr int kNaMaxCols = 8;
constexpr size_t al256(size_t b) { return (b + 255) & ~(size_t)255; }
constexpr size_t kNaBytes = al256(sizeof(cdbl) * (size_t)kNaMaxN * kNaMaxN) + 2 * al256(sizeof(int) * (size_t)kNaMaxN) +
                            4 * al256(sizeof(cdbl) * (size_t)kNaMaxN * kNaMaxCols);
static __device__ __align__(256) unsigned char scratch[kNaBytes];

// Device address of `scratch` on device `dev` (which must be current): the
// array's name is a host-side stand-in, so its address is looked up once per
// device and only that address is given to copies and kernels.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};
void* scratch_on(int dev) {
    if (dev < 0 || dev >= kMaxDevices) throw Error(HBEM_EDEVICE, "device index beyond the library's table");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        HIPC(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
        if (!p) throw Error(HBEM_EDEVICE, "no device address for the work array");
        g_scratch_ptr[dev] = p;
    }
    return g_scratch_ptr[dev];
}

// Free device memory usable by a call that may regrow buffers currently
// holding `held` bytes, less a reserve of 2 GiB or 3% of the device.
size_t usable_device_bytes(size_t held) {
    size_t fr = 0, tot = 0;
    HIPC(hipMemGetInfo(&fr, &tot));
    const size_t reserve = std::max<size_t>((size_t)2 << 30, tot / 32);
    const size_t have = fr + held;
    return have > reserve ? have - reserve : 0;
}

}  // namespace

struct hbem_geometry {
    int dev = 0;
    hipStream_t st = nullptr;
    int N = 0, order = 0, ns = 0, nc = 0, flags = 0;
    bool lines = true, proj = true;
    long long npairs = 0;
    std::vector<double> panels, coefs, xg, wg, xc, wc;
    std::vector<double> M, F, Finv;
    std::vector<int> tr;
    // device geometry
    DBuf d_panels, d_coefs, d_xg, d_wg, d_xc, d_wc, d_gtab;
    DBuf d_gent, d_gpidx, d_gflags, d_gngeo;
    DBuf d_aent, d_apidx, d_aflags, d_angeo;
    DBuf d_cent, d_cpidx, d_cflags, d_cngeo;
    int ngen = 0, nadj = 0, ncoi = 0;
    // General pairs in order of separation (the distance of the panels'
    // bounding discs in panel diameters): sep[0] pairs closer than one
    // diameter (or on a trigonometric or spline panel), sep[1] at one to two, sep[2] at two to
    // eight, sep[3] at eight to sixteen, sep[4] farther apart (see far_order);
    // dmax is the largest panel diameter.
    static constexpr int kSep = 5;
    int sep[kSep] = {0, 0, 0, 0, 0};
    double dmax = 0.0;
    // node tables of the separated classes at lower Gauss orders, built on
    // first use: key 64 cls + order
    struct FarSet {
        DBuf xg, wg, gtab, flags, ngeo;
    };
    std::map<int, std::unique_ptr<FarSet>> far;
    DBuf d_M, d_Finvc;
    DCsr s_F, s_FT, s_FinvT, s_Finv;   // F, F^T, Finv^T and Finv of the projection
    // work space
    DBuf w_Ao, w_U, w_rhs, w_flag;
    DBuf w_Tp, w_B, w_mv, w_bpiv, w_bperm, w_bvec, w_d2;   // bordered system of sv_der2 (sweeps)
    DBuf w_T32, w_rb, w_rr, w_rpart, w_flag2, w_rprev;       // mixed-precision dense stage
    DBuf w_kapi, w_kri, w_L, w_T0, w_T1, w_T2, w_piv, w_perm, w_X, w_Y, w_work, w_R, w_u, w_v, w_a, w_b, w_sig, w_dsig, w_kap,
        w_kr, w_cv, w_qi, w_qo, w_tmp, w_out, w_part;
    // profiling (hbem_profile_enable / query)
    bool prof = false;
    EventLog plog;
    std::mutex mu;
    // bytes held by the per-batch work buffers of the sweep
    size_t work_held() const {
        const DBuf* bs[] = {&w_L, &w_part, &w_T0, &w_T1, &w_T2, &w_piv, &w_perm, &w_X, &w_Y, &w_work, &w_u, &w_v, &w_a, &w_b,
                            &w_Tp, &w_B, &w_bpiv, &w_bperm, &w_bvec, &w_T32, &w_rb, &w_rr, &w_rpart, &w_rprev};
        size_t s = 0;
        for (auto* b : bs) s += b->n;
        return s;
    }
};

// drops the multi-instance graphs that name g (defined with the multi-instance sweep)
void multi_forget(hbem_geometry* g);

namespace {

int fail(const std::exception& e) {
    g_err = e.what();
    if (auto* he = dynamic_cast<const Error*>(&e)) return he->code;
    return HBEM_EINVAL;
}

#define API_BEGIN try {
#define API_END \
    }                                              \
    catch (const std::exception& e) { return fail(e); } \
    return HBEM_OK;

// GalerkinBuilder::initialize_constants' validation (galerkin_builder.cpp:899-916)
void check_k(double k_re, double k_im, double c) {
    if (c < 1.0) throw Error(HBEM_EINVAL, "Refraction index must not be smaller than 1");
    if (k_im < 0.0)
        throw Error(HBEM_EINVAL, "Wavenumber does not meet conditions for uniqueness of the solution (Im(k) >= 0)");
    if (!std::isfinite(k_re) || !std::isfinite(k_im)) throw Error(HBEM_EINVAL, "wavenumber must be finite");
}

// k_rp of initialize_constants: the real Hankel path
bool real_positive(double k_re, double k_im) { return k_im == 0.0 && k_re > 0.0; }

// Slot tables of the tdiff layout (AsmParams::tdiff) for nb wavenumbers ks
// with refraction indices c_i (inner) and c_o: blocks of 32 slots, block 2w
// inner and block 2w + 1 outer for wavenumbers 32w .. 32w + 31; slots past
// the last wavenumber repeat it (computed, never written).
void tdiff_slots(const double* ks, int nb, double c_i, double c_o, std::vector<double>& kap, std::vector<double>& kr,
                 std::vector<double>& cv) {
    const int nblk = (nb + 31) / 32;
    kap.resize(64 * (size_t)nblk);
    kr.resize(kap.size());
    cv.resize(kap.size());
    for (size_t q = 0; q < kap.size(); ++q) {
        const int blk = (int)(q >> 5), kt = std::min(nb - 1, (blk >> 1) * 32 + (int)(q & 31));
        const double c = (blk & 1) ? c_o : c_i;
        kap[q] = ks[kt] * std::sqrt(c);
        kr[q] = ks[kt];
        cv[q] = c;
    }
}

// Gauss order of the tensor rule for the separated general pairs of class cls
// (1: one to two panel diameters apart, 2: two to eight, 3: eight to sixteen,
// 4: farther) when kappa times the largest panel diameter is at most kd: the
// lowest order whose single- and double-layer local matrices stay within 1e-13
// of their largest entry of the 11-point rule's -- with a margin of two, 5e-14
// -- for straight panels at the class's least separation, end-on, side by side
// and at 42 random angles, at both ends of the kd interval (the rule's error
// falls as (separation)^(-2m) and with the oscillation kappa d as
// (kappa d)^(2m) / (2m)!); ns when no lower order qualifies.
int far_order(double kd, int cls, int ns) {
    static const double edge[5] = {0.4, 0.65, 1.25, 2.5, 5.0};
    static const int order[5][4] = {{9, 8, 6, 5}, {9, 8, 6, 6}, {9, 8, 7, 7}, {9, 8, 8, 8}, {10, 10, 10, 10}};
    for (int t = 0; t < 5; ++t)
        if (kd <= edge[t]) return std::min(order[t][cls - 1], ns);
    return ns;
}

// The node tables of separation class cls (1 to 4) at Gauss order m < ns,
// built on first use (one setup launch per table).
hbem_geometry::FarSet& far_set(hbem_geometry* g, int cls, int m) {
    auto& slot = g->far[64 * cls + m];
    if (slot) return *slot;
    auto f = std::make_unique<hbem_geometry::FarSet>();
    std::vector<double> xg(m), wg(m);
    hbem::gauss_rule(0.0, 1.0, m, xg.data(), wg.data());
    const double* dxg = upload(f->xg, xg, g->st);
    upload(f->wg, wg, g->st);
    double* gtab = f->gtab.as<double>((size_t)g->N * m * 5);
    hbem::launch_setup_gauss_table((const double*)g->d_panels.p, (const double*)g->d_coefs.p, g->N, dxg, m, gtab, g->st);
    int off = 0;
    for (int s = 0; s < cls; ++s) off += g->sep[s];
    const int cnt = g->sep[cls];
    f->flags.get(std::max(1, cnt));
    double* ngeo = f->ngeo.as<double>(std::max<size_t>(1, (size_t)cnt * m * m * hbem::kNodeStride));
    if (cnt > 0)
        hbem::launch_setup_general((const int*)g->d_gent.p + 3 * (size_t)off, cnt, gtab, m, ngeo,
                                   (unsigned char*)f->flags.p, g->st);
    HIPC(hipGetLastError());
    slot = std::move(f);
    return *slot;
}

// Contexts with HBEM_FLAG_FAR_FIELD: the node tables every order far_order
// gives for these kappa builds now (a batch's order is the order of one of its
// kappa), so that batch sizing sees the memory they take.
void prepare_far(hbem_geometry* g, const std::vector<double>& kap) {
    if (!(g->flags & HBEM_FLAG_FAR_FIELD)) return;
    for (int s = 1; s < hbem_geometry::kSep; ++s) {
        if (g->sep[s] == 0) continue;
        for (double v : kap) {
            const int m = far_order(std::fabs(v) * g->dmax, s, g->ns);
            if (m < g->ns) far_set(g, s, m);
        }
    }
}

// Assemble the local matrices of all pairs for wavenumber slots (kap, kr, cv)
// with imaginary parts (kapi, kri) when any slot is not real positive; returns
// the local buffer.  Wide real batches run on the paired-lane kernel, the rest
// on node-parallel groups of lanes per item.  nkt > 0: the slots are in the
// tdiff layout (tdiff_slots) for nkt wavenumbers, the paired-lane kernel runs
// whatever the width, and the buffer holds the differences T is formed from,
// nkt columns wide.  Contexts with HBEM_FLAG_FAR_FIELD and real slots: the
// separated general pairs take the lower Gauss orders of far_order, chosen for
// the largest kappa of each block of 32 wavenumbers in the tdiff layout (of
// the batch otherwise); a launch per separation class and run of consecutive
// blocks with the same order.
cdbl* assemble_locals(hbem_geometry* g, const std::vector<double>& kap, const std::vector<double>& kr,
                      const std::vector<double>& cv, int D, const std::vector<double>& kapi = {},
                      const std::vector<double>& kri = {}, int nkt = 0) {
    const int nk = (int)kap.size();
    const int nslot = 16 * (D + 1);
    const bool tdiff = nkt > 0;
    if (tdiff && (!kapi.empty() || nk % 64 != 0 || 2 * nkt > nk || 2 * nkt <= nk - 64))
        throw Error(HBEM_EINVAL, "tdiff layout needs real slots in blocks of 64 for its wavenumbers");
    const int nko = tdiff ? nkt : nk;   // columns of the output layouts
    cdbl* L = g->w_L.as<cdbl>((size_t)nko * g->npairs * nslot);
    hbem::AsmParams P;
    std::memset(&P, 0, sizeof(P));
    P.xg = (const double*)g->d_xg.p;
    P.wg = (const double*)g->d_wg.p;
    P.xc = (const double*)g->d_xc.p;
    P.wc = (const double*)g->d_wc.p;
    const double* dkap = upload(g->w_kap, kap, g->st);
    const double* dkr = upload(g->w_kr, kr, g->st);
    const double* dcv = upload(g->w_cv, cv, g->st);
    const bool cplx = !kapi.empty();
    const double* dkapi = cplx ? upload(g->w_kapi, kapi, g->st) : nullptr;
    const double* dkri = cplx ? upload(g->w_kri, kri, g->st) : nullptr;
    P.npairs_total = g->npairs;
    P.ns = g->ns;
    P.nc = g->nc;
    P.qflags = g->flags;
    P.tdiff = tdiff ? 1 : 0;
    P.nkt = nkt;
    // lanes per item: G = 0 selects the paired-lane kernel for wide batches
    // (32 wavenumbers x 2 nodes per wavefront); node groups of G lanes otherwise
    const int Gg = cplx ? 16 : (nk >= 64 || tdiff ? 0 : (nk >= 16 ? 4 : 16));
    const int Gs = ((nk >= 64 || tdiff) && !cplx) ? 0 : 64;
    struct Launch {
        int cls, which, G, nent, nodes;
        const int* ent;
        const double* ngeo;
        const unsigned char* flags;
        const long long* pidx;
        int ns;                       // general pairs: order of the tensor rule
        const double *xg, *wg;
        int q0, nq, kt0;              // slots [q0, q0 + nq) of the batch; first wavenumber column (tdiff)
    };
    const double* dxg = (const double*)g->d_xg.p;
    const double* dwg = (const double*)g->d_wg.p;
    std::vector<Launch> ls;
    const int* gent = (const int*)g->d_gent.p;
    const long long* gpidx = (const long long*)g->d_gpidx.p;
    const unsigned char* gflags = (const unsigned char*)g->d_gflags.p;
    const double* gngeo = (const double*)g->d_gngeo.p;
    if ((g->flags & HBEM_FLAG_FAR_FIELD) && !cplx) {
        // units of slots that share an order: blocks of 32 wavenumbers (64
        // slots) in the tdiff layout, the whole batch otherwise
        const int uslots = tdiff ? 64 : nk, nunit = nk / uslots;
        std::vector<double> ukd(nunit, 0.0);
        for (int u = 0; u < nunit; ++u)
            for (int q = u * uslots; q < (u + 1) * uslots; ++q) ukd[u] = std::max(ukd[u], std::fabs(kap[q]) * g->dmax);
        int off = 0;
        for (int s = 0; s < hbem_geometry::kSep; ++s) {
            const int cnt = g->sep[s];
            auto order_of = [&](int u) { return s == 0 ? g->ns : far_order(ukd[u], s, g->ns); };
            for (int u = 0; u < nunit && cnt > 0;) {
                const int m = order_of(u);
                int v = u + 1;
                while (v < nunit && order_of(v) == m) ++v;
                Launch l{2, 0, Gg, cnt, g->ns * g->ns, gent + 3 * (size_t)off,
                         gngeo + (size_t)off * g->ns * g->ns * hbem::kNodeStride, gflags + off, gpidx + off, g->ns,
                         dxg, dwg, u * uslots, (v - u) * uslots, tdiff ? 32 * u : 0};
                if (m < g->ns) {
                    auto& f = far_set(g, s, m);
                    l.nodes = m * m;
                    l.ngeo = (const double*)f.ngeo.p;
                    l.flags = (const unsigned char*)f.flags.p;
                    l.ns = m;
                    l.xg = (const double*)f.xg.p;
                    l.wg = (const double*)f.wg.p;
                }
                ls.push_back(l);
                u = v;
            }
            off += cnt;
        }
    } else {
        ls.push_back({2, 0, Gg, g->ngen, g->ns * g->ns, gent, gngeo, gflags, gpidx, g->ns, dxg, dwg, 0, nk, 0});
    }
    ls.push_back({1, 1, Gs, g->nadj, 2 * g->nc * g->nc, (const int*)g->d_aent.p, (const double*)g->d_angeo.p,
                  (const unsigned char*)g->d_aflags.p, (const long long*)g->d_apidx.p, g->ns, dxg, dwg, 0, nk, 0});
    ls.push_back({0, 2, Gs, g->ncoi, g->nc * g->nc, (const int*)g->d_cent.p, (const double*)g->d_cngeo.p,
                  (const unsigned char*)g->d_cflags.p, (const long long*)g->d_cpidx.p, g->ns, dxg, dwg, 0, nk, 0});
    for (auto& l : ls) {
        if (l.nent == 0) continue;
        P.nent = l.nent;
        P.ent = l.ent;
        P.ngeo = l.ngeo;
        P.flags = l.flags;
        P.pidx = l.pidx;
        P.ns = l.ns;
        P.xg = l.xg;
        P.wg = l.wg;
        P.nk = l.nq;
        P.kap = dkap + l.q0;
        P.kr = dkr + l.q0;
        P.cv = dcv + l.q0;
        if (cplx) {
            P.kapi = dkapi + l.q0;
            P.kri = dkri + l.q0;
        }
        P.kt0 = l.kt0;
        // general pairs: one chunk; singular pairs: chunks of 2048 nodes (paired-lane
        // kernel, 16 staging rounds; for adjacent pairs each of the two node halves
        // is chunked on its own) or 64 x 48 nodes (one wavefront per item)
        const int chunk_nodes = l.cls == 2 ? l.nodes : (l.G == 0 ? kChunkNodes : (l.G == 1 ? 512 : 64 * 48));
        P.chunk = chunk_nodes;
        P.nchunk = (l.nodes + chunk_nodes - 1) / chunk_nodes;
        if (l.cls == 1 && l.G == 0) P.nchunk = 2 * ((l.nodes / 2 + chunk_nodes - 1) / chunk_nodes);
        cdbl* part = nullptr;
        if (P.nchunk > 1) part = g->w_part.as<cdbl>((size_t)nko * l.nent * P.nchunk * nslot);
        P.out = P.nchunk > 1 ? part : L;
        // the class's kernel alone is timed: the chunk sums are a separate launch
        hipEvent_t e0 = g->prof ? g->plog.begin(g->st) : nullptr;
        hbem::launch_assemble(l.cls, D, P, l.G, g->st);
        HIPC(hipGetLastError());
        if (g->prof) g->plog.end(g->st, l.which, e0, (double)l.nent * l.nq * l.nodes * flops_per_node(D));
        if (P.nchunk > 1) {
            hbem::launch_chunk_sum(part, l.nent, P.nchunk, nko, nslot, l.pidx, g->npairs, L, g->st);
            HIPC(hipGetLastError());
        }
    }
    return L;
}

// Dense stage shared by sv, sv_der and the sweep: given LU factors of the
// unprojected T (batch matrices), the projection factor (sparse F and F^T, or
// none) and R, returns sigma and the singular pair (u, v) of
// T_p = F^{-1} T F^{-T} in d_u / d_v.
struct DenseCtx {
    hipStream_t st;
    int n, batch, nc, q;
    const cdbl* LU;
    long long ms;        // matrix stride of LU
    const int* perm;
    const DCsr* F;       // projection factor, null = no projection
    const DCsr* FT;      // its transpose
    // or, for a batch of zero-padded matrices of several scatterers, the
    // factor of each matrix's scatterer (device arrays of batch entries)
    const hbem::CsrDesc* Fm = nullptr;
    const hbem::CsrDesc* FTm = nullptr;
    cdbl* X;             // n x nc x batch
    cdbl* Y;
    cdbl* work;          // 2 n nc batch: thin-solve scratch, then the projected right-hand sides
    EventLog* tim = nullptr;   // randomized_svd timers (kTim*), or null
    // mixed precision (LU32 set): the factors are FP32 (LU32, stride ms) and
    // every solve with T is refined `refine` times against T64 (the FP64 T,
    // stride ms): residual in FP64, correction from the FP32 factors
    const cflt* LU32 = nullptr;
    const cdbl* T64 = nullptr;
    int refine = 0;         // steps of the final solve of the randomized SVD
    int refine_inner = 0;   // steps of the solves before it
    bool last_solve = false;
    double* rprev = nullptr;   // per matrix: the previous correction's size
    double rtol = 0.0;   // a matrix whose last correction exceeds rtol of its solution is flagged
    int* flag = nullptr; // per matrix
    cdbl* rb = nullptr;  // n nc batch: the right-hand side
    cdbl* rr = nullptr;  // n nc batch: residual, then correction
    cdbl* rpart = nullptr;
};

// Z <- op(T)^{-1} Z with the context's factors (op = T or T^H): FP64 factors,
// or FP32 factors plus `refine` steps of FP64 iterative refinement.
void t_solve(DenseCtx& c, cdbl* Z, bool adj) {
    if (!c.LU32) {
        hbem::lu_solve_thin_batched(c.LU, c.ms, c.perm, c.n, c.batch, Z, c.nc, adj, c.work, c.st);
        return;
    }
    const size_t cnt = (size_t)c.n * c.nc;
    const int steps = c.last_solve ? c.refine : c.refine_inner;
    if (steps > 0) HIPC(hipMemcpyAsync(c.rb, Z, sizeof(cdbl) * cnt * c.batch, hipMemcpyDeviceToDevice, c.st));
    hbem::lu_solve_thin_batched_f32(c.LU32, c.ms, c.perm, c.n, c.batch, Z, c.nc, adj, c.work, c.st);
    // the final solve, whose vectors give sigma, must reach the tolerance; the
    // solves before it only move the subspace and must contract
    const int check_last = c.last_solve ? 2 : 1;
    for (int it = 0; it < steps; ++it) {
        hbem::residual_thin_batched(c.T64, c.ms, Z, c.rb, c.rr, c.n, c.nc, c.batch, adj, c.rpart, c.st);
        hbem::lu_solve_thin_batched_f32(c.LU32, c.ms, c.perm, c.n, c.batch, c.rr, c.nc, adj, c.work, c.st);
        hbem::refine_update_batched(Z, c.rr, (int)cnt, c.batch, c.rtol, it == steps - 1 ? check_last : 0, it == 0,
                                    c.rprev, c.flag, c.st);
    }
}

// Y <- T_p^{-1} X  or  T_p^{-H} X  (T_p^{-1} = F^T T^{-1} F, T_p^{-H} = F^T T^{-H} F); X is kept.
void tp_solve(DenseCtx& c, const cdbl* X, cdbl* Y, bool adj) {
    const long long cols = (long long)c.nc * c.batch;
    if (c.Fm) {
        cdbl* Z = c.work + (size_t)c.n * cols;
        hbem::csr_apply_multi(c.Fm, c.n, X, Z, c.nc, c.batch, c.st);
        t_solve(c, Z, adj);
        hbem::csr_apply_multi(c.FTm, c.n, Z, Y, c.nc, c.batch, c.st);
    } else if (c.F) {
        cdbl* Z = c.work + (size_t)c.n * cols;
        hbem::csr_apply_batched((const int*)c.F->rp.p, (const int*)c.F->ci.p, (const double*)c.F->v.p, c.n, X, Z, cols,
                                c.st);
        t_solve(c, Z, adj);
        hbem::csr_apply_batched((const int*)c.FT->rp.p, (const int*)c.FT->ci.p, (const double*)c.FT->v.p, c.n, Z, Y,
                                cols, c.st);
    } else {
        HIPC(hipMemcpyAsync(Y, X, sizeof(cdbl) * (size_t)c.n * cols, hipMemcpyDeviceToDevice, c.st));
        t_solve(c, Y, adj);
    }
}

// randomized_svd::sv (randsvd.cpp:99-146) on the batch; X holds R replicated
// per matrix.  The power iteration alternates between X and Y.
void rsvd_core(DenseCtx& c, double* d_sigma, cdbl* d_u, cdbl* d_v) {
    hipEvent_t e = c.tim ? c.tim->begin(c.st) : nullptr;
    tp_solve(c, c.X, c.Y, false);                              // Y = T_p^{-1} R
    hbem::orthonormalize_thin_batched(c.Y, c.n, c.batch, c.nc, c.st);   // Q
    if (c.tim) { c.tim->end(c.st, kTimQR, e); e = c.tim->begin(c.st); }
    for (int it = 0; it < c.q; ++it) {
        tp_solve(c, c.Y, c.X, true);
        hbem::orthonormalize_thin_batched(c.X, c.n, c.batch, c.nc, c.st);
        tp_solve(c, c.X, c.Y, false);
        hbem::orthonormalize_thin_batched(c.Y, c.n, c.batch, c.nc, c.st);
    }
    if (c.tim) { c.tim->end(c.st, kTimIter, e); e = c.tim->begin(c.st); }
    // Q = Y, X = T_p^{-H} Q
    c.last_solve = true;
    tp_solve(c, c.Y, c.X, true);
    c.last_solve = false;
    hbem::sigma_from_thin_batched(c.X, c.Y, c.n, c.batch, c.nc, d_sigma, d_u, d_v, c.st);
    if (c.tim) c.tim->end(c.st, kTimSVD, e);
}

void copy_out_square(const cdbl* d, void* h, size_t count, hipStream_t st) {
    if (h) HIPC(hipMemcpyAsync(h, d, count * sizeof(cdbl), hipMemcpyDeviceToHost, st));
}

// T_p = Finv T Finv^T for `count` matrices (in place through tmp: count n^2
// entries, the context's buffer when null)
void project_inplace(hbem_geometry* g, cdbl* T, int count, cdbl* tmp = nullptr) {
    const int n = 2 * g->N;
    const long long s = (long long)n * n;
    if (!tmp) tmp = g->w_tmp.as<cdbl>((size_t)s * count);
    const cdbl* Fi = (const cdbl*)g->d_Finvc.p;
    hbem::zgemm_batched(n, n, n, hbem::mk(1, 0), Fi, n, 0, T, n, s, false, hbem::mk(0, 0), tmp, n, s, count, g->st);
    hbem::zgemm_batched(n, n, n, hbem::mk(1, 0), tmp, n, s, Fi, n, 0, true, hbem::mk(0, 0), T, n, s, count, g->st);
}

// The linear system of tp::direct_second_kind::A_rhs (solvers.cpp:106-131) on
// the device: both wavenumber slots (c_i, c_o) are assembled in one pass, T is
// gen_sol_op's unprojected operator, and rhs = A_o U for the nrhs host columns
// U (2N x nrhs) with A_o = [[-K_o + M/2, V_o], [W_o, K_o^T + M/2]].
void tp_system(hbem_geometry* g, double k_re, double k_im, double c_o, double c_i, const void* U, int nrhs,
               cdbl*& T, cdbl*& rhs) {
    const double si = std::sqrt(c_i), so = std::sqrt(c_o);
    cdbl* L = real_positive(k_re, k_im)
                  ? assemble_locals(g, {k_re * si, k_re * so}, {k_re, k_re}, {c_i, c_o}, 0)
                  : assemble_locals(g, {k_re * si, k_re * so}, {k_re, k_re}, {c_i, c_o}, 0, {k_im * si, k_im * so},
                                    {k_im, k_im});
    const int N = g->N, n = 2 * N;
    const size_t nn = (size_t)n * n;
    const double* dM = (const double*)g->d_M.p;
    T = g->w_T0.as<cdbl>(nn);
    std::vector<int> qi = {0}, qo = {1};
    const int* dqi = upload(g->w_qi, qi, g->st);
    const int* dqo = upload(g->w_qo, qo, g->st);
    hbem::launch_form_T(L, 2, 16, dqi, dqo, 1, N, 0, dM, T, nullptr, nullptr, g->st);
    HIPC(hipGetLastError());
    cdbl* Ao = g->w_Ao.as<cdbl>(nn);
    hbem::launch_form_Ao(L, 2, 16, 1, N, dM, Ao, g->st);
    HIPC(hipGetLastError());
    cdbl* dU = g->w_U.as<cdbl>((size_t)n * nrhs);
    HIPC(hipMemcpyAsync(dU, U, sizeof(cdbl) * (size_t)n * nrhs, hipMemcpyHostToDevice, g->st));
    rhs = g->w_rhs.as<cdbl>((size_t)n * nrhs);
    hbem::zgemm_batched(n, nrhs, n, hbem::mk(1, 0), Ao, n, 0, dU, n, 0, false, hbem::mk(0, 0), rhs, n, 0, 1, g->st);
    HIPC(hipGetLastError());
}

void check_tp_args(hbem_geometry* g, const void* U, int nrhs) {
    if (!g || !U) throw Error(HBEM_EINVAL, "null argument");
    if (nrhs < 1) throw Error(HBEM_EINVAL, "number of right-hand sides must be positive");
}

}  // namespace

extern "C" {

int hbem_last_error(char* buf, size_t len) {
    if (buf && len > 0) {
        std::strncpy(buf, g_err.c_str(), len - 1);
        buf[len - 1] = 0;
    }
    return (int)g_err.size();
}

int hbem_init(int device) {
    API_BEGIN
    int n = 0;
    HIPC(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) throw Error(HBEM_EINVAL, "invalid device index");
    g_device = device;
    HIPC(hipSetDevice(device));
    scratch_on(device);          // resolve hbem_rsvd_sv_noalloc's array on this device now
    API_END
}

int hbem_geometry_create(const double* panels, int npanels, const double* coefs, int ncoefs, int order, int proj,
                         int flags, hbem_geometry_t* out) {
    hbem_geometry* g = nullptr;
    API_BEGIN
    if (!out) throw Error(HBEM_EINVAL, "null output handle");
    if (npanels < 3) throw Error(HBEM_EINVAL, "a closed P1 mesh needs at least 3 panels");
    if (order < 1 || order > 40) throw Error(HBEM_EINVAL, "quadrature order out of range");
    if (flags & ~(HBEM_FLAG_QUIRK_SLP_COINC_DER | HBEM_FLAG_FAR_FIELD)) throw Error(HBEM_EINVAL, "unknown geometry flags");
    for (int i = 0; i < npanels; ++i) {
        int t = (int)panels[8 * i];
        if (t < 0 || t > 3 || panels[8 * i] != (double)t) throw Error(HBEM_EINVAL, "invalid panel type");
        if (t == 2) {
            int c0 = (int)panels[8 * i + 3], nc = (int)panels[8 * i + 4];
            if (c0 < 0 || nc < 1 || c0 + nc > ncoefs) throw Error(HBEM_EINVAL, "trig-curve panel coefficients out of range");
        }
        if (t == 3) {
            const long long c0 = (long long)panels[8 * i + 3], np = (long long)panels[8 * i + 4];
            if (c0 < 0 || np < 1 || c0 + 3 * np + 1 > ncoefs)
                throw Error(HBEM_EINVAL, "spline panel coefficients out of range");
            if (!(panels[8 * i + 1] < panels[8 * i + 2])) throw Error(HBEM_EINVAL, "spline panel needs tmin < tmax");
        }
    }
    g = new hbem_geometry;
    g->dev = g_device;
    HIPC(hipSetDevice(g->dev));
    HIPC(hipStreamCreateWithFlags(&g->st, hipStreamNonBlocking));
    g->N = npanels;
    g->order = order;
    g->ns = order;
    g->nc = order * (order + 1) / 2 - 1;
    g->flags = flags;
    g->proj = proj != 0;
    g->panels.assign(panels, panels + 8 * (size_t)npanels);
    if (ncoefs > 0) g->coefs.assign(coefs, coefs + 3 * (size_t)ncoefs);
    g->lines = true;
    for (int i = 0; i < npanels; ++i) if ((int)panels[8 * i] != 0) g->lines = false;
    g->xg.resize(g->ns); g->wg.resize(g->ns); g->xc.resize(g->nc); g->wc.resize(g->nc);
    hbem::gauss_rule(0.0, 1.0, g->ns, g->xg.data(), g->wg.data());
    hbem::composite_rule(order, g->xc.data(), g->wc.data());
    const int N = npanels;
    g->npairs = (long long)N * (N + 1) / 2;
    // classify pairs (galerkin_builder.cpp:872-897; adjacency test 189-201)
    std::vector<long long> gpi, api, cpi;
    std::vector<int> gent, aent, cent;
    gent.reserve(3 * (size_t)g->npairs);
    gpi.reserve((size_t)g->npairs);
    const double* cptr = g->coefs.empty() ? nullptr : g->coefs.data();
    std::vector<double> ends(4 * (size_t)N), disc(3 * (size_t)N);
    hbem::panel_endpoints(g->panels.data(), cptr, N, ends.data());
    hbem::panel_bounds(g->panels.data(), cptr, N, disc.data());
    for (int i = 0; i < N; ++i) g->dmax = std::max(g->dmax, 2.0 * disc[3 * (size_t)i + 2]);
    // general pairs by separation eta = (distance of the discs) / (larger
    // diameter), a lower bound of the panels' distance in diameters
    std::vector<int> sent[hbem_geometry::kSep];
    std::vector<long long> spi[hbem_geometry::kSep];
    // (line and arc panels only: a trigonometric or spline panel's |gamma'| has
    // complex branch points a few panel lengths away, which bound every rule
    // on the panel whatever the separation)
    auto separation = [&](int i, int j) {
        if ((int)g->panels[8 * (size_t)i] >= 2 || (int)g->panels[8 * (size_t)j] >= 2) return 0;
        const double* a = &disc[3 * (size_t)i];
        const double* b = &disc[3 * (size_t)j];
        const double eta = (std::hypot(a[0] - b[0], a[1] - b[1]) - a[2] - b[2]) / (2.0 * std::max(a[2], b[2]));
        return eta >= 16.0 ? 4 : (eta >= 8.0 ? 3 : (eta >= 2.0 ? 2 : (eta >= 1.0 ? 1 : 0)));
    };
    const double tol = 100.0 * hbem::kDblEps;
    for (int i = 0; i < N; ++i)
        for (int j = 0; j <= i; ++j) {
            long long p = (long long)i * (i + 1) / 2 + j;
            const double* ei = &ends[4 * (size_t)i];
            const double* ej = &ends[4 * (size_t)j];
            if (i == j) { cent.push_back(i); cent.push_back(i); cent.push_back(0); cpi.push_back(p); }
            else if (std::hypot(ei[2] - ej[0], ei[3] - ej[1]) < tol) {            // end(i) == start(j)
                aent.push_back(i); aent.push_back(j); aent.push_back(0); api.push_back(p);
            } else if (std::hypot(ei[0] - ej[2], ei[1] - ej[3]) < tol) {          // start(i) == end(j)
                aent.push_back(i); aent.push_back(j); aent.push_back(1); api.push_back(p);
            } else {
                const int s = separation(i, j);
                sent[s].push_back(i); sent[s].push_back(j); sent[s].push_back(0); spi[s].push_back(p);
            }
        }
    for (int s = 0; s < hbem_geometry::kSep; ++s) {
        g->sep[s] = (int)spi[s].size();
        gent.insert(gent.end(), sent[s].begin(), sent[s].end());
        gpi.insert(gpi.end(), spi[s].begin(), spi[s].end());
        std::vector<int>().swap(sent[s]);
        std::vector<long long>().swap(spi[s]);
    }
    g->ngen = (int)gpi.size();
    g->nadj = (int)api.size();
    g->ncoi = (int)cpi.size();
    hipStream_t st = g->st;
    const double* dpan = upload(g->d_panels, g->panels, st);
    std::vector<double> cf = g->coefs.empty() ? std::vector<double>(3, 0.0) : g->coefs;
    const double* dcoef = upload(g->d_coefs, cf, st);
    const double* dxg = upload(g->d_xg, g->xg, st);
    upload(g->d_wg, g->wg, st);
    const double* dxc = upload(g->d_xc, g->xc, st);
    upload(g->d_wc, g->wc, st);
    double* gtab = g->d_gtab.as<double>((size_t)N * g->ns * 5);
    hbem::launch_setup_gauss_table(dpan, dcoef, N, dxg, g->ns, gtab, st);
    upload(g->d_gent, gent, st);
    upload(g->d_gpidx, gpi, st);
    g->d_gflags.get(std::max(1, g->ngen));
    double* gngeo = g->d_gngeo.as<double>(std::max<size_t>(1, (size_t)g->ngen * g->ns * g->ns * hbem::kNodeStride));
    hbem::launch_setup_general((const int*)g->d_gent.p, g->ngen, gtab, g->ns, gngeo, (unsigned char*)g->d_gflags.p, st);
    upload(g->d_aent, aent, st);
    upload(g->d_apidx, api, st);
    g->d_aflags.get(std::max(1, g->nadj));
    double* angeo = g->d_angeo.as<double>(std::max<size_t>(1, (size_t)g->nadj * 2 * g->nc * g->nc * hbem::kNodeStride));
    hbem::launch_setup_singular(1, (const int*)g->d_aent.p, g->nadj, dpan, dcoef, dxc, g->nc, angeo,
                                (unsigned char*)g->d_aflags.p, st);
    upload(g->d_cent, cent, st);
    upload(g->d_cpidx, cpi, st);
    g->d_cflags.get(std::max(1, g->ncoi));
    double* cngeo = g->d_cngeo.as<double>(std::max<size_t>(1, (size_t)g->ncoi * g->nc * g->nc * hbem::kNodeStride));
    hbem::launch_setup_singular(0, (const int*)g->d_cent.p, g->ncoi, dpan, dcoef, dxc, g->nc, cngeo,
                                (unsigned char*)g->d_cflags.p, st);
    HIPC(hipGetLastError());
    // mass matrix and projection factor (gen_sol_op.cpp:23-43)
    g->M.resize((size_t)N * N);
    hbem::mass_matrix_host(g->panels.data(), N, cptr, g->lines, order, g->M.data());
    upload(g->d_M, g->M, st);
    if (g->proj) {
        const int n = 2 * N;
        std::vector<double> A((size_t)n * n, 0.0);
        for (int j = 0; j < N; ++j)
            for (int i = 0; i < N; ++i) {
                A[i + (size_t)j * n] = g->M[i + (size_t)j * N];
                A[(N + i) + (size_t)(N + j) * n] = g->M[j + (size_t)i * N];
            }
        g->F.resize((size_t)n * n);
        g->tr.resize(n);
        hbem::ldlt_projection_factor(A.data(), n, g->F.data(), g->tr.data());
        g->Finv.resize((size_t)n * n);
        hbem::invert_projection_factor(g->F.data(), g->tr.data(), n, g->Finv.data());
        g->s_F.build(g->F, n, false, st);
        g->s_FT.build(g->F, n, true, st);
        g->s_FinvT.build(g->Finv, n, true, st);
        g->s_Finv.build(g->Finv, n, false, st);
        std::vector<cdbl> Fic((size_t)n * n);
        for (size_t e = 0; e < Fic.size(); ++e) Fic[e] = hbem::mk(g->Finv[e], 0.0);
        upload(g->d_Finvc, Fic, st);
    }
    HIPC(hipStreamSynchronize(st));
    *out = g;
    g = nullptr;
    }
    catch (const std::exception& e) {
        if (g) { if (g->st) (void)hipStreamDestroy(g->st); delete g; }
        return fail(e);
    }
    return HBEM_OK;
}

int hbem_geometry_destroy(hbem_geometry_t g) {
    API_BEGIN
    if (!g) return HBEM_OK;
    multi_forget(g);
    (void)hipStreamSynchronize(g->st);
    hipStream_t st = g->st;
    delete g;
    (void)hipStreamDestroy(st);
    API_END
}

int hbem_geometry_release_work(hbem_geometry_t g) {
    API_BEGIN
    if (!g) throw Error(HBEM_EINVAL, "null geometry");
    multi_forget(g);
    std::lock_guard<std::mutex> lk(g->mu);
    HIPC(hipSetDevice(g->dev));
    HIPC(hipStreamSynchronize(g->st));
    DBuf* bs[] = {&g->w_Ao, &g->w_U, &g->w_rhs, &g->w_flag, &g->w_Tp, &g->w_B, &g->w_mv, &g->w_bpiv, &g->w_bperm,
                  &g->w_bvec, &g->w_d2, &g->w_T32, &g->w_rb, &g->w_rr, &g->w_rpart, &g->w_flag2, &g->w_rprev, &g->w_kapi, &g->w_kri,
                  &g->w_L, &g->w_T0, &g->w_T1, &g->w_T2, &g->w_piv, &g->w_perm, &g->w_X, &g->w_Y, &g->w_work,
                  &g->w_R, &g->w_u, &g->w_v, &g->w_a, &g->w_b, &g->w_sig, &g->w_dsig, &g->w_kap, &g->w_kr, &g->w_cv,
                  &g->w_qi, &g->w_qo, &g->w_tmp, &g->w_out, &g->w_part};
    for (DBuf* b : bs) {
        if (b->p) (void)hipFree(b->p);
        b->p = nullptr;
        b->n = 0;
    }
    API_END
}

int hbem_geometry_info(hbem_geometry_t g, int* npanels, int* order, int* corder, int* lines) {
    API_BEGIN
    if (!g) throw Error(HBEM_EINVAL, "null geometry");
    if (npanels) *npanels = g->N;
    if (order) *order = g->ns;
    if (corder) *corder = g->nc;
    if (lines) *lines = g->lines ? 1 : 0;
    API_END
}

int hbem_assemble_dense(hbem_geometry_t g, double k_re, double k_im, double c, int layers, int d, void* K, void* W,
                        void* V) {
    API_BEGIN
    if (!g) throw Error(HBEM_EINVAL, "null geometry");
    if (d < 0 || d > 2) throw Error(HBEM_EINVAL, "Invalid order of derivative");
    check_k(k_re, k_im, c);
    std::lock_guard<std::mutex> lk(g->mu);
    HIPC(hipSetDevice(g->dev));
    const double sc = std::sqrt(c);
    cdbl* L = real_positive(k_re, k_im) ? assemble_locals(g, {k_re * sc}, {k_re}, {c}, d)
                                        : assemble_locals(g, {k_re * sc}, {k_re}, {c}, d, {k_im * sc}, {k_im});
    const int N = g->N;
    const size_t NN = (size_t)N * N * (d + 1);
    cdbl* out = g->w_out.as<cdbl>(3 * NN);
    hbem::launch_gather_layers(L, 1, 16 * (d + 1), 0, N, d, layers, out, out + NN, out + 2 * NN, g->st);
    HIPC(hipGetLastError());
    if (layers & HBEM_LAYER_DOUBLE) copy_out_square(out, K, NN, g->st);
    if (layers & HBEM_LAYER_HYPERSINGULAR) copy_out_square(out + NN, W, NN, g->st);
    if (layers & HBEM_LAYER_SINGLE) copy_out_square(out + 2 * NN, V, NN, g->st);
    HIPC(hipStreamSynchronize(g->st));
    API_END
}

int hbem_assemble_dense_batch(hbem_geometry_t g, const double* ks, int nk, double c, int layers, int d, void* K,
                              void* W, void* V) {
    API_BEGIN
    if (!g || (nk > 0 && !ks)) throw Error(HBEM_EINVAL, "null argument");
    if (nk < 0) throw Error(HBEM_EINVAL, "negative number of wavenumbers");
    if (d < 0 || d > 2) throw Error(HBEM_EINVAL, "Invalid order of derivative");
    for (int i = 0; i < nk; ++i) check_k(ks[i], 0.0, c);
    std::lock_guard<std::mutex> lk(g->mu);
    HIPC(hipSetDevice(g->dev));
    const double sc = std::sqrt(c);
    const int N = g->N;
    const size_t NN = (size_t)N * N * (d + 1);
    cdbl* out = g->w_out.as<cdbl>(3 * NN);
    // one wavenumber slot per k: its local matrices and chunk partials
    const size_t nslot = 16 * (size_t)(d + 1);
    const size_t part = (size_t)std::max(g->nadj * adj_chunks(g->nc), g->ncoi * coi_chunks(g->nc));
    const size_t per = ((size_t)g->npairs + part) * nslot * sizeof(cdbl);
    {
        std::vector<double> all(nk);
        for (int i = 0; i < nk; ++i) all[i] = ks[i] * sc;
        prepare_far(g, all);
    }
    const int B = sweep_batch(per, usable_device_bytes(g->w_L.n + g->w_part.n), nk);
    for (int k0 = 0; k0 < nk; k0 += B) {
        const int nb = std::min(B, nk - k0);
        std::vector<double> kap(nb), kr(nb), cv(nb, c);
        bool allpos = true;
        for (int b = 0; b < nb; ++b) {
            kap[b] = ks[k0 + b] * sc;
            kr[b] = ks[k0 + b];
            allpos = allpos && ks[k0 + b] > 0.0;
        }
        const std::vector<double> zero(allpos ? 0 : nb, 0.0);
        cdbl* L = assemble_locals(g, kap, kr, cv, d, zero, zero);
        for (int b = 0; b < nb; ++b) {
            hbem::launch_gather_layers(L, nb, 16 * (d + 1), b, N, d, layers, out, out + NN, out + 2 * NN, g->st);
            HIPC(hipGetLastError());
            const size_t off = (size_t)(k0 + b) * NN * sizeof(cdbl);
            if ((layers & HBEM_LAYER_DOUBLE) && K) copy_out_square(out, (char*)K + off, NN, g->st);
            if ((layers & HBEM_LAYER_HYPERSINGULAR) && W) copy_out_square(out + NN, (char*)W + off, NN, g->st);
            if ((layers & HBEM_LAYER_SINGLE) && V) copy_out_square(out + 2 * NN, (char*)V + off, NN, g->st);
        }
    }
    HIPC(hipStreamSynchronize(g->st));
    API_END
}

int hbem_mass_matrix(hbem_geometry_t g, double* M) {
    API_BEGIN
    if (!g || !M) throw Error(HBEM_EINVAL, "null argument");
    std::memcpy(M, g->M.data(), sizeof(double) * g->M.size());
    API_END
}

int hbem_projection_factor(hbem_geometry_t g, double* F, int* transpositions) {
    API_BEGIN
    if (!g) throw Error(HBEM_EINVAL, "null geometry");
    if (!g->proj) throw Error(HBEM_EINVAL, "geometry was created without projection");
    if (F) std::memcpy(F, g->F.data(), sizeof(double) * g->F.size());
    if (transpositions) std::memcpy(transpositions, g->tr.data(), sizeof(int) * g->tr.size());
    API_END
}

int hbem_gen_sol_op(hbem_geometry_t g, double k_re, double k_im, double c_o, double c_i, int d, void* T, void* Td,
                    void* Tdd) {
    API_BEGIN
    if (!g) throw Error(HBEM_EINVAL, "null geometry");
    if (d < 0 || d > 2) throw Error(HBEM_EINVAL, "Invalid order of derivative");
    check_k(k_re, k_im, c_i);
    check_k(k_re, k_im, c_o);
    std::lock_guard<std::mutex> lk(g->mu);
    HIPC(hipSetDevice(g->dev));
    const double si = std::sqrt(c_i), so = std::sqrt(c_o);
    cdbl* L = real_positive(k_re, k_im)
                  ? assemble_locals(g, {k_re * si, k_re * so}, {k_re, k_re}, {c_i, c_o}, d)
                  : assemble_locals(g, {k_re * si, k_re * so}, {k_re, k_re}, {c_i, c_o}, d, {k_im * si, k_im * so},
                                    {k_im, k_im});
    const int n = 2 * g->N;
    const size_t nn = (size_t)n * n;
    cdbl* T0 = g->w_T0.as<cdbl>(nn);
    cdbl* T1 = d >= 1 ? g->w_T1.as<cdbl>(nn) : nullptr;
    cdbl* T2 = d >= 2 ? g->w_T2.as<cdbl>(nn) : nullptr;
    std::vector<int> qi = {0}, qo = {1};
    const int* dqi = upload(g->w_qi, qi, g->st);
    const int* dqo = upload(g->w_qo, qo, g->st);
    hbem::launch_form_T(L, 2, 16 * (d + 1), dqi, dqo, 1, g->N, d, (const double*)g->d_M.p, T0, T1, T2, g->st);
    HIPC(hipGetLastError());
    cdbl* Ts[3] = {T0, T1, T2};
    void* hs[3] = {T, Td, Tdd};
    for (int t = 0; t <= d; ++t) {
        if (!hs[t]) continue;
        if (g->proj) project_inplace(g, Ts[t], 1);
        copy_out_square(Ts[t], hs[t], nn, g->st);
    }
    HIPC(hipStreamSynchronize(g->st));
    API_END
}

int hbem_projection(hbem_geometry_t g, const void* T, void* Tp) {
    API_BEGIN
    if (!g || !T || !Tp) throw Error(HBEM_EINVAL, "null argument");
    if (!g->proj) throw Error(HBEM_EINVAL, "geometry was created without projection");
    std::lock_guard<std::mutex> lk(g->mu);
    HIPC(hipSetDevice(g->dev));
    const int n = 2 * g->N;
    const size_t nn = (size_t)n * n;
    cdbl* d = g->w_T0.as<cdbl>(nn);
    HIPC(hipMemcpyAsync(d, T, nn * sizeof(cdbl), hipMemcpyHostToDevice, g->st));
    project_inplace(g, d, 1);
    copy_out_square(d, Tp, nn, g->st);
    HIPC(hipStreamSynchronize(g->st));
    API_END
}

int hbem_tp_A_rhs(hbem_geometry_t g, double k_re, double k_im, double c_o, double c_i, const void* u_inc_N, int nrhs,
                  void* A, void* rhs) {
    API_BEGIN
    check_tp_args(g, u_inc_N, nrhs);
    check_k(k_re, k_im, c_i);
    check_k(k_re, k_im, c_o);
    std::lock_guard<std::mutex> lk(g->mu);
    HIPC(hipSetDevice(g->dev));
    cdbl *T = nullptr, *r = nullptr;
    tp_system(g, k_re, k_im, c_o, c_i, u_inc_N, nrhs, T, r);
    const int n = 2 * g->N;
    if (A) {
        if (g->proj) project_inplace(g, T, 1);
        copy_out_square(T, A, (size_t)n * n, g->st);
    }
    if (rhs) copy_out_square(r, rhs, (size_t)n * nrhs, g->st);
    HIPC(hipStreamSynchronize(g->st));
    API_END
}

int hbem_tp_solve(hbem_geometry_t g, double k_re, double k_im, double c_o, double c_i, const void* u_inc_N, int nrhs,
                  void* sol) {
    API_BEGIN
    check_tp_args(g, u_inc_N, nrhs);
    if (!sol) throw Error(HBEM_EINVAL, "null argument");
    check_k(k_re, k_im, c_i);
    check_k(k_re, k_im, c_o);
    std::lock_guard<std::mutex> lk(g->mu);
    HIPC(hipSetDevice(g->dev));
    cdbl *T = nullptr, *r = nullptr;
    tp_system(g, k_re, k_im, c_o, c_i, u_inc_N, nrhs, T, r);
    const int n = 2 * g->N;
    const size_t nn = (size_t)n * n;
    // solve (solvers.cpp:133-152): finiteness of rhs and T, then zgesv
    int* flag = g->w_flag.as<int>(3);
    HIPC(hipMemsetAsync(flag, 0, 3 * sizeof(int), g->st));
    hbem::flag_nonfinite((const double*)r, 2LL * n * nrhs, flag, g->st);
    hbem::flag_nonfinite((const double*)T, 2LL * (long long)nn, flag + 1, g->st);
    HIPC(hipGetLastError());
    int hf[3] = {0, 0, 0};
    HIPC(hipMemcpyAsync(hf, flag, 2 * sizeof(int), hipMemcpyDeviceToHost, g->st));
    HIPC(hipStreamSynchronize(g->st));
    if (hf[0]) throw Error(HBEM_ENUMERIC, "Failed to solve for traces, right-hand side is undefined");
    if (hf[1]) throw Error(HBEM_ENUMERIC, "Failed to solve for traces, solutions operator is undefined");
    int* piv = g->w_piv.as<int>(n);
    int* perm = g->w_perm.as<int>(n);
    hbem::lu_factor_batched(T, n, 1, piv, perm, g->st);
    hbem::flag_zero_diag(T, n, flag + 2, g->st);
    cdbl* work = g->w_work.as<cdbl>((size_t)n * 8);
    for (int j = 0; j < nrhs; j += 8) {
        const int nc = std::min(8, nrhs - j);
        hbem::lu_solve_thin_batched(T, (long long)n * n, perm, n, 1, r + (size_t)j * n, nc, false, work, g->st);
    }
    HIPC(hipGetLastError());
    HIPC(hipMemcpyAsync(hf + 2, flag + 2, sizeof(int), hipMemcpyDeviceToHost, g->st));
    HIPC(hipStreamSynchronize(g->st));
    if (hf[2]) throw Error(HBEM_ENUMERIC, "zgesv: LU is singular");
    copy_out_square(r, sol, (size_t)n * nrhs, g->st);
    HIPC(hipStreamSynchronize(g->st));
    API_END
}

}  // extern "C"

namespace {

// Standalone dense context for randomized_svd::* on caller matrices.  The
// reference keeps the randomized_svd timers in file-scope statics shared by
// every call (randsvd.cpp:27-47); here they are this context's event log,
// filled while benchmarking is on and read when a timer is queried.
struct Standalone {
    std::mutex mu;
    hipStream_t st = nullptr;
    DBuf T, T1, T2, piv, perm, X, Y, work, u, v, sig, out, W, Wpiv, Wperm, p, p2, ud, tmp, Wq, cpn, cps, cpl;
    bool bench = false;
    EventLog tim;
    void init() {
        if (!st) {
            HIPC(hipSetDevice(g_device));
            HIPC(hipStreamCreateWithFlags(&st, hipStreamNonBlocking));
        }
    }
};
Standalone g_sa;

// The eigenvector of the Wielandt matrix W = [[0, T], [T^H, 0]] for sigma as
// randomized_svd::sv_der / sv_der2 take it: the last column of Q of the
// column-pivoted QR of W - sigma I (randsvd.cpp:160-162).  That column is
// orthogonal to every column of the Hermitian W - sigma I but the last
// pivot p, so it is (W - sigma I)^{-1} e_p normalised (up to a phase, which
// neither sigma' nor sigma'' sees): the pivot order comes from
// colpiv_last_pivot_batched, the solve from the blocked LU.  T is the
// caller's host matrix; s.T is overwritten.
void reference_wielandt_vector(Standalone& s, const void* T, int n, double sigma, std::vector<cdbl>& x) {
    const size_t nn = (size_t)n * n;
    const int m2 = 2 * n;
    const size_t mm = (size_t)m2 * m2;
    cdbl* dT = s.T.as<cdbl>(nn);
    HIPC(hipMemcpyAsync(dT, T, nn * sizeof(cdbl), hipMemcpyHostToDevice, s.st));
    cdbl* Wq = s.Wq.as<cdbl>(mm);
    hbem::wielandt_build(dT, n, sigma, Wq, s.st);
    int* dl = s.cpl.as<int>(1);
    hbem::colpiv_last_pivot_batched(Wq, (long long)mm, m2, 1, s.cpn.as<double>(2 * (size_t)m2),
                                    s.cps.as<int>(2 * (size_t)m2), dl, s.st);
    int p = 0;
    HIPC(hipMemcpyAsync(&p, dl, sizeof(int), hipMemcpyDeviceToHost, s.st));
    HIPC(hipStreamSynchronize(s.st));
    if (p < 0 || p >= m2) throw Error(HBEM_ENUMERIC, "pivoted QR of the Wielandt matrix failed");
    hbem::wielandt_build(dT, n, sigma, Wq, s.st);
    int* wpiv = s.Wpiv.as<int>(m2);
    int* wperm = s.Wperm.as<int>(m2);
    hbem::lu_factor_batched(Wq, m2, 1, wpiv, wperm, s.st);
    std::vector<cdbl> e(m2, hbem::mk(0, 0));
    e[p] = hbem::mk(1, 0);
    cdbl* de = s.ud.as<cdbl>(m2);
    HIPC(hipMemcpyAsync(de, e.data(), m2 * sizeof(cdbl), hipMemcpyHostToDevice, s.st));
    cdbl* wk = s.work.as<cdbl>(2 * (size_t)m2);
    hbem::lu_solve_thin_batched(Wq, (long long)mm, wperm, m2, 1, de, 1, false, wk, s.st);
    x.resize(m2);
    HIPC(hipMemcpyAsync(x.data(), de, m2 * sizeof(cdbl), hipMemcpyDeviceToHost, s.st));
    HIPC(hipStreamSynchronize(s.st));
    double nrm = 0.0;
    for (const auto& z : x) nrm += z.x * z.x + z.y * z.y;
    nrm = std::sqrt(nrm);
    if (!(nrm > 0.0) || !std::isfinite(nrm)) throw Error(HBEM_ENUMERIC, "singular Wielandt matrix");
    for (auto& z : x) z = hbem::mk(z.x / nrm, z.y / nrm);
}

void check_dense_args(const void* T, int n, const void* R, int nc, int q) {
    if (!T || !R) throw Error(HBEM_EINVAL, "null matrix");
    if (n < 1) throw Error(HBEM_EINVAL, "matrix order must be positive");
    if (nc < 1 || nc > 8 || nc > n) throw Error(HBEM_EINVAL, "number of random columns must be in [1, min(8, n)]");
    if (q < 0) throw Error(HBEM_EINVAL, "number of subspace iterations must be >= 0");
}

// Runs LU + rsvd on one host matrix; leaves LU/u/v on the device.
double standalone_sv(const void* T, int n, const void* R, int nc, int q, bool want_vectors) {
    Standalone& s = g_sa;
    s.init();
    const size_t nn = (size_t)n * n;
    cdbl* dT = s.T.as<cdbl>(nn);
    HIPC(hipMemcpyAsync(dT, T, nn * sizeof(cdbl), hipMemcpyHostToDevice, s.st));
    int* piv = s.piv.as<int>(n);
    int* perm = s.perm.as<int>(n);
    hipEvent_t e = s.bench ? s.tim.begin(s.st) : nullptr;
    hbem::lu_factor_batched(dT, n, 1, piv, perm, s.st);
    if (s.bench) s.tim.end(s.st, kTimLU, e, (8.0 / 3.0) * (double)n * n * n);
    DenseCtx c;
    c.st = s.st; c.n = n; c.batch = 1; c.nc = nc; c.q = q; c.LU = dT; c.ms = (long long)n * n; c.perm = perm;
    c.F = nullptr; c.FT = nullptr;
    c.tim = s.bench ? &s.tim : nullptr;
    c.X = s.X.as<cdbl>((size_t)n * nc);
    c.Y = s.Y.as<cdbl>((size_t)n * nc);
    c.work = s.work.as<cdbl>(2 * (size_t)n * nc);
    HIPC(hipMemcpyAsync(c.X, R, (size_t)n * nc * sizeof(cdbl), hipMemcpyHostToDevice, s.st));
    double* dsig = s.sig.as<double>(1);
    cdbl* du = want_vectors ? s.u.as<cdbl>(n) : nullptr;
    cdbl* dv = want_vectors ? s.v.as<cdbl>(n) : nullptr;
    rsvd_core(c, dsig, du, dv);
    HIPC(hipGetLastError());
    double sigma = 0.0;
    HIPC(hipMemcpyAsync(&sigma, dsig, sizeof(double), hipMemcpyDeviceToHost, s.st));
    HIPC(hipStreamSynchronize(s.st));
    return sigma;
}

}  // namespace

extern "C" {

int hbem_rsvd_sv(const void* T, int n, const void* R, int nc, int q, double* sigma) {
    API_BEGIN
    check_dense_args(T, n, R, nc, q);
    if (!sigma) throw Error(HBEM_EINVAL, "null output");
    std::lock_guard<std::mutex> lk(g_sa.mu);
    *sigma = standalone_sv(T, n, R, nc, q, false);
    API_END
}

// The kernels of hbem_rsvd_sv with every work buffer carved out of this
// device's `scratch`: a device-to-device copy of T, the blocked LU, the
// randomized iteration, sigma written to the caller's device double.  Only
// enqueues work on `stream`; nothing here allocates or waits.  Every buffer is
// written before it is read, so the array needs no clearing between calls.
int hbem_rsvd_sv_noalloc(const void* T_dev, int n, const void* R_dev, int nc, int q, double* sigma_dev, void* stream) {
    API_BEGIN
    check_dense_args(T_dev, n, R_dev, nc, q);
    if (!sigma_dev) throw Error(HBEM_EINVAL, "null output");
    if (n > kNaMaxN) throw Error(HBEM_EINVAL, "matrix order exceeds the no-allocation limit of 2048");
    int dev = 0;
    HIPC(hipGetDevice(&dev));
    unsigned char* base = (unsigned char*)scratch_on(dev);
    const hipStream_t st = (hipStream_t)stream;
    const size_t nn = (size_t)n * n;
    cdbl* LU = (cdbl*)base;
    base += al256(sizeof(cdbl) * nn);
    int* piv = (int*)base;
    base += al256(sizeof(int) * (size_t)n);
    int* perm = (int*)base;
    base += al256(sizeof(int) * (size_t)n);
    const size_t thin = al256(sizeof(cdbl) * (size_t)n * nc);
    DenseCtx c;
    c.st = st; c.n = n; c.batch = 1; c.nc = nc; c.q = q; c.LU = LU; c.ms = (long long)nn; c.perm = perm;
    c.F = nullptr; c.FT = nullptr;
    c.X = (cdbl*)base;
    c.Y = (cdbl*)(base + thin);
    c.work = (cdbl*)(base + 2 * thin);
    HIPC(hipMemcpyAsync(LU, T_dev, nn * sizeof(cdbl), hipMemcpyDeviceToDevice, st));
    hbem::lu_factor_batched(LU, n, 1, piv, perm, st);
    HIPC(hipMemcpyAsync(c.X, R_dev, (size_t)n * nc * sizeof(cdbl), hipMemcpyDeviceToDevice, st));
    rsvd_core(c, sigma_dev, nullptr, nullptr);
    HIPC(hipGetLastError());
    API_END
}

int hbem_rsvd_sv_der(const void* T, const void* Td, int n, const void* R, int nc, int q, double* out2) {
    API_BEGIN
    check_dense_args(T, n, R, nc, q);
    if (!Td || !out2) throw Error(HBEM_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(g_sa.mu);
    Standalone& s = g_sa;
    double sigma = standalone_sv(T, n, R, nc, q, true);
    const size_t nn = (size_t)n * n;
    const int m2 = 2 * n;
    cdbl* dT1 = s.T1.as<cdbl>(nn);
    HIPC(hipMemcpyAsync(dT1, Td, nn * sizeof(cdbl), hipMemcpyHostToDevice, s.st));
    std::vector<cdbl> x;
    reference_wielandt_vector(s, T, n, sigma, x);
    // sigma' = Re(x^H [T' x_2; T'^H x_1]) (randsvd.cpp:163-167)
    cdbl* dx = s.p.as<cdbl>(m2);
    cdbl* dp = s.p2.as<cdbl>(m2);
    HIPC(hipMemcpyAsync(dx, x.data(), m2 * sizeof(cdbl), hipMemcpyHostToDevice, s.st));
    hbem::gemv_batched(dT1, 0, dx + n, dp, n, 1, false, s.st);
    hbem::gemv_batched(dT1, 0, dx, dp + n, n, 1, true, s.st);
    std::vector<cdbl> hp(m2);
    HIPC(hipMemcpyAsync(hp.data(), dp, m2 * sizeof(cdbl), hipMemcpyDeviceToHost, s.st));
    HIPC(hipStreamSynchronize(s.st));
    double der = 0.0;
    for (int l = 0; l < m2; ++l) der += x[l].x * hp[l].x + x[l].y * hp[l].y;
    out2[0] = sigma;
    out2[1] = der;
    API_END
}

// randomized_svd::sv_der2 (randsvd.cpp:171-216): the eigenvector x of the
// Wielandt matrix W = [[0,T],[T^H,0]] for sigma as the reference takes it
// (reference_wielandt_vector); the bordered system follows the reference.
int hbem_rsvd_sv_der2(const void* T, const void* Td, const void* Tdd, int n, const void* R, int nc, int q,
                      double* out3) {
    API_BEGIN
    check_dense_args(T, n, R, nc, q);
    if (!Td || !Tdd || !out3) throw Error(HBEM_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(g_sa.mu);
    Standalone& s = g_sa;
    double sigma = standalone_sv(T, n, R, nc, q, true);
    const size_t nn = (size_t)n * n;
    const int m2 = 2 * n;
    cdbl* dT1 = s.T1.as<cdbl>(nn);
    cdbl* dT2 = s.T2.as<cdbl>(nn);
    HIPC(hipMemcpyAsync(dT1, Td, nn * sizeof(cdbl), hipMemcpyHostToDevice, s.st));
    HIPC(hipMemcpyAsync(dT2, Tdd, nn * sizeof(cdbl), hipMemcpyHostToDevice, s.st));
    std::vector<cdbl> x;
    reference_wielandt_vector(s, T, n, sigma, x);
    // sigma' = Re(x^H [T' x_2; T'^H x_1]); then u = x normalised by its
    // largest-modulus entry (randsvd.cpp:185-198)
    cdbl* dx = s.p.as<cdbl>(m2);
    cdbl* dpx = s.p2.as<cdbl>(m2);
    HIPC(hipMemcpyAsync(dx, x.data(), m2 * sizeof(cdbl), hipMemcpyHostToDevice, s.st));
    hbem::gemv_batched(dT1, 0, dx + n, dpx, n, 1, false, s.st);
    hbem::gemv_batched(dT1, 0, dx, dpx + n, n, 1, true, s.st);
    std::vector<cdbl> hpx(m2);
    HIPC(hipMemcpyAsync(hpx.data(), dpx, m2 * sizeof(cdbl), hipMemcpyDeviceToHost, s.st));
    HIPC(hipStreamSynchronize(s.st));
    double der = 0.0;
    for (int l = 0; l < m2; ++l) der += x[l].x * hpx[l].x + x[l].y * hpx[l].y;
    double best = 0.0;
    int m = 5 < m2 ? 5 : 0;
    for (int l = 0; l < m2; ++l) {
        double a = std::sqrt(x[l].x * x[l].x + x[l].y * x[l].y);
        if (a > best) { best = a; m = l; }
    }
    const std::complex<double> xm(x[m].x, x[m].y);
    std::vector<cdbl> u(m2);
    for (int l = 0; l < m2; ++l) {
        std::complex<double> z = std::complex<double>(x[l].x, x[l].y) / xm;
        u[l] = hbem::mk(z.real(), z.imag());
    }
    cdbl* W = s.W.as<cdbl>((size_t)m2 * m2);
    // s.T holds the LU factors of T; W is built from the caller's T
    HIPC(hipMemcpyAsync(s.T.p, T, nn * sizeof(cdbl), hipMemcpyHostToDevice, s.st));
    hbem::wielandt_build((const cdbl*)s.T.p, n, sigma, W, s.st);
    cdbl* du = s.p.as<cdbl>(m2);
    HIPC(hipMemcpyAsync(du, u.data(), m2 * sizeof(cdbl), hipMemcpyHostToDevice, s.st));
    std::vector<cdbl> negu(m2);
    for (int l = 0; l < m2; ++l) negu[l] = hbem::mk(-u[l].x, -u[l].y);
    HIPC(hipMemcpyAsync(W + (size_t)m * m2, negu.data(), m2 * sizeof(cdbl), hipMemcpyHostToDevice, s.st));
    int* wpiv = s.Wpiv.as<int>(m2);
    int* wperm = s.Wperm.as<int>(m2);
    hbem::lu_factor_batched(W, m2, 1, wpiv, wperm, s.st);
    // p = [T' u2; T'^H u1]
    cdbl* p = s.p2.as<cdbl>(m2);
    hbem::gemv_batched(dT1, 0, du + n, p, n, 1, false, s.st);
    hbem::gemv_batched(dT1, 0, du, p + n, n, 1, true, s.st);
    cdbl* wk = s.work.as<cdbl>(2 * (size_t)m2);
    hbem::lu_solve_thin_batched(W, (long long)m2 * m2, wperm, m2, 1, p, 1, false, wk, s.st);   // p <- W^{-1} p
    std::vector<cdbl> ud(m2);
    HIPC(hipMemcpyAsync(ud.data(), p, m2 * sizeof(cdbl), hipMemcpyDeviceToHost, s.st));
    HIPC(hipStreamSynchronize(s.st));
    for (auto& z : ud) z = hbem::mk(-z.x, -z.y);   // u_der = -W^{-1} p
    const std::complex<double> t(ud[m].x, ud[m].y);
    ud[m] = hbem::mk(0, 0);
    cdbl* dud = s.ud.as<cdbl>(m2);
    HIPC(hipMemcpyAsync(dud, ud.data(), m2 * sizeof(cdbl), hipMemcpyHostToDevice, s.st));
    cdbl* pp = s.tmp.as<cdbl>(2 * (size_t)m2);
    cdbl* p2 = pp + m2;
    hbem::gemv_batched(dT1, 0, dud + n, pp, n, 1, false, s.st);
    hbem::gemv_batched(dT1, 0, dud, pp + n, n, 1, true, s.st);
    hbem::gemv_batched(dT2, 0, du + n, p2, n, 1, false, s.st);
    hbem::gemv_batched(dT2, 0, du, p2 + n, n, 1, true, s.st);
    std::vector<cdbl> hp(m2), hp2(m2);
    HIPC(hipMemcpyAsync(hp.data(), pp, m2 * sizeof(cdbl), hipMemcpyDeviceToHost, s.st));
    HIPC(hipMemcpyAsync(hp2.data(), p2, m2 * sizeof(cdbl), hipMemcpyDeviceToHost, s.st));
    HIPC(hipStreamSynchronize(s.st));
    std::vector<cdbl> rhs(m2);
    for (int l = 0; l < m2; ++l) {
        std::complex<double> v = std::complex<double>(hp2[l].x, hp2[l].y) +
                                 2.0 * (std::complex<double>(hp[l].x, hp[l].y) - t * std::complex<double>(ud[l].x, ud[l].y));
        rhs[l] = hbem::mk(v.real(), v.imag());
    }
    HIPC(hipMemcpyAsync(p, rhs.data(), m2 * sizeof(cdbl), hipMemcpyHostToDevice, s.st));
    hbem::lu_solve_thin_batched(W, (long long)m2 * m2, wperm, m2, 1, p, 1, false, wk, s.st);
    cdbl sol;
    HIPC(hipMemcpyAsync(&sol, p + m, sizeof(cdbl), hipMemcpyDeviceToHost, s.st));
    HIPC(hipStreamSynchronize(s.st));
    out3[0] = sigma;
    out3[1] = der;
    out3[2] = -sol.x;
    API_END
}

}  // extern "C"

namespace {

// out_head = T_p^(t) x_tail, out_tail = T_p^(t)H x_head for the 2n-vectors X of
// the batch (2n x nb), with T_p^(t) = F^{-1} Tt F^{-T} (Tt: n x n x nb) or Tt
// itself without projection.  h: scratch of 4 n nb complex.
void apply_pair_op(hbem_geometry* g, const cdbl* Tt, const cdbl* X, cdbl* out, int nb, cdbl* h) {
    const int n = 2 * g->N;
    const long long nn = (long long)n * n, m2 = 2LL * n;
    cdbl *xh = h, *xt = h + (size_t)n * nb, *z = h + 2 * (size_t)n * nb, *y = h + 3 * (size_t)n * nb;
    hbem::copy_blocks_batched(X, m2, 0, xh, n, 0, n, nb, g->st);
    hbem::copy_blocks_batched(X, m2, n, xt, n, 0, n, nb, g->st);
    for (int part = 0; part < 2; ++part) {
        const cdbl* src = part == 0 ? xt : xh;
        const bool adj = part == 1;
        const long long off = part == 0 ? 0 : n;
        if (g->proj) {
            const DCsr& A = g->s_FinvT;
            const DCsr& Bm = g->s_Finv;
            hbem::csr_apply_batched((const int*)A.rp.p, (const int*)A.ci.p, (const double*)A.v.p, n, src, z, nb, g->st);
            hbem::gemv_batched(Tt, nn, z, y, n, nb, adj, g->st);
            hbem::csr_apply_batched((const int*)Bm.rp.p, (const int*)Bm.ci.p, (const double*)Bm.v.p, n, y, z, nb, g->st);
            hbem::copy_blocks_batched(z, n, 0, out, m2, off, n, nb, g->st);
        } else {
            hbem::gemv_batched(Tt, nn, src, y, n, nb, adj, g->st);
            hbem::copy_blocks_batched(y, n, 0, out, m2, off, n, nb, g->st);
        }
    }
}

// The sweep shared by hbem_sigma_sweep and hbem_sigma_sweep_der2: sigma for
// every k, dsigma/dk when dsigma is given (operators of order D >= 1), and
// d2sigma/dk2 when d2sigma is given (D = 2).
// sigma of the matrices idx of a batch from their FP64 T (T0, stride n^2) with
// the FP64 factorisation: the fallback of the mixed-precision sweep
void sweep_fp64_subset(hbem_geometry* g, const cdbl* T0, const std::vector<int>& idx, const cdbl* dR, int nc, int q,
                       double* sigma_host) {
    const int n = 2 * g->N, m = (int)idx.size();
    const size_t nn = (size_t)n * n;
    if (m == 0) return;
    cdbl* Tf = g->w_T1.as<cdbl>(nn * m);
    for (int j = 0; j < m; ++j)
        HIPC(hipMemcpyAsync(Tf + j * nn, T0 + (size_t)idx[j] * nn, nn * sizeof(cdbl), hipMemcpyDeviceToDevice, g->st));
    int* piv = g->w_bpiv.as<int>((size_t)n * m);
    int* perm = g->w_bperm.as<int>((size_t)n * m);
    hbem::lu_factor_batched(Tf, n, m, piv, perm, g->st);
    DenseCtx c;
    c.st = g->st; c.n = n; c.batch = m; c.nc = nc; c.q = q; c.LU = Tf; c.ms = (long long)nn; c.perm = perm;
    c.F = g->proj ? &g->s_F : nullptr;
    c.FT = g->proj ? &g->s_FT : nullptr;
    c.X = g->w_X.as<cdbl>((size_t)n * nc * m);
    c.Y = g->w_Y.as<cdbl>((size_t)n * nc * m);
    c.work = g->w_work.as<cdbl>(2 * (size_t)n * nc * m);
    hbem::replicate_batched(dR, (long long)n * nc, m, c.X, g->st);
    double* dsig = g->w_d2.as<double>(m);
    rsvd_core(c, dsig, nullptr, nullptr);
    std::vector<double> out(m);
    HIPC(hipMemcpyAsync(out.data(), dsig, m * sizeof(double), hipMemcpyDeviceToHost, g->st));
    HIPC(hipStreamSynchronize(g->st));
    for (int j = 0; j < m; ++j) sigma_host[idx[j]] = out[j];
}

// mixed >= 0 (D = 0 only): the dense stage in mixed precision with `mixed`
// refinement steps per solve; fp64_used (may be null) marks the k recomputed
// in FP64.
void sweep_run(hbem_geometry* g, const double* ks, int nk, double c_o, double c_i, const void* R, int nc, int q, int D,
               double* sigma, double* dsigma, double* d2sigma, int mixed = -1, int* fp64_used = nullptr,
               int mixed_inner = -1) {
    const int n = 2 * g->N;
    const size_t nn = (size_t)n * n;
    const bool mx = mixed >= 0 && D == 0 && !d2sigma;
    // Batch of B wavenumbers -> slots in blocks of 32 consecutive k of one
    // refraction index: each wavefront of the assembly sees a single index
    // and neighbouring k, so the arguments kappa r of its lanes stay close
    // (fewer divergent Bessel branches).  B <= 256 gives the dense stage (one
    // workgroup per matrix in the panel factorisation and the thin solves)
    // up to one matrix per CU.  B is bounded by the free device memory: local
    // matrices, chunk partials and T, T', T'' per wavenumber (and the
    // bordered systems of the sigma'' stage).
    const bool tdiff = sweep_tdiff(ks, nk);
    const size_t per = sweep_bytes_per_k(g->npairs, g->nadj, g->ncoi, g->nc, n, D, nc, tdiff ? 1 : 2) +
                       (d2sigma ? der2_bytes_per_k(n) : 0) +
                       (mx ? mixed_bytes_per_k(n, nc) : 0);
    {
        std::vector<double> all;
        for (int i = 0; i < nk; ++i)
            if (ks[i] > 0.0) { all.push_back(ks[i] * std::sqrt(c_i)); all.push_back(ks[i] * std::sqrt(c_o)); }
        prepare_far(g, all);
    }
    const int B = sweep_batch(per, usable_device_bytes(g->work_held()), nk);
    cdbl* dR = g->w_R.as<cdbl>((size_t)n * nc);
    HIPC(hipMemcpyAsync(dR, R, (size_t)n * nc * sizeof(cdbl), hipMemcpyHostToDevice, g->st));
    for (int k0 = 0; k0 < nk; k0 += B) {
        const int nb = std::min(B, nk - k0);
        bool allpos = true;
        for (int b = 0; b < nb; ++b) allpos = allpos && ks[k0 + b] > 0.0;
        std::vector<double> kap, kr, cv;
        std::vector<int> qi(nb), qo(nb);
        if (tdiff) {
            tdiff_slots(ks + k0, nb, c_i, c_o, kap, kr, cv);
            for (int b = 0; b < nb; ++b) qi[b] = b;
        } else {
            // inner slots 0..nb-1, outer nb..2nb-1
            kap.resize(2 * nb); kr.resize(2 * nb); cv.resize(2 * nb);
            for (int b = 0; b < nb; ++b) {
                const double k = ks[k0 + b];
                kap[b] = k * std::sqrt(c_i); kr[b] = k; cv[b] = c_i;
                kap[nb + b] = k * std::sqrt(c_o); kr[nb + b] = k; cv[nb + b] = c_o;
                qi[b] = b;
                qo[b] = nb + b;
            }
        }
        const std::vector<double> zero(allpos ? 0 : kap.size(), 0.0);
        cdbl* L = assemble_locals(g, kap, kr, cv, D, zero, zero, tdiff ? nb : 0);
        const int lcols = tdiff ? nb : 2 * nb;   // columns of the local buffer
        cdbl* T0 = g->w_T0.as<cdbl>(nn * nb);
        cdbl* T1 = D >= 1 ? g->w_T1.as<cdbl>(nn * nb) : nullptr;
        cdbl* T2 = D >= 2 ? g->w_T2.as<cdbl>(nn * nb) : nullptr;
        const int* dqi = upload(g->w_qi, qi, g->st);
        const int* dqo = tdiff ? nullptr : upload(g->w_qo, qo, g->st);
        // mixed precision: T formed in FP64 (kept for the residuals) and in FP32
        // (factorised on the FP32 matrix cores) in one pass
        cflt* T32 = mx ? g->w_T32.as<cflt>(nn * nb) : nullptr;
        hipEvent_t e0 = g->prof ? g->plog.begin(g->st) : nullptr;
        hbem::launch_form_T(L, lcols, 16 * (D + 1), dqi, dqo, nb, g->N, D, (const double*)g->d_M.p, T0, T1, T2, g->st, 0,
                            T32);
        if (g->prof) g->plog.end(g->st, kProfFormT, e0);
        int* piv = g->w_piv.as<int>((size_t)n * nb);
        int* perm = g->w_perm.as<int>((size_t)n * nb);
        if (g->prof) e0 = g->plog.begin(g->st);
        GemmTimer gt(&g->plog, kProfGemm);
        if (mx) {
            hbem::lu_factor_batched_f32(T32, n, nb, piv, perm, g->st, g->prof ? &gt : nullptr);
        } else {
            hbem::lu_factor_batched(T0, n, nb, piv, perm, g->st, g->prof ? &gt : nullptr);
        }
        if (g->prof) {
            g->plog.end(g->st, kProfLU, e0, (8.0 / 3.0) * (double)n * n * n * nb);
            e0 = g->plog.begin(g->st);
        }
        DenseCtx c;
        c.st = g->st; c.n = n; c.batch = nb; c.nc = nc; c.q = q; c.LU = T0; c.ms = (long long)nn; c.perm = perm;
        c.F = g->proj ? &g->s_F : nullptr;
        c.FT = g->proj ? &g->s_FT : nullptr;
        c.X = g->w_X.as<cdbl>((size_t)n * nc * nb);
        c.Y = g->w_Y.as<cdbl>((size_t)n * nc * nb);
        c.work = g->w_work.as<cdbl>(2 * (size_t)n * nc * nb);
        int* flag = nullptr;
        if (mx) {
            c.LU32 = T32;
            c.T64 = T0;
            c.refine = mixed;
            c.refine_inner = mixed_inner < 0 ? mixed : mixed_inner;
            c.rtol = kMixedTol;
            flag = g->w_flag2.as<int>(nb);
            hbem::fill_int(flag, nb, mixed > 0 ? 0 : 1, g->st);   // no refinement: nothing certifies the solves
            c.flag = flag;
            c.rprev = g->w_rprev.as<double>(nb);
            c.rb = g->w_rb.as<cdbl>((size_t)n * nc * nb);
            c.rr = g->w_rr.as<cdbl>((size_t)n * nc * nb);
            c.rpart = g->w_rpart.as<cdbl>(hbem::residual_part_count(n, nc, nb));
        }
        hbem::replicate_batched(dR, (long long)n * nc, nb, c.X, g->st);
        double* dsig = g->w_sig.as<double>(nb);
        cdbl* du = D ? g->w_u.as<cdbl>((size_t)n * nb) : nullptr;
        cdbl* dv = D ? g->w_v.as<cdbl>((size_t)n * nb) : nullptr;
        rsvd_core(c, dsig, du, dv);
        HIPC(hipMemcpyAsync(sigma + k0, dsig, nb * sizeof(double), hipMemcpyDeviceToHost, g->st));
        if (mx) {
            // k whose refinement did not converge: sigma again with the FP64 factorisation
            std::vector<int> hf(nb);
            HIPC(hipMemcpyAsync(hf.data(), flag, nb * sizeof(int), hipMemcpyDeviceToHost, g->st));
            HIPC(hipStreamSynchronize(g->st));
            std::vector<int> idx;
            for (int b = 0; b < nb; ++b) {
                if (hf[b]) idx.push_back(b);
                if (fp64_used) fp64_used[k0 + b] = hf[b] ? 1 : 0;
            }
            sweep_fp64_subset(g, T0, idx, dR, nc, q, sigma + k0);
        }
        if (D) {
            cdbl* a = g->w_a.as<cdbl>((size_t)n * nb);
            cdbl* bb = g->w_b.as<cdbl>((size_t)n * nb);
            if (g->proj) {
                const DCsr& S = g->s_FinvT;
                hbem::csr_apply_batched((const int*)S.rp.p, (const int*)S.ci.p, (const double*)S.v.p, n, du, a, nb, g->st);
                hbem::csr_apply_batched((const int*)S.rp.p, (const int*)S.ci.p, (const double*)S.v.p, n, dv, bb, nb, g->st);
            } else {
                HIPC(hipMemcpyAsync(a, du, (size_t)n * nb * sizeof(cdbl), hipMemcpyDeviceToDevice, g->st));
                HIPC(hipMemcpyAsync(bb, dv, (size_t)n * nb * sizeof(cdbl), hipMemcpyDeviceToDevice, g->st));
            }
            double* dd = g->w_dsig.as<double>((size_t)nb * (1 + hbem::kBilinearParts));
            hbem::bilinear_re_batched(T1, (long long)nn, a, bb, n, nb, dd, dd + nb, g->st);
            if (dsigma) HIPC(hipMemcpyAsync(dsigma + k0, dd, nb * sizeof(double), hipMemcpyDeviceToHost, g->st));
        }
        if (d2sigma) {
            // randomized_svd::sv_der2's bordered system (randsvd.cpp:171-216) for
            // every k of the batch, on T_p formed explicitly: T again from the
            // local matrices (T0 holds its LU factors), projected through the
            // bordered matrix's storage.
            const int m2 = 2 * n;
            cdbl* Tp = g->w_Tp.as<cdbl>(nn * nb);
            cdbl* Bw = g->w_B.as<cdbl>(4 * nn * nb);
            hbem::launch_form_T(L, lcols, 16 * (D + 1), dqi, dqo, nb, g->N, 0, (const double*)g->d_M.p, Tp, nullptr,
                                nullptr, g->st);
            if (g->proj) project_inplace(g, Tp, nb, Bw);
            int* mv = g->w_mv.as<int>(nb);
            cdbl* vec = g->w_bvec.as<cdbl>(16 * (size_t)n * nb);
            cdbl *X = vec, *P = vec + (size_t)m2 * nb, *Pd = vec + 2 * (size_t)m2 * nb, *P2 = vec + 3 * (size_t)m2 * nb;
            cdbl *h = vec + 4 * (size_t)m2 * nb, *wk = vec + 6 * (size_t)m2 * nb, *tv = vec + 7 * (size_t)m2 * nb;
            hbem::border_prep_batched(du, dv, n, nb, mv, X, g->st);
            hbem::border_build_batched(Tp, n, nb, dsig, mv, X, Bw, g->st);
            int* bpiv = g->w_bpiv.as<int>((size_t)m2 * nb);
            int* bperm = g->w_bperm.as<int>((size_t)m2 * nb);
            hbem::lu_factor_batched(Bw, m2, nb, bpiv, bperm, g->st);
            apply_pair_op(g, T1, X, P, nb, h);                       // p = [T_p' u_2; T_p'^H u_1]
            hbem::lu_solve_thin_batched(Bw, 4LL * (long long)nn, bperm, m2, nb, P, 1, false, wk, g->st);
            hbem::border_negate_pick_batched(P, n, nb, mv, tv, g->st);   // P = u', tv = t
            apply_pair_op(g, T1, P, Pd, nb, h);                      // [T_p' u'_2; T_p'^H u'_1]
            apply_pair_op(g, T2, X, P2, nb, h);                      // [T_p'' u_2; T_p''^H u_1]
            hbem::border_rhs_batched(Pd, P2, P, tv, n, nb, Pd, g->st);
            hbem::lu_solve_thin_batched(Bw, 4LL * (long long)nn, bperm, m2, nb, Pd, 1, false, wk, g->st);
            double* d2 = g->w_d2.as<double>(nb);
            hbem::border_final_batched(Pd, n, nb, mv, d2, g->st);
            HIPC(hipMemcpyAsync(d2sigma + k0, d2, nb * sizeof(double), hipMemcpyDeviceToHost, g->st));
        }
        if (g->prof) g->plog.end(g->st, kProfRsvd, e0);
        HIPC(hipGetLastError());
    }
    HIPC(hipStreamSynchronize(g->st));
}

void check_sweep_args(hbem_geometry* g, const double* ks, int nk, double c_o, double c_i, const void* R, int nc, int q,
                      const double* sigma) {
    if (!g || !ks || !R || !sigma) throw Error(HBEM_EINVAL, "null argument");
    if (nk < 0) throw Error(HBEM_EINVAL, "negative number of wavenumbers");
    check_dense_args(ks, 2 * g->N, R, nc, q);
    for (int i = 0; i < nk; ++i) { check_k(ks[i], 0.0, c_i); check_k(ks[i], 0.0, c_o); }
}

}  // namespace

extern "C" {

int hbem_sigma_sweep(hbem_geometry_t g, const double* ks, int nk, double c_o, double c_i, const void* R, int nc, int q,
                     int want_der, double* sigma, double* dsigma) {
    API_BEGIN
    check_sweep_args(g, ks, nk, c_o, c_i, R, nc, q, sigma);
    if (want_der && !dsigma) throw Error(HBEM_EINVAL, "null dsigma");
    std::lock_guard<std::mutex> lk(g->mu);
    HIPC(hipSetDevice(g->dev));
    const int D = want_der ? (want_der >= 2 ? 2 : 1) : 0;
    sweep_run(g, ks, nk, c_o, c_i, R, nc, q, D, sigma, D ? dsigma : nullptr, nullptr);
    API_END
}

int hbem_sigma_sweep_der2(hbem_geometry_t g, const double* ks, int nk, double c_o, double c_i, const void* R, int nc,
                          int q, double* sigma, double* dsigma, double* d2sigma) {
    API_BEGIN
    check_sweep_args(g, ks, nk, c_o, c_i, R, nc, q, sigma);
    if (!dsigma || !d2sigma) throw Error(HBEM_EINVAL, "null argument");
    std::lock_guard<std::mutex> lk(g->mu);
    HIPC(hipSetDevice(g->dev));
    sweep_run(g, ks, nk, c_o, c_i, R, nc, q, 2, sigma, dsigma, d2sigma);
    API_END
}

int hbem_sigma_sweep_mixed(hbem_geometry_t g, const double* ks, int nk, double c_o, double c_i, const void* R, int nc,
                           int q, int refine, int refine_inner, double* sigma, int* fp64_used) {
    API_BEGIN
    check_sweep_args(g, ks, nk, c_o, c_i, R, nc, q, sigma);
    if (refine < 0 || refine > 20 || refine_inner > 20)
        throw Error(HBEM_EINVAL, "number of refinement steps must be in [0, 20]");
    std::lock_guard<std::mutex> lk(g->mu);
    HIPC(hipSetDevice(g->dev));
    sweep_run(g, ks, nk, c_o, c_i, R, nc, q, 0, sigma, nullptr, nullptr, refine, fp64_used, refine_inner);
    API_END
}

}  // extern "C"

namespace {

// Work space of hbem_sigma_sweep_multi (one per process; calls are serialised),
// and the hipGraphs of its device phases, one per subset shape: a subset's
// device work (entry concatenation, assembly, T, LU, randomized SVD) is
// captured on first use and replayed by later calls with the same instances,
// sizes and buffers.  Everything a replay reads from the host side is
// uploaded before the launch into buffers the graph reads in place.
struct GraphEntry {
    std::vector<const void*> key;
    hipGraphExec_t exec = nullptr;
};
struct MultiWork {
    std::mutex mu;
    DBuf ent[3], flags[3], pidx[3], einst[3], igeo[3], ibase[3], kap, kr, cv, L, part[3], T, piv, perm, X, Y, work, sig,
        qi, desc, descT, roff, nrow, Rall;
    std::vector<GraphEntry> graphs;
    void drop_graphs() {
        for (auto& e : graphs) (void)hipGraphExecDestroy(e.exec);
        graphs.clear();
    }
};
MultiWork g_multi;

}  // namespace

// Graphs that name a geometry being destroyed (or released) are dropped: a
// later context may reuse its addresses.
void multi_forget(hbem_geometry* g) {
    std::lock_guard<std::mutex> lk(g_multi.mu);
    auto& v = g_multi.graphs;
    for (size_t i = 0; i < v.size();) {
        if (std::find(v[i].key.begin(), v[i].key.end(), (const void*)g) != v[i].key.end()) {
            (void)hipGraphExecDestroy(v[i].exec);
            v.erase(v.begin() + i);
        } else {
            ++i;
        }
    }
}

namespace {

bool graphs_enabled() {
    const char* e = std::getenv("HBEM_NO_GRAPH");
    return !(e && e[0] && e[0] != '0');
}

// One bucket subset of hbem_sigma_sweep_multi: instances `ids` (all padded to
// order n_pad), their nk wavenumbers each.  The panel pairs of every class of
// all these instances are assembled by one launch; T of instance i, k j is
// matrix i_local nk + j of one batch of order n_pad (identity in the
// padding), factorised and reduced to sigma by the batched dense stage.
void multi_subset(const std::vector<hbem_geometry*>& G, const std::vector<int>& ids, int n_pad, const double* ks, int nk,
                  double c_o, const double* c_i, const void* const* R, int nc, int q, double* sigma) {
    MultiWork& w = g_multi;
    hbem_geometry* g0 = G[ids[0]];
    const hipStream_t st = g0->st;
    const int m = (int)ids.size();
    const int nslot = 16, nks = 64 * ((nk + 31) / 32);   // tdiff slots per instance
    const size_t nn = (size_t)n_pad * n_pad;
    const int batch = m * nk;
    // ---- host side: every table the device phase reads, uploaded first ----
    // slots of instance i: its nk wavenumbers in the tdiff layout (blocks of
    // 32 with c_i(i), then c_o); the local buffer holds the T differences, nk
    // columns
    std::vector<double> kap, kr, cv;
    std::vector<long long> pair_off(m + 1, 0);
    for (int a = 0; a < m; ++a) {
        const int i = ids[a];
        std::vector<double> ka, kb, kc;
        tdiff_slots(ks + (size_t)i * nk, nk, c_i[i], c_o, ka, kb, kc);
        kap.insert(kap.end(), ka.begin(), ka.end());
        kr.insert(kr.end(), kb.begin(), kb.end());
        cv.insert(cv.end(), kc.begin(), kc.end());
        pair_off[a + 1] = pair_off[a] + G[i]->npairs;
    }
    const double* dkap = upload(w.kap, kap, st);
    const double* dkr = upload(w.kr, kr, st);
    const double* dcv = upload(w.cv, cv, st);
    struct ClassPlan {
        long long ne = 0;
        std::vector<long long> base;
        int nodes = 0, nchunk = 1;
    } cp[3];
    const double* const* digeo[3] = {nullptr, nullptr, nullptr};
    const long long* dibase[3] = {nullptr, nullptr, nullptr};
    for (int cls = 0; cls < 3; ++cls) {
        ClassPlan& c = cp[cls];
        c.base.assign(m + 1, 0);
        std::vector<const double*> geo(m);
        for (int a = 0; a < m; ++a) {
            hbem_geometry* g = G[ids[a]];
            const int ne = cls == 2 ? g->ngen : cls == 1 ? g->nadj : g->ncoi;
            c.base[a + 1] = c.base[a] + ne;
            geo[a] = (const double*)(cls == 2 ? g->d_gngeo.p : cls == 1 ? g->d_angeo.p : g->d_cngeo.p);
        }
        c.ne = c.base[m];
        c.nodes = cls == 2 ? g0->ns * g0->ns : cls == 1 ? 2 * g0->nc * g0->nc : g0->nc * g0->nc;
        c.nchunk = cls == 2 ? 1 : cls == 1 ? (int)adj_chunks(g0->nc) : (int)coi_chunks(g0->nc);
        digeo[cls] = upload(w.igeo[cls], geo, st);
        dibase[cls] = upload(w.ibase[cls], c.base, st);
    }
    std::vector<int> qi(nk);
    for (int j = 0; j < nk; ++j) qi[j] = j;
    const int* dqi = upload(w.qi, qi, st);
    std::vector<hbem::CsrDesc> desc(batch), descT(batch);
    std::vector<long long> roff(batch);
    std::vector<int> nrow(batch);
    std::vector<cdbl> Rall;
    for (int a = 0; a < m; ++a) {
        hbem_geometry* g = G[ids[a]];
        const int n = 2 * g->N;
        const long long off = (long long)Rall.size();
        const cdbl* Rh = (const cdbl*)R[ids[a]];
        Rall.insert(Rall.end(), Rh, Rh + (size_t)n * nc);
        for (int j = 0; j < nk; ++j) {
            const int b = a * nk + j;
            if (g->proj) {
                desc[b] = {(const int*)g->s_F.rp.p, (const int*)g->s_F.ci.p, (const double*)g->s_F.v.p, n};
                descT[b] = {(const int*)g->s_FT.rp.p, (const int*)g->s_FT.ci.p, (const double*)g->s_FT.v.p, n};
            }
            roff[b] = off;
            nrow[b] = n;
        }
    }
    const hbem::CsrDesc* dF = g0->proj ? upload(w.desc, desc, st) : nullptr;
    const hbem::CsrDesc* dFT = g0->proj ? upload(w.descT, descT, st) : nullptr;
    const cdbl* dRall = upload(w.Rall, Rall, st);
    const long long* droff = upload(w.roff, roff, st);
    const int* dnrow = upload(w.nrow, nrow, st);
    // ---- every device buffer at its final size before the device phase ----
    cdbl* L = w.L.as<cdbl>((size_t)pair_off[m] * nslot * nk);
    int* ent[3];
    unsigned char* fl[3];
    long long* pidx[3];
    int* einst[3];
    cdbl* part[3];
    for (int cls = 0; cls < 3; ++cls) {
        const size_t ne = (size_t)std::max<long long>(cp[cls].ne, 1);
        ent[cls] = w.ent[cls].as<int>(3 * ne);
        fl[cls] = w.flags[cls].as<unsigned char>(ne);
        pidx[cls] = w.pidx[cls].as<long long>(ne);
        einst[cls] = w.einst[cls].as<int>(ne);
        part[cls] = cp[cls].nchunk > 1 ? w.part[cls].as<cdbl>((size_t)nk * ne * cp[cls].nchunk * nslot) : nullptr;
    }
    cdbl* T = w.T.as<cdbl>(nn * batch);
    int* piv = w.piv.as<int>((size_t)n_pad * batch);
    int* perm = w.perm.as<int>((size_t)n_pad * batch);
    cdbl* X = w.X.as<cdbl>((size_t)n_pad * nc * batch);
    cdbl* Y = w.Y.as<cdbl>((size_t)n_pad * nc * batch);
    cdbl* work = w.work.as<cdbl>(2 * (size_t)n_pad * nc * batch);
    double* dsig = w.sig.as<double>(batch);
    HIPC(hipStreamSynchronize(st));   // uploads done; nothing below reads host memory
    // ---- device phase ----
    auto device_phase = [&]() {
        hbem::AsmParams P;
        std::memset(&P, 0, sizeof(P));
        P.xg = (const double*)g0->d_xg.p;
        P.wg = (const double*)g0->d_wg.p;
        P.xc = (const double*)g0->d_xc.p;
        P.wc = (const double*)g0->d_wc.p;
        P.nk = nks;
        P.kap = dkap;
        P.kr = dkr;
        P.cv = dcv;
        P.npairs_total = pair_off[m];
        P.ns = g0->ns;
        P.nc = g0->nc;
        P.qflags = g0->flags;
        P.tdiff = 1;
        P.nkt = nk;
        for (int cls = 2; cls >= 0; --cls) {
            const ClassPlan& c = cp[cls];
            if (c.ne == 0) continue;
            for (int a = 0; a < m; ++a) {
                hbem_geometry* g = G[ids[a]];
                const long long cnt = c.base[a + 1] - c.base[a];
                if (cnt == 0) continue;
                const DBuf& de = cls == 2 ? g->d_gent : cls == 1 ? g->d_aent : g->d_cent;
                const DBuf& df = cls == 2 ? g->d_gflags : cls == 1 ? g->d_aflags : g->d_cflags;
                const DBuf& dp = cls == 2 ? g->d_gpidx : cls == 1 ? g->d_apidx : g->d_cpidx;
                hbem::copy_bytes(ent[cls] + 3 * c.base[a], de.p, 3 * cnt * sizeof(int), st);
                hbem::copy_bytes(fl[cls] + c.base[a], df.p, (size_t)cnt, st);
                hbem::copy_bytes(pidx[cls] + c.base[a], dp.p, cnt * sizeof(long long), st);
                hbem::add_offset_ll(pidx[cls] + c.base[a], cnt, pair_off[a], st);
                hbem::fill_int(einst[cls] + c.base[a], (int)cnt, a, st);
            }
            P.igeo = digeo[cls];
            P.ibase = dibase[cls];
            P.einst = einst[cls];
            P.ent = ent[cls];
            P.flags = fl[cls];
            P.pidx = pidx[cls];
            P.nent = (int)c.ne;
            P.ngeo = nullptr;
            P.chunk = cls == 2 ? c.nodes : kChunkNodes;
            P.nchunk = c.nchunk;
            P.out = c.nchunk > 1 ? part[cls] : L;
            hbem::launch_assemble(cls, 0, P, 0, st);
            if (c.nchunk > 1)
                hbem::launch_chunk_sum(part[cls], (int)c.ne, c.nchunk, nk, nslot, pidx[cls], pair_off[m], L, st);
        }
        for (int a = 0; a < m; ++a) {
            hbem_geometry* g = G[ids[a]];
            const int n = 2 * g->N;
            hbem::launch_form_T(L + (size_t)pair_off[a] * nslot * nk, nk, nslot, dqi, nullptr, nk, g->N, 0,
                                (const double*)g->d_M.p, T + (size_t)a * nk * nn, nullptr, nullptr, st, n_pad);
            if (n < n_pad) hbem::launch_pad_identity(T + (size_t)a * nk * nn, n, n_pad, nk, st);
        }
        hbem::lu_factor_batched(T, n_pad, batch, piv, perm, st);
        DenseCtx c;
        c.st = st; c.n = n_pad; c.batch = batch; c.nc = nc; c.q = q; c.LU = T; c.ms = (long long)nn; c.perm = perm;
        c.F = nullptr;
        c.FT = nullptr;
        c.Fm = dF;
        c.FTm = dFT;
        c.X = X;
        c.Y = Y;
        c.work = work;
        hbem::replicate_padded(dRall, droff, dnrow, n_pad, nc, batch, c.X, st);
        rsvd_core(c, dsig, nullptr, nullptr);
    };
    if (graphs_enabled()) {
        std::vector<const void*> key;
        for (int a = 0; a < m; ++a) key.push_back(G[ids[a]]);
        key.push_back((const void*)(size_t)nk);
        key.push_back((const void*)(size_t)nc);
        key.push_back((const void*)(size_t)q);
        key.push_back((const void*)(size_t)n_pad);
        const void* bufs[] = {dkap, dkr, dcv, digeo[0], digeo[1], digeo[2], dibase[0], dibase[1], dibase[2], dqi,
                              dF, dFT, dRall, droff, dnrow, L, T, piv, perm, X, Y, work, dsig,
                              ent[0], ent[1], ent[2], fl[0], fl[1], fl[2], pidx[0], pidx[1], pidx[2], einst[0],
                              einst[1], einst[2], part[0], part[1], part[2]};
        key.insert(key.end(), std::begin(bufs), std::end(bufs));
        GraphEntry* ge = nullptr;
        for (auto& e : w.graphs)
            if (e.key == key) ge = &e;
        if (!ge) {
            hipGraph_t graph = nullptr;
            HIPC(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
            device_phase();
            const hipError_t ec = hipStreamEndCapture(st, &graph);
            HIPC(ec);
            hipGraphExec_t exec = nullptr;
            const hipError_t ei = hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0);
            (void)hipGraphDestroy(graph);
            HIPC(ei);
            if (w.graphs.size() >= 16) w.drop_graphs();
            w.graphs.push_back({key, exec});
            ge = &w.graphs.back();
        }
        HIPC(hipGraphLaunch(ge->exec, st));
    } else {
        device_phase();
    }
    HIPC(hipGetLastError());
    std::vector<double> out(batch);
    HIPC(hipMemcpyAsync(out.data(), dsig, batch * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPC(hipStreamSynchronize(st));
    for (int a = 0; a < m; ++a)
        for (int j = 0; j < nk; ++j) sigma[(size_t)ids[a] * nk + j] = out[(size_t)a * nk + j];
}

}  // namespace

extern "C" {

int hbem_sigma_sweep_multi(const hbem_geometry_t* geoms, int ninst, const double* ks, int nk, double c_o,
                           const double* c_i, const void* const* R, int nc, int q, int bucket, double* sigma) {
    API_BEGIN
    if (!geoms || !ks || !c_i || !R || !sigma) throw Error(HBEM_EINVAL, "null argument");
    if (ninst < 1 || nk < 1) throw Error(HBEM_EINVAL, "need at least one instance and one wavenumber");
    if (bucket < 0 || bucket % 32 != 0 || bucket > 4096) throw Error(HBEM_EINVAL, "bucket must be a multiple of 32 in [0, 4096]");
    const int gran = bucket == 0 ? 256 : bucket;
    std::vector<hbem_geometry*> G(geoms, geoms + ninst);
    for (int i = 0; i < ninst; ++i) {
        if (!G[i] || !R[i]) throw Error(HBEM_EINVAL, "null geometry or test matrix");
        if (G[i]->ns != G[0]->ns || G[i]->proj != G[0]->proj || G[i]->flags != G[0]->flags || G[i]->dev != G[0]->dev)
            throw Error(HBEM_EINVAL, "instances must share quadrature order, projection, flags and device");
        if (G[i]->flags & HBEM_FLAG_FAR_FIELD)
            throw Error(HBEM_EINVAL, "the multi-instance sweep takes contexts without HBEM_FLAG_FAR_FIELD");
        check_dense_args(ks, 2 * G[i]->N, R[i], nc, q);
        for (int j = 0; j < nk; ++j) { check_k(ks[(size_t)i * nk + j], 0.0, c_i[i]); check_k(ks[(size_t)i * nk + j], 0.0, c_o); }
    }
    // every instance's context is locked for the call (in address order), then the work space
    std::vector<hbem_geometry*> order(G);
    std::sort(order.begin(), order.end());
    order.erase(std::unique(order.begin(), order.end()), order.end());
    std::vector<std::unique_lock<std::mutex>> locks;
    for (auto* g : order) locks.emplace_back(g->mu);
    std::lock_guard<std::mutex> lk(g_multi.mu);
    HIPC(hipSetDevice(G[0]->dev));
    // buckets of equal padded order, in increasing order; a bucket is split
    // into subsets whose work space fits the free device memory
    std::vector<std::pair<int, int>> key(ninst);
    for (int i = 0; i < ninst; ++i) key[i] = {(2 * G[i]->N + gran - 1) / gran * gran, i};
    std::sort(key.begin(), key.end());
    size_t held = 0;
    {
        MultiWork& w = g_multi;
        const DBuf* bs[] = {&w.L, &w.part[0], &w.part[1], &w.part[2], &w.T, &w.piv, &w.perm, &w.X, &w.Y, &w.work};
        for (auto* b : bs) held += b->n;
    }
    const size_t avail = usable_device_bytes(held);
    size_t a = 0;
    while (a < key.size()) {
        const int n_pad = key[a].first;
        std::vector<int> ids;
        size_t used = 0;
        while (a < key.size() && key[a].first == n_pad) {
            hbem_geometry* g = G[key[a].second];
            const size_t part = (size_t)std::max(g->nadj * adj_chunks(g->nc), g->ncoi * coi_chunks(g->nc));
            const size_t need = (size_t)nk * ((size_t)g->npairs + part) * 16 * sizeof(cdbl) +
                                (size_t)nk * ((size_t)n_pad * n_pad * sizeof(cdbl) + (size_t)n_pad * (8 + (4 * nc + 4) * sizeof(cdbl)));
            if (!ids.empty() && used + need > avail) break;
            ids.push_back(key[a].second);
            used += need;
            ++a;
        }
        multi_subset(G, ids, n_pad, ks, nk, c_o, c_i, R, nc, q, sigma);
    }
    API_END
}

int hbem_profile_enable(hbem_geometry_t g, int enable) {
    API_BEGIN
    if (!g) throw Error(HBEM_EINVAL, "null geometry");
    std::lock_guard<std::mutex> lk(g->mu);
    g->plog.reset();
    g->prof = enable != 0;
    API_END
}

int hbem_profile_query(hbem_geometry_t g, int which, double* ms, long long* launches, double* flops) {
    API_BEGIN
    if (!g || which < 0 || which > kProfGemm) throw Error(HBEM_EINVAL, "invalid argument");
    std::lock_guard<std::mutex> lk(g->mu);
    g->plog.resolve();
    if (ms) *ms = g->plog.ms[which];
    if (launches) *launches = g->plog.n[which];
    if (flops) *flops = g->plog.flops[which];
    API_END
}

int hbem_sweep_batch_size(int npanels, int order, int d, int nc, unsigned long long avail_bytes, int nk, int* B) {
    API_BEGIN
    if (!B) throw Error(HBEM_EINVAL, "null output");
    if (npanels < 3 || order < 1 || order > 40 || d < 0 || d > 2 || nc < 1 || nk < 0)
        throw Error(HBEM_EINVAL, "invalid argument");
    // a closed curve of N panels: N coinciding, N adjacent pairs
    const long long N = npanels;
    const int ncr = order * (order + 1) / 2 - 1;
    *B = sweep_batch(sweep_bytes_per_k(N * (N + 1) / 2, N, N, ncr, 2 * npanels, d, nc, nk >= 32 ? 1 : 2),
                     (size_t)avail_bytes, nk);
    API_END
}

int hbem_rsvd_benchmarking(int yes) {
    API_BEGIN
    std::lock_guard<std::mutex> lk(g_sa.mu);
    g_sa.bench = yes != 0;
    API_END
}

int hbem_rsvd_timers(double* seconds4) {
    API_BEGIN
    if (!seconds4) throw Error(HBEM_EINVAL, "null output");
    std::lock_guard<std::mutex> lk(g_sa.mu);
    g_sa.tim.resolve();
    for (int i = 0; i < 4; ++i) seconds4[i] = g_sa.tim.ms[i] * 1e-3;
    API_END
}

int hbem_rsvd_reset_timer(void) {
    API_BEGIN
    std::lock_guard<std::mutex> lk(g_sa.mu);
    g_sa.tim.reset();
    API_END
}

}  // extern "C"
