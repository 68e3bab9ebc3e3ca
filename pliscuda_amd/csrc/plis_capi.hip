// plis_capi.hip -- host side of libplis_hip.so: the C ABI declared in
// include/plis.h.  Builds the canonical device layout of a batch of
// realizations, launches the two step kernels (optionally replayed as a
// hipGraph), and keeps every control decision on the device: the host only
// polls a done-counter once per batch of iterations (the reference makes at
// least seven blocking host<->device transactions per step,
// PLISCuda/src/driver.cu:259-376).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/plis.h"
#include "plis_learned.hip"

using namespace plis;

static thread_local std::string g_create_err;

// Scratch of plis_mlp_forward_noalloc: device staging for the feature rows
// and the outputs of one call, [rows][12] then [rows][2] floats.  One array
// per GPU, held until the process ends; every call on a GPU shares it.  In
// host code `scratch` names a host-side stand-in: kernels and copies use the
// device address taken by hipGetSymbolAddress in plis_create instead.
constexpr uint32_t kNoAllocRows = 16384;
constexpr size_t kNoAllocFloats = (size_t)kNoAllocRows * 14;
static __device__ __align__(256) float scratch[kNoAllocFloats];
constexpr int kMaxDevices = 64;
static float* g_scratch_ptr[kMaxDevices] = {};
static std::mutex g_scratch_mu;

// Device address of this device's scratch, taken once per device with the
// device current (plis_create calls it).  nullptr when the lookup fails.
static float* scratch_ptr(int device, std::string& err) {
    if (device < 0 || device >= kMaxDevices) {
        err = "device index out of range for the no-allocation scratch";
        return nullptr;
    }
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
            return nullptr;
        }
        g_scratch_ptr[device] = static_cast<float*>(p);
    }
    return g_scratch_ptr[device];
}

constexpr int kGraphSizes = 7;

struct plis_ctx {
    hipStream_t stream = nullptr;
    int device = 0;
    int cus = 0;
    std::string err;
    NetHeader nh{};
    int variant = 0;  // index into kVariantS / kVariantR
    int32_t S = 0, R = 0, C = 1;
    uint32_t nreal = 0, ntiles = 0;
    uint64_t total_counts = 0, total_sub = 0;
    uint64_t model_hash = 0;  // network tables, every realization's CSR and compartments
    std::vector<RealInfo> info;
    std::vector<RealCtl> ctl_host;
    // device buffers
    int32_t* d_net_i = nullptr;
    double* d_net_d = nullptr;
    RealInfo* d_info = nullptr;
    RealCtl* d_ctl = nullptr;
    TileEnt* d_tiles = nullptr;
    uint32_t *d_x0 = nullptr, *d_x1 = nullptr;
    uint32_t *d_rp = nullptr, *d_col = nullptr, *d_irp = nullptr, *d_icol = nullptr;
    double* d_wsum = nullptr;  // per subvolume, row_ptr layout
    uint32_t* d_frp = nullptr;  // forward CSR (symmetric realizations): row pointers
    uint64_t* d_fcw = nullptr;  // its entries: neighbour id, row's index in its tile, weight
    unsigned long long* d_gdelta = nullptr;        // leap balances, counts layout
    float *d_w = nullptr, *d_iw = nullptr;
    uint64_t *d_cw = nullptr, *d_icw = nullptr;  // interleaved id + weight
    int32_t* d_comp = nullptr;
    PartA* d_pa_t = nullptr;
    PartA* d_pa_c = nullptr;      // decide_kernel's chunk partials
    PartB* d_pb_c = nullptr;      // commit_decide_kernel's chunk partials of the leap sums
    uint32_t decide_chunks = 1;   // decide_kernel's blocks per realization
    uint32_t* d_dcnt = nullptr;   // decide_kernel's per-realization chunk counters
    PartB* d_pb_t = nullptr;
    uint32_t* d_cnt = nullptr;   // [0] realizations done, [1] redo-list length, [2..3] redone tiles (u64)
    uint32_t* d_redo = nullptr;  // redo list of the leap pass (tile ids)
    uint32_t* h_ndone = nullptr;  // pinned
    double eps = 0.03;
    int32_t nc = 10;
    uint64_t seed = 0;
    uint32_t wpb = 4, blocks = 8;
    int use_graph = 1;
    bool edge_parallel = false;  // every realization symmetric, tiles < 2^27 forward entries
    hipEvent_t* split_ev = nullptr;  // profiling: events between the leap pass's kernels
    // captured step sequences of 1, 2, 4, ..., 64 iterations (index log2 n)
    hipGraphExec_t graphs[kGraphSizes] = {};
    std::vector<void*> allocs;
    // post-leap leap condition (theta = 0: not evaluated) and learned tau
    double theta = 0.0, viol_rate_max = 1.0, kappa = 1.0;
    int learned_on = 0;
    int mlp_act = 1;  // 0 ReLU, 1 GELU
    // blocks per CU of the learned tau pass and of the learned leap-pass end
    uint32_t learned_bpc = 1, finish_bpc = 4;
    float *d_w32 = nullptr, *d_m = nullptr, *d_v = nullptr, *d_norm = nullptr, *d_grad = nullptr;
    __bf16* d_q = nullptr;
    float* d_partial = nullptr;
    uint32_t train_blocks = 0;
    int32_t adam_t = 0;
    float *d_rf = nullptr, *d_ry = nullptr, *d_rc = nullptr;  // training records on the device
    uint64_t rec_cap = 0;
};

#define HIPCHK(ctx, call)                                                              \
    do {                                                                               \
        hipError_t e_ = (call);                                                        \
        if (e_ != hipSuccess) {                                                        \
            (ctx)->err = std::string(#call) + ": " + hipGetErrorString(e_);            \
            return PLIS_ERR_HIP;                                                       \
        }                                                                              \
    } while (0)

template <typename T>
static int dev_alloc(plis_ctx* c, T** p, size_t n) {
    size_t bytes = std::max<size_t>(n, 1) * sizeof(T);
    HIPCHK(c, hipMalloc((void**)p, bytes));
    c->allocs.push_back(*p);
    return 0;
}

template <typename T>
static int upload(plis_ctx* c, T** p, const std::vector<T>& v) {
    if (int rc = dev_alloc(c, p, v.size())) return rc;
    if (!v.empty()) HIPCHK(c, hipMemcpy(*p, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice));
    return 0;
}

static int fail_create(const std::string& m) {
    g_create_err = m;
    return PLIS_ERR_ARG;
}

// 64-bit FNV-1a over the bytes of the model a checkpoint belongs to.
static void hash_bytes(uint64_t& h, const void* p, size_t n) {
    const unsigned char* b = static_cast<const unsigned char*>(p);
    for (size_t i = 0; i < n; ++i) {
        h ^= b[i];
        h *= 0x100000001B3ull;
    }
}
template <typename T>
static void hash_vec(uint64_t& h, const std::vector<T>& v) {
    const uint64_t n = v.size();
    hash_bytes(h, &n, sizeof n);
    if (!v.empty()) hash_bytes(h, v.data(), v.size() * sizeof(T));
}

// ---------------------------------------------------------------------------
// canonical CSR: self loops dropped, neighbours ascending, duplicate
// neighbours merged (weights summed in double in input order, then rounded)
// ---------------------------------------------------------------------------
struct Csr {
    std::vector<uint32_t> rp, col;
    std::vector<float> w;
};

static bool canonical_csr(const plis_topology& t, Csr& out, std::string& err) {
    const uint32_t N = t.n_subvolumes;
    out.rp.assign(N + 1, 0);
    out.col.clear();
    out.w.clear();
    out.col.reserve(t.n_edges);
    out.w.reserve(t.n_edges);
    struct Nb { uint32_t id, pos; float w; };
    std::vector<Nb> tmp;
    for (uint32_t v = 0; v < N; ++v) {
        const uint32_t b = t.row_ptr[v], e = t.row_ptr[v + 1];
        if (e < b || e > t.n_edges) { err = "row_ptr not monotone or beyond n_edges"; return false; }
        tmp.clear();
        for (uint32_t q = b; q < e; ++q) {
            const uint32_t u = t.col[q];
            if (u >= N) { err = "neighbour id >= n_subvolumes"; return false; }
            if (u == v) continue;
            const float wq = t.weight ? t.weight[q] : 1.0f;
            if (!(wq > 0.0f) || !std::isfinite(wq)) { err = "edge weight must be positive and finite"; return false; }
            tmp.push_back({u, q, wq});
        }
        std::sort(tmp.begin(), tmp.end(), [](const Nb& x, const Nb& y) {
            return x.id != y.id ? x.id < y.id : x.pos < y.pos;
        });
        out.rp[v] = (uint32_t)out.col.size();
        for (size_t i = 0; i < tmp.size();) {
            size_t i2 = i;
            double acc = 0.0;
            while (i2 < tmp.size() && tmp[i2].id == tmp[i].id) acc += (double)tmp[i2++].w;
            out.col.push_back(tmp[i].id);
            out.w.push_back(i2 - i == 1 ? tmp[i].w : (float)acc);
            i = i2;
        }
    }
    out.rp[N] = (uint32_t)out.col.size();
    return true;
}

static bool is_symmetric(const Csr& g, uint32_t N) {
    for (uint32_t v = 0; v < N; ++v)
        for (uint32_t q = g.rp[v]; q < g.rp[v + 1]; ++q) {
            const uint32_t u = g.col[q];
            auto b = g.col.begin() + g.rp[u], e = g.col.begin() + g.rp[u + 1];
            auto it = std::lower_bound(b, e, v);
            if (it == e || *it != v) return false;
            if (g.w[it - g.col.begin()] != g.w[q]) return false;
        }
    return true;
}

static void transpose(const Csr& g, uint32_t N, Csr& in) {
    in.rp.assign(N + 1, 0);
    for (uint32_t q = 0; q < g.rp[N]; ++q) in.rp[g.col[q] + 1]++;
    for (uint32_t v = 0; v < N; ++v) in.rp[v + 1] += in.rp[v];
    in.col.assign(g.rp[N], 0);
    in.w.assign(g.rp[N], 0.0f);
    std::vector<uint32_t> fill(N, 0);
    for (uint32_t u = 0; u < N; ++u)
        for (uint32_t q = g.rp[u]; q < g.rp[u + 1]; ++q) {
            const uint32_t v = g.col[q];
            const uint32_t pos = in.rp[v] + fill[v]++;
            in.col[pos] = u;
            in.w[pos] = g.w[q];
        }
}

// ---------------------------------------------------------------------------
// network blob
// ---------------------------------------------------------------------------
static bool build_network(const plis_network& net, NetHeader& h, std::vector<int32_t>& I,
                          std::vector<double>& D, std::string& err) {
    const int S = net.n_species, R = net.n_reactions, C = net.n_compartments;
    if (S < 1 || S > kMaxS) { err = "n_species must be in [1, 64]"; return false; }
    if (R < 0 || R > kMaxR) { err = "n_reactions must be in [0, 128]"; return false; }
    if (C < 1) { err = "n_compartments must be >= 1"; return false; }
    if (!net.k || !net.d || (R > 0 && (!net.reactants || !net.products))) { err = "null network array"; return false; }
    h = NetHeader{};
    h.S = S; h.R = R; h.C = C;
    I.clear();
    D.clear();
    auto put = [&](int32_t& off, size_t n) { off = (int32_t)I.size(); I.resize(I.size() + std::max<size_t>(n, 1), 0); };
    put(h.off_nu, (size_t)S * R);
    put(h.off_rs, (size_t)R * 3);
    put(h.off_rr, (size_t)R * 3);
    put(h.off_nrs, R);
    put(h.off_order, R);
    put(h.off_hor, S);
    put(h.off_mult, S);
    for (int j = 0; j < R; ++j) {
        int ord = 0, n = 0;
        for (int s = 0; s < S; ++s) {
            const int r = net.reactants[s * R + j], p = net.products[s * R + j];
            if (r < 0 || p < 0) { err = "negative stoichiometric coefficient"; return false; }
            I[h.off_nu + s * R + j] = p - r;
            if (r > 0) {
                if (n >= 3) { err = "reaction order above 3"; return false; }
                I[h.off_rs + j * 3 + n] = s;
                I[h.off_rr + j * 3 + n] = r;
                ++n;
                ord += r;
            }
        }
        if (ord > 3) { err = "reaction order above 3"; return false; }
        I[h.off_nrs + j] = n;
        I[h.off_order + j] = ord;
    }
    for (int s = 0; s < S; ++s) {
        int hor = 0, mult = 0;
        for (int j = 0; j < R; ++j) {
            const int r = net.reactants[s * R + j];
            if (r <= 0) continue;
            const int o = I[h.off_order + j];
            if (o > hor) { hor = o; mult = r; }
            else if (o == hor && r > mult) mult = r;
        }
        bool df = false;
        for (int c = 0; c < C; ++c) df = df || net.d[c * S + s] > 0.0;
        if (df && hor == 0) { hor = 1; mult = 1; }
        I[h.off_hor + s] = hor;
        I[h.off_mult + s] = mult;
    }
    h.ints_total = (int32_t)I.size();
    h.off_k = 0;
    D.assign(net.k, net.k + (size_t)C * R);
    h.off_d = (int32_t)D.size();
    D.insert(D.end(), net.d, net.d + (size_t)C * S);
    for (double v : D)
        if (!(v >= 0.0) || !std::isfinite(v)) { err = "rate constants must be finite and >= 0"; return false; }
    if (D.empty()) D.push_back(0.0);
    h.dbl_total = (int32_t)D.size();
    return true;
}

static KernelArgs make_args(plis_ctx* c) {
    KernelArgs a{};
    a.net_i = c->d_net_i;
    a.net_d = c->d_net_d;
    a.nh = c->nh;
    a.info = c->d_info;
    a.ctl = c->d_ctl;
    a.tiles = c->d_tiles;
    a.ntiles = c->ntiles;
    a.nreal = c->nreal;
    a.x0 = c->d_x0;
    a.x1 = c->d_x1;
    a.rp = c->d_rp;
    a.wsum = c->d_wsum;
    a.col = c->d_col;
    a.w = c->d_w;
    a.irp = c->d_irp;
    a.icol = c->d_icol;
    a.iw = c->d_iw;
    a.cw = c->d_cw;
    a.icw = c->d_icw;
    a.fcw = c->d_fcw;
    a.comp = c->d_comp;
    a.partA_tile = c->d_pa_t;
    a.partA_chunk = c->d_pa_c;
    a.partB_chunk = c->d_pb_c;
    a.decide_cnt = c->d_dcnt;
    a.decide_chunks = c->decide_chunks;
    a.partB_tile = c->d_pb_t;
    a.ndone = c->d_cnt;
    a.redo_count = c->d_cnt + 1;
    a.redo_tiles = c->d_redo;
    a.frp = c->d_frp;

    a.gdelta = c->d_gdelta;
    a.redo_total = reinterpret_cast<unsigned long long*>(c->d_cnt + 2);
    a.epsilon = c->eps;
    a.n_critical = c->nc;
    a.seed_lo = (uint32_t)c->seed;
    a.seed_hi = (uint32_t)(c->seed >> 32);
    a.waves_per_block = c->wpb;
    a.theta = c->theta;
    a.viol_rate_max = c->viol_rate_max;
    a.kappa = c->kappa;
    a.learned_on = c->learned_on;
    a.judge_on = c->theta > 0.0 ? 1 : 0;
    a.props_in_finish = (c->learned_on && c->edge_parallel) ? 1 : 0;
    a.mlp_act = c->mlp_act;
    a.mlp_w32 = c->d_w32;
    a.mlp_q = c->d_q;
    a.mlp_norm = c->d_norm;
    return a;
}

static const int kVariantS[] = {1, 3, 5, 8, 64};
static const int kVariantR[] = {4, 4, 8, 16, 128};

// Kernel groups: 0 tau pass (learned and/or Cao), 1 decide, 2 leap pass,
// 3 commit; 4 tau pass on the counts the leap pass wrote and 6 commit +
// decide (the folded order of plain tau); 5 = training records of one
// realization.  In a learned
// context the end of the edge-parallel leap pass also proposes the next
// step's tau (learned_finish_kernel); learned_tau_kernel covers the steps
// whose proposals it could not make.
template <int MS, int MR>
static void launch_variant(plis_ctx* c, int which, const KernelArgs& a) {
    const dim3 grid(c->blocks), block(64 * c->wpb), rgrid(c->nreal), rblock(1024), dblock(decide_threads<MS>());
    const dim3 dgrid(c->nreal * c->decide_chunks);
    const bool mc = c->nh.C > 1;  // several compartments: per-subvolume rate tables
    const bool gelu = c->mlp_act == 1;
    // learned kernels stage the network in LDS once per block: fewer, longer-lived blocks
    // the redo list is usually empty or short: two blocks per CU walk it
    const dim3 rgrid2(std::max<uint32_t>(8u, std::min<uint32_t>(c->blocks, ((uint32_t)c->cus * 2u + 7u) & ~7u)));
    const dim3 fgrid(std::max<uint32_t>(8u, std::min<uint32_t>(c->blocks, ((uint32_t)c->cus * c->finish_bpc + 7u) & ~7u)));
    const dim3 lgrid(std::max<uint32_t>(8u, std::min<uint32_t>(c->blocks, ((uint32_t)c->cus * c->learned_bpc + 7u) & ~7u)));
    switch (which) {
        case 0:
            if constexpr (MS <= 8) {
                if (c->learned_on) {
                    // one pass covers fallbacks and refreshes; in steady state
                    // it skips every tile, so it runs on a small grid
                    if (mc && gelu) hipLaunchKernelGGL((learned_tau_kernel<MS, MR, true, 1>), lgrid, block, 0, c->stream, a);
                    else if (mc) hipLaunchKernelGGL((learned_tau_kernel<MS, MR, true, 0>), lgrid, block, 0, c->stream, a);
                    else if (gelu) hipLaunchKernelGGL((learned_tau_kernel<MS, MR, false, 1>), lgrid, block, 0, c->stream, a);
                    else hipLaunchKernelGGL((learned_tau_kernel<MS, MR, false, 0>), lgrid, block, 0, c->stream, a);
                    break;
                }
            }
            if (mc) hipLaunchKernelGGL((phase_a_kernel<MS, MR, true>), grid, block, 0, c->stream, a);
            else hipLaunchKernelGGL((phase_a_kernel<MS, MR, false>), grid, block, 0, c->stream, a);
            break;
        case 1:
            hipLaunchKernelGGL((decide_kernel<MS, MR>), dgrid, dblock, 0, c->stream, a);
            break;
        case 2:
            if constexpr (MS <= 8) {
                if (c->edge_parallel) {
                    if (mc) {
                        hipLaunchKernelGGL((phase_b_kernel<MS, MR, true, true>), grid, block, 0, c->stream, a);
                        if (c->split_ev) (void)hipEventRecord(c->split_ev[0], c->stream);
                        hipLaunchKernelGGL((phase_b_redo_kernel<MS, MR, true>), rgrid2, block, 0, c->stream, a);
                        if (c->split_ev) (void)hipEventRecord(c->split_ev[1], c->stream);
                        if (!c->learned_on)
                            hipLaunchKernelGGL((phase_b_finish_kernel<MS, MR, true>), grid, block, 0, c->stream, a);
                        else if (gelu)
                            hipLaunchKernelGGL((learned_finish_kernel<MS, MR, true, 1>), fgrid, block, 0, c->stream, a);
                        else
                            hipLaunchKernelGGL((learned_finish_kernel<MS, MR, true, 0>), fgrid, block, 0, c->stream, a);
                    } else {
                        hipLaunchKernelGGL((phase_b_kernel<MS, MR, true, false>), grid, block, 0, c->stream, a);
                        if (c->split_ev) (void)hipEventRecord(c->split_ev[0], c->stream);
                        hipLaunchKernelGGL((phase_b_redo_kernel<MS, MR, false>), rgrid2, block, 0, c->stream, a);
                        if (c->split_ev) (void)hipEventRecord(c->split_ev[1], c->stream);
                        if (!c->learned_on)
                            hipLaunchKernelGGL((phase_b_finish_kernel<MS, MR, false>), grid, block, 0, c->stream, a);
                        else if (gelu)
                            hipLaunchKernelGGL((learned_finish_kernel<MS, MR, false, 1>), fgrid, block, 0, c->stream, a);
                        else
                            hipLaunchKernelGGL((learned_finish_kernel<MS, MR, false, 0>), fgrid, block, 0, c->stream, a);
                    }
                    break;
                }
            }
            if (mc) hipLaunchKernelGGL((phase_b_kernel<MS, MR, false, true>), grid, block, 0, c->stream, a);
            else hipLaunchKernelGGL((phase_b_kernel<MS, MR, false, false>), grid, block, 0, c->stream, a);
            break;
        case 3:
            hipLaunchKernelGGL(commit_kernel, rgrid, rblock, 0, c->stream, a);
            break;
        case 4: {  // folded order: the tau pass on the counts the leap pass wrote
            KernelArgs as = a;
            as.tau_spec = 1u;
            if (mc) hipLaunchKernelGGL((phase_a_kernel<MS, MR, true>), grid, block, 0, c->stream, as);
            else hipLaunchKernelGGL((phase_a_kernel<MS, MR, false>), grid, block, 0, c->stream, as);
            break;
        }
        case 6:  // folded order: commit and decide
            hipLaunchKernelGGL((commit_decide_kernel<MS, MR>), dgrid, dblock, 0, c->stream, a);
            break;
        case 5:
            if constexpr (MS <= 8) hipLaunchKernelGGL((records_kernel<MS, MR>), grid, block, 0, c->stream, a);
            break;
        default:
            break;
    }
}

static void launch_kernel(plis_ctx* c, int which, const KernelArgs& a) {
    switch (c->variant) {
        case 0: launch_variant<1, 4>(c, which, a); break;
        case 1: launch_variant<3, 4>(c, which, a); break;
        case 2: launch_variant<5, 8>(c, which, a); break;
        case 3: launch_variant<8, 16>(c, which, a); break;
        default: launch_variant<64, 128>(c, which, a); break;
    }
}

// The kernels of one iteration.  Plain tau (the folded order): the leap
// pass, the tau pass on the counts it wrote (used if the step is accepted),
// then one kernel that commits the step or leaps back and decides the next
// step.  A learned context: tau pass, decide, leap pass, commit.
static bool folded(const plis_ctx* c) { return !c->learned_on; }

static const int kFoldedStep[3] = {2, 4, 6};

static void launch_step(plis_ctx* c, const KernelArgs& a) {
    if (folded(c))
        for (int k : kFoldedStep) launch_kernel(c, k, a);
    else
        for (int k = 0; k < 4; ++k) launch_kernel(c, k, a);
}

// Before the first iteration of a call in the folded order: the tau pass on
// the current counts and the decision, which each folded iteration makes
// for the next one at its end.  Recomputing a decision that is already in
// place gives the same values (every input is the same), so a call may
// always start with it.
static void launch_prologue(plis_ctx* c, const KernelArgs& a) {
    if (!folded(c)) return;
    launch_kernel(c, 0, a);
    launch_kernel(c, 1, a);
}

// The captured graphs hold the kernel arguments and launch shape by value:
// every change of either drops them.
static void drop_graphs(plis_ctx* c) {
    for (auto& g : c->graphs)
        if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
}

static void set_shape(plis_ctx* c, uint32_t wpb, uint32_t bpc) {
    c->wpb = std::max(1u, std::min(4u, wpb));
    const uint32_t need = (c->ntiles + c->wpb - 1) / c->wpb;
    uint32_t blocks = std::min<uint32_t>((uint32_t)c->cus * std::max(1u, bpc), need);
    blocks = std::max(8u, (blocks + 7u) & ~7u);
    c->blocks = blocks;
    drop_graphs(c);
}

static int launch_iters(plis_ctx* c, uint32_t n) {
    const KernelArgs a = make_args(c);
    for (uint32_t i = 0; i < n; ++i) launch_step(c, a);
    HIPCHK(c, hipGetLastError());
    return 0;
}

static int prologue(plis_ctx* c) {
    launch_prologue(c, make_args(c));
    HIPCHK(c, hipGetLastError());
    return 0;
}

// n iterations as replays of captured graphs of 64, 32, ..., 1 iterations
// (binary decomposition of n).  The first replay captures every size at once,
// so a later call of another length does not capture inside a timed region.
static int graph_iters(plis_ctx* c, uint32_t n) {
    if (!c->use_graph) return launch_iters(c, n);
    if (!c->graphs[0]) {
        const KernelArgs a = make_args(c);
        for (int k = 0; k < kGraphSizes; ++k) {
            hipGraph_t g;
            HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
            for (uint32_t i = 0; i < (1u << k); ++i) launch_step(c, a);
            HIPCHK(c, hipStreamEndCapture(c->stream, &g));
            const hipError_t e = hipGraphInstantiate(&c->graphs[k], g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (e != hipSuccess) {
                c->graphs[k] = nullptr;
                drop_graphs(c);
                c->err = std::string("hipGraphInstantiate: ") + hipGetErrorString(e);
                return PLIS_ERR_HIP;
            }
        }
    }
    while (n > 0) {
        int k = kGraphSizes - 1;
        while ((1u << k) > n) --k;
        HIPCHK(c, hipGraphLaunch(c->graphs[k], c->stream));
        n -= 1u << k;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
extern "C" int plis_device_count(char* name, int32_t name_len) {
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    if (name && name_len > 0) {
        name[0] = 0;
        hipDeviceProp_t p;
        if (n > 0 && hipGetDeviceProperties(&p, 0) == hipSuccess) {
            std::snprintf(name, (size_t)name_len, "%s (%s)", p.name, p.gcnArchName);
        }
    }
    return n;
}

extern "C" const char* plis_last_error(const plis_ctx* ctx) {
    return ctx ? ctx->err.c_str() : g_create_err.c_str();
}

extern "C" void plis_destroy(plis_ctx* c) {
    if (!c) return;
    // teardown: release everything even if an earlier call failed
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_graphs(c);
    for (void* p : c->allocs) (void)hipFree(p);
    if (c->h_ndone) (void)hipHostFree(c->h_ndone);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

extern "C" int plis_create(const plis_network* net, int32_t n_real, const plis_topology* topos,
                           const uint32_t* const* x0, const uint32_t* ids, const plis_params* params,
                           plis_ctx** out) {
    if (!out) return fail_create("out is NULL");
    *out = nullptr;
    if (!net || !topos || !x0 || !params || n_real < 1) return fail_create("null argument or n_real < 1");
    if (!(params->epsilon > 0.0) || params->n_critical < 0) return fail_create("epsilon must be > 0, n_critical >= 0");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        g_create_err = "no HIP device available";
        return PLIS_ERR_NODEV;
    }
    NetHeader nh;
    std::vector<int32_t> I;
    std::vector<double> D;
    std::string err;
    if (!build_network(*net, nh, I, D, err)) return fail_create(err);
    const int S = nh.S;

    plis_ctx* c = new plis_ctx();
    c->nh = nh;
    c->S = nh.S;
    c->R = nh.R;
    c->C = nh.C;
    c->model_hash = 0xCBF29CE484222325ull;
    hash_vec(c->model_hash, I);
    hash_vec(c->model_hash, D);
    c->variant = 4;
    for (int v = 3; v >= 0; --v)
        if (nh.S <= kVariantS[v] && nh.R <= kVariantR[v]) c->variant = v;
    c->eps = params->epsilon;
    c->nc = params->n_critical;
    c->seed = params->seed;
    c->use_graph = params->use_graph;
    c->nreal = (uint32_t)n_real;

    // host-side layout
    std::vector<uint32_t> hx, hrp, hcol, hirp, hicol, hfrp;
    std::vector<float> hw, hiw;
    std::vector<uint64_t> hcw, hicw, hfcw;  // interleaved id + weight bits
    auto idw = [](uint32_t u, float w) {
        uint32_t b;
        std::memcpy(&b, &w, 4);
        return (uint64_t)u | ((uint64_t)b << 32);
    };
    std::vector<double> hwsum;
    std::vector<int32_t> hcomp;
    std::vector<TileEnt> tiles;
    bool any_weighted = false, any_asym = false, huge_tile = false;
    uint64_t xoff = 0, eoff = 0, rpoff = 0, ieoff = 0, irpoff = 0, coff = 0;
    uint32_t tile = 0;
    for (int32_t r = 0; r < n_real; ++r) {
        const plis_topology& t = topos[r];
        const uint32_t N = t.n_subvolumes;
        const uint32_t id = ids ? ids[r] : (uint32_t)r;
        if (N < 1 || !t.row_ptr || (t.n_edges && !t.col) || !x0[r]) {
            delete c;
            return fail_create("realization " + std::to_string(r) + ": empty topology or null array");
        }
        if (id > 0xFFFFu) { delete c; return fail_create("realization id must be < 65536"); }
        Csr g;
        if (!canonical_csr(t, g, err)) { delete c; return fail_create("realization " + std::to_string(r) + ": " + err); }
        bool weighted = false;
        for (float wv : g.w) weighted = weighted || wv != 1.0f;
        const bool sym = is_symmetric(g, N);
        hash_vec(c->model_hash, g.rp);
        hash_vec(c->model_hash, g.col);
        hash_vec(c->model_hash, g.w);
        RealInfo ri{};
        ri.x_off = xoff;
        ri.e_off = eoff;
        ri.rp_off = rpoff;
        ri.N = N;
        ri.id = id;
        ri.flags = (sym ? 1u : 0u) | (weighted ? 2u : 0u);
        any_weighted = any_weighted || weighted;
        hrp.insert(hrp.end(), g.rp.begin(), g.rp.end());
        hcol.insert(hcol.end(), g.col.begin(), g.col.end());
        // forward CSR: each undirected edge once, in the row of its lower end
        ri.fe_off = hfcw.size();
        for (uint32_t v = 0; v < N; ++v) {
            hfrp.push_back((uint32_t)(hfcw.size() - ri.fe_off));
            for (uint32_t q = g.rp[v]; q < g.rp[v + 1]; ++q)
                if (g.col[q] > v) hfcw.push_back(idw((g.col[q] & kFwdIdMask) | ((v % kTile) << 26), g.w[q]));
        }
        hfrp.push_back((uint32_t)(hfcw.size() - ri.fe_off));
        if (N > kFwdIdMask) huge_tile = true;  // ids take 26 bits of a forward entry
        for (uint32_t v = 0; v < N; v += kTile)  // entry offsets of a tile's queue codes
            if (hfrp[rpoff + std::min(v + (uint32_t)kTile, N)] - hfrp[rpoff + v] >= kMaxTileEdges) huge_tile = true;
        hw.insert(hw.end(), g.w.begin(), g.w.end());
        for (size_t q = 0; q < g.col.size(); ++q) hcw.push_back(idw(g.col[q], g.w[q]));
        // W_v = sum of out-edge weights, summed in CSR order in double
        for (uint32_t v = 0; v < N; ++v) {
            double W = 0.0;
            for (uint32_t q = g.rp[v]; q < g.rp[v + 1]; ++q) W += (double)g.w[q];
            hwsum.push_back(W);
        }
        hwsum.push_back(0.0);
        if (!sym) {
            any_asym = true;
            Csr in;
            transpose(g, N, in);
            ri.ie_off = ieoff;
            ri.irp_off = irpoff;
            hirp.insert(hirp.end(), in.rp.begin(), in.rp.end());
            hicol.insert(hicol.end(), in.col.begin(), in.col.end());
            hiw.insert(hiw.end(), in.w.begin(), in.w.end());
            for (size_t q = 0; q < in.col.size(); ++q) hicw.push_back(idw(in.col[q], in.w[q]));
            ieoff += in.col.size();
            irpoff += N + 1;
        } else {
            ri.ie_off = eoff;
            ri.irp_off = rpoff;
        }
        if (nh.C > 1) {
            ri.comp_off = coff;
            for (uint32_t v = 0; v < N; ++v) {
                const int32_t cv = t.compartment ? t.compartment[v] : 0;
                if (cv < 0 || cv >= nh.C) { delete c; return fail_create("compartment id out of range"); }
                hcomp.push_back(cv);
            }
            hash_bytes(c->model_hash, hcomp.data() + coff, (size_t)N * sizeof(int32_t));
            coff += N;
        }
        hx.insert(hx.end(), x0[r], x0[r] + (size_t)S * N);
        ri.tile_begin = tile;
        ri.ntiles = (N + kTile - 1) / kTile;
        for (uint32_t q = 0; q < ri.ntiles; ++q) tiles.push_back(TileEnt{(uint32_t)r, q * (uint32_t)kTile});
        tile += ri.ntiles;
        xoff += (uint64_t)S * N;
        eoff += g.col.size();
        rpoff += N + 1;
        c->info.push_back(ri);
        c->total_sub += N;
    }
    c->ntiles = tile;
    // decide_kernel: one block per 1,024 tile partials of the largest
    // realization (up to kDecideChunks) once it has 8,192 tiles or more;
    // below that the chunks' fences cost more than one block's reads
    // (Schlogl 64^3, 4,096 tiles: 43.7 -> 46.1 us per step with 4 chunks)
    {
        uint32_t maxt = 0;
        for (uint32_t r = 0; r < c->nreal; ++r) maxt = std::max(maxt, c->info[r].ntiles);
        c->decide_chunks = maxt >= 8192u ? std::min(kDecideChunks, (maxt + 1023u) / 1024u) : 1u;
    }
    c->total_counts = xoff;

    if (hipSetDevice(0) != hipSuccess) { delete c; g_create_err = "hipSetDevice failed"; return PLIS_ERR_NODEV; }
    c->device = 0;
    {
        std::string serr;
        if (!scratch_ptr(c->device, serr)) {
            delete c;
            g_create_err = serr;
            return PLIS_ERR_HIP;
        }
    }
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, 0) != hipSuccess) { delete c; g_create_err = "hipGetDeviceProperties failed"; return PLIS_ERR_HIP; }
    c->cus = prop.multiProcessorCount;
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        g_create_err = "hipStreamCreate failed";
        return PLIS_ERR_HIP;
    }
    c->ctl_host.resize(n_real);
    for (auto& rc : c->ctl_host) {
        rc = RealCtl{};
        rc.t_end = __builtin_huge_val();
        rc.t_pause = __builtin_huge_val();
        rc.done = 1u;
        rc.ev_idx = -1;
    }
    int rc = 0;
    std::vector<uint32_t> zeros_cnt(4, 0u);
    if ((rc = upload(c, &c->d_net_i, I)) || (rc = upload(c, &c->d_net_d, D)) ||
        (rc = upload(c, &c->d_info, c->info)) || (rc = upload(c, &c->d_ctl, c->ctl_host)) ||
        (rc = upload(c, &c->d_tiles, tiles)) || (rc = upload(c, &c->d_x0, hx)) ||
        (rc = dev_alloc(c, &c->d_x1, hx.size())) || (rc = upload(c, &c->d_rp, hrp)) ||
        (rc = upload(c, &c->d_wsum, hwsum)) ||
        (rc = upload(c, &c->d_col, hcol)) ||
        (rc = upload(c, &c->d_w, any_weighted ? hw : std::vector<float>())) ||
        (rc = upload(c, &c->d_irp, hirp)) || (rc = upload(c, &c->d_icol, hicol)) ||
        (rc = upload(c, &c->d_iw, hiw)) || (rc = upload(c, &c->d_comp, hcomp)) ||
        (rc = upload(c, &c->d_cw, hcw)) || (rc = upload(c, &c->d_icw, hicw)) ||
        (rc = dev_alloc(c, &c->d_pa_t, c->ntiles)) || (rc = dev_alloc(c, &c->d_pb_t, c->ntiles)) ||
        (rc = dev_alloc(c, &c->d_redo, c->ntiles)) ||
        (rc = dev_alloc(c, &c->d_pa_c, (size_t)c->nreal * kDecideChunks)) ||
        (rc = dev_alloc(c, &c->d_pb_c, (size_t)c->nreal * kDecideChunks)) ||
        (rc = upload(c, &c->d_dcnt, std::vector<uint32_t>(c->nreal, 0u))) ||
        (rc = upload(c, &c->d_cnt, zeros_cnt))) {
        g_create_err = c->err;
        plis_destroy(c);
        return rc;
    }
    // the forward CSR and the zeroed balances serve the edge-parallel leap
    const bool ep = !any_asym && !huge_tile;
    if (ep && ((rc = upload(c, &c->d_frp, hfrp)) || (rc = upload(c, &c->d_fcw, hfcw)) ||
                      (rc = upload(c, &c->d_gdelta, std::vector<unsigned long long>(hx.size(), 0ull))))) {
        g_create_err = c->err;
        plis_destroy(c);
        return rc;
    }
    c->edge_parallel = ep;
    if (hipHostMalloc((void**)&c->h_ndone, sizeof(uint32_t)) != hipSuccess) {
        g_create_err = "hipHostMalloc failed";
        plis_destroy(c);
        return PLIS_ERR_HIP;
    }
    set_shape(c, params->waves_per_block > 0 ? (uint32_t)params->waves_per_block : 4u,
              params->blocks_per_cu > 0 ? (uint32_t)params->blocks_per_cu : 4u);
    *out = c;
    return PLIS_OK;
}

static int pull_ctl(plis_ctx* c) {
    HIPCHK(c, hipMemcpyAsync(c->ctl_host.data(), c->d_ctl, c->ctl_host.size() * sizeof(RealCtl),
                             hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

// The tile partials no longer hold proposals for the current state (the
// host changed the state, the network or the tau mode).
static int invalidate_props(plis_ctx* c) {
    if (int rc = pull_ctl(c)) return rc;
    for (auto& q : c->ctl_host) q.prop_ok = 0u;
    HIPCHK(c, hipMemcpyAsync(c->d_ctl, c->ctl_host.data(), c->ctl_host.size() * sizeof(RealCtl),
                             hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

// Arm every realization for (t_end, t_pause, step_limit); returns how many
// are already done through *ndone_out.
static int arm(plis_ctx* c, double t_end, double t_pause, uint64_t step_limit, uint32_t* ndone_out) {
    if (int rc = pull_ctl(c)) return rc;
    uint32_t ndone = 0;
    for (auto& r : c->ctl_host) {
        r.t_end = t_end;
        r.t_pause = t_pause;
        r.step_limit = step_limit;
        const bool stop = r.status == kAbsorbed || r.status < 0 || r.t >= t_end || r.steps >= step_limit ||
                          r.t >= t_pause;
        r.done = stop ? 1u : 0u;
        if (!stop) r.status = kRunning;
        else if (r.status == kRunning && (r.t >= t_end || r.steps >= step_limit)) r.status = kDone;
        ndone += r.done;
    }
    HIPCHK(c, hipMemcpyAsync(c->d_ctl, c->ctl_host.data(), c->ctl_host.size() * sizeof(RealCtl),
                             hipMemcpyHostToDevice, c->stream));
    uint32_t* d_ndone = c->d_cnt;
    HIPCHK(c, hipMemcpyAsync(d_ndone, &ndone, sizeof(uint32_t), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    *ndone_out = ndone;
    return 0;
}

extern "C" int plis_arm(plis_ctx* c, double t_end, double t_pause, uint64_t step_limit) {
    if (!c) return PLIS_ERR_ARG;
    uint32_t nd = 0;
    return arm(c, t_end, t_pause, step_limit, &nd);
}

extern "C" int plis_run_until(plis_ctx* c, double t_end, double t_pause, uint64_t step_limit) {
    if (!c) return PLIS_ERR_ARG;
    uint32_t ndone = 0;
    if (int rc = arm(c, t_end, t_pause, step_limit, &ndone)) return rc;
    uint32_t* d_ndone = c->d_cnt;
    uint32_t k = 2;
    uint64_t polls = 0, last_progress = 0;
    auto progress = [&]() {
        uint64_t p = 0;
        for (const auto& r : c->ctl_host) p += r.steps + r.leapbacks;
        return p;
    };
    last_progress = progress();
    if (ndone < c->nreal)
        if (int rc = prologue(c)) return rc;
    while (ndone < c->nreal) {
        if (int rc = graph_iters(c, k)) return rc;
        HIPCHK(c, hipMemcpyAsync(c->h_ndone, d_ndone, sizeof(uint32_t), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        ndone = *c->h_ndone;
        if (k < 64) k *= 2;
        // every iteration of an armed realization either accepts a step or
        // leaps back; a batch of iterations without either means the control
        // state is inconsistent, so stop instead of spinning
        if (++polls % 8 == 0 && ndone < c->nreal) {
            if (int rc = pull_ctl(c)) return rc;
            const uint64_t p = progress();
            if (p == last_progress) {
                c->err = "no realization advanced over a batch of iterations";
                return PLIS_ERR_STATE;
            }
            last_progress = p;
        }
    }
    if (int rc = pull_ctl(c)) return rc;
    for (auto& r : c->ctl_host)
        if (r.status == kErrLeapback) {
            c->err = "a step needed more than 63 leap-backs";
            return PLIS_ERR_LEAPBACK;
        }
    return PLIS_OK;
}

extern "C" int plis_run(plis_ctx* c, double t_end, uint64_t step_limit) {
    return plis_run_until(c, t_end, __builtin_huge_val(), step_limit);
}

extern "C" int plis_iterate(plis_ctx* c, uint32_t n) {
    if (!c) return PLIS_ERR_ARG;
    if (n > 0)
        if (int rc = prologue(c)) return rc;
    if (int rc = graph_iters(c, n)) return rc;
    return PLIS_OK;
}

extern "C" int plis_profile_kernels(plis_ctx* c, uint32_t n, double* out) {
    if (!c || !out) return PLIS_ERR_ARG;
    // per iteration: before/after each of the 4 phases, and 2 events
    // between the leap pass's kernels
    // (the folded order launches the leap pass, the tau pass and commit +
    // decide; its decide slot stays empty and its commit slot times the
    // merged kernel)
    std::vector<hipEvent_t> ev(10 * (size_t)n + 2);
    for (auto& e : ev) HIPCHK(c, hipEventCreate(&e));
    const KernelArgs a = make_args(c);
    if (n > 0)
        if (int rc = prologue(c)) return rc;
    const bool fold = folded(c);
    HIPCHK(c, hipEventRecord(ev[0], c->stream));
    for (uint32_t i = 0; i < n; ++i)
        for (int j = 0; j < 4; ++j) {
            // slot k of the output, kernel group `which`, in launch order
            const int k = fold ? (j == 0 ? 2 : j == 1 ? 0 : j == 2 ? 1 : 3) : j;
            const int which = fold ? (k == 0 ? 4 : k == 3 ? 6 : k) : k;
            hipEvent_t* e = &ev[1 + 10 * (size_t)i];
            HIPCHK(c, hipEventRecord(e[2 * k], c->stream));
            c->split_ev = k == 2 ? &e[8] : nullptr;
            if (!(fold && k == 1)) launch_kernel(c, which, a);
            c->split_ev = nullptr;
            HIPCHK(c, hipEventRecord(e[2 * k + 1], c->stream));
        }
    HIPCHK(c, hipEventRecord(ev[10 * (size_t)n + 1], c->stream));
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipEventSynchronize(ev[10 * (size_t)n + 1]));
    float tot = 0.f;
    double t[7] = {0, 0, 0, 0, 0, 0, 0};
    HIPCHK(c, hipEventElapsedTime(&tot, ev[0], ev[10 * (size_t)n + 1]));
    for (uint32_t i = 0; i < n; ++i) {
        hipEvent_t* e = &ev[1 + 10 * (size_t)i];
        float x = 0.f;
        for (int k = 0; k < 4; ++k) {
            HIPCHK(c, hipEventElapsedTime(&x, e[2 * k], e[2 * k + 1]));
            t[k] += x;
        }
        if (c->edge_parallel) {
            HIPCHK(c, hipEventElapsedTime(&x, e[4], e[8]));
            t[4] += x;
            HIPCHK(c, hipEventElapsedTime(&x, e[8], e[9]));
            t[5] += x;
            HIPCHK(c, hipEventElapsedTime(&x, e[9], e[5]));
            t[6] += x;
        }
    }
    if (!c->edge_parallel) t[4] = t[2];  // one leap kernel
    for (auto& e : ev) (void)hipEventDestroy(e);
    out[0] = tot;
    for (int k = 0; k < 7; ++k) out[1 + k] = t[k];
    return PLIS_OK;
}

extern "C" int plis_profile(plis_ctx* c, uint32_t n, double* out) {
    if (!out) return PLIS_ERR_ARG;
    double o[8];
    const int rc = plis_profile_kernels(c, n, o);
    if (rc == PLIS_OK)
        for (int k = 0; k < 5; ++k) out[k] = o[k];
    return rc;
}

extern "C" int plis_sync(plis_ctx* c) {
    if (!c) return PLIS_ERR_ARG;
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PLIS_OK;
}

extern "C" int plis_set_state(plis_ctx* c, int32_t r, const uint32_t* x) {
    if (!c || !x || r < 0 || (uint32_t)r >= c->nreal) return PLIS_ERR_ARG;
    if (int rc = pull_ctl(c)) return rc;
    const RealInfo& ri = c->info[r];
    uint32_t* X = c->ctl_host[r].cur ? c->d_x1 : c->d_x0;
    HIPCHK(c, hipMemcpy(X + ri.x_off, x, (size_t)c->S * ri.N * sizeof(uint32_t), hipMemcpyHostToDevice));
    return invalidate_props(c);
}

extern "C" int plis_get_state(plis_ctx* c, int32_t r, uint32_t* x) {
    if (!c || !x || r < 0 || (uint32_t)r >= c->nreal) return PLIS_ERR_ARG;
    if (int rc = pull_ctl(c)) return rc;
    const RealInfo& ri = c->info[r];
    const uint32_t* X = c->ctl_host[r].cur ? c->d_x1 : c->d_x0;
    HIPCHK(c, hipMemcpy(x, X + ri.x_off, (size_t)c->S * ri.N * sizeof(uint32_t), hipMemcpyDeviceToHost));
    return PLIS_OK;
}

// ---- checkpoint / resume ----
// Blob: CkptHeader | N[nreal] | id[nreal] | RealCtl[nreal] | counts (each
// realization's current buffer, species-major, realizations in order).
struct CkptHeader {
    char magic[8];
    uint32_t version, nreal;
    int32_t S, R;
    uint64_t total_counts, seed;
    double eps;
    int32_t nc;
    uint32_t ctl_bytes;
    uint64_t model_hash;  // reaction tables, constants, every CSR and compartment table
};
constexpr char kCkptMagic[8] = {'P', 'L', 'I', 'S', 'C', 'K', 'P', 'T'};

static uint64_t ckpt_bytes(const plis_ctx* c) {
    return sizeof(CkptHeader) + 2ull * c->nreal * sizeof(uint32_t) + (uint64_t)c->nreal * sizeof(RealCtl) +
           c->total_counts * sizeof(uint32_t);
}

extern "C" int plis_checkpoint_size(plis_ctx* c, uint64_t* bytes) {
    if (!c || !bytes) return PLIS_ERR_ARG;
    *bytes = ckpt_bytes(c);
    return PLIS_OK;
}

extern "C" int plis_checkpoint(plis_ctx* c, void* buf, uint64_t bytes) {
    if (!c || !buf) return PLIS_ERR_ARG;
    if (bytes < ckpt_bytes(c)) {
        c->err = "plis_checkpoint: buffer smaller than plis_checkpoint_size";
        return PLIS_ERR_ARG;
    }
    if (int rc = pull_ctl(c)) return rc;
    char* p = static_cast<char*>(buf);
    CkptHeader h{};
    std::memcpy(h.magic, kCkptMagic, sizeof h.magic);
    h.version = 3u;
    h.nreal = c->nreal;
    h.S = c->S;
    h.R = c->R;
    h.total_counts = c->total_counts;
    h.seed = c->seed;
    h.eps = c->eps;
    h.nc = c->nc;
    h.ctl_bytes = (uint32_t)sizeof(RealCtl);
    h.model_hash = c->model_hash;
    std::memcpy(p, &h, sizeof h);
    p += sizeof h;
    for (uint32_t r = 0; r < c->nreal; ++r) std::memcpy(p + 4ull * r, &c->info[r].N, 4);
    p += 4ull * c->nreal;
    for (uint32_t r = 0; r < c->nreal; ++r) std::memcpy(p + 4ull * r, &c->info[r].id, 4);
    p += 4ull * c->nreal;
    std::memcpy(p, c->ctl_host.data(), c->nreal * sizeof(RealCtl));
    p += c->nreal * sizeof(RealCtl);
    for (uint32_t r = 0; r < c->nreal; ++r) {
        const RealInfo& ri = c->info[r];
        const uint32_t* X = c->ctl_host[r].cur ? c->d_x1 : c->d_x0;
        HIPCHK(c, hipMemcpy(p + 4ull * ri.x_off, X + ri.x_off, (size_t)c->S * ri.N * sizeof(uint32_t),
                            hipMemcpyDeviceToHost));
    }
    return PLIS_OK;
}

extern "C" int plis_restore(plis_ctx* c, const void* buf, uint64_t bytes) {
    if (!c || !buf) return PLIS_ERR_ARG;
    const char* p = static_cast<const char*>(buf);
    CkptHeader h;
    if (bytes < sizeof h) {
        c->err = "plis_restore: blob too short";
        return PLIS_ERR_ARG;
    }
    std::memcpy(&h, p, sizeof h);
    if (std::memcmp(h.magic, kCkptMagic, sizeof h.magic) != 0 || h.version != 3u ||
        h.ctl_bytes != sizeof(RealCtl)) {
        c->err = "plis_restore: not a checkpoint of this library version";
        return PLIS_ERR_ARG;
    }
    if (h.nreal != c->nreal || h.S != c->S || h.R != c->R || h.total_counts != c->total_counts ||
        h.seed != c->seed || h.eps != c->eps || h.nc != c->nc || bytes < ckpt_bytes(c)) {
        c->err = "plis_restore: checkpoint of a different model, size, seed, epsilon or n_c";
        return PLIS_ERR_ARG;
    }
    if (h.model_hash != c->model_hash) {
        c->err = "plis_restore: checkpoint of a different network, topology, edge weights or compartments";
        return PLIS_ERR_ARG;
    }
    p += sizeof h;
    for (uint32_t r = 0; r < c->nreal; ++r) {
        uint32_t N, id;
        std::memcpy(&N, p + 4ull * r, 4);
        std::memcpy(&id, p + 4ull * (c->nreal + r), 4);
        if (N != c->info[r].N || id != c->info[r].id) {
            c->err = "plis_restore: realization " + std::to_string(r) + " differs in size or id";
            return PLIS_ERR_ARG;
        }
    }
    p += 8ull * c->nreal;
    std::vector<RealCtl> ctl(c->nreal);
    std::memcpy(ctl.data(), p, c->nreal * sizeof(RealCtl));
    p += c->nreal * sizeof(RealCtl);
    for (auto& q : ctl) {
        q.cur = 0u;  // counts go to buffer 0
        q.prop_ok = 0u;  // the blob holds no tile partials
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    HIPCHK(c, hipMemcpy(c->d_x0, p, c->total_counts * sizeof(uint32_t), hipMemcpyHostToDevice));
    HIPCHK(c, hipMemcpy(c->d_ctl, ctl.data(), ctl.size() * sizeof(RealCtl), hipMemcpyHostToDevice));
    c->ctl_host = ctl;
    return PLIS_OK;
}

extern "C" int plis_get_stats(plis_ctx* c, int32_t r, plis_stats* st) {
    if (!c || !st || r < 0 || (uint32_t)r >= c->nreal) return PLIS_ERR_ARG;
    if (int rc = pull_ctl(c)) return rc;
    const RealCtl& q = c->ctl_host[r];
    st->t = q.t;
    st->tau = q.tau;
    st->steps = q.steps;
    st->leapbacks = q.leapbacks;
    st->firings = q.firings;
    st->subvolume_leaps = q.subvolume_leaps;
    st->crit_events = q.crit_events;
    st->status = q.status;
    st->pad = 0;
    st->learned_steps = q.learned_steps;
    st->learned_rejects = q.learned_rejects;
    st->violations = q.violations;
    return PLIS_OK;
}

extern "C" int plis_debug_tau(plis_ctx* c, double* tp, double* a0c, double* t2) {
    if (!c || !tp || !a0c || !t2) return PLIS_ERR_ARG;
    double *d_tp, *d_a0c, *d_t2;
    const size_t n = c->total_sub;
    HIPCHK(c, hipMalloc((void**)&d_tp, n * sizeof(double)));
    HIPCHK(c, hipMalloc((void**)&d_a0c, n * sizeof(double)));
    HIPCHK(c, hipMalloc((void**)&d_t2, n * sizeof(double)));
    KernelArgs a = make_args(c);
    a.dbg_tp = d_tp;
    a.dbg_a0c = d_a0c;
    a.dbg_t2 = d_t2;
    a.learned_on = 0;
    const int lo = c->learned_on;
    c->learned_on = 0;
    launch_kernel(c, 0, a);
    c->learned_on = lo;
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipMemcpy(tp, d_tp, n * sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(a0c, d_a0c, n * sizeof(double), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(t2, d_t2, n * sizeof(double), hipMemcpyDeviceToHost);
    (void)hipFree(d_tp);
    (void)hipFree(d_a0c);
    (void)hipFree(d_t2);
    if (e != hipSuccess) {
        c->err = std::string("plis_debug_tau: ") + hipGetErrorString(e);
        return PLIS_ERR_HIP;
    }
    return PLIS_OK;
}

extern "C" int plis_launch_shape(plis_ctx* c, int32_t* out) {
    if (!c || !out) return PLIS_ERR_ARG;
    out[0] = (int32_t)c->wpb;
    out[1] = (int32_t)c->blocks;
    out[2] = (int32_t)c->ntiles;
    return PLIS_OK;
}

// Sentinels of plis_debug_poison: harmless if read (no index leaves its
// range), but a read changes the trajectory -- a tiny tau in the tau
// partials, firings and a negative flag in the leap partials, tile 0 in the
// redo list (its open draws would be added twice), huge counts in the
// non-current count buffers.
constexpr uint32_t kPoisonCount = 0x7FFFFFF0u;

__global__ __launch_bounds__(256) void poison_parts_kernel(PartA* pa, PartB* pb, uint32_t* redo, uint32_t ntiles,
                                                           PartA* pc, PartB* pbc, uint32_t nchunks) {
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < ntiles; t += gridDim.x * blockDim.x) {
        pa[t] = PartA{1e-300, 1e-300, 0u, 0xDEADu};
        pb[t] = PartB{1ull << 40, 1u, 7u};
        redo[t] = 0u;
    }
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < nchunks; t += gridDim.x * blockDim.x) {
        pc[t] = PartA{1e-300, 1e-300, 0u, 0xDEADu};
        pbc[t] = PartB{1ull << 40, 1u, 7u};
    }
}

extern "C" int plis_debug_poison(plis_ctx* c) {
    if (!c) return PLIS_ERR_ARG;
    if (int rc = pull_ctl(c)) return rc;
    for (uint32_t r = 0; r < c->nreal; ++r) {
        const RealInfo& ri = c->info[r];
        uint32_t* other = c->ctl_host[r].cur ? c->d_x0 : c->d_x1;
        HIPCHK(c, hipMemsetD32Async(other + ri.x_off, (int)kPoisonCount, (size_t)c->S * ri.N, c->stream));
    }
    hipLaunchKernelGGL(poison_parts_kernel, dim3(std::min<uint32_t>((c->ntiles + 255) / 256, 1024u)), dim3(256), 0,
                       c->stream, c->d_pa_t, c->d_pb_t, c->d_redo, c->ntiles, c->d_pa_c, c->d_pb_c,
                       c->nreal * kDecideChunks);
    HIPCHK(c, hipGetLastError());
    // the partials no longer hold any proposal
    return invalidate_props(c);
}

extern "C" int plis_redo_tiles(plis_ctx* c, uint64_t* total) {
    if (!c || !total) return PLIS_ERR_ARG;
    HIPCHK(c, hipMemcpyAsync(total, c->d_cnt + 2, sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PLIS_OK;
}

extern "C" int plis_autotune(plis_ctx* c, uint32_t iters) {
    if (!c) return PLIS_ERR_ARG;
    if (iters < 1) iters = 1;
    const uint32_t wpbs[] = {1, 2, 4};
    const uint32_t bpcs[] = {1, 2, 4, 8, 16};
    double best = 1e300;
    uint32_t bw = c->wpb, bb = 4;
    std::vector<double> prof(5);
    for (uint32_t w : wpbs)
        for (uint32_t b : bpcs) {
            set_shape(c, w, b);
            if (int rc = plis_profile(c, 1, prof.data())) return rc;  // warm
            if (int rc = plis_profile(c, iters, prof.data())) return rc;
            if (prof[1] + prof[3] < best) {
                best = prof[1] + prof[3];
                bw = w;
                bb = b;
            }
        }
    set_shape(c, bw, bb);
    return PLIS_OK;
}

// ---------------------------------------------------------------------------
// learned tau: network state, training on the GPU, records, step mode
// ---------------------------------------------------------------------------
static int mlp_alloc(plis_ctx* c) {
    if (c->d_w32) return 0;
    int rc;
    std::vector<float> zero(kPTotal + 1, 0.0f), norm(26, 0.0f);
    for (int k = 12; k < 24; ++k) norm[k] = 1.0f;
    norm[25] = 1.0f;
    c->train_blocks = 1024;  // one 64-sample tile per block at batch 65,536
    if ((rc = upload(c, &c->d_w32, zero)) || (rc = upload(c, &c->d_m, zero)) || (rc = upload(c, &c->d_v, zero)) ||
        (rc = upload(c, &c->d_grad, zero)) || (rc = upload(c, &c->d_norm, norm)) ||
        (rc = dev_alloc(c, &c->d_q, (size_t)kQTotal)) ||
        (rc = dev_alloc(c, &c->d_partial, (size_t)c->train_blocks * (kPTotal + 1))))
        return rc;
    hipLaunchKernelGGL(mlp_pack_kernel, dim3((kQTotal + 255) / 256), dim3(256), 0, c->stream, c->d_w32, c->d_q);
    HIPCHK(c, hipGetLastError());
    return 0;
}

extern "C" int plis_mlp_set(plis_ctx* c, const float* w32, const float* norm) {
    if (!c || !w32 || !norm) return PLIS_ERR_ARG;
    if (int rc = mlp_alloc(c)) return rc;
    for (int k = 0; k < 12; ++k)
        if (!(norm[12 + k] > 0.0f)) {
            c->err = "feature scales must be > 0";
            return PLIS_ERR_ARG;
        }
    if (!(norm[25] > 0.0f)) {
        c->err = "target scale must be > 0";
        return PLIS_ERR_ARG;
    }
    std::vector<float> w(w32, w32 + kPTotal);
    for (int i = 12 * kNh; i < kNin * kNh; ++i) w[i] = 0.0f;  // padded inputs
    std::vector<float> zero(kPTotal, 0.0f);
    HIPCHK(c, hipMemcpyAsync(c->d_w32, w.data(), kPTotal * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_m, zero.data(), kPTotal * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_v, zero.data(), kPTotal * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_norm, norm, 26 * sizeof(float), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(mlp_pack_kernel, dim3((kQTotal + 255) / 256), dim3(256), 0, c->stream, c->d_w32, c->d_q);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->adam_t = 0;
    return invalidate_props(c);
}

extern "C" int plis_mlp_get(plis_ctx* c, float* w32) {
    if (!c || !w32) return PLIS_ERR_ARG;
    if (!c->d_w32) {
        c->err = "no network set";
        return PLIS_ERR_STATE;
    }
    HIPCHK(c, hipMemcpyAsync(w32, c->d_w32, kPTotal * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PLIS_OK;
}

extern "C" int plis_mlp_forward(plis_ctx* c, const float* feats, uint32_t n, float* out) {
    if (!c || !feats || !out || n == 0) return PLIS_ERR_ARG;
    if (!c->d_w32) {
        c->err = "no network set";
        return PLIS_ERR_STATE;
    }
    float *d_f = nullptr, *d_o = nullptr;
    HIPCHK(c, hipMalloc((void**)&d_f, (size_t)n * 12 * sizeof(float)));
    hipError_t e = hipMalloc((void**)&d_o, (size_t)n * 2 * sizeof(float));
    if (e == hipSuccess) e = hipMemcpy(d_f, feats, (size_t)n * 12 * sizeof(float), hipMemcpyHostToDevice);
    if (e == hipSuccess) {
        const MlpDev m{c->d_w32, c->d_q, c->d_norm};
        const uint32_t tiles = (n + 63) / 64;
        if (c->mlp_act == 1)
            hipLaunchKernelGGL(mlp_infer_kernel<1>, dim3(std::min<uint32_t>(tiles, 4096)), dim3(64), 0, c->stream, m,
                               d_f, n, d_o);
        else
            hipLaunchKernelGGL(mlp_infer_kernel<0>, dim3(std::min<uint32_t>(tiles, 4096)), dim3(64), 0, c->stream, m,
                               d_f, n, d_o);
        e = hipGetLastError();
    }
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipMemcpy(out, d_o, (size_t)n * 2 * sizeof(float), hipMemcpyDeviceToHost);
    (void)hipFree(d_f);
    if (d_o) (void)hipFree(d_o);
    if (e != hipSuccess) {
        c->err = std::string("plis_mlp_forward: ") + hipGetErrorString(e);
        return PLIS_ERR_HIP;
    }
    return PLIS_OK;
}

// plis_mlp_forward without any allocation: the feature rows and outputs are
// staged through this device's scratch (at most kNoAllocRows rows).  Every
// element the kernel reads is written by the same call first, so nothing
// needs clearing; calls that overlap on one GPU race on the scratch.
extern "C" int plis_mlp_forward_noalloc(plis_ctx* c, const float* feats, uint32_t n, float* out) {
    if (!c || !feats || !out || n == 0) return PLIS_ERR_ARG;
    if (n > kNoAllocRows) {
        c->err = "plis_mlp_forward_noalloc: at most " + std::to_string(kNoAllocRows) + " rows per call, got " +
                 std::to_string(n);
        return PLIS_ERR_ARG;
    }
    if (!c->d_w32) {
        c->err = "no network set";
        return PLIS_ERR_STATE;
    }
    std::string serr;
    float* base = scratch_ptr(c->device, serr);
    if (!base) {
        c->err = serr;
        return PLIS_ERR_HIP;
    }
    float* d_f = base;
    float* d_o = base + (size_t)kNoAllocRows * 12;
    HIPCHK(c, hipMemcpyAsync(d_f, feats, (size_t)n * 12 * sizeof(float), hipMemcpyHostToDevice, c->stream));
    const MlpDev m{c->d_w32, c->d_q, c->d_norm};
    const uint32_t tiles = (n + 63) / 64;
    if (c->mlp_act == 1)
        hipLaunchKernelGGL(mlp_infer_kernel<1>, dim3(tiles), dim3(64), 0, c->stream, m, d_f, n, d_o);
    else
        hipLaunchKernelGGL(mlp_infer_kernel<0>, dim3(tiles), dim3(64), 0, c->stream, m, d_f, n, d_o);
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(out, d_o, (size_t)n * 2 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PLIS_OK;
}

static int upload_records(plis_ctx* c, const float* feats, const float* ylog, const float* ycrit, uint64_t n) {
    if (n > c->rec_cap) {
        for (float* p : {c->d_rf, c->d_ry, c->d_rc})
            if (p) {
                c->allocs.erase(std::remove(c->allocs.begin(), c->allocs.end(), (void*)p), c->allocs.end());
                (void)hipFree(p);
            }
        c->d_rf = c->d_ry = c->d_rc = nullptr;
        int rc;
        if ((rc = dev_alloc(c, &c->d_rf, n * 12)) || (rc = dev_alloc(c, &c->d_ry, n)) ||
            (rc = dev_alloc(c, &c->d_rc, n)))
            return rc;
        c->rec_cap = n;
    }
    HIPCHK(c, hipMemcpyAsync(c->d_rf, feats, n * 12 * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_ry, ylog, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(c->d_rc, ycrit, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    return 0;
}

static void train_launch(plis_ctx* c, uint64_t n, uint32_t batch, uint32_t step, uint64_t seed, int full,
                         int32_t adam_t, float lr) {
    TrainArgs ta;
    ta.m = MlpDev{c->d_w32, c->d_q, c->d_norm};
    ta.feats = c->d_rf;
    ta.ylog = c->d_ry;
    ta.ycrit = c->d_rc;
    ta.n = n;
    ta.batch = batch;
    ta.step = step;
    ta.seed_lo = (uint32_t)seed;
    ta.seed_hi = (uint32_t)(seed >> 32);
    ta.full_pass = full;
    ta.partial = c->d_partial;
    const uint32_t blocks = std::min<uint32_t>(c->train_blocks, batch / 64);
    if (c->mlp_act == 1) hipLaunchKernelGGL(mlp_train_kernel<1>, dim3(blocks), dim3(64), 0, c->stream, ta);
    else hipLaunchKernelGGL(mlp_train_kernel<0>, dim3(blocks), dim3(64), 0, c->stream, ta);
    AdamArgs aa{c->d_w32, c->d_m, c->d_v, c->d_grad, c->d_partial, blocks, adam_t, lr, 0.9f, 0.999f, 1e-8f};
    hipLaunchKernelGGL(mlp_adam_kernel, dim3((kPTotal + 1 + kAdamParams - 1) / kAdamParams), dim3(256), 0,
                       c->stream, aa);
    if (adam_t > 0)
        hipLaunchKernelGGL(mlp_pack_kernel, dim3((kQTotal + 255) / 256), dim3(256), 0, c->stream, c->d_w32, c->d_q);
}

extern "C" int plis_mlp_loss_grad(plis_ctx* c, const float* feats, const float* ylog, const float* ycrit, uint32_t n,
                                  float* grad, double* loss) {
    if (!c || !feats || !ylog || !ycrit || !grad || n == 0 || n % 64 != 0) {
        if (c) c->err = "n must be a positive multiple of 64";
        return PLIS_ERR_ARG;
    }
    if (!c->d_w32) {
        c->err = "no network set";
        return PLIS_ERR_STATE;
    }
    if (int rc = upload_records(c, feats, ylog, ycrit, n)) return rc;
    train_launch(c, n, n, 0, 0, 1, 0, 0.0f);
    HIPCHK(c, hipGetLastError());
    std::vector<float> g(kPTotal + 1);
    HIPCHK(c, hipMemcpyAsync(g.data(), c->d_grad, (kPTotal + 1) * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    std::copy(g.begin(), g.begin() + kPTotal, grad);
    if (loss) *loss = (double)g[kPTotal] / (double)n;
    return PLIS_OK;
}

extern "C" int plis_mlp_train(plis_ctx* c, const float* feats, const float* ylog, const float* ycrit, uint64_t n,
                              uint32_t steps, uint32_t batch, float lr, uint64_t seed, float* losses) {
    if (!c || !feats || !ylog || !ycrit || n == 0 || batch == 0 || batch % 64 != 0) {
        if (c) c->err = "batch must be a positive multiple of 64";
        return PLIS_ERR_ARG;
    }
    if (!c->d_w32) {
        c->err = "no network set";
        return PLIS_ERR_STATE;
    }
    if (int rc = upload_records(c, feats, ylog, ycrit, n)) return rc;
    float* d_losses = nullptr;
    if (losses) HIPCHK(c, hipMalloc((void**)&d_losses, (size_t)steps * sizeof(float)));
    for (uint32_t t = 0; t < steps; ++t) {
        train_launch(c, n, batch, t, seed, 0, ++c->adam_t, lr);
        if (d_losses)
            (void)hipMemcpyAsync(d_losses + t, c->d_grad + kPTotal, sizeof(float), hipMemcpyDeviceToDevice, c->stream);
    }
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess && losses) {
        e = hipMemcpy(losses, d_losses, (size_t)steps * sizeof(float), hipMemcpyDeviceToHost);
        for (uint32_t t = 0; t < steps; ++t) losses[t] /= (float)batch;
    }
    if (d_losses) (void)hipFree(d_losses);
    if (e != hipSuccess) {
        c->err = std::string("plis_mlp_train: ") + hipGetErrorString(e);
        return PLIS_ERR_HIP;
    }
    return invalidate_props(c);
}

extern "C" int plis_records(plis_ctx* c, int32_t r, float* feats, float* ylog, float* crit) {
    if (!c || !feats || !ylog || !crit || r < 0 || (uint32_t)r >= c->nreal) return PLIS_ERR_ARG;
    if (c->variant > 3) {
        c->err = "learned tau supports networks of at most 8 species and 16 reactions";
        return PLIS_ERR_ARG;
    }
    const uint32_t N = c->info[r].N;
    float *df, *dy, *dc;
    HIPCHK(c, hipMalloc((void**)&df, (size_t)N * 12 * sizeof(float)));
    HIPCHK(c, hipMalloc((void**)&dy, (size_t)N * sizeof(float)));
    HIPCHK(c, hipMalloc((void**)&dc, (size_t)N * sizeof(float)));
    KernelArgs a = make_args(c);
    a.rec_feats = df;
    a.rec_ylog = dy;
    a.rec_crit = dc;
    a.rec_real = (uint32_t)r;
    launch_kernel(c, 5, a);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(c->stream);
    if (e == hipSuccess) e = hipMemcpy(feats, df, (size_t)N * 12 * sizeof(float), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(ylog, dy, (size_t)N * sizeof(float), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(crit, dc, (size_t)N * sizeof(float), hipMemcpyDeviceToHost);
    (void)hipFree(df);
    (void)hipFree(dy);
    (void)hipFree(dc);
    if (e != hipSuccess) {
        c->err = std::string("plis_records: ") + hipGetErrorString(e);
        return PLIS_ERR_HIP;
    }
    return PLIS_OK;
}

extern "C" int plis_set_tau_mode(plis_ctx* c, int32_t learned, double kappa, double theta, double viol_rate_max) {
    if (!c || !(kappa > 0.0) || !(theta >= 0.0) || !(viol_rate_max >= 0.0)) return PLIS_ERR_ARG;
    if (learned && !(theta > 0.0)) {
        c->err = "learned tau needs the judge: theta > 0";
        return PLIS_ERR_ARG;
    }
    if (learned && (c->variant > 3 || !c->d_w32)) {
        c->err = c->d_w32 ? "learned tau supports networks of at most 8 species and 16 reactions"
                          : "set the network (plis_mlp_set) before enabling learned tau";
        return PLIS_ERR_STATE;
    }
    c->learned_on = learned ? 1 : 0;
    c->kappa = kappa;
    c->theta = theta;
    c->viol_rate_max = viol_rate_max;
    drop_graphs(c);
    if (int rc = pull_ctl(c)) return rc;
    for (auto& q : c->ctl_host) {
        q.fallback = 0u;
        q.prop_ok = 0u;
    }
    HIPCHK(c, hipMemcpyAsync(c->d_ctl, c->ctl_host.data(), c->ctl_host.size() * sizeof(RealCtl),
                             hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return PLIS_OK;
}

extern "C" int plis_mlp_activation(plis_ctx* c, int32_t act) {
    if (!c || (act != 0 && act != 1)) return PLIS_ERR_ARG;
    c->mlp_act = act;
    drop_graphs(c);
    return invalidate_props(c);
}
