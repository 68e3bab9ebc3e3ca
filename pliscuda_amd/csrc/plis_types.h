// plis_types.h -- device-resident layout of a batch of RDME realizations.
//
// HBM layout (one context = one batch of realizations sharing a reaction network):
//   counts  : two ping-pong uint32 buffers; realization r occupies
//             [x_off, x_off + S*N) species-major ([s*N + v], coalesced per
//             species like the reference's GET_SPI, PLISCuda/include/cuda/constants.cuh:27)
//   topology: canonical CSR per realization (row_ptr relative to e_off, col
//             local ids, optional float weights); a separate in-CSR only
//             when the graph is not symmetric
//   tiles   : 64 consecutive subvolumes of one realization = one wavefront;
//             64 tiles = one reduction group
//   control : one RealCtl per realization (step, clocks, tau, event, stats)
#pragma once
#include <stdint.h>

namespace plis {

constexpr int kTile = 64;        // subvolumes per tile (one wavefront)
constexpr uint32_t kFwdIdMask = (1u << 26) - 1u;  // neighbour id bits of a forward entry
constexpr int kMaxS = 64;
constexpr int kMaxR = 128;

struct NetHeader {
    int32_t S, R, C, pad;
    // offsets (in units of their element type) into the network blob
    int32_t off_nu;      // int32 [s*R + j]
    int32_t off_rs;      // int32 [j*3 + i] reactant species
    int32_t off_rr;      // int32 [j*3 + i] reactant multiplicity
    int32_t off_nrs;     // int32 [j]
    int32_t off_order;   // int32 [j]
    int32_t off_hor;     // int32 [s]
    int32_t off_mult;    // int32 [s]
    int32_t off_k;       // double [c*R + j] (offset in doubles)
    int32_t off_d;       // double [c*S + s] (offset in doubles)
    int32_t ints_total, dbl_total;
};

struct RealInfo {
    uint64_t x_off;      // counts offset (elements)
    uint64_t e_off;      // out-CSR edge base
    uint64_t ie_off;     // in-CSR edge base (== e_off when symmetric)
    uint64_t rp_off;     // out row_ptr base (N+1 entries)
    uint64_t irp_off;    // in row_ptr base
    uint64_t comp_off;   // compartment ids base (if C > 1)
    uint64_t fe_off;     // forward-CSR edge base (entries v -> u with u > v; symmetric only)
    uint32_t N, id;
    uint32_t tile_begin, ntiles;
    uint32_t flags;      // bit0: symmetric (in-CSR aliases out-CSR), bit1: weighted
    uint32_t pad;
};

enum CtlStatus : int32_t { kRunning = 0, kDone = 1, kAbsorbed = 2, kErrLeapback = -3, kErrOverflow = -4 };

struct RealCtl {
    double t, t_end, t_pause;
    double tau_p, tau2, tau;
    uint64_t step_limit;
    uint64_t steps, leapbacks, firings, subvolume_leaps, crit_events;
    uint32_t step, attempt, vstar, ev_dst;
    int32_t crit, last, ev_kind, ev_idx;
    uint32_t done, retry, cur;
    int32_t status;
    // learned tau: this step's tau came from the network / the step after a
    // rejected proposal uses Cao's formula
    uint32_t learned, fallback;
    uint64_t learned_steps, learned_rejects, violations;
    // the tile partials hold the network's proposals for the current state
    // (written by the leap pass that produced it); cleared whenever the
    // state or the network changes otherwise
    uint32_t prop_ok, pad2;
};

struct TileEnt {
    uint32_t r, v0;
};

// Per-step partials.
struct PartA {
    double tp, t2;
    uint32_t v2, pad;
};
struct PartB {
    unsigned long long fire;
    uint32_t bad, viol;
};

struct KernelArgs {
    const int32_t* net_i;
    const double* net_d;
    NetHeader nh;
    const RealInfo* info;
    RealCtl* ctl;
    const TileEnt* tiles;
    uint32_t ntiles;
    uint32_t nreal;
    uint32_t* x0;
    uint32_t* x1;
    const uint32_t* rp;
    const double* wsum;       // sum of out-edge weights per subvolume (row_ptr layout)
    const uint32_t* col;
    const float* w;
    const uint32_t* irp;
    const uint32_t* icol;
    const float* iw;
    // the same edges interleaved, one 8-byte load per edge: neighbour id in
    // the low word, weight's float bits in the high word (1.0 unweighted) --
    // cw parallel to col / w, icw to icol / iw
    const uint64_t* cw;
    const uint64_t* icw;
    const int32_t* comp;
    PartA* partA_tile;
    PartA* partA_chunk;      // decide_kernel: one partial per (realization, chunk of its tiles)
    PartB* partB_chunk;      // commit_decide_kernel: the same for the leap pass's sums
    uint32_t* decide_cnt;    // decide_kernel: chunks done per realization (zero between steps)
    uint32_t decide_chunks;  // decide_kernel: blocks per realization (1 .. kDecideChunks)
    PartB* partB_tile;
    uint32_t* ndone;
    uint32_t* redo_count;    // tiles phase B left to the redo pass (reset by decide)
    uint32_t* redo_tiles;
    // forward CSR (row_ptr layout of rp; entries v -> u with u > v) and the
    // leap pass's global per-subvolume balances (counts layout, zero between
    // passes).  An entry is one 8-byte word: u in bits 0..25, the index of
    // its row v within v's tile in bits 26..31, the weight's float bits in
    // the high word (1.0 unweighted); realizations of the edge-parallel
    // pass have fewer than 2^26 subvolumes.
    const uint32_t* frp;
    const uint64_t* fcw;
    unsigned long long* gdelta;
    unsigned long long* redo_total;  // tiles redone since the context was made
    double epsilon;
    int32_t n_critical;
    uint32_t seed_lo, seed_hi;
    // debug outputs of phase A (NULL in normal runs): per-subvolume tau', a0c, tau2
    double* dbg_tp;
    double* dbg_a0c;
    double* dbg_t2;
    uint32_t waves_per_block;
    // post-leap leap condition (propensity form, tolerance theta) and
    // learned tau; judge_on: evaluate the condition in the leap pass
    double theta;
    double viol_rate_max;    // accepted fraction of violating subvolumes (learned steps)
    double kappa;            // multiplier of the learned proposal
    int32_t learned_on;
    int32_t judge_on;
    int32_t props_in_finish; // the leap pass proposes the next step's tau
    int32_t mlp_act;         // activation of the network: 0 ReLU, 1 GELU
    // the tau pass of an iteration in the folded order: it runs on the counts
    // the leap pass just wrote, for the step after this one
    uint32_t tau_spec;
    const float* mlp_w32;    // fp32 master blob
    const void* mlp_q;       // packed bf16 copies
    const float* mlp_norm;   // feature / target standardisation
    // records output (records kernel)
    float* rec_feats;
    float* rec_ylog;
    float* rec_crit;         // trial-leap reject label
    uint32_t rec_real;
};

}  // namespace plis
