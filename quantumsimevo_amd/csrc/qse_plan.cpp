// qse_plan.cpp — host fusion planner: turns a gate list (qse_op, the C form of
// the reference's GateCommand list replayed by Circuit::execute,
// Circuit.cpp:64-120) into LDS-tiled passes for the gfx950 fused kernel.
//
// Phase 1 (passes): grow the current pass's set of tile qubits T (|T| <= k,
// always holding the low qubits 0..L-1) and give the pass every gate whose
// non-diagonal qubits lie in T. In gate order (PLAN_DAG clear), a gate that is
// non-diagonal on a qubit outside T, when T is full, closes the pass. With
// PLAN_DAG the gates are list-scheduled over their dependency graph instead
// (gates commute when the qubits they share are diagonal in both): the pass
// takes every ready gate that fits T, and T grows by the qubits of the
// earliest ready gate that does not, while there is room; later gates of a
// light cone inside T thus join the pass out of order. Diagonal gates (Z,
// phase, RZ, cphase, mcphase) never close a pass. A SWAP of two tile qubits is
// a relabelling of tile bits (no data movement until the pass's store, which
// writes each tile bit to its final qubit). Inside a pass the gates are
// ordered the same way over register sets of 4 target qubits, so that the
// stage builder below packs them into few stages.
// Probes (build_reverse_plan): the reverse sweep of the adjoint method runs
// as joint passes over (psi, lambda); a parameterised gate's generator G is
// probed, Im<lambda|G|psi> at the point after the gate, before the gate is
// un-applied. X/Y probes are scheduled like a non-diagonal gate on their
// qubit (a register bit when evaluated), Z-string and projector probes like
// a diagonal gate (they commute with every diagonal gate applied to both
// states).
// Phase 2 (stages): inside a pass, consecutive non-diagonal gates share a
// stage while their targets fit in 4 register bits. Diagonal gates are kept
// pending as monomials (d applied where all bits of a mask are 1) and flushed
// right before the next non-diagonal gate on one of their qubits, or at the
// end of the pass; flushed monomials are merged by register mask and by
// thread/external condition, and conditions on external qubits are folded into
// per-tile factors computed once per tile.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <complex>
#include <map>
#include <set>
#include <string>
#include <vector>

#include "qse_fused.hpp"
#include "qse_internal.hpp"

namespace qse {

// Angle of op: theta, or theta + scale * params[param] for a parameterised op.
double op_angle(const qse_op& op, const double* params) {
    if (op.param >= 0 && params) return op.theta + op.scale * params[op.param];
    return op.theta;
}

int validate_ops(int n, const qse_op* ops, int nops, const double* params) {
    for (int k = 0; k < nops; ++k) {
        const qse_op& o = ops[k];
        auto badq = [&](int q) { return q < 0 || q >= n; };
        switch (o.kind) {
            case QSE_GATE_IDENTITY: break;
            case QSE_GATE_X: case QSE_GATE_Y: case QSE_GATE_Z: case QSE_GATE_H: case QSE_GATE_PHASE:
            case QSE_GATE_RX: case QSE_GATE_RY: case QSE_GATE_RZ: case QSE_GATE_MATRIX:
                if (badq(o.target)) return set_error(QSE_ERR_INVALID_ARG, "op " + std::to_string(k) + ": bad target");
                break;
            case QSE_GATE_CNOT: case QSE_GATE_CPHASE: case QSE_GATE_SWAP:
                if (badq(o.target) || badq(o.control))
                    return set_error(QSE_ERR_INVALID_ARG, "op " + std::to_string(k) + ": bad qubit");
                break;
            case QSE_GATE_MCPHASE:
                if (badq(o.target) || (n < 64 && (o.controls >> n)))
                    return set_error(QSE_ERR_INVALID_ARG, "op " + std::to_string(k) + ": bad qubit");
                break;
            case QSE_GATE_MEASURE: case QSE_GATE_RESET_QUBIT:
                return set_error(QSE_ERR_UNSUPPORTED, "op " + std::to_string(k) +
                                                          ": MEASURE and RESET_QUBIT run through qse_run_measured");
            default:
                return set_error(QSE_ERR_UNSUPPORTED, "op " + std::to_string(k) + ": unsupported gate kind " +
                                                          std::to_string(o.kind));
        }
        if (o.param >= 0 && !params)
            return set_error(QSE_ERR_INVALID_ARG, "op " + std::to_string(k) + " is parameterised but params is NULL");
    }
    return QSE_OK;
}

using cd = std::complex<double>;

namespace {

enum ItemType { IT_DENSE, IT_MONO, IT_SWAP, IT_PROBE };

struct Item {
    ItemType t;
    int q = -1, q2 = -1;   // dense target / swap qubits; X/Y probe qubit
    uint32_t dtype = FOP_DENSE;
    uint64_t cmask = 0;    // dense: control qubits
    cd m[4];
    uint64_t qmask = 0;    // mono: qubits that must be 1; Z-string/projector probe: its qubits
    cd d;
    uint32_t pkind = 0;    // probe: ProbeKind
    int pid = -1;          // probe: caller's id
};

bool probe_nd(const Item& it) { return it.t == IT_PROBE && (it.pkind == PROBE_X || it.pkind == PROBE_Y); }

struct Mono {
    uint64_t lmask;  // tile local bits
    uint64_t emask;  // external qubits
    cd d;
};

struct SOp {
    bool mono;
    // dense
    uint32_t dtype;
    int lt;
    uint64_t lctl, ectl;
    cd m[4];
    // mono
    std::vector<Mono> monos;
    // probe (dtype == FOP_PROBE): kind, id; X/Y: lt; Z-string/projector: tile local bits in lctl, external in ectl
    uint32_t pkind = 0;
    int pid = -1;
};

struct SStage {
    std::vector<int> R;  // register local bits, position order
    std::vector<SOp> ops;
};

}  // namespace

// CNOT(c,t) RZ(theta, t) CNOT(c,t) with c != t is exp(-i theta Z_c Z_t / 2): a
// diagonal (ZZ) phase, the cost-layer pattern of QAOA (QAOA.cpp:128-133).
bool is_zz_triple(const qse_op* ops, int nops, int k) {
    if (k + 2 >= nops) return false;
    const qse_op &a = ops[k], &b = ops[k + 1], &c = ops[k + 2];
    return a.kind == QSE_GATE_CNOT && c.kind == QSE_GATE_CNOT && b.kind == QSE_GATE_RZ && a.control == c.control &&
           a.target == c.target && a.control != a.target && b.target == a.target;
}

namespace {

void to_items(const qse_op* ops, int nops, const double* params, std::vector<Item>& out) {
    const double s = 0.7071067811865475244;
    for (int k = 0; k < nops; ++k) {
        const qse_op& o = ops[k];
        const double th = op_angle(o, params);
        Item it;
        if (is_zz_triple(ops, nops, k)) {
            // phase exp(-i th/2) on even parity of (c, t), exp(+i th/2) on odd:
            // exp(-i th/2) * exp(i th x_c) * exp(i th x_t) * exp(-2 i th x_c x_t)
            const double z = op_angle(ops[k + 1], params);
            const uint64_t cb = 1ULL << o.control, tb2 = 1ULL << o.target;
            auto push = [&](uint64_t mask, cd d) {
                Item mi;
                mi.t = IT_MONO;
                mi.qmask = mask;
                mi.d = d;
                out.push_back(mi);
            };
            push(0, cd(std::cos(z / 2), -std::sin(z / 2)));
            push(cb, cd(std::cos(z), std::sin(z)));
            push(tb2, cd(std::cos(z), std::sin(z)));
            push(cb | tb2, cd(std::cos(2 * z), -std::sin(2 * z)));
            k += 2;
            continue;
        }
        auto dense = [&](uint32_t dt, cd a, cd b, cd c, cd d) {
            it.t = IT_DENSE;
            it.q = o.target;
            it.dtype = dt;
            it.m[0] = a; it.m[1] = b; it.m[2] = c; it.m[3] = d;
            out.push_back(it);
        };
        auto mono = [&](uint64_t mask, cd d) {
            Item mi;
            mi.t = IT_MONO;
            mi.qmask = mask;
            mi.d = d;
            out.push_back(mi);
        };
        const uint64_t tb = 1ULL << o.target;
        switch (o.kind) {
            case QSE_GATE_IDENTITY: break;
            case QSE_GATE_H: dense(FOP_H, s, s, s, -s); break;
            case QSE_GATE_X: dense(FOP_XSWAP, 0, 1, 1, 0); break;
            case QSE_GATE_Y: dense(FOP_DENSE, 0, cd(0, -1), cd(0, 1), 0); break;
            case QSE_GATE_RX: {
                const double c = std::cos(th / 2), sn = std::sin(th / 2);
                dense(FOP_DENSE, c, cd(0, -sn), cd(0, -sn), c);
                break;
            }
            case QSE_GATE_RY: {
                const double c = std::cos(th / 2), sn = std::sin(th / 2);
                dense(FOP_DENSE, c, -sn, sn, c);
                break;
            }
            case QSE_GATE_MATRIX:
                dense(FOP_DENSE, cd(o.m[0], o.m[1]), cd(o.m[2], o.m[3]), cd(o.m[4], o.m[5]), cd(o.m[6], o.m[7]));
                break;
            case QSE_GATE_CNOT:
                if (o.control == o.target) break;  // no-op in the reference loop
                it.cmask = 1ULL << o.control;
                dense(FOP_XSWAP, 0, 1, 1, 0);
                break;
            case QSE_GATE_Z: mono(tb, -1.0); break;
            case QSE_GATE_PHASE: mono(tb, cd(std::cos(th), std::sin(th))); break;
            case QSE_GATE_RZ:
                mono(0, cd(std::cos(th / 2), -std::sin(th / 2)));
                mono(tb, cd(std::cos(th), std::sin(th)));
                break;
            case QSE_GATE_CPHASE: mono(tb | (1ULL << o.control), cd(std::cos(th), std::sin(th))); break;
            case QSE_GATE_MCPHASE: mono(tb | o.controls, cd(std::cos(th), std::sin(th))); break;
            case QSE_GATE_SWAP:
                if (o.target == o.control) break;
                it.t = IT_SWAP;
                it.q = o.target;
                it.q2 = o.control;
                out.push_back(it);
                break;
            default: break;  // validated earlier
        }
    }
}

struct PassItems {
    uint64_t T = 0;
    std::vector<Item> items;
};

int popc(uint64_t x) { return __builtin_popcountll(x); }

}  // namespace


namespace {

// Qubits on which an item is not diagonal (must be tile/register qubits) and
// qubits on which it is diagonal (controls and phase masks).
uint64_t nd_mask(const Item& it) {
    if (it.t == IT_DENSE || probe_nd(it)) return 1ULL << it.q;
    if (it.t == IT_SWAP) return (1ULL << it.q) | (1ULL << it.q2);
    return 0;
}
uint64_t dg_mask(const Item& it) {
    if (it.t == IT_DENSE) return it.cmask;
    if (it.t == IT_MONO || (it.t == IT_PROBE && !probe_nd(it))) return it.qmask;
    return 0;
}

// Dependency graph of an item list: item j precedes item i (j < i) when they
// share a qubit on which at least one of them is not diagonal.
struct Dag {
    std::vector<std::vector<int>> succ;
    std::vector<int> indeg;
};

void build_dag(const std::vector<Item>& v, int n, Dag& D) {
    const int N = (int)v.size();
    D.succ.assign(N, {});
    D.indeg.assign(N, 0);
    std::vector<int> lastND(n, -1), mark(N, -1);
    std::vector<std::vector<int>> diagSince(n);
    for (int i = 0; i < N; ++i) {
        auto dep = [&](int j) {
            if (j >= 0 && mark[j] != i) {
                mark[j] = i;
                D.succ[j].push_back(i);
                ++D.indeg[i];
            }
        };
        const uint64_t nd = nd_mask(v[i]), dg = dg_mask(v[i]);
        for (int q = 0; q < n; ++q) {
            if ((nd >> q) & 1) {
                dep(lastND[q]);
                for (int j : diagSince[q]) dep(j);
                lastND[q] = i;
                diagSince[q].clear();
            } else if ((dg >> q) & 1) {
                dep(lastND[q]);
                diagSince[q].push_back(i);
            }
        }
    }
}

// List scheduling of `v` into groups: a group has a qubit set S (starting at
// `start`, at most `cap` qubits) and takes every ready item whose required
// qubits (`need`) lie in S; S grows by the requirement of the earliest ready
// item that does not fit while |S| stays <= cap. Returns the groups in order;
// their concatenation is a topological order of `v`.
// With `lookahead`, S instead grows by the requirement of the ready item
// whose qubits make the most items executable per added qubit (the items
// that would run in S plus those qubits, given what has run so far).
template <class Need>
void list_schedule(const std::vector<Item>& v, int n, uint64_t start, int cap, Need need,
                   std::vector<std::pair<uint64_t, std::vector<int>>>& groups, bool lookahead = false) {
    Dag D;
    build_dag(v, n, D);
    std::vector<int> deg, stack;
    auto closure = [&](const std::set<int>& ready, uint64_t S2) {
        deg = D.indeg;
        stack.clear();
        for (int i : ready)
            if ((need(v[i]) & ~S2) == 0) stack.push_back(i);
        int cnt = 0;
        while (!stack.empty()) {
            const int i = stack.back();
            stack.pop_back();
            ++cnt;
            for (int s : D.succ[i])
                if (--deg[s] == 0 && (need(v[s]) & ~S2) == 0) stack.push_back(s);
        }
        return cnt;
    };
    std::set<int> ready;
    for (int i = 0; i < (int)v.size(); ++i)
        if (D.indeg[i] == 0) ready.insert(i);
    size_t done = 0;
    while (done < v.size()) {
        uint64_t S = start;
        std::vector<int> g;
        for (;;) {
            bool any = false;
            for (auto it = ready.begin(); it != ready.end();) {
                const int i = *it;
                if ((need(v[i]) & ~S) == 0) {
                    it = ready.erase(it);
                    g.push_back(i);
                    ++done;
                    any = true;
                    for (int s : D.succ[i])
                        if (--D.indeg[s] == 0) ready.insert(s);  // visited in this sweep when s > i
                } else {
                    ++it;
                }
            }
            if (any) continue;
            bool grown = false;
            if (lookahead) {
                uint64_t best_m = 0;
                double best = -1.0;
                std::vector<uint64_t> seen;
                for (int i : ready) {
                    const uint64_t m = need(v[i]) & ~S;
                    if (m == 0 || popc(S | m) > cap || std::find(seen.begin(), seen.end(), m) != seen.end()) continue;
                    seen.push_back(m);
                    const double sc = (double)closure(ready, S | m) / (double)popc(m);
                    if (sc > best) {
                        best = sc;
                        best_m = m;
                    }
                }
                if (best_m) {
                    S |= best_m;
                    grown = true;
                }
            } else {
                for (int i : ready) {
                    const uint64_t m = need(v[i]) & ~S;
                    if (popc(S | m) <= cap) {
                        S |= m;
                        grown = true;
                        break;
                    }
                }
            }
            if (!grown) break;
        }
        groups.push_back({S, std::move(g)});
    }
}

// Reorder the items of one pass so that consecutive runs use at most
// FUSE_REG target qubits (the stage builder then needs few stages). SWAPs
// are tile-bit relabellings in phase 2 and need no register.
void order_for_stages(std::vector<Item>& items, int n) {
    std::vector<std::pair<uint64_t, std::vector<int>>> groups;
    list_schedule(items, n, 0, FUSE_REG,
                  [](const Item& it) { return (it.t == IT_DENSE || probe_nd(it)) ? nd_mask(it) : 0ULL; }, groups);
    std::vector<Item> out;
    out.reserve(items.size());
    for (auto& g : groups)
        for (int i : g.second) out.push_back(items[i]);
    items.swap(out);
}

}  // namespace

// Choose (k, L) for an n-qubit state of precision prec given the context
// settings (0 = default). The defaults are the shapes the tuner below picks
// most often: contiguous runs of 2^L amplitudes of at least 128 bytes.
void fusion_shape(const qse_ctx* ctx, int n, int prec, int* k_out, int* L_out) {
    int k = ctx ? ctx->tile_bits[prec] : 0;
    int L = ctx ? ctx->low_bits[prec] : 0;
    if (k <= 0) k = (prec == QSE_C128) ? 12 : 13;
    if (L <= 0) L = (prec == QSE_C128) ? 3 : 4;
    k = std::min(k, FUSE_MAX_K);
    k = std::min(k, n - 2);
    k = std::max(k, FUSE_MIN_K);
    L = std::min(L, k - FUSE_REG);
    L = std::max(L, 0);
    *k_out = k;
    *L_out = L;
}

// PLAN_RELABEL: SWAPs become a logical->physical qubit map (no data movement);
// the last pass then stores out of place with the inverse map so the state
// ends in the canonical layout. Otherwise SWAPs are in-tile relabellings (in
// place). PLAN_DAG: passes by list scheduling over the dependency graph.
namespace {

// Smallest state (qubits) whose tiles are numbered far apart (see plan_items).
constexpr int TILE_ORDER_FAR_MIN_QUBITS = 27;

int plan_items(int n, int k, int L, const std::vector<Item>& items, HostPlan& P, int flags) {
    const bool relabel = (flags & PLAN_RELABEL) != 0;
    const uint64_t lowT = (L > 0) ? ((1ULL << L) - 1) : 0;
    int phys[64];
    for (int q = 0; q < 64; ++q) phys[q] = q;
    auto map_mask = [&](uint64_t m) {
        uint64_t r = 0;
        for (int q = 0; q < n; ++q)
            if ((m >> q) & 1) r |= 1ULL << phys[q];
        return r;
    };
    // ---- phase 1: passes ----
    std::vector<Item> mapped;
    mapped.reserve(items.size());
    for (Item it : items) {
        if (relabel) {
            if (it.t == IT_SWAP) {
                std::swap(phys[it.q], phys[it.q2]);
                continue;
            }
            if (it.t == IT_DENSE || probe_nd(it)) {
                it.q = phys[it.q];
                it.cmask = map_mask(it.cmask);
            } else {
                it.qmask = map_mask(it.qmask);
            }
        }
        mapped.push_back(it);
    }
    std::vector<PassItems> passes;
    if (flags & PLAN_DAG) {
        std::vector<std::pair<uint64_t, std::vector<int>>> groups;
        list_schedule(mapped, n, lowT, k, nd_mask, groups, (flags & PLAN_LOOKAHEAD) != 0);
        for (auto& g : groups) {
            PassItems pi;
            pi.T = g.first;
            for (int i : g.second) pi.items.push_back(mapped[i]);
            passes.push_back(std::move(pi));
        }
        if (passes.empty()) passes.push_back({lowT, {}});
    } else {
        passes.push_back({lowT, {}});
        for (const Item& it : mapped) {
            PassItems* cur = &passes.back();
            const uint64_t need = nd_mask(it);
            if (need) {
                const uint64_t miss = need & ~cur->T;
                if (popc(cur->T | miss) > k) {
                    passes.push_back({lowT | need, {}});
                    cur = &passes.back();
                } else {
                    cur->T |= miss;
                }
            }
            cur->items.push_back(it);
        }
    }
    for (PassItems& pi : passes) order_for_stages(pi.items, n);
    bool permuted = false;
    for (int q = 0; q < n; ++q) permuted |= (phys[q] != q);
    if (relabel && permuted) {
        // the destination's low qubits 0..L-1 come from phys[0..L-1]: they must be
        // tile qubits of the last pass for coalesced writes
        uint64_t need = 0;
        for (int q = 0; q < L; ++q) need |= 1ULL << phys[q];
        if (popc(lowT | need) > k)
            return set_error(QSE_ERR_UNSUPPORTED, "relabelled plan: read and write runs do not fit one tile");
        PassItems* cur = &passes.back();
        if (popc(cur->T | need) > k) passes.push_back({lowT | need, {}});
        else cur->T |= need;
    }
    if (passes.size() > 1 && passes.back().items.empty() && !(relabel && permuted)) passes.pop_back();
    // PLAN_CONTIG: a pass may store out of place in another qubit order. cur[q]
    // is where the qubit that phase 1 calls q sits in the state when the pass
    // being built runs; each pass's tile and items are moved there first.
    const bool contig = (flags & PLAN_CONTIG) != 0;
    int cur[64];
    for (int q = 0; q < 64; ++q) cur[q] = q;
    auto cmap = [&](uint64_t m) {
        uint64_t r = 0;
        for (int q = 0; q < n; ++q)
            if ((m >> q) & 1) r |= 1ULL << cur[q];
        return r;
    };
    // ---- phase 2: per pass, fill T to k qubits and build stages ----
    P.k = k;
    for (PassItems& pi : passes) {
        if (contig) {
            pi.T = cmap(pi.T);
            for (Item& it : pi.items) {
                if (it.q >= 0) it.q = cur[it.q];
                if (it.q2 >= 0) it.q2 = cur[it.q2];
                it.cmask = cmap(it.cmask);
                it.qmask = cmap(it.qmask);
            }
        }
        for (int q = 0; q < n && popc(pi.T) < k; ++q) pi.T |= 1ULL << q;
        FPass fp{};
        fp.k = k;
        fp.next = n - k;
        std::vector<int> tq, epos;
        for (int q = 0; q < n; ++q) ((pi.T >> q) & 1 ? tq : epos).push_back(q);
        // Tile index bit 0 is the lowest external qubit (the partner of a paired
        // tile, adjacent runs); bits 1.. take the others from the highest down, so
        // workgroups running at the same time (consecutive tile indices) work far
        // apart in the state. Measured on 30-qubit complex128 passes against the
        // ascending order: QFT 6.97 -> 6.86 ms, layered circuit 7.50 -> 7.37 ms per
        // pass; interleaving low and high qubits or a rotation was slower (7.27 and
        // 7.13 ms on QFT). Only for states of 27 qubits or more: on 24-qubit
        // complex128 states the far-apart order was slower (QAOA-24 adjoint
        // un-application passes 0.10 -> 0.15 ms; 64 gradients 0.66 -> 0.79 s).
        // Why is not established; it is not reuse of the Infinity Cache between
        // passes (walking the tiles backwards on every other pass, which would
        // find what such a cache kept of the previous pass, changed nothing at
        // 24-25 qubits).
        if (epos.size() > 2 && n >= TILE_ORDER_FAR_MIN_QUBITS) std::reverse(epos.begin() + 1, epos.end());
        std::vector<int> loc(64, -1);
        for (int b = 0; b < k; ++b) {
            fp.tq[b] = (uint32_t)tq[b];
            loc[tq[b]] = b;
        }
        for (size_t e = 0; e < epos.size(); ++e) fp.epos[e] = (uint32_t)epos[e];

        std::vector<SStage> stages(1);
        std::vector<Mono> pending;
        auto flush = [&](SStage& st, int lt_or_all) {
            std::vector<Mono> take, keep;
            for (auto& m : pending) {
                if (lt_or_all < 0 || ((m.lmask >> lt_or_all) & 1)) take.push_back(m);
                else keep.push_back(m);
            }
            pending.swap(keep);
            if (take.empty()) return;
            SOp op;
            op.mono = true;
            op.monos = std::move(take);
            st.ops.push_back(std::move(op));
        };
        for (const Item& it : pi.items) {
            if (it.t == IT_PROBE && !probe_nd(it)) {
                // diagonal probe: commutes with the pending monomials, evaluated in the current stage
                SOp op;
                op.mono = false;
                op.dtype = FOP_PROBE;
                op.pkind = it.pkind;
                op.pid = it.pid;
                op.lt = -1;
                op.lctl = 0;
                for (int q = 0; q < n; ++q)
                    if (((it.qmask >> q) & 1) && loc[q] >= 0) op.lctl |= 1ULL << loc[q];
                op.ectl = it.qmask & ~pi.T;
                stages.back().ops.push_back(op);
            } else if (it.t == IT_MONO) {
                uint64_t lm = 0;
                for (int q = 0; q < n; ++q)
                    if (((it.qmask >> q) & 1) && loc[q] >= 0) lm |= 1ULL << loc[q];
                pending.push_back({lm, it.qmask & ~pi.T, it.d});
            } else if (it.t == IT_SWAP) {
                std::swap(loc[it.q], loc[it.q2]);
            } else {
                const int lt = loc[it.q];
                SStage* st = &stages.back();
                if (std::find(st->R.begin(), st->R.end(), lt) == st->R.end()) {
                    if ((int)st->R.size() == FUSE_REG) {
                        stages.emplace_back();
                        st = &stages.back();
                    }
                    st->R.push_back(lt);
                }
                flush(*st, lt);
                SOp op;
                op.mono = false;
                op.dtype = it.t == IT_PROBE ? (uint32_t)FOP_PROBE : it.dtype;
                op.pkind = it.pkind;
                op.pid = it.pid;
                op.lt = lt;
                op.lctl = 0;
                op.ectl = 0;
                for (int q = 0; q < n; ++q)
                    if ((it.cmask >> q) & 1) {
                        if (loc[q] >= 0) op.lctl |= 1ULL << loc[q];
                        else op.ectl |= 1ULL << q;
                    }
                for (int j = 0; j < 4; ++j) op.m[j] = it.m[j];
                st->ops.push_back(op);
            }
        }
        flush(stages.back(), -1);
        // store maps: destination qubit of each source qubit
        int dest[64];
        for (int q = 0; q < n; ++q) dest[q] = q;
        bool oop = false;
        if (&pi == &passes.back()) {
            // the last pass restores the canonical order: logical qubit q sits at
            // cur[phys[q]] (phys is the identity without relabelling)
            for (int q = 0; q < n; ++q) dest[cur[phys[q]]] = q;
            for (int q = 0; q < n; ++q) oop |= dest[q] != q;
        } else if (contig && (pi.T & lowT) == lowT) {
            // the next pass's tile qubits go to positions L.., after the low
            // qubits 0..L-1 (tile qubits of this pass and of the next: this
            // pass's writes keep their runs of 2^L amplitudes), then the others;
            // each group in its present order
            const uint64_t Tn = cmap((&pi + 1)->T) & ~lowT;
            int perm[64], np = L;
            for (int q = 0; q < L; ++q) perm[q] = q;
            for (int q = L; q < n; ++q)
                if ((Tn >> q) & 1) perm[q] = np++;
            for (int q = L; q < n; ++q)
                if (!((Tn >> q) & 1)) perm[q] = np++;
            for (int q = 0; q < n; ++q) oop |= perm[q] != q;
            if (oop) {
                for (int q = 0; q < n; ++q) dest[q] = perm[q];
                for (int q = 0; q < n; ++q) cur[q] = perm[cur[q]];
            }
        }
        fp.out_of_place = oop ? 1 : 0;
        {
            // store position c: tile qubits ordered by destination qubit
            std::vector<std::pair<int, int>> byd;  // (dest qubit, local bit)
            for (int b = 0; b < k; ++b) byd.push_back({dest[tq[b]], loc[tq[b]]});
            std::sort(byd.begin(), byd.end());
            for (int c = 0; c < k; ++c) {
                fp.so_q[c] = (uint32_t)byd[c].first;
                fp.so_l[c] = (uint32_t)byd[c].second;
            }
            for (size_t e = 0; e < epos.size(); ++e) fp.eout[e] = (uint32_t)dest[epos[e]];
        }

        // ---- emit stages ----
        fp.stage_begin = (uint32_t)P.stages.size();
        fp.slot_begin = (uint32_t)P.slots.size();
        fp.probe_base = (int32_t)P.probe_ids.size();
        int nslots = 0, nprobes = 0;
        for (SStage& st : stages) {
            // complete the register set with unused local bits
            for (int l = 0; l < k && (int)st.R.size() < FUSE_REG; ++l)
                if (std::find(st.R.begin(), st.R.end(), l) == st.R.end()) st.R.push_back(l);
            int pos[64];
            for (int l = 0; l < 64; ++l) pos[l] = -1;
            uint64_t Rmask = 0;
            for (int p = 0; p < FUSE_REG; ++p) {
                pos[st.R[p]] = p;
                Rmask |= 1ULL << st.R[p];
            }
            // thread bits: first up to 4 with distinct (bit mod 4) for conflict-free LDS banks
            std::vector<int> rest, tm;
            for (int l = 0; l < k; ++l)
                if (pos[l] < 0) rest.push_back(l);
            int used_res = 0;
            for (int l : rest)
                if ((int)tm.size() < 4 && !((used_res >> (l & 3)) & 1)) {
                    tm.push_back(l);
                    used_res |= 1 << (l & 3);
                }
            for (int l : rest)
                if (std::find(tm.begin(), tm.end(), l) == tm.end()) tm.push_back(l);
            FStage fs{};
            for (int p = 0; p < FUSE_REG; ++p) fs.rmap[p] = (uint32_t)st.R[p];
            for (size_t t = 0; t < tm.size(); ++t) fs.tmap[t] = (uint32_t)tm[t];
            fs.op_begin = (uint32_t)P.ops.size();
            for (SOp& so : st.ops) {
                if (!so.mono && so.dtype == FOP_PROBE) {
                    FOp f{};
                    f.type = FOP_PROBE;
                    f.mode = so.pkind;
                    if (so.lt >= 0) {
                        f.p = (uint32_t)pos[so.lt];
                    } else {
                        uint32_t rm = 0;
                        for (int l = 0; l < k; ++l)
                            if (((so.lctl >> l) & 1) && pos[l] >= 0) rm |= 1u << pos[l];
                        f.rctl = rm;
                        f.tctl = (uint32_t)(so.lctl & ~Rmask);
                        f.ectl = so.ectl;
                    }
                    f.term_begin = (uint32_t)nprobes++;
                    P.probe_ids.push_back(so.pid);
                    P.ops.push_back(f);
                    continue;
                }
                if (!so.mono) {
                    FOp f{};
                    f.type = so.dtype;
                    if (so.lctl != 0 || so.ectl != 0) {
                        // controlled: the general formula (controls evaluated per pair/thread)
                        f.type = (so.dtype == FOP_XSWAP) ? FOP_XSWAP : FOP_DENSE;
                    } else if (so.dtype == FOP_DENSE) {
                        auto im0 = [](cd z) { return z.imag() == 0.0; };
                        auto re0 = [](cd z) { return z.real() == 0.0; };
                        if (im0(so.m[0]) && im0(so.m[1]) && im0(so.m[2]) && im0(so.m[3])) f.type = FOP_REAL;
                        else if (im0(so.m[0]) && im0(so.m[3]) && re0(so.m[1]) && re0(so.m[2])) f.type = FOP_RXLIKE;
                    }
                    f.p = (uint32_t)pos[so.lt];
                    uint32_t rctl = 0;
                    for (int l = 0; l < k; ++l)
                        if (((so.lctl >> l) & 1) && pos[l] >= 0) rctl |= 1u << pos[l];
                    f.rctl = (uint32_t)rctl;
                    f.tctl = (uint32_t)(so.lctl & ~Rmask);
                    f.ectl = so.ectl;
                    for (int j = 0; j < 4; ++j) {
                        f.m[2 * j] = so.m[j].real();
                        f.m[2 * j + 1] = so.m[j].imag();
                    }
                    P.ops.push_back(f);
                    continue;
                }
                // group monomials by register mask
                std::map<uint32_t, std::vector<const Mono*>> byR;
                for (const Mono& m : so.monos) {
                    uint32_t mr = 0;
                    for (int l = 0; l < k; ++l)
                        if (((m.lmask >> l) & 1) && pos[l] >= 0) mr |= 1u << pos[l];
                    byR[mr].push_back(&m);
                }
                for (auto& kv : byR) {
                    FOp f{};
                    f.type = FOP_MONO;
                    f.rctl = (uint32_t)kv.first;
                    f.term_begin = (uint32_t)P.terms.size();
                    std::map<uint32_t, cd> cst;                         // tmask -> product
                    std::map<uint32_t, std::vector<const Mono*>> ext;  // tmask -> external monomials
                    for (const Mono* m : kv.second) {
                        const uint32_t tmask = (uint32_t)(m->lmask & ~Rmask);
                        if (m->emask == 0) {
                            auto itc = cst.find(tmask);
                            if (itc == cst.end()) cst[tmask] = m->d;
                            else itc->second *= m->d;
                        } else {
                            ext[tmask].push_back(m);
                        }
                    }
                    // Constant single-thread-bit terms become a per-thread-bit factor
                    // vector (one select + multiply per set bit of the thread id), the
                    // tmask-0 constant goes to m[0..1]; the rest stay terms.
                    f.mode = 1;
                    f.alpha_begin = FUSE_NO_ALPHA;
                    std::vector<cd> alpha(k - FUSE_REG, cd(1.0, 0.0));
                    uint32_t amask = 0;
                    cd c0(1.0, 0.0);
                    for (auto& c : cst) {
                        const uint32_t tmask = c.first;
                        if (tmask == 0) {
                            c0 *= c.second;
                            continue;
                        }
                        if (popc(tmask) == 1) {
                            const int l = __builtin_ctz(tmask);
                            const int t = (int)(std::find(tm.begin(), tm.end(), l) - tm.begin());
                            alpha[t] *= c.second;
                            amask |= 1u << t;
                            continue;
                        }
                        FTerm t{};
                        t.tmask = tmask;
                        t.slot = -1;
                        t.emask = 0;
                        t.d[0] = c.second.real();
                        t.d[1] = c.second.imag();
                        t.ang = std::arg(c.second);
                        P.terms.push_back(t);
                    }
                    f.m[0] = c0.real();
                    f.m[1] = c0.imag();
                    f.tctl = amask;  // MONO: thread-id bits with a factor
                    if (amask) {
                        f.alpha_begin = (uint32_t)P.alphas.size();
                        for (const cd& a : alpha) {
                            P.alphas.push_back(a.real());
                            P.alphas.push_back(a.imag());
                        }
                    }
                    for (auto& e : ext) {
                        if (nslots < FUSE_MAX_SLOTS) {
                            FSlot sl;
                            sl.begin = (uint32_t)P.eterms.size();
                            for (const Mono* m : e.second) {
                                FExtTerm et;
                                et.emask = m->emask;
                                et.ang = std::arg(m->d);
                                P.eterms.push_back(et);
                            }
                            sl.end = (uint32_t)P.eterms.size();
                            P.slots.push_back(sl);
                            FTerm t{};
                            t.tmask = e.first;
                            t.slot = nslots++;
                            t.emask = 0;
                            t.d[0] = 1.0;
                            t.d[1] = 0.0;
                            P.terms.push_back(t);
                        } else {
                            for (const Mono* m : e.second) {
                                FTerm t{};
                                t.tmask = e.first;
                                t.slot = -1;
                                t.emask = m->emask;
                                t.d[0] = m->d.real();
                                t.d[1] = m->d.imag();
                                t.ang = std::arg(m->d);
                                P.terms.push_back(t);
                            }
                        }
                    }
                    f.term_end = (uint32_t)P.terms.size();
                    P.ops.push_back(f);
                }
            }
            fs.op_end = (uint32_t)P.ops.size();
            P.stages.push_back(fs);
        }
        fp.nstages = (int)stages.size();
        fp.nslots = nslots;
        fp.nprobes = nprobes;
        P.passes.push_back(fp);
    }
    return QSE_OK;
}

}  // namespace

int build_plan(int n, int k, int L, const qse_op* ops, int nops, const double* params, HostPlan& P, int flags) {
    std::vector<Item> items;
    to_items(ops, nops, params, items);
    return plan_items(n, k, L, items, P, flags);
}


qse_op dagger(const qse_op& o, const double* params);

int build_reverse_plan(int n, int k, int L, const qse_op* ops, int nops, const double* params,
                       std::vector<HostPlan>& chunks, std::vector<ProbeRef>& refs) {
    chunks.clear();
    refs.clear();
    int kmin = -1;
    for (int i = 0; i < nops; ++i)
        if (ops[i].param >= 0) {
            kmin = i;
            break;
        }
    if (kmin < 0) return QSE_OK;
    std::vector<Item> items;
    auto probe = [&](uint32_t kind, int q, uint64_t qmask, int param, double coeff) {
        Item it;
        it.t = IT_PROBE;
        it.pkind = kind;
        it.q = q;
        it.qmask = qmask;
        it.pid = (int)refs.size();
        refs.push_back({param, coeff});
        items.push_back(it);
    };
    auto unapply = [&](const qse_op* seq, int cnt) {  // seq: ops in application order
        std::vector<qse_op> d;
        for (int i = 0; i < cnt; ++i) d.push_back(dagger(seq[i], params));
        to_items(d.data(), cnt, nullptr, items);
    };
    for (int i = nops - 1; i >= kmin; --i) {
        const qse_op& o = ops[i];
        if (i >= 2 && is_zz_triple(ops, nops, i - 2) && ops[i].param < 0 && ops[i - 2].param < 0) {
            // CNOT RZ CNOT = exp(-i theta Z_c Z_t / 2): probed after the triple
            const qse_op& rz = ops[i - 1];
            if (rz.param >= 0) probe(PROBE_ZSIGN, -1, (1ULL << o.control) | (1ULL << o.target), rz.param, rz.scale);
            if (i - 2 > kmin) {
                const qse_op rev[3] = {ops[i], ops[i - 1], ops[i - 2]};
                unapply(rev, 3);
            }
            i -= 2;
            continue;
        }
        if (o.param >= 0) {
            const uint64_t tb = 1ULL << o.target;
            switch (o.kind) {
                case QSE_GATE_RX: probe(PROBE_X, o.target, 0, o.param, o.scale); break;
                case QSE_GATE_RY: probe(PROBE_Y, o.target, 0, o.param, o.scale); break;
                case QSE_GATE_RZ: probe(PROBE_ZSIGN, -1, tb, o.param, o.scale); break;
                // PHASE and CPHASE = exp(i theta P): dE/dtheta = -2 Im<lambda|P|psi>, P the projector
                case QSE_GATE_PHASE: probe(PROBE_PROJ, -1, tb, o.param, -2.0 * o.scale); break;
                case QSE_GATE_CPHASE:
                    probe(PROBE_PROJ, -1, tb | (1ULL << o.control), o.param, -2.0 * o.scale);
                    break;
                default:
                    return set_error(QSE_ERR_UNSUPPORTED,
                                     "adjoint: parameters are supported on RX, RY, RZ, PHASE and CPHASE only");
            }
        }
        if (i > kmin) unapply(&o, 1);
    }
    // at most FUSE_MAX_PROBES accumulators per pass: split the sweep where needed
    std::vector<std::vector<Item>> parts(1);
    {
        HostPlan W;
        int rc = plan_items(n, k, L, items, W, PLAN_DAG);
        if (rc) return rc;
        bool fits = true;
        for (const FPass& fp : W.passes) fits &= fp.nprobes <= FUSE_MAX_PROBES;
        if (fits) {
            chunks.push_back(std::move(W));
            return QSE_OK;
        }
        int cnt = 0;
        for (const Item& it : items) {
            if (it.t == IT_PROBE && cnt == FUSE_MAX_PROBES) {
                parts.emplace_back();
                cnt = 0;
            }
            cnt += it.t == IT_PROBE ? 1 : 0;
            parts.back().push_back(it);
        }
    }
    for (const auto& part : parts) {
        HostPlan C;
        const int rc = plan_items(n, k, L, part, C, PLAN_DAG);
        if (rc) return rc;
        chunks.push_back(std::move(C));
    }
    return QSE_OK;
}

bool plan_out_of_place(const HostPlan& P) {
    for (const FPass& fp : P.passes)
        if (fp.out_of_place) return true;
    return false;
}

// Smallest state (qubits) for which the planner tries PLAN_CONTIG: the states
// where it was measured (27-33 qubits). Smaller states keep in-place plans and
// need no second buffer.
constexpr int CONTIG_MIN_QUBITS = 27;

// The best of the planner variants: fewest passes, then fewest stages, then
// in place (no second buffer). allow_oop = false keeps to in-place plans.
// From CONTIG_MIN_QUBITS on, the best plan is built again with PLAN_CONTIG:
// every pass but the last stores the state in the order that makes the next
// pass's tile the contiguous low qubits, so that pass reads whole
// 2^k-amplitude runs and, with 4 or more stages, loads its first stage and
// stores its last straight from registers (the pass before pays with writes in
// runs of 2^L amplitudes). Measured on the 30-qubit random layered circuit
// (14 passes of 1-9 stages): 120 -> 112 ms, the passes of 8-9 stages 10.8 ->
// 8.3 and 13.3 -> 11.5 ms; QFT-30 (3 passes of 3 stages) unchanged within the
// box-to-box spread (20.0-20.2 ms per QFT either way). The rebuilt plan is
// taken when it has as many passes and no more stages.
int choose_plan(int n, int k, int L, const qse_op* ops, int nops, const double* params, HostPlan& P,
                bool allow_oop) {
    bool have = false;
    long best_cost = 0;
    bool best_oop = false;
    int best_flags = 0;
    for (int flags : {int(PLAN_DAG), 0, int(PLAN_RELABEL | PLAN_DAG), int(PLAN_RELABEL), int(PLAN_DAG | PLAN_LOOKAHEAD),
                      int(PLAN_RELABEL | PLAN_DAG | PLAN_LOOKAHEAD)}) {
        if (!allow_oop && (flags & PLAN_RELABEL)) continue;
        HostPlan C;
        const int rc = build_plan(n, k, L, ops, nops, params, C, flags);
        if (rc == QSE_ERR_UNSUPPORTED) continue;
        if (rc) return rc;
        const bool oop = plan_out_of_place(C);
        const long cost = (long)C.passes.size() * 100000L + (long)C.stages.size();
        if (!have || cost < best_cost || (cost == best_cost && best_oop && !oop)) {
            P = std::move(C);
            best_cost = cost;
            best_oop = oop;
            best_flags = flags;
            have = true;
        }
    }
    if (!have) return set_error(QSE_ERR_UNSUPPORTED, "no fusion plan for this gate list");
    if (allow_oop && n >= CONTIG_MIN_QUBITS && P.passes.size() > 1) {
        HostPlan C;
        const int rc = build_plan(n, k, L, ops, nops, params, C, best_flags | PLAN_CONTIG);
        if (rc == QSE_OK && C.passes.size() == P.passes.size() && C.stages.size() <= P.stages.size())
            P = std::move(C);
        else if (rc && rc != QSE_ERR_UNSUPPORTED)
            return rc;
    }
    return QSE_OK;
}


// ---- tile shape tuner ----
// Time of one pass relative to the fastest shape of the precision, measured on
// MI355X (30-qubit QFT and random layered circuits, run-time specialised
// passes): short contiguous runs (small L) cost HBM efficiency, 13-qubit
// complex128 tiles leave one workgroup per CU (its stages then do not overlap
// another workgroup's memory traffic), complex64 runs below 16 amplitudes
// (128 bytes) lose most of the bandwidth and are not candidates.
struct ShapeCost {
    int k, L;
    double cost;
};
const ShapeCost kShapesC128[] = {{12, 3, 1.10}, {12, 4, 1.04}, {12, 5, 1.02}, {12, 6, 1.00},
                                 {11, 3, 1.12}, {11, 4, 1.04}, {11, 5, 1.015}, {11, 6, 1.00},
                                 {13, 3, 1.50}, {13, 4, 1.43}, {13, 5, 1.30}, {13, 6, 1.20}};
const ShapeCost kShapesC64[] = {{13, 4, 1.24}, {13, 5, 1.22}, {13, 6, 1.10},
                                {12, 4, 1.11}, {12, 5, 1.01}, {12, 6, 1.00}};

// The plan run_fused executes: with a tile size set on the context, that
// shape; otherwise the candidate shape of least estimated time (passes x
// relative pass cost, stages as a tie-break).
std::vector<char> gate_structure_key(int n, int prec, const qse_op* ops, int nops) {
    const int head[2] = {n, prec};
    std::vector<char> key(sizeof(head) + sizeof(qse_op) * (size_t)nops);
    std::memcpy(key.data(), head, sizeof(head));
    char* w = key.data() + sizeof(head);
    for (int i = 0; i < nops; ++i) {
        qse_op o = ops[i];
        o.theta = 0.0;
        o.scale = 0.0;
        std::fill(o.m, o.m + 8, 0.0);
        std::memcpy(w + sizeof(qse_op) * (size_t)i, &o, sizeof(qse_op));
    }
    return key;
}

const ShapeEntry* find_shape(const qse_ctx* ctx, const std::vector<char>& key) {
    if (!ctx) return nullptr;
    for (const ShapeEntry& e : ctx->shape_cache)
        if (e.key == key) return &e;
    return nullptr;
}

int rank_shapes(int n, int prec, const qse_op* ops, int nops, const double* params, bool allow_oop,
                std::vector<ShapeCandidate>& out) {
    out.clear();
    const ShapeCost* shapes = (prec == QSE_C128) ? kShapesC128 : kShapesC64;
    const int ns = (prec == QSE_C128) ? (int)(sizeof(kShapesC128) / sizeof(ShapeCost))
                                      : (int)(sizeof(kShapesC64) / sizeof(ShapeCost));
    for (int i = 0; i < ns; ++i) {
        const int k = shapes[i].k, L = shapes[i].L;
        if (k > n - 2 || L > k - FUSE_REG) continue;
        HostPlan C;
        const int rc = choose_plan(n, k, L, ops, nops, params, C, allow_oop);
        if (rc) return rc;
        out.push_back({k, L, (int)C.passes.size(),
                       (double)C.passes.size() * shapes[i].cost + 0.001 * (double)C.stages.size()});
    }
    std::stable_sort(out.begin(), out.end(),
                     [](const ShapeCandidate& a, const ShapeCandidate& b) { return a.est < b.est; });
    return QSE_OK;
}

void remember_shape(qse_ctx* ctx, std::vector<char> key, int k, int L) {
    for (ShapeEntry& e : ctx->shape_cache)
        if (e.key == key) {
            e.k = k;
            e.L = L;
            e.stamp = ++ctx->prog_clock;
            return;
        }
    if (ctx->shape_cache.size() >= 256) {
        size_t oldest = 0;
        for (size_t i = 1; i < ctx->shape_cache.size(); ++i)
            if (ctx->shape_cache[i].stamp < ctx->shape_cache[oldest].stamp) oldest = i;
        ctx->shape_cache.erase(ctx->shape_cache.begin() + oldest);
    }
    ShapeEntry e;
    e.key = std::move(key);
    e.k = k;
    e.L = L;
    e.stamp = ++ctx->prog_clock;
    ctx->shape_cache.push_back(std::move(e));
}

int tuned_plan(const qse_ctx* ctx, int n, int prec, const qse_op* ops, int nops, const double* params, HostPlan& P,
               bool allow_oop, int* k_out, int* L_out) {
    int k0, L0;
    fusion_shape(ctx, n, prec, &k0, &L0);
    const bool fixed = ctx && ctx->tile_bits[prec] > 0;
    if (!fixed && ctx && !ctx->shape_cache.empty()) {
        // a shape measured by qse_autotune (or tuned before) for this gate-list structure
        const ShapeEntry* e = find_shape(ctx, gate_structure_key(n, prec, ops, nops));
        if (e && e->k <= n && e->L <= e->k - FUSE_REG) {
            *k_out = e->k;
            *L_out = e->L;
            return choose_plan(n, e->k, e->L, ops, nops, params, P, allow_oop);
        }
    }
    if (!fixed) {
        const ShapeCost* shapes = (prec == QSE_C128) ? kShapesC128 : kShapesC64;
        const int ns = (prec == QSE_C128) ? (int)(sizeof(kShapesC128) / sizeof(ShapeCost))
                                          : (int)(sizeof(kShapesC64) / sizeof(ShapeCost));
        bool have = false;
        double best = 0;
        for (int i = 0; i < ns; ++i) {
            const int k = shapes[i].k, L = shapes[i].L;
            if (k > n - 2 || L > k - FUSE_REG) continue;
            HostPlan C;
            const int rc = choose_plan(n, k, L, ops, nops, params, C, allow_oop);
            if (rc) return rc;
            const double cost = (double)C.passes.size() * shapes[i].cost + 0.001 * (double)C.stages.size();
            if (!have || cost < best) {
                P = std::move(C);
                best = cost;
                have = true;
                *k_out = k;
                *L_out = L;
            }
        }
        if (have) return QSE_OK;
    }
    *k_out = k0;
    *L_out = L0;
    return choose_plan(n, k0, L0, ops, nops, params, P, allow_oop);
}

static size_t al256(size_t v) { return (v + 255) & ~(size_t)255; }

void serialize(const HostPlan& P, Serialized& S) {
    const size_t sz[7] = {sizeof(FPass) * P.passes.size(), sizeof(FStage) * P.stages.size(),
                          sizeof(FOp) * P.ops.size(),       sizeof(FTerm) * P.terms.size(),
                          sizeof(FSlot) * P.slots.size(),   sizeof(FExtTerm) * P.eterms.size(),
                          sizeof(double) * P.alphas.size()};
    const void* src[7] = {P.passes.data(), P.stages.data(), P.ops.data(),   P.terms.data(),
                          P.slots.data(),  P.eterms.data(), P.alphas.data()};
    size_t tot = 0;
    for (int i = 0; i < 7; ++i) {
        S.off[i] = tot;
        tot += al256(std::max<size_t>(sz[i], 1));
    }
    S.bytes.assign(tot, 0);
    for (int i = 0; i < 7; ++i)
        if (sz[i]) std::memcpy(S.bytes.data() + S.off[i], src[i], sz[i]);
}

std::vector<char> structure_key(const HostPlan& P) {
    HostPlan Z = P;
    for (FOp& o : Z.ops) std::fill(o.m, o.m + 8, 0.0);
    for (FTerm& t : Z.terms) t.d[0] = t.d[1] = t.ang = 0.0;
    for (FExtTerm& e : Z.eterms) e.ang = 0.0;
    std::fill(Z.alphas.begin(), Z.alphas.end(), 0.0);
    Serialized S;
    serialize(Z, S);
    return std::move(S.bytes);
}

}  // namespace qse
