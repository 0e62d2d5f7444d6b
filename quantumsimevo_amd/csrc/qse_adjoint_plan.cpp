// qse_adjoint_plan.cpp — generator terms, commuting blocks and reverse-sweep
// segments of the adjoint method (see qse_adjoint_plan.hpp).
//
// Generators (every parameterised kind has two eigenvalues one apart):
//   RX/RY/RZ = exp(-i theta G/2):  dE/dtheta = Im<lambda|G|psi>,  G = X, Y, Z
//   CNOT(c,t) RZ(theta,t) CNOT(c,t) (the QAOA cost gate, QAOA.cpp:128-133): G = Z_c Z_t
//   PHASE = diag(1, e^{i theta}):  dE/dtheta = -2 Im<lambda|P1|psi> = -Im<lambda|psi> + Im<lambda|Z|psi>
//   CPHASE: -2 Im<lambda|P11|psi> = -(Im<l|psi> - Im<l|Z_c|psi> - Im<l|Z_t|psi> + Im<l|Z_cZ_t|psi>)/2
#include "qse_adjoint_plan.hpp"

#include <string>

namespace qse {

int set_error(int status, const std::string& msg);
double op_angle(const qse_op& op, const double* params);
bool is_zz_triple(const qse_op* ops, int nops, int k);

int build_items(const qse_op* ops, int nops, std::vector<AdjItem>& items) {
    for (int k = 0; k < nops; ++k) {
        const qse_op& o = ops[k];
        AdjItem it;
        it.first = k;
        it.count = 1;
        const uint64_t tb = 1ULL << o.target;
        const double sc = o.scale;
        if (is_zz_triple(ops, nops, k)) {
            const qse_op& rz = ops[k + 1];
            it.count = 3;
            it.dg = tb | (1ULL << o.control);
            if (rz.param >= 0) it.diag.push_back({it.dg, {rz.param, rz.scale}});
            items.push_back(it);
            k += 2;
            continue;
        }
        switch (o.kind) {
            case QSE_GATE_IDENTITY: continue;
            case QSE_GATE_H: case QSE_GATE_X: case QSE_GATE_Y: case QSE_GATE_MATRIX: it.nd = tb; break;
            case QSE_GATE_RX: it.nd = tb; break;
            case QSE_GATE_RY: it.nd = tb; break;
            case QSE_GATE_Z: case QSE_GATE_PHASE: case QSE_GATE_RZ: it.dg = tb; break;
            case QSE_GATE_CNOT: it.nd = tb; it.dg = 1ULL << o.control; break;
            case QSE_GATE_CPHASE: it.dg = tb | (1ULL << o.control); break;
            case QSE_GATE_MCPHASE: it.dg = tb | o.controls; break;
            case QSE_GATE_SWAP: it.nd = tb | (1ULL << o.control); break;
            default: return set_error(QSE_ERR_UNSUPPORTED, "adjoint: unsupported gate kind");
        }
        if (o.param >= 0) {
            switch (o.kind) {
                case QSE_GATE_RX: it.pair_q = o.target; it.pair_type = 0; it.pair = {o.param, sc}; break;
                case QSE_GATE_RY: it.pair_q = o.target; it.pair_type = 1; it.pair = {o.param, sc}; break;
                case QSE_GATE_RZ: it.diag.push_back({tb, {o.param, sc}}); break;
                case QSE_GATE_PHASE:
                    it.diag.push_back({0, {o.param, -sc}});
                    it.diag.push_back({tb, {o.param, sc}});
                    break;
                case QSE_GATE_CPHASE: {
                    if (o.control == o.target) {
                        it.diag.push_back({0, {o.param, -sc}});
                        it.diag.push_back({tb, {o.param, sc}});
                    } else {
                        const uint64_t cb = 1ULL << o.control;
                        it.diag.push_back({0, {o.param, -0.5 * sc}});
                        it.diag.push_back({cb, {o.param, 0.5 * sc}});
                        it.diag.push_back({tb, {o.param, 0.5 * sc}});
                        it.diag.push_back({cb | tb, {o.param, -0.5 * sc}});
                    }
                    break;
                }
                default:
                    return set_error(QSE_ERR_UNSUPPORTED,
                                     "adjoint: parameters are supported on RX, RY, RZ, PHASE and CPHASE only");
            }
        }
        items.push_back(it);
    }
    return QSE_OK;
}

void build_blocks(const std::vector<AdjItem>& items, std::vector<std::vector<int>>& blocks) {
    // A block is a set of pairwise-commuting items of the remaining circuit each
    // of which commutes with every later remaining item outside the block: all
    // can be moved to the end together. Scanning from the end, an item joins
    // when it commutes with the block and with the items passed over so far
    // (conservatively through their qubit-mask unions); the scan stops once the
    // passed-over items touch every qubit non-diagonally. Blocks without
    // parameters cost nothing in the reverse sweep (they merge into the next
    // un-application), so each round first takes a block of parameter-free
    // items only, and a block with parameterised items only when no
    // parameter-free item can be moved to the end: a layer of rotations is then
    // one block instead of being split among the gates around it.
    auto commute = [](uint64_t nd, uint64_t dg, uint64_t ndS, uint64_t dgS) {
        return (nd & (ndS | dgS)) == 0 && (dg & ndS) == 0;
    };
    auto has_terms = [](const AdjItem& it) { return !it.diag.empty() || it.pair_q >= 0; };
    uint64_t all = 0;
    for (const AdjItem& it : items) all |= it.nd | it.dg;
    std::vector<int> rem(items.size());
    for (size_t i = 0; i < items.size(); ++i) rem[i] = (int)i;
    while (!rem.empty()) {
        std::vector<int> blk, passed;  // both from the end
        int i = -1;
        for (int pass = 0; pass < 2 && blk.empty(); ++pass) {
            const bool free_only = pass == 0;
            blk.clear();
            passed.clear();
            uint64_t ndB = 0, dgB = 0, ndP = 0, dgP = 0;
            for (i = (int)rem.size() - 1; i >= 0; --i) {
                const AdjItem& it = items[rem[i]];
                if ((!free_only || !has_terms(it)) && commute(it.nd, it.dg, ndB, dgB) &&
                    commute(it.nd, it.dg, ndP, dgP)) {
                    blk.push_back(rem[i]);
                    ndB |= it.nd;
                    dgB |= it.dg;
                    continue;
                }
                passed.push_back(rem[i]);
                ndP |= it.nd;
                dgP |= it.dg;
                if ((ndP & all) == all) {  // every qubit has a passed-over non-diagonal item: none can join
                    --i;
                    break;
                }
            }
        }
        // remaining circuit: the unvisited items rem[0..i], then the passed-over ones in circuit order
        std::vector<int> next(rem.begin(), rem.begin() + (i + 1));
        next.insert(next.end(), passed.rbegin(), passed.rend());
        blocks.push_back(std::move(blk));  // items from the end, as the sweep un-applies them
        rem.swap(next);
    }
}

qse_op dagger(const qse_op& o, const double* params) {
    qse_op d = o;
    d.param = -1;
    d.scale = 0.0;
    switch (o.kind) {
        case QSE_GATE_PHASE: case QSE_GATE_RX: case QSE_GATE_RY: case QSE_GATE_RZ: case QSE_GATE_CPHASE:
        case QSE_GATE_MCPHASE:
            d.theta = -op_angle(o, params);
            break;
        case QSE_GATE_MATRIX:
            // conjugate transpose
            d.m[0] = o.m[0]; d.m[1] = -o.m[1];
            d.m[2] = o.m[4]; d.m[3] = -o.m[5];
            d.m[4] = o.m[2]; d.m[5] = -o.m[3];
            d.m[6] = o.m[6]; d.m[7] = -o.m[7];
            break;
        default: break;  // H, X, Y, Z, CNOT, SWAP: self-inverse
    }
    return d;
}

namespace {

qse_op gate1(int kind, int q, double theta) {
    qse_op o{};
    o.kind = kind;
    o.target = q;
    o.control = -1;
    o.param = -1;
    o.theta = theta;
    return o;
}

}  // namespace

int adjoint_segments(const qse_op* ops, int nops, const double* params, std::vector<AdjSegment>& segs,
                     std::vector<qse_op>& prefix, bool pair_mode) {
    std::vector<AdjItem> items;
    const int rc = build_items(ops, nops, items);
    if (rc) return rc;
    std::vector<std::vector<int>> blocks;
    build_blocks(items, blocks);
    struct Pair {
        int q, type;
        GradTerm t;
    };
    std::vector<std::vector<Pair>> pairs(1);
    segs.assign(1, AdjSegment());
    for (const std::vector<int>& blk : blocks) {
        bool terms = false;
        for (int idx : blk) terms |= !items[idx].diag.empty() || items[idx].pair_q >= 0;
        if (terms) {
            // the segment's evaluation point must be this block's output
            if (!segs.back().unapply.empty()) {
                segs.emplace_back();
                pairs.emplace_back();
            }
            AdjSegment& s = segs.back();
            for (int idx : blk) {
                const AdjItem& it = items[idx];
                for (const auto& d : it.diag) s.diag.push_back(d);
                if (it.pair_q >= 0) pairs.back().push_back({it.pair_q, it.pair_type, it.pair});
            }
        }
        // un-apply: blk lists items from the end; each item's ops reversed
        AdjSegment& s = segs.back();
        for (int idx : blk) {
            const AdjItem& it = items[idx];
            for (int k = it.count - 1; k >= 0; --k) s.unapply.push_back(dagger(ops[it.first + k], params));
        }
        // forward order: this block precedes the blocks already in the segment
        std::vector<qse_op> fwd;
        for (int i = (int)blk.size() - 1; i >= 0; --i) {
            const AdjItem& it = items[blk[i]];
            for (int k = 0; k < it.count; ++k) fwd.push_back(ops[it.first + k]);
        }
        s.forward.insert(s.forward.begin(), fwd.begin(), fwd.end());
    }
    // gates before the first parameterised block are never un-applied
    AdjSegment& last = segs.back();
    if (last.diag.empty() && pairs.back().empty()) {
        prefix = last.forward;
        segs.pop_back();
        pairs.pop_back();
    } else {
        prefix.clear();
    }
    if (!segs.empty()) {
        prefix.insert(prefix.end(), segs.back().forward.begin(), segs.back().forward.end());
        segs.back().forward.clear();
        segs.back().unapply.clear();
    }
    // X/Y generators -> Z-strings in the frame rotated by V (H for X, H S^dag for Y),
    // or, in pair mode, kept as pair terms of the unrotated state
    const double half_pi = 1.5707963267948966192;
    std::vector<std::vector<qse_op>> vdag(segs.size());
    for (size_t g = 0; g < segs.size(); ++g) {
        uint64_t done = 0;
        for (const Pair& p : pairs[g]) {
            if (pair_mode) {
                segs[g].pair.push_back({p.q, p.type, p.t});
                continue;
            }
            segs[g].diag.push_back({1ULL << p.q, p.t});
            if ((done >> p.q) & 1) continue;  // one rotation per qubit (same-qubit terms share a type)
            done |= 1ULL << p.q;
            if (p.type == 1) {
                segs[g].pre.push_back(gate1(QSE_GATE_PHASE, p.q, -half_pi));
                segs[g].pre.push_back(gate1(QSE_GATE_H, p.q, 0.0));
                vdag[g].push_back(gate1(QSE_GATE_H, p.q, 0.0));
                vdag[g].push_back(gate1(QSE_GATE_PHASE, p.q, half_pi));
            } else {
                segs[g].pre.push_back(gate1(QSE_GATE_H, p.q, 0.0));
                vdag[g].push_back(gate1(QSE_GATE_H, p.q, 0.0));
            }
        }
    }
    for (size_t g = 0; g + 1 < segs.size(); ++g) {
        std::vector<qse_op> u = vdag[g];
        u.insert(u.end(), segs[g].unapply.begin(), segs[g].unapply.end());
        u.insert(u.end(), segs[g + 1].pre.begin(), segs[g + 1].pre.end());
        segs[g].unapply.swap(u);
    }
    return QSE_OK;
}

}  // namespace qse
