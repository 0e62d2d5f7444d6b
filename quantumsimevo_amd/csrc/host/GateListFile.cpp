// GateListFile.cpp — text gate lists (format in include/quantumsimevo/GateListFile.h).
#include "quantumsimevo/GateListFile.h"

#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <stdexcept>

namespace qse_gates {

namespace {

struct KindName {
    int kind;
    const char* name;
};
const KindName kNames[] = {{QSE_GATE_IDENTITY, "I"}, {QSE_GATE_X, "X"},         {QSE_GATE_Y, "Y"},
                           {QSE_GATE_Z, "Z"},        {QSE_GATE_H, "H"},         {QSE_GATE_PHASE, "PHASE"},
                           {QSE_GATE_RX, "RX"},      {QSE_GATE_RY, "RY"},       {QSE_GATE_RZ, "RZ"},
                           {QSE_GATE_CNOT, "CNOT"},  {QSE_GATE_CPHASE, "CPHASE"}, {QSE_GATE_MCPHASE, "MCPHASE"},
                           {QSE_GATE_SWAP, "SWAP"},  {QSE_GATE_MATRIX, "MATRIX"}, {QSE_GATE_MEASURE, "MEASURE"},
                           {QSE_GATE_RESET_QUBIT, "RESET"}};

const char* kind_name(int kind) {
    for (const auto& k : kNames)
        if (k.kind == kind) return k.name;
    throw std::runtime_error("gate list: gate kind " + std::to_string(kind) + " has no text form");
}

std::string num(double v) {
    char b[40];
    std::snprintf(b, sizeof(b), "%.17g", v);
    return b;
}

bool one_qubit_angle(int kind) {
    return kind == QSE_GATE_PHASE || kind == QSE_GATE_RX || kind == QSE_GATE_RY || kind == QSE_GATE_RZ;
}

}  // namespace

std::string write(int n, const qse_op* ops, size_t nops) {
    std::ostringstream o;
    o << "qse-gates 1\nqubits " << n << "\n";
    for (size_t i = 0; i < nops; ++i) {
        const qse_op& g = ops[i];
        o << kind_name(g.kind);
        switch (g.kind) {
            case QSE_GATE_CNOT: case QSE_GATE_SWAP:
                o << ' ' << (g.kind == QSE_GATE_CNOT ? g.control : g.target) << ' '
                  << (g.kind == QSE_GATE_CNOT ? g.target : g.control);
                break;
            case QSE_GATE_CPHASE:
                o << ' ' << g.control << ' ' << g.target << ' ' << num(g.theta);
                break;
            case QSE_GATE_MCPHASE: {
                o << ' ';
                bool first = true;
                for (int q = 0; q < 64; ++q)
                    if ((g.controls >> q) & 1) {
                        o << (first ? "" : ",") << q;
                        first = false;
                    }
                if (first) throw std::runtime_error("gate list: MCPHASE without controls");
                o << ' ' << g.target << ' ' << num(g.theta);
                break;
            }
            case QSE_GATE_MATRIX:
                o << ' ' << g.target;
                for (int j = 0; j < 8; ++j) o << ' ' << num(g.m[j]);
                break;
            case QSE_GATE_MEASURE:
                if (g.control < 0) throw std::runtime_error("gate list: MEASURE without a result slot");
                o << ' ' << g.target << ' ' << g.control;
                break;
            default:
                o << ' ' << g.target;
                if (one_qubit_angle(g.kind)) o << ' ' << num(g.theta);
                break;
        }
        if (g.param >= 0) {
            if (!one_qubit_angle(g.kind) && g.kind != QSE_GATE_CPHASE)
                throw std::runtime_error(std::string("gate list: ") + kind_name(g.kind) + " cannot carry a parameter");
            o << " param " << g.param << " scale " << num(g.scale);
        }
        o << '\n';
    }
    return o.str();
}

namespace {

[[noreturn]] void fail(int line, const std::string& what) {
    throw std::runtime_error("gate list line " + std::to_string(line) + ": " + what);
}

long to_int(const std::string& s, int line) {
    if (s.empty()) fail(line, "missing integer");
    char* end = nullptr;
    errno = 0;
    const long v = std::strtol(s.c_str(), &end, 10);
    if (errno || *end) fail(line, "not an integer: '" + s + "'");
    return v;
}

double to_double(const std::string& s, int line) {
    if (s.empty()) fail(line, "missing number");
    char* end = nullptr;
    errno = 0;
    const double v = std::strtod(s.c_str(), &end);
    if (errno == ERANGE || *end) fail(line, "not a number: '" + s + "'");
    return v;
}

}  // namespace

void read(const std::string& text, int* n_out, std::vector<qse_op>& ops) {
    std::istringstream in(text);
    std::string raw;
    int line = 0, n = -1;
    bool header = false;
    ops.clear();
    while (std::getline(in, raw)) {
        ++line;
        const size_t hash = raw.find('#');
        if (hash != std::string::npos) raw.resize(hash);
        std::istringstream ls(raw);
        std::vector<std::string> f;
        for (std::string w; ls >> w;) f.push_back(w);
        if (f.empty()) continue;
        if (!header) {
            if (f.size() != 2 || f[0] != "qse-gates") fail(line, "expected 'qse-gates 1'");
            if (to_int(f[1], line) != 1) fail(line, "unsupported format version " + f[1]);
            header = true;
            continue;
        }
        if (n < 0) {
            if (f.size() != 2 || f[0] != "qubits") fail(line, "expected 'qubits <n>'");
            const long v = to_int(f[1], line);
            if (v < 1 || v > 63) fail(line, "qubit count outside [1, 63]");
            n = (int)v;
            continue;
        }
        qse_op g{};
        g.control = -1;
        g.param = -1;
        int kind = -1;
        for (const auto& k : kNames)
            if (f[0] == k.name) kind = k.kind;
        if (kind < 0) fail(line, "unknown gate '" + f[0] + "'");
        g.kind = kind;
        auto qubit = [&](const std::string& s) {
            const long q = to_int(s, line);
            if (q < 0 || q >= n) fail(line, "qubit " + s + " outside [0, " + std::to_string(n) + ")");
            return (int32_t)q;
        };
        // split off an optional parameter binding
        size_t nf = f.size();
        if (nf >= 5 && f[nf - 4] == "param" && f[nf - 2] == "scale") {
            const long p = to_int(f[nf - 3], line);
            if (p < 0 || p > 0x7fffffff) fail(line, "parameter index outside [0, 2^31)");
            g.param = (int32_t)p;
            g.scale = to_double(f[nf - 1], line);
            nf -= 4;
            if (!one_qubit_angle(kind) && kind != QSE_GATE_CPHASE) fail(line, f[0] + " cannot carry a parameter");
        }
        auto want = [&](size_t count) {
            if (nf != count) fail(line, f[0] + " takes " + std::to_string(count - 1) + " fields");
        };
        switch (kind) {
            case QSE_GATE_CNOT:
                want(3);
                g.control = qubit(f[1]);
                g.target = qubit(f[2]);
                break;
            case QSE_GATE_SWAP:
                want(3);
                g.target = qubit(f[1]);
                g.control = qubit(f[2]);
                break;
            case QSE_GATE_CPHASE:
                want(4);
                g.control = qubit(f[1]);
                g.target = qubit(f[2]);
                g.theta = to_double(f[3], line);
                break;
            case QSE_GATE_MCPHASE: {
                want(4);
                std::istringstream cs(f[1]);
                for (std::string c; std::getline(cs, c, ',');) {
                    const int q = qubit(c);
                    if ((g.controls >> q) & 1) fail(line, "repeated control qubit " + c);
                    g.controls |= 1ULL << q;
                }
                g.target = qubit(f[2]);
                g.theta = to_double(f[3], line);
                break;
            }
            case QSE_GATE_MATRIX:
                want(10);
                g.target = qubit(f[1]);
                for (int j = 0; j < 8; ++j) g.m[j] = to_double(f[2 + j], line);
                break;
            case QSE_GATE_MEASURE: {
                want(3);
                g.target = qubit(f[1]);
                const long slot = to_int(f[2], line);
                if (slot < 0 || slot > 0x7fffffff) fail(line, "result slot outside [0, 2^31)");
                g.control = (int32_t)slot;
                break;
            }
            default:
                if (one_qubit_angle(kind)) {
                    want(3);
                    g.target = qubit(f[1]);
                    g.theta = to_double(f[2], line);
                } else {
                    want(2);
                    g.target = qubit(f[1]);
                }
                break;
        }
        ops.push_back(g);
    }
    if (!header) throw std::runtime_error("gate list: empty input (no 'qse-gates 1' header)");
    if (n < 0) throw std::runtime_error("gate list: no 'qubits <n>' line");
    *n_out = n;
}

void save(const std::string& path, int n, const qse_op* ops, size_t nops) {
    const std::string text = write(n, ops, nops);
    std::ofstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("gate list: cannot write " + path);
    f << text;
    if (!f) throw std::runtime_error("gate list: write failed for " + path);
}

void load(const std::string& path, int* n, std::vector<qse_op>& ops) {
    std::ifstream f(path, std::ios::binary);
    if (!f) throw std::runtime_error("gate list: cannot read " + path);
    std::ostringstream s;
    s << f.rdbuf();
    read(s.str(), n, ops);
}

}  // namespace qse_gates
