// GateListFile.h — a stable text format for gate lists (no counterpart in the
// reference, whose circuits exist only as GateCommand vectors built in code,
// Structurs.h:18-26, Circuit.cpp:9-62).
//
// One file holds one circuit:
//
//   qse-gates 1
//   qubits <n>
//   <gate> ...            one gate per line, in application order
//
// Gate lines (qubit indices are decimal, angles are written with 17
// significant digits so that they read back to the same double):
//
//   I q | X q | Y q | Z q | H q
//   PHASE q theta | RX q theta | RY q theta | RZ q theta      [param p scale s]
//   CNOT control target
//   CPHASE control target theta                               [param p scale s]
//   MCPHASE c1,c2,... target theta
//   SWAP q0 q1
//   MATRIX q re00 im00 re01 im01 re10 im10 re11 im11           (row-major 2x2)
//   MEASURE q slot        mid-circuit measurement of qubit q, outcome to slot
//   RESET q               mid-circuit reset of qubit q to |0>
//
// "param p scale s" binds the angle to a parameter vector: the gate's angle is
// theta + s * params[p] (the qse_op fields of the C ABI, used by the adjoint
// and batched runs). '#' starts a comment; blank lines are ignored.
#ifndef QSE_GATE_LIST_FILE_H
#define QSE_GATE_LIST_FILE_H
#include <string>
#include <vector>

#include "../qse_hip.h"

namespace qse_gates {

// Text of a gate list of n qubits.
std::string write(int n, const qse_op* ops, size_t nops);

// Parse a gate list; throws std::runtime_error naming the line on malformed
// input (unknown gate, missing or extra field, qubit outside [0, n), duplicate
// or equal qubits where the gate needs distinct ones).
void read(const std::string& text, int* n, std::vector<qse_op>& ops);

// File helpers around write/read.
void save(const std::string& path, int n, const qse_op* ops, size_t nops);
void load(const std::string& path, int* n, std::vector<qse_op>& ops);

}  // namespace qse_gates

#endif
