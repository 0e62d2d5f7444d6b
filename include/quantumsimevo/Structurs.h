// Structurs.h — shared types of the state-vector simulator: the same names,
// numbering and layouts as the reference's QuantumSimEvo/Structurs.h:7-36
// (Complex, GateType, GateCommand, Gate2x2), so code written against the
// reference compiles unchanged.
#ifndef QSE_STRUCTURS_H
#define QSE_STRUCTURS_H
#include <complex>
#include <cstddef>
#include <string>
#include <vector>

using Complex = std::complex<double>;

// Same order as the reference enum (Structurs.h:10): the numeric values of the
// reference's members are the qse_op kinds of the C ABI (include/qse_hip.h).
// MATRIX, a generic 2x2 queued with Circuit::gate (ICircuitUnitaryOperation::
// applyGate), comes after them, so the reference's numbering is unchanged; code
// written against the reference enum sees it as an unknown gate type.
enum class GateType { IDENTITY, X, Y, Z, H, PHASE, RX, RY, RZ, CNOT, CPHASE, MCPHASE, SWAP, MEASURE, RESET_QUBIT,
                      MATRIX };

struct GateCommand {
    GateType type;
    size_t target;
    size_t control;                    // CNOT, CPHASE control; SWAP second qubit
    double theta;                      // rotation / phase angle
    std::string description;
    size_t resultIdx = 0;              // MEASURE: result slot; MATRIX: the Circuit's matrix table entry
    std::vector<size_t> controls_list; // MCPHASE controls
};

struct Gate2x2 {
    Complex data[2][2];

    Gate2x2() = default;
    Gate2x2(Complex a, Complex b, Complex c, Complex d) {
        data[0][0] = a;
        data[0][1] = b;
        data[1][0] = c;
        data[1][1] = d;
    }
    Complex* operator[](int row) { return data[row]; }
    const Complex* operator[](int row) const { return data[row]; }
};

#endif
