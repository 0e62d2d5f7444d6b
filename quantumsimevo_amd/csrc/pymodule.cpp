// pymodule.cpp — pybind11 module `_quantumsimevo`: the reference's Python API
// (QSEPythonModule/QuantumSimEvoPy.cpp:8-41: Circuit with snake_case gate
// methods, execute/measure/sample/reset/print_state) on the HIP backend, plus
// state I/O, expectation values and seeded readout.
#include <algorithm>
#include <cstring>
#include <thread>
#include <vector>
#include <pybind11/complex.h>
#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include "quantumsimevo/Circuit.h"
#include "quantumsimevo/GateListFile.h"
#include "quantumsimevo/CircuitUnitaryOperationHIP.h"
#include "quantumsimevo/HardwareCheck.h"
#include "quantumsimevo/Noise.h"

namespace py = pybind11;

static int parse_precision(const py::object& p) {
    if (py::isinstance<py::str>(p)) {
        const std::string s = p.cast<std::string>();
        if (s == "c128" || s == "complex128") return QSE_C128;
        if (s == "c64" || s == "complex64") return QSE_C64;
        throw std::invalid_argument("precision must be 'c128' or 'c64'");
    }
    return p.cast<int>();
}

// "deferred" (the reference's frame after the last gate) or "interleaved".
static NoiseMode parse_noise_mode(const std::string& mode) {
    if (mode == "deferred") return NoiseMode::Deferred;
    if (mode == "interleaved") return NoiseMode::Interleaved;
    throw std::invalid_argument("noise_mode must be 'deferred' or 'interleaved'");
}

PYBIND11_MODULE(_quantumsimevo, m) {
    m.doc() = "QuantumSimEvo circuits on the MI355X (gfx950) HIP backend";

    py::class_<Circuit>(m, "Circuit")
        .def(py::init([](size_t n, bool noise, py::object precision, const std::string& noise_mode) {
                 auto* c = new Circuit(n, noise, parse_precision(precision));
                 c->setNoiseMode(parse_noise_mode(noise_mode));
                 return c;
             }),
             py::arg("num_qubits"), py::arg("simulate_pauli_noise") = false, py::arg("precision") = "c128",
             py::arg("noise_mode") = "deferred")
        .def("set_noise_mode", [](Circuit& c, const std::string& mode) { c.setNoiseMode(parse_noise_mode(mode)); },
             py::arg("mode"))
        .def_property_readonly("noise_mode", [](const Circuit& c) {
            return std::string(c.getNoiseMode() == NoiseMode::Interleaved ? "interleaved" : "deferred");
        })
        // single-qubit gates (QuantumSimEvoPy.cpp:16-24)
        .def("hadamard", &Circuit::hadamard, py::arg("q"))
        .def("pauli_x", &Circuit::pauliX, py::arg("q"))
        .def("pauli_y", &Circuit::pauliY, py::arg("q"))
        .def("pauli_z", &Circuit::pauliZ, py::arg("q"))
        .def("phase", &Circuit::phase, py::arg("theta"), py::arg("q"))
        .def("rotate_x", &Circuit::rotateX, py::arg("theta"), py::arg("q"))
        .def("rotate_y", &Circuit::rotateY, py::arg("theta"), py::arg("q"))
        .def("rotate_z", &Circuit::rotateZ, py::arg("theta"), py::arg("q"))
        // multi-qubit gates (QuantumSimEvoPy.cpp:26-32)
        .def("cnot", &Circuit::cnot, py::arg("control"), py::arg("target"))
        .def("cphase", &Circuit::cphase, py::arg("theta"), py::arg("control"), py::arg("target"))
        .def("mcphase", &Circuit::mcphase, py::arg("theta"), py::arg("controls"), py::arg("target"))
        .def("swap", &Circuit::swap, py::arg("q0"), py::arg("q1"))
        // execution & measurement (QuantumSimEvoPy.cpp:34-38)
        .def("execute", &Circuit::execute, py::arg("print_steps") = false)
        .def("measure", &Circuit::measure)
        .def("sample", &Circuit::sample, py::arg("num_shots"))
        .def("reset", &Circuit::reset)
        .def("print_state", &Circuit::printState)
        // extensions
        .def("gate", [](Circuit& c, py::array_t<std::complex<double>, py::array::c_style | py::array::forcecast> mat,
                        size_t q) {
                 if (mat.size() != 4) throw std::invalid_argument("gate matrix must have 4 entries");
                 auto r = mat.unchecked();
                 const std::complex<double>* p = mat.data();
                 (void)r;
                 c.gate(Gate2x2(p[0], p[1], p[2], p[3]), q);
             },
             py::arg("matrix"), py::arg("q"))
        .def("set_seed", &Circuit::setSeed, py::arg("seed"))
        .def("state", [](Circuit& c) {
            std::vector<Complex> s = c.state();
            py::array_t<std::complex<double>> out(s.size());
            std::copy(s.begin(), s.end(), out.mutable_data());
            return out;
        })
        .def("set_state", [](Circuit& c, py::array_t<std::complex<double>, py::array::c_style | py::array::forcecast> a) {
                 std::vector<Complex> v(a.data(), a.data() + a.size());
                 c.setState(v);
             },
             py::arg("amplitudes"))
        .def("expectation_z", &Circuit::expectationZ, py::arg("masks"), py::arg("weights"))
        .def("measure_qubit", &Circuit::measureQubit, py::arg("q"))
        .def("reset_qubit", &Circuit::resetQubit, py::arg("q"))
        .def("queue_measure", &Circuit::queueMeasure, py::arg("q"))
        .def("queue_reset", &Circuit::queueReset, py::arg("q"))
        .def_property_readonly("measurement_results", &Circuit::measurementResults)
        // the Circuit's StateVector host buffer as it stands (Circuit::getStateVector
        // does not flush, Circuit.h:53)
        .def("host_state", [](Circuit& c) {
            StateVector& sv = c.getStateVector();
            py::array_t<std::complex<double>> out(sv.size());
            std::copy(sv.data().begin(), sv.data().end(), out.mutable_data());
            return out;
        })
        .def_property_readonly("num_qubits", &Circuit::numQubits)
        .def_property_readonly("num_commands", [](const Circuit& c) { return c.commands().size(); })
        .def("noise_stats", [](Circuit& c) {
            return py::dict(py::arg("bit_flips") = c.noise().getTotalBitFlips(),
                            py::arg("phase_flips") = c.noise().getTotalPhaseFlips(),
                            py::arg("x_mask") = c.noise().bitFlipMask(),
                            py::arg("z_mask") = c.noise().phaseFlipMask());
        })
        .def("set_noise_rate",
             [](Circuit& c, int gate_type, double rate) { c.noise().setRate((GateType)gate_type, rate); },
             py::arg("gate_type"), py::arg("rate"))
        .def("set_fusion", [](Circuit& c, bool on) { c.backend().setFusion(on); }, py::arg("fused"))
        // gate-list files (GateListFile.h)
        .def("save_gates", &Circuit::saveGates, py::arg("path"))
        .def("load_gates", &Circuit::loadGates, py::arg("path"))
        .def("gate_list", [](const Circuit& c) {
            const std::vector<qse_op> ops = c.gateList();
            return py::bytes(reinterpret_cast<const char*>(ops.data()), ops.size() * sizeof(qse_op));
        });

    // The reference's StateVector and backend interface as they stand
    // (StateVector.h:17-47, ICircuitUnitaryOperation.h:6-35), for driving the
    // device-mirror contract (flush / markHostDirty / rebinding) directly.
    py::class_<StateVector>(m, "StateVector")
        .def(py::init<size_t>(), py::arg("num_qubits"))
        .def_property_readonly("num_qubits", &StateVector::qubits)
        .def_property_readonly("size", &StateVector::size)
        .def_property_readonly("host_allocated", &StateVector::hostAllocated)
        // copy of the host buffer (allocating it if needed); no device access
        .def("host_data", [](StateVector& sv) {
            py::array_t<std::complex<double>> out(sv.size());
            std::copy(sv.data().begin(), sv.data().end(), out.mutable_data());
            return out;
        })
        // overwrite the host buffer (the caller then marks the host dirty, as in Circuit::reset)
        .def("set_host_data",
             [](StateVector& sv, py::array_t<std::complex<double>, py::array::c_style | py::array::forcecast> a) {
                 if ((size_t)a.size() != sv.size()) throw std::invalid_argument("size mismatch");
                 std::copy(a.data(), a.data() + a.size(), sv.data().begin());
             },
             py::arg("amplitudes"))
        .def("reset", &StateVector::reset);
    py::class_<CircuitUnitaryOperationHIP>(m, "HIPBackend")
        .def(py::init([](int device, py::object precision) {
                 return new CircuitUnitaryOperationHIP(device, parse_precision(precision));
             }),
             py::arg("device") = 0, py::arg("precision") = "c128")
        .def("apply_hadamard", &CircuitUnitaryOperationHIP::applyHadamard, py::keep_alive<1, 2>())
        .def("apply_pauli_x", &CircuitUnitaryOperationHIP::applyPauliX, py::keep_alive<1, 2>())
        .def("apply_pauli_y", &CircuitUnitaryOperationHIP::applyPauliY, py::keep_alive<1, 2>())
        .def("apply_pauli_z", &CircuitUnitaryOperationHIP::applyPauliZ, py::keep_alive<1, 2>())
        .def("apply_phase", &CircuitUnitaryOperationHIP::applyPhase, py::keep_alive<1, 2>())
        .def("apply_rotate_x", &CircuitUnitaryOperationHIP::applyRotateX, py::keep_alive<1, 2>())
        .def("apply_rotate_y", &CircuitUnitaryOperationHIP::applyRotateY, py::keep_alive<1, 2>())
        .def("apply_rotate_z", &CircuitUnitaryOperationHIP::applyRotateZ, py::keep_alive<1, 2>())
        .def("apply_cnot", &CircuitUnitaryOperationHIP::applyCNOT, py::keep_alive<1, 2>())
        .def("apply_cphase", &CircuitUnitaryOperationHIP::applyCPhase, py::keep_alive<1, 2>())
        .def("apply_mcphase", &CircuitUnitaryOperationHIP::applyMCPhase, py::keep_alive<1, 2>())
        .def("apply_swap", &CircuitUnitaryOperationHIP::applySwap, py::keep_alive<1, 2>())
        .def("flush", [](CircuitUnitaryOperationHIP& b) { b.flush(); })
        .def("mark_host_dirty", [](CircuitUnitaryOperationHIP& b) { b.markHostDirty(); })
        .def("set_print_steps", [](CircuitUnitaryOperationHIP& b, bool on) { b.setPrintSteps(on); }, py::arg("on"));

    m.def("device_report", [](int device) {
              const DeviceReport r = HardwareCheck::report(device);
              return py::dict(py::arg("available") = r.available, py::arg("name") = r.name,
                              py::arg("compute_units") = r.computeUnits, py::arg("total_bytes") = r.totalBytes,
                              py::arg("free_bytes") = r.freeBytes, py::arg("max_qubits_c128") = r.maxQubitsC128,
                              py::arg("max_qubits_c64") = r.maxQubitsC64, py::arg("host_threads") = r.hostThreads);
          },
          py::arg("device") = 0);
    m.def("print_report", &HardwareCheck::printReport, py::arg("device") = 0);
    // Gate-list text (GateListFile.h) from / to raw qse_op records (104 bytes each).
    m.def("write_gate_list",
          [](int n, py::bytes records) {
              const std::string b = records;
              if (b.size() % sizeof(qse_op)) throw std::invalid_argument("records are not whole qse_op structs");
              std::vector<qse_op> ops(b.size() / sizeof(qse_op));
              if (!ops.empty()) std::memcpy(ops.data(), b.data(), b.size());
              return qse_gates::write(n, ops.data(), ops.size());
          },
          py::arg("n"), py::arg("records"));
    m.def("read_gate_list",
          [](const std::string& text) {
              int n = 0;
              std::vector<qse_op> ops;
              qse_gates::read(text, &n, ops);
              return py::make_tuple(n, py::bytes(reinterpret_cast<const char*>(ops.data()), ops.size() * sizeof(qse_op)));
          },
          py::arg("text"));
    // Pauli frames of many noise trajectories: per seed, the Noise draw process
    // (Noise.cpp:21-32, two draws per eligible gate in replay order) over the
    // gate list (GateType kinds, targets); rate >= 0 replaces every eligible
    // type's rate. Returns (bit-flip masks, phase-flip masks), one per seed.
    m.def("noise_frames",
          [](size_t n, py::array_t<int, py::array::c_style | py::array::forcecast> kinds,
             py::array_t<int, py::array::c_style | py::array::forcecast> targets,
             py::array_t<unsigned int, py::array::c_style | py::array::forcecast> seeds, double rate,
             double multiplier) {
              if (kinds.size() != targets.size()) throw std::invalid_argument("kinds and targets differ in length");
              if (n < 1 || n > 64) throw std::invalid_argument("1..64 qubits");
              const int* k = kinds.data();
              const int* t = targets.data();
              const unsigned int* sd = seeds.data();
              const ssize_t ng = kinds.size(), nt = seeds.size();
              py::array_t<uint64_t> xs(nt), zs(nt);
              uint64_t* x = xs.mutable_data();
              uint64_t* z = zs.mutable_data();
              {
                  py::gil_scoped_release nogil;
                  // trajectories are independent draw processes: split them over host threads
                  auto work = [&](ssize_t j0, ssize_t j1) {
                      for (ssize_t j = j0; j < j1; ++j) {
                          Noise nz(n, sd[j], multiplier);
                          if (rate >= 0.0)
                              for (GateType g : {GateType::X, GateType::Y, GateType::Z, GateType::H, GateType::CNOT,
                                                 GateType::PHASE, GateType::RX, GateType::RY, GateType::RZ})
                                  nz.setRate(g, rate);
                          for (ssize_t i = 0; i < ng; ++i) nz.accumulateError((GateType)k[i], (size_t)t[i]);
                          x[j] = nz.bitFlipMask();
                          z[j] = nz.phaseFlipMask();
                      }
                  };
                  const ssize_t nthr = std::min<ssize_t>(
                      {nt / 16 + 1, (ssize_t)std::max(1u, std::thread::hardware_concurrency()), (ssize_t)16});
                  std::vector<std::thread> pool;
                  for (ssize_t w = 1; w < nthr; ++w) pool.emplace_back(work, nt * w / nthr, nt * (w + 1) / nthr);
                  work(0, nt / nthr);
                  for (auto& th : pool) th.join();
              }
              return py::make_tuple(xs, zs);
          },
          py::arg("n"), py::arg("kinds"), py::arg("targets"), py::arg("seeds"), py::arg("rate") = -1.0,
          py::arg("multiplier") = 1.0);
    // The same draws per trajectory as noise_frames, as insertions for the
    // interleaved mode: trajectory j's flips are (gate[i], code[i]) for i in
    // [offsets[j], offsets[j+1]), gates ascending, code bit 0 = X and bit 1 = Z on
    // that gate's target, acting right after the gate.
    m.def("noise_insertions",
          [](size_t n, py::array_t<int, py::array::c_style | py::array::forcecast> kinds,
             py::array_t<int, py::array::c_style | py::array::forcecast> targets,
             py::array_t<unsigned int, py::array::c_style | py::array::forcecast> seeds, double rate,
             double multiplier) {
              if (kinds.size() != targets.size()) throw std::invalid_argument("kinds and targets differ in length");
              if (n < 1 || n > 64) throw std::invalid_argument("1..64 qubits");
              const int* k = kinds.data();
              const int* t = targets.data();
              const unsigned int* sd = seeds.data();
              const ssize_t ng = kinds.size(), nt = seeds.size();
              std::vector<std::vector<std::pair<int32_t, int8_t>>> per(nt);
              {
                  py::gil_scoped_release nogil;
                  auto work = [&](ssize_t j0, ssize_t j1) {
                      for (ssize_t j = j0; j < j1; ++j) {
                          Noise nz(n, sd[j], multiplier);
                          if (rate >= 0.0)
                              for (GateType g : {GateType::X, GateType::Y, GateType::Z, GateType::H, GateType::CNOT,
                                                 GateType::PHASE, GateType::RX, GateType::RY, GateType::RZ})
                                  nz.setRate(g, rate);
                          for (ssize_t i = 0; i < ng; ++i) {
                              const int d = nz.drawError((GateType)k[i], (size_t)t[i]);
                              if (d) per[j].push_back({(int32_t)i, (int8_t)d});
                          }
                      }
                  };
                  const ssize_t nthr = std::min<ssize_t>(
                      {nt / 16 + 1, (ssize_t)std::max(1u, std::thread::hardware_concurrency()), (ssize_t)16});
                  std::vector<std::thread> pool;
                  for (ssize_t w = 1; w < nthr; ++w) pool.emplace_back(work, nt * w / nthr, nt * (w + 1) / nthr);
                  work(0, nt / nthr);
                  for (auto& th : pool) th.join();
              }
              size_t total = 0;
              for (const auto& v : per) total += v.size();
              py::array_t<int64_t> offsets(nt + 1);
              py::array_t<int32_t> gate(total);
              py::array_t<int8_t> code(total);
              int64_t* o = offsets.mutable_data();
              int32_t* g = gate.mutable_data();
              int8_t* c = code.mutable_data();
              o[0] = 0;
              for (ssize_t j = 0; j < nt; ++j) {
                  for (size_t i = 0; i < per[j].size(); ++i) {
                      g[o[j] + i] = per[j][i].first;
                      c[o[j] + i] = per[j][i].second;
                  }
                  o[j + 1] = o[j] + (int64_t)per[j].size();
              }
              return py::make_tuple(offsets, gate, code);
          },
          py::arg("n"), py::arg("kinds"), py::arg("targets"), py::arg("seeds"), py::arg("rate") = -1.0,
          py::arg("multiplier") = 1.0);
}
