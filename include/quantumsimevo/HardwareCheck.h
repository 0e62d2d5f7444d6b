// HardwareCheck.h — host/device report (reference QuantumSimEvo/HardwareCheck.h:12-30,
// HardwareCheck.cpp:15-98) for the HIP device: name, compute units, HBM, and the
// largest state that fits (complex128 and complex64).
#pragma once
#include <cstddef>
#include <string>

struct DeviceReport {
    bool available = false;
    std::string name;
    int computeUnits = 0;
    size_t totalBytes = 0, freeBytes = 0;
    int maxQubitsC128 = 0, maxQubitsC64 = 0;
    unsigned hostThreads = 0;
};

class HardwareCheck {
public:
    static DeviceReport report(int device = 0);
    static void printReport(int device = 0);
};
