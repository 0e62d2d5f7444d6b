// HardwareCheck.cpp — device report (reference HardwareCheck.cpp:35-98: device
// name, memory, largest state). maxQubits = floor(log2(0.8 * HBM / bytes per amplitude)).
#include "quantumsimevo/HardwareCheck.h"

#include <cmath>
#include <iostream>
#include <thread>

#include "qse_hip.h"

DeviceReport HardwareCheck::report(int device) {
    DeviceReport r;
    r.hostThreads = std::thread::hardware_concurrency();
    int count = 0;
    if (qse_device_count(&count) != QSE_OK || device >= count) return r;
    char name[256] = {0};
    if (qse_device_info(device, name, sizeof(name), &r.computeUnits, &r.totalBytes) != QSE_OK) return r;
    qse_device_mem_info(device, &r.freeBytes, nullptr);
    r.available = true;
    r.name = name;
    r.maxQubitsC128 = (int)std::floor(std::log2(0.8 * (double)r.totalBytes / 16.0));
    r.maxQubitsC64 = (int)std::floor(std::log2(0.8 * (double)r.totalBytes / 8.0));
    return r;
}

void HardwareCheck::printReport(int device) {
    const DeviceReport r = report(device);
    std::cout << "Host threads: " << r.hostThreads << "\n";
    if (!r.available) {
        std::cout << "HIP device " << device << ": not available\n";
        return;
    }
    std::cout << "HIP device " << device << ": " << r.name << ", " << r.computeUnits << " CUs, "
              << r.totalBytes / double(1ULL << 30) << " GiB HBM (" << r.freeBytes / double(1ULL << 30) << " GiB free)\n"
              << "Max qubits: " << r.maxQubitsC128 << " (complex128), " << r.maxQubitsC64 << " (complex64)\n";
}
