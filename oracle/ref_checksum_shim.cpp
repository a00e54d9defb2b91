// Exposes the reference's calculateChecksum (io/FileUtils.cuh:5-17, compiled
// from the reference's own file by oracle/Makefile) under a C name for ctypes.
// Test infrastructure only.
#include <cstddef>
#include <cstdint>
uint32_t calculateChecksum(const void* data_to_check, size_t size_of_data);
extern "C" uint32_t ref_calculate_checksum(const void* p, size_t n) { return calculateChecksum(p, n); }
