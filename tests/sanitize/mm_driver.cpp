// Sanitizer harness for the Matrix Market reader (cel_v3_amd/csrc/mmio.cpp),
// host code built with AddressSanitizer and UndefinedBehaviorSanitizer.
// For every file named on the command line it calls cel_gpu_mm_info and, on
// success, cel_gpu_mm_read with output arrays of exactly the needed length,
// one entry short, and room for the symmetric expansion; every result code
// is printed.  A memory error or undefined behaviour aborts the run with the
// sanitizer's report.
#include <cstdio>
#include <string>
#include <vector>

#include "../../include/cel_gpu.h"

namespace cel {
// the library's error sink (capi.hip keeps the message for cel_gpu_last_error)
std::string last_msg;
int fail(int code, const std::string& msg)
{
    last_msg = msg;
    return code;
}
}  // namespace cel

static int read_into(const char* path, int flags, int cap)
{
    // arrays of exactly `cap` entries, so an overrun is a heap overflow
    std::vector<int> rows((size_t)cap), cols((size_t)cap);
    std::vector<double> vals((size_t)cap);
    int got = -1;
    int rc = cel_gpu_mm_read(path, flags, cap, rows.data(), cols.data(), vals.data(), &got);
    if (rc == CEL_OK && (got < 0 || got > cap)) {
        std::printf("%s: read returned %d entries for capacity %d\n", path, got, cap);
        return 99;
    }
    return rc;
}

int main(int argc, char** argv)
{
    int bad = 0;
    for (int i = 1; i < argc; ++i) {
        const char* path = argv[i];
        int rep = -1, field = -1, symm = -1, nr = -1, nc = -1, nnz = -1;
        int rc = cel_gpu_mm_info(path, &rep, &field, &symm, &nr, &nc, &nnz);
        std::printf("%s info %d", path, rc);
        if (rc == CEL_OK && nnz >= 0 && nnz <= (1 << 22)) {
            int r0 = read_into(path, CEL_MM_AS_STORED, nnz);
            int r1 = nnz > 0 ? read_into(path, CEL_MM_AS_STORED, nnz - 1) : 0;
            int r2 = read_into(path, CEL_MM_EXPAND_SYMMETRIC, nnz);
            int r3 = read_into(path, CEL_MM_EXPAND_SYMMETRIC, 2 * nnz);
            std::printf(" read %d %d %d %d", r0, r1, r2, r3);
            bad += (r0 == 99) + (r1 == 99) + (r2 == 99) + (r3 == 99);
        }
        std::printf("\n");
    }
    return bad ? 1 : 0;
}
