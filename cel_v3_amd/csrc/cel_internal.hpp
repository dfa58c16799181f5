// cel_internal.hpp -- shared declarations of the gfx950 library behind the
// C ABI in include/cel_gpu.h.  Host code is C++17, device code is HIP for
// CDNA4 (64-lane wavefronts).  Nothing here is visible outside the library.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/cel_gpu.h"

namespace cel {

// ---------------------------------------------------------------- errors
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);

#define CEL_HIP(call)                                                              \
    do {                                                                           \
        hipError_t e_ = (call);                                                    \
        if (e_ != hipSuccess)                                                      \
            return ::cel::fail(CEL_ERR_HIP, std::string(#call) + ": " +            \
                                                hipGetErrorString(e_));            \
    } while (0)

#define CEL_TRY(call)                      \
    do {                                   \
        int s_ = (call);                   \
        if (s_ < 0) return s_;             \
    } while (0)

// ---------------------------------------------------------------- runtime
struct Runtime {
    int device = -1;
    hipStream_t stream = nullptr;
    int num_cus = 0;
    bool ready = false;
};
Runtime& rt();
int ensure_init();

// Blocks of `kernel` (256 threads) resident on the whole GPU at once; a
// persistent grid never exceeds this, so no block waits for a second wave.
int resident_grid(const void* kernel, long long work_items);

// Device buffer owned by the library.
template <typename T>
struct DevBuf {
    T* p = nullptr;
    size_t n = 0;
    DevBuf() = default;
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    ~DevBuf() { release(); }
    int alloc(size_t count) {
        if (count == n && p) return CEL_OK;
        release();
        if (count == 0) return CEL_OK;
        if (hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)) != hipSuccess) {
            p = nullptr;
            return fail(CEL_ERR_OOM, "hipMalloc of " + std::to_string(count * sizeof(T)) + " bytes failed");
        }
        n = count;
        return CEL_OK;
    }
    // at least `count` elements, keeping a larger allocation (a workspace
    // reused from call to call)
    int reserve(size_t count) {
        if (p && n >= count) return CEL_OK;
        return alloc(count);
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
};

// The events, pinned host buffers and graph one host routine creates on a
// stream, released on every return path once the stream has drained (so no
// copy still targets a pinned buffer).
struct StreamScope {
    hipStream_t s;
    std::vector<hipEvent_t> ev;
    std::vector<void*> pinned;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    explicit StreamScope(hipStream_t st) : s(st) {}
    StreamScope(const StreamScope&) = delete;
    StreamScope& operator=(const StreamScope&) = delete;
    ~StreamScope()
    {
        (void)hipStreamSynchronize(s);
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        for (hipEvent_t e : ev) (void)hipEventDestroy(e);
        for (void* q : pinned) (void)hipHostFree(q);
    }
    int event(hipEvent_t& e, unsigned flags = hipEventDefault)
    {
        if (hipEventCreateWithFlags(&e, flags) != hipSuccess) return fail(CEL_ERR_HIP, "hipEventCreate failed");
        ev.push_back(e);
        return CEL_OK;
    }
    template <typename T>
    int host_alloc(T*& p, size_t count)
    {
        void* q = nullptr;
        if (hipHostMalloc(&q, sizeof(T) * (count > 0 ? count : 1), hipHostMallocDefault) != hipSuccess)
            return fail(CEL_ERR_OOM, "hipHostMalloc failed");
        pinned.push_back(q);
        p = static_cast<T*>(q);
        return CEL_OK;
    }
};

// ---------------------------------------------------------------- matrix
struct Workspace;  // PCG work vectors, owned by the matrix handle (pcg.hip)
// One halo block: entries of local rows whose columns belong to another
// rank's range, stored row-sorted.  Column indices are positions in that
// block's receive buffer (ghost slot), so the halo SpMV reads the packed
// face values directly instead of a full-length copy of the neighbour's
// vector.
struct HaloBlock {
    int proc = -1;        // owning rank of the columns
    int nnz = 0;
    int col_offset = 0;   // vtxdist(proc+1)-1, as the reference stores it
    int proc_len = 0;     // rows owned by proc
    int recv_len = 0;     // number of distinct referenced columns
    int recv_base = 0;    // offset of this block's slots in the ghost region
    // host copies kept for getters (reference layout, 1-based, global ids)
    std::vector<int> h_rows, h_cols;
    std::vector<double> h_vals;
    std::vector<int> h_xadj;       // cel_sp_mat_to_csr of this block
    std::vector<int> h_recv_perm;  // ascending local column ids (1-based)
};

struct Mat {
    int n = 0;          // local rows
    int ncols = 0;      // local columns (== n for a square local block)
    int nnz = 0;        // local block entries
    int row_offset = 0; // global row of local row 1 is row_offset+1
    int col_offset = 0; // subtracted from global columns of the local block

    // local block, CSR, 0-based on the device
    DevBuf<int> xadj, cols;
    DevBuf<double> vals;
    DevBuf<float> vals32;  // FP32 copy for the mixed-precision inner solve

    // diagonal, Jacobi vector, row scaling
    DevBuf<double> diag, jac, scaled_diag;
    DevBuf<float> jac32;
    bool has_diag = false, has_jac = false, jac_row_scaled = false, is_scaled = false;

    // JAD (0-based on the device)
    int jad_mmax = 0, jad_ilo = 0, jad_ndef = 0, jad_mode = -1;
    long long jad_nnz = 0;
    DevBuf<int> jad_perm, jad_begin, jad_cols;
    DevBuf<double> jad_vals;
    DevBuf<float> jad_vals32;
    // the JAD columns renumbered into the permuted row order (pinv[cols]),
    // for solves that keep their vectors in that order; built on first use
    DevBuf<int> jad_pcols;

    // DIA (cel_gpu_mat_to_dia): values [ndiag][n], presence mask per row
    int dia_nd = 0;                 // 0: no DIA form
    int dia_off[32] = {};           // ascending column - row offsets
    DevBuf<double> dia_vals;
    DevBuf<float> dia_vals32;
    DevBuf<uint32_t> dia_mask;      // presence bits per row (more than 8 diagonals)
    DevBuf<uint8_t> dia_mask8;      // the same in one byte (at most 8 diagonals)
    std::vector<int> h_jad_begin;  // host copy (0-based starts, mmax+1)

    // halo (distributed block split of one rank)
    std::vector<HaloBlock> halo;
    int ghost_len = 0;             // sum of recv_len over halo blocks
    DevBuf<int> halo_xadj;         // CSR over local rows of the merged halo entries (n+1)
    DevBuf<int> halo_rows;         // rows having halo entries (compact list)
    DevBuf<int> halo_rowq;         // per local row: its compact index, or -1
    DevBuf<uint8_t> halo_flag;     // per local row: min(halo entries, 255); [n] = 0
    DevBuf<int> halo_wq;           // per 64-row group g: halo rows below row 64 g
    DevBuf<int> halo_ptr;          // per compact row: start into halo_cols/vals (+1 end)
    DevBuf<int> halo_cols;         // ghost slot per entry, in block order
    DevBuf<double> halo_vals;
    int halo_nrows = 0;
    DevBuf<int> send_perm;         // 0-based local indices packed per send list
    std::vector<int> send_len, send_base, send_proc;
    int send_total = 0;
    int nranks = 1, my_rank = 0;
    std::vector<int> vtxdist;      // 1-based starts, nranks+1
    std::vector<double> h_diag_split;  // diagonal captured during the split
    std::vector<int> h_local_rows, h_local_cols;  // local block in input order (global ids)
    std::vector<double> h_local_vals;

    Workspace* ws = nullptr;
};

void pcg_release_workspace(Mat* A);

}  // namespace cel

// the opaque handle of the C ABI
struct cel_gpu_mat {
    cel::Mat m;
};

namespace cel {

// ---------------------------------------------------------------- kernels (host launchers)
// All launches go to rt().stream.
int launch_gen_poisson7(int nx, int ny, int nz, const double* kappa, int* xadj, int* cols,
                        double* vals);
// skip (may be null): the launch returns at once when *skip is non-zero on
// the device (for launches captured in a graph whose later replays must do
// nothing)
// coo = 1: the COO product of the same (row-sorted) entries, y(row) += a x
// entry by entry (from y itself when accumulating, else from 0)
int launch_csr_spmv(const Mat& A, const double* x, double* y, int accumulate,
                    const int* skip = nullptr, int coo = 0);
int launch_csr_spmv_f32(const Mat& A, const float* x, float* y, int accumulate, int coo = 0);
int launch_jad_spmv(const Mat& A, const double* x, double* y, int accumulate,
                    const int* skip = nullptr);
int launch_jad_spmv_f32(const Mat& A, const float* x, float* y, int accumulate);
int launch_dia_spmv(const Mat& A, const double* x, double* y, int accumulate,
                    const int* skip = nullptr);
int launch_dia_spmv_f32(Mat& A, const float* x, float* y, int accumulate);
// FP64 products fused with the dot product that follows them in the GMRES
// loop, over G fixed row chunks of `chunk` rows (a multiple of 256): block b
// covers rows [b*chunk, b*chunk + chunk) and writes part[b].
//   kFuseDot:   y = A x,        part[b] = sum of aux_i * y_i over chunk b
//   kFuseResid: y = aux - A x,  part[b] = sum of y_i * y_i over chunk b
// y holds the plain product's values bit for bit, and each thread adds its
// rows in ascending order before the block sum, so part[b] equals what a
// separate chunked dot kernel forms over the same rows.
enum { kFuseDot = 1, kFuseResid = 2 };
int launch_csr_spmv_fused(const Mat& A, int fuse, long long chunk, int G, const double* x,
                          double* y, const double* aux, double* part, const int* skip);
int launch_dia_spmv_fused(const Mat& A, int fuse, long long chunk, int G, const double* x,
                          double* y, const double* aux, double* part, const int* skip);
int dia_prepare_f32(Mat& A);
// Sets bit `bit` (= dia_nd) of the DIA row masks on the rows with halo
// entries and clears it elsewhere; fails when the masks have no spare bit.
int launch_halo_spmv(const Mat& A, const double* ghost, double* y);
int launch_face_pack(const Mat& A, const double* x, double* sendbuf);
int launch_extract_diag(Mat& A);
int launch_set_jacobi(Mat& A, int row_scaled);
int launch_scale_rows(Mat& A);
int launch_scale_vec(const Mat& A, double* b);
int launch_to_f32(const double* src, float* dst, size_t n);
int launch_coo_to_csr(int nrows, int nnz, const int* rows0, int* xadj);  // rows 0-based local
// Frees the workspace the COO sort and row-offset routines keep between calls.
void coo_release_workspace();
int jad_convert_device(Mat& A, int mode);
// Builds A.jad_pcols (once per conversion).
int jad_prepare_permuted(Mat& A);
int launch_sddmm(const Mat& A, const double* lambda, const double* x, double* g);
int launch_kappa_grad(int nx, int ny, int nz, const double* kappa, const double* x,
                      const double* lambda, double* grad);

// bytes one y = A x moves, counted algorithmically (index + value streams,
// row pointers, one read of x and one write of y)
double spmv_bytes(const Mat& A, int format, int precision);

}  // namespace cel
